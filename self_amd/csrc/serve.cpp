// serve.cpp -- the request loop: shallow-water integrations and CGSEM
// elliptic solves of different mesh sizes and degrees, admitted as they
// arrive, run concurrently on the device, with per-request latency.
//
// No counterpart in the reference (its drivers run one model to
// completion).  Design:
//  * submit() enqueues a request (the caller's host arrays stay valid until
//    wait() returns for it) and stamps its arrival time;
//  * W worker threads each own a model at a time; a worker takes the
//    oldest queued request whose device-memory estimate fits in the free
//    HBM left after the bytes reserved by running requests (admission
//    against free HBM; requests that do not fit wait for memory to free);
//  * each model has its own HIP stream (default priority), so the kernels
//    of concurrent requests overlap on the device;
//  * continuous batching: when a worker takes a shallow-water request, the
//    queued requests that can share its ensemble (same mesh, degree,
//    precision, steps, constants and bathymetry) join it, up to 8 members
//    and as memory allows, and advance in one fused launch per RK stage,
//    each with its own state and time step; further queued shallow-water
//    requests of the same precision and step count, on other meshes or
//    degrees, form up to 7 more such groups, and all groups advance
//    together in one persistent launch per RK stage
//    (self_sw2d_forward_step_rk3_multi) -- for small models only (at most
//    16384 nodes per variable), where the stages are launch-bound;
//  * results are written to the caller's output array; wait() returns the
//    request's status, queue time, run time, latency and ensemble size.
// Every kernel of both models is deterministic (fixed reduction trees, no
// atomics on values), so a request's result is bitwise the same alone and
// in any mix.  Meshes and nodal storage are cached per (kind, nX, N).

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct MeshCache {
    int N = 0, nX = 0, nE = 0, nEdges = 0;
    std::vector<double> nodes, w, dS, dP, lW, lE, lS, lN;
    std::vector<int> edges, elemNodes;
    std::vector<double> x, y, dxds, dxdp, dyds, dydp, J, nHat, nLen;
};

struct Job {
    long long id = 0;
    self_request req{};
    size_t bytes = 0;
    uint64_t inHash[2] = {0, 0};  // SW: two independent hashes of the bathymetry
    Clock::time_point arrive, start, end;
    bool done = false;
    self_request_result res{};
};

}  // namespace

struct self_server_s {
    int device = 0;
    double memFraction = 0.9;
    size_t reserved = 0, budget = 0;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<std::shared_ptr<Job>> queue;
    std::unordered_map<long long, std::shared_ptr<Job>> jobs;
    // host meshes by (kind, nX, N, bit pattern of the SW domain side L; 0 for CGSEM)
    std::map<std::tuple<int, int, int, uint64_t>, std::shared_ptr<MeshCache>> meshes;
    std::vector<std::thread> workers;
    long long nextId = 1;
    int running = 0;
    bool stop = false;
};

namespace {

// host mesh + nodal storage for a request kind: SW (Gauss, periodic square
// of side L) or CGSEM (Gauss-Lobatto, walled unit square)
std::shared_ptr<MeshCache> mesh_for(self_server_s *s, int kind, int nX, int N, double L)
{
    const bool sw = kind == SELF_REQ_SW2D;
    uint64_t Lbits = 0;  // the SW mesh is scaled by L; the CGSEM mesh is the unit square
    if (sw) std::memcpy(&Lbits, &L, sizeof(Lbits));
    const auto key = std::make_tuple(kind, nX, N, Lbits);
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->meshes.find(key);
        if (it != s->meshes.end()) return it->second;
    }
    auto m = std::make_shared<MeshCache>();
    const int n1 = N + 1, np = n1 * n1;
    m->N = N;
    m->nX = nX;
    m->nodes.resize(n1); m->w.resize(n1); m->dS.resize(np); m->dP.resize(np);
    m->lW.resize(n1); m->lE.resize(n1); m->lS.resize(n1); m->lN.resize(n1);
    if (self_nodal2d_build(N, sw ? SELF_GAUSS : SELF_GAUSS_LOBATTO, sw ? SELF_DG : SELF_CG, 0, m->nodes.data(),
                           m->w.data(), m->dS.data(), m->dP.data(), m->lW.data(), m->lE.data(), m->lS.data(),
                           m->lN.data()) != SELF_OK)
        return nullptr;
    const int per = sw ? 1 : 0;
    if (self_quadmesh_counts(nX, nX, per, per, &m->nE, &m->nEdges) != SELF_OK) return nullptr;
    const size_t nv = (size_t)m->nE * np, ns = (size_t)m->nE * 4 * n1;
    m->edges.resize(6 * (size_t)m->nEdges);
    for (auto *v : {&m->x, &m->y, &m->dxds, &m->dxdp, &m->dyds, &m->dydp, &m->J}) v->resize(nv);
    m->nHat.resize(2 * ns);
    m->nLen.resize(ns);
    if (self_quadmesh_structured(nX, nX, N, per, per, sw ? L : 1.0, sw ? L : 1.0, 0.0, m->nodes.data(),
                                 m->edges.data(), m->x.data(), m->y.data(), m->dxds.data(), m->dxdp.data(),
                                 m->dyds.data(), m->dydp.data(), m->J.data(), m->nHat.data(),
                                 m->nLen.data()) != SELF_OK)
        return nullptr;
    if (!sw) {  // corner node ids of the walled default mesh: SW, SE, NE, NW (1-based)
        m->elemNodes.resize(4 * (size_t)m->nE);
        for (int iY = 1; iY <= nX; ++iY)
            for (int iX = 1; iX <= nX; ++iX) {
                const int e = iX + (iY - 1) * nX - 1, n0 = iX + (iY - 1) * (nX + 1);
                int *q = &m->elemNodes[4 * (size_t)e];
                q[0] = n0; q[1] = n0 + 1; q[2] = n0 + 1 + nX + 1; q[3] = n0 + nX + 1;
            }
    }
    std::lock_guard<std::mutex> g(s->mu);
    s->meshes[key] = m;
    return m;
}

// two independent 64-bit hashes of an array's bytes (word-wise)
void hash_words(const double *a, size_t n, uint64_t out[2])
{
    uint64_t h0 = 0xcbf29ce484222325ull, h1 = 0x9e3779b97f4a7c15ull;
    for (size_t k = 0; k < n; ++k) {
        uint64_t w;
        std::memcpy(&w, a + k, sizeof(w));
        h0 = (h0 ^ w) * 0x100000001b3ull;
        h1 = (h1 + w) * 0xff51afd7ed558ccdull;
        h1 ^= h1 >> 29;
    }
    out[0] = h0;
    out[1] = h1;
}

// Shallow-water requests that can advance as members of one ensemble: the
// same mesh, degree, precision, step count, constants and bathymetry (each
// member keeps its own state and time step)
bool sw_batchable(const Job &a, const Job &b)
{
    const self_request &x = a.req, &y = b.req;
    return x.kind == SELF_REQ_SW2D && y.kind == SELF_REQ_SW2D && x.nX == y.nX && x.N == y.N &&
           x.precision == y.precision && x.steps == y.steps && x.L == y.L && x.g == y.g && x.f0 == y.f0 &&
           a.inHash[0] == b.inHash[0] && a.inHash[1] == b.inHash[1];
}

constexpr int kMaxMembers = 8;  // ensemble size at which the fused stage stops gaining (C2: 6.4 us/member at 8)
constexpr int kMaxModels = 8;   // groups per persistent launch (the launch's argument block holds 8)
// models of at most this many nodes per variable join a persistent launch:
// there the launch-bound stages of small models gain (2.1x for eight 16^2
// models); larger models run faster in their own launches, side by side
// on the workers' streams
constexpr long kMaxGroupedNodes = 16384;

bool sw_groupable(const self_request &r)
{
    return r.kind == SELF_REQ_SW2D && (long)r.nX * r.nX * (r.N + 1) * (r.N + 1) <= kMaxGroupedNodes;
}

// device bytes a request needs (state, geometry, work vectors), with margin
size_t estimate_bytes(const self_request &r)
{
    const size_t n1 = r.N + 1, nE = (size_t)r.nX * r.nX, nodes = nE * n1 * n1;
    if (r.kind == SELF_REQ_SW2D)  // q, G, tend, bathy(3), plv, relax(4), metrics(5), traces, normals
        return (size_t)(1.25 * (double)(nodes * 8 * 24 + nE * 4 * n1 * 8 * 18)) + (1u << 20);
    // CGSEM: local and global vectors, geometry, Krylov work vectors
    return (size_t)(1.25 * (double)(nodes * 8 * 22)) + (1u << 20);
}

// the model of one group of batchable shallow-water requests (one request:
// a plain model; several: an ensemble, one member each)
int sw_group_model(self_server_s *s, std::vector<std::shared_ptr<Job>> &jobs, self_sw2d_t *out)
{
    const self_request &r = jobs[0]->req;
    const int M = (int)jobs.size();
    auto m = mesh_for(s, SELF_REQ_SW2D, r.nX, r.N, r.L);
    SELF_REQUIRE(m, SELF_ERR_OTHER, "request %lld: mesh set-up failed", jobs[0]->id);
    self_sw2d_desc d{r.N, m->nE, m->nEdges, r.precision, SELF_FUSED, r.g, r.dt, 0.0};
    self_sw2d_t h = nullptr;
    int rc = self_sw2d_create(&d, &h);
    if (rc) return rc;
    // the workers run their models side by side and allocate while others
    // run: stages are launched from the host, not captured as graphs
    rc = self_sw2d_set_graph(h, 0);
    const size_t nv = (size_t)m->nE * (r.N + 1) * (r.N + 1);
    std::vector<double> plv(nv, r.f0);
    if (!rc) rc = self_sw2d_set_nodal(h, m->dS.data(), m->dP.data(), m->w.data(), m->w.data(), m->lW.data(),
                                      m->lE.data(), m->lS.data(), m->lN.data());
    if (!rc) rc = self_sw2d_set_geometry(h, m->edges.data(), m->dxds.data(), m->dxdp.data(), m->dyds.data(),
                                         m->dydp.data(), m->J.data(), m->nHat.data(), m->nLen.data());
    if (!rc) rc = self_sw2d_set_field(h, SELF_SW_BATHYMETRY, r.in1);
    if (!rc) rc = self_sw2d_set_field(h, SELF_SW_PLANETARY_VORTICITY, plv.data());
    if (!rc) rc = self_sw2d_calculate_bathymetry_gradient(h);
    if (M == 1) {
        if (!rc) rc = self_sw2d_set_field(h, SELF_SW_SOLUTION, r.in0);
    } else {
        if (!rc) rc = self_sw2d_set_members(h, M);
        for (int k = 0; k < M && !rc; ++k) {
            rc = self_sw2d_set_member_solution(h, k, jobs[k]->req.in0);
            if (!rc) rc = self_sw2d_set_member_time_step(h, k, jobs[k]->req.dt);
        }
    }
    if (rc) {
        self_sw2d_destroy(h);
        return rc;
    }
    *out = h;
    return SELF_OK;
}

int sw_group_output(self_sw2d_t h, std::vector<std::shared_ptr<Job>> &jobs)
{
    int rc = SELF_OK;
    if (jobs.size() == 1) return self_sw2d_get_field(h, SELF_SW_SOLUTION, jobs[0]->req.out);
    for (size_t k = 0; k < jobs.size() && !rc; ++k) rc = self_sw2d_get_member_solution(h, (int)k, jobs[k]->req.out);
    return rc;
}

// groups of shallow-water requests (each an ensemble of batchable requests)
// advanced together: one group steps alone, several in one persistent
// launch per RK stage
int run_sw_groups(self_server_s *s, std::vector<std::vector<std::shared_ptr<Job>>> &groups)
{
    std::vector<self_sw2d_t> hs;
    int rc = SELF_OK;
    for (auto &g : groups) {
        self_sw2d_t h = nullptr;
        rc = sw_group_model(s, g, &h);
        if (rc) break;
        hs.push_back(h);
    }
    const int steps = groups[0][0]->req.steps;
    if (!rc) rc = hs.size() == 1 ? self_sw2d_forward_step_rk3(hs[0], 0.0, steps)
                                 : self_sw2d_forward_step_rk3_multi(hs.data(), (int)hs.size(), steps, nullptr);
    for (size_t k = 0; k < hs.size() && !rc; ++k) rc = sw_group_output(hs[k], groups[k]);
    for (auto h : hs) self_sw2d_destroy(h);
    return rc;
}

int run_cgsem(self_server_s *s, Job &j)
{
    const self_request &r = j.req;
    auto m = mesh_for(s, SELF_REQ_CGSEM2D, r.nX, r.N, 1.0);
    SELF_REQUIRE(m, SELF_ERR_OTHER, "request %lld: mesh set-up failed", j.id);
    const int nNodes = (r.nX + 1) * (r.nX + 1);
    self_cgsem2d_t h = nullptr;
    int rc = self_cgsem2d_create(r.N, m->nE, m->nEdges, nNodes, m->edges.data(), m->elemNodes.data(),
                                 m->dxds.data(), m->dxdp.data(), m->dyds.data(), m->dydp.data(), m->J.data(), &h);
    if (rc) return rc;
    const size_t nv = (size_t)m->nE * (r.N + 1) * (r.N + 1);
    std::vector<double> zero(nv, 0.0);
    if (!rc) rc = self_cgsem2d_set_flux_coefficient(h, r.in0);
    if (!rc) rc = self_cgsem2d_set_source(h, r.in1);
    if (!rc) rc = self_cgsem2d_set_dirichlet(h, r.in2 ? r.in2 : zero.data());
    if (!rc) rc = self_cgsem2d_set_solution(h, zero.data());
    int ioerr = 0, iters = 0;
    if (!rc) {
        if (r.precision == SELF_MIXED) {  // FP32 PCG with FP64 reliable updates
            int updates = 0;
            double rel = 0.0;
            rc = self_cgsem2d_solve_reliable(h, r.precond, r.tol, 0.2, r.maxIters, 400, &ioerr, &updates, &iters,
                                             &rel, nullptr);
        } else {
            rc = self_cgsem2d_solve(h, SELF_KRYLOV_CGSEM_PCG, r.precond, r.tol, r.maxIters, 30, 0, &ioerr, &iters,
                                    nullptr, 0, nullptr);
        }
    }
    if (!rc) rc = self_cgsem2d_get_solution(h, r.out);
    j.res.iterations = iters;
    j.res.ioerr = ioerr;
    self_cgsem2d_destroy(h);
    return rc;
}

void worker_main(self_server_s *s)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        std::vector<std::shared_ptr<Job>> batch;
        std::vector<std::vector<std::shared_ptr<Job>>> groups;
        {
            std::unique_lock<std::mutex> g(s->mu);
            for (;;) {
                if (s->stop && s->queue.empty()) return;
                // oldest queued request that fits in the free memory; a request
                // larger than the whole budget runs alone
                for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
                    const size_t b = (*it)->bytes;
                    if (s->reserved + b <= s->budget || (s->running == 0 && s->reserved == 0)) {
                        batch.push_back(*it);
                        s->queue.erase(it);
                        break;
                    }
                }
                if (!batch.empty()) break;
                s->cvWork.wait(g);
            }
            size_t bytes = batch[0]->bytes;
            // continuous batching: queued shallow-water requests that can run
            // as members of the same ensemble join it while memory allows
            auto pull_members = [&](std::vector<std::shared_ptr<Job>> &grp) {
                for (auto it = s->queue.begin(); it != s->queue.end() && (int)grp.size() < kMaxMembers;) {
                    if (sw_batchable(*grp[0], **it) && s->reserved + bytes + (*it)->bytes <= s->budget) {
                        bytes += (*it)->bytes;
                        grp.push_back(*it);
                        it = s->queue.erase(it);
                    } else {
                        ++it;
                    }
                }
            };
            pull_members(batch);
            groups.push_back(std::move(batch));
            // more groups of the same precision and step count (other meshes
            // or degrees) join the same persistent launches
            const self_request &r0 = groups[0][0]->req;
            if (sw_groupable(r0)) {
                bool took = true;
                while (took && (int)groups.size() < kMaxModels) {
                    took = false;
                    for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
                        const self_request &c = (*it)->req;
                        if (sw_groupable(c) && c.precision == r0.precision && c.steps == r0.steps &&
                            s->reserved + bytes + (*it)->bytes <= s->budget) {
                            std::vector<std::shared_ptr<Job>> grp{*it};
                            bytes += (*it)->bytes;
                            s->queue.erase(it);
                            pull_members(grp);
                            groups.push_back(std::move(grp));
                            took = true;
                            break;
                        }
                    }
                }
            }
            s->reserved += bytes;
            for (auto &g : groups) s->running += (int)g.size();
        }
        const Clock::time_point start = Clock::now();
        int rc;
        if (groups[0][0]->req.kind == SELF_REQ_SW2D) rc = run_sw_groups(s, groups);
        else rc = run_cgsem(s, *groups[0][0]);
        const Clock::time_point end = Clock::now();
        {
            std::lock_guard<std::mutex> g(s->mu);
            for (auto &grp : groups)
                for (auto &j : grp) {
                    j->start = start;
                    j->end = end;
                    j->res.status = rc;
                    j->res.queue_ms = ms_between(j->arrive, j->start);
                    j->res.run_ms = ms_between(j->start, j->end);
                    j->res.latency_ms = ms_between(j->arrive, j->end);
                    j->res.bytes = j->bytes;
                    j->res.members = (int)grp.size();
                    j->res.models = (int)groups.size();
                    j->done = true;
                    s->reserved -= j->bytes;
                    s->running--;
                }
        }
        s->cvDone.notify_all();
        s->cvWork.notify_all();
    }
}

}  // namespace

extern "C" {

int self_serve_create(int workers, double memFraction, self_server_t *out)
try {
    SELF_REQUIRE(out && workers >= 1 && workers <= 64 && memFraction > 0.0 && memFraction <= 1.0, SELF_ERR_ARG,
                 "self_serve_create: workers 1..64, memFraction in (0, 1]");
    auto *s = new self_server_s();
    SELF_HIP(hipGetDevice(&s->device));
    size_t fr = 0, tot = 0;
    SELF_HIP(hipMemGetInfo(&fr, &tot));
    s->memFraction = memFraction;
    s->budget = (size_t)(memFraction * (double)fr);
    for (int k = 0; k < workers; ++k) s->workers.emplace_back(worker_main, s);
    *out = s;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_serve_submit(self_server_t s, const self_request *req, long long *id)
{
    SELF_REQUIRE(s && req && id && req->in0 && req->out, SELF_ERR_ARG, "self_serve_submit: null argument");
    SELF_REQUIRE(req->kind == SELF_REQ_SW2D || req->kind == SELF_REQ_CGSEM2D, SELF_ERR_ARG, "unknown request kind %d",
                 req->kind);
    SELF_REQUIRE(req->nX >= 1 && req->N >= 1 && req->N <= 11, SELF_ERR_ARG, "request: nX >= 1, N in 1..11");
    SELF_REQUIRE(req->kind != SELF_REQ_SW2D || (req->in1 && req->nX >= 3 && req->steps >= 0), SELF_ERR_ARG,
                 "shallow-water request needs bathymetry, nX >= 3 (periodic) and steps >= 0");
    SELF_REQUIRE(req->kind != SELF_REQ_CGSEM2D || (req->in1 && req->maxIters >= 1 && req->tol > 0), SELF_ERR_ARG,
                 "CGSEM request needs a source, maxIters >= 1 and tol > 0");
    auto j = std::make_shared<Job>();
    j->req = *req;
    j->bytes = estimate_bytes(*req);
    if (req->kind == SELF_REQ_SW2D)
        hash_words(req->in1, 3 * (size_t)req->nX * req->nX * (req->N + 1) * (req->N + 1), j->inHash);
    j->arrive = Clock::now();
    {
        std::lock_guard<std::mutex> g(s->mu);
        SELF_REQUIRE(!s->stop, SELF_ERR_STATE, "server is shutting down");
        j->id = s->nextId++;
        s->jobs[j->id] = j;
        s->queue.push_back(j);
        *id = j->id;
    }
    s->cvWork.notify_all();
    return SELF_OK;
}

int self_serve_wait(self_server_t s, long long id, self_request_result *res)
{
    SELF_REQUIRE(s && res, SELF_ERR_ARG, "self_serve_wait: null argument");
    std::shared_ptr<Job> j;
    {
        std::unique_lock<std::mutex> g(s->mu);
        auto it = s->jobs.find(id);
        SELF_REQUIRE(it != s->jobs.end(), SELF_ERR_ARG, "unknown request id %lld", id);
        j = it->second;
        s->cvDone.wait(g, [&] { return j->done; });
        *res = j->res;
        s->jobs.erase(id);
    }
    return SELF_OK;
}

int self_serve_stats(self_server_t s, int *queued, int *running, size_t *reservedBytes, size_t *budgetBytes)
{
    SELF_REQUIRE(s, SELF_ERR_ARG, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    if (queued) *queued = (int)s->queue.size();
    if (running) *running = s->running;
    if (reservedBytes) *reservedBytes = s->reserved;
    if (budgetBytes) *budgetBytes = s->budget;
    return SELF_OK;
}

int self_serve_destroy(self_server_t s)
{
    if (!s) return SELF_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cvWork.notify_all();
    for (auto &t : s->workers) t.join();
    delete s;
    return SELF_OK;
}

}  // extern "C"

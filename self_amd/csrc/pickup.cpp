// pickup.cpp -- the reference's restart ("pickup") files.
//
// WritePickup/ReadPickup of the shallow-water and advection-diffusion
// models (ConservativeShallowWaterClass.f90:2013-2125,
// AdvectionClass.f90:2040-2171) open '<Model>.<iter, 10 digits>.pickup' as
// an unformatted direct-access file, CONVERT='big_endian', RECL = 8 (N+1)^2
// bytes, and store one record per element field, elements in order:
//   shallow water: U, V, H, bathymetry h, planetary vorticity f   (5 records)
//   advection:     c(1..nEq), relaxation field(1..nEq),
//                  relaxation factor(1..nEq), u, v              (3 nEq + 2)
// each record one (0:N,0:N) array in column-major order.  Direct-access
// files carry no record markers, so the file is exactly the records.
// ReadPickup of the shallow-water model reads only h of the bathymetry
// (its gradient comes back zero, :2088; Build recomputes it, :224-228).
//
// Host arrays here use the C ABI's layouts: element slowest, per element
// the reference's (0:N,0:N,1:nf) array.

#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "self_amd.h"

namespace {

using namespace self;

inline uint64_t bswap64(uint64_t v)
{
    return __builtin_bswap64(v);
}

// big-endian <-> host doubles
void to_be(const double *src, unsigned char *dst, size_t n)
{
    for (size_t k = 0; k < n; ++k) {
        uint64_t w;
        std::memcpy(&w, src + k, 8);
        w = bswap64(w);
        std::memcpy(dst + 8 * k, &w, 8);
    }
}

void from_be(const unsigned char *src, double *dst, size_t n)
{
    for (size_t k = 0; k < n; ++k) {
        uint64_t w;
        std::memcpy(&w, src + 8 * k, 8);
        w = bswap64(w);
        std::memcpy(dst + k, &w, 8);
    }
}

// one element's records: (pointer to its first (0:N,0:N) array, count)
struct Piece {
    const double *src;  // write side
    double *dst;        // read side
    int count;
};

struct File {
    std::FILE *f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

int write_file(const char *path, size_t np, int nElems, int recPerElem,
               const std::vector<std::vector<Piece>> &perElem)
{
    File fh;
    fh.f = std::fopen(path, "wb");
    SELF_REQUIRE(fh.f, SELF_ERR_ARG, "pickup: cannot create %s (%s)", path, std::strerror(errno));
    std::vector<unsigned char> buf((size_t)recPerElem * np * 8);
    for (int e = 0; e < nElems; ++e) {
        size_t off = 0;
        for (const Piece &p : perElem[e]) {
            to_be(p.src, buf.data() + off, (size_t)p.count * np);
            off += (size_t)p.count * np * 8;
        }
        SELF_REQUIRE(std::fwrite(buf.data(), 1, off, fh.f) == off, SELF_ERR_OTHER, "pickup: short write to %s",
                     path);
    }
    SELF_REQUIRE(std::fclose(fh.f) == 0, SELF_ERR_OTHER, "pickup: closing %s failed", path);
    fh.f = nullptr;
    return SELF_OK;
}

int read_file(const char *path, size_t np, int nElems, int recPerElem, const std::vector<std::vector<Piece>> &perElem)
{
    File fh;
    fh.f = std::fopen(path, "rb");
    SELF_REQUIRE(fh.f, SELF_ERR_ARG, "pickup: cannot open %s (%s)", path, std::strerror(errno));
    SELF_REQUIRE(std::fseek(fh.f, 0, SEEK_END) == 0, SELF_ERR_OTHER, "pickup: cannot seek %s", path);
    const long size = std::ftell(fh.f);
    const size_t want = (size_t)nElems * recPerElem * np * 8;
    SELF_REQUIRE(size >= 0 && (size_t)size == want, SELF_ERR_ARG,
                 "pickup: %s holds %ld bytes, %d elements x %d records of %zu float64 need %zu", path, size, nElems,
                 recPerElem, np, want);
    std::rewind(fh.f);
    std::vector<unsigned char> buf((size_t)recPerElem * np * 8);
    for (int e = 0; e < nElems; ++e) {
        SELF_REQUIRE(std::fread(buf.data(), 1, buf.size(), fh.f) == buf.size(), SELF_ERR_OTHER,
                     "pickup: short read from %s", path);
        size_t off = 0;
        for (const Piece &p : perElem[e]) {
            from_be(buf.data() + off, p.dst, (size_t)p.count * np);
            off += (size_t)p.count * np * 8;
        }
    }
    return SELF_OK;
}

int check_sizes(int N, int nElems, int nTracers)
{
    SELF_REQUIRE(N >= 1 && N <= 64 && nElems >= 1 && nTracers >= 1, SELF_ERR_ARG,
                 "pickup: N=%d nElems=%d nTracers=%d", N, nElems, nTracers);
    return SELF_OK;
}

}  // namespace

extern "C" {

int self_pickup_filename(const char *model, int iter, char *out, int outLen)
try {
    SELF_REQUIRE(model && out && outLen > 0, SELF_ERR_ARG, "self_pickup_filename: null argument");
    SELF_REQUIRE(iter >= 0, SELF_ERR_ARG, "pickup iterate %d: WRITE(iterChar,'(I10.10)') needs iter >= 0", iter);
    const int n = std::snprintf(out, (size_t)outLen, "%s.%010d.pickup", model, iter);
    SELF_REQUIRE(n > 0 && n < outLen, SELF_ERR_ARG, "self_pickup_filename: buffer of %d bytes too short", outLen);
    return SELF_OK;
}
SELF_CATCH_ALL

int self_sw2d_pickup_write(const char *path, int N, int nElems, const double *sol, const double *bathy,
                           const double *plv)
try {
    SELF_REQUIRE(path && sol && bathy && plv, SELF_ERR_ARG, "self_sw2d_pickup_write: null argument");
    if (int rc = check_sizes(N, nElems, 1)) return rc;
    const size_t np = (size_t)(N + 1) * (N + 1);
    std::vector<std::vector<Piece>> pe(nElems);
    for (int e = 0; e < nElems; ++e)
        pe[e] = {{sol + (size_t)e * 3 * np, nullptr, 3}, {bathy + (size_t)e * 3 * np, nullptr, 1},
                 {plv + (size_t)e * np, nullptr, 1}};
    return write_file(path, np, nElems, 5, pe);
}
SELF_CATCH_ALL

int self_sw2d_pickup_read(const char *path, int N, int nElems, double *sol, double *bathy, double *plv)
try {
    SELF_REQUIRE(path && sol && bathy && plv, SELF_ERR_ARG, "self_sw2d_pickup_read: null argument");
    if (int rc = check_sizes(N, nElems, 1)) return rc;
    const size_t np = (size_t)(N + 1) * (N + 1);
    std::memset(bathy, 0, sizeof(double) * 3 * np * nElems);  // bathy = ZERO (:2088)
    std::vector<std::vector<Piece>> pe(nElems);
    for (int e = 0; e < nElems; ++e)
        pe[e] = {{nullptr, sol + (size_t)e * 3 * np, 3}, {nullptr, bathy + (size_t)e * 3 * np, 1},
                 {nullptr, plv + (size_t)e * np, 1}};
    return read_file(path, np, nElems, 5, pe);
}
SELF_CATCH_ALL

int self_adv2d_pickup_write(const char *path, int N, int nElems, int nTracers, const double *c, const double *relax,
                            const double *rFac, const double *vel)
try {
    SELF_REQUIRE(path && c && relax && rFac && vel, SELF_ERR_ARG, "self_adv2d_pickup_write: null argument");
    if (int rc = check_sizes(N, nElems, nTracers)) return rc;
    const size_t np = (size_t)(N + 1) * (N + 1), q = (size_t)nTracers;
    std::vector<std::vector<Piece>> pe(nElems);
    for (int e = 0; e < nElems; ++e)
        pe[e] = {{c + e * q * np, nullptr, nTracers}, {relax + e * q * np, nullptr, nTracers},
                 {rFac + e * q * np, nullptr, nTracers}, {vel + (size_t)e * 2 * np, nullptr, 2}};
    return write_file(path, np, nElems, 3 * nTracers + 2, pe);
}
SELF_CATCH_ALL

int self_adv2d_pickup_read(const char *path, int N, int nElems, int nTracers, double *c, double *relax, double *rFac,
                           double *vel)
try {
    SELF_REQUIRE(path && c && relax && rFac && vel, SELF_ERR_ARG, "self_adv2d_pickup_read: null argument");
    if (int rc = check_sizes(N, nElems, nTracers)) return rc;
    const size_t np = (size_t)(N + 1) * (N + 1), q = (size_t)nTracers;
    std::vector<std::vector<Piece>> pe(nElems);
    for (int e = 0; e < nElems; ++e)
        pe[e] = {{nullptr, c + e * q * np, nTracers}, {nullptr, relax + e * q * np, nTracers},
                 {nullptr, rFac + e * q * np, nTracers}, {nullptr, vel + (size_t)e * 2 * np, 2}};
    return read_file(path, np, nElems, 3 * nTracers + 2, pe);
}
SELF_CATCH_ALL

}  // extern "C"

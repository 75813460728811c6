// tuning.cpp -- the process-wide table of launch configurations chosen by
// timing (autotuning).  An entry is keyed by the model kind, the
// dimension, the polynomial degree, the arithmetic (0 FP64, 1 mixed) and a
// size class (floor(log2(elements))); it holds up to four integers of
// configuration and the time per step that chose them.  Models consult the
// table before timing candidates, so one tuning serves every model of the
// same key in the process; the table can be saved to and loaded from a
// text file (one entry per line: kind dim N dtype sizeClass c0 c1 c2 c3 ms).

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "common.hpp"

namespace self {

namespace {

using Key = std::tuple<std::string, int, int, int, int>;
struct Entry {
    int cfg[4] = {0, 0, 0, 0};
    double ms = 0.0;
};

std::mutex &table_mutex()
{
    static std::mutex m;
    return m;
}

std::map<Key, Entry> &table()
{
    static std::map<Key, Entry> t;
    return t;
}

}  // namespace

int tune_size_class(size_t nElems) { return nElems ? (int)std::floor(std::log2((double)nElems)) : 0; }

bool tune_lookup(const char *kind, int dim, int N, int dtype, int sizeClass, int cfg[4], double *ms)
{
    std::lock_guard<std::mutex> g(table_mutex());
    auto it = table().find(Key(kind, dim, N, dtype, sizeClass));
    if (it == table().end()) return false;
    std::memcpy(cfg, it->second.cfg, sizeof(int) * 4);
    if (ms) *ms = it->second.ms;
    return true;
}

void tune_record(const char *kind, int dim, int N, int dtype, int sizeClass, const int cfg[4], double ms)
{
    std::lock_guard<std::mutex> g(table_mutex());
    Entry e;
    std::memcpy(e.cfg, cfg, sizeof(int) * 4);
    e.ms = ms;
    table()[Key(kind, dim, N, dtype, sizeClass)] = e;
}

}  // namespace self

using namespace self;

extern "C" {

int self_tune_count(int *n)
{
    SELF_REQUIRE(n, SELF_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(table_mutex());
    *n = (int)table().size();
    return SELF_OK;
}

int self_tune_entry(int index, char *kind, int kindLen, int *key4, int *cfg4, double *ms)
{
    SELF_REQUIRE(kind && kindLen > 0 && key4 && cfg4 && ms, SELF_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(table_mutex());
    SELF_REQUIRE(index >= 0 && index < (int)table().size(), SELF_ERR_ARG, "entry %d of %zu", index, table().size());
    auto it = table().begin();
    std::advance(it, index);
    std::snprintf(kind, (size_t)kindLen, "%s", std::get<0>(it->first).c_str());
    key4[0] = std::get<1>(it->first);
    key4[1] = std::get<2>(it->first);
    key4[2] = std::get<3>(it->first);
    key4[3] = std::get<4>(it->first);
    std::memcpy(cfg4, it->second.cfg, sizeof(int) * 4);
    *ms = it->second.ms;
    return SELF_OK;
}

int self_tune_clear(void)
{
    std::lock_guard<std::mutex> g(table_mutex());
    table().clear();
    return SELF_OK;
}

int self_tune_save(const char *path)
{
    SELF_REQUIRE(path, SELF_ERR_ARG, "null path");
    std::lock_guard<std::mutex> g(table_mutex());
    FILE *f = std::fopen(path, "w");
    SELF_REQUIRE(f, SELF_ERR_OTHER, "cannot write %s", path);
    for (const auto &kv : table())
        std::fprintf(f, "%s %d %d %d %d %d %d %d %d %.17g\n", std::get<0>(kv.first).c_str(), std::get<1>(kv.first),
                     std::get<2>(kv.first), std::get<3>(kv.first), std::get<4>(kv.first), kv.second.cfg[0],
                     kv.second.cfg[1], kv.second.cfg[2], kv.second.cfg[3], kv.second.ms);
    const bool ok = std::fclose(f) == 0;
    SELF_REQUIRE(ok, SELF_ERR_OTHER, "error writing %s", path);
    return SELF_OK;
}

int self_tune_load(const char *path)
{
    SELF_REQUIRE(path, SELF_ERR_ARG, "null path");
    FILE *f = std::fopen(path, "r");
    SELF_REQUIRE(f, SELF_ERR_OTHER, "cannot read %s", path);
    char kind[64];
    int k[4], c[4];
    double ms;
    int n = 0, bad = 0;
    while (true) {
        const int r = std::fscanf(f, "%63s %d %d %d %d %d %d %d %d %lg", kind, &k[0], &k[1], &k[2], &k[3], &c[0],
                                  &c[1], &c[2], &c[3], &ms);
        if (r == EOF) break;
        if (r != 10) { bad = 1; break; }
        tune_record(kind, k[0], k[1], k[2], k[3], c, ms);
        ++n;
    }
    std::fclose(f);
    SELF_REQUIRE(!bad, SELF_ERR_ARG, "%s: malformed entry after %d lines", path, n);
    return SELF_OK;
}

}  // extern "C"

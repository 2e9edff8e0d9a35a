// Event-pair timing of kernel classes (see prof.hpp) and its C-ABI readout.
#include "prof.hpp"
#include "common.hpp"
#include "../../include/scirs_gpu.h"

#include <mutex>
#include <vector>

namespace {
struct Rec {
    int cat;
    hipEvent_t e0, e1;
    double flops, bytes;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
} // namespace

namespace scirs {
bool prof_enabled() { return g_on; }

ProfScope::ProfScope(int cat_, double flops_, double bytes_, hipStream_t s_)
    : cat(cat_), flops(flops_), bytes(bytes_), s(s_)
{
    if (!g_on) return;
    std::lock_guard<std::mutex> lk(g_mu);
    e0 = take_event();
    if (e0) (void)hipEventRecord(e0, s);
}

ProfScope::~ProfScope()
{
    if (!g_on || !e0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t e1 = take_event();
    if (!e1) return;
    (void)hipEventRecord(e1, s);
    g_recs.push_back(Rec{cat, e0, e1, flops, bytes});
}
} // namespace scirs

extern "C" int scirs_gpu_prof_begin(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &r : g_recs) { g_pool.push_back(r.e0); g_pool.push_back(r.e1); }
    g_recs.clear();
    g_on = true;
    return SCIRS_OK;
}

extern "C" int scirs_gpu_prof_end(scirs_prof_stats *out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = false;
    if (!out) return SCIRS_ERR_ARG;
    for (int c = 0; c < SCIRS_PROF_NCAT; ++c) {
        out[c].launches = 0;
        out[c].total_ms = 0;
        out[c].flops = 0;
        out[c].bytes = 0;
    }
    for (auto &r : g_recs) {
        if (hipEventSynchronize(r.e1) != hipSuccess) continue;
        float ms = 0;
        if (hipEventElapsedTime(&ms, r.e0, r.e1) != hipSuccess) continue;
        scirs_prof_stats &o = out[r.cat];
        o.launches += 1;
        o.total_ms += ms;
        o.flops += r.flops;
        o.bytes += r.bytes;
    }
    for (auto &r : g_recs) { g_pool.push_back(r.e0); g_pool.push_back(r.e1); }
    g_recs.clear();
    return SCIRS_OK;
}

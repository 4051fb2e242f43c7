// gt_profile.cpp -- optional HIP-event timing of individual kernel launches.
//
// When enabled (gt_profile_enable(1)) every launch site brackets its kernel
// with a pair of events recorded on the launch stream; gt_profile_read()
// waits for the recorded pairs of one kernel id and returns their summed
// duration.  Disabled (the default) the hooks cost one relaxed atomic load.
#include <atomic>
#include <mutex>
#include <vector>

#include "gt_common.h"

namespace gt {

namespace {
struct Rec {
    int id;
    hipEvent_t a, b;
};
std::atomic<int> g_on{0};
std::atomic<unsigned> g_mask{~0u};  // kernel ids bracketed while enabled

bool selected(int id)
{
    return g_on.load(std::memory_order_relaxed) && id >= 0 && id < 32 &&
           ((g_mask.load(std::memory_order_relaxed) >> id) & 1u);
}
std::mutex g_mu;
std::vector<Rec> g_done;                 // ended, not read
std::vector<hipEvent_t> g_pool;
thread_local std::vector<Rec> t_stack;   // per-thread begin stack

hipEvent_t get_event()
{
    hipEvent_t e = nullptr;
    {
        std::lock_guard<std::mutex> l(g_mu);
        if (!g_pool.empty()) {
            e = g_pool.back();
            g_pool.pop_back();
        }
    }
    if (!e && hipEventCreate(&e) != hipSuccess) e = nullptr;
    return e;
}
}  // namespace

void prof_begin(hipStream_t st, int id)
{
    if (!selected(id)) return;
    hipEvent_t a = get_event(), b = get_event();
    if (!a || !b) return;
    hipEventRecord(a, st);
    t_stack.push_back({id, a, b});
}

void prof_end(hipStream_t st, int id)
{
    if (!selected(id) || t_stack.empty()) return;
    Rec r = t_stack.back();
    t_stack.pop_back();
    if (r.id != id) return;
    hipEventRecord(r.b, st);
    std::lock_guard<std::mutex> l(g_mu);
    g_done.push_back(r);
}

bool prof_ext_begin(int id, hipEvent_t *a, hipEvent_t *b)
{
    *a = *b = nullptr;
    if (!selected(id)) return false;
    hipEvent_t ea = get_event(), eb = get_event();
    if (!ea || !eb) return false;
    *a = ea;
    *b = eb;
    return true;
}

void prof_ext_end(int id, hipEvent_t a, hipEvent_t b)
{
    std::lock_guard<std::mutex> l(g_mu);
    g_done.push_back({id, a, b});
}

}  // namespace gt

using namespace gt;

extern "C" int gt_profile_enable(int on)
{
    g_on.store(on ? 1 : 0);
    return 0;
}

extern "C" int gt_profile_select(unsigned mask)
{
    g_mask.store(mask);
    return 0;
}

extern "C" int gt_profile_read(int kernel_id, double *total_ms, long *count)
{
    GT_CHECK_ARG(total_ms && count, "gt_profile_read: NULL output");
    std::vector<Rec> mine;
    {
        std::lock_guard<std::mutex> l(g_mu);
        std::vector<Rec> keep;
        for (auto &r : g_done) (r.id == kernel_id ? mine : keep).push_back(r);
        g_done.swap(keep);
    }
    double tot = 0.0;
    long n = 0;
    for (auto &r : mine) {
        GT_HIP(hipEventSynchronize(r.b));
        float ms = 0.0f;
        GT_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        tot += ms;
        n++;
    }
    {
        std::lock_guard<std::mutex> l(g_mu);
        for (auto &r : mine) {
            g_pool.push_back(r.a);
            g_pool.push_back(r.b);
        }
    }
    *total_ms = tot;
    *count = n;
    return 0;
}

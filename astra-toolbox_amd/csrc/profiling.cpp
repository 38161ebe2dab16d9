// Optional per-kernel timing: when enabled, the projector launchers record
// a hipEvent pair around every hot kernel launch on the launch's own stream,
// together with the algorithmic work units (ray-slice samples or
// voxel-angle updates) the launch performs.  Read back with
// astra_amd_profile_read().
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"

namespace amx {

namespace {
struct Rec {
    std::string name;
    hipEvent_t a, b;
    double units;
};
std::atomic<bool> g_on{false};
std::mutex g_mu;
std::vector<Rec> g_pending;
struct Acc {
    double ms = 0, units = 0;
    unsigned long long launches = 0;
};
std::map<std::string, Acc> g_acc;

void drain_locked() {
    for (Rec &r : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            Acc &x = g_acc[r.name];
            x.ms += ms;
            x.units += r.units;
            x.launches++;
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_pending.clear();
    (void)hipGetLastError();
}
}  // namespace

bool prof_on() { return g_on.load(std::memory_order_relaxed); }

ProfToken prof_begin(const char *name, double units, hipStream_t s) {
    ProfToken t{};
    if (!prof_on()) return t;
    Rec r{name, nullptr, nullptr, units};
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) {
        (void)hipGetLastError();
        return t;
    }
    (void)hipEventRecord(r.a, s);
    std::lock_guard<std::mutex> lk(g_mu);
    g_pending.push_back(r);
    t.index = (long)g_pending.size() - 1;
    t.active = true;
    return t;
}

void prof_end(const ProfToken &t, hipStream_t s) {
    if (!t.active) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (t.index >= 0 && t.index < (long)g_pending.size()) (void)hipEventRecord(g_pending[t.index].b, s);
}

void prof_enable(bool on) {
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    g_acc.clear();
    g_on.store(on);
}

bool prof_read(const char *name, double &ms, double &units, unsigned long long &launches) {
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    auto it = g_acc.find(name);
    if (it == g_acc.end()) {
        ms = units = 0;
        launches = 0;
        return false;
    }
    ms = it->second.ms;
    units = it->second.units;
    launches = it->second.launches;
    return true;
}

}  // namespace amx

// Stage timing: pairs of hipEvents around each stage, resolved when read.
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pylians_hip.h"
#include "profile.hpp"

namespace pyl {

namespace {
struct Pending {
    std::string name;
    hipEvent_t a, b;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Pending> g_pending;
std::map<std::string, std::pair<long long, double>> g_acc;  // name -> (count, total ms)
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

bool profiling_enabled() { return g_on; }

StageTimer::StageTimer(const char* n, hipStream_t s) : name(n), st(s) {
    if (!g_on) return;
    // no events inside a stream capture: they would become graph nodes that are never resolved
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    std::lock_guard<std::mutex> lk(g_mu);
    e0 = take_event();
    if (e0 && hipEventRecord(e0, st) == hipSuccess) live = true;
}

void StageTimer::stop() {
    if (!live) return;
    live = false;
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t e1 = take_event();
    if (!e1 || hipEventRecord(e1, st) != hipSuccess) return;
    g_pending.push_back({name, e0, e1});
}

static void resolve_locked() {
    for (auto& p : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto& a = g_acc[p.name];
            a.first += 1;
            a.second += ms;
        }
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
}

}  // namespace pyl

using namespace pyl;

extern "C" int pyl_profile(int enable) {
    std::lock_guard<std::mutex> lk(g_mu);
    resolve_locked();
    g_acc.clear();
    g_on = enable != 0;
    return PYL_OK;
}

extern "C" int pyl_profile_read(char* buf, int len) {
    std::lock_guard<std::mutex> lk(g_mu);
    resolve_locked();
    std::string s;
    for (auto& kv : g_acc) {
        char line[256];
        snprintf(line, sizeof(line), "%s %lld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
        s += line;
    }
    if (buf && len > 0) {
        snprintf(buf, (size_t)len, "%s", s.c_str());
    }
    return (int)s.size();
}

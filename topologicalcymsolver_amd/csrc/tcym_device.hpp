// Stream, allocations, timed launches and error state shared by the
// single-lattice context (tcym_api.hip) and the batch (tcym_batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "tcym.h"
#include "tcym_backend.hpp"

namespace tcym {

struct DeviceQueue {
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    int device = 0;
    hipStream_t stream = nullptr;
    long long bytes = 0;
    std::string err;
    bool timing = false;
    bool capturing = false;      // launches are being recorded into a HIP graph
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, std::pair<double, long long>> stats;
    hipError_t last = hipSuccess;
    std::vector<void*> owned;    // device buffers freed by release()

    hipEvent_t ev() {
        if (!free_events.empty()) { hipEvent_t e = free_events.back(); free_events.pop_back(); return e; }
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    // one kernel launch, bracketed by timing events unless a graph is being captured
    void launch(const char* name, const Launch& f) {
        if (last != hipSuccess) return;
        if (timing && !capturing) {
            hipEvent_t a = ev(), b = ev();
            hipEventRecord(a, stream);
            f(stream);
            hipEventRecord(b, stream);
            pending.push_back({name, a, b});
        } else {
            f(stream);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { last = e; err = std::string("kernel ") + name + ": " + hipGetErrorString(e); }
    }
    Launcher L() { return [this](const char* n, const Launch& f) { launch(n, f); }; }

    int fail(int code, const std::string& m) { err = m; return code; }
    int hip(hipError_t e, const char* what) {
        if (e == hipSuccess) return TCYM_OK;
        last = e;
        err = std::string(what) + ": " + hipGetErrorString(e);
        return TCYM_E_HIP;
    }
    int status() { return last == hipSuccess ? TCYM_OK : TCYM_E_HIP; }
    int sync() {
        if (last != hipSuccess) return TCYM_E_HIP;
        return hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    // zero-filled device buffer, allocated once (*p stays if already set)
    int alloc(void** p, size_t n, const char* what) {
        if (*p) return TCYM_OK;
        int r = hip(hipMalloc(p, n), what);
        if (r == TCYM_OK) {
            bytes += (long long)n;
            owned.push_back(*p);
            r = hip(hipMemsetAsync(*p, 0, n, stream), what);
        }
        return r;
    }
    // fold finished timing events into the per-kernel statistics
    int resolve() {
        int r = sync();
        if (r) return r;
        for (auto& q : pending) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, q.a, q.b);
            auto& s = stats[q.name];
            s.first += ms;
            s.second += 1;
            free_events.push_back(q.a);
            free_events.push_back(q.b);
        }
        pending.clear();
        return TCYM_OK;
    }
    void release() {
        if (stream) hipStreamSynchronize(stream);
        for (auto& q : pending) { hipEventDestroy(q.a); hipEventDestroy(q.b); }
        pending.clear();
        for (auto e : free_events) hipEventDestroy(e);
        free_events.clear();
        for (void* p : owned) hipFree(p);
        owned.clear();
        if (stream) hipStreamDestroy(stream);
        stream = nullptr;
    }
};

}  // namespace tcym

// Per-kernel timing with HIP events on the launch stream.
//
// When enabled, every kernel of a draw or resolve is launched with
// hipExtLaunchKernel and a start and a stop event that the runtime stamps at
// the dispatch's own start and end, so a kernel's time is its execution (what
// rocprofv3 --kernel-trace reports), not the gap between markers recorded
// around it; collect() waits for them and sums the elapsed time per kernel
// name. Launches that are not single kernels use a Scope of two events on the
// stream. This is the device counterpart of the reference's
// DrawTime/ResolveTime perf counters (Rasterizer.h:366-379).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

namespace glimp {

class Profiler {
public:
    bool enabled = false;

    struct Scope {
        Profiler* prof;
        size_t slot;
        hipStream_t stream;
        Scope(Profiler* p, const char* name, hipStream_t s) : prof(p), slot((size_t)-1), stream(s) {
            if (prof && prof->enabled) slot = prof->begin(name, s);
        }
        ~Scope() {
            if (slot != (size_t)-1) prof->end(slot, stream);
        }
    };

    struct Stat {
        unsigned long long launches = 0;
        double total_ms = 0.0;
    };

    ~Profiler() {
        for (auto& p : pending_) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : free_) (void)hipEventDestroy(e);
    }

    size_t begin(const char* name, hipStream_t s) {
        Pending p;
        p.name = name;
        p.a = acquire();
        p.b = acquire();
        (void)hipEventRecord(p.a, s);
        pending_.push_back(p);
        return pending_.size() - 1;
    }

    void end(size_t slot, hipStream_t s) { (void)hipEventRecord(pending_[slot].b, s); }

    // Events for one hipExtLaunchKernel dispatch of kernel `name`.
    void dispatch_events(const char* name, hipEvent_t& a, hipEvent_t& b) {
        Pending p;
        p.name = name;
        p.a = acquire();
        p.b = acquire();
        pending_.push_back(p);
        a = p.a;
        b = p.b;
    }

    // Waits for all recorded events and folds them into the statistics.
    hipError_t collect() {
        for (auto& p : pending_) {
            hipError_t e = hipEventSynchronize(p.b);
            if (e != hipSuccess) return e;
            float ms = 0.0f;
            e = hipEventElapsedTime(&ms, p.a, p.b);
            if (e != hipSuccess) return e;
            Stat& st = stats_[p.name];
            st.launches++;
            st.total_ms += ms;
            free_.push_back(p.a);
            free_.push_back(p.b);
        }
        pending_.clear();
        return hipSuccess;
    }

    const std::map<std::string, Stat>& stats() const { return stats_; }
    void reset() { stats_.clear(); }

private:
    struct Pending {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> free_;
    std::map<std::string, Stat> stats_;

    hipEvent_t acquire() {
        if (!free_.empty()) {
            hipEvent_t e = free_.back();
            free_.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
};

// Launches `kernel` on `s`; timed per dispatch when `prof` is enabled (never
// during stream capture: captured launches pass no profiler).
template <typename F, typename... Args>
inline void launch_kernel(Profiler* prof, const char* name, F kernel, dim3 grid, dim3 block, uint32_t shmem,
                          hipStream_t s, Args... args) {
    if (prof && prof->enabled) {
        hipEvent_t a, b;
        prof->dispatch_events(name, a, b);
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, a, b, 0u, args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
    }
}

}  // namespace glimp

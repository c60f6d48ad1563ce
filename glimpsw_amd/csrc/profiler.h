// Per-kernel timing with HIP events on the launch stream.
//
// When enabled, every kernel launch of a draw or resolve is bracketed by two
// events recorded on the stream it runs on; read() waits for them and sums
// the elapsed time per kernel name. This is the device counterpart of the
// reference's DrawTime/ResolveTime perf counters (Rasterizer.h:366-379).
#pragma once

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

}  // namespace glimp

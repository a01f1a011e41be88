// Per-kernel timing with HIP events on the launching stream.
//
// A KernelTimer installed on the calling host thread (KernelTimer::Install)
// makes every FLIP2_TIMED scope record an event before and after the launch
// it encloses; with no timer installed a scope costs one thread-local load.
// Events come from a pool that grows on demand and are read only by
// collect(), so a timed sequence of launches runs without host waits.  A
// filter restricts the recording to some names (e.g. the dominant kernel of
// a step, so that the other launches run without the events' gaps).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <set>
#include <string>
#include <vector>

namespace flip2 {

class KernelTimer {
public:
    KernelTimer() = default;
    ~KernelTimer();
    KernelTimer(const KernelTimer&) = delete;
    KernelTimer& operator=(const KernelTimer&) = delete;

    // Only the listed names are recorded (an empty list records all).
    void setFilter(const std::vector<std::string>& names);
    // Record one launch in every `every` of each name (1: all).  The events
    // around a launch cost the stream a few microseconds, so a timed run of
    // many short launches samples them.
    void setSampleEvery(uint32_t every) { every_ = every ? every : 1u; seen_.clear(); }
    // Whether the next launch named `name` is recorded (counts it).
    bool wants(const char* name);
    // Record the start / stop event of one launch named `name` on stream s.
    void start(const char* name, hipStream_t s);
    void stop(hipStream_t s);
    // Wait for the recorded events and add each interval to its name.
    void collect();
    void reset();

    struct Entry {
        double ms = 0.0;
        uint64_t launches = 0;
    };
    const std::map<std::string, Entry>& totals() const { return totals_; }

    // The calling thread's installed timer (null: timing off).
    static KernelTimer*& current();
    struct Install {
        KernelTimer* prev;
        explicit Install(KernelTimer* t) : prev(current()) { current() = t; }
        ~Install() { current() = prev; }
    };

private:
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    hipEvent_t take();
    std::vector<hipEvent_t> free_;
    std::vector<hipEvent_t> all_;
    std::vector<Pending> pending_;
    std::set<std::string> filter_;
    std::map<std::string, Entry> totals_;
    std::map<std::string, uint64_t> seen_;
    uint32_t every_ = 1;
    bool open_ = false;
};

// Kernel name of a launch as written at the call site, without enclosing
// parentheses and template arguments ("(k<A, B>)" -> "k"): the name rocprof
// reports before its argument list.
std::string kernel_base_name(const char* written);

// Records the enclosed launch(es) under `name` when a timer is installed.
class TimedScope {
public:
    TimedScope(const char* name, hipStream_t s) : s_(s) {
        KernelTimer* t = KernelTimer::current();
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        // launches captured into a graph are not timed (an event recorded
        // during capture becomes a graph node, not a timestamp)
        if (t && hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
            const std::string n = kernel_base_name(name);
            if (t->wants(n.c_str())) {
                t_ = t;
                t->start(n.c_str(), s);
            }
        }
    }
    ~TimedScope() {
        if (t_) t_->stop(s_);
    }
    TimedScope(const TimedScope&) = delete;
    TimedScope& operator=(const TimedScope&) = delete;

private:
    KernelTimer* t_ = nullptr;
    hipStream_t s_;
};

#define FLIP2_TIMED_CAT2(a, b) a##b
#define FLIP2_TIMED_CAT(a, b) FLIP2_TIMED_CAT2(a, b)
#define FLIP2_TIMED(name, stream) ::flip2::TimedScope FLIP2_TIMED_CAT(flip2_timed_, __LINE__)((name), (stream))

// Every kernel launch of the library: hipLaunchKernelGGL, recorded under the
// kernel's name when a KernelTimer is installed on this thread.
#define FLIP2_KERNEL(kernel, grid, block, shmem, stream, ...)                    \
    do {                                                                         \
        ::flip2::TimedScope flip2_kernel_scope_(#kernel, (stream));              \
        hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ##__VA_ARGS__);   \
    } while (0)

}  // namespace flip2

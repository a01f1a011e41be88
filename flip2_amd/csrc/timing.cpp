// Per-kernel HIP-event timing (see timing.hpp).
#include "timing.hpp"

#include "common.hpp"

namespace flip2 {

std::string kernel_base_name(const char* written) {
    std::string s(written);
    size_t a = s.find_first_not_of("( ");
    if (a == std::string::npos) return s;
    size_t b = s.find_first_of("<) ", a);
    return s.substr(a, b == std::string::npos ? std::string::npos : b - a);
}

KernelTimer*& KernelTimer::current() {
    static thread_local KernelTimer* t = nullptr;
    return t;
}

KernelTimer::~KernelTimer() {
    for (hipEvent_t e : all_) (void)hipEventDestroy(e);
}

void KernelTimer::setFilter(const std::vector<std::string>& names) {
    filter_.clear();
    filter_.insert(names.begin(), names.end());
    seen_.clear();
}

bool KernelTimer::wants(const char* name) {
    // nested scopes (a group around its kernels) record the outer one only
    if (open_) return false;
    if (!filter_.empty() && filter_.count(name) == 0) return false;
    if (every_ == 1u) return true;
    return (seen_[name]++ % every_) == 0u;
}

hipEvent_t KernelTimer::take() {
    if (free_.empty()) {
        hipEvent_t e;
        FLIP2_HIP(hipEventCreate(&e));
        all_.push_back(e);
        return e;
    }
    hipEvent_t e = free_.back();
    free_.pop_back();
    return e;
}

void KernelTimer::start(const char* name, hipStream_t s) {
    Pending p;
    p.name = name;
    p.a = take();
    p.b = take();
    FLIP2_HIP(hipEventRecord(p.a, s));
    pending_.push_back(std::move(p));
    open_ = true;
}

void KernelTimer::stop(hipStream_t s) {
    FLIP2_HIP(hipEventRecord(pending_.back().b, s));
    open_ = false;
}

void KernelTimer::collect() {
    for (Pending& p : pending_) {
        FLIP2_HIP(hipEventSynchronize(p.b));
        float ms = 0.f;
        FLIP2_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        Entry& e = totals_[p.name];
        e.ms += ms;
        e.launches += 1;
        free_.push_back(p.a);
        free_.push_back(p.b);
    }
    pending_.clear();
}

void KernelTimer::reset() {
    collect();
    totals_.clear();
    seen_.clear();
}

}  // namespace flip2

// Device manager and device vector.
//
// DeviceManager replaces the reference's CudaDeviceManager
// (engine/cudaDeviceManager.hu:26-107): device selection, properties, one
// stream, and a pool allocator with the same setPoolSize / cuMalloc / cuFree
// surface (bool true on failure, as the reference returns).
//
// DeviceVec<T> replaces CudaVec<T> (engine/cudaVec.hu:17-53, cudaVec.cu) with
// the same method names: host mirror operator[], upload/download, devPtr,
// size, resize / resizeAsync, swapDevicePtr(Async), zeroDevice(Async), clear.
// Unlike the reference, which frees and reallocates on every resizeAsync
// (particles.cu:326-351), the buffer keeps its capacity and only grows, so a
// steady-state step performs no device allocation.  Errors throw
// flip2::Error instead of calling exit().
#pragma once

#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <iostream>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace flip2 {

class DeviceManager {
public:
    explicit DeviceManager(int devID = 0) : deviceID_(devID) {
        int count = 0;
        FLIP2_HIP(hipGetDeviceCount(&count));
        FLIP2_REQUIRE(devID >= 0 && devID < count, kErrArgument, "device " + std::to_string(devID) + " not available");
        FLIP2_HIP(hipSetDevice(devID));
        FLIP2_HIP(hipGetDeviceProperties(&prop_, devID));
    }
    ~DeviceManager() {
        if (pool_) (void)hipFree(pool_);
    }
    DeviceManager(const DeviceManager&) = delete;
    DeviceManager& operator=(const DeviceManager&) = delete;

    int device() const { return deviceID_; }
    const hipDeviceProp_t& properties() const { return prop_; }

    // setPoolSize (cudaDeviceManager.hu:38-46): true on failure.
    bool setPoolSize(unsigned long long size) {
        if (pool_) return true;
        if (hipMalloc((void**)&pool_, size) != hipSuccess) {
            pool_ = nullptr;
            return true;
        }
        poolSize_ = size;
        allocations_.clear();
        return false;
    }

    // First fit over the gaps between live allocations, 256-byte aligned
    // (cudaDeviceManager.hu:48-71 bump-then-first-fit).  True on failure.
    template <typename T>
    bool cuMalloc(T*& ptr, unsigned long long count) {
        if (!pool_) return true;
        const unsigned long long bytes = ((sizeof(T) * count) + 255ull) & ~255ull;
        unsigned long long prevEnd = 0;
        for (const auto& a : allocations_) {
            if (a.first - prevEnd >= bytes) break;
            prevEnd = a.first + a.second;
        }
        if (prevEnd + bytes > poolSize_) return true;
        allocations_[prevEnd] = bytes;
        ptr = reinterpret_cast<T*>(pool_ + prevEnd);
        return false;
    }

    // cuFree (cudaDeviceManager.hu:73-91): true if ptr is not a live allocation.
    template <typename T>
    bool cuFree(T* ptr) {
        if (!pool_) return true;
        const char* c = reinterpret_cast<const char*>(ptr);
        if (c < pool_ || c >= pool_ + poolSize_) return true;
        auto it = allocations_.find((unsigned long long)(c - pool_));
        if (it == allocations_.end()) return true;
        allocations_.erase(it);
        return false;
    }

    unsigned long long poolBytesInUse() const {
        unsigned long long s = 0;
        for (const auto& a : allocations_) s += a.second;
        return s;
    }

private:
    int deviceID_;
    hipDeviceProp_t prop_{};
    char* pool_ = nullptr;
    unsigned long long poolSize_ = 0;
    std::map<unsigned long long, unsigned long long> allocations_;  // offset -> bytes
};

struct MemoryCounter {
    static size_t& bytes() {
        static size_t b = 0;
        return b;
    }
};

template <typename T>
class DeviceVec {
public:
    DeviceVec() = default;
    explicit DeviceVec(uint32_t size) { resize(size); }
    ~DeviceVec() { release(); }
    DeviceVec(const DeviceVec&) = delete;
    DeviceVec& operator=(const DeviceVec&) = delete;

    // Synchronous resize with host mirror (cudaVec.cu:20-28).  Contents are
    // not preserved when the capacity grows.
    void resize(uint32_t size) {
        reserve(size);
        numElements_ = size;
        host_.resize(size);
    }
    // Stream-ordered resize (cudaVec.cu:30-37); the host mirror is untouched.
    void resizeAsync(uint32_t size, hipStream_t) {
        reserve(size);
        numElements_ = size;
    }
    T& operator[](uint32_t i) { return host_[i]; }
    const T& operator[](uint32_t i) const { return host_[i]; }

    void upload() {
        FLIP2_REQUIRE(host_.size() == numElements_, kErrState, "upload: host mirror size differs from device size");
        if (numElements_) FLIP2_HIP(hipMemcpy(d_, host_.data(), bytes(), hipMemcpyHostToDevice));
    }
    void upload(hipStream_t s) {
        FLIP2_REQUIRE(host_.size() == numElements_, kErrState, "upload: host mirror size differs from device size");
        if (numElements_) FLIP2_HIP(hipMemcpyAsync(d_, host_.data(), bytes(), hipMemcpyHostToDevice, s));
    }
    void download() {
        host_.resize(numElements_);
        if (numElements_) FLIP2_HIP(hipMemcpy(host_.data(), d_, bytes(), hipMemcpyDeviceToHost));
    }
    void download(hipStream_t s) {
        host_.resize(numElements_);
        if (numElements_) FLIP2_HIP(hipMemcpyAsync(host_.data(), d_, bytes(), hipMemcpyDeviceToHost, s));
    }
    void uploadFrom(const T* src, uint32_t count, hipStream_t s) {
        FLIP2_REQUIRE(count <= numElements_, kErrRange, "uploadFrom: count exceeds size");
        if (count) FLIP2_HIP(hipMemcpyAsync(d_, src, (size_t)count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void downloadTo(T* dst, uint32_t count, hipStream_t s) const {
        FLIP2_REQUIRE(count <= numElements_, kErrRange, "downloadTo: count exceeds size");
        if (count) FLIP2_HIP(hipMemcpyAsync(dst, d_, (size_t)count * sizeof(T), hipMemcpyDeviceToHost, s));
    }

    // Exchanges the device buffers of two vectors (the reference adopts a raw
    // pointer and frees its own, cudaVec.cu:107-123).
    void swapDevice(DeviceVec& o) {
        std::swap(d_, o.d_);
        std::swap(capacity_, o.capacity_);
        std::swap(numElements_, o.numElements_);
    }

    T* devPtr() { return d_; }
    const T* devPtr() const { return d_; }
    uint32_t size() const { return numElements_; }
    size_t bytes() const { return (size_t)numElements_ * sizeof(T); }

    void zeroDevice() {
        if (numElements_) FLIP2_HIP(hipMemset(d_, 0, bytes()));
    }
    void zeroDeviceAsync(hipStream_t s) {
        if (numElements_) FLIP2_HIP(hipMemsetAsync(d_, 0, bytes(), s));
    }
    void clear() {
        release();
        host_.clear();
    }
    void print() const {
        for (uint32_t i = 0; i < host_.size(); ++i) std::cout << +host_[i] << ", ";
        std::cout << std::endl;
    }
    std::vector<T>& host() { return host_; }

private:
    void reserve(uint32_t size) {
        if (size <= capacity_ && d_) return;
        release();
        uint64_t cap = std::max<uint64_t>(size, 1);
        if (size > 1024) cap = (uint64_t)size + size / 8;  // headroom for per-step growth
        void* p = nullptr;
        FLIP2_HIP(hipMalloc(&p, cap * sizeof(T)));
        d_ = reinterpret_cast<T*>(p);
        capacity_ = (uint32_t)std::min<uint64_t>(cap, 0xffffffffull);
        MemoryCounter::bytes() += (size_t)capacity_ * sizeof(T);
    }
    void release() {
        if (d_) {
            (void)hipFree(d_);
            MemoryCounter::bytes() -= (size_t)capacity_ * sizeof(T);
        }
        d_ = nullptr;
        capacity_ = 0;
        numElements_ = 0;
    }

    T* d_ = nullptr;
    uint32_t capacity_ = 0;
    uint32_t numElements_ = 0;
    std::vector<T> host_;
};

}  // namespace flip2

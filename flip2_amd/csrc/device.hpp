// Device manager (an arena allocator per owner) and device vector.
//
// DeviceManager replaces the reference's CudaDeviceManager
// (engine/cudaDeviceManager.hu:26-107): device selection and properties and a
// pool with the same setPoolSize / cuMalloc / cuFree surface (bool true on
// failure, as the reference returns).  Beyond the reference, the pool is the
// allocator of the library's device vectors: it is a list of chunks, each
// handed out first-fit in 256-byte units, and it grows by a new chunk (at
// least half of what it already holds) only when no gap fits.  Freed
// regions go back to the pool, so a simulation whose node or voxel count
// moves up and down re-uses its memory instead of calling hipFree/hipMalloc.
//
// DeviceVec<T> replaces CudaVec<T> (engine/cudaVec.hu:17-53, cudaVec.cu) with
// the same method names: host mirror operator[], upload/download, devPtr,
// size, resize / resizeAsync, swapDevicePtr(Async), zeroDevice(Async),
// clear(Async), getMax.  A vector built with a DeviceManager allocates from
// its pool; one built without uses hipMalloc.  Unlike the reference, which
// frees and reallocates on every resizeAsync (particles.cu:326-351), a vector
// keeps its capacity and only grows (with 1/8 headroom), so a steady-state
// step performs no device allocation.  Errors throw flip2::Error instead of
// calling exit().
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
        for (auto& c : chunks_) (void)hipFree(c.base);
    }
    DeviceManager(const DeviceManager&) = delete;
    DeviceManager& operator=(const DeviceManager&) = delete;

    int device() const { return deviceID_; }
    const hipDeviceProp_t& properties() const { return prop_; }

    // The owner's stream: a freed region is handed out again only after the
    // work queued on it has finished.  The owner clears it (null) once that
    // stream is synchronised for good, before destroying it; without a stream
    // the owner orders its own work.
    void setStream(hipStream_t s) { stream_ = s; }

    // setPoolSize (cudaDeviceManager.hu:38-46): reserves the first chunk.
    // True on failure, or when the pool already holds memory.
    bool setPoolSize(unsigned long long size) {
        if (!chunks_.empty() || size == 0) return true;
        return !addChunk(size);
    }

    // cuMalloc (cudaDeviceManager.hu:48-71): first fit over the pool's gaps,
    // without growing it.  True on failure.
    template <typename T>
    bool cuMalloc(T*& ptr, unsigned long long count) {
        void* p = fit(sizeof(T) * count);
        if (!p) return true;
        ptr = reinterpret_cast<T*>(p);
        return false;
    }

    // cuFree (cudaDeviceManager.hu:73-91): true if ptr is not a live allocation.
    template <typename T>
    bool cuFree(T* ptr) {
        return !unfit(reinterpret_cast<const char*>(ptr));
    }

    // The vectors' allocator: first fit, growing the pool when nothing fits.
    void* allocate(size_t bytes) {
        if (void* p = fit(bytes)) return p;
        const size_t want = std::max<size_t>({round(bytes), kMinChunk, reserved_ / 2});
        if (!addChunk(want)) {
            // a smaller chunk may still fit where the geometric one did not
            if (want == round(bytes) || !addChunk(round(bytes)))
                throw Error(kErrNoMemory, "device pool: cannot reserve " + std::to_string(bytes) + " bytes");
        }
        void* p = fit(bytes);
        FLIP2_REQUIRE(p != nullptr, kErrInternal, "device pool: new chunk does not fit the request");
        return p;
    }
    // Returns a region to the pool once the owner's stream has finished with
    // it.  Called from destructors: reports nothing.
    void release(void* p) {
        if (stream_) (void)hipStreamSynchronize(stream_);
        (void)unfit(reinterpret_cast<const char*>(p));
    }
    bool owns(const void* p) const {
        const char* c = reinterpret_cast<const char*>(p);
        for (const auto& ch : chunks_)
            if (c >= ch.base && c < ch.base + ch.size) return true;
        return false;
    }
    // Frees the chunks that hold no live allocation.
    void trim() {
        if (stream_) FLIP2_HIP(hipStreamSynchronize(stream_));
        for (size_t i = chunks_.size(); i-- > 0;) {
            if (chunks_[i].live.empty()) {
                (void)hipFree(chunks_[i].base);
                reserved_ -= chunks_[i].size;
                chunks_.erase(chunks_.begin() + (long)i);
            }
        }
    }

    unsigned long long poolBytesInUse() const {
        unsigned long long s = 0;
        for (const auto& c : chunks_)
            for (const auto& a : c.live) s += a.second;
        return s;
    }
    unsigned long long poolBytesReserved() const { return reserved_; }
    // hipMalloc calls the pool has made (its chunks, including trimmed ones).
    unsigned long long chunkAllocations() const { return chunkAllocs_; }

private:
    static constexpr size_t kAlign = 256;
    static constexpr size_t kMinChunk = 64ull << 20;
    struct Chunk {
        char* base;
        size_t size;
        std::map<size_t, size_t> live;  // offset -> bytes
    };
    static size_t round(size_t bytes) { return std::max<size_t>(kAlign, (bytes + kAlign - 1) & ~(kAlign - 1)); }

    bool addChunk(size_t bytes) {
        bytes = round(bytes);
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        chunks_.push_back({reinterpret_cast<char*>(p), bytes, {}});
        reserved_ += bytes;
        ++chunkAllocs_;
        return true;
    }
    void* fit(size_t bytes) {
        const size_t need = round(bytes);
        for (auto& c : chunks_) {
            size_t prevEnd = 0;
            bool found = false;
            for (const auto& a : c.live) {
                if (a.first - prevEnd >= need) {
                    found = true;
                    break;
                }
                prevEnd = a.first + a.second;
            }
            if (found || c.size - prevEnd >= need) {
                c.live[prevEnd] = need;
                return c.base + prevEnd;
            }
        }
        return nullptr;
    }
    bool unfit(const char* p) {
        for (auto& c : chunks_) {
            if (p < c.base || p >= c.base + c.size) continue;
            auto it = c.live.find((size_t)(p - c.base));
            if (it == c.live.end()) return false;
            c.live.erase(it);
            return true;
        }
        return false;
    }

    int deviceID_;
    hipDeviceProp_t prop_{};
    hipStream_t stream_ = nullptr;
    std::vector<Chunk> chunks_;
    size_t reserved_ = 0;
    unsigned long long chunkAllocs_ = 0;
};

struct MemoryCounter {
    static size_t& bytes() {
        static size_t b = 0;
        return b;
    }
};

// CudaVec::getMax (cudaVec.cu:174-233): the reference's tree of 512-element
// tiles that keeps max(tile[0], tile[256]) per tile, level by level (only
// those two elements of each tile are compared), padding short levels with
// 0; with `absolute` the comparison is on magnitudes (identity for unsigned
// and char) and the signed element is returned.  Defined for double, float,
// uint32_t, int32_t, char and int8_t.  `scratch`: device memory for one T.
template <typename T>
T sampled_max(const T* d, uint32_t n, bool absolute, T* scratch, hipStream_t s);

template <typename T>
class DeviceVec {
public:
    DeviceVec() = default;
    explicit DeviceVec(DeviceManager& pool) : pool_(&pool) {}
    explicit DeviceVec(uint32_t size) { resize(size); }
    DeviceVec(uint32_t size, DeviceManager& pool) : pool_(&pool) { resize(size); }
    ~DeviceVec() { release(); }
    DeviceVec(const DeviceVec&) = delete;
    DeviceVec& operator=(const DeviceVec&) = delete;

    // Allocate later growth from this pool (null: hipMalloc).
    void setPool(DeviceManager* pool) { pool_ = pool; }
    DeviceManager* pool() const { return pool_; }

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

    // Exchanges the device buffers of two vectors (each buffer keeps the
    // allocator it came from).
    void swapDevice(DeviceVec& o) {
        std::swap(d_, o.d_);
        std::swap(owner_, o.owner_);
        std::swap(capacity_, o.capacity_);
        std::swap(numElements_, o.numElements_);
    }
    // swapDevicePtr / swapDevicePtrAsync (cudaVec.cu:107-124): adopt a device
    // buffer of size() elements and free the current one.  The adopted
    // pointer must come from this vector's pool or from hipMalloc; it is
    // freed the same way later.
    void swapDevicePtr(T* devPtr) {
        if (devPtr == d_) return;
        const uint32_t n = numElements_;
        release();
        adopt(devPtr, n);
    }
    void swapDevicePtrAsync(T* devPtr, hipStream_t s) {
        if (devPtr == d_) return;
        if (d_ && !owner_) FLIP2_HIP(hipStreamSynchronize(s));
        swapDevicePtr(devPtr);
    }

    T* devPtr() { return d_; }
    const T* devPtr() const { return d_; }
    uint32_t size() const { return numElements_; }
    uint32_t capacity() const { return capacity_; }
    size_t bytes() const { return (size_t)numElements_ * sizeof(T); }

    void zeroDevice() {
        if (numElements_) FLIP2_HIP(hipMemset(d_, 0, bytes()));
    }
    void zeroDeviceAsync(hipStream_t s) {
        if (numElements_) FLIP2_HIP(hipMemsetAsync(d_, 0, bytes(), s));
    }
    // getMax (cudaVec.cu:205-233): see sampled_max.  Synchronises s.
    T getMax(hipStream_t s, bool absolute = false) {
        FLIP2_REQUIRE(numElements_ > 0 && d_, kErrState, "getMax of an empty vector");
        T* scratch = nullptr;
        if (pool_) scratch = reinterpret_cast<T*>(pool_->allocate(sizeof(T)));
        else FLIP2_HIP(hipMallocAsync((void**)&scratch, sizeof(T), s));
        T out{};
        try {
            out = sampled_max<T>(d_, numElements_, absolute, scratch, s);
        } catch (...) {
            if (pool_) pool_->release(scratch);
            else (void)hipFreeAsync(scratch, s);
            throw;
        }
        if (pool_) pool_->release(scratch);
        else FLIP2_HIP(hipFreeAsync(scratch, s));
        return out;
    }
    void clear() {
        release();
        host_.clear();
    }
    void clearAsync(hipStream_t s) {
        if (d_ && !owner_) FLIP2_HIP(hipStreamSynchronize(s));
        clear();
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
        cap = std::min<uint64_t>(cap, 0xffffffffull);
        void* p = nullptr;
        if (pool_) {
            p = pool_->allocate(cap * sizeof(T));
        } else {
            FLIP2_HIP(hipMalloc(&p, cap * sizeof(T)));
        }
        d_ = reinterpret_cast<T*>(p);
        owner_ = pool_;
        capacity_ = (uint32_t)cap;
        MemoryCounter::bytes() += (size_t)capacity_ * sizeof(T);
    }
    void adopt(T* p, uint32_t n) {
        d_ = p;
        owner_ = (p && pool_ && pool_->owns(p)) ? pool_ : nullptr;
        capacity_ = p ? n : 0;
        numElements_ = p ? n : 0;
        MemoryCounter::bytes() += (size_t)capacity_ * sizeof(T);
    }
    void release() {
        if (d_) {
            if (owner_) owner_->release(d_);
            else (void)hipFree(d_);
            MemoryCounter::bytes() -= (size_t)capacity_ * sizeof(T);
        }
        d_ = nullptr;
        owner_ = nullptr;
        capacity_ = 0;
        numElements_ = 0;
    }

    T* d_ = nullptr;
    DeviceManager* pool_ = nullptr;   // allocator of future growth
    DeviceManager* owner_ = nullptr;  // allocator of the current buffer (null: hipMalloc)
    uint32_t capacity_ = 0;
    uint32_t numElements_ = 0;
    std::vector<T> host_;
};

}  // namespace flip2

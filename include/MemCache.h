// Drop-in device memory pool with the reference's interface
// (reference: include/MemCache.h:1-24, MemCache.cpp:5-129).
//
// Size classes {32 B, 1 KiB, 64 KiB, 128 KiB, 256 KiB, 512 KiB, 1 MiB, 2 MiB,
// 4 MiB, 8 MiB, 24 MiB, 512 MiB} as in MemCache.cpp:5-6: a request takes the
// smallest class >= its size.  release must be given a size in the same
// class as the request (the reference's contract, utils.cuh:33-39).
//
// Differences from the reference, all additive:
//   - nothing is preallocated until initAbtrtMem() is called (the reference
//     preallocates about 3.1 GiB, twice; MemCache.cpp:60-76,106-116);
//   - a mutex makes the pool usable from several host threads;
//   - a byte ledger (bytesInUse / bytesCached / peak) and an optional budget
//     on the bytes in use, which the serving path sets from its admission
//     budget;
//   - requests above 512 MiB are rounded up to 64 MiB and cached like the
//     classes (the reference cudaMalloc/cudaFrees them every time, and every
//     free synchronises the device); when hipMalloc fails the whole cache is
//     returned to the device and the allocation retried once;
//   - releaseCache() returns every cached block.
#pragma once
#include <cstddef>
#include <list>
#include <mutex>
#include <stack>
#include <unordered_map>
#include <vector>

class MemCache {
public:
    void* getMemByAbtrtSize(size_t size);
    void* getMemBySize(size_t size);

    void releaseAbtrtMem(void* ptr, size_t size);
    void releaseMem(void* ptr, size_t size);
    void initAbtrtMem();
    void releaseCache();

    void* getPinnedPage();
    void releasePinnedPage(void* mem);
    void report();

    // ledger
    size_t bytesInUse() const { return inUse_; }
    size_t bytesCached() const { return cached_; }
    size_t peakBytes() const { return peak_; }
    void resetPeak() { peak_ = inUse_; }
    // budget for admission control (0 = unlimited): an allocation that would
    // take the bytes in use past it fails (returns nullptr)
    void setBudget(size_t bytes) { budget_ = bytes; }
    size_t budget() const { return budget_; }
    // bytes a request of `size` occupies in the pool
    static size_t classBytes(size_t size);

private:
    std::mutex mu_;
    std::unordered_map<size_t, std::list<void*>> unordered_mp;
    std::vector<std::stack<void*>> stkInVec;
    // blocks above the largest class, kept by their size rounded up to 64 MiB
    std::unordered_map<size_t, std::vector<void*>> large_;
    void* m_pinnedPage{nullptr};
    bool inited_{false};
    bool preallocated_{false};
    size_t inUse_{0}, cached_{0}, peak_{0}, budget_{0};
    void ensureClasses();
    void freeCachedLocked();
};

MemCache& getMemCacheRef();

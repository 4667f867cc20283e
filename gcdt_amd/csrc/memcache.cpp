// Size-class device pool; see include/MemCache.h.
#include "../../include/MemCache.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

static const size_t kSizes[] = {32, 1024, 65536, 131072, 262144, 524288, 1048576, 2097152,
                                4194304, 8388608, 25165824, 536870912};
static const int kInitCount[] = {32, 32, 16, 16, 16, 16, 16, 8, 8, 4, 1, 6};
static const int kClasses = sizeof(kSizes) / sizeof(kSizes[0]);

static int classOf(size_t size)
{
    const size_t* p = std::lower_bound(kSizes, kSizes + kClasses, size);
    return p == kSizes + kClasses ? -1 : (int)(p - kSizes);
}

// Requests above the largest class are rounded up to 64 MiB and cached too:
// freeing multi-GB buffers on every solve would synchronise the device in
// the middle of the pipelined rounds and pay the mapping cost again.
static const size_t kLargeStep = size_t(64) << 20;
static size_t largeBytes(size_t size) { return (size + kLargeStep - 1) / kLargeStep * kLargeStep; }

size_t MemCache::classBytes(size_t size)
{
    int c = classOf(size);
    return c < 0 ? largeBytes(size) : kSizes[c];
}

// Returns every cached block to the device (the caller holds mu_).
void MemCache::freeCachedLocked()
{
    for (auto& kv : large_)
        for (void* p : kv.second) (void)hipFree(p);
    large_.clear();
    for (auto& s : stkInVec)
        while (!s.empty()) {
            (void)hipFree(s.top());
            s.pop();
        }
    cached_ = 0;
}

void MemCache::ensureClasses()
{
    if (!inited_) {
        stkInVec.resize(kClasses);
        inited_ = true;
    }
}

void* MemCache::getMemByAbtrtSize(size_t size)
{
    std::lock_guard<std::mutex> g(mu_);
    ensureClasses();
    void* mem = nullptr;
    int c = classOf(size);
    size_t bytes = c < 0 ? largeBytes(size) : kSizes[c];
    // the budget bounds the bytes in use (admission control sets it)
    if (budget_ && inUse_ + bytes > budget_) return nullptr;
    if (c >= 0 && !stkInVec[c].empty()) {
        mem = stkInVec[c].top();
        stkInVec[c].pop();
        cached_ -= bytes;
    } else if (c < 0 && !large_[bytes].empty()) {
        mem = large_[bytes].back();
        large_[bytes].pop_back();
        cached_ -= bytes;
    } else if (hipMalloc(&mem, bytes == 0 ? 1 : bytes) != hipSuccess) {
        // out of device memory: give the cache back and try once more
        (void)hipGetLastError();
        freeCachedLocked();
        if (hipMalloc(&mem, bytes == 0 ? 1 : bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    inUse_ += bytes;
    peak_ = std::max(peak_, inUse_);
    return mem;
}

void MemCache::releaseAbtrtMem(void* ptr, size_t size)
{
    if (!ptr) return;
    std::lock_guard<std::mutex> g(mu_);
    ensureClasses();
    int c = classOf(size);
    size_t bytes = c < 0 ? largeBytes(size) : kSizes[c];
    inUse_ -= std::min(inUse_, bytes);
    if (c < 0) large_[bytes].push_back(ptr);
    else stkInVec[c].push(ptr);
    cached_ += bytes;
}

void* MemCache::getMemBySize(size_t size)
{
    std::lock_guard<std::mutex> g(mu_);
    auto it = unordered_mp.find(size);
    if (it == unordered_mp.end() || it->second.empty()) return nullptr;
    void* p = it->second.back();
    it->second.pop_back();
    return p;
}

void MemCache::releaseMem(void* ptr, size_t size)
{
    std::lock_guard<std::mutex> g(mu_);
    unordered_mp[size].push_back(ptr);
}

void MemCache::initAbtrtMem()
{
    std::lock_guard<std::mutex> g(mu_);
    ensureClasses();
    if (preallocated_) return;
    preallocated_ = true;
    for (int c = 0; c < kClasses; c++) {
        for (int j = 0; j < kInitCount[c]; j++) {
            void* mem = nullptr;
            if (hipMalloc(&mem, kSizes[c]) != hipSuccess) return;
            stkInVec[c].push(mem);
            cached_ += kSizes[c];
        }
    }
}

void MemCache::releaseCache()
{
    std::lock_guard<std::mutex> g(mu_);
    for (auto& kv : unordered_mp)
        for (void* p : kv.second) (void)hipFree(p);
    unordered_mp.clear();
    freeCachedLocked();
    if (m_pinnedPage) (void)hipHostFree(m_pinnedPage);
    m_pinnedPage = nullptr;
}

void* MemCache::getPinnedPage() { return m_pinnedPage; }
void MemCache::releasePinnedPage(void* mem) { m_pinnedPage = mem; }

void MemCache::report()
{
    std::lock_guard<std::mutex> g(mu_);
    ensureClasses();
    printf("pool report: in use %zu B, cached %zu B, peak %zu B\n", inUse_, cached_, peak_);
    for (int c = 0; c < kClasses; c++) printf("%d (%zu B): %zu free\n", c, kSizes[c], stkInVec[c].size());
}

MemCache& getMemCacheRef()
{
    static MemCache cache;
    return cache;
}

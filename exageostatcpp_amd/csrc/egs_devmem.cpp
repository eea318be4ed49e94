// Size-class cache of device blocks (see egs_devmem.hpp).
#include "egs_devmem.hpp"

#include <algorithm>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace egs {

namespace {

constexpr int kMaxDev = 64;

struct Block {
    int device;
    int64_t bytes;
};

struct Cache {
    std::mutex mu;
    int64_t limit = 0;
    int64_t cached[kMaxDev] = {};
    int64_t live[kMaxDev] = {};
    std::unordered_map<int64_t, std::vector<void *>> free_blocks[kMaxDev];
    std::unordered_map<void *, Block> owner;
};

Cache &cache() {
    static Cache *c = new Cache();   // never destroyed: blocks may outlive static teardown
    return *c;
}

int current_device() {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kMaxDev) return 0;
    return d;
}

// Frees cached blocks of device d until live + cached <= keep (mu held, d current).
void trim(Cache &c, int d, int64_t keep) {
    for (auto it = c.free_blocks[d].begin(); it != c.free_blocks[d].end() && c.live[d] + c.cached[d] > keep;) {
        auto &v = it->second;
        while (!v.empty() && c.live[d] + c.cached[d] > keep) {
            hipFree(v.back());
            v.pop_back();
            c.cached[d] -= it->first;
        }
        it = v.empty() ? c.free_blocks[d].erase(it) : std::next(it);
    }
}

}  // namespace

int64_t dev_class_bytes(size_t bytes) {
    // 256-byte granules up to 64 KiB, then eight classes per octave (at most
    // 1/8 over the request)
    const int64_t b = (int64_t) std::max<size_t>(bytes, 1);
    if (b <= (int64_t(1) << 16)) return (b + 255) & ~int64_t(255);
    int e = 63 - __builtin_clzll((unsigned long long) b);   // 2^e <= b < 2^(e+1)
    const int64_t step = int64_t(1) << (e - 3);
    return (b + step - 1) / step * step;
}

hipError_t dev_alloc(void **ptr, size_t bytes, int64_t *owner_bytes) {
    Cache &c = cache();
    const int d = current_device();
    const int64_t cls = dev_class_bytes(bytes);
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.free_blocks[d].find(cls);
    if (it != c.free_blocks[d].end() && !it->second.empty()) {
        *ptr = it->second.back();
        it->second.pop_back();
        c.cached[d] -= cls;
    } else {
        // keep live + cached within the limit before taking new memory
        if (c.cached[d] > 0 && c.live[d] + c.cached[d] + cls > c.limit) trim(c, d, std::max<int64_t>(0, c.limit - cls));
        hipError_t e = hipMalloc(ptr, (size_t) cls);
        if (e != hipSuccess && c.cached[d] > 0) {
            (void) hipGetLastError();
            trim(c, d, 0);
            e = hipMalloc(ptr, (size_t) cls);
        }
        if (e != hipSuccess) {
            *ptr = nullptr;
            return e;
        }
    }
    c.owner[*ptr] = Block{d, cls};
    c.live[d] += cls;
    if (owner_bytes) *owner_bytes += cls;
    return hipSuccess;
}

void dev_free(void *ptr, int64_t *owner_bytes) {
    if (!ptr) return;
    Cache &c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.owner.find(ptr);
    if (it == c.owner.end()) {
        hipFree(ptr);   // not from dev_alloc
        return;
    }
    const Block b = it->second;
    c.owner.erase(it);
    c.live[b.device] -= b.bytes;
    if (owner_bytes) *owner_bytes -= b.bytes;
    if (c.live[b.device] + c.cached[b.device] + b.bytes <= c.limit) {
        c.free_blocks[b.device][b.bytes].push_back(ptr);
        c.cached[b.device] += b.bytes;
    } else {
        int cur = 0;
        hipGetDevice(&cur);
        if (cur != b.device) hipSetDevice(b.device);
        hipFree(ptr);
        if (cur != b.device) hipSetDevice(cur);
    }
}

void dev_cache(int64_t limit, int64_t *cached, int64_t *live) {
    Cache &c = cache();
    const int d = current_device();
    std::lock_guard<std::mutex> lk(c.mu);
    if (limit >= 0) {
        c.limit = limit;
        trim(c, d, std::max(limit, c.live[d]));
    }
    if (cached) *cached = c.cached[d];
    if (live) *live = c.live[d];
}

}  // namespace egs

// Device memory of problems: blocks come from hipMalloc in size classes and,
// when released, are kept in a per-device cache for the next allocation of the
// same class instead of going back through hipFree, which waits for the whole
// device (every stream of every context) before it returns.  The cache holds
// at most `limit` bytes together with the live blocks (0, the default: no
// caching, every release is a hipFree).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace egs {

// Bytes a request of `bytes` occupies (its size class).
int64_t dev_class_bytes(size_t bytes);
// Allocates on the current device; adds the class bytes to *owner_bytes.
hipError_t dev_alloc(void **ptr, size_t bytes, int64_t *owner_bytes);
// Releases a block of dev_alloc (null: no-op); subtracts from *owner_bytes.
void dev_free(void *ptr, int64_t *owner_bytes);
// Sets the cache limit (limit < 0: unchanged), releases cached blocks beyond
// it, and reports the cached and live bytes of the current device.
void dev_cache(int64_t limit, int64_t *cached, int64_t *live);

}  // namespace egs

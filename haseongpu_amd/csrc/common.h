// common.h - error handling, device buffers and constants shared by the
// translation units of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/hase_ase.h"

namespace hase {

constexpr int kBlock = 256;
// Buckets of the proposal prefix per sample: bucket k holds the first entry
// whose prefix exceeds k / kPrefixBuckets, so the ray search for a target in
// [k, k+1) / kPrefixBuckets covers only the entries between two buckets.
constexpr uint32_t kPrefixBuckets = 2048;

// message of the last failed call on this thread (hase_last_error)
extern thread_local std::string g_last_error;

struct HaseError : std::runtime_error {
    int code;
    HaseError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                    \
    do {                                                                                   \
        hipError_t err_ = (expr);                                                          \
        if (err_ != hipSuccess)                                                            \
            throw ::hase::HaseError(HASE_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(err_)); \
    } while (0)

// Bump allocator over a caller-provided device region.  While a thread has
// one installed (ArenaScope), every DevBuf of that thread takes its memory
// from it instead of hipMalloc, and never frees it: the region is reused as a
// whole by the next call that installs it.
struct Arena {
    char* base = nullptr;
    size_t capacity = 0;
    size_t used = 0;
    size_t remaining() const { return capacity - used; }
    void* take(size_t bytes)
    {
        const size_t start = (used + 255u) & ~(size_t)255u;
        if (start > capacity || bytes > capacity - start)
            throw HaseError(HASE_E_MEMORY, "request of " + std::to_string(bytes) + " bytes exceeds the " +
                                               std::to_string(capacity) + "-byte no-allocation scratch (" +
                                               std::to_string(capacity - std::min(start, capacity)) +
                                               " bytes left)");
        used = start + bytes;
        return base + start;
    }
};

// the arena of the calling thread, or null (DevBuf then uses hipMalloc)
extern thread_local Arena* g_arena;

struct ArenaScope {
    Arena* prev;
    explicit ArenaScope(Arena* a) : prev(g_arena) { g_arena = a; }
    ~ArenaScope() { g_arena = prev; }
    ArenaScope(const ArenaScope&) = delete;
    ArenaScope& operator=(const ArenaScope&) = delete;
};

// Owning device allocation that grows on demand (or a slice of the thread's
// arena, which it does not own).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = false;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = false;
    }
    void alloc(size_t count)
    {
        if (count <= n && p) return;
        release();
        if (count == 0) return;
        if (g_arena) {
            p = static_cast<T*>(g_arena->take(count * sizeof(T)));
            n = count;
            return;
        }
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw HaseError(HASE_E_MEMORY, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                               " bytes failed: " + hipGetErrorString(e));
        }
        n = count;
        owned = true;
    }
    void upload(const T* src, size_t count, hipStream_t st)
    {
        alloc(count);
        if (count) HIP_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, st));
    }
};

// Runs f, turning every exception into a status code and a thread-local message.
template <class F>
int guarded(F&& f)
{
    try {
        g_last_error.clear();
        f();
        return HASE_OK;
    } catch (const HaseError& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return HASE_E_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return HASE_E_INTERNAL;
    } catch (...) {
        g_last_error = "unknown failure";
        return HASE_E_INTERNAL;
    }
}

}  // namespace hase

// common.h - error handling, device buffers and constants shared by the
// translation units of the library.
#pragma once

#include <hip/hip_runtime.h>

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
constexpr float kPi = 3.14159265358979f;

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

// Owning device allocation that grows on demand.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count)
    {
        if (count <= n && p) return;
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw HaseError(HASE_E_MEMORY, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                               " bytes failed: " + hipGetErrorString(e));
        }
        n = count;
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

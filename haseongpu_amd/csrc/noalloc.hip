// noalloc.hip - hase_calc_phi_ase_noalloc: the one-shot evaluation of
// hase_calc_phi_ase with every device buffer (mesh records, beta copies,
// proposal tables, per-sample sums, work-queue state) carved out of one
// file-scope device array instead of hipMalloc, for callers that cannot
// allocate device memory.
//
// The array lives in this translation unit's code object, so only a process
// that calls this entry point loads it, and each GPU that runs it holds its
// own copy until the process ends.  Host code never passes the array's name
// to a kernel or a copy: the name is a host-side stand-in.  The device
// address of each GPU's copy is looked up once per device
// (hipGetSymbolAddress, with that device current) and only that address is
// handed out.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.h"
#include "context.h"

static __device__ __align__(256) unsigned char scratch[HASE_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_lookup;                    // guards g_scratch
void* g_scratch[kMaxDevices] = {};      // device address of each GPU's copy
std::mutex g_device_lock[kMaxDevices];  // one call at a time per GPU in this process

// Device address of the current device's copy of the scratch array.
void* scratch_address(int device)
{
    std::lock_guard<std::mutex> lock(g_lookup);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw hase::HaseError(HASE_E_DEVICE, "hipGetSymbolAddress returned no address for the scratch");
        g_scratch[device] = ptr;
    }
    return g_scratch[device];
}

}  // namespace

extern "C" int hase_calc_phi_ase_noalloc(const hase_mesh_t* mesh, const hase_experiment_t* experiment,
                                         const hase_compute_t* compute, hase_result_t* result)
{
    return hase::guarded([&] {
        hase::validate_arguments(mesh, experiment, compute);
        if (!result || !result->phi_ase || !result->mse || !result->total_rays || !result->dndt_ase)
            throw hase::HaseError(HASE_E_INVALID, "result arrays phi_ase, mse, total_rays and dndt_ase are required");
        int count = 0;
        HIP_CHECK(hipGetDeviceCount(&count));
        if (count <= 0) throw hase::HaseError(HASE_E_DEVICE, "no HIP device is visible");
        const int device = compute->device;
        if (device < 0 || device >= count || device >= kMaxDevices)
            throw hase::HaseError(HASE_E_INVALID, "device ordinal out of range");
        HIP_CHECK(hipSetDevice(device));
        void* base = scratch_address(device);
        std::lock_guard<std::mutex> lock(g_device_lock[device]);
        hase::Arena arena;
        arena.base = static_cast<char*>(base);
        arena.capacity = HASE_NOALLOC_SCRATCH_BYTES;
        hase::ArenaScope scope(&arena);
        hase::calc_phi_ase(mesh, experiment, compute, result);
    });
}

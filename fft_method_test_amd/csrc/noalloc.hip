// noalloc.hip -- the support map for callers that cannot allocate device memory.
//
// fftperm_support_noalloc classifies and compacts the support exactly as fftperm_set_medium_dev
// (the same count and scan kernels, src/geometry/InterfaceVoxels.cpp:7-43 and
// src/bignonnet/ForceSupport.cpp:9-47), but owns no context and allocates nothing: its scan
// scratch -- two counters and one int per 1024 voxels -- lives in the file-scope device array
// `scratch` below, whose size fixes the largest grid it accepts.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/fftperm.h"
#include "kernels.hpp"

namespace {

constexpr size_t kCountersBytes = 256;  // counts[0] = support voxels, counts[1] = fluid voxels
constexpr unsigned long long kMaxVoxels = FFTPERM_NOALLOC_MAX_VOXELS;
constexpr size_t kScratchBytes = kCountersBytes + (size_t)((kMaxVoxels + 1023) / 1024) * sizeof(int);

// One copy per GPU, held until the process ends.  Host code never uses this name as an address:
// it is a host-side stand-in; the device address comes from hipGetSymbolAddress (per device).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's copy of `scratch`, looked up once per device with that device current.
hipError_t scratch_address(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
    return hipSuccess;
}

int fail(int code, const std::string& msg) {
    fftperm::set_last_error(msg);
    return code;
}

__global__ void k_copy_counts(const unsigned long long* __restrict__ src, unsigned long long* __restrict__ dst) {
    if (threadIdx.x < 2) dst[threadIdx.x] = src[threadIdx.x];
}

}  // namespace

extern "C" {

int fftperm_support_noalloc(uint64_t nx, uint64_t ny, uint64_t nz, const uint8_t* fluid_mask_dev, int support_mode,
                            int64_t* voxel_to_local_dev, int64_t* active_dev, uint64_t active_capacity,
                            uint64_t* counts_dev, void* stream, uint64_t* nb_out, uint64_t* nfluid_out) {
    if (support_mode != FFTPERM_SUPPORT_INTERFACE && support_mode != FFTPERM_SUPPORT_FULLSOLID)
        return fail(FFTPERM_EINVAL, "unknown support mode");
    if (!fluid_mask_dev || (!active_dev && active_capacity > 0))
        return fail(FFTPERM_EINVAL, "null argument");
    if (nx == 0 || ny == 0 || nz == 0) return fail(FFTPERM_EINVAL, "Grid dimensions must be positive");
    if (nx > (1u << 20) || ny > (1u << 20) || nz > (1u << 20) || nx * ny * nz > kMaxVoxels)
        return fail(FFTPERM_ERANGE, "fftperm_support_noalloc: grid larger than FFTPERM_NOALLOC_MAX_VOXELS voxels "
                                    "(the size of its static scratch)");
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return fail(FFTPERM_EDEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (device < 0 || device >= kMaxDevices) return fail(FFTPERM_EDEVICE, "device ordinal out of range");
    unsigned char* base = nullptr;
    e = scratch_address(device, &base);
    if (e != hipSuccess) return fail(FFTPERM_EDEVICE, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
    auto* counts = reinterpret_cast<unsigned long long*>(base);
    auto* bsum = reinterpret_cast<int*>(base + kCountersBytes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    (void)hipGetLastError();  // so that the check below sees only these launches
    fftperm::launch_support_v2l(fluid_mask_dev, (int)nx, (int)ny, (int)nz,
                                support_mode == FFTPERM_SUPPORT_INTERFACE ? 1 : 0, voxel_to_local_dev, active_dev,
                                (long)active_capacity, counts, bsum, st);
    if (counts_dev) hipLaunchKernelGGL(k_copy_counts, dim3(1), dim3(64), 0, st, counts, (unsigned long long*)counts_dev);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(FFTPERM_EDEVICE, std::string("support launch: ") + hipGetErrorString(e));
    if (nb_out || nfluid_out) {
        unsigned long long c[2] = {0, 0};
        e = hipMemcpyAsync(c, counts, sizeof(c), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return fail(FFTPERM_EDEVICE, std::string("count read-back: ") + hipGetErrorString(e));
        if (nb_out) *nb_out = c[0];
        if (nfluid_out) *nfluid_out = c[1];
        if (c[0] > active_capacity)
            return fail(FFTPERM_ERANGE, "fftperm_support_noalloc: the support has more voxels than active_capacity");
    }
    fftperm::set_last_error("");
    return FFTPERM_OK;
}

}  // extern "C"

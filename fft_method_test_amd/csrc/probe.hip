// probe.hip -- HBM copy probes on a context's own FP32 spectrum buffers (measurement hooks).
//
// The passes are HBM-bound, so their fraction of the 8 TB/s nominal peak has two parts: what the
// device reaches at all, and what their access patterns allow.  Three copies of the three padded
// FP32 half spectra (spectrum slot 0 -> slot 1 of a MIXED context) measure both:
//   0  a contiguous stream (16-byte non-temporal loads and stores): the reachable HBM bandwidth;
//   1  the y passes' pattern: tiles of 32 consecutive row elements (256-byte segments) walked down
//      the column (ky) at the row pitch, one tile and component per workgroup;
//   2  the z pass's pattern: tiles of 8 consecutive row elements (64-byte segments) walked down the
//      z column at the plane pitch.
// Each kernel only copies; the rate is (bytes read + bytes written) / time.
#include "kernels.hpp"

namespace fftperm {

namespace {

constexpr int kProbeThreads = 256;

// Block b copies the 16-byte elements [b, b + 1) * kProbeThreads * KS: each thread issues its KS
// loads before its stores.
constexpr int KS = 8;
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kProbeThreads) void k_probe_stream(const f32x4* __restrict__ src,
                                                                 f32x4* __restrict__ dst, long n4) {
    const long i0 = (long)blockIdx.x * kProbeThreads * KS + threadIdx.x;
    if (i0 + (long)(KS - 1) * kProbeThreads < n4) {
        f32x4 v[KS];
#pragma unroll
        for (int k = 0; k < KS; ++k) v[k] = __builtin_nontemporal_load(src + i0 + k * kProbeThreads);
#pragma unroll
        for (int k = 0; k < KS; ++k) __builtin_nontemporal_store(v[k], dst + i0 + k * kProbeThreads);
    } else {
        for (long i = i0; i < n4; i += kProbeThreads) dst[i] = src[i];
    }
}

// Column tiles: tile t of component blockIdx.y covers columns t * TL .. t * TL + TL - 1, column col
// at (col / inner) * outer + col % inner, element e of the column at + e * estride.
template <int TL>
__global__ __launch_bounds__(kProbeThreads) void k_probe_cols(const float2* __restrict__ src,
                                                               float2* __restrict__ dst, long ncols, long inner,
                                                               long outer, long estride, int n, long comp) {
    constexpr int RPI = kProbeThreads / TL;  // column elements per thread step
    constexpr int KC = 16;                     // loads in flight per thread before the stores
    const long col = (long)blockIdx.x * TL + threadIdx.x % TL;
    if (col >= ncols) return;
    const long base = (long)blockIdx.y * comp + (col / inner) * outer + col % inner;
    int e = threadIdx.x / TL;
    for (; e + (KC - 1) * RPI < n; e += KC * RPI) {
        float2 v[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k) v[k] = src[base + (long)(e + k * RPI) * estride];
#pragma unroll
        for (int k = 0; k < KC; ++k) dst[base + (long)(e + k * RPI) * estride] = v[k];
    }
    for (; e < n; e += RPI) dst[base + (long)e * estride] = src[base + (long)e * estride];
}

}  // namespace

// Bytes one probe launch reads plus writes, or -1 for an unknown pattern.
long probe_bytes(int pattern, const Geometry3& g) {
    const long cx = (long)sizeof(float2);
    switch (pattern) {
        case 0: return 2L * 3 * g.Nc * cx;
        case 1:
        case 2: return 2L * 3 * (long)g.nz * g.ny * g.rp * cx;
        default: return -1;
    }
}

void launch_probe(int pattern, const float2* src, float2* dst, const Geometry3& g, hipStream_t st) {
    if (pattern == 0) {
        const long n4 = 3 * g.Nc * (long)sizeof(float2) / (long)sizeof(float4);
        const long blocks = (n4 + (long)kProbeThreads * KS - 1) / ((long)kProbeThreads * KS);
        hipLaunchKernelGGL(k_probe_stream, dim3((unsigned)blocks), dim3(kProbeThreads), 0, st,
                           reinterpret_cast<const f32x4*>(src), reinterpret_cast<f32x4*>(dst), n4);
    } else if (pattern == 1) {
        // y columns (kz, kx < rp): base kz * pz + kx, elements ky * rp
        constexpr int TL = 32;
        const long ncols = (long)g.nz * g.rp;
        hipLaunchKernelGGL(k_probe_cols<TL>, dim3((unsigned)((ncols + TL - 1) / TL), 3), dim3(kProbeThreads), 0, st,
                           src, dst, ncols, (long)g.rp, g.pz, (long)g.rp, g.ny, g.Nc);
    } else {
        // z columns (ky, kx < rp): base ky * rp + kx, elements kz * pz
        constexpr int TL = 8;
        const long ncols = (long)g.ny * g.rp;
        hipLaunchKernelGGL(k_probe_cols<TL>, dim3((unsigned)((ncols + TL - 1) / TL), 3), dim3(kProbeThreads), 0, st,
                           src, dst, ncols, ncols, 0L, g.pz, g.nz, g.Nc);
    }
}

}  // namespace fftperm

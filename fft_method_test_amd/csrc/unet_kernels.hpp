// unet_kernels.hpp -- kernels of the learned warm start: a 3D U-Net in channels-last BF16 with
// FP32 accumulation on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Activations: [V][C] BF16 (V = n^3 voxels, flat index (k n + j) n + i, C channels contiguous),
// C a multiple of 8.  Convolutions are periodic (the cell is periodic), 3x3x3 (27 taps,
// t = (dk+1)*9 + (dj+1)*3 + (di+1)) or 1x1x1, computed as an implicit GEMM:
//   Y[v][co] = sum_k A[v][k] W[co][k],  k = t*Cin + ci,  A[v][k] = X[neighbour(v, t)][ci]
// with the weights stored [Cout][Kp] (Kp = 27 Cin rounded up to 32, zero padded), so that one
// lane's 8 consecutive K entries are one 16-byte load of the activation and of the weights.
// A convolution may read its input from two tensors (channel concatenation [in0 | in1]).
#pragma once

#include "common.hpp"

namespace fftperm {
namespace unet {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// float <-> bf16 bits, round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ v8bf load8(const uint16_t* p) {
    const uint4 u = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(v8bf, u);
}

__device__ __forceinline__ v8bf zero8() { return __builtin_bit_cast(v8bf, make_uint4(0u, 0u, 0u, 0u)); }

__device__ __forceinline__ int wrapn(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Neighbour of voxel (i, j, k) across tap t of a TAPS-point stencil (27: 3x3x3, 1: identity).
template <int TAPS>
__device__ __forceinline__ long neighbour(int i, int j, int k, int n, int t) {
    if constexpr (TAPS == 1) {
        return ((long)k * n + j) * n + i;
    } else {
        const int dk = t / 9 - 1, dj = (t / 3) % 3 - 1, di = t % 3 - 1;
        return ((long)wrapn(k + dk, n) * n + wrapn(j + dj, n)) * n + wrapn(i + di, n);
    }
}

// Forward convolution (also the data gradient, with transposed, tap-flipped weights).
// One wave computes 16 voxels x 16*NT output channels; 4 waves per workgroup.
// out_f32 == false: out BF16 [V][Cout]; true: out FP32 [V][Cout].
template <int TAPS, int NT>
__global__ __launch_bounds__(256) void k_conv_fwd(const uint16_t* __restrict__ in0, int c0,
                                                  const uint16_t* __restrict__ in1, int c1, int n, long V,
                                                  const uint16_t* __restrict__ wt, int Kp,
                                                  const float* __restrict__ bias, int relu, int out_f32,
                                                  void* __restrict__ out) {
    constexpr int COUT = 16 * NT;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long v0 = ((long)blockIdx.x * 4 + wave) * 16;
    if (v0 >= V) return;
    const int row = lane & 15;  // A: voxel row; B: output-channel column
    const int g = lane >> 4;    // K group of 8
    const long v = v0 + row;
    const bool valid = v < V;
    const int i = (int)(v % n);
    const int j = (int)((v / n) % n);
    const int k = (int)(v / ((long)n * n));
    const int cin = c0 + c1;
    const int kmax = TAPS * cin;
    v4f acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = (v4f){0.f, 0.f, 0.f, 0.f};
    const int nchunks = Kp / 32;
    for (int kc = 0; kc < nchunks; ++kc) {
        const int kg = kc * 32 + 8 * g;
        v8bf a = zero8();
        if (valid && kg < kmax) {
            const int t = kg / cin;
            const int ci = kg - t * cin;
            const long nb = neighbour<TAPS>(i, j, k, n, t);
            a = (ci < c0) ? load8(in0 + nb * c0 + ci) : load8(in1 + nb * c1 + (ci - c0));
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const v8bf b = load8(wt + (long)(nt * 16 + row) * Kp + kg);
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nt], 0, 0, 0);
        }
    }
    // D[r][c]: lane holds rows 4g + q (voxels), column row (channel nt*16 + row)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int co = nt * 16 + row;
        const float bb = bias != nullptr ? bias[co] : 0.f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long vo = v0 + 4 * g + q;
            if (vo >= V) continue;
            float y = acc[nt][q] + bb;
            if (relu && y < 0.f) y = 0.f;
            if (out_f32)
                static_cast<float*>(out)[vo * COUT + co] = y;
            else
                static_cast<uint16_t*>(out)[vo * COUT + co] = f2bf(y);
        }
    }
}

// Halo-tiled 3x3x3 convolution: a workgroup computes a brick of 64 (x) x 2 (y) x 2 (z) voxels, one
// wave per x row of 64 voxels (four 16-voxel MFMA tiles, so each weight fragment serves four
// MFMAs).  The brick's neighbourhood (4 x 4 rows of 66 voxels) is staged in LDS 16 channels at a
// time and every tap reads it from there, instead of each wave gathering its 27 neighbours through
// the caches.  K runs over channel chunks, then tap pairs (taps t, t + 1 x 16 channels = one K-32
// MFMA step).  Needs n % 64 == 0 and 16 | c0, c1 (conv_fwd falls back to k_conv_fwd otherwise).
constexpr int kTileX = 64;
constexpr int kTileRows = 16;  // (z - 1 .. z + 2) x (y - 1 .. y + 2)
constexpr int kTileW = kTileX + 2;

template <int NT>
__global__ __launch_bounds__(256) void k_conv3_tile(const uint16_t* __restrict__ in0, int c0,
                                                    const uint16_t* __restrict__ in1, int c1, int n,
                                                    const uint16_t* __restrict__ wt, int Kp,
                                                    const float* __restrict__ bias, int relu, int out_f32,
                                                    void* __restrict__ out) {
    constexpr int COUT = 16 * NT;
    __shared__ uint4 st[kTileRows * kTileW * 2];  // 16 bf16 channels per voxel = two 16-byte words
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int row = lane & 15;  // A: voxel within an M tile; B: output channel within an N tile
    const int g = lane >> 4;    // K group of 8: tap t + (g >> 1), channels 8 (g & 1) .. + 7
    const int ntx = n / kTileX;
    const int bx = blockIdx.x % ntx;
    const int byz = blockIdx.x / ntx;
    const int i0 = bx * kTileX;
    const int j0 = (byz % (n / 2)) * 2;
    const int k0 = (byz / (n / 2)) * 2;
    const int jj = wave & 1, kk = wave >> 1;  // this wave's output row (j0 + jj, k0 + kk)
    const int cin = c0 + c1;
    v4f acc[4][NT];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[m][nt] = (v4f){0.f, 0.f, 0.f, 0.f};
    for (int cc = 0; cc < cin; cc += 16) {
        // stage channels cc .. cc + 15 of the neighbourhood (periodic)
        const uint16_t* src = cc < c0 ? in0 : in1;
        const int cs = cc < c0 ? c0 : c1;
        const int co = cc < c0 ? cc : cc - c0;
        if (cc > 0) __syncthreads();  // every wave has read the previous chunk
        for (int e = threadIdx.x; e < kTileRows * kTileW * 2; e += 256) {
            const int half = e & 1;
            const int xi = (e >> 1) % kTileW;
            const int r = (e >> 1) / kTileW;
            const int x = wrapn(i0 - 1 + xi, n);
            const int y = wrapn(j0 - 1 + (r & 3), n);
            const int z = wrapn(k0 - 1 + (r >> 2), n);
            const long v = ((long)z * n + y) * n + x;
            st[e] = *reinterpret_cast<const uint4*>(src + v * cs + co + 8 * half);
        }
        __syncthreads();
        for (int t0 = 0; t0 < 27; t0 += 2) {
            const int t = t0 + (g >> 1);
            const bool tap = t < 27;
            const int dk = t / 9, dj = (t / 3) % 3, di = t % 3;  // offsets + 1
            v8bf a[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                a[m] = zero8();
                if (tap) {
                    const int r = (kk + dk) * 4 + (jj + dj);
                    const int xi = m * 16 + row + di;
                    a[m] = __builtin_bit_cast(v8bf, st[(r * kTileW + xi) * 2 + (g & 1)]);
                }
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                v8bf b = zero8();
                if (tap) b = load8(wt + (long)(nt * 16 + row) * Kp + (long)t * cin + cc + 8 * (g & 1));
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    acc[m][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b, acc[m][nt], 0, 0, 0);
            }
        }
    }
    // D[r][c]: lane holds rows 4g + q (voxels of M tile m), column row (channel nt*16 + row)
    const long vrow = ((long)(k0 + kk) * n + (j0 + jj)) * n + i0;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int cout = nt * 16 + row;
        const float bb = bias != nullptr ? bias[cout] : 0.f;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const long vo = vrow + m * 16 + 4 * g + q;
                float y = acc[m][nt][q] + bb;
                if (relu && y < 0.f) y = 0.f;
                if (out_f32)
                    static_cast<float*>(out)[vo * COUT + cout] = y;
                else
                    static_cast<uint16_t*>(out)[vo * COUT + cout] = f2bf(y);
            }
        }
    }
}

// Weight gradient: dW[co][k] += sum_v dZ[v][co] A[v][k] over a chunk of voxels per workgroup.
// One wave per workgroup computes the 16x16 tile (co0.., k0..) with the voxel sum as the MFMA K
// dimension (32 voxels per step); FP32 atomics add the chunk's tile into dw.
template <int TAPS>
__global__ __launch_bounds__(64) void k_conv_wgrad(const uint16_t* __restrict__ in0, int c0,
                                                   const uint16_t* __restrict__ in1, int c1, int n, long V,
                                                   const uint16_t* __restrict__ dz, int cout, int Kp,
                                                   long vchunk, float* __restrict__ dw) {
    const int lane = threadIdx.x;
    const int row = lane & 15;
    const int g = lane >> 4;
    const int co0 = blockIdx.x * 16;
    const int k0 = blockIdx.y * 16;
    const long vb = (long)blockIdx.z * vchunk;
    const long ve = min(V, vb + vchunk);
    const int cin = c0 + c1;
    const int kmax = TAPS * cin;
    // B column k = k0 + row: tap and channel fixed for this lane
    const int kk = k0 + row;
    const bool kin = kk < kmax;
    const int t = kin ? kk / cin : 0;
    const int ci = kin ? kk - t * cin : 0;
    const uint16_t* src = (ci < c0) ? in0 : in1;
    const int cs = (ci < c0) ? c0 : c1;
    const int cc = (ci < c0) ? ci : ci - c0;
    v4f acc = (v4f){0.f, 0.f, 0.f, 0.f};
    for (long vs = vb; vs < ve; vs += 32) {
        v8bf a, b;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const long v = vs + 8 * g + q;
            uint16_t av = 0, bv = 0;
            if (v < ve) {
                av = dz[v * cout + co0 + row];
                if (kin) {
                    const int i = (int)(v % n);
                    const int j = (int)((v / n) % n);
                    const int k = (int)(v / ((long)n * n));
                    bv = src[neighbour<TAPS>(i, j, k, n, t) * cs + cc];
                }
            }
            a[q] = __builtin_bit_cast(__bf16, av);
            b[q] = __builtin_bit_cast(__bf16, bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    // D rows 4g + q = output channel co0 + 4g + q, column row = k0 + row
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int co = co0 + 4 * g + q;
        if (co < cout && k0 + row < Kp) atomicAdd(&dw[(long)co * Kp + k0 + row], acc[q]);
    }
}

}  // namespace unet
}  // namespace fftperm

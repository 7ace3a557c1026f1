// passes_plane.hip -- the x and y transforms of an FP32 apply fused into one pass over whole z
// planes (nx = ny = 256).
//
// The five-sweep apply (x R2C, y, fused z + Green, y inverse, x C2R) moves the half spectrum
// through HBM four times outside the z pass.  A 256 x 129 FP32 half-spectrum plane of one
// component is 264 KB: too big for one LDS, but it fits the register file of a CU's 16 waves
// together with a 139 KB LDS staging area.  So one workgroup takes a (z plane, component) pair and
// runs both transforms of the plane on chip:
//
//   forward: gather the plane's rows from the compact vector (line bitmask), 128 complex x FFTs
//            (rows 2m and 2m+1 as z = a + i b), then in two chunks of 64 x frequencies: stage the
//            chunk's spectra in LDS, form the half-spectrum columns (two-for-one split), 64 y FFTs,
//            stage the result by rows and write it with coalesced row stores;
//   inverse: the same in reverse: read each chunk's rows, 64 inverse y FFTs, stage by rows, take
//            each row pair's x spectrum into registers, then 128 inverse x FFTs and the scatter to
//            the compact vector with the support sums.
//
// HBM traffic per apply falls from 3.8 (FP32: 1.96) GB to the compact vectors, the line bitmask
// and the z pass at 256^3.  FP64 planes (528 KB) do not fit a CU and keep the separate passes.
//
// Every transform is a 256-point FFT by a group of 16 lanes (four lines per wave): lane n holds
// elements n + 16 j in registers, a radix-16 DFT over j, the twiddles W_256^{n k1} (16-byte reads
// of a small LDS table; in registers they would push the kernels past the 128 VGPRs that 16 waves
// per CU allow), one 16 x 16 transpose through LDS and a second radix-16 DFT leave element n + 16 k
// in register k: the same layout as the input, so no reordering pass is needed.
//
// Special frequencies: the x spectra of rows 2m and 2m+1 at kx = 0 and kx = 128 are real and sit
// in Z_m[0] and Z_m[128] (real parts for row 2m, imaginary parts for row 2m+1), so the two real
// columns kx = 0 and kx = 128 travel as one complex column through the y FFT and are split after
// it (forward); the inverse packs the Hermitian parts of the two columns, which gives the real
// parts that the C2R takes of those bins (src/fft/FFTWBackend.cpp:292: the real part of the full
// C2C inverse).
#include "fft_engine.hpp"
#include "fft_static.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace fftperm {

namespace {

constexpr int kPN = 256;            // x and y edge of the planes served
constexpr int kXP = 17;             // pitch of a group's 16 x 16 transpose (complex values)
constexpr int kGroupX = 16 * kXP;   // LDS of one group's transpose
constexpr int kZP = 129;            // forward: row-pair pitch of the chunk's x spectra (128 slots)
constexpr int kOP = 66;             // row pitch of a chunk's 64 columns staged by rows (+2 spare columns)
constexpr int kTwP = 18;            // row pitch of the twiddle table (16-byte rows, no bank conflicts)

// W waves per workgroup = 4 W groups of 16 lanes.  A workgroup owns the whole plane (the LDS stage
// allows one per CU), so W trades latency hiding against registers: 16 waves leave 128 VGPRs a lane
// and the plane's spectra then spill; 8 waves leave 256, enough to keep each phase's loads in
// flight while the previous phase computes.
template <int W>
struct PlaneShape {
    static constexpr int THREADS = 64 * W;
    static constexpr int GROUPS = 4 * W;
    static constexpr int XR = 128 / GROUPS;  // line pairs (x transforms) per group
    static constexpr int YC = 64 / GROUPS;   // columns (y transforms) per group and chunk
    static constexpr int RW = kPN / W;       // rows per wave in the row-wise phases
    static constexpr int MAIN = GROUPS * kGroupX > 128 * kZP ? (GROUPS * kGroupX > kPN * kOP ? GROUPS * kGroupX : kPN * kOP)
                                                             : (128 * kZP > kPN * kOP ? 128 * kZP : kPN * kOP);
    static constexpr size_t LDS = (size_t)(MAIN + 16 * kTwP) * sizeof(cx<float>);
    static_assert(XR * GROUPS == 128 && YC * GROUPS == 64, "wave count must divide the plane");
};

// Orders the wave's LDS writes before its later reads (a wave executes its LDS operations in order;
// the fences keep the compiler from moving them across).
__device__ __forceinline__ void plane_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Twiddle table T[n * kTwP + k1] = W_256^{n k1} (conjugated for the inverse), n, k1 < 16, from
// twg[k] = exp(-2 pi i k / 256); returns after a barrier.
template <int SIGN>
__device__ __forceinline__ void plane_twiddles(cx<float>* T, const cx<float>* __restrict__ twg) {
    if (threadIdx.x < 256) {
        const int n = threadIdx.x >> 4, k1 = threadIdx.x & 15;
        cx<float> t = twg[n * k1];
        if (SIGN > 0) t.y = -t.y;
        T[n * kTwP + k1] = t;
    }
    __syncthreads();
}

// 256-point FFT of the group's line: v[j] = x[n + 16 j] in, v[k] = X[n + 16 k] out
// (X[k1 + 16 k2] = sum_n2 W_16^{n2 k2} W_256^{n2 k1} sum_n1 W_16^{n1 k1} x[n2 + 16 n1]).
// X: the group's transpose area, free again when the call returns; T: the lane's twiddle row.
template <int SIGN>
__device__ __forceinline__ void group_fft256(cx<float> (&v)[16], cx<float>* X, const cx<float>* T, int n) {
    dft_small<float, SIGN, 16>(v);
    // an opaque copy of the row address: the compiler would otherwise hoist the table reads out of
    // the callers' loops and hold the 30 values in registers
    uintptr_t ta = reinterpret_cast<uintptr_t>(T);
    asm volatile("" : "+v"(ta));
    const float4* T4 = reinterpret_cast<const float4*>(ta);
#pragma unroll
    for (int k1 = 0; k1 < 16; k1 += 2) {
        const float4 w = T4[k1 >> 1];
        if (k1 > 0) v[k1] = cmul<float>(v[k1], mkc<float>(w.x, w.y));
        v[k1 + 1] = cmul<float>(v[k1 + 1], mkc<float>(w.z, w.w));
    }
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) X[k1 * kXP + n] = v[k1];
    plane_wave_sync();
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[n2] = X[n * kXP + n2];
    plane_wave_sync();
    dft_small<float, SIGN, 16>(v);
}

// The same transform with the transpose done in place in the group's own LDS column col[r * kOP],
// r = 0..255 (a column only this group reads or writes): X(k1, n) goes to row 16 k1 + (n + k1) % 16,
// a skew that keeps both the writes (k1 fixed) and the reads (n fixed) free of bank conflicts.
template <int SIGN>
__device__ __forceinline__ void group_fft256_col(cx<float> (&v)[16], cx<float>* col, const cx<float>* T, int n) {
    dft_small<float, SIGN, 16>(v);
    uintptr_t ta = reinterpret_cast<uintptr_t>(T);
    asm volatile("" : "+v"(ta));
    const float4* T4 = reinterpret_cast<const float4*>(ta);
#pragma unroll
    for (int k1 = 0; k1 < 16; k1 += 2) {
        const float4 w = T4[k1 >> 1];
        if (k1 > 0) v[k1] = cmul<float>(v[k1], mkc<float>(w.x, w.y));
        v[k1 + 1] = cmul<float>(v[k1 + 1], mkc<float>(w.z, w.w));
    }
    plane_wave_sync();
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) col[(16 * k1 + ((n + k1) & 15)) * kOP] = v[k1];
    plane_wave_sync();
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[n2] = col[(16 * n + ((n2 + n) & 15)) * kOP];
    plane_wave_sync();
    dft_small<float, SIGN, 16>(v);
}

// Support words of x line `line` and the compact index of their first voxel; on(j) / loc(j) for
// voxel n + 16 j of the line (word j / 2, bit n + 16 (j % 2)).
struct PlaneLineBits {
    uint32_t word[8];
    int pre[8];
    __device__ __forceinline__ void load(const LineBits& lb, long line) {
        const uint32_t* wb = lb.bits + line * 8;
#pragma unroll
        for (int u = 0; u < 8; ++u) word[u] = wb[u];
        int acc = lb.lstart[line];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            pre[u] = acc;
            acc += __popc(word[u]);
        }
    }
};

__device__ __forceinline__ cx<float> half_a(cx<float> zk, cx<float> zm) {
    return mkc<float>(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
}
__device__ __forceinline__ cx<float> half_b(cx<float> zk, cx<float> zm) {
    return mkc<float>(0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x));
}
// Z = A + i B (k <= N / 2) and conj(A) + i conj(B) (k > N / 2)
__device__ __forceinline__ cx<float> pair_lower(cx<float> a, cx<float> b) { return mkc<float>(a.x - b.y, a.y + b.x); }
__device__ __forceinline__ cx<float> pair_upper(cx<float> a, cx<float> b) { return mkc<float>(a.x + b.y, b.x - a.y); }

// Workgroup (z plane, component c) of a grid of 3 nz workgroups, dealt so that the three
// components of a plane run on one XCD (they read the same compact run and bitmask).  Group G =
// 4 wave + g transforms line pairs m = G + GROUPS t (rows 2m, 2m+1) and columns cs = G + GROUPS u.
template <int W>
__global__ __launch_bounds__(PlaneShape<W>::THREADS) void k_xyfwd_plane(XSource src, cx<float>* __restrict__ S,
                                                                          SpecLayout SL,
                                                                          const cx<float>* __restrict__ twg,
                                                                          const int* skip) {
    using PS = PlaneShape<W>;
    constexpr int XR = PS::XR, YC = PS::YC, GR = PS::GROUPS;
    if (skip != nullptr && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    cx<float>* lds = reinterpret_cast<cx<float>*>(smem_raw);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const int G = 4 * wave + g;
    const int q = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x));
    const int z = q / 3, c = q - 3 * (q / 3);
    // an FP32 source has its mean subtracted in FP32
    const float mc = src.mean != nullptr ? (float)src.mean[c] : 0.f;
    cx<float>* Xg = lds + G * kGroupX;
    const long line0 = (long)z * kPN;
    const uint32_t blo = 1u << n, bhi = 1u << (n + 16);
    const float* __restrict__ src32 = src.compact32;
    const unsigned last = (unsigned)(src.lines.nb - 1);

    // gathers of the rows of round t (line pair G + GR t) into gv[t & 1]: every lane loads, so the
    // loads issue back to back.  An off-support voxel reads the entry of the next support voxel of
    // its line (clamped to the last one): a line the neighbouring lanes fetch anyway, where a fixed
    // dummy entry would be one cache line that every CU hits.  32-bit byte offsets (12 NB < 2^32)
    // keep each address in one VGPR.
    float gv[2][2][16];
    auto gather = [&](int t) {
        const int m = G + GR * t;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            PlaneLineBits lb;
            lb.load(src.lines, line0 + 2 * m + p);
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const uint32_t w = lb.word[j >> 1];
                const uint32_t bit = (j & 1) ? bhi : blo;
                const bool on = (w & bit) != 0u;
                const unsigned loc = min((unsigned)(lb.pre[j >> 1] + __popc(w & (bit - 1u))), last);
                const float raw = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(src32) +
                                                                  4u * (3u * loc + (unsigned)c));
                gv[t & 1][p][j] = on ? raw - mc : 0.f;
            }
        }
    };
    // ---- x transforms, the gathers two rounds ahead
    gather(0);
    if (XR > 1) gather(1);
    cx<float>* T = lds + PS::MAIN;
    plane_twiddles<-1>(T, twg);
    const cx<float>* Tn = T + n * kTwP;
    cx<float> zr[XR][16];
#pragma unroll
    for (int t = 0; t < XR; ++t) {
        cx<float> v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = mkc<float>(gv[t & 1][0][j], gv[t & 1][1][j]);
        if (t + 2 < XR) gather(t + 2);
        group_fft256<-1>(v, Xg, Tn, n);
#pragma unroll
        for (int k = 0; k < 16; ++k) zr[t][k] = v[k];
    }
    __syncthreads();

    // ---- two chunks of x frequencies: chunk 0 = kx 1..63 and the packed (0, 128) column, chunk 1 =
    // kx 64..127.  Slots of Z_m in chunk 0: index i < 64 -> i, 128 -> 64, 193..255 -> i - 128; in
    // chunk 1: 64..127 -> i - 64, 129..192 -> i - 65.
    cx<float>* Zb = lds;
    cx<float>* Ob = lds;
    cx<float>* Sc = S + (long)c * SL.nc + (long)z * SL.pz;
    const bool odd = (n & 1) != 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int t = 0; t < XR; ++t) {
            cx<float>* row = Zb + (G + GR * t) * kZP;
            if (h == 0) {
#pragma unroll
                for (int k = 0; k < 4; ++k) row[n + 16 * k] = zr[t][k];
                if (n == 0) row[64] = zr[t][8];
#pragma unroll
                for (int k = 12; k < 16; ++k)
                    if (k != 12 || n != 0) row[n + 16 * k - 128] = zr[t][k];
            } else {
#pragma unroll
                for (int k = 4; k < 8; ++k) row[n + 16 * k - 64] = zr[t][k];
#pragma unroll
                for (int k = 8; k < 12; ++k)
                    if (k != 8 || n != 0) row[n + 16 * k - 65] = zr[t][k];
                if (n == 0) row[127] = zr[t][12];
            }
        }
        __syncthreads();
        // columns cs of the chunk at rows y = n + 16 j (row pair m = n / 2 + 8 j, half n % 2)
        cx<float> v[YC][16];
#pragma unroll
        for (int u = 0; u < YC; ++u) {
            const int cs = G + GR * u;
            const bool packed = h == 0 && cs == 0;
            const int kx = h == 0 ? cs : cs + 64;
            const int sa = packed ? 0 : (h == 0 ? kx : kx - 64);
            const int sb = packed ? 64 : (h == 0 ? 128 - kx : 191 - kx);
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const cx<float>* zrow = Zb + ((n >> 1) + 8 * j) * kZP;
                const cx<float> zk = zrow[sa], zm = zrow[sb];
                // half_a(zk, zm) on even rows, half_b on odd rows, (zk, zm) parts for the packed
                // column: operand selects instead of branches
                const float p1 = odd ? zk.y : zk.x, q1 = odd ? zm.y : zm.x;
                const float s1 = odd ? zm.x : zk.y, s2 = odd ? zk.x : zm.y;
                const float rx = 0.5f * (p1 + q1), ry = 0.5f * (s1 - s2);
                v[u][j] = mkc<float>(packed ? p1 : rx, packed ? q1 : ry);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < YC; ++u) group_fft256<-1>(v[u], Xg, Tn, n);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < YC; ++u)
#pragma unroll
            for (int k = 0; k < 16; ++k) Ob[(n + 16 * k) * kOP + G + GR * u] = v[u][k];
        __syncthreads();
        // rows out: wave w writes rows w + W it, lane = column of the chunk
#pragma unroll 8
        for (int it = 0; it < PS::RW; ++it) {
            const int ky = wave + W * it;
            const cx<float> val = Ob[ky * kOP + lane];
            cx<float>* dst = Sc + (long)ky * SL.rp;
            if (h == 0 && lane == 0) {
                const cx<float> ym = Ob[((kPN - ky) & (kPN - 1)) * kOP];
                dst[0] = half_a(val, ym);
                dst[kPN / 2] = half_b(val, ym);
            } else {
                dst[h == 0 ? lane : lane + 64] = val;
            }
        }
        __syncthreads();
    }
}

template <int W>
__global__ __launch_bounds__(PlaneShape<W>::THREADS) void k_yxinv_plane(const cx<float>* __restrict__ S, XSink sink,
                                                                          SpecLayout SL,
                                                                          const cx<float>* __restrict__ twg,
                                                                          float inv_n, const int* skip) {
    using PS = PlaneShape<W>;
    constexpr int XR = PS::XR, YC = PS::YC, GR = PS::GROUPS, RW = PS::RW;
    if (skip != nullptr && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    cx<float>* lds = reinterpret_cast<cx<float>*>(smem_raw);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const int G = 4 * wave + g;
    const int q = __builtin_amdgcn_readfirstlane((int)xcd_tile(blockIdx.x, gridDim.x));
    const int z = q / 3, c = q - 3 * (q / 3);
    cx<float>* Xg = lds + G * kGroupX;
    const cx<float>* Sc = S + (long)c * SL.nc + (long)z * SL.pz;
    cx<float>* Ob = lds;

    // ---- the row loads of a chunk are issued together (row ky = wave + W it: chunk 0 column
    // kx = lane, chunk 1 kx = lane + 64; in chunk 0 lanes < RW also fetch kx = N/2 of row
    // wave + W lane): chunk 0's before the twiddle table, chunk 1's during chunk 0's transforms
    cx<float> pre[RW];
    auto rows_in = [&](int h) {
#pragma unroll
        for (int it = 0; it < RW; ++it) pre[it] = Sc[(long)(wave + W * it) * SL.rp + lane + 64 * h];
    };
    rows_in(0);
    cx<float> nyq = mkc<float>(0.f, 0.f);
    if (lane < RW) nyq = Sc[(long)(wave + W * lane) * SL.rp + kPN / 2];
    cx<float>* T = lds + PS::MAIN;
    plane_twiddles<+1>(T, twg);
    const cx<float>* Tn = T + n * kTwP;

    // ---- two chunks: y inverse of 64 columns each, then the row pairs' x spectra into registers:
    // zr[t][j] = Z_m[n + 16 j] of line pair m = G + GR t (every element is assigned by one of the
    // two chunks)
    cx<float> zr[XR][16];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int it = 0; it < RW; ++it) Ob[(wave + W * it) * kOP + lane] = pre[it];
        if (h == 0) {
            // chunk 0: kx = N/2 in spare column 64, then the Hermitian parts of the columns kx = 0
            // and kx = N/2 packed as H0 + i H128 into spare column 65, which column 0 reads
            if (lane < RW) Ob[(wave + W * lane) * kOP + 64] = nyq;
            __syncthreads();
            if (threadIdx.x < kPN) {
                const int ky = threadIdx.x, kym = (kPN - ky) & (kPN - 1);
                const cx<float> h0 = half_a(Ob[ky * kOP], Ob[kym * kOP]);
                const cx<float> h1 = half_a(Ob[ky * kOP + 64], Ob[kym * kOP + 64]);
                Ob[ky * kOP + 65] = pair_lower(h0, h1);
            }
        }
        __syncthreads();
        if (h == 0) rows_in(1);
#pragma unroll
        for (int u = 0; u < YC; ++u) {
            const int cs = G + GR * u;
            const int rc = (h == 0 && cs == 0) ? 65 : cs;
            cx<float> v[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = Ob[(n + 16 * j) * kOP + rc];
            // the column is this group's alone from here: transpose in place, result back by rows
            group_fft256_col<+1>(v, Ob + cs, Tn, n);
#pragma unroll
            for (int k = 0; k < 16; ++k) Ob[(n + 16 * k) * kOP + cs] = v[k];
        }
        __syncthreads();
        // x spectra of the row pairs from this chunk (rows 2m: A, 2m+1: B; kk = k or N - k)
#pragma unroll
        for (int t = 0; t < XR; ++t) {
            const int m = G + GR * t;
            const cx<float>* ra = Ob + (2 * m) * kOP;
            const cx<float>* rb = Ob + (2 * m + 1) * kOP;
            if (h == 0) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {  // k = kk = n + 16 j (n = j = 0: the packed column)
                    const cx<float> a = ra[n + 16 * j], b = rb[n + 16 * j];
                    zr[t][j] = (j == 0 && n == 0) ? mkc<float>(a.x, b.x) : pair_lower(a, b);
                }
                {  // k = 128 (n = 0): the packed column's imaginary parts
                    const cx<float> a = ra[0], b = rb[0];
                    if (n == 0) zr[t][8] = mkc<float>(a.y, b.y);
                }
#pragma unroll
                for (int j = 12; j < 16; ++j) {  // k = n + 16 j > 128, kk = 256 - k (j = 12, n = 0: chunk 1)
                    const int kk = kPN - n - 16 * j;
                    const cx<float> a = ra[kk], b = rb[kk];
                    if (j != 12 || n != 0) zr[t][j] = pair_upper(a, b);
                }
            } else {
#pragma unroll
                for (int j = 4; j < 8; ++j) {  // k = kk = n + 16 j in 64..127
                    const cx<float> a = ra[n + 16 * j - 64], b = rb[n + 16 * j - 64];
                    zr[t][j] = pair_lower(a, b);
                }
#pragma unroll
                for (int j = 8; j < 13; ++j) {  // k = n + 16 j > 128 with kk = 256 - k in 64..127
                    const int kk = kPN - n - 16 * j;   // j = 8, n = 0 (k = 128) and j = 12, n > 0: chunk 0
                    const bool mine = j == 8 ? n != 0 : (j == 12 ? n == 0 : true);
                    const int col = mine ? kk - 64 : 64;  // column 64 is a spare: read, unused
                    const cx<float> a = ra[col], b = rb[col];
                    if (mine) zr[t][j] = pair_upper(a, b);
                }
            }
        }
        __syncthreads();
    }

    // ---- x inverse of the row pairs and the scatter to the compact vector
    const long line0 = (long)z * kPN;
    const uint32_t blo = 1u << n, bhi = 1u << (n + 16);
    float* __restrict__ dst32 = sink.compact32;
    double acc[1] = {0.0};
#pragma unroll
    for (int t = 0; t < XR; ++t) {
        const int m = G + GR * t;
        PlaneLineBits lb[2];
#pragma unroll
        for (int p = 0; p < 2; ++p) lb[p].load(sink.lines, line0 + 2 * m + p);
        cx<float> v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = zr[t][j];
        group_fft256<+1>(v, Xg, Tn, n);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const uint32_t w = lb[p].word[k >> 1];
                const uint32_t bit = (k & 1) ? bhi : blo;
                if ((w & bit) != 0u) {
                    const int loc = lb[p].pre[k >> 1] + __popc(w & (bit - 1u));
                    const float val = (p == 0 ? v[k].x : v[k].y) * inv_n;
                    dst32[(unsigned)(3 * loc + c)] = val;
                    acc[0] += (double)val;
                }
            }
        }
    }
    block_sum<1>(acc, sink.partials != nullptr ? sink.partials + 3L * z + c : nullptr);
}

}  // namespace

bool plane_pass_shape_ok(const Geometry3& g) {
    // the compact-vector byte offsets 4 (3 l + c) are 32-bit
    return g.nx == kPN && g.ny == kPN && 12L * g.N < (1L << 32) &&
           lds_limit() >= std::max(PlaneShape<8>::LDS, PlaneShape<16>::LDS) + 1024;
}

bool plane_src_ok(const XSource& src) {
    const bool no_f0 = src.a_fluid[0] == 0.0 && src.a_fluid[1] == 0.0 && src.a_fluid[2] == 0.0 &&
                       src.a_support[0] == 0.0 && src.a_support[1] == 0.0 && src.a_support[2] == 0.0;
    return src.field == nullptr && src.compact == nullptr && src.compact32 != nullptr && src.lines.bits != nullptr &&
           no_f0;
}

bool plane_sink_ok(const XSink& sink) {
    return sink.field == nullptr && sink.compact == nullptr && sink.compact32 != nullptr && sink.lines.bits != nullptr;
}

long launch_xyfwd_plane(const XSource& src, cx<float>* S, const Geometry3& g, const cx<float>* twx, int waves,
                        const int* skip, hipStream_t st) {
    if (!plane_pass_shape_ok(g) || !plane_src_ok(src)) return -1;
    const long blocks = 3L * g.nz;
    if (waves == 16)
        hipLaunchKernelGGL(k_xyfwd_plane<16>, dim3((unsigned)blocks), dim3(PlaneShape<16>::THREADS),
                           PlaneShape<16>::LDS, st, src, S, spec_layout(g), twx, skip);
    else
        hipLaunchKernelGGL(k_xyfwd_plane<8>, dim3((unsigned)blocks), dim3(PlaneShape<8>::THREADS),
                           PlaneShape<8>::LDS, st, src, S, spec_layout(g), twx, skip);
    return blocks;
}

long launch_yxinv_plane(const cx<float>* S, const XSink& sink, const Geometry3& g, const cx<float>* twx, int waves,
                        const int* skip, hipStream_t st) {
    if (!plane_pass_shape_ok(g) || !plane_sink_ok(sink)) return -1;
    const long blocks = 3L * g.nz;
    const float inv_n = (float)(1.0 / (double)g.N);
    if (waves == 16)
        hipLaunchKernelGGL(k_yxinv_plane<16>, dim3((unsigned)blocks), dim3(PlaneShape<16>::THREADS),
                           PlaneShape<16>::LDS, st, S, sink, spec_layout(g), twx, inv_n, skip);
    else
        hipLaunchKernelGGL(k_yxinv_plane<8>, dim3((unsigned)blocks), dim3(PlaneShape<8>::THREADS),
                           PlaneShape<8>::LDS, st, S, sink, spec_layout(g), twx, inv_n, skip);
    return g.nz;
}

}  // namespace fftperm

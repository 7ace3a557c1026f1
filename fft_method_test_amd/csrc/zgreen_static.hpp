// zgreen_static.hpp -- compile-time specialised fused z passes (forward z FFT, symmetric Green
// multiply, backward z FFT in one HBM round trip) and their launchers, one set per column length.
// Included by kernels_fft.hip (power-of-two lengths) and zgreen_mixed.hip (mixed-radix lengths),
// so that the two sets compile in parallel.
#pragma once

#include "fft_engine.hpp"
#include "fft_static.hpp"
#include "kernels.hpp"
#include "zgreen_reg.hpp"

#include <hip/hip_ext.h>

#include <algorithm>

namespace fftperm {

namespace {
// Launches a fused z-pass kernel; when the caller armed zpass_events(), the launch itself records
// them at the kernel's start and end (hipExtLaunchKernelGGL), so the timing holds the kernel alone,
// without the launch gaps that events recorded around it would add.  The events are consumed.
template <typename... KArgs, typename... Args>
void zlaunch(void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t lds, hipStream_t st, Args... args) {
    KernelEvents& ev = zpass_events();
    if (ev.start == nullptr && ev.stop == nullptr) {
        hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
    } else {
        hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)lds, st, ev.start, ev.stop, 0, args...);
        ev = KernelEvents{};
    }
}
}  // namespace

namespace {

template <typename R, int N>
struct ZShape {
    static constexpr int T = threads_per_line<N>();  // threads per line
    static constexpr int LD = N + 1;                  // LDS pitch (complex values)
    // three components of ZL columns in LDS, at most ~100 KB and 1024 threads
    static constexpr int zl_raw() {
        int z = 16;
        while (z > 1 && ((size_t)(3 * z * LD + N) * sizeof(cx<R>) > 102400u || 3 * z * T > 1024)) z /= 2;
        return z;
    }
    static constexpr int ZL = zl_raw();
};

// Hermitian weight of column kx of the half spectrum in a sum over the full spectrum (0 on the
// row padding, kx > nx / 2).
__device__ __forceinline__ double herm_weight(int kx, int nx) {
    return (2 * kx > nx) ? 0.0 : ((kx == 0 || 2 * kx == nx) ? 1.0 : 2.0);
}

// Symmetric Green multiply g = T f of one frequency (src/bignonnet/GreenOperator.cpp:84-92 with
// G_sym; T = {xx, xy, xz, yy, yz, zz}) and its quadratic form conj(f) . g in FP64, both in a fixed
// operation order (g_i = T_iz f_z + (T_iy f_y + T_ix f_x), products fused in turn), so that every
// z-pass kernel rounds them alike.
template <typename R>
__device__ __forceinline__ void green_mul(R txx, R txy, R txz, R tyy, R tyz, R tzz, cx<R> fx, cx<R> fy, cx<R> fz,
                                          cx<R>& gx, cx<R>& gy, cx<R>& gz) {
#pragma clang fp contract(off)
    gx = mkc<R>(fmar<R>(txz, fz.x, fmar<R>(txy, fy.x, txx * fx.x)), fmar<R>(txz, fz.y, fmar<R>(txy, fy.y, txx * fx.y)));
    gy = mkc<R>(fmar<R>(tyz, fz.x, fmar<R>(tyy, fy.x, txy * fx.x)), fmar<R>(tyz, fz.y, fmar<R>(tyy, fy.y, txy * fx.y)));
    gz = mkc<R>(fmar<R>(tzz, fz.x, fmar<R>(tyz, fy.x, txz * fx.x)), fmar<R>(tzz, fz.y, fmar<R>(tyz, fy.y, txz * fx.y)));
}
template <typename R>
__device__ __forceinline__ double green_quad(cx<R> fx, cx<R> fy, cx<R> fz, cx<R> gx, cx<R> gy, cx<R> gz) {
#pragma clang fp contract(off)
    double q = (double)fx.x * (double)gx.x;
    q = fmar<double>((double)fx.y, (double)gx.y, q);
    q = fmar<double>((double)fy.x, (double)gy.x, q);
    q = fmar<double>((double)fy.y, (double)gy.y, q);
    q = fmar<double>((double)fz.x, (double)gz.x, q);
    q = fmar<double>((double)fz.y, (double)gz.y, q);
    return q;
}
template <typename R>
__device__ __forceinline__ void green_quad_add(double& acc, double w, cx<R> fx, cx<R> fy, cx<R> fz, cx<R> gx,
                                               cx<R> gy, cx<R> gz) {
#pragma clang fp contract(off)
    acc = acc + w * green_quad<R>(fx, fy, fz, gx, gy, gz);
}

// Fused z pass: lines 0..3*ZL-1 of the LDS buffer hold components x | y | z of ZL columns.
// OS = threads per butterfly thread (1 or 2): extra threads only help the memory phases
template <typename R, int N, int ZL, int OS>
constexpr int zgreen_threads() {
    return OS * 3 * ZL * ZShape<R, N>::T;
}

// Green planes a z column reads: FP32 tensors are read for the planes kz <= N / 2 only, and plane
// N - kz takes plane kz with the xz and yz components negated (G(kx, ky, -kz) = Rz G(kx, ky, kz) Rz),
// so a launch moves half of the tensor and each value is read once.  The stored G_sym obeys that
// mirror to 2e-8 .. 8e-8 of max |G| (the truncated image sum is symmetric under the full inversion
// only), below the FP32 rounding of the stored values; FP64 reads every plane.
template <typename R, int N>
constexpr bool zgreen_mirror() {
    return sizeof(R) == 4;
}
template <typename R, int N>
constexpr int zgreen_planes() {
    return zgreen_mirror<R, N>() ? N / 2 + 1 : N;
}

// The six Green components of plane p of a z column whose first tensor element is at gc.
// STREAM: non-temporal loads.
template <typename R, bool STREAM>
__device__ __forceinline__ void load_green_plane(const R* __restrict__ G6, long gc, int p, long pitch, R (&g)[6]) {
    const R* gp = G6 + gc + 6L * p * pitch;
    if constexpr (STREAM) {
#pragma unroll
        for (int q = 0; q < 6; ++q) g[q] = __builtin_nontemporal_load(gp + q);
    } else {
#pragma unroll
        for (int q = 0; q < 6; ++q) g[q] = gp[q];
    }
}

// Symmetric Green multiply (src/bignonnet/GreenOperator.cpp:84-92 with G_sym) of frequency e of
// the column whose components are at bx, by, bz (LDS), with the tensor t (xz and yz negated when
// flip: the mirrored plane), and the quadratic form conj(f) . G f of the forward spectrum added to
// q when asked for.
template <typename R>
__device__ __forceinline__ void green_step(cx<R>* bx, cx<R>* by, cx<R>* bz, int e, const R (&t)[6], bool flip,
                                           bool want_q, double& q) {
    const R txz = flip ? -t[2] : t[2], tyz = flip ? -t[4] : t[4];
    const cx<R> fx = bx[e], fy = by[e], fz = bz[e];
    cx<R> gx, gy, gz;
    green_mul<R>(t[0], t[1], txz, t[3], tyz, t[5], fx, fy, fz, gx, gy, gz);
    bx[e] = gx;
    by[e] = gy;
    bz[e] = gz;
    if (want_q) green_quad_add<R>(q, 1.0, fx, fy, fz, gx, gy, gz);
}

// The Green step of a thread's planes p = threadIdx.x / ZL + m * MT (m < KP) of column mci: the
// frequency p and, with the mirror, N - p.
template <typename R, int N, int ZL, int MT, int KP>
__device__ __forceinline__ void green_tile_step(cx<R>* buf, int LD, int mci, bool mown, const R (&gt)[KP][6],
                                                bool want_q, double& q) {
    constexpr int NP = zgreen_planes<R, N>();
    cx<R>* bx = buf + (0 * ZL + mci) * LD;
    cx<R>* by = buf + (1 * ZL + mci) * LD;
    cx<R>* bz = buf + (2 * ZL + mci) * LD;
#pragma unroll
    for (int m = 0; m < KP; ++m) {
        const int p = threadIdx.x / ZL + m * MT;
        if (!mown || p >= NP) continue;
        green_step<R>(bx, by, bz, p, gt[m], false, want_q, q);
        if (zgreen_mirror<R, N>() && p != 0 && 2 * p != N) green_step<R>(bx, by, bz, N - p, gt[m], true, want_q, q);
    }
}

// STREAM: the Green tensor is read with non-temporal loads, so that it does not displace the
// spectra from the last-level cache (zpass_streams_green).
template <typename R, int N, int ZL, int OS, bool STREAM>
__global__ __launch_bounds__((zgreen_threads<R, N, ZL, OS>())) void k_zgreen_s(cx<R>* __restrict__ S,
                                                                                 const R* __restrict__ G6, long ncols,
                                                                                 long pitch, long Nc, int nx, int rp,
                                                                                 const cx<R>* __restrict__ twg,
                                                                                 double* __restrict__ qpart,
                                                                                 const int* skip) {
    if (skip != nullptr && *skip) return;
    using SH = ZShape<R, N>;
    constexpr int LD = SH::LD, THREADS = zgreen_threads<R, N, ZL, OS>(), L3 = 3 * ZL;
    constexpr int EST = THREADS / L3;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long col0 = xcd_tile(blockIdx.x, gridDim.x) * ZL;
    const int ncl = (int)min((long)ZL, ncols - col0);
    // loads: thread t serves line t % L3 = comp * ZL + ci, elements t / L3 + k * EST
    const int lt = threadIdx.x % L3;
    const int comp = lt / ZL;
    const int ci = lt % ZL;
    const bool own = ci < ncl;
    constexpr int KL = (N + EST - 1) / EST;  // a thread's last slot idles when EST does not divide N
    constexpr bool EXACT = N % EST == 0;
    const cx<R>* Sc = S + comp * Nc + col0 + ci;
    cx<R> ld[KL];
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int e = threadIdx.x / L3 + k * EST;
        ld[k] = mkc<R>((R)0, (R)0);
        if (own && (EXACT || e < N)) ld[k] = Sc[(long)e * pitch];
    }
    // the Green tensor of the frequencies this thread multiplies, fetched now so that its latency
    // overlaps the column loads and the forward transform
    constexpr int MT = THREADS / ZL;
    constexpr int NP = zgreen_planes<R, N>();
    constexpr int KP = (NP + MT - 1) / MT;
    const int mci = threadIdx.x % ZL;
    const bool mown = mci < ncl;
    R gt[KP][6];
    {
        const long gc = 6L * (col0 + mci);
#pragma unroll
        for (int m = 0; m < KP; ++m) {
            const int p = threadIdx.x / ZL + m * MT;
#pragma unroll
            for (int q = 0; q < 6; ++q) gt[m][q] = (R)0;
            if (mown && p < NP) load_green_plane<R, STREAM>(G6, gc, p, pitch, gt[m]);
        }
    }
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int e = threadIdx.x / L3 + k * EST;
        if (EXACT || e < N) buf[lt * LD + e] = ld[k];
    }
    __syncthreads();
    sfft_lines<R, -1, N, L3, SH::T, LD>(buf, tw, L3);
    // symmetric Green multiply and the quadratic form of the forward spectrum when asked for
    {
        double q[1] = {0.0};
        green_tile_step<R, N, ZL, MT, KP>(buf, LD, mci, mown, gt, qpart != nullptr, q[0]);
        if (qpart != nullptr) {
            q[0] *= herm_weight((int)((col0 + mci) % rp), nx);
            block_sum<1>(q, qpart + blockIdx.x);
        }
    }
    __syncthreads();
    sfft_lines<R, +1, N, L3, SH::T, LD>(buf, tw, L3);
    cx<R>* Sw = S + comp * Nc + col0 + ci;
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int e = threadIdx.x / L3 + k * EST;
        if (own && (EXACT || e < N)) Sw[(long)e * pitch] = buf[lt * LD + e];
    }
}

// Several right-hand sides per launch (the multi-RHS solve's one-launch z pass): the tile's Green
// tensor stays in registers across the sides, read once per launch.  A separate kernel: holding
// the tensor across the loop raises the register count (72 -> 118 VGPRs at 256, FP32, 8 x 2), which
// would cost the single-side kernel half its occupancy.
template <typename R, int N, int ZL, int OS>
__global__ __launch_bounds__((zgreen_threads<R, N, ZL, OS>())) void k_zgreen_sb(cx<R>* __restrict__ S,
                                                                                  const R* __restrict__ G6, long ncols,
                                                                                  long pitch, long Nc, int nx, int rp,
                                                                                  const cx<R>* __restrict__ twg,
                                                                                  double* __restrict__ qpart,
                                                                                  const int* skip, ZBatch zb) {
    using SH = ZShape<R, N>;
    constexpr int LD = SH::LD, THREADS = zgreen_threads<R, N, ZL, OS>(), L3 = 3 * ZL;
    constexpr int EST = THREADS / L3;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    const long col0 = xcd_tile(blockIdx.x, gridDim.x) * ZL;
    const int ncl = (int)min((long)ZL, ncols - col0);
    // loads: thread t serves line t % L3 = comp * ZL + ci, elements t / L3 + k * EST
    const int lt = threadIdx.x % L3;
    const int comp = lt / ZL;
    const int ci = lt % ZL;
    const bool own = ci < ncl;
    constexpr int KL = (N + EST - 1) / EST;  // a thread's last slot idles when EST does not divide N
    constexpr bool EXACT = N % EST == 0;
    constexpr int MT = THREADS / ZL;
    constexpr int NP = zgreen_planes<R, N>();
    constexpr int KP = (NP + MT - 1) / MT;
    const int mci = threadIdx.x % ZL;
    const bool mown = mci < ncl;
    // the Green tensor of the frequencies this thread multiplies stays in registers for every
    // right-hand side of the launch: one read of G per launch
    bool gloaded = false;
    R gt[KP][6];
    for (int b = 0; b < zb.nrhs; ++b) {
        if (skip != nullptr && skip[b * zb.skip_stride]) continue;
        cx<R>* Sb = S + b * zb.rhs_stride;
        const cx<R>* Sc = Sb + comp * Nc + col0 + ci;
        cx<R> ld[KL];
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const int e = threadIdx.x / L3 + k * EST;
            ld[k] = mkc<R>((R)0, (R)0);
            if (own && (EXACT || e < N)) ld[k] = Sc[(long)e * pitch];
        }
        if (!gloaded) {
            // fetched after the first side's column loads are issued, so that its latency overlaps
            // them and the forward transform; the twiddles go to LDS meanwhile
            const long gc = 6L * (col0 + mci);
#pragma unroll
            for (int m = 0; m < KP; ++m) {
                const int p = threadIdx.x / ZL + m * MT;
#pragma unroll
                for (int q = 0; q < 6; ++q) gt[m][q] = (R)0;
                if (mown && p < NP) load_green_plane<R, false>(G6, gc, p, pitch, gt[m]);
            }
            for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
            gloaded = true;
        } else {
            __syncthreads();  // the previous side's stores have read the buffer
        }
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const int e = threadIdx.x / L3 + k * EST;
            if (EXACT || e < N) buf[lt * LD + e] = ld[k];
        }
        __syncthreads();
        sfft_lines<R, -1, N, L3, SH::T, LD>(buf, tw, L3);
        // symmetric Green multiply and the quadratic form of the forward spectrum when asked for
        {
            double q[1] = {0.0};
            green_tile_step<R, N, ZL, MT, KP>(buf, LD, mci, mown, gt, qpart != nullptr, q[0]);
            if (qpart != nullptr) {
                q[0] *= herm_weight((int)((col0 + mci) % rp), nx);
                block_sum<1>(q, qpart + b * zb.qstride + blockIdx.x);
            }
        }
        __syncthreads();
        sfft_lines<R, +1, N, L3, SH::T, LD>(buf, tw, L3);
        cx<R>* Sw = Sb + comp * Nc + col0 + ci;
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const int e = threadIdx.x / L3 + k * EST;
            if (own && (EXACT || e < N)) Sw[(long)e * pitch] = buf[lt * LD + e];
        }
    }
}

// Component-sequential variant: the LDS holds one component of the ZL-column tile at a time
// (ZL lines instead of 3 ZL), so wide tiles -- rows of 16 columns (128 B in FP32, 256 B in FP64)
// -- fit several workgroups per CU.  Every thread keeps its share of all three components in
// registers across the six transforms (three forward, three inverse) and does the Green
// multiply on them; the tensor is read at the multiply.
template <int N, int ZL, int OS>
constexpr int zgreen_c_threads() {
    return OS * ZL * threads_per_line<N>();
}

template <typename R, int N, int ZL, int OS>
__global__ __launch_bounds__((zgreen_c_threads<N, ZL, OS>())) void k_zgreen_c(cx<R>* __restrict__ S,
                                                                               const R* __restrict__ G6, long ncols,
                                                                               long pitch, long Nc, int nx, int rp,
                                                                               const cx<R>* __restrict__ twg,
                                                                               double* __restrict__ qpart,
                                                                               const int* skip) {
    if (skip != nullptr && *skip) return;
    constexpr int T = threads_per_line<N>(), LD = N + 1, THREADS = zgreen_c_threads<N, ZL, OS>();
    constexpr int EST = THREADS / ZL;  // element stride of a thread within its column
    constexpr int KL = N / EST;        // elements per thread per component
    static_assert(THREADS % ZL == 0 && N % EST == 0, "column elements split evenly over the threads");
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long col0 = xcd_tile(blockIdx.x, gridDim.x) * ZL;
    const int ncl = (int)min((long)ZL, ncols - col0);
    const int ci = threadIdx.x % ZL;
    const int e0 = threadIdx.x / ZL;
    const bool own = ci < ncl;
    const long colg = col0 + ci;
    cx<R> f[3][KL];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            f[c][k] = mkc<R>((R)0, (R)0);
            if (own) f[c][k] = S[c * Nc + colg + (long)(e0 + k * EST) * pitch];
        }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (c > 0) __syncthreads();  // the previous component has been read back
#pragma unroll
        for (int k = 0; k < KL; ++k) buf[ci * LD + e0 + k * EST] = f[c][k];
        __syncthreads();
        sfft_lines<R, -1, N, ZL, T, LD>(buf, tw, ZL);
#pragma unroll
        for (int k = 0; k < KL; ++k) f[c][k] = buf[ci * LD + e0 + k * EST];
    }
    // symmetric Green multiply (src/bignonnet/GreenOperator.cpp:84-92 with G_sym) and the quadratic
    // form conj(f) . G f of the forward spectrum when asked for
    double q[1] = {0.0};
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const long e = e0 + k * EST;
        R t6[6];
        const R* gp = G6 + 6L * (colg + e * pitch);
#pragma unroll
        for (int m = 0; m < 6; ++m) t6[m] = own ? gp[m] : (R)0;
        const R txx = t6[0], txy = t6[1], txz = t6[2], tyy = t6[3], tyz = t6[4], tzz = t6[5];
        const cx<R> fx = f[0][k], fy = f[1][k], fz = f[2][k];
        cx<R> gx, gy, gz;
        green_mul<R>(txx, txy, txz, tyy, tyz, tzz, fx, fy, fz, gx, gy, gz);
        if (qpart != nullptr)
            green_quad_add<R>(q[0], 1.0, fx, fy, fz, gx, gy, gz);
        f[0][k] = gx;
        f[1][k] = gy;
        f[2][k] = gz;
        // two frequencies per scheduling group: the tensor loads are not all hoisted to the top,
        // which would hold 6 KL tensor values next to the 3 KL spectrum values
        if (k % 2 == 1) __builtin_amdgcn_sched_barrier(0);
    }
    if (qpart != nullptr) {
        q[0] = own ? q[0] * herm_weight((int)(colg % rp), nx) : 0.0;
        block_sum<1>(q, qpart + blockIdx.x);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KL; ++k) buf[ci * LD + e0 + k * EST] = f[c][k];
        __syncthreads();
        sfft_lines<R, +1, N, ZL, T, LD>(buf, tw, ZL);
        cx<R>* Sc = S + c * Nc + colg;
#pragma unroll
        for (int k = 0; k < KL; ++k)
            if (own) Sc[(long)(e0 + k * EST) * pitch] = buf[ci * LD + e0 + k * EST];
    }
}

template <typename R, int N, int ZL, int OS>
long launch_zgreen_c(cx<R>* S, const R* G6, long ncols, const Geometry3& g, const cx<R>* tw, double* qpart,
                     const int* skip, hipStream_t st) {
    constexpr int threads = zgreen_c_threads<N, ZL, OS>();
    constexpr size_t lds = (size_t)(N + ZL * (N + 1)) * sizeof(cx<R>);
    if constexpr (threads > 1024 || threads < 64 || lds > 100u * 1024u || N % (OS * threads_per_line<N>()) != 0) {
        return -1;
    } else {
        if (lds > lds_limit()) return -1;
        const long blocks = (ncols + ZL - 1) / ZL;
        zlaunch((k_zgreen_c<R, N, ZL, OS>), dim3((unsigned)blocks), dim3(threads), lds, st, S, G6, ncols,
                           g.pz, g.Nc, g.nx, g.rp, tw, qpart, skip);
        return blocks;
    }
}

// Persistent variant: a grid of a few workgroups per CU walks the column tiles; before a tile's
// transforms start, each thread issues the loads of its share of the NEXT tile into registers, so
// one workgroup overlaps HBM latency with its own LDS / FFT phases (the all-in-LDS kernel needs
// ~100 KB of LDS in FP64, which leaves one workgroup per CU and nothing to overlap with).  Tiles
// are dealt per XCD: the workgroups of one XCD (blockIdx % 8) take consecutive tiles, whose rows
// share L2 lines when a tile row is shorter than 128 bytes.
template <int N>
constexpr int zgreen_p_t() {
    return N / PlanInfo<typename SmallRadixPlan<N>::type>::max_radix();
}

template <typename R, int N, int ZL>
constexpr int zgreen_p_threads() {
    return 3 * ZL * zgreen_p_t<N>();
}

template <typename R, int N, int ZL>
__global__ __launch_bounds__((zgreen_p_threads<R, N, ZL>())) void k_zgreen_p(cx<R>* __restrict__ S,
                                                                      const R* __restrict__ G6, long ncols, long pitch,
                                                                      long Nc, int nx, int rp,
                                                                      const cx<R>* __restrict__ twg,
                                                                      double* __restrict__ qpart, const int* skip) {
    if (skip != nullptr && *skip) return;
    using PLAN = typename SmallRadixPlan<N>::type;
    constexpr int T = zgreen_p_t<N>();
    constexpr int LD = N + 1, L3 = 3 * ZL, THREADS = L3 * T;
    constexpr int EST = THREADS / L3;  // = T
    constexpr int KL = N / EST;
    constexpr int MT = THREADS / ZL;
    constexpr int KM = (N + MT - 1) / MT;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long ntiles = (ncols + ZL - 1) / ZL;
    // tiles of this workgroup: XCD x = blockIdx % 8 owns tiles [x*per, min((x+1)*per, ntiles)),
    // walked by the workgroups of that XCD in turn
    const int nx8 = (int)(gridDim.x / 8);
    const long per = (ntiles + 7) / 8;
    const int xcd = (int)(blockIdx.x % 8), slot = (int)(blockIdx.x / 8);
    const long tbeg = xcd * per, tend = min(ntiles, (long)(xcd + 1) * per);
    const int lt = threadIdx.x % L3;
    const int comp = lt / ZL;
    const int ci = lt % ZL;
    const int mci = threadIdx.x % ZL;
    auto fetch = [&](long tile, cx<R> (&ld)[KL]) {
        const long col0 = tile * ZL;
        const bool own = col0 + ci < ncols;
        const cx<R>* Sc = S + comp * Nc + col0 + ci;
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            ld[k] = mkc<R>((R)0, (R)0);
            if (own) ld[k] = Sc[(long)(threadIdx.x / L3 + k * EST) * pitch];
        }
    };
    double q[1] = {0.0};
    cx<R> ld[KL];
    long tile = tbeg + slot;
    if (tile < tend) fetch(tile, ld);
    __syncthreads();  // twiddles
    for (; tile < tend; tile += nx8) {
        const long col0 = tile * ZL;
        const int ncl = (int)min((long)ZL, ncols - col0);
#pragma unroll
        for (int k = 0; k < KL; ++k) buf[lt * LD + threadIdx.x / L3 + k * EST] = ld[k];
        // this tile's Green tensor, then the next tile's columns, in flight during the transforms
        const bool mown = mci < ncl;
        R gt[KM][6];
        {
            const R* Gc = G6 + 6L * (col0 + mci);
#pragma unroll
            for (int m = 0; m < KM; ++m) {
                const int e = threadIdx.x / ZL + m * MT;
#pragma unroll
                for (int qq = 0; qq < 6; ++qq) gt[m][qq] = (R)0;
                if (mown && e < N) {
                    const R* gp = Gc + 6L * e * pitch;
#pragma unroll
                    for (int qq = 0; qq < 6; ++qq) gt[m][qq] = gp[qq];
                }
            }
        }
        const long next = tile + nx8;
        if (next < tend) fetch(next, ld);
        __syncthreads();
        sfft_lines_plan<R, -1, N, L3, T, LD, PLAN>(buf, tw, L3);
#pragma unroll
        for (int m = 0; m < KM; ++m) {
            const int e = threadIdx.x / ZL + m * MT;
            if (!mown || e >= N) continue;
            const R txx = gt[m][0], txy = gt[m][1], txz = gt[m][2];
            const R tyy = gt[m][3], tyz = gt[m][4], tzz = gt[m][5];
            cx<R>* px = buf + (0 * ZL + mci) * LD + e;
            cx<R>* py = buf + (1 * ZL + mci) * LD + e;
            cx<R>* pz = buf + (2 * ZL + mci) * LD + e;
            const cx<R> fx = *px, fy = *py, fz = *pz;
            cx<R> gx, gy, gz;
            green_mul<R>(txx, txy, txz, tyy, tyz, tzz, fx, fy, fz, gx, gy, gz);
            *px = gx;
            *py = gy;
            *pz = gz;
            if (qpart != nullptr)
                green_quad_add<R>(q[0], herm_weight((int)((col0 + mci) % rp), nx), fx, fy, fz, gx, gy, gz);
        }
        __syncthreads();
        sfft_lines_plan<R, +1, N, L3, T, LD, PLAN>(buf, tw, L3);
        cx<R>* Sw = S + comp * Nc + col0 + ci;
        const bool own = ci < ncl;
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const int e = threadIdx.x / L3 + k * EST;
            if (own) Sw[(long)e * pitch] = buf[lt * LD + e];
        }
        __syncthreads();  // the next tile overwrites the buffer
    }
    if (qpart != nullptr) block_sum<1>(q, qpart + blockIdx.x);
}

template <typename R, int N, int ZL>
long launch_zgreen_p(cx<R>* S, const R* G6, long ncols, const Geometry3& g, const cx<R>* tw, double* qpart,
                     const int* skip, hipStream_t st) {
    if constexpr (!is_pow2(N)) {  // radix <= 8 plans exist for the powers of two only
        return -1;
    } else if constexpr (zgreen_p_threads<R, N, ZL>() > 1024) {
        return -1;
    } else {
        constexpr int threads = zgreen_p_threads<R, N, ZL>();
        constexpr size_t lds = (size_t)(N + 3 * ZL * (N + 1)) * sizeof(cx<R>);
        if (lds > lds_limit()) return -1;
        const long ntiles = (ncols + ZL - 1) / ZL;
        // as many workgroups as fit the device at once (LDS-limited), a multiple of 8 (XCDs)
        const long per_cu = std::max<long>(1, (long)(lds_per_cu() / lds));
        long blocks = std::min<long>((long)device_cus() * per_cu, ((ntiles + 7) / 8) * 8);
        blocks = std::max<long>(8, blocks / 8 * 8);
        zlaunch((k_zgreen_p<R, N, ZL>), dim3((unsigned)blocks), dim3(threads), lds, st, S, G6, ncols, g.pz,
                           g.Nc, g.nx, g.rp, tw, qpart, skip);
        return blocks;
    }
}

// Register-resident variant (zgreen_reg.hpp): ZL columns, N/4 threads per column.
template <typename R, int N, int ZL>
__global__ __launch_bounds__(ZL * N / 4) void k_zgreen_r(cx<R>* __restrict__ S, const R* __restrict__ G6, long ncols,
                                                         long pitch, long Nc, int nx, int rp,
                                                         const cx<R>* __restrict__ twg,
                                                         double* __restrict__ qpart, const int* skip) {
    if (skip != nullptr && *skip) return;
    constexpr int THREADS = ZL * N / 4, LD = N + 1, Q = N / 4;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long col0 = xcd_tile(blockIdx.x, gridDim.x) * ZL;
    const int ncl = (int)min((long)ZL, ncols - col0);
    const int ci = threadIdx.x % ZL;
    const int j = threadIdx.x / ZL;
    const bool own = ci < ncl;
    const long colg = col0 + ci;
    // element offsets inside one component (Nc < 2^31 for every grid this kernel serves) off a
    // wave-uniform component base: one 32-bit offset per lane instead of a 64-bit address per access
    const int nc = (int)pitch;
    const int off0 = (int)colg + j * nc;
    // element j + r N/4 of the three components
    cx<R> f[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const cx<R>* Sc = S + c * Nc;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            f[c][r] = mkc<R>((R)0, (R)0);
            if (own) f[c][r] = Sc[off0 + r * Q * nc];
        }
    }
    __syncthreads();
    cx<R>* col = buf + ci * LD;
#pragma unroll
    for (int c = 0; c < 3; ++c) rfft_fwd<R, N>(f[c], col, tw, j);
    // symmetric Green multiply (src/bignonnet/GreenOperator.cpp:84-92 with G_sym) on registers; the
    // tensor is read here rather than up front, which keeps the workgroup's register footprint
    // small enough for several workgroups per CU (they hide each other's load latency)
    double qf[1] = {0.0};
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const cx<R> fx = f[0][r], fy = f[1][r], fz = f[2][r];
        R t6[6];
        const R* gp = G6 + 6L * (off0 + r * Q * nc);
#pragma unroll
        for (int q = 0; q < 6; ++q) t6[q] = own ? gp[q] : (R)0;
        const R txx = t6[0], txy = t6[1], txz = t6[2], tyy = t6[3], tyz = t6[4], tzz = t6[5];
        cx<R> gx, gy, gz;
        green_mul<R>(txx, txy, txz, tyy, tyz, tzz, fx, fy, fz, gx, gy, gz);
        if (qpart != nullptr)
            green_quad_add<R>(qf[0], 1.0, fx, fy, fz, gx, gy, gz);
        f[0][r] = gx;
        f[1][r] = gy;
        f[2][r] = gz;
        // two frequencies per scheduling group: the tensor loads of the next pair are not hoisted
        // above this point, which caps the live registers (FP64: 4 x 6 tensor values would not fit
        // next to the 12 spectrum values at two workgroups per CU)
        if (r % 2 == 1) __builtin_amdgcn_sched_barrier(0);
    }
    if (qpart != nullptr) {
        qf[0] = own ? qf[0] * herm_weight((int)(colg % rp), nx) : 0.0;
        block_sum<1>(qf, qpart + blockIdx.x);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cx<R>* Sc = S + c * Nc;
        rfft_bwd<R, N>(f[c], col, tw, j, [&](int e, cx<R> v) {
            if (own) Sc[(int)colg + e * nc] = v;
        });
    }
}

template <typename R, int N, int ZL>
constexpr size_t zgreen_r_lds() {
    return (size_t)(N + ZL * (N + 1)) * sizeof(cx<R>);
}

template <typename R, int N, int ZL>
long launch_zgreen_r(cx<R>* S, const R* G6, long ncols, const Geometry3& g, const cx<R>* tw, double* qpart,
                     const int* skip, hipStream_t st) {
    constexpr int threads = ZL * N / 4;
    if constexpr (!is_pow2(N) || threads > 1024 || threads < 64) {  // radix-4 stages: powers of two
        return -1;
    } else {
        constexpr size_t lds = zgreen_r_lds<R, N, ZL>();
        if (lds > lds_limit()) return -1;
        const long blocks = (ncols + ZL - 1) / ZL;
        zlaunch((k_zgreen_r<R, N, ZL>), dim3((unsigned)blocks), dim3(threads), lds, st, S, G6, ncols, g.pz,
                           g.Nc, g.nx, g.rp, tw, qpart, skip);
        return blocks;
    }
}

template <typename R, int N, int ZL>
constexpr size_t zgreen_lds() {
    return (size_t)(N + 3 * ZL * ZShape<R, N>::LD) * sizeof(cx<R>);
}

template <typename R, int N, int ZL, int OS>
long launch_zgreen_zl(cx<R>* S, const R* G6, long ncols, const Geometry3& g, const cx<R>* tw, double* qpart,
                      const int* skip, hipStream_t st, ZBatch zb = kOneRhs) {
    if constexpr (zgreen_threads<R, N, ZL, OS>() > 1024) {
        return -1;
    } else {
        constexpr size_t lds = zgreen_lds<R, N, ZL>();
        if (lds > lds_limit()) return -1;
        const long blocks = (ncols + ZL - 1) / ZL;
        constexpr int threads = zgreen_threads<R, N, ZL, OS>();
        if (zb.nrhs == 1) {
            // the FP32 operator's tensor is streamed where the spectra fit the last-level cache alone
            if (sizeof(R) == 4 && zpass_streams_green<R>(g))
                zlaunch((k_zgreen_s<R, N, ZL, OS, sizeof(R) == 4>), dim3((unsigned)blocks), dim3(threads), lds, st,
                        S, G6, ncols, g.pz, g.Nc, g.nx, g.rp, tw, qpart, skip);
            else
                zlaunch((k_zgreen_s<R, N, ZL, OS, false>), dim3((unsigned)blocks), dim3(threads), lds, st, S, G6,
                        ncols, g.pz, g.Nc, g.nx, g.rp, tw, qpart, skip);
        } else if constexpr (sizeof(R) == 4) {  // the multi-RHS solve runs the FP32 operator
            zlaunch((k_zgreen_sb<R, N, ZL, OS>), dim3((unsigned)blocks), dim3(threads), lds, st, S, G6,
                               ncols, g.pz, g.Nc, g.nx, g.rp, tw, qpart, skip, zb);
        } else {
            return -1;
        }
        return blocks;
    }
}

// variant code = 10 * ZL + OS for the all-in-LDS kernel (OS = 1, 2 threads per butterfly thread),
// 10 * ZL + 4 for the register-resident radix-4 kernel, 10 * ZL + 5 for the persistent kernel,
// 10 * ZL + 6 / + 7 / + 8 for the component-sequential kernel with OS = 1 / 2 / 4
constexpr int kZgreenCodes[] = {11, 12, 21, 22, 41, 42, 81, 82, 161, 162, 44, 84, 164, 324, 45, 85, 165,
                                86, 87, 88, 166, 167, 168, 326, 327};

// zb.nrhs > 1: only the all-in-LDS kernel serves several right-hand sides in one launch; the
// other codes return -1 for it.
template <typename R, int N>
long launch_zgreen_n(int code, cx<R>* S, const R* G6, long ncols, const Geometry3& g, const cx<R>* tw,
                     double* qpart, const int* skip, hipStream_t st, ZBatch zb = kOneRhs) {
    if (zb.nrhs > 1 && code % 10 != 1 && code % 10 != 2) return -1;
    switch (code) {
        case 11: return launch_zgreen_zl<R, N, 1, 1>(S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 12: return launch_zgreen_zl<R, N, 1, 2>(S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 21: return launch_zgreen_zl<R, N, 2, 1>(S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 22: return launch_zgreen_zl<R, N, 2, 2>(S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 41: return launch_zgreen_zl<R, N, 4, 1>(S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 42: return launch_zgreen_zl<R, N, 4, 2>(S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 81: return launch_zgreen_zl<R, N, 8, 1>(S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 82: return launch_zgreen_zl<R, N, 8, 2>(S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 161: return launch_zgreen_zl<R, N, 16, 1>(S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 162: return launch_zgreen_zl<R, N, 16, 2>(S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 44: return launch_zgreen_r<R, N, 4>(S, G6, ncols, g, tw, qpart, skip, st);
        case 84: return launch_zgreen_r<R, N, 8>(S, G6, ncols, g, tw, qpart, skip, st);
        case 164: return launch_zgreen_r<R, N, 16>(S, G6, ncols, g, tw, qpart, skip, st);
        case 324: return launch_zgreen_r<R, N, 32>(S, G6, ncols, g, tw, qpart, skip, st);
        case 45: return launch_zgreen_p<R, N, 4>(S, G6, ncols, g, tw, qpart, skip, st);
        case 85: return launch_zgreen_p<R, N, 8>(S, G6, ncols, g, tw, qpart, skip, st);
        case 165: return launch_zgreen_p<R, N, 16>(S, G6, ncols, g, tw, qpart, skip, st);
        case 86: return launch_zgreen_c<R, N, 8, 1>(S, G6, ncols, g, tw, qpart, skip, st);
        case 87: return launch_zgreen_c<R, N, 8, 2>(S, G6, ncols, g, tw, qpart, skip, st);
        case 88: return launch_zgreen_c<R, N, 8, 4>(S, G6, ncols, g, tw, qpart, skip, st);
        case 168: return launch_zgreen_c<R, N, 16, 4>(S, G6, ncols, g, tw, qpart, skip, st);
        case 166: return launch_zgreen_c<R, N, 16, 1>(S, G6, ncols, g, tw, qpart, skip, st);
        case 167: return launch_zgreen_c<R, N, 16, 2>(S, G6, ncols, g, tw, qpart, skip, st);
        case 326: return launch_zgreen_c<R, N, 32, 1>(S, G6, ncols, g, tw, qpart, skip, st);
        case 327: return launch_zgreen_c<R, N, 32, 2>(S, G6, ncols, g, tw, qpart, skip, st);
        default: return -1;
    }
}

}  // namespace

}  // namespace fftperm

// kernels_fft.hip -- the five sweeps of the Green-operator apply on the R2C half spectrum.
//
// Layout in HBM: real fields are SoA by component, flat index (z*ny + y)*nx + x
// (src/core/Grid3D.hpp:41-43).  The spectrum S holds three components of the
// nz*ny*nxh frequencies (nxh = nx/2 + 1) in the padded layout of Geometry3:
// component c, frequency (kz, ky, kx) at c*Nc + kz*pz + ky*rp + kx.
//
//   x-forward : real x-lines -> complex lines -> n-point FFT -> first nxh bins
//               (source values built on the fly from the support map: fused scatter)
//   y-pass    : strided columns along y, one component per blockIdx.y
//   z-green   : all three components of a tile of z-columns in LDS: forward FFT,
//               symmetric 3x3 Green multiply, backward FFT (one HBM round trip)
//   x-inverse : Hermitian completion of each x-line, backward FFT, real part / N
//               (sink: fused gather to the compact vector and/or field write,
//               per-block partial sums for the support/domain means)
#include "fft_engine.hpp"
#include "fft_static.hpp"
#include "kernels.hpp"
#include "zgreen_reg.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace fftperm {

KernelEvents& zpass_events() {
    thread_local KernelEvents ev;
    return ev;
}

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

size_t g_lds_max = 64 * 1024;  // per-workgroup LDS limit of the device (set at context creation)
int g_cus = 256;               // compute units of the device (set at context creation)
size_t g_lds_cu = 160 * 1024;  // LDS per compute unit

size_t lds_single() { return std::min<size_t>(72 * 1024, g_lds_max); }
size_t lds_triple() { return std::min<size_t>(104 * 1024, g_lds_max); }

int floor_pow2(long v) {
    int p = 1;
    while ((long)p * 2 <= v) p *= 2;
    return p;
}

PassCfg make_cfg(const LinePlan& p, long nunits, int ncomp, size_t budget, size_t csize) {
    const int n = p.n;
    const int maxbf = std::max(1, p.maxbf);
    long lines = 1024 / (ncomp * maxbf);
    const size_t per_line = (size_t)(n + 1) * csize * ncomp;
    const size_t tw = (size_t)n * csize;
    long by_lds = budget > tw ? (long)((budget - tw) / per_line) : 1;
    lines = std::min(lines, std::max(1L, by_lds));
    lines = floor_pow2(std::max(1L, lines));
    if (lines > nunits) lines = std::max(1L, nunits);
    PassCfg c;
    c.lines = (int)lines;
    int thr = ncomp * (int)lines * maxbf;
    thr = ((thr + 63) / 64) * 64;
    thr = std::max(thr, 256);
    thr = std::min(thr, 1024);
    c.threads = thr;
    c.lds = tw + per_line * (size_t)lines;
    c.blocks = (nunits + lines - 1) / lines;
    return c;
}

}  // namespace

PassCfg cfg_xpass(const LinePlan& p, long nlines) { return make_cfg(p, nlines, 1, lds_single(), 16); }
PassCfg cfg_colpass(const LinePlan& p, long ncols) { return make_cfg(p, ncols, 1, lds_single(), 16); }
PassCfg cfg_zgreen(const LinePlan& p, long ncols) { return make_cfg(p, ncols, 3, lds_triple(), 16); }

void set_lds_limit(size_t bytes) { g_lds_max = bytes; }
void set_device_shape(int cus, size_t lds_per_cu) {
    if (cus > 0) g_cus = cus;
    if (lds_per_cu > 0) g_lds_cu = lds_per_cu;
}
size_t lds_limit() { return g_lds_max; }


// =====================================================================================
// Compile-time specialised fused z pass for N in {32, 64, 128, 256, 512}; the specialised x and
// column passes live in passes_x.hip and passes_col.hip.
// =====================================================================================
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

// Right-hand sides of one fused z-pass launch: side b's spectrum starts at b * rhs_stride, its
// quadratic-form partials at b * qstride, and skip[b * skip_stride] set skips it.
struct ZBatch {
    int nrhs;
    int skip_stride;
    long rhs_stride;
    long qstride;
};
constexpr ZBatch kOneRhs{1, 0, 0, 0};

template <typename R, int N, int ZL, int OS>
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
    constexpr int KL = N / EST;
    static_assert(N % EST == 0, "column elements split evenly over the load threads");
    const cx<R>* Sc = S + comp * Nc + col0 + ci;
    cx<R> ld[KL];
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        ld[k] = mkc<R>((R)0, (R)0);
        if (own) ld[k] = Sc[(long)(threadIdx.x / L3 + k * EST) * pitch];
    }
    // the Green tensor of the frequencies this thread multiplies, fetched now so that its latency
    // overlaps the column loads and the forward transform
    constexpr int MT = THREADS / ZL;
    constexpr int KM = (N + MT - 1) / MT;
    const int mci = threadIdx.x % ZL;
    const bool mown = mci < ncl;
    R gt[KM][6];
    {
        const R* Gc = G6 + 6L * (col0 + mci);
#pragma unroll
        for (int m = 0; m < KM; ++m) {
            const int e = threadIdx.x / ZL + m * MT;
#pragma unroll
            for (int q = 0; q < 6; ++q) gt[m][q] = (R)0;
            if (mown && e < N) {
                const R* gp = Gc + 6L * e * pitch;
#pragma unroll
                for (int q = 0; q < 6; ++q) gt[m][q] = gp[q];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < KL; ++k) buf[lt * LD + threadIdx.x / L3 + k * EST] = ld[k];
    __syncthreads();
    sfft_lines<R, -1, N, L3, SH::T, LD>(buf, tw, L3);
    // symmetric Green multiply (src/bignonnet/GreenOperator.cpp:84-92 with G_sym), and the
    // quadratic form conj(f) . G f of the forward spectrum when asked for
    {
        double q[1] = {0.0};
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
                green_quad_add<R>(q[0], 1.0, fx, fy, fz, gx, gy, gz);
        }
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
        if (own) Sw[(long)e * pitch] = buf[lt * LD + e];
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
    constexpr int KL = N / EST;
    static_assert(N % EST == 0, "column elements split evenly over the load threads");
    constexpr int MT = THREADS / ZL;
    constexpr int KM = (N + MT - 1) / MT;
    const int mci = threadIdx.x % ZL;
    const bool mown = mci < ncl;
    // the Green tensor of the frequencies this thread multiplies stays in registers for every
    // right-hand side of the launch: one read of G per launch
    bool gloaded = false;
    R gt[KM][6];
    for (int b = 0; b < zb.nrhs; ++b) {
        if (skip != nullptr && skip[b * zb.skip_stride]) continue;
        cx<R>* Sb = S + b * zb.rhs_stride;
        const cx<R>* Sc = Sb + comp * Nc + col0 + ci;
        cx<R> ld[KL];
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            ld[k] = mkc<R>((R)0, (R)0);
            if (own) ld[k] = Sc[(long)(threadIdx.x / L3 + k * EST) * pitch];
        }
        if (!gloaded) {
            // fetched after the first side's column loads are issued, so that its latency overlaps
            // them and the forward transform; the twiddles go to LDS meanwhile
            const R* Gc = G6 + 6L * (col0 + mci);
#pragma unroll
            for (int m = 0; m < KM; ++m) {
                const int e = threadIdx.x / ZL + m * MT;
#pragma unroll
                for (int q = 0; q < 6; ++q) gt[m][q] = (R)0;
                if (mown && e < N) {
                    const R* gp = Gc + 6L * e * pitch;
#pragma unroll
                    for (int q = 0; q < 6; ++q) gt[m][q] = gp[q];
                }
            }
            for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
            gloaded = true;
        } else {
            __syncthreads();  // the previous side's stores have read the buffer
        }
#pragma unroll
        for (int k = 0; k < KL; ++k) buf[lt * LD + threadIdx.x / L3 + k * EST] = ld[k];
        __syncthreads();
        sfft_lines<R, -1, N, L3, SH::T, LD>(buf, tw, L3);
        // symmetric Green multiply (src/bignonnet/GreenOperator.cpp:84-92 with G_sym), and the
        // quadratic form conj(f) . G f of the forward spectrum when asked for
        {
            double q[1] = {0.0};
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
                    green_quad_add<R>(q[0], 1.0, fx, fy, fz, gx, gy, gz);
            }
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
            if (own) Sw[(long)e * pitch] = buf[lt * LD + e];
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
    if constexpr (threads > 1024 || threads < 64 || lds > 100u * 1024u) {
        return -1;
    } else {
        if (lds > g_lds_max) return -1;
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
    constexpr int threads = zgreen_p_threads<R, N, ZL>();
    if constexpr (threads > 1024) {
        return -1;
    } else {
        constexpr size_t lds = (size_t)(N + 3 * ZL * (N + 1)) * sizeof(cx<R>);
        if (lds > g_lds_max) return -1;
        const long ntiles = (ncols + ZL - 1) / ZL;
        // as many workgroups as fit the device at once (LDS-limited), a multiple of 8 (XCDs)
        const long per_cu = std::max<long>(1, (long)(g_lds_cu / lds));
        long blocks = std::min<long>((long)g_cus * per_cu, ((ntiles + 7) / 8) * 8);
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
    if constexpr (threads > 1024 || threads < 64) {
        return -1;
    } else {
        constexpr size_t lds = zgreen_r_lds<R, N, ZL>();
        if (lds > g_lds_max) return -1;
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
        if (lds > g_lds_max) return -1;
        const long blocks = (ncols + ZL - 1) / ZL;
        constexpr int threads = zgreen_threads<R, N, ZL, OS>();
        if (zb.nrhs == 1) {
            zlaunch((k_zgreen_s<R, N, ZL, OS>), dim3((unsigned)blocks), dim3(threads), lds, st, S, G6,
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

template <typename R>
long static_zgreen_code(int code, cx<R>* S, const R* G6, long ncols, const Geometry3& g, const cx<R>* tw,
                        double* qpart, const int* skip, hipStream_t st, ZBatch zb = kOneRhs) {
    switch (g.nz) {
        case 32: return launch_zgreen_n<R, 32>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 64: return launch_zgreen_n<R, 64>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 128: return launch_zgreen_n<R, 128>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 256: return launch_zgreen_n<R, 256>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 512: return launch_zgreen_n<R, 512>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        default: return -1;
    }
}

// untuned: the widest tile that fits ~100 KB of LDS, one thread per butterfly
int zgreen_default_code(int n, int prec) {
    switch (n) {
        case 32: return 10 * (prec ? ZShape<float, 32>::ZL : ZShape<double, 32>::ZL) + 1;
        case 64: return 10 * (prec ? ZShape<float, 64>::ZL : ZShape<double, 64>::ZL) + 1;
        case 128: return 10 * (prec ? ZShape<float, 128>::ZL : ZShape<double, 128>::ZL) + 1;
        case 256: return 10 * (prec ? ZShape<float, 256>::ZL : ZShape<double, 256>::ZL) + 1;
        case 512: return 10 * (prec ? ZShape<float, 512>::ZL : ZShape<double, 512>::ZL) + 1;
        default: return 0;
    }
}

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, long>, int> g_tuned;
std::map<std::tuple<int, int, int, long>, std::vector<int>> g_live;  // variants that launched
long g_tune_gen = 0;

template <typename R>
constexpr int prec_of() {
    return sizeof(R) == 8 ? 0 : 1;
}

// Times `launch` (which returns < 0 when the variant does not exist) for every candidate and
// records the fastest under (kind, prec, n, count).  Every candidate runs once first (existence
// check and warm-up, so that the clocks are up before any timing); then kTuneRounds rounds visit
// the candidates in turn, each timing kTuneReps launches, and a candidate's time is its best
// round -- interleaving spreads clock and cache drift over all candidates instead of charging it
// to whichever is timed first.
constexpr int kTuneRounds = 3;
constexpr int kTuneReps = 3;

template <typename F>
std::vector<std::pair<int, float>> tune_candidates(int kind, int prec, int n, long count, const std::vector<int>& cands,
                                                   F&& launch, int (*code_of)(int, int), hipStream_t st) {
    std::vector<std::pair<int, float>> timings;
    if (!static_passes_ok() || tuned_variant(kind, prec, n, count) >= 0) return timings;
    std::vector<int> live;
    for (int v : cands)
        if (launch(v) >= 0) live.push_back(v);
    if (live.empty()) return timings;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return timings;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return timings;
    }
    std::vector<float> best_of(live.size(), 1e30f);
    bool ok = true;
    for (int round = 0; round < kTuneRounds && ok; ++round) {
        for (size_t i = 0; i < live.size(); ++i) {
            (void)hipEventRecord(a, st);
            for (int rep = 0; rep < kTuneReps; ++rep) launch(live[i]);
            (void)hipEventRecord(b, st);
            if (hipEventSynchronize(b) != hipSuccess) {
                ok = false;
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            best_of[i] = std::min(best_of[i], ms / (float)kTuneReps);
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!ok) return timings;
    size_t best = 0;
    for (size_t i = 0; i < live.size(); ++i) {
        timings.emplace_back(code_of(n, live[i]), best_of[i]);
        if (best_of[i] < best_of[best]) best = i;
    }
    set_tuned_variant(kind, prec, n, count, live[best]);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_live[std::make_tuple(kind, prec, n, count)] = live;
    }
    return timings;
}

int zgreen_code_identity(int, int v) { return v; }

std::vector<int> pass_candidates(int count = kPassVariants) {
    std::vector<int> v;
    for (int i = 0; i < count; ++i) v.push_back(i);
    return v;
}

int pass_variant_or_default(int kind, int prec, int n, long count) {
    const int v = tuned_variant(kind, prec, n, count);
    return v < 0 ? 0 : v;
}

}  // namespace

// The specialised kernels need their static LDS budget; the runtime kernels cover devices with
// less LDS per workgroup and every other length.
bool static_passes_ok() { return g_lds_max >= 102400u + 8192u; }

int tuned_variant(int kind, int prec, int n, long count) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(std::make_tuple(kind, prec, n, count));
    return it == g_tuned.end() ? -1 : it->second;
}

void set_tuned_variant(int kind, int prec, int n, long count, int v) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[std::make_tuple(kind, prec, n, count)] = v;
    ++g_tune_gen;
}

std::vector<int> tuned_candidates(int kind, int prec, int n, long count) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_live.find(std::make_tuple(kind, prec, n, count));
    return it == g_live.end() ? std::vector<int>() : it->second;
}

long tuning_generation() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return g_tune_gen;
}

void clear_tuned_variant(int kind, int prec, int n, long count) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.erase(std::make_tuple(kind, prec, n, count));
    ++g_tune_gen;
}

int variant_of_code(int kind, int n, int code) {
    if (kind == kKindZgreen) {
        for (int c : kZgreenCodes)
            if (c == code) return c;
        return -1;
    }
    if (kind == kKindCol) {
        for (int v = 0; v < kColVariants; ++v)
            if (col_variant_code(n, v) == code && code != 0) return v;
        return -1;
    }
    for (int v = 0; v < kXPassVariants; ++v)
        if (pass_variant_code(n, v) == code && code != 0) return v;
    return -1;
}

int variant_code(int kind, int prec, int n, long count) {
    if (!static_passes_ok()) return 0;
    if (kind == kKindZgreen) {
        const int c = tuned_variant(kind, prec, n, count);
        return c < 0 ? zgreen_default_code(n, prec) : c;
    }
    const int v = pass_variant_or_default(kind, prec, n, count);
    return kind == kKindCol ? col_variant_code(n, v) : pass_variant_code(n, v);
}

template <typename R>
std::vector<std::pair<int, float>> autotune_zgreen(cx<R>* S, const R* G6, const Geometry3& g, const FFTTables<R>& t,
                                                   hipStream_t st) {
    const long key = (long)g.ny * g.nxh;   // registry count: the frequency columns
    const long ncols = (long)g.ny * g.rp;  // columns a launch covers: every padded row
    (void)hipMemsetAsync(S, 0, 3 * (size_t)g.Nc * sizeof(cx<R>), st);
    std::vector<int> cands(std::begin(kZgreenCodes), std::end(kZgreenCodes));
    return tune_candidates(
        kKindZgreen, prec_of<R>(), g.nz, key, cands,
        [&](int code) { return static_zgreen_code<R>(code, S, G6, ncols, g, t.twz, nullptr, nullptr, st); },
        zgreen_code_identity, st);
}

template <typename R>
std::vector<std::pair<int, float>> autotune_col(cx<R>* S, const Geometry3& g, const FFTTables<R>& t, hipStream_t st) {
    const long key = (long)g.nz * g.nxh;   // registry count: the frequency columns
    const long ncols = (long)g.nz * g.rp;  // columns a launch covers: every padded row
    (void)hipMemsetAsync(S, 0, 3 * (size_t)g.Nc * sizeof(cx<R>), st);
    return tune_candidates(
        kKindCol, prec_of<R>(), g.ny, key, pass_candidates(kColVariants),
        [&](int v) {
            return static_col<R>(v, g.ny, -1, S, ncols, (long)g.rp, g.pz, (long)g.rp, g.Nc, t.twy, nullptr, st);
        },
        col_variant_code, st);
}

template <typename R>
std::vector<std::pair<int, float>> autotune_x(const XSource& src, const XSink& sink, cx<R>* S, const Geometry3& g,
                                              const FFTTables<R>& t, hipStream_t st) {
    const long nlines = (long)g.ny * g.nz;
    auto tf = tune_candidates(
        kKindXfwd, prec_of<R>(), g.nx, nlines, pass_candidates(kXPassVariants),
        [&](int v) { return static_xfwd<R>(v, src, S, g, t.twx, nullptr, st); }, pass_variant_code, st);
    auto ti = tune_candidates(
        kKindXinv, prec_of<R>(), g.nx, nlines, pass_candidates(kXPassVariants),
        [&](int v) { return static_xinv<R>(v, S, sink, g, t.twx, nullptr, st); }, pass_variant_code, st);
    tf.insert(tf.end(), ti.begin(), ti.end());
    return tf;
}

template std::vector<std::pair<int, float>> autotune_zgreen<double>(cx<double>*, const double*, const Geometry3&,
                                                                    const FFTTables<double>&, hipStream_t);
template std::vector<std::pair<int, float>> autotune_zgreen<float>(cx<float>*, const float*, const Geometry3&,
                                                                   const FFTTables<float>&, hipStream_t);
template std::vector<std::pair<int, float>> autotune_col<double>(cx<double>*, const Geometry3&,
                                                                 const FFTTables<double>&, hipStream_t);
template std::vector<std::pair<int, float>> autotune_col<float>(cx<float>*, const Geometry3&, const FFTTables<float>&,
                                                                hipStream_t);
template std::vector<std::pair<int, float>> autotune_x<double>(const XSource&, const XSink&, cx<double>*,
                                                               const Geometry3&, const FFTTables<double>&,
                                                               hipStream_t);
template std::vector<std::pair<int, float>> autotune_x<float>(const XSource&, const XSink&, cx<float>*,
                                                              const Geometry3&, const FFTTables<float>&, hipStream_t);

long xinv_max_blocks(const Geometry3& g) {
    // every variant and the runtime kernel put at least one line in a workgroup
    return (long)g.ny * g.nz;
}

long zgreen_max_blocks(const Geometry3& g) { return (long)g.ny * g.rp; }

// ------------------------------------------------------------------ x forward
template <typename R>
__global__ void k_xfwd(XSource src, cx<R>* __restrict__ S, int nx, long nlines, int nxh, long N, SpecLayout SL,
                       LinePlan pl, const cx<R>* __restrict__ twg, int lpw, const int* skip) {
    if (skip != nullptr && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + nx;
    const int ld = nx + 1;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) tw[i] = twg[i];
    const long line0 = (long)blockIdx.x * lpw;
    const int nl = (int)min((long)lpw, nlines - line0);
    const int cnt = nl * nx;
    for (int c = 0; c < 3; ++c) {
        const double af = src.a_fluid[c];
        const double as = src.a_support[c];
        const double mc = (src.mean != nullptr) ? src.mean[c] : 0.0;
        for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
            const int line = idx / nx;
            const int i = idx - line * nx;
            const long vox = (line0 + line) * nx + i;
            double val;
            if (src.field != nullptr) {
                val = src.field[c * N + vox];
            } else {
                const int code = src.code[vox];
                const double f0 = (code == -1) ? af : ((code >= 0) ? as : 0.0);
                double fx = 0.0;
                if (src.compact != nullptr && code >= 0) fx = src.compact[3L * code + c] - mc;
                if (src.compact32 != nullptr && code >= 0) fx = (double)src.compact32[3L * code + c] - mc;
                val = f0 + fx;
            }
            buf[line * ld + i] = mkc<R>((R)val, (R)0);
        }
        __syncthreads();
        fft_lines<R, -1>(buf, ld, nl, pl, tw);
        const int cnth = nl * nxh;
        cx<R>* Sc = S + c * SL.nc;
        for (int idx = threadIdx.x; idx < cnth; idx += blockDim.x) {
            const int line = idx / nxh;
            const int kx = idx - line * nxh;
            Sc[spec_line(SL, line0 + line) + kx] = buf[line * ld + kx];
        }
        __syncthreads();
    }
}

template <typename R>
void launch_xfwd(const XSource& src, cx<R>* S, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                 hipStream_t st) {
    const long nlines = (long)g.ny * g.nz;
    if (static_passes_ok() &&
        static_xfwd<R>(pass_variant_or_default(kKindXfwd, prec_of<R>(), g.nx, nlines), src, S, g, t.twx, skip, st) >= 0)
        return;
    PassCfg c = cfg_xpass(t.px, nlines);
    hipLaunchKernelGGL(k_xfwd<R>, dim3((unsigned)c.blocks), dim3(c.threads), c.lds, st, src, S, g.nx, nlines,
                       g.nxh, g.N, spec_layout(g), t.px, t.twx, c.lines, skip);
}

// ------------------------------------------------------------------ x inverse
template <typename R>
__global__ void k_xinv(const cx<R>* __restrict__ S, XSink sink, int nx, long nlines, int nxh, long N, SpecLayout SL,
                       LinePlan pl, const cx<R>* __restrict__ twg, int lpw, double inv_n, const int* skip) {
    if (skip != nullptr && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + nx;
    const int ld = nx + 1;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) tw[i] = twg[i];
    const long line0 = (long)blockIdx.x * lpw;
    const int nl = (int)min((long)lpw, nlines - line0);
    double acc[kXinvParts] = {0, 0, 0};
    for (int c = 0; c < 3; ++c) {
        const cx<R>* Sc = S + c * SL.nc;
        const int cnth = nl * nxh;
        for (int idx = threadIdx.x; idx < cnth; idx += blockDim.x) {
            const int line = idx / nxh;
            const int kx = idx - line * nxh;
            buf[line * ld + kx] = Sc[spec_line(SL, line0 + line) + kx];
        }
        __syncthreads();
        const int nm = nx - nxh;
        for (int idx = threadIdx.x; idx < nl * nm; idx += blockDim.x) {
            const int line = idx / nm;
            const int kx = nxh + (idx - line * nm);
            buf[line * ld + kx] = conjc<R>(buf[line * ld + (nx - kx)]);
        }
        __syncthreads();
        fft_lines<R, +1>(buf, ld, nl, pl, tw);
        const int cnt = nl * nx;
        for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
            const int line = idx / nx;
            const int i = idx - line * nx;
            const long vox = (line0 + line) * nx + i;
            const double val = (double)buf[line * ld + i].x * inv_n;
            if (sink.field != nullptr) sink.field[c * N + vox] = val;
            const int code = sink.code[vox];
            if (code >= 0) {
                if (sink.compact != nullptr) sink.compact[3L * code + c] = val;
                if (sink.compact32 != nullptr) sink.compact32[3L * code + c] = (float)val;
                if (c == 0) acc[0] += val;
                if (c == 1) acc[1] += val;
                if (c == 2) acc[2] += val;
            }
        }
        __syncthreads();
    }
    if (sink.partials != nullptr) {
        block_sum<kXinvParts>(acc, sink.partials + (long)kXinvParts * blockIdx.x);
    }
}

template <typename R>
long launch_xinv(const cx<R>* S, const XSink& sink, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                 hipStream_t st) {
    const long nlines = (long)g.ny * g.nz;
    if (static_passes_ok()) {
        const long b =
            static_xinv<R>(pass_variant_or_default(kKindXinv, prec_of<R>(), g.nx, nlines), S, sink, g, t.twx, skip, st);
        if (b >= 0) return b;
    }
    PassCfg c = cfg_xpass(t.px, nlines);
    const double inv_n = 1.0 / (double)g.N;
    hipLaunchKernelGGL(k_xinv<R>, dim3((unsigned)c.blocks), dim3(c.threads), c.lds, st, S, sink, g.nx, nlines,
                       g.nxh, g.N, spec_layout(g), t.px, t.twx, c.lines, inv_n, skip);
    return c.blocks;
}

// ------------------------------------------------------------- column passes
// Column col of a component: base = (col / inner) * outer_stride + col % inner,
// element e at base + e * estride.
template <typename R, int SIGN>
__global__ void k_colpass(cx<R>* __restrict__ S, int n, long ncols, long inner, long outer_stride, long estride,
                          long comp_stride, LinePlan pl, const cx<R>* __restrict__ twg, int tl,
                          const int* skip, int skip_stride) {
    if (skip != nullptr && skip[(blockIdx.y / 3) * skip_stride]) return;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + n;
    const int ld = n + 1;
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
    const long col0 = (long)blockIdx.x * tl;
    const int ncl = (int)min((long)tl, ncols - col0);
    cx<R>* Sc = S + blockIdx.y * comp_stride;
    const int cnt = ncl * n;
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
        const int e = idx / ncl;
        const int ci = idx - e * ncl;
        const long col = col0 + ci;
        const long base = (col / inner) * outer_stride + (col % inner);
        buf[ci * ld + e] = Sc[base + e * estride];
    }
    __syncthreads();
    fft_lines<R, SIGN>(buf, ld, ncl, pl, tw);
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
        const int e = idx / ncl;
        const int ci = idx - e * ncl;
        const long col = col0 + ci;
        const long base = (col / inner) * outer_stride + (col % inner);
        Sc[base + e * estride] = buf[ci * ld + e];
    }
}

template <typename R>
void launch_ypass(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                  hipStream_t st) {
    launch_ypass_n<R>(S, sign, g, t, 1, skip, 0, st);
}

template <typename R>
void launch_ypass_n(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, int nrhs, const int* skip,
                    int skip_stride, hipStream_t st) {
    // column (kz, kx) of every padded row, kx < rp: base kz * pz + kx, elements ky * rp.  Covering
    // the row padding (zeros in, zeros out) keeps every tile of 8 columns on a 64-byte boundary.
    // Component c of right-hand side b is spectrum component 3 b + c.
    const long key = (long)g.nz * g.nxh;
    const long ncols = (long)g.nz * g.rp;
    const int ncomp = 3 * nrhs;
    if (static_passes_ok() &&
        static_col<R>(pass_variant_or_default(kKindCol, prec_of<R>(), g.ny, key), g.ny, sign, S, ncols, (long)g.rp,
                      g.pz, (long)g.rp, g.Nc, t.twy, skip, st, ncomp, skip_stride) >= 0)
        return;
    PassCfg c = cfg_colpass(t.py, ncols);
    dim3 grid((unsigned)c.blocks, (unsigned)ncomp);
    if (sign < 0)
        hipLaunchKernelGGL((k_colpass<R, -1>), grid, dim3(c.threads), c.lds, st, S, g.ny, ncols, (long)g.rp, g.pz,
                           (long)g.rp, g.Nc, t.py, t.twy, c.lines, skip, skip_stride);
    else
        hipLaunchKernelGGL((k_colpass<R, +1>), grid, dim3(c.threads), c.lds, st, S, g.ny, ncols, (long)g.rp, g.pz,
                           (long)g.rp, g.Nc, t.py, t.twy, c.lines, skip, skip_stride);
}

template <typename R>
void launch_zpass(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                  hipStream_t st) {
    // column (ky, kx) of every padded row: base ky * rp + kx, elements kz * pz (padding stays zero)
    const long ncols = (long)g.ny * g.rp;
    if (static_passes_ok() && static_col<R>(0, g.nz, sign, S, ncols, ncols, 0L, g.pz, g.Nc, t.twz, skip, st) >= 0)
        return;
    PassCfg c = cfg_colpass(t.pz, ncols);
    dim3 grid((unsigned)c.blocks, 3);
    if (sign < 0)
        hipLaunchKernelGGL((k_colpass<R, -1>), grid, dim3(c.threads), c.lds, st, S, g.nz, ncols, ncols, 0L, g.pz,
                           g.Nc, t.pz, t.twz, c.lines, skip, 0);
    else
        hipLaunchKernelGGL((k_colpass<R, +1>), grid, dim3(c.threads), c.lds, st, S, g.nz, ncols, ncols, 0L, g.pz,
                           g.Nc, t.pz, t.twz, c.lines, skip, 0);
}

// ------------------------------------------------------------- fused z + Green
template <typename R>
__global__ void k_zgreen(cx<R>* __restrict__ S, const R* __restrict__ G6, int nz, long ncols, long pitch, long Nc,
                         int nx, int rp, LinePlan pl, const cx<R>* __restrict__ twg, int tl,
                         double* __restrict__ qpart, const int* skip) {
    if (skip != nullptr && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + nz;
    const int ld = nz + 1;
    for (int i = threadIdx.x; i < nz; i += blockDim.x) tw[i] = twg[i];
    const long col0 = (long)blockIdx.x * tl;
    const int ncl = (int)min((long)tl, ncols - col0);
    const int cnt = ncl * nz;
    for (int c = 0; c < 3; ++c) {
        const cx<R>* Sc = S + c * Nc;
        cx<R>* B = buf + c * ncl * ld;
        for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
            const int e = idx / ncl;
            const int ci = idx - e * ncl;
            B[ci * ld + e] = Sc[col0 + ci + e * pitch];
        }
    }
    __syncthreads();
    fft_lines<R, -1>(buf, ld, 3 * ncl, pl, tw);
    // Symmetric Green multiply, as GreenOperator::apply_spectral_green_tensor
    // (src/bignonnet/GreenOperator.cpp:84-92) with the symmetrised tensor, plus the weighted
    // quadratic form conj(f) . G f when asked for.
    double q[1] = {0.0};
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
        const int e = idx / ncl;
        const int ci = idx - e * ncl;
        const R* gp = G6 + 6L * (col0 + ci + e * pitch);
        const R txx = gp[0], txy = gp[1], txz = gp[2];
        const R tyy = gp[3], tyz = gp[4], tzz = gp[5];
        cx<R>* px = buf + (0 * ncl + ci) * ld + e;
        cx<R>* py = buf + (1 * ncl + ci) * ld + e;
        cx<R>* pz = buf + (2 * ncl + ci) * ld + e;
        const cx<R> fx = *px, fy = *py, fz = *pz;
        cx<R> gx, gy, gz;
        green_mul<R>(txx, txy, txz, tyy, tyz, tzz, fx, fy, fz, gx, gy, gz);
        *px = gx;
        *py = gy;
        *pz = gz;
        if (qpart != nullptr)
            green_quad_add<R>(q[0], herm_weight((int)((col0 + ci) % rp), nx), fx, fy, fz, gx, gy, gz);
    }
    if (qpart != nullptr) block_sum<1>(q, qpart + blockIdx.x);
    __syncthreads();
    fft_lines<R, +1>(buf, ld, 3 * ncl, pl, tw);
    for (int c = 0; c < 3; ++c) {
        cx<R>* Sc = S + c * Nc;
        const cx<R>* B = buf + c * ncl * ld;
        for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
            const int e = idx / ncl;
            const int ci = idx - e * ncl;
            Sc[col0 + ci + e * pitch] = B[ci * ld + e];
        }
    }
}

template <typename R>
long launch_zgreen(cx<R>* S, const R* G6, const Geometry3& g, const FFTTables<R>& t, double* qpart, const int* skip,
                   hipStream_t st) {
    const long key = (long)g.ny * g.nxh;   // registry count: the frequency columns
    const long ncols = (long)g.ny * g.rp;  // columns a launch covers: every padded row
    if (static_passes_ok()) {
        int code = tuned_variant(kKindZgreen, prec_of<R>(), g.nz, key);
        if (code < 0) code = zgreen_default_code(g.nz, prec_of<R>());
        const long b = static_zgreen_code<R>(code, S, G6, ncols, g, t.twz, qpart, skip, st);
        if (b >= 0) return b;
    }
    PassCfg c = cfg_zgreen(t.pz, ncols);
    zlaunch(k_zgreen<R>, dim3((unsigned)c.blocks), dim3(c.threads), c.lds, st, S, G6, g.nz, ncols, g.pz,
                       g.Nc, g.nx, g.rp, t.pz, t.twz, c.lines, qpart, skip);
    return c.blocks;
}

template <typename R>
long launch_zgreen_batch(cx<R>* S, const R* G6, const Geometry3& g, const FFTTables<R>& t, double* qpart,
                         long qstride, int nrhs, const int* skip, int skip_stride, hipStream_t st) {
    if (nrhs == 1) return launch_zgreen<R>(S, G6, g, t, qpart, skip, st);
    const long key = (long)g.ny * g.nxh;
    const long ncols = (long)g.ny * g.rp;
    const long rhs_stride = 3 * g.Nc;
    if (static_passes_ok()) {
        int code = tuned_variant(kKindZgreen, prec_of<R>(), g.nz, key);
        if (code < 0) code = zgreen_default_code(g.nz, prec_of<R>());
        const ZBatch zb{nrhs, skip_stride, rhs_stride, qstride};
        const long b = static_zgreen_code<R>(code, S, G6, ncols, g, t.twz, qpart, skip, st, zb);
        if (b >= 0) return b;
    }
    // the tuned kernel serves one side per launch: one launch per side
    long blocks = 0;
    for (int b = 0; b < nrhs; ++b)
        blocks = launch_zgreen<R>(S + b * rhs_stride, G6, g, t, qpart != nullptr ? qpart + b * qstride : nullptr,
                                  skip != nullptr ? skip + b * skip_stride : nullptr, st);
    return blocks;
}

// ------------------------------------------------------- spectrum utilities
template <typename R>
__global__ void k_expand_full(const cx<R>* __restrict__ S, double2* __restrict__ F, int nx, int ny, int nz,
                              int nxh, long N, SpecLayout SL) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int c = blockIdx.y;
    const int kx = (int)(i % nx);
    const long r = i / nx;
    const int ky = (int)(r % ny);
    const int kz = (int)(r / ny);
    double2 v;
    if (kx < nxh) {
        const cx<R> s = S[c * SL.nc + (long)kz * SL.pz + (long)ky * SL.rp + kx];
        v.x = (double)s.x;
        v.y = (double)s.y;
    } else {
        const int mz = (nz - kz) % nz, my = (ny - ky) % ny, mx = nx - kx;
        const cx<R> s = S[c * SL.nc + (long)mz * SL.pz + (long)my * SL.rp + mx];
        v.x = (double)s.x;
        v.y = -(double)s.y;
    }
    F[c * N + i] = v;
}

template <typename R>
void launch_expand_full(const cx<R>* S, double2* F, const Geometry3& g, hipStream_t st) {
    dim3 grid((unsigned)((g.N + 255) / 256), 3);
    hipLaunchKernelGGL(k_expand_full<R>, grid, dim3(256), 0, st, S, F, g.nx, g.ny, g.nz, g.nxh, g.N, spec_layout(g));
}

template <typename R>
__global__ void k_hermitian_half(const double2* __restrict__ X, cx<R>* __restrict__ S, int nx, int ny, int nz,
                                 int nxh, long N, long Nh, SpecLayout SL) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nh) return;
    const int c = blockIdx.y;
    const int kx = (int)(i % nxh);
    const long r = i / nxh;
    const int ky = (int)(r % ny);
    const int kz = (int)(r / ny);
    const double2 a = X[c * N + ((long)kz * ny + ky) * nx + kx];
    const int mz = (nz - kz) % nz, my = (ny - ky) % ny, mx = (nx - kx) % nx;
    const double2 b = X[c * N + ((long)mz * ny + my) * nx + mx];
    S[c * SL.nc + (long)kz * SL.pz + (long)ky * SL.rp + kx] = mkc<R>((R)(0.5 * (a.x + b.x)), (R)(0.5 * (a.y - b.y)));
}

template <typename R>
void launch_hermitian_half(const double2* X, cx<R>* S, const Geometry3& g, hipStream_t st) {
    dim3 grid((unsigned)((g.Nh + 255) / 256), 3);
    hipLaunchKernelGGL(k_hermitian_half<R>, grid, dim3(256), 0, st, X, S, g.nx, g.ny, g.nz, g.nxh, g.N, g.Nh,
                       spec_layout(g));
}

__global__ void k_spectral_green_full(const double* __restrict__ T6, const double2* __restrict__ F,
                                      double2* __restrict__ V, long N) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double txx = T6[i], txy = T6[N + i], txz = T6[2 * N + i];
    const double tyy = T6[3 * N + i], tyz = T6[4 * N + i], tzz = T6[5 * N + i];
    const double2 fx = F[i], fy = F[N + i], fz = F[2 * N + i];
    double2 v;
    v.x = txx * fx.x + txy * fy.x + txz * fz.x;
    v.y = txx * fx.y + txy * fy.y + txz * fz.y;
    V[i] = v;
    v.x = txy * fx.x + tyy * fy.x + tyz * fz.x;
    v.y = txy * fx.y + tyy * fy.y + tyz * fz.y;
    V[N + i] = v;
    v.x = txz * fx.x + tyz * fy.x + tzz * fz.x;
    v.y = txz * fx.y + tyz * fy.y + tzz * fz.y;
    V[2 * N + i] = v;
}

void launch_spectral_green_full(const double* T6, const double2* F, double2* V, long N, hipStream_t st) {
    hipLaunchKernelGGL(k_spectral_green_full, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, T6, F, V, N);
}

#define FFTPERM_INST(R)                                                                                       \
    template void launch_xfwd<R>(const XSource&, cx<R>*, const Geometry3&, const FFTTables<R>&, const int*,   \
                                 hipStream_t);                                                                \
    template long launch_xinv<R>(const cx<R>*, const XSink&, const Geometry3&, const FFTTables<R>&,           \
                                 const int*, hipStream_t);                                                    \
    template void launch_ypass<R>(cx<R>*, int, const Geometry3&, const FFTTables<R>&, const int*,             \
                                  hipStream_t);                                                               \
    template void launch_zpass<R>(cx<R>*, int, const Geometry3&, const FFTTables<R>&, const int*,             \
                                  hipStream_t);                                                               \
    template long launch_zgreen<R>(cx<R>*, const R*, const Geometry3&, const FFTTables<R>&, double*,          \
                                   const int*, hipStream_t);                                                  \
    template long launch_zgreen_batch<R>(cx<R>*, const R*, const Geometry3&, const FFTTables<R>&, double*,    \
                                         long, int, const int*, int, hipStream_t);                            \
    template void launch_ypass_n<R>(cx<R>*, int, const Geometry3&, const FFTTables<R>&, int, const int*, int, \
                                    hipStream_t);                                                             \
    template void launch_expand_full<R>(const cx<R>*, double2*, const Geometry3&, hipStream_t);               \
    template void launch_hermitian_half<R>(const double2*, cx<R>*, const Geometry3&, hipStream_t);

FFTPERM_INST(double)
FFTPERM_INST(float)

}  // namespace fftperm

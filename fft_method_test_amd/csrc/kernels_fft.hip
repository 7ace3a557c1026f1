// kernels_fft.hip -- the five sweeps of the Green-operator apply on the R2C half spectrum.
//
// Layout in HBM: real fields are SoA by component, flat index (z*ny + y)*nx + x
// (src/core/Grid3D.hpp:41-43).  The spectrum S holds three components of
// nz*ny*nxh complex values (nxh = nx/2 + 1), flat index (kz*ny + ky)*nxh + kx.
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

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace fftperm {

namespace {

size_t g_lds_max = 64 * 1024;  // per-workgroup LDS limit of the device (set at context creation)

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
size_t lds_limit() { return g_lds_max; }


// =====================================================================================
// Compile-time specialised passes for N in {32, 64, 128, 256, 512}.
// =====================================================================================
namespace {

template <typename R, int N>
struct Shape {
    static constexpr int T = threads_per_line<N>();                  // threads per line
    static constexpr int LD = N + 1;                                  // LDS pitch (complex values)
    static constexpr int LINES = (256 / T) > 0 ? (256 / T) : 1;       // x / column passes
    // twice the butterfly threads: the extra waves keep loads and stores in flight while other
    // workgroups of the CU are in their LDS stages
    static constexpr int THREADS = 2 * LINES * T <= 1024 ? 2 * LINES * T : LINES * T;
    static constexpr size_t LDS1 = (size_t)(N + LINES * LD) * sizeof(cx<R>);
    // fused z pass: three components of ZL columns in LDS, at most ~100 KB and 1024 threads
    static constexpr int zl_raw() {
        int z = 16;
        while (z > 1 && ((size_t)(3 * z * LD + N) * sizeof(cx<R>) > 102400u || 3 * z * T > 1024)) z /= 2;
        return z;
    }
    static constexpr int ZL = zl_raw();
    static constexpr int ZTHREADS = 3 * ZL * T;
    static constexpr size_t LDSZ = (size_t)(N + 3 * ZL * LD) * sizeof(cx<R>);
};

// Real x-lines are transformed two at a time as one complex line z = a + i b
// ("two for one"): round A packs components x and y of each line, round B packs component z of
// lines 2m and 2m+1.  The half spectra follow from A[k] = (Z[k] + conj Z[N-k]) / 2 and
// B[k] = (Z[k] - conj Z[N-k]) / 2i, which halves the FFT and LDS work of the R2C pass.
template <typename R, int N>
__global__ __launch_bounds__((Shape<R, N>::THREADS)) void k_xfwd_s(XSource src, cx<R>* __restrict__ S, long nlines,
                                                                     long N3, long Nh, const cx<R>* __restrict__ twg,
                                                                     const int* skip) {
    if (skip != nullptr && *skip) return;
    using SH = Shape<R, N>;
    constexpr int LINES = SH::LINES, LD = SH::LD, THREADS = SH::THREADS, NXH = N / 2 + 1;
    constexpr int KE = LINES * N / THREADS;  // elements per thread per component
    static_assert((LINES * N) % THREADS == 0 && LINES % 2 == 0, "element split");
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long line0 = xcd_tile(blockIdx.x, gridDim.x) * LINES;
    const int nl = (int)min((long)LINES, nlines - line0);
    // every global load of the block is issued up front: the support code once per voxel, then
    // the three compact components (or the three field components) of each owned voxel
    double v[3][KE];
    if (src.field != nullptr) {
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const int idx = threadIdx.x + k * THREADS;
            const int line = idx / N;
            const long vox = (line0 + line) * N + idx % N;
            const bool ok = line < nl;
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c][k] = ok ? src.field[c * N3 + vox] : 0.0;
        }
    } else {
        int code[KE];
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const int idx = threadIdx.x + k * THREADS;
            const int line = idx / N;
            code[k] = (line < nl) ? src.code[(line0 + line) * N + idx % N] : -2;
        }
        double mc[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) mc[c] = (src.mean != nullptr) ? src.mean[c] : 0.0;
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const int cd = code[k];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double f0 = (cd == -1) ? src.a_fluid[c] : ((cd >= 0) ? src.a_support[c] : 0.0);
                double fx = 0.0;
                if (src.compact != nullptr && cd >= 0) fx = src.compact[3L * cd + c] - mc[c];
                v[c][k] = f0 + fx;
            }
        }
    }
    // round A: z = x + i y per line
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        buf[(idx / N) * LD + idx % N] = mkc<R>((R)v[0][k], (R)v[1][k]);
    }
    __syncthreads();
    sfft_lines<R, -1, N, LINES, SH::T, LD>(buf, tw, nl);
    for (int idx = threadIdx.x; idx < LINES * NXH; idx += THREADS) {
        const int line = idx / NXH;
        const int kx = idx % NXH;
        if (line < nl) {
            const cx<R> zk = buf[line * LD + kx];
            const cx<R> zm = buf[line * LD + (N - kx) % N];
            const long o = (line0 + line) * NXH + kx;
            S[o] = mkc<R>((R)0.5 * (zk.x + zm.x), (R)0.5 * (zk.y - zm.y));
            S[Nh + o] = mkc<R>((R)0.5 * (zk.y + zm.y), (R)0.5 * (zm.x - zk.x));
        }
    }
    __syncthreads();
    // round B: z = z_{2m} + i z_{2m+1} (component z of two adjacent lines)
    R* bufr = reinterpret_cast<R*>(buf);
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int line = idx / N;
        bufr[2 * ((line >> 1) * LD + idx % N) + (line & 1)] = (R)v[2][k];
    }
    __syncthreads();
    const int nl2 = (nl + 1) / 2;
    sfft_lines<R, -1, N, LINES, SH::T, LD>(buf, tw, nl2);
    cx<R>* Sz = S + 2 * Nh;
    for (int idx = threadIdx.x; idx < (LINES / 2) * NXH; idx += THREADS) {
        const int m = idx / NXH;
        const int kx = idx % NXH;
        const cx<R> zk = buf[m * LD + kx];
        const cx<R> zm = buf[m * LD + (N - kx) % N];
        const int la = 2 * m, lb = 2 * m + 1;
        if (la < nl) Sz[(line0 + la) * NXH + kx] = mkc<R>((R)0.5 * (zk.x + zm.x), (R)0.5 * (zk.y - zm.y));
        if (lb < nl) Sz[(line0 + lb) * NXH + kx] = mkc<R>((R)0.5 * (zk.y + zm.y), (R)0.5 * (zm.x - zk.x));
    }
}

// Z[k] = A + iB and Z[N-k] = conj(A) + i conj(B) for 0 < k < N - k; the self-conjugate bins keep
// their real parts only.
template <typename R, int N, int LD>
__device__ __forceinline__ void xinv_put(cx<R>* buf, int line, int kx, cx<R> a, cx<R> b) {
    const bool self = (kx == 0 || 2 * kx == N);
    if (self) {
        a.y = (R)0;
        b.y = (R)0;
    }
    buf[line * LD + kx] = mkc<R>(a.x - b.y, a.y + b.x);
    if (!self) buf[line * LD + (N - kx)] = mkc<R>(a.x + b.y, b.x - a.y);
}

// C2R of two half spectra at once: Z[k] = A[k] + i B[k] over the Hermitian completions of A and
// B, then a = Re ifft(Z), b = Im ifft(Z).  The imaginary parts of the self-conjugate bins (0 and
// N/2) are dropped, which is what taking the real part of a full C2C inverse does
// (src/fft/FFTWBackend.cpp:292).
template <typename R, int N>
__global__ __launch_bounds__((Shape<R, N>::THREADS)) void k_xinv_s(const cx<R>* __restrict__ S, XSink sink, long nlines,
                                                                     long N3, long Nh, const cx<R>* __restrict__ twg,
                                                                     double inv_n, const int* skip) {
    if (skip != nullptr && *skip) return;
    using SH = Shape<R, N>;
    constexpr int LINES = SH::LINES, LD = SH::LD, THREADS = SH::THREADS, NXH = N / 2 + 1;
    constexpr int KE = LINES * N / THREADS;
    constexpr int KA = (LINES * NXH + THREADS - 1) / THREADS;        // round A loads per thread
    constexpr int KB = ((LINES / 2) * NXH + THREADS - 1) / THREADS;  // round B loads per thread
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long line0 = xcd_tile(blockIdx.x, gridDim.x) * LINES;
    const int nl = (int)min((long)LINES, nlines - line0);
    const cx<R> zero = mkc<R>((R)0, (R)0);
    // all global loads up front: support codes, the x and y half spectra, component z, then the v
    // values of the alpha dot product (these depend on the codes, so they go last)
    int code[KE];
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int line = idx / N;
        code[k] = (line < nl) ? sink.code[(line0 + line) * N + idx % N] : -2;
    }
    cx<R> ha[KA], hb[KA];
#pragma unroll
    for (int k = 0; k < KA; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int line = idx / NXH;
        const bool ok = idx < LINES * NXH && line < nl;
        const long o = (line0 + line) * NXH + idx % NXH;
        ha[k] = ok ? S[o] : zero;
        hb[k] = ok ? S[Nh + o] : zero;
    }
    cx<R> ga[KB], gb[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int m = idx / NXH;
        const int kx = idx % NXH;
        const bool in = idx < (LINES / 2) * NXH;
        ga[k] = (in && 2 * m < nl) ? S[2 * Nh + (line0 + 2 * m) * NXH + kx] : zero;
        gb[k] = (in && 2 * m + 1 < nl) ? S[2 * Nh + (line0 + 2 * m + 1) * NXH + kx] : zero;
    }
    double dv[3][KE];
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int cd = code[k];
#pragma unroll
        for (int c = 0; c < 3; ++c) dv[c][k] = (sink.dotv != nullptr && cd >= 0) ? sink.dotv[3L * cd + c] : 0.0;
    }
    double out[3][KE];
    // round A: components x and y
#pragma unroll
    for (int k = 0; k < KA; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        if (idx < LINES * NXH) xinv_put<R, N, LD>(buf, idx / NXH, idx % NXH, ha[k], hb[k]);
    }
    __syncthreads();
    sfft_lines<R, +1, N, LINES, SH::T, LD>(buf, tw, nl);
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const cx<R> z = buf[(idx / N) * LD + idx % N];
        out[0][k] = (double)z.x * inv_n;
        out[1][k] = (double)z.y * inv_n;
    }
    __syncthreads();
    // round B: component z of lines 2m (real part) and 2m+1 (imaginary part)
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        if (idx < (LINES / 2) * NXH) xinv_put<R, N, LD>(buf, idx / NXH, idx % NXH, ga[k], gb[k]);
    }
    __syncthreads();
    sfft_lines<R, +1, N, LINES, SH::T, LD>(buf, tw, (nl + 1) / 2);
    const R* bufr = reinterpret_cast<const R*>(buf);
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int line = idx / N;
        out[2][k] = (double)bufr[2 * ((line >> 1) * LD + idx % N) + (line & 1)] * inv_n;
    }
    double acc[kXinvParts] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int line = idx / N;
        if (line >= nl) continue;
        const long vox = (line0 + line) * N + idx % N;
        if (sink.field != nullptr) {
#pragma unroll
            for (int c = 0; c < 3; ++c) sink.field[c * N3 + vox] = out[c][k];
        }
        const int cd = code[k];
        if (cd >= 0) {
            if (sink.compact != nullptr) {
#pragma unroll
                for (int c = 0; c < 3; ++c) sink.compact[3L * cd + c] = out[c][k];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[6] += dv[c][k] * out[c][k];
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] += out[c][k];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[3 + c] += out[c][k];
    }
    if (sink.partials != nullptr) block_sum<kXinvParts>(acc, sink.partials + (long)kXinvParts * blockIdx.x);
}

// Column pass: column col has base (col / inner) * outer_stride + col % inner, element e at
// base + e * estride.  Thread t always serves column t % LINES in the load/store loops.
template <typename R, int N, int SIGN>
__global__ __launch_bounds__((Shape<R, N>::THREADS)) void k_col_s(cx<R>* __restrict__ S, long ncols, long inner,
                                                                  long outer_stride, long estride, long comp_stride,
                                                                  const cx<R>* __restrict__ twg, const int* skip) {
    if (skip != nullptr && *skip) return;
    using SH = Shape<R, N>;
    constexpr int LINES = SH::LINES, LD = SH::LD, THREADS = SH::THREADS;
    constexpr int EST = THREADS / LINES;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long col0 = xcd_tile(blockIdx.x, gridDim.x) * LINES;
    const int ncl = (int)min((long)LINES, ncols - col0);
    cx<R>* Sc = S + blockIdx.y * comp_stride;
    const int ci = threadIdx.x % LINES;
    const long col = col0 + ci;
    const unsigned cu = (unsigned)col, iu = (unsigned)inner;
    const long base = (long)(cu / iu) * outer_stride + (long)(cu % iu);
    const bool own = ci < ncl;
#pragma unroll 4
    for (int e = threadIdx.x / LINES; e < N; e += EST)
        if (own) buf[ci * LD + e] = Sc[base + e * estride];
    __syncthreads();
    sfft_lines<R, SIGN, N, LINES, SH::T, LD>(buf, tw, ncl);
#pragma unroll 4
    for (int e = threadIdx.x / LINES; e < N; e += EST)
        if (own) Sc[base + e * estride] = buf[ci * LD + e];
}

// Fused z pass: lines 0..3*ZL-1 of the LDS buffer hold components x | y | z of ZL columns.
// OS = threads per butterfly thread (1 or 2): extra threads only help the memory phases
template <typename R, int N, int ZL, int OS>
constexpr int zgreen_threads() {
    return OS * 3 * ZL * Shape<R, N>::T;
}

template <typename R, int N, int ZL, int OS>
__global__ __launch_bounds__((zgreen_threads<R, N, ZL, OS>())) void k_zgreen_s(cx<R>* __restrict__ S,
                                                                                 const R* __restrict__ G6, long ncols,
                                                                                 long Nh,
                                                                                 const cx<R>* __restrict__ twg,
                                                                                 const int* skip) {
    if (skip != nullptr && *skip) return;
    using SH = Shape<R, N>;
    constexpr int LD = SH::LD, THREADS = zgreen_threads<R, N, ZL, OS>(), L3 = 3 * ZL;
    constexpr int EST = THREADS / L3;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long col0 = xcd_tile(blockIdx.x, gridDim.x) * ZL;
    const int ncl = (int)min((long)ZL, ncols - col0);
    // loads: thread t serves line t % L3 = comp * ZL + ci
    const int lt = threadIdx.x % L3;
    const int comp = lt / ZL;
    const int ci = lt % ZL;
    const bool own = ci < ncl;
    const cx<R>* Sc = S + comp * Nh + col0 + ci;
#pragma unroll 4
    for (int e = threadIdx.x / L3; e < N; e += EST)
        buf[lt * LD + e] = own ? Sc[e * ncols] : mkc<R>((R)0, (R)0);
    __syncthreads();
    sfft_lines<R, -1, N, L3, SH::T, LD>(buf, tw, L3);
    // symmetric Green multiply (src/bignonnet/GreenOperator.cpp:84-92 with G_sym)
    {
        constexpr int MT = THREADS / ZL;
        const int mci = threadIdx.x % ZL;
        const bool mown = mci < ncl;
        const R* Gc = G6 + 6L * (col0 + mci);
#pragma unroll 2
        for (int e = threadIdx.x / ZL; e < N; e += MT) {
            if (!mown) continue;
            const R* gp = Gc + 6L * e * ncols;
            const R txx = gp[0], txy = gp[1], txz = gp[2];
            const R tyy = gp[3], tyz = gp[4], tzz = gp[5];
            cx<R>* px = buf + (0 * ZL + mci) * LD + e;
            cx<R>* py = buf + (1 * ZL + mci) * LD + e;
            cx<R>* pz = buf + (2 * ZL + mci) * LD + e;
            const cx<R> fx = *px, fy = *py, fz = *pz;
            *px = mkc<R>(txx * fx.x + txy * fy.x + txz * fz.x, txx * fx.y + txy * fy.y + txz * fz.y);
            *py = mkc<R>(txy * fx.x + tyy * fy.x + tyz * fz.x, txy * fx.y + tyy * fy.y + tyz * fz.y);
            *pz = mkc<R>(txz * fx.x + tyz * fy.x + tzz * fz.x, txz * fx.y + tyz * fy.y + tzz * fz.y);
        }
    }
    __syncthreads();
    sfft_lines<R, +1, N, L3, SH::T, LD>(buf, tw, L3);
    cx<R>* Sw = S + comp * Nh + col0 + ci;
#pragma unroll 4
    for (int e = threadIdx.x / L3; e < N; e += EST)
        if (own) Sw[e * ncols] = buf[lt * LD + e];
}

static_assert(Shape<double, 256>::LINES == Shape<float, 256>::LINES, "x-pass shape independent of precision");

template <typename R>
bool static_xfwd(int n, const XSource& src, cx<R>* S, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                 hipStream_t st) {
    const long nlines = (long)g.ny * g.nz;
#define FFTPERM_XF(NN)                                                                                          \
    case NN: {                                                                                                  \
        using SH = Shape<R, NN>;                                                                                \
        const long blocks = (nlines + SH::LINES - 1) / SH::LINES;                                              \
        hipLaunchKernelGGL((k_xfwd_s<R, NN>), dim3((unsigned)blocks), dim3(SH::THREADS), SH::LDS1, st, src, S,  \
                           nlines, g.N, g.Nh, t.twx, skip);                                                     \
        return true;                                                                                            \
    }
    switch (n) { FFTPERM_XF(32) FFTPERM_XF(64) FFTPERM_XF(128) FFTPERM_XF(256) FFTPERM_XF(512) default: return false; }
#undef FFTPERM_XF
}

template <typename R>
bool static_xinv(int n, const cx<R>* S, const XSink& sink, const Geometry3& g, const FFTTables<R>& t,
                 const int* skip, hipStream_t st) {
    const long nlines = (long)g.ny * g.nz;
    const double inv_n = 1.0 / (double)g.N;
#define FFTPERM_XI(NN)                                                                                          \
    case NN: {                                                                                                  \
        using SH = Shape<R, NN>;                                                                                \
        const long blocks = (nlines + SH::LINES - 1) / SH::LINES;                                              \
        hipLaunchKernelGGL((k_xinv_s<R, NN>), dim3((unsigned)blocks), dim3(SH::THREADS), SH::LDS1, st, S, sink, \
                           nlines, g.N, g.Nh, t.twx, inv_n, skip);                                              \
        return true;                                                                                            \
    }
    switch (n) { FFTPERM_XI(32) FFTPERM_XI(64) FFTPERM_XI(128) FFTPERM_XI(256) FFTPERM_XI(512) default: return false; }
#undef FFTPERM_XI
}

template <typename R>
bool static_col(int n, int sign, cx<R>* S, long ncols, long inner, long outer, long estride, long comp_stride,
                const cx<R>* tw, const int* skip, hipStream_t st) {
#define FFTPERM_CO(NN)                                                                                          \
    case NN: {                                                                                                  \
        using SH = Shape<R, NN>;                                                                                \
        const dim3 grid((unsigned)((ncols + SH::LINES - 1) / SH::LINES), 3);                                    \
        if (sign < 0)                                                                                           \
            hipLaunchKernelGGL((k_col_s<R, NN, -1>), grid, dim3(SH::THREADS), SH::LDS1, st, S, ncols, inner,    \
                               outer, estride, comp_stride, tw, skip);                                          \
        else                                                                                                    \
            hipLaunchKernelGGL((k_col_s<R, NN, +1>), grid, dim3(SH::THREADS), SH::LDS1, st, S, ncols, inner,    \
                               outer, estride, comp_stride, tw, skip);                                          \
        return true;                                                                                            \
    }
    switch (n) { FFTPERM_CO(32) FFTPERM_CO(64) FFTPERM_CO(128) FFTPERM_CO(256) FFTPERM_CO(512) default: return false; }
#undef FFTPERM_CO
}

template <typename R, int N, int ZL>
constexpr size_t zgreen_lds() {
    return (size_t)(N + 3 * ZL * Shape<R, N>::LD) * sizeof(cx<R>);
}

template <typename R, int N, int ZL>
bool zgreen_fits() {
    return zgreen_lds<R, N, ZL>() <= g_lds_max;
}

template <typename R, int N, int ZL, int OS>
bool launch_zgreen_zl(cx<R>* S, const R* G6, long ncols, long Nh, const cx<R>* tw, const int* skip, hipStream_t st) {
    if constexpr (zgreen_threads<R, N, ZL, OS>() > 1024) {
        return false;
    } else {
        if (!zgreen_fits<R, N, ZL>()) return false;
        const long blocks = (ncols + ZL - 1) / ZL;
        constexpr int threads = zgreen_threads<R, N, ZL, OS>();
        constexpr size_t lds = zgreen_lds<R, N, ZL>();
        hipLaunchKernelGGL((k_zgreen_s<R, N, ZL, OS>), dim3((unsigned)blocks), dim3(threads), lds, st, S, G6, ncols,
                           Nh, tw, skip);
        return true;
    }
}

// Autotuned choice of the fused z-pass tile width ZL (columns per workgroup), keyed by
// (precision, N, columns); filled by autotune_zgreen, read by every launch.
std::mutex g_tune_mu;
std::map<std::tuple<int, int, long>, int> g_zl_choice;

int zl_choice(int prec, int n, long ncols) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_zl_choice.find(std::make_tuple(prec, n, ncols));
    return it == g_zl_choice.end() ? 0 : it->second;
}

// variant code = 10 * ZL + OS
template <typename R, int N>
bool launch_zgreen_n(int code, cx<R>* S, const R* G6, long ncols, long Nh, const cx<R>* tw, const int* skip,
                     hipStream_t st) {
    switch (code) {
        case 11: return launch_zgreen_zl<R, N, 1, 1>(S, G6, ncols, Nh, tw, skip, st);
        case 12: return launch_zgreen_zl<R, N, 1, 2>(S, G6, ncols, Nh, tw, skip, st);
        case 21: return launch_zgreen_zl<R, N, 2, 1>(S, G6, ncols, Nh, tw, skip, st);
        case 22: return launch_zgreen_zl<R, N, 2, 2>(S, G6, ncols, Nh, tw, skip, st);
        case 41: return launch_zgreen_zl<R, N, 4, 1>(S, G6, ncols, Nh, tw, skip, st);
        case 42: return launch_zgreen_zl<R, N, 4, 2>(S, G6, ncols, Nh, tw, skip, st);
        case 81: return launch_zgreen_zl<R, N, 8, 1>(S, G6, ncols, Nh, tw, skip, st);
        case 82: return launch_zgreen_zl<R, N, 8, 2>(S, G6, ncols, Nh, tw, skip, st);
        case 161: return launch_zgreen_zl<R, N, 16, 1>(S, G6, ncols, Nh, tw, skip, st);
        case 162: return launch_zgreen_zl<R, N, 16, 2>(S, G6, ncols, Nh, tw, skip, st);
        default: return false;
    }
}

template <typename R>
bool static_zgreen_zl(int n, int zl, cx<R>* S, const R* G6, long ncols, long Nh, const cx<R>* tw, const int* skip,
                      hipStream_t st) {
    switch (n) {
        case 32: return launch_zgreen_n<R, 32>(zl, S, G6, ncols, Nh, tw, skip, st);
        case 64: return launch_zgreen_n<R, 64>(zl, S, G6, ncols, Nh, tw, skip, st);
        case 128: return launch_zgreen_n<R, 128>(zl, S, G6, ncols, Nh, tw, skip, st);
        case 256: return launch_zgreen_n<R, 256>(zl, S, G6, ncols, Nh, tw, skip, st);
        case 512: return launch_zgreen_n<R, 512>(zl, S, G6, ncols, Nh, tw, skip, st);
        default: return false;
    }
}

template <typename R>
bool static_zgreen(int n, cx<R>* S, const R* G6, long ncols, long Nh, const cx<R>* tw, const int* skip,
                   hipStream_t st) {
    int code = zl_choice(sizeof(R) == 8 ? 0 : 1, n, ncols);
    if (code == 0) {
        // untuned: the widest tile that fits ~100 KB of LDS, one thread per butterfly
        switch (n) {
            case 32: code = 10 * Shape<R, 32>::ZL + 1; break;
            case 64: code = 10 * Shape<R, 64>::ZL + 1; break;
            case 128: code = 10 * Shape<R, 128>::ZL + 1; break;
            case 256: code = 10 * Shape<R, 256>::ZL + 1; break;
            case 512: code = 10 * Shape<R, 512>::ZL + 1; break;
            default: return false;
        }
    }
    return static_zgreen_zl<R>(n, code, S, G6, ncols, Nh, tw, skip, st);
}

// The specialised kernels need their static LDS budget; fall back to the runtime kernels on
// devices with less LDS per workgroup.
bool static_ok() { return g_lds_max >= 102400u + 8192u; }

}  // namespace

template <typename R>
std::vector<std::pair<int, float>> autotune_zgreen(cx<R>* S, const R* G6, const Geometry3& g, const FFTTables<R>& t,
                                                   hipStream_t st) {
    std::vector<std::pair<int, float>> timings;
    const long ncols = (long)g.ny * g.nxh;
    const int prec = sizeof(R) == 8 ? 0 : 1;
    if (!static_ok() || zl_choice(prec, g.nz, ncols) != 0) return timings;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return timings;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return timings;
    }
    (void)hipMemsetAsync(S, 0, 3 * (size_t)g.Nh * sizeof(cx<R>), st);
    int best = 0;
    float best_ms = 1e30f;
    for (int zl : {11, 12, 21, 22, 41, 42, 81, 82, 161, 162}) {
        if (!static_zgreen_zl<R>(g.nz, zl, S, G6, ncols, g.Nh, t.twz, nullptr, st)) continue;  // warm-up
        (void)hipEventRecord(a, st);
        for (int rep = 0; rep < 3; ++rep) static_zgreen_zl<R>(g.nz, zl, S, G6, ncols, g.Nh, t.twz, nullptr, st);
        (void)hipEventRecord(b, st);
        if (hipEventSynchronize(b) != hipSuccess) break;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        ms /= 3.f;
        timings.emplace_back(zl, ms);
        if (ms < best_ms) {
            best_ms = ms;
            best = zl;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (best != 0) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_zl_choice[std::make_tuple(prec, g.nz, ncols)] = best;
    }
    return timings;
}

int zgreen_tile(int prec, int n, long ncols) { return zl_choice(prec, n, ncols); }

template std::vector<std::pair<int, float>> autotune_zgreen<double>(cx<double>*, const double*, const Geometry3&,
                                                                    const FFTTables<double>&, hipStream_t);
template std::vector<std::pair<int, float>> autotune_zgreen<float>(cx<float>*, const float*, const Geometry3&,
                                                                   const FFTTables<float>&, hipStream_t);

long xinv_blocks(const Geometry3& g, const LinePlan& px) {
    // exactly the number of blocks (= partial sums) the x-inverse launcher will use
    const long nlines = (long)g.ny * g.nz;
    if (static_ok()) {
        switch (g.nx) {
            case 32: return (nlines + Shape<double, 32>::LINES - 1) / Shape<double, 32>::LINES;
            case 64: return (nlines + Shape<double, 64>::LINES - 1) / Shape<double, 64>::LINES;
            case 128: return (nlines + Shape<double, 128>::LINES - 1) / Shape<double, 128>::LINES;
            case 256: return (nlines + Shape<double, 256>::LINES - 1) / Shape<double, 256>::LINES;
            case 512: return (nlines + Shape<double, 512>::LINES - 1) / Shape<double, 512>::LINES;
            default: break;
        }
    }
    return cfg_xpass(px, nlines).blocks;
}

// ------------------------------------------------------------------ x forward
template <typename R>
__global__ void k_xfwd(XSource src, cx<R>* __restrict__ S, int nx, long nlines, int nxh, long N, long Nh,
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
                val = f0 + fx;
            }
            buf[line * ld + i] = mkc<R>((R)val, (R)0);
        }
        __syncthreads();
        fft_lines<R, -1>(buf, ld, nl, pl, tw);
        const int cnth = nl * nxh;
        cx<R>* Sc = S + c * Nh;
        for (int idx = threadIdx.x; idx < cnth; idx += blockDim.x) {
            const int line = idx / nxh;
            const int kx = idx - line * nxh;
            Sc[(line0 + line) * nxh + kx] = buf[line * ld + kx];
        }
        __syncthreads();
    }
}

template <typename R>
void launch_xfwd(const XSource& src, cx<R>* S, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                 hipStream_t st) {
    if (static_ok() && static_xfwd<R>(g.nx, src, S, g, t, skip, st)) return;
    const long nlines = (long)g.ny * g.nz;
    PassCfg c = cfg_xpass(t.px, nlines);
    hipLaunchKernelGGL(k_xfwd<R>, dim3((unsigned)c.blocks), dim3(c.threads), c.lds, st, src, S, g.nx, nlines,
                       g.nxh, g.N, g.Nh, t.px, t.twx, c.lines, skip);
}

// ------------------------------------------------------------------ x inverse
template <typename R>
__global__ void k_xinv(const cx<R>* __restrict__ S, XSink sink, int nx, long nlines, int nxh, long N, long Nh,
                       LinePlan pl, const cx<R>* __restrict__ twg, int lpw, double inv_n, const int* skip) {
    if (skip != nullptr && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + nx;
    const int ld = nx + 1;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) tw[i] = twg[i];
    const long line0 = (long)blockIdx.x * lpw;
    const int nl = (int)min((long)lpw, nlines - line0);
    double acc[kXinvParts] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 3; ++c) {
        const cx<R>* Sc = S + c * Nh;
        const int cnth = nl * nxh;
        for (int idx = threadIdx.x; idx < cnth; idx += blockDim.x) {
            const int line = idx / nxh;
            const int kx = idx - line * nxh;
            buf[line * ld + kx] = Sc[(line0 + line) * nxh + kx];
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
                if (sink.dotv != nullptr) acc[6] += sink.dotv[3L * code + c] * val;
                if (c == 0) acc[0] += val;
                if (c == 1) acc[1] += val;
                if (c == 2) acc[2] += val;
            }
            if (c == 0) acc[3] += val;
            if (c == 1) acc[4] += val;
            if (c == 2) acc[5] += val;
        }
        __syncthreads();
    }
    if (sink.partials != nullptr) {
        block_sum<kXinvParts>(acc, sink.partials + (long)kXinvParts * blockIdx.x);
    }
}

template <typename R>
void launch_xinv(const cx<R>* S, const XSink& sink, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                 hipStream_t st) {
    if (static_ok() && static_xinv<R>(g.nx, S, sink, g, t, skip, st)) return;
    const long nlines = (long)g.ny * g.nz;
    PassCfg c = cfg_xpass(t.px, nlines);
    const double inv_n = 1.0 / (double)g.N;
    hipLaunchKernelGGL(k_xinv<R>, dim3((unsigned)c.blocks), dim3(c.threads), c.lds, st, S, sink, g.nx, nlines,
                       g.nxh, g.N, g.Nh, t.px, t.twx, c.lines, inv_n, skip);
}

// ------------------------------------------------------------- column passes
// Column col of a component: base = (col / inner) * outer_stride + col % inner,
// element e at base + e * estride.
template <typename R, int SIGN>
__global__ void k_colpass(cx<R>* __restrict__ S, int n, long ncols, long inner, long outer_stride, long estride,
                          long comp_stride, LinePlan pl, const cx<R>* __restrict__ twg, int tl,
                          const int* skip) {
    if (skip != nullptr && *skip) return;
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
    const long ncols = (long)g.nz * g.nxh;
    if (static_ok() && static_col<R>(g.ny, sign, S, ncols, (long)g.nxh, (long)g.ny * g.nxh, (long)g.nxh, g.Nh, t.twy,
                                     skip, st))
        return;
    PassCfg c = cfg_colpass(t.py, ncols);
    dim3 grid((unsigned)c.blocks, 3);
    if (sign < 0)
        hipLaunchKernelGGL((k_colpass<R, -1>), grid, dim3(c.threads), c.lds, st, S, g.ny, ncols, (long)g.nxh,
                           (long)g.ny * g.nxh, (long)g.nxh, g.Nh, t.py, t.twy, c.lines, skip);
    else
        hipLaunchKernelGGL((k_colpass<R, +1>), grid, dim3(c.threads), c.lds, st, S, g.ny, ncols, (long)g.nxh,
                           (long)g.ny * g.nxh, (long)g.nxh, g.Nh, t.py, t.twy, c.lines, skip);
}

template <typename R>
void launch_zpass(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                  hipStream_t st) {
    const long ncols = (long)g.ny * g.nxh;
    if (static_ok() && static_col<R>(g.nz, sign, S, ncols, ncols, 0L, ncols, g.Nh, t.twz, skip, st)) return;
    PassCfg c = cfg_colpass(t.pz, ncols);
    dim3 grid((unsigned)c.blocks, 3);
    if (sign < 0)
        hipLaunchKernelGGL((k_colpass<R, -1>), grid, dim3(c.threads), c.lds, st, S, g.nz, ncols, ncols, 0L,
                           ncols, g.Nh, t.pz, t.twz, c.lines, skip);
    else
        hipLaunchKernelGGL((k_colpass<R, +1>), grid, dim3(c.threads), c.lds, st, S, g.nz, ncols, ncols, 0L,
                           ncols, g.Nh, t.pz, t.twz, c.lines, skip);
}

// ------------------------------------------------------------- fused z + Green
template <typename R>
__global__ void k_zgreen(cx<R>* __restrict__ S, const R* __restrict__ G6, int nz, long ncols, long Nh,
                         LinePlan pl, const cx<R>* __restrict__ twg, int tl, const int* skip) {
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
        const cx<R>* Sc = S + c * Nh;
        cx<R>* B = buf + c * ncl * ld;
        for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
            const int e = idx / ncl;
            const int ci = idx - e * ncl;
            B[ci * ld + e] = Sc[col0 + ci + e * ncols];
        }
    }
    __syncthreads();
    fft_lines<R, -1>(buf, ld, 3 * ncl, pl, tw);
    // Symmetric Green multiply, as GreenOperator::apply_spectral_green_tensor
    // (src/bignonnet/GreenOperator.cpp:84-92) with the symmetrised tensor.
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
        const int e = idx / ncl;
        const int ci = idx - e * ncl;
        const R* gp = G6 + 6L * (col0 + ci + e * ncols);
        const R txx = gp[0], txy = gp[1], txz = gp[2];
        const R tyy = gp[3], tyz = gp[4], tzz = gp[5];
        cx<R>* px = buf + (0 * ncl + ci) * ld + e;
        cx<R>* py = buf + (1 * ncl + ci) * ld + e;
        cx<R>* pz = buf + (2 * ncl + ci) * ld + e;
        const cx<R> fx = *px, fy = *py, fz = *pz;
        *px = mkc<R>(txx * fx.x + txy * fy.x + txz * fz.x, txx * fx.y + txy * fy.y + txz * fz.y);
        *py = mkc<R>(txy * fx.x + tyy * fy.x + tyz * fz.x, txy * fx.y + tyy * fy.y + tyz * fz.y);
        *pz = mkc<R>(txz * fx.x + tyz * fy.x + tzz * fz.x, txz * fx.y + tyz * fy.y + tzz * fz.y);
    }
    __syncthreads();
    fft_lines<R, +1>(buf, ld, 3 * ncl, pl, tw);
    for (int c = 0; c < 3; ++c) {
        cx<R>* Sc = S + c * Nh;
        const cx<R>* B = buf + c * ncl * ld;
        for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
            const int e = idx / ncl;
            const int ci = idx - e * ncl;
            Sc[col0 + ci + e * ncols] = B[ci * ld + e];
        }
    }
}

template <typename R>
void launch_zgreen(cx<R>* S, const R* G6, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                   hipStream_t st) {
    const long ncols = (long)g.ny * g.nxh;
    if (static_ok() && static_zgreen<R>(g.nz, S, G6, ncols, g.Nh, t.twz, skip, st)) return;
    PassCfg c = cfg_zgreen(t.pz, ncols);
    hipLaunchKernelGGL(k_zgreen<R>, dim3((unsigned)c.blocks), dim3(c.threads), c.lds, st, S, G6, g.nz, ncols,
                       g.Nh, t.pz, t.twz, c.lines, skip);
}

// ------------------------------------------------------- spectrum utilities
template <typename R>
__global__ void k_expand_full(const cx<R>* __restrict__ S, double2* __restrict__ F, int nx, int ny, int nz,
                              int nxh, long N, long Nh) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int c = blockIdx.y;
    const int kx = (int)(i % nx);
    const long r = i / nx;
    const int ky = (int)(r % ny);
    const int kz = (int)(r / ny);
    double2 v;
    if (kx < nxh) {
        const cx<R> s = S[c * Nh + ((long)kz * ny + ky) * nxh + kx];
        v.x = (double)s.x;
        v.y = (double)s.y;
    } else {
        const int mz = (nz - kz) % nz, my = (ny - ky) % ny, mx = nx - kx;
        const cx<R> s = S[c * Nh + ((long)mz * ny + my) * nxh + mx];
        v.x = (double)s.x;
        v.y = -(double)s.y;
    }
    F[c * N + i] = v;
}

template <typename R>
void launch_expand_full(const cx<R>* S, double2* F, const Geometry3& g, hipStream_t st) {
    dim3 grid((unsigned)((g.N + 255) / 256), 3);
    hipLaunchKernelGGL(k_expand_full<R>, grid, dim3(256), 0, st, S, F, g.nx, g.ny, g.nz, g.nxh, g.N, g.Nh);
}

template <typename R>
__global__ void k_hermitian_half(const double2* __restrict__ X, cx<R>* __restrict__ S, int nx, int ny, int nz,
                                 int nxh, long N, long Nh) {
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
    S[c * Nh + i] = mkc<R>((R)(0.5 * (a.x + b.x)), (R)(0.5 * (a.y - b.y)));
}

template <typename R>
void launch_hermitian_half(const double2* X, cx<R>* S, const Geometry3& g, hipStream_t st) {
    dim3 grid((unsigned)((g.Nh + 255) / 256), 3);
    hipLaunchKernelGGL(k_hermitian_half<R>, grid, dim3(256), 0, st, X, S, g.nx, g.ny, g.nz, g.nxh, g.N, g.Nh);
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
    template void launch_xinv<R>(const cx<R>*, const XSink&, const Geometry3&, const FFTTables<R>&,           \
                                 const int*, hipStream_t);                                                    \
    template void launch_ypass<R>(cx<R>*, int, const Geometry3&, const FFTTables<R>&, const int*,             \
                                  hipStream_t);                                                               \
    template void launch_zpass<R>(cx<R>*, int, const Geometry3&, const FFTTables<R>&, const int*,             \
                                  hipStream_t);                                                               \
    template void launch_zgreen<R>(cx<R>*, const R*, const Geometry3&, const FFTTables<R>&, const int*,       \
                                   hipStream_t);                                                              \
    template void launch_expand_full<R>(const cx<R>*, double2*, const Geometry3&, hipStream_t);               \
    template void launch_hermitian_half<R>(const double2*, cx<R>*, const Geometry3&, hipStream_t);

FFTPERM_INST(double)
FFTPERM_INST(float)

}  // namespace fftperm

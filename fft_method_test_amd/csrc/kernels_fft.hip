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
#include "kernels.hpp"

#include <algorithm>

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

long xinv_blocks(const Geometry3& g, const LinePlan& px) {
    return cfg_xpass(px, (long)g.ny * g.nz).blocks;
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
    double acc[6] = {0, 0, 0, 0, 0, 0};
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
                acc[c] += val;
            }
            acc[3 + c] += val;
        }
        __syncthreads();
    }
    if (sink.partials != nullptr) {
        block_sum<6>(acc, sink.partials + 6L * blockIdx.x);
    }
}

template <typename R>
void launch_xinv(const cx<R>* S, const XSink& sink, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                 hipStream_t st) {
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
        const long gi = col0 + ci + e * ncols;
        const R txx = G6[gi], txy = G6[Nh + gi], txz = G6[2 * Nh + gi];
        const R tyy = G6[3 * Nh + gi], tyz = G6[4 * Nh + gi], tzz = G6[5 * Nh + gi];
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

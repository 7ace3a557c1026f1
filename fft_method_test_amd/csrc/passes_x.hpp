// passes_x.hpp -- compile-time specialised x passes (R2C with fused scatter, C2R with fused
// gather and partial sums) and their launchers, in kPassVariants tile shapes per x-line length
// (lines per workgroup x threads per butterfly thread) plus the merged-round and wave-per-line-pair
// shapes.  Included by passes_x.hip (power-of-two lengths) and passes_x_mixed.hip (mixed-radix
// lengths), so that the two sets compile in parallel.
//
// Real x-lines are transformed two at a time as one complex line z = a + i b ("two for one"):
// round A packs components x and y of each line, round B packs component z of lines 2m and
// 2m+1.  Forward: A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = (Z[k] - conj Z[N-k]) / 2i.  Inverse:
// Z = A + iB over the Hermitian completions of A and B, a = Re ifft(Z), b = Im ifft(Z).  This
// halves the FFT and LDS work of both x passes against one complex transform per real line.
#pragma once

#include "fft_static.hpp"
#include "kernels.hpp"

namespace fftperm {

namespace {

// LDS after the transform buffer: the tile's support words, their in-line prefix counts and the
// line starts (line-bitmask path)
template <int N, int LINES>
constexpr size_t bits_lds() {
    return (size_t)LINES * (2 * ((N + 31) / 32) + 1) * sizeof(int);
}

// The tile's support words and line starts, in two phases so that kernels can issue other loads in
// between: fetch the thread's share into registers first (waiting for them then does not wait for
// loads issued later), then store them to LDS (sb: LINES*W words, sl: LINES starts) and compute
// the in-line word prefixes (sp: LINES*W); store() returns after a barrier.
template <int N, int LINES, int THREADS>
struct TileBitsRegs {
    static constexpr int W = (N + 31) / 32;
    static constexpr int PW = (LINES * W + THREADS - 1) / THREADS;
    uint32_t word[PW];
    int start;
    __device__ __forceinline__ void fetch(const LineBits& lb, long line0, int nl) {
#pragma unroll
        for (int q = 0; q < PW; ++q) {
            const int t = threadIdx.x + q * THREADS;
            const int line = t / W;
            word[q] = (t < LINES * W && line < nl) ? lb.bits[(line0 + line) * W + t % W] : 0u;
        }
        start = ((int)threadIdx.x < nl) ? lb.lstart[line0 + threadIdx.x] : 0;
    }
    __device__ __forceinline__ void store(uint32_t* sb, int* sp, int* sl) const {
#pragma unroll
        for (int q = 0; q < PW; ++q) {
            const int t = threadIdx.x + q * THREADS;
            if (t < LINES * W) sb[t] = word[q];
        }
        if ((int)threadIdx.x < LINES) sl[threadIdx.x] = start;
        __syncthreads();
        for (int t = threadIdx.x; t < LINES; t += THREADS) {
            int acc = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                sp[t * W + w] = acc;
                acc += __popc(sb[t * W + w]);
            }
        }
        __syncthreads();
    }
};

// support index of voxel i of tile line `line` (-1 off the support)
template <int N>
__device__ __forceinline__ int tile_local(const uint32_t* sb, const int* sp, const int* sl, int line, int i) {
    constexpr int W = (N + 31) / 32;
    const int w = i >> 5, b = i & 31;
    const uint32_t word = sb[line * W + w];
    if (((word >> b) & 1u) == 0u) return -1;
    return sl[line] + sp[line * W + w] + __popc(word & ((1u << b) - 1u));
}

// The three source values of the tile's LINES * N voxels, v[c][k] for element
// idx = threadIdx.x + k * THREADS (tile line idx / N, voxel idx % N; zero past the tile).  Every
// global load of the block is issued up front: the support words (or codes) once per voxel, then
// the three compact components (or the three field components) of each owned voxel.  With the
// line bitmask and no background force, the tile's support indices form one contiguous run
// [cs, ce) of the compact vector: it is staged through `stage` (cap doubles of LDS, the transform
// buffer) with coalesced loads, then each voxel reads its three values from there; runs larger
// than the stage fall back to direct gathers.  `bits`: LDS for the tile's support words.  Returns
// after a barrier once the stage may be overwritten.
template <int N, int LINES, int THREADS, int KE>
__device__ __forceinline__ void xfwd_load(const XSource& src, long line0, int nl, long nlines, long N3, double* stage,
                                          long cap, unsigned char* bits, double (&v)[3][KE]) {
    const bool no_f0 = src.a_fluid[0] == 0.0 && src.a_fluid[1] == 0.0 && src.a_fluid[2] == 0.0 &&
                       src.a_support[0] == 0.0 && src.a_support[1] == 0.0 && src.a_support[2] == 0.0;
    if (src.field == nullptr && src.lines.bits != nullptr && no_f0) {
        constexpr int W = (N + 31) / 32;
        uint32_t* sb = reinterpret_cast<uint32_t*>(bits);
        int* sp = reinterpret_cast<int*>(sb + LINES * W);
        int* sl = sp + LINES * W;
        // the support words are requested first, so their latency overlaps the compact-run loads
        TileBitsRegs<N, LINES, THREADS> tb;
        tb.fetch(src.lines, line0, nl);
        const long cs = src.lines.lstart[line0];
        const long ce = (line0 + nl < nlines) ? (long)src.lines.lstart[line0 + nl] : src.lines.nb;
        const long cnt3 = 3 * (ce - cs);
        const bool staged = cnt3 <= cap;
        const bool has = src.compact != nullptr || src.compact32 != nullptr;
        if (staged && src.compact != nullptr) {
            const double* cc = src.compact + 3 * cs;
#pragma unroll 4
            for (long e = threadIdx.x; e < cnt3; e += THREADS) stage[e] = cc[e];
        } else if (staged && src.compact32 != nullptr) {
            const float* cc = src.compact32 + 3 * cs;
#pragma unroll 4
            for (long e = threadIdx.x; e < cnt3; e += THREADS) stage[e] = (double)cc[e];
        }
        tb.store(sb, sp, sl);
        double mc[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) mc[c] = (src.mean != nullptr) ? src.mean[c] : 0.0;
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const int idx = threadIdx.x + k * THREADS;
            const int line = idx / N;
            const int l = (idx < LINES * N && line < nl && has) ? tile_local<N>(sb, sp, sl, line, idx % N) : -1;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double raw = 0.0;
                if (l >= 0)
                    raw = staged ? stage[3 * (l - cs) + c]
                                 : (src.compact != nullptr ? src.compact[3L * l + c]
                                                           : (double)src.compact32[3L * l + c]);
                v[c][k] = (l < 0) ? 0.0 : raw - mc[c];
            }
        }
        __syncthreads();  // the stage is the transform buffer
    } else if (src.field != nullptr) {
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const int idx = threadIdx.x + k * THREADS;
            const int line = idx / N;
            const long vox = (line0 + line) * N + idx % N;
            const bool ok = idx < LINES * N && line < nl;
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c][k] = ok ? src.field[c * N3 + vox] : 0.0;
        }
    } else {
        int code[KE];
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const int idx = threadIdx.x + k * THREADS;
            const int line = idx / N;
            code[k] = (idx < LINES * N && line < nl) ? src.code[(line0 + line) * N + idx % N] : -2;
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
                if (src.compact32 != nullptr && cd >= 0) fx = (double)src.compact32[3L * cd + c] - mc[c];
                v[c][k] = f0 + fx;
            }
        }
    }
}

template <typename R, int N, int LINES, int OS>
__global__ __launch_bounds__((PassShape<R, N, LINES, OS>::THREADS)) void k_xfwd_s(XSource src, cx<R>* __restrict__ S,
                                                                                long nlines, long N3, SpecLayout SL,
                                                                                const cx<R>* __restrict__ twg,
                                                                                const int* skip) {
    if (skip != nullptr && *skip) return;
    using SH = PassShape<R, N, LINES, OS>;
    constexpr int LD = SH::LD, THREADS = SH::THREADS, NXH = N / 2 + 1;
    constexpr int KE = (LINES * N + THREADS - 1) / THREADS;  // elements per thread per component
    constexpr bool EXACT = (LINES * N) % THREADS == 0;        // else the last slot idles in some threads
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long line0 = xcd_tile(blockIdx.x, gridDim.x) * LINES;
    const int nl = (int)min((long)LINES, nlines - line0);
    double v[3][KE];
    constexpr long CAP = (long)(LINES * LD * sizeof(cx<R>) / sizeof(double));
    xfwd_load<N, LINES, THREADS, KE>(src, line0, nl, nlines, N3, reinterpret_cast<double*>(buf), CAP,
                                     reinterpret_cast<unsigned char*>(buf + LINES * LD), v);
    // round A: z = x + i y per line
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        if (EXACT || idx < LINES * N) buf[(idx / N) * LD + idx % N] = mkc<R>((R)v[0][k], (R)v[1][k]);
    }
    __syncthreads();
    sfft_lines<R, -1, N, LINES, SH::T, LD>(buf, tw, nl);
    for (int idx = threadIdx.x; idx < LINES * NXH; idx += THREADS) {
        const int line = idx / NXH;
        const int kx = idx % NXH;
        if (line < nl) {
            const cx<R> zk = buf[line * LD + kx];
            const cx<R> zm = buf[line * LD + (N - kx) % N];
            const long o = spec_line(SL, line0 + line) + kx;
            S[o] = mkc<R>((R)0.5 * (zk.x + zm.x), (R)0.5 * (zk.y - zm.y));
            S[SL.nc + o] = mkc<R>((R)0.5 * (zk.y + zm.y), (R)0.5 * (zm.x - zk.x));
        }
    }
    __syncthreads();
    // round B: z = z_{2m} + i z_{2m+1} (component z of two adjacent lines)
    R* bufr = reinterpret_cast<R*>(buf);
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int line = idx / N;
        if (EXACT || idx < LINES * N) bufr[2 * ((line >> 1) * LD + idx % N) + (line & 1)] = (R)v[2][k];
    }
    __syncthreads();
    const int nl2 = (nl + 1) / 2;
    sfft_lines<R, -1, N, LINES, SH::T, LD>(buf, tw, nl2);
    cx<R>* Sz = S + 2 * SL.nc;
    for (int idx = threadIdx.x; idx < (LINES / 2) * NXH; idx += THREADS) {
        const int m = idx / NXH;
        const int kx = idx % NXH;
        const cx<R> zk = buf[m * LD + kx];
        const cx<R> zm = buf[m * LD + (N - kx) % N];
        const int la = 2 * m, lb = 2 * m + 1;
        if (la < nl) Sz[spec_line(SL, line0 + la) + kx] = mkc<R>((R)0.5 * (zk.x + zm.x), (R)0.5 * (zk.y - zm.y));
        if (lb < nl) Sz[spec_line(SL, line0 + lb) + kx] = mkc<R>((R)0.5 * (zk.y + zm.y), (R)0.5 * (zm.x - zk.x));
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

// Writes one real output value of component c at voxel vox / support index cd and accumulates
// the support sum of c.
__device__ __forceinline__ void xinv_emit(const XSink& sink, long N3, int c, long vox, int cd, double val,
                                          double (&acc)[kXinvParts]) {
    if (sink.field != nullptr) sink.field[c * N3 + vox] = val;
    if (cd >= 0) {
        if (sink.compact != nullptr) sink.compact[3L * cd + c] = val;
        if (sink.compact32 != nullptr) sink.compact32[3L * cd + c] = (float)val;
        acc[c] += val;
    }
}

// C2R of two half spectra at once.  The imaginary parts of the self-conjugate bins (0 and N/2)
// are dropped, which is what taking the real part of a full C2C inverse does
// (src/fft/FFTWBackend.cpp:292).  Components x and y are written out as soon as round A has
// produced them, so only the codes and the round-B inputs stay in registers across a transform.
template <typename R, int N, int LINES, int OS>
__global__ __launch_bounds__((PassShape<R, N, LINES, OS>::THREADS)) void k_xinv_s(const cx<R>* __restrict__ S,
                                                                                XSink sink, long nlines, long N3,
                                                                                SpecLayout SL,
                                                                                const cx<R>* __restrict__ twg,
                                                                                double inv_n, const int* skip) {
    if (skip != nullptr && *skip) return;
    using SH = PassShape<R, N, LINES, OS>;
    constexpr int LD = SH::LD, THREADS = SH::THREADS, NXH = N / 2 + 1;
    constexpr int KE = (LINES * N + THREADS - 1) / THREADS;
    constexpr bool EXACT = (LINES * N) % THREADS == 0;
    constexpr int KA = (LINES * NXH + THREADS - 1) / THREADS;        // round A loads per thread
    constexpr int KB = ((LINES / 2) * NXH + THREADS - 1) / THREADS;  // round B loads per thread
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long line0 = xcd_tile(blockIdx.x, gridDim.x) * LINES;
    const int nl = (int)min((long)LINES, nlines - line0);
    // all global loads up front: support codes (or the tile's support words), the x and y half
    // spectra, then component z
    const bool use_bits = sink.lines.bits != nullptr;
    TileBitsRegs<N, LINES, THREADS> tb;
    if (use_bits) tb.fetch(sink.lines, line0, nl);
    int code[KE];
    if (!use_bits) {
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const int idx = threadIdx.x + k * THREADS;
            const int line = idx / N;  // >= LINES > nl - 1 for an idle slot
            code[k] = (line < nl) ? sink.code[(line0 + line) * N + idx % N] : -2;
        }
    }
    cx<R> ha[KA], hb[KA];
#pragma unroll
    for (int k = 0; k < KA; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int line = idx / NXH;
        const bool ok = idx < LINES * NXH && line < nl;
        const long o = spec_line(SL, line0 + line) + idx % NXH;
        // predicated loads into zeroed registers (a select between the spectrum and a zero
        // constant in memory would become a flat load through a private stack slot)
        ha[k] = mkc<R>((R)0, (R)0);
        hb[k] = mkc<R>((R)0, (R)0);
        if (ok) {
            ha[k] = S[o];
            hb[k] = S[SL.nc + o];
        }
    }
    cx<R> ga[KB], gb[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int m = idx / NXH;
        const int kx = idx % NXH;
        const bool in = idx < (LINES / 2) * NXH;
        ga[k] = mkc<R>((R)0, (R)0);
        gb[k] = mkc<R>((R)0, (R)0);
        if (in && 2 * m < nl) ga[k] = S[2 * SL.nc + spec_line(SL, line0 + 2 * m) + kx];
        if (in && 2 * m + 1 < nl) gb[k] = S[2 * SL.nc + spec_line(SL, line0 + 2 * m + 1) + kx];
    }
    if (use_bits) {
        constexpr int W = (N + 31) / 32;
        uint32_t* sb = reinterpret_cast<uint32_t*>(buf + LINES * LD);
        int* sp = reinterpret_cast<int*>(sb + LINES * W);
        int* sl = sp + LINES * W;
        tb.store(sb, sp, sl);
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const int idx = threadIdx.x + k * THREADS;
            const int line = idx / N;
            code[k] = (line < nl) ? tile_local<N>(sb, sp, sl, line, idx % N) : -2;
        }
    }
    double acc[kXinvParts] = {0, 0, 0};
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
        const int line = idx / N;
        const cx<R> z = (EXACT || line < LINES) ? buf[line * LD + idx % N] : mkc<R>((R)0, (R)0);
        if (line < nl) {
            const long vox = (line0 + line) * N + idx % N;
            xinv_emit(sink, N3, 0, vox, code[k], (double)z.x * inv_n, acc);
            xinv_emit(sink, N3, 1, vox, code[k], (double)z.y * inv_n, acc);
        }
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
        const double val =
            (EXACT || line < LINES) ? (double)bufr[2 * ((line >> 1) * LD + idx % N) + (line & 1)] * inv_n : 0.0;
        if (line < nl) xinv_emit(sink, N3, 2, (line0 + line) * N + idx % N, code[k], val, acc);
    }
    if (sink.partials != nullptr) block_sum<kXinvParts>(acc, sink.partials + (long)kXinvParts * blockIdx.x);
}

// Merged-round x forward: the tile's LINES real lines of the three components go through one
// transform of LT = 3 LINES / 2 complex lines (line l < LINES: x + i y of tile line l; line
// LINES + m: z of tile lines 2m (real part) and 2m + 1 (imaginary part)), then every output is
// split and stored in one pass.
template <typename R, int N, int LINES, int OS>
__global__ __launch_bounds__((MergedShape<R, N, LINES, OS>::THREADS)) void k_xfwd_m(XSource src,
                                                                                 cx<R>* __restrict__ S, long nlines,
                                                                                 long N3, SpecLayout SL,
                                                                                 const cx<R>* __restrict__ twg,
                                                                                 const int* skip) {
    if (skip != nullptr && *skip) return;
    using SH = MergedShape<R, N, LINES, OS>;
    constexpr int LD = SH::LD, THREADS = SH::THREADS, NXH = N / 2 + 1, LT = SH::LT;
    constexpr int KE = (LINES * N + THREADS - 1) / THREADS;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long line0 = xcd_tile(blockIdx.x, gridDim.x) * LINES;
    const int nl = (int)min((long)LINES, nlines - line0);
    double v[3][KE];
    constexpr long CAP = (long)(LT * LD * sizeof(cx<R>) / sizeof(double));
    xfwd_load<N, LINES, THREADS, KE>(src, line0, nl, nlines, N3, reinterpret_cast<double*>(buf), CAP,
                                     reinterpret_cast<unsigned char*>(buf + LT * LD), v);
    R* bufr = reinterpret_cast<R*>(buf);
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        if (idx < LINES * N) {
            const int line = idx / N, i = idx % N;
            buf[line * LD + i] = mkc<R>((R)v[0][k], (R)v[1][k]);
            bufr[2 * ((LINES + (line >> 1)) * LD + i) + (line & 1)] = (R)v[2][k];
        }
    }
    __syncthreads();
    sfft_lines<R, -1, N, LT, SH::T, LD>(buf, tw, LT);
    cx<R>* Sz = S + 2 * SL.nc;
    for (int idx = threadIdx.x; idx < LT * NXH; idx += THREADS) {
        const int line = idx / NXH;
        const int kx = idx % NXH;
        const cx<R> zk = buf[line * LD + kx];
        const cx<R> zm = buf[line * LD + (N - kx) % N];
        const cx<R> a = mkc<R>((R)0.5 * (zk.x + zm.x), (R)0.5 * (zk.y - zm.y));
        const cx<R> b = mkc<R>((R)0.5 * (zk.y + zm.y), (R)0.5 * (zm.x - zk.x));
        if (line < LINES) {
            if (line < nl) {
                const long o = spec_line(SL, line0 + line) + kx;
                S[o] = a;
                S[SL.nc + o] = b;
            }
        } else {
            const int la = 2 * (line - LINES), lb = la + 1;
            if (la < nl) Sz[spec_line(SL, line0 + la) + kx] = a;
            if (lb < nl) Sz[spec_line(SL, line0 + lb) + kx] = b;
        }
    }
}

// Writes the three output values of voxel vox / support index cd and accumulates the support
// sums of x, y, z.
__device__ __forceinline__ void xinv_emit3(const XSink& sink, long N3, long vox, int cd, double a, double b, double c,
                                           double (&acc)[kXinvParts]) {
    if (sink.field != nullptr) {
        sink.field[vox] = a;
        sink.field[N3 + vox] = b;
        sink.field[2 * N3 + vox] = c;
    }
    if (cd >= 0) {
        if (sink.compact != nullptr) {
            sink.compact[3L * cd] = a;
            sink.compact[3L * cd + 1] = b;
            sink.compact[3L * cd + 2] = c;
        }
        if (sink.compact32 != nullptr) {
            sink.compact32[3L * cd] = (float)a;
            sink.compact32[3L * cd + 1] = (float)b;
            sink.compact32[3L * cd + 2] = (float)c;
        }
        acc[0] += a;
        acc[1] += b;
        acc[2] += c;
    }
}

// Merged-round x inverse: the x and y half spectra of the tile's lines and the z half spectra of
// its line pairs fill LT = 3 LINES / 2 transform lines, one inverse transform, and every voxel's
// three components are written together.
template <typename R, int N, int LINES, int OS>
__global__ __launch_bounds__((MergedShape<R, N, LINES, OS>::THREADS)) void k_xinv_m(const cx<R>* __restrict__ S,
                                                                                 XSink sink, long nlines, long N3,
                                                                                 SpecLayout SL,
                                                                                 const cx<R>* __restrict__ twg,
                                                                                 double inv_n, const int* skip) {
    if (skip != nullptr && *skip) return;
    using SH = MergedShape<R, N, LINES, OS>;
    constexpr int LD = SH::LD, THREADS = SH::THREADS, NXH = N / 2 + 1, LT = SH::LT;
    constexpr int KE = (LINES * N + THREADS - 1) / THREADS;
    constexpr int KA = (LINES * NXH + THREADS - 1) / THREADS;        // x / y spectrum loads per thread
    constexpr int KB = ((LINES / 2) * NXH + THREADS - 1) / THREADS;  // z spectrum pair loads per thread
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + N;
    for (int i = threadIdx.x; i < N; i += THREADS) tw[i] = twg[i];
    const long line0 = xcd_tile(blockIdx.x, gridDim.x) * LINES;
    const int nl = (int)min((long)LINES, nlines - line0);
    const bool use_bits = sink.lines.bits != nullptr;
    TileBitsRegs<N, LINES, THREADS> tb;
    if (use_bits) tb.fetch(sink.lines, line0, nl);
    int code[KE];
    if (!use_bits) {
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const int idx = threadIdx.x + k * THREADS;
            const int line = idx / N;
            code[k] = (idx < LINES * N && line < nl) ? sink.code[(line0 + line) * N + idx % N] : -2;
        }
    }
    cx<R> ha[KA], hb[KA];
#pragma unroll
    for (int k = 0; k < KA; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int line = idx / NXH;
        const bool ok = idx < LINES * NXH && line < nl;
        const long o = spec_line(SL, line0 + line) + idx % NXH;
        ha[k] = mkc<R>((R)0, (R)0);
        hb[k] = mkc<R>((R)0, (R)0);
        if (ok) {
            ha[k] = S[o];
            hb[k] = S[SL.nc + o];
        }
    }
    cx<R> ga[KB], gb[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int m = idx / NXH;
        const int kx = idx % NXH;
        const bool in = idx < (LINES / 2) * NXH;
        ga[k] = mkc<R>((R)0, (R)0);
        gb[k] = mkc<R>((R)0, (R)0);
        if (in && 2 * m < nl) ga[k] = S[2 * SL.nc + spec_line(SL, line0 + 2 * m) + kx];
        if (in && 2 * m + 1 < nl) gb[k] = S[2 * SL.nc + spec_line(SL, line0 + 2 * m + 1) + kx];
    }
    if (use_bits) {
        constexpr int W = (N + 31) / 32;
        uint32_t* sb = reinterpret_cast<uint32_t*>(buf + LT * LD);
        int* sp = reinterpret_cast<int*>(sb + LINES * W);
        int* sl = sp + LINES * W;
        tb.store(sb, sp, sl);
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const int idx = threadIdx.x + k * THREADS;
            const int line = idx / N;
            code[k] = (idx < LINES * N && line < nl) ? tile_local<N>(sb, sp, sl, line, idx % N) : -2;
        }
    }
#pragma unroll
    for (int k = 0; k < KA; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        if (idx < LINES * NXH) xinv_put<R, N, LD>(buf, idx / NXH, idx % NXH, ha[k], hb[k]);
    }
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        if (idx < (LINES / 2) * NXH) xinv_put<R, N, LD>(buf, LINES + idx / NXH, idx % NXH, ga[k], gb[k]);
    }
    __syncthreads();
    sfft_lines<R, +1, N, LT, SH::T, LD>(buf, tw, LT);
    const R* bufr = reinterpret_cast<const R*>(buf);
    double acc[kXinvParts] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int idx = threadIdx.x + k * THREADS;
        const int line = idx / N;
        if (idx < LINES * N && line < nl) {
            const int i = idx % N;
            const cx<R> z = buf[line * LD + i];
            const double zz = (double)bufr[2 * ((LINES + (line >> 1)) * LD + i) + (line & 1)];
            xinv_emit3(sink, N3, (line0 + line) * N + i, code[k], (double)z.x * inv_n, (double)z.y * inv_n,
                       zz * inv_n, acc);
        }
    }
    if (sink.partials != nullptr) block_sum<kXinvParts>(acc, sink.partials + (long)kXinvParts * blockIdx.x);
}

// ------------------------------------------------------- wave-per-line-pair x passes
// One wave transforms the three complex lines of an x-line pair (x + i y of each line, then z of
// the pair) one after the other.  Lane j holds elements j + 64 r (r < RAD = N / 64) in registers
// and the radix-RAD Stockham stages (N = RAD^S: 128 = 2^7, 256 = 4^4, 512 = 8^3) exchange through
// a per-wave LDS line, so after the twiddle table is loaded the waves never wait for each other.
// The support of a line is read as uniform words (one x line per wave at a time) and the compact
// vector is gathered straight into registers.
constexpr int kWaveLines = 8;  // x lines per workgroup: 4 waves x 2

template <int N>
constexpr bool wave_pass_ok() {
    return N == 128 || N == 256 || N == 512;
}

// LDS index of element e of a wave's line or of the twiddle table: one pad slot after every 16
// elements keeps the Stockham writes (element strides 4 and 16 across lanes at N = 256), the
// consecutive reads and the twiddle reads of stride r < 4 free of bank conflicts.
__device__ __forceinline__ int lpad(int e) { return e + (e >> 4); }
template <int N>
constexpr int lpad_size() {
    return N + N / 16;
}

// Orders the wave's LDS writes before its later reads (a wave executes its LDS operations in
// order; the fences keep the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename R, int SIGN, int N, int NS, bool KEEP>
__device__ __forceinline__ void wave_fft_stage(cx<R> (&v)[N / 64], cx<R>* L, const cx<R>* tw, int lane) {
    constexpr int RAD = N / 64;
    if constexpr (NS < N) {
        const int k = lane % NS;
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 0; r < RAD; ++r) v[r] = L[lpad(lane + r * 64)];
            constexpr int TS = N / (NS * RAD);
#pragma unroll
            for (int r = 1; r < RAD; ++r) {
                cx<R> w = tw[lpad(r * k * TS)];
                if (SIGN > 0) w.y = -w.y;
                v[r] = cmul<R>(v[r], w);
            }
            wave_lds_sync();
        }
        dft_small<R, SIGN, RAD>(v);
        // the last stage (NS = N / RAD = 64) writes element lane + 64 r from v[r]: with KEEP the
        // spectrum stays in the registers, already in that order
        if constexpr (KEEP && NS * RAD == N) return;
        const int base = (lane / NS) * NS * RAD + k;
#pragma unroll
        for (int r = 0; r < RAD; ++r) L[lpad(base + r * NS)] = v[r];
        wave_lds_sync();
        wave_fft_stage<R, SIGN, N, NS * RAD, KEEP>(v, L, tw, lane);
    }
}

// Transform of one line given as v[r] = x[lane + 64 r]; the spectrum is left in L in natural order
// (element k at L[lpad(k)]; the twiddle table padded the same way), or with KEEP in the registers
// (v[r] = X[lane + 64 r]; L then holds nothing of it).
template <typename R, int SIGN, int N, bool KEEP = false>
__device__ __forceinline__ void wave_fft(cx<R> (&v)[N / 64], cx<R>* L, const cx<R>* tw, int lane) {
    wave_fft_stage<R, SIGN, N, 1, KEEP>(v, L, tw, lane);
}

// Support words of x line `line` (uniform across the wave) and their in-line prefix counts;
// loc(r) = support index of voxel lane + 64 r, or -1.  Voxel lane + 64 r is bit `lane` of the 64-bit
// word pair (2 r, 2 r + 1), so with the lane's own masks (LaneBits64, computed once per kernel) the
// index is two masked popcounts of uniform words: no per-lane choice between the two words.
struct LaneBits64 {
    uint32_t bit_lo, bit_hi, below_lo, below_hi;
};
__device__ __forceinline__ LaneBits64 lane_bits64(int lane) {
    const uint64_t bit = 1ull << lane, below = bit - 1ull;
    return LaneBits64{(uint32_t)bit, (uint32_t)(bit >> 32), (uint32_t)below, (uint32_t)(below >> 32)};
}

template <int N>
struct WaveLineBits {
    static constexpr int W = N / 32;
    uint32_t word[W];
    int pre[W];
    int start;
    __device__ __forceinline__ void load(const LineBits& lb, long line) {
        const uint32_t* wb = lb.bits + line * W;
#pragma unroll
        for (int u = 0; u < W; ++u) word[u] = wb[u];
        start = lb.lstart[line];
        int acc = 0;
#pragma unroll
        for (int u = 0; u < W; ++u) {
            pre[u] = acc;
            acc += __popc(word[u]);
        }
    }
    __device__ __forceinline__ int loc(const LaneBits64& m, int r) const {
        const uint32_t wl = word[2 * r], wh = word[2 * r + 1];
        const bool on = ((wl & m.bit_lo) | (wh & m.bit_hi)) != 0u;
        const int cnt = start + pre[2 * r] + __popc(wl & m.below_lo) + __popc(wh & m.below_hi);
        return on ? cnt : -1;
    }
};

// Two-for-one split of the wave's spectrum Z (natural order in L, and v[r] = Z[lane + 64 r] as the
// last stage left it) into the half spectra of its real and imaginary parts, stored at rows oa
// (real part) and ob (imaginary part; ob < 0: none).
template <typename R, int N>
__device__ __forceinline__ void wave_store_halves(const cx<R>* L, const cx<R> (&v)[N / 64], cx<R>* Sa, cx<R>* Sb,
                                                  bool has_b, int lane) {
    constexpr int NXH = N / 2 + 1;
    static_assert((NXH + 63) / 64 <= N / 64, "the kept half lies in the registers");
#pragma unroll
    for (int q = 0; q < (NXH + 63) / 64; ++q) {
        const int kx = lane + 64 * q;
        if (kx < NXH) {
            const cx<R> zk = v[q];
            const cx<R> zm = L[lpad((N - kx) % N)];
            Sa[kx] = mkc<R>((R)0.5 * (zk.x + zm.x), (R)0.5 * (zk.y - zm.y));
            if (has_b) Sb[kx] = mkc<R>((R)0.5 * (zk.y + zm.y), (R)0.5 * (zm.x - zk.x));
        }
    }
}

// Copies the run src[0, cnt) (cnt <= CAP, uniform across the wave) into the wave's LDS line with
// every global load issued before the first LDS store waits on one.
template <typename T, int CAP>
__device__ __forceinline__ void wave_stage_run(const T* __restrict__ src, int cnt, T* dst, int lane) {
    constexpr int K = (CAP + 63) / 64;
    T t[K];
    const int it = (cnt + 63) >> 6;
#pragma unroll
    for (int j = 0; j < K; ++j)
        if (j < it) t[j] = src[min(lane + 64 * j, cnt - 1)];
#pragma unroll
    for (int j = 0; j < K; ++j)
        if (j < it && lane + 64 * j < cnt) dst[lane + 64 * j] = t[j];
}

template <typename R, int N>
__global__ __launch_bounds__(256) void k_xfwd_w(XSource src, cx<R>* __restrict__ S, long nlines, SpecLayout SL,
                                                const cx<R>* __restrict__ twg, const int* skip) {
    if (skip != nullptr && *skip) return;
    constexpr int RAD = N / 64;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    // the wave index is uniform: the line numbers, support words and run bounds below are too
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const LaneBits64 lmask = lane_bits64(lane);
    cx<R>* L = tw + lpad_size<N>() + wave * lpad_size<N>();
    for (int i = threadIdx.x; i < N; i += 256) tw[lpad(i)] = twg[i];
    __syncthreads();
    const long line0 = __builtin_amdgcn_readfirstlane((int)(xcd_tile(blockIdx.x, gridDim.x) * kWaveLines)) + 2 * wave;
    if (line0 >= nlines) return;
    const bool two = line0 + 1 < nlines;
    WaveLineBits<N> wb[2];
    wb[0].load(src.lines, line0);
    if (two) wb[1].load(src.lines, line0 + 1);
    double mc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) mc[c] = (src.mean != nullptr) ? src.mean[c] : 0.0;
    // The pair's support voxels are one contiguous run [cs, ce) of the compact vector: it is staged
    // through the wave's LDS line with coalesced loads when it fits (else gathered directly).
    const long nb = src.lines.nb;
    const long cs = wb[0].start;
    const long ce = (line0 + 2 < nlines) ? (long)src.lines.lstart[line0 + 2] : nb;
    const long cnt3 = 3 * (ce - cs);
    float* st32 = reinterpret_cast<float*>(L);
    double* st64 = reinterpret_cast<double*>(L);
    constexpr int cap32 = (int)(lpad_size<N>() * sizeof(cx<R>) / sizeof(float));
    constexpr int cap64 = (int)(lpad_size<N>() * sizeof(cx<R>) / sizeof(double));
    const bool staged32 = src.compact32 != nullptr && cnt3 <= cap32;
    const bool staged64 = src.compact32 == nullptr && src.compact != nullptr && cnt3 <= cap64;
    if (staged32)
        wave_stage_run<float, cap32>(src.compact32 + 3 * cs, (int)cnt3, st32, lane);
    else if (staged64)
        wave_stage_run<double, cap64>(src.compact + 3 * cs, (int)cnt3, st64, lane);
    wave_lds_sync();
    // every source value of the pair first: v[a][c][r] of voxel lane + 64 r of line line0 + a
    R val[2][3][RAD];
    // from the stage: every lane reads an in-range slot and off-support voxels select zero
    auto gather_staged = [&](auto&& read) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
#pragma unroll
            for (int r = 0; r < RAD; ++r) {
                const int l = (a == 0 || two) ? wb[a].loc(lmask, r) : -1;
                const int e = l >= 0 ? 3 * (l - (int)cs) : 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const R x = read(e + c, c);
                    val[a][c][r] = l >= 0 ? x : (R)0;
                }
            }
        }
    };
    if (staged32) {
        // an FP32 source has its mean subtracted in FP32 (fewer conversions and registers)
        const R mr[3] = {(R)mc[0], (R)mc[1], (R)mc[2]};
        gather_staged([&](int e, int c) { return (R)st32[e] - mr[c]; });
    } else if (staged64) {
        gather_staged([&](int e, int c) { return (R)(st64[e] - mc[c]); });
    } else {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
#pragma unroll
            for (int r = 0; r < RAD; ++r) {
                const int l = (a == 0 || two) ? wb[a].loc(lmask, r) : -1;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    if (src.compact32 != nullptr) {
                        const float raw = (l >= 0) ? src.compact32[3L * l + c] : 0.0f;
                        val[a][c][r] = (l < 0) ? (R)0 : (R)raw - (R)mc[c];
                    } else {
                        double raw = 0.0;
                        if (l >= 0 && src.compact != nullptr) raw = src.compact[3L * l + c];
                        val[a][c][r] = (l < 0) ? (R)0 : (R)(raw - mc[c]);
                    }
                }
            }
        }
    }
    wave_lds_sync();  // the stage is the transform line
    const long oa = spec_line(SL, line0), ob = spec_line(SL, line0 + (two ? 1 : 0));
    cx<R> v[RAD];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        if (a == 1 && !two) break;
#pragma unroll
        for (int r = 0; r < RAD; ++r) v[r] = mkc<R>(val[a][0][r], val[a][1][r]);
        wave_fft<R, -1, N>(v, L, tw, lane);
        const long o = a == 0 ? oa : ob;
        wave_store_halves<R, N>(L, v, S + o, S + SL.nc + o, true, lane);
        wave_lds_sync();
    }
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[r] = mkc<R>(val[0][2][r], val[1][2][r]);
    wave_fft<R, -1, N>(v, L, tw, lane);
    wave_store_halves<R, N>(L, v, S + 2 * SL.nc + oa, S + 2 * SL.nc + ob, two, lane);
}

// Z = A + i B over the Hermitian completions of the half-spectrum rows A and B at element
// lane + 64 r (self-conjugate bins keep their real parts only).
template <typename R, int N>
__device__ __forceinline__ void wave_load_pair(const cx<R>* __restrict__ Ar, const cx<R>* __restrict__ Br, bool has_b,
                                               cx<R> (&v)[N / 64], int lane) {
#pragma unroll
    for (int r = 0; r < N / 64; ++r) {
        const int k = lane + 64 * r;
        const bool upper = 2 * k > N;
        const int kk = upper ? N - k : k;
        cx<R> a = Ar[kk];
        cx<R> b = mkc<R>((R)0, (R)0);
        if (has_b) b = Br[kk];
        if (kk == 0 || 2 * kk == N) {
            a.y = (R)0;
            b.y = (R)0;
        }
        v[r] = upper ? mkc<R>(a.x + b.y, b.x - a.y) : mkc<R>(a.x - b.y, a.y + b.x);
    }
}

// Bytes of a wave's output stage in the x-inverse pass: one transform line.
template <typename R, int N>
constexpr int wave_ostage_bytes() {
    return (int)(lpad_size<N>() * sizeof(cx<R>));
}

// Waves per SIMD the compiler must allow: the FP32 256 x-inverse fits 80 VGPRs without spilling
// (6 waves instead of 5); the others keep the compiler's choice.
template <typename R, int N>
constexpr int xinv_w_min_waves() {
    return sizeof(R) == 4 && N == 256 ? 6 : 1;
}

template <typename R, int N>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(xinv_w_min_waves<R, N>(), 8))) void k_xinv_w(const cx<R>* __restrict__ S, XSink sink, long nlines, SpecLayout SL,
                                                const cx<R>* __restrict__ twg, double inv_n, const int* skip) {
    if (skip != nullptr && *skip) return;
    constexpr int RAD = N / 64;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    // the wave index is uniform: the line numbers, support words and run bounds below are too
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const LaneBits64 lmask = lane_bits64(lane);
    cx<R>* L = tw + lpad_size<N>() + wave * lpad_size<N>();
    R* ost = reinterpret_cast<R*>(reinterpret_cast<unsigned char*>(tw + 5 * lpad_size<N>()) +
                                  wave * wave_ostage_bytes<R, N>());
    for (int i = threadIdx.x; i < N; i += 256) tw[lpad(i)] = twg[i];
    __syncthreads();
    const long line0 = __builtin_amdgcn_readfirstlane((int)(xcd_tile(blockIdx.x, gridDim.x) * kWaveLines)) + 2 * wave;
    const bool one = line0 < nlines;
    const bool two = line0 + 1 < nlines;
    double acc[kXinvParts] = {0, 0, 0};
    if (one) {
        const long oa = spec_line(SL, line0), ob = spec_line(SL, line0 + (two ? 1 : 0));
        // every spectrum value of the pair first
        cx<R> z[3][RAD];
        wave_load_pair<R, N>(S + oa, S + SL.nc + oa, true, z[0], lane);
        if (two) wave_load_pair<R, N>(S + ob, S + SL.nc + ob, true, z[1], lane);
        wave_load_pair<R, N>(S + 2 * SL.nc + oa, S + 2 * SL.nc + ob, two, z[2], lane);
        int loc[2][RAD];
        int cs = 0;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            WaveLineBits<N> wb;
            const bool ok = a == 0 || two;
            if (ok) wb.load(sink.lines, line0 + a);
            if (a == 0) cs = wb.start;
#pragma unroll
            for (int r = 0; r < RAD; ++r) loc[a][r] = ok ? wb.loc(lmask, r) : -1;
        }
        // The pair's support voxels are one contiguous run [cs, ce) of the compact vector.  A sink of
        // the transform's precision whose run fits the wave's output stage takes the values there
        // (off-support voxels write a dummy slot) and then gets one coalesced copy of the run; the
        // support sums are taken during that copy.  Other sinks are written voxel by voxel.
        R* sdst = sizeof(R) == 4 ? reinterpret_cast<R*>(sink.compact32) : reinterpret_cast<R*>(sink.compact);
        if (sizeof(R) == 8 && sink.compact32 != nullptr) sdst = nullptr;
        const int ce = (line0 + 2 < nlines) ? sink.lines.lstart[line0 + 2] : (int)sink.lines.nb;
        const int cnt3 = 3 * (ce - cs);
        constexpr int cap = wave_ostage_bytes<R, N>() / (int)sizeof(R) - 3;  // slots cap..cap+2: dummy
        const R inv_r = (R)inv_n;
        if (sdst != nullptr && cnt3 <= cap) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int r = 0; r < RAD; ++r) loc[a][r] = loc[a][r] >= 0 ? 3 * (loc[a][r] - cs) : cap;
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                if (f == 1 && !two) continue;
                wave_fft<R, +1, N, true>(z[f], L, tw, lane);
#pragma unroll
                for (int r = 0; r < RAD; ++r) {
                    const cx<R> t = z[f][r];
                    if (f < 2) {
                        ost[loc[f][r]] = t.x * inv_r;
                        ost[loc[f][r] + 1] = t.y * inv_r;
                    } else {
                        ost[loc[0][r] + 2] = t.x * inv_r;
                        if (two) ost[loc[1][r] + 2] = t.y * inv_r;
                    }
                }
                wave_lds_sync();
            }
            // element e of the run is component e % 3; lane + 64 j has component (lane + j) % 3
            constexpr int K = (cap + 63) / 64;
            const int it = (cnt3 + 63) >> 6;
            R* dst = sdst + 3L * cs;
            double a3[3] = {0, 0, 0};
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const int e = lane + 64 * j;
                if (j < it && e < cnt3) {
                    const R v = ost[e];
                    dst[e] = v;
                    a3[j % 3] += (double)v;
                }
            }
            const int c0 = lane % 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int k = (c + 3 - c0) % 3;
                acc[c] += k == 0 ? a3[0] : (k == 1 ? a3[1] : a3[2]);
            }
        } else {
            // an FP32 sink takes the value scaled in the transform's precision (one multiply, no
            // round trip through FP64); the support sums stay FP64
            auto emit = [&](int c, int l, R t) {
                if (l >= 0) {
                    if (sink.compact32 != nullptr) {
                        const float v = (float)(t * inv_r);
                        sink.compact32[3L * l + c] = v;
                        acc[c] += (double)v;
                    } else {
                        const double v = (double)t * inv_n;
                        if (sink.compact != nullptr) sink.compact[3L * l + c] = v;
                        acc[c] += v;
                    }
                }
            };
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                if (f == 1 && !two) continue;
                wave_fft<R, +1, N, true>(z[f], L, tw, lane);
#pragma unroll
                for (int r = 0; r < RAD; ++r) {
                    const cx<R> t = z[f][r];
                    if (f < 2) {
                        emit(0, loc[f][r], t.x);
                        emit(1, loc[f][r], t.y);
                    } else {
                        emit(2, loc[0][r], t.x);
                        if (two) emit(2, loc[1][r], t.y);
                    }
                }
                wave_lds_sync();
            }
        }
    }
    if (sink.partials != nullptr) block_sum<kXinvParts>(acc, sink.partials + (long)kXinvParts * blockIdx.x);
}

template <typename R, int N>
constexpr size_t wave_pass_lds() {
    return (size_t)(5 * lpad_size<N>()) * sizeof(cx<R>);
}
// the x-inverse adds an output stage per wave
template <typename R, int N>
constexpr size_t wave_xinv_lds() {
    return wave_pass_lds<R, N>() + 4 * (size_t)wave_ostage_bytes<R, N>();
}

// The wave passes serve the solver's applies: a compact (FP64 or FP32) source without background
// force and a compact sink, both over the line bitmask; other calls take the 8 x 2 tile.
inline bool wave_src_ok(const XSource& src) {
    const bool no_f0 = src.a_fluid[0] == 0.0 && src.a_fluid[1] == 0.0 && src.a_fluid[2] == 0.0 &&
                       src.a_support[0] == 0.0 && src.a_support[1] == 0.0 && src.a_support[2] == 0.0;
    return src.field == nullptr && src.lines.bits != nullptr && no_f0;
}
inline bool wave_sink_ok(const XSink& sink) { return sink.field == nullptr && sink.lines.bits != nullptr; }

template <typename R, int N, int V>
long xfwd_variant(const XSource& src, cx<R>* S, const Geometry3& g, const cx<R>* tw, const int* skip,
                  hipStream_t st) {
    constexpr int LINES = xpass_lines<N>(V), OS = xpass_os(V);
    const long nlines = (long)g.ny * g.nz;
    const long blocks = (nlines + LINES - 1) / LINES;
    if constexpr (xpass_wave(V)) {
        if constexpr (!wave_pass_ok<N>()) {
            return -1;
        } else {
            if (!wave_src_ok(src)) return xfwd_variant<R, N, 2>(src, S, g, tw, skip, st);
            constexpr size_t lds = wave_pass_lds<R, N>();
            if (lds > lds_limit()) return -1;
            hipLaunchKernelGGL((k_xfwd_w<R, N>), dim3((unsigned)blocks), dim3(256), lds, st, src, S, nlines,
                               spec_layout(g), tw, skip);
            return blocks;
        }
    } else if constexpr (xpass_merged(V)) {
        using SH = MergedShape<R, N, LINES, OS>;
        if constexpr (!SH::valid) {
            return -1;
        } else {
            constexpr size_t lds = SH::LDS + bits_lds<N, LINES>();
            if (lds > lds_limit()) return -1;
            hipLaunchKernelGGL((k_xfwd_m<R, N, LINES, OS>), dim3((unsigned)blocks), dim3(SH::THREADS), lds, st, src,
                               S, nlines, g.N, spec_layout(g), tw, skip);
            return blocks;
        }
    } else {
        using SH = PassShape<R, N, LINES, OS>;
        if constexpr (!SH::valid) {
            return -1;
        } else {
            constexpr size_t lds = SH::LDS + bits_lds<N, LINES>();
            if (lds > lds_limit()) return -1;
            hipLaunchKernelGGL((k_xfwd_s<R, N, LINES, OS>), dim3((unsigned)blocks), dim3(SH::THREADS), lds, st, src,
                               S, nlines, g.N, spec_layout(g), tw, skip);
            return blocks;
        }
    }
}

template <typename R, int N, int V>
long xinv_variant(const cx<R>* S, const XSink& sink, const Geometry3& g, const cx<R>* tw, const int* skip,
                  hipStream_t st) {
    constexpr int LINES = xpass_lines<N>(V), OS = xpass_os(V);
    const long nlines = (long)g.ny * g.nz;
    const long blocks = (nlines + LINES - 1) / LINES;
    const double inv_n = 1.0 / (double)g.N;
    if constexpr (xpass_wave(V)) {
        if constexpr (!wave_pass_ok<N>()) {
            return -1;
        } else {
            if (!wave_sink_ok(sink)) return xinv_variant<R, N, 2>(S, sink, g, tw, skip, st);
            constexpr size_t lds = wave_xinv_lds<R, N>();
            if (lds > lds_limit()) return -1;
            hipLaunchKernelGGL((k_xinv_w<R, N>), dim3((unsigned)blocks), dim3(256), lds, st, S, sink, nlines,
                               spec_layout(g), tw, inv_n, skip);
            return blocks;
        }
    } else if constexpr (xpass_merged(V)) {
        using SH = MergedShape<R, N, LINES, OS>;
        if constexpr (!SH::valid) {
            return -1;
        } else {
            constexpr size_t lds = SH::LDS + bits_lds<N, LINES>();
            if (lds > lds_limit()) return -1;
            hipLaunchKernelGGL((k_xinv_m<R, N, LINES, OS>), dim3((unsigned)blocks), dim3(SH::THREADS), lds, st, S,
                               sink, nlines, g.N, spec_layout(g), tw, inv_n, skip);
            return blocks;
        }
    } else {
        using SH = PassShape<R, N, LINES, OS>;
        if constexpr (!SH::valid) {
            return -1;
        } else {
            constexpr size_t lds = SH::LDS + bits_lds<N, LINES>();
            if (lds > lds_limit()) return -1;
            hipLaunchKernelGGL((k_xinv_s<R, N, LINES, OS>), dim3((unsigned)blocks), dim3(SH::THREADS), lds, st, S,
                               sink, nlines, g.N, spec_layout(g), tw, inv_n, skip);
            return blocks;
        }
    }
}

template <typename R, int N>
long xfwd_n(int v, const XSource& src, cx<R>* S, const Geometry3& g, const cx<R>* tw, const int* skip,
            hipStream_t st) {
    switch (v) {
        case 0: return xfwd_variant<R, N, 0>(src, S, g, tw, skip, st);
        case 1: return xfwd_variant<R, N, 1>(src, S, g, tw, skip, st);
        case 2: return xfwd_variant<R, N, 2>(src, S, g, tw, skip, st);
        case 3: return xfwd_variant<R, N, 3>(src, S, g, tw, skip, st);
        case 4: return xfwd_variant<R, N, 4>(src, S, g, tw, skip, st);
        case 5: return xfwd_variant<R, N, 5>(src, S, g, tw, skip, st);
        case 6: return xfwd_variant<R, N, 6>(src, S, g, tw, skip, st);
        case 7: return xfwd_variant<R, N, 7>(src, S, g, tw, skip, st);
        case 8: return xfwd_variant<R, N, 8>(src, S, g, tw, skip, st);
        default: return -1;
    }
}

template <typename R, int N>
long xinv_n(int v, const cx<R>* S, const XSink& sink, const Geometry3& g, const cx<R>* tw, const int* skip,
            hipStream_t st) {
    switch (v) {
        case 0: return xinv_variant<R, N, 0>(S, sink, g, tw, skip, st);
        case 1: return xinv_variant<R, N, 1>(S, sink, g, tw, skip, st);
        case 2: return xinv_variant<R, N, 2>(S, sink, g, tw, skip, st);
        case 3: return xinv_variant<R, N, 3>(S, sink, g, tw, skip, st);
        case 4: return xinv_variant<R, N, 4>(S, sink, g, tw, skip, st);
        case 5: return xinv_variant<R, N, 5>(S, sink, g, tw, skip, st);
        case 6: return xinv_variant<R, N, 6>(S, sink, g, tw, skip, st);
        case 7: return xinv_variant<R, N, 7>(S, sink, g, tw, skip, st);
        case 8: return xinv_variant<R, N, 8>(S, sink, g, tw, skip, st);
        default: return -1;
    }
}

// 10 * lines + threads per butterfly thread; merged-round x variants add 5 (last digit 6, 7 or 9);
// the wave-per-line-pair variant is 88, for the lengths it serves
template <int N>
int code_of(int v) {
    if (xpass_wave(v)) return wave_pass_ok<N>() ? 88 : 0;
    return 10 * xpass_lines<N>(v) + xpass_os(v) + (xpass_merged(v) ? 5 : 0);
}

}  // namespace

}  // namespace fftperm

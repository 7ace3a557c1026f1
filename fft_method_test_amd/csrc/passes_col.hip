// passes_col.hip -- compile-time specialised column passes (the y transform of all three
// components; the z transform of the standalone FFT hooks) for the column lengths 32 ... 512 and
// 80, 96, 112, 144, 160, in kColVariants tile shapes each.
//
// Column col has base (col / inner) * outer_stride + col % inner and element e at
// base + e * estride; a workgroup takes LINES consecutive columns, so each row of its tile is
// LINES contiguous complex values in HBM.  Thread t always serves column t % LINES in the load
// and store loops.  Grid row y transforms component y (at y * comp_stride); a launch over several
// right-hand sides has 3 components per side and skips side b when skip[b * skip_stride] is set.
#include "fft_static.hpp"
#include "kernels.hpp"

namespace fftperm {

namespace {

template <typename R, int N, int LINES, int OS, int SIGN>
__global__ __launch_bounds__((PassShape<R, N, LINES, OS>::THREADS)) void k_col_s(cx<R>* __restrict__ S, long ncols,
                                                                               long inner, long outer_stride,
                                                                               long estride, long comp_stride,
                                                                               const cx<R>* __restrict__ twg,
                                                                               const int* skip, int skip_stride) {
    // grid row blockIdx.y = component; components 3b .. 3b + 2 belong to right-hand side b
    if (skip != nullptr && skip[(blockIdx.y / 3) * skip_stride]) return;
    using SH = PassShape<R, N, LINES, OS>;
    constexpr int LD = SH::LD, THREADS = SH::THREADS;
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
    // all loads of the thread in flight at once, then one LDS fill; a thread's last element slot
    // idles when EST does not divide the column length (mixed-radix lengths)
    constexpr int KL = (N + EST - 1) / EST;
    constexpr bool EXACT = N % EST == 0;
    cx<R> ld[KL];
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int e = threadIdx.x / LINES + k * EST;
        ld[k] = mkc<R>((R)0, (R)0);
        if (own && (EXACT || e < N)) ld[k] = Sc[base + (long)e * estride];
    }
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int e = threadIdx.x / LINES + k * EST;
        if (own && (EXACT || e < N)) buf[ci * LD + e] = ld[k];
    }
    __syncthreads();
    sfft_lines<R, SIGN, N, LINES, SH::T, LD>(buf, tw, ncl);
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int e = threadIdx.x / LINES + k * EST;
        if (own && (EXACT || e < N)) Sc[base + (long)e * estride] = buf[ci * LD + e];
    }
}

template <typename R, int N, int V>
long col_variant(int sign, cx<R>* S, long ncols, long inner, long outer, long estride, long comp_stride,
                 const cx<R>* tw, const int* skip, int ncomp, int skip_stride, hipStream_t st) {
    constexpr int LINES = col_lines<N>(V), OS = col_os(V);
    using SH = PassShape<R, N, LINES, OS>;
    if constexpr (!SH::valid) {
        return -1;
    } else {
        if (SH::LDS > lds_limit()) return -1;
        const long blocks = (ncols + LINES - 1) / LINES;
        const dim3 grid((unsigned)blocks, (unsigned)ncomp);
        if (sign < 0)
            hipLaunchKernelGGL((k_col_s<R, N, LINES, OS, -1>), grid, dim3(SH::THREADS), SH::LDS, st, S, ncols, inner,
                               outer, estride, comp_stride, tw, skip, skip_stride);
        else
            hipLaunchKernelGGL((k_col_s<R, N, LINES, OS, +1>), grid, dim3(SH::THREADS), SH::LDS, st, S, ncols, inner,
                               outer, estride, comp_stride, tw, skip, skip_stride);
        return blocks;
    }
}

template <typename R, int N>
long col_n(int v, int sign, cx<R>* S, long ncols, long inner, long outer, long estride, long comp_stride,
           const cx<R>* tw, const int* skip, int ncomp, int skip_stride, hipStream_t st) {
    switch (v) {
        case 0:
            return col_variant<R, N, 0>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp,
                                             skip_stride, st);
        case 1:
            return col_variant<R, N, 1>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp,
                                             skip_stride, st);
        case 2:
            return col_variant<R, N, 2>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp,
                                             skip_stride, st);
        case 3:
            return col_variant<R, N, 3>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp,
                                             skip_stride, st);
        case 4:
            return col_variant<R, N, 4>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp,
                                             skip_stride, st);
        case 5:
            return col_variant<R, N, 5>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp,
                                             skip_stride, st);
        case 6:
            return col_variant<R, N, 6>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp,
                                             skip_stride, st);
        default: return -1;
    }
}

}  // namespace

template <typename R>
long static_col(int v, int n, int sign, cx<R>* S, long ncols, long inner, long outer, long estride, long comp_stride,
                const cx<R>* tw, const int* skip, hipStream_t st, int ncomp, int skip_stride) {
    switch (n) {
        case 32:
            return col_n<R, 32>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp, skip_stride,
                                    st);
        case 64:
            return col_n<R, 64>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp, skip_stride,
                                    st);
        case 128:
            return col_n<R, 128>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp, skip_stride,
                                    st);
        case 256:
            return col_n<R, 256>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp, skip_stride,
                                    st);
        case 512:
            return col_n<R, 512>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp, skip_stride,
                                    st);
        case 80:
            return col_n<R, 80>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp, skip_stride,
                                st);
        case 96:
            return col_n<R, 96>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp, skip_stride,
                                st);
        case 112:
            return col_n<R, 112>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp, skip_stride,
                                 st);
        case 144:
            return col_n<R, 144>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp, skip_stride,
                                 st);
        case 160:
            return col_n<R, 160>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, ncomp, skip_stride,
                                 st);
        default: return -1;
    }
}

template <int N>
int col_code_of(int v) {
    return 10 * col_lines<N>(v) + col_os(v);
}

int col_variant_code(int n, int v) {
    if (v < 0 || v >= kColVariants) return 0;
    switch (n) {
        case 32: return col_code_of<32>(v);
        case 64: return col_code_of<64>(v);
        case 128: return col_code_of<128>(v);
        case 256: return col_code_of<256>(v);
        case 512: return col_code_of<512>(v);
        case 80: return col_code_of<80>(v);
        case 96: return col_code_of<96>(v);
        case 112: return col_code_of<112>(v);
        case 144: return col_code_of<144>(v);
        case 160: return col_code_of<160>(v);
        default: return 0;
    }
}

template long static_col<double>(int, int, int, cx<double>*, long, long, long, long, long, const cx<double>*,
                                 const int*, hipStream_t, int, int);
template long static_col<float>(int, int, int, cx<float>*, long, long, long, long, long, const cx<float>*,
                                const int*, hipStream_t, int, int);

}  // namespace fftperm

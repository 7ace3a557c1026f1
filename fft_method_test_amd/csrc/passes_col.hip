// passes_col.hip -- compile-time specialised column passes (the y transform of all three
// components; the z transform of the standalone FFT hooks) for column lengths 32 ... 512, in
// kPassVariants tile shapes each.
//
// Column col has base (col / inner) * outer_stride + col % inner and element e at
// base + e * estride; a workgroup takes LINES consecutive columns, so each row of its tile is
// LINES contiguous complex values in HBM.  Thread t always serves column t % LINES in the load
// and store loops.
#include "fft_static.hpp"
#include "kernels.hpp"

namespace fftperm {

namespace {

template <typename R, int N, int LINES, int OS, int SIGN>
__global__ __launch_bounds__((PassShape<R, N, LINES, OS>::THREADS)) void k_col_s(cx<R>* __restrict__ S, long ncols,
                                                                               long inner, long outer_stride,
                                                                               long estride, long comp_stride,
                                                                               const cx<R>* __restrict__ twg,
                                                                               const int* skip) {
    if (skip != nullptr && *skip) return;
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
    // all loads of the thread in flight at once, then one LDS fill
    constexpr int KL = N / EST;
    static_assert(N % EST == 0, "column elements split evenly over the load threads");
    cx<R> ld[KL];
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        ld[k] = mkc<R>((R)0, (R)0);
        if (own) ld[k] = Sc[base + (long)(threadIdx.x / LINES + k * EST) * estride];
    }
#pragma unroll
    for (int k = 0; k < KL; ++k)
        if (own) buf[ci * LD + threadIdx.x / LINES + k * EST] = ld[k];
    __syncthreads();
    sfft_lines<R, SIGN, N, LINES, SH::T, LD>(buf, tw, ncl);
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int e = threadIdx.x / LINES + k * EST;
        if (own) Sc[base + (long)e * estride] = buf[ci * LD + e];
    }
}

template <typename R, int N, int V>
long col_variant(int sign, cx<R>* S, long ncols, long inner, long outer, long estride, long comp_stride,
                 const cx<R>* tw, const int* skip, hipStream_t st) {
    constexpr int LINES = col_lines<N>(V), OS = col_os(V);
    using SH = PassShape<R, N, LINES, OS>;
    if constexpr (!SH::valid) {
        return -1;
    } else {
        if (SH::LDS > lds_limit()) return -1;
        const long blocks = (ncols + LINES - 1) / LINES;
        const dim3 grid((unsigned)blocks, 3);
        if (sign < 0)
            hipLaunchKernelGGL((k_col_s<R, N, LINES, OS, -1>), grid, dim3(SH::THREADS), SH::LDS, st, S, ncols, inner,
                               outer, estride, comp_stride, tw, skip);
        else
            hipLaunchKernelGGL((k_col_s<R, N, LINES, OS, +1>), grid, dim3(SH::THREADS), SH::LDS, st, S, ncols, inner,
                               outer, estride, comp_stride, tw, skip);
        return blocks;
    }
}

template <typename R, int N>
long col_n(int v, int sign, cx<R>* S, long ncols, long inner, long outer, long estride, long comp_stride,
           const cx<R>* tw, const int* skip, hipStream_t st) {
    switch (v) {
        case 0: return col_variant<R, N, 0>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
        case 1: return col_variant<R, N, 1>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
        case 2: return col_variant<R, N, 2>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
        case 3: return col_variant<R, N, 3>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
        case 4: return col_variant<R, N, 4>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
        case 5: return col_variant<R, N, 5>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
        case 6: return col_variant<R, N, 6>(sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
        default: return -1;
    }
}

}  // namespace

template <typename R>
long static_col(int v, int n, int sign, cx<R>* S, long ncols, long inner, long outer, long estride, long comp_stride,
                const cx<R>* tw, const int* skip, hipStream_t st) {
    switch (n) {
        case 32: return col_n<R, 32>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
        case 64: return col_n<R, 64>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
        case 128: return col_n<R, 128>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
        case 256: return col_n<R, 256>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
        case 512: return col_n<R, 512>(v, sign, S, ncols, inner, outer, estride, comp_stride, tw, skip, st);
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
        default: return 0;
    }
}

template long static_col<double>(int, int, int, cx<double>*, long, long, long, long, long, const cx<double>*,
                                 const int*, hipStream_t);
template long static_col<float>(int, int, int, cx<float>*, long, long, long, long, long, const cx<float>*,
                                const int*, hipStream_t);

}  // namespace fftperm

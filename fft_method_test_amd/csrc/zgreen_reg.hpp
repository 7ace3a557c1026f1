// zgreen_reg.hpp -- fused z pass with register-resident spectra (radix-4 Stockham).
//
// A workgroup owns ZL z-columns; thread t serves column ci = t % ZL and butterfly slot
// j = t / ZL in [0, N/4).  Every stage is radix 4 (one butterfly per thread), except a final
// radix-2 stage when log2 N is odd (two butterflies per thread, j and j + N/4).  The last
// forward stage leaves the thread holding frequencies {j + r N/4, r = 0..3} of its column,
// which is exactly the input set of the first backward stage, so the spectrum never returns to
// LDS between the transforms: the symmetric Green multiply (and the quadratic form) run on
// registers.  The three components pass through one shared single-component LDS buffer in
// turn, so a workgroup needs ZL*(N+1) complex values of LDS instead of 3*ZL*(N+1), which lets
// several workgroups share a CU and overlap each other's memory and transform phases.
#pragma once

#include "fft_engine.hpp"
#include "fft_static.hpp"

namespace fftperm {

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// Radix of the last stage and span before it: stages 4, 4, ..., 4 (, 2).
template <int N>
struct RegPlan {
    static constexpr int LOG = ilog2c(N);
    static constexpr int LAST = (LOG % 2 == 1) ? 2 : 4;
    static constexpr int NS_LAST = N / LAST;
    static_assert((1 << LOG) == N && N >= 16, "power-of-two column length >= 16");
};

// Twiddled radix-RAD butterflies of one stage for the thread's butterflies jb = j + q N/4
// (q < 4/RAD), inputs taken from (and outputs returned in) v[q][r].
template <typename R, int SIGN, int N, int RAD, int NS>
__device__ __forceinline__ void rbf(cx<R> (&v)[4 / RAD][RAD], const cx<R>* tw, int j) {
    constexpr int TS = N / (NS * RAD);
#pragma unroll
    for (int q = 0; q < 4 / RAD; ++q) {
        const int jb = j + q * (N / 4);
        const int k = jb % NS;
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < RAD; ++r) {
                cx<R> w = tw[r * k * TS];
                if (SIGN > 0) w.y = -w.y;
                v[q][r] = cmul<R>(v[q][r], w);
            }
        }
        dft_small<R, SIGN, RAD>(v[q]);
    }
}

template <typename R, int N, int RAD, int NS>
__device__ __forceinline__ void rload(cx<R> (&v)[4 / RAD][RAD], const cx<R>* col, int j) {
    constexpr int M = N / RAD;
#pragma unroll
    for (int q = 0; q < 4 / RAD; ++q)
#pragma unroll
        for (int r = 0; r < RAD; ++r) v[q][r] = col[j + q * (N / 4) + r * M];
}

template <typename R, int N, int RAD, int NS>
__device__ __forceinline__ void rstore(const cx<R> (&v)[4 / RAD][RAD], cx<R>* col, int j) {
#pragma unroll
    for (int q = 0; q < 4 / RAD; ++q) {
        const int jb = j + q * (N / 4);
        const int k = jb % NS;
        const int base = (jb / NS) * NS * RAD + k;
#pragma unroll
        for (int r = 0; r < RAD; ++r) col[base + r * NS] = v[q][r];
    }
}

// Middle radix-4 stages NS = NS0, 4 NS0, ... while 4 NS < NS_LAST: LDS -> LDS.
template <typename R, int SIGN, int N, int NS>
__device__ __forceinline__ void rmiddle(cx<R>* col, const cx<R>* tw, int j) {
    if constexpr (NS < RegPlan<N>::NS_LAST) {
        cx<R> v[1][4];
        rload<R, N, 4, NS>(v, col, j);
        rbf<R, SIGN, N, 4, NS>(v, tw, j);
        __syncthreads();
        rstore<R, N, 4, NS>(v, col, j);
        __syncthreads();
        rmiddle<R, SIGN, N, NS * 4>(col, tw, j);
    }
}

// Forward transform of one column held as registers x[r] = element j + r N/4; returns the
// spectrum in x[r] = frequency j + r N/4.  col = this thread's column in the shared buffer.
template <typename R, int N>
__device__ __forceinline__ void rfft_fwd(cx<R> (&x)[4], cx<R>* col, const cx<R>* tw, int j) {
    using P = RegPlan<N>;
    {
        cx<R> v[1][4] = {{x[0], x[1], x[2], x[3]}};
        rbf<R, -1, N, 4, 1>(v, tw, j);
        rstore<R, N, 4, 1>(v, col, j);
        __syncthreads();
    }
    rmiddle<R, -1, N, 4>(col, tw, j);
    constexpr int RL = P::LAST, NSL = P::NS_LAST;
    cx<R> v[4 / RL][RL];
    rload<R, N, RL, NSL>(v, col, j);
    rbf<R, -1, N, RL, NSL>(v, tw, j);
    // outputs of butterfly q, index r sit at frequency j + q N/4 + r NSL
#pragma unroll
    for (int q = 0; q < 4 / RL; ++q)
#pragma unroll
        for (int r = 0; r < RL; ++r) x[q + r * (NSL / (N / 4))] = v[q][r];
    __syncthreads();  // the buffer is rewritten by the next column transform
}

// Backward transform of a spectrum held as x[r] = frequency j + r N/4; the outputs (element
// j + q N/4 + r NSL of the last stage) go to `emit(e, value)`.
template <typename R, int N, typename F>
__device__ __forceinline__ void rfft_bwd(const cx<R> (&x)[4], cx<R>* col, const cx<R>* tw, int j, F&& emit) {
    using P = RegPlan<N>;
    {
        cx<R> v[1][4] = {{x[0], x[1], x[2], x[3]}};
        rbf<R, +1, N, 4, 1>(v, tw, j);
        rstore<R, N, 4, 1>(v, col, j);
        __syncthreads();
    }
    rmiddle<R, +1, N, 4>(col, tw, j);
    constexpr int RL = P::LAST, NSL = P::NS_LAST;
    cx<R> v[4 / RL][RL];
    rload<R, N, RL, NSL>(v, col, j);
    rbf<R, +1, N, RL, NSL>(v, tw, j);
#pragma unroll
    for (int q = 0; q < 4 / RL; ++q)
#pragma unroll
        for (int r = 0; r < RL; ++r) emit(j + q * (N / 4) + r * NSL, v[q][r]);
    __syncthreads();
}

}  // namespace fftperm

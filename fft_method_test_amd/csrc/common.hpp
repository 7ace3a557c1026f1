// common.hpp -- shared device/host definitions for the fftperm HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fftperm {

// Last-level cache (Infinity Cache) of the device: 256 MiB.  Passes whose data would not stay
// cached beside the spectra of the next pass read or write it with non-temporal accesses.
constexpr long kLastLevelCacheBytes = 256L << 20;

// Complex values are stored interleaved (re, im) as HIP's double2 / float2.
template <typename R> struct CxOf;
template <> struct CxOf<double> { using T = double2; };
template <> struct CxOf<float> { using T = float2; };
template <typename R> using cx = typename CxOf<R>::T;

template <typename R>
__device__ __forceinline__ cx<R> mkc(R re, R im) {
    cx<R> c;
    c.x = re;
    c.y = im;
    return c;
}
// The arithmetic of the transforms and of the Green multiply fixes its own rounding: products that
// are fused say so (fmar), and contraction is off for everything else, so that two kernels built
// from the same helpers round identically (the compiler would otherwise choose per kernel which
// product of a * b - c * d to fuse).
template <typename R>
__device__ __forceinline__ R fmar(R a, R b, R c);
template <>
__device__ __forceinline__ float fmar<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <>
__device__ __forceinline__ double fmar<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) {
#pragma clang fp contract(off)
    return mkc<R>(a.x + b.x, a.y + b.y);
}
template <typename R>
__device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) {
#pragma clang fp contract(off)
    return mkc<R>(a.x - b.x, a.y - b.y);
}
// (a.x b.x - a.y b.y, a.x b.y + a.y b.x) with the first product of each part fused
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
#pragma clang fp contract(off)
    return mkc<R>(fmar<R>(a.x, b.x, -(a.y * b.y)), fmar<R>(a.x, b.y, a.y * b.x));
}
template <typename R>
__device__ __forceinline__ cx<R> conjc(cx<R> a) { return mkc<R>(a.x, -a.y); }
// complex value converted to the precision T (rounded to nearest when narrowing)
template <typename T, typename C>
__device__ __forceinline__ cx<T> cvtc(C a) { return mkc<T>((T)a.x, (T)a.y); }

// One line-transform plan: n = prod(rad[0..nst)), Stockham stages in that order.
struct LinePlan {
    int n;
    int nst;
    int rad[8];
    int maxbf;  // max over stages of n / rad[s] (butterflies per line)
};

constexpr int kMaxWaves = 16;  // 1024 threads per workgroup at most

// Deterministic block sum of NV doubles (wave64 shuffles + LDS).  Result valid in
// thread 0 (written to out[]) when out != nullptr; every thread must call it.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* out) {
    __shared__ double red[NV][kMaxWaves];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = v[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        v[i] = s;
    }
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) red[i][wid] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0 && out != nullptr) {
        const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += red[i][w];
            out[i] = s;
        }
    }
    __syncthreads();
}

}  // namespace fftperm

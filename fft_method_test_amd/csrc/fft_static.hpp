// fft_static.hpp -- compile-time specialised line transforms for the hot grid sizes (the powers of
// two 32 ... 512 and the mixed-radix edges 80, 96, 112, 144, 160 of the uneven-size batch).
//
// Same Stockham algorithm as fft_engine.hpp, with the line length N, the radix sequence,
// the number of lines per workgroup (LINES) and the workgroup size (THREADS) known at
// compile time, so index arithmetic folds to shifts and every register array has a
// static size.  Butterflies are mapped line-fastest (thread t -> line t % LINES), which
// with an LDS pitch of N + 1 complex values keeps the eight lanes that share an LDS
// cycle on distinct banks for both the strided reads and the Stockham writes.  The
// twiddle index is uniform across a group of LINES lanes and is served as an LDS
// broadcast.
#pragma once

#include "fft_engine.hpp"

namespace fftperm {

template <int... Rs>
struct Radices {};

template <int N>
struct StaticPlan;
template <> struct StaticPlan<16> { using type = Radices<16>; };
template <> struct StaticPlan<32> { using type = Radices<8, 4>; };
template <> struct StaticPlan<64> { using type = Radices<8, 8>; };
template <> struct StaticPlan<128> { using type = Radices<16, 8>; };
template <> struct StaticPlan<256> { using type = Radices<16, 16>; };
template <> struct StaticPlan<512> { using type = Radices<8, 8, 8>; };
// Mixed-radix lengths (the uneven-size batch): a radix-16 stage, then the odd / mixed radix.
template <> struct StaticPlan<80> { using type = Radices<16, 5>; };
template <> struct StaticPlan<96> { using type = Radices<16, 6>; };
template <> struct StaticPlan<112> { using type = Radices<16, 7>; };
template <> struct StaticPlan<144> { using type = Radices<16, 9>; };
template <> struct StaticPlan<160> { using type = Radices<16, 10>; };

constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

template <typename P>
struct PlanInfo;
template <int... Rs>
struct PlanInfo<Radices<Rs...>> {
    static constexpr int min_radix() {
        int m = 1 << 30;
        for (int r : {Rs...}) m = r < m ? r : m;
        return m;
    }
    static constexpr int max_radix() {
        int m = 0;
        for (int r : {Rs...}) m = r > m ? r : m;
        return m;
    }
    // Threads per line for length n: the smallest stage's butterfly count when every stage's count
    // is a multiple of it (power-of-two plans: each thread owns M_s / T butterflies of stage s),
    // else the largest count (mixed plans: one butterfly per thread, idle threads in the stages
    // with fewer butterflies).
    static constexpr int threads(int n) {
        int lo = 1 << 30, hi = 0;
        for (int r : {Rs...}) {
            lo = n / r < lo ? n / r : lo;
            hi = n / r > hi ? n / r : hi;
        }
        for (int r : {Rs...})
            if ((n / r) % lo != 0) return hi;
        return lo;
    }
};

// One stage: each thread owns K = ceil((N/RAD)/T) butterflies of one line (the last ones idle when
// T does not divide the stage's butterfly count).  Values are stored in LDS
// as cx<R> and computed in registers as cx<RC> with twiddles from a cx<RT> table; RC = double over
// R = float keeps the FP32 storage but rounds once per stage instead of once per butterfly level.
template <typename R, int SIGN, int N, int RAD, int NS, int LINES, int T, int LD, typename RC, typename RT>
__device__ __forceinline__ void sstage(cx<R>* __restrict__ buf, const cx<RT>* __restrict__ tw, int nlines) {
    constexpr int M = N / RAD;
    constexpr int K = (M + T - 1) / T;
    constexpr bool EVEN = M % T == 0;
    constexpr int TS = N / (NS * RAD);
    const int ci = threadIdx.x % LINES;
    const int jt = threadIdx.x / LINES;
    const bool act = ci < nlines && jt < T;  // workgroups may carry extra threads for the memory phases
    cx<RC> v[K][RAD];
    cx<R>* L = buf + ci * LD;
    if (act) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int j = jt + q * T;
            if (!EVEN && j >= M) continue;
            const int k = j % NS;
#pragma unroll
            for (int r = 0; r < RAD; ++r) v[q][r] = cvtc<RC>(L[j + r * M]);
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < RAD; ++r) {
                    cx<RC> w = cvtc<RC>(tw[r * k * TS]);
                    if (SIGN > 0) w.y = -w.y;
                    v[q][r] = cmul<RC>(v[q][r], w);
                }
            }
            dft_small<RC, SIGN, RAD>(v[q]);
        }
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int j = jt + q * T;
            if (!EVEN && j >= M) continue;
            const int k = j % NS;
            const int base = (j / NS) * NS * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; ++r) L[base + r * NS] = cvtc<R>(v[q][r]);
        }
    }
    __syncthreads();
}

template <typename R, typename RC, typename RT, int SIGN, int N, int LINES, int T, int LD, int NS>
__device__ __forceinline__ void sstages(cx<R>*, const cx<RT>*, int, Radices<>) {}

template <typename R, typename RC, typename RT, int SIGN, int N, int LINES, int T, int LD, int NS, int R0, int... Rs>
__device__ __forceinline__ void sstages(cx<R>* buf, const cx<RT>* tw, int nlines, Radices<R0, Rs...>) {
    sstage<R, SIGN, N, R0, NS, LINES, T, LD, RC, RT>(buf, tw, nlines);
    sstages<R, RC, RT, SIGN, N, LINES, T, LD, NS * R0>(buf, tw, nlines, Radices<Rs...>{});
}

// Transform `nlines` (<= LINES) lines of length N held in LDS at pitch LD.
// T threads per line do the butterflies (threads LINES*T and above idle through the stages).
// Caller synchronised after filling buf / tw.  RC / RT: compute and twiddle precision (default R).
template <typename R, int SIGN, int N, int LINES, int T, int LD, typename RC = R, typename RT = R>
__device__ __forceinline__ void sfft_lines(cx<R>* buf, const cx<RT>* tw, int nlines) {
    sstages<R, RC, RT, SIGN, N, LINES, T, LD, 1>(buf, tw, nlines, typename StaticPlan<N>::type{});
}

// The same with an explicit radix plan (PLAN = Radices<...>, product N).
template <typename R, int SIGN, int N, int LINES, int T, int LD, typename PLAN, typename RC = R, typename RT = R>
__device__ __forceinline__ void sfft_lines_plan(cx<R>* buf, const cx<RT>* tw, int nlines) {
    sstages<R, RC, RT, SIGN, N, LINES, T, LD, 1>(buf, tw, nlines, PLAN{});
}

// Radix <= 8 plans: smaller butterflies (32 live values in FP64 instead of 64) for kernels that
// keep other data in registers across the transform.
template <int N>
struct SmallRadixPlan;
template <> struct SmallRadixPlan<32> { using type = Radices<8, 4>; };
template <> struct SmallRadixPlan<64> { using type = Radices<8, 8>; };
template <> struct SmallRadixPlan<128> { using type = Radices<8, 4, 4>; };
template <> struct SmallRadixPlan<256> { using type = Radices<8, 8, 4>; };
template <> struct SmallRadixPlan<512> { using type = Radices<8, 8, 8>; };

// Threads per line (PlanInfo::threads): for the power-of-two plans the butterfly count of the
// largest-radix stage, so that every stage's butterflies split evenly over them; 16 for the
// mixed-radix plans.
template <int N>
constexpr int threads_per_line() {
    return PlanInfo<typename StaticPlan<N>::type>::threads(N);
}

// Tile variants of the x and column passes.  Variant v in [0, kPassVariants) puts LINES lines in
// a workgroup with OS threads per butterfly thread (the extra threads only widen the memory
// phases); code = 10 * LINES + OS names it in timings.  LINES starts from the count that gives
// 256 butterfly threads and is halved twice.
constexpr int kPassVariants = 5;

template <int N>
constexpr int pass_base_lines() {
    return (256 / threads_per_line<N>()) > 2 ? 256 / threads_per_line<N>() : 2;
}

template <int N>
constexpr int pass_lines(int v) {
    return v < 2 ? pass_base_lines<N>() : (v < 4 ? pass_base_lines<N>() / 2 : pass_base_lines<N>() / 4);
}

constexpr int pass_os(int v) { return (v == 0 || v == 2) ? 2 : 4; }

template <typename R, int N, int LINES, int OS>
struct PassShape {
    static constexpr int T = threads_per_line<N>();
    static constexpr int LD = N + 1;
    static constexpr int THREADS = OS * LINES * T;
    static constexpr size_t LDS = (size_t)(N + LINES * LD) * sizeof(cx<R>);
    // every thread owns ceil(LINES N / THREADS) elements of the tile (the last ones idle when the
    // tile does not split evenly: mixed-radix lengths)
    static constexpr bool valid = LINES >= 2 && LINES % 2 == 0 && THREADS >= 64 && THREADS <= 1024 &&
                                  LDS <= 100u * 1024u;
};

// Merged x-pass variants v in [kPassVariants, kXPassVariants): the two two-for-one rounds of a tile
// of LINES real lines (x + i y of every line, z of line pairs) share one transform of
// LT = 3 LINES / 2 complex lines, so a tile has one load phase, one transform and one store phase.
// LINES = the base count, halved, quartered; OS = 1, 2, 4 keeps 1.5 * base * T threads.
// The last x variant is the wave-per-line-pair pass (passes_x.hip): 8 lines per workgroup, one
// wave per line pair, code 88.
constexpr int kXPassVariants = kPassVariants + 4;

constexpr bool xpass_wave(int v) { return v == kPassVariants + 3; }
constexpr bool xpass_merged(int v) { return v >= kPassVariants && !xpass_wave(v); }

template <int N>
constexpr int xpass_lines(int v) {
    return xpass_wave(v) ? 8
                         : (v < kPassVariants ? pass_lines<N>(v) : (pass_base_lines<N>() >> (v - kPassVariants)));
}

constexpr int xpass_os(int v) {
    return xpass_wave(v) ? 8 : (v < kPassVariants ? pass_os(v) : (1 << (v - kPassVariants)));
}

template <typename R, int N, int LINES, int OS>
struct MergedShape {
    static constexpr int T = threads_per_line<N>();
    static constexpr int LT = 3 * LINES / 2;
    static constexpr int LD = N + 1;
    static constexpr int THREADS = OS * LT * T;
    static constexpr size_t LDS = (size_t)(N + LT * LD) * sizeof(cx<R>);
    static constexpr bool valid = LINES >= 2 && LINES % 2 == 0 && THREADS >= 64 && THREADS <= 1024 &&
                                  THREADS % 64 == 0 && LDS <= 100u * 1024u;
};

// Column-pass variants v in [kPassVariants, kColVariants): twice the base line count (rows of
// 2 * base contiguous values, 256 bytes in FP32 at n = 256) with OS = 1 or 2; in FP64 their LDS
// exceeds the budget and they are skipped.
constexpr int kColVariants = kPassVariants + 2;

template <int N>
constexpr int col_lines(int v) {
    return v < kPassVariants ? pass_lines<N>(v) : 2 * pass_base_lines<N>();
}

constexpr int col_os(int v) { return v < kPassVariants ? pass_os(v) : (v == kPassVariants ? 1 : 2); }

// Block index remap so that consecutive tiles land on the same XCD (workgroups are dealt to
// the 8 XCDs round-robin; each XCD has its own L2).
__device__ __forceinline__ long xcd_tile(long b, long nblocks) {
    constexpr int kXcd = 8;
    const long per = nblocks / kXcd;
    if (per == 0 || b >= per * kXcd) return b;
    return (b % kXcd) * per + b / kXcd;
}

}  // namespace fftperm

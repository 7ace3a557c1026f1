// fft_engine.hpp -- in-LDS Stockham line transforms for gfx950.
//
// A workgroup holds `nlines` lines of length n in LDS (pitch ld complex values)
// and transforms them together, one radix stage at a time.  In each stage every
// active thread owns exactly one butterfly: it reads its `RAD` inputs, applies the
// stage twiddles, performs an in-register RAD-point DFT, waits at a barrier, then
// writes its outputs in Stockham (self-sorting) order.  The host sizes the
// workgroup so that nlines * (n / RAD) <= blockDim for every stage, which is what
// makes the single-buffer in-place update safe.
//
// Twiddles come from a table tw[k] = exp(-2 pi i k / n) held in LDS (computed on
// the host in extended precision and rounded once).  SIGN = -1 is the forward
// (unnormalised, e^{-i...}) transform, SIGN = +1 the backward one.
#pragma once

#include "common.hpp"

namespace fftperm {

// cos / sin of 2 pi m / 16, m = 0..15.
__device__ constexpr double kCos16[16] = {
    1.0,
    0.92387953251128675613,
    0.70710678118654752440,
    0.38268343236508977173,
    0.0,
    -0.38268343236508977173,
    -0.70710678118654752440,
    -0.92387953251128675613,
    -1.0,
    -0.92387953251128675613,
    -0.70710678118654752440,
    -0.38268343236508977173,
    0.0,
    0.38268343236508977173,
    0.70710678118654752440,
    0.92387953251128675613};
__device__ constexpr double kSin16[16] = {
    0.0,
    0.38268343236508977173,
    0.70710678118654752440,
    0.92387953251128675613,
    1.0,
    0.92387953251128675613,
    0.70710678118654752440,
    0.38268343236508977173,
    0.0,
    -0.38268343236508977173,
    -0.70710678118654752440,
    -0.92387953251128675613,
    -1.0,
    -0.92387953251128675613,
    -0.70710678118654752440,
    -0.38268343236508977173};

// v * W_16^E with W = exp(SIGN * 2 pi i / 16); exact for multiples of 4.
template <typename R, int SIGN, int E>
__device__ __forceinline__ cx<R> tw16(cx<R> v) {
    constexpr int e = E & 15;
    if constexpr (e == 0) {
        return v;
    } else if constexpr (e == 4) {  // W^4 = SIGN * i
        return SIGN < 0 ? mkc<R>(v.y, -v.x) : mkc<R>(-v.y, v.x);
    } else if constexpr (e == 8) {
        return mkc<R>(-v.x, -v.y);
    } else if constexpr (e == 12) {
        return SIGN < 0 ? mkc<R>(-v.y, v.x) : mkc<R>(v.y, -v.x);
    } else {
        const R c = (R)kCos16[e];
        const R s = (R)(SIGN < 0 ? -kSin16[e] : kSin16[e]);
        return cmul<R>(v, mkc<R>(c, s));
    }
}

template <typename R, int SIGN>
__device__ __forceinline__ void dft2(cx<R>* v) {
    const cx<R> a = v[0], b = v[1];
    v[0] = cadd<R>(a, b);
    v[1] = csub<R>(a, b);
}

template <typename R, int SIGN>
__device__ __forceinline__ void dft4(cx<R>* v) {
    const cx<R> t0 = cadd<R>(v[0], v[2]);
    const cx<R> t1 = csub<R>(v[0], v[2]);
    const cx<R> t2 = cadd<R>(v[1], v[3]);
    const cx<R> t3 = csub<R>(v[1], v[3]);
    // u = W_4 * t3, W_4 = SIGN * i
    const cx<R> u = SIGN < 0 ? mkc<R>(t3.y, -t3.x) : mkc<R>(-t3.y, t3.x);
    v[0] = cadd<R>(t0, t2);
    v[1] = cadd<R>(t1, u);
    v[2] = csub<R>(t0, t2);
    v[3] = csub<R>(t1, u);
}

template <typename R, int SIGN, int RAD>
__device__ __forceinline__ void dft_small(cx<R>* v);

// RAD = P * Q:  X[k1 + P k2] = sum_n2 W_Q^{n2 k2} W_RAD^{n2 k1} sum_n1 x[Q n1 + n2] W_P^{n1 k1}.
template <typename R, int SIGN, int P, int Q>
__device__ __forceinline__ void dft_pq(cx<R>* v) {
    cx<R> y[P * Q];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
        cx<R> t[P];
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) t[n1] = v[Q * n1 + n2];
        dft_small<R, SIGN, P>(t);
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) y[n2 * P + k1] = t[k1];
    }
    // twiddles W_{PQ}^{n2 k1} = W_16^{n2 k1 16/(PQ)}
#define FFTPERM_TW(N2, K1)                                                                  \
    if constexpr ((N2) < Q && (K1) < P)                                                     \
        y[(N2) * P + (K1)] = tw16<R, SIGN, ((N2) * (K1) * (16 / (P * Q)))>(y[(N2) * P + (K1)]);
    FFTPERM_TW(1, 1) FFTPERM_TW(1, 2) FFTPERM_TW(1, 3)
    FFTPERM_TW(2, 1) FFTPERM_TW(2, 2) FFTPERM_TW(2, 3)
    FFTPERM_TW(3, 1) FFTPERM_TW(3, 2) FFTPERM_TW(3, 3)
#undef FFTPERM_TW
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
        cx<R> t[Q];
#pragma unroll
        for (int n2 = 0; n2 < Q; ++n2) t[n2] = y[n2 * P + k1];
        dft_small<R, SIGN, Q>(t);
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = t[k2];
    }
}

// cos and sin of 2 pi e / D for the odd and mixed radices (D = 3, 5, 6, 7, 9, 10; e <= D / 2; the
// others follow from cos(2 pi (D - e) / D) = cos(2 pi e / D), sin(2 pi (D - e) / D) = -sin(...)).
// Decimal expansions to 24 places, rounded once to the working precision.
__host__ __device__ constexpr double root_cos(int d, int e) {
    switch (d * 16 + e) {
        case 3 * 16 + 1: return -0.5;
        case 5 * 16 + 1: return 0.309016994374947424102293;
        case 5 * 16 + 2: return -0.809016994374947424102293;
        case 6 * 16 + 1: return 0.5;
        case 6 * 16 + 2: return -0.5;
        case 7 * 16 + 1: return 0.623489801858733530525005;
        case 7 * 16 + 2: return -0.222520933956314404288903;
        case 7 * 16 + 3: return -0.900968867902419126236102;
        case 9 * 16 + 1: return 0.766044443118978035202393;
        case 9 * 16 + 2: return 0.173648177666930348851717;
        case 9 * 16 + 3: return -0.5;
        case 9 * 16 + 4: return -0.939692620785908384054109;
        case 10 * 16 + 1: return 0.809016994374947424102293;
        case 10 * 16 + 2: return 0.309016994374947424102293;
        case 10 * 16 + 3: return -0.309016994374947424102293;
        case 10 * 16 + 4: return -0.809016994374947424102293;
        case 3 * 16: case 5 * 16: case 6 * 16: case 7 * 16: case 9 * 16: case 10 * 16: return 1.0;
        default: return __builtin_nan("");  // not tabulated
    }
}
__host__ __device__ constexpr double root_sin(int d, int e) {
    switch (d * 16 + e) {
        case 3 * 16 + 1: return 0.866025403784438646763723;
        case 5 * 16 + 1: return 0.951056516295153572116439;
        case 5 * 16 + 2: return 0.587785252292473129168706;
        case 6 * 16 + 1: return 0.866025403784438646763723;
        case 6 * 16 + 2: return 0.866025403784438646763723;
        case 7 * 16 + 1: return 0.781831482468029808708445;
        case 7 * 16 + 2: return 0.974927912181823607018132;
        case 7 * 16 + 3: return 0.433883739117558120475768;
        case 9 * 16 + 1: return 0.642787609686539326322643;
        case 9 * 16 + 2: return 0.984807753012208059366743;
        case 9 * 16 + 3: return 0.866025403784438646763723;
        case 9 * 16 + 4: return 0.342020143325668733044100;
        case 10 * 16 + 1: return 0.587785252292473129168706;
        case 10 * 16 + 2: return 0.951056516295153572116439;
        case 10 * 16 + 3: return 0.951056516295153572116439;
        case 10 * 16 + 4: return 0.587785252292473129168706;
        case 3 * 16: case 5 * 16: case 6 * 16: case 7 * 16: case 9 * 16: case 10 * 16: return 0.0;
        default: return __builtin_nan("");  // not tabulated
    }
}
// cos / sin of 2 pi e / D for any e (reduced to 0 <= e <= D / 2 by the symmetries above)
__host__ __device__ constexpr double root_cos_any(int d, int e) {
    e %= d;
    return 2 * e == d ? -1.0 : (2 * e < d ? root_cos(d, e) : root_cos(d, d - e));
}
__host__ __device__ constexpr double root_sin_any(int d, int e) {
    e %= d;
    return 2 * e == d ? 0.0 : (2 * e < d ? root_sin(d, e) : -root_sin(d, d - e));
}

// v * W_D^E with W_D = exp(SIGN * 2 pi i / D); exact for E = 0 and 2 E = D.
template <typename R, int SIGN, int D, int E>
__device__ __forceinline__ cx<R> twd(cx<R> v) {
    constexpr int e = E % D;
    if constexpr (e == 0) {
        return v;
    } else if constexpr (2 * e == D) {
        return mkc<R>(-v.x, -v.y);
    } else {
        constexpr double c = root_cos_any(D, e);
        constexpr double s = SIGN < 0 ? -root_sin_any(D, e) : root_sin_any(D, e);
        static_assert(c == c && s == s, "root of unity not tabulated");
        return cmul<R>(v, mkc<R>((R)c, (R)s));
    }
}

// Odd prime radix P (3, 5, 7) with the pairs a_m = v_m + v_{P-m}, b_m = v_m - v_{P-m}:
// X_k = v_0 + sum_m cos(2 pi m k / P) a_m + i SIGN sum_m sin(2 pi m k / P) b_m and X_{P-k} the same
// with the sine sum negated.  Products fused in a fixed order (fmar), contraction off otherwise.
template <typename R, int SIGN, int P>
__device__ __forceinline__ void dft_prime(cx<R>* v) {
#pragma clang fp contract(off)
    constexpr int H = (P - 1) / 2;
    cx<R> a[H + 1], b[H + 1];
#pragma unroll
    for (int m = 1; m <= H; ++m) {
        a[m] = cadd<R>(v[m], v[P - m]);
        b[m] = csub<R>(v[m], v[P - m]);
    }
    cx<R> x0 = v[0];
#pragma unroll
    for (int m = 1; m <= H; ++m) x0 = cadd<R>(x0, a[m]);
    cx<R> o[P];
    o[0] = x0;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        cx<R> c = v[0];
        cx<R> s = mkc<R>((R)0, (R)0);
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            const R cw = (R)root_cos_any(P, m * k);
            const R sw = (R)root_sin_any(P, m * k);
            c = mkc<R>(fmar<R>(cw, a[m].x, c.x), fmar<R>(cw, a[m].y, c.y));
            s = (m == 1) ? mkc<R>(sw * b[m].x, sw * b[m].y)
                         : mkc<R>(fmar<R>(sw, b[m].x, s.x), fmar<R>(sw, b[m].y, s.y));
        }
        // i SIGN s
        const cx<R> is = SIGN < 0 ? mkc<R>(s.y, -s.x) : mkc<R>(-s.y, s.x);
        o[k] = cadd<R>(c, is);
        o[P - k] = csub<R>(c, is);
    }
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = o[k];
}

// Mixed radix P * Q (6 = 3 x 2, 9 = 3 x 3, 10 = 5 x 2) by the same split as dft_pq, with the
// inner twiddles W_{PQ}^{n2 k1} from the root tables.
template <typename R, int SIGN, int P, int Q>
__device__ __forceinline__ void dft_pq_mixed(cx<R>* v) {
    cx<R> y[P * Q];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
        cx<R> t[P];
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) t[n1] = v[Q * n1 + n2];
        dft_small<R, SIGN, P>(t);
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) y[n2 * P + k1] = t[k1];
    }
#define FFTPERM_TWM(N2, K1)                                                                  \
    if constexpr ((N2) < Q && (K1) < P)                                                      \
        y[(N2) * P + (K1)] = twd<R, SIGN, P * Q, (N2) * (K1)>(y[(N2) * P + (K1)]);
    FFTPERM_TWM(1, 1) FFTPERM_TWM(1, 2) FFTPERM_TWM(1, 3) FFTPERM_TWM(1, 4)
    FFTPERM_TWM(2, 1) FFTPERM_TWM(2, 2) FFTPERM_TWM(2, 3) FFTPERM_TWM(2, 4)
#undef FFTPERM_TWM
    static_assert(Q <= 3 && P <= 5, "inner twiddles enumerated for n2 < 3, k1 < 5");
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
        cx<R> t[Q];
#pragma unroll
        for (int n2 = 0; n2 < Q; ++n2) t[n2] = y[n2 * P + k1];
        dft_small<R, SIGN, Q>(t);
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = t[k2];
    }
}

template <typename R, int SIGN, int RAD>
__device__ __forceinline__ void dft_small(cx<R>* v) {
    if constexpr (RAD == 1) {
    } else if constexpr (RAD == 2) {
        dft2<R, SIGN>(v);
    } else if constexpr (RAD == 4) {
        dft4<R, SIGN>(v);
    } else if constexpr (RAD == 8) {
        dft_pq<R, SIGN, 4, 2>(v);
    } else if constexpr (RAD == 16) {
        dft_pq<R, SIGN, 4, 4>(v);
    } else if constexpr (RAD == 3 || RAD == 5 || RAD == 7) {
        dft_prime<R, SIGN, RAD>(v);
    } else if constexpr (RAD == 6) {
        dft_pq_mixed<R, SIGN, 3, 2>(v);
    } else if constexpr (RAD == 9) {
        dft_pq_mixed<R, SIGN, 3, 3>(v);
    } else if constexpr (RAD == 10) {
        dft_pq_mixed<R, SIGN, 5, 2>(v);
    } else {
        static_assert(RAD == 1, "radix without an in-register DFT");
    }
}

// Odd prime radices up to 13: direct O(RAD^2) DFT with roots from the line's twiddle table
// (n is a multiple of RAD, so W_RAD^m = tw[m * n / RAD]).
template <typename R, int SIGN, int RAD>
__device__ __forceinline__ void dft_odd(cx<R>* v, const cx<R>* tw, int n) {
    const int step = n / RAD;
    cx<R> o[RAD];
#pragma unroll
    for (int q = 0; q < RAD; ++q) {
        cx<R> acc = v[0];
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
            cx<R> w = tw[((r * q) % RAD) * step];
            if (SIGN > 0) w.y = -w.y;
            acc = cadd<R>(acc, cmul<R>(v[r], w));
        }
        o[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < RAD; ++q) v[q] = o[q];
}

template <typename R, int SIGN, int RAD>
__device__ __forceinline__ void stockham_stage(cx<R>* buf, int ld, int nlines, int n, int Ns,
                                               const cx<R>* tw) {
    const int m = n / RAD;
    const int total = nlines * m;
    const int t = threadIdx.x;
    cx<R> v[RAD];
    int line = 0, j = 0, k = 0;
    const bool act = t < total;
    if (act) {
        line = t / m;
        j = t - line * m;
        k = j % Ns;
        const cx<R>* L = buf + line * ld;
        const int tstride = n / (Ns * RAD);
#pragma unroll
        for (int r = 0; r < RAD; ++r) v[r] = L[j + r * m];
        if (k != 0) {
#pragma unroll
            for (int r = 1; r < RAD; ++r) {
                cx<R> w = tw[r * k * tstride];
                if (SIGN > 0) w.y = -w.y;
                v[r] = cmul<R>(v[r], w);
            }
        }
        if constexpr (RAD == 3 || RAD == 5 || RAD == 7 || RAD == 11 || RAD == 13) {
            dft_odd<R, SIGN, RAD>(v, tw, n);
        } else {
            dft_small<R, SIGN, RAD>(v);
        }
    }
    __syncthreads();
    if (act) {
        cx<R>* L = buf + line * ld;
        const int base = (j / Ns) * Ns * RAD + k;
#pragma unroll
        for (int r = 0; r < RAD; ++r) L[base + r * Ns] = v[r];
    }
    __syncthreads();
}

// A stage of prime radix RAD > 16: thread t computes output q of butterfly j of its line as the
// direct sum o_q = sum_r W_RAD^{r q} (W^{r k} x_{j + r m}) over the RAD twiddled inputs (one thread
// per output element: nlines * n threads).
template <typename R, int SIGN>
__device__ void stockham_stage_prime(cx<R>* buf, int ld, int nlines, int n, int Ns, int RAD, const cx<R>* tw) {
    const int m = n / RAD;
    const int t = threadIdx.x;
    const bool act = t < nlines * n;
    int line = 0, j = 0, k = 0, q = 0;
    cx<R> acc = mkc<R>((R)0, (R)0);
    if (act) {
        line = t / n;
        const int e = t - line * n;
        j = e / RAD;
        q = e - j * RAD;
        k = j % Ns;
        const cx<R>* L = buf + line * ld;
        const int tstride = n / (Ns * RAD);
        for (int r = 0; r < RAD; ++r) {
            cx<R> x = L[j + r * m];
            if (k != 0 && r != 0) {
                cx<R> w = tw[r * k * tstride];
                if (SIGN > 0) w.y = -w.y;
                x = cmul<R>(x, w);
            }
            cx<R> wq = tw[((r * q) % RAD) * m];
            if (SIGN > 0) wq.y = -wq.y;
            acc = cadd<R>(acc, cmul<R>(x, wq));
        }
    }
    __syncthreads();
    if (act) buf[line * ld + (j / Ns) * Ns * RAD + k + q * Ns] = acc;
    __syncthreads();
}

// Transform nlines lines in LDS.  Caller has synchronised after filling buf and tw.
template <typename R, int SIGN>
__device__ void fft_lines(cx<R>* buf, int ld, int nlines, const LinePlan& pl, const cx<R>* tw) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int rad = pl.rad[s];
        if (rad > 16) {
            stockham_stage_prime<R, SIGN>(buf, ld, nlines, pl.n, Ns, rad, tw);
            Ns *= rad;
            continue;
        }
        switch (rad) {
            case 2: stockham_stage<R, SIGN, 2>(buf, ld, nlines, pl.n, Ns, tw); break;
            case 3: stockham_stage<R, SIGN, 3>(buf, ld, nlines, pl.n, Ns, tw); break;
            case 4: stockham_stage<R, SIGN, 4>(buf, ld, nlines, pl.n, Ns, tw); break;
            case 5: stockham_stage<R, SIGN, 5>(buf, ld, nlines, pl.n, Ns, tw); break;
            case 7: stockham_stage<R, SIGN, 7>(buf, ld, nlines, pl.n, Ns, tw); break;
            case 8: stockham_stage<R, SIGN, 8>(buf, ld, nlines, pl.n, Ns, tw); break;
            case 11: stockham_stage<R, SIGN, 11>(buf, ld, nlines, pl.n, Ns, tw); break;
            case 13: stockham_stage<R, SIGN, 13>(buf, ld, nlines, pl.n, Ns, tw); break;
            default: stockham_stage<R, SIGN, 16>(buf, ld, nlines, pl.n, Ns, tw); break;
        }
        Ns *= rad;
    }
}

}  // namespace fftperm

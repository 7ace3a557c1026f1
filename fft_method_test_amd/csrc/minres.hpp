// minres.hpp -- device-resident MINRES state and the launchers of its vector kernels.
//
// MinresSolver::solve (src/solver/MinresSolver.cpp:46-218) is split into kernels whose
// scalar recurrences (alpha, beta, Givens rotation, phi_bar) are computed on the device
// by single-block "finish" kernels, so an iteration needs no host synchronisation.
// Every kernel of an iteration returns immediately once `done` is set: iterations
// launched past convergence are no-ops, and the iteration count equals the reference's.
#pragma once

#include "common.hpp"

namespace fftperm {

struct MinresState {
    // recurrence (names of src/solver/MinresSolver.cpp:112-121)
    double beta, cs_old, sn_old, cs, sn, phi_bar;
    double rhs_norm, tol;
    double alpha;
    // coefficients of the current iteration's vector update
    double beta_new, c_eps, c_delta, c_rho, c_xcoef, c_inv_beta, c_inv_rho;
    int v_zero, w_zero;
    double rel_res;
    double mean_in[3];   // support mean of v (input projection of the next apply)
    double vsum[3];      // per-component sums of v
    double mean_out[3];  // support mean of the gathered apply output
    long long iterations;
    long long maxit;
    int done;
    int pending_done;
    int converged;
    int pad_;
};

// z = z - mean_out ; partials: sum v.z                     (Algorithm 1 step 17 + alpha)
void launch_mr_sub_dot(double* z, const double* v, MinresState* st, long nb, double* partials, hipStream_t s);
void launch_mr_finish_alpha(const double* partials, MinresState* st, hipStream_t s);
// z = (z_raw - mean_out) - (alpha v + beta v_old) ; partials: sum z^2
void launch_mr_axpy_norm(double* z, const double* v, const double* v_old, MinresState* st, long nb,
                         double* partials, hipStream_t s);
// beta_new, Givens rotation, phi_bar, relative residual estimate, stop flags, history
void launch_mr_givens(const double* partials, MinresState* st, double* history, hipStream_t s);
// v_new = z / beta_new (into the buffer that held v_old), w_next = (v - eps w_prev - delta w) / rho
// (over w_prev), x += (cs_new phi_bar) w_next; the host then swaps (v_old, v) and (w_prev, w).
// partials: per-component sums of v_new
void launch_mr_update(double* x, const double* v, const double* z, double* w_prev_io, const double* w,
                      double* v_new, MinresState* st, long nb, double* partials, hipStream_t s);
// vsum = sums, mean_in = sums / nb ; done = pending_done
void launch_mr_finish_mean_in(const double* partials, MinresState* st, double inv_nb, hipStream_t s);
// mean_out = support sums of the x-inverse partials / nb ; alpha = (sum of the fused z-pass
// quadratic-form partials) / N
void launch_mr_finish_mean_out(const double* xpart, int nxparts, const double* zpart, int nzparts, MinresState* st,
                               double inv_nb, double inv_n, hipStream_t s);

// Conjugate gradients with reliable updates (the MIXED-mode solver): the iterations apply the FP32
// operator and update x and the recursive residual r; whenever ||r|| drops below delta times its
// largest value since the last update, or below tol ||b||, the iteration pauses (done = pause = 1)
// and the host replaces r by the FP64 true residual of the accumulated solution.
struct CgState {
    double rr;           // r.r of the recursive residual the current beta / alpha use
    double rr_new;       // r.r after the latest update
    double alpha, beta, pap;
    double mean_in[3];   // support mean of p (input projection of the next apply)
    double mean_out[3];  // support mean of the raw apply output
    double rnorm_max;    // largest ||r|| since the last reliable update
    double rhs_norm, tol, delta;
    double rel_res;      // ||r|| / ||b|| of the recursive residual
    long long iterations;
    long long maxit;
    int done;            // every CG kernel returns at once while set
    int pause;           // 1: the host owes a reliable update / convergence check
    int breakdown;       // p.A p <= 0
    int pad_;
};

// mean_out, p.A p (Parseval partials of the fused z pass / N), alpha = rr / pAp
void launch_cg_alpha(const double* xpart, int nxparts, const double* zpart, int nzparts, CgState* st, double inv_nb,
                     double inv_n, hipStream_t s);
// x += alpha p ; r -= alpha (q - mean_out) ; partials: r.r
void launch_cg_update(double* x, double* r, const double* p, const double* q, CgState* st, long nb, double* partials,
                      hipStream_t s);
// rr_new, relative residual, history, pause / stop decisions, beta = rr_new / rr
void launch_cg_beta(const double* partials, CgState* st, double* history, hipStream_t s);
// p = r + beta p ; partials: per-component sums of p
void launch_cg_direction(double* p, const double* r, CgState* st, long nb, double* partials, hipStream_t s);
// mean_in = sums / nb
void launch_cg_finish_mean_in(const double* partials, CgState* st, double inv_nb, hipStream_t s);

// Conjugate residuals with reliable updates: for symmetric A the iterates of CR minimise the residual
// norm over the Krylov space, as MINRES's do, but CR carries the residual r explicitly, so it can be
// replaced by the FP64 true residual without a restart.  Per iteration: q = A r (FP32 apply,
// r.A r from the fused z pass), beta = r.Ar / previous r.Ar, p = r + beta p, Ap = (q - mean) +
// beta Ap, alpha = r.Ar / Ap.Ap, x += alpha p, r -= alpha Ap.
struct CrState {
    double s;            // r.A r of the current residual (beta's denominator next time)
    double alpha, beta, apap, rr;
    double mean_in[3];   // support mean of r (input projection of the next apply)
    double mean_out[3];  // support mean of the raw apply output
    double rnorm_max;    // largest ||r|| since the last reliable update
    double rhs_norm, tol, delta;
    double rel_res;      // ||r|| / ||b|| of the recursive residual
    long long iterations;
    long long maxit;
    long long nb;        // support voxels of this side when the sides of a launch differ (0: the launch's)
    int done;            // every CR kernel returns at once while set
    int pause;           // 1: the host owes a reliable update / convergence check
    int breakdown;       // r.A r <= 0 or Ap.Ap <= 0
    int first;           // 1: beta = 0 (p = r, Ap = A r)
    int pad_;
};

// mean_out, s_new = r.A r (Parseval partials / N), beta = first ? 0 : s_new / s
void launch_cr_beta(const double* xpart, int nxparts, const double* zpart, int nzparts, CrState* st, double inv_nb,
                    double inv_n, hipStream_t s);
// p = r + beta p ; Ap = (q - mean_out) + beta Ap ; partials: Ap.Ap
void launch_cr_direction(double* p, double* ap, const double* r, const double* q, CrState* st, long nb,
                         double* partials, hipStream_t s);
// alpha = s / Ap.Ap
void launch_cr_alpha(const double* partials, CrState* st, hipStream_t s);
// x += alpha p ; r -= alpha Ap ; partials: r.r and the per-component sums of r
void launch_cr_update(double* x, double* r, const double* p, const double* ap, CrState* st, long nb,
                      double* partials, hipStream_t s);
// rr, relative residual, history, mean_in, pause / stop decisions
void launch_cr_check(const double* partials, CrState* st, double* history, double inv_nb, hipStream_t s);

// The same CR steps on FP32 Krylov vectors (x, r, p, Ap and the apply output in FP32; scalars,
// reductions and the arithmetic of each element in FP64).  stream: non-temporal vector accesses
// (cr32_streams).
void launch_cr32_direction(float* p, float* ap, const float* r, const float* q, CrState* st, long nb,
                           double* partials, hipStream_t s, bool stream = false);
void launch_cr32_update(float* x, float* r, const float* p, const float* ap, CrState* st, long nb, double* partials,
                        hipStream_t s, bool stream = false);
// Batched forms over nrhs right-hand sides: RHS b uses st[b], the vectors at b * vstride (a
// multiple of 4 elements), the partial rows at b * pstride / xstride / zstride and the history at
// b * hstride.  The single-RHS launchers above are the nrhs = 1 case.
void launch_cr_beta_n(const double* xpart, int nxparts, long xstride, const double* zpart, int nzparts, long zstride,
                      CrState* st, int nrhs, double inv_nb, double inv_n, hipStream_t s);
void launch_cr_alpha_n(const double* partials, long pstride, CrState* st, int nrhs, hipStream_t s);
void launch_cr_check_n(const double* partials, long pstride, CrState* st, double* history, long hstride, int nrhs,
                       double inv_nb, hipStream_t s);
void launch_cr32_direction_n(float* p, float* ap, const float* r, const float* q, CrState* st, long nb, long vstride,
                             double* partials, long pstride, int nrhs, hipStream_t s, bool stream = false);
void launch_cr32_update_n(float* x, float* r, const float* p, const float* ap, CrState* st, long nb, long vstride,
                          double* partials, long pstride, int nrhs, hipStream_t s, bool stream = false);
// Whether the FP32 CR vector kernels stream their vectors: when the vectors one launch touches
// (five arrays per side) and the spectra of the passes do not fit the last-level cache together,
// non-temporal vector accesses keep the spectra cached for the next apply (256^3, three sides:
// 466 MB of vectors, 214 MB of spectra; the first side's y-forward pass 83 -> 66 us).
inline bool cr32_streams(long nb, int nrhs, long spectra_bytes) {
    const long vectors = 5L * nrhs * 3 * nb * (long)sizeof(float);
    return vectors + spectra_bytes > kLastLevelCacheBytes;
}
// xacc += x32, x32 = 0
void launch_accumulate_f32(double* xacc, float* x32, long n, hipStream_t s);
// out = (float) a
void launch_to_f32(const double* a, float* out, long n, hipStream_t s);
// deterministic values in [-1, 1) (the input of the timing applies)
void launch_fill_pattern(float* p, long n, hipStream_t s);
void launch_fill_pattern(double* p, long n, hipStream_t s);

// Generic helpers
void launch_dot_partials(const double* a, const double* b, long n, double* partials, hipStream_t s);
void launch_axpby(double a, const double* x, double b, const double* y, double* out, long n, hipStream_t s);
void launch_scale_copy(const double* x, double scale_div, double* out, long n, hipStream_t s);

}  // namespace fftperm

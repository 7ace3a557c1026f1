/*
 * oracle.h -- CPU FP64 restatement of the FFT permeability hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in fft_method_test_amd/: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  Nothing in the product path
 * links, loads or calls it.
 *
 * It restates, in plain C and in the same order of operations, the algorithm
 * of lcolabella/FFT_Method_Test for the Bignonnet variational Stokes solver
 * and the Brinkman penalisation loop.  Each function cites the reference
 * file:line it follows.  The transform is either a project-written 3D C2C FFT
 * (what the reference gets from FFTW3, src/fft/FFTWBackend.cpp:203-329) or the
 * identity transform that the reference's own FFTWBackend.cpp performs when it
 * is compiled without FFTW (its #else branches at :309-315 and :374-380); the
 * identity mode exists so that the oracle can be compared number for number
 * with the reference code itself, built from its sources into oracle/_ref.
 *
 * Pinning (which check pins which function) is listed in the header comment
 * of oracle.c.
 */
#ifndef FFTPERM_ORACLE_H
#define FFTPERM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_SUPPORT_FULLSOLID = 0, ORC_SUPPORT_INTERFACE = 1 };
enum { ORC_TRANSFORM_FFT = 0, ORC_TRANSFORM_IDENTITY = 1 };

typedef struct orc_ctx orc_ctx;

/* Grid3D + GreenOperator set-up (src/core/Grid3D.hpp:14-21,
 * src/bignonnet/GreenOperator.cpp:27-52).  lx/ly/lz <= 0 means h*n. */
orc_ctx* orc_create(size_t nx, size_t ny, size_t nz, double hx, double hy, double hz,
                    double lx, double ly, double lz, double mu, int p_radius,
                    int transform);
void orc_destroy(orc_ctx* c);
const char* orc_last_error(void);
size_t orc_total_size(const orc_ctx* c);

/* BinaryMedium + ForceSupport (src/geometry/BinaryMedium.cpp:7-43,
 * src/bignonnet/ForceSupport.cpp:9-25). Returns 0 on success. */
int orc_set_medium(orc_ctx* c, const uint8_t* fluid_mask, int support_mode, uint64_t* nb_out);
/* active voxels (NB, ascending) and voxel_to_local (N, -1 = kInvalid). */
int orc_get_support(const orc_ctx* c, uint64_t* active, int64_t* voxel_to_local);

/* Energy-consistent Green tensor, reference layout: 6 arrays of N (xx,xy,xz,yy,yz,zz),
 * flat index over the unshifted frequency k in [0,n)^3. */
void orc_green_tensor(const orc_ctx* c, double* t6);

/* Standalone 3D C2C FFT over an (nz,ny,nx) row-major complex array (interleaved re,im).
 * sign = -1 forward (FFTW_FORWARD), +1 backward; unnormalised. */
int orc_fft3d(size_t nx, size_t ny, size_t nz, double* data, int sign);

/* GreenOperator::apply: u = Re(IFFT(G . FFT f)) (3N SoA in/out). */
void orc_green_apply(orc_ctx* c, const double* f, double* u);
/* GreenOperator::apply_spectral_green_tensor CPU loop (GreenOperator.cpp:80-92): V = T F. */
void orc_spectral_green(size_t N, const double* t6, const double* F, double* V);
/* BackgroundForce::build (3N SoA out). */
void orc_background_force(const orc_ctx* c, const double grad[3], double* f0);
/* VariationalRHS::build (L = 3 NB AoS out). */
void orc_rhs(orc_ctx* c, const double grad[3], double* b);
/* VariationalOperator::apply (L -> L). */
void orc_operator_apply(orc_ctx* c, const double* x, double* y);

typedef struct {
    uint64_t iterations;
    double final_rel_residual;
    int converged;
    uint64_t history_len;
} orc_solve_info;

/* MinresSolver::solve on the variational operator.  x0 may be NULL (empty guess).
 * history (maxit + 1 entries) may be NULL. */
/* Order of the MINRES dot products and support means: 0/1 serial, T > 1 the order of T
 * OpenMP threads (static schedule) as in the reference built with OpenMP.  Process-wide. */
void orc_set_reduction_chunks(int chunks);

int orc_minres(orc_ctx* c, const double* b, const double* x0, double tol, uint64_t maxit,
               double* x, double* history, orc_solve_info* info);

/* MinresSolver::solve on a dense symmetric matrix (row-major n x n), for tests. */
int orc_minres_dense(const double* A, uint64_t n, const double* b, const double* x0,
                     double tol, uint64_t maxit, double* x, double* history,
                     orc_solve_info* info);

/* VelocityRecovery::recover + average_over_domain.  u (3N) may be NULL. */
void orc_recover(orc_ctx* c, const double grad[3], const double* x, double* u, double avg[3]);

/* solve_one_direction (src/app/PermeabilityRunner.cpp:173-242) + the K column of
 * :772-777 with src/post/PermeabilityTensor.cpp:11-18: kcol[r] = K[r][col]. */
int orc_solve_direction(orc_ctx* c, const double grad[3], double tol, uint64_t maxit,
                        double avg[3], orc_solve_info* info);
/* Full tensor: K (row-major 9) = +mu <v_r>^(c) for the three unit gradients. */
int orc_solve_tensor(orc_ctx* c, double tol, uint64_t maxit, double K[9], orc_solve_info info[3]);

/* Diagnostics::residual_norm and relative_symmetry_gap (src/post/Diagnostics.cpp:26-64). */
double orc_residual_norm(orc_ctx* c, const double* x, const double* b);
double orc_symmetry_gap(orc_ctx* c, const double* x, const double* y);

/* Brinkman penalisation (src/brinkman/FluidBrinkman.cpp:282-507).
 * alpha / viscosity: per-voxel coefficient fields (buildCoefficientFields :152-191).
 * u (3N SoA) may be NULL. */
typedef struct {
    int iterations;
    double residual;
    int converged;
} orc_brinkman_info;

int orc_brinkman_solve(size_t nx, size_t ny, size_t nz, const double* alpha,
                       const double* viscosity, double relaxation, double tol,
                       int max_iterations, const double force[3], double U[3], double* u,
                       orc_brinkman_info* info);

#ifdef __cplusplus
}
#endif
#endif

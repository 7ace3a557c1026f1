/*
 * fftperm.h -- C ABI of the MI355X (gfx950) FFT permeability library (libfftperm.so).
 *
 * Drop-in boundary for the hot path of lcolabella/FFT_Method_Test: the Bignonnet
 * variational Stokes solver (support map, energy-consistent Green operator applied
 * by 3D FFTs, MINRES on the interface forces, velocity recovery, K assembly) and the
 * Brinkman penalisation loop.  Plain C types only; no exception crosses this ABI.
 *
 * Conventions (as in the reference):
 *  - grid flat index (k*ny + j)*nx + i                    (src/core/Grid3D.hpp:41-43)
 *  - vector fields are SoA by component: [x(N) | y(N) | z(N)]  (src/core/Field3D.hpp:16-114)
 *  - compact support vectors are AoS: [x0,y0,z0,x1,...], L = 3*NB
 *                                                          (src/bignonnet/TrialForceField.cpp:31)
 *  - complex spectra are interleaved (re, im) doubles, full spectrum of N per component
 *  - fluid mask: 1 = fluid, 0 = solid                     (src/geometry/BinaryMedium.cpp:14-20)
 * Pointers are host memory unless the name ends in _dev (device memory of the context's GPU).
 *
 * Return value: FFTPERM_OK (0) or an error code; fftperm_last_error() gives the message
 * of the last failure on the calling thread.  The codes map the reference's exceptions:
 * std::invalid_argument -> FFTPERM_EINVAL, std::out_of_range -> FFTPERM_ERANGE,
 * std::runtime_error -> FFTPERM_ERUNTIME; HIP failures -> FFTPERM_EDEVICE.
 *
 * Threading: contexts are independent (own HIP stream and buffers); one context must
 * not be used from two threads at once.
 */
#ifndef FFTPERM_H
#define FFTPERM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FFTPERM_OK = 0,
    FFTPERM_EINVAL = 1,
    FFTPERM_ERANGE = 2,
    FFTPERM_ERUNTIME = 3,
    FFTPERM_EDEVICE = 4,
    FFTPERM_EUNSUPPORTED = 5
};

/* SupportMode (src/core/Types.hpp:26-29) */
enum { FFTPERM_SUPPORT_FULLSOLID = 0, FFTPERM_SUPPORT_INTERFACE = 1 };
/* FFTPERM_PRECISION_FP64: every step in FP64 (the reference with FFT_METHOD_PRECISION=DOUBLE).
 * FFTPERM_PRECISION_MIXED: FFTs and Green multiply in FP32 inside the iterative solve; Krylov
 * vectors, inner products, recurrence scalars, true residuals, RHS, velocity recovery and
 * averages in FP64.  The solve ends on the FP64 true residual ||b - A x|| <= tol ||b||. */
enum { FFTPERM_PRECISION_FP64 = 0, FFTPERM_PRECISION_MIXED = 1 };
/* MIXED-mode solvers (fftperm_set_mixed_solver):
 * FFTPERM_MIXED_CR_RELIABLE: conjugate residuals (residual-minimising like MINRES, with an explicit
 *   residual) on the FP32 operator with reliable updates: the recursive residual is replaced by the
 *   FP64 true residual whenever it has fallen by the factor reliable_delta since the last
 *   replacement, keeping the search direction; FP64 Krylov vectors.
 * FFTPERM_MIXED_CR_FP32_VECTORS: the same with its iterated vectors (group solution, residual,
 *   direction and its image) in FP32; the accumulated solution, b, true residuals and every scalar
 *   stay FP64.  The default.
 * FFTPERM_MIXED_CG_RELIABLE: conjugate gradients with the same reliable updates (FP64 vectors).
 * FFTPERM_MIXED_MINRES_REFINED: iterative refinement around restarted FP32-operator MINRES sweeps
 *   (fftperm_set_refinement). */
enum {
    FFTPERM_MIXED_CG_RELIABLE = 0,
    FFTPERM_MIXED_MINRES_REFINED = 1,
    FFTPERM_MIXED_CR_RELIABLE = 2,
    FFTPERM_MIXED_CR_FP32_VECTORS = 3
};

typedef struct fftperm_ctx fftperm_ctx;

/* Grid3D(nx,ny,nz,hx,hy,hz,lx,ly,lz) (src/core/Grid3D.hpp:14-24): l <= 0 means h * n. */
typedef struct {
    uint64_t nx, ny, nz;
    double hx, hy, hz;
    double lx, ly, lz;
} fftperm_grid;

/* SolverResult (src/solver/SolverResult.hpp:10-16) without the solution vector. */
typedef struct {
    uint64_t iterations;        /* MINRES iterations (summed over refinement sweeps in MIXED) */
    double final_rel_residual;  /* recurrence estimate |phi_bar| / ||b|| (FP64), or the true
                                   relative residual of the last refinement sweep (MIXED) */
    int converged;
    uint64_t history_len;       /* entries written to the history array */
    uint64_t refinement_sweeps; /* 0 in FP64 mode */
} fftperm_solve_info;

/* Last error message of the calling thread ("" if none). */
const char* fftperm_last_error(void);

/* Human-readable device description (IFFTBackend::device_info, src/fft/IFFTBackend.hpp:48-49). */
int fftperm_device_info(char* buf, size_t len);

/* GreenOperator construction (src/bignonnet/GreenOperator.cpp:27-52) + IFFTBackend::initialize
 * (src/fft/IFFTBackend.hpp:16): plans the FFT passes and precomputes the symmetrised
 * energy-consistent Green tensor once per context (the reference recomputes it per load case,
 * src/app/PermeabilityRunner.cpp:186-198).  mu <= 0 -> FFTPERM_EINVAL.  Grid edges must factor
 * into 2, 3, 5, 7, 11, 13 (FFTPERM_EUNSUPPORTED otherwise). */
int fftperm_create(const fftperm_grid* grid, double mu, int p_radius, int precision, fftperm_ctx** out);
int fftperm_destroy(fftperm_ctx* ctx);

/* BinaryMedium + ForceSupport (src/geometry/BinaryMedium.cpp:7-12, src/bignonnet/ForceSupport.cpp:9-25):
 * uploads the fluid mask and builds the support on the GPU.  An empty support is not an error
 * here; solves reject it as the runner does (src/app/PermeabilityRunner.cpp:523-526). */
int fftperm_set_medium(fftperm_ctx* ctx, const uint8_t* fluid_mask, int support_mode, uint64_t* nb_out);
int fftperm_set_medium_dev(fftperm_ctx* ctx, const uint8_t* fluid_mask_dev, int support_mode, uint64_t* nb_out);

/* No counterpart in the reference.  The support map of fftperm_set_medium_dev + fftperm_get_support
 * (ForceSupport, src/bignonnet/ForceSupport.cpp:9-47; InterfaceVoxels::compute_interface_only,
 * src/geometry/InterfaceVoxels.cpp:7-43) for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  It needs no context and
 * allocates nothing; its scan scratch (two counters and one int per 1024 voxels, 4 MiB + 256 B) is a
 * static device array in the library, one copy per GPU held until the process ends, which bounds the
 * grid to FFTPERM_NOALLOC_MAX_VOXELS voxels (FFTPERM_ERANGE beyond).
 *   fluid_mask_dev     N = nx*ny*nz bytes on the current device, 1 = fluid
 *   voxel_to_local_dev N int64 (may be NULL): local index on the support, -1 elsewhere (kInvalid)
 *   active_dev         the ascending support voxels; at most active_capacity entries are written
 *   counts_dev         2 uint64 on the device (may be NULL): [support voxels, fluid voxels]
 *   stream             hipStream_t to launch on (NULL: the default stream)
 *   nb_out, nfluid_out host (may be NULL): when either is given the call waits for the stream, reads
 *                      the counts and returns FFTPERM_ERANGE if the support exceeds active_capacity
 * Every call on a GPU shares that GPU's scratch and clears its counters itself; calls that overlap
 * on one GPU (two host threads, or two streams in flight) race and must be serialised by the caller. */
#define FFTPERM_NOALLOC_MAX_VOXELS 1073741824ull /* 1024^3 */
int fftperm_support_noalloc(uint64_t nx, uint64_t ny, uint64_t nz, const uint8_t* fluid_mask_dev, int support_mode,
                            int64_t* voxel_to_local_dev, int64_t* active_dev, uint64_t active_capacity,
                            uint64_t* counts_dev, void* stream, uint64_t* nb_out, uint64_t* nfluid_out);

/* BinaryMedium::porosity (src/geometry/BinaryMedium.cpp:34-43). */
int fftperm_porosity(fftperm_ctx* ctx, double* porosity);

/* ForceSupport::active_voxels (ascending) and voxel_to_local (-1 where the reference stores
 * kInvalid = SIZE_MAX) (src/bignonnet/ForceSupport.hpp:22-29).  Either pointer may be NULL. */
int fftperm_get_support(fftperm_ctx* ctx, uint64_t* active, int64_t* voxel_to_local);

/* GreenOperator::tensor_components for every index: 6 arrays of N (xx,xy,xz,yy,yz,zz) over the
 * unshifted frequencies (src/bignonnet/GreenOperator.cpp:95-204). */
int fftperm_green_tensor(fftperm_ctx* ctx, double* t6);

/* IFFTBackend::forward: real 3N -> full complex spectrum 3N (unnormalised, e^{-i}),
 * (src/fft/IFFTBackend.hpp:18-19, src/fft/FFTWBackend.cpp:203-264). */
int fftperm_fft_forward(fftperm_ctx* ctx, const double* real3, double* spec3);
/* IFFTBackend::inverse: Re(backward C2C) / N (src/fft/IFFTBackend.hpp:21-22, FFTWBackend.cpp:266-329). */
int fftperm_fft_inverse(fftperm_ctx* ctx, const double* spec3, double* real3);
/* IFFTBackend::apply_spectral_green_tensor (src/fft/IFFTBackend.hpp:26-43): V = T F per frequency,
 * T = 6 arrays of N, F and V full complex spectra of 3N. */
int fftperm_spectral_green(fftperm_ctx* ctx, const double* t6, const double* force_spec, double* vel_spec);

/* GreenOperator::apply (src/bignonnet/GreenOperator.cpp:54-65): u = Re IFFT(G FFT f), 3N -> 3N. */
int fftperm_green_apply(fftperm_ctx* ctx, const double* f, double* u);
/* BackgroundForce::build (src/bignonnet/BackgroundForce.cpp:14-60): f0 (3N). */
int fftperm_background_force(fftperm_ctx* ctx, const double grad[3], double* f0);
/* VariationalRHS::build (src/bignonnet/VariationalRHS.cpp:15-31): b (L). */
int fftperm_rhs(fftperm_ctx* ctx, const double grad[3], double* b);
/* VariationalOperator::apply (src/bignonnet/VariationalOperator.cpp:19-41): y = A x (L -> L).
 * FP64 in every precision mode. */
int fftperm_operator_apply(fftperm_ctx* ctx, const double* x, double* y);
/* The same operator through the FP32 five-sweep apply that MIXED-mode MINRES sweeps use (FP32
 * spectrum and Green tensor, FP64 compact vectors and sums).  No counterpart in the reference: a
 * hook to measure the FP32 operator's error against fftperm_operator_apply.  MIXED contexts only
 * (FFTPERM_EINVAL otherwise). */
int fftperm_operator_apply_fp32(fftperm_ctx* ctx, const double* x, double* y);
/* The FP32 apply as the default MIXED solver (conjugate residuals on FP32 vectors) iterates it: x
 * rounded to an FP32 compact vector (its support mean taken in FP64), the FP32 sweeps, the FP32
 * compact output widened to y and centred.  On 256 x 256 x nz grids this is the apply whose x and
 * y passes may run fused over z planes (FFTPERM_PASS_PLANE).  No counterpart in the reference;
 * MIXED contexts only (FFTPERM_EINVAL otherwise). */
int fftperm_operator_apply_fp32v(fftperm_ctx* ctx, const double* x, double* y);
/* Host-only check of a grid shape against the FFT plans (touches no device).  Every length is
 * planned, as by FFTW, the reference's backend (src/fft/FFTWBackend.cpp:176-192): factors 2..16 as
 * Stockham radices, any other prime as a stage of direct DFT sums with one thread per output.  A plan
 * needs at most eight stages and at most 1024 threads per line in a stage (so an edge with a prime
 * factor above 16 is at most 1024), and the z edge's plan three times that (the fused z pass
 * transforms three components together: an edge with a prime factor above 16 is at most 341 along
 * z); otherwise FFTPERM_EUNSUPPORTED with a message naming the edge.  A non-positive edge ->
 * FFTPERM_EINVAL.  The context creation applies the same rules plus the device's LDS limits. */
int fftperm_check_edges(int nx, int ny, int nz);
/* HBM copy probe (a measurement hook, no reference counterpart): reps copies of the context's three
 * padded FP32 half spectra (spectrum slot 0 -> slot 1) in an access pattern: 0 contiguous 16-byte
 * loads and stores (the bandwidth the device reaches), 1 the y passes' pattern (32-element row
 * segments down the columns), 2 the z pass's pattern (8-element row segments down the z columns).
 * *gbps = (bytes read + written per copy) / (HIP-event time / reps).  Overwrites the spectrum
 * slots (no solver state lives there between calls).  MIXED contexts only; an unknown pattern or
 * reps < 1 -> FFTPERM_EINVAL. */
int fftperm_probe_copy(fftperm_ctx* ctx, int pattern, int reps, double* gbps);

/* MinresSolver(tol, maxit).solve(op, b, x0) (src/solver/MinresSolver.cpp:46-218) on the variational
 * operator.  x0 may be NULL (the reference's empty guess).  history (maxit + 1 entries) may be NULL.
 * tol <= 0 -> FFTPERM_EINVAL.  In MIXED mode the context's mixed solver runs instead (CR with
 * reliable updates on FP32 vectors by default) until the FP64 true residual ||b - A x|| <= tol ||b|| (or maxit
 * iterations in total); history then holds the true residuals at each FP64 check and
 * info->refinement_sweeps counts those checks after the initial one. */
int fftperm_minres(fftperm_ctx* ctx, const double* b, const double* x0, double tol, uint64_t maxit,
                   double* x, double* history, fftperm_solve_info* info);

/* MIXED-mode precision schedule (no counterpart in the reference, which builds one precision per
 * binary).  Every sweep is a MINRES run on r = b - A x to the recurrence residual
 * max(inner_floor, 0.5 tol / r), r = the current FP64 true relative residual; sweep 1 applies the
 * FP64 operator while its recurrence residual is above fp32_below and the FP32 operator below it
 * (0: FP64 throughout, >= 1: FP32 throughout), later sweeps FP32 throughout; max_sweeps sweeps at
 * most.  Defaults 2 (FP32 throughout), 3e-6, 64.  fp32_below < 0, inner_floor outside (0, 1) or
 * max_sweeps == 0 -> FFTPERM_EINVAL. */
int fftperm_set_refinement(fftperm_ctx* ctx, double fp32_below, double inner_floor, uint32_t max_sweeps);

/* MIXED-mode solver choice (no counterpart in the reference): FFTPERM_MIXED_CR_FP32_VECTORS
 * (default), FFTPERM_MIXED_CR_RELIABLE or FFTPERM_MIXED_CG_RELIABLE (reliable_delta in (0, 1), default 0.1),
 * or FFTPERM_MIXED_MINRES_REFINED (reliable_delta ignored).  Anything else -> FFTPERM_EINVAL. */
int fftperm_set_mixed_solver(fftperm_ctx* ctx, int solver, double reliable_delta);

/* VelocityRecovery::recover + average_over_domain (src/bignonnet/VelocityRecovery.cpp:12-47):
 * u (3N, may be NULL) and its domain average avg[3]. */
int fftperm_recover(fftperm_ctx* ctx, const double grad[3], const double* x, double* u, double avg[3]);

/* solve_one_direction (src/app/PermeabilityRunner.cpp:173-242): RHS, MINRES (x0 empty unless given),
 * recovery, domain average.  u (3N) may be NULL. */
int fftperm_solve_direction(fftperm_ctx* ctx, const double grad[3], double tol, uint64_t maxit,
                            const double* x0, double avg[3], double* u, fftperm_solve_info* info);

/* PermeabilityRunner::run K assembly for gradient=all (src/app/PermeabilityRunner.cpp:740-777 with
 * src/post/PermeabilityTensor.cpp:11-18): K (row-major 9) = +mu <v_r>^(c).  info[3] may be NULL.
 * The reference solves the three load directions one after the other (:186-198); with the MIXED
 * default solver (FFTPERM_MIXED_CR_FP32_VECTORS) this library iterates them together (see
 * fftperm_set_tensor_batch), with identical results. */
int fftperm_solve_tensor(fftperm_ctx* ctx, double tol, uint64_t maxit, double K[9], fftperm_solve_info* info);

/* Multi-RHS tensor solves (no counterpart in the reference).  With the MIXED solver
 * FFTPERM_MIXED_CR_FP32_VECTORS, fftperm_solve_tensor and fftperm_solve_tensor_warm run the three
 * directions' CR iterations in lockstep: one launch of each vector and scalar kernel serves the
 * three.  mode 1 (the default): each direction's five FFT passes run back to back on one spectrum
 * buffer that the three share (it stays in the last-level cache from one pass and one direction to
 * the next); mode 2: one launch per pass for the
 * three (one y launch for the nine spectrum components, one fused z launch reading the Green tensor
 * once); mode 0: one direction after the other.  Every direction keeps its own recurrence, reliable
 * updates and stopping test, so iterations, residuals and K are bit-identical in all modes.  Other
 * precision modes and solvers ignore it.  mode outside 0 .. 2 -> FFTPERM_EINVAL. */
int fftperm_set_tensor_batch(fftperm_ctx* ctx, int mode);

/* Batched solves of many media of one grid shape (no counterpart in the reference, which solves
 * one medium per run; the building block of the uneven-size batch: one call per size class).  The
 * count media (fluid_masks: count host pointers to N bytes each) are solved for one load direction
 * grad together: every Krylov iteration advances all of them, the y and z FFT passes of an iteration
 * are one launch for all media (the fused z pass reads the Green tensor once per column tile for
 * all of them), the vector and scalar kernels one launch each, and the iterations replay captured
 * hipGraphs.  Each medium keeps its own support map, compact vectors, CR state, reliable FP64
 * updates and stopping test (FFTPERM_MIXED_CR_FP32_VECTORS on a MIXED context, which is required;
 * FFTPERM_EINVAL otherwise), so its result is the one fftperm_solve_direction gives for it.
 * k_out (count x 3) receives each medium's column +mu <v> (K column of grad), info (count, may be
 * NULL) each solve.  A medium with an empty support gets k = 0 and converged = 0.  The context is
 * left without a medium.  count in 1 .. FFTPERM_BATCH_MAX. */
#define FFTPERM_BATCH_MAX 64
int fftperm_solve_batch(fftperm_ctx* ctx, uint32_t count, const uint8_t* const* fluid_masks, int support_mode,
                        const double grad[3], double tol, uint64_t maxit, double* k_out, fftperm_solve_info* info);

/* Diagnostics::residual_norm / relative_symmetry_gap (src/post/Diagnostics.cpp:26-64), FP64. */
int fftperm_residual_norm(fftperm_ctx* ctx, const double* x, const double* b, double* out);
int fftperm_symmetry_gap(fftperm_ctx* ctx, const double* x, const double* y, double* out);

/* Live kernel timing (no counterpart in the reference; a measurement hook).  While enabled, the
 * context times every fused z-pass launch (z forward FFT, Green multiply, z backward FFT) with HIP
 * events that the launch itself records at the kernel's start and end (hipExtLaunchKernelGGL; the
 * one-launch z pass of tensor-batch mode 2 with events recorded around it), and every five-sweep
 * Green apply (of one right-hand side, or of all sides of a multi-RHS iteration) with events
 * around it on the context's stream; Krylov applies that ran after convergence (no-ops) are not
 * counted.  fftperm_profile(ctx, mode) resets and sets the mode: 0 off; 1 every Krylov
 * iteration is timed (launched eagerly, no graph replay); 2 sampled: one chunk of iterations in
 * eight is timed and the others replay graphs without events.  Other modes -> FFTPERM_EINVAL. */
typedef struct {
    double zgreen_ms[2];         /* [0] FP64, [1] FP32 (mixed) : summed durations */
    uint64_t zgreen_launches[2];
    double apply_ms[2];
    uint64_t apply_launches[2];
    uint64_t N, Nh, nb;          /* grid size, half-spectrum size, support voxels */
    int zgreen_tile[2];          /* fused z-pass variant in use: 10*columns + threads per butterfly [FP64, FP32]; 0 = runtime kernel */
    int xfwd_tile[2];            /* x-forward variant in use: 10*lines + threads per butterfly [FP64, FP32]; 0 = runtime kernel;
                                    1000 (FP32) = x and y passes fused over z planes (the x, x-inverse and y codes all read 1000) */
    int xinv_tile[2];            /* x-inverse variant in use, same code */
    int ypass_tile[2];           /* y-pass variant in use, same code */
    uint64_t zgreen_rhs[2];      /* right-hand sides the counted z-pass launches transformed (a multi-RHS launch counts each side it served) */
    uint64_t apply_rhs[2];       /* right-hand sides the counted applies served */
} fftperm_profile_stats;

/* Pin the tile variant a pass uses on this context's grid shape, process-wide (no counterpart in the
 * reference; the autotuner chooses these otherwise).  kind: FFTPERM_PASS_XFWD, _XINV, _YPASS,
 * _ZGREEN; precision: FFTPERM_PRECISION_FP64 or _MIXED (the FP32 transforms); code as reported in
 * fftperm_profile_stats (10 * lines or columns per workgroup + threads per butterfly thread, or + 4
 * for the register-resident z pass); code 0 returns the choice to the autotuner.  An unknown code
 * -> FFTPERM_EINVAL; a known code that does not fit the device runs the generic kernel.
 * kind FFTPERM_PASS_PLANE (precision _MIXED only): code 1 = separate x and y passes, code 2 / 3 = the
 * FP32 x and y passes of the solver's applies fused over z planes with 8 / 16 waves per workgroup
 * (grids with nx = ny = 256; other shapes, sources and sinks keep the separate passes). */
enum { FFTPERM_PASS_XFWD = 0, FFTPERM_PASS_XINV = 1, FFTPERM_PASS_YPASS = 2, FFTPERM_PASS_ZGREEN = 3,
       FFTPERM_PASS_PLANE = 4 };
int fftperm_pin_variant(fftperm_ctx* ctx, int kind, int precision, int code);

/* MINRES iterations replay hipGraphs of two captured iterations (default on; no counterpart in the
 * reference).  enable = 0 launches every kernel from the host instead; results are identical.
 * Iterations that fftperm_profile times launch eagerly (the live timing records events). */
int fftperm_set_graphs(fftperm_ctx* ctx, int enable);

/* Device time of the operator apply (no counterpart in the reference; a measurement hook): reps
 * back-to-back five-sweep applies of the solver's form on a patterned input held on the device
 * (values in [-1, 1) from a hash of the index; FP32 compact vectors for precision
 * FFTPERM_PRECISION_MIXED, which needs a MIXED context; FP64 otherwise), after one untimed apply;
 * *ms_per_apply = HIP-event time / reps.  The context's iterate is zero again afterwards.  Needs a
 * medium with a non-empty support; reps < 1 -> FFTPERM_EINVAL. */
int fftperm_time_apply(fftperm_ctx* ctx, int precision, int reps, double* ms_per_apply);

int fftperm_profile(fftperm_ctx* ctx, int enable);
int fftperm_profile_read(fftperm_ctx* ctx, fftperm_profile_stats* out);

/* ------------------------------------------------------------------ learned warm start
 * No counterpart in the reference.  A 4-level 3D U-Net (16 base channels, BF16 weights and
 * activations, FP32 accumulation on the matrix cores) maps a periodic fluid mask to the interface
 * force densities of the three unit load directions of the variational problem in one forward pass
 * (output channels 3d, 3d + 1, 3d + 2: the x, y, z components for load direction d).
 * Grids: cubic, edge a positive multiple of 8. */
typedef struct fftperm_unet fftperm_unet;

int fftperm_unet_create(uint64_t seed, fftperm_unet** out);
int fftperm_unet_destroy(fftperm_unet* net);
int fftperm_unet_num_params(fftperm_unet* net, uint64_t* count);
/* flat FP32 parameters: per layer the weights [cout][27 or 1 taps][cin] then the biases [cout] */
int fftperm_unet_get_params(fftperm_unet* net, float* params);
int fftperm_unet_set_params(fftperm_unet* net, const float* params);
/* Forward, mean squared error over the support voxels of the nine predicted components, backward;
 * with lr > 0 one Adam step (beta 0.9 / 0.999).  fluid_mask_dev (n^3 uint8, fluid = 1) and
 * target_dev (n^3 x 9 FP32: per voxel the force densities of load directions x, y, z, three
 * components each, on support voxels) are device pointers.  loss and grads (num_params FP32, may
 * be NULL) receive the loss and its gradient before the step. */
int fftperm_unet_loss_grad(fftperm_unet* net, const uint8_t* fluid_mask_dev, const float* target_dev, uint32_t n,
                           float lr, double* loss, float* grads);
/* Prediction for the three load directions: field_dev (device) n^3 x 9 FP32, layout as the target. */
int fftperm_unet_predict(fftperm_unet* net, const uint8_t* fluid_mask_dev, uint32_t n, float* field_dev);
/* fftperm_solve_tensor with the network's prediction (one forward pass for the three directions;
 * direction d's density p gathered on the support) seeding each direction's solve: x0 = alpha p with
 * alpha = <b, A p> / ||A p||^2 (the minimal-residual scale of the predicted shape, one FP64 apply).  MINRES stays the judge: a guess whose FP64 true residual is
 * not below ||b|| is dropped and that direction starts cold.  used_guess[3] (may be NULL) reports
 * the choice. */
int fftperm_solve_tensor_warm(fftperm_ctx* ctx, fftperm_unet* net, double tol, uint64_t maxit, double K[9],
                              fftperm_solve_info* info, int* used_guess);

/* Brinkman penalisation: common::fluid::solveBrinkman (src/brinkman/FluidBrinkman.cpp:282-507) for
 * one forcing vector on an (nx,ny,nz) grid.  alpha / viscosity: per-voxel coefficient fields
 * (buildCoefficientFields, :152-191: alpha = penalty in solid and 0 in fluid; viscosity = the fluid's
 * in fluid and 1 in solid).  precision: FFTPERM_PRECISION_FP64 or FFTPERM_PRECISION_MIXED (FP32
 * pressure-Poisson sweeps, FP64 velocity/pressure/residual).  U[3] = superficial velocity; u (3N) may
 * be NULL.  Non-finite residual -> FFTPERM_ERUNTIME, as the reference throws (:473-477). */
typedef struct {
    int iterations;
    double residual;
    int converged;
} fftperm_brinkman_info;

int fftperm_brinkman_solve(uint64_t nx, uint64_t ny, uint64_t nz, const double* alpha, const double* viscosity,
                           double relaxation, double tol, int max_iterations, const double force[3], int precision,
                           double U[3], double* u, fftperm_brinkman_info* info);

#ifdef __cplusplus
}
#endif
#endif

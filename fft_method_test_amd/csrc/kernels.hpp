// kernels.hpp -- launchers of the fftperm device kernels (definitions in kernels_*.hip).
#pragma once

#include "common.hpp"

#include <utility>
#include <vector>

namespace fftperm {

// ---------------------------------------------------------------- FFT passes
// Source of the x-forward pass (one real value per voxel and component):
//   field != null : value = field[c*N + v]                       (raw Green apply)
//   otherwise     : value = f0 + fx with
//                   f0 = a_fluid[c] on fluid voxels, a_support[c] on support voxels, else 0
//                   fx = compact[3*l + c] - mean[c] on support voxel l (when compact != null)
// `code` holds the support map: l >= 0 support local index, -1 fluid, -2 other solid.
struct XSource {
    const int32_t* code;
    const double* compact;
    const double* mean;  // device [3], may be null (= 0)
    const double* field;
    double a_fluid[3];
    double a_support[3];
};

// Sink of the x-inverse pass.  value = Re(line) * inv_n.
//   compact  != null : compact[3*l + c] = value on support voxels
//   field    != null : field[c*N + v] = value
//   partials != null : per-block sums (kXinvParts wide): {support sums x,y,z, all-voxel sums x,y,z,
//                      sum over the support of dotv[3l+c] * value (when dotv != null), 0}
struct XSink {
    const int32_t* code;
    double* compact;
    double* field;
    double* partials;
    const double* dotv;
};
constexpr int kXinvParts = 8;

struct PassCfg {
    int lines;
    int threads;
    size_t lds;
    long blocks;
};

struct Geometry3 {
    int nx, ny, nz, nxh;
    long N, Nh;
};

template <typename R>
struct FFTTables {
    LinePlan px, py, pz;
    const cx<R>* twx;
    const cx<R>* twy;
    const cx<R>* twz;
};

PassCfg cfg_xpass(const LinePlan& p, long nlines);
PassCfg cfg_colpass(const LinePlan& p, long ncols);
PassCfg cfg_zgreen(const LinePlan& p, long ncols);
void set_lds_limit(size_t bytes);
size_t lds_limit();

template <typename R>
void launch_xfwd(const XSource& src, cx<R>* S, const Geometry3& g, const FFTTables<R>& t,
                 const int* skip, hipStream_t st);
template <typename R>
void launch_xinv(const cx<R>* S, const XSink& sink, const Geometry3& g, const FFTTables<R>& t,
                 const int* skip, hipStream_t st);
// y-axis transform of all three components (sign -1 forward, +1 backward).
template <typename R>
void launch_ypass(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                  hipStream_t st);
// z-axis transform without the Green multiply (standalone FFT hooks).
template <typename R>
void launch_zpass(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                  hipStream_t st);
// Fused z forward transform, symmetric Green multiply, z backward transform.
// G6 = the symmetrised tensor on the half spectrum, AoS: G6[6*i + g], g = xx,xy,xz,yy,yz,zz.
template <typename R>
void launch_zgreen(cx<R>* S, const R* G6, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                   hipStream_t st);
long xinv_blocks(const Geometry3& g, const LinePlan& px);
// Times every tile width of the fused z pass that fits the device on the given buffers and
// caches the fastest for (precision, nz, columns); returns (width, ms) pairs (empty when the
// size has no specialised kernel or was tuned before).  Overwrites S.
// Autotuned z-pass variant (10 * columns per workgroup + threads per butterfly) for (precision
// 0 = FP64 / 1 = FP32, nz, columns), 0 if untuned.
int zgreen_tile(int prec, int n, long ncols);
template <typename R>
std::vector<std::pair<int, float>> autotune_zgreen(cx<R>* S, const R* G6, const Geometry3& g,
                                                   const FFTTables<R>& t, hipStream_t st);

// ------------------------------------------------------------ setup kernels
// Support classification + ascending compaction.  Fills code (N) and active (NB);
// nb_out / nf_out are device counters (uint64).  scratch must hold scan_scratch_bytes(N).
size_t scan_scratch_bytes(long N);
void launch_support(const uint8_t* mask, int nx, int ny, int nz, int interface_only, int32_t* code,
                    int64_t* active, unsigned long long* counts /*[2]: nb, nf*/, void* scratch,
                    hipStream_t st);

// Green tensor tables: per axis, for k in [0,n) and p in [-P,P]: s2 = sinc^2, q.
struct GreenAxisTables {
    const double* s2x;
    const double* qx;
    const double* s2y;
    const double* qy;
    const double* s2z;
    const double* qz;
    int P;
    double mu;
};
// Half-spectrum symmetrised tensor 0.5 (G(k) + G(-k)), AoS G6[6*i + g], in double (and float if
// G6f != null).
void launch_green_sym(const GreenAxisTables& t, int nx, int ny, int nz, double* G6, float* G6f,
                      hipStream_t st);
// Reference-layout tensor over the full unshifted spectrum (test hook).
void launch_green_full(const GreenAxisTables& t, int nx, int ny, int nz, double* T6, hipStream_t st);

// Hermitian expansion of the half spectrum to the full spectrum (IFFTBackend::forward layout).
template <typename R>
void launch_expand_full(const cx<R>* S, double2* F, const Geometry3& g, hipStream_t st);
// Half-spectrum H(k) = (X(k) + conj(X(-k))) / 2 from a full spectrum X (IFFTBackend::inverse).
template <typename R>
void launch_hermitian_half(const double2* X, cx<R>* S, const Geometry3& g, hipStream_t st);
// Pointwise spectral Green multiply over full spectra (apply_spectral_green_tensor hook).
void launch_spectral_green_full(const double* T6, const double2* F, double2* V, long N, hipStream_t st);

// ------------------------------------------------------------ reductions
constexpr int kRedBlocks = 4096;
constexpr int kRedThreads = 256;
// Sum of nparts consecutive NV-vectors in partials -> out[NV] (single block, fixed order).
void launch_finish_sum(const double* partials, int nparts, int nv, double scale, double* out,
                       const int* skip, hipStream_t st);
// Per-component sums of an AoS compact vector (L = 3*nb) -> partials (kRedBlocks x 3).
void launch_compact_sums(const double* x, long nb, double* partials, const int* skip, hipStream_t st);
// y = sign * (x - mean) (AoS compact).
void launch_sub_mean(const double* x, const double* mean, double* y, long nb, double sign, const int* skip,
                     hipStream_t st);
// Field shift + sum: f[c*N + i] += s_c; partial sums (kRedBlocks x 3) of the shifted field.
void launch_field_shift_sum(double* f, double s0, double s1, double s2, long N, double* partials,
                            hipStream_t st);
// Support means of a field: partial sums over support voxels (active list).
void launch_support_sums(const double* f, const int64_t* active, long nb, long N, double* partials,
                         hipStream_t st);
struct Coef3 {
    double v[3];
};
// BackgroundForce::build into field (3N): a_fluid on fluid voxels, a_support on support voxels.
void launch_background_force(const int32_t* code, long N, Coef3 a_fluid, Coef3 a_support, double* field,
                             hipStream_t st);

}  // namespace fftperm

// kernels.hpp -- launchers of the fftperm device kernels (definitions in kernels_*.hip).
#pragma once

#include "common.hpp"

#include <string>

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
// Line-compressed support map (specialised x passes): bit i % 32 of word i / 32 of line l is set
// when voxel (l, i) is on the support (lines padded to ceil(nx / 32) words); lstart[l] = number of
// support voxels before line l, nb the total.  The support index of a voxel is lstart[line] + the
// set bits before it in its line, which is what `code` stores.
struct LineBits {
    const uint32_t* bits;
    const int32_t* lstart;
    long nb;
};

// With a_fluid = a_support = 0 and `lines.bits` set, the specialised forward pass reads the
// support from the line bitmask instead of `code` and stages each tile's contiguous run of the
// compact vector through LDS.
struct XSource {
    const int32_t* code;
    const double* compact;
    const float* compact32;  // FP32 compact vector (instead of compact; at most one of the two)
    const double* mean;  // device [3], may be null (= 0)
    const double* field;
    double a_fluid[3];
    double a_support[3];
    LineBits lines;
};

// Sink of the x-inverse pass.  value = Re(line) * inv_n.
//   compact  != null : compact[3*l + c] = value on support voxels
//   field    != null : field[c*N + v] = value
//   partials != null : per-block sums (kXinvParts wide): the support sums of x, y, z
struct XSink {
    const int32_t* code;
    double* compact;
    float* compact32;  // FP32 compact output (instead of compact; at most one of the two)
    double* field;
    double* partials;
    LineBits lines;  // when set, the specialised inverse pass reads the support from it
};
constexpr int kXinvParts = 3;

struct PassCfg {
    int lines;
    int threads;
    size_t lds;
    long blocks;
};

// Half-spectrum layout: element (c, kz, ky, kx), kx < nxh, at c * Nc + kz * pz + ky * rp + kx.  The
// row pitch rp rounds nxh up to a multiple of 8 (rows start on 64-byte boundaries in FP32, 128 in
// FP64) and the plane pitch pz = ny * rp + 16 breaks the power-of-two plane stride; both lift the
// HBM throughput of the strided column passes (a probe of the y and z access patterns at 256^3:
// 3.4 -> 4.2 TB/s and 2.2 -> 3.8 TB/s in FP32).  Padding elements hold zeros: the x passes
// never touch them, the y and z passes transform the row padding's zeros back into zeros (so that
// their column tiles stay aligned) and the z pass multiplies them by the Green tensor, which shares
// the layout (6 reals per element, zero in the padding); the plane padding is never accessed.  Nh = nz * ny * nxh counts the frequencies.
struct Geometry3 {
    int nx, ny, nz, nxh;
    long N, Nh;
    int rp;     // row pitch (elements)
    long pz;    // plane pitch (elements)
    long Nc;    // component stride = nz * pz
};

// Row pitch and plane pitch of the padded layout for a grid (set by the context).
inline int spec_row_pitch(int nxh) { return (nxh + 7) / 8 * 8; }
inline long spec_plane_pitch(int ny, int rp) { return (long)ny * rp + 16; }

// The fused z pass reads the Green tensor with non-temporal loads when a context's three half
// spectra fit the last-level cache but do not fit next to the tensor it reads (FP32: half of the
// planes, the others by mirror symmetry), so that the spectra the pass writes are still cached
// when the y-inverse pass reads them (FP32 at 256^3: 214 MB of spectra beside 107 MB of tensor).
template <typename R>
inline bool zpass_streams_green(const Geometry3& g) {
    const long spectra = 3 * g.Nc * 2 * (long)sizeof(R);
    const long green = (sizeof(R) == 4 ? 3 : 6) * g.Nc * (long)sizeof(R);
    return spectra <= kLastLevelCacheBytes && spectra + green > kLastLevelCacheBytes;
}

// The spectrum layout as the x passes see it: x line l = kz * ny + ky starts at kz * pz + ky * rp.
struct SpecLayout {
    long nc;  // component stride
    long pz;  // plane pitch
    int rp;   // row pitch
    int ny;
};

inline SpecLayout spec_layout(const Geometry3& g) { return SpecLayout{g.Nc, g.pz, g.rp, g.ny}; }

__device__ __forceinline__ long spec_line(const SpecLayout& L, long line) {
    const int l = (int)line;  // ny * nz < 2^31 x lines
    return (long)(l / L.ny) * L.pz + (long)(l % L.ny) * L.rp;
}

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
// compute units and LDS per compute unit of the device (sizes the persistent z-pass grid)
void set_device_shape(int cus, size_t lds_per_cu);
size_t lds_limit();
int device_cus();
size_t lds_per_cu();

template <typename R>
void launch_xfwd(const XSource& src, cx<R>* S, const Geometry3& g, const FFTTables<R>& t,
                 const int* skip, hipStream_t st);
// Returns the number of workgroups (= partial-sum rows written when sink.partials != null).
template <typename R>
long launch_xinv(const cx<R>* S, const XSink& sink, const Geometry3& g, const FFTTables<R>& t,
                 const int* skip, hipStream_t st);
// y-axis transform of all three components (sign -1 forward, +1 backward).
template <typename R>
void launch_ypass(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                  hipStream_t st);
// The same over nrhs right-hand sides: side b's three components are spectrum components 3 b ..
// 3 b + 2 (at (3 b + c) * Nc); skip[b * skip_stride] set skips side b.
template <typename R>
void launch_ypass_n(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, int nrhs, const int* skip,
                    int skip_stride, hipStream_t st);
// z-axis transform without the Green multiply (standalone FFT hooks).
template <typename R>
void launch_zpass(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                  hipStream_t st);
// Fused z forward transform, symmetric Green multiply, z backward transform.
// G6 = the symmetrised tensor on the half spectrum, AoS: G6[6*i + g], g = xx,xy,xz,yy,yz,zz.
// qpart != null: one partial per workgroup of the Hermitian-weighted quadratic form
// sum_k w(kx) Re(conj(f(k)) . G(k) f(k)) over the forward spectrum f (w = 1 on the
// self-conjugate planes kx = 0 and 2 kx = nx, else 2); divided by N it is <u, G u> over the
// grid for the real field u that was transformed (Parseval).  Returns the workgroup count.
template <typename R>
long launch_zgreen(cx<R>* S, const R* G6, const Geometry3& g, const FFTTables<R>& t, double* qpart,
                   const int* skip, hipStream_t st);
// The same over nrhs right-hand sides laid out as for launch_ypass_n, side b's partials at
// qpart + b * qstride; the tensor is read once per launch when the tuned kernel serves several
// sides (else one launch per side).  Returns the workgroup count of one side.
template <typename R>
long launch_zgreen_batch(cx<R>* S, const R* G6, const Geometry3& g, const FFTTables<R>& t, double* qpart,
                         long qstride, int nrhs, const int* skip, int skip_stride, hipStream_t st);
// Events that the next fused z-pass launch of the calling thread records at the kernel's own start
// and end (armed by a caller that times it; consumed by that launch).
struct KernelEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
KernelEvents& zpass_events();
// Right-hand sides of one fused z-pass launch: side b's spectrum starts at b * rhs_stride, its
// quadratic-form partials at b * qstride, and skip[b * skip_stride] set skips it.
struct ZBatch {
    int nrhs;
    int skip_stride;
    long rhs_stride;
    long qstride;
};
constexpr ZBatch kOneRhs{1, 0, 0, 0};
// The compile-time specialised passes of the mixed-radix lengths (80, 96, 112, 144, 160), built in
// their own translation units; each returns -1 when the variant does not exist at that length.
bool mixed_length(int n);
template <typename R>
long static_zgreen_mixed(int code, cx<R>* S, const R* G6, long ncols, const Geometry3& g, const cx<R>* tw,
                         double* qpart, const int* skip, hipStream_t st, ZBatch zb);
int zgreen_default_code_mixed(int n, int prec);
template <typename R>
long static_xfwd_mixed(int v, const XSource& src, cx<R>* S, const Geometry3& g, const cx<R>* tw, const int* skip,
                       hipStream_t st);
template <typename R>
long static_xinv_mixed(int v, const cx<R>* S, const XSink& sink, const Geometry3& g, const cx<R>* tw, const int* skip,
                       hipStream_t st);
int pass_variant_code_mixed(int n, int v);
// FP32 x and y transforms fused over whole z planes (nx = ny = 256; passes_plane.hip): the
// forward reads the compact FP32 source over the line bitmask and writes the y-transformed half
// spectrum, the inverse reads it and writes the compact FP32 sink and one partial row per z plane
// (component c's support sum at partials[3 z + c]).  Each returns its workgroup count (the inverse:
// the partial rows), or -1 when the shape, source or sink is not one it serves.
bool plane_pass_shape_ok(const Geometry3& g);
bool plane_src_ok(const XSource& src);
bool plane_sink_ok(const XSink& sink);
// waves: 8 or 16 waves per workgroup (one workgroup per CU either way).
long launch_xyfwd_plane(const XSource& src, cx<float>* S, const Geometry3& g, const cx<float>* twx, int waves,
                        const int* skip, hipStream_t st);
long launch_yxinv_plane(const cx<float>* S, const XSink& sink, const Geometry3& g, const cx<float>* twx, int waves,
                        const int* skip, hipStream_t st);
// HBM copy probes over the three padded FP32 half spectra (probe.hip): pattern 0 contiguous, 1 the
// y passes' 32-element column tiles, 2 the z pass's 8-element column tiles; src -> dst, both 3 Nc
// complex.  probe_bytes: bytes one launch reads plus writes (-1: unknown pattern).
long probe_bytes(int pattern, const Geometry3& g);
void launch_probe(int pattern, const float2* src, float2* dst, const Geometry3& g, hipStream_t st);
// Upper bounds of the workgroup counts of launch_xinv / launch_zgreen over every variant.
long xinv_max_blocks(const Geometry3& g);
long zgreen_max_blocks(const Geometry3& g);

// Specialised passes (x-line / column lengths 32 ... 512) in kPassVariants tile shapes; each
// returns the workgroup count, or -1 when the size or variant has no kernel on this device.
bool static_passes_ok();
template <typename R>
long static_xfwd(int v, const XSource& src, cx<R>* S, const Geometry3& g, const cx<R>* tw, const int* skip,
                 hipStream_t st);
template <typename R>
long static_xinv(int v, const cx<R>* S, const XSink& sink, const Geometry3& g, const cx<R>* tw, const int* skip,
                 hipStream_t st);
// ncomp components (grid rows); skip[(component / 3) * skip_stride] set skips a component's side.
template <typename R>
long static_col(int v, int n, int sign, cx<R>* S, long ncols, long inner, long outer, long estride, long comp_stride,
                const cx<R>* tw, const int* skip, hipStream_t st, int ncomp = 3, int skip_stride = 0);
// 10 * lines per workgroup + threads per butterfly thread of variant v at length n (0 if none).
int pass_variant_code(int n, int v);  // x passes (kXPassVariants)
int col_variant_code(int n, int v);   // column passes (kColVariants)

// Autotuner registry, keyed by (pass kind, precision 0 = FP64 / 1 = FP32, transform length,
// line or column count).  Values: pass variant index for the x and column passes, variant code
// (10 * columns per workgroup + threads per butterfly thread) for the fused z pass; -1 untuned.
// kKindPlane (keyed by nx and ny * nz): the FP32 x and y passes fused over z planes with 8 (1) or
// 16 (2) waves per workgroup; 0 or untuned = separate passes.
enum PassKind { kKindXfwd = 0, kKindXinv = 1, kKindCol = 2, kKindZgreen = 3, kKindPlane = 4 };
int tuned_variant(int kind, int prec, int n, long count);
void set_tuned_variant(int kind, int prec, int n, long count, int v);
void clear_tuned_variant(int kind, int prec, int n, long count);
// Registry values of the variants the autotuner could launch for this key (empty: not tuned in
// this process, or no specialised kernel).
std::vector<int> tuned_candidates(int kind, int prec, int n, long count);
// Incremented by every change of the registry (captured launch sequences depend on it).
long tuning_generation();
// Registry value of a report code (pass variant index or z-pass code), -1 when unknown.
int variant_of_code(int kind, int n, int code);
// Report code of the variant a launch of this kind and size uses now (0 = runtime kernel).
int variant_code(int kind, int prec, int n, long count);

// Each autotuner times every variant that fits the device on the given buffers (3 launches
// after a warm-up), records the fastest and returns (code, ms) pairs; empty when the size has
// no specialised kernel or was tuned before in this process.  They overwrite S (and the sink
// outputs of the x-inverse).
template <typename R>
std::vector<std::pair<int, float>> autotune_zgreen(cx<R>* S, const R* G6, const Geometry3& g,
                                                   const FFTTables<R>& t, hipStream_t st);
template <typename R>
std::vector<std::pair<int, float>> autotune_col(cx<R>* S, const Geometry3& g, const FFTTables<R>& t,
                                                hipStream_t st);
template <typename R>
std::vector<std::pair<int, float>> autotune_x(const XSource& src, const XSink& sink, cx<R>* S, const Geometry3& g,
                                              const FFTTables<R>& t, hipStream_t st);

// ------------------------------------------------------------ setup kernels
// Support classification + ascending compaction.  Fills code (N) and active (NB);
// nb_out / nf_out are device counters (uint64).  scratch must hold scan_scratch_bytes(N).
size_t scan_scratch_bytes(long N);
// Line bitmask + line starts from the support map (bits: nlines * ceil(nx / 32) words; lstart:
// nlines ints, scanned in place; total: device uint64 receiving the support count).
void launch_line_bits(const int32_t* code, int nx, long nlines, uint32_t* bits, int32_t* lstart,
                      unsigned long long* total, hipStream_t st);
void launch_support(const uint8_t* mask, int nx, int ny, int nz, int interface_only, int32_t* code,
                    int64_t* active, unsigned long long* counts /*[2]: nb, nf*/, void* scratch,
                    hipStream_t st);
// The same classification writing the reference's voxel_to_local (int64, -1 off the support; may be
// null) and at most `capacity` entries of the active list.  counts[2] (nb, nf) are cleared by the
// launch itself (a kernel, no memset); bsum holds support_scan_blocks(N) ints.
long support_scan_blocks(long N);
// Sets the calling thread's fftperm_last_error() message (api.hip).
void set_last_error(const std::string& msg);
void launch_support_v2l(const uint8_t* mask, int nx, int ny, int nz, int interface_only, int64_t* voxel_to_local,
                        int64_t* active, long capacity, unsigned long long* counts, int* bsum, hipStream_t st);

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
// Half-spectrum symmetrised tensor 0.5 (G(k) + G(-k)), AoS G6[6*e + g] at the padded spectrum index
// e = kz * pz + ky * rp + kx, in double (and float if G6f != null).  The padding is left untouched.
void launch_green_sym(const GreenAxisTables& t, int nx, int ny, int nz, int rp, long pz, double* G6, float* G6f,
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

// Right-hand sides a multi-RHS solve carries at most (the three load directions of a tensor).
constexpr int kMaxRhs = 3;

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

// kernels_fft.hip -- the five sweeps of the Green-operator apply on the R2C half spectrum.
//
// Layout in HBM: real fields are SoA by component, flat index (z*ny + y)*nx + x
// (src/core/Grid3D.hpp:41-43).  The spectrum S holds three components of the
// nz*ny*nxh frequencies (nxh = nx/2 + 1) in the padded layout of Geometry3:
// component c, frequency (kz, ky, kx) at c*Nc + kz*pz + ky*rp + kx.
//
//   x-forward : real x-lines -> complex lines -> n-point FFT -> first nxh bins
//               (source values built on the fly from the support map: fused scatter)
//   y-pass    : strided columns along y, one component per blockIdx.y
//   z-green   : all three components of a tile of z-columns in LDS: forward FFT,
//               symmetric 3x3 Green multiply, backward FFT (one HBM round trip)
//   x-inverse : Hermitian completion of each x-line, backward FFT, real part / N
//               (sink: fused gather to the compact vector and/or field write,
//               per-block partial sums for the support/domain means)
#include "zgreen_static.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace fftperm {

KernelEvents& zpass_events() {
    thread_local KernelEvents ev;
    return ev;
}


namespace {

size_t g_lds_max = 64 * 1024;  // per-workgroup LDS limit of the device (set at context creation)
int g_cus = 256;               // compute units of the device (set at context creation)
size_t g_lds_cu = 160 * 1024;  // LDS per compute unit

size_t lds_single() { return std::min<size_t>(72 * 1024, g_lds_max); }
size_t lds_triple() { return std::min<size_t>(104 * 1024, g_lds_max); }

int floor_pow2(long v) {
    int p = 1;
    while ((long)p * 2 <= v) p *= 2;
    return p;
}

PassCfg make_cfg(const LinePlan& p, long nunits, int ncomp, size_t budget, size_t csize) {
    const int n = p.n;
    const int maxbf = std::max(1, p.maxbf);
    long lines = 1024 / (ncomp * maxbf);
    const size_t per_line = (size_t)(n + 1) * csize * ncomp;
    const size_t tw = (size_t)n * csize;
    long by_lds = budget > tw ? (long)((budget - tw) / per_line) : 1;
    lines = std::min(lines, std::max(1L, by_lds));
    lines = floor_pow2(std::max(1L, lines));
    if (lines > nunits) lines = std::max(1L, nunits);
    PassCfg c;
    c.lines = (int)lines;
    int thr = ncomp * (int)lines * maxbf;
    thr = ((thr + 63) / 64) * 64;
    thr = std::max(thr, 256);
    thr = std::min(thr, 1024);
    c.threads = thr;
    c.lds = tw + per_line * (size_t)lines;
    c.blocks = (nunits + lines - 1) / lines;
    return c;
}

}  // namespace

PassCfg cfg_xpass(const LinePlan& p, long nlines) { return make_cfg(p, nlines, 1, lds_single(), 16); }
PassCfg cfg_colpass(const LinePlan& p, long ncols) { return make_cfg(p, ncols, 1, lds_single(), 16); }
PassCfg cfg_zgreen(const LinePlan& p, long ncols) { return make_cfg(p, ncols, 3, lds_triple(), 16); }

void set_lds_limit(size_t bytes) { g_lds_max = bytes; }
void set_device_shape(int cus, size_t lds_per_cu) {
    if (cus > 0) g_cus = cus;
    if (lds_per_cu > 0) g_lds_cu = lds_per_cu;
}
size_t lds_limit() { return g_lds_max; }
int device_cus() { return g_cus; }
size_t lds_per_cu() { return g_lds_cu; }


namespace {

template <typename R>
long static_zgreen_code(int code, cx<R>* S, const R* G6, long ncols, const Geometry3& g, const cx<R>* tw,
                        double* qpart, const int* skip, hipStream_t st, ZBatch zb = kOneRhs) {
    switch (g.nz) {
        case 32: return launch_zgreen_n<R, 32>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 64: return launch_zgreen_n<R, 64>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 128: return launch_zgreen_n<R, 128>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 256: return launch_zgreen_n<R, 256>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 512: return launch_zgreen_n<R, 512>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        default: return mixed_length(g.nz) ? static_zgreen_mixed<R>(code, S, G6, ncols, g, tw, qpart, skip, st, zb)
                                           : -1;
    }
}

// untuned: the widest tile that fits ~100 KB of LDS, one thread per butterfly
int zgreen_default_code(int n, int prec) {
    switch (n) {
        case 32: return 10 * (prec ? ZShape<float, 32>::ZL : ZShape<double, 32>::ZL) + 1;
        case 64: return 10 * (prec ? ZShape<float, 64>::ZL : ZShape<double, 64>::ZL) + 1;
        case 128: return 10 * (prec ? ZShape<float, 128>::ZL : ZShape<double, 128>::ZL) + 1;
        case 256: return 10 * (prec ? ZShape<float, 256>::ZL : ZShape<double, 256>::ZL) + 1;
        case 512: return 10 * (prec ? ZShape<float, 512>::ZL : ZShape<double, 512>::ZL) + 1;
        default: return mixed_length(n) ? zgreen_default_code_mixed(n, prec) : 0;
    }
}

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, long>, int> g_tuned;
std::map<std::tuple<int, int, int, long>, std::vector<int>> g_live;  // variants that launched
long g_tune_gen = 0;

template <typename R>
constexpr int prec_of() {
    return sizeof(R) == 8 ? 0 : 1;
}

// Times `launch` (which returns < 0 when the variant does not exist) for every candidate and
// records the fastest under (kind, prec, n, count).  Every candidate runs once first (existence
// check and warm-up, so that the clocks are up before any timing); then kTuneRounds rounds visit
// the candidates in turn, each timing kTuneReps launches, and a candidate's time is its best
// round -- interleaving spreads clock and cache drift over all candidates instead of charging it
// to whichever is timed first.
constexpr int kTuneRounds = 3;
constexpr int kTuneReps = 3;

template <typename F>
std::vector<std::pair<int, float>> tune_candidates(int kind, int prec, int n, long count, const std::vector<int>& cands,
                                                   F&& launch, int (*code_of)(int, int), hipStream_t st) {
    std::vector<std::pair<int, float>> timings;
    if (!static_passes_ok() || tuned_variant(kind, prec, n, count) >= 0) return timings;
    std::vector<int> live;
    for (int v : cands)
        if (launch(v) >= 0) live.push_back(v);
    if (live.empty()) return timings;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return timings;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return timings;
    }
    std::vector<float> best_of(live.size(), 1e30f);
    bool ok = true;
    for (int round = 0; round < kTuneRounds && ok; ++round) {
        for (size_t i = 0; i < live.size(); ++i) {
            (void)hipEventRecord(a, st);
            for (int rep = 0; rep < kTuneReps; ++rep) launch(live[i]);
            (void)hipEventRecord(b, st);
            if (hipEventSynchronize(b) != hipSuccess) {
                ok = false;
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            best_of[i] = std::min(best_of[i], ms / (float)kTuneReps);
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!ok) return timings;
    size_t best = 0;
    for (size_t i = 0; i < live.size(); ++i) {
        timings.emplace_back(code_of(n, live[i]), best_of[i]);
        if (best_of[i] < best_of[best]) best = i;
    }
    set_tuned_variant(kind, prec, n, count, live[best]);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_live[std::make_tuple(kind, prec, n, count)] = live;
    }
    return timings;
}

int zgreen_code_identity(int, int v) { return v; }

std::vector<int> pass_candidates(int count = kPassVariants) {
    std::vector<int> v;
    for (int i = 0; i < count; ++i) v.push_back(i);
    return v;
}

int pass_variant_or_default(int kind, int prec, int n, long count) {
    const int v = tuned_variant(kind, prec, n, count);
    return v < 0 ? 0 : v;
}

}  // namespace

// The specialised kernels need their static LDS budget; the runtime kernels cover devices with
// less LDS per workgroup and every other length.
bool static_passes_ok() { return g_lds_max >= 102400u + 8192u; }

int tuned_variant(int kind, int prec, int n, long count) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(std::make_tuple(kind, prec, n, count));
    return it == g_tuned.end() ? -1 : it->second;
}

void set_tuned_variant(int kind, int prec, int n, long count, int v) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[std::make_tuple(kind, prec, n, count)] = v;
    ++g_tune_gen;
}

std::vector<int> tuned_candidates(int kind, int prec, int n, long count) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_live.find(std::make_tuple(kind, prec, n, count));
    return it == g_live.end() ? std::vector<int>() : it->second;
}

long tuning_generation() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return g_tune_gen;
}

void clear_tuned_variant(int kind, int prec, int n, long count) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.erase(std::make_tuple(kind, prec, n, count));
    ++g_tune_gen;
}

int variant_of_code(int kind, int n, int code) {
    if (kind == kKindZgreen) {
        for (int c : kZgreenCodes)
            if (c == code) return c;
        return -1;
    }
    if (kind == kKindCol) {
        for (int v = 0; v < kColVariants; ++v)
            if (col_variant_code(n, v) == code && code != 0) return v;
        return -1;
    }
    for (int v = 0; v < kXPassVariants; ++v)
        if (pass_variant_code(n, v) == code && code != 0) return v;
    return -1;
}

int variant_code(int kind, int prec, int n, long count) {
    if (!static_passes_ok()) return 0;
    if (kind == kKindZgreen) {
        const int c = tuned_variant(kind, prec, n, count);
        return c < 0 ? zgreen_default_code(n, prec) : c;
    }
    const int v = pass_variant_or_default(kind, prec, n, count);
    return kind == kKindCol ? col_variant_code(n, v) : pass_variant_code(n, v);
}

template <typename R>
std::vector<std::pair<int, float>> autotune_zgreen(cx<R>* S, const R* G6, const Geometry3& g, const FFTTables<R>& t,
                                                   hipStream_t st) {
    const long key = (long)g.ny * g.nxh;   // registry count: the frequency columns
    const long ncols = (long)g.ny * g.rp;  // columns a launch covers: every padded row
    (void)hipMemsetAsync(S, 0, 3 * (size_t)g.Nc * sizeof(cx<R>), st);
    std::vector<int> cands(std::begin(kZgreenCodes), std::end(kZgreenCodes));
    return tune_candidates(
        kKindZgreen, prec_of<R>(), g.nz, key, cands,
        [&](int code) { return static_zgreen_code<R>(code, S, G6, ncols, g, t.twz, nullptr, nullptr, st); },
        zgreen_code_identity, st);
}

template <typename R>
std::vector<std::pair<int, float>> autotune_col(cx<R>* S, const Geometry3& g, const FFTTables<R>& t, hipStream_t st) {
    const long key = (long)g.nz * g.nxh;   // registry count: the frequency columns
    const long ncols = (long)g.nz * g.rp;  // columns a launch covers: every padded row
    (void)hipMemsetAsync(S, 0, 3 * (size_t)g.Nc * sizeof(cx<R>), st);
    return tune_candidates(
        kKindCol, prec_of<R>(), g.ny, key, pass_candidates(kColVariants),
        [&](int v) {
            return static_col<R>(v, g.ny, -1, S, ncols, (long)g.rp, g.pz, (long)g.rp, g.Nc, t.twy, nullptr, st);
        },
        col_variant_code, st);
}

template <typename R>
std::vector<std::pair<int, float>> autotune_x(const XSource& src, const XSink& sink, cx<R>* S, const Geometry3& g,
                                              const FFTTables<R>& t, hipStream_t st) {
    const long nlines = (long)g.ny * g.nz;
    auto tf = tune_candidates(
        kKindXfwd, prec_of<R>(), g.nx, nlines, pass_candidates(kXPassVariants),
        [&](int v) { return static_xfwd<R>(v, src, S, g, t.twx, nullptr, st); }, pass_variant_code, st);
    auto ti = tune_candidates(
        kKindXinv, prec_of<R>(), g.nx, nlines, pass_candidates(kXPassVariants),
        [&](int v) { return static_xinv<R>(v, S, sink, g, t.twx, nullptr, st); }, pass_variant_code, st);
    tf.insert(tf.end(), ti.begin(), ti.end());
    return tf;
}

template std::vector<std::pair<int, float>> autotune_zgreen<double>(cx<double>*, const double*, const Geometry3&,
                                                                    const FFTTables<double>&, hipStream_t);
template std::vector<std::pair<int, float>> autotune_zgreen<float>(cx<float>*, const float*, const Geometry3&,
                                                                   const FFTTables<float>&, hipStream_t);
template std::vector<std::pair<int, float>> autotune_col<double>(cx<double>*, const Geometry3&,
                                                                 const FFTTables<double>&, hipStream_t);
template std::vector<std::pair<int, float>> autotune_col<float>(cx<float>*, const Geometry3&, const FFTTables<float>&,
                                                                hipStream_t);
template std::vector<std::pair<int, float>> autotune_x<double>(const XSource&, const XSink&, cx<double>*,
                                                               const Geometry3&, const FFTTables<double>&,
                                                               hipStream_t);
template std::vector<std::pair<int, float>> autotune_x<float>(const XSource&, const XSink&, cx<float>*,
                                                              const Geometry3&, const FFTTables<float>&, hipStream_t);

long xinv_max_blocks(const Geometry3& g) {
    // every variant and the runtime kernel put at least one line in a workgroup
    return (long)g.ny * g.nz;
}

long zgreen_max_blocks(const Geometry3& g) { return (long)g.ny * g.rp; }

// ------------------------------------------------------------------ x forward
template <typename R>
__global__ void k_xfwd(XSource src, cx<R>* __restrict__ S, int nx, long nlines, int nxh, long N, SpecLayout SL,
                       LinePlan pl, const cx<R>* __restrict__ twg, int lpw, const int* skip) {
    if (skip != nullptr && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + nx;
    const int ld = nx + 1;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) tw[i] = twg[i];
    const long line0 = (long)blockIdx.x * lpw;
    const int nl = (int)min((long)lpw, nlines - line0);
    const int cnt = nl * nx;
    for (int c = 0; c < 3; ++c) {
        const double af = src.a_fluid[c];
        const double as = src.a_support[c];
        const double mc = (src.mean != nullptr) ? src.mean[c] : 0.0;
        for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
            const int line = idx / nx;
            const int i = idx - line * nx;
            const long vox = (line0 + line) * nx + i;
            double val;
            if (src.field != nullptr) {
                val = src.field[c * N + vox];
            } else {
                const int code = src.code[vox];
                const double f0 = (code == -1) ? af : ((code >= 0) ? as : 0.0);
                double fx = 0.0;
                if (src.compact != nullptr && code >= 0) fx = src.compact[3L * code + c] - mc;
                if (src.compact32 != nullptr && code >= 0) fx = (double)src.compact32[3L * code + c] - mc;
                val = f0 + fx;
            }
            buf[line * ld + i] = mkc<R>((R)val, (R)0);
        }
        __syncthreads();
        fft_lines<R, -1>(buf, ld, nl, pl, tw);
        const int cnth = nl * nxh;
        cx<R>* Sc = S + c * SL.nc;
        for (int idx = threadIdx.x; idx < cnth; idx += blockDim.x) {
            const int line = idx / nxh;
            const int kx = idx - line * nxh;
            Sc[spec_line(SL, line0 + line) + kx] = buf[line * ld + kx];
        }
        __syncthreads();
    }
}

template <typename R>
void launch_xfwd(const XSource& src, cx<R>* S, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                 hipStream_t st) {
    const long nlines = (long)g.ny * g.nz;
    if (static_passes_ok() &&
        static_xfwd<R>(pass_variant_or_default(kKindXfwd, prec_of<R>(), g.nx, nlines), src, S, g, t.twx, skip, st) >= 0)
        return;
    PassCfg c = cfg_xpass(t.px, nlines);
    hipLaunchKernelGGL(k_xfwd<R>, dim3((unsigned)c.blocks), dim3(c.threads), c.lds, st, src, S, g.nx, nlines,
                       g.nxh, g.N, spec_layout(g), t.px, t.twx, c.lines, skip);
}

// ------------------------------------------------------------------ x inverse
template <typename R>
__global__ void k_xinv(const cx<R>* __restrict__ S, XSink sink, int nx, long nlines, int nxh, long N, SpecLayout SL,
                       LinePlan pl, const cx<R>* __restrict__ twg, int lpw, double inv_n, const int* skip) {
    if (skip != nullptr && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + nx;
    const int ld = nx + 1;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) tw[i] = twg[i];
    const long line0 = (long)blockIdx.x * lpw;
    const int nl = (int)min((long)lpw, nlines - line0);
    double acc[kXinvParts] = {0, 0, 0};
    for (int c = 0; c < 3; ++c) {
        const cx<R>* Sc = S + c * SL.nc;
        const int cnth = nl * nxh;
        for (int idx = threadIdx.x; idx < cnth; idx += blockDim.x) {
            const int line = idx / nxh;
            const int kx = idx - line * nxh;
            buf[line * ld + kx] = Sc[spec_line(SL, line0 + line) + kx];
        }
        __syncthreads();
        const int nm = nx - nxh;
        for (int idx = threadIdx.x; idx < nl * nm; idx += blockDim.x) {
            const int line = idx / nm;
            const int kx = nxh + (idx - line * nm);
            buf[line * ld + kx] = conjc<R>(buf[line * ld + (nx - kx)]);
        }
        __syncthreads();
        fft_lines<R, +1>(buf, ld, nl, pl, tw);
        const int cnt = nl * nx;
        for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
            const int line = idx / nx;
            const int i = idx - line * nx;
            const long vox = (line0 + line) * nx + i;
            const double val = (double)buf[line * ld + i].x * inv_n;
            if (sink.field != nullptr) sink.field[c * N + vox] = val;
            const int code = sink.code[vox];
            if (code >= 0) {
                if (sink.compact != nullptr) sink.compact[3L * code + c] = val;
                if (sink.compact32 != nullptr) sink.compact32[3L * code + c] = (float)val;
                if (c == 0) acc[0] += val;
                if (c == 1) acc[1] += val;
                if (c == 2) acc[2] += val;
            }
        }
        __syncthreads();
    }
    if (sink.partials != nullptr) {
        block_sum<kXinvParts>(acc, sink.partials + (long)kXinvParts * blockIdx.x);
    }
}

template <typename R>
long launch_xinv(const cx<R>* S, const XSink& sink, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                 hipStream_t st) {
    const long nlines = (long)g.ny * g.nz;
    if (static_passes_ok()) {
        const long b =
            static_xinv<R>(pass_variant_or_default(kKindXinv, prec_of<R>(), g.nx, nlines), S, sink, g, t.twx, skip, st);
        if (b >= 0) return b;
    }
    PassCfg c = cfg_xpass(t.px, nlines);
    const double inv_n = 1.0 / (double)g.N;
    hipLaunchKernelGGL(k_xinv<R>, dim3((unsigned)c.blocks), dim3(c.threads), c.lds, st, S, sink, g.nx, nlines,
                       g.nxh, g.N, spec_layout(g), t.px, t.twx, c.lines, inv_n, skip);
    return c.blocks;
}

// ------------------------------------------------------------- column passes
// Column col of a component: base = (col / inner) * outer_stride + col % inner,
// element e at base + e * estride.
template <typename R, int SIGN>
__global__ void k_colpass(cx<R>* __restrict__ S, int n, long ncols, long inner, long outer_stride, long estride,
                          long comp_stride, LinePlan pl, const cx<R>* __restrict__ twg, int tl,
                          const int* skip, int skip_stride) {
    if (skip != nullptr && skip[(blockIdx.y / 3) * skip_stride]) return;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + n;
    const int ld = n + 1;
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
    const long col0 = (long)blockIdx.x * tl;
    const int ncl = (int)min((long)tl, ncols - col0);
    cx<R>* Sc = S + blockIdx.y * comp_stride;
    const int cnt = ncl * n;
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
        const int e = idx / ncl;
        const int ci = idx - e * ncl;
        const long col = col0 + ci;
        const long base = (col / inner) * outer_stride + (col % inner);
        buf[ci * ld + e] = Sc[base + e * estride];
    }
    __syncthreads();
    fft_lines<R, SIGN>(buf, ld, ncl, pl, tw);
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
        const int e = idx / ncl;
        const int ci = idx - e * ncl;
        const long col = col0 + ci;
        const long base = (col / inner) * outer_stride + (col % inner);
        Sc[base + e * estride] = buf[ci * ld + e];
    }
}

template <typename R>
void launch_ypass(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                  hipStream_t st) {
    launch_ypass_n<R>(S, sign, g, t, 1, skip, 0, st);
}

template <typename R>
void launch_ypass_n(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, int nrhs, const int* skip,
                    int skip_stride, hipStream_t st) {
    // column (kz, kx) of every padded row, kx < rp: base kz * pz + kx, elements ky * rp.  Covering
    // the row padding (zeros in, zeros out) keeps every tile of 8 columns on a 64-byte boundary.
    // Component c of right-hand side b is spectrum component 3 b + c.
    const long key = (long)g.nz * g.nxh;
    const long ncols = (long)g.nz * g.rp;
    const int ncomp = 3 * nrhs;
    if (static_passes_ok() &&
        static_col<R>(pass_variant_or_default(kKindCol, prec_of<R>(), g.ny, key), g.ny, sign, S, ncols, (long)g.rp,
                      g.pz, (long)g.rp, g.Nc, t.twy, skip, st, ncomp, skip_stride) >= 0)
        return;
    PassCfg c = cfg_colpass(t.py, ncols);
    dim3 grid((unsigned)c.blocks, (unsigned)ncomp);
    if (sign < 0)
        hipLaunchKernelGGL((k_colpass<R, -1>), grid, dim3(c.threads), c.lds, st, S, g.ny, ncols, (long)g.rp, g.pz,
                           (long)g.rp, g.Nc, t.py, t.twy, c.lines, skip, skip_stride);
    else
        hipLaunchKernelGGL((k_colpass<R, +1>), grid, dim3(c.threads), c.lds, st, S, g.ny, ncols, (long)g.rp, g.pz,
                           (long)g.rp, g.Nc, t.py, t.twy, c.lines, skip, skip_stride);
}

template <typename R>
void launch_zpass(cx<R>* S, int sign, const Geometry3& g, const FFTTables<R>& t, const int* skip,
                  hipStream_t st) {
    // column (ky, kx) of every padded row: base ky * rp + kx, elements kz * pz (padding stays zero)
    const long ncols = (long)g.ny * g.rp;
    if (static_passes_ok() && static_col<R>(0, g.nz, sign, S, ncols, ncols, 0L, g.pz, g.Nc, t.twz, skip, st) >= 0)
        return;
    PassCfg c = cfg_colpass(t.pz, ncols);
    dim3 grid((unsigned)c.blocks, 3);
    if (sign < 0)
        hipLaunchKernelGGL((k_colpass<R, -1>), grid, dim3(c.threads), c.lds, st, S, g.nz, ncols, ncols, 0L, g.pz,
                           g.Nc, t.pz, t.twz, c.lines, skip, 0);
    else
        hipLaunchKernelGGL((k_colpass<R, +1>), grid, dim3(c.threads), c.lds, st, S, g.nz, ncols, ncols, 0L, g.pz,
                           g.Nc, t.pz, t.twz, c.lines, skip, 0);
}

// ------------------------------------------------------------- fused z + Green
template <typename R>
__global__ void k_zgreen(cx<R>* __restrict__ S, const R* __restrict__ G6, int nz, long ncols, long pitch, long Nc,
                         int nx, int rp, LinePlan pl, const cx<R>* __restrict__ twg, int tl,
                         double* __restrict__ qpart, const int* skip) {
    if (skip != nullptr && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    cx<R>* tw = reinterpret_cast<cx<R>*>(smem_raw);
    cx<R>* buf = tw + nz;
    const int ld = nz + 1;
    for (int i = threadIdx.x; i < nz; i += blockDim.x) tw[i] = twg[i];
    const long col0 = (long)blockIdx.x * tl;
    const int ncl = (int)min((long)tl, ncols - col0);
    const int cnt = ncl * nz;
    for (int c = 0; c < 3; ++c) {
        const cx<R>* Sc = S + c * Nc;
        cx<R>* B = buf + c * ncl * ld;
        for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
            const int e = idx / ncl;
            const int ci = idx - e * ncl;
            B[ci * ld + e] = Sc[col0 + ci + e * pitch];
        }
    }
    __syncthreads();
    fft_lines<R, -1>(buf, ld, 3 * ncl, pl, tw);
    // Symmetric Green multiply, as GreenOperator::apply_spectral_green_tensor
    // (src/bignonnet/GreenOperator.cpp:84-92) with the symmetrised tensor, plus the weighted
    // quadratic form conj(f) . G f when asked for.
    double q[1] = {0.0};
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
        const int e = idx / ncl;
        const int ci = idx - e * ncl;
        const R* gp = G6 + 6L * (col0 + ci + e * pitch);
        const R txx = gp[0], txy = gp[1], txz = gp[2];
        const R tyy = gp[3], tyz = gp[4], tzz = gp[5];
        cx<R>* px = buf + (0 * ncl + ci) * ld + e;
        cx<R>* py = buf + (1 * ncl + ci) * ld + e;
        cx<R>* pz = buf + (2 * ncl + ci) * ld + e;
        const cx<R> fx = *px, fy = *py, fz = *pz;
        cx<R> gx, gy, gz;
        green_mul<R>(txx, txy, txz, tyy, tyz, tzz, fx, fy, fz, gx, gy, gz);
        *px = gx;
        *py = gy;
        *pz = gz;
        if (qpart != nullptr)
            green_quad_add<R>(q[0], herm_weight((int)((col0 + ci) % rp), nx), fx, fy, fz, gx, gy, gz);
    }
    if (qpart != nullptr) block_sum<1>(q, qpart + blockIdx.x);
    __syncthreads();
    fft_lines<R, +1>(buf, ld, 3 * ncl, pl, tw);
    for (int c = 0; c < 3; ++c) {
        cx<R>* Sc = S + c * Nc;
        const cx<R>* B = buf + c * ncl * ld;
        for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
            const int e = idx / ncl;
            const int ci = idx - e * ncl;
            Sc[col0 + ci + e * pitch] = B[ci * ld + e];
        }
    }
}

template <typename R>
long launch_zgreen(cx<R>* S, const R* G6, const Geometry3& g, const FFTTables<R>& t, double* qpart, const int* skip,
                   hipStream_t st) {
    const long key = (long)g.ny * g.nxh;   // registry count: the frequency columns
    const long ncols = (long)g.ny * g.rp;  // columns a launch covers: every padded row
    if (static_passes_ok()) {
        int code = tuned_variant(kKindZgreen, prec_of<R>(), g.nz, key);
        if (code < 0) code = zgreen_default_code(g.nz, prec_of<R>());
        const long b = static_zgreen_code<R>(code, S, G6, ncols, g, t.twz, qpart, skip, st);
        if (b >= 0) return b;
    }
    PassCfg c = cfg_zgreen(t.pz, ncols);
    zlaunch(k_zgreen<R>, dim3((unsigned)c.blocks), dim3(c.threads), c.lds, st, S, G6, g.nz, ncols, g.pz,
                       g.Nc, g.nx, g.rp, t.pz, t.twz, c.lines, qpart, skip);
    return c.blocks;
}

template <typename R>
long launch_zgreen_batch(cx<R>* S, const R* G6, const Geometry3& g, const FFTTables<R>& t, double* qpart,
                         long qstride, int nrhs, const int* skip, int skip_stride, hipStream_t st) {
    if (nrhs == 1) return launch_zgreen<R>(S, G6, g, t, qpart, skip, st);
    const long key = (long)g.ny * g.nxh;
    const long ncols = (long)g.ny * g.rp;
    const long rhs_stride = 3 * g.Nc;
    if (static_passes_ok()) {
        int code = tuned_variant(kKindZgreen, prec_of<R>(), g.nz, key);
        if (code < 0) code = zgreen_default_code(g.nz, prec_of<R>());
        const ZBatch zb{nrhs, skip_stride, rhs_stride, qstride};
        const long b = static_zgreen_code<R>(code, S, G6, ncols, g, t.twz, qpart, skip, st, zb);
        if (b >= 0) return b;
    }
    // the tuned kernel serves one side per launch: one launch per side
    long blocks = 0;
    for (int b = 0; b < nrhs; ++b)
        blocks = launch_zgreen<R>(S + b * rhs_stride, G6, g, t, qpart != nullptr ? qpart + b * qstride : nullptr,
                                  skip != nullptr ? skip + b * skip_stride : nullptr, st);
    return blocks;
}

// ------------------------------------------------------- spectrum utilities
template <typename R>
__global__ void k_expand_full(const cx<R>* __restrict__ S, double2* __restrict__ F, int nx, int ny, int nz,
                              int nxh, long N, SpecLayout SL) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int c = blockIdx.y;
    const int kx = (int)(i % nx);
    const long r = i / nx;
    const int ky = (int)(r % ny);
    const int kz = (int)(r / ny);
    double2 v;
    if (kx < nxh) {
        const cx<R> s = S[c * SL.nc + (long)kz * SL.pz + (long)ky * SL.rp + kx];
        v.x = (double)s.x;
        v.y = (double)s.y;
    } else {
        const int mz = (nz - kz) % nz, my = (ny - ky) % ny, mx = nx - kx;
        const cx<R> s = S[c * SL.nc + (long)mz * SL.pz + (long)my * SL.rp + mx];
        v.x = (double)s.x;
        v.y = -(double)s.y;
    }
    F[c * N + i] = v;
}

template <typename R>
void launch_expand_full(const cx<R>* S, double2* F, const Geometry3& g, hipStream_t st) {
    dim3 grid((unsigned)((g.N + 255) / 256), 3);
    hipLaunchKernelGGL(k_expand_full<R>, grid, dim3(256), 0, st, S, F, g.nx, g.ny, g.nz, g.nxh, g.N, spec_layout(g));
}

template <typename R>
__global__ void k_hermitian_half(const double2* __restrict__ X, cx<R>* __restrict__ S, int nx, int ny, int nz,
                                 int nxh, long N, long Nh, SpecLayout SL) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nh) return;
    const int c = blockIdx.y;
    const int kx = (int)(i % nxh);
    const long r = i / nxh;
    const int ky = (int)(r % ny);
    const int kz = (int)(r / ny);
    const double2 a = X[c * N + ((long)kz * ny + ky) * nx + kx];
    const int mz = (nz - kz) % nz, my = (ny - ky) % ny, mx = (nx - kx) % nx;
    const double2 b = X[c * N + ((long)mz * ny + my) * nx + mx];
    S[c * SL.nc + (long)kz * SL.pz + (long)ky * SL.rp + kx] = mkc<R>((R)(0.5 * (a.x + b.x)), (R)(0.5 * (a.y - b.y)));
}

template <typename R>
void launch_hermitian_half(const double2* X, cx<R>* S, const Geometry3& g, hipStream_t st) {
    dim3 grid((unsigned)((g.Nh + 255) / 256), 3);
    hipLaunchKernelGGL(k_hermitian_half<R>, grid, dim3(256), 0, st, X, S, g.nx, g.ny, g.nz, g.nxh, g.N, g.Nh,
                       spec_layout(g));
}

__global__ void k_spectral_green_full(const double* __restrict__ T6, const double2* __restrict__ F,
                                      double2* __restrict__ V, long N) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double txx = T6[i], txy = T6[N + i], txz = T6[2 * N + i];
    const double tyy = T6[3 * N + i], tyz = T6[4 * N + i], tzz = T6[5 * N + i];
    const double2 fx = F[i], fy = F[N + i], fz = F[2 * N + i];
    double2 v;
    v.x = txx * fx.x + txy * fy.x + txz * fz.x;
    v.y = txx * fx.y + txy * fy.y + txz * fz.y;
    V[i] = v;
    v.x = txy * fx.x + tyy * fy.x + tyz * fz.x;
    v.y = txy * fx.y + tyy * fy.y + tyz * fz.y;
    V[N + i] = v;
    v.x = txz * fx.x + tyz * fy.x + tzz * fz.x;
    v.y = txz * fx.y + tyz * fy.y + tzz * fz.y;
    V[2 * N + i] = v;
}

void launch_spectral_green_full(const double* T6, const double2* F, double2* V, long N, hipStream_t st) {
    hipLaunchKernelGGL(k_spectral_green_full, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, T6, F, V, N);
}

#define FFTPERM_INST(R)                                                                                       \
    template void launch_xfwd<R>(const XSource&, cx<R>*, const Geometry3&, const FFTTables<R>&, const int*,   \
                                 hipStream_t);                                                                \
    template long launch_xinv<R>(const cx<R>*, const XSink&, const Geometry3&, const FFTTables<R>&,           \
                                 const int*, hipStream_t);                                                    \
    template void launch_ypass<R>(cx<R>*, int, const Geometry3&, const FFTTables<R>&, const int*,             \
                                  hipStream_t);                                                               \
    template void launch_zpass<R>(cx<R>*, int, const Geometry3&, const FFTTables<R>&, const int*,             \
                                  hipStream_t);                                                               \
    template long launch_zgreen<R>(cx<R>*, const R*, const Geometry3&, const FFTTables<R>&, double*,          \
                                   const int*, hipStream_t);                                                  \
    template long launch_zgreen_batch<R>(cx<R>*, const R*, const Geometry3&, const FFTTables<R>&, double*,    \
                                         long, int, const int*, int, hipStream_t);                            \
    template void launch_ypass_n<R>(cx<R>*, int, const Geometry3&, const FFTTables<R>&, int, const int*, int, \
                                    hipStream_t);                                                             \
    template void launch_expand_full<R>(const cx<R>*, double2*, const Geometry3&, hipStream_t);               \
    template void launch_hermitian_half<R>(const double2*, cx<R>*, const Geometry3&, hipStream_t);

FFTPERM_INST(double)
FFTPERM_INST(float)

}  // namespace fftperm

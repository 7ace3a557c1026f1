// api.hip -- C ABI of libfftperm.so: contexts, planning, and the solver drivers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <set>
#include <vector>

#include "../../include/fftperm.h"
#include "brinkman.hpp"
#include "kernels.hpp"
#include "minres.hpp"
#include "unet.hpp"

using namespace fftperm;

namespace {

thread_local std::string g_last_error;

struct Err : std::runtime_error {
    int code;
    Err(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            throw Err(FFTPERM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));               \
    } while (0)

}  // namespace

// The calling thread's message for fftperm_last_error(), for entry points outside this file.
void fftperm::set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {

template <typename F>
int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return FFTPERM_OK;
    } catch (const Err& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return FFTPERM_ERUNTIME;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return FFTPERM_ERUNTIME;
    }
}

template <typename T>
T* dalloc(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    HIPCHK(hipMalloc(&p, count * sizeof(T)));
    return static_cast<T*>(p);
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree(const_cast<void*>(static_cast<const void*>(p)));
    p = nullptr;
}

// Radix plan for a line of length n: powers of two split into balanced radix-16/8/4/2
// stages, then the odd prime factors 3, 5, 7, 11, 13.
LinePlan make_plan(int n) {
    LinePlan p;
    std::memset(&p, 0, sizeof(p));
    p.n = n;
    int m = n, k = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++k;
    }
    std::vector<int> rad;
    if (k > 0) {
        const int ns = (k + 3) / 4;
        for (int s = 0; s < ns; ++s) {
            const int bits = k / ns + (s < k % ns ? 1 : 0);
            rad.push_back(1 << bits);
        }
    }
    for (int f : {3, 5, 7, 11, 13}) {
        while (m % f == 0) {
            rad.push_back(f);
            m /= f;
        }
    }
    // any other prime factor is a stage of its own, computed as direct DFT sums with one thread
    // per output (FFTW, the reference's backend, plans every length: src/fft/FFTWBackend.cpp:176-192)
    for (int f = 17; m > 1; f += 2) {
        if ((long)f * f > m) f = m;
        while (m % f == 0) {
            rad.push_back(f);
            m /= f;
        }
    }
    if (rad.size() > 8) throw Err(FFTPERM_EUNSUPPORTED, "grid edge " + std::to_string(n) + " needs too many stages");
    p.nst = (int)rad.size();
    p.maxbf = 1;
    for (int s = 0; s < p.nst; ++s) {
        p.rad[s] = rad[s];
        // threads a line needs in this stage: one per butterfly, one per output for a direct-sum prime
        p.maxbf = std::max(p.maxbf, rad[s] > 16 ? n : n / rad[s]);
    }
    if (p.maxbf > 1024)
        throw Err(FFTPERM_EUNSUPPORTED, "grid edge " + std::to_string(n) +
                                            " needs more than 1024 threads per line in one FFT stage");
    return p;
}

// The runtime fused z pass transforms three components of a column in one workgroup: its plan
// must fit three lines' threads into 1024.
void check_z_threads(const LinePlan& pz) {
    if (3 * pz.maxbf > 1024)
        throw Err(FFTPERM_EUNSUPPORTED, "grid edge " + std::to_string(pz.n) +
                                            " along z needs more than 1024 threads for the fused z pass");
}

template <typename R>
cx<R>* make_twiddles(int n) {
    std::vector<cx<R>> h(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < n; ++k) {
        const long double a = 2.0L * pi * (long double)k / (long double)n;
        h[k].x = (R)cosl(a);
        h[k].y = (R)(-sinl(a));
    }
    cx<R>* d = dalloc<cx<R>>(n);
    HIPCHK(hipMemcpy(d, h.data(), n * sizeof(cx<R>), hipMemcpyHostToDevice));
    return d;
}

constexpr double kPiRef = 3.14159265358979323846;

// sinc and q_component of src/bignonnet/GreenOperator.cpp:13-23, evaluated with the same
// double expressions, tabulated per axis for k in [0,n) and p in [-P,P].
void green_axis_table(int n, double period, int P, std::vector<double>& s2, std::vector<double>& q) {
    const int W = 2 * P + 1;
    s2.assign((size_t)n * W, 0.0);
    q.assign((size_t)n * W, 0.0);
    for (int k = 0; k < n; ++k) {
        for (int p = -P; p <= P; ++p) {
            const long long m = (long long)k + (long long)p * (long long)n;
            const double x = kPiRef * (double)m / (double)n;
            const double ax = std::fabs(x);
            const double s = (ax < 1e-12) ? 1.0 : std::sin(x) / x;
            s2[(size_t)k * W + (p + P)] = s * s;
            q[(size_t)k * W + (p + P)] = (2.0 * kPiRef * (double)m) / period;
        }
    }
}

}  // namespace

struct fftperm_ctx {
    Geometry3 g{};
    double h[3]{};
    double L[3]{};
    double mu = 1.0;
    int P = 2;
    int precision = FFTPERM_PRECISION_FP64;
    double mixed_switch = 2.0;      // recurrence residual below which sweep 1 applies the FP32 operator
    double mixed_floor = 3e-6;      // inner recurrence-residual floor of a refinement sweep
    uint32_t mixed_sweeps = 64;     // sweeps per solve at most
    uint64_t fp32_iters = 0;        // iterations with the FP32 operator (whole context)
    int mixed_solver = FFTPERM_MIXED_CR_FP32_VECTORS;
    double reliable_delta = 0.1;    // CG: reliable update when ||r|| falls below this x its last peak
    // MIXED CR32 tensor solves: 0 one direction after the other; 1 the three in lockstep, each
    // direction's five passes run back to back (its spectrum stays in the last-level cache from one
    // pass to the next); 2 the three in lockstep with one launch per pass for all three
    int tensor_batch = 1;
    int device = 0;
    hipStream_t s = nullptr;

    FFTTables<double> t64{};
    FFTTables<float> t32{};
    double2* S64 = nullptr;
    float2* S32 = nullptr;
    double* G64 = nullptr;
    float* G32 = nullptr;
    double* gtab = nullptr;  // per-axis Green tables (6 arrays)
    std::vector<std::pair<int, float>> tune64, tune32;  // autotune timings (variant code, ms) of this context
    GreenAxisTables gt{};

    // medium
    bool has_medium = false;
    uint64_t nb = 0, nf = 0;
    uint8_t* mask = nullptr;
    int32_t* code = nullptr;
    int64_t* active = nullptr;
    uint32_t* lbits = nullptr;           // line bitmask of the support (specialised x passes)
    int32_t* lstart = nullptr;           // support voxels before each x line
    unsigned long long* lcount = nullptr;
    void* scan_scratch = nullptr;
    unsigned long long* counts = nullptr;

    // reductions
    double* part_x = nullptr;  // kXinvParts x x-inverse workgroups
    double* part_z = nullptr;  // one quadratic-form partial per fused z-pass workgroup
    double* part = nullptr;    // 4 kRedBlocks per right-hand side (kMaxRhs sides)
    double* small = nullptr;   // 32 scalars: [0,3) input mean, [4,10) x-inverse sums, [12,15) average, 16 norm
    long xparts = 0;           // partial rows written by the last x-inverse launch
    long zparts = 0;           // partials written by the last fused z-pass launch

    // Krylov vectors (L each)
    size_t Lcap = 0;
    double *x = nullptr, *v_old = nullptr, *v = nullptr, *z = nullptr, *w_prev = nullptr, *w = nullptr;
    double *b = nullptr, *tmp = nullptr, *r = nullptr, *dx = nullptr;
    // FP32 Krylov vectors of the mixed CR solver (allocated on first use), kMaxRhs slots of vs32
    // elements each (slot 0 serves the single-RHS solves)
    size_t L32cap = 0;
    long vs32 = 0;
    float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *ap32 = nullptr, *q32 = nullptr;
    // right-hand sides and FP64 accumulated solutions of a multi-RHS solve (kMaxRhs x L each)
    size_t tcap = 0;
    double *tb = nullptr, *txa = nullptr;
    MinresState* st = nullptr;
    CgState* cst = nullptr;
    CrState* rst = nullptr;
    double* hist = nullptr;
    size_t hist_cap = 0;

    // hipGraphs of two MINRES iterations (two, so that the v / w pointer rotation ends where it
    // started), keyed by everything a capture bakes in: precision, buffer addresses, sizes
    struct GraphKey {
        int fp32;
        int solver;  // 0 MINRES, 1 CG, 2 CR, 3 CR on FP32 vectors, 10 * sides for the multi-RHS CR
        const void* p[9];
        long nb, tuning;
        bool operator==(const GraphKey& o) const {
            if (fp32 != o.fp32 || solver != o.solver || nb != o.nb || tuning != o.tuning) return false;
            for (int i = 0; i < 9; ++i)
                if (p[i] != o.p[i]) return false;
            return true;
        }
    };
    struct GraphSlot {
        GraphKey key;
        hipGraphExec_t exec;
    };
    bool use_graphs = true;
    std::vector<GraphSlot> graphs;

    // live kernel timing (fftperm_profile): HIP events on this context's stream around the
    // fused z-pass and around the whole five-sweep apply
    struct ProfRec {
        hipEvent_t a, b, c, d;  // before x-fwd, before z-green, after z-green, after x-inv
        long long tag;          // Krylov iteration index, or -1 (always a real apply)
        int fp32;
        int batch = 0;          // lockstep record of sides 0 .. batch - 1 (0: not a lockstep record)
        int side = -1;          // per-side record of a multi-RHS iteration (-1: none)
    };
    bool prof = false;
    bool prof_sampled = false;  // fftperm_profile(ctx, 2): one chunk of iterations in kProfSamplePeriod is timed
    long long prof_chunk = 0;
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> ev_free;
    double prof_z_ms[2] = {0, 0}, prof_apply_ms[2] = {0, 0};
    long long prof_z_n[2] = {0, 0}, prof_apply_n[2] = {0, 0};
    long long prof_rhs[2] = {0, 0};  // right-hand sides served by the counted applies
    long long cur_tag = -1;

    hipEvent_t take_event() {
        if (!ev_free.empty()) {
            hipEvent_t e = ev_free.back();
            ev_free.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    // Accumulate recorded applies; Krylov applies launched after convergence (tag >= valid_below)
    // were no-ops and are dropped.  Multi-RHS records: side j ran d[j] iterations of the segment,
    // so iteration t served the sides with t < d[j] (none: a no-op).
    void prof_resolve(long long valid_below, const long long* d = nullptr, int nd = 0) {
        if (prof_pending.empty()) return;
        HIPCHK(hipStreamSynchronize(s));
        for (const ProfRec& r : prof_pending) {
            long long rhs = 1;
            if (r.tag >= 0 && r.batch > 0) {
                rhs = 0;
                for (int j = 0; j < nd; ++j) rhs += (r.tag < d[j]) ? 1 : 0;
            } else if (r.tag >= 0 && r.side >= 0) {
                rhs = (r.side < nd && r.tag < d[r.side]) ? 1 : 0;
            } else if (r.tag >= valid_below) {
                rhs = 0;
            }
            if (rhs > 0) {
                float zms = 0.f, ams = 0.f;
                HIPCHK(hipEventElapsedTime(&zms, r.b, r.c));
                HIPCHK(hipEventElapsedTime(&ams, r.a, r.d));
                prof_z_ms[r.fp32] += zms;
                prof_apply_ms[r.fp32] += ams;
                prof_z_n[r.fp32] += 1;
                prof_apply_n[r.fp32] += 1;
                prof_rhs[r.fp32] += rhs;
            }
            ev_free.push_back(r.a);
            ev_free.push_back(r.b);
            ev_free.push_back(r.c);
            ev_free.push_back(r.d);
        }
        prof_pending.clear();
    }

    // lazily allocated hook buffers
    double* field = nullptr;   // 3N
    double2* full_a = nullptr; // 3N complex
    double* t6 = nullptr;      // 6N

    ~fftperm_ctx() {
        if (s) (void)hipStreamSynchronize(s);
        for (GraphSlot& g : graphs) (void)hipGraphExecDestroy(g.exec);
        dfree(t64.twx); dfree(t64.twy); dfree(t64.twz);
        dfree(t32.twx); dfree(t32.twy); dfree(t32.twz);
        dfree(S64); dfree(S32); dfree(G64); dfree(G32); dfree(gtab);
        dfree(mask); dfree(code); dfree(active); dfree(scan_scratch); dfree(counts);
        dfree(lbits); dfree(lstart); dfree(lcount);
        dfree(part_x); dfree(part_z); dfree(part); dfree(small);
        dfree(x); dfree(v_old); dfree(v); dfree(z); dfree(w_prev); dfree(w);
        dfree(b); dfree(tmp); dfree(r); dfree(dx);
        dfree(x32); dfree(r32); dfree(p32); dfree(ap32); dfree(q32); dfree(tb); dfree(txa);
        dfree(st); dfree(cst); dfree(rst); dfree(hist); dfree(field); dfree(full_a); dfree(t6);
        for (const ProfRec& r : prof_pending) {
            (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b);
            (void)hipEventDestroy(r.c); (void)hipEventDestroy(r.d);
        }
        for (hipEvent_t e : ev_free) (void)hipEventDestroy(e);
        if (s) (void)hipStreamDestroy(s);
    }

    // no copies
    fftperm_ctx() = default;
    fftperm_ctx(const fftperm_ctx&) = delete;
    fftperm_ctx& operator=(const fftperm_ctx&) = delete;

    long L3() const { return 3L * (long)nb; }

    void sync() {
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    }

    void ensure_vectors() {
        const size_t L = (size_t)std::max(1L, L3());
        if (L <= Lcap) return;
        dfree(x); dfree(v_old); dfree(v); dfree(z); dfree(w_prev); dfree(w);
        dfree(b); dfree(tmp); dfree(r); dfree(dx);
        x = dalloc<double>(L); v_old = dalloc<double>(L); v = dalloc<double>(L); z = dalloc<double>(L);
        w_prev = dalloc<double>(L); w = dalloc<double>(L); b = dalloc<double>(L); tmp = dalloc<double>(L);
        r = dalloc<double>(L); dx = dalloc<double>(L);
        Lcap = L;
    }

    void ensure_vectors32() {
        const size_t L = (size_t)std::max(1L, L3());
        if (L <= L32cap) return;
        dfree(x32); dfree(r32); dfree(p32); dfree(ap32); dfree(q32);
        // slots start on 256-byte boundaries (the vector kernels move 16-byte quads)
        const long vs = (long)((L + 63) / 64 * 64);
        const size_t n = (size_t)kMaxRhs * (size_t)vs;
        x32 = dalloc<float>(n); r32 = dalloc<float>(n); p32 = dalloc<float>(n); ap32 = dalloc<float>(n);
        q32 = dalloc<float>(n);
        vs32 = vs;
        L32cap = L;
    }

    void ensure_tensor_vectors() {
        const size_t L = (size_t)std::max(1L, L3());
        if (L <= tcap) return;
        dfree(tb); dfree(txa);
        tb = dalloc<double>((size_t)kMaxRhs * L);
        txa = dalloc<double>((size_t)kMaxRhs * L);
        tcap = L;
    }

    void ensure_hist(size_t n) {
        if (n <= hist_cap) return;
        dfree(hist);
        hist = dalloc<double>(n);
        hist_cap = n;
    }

    double* ensure_field() {
        if (!field) field = dalloc<double>(3 * (size_t)g.N);
        return field;
    }

    void require_medium() const {
        if (!has_medium) throw Err(FFTPERM_ERUNTIME, "no medium set on this context");
    }
    void require_support() const {
        require_medium();
        if (nb == 0) throw Err(FFTPERM_ERUNTIME, "Support set is empty; cannot run variational solve");
    }

    template <typename R>
    void chain(const XSource& src, const XSink& sink, const int* skip, double* qpart = nullptr);

    double host_scalar(const double* d) {
        double h;
        HIPCHK(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, s));
        sync();
        return h;
    }
};

// Waves per workgroup of the FP32 x and y passes fused over z planes when they are on for this
// grid shape (autotuner registry), else 0.
static int plane_on(const fftperm_ctx* c) {
    const Geometry3& g = c->g;
    const int v = tuned_variant(kKindPlane, 1, g.nx, (long)g.ny * g.nz);
    return (v == 1 || v == 2) && plane_pass_shape_ok(g) ? 8 * v : 0;
}

template <typename R, typename S_, typename G_, typename T_>
static void run_chain(fftperm_ctx* c, S_* S, const G_* G, const T_& t, const XSource& src, const XSink& sink,
                      const int* skip, double* qpart, int fp32) {
    fftperm_ctx::ProfRec rec{};
    if (c->prof) {
        rec.a = c->take_event();
        rec.b = c->take_event();
        rec.c = c->take_event();
        rec.d = c->take_event();
        rec.tag = c->cur_tag;
        rec.fp32 = fp32;
        HIPCHK(hipEventRecord(rec.a, c->s));
    }
    const int plane = fp32 && plane_src_ok(src) && plane_sink_ok(sink) ? plane_on(c) : 0;
    if (plane)
        launch_xyfwd_plane(src, reinterpret_cast<cx<float>*>(S), c->g, reinterpret_cast<const cx<float>*>(t.twx), plane,
                           skip, c->s);
    else {
        launch_xfwd<R>(src, S, c->g, t, skip, c->s);
        launch_ypass<R>(S, -1, c->g, t, skip, c->s);
    }
    // the z-pass launch records rec.b / rec.c at its own start and end
    if (c->prof) zpass_events() = KernelEvents{rec.b, rec.c};
    c->zparts = launch_zgreen<R>(S, G, c->g, t, qpart, skip, c->s);
    if (plane)
        c->xparts = launch_yxinv_plane(reinterpret_cast<const cx<float>*>(S), sink, c->g,
                                       reinterpret_cast<const cx<float>*>(t.twx), plane, skip, c->s);
    else {
        launch_ypass<R>(S, +1, c->g, t, skip, c->s);
        c->xparts = launch_xinv<R>(S, sink, c->g, t, skip, c->s);
    }
    if (c->prof) {
        HIPCHK(hipEventRecord(rec.d, c->s));
        c->prof_pending.push_back(rec);
    }
}

template <>
void fftperm_ctx::chain<double>(const XSource& src, const XSink& sink, const int* skip, double* qpart) {
    run_chain<double>(this, S64, G64, t64, src, sink, skip, qpart, 0);
}

template <>
void fftperm_ctx::chain<float>(const XSource& src, const XSink& sink, const int* skip, double* qpart) {
    run_chain<float>(this, S32, G32, t32, src, sink, skip, qpart, 1);
}

namespace {

LineBits line_bits(const fftperm_ctx* c) {
    LineBits lb{};
    if (c->has_medium && c->lbits != nullptr) {
        lb.bits = c->lbits;
        lb.lstart = c->lstart;
        lb.nb = (long)c->nb;
    }
    return lb;
}

XSource src_none(const fftperm_ctx* c) {
    XSource sr;
    std::memset(&sr, 0, sizeof(sr));
    sr.code = c->code;
    sr.lines = line_bits(c);
    return sr;
}

// Sampled live timing: with fftperm_profile(ctx, 2) one chunk of Krylov iterations in
// kProfSamplePeriod runs eagerly with events and the others replay graphs without them (events
// cost ~3% of a 256^3 solve and rule out graph replay); with fftperm_profile(ctx, 1) every chunk is
// timed.  ChunkTiming decides for the chunk about to run and switches the event recording off for
// an untimed one (graph capture included) until it goes out of scope.
constexpr long long kProfSamplePeriod = 8;
struct ChunkTiming {
    fftperm_ctx* c;
    bool saved;
    bool timed;
    explicit ChunkTiming(fftperm_ctx* ctx) : c(ctx), saved(ctx->prof) {
        timed = ctx->prof && (!ctx->prof_sampled || (ctx->prof_chunk++ % kProfSamplePeriod) == 0);
        if (!timed) ctx->prof = false;
    }
    ~ChunkTiming() { c->prof = saved; }
    bool graphs() const { return c->use_graphs && !timed; }
};

XSink sink_none(const fftperm_ctx* c) {
    XSink sk;
    std::memset(&sk, 0, sizeof(sk));
    sk.code = c->code;
    sk.lines = line_bits(c);
    return sk;
}

void check_ctx(const fftperm_ctx* c) {
    if (!c) throw Err(FFTPERM_EINVAL, "null context");
    HIPCHK(hipSetDevice(c->device));
}

// f0 coefficients of BackgroundForce::build (src/bignonnet/BackgroundForce.cpp:34-35).
void f0_coefficients(const fftperm_ctx* c, const double grad[3], double af[3], double as[3]) {
    const double c_fluid = -1.0;
    const double c_support = (c->nb == 0) ? 0.0 : -(double)c->nf / (double)c->nb;
    for (int k = 0; k < 3; ++k) {
        af[k] = c_fluid * grad[k];
        as[k] = c_support * grad[k];
    }
}

// Device operator apply y = A x (FP64 or FP32 transform chain), all pointers device.
template <typename R>
void device_apply(fftperm_ctx* c, const double* xin, double* yout) {
    const long nb = (long)c->nb;
    const double inv_nb = 1.0 / (double)nb;
    launch_compact_sums(xin, nb, c->part, nullptr, c->s);
    launch_finish_sum(c->part, kRedBlocks, 3, inv_nb, c->small, nullptr, c->s);
    XSource sr = src_none(c);
    sr.compact = xin;
    sr.mean = c->small;
    XSink sk = sink_none(c);
    sk.compact = yout;
    sk.partials = c->part_x;
    c->chain<R>(sr, sk, nullptr);
    launch_finish_sum(c->part_x, (int)c->xparts, kXinvParts, inv_nb, c->small + 4, nullptr, c->s);
    launch_sub_mean(yout, c->small + 4, yout, nb, 1.0, nullptr, c->s);
}

double device_norm(fftperm_ctx* c, const double* a, long n) {
    launch_dot_partials(a, a, n, c->part, c->s);
    launch_finish_sum(c->part, kRedBlocks, 1, 1.0, c->small + 16, nullptr, c->s);
    const double d = c->host_scalar(c->small + 16);
    return std::sqrt(std::max(0.0, d));
}

// b (device) = VariationalRHS::build(grad)  (src/bignonnet/VariationalRHS.cpp:15-31), FP64.
void device_rhs(fftperm_ctx* c, const double grad[3], double* bout) {
    const long nb = (long)c->nb;
    XSource sr = src_none(c);
    f0_coefficients(c, grad, sr.a_fluid, sr.a_support);
    XSink sk = sink_none(c);
    sk.compact = c->tmp;
    sk.partials = c->part_x;
    c->chain<double>(sr, sk, nullptr);
    launch_finish_sum(c->part_x, (int)c->xparts, kXinvParts, 1.0 / (double)nb, c->small + 4, nullptr, c->s);
    launch_sub_mean(c->tmp, c->small + 4, bout, nb, -1.0, nullptr, c->s);
}

// One MINRES iteration on device; every kernel is a no-op once st->done is set.
template <typename R>
void minres_iteration(fftperm_ctx* c) {
    const long nb = (long)c->nb;
    const double inv_nb = 1.0 / (double)nb;
    MinresState* st = c->st;
    XSource sr = src_none(c);
    sr.compact = c->v;
    sr.mean = st->mean_in;
    XSink sk = sink_none(c);
    sk.compact = c->z;
    sk.partials = c->part_x;
    // alpha = <P_B v, G scatter(P_B v)> comes from the fused z pass (Parseval)
    c->chain<R>(sr, sk, &st->done, c->part_z);
    launch_mr_finish_mean_out(c->part_x, (int)c->xparts, c->part_z, (int)c->zparts, st, inv_nb,
                              1.0 / (double)c->g.N, c->s);
    launch_mr_axpy_norm(c->z, c->v, c->v_old, st, nb, c->part, c->s);
    launch_mr_givens(c->part, st, c->hist, c->s);
    launch_mr_update(c->x, c->v, c->z, c->w_prev, c->w, c->v_old, st, nb, c->part, c->s);
    launch_mr_finish_mean_in(c->part, st, inv_nb, c->s);
    // v_old <- v, v <- v_new (written over v_old); w_prev <- w, w <- w_next (written over w_prev)
    std::swap(c->v_old, c->v);
    std::swap(c->w_prev, c->w);
}

// Graph of two MINRES iterations for the current buffers (captured on first use).
template <typename R>
hipGraphExec_t minres_graph(fftperm_ctx* c) {
    fftperm_ctx::GraphKey k{};
    k.fp32 = sizeof(R) == 4 ? 1 : 0;
    const void* ptrs[9] = {c->x, c->v_old, c->v, c->z, c->w_prev, c->w, c->st, c->hist, c->part_x};
    for (int i = 0; i < 9; ++i) k.p[i] = ptrs[i];
    k.nb = (long)c->nb;
    k.tuning = tuning_generation();  // a pinned or re-tuned variant changes the captured launches
    for (const auto& g : c->graphs)
        if (g.key == k) return g.exec;
    if (c->graphs.size() >= 8) {
        HIPCHK(hipGraphExecDestroy(c->graphs.front().exec));
        c->graphs.erase(c->graphs.begin());
    }
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    minres_iteration<R>(c);
    minres_iteration<R>(c);
    HIPCHK(hipStreamEndCapture(c->s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    c->graphs.push_back({k, exec});
    return exec;
}

struct MinresOut {
    uint64_t iterations = 0;
    double final_rel = 1.0;
    int converged = 0;
    uint64_t hist_len = 0;
};

// MinresSolver::solve (src/solver/MinresSolver.cpp:46-218) with the device operator.
// bdev: rhs; x0dev: initial guess or null; result in c->x; history in c->hist[0..hist_len).
// Precision schedule of the operator applies: FP64 while the recurrence relative residual is
// above fp32_below, FP32 from the first chunk boundary at or below it (fp32_below <= 0: FP64
// throughout; >= 1: FP32 throughout).  The initial residual and every Lanczos scalar stay FP64.
// *fp32_iters (may be null) receives the number of iterations that used the FP32 operator.
MinresOut device_minres(fftperm_ctx* c, const double* bdev, const double* x0dev, double tol, uint64_t maxit,
                        double fp32_below, uint64_t* fp32_iters) {
    if (!(tol > 0.0)) throw Err(FFTPERM_EINVAL, "MinresSolver tolerance must be positive");
    const long L = c->L3();
    const long nb = (long)c->nb;
    MinresOut out;
    c->ensure_hist((size_t)maxit + 2);
    if (x0dev)
        HIPCHK(hipMemcpyAsync(c->x, x0dev, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    else
        HIPCHK(hipMemsetAsync(c->x, 0, L * sizeof(double), c->s));
    const double rhs_raw_norm = device_norm(c, bdev, L);
    if (rhs_raw_norm == 0.0) {
        const double zero = 0.0;
        HIPCHK(hipMemcpyAsync(c->hist, &zero, sizeof(double), hipMemcpyHostToDevice, c->s));
        c->sync();
        out.converged = 1;
        out.final_rel = 0.0;
        out.hist_len = 1;
        return out;
    }
    const double rhs_norm = std::max(rhs_raw_norm, DBL_MIN);
    // r0 = b - A x0 (A 0 = 0 exactly, so the empty guess gives r0 = b)
    bool use32 = fp32_below >= 1.0;
    if (x0dev) {
        if (use32)
            device_apply<float>(c, c->x, c->z);
        else
            device_apply<double>(c, c->x, c->z);
        launch_axpby(1.0, bdev, -1.0, c->z, c->z, L, c->s);
    } else {
        HIPCHK(hipMemcpyAsync(c->z, bdev, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    }
    const double beta1 = device_norm(c, c->z, L);
    const double rel0 = beta1 / rhs_norm;
    HIPCHK(hipMemcpyAsync(c->hist, &rel0, sizeof(double), hipMemcpyHostToDevice, c->s));
    out.final_rel = rel0;
    out.hist_len = 1;
    if (rel0 <= tol || beta1 == 0.0) {
        c->sync();
        out.converged = 1;
        return out;
    }
    launch_scale_copy(c->z, beta1, c->v, L, c->s);
    HIPCHK(hipMemsetAsync(c->v_old, 0, L * sizeof(double), c->s));
    HIPCHK(hipMemsetAsync(c->w_prev, 0, L * sizeof(double), c->s));
    HIPCHK(hipMemsetAsync(c->w, 0, L * sizeof(double), c->s));
    MinresState h;
    std::memset(&h, 0, sizeof(h));
    h.beta = beta1;
    h.cs_old = -1.0;
    h.sn_old = 0.0;
    h.cs = -1.0;
    h.sn = 0.0;
    h.phi_bar = beta1;
    h.rhs_norm = rhs_norm;
    h.tol = tol;
    h.rel_res = rel0;
    h.iterations = 0;
    h.maxit = (long long)maxit;
    h.done = (maxit == 0) ? 1 : 0;
    HIPCHK(hipMemcpyAsync(c->st, &h, sizeof(h), hipMemcpyHostToDevice, c->s));
    launch_compact_sums(c->v, nb, c->part, nullptr, c->s);
    launch_finish_sum(c->part, kRedBlocks, 3, 1.0 / (double)nb, c->st->mean_in, nullptr, c->s);
    launch_finish_sum(c->part, kRedBlocks, 3, 1.0, c->st->vsum, nullptr, c->s);
    const uint64_t chunk = (c->g.N >= (1L << 21)) ? 4 : 32;
    uint64_t launched = 0, switched_at = use32 ? 0 : UINT64_MAX;
    if (!use32 && fp32_below > 0.0 && rel0 <= fp32_below) {
        use32 = true;
        switched_at = 0;
    }
    int done = h.done;
    while (!done && launched < maxit) {
        const uint64_t k = std::min<uint64_t>(chunk, maxit - launched);
        // pairs of iterations replay a captured graph (not while live timing records events)
        const ChunkTiming timing(c);
        const bool graph = timing.graphs();
        uint64_t i = 0;
        if (graph) {
            for (; i + 2 <= k; i += 2) {
                hipGraphExec_t ex = use32 ? minres_graph<float>(c) : minres_graph<double>(c);
                HIPCHK(hipGraphLaunch(ex, c->s));
            }
        }
        for (; i < k; ++i) {
            c->cur_tag = (long long)(launched + i);
            if (use32)
                minres_iteration<float>(c);
            else
                minres_iteration<double>(c);
        }
        c->cur_tag = -1;
        launched += k;
        if (!use32 && fp32_below > 0.0) {
            HIPCHK(hipMemcpyAsync(&h, c->st, sizeof(h), hipMemcpyDeviceToHost, c->s));
            c->sync();
            done = h.done;
            if (!done && h.rel_res <= fp32_below) {
                use32 = true;
                switched_at = launched;
            }
        } else {
            HIPCHK(hipMemcpyAsync(&done, &c->st->done, sizeof(int), hipMemcpyDeviceToHost, c->s));
            c->sync();
        }
    }
    HIPCHK(hipMemcpyAsync(&h, c->st, sizeof(h), hipMemcpyDeviceToHost, c->s));
    c->sync();
    c->prof_resolve(h.iterations);
    out.iterations = (uint64_t)h.iterations;
    out.final_rel = (h.iterations > 0) ? h.rel_res : rel0;
    out.converged = h.converged || (out.final_rel <= tol);
    out.hist_len = (uint64_t)h.iterations + 1;
    if (fp32_iters)
        *fp32_iters = (switched_at == UINT64_MAX || (uint64_t)h.iterations <= switched_at)
                          ? 0
                          : (uint64_t)h.iterations - switched_at;
    return out;
}

// Mixed precision: FP64 iterative refinement around MINRES sweeps on r = b - A x, each to the
// recurrence residual max(c->mixed_floor, 0.5 tol / r) with r the FP64 true relative residual.
// Sweep 1 may start with FP64 operator applies and switch to FP32 once its recurrence residual
// is below c->mixed_switch (default: FP32 throughout; measured at 256^3, an FP64 head start does
// not remove the second sweep, because the FP32-storage operator error sets the attainable
// residual whichever iterations it enters); later sweeps apply the FP32 operator throughout.
MinresOut device_minres_mixed(fftperm_ctx* c, const double* bdev, const double* x0dev, double tol, uint64_t maxit,
                              uint64_t* sweeps_out, std::vector<double>* hist_out) {
    if (!(tol > 0.0)) throw Err(FFTPERM_EINVAL, "MinresSolver tolerance must be positive");
    const long L = c->L3();
    MinresOut out;
    const double bnorm = device_norm(c, bdev, L);
    double* xacc = c->dx;  // accumulated solution
    if (x0dev) {
        if (x0dev != xacc)
            HIPCHK(hipMemcpyAsync(xacc, x0dev, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    } else {
        HIPCHK(hipMemsetAsync(xacc, 0, L * sizeof(double), c->s));
    }
    bool xacc_zero = (x0dev == nullptr);
    if (bnorm == 0.0) {
        HIPCHK(hipMemcpyAsync(c->x, xacc, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
        c->sync();
        out.converged = 1;
        out.final_rel = 0.0;
        if (hist_out) hist_out->push_back(0.0);
        out.hist_len = 1;
        *sweeps_out = 0;
        return out;
    }
    // r = b - A64 x
    auto true_residual = [&]() {
        if (!xacc_zero) {
            device_apply<double>(c, xacc, c->tmp);
            launch_axpby(1.0, bdev, -1.0, c->tmp, c->r, L, c->s);
        } else {
            HIPCHK(hipMemcpyAsync(c->r, bdev, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
        }
        return device_norm(c, c->r, L) / bnorm;
    };
    double rel = true_residual();
    if (hist_out) hist_out->push_back(rel);
    uint64_t sweeps = 0;
    while (rel > tol && out.iterations < maxit && sweeps < c->mixed_sweeps) {
        const bool first = (sweeps == 0);
        const double itol = std::max(c->mixed_floor, 0.5 * tol / rel);
        uint64_t it32 = 0;
        MinresOut in = device_minres(c, c->r, nullptr, itol, maxit - out.iterations, first ? c->mixed_switch : 2.0,
                                     &it32);
        c->fp32_iters += it32;
        out.iterations += in.iterations;
        ++sweeps;
        // x += dx (MINRES result in c->x)
        launch_axpby(1.0, xacc, 1.0, c->x, xacc, L, c->s);
        xacc_zero = false;
        rel = true_residual();
        if (hist_out) hist_out->push_back(rel);
        if (in.iterations == 0) break;
    }
    HIPCHK(hipMemcpyAsync(c->x, xacc, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    c->sync();
    out.final_rel = rel;
    out.converged = rel <= tol ? 1 : 0;
    out.hist_len = hist_out ? hist_out->size() : 0;
    *sweeps_out = sweeps;
    return out;
}

// One CG iteration on the device (p = c->v, q = c->z, x = c->x group solution, r = c->r); every
// kernel is a no-op once cst->done is set.
template <typename R>
void cg_iteration(fftperm_ctx* c) {
    const long nb = (long)c->nb;
    const double inv_nb = 1.0 / (double)nb;
    CgState* st = c->cst;
    XSource sr = src_none(c);
    sr.compact = c->v;
    sr.mean = st->mean_in;
    XSink sk = sink_none(c);
    sk.compact = c->z;
    sk.partials = c->part_x;
    // p.A p comes from the fused z pass (Parseval)
    c->chain<R>(sr, sk, &st->done, c->part_z);
    launch_cg_alpha(c->part_x, (int)c->xparts, c->part_z, (int)c->zparts, st, inv_nb, 1.0 / (double)c->g.N, c->s);
    launch_cg_update(c->x, c->r, c->v, c->z, st, nb, c->part, c->s);
    launch_cg_beta(c->part, st, c->hist, c->s);
    launch_cg_direction(c->v, c->r, st, nb, c->part, c->s);
    launch_cg_finish_mean_in(c->part, st, inv_nb, c->s);
}

template <typename R>
hipGraphExec_t cg_graph(fftperm_ctx* c) {
    fftperm_ctx::GraphKey k{};
    k.fp32 = sizeof(R) == 4 ? 1 : 0;
    k.solver = 1;
    const void* ptrs[9] = {c->x, c->r, c->v, c->z, c->cst, c->hist, c->part_x, c->part_z, c->part};
    for (int i = 0; i < 9; ++i) k.p[i] = ptrs[i];
    k.nb = (long)c->nb;
    k.tuning = tuning_generation();
    for (const auto& g : c->graphs)
        if (g.key == k) return g.exec;
    if (c->graphs.size() >= 8) {
        HIPCHK(hipGraphExecDestroy(c->graphs.front().exec));
        c->graphs.erase(c->graphs.begin());
    }
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    cg_iteration<R>(c);
    cg_iteration<R>(c);
    HIPCHK(hipStreamEndCapture(c->s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    c->graphs.push_back({k, exec});
    return exec;
}

// MIXED-mode solve by conjugate gradients with reliable updates: the iterations apply the operator
// in precision R (FP32 in MIXED mode) with FP64 vectors; the recursive residual r drifts from the
// true one by the accumulated FP32 apply errors, so whenever ||r|| falls below reliable_delta times
// its peak since the last update (or below tol ||b||) the group solution is added to the FP64
// accumulator and r is replaced by b - A64 x_acc.  The solve ends when that true residual meets
// tol, or after maxit iterations.  Search directions survive the updates, so no Krylov
// information is thrown away (unlike restarted refinement sweeps).  The variational operator is
// symmetric positive semi-definite and b lies in its range, so CG applies.
// Result in c->x; *updates_out = reliable updates done; hist_out: the true residuals.
template <typename R>
MinresOut device_cg_reliable(fftperm_ctx* c, const double* bdev, const double* x0dev, double tol, uint64_t maxit,
                             uint64_t* updates_out, std::vector<double>* hist_out) {
    if (!(tol > 0.0)) throw Err(FFTPERM_EINVAL, "MinresSolver tolerance must be positive");
    const long L = c->L3();
    const long nb = (long)c->nb;
    const double inv_nb = 1.0 / (double)nb;
    MinresOut out;
    *updates_out = 0;
    const double bnorm = device_norm(c, bdev, L);
    double* xacc = c->dx;
    if (x0dev) {
        if (x0dev != xacc)
            HIPCHK(hipMemcpyAsync(xacc, x0dev, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    } else {
        HIPCHK(hipMemsetAsync(xacc, 0, L * sizeof(double), c->s));
    }
    bool xacc_zero = (x0dev == nullptr);
    auto finish = [&](double rel) {
        HIPCHK(hipMemcpyAsync(c->x, xacc, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
        c->sync();
        out.final_rel = rel;
        out.converged = rel <= tol ? 1 : 0;
        out.hist_len = hist_out ? hist_out->size() : 0;
    };
    if (bnorm == 0.0) {
        if (hist_out) hist_out->push_back(0.0);
        finish(0.0);
        out.converged = 1;
        return out;
    }
    // r = b - A64 x_acc, its norm relative to ||b||
    auto true_residual = [&]() {
        if (!xacc_zero) {
            device_apply<double>(c, xacc, c->tmp);
            launch_axpby(1.0, bdev, -1.0, c->tmp, c->r, L, c->s);
        } else {
            HIPCHK(hipMemcpyAsync(c->r, bdev, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
        }
        return device_norm(c, c->r, L) / bnorm;
    };
    double rel = true_residual();
    if (hist_out) hist_out->push_back(rel);
    if (rel <= tol || maxit == 0) {
        finish(rel);
        return out;
    }
    c->ensure_hist((size_t)maxit + 2);
    HIPCHK(hipMemsetAsync(c->x, 0, L * sizeof(double), c->s));
    HIPCHK(hipMemcpyAsync(c->v, c->r, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    CgState h;
    std::memset(&h, 0, sizeof(h));
    const double rn0 = rel * bnorm;
    h.rr = rn0 * rn0;
    h.rnorm_max = rn0;
    h.rhs_norm = bnorm;
    h.tol = tol;
    h.delta = c->reliable_delta;
    h.maxit = (long long)maxit;
    HIPCHK(hipMemcpyAsync(c->cst, &h, sizeof(h), hipMemcpyHostToDevice, c->s));
    launch_compact_sums(c->v, nb, c->part, nullptr, c->s);
    launch_finish_sum(c->part, kRedBlocks, 3, inv_nb, c->cst->mean_in, nullptr, c->s);
    const uint64_t chunk = (c->g.N >= (1L << 21)) ? 4 : 32;
    long long seg_start = 0;  // iterations done before the current segment (between updates)
    for (;;) {
        // iterate until the device pauses (update owed) or stops (maxit, breakdown); live timing
        // tags each eager launch with its index in the segment so that no-op launches are dropped
        int done = 0;
        long long seg_launch = 0;
        while (!done) {
            const ChunkTiming timing(c);
            const bool graph = timing.graphs();
            uint64_t i = 0;
            if (graph)
                for (; i + 2 <= chunk; i += 2, seg_launch += 2) HIPCHK(hipGraphLaunch(cg_graph<R>(c), c->s));
            for (; i < chunk; ++i) {
                c->cur_tag = seg_launch++;
                cg_iteration<R>(c);
            }
            c->cur_tag = -1;
            HIPCHK(hipMemcpyAsync(&done, &c->cst->done, sizeof(int), hipMemcpyDeviceToHost, c->s));
            c->sync();
        }
        HIPCHK(hipMemcpyAsync(&h, c->cst, sizeof(h), hipMemcpyDeviceToHost, c->s));
        c->sync();
        c->prof_resolve(h.iterations - seg_start);
        seg_start = h.iterations;
        // fold the group solution into the accumulator and take the FP64 true residual
        launch_axpby(1.0, xacc, 1.0, c->x, xacc, L, c->s);
        xacc_zero = false;
        HIPCHK(hipMemsetAsync(c->x, 0, L * sizeof(double), c->s));
        rel = true_residual();
        if (hist_out) hist_out->push_back(rel);
        ++*updates_out;
        if (rel <= tol || !h.pause || h.breakdown || h.iterations >= h.maxit) break;
        // continue from the true residual with the same search direction
        const double rn = rel * bnorm;
        h.beta = (rn * rn) / h.rr;
        h.rr = rn * rn;
        h.rnorm_max = rn;
        h.done = 0;
        h.pause = 0;
        HIPCHK(hipMemcpyAsync(c->cst, &h, sizeof(h), hipMemcpyHostToDevice, c->s));
        launch_cg_direction(c->v, c->r, c->cst, nb, c->part, c->s);
        launch_cg_finish_mean_in(c->part, c->cst, inv_nb, c->s);
    }
    c->prof_resolve(LLONG_MAX);
    out.iterations = (uint64_t)h.iterations;
    if (sizeof(R) == 4) c->fp32_iters += (uint64_t)h.iterations;
    finish(rel);
    return out;
}

// One CR iteration on the device (r = c->r, q = c->z, p = c->v, Ap = c->w, x = c->x group
// solution); every kernel is a no-op once rst->done is set.
template <typename R>
void cr_iteration(fftperm_ctx* c) {
    const long nb = (long)c->nb;
    const double inv_nb = 1.0 / (double)nb;
    CrState* st = c->rst;
    XSource sr = src_none(c);
    sr.compact = c->r;
    sr.mean = st->mean_in;
    XSink sk = sink_none(c);
    sk.compact = c->z;
    sk.partials = c->part_x;
    // r.A r comes from the fused z pass (Parseval)
    c->chain<R>(sr, sk, &st->done, c->part_z);
    launch_cr_beta(c->part_x, (int)c->xparts, c->part_z, (int)c->zparts, st, inv_nb, 1.0 / (double)c->g.N, c->s);
    launch_cr_direction(c->v, c->w, c->r, c->z, st, nb, c->part, c->s);
    launch_cr_alpha(c->part, st, c->s);
    launch_cr_update(c->x, c->r, c->v, c->w, st, nb, c->part, c->s);
    launch_cr_check(c->part, st, c->hist, inv_nb, c->s);
}

// The same on FP32 vectors (x32, r32, p32, ap32, q32).
template <typename R>
void cr32_iteration(fftperm_ctx* c) {
    const long nb = (long)c->nb;
    const double inv_nb = 1.0 / (double)nb;
    CrState* st = c->rst;
    XSource sr = src_none(c);
    sr.compact32 = c->r32;
    sr.mean = st->mean_in;
    XSink sk = sink_none(c);
    sk.compact32 = c->q32;
    sk.partials = c->part_x;
    c->chain<R>(sr, sk, &st->done, c->part_z);
    launch_cr_beta(c->part_x, (int)c->xparts, c->part_z, (int)c->zparts, st, inv_nb, 1.0 / (double)c->g.N, c->s);
    const bool stream = cr32_streams(nb, 1, 3 * c->g.Nc * (long)sizeof(float2));
    launch_cr32_direction(c->p32, c->ap32, c->r32, c->q32, st, nb, c->part, c->s, stream);
    launch_cr_alpha(c->part, st, c->s);
    launch_cr32_update(c->x32, c->r32, c->p32, c->ap32, st, nb, c->part, c->s, stream);
    launch_cr_check(c->part, st, c->hist, inv_nb, c->s);
}

template <typename R, bool V32>
hipGraphExec_t cr_graph(fftperm_ctx* c) {
    fftperm_ctx::GraphKey k{};
    k.fp32 = sizeof(R) == 4 ? 1 : 0;
    k.solver = V32 ? 3 : 2;
    const void* ptrs[9] = {V32 ? (const void*)c->x32 : c->x, V32 ? (const void*)c->r32 : c->r,
                           V32 ? (const void*)c->p32 : c->v, V32 ? (const void*)c->q32 : c->z,
                           V32 ? (const void*)c->ap32 : c->w, c->rst, c->hist, c->part_x, c->part};
    for (int i = 0; i < 9; ++i) k.p[i] = ptrs[i];
    k.nb = (long)c->nb;
    k.tuning = tuning_generation();
    for (const auto& g : c->graphs)
        if (g.key == k) return g.exec;
    if (c->graphs.size() >= 8) {
        HIPCHK(hipGraphExecDestroy(c->graphs.front().exec));
        c->graphs.erase(c->graphs.begin());
    }
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < 2; ++i) {
        if (V32)
            cr32_iteration<R>(c);
        else
            cr_iteration<R>(c);
    }
    HIPCHK(hipStreamEndCapture(c->s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    c->graphs.push_back({k, exec});
    return exec;
}

// MIXED-mode solve by conjugate residuals with reliable updates: the same FP64 accumulator and
// true-residual replacement as device_cg_reliable, around the CR recurrence (residual-minimising
// like MINRES, and carrying r explicitly).  The direction p and its image Ap survive the
// replacements.  V32: the iterated vectors (group solution, r, p, Ap, apply output) are FP32; the
// accumulator, b, the true residual and every scalar stay FP64.  Result in c->x;
// *updates_out = FP64 checks after the initial one; hist_out: the true residuals.
template <typename R, bool V32>
MinresOut device_cr_reliable(fftperm_ctx* c, const double* bdev, const double* x0dev, double tol, uint64_t maxit,
                             uint64_t* updates_out, std::vector<double>* hist_out) {
    if (!(tol > 0.0)) throw Err(FFTPERM_EINVAL, "MinresSolver tolerance must be positive");
    const long L = c->L3();
    const long nb = (long)c->nb;
    const double inv_nb = 1.0 / (double)nb;
    MinresOut out;
    *updates_out = 0;
    const double bnorm = device_norm(c, bdev, L);
    double* xacc = c->dx;
    if (x0dev) {
        if (x0dev != xacc)
            HIPCHK(hipMemcpyAsync(xacc, x0dev, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    } else {
        HIPCHK(hipMemsetAsync(xacc, 0, L * sizeof(double), c->s));
    }
    bool xacc_zero = (x0dev == nullptr);
    auto finish = [&](double rel) {
        HIPCHK(hipMemcpyAsync(c->x, xacc, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
        c->sync();
        out.final_rel = rel;
        out.converged = rel <= tol ? 1 : 0;
        out.hist_len = hist_out ? hist_out->size() : 0;
    };
    if (bnorm == 0.0) {
        if (hist_out) hist_out->push_back(0.0);
        finish(0.0);
        out.converged = 1;
        return out;
    }
    auto true_residual = [&]() {
        if (!xacc_zero) {
            device_apply<double>(c, xacc, c->tmp);
            launch_axpby(1.0, bdev, -1.0, c->tmp, c->r, L, c->s);
        } else {
            HIPCHK(hipMemcpyAsync(c->r, bdev, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
        }
        return device_norm(c, c->r, L) / bnorm;
    };
    auto r_mean = [&]() {
        launch_compact_sums(c->r, nb, c->part, nullptr, c->s);
        launch_finish_sum(c->part, kRedBlocks, 3, inv_nb, c->rst->mean_in, nullptr, c->s);
    };
    double rel = true_residual();
    if (hist_out) hist_out->push_back(rel);
    if (rel <= tol || maxit == 0) {
        finish(rel);
        return out;
    }
    c->ensure_hist((size_t)maxit + 2);
    if (V32) {
        c->ensure_vectors32();
        HIPCHK(hipMemsetAsync(c->x32, 0, L * sizeof(float), c->s));
        launch_to_f32(c->r, c->r32, L, c->s);
    } else {
        HIPCHK(hipMemsetAsync(c->x, 0, L * sizeof(double), c->s));
    }
    CrState h;
    std::memset(&h, 0, sizeof(h));
    h.rnorm_max = rel * bnorm;
    h.rhs_norm = bnorm;
    h.tol = tol;
    h.delta = c->reliable_delta;
    h.maxit = (long long)maxit;
    h.first = 1;
    HIPCHK(hipMemcpyAsync(c->rst, &h, sizeof(h), hipMemcpyHostToDevice, c->s));
    r_mean();
    const uint64_t chunk = (c->g.N >= (1L << 21)) ? 4 : 32;
    long long seg_start = 0;
    for (;;) {
        int done = 0;
        long long seg_launch = 0;
        while (!done) {
            const ChunkTiming timing(c);
            const bool graph = timing.graphs();
            uint64_t i = 0;
            if (graph)
                for (; i + 2 <= chunk; i += 2, seg_launch += 2) HIPCHK(hipGraphLaunch(cr_graph<R, V32>(c), c->s));
            for (; i < chunk; ++i) {
                c->cur_tag = seg_launch++;
                if (V32)
                    cr32_iteration<R>(c);
                else
                    cr_iteration<R>(c);
            }
            c->cur_tag = -1;
            HIPCHK(hipMemcpyAsync(&done, &c->rst->done, sizeof(int), hipMemcpyDeviceToHost, c->s));
            c->sync();
        }
        HIPCHK(hipMemcpyAsync(&h, c->rst, sizeof(h), hipMemcpyDeviceToHost, c->s));
        c->sync();
        c->prof_resolve(h.iterations - seg_start);
        seg_start = h.iterations;
        if (V32) {
            launch_accumulate_f32(xacc, c->x32, L, c->s);
        } else {
            launch_axpby(1.0, xacc, 1.0, c->x, xacc, L, c->s);
            HIPCHK(hipMemsetAsync(c->x, 0, L * sizeof(double), c->s));
        }
        xacc_zero = false;
        rel = true_residual();
        if (hist_out) hist_out->push_back(rel);
        ++*updates_out;
        if (rel <= tol || !h.pause || h.breakdown || h.iterations >= h.maxit) break;
        // continue from the true residual; p, Ap and r.A r of the last apply carry over
        const double rn = rel * bnorm;
        h.rr = rn * rn;
        h.rnorm_max = rn;
        h.done = 0;
        h.pause = 0;
        HIPCHK(hipMemcpyAsync(c->rst, &h, sizeof(h), hipMemcpyHostToDevice, c->s));
        r_mean();
        if (V32) launch_to_f32(c->r, c->r32, L, c->s);
    }
    c->prof_resolve(LLONG_MAX);
    out.iterations = (uint64_t)h.iterations;
    if (sizeof(R) == 4) c->fp32_iters += (uint64_t)h.iterations;
    finish(rel);
    return out;
}

// ------------------------------------------------------------------ multi-RHS CR on FP32 vectors
// The three load directions of a tensor iterate in lockstep (src/app/PermeabilityRunner.cpp:186-198
// solves them one after the other): side b owns CR state rst[b], FP32 vector slot b (at b * vs32),
// partial rows b and, when the passes serve all sides per launch, spectrum slot b (components
// 3b .. 3b + 2 of S32); every kernel skips a side whose state is done, so each side runs exactly
// the single-RHS recurrence.

// int stride between the done flags of consecutive CR states
constexpr int kCrDoneStride = (int)(sizeof(CrState) / sizeof(int));
static_assert(sizeof(CrState) % sizeof(int) == 0, "CR states are laid out in whole ints");

long cr_xstride(const fftperm_ctx* c) { return (long)kXinvParts * xinv_max_blocks(c->g); }
long cr_zstride(const fftperm_ctx* c) { return zgreen_max_blocks(c->g); }

// The FP32 five-sweep apply q_b = gather(G scatter(r_b - mean_b)) of nrhs sides: side after side
// (tensor_batch 1), or pass after pass over all sides (tensor_batch 2: one y launch for the nine
// components, one z launch reading the Green tensor once).
void chain32_batch(fftperm_ctx* c, int nrhs) {
    const Geometry3& g = c->g;
    CrState* st = c->rst;
    const long spec = 3 * g.Nc;
    auto source = [&](int b) {
        XSource sr = src_none(c);
        sr.compact32 = c->r32 + b * c->vs32;
        sr.mean = st[b].mean_in;
        return sr;
    };
    auto sink = [&](int b) {
        XSink sk = sink_none(c);
        sk.compact32 = c->q32 + b * c->vs32;
        sk.partials = c->part_x + b * cr_xstride(c);
        return sk;
    };
    auto prof_begin = [&](int batch, int side) {
        fftperm_ctx::ProfRec rec{};
        rec.a = c->take_event();
        rec.b = c->take_event();
        rec.c = c->take_event();
        rec.d = c->take_event();
        rec.tag = c->cur_tag;
        rec.fp32 = 1;
        rec.batch = batch;
        rec.side = side;
        HIPCHK(hipEventRecord(rec.a, c->s));
        return rec;
    };
    if (c->tensor_batch != 2) {
        // The sides' chains run one after the other and a chain consumes its spectra before the
        // next starts, so every side uses the first side's spectrum buffer: it stays in the
        // last-level cache from one chain to the next (FP32 at 256^3: 214 MB).
        for (int b = 0; b < nrhs; ++b) {
            const int* skip = &st[b].done;
            cx<float>* S = c->S32;
            fftperm_ctx::ProfRec rec{};
            if (c->prof) rec = prof_begin(0, b);
            const int plane = plane_src_ok(source(b)) && plane_sink_ok(sink(b)) ? plane_on(c) : 0;
            if (plane)
                launch_xyfwd_plane(source(b), S, g, c->t32.twx, plane, skip, c->s);
            else {
                launch_xfwd<float>(source(b), S, g, c->t32, skip, c->s);
                launch_ypass<float>(S, -1, g, c->t32, skip, c->s);
            }
            if (c->prof) zpass_events() = KernelEvents{rec.b, rec.c};
            c->zparts = launch_zgreen<float>(S, c->G32, g, c->t32, c->part_z + b * cr_zstride(c), skip, c->s);
            if (plane)
                c->xparts = launch_yxinv_plane(S, sink(b), g, c->t32.twx, plane, skip, c->s);
            else {
                launch_ypass<float>(S, +1, g, c->t32, skip, c->s);
                c->xparts = launch_xinv<float>(S, sink(b), g, c->t32, skip, c->s);
            }
            if (c->prof) {
                HIPCHK(hipEventRecord(rec.d, c->s));
                c->prof_pending.push_back(rec);
            }
        }
        return;
    }
    fftperm_ctx::ProfRec rec{};
    if (c->prof) rec = prof_begin(nrhs, -1);
    for (int b = 0; b < nrhs; ++b) launch_xfwd<float>(source(b), c->S32 + b * spec, g, c->t32, &st[b].done, c->s);
    launch_ypass_n<float>(c->S32, -1, g, c->t32, nrhs, &st[0].done, kCrDoneStride, c->s);
    if (c->prof) HIPCHK(hipEventRecord(rec.b, c->s));
    c->zparts = launch_zgreen_batch<float>(c->S32, c->G32, g, c->t32, c->part_z, cr_zstride(c), nrhs, &st[0].done,
                                           kCrDoneStride, c->s);
    if (c->prof) HIPCHK(hipEventRecord(rec.c, c->s));
    launch_ypass_n<float>(c->S32, +1, g, c->t32, nrhs, &st[0].done, kCrDoneStride, c->s);
    for (int b = 0; b < nrhs; ++b)
        c->xparts = launch_xinv<float>(c->S32 + b * spec, sink(b), g, c->t32, &st[b].done, c->s);
    if (c->prof) {
        HIPCHK(hipEventRecord(rec.d, c->s));
        c->prof_pending.push_back(rec);
    }
}

// One lockstep CR iteration of nrhs sides (the steps of cr32_iteration, one launch each; no
// residual history).
void cr32_batch_iteration(fftperm_ctx* c, int nrhs) {
    const long nb = (long)c->nb;
    const double inv_nb = 1.0 / (double)nb;
    const long pstride = 4L * kRedBlocks;
    chain32_batch(c, nrhs);
    launch_cr_beta_n(c->part_x, (int)c->xparts, cr_xstride(c), c->part_z, (int)c->zparts, cr_zstride(c), c->rst, nrhs,
                     inv_nb, 1.0 / (double)c->g.N, c->s);
    // tensor-batch mode 2 keeps a spectrum per side
    const long spectra = 3 * c->g.Nc * (long)sizeof(float2) * (c->tensor_batch == 2 ? nrhs : 1);
    const bool stream = cr32_streams(nb, nrhs, spectra);
    launch_cr32_direction_n(c->p32, c->ap32, c->r32, c->q32, c->rst, nb, c->vs32, c->part, pstride, nrhs, c->s,
                            stream);
    launch_cr_alpha_n(c->part, pstride, c->rst, nrhs, c->s);
    launch_cr32_update_n(c->x32, c->r32, c->p32, c->ap32, c->rst, nb, c->vs32, c->part, pstride, nrhs, c->s, stream);
    launch_cr_check_n(c->part, pstride, c->rst, nullptr, 0, nrhs, inv_nb, c->s);
}

hipGraphExec_t cr32_batch_graph(fftperm_ctx* c, int nrhs) {
    fftperm_ctx::GraphKey k{};
    k.fp32 = 1;
    k.solver = 10 * nrhs + c->tensor_batch;
    const void* ptrs[9] = {c->x32, c->r32, c->p32, c->q32, c->ap32, c->rst, c->part_z, c->part_x, c->part};
    for (int i = 0; i < 9; ++i) k.p[i] = ptrs[i];
    k.nb = (long)c->nb;
    k.tuning = tuning_generation();
    for (const auto& g : c->graphs)
        if (g.key == k) return g.exec;
    if (c->graphs.size() >= 8) {
        HIPCHK(hipGraphExecDestroy(c->graphs.front().exec));
        c->graphs.erase(c->graphs.begin());
    }
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < 2; ++i) cr32_batch_iteration(c, nrhs);
    HIPCHK(hipStreamEndCapture(c->s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    c->graphs.push_back({k, exec});
    return exec;
}

// device_cr_reliable<float, true> for nrhs sides at once: side b solves A x = bdev[b] from the
// guess already in xacc[b] (has_x0[b]) or from zero, into xacc[b].  Each side pauses, takes its
// FP64 true residual and resumes on its own; the others keep iterating meanwhile.
void device_cr32_batch(fftperm_ctx* c, int nrhs, const double* const* bdev, double* const* xacc, const bool* has_x0,
                       double tol, uint64_t maxit, MinresOut* outs, uint64_t* updates) {
    if (!(tol > 0.0)) throw Err(FFTPERM_EINVAL, "MinresSolver tolerance must be positive");
    if (nrhs < 1 || nrhs > kMaxRhs) throw Err(FFTPERM_EINVAL, "unsupported right-hand side count");
    const long L = c->L3();
    const long nb = (long)c->nb;
    const double inv_nb = 1.0 / (double)nb;
    c->ensure_vectors32();
    CrState h[kMaxRhs];
    double bnorm[kMaxRhs];
    bool active[kMaxRhs];
    long long seg_start[kMaxRhs];
    // the FP64 true residual r = b - A xacc of side b into c->r; returns ||r|| / ||b||
    auto true_residual = [&](int b, bool xzero) {
        if (!xzero) {
            device_apply<double>(c, xacc[b], c->tmp);
            launch_axpby(1.0, bdev[b], -1.0, c->tmp, c->r, L, c->s);
        } else {
            HIPCHK(hipMemcpyAsync(c->r, bdev[b], L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
        }
        return device_norm(c, c->r, L) / bnorm[b];
    };
    // side b continues from c->r: mean of r into its state, r into its FP32 slot
    auto load_residual = [&](int b) {
        launch_compact_sums(c->r, nb, c->part, nullptr, c->s);
        launch_finish_sum(c->part, kRedBlocks, 3, inv_nb, c->rst[b].mean_in, nullptr, c->s);
        launch_to_f32(c->r, c->r32 + b * c->vs32, L, c->s);
    };
    for (int b = 0; b < nrhs; ++b) {
        updates[b] = 0;
        outs[b] = MinresOut{};
        active[b] = false;
        seg_start[b] = 0;
        std::memset(&h[b], 0, sizeof(CrState));
        h[b].done = 1;  // sides that do not iterate stay skipped
        HIPCHK(hipMemcpyAsync(&c->rst[b], &h[b], sizeof(CrState), hipMemcpyHostToDevice, c->s));
        if (!has_x0[b]) HIPCHK(hipMemsetAsync(xacc[b], 0, L * sizeof(double), c->s));
        bnorm[b] = device_norm(c, bdev[b], L);
        if (bnorm[b] == 0.0) {
            outs[b].final_rel = 0.0;
            outs[b].converged = 1;
            continue;
        }
        const double rel = true_residual(b, !has_x0[b]);
        if (rel <= tol || maxit == 0) {
            outs[b].final_rel = rel;
            outs[b].converged = rel <= tol ? 1 : 0;
            continue;
        }
        HIPCHK(hipMemsetAsync(c->x32 + b * c->vs32, 0, L * sizeof(float), c->s));
        h[b].done = 0;
        h[b].rnorm_max = rel * bnorm[b];
        h[b].rhs_norm = bnorm[b];
        h[b].tol = tol;
        h[b].delta = c->reliable_delta;
        h[b].maxit = (long long)maxit;
        h[b].first = 1;
        outs[b].final_rel = rel;
        HIPCHK(hipMemcpyAsync(&c->rst[b], &h[b], sizeof(CrState), hipMemcpyHostToDevice, c->s));
        load_residual(b);
        active[b] = true;
    }
    const uint64_t chunk = (c->g.N >= (1L << 21)) ? 4 : 32;
    for (;;) {
        bool any = false;
        for (int b = 0; b < nrhs; ++b) any = any || active[b];
        if (!any) break;
        // iterate until an active side stops (pause for its reliable update, or its end)
        bool stop = false;
        long long seg_launch = 0;
        while (!stop) {
            const ChunkTiming timing(c);
            const bool graph = timing.graphs();
            uint64_t i = 0;
            if (graph)
                for (; i + 2 <= chunk; i += 2, seg_launch += 2)
                    HIPCHK(hipGraphLaunch(cr32_batch_graph(c, nrhs), c->s));
            for (; i < chunk; ++i) {
                c->cur_tag = seg_launch++;
                cr32_batch_iteration(c, nrhs);
            }
            c->cur_tag = -1;
            HIPCHK(hipMemcpyAsync(h, c->rst, nrhs * sizeof(CrState), hipMemcpyDeviceToHost, c->s));
            c->sync();
            for (int b = 0; b < nrhs; ++b) stop = stop || (active[b] && h[b].done);
        }
        long long d[kMaxRhs];
        for (int b = 0; b < nrhs; ++b) d[b] = active[b] ? h[b].iterations - seg_start[b] : 0;
        c->prof_resolve(0, d, nrhs);
        for (int b = 0; b < nrhs; ++b) seg_start[b] = h[b].iterations;
        for (int b = 0; b < nrhs; ++b) {
            if (!active[b] || !h[b].done) continue;
            launch_accumulate_f32(xacc[b], c->x32 + b * c->vs32, L, c->s);
            const double rel = true_residual(b, false);
            outs[b].final_rel = rel;
            ++updates[b];
            if (rel <= tol || !h[b].pause || h[b].breakdown || h[b].iterations >= h[b].maxit) {
                active[b] = false;  // its state stays done: every later kernel skips it
                outs[b].iterations = (uint64_t)h[b].iterations;
                outs[b].converged = rel <= tol ? 1 : 0;
                c->fp32_iters += (uint64_t)h[b].iterations;
                continue;
            }
            // continue from the true residual; p, Ap and r.A r of the last apply carry over
            const double rn = rel * bnorm[b];
            h[b].rr = rn * rn;
            h[b].rnorm_max = rn;
            h[b].done = 0;
            h[b].pause = 0;
            HIPCHK(hipMemcpyAsync(&c->rst[b], &h[b], sizeof(CrState), hipMemcpyHostToDevice, c->s));
            load_residual(b);
        }
    }
    c->prof_resolve(LLONG_MAX);
    c->sync();
}

bool use_tensor_batch(const fftperm_ctx* c) {
    return c->tensor_batch != 0 && c->precision == FFTPERM_PRECISION_MIXED &&
           c->mixed_solver == FFTPERM_MIXED_CR_FP32_VECTORS;
}

MinresOut run_minres(fftperm_ctx* c, const double* bdev, const double* x0dev, double tol, uint64_t maxit,
                     uint64_t* sweeps, std::vector<double>* mixed_hist) {
    *sweeps = 0;
    if (c->precision == FFTPERM_PRECISION_MIXED && c->mixed_solver == FFTPERM_MIXED_CR_RELIABLE)
        return device_cr_reliable<float, false>(c, bdev, x0dev, tol, maxit, sweeps, mixed_hist);
    if (c->precision == FFTPERM_PRECISION_MIXED && c->mixed_solver == FFTPERM_MIXED_CR_FP32_VECTORS)
        return device_cr_reliable<float, true>(c, bdev, x0dev, tol, maxit, sweeps, mixed_hist);
    if (c->precision == FFTPERM_PRECISION_MIXED && c->mixed_solver == FFTPERM_MIXED_CG_RELIABLE)
        return device_cg_reliable<float>(c, bdev, x0dev, tol, maxit, sweeps, mixed_hist);
    if (c->precision == FFTPERM_PRECISION_MIXED) return device_minres_mixed(c, bdev, x0dev, tol, maxit, sweeps, mixed_hist);
    return device_minres(c, bdev, x0dev, tol, maxit, 0.0, nullptr);
}

// VelocityRecovery::recover + average_over_domain (src/bignonnet/VelocityRecovery.cpp:12-47), FP64.
void device_recover(fftperm_ctx* c, const double grad[3], const double* xdev, double avg[3]) {
    const long nb = (long)c->nb;
    const double inv_nb = 1.0 / (double)nb;
    double* field = c->ensure_field();
    launch_compact_sums(xdev, nb, c->part, nullptr, c->s);
    launch_finish_sum(c->part, kRedBlocks, 3, inv_nb, c->small, nullptr, c->s);
    XSource sr = src_none(c);
    f0_coefficients(c, grad, sr.a_fluid, sr.a_support);
    sr.compact = xdev;
    sr.mean = c->small;
    XSink sk = sink_none(c);
    sk.field = field;
    sk.partials = c->part_x;
    c->chain<double>(sr, sk, nullptr);
    launch_finish_sum(c->part_x, (int)c->xparts, kXinvParts, inv_nb, c->small + 4, nullptr, c->s);
    double mb[3];
    HIPCHK(hipMemcpyAsync(mb, c->small + 4, 3 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    c->sync();
    launch_field_shift_sum(field, -mb[0], -mb[1], -mb[2], c->g.N, c->part, c->s);
    launch_finish_sum(c->part, kRedBlocks, 3, 1.0 / (double)c->g.N, c->small + 12, nullptr, c->s);
    HIPCHK(hipMemcpyAsync(avg, c->small + 12, 3 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    c->sync();
    c->prof_resolve(LLONG_MAX);
}

void fill_info(fftperm_solve_info* info, const MinresOut& o, uint64_t sweeps) {
    if (!info) return;
    info->iterations = o.iterations;
    info->final_rel_residual = o.final_rel;
    info->converged = o.converged;
    info->history_len = o.hist_len;
    info->refinement_sweeps = sweeps;
}

// Upload an initial guess where the solver of the context's precision mode expects it:
// FP64 MINRES copies it into x itself (tmp is free), the refinement accumulates in dx.
const double* stage_guess(fftperm_ctx* c, const double* x0host) {
    double* dst = (c->precision == FFTPERM_PRECISION_MIXED) ? c->dx : c->tmp;
    HIPCHK(hipMemcpyAsync(dst, x0host, c->L3() * sizeof(double), hipMemcpyHostToDevice, c->s));
    return dst;
}

void solve_direction_dev(fftperm_ctx* c, const double grad[3], double tol, uint64_t maxit, const double* x0host,
                         double avg[3], double* uhost, fftperm_solve_info* info) {
    c->require_support();
    c->ensure_vectors();
    device_rhs(c, grad, c->b);
    const double* x0dev = x0host ? stage_guess(c, x0host) : nullptr;
    uint64_t sweeps = 0;
    MinresOut o = run_minres(c, c->b, x0dev, tol, maxit, &sweeps, nullptr);
    device_recover(c, grad, c->x, avg);
    if (uhost) {
        HIPCHK(hipMemcpyAsync(uhost, c->field, 3 * c->g.N * sizeof(double), hipMemcpyDeviceToHost, c->s));
        c->sync();
    }
    fill_info(info, o, sweeps);
}

// First medium of a grid shape in this process: time the x-pass variants on its support map
// (scatter from and gather to compact vectors, as in the solver) and keep the fastest.
void tune_x_passes(fftperm_ctx* c) {
    if (!static_passes_ok()) return;
    const Geometry3& g = c->g;
    const long nlines = (long)g.ny * g.nz;
    auto untuned = [&](int prec) {
        return tuned_variant(kKindXfwd, prec, g.nx, nlines) < 0 || tuned_variant(kKindXinv, prec, g.nx, nlines) < 0;
    };
    const bool need64 = untuned(0);
    const bool need32 = c->precision == FFTPERM_PRECISION_MIXED && untuned(1);
    if (!need64 && !need32) return;
    c->ensure_vectors();
    launch_fill_pattern(c->x, c->L3(), c->s);
    XSource sr = src_none(c);  // f0 = 0: the source of every MINRES apply
    sr.compact = c->x;
    XSink sk = sink_none(c);
    sk.compact = c->tmp;
    sk.partials = c->part_x;
    if (need64) {
        auto t = autotune_x<double>(sr, sk, c->S64, g, c->t64, c->s);
        c->tune64.insert(c->tune64.end(), t.begin(), t.end());
    }
    if (need32) {
        auto t = autotune_x<float>(sr, sk, c->S32, g, c->t32, c->s);
        c->tune32.insert(c->tune32.end(), t.begin(), t.end());
    }
    HIPCHK(hipMemsetAsync(c->x, 0, c->L3() * sizeof(double), c->s));
    c->sync();
}

// Second tuning stage, once per grid shape and precision in a process, on the first medium: the
// standalone timings above see each pass alone on a buffer it has just touched, while in the
// solver a pass reads what the previous pass left in the last-level cache.  So every launchable
// variant of each pass kind is timed again inside back-to-back five-sweep applies of the solver's
// own form (FP32 compact source and sink for the FP32 operator, Parseval partials on), one kind
// after the other with the others at their current choice, best of two interleaved rounds.
std::mutex g_chain_mu;
std::set<std::tuple<int, int, int, int>> g_chain_tuned;

// Milliseconds per five-sweep apply of the solver's form (FP32 compact source and sink for the FP32
// operator, FP64 for FP64; Parseval partials on) over reps back-to-back applies of a patterned
// input (launch_fill_pattern), after one untimed apply.
template <typename R>
float time_chain(fftperm_ctx* c, int reps) {
    const long L = c->L3();
    c->ensure_vectors();
    XSource sr = src_none(c);
    XSink sk = sink_none(c);
    sk.partials = c->part_x;
    if (sizeof(R) == 4) {
        c->ensure_vectors32();
        launch_fill_pattern(c->x32, L, c->s);
        sr.compact32 = c->x32;
        sk.compact32 = c->q32;
    } else {
        launch_fill_pattern(c->x, L, c->s);
        sr.compact = c->x;
        sk.compact = c->tmp;
    }
    hipEvent_t a = c->take_event(), b = c->take_event();
    c->chain<R>(sr, sk, nullptr, c->part_z);
    HIPCHK(hipEventRecord(a, c->s));
    for (int rep = 0; rep < reps; ++rep) c->chain<R>(sr, sk, nullptr, c->part_z);
    HIPCHK(hipEventRecord(b, c->s));
    // the solvers start from a zero iterate
    if (sizeof(R) == 4)
        HIPCHK(hipMemsetAsync(c->x32, 0, L * sizeof(float), c->s));
    else
        HIPCHK(hipMemsetAsync(c->x, 0, L * sizeof(double), c->s));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    c->ev_free.push_back(a);
    c->ev_free.push_back(b);
    return ms / (float)std::max(1, reps);
}

template <typename R>
void tune_chain_prec(fftperm_ctx* c) {
    const Geometry3& g = c->g;
    const int prec = sizeof(R) == 4 ? 1 : 0;
    {
        std::lock_guard<std::mutex> lk(g_chain_mu);
        if (!g_chain_tuned.insert(std::make_tuple(g.nx, g.ny, g.nz, prec)).second) return;
    }
    struct Key {
        int kind, n;
        long count;
    };
    const Key keys[4] = {{kKindZgreen, g.nz, (long)g.ny * g.nxh},
                         {kKindCol, g.ny, (long)g.nz * g.nxh},
                         {kKindXfwd, g.nx, (long)g.ny * g.nz},
                         {kKindXinv, g.nx, (long)g.ny * g.nz}};
    constexpr int kRounds = 2, kReps = 2;
    for (const Key& k : keys) {
        const std::vector<int> cands = tuned_candidates(k.kind, prec, k.n, k.count);
        if (cands.size() < 2) continue;
        std::vector<float> best_of(cands.size(), 1e30f);
        for (int round = 0; round < kRounds; ++round) {
            for (size_t i = 0; i < cands.size(); ++i) {
                set_tuned_variant(k.kind, prec, k.n, k.count, cands[i]);
                best_of[i] = std::min(best_of[i], time_chain<R>(c, kReps));
            }
        }
        size_t best = 0;
        for (size_t i = 1; i < cands.size(); ++i)
            if (best_of[i] < best_of[best]) best = i;
        if (k.kind == kKindCol) {
            // column tiles within 1% of the fastest are a tie here (the timing noise of back-to-back
            // applies); the widest of them reads the longest contiguous row segments, which is what
            // counts when the spectrum comes from HBM rather than the last-level cache
            const float bound = best_of[best] * 1.01f;
            for (size_t i = 0; i < cands.size(); ++i)
                if (best_of[i] <= bound && col_variant_code(k.n, cands[i]) / 10 > col_variant_code(k.n, cands[best]) / 10)
                    best = i;
        }
        set_tuned_variant(k.kind, prec, k.n, k.count, cands[best]);
    }
    if (prec == 1 && plane_pass_shape_ok(g)) {
        // FP32 x and y passes fused over z planes against the separate passes just tuned
        const long pcount = (long)g.ny * g.nz;
        float best_of[3] = {1e30f, 1e30f, 1e30f};
        for (int round = 0; round < kRounds; ++round) {
            for (int v = 0; v < 3; ++v) {
                set_tuned_variant(kKindPlane, prec, g.nx, pcount, v);
                best_of[v] = std::min(best_of[v], time_chain<R>(c, kReps));
            }
        }
        int best = 0;
        for (int v = 1; v < 3; ++v)
            if (best_of[v] < best_of[best]) best = v;
        set_tuned_variant(kKindPlane, prec, g.nx, pcount, best);
    }
    c->sync();
}

void tune_chain(fftperm_ctx* c) {
    if (!static_passes_ok() || c->nb == 0) return;
    const bool prof = c->prof;
    c->prof = false;  // the tuning applies are not the solver's
    tune_chain_prec<double>(c);
    if (c->precision == FFTPERM_PRECISION_MIXED) tune_chain_prec<float>(c);
    c->prof = prof;
}

void set_medium_impl(fftperm_ctx* c, const uint8_t* mask, bool host, int mode, uint64_t* nb_out) {
    if (mode != FFTPERM_SUPPORT_INTERFACE && mode != FFTPERM_SUPPORT_FULLSOLID)
        throw Err(FFTPERM_EINVAL, "unknown support mode");
    if (!mask) throw Err(FFTPERM_EINVAL, "null mask");
    const long N = c->g.N;
    if (!c->mask) c->mask = dalloc<uint8_t>(N);
    if (!c->code) c->code = dalloc<int32_t>(N);
    if (!c->active) c->active = dalloc<int64_t>(N);
    if (!c->scan_scratch) c->scan_scratch = dalloc<unsigned char>(scan_scratch_bytes(N));
    if (!c->counts) c->counts = dalloc<unsigned long long>(2);
    HIPCHK(hipMemcpyAsync(c->mask, mask, N, host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipMemsetAsync(c->counts, 0, 2 * sizeof(unsigned long long), c->s));
    launch_support(c->mask, c->g.nx, c->g.ny, c->g.nz, mode == FFTPERM_SUPPORT_INTERFACE ? 1 : 0, c->code, c->active,
                   c->counts, c->scan_scratch, c->s);
    const long nlines = (long)c->g.ny * c->g.nz;
    if (!c->lbits) c->lbits = dalloc<uint32_t>((size_t)nlines * ((c->g.nx + 31) / 32));
    if (!c->lstart) c->lstart = dalloc<int32_t>((size_t)nlines);
    if (!c->lcount) c->lcount = dalloc<unsigned long long>(1);
    launch_line_bits(c->code, c->g.nx, nlines, c->lbits, c->lstart, c->lcount, c->s);
    unsigned long long cnt[2];
    HIPCHK(hipMemcpyAsync(cnt, c->counts, sizeof(cnt), hipMemcpyDeviceToHost, c->s));
    c->sync();
    c->nb = cnt[0];
    c->nf = cnt[1];
    c->has_medium = true;
    if (nb_out) *nb_out = c->nb;
    if (c->nb > 0) {
        tune_x_passes(c);
        tune_chain(c);
    }
}

// ------------------------------------------------------------ batches of media of one shape
// The uneven-size batch solves many media of a grid shape together: every Krylov iteration
// advances all of them, the y and z FFT passes of an iteration are one launch for all media (the
// z pass reads the Green tensor once per tile for all of them), the vector and scalar kernels are
// one launch each, and the iterations replay captured hipGraphs.  Each medium keeps its own support
// map, compact vectors, CR state, reliable updates and stopping test, so its result is the one its
// own solve_one_direction gives.

// A medium's support map, owned by the batch (the context's medium fields alias it while the
// batch works on that medium).
struct MediumSlot {
    uint8_t* mask = nullptr;
    int32_t* code = nullptr;
    int64_t* active = nullptr;
    uint32_t* lbits = nullptr;
    int32_t* lstart = nullptr;
    unsigned long long* lcount = nullptr;
    uint64_t nb = 0, nf = 0;
};

struct BatchBufs {
    int sides = 0;
    long vs = 0;       // FP32 vector slot stride (>= the largest 3 NB, 256-byte aligned)
    long Lmax = 0;     // FP64 vector slot stride
    long spec = 0;     // spectrum slot stride (3 Nc)
    float2* S = nullptr;
    float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *ap32 = nullptr, *q32 = nullptr;
    double *b = nullptr, *xacc = nullptr;
    double *part_x = nullptr, *part_z = nullptr, *part = nullptr;
    CrState* st = nullptr;
    long xstride = 0, zstride = 0, pstride = 0;
    ~BatchBufs() {
        dfree(S); dfree(x32); dfree(r32); dfree(p32); dfree(ap32); dfree(q32); dfree(b); dfree(xacc);
        dfree(part_x); dfree(part_z); dfree(part); dfree(st);
    }
};

// The context's medium fields point at slot m (non-owning).
void select_medium(fftperm_ctx* c, const MediumSlot& m) {
    c->mask = m.mask;
    c->code = m.code;
    c->active = m.active;
    c->lbits = m.lbits;
    c->lstart = m.lstart;
    c->lcount = m.lcount;
    c->nb = m.nb;
    c->nf = m.nf;
    c->has_medium = true;
}

void release_media(fftperm_ctx* c, std::vector<MediumSlot>& slots) {
    // the context's medium fields alias a slot: clear them before the slots go
    c->mask = nullptr; c->code = nullptr; c->active = nullptr;
    c->lbits = nullptr; c->lstart = nullptr; c->lcount = nullptr;
    c->nb = c->nf = 0;
    c->has_medium = false;
    for (MediumSlot& m : slots) {
        dfree(m.mask); dfree(m.code); dfree(m.active); dfree(m.lbits); dfree(m.lstart); dfree(m.lcount);
    }
    slots.clear();
}

// One lockstep CR iteration over the batch (the steps of cr32_batch_iteration with per-medium
// support maps).
void cr32_multi_iteration(fftperm_ctx* c, BatchBufs& B, const std::vector<MediumSlot>& slots) {
    const Geometry3& g = c->g;
    const int n = B.sides;
    for (int b = 0; b < n; ++b) {
        const MediumSlot& m = slots[b];
        XSource sr;
        std::memset(&sr, 0, sizeof(sr));
        sr.code = m.code;
        sr.lines = LineBits{m.lbits, m.lstart, (long)m.nb};
        sr.compact32 = B.r32 + b * B.vs;
        sr.mean = B.st[b].mean_in;
        launch_xfwd<float>(sr, B.S + b * B.spec, g, c->t32, &B.st[b].done, c->s);
    }
    launch_ypass_n<float>(B.S, -1, g, c->t32, n, &B.st[0].done, kCrDoneStride, c->s);
    const long zparts = launch_zgreen_batch<float>(B.S, c->G32, g, c->t32, B.part_z, B.zstride, n, &B.st[0].done,
                                                   kCrDoneStride, c->s);
    launch_ypass_n<float>(B.S, +1, g, c->t32, n, &B.st[0].done, kCrDoneStride, c->s);
    long xparts = 0;
    for (int b = 0; b < n; ++b) {
        const MediumSlot& m = slots[b];
        XSink sk;
        std::memset(&sk, 0, sizeof(sk));
        sk.code = m.code;
        sk.lines = LineBits{m.lbits, m.lstart, (long)m.nb};
        sk.compact32 = B.q32 + b * B.vs;
        sk.partials = B.part_x + b * B.xstride;
        xparts = launch_xinv<float>(B.S + b * B.spec, sk, g, c->t32, &B.st[b].done, c->s);
    }
    const long nbmax = B.vs / 3;
    launch_cr_beta_n(B.part_x, (int)xparts, B.xstride, B.part_z, (int)zparts, B.zstride, B.st, n, 0.0,
                     1.0 / (double)g.N, c->s);
    launch_cr32_direction_n(B.p32, B.ap32, B.r32, B.q32, B.st, nbmax, B.vs, B.part, B.pstride, n, c->s);
    launch_cr_alpha_n(B.part, B.pstride, B.st, n, c->s);
    launch_cr32_update_n(B.x32, B.r32, B.p32, B.ap32, B.st, nbmax, B.vs, B.part, B.pstride, n, c->s);
    launch_cr_check_n(B.part, B.pstride, B.st, nullptr, 0, n, 0.0, c->s);
}

// device_cr_reliable<float, true> for every medium of the batch at once (from zero guesses).
void device_cr32_multi(fftperm_ctx* c, BatchBufs& B, const std::vector<MediumSlot>& slots, double tol,
                       uint64_t maxit, std::vector<MinresOut>& outs) {
    const int n = B.sides;
    std::vector<CrState> h(n);
    std::vector<double> bnorm(n, 0.0);
    std::vector<char> active(n, 0);
    outs.assign(n, MinresOut{});
    auto L_of = [&](int b) { return 3L * (long)slots[b].nb; };
    auto true_residual = [&](int b, bool xzero) {
        select_medium(c, slots[b]);
        const long L = L_of(b);
        double* xb = B.xacc + b * B.Lmax;
        double* bb = B.b + b * B.Lmax;
        if (!xzero) {
            device_apply<double>(c, xb, c->tmp);
            launch_axpby(1.0, bb, -1.0, c->tmp, c->r, L, c->s);
        } else {
            HIPCHK(hipMemcpyAsync(c->r, bb, L * sizeof(double), hipMemcpyDeviceToDevice, c->s));
        }
        return device_norm(c, c->r, L) / bnorm[b];
    };
    auto load_residual = [&](int b) {
        const long nb = (long)slots[b].nb;
        launch_compact_sums(c->r, nb, c->part, nullptr, c->s);
        launch_finish_sum(c->part, kRedBlocks, 3, 1.0 / (double)nb, B.st[b].mean_in, nullptr, c->s);
        launch_to_f32(c->r, B.r32 + b * B.vs, 3 * nb, c->s);
    };
    for (int b = 0; b < n; ++b) {
        std::memset(&h[b], 0, sizeof(CrState));
        h[b].done = 1;
        h[b].nb = (long long)slots[b].nb;
        HIPCHK(hipMemcpyAsync(&B.st[b], &h[b], sizeof(CrState), hipMemcpyHostToDevice, c->s));
        if (slots[b].nb == 0) {
            outs[b].final_rel = 1.0;
            continue;
        }
        const long L = L_of(b);
        HIPCHK(hipMemsetAsync(B.xacc + b * B.Lmax, 0, L * sizeof(double), c->s));
        HIPCHK(hipMemsetAsync(B.x32 + b * B.vs, 0, B.vs * sizeof(float), c->s));
        bnorm[b] = device_norm(c, B.b + b * B.Lmax, L);
        if (bnorm[b] == 0.0) {
            outs[b].final_rel = 0.0;
            outs[b].converged = 1;
            continue;
        }
        const double rel = true_residual(b, true);
        if (rel <= tol || maxit == 0) {
            outs[b].final_rel = rel;
            outs[b].converged = rel <= tol ? 1 : 0;
            continue;
        }
        h[b].done = 0;
        h[b].rnorm_max = rel * bnorm[b];
        h[b].rhs_norm = bnorm[b];
        h[b].tol = tol;
        h[b].delta = c->reliable_delta;
        h[b].maxit = (long long)maxit;
        h[b].first = 1;
        outs[b].final_rel = rel;
        HIPCHK(hipMemcpyAsync(&B.st[b], &h[b], sizeof(CrState), hipMemcpyHostToDevice, c->s));
        load_residual(b);
        active[b] = 1;
    }
    // two iterations per graph replay; the capture bakes in this batch's buffers and support maps
    hipGraphExec_t exec = nullptr;
    if (c->use_graphs) {
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < 2; ++i) cr32_multi_iteration(c, B, slots);
        HIPCHK(hipStreamEndCapture(c->s, &graph));
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(e);
    }
    struct GraphGuard {
        hipGraphExec_t e;
        ~GraphGuard() {
            if (e) (void)hipGraphExecDestroy(e);
        }
    } guard{exec};
    const uint64_t chunk = 8;
    for (;;) {
        bool any = false;
        for (int b = 0; b < n; ++b) any = any || active[b];
        if (!any) break;
        bool stop = false;
        while (!stop) {
            if (exec) {
                for (uint64_t i = 0; i < chunk; i += 2) HIPCHK(hipGraphLaunch(exec, c->s));
            } else {
                for (uint64_t i = 0; i < chunk; ++i) cr32_multi_iteration(c, B, slots);
            }
            HIPCHK(hipMemcpyAsync(h.data(), B.st, n * sizeof(CrState), hipMemcpyDeviceToHost, c->s));
            c->sync();
            for (int b = 0; b < n; ++b) stop = stop || (active[b] && h[b].done);
        }
        for (int b = 0; b < n; ++b) {
            if (!active[b] || !h[b].done) continue;
            const long L = L_of(b);
            launch_accumulate_f32(B.xacc + b * B.Lmax, B.x32 + b * B.vs, L, c->s);
            const double rel = true_residual(b, false);
            outs[b].final_rel = rel;
            if (rel <= tol || !h[b].pause || h[b].breakdown || h[b].iterations >= h[b].maxit) {
                active[b] = 0;
                outs[b].iterations = (uint64_t)h[b].iterations;
                outs[b].converged = rel <= tol ? 1 : 0;
                c->fp32_iters += (uint64_t)h[b].iterations;
                continue;
            }
            const double rn = rel * bnorm[b];
            h[b].rr = rn * rn;
            h[b].rnorm_max = rn;
            h[b].done = 0;
            h[b].pause = 0;
            HIPCHK(hipMemcpyAsync(&B.st[b], &h[b], sizeof(CrState), hipMemcpyHostToDevice, c->s));
            load_residual(b);
        }
    }
    c->sync();
}

void solve_batch_impl(fftperm_ctx* c, uint32_t count, const uint8_t* const* masks, int support_mode,
                      const double grad[3], double tol, uint64_t maxit, double* k_out, fftperm_solve_info* info) {
    const Geometry3& g = c->g;
    std::vector<MediumSlot> slots;
    struct Release {
        fftperm_ctx* c;
        std::vector<MediumSlot>& s;
        ~Release() { release_media(c, s); }
    } release{c, slots};
    // 1. the support map of every medium, owned by its slot
    c->has_medium = false;
    dfree(c->mask); dfree(c->code); dfree(c->active); dfree(c->lbits); dfree(c->lstart); dfree(c->lcount);
    uint64_t nbmax = 0;
    for (uint32_t b = 0; b < count; ++b) {
        if (!masks[b]) throw Err(FFTPERM_EINVAL, "null mask");
        uint64_t nb = 0;
        set_medium_impl(c, masks[b], true, support_mode, &nb);
        MediumSlot m;
        m.mask = c->mask; m.code = c->code; m.active = c->active;
        m.lbits = c->lbits; m.lstart = c->lstart; m.lcount = c->lcount;
        m.nb = c->nb; m.nf = c->nf;
        c->mask = nullptr; c->code = nullptr; c->active = nullptr;
        c->lbits = nullptr; c->lstart = nullptr; c->lcount = nullptr;
        slots.push_back(m);
        nbmax = std::max(nbmax, nb);
    }
    // 2. batch buffers
    BatchBufs B;
    B.sides = (int)count;
    const long Lm = std::max(1L, 3L * (long)nbmax);
    B.vs = (Lm + 63) / 64 * 64;
    B.Lmax = B.vs;
    B.spec = 3 * g.Nc;
    B.S = dalloc<float2>((size_t)count * (size_t)B.spec);
    HIPCHK(hipMemsetAsync(B.S, 0, (size_t)count * (size_t)B.spec * sizeof(float2), c->s));
    const size_t nv = (size_t)count * (size_t)B.vs;
    B.x32 = dalloc<float>(nv); B.r32 = dalloc<float>(nv); B.p32 = dalloc<float>(nv); B.ap32 = dalloc<float>(nv);
    B.q32 = dalloc<float>(nv);
    // lanes past a medium's 3 NB stay zero: the vector kernels stop at each medium's length
    for (float* v : {B.x32, B.r32, B.p32, B.ap32, B.q32}) HIPCHK(hipMemsetAsync(v, 0, nv * sizeof(float), c->s));
    B.b = dalloc<double>((size_t)count * (size_t)B.Lmax);
    B.xacc = dalloc<double>((size_t)count * (size_t)B.Lmax);
    B.xstride = (long)kXinvParts * xinv_max_blocks(g);
    B.zstride = zgreen_max_blocks(g);
    B.pstride = 4L * kRedBlocks;
    B.part_x = dalloc<double>((size_t)count * (size_t)B.xstride);
    B.part_z = dalloc<double>((size_t)count * (size_t)B.zstride);
    B.part = dalloc<double>((size_t)count * (size_t)B.pstride);
    B.st = dalloc<CrState>(count);
    // 3. right-hand sides
    for (uint32_t b = 0; b < count; ++b) {
        if (slots[b].nb == 0) continue;
        select_medium(c, slots[b]);
        c->ensure_vectors();
        device_rhs(c, grad, B.b + b * B.Lmax);
    }
    // 4. the lockstep solve
    std::vector<MinresOut> outs;
    device_cr32_multi(c, B, slots, tol, maxit, outs);
    // 5. velocity recovery and domain average per medium: K column = +mu <v>
    for (uint32_t b = 0; b < count; ++b) {
        double avg[3] = {0.0, 0.0, 0.0};
        if (slots[b].nb > 0) {
            select_medium(c, slots[b]);
            device_recover(c, grad, B.xacc + b * B.Lmax, avg);
        }
        for (int r = 0; r < 3; ++r) k_out[3 * b + r] = c->mu * avg[r];
        if (info) {
            fill_info(&info[b], outs[b], 0);
            if (slots[b].nb == 0) info[b].converged = 0;
        }
    }
    c->sync();
}

}  // namespace

// ============================================================== C ABI
extern "C" {

const char* fftperm_last_error(void) { return g_last_error.c_str(); }

int fftperm_device_info(char* buf, size_t len) {
    return guarded([&] {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        hipDeviceProp_t p;
        HIPCHK(hipGetDeviceProperties(&p, dev));
        if (buf && len) std::snprintf(buf, len, "%s (%s), %d CUs, %.1f GiB", p.name, p.gcnArchName,
                                      p.multiProcessorCount, (double)p.totalGlobalMem / (1024.0 * 1024 * 1024));
    });
}

int fftperm_create(const fftperm_grid* grid, double mu, int p_radius, int precision, fftperm_ctx** out) {
    return guarded([&] {
        if (!grid || !out) throw Err(FFTPERM_EINVAL, "null argument");
        *out = nullptr;
        if (grid->nx == 0 || grid->ny == 0 || grid->nz == 0)
            throw Err(FFTPERM_EINVAL, "Grid dimensions must be positive");
        if (!(mu > 0.0)) throw Err(FFTPERM_EINVAL, "GreenOperator viscosity must be positive");
        if (precision != FFTPERM_PRECISION_FP64 && precision != FFTPERM_PRECISION_MIXED)
            throw Err(FFTPERM_EINVAL, "unknown precision mode");
        if (grid->nx > (1u << 20) || grid->ny > (1u << 20) || grid->nz > (1u << 20) ||
            grid->nx * grid->ny * grid->nz > (uint64_t)INT32_MAX)
            throw Err(FFTPERM_EUNSUPPORTED, "grid too large (N must fit in int32)");
        auto c = std::make_unique<fftperm_ctx>();
        HIPCHK(hipGetDevice(&c->device));
        {
            hipDeviceProp_t prop;
            HIPCHK(hipGetDeviceProperties(&prop, c->device));
            set_lds_limit(prop.sharedMemPerBlock);
            set_device_shape(prop.multiProcessorCount, prop.maxSharedMemoryPerMultiProcessor);
        }
        HIPCHK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
        Geometry3& g = c->g;
        g.nx = (int)grid->nx;
        g.ny = (int)grid->ny;
        g.nz = (int)grid->nz;
        g.nxh = g.nx / 2 + 1;
        g.N = (long)g.nx * g.ny * g.nz;
        g.Nh = (long)g.nxh * g.ny * g.nz;
        g.rp = spec_row_pitch(g.nxh);
        g.pz = spec_plane_pitch(g.ny, g.rp);
        g.Nc = (long)g.nz * g.pz;
        c->h[0] = grid->hx; c->h[1] = grid->hy; c->h[2] = grid->hz;
        // Grid3D: l = h * n unless given (src/core/Grid3D.hpp:19-21)
        c->L[0] = grid->lx > 0.0 ? grid->lx : grid->hx * (double)g.nx;
        c->L[1] = grid->ly > 0.0 ? grid->ly : grid->hy * (double)g.ny;
        c->L[2] = grid->lz > 0.0 ? grid->lz : grid->hz * (double)g.nz;
        c->mu = mu;
        c->P = p_radius < 0 ? 0 : p_radius;
        c->precision = precision;
        c->t64.px = make_plan(g.nx);
        c->t64.py = make_plan(g.ny);
        c->t64.pz = make_plan(g.nz);
        for (const LinePlan* lp : {&c->t64.px, &c->t64.py, &c->t64.pz}) {
            const PassCfg pc = cfg_zgreen(*lp, 1);
            if (pc.lds > lds_limit())
                throw Err(FFTPERM_EUNSUPPORTED, "grid edge " + std::to_string(lp->n) + " needs more LDS than the device has");
        }
        check_z_threads(c->t64.pz);
        c->t64.twx = make_twiddles<double>(g.nx);
        c->t64.twy = make_twiddles<double>(g.ny);
        c->t64.twz = make_twiddles<double>(g.nz);
        // padded half spectra: the padding stays zero (the passes read it as zero spectrum and never
        // write it), so the buffers are cleared once here
        c->S64 = dalloc<double2>(3 * (size_t)g.Nc);
        c->G64 = dalloc<double>(6 * (size_t)g.Nc);
        HIPCHK(hipMemsetAsync(c->S64, 0, 3 * (size_t)g.Nc * sizeof(double2), c->s));
        HIPCHK(hipMemsetAsync(c->G64, 0, 6 * (size_t)g.Nc * sizeof(double), c->s));
        if (precision == FFTPERM_PRECISION_MIXED) {
            c->t32.px = c->t64.px;
            c->t32.py = c->t64.py;
            c->t32.pz = c->t64.pz;
            c->t32.twx = make_twiddles<float>(g.nx);
            c->t32.twy = make_twiddles<float>(g.ny);
            c->t32.twz = make_twiddles<float>(g.nz);
            // kMaxRhs spectra: the multi-RHS solve transforms the three directions side by side
            c->S32 = dalloc<float2>(3 * (size_t)kMaxRhs * (size_t)g.Nc);
            c->G32 = dalloc<float>(6 * (size_t)g.Nc);
            HIPCHK(hipMemsetAsync(c->S32, 0, 3 * (size_t)kMaxRhs * (size_t)g.Nc * sizeof(float2), c->s));
            HIPCHK(hipMemsetAsync(c->G32, 0, 6 * (size_t)g.Nc * sizeof(float), c->s));
        }
        // per-axis Green tables
        std::vector<double> s2x, qx, s2y, qy, s2z, qz;
        green_axis_table(g.nx, c->L[0], c->P, s2x, qx);
        green_axis_table(g.ny, c->L[1], c->P, s2y, qy);
        green_axis_table(g.nz, c->L[2], c->P, s2z, qz);
        const size_t W = 2 * (size_t)c->P + 1;
        const size_t tot = (2 * (g.nx + g.ny + g.nz)) * W;
        c->gtab = dalloc<double>(tot);
        std::vector<double> host(tot);
        size_t off = 0;
        auto put = [&](const std::vector<double>& v) {
            std::copy(v.begin(), v.end(), host.begin() + off);
            const double* p = c->gtab + off;
            off += v.size();
            return p;
        };
        c->gt.s2x = put(s2x);
        c->gt.qx = put(qx);
        c->gt.s2y = put(s2y);
        c->gt.qy = put(qy);
        c->gt.s2z = put(s2z);
        c->gt.qz = put(qz);
        c->gt.P = c->P;
        c->gt.mu = mu;
        HIPCHK(hipMemcpy(c->gtab, host.data(), tot * sizeof(double), hipMemcpyHostToDevice));
        launch_green_sym(c->gt, g.nx, g.ny, g.nz, g.rp, g.pz, c->G64, c->G32, c->s);
        HIPCHK(hipGetLastError());
        // autotune the fused z pass tile width on this grid (cached per process)
        c->tune64 = autotune_zgreen<double>(c->S64, c->G64, g, c->t64, c->s);
        if (precision == FFTPERM_PRECISION_MIXED) c->tune32 = autotune_zgreen<float>(c->S32, c->G32, g, c->t32, c->s);
        HIPCHK(hipGetLastError());
        {
            auto t = autotune_col<double>(c->S64, g, c->t64, c->s);
            c->tune64.insert(c->tune64.end(), t.begin(), t.end());
        }
        if (precision == FFTPERM_PRECISION_MIXED) {
            auto t = autotune_col<float>(c->S32, g, c->t32, c->s);
            c->tune32.insert(c->tune32.end(), t.begin(), t.end());
        }
        HIPCHK(hipGetLastError());
        c->part_x = dalloc<double>((size_t)kMaxRhs * kXinvParts * xinv_max_blocks(g));
        c->part_z = dalloc<double>((size_t)kMaxRhs * zgreen_max_blocks(g));
        c->part = dalloc<double>((size_t)kMaxRhs * 4 * kRedBlocks);
        c->small = dalloc<double>(32);
        c->st = dalloc<MinresState>(1);
        c->cst = dalloc<CgState>(1);
        c->rst = dalloc<CrState>(kMaxRhs);
        c->sync();
        *out = c.release();
    });
}

int fftperm_destroy(fftperm_ctx* ctx) {
    return guarded([&] {
        if (ctx) {
            HIPCHK(hipSetDevice(ctx->device));
            delete ctx;
        }
    });
}

int fftperm_set_medium(fftperm_ctx* ctx, const uint8_t* fluid_mask, int support_mode, uint64_t* nb_out) {
    return guarded([&] {
        check_ctx(ctx);
        set_medium_impl(ctx, fluid_mask, true, support_mode, nb_out);
    });
}

int fftperm_set_medium_dev(fftperm_ctx* ctx, const uint8_t* fluid_mask_dev, int support_mode, uint64_t* nb_out) {
    return guarded([&] {
        check_ctx(ctx);
        set_medium_impl(ctx, fluid_mask_dev, false, support_mode, nb_out);
    });
}

int fftperm_porosity(fftperm_ctx* ctx, double* porosity) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->require_medium();
        if (!porosity) throw Err(FFTPERM_EINVAL, "null output");
        *porosity = (double)ctx->nf / (double)ctx->g.N;
    });
}

int fftperm_get_support(fftperm_ctx* ctx, uint64_t* active, int64_t* voxel_to_local) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->require_medium();
        if (active && ctx->nb)
            HIPCHK(hipMemcpyAsync(active, ctx->active, ctx->nb * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->s));
        if (voxel_to_local) {
            std::vector<int32_t> code(ctx->g.N);
            HIPCHK(hipMemcpyAsync(code.data(), ctx->code, ctx->g.N * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->s));
            ctx->sync();
            for (long i = 0; i < ctx->g.N; ++i) voxel_to_local[i] = code[i] >= 0 ? (int64_t)code[i] : -1;
        }
        ctx->sync();
    });
}

int fftperm_green_tensor(fftperm_ctx* ctx, double* t6) {
    return guarded([&] {
        check_ctx(ctx);
        if (!t6) throw Err(FFTPERM_EINVAL, "null output");
        if (!ctx->t6) ctx->t6 = dalloc<double>(6 * (size_t)ctx->g.N);
        launch_green_full(ctx->gt, ctx->g.nx, ctx->g.ny, ctx->g.nz, ctx->t6, ctx->s);
        HIPCHK(hipMemcpyAsync(t6, ctx->t6, 6 * ctx->g.N * sizeof(double), hipMemcpyDeviceToHost, ctx->s));
        ctx->sync();
    });
}

int fftperm_fft_forward(fftperm_ctx* ctx, const double* real3, double* spec3) {
    return guarded([&] {
        check_ctx(ctx);
        if (!real3 || !spec3) throw Err(FFTPERM_EINVAL, "null argument");
        const long N = ctx->g.N;
        double* f = ctx->ensure_field();
        if (!ctx->full_a) ctx->full_a = dalloc<double2>(3 * (size_t)N);
        HIPCHK(hipMemcpyAsync(f, real3, 3 * N * sizeof(double), hipMemcpyHostToDevice, ctx->s));
        XSource sr;
        std::memset(&sr, 0, sizeof(sr));
        sr.field = f;
        launch_xfwd<double>(sr, ctx->S64, ctx->g, ctx->t64, nullptr, ctx->s);
        launch_ypass<double>(ctx->S64, -1, ctx->g, ctx->t64, nullptr, ctx->s);
        launch_zpass<double>(ctx->S64, -1, ctx->g, ctx->t64, nullptr, ctx->s);
        launch_expand_full<double>(ctx->S64, ctx->full_a, ctx->g, ctx->s);
        HIPCHK(hipMemcpyAsync(spec3, ctx->full_a, 3 * N * sizeof(double2), hipMemcpyDeviceToHost, ctx->s));
        ctx->sync();
    });
}

int fftperm_fft_inverse(fftperm_ctx* ctx, const double* spec3, double* real3) {
    return guarded([&] {
        check_ctx(ctx);
        if (!real3 || !spec3) throw Err(FFTPERM_EINVAL, "null argument");
        const long N = ctx->g.N;
        double* f = ctx->ensure_field();
        if (!ctx->full_a) ctx->full_a = dalloc<double2>(3 * (size_t)N);
        if (!ctx->code) {  // the sink needs a support map; an all "-2" map is fine without a medium
            ctx->code = dalloc<int32_t>(N);
            HIPCHK(hipMemsetAsync(ctx->code, 0xFF, N * sizeof(int32_t), ctx->s));
        }
        HIPCHK(hipMemcpyAsync(ctx->full_a, spec3, 3 * N * sizeof(double2), hipMemcpyHostToDevice, ctx->s));
        launch_hermitian_half<double>(ctx->full_a, ctx->S64, ctx->g, ctx->s);
        launch_zpass<double>(ctx->S64, +1, ctx->g, ctx->t64, nullptr, ctx->s);
        launch_ypass<double>(ctx->S64, +1, ctx->g, ctx->t64, nullptr, ctx->s);
        XSink sk;
        std::memset(&sk, 0, sizeof(sk));
        sk.code = ctx->code;
        sk.field = f;
        launch_xinv<double>(ctx->S64, sk, ctx->g, ctx->t64, nullptr, ctx->s);
        HIPCHK(hipMemcpyAsync(real3, f, 3 * N * sizeof(double), hipMemcpyDeviceToHost, ctx->s));
        ctx->sync();
    });
}

int fftperm_spectral_green(fftperm_ctx* ctx, const double* t6, const double* force_spec, double* vel_spec) {
    return guarded([&] {
        check_ctx(ctx);
        if (!t6 || !force_spec || !vel_spec) throw Err(FFTPERM_EINVAL, "null argument");
        const long N = ctx->g.N;
        if (!ctx->t6) ctx->t6 = dalloc<double>(6 * (size_t)N);
        if (!ctx->full_a) ctx->full_a = dalloc<double2>(3 * (size_t)N);
        double2* fb = dalloc<double2>(3 * (size_t)N);
        try {
            HIPCHK(hipMemcpyAsync(ctx->t6, t6, 6 * N * sizeof(double), hipMemcpyHostToDevice, ctx->s));
            HIPCHK(hipMemcpyAsync(ctx->full_a, force_spec, 3 * N * sizeof(double2), hipMemcpyHostToDevice, ctx->s));
            launch_spectral_green_full(ctx->t6, ctx->full_a, fb, N, ctx->s);
            HIPCHK(hipMemcpyAsync(vel_spec, fb, 3 * N * sizeof(double2), hipMemcpyDeviceToHost, ctx->s));
            ctx->sync();
        } catch (...) {
            dfree(fb);
            throw;
        }
        dfree(fb);
    });
}

int fftperm_green_apply(fftperm_ctx* ctx, const double* fin, double* uout) {
    return guarded([&] {
        check_ctx(ctx);
        if (!fin || !uout) throw Err(FFTPERM_EINVAL, "null argument");
        const long N = ctx->g.N;
        double* f = ctx->ensure_field();
        if (!ctx->code) {
            ctx->code = dalloc<int32_t>(N);
            HIPCHK(hipMemsetAsync(ctx->code, 0xFF, N * sizeof(int32_t), ctx->s));
        }
        HIPCHK(hipMemcpyAsync(f, fin, 3 * N * sizeof(double), hipMemcpyHostToDevice, ctx->s));
        XSource sr;
        std::memset(&sr, 0, sizeof(sr));
        sr.field = f;
        XSink sk;
        std::memset(&sk, 0, sizeof(sk));
        sk.code = ctx->code;
        sk.field = f;  // the x-inverse pass reads only the spectrum, so in-place output is safe
        ctx->chain<double>(sr, sk, nullptr);
        HIPCHK(hipMemcpyAsync(uout, f, 3 * N * sizeof(double), hipMemcpyDeviceToHost, ctx->s));
        ctx->sync();
    });
}

int fftperm_background_force(fftperm_ctx* ctx, const double grad[3], double* f0) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->require_medium();
        if (!grad || !f0) throw Err(FFTPERM_EINVAL, "null argument");
        double* f = ctx->ensure_field();
        Coef3 af, as;
        f0_coefficients(ctx, grad, af.v, as.v);
        launch_background_force(ctx->code, ctx->g.N, af, as, f, ctx->s);
        HIPCHK(hipMemcpyAsync(f0, f, 3 * ctx->g.N * sizeof(double), hipMemcpyDeviceToHost, ctx->s));
        ctx->sync();
    });
}

int fftperm_rhs(fftperm_ctx* ctx, const double grad[3], double* b) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->require_support();
        if (!grad || !b) throw Err(FFTPERM_EINVAL, "null argument");
        ctx->ensure_vectors();
        device_rhs(ctx, grad, ctx->b);
        HIPCHK(hipMemcpyAsync(b, ctx->b, ctx->L3() * sizeof(double), hipMemcpyDeviceToHost, ctx->s));
        ctx->sync();
    });
}

int fftperm_operator_apply(fftperm_ctx* ctx, const double* x, double* y) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->require_support();
        if (!x || !y) throw Err(FFTPERM_EINVAL, "null argument");
        ctx->ensure_vectors();
        const long L = ctx->L3();
        HIPCHK(hipMemcpyAsync(ctx->r, x, L * sizeof(double), hipMemcpyHostToDevice, ctx->s));
        device_apply<double>(ctx, ctx->r, ctx->tmp);
        HIPCHK(hipMemcpyAsync(y, ctx->tmp, L * sizeof(double), hipMemcpyDeviceToHost, ctx->s));
        ctx->sync();
    });
}

int fftperm_operator_apply_fp32(fftperm_ctx* ctx, const double* x, double* y) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->require_support();
        if (!x || !y) throw Err(FFTPERM_EINVAL, "null argument");
        if (ctx->precision != FFTPERM_PRECISION_MIXED)
            throw Err(FFTPERM_EINVAL, "the FP32 operator exists in MIXED contexts only");
        ctx->ensure_vectors();
        const long L = ctx->L3();
        HIPCHK(hipMemcpyAsync(ctx->r, x, L * sizeof(double), hipMemcpyHostToDevice, ctx->s));
        device_apply<float>(ctx, ctx->r, ctx->tmp);
        HIPCHK(hipMemcpyAsync(y, ctx->tmp, L * sizeof(double), hipMemcpyDeviceToHost, ctx->s));
        ctx->sync();
    });
}

int fftperm_probe_copy(fftperm_ctx* ctx, int pattern, int reps, double* gbps) {
    return guarded([&] {
        check_ctx(ctx);
        if (!gbps) throw Err(FFTPERM_EINVAL, "null output");
        if (reps < 1) throw Err(FFTPERM_EINVAL, "reps must be >= 1");
        if (ctx->precision != FFTPERM_PRECISION_MIXED)
            throw Err(FFTPERM_EINVAL, "the copy probes use the FP32 spectrum slots of a MIXED context");
        const long bytes = probe_bytes(pattern, ctx->g);
        if (bytes < 0) throw Err(FFTPERM_EINVAL, "unknown probe pattern " + std::to_string(pattern));
        const float2* src = ctx->S32;
        float2* dst = ctx->S32 + 3 * ctx->g.Nc;
        launch_probe(pattern, src, dst, ctx->g, ctx->s);  // untimed first launch
        hipEvent_t a = ctx->take_event(), b = ctx->take_event();
        HIPCHK(hipEventRecord(a, ctx->s));
        for (int r = 0; r < reps; ++r) launch_probe(pattern, src, dst, ctx->g, ctx->s);
        HIPCHK(hipEventRecord(b, ctx->s));
        HIPCHK(hipEventSynchronize(b));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, a, b));
        ctx->ev_free.push_back(a);
        ctx->ev_free.push_back(b);
        *gbps = (double)bytes / ((double)ms * 1e-3 / reps) / 1e9;
    });
}

int fftperm_check_edges(int nx, int ny, int nz) {
    return guarded([&] {
        for (int n : {nx, ny, nz})
            if (n <= 0) throw Err(FFTPERM_EINVAL, "Grid dimensions must be positive");
        (void)make_plan(nx);
        (void)make_plan(ny);
        check_z_threads(make_plan(nz));
    });
}

int fftperm_operator_apply_fp32v(fftperm_ctx* ctx, const double* x, double* y) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->require_support();
        if (!x || !y) throw Err(FFTPERM_EINVAL, "null argument");
        if (ctx->precision != FFTPERM_PRECISION_MIXED)
            throw Err(FFTPERM_EINVAL, "the FP32 operator exists in MIXED contexts only");
        ctx->ensure_vectors();
        ctx->ensure_vectors32();
        const long nb = (long)ctx->nb;
        const long L = ctx->L3();
        const double inv_nb = 1.0 / (double)nb;
        HIPCHK(hipMemcpyAsync(ctx->r, x, L * sizeof(double), hipMemcpyHostToDevice, ctx->s));
        // the support mean of x in FP64, x itself rounded to an FP32 compact vector
        launch_compact_sums(ctx->r, nb, ctx->part, nullptr, ctx->s);
        launch_finish_sum(ctx->part, kRedBlocks, 3, inv_nb, ctx->small, nullptr, ctx->s);
        launch_to_f32(ctx->r, ctx->x32, L, ctx->s);
        XSource sr = src_none(ctx);
        sr.compact32 = ctx->x32;
        sr.mean = ctx->small;
        XSink sk = sink_none(ctx);
        sk.compact32 = ctx->q32;
        sk.partials = ctx->part_x;
        ctx->chain<float>(sr, sk, nullptr);
        launch_finish_sum(ctx->part_x, (int)ctx->xparts, kXinvParts, inv_nb, ctx->small + 4, nullptr, ctx->s);
        // y = the FP32 output widened, minus its support mean
        HIPCHK(hipMemsetAsync(ctx->tmp, 0, L * sizeof(double), ctx->s));
        launch_accumulate_f32(ctx->tmp, ctx->q32, L, ctx->s);
        launch_sub_mean(ctx->tmp, ctx->small + 4, ctx->tmp, nb, 1.0, nullptr, ctx->s);
        HIPCHK(hipMemcpyAsync(y, ctx->tmp, L * sizeof(double), hipMemcpyDeviceToHost, ctx->s));
        ctx->sync();
    });
}

int fftperm_minres(fftperm_ctx* ctx, const double* b, const double* x0, double tol, uint64_t maxit, double* x,
                   double* history, fftperm_solve_info* info) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->require_support();
        if (!b || !x) throw Err(FFTPERM_EINVAL, "null argument");
        ctx->ensure_vectors();
        const long L = ctx->L3();
        HIPCHK(hipMemcpyAsync(ctx->b, b, L * sizeof(double), hipMemcpyHostToDevice, ctx->s));
        const double* x0dev = x0 ? stage_guess(ctx, x0) : nullptr;
        uint64_t sweeps = 0;
        std::vector<double> mh;
        MinresOut o;
        o = run_minres(ctx, ctx->b, x0dev, tol, maxit, &sweeps, &mh);
        HIPCHK(hipMemcpyAsync(x, ctx->x, L * sizeof(double), hipMemcpyDeviceToHost, ctx->s));
        if (history) {
            if (ctx->precision == FFTPERM_PRECISION_MIXED) {
                std::copy(mh.begin(), mh.end(), history);
            } else {
                HIPCHK(hipMemcpyAsync(history, ctx->hist, o.hist_len * sizeof(double), hipMemcpyDeviceToHost,
                                      ctx->s));
            }
        }
        ctx->sync();
        fill_info(info, o, sweeps);
    });
}

int fftperm_set_refinement(fftperm_ctx* ctx, double fp32_below, double inner_floor, uint32_t max_sweeps) {
    return guarded([&] {
        check_ctx(ctx);
        if (!(fp32_below >= 0.0)) throw Err(FFTPERM_EINVAL, "fp32_below must be >= 0");
        if (!(inner_floor > 0.0 && inner_floor < 1.0)) throw Err(FFTPERM_EINVAL, "inner_floor must lie in (0, 1)");
        if (max_sweeps == 0) throw Err(FFTPERM_EINVAL, "max_sweeps must be positive");
        ctx->mixed_switch = fp32_below;
        ctx->mixed_floor = inner_floor;
        ctx->mixed_sweeps = max_sweeps;
    });
}

int fftperm_set_mixed_solver(fftperm_ctx* ctx, int solver, double reliable_delta) {
    return guarded([&] {
        check_ctx(ctx);
        if (solver != FFTPERM_MIXED_CG_RELIABLE && solver != FFTPERM_MIXED_MINRES_REFINED &&
            solver != FFTPERM_MIXED_CR_RELIABLE && solver != FFTPERM_MIXED_CR_FP32_VECTORS)
            throw Err(FFTPERM_EINVAL, "unknown mixed solver");
        const bool reliable = solver != FFTPERM_MIXED_MINRES_REFINED;
        if (reliable && !(reliable_delta > 0.0 && reliable_delta < 1.0))
            throw Err(FFTPERM_EINVAL, "reliable_delta must lie in (0, 1)");
        ctx->mixed_solver = solver;
        if (reliable) ctx->reliable_delta = reliable_delta;
    });
}

int fftperm_recover(fftperm_ctx* ctx, const double grad[3], const double* x, double* u, double avg[3]) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->require_support();
        if (!grad || !x || !avg) throw Err(FFTPERM_EINVAL, "null argument");
        ctx->ensure_vectors();
        HIPCHK(hipMemcpyAsync(ctx->tmp, x, ctx->L3() * sizeof(double), hipMemcpyHostToDevice, ctx->s));
        device_recover(ctx, grad, ctx->tmp, avg);
        if (u) {
            HIPCHK(hipMemcpyAsync(u, ctx->field, 3 * ctx->g.N * sizeof(double), hipMemcpyDeviceToHost, ctx->s));
            ctx->sync();
        }
    });
}

int fftperm_solve_direction(fftperm_ctx* ctx, const double grad[3], double tol, uint64_t maxit, const double* x0,
                            double avg[3], double* u, fftperm_solve_info* info) {
    return guarded([&] {
        check_ctx(ctx);
        if (!grad || !avg) throw Err(FFTPERM_EINVAL, "null argument");
        solve_direction_dev(ctx, grad, tol, maxit, x0, avg, u, info);
    });
}

// The three directions of a tensor from the right-hand sides in c->tb (guesses in c->txa where
// has_x0) by the multi-RHS CR, then recovery and K (row-major, K[r][col] = mu <v_r>^(col)).
void solve_tensor_batch(fftperm_ctx* c, const bool has_x0[3], double tol, uint64_t maxit, double K[9],
                        fftperm_solve_info* info) {
    const long L = c->L3();
    const double* bptr[3];
    double* xptr[3];
    for (int b = 0; b < 3; ++b) {
        bptr[b] = c->tb + b * L;
        xptr[b] = c->txa + b * L;
    }
    MinresOut outs[3];
    uint64_t updates[3];
    device_cr32_batch(c, 3, bptr, xptr, has_x0, tol, maxit, outs, updates);
    for (int col = 0; col < 3; ++col) {
        double g[3] = {0.0, 0.0, 0.0};
        g[col] = 1.0;
        double avg[3];
        device_recover(c, g, xptr[col], avg);
        fill_info(info ? &info[col] : nullptr, outs[col], updates[col]);
        for (int r = 0; r < 3; ++r) K[r * 3 + col] = c->mu * avg[r];
    }
}

int fftperm_solve_tensor(fftperm_ctx* ctx, double tol, uint64_t maxit, double K[9], fftperm_solve_info* info) {
    return guarded([&] {
        check_ctx(ctx);
        if (!K) throw Err(FFTPERM_EINVAL, "null argument");
        if (use_tensor_batch(ctx)) {
            ctx->require_support();
            ctx->ensure_vectors();
            ctx->ensure_tensor_vectors();
            for (int col = 0; col < 3; ++col) {
                double g[3] = {0.0, 0.0, 0.0};
                g[col] = 1.0;
                device_rhs(ctx, g, ctx->tb + col * ctx->L3());
            }
            const bool none[3] = {false, false, false};
            solve_tensor_batch(ctx, none, tol, maxit, K, info);
            return;
        }
        for (int col = 0; col < 3; ++col) {
            double g[3] = {0.0, 0.0, 0.0};
            g[col] = 1.0;
            double avg[3];
            solve_direction_dev(ctx, g, tol, maxit, nullptr, avg, nullptr, info ? &info[col] : nullptr);
            for (int r = 0; r < 3; ++r) {
                const double colK = -ctx->mu * avg[r];
                K[r * 3 + col] = -colK;
            }
        }
    });
}

int fftperm_solve_batch(fftperm_ctx* ctx, uint32_t count, const uint8_t* const* fluid_masks, int support_mode,
                        const double grad[3], double tol, uint64_t maxit, double* k_out, fftperm_solve_info* info) {
    return guarded([&] {
        check_ctx(ctx);
        if (!fluid_masks || !grad || !k_out) throw Err(FFTPERM_EINVAL, "null argument");
        if (count == 0 || count > FFTPERM_BATCH_MAX) throw Err(FFTPERM_EINVAL, "batch size out of range");
        if (!(tol > 0.0)) throw Err(FFTPERM_EINVAL, "MinresSolver tolerance must be positive");
        if (support_mode != FFTPERM_SUPPORT_INTERFACE && support_mode != FFTPERM_SUPPORT_FULLSOLID)
            throw Err(FFTPERM_EINVAL, "unknown support mode");
        if (ctx->precision != FFTPERM_PRECISION_MIXED || ctx->mixed_solver != FFTPERM_MIXED_CR_FP32_VECTORS)
            throw Err(FFTPERM_EINVAL, "batched solves need a MIXED context with the FP32-vector CR solver");
        solve_batch_impl(ctx, count, fluid_masks, support_mode, grad, tol, maxit, k_out, info);
    });
}

int fftperm_set_tensor_batch(fftperm_ctx* ctx, int mode) {
    return guarded([&] {
        check_ctx(ctx);
        if (mode < 0 || mode > 2) throw Err(FFTPERM_EINVAL, "tensor batch mode must be 0, 1 or 2");
        ctx->tensor_batch = mode;
    });
}

int fftperm_residual_norm(fftperm_ctx* ctx, const double* x, const double* b, double* out) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->require_support();
        if (!x || !b || !out) throw Err(FFTPERM_EINVAL, "null argument");
        ctx->ensure_vectors();
        const long L = ctx->L3();
        HIPCHK(hipMemcpyAsync(ctx->r, x, L * sizeof(double), hipMemcpyHostToDevice, ctx->s));
        HIPCHK(hipMemcpyAsync(ctx->dx, b, L * sizeof(double), hipMemcpyHostToDevice, ctx->s));
        device_apply<double>(ctx, ctx->r, ctx->tmp);
        launch_axpby(1.0, ctx->tmp, -1.0, ctx->dx, ctx->tmp, L, ctx->s);
        *out = device_norm(ctx, ctx->tmp, L);
    });
}

int fftperm_symmetry_gap(fftperm_ctx* ctx, const double* x, const double* y, double* out) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->require_support();
        if (!x || !y || !out) throw Err(FFTPERM_EINVAL, "null argument");
        ctx->ensure_vectors();
        const long L = ctx->L3();
        double* dxv = ctx->r;
        double* dyv = ctx->dx;
        double* ax = ctx->tmp;
        double* ay = ctx->z;
        HIPCHK(hipMemcpyAsync(dxv, x, L * sizeof(double), hipMemcpyHostToDevice, ctx->s));
        HIPCHK(hipMemcpyAsync(dyv, y, L * sizeof(double), hipMemcpyHostToDevice, ctx->s));
        device_apply<double>(ctx, dxv, ax);
        device_apply<double>(ctx, dyv, ay);
        auto dot = [&](const double* a, const double* b2) {
            launch_dot_partials(a, b2, L, ctx->part, ctx->s);
            launch_finish_sum(ctx->part, kRedBlocks, 1, 1.0, ctx->small + 16, nullptr, ctx->s);
            return ctx->host_scalar(ctx->small + 16);
        };
        const double xay = dot(dxv, ay);
        const double axy = dot(ax, dyv);
        const double nx = std::sqrt(std::max(0.0, dot(dxv, dxv)));
        const double ny = std::sqrt(std::max(0.0, dot(dyv, dyv)));
        const double denom = std::max({1.0, std::fabs(xay), std::fabs(axy), nx * ny});
        *out = std::fabs(xay - axy) / denom;
    });
}

int fftperm_pin_variant(fftperm_ctx* ctx, int kind, int precision, int code) {
    return guarded([&] {
        check_ctx(ctx);
        if (kind < kKindXfwd || kind > kKindPlane) throw Err(FFTPERM_EINVAL, "unknown pass kind");
        if (precision != FFTPERM_PRECISION_FP64 && precision != FFTPERM_PRECISION_MIXED)
            throw Err(FFTPERM_EINVAL, "unknown precision mode");
        const Geometry3& g = ctx->g;
        const int prec = precision == FFTPERM_PRECISION_MIXED ? 1 : 0;
        if (kind == kKindPlane) {
            const long pcount = (long)g.ny * g.nz;
            if (code == 0)
                clear_tuned_variant(kind, prec, g.nx, pcount);
            else if (code < 1 || code > 3)
                throw Err(FFTPERM_EINVAL, "unknown plane-pass code " + std::to_string(code));
            else if (prec != 1)
                throw Err(FFTPERM_EINVAL, "the fused plane passes are FP32 (precision MIXED)");
            else
                set_tuned_variant(kind, prec, g.nx, pcount, code - 1);
            return;
        }
        int n = 0;
        long count = 0;
        switch (kind) {
            case kKindXfwd:
            case kKindXinv: n = g.nx; count = (long)g.ny * g.nz; break;
            case kKindCol: n = g.ny; count = (long)g.nz * g.nxh; break;
            default: n = g.nz; count = (long)g.ny * g.nxh; break;
        }
        if (code == 0) {
            clear_tuned_variant(kind, prec, n, count);
            return;
        }
        const int v = variant_of_code(kind, n, code);
        if (v < 0) throw Err(FFTPERM_EINVAL, "unknown variant code " + std::to_string(code));
        set_tuned_variant(kind, prec, n, count, v);
    });
}

int fftperm_set_graphs(fftperm_ctx* ctx, int enable) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->use_graphs = enable != 0;
    });
}

int fftperm_time_apply(fftperm_ctx* ctx, int precision, int reps, double* ms_per_apply) {
    return guarded([&] {
        check_ctx(ctx);
        if (!ms_per_apply) throw Err(FFTPERM_EINVAL, "null output");
        if (reps < 1) throw Err(FFTPERM_EINVAL, "reps must be >= 1");
        if (precision != FFTPERM_PRECISION_FP64 && precision != FFTPERM_PRECISION_MIXED)
            throw Err(FFTPERM_EINVAL, "unknown precision mode");
        if (precision == FFTPERM_PRECISION_MIXED && ctx->precision != FFTPERM_PRECISION_MIXED)
            throw Err(FFTPERM_EINVAL, "the FP32 operator needs a MIXED context");
        ctx->require_support();
        const bool prof = ctx->prof;
        ctx->prof = false;
        const float ms = precision == FFTPERM_PRECISION_MIXED ? time_chain<float>(ctx, reps)
                                                               : time_chain<double>(ctx, reps);
        ctx->prof = prof;
        *ms_per_apply = (double)ms;
    });
}

int fftperm_profile(fftperm_ctx* ctx, int enable) {
    return guarded([&] {
        check_ctx(ctx);
        if (enable < 0 || enable > 2) throw Err(FFTPERM_EINVAL, "profile mode must be 0, 1 or 2");
        ctx->prof_resolve(LLONG_MAX);
        ctx->prof = enable != 0;
        ctx->prof_sampled = enable == 2;
        ctx->prof_chunk = 0;
        for (int i = 0; i < 2; ++i) {
            ctx->prof_z_ms[i] = ctx->prof_apply_ms[i] = 0.0;
            ctx->prof_z_n[i] = ctx->prof_apply_n[i] = 0;
            ctx->prof_rhs[i] = 0;
        }
    });
}

int fftperm_profile_read(fftperm_ctx* ctx, fftperm_profile_stats* out) {
    return guarded([&] {
        check_ctx(ctx);
        if (!out) throw Err(FFTPERM_EINVAL, "null output");
        ctx->prof_resolve(LLONG_MAX);
        for (int i = 0; i < 2; ++i) {
            out->zgreen_ms[i] = ctx->prof_z_ms[i];
            out->zgreen_launches[i] = (uint64_t)ctx->prof_z_n[i];
            out->apply_ms[i] = ctx->prof_apply_ms[i];
            out->apply_launches[i] = (uint64_t)ctx->prof_apply_n[i];
            out->zgreen_rhs[i] = (uint64_t)ctx->prof_rhs[i];
            out->apply_rhs[i] = (uint64_t)ctx->prof_rhs[i];
        }
        out->N = (uint64_t)ctx->g.N;
        out->Nh = (uint64_t)ctx->g.Nh;
        out->nb = ctx->nb;
        const Geometry3& g = ctx->g;
        for (int i = 0; i < 2; ++i) {
            out->zgreen_tile[i] = variant_code(kKindZgreen, i, g.nz, (long)g.ny * g.nxh);
            out->xfwd_tile[i] = variant_code(kKindXfwd, i, g.nx, (long)g.ny * g.nz);
            out->xinv_tile[i] = variant_code(kKindXinv, i, g.nx, (long)g.ny * g.nz);
            out->ypass_tile[i] = variant_code(kKindCol, i, g.ny, (long)g.nz * g.nxh);
        }
        if (plane_on(ctx)) out->xfwd_tile[1] = out->xinv_tile[1] = out->ypass_tile[1] = 1000;
    });
}

int fftperm_unet_create(uint64_t seed, fftperm_unet** out) {
    return guarded([&] {
        if (!out) throw Err(FFTPERM_EINVAL, "null argument");
        *out = reinterpret_cast<fftperm_unet*>(unet::create(seed));
    });
}

int fftperm_unet_destroy(fftperm_unet* net) {
    return guarded([&] { unet::destroy(reinterpret_cast<unet::Net*>(net)); });
}

int fftperm_unet_num_params(fftperm_unet* net, uint64_t* count) {
    return guarded([&] {
        if (!net || !count) throw Err(FFTPERM_EINVAL, "null argument");
        *count = (uint64_t)unet::num_params(reinterpret_cast<unet::Net*>(net));
    });
}

int fftperm_unet_get_params(fftperm_unet* net, float* params) {
    return guarded([&] {
        if (!net || !params) throw Err(FFTPERM_EINVAL, "null argument");
        unet::get_params(reinterpret_cast<unet::Net*>(net), params);
    });
}

int fftperm_unet_set_params(fftperm_unet* net, const float* params) {
    return guarded([&] {
        if (!net || !params) throw Err(FFTPERM_EINVAL, "null argument");
        unet::set_params(reinterpret_cast<unet::Net*>(net), params);
    });
}

namespace {
void check_unet_edge(uint32_t n) {
    if (n < 8 || n % 8 != 0) throw Err(FFTPERM_EUNSUPPORTED, "U-Net grid edge must be a positive multiple of 8");
}
}  // namespace

int fftperm_unet_loss_grad(fftperm_unet* net, const uint8_t* fluid_mask_dev, const float* target_dev, uint32_t n,
                           float lr, double* loss, float* grads) {
    return guarded([&] {
        if (!net || !fluid_mask_dev || !target_dev) throw Err(FFTPERM_EINVAL, "null argument");
        check_unet_edge(n);
        auto* u = reinterpret_cast<unet::Net*>(net);
        // the Adam step reads the gradient buffers without changing them, so one pass serves both
        const double l = unet::loss_grad(u, fluid_mask_dev, target_dev, (int)n, lr > 0.f, lr);
        if (grads) unet::get_grads(u, grads);
        if (loss) *loss = l;
    });
}

int fftperm_unet_predict(fftperm_unet* net, const uint8_t* fluid_mask_dev, uint32_t n, float* field_dev) {
    return guarded([&] {
        if (!net || !fluid_mask_dev || !field_dev) throw Err(FFTPERM_EINVAL, "null argument");
        check_unet_edge(n);
        auto* u = reinterpret_cast<unet::Net*>(net);
        unet::predict(u, fluid_mask_dev, (int)n);
        unet::pred_field(u, field_dev);
    });
}

int fftperm_solve_tensor_warm(fftperm_ctx* ctx, fftperm_unet* net, double tol, uint64_t maxit, double K[9],
                              fftperm_solve_info* info, int* used_guess) {
    return guarded([&] {
        check_ctx(ctx);
        if (!net || !K) throw Err(FFTPERM_EINVAL, "null argument");
        const Geometry3& g = ctx->g;
        if (g.nx != g.ny || g.nx != g.nz) throw Err(FFTPERM_EUNSUPPORTED, "the warm start needs a cubic grid");
        check_unet_edge((uint32_t)g.nx);
        ctx->require_support();
        ctx->ensure_vectors();
        auto* u = reinterpret_cast<unet::Net*>(net);
        const long L = ctx->L3();
        const bool batch = use_tensor_batch(ctx);
        if (batch) ctx->ensure_tensor_vectors();
        bool has_x0[3] = {false, false, false};
        // one forward pass predicts the force densities of the three load directions
        unet::predict(u, ctx->mask, g.nx);
        for (int col = 0; col < 3; ++col) {
            double grad[3] = {0.0, 0.0, 0.0};
            grad[col] = 1.0;
            double* bvec = batch ? ctx->tb + col * L : ctx->b;
            device_rhs(ctx, grad, bvec);
            ctx->sync();
            double* x0 = batch ? ctx->txa + col * L : (ctx->precision == FFTPERM_PRECISION_MIXED) ? ctx->dx : ctx->tmp;
            unet::gather_guess(u, col, ctx->active, (long)ctx->nb, x0);
            // the network fixes the shape of the density, the scale comes from the minimal-residual
            // step along it: x0 = alpha p, alpha = <b, A p> / ||A p||^2, so ||b - A x0|| <= ||b||;
            // the guess is kept only when it lowers the FP64 true residual below ||b||
            const double bnorm = device_norm(ctx, bvec, L);
            device_apply<double>(ctx, x0, ctx->r);
            auto dot = [&](const double* a, const double* bb) {
                launch_dot_partials(a, bb, L, ctx->part, ctx->s);
                launch_finish_sum(ctx->part, kRedBlocks, 1, 1.0, ctx->small + 16, nullptr, ctx->s);
                return ctx->host_scalar(ctx->small + 16);
            };
            const double bap = dot(bvec, ctx->r);
            const double apap = dot(ctx->r, ctx->r);
            const double alpha = (apap > 0.0 && std::isfinite(apap)) ? bap / apap : 0.0;
            launch_axpby(alpha, x0, 0.0, x0, x0, L, ctx->s);
            launch_axpby(1.0, bvec, -alpha, ctx->r, ctx->r, L, ctx->s);
            const double rnorm = device_norm(ctx, ctx->r, L);
            const bool use = alpha != 0.0 && std::isfinite(rnorm) && rnorm < bnorm;
            if (used_guess) used_guess[col] = use ? 1 : 0;
            has_x0[col] = use;
            if (batch) continue;  // the three directions are solved together below
            uint64_t sweeps = 0;
            MinresOut o = run_minres(ctx, ctx->b, use ? x0 : nullptr, tol, maxit, &sweeps, nullptr);
            double avg[3];
            device_recover(ctx, grad, ctx->x, avg);
            fill_info(info ? &info[col] : nullptr, o, sweeps);
            for (int r = 0; r < 3; ++r) K[r * 3 + col] = ctx->mu * avg[r];
        }
        if (batch) solve_tensor_batch(ctx, has_x0, tol, maxit, K, info);
    });
}

int fftperm_brinkman_solve(uint64_t nx, uint64_t ny, uint64_t nz, const double* alpha, const double* viscosity,
                           double relaxation, double tol, int max_iterations, const double force[3], int precision,
                           double U[3], double* u, fftperm_brinkman_info* info) {
    return guarded([&] {
        if (!alpha || !viscosity || !force || !U) throw Err(FFTPERM_EINVAL, "null argument");
        if (nx == 0 || ny == 0 || nz == 0) throw Err(FFTPERM_EINVAL, "Grid dimensions must be positive");
        if (precision != FFTPERM_PRECISION_FP64 && precision != FFTPERM_PRECISION_MIXED)
            throw Err(FFTPERM_EINVAL, "unknown precision mode");
        BrinkmanResult res;
        const int rc = brinkman_run((int)nx, (int)ny, (int)nz, alpha, viscosity, relaxation, tol, max_iterations,
                                    force, precision == FFTPERM_PRECISION_MIXED, U, u, &res);
        if (info) {
            info->iterations = res.iterations;
            info->residual = res.residual;
            info->converged = res.converged;
        }
        if (rc == 2)
            throw Err(FFTPERM_ERUNTIME, "Brinkman solver became unstable: non-finite residual at iteration " +
                                            std::to_string(res.iterations));
        if (rc == 3) throw Err(FFTPERM_ERUNTIME, "Brinkman solver produced non-finite macroscopic velocity");
        if (rc != 0) throw Err(FFTPERM_EDEVICE, res.message);
    });
}

}  // extern "C"

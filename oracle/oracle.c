/*
 * oracle.c -- CPU FP64 restatement of the FFT permeability hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker, never the product.
 *
 * Every function follows the reference file:line cited above it, in the same
 * order of floating-point operations, with serial reductions (the order of the
 * reference when it is built without OpenMP).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *  - support/interface voxels: bit-exact against the reference code built into
 *    oracle/_ref, against brute force on tiny masks, and against the hand
 *    values of tests/test_interface_voxels.cpp and test_support_mapping.cpp.
 *  - Green tensor: bit-exact against the reference's GreenOperator built into
 *    oracle/_ref; invariants of tests/test_projector_green.cpp; closed-form
 *    single-term values computed independently in numpy.
 *  - 3D FFT: against numpy.fft.fftn / ifftn and a brute-force O(N^2) DFT.
 *  - Green apply with the FFT: against numpy (reference Green tensor from
 *    oracle/_ref, numpy.fft full C2C, real part), and the dense real-space
 *    convolution matrix on 4^3.
 *  - f0, RHS, operator, MINRES, recovery, K assembly, residual/symmetry
 *    diagnostics: number for number against the reference code in identity
 *    transform mode (oracle/_ref); MINRES also against scipy.sparse.linalg.minres
 *    and pinv(A) b on dense matrices.
 *  - Brinkman: number for number against the reference's solveBrinkman built
 *    into oracle/_ref; slit-channel (plane Poiseuille) physics check.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

static __thread char g_err[512];

static void set_err(const char* msg) {
    snprintf(g_err, sizeof(g_err), "%s", msg);
}

const char* orc_last_error(void) { return g_err; }

/* ------------------------------------------------------------------------- */
/* 1D mixed-radix Stockham FFT (what FFTW3 supplies to the reference;        */
/* src/fft/FFTWBackend.cpp:123-139 plans C2C over (nz,ny,nx)).  Pinned by the */
/* DFT definition: numpy.fft and brute force.                                */
/* ------------------------------------------------------------------------- */

typedef struct {
    size_t n;
    int nf;
    size_t fac[64];
    double* tw; /* tw[2k] = cos(2 pi k/n), tw[2k+1] = -sin(2 pi k/n): forward root e^{-2 pi i k/n} */
} fft1d;

static int fft1d_init(fft1d* p, size_t n) {
    memset(p, 0, sizeof(*p));
    p->n = n;
    size_t m = n;
    while (m % 4 == 0 && m > 1) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    for (size_t f = 3; m > 1; f += 2) {
        while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
        if (f * f > m && m > 1) { p->fac[p->nf++] = m; m = 1; }
    }
    p->tw = (double*)malloc(2 * n * sizeof(double));
    if (!p->tw) return -1;
    for (size_t k = 0; k < n; ++k) {
        long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        p->tw[2 * k] = (double)cosl(a);
        p->tw[2 * k + 1] = -(double)sinl(a);
    }
    return 0;
}

static void fft1d_free(fft1d* p) {
    free(p->tw);
    p->tw = NULL;
}

/* In-place transform of one contiguous complex line (interleaved), using work (2n doubles). */
static void fft1d_exec(const fft1d* p, double* x, double* work, double* vbuf, int sign) {
    const size_t n = p->n;
    if (n <= 1) return;
    double* a = x;
    double* b = work;
    const double s = (sign > 0) ? -1.0 : 1.0; /* multiplies the imaginary part of roots */
    size_t Ns = 1;
    for (int st = 0; st < p->nf; ++st) {
        const size_t R = p->fac[st];
        const size_t m = n / R;
        const size_t tstride = n / (Ns * R);
        for (size_t j = 0; j < m; ++j) {
            const size_t k = j % Ns;
            double* v = vbuf; /* 2R doubles in, 2R doubles out at vbuf + 2R */
            for (size_t r = 0; r < R; ++r) {
                const double xr = a[2 * (j + r * m)];
                const double xi = a[2 * (j + r * m) + 1];
                const size_t e = r * k * tstride;
                const double wr = p->tw[2 * e];
                const double wi = s * p->tw[2 * e + 1];
                v[2 * r] = xr * wr - xi * wi;
                v[2 * r + 1] = xr * wi + xi * wr;
            }
            double* o = vbuf + 2 * R;
            if (R == 2) {
                o[0] = v[0] + v[2];
                o[1] = v[1] + v[3];
                o[2] = v[0] - v[2];
                o[3] = v[1] - v[3];
            } else if (R == 4) {
                /* forward: W4 = -i ; backward: +i */
                const double t0r = v[0] + v[4], t0i = v[1] + v[5];
                const double t1r = v[0] - v[4], t1i = v[1] - v[5];
                const double t2r = v[2] + v[6], t2i = v[3] + v[7];
                const double t3r = v[2] - v[6], t3i = v[3] - v[7];
                /* (-i * s') * t3 where forward uses -i */
                double ur, ui;
                if (sign < 0) { ur = t3i; ui = -t3r; } else { ur = -t3i; ui = t3r; }
                o[0] = t0r + t2r; o[1] = t0i + t2i;
                o[2] = t1r + ur;  o[3] = t1i + ui;
                o[4] = t0r - t2r; o[5] = t0i - t2i;
                o[6] = t1r - ur;  o[7] = t1i - ui;
            } else {
                for (size_t q = 0; q < R; ++q) {
                    double accr = 0.0, acci = 0.0;
                    for (size_t r = 0; r < R; ++r) {
                        const size_t e = ((r * q) % R) * m;
                        const double wr = p->tw[2 * e];
                        const double wi = s * p->tw[2 * e + 1];
                        accr += v[2 * r] * wr - v[2 * r + 1] * wi;
                        acci += v[2 * r] * wi + v[2 * r + 1] * wr;
                    }
                    o[2 * q] = accr;
                    o[2 * q + 1] = acci;
                }
            }
            const size_t base = (j / Ns) * Ns * R + k;
            for (size_t r = 0; r < R; ++r) {
                b[2 * (base + r * Ns)] = o[2 * r];
                b[2 * (base + r * Ns) + 1] = o[2 * r + 1];
            }
        }
        double* t = a; a = b; b = t;
        Ns *= R;
    }
    if (a != x) memcpy(x, a, 2 * n * sizeof(double));
}

#define ORC_COLBLOCK 16

/* Transform all lines along one axis of an (nz,ny,nx) complex array. */
static void fft_axis(const fft1d* p, double* data, size_t nx, size_t ny, size_t nz, int axis, int sign) {
    const size_t n = p->n;
    if (n <= 1) return;
    size_t nlines, stride;
    if (axis == 0) { nlines = ny * nz; stride = 1; }
    else if (axis == 1) { nlines = nx * nz; stride = nx; }
    else { nlines = nx * ny; stride = nx * ny; }
    (void)stride;
    /* Lines are grouped in blocks of up to ORC_COLBLOCK consecutive x positions (for y/z axes). */
    const size_t nblocks = (axis == 0) ? nlines : ((nx + ORC_COLBLOCK - 1) / ORC_COLBLOCK) * (axis == 1 ? nz : ny);
    /* Lines are independent, so the thread count never changes the result; grids below 32^3 run
     * serially, where a parallel region costs more than the lines it splits. */
#ifdef _OPENMP
#pragma omp parallel if (nx * ny * nz >= 32768)
#endif
    {
        double* line = (double*)malloc(2 * n * ORC_COLBLOCK * sizeof(double));
        double* work = (double*)malloc(2 * n * sizeof(double));
        double* vbuf = (double*)malloc(4 * n * sizeof(double) + 64);
        double* col = (double*)malloc(2 * n * sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (long long ib = 0; ib < (long long)nblocks; ++ib) {
            if (axis == 0) {
                double* l = data + 2 * (size_t)ib * nx;
                fft1d_exec(p, l, work, vbuf, sign);
                continue;
            }
            const size_t nxb = (nx + ORC_COLBLOCK - 1) / ORC_COLBLOCK;
            const size_t xb = (size_t)ib % nxb;
            const size_t other = (size_t)ib / nxb; /* z for axis 1, y for axis 0... */
            const size_t x0 = xb * ORC_COLBLOCK;
            const size_t xw = (x0 + ORC_COLBLOCK <= nx) ? ORC_COLBLOCK : nx - x0;
            size_t base, estride;
            if (axis == 1) { base = other * ny * nx + x0; estride = nx; }
            else { base = other * nx + x0; estride = nx * ny; }
            /* gather */
            for (size_t e = 0; e < n; ++e) {
                const double* src = data + 2 * (base + e * estride);
                for (size_t c = 0; c < xw; ++c) {
                    line[2 * (c * n + e)] = src[2 * c];
                    line[2 * (c * n + e) + 1] = src[2 * c + 1];
                }
            }
            for (size_t c = 0; c < xw; ++c) {
                memcpy(col, line + 2 * c * n, 2 * n * sizeof(double));
                fft1d_exec(p, col, work, vbuf, sign);
                memcpy(line + 2 * c * n, col, 2 * n * sizeof(double));
            }
            for (size_t e = 0; e < n; ++e) {
                double* dst = data + 2 * (base + e * estride);
                for (size_t c = 0; c < xw; ++c) {
                    dst[2 * c] = line[2 * (c * n + e)];
                    dst[2 * c + 1] = line[2 * (c * n + e) + 1];
                }
            }
        }
        free(line);
        free(work);
        free(vbuf);
        free(col);
    }
}

int orc_fft3d(size_t nx, size_t ny, size_t nz, double* data, int sign) {
    fft1d px, py, pz;
    if (fft1d_init(&px, nx) || fft1d_init(&py, ny) || fft1d_init(&pz, nz)) {
        set_err("orc_fft3d: out of memory");
        return -1;
    }
    fft_axis(&px, data, nx, ny, nz, 0, sign);
    fft_axis(&py, data, nx, ny, nz, 1, sign);
    fft_axis(&pz, data, nx, ny, nz, 2, sign);
    fft1d_free(&px);
    fft1d_free(&py);
    fft1d_free(&pz);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Context                                                                   */
/* ------------------------------------------------------------------------- */

struct orc_ctx {
    size_t nx, ny, nz, N;
    double hx, hy, hz, lx, ly, lz;
    double mu;
    int p_radius;
    int transform;
    fft1d px, py, pz;
    uint8_t* mask;
    int has_medium;
    uint64_t* active;
    size_t nb;
    int64_t* v2l;
    size_t nfluid;
    double* G;        /* 6N reference Green tensor */
    double* spec_f;   /* 3 x 2N */
    double* spec_v;   /* 3 x 2N */
    double* full_a;   /* 3N */
    double* full_b;   /* 3N */
    double* cin;      /* L */
    double* cout;     /* L */
};

size_t orc_total_size(const orc_ctx* c) { return c->N; }

/* src/bignonnet/GreenOperator.cpp:13-19 */
static double sinc_ref(double x) {
    const double ax = fabs(x);
    if (ax < 1e-12) return 1.0;
    return sin(x) / x;
}

/* src/bignonnet/GreenOperator.cpp:21-23 */
static double q_component(long long m, double period_length) {
    return (2.0 * ORC_PI * (double)m) / period_length;
}

/* GreenOperator::precompute_energy_consistent_tensor, src/bignonnet/GreenOperator.cpp:95-178 */
static void precompute_green(orc_ctx* c) {
    const size_t N = c->N, nx = c->nx, ny = c->ny, nz = c->nz;
    const int P = c->p_radius;
    double* gxx_a = c->G + 0 * N;
    double* gxy_a = c->G + 1 * N;
    double* gxz_a = c->G + 2 * N;
    double* gyy_a = c->G + 3 * N;
    double* gyz_a = c->G + 4 * N;
    double* gzz_a = c->G + 5 * N;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (long long sidx = 0; sidx < (long long)N; ++sidx) {
        const size_t idx = (size_t)sidx;
        const size_t ki = idx % nx;
        const size_t kj = (idx / nx) % ny;
        const size_t kk = idx / (nx * ny);
        if (ki == 0 && kj == 0 && kk == 0) {
            gxx_a[idx] = gxy_a[idx] = gxz_a[idx] = 0.0;
            gyy_a[idx] = gyz_a[idx] = gzz_a[idx] = 0.0;
            continue;
        }
        double gxx = 0.0, gxy = 0.0, gxz = 0.0, gyy = 0.0, gyz = 0.0, gzz = 0.0;
        for (int px = -P; px <= P; ++px) {
            for (int py = -P; py <= P; ++py) {
                for (int pz = -P; pz <= P; ++pz) {
                    const long long mx = (long long)ki + (long long)px * (long long)nx;
                    const long long my = (long long)kj + (long long)py * (long long)ny;
                    const long long mz = (long long)kk + (long long)pz * (long long)nz;
                    const double sx = sinc_ref(ORC_PI * (double)mx / (double)nx);
                    const double sy = sinc_ref(ORC_PI * (double)my / (double)ny);
                    const double sz = sinc_ref(ORC_PI * (double)mz / (double)nz);
                    const double weight = (sx * sx) * (sy * sy) * (sz * sz);
                    if (weight == 0.0) continue;
                    const double qx = q_component(mx, c->lx);
                    const double qy = q_component(my, c->ly);
                    const double qz = q_component(mz, c->lz);
                    const double q2 = qx * qx + qy * qy + qz * qz;
                    if (q2 <= DBL_EPSILON) continue;
                    const double inv_mu_q2 = 1.0 / (c->mu * q2);
                    const double inv_q2 = 1.0 / q2;
                    const double p11 = 1.0 - qx * qx * inv_q2;
                    const double p12 = -qx * qy * inv_q2;
                    const double p13 = -qx * qz * inv_q2;
                    const double p22 = 1.0 - qy * qy * inv_q2;
                    const double p23 = -qy * qz * inv_q2;
                    const double p33 = 1.0 - qz * qz * inv_q2;
                    const double fac = weight * inv_mu_q2;
                    gxx += fac * p11;
                    gxy += fac * p12;
                    gxz += fac * p13;
                    gyy += fac * p22;
                    gyz += fac * p23;
                    gzz += fac * p33;
                }
            }
        }
        gxx_a[idx] = gxx;
        gxy_a[idx] = gxy;
        gxz_a[idx] = gxz;
        gyy_a[idx] = gyy;
        gyz_a[idx] = gyz;
        gzz_a[idx] = gzz;
    }
}

orc_ctx* orc_create(size_t nx, size_t ny, size_t nz, double hx, double hy, double hz,
                    double lx, double ly, double lz, double mu, int p_radius, int transform) {
    if (nx == 0 || ny == 0 || nz == 0) {
        set_err("Grid dimensions must be positive");
        return NULL;
    }
    if (!(mu > 0.0)) {
        set_err("GreenOperator viscosity must be positive");
        return NULL;
    }
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    if (!c) return NULL;
    c->nx = nx; c->ny = ny; c->nz = nz; c->N = nx * ny * nz;
    c->hx = hx; c->hy = hy; c->hz = hz;
    /* src/core/Grid3D.hpp:19-21 */
    c->lx = lx > 0.0 ? lx : hx * (double)nx;
    c->ly = ly > 0.0 ? ly : hy * (double)ny;
    c->lz = lz > 0.0 ? lz : hz * (double)nz;
    c->mu = mu;
    c->p_radius = p_radius < 0 ? 0 : p_radius;
    c->transform = transform;
    const size_t N = c->N;
    c->G = (double*)malloc(6 * N * sizeof(double));
    c->spec_f = (double*)malloc(6 * N * sizeof(double));
    c->spec_v = (double*)malloc(6 * N * sizeof(double));
    c->full_a = (double*)malloc(3 * N * sizeof(double));
    c->full_b = (double*)malloc(3 * N * sizeof(double));
    if (!c->G || !c->spec_f || !c->spec_v || !c->full_a || !c->full_b ||
        fft1d_init(&c->px, nx) || fft1d_init(&c->py, ny) || fft1d_init(&c->pz, nz)) {
        set_err("orc_create: out of memory");
        orc_destroy(c);
        return NULL;
    }
    precompute_green(c);
    return c;
}

void orc_destroy(orc_ctx* c) {
    if (!c) return;
    fft1d_free(&c->px);
    fft1d_free(&c->py);
    fft1d_free(&c->pz);
    free(c->mask);
    free(c->active);
    free(c->v2l);
    free(c->G);
    free(c->spec_f);
    free(c->spec_v);
    free(c->full_a);
    free(c->full_b);
    free(c->cin);
    free(c->cout);
    free(c);
}

/* InterfaceVoxels::compute_interface_only, src/geometry/InterfaceVoxels.cpp:7-43,
 * with Grid3D::periodic_neighbor (src/core/Grid3D.hpp:54-93).  Ascending by construction. */
static size_t interface_voxels(const orc_ctx* c, uint64_t* out) {
    const size_t nx = c->nx, ny = c->ny, nz = c->nz;
    size_t cnt = 0;
    for (size_t flat = 0; flat < c->N; ++flat) {
        if (c->mask[flat] != 0) continue; /* not solid */
        const long long i = (long long)(flat % nx);
        const long long j = (long long)((flat / nx) % ny);
        const long long k = (long long)(flat / (nx * ny));
        int iface = 0;
        for (int dk = -1; dk <= 1 && !iface; ++dk) {
            for (int dj = -1; dj <= 1 && !iface; ++dj) {
                for (int di = -1; di <= 1; ++di) {
                    if (di == 0 && dj == 0 && dk == 0) continue;
                    long long a = (i + di) % (long long)nx; if (a < 0) a += (long long)nx;
                    long long b = (j + dj) % (long long)ny; if (b < 0) b += (long long)ny;
                    long long d = (k + dk) % (long long)nz; if (d < 0) d += (long long)nz;
                    const size_t nidx = ((size_t)d * ny + (size_t)b) * nx + (size_t)a;
                    if (c->mask[nidx] != 0) { iface = 1; break; }
                }
            }
        }
        if (iface) out[cnt++] = flat;
    }
    return cnt;
}

/* BinaryMedium (src/geometry/BinaryMedium.cpp:7-12) + ForceSupport ctor (src/bignonnet/ForceSupport.cpp:9-25). */
int orc_set_medium(orc_ctx* c, const uint8_t* fluid_mask, int support_mode, uint64_t* nb_out) {
    const size_t N = c->N;
    free(c->mask); free(c->active); free(c->v2l); free(c->cin); free(c->cout);
    c->mask = (uint8_t*)malloc(N);
    c->active = (uint64_t*)malloc((N ? N : 1) * sizeof(uint64_t));
    c->v2l = (int64_t*)malloc(N * sizeof(int64_t));
    if (!c->mask || !c->active || !c->v2l) { set_err("orc_set_medium: out of memory"); return -1; }
    memcpy(c->mask, fluid_mask, N);
    size_t nf = 0;
    for (size_t i = 0; i < N; ++i) nf += (c->mask[i] != 0) ? 1 : 0;
    c->nfluid = nf;
    if (support_mode == ORC_SUPPORT_INTERFACE) {
        c->nb = interface_voxels(c, c->active);
    } else {
        size_t cnt = 0;
        for (size_t i = 0; i < N; ++i) if (c->mask[i] == 0) c->active[cnt++] = i;
        c->nb = cnt;
    }
    for (size_t i = 0; i < N; ++i) c->v2l[i] = -1;
    for (size_t l = 0; l < c->nb; ++l) c->v2l[c->active[l]] = (int64_t)l;
    const size_t L = 3 * c->nb;
    c->cin = (double*)malloc((L ? L : 1) * sizeof(double));
    c->cout = (double*)malloc((L ? L : 1) * sizeof(double));
    c->has_medium = 1;
    if (nb_out) *nb_out = c->nb;
    return 0;
}

int orc_get_support(const orc_ctx* c, uint64_t* active, int64_t* voxel_to_local) {
    if (!c->has_medium) { set_err("no medium"); return -1; }
    if (active) memcpy(active, c->active, c->nb * sizeof(uint64_t));
    if (voxel_to_local) memcpy(voxel_to_local, c->v2l, c->N * sizeof(int64_t));
    return 0;
}

void orc_green_tensor(const orc_ctx* c, double* t6) {
    memcpy(t6, c->G, 6 * c->N * sizeof(double));
}

/* IFFTBackend::forward as in src/fft/FFTWBackend.cpp:203-264 (per component: real ->
 * complex staging, C2C forward, copy out); identity mode follows its #else branch. */
static void backend_forward(orc_ctx* c, const double* real3, double* spec3) {
    const size_t N = c->N;
    for (int comp = 0; comp < 3; ++comp) {
        const double* src = real3 + comp * N;
        double* dst = spec3 + 2 * comp * N;
        for (size_t i = 0; i < N; ++i) { dst[2 * i] = src[i]; dst[2 * i + 1] = 0.0; }
        if (c->transform == ORC_TRANSFORM_FFT) {
            fft_axis(&c->px, dst, c->nx, c->ny, c->nz, 0, -1);
            fft_axis(&c->py, dst, c->nx, c->ny, c->nz, 1, -1);
            fft_axis(&c->pz, dst, c->nx, c->ny, c->nz, 2, -1);
        }
    }
}

/* IFFTBackend::inverse, src/fft/FFTWBackend.cpp:266-329: C2C backward then real part x (1/N);
 * identity mode: real part only (:374-380). */
static void backend_inverse(orc_ctx* c, double* spec3, double* real3) {
    const size_t N = c->N;
    const double inv_n = 1.0 / (double)N;
    for (int comp = 0; comp < 3; ++comp) {
        double* src = spec3 + 2 * comp * N;
        double* dst = real3 + comp * N;
        if (c->transform == ORC_TRANSFORM_FFT) {
            fft_axis(&c->px, src, c->nx, c->ny, c->nz, 0, +1);
            fft_axis(&c->py, src, c->nx, c->ny, c->nz, 1, +1);
            fft_axis(&c->pz, src, c->nx, c->ny, c->nz, 2, +1);
            for (size_t i = 0; i < N; ++i) dst[i] = src[2 * i] * inv_n;
        } else {
            for (size_t i = 0; i < N; ++i) dst[i] = src[2 * i];
        }
    }
}

/* GreenOperator::apply_spectral_green_tensor's CPU loop, src/bignonnet/GreenOperator.cpp:80-92:
 * per frequency, V = T F with T real symmetric (6 arrays of N: xx, xy, xz, yy, yz, zz) and F, V
 * complex 3-vectors (3 interleaved spectra of N).  A real T times a complex component is the
 * product with the real and imaginary parts separately (std::complex<double> * double). */
void orc_spectral_green(size_t N, const double* t6, const double* F, double* V) {
    const double* fxs = F;
    const double* fys = F + 2 * N;
    const double* fzs = F + 4 * N;
    double* vx = V;
    double* vy = V + 2 * N;
    double* vz = V + 4 * N;
    const double *txx = t6, *txy = t6 + N, *txz = t6 + 2 * N;
    const double *tyy = t6 + 3 * N, *tyz = t6 + 4 * N, *tzz = t6 + 5 * N;
    for (size_t i = 0; i < N; ++i) {
        for (int part = 0; part < 2; ++part) {
            const double fx = fxs[2 * i + part], fy = fys[2 * i + part], fz = fzs[2 * i + part];
            vx[2 * i + part] = txx[i] * fx + txy[i] * fy + txz[i] * fz;
            vy[2 * i + part] = txy[i] * fx + tyy[i] * fy + tyz[i] * fz;
            vz[2 * i + part] = txz[i] * fx + tyz[i] * fy + tzz[i] * fz;
        }
    }
}

/* GreenOperator::apply + apply_spectral_green_tensor, src/bignonnet/GreenOperator.cpp:54-93. */
void orc_green_apply(orc_ctx* c, const double* f, double* u) {
    backend_forward(c, f, c->spec_f);
    orc_spectral_green(c->N, c->G, c->spec_f, c->spec_v);
    backend_inverse(c, c->spec_v, u);
}

/* BackgroundForce::build, src/bignonnet/BackgroundForce.cpp:14-60. */
void orc_background_force(const orc_ctx* c, const double grad[3], double* f0) {
    const size_t N = c->N;
    memset(f0, 0, 3 * N * sizeof(double));
    const size_t nf = c->nfluid;
    const size_t nb = c->nb;
    if (nb == 0) return;
    const double c_fluid = -1.0;
    const double c_support = -(double)nf / (double)nb;
    for (size_t i = 0; i < N; ++i) {
        if (c->mask[i] != 0) {
            f0[i] += c_fluid * grad[0];
            f0[N + i] += c_fluid * grad[1];
            f0[2 * N + i] += c_fluid * grad[2];
        }
    }
    for (size_t l = 0; l < nb; ++l) {
        const size_t v = c->active[l];
        f0[v] += c_support * grad[0];
        f0[N + v] += c_support * grad[1];
        f0[2 * N + v] += c_support * grad[2];
    }
}

/* Reduction order of the reference's OpenMP reductions (MinresSolver.cpp:13-22 dot,
 * TrialForceField.cpp:42-64 mean_vector, both `#pragma omp parallel for reduction(+:...)`).
 * 0 or 1: serial order (a build without PERMEABILITY_USE_OPENMP, the default here).  T > 1:
 * the order of T threads under the static schedule -- contiguous blocks, the first n % T of
 * length n / T + 1, each summed in index order, the partial sums then added in thread order.
 * Used to measure how far two correct FP64 runs of the reference drift apart when only the
 * thread count changes. */
static int g_red_chunks = 0;

void orc_set_reduction_chunks(int chunks) { g_red_chunks = chunks > 1 ? chunks : 0; }

static void chunk_bounds(size_t n, int t, int T, size_t* lo, size_t* hi) {
    const size_t q = n / (size_t)T, r = n % (size_t)T, ut = (size_t)t;
    *lo = ut * q + (ut < r ? ut : r);
    *hi = *lo + q + (ut < r ? 1 : 0);
}

/* TrialForceField::mean_vector / subtract_mean, src/bignonnet/TrialForceField.cpp:42-80. */
static void subtract_mean(double* values, size_t nb) {
    if (nb == 0) return;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (g_red_chunks > 1) {
        for (int t = 0; t < g_red_chunks; ++t) {
            size_t lo, hi;
            chunk_bounds(nb, t, g_red_chunks, &lo, &hi);
            double px = 0.0, py = 0.0, pz = 0.0;
            for (size_t n = lo; n < hi; ++n) {
                px += values[3 * n + 0];
                py += values[3 * n + 1];
                pz += values[3 * n + 2];
            }
            sx += px; sy += py; sz += pz;
        }
    } else {
        for (size_t n = 0; n < nb; ++n) {
            sx += values[3 * n + 0];
            sy += values[3 * n + 1];
            sz += values[3 * n + 2];
        }
    }
    const double inv = 1.0 / (double)nb;
    const double m0 = sx * inv, m1 = sy * inv, m2 = sz * inv;
    for (size_t n = 0; n < nb; ++n) {
        values[3 * n + 0] -= m0;
        values[3 * n + 1] -= m1;
        values[3 * n + 2] -= m2;
    }
}

/* scatter_to_full / gather_from_full, src/bignonnet/TrialForceField.cpp:97-125. */
static void scatter(const orc_ctx* c, const double* compact, double* full) {
    const size_t N = c->N;
    for (size_t n = 0; n < c->nb; ++n) {
        const size_t v = c->active[n];
        full[v] = compact[3 * n + 0];
        full[N + v] = compact[3 * n + 1];
        full[2 * N + v] = compact[3 * n + 2];
    }
}

static void gather(const orc_ctx* c, const double* full, double* compact) {
    const size_t N = c->N;
    for (size_t n = 0; n < c->nb; ++n) {
        const size_t v = c->active[n];
        compact[3 * n + 0] = full[v];
        compact[3 * n + 1] = full[N + v];
        compact[3 * n + 2] = full[2 * N + v];
    }
}

/* VariationalRHS::build, src/bignonnet/VariationalRHS.cpp:15-31. */
void orc_rhs(orc_ctx* c, const double grad[3], double* b) {
    const size_t L = 3 * c->nb;
    orc_background_force(c, grad, c->full_a);
    orc_green_apply(c, c->full_a, c->full_b);
    gather(c, c->full_b, c->cout);
    subtract_mean(c->cout, c->nb);
    for (size_t i = 0; i < L; ++i) b[i] = -c->cout[i];
}

/* VariationalOperator::apply, src/bignonnet/VariationalOperator.cpp:19-41. */
void orc_operator_apply(orc_ctx* c, const double* x, double* y) {
    const size_t L = 3 * c->nb;
    memcpy(c->cin, x, L * sizeof(double));
    subtract_mean(c->cin, c->nb);
    memset(c->full_a, 0, 3 * c->N * sizeof(double));
    scatter(c, c->cin, c->full_a);
    orc_green_apply(c, c->full_a, c->full_b);
    gather(c, c->full_b, c->cout);
    subtract_mean(c->cout, c->nb);
    memcpy(y, c->cout, L * sizeof(double));
}

/* ------------------------------------------------------------------------- */
/* MINRES, src/solver/MinresSolver.cpp:13-218 (serial dot order).            */
/* ------------------------------------------------------------------------- */

typedef void (*orc_apply_fn)(void* op, const double* in, double* out);

static double vdot(const double* a, const double* b, size_t n) {
    double s = 0.0;
    if (g_red_chunks > 1) {
        for (int t = 0; t < g_red_chunks; ++t) {
            size_t lo, hi;
            chunk_bounds(n, t, g_red_chunks, &lo, &hi);
            double p = 0.0;
            for (size_t i = lo; i < hi; ++i) p += a[i] * b[i];
            s += p;
        }
        return s;
    }
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static double vnorm2(const double* v, size_t n) {
    const double d = vdot(v, v, n);
    return sqrt(d > 0.0 ? d : 0.0);
}

static int minres_generic(orc_apply_fn apply, void* op, size_t n, const double* rhs,
                          const double* x0, double tol, uint64_t maxit, double* xout,
                          double* history, orc_solve_info* info) {
    if (!(tol > 0.0)) { set_err("MinresSolver tolerance must be positive"); return -1; }
    double* x = (double*)calloc(n ? n : 1, sizeof(double));
    double* v_old = (double*)calloc(n ? n : 1, sizeof(double));
    double* v = (double*)calloc(n ? n : 1, sizeof(double));
    double* z = (double*)calloc(n ? n : 1, sizeof(double));
    double* w_prev = (double*)calloc(n ? n : 1, sizeof(double));
    double* w = (double*)calloc(n ? n : 1, sizeof(double));
    double* w_next = (double*)calloc(n ? n : 1, sizeof(double));
    if (!x || !v_old || !v || !z || !w_prev || !w || !w_next) { set_err("minres: out of memory"); return -1; }
    if (x0) memcpy(x, x0, n * sizeof(double));
    uint64_t hl = 0;
    orc_solve_info r = {0, 1.0, 0, 0};

    const double rhs_raw_norm = vnorm2(rhs, n);
    if (rhs_raw_norm == 0.0) {
        r.converged = 1; r.iterations = 0; r.final_rel_residual = 0.0;
        if (history) history[hl] = 0.0;
        hl++;
        goto done;
    }
    {
        const double rhs_norm = rhs_raw_norm > DBL_MIN ? rhs_raw_norm : DBL_MIN;
        apply(op, x, z);
        for (size_t i = 0; i < n; ++i) z[i] = rhs[i] - z[i];
        const double beta1 = vnorm2(z, n);
        const double rel0 = beta1 / rhs_norm;
        if (history) history[hl] = rel0;
        hl++;
        r.final_rel_residual = rel0;
        if (rel0 <= tol || beta1 == 0.0) {
            r.converged = 1; r.iterations = 0;
            goto done;
        }
        for (size_t i = 0; i < n; ++i) v[i] = z[i] / beta1;
        double beta = beta1;
        double cs_old = -1.0, sn_old = 0.0, cs = -1.0, sn = 0.0;
        double phi_bar = beta1;
        const double kBreakdownEps = 1e-14;
        double rel_res = rel0;
        for (uint64_t k = 0; k < maxit; ++k) {
            apply(op, v, z);
            const double alpha = vdot(v, z, n);
            for (size_t i = 0; i < n; ++i) z[i] -= alpha * v[i] + beta * v_old[i];
            const double beta_new = vnorm2(z, n);
            memcpy(v_old, v, n * sizeof(double));
            if (beta_new > kBreakdownEps) {
                for (size_t i = 0; i < n; ++i) v[i] = z[i] / beta_new;
            } else {
                memset(v, 0, n * sizeof(double));
            }
            const double eps = sn_old * beta;
            const double delta_raw = -cs_old * beta;
            const double delta = cs * delta_raw + sn * alpha;
            const double phi = sn * delta_raw - cs * alpha;
            const double rho = hypot(phi, beta_new);
            double cs_new = 1.0, sn_new = 0.0;
            if (rho > kBreakdownEps) { cs_new = phi / rho; sn_new = beta_new / rho; }
            if (rho > kBreakdownEps) {
                for (size_t i = 0; i < n; ++i) w_next[i] = (v_old[i] - eps * w_prev[i] - delta * w[i]) / rho;
            } else {
                memset(w_next, 0, n * sizeof(double));
            }
            {
                const double a = cs_new * phi_bar;
                for (size_t i = 0; i < n; ++i) x[i] += a * w_next[i];
            }
            phi_bar *= sn_new;
            memcpy(w_prev, w, n * sizeof(double));
            memcpy(w, w_next, n * sizeof(double));
            cs_old = cs; sn_old = sn; cs = cs_new; sn = sn_new; beta = beta_new;
            rel_res = fabs(phi_bar) / rhs_norm;
            if (history) history[hl] = rel_res;
            hl++;
            r.final_rel_residual = rel_res;
            r.iterations = k + 1;
            if (rel_res <= tol) { r.converged = 1; break; }
            if (beta_new < kBreakdownEps) { r.converged = (rel_res <= tol); break; }
        }
        if (!r.converged) r.converged = (r.final_rel_residual <= tol);
    }
done:
    r.history_len = hl;
    memcpy(xout, x, n * sizeof(double));
    if (info) *info = r;
    free(x); free(v_old); free(v); free(z); free(w_prev); free(w); free(w_next);
    return 0;
}

static void op_apply_thunk(void* op, const double* in, double* out) {
    orc_operator_apply((orc_ctx*)op, in, out);
}

int orc_minres(orc_ctx* c, const double* b, const double* x0, double tol, uint64_t maxit,
               double* x, double* history, orc_solve_info* info) {
    return minres_generic(op_apply_thunk, c, 3 * c->nb, b, x0, tol, maxit, x, history, info);
}

typedef struct { const double* A; size_t n; } dense_op;

static void dense_apply(void* op, const double* in, double* out) {
    const dense_op* d = (const dense_op*)op;
    for (size_t i = 0; i < d->n; ++i) {
        double s = 0.0;
        for (size_t j = 0; j < d->n; ++j) s += d->A[i * d->n + j] * in[j];
        out[i] = s;
    }
}

int orc_minres_dense(const double* A, uint64_t n, const double* b, const double* x0,
                     double tol, uint64_t maxit, double* x, double* history, orc_solve_info* info) {
    dense_op d = {A, (size_t)n};
    return minres_generic(dense_apply, &d, (size_t)n, b, x0, tol, maxit, x, history, info);
}

/* VelocityRecovery::recover + average_over_domain, src/bignonnet/VelocityRecovery.cpp:12-47
 * and src/core/Field3D.hpp:58-90. */
void orc_recover(orc_ctx* c, const double grad[3], const double* x, double* u_out, double avg[3]) {
    const size_t N = c->N, nb = c->nb;
    memcpy(c->cin, x, 3 * nb * sizeof(double));
    subtract_mean(c->cin, nb);
    double* f = c->full_a;
    orc_background_force(c, grad, f);
    double* fx = c->full_b;
    memset(fx, 0, 3 * N * sizeof(double));
    scatter(c, c->cin, fx);
    for (size_t i = 0; i < 3 * N; ++i) f[i] += 1.0 * fx[i];
    double* u = c->full_b;
    orc_green_apply(c, f, u);
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    if (nb > 0) {
        for (size_t l = 0; l < nb; ++l) {
            const size_t v = c->active[l];
            m0 += u[v]; m1 += u[N + v]; m2 += u[2 * N + v];
        }
        const double inv = 1.0 / (double)nb;
        m0 *= inv; m1 *= inv; m2 *= inv;
    }
    const double vp0 = -m0, vp1 = -m1, vp2 = -m2;
    for (size_t i = 0; i < N; ++i) {
        u[i] += vp0; u[N + i] += vp1; u[2 * N + i] += vp2;
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (size_t i = 0; i < N; ++i) { s0 += u[i]; s1 += u[N + i]; s2 += u[2 * N + i]; }
    const double inv = 1.0 / (double)N;
    avg[0] = s0 * inv; avg[1] = s1 * inv; avg[2] = s2 * inv;
    if (u_out) memcpy(u_out, u, 3 * N * sizeof(double));
}

/* solve_one_direction, src/app/PermeabilityRunner.cpp:173-242 (x0 empty). */
int orc_solve_direction(orc_ctx* c, const double grad[3], double tol, uint64_t maxit,
                        double avg[3], orc_solve_info* info) {
    if (!c->has_medium || c->nb == 0) { set_err("Support set is empty; cannot run variational solve"); return -1; }
    const size_t L = 3 * c->nb;
    double* b = (double*)malloc(L * sizeof(double));
    double* x = (double*)malloc(L * sizeof(double));
    if (!b || !x) { free(b); free(x); set_err("out of memory"); return -1; }
    orc_rhs(c, grad, b);
    int rc = orc_minres(c, b, NULL, tol, maxit, x, NULL, info);
    if (rc == 0) orc_recover(c, grad, x, NULL, avg);
    free(b); free(x);
    return rc;
}

/* PermeabilityRunner::run K assembly, src/app/PermeabilityRunner.cpp:757-777 with
 * PermeabilityTensor::set_column, src/post/PermeabilityTensor.cpp:11-18. */
int orc_solve_tensor(orc_ctx* c, double tol, uint64_t maxit, double K[9], orc_solve_info info[3]) {
    for (int col = 0; col < 3; ++col) {
        double g[3] = {0.0, 0.0, 0.0};
        g[col] = 1.0;
        double avg[3];
        int rc = orc_solve_direction(c, g, tol, maxit, avg, info ? &info[col] : NULL);
        if (rc) return rc;
        for (int r = 0; r < 3; ++r) {
            const double colK = -c->mu * avg[r];
            K[r * 3 + col] = -colK;
        }
    }
    return 0;
}

/* Diagnostics::residual_norm, src/post/Diagnostics.cpp:26-42. */
double orc_residual_norm(orc_ctx* c, const double* x, const double* b) {
    const size_t L = 3 * c->nb;
    double* ax = (double*)malloc((L ? L : 1) * sizeof(double));
    orc_operator_apply(c, x, ax);
    double s = 0.0;
    for (size_t i = 0; i < L; ++i) { const double r = ax[i] - b[i]; s += r * r; }
    free(ax);
    return sqrt(s);
}

/* Diagnostics::relative_symmetry_gap, src/post/Diagnostics.cpp:44-64. */
double orc_symmetry_gap(orc_ctx* c, const double* x, const double* y) {
    const size_t L = 3 * c->nb;
    double* ax = (double*)malloc((L ? L : 1) * sizeof(double));
    double* ay = (double*)malloc((L ? L : 1) * sizeof(double));
    orc_operator_apply(c, x, ax);
    orc_operator_apply(c, y, ay);
    double xay = 0.0, axy = 0.0;
    for (size_t i = 0; i < L; ++i) { xay += x[i] * ay[i]; axy += ax[i] * y[i]; }
    const double nxn = vnorm2(x, L), nyn = vnorm2(y, L);
    double denom = 1.0;
    if (fabs(xay) > denom) denom = fabs(xay);
    if (fabs(axy) > denom) denom = fabs(axy);
    if (nxn * nyn > denom) denom = nxn * nyn;
    free(ax); free(ay);
    return fabs(xay - axy) / denom;
}

/* ------------------------------------------------------------------------- */
/* Brinkman penalisation, src/brinkman/FluidBrinkman.cpp:24-33, 233-258, 282-507. */
/* ------------------------------------------------------------------------- */

static size_t wrap_m(size_t i, size_t n) { return (i + n - 1U) % n; }
static size_t wrap_p(size_t i, size_t n) { return (i + 1U) % n; }

int orc_brinkman_solve(size_t nx, size_t ny, size_t nz, const double* alpha,
                       const double* viscosity, double relaxation, double tol,
                       int max_iterations, const double force[3], double U[3], double* u_out,
                       orc_brinkman_info* info) {
    const size_t n = nx * ny * nz;
    double* ux = (double*)calloc(n, sizeof(double));
    double* uy = (double*)calloc(n, sizeof(double));
    double* uz = (double*)calloc(n, sizeof(double));
    double* pressure = (double*)calloc(n, sizeof(double));
    double* sx = (double*)calloc(n, sizeof(double));
    double* sy = (double*)calloc(n, sizeof(double));
    double* sz = (double*)calloc(n, sizeof(double));
    double* divergence = (double*)calloc(n, sizeof(double));
    double* phi = (double*)calloc(n, sizeof(double));
    double* phin = (double*)calloc(n, sizeof(double));
    if (!ux || !uy || !uz || !pressure || !sx || !sy || !sz || !divergence || !phi || !phin) {
        set_err("brinkman: out of memory");
        return -1;
    }
    orc_brinkman_info r = {0, 0.0, 0};
    int rc = 0;
    const int pressureIterationsPerStep = 35;
#define IDX(x, y, z) ((z) * nx * ny + (y) * nx + (x))
    for (int iter = 0; iter < max_iterations; ++iter) {
        double divergenceL2 = 0.0;
        for (size_t z = 0; z < nz; ++z) {
            const size_t zm = wrap_m(z, nz), zp = wrap_p(z, nz);
            for (size_t y = 0; y < ny; ++y) {
                const size_t ym = wrap_m(y, ny), yp = wrap_p(y, ny);
                for (size_t x = 0; x < nx; ++x) {
                    const size_t xm = wrap_m(x, nx), xp = wrap_p(x, nx);
                    const size_t id = IDX(x, y, z);
                    const size_t iXm = IDX(xm, y, z), iXp = IDX(xp, y, z);
                    const size_t iYm = IDX(x, ym, z), iYp = IDX(x, yp, z);
                    const size_t iZm = IDX(x, y, zm), iZp = IDX(x, y, zp);
                    const double mu = viscosity[id];
                    const double al = alpha[id];
                    const double lapUx = ux[iXm] + ux[iXp] + ux[iYm] + ux[iYp] + ux[iZm] + ux[iZp];
                    const double lapUy = uy[iXm] + uy[iXp] + uy[iYm] + uy[iYp] + uy[iZm] + uy[iZp];
                    const double lapUz = uz[iXm] + uz[iXp] + uz[iYm] + uz[iYp] + uz[iZm] + uz[iZp];
                    const double dpdx = 0.5 * (pressure[iXp] - pressure[iXm]);
                    const double dpdy = 0.5 * (pressure[iYp] - pressure[iYm]);
                    const double dpdz = 0.5 * (pressure[iZp] - pressure[iZm]);
                    const double denom = al + 6.0 * mu;
                    const double jacUx = (force[0] - dpdx + mu * lapUx) / denom;
                    const double jacUy = (force[1] - dpdy + mu * lapUy) / denom;
                    const double jacUz = (force[2] - dpdz + mu * lapUz) / denom;
                    sx[id] = relaxation * jacUx + (1.0 - relaxation) * ux[id];
                    sy[id] = relaxation * jacUy + (1.0 - relaxation) * uy[id];
                    sz[id] = relaxation * jacUz + (1.0 - relaxation) * uz[id];
                }
            }
        }
        for (size_t z = 0; z < nz; ++z) {
            const size_t zm = wrap_m(z, nz), zp = wrap_p(z, nz);
            for (size_t y = 0; y < ny; ++y) {
                const size_t ym = wrap_m(y, ny), yp = wrap_p(y, ny);
                for (size_t x = 0; x < nx; ++x) {
                    const size_t xm = wrap_m(x, nx), xp = wrap_p(x, nx);
                    const size_t id = IDX(x, y, z);
                    const double div = 0.5 * (sx[IDX(xp, y, z)] - sx[IDX(xm, y, z)]) +
                                       0.5 * (sy[IDX(x, yp, z)] - sy[IDX(x, ym, z)]) +
                                       0.5 * (sz[IDX(x, y, zp)] - sz[IDX(x, y, zm)]);
                    divergence[id] = div;
                    divergenceL2 += div * div;
                }
            }
        }
        memset(phi, 0, n * sizeof(double));
        for (int k = 0; k < pressureIterationsPerStep; ++k) {
            for (size_t z = 0; z < nz; ++z) {
                const size_t zm = wrap_m(z, nz), zp = wrap_p(z, nz);
                for (size_t y = 0; y < ny; ++y) {
                    const size_t ym = wrap_m(y, ny), yp = wrap_p(y, ny);
                    for (size_t x = 0; x < nx; ++x) {
                        const size_t xm = wrap_m(x, nx), xp = wrap_p(x, nx);
                        const size_t id = IDX(x, y, z);
                        const double neighbor = phi[IDX(xm, y, z)] + phi[IDX(xp, y, z)] + phi[IDX(x, ym, z)] +
                                                phi[IDX(x, yp, z)] + phi[IDX(x, y, zm)] + phi[IDX(x, y, zp)];
                        phin[id] = (neighbor - divergence[id]) / 6.0;
                    }
                }
            }
            double* t = phi; phi = phin; phin = t;
        }
        for (size_t z = 0; z < nz; ++z) {
            const size_t zm = wrap_m(z, nz), zp = wrap_p(z, nz);
            for (size_t y = 0; y < ny; ++y) {
                const size_t ym = wrap_m(y, ny), yp = wrap_p(y, ny);
                for (size_t x = 0; x < nx; ++x) {
                    const size_t xm = wrap_m(x, nx), xp = wrap_p(x, nx);
                    const size_t id = IDX(x, y, z);
                    const double dphidx = 0.5 * (phi[IDX(xp, y, z)] - phi[IDX(xm, y, z)]);
                    const double dphidy = 0.5 * (phi[IDX(x, yp, z)] - phi[IDX(x, ym, z)]);
                    const double dphidz = 0.5 * (phi[IDX(x, y, zp)] - phi[IDX(x, y, zm)]);
                    ux[id] = sx[id] - dphidx;
                    uy[id] = sy[id] - dphidy;
                    uz[id] = sz[id] - dphidz;
                    pressure[id] += phi[id];
                }
            }
        }
        r.iterations = iter + 1;
        r.residual = sqrt(divergenceL2 / (double)n);
        if (!isfinite(r.residual)) {
            set_err("Brinkman solver became unstable: non-finite residual");
            rc = -2;
            break;
        }
        if (r.residual < tol) { r.converged = 1; break; }
    }
#undef IDX
    if (rc == 0) {
        double sumX = 0.0, sumY = 0.0, sumZ = 0.0;
        for (size_t i = 0; i < n; ++i) { sumX += ux[i]; sumY += uy[i]; sumZ += uz[i]; }
        U[0] = sumX / (double)n;
        U[1] = sumY / (double)n;
        U[2] = sumZ / (double)n;
        if (!isfinite(U[0]) || !isfinite(U[1]) || !isfinite(U[2])) {
            set_err("Brinkman solver produced non-finite macroscopic velocity");
            rc = -2;
        }
        if (u_out) {
            memcpy(u_out, ux, n * sizeof(double));
            memcpy(u_out + n, uy, n * sizeof(double));
            memcpy(u_out + 2 * n, uz, n * sizeof(double));
        }
    }
    if (info) *info = r;
    free(ux); free(uy); free(uz); free(pressure); free(sx); free(sy); free(sz);
    free(divergence); free(phi); free(phin);
    return rc;
}

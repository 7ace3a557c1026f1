// brinkman.hip -- Brinkman penalisation loop (src/brinkman/FluidBrinkman.cpp:282-507) on the GPU.
//
// One iteration = relaxed Jacobi momentum update, central divergence (+ its L2 sum),
// 35 Jacobi sweeps of the pressure-correction Poisson problem from phi = 0, projection
// and pressure update.  Fields are SoA FP64 arrays with the flat index
// z*nx*ny + y*nx + x (:24-26); neighbours wrap periodically (:28-33).  In mixed mode the
// Poisson sweeps run on FP32 copies of div and phi; velocity, pressure, divergence and
// the residual stay FP64.
#include "brinkman.hpp"
#include "common.hpp"

#include <cmath>
#include <vector>

namespace fftperm {

namespace {

constexpr int kBThreads = 256;
constexpr int kBRedBlocks = 2048;

struct Nbr {
    long id, xm, xp, ym, yp, zm, zp;
};

__device__ __forceinline__ Nbr neighbours(long i, int nx, int ny, int nz) {
    const int x = (int)(i % nx);
    const long r = i / nx;
    const int y = (int)(r % ny);
    const int z = (int)(r / ny);
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x + 1 == nx ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y + 1 == ny ? 0 : y + 1;
    const int zm = z == 0 ? nz - 1 : z - 1, zp = z + 1 == nz ? 0 : z + 1;
    const long pl = (long)nx * ny;
    const long row = (long)y * nx, pln = (long)z * pl;
    Nbr n;
    n.id = i;
    n.xm = pln + row + xm;
    n.xp = pln + row + xp;
    n.ym = pln + (long)ym * nx + x;
    n.yp = pln + (long)yp * nx + x;
    n.zm = (long)zm * pl + row + x;
    n.zp = (long)zp * pl + row + x;
    return n;
}

// :313-369
__global__ void k_momentum(const double* __restrict__ ux, const double* __restrict__ uy,
                           const double* __restrict__ uz, const double* __restrict__ p,
                           const double* __restrict__ alpha, const double* __restrict__ visc, double* __restrict__ sx,
                           double* __restrict__ sy, double* __restrict__ sz, double f0, double f1, double f2,
                           double relax, int nx, int ny, int nz, long N) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        const Nbr n = neighbours(i, nx, ny, nz);
        const double mu = visc[i];
        const double al = alpha[i];
        const double lapUx = ux[n.xm] + ux[n.xp] + ux[n.ym] + ux[n.yp] + ux[n.zm] + ux[n.zp];
        const double lapUy = uy[n.xm] + uy[n.xp] + uy[n.ym] + uy[n.yp] + uy[n.zm] + uy[n.zp];
        const double lapUz = uz[n.xm] + uz[n.xp] + uz[n.ym] + uz[n.yp] + uz[n.zm] + uz[n.zp];
        const double dpdx = 0.5 * (p[n.xp] - p[n.xm]);
        const double dpdy = 0.5 * (p[n.yp] - p[n.ym]);
        const double dpdz = 0.5 * (p[n.zp] - p[n.zm]);
        const double denom = al + 6.0 * mu;
        const double jx = (f0 - dpdx + mu * lapUx) / denom;
        const double jy = (f1 - dpdy + mu * lapUy) / denom;
        const double jz = (f2 - dpdz + mu * lapUz) / denom;
        sx[i] = relax * jx + (1.0 - relax) * ux[i];
        sy[i] = relax * jy + (1.0 - relax) * uy[i];
        sz[i] = relax * jz + (1.0 - relax) * uz[i];
    }
}

// :371-400
template <typename P>
__global__ void k_divergence(const double* __restrict__ sx, const double* __restrict__ sy,
                             const double* __restrict__ sz, double* __restrict__ div, P* __restrict__ divp, int nx,
                             int ny, int nz, long N, double* __restrict__ partials) {
    double acc[1] = {0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        const Nbr n = neighbours(i, nx, ny, nz);
        const double d = 0.5 * (sx[n.xp] - sx[n.xm]) + 0.5 * (sy[n.yp] - sy[n.ym]) + 0.5 * (sz[n.zp] - sz[n.zm]);
        div[i] = d;
        if (divp != nullptr) divp[i] = (P)d;
        acc[0] += d * d;
    }
    block_sum<1>(acc, partials + blockIdx.x);
}

// :402-433; `first` = the sweep from phi = 0
template <typename P>
__global__ void k_jacobi(const P* __restrict__ phi, const P* __restrict__ div, P* __restrict__ phin, int first,
                         int nx, int ny, int nz, long N) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        P nb = (P)0;
        if (!first) {
            const Nbr n = neighbours(i, nx, ny, nz);
            nb = phi[n.xm] + phi[n.xp] + phi[n.ym] + phi[n.yp] + phi[n.zm] + phi[n.zp];
        }
        phin[i] = (nb - div[i]) / (P)6.0;
    }
}

// :435-466
template <typename P>
__global__ void k_project(const double* __restrict__ sx, const double* __restrict__ sy,
                          const double* __restrict__ sz, const P* __restrict__ phi, double* __restrict__ ux,
                          double* __restrict__ uy, double* __restrict__ uz, double* __restrict__ p, int nx, int ny,
                          int nz, long N) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        const Nbr n = neighbours(i, nx, ny, nz);
        const double dx = 0.5 * ((double)phi[n.xp] - (double)phi[n.xm]);
        const double dy = 0.5 * ((double)phi[n.yp] - (double)phi[n.ym]);
        const double dz = 0.5 * ((double)phi[n.zp] - (double)phi[n.zm]);
        ux[i] = sx[i] - dx;
        uy[i] = sy[i] - dy;
        uz[i] = sz[i] - dz;
        p[i] += (double)phi[i];
    }
}

__global__ void k_bfinish(const double* __restrict__ partials, int nparts, int nv, double* __restrict__ out) {
    double acc[3] = {0, 0, 0};
    for (int q = threadIdx.x; q < nparts; q += blockDim.x)
        for (int j = 0; j < nv; ++j) acc[j] += partials[(long)q * nv + j];
    __shared__ double res[3];
    block_sum<3>(acc, res);
    if (threadIdx.x < (unsigned)nv) out[threadIdx.x] = res[threadIdx.x];
}

// computeSuperficialVelocity (:233-258): plain sums over all voxels
__global__ void k_vel_sums(const double* __restrict__ ux, const double* __restrict__ uy,
                           const double* __restrict__ uz, long N, double* __restrict__ partials) {
    double acc[3] = {0, 0, 0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        acc[0] += ux[i];
        acc[1] += uy[i];
        acc[2] += uz[i];
    }
    block_sum<3>(acc, partials + 3L * blockIdx.x);
}

struct DevBuf {
    std::vector<void*> ptrs;
    ~DevBuf() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T* alloc(long n, hipError_t& err) {
        void* p = nullptr;
        if (err == hipSuccess) err = hipMalloc(&p, (size_t)std::max(1L, n) * sizeof(T));
        if (err == hipSuccess) ptrs.push_back(p);
        return static_cast<T*>(p);
    }
};

template <typename P>
int run(int nx, int ny, int nz, const double* alpha_h, const double* visc_h, double relax, double tol, int maxit,
        const double force[3], double U[3], double* u_h, BrinkmanResult* res) {
    const long N = (long)nx * ny * nz;
    hipError_t err = hipSuccess;
    DevBuf db;
    double* ux = db.alloc<double>(N, err);
    double* uy = db.alloc<double>(N, err);
    double* uz = db.alloc<double>(N, err);
    double* p = db.alloc<double>(N, err);
    double* sx = db.alloc<double>(N, err);
    double* sy = db.alloc<double>(N, err);
    double* sz = db.alloc<double>(N, err);
    double* div = db.alloc<double>(N, err);
    double* alpha = db.alloc<double>(N, err);
    double* visc = db.alloc<double>(N, err);
    double* partials = db.alloc<double>(3L * kBRedBlocks, err);
    double* scal = db.alloc<double>(4, err);
    const bool mixed = sizeof(P) == sizeof(float);
    P* divp = mixed ? db.alloc<P>(N, err) : reinterpret_cast<P*>(div);
    P* phi = db.alloc<P>(N, err);
    P* phin = db.alloc<P>(N, err);
    hipStream_t s = nullptr;
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    auto fail = [&](hipError_t e) {
        res->message = std::string("HIP error in Brinkman solver: ") + hipGetErrorString(e);
        if (s) (void)hipStreamDestroy(s);
        return 1;
    };
    if (err != hipSuccess) return fail(err);
    if ((err = hipMemcpyAsync(alpha, alpha_h, N * sizeof(double), hipMemcpyHostToDevice, s)) != hipSuccess) return fail(err);
    if ((err = hipMemcpyAsync(visc, visc_h, N * sizeof(double), hipMemcpyHostToDevice, s)) != hipSuccess) return fail(err);
    for (double* a : {ux, uy, uz, p})
        if ((err = hipMemsetAsync(a, 0, N * sizeof(double), s)) != hipSuccess) return fail(err);
    const unsigned grid = (unsigned)std::min<long>((N + kBThreads - 1) / kBThreads, 65536);
    const int pressureIterationsPerStep = 35;
    int rc = 0;
    for (int iter = 0; iter < maxit; ++iter) {
        hipLaunchKernelGGL(k_momentum, dim3(grid), dim3(kBThreads), 0, s, ux, uy, uz, p, alpha, visc, sx, sy, sz,
                           force[0], force[1], force[2], relax, nx, ny, nz, N);
        hipLaunchKernelGGL(k_divergence<P>, dim3(kBRedBlocks), dim3(kBThreads), 0, s, sx, sy, sz, div,
                           mixed ? divp : (P*)nullptr, nx, ny, nz, N, partials);
        hipLaunchKernelGGL(k_bfinish, dim3(1), dim3(kBThreads), 0, s, partials, kBRedBlocks, 1, scal);
        const P* divsrc = mixed ? divp : reinterpret_cast<const P*>(div);
        for (int k = 0; k < pressureIterationsPerStep; ++k) {
            hipLaunchKernelGGL(k_jacobi<P>, dim3(grid), dim3(kBThreads), 0, s, phi, divsrc, phin, k == 0 ? 1 : 0,
                               nx, ny, nz, N);
            std::swap(phi, phin);
        }
        hipLaunchKernelGGL(k_project<P>, dim3(grid), dim3(kBThreads), 0, s, sx, sy, sz, phi, ux, uy, uz, p, nx, ny,
                           nz, N);
        double l2 = 0.0;
        if ((err = hipMemcpyAsync(&l2, scal, sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) return fail(err);
        if ((err = hipStreamSynchronize(s)) != hipSuccess) return fail(err);
        res->iterations = iter + 1;
        res->residual = std::sqrt(l2 / (double)N);
        if (!std::isfinite(res->residual)) {
            rc = 2;
            break;
        }
        if (res->residual < tol) {
            res->converged = 1;
            break;
        }
    }
    if (rc == 0) {
        hipLaunchKernelGGL(k_vel_sums, dim3(kBRedBlocks), dim3(kBThreads), 0, s, ux, uy, uz, N, partials);
        hipLaunchKernelGGL(k_bfinish, dim3(1), dim3(kBThreads), 0, s, partials, kBRedBlocks, 3, scal);
        double sums[3];
        if ((err = hipMemcpyAsync(sums, scal, sizeof(sums), hipMemcpyDeviceToHost, s)) != hipSuccess) return fail(err);
        if (u_h) {
            if ((err = hipMemcpyAsync(u_h, ux, N * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) return fail(err);
            if ((err = hipMemcpyAsync(u_h + N, uy, N * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) return fail(err);
            if ((err = hipMemcpyAsync(u_h + 2 * N, uz, N * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) return fail(err);
        }
        if ((err = hipStreamSynchronize(s)) != hipSuccess) return fail(err);
        for (int k = 0; k < 3; ++k) U[k] = sums[k] / (double)N;
        if (!std::isfinite(U[0]) || !std::isfinite(U[1]) || !std::isfinite(U[2])) rc = 3;
    }
    if ((err = hipGetLastError()) != hipSuccess) return fail(err);
    (void)hipStreamDestroy(s);
    return rc;
}

}  // namespace

int brinkman_run(int nx, int ny, int nz, const double* alpha, const double* viscosity, double relaxation, double tol,
                 int max_iterations, const double force[3], bool mixed, double U[3], double* u, BrinkmanResult* res) {
    if (mixed) return run<float>(nx, ny, nz, alpha, viscosity, relaxation, tol, max_iterations, force, U, u, res);
    return run<double>(nx, ny, nz, alpha, viscosity, relaxation, tol, max_iterations, force, U, u, res);
}

}  // namespace fftperm

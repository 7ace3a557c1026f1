// brinkman.hip -- Brinkman penalisation loop (src/brinkman/FluidBrinkman.cpp:282-507) on the GPU.
//
// One iteration = relaxed Jacobi momentum update, central divergence (+ its L2 sum),
// 35 Jacobi sweeps of the pressure-correction Poisson problem from phi = 0, projection
// and pressure update.  Fields are SoA FP64 arrays with the flat index
// z*nx*ny + y*nx + x (:24-26); neighbours wrap periodically (:28-33).  In mixed mode the
// Poisson sweeps run on FP32 copies of div and phi; velocity, pressure, divergence and
// the residual stay FP64.
//
// Layout of the work: the first sweep (from phi = 0) is fused into the divergence kernel, the
// other 34 run two at a time in one pass that marches through z with phi, div and the
// intermediate sweep held in LDS plane rings, so a pair of sweeps reads phi and div once and
// writes phi once.
// Every sweep keeps the reference's per-voxel arithmetic, (sum of the six neighbours in the order
// -x, +x, -y, +y, -z, +z  minus div) / 6, so the result is independent of the blocking.  The host
// does not wait for each iteration: a device-side check after each projection records the
// residual and sets a stop flag when the reference would break (:469-499), later iterations of
// the chunk skip all their work, and chunks of iterations replay a captured hipGraph.
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

// Workgroups are dealt to the 8 XCDs round-robin and each XCD has its own L2: remap the block
// index so that each XCD takes one contiguous range of the work (consecutive rows or tiles), whose
// neighbour rows and halo planes then stay in its own L2.
__device__ __forceinline__ long xcd_block(long b, long nblocks) {
    constexpr int kXcd = 8;
    const long per = nblocks / kXcd;
    if (per == 0 || b >= per * kXcd) return b;
    return (b % kXcd) * per + b / kXcd;
}

// Walks the voxels row by row: workgroup b (XCD-remapped) takes a contiguous range of x rows
// (y, z) (the row coordinates are uniform across the workgroup), its threads the x positions of a
// row; the six periodic neighbours (:28-33) come from the row coordinates, with no per-voxel
// division.
template <typename F>
__device__ __forceinline__ void for_each_voxel(int nx, int ny, int nz, F&& f) {
    const long rows = (long)ny * nz;
    const long per = (rows + gridDim.x - 1) / gridDim.x;
    const long r0 = xcd_block(blockIdx.x, gridDim.x) * per;
    const long r1 = r0 + per < rows ? r0 + per : rows;
    for (long r = r0; r < r1; ++r) {
        const int y = (int)(r % ny), z = (int)(r / ny);
        const int ym = y == 0 ? ny - 1 : y - 1, yp = y + 1 == ny ? 0 : y + 1;
        const int zm = z == 0 ? nz - 1 : z - 1, zp = z + 1 == nz ? 0 : z + 1;
        const long base = r * nx;
        const long oym = ((long)z * ny + ym) * nx, oyp = ((long)z * ny + yp) * nx;
        const long ozm = ((long)zm * ny + y) * nx, ozp = ((long)zp * ny + y) * nx;
        for (int x = threadIdx.x; x < nx; x += blockDim.x) {
            const int xm = x == 0 ? nx - 1 : x - 1, xp = x + 1 == nx ? 0 : x + 1;
            Nbr n;
            n.id = base + x;
            n.xm = base + xm;
            n.xp = base + xp;
            n.ym = oym + x;
            n.yp = oyp + x;
            n.zm = ozm + x;
            n.zp = ozp + x;
            f(n);
        }
    }
}

// :313-369
__global__ void k_momentum(const double* __restrict__ ux, const double* __restrict__ uy,
                           const double* __restrict__ uz, const double* __restrict__ p,
                           const double* __restrict__ alpha, const double* __restrict__ visc, double* __restrict__ sx,
                           double* __restrict__ sy, double* __restrict__ sz, double f0, double f1, double f2,
                           double relax, int nx, int ny, int nz, const int* __restrict__ done) {
    if (*done) return;
    for_each_voxel(nx, ny, nz, [&](const Nbr& n) {
        const long i = n.id;
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
    });
}

// :371-400, plus the first pressure sweep (:402-433 from phi = 0): phi1 = (0 - div) / 6
template <typename P>
__global__ void k_divergence(const double* __restrict__ sx, const double* __restrict__ sy,
                             const double* __restrict__ sz, double* __restrict__ div, P* __restrict__ divp,
                             P* __restrict__ phi1, int nx, int ny, int nz, double* __restrict__ partials,
                             const int* __restrict__ done) {
    if (*done) return;
    double acc[1] = {0.0};
    for_each_voxel(nx, ny, nz, [&](const Nbr& n) {
        const long i = n.id;
        const double d = 0.5 * (sx[n.xp] - sx[n.xm]) + 0.5 * (sy[n.yp] - sy[n.ym]) + 0.5 * (sz[n.zp] - sz[n.zm]);
        div[i] = d;
        const P dp = (P)d;
        if (divp != nullptr) divp[i] = dp;
        phi1[i] = ((P)0 - dp) / (P)6.0;
        acc[0] += d * d;
    });
    block_sum<1>(acc, partials + blockIdx.x);
}

// Two Jacobi sweeps (:402-433) in one pass.  A workgroup owns a TX x TY column of the x-y plane
// and ZC consecutive planes, and marches through z keeping three planes of phi (with a two-voxel
// x-y halo), three of div and three of the intermediate sweep (one-voxel halo) in LDS rings: at
// step q it loads phi plane q + 1 and div plane q, computes the first sweep on plane q and the
// second on plane q - 1, which it writes.  Every LDS slot holds the periodic value of its wrapped
// voxel, so columns that stick out of small grids stay exact (their outside voxels are not
// written).
template <int TX, int TY>
struct JCol {
    static constexpr int AX = TX + 4, AY = TY + 4;  // phi plane with the two-voxel halo
    static constexpr int CX = TX + 2, CY = TY + 2;  // first sweep and div plane with the one-voxel halo
    static constexpr int NA = AX * AY, NC = CX * CY, NT = TX * TY;
};

__device__ __forceinline__ int wrapm(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

template <typename P, int TX, int TY, int ZC>
__global__ __launch_bounds__(256) void k_jacobi2(const P* __restrict__ phi, const P* __restrict__ div,
                                                 P* __restrict__ phin, int nx, int ny, int nz, int tiles_x,
                                                 int tiles_y, const int* __restrict__ done) {
    if (*done) return;
    using J = JCol<TX, TY>;
    constexpr int T = 256;
    constexpr int KA = (J::NA + T - 1) / T, KC = (J::NC + T - 1) / T;
    __shared__ P f0[3][J::NA];
    __shared__ P f1[3][J::NC];
    __shared__ P dv[3][J::NC];
    __shared__ int xt[J::AX], yt[J::AY];
    const long blk = xcd_block(blockIdx.x, gridDim.x);
    const int bx = (int)(blk % tiles_x);
    const int by = (int)((blk / tiles_x) % tiles_y);
    const int bz = (int)(blk / ((long)tiles_x * tiles_y));
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * ZC;
    const int z1 = min(z0 + ZC, nz);  // planes [z0, z1) are written
    const long plane = (long)nx * ny;
    for (int t = threadIdx.x; t < J::AX; t += T) xt[t] = wrapm(x0 - 2 + t, nx);
    for (int t = threadIdx.x; t < J::AY; t += T) yt[t] = wrapm(y0 - 2 + t, ny) * nx;
    __syncthreads();
    // ring slot of plane q (q >= z0 - 2)
    auto slot = [&](int q) { return (q - z0 + 3) % 3; };
    // each thread's share of a phi plane (two-voxel halo) and of a div plane (one-voxel halo),
    // fetched into registers one step ahead so that the loads overlap the sweeps of the step
    P ra[KA], rc[KC];
    auto fetch = [&](int qa, int qc) {
        const P* sa = phi + (long)wrapm(qa, nz) * plane;
        const P* sc = div + (long)wrapm(qc, nz) * plane;
#pragma unroll
        for (int k = 0; k < KA; ++k) {
            const int t = threadIdx.x + k * T;
            if (t < J::NA) ra[k] = sa[yt[t / J::AX] + xt[t % J::AX]];
        }
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int t = threadIdx.x + k * T;
            if (t < J::NC) rc[k] = sc[yt[t / J::CX + 1] + xt[t % J::CX + 1]];
        }
    };
    auto store = [&](int qa, int qc) {
        P* da = f0[slot(qa)];
        P* dc = dv[slot(qc)];
#pragma unroll
        for (int k = 0; k < KA; ++k) {
            const int t = threadIdx.x + k * T;
            if (t < J::NA) da[t] = ra[k];
        }
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int t = threadIdx.x + k * T;
            if (t < J::NC) dc[t] = rc[k];
        }
    };
    // phi planes z0 - 2 and z0 - 1 (the div plane staged with the first is not used)
    fetch(z0 - 2, z0 - 2);
    store(z0 - 2, z0 - 2);
    fetch(z0 - 1, z0 - 2);
    store(z0 - 1, z0 - 2);
    fetch(z0, z0 - 1);
    for (int q = z0 - 1; q <= z1; ++q) {
        store(q + 1, q);  // phi plane q + 1, div plane q
        __syncthreads();
        if (q < z1) fetch(q + 2, q + 1);
        {
            // first sweep on plane q, over the one-voxel-halo region
            const P* pm = f0[slot(q - 1)];
            const P* pc = f0[slot(q)];
            const P* pp = f0[slot(q + 1)];
            const P* d = dv[slot(q)];
            P* o = f1[slot(q)];
            for (int t = threadIdx.x; t < J::NC; t += T) {
                const int lx = t % J::CX, ly = t / J::CX;
                const int a = (ly + 1) * J::AX + (lx + 1);
                const P nb = pc[a - 1] + pc[a + 1] + pc[a - J::AX] + pc[a + J::AX] + pm[a] + pp[a];
                o[t] = (nb - d[t]) / (P)6.0;
            }
        }
        __syncthreads();
        const int zo = q - 1;  // second sweep on plane q - 1
        if (zo >= z0) {
            const P* pm = f1[slot(zo - 1)];
            const P* pc = f1[slot(zo)];
            const P* pp = f1[slot(zo + 1)];
            const P* d = dv[slot(zo)];
            P* out = phin + (long)zo * plane;
            for (int t = threadIdx.x; t < J::NT; t += T) {
                const int lx = t % TX, ly = t / TX;
                const int gx = x0 + lx, gy = y0 + ly;
                if (gx >= nx || gy >= ny) continue;
                const int c = (ly + 1) * J::CX + (lx + 1);
                const P nb = pc[c - 1] + pc[c + 1] + pc[c - J::CX] + pc[c + J::CX] + pm[c] + pp[c];
                out[(long)gy * nx + gx] = (nb - d[c]) / (P)6.0;
            }
        }
    }
}

// One Jacobi sweep (:402-433), for an odd count of sweeps after the fused first one.
template <typename P>
__global__ void k_jacobi(const P* __restrict__ phi, const P* __restrict__ div, P* __restrict__ phin, int nx, int ny,
                         int nz, const int* __restrict__ done) {
    if (*done) return;
    for_each_voxel(nx, ny, nz, [&](const Nbr& n) {
        const P nb = phi[n.xm] + phi[n.xp] + phi[n.ym] + phi[n.yp] + phi[n.zm] + phi[n.zp];
        phin[n.id] = (nb - div[n.id]) / (P)6.0;
    });
}

// :435-466
template <typename P>
__global__ void k_project(const double* __restrict__ sx, const double* __restrict__ sy,
                          const double* __restrict__ sz, const P* __restrict__ phi, double* __restrict__ ux,
                          double* __restrict__ uy, double* __restrict__ uz, double* __restrict__ p, int nx, int ny,
                          int nz, const int* __restrict__ done) {
    if (*done) return;
    for_each_voxel(nx, ny, nz, [&](const Nbr& n) {
        const long i = n.id;
        const double dx = 0.5 * ((double)phi[n.xp] - (double)phi[n.xm]);
        const double dy = 0.5 * ((double)phi[n.yp] - (double)phi[n.ym]);
        const double dz = 0.5 * ((double)phi[n.zp] - (double)phi[n.zm]);
        ux[i] = sx[i] - dx;
        uy[i] = sy[i] - dy;
        uz[i] = sz[i] - dz;
        p[i] += (double)phi[i];
    });
}

// Device-side stop rule of one iteration (:468-499): after the projection, the residual
// sqrt(sum div^2 / N) of this iteration is recorded; a non-finite residual (state 2) or one below
// tol (state 1) stops the loop -- later kernels of the chunk see done != 0 and return.
struct BState {
    int done;        // 0 running, 1 converged, 2 non-finite residual
    int iterations;  // iterations completed
    double residual;
};

__global__ void k_bcheck(const double* __restrict__ l2, BState* __restrict__ st, double tol, double inv_n) {
    if (threadIdx.x != 0 || st->done) return;
    const double r = sqrt(*l2 * inv_n);
    st->iterations += 1;
    st->residual = r;
    if (!isfinite(r))
        st->done = 2;
    else if (r < tol)
        st->done = 1;
}

__global__ void k_bfinish(const double* __restrict__ partials, int nparts, int nv, double* __restrict__ out,
                          const int* __restrict__ done) {
    if (done != nullptr && *done) return;
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

constexpr int kJTX = 32, kJTY = 8, kJZC = 16;  // column of the fused sweep pairs: x-y tile, planes
constexpr int kChunk = 8;                     // iterations per captured graph / host check

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
    BState* bst = db.alloc<BState>(1, err);
    const bool mixed = sizeof(P) == sizeof(float);
    P* divp = mixed ? db.alloc<P>(N, err) : reinterpret_cast<P*>(div);
    P* phi = db.alloc<P>(N, err);
    P* phin = db.alloc<P>(N, err);
    hipStream_t s = nullptr;
    hipGraphExec_t chunk_exec = nullptr;
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    auto fail = [&](hipError_t e) {
        res->message = std::string("HIP error in Brinkman solver: ") + hipGetErrorString(e);
        if (chunk_exec) (void)hipGraphExecDestroy(chunk_exec);
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        return 1;
    };
    if (err != hipSuccess) return fail(err);
    if ((err = hipMemcpyAsync(alpha, alpha_h, N * sizeof(double), hipMemcpyHostToDevice, s)) != hipSuccess) return fail(err);
    if ((err = hipMemcpyAsync(visc, visc_h, N * sizeof(double), hipMemcpyHostToDevice, s)) != hipSuccess) return fail(err);
    for (double* a : {ux, uy, uz, p})
        if ((err = hipMemsetAsync(a, 0, N * sizeof(double), s)) != hipSuccess) return fail(err);
    if ((err = hipMemsetAsync(bst, 0, sizeof(BState), s)) != hipSuccess) return fail(err);
    const unsigned grid = (unsigned)std::min<long>((long)ny * nz, 65536);  // row-walking kernels
    const int tiles_x = (nx + kJTX - 1) / kJTX, tiles_y = (ny + kJTY - 1) / kJTY, tiles_z = (nz + kJZC - 1) / kJZC;
    const unsigned jgrid = (unsigned)((long)tiles_x * tiles_y * tiles_z);
    const int pressureIterationsPerStep = 35;  // the first one fused into the divergence kernel
    const int pairs = (pressureIterationsPerStep - 1) / 2, single = (pressureIterationsPerStep - 1) % 2;
    const int* done = &bst->done;
    const P* divsrc = mixed ? divp : reinterpret_cast<const P*>(div);
    // one iteration; the phi buffers alternate the same way every iteration, so the launches of a
    // chunk replay unchanged
    auto iteration = [&]() {
        hipLaunchKernelGGL(k_momentum, dim3(grid), dim3(kBThreads), 0, s, ux, uy, uz, p, alpha, visc, sx, sy, sz,
                           force[0], force[1], force[2], relax, nx, ny, nz, done);
        hipLaunchKernelGGL(k_divergence<P>, dim3(kBRedBlocks), dim3(kBThreads), 0, s, sx, sy, sz, div,
                           mixed ? divp : (P*)nullptr, phi, nx, ny, nz, partials, done);
        hipLaunchKernelGGL(k_bfinish, dim3(1), dim3(kBThreads), 0, s, partials, kBRedBlocks, 1, scal, done);
        P* a = phi;
        P* b = phin;
        for (int k = 0; k < pairs; ++k) {
            hipLaunchKernelGGL((k_jacobi2<P, kJTX, kJTY, kJZC>), dim3(jgrid), dim3(256), 0, s, a, divsrc, b, nx, ny, nz,
                               tiles_x, tiles_y, done);
            std::swap(a, b);
        }
        if (single) {
            hipLaunchKernelGGL(k_jacobi<P>, dim3(grid), dim3(kBThreads), 0, s, a, divsrc, b, nx, ny, nz, done);
            std::swap(a, b);
        }
        hipLaunchKernelGGL(k_project<P>, dim3(grid), dim3(kBThreads), 0, s, sx, sy, sz, a, ux, uy, uz, p, nx, ny,
                           nz, done);
        hipLaunchKernelGGL(k_bcheck, dim3(1), dim3(64), 0, s, scal, bst, tol, 1.0 / (double)N);
    };
    BState hs{0, 0, 0.0};
    auto read_state = [&]() {
        if ((err = hipMemcpyAsync(&hs, bst, sizeof(BState), hipMemcpyDeviceToHost, s)) != hipSuccess) return false;
        return (err = hipStreamSynchronize(s)) == hipSuccess;
    };
    int it = 0;
    const int full = maxit / kChunk;
    if (full > 0) {
        hipGraph_t graph = nullptr;
        if ((err = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess) return fail(err);
        for (int c = 0; c < kChunk; ++c) iteration();
        if ((err = hipStreamEndCapture(s, &graph)) != hipSuccess) return fail(err);
        err = hipGraphInstantiate(&chunk_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (err != hipSuccess) return fail(err);
        for (int c = 0; c < full; ++c) {
            if ((err = hipGraphLaunch(chunk_exec, s)) != hipSuccess) return fail(err);
            it += kChunk;
            if (!read_state()) return fail(err);
            if (hs.done) break;
        }
    }
    if (!hs.done) {
        for (; it < maxit; ++it) iteration();
        if (!read_state()) return fail(err);
    }
    if ((err = hipGetLastError()) != hipSuccess) return fail(err);
    res->iterations = hs.iterations;
    res->residual = hs.residual;
    res->converged = hs.done == 1 ? 1 : 0;
    int rc = hs.done == 2 ? 2 : 0;
    if (rc == 0) {
        hipLaunchKernelGGL(k_vel_sums, dim3(kBRedBlocks), dim3(kBThreads), 0, s, ux, uy, uz, N, partials);
        hipLaunchKernelGGL(k_bfinish, dim3(1), dim3(kBThreads), 0, s, partials, kBRedBlocks, 3, scal, (const int*)nullptr);
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
    if (chunk_exec) (void)hipGraphExecDestroy(chunk_exec);
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

// kernels_setup.hip -- support detection/compaction and the energy-consistent Green tensor.
#include "kernels.hpp"

#include <algorithm>
#include <cfloat>

namespace fftperm {

// ------------------------------------------------------------------ support
// InterfaceVoxels::compute_interface_only (src/geometry/InterfaceVoxels.cpp:7-43): a solid
// voxel is on the interface when any of its 26 periodic neighbours is fluid.  FullSolid
// (src/bignonnet/ForceSupport.cpp:13-18): every solid voxel.  The flat-index ascending order
// of the reference falls out of an exclusive prefix sum over the flags.
namespace {
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanChunk = kScanThreads * kScanItems;

__device__ __forceinline__ int wrapi(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

__device__ __forceinline__ int support_flag(const uint8_t* __restrict__ mask, int nx, int ny, int nz, long v,
                                            int interface_only) {
    if (mask[v] != 0) return 0;
    if (!interface_only) return 1;
    const int i = (int)(v % nx);
    const long r = v / nx;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    for (int dk = -1; dk <= 1; ++dk) {
        const int kk = wrapi(k + dk, nz);
        for (int dj = -1; dj <= 1; ++dj) {
            const int jj = wrapi(j + dj, ny);
            const long rowb = ((long)kk * ny + jj) * nx;
            for (int di = -1; di <= 1; ++di) {
                if (di == 0 && dj == 0 && dk == 0) continue;
                const int ii = wrapi(i + di, nx);
                if (mask[rowb + ii] != 0) return 1;
            }
        }
    }
    return 0;
}

// Exclusive block scan of one int per thread; returns the block total.
__device__ int block_exclusive_scan(int v, int* out_excl) {
    __shared__ int sh[kScanThreads];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int t = (threadIdx.x >= (unsigned)off) ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    const int incl = sh[threadIdx.x];
    const int total = sh[kScanThreads - 1];
    __syncthreads();
    *out_excl = incl - v;
    return total;
}
}  // namespace

__global__ void k_support_count(const uint8_t* __restrict__ mask, int nx, int ny, int nz, long N,
                                int interface_only, int* __restrict__ bsum, unsigned long long* __restrict__ nf) {
    const long base = (long)blockIdx.x * kScanChunk + (long)threadIdx.x * kScanItems;
    int cnt = 0, fl = 0;
    for (int q = 0; q < kScanItems; ++q) {
        const long v = base + q;
        if (v < N) {
            cnt += support_flag(mask, nx, ny, nz, v, interface_only);
            fl += (mask[v] != 0) ? 1 : 0;
        }
    }
    int excl;
    const int tot = block_exclusive_scan(cnt, &excl);
    int exf;
    const int totf = block_exclusive_scan(fl, &exf);
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = tot;
        atomicAdd(nf, (unsigned long long)totf);
    }
}

__global__ void k_scan_blocks(int* __restrict__ bsum, long nblocks, unsigned long long* __restrict__ nb) {
    // single block: exclusive scan of bsum in place (as 64-bit running carry)
    __shared__ long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (long c0 = 0; c0 < nblocks; c0 += kScanThreads) {
        const long i = c0 + threadIdx.x;
        const int v = (i < nblocks) ? bsum[i] : 0;
        int excl;
        const int tot = block_exclusive_scan(v, &excl);
        if (i < nblocks) bsum[i] = (int)(carry + excl);
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *nb = (unsigned long long)carry;
}

__global__ void k_support_write(const uint8_t* __restrict__ mask, int nx, int ny, int nz, long N,
                                int interface_only, const int* __restrict__ bsum, int32_t* __restrict__ code,
                                int64_t* __restrict__ active) {
    const long base = (long)blockIdx.x * kScanChunk + (long)threadIdx.x * kScanItems;
    int f[kScanItems];
    int cnt = 0;
    for (int q = 0; q < kScanItems; ++q) {
        const long v = base + q;
        f[q] = (v < N) ? support_flag(mask, nx, ny, nz, v, interface_only) : 0;
        cnt += f[q];
    }
    int excl;
    block_exclusive_scan(cnt, &excl);
    int local = bsum[blockIdx.x] + excl;
    for (int q = 0; q < kScanItems; ++q) {
        const long v = base + q;
        if (v >= N) break;
        if (f[q]) {
            code[v] = local;
            active[local] = v;
            ++local;
        } else {
            code[v] = (mask[v] != 0) ? -1 : -2;
        }
    }
}

// The support map in the reference's own form: voxel_to_local (int64, -1 = kInvalid,
// src/bignonnet/ForceSupport.hpp:22-29) instead of the int32 code map, and the ascending active list
// bounded by the caller's capacity (entries at and beyond it are not written; the caller compares
// the count against the capacity).
__global__ void k_support_write_v2l(const uint8_t* __restrict__ mask, int nx, int ny, int nz, long N,
                                    int interface_only, const int* __restrict__ bsum,
                                    int64_t* __restrict__ voxel_to_local, int64_t* __restrict__ active,
                                    long capacity) {
    const long base = (long)blockIdx.x * kScanChunk + (long)threadIdx.x * kScanItems;
    int f[kScanItems];
    int cnt = 0;
    for (int q = 0; q < kScanItems; ++q) {
        const long v = base + q;
        f[q] = (v < N) ? support_flag(mask, nx, ny, nz, v, interface_only) : 0;
        cnt += f[q];
    }
    int excl;
    block_exclusive_scan(cnt, &excl);
    long local = (long)bsum[blockIdx.x] + excl;
    for (int q = 0; q < kScanItems; ++q) {
        const long v = base + q;
        if (v >= N) break;
        if (f[q]) {
            if (voxel_to_local) voxel_to_local[v] = local;
            if (local < capacity) active[local] = v;
            ++local;
        } else if (voxel_to_local) {
            voxel_to_local[v] = -1;
        }
    }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int n, unsigned long long value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void k_line_bits(const int32_t* __restrict__ code, int nx, long nlines, int W, uint32_t* __restrict__ bits,
                            int* __restrict__ counts) {
    // one thread per word; the W threads of a line add their popcounts into counts[line]
    for (long wid = (long)blockIdx.x * blockDim.x + threadIdx.x; wid < nlines * W; wid += (long)gridDim.x * blockDim.x) {
        const long line = wid / W;
        const int w = (int)(wid % W);
        uint32_t b = 0;
        for (int q = 0; q < 32; ++q) {
            const int i = w * 32 + q;
            if (i < nx && code[line * nx + i] >= 0) b |= 1u << q;
        }
        bits[wid] = b;
        if (b != 0) atomicAdd(&counts[line], __popc(b));
    }
}

void launch_line_bits(const int32_t* code, int nx, long nlines, uint32_t* bits, int32_t* lstart,
                      unsigned long long* total, hipStream_t st) {
    const int W = (nx + 31) / 32;
    (void)hipMemsetAsync(lstart, 0, nlines * sizeof(int32_t), st);
    const long words = nlines * W;
    hipLaunchKernelGGL(k_line_bits, dim3((unsigned)std::min<long>((words + 255) / 256, 65536)), dim3(256), 0, st,
                       code, nx, nlines, W, bits, lstart);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanThreads), 0, st, lstart, nlines, total);
}

size_t scan_scratch_bytes(long N) {
    const long nblocks = (N + kScanChunk - 1) / kScanChunk;
    return (size_t)nblocks * sizeof(int) + 256;
}

void launch_support(const uint8_t* mask, int nx, int ny, int nz, int interface_only, int32_t* code,
                    int64_t* active, unsigned long long* counts, void* scratch, hipStream_t st) {
    const long N = (long)nx * ny * nz;
    const long nblocks = (N + kScanChunk - 1) / kScanChunk;
    int* bsum = static_cast<int*>(scratch);
    hipLaunchKernelGGL(k_support_count, dim3((unsigned)nblocks), dim3(kScanThreads), 0, st, mask, nx, ny, nz, N,
                       interface_only, bsum, counts + 1);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanThreads), 0, st, bsum, nblocks, counts);
    hipLaunchKernelGGL(k_support_write, dim3((unsigned)nblocks), dim3(kScanThreads), 0, st, mask, nx, ny, nz, N,
                       interface_only, bsum, code, active);
}

long support_scan_blocks(long N) { return (N + kScanChunk - 1) / kScanChunk; }

void launch_support_v2l(const uint8_t* mask, int nx, int ny, int nz, int interface_only, int64_t* voxel_to_local,
                        int64_t* active, long capacity, unsigned long long* counts, int* bsum, hipStream_t st) {
    const long N = (long)nx * ny * nz;
    const long nblocks = support_scan_blocks(N);
    // the counters start at zero on every call: a fill kernel, not a memset node, so that the
    // sequence can sit inside a captured graph
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, counts, 2, 0ull);
    hipLaunchKernelGGL(k_support_count, dim3((unsigned)nblocks), dim3(kScanThreads), 0, st, mask, nx, ny, nz, N,
                       interface_only, bsum, counts + 1);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanThreads), 0, st, bsum, nblocks, counts);
    hipLaunchKernelGGL(k_support_write_v2l, dim3((unsigned)nblocks), dim3(kScanThreads), 0, st, mask, nx, ny, nz, N,
                       interface_only, bsum, voxel_to_local, active, capacity);
}

// ------------------------------------------------------------ Green tensor
// GreenOperator::precompute_energy_consistent_tensor (src/bignonnet/GreenOperator.cpp:95-178).
// sinc^2 and q come from per-axis tables computed on the host with the reference's
// expressions; the p-sum keeps the px -> py -> pz order and the same operation order,
// with contraction disabled so that the products and quotients round as in the reference.
__device__ void green_at(const GreenAxisTables& t, int ki, int kj, int kk, double out[6]) {
#pragma clang fp contract(off)
    for (int g = 0; g < 6; ++g) out[g] = 0.0;
    if (ki == 0 && kj == 0 && kk == 0) return;
    const int P = t.P;
    const int W = 2 * P + 1;
    double gxx = 0.0, gxy = 0.0, gxz = 0.0, gyy = 0.0, gyz = 0.0, gzz = 0.0;
    for (int px = 0; px < W; ++px) {
        const double sx2 = t.s2x[ki * W + px];
        const double qx = t.qx[ki * W + px];
        for (int py = 0; py < W; ++py) {
            const double sy2 = t.s2y[kj * W + py];
            const double qy = t.qy[kj * W + py];
            const double wxy = sx2 * sy2;
            for (int pz = 0; pz < W; ++pz) {
                const double sz2 = t.s2z[kk * W + pz];
                const double weight = wxy * sz2;
                if (weight == 0.0) continue;
                const double qz = t.qz[kk * W + pz];
                const double q2 = qx * qx + qy * qy + qz * qz;
                if (q2 <= DBL_EPSILON) continue;
                const double inv_mu_q2 = 1.0 / (t.mu * q2);
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
    out[0] = gxx;
    out[1] = gxy;
    out[2] = gxz;
    out[3] = gyy;
    out[4] = gyz;
    out[5] = gzz;
}

__global__ void k_green_sym(GreenAxisTables t, int nx, int ny, int nz, int nxh, long Nh, int rp, long pz,
                            double* __restrict__ G6, float* __restrict__ G6f) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nh) return;
    const int kx = (int)(i % nxh);
    const long r = i / nxh;
    const int ky = (int)(r % ny);
    const int kz = (int)(r / ny);
    double a[6], b[6];
    green_at(t, kx, ky, kz, a);
    green_at(t, (nx - kx) % nx, (ny - ky) % ny, (nz - kz) % nz, b);
    // AoS: the six components of one frequency are contiguous (one 48-byte run per frequency), at
    // the spectrum's padded element index kz * pz + ky * rp + kx
    const long e = (long)kz * pz + (long)ky * rp + kx;
    for (int g = 0; g < 6; ++g) {
        const double s = 0.5 * (a[g] + b[g]);
        G6[6 * e + g] = s;
        if (G6f != nullptr) G6f[6 * e + g] = (float)s;
    }
}

void launch_green_sym(const GreenAxisTables& t, int nx, int ny, int nz, int rp, long pz, double* G6, float* G6f,
                      hipStream_t st) {
    const int nxh = nx / 2 + 1;
    const long Nh = (long)nxh * ny * nz;
    hipLaunchKernelGGL(k_green_sym, dim3((unsigned)((Nh + 127) / 128)), dim3(128), 0, st, t, nx, ny, nz, nxh, Nh, rp,
                       pz, G6, G6f);
}

__global__ void k_green_full(GreenAxisTables t, int nx, int ny, int nz, long N, double* __restrict__ T6) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int kx = (int)(i % nx);
    const long r = i / nx;
    const int ky = (int)(r % ny);
    const int kz = (int)(r / ny);
    double a[6];
    green_at(t, kx, ky, kz, a);
    for (int g = 0; g < 6; ++g) T6[g * N + i] = a[g];
}

void launch_green_full(const GreenAxisTables& t, int nx, int ny, int nz, double* T6, hipStream_t st) {
    const long N = (long)nx * ny * nz;
    hipLaunchKernelGGL(k_green_full, dim3((unsigned)((N + 127) / 128)), dim3(128), 0, st, t, nx, ny, nz, N, T6);
}

}  // namespace fftperm

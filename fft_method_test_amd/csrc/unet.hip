// unet.hip -- learned warm start: a 4-level 3D U-Net (16 base channels, BF16 weights and
// activations, FP32 accumulation) that maps a periodic binary medium to the interface force
// densities of the three unit load directions of the variational problem (one forward pass for
// the whole tensor solve), with its own training kernels (backward data / weight passes, Adam)
// and the C ABI of include/fftperm.h.
//
//   features x0 [n^3][8]: solid indicator, interface-support indicator, 6 zero channels
//   level 0 (n)   : conv 8->16, conv 16->16                 -> e0 (skip)
//   level 1 (n/2) : avgpool, conv 16->32, conv 32->32        -> e1 (skip)
//   level 2 (n/4) : avgpool, conv 32->64, conv 64->64        -> e2 (skip)
//   level 3 (n/8) : avgpool, conv 64->128, conv 128->128
//   decoder       : nearest upsample, concat skip, conv (C+skip)->C_l, conv C_l->C_l per level
//   head          : 1x1x1 conv 16->16 (channels 3d..3d+2 = force density of load direction d, FP32)
// Every 3x3x3 convolution is periodic and followed by ReLU.  The loss is the mean squared error
// over the support voxels (channel 1 of the features) of the nine predicted components.
#include "kernels.hpp"
#include "unet.hpp"
#include "unet_kernels.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace fftperm {
namespace unet {

// ------------------------------------------------------------------ elementwise kernels
constexpr int kEw = 256;

inline unsigned ew_grid(long n) { return (unsigned)std::min<long>((n + kEw - 1) / kEw, 65535L * 8); }

__global__ void k_features(const uint8_t* __restrict__ fluid, int n, long V, uint16_t* __restrict__ x) {
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
        const int i = (int)(v % n), j = (int)((v / n) % n), k = (int)(v / ((long)n * n));
        const bool solid = fluid[v] == 0;
        bool iface = false;
        if (solid) {
            for (int t = 0; t < 27 && !iface; ++t)
                if (t != 13 && fluid[neighbour<27>(i, j, k, n, t)] != 0) iface = true;
        }
        uint16_t* o = x + v * 8;
        o[0] = solid ? f2bf(1.f) : 0;
        o[1] = iface ? f2bf(1.f) : 0;
#pragma unroll
        for (int c = 2; c < 8; ++c) o[c] = 0;
    }
}

// y = dy * (x > 0) over count BF16 values
__global__ void k_relu_bwd(const uint16_t* __restrict__ dy, const uint16_t* __restrict__ y, uint16_t* __restrict__ dz,
                           long count) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (long)gridDim.x * blockDim.x)
        dz[e] = (bf2f(y[e]) > 0.f) ? dy[e] : (uint16_t)0;
}

// 2x2x2 average pool, fine edge n -> n/2
__global__ void k_pool_fwd(const uint16_t* __restrict__ x, int n, int C, uint16_t* __restrict__ y) {
    const int m = n / 2;
    const long Vc = (long)m * m * m;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < Vc * C; e += (long)gridDim.x * blockDim.x) {
        const long vc = e / C;
        const int c = (int)(e % C);
        const int i = (int)(vc % m), j = (int)((vc / m) % m), k = (int)(vc / ((long)m * m));
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const long vf = ((long)(2 * k + (q >> 2)) * n + (2 * j + ((q >> 1) & 1))) * n + (2 * i + (q & 1));
            s += bf2f(x[vf * C + c]);
        }
        y[e] = f2bf(0.125f * s);
    }
}

// gradient of the average pool: dx_fine = dy_coarse / 8 (+ addend when given)
__global__ void k_pool_bwd(const uint16_t* __restrict__ dy, int n, int C, const uint16_t* __restrict__ addend,
                           uint16_t* __restrict__ dx) {
    const int m = n / 2;
    const long V = (long)n * n * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < V * C; e += (long)gridDim.x * blockDim.x) {
        const long vf = e / C;
        const int c = (int)(e % C);
        const int i = (int)(vf % n), j = (int)((vf / n) % n), k = (int)(vf / ((long)n * n));
        const long vc = ((long)(k / 2) * m + j / 2) * m + i / 2;
        float s = 0.125f * bf2f(dy[vc * C + c]);
        if (addend != nullptr) s += bf2f(addend[e]);
        dx[e] = f2bf(s);
    }
}

// nearest upsample, coarse edge m -> 2m
__global__ void k_up_fwd(const uint16_t* __restrict__ x, int m, int C, uint16_t* __restrict__ y) {
    const int n = 2 * m;
    const long V = (long)n * n * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < V * C; e += (long)gridDim.x * blockDim.x) {
        const long vf = e / C;
        const int c = (int)(e % C);
        const int i = (int)(vf % n), j = (int)((vf / n) % n), k = (int)(vf / ((long)n * n));
        y[e] = x[(((long)(k / 2) * m + j / 2) * m + i / 2) * C + c];
    }
}

// gradient of the upsample: dx_coarse = sum over the 8 children
__global__ void k_up_bwd(const uint16_t* __restrict__ dy, int m, int C, uint16_t* __restrict__ dx) {
    const int n = 2 * m;
    const long Vc = (long)m * m * m;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < Vc * C; e += (long)gridDim.x * blockDim.x) {
        const long vc = e / C;
        const int c = (int)(e % C);
        const int i = (int)(vc % m), j = (int)((vc / m) % m), k = (int)(vc / ((long)m * m));
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const long vf = ((long)(2 * k + (q >> 2)) * n + (2 * j + ((q >> 1) & 1))) * n + (2 * i + (q & 1));
            s += bf2f(dy[vf * C + c]);
        }
        dx[e] = f2bf(s);
    }
}

// d[V][C0 + C1] -> d0 [V][C0] (written), d1 [V][C1] (written)
__global__ void k_split(const uint16_t* __restrict__ d, long V, int C0, int C1, uint16_t* __restrict__ d0,
                        uint16_t* __restrict__ d1) {
    const int C = C0 + C1;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < V * C; e += (long)gridDim.x * blockDim.x) {
        const long v = e / C;
        const int c = (int)(e % C);
        if (c < C0)
            d0[v * C0 + c] = d[e];
        else
            d1[v * C1 + (c - C0)] = d[e];
    }
}

// masked MSE over the support (feature channel 1) of pred channels 0..8 (direction d's three
// components at 3d..3d+2) against target [V][9]; dpred [V][16] BF16 = d loss / d pred (0 off the
// support and on channels 9..15); partial sums of (squared error, support count) per block.
__global__ void k_loss(const float* __restrict__ pred, const float* __restrict__ target,
                       const uint16_t* __restrict__ feat, long V, float inv_count, uint16_t* __restrict__ dpred,
                       double* __restrict__ part) {
    double acc[2] = {0.0, 0.0};
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
        const bool sup = bf2f(feat[v * 8 + 1]) > 0.5f;
        uint16_t* dp = dpred + v * 16;
#pragma unroll
        for (int c = 0; c < 16; ++c) dp[c] = 0;
        if (sup) {
            acc[1] += 1.0;
#pragma unroll
            for (int c = 0; c < kOutputs; ++c) {
                const float d = pred[v * 16 + c] - target[v * kOutputs + c];
                acc[0] += (double)d * d;
                if (inv_count > 0.f) dp[c] = f2bf(2.f * d * inv_count);
            }
        }
    }
    block_sum<2>(acc, part + 2L * blockIdx.x);
}

// support voxel count of the features (channel 1), partial per block
__global__ void k_count_support(const uint16_t* __restrict__ feat, long V, double* __restrict__ part) {
    double acc[1] = {0.0};
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x)
        acc[0] += bf2f(feat[v * 8 + 1]) > 0.5f ? 1.0 : 0.0;
    block_sum<1>(acc, part + blockIdx.x);
}

// db[co] += sum_v dz[v][co]
__global__ void k_dbias(const uint16_t* __restrict__ dz, long V, int C, float* __restrict__ db) {
    const int c = threadIdx.x % C;
    const int lanes = blockDim.x / C;
    const int li = threadIdx.x / C;
    float s = 0.f;
    if (li < lanes)
        for (long v = (long)blockIdx.x * lanes + li; v < V; v += (long)gridDim.x * lanes) s += bf2f(dz[v * C + c]);
    atomicAdd(&db[c], s);
}

// dgrad weights: wd[ci][t' * cout + co] = w[co][t * cin + ci], t' = taps - 1 - t
__global__ void k_flip(const uint16_t* __restrict__ w, int taps, int cin, int cout, int Kp, int Kpd,
                       uint16_t* __restrict__ wd) {
    const long tot = (long)cin * Kpd;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int ci = (int)(e / Kpd);
        const int kd = (int)(e % Kpd);
        uint16_t val = 0;
        if (kd < taps * cout) {
            const int tp = kd / cout, co = kd % cout;
            const int t = taps - 1 - tp;
            val = w[(long)co * Kp + t * cin + ci];
        }
        wd[e] = val;
    }
}

// Adam on FP32 master weights; refreshes the BF16 copy
__global__ void k_adam(float* __restrict__ w, uint16_t* __restrict__ wb, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ s, long count, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (long)gridDim.x * blockDim.x) {
        const float ge = g[e];
        const float me = b1 * m[e] + (1.f - b1) * ge;
        const float se = b2 * s[e] + (1.f - b2) * ge * ge;
        m[e] = me;
        s[e] = se;
        const float we = w[e] - lr * (me * c1) / (sqrtf(se * c2) + eps);
        w[e] = we;
        if (wb != nullptr) wb[e] = f2bf(we);
    }
}

__global__ void k_to_bf16(const float* __restrict__ x, long count, uint16_t* __restrict__ y) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (long)gridDim.x * blockDim.x)
        y[e] = f2bf(x[e]);
}

// The predicted force density of load direction `dir` (channels 3 dir .. 3 dir + 2) gathered on
// the support (through `active`) into the compact AoS vector, in double.
__global__ void k_gather_direction(const float* __restrict__ pred, int dir, const int64_t* __restrict__ active,
                                   long nb, double* __restrict__ x0) {
    for (long l = (long)blockIdx.x * blockDim.x + threadIdx.x; l < nb; l += (long)gridDim.x * blockDim.x) {
        const float* f = pred + active[l] * 16 + 3 * dir;
        x0[3 * l + 0] = f[0];
        x0[3 * l + 1] = f[1];
        x0[3 * l + 2] = f[2];
    }
}

__global__ void k_pred_field(const float* __restrict__ pred, long V, float* __restrict__ out9) {
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < kOutputs; ++c) out9[v * kOutputs + c] = pred[v * 16 + c];
}

// ------------------------------------------------------------------ launchers
template <int TAPS>
void conv_fwd_t(const uint16_t* in0, int c0, const uint16_t* in1, int c1, int n, const uint16_t* wt, int Kp, int cout,
                const float* bias, bool relu, bool out_f32, void* out, hipStream_t st) {
    const long V = (long)n * n * n;
    const dim3 grid((unsigned)((V + 63) / 64));
#define UNET_FWD(NT)                                                                                          \
    case NT:                                                                                                  \
        hipLaunchKernelGGL((k_conv_fwd<TAPS, NT>), grid, dim3(256), 0, st, in0, c0, in1, c1, n, V, wt, Kp, bias, \
                           relu ? 1 : 0, out_f32 ? 1 : 0, out);                                               \
        break;
    switch (cout / 16) {
        UNET_FWD(1) UNET_FWD(2) UNET_FWD(3) UNET_FWD(4) UNET_FWD(6) UNET_FWD(8) UNET_FWD(12)
        default: throw std::runtime_error("unsupported output channel count " + std::to_string(cout));
    }
#undef UNET_FWD
}

// The halo-tiled kernel where the shape allows it (n a multiple of 64, 16 | c0, c1, cout <= 64).
bool conv3_tile(const uint16_t* in0, int c0, const uint16_t* in1, int c1, int n, const uint16_t* wt, int Kp,
                int cout, const float* bias, bool relu, bool out_f32, void* out, hipStream_t st) {
    if (n % kTileX != 0 || c0 % 16 != 0 || c1 % 16 != 0 || c0 == 0) return false;
    const long blocks = (long)(n / kTileX) * (n / 2) * (n / 2);
#define UNET_TILE(NT)                                                                                        \
    case NT:                                                                                                 \
        hipLaunchKernelGGL((k_conv3_tile<NT>), dim3((unsigned)blocks), dim3(256), 0, st, in0, c0, in1, c1, n, wt, \
                           Kp, bias, relu ? 1 : 0, out_f32 ? 1 : 0, out);                                      \
        return true;
    switch (cout / 16) {
        UNET_TILE(1) UNET_TILE(2) UNET_TILE(4)
        default: return false;
    }
#undef UNET_TILE
}

void conv_fwd(int taps, const uint16_t* in0, int c0, const uint16_t* in1, int c1, int n, const uint16_t* wt, int Kp,
              int cout, const float* bias, bool relu, bool out_f32, void* out, hipStream_t st) {
    if (taps == 27) {
        if (cout % 16 == 0 && conv3_tile(in0, c0, in1, c1, n, wt, Kp, cout, bias, relu, out_f32, out, st)) return;
        conv_fwd_t<27>(in0, c0, in1, c1, n, wt, Kp, cout, bias, relu, out_f32, out, st);
    } else {
        conv_fwd_t<1>(in0, c0, in1, c1, n, wt, Kp, cout, bias, relu, out_f32, out, st);
    }
}

void conv_wgrad(int taps, const uint16_t* in0, int c0, const uint16_t* in1, int c1, int n, const uint16_t* dz,
                int cout, int Kp, float* dw, hipStream_t st) {
    const long V = (long)n * n * n;
    const long vchunk = std::max<long>(512, ((V / 64) + 31) / 32 * 32);  // at most ~64 chunks
    const dim3 grid((unsigned)(cout / 16), (unsigned)(Kp / 16), (unsigned)((V + vchunk - 1) / vchunk));
    if (taps == 27)
        hipLaunchKernelGGL((k_conv_wgrad<27>), grid, dim3(64), 0, st, in0, c0, in1, c1, n, V, dz, cout, Kp, vchunk, dw);
    else
        hipLaunchKernelGGL((k_conv_wgrad<1>), grid, dim3(64), 0, st, in0, c0, in1, c1, n, V, dz, cout, Kp, vchunk, dw);
}

inline int round32(int x) { return (x + 31) / 32 * 32; }

// ------------------------------------------------------------------ network
struct Layer {
    int taps, cin0, cin1, cout, Kp, Kpd;
    bool relu;
    long nw() const { return (long)cout * Kp; }
    // parameters and their optimiser state (device)
    float *w = nullptr, *b = nullptr, *gw = nullptr, *gb = nullptr, *mw = nullptr, *sw = nullptr, *mb = nullptr,
          *sb = nullptr;
    uint16_t *wb = nullptr, *wd = nullptr;  // BF16 weights [cout][Kp], dgrad weights [cin][Kpd]
};

struct Acts {
    int n = 0;
    // per tensor: BF16 [V_level][C]
    std::vector<uint16_t*> t;
    float* pred = nullptr;   // [V][16]
    float* tgt = nullptr;    // [V][kOutputs]
    uint8_t* mask = nullptr; // swapped medium
    double* part = nullptr;
    uint16_t* g1 = nullptr;  // gradient scratch, largest tensor
    uint16_t* g2 = nullptr;
    uint16_t* g3 = nullptr;
    uint16_t* dpred = nullptr;
    uint16_t* ds[3] = {nullptr, nullptr, nullptr};  // skip gradients of e0, e1, e2
};

template <typename T>
T* dmalloc(size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    return static_cast<T*>(p);
}
template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// tensor slots of one forward pass
enum Slot {
    X0, E0A, E0, P1, E1A, E1, P2, E2A, E2, P3, B3A, B3, U2, D2A, D2, U1, D1A, D1, U0, D0A, D0, NSLOT
};

struct Net {
    std::vector<Layer> L;
    hipStream_t s = nullptr;
    int step = 0;
    Acts a;
    int chan[NSLOT];
    int lev[NSLOT];

    Net() {
        const int c0 = 16, c1 = 32, c2 = 64, c3 = 128;
        auto add = [&](int taps, int ci0, int ci1, int co, bool relu) {
            Layer l;
            l.taps = taps;
            l.cin0 = ci0;
            l.cin1 = ci1;
            l.cout = co;
            l.relu = relu;
            l.Kp = round32(taps * (ci0 + ci1));
            l.Kpd = round32(taps * co);
            L.push_back(l);
        };
        add(27, 8, 0, c0, true);     // 0  e0a
        add(27, c0, 0, c0, true);    // 1  e0
        add(27, c0, 0, c1, true);    // 2  e1a
        add(27, c1, 0, c1, true);    // 3  e1
        add(27, c1, 0, c2, true);    // 4  e2a
        add(27, c2, 0, c2, true);    // 5  e2
        add(27, c2, 0, c3, true);    // 6  b3a
        add(27, c3, 0, c3, true);    // 7  b3
        add(27, c3, c2, c2, true);   // 8  d2a = conv(cat(u2, e2))
        add(27, c2, 0, c2, true);    // 9  d2
        add(27, c2, c1, c1, true);   // 10 d1a
        add(27, c1, 0, c1, true);    // 11 d1
        add(27, c1, c0, c0, true);   // 12 d0a
        add(27, c0, 0, c0, true);    // 13 d0
        add(1, c0, 0, 16, false);    // 14 head (channels 0..2 used)
        const int ch[NSLOT] = {8, c0, c0, c0, c1, c1, c1, c2, c2, c2, c3, c3, c3, c2, c2, c2, c1, c1, c1, c0, c0};
        const int lv[NSLOT] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 2, 1, 1, 1, 0, 0, 0};
        for (int i = 0; i < NSLOT; ++i) {
            chan[i] = ch[i];
            lev[i] = lv[i];
        }
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) throw std::runtime_error("stream");
        for (Layer& l : L) {
            l.w = dmalloc<float>(l.nw());
            l.b = dmalloc<float>(l.cout);
            l.gw = dmalloc<float>(l.nw());
            l.gb = dmalloc<float>(l.cout);
            l.mw = dmalloc<float>(l.nw());
            l.sw = dmalloc<float>(l.nw());
            l.mb = dmalloc<float>(l.cout);
            l.sb = dmalloc<float>(l.cout);
            l.wb = dmalloc<uint16_t>(l.nw());
            l.wd = dmalloc<uint16_t>((long)(l.cin0 + l.cin1) * l.Kpd);
        }
    }

    ~Net() {
        if (s) (void)hipStreamSynchronize(s);
        for (Layer& l : L) {
            dfree(l.w); dfree(l.b); dfree(l.gw); dfree(l.gb); dfree(l.mw); dfree(l.sw); dfree(l.mb); dfree(l.sb);
            dfree(l.wb); dfree(l.wd);
        }
        free_acts();
        if (s) (void)hipStreamDestroy(s);
    }

    void free_acts() {
        for (auto& p : a.t) dfree(p);
        a.t.clear();
        dfree(a.pred); dfree(a.tgt); dfree(a.mask); dfree(a.part); dfree(a.g1); dfree(a.g2); dfree(a.g3);
        dfree(a.dpred);
        for (auto& p : a.ds) dfree(p);
        a.n = 0;
    }

    static long vox(int n, int level) {
        const long m = n >> level;
        return m * m * m;
    }

    void ensure_acts(int n) {
        if (a.n == n) return;
        free_acts();
        long gmax = 0;
        for (int i = 0; i < NSLOT; ++i) {
            a.t.push_back(dmalloc<uint16_t>(vox(n, lev[i]) * chan[i]));
            gmax = std::max(gmax, vox(n, lev[i]) * chan[i]);
        }
        // concat gradients are the widest per level: 192 * (n/4)^3, 96 * (n/2)^3, 48 n^3
        gmax = std::max({gmax, 48 * vox(n, 0), 96 * vox(n, 1), 192 * vox(n, 2), 16 * vox(n, 0)});
        a.g1 = dmalloc<uint16_t>(gmax);
        a.g2 = dmalloc<uint16_t>(gmax);
        a.g3 = dmalloc<uint16_t>(gmax);
        a.pred = dmalloc<float>(vox(n, 0) * 16);
        a.tgt = dmalloc<float>(vox(n, 0) * kOutputs);
        a.mask = dmalloc<uint8_t>(vox(n, 0));
        a.dpred = dmalloc<uint16_t>(vox(n, 0) * 16);
        a.part = dmalloc<double>(2 * 1024);
        a.ds[0] = dmalloc<uint16_t>(vox(n, 0) * chan[E0]);
        a.ds[1] = dmalloc<uint16_t>(vox(n, 1) * chan[E1]);
        a.ds[2] = dmalloc<uint16_t>(vox(n, 2) * chan[E2]);
        a.n = n;
    }

    long num_params() const {
        long p = 0;
        for (const Layer& l : L) p += (long)l.cout * l.taps * (l.cin0 + l.cin1) + l.cout;
        return p;
    }

    // He-uniform initialisation of the convolutions (head: small), biases 0
    void init(uint64_t seed) {
        std::mt19937_64 rng(seed);
        for (size_t li = 0; li < L.size(); ++li) {
            Layer& l = L[li];
            const int cin = l.cin0 + l.cin1;
            const double fan_in = (double)l.taps * cin;
            const double lim = (li + 1 == L.size()) ? 1e-2 : std::sqrt(6.0 / fan_in);
            std::uniform_real_distribution<double> U(-lim, lim);
            std::vector<float> w(l.nw(), 0.f);
            for (int co = 0; co < l.cout; ++co)
                for (int k = 0; k < l.taps * cin; ++k) w[(long)co * l.Kp + k] = (float)U(rng);
            std::vector<float> b(l.cout, 0.f);
            upload(l, w.data(), b.data());
        }
        step = 0;
    }

    void upload(Layer& l, const float* w, const float* b) {
        check(hipMemcpyAsync(l.w, w, l.nw() * sizeof(float), hipMemcpyHostToDevice, s));
        check(hipMemcpyAsync(l.b, b, l.cout * sizeof(float), hipMemcpyHostToDevice, s));
        check(hipMemsetAsync(l.mw, 0, l.nw() * sizeof(float), s));
        check(hipMemsetAsync(l.sw, 0, l.nw() * sizeof(float), s));
        check(hipMemsetAsync(l.mb, 0, l.cout * sizeof(float), s));
        check(hipMemsetAsync(l.sb, 0, l.cout * sizeof(float), s));
        hipLaunchKernelGGL(k_to_bf16, dim3(ew_grid(l.nw())), dim3(kEw), 0, s, l.w, l.nw(), l.wb);
        refresh_dgrad_weights(l);
        check(hipStreamSynchronize(s));
    }

    void refresh_dgrad_weights(Layer& l) {
        const int cin = l.cin0 + l.cin1;
        hipLaunchKernelGGL(k_flip, dim3(ew_grid((long)cin * l.Kpd)), dim3(kEw), 0, s, l.wb, l.taps, cin, l.cout, l.Kp,
                           l.Kpd, l.wd);
    }

    static void check(hipError_t e) {
        if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
    }

    // Parameters in a flat FP32 array: per layer the weights [cout][taps][cin] then the biases.
    void get_params(float* out) {
        long off = 0;
        for (Layer& l : L) {
            const int cin = l.cin0 + l.cin1;
            std::vector<float> w(l.nw()), b(l.cout);
            check(hipMemcpyAsync(w.data(), l.w, l.nw() * sizeof(float), hipMemcpyDeviceToHost, s));
            check(hipMemcpyAsync(b.data(), l.b, l.cout * sizeof(float), hipMemcpyDeviceToHost, s));
            check(hipStreamSynchronize(s));
            for (int co = 0; co < l.cout; ++co)
                for (int k = 0; k < l.taps * cin; ++k) out[off++] = w[(long)co * l.Kp + k];
            for (int co = 0; co < l.cout; ++co) out[off++] = b[co];
        }
    }

    void set_params(const float* in) {
        long off = 0;
        for (Layer& l : L) {
            const int cin = l.cin0 + l.cin1;
            std::vector<float> w(l.nw(), 0.f), b(l.cout);
            for (int co = 0; co < l.cout; ++co)
                for (int k = 0; k < l.taps * cin; ++k) w[(long)co * l.Kp + k] = in[off++];
            for (int co = 0; co < l.cout; ++co) b[co] = in[off++];
            upload(l, w.data(), b.data());
        }
        step = 0;
    }

    void get_grads(float* out) {
        long off = 0;
        for (Layer& l : L) {
            const int cin = l.cin0 + l.cin1;
            std::vector<float> w(l.nw()), b(l.cout);
            check(hipMemcpyAsync(w.data(), l.gw, l.nw() * sizeof(float), hipMemcpyDeviceToHost, s));
            check(hipMemcpyAsync(b.data(), l.gb, l.cout * sizeof(float), hipMemcpyDeviceToHost, s));
            check(hipStreamSynchronize(s));
            for (int co = 0; co < l.cout; ++co)
                for (int k = 0; k < l.taps * cin; ++k) out[off++] = w[(long)co * l.Kp + k];
            for (int co = 0; co < l.cout; ++co) out[off++] = b[co];
        }
    }

    uint16_t* T(Slot i) { return a.t[i]; }

    void conv(int li, const uint16_t* in0, const uint16_t* in1, int n, void* out, bool f32 = false) {
        const Layer& l = L[li];
        conv_fwd(l.taps, in0, l.cin0, in1, l.cin1, n, l.wb, l.Kp, l.cout, l.b, l.relu, f32, out, s);
    }

    // forward on the medium in a.mask (fluid = 1), edge n (multiple of 8)
    void forward(int n) {
        const long V = vox(n, 0);
        hipLaunchKernelGGL(k_features, dim3(ew_grid(V)), dim3(kEw), 0, s, a.mask, n, V, T(X0));
        const int n1 = n / 2, n2 = n / 4, n3 = n / 8;
        conv(0, T(X0), nullptr, n, T(E0A));
        conv(1, T(E0A), nullptr, n, T(E0));
        pool(T(E0), n, chan[E0], T(P1));
        conv(2, T(P1), nullptr, n1, T(E1A));
        conv(3, T(E1A), nullptr, n1, T(E1));
        pool(T(E1), n1, chan[E1], T(P2));
        conv(4, T(P2), nullptr, n2, T(E2A));
        conv(5, T(E2A), nullptr, n2, T(E2));
        pool(T(E2), n2, chan[E2], T(P3));
        conv(6, T(P3), nullptr, n3, T(B3A));
        conv(7, T(B3A), nullptr, n3, T(B3));
        up(T(B3), n3, chan[B3], T(U2));
        conv(8, T(U2), T(E2), n2, T(D2A));
        conv(9, T(D2A), nullptr, n2, T(D2));
        up(T(D2), n2, chan[D2], T(U1));
        conv(10, T(U1), T(E1), n1, T(D1A));
        conv(11, T(D1A), nullptr, n1, T(D1));
        up(T(D1), n1, chan[D1], T(U0));
        conv(12, T(U0), T(E0), n, T(D0A));
        conv(13, T(D0A), nullptr, n, T(D0));
        conv(14, T(D0), nullptr, n, a.pred, true);
    }

    void pool(const uint16_t* x, int n, int C, uint16_t* y) {
        const long cnt = vox(n, 1) * C;
        hipLaunchKernelGGL(k_pool_fwd, dim3(ew_grid(cnt)), dim3(kEw), 0, s, x, n, C, y);
    }
    void up(const uint16_t* x, int m, int C, uint16_t* y) {
        const long cnt = 8 * (long)m * m * m * C;
        hipLaunchKernelGGL(k_up_fwd, dim3(ew_grid(cnt)), dim3(kEw), 0, s, x, m, C, y);
    }

    // loss (and dpred when grad) against a.tgt; returns the mean squared error over the support
    double loss(int n, bool grad) {
        const long V = vox(n, 0);
        const unsigned blocks = 1024;
        hipLaunchKernelGGL(k_count_support, dim3(blocks), dim3(kEw), 0, s, T(X0), V, a.part);
        std::vector<double> part(2 * blocks);
        check(hipMemcpyAsync(part.data(), a.part, blocks * sizeof(double), hipMemcpyDeviceToHost, s));
        check(hipStreamSynchronize(s));
        double cnt = 0.0;
        for (unsigned i = 0; i < blocks; ++i) cnt += part[i];
        const float inv = (grad && cnt > 0) ? (float)(1.0 / ((double)kOutputs * cnt)) : 0.f;
        hipLaunchKernelGGL(k_loss, dim3(blocks), dim3(kEw), 0, s, a.pred, a.tgt, T(X0), V, inv, a.dpred, a.part);
        check(hipMemcpyAsync(part.data(), a.part, 2 * blocks * sizeof(double), hipMemcpyDeviceToHost, s));
        check(hipStreamSynchronize(s));
        double se = 0.0;
        for (unsigned i = 0; i < blocks; ++i) se += part[2 * i];
        return cnt > 0 ? se / ((double)kOutputs * cnt) : 0.0;
    }

    // gradient of layer li given dY (pre-ReLU-masked inside): accumulates gw, gb; writes dX to
    // dx (when non-null) as [V][cin0 + cin1]
    void layer_bwd(int li, const uint16_t* in0, const uint16_t* in1, const uint16_t* y, const uint16_t* dy, int n,
                   uint16_t* dz, uint16_t* dx) {
        Layer& l = L[li];
        const long V = vox(n, 0);
        const uint16_t* d = dy;
        if (l.relu) {
            hipLaunchKernelGGL(k_relu_bwd, dim3(ew_grid(V * l.cout)), dim3(kEw), 0, s, dy, y, dz, V * l.cout);
            d = dz;
        }
        conv_wgrad(l.taps, in0, l.cin0, in1, l.cin1, n, d, l.cout, l.Kp, l.gw, s);
        const int lanes = std::max(1, 256 / l.cout);
        hipLaunchKernelGGL(k_dbias, dim3((unsigned)std::min<long>(1024, (V + lanes - 1) / lanes)), dim3(lanes * l.cout),
                           0, s, d, V, l.cout, l.gb);
        if (dx != nullptr) {
            const int cin = l.cin0 + l.cin1;
            conv_fwd(l.taps, d, l.cout, nullptr, 0, n, l.wd, l.Kpd, cin, nullptr, false, false, dx, s);
        }
    }

    void zero_grads() {
        for (Layer& l : L) {
            check(hipMemsetAsync(l.gw, 0, l.nw() * sizeof(float), s));
            check(hipMemsetAsync(l.gb, 0, l.cout * sizeof(float), s));
        }
    }

    // full backward from a.dpred; gradients accumulate into gw / gb
    void backward(int n) {
        const int n1 = n / 2, n2 = n / 4, n3 = n / 8;
        uint16_t *g1 = a.g1, *g2 = a.g2, *g3 = a.g3;
        auto split = [&](const uint16_t* d, long V, int C0, int C1, uint16_t* d0, uint16_t* d1) {
            hipLaunchKernelGGL(k_split, dim3(ew_grid(V * (C0 + C1))), dim3(kEw), 0, s, d, V, C0, C1, d0, d1);
        };
        // the decoder concat gives every skip tensor e_l a second gradient contribution (kept in
        // a.ds[l] until the encoder's pool backward adds it); an upsampled tensor U_l is dead once
        // the weight gradient of its consumer has been taken, so its slot holds dU_l
        layer_bwd(14, T(D0), nullptr, nullptr, a.dpred, n, nullptr, g1);  // g1 = dD0 [n^3][16]
        layer_bwd(13, T(D0A), nullptr, T(D0), g1, n, g2, g3);              // g3 = dD0A [n^3][16]
        layer_bwd(12, T(U0), T(E0), T(D0A), g3, n, g2, g1);                // g1 = dcat [n^3][48]
        split(g1, vox(n, 0), chan[U0], chan[E0], T(U0), a.ds[0]);
        hipLaunchKernelGGL(k_up_bwd, dim3(ew_grid(vox(n1, 0) * chan[D1])), dim3(kEw), 0, s, T(U0), n1, chan[D1], g2);
        // g2 = dD1 [(n/2)^3][32]
        layer_bwd(11, T(D1A), nullptr, T(D1), g2, n1, g3, g1);             // g1 = dD1A
        layer_bwd(10, T(U1), T(E1), T(D1A), g1, n1, g3, g2);               // g2 = dcat1 [96]
        split(g2, vox(n1, 0), chan[U1], chan[E1], T(U1), a.ds[1]);
        hipLaunchKernelGGL(k_up_bwd, dim3(ew_grid(vox(n2, 0) * chan[D2])), dim3(kEw), 0, s, T(U1), n2, chan[D2], g1);
        // g1 = dD2 [(n/4)^3][64]
        layer_bwd(9, T(D2A), nullptr, T(D2), g1, n2, g3, g2);              // g2 = dD2A
        layer_bwd(8, T(U2), T(E2), T(D2A), g2, n2, g3, g1);                // g1 = dcat2 [192]
        split(g1, vox(n2, 0), chan[U2], chan[E2], T(U2), a.ds[2]);
        hipLaunchKernelGGL(k_up_bwd, dim3(ew_grid(vox(n3, 0) * chan[B3])), dim3(kEw), 0, s, T(U2), n3, chan[B3], g2);
        // g2 = dB3 [(n/8)^3][128]
        layer_bwd(7, T(B3A), nullptr, T(B3), g2, n3, g3, g1);              // g1 = dB3A
        layer_bwd(6, T(P3), nullptr, T(B3A), g1, n3, g3, g2);              // g2 = dP3 [(n/8)^3][64]
        // dE2 = pool_bwd(dP3) + dskip2
        hipLaunchKernelGGL(k_pool_bwd, dim3(ew_grid(vox(n2, 0) * chan[E2])), dim3(kEw), 0, s, g2, n2, chan[E2], a.ds[2],
                           g1);                                            // g1 = dE2
        layer_bwd(5, T(E2A), nullptr, T(E2), g1, n2, g3, g2);              // g2 = dE2A
        layer_bwd(4, T(P2), nullptr, T(E2A), g2, n2, g3, g1);              // g1 = dP2
        hipLaunchKernelGGL(k_pool_bwd, dim3(ew_grid(vox(n1, 0) * chan[E1])), dim3(kEw), 0, s, g1, n1, chan[E1], a.ds[1],
                           g2);                                            // g2 = dE1
        layer_bwd(3, T(E1A), nullptr, T(E1), g2, n1, g3, g1);              // g1 = dE1A
        layer_bwd(2, T(P1), nullptr, T(E1A), g1, n1, g3, g2);              // g2 = dP1
        hipLaunchKernelGGL(k_pool_bwd, dim3(ew_grid(vox(n, 0) * chan[E0])), dim3(kEw), 0, s, g2, n, chan[E0], a.ds[0],
                           g1);                                            // g1 = dE0
        layer_bwd(1, T(E0A), nullptr, T(E0), g1, n, g3, g2);               // g2 = dE0A
        layer_bwd(0, T(X0), nullptr, T(E0A), g2, n, g3, nullptr);
    }

    void adam(float lr) {
        ++step;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float c1 = 1.f / (1.f - std::pow(b1, (float)step));
        const float c2 = 1.f / (1.f - std::pow(b2, (float)step));
        for (Layer& l : L) {
            hipLaunchKernelGGL(k_adam, dim3(ew_grid(l.nw())), dim3(kEw), 0, s, l.w, l.wb, l.gw, l.mw, l.sw, l.nw(), lr,
                               b1, b2, eps, c1, c2);
            hipLaunchKernelGGL(k_adam, dim3(1), dim3(kEw), 0, s, l.b, (uint16_t*)nullptr, l.gb, l.mb, l.sb,
                               (long)l.cout, lr, b1, b2, eps, c1, c2);
            refresh_dgrad_weights(l);
        }
    }
};

Net* create(uint64_t seed) {
    auto n = std::make_unique<Net>();
    n->init(seed);
    return n.release();
}

void destroy(Net* n) { delete n; }

long num_params(Net* n) { return n->num_params(); }
void get_params(Net* n, float* out) { n->get_params(out); }
void set_params(Net* n, const float* in) { n->set_params(in); }
void get_grads(Net* n, float* out) { n->get_grads(out); }

static void check_edge(int n) {
    if (n < 8 || n % 8 != 0) throw std::invalid_argument("U-Net grid edge must be a positive multiple of 8");
}

void predict(Net* net, const uint8_t* mask_dev, int n) {
    check_edge(n);
    net->ensure_acts(n);
    Net::check(hipMemcpyAsync(net->a.mask, mask_dev, Net::vox(n, 0), hipMemcpyDeviceToDevice, net->s));
    net->forward(n);
    Net::check(hipGetLastError());
}

void pred_field(Net* net, float* out9) {
    const long V = Net::vox(net->a.n, 0);
    hipLaunchKernelGGL(k_pred_field, dim3(ew_grid(V)), dim3(kEw), 0, net->s, net->a.pred, V, out9);
    Net::check(hipStreamSynchronize(net->s));
}

void gather_guess(Net* net, int dir, const int64_t* active, long nb, double* x0) {
    hipLaunchKernelGGL(k_gather_direction, dim3(ew_grid(nb)), dim3(kEw), 0, net->s, net->a.pred, dir, active, nb,
                       x0);
    Net::check(hipStreamSynchronize(net->s));
}

double loss_grad(Net* net, const uint8_t* mask_dev, const float* target_dev, int n, bool update, float lr) {
    check_edge(n);
    net->ensure_acts(n);
    const long V = Net::vox(n, 0);
    Net::check(hipMemcpyAsync(net->a.mask, mask_dev, V, hipMemcpyDeviceToDevice, net->s));
    Net::check(hipMemcpyAsync(net->a.tgt, target_dev, V * kOutputs * sizeof(float), hipMemcpyDeviceToDevice, net->s));
    net->forward(n);
    const double l = net->loss(n, true);
    net->zero_grads();
    net->backward(n);
    if (update) net->adam(lr);
    Net::check(hipGetLastError());
    Net::check(hipStreamSynchronize(net->s));
    return l;
}

hipStream_t stream(Net* n) { return n->s; }

}  // namespace unet
}  // namespace fftperm

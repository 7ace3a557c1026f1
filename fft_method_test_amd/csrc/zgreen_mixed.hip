// zgreen_mixed.hip -- the compile-time specialised fused z passes (zgreen_static.hpp) at the
// mixed-radix column lengths 80, 96, 112, 144 and 160, in their own translation unit so that they
// compile beside the power-of-two set.  The register-resident (radix-4) and persistent (radix <= 8)
// shapes exist for the powers of two only and return -1 here.
#include "zgreen_static.hpp"

namespace fftperm {

bool mixed_length(int n) { return n == 80 || n == 96 || n == 112 || n == 144 || n == 160; }

template <typename R>
long static_zgreen_mixed(int code, cx<R>* S, const R* G6, long ncols, const Geometry3& g, const cx<R>* tw,
                         double* qpart, const int* skip, hipStream_t st, ZBatch zb) {
    switch (g.nz) {
        case 80: return launch_zgreen_n<R, 80>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 96: return launch_zgreen_n<R, 96>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 112: return launch_zgreen_n<R, 112>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 144: return launch_zgreen_n<R, 144>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        case 160: return launch_zgreen_n<R, 160>(code, S, G6, ncols, g, tw, qpart, skip, st, zb);
        default: return -1;
    }
}

// untuned: the widest all-in-LDS tile that fits ~100 KB of LDS, one thread per butterfly
int zgreen_default_code_mixed(int n, int prec) {
    switch (n) {
        case 80: return 10 * (prec ? ZShape<float, 80>::ZL : ZShape<double, 80>::ZL) + 1;
        case 96: return 10 * (prec ? ZShape<float, 96>::ZL : ZShape<double, 96>::ZL) + 1;
        case 112: return 10 * (prec ? ZShape<float, 112>::ZL : ZShape<double, 112>::ZL) + 1;
        case 144: return 10 * (prec ? ZShape<float, 144>::ZL : ZShape<double, 144>::ZL) + 1;
        case 160: return 10 * (prec ? ZShape<float, 160>::ZL : ZShape<double, 160>::ZL) + 1;
        default: return 0;
    }
}

template long static_zgreen_mixed<double>(int, cx<double>*, const double*, long, const Geometry3&, const cx<double>*,
                                          double*, const int*, hipStream_t, ZBatch);
template long static_zgreen_mixed<float>(int, cx<float>*, const float*, long, const Geometry3&, const cx<float>*,
                                         double*, const int*, hipStream_t, ZBatch);

}  // namespace fftperm

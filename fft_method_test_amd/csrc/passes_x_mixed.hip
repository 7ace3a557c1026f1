// passes_x_mixed.hip -- the compile-time specialised x passes (passes_x.hpp) at the mixed-radix
// x-line lengths 80, 96, 112, 144 and 160 (StaticPlan: a radix-16 stage, then radix 5, 6, 7, 9 or
// 10), in their own translation unit so that they compile beside the power-of-two set.  The
// wave-per-line-pair shape does not exist at these lengths (its lanes hold N / 64 elements).
#include "passes_x.hpp"

namespace fftperm {

template <typename R>
long static_xfwd_mixed(int v, const XSource& src, cx<R>* S, const Geometry3& g, const cx<R>* tw, const int* skip,
                       hipStream_t st) {
    switch (g.nx) {
        case 80: return xfwd_n<R, 80>(v, src, S, g, tw, skip, st);
        case 96: return xfwd_n<R, 96>(v, src, S, g, tw, skip, st);
        case 112: return xfwd_n<R, 112>(v, src, S, g, tw, skip, st);
        case 144: return xfwd_n<R, 144>(v, src, S, g, tw, skip, st);
        case 160: return xfwd_n<R, 160>(v, src, S, g, tw, skip, st);
        default: return -1;
    }
}

template <typename R>
long static_xinv_mixed(int v, const cx<R>* S, const XSink& sink, const Geometry3& g, const cx<R>* tw, const int* skip,
                       hipStream_t st) {
    switch (g.nx) {
        case 80: return xinv_n<R, 80>(v, S, sink, g, tw, skip, st);
        case 96: return xinv_n<R, 96>(v, S, sink, g, tw, skip, st);
        case 112: return xinv_n<R, 112>(v, S, sink, g, tw, skip, st);
        case 144: return xinv_n<R, 144>(v, S, sink, g, tw, skip, st);
        case 160: return xinv_n<R, 160>(v, S, sink, g, tw, skip, st);
        default: return -1;
    }
}

int pass_variant_code_mixed(int n, int v) {
    if (v < 0 || v >= kXPassVariants) return 0;
    switch (n) {
        case 80: return code_of<80>(v);
        case 96: return code_of<96>(v);
        case 112: return code_of<112>(v);
        case 144: return code_of<144>(v);
        case 160: return code_of<160>(v);
        default: return 0;
    }
}

template long static_xfwd_mixed<double>(int, const XSource&, cx<double>*, const Geometry3&, const cx<double>*,
                                        const int*, hipStream_t);
template long static_xfwd_mixed<float>(int, const XSource&, cx<float>*, const Geometry3&, const cx<float>*,
                                       const int*, hipStream_t);
template long static_xinv_mixed<double>(int, const cx<double>*, const XSink&, const Geometry3&, const cx<double>*,
                                        const int*, hipStream_t);
template long static_xinv_mixed<float>(int, const cx<float>*, const XSink&, const Geometry3&, const cx<float>*,
                                       const int*, hipStream_t);

}  // namespace fftperm

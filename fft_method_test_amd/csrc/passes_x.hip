// passes_x.hip -- the compile-time specialised x passes (passes_x.hpp) at the power-of-two x-line
// lengths 32, 64, 128, 256 and 512; the mixed-radix lengths are dispatched to passes_x_mixed.hip.
#include "passes_x.hpp"

namespace fftperm {

template <typename R>
long static_xfwd(int v, const XSource& src, cx<R>* S, const Geometry3& g, const cx<R>* tw, const int* skip,
                 hipStream_t st) {
    switch (g.nx) {
        case 32: return xfwd_n<R, 32>(v, src, S, g, tw, skip, st);
        case 64: return xfwd_n<R, 64>(v, src, S, g, tw, skip, st);
        case 128: return xfwd_n<R, 128>(v, src, S, g, tw, skip, st);
        case 256: return xfwd_n<R, 256>(v, src, S, g, tw, skip, st);
        case 512: return xfwd_n<R, 512>(v, src, S, g, tw, skip, st);
        default: return mixed_length(g.nx) ? static_xfwd_mixed<R>(v, src, S, g, tw, skip, st) : -1;
    }
}

template <typename R>
long static_xinv(int v, const cx<R>* S, const XSink& sink, const Geometry3& g, const cx<R>* tw, const int* skip,
                 hipStream_t st) {
    switch (g.nx) {
        case 32: return xinv_n<R, 32>(v, S, sink, g, tw, skip, st);
        case 64: return xinv_n<R, 64>(v, S, sink, g, tw, skip, st);
        case 128: return xinv_n<R, 128>(v, S, sink, g, tw, skip, st);
        case 256: return xinv_n<R, 256>(v, S, sink, g, tw, skip, st);
        case 512: return xinv_n<R, 512>(v, S, sink, g, tw, skip, st);
        default: return mixed_length(g.nx) ? static_xinv_mixed<R>(v, S, sink, g, tw, skip, st) : -1;
    }
}

int pass_variant_code(int n, int v) {
    if (v < 0 || v >= kXPassVariants) return 0;
    switch (n) {
        case 32: return code_of<32>(v);
        case 64: return code_of<64>(v);
        case 128: return code_of<128>(v);
        case 256: return code_of<256>(v);
        case 512: return code_of<512>(v);
        default: return mixed_length(n) ? pass_variant_code_mixed(n, v) : 0;
    }
}

template long static_xfwd<double>(int, const XSource&, cx<double>*, const Geometry3&, const cx<double>*, const int*,
                                  hipStream_t);
template long static_xfwd<float>(int, const XSource&, cx<float>*, const Geometry3&, const cx<float>*, const int*,
                                 hipStream_t);
template long static_xinv<double>(int, const cx<double>*, const XSink&, const Geometry3&, const cx<double>*,
                                  const int*, hipStream_t);
template long static_xinv<float>(int, const cx<float>*, const XSink&, const Geometry3&, const cx<float>*, const int*,
                                 hipStream_t);

}  // namespace fftperm

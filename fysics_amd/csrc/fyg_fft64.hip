// fyg_fft64.hip -- the fp64 instantiation of the deconvolution passes
// (fyg_fft_engine.hpp): the all-FP64 variant of deconv_radial
// (paganin.c:152-293) with the LUT weights evaluated in double, for the
// precision study and its speed comparison.
#include "fyg_fft_engine.hpp"

extern "C" int fyg_deconv_run_batch_f64(const fyg_deconv_args *a, int n, fyg_stream *s) {
    return deconv_batch<double>(a, n, s);
}

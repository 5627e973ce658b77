// FP32 (float2) instantiations of the pencil FFT passes.
#include "fft_pass.hpp"

int lg::launch_f32(const hgrid_pass *p, long L, int kind) { return lg::launch_prec<float2>(p, L, kind); }

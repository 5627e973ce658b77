// FP64 (double2) instantiations of the pencil FFT passes.
#include "fft_pass.hpp"

int lg::launch_f64(const hgrid_pass *p, long L, int kind) { return lg::launch_prec<double2>(p, L, kind); }

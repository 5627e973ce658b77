// Pencil FFT passes for FP32 storage with FP64 arithmetic: float2 loads are widened to
// double2, every butterfly, twiddle and phase factor is FP64, and only the final store
// rounds.  The rounding then follows the data, so the FP32 grid has no fixed per-element
// error in its transforms (the plain FP32 passes round twiddles and DFT constants the same
// way every step).
#include "fft_pass.hpp"

int lg::launch_f32s_f64(const hgrid_pass *p, long L, int kind) { return lg::launch_prec<double2, float2>(p, L, kind); }

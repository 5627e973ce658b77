// C-ABI of the pencil FFT passes (see fft_pass.hpp for the kernels).
#include "fft_pass.hpp"

extern "C" int hgrid_fft_supported(long n) {
  if (n >= 2 && n <= 1024 && (n & (n - 1)) == 0) return 1;
  switch (n) {
    case 3: case 6: case 12: case 24: case 48: case 96: case 192: case 384: return 1;
    default: return 0;
  }
}

static long axis_len(long n0, long n1, long n2, int axis) { return axis == 0 ? n0 : axis == 1 ? n1 : n2; }

extern "C" long hgrid_pass_blocks(int prec, long n0, long n1, long n2, int axis) {
  const long L = axis_len(n0, n1, n2, axis);
  // the workgroup geometry follows the arithmetic type (double2 for FP32 storage with FP64 arithmetic)
  const bool contig = axis == 2;
  const long P = prec == HGRID_C64 ? (contig ? lg::pencils_per_block<float2, true>(L, axis) : lg::pencils_per_block<float2, false>(L, axis))
                                   : (contig ? lg::pencils_per_block<double2, true>(L, axis) : lg::pencils_per_block<double2, false>(L, axis));
  if (P <= 0) return 0;
  const long npencils = (n0 * n1 * n2) / L;
  return (npencils + P - 1) / P;
}

extern "C" int hgrid_fft_pass(const hgrid_pass *p) {
  if (!p || p->axis < 0 || p->axis > 2) return -1;
  const long L = axis_len(p->n0, p->n1, p->n2, p->axis);
  if (!hgrid_fft_supported(L)) return -1;
  // a flat element offset must fit the 32-bit buffer offsets
  const long ntot = p->n0 * p->n1 * p->n2;
  const long esize = p->prec == HGRID_C128 ? 16 : 8;
  if (ntot * esize > 0xffffffffL) return -1;
  int kind;
  if (p->flags & HGRID_PF_MID_KIN) {
    kind = (p->flags & HGRID_PF_KIN_REAL) ? lg::KIND_KIR : lg::KIND_KIN;
    // FP32 real factors at the lengths where the kernel with both factor paths runs better
    if (kind == lg::KIND_KIR && p->prec == HGRID_C64 && lg::kib_length(L)) kind = lg::KIND_KIB;
  }
  else if (p->flags & HGRID_PF_MID_POT) kind = lg::KIND_POT;
  else if (p->flags & HGRID_PF_MID_GPE) kind = p->c_im == 0.0 ? lg::KIND_GPR : lg::KIND_GPE;
  else kind = lg::KIND_FFT;
  // the read-only start-of-step pass: inverse transform, then density / V_eff / reductions
  constexpr unsigned PRE_ONLY = HGRID_PF_PRE_AFTER | HGRID_PF_NOSTORE;
  constexpr unsigned PRE_KEEP = HGRID_PF_PRESCALE | HGRID_PF_PRE_RHO | HGRID_PF_PRE_RED | HGRID_PF_PRE_VEFF;
  if (kind == lg::KIND_FFT && p->axis == 2 && (p->flags & PRE_ONLY) == PRE_ONLY && !(p->flags & ~(PRE_ONLY | PRE_KEEP)))
    kind = lg::KIND_PRE;
  // a plain transform (pending scale, post factor and output norm at most)
  if (kind == lg::KIND_FFT && !(p->flags & ~(HGRID_PF_PRESCALE | HGRID_PF_POST_RED))) kind = lg::KIND_FY;
  const unsigned skip = p->flags & (HGRID_PF_POT_NOFFT1 | HGRID_PF_POT_NOFFT2);
  if (skip && (kind != lg::KIND_POT || skip == (HGRID_PF_POT_NOFFT1 | HGRID_PF_POT_NOFFT2))) return -1;
  const bool one = kind == lg::KIND_FFT || kind == lg::KIND_PRE || kind == lg::KIND_FY;
  if (one && p->nfft != 1) return -1;
  if (!one && p->nfft != (skip ? 1 : 2)) return -1;
  if (p->prec == HGRID_C128) return lg::launch_f64(p, L, kind);
  if (p->prec == HGRID_C64) return lg::launch_f32(p, L, kind);
  if (p->prec == HGRID_C64X) return lg::launch_f32s_f64(p, L, kind);
  return -1;
}

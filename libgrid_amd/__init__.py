"""Python binding of libgrid.so (the gfx950 build of libgrid's FFT split-operator path).

The product is the C library ``libgrid_amd/lib/libgrid.so`` with the C API of
``include/grid.h``; this module only loads it with ctypes, declares the prototypes and
offers numpy helpers to move grid values in and out.  It performs no computation itself:
if the library is missing, :func:`load` raises.

Reference API mirrored: src/grid_wf.c, src/grid_wf_fft.c, src/cgrid.c, src/rgrid.c,
src/ngrid.c (function names, argument order and meaning are the reference's).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "lib", "libgrid.so")

WF_DIRICHLET_BOUNDARY = 1
WF_NEUMANN_BOUNDARY = 2
WF_PERIODIC_BOUNDARY = 3
WF_2ND_ORDER_FFT = 1
WF_4TH_ORDER_FFT = 2
WF_2ND_ORDER_CFFT = 3
WF_4TH_ORDER_CFFT = 4
WF_2ND_ORDER_CN = 5
WF_4TH_ORDER_CN = 6
WF_1ST_ORDER_EULER = 7
GRID_PREC_DOUBLE = 0
GRID_PREC_SINGLE = 1


class c_complex(C.Structure):
    """C99 ``double _Complex`` passed or returned by value (two SSE doubles on x86-64)."""

    _fields_ = [("re", C.c_double), ("im", C.c_double)]

    @classmethod
    def of(cls, z) -> "c_complex":
        z = complex(z)
        return cls(z.real, z.imag)

    def value(self) -> complex:
        return complex(self.re, self.im)


class cgrid(C.Structure):
    """Layout of ``cgrid`` in include/grid.h."""

    _fields_ = [
        ("value", C.c_void_p),
        ("id", C.c_char * 32),
        ("grid_len", C.c_size_t),
        ("nx", C.c_long),
        ("ny", C.c_long),
        ("nz", C.c_long),
        ("step", C.c_double),
        ("x0", C.c_double),
        ("y0", C.c_double),
        ("z0", C.c_double),
        ("kx0", C.c_double),
        ("ky0", C.c_double),
        ("kz0", C.c_double),
        ("omega", C.c_double),
        ("value_outside", C.c_void_p),
        ("outside_params_ptr", C.c_void_p),
        ("default_outside_params", c_complex),
        ("host_lock", C.c_char),
        ("fft_norm", C.c_double),
        ("fft_norm2", C.c_double),
        ("flag", C.c_char),
        ("dev", C.c_void_p),
    ]


class rgrid(C.Structure):
    """Layout of ``rgrid`` in include/grid.h."""

    _fields_ = [
        ("value", C.c_void_p),
        ("id", C.c_char * 32),
        ("grid_len", C.c_size_t),
        ("nx", C.c_long),
        ("ny", C.c_long),
        ("nz", C.c_long),
        ("nz2", C.c_long),
        ("step", C.c_double),
        ("x0", C.c_double),
        ("y0", C.c_double),
        ("z0", C.c_double),
        ("kx0", C.c_double),
        ("ky0", C.c_double),
        ("kz0", C.c_double),
        ("value_outside", C.c_void_p),
        ("outside_params_ptr", C.c_void_p),
        ("default_outside_params", C.c_double),
        ("host_lock", C.c_char),
        ("fft_norm", C.c_double),
        ("fft_norm2", C.c_double),
        ("flag", C.c_char),
        ("dev", C.c_void_p),
    ]


class wf(C.Structure):
    """Layout of ``wf`` in include/grid.h."""

    _fields_ = [
        ("grid", C.POINTER(cgrid)),
        ("mass", C.c_double),
        ("norm", C.c_double),
        ("cfft_width", C.c_double),
        ("boundary", C.c_char),
        ("propagator", C.c_char),
        ("cworkspace", C.POINTER(cgrid)),
        ("cworkspace2", C.POINTER(cgrid)),
        ("cworkspace3", C.POINTER(cgrid)),
        ("ts_func", C.c_void_p),
        ("lx", C.c_long),
        ("hx", C.c_long),
        ("ly", C.c_long),
        ("hy", C.c_long),
        ("lz", C.c_long),
        ("hz", C.c_long),
        ("kmax", C.c_double),
    ]


class hgrid_pass(C.Structure):
    """Layout of ``hgrid_pass`` in include/grid_hip.h (one pencil-FFT pass)."""

    _fields_ = [
        ("in_", C.c_void_p),
        ("out", C.c_void_p),
        ("prec", C.c_int),
        ("n0", C.c_long),
        ("n1", C.c_long),
        ("n2", C.c_long),
        ("axis", C.c_int),
        ("nfft", C.c_int),
        ("dir1", C.c_int),
        ("dir2", C.c_int),
        ("flags", C.c_uint),
        ("tw", C.c_void_p),
        ("scale", C.c_void_p),
        ("post", C.c_double),
        ("mid", C.c_double),
        ("ktab_a", C.c_void_p),
        ("ktab_1", C.c_void_p),
        ("ktab_2", C.c_void_p),
        ("k2_a", C.c_void_p),
        ("k2_1", C.c_void_p),
        ("k2_2", C.c_void_p),
        ("kc_a", C.c_void_p),
        ("kc_1", C.c_void_p),
        ("kc_2", C.c_void_p),
        ("kcut2", C.c_double),
        ("pot", C.c_void_p),
        ("vext", C.c_void_p),
        ("rho", C.c_void_p),
        ("g", C.c_double),
        ("c_re", C.c_double),
        ("c_im", C.c_double),
        ("cons", C.c_double),
        ("rho_out", C.c_void_p),
        ("partials", C.c_void_p),
        ("k2_dk", C.c_double),
    ]


PF_PRESCALE = 0x0001
PF_PRE_RHO = 0x0002
PF_PRE_RED = 0x0004
PF_MID_KIN = 0x0008
PF_MID_KRED = 0x0010
PF_MID_POT = 0x0020
PF_MID_GPE = 0x0040
PF_POST_RED = 0x0080
PF_NOSTORE = 0x0100
PF_PRE_VEFF = 0x0200
PF_KIN_REAL = 0x0400
PF_PRE_AFTER = 0x0800
HGRID_C128 = 0
HGRID_C64 = 1
HGRID_C64X = 2

CG = C.POINTER(cgrid)
RG = C.POINTER(rgrid)
WF = C.POINTER(wf)
D = C.c_double
L = C.c_long
V = None
VP = C.c_void_p
CHR = C.c_char
CX = c_complex

# name: (restype, argtypes)
PROTOTYPES = {
    "grid_threads_init": (V, [L]),
    "grid_threads": (L, []),
    "grid_hip_enable": (V, [CHR, C.c_int, C.POINTER(C.c_int)]),
    "grid_hip_status": (CHR, []),
    "grid_hip_sync": (V, []),
    "cgrid_alloc": (CG, [L, L, L, D, VP, VP, C.c_char_p]),
    "cgrid_clone": (CG, [CG, C.c_char_p]),
    "cgrid_free": (V, [CG]),
    "cgrid_claim": (V, [CG]),
    "cgrid_release": (V, [CG]),
    "cgrid_set_origin": (V, [CG, D, D, D]),
    "cgrid_shift_origin": (V, [CG, D, D, D]),
    "cgrid_set_momentum": (V, [CG, D, D, D]),
    "cgrid_set_rotation": (V, [CG, D]),
    "cgrid_map": (V, [CG, VP, VP]),
    "cgrid_smooth_map": (V, [CG, VP, VP, L]),
    "cgrid_value": (CX, [CG, D, D, D]),
    "cgrid_extrapolate": (V, [CG, CG]),
    "cgrid_write": (V, [CG, VP]),
    "cgrid_read": (CG, [CG, VP]),
    "cgrid_write_grid": (V, [C.c_char_p, CG]),
    "cgrid_zero": (V, [CG]),
    "cgrid_constant": (V, [CG, CX]),
    "cgrid_copy": (V, [CG, CG]),
    "cgrid_multiply": (V, [CG, CX]),
    "cgrid_integral_of_square": (D, [CG]),
    "cgrid_integral_of_conjugate_product": (CX, [CG, CG]),
    "cgrid_fft": (V, [CG]),
    "cgrid_inverse_fft": (V, [CG]),
    "cgrid_scaled_inverse_fft": (V, [CG, CX]),
    "cgrid_inverse_fft_norm": (V, [CG]),
    "cgrid_inverse_fft_norm2": (V, [CG]),
    "cgrid_value_at_index": (CX, [CG, L, L, L]),
    "cgrid_value_to_index": (V, [CG, L, L, L, CX]),
    "cgrid_fft_laplace_expectation_value": (D, [CG]),
    "cgrid_add_scaled": (V, [CG, CX, CG]),
    "cgrid_set_precision": (V, [CG, CHR]),
    "cgrid_get_precision": (CHR, [CG]),
    "cgrid_host_sync": (V, [CG]),
    "cgrid_host_modified": (V, [CG]),
    "rgrid_alloc": (RG, [L, L, L, D, VP, VP, C.c_char_p]),
    "rgrid_free": (V, [RG]),
    "rgrid_map": (V, [RG, VP, VP]),
    "rgrid_smooth_map": (V, [RG, VP, VP, L]),
    "rgrid_value": (D, [RG, D, D, D]),
    "rgrid_extrapolate": (V, [RG, RG]),
    "rgrid_write": (V, [RG, VP]),
    "rgrid_read": (RG, [RG, VP]),
    "rgrid_write_grid": (V, [C.c_char_p, RG]),
    "rgrid_zero": (V, [RG]),
    "rgrid_constant": (V, [RG, D]),
    "rgrid_value_at_index": (D, [RG, L, L, L]),
    "rgrid_value_to_index": (V, [RG, L, L, L, D]),
    "rgrid_integral": (D, [RG]),
    "rgrid_integral_of_square": (D, [RG]),
    "rgrid_host_sync": (V, [RG]),
    "rgrid_host_modified": (V, [RG]),
    "grid_grid_expectation_value": (D, [CG, RG]),
    "grid_real_to_complex_re": (V, [CG, RG]),
    "grid_wf_alloc": (WF, [L, L, L, D, D, CHR, CHR, C.c_char_p]),
    "grid_wf_clone": (WF, [WF, C.c_char_p]),
    "grid_wf_free": (V, [WF]),
    "grid_wf_propagate": (V, [WF, CG, CX]),
    "grid_wf_propagate_potential": (V, [WF, CX, CG, D]),
    "grid_wf_propagate_kinetic_fft": (V, [WF, CX]),
    "grid_wf_propagate_kinetic_cfft": (V, [WF, CX]),
    "grid_wf_boundary": (V, [WF, WF, L, L, L, L, L, L]),
    "grid_wf_absorb": (D, [L, L, L, L, L, L, L, L, L]),
    "grid_wf_merge": (V, [WF, CG, CG]),
    "grid_wf_density": (V, [WF, RG]),
    "grid_wf_zero": (V, [WF]),
    "grid_wf_constant": (V, [WF, CX]),
    "grid_wf_map": (V, [WF, VP, VP]),
    "grid_wf_norm": (D, [WF]),
    "grid_wf_normalize": (D, [WF]),
    "grid_wf_overlap": (CX, [WF, WF]),
    "grid_wf_energy": (D, [WF, RG]),
    "grid_wf_kinetic_energy": (D, [WF]),
    "grid_wf_potential_energy": (D, [WF, RG]),
    "grid_wf_energy_fft": (D, [WF, RG]),
    "grid_wf_kinetic_energy_fft": (D, [WF]),
    "grid_wf_energy_cn": (D, [WF, RG]),
    "grid_wf_kinetic_energy_cn": (D, [WF]),
    "grid_wf_analyze_method": (V, [CHR]),
    "grid_wf_gpe_propagate": (V, [WF, RG, D, CX, CHR]),
    "grid_wf_gpe_last_energy": (D, [WF]),
    "grid_wf_gpe_energy": (D, [WF, RG, D]),
    "grid_wf_gpe_relax": (L, [WF, RG, D, D, D, L, C.POINTER(D)]),
    "grid_wf_gpe_relax_extrapolated": (L, [WF, RG, D, D, D, L, C.POINTER(C.c_double), L, C.POINTER(L)]),
    "grid_wf_batch2d_relax": (D, [L, C.POINTER(WF), C.POINTER(RG), D, D, L, C.POINTER(D), C.POINTER(L)]),
    "grid_batch2d_autotune": (C.c_char_p, [L, C.POINTER(WF), C.POINTER(RG), D, L]),
    "grid_batch2d_solve_times": (L, [C.POINTER(D), C.POINTER(D)]),
    "grid_batch2d_set_straggler_cap": (V, [L]),
    "grid_batch2d_set_extrapolation": (V, [L]),
    "grid_batch2d_last_extrapolations": (L, []),
    "grid_hip_graphs": (V, [C.c_int]),
    "grid_hip_merged_relaxation": (V, [C.c_int]),
    "grid_hip_pass_timing": (V, [C.c_int]),
    "grid_hip_pass_times": (L, [L, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_uint), C.POINTER(L),
                                C.POINTER(C.c_int), C.POINTER(D)]),
    "grid_batch2d_set_straggler_extrapolation": (V, [L]),
    "grid_batch2d_last_phase2": (V, [C.POINTER(L), C.POINTER(D)]),
    "grid_surrogate_alloc": (C.c_void_p, [L, L, D, L]),
    "grid_surrogate_free": (V, [C.c_void_p]),
    "grid_surrogate_nparams": (L, [C.c_void_p]),
    "grid_surrogate_side": (L, [C.c_void_p]),
    "grid_surrogate_set_params": (V, [C.c_void_p, C.POINTER(D)]),
    "grid_surrogate_get_params": (V, [C.c_void_p, C.POINTER(D)]),
    "grid_surrogate_gradient": (D, [C.c_void_p, L, C.POINTER(D), C.POINTER(D), C.POINTER(D)]),
    "grid_surrogate_loss": (D, [C.c_void_p, L, C.POINTER(D), C.POINTER(D)]),
    "grid_surrogate_predict": (V, [C.c_void_p, L, C.POINTER(D), C.POINTER(D)]),
    "grid_surrogate_set_training_data": (V, [C.c_void_p, L, C.POINTER(D), C.POINTER(D)]),
    "grid_surrogate_train": (D, [C.c_void_p, L, D, L, C.POINTER(D)]),
    "grid_surrogate_train_ex": (D, [C.c_void_p, L, D, L, C.c_int, C.POINTER(D)]),
    "grid_surrogate_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "grid_surrogate_load": (C.c_void_p, [C.c_char_p, L]),
    "grid_wf_warm_start": (V, [WF, RG, C.c_void_p, D]),
    "grid_wf_gaussian_start": (V, [WF, D]),
    "grid_batch2d_record_history": (V, [L]),
    "grid_batch2d_history": (L, [L, C.POINTER(D), L]),
    "grid_batch2d_set_tile": (V, [C.c_int, C.c_int, C.c_int]),
    "grid_batch2d_get_tile": (V, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    # thin C-ABI (grid_hip.h) used for timing
    "hgrid_init": (C.c_int, [C.c_int]),
    "hgrid_last_error": (C.c_char_p, []),
    "hgrid_device_name": (C.c_char_p, []),
    "hgrid_sync": (C.c_int, []),
    "hgrid_event_create": (C.c_int, [C.POINTER(VP)]),
    "hgrid_event_destroy": (C.c_int, [VP]),
    "hgrid_event_record": (C.c_int, [VP]),
    "hgrid_event_elapsed_ms": (C.c_int, [VP, VP, C.POINTER(C.c_float)]),
    "hgrid_mem_info": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "hgrid_malloc": (C.c_int, [C.POINTER(VP), C.c_size_t]),
    "hgrid_free": (C.c_int, [VP]),
    "hgrid_nn_symgather": (C.c_int, [VP, VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "hgrid_b2d_cus": (C.c_int, []),
    "hgrid_memcpy_h2d": (C.c_int, [VP, VP, C.c_size_t]),
    "hgrid_memcpy_d2h": (C.c_int, [VP, VP, C.c_size_t]),
    "hgrid_memcpy_d2d": (C.c_int, [VP, VP, C.c_size_t]),
    "hgrid_fft_pass": (C.c_int, [C.POINTER(hgrid_pass)]),
    "hgrid_stream_copy": (C.c_int, [VP, VP, C.c_size_t]),
    "hgrid_pass_blocks": (L, [C.c_int, L, L, L, C.c_int]),
    "hgrid_reduce_scratch": (L, [L]),
    "hgrid_sum_norm2": (C.c_int, [VP, VP, C.c_int, L, VP, VP]),
    "hgrid_sum_norm2_noalloc": (C.c_int, [VP, VP, C.c_int, L, VP]),
}


def build(quiet: bool = True) -> None:
    """Compile libgrid.so (hipcc for gfx950 + gcc) with libgrid_amd/Makefile."""
    cmd = ["make", "-C", HERE, "-j8"]
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("libgrid build failed:\n" + out.stdout[-4000:] + out.stderr[-4000:])
    if not quiet:
        print(out.stdout)


_lib = None


def load() -> C.CDLL:
    """Load libgrid.so and declare its prototypes.  Raises if the library was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `make -C libgrid_amd` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    for name, (res, args) in PROTOTYPES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def boundary(name: str) -> int:
    """Address of a reference boundary function (CGRID_PERIODIC_BOUNDARY etc.)."""
    lib = load()
    return C.cast(getattr(lib, name), C.c_void_p).value


def set_cgrid(g, a: np.ndarray) -> None:
    """Copy a complex128 [nx][ny][nz] array into grid->value and mark the host copy current."""
    lib = load()
    c = g.contents
    a = np.ascontiguousarray(a, dtype=np.complex128).reshape(c.nx, c.ny, c.nz)
    C.memmove(c.value, a.ctypes.data, a.nbytes)
    lib.cgrid_host_modified(g)


def get_cgrid(g) -> np.ndarray:
    """Copy grid values (device copy first, if newer) to a new complex128 array."""
    lib = load()
    lib.cgrid_host_sync(g)
    c = g.contents
    n = c.nx * c.ny * c.nz
    buf = (C.c_double * (2 * n)).from_address(c.value)
    return np.frombuffer(buf, dtype=np.complex128).reshape(c.nx, c.ny, c.nz).copy()


def set_rgrid(g, a: np.ndarray) -> None:
    """Copy a float64 [nx][ny][nz] array into the nz2-padded rows of grid->value."""
    lib = load()
    c = g.contents
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(c.nx, c.ny, c.nz)
    n = c.nx * c.ny * c.nz2
    host = np.frombuffer((C.c_double * n).from_address(c.value), dtype=np.float64).reshape(c.nx, c.ny, c.nz2)
    host[:, :, : c.nz] = a
    host[:, :, c.nz :] = 0.0
    lib.rgrid_host_modified(g)


def get_rgrid(g) -> np.ndarray:
    lib = load()
    lib.rgrid_host_sync(g)
    c = g.contents
    n = c.nx * c.ny * c.nz2
    host = np.frombuffer((C.c_double * n).from_address(c.value), dtype=np.float64).reshape(c.nx, c.ny, c.nz2)
    return host[:, :, : c.nz].copy()


class HipTimer:
    """HIP events recorded on the library stream (the stream every libgrid kernel runs on)."""

    def __init__(self):
        lib = load()
        self.lib = lib
        self.a = C.c_void_p()
        self.b = C.c_void_p()
        if lib.hgrid_event_create(C.byref(self.a)) or lib.hgrid_event_create(C.byref(self.b)):
            raise RuntimeError(lib.hgrid_last_error().decode())

    def start(self):
        self.lib.hgrid_event_record(self.a)

    def stop(self) -> float:
        self.lib.hgrid_event_record(self.b)
        ms = C.c_float()
        if self.lib.hgrid_event_elapsed_ms(self.a, self.b, C.byref(ms)):
            raise RuntimeError(self.lib.hgrid_last_error().decode())
        return float(ms.value)

    def close(self):
        self.lib.hgrid_event_destroy(self.a)
        self.lib.hgrid_event_destroy(self.b)

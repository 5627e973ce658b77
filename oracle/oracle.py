"""ORACLE — test infrastructure only.

ctypes binding of ``oracle/liboracle.so`` (the C/OpenMP restatement of the reference hot
path in ``oracle/oracle.c``).  Imported by tests/, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py`` — as the checker or the CPU baseline, never as the thing
measured or shipped.  Arrays are numpy complex128 / float64 in row-major [nx][ny][nz].
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None
P = C.c_void_p
D = C.c_double
L = C.c_long
I = C.c_int

_PROTOS = {
    "or_fft3d": (None, [P, L, L, L, I]),
    "or_cgrid_fft": (None, [P, L, L, L]),
    "or_cgrid_inverse_fft": (None, [P, L, L, L]),
    "or_cgrid_inverse_fft_norm": (None, [P, L, L, L]),
    "or_cgrid_multiply": (None, [P, L, D, D]),
    "or_integral_of_square": (D, [P, L, L, L, D]),
    "or_integral_of_conjugate_product": (None, [P, P, L, L, L, D, P]),
    "or_fft_laplace_expectation_value": (D, [P, L, L, L, D]),
    "or_grid_expectation_value": (D, [P, P, L, L, L, D]),
    "or_propagate_kinetic_fft": (None, [P, L, L, L, D, D, D, D, D, D, D]),
    "or_propagate_potential": (None, [P, P, L, D, D, D]),
    "or_wf_propagate": (None, [P, P, L, L, L, D, D, D, D, D, D, D]),
    "or_propagate_kinetic_cfft": (None, [P, L, L, L, D, D, D, D, D, D, D, D]),
    "or_wf_propagate_cfft": (None, [P, P, L, L, L, D, D, D, D, D, D, D, D]),
    "or_wf_propagate_4th": (None, [P, P, L, L, L, D, D, D, D, D, D, D, D]),
    "or_wf_absorb": (D, [L, L, L, L, L, L, L, L, L]),
    "or_wf_merge": (None, [P, P, P, L, L, L, L, L, L, L, L, L]),
    "or_wf_propagate_absorb": (None, [P, P, L, L, L, D, D, D, D, D, D, D, D, L, L, L, L, L, L]),
    "or_wf_normalize": (D, [P, L, L, L, D, D]),
    "or_kinetic_energy_fft": (D, [P, L, L, L, D, D]),
    "or_energy_fft": (D, [P, P, L, L, L, D, D]),
    "or_kinetic_energy_fd_periodic": (D, [P, L, L, L, D, D]),
    "or_kinetic_energy_fd": (D, [P, L, L, L, D, D, I, D, D]),
    "or_gpe_step": (None, [P, P, D, L, L, L, D, D, D, D, D, D, D, I, D, D]),
    "or_gpe_energy": (D, [P, P, D, L, L, L, D, D]),
    "or_num_threads": (I, []),
    "or_set_num_threads": (None, [I]),
}


def build() -> None:
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout[-3000:] + out.stderr[-3000:])


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _c(a: np.ndarray) -> np.ndarray:
    if a.dtype != np.complex128 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("oracle arrays must be C-contiguous complex128")
    return a


def _r(a: np.ndarray) -> np.ndarray:
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("oracle arrays must be C-contiguous float64")
    return a


def fft3d(a: np.ndarray, sign: int) -> np.ndarray:
    """Unnormalised 3-D DFT of a copy (sign -1 forward, +1 backward)."""
    b = _c(np.array(a, dtype=np.complex128, order="C", copy=True))
    nx, ny, nz = b.shape
    lib().or_fft3d(b.ctypes.data, nx, ny, nz, sign)
    return b


def inverse_fft_norm(a: np.ndarray) -> np.ndarray:
    b = _c(np.array(a, dtype=np.complex128, order="C", copy=True))
    lib().or_cgrid_inverse_fft_norm(b.ctypes.data, *b.shape)
    return b


def integral_of_square(a, step) -> float:
    a = _c(a)
    return lib().or_integral_of_square(a.ctypes.data, *a.shape, step)


def integral_of_conjugate_product(a, b, step) -> complex:
    a, b = _c(a), _c(b)
    out = np.zeros(2)
    lib().or_integral_of_conjugate_product(a.ctypes.data, b.ctypes.data, *a.shape, step, out.ctypes.data)
    return complex(out[0], out[1])


def laplace_expectation(ak, step) -> float:
    ak = _c(ak)
    return lib().or_fft_laplace_expectation_value(ak.ctypes.data, *ak.shape, step)


def grid_expectation(a, v, step) -> float:
    a, v = _c(a), _r(v)
    return lib().or_grid_expectation_value(a.ctypes.data, v.ctypes.data, *a.shape, step)


def propagate_kinetic_fft(ak, step, mass, k0, time) -> np.ndarray:
    b = _c(np.array(ak, dtype=np.complex128, order="C", copy=True))
    t = complex(time)
    lib().or_propagate_kinetic_fft(b.ctypes.data, *b.shape, step, mass, k0[0], k0[1], k0[2], t.real, t.imag)
    return b


def propagate_kinetic_cfft(ak, step, mass, k0, time, kmax) -> np.ndarray:
    b = _c(np.array(ak, dtype=np.complex128, order="C", copy=True))
    t = complex(time)
    lib().or_propagate_kinetic_cfft(b.ctypes.data, *b.shape, step, mass, k0[0], k0[1], k0[2], t.real, t.imag, kmax)
    return b


def propagate_potential(a, pot, tstep, cons) -> np.ndarray:
    b = _c(np.array(a, dtype=np.complex128, order="C", copy=True))
    pot = _c(np.ascontiguousarray(pot, dtype=np.complex128))
    t = complex(tstep)
    lib().or_propagate_potential(b.ctypes.data, pot.ctypes.data, b.size, t.real, t.imag, cons)
    return b


def wf_propagate(a, pot, step, mass, k0, time) -> np.ndarray:
    """One reference grid_wf_propagate (WF_2ND_ORDER_FFT) step on a copy; pot may be None."""
    b = _c(np.array(a, dtype=np.complex128, order="C", copy=True))
    t = complex(time)
    pp = None
    if pot is not None:
        pot = _c(np.ascontiguousarray(pot, dtype=np.complex128))
        pp = pot.ctypes.data
    lib().or_wf_propagate(b.ctypes.data, pp, *b.shape, step, mass, k0[0], k0[1], k0[2], t.real, t.imag)
    return b


def wf_propagate_inplace(b: np.ndarray, pot, step, mass, k0, time) -> None:
    t = complex(time)
    pp = None if pot is None else _c(pot).ctypes.data
    lib().or_wf_propagate(_c(b).ctypes.data, pp, *b.shape, step, mass, k0[0], k0[1], k0[2], t.real, t.imag)


def wf_propagate_cfft_inplace(b: np.ndarray, pot, step, mass, k0, time, kmax) -> None:
    """One reference grid_wf_propagate (WF_2ND_ORDER_CFFT) step in place; pot may be None."""
    t = complex(time)
    pp = None if pot is None else _c(pot).ctypes.data
    lib().or_wf_propagate_cfft(_c(b).ctypes.data, pp, *b.shape, step, mass, k0[0], k0[1], k0[2], t.real, t.imag, kmax)


def wf_propagate_4th_inplace(b: np.ndarray, pot, step, mass, k0, time, kmax=None) -> None:
    """One reference grid_wf_propagate WF_4TH_ORDER_FFT step in place (WF_4TH_ORDER_CFFT when
    kmax is given); pot may be None."""
    t = complex(time)
    pp = None if pot is None else _c(pot).ctypes.data
    lib().or_wf_propagate_4th(_c(b).ctypes.data, pp, *b.shape, step, mass, k0[0], k0[1], k0[2], t.real, t.imag,
                              -1.0 if kmax is None else kmax)


def wf_absorb(i, j, k, bounds) -> float:
    """grid_wf_absorb at (i, j, k); bounds = (lx, hx, ly, hy, lz, hz)."""
    return lib().or_wf_absorb(i, j, k, *bounds)


def wf_merge(wr, wi, bounds) -> np.ndarray:
    wr, wi = _c(wr), _c(wi)
    out = np.empty_like(wr)
    lib().or_wf_merge(out.ctypes.data, wr.ctypes.data, wi.ctypes.data, *wr.shape, *bounds)
    return out


def wf_propagate_absorb_inplace(b: np.ndarray, pot, step, mass, k0, time, bounds, kmax=None) -> None:
    """One reference grid_wf_propagate step with an absorbing boundary (2nd order; CFFT when
    kmax is given) in place; pot may be None."""
    t = complex(time)
    pp = None if pot is None else _c(pot).ctypes.data
    lib().or_wf_propagate_absorb(_c(b).ctypes.data, pp, *b.shape, step, mass, k0[0], k0[1], k0[2], t.real, t.imag,
                                 -1.0 if kmax is None else kmax, *bounds)


def wf_normalize(a, step, target=1.0):
    b = _c(np.array(a, dtype=np.complex128, order="C", copy=True))
    n = lib().or_wf_normalize(b.ctypes.data, *b.shape, step, target)
    return b, n


def kinetic_energy_fft(a, step, mass) -> float:
    a = _c(a)
    return lib().or_kinetic_energy_fft(a.ctypes.data, *a.shape, step, mass)


def energy_fft(a, v, step, mass) -> float:
    a = _c(a)
    vp = None if v is None else _r(v).ctypes.data
    return lib().or_energy_fft(a.ctypes.data, vp, *a.shape, step, mass)


def kinetic_energy_fd_periodic(a, step, mass) -> float:
    a = _c(a)
    return lib().or_kinetic_energy_fd_periodic(a.ctypes.data, *a.shape, step, mass)


# boundary rules of the finite-difference energy (src/cgrid_bc.c:21-86)
BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC = 0, 1, 2


def kinetic_energy_fd(a, step, mass, bc, dirichlet=0j) -> float:
    """grid_wf_kinetic_energy_cn with the given boundary rule and Dirichlet value."""
    a = _c(a)
    d = complex(dirichlet)
    return lib().or_kinetic_energy_fd(a.ctypes.data, *a.shape, step, mass, bc, d.real, d.imag)


# ---- host-side grid mapping and interpolation (numpy restatements) ---------------------------

def axis_coords(n, step, origin=0.0):
    """Coordinates of the n points of an axis: (i - n/2) step - origin (src/cgrid.c:793-815)."""
    return (np.arange(n, dtype=np.int64) - n // 2).astype(np.float64) * step - origin


def cell_rule(step, ns):
    """Offsets and hat weights of the ns samples per axis of cgrid_smooth_map
    (linearly_weighted_integralc, src/cprivate.h:35-68)."""
    sub = 2.0 * step / float(ns)
    off = np.array([-step + (float(m) + 0.5) * sub for m in range(ns)])
    return off, 1.0 - np.abs(off / step)


def smooth_map(func, shape, step, origin=(0.0, 0.0, 0.0), ns=1):
    """cgrid_smooth_map / rgrid_smooth_map (src/cgrid.c:832-860, src/rgrid.c:846-870): func is
    vectorised over coordinate arrays; each value is sum_m w_m f(r + s_m) / sum_m w_m over the
    ns^3 samples, accumulated in the reference's loop order (x samples outermost)."""
    x = axis_coords(shape[0], step, origin[0])[:, None, None]
    y = axis_coords(shape[1], step, origin[1])[None, :, None]
    z = axis_coords(shape[2], step, origin[2])[None, None, :]
    if ns == 1:
        return np.broadcast_to(func(x, y, z), shape).copy()
    off, w = cell_rule(step, ns)
    acc = 0.0
    wacc = 0.0
    for a in range(ns):
        for b in range(ns):
            for c in range(ns):
                wt = w[a] * w[b] * w[c]
                acc = acc + wt * func(x + off[a], y + off[b], z + off[c])
                wacc += wt
    return np.broadcast_to(acc / wacc, shape).copy()


def _axis_cell(u, origin, step, n):
    """Lower cell index (shifted by n/2) and fraction of coordinate u (src/cgrid.c:2070-2090)."""
    t = (u + origin) / step
    c = np.trunc(t).astype(np.int64)
    c = np.where(t < 0, c - 1, c)
    return c + n // 2, t - c


def value_at(a, i, j, k, bc, dirichlet=0.0):
    """a[i, j, k] with the boundary function's value outside the grid (src/cgrid_bc.c:21-86);
    a Neumann mirror that stays outside (n < 3) takes the Dirichlet value."""
    nx, ny, nz = a.shape
    i, j, k = np.broadcast_arrays(np.asarray(i), np.asarray(j), np.asarray(k))
    if bc == BC_PERIODIC:
        return a[i % nx, j % ny, k % nz]
    if bc == BC_NEUMANN:
        def mirror(m, n):
            m = np.where(m < 0, -m, m)
            return np.where(m > n - 1, 2 * n - m - 2, m)
        i, j, k = mirror(i, nx), mirror(j, ny), mirror(k, nz)
    inside = (i >= 0) & (j >= 0) & (k >= 0) & (i < nx) & (j < ny) & (k < nz)
    out = np.full(i.shape, dirichlet, dtype=a.dtype)
    out[inside] = a[i[inside], j[inside], k[inside]]
    return out


def interpolate(a, step, origin, x, y, z, bc, dirichlet=0.0):
    """cgrid_value (complex a, nested weights, src/cgrid.c:2060-2104) or rgrid_value (real a,
    expanded weights, src/rgrid.c:2208-2256) at coordinate arrays x, y, z."""
    nx, ny, nz = a.shape
    i, fx = _axis_cell(np.asarray(x, dtype=np.float64), origin[0], step, nx)
    j, fy = _axis_cell(np.asarray(y, dtype=np.float64), origin[1], step, ny)
    k, fz = _axis_cell(np.asarray(z, dtype=np.float64), origin[2], step, nz)
    gx, gy, gz = 1.0 - fx, 1.0 - fy, 1.0 - fz

    def f(di, dj, dk):
        return value_at(a, i + di, j + dj, k + dk, bc, dirichlet)

    if np.iscomplexobj(a):
        return gx * (gy * (gz * f(0, 0, 0) + fz * f(0, 0, 1)) + fy * (gz * f(0, 1, 0) + fz * f(0, 1, 1))) + \
            fx * (gy * (gz * f(1, 0, 0) + fz * f(1, 0, 1)) + fy * (gz * f(1, 1, 0) + fz * f(1, 1, 1)))
    return gx * gy * gz * f(0, 0, 0) + fx * gy * gz * f(1, 0, 0) + gx * fy * gz * f(0, 1, 0) + gx * gy * fz * f(0, 0, 1) + \
        fx * fy * gz * f(1, 1, 0) + fx * gy * fz * f(1, 0, 1) + gx * fy * fz * f(0, 1, 1) + fx * fy * fz * f(1, 1, 1)


def extrapolate(src, src_step, dst_shape, dst_step, dst_origin=(0.0, 0.0, 0.0), src_origin=(0.0, 0.0, 0.0), bc=BC_PERIODIC):
    """cgrid_extrapolate / rgrid_extrapolate (src/cgrid.c:2126-2150, src/rgrid.c:2267-2288): the
    source interpolated at every destination coordinate."""
    x = axis_coords(dst_shape[0], dst_step, dst_origin[0])[:, None, None]
    y = axis_coords(dst_shape[1], dst_step, dst_origin[1])[None, :, None]
    z = axis_coords(dst_shape[2], dst_step, dst_origin[2])[None, None, :]
    x, y, z = np.broadcast_arrays(x, y, z)
    return interpolate(src, src_step, src_origin, x, y, z, bc)


def gpe_step_inplace(b, vext, g, step, mass, k0, time, normalize=True, target=1.0, kmax=None) -> None:
    """One GPE step composed of reference calls; kmax selects the CFFT propagator."""
    t = complex(time)
    lib().or_gpe_step(_c(b).ctypes.data, _r(vext).ctypes.data, g, *b.shape, step, mass, k0[0], k0[1], k0[2], t.real,
                      t.imag, 1 if normalize else 0, target, -1.0 if kmax is None else kmax)


def gpe_energy(a, vext, g, step, mass) -> float:
    a = _c(a)
    return lib().or_gpe_energy(a.ctypes.data, _r(vext).ctypes.data, g, *a.shape, step, mass)


def num_threads() -> int:
    return lib().or_num_threads()


def set_num_threads(n: int) -> None:
    lib().or_set_num_threads(n)

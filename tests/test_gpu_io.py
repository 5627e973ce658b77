"""GPU checks of the boundary as a C program sees it: grids mapped on the host and then used on
the device, binary grid files written from device-resident (deferred, pending-scaled) states and
read back into a running propagation, and a complete C program in the shape of the reference's
examples/wavepacket.c, compiled against include/grid.h, linked with libgrid.so and run, with its
outputs compared against the oracle."""
import ctypes as C
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import libgrid_amd as LG
from conftest import rel_l2
from libgrid_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-12


def case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mapfuncs(tmp_path_factory):
    d = tmp_path_factory.mktemp("mapfuncs")
    so = str(d / "libmapfuncs.so")
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-o", so, os.path.join(ROOT, "tests", "c", "mapfuncs.c")], check=True)
    return C.CDLL(so)


@pytest.fixture(scope="module")
def libc():
    c = C.CDLL(None)
    c.fopen.restype = C.c_void_p
    c.fopen.argtypes = [C.c_char_p, C.c_char_p]
    c.fclose.argtypes = [C.c_void_p]
    return c


def read_grid_file(path):
    raw = open(path, "rb").read()
    nx, ny, nz = struct.unpack("<3q", raw[:24])
    (step,) = struct.unpack("<d", raw[24:32])
    return (nx, ny, nz), step, np.frombuffer(raw[32:], dtype="<c16").reshape(nx, ny, nz)


def test_map_reaches_device_after_device_work(lg, orc, mapfuncs):
    """cgrid_map / grid_wf_map write the host copy; the next device call must use the mapped
    values, also when the device held a deferred, pending-scaled state from earlier steps."""
    shape, step, a = (32, 24, 16), 0.3, 0.37
    case(f"map -> device {shape}")
    x, y, z = np.meshgrid(*[orc.axis_coords(n, step, 0.0) for n in shape], indexing="ij")
    w = lg.grid_wf_alloc(*shape, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"m")
    arg = C.c_double(a)
    fn = C.cast(mapfuncs.map_poly, C.c_void_p).value
    lg.grid_wf_map(w, fn, C.byref(arg))
    want = (1.0 + a * x - 0.2 * y + 0.1 * z + 0.05 * x * x + 0.02 * y * z) + 1j * (0.1 * x * y + 0.03 * z * z)
    assert abs(lg.grid_wf_norm(w) - orc.integral_of_square(want, step)) < 1e-13 * orc.integral_of_square(want, step)
    # device work that leaves the state transformed along z with a pending scale
    vp = lg.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"v")
    lg.cgrid_map(vp, C.cast(mapfuncs.map_xy, C.c_void_p).value, None)
    lg.grid_wf_propagate(w, vp, LG.c_complex.of(0.01 - 0.002j))
    lg.grid_wf_normalize(w)
    lg.grid_wf_propagate(w, vp, LG.c_complex.of(0.01 - 0.002j))
    ref = want.copy()
    pot = x + 1j * y
    orc.wf_propagate_inplace(ref, pot, step, 1.0, (0, 0, 0), 0.01 - 0.002j)
    ref, _ = orc.wf_normalize(ref, step)
    orc.wf_propagate_inplace(ref, pot, step, 1.0, (0, 0, 0), 0.01 - 0.002j)
    assert rel_l2(LG.get_cgrid(w.contents.grid), ref) < TOL
    lg.grid_wf_propagate(w, vp, LG.c_complex.of(0.01 - 0.002j))  # device copy canonical again
    lg.grid_wf_map(w, C.cast(mapfuncs.map_z, C.c_void_p).value, None)
    fresh = z + 1j * (x - y)
    assert abs(lg.grid_wf_norm(w) - orc.integral_of_square(fresh, step)) < 1e-13 * orc.integral_of_square(fresh, step)
    lg.cgrid_fft(w.contents.grid)
    assert rel_l2(LG.get_cgrid(w.contents.grid), orc.fft3d(fresh, -1)) < TOL
    lg.grid_wf_free(w)
    lg.cgrid_free(vp)


@pytest.mark.parametrize("prec", [b"\x00", b"\x01"])
def test_write_read_round_trip_mid_propagation(lg, orc, libc, tmp_path, prec):
    """cgrid_write of a state held on the device transformed along z with a pending
    renormalisation writes the real-space state (header [nx, ny, nz, step], then the values);
    cgrid_read into another wavefunction continues the propagation exactly where the first is."""
    shape, step, dt = (32, 32, 64), 0.25, 0.04
    case(f"write/read mid-propagation {shape} prec={prec}")
    psi = W.normalize(W.gaussian_packet(shape, step, 1.5, (0.3, -0.2, 0.5), (0.2, 0.1, -0.3)), step)
    v = W.harmonic(shape, step)
    lg.cgrid_write.argtypes = [LG.CG, C.c_void_p]
    lg.cgrid_read.argtypes = [LG.CG, C.c_void_p]
    lg.cgrid_read.restype = LG.CG
    w = lg.grid_wf_alloc(*shape, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"a")
    LG.set_cgrid(w.contents.grid, psi)
    lg.cgrid_set_precision(w.contents.grid, prec)
    vr = lg.rgrid_alloc(*shape, step, LG.boundary("rgrid_value_outside_periodic"), None, b"v")
    LG.set_rgrid(vr, v)
    ref = psi.copy()
    for _ in range(3):
        lg.grid_wf_gpe_propagate(w, vr, 0.05, LG.c_complex.of(-1j * dt), b"\x01")
        orc.gpe_step_inplace(ref, v, 0.05, step, 1.0, (0, 0, 0), -1j * dt)
    path = str(tmp_path / "mid.bin")
    fp = libc.fopen(path.encode(), b"wb")
    lg.cgrid_write(w.contents.grid, fp)
    libc.fclose(fp)
    dims, st, vals = read_grid_file(path)
    tol = TOL if prec == b"\x00" else 2e-6
    assert dims == shape and st == step
    assert rel_l2(vals, ref) < tol
    w2 = lg.grid_wf_alloc(*shape, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"b")
    fp = libc.fopen(path.encode(), b"rb")
    lg.cgrid_read(w2.contents.grid, fp)
    libc.fclose(fp)
    lg.cgrid_set_precision(w2.contents.grid, prec)
    for _ in range(3):
        lg.grid_wf_gpe_propagate(w, vr, 0.05, LG.c_complex.of(-1j * dt), b"\x01")
        lg.grid_wf_gpe_propagate(w2, vr, 0.05, LG.c_complex.of(-1j * dt), b"\x01")
        orc.gpe_step_inplace(ref, v, 0.05, step, 1.0, (0, 0, 0), -1j * dt)
    a1, a2 = LG.get_cgrid(w.contents.grid), LG.get_cgrid(w2.contents.grid)
    if prec == b"\x00":
        assert rel_l2(a2, ref) < TOL
        assert rel_l2(a1, a2) < 1e-14
    else:
        # the file holds the FP32 state widened to FP64; reading it back rounds to the same FP32 values
        assert rel_l2(a2, a1) < 1e-6
        assert rel_l2(a2, ref) < 4e-6
    for x in (w, w2):
        lg.grid_wf_free(x)
    lg.rgrid_free(vr)


def test_wavepacket_c_program(lg, orc, tmp_path):
    """The C program tests/c/wavepacket_driver.c, built with gcc against include/grid.h (as
    <grid/grid.h>) and linked with libgrid.so, runs the 128^3 packet-in-a-trap workload (100
    steps of dt = 0.05) through the reference's calls: its printed norms stay within 1e-13 per
    step, the density files it writes and its final wavefunction match the oracle."""
    psi0, v, prm = W.wavepacket_harmonic(seed=0)
    step, dt, iters = prm["step"], prm["dt"], prm["steps"]
    case(f"C program wavepacket 128^3, {iters} steps, {prm}")
    inc = tmp_path / "inc"
    inc.mkdir()
    os.symlink(os.path.join(ROOT, "include"), str(inc / "grid"))
    exe = str(tmp_path / "wavepacket_driver")
    libdir = os.path.join(ROOT, "libgrid_amd", "lib")
    subprocess.run(["gcc", "-std=gnu11", "-O2", "-I" + str(inc), "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "wavepacket_driver.c"), "-o", exe, "-L" + libdir, "-lgrid",
                    "-Wl,-rpath," + libdir, "-lm"], check=True)
    prefix = str(tmp_path / "out")
    kx, ky, kz = prm["momentum"]
    xc, yc, zc = prm["centre"]
    args = [exe, "4", "128", repr(step), repr(dt), str(iters), repr(kx), repr(ky), repr(kz), repr(prm["width"]),
            repr(xc), repr(yc), repr(zc), repr(prm["k"]), prefix]
    print("[run] " + " ".join(args), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    norms = [float(line.split("=")[1]) for line in p.stdout.splitlines() if "norm =" in line]
    assert len(norms) == iters + 1
    assert abs(norms[0] - 1.0) < 1e-14
    assert max(abs(a - b) for a, b in zip(norms[1:], norms[:-1])) <= 1e-13
    pot = v.astype(np.complex128)
    ref = psi0.copy()

    def density_file(l):
        raw = open(f"{prefix}-{l}.grd", "rb").read()
        vals = np.frombuffer(raw[32:], dtype="<f8").reshape(128, 128, 130)[:, :, :128]
        return vals

    d0 = density_file(0)
    assert np.max(np.abs(d0 - np.abs(ref) ** 2)) < 1e-13 * np.max(np.abs(ref) ** 2)
    for s in range(iters):
        if s == iters - 1:
            dl = density_file(s)
            assert np.max(np.abs(dl - np.abs(ref) ** 2)) < 1e-11 * np.max(np.abs(ref) ** 2)
        orc.wf_propagate_inplace(ref, pot, step, 1.0, (0, 0, 0), dt)
    dims, st, final = read_grid_file(prefix + ".final.bin")
    assert dims == (128, 128, 128) and st == step
    err = rel_l2(final, ref)
    print(f"C program: final psi rel L2 vs oracle {err:.2e}, norms {norms[0]:.15f} .. {norms[-1]:.15f}")
    assert err < TOL

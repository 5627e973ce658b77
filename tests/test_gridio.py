"""CPU checks of the host-side grid I/O (libgrid_amd/csrc/host/gridio.c): coordinates of
cgrid_map / rgrid_map / grid_wf_map (reference src/cgrid.c:793-815), the cell-averaged
cgrid_smooth_map / rgrid_smooth_map (src/cgrid.c:832-860), the binary grid files of
cgrid_write / cgrid_read / rgrid_write / rgrid_read (src/cgrid.c:360-420, src/rgrid.c:339-398)
including interpolation onto a grid of another size (cgrid_extrapolate, src/cgrid.c:2126), and
the .grd writers with their ASCII cuts (src/cgrid.c:432-504, src/rgrid.c:505-566).

All of it is host arithmetic, as in the reference, so no GPU is needed: grids are allocated,
mapped, written and read on the host, and compared with the oracle's numpy restatements
(oracle/oracle.py), which tests/test_oracle.py pins."""
import ctypes as C
import os
import struct
import subprocess

import numpy as np
import pytest

import libgrid_amd as LG

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def mapfuncs(tmp_path_factory):
    d = tmp_path_factory.mktemp("mapfuncs")
    so = str(d / "libmapfuncs.so")
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-o", so, os.path.join(ROOT, "tests", "c", "mapfuncs.c")], check=True)
    return C.CDLL(so)


def fptr(lib, name):
    return C.cast(getattr(lib, name), C.c_void_p).value


def np_poly(a):
    return lambda x, y, z: (1.0 + a * x - 0.2 * y + 0.1 * z + 0.05 * x * x + 0.02 * y * z) + 1j * (0.1 * x * y + 0.03 * z * z)


def np_rpoly(a):
    return lambda x, y, z: 1.0 + a * x * x - 0.2 * y + 0.1 * z * y + 0.3 * z


def host_cgrid(lib, g):
    c = g.contents
    n = c.nx * c.ny * c.nz
    return np.frombuffer((C.c_double * (2 * n)).from_address(c.value), dtype=np.complex128).reshape(c.nx, c.ny, c.nz).copy()


def host_rgrid(lib, g):
    c = g.contents
    n = c.nx * c.ny * c.nz2
    return np.frombuffer((C.c_double * n).from_address(c.value), dtype=np.float64).reshape(c.nx, c.ny, c.nz2)[:, :, : c.nz].copy()


@pytest.mark.parametrize("shape,origin", [((5, 6, 7), (0.0, 0.0, 0.0)), ((4, 1, 9), (0.3, -1.2, 0.45)), ((1, 8, 3), (0.0, 2.0, -0.5))])
def test_map_coordinates(orc, mapfuncs, shape, origin):
    """value(i, j, k) = f((i - nx/2) step - x0, (j - ny/2) step - y0, (k - nz/2) step - z0)."""
    lib = LG.load()
    step = 0.35
    g = lib.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"m")
    lib.cgrid_set_origin(g, *origin)
    x, y, z = np.meshgrid(*[orc.axis_coords(n, step, o) for n, o in zip(shape, origin)], indexing="ij")
    lib.cgrid_map(g, fptr(mapfuncs, "map_xy"), None)
    assert np.array_equal(host_cgrid(lib, g), x + 1j * y)
    lib.cgrid_map(g, fptr(mapfuncs, "map_z"), None)
    assert np.array_equal(host_cgrid(lib, g), z + 1j * (x - y))
    lib.cgrid_free(g)
    r = lib.rgrid_alloc(*shape, step, LG.boundary("rgrid_value_outside_periodic"), None, b"r")
    r.contents.x0, r.contents.y0, r.contents.z0 = origin
    lib.rgrid_map(r, fptr(mapfuncs, "rmap_x"), None)
    assert np.array_equal(host_rgrid(lib, r), x)
    lib.rgrid_map(r, fptr(mapfuncs, "rmap_yz"), None)
    assert np.array_equal(host_rgrid(lib, r), 1000.0 * y + z)
    lib.rgrid_free(r)
    w = lib.grid_wf_alloc(*shape, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"w")
    lib.cgrid_set_origin(w.contents.grid, *origin)
    lib.grid_wf_map(w, fptr(mapfuncs, "map_xy"), None)
    assert np.array_equal(host_cgrid(lib, w.contents.grid), x + 1j * y)
    lib.grid_wf_free(w)


@pytest.mark.parametrize("ns", [1, 2, 3, 5])
def test_smooth_map_matches_oracle(orc, mapfuncs, ns):
    lib = LG.load()
    shape, step, origin, a = (6, 5, 8), 0.3, (0.2, -0.1, 0.05), 0.37
    arg = C.c_double(a)
    g = lib.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"s")
    lib.cgrid_set_origin(g, *origin)
    lib.cgrid_smooth_map(g, fptr(mapfuncs, "map_poly"), C.byref(arg), ns)
    want = orc.smooth_map(np_poly(a), shape, step, origin, ns)
    assert np.max(np.abs(host_cgrid(lib, g) - want)) < 1e-14
    lib.cgrid_free(g)
    r = lib.rgrid_alloc(*shape, step, LG.boundary("rgrid_value_outside_periodic"), None, b"r")
    r.contents.x0, r.contents.y0, r.contents.z0 = origin
    lib.rgrid_smooth_map(r, fptr(mapfuncs, "rmap_poly"), C.byref(arg), ns)
    want = orc.smooth_map(np_rpoly(a), shape, step, origin, ns)
    assert np.max(np.abs(host_rgrid(lib, r) - want)) < 1e-14
    lib.rgrid_free(r)


def read_grid_file(path, real=False):
    raw = open(path, "rb").read()
    nx, ny, nz = struct.unpack("<3q", raw[:24])
    (step,) = struct.unpack("<d", raw[24:32])
    if real:
        nz2 = 2 * (nz // 2 + 1)
        vals = np.frombuffer(raw[32:], dtype="<f8").reshape(nx, ny, nz2)
    else:
        vals = np.frombuffer(raw[32:], dtype="<c16").reshape(nx, ny, nz)
    return (nx, ny, nz), step, vals


def write_grid_file(path, shape, step, vals):
    with open(path, "wb") as f:
        f.write(struct.pack("<3q", *shape))
        f.write(struct.pack("<d", step))
        f.write(np.ascontiguousarray(vals).tobytes())


def test_cgrid_write_read_layout_and_round_trip(tmp_path):
    """[INT nx][INT ny][INT nz][REAL step][nx*ny*nz complex] little-endian; read restores the
    values bit for bit; a file written by another program in that layout reads back as written."""
    lib = LG.load()
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    lib.cgrid_write.argtypes = [LG.CG, C.c_void_p]
    lib.cgrid_read.argtypes = [LG.CG, C.c_void_p]
    lib.cgrid_read.restype = LG.CG
    shape, step = (3, 5, 4), 0.125
    a = np.random.default_rng(1).standard_normal(shape) + 1j * np.random.default_rng(2).standard_normal(shape)
    g = lib.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"w")
    LG.set_cgrid(g, a)
    path = str(tmp_path / "g.bin")
    fp = libc.fopen(path.encode(), b"wb")
    lib.cgrid_write(g, fp)
    libc.fclose(fp)
    dims, st, vals = read_grid_file(path)
    assert dims == shape and st == step and np.array_equal(vals, a)
    assert os.path.getsize(path) == 32 + 16 * a.size
    # read into a fresh grid (NULL: allocated from the header) and into an existing one
    b = a * (0.5 - 0.25j)
    write_grid_file(path, shape, step, b)
    fp = libc.fopen(path.encode(), b"rb")
    h = lib.cgrid_read(None, fp)
    libc.fclose(fp)
    assert (h.contents.nx, h.contents.ny, h.contents.nz, h.contents.step) == (*shape, step)
    assert np.array_equal(host_cgrid(lib, h), b)
    fp = libc.fopen(path.encode(), b"rb")
    lib.cgrid_read(g, fp)
    libc.fclose(fp)
    assert np.array_equal(host_cgrid(lib, g), b)
    lib.cgrid_free(g)
    lib.cgrid_free(h)


@pytest.mark.parametrize("bc,bcname", [(2, "cgrid_value_outside_periodic"), (1, "cgrid_value_outside_neumann")])
def test_cgrid_read_interpolates_other_sizes(orc, tmp_path, bc, bcname):
    """A file grid of another size and step is interpolated onto the grid in memory with the
    grid's boundary rule (src/cgrid.c:403-416 with cgrid_extrapolate)."""
    lib = LG.load()
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    lib.cgrid_read.argtypes = [LG.CG, C.c_void_p]
    lib.cgrid_read.restype = LG.CG
    src_shape, src_step = (8, 6, 10), 0.3
    src = np.random.default_rng(3).standard_normal(src_shape) + 1j * np.random.default_rng(4).standard_normal(src_shape)
    path = str(tmp_path / "s.bin")
    write_grid_file(path, src_shape, src_step, src)
    dst_shape, dst_step, origin = (11, 7, 13), 0.21, (0.05, -0.1, 0.2)
    g = lib.cgrid_alloc(*dst_shape, dst_step, LG.boundary(bcname), None, b"d")
    lib.cgrid_set_origin(g, *origin)
    fp = libc.fopen(path.encode(), b"rb")
    lib.cgrid_read(g, fp)
    libc.fclose(fp)
    want = orc.extrapolate(src, src_step, dst_shape, dst_step, dst_origin=origin, bc=bc)
    assert np.max(np.abs(host_cgrid(lib, g) - want)) < 1e-14
    lib.cgrid_free(g)


def test_rgrid_write_read_and_interpolation(orc, tmp_path):
    """rgrid files keep the nz2-padded rows; same-size reads are exact, other sizes interpolate."""
    lib = LG.load()
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    lib.rgrid_write.argtypes = [LG.RG, C.c_void_p]
    lib.rgrid_read.argtypes = [LG.RG, C.c_void_p]
    lib.rgrid_read.restype = LG.RG
    shape, step = (4, 3, 5), 0.2
    a = np.random.default_rng(7).standard_normal(shape)
    r = lib.rgrid_alloc(*shape, step, LG.boundary("rgrid_value_outside_periodic"), None, b"r")
    LG.set_rgrid(r, a)
    path = str(tmp_path / "r.bin")
    fp = libc.fopen(path.encode(), b"wb")
    lib.rgrid_write(r, fp)
    libc.fclose(fp)
    dims, st, vals = read_grid_file(path, real=True)
    assert dims == shape and st == step
    assert np.array_equal(vals[:, :, :5], a) and np.all(vals[:, :, 5:] == 0.0)
    fp = libc.fopen(path.encode(), b"rb")
    h = lib.rgrid_read(None, fp)
    libc.fclose(fp)
    assert np.array_equal(host_rgrid(lib, h), a)
    d = lib.rgrid_alloc(9, 5, 8, 0.11, LG.boundary("rgrid_value_outside_periodic"), None, b"d")
    fp = libc.fopen(path.encode(), b"rb")
    lib.rgrid_read(d, fp)
    libc.fclose(fp)
    want = orc.extrapolate(a, step, (9, 5, 8), 0.11)
    assert np.max(np.abs(host_rgrid(lib, d) - want)) < 1e-14
    for x in (r, h, d):
        lib.rgrid_free(x)


def test_write_grid_files(tmp_path):
    """cgrid_write_grid / rgrid_write_grid: base.grd in the binary layout and the cuts through the
    middle point, one '%le %le[ %le]' line per point with the coordinate (m - n/2) step."""
    lib = LG.load()
    shape, step = (6, 5, 4), 0.25
    a = np.random.default_rng(9).standard_normal(shape) + 1j * np.random.default_rng(10).standard_normal(shape)
    g = lib.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"c")
    LG.set_cgrid(g, a)
    base = str(tmp_path / "cg")
    lib.cgrid_write_grid.argtypes = [C.c_char_p, LG.CG]
    lib.cgrid_write_grid(base.encode(), g)
    dims, st, vals = read_grid_file(base + ".grd")
    assert dims == shape and st == step and np.array_equal(vals, a)
    mid = [n // 2 for n in shape]
    for axis, ext in enumerate("xyz"):
        lines = open(base + "." + ext).read().splitlines()
        assert len(lines) == shape[axis]
        for m, line in enumerate(lines):
            idx = list(mid)
            idx[axis] = m
            v = a[tuple(idx)]
            assert line == "%e %e %e" % ((m - shape[axis] // 2) * step, v.real, v.imag)
    lib.cgrid_free(g)
    r = lib.rgrid_alloc(*shape, step, LG.boundary("rgrid_value_outside_periodic"), None, b"r")
    LG.set_rgrid(r, a.real)
    base = str(tmp_path / "rg")
    lib.rgrid_write_grid.argtypes = [C.c_char_p, LG.RG]
    lib.rgrid_write_grid(base.encode(), r)
    dims, st, vals = read_grid_file(base + ".grd", real=True)
    assert dims == shape and np.array_equal(vals[:, :, :4], a.real)
    for axis, ext in enumerate("xyz"):
        lines = open(base + "." + ext).read().splitlines()
        for m, line in enumerate(lines):
            idx = list(mid)
            idx[axis] = m
            assert line == "%e %e" % ((m - shape[axis] // 2) * step, a.real[tuple(idx)])
    lib.rgrid_free(r)

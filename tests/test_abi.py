"""CPU checks of the C-ABI: libgrid.so loads, exports every function declared in include/*.h,
and the ctypes mirrors of the public structs have the C layout.  No GPU call is made."""
import ctypes as C
import os
import re
import subprocess
import sys
import tempfile

import pytest

import libgrid_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "grid.h"), os.path.join(ROOT, "include", "grid_hip.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//.*", "", text)
    names = set()
    for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_ \*]*?)\b([a-z][a-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
        name = m.group(2)
        if name in ("if", "while", "sizeof", "return", "complex"):
            continue
        names.add(name)
    return names


def test_library_loads_and_exports_all_declared_symbols():
    lib = libgrid_amd.load()
    missing = []
    total = 0
    for h in HEADERS:
        for name in sorted(declared_functions(h)):
            total += 1
            if not hasattr(lib, name):
                missing.append(name)
    assert total > 80
    assert not missing, missing


def test_prototype_table_covers_only_exported_names():
    lib = libgrid_amd.load()
    for name in libgrid_amd.PROTOTYPES:
        assert hasattr(lib, name), name


def test_struct_layout_matches_c():
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "grid.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(cgrid), offsetof(cgrid, fft_norm), offsetof(cgrid, dev),
         sizeof(rgrid), offsetof(rgrid, nz2), sizeof(wf));
  printf("%zu %zu %zu\n", offsetof(wf, mass), offsetof(wf, ts_func), offsetof(wf, kmax));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=gnu11", "-I" + os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    vals = list(map(int, out))
    assert vals[0] == C.sizeof(libgrid_amd.cgrid)
    assert vals[1] == libgrid_amd.cgrid.fft_norm.offset
    assert vals[2] == libgrid_amd.cgrid.dev.offset
    assert vals[3] == C.sizeof(libgrid_amd.rgrid)
    assert vals[4] == libgrid_amd.rgrid.nz2.offset
    assert vals[5] == C.sizeof(libgrid_amd.wf)
    assert vals[6] == libgrid_amd.wf.mass.offset
    assert vals[7] == libgrid_amd.wf.ts_func.offset
    assert vals[8] == libgrid_amd.wf.kmax.offset


def test_no_gpu_means_loud_failure():
    """Without a HIP device every computational call aborts with a libgrid message (no CPU path)."""
    import torch  # noqa: F401  (only to know whether a GPU exists here)

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    code = (
        "import libgrid_amd, ctypes; lib = libgrid_amd.load();"
        "g = lib.cgrid_alloc(4, 4, 4, 0.1, None, None, b'x'); lib.cgrid_fft(g)"
    )
    p = subprocess.run(["python", "-c", code], capture_output=True, text=True, cwd=ROOT,
                       env=dict(os.environ, HIP_VISIBLE_DEVICES="", PYTHONPATH=ROOT))
    assert p.returncode != 0
    assert "libgrid:" in p.stderr


def test_absorb_amplitude_matches_oracle(orc):
    """grid_wf_absorb is host arithmetic (no GPU call): every point of a 12 x 9 x 10 grid for
    several boundary boxes, bit for bit against the oracle's restatement."""
    lib = libgrid_amd.load()
    for b in [(3, 8, 2, 6, 2, 7), (1, 10, 4, 4, 3, 5), (5, 6, 1, 7, 4, 5)]:
        for i in range(12):
            for j in range(9):
                for k in range(10):
                    assert lib.grid_wf_absorb(i, j, k, *b) == orc.wf_absorb(i, j, k, b), (b, i, j, k)


def test_every_struct_mirror_matches_c_field_by_field():
    """sizeof and the offset of every field of each ctypes mirror (cgrid, rgrid, wf, hgrid_pass)
    against the C compiler's, so a field added on one side only fails here, on the CPU."""
    mirrors = {"cgrid": libgrid_amd.cgrid, "rgrid": libgrid_amd.rgrid, "wf": libgrid_amd.wf,
               "hgrid_pass": libgrid_amd.hgrid_pass}
    lines = []
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            field = f[0].rstrip("_")
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {field}));')
    src = "#include <stddef.h>\n#include <stdio.h>\n#include \"grid.h\"\n#include \"grid_hip.h\"\n" \
          "int main(void) {\n" + "\n".join(lines) + "\n  return 0;\n}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=gnu11", "-I" + os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.splitlines()
    for line in out:
        cname, field, val = line.split()
        py = mirrors[cname]
        want = C.sizeof(py) if field == "sizeof" else getattr(py, field).offset
        assert int(val) == want, (cname, field, int(val), want)


def test_wf_boundary_fields():
    """grid_wf_boundary (reference src/grid_wf.c:99-120) is host bookkeeping: it sets the box and
    ts_func on the wavefunction and ts_func on the predict wavefunction; all bounds 0 clears
    them, and of the predict wavefunction's bounds only lx, as the reference does."""
    lib = libgrid_amd.load()
    w = lib.grid_wf_alloc(4, 4, 4, 0.1, 1.0, libgrid_amd.WF_PERIODIC_BOUNDARY, libgrid_amd.WF_2ND_ORDER_FFT, b"w")
    p = lib.grid_wf_alloc(4, 4, 4, 0.1, 1.0, libgrid_amd.WF_PERIODIC_BOUNDARY, libgrid_amd.WF_2ND_ORDER_FFT, b"p")
    try:
        p.contents.ly = 7
        lib.grid_wf_boundary(w, p, 1, 2, 3, 4, 5, 6)
        c = w.contents
        assert (c.lx, c.hx, c.ly, c.hy, c.lz, c.hz) == (1, 2, 3, 4, 5, 6)
        assert c.ts_func and p.contents.ts_func and p.contents.lx == 0 and p.contents.ly == 7
        p.contents.lx = 9
        lib.grid_wf_boundary(w, p, 0, 0, 0, 0, 0, 0)
        assert (c.lx, c.hx, c.ly, c.hy, c.lz, c.hz) == (0,) * 6 and not c.ts_func
        assert not p.contents.ts_func and p.contents.lx == 0 and p.contents.ly == 7
    finally:
        lib.grid_wf_free(w)
        lib.grid_wf_free(p)


@pytest.mark.parametrize("propagator,setup,message", [
    (5, "", "Propagator 5 is not implemented"),
    (2, "lib.grid_wf_boundary(w, None, 1, 2, 1, 2, 1, 2);", "4th order propagator not available for FFT absorbing BC"),
])
def test_unsupported_propagator_combinations_abort(propagator, setup, message):
    """Propagators this build does not implement, and the 4th-order step with an absorbing
    boundary (which the reference refuses too, src/grid_wf.c:434-437), abort with a libgrid
    message before any GPU work."""
    code = (
        "import libgrid_amd as LG; lib = LG.load();"
        f"w = lib.grid_wf_alloc(4, 4, 4, 0.1, 1.0, LG.WF_PERIODIC_BOUNDARY, {propagator}, b'w');"
        f"{setup}"
        "lib.grid_wf_propagate(w, None, LG.c_complex.of(0.01))"
    )
    p = subprocess.run(["python", "-c", code], capture_output=True, text=True, cwd=ROOT,
                       env=dict(os.environ, HIP_VISIBLE_DEVICES="", PYTHONPATH=ROOT))
    assert p.returncode != 0
    assert "libgrid:" in p.stderr and message in p.stderr


def test_gpu_fft_test_shapes_are_in_the_transform_set():
    """Every shape the GPU FFT parity test runs has axis lengths the library transforms (a length
    outside the set aborts on the host before any GPU work, which would end the whole GPU test
    process), and lengths outside the set are refused with the library's message."""
    from test_gpu_parity import FFT32_SHAPES, FFT_SHAPES

    lib = libgrid_amd.load()
    for shape in FFT_SHAPES + FFT32_SHAPES:
        for n in shape:
            assert n == 1 or lib.hgrid_fft_supported(n), (shape, n)
    code = ("import libgrid_amd as LG; lib = LG.load(); "
            "g = lib.cgrid_alloc(3, 5, 512, 0.2, LG.boundary('cgrid_value_outside_periodic'), None, b'g'); "
            "lib.cgrid_fft(g)")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode != 0
    assert "libgrid: FFT length 5 is not supported" in r.stderr

"""Learned warm start for imaginary-time ground states: training data from the library's own
solver, training through the C API, and the comparison with the plain Gaussian start.

Geometry: a periodic box of side BOX bohr; the network works on a COARSE^3 grid (step
BOX / COARSE), wavefunctions are relaxed on FINE^3 (step BOX / FINE).  Coarse points coincide
with every (FINE / COARSE)-th fine point, so the network sees the fine potential sampled there.
Potentials are seeded sums of 4-10 attractive Gaussians (workloads.wells_3d).  Relaxation is
grid_wf_gpe_relax with g = 0 (linear Schroedinger), time step DT, windowed energy rule; the
converged solver energy, never the network, decides the result.

Run as a module to build a network:
    python -m libgrid_amd.warmstart train --out weights.bin [--count 2000 --epochs 30]
"""
import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

from . import WF_2ND_ORDER_FFT, WF_PERIODIC_BOUNDARY, boundary, get_cgrid, load, set_rgrid
from . import workloads as W

BOX = 32.0
COARSE = 64
FINE = 256
DT = 0.1
GAUSS_WIDTH = 0.15 * BOX  # plain start: exp(-r^2 / (2 w^2)) at the box centre
VSCALE = 2.0              # network input: 2 V (V in [-0.5 * wells, 0] Eh)
TRAIN_SEED = 20240101     # potentials 0 .. count-1 of this stream train; HELDOUT_SEED tests
HELDOUT_SEED = 777
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet64.bin")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def log(msg):
    print(msg, flush=True)


class Problem:
    """One wavefunction + potential pair on an n^3 grid of the box."""

    def __init__(self, lg, n, v):
        self.lg, self.n = lg, n
        step = BOX / n
        self.w = lg.grid_wf_alloc(n, n, n, step, 1.0, WF_PERIODIC_BOUNDARY, WF_2ND_ORDER_FFT, b"ws")
        self.r = lg.rgrid_alloc(n, n, n, step, boundary("rgrid_value_outside_periodic"), None, b"v")
        set_rgrid(self.r, v)

    def free(self):
        self.lg.grid_wf_free(self.w)
        self.lg.rgrid_free(self.r)

    def start(self, net):
        if net is None:
            self.lg.grid_wf_gaussian_start(self.w, GAUSS_WIDTH)
        else:
            self.lg.grid_wf_warm_start(self.w, self.r, net, GAUSS_WIDTH)

    def relax(self, net, fp32_tol=None, tol=1e-10, max_steps=200000, fp32_mode=b"\x01", stride=0, stride32=0):
        """Timed: start construction, optional FP32-storage relaxation to fp32_tol (fp32_mode 1:
        FP32 arithmetic, 2: FP64 arithmetic), then FP64 to tol.  stride > 0 runs the FP64 stage
        and stride32 > 0 the FP32 stage with vector extrapolation every that many steps
        (grid_wf_gpe_relax_extrapolated)."""
        lg = self.lg
        g = self.w.contents.grid
        e = C.c_double()
        lg.cgrid_set_precision(g, b"\x00")
        self.start(None)                     # untimed: uploads V, builds tables and plans
        lg.grid_wf_gpe_energy(self.w, self.r, 0.0)
        lg.grid_hip_sync()
        t0 = time.perf_counter()
        s32 = 0
        if fp32_tol is not None:
            lg.cgrid_set_precision(g, fp32_mode)
        self.start(net)
        nx = C.c_long(0)
        n32 = C.c_long(0)
        if fp32_tol is not None:
            if stride32:
                s32 = lg.grid_wf_gpe_relax_extrapolated(self.w, self.r, 0.0, DT, fp32_tol, max_steps, C.byref(e),
                                                        stride32, C.byref(n32))
            else:
                s32 = lg.grid_wf_gpe_relax(self.w, self.r, 0.0, DT, fp32_tol, max_steps, C.byref(e))
            lg.cgrid_set_precision(g, b"\x00")
        if stride:
            s64 = lg.grid_wf_gpe_relax_extrapolated(self.w, self.r, 0.0, DT, tol, max_steps, C.byref(e), stride,
                                                    C.byref(nx))
        else:
            s64 = lg.grid_wf_gpe_relax(self.w, self.r, 0.0, DT, tol, max_steps, C.byref(e))
        lg.grid_hip_sync()
        sec = time.perf_counter() - t0
        energy = lg.grid_wf_gpe_energy(self.w, self.r, 0.0)
        return dict(seconds=sec, fp32_steps=int(s32), fp64_steps=int(s64), energy=energy,
                    converged=bool(s64 < max_steps), extrapolations=int(nx.value + n32.value))


def potentials(seed, count, n):
    prm = W.wells_3d_params(seed, count, BOX)
    return prm, [W.wells_3d(p, n, BOX / n) for p in prm]


def ground_state_amplitude(lg, v, tol=1e-7, max_steps=20000):
    """|psi_0| on the coarse grid from the Gaussian start (FP32 relaxation), scaled to unit
    mean square; also returns the steps taken."""
    p = Problem(lg, v.shape[0], v)
    try:
        e = C.c_double()
        lg.cgrid_set_precision(p.w.contents.grid, b"\x01")
        p.start(None)
        steps = lg.grid_wf_gpe_relax(p.w, p.r, 0.0, DT, tol, max_steps, C.byref(e))
        a = np.abs(get_cgrid(p.w.contents.grid))
        a /= np.sqrt(np.mean(a * a))
        return a, int(steps), e.value
    finally:
        p.free()


def make_dataset(lg, seed, count, n=COARSE, tol=1e-7, max_steps=20000, every=100):
    _, vs = potentials(seed, count, n)
    V = np.stack(vs)
    T = np.empty_like(V)
    steps = np.empty(count, dtype=np.int64)
    t0 = time.perf_counter()
    for i, v in enumerate(vs):
        T[i], steps[i], _ = ground_state_amplitude(lg, v, tol, max_steps)
        if (i + 1) % every == 0:
            log(f"  ground states {i + 1}/{count} ({time.perf_counter() - t0:.0f} s, median steps "
                f"{int(np.median(steps[:i + 1]))})")
    return V, T, steps


def train(lg, V, T, epochs=30, lr=1e-3, batch=8, seed=0, budget=None, augment=True):
    """Adam over (V, T); with a time budget (seconds) the epoch count is cut to what fits,
    from the measured time of a few minibatches."""
    net = lg.grid_surrogate_alloc(V.shape[1], batch, VSCALE, seed)
    n = V.shape[0]
    if budget:
        vb, tb = np.ascontiguousarray(V[:batch]), np.ascontiguousarray(T[:batch])
        lg.grid_surrogate_gradient(net, batch, _dp(vb), _dp(tb), None)
        lg.grid_hip_sync()
        t0 = time.perf_counter()
        for _ in range(3):
            lg.grid_surrogate_gradient(net, batch, _dp(vb), _dp(tb), None)
        lg.grid_hip_sync()
        per_batch = (time.perf_counter() - t0) / 3
        per_epoch = per_batch * ((n + batch - 1) // batch)
        fit = max(1, int(budget / per_epoch))
        log(f"  {per_batch * 1e3:.1f} ms per minibatch of {batch} (forward + backward), {per_epoch:.1f} s per epoch; "
            f"epochs {epochs} -> {min(epochs, fit)}")
        epochs = min(epochs, fit)
    lg.grid_surrogate_set_training_data(net, n, _dp(np.ascontiguousarray(V)), _dp(np.ascontiguousarray(T)))
    hist = np.zeros(epochs)
    t0 = time.perf_counter()
    lg.grid_surrogate_train_ex(net, epochs, lr, seed, (1 if augment else 0) | 2, _dp(hist))
    lg.grid_hip_sync()
    return net, hist[:epochs], time.perf_counter() - t0


def overlap(a, b):
    return float(np.sum(a * b) / np.sqrt(np.sum(a * a) * np.sum(b * b)))


EXTRAP_STRIDE = 64
MODES = {  # name: (use the network, FP32-storage stage tolerance or None, its arithmetic[, FP64 extrapolation stride])
    "gauss64": (False, None, None),
    "net64": (True, None, None),
    # the FP64 stage with vector extrapolation (grid_wf_gpe_relax_extrapolated)
    "gauss64e": (False, None, None, EXTRAP_STRIDE),
    "net32e": (True, 3e-8, b"\x01", EXTRAP_STRIDE),
    "net32ee": (True, 3e-8, b"\x01", EXTRAP_STRIDE, EXTRAP_STRIDE),  # both stages extrapolated
    # FP32 storage and arithmetic to 3e-8 Eh, then FP64.  The FP32 state's energy wanders by
    # about 1e-8 around the FP64 one, so a tighter FP32 stage mostly waits on noise.
    "gauss32": (False, 3e-8, b"\x01"),
    "net32": (True, 3e-8, b"\x01"),
    "net32_1e8": (True, 1e-8, b"\x01"),
    "net32_1e7": (True, 1e-7, b"\x01"),
    "gauss32x": (False, 1e-10, b"\x02"),  # FP32 storage with FP64 arithmetic, then FP64 confirms
    "net32x": (True, 1e-10, b"\x02"),
}


# A fixed subset of the 200 held-out potentials, every 8th (#0, #8, ..., #192): the same draws
# of the seed's stream whatever the caller, spread over the whole set
HELDOUT_SUBSET = tuple(range(0, 200, 8))


def evaluate(lg, net, count, n=FINE, modes=("gauss64", "net32", "net64"), seed=HELDOUT_SEED, first=0,
             max_steps=200000, tol=1e-10, every=1, indices=None):
    """Relax held-out potentials first .. first+count-1 of the seed's stream (or the given
    indices) on n^3 from each start; returns per-potential records.  The converged energies of
    all modes must agree."""
    idx = list(indices) if indices is not None else list(range(first, first + count))
    prm = W.wells_3d_params(seed, max(idx) + 1, BOX)
    out = []
    for i in idx:
        v = W.wells_3d(prm[i], n, BOX / n)
        p = Problem(lg, n, v)
        rec = dict(index=i, wells=len(prm[i]["depth"]))
        try:
            for m in modes:
                use, t32, mode32, *rest = MODES[m]
                rec[m] = p.relax(net if use else None, fp32_tol=t32, tol=tol, max_steps=max_steps,
                                 fp32_mode=mode32 or b"\x01", stride=rest[0] if rest else 0,
                                 stride32=rest[1] if len(rest) > 1 else 0)
        finally:
            p.free()
        es = [rec[m]["energy"] for m in modes]
        rec["energy_spread"] = float(max(es) - min(es))
        out.append(rec)
        if (len(out) % every) == 0:
            log(f"  potential {i}: " + "; ".join(
                f"{m} {rec[m]['seconds']:.3f} s ({rec[m]['fp32_steps']}+{rec[m]['fp64_steps']} steps)" for m in modes)
                + f"; E {rec[modes[0]]['energy']:.12f}, spread {rec['energy_spread']:.1e}")
    return out


def summarize(recs, modes):
    s = {}
    for m in modes:
        s[m] = dict(seconds_total=float(sum(r[m]["seconds"] for r in recs)),
                    fp32_steps_mean=float(np.mean([r[m]["fp32_steps"] for r in recs])),
                    fp64_steps_mean=float(np.mean([r[m]["fp64_steps"] for r in recs])),
                    converged=int(sum(r[m]["converged"] for r in recs)),
                    extrapolations_mean=float(np.mean([r[m].get("extrapolations", 0) for r in recs])))
    base = modes[0]
    for m in modes[1:]:
        ratios = [r[base]["seconds"] / r[m]["seconds"] for r in recs]
        s[m]["speedup_total"] = s[base]["seconds_total"] / s[m]["seconds_total"]
        s[m]["speedup_median"] = float(np.median(ratios))
    s["energy_spread_max"] = float(max(r["energy_spread"] for r in recs))
    s["count"] = len(recs)
    return s


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--count", type=int, default=2000)
    t.add_argument("--heldout", type=int, default=200)
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--budget", type=float, default=None, help="training time budget in seconds")
    t.add_argument("--tol", type=float, default=1e-6, help="energy tolerance of the training ground states")
    t.add_argument("--max-steps", type=int, default=8000)
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", required=True)
    e.add_argument("--count", type=int, default=200)
    e.add_argument("--first", type=int, default=0)
    e.add_argument("--n", type=int, default=FINE)
    e.add_argument("--modes", default="gauss64,net32ee,net64")
    e.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    lg = load()
    if a.cmd == "evaluate":
        import json
        net = lg.grid_surrogate_load(a.weights.encode(), 1)
        if not net:
            raise SystemExit(f"cannot read {a.weights}")
        modes = tuple(a.modes.split(","))
        log(f"evaluating {a.count} held-out potentials (seed {HELDOUT_SEED}) on {a.n}^3: {modes}")
        recs = evaluate(lg, net, a.count, a.n, modes, first=a.first)
        summ = summarize(recs, modes)
        log(json.dumps(summ, indent=1))
        if a.json:
            with open(a.json, "w") as f:
                json.dump(dict(summary=summ, records=recs), f, indent=1)
        lg.grid_surrogate_free(net)
    if a.cmd == "train":
        log(f"training set: {a.count} potentials (seed {TRAIN_SEED}), {COARSE}^3, box {BOX} bohr")
        V, T, st = make_dataset(lg, TRAIN_SEED, a.count, tol=a.tol, max_steps=a.max_steps)
        log(f"  relaxation steps: median {int(np.median(st))}, max {st.max()}, capped {int(np.sum(st >= a.max_steps))}")
        Vh, Th, sth = make_dataset(lg, HELDOUT_SEED, a.heldout, tol=a.tol, max_steps=a.max_steps)
        net, hist, sec = train(lg, V, T, a.epochs, a.lr, a.batch, budget=a.budget)
        log(f"trained {len(hist)} epochs in {sec:.1f} s; loss per epoch {np.array2string(hist, precision=4)}")
        lh = lg.grid_surrogate_loss(net, len(Vh), _dp(Vh), _dp(Th))
        pred = np.zeros_like(Th)
        lg.grid_surrogate_predict(net, len(Vh), _dp(Vh), _dp(pred))
        ov = [overlap(pred[i], Th[i]) for i in range(len(Vh))]
        log(f"held-out loss {lh:.4f}; overlap <net|psi0> median {np.median(ov):.4f} min {np.min(ov):.4f}")
        if lg.grid_surrogate_save(net, a.out.encode()) != 0:
            raise SystemExit(f"cannot write {a.out}")
        log(f"saved {a.out}")
        lg.grid_surrogate_free(net)
    return 0


if __name__ == "__main__":
    sys.exit(main())

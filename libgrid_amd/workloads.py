"""Seeded synthetic inputs for the split-operator workloads (numpy only).

Coordinates follow the reference's grid mapping x = (i - nx/2) * step - x0 (src/cgrid.c:793-815,
src/rgrid.c:806).  Every drawn parameter is returned in a dict so reports can print it.

Workloads:
  wavepacket_harmonic  128^3 complex128, step 0.2, mass 1: Gaussian packet of width 1.5 bohr
                       (functional form of examples/wavepacket.c:159-180) in V = 0.5*0.01*r^2,
                       centre U(-2,2)^3, momentum U(-1,1)^3; 100 real-time steps of dt = 0.05.
  gpe_barrier          256^3 complex128, step 0.25: V = 0.5*0.01*r^2 + 0.05*exp(-r^2/2^2), g = 0.1,
                       Gaussian (sigma = 1/sqrt(m*omega)) times exp(i theta), theta ~ U[0, 2pi) per
                       point then a 3x3x3 periodic box average; imaginary-time relaxation.
  vortex_ring          512^3, step 0.15: condensate envelope exp(-r^2 / (2 sigma^2)), sigma = 10, in
                       V = 0.5*0.01*r^2 with g = 0.1, times a vortex ring of radius R ~ U[4, 8] bohr
                       about an axis drawn uniformly on the sphere (core profile tanh(d / xi),
                       xi = 1, phase winding atan2(axial offset, radial offset - R)); real time
                       dt = 0.02, 500 steps.
  wells_2d             256 independent 2-D problems (1 x n x n, n uniform in {64, 96, 128, 192, 256},
                       step 0.2): V = -sum_w depth_w exp(-|r - c_w|^2 / (2 width_w^2)) with 3..8
                       wells, depth U[0.05, 0.5], width U[0.5, 3], centres uniform in the middle 60%
                       of the box; start state a Gaussian of standard deviation 0.15 * box at the
                       centre; imaginary-time relaxation.
"""
from __future__ import annotations

import numpy as np


def axes(n: int, step: float, origin: float = 0.0) -> np.ndarray:
    return (np.arange(n) - n // 2) * step - origin


def coords(shape, step):
    nx, ny, nz = shape
    return np.meshgrid(axes(nx, step), axes(ny, step), axes(nz, step), indexing="ij", sparse=True)


def normalize(psi: np.ndarray, step: float) -> np.ndarray:
    dv = step ** sum(1 for n in psi.shape if n > 1)
    return psi / np.sqrt(np.sum(np.abs(psi) ** 2) * dv)


def gaussian_packet(shape, step, width, centre, momentum) -> np.ndarray:
    """exp(-|r - rc|^2 / w^2 + i p.(r - rc)), the form of examples/wavepacket.c:173-180."""
    x, y, z = coords(shape, step)
    dx, dy, dz = x - centre[0], y - centre[1], z - centre[2]
    arg = -(dx * dx + dy * dy + dz * dz) / width**2 + 1j * (momentum[0] * dx + momentum[1] * dy + momentum[2] * dz)
    return np.exp(arg).astype(np.complex128)


def harmonic(shape, step, k=0.01) -> np.ndarray:
    x, y, z = coords(shape, step)
    return (0.5 * k * (x * x + y * y + z * z)).astype(np.float64) * np.ones(shape)


def wavepacket_harmonic(seed: int = 0, n: int = 128, step: float = 0.2):
    rng = np.random.default_rng(seed)
    centre = rng.uniform(-2.0, 2.0, 3)
    momentum = rng.uniform(-1.0, 1.0, 3)
    shape = (n, n, n)
    psi = normalize(gaussian_packet(shape, step, 1.5, centre, momentum), step)
    v = harmonic(shape, step)
    params = dict(shape=shape, step=step, mass=1.0, width=1.5, centre=centre.tolist(), momentum=momentum.tolist(),
                  dt=0.05, steps=100, k=0.01)
    return psi, v, params


def box_average3(theta: np.ndarray) -> np.ndarray:
    """3x3x3 periodic box average (the smoothing of the random phase field)."""
    out = np.zeros_like(theta)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                out += np.roll(theta, (dx, dy, dz), axis=(0, 1, 2))
    return out / 27.0


def gpe_barrier(seed: int = 0, n: int = 256, step: float = 0.25, omega: float = 0.1, barrier: float = 0.05,
                barrier_width: float = 2.0, g: float = 0.1):
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    x, y, z = coords(shape, step)
    r2 = x * x + y * y + z * z
    vext = (0.5 * omega**2 * r2 + barrier * np.exp(-r2 / barrier_width**2)).astype(np.float64)
    vext = np.broadcast_to(vext, shape).copy()
    theta = rng.uniform(0.0, 2.0 * np.pi, shape)
    theta = box_average3(theta)
    sigma2 = 1.0 / omega  # 1 / (m omega), m = 1
    psi = np.exp(-r2 / (2.0 * sigma2)) * np.exp(1j * theta)
    psi = normalize(psi.astype(np.complex128), step)
    params = dict(shape=shape, step=step, mass=1.0, g=g, omega=omega, barrier=barrier, barrier_width=barrier_width,
                  dt=0.05, tol=1e-10)
    return psi, vext, params


def vortex_ring(seed: int = 0, n: int = 512, step: float = 0.15, sigma: float = 10.0, xi: float = 1.0,
                omega: float = 0.1, g: float = 0.1):
    rng = np.random.default_rng(seed)
    radius = float(rng.uniform(4.0, 8.0))
    v = rng.standard_normal(3)
    axis = v / np.linalg.norm(v)
    shape = (n, n, n)
    x, y, z = coords(shape, step)
    r2 = x * x + y * y + z * z
    zp = x * axis[0] + y * axis[1] + z * axis[2]               # axial coordinate
    rho = np.sqrt(np.maximum(r2 - zp * zp, 0.0))               # distance from the axis
    d = np.sqrt((rho - radius) ** 2 + zp * zp)                 # distance from the ring core
    psi = np.exp(-r2 / (2 * sigma * sigma)) * np.tanh(d / xi) * np.exp(1j * np.arctan2(zp, rho - radius))
    psi = normalize(psi.astype(np.complex128), step)
    vext = np.broadcast_to((0.5 * omega**2 * r2).astype(np.float64), shape).copy()
    params = dict(shape=shape, step=step, mass=1.0, g=g, omega=omega, radius=radius, axis=axis.tolist(), sigma=sigma,
                  xi=xi, dt=0.02, steps=500)
    return psi, vext, params


def wells_2d(seed: int = 0, count: int = 256, step: float = 0.2, sizes=(64, 96, 128, 192, 256)):
    """List of (psi, v, params) for the batched 2-D ground-state solves."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(sizes))
        box = n * step
        nw = int(rng.integers(3, 9))
        depth = rng.uniform(0.05, 0.5, nw)
        width = rng.uniform(0.5, 3.0, nw)
        centre = rng.uniform(-0.3 * box, 0.3 * box, (nw, 2))
        y = axes(n, step)[:, None]
        z = axes(n, step)[None, :]
        v = np.zeros((n, n))
        for d, wd, c in zip(depth, width, centre):
            v -= d * np.exp(-((y - c[0]) ** 2 + (z - c[1]) ** 2) / (2 * wd * wd))
        sig = 0.15 * box
        psi = np.exp(-(y * y + z * z) / (2 * sig * sig)).astype(np.complex128)
        psi = normalize(psi.reshape(1, n, n), step)
        out.append((psi, v.reshape(1, n, n), dict(n=n, step=step, wells=nw, depth=depth.tolist(), width=width.tolist(),
                                                 centre=centre.tolist())))
    return out


def wells_3d_params(seed: int, count: int, box: float):
    """Seeded parameters of 3-D attractive Gaussian wells: 4-10 wells, depth U[0.05, 0.5] Eh,
    width U[0.5, 3] bohr, centres uniform within the middle 60% of the box (per axis)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nw = int(rng.integers(4, 11))
        out.append(dict(depth=rng.uniform(0.05, 0.5, nw), width=rng.uniform(0.5, 3.0, nw),
                        centre=rng.uniform(-0.3 * box, 0.3 * box, (nw, 3)), box=box))
    return out


def wells_3d(prm, n: int, step: float):
    """V on an n^3 grid (coordinates (i - n/2) step, as cgrid_map): sum of -depth exp(-d^2 / (2
    width^2)) with d the minimum-image distance in the periodic box of side n step, so V is
    smooth across the boundary.  Grids of the same box that share every f-th point sample the
    same function there."""
    box = n * step
    assert abs(box - prm["box"]) < 1e-9 * box
    x = axes(n, step)
    v = np.zeros((n, n, n))
    for d, w, c in zip(prm["depth"], prm["width"], prm["centre"]):
        f = []
        for a in range(3):
            dx = (x - c[a] + 0.5 * box) % box - 0.5 * box
            f.append(np.exp(-dx * dx / (2 * w * w)))
        v -= d * (f[0][:, None, None] * f[1][None, :, None] * f[2][None, None, :])
    return v

"""Synthetic HeatFlow grids used by the benchmark and the parity tests.

Each builder returns a `System` (uniform per-axis cell lengths, per-cell material fields) plus
a dict describing the run (steps, tolerance, seeds).  Random fields use
numpy.random.Generator(PCG64(seed)); the seeds are part of the returned description.

  c1  16^3 Cattaneo, kappa=150, one centre heater 1e-4 W, x faces 300 K         (100 steps)
  c2  256^3 two-phase Bernoulli(0.3) voxels, 32 heaters x 1e-5 W               (50 steps)
  c3  512^3 log-normal kappa=exp(N(ln 50, 1)), Gaussian heating, 6 Dirichlet faces (20 steps)
  c4  128^3 kappa=100(1+0.1U(-1,1)), 8 heaters x 1e-5 W, 64 probes             (200 steps)
  c5  384^3 layered slabs + spherical inclusions, 32 heaters, 256 probes          (500 steps)
Any builder takes `n` to scale the grid down (parity tests) with the same physics.
"""
from __future__ import annotations

import numpy as np

from .system import Params, System

RHO_A, CP_A = 2330.0, 700.0          # rho c_p = 1.631e6 J/(m^3 K)
RHO_B, CP_B = 1200.0, 1500.0


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def c1(n: int = 16, icattaneo: int = 1):
    """16^3, 1 um cells, kappa=150, tau=1e-12, x faces 300 K (kappa_b=150), y/z insulating,
    one heater at 1-based (n/2+1, n/2+1, n/2+1), power_in=1e-4 W, dt=1e-10, 100 steps."""
    L = n * 1.0e-6
    p = Params(nx=n, ny=n, nz=n, time_step=1.0e-10, icattaneo=icattaneo, isteady=0,
               kappa_bound=(150.0, 150.0, 0.0, 0.0, 0.0, 0.0), t_bath=(300.0,) * 6,
               power_in=1.0e-4, t_system=300.0, ntime=100)
    ih = np.zeros((n, n, n), np.int64)
    c = n // 2
    ih[c, c, c] = 1
    s = System.uniform(p, (L, L, L), kappa=150.0, rho=RHO_A, heat_capacity=CP_A, tau=1.0e-12,
                       iheater=ih.reshape(-1))
    return s, dict(name="c1", steps=100, tol=1.0e-10, heater=(c, c, c))


def c2(n: int = 256, nheat: int = 32, seed_mask: int = 2, seed_heat: int = 21):
    """Two-phase voxels: Bernoulli(0.3) -> phase B (kappa 0.2, rho 1200, c_p 1500), else phase
    A (kappa 150, rho 2330, c_p 700); 0.5 um cells; tau=1e-12; x faces 300 K (kappa_b=150),
    others insulating; `nheat` distinct random heater cells, 1e-5 W each; dt=5e-11, 50 steps."""
    N = n ** 3
    L = n * 0.5e-6
    mask = _rng(seed_mask).random(N) < 0.3
    kappa = np.where(mask, 0.2, 150.0)
    rho = np.where(mask, RHO_B, RHO_A)
    cp = np.where(mask, CP_B, CP_A)
    heat = _rng(seed_heat).choice(N, size=min(nheat, N), replace=False)
    ih = np.zeros(N, np.int64)
    ih[heat] = 1
    p = Params(nx=n, ny=n, nz=n, time_step=5.0e-11, icattaneo=1, isteady=0,
               kappa_bound=(150.0, 150.0, 0.0, 0.0, 0.0, 0.0), t_bath=(300.0,) * 6,
               power_in=1.0e-5 * len(heat), t_system=300.0, ntime=50)
    s = System.uniform(p, (L, L, L), kappa=kappa, rho=rho, heat_capacity=cp, tau=1.0e-12,
                       iheater=ih)
    return s, dict(name="c2", steps=50, tol=1.0e-10, seeds=dict(mask=seed_mask, heat=seed_heat))


def gaussian_qdens(n: int, sigma: float = 32.0, peak: float = 1.0e12) -> np.ndarray:
    """peak * exp(-|r - r0|^2 / (2 sigma^2)), r in cells, r0 = ((n-1)/2,)*3 (x fastest)."""
    r = np.arange(n, dtype=np.float64) - (n - 1) / 2.0
    g = np.exp(-(r * r) / (2.0 * sigma * sigma))
    return (peak * g[:, None, None] * g[None, :, None] * g[None, None, :]).reshape(-1)


def c3(n: int = 512, seed: int = 3, sigma: float = 32.0):
    """Log-normal kappa = exp(N(ln 50, 1)) i.i.d. per cell, rho c_p = 1.631e6, tau=1e-12,
    0.5 um cells, Dirichlet 300 K on all six faces (kappa_b=50), Gaussian heating
    (sigma=32 cells, peak 1e12 W/m^3) as a per-cell power density; dt=5e-11, 20 steps."""
    N = n ** 3
    L = n * 0.5e-6
    kappa = np.exp(_rng(seed).normal(np.log(50.0), 1.0, N))
    p = Params(nx=n, ny=n, nz=n, time_step=5.0e-11, icattaneo=1, isteady=0,
               kappa_bound=(50.0,) * 6, t_bath=(300.0,) * 6, t_system=300.0, ntime=20)
    s = System.uniform(p, (L, L, L), kappa=kappa, rho=RHO_A, heat_capacity=CP_A, tau=1.0e-12,
                       qdens_extra=gaussian_qdens(n, sigma * n / 512.0))
    return s, dict(name="c3", steps=20, tol=1.0e-10, seeds=dict(kappa=seed))


def near_cells(n: int, centres: np.ndarray, radius: int) -> np.ndarray:
    """Linear indices of the cells within Chebyshev distance `radius` of any centre (n^3 grid,
    x fastest), sorted."""
    out = set()
    for c in centres:
        x, y, z = int(c % n), int((c // n) % n), int(c // (n * n))
        for dz in range(-radius, radius + 1):
            for dy in range(-radius, radius + 1):
                for dx in range(-radius, radius + 1):
                    X, Y, Z = x + dx, y + dy, z + dz
                    if 0 <= X < n and 0 <= Y < n and 0 <= Z < n:
                        out.add(X + n * (Y + n * Z))
    return np.array(sorted(out), np.int64)


def c4(n: int = 128, icattaneo: int = 1, nheat: int = 8, nprobe: int = 64, seed_k: int = 4,
       seed_kstar: int = 40, seed_probe: int = 41, seed_heat: int = 42, t_ref: float = 300.0):
    """kappa = 100 (1 + 0.1 U(-1,1)) (trainable), 1 um cells, rho c_p = 1.631e6, BCs as c1,
    `nheat` heaters of 1e-5 W each entering as per-cell power density P/V (trainable),
    `nprobe` probe cells drawn from the heaters' neighbourhoods (Chebyshev radius 2: over 200
    steps of 0.1 ns heat diffuses about one cell, so farther probes would not see the heaters),
    targets from a ground-truth kappa* (seed_kstar); dt=1e-10, 200 steps, checkpoints every 20.
    `t_ref` is the bath and initial temperature (300 K); the operator is linear, so t_ref = 0
    gives the same field minus 300 K, resolved to ~1e-17 K instead of ~6e-14 K.
    Returns the system (with kappa) and kappa*, probes, powers."""
    N = n ** 3
    L = n * 1.0e-6
    kappa = 100.0 * (1.0 + 0.1 * _rng(seed_k).uniform(-1.0, 1.0, N))
    kstar = 100.0 * (1.0 + 0.1 * _rng(seed_kstar).uniform(-1.0, 1.0, N))
    # heaters away from the domain faces
    lo, hi = (2, n - 2) if n > 6 else (0, n)
    side = hi - lo
    hsel = _rng(seed_heat).choice(side ** 3, size=min(nheat, side ** 3), replace=False)
    hx, hy, hz = hsel % side + lo, (hsel // side) % side + lo, hsel // (side * side) + lo
    heat = (hx + n * (hy + n * hz)).astype(np.int64)
    near = near_cells(n, heat, 2)
    probes = _rng(seed_probe).choice(near, size=min(nprobe, len(near)), replace=False).astype(np.int64)
    vol = (L / n) ** 3
    power = np.zeros(N)
    power[heat] = 1.0e-5
    p = Params(nx=n, ny=n, nz=n, time_step=1.0e-10, icattaneo=icattaneo, isteady=0,
               kappa_bound=(150.0, 150.0, 0.0, 0.0, 0.0, 0.0), t_bath=(t_ref,) * 6,
               t_system=t_ref, ntime=200)
    s = System.uniform(p, (L, L, L), kappa=kappa, rho=RHO_A, heat_capacity=CP_A,
                       tau=1.0e-12 if icattaneo else 0.0, qdens_extra=power / vol)
    return s, dict(name="c4", steps=200, ckpt=20, kstar=kstar, probes=probes, heaters=heat,
                   power=power, volume=vol, near=near,
                   seeds=dict(kappa=seed_k, kstar=seed_kstar, probes=seed_probe, heat=seed_heat))


def c5(n: int = 384, nheat: int = 32, nprobe: int = 256, seed_geom: int = 5,
       seed_heat: int = 52, seed_probe: int = 51):
    """Layered stack: 8 z-slabs of n/8 planes with kappa cycling 150, 1.4, 30 W/(m K), plus
    seeded spherical inclusions (radius 4..12 cells) of kappa 0.2 until 5% of the volume;
    rho c_p = 1.631e6, tau = 1e-12, 0.5 um cells, Dirichlet 300 K on all six faces
    (kappa_b = 50), `nheat` heaters of 1e-5 W (per-cell power density), `nprobe` probes in the
    heaters' neighbourhoods (radius 2); dt = 5e-11 s, 500 steps."""
    N = n ** 3
    L = n * 0.5e-6
    kap = np.empty((n, n, n))                      # (z, y, x)
    vals = (150.0, 1.4, 30.0)
    edges = np.linspace(0, n, 9).astype(int)
    for s in range(8):
        kap[edges[s]:edges[s + 1]] = vals[s % 3]
    rng = _rng(seed_geom)
    incl = np.zeros((n, n, n), bool)
    target = int(0.05 * N)
    count = 0
    while count < target:
        c = rng.integers(0, n, 3)
        r = int(rng.integers(4, 13))
        lo = np.maximum(c - r, 0)
        hi = np.minimum(c + r + 1, n)
        zz, yy, xx = np.ogrid[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        ball = (zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2 <= r * r
        sub = incl[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        count += int(np.count_nonzero(ball & ~sub))
        sub |= ball
    kap[incl] = 0.2
    kappa = kap.reshape(-1)
    lo_, hi_ = (2, n - 2) if n > 6 else (0, n)
    side = hi_ - lo_
    hsel = _rng(seed_heat).choice(side ** 3, size=min(nheat, side ** 3), replace=False)
    hx, hy, hz = hsel % side + lo_, (hsel // side) % side + lo_, hsel // (side * side) + lo_
    heat = (hx + n * (hy + n * hz)).astype(np.int64)
    near = near_cells(n, heat, 2)
    probes = _rng(seed_probe).choice(near, size=min(nprobe, len(near)), replace=False).astype(np.int64)
    vol = (L / n) ** 3
    power = np.zeros(N)
    power[heat] = 1.0e-5
    p = Params(nx=n, ny=n, nz=n, time_step=5.0e-11, icattaneo=1, isteady=0,
               kappa_bound=(50.0,) * 6, t_bath=(300.0,) * 6, t_system=300.0, ntime=500)
    s = System.uniform(p, (L, L, L), kappa=kappa, rho=RHO_A, heat_capacity=CP_A, tau=1.0e-12,
                       qdens_extra=power / vol)
    return s, dict(name="c5", steps=500, probes=probes, heaters=heat, power=power, volume=vol,
                   near=near, inclusion_fraction=count / N,
                   seeds=dict(geometry=seed_geom, heat=seed_heat, probes=seed_probe))

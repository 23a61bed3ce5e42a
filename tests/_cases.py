"""Shared system builders for the CPU (oracle) and GPU parity tests."""
from __future__ import annotations

import json
import os

import numpy as np

from heatflow_amd.system import Params, System

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def random_system(seed: int, nx: int, ny: int, nz: int, periodic=(False, False, False),
                  cg=(False,) * 6, icattaneo: int = 1, isteady: int = 0, nheat: int = 3,
                  modes=(1,), em: float = 0.0, br: float = 0.7, equal_frac: float = 0.2,
                  dt: float = 1.0e-10, tiny_cells: int = 0, qextra: bool = False) -> System:
    """Random per-cell materials with every path feature switchable.  `equal_frac` of the
    cells share kappa=150 (exercises the BR switch on equal neighbours); `tiny_cells` cells get
    kappa=0 so their couplings fall under TINY and are dropped from the COO."""
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    p = Params(nx=nx, ny=ny, nz=nz, time_step=dt, icattaneo=icattaneo, isteady=isteady,
               periodic=periodic, cg=cg, t_bath_cg=5.0,
               kappa_bound=tuple(rng.uniform(0, 200, 6)), t_bath=tuple(rng.uniform(280, 320, 6)),
               br=br, power_in=1e-4, freq=1.0e9, heated_steps=3, t_bath_rad=290.0)
    ih = np.zeros(N, np.int64)
    if nheat:
        ih[rng.choice(N, min(nheat, N), replace=False)] = rng.choice(modes, min(nheat, N))
    kap = rng.uniform(1, 200, N)
    kap[rng.random(N) < equal_frac] = 150.0
    if tiny_cells:
        kap[rng.choice(N, tiny_cells, replace=False)] = 1.0e-30
    qe = rng.uniform(0, 1e12, N) if qextra else None
    return System.uniform(p, (nx * 1e-6, ny * 1.3e-6, nz * 0.7e-6), kappa=kap,
                          rho=rng.uniform(1000, 3000, N), heat_capacity=rng.uniform(500, 1500, N),
                          tau=1e-12, em=em, iheater=ih, qdens_extra=qe)


def random_state(seed: int, n: int):
    rng = np.random.default_rng(seed + 99)
    tp = rng.uniform(290, 310, n)
    tpp = tp + rng.uniform(-1, 1, n)
    return tp, tpp


def manual_cases():
    with open(os.path.join(GOLDEN, "manual_cases.json")) as f:
        d = json.load(f)
    out = []
    for c in d["cases"]:
        p = Params(nx=3, ny=1, nz=1, time_step=1.0, icattaneo=0, isteady=1,
                   kappa_bound=tuple(float(v) for v in d["kappa_bound"]),
                   t_bath=(d["t_bath"],) * 6, power_in=c["qdens_centre"], t_system=d["t_bath"])
        ih = np.array([0, 1 if c["qdens_centre"] else 0, 0], np.int64)
        s = System.uniform(p, (3.0, 1.0, 1.0), kappa=np.array(c["kappa"], float), rho=1.0,
                           heat_capacity=1.0, tau=0.0, iheater=ih)
        out.append((c, s))
    return out


def read_tempdis(path: str) -> np.ndarray:
    with open(path) as f:
        return np.array(f.read().split(), np.float64)


# geometry variants covering the reference's sparsity rules
GEOMETRIES = [
    dict(nx=4, ny=3, nz=5),
    dict(nx=5, ny=4, nz=3, periodic=(True, False, True)),
    dict(nx=6, ny=1, nz=4),
    dict(nx=5, ny=5, nz=1, cg=(True, False, False, True, False, False)),
    dict(nx=4, ny=4, nz=4, em=0.5, modes=(1, 2, 3, 4, 5, 6)),
    dict(nx=3, ny=3, nz=3, periodic=(True, True, True)),
    dict(nx=7, ny=1, nz=1),
    dict(nx=6, ny=5, nz=4, tiny_cells=6),
    dict(nx=5, ny=3, nz=3, periodic=(False, True, False)),
    dict(nx=4, ny=5, nz=3, icattaneo=0, qextra=True),
]


_M1, _M2, _M3 = np.uint64(0x9e3779b97f4a7c15), np.uint64(0xbf58476d1ce4e5b9), np.uint64(0x94d049bb133111eb)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z + _M1
    z = (z ^ (z >> np.uint64(30))) * _M2
    z = (z ^ (z >> np.uint64(27))) * _M3
    return z ^ (z >> np.uint64(31))


def coo_hash(val: np.ndarray, irow: np.ndarray, jcol: np.ndarray, chunk: int = 1 << 25,
             k0: int = 0) -> int:
    """The order-sensitive COO hash of the oracle (hfo_coo_hash, hf_oracle.h): the sum over
    positions k of mix(mix(mix(mix(k) + irow) + jcol) + value bits), mod 2^64, in chunks; the
    arrays hold positions k0, k0 + 1, ..."""
    total = np.uint64(0)
    with np.errstate(over="ignore"):
        for a in range(0, len(val), chunk):
            b = min(a + chunk, len(val))
            k = np.arange(k0 + a, k0 + b, dtype=np.uint64)
            h = _mix64(_mix64(_mix64(_mix64(k) + irow[a:b].view(np.uint64))
                              + jcol[a:b].view(np.uint64)) + val[a:b].view(np.uint64))
            total = total + np.sum(h, dtype=np.uint64)
    return int(total)

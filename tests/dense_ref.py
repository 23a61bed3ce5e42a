"""Dense FP64 torch restatement of the HeatFlow time step, differentiable by torch autograd.

Test infrastructure: an independent check of the hand-written adjoint on small grids.  It
assembles H densely from kappa with the harmonic-mean face conductance of hmatrixfunc
(src/heatflow/mod_hmatrix.f90:288-314, 357-385; uniform per-axis lengths) and steps
H T^{n+1} = S^n (src/heatflow/mod_evolve.f90:112-119) with an LU solve.  Periodic axes couple
the last cell of a line to the first (mod_hmatrix.f90:196-246) and have no bath faces; a CG face
keeps its bath conductance on the diagonal and assigns B_I = T_BathCG / (V n_a n_b) in the face
order x1, xN, y1, yN, z1, zN (mod_boundary.f90:43-66, 81-82), so a later face's bath term adds
to it and a later CG face replaces it.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch


# boundary variants of the adjoint checks: periodic axes (wrap couplings, no bath faces) and CG
# faces (B assigned in face order; a later bath face adds to it)
DENSE_BOUNDARIES = [
    dict(),
    dict(periodic=(True, False, False)),
    dict(periodic=(False, True, True)),
    dict(cg=(True, False, False, True, False, False), t_bath_cg=2.0e-9),
    dict(periodic=(False, False, True), cg=(False, True, True, False, False, False),
         t_bath_cg=5.0e-10),
]


def with_boundary(s, bi):
    """The system s with the boundary variant DENSE_BOUNDARIES[bi]."""
    from heatflow_amd.system import System
    return System(params=dataclasses.replace(s.params, **DENSE_BOUNDARIES[bi]), cells=s.cells)


def dense_run(system, kappa: torch.Tensor, power: torch.Tensor, nsteps: int,
              t_init: float | None = None) -> torch.Tensor:
    p, c = system.params, system.cells
    nx, ny, nz = p.nx, p.ny, p.nz
    N = p.n
    dt = p.time_step
    L = (float(c.len_x[0]), float(c.len_y[0]), float(c.len_z[0]))
    vol = torch.as_tensor(c.volume, dtype=torch.float64)
    rc = torch.as_tensor(c.rho * c.heat_capacity, dtype=torch.float64)
    tau_s = torch.as_tensor(c.tau_scaled(dt), dtype=torch.float64)
    idx = np.arange(N)
    x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    H = torch.zeros((N, N), dtype=torch.float64)
    B = torch.zeros(N, dtype=torch.float64)
    strides = (1, nx, nx * ny)
    coords = (x, y, z)
    dims = (nx, ny, nz)
    periodic = tuple(bool(v) for v in p.periodic)
    cg = tuple(bool(v) for v in p.cg)
    vol0 = float(c.volume[0])
    for ax in range(3):
        m = coords[ax] < dims[ax] - 1
        I = torch.as_tensor(idx[m])
        J = I + strides[ax]
        if periodic[ax]:
            mw = coords[ax] == dims[ax] - 1
            I = torch.cat([I, torch.as_tensor(idx[mw])])
            J = torch.cat([J, torch.as_tensor(idx[mw] - (dims[ax] - 1) * strides[ax])])
        ki, kj = kappa[I], kappa[J]
        G = 2.0 * ki * kj / ((ki + kj) * L[ax] * L[ax])
        G = torch.where(ki != kj, G * p.br, G)
        H = H.index_put((I, J), G, accumulate=True)
        H = H.index_put((J, I), G, accumulate=True)
        H = H.index_put((I, I), -G, accumulate=True)
        H = H.index_put((J, J), -G, accumulate=True)
    for f in range(6):
        ax, side = f // 2, f % 2
        if periodic[ax]:
            continue
        mb = coords[ax] == (0 if side == 0 else dims[ax] - 1)
        Ib = torch.as_tensor(idx[mb])
        kb = p.kappa_bound[f]
        k = kappa[Ib]
        gb = 2.0 * kb * k / ((kb + k) * L[ax] * L[ax])
        gb = torch.where(k != kb, gb * p.br, gb)
        H = H.index_put((Ib, Ib), -gb, accumulate=True)
        if cg[f]:
            other = [dims[a] for a in range(3) if a != ax]
            val = p.t_bath_cg / (vol0 * other[0] * other[1])
            B = B.index_put((Ib,), torch.full((len(Ib),), val, dtype=torch.float64))
        else:
            B = B.index_put((Ib,), gb * p.t_bath[f], accumulate=True)
    cvec = tau_s * rc if p.icattaneo == 1 else torch.zeros(N, dtype=torch.float64)
    alpha = cvec + rc / dt
    H = H - torch.diag(alpha)
    qd = power / vol
    lu, piv = torch.linalg.lu_factor(H)
    t0 = p.t_system if t_init is None else t_init
    tp = torch.full((N,), t0, dtype=torch.float64)
    tpp = tp.clone()
    for _ in range(nsteps):
        S = -(tp * rc / dt) - qd - B + cvec * (tpp - 2.0 * tp)
        tn = torch.linalg.lu_solve(lu, piv, S.unsqueeze(1)).squeeze(1)
        tpp, tp = tp, tn
    return tp

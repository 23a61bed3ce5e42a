"""Dense FP64 torch restatement of the HeatFlow time step, differentiable by torch autograd.

Test infrastructure: an independent check of the hand-written adjoint on small grids.  It
assembles H densely from kappa with the harmonic-mean face conductance of hmatrixfunc
(src/heatflow/mod_hmatrix.f90:288-314, 357-385; uniform per-axis lengths, no periodic axes, no
CG faces) and steps H T^{n+1} = S^n (src/heatflow/mod_evolve.f90:112-119) with an LU solve.
"""
from __future__ import annotations

import numpy as np
import torch


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
    for ax in range(3):
        m = coords[ax] < dims[ax] - 1
        I = torch.as_tensor(idx[m])
        J = I + strides[ax]
        ki, kj = kappa[I], kappa[J]
        G = 2.0 * ki * kj / ((ki + kj) * L[ax] * L[ax])
        G = torch.where(ki != kj, G * p.br, G)
        H = H.index_put((I, J), G, accumulate=True)
        H = H.index_put((J, I), G, accumulate=True)
        H = H.index_put((I, I), -G, accumulate=True)
        H = H.index_put((J, J), -G, accumulate=True)
        for side in (0, 1):
            f = 2 * ax + side
            mb = coords[ax] == (0 if side == 0 else dims[ax] - 1)
            Ib = torch.as_tensor(idx[mb])
            kb = p.kappa_bound[f]
            k = kappa[Ib]
            gb = 2.0 * kb * k / ((kb + k) * L[ax] * L[ax])
            gb = torch.where(k != kb, gb * p.br, gb)
            H = H.index_put((Ib, Ib), -gb, accumulate=True)
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

"""GPU adjoint: the hand-written reverse-time gradient against dense FP64 torch autograd
(tests/dense_ref.py, itself pinned to the oracle in test_oracle.py) and against central finite
differences of the GPU forward."""
import numpy as np
import pytest
import torch

from heatflow_amd import Simulation, workloads
from heatflow_amd.autograd import heatflow_run

from dense_ref import dense_run

pytestmark = pytest.mark.gpu


def _loss_np(T, probes, targets):
    return float(np.sum((T[probes] - targets) ** 2))


@pytest.mark.parametrize("icat", [0, 1])
def test_adjoint_matches_dense_autograd(icat):
    """12^3 copy of the 128^3 gradient workload, 20 steps, checkpoints every 6 (a partial last
    segment): dL/dkappa over all cells and dL/dP at the heaters."""
    s, m = workloads.c4(12, icattaneo=icat, nheat=4, nprobe=24)
    steps, ckpt = 20, 6
    # targets from the ground-truth kappa* (dense forward)
    with torch.no_grad():
        tstar = dense_run(s, torch.as_tensor(m["kstar"]), torch.as_tensor(m["power"]), steps)
    targets = tstar.numpy()[m["probes"]] - 0.01   # a well-conditioned loss: T - t ~ 1e-2
    k = torch.tensor(s.cells.kappa, requires_grad=True)
    pw = torch.tensor(m["power"], requires_grad=True)
    T = dense_run(s, k, pw, steps)
    loss = ((T[torch.as_tensor(m["probes"])] - torch.as_tensor(targets)) ** 2).sum()
    loss.backward()
    with Simulation(s, tol=1e-32) as sim:
        sim.initial_evolve()
        L, gk, gp = sim.grad(1, steps, ckpt, m["probes"], targets)
    assert loss.item() > 1e-4
    # T - t ~ 1e-2: the loss carries the temperatures' round-off (~1e-14 relative) x 1e4
    assert abs(L - loss.item()) <= 1e-7 * abs(loss.item())
    gk_ref = k.grad.numpy()
    gp_ref = pw.grad.numpy()
    scale = np.abs(gk_ref).max()
    assert np.abs(gk - gk_ref).max() <= 1e-7 * scale
    h = m["heaters"]
    np.testing.assert_allclose(gp[h], gp_ref[h], rtol=1e-7)


def test_autograd_function_on_device():
    """heatflow_run (torch.autograd.Function) on cuda tensors gives the same gradients as the
    probe-loss convenience entry point."""
    s, m = workloads.c4(16, nprobe=16)
    steps, ckpt = 30, 7
    sim = Simulation(s, tol=1e-32)
    try:
        dev = torch.device("cuda:0")
        k = torch.tensor(s.cells.kappa, device=dev, requires_grad=True)
        pw = torch.tensor(m["power"], device=dev, requires_grad=True)
        targets = torch.full((len(m["probes"]),), 300.01, dtype=torch.float64, device=dev)
        pr = torch.as_tensor(m["probes"], device=dev)
        T = heatflow_run(k, pw, sim, steps, ckpt)
        loss = ((T[pr] - targets) ** 2).sum()
        loss.backward()
        sim.initial_evolve()
        L, gk, gp = sim.grad(1, steps, ckpt, m["probes"], targets.cpu().numpy())
        assert abs(L - loss.item()) <= 1e-12 * abs(L)
        np.testing.assert_allclose(k.grad.cpu().numpy(), gk, rtol=1e-12,
                                   atol=1e-12 * np.abs(gk).max())
        np.testing.assert_allclose(pw.grad.cpu().numpy(), gp, rtol=1e-12,
                                   atol=1e-12 * np.abs(gp).max())
    finally:
        sim.close()


@pytest.mark.parametrize("icat", [0, 1])
def test_adjoint_central_fd(icat):
    """Central differences (relative step 1e-6) of the GPU forward, 32^3, 40 steps, on 16
    seeded cells next to the heaters and on 4 heater powers.  A central difference of L carries
    the forward's round-off: |dL_noise| <= 2 sum|T - t| * dT with dT ~ 1e-13 * 300 K (the
    FP64 solve converges to ~1e-16 relative; 1e-13 leaves margin), so the bar for each cell is
    1e-6 relative plus that floor divided by 2h."""
    s, m = workloads.c4(32, icattaneo=icat, nprobe=64)
    steps, ckpt = 40, 10
    probes = m["probes"]
    sim = Simulation(s, tol=1e-32)
    try:
        sim.initial_evolve()
        sim.run(1, steps)
        T0, _ = sim.get_state()
        targets = T0[probes] - 0.01
        sim.initial_evolve()
        L, gk, gp = sim.grad(1, steps, ckpt, probes, targets)
        near1 = workloads.near_cells(32, m["heaters"], 1)
        cells = np.random.default_rng(43).choice(near1, 16, replace=False)
        kap = s.cells.kappa.copy()
        vol = s.cells.volume

        def loss_at(kv, pv):
            sim.set_fields(kv, pv / vol)
            sim.sparse_Hmatrix()
            sim.initial_evolve()
            sim.run(1, steps)
            T, _ = sim.get_state()
            return _loss_np(T, probes, targets), T

        L0, T0b = loss_at(kap, m["power"])
        assert abs(L0 - L) <= 1e-12 * L
        noise = 2.0 * np.sum(np.abs(T0b[probes] - targets)) * 1e-13 * 300.0
        for c in cells:
            h = 1e-6 * kap[c]
            kp, km = kap.copy(), kap.copy()
            kp[c] += h
            km[c] -= h
            fd = (loss_at(kp, m["power"])[0] - loss_at(km, m["power"])[0]) / (2 * h)
            assert abs(fd - gk[c]) <= 1e-6 * abs(gk[c]) + noise / (2 * h), (c, fd, gk[c])
        for hc in m["heaters"][:4]:
            hp = 1e-6 * m["power"][hc]
            pp, pm = m["power"].copy(), m["power"].copy()
            pp[hc] += hp
            pm[hc] -= hp
            fdp = (loss_at(kap, pp)[0] - loss_at(kap, pm)[0]) / (2 * hp)
            assert abs(fdp - gp[hc]) <= 1e-6 * abs(gp[hc]) + noise / (2 * hp), (hc, fdp, gp[hc])
    finally:
        sim.close()

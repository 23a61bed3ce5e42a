"""GPU adjoint: the hand-written reverse-time gradient against dense FP64 torch autograd
(tests/dense_ref.py, itself pinned to the oracle in test_oracle.py) and against central finite
differences of the GPU forward."""
import numpy as np
import pytest
import torch

from heatflow_amd import HeatFlowError, Simulation, workloads
from heatflow_amd.autograd import heatflow_run

from _cases import random_system
from dense_ref import DENSE_BOUNDARIES, dense_run, radiative_system, with_boundary

pytestmark = pytest.mark.gpu


def _loss_np(T, probes, targets):
    return float(np.sum((T[probes] - targets) ** 2))


@pytest.mark.parametrize("bi", range(len(DENSE_BOUNDARIES)))
@pytest.mark.parametrize("icat", [0, 1])
def test_adjoint_matches_dense_autograd(icat, bi):
    """12^3 copy of the 128^3 gradient workload, 20 steps, checkpoints every 6 (a partial last
    segment): dL/dkappa over all cells and dL/dP at the heaters.  Boundary variants: Dirichlet
    baths, periodic axes (the flat adjoint kernel with its wrap terms) and CG faces (bath terms
    that a CG face's assignment removes from B carry no dB/dkappa)."""
    s, m = workloads.c4(12, icattaneo=icat, nheat=4, nprobe=24)
    s = with_boundary(s, bi)
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


def test_autograd_refuses_a_stale_graph():
    """A second heatflow_run on the same Simulation replaces the checkpoints the first graph's
    backward would need: that backward raises; the second graph's still works."""
    s, m = workloads.c4(12, nprobe=8)
    with Simulation(s, tol=1e-32) as sim:
        k = torch.tensor(s.cells.kappa, requires_grad=True)
        pw = torch.tensor(m["power"], requires_grad=True)
        pr = torch.as_tensor(m["probes"])
        T1 = heatflow_run(k, pw, sim, 6, 3)
        T2 = heatflow_run(k, pw, sim, 6, 3)
        with pytest.raises(RuntimeError):
            T1[pr].sum().backward()
        T2[pr].sum().backward()
        assert k.grad is not None and torch.any(k.grad != 0)


@pytest.mark.parametrize("icat", [0, 1])
def test_adjoint_central_fd(icat):
    """Central differences of the GPU forward, 32^3, 40 steps: the 16 cells with the largest
    |dL/dkappa|, kappa along a seeded direction, and 4 heater powers, each to 1e-6 relative;
    targets come from the ground-truth kappa*.

    The check runs on the system with temperatures measured from the bath (baths and initial
    state 0 K; the operator is linear, so the field is the 300 K field minus 300 K, checked
    here to 1e-10 K).  There the loss, sum (T - T*)^2 with T - T* ~ 1e-6..1e-5 K, is resolved
    to ~1e-17 K; at 300 K FP64 resolves T only to 6e-14 K, which bounds how well the loss and
    its gradient are defined at all.  The 300 K adjoint gradient is compared with the 0 K one
    (measured ~1e-9 apart; bar 1e-7).  Relative step 1e-5 for kappa (the h^2 truncation is
    ~1e-10 there; a 1e-6 step moves the loss by a few ulps of T at 40 steps), 1e-3 for the
    powers (T is affine in them, so the difference is exact up to round-off)."""
    s, m = workloads.c4(32, icattaneo=icat, nprobe=64)
    s0, _ = workloads.c4(32, icattaneo=icat, nprobe=64, t_ref=0.0)
    steps, ckpt = 40, 10
    probes = m["probes"]
    kap = s.cells.kappa.copy()
    vol = s.cells.volume
    sim = Simulation(s, tol=1e-32)
    sim0 = Simulation(s0, tol=1e-32)

    def final_t(sm, kv, pv):
        sm.set_fields(kv, pv / vol)
        sm.sparse_Hmatrix()
        sm.initial_evolve()
        sm.run(1, steps)
        return sm.get_state()[0]

    try:
        targets = final_t(sim, m["kstar"], m["power"])[probes]
        targets0 = final_t(sim0, m["kstar"], m["power"])[probes]
        T300 = final_t(sim, kap, m["power"])
        T0 = final_t(sim0, kap, m["power"])
        assert np.abs(T0 - (T300 - 300.0)).max() <= 1e-10
        sim.initial_evolve()
        L, gk, gp = sim.grad(1, steps, ckpt, probes, targets)
        assert abs(_loss_np(T300, probes, targets) - L) <= 1e-12 * L
        sim0.initial_evolve()
        L0, gk0, gp0 = sim0.grad(1, steps, ckpt, probes, targets0)
        assert abs(_loss_np(T0, probes, targets0) - L0) <= 1e-12 * L0
        # 300 K vs 0 K: the seed 2 (T - T*) at 300 K carries the 300 K field's round-off, dT
        dT = float(np.abs(T0[probes] - (T300[probes] - 300.0)).max()
                   + np.abs(targets0 - (targets - 300.0)).max())
        h = m["heaters"]
        ek = np.abs(gk - gk0).max() / np.abs(gk0).max()
        ep = np.abs(gp[h] - gp0[h]).max() / np.abs(gp0[h]).max()
        print(f"300 K vs 0 K: dT {dT:.3e} K, loss {abs(L - L0) / L0:.3e}, "
              f"dL/dkappa {ek:.3e}, dL/dP {ep:.3e}", flush=True)
        bound = 2 * np.abs(T0[probes] - targets0).sum() * dT + len(probes) * dT * dT
        assert abs(L - L0) <= bound + 1e-12 * L0
        assert ek <= 1e-7 and ep <= 1e-7

        def loss0(kv, pv):
            return _loss_np(final_t(sim0, kv, pv), probes, targets0)

        for c in np.argsort(-np.abs(gk0))[:16]:
            hk = 1e-5 * kap[c]
            kp, km = kap.copy(), kap.copy()
            kp[c] += hk
            km[c] -= hk
            fd = (loss0(kp, m["power"]) - loss0(km, m["power"])) / (2 * hk)
            assert abs(fd - gk0[c]) <= 1e-6 * abs(gk0[c]), (c, fd, gk0[c])
        u = np.random.default_rng(44).uniform(-1.0, 1.0, kap.size)
        eps = 1e-5
        fd = (loss0(kap * (1 + eps * u), m["power"]) - loss0(kap * (1 - eps * u), m["power"])) / (2 * eps)
        dd = float(np.dot(gk0, kap * u))
        assert abs(fd - dd) <= 1e-6 * abs(dd), (fd, dd)
        for hc in h[:4]:
            hp = 1e-3 * m["power"][hc]
            pp, pm = m["power"].copy(), m["power"].copy()
            pp[hc] += hp
            pm[hc] -= hp
            fdp = (loss0(kap, pp) - loss0(kap, pm)) / (2 * hp)
            assert abs(fdp - gp0[hc]) <= 1e-6 * abs(gp0[hc]), (hc, fdp, gp0[hc])
    finally:
        sim.close()
        sim0.close()


def fd_random_cells(n: int, probes, seed: int = 43, count: int = 16) -> np.ndarray:
    """`count` cells drawn with `seed` from the cells within one cell of a probe (the support of
    dL/dkappa: farther cells move the probes' temperatures by less than the loss resolves)."""
    near = workloads.near_cells(n, np.asarray(probes), 1)
    return np.sort(np.random.default_rng(seed).choice(near, size=count, replace=False))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("icat", [1, 0])
def test_adjoint_fd_random_cells_128(icat):
    """The 128^3 / 200-step gradient workload as configured (checkpoints every 20 steps, 8
    heaters, 64 probes, Cattaneo and Fourier): dL/dkappa on 16 random cells (seed 43, near the
    probes) against central differences of the GPU forward at a relative step of 1e-6: the
    16-cell gradient to 1e-6 relative (max |fd - g| over max |g|; measured 7e-8 Cattaneo,
    1.7e-7 Fourier), and each cell whose gradient is at least 1e-3 of the largest to 1e-5 of
    its own (measured up to 1.8e-6); differences taken on the bath-origin system as in
    test_adjoint_central_fd.  The random cells' gradients span four orders of magnitude
    (8e-15 .. 2.5e-11): a 1e-6 step in the smallest moves the loss by ~1e-24, near its
    round-off, where the relative error of the difference itself reaches ~4e-5."""
    n, steps, ckpt = 128, 200, 20
    s0, m = workloads.c4(n, icattaneo=icat, t_ref=0.0)
    probes = m["probes"]
    kap = s0.cells.kappa.copy()
    vol = s0.cells.volume
    cells = fd_random_cells(n, probes)
    with Simulation(s0, tol=1e-32) as sim:
        def final_t(kv, pv):
            sim.set_fields(kv, pv / vol)
            sim.sparse_Hmatrix()
            sim.initial_evolve()
            sim.run(1, steps)
            return sim.get_state()[0]

        targets = final_t(m["kstar"], m["power"])[probes]
        final_t(kap, m["power"])
        sim.initial_evolve()
        L, gk, gp = sim.grad(1, steps, ckpt, probes, targets)
        fds = []
        for c in cells:
            hk = 1e-6 * kap[c]
            kp, km = kap.copy(), kap.copy()
            kp[c] += hk
            km[c] -= hk
            lp = _loss_np(final_t(kp, m["power"]), probes, targets)
            lm = _loss_np(final_t(km, m["power"]), probes, targets)
            fds.append((lp - lm) / (2 * hk))
        fd, g = np.array(fds), gk[cells]
        gmax = np.abs(g).max()
        rel = np.abs(fd - g) / np.abs(g)
        big = np.abs(g) >= 1e-3 * gmax
        print(f"128^3 icattaneo={icat}: FD on {len(cells)} random cells, |g| "
              f"{np.abs(g).min():.3e}..{gmax:.3e}; max |fd - g| / max |g| "
              f"{np.abs(fd - g).max() / gmax:.3e}; per cell (|g| >= 1e-3 max, {big.sum()} "
              f"cells) max {rel[big].max():.3e}; all cells max {rel.max():.3e}", flush=True)
        assert np.abs(fd - g).max() <= 1e-6 * gmax
        assert big.sum() >= 8 and rel[big].max() <= 1e-5, rel


def test_adjoint_mixed_precision_matches_fp64():
    """FP32-inner / FP64-refined forward and adjoint solves give the FP64 gradient to the
    solver tolerance (layered-slab workload with inclusions, scaled to 24^3, 40 steps).

    Both modes stop at the same scaled residual, 1e-13.  The loss difference is bounded by
    the measured probe temperature difference dT: |L32 - L64| <= 2 sum|T64 - t| dT + m dT^2."""
    s, m = workloads.c5(24, nheat=4, nprobe=40)
    probes = m["probes"]
    steps = 40
    with Simulation(s, precision="fp64", tol=1e-13) as sim:
        sim.initial_evolve()
        sim.run(1, steps)
        t_end, _ = sim.get_state()
    targets = t_end[probes] - 0.01
    out = {}
    for prec in ("fp32_ir", "fp64"):
        with Simulation(s, precision=prec, tol=1e-13) as sim:
            sim.initial_evolve()
            loss, gk, gp = sim.grad(1, steps, 7, probes, targets)
            tp, _ = sim.get_state()
            out[prec] = (loss, gk, gp, tp)
    L32, gk32, gp32, T32 = out["fp32_ir"]
    L64, gk64, gp64, T64 = out["fp64"]
    assert np.max(np.abs(T32 - T64) / np.abs(T64)) <= 1e-10
    dT = float(np.abs(T32[probes] - T64[probes]).max())
    bound = 2 * np.abs(T64[probes] - targets).sum() * dT + len(probes) * dT ** 2
    assert abs(L32 - L64) <= bound + 1e-14 * abs(L64)
    assert np.abs(gk32 - gk64).max() <= 1e-6 * np.abs(gk64).max()
    h = m["heaters"]
    assert np.abs(gp32[h] - gp64[h]).max() <= 1e-6 * np.abs(gp64[h]).max()


@pytest.mark.parametrize("prec", ["fp64", "fp32_ir"])
def test_adjoint_uniform_field_reads_are_exact(prec):
    """The 128^3 gradient workload has uniform rho c_p, tau' rho c_p and cell volume; the
    adjoint kernel (and the forward step kernels) then read one value for every cell, and debug
    bit 3 keeps the per-cell reads.  Loss and both gradients come out bit-identical."""
    s, m = workloads.c4(16, nprobe=16)
    targets = np.full(len(m["probes"]), 300.0)
    out = []
    for dbg in (0, 8):
        with Simulation(s, precision=prec, tol=1e-13) as sim:
            sim.set_debug(dbg)
            sim.initial_evolve()
            out.append(sim.grad(1, 12, 4, m["probes"], targets))
    (la, ka, pa), (lb, kb, pb) = out
    assert la == lb
    assert np.array_equal(ka.view(np.int64), kb.view(np.int64))
    assert np.array_equal(pa.view(np.int64), pb.view(np.int64))


def test_adjoint_tolerance_apart():
    """hf_set_adjoint_tol gives the adjoint solves their own tolerance.  With the forward at the
    reference tolerance and the adjoint at 1e-10 (scaled residual), the loss is unchanged (it
    comes from the forward alone) and the gradients stay within 1e-6 of those with the adjoint
    at the reference tolerance; a negative tolerance is refused."""
    s, m = workloads.c4(16, nprobe=16)
    targets = np.full(len(m["probes"]), 300.0)
    out = []
    for atol in (0.0, 1e-10):
        with Simulation(s, tol=1e-32) as sim:
            sim.set_adjoint_tol(atol)
            sim.initial_evolve()
            out.append(sim.grad(1, 12, 4, m["probes"], targets))
    (la, ka, pa), (lb, kb, pb) = out
    assert la == lb
    dk = np.abs(ka - kb).max() / np.abs(ka).max()
    h = m["heaters"]
    dp = np.abs(pa[h] - pb[h]).max() / np.abs(pa[h]).max()
    assert 0.0 < dk <= 1e-6 and dp <= 1e-6, (dk, dp)
    with Simulation(s) as sim:
        with pytest.raises(HeatFlowError):
            sim.set_adjoint_tol(-1.0)



ADJ_KERNEL_CASES = [
    dict(nx=130, ny=9, nz=5),                                    # tiles cut in x and y
    dict(nx=130, ny=9, nz=5, icattaneo=0, cg=(True, False, False, True, False, True)),
    dict(nx=6, ny=1, nz=4),
    dict(nx=7, ny=1, nz=1),
    dict(nx=5, ny=5, nz=1, cg=(True, False, False, True, False, False)),
    dict(nx=20, ny=17, nz=70, isteady=1),
]


@pytest.mark.parametrize("ci", range(len(ADJ_KERNEL_CASES)))
def test_adjoint_step_kernels_agree(ci):
    """The adjoint step runs as a z-march (the neighbours' lambda, T and kappa staged in LDS) on
    grids without a periodic axis; debug bit 4 runs the flat per-cell kernel.  Both sum the same
    face terms in the same order, so loss and gradients come out bit-identical."""
    s = random_system(800 + ci, **ADJ_KERNEL_CASES[ci])
    rng = np.random.default_rng(ci)
    probes = rng.choice(s.params.n, min(s.params.n, 12), replace=False)
    targets = np.full(len(probes), 300.0)
    out = []
    for dbg in (0, 16):
        with Simulation(s, precision="fp64", tol=1e-14) as sim:
            sim.set_debug(dbg)
            sim.initial_evolve()
            out.append(sim.grad(1, 10, 4, probes, targets))
    (la, ka, pa), (lb, kb, pb) = out
    assert la == lb
    assert np.array_equal(ka.view(np.int64), kb.view(np.int64))
    assert np.array_equal(pa.view(np.int64), pb.view(np.int64))
    assert np.any(ka != 0.0)


@pytest.mark.parametrize("prec", ["fp64", "fp32_ir"])
def test_adjoint_guess_orders_agree(prec):
    """The adjoint solves start from the polynomial extrapolation through the last six
    solutions (fewer at first); debug bit 5 keeps the linear 2 lambda^{n+2} - lambda^{n+3}.  The
    guess changes where each solve starts, not what it converges to: loss and gradients agree to
    the solve tolerance (here 1e-13 on the scaled residual; gradients within 1e-9 of max |g|)."""
    s, m = workloads.c4(32, nprobe=32)
    targets = np.full(len(m["probes"]), 300.0)
    out, iters = [], []
    for dbg in (0, 32):
        with Simulation(s, precision=prec, tol=1e-13) as sim:
            sim.set_debug(dbg)
            sim.initial_evolve()
            out.append(sim.grad(1, 40, 8, m["probes"], targets))
            st = sim.stats()
            assert st.adjoint_solves == 40
            iters.append(st.adjoint_iters)
    (la, ka, pa), (lb, kb, pb) = out
    assert iters[0] < iters[1]   # the six-point guess starts the solves closer
    assert abs(la - lb) <= 1e-9 * abs(lb)
    assert np.max(np.abs(ka - kb)) <= 1e-9 * np.max(np.abs(kb))
    assert np.max(np.abs(pa - pb)) <= 1e-9 * np.max(np.abs(pb))
    assert not np.array_equal(ka, kb)   # the two guesses did run


@pytest.mark.parametrize("prec", ["fp64", "fp32_ir"])
def test_adjoint_checkpoint_spacing_is_exact(prec):
    """The reverse pass recomputes each segment from its checkpoint except the last step, whose
    state is the next checkpoint (or the final state); with checkpoints every step (spacing 1)
    every state is kept once and nothing is recomputed, and spacing 0 picks that when it fits in
    device memory (here it does).  FP64 steps give the same bits from a restored state as from
    the carried residual, so loss and gradients are bit-identical for every spacing (1, 0, 4, 5
    with a partial last segment, 13 = all steps in one segment); mixed-mode steps agree to the
    tolerance (1e-13 here; gradients within 1e-9 of max |g|)."""
    s, m = workloads.c4(16, icattaneo=1, nprobe=16)
    targets = np.full(len(m["probes"]), 300.0)
    out = []
    for k in (1, 0, 4, 5, 13):
        with Simulation(s, precision=prec, tol=1e-13) as sim:
            sim.initial_evolve()
            out.append(sim.grad(1, 13, k, m["probes"], targets))
    la, ka, pa = out[0]
    for lb, kb, pb in out[1:]:
        if prec == "fp64":
            assert la == lb
            assert np.array_equal(ka.view(np.int64), kb.view(np.int64))
            assert np.array_equal(pa.view(np.int64), pb.view(np.int64))
        else:
            assert abs(la - lb) <= 1e-9 * abs(lb)
            assert np.max(np.abs(ka - kb)) <= 1e-9 * np.max(np.abs(kb))
            assert np.max(np.abs(pa - pb)) <= 1e-9 * np.max(np.abs(pb))


def test_adjoint_every_state_fallback():
    """Checkpoint spacing 0 keeps every state when they fit in device memory (nothing
    recomputed: 13 steps counted for 13); when that allocation fails after all (forced by debug
    bit 10) it falls back to ceil(sqrt(13)) = 4, whose reverse sweep recomputes 9 steps.  FP64
    recomputation is exact, so all three give the same bits as spacing 1."""
    s, m = workloads.c4(16, icattaneo=1, nprobe=16)
    targets = np.full(len(m["probes"]), 300.0)
    out, steps = [], []
    for k, dbg in ((1, 0), (0, 0), (0, 1024), (4, 0)):
        with Simulation(s, precision="fp64", tol=1e-13) as sim:
            sim.set_debug(dbg)
            sim.initial_evolve()
            sim.reset_stats()
            out.append(sim.grad(1, 13, k, m["probes"], targets))
            steps.append(sim.stats().steps)
    print(f"steps counted: {steps}", flush=True)
    assert steps == [13, 13, 22, 22]
    la, ka, pa = out[0]
    for lb, kb, pb in out[1:]:
        assert la == lb
        assert np.array_equal(ka.view(np.int64), kb.view(np.int64))
        assert np.array_equal(pa.view(np.int64), pb.view(np.int64))


@pytest.mark.parametrize("prec", ["fp64", "fp32_ir"])
@pytest.mark.parametrize("icat", [0, 1])
def test_adjoint_radiative_source_matches_dense(icat, prec):
    """A heated system whose radiative source -em L1 L2 sigma (T^n^4 - T_Bath^4) / V_heated
    (src/heatflow/mod_heating.f90:121-123, 164-167) weighs in every step: the adjoint carries its
    derivative 4 em L1 L2 sigma (T^n)^3 / V_heated into the adjoint history.  Loss and gradients
    against dense FP64 autograd (which differentiates T^4 itself); the same autograd with the
    term's T held constant is 60-90% away, so a missing term cannot pass."""
    s, probes = radiative_system(8, icattaneo=icat)
    steps, ckpt = 10, 3
    k = torch.tensor(s.cells.kappa, requires_grad=True)
    pw = torch.zeros(s.params.n, dtype=torch.float64, requires_grad=True)
    T = dense_run(s, k, pw, steps)
    targets = T.detach().numpy()[probes] - 0.01
    loss = ((T[torch.as_tensor(probes)] - torch.as_tensor(targets)) ** 2).sum()
    loss.backward()
    gk_ref, gp_ref = k.grad.numpy(), pw.grad.numpy()
    k2 = torch.tensor(s.cells.kappa, requires_grad=True)
    pw2 = torch.zeros(s.params.n, dtype=torch.float64, requires_grad=True)
    T2 = dense_run(s, k2, pw2, steps, radiative_grad=False)
    ((T2[torch.as_tensor(probes)] - torch.as_tensor(targets)) ** 2).sum().backward()
    tol = 1e-32 if prec == "fp64" else 1e-13
    with Simulation(s, precision=prec, tol=tol) as sim:
        sim.initial_evolve()
        L, gk, gp = sim.grad(1, steps, ckpt, probes, targets)
    bar = 1e-7 if prec == "fp64" else 1e-6
    assert abs(L - loss.item()) <= bar * abs(loss.item())
    ek = np.abs(gk - gk_ref).max() / np.abs(gk_ref).max()
    ep = np.abs(gp - gp_ref).max() / np.abs(gp_ref).max()
    print(f"radiative adjoint {prec} icat={icat}: dL/dkappa {ek:.3e}, dL/dP {ep:.3e}", flush=True)
    assert ek <= bar and ep <= bar
    miss = np.abs(k2.grad.numpy() - gk_ref).max() / np.abs(gk_ref).max()
    assert miss > 0.1


def test_adjoint_refuses_mixed_uses():
    """The checkpoints belong to the fields, operator and solver controls of their forward run:
    re-uploading, re-assembling or changing the solver in between makes hf_adjoint fail with
    HF_ERR_STATE instead of reversing another system.  Non-uniform cell lengths are refused."""
    s, m = workloads.c4(12, nprobe=8)
    seed = np.zeros(s.params.n)
    seed[m["probes"]] = 1.0
    for change in ("assemble", "solver", "upload"):
        with Simulation(s, tol=1e-32) as sim:
            sim.initial_evolve()
            sim.forward_checkpointed(1, 6, 3)
            sim.adjoint(seed)   # valid
            if change == "assemble":
                sim.sparse_Hmatrix()
            elif change == "solver":
                sim.set_solver(1e-13, 50000, "fp64")
            else:
                sim.upload(s)
            with pytest.raises(HeatFlowError) as ei:
                sim.adjoint(seed)
            assert ei.value.code == 13
    s2 = random_system(5, 6, 5, 4)
    s2.cells.len_x[7] *= 1.5
    with Simulation(s2, tol=1e-32) as sim:
        sim.initial_evolve()
        with pytest.raises(HeatFlowError) as ei:
            sim.forward_checkpointed(1, 4, 2)
        assert ei.value.code == 10
    with Simulation(s, tol=1e-32) as sim:   # a negative spacing or no steps
        sim.initial_evolve()
        for nsteps, k in ((4, -1), (0, 2)):
            with pytest.raises(HeatFlowError) as ei:
                sim.forward_checkpointed(1, nsteps, k)
            assert ei.value.code == 10

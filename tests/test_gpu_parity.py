"""GPU parity: the gfx950 path through the C-ABI against the oracle (which is pinned to the
reference build and the reference's golden vectors in test_oracle.py)."""
import numpy as np
import pytest

import oracle
from heatflow_amd import HeatFlowError, Simulation, workloads
from heatflow_amd.system import Params, System, read_inputs

from _cases import GEOMETRIES, manual_cases, random_state, random_system, read_tempdis

pytestmark = pytest.mark.gpu

# FP64 path at the reference's tolerance (1e-32 on the recursive residual): temperatures agree
# with the oracle to round-off; the bar of the port is 1e-10 relative.
RTOL_FP64 = 1e-12


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


# ---- assembly: bit-exact sparsity, indexing and values -------------------------------------

@pytest.mark.parametrize("gi", range(len(GEOMETRIES)))
def test_coo_export_bit_exact(gi):
    s = random_system(300 + gi, **GEOMETRIES[gi])
    o = oracle.OracleSystem(s)
    n, L, cnt, val, irow, jcol = o.coo()
    with Simulation(s) as sim:
        gn, gL, gcnt, gval, girow, gjcol = sim.export_coo()
    assert (gn, gL, gcnt) == (n, L, cnt)
    assert np.array_equal(girow, irow) and np.array_equal(gjcol, jcol)
    assert np.array_equal(bits(gval), bits(val))


@pytest.mark.parametrize("n", [16, 32])
def test_coo_export_c1_grid(n):
    s, _ = workloads.c1(n)
    o = oracle.OracleSystem(s)
    ref = o.coo()
    with Simulation(s) as sim:
        got = sim.export_coo()
    assert got[:3] == ref[:3]
    for a, b in zip(got[3:], ref[3:]):
        assert np.array_equal(bits(a) if a.dtype == np.float64 else a,
                              bits(b) if b.dtype == np.float64 else b)


def test_coo_export_two_phase_tiny_dropped():
    """Two-phase voxels with near-zero conductivity cells: couplings under TINY are dropped,
    count < len and the tail keeps the (0, -2, -1) sentinel."""
    s = random_system(9, 8, 7, 6, tiny_cells=40)
    o = oracle.OracleSystem(s)
    n, L, cnt, val, irow, jcol = o.coo()
    assert cnt < L
    with Simulation(s) as sim:
        gn, gL, gcnt, gval, girow, gjcol = sim.export_coo()
    assert gcnt == cnt and gL == L
    assert np.all(girow[cnt:] == -2) and np.all(gjcol[cnt:] == -1) and np.all(gval[cnt:] == 0)
    assert np.array_equal(girow, irow) and np.array_equal(bits(gval), bits(val))


@pytest.mark.parametrize("gi", range(len(GEOMETRIES)))
def test_boundary_and_diag_bit_exact(gi):
    s = random_system(400 + gi, **GEOMETRIES[gi])
    o = oracle.OracleSystem(s)
    n, L, cnt, val, irow, jcol = o.coo()
    diag = np.zeros(n)
    m = (irow == jcol) & (irow > 0)
    diag[irow[m] - 1] = val[m]
    with Simulation(s) as sim:
        assert np.array_equal(bits(sim.vector("B")), bits(o.boundary()))
        assert np.array_equal(bits(sim.vector("diag")), bits(diag))


# ---- right-hand side -------------------------------------------------------------------------

@pytest.mark.parametrize("gi", [0, 1, 3, 4, 9])
def test_rhs_S_matches_oracle(gi):
    """S of one step (mod_evolve.f90:112-119): bit-exact for constant heaters; the radiative
    T^4 term (device pow vs host libm) is held to 1 ulp-scale relative error."""
    s = random_system(500 + gi, **GEOMETRIES[gi])
    o = oracle.OracleSystem(s)
    p, c = s.params, s.cells
    tp, tpp = random_state(gi, p.n)
    q, qd, hv = o.heater(2, tp)
    if s.cells.qdens_extra is not None:
        qd = qd + s.cells.qdens_extra
    B = o.boundary()
    rc_ = c.rho * c.heat_capacity
    S = -(p.inverse_time * tp * rc_) - qd - B
    if p.icattaneo == 1:
        S = S + o.scat(tp, tpp)
    with Simulation(s) as sim:
        sim.set_debug(1)
        sim.set_state(tp, tpp)
        sim.simulate(2)
        gS = sim.vector("S")
        gQ = sim.vector("Qdens")
    if GEOMETRIES[gi].get("em", 0.0) == 0.0:
        assert np.array_equal(bits(gS), bits(S))
        assert np.array_equal(bits(gQ), bits(qd))
    else:
        np.testing.assert_allclose(gS, S, rtol=1e-15)
        np.testing.assert_allclose(gQ, qd, rtol=1e-13, atol=1e-3)


# ---- whole steps against the oracle ------------------------------------------------------------

def test_golden_tempdis_on_gpu(golden_dir):
    s = read_inputs(golden_dir)
    gold = read_tempdis(f"{golden_dir}/TempDis.dat")
    _, otp, _, _ = oracle.OracleSystem(s).run(s.params.ntime)
    with Simulation(s) as sim:
        sim.initial_evolve()
        sim.run(1, s.params.ntime)
        tp, tpp = sim.get_state()
    assert np.abs(tp - gold).max() <= 5.0e-7 + 1e-12
    assert relerr(tp, otp) <= RTOL_FP64


@pytest.mark.parametrize("idx", range(4))
def test_manual_cases_on_gpu(idx):
    case, s = manual_cases()[idx]
    with Simulation(s) as sim:
        n, L, cnt, val, irow, jcol = sim.export_coo()
        H = np.zeros((n, n))
        m = irow > 0
        np.add.at(H, (irow[m] - 1, jcol[m] - 1), val[m])
        np.testing.assert_allclose(H, case["H"], rtol=0, atol=1e-14)
        sim.initial_evolve(200.0)
        sim.simulate(1)
        tp, _ = sim.get_state()
    np.testing.assert_allclose(tp, case["T"], rtol=1e-12)


@pytest.mark.parametrize("gi", range(len(GEOMETRIES)))
@pytest.mark.parametrize("dbg", [0, 64])
def test_steps_match_oracle(gi, dbg):
    """Four steps from a random state on every geometry (periodic axes, CG faces, 1-cell axes,
    heater modes 1-6 with radiation, near-zero conductivities, Fourier with a per-cell source)
    against the oracle: the persistent small-grid path (debug 0) and the multi-kernel path
    (debug bit 6)."""
    s = random_system(600 + gi, **GEOMETRIES[gi])
    tp0, tpp0 = random_state(gi, s.params.n)
    rc, otp, otpp, oit = oracle.OracleSystem(s).simulate_steps([1, 2, 3, 4], tp0, tpp0)
    assert rc == 0
    with Simulation(s) as sim:
        sim.set_debug(dbg)
        sim.set_state(tp0, tpp0)
        its = [sim.simulate(k)[0] for k in (1, 2, 3, 4)]
        tp, tpp = sim.get_state()
    assert relerr(tp, otp) <= RTOL_FP64 and relerr(tpp, otpp) <= RTOL_FP64
    assert all(i > 0 for i in its)


FUSED64_CASES = [dict(GEOMETRIES[i]) for i in (0, 1, 2, 3, 5, 6, 7, 8, 9)] + [
    dict(nx=5, ny=4, nz=6, isteady=1)]


def _fp64_runs(s, tp0, tpp0, nsteps):
    """(T^n, T^{n-1}, stats) of the fused FP64 step and of the unfused one (debug bit 0), both
    on the multi-kernel path (debug bit 6)."""
    out = []
    for dbg in (64, 65):
        with Simulation(s) as sim:
            sim.set_debug(dbg)
            sim.set_state(tp0, tpp0)
            sim.run(1, nsteps)
            tp, tpp = sim.get_state()
            out.append((tp, tpp, sim.stats()))
    return out


@pytest.mark.parametrize("ci", range(len(FUSED64_CASES)))
def test_fp64_fused_step_bit_exact(ci):
    """The FP64 step fuses k_finalize with the next step's right-hand side (k_step64_z) and
    carries that residual to the next step; debug bit 0 runs the kernels apart.  Same bits over
    6 steps on every geometry without radiation (periodic axes, 1-cell axes, Fourier, steady,
    near-zero conductivities), and both match the oracle."""
    s = random_system(700 + ci, **FUSED64_CASES[ci])
    tp0, tpp0 = random_state(ci, s.params.n)
    (ta, tpa, sta), (tb, tpb, _) = _fp64_runs(s, tp0, tpp0, 6)
    assert np.array_equal(bits(ta), bits(tb)) and np.array_equal(bits(tpa), bits(tpb))
    rc, otp, _, _ = oracle.OracleSystem(s).simulate_steps(list(range(1, 7)), tp0, tpp0)
    assert rc == 0
    assert relerr(ta, otp) <= RTOL_FP64


@pytest.mark.parametrize("heating", ["uniform", "gaussian"])
def test_fp64_fused_step_tiny_guess(heating):
    """The fused FP64 pass forms the next step's residual for the last TINY outcome seen (first
    guess: TINY).  Uniform heating of an insulated grid makes the first guess miss once (the
    next step starts from the state); Gaussian heating with 300 K faces keeps TINY and the
    guess never misses.  Bit-identical to the unfused kernels in both cases."""
    s, _ = workloads.c3(16)
    p = s.params
    if heating == "uniform":
        p.kappa_bound = (0.0,) * 6
        s.cells.qdens_extra = np.full(p.n, 1.0e12)
    t0 = np.full(p.n, p.t_system)
    (ta, tpa, sta), (tb, tpb, _) = _fp64_runs(s, t0, t0, 10)
    assert sta.tiny_misses == (1 if heating == "uniform" else 0)
    assert np.array_equal(bits(ta), bits(tb)) and np.array_equal(bits(tpa), bits(tpb))


def test_c1_hundred_steps():
    """16^3 Cattaneo, one centre heater, 100 steps at the reference tolerance."""
    s, meta = workloads.c1(16)
    rc, otp, otpp, _ = oracle.OracleSystem(s).run(100)
    assert rc == 0
    with Simulation(s) as sim:
        sim.initial_evolve()
        sim.run(1, 100)
        tp, tpp = sim.get_state()
    assert relerr(tp, otp) <= RTOL_FP64
    assert tp.max() > 300.0 + 1e-3        # the heater did heat


def test_c1_tol_1e12_within_gate():
    """A relative residual of 1e-10 is not enough for a 1e-10 temperature error (local errors
    around the heater exceed the 2-norm test by ~50x over 100 steps); at the benchmark
    tolerance 1e-12 the temperatures stay within the 1e-10 relative gate of the
    reference-tolerance oracle."""
    s, meta = workloads.c1(16)
    rc, otp, _, _ = oracle.OracleSystem(s).run(100)
    with Simulation(s, tol=1e-12) as sim:
        sim.initial_evolve()
        sim.run(1, 100)
        tp, _ = sim.get_state()
    assert relerr(tp, otp) <= 1e-10


def test_two_phase_c2_small():
    s, meta = workloads.c2(24, nheat=6)
    rc, otp, _, _ = oracle.OracleSystem(s).run(10)
    assert rc == 0
    with Simulation(s) as sim:
        sim.initial_evolve()
        sim.run(1, 10)
        tp, _ = sim.get_state()
    assert relerr(tp, otp) <= RTOL_FP64


def test_linbcg_dense():
    s = random_system(21, 5, 4, 4)
    o = oracle.OracleSystem(s)
    n, L, cnt, val, irow, jcol = o.coo()
    H = np.zeros((n, n))
    m = irow > 0
    np.add.at(H, (irow[m] - 1, jcol[m] - 1), val[m])
    b = np.random.default_rng(0).uniform(-1e20, 1e20, n)
    with Simulation(s) as sim:
        x, it, err = sim.linbcg(b, np.zeros(n))
    xd = np.linalg.solve(H, b)
    np.testing.assert_allclose(x, xd, rtol=1e-12, atol=1e-12 * np.abs(xd).max())
    assert it > 0


def _dense_h_solution(o, b):
    n, L, cnt, val, irow, jcol = o.coo()
    H = np.zeros((n, n))
    m = irow > 0
    np.add.at(H, (irow[m] - 1, jcol[m] - 1), val[m])
    return H, np.linalg.solve(H, b)


@pytest.mark.parametrize("itol,tol", [(1, 1e-4), (1, 1e-8), (1, 1e-12), (2, 1e-17), (2, 1e-19),
                                      (2, 1e-21)])
def test_linbcg_stop_rule_matches_oracle(itol, tol):
    """hf_solve stops on linbcg's own test (mod_Sparse.f90:418-425): itol 1 on ||r|| / ||b||,
    itol 2 on ||r / diag H|| / ||b||, after the first iteration that meets it.  On a steady
    system (no mass term; 14 to 36 iterations) the iteration count is the oracle's (BiCG on the
    COO H, which the reference build reproduces) within one, err <= tol, and the same solve
    capped one iteration earlier has err > tol."""
    s = random_system(21, 6, 5, 4, isteady=1)
    o = oracle.OracleSystem(s)
    b = np.random.default_rng(0).uniform(-1e3, 1e3, s.params.n)
    x0 = np.zeros(s.params.n)
    xo, ito, erro = o.linbcg(b, x0, itol=itol, tol=tol)
    with Simulation(s) as sim:
        x, it, err = sim.linbcg(b, x0, itol=itol, tol=tol)
        _, it_prev, err_prev = sim.linbcg(b, x0, itol=itol, tol=tol, itmax=it - 1)
    assert err <= tol < err_prev
    assert it_prev == it - 1
    assert abs(it - ito) <= 1, (it, ito)
    # the solution's error is what the stop rule allows: ||x - x*|| <= ||H^-1 W|| tol ||b||
    # (W = I for itol 1, diag H for itol 2; the recursive residual is the true one to round-off)
    H, xd = _dense_h_solution(o, b)
    W = np.eye(len(b)) if itol == 1 else np.diag(np.diag(H))
    bound = np.linalg.norm(np.linalg.solve(H, W), 2) * tol * np.linalg.norm(b)
    assert np.linalg.norm(x - xd) <= 2 * bound + 1e-12 * np.linalg.norm(xd)
    assert np.linalg.norm(xo - xd) <= 2 * bound + 1e-12 * np.linalg.norm(xd)


def test_linbcg_rejects_unknown_itol():
    s = random_system(21, 5, 4, 4)
    with Simulation(s) as sim:
        with pytest.raises(HeatFlowError) as ei:
            sim.linbcg(np.ones(s.params.n), np.zeros(s.params.n), itol=3)
        assert ei.value.code == 10


# ---- failure handling (the reference's fatal NaN guards) -----------------------------------

def _nan_system(kind):
    p = Params(nx=4, ny=3, nz=3, time_step=1e-10, icattaneo=1, kappa_bound=(150.0,) * 6,
               t_bath=(300.0,) * 6, power_in=1e-4)
    kap = np.full(36, 150.0)
    tau = np.full(36, 1e-12)
    ih = np.zeros(36, np.int64)
    ih[5] = 1
    if kind == "B":           # kappa = kappa_b = 0 on a boundary face: 0/0 in B
        p.kappa_bound = (0.0, 150.0, 150.0, 150.0, 150.0, 150.0)
        kap[0] = 0.0
    elif kind == "Q":         # NaN heater power
        p.power_in = float("nan")
    elif kind == "SCAT":      # NaN relaxation time: S_CAT is NaN first
        tau[7] = float("nan")
    return System.uniform(p, (4e-6, 3e-6, 3e-6), kappa=kap, rho=2330.0, heat_capacity=700.0,
                          tau=tau, iheater=ih)


@pytest.mark.parametrize("kind,code", [("B", 1), ("Q", 2), ("SCAT", 4)])
def test_nan_guards(kind, code):
    s = _nan_system(kind)
    with Simulation(s) as sim:
        sim.initial_evolve(300.0)
        with pytest.raises(HeatFlowError) as ei:
            sim.simulate(1)
    assert ei.value.code == code


def test_state_errors():
    s = random_system(1, 4, 4, 4)
    with Simulation(s) as sim:
        with pytest.raises(HeatFlowError) as ei:
            sim.simulate(1)                       # no state yet
        assert ei.value.code == 13


# ---- mixed precision ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [24, 40])
def test_fp32_ir_matches_fp64(n):
    """FP32 inner CG inside FP64 refinement at the 1e-10 gate against the FP64 path run to the
    reference tolerance; log-normal kappa with Gaussian heating (the 512^3 workload, scaled)."""
    s, meta = workloads.c3(n)
    with Simulation(s, tol=1e-32) as ref:
        ref.initial_evolve()
        ref.run(1, 20)
        t64, _ = ref.get_state()
    with Simulation(s, precision="fp32_ir", tol=1e-12) as sim:
        sim.initial_evolve()
        sim.run(1, 20)
        t32, _ = sim.get_state()
        st = sim.stats()
    err = np.abs(t32 - t64) / np.abs(t64)
    assert err.max() <= 1e-10
    assert st.inner_iters > 0 and st.outer_iters >= 20


# GEOMETRIES plus a steady-state case and an even nx with periodic x and y (the FP32 stencil
# kernels take two x-neighbouring cells per thread when nx is even, including the wrap pair)
MIXED_GEOMETRIES = GEOMETRIES + [dict(nx=5, ny=4, nz=3, isteady=1),
                                 dict(nx=6, ny=4, nz=3, periodic=(True, True, False)),
                                 dict(nx=8, ny=3, nz=4, cg=(False, True, True, False, False, True))]


@pytest.mark.parametrize("gi", range(len(MIXED_GEOMETRIES)))
def test_fp32_ir_steps_match_oracle(gi):
    """Mixed precision on every geometry (periodic axes, CG faces, TINY-dropped couplings,
    heater modes 1-6 with radiation, Fourier, steady state), 6 steps in one hf_run from a
    random state, against the oracle at the reference tolerance.  Per-step constant heating
    takes the fused step (refinement residual + next right-hand side in one pass, residual
    carried between steps); time-dependent heating or radiation the unfused one."""
    s = random_system(700 + gi, **MIXED_GEOMETRIES[gi])
    tp0, tpp0 = random_state(gi, s.params.n)
    steps = [1, 2, 3, 4, 5, 6]
    rc, otp, otpp, _ = oracle.OracleSystem(s).simulate_steps(steps, tp0, tpp0)
    assert rc == 0
    with Simulation(s, precision="fp32_ir", tol=1e-14) as sim:
        sim.set_state(tp0, tpp0)
        sim.run(1, len(steps))
        tp, tpp = sim.get_state()
        st = sim.stats()
    assert relerr(tp, otp) <= 1e-11 and relerr(tpp, otpp) <= 1e-11
    assert st.outer_iters >= len(steps)


def test_fp32_ir_carried_residual_matches_fresh_start():
    """hf_run carries each step's initial residual from the previous step's fused pass; with
    the state set again before every step it is recomputed from the state by k_rhs_res.  Both
    match the oracle (log-normal kappa, Gaussian heating, 20^3, 12 steps) and each other."""
    s, _ = workloads.c3(20)
    n = s.params.n
    t0 = np.full(n, s.params.t_system)
    rc, otp, _, _ = oracle.OracleSystem(s).simulate_steps(list(range(1, 13)), t0, t0)
    assert rc == 0
    with Simulation(s, precision="fp32_ir", tol=1e-14) as a:
        a.initial_evolve()
        a.run(1, 12)
        ta, _ = a.get_state()
    with Simulation(s, precision="fp32_ir", tol=1e-14) as b:
        b.initial_evolve()
        for k in range(1, 13):
            b.simulate(k)
            tp, tpp = b.get_state()
            b.set_state(tp, tpp)
        tb, _ = b.get_state()
    assert relerr(ta, otp) <= 1e-12
    assert relerr(tb, otp) <= 1e-12
    assert relerr(ta, tb) <= 1e-12


def test_fp32_ir_fields_changed_between_runs():
    """The fused pass reads Qdens + B from a sum formed once per change of either.  A run, a
    new kappa and extra heating (new B and Qdens after re-assembly), then a second run from the
    state the first left: the second run matches the oracle built on the new fields."""
    s, _ = workloads.c3(18)
    n = s.params.n
    rng = np.random.default_rng(11)
    old = oracle.OracleSystem(s)   # keeps the first fields' arrays
    with Simulation(s, precision="fp32_ir", tol=1e-14) as sim:
        sim.initial_evolve()
        sim.run(1, 6)
        tp, tpp = sim.get_state()
        s.cells.kappa = s.cells.kappa * rng.uniform(0.5, 2.0, n)
        s.cells.qdens_extra = rng.uniform(0.0, 5.0e11, n)
        sim.set_fields(s.cells.kappa, s.cells.qdens_extra)
        sim.sparse_Hmatrix()
        sim.run(7, 6)
        ta, _ = sim.get_state()
    rc, otp, _, _ = oracle.OracleSystem(s).simulate_steps(list(range(7, 13)), tp, tpp)
    assert rc == 0
    assert relerr(ta, otp) <= 1e-12
    # the same steps on the first fields end far from these (a stale sum would show)
    rc, otp_old, _, _ = old.simulate_steps(list(range(7, 13)), tp, tpp)
    assert rc == 0
    assert relerr(ta, otp_old) > 1e-9


@pytest.mark.parametrize("heating", ["uniform", "gaussian"])
def test_fp32_ir_tiny_guess(heating):
    """The fused pass forms the next step's residual for the last TINY outcome seen (first
    guess: TINY).  Uniform heating of an insulated log-normal grid raises every cell each
    step, so TINY holds only at step 1 (x0 = T^n from a flat start): the first guess misses
    once and the next step starts from the state.  Gaussian heating with 300 K faces leaves
    cells that do not warm, so TINY holds every step and the guess never misses.  Both match
    the oracle at the reference tolerance."""
    s, _ = workloads.c3(16)
    p = s.params
    if heating == "uniform":
        p.kappa_bound = (0.0,) * 6
        s.cells.qdens_extra = np.full(p.n, 1.0e12)
    t0 = np.full(p.n, p.t_system)
    rc, otp, _, _ = oracle.OracleSystem(s).simulate_steps(list(range(1, 11)), t0, t0)
    assert rc == 0
    with Simulation(s, precision="fp32_ir", tol=1e-14) as sim:
        sim.initial_evolve()
        sim.run(1, 10)
        tp, _ = sim.get_state()
        st = sim.stats()
    assert st.tiny_misses == (1 if heating == "uniform" else 0)
    assert relerr(tp, otp) <= 1e-12


def test_fp32_ir_refinement_sweeps():
    """A loose inner tolerance (1e-2) forces several refinement sweeps per step, so the fused
    pass also runs from a refined iterate (not only from the state's initial guess); the
    result still meets the oracle at the reference tolerance.  The pass stores a step's
    residual only for sweeps the previous step followed with another; the first step (and any
    step needing more sweeps than the last) runs a pass again to store it."""
    s, _ = workloads.c3(24)
    n = s.params.n
    t0 = np.full(n, s.params.t_system)
    rc, otp, _, _ = oracle.OracleSystem(s).simulate_steps(list(range(1, 9)), t0, t0)
    assert rc == 0
    with Simulation(s, precision="fp32_ir", tol=1e-14, inner_tol=1e-2) as sim:
        sim.initial_evolve()
        sim.run(1, 8)
        tp, _ = sim.get_state()
        st = sim.stats()
    assert st.outer_iters >= 2 * 8
    assert 1 <= st.residual_reruns < st.outer_iters - 8
    assert relerr(tp, otp) <= 1e-12


def test_fp32_ir_alpha_reuse():
    """After a step whose inner solve ended on its first iteration (y = alpha r^0) in one sweep,
    the next carried step takes y = alpha r^0 with that alpha and no inner solve, and keeps it
    only if its FP64 residual meets half the inner solve's target.  Over 30 steps of the
    log-normal workload at 256^3 most steps reuse alpha; with reuse and without it (debug bit 2)
    the temperatures meet the 1e-10 gate against the FP64 path at the reference tolerance."""
    s, _ = workloads.c3(256)
    with Simulation(s, tol=1e-32) as ref:
        ref.initial_evolve()
        ref.run(1, 30)
        otp, _ = ref.get_state()
    res = {}
    for dbg in (0, 4):
        with Simulation(s, precision="fp32_ir", tol=1e-12) as sim:
            sim.set_debug(dbg)
            sim.initial_evolve()
            sim.run(1, 30)
            res[dbg] = (sim.get_state()[0], sim.stats())
    (ta, sa), (tb, sb) = res[0], res[4]
    assert sa.alpha_reuses >= 15, (sa.alpha_reuses, sa.reuse_misses)
    assert sb.alpha_reuses == 0 and sa.inner_iters < sb.inner_iters
    assert relerr(ta, otp) <= 1e-10 and relerr(tb, otp) <= 1e-10


def test_fp32_ir_alpha_reuse_miss():
    """A reused alpha whose sweep misses the target (40^3, where the later steps' one-iteration
    residuals sit near the inner target) is followed by ordinary sweeps in the same step and a
    pause of reuse; the step still meets the tolerance and the run the 1e-10 gate."""
    s, _ = workloads.c3(40)
    t0 = np.full(s.params.n, s.params.t_system)
    with Simulation(s, tol=1e-32) as ref:
        ref.initial_evolve()
        ref.run(1, 20)
        t64, _ = ref.get_state()
    with Simulation(s, precision="fp32_ir", tol=1e-12) as sim:
        sim.initial_evolve()
        errs = [sim.simulate(k)[1] for k in range(1, 21)]
        t32, _ = sim.get_state()
        st = sim.stats()
    assert st.reuse_misses >= 1 and st.alpha_reuses == st.reuse_misses, (st.alpha_reuses,)
    assert st.outer_iters > st.steps   # the missed step took a further sweep
    assert max(errs) <= 1e-12
    assert relerr(t32, t64) <= 1e-10


def _mixed_run(s, nsteps, debug, tol=1e-12):
    with Simulation(s, precision="fp32_ir", tol=tol) as sim:
        sim.set_debug(debug)
        sim.initial_evolve()
        sim.set_timing(True)
        sim.run(1, nsteps)
        kt = sim.kernel_times()
        st = sim.stats()
        stats = (st.steps, st.outer_iters, st.inner_iters, st.first_iter_stops, st.tiny_misses,
                 st.residual_reruns, st.alpha_reuses, st.reuse_misses)
        return sim.get_state(), stats, kt


@pytest.mark.parametrize("n", [24, 40])
def test_fp32_ir_pairs_bit_identical_with_frozen_alpha(n):
    """Two reusing steps in one pass (k_ir_step2_z) do the per-cell operations of two
    single-step passes in the same order and rounding; only the block sums are grouped by other
    tiles.  With the reused step length frozen (debug bit 8) and the same decisions, the two
    paths give identical bits and statistics: 40 steps of the log-normal workload at tol 1e-11,
    pairs on and off (debug bit 7); at least 15 pairs ran."""
    s, _ = workloads.c3(n)
    (ta, tpa), sa, ka = _mixed_run(s, 40, 256, tol=1e-11)
    (tb, tpb), sb, kb = _mixed_run(s, 40, 256 | 128, tol=1e-11)
    print(f"pairs {n}^3 frozen: launches {ka.get('k_ir_step2', (0, 0, 0))[1]}, stats {sa}",
          flush=True)
    assert ka.get("k_ir_step2", (0.0, 0, 0.0))[1] >= 15 and "k_ir_step2" not in kb
    assert sa == sb
    assert np.array_equal(bits(ta), bits(tb)) and np.array_equal(bits(tpa), bits(tpb))


def test_fp32_ir_pair_after_first_iteration_stop():
    """A carried step whose inner solve ends on its first iteration (y = alpha r^0, with that
    solve's alpha) runs as the first step of a two-step pass whose second step reuses the same
    alpha.  In the log-normal workload at 24^3 and tol 1e-11 the second step is the first whose
    solve stops there, and every later step reuses alpha: 39 steps run as one single step and
    19 pairs (the first of them starting from a solve), where pairs of reusing steps alone
    would leave steps 1, 2 and 39 single.  With alpha frozen (debug bit 8) the result and the
    statistics are those of the same 39 steps without pairs (debug bit 7), bit for bit."""
    s, _ = workloads.c3(24)
    (ta, tpa), sa, ka = _mixed_run(s, 39, 256, tol=1e-11)
    (tb, tpb), sb, kb = _mixed_run(s, 39, 256 | 128, tol=1e-11)
    n2, n1 = ka.get("k_ir_step2", (0.0, 0, 0.0))[1], ka.get("k_ir_step", (0.0, 0, 0.0))[1]
    print(f"solve then pair: pairs {n2}, single passes {n1}, stats {sa} / {sb}", flush=True)
    assert sa == sb and sa[3] >= 1
    assert (n2, n1) == (19, 1)
    assert np.array_equal(bits(ta), bits(tb)) and np.array_equal(bits(tpa), bits(tpb))


@pytest.mark.parametrize("case", ["gaussian", "uniform_insulated", "fourier"])
def test_fp32_ir_pairs_match_oracle(case):
    """Pairs with the step length updated after each pair (the default) against the oracle and
    the single-step path: 40 steps of the log-normal workload at 24^3 and tol 1e-11 with
    Gaussian heating, with uniform heating of an insulated grid, and Fourier (no tau term).  At
    this tolerance (pairs run only where one inner iteration meets it) 40 steps of either path
    are about 3e-10 from the oracle; the pairs add no error of their own."""
    s, _ = workloads.c3(24)
    p = s.params
    if case == "uniform_insulated":
        p.kappa_bound = (0.0,) * 6
        s.cells.qdens_extra = np.full(p.n, 1.0e12)
    if case == "fourier":
        p.icattaneo = 0
    t0 = np.full(p.n, p.t_system)
    rc, otp, otpp, _ = oracle.OracleSystem(s).simulate_steps(list(range(1, 41)), t0, t0)
    assert rc == 0
    (ta, tpa), sa, ka = _mixed_run(s, 40, 0, tol=1e-11)
    (tb, tpb), sb, kb = _mixed_run(s, 40, 128, tol=1e-11)
    print(f"pairs {case}: launches {ka.get('k_ir_step2', (0, 0, 0))[1]}, stats {sa} / {sb}, "
          f"err {relerr(ta, otp):.3g} / {relerr(tb, otp):.3g}", flush=True)
    assert ka.get("k_ir_step2", (0.0, 0, 0.0))[1] >= 1
    ea, eb = relerr(ta, otp), relerr(tb, otp)
    assert ea <= 1e-9 and relerr(tpa, otpp) <= 1e-9
    assert ea <= 1.2 * eb + 1e-12
    assert relerr(ta, tb) <= 1e-10


def _lognormal_box(nx, ny, nz, seed=3, sigma=4.0, dt=5.0e-11):
    """The log-normal workload on an nx x ny x nz grid: kappa = exp(N(ln 50, 1)) per cell, 0.5 um
    cells, Dirichlet 300 K on six faces, a Gaussian heating of sigma cells at the centre."""
    p = Params(nx=nx, ny=ny, nz=nz, time_step=dt, icattaneo=1, isteady=0,
               kappa_bound=(50.0,) * 6, t_bath=(300.0,) * 6, t_system=300.0, ntime=40)
    kappa = np.exp(np.random.default_rng(seed).normal(np.log(50.0), 1.0, p.n))
    g = [np.exp(-((np.arange(m) - (m - 1) / 2.0) ** 2) / (2.0 * sigma * sigma))
         for m in (nx, ny, nz)]
    q = 1.0e12 * g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
    return System.uniform(p, (nx * 0.5e-6, ny * 0.5e-6, nz * 0.5e-6), kappa=kappa,
                          rho=workloads.RHO_A, heat_capacity=workloads.CP_A, tau=1.0e-12,
                          qdens_extra=q.reshape(-1))


def test_fp32_ir_pairs_across_tiles():
    """The two-step pass on a grid of several of its 128 x 8 tiles a plane: 300 x 20 x 10 cells
    (3 x 3 tiles, the last ones 44 and 4 cells wide), so ring cells lie on tile edges inside the
    grid and on partial tiles at its far edges.  40 steps of the log-normal workload at tol
    1e-11 with pairs on and off (debug bit 7) end as close to the oracle as each other; with the
    step length frozen (debug bit 8) they give the same bits."""
    s = _lognormal_box(300, 20, 10)
    t0 = np.full(s.params.n, s.params.t_system)
    rc, otp, otpp, _ = oracle.OracleSystem(s).simulate_steps(list(range(1, 41)), t0, t0)
    assert rc == 0
    (ta, tpa), sa, ka = _mixed_run(s, 40, 0, tol=1e-11)
    (tb, tpb), sb, _ = _mixed_run(s, 40, 128, tol=1e-11)
    (tc, tpc), sc_, kc = _mixed_run(s, 40, 256, tol=1e-11)
    (td, tpd), sd, _ = _mixed_run(s, 40, 256 | 128, tol=1e-11)
    n2, n2c = ka.get("k_ir_step2", (0.0, 0, 0.0))[1], kc.get("k_ir_step2", (0.0, 0, 0.0))[1]
    ea, eb = relerr(ta, otp), relerr(tb, otp)
    print(f"pairs 300x20x10: launches {n2} / {n2c} (frozen), stats {sa} / {sb}, "
          f"err {ea:.3g} / {eb:.3g}", flush=True)
    assert n2 >= 10 and n2c >= 10
    assert ea <= 1e-9 and relerr(tpa, otpp) <= 1e-9
    assert ea <= 1.2 * eb + 1e-12 and relerr(ta, tb) <= 1e-10
    assert sc_ == sd
    assert np.array_equal(bits(tc), bits(td)) and np.array_equal(bits(tpc), bits(tpd))


# a shorter time step for the thin grids: on two planes every cell touches a Dirichlet face, and
# at 5e-11 s one inner iteration rarely meets the reuse test there (reuse then pauses and few
# pairs run); at 1e-11 s the step is mass-dominated enough for every shape to pair
DT_THIN = 1.0e-11


@pytest.mark.parametrize("shape", [(40, 30, 2), (20, 5, 30), (130, 9, 3)])
def test_fp32_ir_pairs_thin_grids(shape):
    """The two-step pass where its z-march and tiles meet the grid's edges: two planes only (the
    pipeline's fill and drain trips cover the whole grid), tiles with rows outside the grid
    (ny = 5 < 8), and one cell past a full tile in x and in y (130 x 9).  Pairs run, the steps
    meet the oracle, and with the step length frozen (debug bit 8) pairs give the bits of single
    steps (debug bit 7)."""
    s = _lognormal_box(*shape, sigma=2.0, dt=DT_THIN)
    t0 = np.full(s.params.n, s.params.t_system)
    rc, otp, _, _ = oracle.OracleSystem(s).simulate_steps(list(range(1, 41)), t0, t0)
    assert rc == 0
    (ta, _), sa, ka = _mixed_run(s, 40, 256, tol=1e-11)
    (tb, _), sb, _ = _mixed_run(s, 40, 256 | 128, tol=1e-11)
    n2 = ka.get("k_ir_step2", (0.0, 0, 0.0))[1]
    print(f"pairs {shape}: launches {n2}, stats {sa}, err {relerr(ta, otp):.3g}", flush=True)
    assert n2 >= 10
    assert relerr(ta, otp) <= 1e-9
    assert sa == sb and np.array_equal(bits(ta), bits(tb))


@pytest.mark.parametrize("icat", [1, 0])
def test_fp32_ir_pairs_per_cell_heat_capacity(icat):
    """The two-step pass's per-cell variants (rho c_p and tau' rho c_p read for every cell
    rather than once: rho varied by up to 0.1% per cell), Cattaneo and Fourier: 40 steps at
    24^3 and tol 1e-11 pair, meet the oracle, and with the step length frozen (debug bit 8)
    give the bits of single steps (debug bit 7)."""
    s, _ = workloads.c3(24)
    s.params.icattaneo = icat
    s.cells.rho = s.cells.rho * (1.0 + 1e-3 * np.random.default_rng(11).random(s.params.n))
    t0 = np.full(s.params.n, s.params.t_system)
    rc, otp, _, _ = oracle.OracleSystem(s).simulate_steps(list(range(1, 41)), t0, t0)
    assert rc == 0
    (ta, tpa), sa, ka = _mixed_run(s, 40, 256, tol=1e-11)
    (tb, tpb), sb, _ = _mixed_run(s, 40, 256 | 128, tol=1e-11)
    n2 = ka.get("k_ir_step2", (0.0, 0, 0.0))[1]
    print(f"pairs per-cell rho_c icat {icat}: launches {n2}, stats {sa}, "
          f"err {relerr(ta, otp):.3g}", flush=True)
    assert n2 >= 10
    assert relerr(ta, otp) <= 1e-9
    assert sa == sb
    assert np.array_equal(bits(ta), bits(tb)) and np.array_equal(bits(tpa), bits(tpb))


def test_fp32_ir_pairs_with_misses():
    """The 48^3 workload at tol 1e-12, where one-iteration residuals sit near the inner target:
    a pair whose step misses its test falls back to single steps and pauses reuse.  The 24
    steps with pairs on and off (debug bit 7) end as far from FP64 at the reference tolerance
    as each other, within 2e-10 of it (1.17e-10 and 1.11e-10 measured)."""
    s, _ = workloads.c3(48)
    with Simulation(s, tol=1e-32) as ref:
        ref.initial_evolve()
        ref.run(1, 24)
        t64, _ = ref.get_state()
    (ta, _), sa, ka = _mixed_run(s, 24, 0)
    (tb, _), sb, _ = _mixed_run(s, 24, 128)
    ea, eb = relerr(ta, t64), relerr(tb, t64)
    print(f"pairs 48^3: launches {ka.get('k_ir_step2', (0, 0, 0))[1]}, stats {sa} / {sb}, "
          f"err {ea:.3g} / {eb:.3g}", flush=True)
    assert sa[0] == 24 and sb[0] == 24
    assert ka.get("k_ir_step2", (0.0, 0, 0.0))[1] >= 1
    assert ea <= 2e-10 and ea <= 1.2 * eb + 1e-12


def test_fp32_ir_pair_fallbacks():
    """Both ways out of a two-step pass whose first step misses its test: a pair started from a
    first-iteration solve leaves the step to its single pass, and a pair of reusing steps
    reruns step n alone from the state and pauses reuse.  Debug bit 9 fails that test in every
    pair, so that each pair falls back: 40 steps of the log-normal workload at 24^3 and tol
    1e-11 take both fallbacks (pair launches beyond the reuse misses are solve-started pairs)
    and end as close to the oracle as the same run without pairs (debug bit 7)."""
    s, _ = workloads.c3(24)
    p = s.params
    t0 = np.full(p.n, p.t_system)
    rc, otp, otpp, _ = oracle.OracleSystem(s).simulate_steps(list(range(1, 41)), t0, t0)
    assert rc == 0
    (ta, tpa), sa, ka = _mixed_run(s, 40, 512, tol=1e-11)
    (tb, tpb), sb, kb = _mixed_run(s, 40, 128, tol=1e-11)
    n2 = ka.get("k_ir_step2", (0.0, 0, 0.0))[1]
    ea, eb = relerr(ta, otp), relerr(tb, otp)
    print(f"pair fallbacks: pair launches {n2}, stats {sa} / {sb}, err {ea:.3g} / {eb:.3g}",
          flush=True)
    assert sa[0] == 40 and sb[0] == 40
    assert sa[7] >= 1 and sa[6] == sa[7]   # every reusing pair missed and was redone
    assert n2 > sa[7]                      # and some pairs started from a solve
    assert ea <= 1e-9 and relerr(tpa, otpp) <= 1e-9
    assert ea <= 1.2 * eb + 1e-12
    assert relerr(ta, tb) <= 1e-10


@pytest.mark.parametrize("prec", ["fp64", "fp32_ir"])
def test_uniform_heat_capacity_reads_are_exact(prec):
    """Where rho c_p and tau' rho c_p are the same in every cell (found at upload), the step
    kernels read one value for all cells; debug bit 3 keeps the per-cell reads.  The values
    are the same, so 12 steps of the log-normal workload (uniform rho, c_p, tau) give the same
    bits either way."""
    s, _ = workloads.c3(24)
    out = []
    for dbg in (0, 8):
        with Simulation(s, precision=prec, tol=1e-14) as sim:
            sim.set_debug(dbg)
            sim.initial_evolve()
            sim.run(1, 12)
            out.append(sim.get_state())
    assert np.array_equal(bits(out[0][0]), bits(out[1][0]))
    assert np.array_equal(bits(out[0][1]), bits(out[1][1]))


@pytest.mark.parametrize("prec", ["fp64", "fp32_ir"])
def test_first_iteration_stop_is_exact(prec):
    """A CG solve whose first iteration already meets the tolerance stops there: the next
    residual norm is predicted as alpha^2 q.q - r.r and y = alpha p is left to its consumer,
    which forms it with k_cg_b's own product.  With that stop turned off (debug bit 1) the
    same iteration runs in full, so the temperatures must agree bit for bit; over 20 steps
    at least 5 solves take the stop (inner tolerance 0.5 for the mixed inner solves, so that
    nearly every inner solve stops after one iteration and the refinement does the rest).
    Debug bit 2 keeps every mixed step solving (no reuse of the last step length)."""
    s, _ = workloads.c3(24)
    out = {}
    for dbg in (4, 6):
        with Simulation(s, precision=prec, tol=1e-12,
                        inner_tol=0.5 if prec == "fp32_ir" else 1e-6) as sim:
            sim.set_debug(dbg)
            sim.initial_evolve()
            its = [sim.simulate(k)[0] for k in range(1, 21)]
            out[dbg] = (sim.get_state()[0], its, sim.stats().first_iter_stops)
    t0, its0, n0 = out[4]
    t2, its2, n2 = out[6]
    assert its0 == its2
    assert n0 >= 5 and n2 == 0, (n0, n2, its0)
    assert np.array_equal(bits(t0), bits(t2))


@pytest.mark.parametrize("prec", ["fp32_ir", "fp64"])
def test_pair_kernels_match_scalar(monkeypatch, prec):
    """The CG kernels take several x-neighbouring cells per thread: four in FP32 when nx is a
    multiple of 4 (HF_NO_QUADS: two), two in FP64 and in FP32 when nx is even (HF_NO_PAIRS:
    one).  The sums per cell are the same; only the order of the per-block partial sums
    differs, so all runs meet the tolerance and agree to it (log-normal kappa, Gaussian
    heating, 24^3 and a periodic 8 x 6 x 5 grid)."""
    variants = [{}, {"HF_NO_QUADS": "1"}, {"HF_NO_PAIRS": "1"}]
    if prec == "fp64":
        variants = variants[1:]   # no FP64 quads (and debug bit 6: not the persistent path)
    for s in (workloads.c3(24)[0], random_system(31, 8, 6, 5, periodic=(True, False, True))):
        out = []
        for env in variants:
            for k in ("HF_NO_QUADS", "HF_NO_PAIRS"):
                if k in env:
                    monkeypatch.setenv(k, env[k])
                else:
                    monkeypatch.delenv(k, raising=False)
            with Simulation(s, precision=prec, tol=1e-14) as sim:
                sim.set_debug(64)
                sim.initial_evolve()
                sim.run(1, 10)
                out.append(sim.get_state()[0])
        for o in out[1:]:
            assert relerr(out[0], o) <= 1e-12


@pytest.mark.parametrize("prec", ["fp32_ir", "fp64"])
def test_timed_steps_do_not_depend_on_earlier_calls(prec):
    """hf_set_state starts the step heuristics afresh (carried residual, TINY guess, sweep and
    step-length-reuse history, CG batch sizes): 20 steps from the same state after 0, 2 or 5
    earlier steps give identical statistics and bit-identical temperatures (the 512^3
    mixed-precision workload's physics at 32^3, in the step-length-reuse regime)."""
    s, _ = workloads.c3(32)
    out = []
    for warm in (0, 2, 5):
        with Simulation(s, precision=prec, tol=1e-12) as sim:
            sim.initial_evolve()
            if warm:
                sim.run(1, warm)
            sim.initial_evolve()
            sim.reset_stats()
            sim.run(1, 20)
            st = sim.stats()
            stats = (st.steps, st.outer_iters, st.inner_iters, st.first_iter_stops,
                     st.tiny_misses, st.residual_reruns, st.alpha_reuses, st.reuse_misses)
            out.append((stats, sim.get_state()[0]))
    for stats, tp in out[1:]:
        assert stats == out[0][0]
        assert np.array_equal(tp.view(np.int64), out[0][1].view(np.int64))
    if prec == "fp32_ir":
        assert out[0][0][6] > 0   # the run reused the step length

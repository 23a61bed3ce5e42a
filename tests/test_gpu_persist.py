"""The persistent small-grid path: FP64 steps of grids up to 8192 cells run in one workgroup,
up to 256 steps per launch, with the CG decisions taken on the device (hf_persist.hip).
Checked against the oracle, against the multi-kernel path (debug bit 6) and against itself
one step per call."""
import numpy as np
import pytest

import oracle
from heatflow_amd import HeatFlowError, Simulation, workloads
from heatflow_amd.system import Params, System

from _cases import GEOMETRIES, random_state, random_system

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def test_run_equals_single_steps_across_launches():
    """300 steps of the 16^3 Cattaneo workload (one centre heater) in one hf_run call (two
    launches: 256 + 44 steps) and in 300 hf_step calls: the same bits, the same iteration
    counts; within the 1e-10 gate of the oracle at the reference tolerance (at tol 1e-14: the
    2-norm stop test at 1e-12 lets the error near the heater grow to 1.3e-10 over 300 steps)."""
    s, _ = workloads.c1(16)
    out = []
    for single in (False, True):
        with Simulation(s, tol=1e-14) as sim:
            sim.initial_evolve()
            sim.reset_stats()
            if single:
                its = sum(sim.simulate(k)[0] for k in range(1, 301))
            else:
                its = sim.run(1, 300)
            st = sim.stats()
            out.append((sim.get_state(), its, st.steps, st.outer_iters))
    (ta, tpa), ia, sa, oa = out[0]
    (tb, tpb), ib, sb, ob = out[1]
    assert np.array_equal(bits(ta), bits(tb)) and np.array_equal(bits(tpa), bits(tpb))
    assert (ia, sa, oa) == (ib, sb, ob) and sa == 300
    rc, otp, _, _ = oracle.OracleSystem(s).run(300)
    assert rc == 0
    assert relerr(ta, otp) <= 1e-10
    assert ta.max() > 300.0 + 1e-3


@pytest.mark.parametrize("gi", range(len(GEOMETRIES)))
def test_persistent_matches_multi_kernel_path(gi):
    """Eight steps on every geometry: the persistent path against the multi-kernel path (debug
    bit 6) at tol 1e-14 -- the same solver, with dot products grouped differently -- to 1e-12,
    with the same total of CG iterations to within one per step."""
    s = random_system(900 + gi, **GEOMETRIES[gi])
    tp0, tpp0 = random_state(gi + 40, s.params.n)
    res = []
    for dbg in (0, 64):
        with Simulation(s, tol=1e-14) as sim:
            sim.set_debug(dbg)
            sim.set_state(tp0, tpp0)
            its = sim.run(1, 8)
            res.append((sim.get_state()[0], its))
    assert relerr(res[0][0], res[1][0]) <= 1e-12
    assert abs(res[0][1] - res[1][1]) <= 8


def test_path_switch_keeps_the_state():
    """Five persistent steps, then five multi-kernel ones (debug bit 6 set between runs): the
    second run starts from the first one's T^n, T^{n-1} and TINY outcome; the ten steps match
    the oracle's ten (a 12^3 grid with Gaussian heating and log-normal kappa)."""
    s, _ = workloads.c3(12)
    with Simulation(s) as sim:
        sim.initial_evolve()
        sim.run(1, 5)
        sim.set_debug(64)
        sim.run(6, 5)
        tp, tpp = sim.get_state()
    rc, otp, otpp, _ = oracle.OracleSystem(s).run(10)
    assert rc == 0
    assert relerr(tp, otp) <= 1e-12 and relerr(tpp, otpp) <= 1e-12


def test_time_dependent_heating_and_nan_inside_a_run():
    """Heater modes 2-6 recompute Qdens every step inside the persistent run (per-step powers
    from the host): 12 steps against the oracle.  A NaN frequency makes the AC heater's power
    NaN: the run stops at step 1 with the reference's Q guard (code 2) and leaves the state."""
    s = random_system(77, **GEOMETRIES[4])
    tp0, tpp0 = random_state(5, s.params.n)
    rc, otp, _, _ = oracle.OracleSystem(s).simulate_steps(list(range(1, 13)), tp0, tpp0)
    assert rc == 0
    with Simulation(s) as sim:
        sim.set_state(tp0, tpp0)
        sim.run(1, 12)
        tp, _ = sim.get_state()
    assert relerr(tp, otp) <= 1e-12
    p = Params(nx=4, ny=3, nz=3, time_step=1e-10, icattaneo=1, kappa_bound=(150.0,) * 6,
               t_bath=(300.0,) * 6, power_in=1e-4, freq=float("nan"))
    ih = np.zeros(36, np.int64)
    ih[5] = 3
    sn = System.uniform(p, (4e-6, 3e-6, 3e-6), kappa=150.0, rho=2330.0, heat_capacity=700.0,
                        tau=1e-12, iheater=ih)
    with Simulation(sn) as sim:
        sim.initial_evolve(300.0)
        with pytest.raises(HeatFlowError) as ei:
            sim.run(1, 3)
        assert ei.value.code == 2
        assert sim.stats().steps == 0
        assert np.all(sim.get_state()[0] == 300.0)


def test_kernel_timing_reports_the_run():
    """With timing on, a persistent run is one timed launch per 256 steps (k_run_small)."""
    s, _ = workloads.c1(16)
    with Simulation(s, tol=1e-12) as sim:
        sim.initial_evolve()
        sim.set_timing(True)
        sim.run(1, 20)
        kt = sim.kernel_times()
    assert set(kt) == {"k_run_small"}
    ms, launches, by = kt["k_run_small"]
    assert launches == 1 and ms > 0 and by > 0

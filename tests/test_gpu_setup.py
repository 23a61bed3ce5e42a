"""Setup on the device: the per-cell fields set_global_variables derives from material ids
(src/heatflow/mod_setup.f90:34-68 over material(), src/heatflow/mod_material.f90:54-265) and
the stability check (mod_setup.f90:158-199), against the oracle's restatement of both (itself
pinned to the reference's own material() in test_oracle.py)."""
import numpy as np
import pytest

import oracle
from heatflow_amd import HeatFlowError, Simulation, _lib
from heatflow_amd.system import Cells, Geometry, Params, System

pytestmark = pytest.mark.gpu

BUILTIN_IDS = [100, 140, 3001, 3002, 3003, 3004, 9001, 9002, 9003, 310, 320, 330, 340, 350]
FILE_MATS = {1: dict(kappa=0.541, rho=997.0, heat_capacity=4200.0, tau=1e-12, em=0.0),
             7: dict(kappa=2.5, rho=1000.0, heat_capacity=900.0, tau=3e-9, em=0.7),
             42: dict(kappa=0.1, rho=10.0, heat_capacity=5.0, tau=0.0, em=0.25)}


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


def geometry(nx=64, ny=48, nz=20, seed=5):
    """Random ids over every built-in and three file materials, with runs that make the tau/em
    carry cross chunk boundaries: cells 0..99 set neither (they keep 0), 4000..13099 are
    silicon (sets neither, spans two 4096-cell chunk edges), and 13100 sets both."""
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    ids = rng.choice(BUILTIN_IDS + list(FILE_MATS), size=N)
    ids[:100] = rng.choice([140, 3001, 310], size=100)
    if N > 13100:
        ids[4000:13100] = 140
        ids[13100] = 7
    heat = np.zeros(N, np.int64)
    heat[rng.choice(N, 6, replace=False)] = 1
    return Geometry(nx, ny, nz, (0.64, 0.48, 0.4), ids, heat)


def params(g, dt=1e-3, **kw):
    return Params(nx=g.nx, ny=g.ny, nz=g.nz, time_step=dt, icattaneo=1,
                  kappa_bound=(2.0,) * 6, t_bath=(300.0, 310.0, 305.0, 300.0, 302.0, 300.0),
                  power_in=50.0, heated_steps=100, t_bath_rad=290.0, t_system=300.0, **kw)


def oracle_system(p, g):
    k, r, h, t, e = oracle.materials(g.imaterial_type, FILE_MATS)
    n = g.nx * g.ny * g.nz
    lx, ly, lz = g.cell_lengths
    return System(p, Cells(kappa=k, rho=r, heat_capacity=h, tau=t, em=e, len_x=np.full(n, lx),
                           len_y=np.full(n, ly), len_z=np.full(n, lz),
                           iheater=g.iheater.copy()))


def test_upload_materials_fields_and_steps():
    """hf_upload_materials over 61440 cells against the host upload of the oracle's material
    loop: kappa, rho c_p, tau/dt^2, the radiative coefficient and the volume bit-exact (the
    carry of tau and em across chunks included), the stability verdict on those fields, then
    6 steps bit-exact between the two contexts and within 1e-10 of the oracle's run."""
    g = geometry()
    p = params(g)
    ref = oracle_system(p, g)
    assert ref.cells.tau[:100].max() == 0.0 and ref.cells.em[:100].max() == 0.0
    assert ref.cells.tau[13099] != ref.cells.tau[13100]
    ocell, oval, _ = oracle.OracleSystem(ref).stability()
    assert ocell >= 0          # air (100: kappa = rho c_p = 1) is far from stable at dt 1e-3
    with Simulation.from_geometry(p, g, FILE_MATS) as a, Simulation(ref) as b:
        for name in ("kappa", "rhoc", "tau_s", "ells", "volume", "diag", "cx", "cy", "cz", "B"):
            assert np.array_equal(bits(a.vector(name)), bits(b.vector(name))), name
        assert a.stability() == (ocell, oval)
        for sim in (a, b):
            sim.initial_evolve()
            sim.run(1, 6)
        ta, tpa = a.get_state()
        tb, tpb = b.get_state()
    assert np.array_equal(bits(ta), bits(tb)) and np.array_equal(bits(tpa), bits(tpb))
    rc, otp, _, _ = oracle.OracleSystem(ref).run(6)
    assert rc == 0
    assert np.max(np.abs(ta - otp) / np.abs(otp)) <= 1e-10


def test_upload_materials_small_grid_and_device_ids():
    """A grid smaller than one block (3 x 2 x 2), ids given as a device tensor."""
    import torch
    g = Geometry(3, 2, 2, (0.3, 0.2, 0.2), np.array([3001, 320, 140, 42, 330, 1, 350, 100,
                                                      9001, 140, 7, 310], np.int64),
                 np.zeros(12, np.int64))
    p = params(g)
    ref = oracle_system(p, g)
    with Simulation(System(p, None)) as a, Simulation(ref) as b:
        dev_ids = torch.as_tensor(g.imaterial_type, device="cuda:0")
        fm = (_lib.HfMaterial * 3)()
        for k, (idx, m) in enumerate(sorted(FILE_MATS.items())):
            fm[k].index = idx
            for key in ("kappa", "rho", "heat_capacity", "tau", "em"):
                setattr(fm[k], key, m[key])
        import ctypes as C
        lx, ly, lz = g.cell_lengths
        torch.cuda.synchronize()
        _lib.check(a.lib.hf_upload_materials(a.ctx, C.c_void_p(dev_ids.data_ptr()), None, 3, fm,
                                             lx, ly, lz), a.ctx)
        for name in ("kappa", "rhoc", "tau_s", "ells", "volume"):
            assert np.array_equal(bits(a.vector(name)), bits(b.vector(name))), name


@pytest.mark.parametrize("bad", [0, -3, 5, 141])
def test_upload_materials_unknown_id(bad):
    """An id material() does not know (non-positive, a file index not in mat.in, an unknown
    built-in) is HF_ERR_ARG naming the first such cell, where the reference prints and exits."""
    g = geometry(16, 8, 4)
    g.imaterial_type[300] = bad
    g.imaterial_type[400] = bad
    with pytest.raises(ValueError):
        oracle.materials(g.imaterial_type, FILE_MATS)
    with pytest.raises(HeatFlowError) as e:
        Simulation.from_geometry(params(g), g, FILE_MATS)
    assert e.value.code == _lib.HF_ERR_ARG
    assert f"{bad} (cell 300)" in str(e.value)


def _stab_system(dt, kappa=None, kb=0.0):
    nx, ny, nz = 40, 30, 12
    n = nx * ny * nz
    k = np.ones(n) if kappa is None else kappa
    return System(Params(nx=nx, ny=ny, nz=nz, time_step=dt, kappa_bound=(kb,) * 6),
                  Cells(kappa=k, rho=np.full(n, 2.0), heat_capacity=np.full(n, 0.5),
                        tau=np.zeros(n), em=np.zeros(n), len_x=np.full(n, 1.0),
                        len_y=np.full(n, 0.5), len_z=np.full(n, 2.0),
                        iheater=np.zeros(n, np.int64)))


def test_check_stability_matches_oracle():
    """hf_check_stability against the oracle's stability over 14400 cells: a var between the
    double 1/12 and the reference's single-precision 1/12 passes; the first of two unstable
    interior cells is reported with its value; kappaBoundx1 fails on the low faces only (an
    interior cell with a large kappa behind a failing boundary is not the one reported)."""
    s = 5.25
    thr32 = float(np.float32(1.0) / np.float32(12.0))
    cases = {"between": _stab_system((1.0 / 12.0 + thr32) / 2 / s)}
    n = 40 * 30 * 12
    kap = np.ones(n)
    first = 7 + 40 * (9 + 30 * 5)
    kap[first] = 30.0
    kap[first + 4000] = 50.0
    cases["interior"] = _stab_system(0.05 / s, kap)
    cases["boundary"] = _stab_system(0.05 / s, kap, kb=20.0)
    cases["stable"] = _stab_system(1e-4, kap)
    for name, sysm in cases.items():
        ocell, oval, _ = oracle.OracleSystem(sysm).stability()
        with Simulation(sysm, assemble=False) as sim:
            cell, val = sim.stability()
            if ocell < 0:
                sim.check_stability()
            else:
                with pytest.raises(HeatFlowError) as e:
                    sim.check_stability()
                assert e.value.code == _lib.HF_ERR_UNSTABLE
        print(f"stability {name}: cell {cell} value {val:.6g} (oracle {ocell}, {oval:.6g})",
              flush=True)
        assert cell == ocell
        if ocell >= 0:
            assert val == oval
    assert [oracle.OracleSystem(cases[c]).stability()[0] for c in ("between", "interior",
                                                                    "boundary")] == [-1, first, 0]

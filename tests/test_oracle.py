"""The oracle pinned against the reference: golden vectors, the manual's closed forms, the
reference build itself, and invariants the mathematics fixes.  CPU only."""
import numpy as np
import pytest

import oracle
from heatflow_amd.system import read_inputs

from _cases import GEOMETRIES, manual_cases, random_state, random_system, read_tempdis


def dense_from_coo(n, irow, jcol, val):
    H = np.zeros((n, n))
    ok = irow > 0
    np.add.at(H, (irow[ok] - 1, jcol[ok] - 1), val[ok])
    return H


# ---- golden vector: the reference's own TempDis.dat ----------------------------------------

def test_golden_tempdis(golden_dir):
    s = read_inputs(golden_dir)
    assert (s.params.nx, s.params.ny, s.params.nz, s.params.ntime) == (25, 25, 1, 10)
    o = oracle.OracleSystem(s)
    rc, tp, tpp, iters = o.run(s.params.ntime)
    gold = read_tempdis(f"{golden_dir}/TempDis.dat")
    assert rc == 0 and gold.shape == tp.shape
    # ES16.8 keeps 9 significant digits: the print floor is 5e-7 K at 280 K
    assert np.abs(tp - gold).max() <= 5.0e-7 + 1e-12
    # a wrong step count is far off (9 or 11 steps differ by > 1e-3 K)
    rc9, tp9, _, _ = o.run(9)
    assert np.abs(tp9 - gold).max() > 1e-4


# ---- the manual's chapter-6 cases (closed forms) ---------------------------------------------

@pytest.mark.parametrize("idx", range(4))
def test_manual_cases(idx):
    case, s = manual_cases()[idx]
    o = oracle.OracleSystem(s)
    n, L, cnt, val, irow, jcol = o.coo()
    H = dense_from_coo(n, irow, jcol, val)
    np.testing.assert_allclose(H, np.array(case["H"]), rtol=0, atol=1e-14)
    np.testing.assert_allclose(o.boundary(), case["B"], rtol=0, atol=1e-12)
    rc, tp, tpp, _ = o.run(1, t_system=200.0)
    assert rc == 0
    np.testing.assert_allclose(tp, case["T"], rtol=1e-12)


# ---- the oracle against the reference build ---------------------------------------------------

needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="reference build oracle/_ref/libhfref.so not present")


@needs_ref
@pytest.mark.parametrize("gi", range(len(GEOMETRIES)))
def test_oracle_matches_reference_build(gi):
    g = dict(GEOMETRIES[gi])
    s = random_system(100 + gi, **g)
    o = oracle.OracleSystem(s)
    r = oracle.RefSystem(s)
    n, L, cnt, val, irow, jcol = o.coo()
    rn, rL, rv, ri, rj = r.coo()
    assert L == rL
    assert np.array_equal(irow, ri) and np.array_equal(jcol, rj)
    assert np.array_equal(val.view(np.int64), rv.view(np.int64))      # bit-exact values
    assert np.array_equal(o.boundary(), r.boundary())
    tp, tpp = random_state(gi, n)
    r.set_state(tp, tpp)
    q, qd, hv = o.heater(2, tp)
    rq, rqd, rhv = r.heater(2)
    assert np.array_equal(q, rq) and np.array_equal(qd, rqd) and hv == rhv
    assert np.array_equal(o.scat(tp, tpp), r.scat())
    if s.cells.qdens_extra is None:
        rc, otp, otpp, _ = o.simulate_steps([1, 2, 3], tp, tpp)
        r.set_state(tp, tpp)
        for it in (1, 2, 3):
            r.simulate(it)
        rtp, rtpp = r.get_state()
        assert rc == 0
        assert np.array_equal(otp, rtp) and np.array_equal(otpp, rtpp)


@pytest.mark.skipif(not oracle.ref_omp_available(),
                    reason="OpenMP reference build oracle/_ref/libhfref_omp.so not present")
def test_reference_openmp_build_matches_serial():
    """The OpenMP build of the reference (the CPU baseline's parallel leg) solves the same
    steps: its atomic SpMV only reorders sums, so temperatures agree to round-off."""
    s = random_system(11, 12, 10, 9, nheat=4)
    tp0, tpp0 = random_state(3, s.params.n)
    out = []
    for omp in (False, True):
        r = oracle.RefSystem(s, omp=omp)
        r.set_state(tp0, tpp0)
        for k in (1, 2, 3):
            r.simulate(k)
        out.append(r.get_state()[0])
    np.testing.assert_allclose(out[1], out[0], rtol=1e-12)


@needs_ref
def test_oracle_matches_reference_qextra():
    s = random_system(7, 5, 4, 3, qextra=True)
    o = oracle.OracleSystem(s)
    r = oracle.RefSystem(s)
    tp, tpp = random_state(7, s.params.n)
    rc, otp, otpp, _ = o.simulate_steps([1, 2], tp, tpp)
    r.set_state(tp, tpp)
    for it in (1, 2):
        r.simulate_qextra(it, s.cells.qdens_extra)
    rtp, _ = r.get_state()
    assert rc == 0 and np.array_equal(otp, rtp)


# ---- invariants ----------------------------------------------------------------------------

def _gb(kb, k, L, br):
    v = (2.0 * kb * k / (kb + k)) / (L * L)
    return np.where(k != kb, v * br, v)


@pytest.mark.parametrize("gi", [0, 2, 7])
def test_h_symmetric_and_row_sums(gi):
    """H is symmetric; each row sums to -alpha - (bath conductances of its boundary faces)."""
    s = random_system(200 + gi, **GEOMETRIES[gi])
    p, c = s.params, s.cells
    o = oracle.OracleSystem(s)
    n, L, cnt, val, irow, jcol = o.coo()
    assert cnt <= L and np.all(irow[cnt:] == -2) and np.all(jcol[cnt:] == -1)
    H = dense_from_coo(n, irow, jcol, val)
    assert np.array_equal(H, H.T)
    tau = c.tau_scaled(p.time_step) if p.icattaneo else 0.0
    rc = c.rho * c.heat_capacity
    alpha = tau * rc + rc / p.time_step
    x = np.arange(n) % p.nx
    y = (np.arange(n) // p.nx) % p.ny
    z = np.arange(n) // (p.nx * p.ny)
    gsum = np.zeros(n)
    for ax, (pos, m, Lv) in enumerate(((x, p.nx, c.len_x), (y, p.ny, c.len_y), (z, p.nz, c.len_z))):
        for side in (0, 1):
            at = pos == (0 if side == 0 else m - 1)
            gsum += np.where(at, _gb(p.kappa_bound[2 * ax + side], c.kappa, Lv, p.br), 0.0)
    # dropped (|H| < TINY) couplings still enter the diagonal: add them back
    dropped = np.zeros(n)
    for i in range(n):
        for j in range(n):
            if i != j and H[i, j] == 0.0:
                hij = o.hmatrixfunc(i + 1, j + 1)
                if hij != 0.0:
                    dropped[i] += hij
    rows = H.sum(axis=1) + dropped
    np.testing.assert_allclose(rows, -alpha - gsum, rtol=1e-12, atol=1e-9 * np.abs(alpha).max())


def test_coo_against_brute_force_hmatrixfunc():
    """Every non-zero hmatrixfunc(i, j), i >= j, over all pairs of a 3x3x3 periodic-x grid is in
    the COO with the same bits and nothing else is; the upper triangle is its mirror
    (mod_setup.f90:138-145), equal to hmatrixfunc(j, i) up to the rounding of the reversed
    harmonic-mean operand order."""
    s = random_system(5, 3, 3, 3, periodic=(True, False, False))
    o = oracle.OracleSystem(s)
    n, L, cnt, val, irow, jcol = o.coo()
    H = dense_from_coo(n, irow, jcol, val)
    assert np.array_equal(H, H.T)
    for i in range(1, n + 1):
        for j in range(1, i + 1):
            h = o.hmatrixfunc(i, j)
            if abs(h) >= 1e-12:
                assert H[i - 1, j - 1] == h
            else:
                assert H[i - 1, j - 1] == 0.0
            hu = o.hmatrixfunc(j, i)
            assert abs(hu - H[j - 1, i - 1]) <= 4 * np.spacing(abs(hu) + 1e-300)


def test_uniform_fixed_point():
    """Uniform material at the bath temperature, no heating: T stays at the bath value."""
    from heatflow_amd.system import Params, System
    p = Params(nx=6, ny=5, nz=4, time_step=1e-10, icattaneo=1, kappa_bound=(150.0,) * 6,
               t_bath=(300.0,) * 6, t_system=300.0)
    s = System.uniform(p, (6e-6, 5e-6, 4e-6), kappa=150.0, rho=2330.0, heat_capacity=700.0,
                       tau=1e-12)
    rc, tp, tpp, _ = oracle.OracleSystem(s).run(5)
    assert rc == 0
    np.testing.assert_allclose(tp, 300.0, rtol=1e-13)


@pytest.mark.parametrize("icat", [0, 1])
def test_energy_balance_insulated(icat):
    """All faces insulating: summing H T^{n+1} = S^n over cells, the couplings cancel and
    sum[a (T^{n+1}-T^n) + c (T^{n+1} - 2T^n + T^{n-1})] = sum Qdens exactly (to solver
    tolerance), a = rho c_p / dt, c = tau rho c_p / dt^2."""
    s = random_system(11, 5, 4, 3, icattaneo=icat, nheat=2)
    p = s.params
    p.kappa_bound = (0.0,) * 6
    o = oracle.OracleSystem(s)
    tp0, tpp0 = random_state(3, p.n)
    rc, tp1, _, _ = o.simulate_steps([1], tp0, tpp0)
    assert rc == 0
    c = s.cells
    rc_ = c.rho * c.heat_capacity
    a = rc_ / p.time_step
    cc = c.tau_scaled(p.time_step) * rc_ if icat else 0.0 * rc_
    q, qd, hv = o.heater(1, tp0)
    lhs = np.sum(a * (tp1 - tp0) + cc * (tp1 - 2 * tp0 + tpp0))
    assert abs(lhs - qd.sum()) <= 1e-9 * max(abs(qd.sum()), np.abs(a * tp0).sum() * 1e-6)


def test_linbcg_matches_dense_solve():
    s = random_system(21, 5, 4, 4)
    o = oracle.OracleSystem(s)
    n, L, cnt, val, irow, jcol = o.coo()
    H = dense_from_coo(n, irow, jcol, val)
    b = np.random.default_rng(0).uniform(-1e20, 1e20, n)
    x = np.zeros(n)
    import ctypes as C
    it, err = C.c_int64(), C.c_double()
    oracle.olib().hfo_linbcg(C.c_int64(n), C.c_int64(L), oracle._p(val), oracle._p(irow),
                             oracle._p(jcol), oracle._p(b), oracle._p(x), C.c_int64(1),
                             C.c_double(1e-32), C.c_int64(50000), C.byref(it), C.byref(err))
    xd = np.linalg.solve(H, b)
    np.testing.assert_allclose(x, xd, rtol=1e-12, atol=1e-12 * np.abs(xd).max())


# ---- the dense torch restatement used to check the adjoint ----------------------------------

@pytest.mark.parametrize("bi", range(5))
@pytest.mark.parametrize("icat", [0, 1])
def test_dense_reference_matches_oracle(icat, bi):
    """tests/dense_ref.py (the autograd reference of the adjoint tests) steps the same physics
    as the oracle: per-cell power P enters as Qdens = P / V; periodic axes and CG faces as the
    oracle has them."""
    import torch
    from dense_ref import dense_run, with_boundary
    from heatflow_amd import workloads
    s, m = workloads.c4(6, icattaneo=icat, nheat=3, nprobe=5)
    s = with_boundary(s, bi)
    rc, tp, _, _ = oracle.OracleSystem(s).run(7)
    assert rc == 0
    k = torch.as_tensor(s.cells.kappa)
    pw = torch.as_tensor(m["power"])
    td = dense_run(s, k, pw, 7).numpy()
    np.testing.assert_allclose(td, tp, rtol=1e-12)


@pytest.mark.parametrize("icat", [0, 1])
def test_dense_reference_radiative_matches_oracle(icat):
    """The dense reference's heaters and radiative source (Q = power_in - em L1 L2 sigma
    (T^4 - T_Bath^4), Qdens = Q / V_heated, mod_heating.f90:121-123, 164-167) step the same
    temperatures as the oracle on a system where radiation cools it by ~9 K in 10 steps."""
    import torch
    from dense_ref import dense_run, radiative_system
    s, _ = radiative_system(6, icattaneo=icat)
    rc, tp, _, _ = oracle.OracleSystem(s).run(10)
    assert rc == 0
    td = dense_run(s, torch.as_tensor(s.cells.kappa), torch.zeros(s.params.n,
                   dtype=torch.float64), 10).numpy()
    assert tp.min() < 295.0
    np.testing.assert_allclose(td, tp, rtol=1e-12)


def test_coo_hash_streamed_matches_stored():
    """The oracle's streamed COO hash (nothing stored; used where the COO does not fit in
    memory) equals the numpy hash of its stored COO, entry hash by entry hash and in total,
    on periodic, CG-face and TINY-dropping geometries; swapping two entries or flipping one
    value bit changes it."""
    from _cases import GEOMETRIES, coo_hash, random_system
    for gi in (0, 1, 3, 7):
        s = random_system(300 + gi, **GEOMETRIES[gi])
        o = oracle.OracleSystem(s)
        n, L, cnt, val, irow, jcol = o.coo()
        L2, cnt2, h = o.coo_hash()
        assert (L2, cnt2) == (L, cnt)
        assert coo_hash(val, irow, jcol) == h
        for k in (0, cnt // 2, L - 1):
            e = oracle.olib().hfo_coo_entry_hash(k, int(irow[k]), int(jcol[k]), float(val[k]))
            assert coo_hash(val[k:k + 1], irow[k:k + 1], jcol[k:k + 1], k0=k) == e
        v2 = val.copy()
        v2[[1, 2]] = v2[[2, 1]]
        i2 = irow.copy()
        i2[[1, 2]] = i2[[2, 1]]
        j2 = jcol.copy()
        j2[[1, 2]] = j2[[2, 1]]
        assert coo_hash(v2, i2, j2) != h
        v3 = val.copy()
        v3.view(np.uint64)[5] ^= np.uint64(1)
        assert coo_hash(v3, irow, jcol) != h


# ---- setup: material() over the grid and the stability check ------------------------------

BUILTIN_IDS = [100, 140, 3001, 3002, 3003, 3004, 9001, 9002, 9003, 310, 320, 330, 340, 350]
FILE_MATS = {1: dict(kappa=0.541, rho=997.0, heat_capacity=4200.0, tau=1e-12, em=0.0),
             7: dict(kappa=2.5, rho=1000.0, heat_capacity=900.0, tau=3e-9, em=0.7),
             42: dict(kappa=0.1, rho=10.0, heat_capacity=5.0, tau=0.0, em=0.25)}


def test_materials_known_answers():
    """The reference's own material test (test/test_mod_material.f90:22-35): silicon (140) has
    kappa 130, heat_capacity 4200, rho 2328 to 1e-6; the built-ins are the single-precision
    roundings of their literals (0.19 is not 0.19 in double); tau carries from the last material
    that set it, em only from file materials."""
    k, r, h, t, e = oracle.materials([140], {})
    assert abs(k[0] - 130) < 1e-6 and abs(h[0] - 4200) < 1e-6 and abs(r[0] - 2328) < 1e-6
    k, *_ = oracle.materials([3001, 9002], {})
    assert k[0] == float(np.float32(0.19)) != 0.19 and k[1] == float(np.float32(0.209))
    _, _, _, t, e = oracle.materials([140, 320, 140, 7, 330, 350, 100, 3001], FILE_MATS)
    assert list(t) == [0.0, 1.0, 1.0, 3e-9, 3e-9, 1.0, 0.0, 0.0]
    assert list(e) == [0.0, 0.0, 0.0, 0.7, 0.7, 0.7, 0.7, 0.7]
    for bad in ([0], [-1], [5], [141]):
        with pytest.raises(ValueError):
            oracle.materials(bad, FILE_MATS)


@needs_ref
def test_materials_match_reference_build():
    """The oracle's material loop against the reference's own material() (ref harness), over
    5000 ids drawn from every built-in and three file materials: bit-exact."""
    rng = np.random.default_rng(11)
    ids = rng.choice(BUILTIN_IDS + list(FILE_MATS), size=5000)
    ids[:3] = [3001, 140, 9003]          # cells before the first tau / em assignment
    got = oracle.materials(ids, FILE_MATS)
    ref = oracle.ref_materials(ids, [dict(index=i, **m) for i, m in FILE_MATS.items()])
    for a, b in zip(got, ref):
        assert np.array_equal(a.view(np.int64), b.view(np.int64))


def _stability_system(n=(6, 5, 4), dt=1.0, kappa=None, kb=0.0, rhoc=(1.0, 1.0)):
    from heatflow_amd.system import Cells, Params, System
    nx, ny, nz = n
    N = nx * ny * nz
    p = Params(nx=nx, ny=ny, nz=nz, time_step=dt, kappa_bound=(kb,) * 6)
    k = np.full(N, 1.0) if kappa is None else np.asarray(kappa, np.float64)
    c = Cells(kappa=k, rho=np.full(N, rhoc[0]), heat_capacity=np.full(N, rhoc[1]),
              tau=np.zeros(N), em=np.zeros(N), len_x=np.full(N, 1.0), len_y=np.full(N, 0.5),
              len_z=np.full(N, 2.0), iheater=np.zeros(N, np.int64))
    return System(p, c)


def test_stability_closed_forms():
    """setup::stability (mod_setup.f90:158-199): var = dt kappa/(rho c_p) (1 + 4 + 1/4) for
    cells 1 x 0.5 x 2 against the single-precision 1/12.  A var between the double 1/12 and
    float(1/12) passes (pins the constant); the first failing cell is the smallest index in
    loop order; kappaBoundx1 is tested on the x1, y1 and z1 faces only."""
    s = 5.25
    thr32 = float(np.float32(1.0) / np.float32(12.0))
    assert thr32 > 1.0 / 12.0
    dt_mid = (1.0 / 12.0 + thr32) / 2 / s            # var strictly between the two constants
    assert oracle.OracleSystem(_stability_system(dt=dt_mid)).stability()[0] == -1
    dt_fail = thr32 * 1.0000001 / s
    cell, v, b = oracle.OracleSystem(_stability_system(dt=dt_fail)).stability()
    assert cell == 0 and not b and v > thr32
    # one interior cell 40 times more conductive fails alone, at its own index
    N = 120
    kap = np.full(N, 1.0)
    hot = 2 + 6 * (3 + 5 * 2)                          # (x, y, z) = (2, 3, 2), 0-based
    kap[hot] = 40.0
    dt = 0.05 / s
    cell, v, b = oracle.OracleSystem(_stability_system(dt=dt, kappa=kap)).stability()
    assert cell == hot and not b and abs(v - dt * 40.0 * s) <= 1e-15
    # kappaBoundx1 failing: cell 0 (on every low face) is the first
    cell, v, b = oracle.OracleSystem(_stability_system(dt=dt, kb=40.0)).stability()
    assert cell == 0 and b and abs(v - dt * 40.0 * s) <= 1e-15
    # the reference's own stability test case (test/test_mod_setup.f90:117-142: unit cell,
    # kappa = rho = c_p = 1, dt = 1e4) fails the condition: var = 3e4 (the reference stops)
    from heatflow_amd.system import Cells, Params, System
    one = np.ones(8)
    sysr = System(Params(nx=2, ny=2, nz=2, time_step=1e4),
                  Cells(kappa=one, rho=one, heat_capacity=one, tau=0 * one, em=0 * one,
                        len_x=one, len_y=one, len_z=one, iheater=np.zeros(8, np.int64)))
    cell, v, b = oracle.OracleSystem(sysr).stability()
    assert cell == 0 and v == 3e4 and not b

"""Oracle for HeatFlow's time-step path -- TEST INFRASTRUCTURE ONLY.

Two CPU implementations, used only by tests/, __graft_entry__.smoke() and the cpu_baseline leg
of bench.py, and only as the checker (never as the thing measured or shipped):

  C     liboracle.so: hf_oracle.c, a serial restatement of the reference's algorithm
        (mod_hmatrix, mod_setup, mod_boundary, mod_heating, mod_cattaneo, mod_evolve,
        mod_Sparse), each function citing the reference lines it follows.
  Ref   _ref/libhfref.so: the reference's own Fortran modules compiled from the HeatFlow
        sources by oracle/Makefile (amdflang -O3), driven through ref_harness.f90.  Built only
        where the HeatFlow checkout exists; the built file travels with the tree.

Parity pins (see tests/test_oracle.py): the C oracle is checked against the Ref build on random
inputs (COO bit-exact, B/Q/S_cat bit-exact, temperatures), against the reference's golden
test/data/TempDis.dat (10 Fourier steps, 25x25x1), against the HeatFlow manual's closed-form
3-cell cases, and against dense numpy solves and conservation invariants.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhfref.so")
REF_OMP_SO = os.path.join(HERE, "_ref", "libhfref_omp.so")   # the same with -fopenmp
REF_EXE = os.path.join(HERE, "_ref", "heatflow_ref")
REF_OMP_EXE = os.path.join(HERE, "_ref", "heatflow_ref_omp")   # the same with -fopenmp
REF_SRC = "/root/reference"


def build(ref: bool = True) -> None:
    """Builds liboracle.so, and the reference build when the HeatFlow sources are present."""
    targets = ["oracle"]
    if ref and os.path.isdir(os.path.join(REF_SRC, "src", "heatflow")):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)


class _Sys(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("time_step", C.c_double),
        ("icattaneo", C.c_int64), ("isteady", C.c_int64),
        ("periodic", C.c_int32 * 3),
        ("kappa_bound", C.c_double * 6), ("t_bath", C.c_double * 6),
        ("br", C.c_double),
        ("cg", C.c_int32 * 6),
        ("t_bath_cg", C.c_double), ("power_in", C.c_double), ("freq", C.c_double),
        ("heated_steps", C.c_int64), ("t_bath_rad", C.c_double),
        ("kappa", C.c_void_p), ("rho", C.c_void_p), ("heat_capacity", C.c_void_p),
        ("tau", C.c_void_p), ("em", C.c_void_p),
        ("len_x", C.c_void_p), ("len_y", C.c_void_p), ("len_z", C.c_void_p),
        ("volume", C.c_void_p), ("iheater", C.c_void_p), ("qdens_extra", C.c_void_p),
    ]


class _Solver(C.Structure):
    _fields_ = [("itol", C.c_int64), ("tol", C.c_double), ("itmax", C.c_int64)]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


_olib = None


def olib() -> C.CDLL:
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        _olib = C.CDLL(ORACLE_SO)
        _olib.hfo_hmatrixfunc.restype = C.c_double
        _olib.hfo_coo_len.restype = C.c_int64
        _olib.hfo_sparse_hmatrix.restype = C.c_int64
        _olib.hfo_heater.restype = C.c_double
        _olib.hfo_coo_hash.restype = C.c_int64
        _olib.hfo_stability.restype = C.c_int64
        _olib.hfo_coo_entry_hash.restype = C.c_uint64
        _olib.hfo_coo_entry_hash.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_double]
    return _olib


class OracleSystem:
    """A System bound to the C oracle (keeps the arrays alive)."""

    def __init__(self, system, qdens_extra: Optional[np.ndarray] = None):
        p, c = system.params, system.cells
        self.n = p.n
        f = lambda a: np.ascontiguousarray(a, np.float64)
        self.arrs = dict(
            kappa=f(c.kappa), rho=f(c.rho), heat_capacity=f(c.heat_capacity),
            tau=f(c.tau_scaled(p.time_step)), em=f(c.em), len_x=f(c.len_x), len_y=f(c.len_y),
            len_z=f(c.len_z), volume=f(c.volume),
            iheater=np.ascontiguousarray(c.iheater, np.int64))
        qe = qdens_extra if qdens_extra is not None else c.qdens_extra
        self.arrs["qdens_extra"] = None if qe is None else f(qe)
        s = _Sys()
        s.nx, s.ny, s.nz = p.nx, p.ny, p.nz
        s.time_step = p.time_step
        s.icattaneo, s.isteady = p.icattaneo, p.isteady
        for a in range(3):
            s.periodic[a] = int(bool(p.periodic[a]))
        for k in range(6):
            s.kappa_bound[k] = p.kappa_bound[k]
            s.t_bath[k] = p.t_bath[k]
            s.cg[k] = int(bool(p.cg[k]))
        s.br, s.t_bath_cg, s.power_in, s.freq = p.br, p.t_bath_cg, p.power_in, p.freq
        s.heated_steps, s.t_bath_rad = p.heated_steps, p.t_bath_rad
        for k, v in self.arrs.items():
            setattr(s, k, _p(v))
        self.s = s
        self.params = p

    def hmatrixfunc(self, i: int, j: int) -> float:
        return olib().hfo_hmatrixfunc(C.byref(self.s), C.c_int64(i), C.c_int64(j))

    def coo(self):
        L = olib().hfo_coo_len(C.byref(self.s))
        val = np.empty(L)
        irow = np.empty(L, np.int64)
        jcol = np.empty(L, np.int64)
        cnt = C.c_int64()
        olib().hfo_sparse_hmatrix(C.byref(self.s), _p(val), _p(irow), _p(jcol), C.byref(cnt))
        return self.n, L, cnt.value, val, irow, jcol

    def coo_hash(self):
        """(len, count, hash) of the COO sparse_Hmatrix would write, streamed (nothing stored):
        sum over positions k of mix(k, irow, jcol, value bits) mod 2^64 (hf_oracle.h)."""
        h = C.c_uint64()
        cnt = C.c_int64()
        L = olib().hfo_coo_hash(C.byref(self.s), C.byref(h), C.byref(cnt))
        return L, cnt.value, h.value

    def stability(self):
        """setup::stability over the cells: (first failing cell or -1, its value, whether the
        boundary test failed)."""
        v = C.c_double()
        b = C.c_int32()
        cell = olib().hfo_stability(C.byref(self.s), C.byref(v), C.byref(b))
        return cell, v.value, bool(b.value)

    def boundary(self) -> np.ndarray:
        B = np.empty(self.n)
        olib().hfo_boundary(C.byref(self.s), _p(B))
        return B

    def heater(self, itime: int, temp_p: np.ndarray):
        Q = np.empty(self.n)
        Qd = np.empty(self.n)
        tp = np.ascontiguousarray(temp_p, np.float64)
        hv = olib().hfo_heater(C.byref(self.s), C.c_int64(itime), _p(tp), _p(Q), _p(Qd))
        return Q, Qd, hv

    def scat(self, temp_p, temp_pp) -> np.ndarray:
        out = np.empty(self.n)
        a = np.ascontiguousarray(temp_p, np.float64)
        b = np.ascontiguousarray(temp_pp, np.float64)
        olib().hfo_scat(C.byref(self.s), _p(a), _p(b), _p(out))
        return out

    def run(self, ntime: int, temp_p=None, temp_pp=None, t_system: Optional[float] = None,
            tol: float = 1e-32, itmax: int = 50000, itol: int = 1):
        """Runs ntime steps (simulate(1..ntime)); returns (rc, T_p, T_pp, total_iters)."""
        init = temp_p is None
        tp = np.empty(self.n) if init else np.array(temp_p, np.float64)
        tpp = np.empty(self.n) if init else np.array(temp_pp if temp_pp is not None else temp_p,
                                                    np.float64)
        ts = self.params.t_system if t_system is None else t_system
        sv = _Solver(itol, tol, itmax)
        tot = C.c_int64()
        rc = olib().hfo_run(C.byref(self.s), C.byref(sv), C.c_int64(ntime), C.c_double(ts),
                            C.c_int(1 if init else 0), _p(tp), _p(tpp), C.byref(tot))
        return rc, tp, tpp, tot.value

    def linbcg(self, b, x, itol: int = 1, tol: float = 1e-32, itmax: int = 50000):
        """linbcg (mod_Sparse.f90:372-434) on this system's COO H; returns (x, iter, err)."""
        n, L, cnt, val, irow, jcol = self.coo()
        bb = np.ascontiguousarray(b, np.float64)
        xx = np.array(x, np.float64)
        it, err = C.c_int64(), C.c_double()
        olib().hfo_linbcg(C.c_int64(n), C.c_int64(L), _p(val), _p(irow), _p(jcol), _p(bb),
                          _p(xx), C.c_int64(itol), C.c_double(tol), C.c_int64(itmax),
                          C.byref(it), C.byref(err))
        return xx, it.value, err.value

    def simulate_steps(self, itimes, temp_p, temp_pp, tol: float = 1e-32, itmax: int = 50000):
        """simulate(itime) for each itime in order, from an explicit state; returns
        (rc, T_p, T_pp, iters per step)."""
        n, L, cnt, val, irow, jcol = self.coo()
        tp = np.array(temp_p, np.float64)
        tpp = np.array(temp_pp, np.float64)
        sv = _Solver(1, tol, itmax)
        iters = []
        for it in itimes:
            k = C.c_int64()
            rc = olib().hfo_simulate(C.byref(self.s), C.byref(sv), C.c_int64(L), _p(val),
                                     _p(irow), _p(jcol), C.c_int64(it), _p(tp), _p(tpp),
                                     C.byref(k), None)
            iters.append(k.value)
            if rc != 0:
                return rc, tp, tpp, iters
        return 0, tp, tpp, iters


# ---- the reference build (Fortran modules + ref_harness.f90) ------------------------------

# simulate() and linbcg() keep their N-vectors as automatic arrays on the stack
# (src/heatflow/mod_evolve.f90:47, src/heatflow/mod_Sparse.f90:384): every call into the
# reference build runs on a thread with a large stack.
REF_STACK = 2 << 30


def _big_stack(fn, *args):
    import threading
    out, err = {}, {}

    def target():
        try:
            out["v"] = fn(*args)
        except BaseException as e:  # re-raised in the caller
            err["e"] = e

    old = threading.stack_size()
    threading.stack_size(REF_STACK)
    try:
        t = threading.Thread(target=target)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "e" in err:
        raise err["e"]
    return out.get("v")


def ref_available() -> bool:
    return os.path.exists(REF_SO)


_rlibs = {}


def rlib(omp: bool = False) -> C.CDLL:
    """The reference build: serial, or with the reference's OpenMP directives (-fopenmp; its
    SpMV loops, mod_Sparse.f90:306-336, run in parallel with atomic updates)."""
    if omp not in _rlibs:
        path = REF_OMP_SO if omp else REF_SO
        if not os.path.exists(path):
            raise RuntimeError("reference build missing: run `make -C oracle ref`")
        _rlibs[omp] = C.CDLL(path)
    return _rlibs[omp]


def ref_omp_available() -> bool:
    return os.path.exists(REF_OMP_SO)


# material()'s built-in materials (src/heatflow/mod_material.f90:116-265): index ->
# (kappa, rho, heat_capacity, tau or None where the case leaves tau as it was).  The literals
# there are default REAL, so each value is the single-precision rounding of the decimal.
_BUILTIN_MATERIALS = {
    100: (1.0, 1.0, 1.0, 0.0),            # :118-126
    140: (130.0, 2328.0, 4200.0, None),   # :136-144
    3001: (0.19, 941.0, 840.0, None),     # :146-153
    3002: (1.42, 1700.0, 1100.0, None),   # :155-161
    3003: (1.59, 1800.0, 1400.0, None),   # :163-169
    3004: (1.25, 2300.0, 2300.0, None),   # :171-177
    9001: (0.6, 1090.0, 3421.0, None),    # :179-185
    9002: (0.209, 1109.0, 3391.0, None),  # :187-193
    9003: (0.317, 911.0, 2348.0, None),   # :195-201
    310: (130.0, 2329.0, 700.0, None),    # :203-209
    320: (1.4, 2400.0, 720.0, 1.0),       # :211-218
    330: (427.0, 10524.0, 236.0, None),   # :220-226
    340: (0.2, 1153.0, 1500.0, 1.0),      # :228-235
    350: (10.2, 7140.0, 361.0, 1.0),      # :237-244
}


def materials(ids, file_materials=None):
    """set_global_variables' per-cell loop over material() (src/heatflow/mod_setup.f90:55-68,
    src/heatflow/mod_material.f90:54-265), cell by cell in loop order: (kappa, rho,
    heat_capacity, tau, em) per cell, tau unscaled.  file_materials: {index: {key: value}} as
    read_mat returns mat.in.  material() assigns tau and em for file materials, tau for the
    built-ins that set it and em for none of them; a value it leaves alone is the one the
    previous call left in set_global_variables' locals (0 before any call: the reference
    leaves them uninitialised).  Unknown or non-positive ids raise ValueError (material()
    prints and exits, :76-80, :104-107, :261-264)."""
    f32 = lambda v: float(np.float32(v))
    fm = file_materials or {}
    ids = np.asarray(ids, np.int64).reshape(-1)
    out = np.zeros((5, ids.size))
    tau = em = 0.0
    for i, mid in enumerate(ids.tolist()):
        if mid <= 0:
            raise ValueError(f"imaterial_type not recognized: {mid}")
        if mid <= 99:                                          # :86-108
            if mid not in fm:
                raise ValueError(f"imaterial_type not recognized: {mid}")
            m = fm[mid]
            kappa, rho, hc, tau, em = (float(m[k]) for k in
                                       ("kappa", "rho", "heat_capacity", "tau", "em"))
        else:                                                  # :115-265
            if mid not in _BUILTIN_MATERIALS:
                raise ValueError(f"imaterial_type not recognized: {mid}")
            k, r, h, t = _BUILTIN_MATERIALS[mid]
            kappa, rho, hc = f32(k), f32(r), f32(h)
            if t is not None:
                tau = f32(t)
        out[:, i] = (kappa, rho, hc, tau, em)
    return tuple(out)


def ref_materials(ids, file_materials=()):
    """The reference's own material() (src/heatflow/mod_material.f90:40-265) over the ids in
    order, as set_global_variables' loop calls it (mod_setup.f90:55-68): per id (kappa, rho,
    heat_capacity, tau, em) with tau unscaled.  file_materials: dicts with index, kappa, rho,
    heat_capacity, tau, em (mat.in's entries)."""
    ids = np.ascontiguousarray(ids, np.int64)
    fm = list(file_materials)
    col = lambda k, dt=np.float64: np.ascontiguousarray([m[k] for m in fm] or [0], dt)
    fidx, fk, fr, fh, ft, fe = (col("index", np.int64), col("kappa"), col("rho"),
                                col("heat_capacity"), col("tau"), col("em"))
    out = [np.empty(len(ids)) for _ in range(5)]
    _big_stack(rlib().hfr_materials, C.c_int64(len(ids)), _p(ids), C.c_int64(len(fm)),
               _p(fidx), _p(fk), _p(fr), _p(fh), _p(ft), _p(fe), *[_p(a) for a in out])
    return tuple(out)


class RefSystem:
    """The reference modules set up with a System (one at a time: module globals)."""

    def __init__(self, system, omp: bool = False):
        p, c = system.params, system.cells
        self.lib = rlib(omp)
        self.n = p.n
        self.params = p
        f = lambda a: np.ascontiguousarray(a, np.float64)
        ip = np.array([p.nx, p.ny, p.nz, p.icattaneo, p.isteady, *[int(bool(v)) for v in p.periodic],
                       *[int(bool(v)) for v in p.cg], p.heated_steps], np.int64)
        dp = np.array([p.time_step, *p.kappa_bound, *p.t_bath, p.br, p.t_bath_cg, p.power_in,
                       p.freq, p.t_bath_rad], np.float64)
        self.keep = [ip, dp, f(c.kappa), f(c.rho), f(c.heat_capacity), f(c.tau), f(c.em),
                     f(c.len_x), f(c.len_y), f(c.len_z),
                     np.ascontiguousarray(c.iheater, np.int64)]
        _big_stack(self.lib.hfr_setup, *[_p(a) for a in self.keep])

    def coo(self):
        n, L = C.c_int64(), C.c_int64()
        self.lib.hfr_coo_sizes(C.byref(n), C.byref(L))
        val = np.empty(L.value)
        irow = np.empty(L.value, np.int64)
        jcol = np.empty(L.value, np.int64)
        self.lib.hfr_coo(_p(val), _p(irow), _p(jcol))
        return n.value, L.value, val, irow, jcol

    def set_state(self, tp, tpp):
        a = np.ascontiguousarray(tp, np.float64)
        b = np.ascontiguousarray(tpp, np.float64)
        self.lib.hfr_set_state(_p(a), _p(b))

    def get_state(self):
        a = np.empty(self.n)
        b = np.empty(self.n)
        self.lib.hfr_get_state(_p(a), _p(b))
        return a, b

    def simulate(self, itime: int):
        _big_stack(self.lib.hfr_simulate, C.c_int64(itime))

    def simulate_qextra(self, itime: int, qextra, tol: float = 1e-32) -> int:
        q = np.ascontiguousarray(qextra, np.float64)
        it = C.c_int64()
        _big_stack(self.lib.hfr_simulate_qextra, C.c_int64(itime), _p(q), C.c_double(tol),
                   C.byref(it))
        return it.value

    def boundary(self):
        b = np.empty(self.n)
        _big_stack(self.lib.hfr_boundary, _p(b))
        return b

    def heater(self, itime: int):
        q = np.empty(self.n)
        qd = np.empty(self.n)
        hv = C.c_double()
        _big_stack(self.lib.hfr_heater, C.c_int64(itime), _p(q), _p(qd), C.byref(hv))
        return q, qd, hv.value

    def scat(self):
        s = np.empty(self.n)
        _big_stack(self.lib.hfr_scat, _p(s))
        return s

    def linbcg(self, b, x, tol=1e-32, itmax=50000, itol=1):
        bb = np.ascontiguousarray(b, np.float64)
        xx = np.array(x, np.float64)
        it, err = C.c_int64(), C.c_double()
        _big_stack(self.lib.hfr_linbcg, _p(bb), _p(xx), C.c_int64(itol), C.c_double(tol),
                   C.c_int64(itmax), C.byref(it), C.byref(err))
        return xx, it.value, err.value

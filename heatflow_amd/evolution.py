"""Python mirror of HeatFlow's time-step interface over the gfx950 library.

`Simulation` exposes the reference's procedure names on one device context:
  sparse_Hmatrix()        setup::sparse_Hmatrix      src/heatflow/mod_setup.f90:89-152
  initial_evolve(...)     INITIAL::initial_evolve    src/heatflow/mod_init_evolve.f90:34-51
  simulate(itime)         evolution::simulate        src/heatflow/mod_evolve.f90:45-177
  linbcg(b, x, ...)       solver::linbcg             src/heatflow/mod_Sparse.f90:372-434
  boundary(), heater()    read back B and Qdens of the device path
Errors: the six NaN guards of simulate() raise HeatFlowError with codes 1..6 (the Fortran
program prints "fatal error: NAN in <...>" and stops with status 1).  Every call runs on the
GPU; without the built library construction raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import HF_PREC_FP32_IR, HF_PREC_FP64, HeatFlowError, HfParams, HfStats, check
from .system import Geometry, Params, System

PRECISIONS = {"fp64": HF_PREC_FP64, "fp32_ir": HF_PREC_FP32_IR}


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _anyptr(a):
    """Pointer to a contiguous float64 array: numpy (host) or torch tensor (host or device).
    The library copies with hipMemcpyDefault, so either works."""
    if a is None:
        return None, None
    if hasattr(a, "data_ptr"):
        import torch
        t = a.detach()
        if t.dtype != torch.float64:
            t = t.double()
        t = t.contiguous()
        if t.is_cuda:
            torch.cuda.current_stream(t.device).synchronize()
        return C.c_void_p(t.data_ptr()), t
    arr = _f64(a)
    return _ptr(arr), arr


def make_params(system: System, precision: str = "fp64", tol: float = 1.0e-32,
                itmax: int = 50000, inner_tol: float = 1.0e-6, inner_itmax: int = 200) -> HfParams:
    lib = _lib.load()
    p = HfParams()
    lib.hf_params_default(C.byref(p))
    q = system.params
    p.nx, p.ny, p.nz = q.nx, q.ny, q.nz
    p.time_step = q.time_step
    p.icattaneo = int(q.icattaneo)
    p.isteady = int(q.isteady)
    for a in range(3):
        p.periodic[a] = int(bool(q.periodic[a]))
    for f in range(6):
        p.cg[f] = int(bool(q.cg[f]))
        p.kappa_bound[f] = float(q.kappa_bound[f])
        p.t_bath[f] = float(q.t_bath[f])
    p.br = q.br
    p.t_bath_cg = q.t_bath_cg
    p.power_in = q.power_in
    p.freq = q.freq
    p.t_bath_rad = q.t_bath_rad
    p.heated_steps = int(q.heated_steps)
    p.tol = tol
    p.itmax = itmax
    p.precision = PRECISIONS[precision]
    p.inner_tol = inner_tol
    p.inner_itmax = inner_itmax
    return p


class Simulation:
    """One HeatFlow grid on the current HIP device."""

    def __init__(self, system: System, precision: str = "fp64", tol: float = 1.0e-32,
                 itmax: int = 50000, inner_tol: float = 1.0e-6, inner_itmax: int = 200,
                 assemble: bool = True):
        self.lib = _lib.load()
        self.forward_generation = 0
        self.system = system
        self.n = system.params.n
        self.precision = precision
        p = make_params(system, precision, tol, itmax, inner_tol, inner_itmax)
        ctx = C.c_void_p()
        check(self.lib.hf_create(C.byref(ctx), C.byref(p)), None)
        self.ctx = ctx
        if system.cells is not None:
            self.upload(system)
            if assemble:
                self.sparse_Hmatrix()

    @classmethod
    def from_geometry(cls, params: Params, geometry: Geometry, file_materials: dict = None,
                      check_stability: bool = False, assemble: bool = True, **solver):
        """set_global_variables on the device (src/heatflow/mod_setup.f90:34-68): the per-cell
        kappa, rho, heat_capacity, tau/dt^2 and em come from material() over the geometry's
        material ids (built-in materials, index > 99, and file materials {index: {key: value}}
        as read_mat returns them), with stability() over every cell when check_stability
        (the reference's _Check_Stability), then sparse_Hmatrix unless assemble is False.
        `system.cells` stays None: the fields live on the device (Simulation.vector)."""
        if (geometry.nx, geometry.ny, geometry.nz) != (params.nx, params.ny, params.nz):
            raise ValueError("geometry and params disagree on the grid")
        sim = cls(System(params, None), assemble=False, **solver)
        try:
            sim.upload_materials(geometry, file_materials or {})
            if check_stability:
                sim.check_stability()
            if assemble:
                sim.sparse_Hmatrix()
        except BaseException:
            sim.close()
            raise
        return sim

    # -- setup --------------------------------------------------------------------------------
    def upload(self, system: System, qdens_extra: Optional[np.ndarray] = None) -> None:
        c = system.cells
        self._keep = [
            _f64(c.kappa), _f64(c.rho), _f64(c.heat_capacity),
            _f64(c.tau_scaled(system.params.time_step)), _f64(c.em), _f64(c.len_x),
            _f64(c.len_y), _f64(c.len_z), _f64(c.volume),
            np.ascontiguousarray(c.iheater, dtype=np.int64),
        ]
        qe = qdens_extra if qdens_extra is not None else c.qdens_extra
        qe = None if qe is None else _f64(qe)
        check(self.lib.hf_upload_cells(self.ctx, *[_ptr(a) for a in self._keep], _ptr(qe)),
              self.ctx)

    def upload_materials(self, geometry: Geometry, file_materials: dict) -> None:
        """Per-cell fields from material ids on the device (hf_upload_materials)."""
        keys = ("kappa", "rho", "heat_capacity", "tau", "em")
        fm = (_lib.HfMaterial * max(len(file_materials), 1))()
        for k, (idx, m) in enumerate(sorted(file_materials.items())):
            fm[k].index = int(idx)
            for key in keys:
                setattr(fm[k], key, float(m[key]))
        ids = np.ascontiguousarray(geometry.imaterial_type, np.int64)
        heat = np.ascontiguousarray(geometry.iheater, np.int64)
        if ids.size != self.n or heat.size != self.n:
            raise ValueError("geometry arrays must hold nx*ny*nz cells")
        lx, ly, lz = geometry.cell_lengths
        check(self.lib.hf_upload_materials(self.ctx, _ptr(ids), _ptr(heat), len(file_materials),
                                           fm, lx, ly, lz), self.ctx)

    def stability(self):
        """setup::stability over every cell (mod_setup.f90:158-199): (first failing cell in
        loop order or -1, its failing var_stability or 0)."""
        cell, value = C.c_int64(-1), C.c_double(0.0)
        rc = self.lib.hf_check_stability(self.ctx, C.byref(cell), C.byref(value))
        if rc not in (_lib.HF_OK, _lib.HF_ERR_UNSTABLE):
            check(rc, self.ctx)
        return cell.value, value.value

    def check_stability(self) -> None:
        """As the reference's _Check_Stability does: raises HeatFlowError (HF_ERR_UNSTABLE)
        naming the first failing cell, where the reference prints and stops."""
        check(self.lib.hf_check_stability(self.ctx, None, None), self.ctx)

    def sparse_Hmatrix(self) -> None:
        check(self.lib.hf_assemble(self.ctx), self.ctx)

    def export_coo(self):
        """(n, len, count, val, irow, jcol) in the reference's sprs2_dp layout (1-based)."""
        n, ln, cnt = C.c_int64(), C.c_int64(), C.c_int64()
        check(self.lib.hf_export_coo(self.ctx, C.byref(n), C.byref(ln), C.byref(cnt), None, None,
                                     None), self.ctx)
        val = np.empty(ln.value, np.float64)
        irow = np.empty(ln.value, np.int64)
        jcol = np.empty(ln.value, np.int64)
        check(self.lib.hf_export_coo(self.ctx, C.byref(n), C.byref(ln), C.byref(cnt), _ptr(val),
                                     _ptr(irow), _ptr(jcol)), self.ctx)
        return n.value, ln.value, cnt.value, val, irow, jcol

    # -- state --------------------------------------------------------------------------------
    def initial_evolve(self, t_system: Optional[float] = None, temp_p=None, temp_pp=None) -> None:
        if temp_p is None:
            ts = self.system.params.t_system if t_system is None else t_system
            temp_p = np.full(self.n, ts)
        if temp_pp is None:
            temp_pp = temp_p
        self.set_state(temp_p, temp_pp)

    def set_state(self, temp_p, temp_pp) -> None:
        a, b = _f64(temp_p), _f64(temp_pp)
        check(self.lib.hf_set_state(self.ctx, _ptr(a), _ptr(b)), self.ctx)

    def get_state(self):
        a = np.empty(self.n)
        b = np.empty(self.n)
        check(self.lib.hf_get_state(self.ctx, _ptr(a), _ptr(b)), self.ctx)
        return a, b

    # -- time stepping ------------------------------------------------------------------------
    def simulate(self, itime: int):
        it, err = C.c_int64(), C.c_double()
        check(self.lib.hf_step(self.ctx, int(itime), C.byref(it), C.byref(err)), self.ctx)
        return it.value, err.value

    def run(self, itime0: int, nsteps: int) -> int:
        tot = C.c_int64()
        check(self.lib.hf_run(self.ctx, int(itime0), int(nsteps), C.byref(tot)), self.ctx)
        return tot.value

    def linbcg(self, b, x, itol: int = 1, tol: float = 1.0e-32, itmax: int = 50000):
        bb, xx = _f64(b), _f64(x).copy()
        it, err = C.c_int64(), C.c_double()
        check(self.lib.hf_solve(self.ctx, _ptr(bb), _ptr(xx), itol, tol, itmax, C.byref(it),
                                C.byref(err)), self.ctx)
        return xx, it.value, err.value

    def vector(self, which: str) -> np.ndarray:
        ids = {"B": 0, "Qdens": 1, "S": 2, "diag": 3, "cx": 4, "cy": 5, "cz": 6, "kappa": 7,
               "rhoc": 8, "tau_s": 9, "ells": 10, "volume": 11}
        out = np.empty(self.n)
        check(self.lib.hf_get_vector(self.ctx, ids[which], _ptr(out)), self.ctx)
        return out

    def set_debug(self, flags: int) -> None:
        check(self.lib.hf_set_debug(self.ctx, flags), self.ctx)

    def set_solver(self, tol: float, itmax: int, precision: str = "fp64",
                   inner_tol: float = 1.0e-6, inner_itmax: int = 200) -> None:
        check(self.lib.hf_set_solver(self.ctx, tol, itmax, PRECISIONS[precision], inner_tol,
                                     inner_itmax), self.ctx)
        self.precision = precision

    def stats(self) -> HfStats:
        st = HfStats()
        check(self.lib.hf_get_stats(self.ctx, C.byref(st)), self.ctx)
        return st

    def reset_stats(self) -> None:
        check(self.lib.hf_reset_stats(self.ctx), self.ctx)

    def grad(self, itime0: int, nsteps: int, ckpt_every: int, probes, targets, weights=None):
        """Forward nsteps with checkpoints, then the reverse-time adjoint.  Returns
        (loss, dL/dkappa, dL/dP) where P_I enters as Qdens_I += P_I / volume_I.  ckpt_every: the
        checkpoint spacing; 1 keeps every state (nothing recomputed), 0 picks 1 when that fits
        in free device memory, else ceil(sqrt(nsteps))."""
        pi = np.ascontiguousarray(probes, dtype=np.int64)
        tg = _f64(targets)
        w = None if weights is None else _f64(weights)
        gk = np.empty(self.n)
        gp = np.empty(self.n)
        loss = C.c_double()
        check(self.lib.hf_grad(self.ctx, int(itime0), int(nsteps), int(ckpt_every), len(pi),
                               _ptr(pi), _ptr(tg), _ptr(w), C.byref(loss), _ptr(gk), _ptr(gp)),
              self.ctx)
        return loss.value, gk, gp

    def set_timing(self, on: bool) -> None:
        check(self.lib.hf_set_timing(self.ctx, int(bool(on))), self.ctx)

    def kernel_times(self) -> dict:
        """{kernel name: (ms total, launches, algorithmic bytes)} since timing was enabled."""
        k = len(_lib.KERNELS)
        ms = np.zeros(k)
        n = np.zeros(k, np.int64)
        by = np.zeros(k)
        check(self.lib.hf_kernel_times(self.ctx, _ptr(ms), _ptr(n), _ptr(by)), self.ctx)
        return {name: (float(ms[i]), int(n[i]), float(by[i]))
                for i, name in enumerate(_lib.KERNELS) if n[i] > 0}

    def set_fields(self, kappa=None, qdens_extra=None) -> None:
        """Replaces kappa and/or the per-cell power density (numpy or torch); re-assemble next."""
        pk, keep_k = _anyptr(kappa)
        pq, keep_q = _anyptr(qdens_extra)
        check(self.lib.hf_set_fields(self.ctx, pk, pq), self.ctx)

    def set_adjoint_tol(self, tol: float) -> None:
        """Scaled-residual tolerance of the adjoint solves (0: the forward tolerance)."""
        check(self.lib.hf_set_adjoint_tol(self.ctx, float(tol)), self.ctx)

    def forward_checkpointed(self, itime0: int, nsteps: int, ckpt_every: int) -> None:
        # every forward replaces the context's checkpoints: the count tells an autograd graph
        # whether its own are still there
        self.forward_generation += 1
        check(self.lib.hf_forward_checkpointed(self.ctx, int(itime0), int(nsteps),
                                               int(ckpt_every)), self.ctx)

    def adjoint(self, dl_dt_final, out_kappa=None, out_power=None):
        """Reverse sweep of the last forward_checkpointed run.  Outputs may be numpy arrays or
        contiguous float64 torch tensors (filled in place); numpy arrays are returned if None."""
        ps, keep = _anyptr(dl_dt_final)
        if out_kappa is None:
            out_kappa = np.empty(self.n)
        if out_power is None:
            out_power = np.empty(self.n)
        pk = C.c_void_p(out_kappa.data_ptr()) if hasattr(out_kappa, "data_ptr") else _ptr(out_kappa)
        pp = C.c_void_p(out_power.data_ptr()) if hasattr(out_power, "data_ptr") else _ptr(out_power)
        check(self.lib.hf_adjoint(self.ctx, ps, pk, pp), self.ctx)
        return out_kappa, out_power

    def state_into(self, temp_p, temp_pp=None) -> None:
        """Copies T^n (and T^{n-1}) into preallocated float64 torch tensors or numpy arrays."""
        pa = C.c_void_p(temp_p.data_ptr()) if hasattr(temp_p, "data_ptr") else _ptr(temp_p)
        pb = None
        if temp_pp is not None:
            pb = C.c_void_p(temp_pp.data_ptr()) if hasattr(temp_pp, "data_ptr") else _ptr(temp_pp)
        check(self.lib.hf_get_state(self.ctx, pa, pb), self.ctx)

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.hf_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

"""Benchmark of HeatFlow's time step on one MI355X.

Metric: timesteps/s of the FP32-inner / FP64-refined solver on the 512^3 log-normal-conductivity
Cattaneo workload (20 steps, Gaussian heating, Dirichlet faces), next to the plain FP64 path at
the same tolerance and at the reference's own tolerance (1e-32), and the mixed path with the
reuse of the inner step length off, with the max/rms relative temperature error of each (and
the error on the heating signal T - 300 K) against the FP64 run at the reference tolerance.
Also reported: the same grid at a 1000x time step (about 60 CG iterations per step instead of
one); the forward+adjoint gradient at 128^3 / 200 steps with central finite differences; the
mixed-precision forward+adjoint at 384^3 / 500 steps; the 16^3 workload on the persistent
small-grid path and on the multi-kernel path; the 256^3 two-phase workload in FP64; and the
reference Fortran timed on the host (its modules on a 128^3 sample of the 512^3 workload, and
its program and the Fortran drop-in on the 16^3 input files; `--ref-c2` times them on the
256^3 input files).

A "step" is one time step (evolution::simulate, src/heatflow/mod_evolve.f90:45-177) of the
whole grid.  W untimed warm-up steps run first, then the state is reset and the K timed steps
run from the initial condition; the timed region is bracketed by torch.cuda.synchronize().

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, 8 TB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sync():
    import torch
    torch.cuda.synchronize()


def timed_run(sim, steps: int, warmup: int, timing: bool = True):
    """warm-up, reset, then `steps` timed steps (with kernel timing unless `timing` is False:
    its events cost a few microseconds a launch, which counts on small grids); returns
    (seconds, stats, kernel times, final state)."""
    sim.initial_evolve()
    if warmup:
        sim.run(1, warmup)
    sim.initial_evolve()
    sim.reset_stats()
    sim.set_timing(timing)
    sync()
    t0 = time.perf_counter()
    sim.run(1, steps)
    sync()
    dt = time.perf_counter() - t0
    kt = sim.kernel_times() if timing else {}
    sim.set_timing(False)
    st = sim.stats()
    tp, _ = sim.get_state()
    return dt, st, kt, tp


def kernel_summary(kt: dict, n_cells: int) -> dict:
    out = {}
    for name, (ms, launches, by) in kt.items():
        out[name] = dict(ms_total=round(ms, 4), launches=launches,
                         us_per_launch=round(1000.0 * ms / launches, 3),
                         bytes_per_cell=round(by / launches / n_cells, 2),
                         gbs=round(by / (ms * 1e-3) / 1e9, 1))
    return out


def roofline(kt: dict) -> dict:
    name = max(kt, key=lambda k: kt[k][0])
    ms, launches, by = kt[name]
    per_launch_bytes = by / launches
    avg_s = ms * 1e-3 / launches
    ach = per_launch_bytes / avg_s / 1e9
    return dict(kernel=name, bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(ach / HBM_PEAK_GBS, 4), traffic=None,
                algorithmic_bytes_per_launch=per_launch_bytes,
                avg_launch_us=round(avg_s * 1e6, 2))


def measured_traffic(kernel: str, n: int):
    """HBM bytes per launch of `kernel` at n^3 from the committed rocprofv3 counter summaries
    (FETCH_SIZE / WRITE_SIZE passes, scaled by a copy of known size), or None."""
    for mode in ("fp32-ir", "fp64"):
        path = os.path.join(ROOT, "profiles", f"hbm_traffic_c3-{n}_{mode}.json")
        if not os.path.exists(path):
            continue
        with open(path) as f:
            d = json.load(f)
        if d.get("n") != n:
            continue
        entries = d.get("kernels", {d.get("kernel"): d})
        if kernel in entries:
            return entries[kernel]["bytes_per_launch"]
    return None


def rel_errors(a, ref):
    e = np.abs(a - ref) / np.abs(ref)
    return float(e.max()), float(np.sqrt(np.mean(e * e)))


def err_summary(a, ref, t_bath: float = 300.0) -> dict:
    """max / rms error relative to T, and the largest error on the signal T - T_bath relative
    to that signal's peak (the ~300 K offset hides the error of the heating response)."""
    mx, rms = rel_errors(a, ref)
    peak = float(np.abs(ref - t_bath).max())
    return dict(max=mx, rms=rms, signal=float(np.abs(a - ref).max()) / peak if peak > 0 else None,
                signal_peak_K=peak)


def _set_param(path: str, key: str, value) -> None:
    with open(path) as f:
        lines = f.readlines()
    with open(path, "w") as f:
        for ln in lines:
            if ln.split("=", 1)[0].strip() == key:
                ln = f"    {key} = {value}\n"
            f.write(ln)


def program_rate(s, exe: str, workdir: str, n1: int, n2: int, threads: int = 1,
                 inputs_ready: bool = False, repeats: int = 1) -> dict:
    """timesteps/s of a HeatFlow program (the reference, or the drop-in) from the wall-clock
    difference of two runs of n1 < n2 steps on the same inputs, which removes input parsing,
    setup and assembly (the fastest of `repeats` runs of each).  Output off (IVERB 0, no text
    output); stack unlimited (the reference's automatic arrays); `threads` OpenMP threads,
    bound spread."""
    import resource
    import subprocess
    from heatflow_amd.system import write_inputs
    inp = os.path.join(workdir, "inputs")
    if not inputs_ready:
        write_inputs(s, inp, ntime=n1)
    os.makedirs(os.path.join(workdir, "outputs"), exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread")

    def unlimited_stack():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    times = []
    for nt in (n1, n2):
        _set_param(os.path.join(inp, "param.in"), "ntime", nt)
        best = None
        for _ in range(repeats):
            t0 = time.perf_counter()
            p = subprocess.run([exe], cwd=workdir, env=env, preexec_fn=unlimited_stack,
                               capture_output=True, text=True)
            t = time.perf_counter() - t0
            if p.returncode != 0:
                raise RuntimeError(f"{exe} failed ({p.returncode}): {p.stdout[-500:]} "
                                   f"{p.stderr[-500:]}")
            best = t if best is None else min(best, t)
        times.append(best)
        log(f"[bench] {os.path.basename(exe)} {nt} steps: {times[-1]:.3f} s")
    per = (times[1] - times[0]) / (n2 - n1)
    return dict(value=round(1.0 / per, 4) if per > 0 else None, unit="timesteps/s",
                cores=threads, seconds=[round(t, 3) for t in times], steps=[n1, n2])


def _omp_threads() -> int:
    """Threads the reference's OpenMP build runs with (omp_get_max_threads of the libomp it
    links)."""
    import oracle
    return int(oracle.rlib(True).omp_get_max_threads())


def cpu_baseline(n_full: int, tol: float, sample_n: int, sample_steps: int) -> dict:
    """The reference Fortran path (oracle/_ref, amdflang -O3) on the same physics at a reduced
    grid, serial and with the reference's OpenMP directives on this host's threads; scaled to
    timesteps/s at n_full^3 by cell count (iterations per step do not grow with N for this
    mass-dominated operator).  The faster build is the reported value.  Also times 2 steps at
    the reference's own tolerance, 1e-32.  Falls back to the C restatement (serial)."""
    import oracle
    from heatflow_amd import workloads
    s, _ = workloads.c3(sample_n)
    n = s.params.n
    tp = np.full(n, 300.0)

    def ref_rate(omp: bool, steps: int, t: float):
        r = oracle.RefSystem(s, omp=omp)
        r.set_state(tp, tp)
        t0 = time.perf_counter()
        its = [r.simulate_qextra(k, s.cells.qdens_extra, t) for k in range(1, steps + 1)]
        dt = time.perf_counter() - t0
        return n * steps / dt / float(n_full) ** 3, dt, sum(its)

    if not oracle.ref_available():
        o = oracle.OracleSystem(s)
        t0 = time.perf_counter()
        rc, _, _, tot = o.run(sample_steps, tol=tol)
        dt = time.perf_counter() - t0
        return dict(value=round(n * sample_steps / dt / float(n_full) ** 3, 6),
                    unit="timesteps/s", cores=1, kind="port",
                    sample=(f"C restatement on the same workload at {sample_n}^3, {sample_steps} "
                            f"steps, tol {tol:g}: {dt:.2f} s, {tot} solver iterations; scaled by "
                            f"cell count to {n_full}^3"), sample_seconds=round(dt, 3))
    v1, dt1, it1 = ref_rate(False, sample_steps, tol)
    out = dict(serial=dict(value=round(v1, 6), cores=1, seconds=round(dt1, 3), iterations=it1))
    best, cores, dtb, itb, label = v1, 1, dt1, it1, "serial"
    if oracle.ref_omp_available():
        threads = _omp_threads()
        v2, dt2, it2 = ref_rate(True, sample_steps, tol)
        out["openmp"] = dict(value=round(v2, 6), cores=threads, seconds=round(dt2, 3),
                             iterations=it2)
        if v2 > v1:
            best, cores, dtb, itb, label = v2, threads, dt2, it2, f"OpenMP, {threads} threads"
    v3, dt3, it3 = ref_rate(cores > 1, 2, 1e-32)
    out["reference_tol_1e-32"] = dict(value=round(v3, 6), steps=2, seconds=round(dt3, 3),
                                      iterations=it3)
    return dict(value=round(best, 6), unit="timesteps/s", cores=cores, kind="reference",
                sample=(f"reference Fortran modules (amdflang -O3, COO BiCG, {label}) on the "
                        f"same workload at {sample_n}^3, {sample_steps} steps, tol {tol:g}: "
                        f"{dtb:.2f} s, {itb} solver iterations; scaled by cell count to "
                        f"{n_full}^3"),
                sample_seconds=round(dtb, 3), builds=out)


DROPIN_EXE = os.path.join(ROOT, "heatflow_amd", "fortran", "_build", "heatflow_hip")


def c1_bench(steps: int, warmup: int, cpu: bool) -> dict:
    """The 16^3 Cattaneo workload (1 um cells, one centre heater of 1e-4 W, x faces 300 K, y/z
    insulating, dt 1e-10 s), 100 steps in FP64 at the configuration's CG tolerance 1e-10 and at
    1e-12 (which keeps the temperatures within 1e-10 of the reference tolerance's), on the
    persistent small-grid path and on the multi-kernel path (debug bit 6); errors against FP64
    at 1e-32.  Kernel timing is off (its events would add to the latency measured here).
    With `cpu`: the reference program (serial and OpenMP) and the Fortran drop-in timed on the
    same input files by the two-run difference (100 and 1000 steps, fastest of three runs
    each), at the reference's own tolerance 1e-32."""
    import shutil
    import tempfile
    from heatflow_amd import Simulation, workloads
    s, _ = workloads.c1(16)
    out, states = {}, {}
    for label, dbg, tol in (("persistent", 0, 1e-10), ("persistent", 0, 1e-12),
                            ("multi_kernel", 64, 1e-10), ("multi_kernel", 64, 1e-12),
                            ("persistent", 0, 1e-32)):
        key = f"{label}@{tol:g}"
        with Simulation(s, tol=tol) as sim:
            sim.set_debug(dbg)
            dt, st, _, tp = timed_run(sim, steps, warmup, timing=False)
        states[key] = tp
        out[key] = dict(timesteps_per_s=steps / dt, ms_per_step=1000.0 * dt / steps,
                        cg_iters=int(st.outer_iters), cg_iters_per_s=st.outer_iters / dt)
        log(f"[bench] c1 {key}: {steps / dt:.1f} steps/s, {st.outer_iters} CG iterations")
    ref = states["persistent@1e-32"]
    for key in out:
        if key != "persistent@1e-32":
            out[key]["err_vs_fp64_reftol"] = err_summary(states[key], ref)
    out["persistent_over_multi_kernel"] = {
        f"{t:g}": out[f"persistent@{t:g}"]["timesteps_per_s"]
        / out[f"multi_kernel@{t:g}"]["timesteps_per_s"] for t in (1e-10, 1e-12)}
    if cpu:
        import oracle
        progs = [("reference_serial", oracle.REF_EXE, 1),
                 ("reference_openmp", oracle.REF_OMP_EXE, _host_threads()),
                 ("dropin_gpu", DROPIN_EXE, 1)]
        for label, exe, thr in progs:
            if not os.path.exists(exe):
                out[label] = None
                continue
            d = tempfile.mkdtemp(prefix="hf_c1_")
            try:
                out[label] = program_rate(s, exe, d, 100, 1000, threads=thr, repeats=3)
            finally:
                shutil.rmtree(d, ignore_errors=True)
    return out


def _host_threads() -> int:
    """OpenMP threads for the reference's parallel build: OMP_NUM_THREADS when set (the GPU
    box sets its CPU share), else the usable cores."""
    v = os.environ.get("OMP_NUM_THREADS")
    return int(v) if v and v.isdigit() else len(os.sched_getaffinity(0))


def c2_bench(n: int, steps: int, warmup: int) -> dict:
    """The 256^3 two-phase voxel workload (Bernoulli(0.3) phase B kappa 0.2 / rho 1200 / c_p
    1500 in phase A kappa 150 / rho 2330 / c_p 700, 32 heaters of 1e-5 W, x faces 300 K,
    dt 5e-11 s), 50 steps in FP64 at the configuration's CG tolerance 1e-10, at 1e-12 and at
    1e-14; errors against FP64 at 1e-32; kernel table and roofline of each."""
    from heatflow_amd import Simulation, workloads
    s, _ = workloads.c2(n)
    N = s.params.n
    out, states = {}, {}
    for tol in (1e-10, 1e-12, 1e-14, 1e-32):
        key = f"fp64@{tol:g}"
        with Simulation(s, tol=tol) as sim:
            dt, st, kt, tp = timed_run(sim, steps, warmup)
        states[key] = tp
        out[key] = dict(timesteps_per_s=steps / dt, ms_per_step=1000.0 * dt / steps,
                        cg_iters=int(st.outer_iters), cg_iters_per_s=st.outer_iters / dt,
                        kernels=kernel_summary(kt, N), roofline=roofline(kt))
        log(f"[bench] c2 {key}: {steps / dt:.2f} steps/s, {st.outer_iters} CG iterations, "
            f"roofline {out[key]['roofline']}")
    for key in out:
        if key != "fp64@1e-32":
            out[key]["err_vs_fp64_reftol"] = err_summary(states[key], states["fp64@1e-32"])
    return out


def c2_reference(n: int, n1: int, n2: int) -> dict:
    """The reference program (serial and OpenMP) and the Fortran drop-in on the 256^3 two-phase
    workload's input files (a 16.8M-token system.in), timed by the two-run difference of n1 and
    n2 steps at the reference's tolerance 1e-32; the drop-in by the difference of 1 and 51
    steps, fastest of two runs each (its steps take milliseconds against ~14 s of set-up)."""
    import shutil
    import tempfile
    import oracle
    from heatflow_amd import workloads
    from heatflow_amd.system import write_inputs
    s, _ = workloads.c2(n)
    out = {}
    d = tempfile.mkdtemp(prefix="hf_c2_")
    try:
        t0 = time.perf_counter()
        write_inputs(s, os.path.join(d, "inputs"), ntime=n1)
        log(f"[bench] c2 inputs written in {time.perf_counter() - t0:.1f} s")
        for label, exe, thr, a, b, rep in (
                ("dropin_gpu", DROPIN_EXE, 1, 1, 51, 2),
                ("reference_openmp", oracle.REF_OMP_EXE, _host_threads(), n1, n2, 1),
                ("reference_serial", oracle.REF_EXE, 1, n1, n2, 1)):
            if os.path.exists(exe):
                out[label] = program_rate(s, exe, d, a, b, threads=thr, inputs_ready=True,
                                          repeats=rep)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def per_cell_rc_bench(s, steps: int, warmup: int, tol: float) -> dict:
    """The FP32-IR step on the same grid with rho varied by up to 0.1% per cell.  The workload's
    rho c_p and tau' rho c_p are uniform, which the step kernels use (one value read for every
    cell, 16 B/cell less); this leg times the per-cell reads of a general system."""
    from heatflow_amd import Simulation
    c = s.cells
    rho0 = c.rho
    rng = np.random.default_rng(7)
    c.rho = rho0 * (1.0 + 1.0e-3 * rng.uniform(-1.0, 1.0, rho0.shape[0]))
    try:
        sim = Simulation(s, precision="fp32_ir", tol=tol)
        try:
            dt, st, kt, _ = timed_run(sim, steps, warmup)
        finally:
            sim.close()
    finally:
        c.rho = rho0
    ks = kernel_summary(kt, s.params.n)
    out = dict(timesteps_per_s=steps / dt, ms_per_step=1000.0 * dt / steps,
               cg_iters=int(st.inner_iters), k_ir_step=ks.get("k_ir_step"))
    log(f"[bench] per-cell rho c_p fp32_ir@{tol:g}: {steps / dt:.2f} steps/s, "
        f"k_ir_step {out['k_ir_step']}")
    return out


def stiff_bench(s, steps: int, tol: float, dt_factor: float = 1000.0) -> dict:
    """The 512^3 workload at a 1000x longer time step (5e-8 s): about 60 CG iterations per step
    instead of one, so the solve, not the per-step passes, dominates.  Both precisions at
    `tol`, errors against FP64 at the reference tolerance."""
    import copy
    from heatflow_amd import Simulation
    st_sys = copy.copy(s)
    st_sys.params = copy.copy(s.params)
    st_sys.params.time_step = s.params.time_step * dt_factor
    out = dict(time_step=st_sys.params.time_step, steps=steps, tol=tol)
    states = {}
    for prec, t in (("fp32_ir", tol), ("fp64", tol), ("fp64", 1e-32)):
        key = f"{prec}@{t:g}"
        sim = Simulation(st_sys, precision=prec, tol=t)
        try:
            dt, st, kt, tp = timed_run(sim, steps, 1)
        finally:
            sim.close()
        states[key] = tp
        iters = st.inner_iters if prec == "fp32_ir" else st.outer_iters
        out[key] = dict(timesteps_per_s=steps / dt, cg_iters=int(iters),
                        refine_sweeps=int(st.outer_iters) if prec == "fp32_ir" else None,
                        cg_iters_per_s=iters / dt, kernels=kernel_summary(kt, st_sys.params.n))
        log(f"[bench] stiff {key}: {steps / dt:.2f} steps/s, {iters} CG iterations")
    ref = states["fp64@1e-32"]
    for key in (f"fp32_ir@{tol:g}", f"fp64@{tol:g}"):
        mx, rms = rel_errors(states[key], ref)
        out[key]["err_vs_fp64_reftol"] = dict(max=mx, rms=rms)
    out["fp32_ir_over_fp64"] = (out[f"fp32_ir@{tol:g}"]["timesteps_per_s"]
                                / out[f"fp64@{tol:g}"]["timesteps_per_s"])
    return out


def grad_bench(n: int, steps: int, ckpt: int, fd_cells: int = 16) -> dict:
    """Forward+adjoint gradient of the 128^3 / 200-step workload (Fourier and Cattaneo, FP64):
    loss = sum over the probes of (T - T*)^2 at the last step, T* from the same run with the
    seeded ground-truth conductivity kappa*.  Timed: one hf_grad call (forward with checkpoints
    every `ckpt` steps, reverse sweep with recomputation) on the 300 K system.

    Checked against central finite differences (relative step 1e-6) of the GPU forward:
      - kappa on `fd_cells` random cells (seed 43) within one cell of a probe;
      - kappa along a seeded direction, kappa (1 + eps u), u ~ U(-1, 1), against
        sum_i g_i kappa_i u_i;
      - 4 heater powers.
    T - T* is ~1e-5 K while FP64 resolves T ~ 300 K only to 6e-14 K, so a difference of the
    loss at 300 K is round-off bound at ~1e-3 relative.  The differences are therefore taken on
    the same system with temperatures measured from the bath (baths and initial state 0 K
    instead of 300 K; the operator is linear, so the field is the same minus 300 K), where the
    loss is resolved to ~1e-17 K.  The 300 K single-cell differences are reported beside them.
    """
    from heatflow_amd import Simulation, workloads
    out = {}
    for icat, label in ((1, "cattaneo"), (0, "fourier")):
        s, m = workloads.c4(n, icattaneo=icat)
        s0, _ = workloads.c4(n, icattaneo=icat, t_ref=0.0)
        kap = s.cells.kappa.copy()
        vol = s.cells.volume
        probes = m["probes"]
        sims = []

        def final_t(sim, kv, pv):
            sim.set_fields(kv, pv / vol)
            sim.sparse_Hmatrix()
            sim.initial_evolve()
            sim.run(1, steps)
            return sim.get_state()[0]

        try:
            sim = Simulation(s, tol=1e-32)
            sims.append(sim)
            targets = final_t(sim, m["kstar"], m["power"])[probes]
            sim.set_fields(kap, m["power"] / vol)
            sim.sparse_Hmatrix()
            sim.initial_evolve()
            sync()
            t0 = time.perf_counter()
            loss, gk, gp = sim.grad(1, steps, ckpt, probes, targets)
            sync()
            secs = time.perf_counter() - t0
            res = dict(seconds=round(secs, 3), loss=loss, targets="T(kappa*) at the probes",
                       grad_kappa_norm=float(np.linalg.norm(gk)),
                       grad_power_norm=float(np.linalg.norm(gp[m["heaters"]])))
            if fd_cells:
                sim0 = Simulation(s0, tol=1e-32)
                sims.append(sim0)
                targets0 = final_t(sim0, m["kstar"], m["power"])[probes]

                def loss_at(sm, tg, kv, pv):
                    return float(np.sum((final_t(sm, kv, pv)[probes] - tg) ** 2))

                def cd(sm, tg, kp, km, pp, pm, h):
                    return (loss_at(sm, tg, kp, pp) - loss_at(sm, tg, km, pm)) / (2 * h)

                def stats(errs):
                    return dict(max_rel_err=float(max(errs)),
                                median_rel_err=float(np.median(errs)))

                def stats_g(fds, gs):
                    fds, gs = np.asarray(fds), np.asarray(gs)
                    gmax = float(np.abs(gs).max())
                    rel = np.abs(fds - gs) / np.abs(gs)
                    big = np.abs(gs) >= 1e-3 * gmax
                    return {"max_abs_err_over_max_g": float(np.abs(fds - gs).max()) / gmax,
                            "max_rel_err_cells_ge_1e-3_max": float(rel[big].max()),
                            "cells_ge_1e-3_max": int(big.sum()),
                            "max_rel_err_all": float(rel.max()),
                            "median_rel_err_all": float(np.median(rel)),
                            "g_range": [float(np.abs(gs).min()), gmax]}

                # `fd_cells` cells drawn with seed 43 from those within one cell of a probe
                # (the support of dL/dkappa)
                near = workloads.near_cells(n, np.asarray(probes), 1)
                cells = np.sort(np.random.default_rng(43).choice(near, size=fd_cells,
                                                                 replace=False))
                f0, f300 = [], []
                for c in cells:
                    h = 1e-6 * kap[c]
                    kp, km = kap.copy(), kap.copy()
                    kp[c] += h
                    km[c] -= h
                    pw = m["power"]
                    f0.append(cd(sim0, targets0, kp, km, pw, pw, h))
                    f300.append(cd(sim, targets, kp, km, pw, pw, h))
                res["fd_kappa_cells"] = {"cells": len(cells),
                                         "choice": "random (seed 43) within one cell of a probe",
                                         "rel_step": 1e-6,
                                         "bath_origin": stats_g(f0, gk[cells]),
                                         "at_300K": stats_g(f300, gk[cells])}
                u = np.random.default_rng(44).uniform(-1.0, 1.0, kap.size)
                dirder = float(np.dot(gk, kap * u))
                eps = 1e-6
                fd = cd(sim0, targets0, kap * (1 + eps * u), kap * (1 - eps * u),
                        m["power"], m["power"], eps)
                res["fd_kappa_directional"] = {"rel_step": eps, "adjoint": dirder, "fd": fd,
                                               "rel_err": abs(fd - dirder) / abs(dirder)}
                ep = []
                for hc in m["heaters"][:4]:
                    hp = 1e-6 * m["power"][hc]
                    pp, pm = m["power"].copy(), m["power"].copy()
                    pp[hc] += hp
                    pm[hc] -= hp
                    ep.append(abs(cd(sim0, targets0, kap, kap, pp, pm, hp) - gp[hc]) / abs(gp[hc]))
                res["fd_power"] = dict(heaters=len(ep), rel_step=1e-6, **stats(ep))
            out[label] = res
            log(f"[bench] grad {n}^3 {label}: {res}")
        finally:
            for sm in sims:
                sm.close()
    return out


def grad_c5_bench(n: int, steps: int) -> dict:
    """The 384^3 layered workload: forward + adjoint with sqrt(steps) checkpoints, FP32-inner /
    FP64-refined and plain FP64 solves both stopping at a scaled residual of 1e-13 (and the
    FP32-IR run once more with the adjoint solves at 1e-10, and once more with every state kept
    when they fit in device memory: checkpoint spacing 0, nothing recomputed), each against the
    FP64 gradient at the reference tolerance 1e-32; device memory held by each context."""
    import torch
    from heatflow_amd import Simulation, workloads
    s, m = workloads.c5(n)
    ckpt = int(np.ceil(np.sqrt(steps)))
    targets = np.full(len(m["probes"]), 300.0)
    res = {}
    grads = {}
    every_bytes = (steps + 1) * s.params.n * 8
    for prec, tol, atol, k in (("fp64", 1e-32, 0.0, ckpt), ("fp32_ir", 1e-13, 0.0, ckpt),
                               ("fp64", 1e-13, 0.0, ckpt), ("fp32_ir", 1e-13, 1e-10, ckpt),
                               ("fp32_ir", 1e-13, 1e-10, 0)):
        key = f"{prec}@{tol:g}" + (f"/adjoint@{atol:g}" if atol else "") + \
            ("/every_state" if k == 0 else "")
        torch.cuda.synchronize()
        free0, total = torch.cuda.mem_get_info()
        sim = Simulation(s, precision=prec, tol=tol)
        try:
            sim.set_adjoint_tol(atol)
            sim.initial_evolve()
            sync()
            t0 = time.perf_counter()
            loss, gk, gp = sim.grad(1, steps, k, m["probes"], targets)
            sync()
            secs = time.perf_counter() - t0
            free1, _ = torch.cuda.mem_get_info()
        finally:
            sim.close()
        grads[key] = (loss, gk, gp[m["heaters"]])
        res[key] = dict(seconds=round(secs, 3), loss=loss, device_bytes=int(free0 - free1))
        if k == 0:
            # spacing 0 keeps every state (about (steps + 1) N doubles) when they fit
            res[key]["kept_every_state"] = bool(free0 - free1 >= every_bytes)
        log(f"[bench] grad c5 {n}^3 {key}: {res[key]}")
    L0, gk0, gp0 = grads["fp64@1e-32"]
    for key in ("fp32_ir@1e-13", "fp64@1e-13", "fp32_ir@1e-13/adjoint@1e-10",
                "fp32_ir@1e-13/adjoint@1e-10/every_state"):
        L, gk, gp = grads[key]
        res[key]["err_vs_fp64_reftol"] = dict(
            loss=abs(L - L0) / abs(L0),
            grad_kappa=float(np.abs(gk - gk0).max() / np.abs(gk0).max()),
            grad_power=float(np.abs(gp - gp0).max() / np.abs(gp0).max()))
    res.update(cells=s.params.n, steps=steps, checkpoint_every=ckpt,
               probes=len(m["probes"]), inclusion_fraction=round(m["inclusion_fraction"], 4))
    return res


def tol_sweep(n: int, steps: int) -> dict:
    """Accuracy against solver tolerance on the layered 384^3 workload: the forward (`steps`
    steps) and the forward+adjoint gradient in FP64 and FP32-IR at scaled-residual tolerances
    1e-12 .. 1e-15, each against FP64 at the reference tolerance 1e-32."""
    from heatflow_amd import Simulation, workloads
    s, m = workloads.c5(n)
    ckpt = int(np.ceil(np.sqrt(steps)))
    targets = np.full(len(m["probes"]), 300.0)
    h = m["heaters"]

    def fwd(prec, tol):
        with Simulation(s, precision=prec, tol=tol) as sim:
            sim.initial_evolve()
            sync()
            t0 = time.perf_counter()
            it = sim.run(1, steps)
            sync()
            dt = time.perf_counter() - t0
            return sim.get_state()[0], dt, it

    def grad(prec, tol):
        with Simulation(s, precision=prec, tol=tol) as sim:
            sim.initial_evolve()
            sync()
            t0 = time.perf_counter()
            L, gk, gp = sim.grad(1, steps, ckpt, m["probes"], targets)
            sync()
            return L, gk, gp[h], time.perf_counter() - t0

    ref, dt, it = fwd("fp64", 1e-32)
    out = dict(cells=s.params.n, steps=steps, checkpoint_every=ckpt,
               forward={"fp64@1e-32": dict(seconds=round(dt, 3), iterations=it)}, gradient={})
    for prec in ("fp64", "fp32_ir"):
        for tol in (1e-12, 1e-13, 1e-14, 1e-15):
            tp, dt, it = fwd(prec, tol)
            mx, rms = rel_errors(tp, ref)
            out["forward"][f"{prec}@{tol:g}"] = dict(seconds=round(dt, 3), iterations=it,
                                                     max_rel_err=mx, rms_rel_err=rms)
            log(f"[sweep] forward {prec}@{tol:g}: {out['forward'][f'{prec}@{tol:g}']}")
    Lr, gkr, gpr, dt = grad("fp64", 1e-32)
    out["gradient"]["fp64@1e-32"] = dict(seconds=round(dt, 3), loss=Lr)
    for prec in ("fp64", "fp32_ir"):
        for tol in (1e-13, 1e-14, 1e-15):
            L, gk, gp, dt = grad(prec, tol)
            out["gradient"][f"{prec}@{tol:g}"] = dict(
                seconds=round(dt, 3), loss_rel_err=abs(L - Lr) / abs(Lr),
                grad_kappa_err=float(np.abs(gk - gkr).max() / np.abs(gkr).max()),
                grad_power_err=float(np.abs(gp - gpr).max() / np.abs(gpr).max()))
            log(f"[sweep] gradient {prec}@{tol:g}: {out['gradient'][f'{prec}@{tol:g}']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-ref", action="store_true", help="skip the FP64 1e-32 runs")
    ap.add_argument("--only", choices=["fp32_ir", "fp64", "fp32_ir_noreuse"], default=None,
                    help="time one mode only (profiling)")
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--grad-n", type=int, default=128)
    ap.add_argument("--c5-n", type=int, default=384)
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--no-stiff", action="store_true")
    ap.add_argument("--stiff-steps", type=int, default=3)
    ap.add_argument("--tol-sweep", action="store_true",
                    help="only run the 384^3 accuracy-vs-tolerance sweep and print it")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 2 GiB device copy (PMC byte-counter calibration)")
    ap.add_argument("--no-c1", action="store_true", help="skip the 16^3 leg")
    ap.add_argument("--no-c2", action="store_true", help="skip the 256^3 two-phase leg")
    ap.add_argument("--c2-n", type=int, default=256)
    ap.add_argument("--c1-only", action="store_true",
                    help="only run the 16^3 leg (GPU paths) and print it (profiling)")
    ap.add_argument("--ref-c2", action="store_true",
                    help="only time the reference program and the drop-in on the 256^3 "
                         "two-phase input files (minutes of CPU) and print that")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU z-slab decomposition is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    from heatflow_amd import Simulation, workloads

    if args.tol_sweep:
        print(json.dumps({"tol_sweep": tol_sweep(args.c5_n, 500)}), flush=True)
        return
    if args.ref_c2:
        print(json.dumps({"c2_reference": c2_reference(args.c2_n, 1, 2)}), flush=True)
        return
    if args.c1_only:
        print(json.dumps({"c1_16": c1_bench(100, 5, cpu=False)}), flush=True)
        return

    if args.copy_probe:
        # known traffic: 2 GiB read + 2 GiB written per copy
        src = torch.empty(1 << 28, dtype=torch.float64, device="cuda")
        src.fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(5):
            dst.copy_(src)
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / 5
        log(f"[bench] copy probe: 2 GiB copy {ms:.3f} ms = {2 * (1 << 31) / (ms * 1e-3) / 1e9:.0f} GB/s")
        del src, dst

    t_gen = time.perf_counter()
    s, meta = workloads.c3(args.n)
    log(f"[bench] workload c3 {args.n}^3 generated in {time.perf_counter() - t_gen:.1f} s")
    N = s.params.n
    K, W = args.steps, args.warmup

    results = {}
    # (precision, tolerance, debug flags, label): debug bit 2 (value 4) turns off the reuse of
    # the inner step length, so that leg compares the precisions with every step solving
    modes = [("fp32_ir", args.tol, 0, "fp32_ir"), ("fp64", args.tol, 0, "fp64")]
    if args.only == "fp32_ir_noreuse":
        modes = [("fp32_ir", args.tol, 4, args.only)]
    elif args.only:
        modes = [(args.only, args.tol, 0, args.only)]
    else:
        modes.append(("fp32_ir", args.tol, 4, "fp32_ir_noreuse"))
        if not args.no_ref:
            modes.append(("fp64", 1e-32, 0, "fp64"))
    states = {}
    for prec, tol, dbg, label in modes:
        key = f"{label}@{tol:g}"
        t_setup = time.perf_counter()
        sim = Simulation(s, precision=prec, tol=tol)
        try:
            setup = time.perf_counter() - t_setup
            sim.set_debug(dbg)
            dt, st, kt, tp = timed_run(sim, K, W)
        finally:
            sim.close()
        states[key] = tp
        iters = st.inner_iters if prec == "fp32_ir" else st.outer_iters
        results[key] = dict(
            timesteps_per_s=K / dt, ms_per_step=1000.0 * dt / K,
            cg_iters=int(iters), refine_sweeps=int(st.outer_iters) if prec == "fp32_ir" else None,
            cg_iters_per_s=iters / dt, setup_s=round(setup, 2),
            kernels=kernel_summary(kt, N), roofline=roofline(kt), last_err=st.last_err)
        log(f"[bench] {key}: {K / dt:.2f} steps/s, {1000 * dt / K:.2f} ms/step, "
            f"{iters} CG iterations, roofline {results[key]['roofline']}")

    ref_key = "fp64@1e-32"
    if ref_key in states:
        for key in results:
            if key != ref_key:
                results[key]["err_vs_fp64_reftol"] = err_summary(states[key], states[ref_key])
    del states

    head = results.get(f"fp32_ir@{args.tol:g}") or next(iter(results.values()))
    fp64 = results.get(f"fp64@{args.tol:g}")
    percell = None if args.only else per_cell_rc_bench(s, K, W, args.tol)
    stiff = None if (args.no_stiff or args.only) else stiff_bench(s, args.stiff_steps, args.tol)
    grad = None if (args.no_grad or args.only) else grad_bench(args.grad_n, 200, 20, fd_cells=16)
    grad5 = None if (args.no_grad or args.no_c5 or args.only) else grad_c5_bench(args.c5_n, 500)
    cpu = None if (args.no_cpu or args.only) else cpu_baseline(args.n, args.tol, args.cpu_n,
                                                                args.cpu_steps)
    c1 = None if (args.no_c1 or args.only) else c1_bench(100, 5, cpu=not args.no_cpu)
    c2 = None if (args.no_c2 or args.only) else c2_bench(args.c2_n, 50, 2)
    for r in results.values():
        r["roofline"]["traffic"] = measured_traffic(r["roofline"]["kernel"], args.n)
    line = {
        "metric": "timesteps/s at 512^3 (FP32-refined vs FP64, 1e-10 rel); fwd+adjoint grad time",
        "value": round(head["timesteps_per_s"], 4),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.only == "fp64" else "f32-inner/f64-refined",
        "data": "synthetic",
        "config": {"workload": f"c3: {args.n}^3 Cattaneo, log-normal kappa exp(N(ln50,1)) seed 3, "
                               f"Gaussian heating sigma=32 cells peak 1e12 W/m^3, Dirichlet 300 K "
                               f"on 6 faces, dt=5e-11, {K} steps, CG tol {args.tol:g}",
                   "cells": N,
                   "uniform_rho_cp": True},
        "fp32_ir_over_fp64": (round(head["timesteps_per_s"] / fp64["timesteps_per_s"], 3)
                              if fp64 and head is not fp64 else None),
        "fp32_ir_noreuse_over_fp64": (
            round(results[f"fp32_ir_noreuse@{args.tol:g}"]["timesteps_per_s"]
                  / fp64["timesteps_per_s"], 3)
            if fp64 and f"fp32_ir_noreuse@{args.tol:g}" in results else None),
        "modes": results,
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
        "grad_128": grad,
        "per_cell_rho_cp": percell,
        "stiff_dt_x1000": stiff,
        "grad_c5_384": grad5,
        "c1_16": c1,
        "c2_256": c2,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

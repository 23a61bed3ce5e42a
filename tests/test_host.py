"""Host-side checks without a GPU: the C-ABI library loads and exports every declared entry
point, argument validation, the input-file reader and the workload builders."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from heatflow_amd import _lib, workloads
from heatflow_amd.system import read_inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "heatflow_hip.h")) as f:
        src = f.read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hf_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 19
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.EXPORTS)
    assert lib.hf_version().decode().startswith("heatflow_hip")


def test_params_struct_layout_and_defaults(tmp_path):
    """The ctypes mirror of hf_params / hf_stats has the C layout (offsets from a C probe
    compiled against include/heatflow_hip.h), and hf_params_default gives the reference's
    defaults and hard-coded solver controls (mod_evolve.f90:142-144, mod_inputs.f90:213-270)."""
    lib = _lib.load()
    fields = [f[0] for f in _lib.HfParams._fields_]
    sfields = [f[0] for f in _lib.HfStats._fields_]
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "heatflow_hip.h"', "int main(void){",
           'printf("%zu %zu\\n", sizeof(hf_params), sizeof(hf_stats));']
    src += [f'printf("%zu\\n", offsetof(hf_params, {f}));' for f in fields]
    src += [f'printf("%zu\\n", offsetof(hf_stats, {f}));' for f in sfields]
    src += ["return 0;}"]
    (tmp_path / "probe.c").write_text("\n".join(src))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    str(tmp_path / "probe.c")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    vals = [int(v) for v in out]
    assert vals[0] == C.sizeof(_lib.HfParams) and vals[1] == C.sizeof(_lib.HfStats)
    assert vals[2:2 + len(fields)] == [getattr(_lib.HfParams, f).offset for f in fields]
    assert vals[2 + len(fields):] == [getattr(_lib.HfStats, f).offset for f in sfields]
    p = _lib.HfParams()
    lib.hf_params_default(C.byref(p))
    assert p.tol == 1e-32 and p.itmax == 50000
    assert p.icattaneo == 1 and p.time_step == 1.0 and p.br == 1.0 and p.freq == 1.0


@pytest.mark.parametrize("dims,periodic", [((1, 4, 4), (0, 0, 0)), ((4, 2, 4), (0, 1, 0)),
                                           ((4, 4, 1), (0, 0, 1)), ((0, 4, 4), (0, 0, 0))])
def test_create_rejects_geometry_the_reference_mishandles(dims, periodic):
    lib = _lib.load()
    p = _lib.HfParams()
    lib.hf_params_default(C.byref(p))
    p.nx, p.ny, p.nz = dims
    for a in range(3):
        p.periodic[a] = periodic[a]
    ctx = C.c_void_p()
    rc = lib.hf_create(C.byref(ctx), C.byref(p))
    assert rc == _lib.HF_ERR_ARG and not ctx.value
    assert lib.hf_last_error(None)


def test_read_inputs_golden(golden_dir):
    s = read_inputs(golden_dir)
    p, c = s.params, s.cells
    assert (p.nx, p.ny, p.nz) == (25, 25, 1)
    assert p.time_step == 1.0 and p.icattaneo == 0 and p.isteady == 0 and p.ntime == 10
    assert p.t_bath == (283.0, 278.0, 278.0, 278.0, 278.0, 278.0)
    assert p.kappa_bound == (0.541,) * 6
    assert np.all(c.kappa == 0.541) and np.all(c.rho == 997.0) and np.all(c.heat_capacity == 4200.0)
    assert np.allclose(c.len_x, 0.002 / 25) and np.allclose(c.len_z, 0.002)
    assert c.iheater.sum() == 0


def test_workload_builders_small():
    s, m = workloads.c1(8)
    assert s.params.n == 512 and s.cells.iheater.sum() == 1
    s, m = workloads.c2(16, nheat=5)
    assert s.cells.iheater.sum() == 5 and set(np.unique(s.cells.kappa)) <= {0.2, 150.0}
    s, m = workloads.c3(16)
    q = s.cells.qdens_extra
    assert q.shape == (16 ** 3,) and q.max() <= 1e12 and q.argmax() > 0
    s, m = workloads.c4(12, nheat=3, nprobe=7)
    assert len(m["probes"]) == 7 and np.count_nonzero(m["power"]) == 3
    assert np.all(np.abs(s.cells.kappa / 100.0 - 1.0) <= 0.1)


def test_builtin_material_table_matches_oracle():
    """hf_builtin_material (a host table, no device) against the oracle's material() over
    every built-in id: the same single-precision roundings, tau assigned by the same ids, em by
    none; unknown ids are HF_ERR_ARG."""
    import oracle
    lib = _lib.load()
    ids = [100, 140, 3001, 3002, 3003, 3004, 9001, 9002, 9003, 310, 320, 330, 340, 350]
    for mid in ids:
        m, st, se = _lib.HfMaterial(), C.c_int32(), C.c_int32()
        assert lib.hf_builtin_material(mid, C.byref(m), C.byref(st), C.byref(se)) == _lib.HF_OK
        # a file material sets tau = 3e-9 and em = 0.5 first: what the built-in leaves stays
        fm = {7: dict(kappa=1.0, rho=1.0, heat_capacity=1.0, tau=3e-9, em=0.5)}
        k, r, h, t, e = oracle.materials([7, mid], fm)
        assert (m.index, m.kappa, m.rho, m.heat_capacity) == (mid, k[1], r[1], h[1])
        assert bool(st.value) == (t[1] != 3e-9) and se.value == 0 and e[1] == 0.5
        if st.value:
            assert m.tau == t[1]
    m = _lib.HfMaterial()
    for bad in (0, 1, 99, 141, 3005):
        assert lib.hf_builtin_material(bad, C.byref(m), None, None) == _lib.HF_ERR_ARG


def test_geometry_text_and_binary_forms(tmp_path, golden_dir):
    """system.in read as read_system reads it (mod_inputs.f90:602-649) and the binary .npz
    form: the same ids, heater flags and cell lengths."""
    from heatflow_amd.system import Geometry, read_system
    g = read_system(os.path.join(golden_dir, "system.in"))
    assert (g.nx, g.ny, g.nz) == (25, 25, 1)
    assert np.all(g.imaterial_type == 1) and g.iheater.sum() == 0
    assert g.cell_lengths == (0.002 / 25.0, 0.002 / 25.0, 0.002)
    rng = np.random.default_rng(3)
    nx, ny, nz = 7, 5, 3
    ids = rng.choice([1, 140, 320], size=nx * ny * nz)
    heat = rng.integers(0, 2, size=nx * ny * nz)
    lines = [f"{nx} {ny} {nz}", "0.7 0.5 0.3"]
    for iz in range(nz):
        lines.append("")
        for iy in range(ny):
            row = range((iz * ny + iy) * nx, (iz * ny + iy + 1) * nx)
            lines.append(" ".join(f"{ids[i]}:{heat[i]}" for i in row))
    (tmp_path / "system.in").write_text("\n".join(lines) + "\n")
    g = read_system(str(tmp_path / "system.in"))
    assert np.array_equal(g.imaterial_type, ids) and np.array_equal(g.iheater, heat)
    g.save(str(tmp_path / "geo.npz"))
    h = Geometry.load(str(tmp_path / "geo.npz"))
    assert (h.nx, h.ny, h.nz, tuple(h.lengths)) == (nx, ny, nz, (0.7, 0.5, 0.3))
    assert np.array_equal(h.imaterial_type, ids) and np.array_equal(h.iheater, heat)
    assert h.cell_lengths == (0.7 / 7.0, 0.5 / 5.0, 0.3 / 3.0)

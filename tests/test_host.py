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

"""Host-side description of a HeatFlow system: the `inputs` module state the time step reads.

`Params` holds the scalar keys of param.in that the path uses (src/heatflow/mod_inputs.f90:
204-360, defaults :213-270, the check_param fix-ups :366-470) and `Cells` the per-cell fields of
grid(:,:,:) (src/heatflow/mod_constructions.f90:25-31) as structure-of-arrays in the
reference's linear order (x fastest).  `read_inputs` reads the reference's own input files
(param.in, mat.in, system.in) for file materials (index <= 99).
"""
from __future__ import annotations

import dataclasses
import os
import re
from typing import Optional, Sequence

import numpy as np

FACES = ("x1", "xN", "y1", "yN", "z1", "zN")


@dataclasses.dataclass
class Params:
    nx: int
    ny: int
    nz: int
    time_step: float = 1.0          # mod_inputs.f90:226
    icattaneo: int = 1              # :229
    isteady: int = 0
    periodic: Sequence[bool] = (False, False, False)
    kappa_bound: Sequence[float] = (0.0,) * 6   # kappaBoundx1, Nx, y1, Ny, z1, Nz
    t_bath: Sequence[float] = (0.0,) * 6        # T_Bathx1, x2, y1, y2, z1, z2
    br: float = 1.0                 # BR
    cg: Sequence[bool] = (False,) * 6           # CG_dir: -x, +x, -y, +y, -z, +z
    t_bath_cg: float = 0.0
    power_in: float = 0.0
    freq: float = 1.0
    heated_steps: int = 0           # heattime
    t_bath_rad: float = 0.0         # T_Bath (radiative term)
    t_system: float = 300.0         # T_System (initial_evolve)
    ntime: int = 10

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def inverse_time(self) -> float:
        return 1.0 / self.time_step


@dataclasses.dataclass
class Cells:
    kappa: np.ndarray
    rho: np.ndarray
    heat_capacity: np.ndarray
    tau: np.ndarray                 # physical relaxation time (s)
    em: np.ndarray
    len_x: np.ndarray
    len_y: np.ndarray
    len_z: np.ndarray
    iheater: np.ndarray             # int64
    qdens_extra: Optional[np.ndarray] = None   # superset: W/m^3 added to Qdens

    @property
    def volume(self) -> np.ndarray:
        # grid%volume = Length(1)*Length(2)*Length(3) (mod_inputs.f90:623)
        return self.len_x * self.len_y * self.len_z

    def tau_scaled(self, time_step: float) -> np.ndarray:
        # grid%tau = tau*inverse_time*inverse_time (mod_setup.f90:62)
        it = 1.0 / time_step
        return self.tau * it * it

    def lin_rhoc(self) -> np.ndarray:
        return self.rho * self.heat_capacity


@dataclasses.dataclass
class System:
    params: Params
    cells: Cells

    @staticmethod
    def uniform(params: Params, lengths, kappa, rho, heat_capacity, tau=0.0, em=0.0,
                iheater=None, qdens_extra=None) -> "System":
        """Builds a system with per-axis uniform cell lengths (Lx/nx, Ly/ny, Lz/nz as
        read_system does, mod_inputs.f90:619-622); scalars broadcast to every cell."""
        n = params.n
        Lx, Ly, Lz = lengths

        def full(v, dtype=np.float64):
            a = np.asarray(v, dtype=dtype)
            return np.broadcast_to(a, (n,)).copy() if a.ndim == 0 else a.reshape(n).astype(dtype)

        cells = Cells(
            kappa=full(kappa), rho=full(rho), heat_capacity=full(heat_capacity),
            tau=full(tau), em=full(em),
            len_x=full(Lx / float(params.nx)), len_y=full(Ly / float(params.ny)),
            len_z=full(Lz / float(params.nz)),
            iheater=full(0 if iheater is None else iheater, np.int64),
            qdens_extra=None if qdens_extra is None else full(qdens_extra),
        )
        return System(params, cells)


# ---- reference input files -----------------------------------------------------------------

def _kv_lines(path: str):
    with open(path) as f:
        for raw in f:
            line = raw.split("!", 1)[0]
            if not line.strip() or "=" not in line:
                continue
            key, val = line[: line.index("=")], line[line.rindex("=") + 1:]
            yield key.strip(), val.strip()


def _logical(v: str) -> bool:
    # assignL: 'T'/'t' anywhere -> .TRUE., then 'F'/'f' anywhere -> .FALSE.
    out = False
    if "T" in v or "t" in v:
        out = True
    if "F" in v or "f" in v:
        out = False
    return out


@dataclasses.dataclass
class Geometry:
    """What system.in holds (read_system, src/heatflow/mod_inputs.f90:602-649): the grid,
    the box lengths and, per cell (x fastest), grid%imaterial_type and grid%iheater.  `save` and
    `load` keep it as a binary .npz file: a 512^3 grid is 134M "id:heater" tokens as text."""
    nx: int
    ny: int
    nz: int
    lengths: Sequence[float]          # Lx, Ly, Lz of the box
    imaterial_type: np.ndarray        # int64, N
    iheater: np.ndarray               # int64, N

    @property
    def cell_lengths(self):
        """grid%Length(1:3) = Lx/nx, Ly/ny, Lz/nz (mod_inputs.f90:619-621)."""
        Lx, Ly, Lz = self.lengths
        return Lx / float(self.nx), Ly / float(self.ny), Lz / float(self.nz)

    def save(self, path: str) -> None:
        np.savez(path, dims=np.array([self.nx, self.ny, self.nz], np.int64),
                 lengths=np.asarray(self.lengths, np.float64),
                 imaterial_type=np.ascontiguousarray(self.imaterial_type, np.int64),
                 iheater=np.ascontiguousarray(self.iheater, np.int64))

    @staticmethod
    def load(path: str) -> "Geometry":
        with np.load(path, allow_pickle=False) as d:
            nx, ny, nz = (int(v) for v in d["dims"])
            return Geometry(nx, ny, nz, tuple(float(v) for v in d["lengths"]),
                            d["imaterial_type"].astype(np.int64), d["iheater"].astype(np.int64))


def read_system(path: str) -> Geometry:
    """system.in as read_system reads it (mod_inputs.f90:602-649)."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines()]
    pos = 0

    def nextline():
        nonlocal pos
        ln = lines[pos]
        pos += 1
        return ln

    nx, ny, nz = (int(v) for v in nextline().split()[:3])
    Lx, Ly, Lz = (float(v.replace("d", "e").replace("D", "e")) for v in nextline().split()[:3])
    mat = np.zeros((nz, ny, nx), np.int64)
    heat = np.zeros((nz, ny, nx), np.int64)
    for iz in range(nz):
        nextline()  # separator line before each z block
        for iy in range(ny):
            toks = nextline().split()[:nx]
            for ix, t in enumerate(toks):
                m, h = t.split(":")
                mat[iz, iy, ix] = int(m)
                heat[iz, iy, ix] = int(h)
    return Geometry(nx, ny, nz, (Lx, Ly, Lz), mat.reshape(-1), heat.reshape(-1))


def read_params(path: str, nx: int, ny: int, nz: int) -> Params:
    """param.in as read_param and check_param read it (mod_inputs.f90:204-470), for the keys
    the time step uses."""
    kv = {}
    for k, v in _kv_lines(path):
        kv[k] = v
    D = lambda k, d: float(kv[k].replace("d", "e").replace("D", "e")) if k in kv else d
    I = lambda k, d: int(kv[k]) if k in kv else d
    t_bath = D("T_Bath", 0.0)
    baths_keys = ("T_Bathx1", "T_Bathx2", "T_Bathy1", "T_Bathy2", "T_Bathz1", "T_Bathz2")
    baths = [D(k, t_bath) for k in baths_keys]
    tcg = D("T_BathCG", 0.0)
    # check_param (:436-460): if any face bath is missing all six take T_Bath
    if not ("T_BathCG" in kv and tcg > 0) and any(k not in kv for k in baths_keys):
        baths = [t_bath] * 6
    kb_keys = ("kappaBoundx1", "kappaBoundNx", "kappaBoundy1", "kappaBoundNy", "kappaBoundz1",
               "kappaBoundNz")
    kb = [D(k, 0.0) for k in kb_keys]
    periodic_s = kv.get("Periodic", "")
    periodic = tuple(c in periodic_s or c.upper() in periodic_s for c in "xyz")
    if not all(periodic) and "kappaBound" in kv and any(k not in kv for k in kb_keys):
        kb = [D("kappaBound", 0.0)] * 6   # :412-426
    cg_s = kv.get("CG_dir", "")
    cg = tuple(tok in cg_s for tok in ("-x", "+x", "-y", "+y", "-z", "+z"))
    params = Params(
        nx=nx, ny=ny, nz=nz, time_step=D("time_step", 1.0), icattaneo=I("icattaneo", 1),
        isteady=I("isteady", 0), periodic=periodic, kappa_bound=tuple(kb), t_bath=tuple(baths),
        br=D("BR", 1.0), cg=cg, t_bath_cg=tcg, power_in=D("power_in", 0.0),
        freq=D("freq", 1.0), heated_steps=I("heattime", 0), t_bath_rad=t_bath,
        t_system=D("T_System", 300.0), ntime=I("ntime", 10))
    return params


def read_mat(path: str) -> dict:
    """mat.in as read_mat reads it (mod_inputs.f90:654-721): {index: {key: value}}."""
    mats = {}
    cur = None
    with open(path) as f:
        for raw in f:
            line = raw.split("!", 1)[0].strip()
            if not line:
                continue
            if re.fullmatch(r"[+-]?\d+", line.split()[0]) and "=" not in line:
                idx = int(line.split()[0])
                if idx != 0:
                    cur = mats.setdefault(idx, {})
                continue
            if "=" in line and cur is not None:
                k, v = line.split("=", 1)
                cur[k.strip()] = float(v.strip().replace("d", "e").replace("D", "e"))
    return mats


def write_inputs(system: System, dirname: str, ntime: Optional[int] = None,
                 write_every: int = 10, **param_overrides) -> None:
    """Writes a System in the reference's input format: <dirname>/param.in, mat.in and
    system.in, as read_param, read_mat and read_system read them (mod_inputs.f90:204-360,
    602-721) -- the inverse of read_inputs.  Each distinct (kappa, rho, heat_capacity, tau, em)
    becomes a file material (at most 99); cell lengths must be uniform per axis and there is no
    per-cell power density (the reference cannot express either).  Output is off (IVERB 0,
    no text output, _Test_Run) unless overridden: param_overrides replace or add param.in keys."""
    p, c = system.params, system.cells
    n = p.n
    if c.qdens_extra is not None:
        raise ValueError("a per-cell power density has no reference input form")
    lens = []
    for a in (c.len_x, c.len_y, c.len_z):
        a = np.asarray(a, np.float64)
        if np.any(a != a[0]):
            raise ValueError("cell lengths must be uniform per axis")
        lens.append(float(a[0]))
    # materials: distinct rows of the five fields, found column by column (a row-wise unique
    # over 16.8M cells sorts 5-wide rows and takes minutes)
    cols = [np.broadcast_to(np.asarray(v, np.float64), (n,)) for v in
            (c.kappa, c.rho, c.heat_capacity, c.tau, c.em)]
    key = np.zeros(n, np.int64)
    vals = []
    for col in cols:
        u, inv = np.unique(col, return_inverse=True)
        key = key * len(u) + inv.reshape(-1)
        vals.append(u)
    ukey, mat_of = np.unique(key, return_inverse=True)
    uniq = np.empty((len(ukey), 5))
    rem = ukey.copy()
    for j in range(4, -1, -1):
        uniq[:, j] = vals[j][rem % len(vals[j])]
        rem //= len(vals[j])
    if len(uniq) > 99:
        raise ValueError(f"{len(uniq)} distinct materials: mat.in holds at most 99")
    os.makedirs(dirname, exist_ok=True)
    per = "".join(ax for ax, on in zip("xyz", p.periodic) if on)
    cg = " ".join(t for t, on in zip(("-x", "+x", "-y", "+y", "-z", "+z"), p.cg) if on)
    kv = {
        "IVERB": "0", "_WriteToTxt": ".False.", "_Test_Run": ".True.",
        "_Percentage_Completion": ".False.", "_RunName": "run",
        "ntime": str(p.ntime if ntime is None else ntime), "write_every": str(write_every),
        "heattime": str(p.heated_steps), "time_step": repr(p.time_step), "freq": repr(p.freq),
        "iboundary": "1", "icattaneo": str(p.icattaneo), "isteady": str(p.isteady),
        "T_System": repr(p.t_system), "T_Bath": repr(p.t_bath_rad),
        "power_in": repr(p.power_in), "BR": repr(p.br), "T_BathCG": repr(p.t_bath_cg),
        "CG_dir": f"'{cg}'" if cg else "none", "Periodic": f"'{per}'", "TempDepProp": "0",
        "_Check_Sparse_Full": ".False.", "_Check_Stability": ".False.",
        "_InputTempDis": ".False.", "_FullRestart": ".False.", "T_BathCC": ".False.",
        "start_ix": "1", "end_ix": str(p.nx), "start_iy": "1", "end_iy": str(p.ny),
        "start_iz": "1", "end_iz": str(p.nz), "kappaBound": "0",
    }
    for k, v in zip(("T_Bathx1", "T_Bathx2", "T_Bathy1", "T_Bathy2", "T_Bathz1", "T_Bathz2"),
                    p.t_bath):
        kv[k] = repr(float(v))
    for k, v in zip(("kappaBoundx1", "kappaBoundNx", "kappaBoundy1", "kappaBoundNy",
                     "kappaBoundz1", "kappaBoundNz"), p.kappa_bound):
        kv[k] = repr(float(v))
    kv.update({k: str(v) for k, v in param_overrides.items()})
    with open(os.path.join(dirname, "param.in"), "w") as f:
        f.writelines(f"    {k} = {v}\n" for k, v in kv.items())
    # each material block ends with a 0 line: read_mat checks and resets its key counts there
    # (mod_inputs.f90:675-684)
    with open(os.path.join(dirname, "mat.in"), "w") as f:
        for i, (k, r, h, t, e) in enumerate(uniq.tolist(), start=1):
            f.write(f"{i}\nheat_capacity = {h!r}\nh_conv = 0\nkappa = {k!r}\n"
                    f"kappa3D = {k!r}\nrho = {r!r}\nsound_speed = 0\ntau = {t!r}\n"
                    f"em = {e!r}\n0\n")
    heat = np.asarray(c.iheater, np.int64).reshape(-1)
    if heat.min() < 0:
        raise ValueError("heater modes are >= 0")
    hspan = int(heat.max()) + 1
    combos, tok_of = np.unique(mat_of.reshape(-1) * hspan + heat, return_inverse=True)
    # "m:h" tokens padded to one width (list-directed reads skip the blanks), one row of x per
    # line, a blank line before each z plane
    words = [f"{k // hspan + 1}:{k % hspan}" for k in combos.tolist()]
    w = max(len(t) for t in words) + 1
    table = np.frombuffer("".join(t.ljust(w) for t in words).encode(), np.uint8).reshape(-1, w)
    L = (lens[0] * p.nx, lens[1] * p.ny, lens[2] * p.nz)
    tok_of = tok_of.reshape(p.nz, p.ny * p.nx)
    with open(os.path.join(dirname, "system.in"), "wb") as f:
        f.write(f"{p.nx} {p.ny} {p.nz}\n{L[0]!r} {L[1]!r} {L[2]!r}\n".encode())
        for iz in range(p.nz):
            rows = table[tok_of[iz]].reshape(p.ny, p.nx * w)
            rows[:, -1] = ord("\n")
            f.write(b"\n")
            f.write(rows.tobytes())


def read_inputs(dirname: str) -> System:
    """Reads <dirname>/param.in, mat.in and system.in like read_all_files
    (src/heatflow/mod_inputs.f90:116-193) for file materials, then applies
    set_global_variables' per-cell material assignment (mod_setup.f90:53-68) on the host.
    (Simulation.from_geometry does that assignment on the device, built-in materials
    included.)"""
    geo = read_system(os.path.join(dirname, "system.in"))
    params = read_params(os.path.join(dirname, "param.in"), geo.nx, geo.ny, geo.nz)
    mats = read_mat(os.path.join(dirname, "mat.in"))
    n = geo.nx * geo.ny * geo.nz
    get = lambda key: np.array([mats[int(m)][key] for m in geo.imaterial_type], np.float64)
    lx, ly, lz = geo.cell_lengths
    cells = Cells(
        kappa=get("kappa"), rho=get("rho"), heat_capacity=get("heat_capacity"), tau=get("tau"),
        em=get("em"), len_x=np.full(n, lx), len_y=np.full(n, ly), len_z=np.full(n, lz),
        iheater=geo.iheater.copy())
    return System(params, cells)

"""Driver-side problem setup: what exfem's Fortran driver does before the step
loop, restated in numpy so that the HIP step library can be fed from the same
problem.in inputs on a machine without the Fortran driver.

Covers:
  * problem.in grammar (io.cpp:113-228) and material conversion
    (model.f90:893-945, formulation.f90:62-217);
  * grid sizing and time step (model.f90:856-889, communicator.f90:1629-1662);
  * the plane-wave box (model.f90:982-1001, 197-241);
  * the initial structured mesh in the reference numbering
    (mesh_utils.f90:64-112), geometric cuts (mesh_utils.f90:720-750,
    model.f90:303-388) -- these reproduce the reference arrays exactly for
    meshes without refinement;
  * a 2:1-balanced quadtree with hanging nodes in the reference CSR layout
    (corner, hanging, corner, ... with ediv flags; mesh_utils.f90:538-547).
    Node and element *numbering* of refined meshes follows this module's own
    rule (the reference's depends on ParMETIS adjacency order);
  * boundary classification arrays: absorbing flags (model.f90:1080-1096),
    DRM layer (model.f90:1015-1045), hanging triples (model.f90:654-675);
  * receiver placement on the top surface (visualization.f90:286-305,
    quadtree.f90:253-291).

All integer arrays produced here use the reference's 1-based values.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

U_TOL = 1.0e-9

F_YOUNG, F_NU, F_SHEAR, F_BULK, F_LAMBDA = 1, 2, 3, 4, 5
F_VP, F_VS = 1, 2

AB_NOABS, AB_ZDIR, AB_XDIR = 0, 1, 2
SE_SHORI, SE_SVERL, SE_SVERR, SE_CORNR, SE_CORNL = 1, 2, 4, 5, 6


class SetupError(RuntimeError):
    pass


# ---------------------------------------------------------------------------
# problem.in
# ---------------------------------------------------------------------------
@dataclass
class ProblemFile:
    width: float = 0.0
    height: float = 0.0
    duration: float = 0.0
    inttime: float = 0.0
    chper: float = 0.0
    nperlen: float = 0.0
    sperlen: float = 0.0
    cmats: list = field(default_factory=list)   # [(c1, c2)]
    vmats: list = field(default_factory=list)   # [(v1, v2, rho)]
    prob: str = "      "
    pparams: list = field(default_factory=list)
    geom: str = "      "
    gparams: list = field(default_factory=list)


def parse_problem(text: str) -> ProblemFile:
    """Section grammar of load_problem_file (io.cpp:113-228)."""
    lines = text.splitlines()
    pf = ProblemFile()
    if not lines or lines[0] != "PROBLEM FILE EXFEM":
        raise SetupError("incorrect header")
    i = 1

    def nxt():
        nonlocal i
        if i >= len(lines):
            raise SetupError("truncated problem file")
        s = lines[i]
        i += 1
        return s

    def six(s):
        return (s + "      ")[:6]

    while i < len(lines):
        sec = nxt()
        if sec == "DOMAIN:":
            v = nxt().split()
            pf.width, pf.height = float(v[0]), float(v[1])
        elif sec == "CH PERIOD:":
            pf.chper = float(nxt().split()[0])
        elif sec == "TIME:":
            v = nxt().split()
            pf.duration, pf.inttime = float(v[0]), float(v[1])
        elif sec == "CONTROL:":
            v = nxt().split()
            pf.nperlen, pf.sperlen = float(v[0]), float(v[1])
        elif sec == "MATERIALS:":
            n = int(nxt().split()[0])
            for _ in range(n):
                t = nxt().split()
                pf.cmats.append((t[0][0], t[2][0]))
                pf.vmats.append((float(t[1]), float(t[3]), float(t[4])))
        elif sec == "PROBLEM:":
            pf.prob = six(nxt())
            v = nxt().split()
            n = int(v[0])
            pf.pparams = [float(x) for x in v[1:1 + n]]
        elif sec == "GEOMETRY:":
            pf.geom = six(nxt())
            v = nxt().split()
            n = int(v[0])
            pf.gparams = [float(x) for x in v[1:1 + n]]
    return pf


# ---------------------------------------------------------------------------
# materials (formulation.f90:62-217, model.f90:893-945)
# ---------------------------------------------------------------------------
def _c_to_p(ch):
    return {"E": F_YOUNG, "v": F_NU, "G": F_SHEAR, "K": F_BULK, "l": F_LAMBDA}.get(ch, 0)


def _c_to_v(ch):
    return {"p": F_VP, "s": F_VS}.get(ch, 0)


def e_to_e(pfrom, pto, param):
    """Mechanical-parameter conversion, same fixed-point sweep order as
    formulation.f90:135-217."""
    frm = [0.0] * 6
    found = [False] * 6
    for k in range(2):
        frm[pfrom[k]] = param[k]
        found[pfrom[k]] = True
    for _ in range(3):
        if all(found[p] for p in pto):
            break
        for j in range(1, 6):
            if found[j]:
                continue
            if j == F_YOUNG:
                if found[F_SHEAR] and found[F_NU]:
                    found[j] = True
                    frm[j] = 2 * (1 + frm[F_NU]) * frm[F_SHEAR]
                elif found[F_BULK] and found[F_NU]:
                    found[j] = True
                    frm[j] = 3 * (1 - 2 * frm[F_NU]) * frm[F_BULK]
                elif found[F_LAMBDA] and found[F_NU]:
                    found[j] = True
                    frm[j] = (1 - 2 * frm[F_NU]) * (1 + frm[F_NU])
                    frm[j] = frm[j] * frm[F_LAMBDA] / frm[F_NU]
            elif j == F_NU:
                if found[F_YOUNG] and found[F_SHEAR]:
                    found[j] = True
                    frm[j] = frm[F_YOUNG] - 2 * frm[F_SHEAR]
                    frm[j] = frm[j] / (2 * frm[F_SHEAR])
                elif found[F_YOUNG] and found[F_BULK]:
                    found[j] = True
                    frm[j] = 3 * frm[F_BULK] - frm[F_YOUNG]
                    frm[j] = frm[j] / (6 * frm[F_BULK])
                elif found[F_BULK] and found[F_SHEAR]:
                    found[j] = True
                    frm[j] = 3 * frm[F_BULK] - 2 * frm[F_SHEAR]
                    frm[j] = frm[j] / (6 * frm[F_BULK] + 2 * frm[F_SHEAR])
                elif found[F_SHEAR] and found[F_LAMBDA]:
                    found[j] = True
                    frm[j] = 2 * (frm[F_LAMBDA] + frm[F_SHEAR])
                    frm[j] = frm[F_LAMBDA] / frm[j]
                elif found[F_BULK] and found[F_LAMBDA]:
                    found[j] = True
                    frm[j] = 3 * frm[F_BULK] - frm[F_LAMBDA]
                    frm[j] = frm[F_LAMBDA] / frm[j]
            elif j == F_SHEAR:
                if found[F_YOUNG] and found[F_NU]:
                    found[j] = True
                    frm[j] = frm[F_YOUNG] / (2 * (frm[F_NU] + 1))
            elif j == F_BULK:
                if found[F_YOUNG] and found[F_NU]:
                    found[j] = True
                    frm[j] = frm[F_YOUNG] / (3 * (1 - 2 * frm[F_NU]))
            elif j == F_LAMBDA:
                if found[F_YOUNG] and found[F_NU]:
                    found[j] = True
                    frm[j] = frm[F_YOUNG] * frm[F_NU]
                    frm[j] = frm[j] / ((1 + frm[F_NU]) * (1 - 2 * frm[F_NU]))
    if not all(found[p] for p in pto):
        raise SetupError("invalid material conversion")
    return [frm[p] for p in pto]


def v_to_e(vp, vs, rho):
    """(mu, lambda) from velocities (formulation.f90:62-71)."""
    g = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * g
    return e_to_e((F_SHEAR, F_LAMBDA), (F_SHEAR, F_LAMBDA), (g, lam))


def setup_materials(cmats, vmats):
    """matconst rows (mu, lambda, rho, vp, vs) per material (model.f90:893-945)."""
    out = []
    for (c1, c2), (v1, v2, rho) in zip(cmats, vmats):
        i1, i2 = _c_to_p(c1), _c_to_p(c2)
        if i1 == 0 or i2 == 0:
            if i2 == F_NU:
                i1, i2, c1, c2, v1, v2 = i2, i1, c2, c1, v2, v1
            if i1 == F_NU:
                i2 = _c_to_v(c2)
                if i2 == 0:
                    raise SetupError("Not recognized material")
                nu = v1
                rel = math.sqrt(2.0 * (nu - 1.0) / (2.0 * nu - 1.0))
                vel = [0.0, 0.0, 0.0]
                vel[i2] = v2
                if i2 == F_VS:
                    vel[F_VP] = rel * vel[F_VS]
                else:
                    vel[F_VS] = vel[F_VP] / rel
                vp, vs = vel[F_VP], vel[F_VS]
                mu, lam = v_to_e(vp, vs, rho)
            elif i1 == 0 and i2 == 0:
                i1, i2 = _c_to_v(c1), _c_to_v(c2)
                if i1 == 0 or i2 == 0:
                    raise SetupError("Not recognized material")
                vel = [0.0, 0.0, 0.0]
                vel[i1] = v1
                vel[i2] = v2
                vp, vs = vel[F_VP], vel[F_VS]
                mu, lam = v_to_e(vp, vs, rho)
            else:
                raise SetupError("Not recognized material")
        else:
            mu, lam = e_to_e((i1, i2), (F_SHEAR, F_LAMBDA), (v1, v2))
            vp = math.sqrt((lam + 2 * mu) / rho)
            vs = math.sqrt(mu / rho)
        out.append((mu, lam, rho, vp, vs))
    return np.array(out, dtype=np.float64)


# ---------------------------------------------------------------------------
# grid, time, plane-wave box
# ---------------------------------------------------------------------------
@dataclass
class Model:
    """The model-module scalars the step loop consumes (model.f90:22-101)."""
    matconst: np.ndarray
    blength: float
    mlength: float
    totcols: int
    totrows: int
    nsteps: int
    deltat: float
    duration: float
    inttime: float
    nperlen: float
    sperlen: float
    probconst: np.ndarray
    prob: str = "      "
    pparams: tuple = ()
    geom: str = "      "
    gparams: tuple = ()
    # plane wave (init_wave)
    wrows: int = 0
    wcols: int = 0
    worigin: tuple = (0.0, 0.0)
    wt0: float = 0.0
    wangle: float = 0.0
    wamp: float = 0.0
    wfreq: float = 0.0
    wvelo: tuple = (0.0, 0.0)
    io_interval: int = 1

    @property
    def is_wave(self):
        return self.prob[:4] == "WAVE"

    @property
    def wave_kind(self):
        return {"WAVEP ": 1, "WAVESV": 2}.get(self.prob, 0)

    @property
    def box(self):
        """(left, right, top) of the plane-wave box (model.f90:205-207)."""
        left = float(self.wcols - 1) * self.blength
        right = float((self.totcols - self.wcols) - 1) * self.blength
        top = float(self.wrows - 1) * self.blength
        return left, right, top


def build_model(pf: ProblemFile) -> Model:
    """setup_problem_params (model.f90:856-889) + init_wave (982-1001)."""
    matconst = setup_materials(pf.cmats, pf.vmats)
    nm = len(matconst)
    probconst = np.zeros((nm, 2))
    length = 0.0
    deltat = math.inf
    for i in range(nm):
        ln = (pf.chper * matconst[i, 4]) / pf.nperlen
        st = (ln / matconst[i, 3]) / pf.sperlen
        probconst[i] = (ln, st)
        if length < ln:
            length = ln
        if deltat > st:
            deltat = st
    m = math.ceil(pf.height / length)
    length = pf.height / float(m)
    n = math.ceil(pf.width / length)
    steps = math.ceil(pf.duration / deltat)
    deltat = pf.duration / float(steps)
    mdl = Model(matconst=matconst, blength=length, mlength=length, totcols=n + 1,
                totrows=m + 1, nsteps=steps, deltat=deltat, duration=deltat * steps,
                inttime=pf.inttime, nperlen=pf.nperlen, sperlen=pf.sperlen,
                probconst=probconst, prob=pf.prob, pparams=tuple(pf.pparams),
                geom=pf.geom, gparams=tuple(pf.gparams))
    if mdl.is_wave:
        init_wave(mdl)
    return mdl


def init_wave(mdl: Model):
    """model.f90:982-1001 (origin depth snapped to the base grid)."""
    pp = mdl.pparams
    mdl.wrows = math.ceil(pp[1] / mdl.blength) + 1
    mdl.wcols = int((mdl.totcols - (1 + math.ceil(pp[0] / mdl.blength))) / 2)
    mat = int(pp[4])
    mdl.wvelo = (float(mdl.matconst[mat - 1, 3]), float(mdl.matconst[mat - 1, 4]))
    mdl.wt0 = pp[5]
    mdl.wangle = pp[6]
    mdl.wamp = pp[7]
    mdl.wfreq = pp[8]
    wzero = math.floor(pp[3] / mdl.blength)
    mdl.worigin = (pp[2], float(wzero) * mdl.blength)


def recalculate_deltat(mdl: Model, elen: np.ndarray, ematprop: np.ndarray):
    """communicator.f90:1629-1662 on one rank: dt = min(dt, min_e (h/vp)/sperlen),
    nsteps = int(duration/dt), io_interval = int(inttime/dt)."""
    mindt = mdl.deltat
    vp = mdl.matconst[ematprop - 1, 3]
    eldt = (elen / vp) / mdl.sperlen
    if eldt.size:
        mindt = min(mindt, float(eldt.min()))
    mdl.deltat = mindt
    mdl.nsteps = int(mdl.duration / mdl.deltat)
    mdl.io_interval = int(mdl.inttime / mdl.deltat)
    mdl.inttime = mdl.deltat * mdl.io_interval
    return mdl


# ---------------------------------------------------------------------------
# region predicates (model.f90:197-241, 397-445)
# ---------------------------------------------------------------------------
def inside_soil(mdl: Model, x, z):
    if not mdl.is_wave:
        return np.zeros(np.broadcast(x, z).shape, dtype=bool)
    left, right, top = mdl.box
    return ((x - left) < -U_TOL) | ((x - right) > U_TOL) | ((z - top) > U_TOL)


def void_soil(mdl: Model, z):
    return (z - mdl.worigin[1]) < -U_TOL


def material_prop(mdl: Model, x, z):
    x = np.asarray(x, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    out = np.ones(np.broadcast(x, z).shape, dtype=np.int32)
    g = mdl.gparams
    if mdl.geom == "VALLEY":
        d = np.sqrt((x - g[0]) ** 2 + (z - g[1]) ** 2)
        out = np.where(d < g[2], 2, 1).astype(np.int32)
    elif mdl.geom == "VARECT":
        inside = (np.abs(x - g[0]) < g[2]) & (np.abs(z - g[1]) < g[3])
        out = np.where(inside, 2, 1).astype(np.int32)
    if mdl.is_wave:
        out = np.where(inside_soil(mdl, x, z), int(mdl.pparams[4]), out).astype(np.int32)
    return out


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
@dataclass
class Mesh:
    eptr: np.ndarray      # int32 (nelems+1), 1-based
    eind: np.ndarray      # int32, 1-based node ids
    ediv: np.ndarray      # uint8 hanging flags
    nodes: np.ndarray     # float64 (nnodes, 2)
    ematprop: np.ndarray  # int32 (nelems,), 1-based

    @property
    def nelems(self):
        return len(self.eptr) - 1

    @property
    def nnodes(self):
        return len(self.nodes)

    def corners(self):
        """(nelems, 4) 1-based non-hanging corners (elemallnhnodes)."""
        keep = self.ediv == 0
        c = self.eind[keep]
        if c.size != 4 * self.nelems:
            raise SetupError("element without exactly 4 non-hanging corners")
        return c.reshape(-1, 4)

    def elength(self):
        c = self.corners()
        return self.nodes[c[:, 2] - 1, 1] - self.nodes[c[:, 0] - 1, 1]


def initial_mesh(mdl: Model) -> Mesh:
    """mesh_utils.f90:64-112 on one rank (frow=fcol=1, lrow=totrows, lcol=totcols)."""
    nr, nc = mdl.totrows, mdl.totcols
    i = np.arange(1, nc + 1)
    j = np.arange(1, nr + 1)
    I, J = np.meshgrid(i, j, indexing="ij")   # node index = (i-1)*nr + j
    nodes = np.stack([(I - 1).astype(np.float64) * mdl.blength,
                      (J - 1).astype(np.float64) * mdl.blength], axis=-1).reshape(-1, 2)
    il = np.arange(1, nc)
    jl = np.arange(1, nr)
    IL, JL = np.meshgrid(il, jl, indexing="ij")
    IL = IL.ravel()
    JL = JL.ravel()
    eind = np.stack([(IL - 1) * nr + JL, IL * nr + JL, IL * nr + (JL + 1),
                     (IL - 1) * nr + (JL + 1)], axis=-1).astype(np.int32).ravel()
    ne = len(IL)
    eptr = (1 + 4 * np.arange(ne + 1)).astype(np.int32)
    cx = ((IL - 1).astype(np.float64) + 0.5) * mdl.blength
    cz = ((JL - 1).astype(np.float64) + 0.5) * mdl.blength
    ematprop = material_prop(mdl, cx, cz)
    return Mesh(eptr, eind, np.zeros(len(eind), np.uint8), nodes, ematprop.astype(np.int32))


def needs_deletion(mdl: Model, mesh: Mesh) -> np.ndarray:
    """model.f90:303-388 (CANYON, RECTAN; soil void) for every element."""
    c = mesh.corners()
    elen = mesh.elength()
    px = mesh.nodes[c - 1, 0]
    pz = mesh.nodes[c - 1, 1]
    cx = px.sum(axis=1) / 4.0
    cz = pz.sum(axis=1) / 4.0
    out = np.zeros(mesh.nelems, dtype=bool)
    big = (elen - mdl.mlength) > U_TOL
    soil = inside_soil(mdl, cx, cz)
    g = mdl.gparams
    if mdl.geom == "CANYON":
        d = np.sqrt((px - g[0]) ** 2 + (pz - g[1]) ** 2)
        geo = ~(d > g[2]).any(axis=1)
    elif mdl.geom == "RECTAN":
        geo = ~((np.abs(px - g[0]) > g[2]) | (np.abs(pz - g[1]) > g[3])).any(axis=1)
    else:
        geo = np.zeros(mesh.nelems, dtype=bool)
    out = np.where(soil, void_soil(mdl, cz), geo)
    out[big] = False
    return out


def cut_geom(mesh: Mesh, mask: np.ndarray) -> Mesh:
    """mesh_utils.f90:720-750: drop masked elements, then drop unreferenced
    nodes keeping the node order (local number = rank in gnode order)."""
    keep_e = ~mask
    lens = np.diff(mesh.eptr)
    row = np.repeat(np.arange(mesh.nelems), lens)
    keep_entries = keep_e[row]
    eind = mesh.eind[keep_entries]
    ediv = mesh.ediv[keep_entries]
    nl = lens[keep_e]
    eptr = np.concatenate([[1], 1 + np.cumsum(nl)]).astype(np.int32)
    used = np.zeros(mesh.nnodes + 1, dtype=bool)
    used[eind] = True
    used[0] = False
    newid = np.cumsum(used).astype(np.int32)
    return Mesh(eptr, newid[eind].astype(np.int32), ediv, mesh.nodes[used[1:]],
                mesh.ematprop[keep_e])


def reference_uniform_mesh(mdl: Model) -> Mesh:
    """Initial mesh + cuts, i.e. the reference's final mesh whenever no element
    needs division (single material, or every material resolved by the base
    grid)."""
    mesh = initial_mesh(mdl)
    return cut_geom(mesh, needs_deletion(mdl, mesh))


def quadtree_mesh(nx0: int, nz0: int, h0: float, nlevels: int,
                  refine: Callable[[np.ndarray, np.ndarray, np.ndarray], np.ndarray],
                  delete: Optional[Callable] = None,
                  material: Optional[Callable] = None) -> Mesh:
    """2:1 edge-balanced quadtree over an nx0 x nz0 grid of size-h0 squares.

    refine(x0, z0, size) -> bool array: leaf wants division.
    delete(corner_x (n,4), corner_z (n,4)) -> bool array: drop the leaf.
    material(cx, cz, size) -> int32 1-based material ids.

    Output follows the reference CSR convention: each row lists the corners
    BL, BR, TR, TL (x right, z down) with a hanging node (ediv=1) inserted
    after the corner that starts the edge it lies on.
    """
    L = nlevels - 1
    F = 1 << L
    nxf, nzf = nx0 * F, nz0 * F
    lev = np.zeros((nxf, nzf), dtype=np.int8)
    hf = h0 / F

    def leaves_at(l):
        s = 1 << (L - l)
        sub = lev[::s, ::s]
        ix, iz = np.nonzero(sub == l)
        return ix * s, iz * s, s

    def split(ix, iz, s):
        for a, b in zip(ix, iz):
            lev[a:a + s, b:b + s] += 1

    for l in range(L):
        ix, iz, s = leaves_at(l)
        if len(ix):
            want = refine(ix * hf, iz * hf, np.full(len(ix), s * hf))
            split(ix[want], iz[want], s)
        # 2:1 balance: any fine cell whose edge-neighbour is >1 level finer
        # forces its leaf to divide; iterate to a fixed point.
        while True:
            nb = np.full_like(lev, -1)
            nb[:-1, :] = np.maximum(nb[:-1, :], lev[1:, :])
            nb[1:, :] = np.maximum(nb[1:, :], lev[:-1, :])
            nb[:, :-1] = np.maximum(nb[:, :-1], lev[:, 1:])
            nb[:, 1:] = np.maximum(nb[:, 1:], lev[:, :-1])
            bad = nb > lev + 1
            if not bad.any():
                break
            # leaves to split, decided on this pass's snapshot of the map
            bx, bz = np.nonzero(bad)
            lv = lev[bx, bz].astype(np.int64)
            s = np.left_shift(1, L - lv)
            todo = np.unique(np.stack([(bx // s) * s, (bz // s) * s, s], axis=1), axis=0)
            for a0, b0, sz in todo:
                lev[a0:a0 + sz, b0:b0 + sz] += 1

    # Enumerate leaves (coarse-cell order, then Morton order inside).
    lx, lz, ls = [], [], []
    for l in range(L + 1):
        ix, iz, s = leaves_at(l)
        lx.append(ix)
        lz.append(iz)
        ls.append(np.full(len(ix), s))
    lx = np.concatenate(lx)
    lz = np.concatenate(lz)
    ls = np.concatenate(ls)
    # ordering key: coarse column, coarse row, then interleaved bits
    cxi, czi = lx // F, lz // F
    fx, fz = lx % F, lz % F
    mort = np.zeros(len(lx), dtype=np.int64)
    for b in range(L):
        mort |= ((fz >> b) & 1).astype(np.int64) << (2 * b)
        mort |= ((fx >> b) & 1).astype(np.int64) << (2 * b + 1)
    order = np.lexsort((mort, czi, cxi))
    lx, lz, ls = lx[order], lz[order], ls[order]

    # Deletion of whole leaves.
    if delete is not None:
        ccx = np.stack([lx, lx + ls, lx + ls, lx], axis=1) * hf
        ccz = np.stack([lz, lz, lz + ls, lz + ls], axis=1) * hf
        keep = ~delete(ccx, ccz)
        lx, lz, ls = lx[keep], lz[keep], ls[keep]

    ne = len(lx)
    # Corner + midpoint candidates on the fine lattice (nodes live there).
    key = lambda a, b: a.astype(np.int64) * (nzf + 1) + b.astype(np.int64)
    cx = np.stack([lx, lx + ls, lx + ls, lx], axis=1)
    cz = np.stack([lz, lz, lz + ls, lz + ls], axis=1)
    corner_keys = key(cx, cz)
    half = ls // 2
    # edge midpoints: bottom (BL->BR), right (BR->TR), top (TR->TL), left (TL->BL)
    mx = np.stack([lx + half, lx + ls, lx + half, lx], axis=1)
    mz = np.stack([lz, lz + half, lz + ls, lz + half], axis=1)
    mid_keys = key(mx, mz)
    corner_set = np.unique(corner_keys)
    has_mid = np.isin(mid_keys, corner_set) & (ls[:, None] > 1)
    # Node numbering: corner nodes in column-major lattice order (x major,
    # z minor), matching the reference's initial numbering on uniform meshes.
    all_nodes = corner_set
    nid = {}
    node_ids = np.searchsorted(all_nodes, corner_keys) + 1
    mid_ids = np.where(has_mid, np.searchsorted(all_nodes, mid_keys) + 1, 0)
    rows = np.zeros((ne, 8), dtype=np.int64)
    flags = np.zeros((ne, 8), dtype=np.uint8)
    rows[:, 0::2] = node_ids
    rows[:, 1::2] = mid_ids
    flags[:, 1::2] = 1
    valid = rows != 0
    lens = valid.sum(axis=1)
    eind = rows[valid].astype(np.int32)
    ediv = flags[valid].astype(np.uint8)
    eptr = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    nzl = nzf + 1
    nodes = np.stack([(all_nodes // nzl).astype(np.float64) * hf,
                      (all_nodes % nzl).astype(np.float64) * hf], axis=1)
    if material is None:
        ematprop = np.ones(ne, dtype=np.int32)
    else:
        ematprop = material((lx + ls / 2.0) * hf, (lz + ls / 2.0) * hf, ls * hf).astype(np.int32)
    return Mesh(eptr, eind, ediv, nodes, ematprop)


# ---------------------------------------------------------------------------
# boundary classification (model.f90)
# ---------------------------------------------------------------------------
def create_absorbingbc(mdl: Model, mesh: Mesh) -> np.ndarray:
    """model.f90:1080-1096."""
    left = 0.0 * mdl.blength
    right = float(mdl.totcols - 1) * mdl.blength
    top = float(mdl.totrows - 1) * mdl.blength
    x = mesh.nodes[:, 0]
    z = mesh.nodes[:, 1]
    xdir = (np.abs(x - left) < U_TOL) | (np.abs(x - right) < U_TOL)
    zdir = np.abs(z - top) < U_TOL
    return np.where(xdir, AB_XDIR, np.where(zdir, AB_ZDIR, AB_NOABS)).astype(np.int32)


def wave_elems(mdl: Model, mesh: Mesh) -> np.ndarray:
    """model.f90:1015-1045 -> (nseis, 2) int32 of (element, kind)."""
    left, right, top = mdl.box
    c = mesh.corners()
    x1, z1 = mesh.nodes[c[:, 0] - 1, 0], mesh.nodes[c[:, 0] - 1, 1]
    x2, z2 = mesh.nodes[c[:, 1] - 1, 0], mesh.nodes[c[:, 1] - 1, 1]
    eq = lambda a, b: np.abs(a - b) < U_TOL
    un = lambda a, b: (a - b) < -U_TOL
    ov = lambda a, b: (a - b) > U_TOL
    k = np.zeros(mesh.nelems, dtype=np.int32)
    c5 = eq(x1, right) & eq(z1, top)
    c6 = eq(x2, left) & eq(z2, top)
    c2 = eq(x2, left) & un(z2, top)
    c4 = eq(x1, right) & un(z1, top)
    c1 = ov(x2, left) & un(x1, right) & eq(z1, top)
    k = np.where(c1, SE_SHORI, k)
    k = np.where(c4, SE_SVERR, k)
    k = np.where(c2, SE_SVERL, k)
    k = np.where(c5, SE_CORNR, k)
    k = np.where(c6, SE_CORNL, k)   # first matching branch wins
    idx = np.nonzero(k)[0]
    return np.stack([idx + 1, k[idx]], axis=1).astype(np.int32)


def setup_hanging(mesh: Mesh) -> np.ndarray:
    """model.f90:654-675 -> (nhang, 3) int32 of (hanging, anchor1, anchor2)."""
    pos = np.nonzero(mesh.ediv)[0]
    if pos.size == 0:
        return np.zeros((0, 3), dtype=np.int32)
    row = np.searchsorted(mesh.eptr - 1, pos, side="right") - 1
    start = mesh.eptr[row] - 1
    end = mesh.eptr[row + 1] - 1
    nxt = np.where(pos + 1 < end, pos + 1, start)
    return np.stack([mesh.eind[pos], mesh.eind[pos - 1], mesh.eind[nxt]], axis=1).astype(np.int32)


def setup_point_source(mdl: Model, mesh: Mesh):
    """model.f90:1051-1078: pparams = (amp, sigma, start, dist, x1, z1, x2,
    z2, ...); the nodes strictly closer than dist to any point, unique and
    ascending."""
    pp = mdl.pparams
    amp, sigma, start, dist = pp[0], pp[1], pp[2], pp[3]
    sel = np.zeros(mesh.nnodes, dtype=bool)
    for j in range(4, len(pp) - 1, 2):
        d = np.sqrt((pp[j] - mesh.nodes[:, 0]) ** 2 + (pp[j + 1] - mesh.nodes[:, 1]) ** 2)
        sel |= d < dist
    return (mdl.prob, amp, sigma, start, (np.nonzero(sel)[0] + 1).astype(np.int32))


def find_highest(mesh: Mesh, xmin: float, xmax: float, dist: float):
    """Receivers on the top surface (visualization.f90:286-305): x = xmin +
    dist*i for i = 1..int((xmax-xmin)/dist)+1, at the top of the highest
    element whose x-range contains x (tolerance u_tol).  Returns (elems
    (1-based, 0 = not found), points (n, 2))."""
    npts = int((xmax - xmin) / dist) + 1
    xs = xmin + dist * np.arange(1, npts + 1, dtype=np.float64)
    c = mesh.corners()
    ex0 = mesh.nodes[c[:, 0] - 1, 0]
    ex1 = mesh.nodes[c[:, 2] - 1, 0]
    ez0 = mesh.nodes[c[:, 0] - 1, 1]
    elems = np.zeros(npts, dtype=np.int32)
    pts = np.zeros((npts, 2))
    order = np.lexsort((-ex0, ez0))   # highest first; ties -> right-hand element
    for i, x in enumerate(xs):
        ok = ~(((x - ex1[order]) > U_TOL) | ((x - ex0[order]) < -U_TOL))
        hit = np.nonzero(ok)[0]
        pts[i, 0] = x
        if hit.size:
            e = order[hit[0]]
            elems[i] = e + 1
            pts[i, 1] = ez0[e]
    return elems, pts


@dataclass
class Case:
    """Everything the step library needs for one site-response instance."""
    model: Model
    mesh: Mesh
    absorbing: np.ndarray
    hanging: np.ndarray
    seismic: np.ndarray
    recv_elem: np.ndarray
    recv_xy: np.ndarray
    recv_ff: np.ndarray
    dt: float
    nsteps: int
    name: str = ""
    elem_props: Optional[np.ndarray] = None   # (nelems, 5): mu, lambda, rho, vp, vs
    point: Optional[tuple] = None             # (prob, amp, sigma, start, nodes 1-based)

    @property
    def nelems(self):
        return self.mesh.nelems


def finish_case(mdl: Model, mesh: Mesh, sheet=None, name="", recalc_dt=True,
                nsteps=None, dt=None) -> Case:
    """problem_setup (model.f90:958-980) + recalculate_deltat + receivers."""
    absorbing = create_absorbingbc(mdl, mesh)
    hanging = setup_hanging(mesh)
    seismic = wave_elems(mdl, mesh) if mdl.is_wave else np.zeros((0, 2), np.int32)
    if recalc_dt:
        recalculate_deltat(mdl, mesh.elength(), mesh.ematprop)
    if sheet is not None:
        relem, rxy = find_highest(mesh, *sheet)
    else:
        relem, rxy = np.zeros(0, np.int32), np.zeros((0, 2))
    rff = inside_soil(mdl, rxy[:, 0], rxy[:, 1]).astype(np.uint8) if len(rxy) else np.zeros(0, np.uint8)
    point = setup_point_source(mdl, mesh) if mdl.prob in ("POINTN", "POINTD") else None
    return Case(model=mdl, mesh=mesh, absorbing=absorbing, hanging=hanging, seismic=seismic,
                recv_elem=relem, recv_xy=rxy, recv_ff=rff,
                dt=mdl.deltat if dt is None else dt,
                nsteps=mdl.nsteps if nsteps is None else nsteps, name=name, point=point)


def case_from_problem_text(text: str, sheet=None, name="") -> Case:
    """A reference problem.in whose mesh needs no refinement, reproduced
    exactly (initial grid + cuts)."""
    pf = parse_problem(text)
    mdl = build_model(pf)
    mesh = reference_uniform_mesh(mdl)
    lengths = mdl.probconst[:, 0]
    if np.any(lengths < mdl.blength - U_TOL):
        raise SetupError("problem needs quadtree refinement; use quadtree_mesh")
    return finish_case(mdl, mesh, sheet=sheet, name=name)

"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
REF_IO_PATH = HERE / "_ref" / "libexfem_io.so"

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C oracle`")
        _lib = C.CDLL(str(LIB_PATH))
        _lib.orc_calcm.restype = C.c_double
    return _lib


class Wave(C.Structure):
    _fields_ = [("kind", C.c_int), ("alpha", C.c_double), ("beta", C.c_double),
                ("freq", C.c_double), ("sine", C.c_double), ("cose", C.c_double),
                ("sinf", C.c_double), ("cosf", C.c_double), ("a", C.c_double),
                ("a1", C.c_double), ("b", C.c_double), ("b1", C.c_double),
                ("ui", C.c_double * 6)]


class Problem(C.Structure):
    _fields_ = [("nelems", C.c_int), ("nnodes", C.c_int),
                ("eptr", C.c_void_p), ("eind", C.c_void_p), ("ediv", C.c_void_p),
                ("nodes", C.c_void_p), ("ematprop", C.c_void_p), ("nummat", C.c_int),
                ("matconst", C.c_void_p), ("absorbing", C.c_void_p), ("nhang", C.c_int),
                ("hanging", C.c_void_p), ("nseis", C.c_int), ("seismic", C.c_void_p),
                ("wave_kind", C.c_int), ("worigin", C.c_double * 2), ("wt0", C.c_double),
                ("wangle", C.c_double), ("wamp", C.c_double), ("wfreq", C.c_double),
                ("walpha", C.c_double), ("wbeta", C.c_double), ("box_left", C.c_double),
                ("box_right", C.c_double), ("box_top", C.c_double), ("dt", C.c_double),
                ("nsteps", C.c_int), ("nrecv", C.c_int), ("recv_elem", C.c_void_p),
                ("recv_xy", C.c_void_p), ("recv_interval", C.c_int), ("hang_mode", C.c_int),
                ("point_kind", C.c_int), ("point_amp", C.c_double), ("point_sigma", C.c_double),
                ("point_start", C.c_double), ("npoint", C.c_int), ("pointnodes", C.c_void_p)]


def _p(a, dt, keep):
    a = np.ascontiguousarray(a, dtype=dt)
    keep.append(a)
    return a.ctypes.data_as(C.c_void_p)


def ricker(amp, sigma, t):
    lib().orc_ricker.restype = C.c_double
    return lib().orc_ricker(C.c_double(amp), C.c_double(sigma), C.c_double(t))


def calck(mu, lam):
    k = np.zeros(64)
    lib().orc_calck(C.c_double(mu), C.c_double(lam), k.ctypes.data_as(C.c_void_p))
    return k.reshape(8, 8, order="F")


def calcsqm():
    m = np.zeros(64)
    lib().orc_calcsqm(m.ctypes.data_as(C.c_void_p))
    return m.reshape(8, 8, order="F")


def multdrm(k, u, e):
    p = np.zeros(8)
    kk = np.asfortranarray(k, dtype=np.float64).ravel(order="F")
    uu = np.ascontiguousarray(u, dtype=np.float64)
    ee = np.ascontiguousarray(e, dtype=np.uint8)
    lib().orc_multdrm(kk.ctypes.data_as(C.c_void_p), uu.ctypes.data_as(C.c_void_p),
                      ee.ctypes.data_as(C.c_void_p), p.ctypes.data_as(C.c_void_p))
    return p


def initplanewaves(kind, mag, alpha, beta, angle, freq):
    w = Wave()
    rc = lib().orc_initplanewaves(C.byref(w), kind, C.c_double(mag), C.c_double(alpha),
                                  C.c_double(beta), C.c_double(angle), C.c_double(freq))
    return rc, w


def planewave(w, t, x1, x3):
    u = np.zeros(6)
    lib().orc_planewave(C.byref(w), C.c_double(t), C.c_double(x1), C.c_double(x3),
                        u.ctypes.data_as(C.c_void_p))
    return u.reshape(2, 3, order="F")


def setup_hanging(mesh):
    keep = []
    n = lib().orc_setup_hanging(mesh.nelems, _p(mesh.eptr, np.int32, keep),
                                _p(mesh.eind, np.int32, keep), _p(mesh.ediv, np.uint8, keep), None)
    out = np.zeros(max(3 * n, 1), np.int32)
    lib().orc_setup_hanging(mesh.nelems, _p(mesh.eptr, np.int32, keep),
                            _p(mesh.eind, np.int32, keep), _p(mesh.ediv, np.uint8, keep),
                            out.ctypes.data_as(C.c_void_p))
    return out[:3 * n].reshape(-1, 3)


def create_absorbingbc(mdl, mesh):
    keep = []
    out = np.zeros(mesh.nnodes, np.int32)
    lib().orc_create_absorbingbc(mesh.nnodes, _p(mesh.nodes.ravel(), np.float64, keep),
                                 C.c_double(mdl.blength), mdl.totcols, mdl.totrows,
                                 out.ctypes.data_as(C.c_void_p))
    return out


def wave_elems(mdl, mesh):
    keep = []
    args = (mesh.nelems, _p(mesh.eptr, np.int32, keep), _p(mesh.eind, np.int32, keep),
            _p(mesh.ediv, np.uint8, keep), _p(mesh.nodes.ravel(), np.float64, keep),
            C.c_double(mdl.blength), mdl.wcols, mdl.wrows, mdl.totcols)
    n = lib().orc_wave_elems(*args, None)
    out = np.zeros(max(2 * n, 1), np.int32)
    lib().orc_wave_elems(*args, out.ctypes.data_as(C.c_void_p))
    return out[:2 * n].reshape(-1, 2)


def problem(case, nsteps=None, recv_interval=1, hang_mode=2, receivers=True):
    """orc_problem for a case; returns (Problem, keep-alive list)."""
    keep = []
    m = case.mesh
    mdl = case.model
    P = Problem()
    P.nelems, P.nnodes = m.nelems, m.nnodes
    P.eptr = _p(m.eptr, np.int32, keep)
    P.eind = _p(m.eind, np.int32, keep)
    P.ediv = _p(m.ediv, np.uint8, keep)
    P.nodes = _p(m.nodes.ravel(), np.float64, keep)
    if case.elem_props is not None:
        props = np.ascontiguousarray(case.elem_props, dtype=np.float64)
        P.ematprop = _p(np.arange(1, m.nelems + 1), np.int32, keep)
        P.nummat = m.nelems
        P.matconst = _p(props.ravel(), np.float64, keep)
    else:
        P.ematprop = _p(m.ematprop, np.int32, keep)
        P.nummat = len(mdl.matconst)
        P.matconst = _p(mdl.matconst.ravel(), np.float64, keep)
    P.absorbing = _p(case.absorbing, np.int32, keep)
    P.nhang = len(case.hanging)
    P.hanging = _p(case.hanging.ravel(), np.int32, keep)
    P.nseis = len(case.seismic)
    P.seismic = _p(case.seismic.ravel(), np.int32, keep)
    P.wave_kind = mdl.wave_kind
    P.worigin[0], P.worigin[1] = mdl.worigin
    P.wt0, P.wangle, P.wamp, P.wfreq = mdl.wt0, mdl.wangle, mdl.wamp, mdl.wfreq
    P.walpha, P.wbeta = mdl.wvelo
    if mdl.is_wave:
        P.box_left, P.box_right, P.box_top = mdl.box
    P.dt = case.dt
    P.nsteps = case.nsteps if nsteps is None else nsteps
    nr = len(case.recv_elem) if receivers else 0
    P.nrecv = nr
    P.recv_elem = _p(case.recv_elem if nr else np.zeros(1), np.int32, keep)
    P.recv_xy = _p(case.recv_xy.ravel() if nr else np.zeros(2), np.float64, keep)
    P.recv_interval = recv_interval
    P.hang_mode = hang_mode
    if case.point is not None:
        kind, amp, sigma, start, nodes = case.point
        P.point_kind = {"POINTN": 1, "POINTD": 2}[kind]
        P.point_amp, P.point_sigma, P.point_start = amp, sigma, start
        P.npoint = len(nodes)
        P.pointnodes = _p(nodes, np.int32, keep)
    return P, keep


def run(case, nsteps=None, recv_interval=1, hang_mode=2, receivers=True, threads=1):
    """init_explicit + nsteps of run_explicit on the CPU (element loop on
    `threads` OpenMP threads, bit-identical results).  Returns
    (u, up, mfull, recv (nsamp, nrecv, 2) or None)."""
    P, keep = problem(case, nsteps, recv_interval, hang_mode, receivers)
    m = case.mesh
    nr = P.nrecv
    nd = 2 * m.nnodes
    u = np.zeros(nd)
    up = np.zeros(nd)
    mf = np.zeros(nd)
    nsamp = (P.nsteps // recv_interval + 1) if (nr and recv_interval > 0) else 0
    rec = np.zeros((max(nsamp, 1), max(nr, 1), 2))
    args = (C.byref(P), u.ctypes.data_as(C.c_void_p), up.ctypes.data_as(C.c_void_p),
            mf.ctypes.data_as(C.c_void_p), rec.ctypes.data_as(C.c_void_p) if nsamp else None)
    rc = lib().orc_run(*args) if threads <= 1 else lib().orc_run_mt(*args, C.c_int(threads))
    if rc < 0:
        raise RuntimeError(f"oracle failed with {rc}")
    return u, up, mf, (rec[:rc] if nsamp else None)


def run_ld(case, nsteps=None, recv_interval=1, receivers=True, threads=1):
    """The loop in 80-bit long double (orc_run_ld): (u, up, recv) as run()
    returns them, without the effective mass."""
    P, keep = problem(case, nsteps, recv_interval, 2, receivers)
    nd = 2 * case.mesh.nnodes
    u = np.zeros(nd)
    up = np.zeros(nd)
    nr = P.nrecv
    nsamp = (P.nsteps // recv_interval + 1) if (nr and recv_interval > 0) else 0
    rec = np.zeros((max(nsamp, 1), max(nr, 1), 2))
    rc = lib().orc_run_ld(C.byref(P), u.ctypes.data_as(C.c_void_p), up.ctypes.data_as(C.c_void_p),
                          rec.ctypes.data_as(C.c_void_p) if nsamp else None, C.c_int(threads))
    if rc < 0:
        raise RuntimeError(f"oracle failed with {rc}")
    return u, up, (rec[:rc] if nsamp else None)


def visu_field(case, u, elems, pts, time=0.0, free_field=True):
    """The loop's receiver sampler on state u (2*nnodes, hanging nodes
    projected) at points pts (n, 2) of 1-based elements elems (0 = not
    found), at visualization time `time`: fields (n, 2).  free_field=False
    samples as if the problem had no plane wave."""
    P, keep = problem(case, receivers=False)
    if not free_field:
        P.wave_kind = 0
    u = np.ascontiguousarray(u, dtype=np.float64)
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((len(pts), 2))
    f = np.zeros(2)
    for i, (e, pt) in enumerate(zip(np.asarray(elems).ravel(), pts)):
        pp = np.ascontiguousarray(pt)
        rc = lib().orc_visu_field(C.byref(P), u.ctypes.data_as(C.c_void_p), C.c_int(int(e)),
                                  pp.ctypes.data_as(C.c_void_p), C.c_double(time),
                                  f.ctypes.data_as(C.c_void_p))
        if rc < 0:
            raise RuntimeError(f"orc_visu_field failed with {rc}")
        out[i] = f
    return out

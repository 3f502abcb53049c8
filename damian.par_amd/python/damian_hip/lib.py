"""ctypes binding of libdamian_hip.so (include/damian_step.h).

Arguments are passed by pointer exactly as a Fortran ISO_C_BINDING caller
would pass them.  Any nonzero status raises DmpError with the library's
message; there is no fallback path -- if the shared library is missing the
import of this module fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .setup import Case

_PKG = Path(__file__).resolve().parents[2]
LIB_PATH = Path(os.environ.get("DMP_LIB", _PKG / "lib" / "libdamian_hip.so"))

DMP_FP64 = 0
DMP_FP32C = 1

EXPORTS = [
    "dmp_create", "dmp_destroy", "dmp_last_error", "dmp_set_mesh", "dmp_set_boundaries",
    "dmp_set_wave", "dmp_set_time", "dmp_set_receivers", "dmp_init", "dmp_run",
    "dmp_get_state", "dmp_get_step", "dmp_export_mesh", "dmp_set_tuning", "dmp_get_tuning",
    "dmp_autotune", "dmp_batch_create", "dmp_batch_destroy", "dmp_batch_run",
    "dmp_batch_get_state", "dmp_batch_set_tuning", "dmp_batch_autotune", "dmp_write_step",
    "dmp_write_mass", "dmp_last_run_ms", "dmp_batch_last_run_ms", "dmp_set_point_source",
    "dmp_write_mesh", "dmp_write_tparams", "dmp_write_rank", "dmp_run_noalloc",
]


class DmpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load_library(path=None):
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} not built: run `python damian.par_amd/build.py` (no CPU fallback exists)")
    return C.CDLL(str(p))


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def _i(v):
    return C.byref(C.c_int(int(v)))


_RUN_ARGTYPES = [C.c_void_p, C.POINTER(C.c_int), C.c_void_p, C.POINTER(C.c_int)]


def _address(a):
    """Address of a writable, C-contiguous numpy array's data (cheaper than
    a.ctypes.data)."""
    return C.addressof(C.c_char.from_buffer(a))


def _run_fn(name):
    """dmp_run / dmp_run_noalloc / dmp_batch_run with their argument types
    declared (handle, &nsteps, recv_out, &nsaved): a run call is on the
    caller's critical path, and declared types convert faster."""
    f = getattr(lib(), name)
    if f.argtypes is None:
        f.restype = C.c_int
        f.argtypes = _RUN_ARGTYPES
    return f


def _d(v):
    return C.byref(C.c_double(float(v)))


def _arr(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(C.c_void_p)


def last_error(h=None):
    buf = C.create_string_buffer(512)
    lib().dmp_last_error(h, buf, _i(512))
    return buf.raw.decode(errors="replace").rstrip()


def _check(rc, h=None):
    if rc != 0:
        raise DmpError(rc, last_error(h))


class Solver:
    """One site-response instance on one GPU."""

    def __init__(self, case: Case, precision=DMP_FP64, device=0, io_interval=1,
                 receivers=True, init=True):
        self.case = case
        self.precision = precision
        self.h = C.c_void_p()
        # run-call arguments, reused (a run call is on the caller's critical path)
        self._n, self._ns = C.c_int(0), C.c_int(0)
        self._pn, self._pns = C.byref(self._n), C.byref(self._ns)
        L = lib()
        _check(L.dmp_create(C.byref(self.h), _i(precision), _i(device)))
        m = case.mesh
        keep = []

        def A(a, dt):
            arr, p = _arr(a, dt)
            keep.append(arr)
            return p

        ne, nn = m.nelems, m.nnodes
        ediv = np.ascontiguousarray(m.ediv.astype(np.bool_))
        keep.append(ediv)
        if case.elem_props is not None:
            props = A(np.ascontiguousarray(case.elem_props, dtype=np.float64).ravel(), np.float64)
            _check(L.dmp_set_mesh(self.h, _i(ne), A(m.eptr, np.int32), A(m.eind, np.int32),
                                  ediv.ctypes.data_as(C.c_void_p), _i(nn),
                                  A(m.nodes.ravel(), np.float64), None, _i(0), None, props),
                   self.h)
        else:
            mc = case.model.matconst
            _check(L.dmp_set_mesh(self.h, _i(ne), A(m.eptr, np.int32), A(m.eind, np.int32),
                                  ediv.ctypes.data_as(C.c_void_p), _i(nn),
                                  A(m.nodes.ravel(), np.float64), A(m.ematprop, np.int32),
                                  _i(len(mc)), A(mc.ravel(), np.float64), None), self.h)
        _check(L.dmp_set_boundaries(self.h, A(case.absorbing, np.int32), _i(len(case.hanging)),
                                    A(case.hanging.ravel(), np.int32), _i(len(case.seismic)),
                                    A(case.seismic.ravel(), np.int32)), self.h)
        mdl = case.model
        if mdl.is_wave:
            left, right, top = mdl.box
            box = np.array([left, right, top], dtype=np.float64)
            org = np.array(mdl.worigin, dtype=np.float64)
            keep += [box, org]
            _check(L.dmp_set_wave(self.h, C.c_char_p(mdl.prob.encode()[:6]),
                                  org.ctypes.data_as(C.c_void_p), _d(mdl.wt0), _d(mdl.wangle),
                                  _d(mdl.wamp), _d(mdl.wfreq), _d(mdl.wvelo[0]),
                                  _d(mdl.wvelo[1]), box.ctypes.data_as(C.c_void_p)), self.h)
        if case.point is not None:
            kind, amp, sigma, start, nodes = case.point
            _check(L.dmp_set_point_source(self.h, C.c_char_p(kind.encode()[:6]), _d(amp), _d(sigma),
                                          _d(start), _i(len(nodes)), A(nodes, np.int32)), self.h)
        _check(L.dmp_set_time(self.h, _d(case.dt), _i(case.nsteps), _i(io_interval)), self.h)
        self.nrecv = len(case.recv_elem) if receivers else 0
        if self.nrecv:
            ff = np.ascontiguousarray(case.recv_ff.astype(np.bool_))
            keep.append(ff)
            _check(L.dmp_set_receivers(self.h, _i(self.nrecv), A(case.recv_elem, np.int32),
                                       A(case.recv_xy.ravel(), np.float64),
                                       ff.ctypes.data_as(C.c_void_p)), self.h)
        self.io_interval = io_interval
        self.step = 0
        if init:
            self.init()

    def init(self):
        _check(lib().dmp_init(self.h), self.h)

    def run(self, nsteps, record=True, noalloc=False):
        """Advance nsteps; returns receiver samples (nsamp, nrecv, 2) or None.
        noalloc: through dmp_run_noalloc (no allocation in the library)."""
        s = self.step
        I = self.io_interval
        if record and self.nrecv and I > 0:
            cap = (s + nsteps) // I - s // I + (1 if s == 0 else 0)
            # every returned row is written by the library (empty, not zeros:
            # short calls are on the caller's critical path)
            out = np.empty((max(cap, 1), self.nrecv, 2))
            ptr = _address(out)
        else:
            out, ptr = None, None
        self._n.value = int(nsteps)
        fn = _run_fn("dmp_run_noalloc" if noalloc else "dmp_run")
        _check(fn(self.h, self._pn, ptr, self._pns), self.h)
        self.step += nsteps
        if out is None:
            return None
        return out[:self._ns.value]

    def last_run_ms(self):
        ms = C.c_float(0)
        _check(lib().dmp_last_run_ms(self.h, C.byref(ms)), self.h)
        return float(ms.value)

    def state(self):
        nn = self.case.mesh.nnodes
        u = np.zeros(2 * nn)
        up = np.zeros(2 * nn)
        mf = np.zeros(2 * nn)
        _check(lib().dmp_get_state(self.h, u.ctypes.data_as(C.c_void_p),
                                   up.ctypes.data_as(C.c_void_p),
                                   mf.ctypes.data_as(C.c_void_p)), self.h)
        return u, up, mf

    def export_mesh(self):
        m = self.case.mesh
        eptr = np.zeros(m.nelems + 1, np.int32)
        eind = np.zeros(len(m.eind), np.int32)
        ediv = np.zeros(len(m.eind), np.bool_)
        hang = np.zeros(max(1, 3 * len(self.case.hanging)), np.int32)
        seis = np.zeros(max(1, 2 * len(self.case.seismic)), np.int32)
        _check(lib().dmp_export_mesh(self.h, eptr.ctypes.data_as(C.c_void_p),
                                     eind.ctypes.data_as(C.c_void_p),
                                     ediv.ctypes.data_as(C.c_void_p),
                                     hang.ctypes.data_as(C.c_void_p),
                                     seis.ctypes.data_as(C.c_void_p)), self.h)
        return (eptr, eind, ediv.astype(np.uint8), hang[:3 * len(self.case.hanging)].reshape(-1, 3),
                seis[:2 * len(self.case.seismic)].reshape(-1, 2))

    def set_tuning(self, block=256, xcd=1, graph=60, steps=4):
        """Threads per block, XCD-aware tile order, steps per captured graph
        (0: eager launches), steps per launch (1..8)."""
        t = (C.c_int * 4)(block, xcd, graph, steps)
        _check(lib().dmp_set_tuning(self.h, t), self.h)

    def tuning(self):
        t = (C.c_int * 4)()
        _check(lib().dmp_get_tuning(self.h, t), self.h)
        return tuple(t)

    def autotune(self, nsteps=60):
        t = (C.c_int * 4)()
        _check(lib().dmp_autotune(self.h, _i(nsteps), t), self.h)
        return tuple(t)

    def close(self):
        if self.h:
            lib().dmp_destroy(C.byref(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several initialised Solvers advanced together, one launch per step block."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        self.b = C.c_void_p()
        self._n, self._ns = C.c_int(0), C.c_int(0)
        self._pn, self._pns = C.byref(self._n), C.byref(self._ns)
        hs = (C.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        _check(lib().dmp_batch_create(C.byref(self.b), _i(len(self.solvers)), hs))
        self.nrecv = sum(s.nrecv for s in self.solvers)
        self.io_interval = self.solvers[0].io_interval
        self.step = 0

    def _err(self, rc):
        if rc != 0:
            raise DmpError(rc, last_error(None))

    def run(self, nsteps, record=True):
        s = self.step
        I = self.io_interval
        if record and self.nrecv and I > 0:
            cap = (s + nsteps) // I - s // I + (1 if s == 0 else 0)
            # every returned row is written by the library (empty, not zeros:
            # short calls are on the caller's critical path)
            out = np.empty((max(cap, 1), self.nrecv, 2))
            ptr = _address(out)
        else:
            out, ptr = None, None
        self._n.value = int(nsteps)
        self._err(_run_fn("dmp_batch_run")(self.b, self._pn, ptr, self._pns))
        self.step += nsteps
        return None if out is None else out[:self._ns.value]

    def last_run_ms(self):
        ms = C.c_float(0)
        self._err(lib().dmp_batch_last_run_ms(self.b, C.byref(ms)))
        return float(ms.value)

    def state(self, q):
        nn = self.solvers[q].case.mesh.nnodes
        u = np.zeros(2 * nn)
        up = np.zeros(2 * nn)
        self._err(lib().dmp_batch_get_state(self.b, _i(q), u.ctypes.data_as(C.c_void_p),
                                            up.ctypes.data_as(C.c_void_p)))
        return u, up

    def set_tuning(self, block=256, xcd=1, graph=60, steps=4):
        t = (C.c_int * 4)(block, xcd, graph, steps)
        self._err(lib().dmp_batch_set_tuning(self.b, t))

    def autotune(self, nsteps=60):
        t = (C.c_int * 4)()
        self._err(lib().dmp_batch_autotune(self.b, _i(nsteps), t))
        return tuple(t)

    def close(self):
        if self.b:
            lib().dmp_batch_destroy(C.byref(self.b))
            self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_step(path, mode, step, u):
    b = str(path).encode()
    u = np.ascontiguousarray(u, dtype=np.float64)
    rc = lib().dmp_write_step(_i(len(b)), C.c_char_p(b), C.c_char_p(mode.encode()), _i(step),
                              _i(len(u)), u.ctypes.data_as(C.c_void_p))
    _check(rc)


def write_mass(path, mass):
    b = str(path).encode()
    m = np.ascontiguousarray(mass, dtype=np.float64)
    rc = lib().dmp_write_mass(_i(len(b)), C.c_char_p(b), _i(len(m)), m.ctypes.data_as(C.c_void_p))
    _check(rc)


def write_mesh(path, eptr, eind, ediv, nodes):
    """mesh.NNN.bin in the reference's format (1-based eptr / eind, ediv
    logical*1, nodes (nnodes, 2))."""
    b = str(path).encode()
    eptr = np.ascontiguousarray(eptr, dtype=np.int32)
    eind = np.ascontiguousarray(eind, dtype=np.int32)
    ediv = np.ascontiguousarray(ediv, dtype=np.uint8)
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    rc = lib().dmp_write_mesh(_i(len(b)), C.c_char_p(b), _i(len(eptr) - 1),
                              eptr.ctypes.data_as(C.c_void_p), eind.ctypes.data_as(C.c_void_p),
                              ediv.ctypes.data_as(C.c_void_p), _i(nodes.size // 2),
                              nodes.ctypes.data_as(C.c_void_p))
    _check(rc)


def write_tparams(path, deltat, iointerval):
    b = str(path).encode()
    rc = lib().dmp_write_tparams(_i(len(b)), C.c_char_p(b), C.byref(C.c_double(deltat)),
                                 _i(iointerval))
    _check(rc)


def write_rank(path, rank):
    b = str(path).encode()
    _check(lib().dmp_write_rank(_i(len(b)), C.c_char_p(b), _i(rank)))

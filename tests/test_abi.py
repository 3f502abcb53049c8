"""C-ABI surface (no GPU compute): exports, error path, output formats
byte-identical with the reference's io.cpp, device-code guards."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from damian_hip import lib

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "damian_step.h").read_text()
    return sorted(set(re.findall(r"^int (dmp_\w+)\(", text, flags=re.M)))


def test_exports_every_declared_symbol():
    L = lib.lib()
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(lib.EXPORTS)


def test_create_without_device_fails_loudly():
    h = C.c_void_p()
    rc = lib.lib().dmp_create(C.byref(h), lib._i(0), lib._i(0))
    if rc == 0:        # a GPU is present on this machine
        lib.lib().dmp_destroy(C.byref(h))
        pytest.skip("HIP device present")
    assert rc == -3
    assert "device" in lib.last_error()


def test_argument_checks():
    L = lib.lib()
    h = C.c_void_p()
    assert L.dmp_create(C.byref(h), lib._i(7), lib._i(0)) == -1    # bad precision
    assert L.dmp_run(None, lib._i(1), None, None) == -1
    assert L.dmp_init(None) == -1


def _ref_io():
    import oracle_ctypes as O
    if not O.REF_IO_PATH.exists():
        pytest.skip("reference io.cpp not built (reference tree absent)")
    return C.CDLL(str(O.REF_IO_PATH))


def test_step_file_matches_reference_writer(tmp_path):
    ref = _ref_io()
    rng = np.random.default_rng(3)
    u1, u2 = rng.standard_normal(17), rng.standard_normal(17)
    ours = tmp_path / "ours.bin"
    lib.write_step(ours, "n", 0, u1)
    lib.write_step(ours, "o", 1, u2)
    theirs = str(tmp_path / "theirs.bin").encode()
    for mode, step, u in ((b"n", 0, u1), (b"o", 1, u2)):
        ref.save_binstep(lib._i(len(theirs)), C.c_char_p(theirs), C.c_char_p(mode), lib._i(step),
                         lib._i(len(u)), np.ascontiguousarray(u).ctypes.data_as(C.c_void_p))
    assert ours.read_bytes() == Path(theirs.decode()).read_bytes()
    # and the reference reader gets the second record back
    pu = C.c_void_p()
    nd, st = C.c_int(0), C.c_int(0)
    name = str(ours).encode()
    ref.load_binstep(lib._i(len(name)), C.c_char_p(name), lib._i(1), C.byref(nd), C.byref(pu),
                     C.byref(st))
    assert st.value == 0 and nd.value == 17
    got = np.ctypeslib.as_array(C.cast(pu, C.POINTER(C.c_double)), shape=(17,))
    np.testing.assert_array_equal(got, u2)


def test_mass_file_matches_reference_writer(tmp_path):
    ref = _ref_io()
    m = np.linspace(1.0, 2.0, 11)
    ours = tmp_path / "m1.bin"
    lib.write_mass(ours, m)
    theirs = str(tmp_path / "m2.bin").encode()
    ref.save_binmass(lib._i(len(theirs)), C.c_char_p(theirs), lib._i(len(m)),
                     m.ctypes.data_as(C.c_void_p))
    assert ours.read_bytes() == Path(theirs.decode()).read_bytes()


def test_mesh_file_matches_reference_writer(tmp_path):
    """mesh.NNN.bin of a quadtree mesh with hanging nodes.  The reference writer
    stores eptr[nelems] entries of eind / ediv, one past the data
    (io.cpp:354, 357); it is given arrays with a spare zero so both files are
    defined, and its reader must get the mesh back."""
    ref = _ref_io()
    from damian_hip import cases
    m = cases.f2(nsteps=1).mesh
    eptr = np.ascontiguousarray(m.eptr, dtype=np.int32)
    eind = np.ascontiguousarray(m.eind, dtype=np.int32)
    ediv = np.ascontiguousarray(m.ediv, dtype=np.uint8)
    nodes = np.ascontiguousarray(m.nodes, dtype=np.float64)
    assert int(np.count_nonzero(ediv)) > 0
    ours = tmp_path / "mesh1.bin"
    lib.write_mesh(ours, eptr, eind, ediv, nodes)
    eind_pad = np.append(eind, np.int32(0))
    ediv_pad = np.append(ediv, np.uint8(0))
    theirs = str(tmp_path / "mesh2.bin").encode()
    ref.save_binmesh(lib._i(len(theirs)), C.c_char_p(theirs), lib._i(len(eptr) - 1),
                     eptr.ctypes.data_as(C.c_void_p), eind_pad.ctypes.data_as(C.c_void_p),
                     ediv_pad.ctypes.data_as(C.c_void_p), lib._i(len(nodes)),
                     nodes.ctypes.data_as(C.c_void_p))
    assert ours.read_bytes() == Path(theirs.decode()).read_bytes()
    ne, nn = C.c_int(0), C.c_int(0)
    pe, pi, pd, pn = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    name = str(ours).encode()
    ref.load_binmesh(lib._i(len(name)), C.c_char_p(name), C.byref(ne), C.byref(pe), C.byref(pi),
                     C.byref(pd), C.byref(nn), C.byref(pn))
    assert ne.value == len(eptr) - 1 and nn.value == len(nodes)
    got_e = np.ctypeslib.as_array(C.cast(pe, C.POINTER(C.c_int)), shape=(len(eptr),))
    got_i = np.ctypeslib.as_array(C.cast(pi, C.POINTER(C.c_int)), shape=(len(eind),))
    got_d = np.ctypeslib.as_array(C.cast(pd, C.POINTER(C.c_uint8)), shape=(len(ediv),))
    got_n = np.ctypeslib.as_array(C.cast(pn, C.POINTER(C.c_double)), shape=(nodes.size,))
    np.testing.assert_array_equal(got_e, eptr)
    np.testing.assert_array_equal(got_i, eind)
    np.testing.assert_array_equal(got_d, ediv)
    np.testing.assert_array_equal(got_n, nodes.ravel())


def test_tparams_file_matches_reference_writer(tmp_path):
    ref = _ref_io()
    ours = tmp_path / "t1.bin"
    lib.write_tparams(ours, 2.5e-4, 40)
    theirs = str(tmp_path / "t2.bin").encode()
    ref.save_bintparams(lib._i(len(theirs)), C.c_char_p(theirs), C.byref(C.c_double(2.5e-4)),
                        lib._i(40))
    assert ours.read_bytes() == Path(theirs.decode()).read_bytes()


def test_rank_file_matches_reference_writer(tmp_path):
    """ranks.bin after several save_rank_file calls, and the reference reader
    walks it (load_rank_file advances its cursor)."""
    ref = _ref_io()
    ours = tmp_path / "r1.bin"
    theirs = str(tmp_path / "r2.bin").encode()
    for r in (0, 3, 1):
        lib.write_rank(ours, r)
        ref.save_rank_file(lib._i(len(theirs)), C.c_char_p(theirs), lib._i(r))
    assert ours.read_bytes() == Path(theirs.decode()).read_bytes()
    name = str(ours).encode()
    got = []
    for _ in range(3):
        n, r = C.c_int(0), C.c_int(-1)
        ref.load_rank_file(lib._i(len(name)), C.c_char_p(name), C.byref(n), C.byref(r))
        assert n.value == 3
        got.append(r.value)
    assert got == [0, 3, 1]


def test_device_code_has_no_scratch_and_no_packed_fp32(tmp_path):
    """The step kernels keep everything in registers/LDS (no private
    segment) and contain no packed-FP32 instructions."""
    out = tmp_path / "k.s"
    cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only",
           "-S", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", str(ROOT / "include"), "-I", str(ROOT / "damian.par_amd" / "csrc"),
           str(ROOT / "damian.par_amd" / "csrc" / "dmp_kernels.hip"), "-o", str(out)]
    if not Path(cmd[0]).exists():
        pytest.skip("hipcc not available")
    subprocess.run(cmd, check=True, capture_output=True)
    asm = out.read_text()
    assert re.findall(r"\.private_segment_fixed_size:\s+(\d+)", asm) and \
        all(v == "0" for v in re.findall(r"\.private_segment_fixed_size:\s+(\d+)", asm))
    assert not re.search(r"v_pk_(add|mul|fma)_f32", asm)


def test_sheet_file_matches_reference_writer(tmp_path):
    from damian_hip import sheets
    ref = _ref_io()
    rng = np.random.default_rng(5)
    pts = rng.uniform(0, 100, (3, 2))
    samples = rng.standard_normal((4, 3, 2))
    ours = tmp_path / "ours.txt"
    sheets.write_sheet(ours, pts, samples, 0.1)
    theirs = str(tmp_path / "theirs.txt").encode()
    n = lib._i(len(theirs))
    p = np.ascontiguousarray(pts.ravel())
    ref.write_sheet_points(n, C.c_char_p(theirs), lib._i(p.size), p.ctypes.data_as(C.c_void_p))
    for row in samples:
        r = np.ascontiguousarray(row.ravel())
        ref.write_sheet_step(n, C.c_char_p(theirs), lib._i(r.size), r.ctypes.data_as(C.c_void_p))
    samp = C.c_double(0.1)
    ref.write_sheet_close(n, C.c_char_p(theirs), C.byref(samp))
    assert ours.read_text() == Path(theirs.decode()).read_text()

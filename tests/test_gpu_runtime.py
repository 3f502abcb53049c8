"""Runtime behaviour of the step library on the GPU: determinism across
launch shapes and graph capture, step-block splitting, batching, integer
array export, error paths."""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from damian_hip import cases, lib
    return cases, lib


@pytest.mark.parametrize("prec", [0, 1])
def test_launch_shapes_bitwise(mods, prec):
    """Block size, XCD remap, steps per launch and graph capture do not change
    a single bit."""
    cases, lib = mods
    case = cases.f2(nsteps=130)
    ref = None
    for tun in [(256, 1, 0, 1), (64, 0, 0, 2), (128, 1, 60, 4), (256, 0, 12, 3), (64, 1, 7, 4),
                (256, 1, 60, 4), (128, 0, 0, 1), (100, 1, 60, 3), (216, 0, 0, 2), (1000, 1, 60, 1),
                (64, 1, 60, 6), (64, 0, 48, 8)]:
        print(f"prec={prec} tuning={tun}", flush=True)
        s = lib.Solver(case, precision=prec)
        s.set_tuning(*tun)
        rec = s.run(130)
        u, up, _ = s.state()
        s.close()
        if ref is None:
            ref = (rec, u, up)
        else:
            assert np.array_equal(rec, ref[0])
            assert np.array_equal(u, ref[1])
            assert np.array_equal(up, ref[2])


@pytest.mark.parametrize("name", ["f3p", "f3s", "pointn", "pointd"])
@pytest.mark.parametrize("prec", [0, 1])
def test_launch_shapes_bitwise_sources(mods, name, prec):
    """Plane-wave forcing and point sources through the one-step kernel (side
    records in registers) and the multi-step kernel (records in LDS): the same
    bits."""
    cases, lib = mods
    case = cases.BUILDERS[name]()
    n = min(case.nsteps, 150)
    ref = None
    for tun in [(256, 1, 0, 1), (64, 0, 0, 2), (128, 1, 60, 4), (100, 1, 60, 3), (64, 1, 30, 1)]:
        print(f"case={name} prec={prec} tuning={tun}", flush=True)
        s = lib.Solver(case, precision=prec)
        s.set_tuning(*tun)
        rec = s.run(n)
        u, up, _ = s.state()
        s.close()
        if ref is None:
            ref = (rec, u, up)
        else:
            assert np.array_equal(rec, ref[0])
            assert np.array_equal(u, ref[1])
            assert np.array_equal(up, ref[2])


def test_tile_cuts_bitwise(mods):
    """Tiles cut at equal work (few tiles, several steps per launch) and tiles
    of fixed size (many tiles, or one step per launch) on a mesh with hanging
    nodes give the same bits."""
    cases, lib = mods
    case = cases.c2(nsteps=90)
    ref = None
    for tun in [(209, 1, 60, 3), (209, 1, 60, 4), (32, 1, 0, 2), (32, 1, 0, 1), (128, 1, 60, 2)]:
        print(f"tuning={tun}", flush=True)
        s = lib.Solver(case, precision=0)
        s.set_tuning(*tun)
        rec = s.run(90)
        u, up, _ = s.state()
        s.close()
        if ref is None:
            ref = (rec, u, up)
        else:
            assert np.array_equal(rec, ref[0])
            assert np.array_equal(u, ref[1])
            assert np.array_equal(up, ref[2])


@pytest.mark.parametrize("tuning", [(128, 1, 0, 3), (128, 1, 12, 3), (256, 0, 30, 1)])
@pytest.mark.parametrize("io", [1, 3])
def test_block_splitting(mods, tuning, io):
    """run(a) + run(b) == run(a + b), receivers concatenate, for eager and
    graph-launched blocks, calls of zero steps and io_interval > 1 (the last
    step of a call is sampled from HBM, the first launch of the next call
    samples it again)."""
    cases, lib = mods
    case = cases.f3("SV", nsteps=200)
    print(f"tuning={tuning} io={io}", flush=True)
    s1 = lib.Solver(case, io_interval=io)
    s1.set_tuning(*tuning)
    r_all = s1.run(200)
    u1, up1, _ = s1.state()
    s2 = lib.Solver(case, io_interval=io)
    s2.set_tuning(*tuning)
    parts = [s2.run(n) for n in (0, 37, 1, 0, 62, 24, 76)]
    u2, up2, _ = s2.state()
    r_cat = np.concatenate([p for p in parts if p is not None and len(p)], axis=0)
    assert r_all.shape[0] == 200 // io + 1
    assert np.array_equal(r_all, r_cat)
    assert np.array_equal(u1, u2) and np.array_equal(up1, up2)


@pytest.mark.parametrize("io", [1, 3])
def test_receiver_chunks(mods, io):
    """Receiver output leaves the device in chunks of about 2 MB while a run
    continues: a 2000-step canyon run (201 receivers, several chunks) equals
    the same steps split into runs of 1, 700 and 1299 steps."""
    cases, lib = mods
    case = cases.c2(nsteps=2000)
    a = lib.Solver(case, precision=1, io_interval=io)
    r_all = a.run(2000)
    u_a, _, _ = a.state()
    a.close()
    b = lib.Solver(case, precision=1, io_interval=io)
    parts = [b.run(n) for n in (1, 700, 1299)]
    u_b, _, _ = b.state()
    b.close()
    r_cat = np.concatenate([p for p in parts if p is not None and len(p)], axis=0)
    assert r_all.shape == (2000 // io + 1, 201, 2)
    assert np.array_equal(r_all, r_cat)
    assert np.array_equal(u_a, u_b)
    assert np.abs(r_all).max() > 0 and np.isfinite(r_all).all()
    # an instance set up for 100 steps has a sample ring of 64 slots: a longer
    # run streams its samples out through it
    case.nsteps = 100
    c = lib.Solver(case, precision=1, io_interval=io)
    r_grow = c.run(2000)
    c.close()
    assert np.array_equal(r_grow, r_all)


def test_autotune_restores_state(mods):
    cases, lib = mods
    case = cases.f1(nsteps=120)
    a = lib.Solver(case)
    ra = a.run(60)
    t = a.autotune(30)
    assert 32 <= t[0] <= 1024 and t[1] in (0, 1) and 0 <= t[2] <= 4096 and 1 <= t[3] <= 8
    ra2 = a.run(60)
    b = lib.Solver(case)
    rb = b.run(120)
    assert np.array_equal(np.concatenate([ra, ra2]), rb)


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_matches_single(mods, prec):
    """Uneven instances packed in one launch per step give the same bits as
    separate runs."""
    cases, lib = mods
    cs = [cases.f1(nsteps=150), cases.f2(nsteps=150), cases.f3("P", nsteps=150)]
    singles = []
    for c in cs:
        s = lib.Solver(c, precision=prec)
        r = s.run(150)
        u, up, _ = s.state()
        singles.append((r, u, up))
        s.close()
    solvers = [lib.Solver(c, precision=prec) for c in cs]
    B = lib.Batch(solvers)
    rb = B.run(150)
    off = 0
    for q, (r, u, up) in enumerate(singles):
        nr = r.shape[1]
        assert np.array_equal(rb[:, off:off + nr], r)
        off += nr
        ub, upb = B.state(q)
        assert np.array_equal(ub, u)
        assert np.array_equal(upb, up)


@pytest.mark.parametrize("name", ["f1", "f2"])
def test_export_mesh_bit_exact(mods, name):
    cases, lib = mods
    case = cases.BUILDERS[name]()
    s = lib.Solver(case)
    eptr, eind, ediv, hang, seis = s.export_mesh()
    m = case.mesh
    assert np.array_equal(eptr, m.eptr)
    assert np.array_equal(eind, m.eind)
    assert np.array_equal(ediv, m.ediv)
    assert np.array_equal(hang, case.hanging)
    assert np.array_equal(seis, case.seismic)


def test_critical_sv_rejected(mods):
    """SV at exactly the critical angle is rejected like pwaves.f90:87-89."""
    cases, lib = mods
    with pytest.raises(lib.DmpError) as ei:
        cases_ = cases.f3("SV", nsteps=10)
        cases_.model.wangle = math.asin(0.5)
        lib.Solver(cases_)
    assert ei.value.code == -6


def test_bad_inputs(mods):
    cases, lib = mods
    case = cases.f1(nsteps=10)
    case.mesh.eind = case.mesh.eind.copy()
    case.mesh.eind[5] = 10 ** 6          # node id out of range
    with pytest.raises(lib.DmpError) as ei:
        lib.Solver(case)
    assert ei.value.code == -1
    case = cases.f1(nsteps=10)
    nodes = case.mesh.nodes.copy()
    nodes[case.mesh.eind[2] - 1, 0] += 1.0   # distort element 1
    case.mesh.nodes = nodes
    with pytest.raises(lib.DmpError) as ei:
        lib.Solver(case)
    assert ei.value.code == -4


def test_zero_and_single_steps(mods):
    cases, lib = mods
    case = cases.f1(nsteps=5)
    s = lib.Solver(case)
    r0 = s.run(0)
    assert r0.shape[0] == 1                 # step-0 sample only
    r1 = s.run(1)
    assert r1.shape[0] == 1
    u, up, _ = s.state()
    assert np.isfinite(u).all()


def test_unsupported_shape_keeps_previous(mods):
    """A launch shape the mesh cannot use is refused (DMP_EUNSUPPORTED) and the
    previous shape stays in force, results unchanged."""
    cases, lib = mods
    case = cases.c2(nsteps=60)
    ref = lib.Solver(case)
    ref.set_tuning(128, 1, 0, 2)
    r_ref = ref.run(40)
    ref.close()
    s = lib.Solver(case)
    s.set_tuning(128, 1, 0, 2)
    with pytest.raises(lib.DmpError) as ei:
        s.set_tuning(1024, 1, 0, 8)
    assert ei.value.code == -4
    assert s.tuning() == (128, 1, 0, 2)
    assert np.array_equal(s.run(40), r_ref)
    s.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_deep_launches_many_tiles(mods, prec):
    """Launches of 5 to 8 steps (those whose tiles fit) on a mesh with far more tiles than run at once
    (canyon quadtree in tiles of 32 and 64 nodes): every launch reads two
    state levels and writes two others while workgroups of the same launch
    start at any time, so the state ring must keep the four apart.  Bitwise
    equal to one step per launch."""
    cases, lib = mods
    case = cases.c2(nsteps=120)
    ref = lib.Solver(case, precision=prec)
    ref.set_tuning(64, 1, 0, 1)
    r_ref = ref.run(120)
    u_ref, up_ref, _ = ref.state()
    ref.close()
    ran = []
    for tun in [(32, 1, 0, 5), (32, 0, 0, 6), (32, 1, 40, 7), (32, 1, 0, 8), (64, 1, 0, 6),
                (64, 1, 48, 8)]:
        s = lib.Solver(case, precision=prec)
        try:
            s.set_tuning(*tun)
        except lib.DmpError as e:
            assert e.code == -4, e   # the layout does not fit: DMP_EUNSUPPORTED
            s.close()
            continue
        print(f"prec={prec} tuning={tun}", flush=True)
        rec = s.run(120)
        u, up, _ = s.state()
        s.close()
        assert np.array_equal(rec, r_ref), tun
        assert np.array_equal(u, u_ref) and np.array_equal(up, up_ref), tun
        ran.append(tun)
    # 5 steps per launch: a six-level ring would have written u_{t0+5} over u_{t0-1}
    assert any(t[3] >= 5 for t in ran), ran


@pytest.mark.parametrize("tuning", [(128, 1, 0, 3), (128, 1, 12, 3)])
def test_no_receivers(mods, tuning):
    """An instance without receivers (the drop-in's use): runs of any length,
    eager and graph-launched, give the same fields as with receivers."""
    cases, lib = mods
    case = cases.f2(nsteps=100)
    a = lib.Solver(case, receivers=False)
    a.set_tuning(*tuning)
    for n in (0, 7, 30, 63):
        assert a.run(n) is None
    u_a, up_a, _ = a.state()
    a.close()
    b = lib.Solver(case)
    b.set_tuning(*tuning)
    b.run(100)
    u_b, up_b, _ = b.state()
    b.close()
    assert np.array_equal(u_a, u_b) and np.array_equal(up_a, up_b)

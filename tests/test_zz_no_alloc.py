"""dmp_run_noalloc (the allocation-free run call) against dmp_run through
the C-ABI: two calls in a row, the second shorter, then another, smaller
instance on the same GPU sharing the library's scratch array; results and
state bit-identical; receivers in no element read zero; a call whose samples
exceed the scratch is refused and leaves the instance where it was."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from damian_hip import cases, lib
    return cases, lib


def pair(lib, case, prec, tuning):
    a = lib.Solver(case, precision=prec)
    b = lib.Solver(case, precision=prec)
    a.set_tuning(*tuning)
    b.set_tuning(*tuning)
    return a, b


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_run(mods, prec):
    cases, lib = mods
    big = cases.c2(nsteps=400)
    small = cases.f2(nsteps=100)
    print(f"noalloc prec={prec}: c2 {big.nelems} elements, then f2 {small.nelems}", flush=True)
    for case, calls, tuning in ((big, (150, 61), (209, 1, 60, 3)), (small, (70, 23), (128, 1, 12, 2))):
        a, b = pair(lib, case, prec, tuning)
        for n in calls:
            ra = a.run(n)
            rb = b.run(n, noalloc=True)
            assert ra.shape == rb.shape and ra.shape[0] > 0
            assert np.array_equal(ra, rb), (case.name, n)
            sa, sb = a.state(), b.state()
            for x, y in zip(sa, sb):
                assert np.array_equal(x, y), (case.name, n)
        assert np.abs(ra).max() > 0
        a.close()
        b.close()


def test_noalloc_unlocated_receivers_read_zero(mods):
    """Receivers in no element (recv_elem 0: another rank's) are sampled as
    zero by both calls, also after a larger instance has filled the scratch
    where their samples would sit."""
    import copy
    cases, lib = mods
    big = cases.c2(nsteps=200)
    a = lib.Solver(big, precision=1)
    a.set_tuning(209, 1, 60, 3)
    assert np.abs(a.run(120, noalloc=True)).max() > 0
    a.close()
    case = copy.deepcopy(cases.f2(nsteps=100))
    case.recv_elem = np.array(case.recv_elem, copy=True)
    gone = np.arange(len(case.recv_elem)) % 3 == 1
    case.recv_elem[gone] = 0
    print(f"noalloc unlocated: {int(gone.sum())} of {len(gone)} receivers", flush=True)
    a, b = pair(lib, case, 1, (128, 1, 12, 2))
    for n in (70, 23):
        ra = a.run(n)
        rb = b.run(n, noalloc=True)
        assert np.array_equal(ra, rb), n
        assert not ra[:, gone].any()
        assert np.abs(ra[:, ~gone]).max() > 0
    a.close()
    b.close()


def test_noalloc_refuses_calls_beyond_the_scratch(mods):
    cases, lib = mods
    case = cases.f2(nsteps=100)
    s = lib.Solver(case)
    s.run(10, noalloc=True)
    u0, up0, _ = s.state()
    slots = (16 << 20) // (16 * len(case.recv_elem))
    with pytest.raises(lib.DmpError) as e:
        s.run(slots, noalloc=True)
    assert e.value.code == -1 and "scratch" in str(e.value)
    s.step -= slots   # the binding counts steps itself; the library did not advance
    u1, up1, _ = s.state()
    assert np.array_equal(u0, u1) and np.array_equal(up0, up1)
    r = s.run(5, noalloc=True)   # still usable
    assert r.shape == (5, len(case.recv_elem), 2)
    s.close()

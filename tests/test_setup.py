"""Driver-side setup: problem.in grammar on the reference's own inputs,
material conversion, grid sizing and the config-1 workload numbers."""
import math

import numpy as np
import pytest

from damian_hip import cases, setup as S

CURRENT = [f"{g}{i}" for g in ("canyon", "valley", "rectan", "varect") for i in range(1, 7)]


def test_parse_reference_inputs(reference_tree):
    for d in CURRENT:
        text = (reference_tree / d / "problem.in").read_text()
        pf = S.parse_problem(text)
        mdl = S.build_model(pf)
        assert pf.prob in ("WAVESV", "WAVEP ")
        assert len(pf.pparams) == 9
        # nu = 1/3 materials: vp = 2 vs (formulation.f90:85-97)
        for mu, lam, rho, vp, vs in mdl.matconst:
            assert abs(vp / vs - 2.0) < 1e-6
            assert abs(mu - rho * vs ** 2) < 1e-6 * mu
        assert mdl.totcols > 1 and mdl.totrows > 1


def test_canyon1_mesh_follows_reference_rules(reference_tree):
    """canyon1: one material, so the reference mesh is the initial grid minus
    the elements whose four corners lie inside the canyon."""
    pf = S.parse_problem((reference_tree / "canyon1" / "problem.in").read_text())
    mdl = S.build_model(pf)
    assert mdl.blength == 25.0 and mdl.totcols == 1601 and mdl.totrows == 841
    mesh = S.reference_uniform_mesh(mdl)
    full = (mdl.totcols - 1) * (mdl.totrows - 1)
    cut = full - mesh.nelems
    # brute force: 25 m cells whose four corners are within 1000 m of (20000, 0)
    xs = np.arange(18950.0, 21050.0, 25.0)
    zs = np.arange(0.0, 1050.0, 25.0)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    ins = np.hypot(X - 20000.0, Z) <= 1000.0
    expect = int((ins[:-1, :-1] & ins[1:, :-1] & ins[1:, 1:] & ins[:-1, 1:]).sum())
    assert cut == expect
    c = mesh.corners()
    inside = np.hypot(mesh.nodes[c - 1, 0] - 20000.0, mesh.nodes[c - 1, 1]) <= 1000.0
    assert not inside.all(axis=1).any()
    # node numbering: rank of the initial global id among surviving nodes
    ix = np.rint(mesh.nodes[:, 0] / 25.0).astype(int)
    iz = np.rint(mesh.nodes[:, 1] / 25.0).astype(int)
    gnode = ix * mdl.totrows + iz + 1
    assert np.all(np.diff(gnode) > 0)


def test_material_conversions():
    m = S.setup_materials([("E", "v")], [(2.0e9, 0.25, 2000.0)])[0]
    mu, lam, rho, vp, vs = m
    assert abs(mu - 2.0e9 / 2.5) < 1e-3
    assert abs(lam - 2.0e9 * 0.25 / (1.25 * 0.5)) < 1e-3
    assert abs(vs - math.sqrt(mu / rho)) < 1e-9
    m = S.setup_materials([("p", "s")], [(1000.0, 500.0, 2000.0)])[0]
    assert tuple(m) == (2000.0 * 500.0 ** 2, 2000.0 * 1000.0 ** 2 - 2 * 2000.0 * 500.0 ** 2, 2000.0,
                        1000.0, 500.0)


def test_config1_numbers():
    c = cases.c1()
    assert c.nelems == 40000 and c.mesh.nnodes == 40501
    assert c.dt == 0.002 and c.model.nsteps == 5000 and c.nsteps == 5000
    assert len(c.recv_elem) == 101 and (c.recv_elem > 0).all()
    np.testing.assert_allclose(c.recv_xy[:, 0], np.arange(101) * 20.0)
    assert len(c.hanging) == 0
    # plane-wave box 1000 m x 300 m around x0 = 1000 (model.f90:988-989, 205-207)
    assert c.model.box == (495.0, 1500.0, 300.0)


def test_config2_numbers():
    c = cases.c2(nsteps=10)
    h = c.mesh.elength()
    assert set(np.unique(h)) == {2.5, 5.0, 10.0, 20.0}
    assert len(c.hanging) > 0
    assert len(c.recv_elem) == 201 and (c.recv_elem > 0).all()
    # refined band: every element touching the free surface is 2.5 m
    top = c.mesh.nodes[c.mesh.corners()[:, 0] - 1, 1] == 0.0
    assert np.all(h[top] == 2.5)
    assert c.dt == 2.5e-4


def test_von_karman_field():
    """For Hurst exponent 1/2 the von Karman correlation function is
    exp(-r / a): the generated field must reproduce it at lags a/2, a, 2a
    (both axes), with zero mean, unit variance, and be fixed by its seed."""
    a = 16.0
    g = cases.von_karman(2048, 2048, 1.0, a, 0.5, seed=11)
    assert abs(g.mean()) < 1e-12 and abs(g.std() - 1.0) < 1e-12
    for r in (8, 16, 32):
        acf = (np.mean(g * np.roll(g, r, axis=1)) + np.mean(g * np.roll(g, r, axis=0))) / 2
        assert abs(acf - math.exp(-r / a)) < 0.03, (r, acf)
    # a rougher field decorrelates faster: Hurst 0.25 at lag a
    g4 = cases.von_karman(2048, 2048, 1.0, a, 0.25, seed=11)
    acf4 = np.mean(g4 * np.roll(g4, 16, axis=1))
    assert acf4 < (np.mean(g * np.roll(g, 16, axis=1)))
    assert np.array_equal(cases.von_karman(64, 32, 2.5, 40.0, 0.5, seed=3003),
                          cases.von_karman(64, 32, 2.5, 40.0, 0.5, seed=3003))
    assert not np.array_equal(cases.von_karman(64, 32, 2.5, 40.0, 0.5, seed=3003),
                              cases.von_karman(64, 32, 2.5, 40.0, 0.5, seed=3004))


def test_config3_numbers():
    """Config 3: 10 points per wavelength at 8 Hz in the sediment (Vs 250 ->
    h <= 3.125 m), lognormal Vs with 10 % spread in the valley only."""
    c = cases.c3(nsteps=10)
    h = c.mesh.elength()
    sed = c.mesh.ematprop == 2
    assert set(np.unique(h)) == {2.5, 5.0, 10.0}
    assert np.all(h[sed] <= 250.0 / (10 * 8.0))
    assert np.all(h[~sed] <= 1500.0 / (10 * 8.0) + 1e-9)
    vs, vp = c.elem_props[:, 4], c.elem_props[:, 3]
    assert np.all(vs[~sed] == 1500.0) and np.all(vp[~sed] == 3000.0) and np.all(vp[sed] == 600.0)
    assert abs(np.log(vs[sed] / 250.0).std() - 0.1) < 0.02
    assert np.all(c.elem_props[:, 1] > 0)            # lambda > 0 everywhere
    assert len(c.hanging) > 0 and len(c.recv_elem) == 102
    # the step follows the reference's rule: finest h / Vp / sperlen
    assert c.dt == 2.5 / 3000.0 / 5.0


def test_config5_numbers():
    """Config 5: every element has 10 points per wavelength at 5 Hz for the
    Vs at its top, Vs follows the depth gradient times a 7 % lognormal field,
    Vp = 2 Vs and rho = 1700 + 0.2 Vs, 1001 receivers, the step rule applied
    with each element's own Vp."""
    c = cases.c5(nsteps=10)
    m = c.mesh
    h = m.elength()
    cc = m.corners() - 1
    ztop = m.nodes[cc, 1].min(axis=1)
    zc = m.nodes[cc, 1].mean(axis=1)
    assert set(np.unique(h)) == {3.125, 6.25, 12.5, 25.0, 50.0}
    assert np.all(h <= (300.0 + 2700.0 * ztop / 5000.0) / 50.0 + 1e-9)
    mu, lam, rho, vp, vs = c.elem_props.T
    assert abs(np.log(vs / (300.0 + 2700.0 * zc / 5000.0)).std() - 0.07) < 0.015
    np.testing.assert_allclose(vp, 2.0 * vs)
    np.testing.assert_allclose(rho, 1700.0 + 0.2 * vs)
    assert np.all(lam > 0)
    assert len(c.recv_elem) == 1001 and len(c.hanging) > 0
    assert c.dt == float(np.min(h / vp)) / 5.0


def test_quadtree_is_two_to_one_balanced():
    c = cases.c2(nsteps=10)
    m = c.mesh
    # every hanging node sits at the midpoint of its anchors, anchors are not hanging
    hn = set(c.hanging[:, 0])
    for h, a, b in c.hanging:
        np.testing.assert_allclose(m.nodes[h - 1], (m.nodes[a - 1] + m.nodes[b - 1]) / 2)
        assert a not in hn and b not in hn

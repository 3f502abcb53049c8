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


def test_quadtree_is_two_to_one_balanced():
    c = cases.c2(nsteps=10)
    m = c.mesh
    # every hanging node sits at the midpoint of its anchors, anchors are not hanging
    hn = set(c.hanging[:, 0])
    for h, a, b in c.hanging:
        np.testing.assert_allclose(m.nodes[h - 1], (m.nodes[a - 1] + m.nodes[b - 1]) / 2)
        assert a not in hn and b not in hn

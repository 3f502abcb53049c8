"""Pins the oracle's receiver sampler (visu_field, visualization.f90:156-181,
with calcfieldat / shapeparams / calcnmat / shapefuncs, formulation.f90:315-363)
by what the mathematics fixes, on the CPU:

* bilinear interpolation on an axis-aligned square reproduces every field
  a + b x + c z + d x z exactly, at points in all four quadrants of the
  natural coordinates, in uniform elements, in fine elements with a hanging
  corner and in coarse elements with hanging nodes on their edges;
* points outside the element are clamped to its boundary (shapeparams);
* element 0 (receiver not found) gives zeros (communicator.f90:1392-1394);
* the free field added is the plane wave's column 1 at t - t0, only where
  inside_soil holds (model.f90:217-235).

Each check is chosen so that a plausible mistake fails it: a swapped
shape-function order, a dropped x z term, a clamp to [0, 1] instead of
[-1, 1], or the free field added inside the box.
"""
import numpy as np
import pytest

import oracle_ctypes as O


@pytest.fixture(scope="module")
def cases():
    from damian_hip import cases
    return cases


def bilinear_state(case, coef):
    """u at every node (hanging nodes included) of the field
    (a + b x + c z + d x z) per component; hanging nodes then take the mean
    of their anchors as set_dispshanging does (model.f90:612-624)."""
    x, z = case.mesh.nodes[:, 0], case.mesh.nodes[:, 1]
    u = np.zeros(2 * case.mesh.nnodes)
    for comp in range(2):
        a, b, c, d = coef[comp]
        u[comp::2] = a + b * x + c * z + d * x * z
    for h, a1, a2 in case.hanging:
        for comp in range(2):
            u[2 * (h - 1) + comp] = (u[2 * (a1 - 1) + comp] + u[2 * (a2 - 1) + comp]) / 2
    return u


def exact(coef, pts):
    out = np.zeros((len(pts), 2))
    for comp in range(2):
        a, b, c, d = coef[comp]
        out[:, comp] = a + b * pts[:, 0] + c * pts[:, 1] + d * pts[:, 0] * pts[:, 1]
    return out


def element_points(case, elems, rs):
    """Physical points of natural coordinates rs (n, 2) in [-1, 1]^2 of the
    given 1-based elements (corner BL = (-1, -1), TR = (1, 1))."""
    c = case.mesh.corners()
    n1 = case.mesh.nodes[c[elems - 1, 0] - 1]
    n3 = case.mesh.nodes[c[elems - 1, 2] - 1]
    return n1 + (rs + 1.0) / 2.0 * (n3 - n1)


def shape_weights(rs, order=(0, 1, 2, 3)):
    r, s = rs[:, 0], rs[:, 1]
    w = np.stack([(1 - r) * (1 - s) / 4, (1 + r) * (1 - s) / 4, (1 + r) * (1 + s) / 4,
                  (1 - r) * (1 + s) / 4], axis=1)
    return w[:, list(order)]


COEF = np.array([[0.3, -1.7e-2, 2.3e-2, 4.1e-4], [-1.1, 2.9e-2, -0.7e-2, -3.3e-4]])
QUADRANTS = np.array([[-0.5, -0.5], [0.5, -0.5], [0.5, 0.5], [-0.5, 0.5],
                      [-0.83, 0.21], [0.37, -0.91], [0.0, 0.0], [0.99, 0.99]])


def pick_elements(case, rng, n=12):
    """Elements of each kind: uniform ones, fine ones with a hanging corner,
    coarse ones with hanging nodes on their edges."""
    m = case.mesh
    c = m.corners()
    hang = set(int(v) for v in case.hanging[:, 0]) if len(case.hanging) else set()
    fine = [e for e in range(m.nelems) if any(int(q) in hang for q in c[e])]
    rows = np.add.reduceat(m.ediv.astype(int), m.eptr[:-1] - 1)
    coarse = list(np.nonzero(rows > 0)[0])
    kinds = {"any": list(rng.choice(m.nelems, n, replace=False))}
    if fine:
        kinds["fine_with_hanging_corner"] = list(rng.choice(fine, min(n, len(fine)), replace=False))
    if coarse:
        kinds["coarse_with_hanging_edges"] = list(rng.choice(coarse, min(n, len(coarse)), replace=False))
    return {k: np.array(v, dtype=int) + 1 for k, v in kinds.items()}


@pytest.mark.parametrize("name", ["f1", "f2"])
def test_bilinear_field_reproduced_in_all_quadrants(cases, name):
    case = cases.BUILDERS[name](nsteps=1)
    rng = np.random.default_rng(5)
    u = bilinear_state(case, COEF)
    kinds = pick_elements(case, rng)
    if name == "f2":
        assert {"fine_with_hanging_corner", "coarse_with_hanging_edges"} <= set(kinds)
    for kind, elems in kinds.items():
        for rs in QUADRANTS:
            rsn = np.tile(rs, (len(elems), 1))
            pts = element_points(case, elems, rsn)
            got = O.visu_field(case, u, elems, pts, free_field=False)
            want = exact(COEF, pts)
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err < 1e-13, (kind, rs, err)
        # random points anywhere inside the element
        rsn = rng.uniform(-1, 1, size=(len(elems), 2))
        pts = element_points(case, elems, rsn)
        got = O.visu_field(case, u, elems, pts, free_field=False)
        assert np.abs(got - exact(COEF, pts)).max() < 1e-13 * np.abs(exact(COEF, pts)).max()


def test_planted_mistakes_are_detected(cases):
    """The same points distinguish the correct weights from a swapped
    shape-function order and from a dropped x z term: a sampler with either
    mistake would fail the test above."""
    case = cases.f2(nsteps=1)
    u = bilinear_state(case, COEF)
    elems = pick_elements(case, np.random.default_rng(5))["any"]
    c = case.mesh.corners()
    for rs in QUADRANTS[:4]:
        rsn = np.tile(rs, (len(elems), 1))
        pts = element_points(case, elems, rsn)
        vals = np.stack([u[2 * (c[elems - 1, j] - 1)] for j in range(4)], axis=1)
        good = (shape_weights(rsn) * vals).sum(axis=1)
        np.testing.assert_allclose(good, exact(COEF, pts)[:, 0], rtol=0, atol=1e-12)
        swapped = (shape_weights(rsn, (0, 1, 3, 2)) * vals).sum(axis=1)
        assert np.abs(swapped - exact(COEF, pts)[:, 0]).max() > 1e-6
    nodz = COEF.copy()
    nodz[:, 3] = 0.0
    pts = element_points(case, elems, np.tile(QUADRANTS[0], (len(elems), 1)))
    assert np.abs(exact(nodz, pts) - exact(COEF, pts)).max() > 1e-6


def test_points_outside_the_element_are_clamped(cases):
    case = cases.f2(nsteps=1)
    u = bilinear_state(case, COEF)
    elems = pick_elements(case, np.random.default_rng(9))["any"]
    outside = np.array([[1.6, -1.3], [-2.0, 0.4], [0.2, 1.9], [-1.4, -1.8], [3.0, 3.0]])
    for rs in outside:
        rsn = np.tile(rs, (len(elems), 1))
        pts = element_points(case, elems, rsn)
        clamped = element_points(case, elems, np.clip(rsn, -1.0, 1.0))
        got = O.visu_field(case, u, elems, pts, free_field=False)
        want = exact(COEF, clamped)
        assert np.abs(got - want).max() < 1e-13 * np.abs(want).max(), rs
        if (rs < 0).any():
            # a clamp of the natural coordinates to [0, 1] would give another value
            wrong = exact(COEF, element_points(case, elems, np.clip(rsn, 0.0, 1.0)))
            assert np.abs(wrong - want).max() > 1e-6


def test_receiver_not_found_gives_zeros(cases):
    case = cases.f1(nsteps=1)
    u = bilinear_state(case, COEF)
    got = O.visu_field(case, u, np.array([0, 0]), np.array([[10.0, 0.0], [3.0, 4.0]]), time=1.3)
    assert np.array_equal(got, np.zeros((2, 2)))


@pytest.mark.parametrize("name", ["f1", "f3s", "f3p"])
def test_free_field_added_only_outside_the_box(cases, name):
    """With the wave on, the sample is the interpolated field plus the plane
    wave's displacement column at (t - t0, x - x0, z - z0), exactly where
    inside_soil holds (the exterior region stores the scattered field)."""
    from damian_hip import setup as S
    case = cases.BUILDERS[name](nsteps=1)
    mdl = case.model
    u = bilinear_state(case, COEF)
    rng = np.random.default_rng(3)
    elems = rng.choice(case.nelems, 60, replace=False) + 1
    pts = element_points(case, elems, rng.uniform(-1, 1, size=(60, 2)))
    t = 1.37
    with_ff = O.visu_field(case, u, elems, pts, time=t)
    without = O.visu_field(case, u, elems, pts, time=t, free_field=False)
    rc, w = O.initplanewaves(mdl.wave_kind, mdl.wamp, mdl.wvelo[0], mdl.wvelo[1], mdl.wangle,
                             mdl.wfreq)
    assert rc == 0
    soil = S.inside_soil(mdl, pts[:, 0], pts[:, 1])
    assert soil.any() and (~soil).any(), "points on both sides of the box"
    for i in range(len(pts)):
        d = with_ff[i] - without[i]
        if soil[i]:
            pw = O.planewave(w, t - mdl.wt0, pts[i, 0] - mdl.worigin[0], pts[i, 1] - mdl.worigin[1])
            np.testing.assert_allclose(d, pw[:, 0], rtol=1e-12, atol=1e-15)
        else:
            assert np.array_equal(d, np.zeros(2))

"""Pins the oracle's element kernels (formulation.f90) without running the
reference: closed forms, invariants and brute force."""
import numpy as np
import pytest

import oracle_ctypes as O


def reference_closed_form_k(m, l):
    """The 8x8 entries listed in formulation.f90:279-312 (column-major), as
    functions of mu = m and lambda = l; used as the known answer."""
    a = (3 * m + l) / 3
    b = (m + l) / 4
    c = -(3 * m + 2 * l) / 6
    d = -(m - l) / 4
    e = -(3 * m + l) / 6
    g = l / 6
    cols = [
        [a, b, c, d, e, -b, g, -d],
        [b, a, -d, g, -b, e, d, c],
        [c, -d, a, -b, g, d, e, b],
        [d, g, -b, a, -d, c, b, e],
        [e, -b, g, -d, a, b, c, d],
        [-b, e, d, c, b, a, -d, g],
        [g, d, e, b, c, -d, a, -b],
        [-d, c, b, e, d, g, -b, a],
    ]
    return np.array(cols).T  # column j of the list is column j of K


@pytest.mark.parametrize("mu,lam", [(1.0, 0.0), (0.0, 1.0), (5e8, 1e9), (2.2e9, 4.4e9), (3.7, 11.3)])
def test_stiffness_matches_reference_closed_form(mu, lam):
    K = O.calck(mu, lam)
    R = reference_closed_form_k(mu, lam)
    np.testing.assert_allclose(K, R, rtol=1e-14, atol=1e-14 * max(mu, lam))


def test_stiffness_invariants():
    K = O.calck(2.0, 3.0)
    np.testing.assert_allclose(K, K.T, atol=1e-15)
    corners = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=float)
    tx = np.tile([1.0, 0.0], 4)
    tz = np.tile([0.0, 1.0], 4)
    rot = np.column_stack([-corners[:, 1], corners[:, 0]]).ravel()
    for r in (tx, tz, rot):
        np.testing.assert_allclose(K @ r, 0.0, atol=1e-14)
    w = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(w) < 1e-12) == 3          # exactly the rigid-body modes
    assert np.all(w > -1e-12)
    # linear in (mu, lambda)
    np.testing.assert_allclose(O.calck(2.0, 3.0), 2.0 * O.calck(1.0, 0.0) + 3.0 * O.calck(0.0, 1.0),
                               atol=1e-14)
    # size independence is structural: K of a square of side h has no h.


@pytest.mark.parametrize("mu,lam", [(1.0, 0.0), (0.0, 1.0), (2.5, 7.0)])
def test_uniform_strain_patch_forces(mu, lam):
    """u_x = x on the unit square (eps_xx = 1): nodal forces equal the edge
    tractions of sigma = (lambda + 2 mu, lambda, 0) lumped to the corners."""
    K = O.calck(mu, lam)
    corners = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=float)
    u = np.column_stack([corners[:, 0], np.zeros(4)]).ravel()
    s = lam + 2 * mu
    f = np.array([-s / 2, -lam / 2, s / 2, -lam / 2, s / 2, lam / 2, -s / 2, lam / 2])
    np.testing.assert_allclose(K @ u, f, atol=1e-14 * (1 + s))
    # pure shear u_x = z: sigma_xz = mu
    u = np.column_stack([corners[:, 1], np.zeros(4)]).ravel()
    f = np.array([-mu / 2, -mu / 2, -mu / 2, mu / 2, mu / 2, mu / 2, mu / 2, -mu / 2])
    np.testing.assert_allclose(K @ u, f, atol=1e-14 * (1 + mu))


def test_consistent_mass_pattern():
    M = O.calcsqm()
    row = np.array([4, 0, 2, 0, 1, 0, 2, 0]) / 9.0      # formulation.f90:258
    np.testing.assert_allclose(M[:, 0], row, atol=1e-16)
    np.testing.assert_allclose(M, M.T)
    # rows sum to one per component: 4 * integral N_i = 1, so m_e * M has the
    # element mass rho*h^2 summed per component
    np.testing.assert_allclose(M[0::2, 0::2].sum(axis=1), 1.0)
    np.testing.assert_allclose(M.sum(), 8.0)


def test_lumped_mass_and_lysmer_closed_forms():
    L = O.lib()
    import ctypes as C
    assert L.orc_calcm(C.c_double(2.5), C.c_double(2000.0)) == 2.5 ** 2 * 2000.0
    res = (C.c_double * 2)()
    L.orc_lysmer(C.c_double(5.0), C.c_double(2000.0), C.c_double(1000.0), C.c_double(500.0), 0, res)
    assert tuple(res) == (1000.0 * 2000.0 * 5.0, 500.0 * 2000.0 * 5.0)   # XDIR: (vp, vs)
    L.orc_lysmer(C.c_double(5.0), C.c_double(2000.0), C.c_double(1000.0), C.c_double(500.0), 1, res)
    assert tuple(res) == (500.0 * 2000.0 * 5.0, 1000.0 * 2000.0 * 5.0)   # ZDIR swapped


def test_multdrm_brute_force():
    """p_i = -sum_{j ext} K_ij u_j for interior i, p_j = +sum_{i int} K_ji u_i
    for exterior j (formulation.f90:378-393), on random inputs and masks."""
    rng = np.random.default_rng(7)
    for _ in range(50):
        K = rng.standard_normal((8, 8))
        u = rng.standard_normal(8)
        e = rng.integers(0, 2, 8).astype(np.uint8)
        p = O.multdrm(K, u, e)
        ext = e.astype(bool)
        q = np.zeros(8)
        q[~ext] = -K[np.ix_(~ext, ext)] @ u[ext]
        q[ext] = K[np.ix_(ext, ~ext)] @ u[~ext]
        np.testing.assert_allclose(p, q, atol=1e-12)

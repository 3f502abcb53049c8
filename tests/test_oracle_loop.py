"""Pins the oracle's loop (explicit.f90:79-159) and setup arrays (model.f90)
against an independent restatement (assembled global matrices in numpy) on
tiny meshes, and against invariants of the Domain Reduction Method."""
import numpy as np
import pytest

import oracle_ctypes as O
from conftest import rel_l2
from damian_hip import cases, setup as S
from test_oracle_elements import reference_closed_form_k

MASK = {2: (1, 0, 0, 1), 4: (0, 1, 1, 0), 1: (0, 0, 1, 1), 6: (1, 0, 1, 1), 5: (0, 1, 1, 1)}


def assembled_run(case, nsteps):
    """u after nsteps with global K, M, C and the constraint transpose P2."""
    m = case.mesh
    mdl = case.model
    nd = 2 * m.nnodes
    a1 = 1.0 / case.dt ** 2
    a2 = 1.0 / (2 * case.dt)
    cr = m.corners() - 1
    h = m.elength()
    mc = case.elem_props if case.elem_props is not None else mdl.matconst[m.ematprop - 1]
    K = np.zeros((nd, nd))
    M = np.zeros(nd)
    Cd = np.zeros(nd)
    Ke = []
    for e in range(m.nelems):
        ke = reference_closed_form_k(mc[e, 0], mc[e, 1])
        Ke.append(ke)
        dofs = np.ravel(np.column_stack([2 * cr[e], 2 * cr[e] + 1]))
        K[np.ix_(dofs, dofs)] += ke
        me = (h[e] / 2) ** 2 * mc[e, 2]
        M[dofs] += me
        for k, n in enumerate(cr[e]):
            ab = case.absorbing[n]
            if ab:
                c = np.array([mc[e, 3], mc[e, 4]]) * mc[e, 2] * h[e]
                if ab == 1:
                    c = c[::-1]
                Cd[2 * n:2 * n + 2] += c
    P = np.eye(nd)          # sum_reactanchors as a matrix (P2: R_a += R_h / 2)
    for hn, x1, x2 in case.hanging - 1:
        T = np.eye(nd)
        for c in range(2):
            T[2 * x1 + c, 2 * hn + c] += 0.5
            T[2 * x2 + c, 2 * hn + c] += 0.5
        P = T @ P
    mfull = P @ (a1 * M + a2 * Cd)
    Mh = O.calcsqm()
    if mdl.is_wave:
        rc, w = O.initplanewaves(mdl.wave_kind, mdl.wamp, *mdl.wvelo, mdl.wangle, mdl.wfreq)
    u = np.zeros(nd)
    up = np.zeros(nd)
    pt = case.point
    if pt is not None:
        kind, amp, sig, start, pnodes = pt
        rk = lambda t: amp * np.exp(-t * t / (4 * sig)) * (t * t - 2 * sig) / (8 * np.sqrt(np.pi) * sig ** 2.5)
        pz = 2 * (pnodes - 1) + 1
        if kind == "POINTD":
            u[pz] = rk(0.0 - start)
            up[pz] = rk(-case.dt - start)
    for i in range(1, nsteps + 1):
        F = np.zeros(nd)
        if pt is not None and kind == "POINTN":
            F[pz] += rk(i * case.dt - start)
        t = i * case.dt - mdl.wt0
        for e, skind in case.seismic:
            e -= 1
            xy = m.nodes[cr[e]]
            ff = [O.planewave(w, t, x - mdl.worigin[0], z - mdl.worigin[1]) for x, z in xy]
            u0 = np.ravel([f[:, 0] for f in ff])
            acc = np.ravel([f[:, 2] for f in ff])
            ext = np.repeat(np.array(MASK[skind], bool), 2)
            me = (h[e] / 2) ** 2 * mc[e, 2]
            p = np.zeros(8)
            for mat, vec, scale in ((Ke[e], u0, 1.0), (Mh, acc, me)):
                p[~ext] -= scale * (mat[np.ix_(~ext, ext)] @ vec[ext])
                p[ext] += scale * (mat[np.ix_(ext, ~ext)] @ vec[~ext])
            dofs = np.ravel(np.column_stack([2 * cr[e], 2 * cr[e] + 1]))
            F[dofs] += p
        ro = P @ (F + M * (2 * a1 * u - a1 * up) + a2 * Cd * up - K @ u)
        up = u
        u = ro / mfull
        if pt is not None and kind == "POINTD":
            u[pz] = rk(i * case.dt - start)
        for hn, x1, x2 in case.hanging - 1:
            u[2 * hn:2 * hn + 2] = (u[2 * x1:2 * x1 + 2] + u[2 * x2:2 * x2 + 2]) / 2
    return u, up, mfull


@pytest.mark.parametrize("name,n", [("f1", 60), ("f2", 60), ("f3p", 40), ("f2het", 50),
                                    ("pointn", 90), ("pointd", 90)])
def test_loop_vs_assembled_restatement(name, n):
    case = cases.BUILDERS[name]()
    u_o, up_o, mf_o, _ = O.run(case, nsteps=n, receivers=False)
    u_a, up_a, mf_a = assembled_run(case, n)
    np.testing.assert_allclose(mf_o, mf_a, rtol=1e-14)
    assert np.abs(u_a).max() > 1e-6
    scale = np.abs(u_a).max()
    np.testing.assert_allclose(u_o, u_a, rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(up_o, up_a, rtol=0, atol=1e-11 * scale)


@pytest.mark.parametrize("name", ["f1", "f2", "f3s", "f2het", "pointn", "pointd"])
def test_threaded_loop_is_bit_identical(name):
    """The threaded oracle (node-centric gather in ascending element order,
    the all-cores CPU baseline) gives the same bits as the element-by-element
    scatter, receivers included."""
    case = cases.BUILDERS[name]()
    n = min(case.nsteps, 120)
    ref = O.run(case, nsteps=n, recv_interval=1)
    for threads in (2, 5):
        got = O.run(case, nsteps=n, recv_interval=1, threads=threads)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


def test_hanging_projection_is_exact():
    case = cases.f2(nsteps=80)
    u, up, mf, _ = O.run(case, receivers=False)
    for hn, x1, x2 in case.hanging - 1:
        for c in range(2):
            assert u[2 * hn + c] == (u[2 * x1 + c] + u[2 * x2 + c]) / 2.0


def test_p1_literal_differs_from_p2():
    """The literal sum_reactanchors (with `red` declared) is not the
    constraint transpose: it changes the effective mass of anchors."""
    case = cases.f2(nsteps=5)
    _, _, mf2, _ = O.run(case, receivers=False, hang_mode=2)
    _, _, mf1, _ = O.run(case, receivers=False, hang_mode=1)
    anchors = np.unique(case.hanging[:, 1:]) - 1
    assert not np.allclose(mf1[2 * anchors], mf2[2 * anchors])
    # P2 conserves the total effective mass over the state-carrying nodes
    act = np.setdiff1d(np.arange(case.mesh.nnodes), case.hanging[:, 0] - 1)
    dt = case.dt
    tot = 2 * (1 / dt ** 2) * np.sum((case.mesh.elength() / 2) ** 2 *
                                      case.model.matconst[case.mesh.ematprop - 1, 2] * 4)
    damp = mf2[np.r_[2 * act, 2 * act + 1]].sum()
    assert damp >= tot * (1 - 1e-12)


def test_setup_arrays_match_oracle():
    for name in ("f1", "f2", "c1"):
        c = cases.BUILDERS[name]()
        assert np.array_equal(O.setup_hanging(c.mesh), c.hanging)
        assert np.array_equal(O.create_absorbingbc(c.model, c.mesh), c.absorbing)
        assert np.array_equal(O.wave_elems(c.model, c.mesh), c.seismic)


def test_setup_arrays_hand_checked():
    """2 x 1 coarse cells, the left one refined: one hanging node on the
    shared edge with the right cell's corners as anchors."""
    def refine(x0, z0, s):
        return x0 < 1.0
    mesh = S.quadtree_mesh(2, 1, 10.0, 2, refine)
    hang = S.setup_hanging(mesh)
    assert hang.shape == (1, 3)
    xy = mesh.nodes[hang[0] - 1]
    np.testing.assert_allclose(xy, [[10.0, 5.0], [10.0, 10.0], [10.0, 0.0]])
    # coarse element row: BL, BR, TR, TL with the hanging node after TL
    coarse = [e for e in range(mesh.nelems) if mesh.eptr[e + 1] - mesh.eptr[e] == 5][0]
    row = mesh.eind[mesh.eptr[coarse] - 1:mesh.eptr[coarse + 1] - 1]
    flags = mesh.ediv[mesh.eptr[coarse] - 1:mesh.eptr[coarse + 1] - 1]
    assert list(flags) == [0, 0, 0, 0, 1]
    np.testing.assert_allclose(mesh.nodes[row - 1],
                               [[10, 0], [20, 0], [20, 10], [10, 10], [10, 5]])


def _f1_at(f0, t0):
    txt = cases.problem_text(200, 100, 2.0, 0.002, 0.1, 10, 2.5, [("p", 1000, "s", 500, 2000)],
                             "WAVESV", [100, 50, 100, 0, 1, t0, 0.0, 1.0, f0])
    case = S.case_from_problem_text(txt, sheet=(-5.0, 200.0, 5.0))
    case.nsteps = int(round((t0 + 0.2) / case.dt))
    return case


def _drm_errors(case):
    mdl = case.model
    u, up, mf, rec = O.run(case, recv_interval=1)
    rc, w = O.initplanewaves(mdl.wave_kind, mdl.wamp, *mdl.wvelo, mdl.wangle, mdl.wfreq)
    t = case.nsteps * case.dt - mdl.wt0
    x, z = case.mesh.nodes[:, 0], case.mesh.nodes[:, 1]
    ff = np.array([O.planewave(w, t, xi - mdl.worigin[0], zi - mdl.worigin[1])[:, 0]
                   for xi, zi in zip(x, z)])
    soil = S.inside_soil(mdl, x, z)
    ui = u.reshape(-1, 2)
    err_in = np.linalg.norm(ui[~soil] - ff[~soil]) / np.linalg.norm(ff[~soil])
    scat = np.linalg.norm(ui[soil]) / np.linalg.norm(ff[soil])
    return err_in, scat, rec


def test_drm_invariants_vertical_sv():
    """Homogeneous half-space: inside the plane-wave box the total field
    equals the free field and outside it the scattered field vanishes, up to
    a discretization error that falls with the square of the frequency;
    surface motion doubles (u_x = 2 r(t - t0), u_z = 0)."""
    e4, s4, _ = _drm_errors(_f1_at(4.0, 0.3))
    case = _f1_at(2.0, 0.6)
    e2, s2, rec = _drm_errors(case)
    assert e2 < 0.03 and s2 < 0.005, (e2, s2)
    assert e2 < e4 / 3 and s2 < s4 / 3, (e4, e2, s4, s2)     # second-order convergence
    mdl = case.model
    found = case.recv_elem > 0
    times = np.arange(rec.shape[0]) * case.dt - mdl.wt0
    a = (np.pi * mdl.wfreq * times) ** 2
    r2 = 2 * (1 - 2 * a) * np.exp(-a)
    ex = np.linalg.norm(rec[:, found, 0] - r2[:, None]) / (np.linalg.norm(r2) * np.sqrt(found.sum()))
    assert ex < 0.03, ex
    assert np.abs(rec[:, found, 1]).max() < 0.01      # vs amplitude 2
    assert np.all(rec[:, ~found] == 0.0)


@pytest.mark.parametrize("name", ["f1", "f2", "f3p", "pointn", "pointd", "f2het"])
def test_long_double_loop_agrees_with_fp64(name):
    """orc_run_ld is orc_run's loop in 80-bit precision: on the short runs of
    the small cases the two differ by FP64 rounding only (the seismograms
    to 1e-13; the final fields of the plane-wave cases are the small residue
    left after the pulse has passed, relative 1e-12)."""
    from damian_hip import cases
    case = cases.BUILDERS[name]()
    u, up, mf, rec = O.run(case, recv_interval=1)
    ul, upl, recl = O.run_ld(case, recv_interval=1, threads=4)
    assert rec.shape == recl.shape
    assert rel_l2(rec, recl) < 1e-13
    assert rel_l2(u, ul) < 1e-12 and rel_l2(up, upl) < 1e-12
    assert not np.array_equal(u, ul), "the extra precision must show"


def test_long_double_loop_threads_bit_identical():
    from damian_hip import cases
    case = cases.f2(nsteps=60)
    a = O.run_ld(case, recv_interval=1, threads=1)
    b = O.run_ld(case, recv_interval=1, threads=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)

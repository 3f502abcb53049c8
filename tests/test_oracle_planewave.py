"""Pins the oracle's free field (pwaves.f90) by physics, not by retyping it:
traction-free surface, the elastic wave equation, the Ricker time
derivatives, vertical-incidence doubling, and the reference's guards."""
import math

import numpy as np
import pytest

import oracle_ctypes as O

ALPHA, BETA = 2000.0, 1000.0


def field(w, t, x, z, col=0):
    u = O.planewave(w, t, x, z)
    return u[:, col]


def grads(w, t, x, z, h):
    """Centered finite-difference gradient of the displacement (col 1)."""
    ux = (field(w, t, x + h, z) - field(w, t, x - h, z)) / (2 * h)
    uz = (field(w, t, x, z + h) - field(w, t, x, z - h)) / (2 * h)
    return ux, uz   # d/dx (u_x, u_z), d/dz (u_x, u_z)


CASES = [(1, 0.0), (1, math.radians(20)), (1, math.radians(45)), (2, 0.0),
         (2, math.radians(15)), (2, math.radians(25))]


@pytest.mark.parametrize("kind,angle", CASES)
def test_traction_free_surface(kind, angle):
    """sigma_xz = sigma_zz = 0 on z = 0 for every x and t (rho = 1, so
    mu = beta^2, lambda = alpha^2 - 2 beta^2)."""
    rc, w = O.initplanewaves(kind, 1.0, ALPHA, BETA, angle, 2.0)
    assert rc == 0
    mu = BETA ** 2
    lam = ALPHA ** 2 - 2 * mu
    h = 0.05
    worst, scale = 0.0, 0.0
    for t in np.linspace(-0.4, 0.4, 17):
        for x in (-300.0, 0.0, 170.0):
            dx, dz = grads(w, t, x, 0.0, h)
            szz = lam * (dx[0] + dz[1]) + 2 * mu * dz[1]
            sxz = mu * (dz[0] + dx[1])
            worst = max(worst, abs(szz), abs(sxz))
            # stress scale of the incident wave alone
            dxi, dzi = grads(w, t, x, 150.0, h)
            scale = max(scale, abs(lam * (dxi[0] + dzi[1]) + 2 * mu * dzi[1]),
                        abs(mu * (dzi[0] + dxi[1])))
    assert scale > 0
    assert worst < 1e-5 * scale


@pytest.mark.parametrize("kind,angle", CASES)
def test_wave_equation_and_acceleration_column(kind, angle):
    """Column 3 is the acceleration of column 1: it equals d2u/dt2 and the
    Navier operator of the displacement (rho = 1)."""
    rc, w = O.initplanewaves(kind, 1.0, ALPHA, BETA, angle, 2.0)
    mu = BETA ** 2
    lam = ALPHA ** 2 - 2 * mu
    h, dt = 0.5, 1e-4
    for t, x, z in [(0.05, 10.0, 120.0), (-0.1, -200.0, 40.0), (0.2, 300.0, 260.0)]:
        a = field(w, t, x, z, 2)
        att = (field(w, t + dt, x, z) - 2 * field(w, t, x, z) + field(w, t - dt, x, z)) / dt ** 2
        np.testing.assert_allclose(att, a, rtol=2e-4, atol=1e-4 * (1 + np.abs(a).max()))
        f = lambda xx, zz: field(w, t, xx, zz)
        uxx = (f(x + h, z) - 2 * f(x, z) + f(x - h, z)) / h ** 2
        uzz = (f(x, z + h) - 2 * f(x, z) + f(x, z - h)) / h ** 2
        uxz = (f(x + h, z + h) - f(x + h, z - h) - f(x - h, z + h) + f(x - h, z - h)) / (4 * h * h)
        nav = np.array([mu * (uxx[0] + uzz[0]) + (lam + mu) * (uxx[0] + uxz[1]),
                        mu * (uxx[1] + uzz[1]) + (lam + mu) * (uxz[0] + uzz[1])])
        np.testing.assert_allclose(nav, a, rtol=2e-3, atol=2e-3 * (1 + np.abs(a).max()))


def ricker(t, f):
    s = (math.pi * f * t) ** 2
    return (1 - 2 * s) * math.exp(-s)


def test_vertical_doubling():
    rc, w = O.initplanewaves(2, 1.0, ALPHA, BETA, 0.0, 2.0)       # SV
    for t in np.linspace(-0.5, 0.5, 11):
        u = field(w, t, 123.0, 0.0)
        assert abs(u[0] - 2 * ricker(t, 2.0)) < 1e-14
        assert abs(u[1]) < 1e-14
    rc, w = O.initplanewaves(1, 1.0, ALPHA, BETA, 0.0, 2.0)       # P (z down)
    for t in np.linspace(-0.5, 0.5, 11):
        u = field(w, t, -40.0, 0.0)
        assert abs(u[1] + 2 * ricker(t, 2.0)) < 1e-14
        assert abs(u[0]) < 1e-14


def test_reflection_energy_flux_sv():
    """Free-surface SV reflection conserves vertical energy flux:
    |b1/b|^2 + (alpha cos e)/(beta cos f) |a1/b|^2 = 1 (pre-critical)."""
    for deg in (5, 12, 20, 27):
        rc, w = O.initplanewaves(2, 1.0, ALPHA, BETA, math.radians(deg), 1.0)
        assert rc == 0
        flux = w.b1 ** 2 + (ALPHA * w.cose) / (BETA * w.cosf) * w.a1 ** 2
        assert abs(flux - 1.0) < 1e-12


def test_reflection_energy_flux_p():
    for deg in (5, 20, 40, 70):
        rc, w = O.initplanewaves(1, 1.0, ALPHA, BETA, math.radians(deg), 1.0)
        assert rc == 0
        flux = w.a1 ** 2 + (BETA * w.cosf) / (ALPHA * w.cose) * w.b1 ** 2
        assert abs(flux - 1.0) < 1e-12


def test_guards():
    # grazing P (cos e below the threshold) and post-critical SV are refused
    assert O.initplanewaves(1, 1.0, ALPHA, BETA, math.pi / 2, 1.0)[0] != 0
    assert O.initplanewaves(2, 1.0, ALPHA, BETA, math.asin(BETA / ALPHA), 1.0)[0] != 0
    assert O.initplanewaves(2, 1.0, ALPHA, BETA, math.radians(40), 1.0)[0] != 0
    # the vertical branch threshold is the single-precision 1e-4 of pwaves.f90:8
    thr = float(np.float32(1e-4))
    rc, w = O.initplanewaves(2, 1.0, ALPHA, BETA, thr * 0.999, 1.0)
    assert rc == 0 and w.sinf == 0.0 and w.a1 == 0.0
    rc, w = O.initplanewaves(2, 1.0, ALPHA, BETA, thr * 1.001, 1.0)
    assert rc == 0 and w.sinf != 0.0

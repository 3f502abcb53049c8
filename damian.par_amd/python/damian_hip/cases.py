"""Named site-response workloads (synthetic inputs, fixed parameters).

  f1       uniform 40 x 20 grid (800 elements), vertical SV, 400 steps
  f2       two-material circular valley, one quadtree level (hanging nodes)
  f3p/f3s  f1 with an oblique P (30 deg) / SV (20 deg) plane wave
  c1       2000 m x 500 m layer, 5 m quads (40,000 elements), SV 0 deg, f0 2 Hz,
           dt 0.002 s, 5000 steps, 101 receivers
  c2       semicircular canyon R = 250 m in 4000 m x 1500 m, Vs 1000, Vp 2000,
           rho 2200; quadtree 20 m -> 2.5 m within 50 m of the free surface
           (4 levels, hanging nodes); SV 29 deg (30 deg is exactly critical and
           rejected like the reference), f0 3 Hz, dt 2.5e-4 s, 20,000 steps,
           201 receivers
  c2u      the same canyon on a uniform 2.5 m grid (no refinement)
  c3       trapezoid sediment valley in bedrock with a von Karman random
           valley Vs (lognormal, 10 %, correlation length 40 m), 10 m -> 2.5 m
           quadtree, vertical P, f0 4 Hz, 40,000 steps, 102 receivers
  c5       20 km x 5 km graded random medium (Vs 300 -> 3000 m/s with depth, von
           Karman 7 %, correlation length 100 m), 50 m -> 3.125 m quadtree, SV 15 deg,
           f0 2.5 Hz, 30,000 steps, 1001 receivers
  valley   trapezoid sediment valley instances for batched sweeps
  c4       the 64 instances of the batched sweep: config-3 valleys scaled with
           their depth (100-400 m), 37k to 503k elements, 13.9M in total
"""
from __future__ import annotations

import math

import numpy as np

from . import setup as S


def problem_text(width, height, duration, inttime, chper, nperlen, sperlen, mats, prob, pparams,
                 geom="NONE  ", gparams=(0.0,)):
    lines = ["PROBLEM FILE EXFEM", "DOMAIN:", f"{width} {height}", "TIME:",
             f"{duration} {inttime}", "CH PERIOD:", f"{chper}", "CONTROL:",
             f"{nperlen} {sperlen}", "MATERIALS:", f"{len(mats)}"]
    lines += [" ".join(str(x) for x in m) for m in mats]
    lines += ["PROBLEM:", prob, " ".join([str(len(pparams))] + [repr(float(p)) for p in pparams])]
    lines += ["GEOMETRY:", geom, " ".join([str(len(gparams))] + [repr(float(g)) for g in gparams])]
    return "\n".join(lines) + "\n"


def _uniform_case(text, sheet, name, nsteps=None):
    case = S.case_from_problem_text(text, sheet=sheet, name=name)
    if nsteps is not None:
        case.nsteps = nsteps
    return case


# Surface receivers of the small cases at x = 0.8 + 4.1 i m (i = 0..48): off
# the 5 m / 10 m node grid, so the samples interpolate between corners
# (visualization.f90:294-299 places them at xmin + dist * i).
OFF_NODE_SHEET = (-3.3, 197.0, 4.1)


def f1(nsteps=400):
    txt = problem_text(200, 100, 0.8, 0.002, 0.1, 10, 2.5, [("p", 1000, "s", 500, 2000)], "WAVESV",
                       [100, 50, 100, 0, 1, 0.3, 0.0, 1.0, 4.0])
    return _uniform_case(txt, OFF_NODE_SHEET, "f1", nsteps)


def f3(kind="P", nsteps=400):
    if kind == "P":
        pp = [100, 50, 100, 0, 1, 0.3, math.radians(30.0), 1.0, 4.0]
        prob = "WAVEP "
    else:
        pp = [100, 50, 100, 0, 1, 0.3, math.radians(20.0), 1.0, 4.0]
        prob = "WAVESV"
    txt = problem_text(200, 100, 0.8, 0.002, 0.1, 10, 2.5, [("p", 1000, "s", 500, 2000)], prob, pp)
    return _uniform_case(txt, OFF_NODE_SHEET, "f3" + kind.lower(), nsteps)


def c1(nsteps=5000, inttime=0.002):
    txt = problem_text(2000, 500, 10, inttime, 0.1, 10, 2.5, [("p", 1000, "s", 500, 2000)],
                       "WAVESV", [1000, 300, 1000, 0, 1, 1, 0.0, 1.0, 2.0])
    return _uniform_case(txt, (-20.0, 1980.0, 20.0), "c1", nsteps)


def _quadtree_case(text, nlevels, refine, delete=None, material=None, sheet=None, name="",
                   nsteps=None, dt=None):
    pf = S.parse_problem(text)
    mdl = S.build_model(pf)
    mesh = S.quadtree_mesh(mdl.totcols - 1, mdl.totrows - 1, mdl.blength, nlevels, refine,
                           delete=delete, material=material)
    case = S.finish_case(mdl, mesh, sheet=sheet, name=name, nsteps=nsteps, dt=dt)
    return case


def f2(nsteps=300):
    """Soft circular valley (R 40 m) in a 200 x 100 m block, 10 m base grid,
    5 m inside and around the valley: one refinement level, hanging nodes."""
    txt = problem_text(200, 100, 0.6, 0.002, 0.2, 10, 5.0,
                       [("p", 1000, "s", 500, 2000), ("p", 600, "s", 250, 1800)], "WAVESV",
                       [120, 50, 100, 0, 1, 0.3, 0.0, 1.0, 3.0], "VALLEY", (100.0, 0.0, 40.0))
    pf = S.parse_problem(txt)
    mdl = S.build_model(pf)

    def refine(x0, z0, s):
        cx, cz = x0 + s / 2, z0 + s / 2
        return np.hypot(cx - 100.0, cz) - s < 45.0

    def material(cx, cz, s):
        return S.material_prop(mdl, cx, cz)

    mesh = S.quadtree_mesh(mdl.totcols - 1, mdl.totrows - 1, mdl.blength, 2, refine,
                           material=material)
    return S.finish_case(mdl, mesh, sheet=OFF_NODE_SHEET, name="f2", nsteps=nsteps)


def _canyon_surface_distance(cx, cz, x0, r):
    """Distance from points to the free surface of a semicircular canyon
    (flat top outside |x - x0| < r, arc of radius r below)."""
    dx = np.abs(cx - x0)
    flat = np.where(dx >= r, np.abs(cz), np.hypot(dx - r, cz))
    arc = np.abs(np.hypot(cx - x0, cz) - r)
    arc = np.where(cz >= 0, arc, np.inf)
    return np.minimum(flat, arc)


def c2(nsteps=20000, uniform=False, angle_deg=29.0, band=50.0, hmin=2.5):
    """Canyon case of the benchmark (see module docstring)."""
    h0 = 20.0
    levels = int(round(math.log2(h0 / hmin))) + 1
    txt = problem_text(4000, 1500, nsteps * 2.5e-4, 2.5e-4, 0.2, 10, 5.0,
                       [("p", 2000, "s", 1000, 2200)], "WAVESV",
                       [1200, 600, 2000, 0, 1, 0.6, math.radians(angle_deg), 1.0, 3.0],
                       "CANYON", (2000.0, 0.0, 250.0))
    pf = S.parse_problem(txt)
    mdl = S.build_model(pf)
    x0, r = 2000.0, 250.0

    def refine(xl, zl, s):
        if uniform:
            return np.ones(len(xl), dtype=bool)
        d = _canyon_surface_distance(xl + s / 2, zl + s / 2, x0, r) - s * math.sqrt(0.5)
        return d < band

    def delete(ccx, ccz):
        return ~(np.hypot(ccx - x0, ccz) > r).any(axis=1)

    mesh = S.quadtree_mesh(mdl.totcols - 1, mdl.totrows - 1, mdl.blength, levels, refine,
                           delete=delete)
    # 201 receivers at x = 0.09 + 19.99 i m (i = 0..200): between the nodes
    # of every mesh level, so each sample interpolates two corners
    case = S.finish_case(mdl, mesh, sheet=(-19.9, 3980.0, 19.99), name="c2u" if uniform else "c2",
                         nsteps=nsteps)
    case.dt = 2.5e-4
    case.nsteps = nsteps
    return case


def valley(depth=200.0, angle_deg=0.0, kind="SV", width=6000.0, height=2000.0, h0=40.0,
           fine=2, nsteps=10000, f0=4.0, top=1000.0, bottom=300.0, box=(3000.0, 1000.0), t0=0.8,
           sheet_half=1000.0):
    """Trapezoid sediment valley (top and bottom widths, given depth) in
    bedrock; quadtree refined inside the valley (fine levels below h0);
    plane wave through a box (width, depth) centred at the surface, passing
    the origin at t0; surface receivers every 20 m within sheet_half of the
    centre."""
    chper = h0 * 10 / 1500.0
    mats = [("p", 3000, "s", 1500, 2400), ("p", 600, "s", 250, 1800)]
    prob = "WAVESV" if kind == "SV" else "WAVEP "
    txt = problem_text(width, height, nsteps * 1e-3, 1e-3, chper, 10, 5.0, mats, prob,
                       [box[0], box[1], width / 2, 0, 1, t0, math.radians(angle_deg), 1.0, f0])
    pf = S.parse_problem(txt)
    mdl = S.build_model(pf)
    xc = width / 2

    def half_width(z):
        return top / 2 - (top / 2 - bottom / 2) * np.clip(z / depth, 0, 1)

    def in_valley(cx, cz):
        return (cz <= depth) & (np.abs(cx - xc) <= half_width(cz))

    def refine(xl, zl, s):
        cx, cz = xl + s / 2, zl + s / 2
        return (cz - s <= depth) & (np.abs(cx - xc) - s <= half_width(np.maximum(cz - s, 0)))

    def material(cx, cz, s):
        return np.where(in_valley(cx, cz), 2, 1).astype(np.int32)

    mesh = S.quadtree_mesh(mdl.totcols - 1, mdl.totrows - 1, mdl.blength, fine + 1, refine,
                           material=material)
    case = S.finish_case(mdl, mesh, sheet=(xc - sheet_half - 20.0, xc + sheet_half, 20.0),
                         name=f"valley_d{depth:.0f}_{kind}{angle_deg:.0f}", nsteps=nsteps)
    return case


def add_von_karman(case, seed, sigma=0.1, a=40.0, nu=0.5):
    """Scale the valley (material 2) Vs of a valley case by a lognormal von
    Karman field (correlation length a, sigma, Hurst nu, seeded), on a 2.5 m
    grid over the valley twice its extent (the periodic wrap stays outside);
    each element takes its centre's value; per-element properties (mu,
    lambda, rho, vp, vs)."""
    m = case.mesh
    c = m.corners() - 1
    cx = m.nodes[c, 0].mean(axis=1)
    cz = m.nodes[c, 1].mean(axis=1)
    sed = m.ematprop == 2
    dx, x0 = 2.5, cx[sed].min() - 20.0
    nx, nz = 2 * int((cx[sed].max() + 20.0 - x0) / dx + 1), 2 * int((cz[sed].max() + 20.0) / dx + 1)
    g = von_karman(nx, nz, dx, a, nu, seed)
    ix = np.clip(np.rint((cx - x0) / dx).astype(int), 0, nx - 1)
    iz = np.clip(np.rint(cz / dx).astype(int), 0, nz - 1)
    mc = case.model.matconst[m.ematprop - 1]
    rho, vp = mc[:, 2], mc[:, 3]
    vs = mc[:, 4] * np.where(sed, np.exp(sigma * g[iz, ix]), 1.0)
    mu = rho * vs ** 2
    case.elem_props = np.stack([mu, rho * vp ** 2 - 2 * mu, rho, vp, vs], axis=1)
    return case


def c4_specs():
    """The 64 sources of the batched sweep: P waves at 0..60 deg in 4 deg
    steps (two depth draws each) and SV waves at 0..28 deg in 4 deg steps
    (four draws each) -- SV at 30 deg and beyond is post-critical in the
    bedrock (Vs/Vp = 0.5), where the reference's free field stops
    (pwaves.f90:87-89); instance k draws its valley depth uniformly from
    100 to 400 m with seed 4000 + k."""
    specs = [("P", a) for a in range(0, 61, 4)] * 2 + [("SV", a) for a in range(0, 29, 4)] * 4
    return [(kind, float(a), float(np.random.default_rng(4000 + k).uniform(100.0, 400.0)))
            for k, (kind, a) in enumerate(specs)]


def c4_valley(d, kind="P", angle=0.0, nsteps=10000, field_seed=3003):
    """A valley of the config-3 type scaled with its depth d: trapezoid 5d
    wide at the top, 1.5d at the bottom, in a 30d x 10d bedrock (d = 200 m
    is config 3's 1000 / 300 / 200 m valley in 6000 x 2000 m), same
    materials, 10 m cells refined to 2.5 m in and around the valley, valley
    Vs scaled by a lognormal von Karman field (40 m, 10 %, Hurst 0.5,
    field_seed), Ricker 4 Hz through a 10d x 3d box (t0 = 0.25 s + 3d /
    1500 m/s, so the pulse enters the box from outside), receivers every
    20 m over the box width.  About 35k (d = 100 m) to 560k (d = 400 m)
    elements."""
    case = valley(depth=d, angle_deg=angle, kind=kind, width=30.0 * d, height=10.0 * d, h0=10.0,
                  fine=2, nsteps=nsteps, f0=4.0, top=5.0 * d, bottom=1.5 * d,
                  box=(10.0 * d, 3.0 * d), t0=0.25 + 3.0 * d / 1500.0, sheet_half=5.0 * d)
    add_von_karman(case, seed=field_seed)
    return case


def c4_instance(k, nsteps=10000):
    """Instance k of the batched sweep: c4_valley with the depth and source
    of c4_specs()[k] and the von Karman seed 3003 + k."""
    kind, angle, d = c4_specs()[k]
    case = c4_valley(d, kind, angle, nsteps, field_seed=3003 + k)
    case.name = f"c4_{k}_d{d:.0f}_{kind}{angle:.0f}"
    return case


def von_karman(nx, nz, dx, a, nu=0.5, seed=0):
    """Zero-mean, unit-variance 2-D Gaussian random field with the von Karman
    spectrum P(k) ~ (1 + k^2 a^2)^-(nu + 1) on an nz x nx periodic grid of
    spacing dx: seeded white noise filtered in the Fourier domain.  For
    nu = 1/2 its correlation function is exp(-r / a)."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((nz, nx))
    kx = 2 * np.pi * np.fft.fftfreq(nx, dx)
    kz = 2 * np.pi * np.fft.fftfreq(nz, dx)
    k2 = kz[:, None] ** 2 + kx[None, :] ** 2
    g = np.fft.ifft2(np.fft.fft2(w) * (1.0 + k2 * a * a) ** (-(nu + 1) / 2)).real
    g -= g.mean()
    return g / g.std()


def c3(nsteps=40000, seed=3003, sigma=0.1, a=40.0, nu=0.5):
    """Heterogeneous sediment valley: the trapezoid valley (top 1000 m,
    bottom 300 m, 200 m deep) in 6000 x 2000 m bedrock (Vs 1500, Vp 3000,
    rho 2400), valley Vs 250, Vp 600, rho 1800 with Vs scaled by a lognormal
    von Karman field (correlation length 40 m, 10 %, Hurst 0.5, seed 3003);
    quadtree from 10 m to 2.5 m inside the valley (at least 10 points per
    wavelength at 8 Hz in both materials), vertical P wave, Ricker 4 Hz, 102
    receivers."""
    case = valley(depth=200.0, angle_deg=0.0, kind="P", h0=10.0, fine=2, nsteps=nsteps, f0=4.0)
    add_von_karman(case, seed, sigma, a, nu)
    case.name = f"c3_vk{seed}"
    return case


def c5(nsteps=30000, seed=5005, sigma=0.07, a=100.0, nu=0.5):
    """Large graded random medium: 20 km x 5 km, Vs = 300 + 2700 z / 5000 m/s,
    Vp = 2 Vs, rho = 1700 + 0.2 Vs, Vs scaled by a lognormal von Karman field
    (correlation length 100 m, 7 %, Hurst 0.5, seed 5005); quadtree from
    50 m down to 3.125 m (5 levels) so that every element has 10 points per
    wavelength at 5 Hz for the Vs at its top; SV plane wave at 15 deg, Ricker
    2.5 Hz, through a 16 km x 4 km DRM box; 1001 surface receivers every
    20 m.  The plane-wave forcing is that of a homogeneous half-space of the
    box material (Vs 3000), so in the graded medium the injected field is
    not an exact solution: the case is a stress test of the step, compared
    between precisions on identical inputs."""
    width, height = 20000.0, 5000.0
    vs_of = lambda z: 300.0 + 2700.0 * np.clip(z, 0.0, height) / height
    txt = problem_text(width, height, nsteps * 1e-3, 1e-3, 50.0 * 10 / 3000.0, 10, 5.0,
                       [("p", 6000, "s", 3000, 2300)], "WAVESV",
                       [16000, 4000, width / 2, 0, 1, 2.0, math.radians(15.0), 1.0, 2.5])
    pf = S.parse_problem(txt)
    mdl = S.build_model(pf)

    def refine(xl, zl, s):
        return s > vs_of(zl) / (10 * 5.0) + 1e-9

    mesh = S.quadtree_mesh(mdl.totcols - 1, mdl.totrows - 1, mdl.blength, 5, refine)
    case = S.finish_case(mdl, mesh, sheet=(-20.0, width - 20.0, 20.0), name=f"c5_vk{seed}",
                         nsteps=nsteps)
    c = mesh.corners() - 1
    cx = mesh.nodes[c, 0].mean(axis=1)
    cz = mesh.nodes[c, 1].mean(axis=1)
    dx = 12.5   # field grid; elements take the value at their centre
    nx, nz = 2 * int(width / dx + 1), 2 * int(height / dx + 1)
    g = von_karman(nx, nz, dx, a, nu, seed)
    ix = np.clip(np.rint(cx / dx).astype(int), 0, nx - 1)
    iz = np.clip(np.rint(cz / dx).astype(int), 0, nz - 1)
    vs = vs_of(cz) * np.exp(sigma * g[iz, ix])
    vp = 2.0 * vs
    rho = 1700.0 + 0.2 * vs
    mu = rho * vs ** 2
    case.elem_props = np.stack([mu, rho * vp ** 2 - 2 * mu, rho, vp, vs], axis=1)
    # the reference's step rule (min over elements of h / vp, over sperlen 5)
    # with each element's own vp
    case.dt = float(np.min(mesh.elength() / vp)) / 5.0
    case.nsteps = nsteps
    return case


BUILDERS = {"f1": f1, "f2": f2, "f3p": lambda **k: f3("P", **k), "f3s": lambda **k: f3("SV", **k),
            "c1": c1, "c2": c2, "c2u": lambda **k: c2(uniform=True, **k), "c3": c3, "c5": c5}


def perturbed(case, sigma=0.1, seed=1234):
    """Per-element heterogeneous copy of a case: Vs and Vp scaled by a seeded
    lognormal factor per element (rho unchanged), passed as per-element
    properties (mu, lambda, rho, vp, vs)."""
    rng = np.random.default_rng(seed)
    mc = case.model.matconst[case.mesh.ematprop - 1]
    f = np.exp(sigma * rng.standard_normal(case.nelems))
    rho = mc[:, 2]
    vp = mc[:, 3] * f
    vs = mc[:, 4] * f
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    case.elem_props = np.stack([mu, lam, rho, vp, vs], axis=1)
    case.name += f"_het{seed}"
    return case


BUILDERS["f1het"] = lambda **k: perturbed(f1(**k))
BUILDERS["f2het"] = lambda **k: perturbed(f2(**k), seed=77)


def canyon_reference(kind="SV"):
    """The reference's canyon1 (SV, 0 deg) / canyon4 (P, 0 deg) problems:
    40 km x 21 km, Vs 1000 (nu = 1/3, so Vp = 2000), rho 1000, 25 m grid,
    semicircular canyon a = 1000 m at x = 20 km, Ricker f0 1 Hz, t0 4 s,
    40 s; receivers every 25 m for x/a in [-2, 2]."""
    nu = "0.333333333333333" if kind == "SV" else "0.333333333"
    prob = "WAVESV" if kind == "SV" else "WAVEP "
    txt = problem_text(40000, 21000, 40, 0.1, 0.25, 10, 4, [("s", 1000, "v", nu, 1000)], prob,
                       [5000, 2000, 3000, 0, 1, 4, 0, 1, 1], "CANYON", (20000.0, 0.0, 1000.0))
    return _uniform_case(txt, (17975.0, 22000.0, 25.0), "canyon1" if kind == "SV" else "canyon4")


def point(kind="POINTN", nsteps=300):
    """f1's grid with a point source instead of the plane wave: sigma-Ricker
    (amp 1e9 for a force, 1e-2 m for a displacement, sigma 0.002, start
    0.15 s) on the nodes within 6 m of (100, 0) and (60, 40)."""
    amp = 1.0e9 if kind == "POINTN" else 1.0e-2
    txt = problem_text(200, 100, 0.8, 0.002, 0.1, 10, 2.5, [("p", 1000, "s", 500, 2000)], kind,
                       [amp, 0.002, 0.15, 6.0, 100.0, 0.0, 60.0, 40.0])
    return _uniform_case(txt, (-5.0, 200.0, 5.0), kind.lower(), nsteps)


BUILDERS["pointn"] = lambda **k: point("POINTN", **k)
BUILDERS["pointd"] = lambda **k: point("POINTD", **k)


def with_buried_receivers(case, n=40, seed=11):
    """Copy of a case with n more receivers at seeded random points strictly
    inside elements (all four quadrants of the natural coordinates): a third
    in elements with a hanging node as a corner, a third in coarse elements
    with hanging nodes on their edges (when the mesh has them), the rest
    anywhere.  The free field is added where inside_soil holds, as for
    surface receivers."""
    import copy
    rng = np.random.default_rng(seed)
    m = case.mesh
    c = m.corners()
    hang_nodes = set(int(v) for v in case.hanging[:, 0]) if len(case.hanging) else set()
    fine = np.array([e for e in range(m.nelems) if any(int(q) in hang_nodes for q in c[e])], dtype=int)
    rows = np.add.reduceat(m.ediv.astype(int), m.eptr[:-1] - 1) if m.nelems else np.zeros(0, int)
    coarse = np.nonzero(rows > 0)[0]
    pools = [p for p in (fine, coarse) if len(p)]
    pick = []
    for i in range(n):
        pool = pools[i % (len(pools) + 1)] if i % (len(pools) + 1) < len(pools) else np.arange(m.nelems)
        pick.append(int(pool[rng.integers(len(pool))]))
    pick = np.array(pick, dtype=int)
    x0 = m.nodes[c[pick, 0] - 1, 0]
    z0 = m.nodes[c[pick, 0] - 1, 1]
    h = m.nodes[c[pick, 2] - 1, 1] - z0
    rs = rng.uniform(0.03, 0.97, size=(n, 2))
    pts = np.stack([x0 + rs[:, 0] * h, z0 + rs[:, 1] * h], axis=1)
    out = copy.copy(case)
    out.recv_elem = np.concatenate([case.recv_elem, (pick + 1).astype(np.int32)])
    out.recv_xy = np.concatenate([case.recv_xy.reshape(-1, 2), pts])
    ff = S.inside_soil(case.model, pts[:, 0], pts[:, 1]).astype(np.uint8) if case.model.is_wave \
        else np.zeros(n, np.uint8)
    out.recv_ff = np.concatenate([case.recv_ff, ff])
    out.name = case.name + "_buried"
    return out

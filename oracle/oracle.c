/*
 * oracle.c -- CPU restatement (FP64) of exfem's explicit elastic-wave step
 * loop: init_explicit + run_explicit (explicit.f90:72-159) with everything
 * they call on one rank.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker / timed CPU baseline.  The
 * product library never links or calls this code.
 *
 * The reference Fortran cannot be built in this image (it needs m4-generated
 * sources, MPI + ParMETIS, and does not compile as shipped: `red` is
 * undeclared at model.f90:643), so this file is a restatement written from the
 * reference's source, pinned as follows (tests/test_oracle_*.py):
 *
 *   orc_calck            derived here by exact integration of B^T D B on the
 *                        unit square; pinned against the closed-form entries of
 *                        formulation.f90:276-313, symmetry, rigid-body null
 *                        space, linearity in (mu, lambda), uniform-strain patch
 *                        forces.
 *   orc_calcsqm          derived by exact integration of N_i N_j; pinned
 *                        against formulation.f90:256-268 and total mass.
 *   orc_calcm/lysmer     closed forms (formulation.f90:219-229, 250-254).
 *   orc_multdrm          brute force on random inputs (formulation.f90:378-393).
 *   orc_initplanewaves/  textbook P-SV free-surface reflection coefficients,
 *   orc_planewave        vertical-incidence surface doubling, traction-free
 *                        surface (finite differences), Ricker identities.
 *   orc_setup_hanging,   brute force on hand-built meshes (model.f90:654-675,
 *   orc_create_absorbingbc, 1080-1096, 1015-1045).
 *   orc_wave_elems
 *   orc_run              1-2 element brute force restated in numpy; the
 *                        homogeneous-half-space DRM invariant (interior total
 *                        field == free field, exterior scattered field ~ 0);
 *                        vertical SV surface motion u_x = 2 r(t - t0), u_z = 0;
 *                        rigid translation / energy checks; hanging projection
 *                        exactness.  End-to-end seismograms of the reference
 *                        program itself are parity unpinned (reference
 *                        unbuildable here, no golden traces shipped).
 *   orc_run_mt           bit-identical to orc_run on every small case
 *                        (receivers, u, u_prev, effective mass).
 *   orc_run_ld           the same loop in long double: agrees with orc_run to
 *                        rounding on short runs of every small case; over
 *                        thousands of steps orc_run drifts from it by the
 *                        accumulated FP64 rounding of the reference formula
 *                        (e.g. 1.8e-12 on config 3 after 3500 steps), while a
 *                        reordered FP64 sum drifts by 6e-14.
 *   orc_visu_field       (the receiver sampler of orc_run) reproduces any
 *                        bilinear field a + b x + c z + d x z exactly at points
 *                        in all four quadrants of uniform, coarse and
 *                        hanging-corner elements, clamps points outside the
 *                        element to its boundary, returns zeros for element 0,
 *                        and adds the free field column 1 at t - t0 only where
 *                        inside_soil holds (visualization.f90:156-181,
 *                        formulation.f90:315-363).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define U_PI 3.14159265358979323846
#define U_TOL 1.0e-9
/* pwaves.f90:8 declares f_wangzero = 0.0001 as a *default real* literal, so the
 * threshold the reference compares against is the single-precision value. */
static const double F_WANGZERO = (double)0.0001f;

static int realequal(double a, double b) { return fabs(a - b) < U_TOL; }
static int realunder(double a, double b) { return (a - b) < -U_TOL; }
static int realover(double a, double b) { return (a - b) > U_TOL; }

/* ------------------------------------------------------------------------- */
/* Element matrices.                                                          */
/* ------------------------------------------------------------------------- */

/* Corner a of the unit square in the reference corner order BL, BR, TR, TL
 * (x right, z down; mesh_utils.f90:101-104): x_a, z_a in {0,1}. */
static const int CX[4] = {0, 1, 1, 0};
static const int CZ[4] = {0, 0, 1, 1};

/* Integral over [0,1] of phi_a(s) phi_b(s), phi_a = s if corner coordinate is
 * 1 else 1-s. */
static double int_pp(int ca, int cb) { return ca == cb ? 1.0 / 3.0 : 1.0 / 6.0; }
/* Sign of d phi_a / ds. */
static double dsign(int c) { return c ? 1.0 : -1.0; }

void orc_calck(double mu, double lam, double k[64])
{
    /* Plane strain, D = [[l+2m, l, 0], [l, l+2m, 0], [0, 0, m]], bilinear
     * shape functions N_a = X_a(x) Z_a(z) on the unit square (the element
     * stiffness of a square is size independent in 2-D). */
    for (int a = 0; a < 4; ++a) {
        for (int b = 0; b < 4; ++b) {
            double xx = dsign(CX[a]) * dsign(CX[b]) * int_pp(CZ[a], CZ[b]); /* N_a,x N_b,x */
            double zz = dsign(CZ[a]) * dsign(CZ[b]) * int_pp(CX[a], CX[b]); /* N_a,z N_b,z */
            double xz = dsign(CX[a]) * dsign(CZ[b]) * 0.25;                  /* N_a,x N_b,z */
            double zx = dsign(CZ[a]) * dsign(CX[b]) * 0.25;                  /* N_a,z N_b,x */
            int ra = 2 * a, rb = 2 * b;
            k[(ra) + 8 * (rb)] = (lam + 2.0 * mu) * xx + mu * zz;
            k[(ra + 1) + 8 * (rb + 1)] = (lam + 2.0 * mu) * zz + mu * xx;
            k[(ra) + 8 * (rb + 1)] = lam * xz + mu * zx;
            k[(ra + 1) + 8 * (rb)] = lam * zx + mu * xz;
        }
    }
}

void orc_calcsqm(double m[64])
{
    /* 4 * integral of N_a N_b over the unit square, same component only
     * (the reference scales it by the lumped element mass rho*h^2/4). */
    memset(m, 0, 64 * sizeof(double));
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            double v = 4.0 * int_pp(CX[a], CX[b]) * int_pp(CZ[a], CZ[b]);
            m[(2 * a) + 8 * (2 * b)] = v;
            m[(2 * a + 1) + 8 * (2 * b + 1)] = v;
        }
}

/* formulation.f90:250-254 */
double orc_calcm(double scl, double rho) { return (scl * scl) * rho; }

/* formulation.f90:219-229 */
void orc_lysmer(double scl, double rho, double vp, double vs, int iszdir, double res[2])
{
    res[0] = vp * rho * scl;
    res[1] = vs * rho * scl;
    if (iszdir) {
        double t = res[0];
        res[0] = res[1];
        res[1] = t;
    }
}

/* formulation.f90:378-393 */
void orc_multdrm(const double k[64], const double u[8], const unsigned char e[8], double p[8])
{
    for (int i = 0; i < 8; ++i) p[i] = 0.0;
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j)
            if (e[j] && !e[i]) {
                p[i] = p[i] - k[i + 8 * j] * u[j];
                p[j] = p[j] + k[j + 8 * i] * u[i];
            }
}

/* formulation.f90:356-363 */
void orc_shapefuncs(double r, double s, double inter[4])
{
    inter[0] = (1 - r) * (1 - s) / 4;
    inter[1] = (1 + r) * (1 - s) / 4;
    inter[2] = (1 + r) * (1 + s) / 4;
    inter[3] = (1 - r) * (1 + s) / 4;
}

/* formulation.f90:53-59 */
double orc_ricker(double amp, double sigma, double t)
{
    double r = exp(-(t * t) / (4 * sigma)) * ((t * t) - 2 * sigma) / (8 * sqrt(U_PI) * pow(sigma, 2.5));
    return r * amp;
}

/* formulation.f90:27-31 */
void orc_recalcstep(double alpha, double length, double sperlen, double *step)
{
    *step = (length / alpha) / sperlen;
}

/* ------------------------------------------------------------------------- */
/* Free field of a homogeneous half-space (pwaves.f90).                       */
/* ------------------------------------------------------------------------- */

int orc_initplanewaves(orc_wave *w, int kind, double mag, double alpha, double beta,
                       double angle, double freq)
{
    memset(w, 0, sizeof(*w));
    w->kind = kind;
    w->alpha = alpha;
    w->beta = beta;
    w->freq = freq;
    if (kind == ORC_WAVE_P) {
        double e = angle, a = mag;
        w->sine = sin(e);
        w->sinf = w->sine * (beta / alpha);
        double f = asin(w->sinf);
        w->cose = cos(e);
        w->cosf = cos(f);
        if (w->cose < F_WANGZERO) {
            return -1; /* grazing incidence (pwaves.f90:32-35) */
        } else if (fabs(w->sine) < F_WANGZERO) {
            w->a = a;
            w->a1 = -w->a;
            w->b1 = 0.0;
            w->cose = 1;
            w->sine = 0;
        } else {
            double sin2e = sin(2 * e), sin2f = sin(2 * f), cos2f = cos(2 * f);
            double r = alpha / beta;
            w->a = a;
            w->a1 = (sin2e * sin2f - (r * r) * (cos2f * cos2f));
            w->a1 = w->a1 / (sin2e * sin2f + (r * r) * (cos2f * cos2f));
            w->a1 = w->a1 * a;
            w->b1 = -(2 * r * sin2e * cos2f);
            w->b1 = w->b1 / (sin2e * sin2f + (r * r) * (cos2f * cos2f));
            w->b1 = w->b1 * a;
        }
        w->ui[0] = w->sine;
        w->ui[1] = -w->cose;
        w->ui[2] = w->sine;
        w->ui[3] = w->cose;
        w->ui[4] = -w->cosf;
        w->ui[5] = w->sinf;
    } else if (kind == ORC_WAVE_SV) {
        double f = angle, b = mag;
        w->sinf = sin(f);
        double fcr = asin(beta / alpha);
        w->sine = w->sinf * (alpha / beta);
        double e = asin(w->sine);
        w->cose = cos(e);
        w->cosf = cos(f);
        if (fabs(w->sinf) < F_WANGZERO) {
            w->b = b;
            w->b1 = -w->b;
            w->a1 = 0.0;
            w->cosf = 1;
            w->sinf = 0;
        } else if ((fcr - f) > F_WANGZERO) {
            double sin2e = sin(2 * e), sin2f = sin(2 * f), cos2f = cos(2 * f);
            double r = alpha / beta;
            w->b = b;
            w->b1 = (sin2e * sin2f - (r * r) * (cos2f * cos2f));
            w->b1 = w->b1 / (sin2e * sin2f + (r * r) * (cos2f * cos2f));
            w->b1 = w->b1 * b;
            w->a1 = r * sin(4.0 * f);
            w->a1 = w->a1 / (sin2e * sin2f + (r * r) * (cos2f * cos2f));
            w->a1 = w->a1 * b;
        } else {
            return -2; /* post-critical SV (pwaves.f90:87-89) */
        }
        w->ui[0] = w->cosf;
        w->ui[1] = w->sinf;
        w->ui[2] = w->sine;
        w->ui[3] = w->cose;
        w->ui[4] = -w->cosf;
        w->ui[5] = w->sinf;
    } else {
        return -3;
    }
    return 0;
}

/* One Ricker wave train evaluated at slowness-projected delay v (pwaves.f90
 * wavep/wavesv bodies): columns displacement, (unused) velocity, acceleration. */
static void ricker3(double freq, double t, double v, double out[3])
{
    double fp = freq * U_PI;
    double s = (fp * (t - v));
    s = s * s;
    double ex = exp(s);
    out[0] = (1.0 - 2.0 * s) / ex;
    out[1] = (2.0 * s - 3.0) * (fp * fp) * (t - v) / ex;
    out[2] = -2.0 * (4.0 * (s * s) - 12.0 * s + 3.0) * (fp * fp) / ex;
}

void orc_planewave(const orc_wave *w, double t, double x1, double x3, double u[6])
{
    double v[3];
    double amp_i;
    if (w->kind == ORC_WAVE_P) {
        v[0] = (x1 * w->sine - x3 * w->cose) / w->alpha;
        v[1] = (x1 * w->sine + x3 * w->cose) / w->alpha;
        v[2] = (x1 * w->sinf + x3 * w->cosf) / w->beta;
        amp_i = w->a;
    } else if (w->kind == ORC_WAVE_SV) {
        v[0] = (x1 * w->sinf - x3 * w->cosf) / w->beta;
        v[1] = (x1 * w->sine + x3 * w->cose) / w->alpha;
        v[2] = (x1 * w->sinf + x3 * w->cosf) / w->beta;
        amp_i = w->b;
    } else {
        for (int i = 0; i < 6; ++i) u[i] = 0.0;
        return;
    }
    double ip[3], rp[3], rs[3];
    ricker3(w->freq, t, v[0], ip);
    ricker3(w->freq, t, v[1], rp);
    ricker3(w->freq, t, v[2], rs);
    for (int c = 0; c < 3; ++c) {
        double ipx = ip[c] * w->ui[0] * amp_i, ipz = ip[c] * w->ui[1] * amp_i;
        double rpx = rp[c] * w->ui[2] * w->a1, rpz = rp[c] * w->ui[3] * w->a1;
        double rsx = rs[c] * w->ui[4] * w->b1, rsz = rs[c] * w->ui[5] * w->b1;
        u[2 * c] = ipx + rpx + rsx;
        u[2 * c + 1] = ipz + rpz + rsz;
    }
}

/* ------------------------------------------------------------------------- */
/* Setup arrays (model.f90).                                                  */
/* ------------------------------------------------------------------------- */

/* model.f90:163-174: the 4 non-hanging entries of an element row, in order. */
static int elemallnhnodes(const int *eptr, const int *eind, const unsigned char *ediv, int ele,
                          int nodes[4])
{
    int n = 0;
    for (int j = eptr[ele] - 1; j < eptr[ele + 1] - 1; ++j)
        if (!ediv[j]) {
            if (n < 4) nodes[n] = eind[j];
            ++n;
        }
    return n == 4 ? 0 : -1;
}

/* model.f90:654-675 */
int orc_setup_hanging(int nelems, const int *eptr, const int *eind, const unsigned char *ediv,
                      int *hanging)
{
    int ind = 0;
    for (int i = 0; i < nelems; ++i)
        for (int j = eptr[i] - 1; j < eptr[i + 1] - 1; ++j)
            if (ediv[j]) {
                if (hanging) {
                    hanging[3 * ind + 0] = eind[j];
                    hanging[3 * ind + 1] = eind[j - 1];
                    hanging[3 * ind + 2] = (j + 1 < eptr[i + 1] - 1) ? eind[j + 1] : eind[eptr[i] - 1];
                }
                ++ind;
            }
    return ind;
}

/* model.f90:1080-1096 */
void orc_create_absorbingbc(int nnodes, const double *nodes, double blength, int totcols,
                            int totrows, int *absorbing)
{
    double left = 0.0 * blength;
    double right = (double)(totcols - 1) * blength;
    double top = (double)(totrows - 1) * blength;
    for (int i = 0; i < nnodes; ++i) {
        absorbing[i] = 0;
        if (realequal(nodes[2 * i], left) || realequal(nodes[2 * i], right))
            absorbing[i] = 2;
        else if (realequal(nodes[2 * i + 1], top))
            absorbing[i] = 1;
    }
}

/* model.f90:1015-1045 (kinds: SHORI=1, SVERL=2, SVERR=4, CORNR=5, CORNL=6) */
int orc_wave_elems(int nelems, const int *eptr, const int *eind, const unsigned char *ediv,
                   const double *nodes, double blength, int wcols, int wrows, int totcols,
                   int *seismic)
{
    double left = (double)(wcols - 1) * blength;
    double right = (double)((totcols - wcols) - 1) * blength;
    double top = (double)(wrows - 1) * blength;
    int cnt = 0;
    for (int i = 0; i < nelems; ++i) {
        int nd[4];
        elemallnhnodes(eptr, eind, ediv, i, nd);
        const double *p1 = &nodes[2 * (nd[0] - 1)], *p2 = &nodes[2 * (nd[1] - 1)];
        int kind = 0;
        if (realequal(p2[0], left) && realequal(p2[1], top))
            kind = 6;
        else if (realequal(p1[0], right) && realequal(p1[1], top))
            kind = 5;
        else if (realequal(p2[0], left) && realunder(p2[1], top))
            kind = 2;
        else if (realequal(p1[0], right) && realunder(p1[1], top))
            kind = 4;
        else if (realover(p2[0], left) && realunder(p1[0], right) && realequal(p1[1], top))
            kind = 1;
        if (kind) {
            if (seismic) {
                seismic[2 * cnt] = i + 1;
                seismic[2 * cnt + 1] = kind;
            }
            ++cnt;
        }
    }
    return cnt;
}

/* ------------------------------------------------------------------------- */
/* The loop.                                                                  */
/* ------------------------------------------------------------------------- */

static void drm_mask(int kind, unsigned char e[8])
{
    /* model.f90:778-793, per node (BL, BR, TR, TL) -> both DOFs */
    int m[4] = {0, 0, 0, 0};
    switch (kind) {
    case 2: m[0] = 1; m[3] = 1; break;             /* SVERL */
    case 4: m[1] = 1; m[2] = 1; break;             /* SVERR */
    case 1: m[2] = 1; m[3] = 1; break;             /* SHORI */
    case 6: m[0] = 1; m[2] = 1; m[3] = 1; break;   /* CORNL */
    case 5: m[1] = 1; m[2] = 1; m[3] = 1; break;   /* CORNR */
    default: break;
    }
    for (int i = 0; i < 4; ++i) e[2 * i] = e[2 * i + 1] = (unsigned char)m[i];
}

/* model.f90:632-647 (P1 literal with `red` declared) or the constraint
 * transpose P2 (communicator.f90:1232 semantics). */
static void sum_reactanchors(const orc_problem *p, double *r)
{
    for (int i = 0; i < p->nhang; ++i) {
        int h = p->hanging[3 * i] - 1, a1 = p->hanging[3 * i + 1] - 1, a2 = p->hanging[3 * i + 2] - 1;
        for (int c = 0; c < 2; ++c) {
            if (p->hang_mode == 1) {
                double red = r[2 * h + c] - (r[2 * a1 + c] + r[2 * a2 + c]) / 2.0;
                r[2 * a1 + c] = r[2 * a1 + c] + red;
                r[2 * a2 + c] = r[2 * a2 + c] + red;
            } else {
                double half = r[2 * h + c] / 2.0;
                r[2 * a1 + c] = r[2 * a1 + c] + half;
                r[2 * a2 + c] = r[2 * a2 + c] + half;
            }
        }
    }
}

/* model.f90:612-624 */
static void set_dispshanging(const orc_problem *p, double *u)
{
    for (int i = 0; i < p->nhang; ++i) {
        int h = p->hanging[3 * i] - 1, a1 = p->hanging[3 * i + 1] - 1, a2 = p->hanging[3 * i + 2] - 1;
        for (int c = 0; c < 2; ++c) u[2 * h + c] = (u[2 * a1 + c] + u[2 * a2 + c]) / 2.0;
    }
}

static int inside_soil(const orc_problem *p, const double pt[2])
{
    if (p->wave_kind == ORC_WAVE_NONE) return 0;
    return realunder(pt[0], p->box_left) || realover(pt[0], p->box_right) ||
           realover(pt[1], p->box_top);
}

/* calcfieldat (formulation.f90:315-363) on the element's 4 non-hanging
 * corners nd (1-based node ids), plus the free field where inside_soil(pt). */
static void sample_element(const orc_problem *p, const orc_wave *w, const int nd[4],
                           const double *u, const double pt[2], double time, double field[2])
{
    const double *n1 = &p->nodes[2 * (nd[0] - 1)], *n3 = &p->nodes[2 * (nd[2] - 1)];
    double par[2];
    par[0] = (pt[0] - n1[0]) / (n3[0] - n1[0]);
    par[1] = (pt[1] - n1[1]) / (n3[1] - n1[1]);
    for (int c = 0; c < 2; ++c) {
        par[c] = 2.0 * par[c] - 1.0;
        if (par[c] < -1.0) par[c] = -1.0;
        if (par[c] > 1.0) par[c] = 1.0;
    }
    double inter[4];
    orc_shapefuncs(par[0], par[1], inter);
    /* matmul(nmat, dofs): sum over the 8 dofs in order; zero entries of nmat
     * contribute exact zeros. */
    double fx = 0.0, fz = 0.0;
    for (int j = 0; j < 4; ++j) {
        fx = fx + inter[j] * u[2 * (nd[j] - 1)];
        fz = fz + inter[j] * u[2 * (nd[j] - 1) + 1];
    }
    field[0] = fx;
    field[1] = fz;
    if (inside_soil(p, pt)) {
        double disps[6];
        orc_planewave(w, time - p->wt0, pt[0] - p->worigin[0], pt[1] - p->worigin[1], disps);
        field[0] = field[0] + disps[0];
        field[1] = field[1] + disps[1];
    }
}

/* visualization.f90:156-181 */
static void visu_field(const orc_problem *p, const orc_wave *w, const int (*corners)[4],
                       const double *u, int ele, const double pt[2], double time,
                       double field[2])
{
    field[0] = field[1] = 0.0;
    if (ele <= 0) return; /* receiver not found: zeros (communicator.f90:1392-1394) */
    sample_element(p, w, corners[ele - 1], u, pt, time, field);
}

/* visu_field for one receiver on a state u (2*nnodes, hanging nodes already
 * projected), at visualization time `time` (m_deltat*step*io_interval,
 * visualization.f90:329): the sampler the loop uses, exported so that it can
 * be pinned on its own. */
int orc_visu_field(const orc_problem *p, const double *u, int ele, const double pt[2],
                   double time, double field[2])
{
    field[0] = field[1] = 0.0;
    if (ele < 0 || ele > p->nelems) return -1;
    orc_wave w;
    memset(&w, 0, sizeof(w));
    if (p->wave_kind != ORC_WAVE_NONE &&
        orc_initplanewaves(&w, p->wave_kind, p->wamp, p->walpha, p->wbeta, p->wangle, p->wfreq) != 0)
        return -2;
    if (ele == 0) return 0;
    int nd[4];
    if (elemallnhnodes(p->eptr, p->eind, p->ediv, ele - 1, nd) != 0) return -3;
    sample_element(p, &w, nd, u, pt, time, field);
    return 0;
}

/* One element's contribution to local dof j (0..7): row j of
 * r = (a3 u - a1 up) m + a2 c.up - K u, in the reference's operation order
 * (the dgemv call adds column 0, 1, ..., 7 to every row). */
static double elem_row(const double *k, const double *lu, const double *lup, double mass,
                       const double *damp, double a1, double a2, double a3, int j)
{
    double r = (a3 * lu[j] - a1 * lup[j]) * mass;
    r = r + a2 * damp[j] * lup[j];
    for (int c = 0; c < 8; ++c) {
        double temp = -1.0 * lu[c];
        r = r + temp * k[j + 8 * c];
    }
    return r;
}

static int run_impl(const orc_problem *p, double *u_out, double *up_out, double *mfull_out,
                    double *recv_out, int nthreads);

int orc_run(const orc_problem *p, double *u_out, double *up_out, double *mfull_out,
            double *recv_out)
{
    return run_impl(p, u_out, up_out, mfull_out, recv_out, 1);
}

/* orc_run with the element loop spread over nthreads OpenMP threads as a
 * node-centric gather: every degree of freedom adds its elements' rows in
 * ascending element order, the order in which the element-by-element scatter
 * adds them, so the result is bit-identical to orc_run. */
int orc_run_mt(const orc_problem *p, double *u_out, double *up_out, double *mfull_out,
               double *recv_out, int nthreads)
{
    return run_impl(p, u_out, up_out, mfull_out, recv_out, nthreads < 1 ? 1 : nthreads);
}

static int run_impl(const orc_problem *p, double *u_out, double *up_out, double *mfull_out,
                    double *recv_out, int nthreads)
{
    const int ne = p->nelems, nn = p->nnodes, nd = 2 * nn;
    const double dt = p->dt;
    const double a1 = 1.0 / (dt * dt), a2 = 1.0 / (2.0 * dt), a3 = 2.0 * a1;
    int rc = 0, nsamp = 0;
    int *nptr = NULL, *nlist = NULL;   /* gather lists (nthreads > 1) */
    double *seisbuf = NULL;            /* layer-element terms (nthreads > 1) */

    double *kmat = (double *)malloc(sizeof(double) * 64 * (size_t)p->nummat);
    double mmat[64];
    int (*corners)[4] = (int (*)[4])malloc(sizeof(int[4]) * (size_t)ne);
    double *mass = (double *)malloc(sizeof(double) * (size_t)ne);
    double *damp = (double *)malloc(sizeof(double) * 8 * (size_t)ne);
    double *u = (double *)calloc((size_t)nd, sizeof(double));
    double *up = (double *)calloc((size_t)nd, sizeof(double));
    double *ro = (double *)calloc((size_t)nd, sizeof(double));
    double *mfull = (double *)calloc((size_t)nd, sizeof(double));
    if (!kmat || !corners || !mass || !damp || !u || !up || !ro || !mfull) {
        rc = -1;
        goto done;
    }
    for (int m = 0; m < p->nummat; ++m)
        orc_calck(p->matconst[5 * m + 0], p->matconst[5 * m + 1], &kmat[64 * m]);
    orc_calcsqm(mmat);

    orc_wave w;
    memset(&w, 0, sizeof(w));
    if (p->wave_kind != ORC_WAVE_NONE) {
        if (orc_initplanewaves(&w, p->wave_kind, p->wamp, p->walpha, p->wbeta, p->wangle,
                               p->wfreq) != 0) {
            rc = -2;
            goto done;
        }
    }

    /* Per-element geometry: getmassof / getdampingof (model.f90:564-604). */
    for (int e = 0; e < ne; ++e) {
        if (elemallnhnodes(p->eptr, p->eind, p->ediv, e, corners[e]) != 0) {
            rc = -3;
            goto done;
        }
        const int *c = corners[e];
        double h = p->nodes[2 * (c[2] - 1) + 1] - p->nodes[2 * (c[0] - 1) + 1]; /* elength */
        const double *mc = &p->matconst[5 * (p->ematprop[e] - 1)];
        mass[e] = orc_calcm(h / 2.0, mc[2]);
        for (int i = 0; i < 8; ++i) damp[8 * e + i] = 0.0;
        for (int i = 0; i < 4; ++i) {
            int ab = p->absorbing[c[i] - 1];
            if (ab == 1) orc_lysmer(h, mc[2], mc[3], mc[4], 1, &damp[8 * e + 2 * i]);
            else if (ab == 2) orc_lysmer(h, mc[2], mc[3], mc[4], 0, &damp[8 * e + 2 * i]);
        }
    }

    /* Gather lists: per node, (element, corner) in ascending element order. */
    if (nthreads > 1) {
        nptr = (int *)calloc((size_t)nn + 1, sizeof(int));
        nlist = (int *)malloc(sizeof(int) * 4 * (size_t)ne);
        if (!nptr || !nlist) {
            rc = -1;
            goto done;
        }
        for (int e = 0; e < ne; ++e)
            for (int j = 0; j < 4; ++j) nptr[corners[e][j]]++;
        for (int n = 0; n < nn; ++n) nptr[n + 1] += nptr[n];
        int *fill = (int *)malloc(sizeof(int) * (size_t)nn);
        if (!fill) {
            rc = -1;
            goto done;
        }
        seisbuf = (double *)malloc(sizeof(double) * 16 * (size_t)(p->nseis > 0 ? p->nseis : 1));
        if (!seisbuf) {
            free(fill);
            rc = -1;
            goto done;
        }
        memcpy(fill, nptr, sizeof(int) * (size_t)nn);
        for (int e = 0; e < ne; ++e)
            for (int j = 0; j < 4; ++j) nlist[fill[corners[e][j] - 1]++] = 4 * e + j;
        free(fill);
    }

    /* init_explicit (explicit.f90:86-107). */
    for (int e = 0; e < ne; ++e)
        for (int i = 0; i < 8; ++i) {
            int d = 2 * (corners[e][i / 2] - 1) + (i % 2);
            mfull[d] = mfull[d] + a2 * damp[8 * e + i] + a1 * mass[e];
        }
    sum_reactanchors(p, mfull);

    /* problem_displacements(0, e_u), problem_displacements(-dt, e_up) (explicit.f90:93-94) */
    if (p->point_kind == 2)
        for (int q = 0; q < p->npoint; ++q) {
            int d = 2 * (p->pointnodes[q] - 1) + 1;
            u[d] = orc_ricker(p->point_amp, p->point_sigma, 0.0 - p->point_start);
            up[d] = orc_ricker(p->point_amp, p->point_sigma, -dt - p->point_start);
        }

    /* Step-0 sample (save_step(0) + visualize). */
    if (recv_out && p->nrecv > 0) {
        for (int r = 0; r < p->nrecv; ++r)
            visu_field(p, &w, (const int(*)[4])corners, u, p->recv_elem[r], &p->recv_xy[2 * r],
                       0.0, &recv_out[2 * r]);
        ++nsamp;
    }

    for (int i = 1; i <= p->nsteps; ++i) {
        memset(ro, 0, sizeof(double) * (size_t)nd);
        /* additive_reactions -> wave_reactions (model.f90:745-800) */
        if (p->wave_kind != ORC_WAVE_NONE && nthreads > 1) {
            /* the layer elements' terms in parallel, added in the same order */
            double outt = (double)i * dt - p->wt0;
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (int s = 0; s < p->nseis; ++s) {
                int ele = p->seismic[2 * s] - 1, bck = p->seismic[2 * s + 1];
                const int *c = corners[ele];
                double uu[8], aa[8];
                unsigned char emask[8];
                for (int j = 0; j < 4; ++j) {
                    const double *x = &p->nodes[2 * (c[j] - 1)];
                    double up6[6];
                    orc_planewave(&w, outt, x[0] - p->worigin[0], x[1] - p->worigin[1], up6);
                    uu[2 * j] = up6[0];
                    uu[2 * j + 1] = up6[1];
                    aa[2 * j] = up6[4];
                    aa[2 * j + 1] = up6[5];
                }
                drm_mask(bck, emask);
                orc_multdrm(&kmat[64 * (p->ematprop[ele] - 1)], uu, emask, &seisbuf[16 * s]);
                orc_multdrm(mmat, aa, emask, &seisbuf[16 * s + 8]);
            }
            for (int s = 0; s < p->nseis; ++s) {
                int ele = p->seismic[2 * s] - 1;
                const int *c = corners[ele];
                for (int j = 0; j < 8; ++j) {
                    int d = 2 * (c[j / 2] - 1) + (j % 2);
                    ro[d] = ro[d] + seisbuf[16 * s + j];
                }
                for (int j = 0; j < 8; ++j) {
                    int d = 2 * (c[j / 2] - 1) + (j % 2);
                    ro[d] = ro[d] + seisbuf[16 * s + 8 + j] * mass[ele];
                }
            }
        } else if (p->wave_kind != ORC_WAVE_NONE) {
            double outt = (double)i * dt - p->wt0;
            for (int s = 0; s < p->nseis; ++s) {
                int ele = p->seismic[2 * s] - 1, bck = p->seismic[2 * s + 1];
                const int *c = corners[ele];
                double uu[8], aa[8], pp[8];
                unsigned char emask[8];
                for (int j = 0; j < 4; ++j) {
                    const double *x = &p->nodes[2 * (c[j] - 1)];
                    double up6[6];
                    orc_planewave(&w, outt, x[0] - p->worigin[0], x[1] - p->worigin[1], up6);
                    uu[2 * j] = up6[0];
                    uu[2 * j + 1] = up6[1];
                    aa[2 * j] = up6[4];
                    aa[2 * j + 1] = up6[5];
                }
                drm_mask(bck, emask);
                orc_multdrm(&kmat[64 * (p->ematprop[ele] - 1)], uu, emask, pp);
                for (int j = 0; j < 8; ++j) {
                    int d = 2 * (c[j / 2] - 1) + (j % 2);
                    ro[d] = ro[d] + pp[j];
                }
                orc_multdrm(mmat, aa, emask, pp);
                for (int j = 0; j < 8; ++j) {
                    int d = 2 * (c[j / 2] - 1) + (j % 2);
                    ro[d] = ro[d] + pp[j] * mass[ele];
                }
            }
        }
        /* Element loop (explicit.f90:129-142). */
        if (nthreads > 1) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (int n = 0; n < nn; ++n) {
                double rx = ro[2 * n], rz = ro[2 * n + 1];
                for (int q = nptr[n]; q < nptr[n + 1]; ++q) {
                    const int e = nlist[q] / 4, jc = nlist[q] % 4;
                    const int *c = corners[e];
                    double lu[8], lup[8];
                    for (int j = 0; j < 8; ++j) {
                        const int d = 2 * (c[j / 2] - 1) + (j % 2);
                        lu[j] = u[d];
                        lup[j] = up[d];
                    }
                    const double *k = &kmat[64 * (p->ematprop[e] - 1)];
                    rx = rx + elem_row(k, lu, lup, mass[e], &damp[8 * e], a1, a2, a3, 2 * jc);
                    rz = rz + elem_row(k, lu, lup, mass[e], &damp[8 * e], a1, a2, a3, 2 * jc + 1);
                }
                ro[2 * n] = rx;
                ro[2 * n + 1] = rz;
            }
        } else
        for (int e = 0; e < ne; ++e) {
            const int *c = corners[e];
            const double *k = &kmat[64 * (p->ematprop[e] - 1)];
            double lu[8], lup[8], lr[8];
            int dofs[8];
            for (int j = 0; j < 8; ++j) {
                dofs[j] = 2 * (c[j / 2] - 1) + (j % 2);
                lu[j] = u[dofs[j]];
                lup[j] = up[dofs[j]];
            }
            for (int j = 0; j < 8; ++j) lr[j] = (a3 * lu[j] - a1 * lup[j]) * mass[e];
            for (int j = 0; j < 8; ++j) lr[j] = lr[j] + a2 * damp[8 * e + j] * lup[j];
            /* dgemv('N', 8, 8, -1, K, 8, u, 1, 1, r, 1), reference BLAS column
             * order: temp = alpha*x(j); y(:) += temp*A(:,j). */
            for (int j = 0; j < 8; ++j) {
                double temp = -1.0 * lu[j];
                for (int ii = 0; ii < 8; ++ii) lr[ii] = lr[ii] + temp * k[ii + 8 * j];
            }
            for (int j = 0; j < 8; ++j) ro[dofs[j]] = ro[dofs[j]] + lr[j];
        }
        /* absolute_reactions -> point_reactions (model.f90:706-715, 809-822) */
        if (p->point_kind == 1)
            for (int q = 0; q < p->npoint; ++q) {
                int d = 2 * (p->pointnodes[q] - 1) + 1;
                ro[d] = ro[d] + orc_ricker(p->point_amp, p->point_sigma, (double)i * dt - p->point_start);
            }
        sum_reactanchors(p, ro);
        memcpy(up, u, sizeof(double) * (size_t)nd);
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
        for (int d = 0; d < nd; ++d) u[d] = ro[d] / mfull[d];
        /* problem_displacements -> point_displacements (model.f90:728-736, 831-843) */
        if (p->point_kind == 2)
            for (int q = 0; q < p->npoint; ++q) {
                int d = 2 * (p->pointnodes[q] - 1) + 1;
                u[d] = orc_ricker(p->point_amp, p->point_sigma, (double)i * dt - p->point_start);
            }
        set_dispshanging(p, u);
        if (recv_out && p->nrecv > 0 && p->recv_interval > 0 && i % p->recv_interval == 0) {
            /* visualize: time = m_deltat*step*io_interval (visualization.f90:329) */
            int k = i / p->recv_interval;
            double time = (dt * (double)k) * (double)p->recv_interval;
            for (int r = 0; r < p->nrecv; ++r)
                visu_field(p, &w, (const int(*)[4])corners, u, p->recv_elem[r], &p->recv_xy[2 * r],
                           time, &recv_out[2 * ((size_t)nsamp * p->nrecv + r)]);
            ++nsamp;
        }
    }
    if (u_out) memcpy(u_out, u, sizeof(double) * (size_t)nd);
    if (up_out) memcpy(up_out, up, sizeof(double) * (size_t)nd);
    if (mfull_out) memcpy(mfull_out, mfull, sizeof(double) * (size_t)nd);
    rc = nsamp;
done:
    free(nptr);
    free(nlist);
    free(seisbuf);
    free(kmat);
    free(corners);
    free(mass);
    free(damp);
    free(u);
    free(up);
    free(ro);
    free(mfull);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* The same loop in 80-bit long double (x86 extended precision, 64-bit        */
/* significand): explicit.f90:86-153 with every state array, element matrix,  */
/* mass, damping and sum in long double, the plane-wave and Ricker values     */
/* evaluated in double as orc_run does and widened, receivers sampled by      */
/* visu_field from u rounded to double.  The per-step rounding of the         */
/* reference's FP64 formula (a3 u - a1 up) m + a2 c up - K u, summed and      */
/* divided by e_mfull, accumulates over thousands of steps; this evaluation   */
/* carries 11 more bits, so it stands for the exact solution of the discrete  */
/* scheme when FP64 implementations of the loop are compared over long runs. */
/* The element loop is a node-centric gather: every degree of freedom adds    */
/* its elements' rows in ascending element order (the scatter's order), over  */
/* nthreads OpenMP threads.                                                   */
/* ------------------------------------------------------------------------- */
typedef long double ld_t;

static void multdrm_ld(const ld_t k[64], const ld_t u[8], const unsigned char e[8], ld_t p[8])
{
    for (int i = 0; i < 8; ++i) p[i] = 0.0L;
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j)
            if (e[j] && !e[i]) {
                p[i] = p[i] - k[i + 8 * j] * u[j];
                p[j] = p[j] + k[j + 8 * i] * u[i];
            }
}

int orc_run_ld(const orc_problem *p, double *u_out, double *up_out, double *recv_out, int nthreads)
{
    const int ne = p->nelems, nn = p->nnodes, nd = 2 * nn;
    const double dtd = p->dt;
    const ld_t dt = dtd;
    const ld_t a1 = 1.0L / (dt * dt), a2 = 1.0L / (2.0L * dt), a3 = 2.0L * a1;
    int rc = 0, nsamp = 0;
    if (nthreads < 1) nthreads = 1;
    ld_t *kmat = (ld_t *)malloc(sizeof(ld_t) * 64 * (size_t)p->nummat);
    ld_t mmat[64];
    int (*corners)[4] = (int (*)[4])malloc(sizeof(int[4]) * (size_t)ne);
    ld_t *mass = (ld_t *)malloc(sizeof(ld_t) * (size_t)ne);
    ld_t *damp = (ld_t *)malloc(sizeof(ld_t) * 8 * (size_t)ne);
    ld_t *u = (ld_t *)calloc((size_t)nd, sizeof(ld_t));
    ld_t *up = (ld_t *)calloc((size_t)nd, sizeof(ld_t));
    ld_t *ro = (ld_t *)calloc((size_t)nd, sizeof(ld_t));
    ld_t *mfull = (ld_t *)calloc((size_t)nd, sizeof(ld_t));
    double *ud = (double *)calloc((size_t)nd, sizeof(double));
    int *nptr = (int *)calloc((size_t)nn + 1, sizeof(int));
    int *nlist = (int *)malloc(sizeof(int) * 4 * (size_t)(ne > 0 ? ne : 1));
    if (!kmat || !corners || !mass || !damp || !u || !up || !ro || !mfull || !ud || !nptr || !nlist) {
        rc = -1;
        goto done;
    }
    {
        /* K(mu, lambda) = mu K(1, 0) + lambda K(0, 1) (linear in the moduli) */
        double kmu[64], klam[64], md[64];
        orc_calck(1.0, 0.0, kmu);
        orc_calck(0.0, 1.0, klam);
        for (int m = 0; m < p->nummat; ++m)
            for (int i = 0; i < 64; ++i)
                kmat[64 * m + i] = (ld_t)p->matconst[5 * m] * kmu[i] + (ld_t)p->matconst[5 * m + 1] * klam[i];
        orc_calcsqm(md);
        for (int i = 0; i < 64; ++i) mmat[i] = md[i];
    }
    orc_wave w;
    memset(&w, 0, sizeof(w));
    if (p->wave_kind != ORC_WAVE_NONE &&
        orc_initplanewaves(&w, p->wave_kind, p->wamp, p->walpha, p->wbeta, p->wangle, p->wfreq) != 0) {
        rc = -2;
        goto done;
    }
    for (int e = 0; e < ne; ++e) {
        if (elemallnhnodes(p->eptr, p->eind, p->ediv, e, corners[e]) != 0) {
            rc = -3;
            goto done;
        }
        const int *c = corners[e];
        const ld_t h = (ld_t)p->nodes[2 * (c[2] - 1) + 1] - (ld_t)p->nodes[2 * (c[0] - 1) + 1];
        const double *mc = &p->matconst[5 * (p->ematprop[e] - 1)];
        mass[e] = (h / 2.0L) * (h / 2.0L) * (ld_t)mc[2];
        for (int i = 0; i < 8; ++i) damp[8 * e + i] = 0.0L;
        for (int i = 0; i < 4; ++i) {
            const int ab = p->absorbing[c[i] - 1];
            const ld_t cp = (ld_t)mc[2] * (ld_t)mc[3] * h, cs = (ld_t)mc[2] * (ld_t)mc[4] * h;
            if (ab == 1) {          /* ZDIR: (rho vs h, rho vp h) */
                damp[8 * e + 2 * i] = cs;
                damp[8 * e + 2 * i + 1] = cp;
            } else if (ab == 2) {   /* XDIR: (rho vp h, rho vs h) */
                damp[8 * e + 2 * i] = cp;
                damp[8 * e + 2 * i + 1] = cs;
            }
        }
    }
    for (int e = 0; e < ne; ++e)
        for (int j = 0; j < 4; ++j) nptr[corners[e][j]]++;
    for (int n = 0; n < nn; ++n) nptr[n + 1] += nptr[n];
    {
        int *fill = (int *)malloc(sizeof(int) * (size_t)(nn > 0 ? nn : 1));
        if (!fill) {
            rc = -1;
            goto done;
        }
        memcpy(fill, nptr, sizeof(int) * (size_t)nn);
        for (int e = 0; e < ne; ++e)
            for (int j = 0; j < 4; ++j) nlist[fill[corners[e][j] - 1]++] = 4 * e + j;
        free(fill);
    }
    for (int e = 0; e < ne; ++e)
        for (int i = 0; i < 8; ++i) {
            const int d = 2 * (corners[e][i / 2] - 1) + (i % 2);
            mfull[d] = mfull[d] + a2 * damp[8 * e + i] + a1 * mass[e];
        }
    for (int i = 0; i < p->nhang; ++i) {   /* sum_reactanchors, P2 */
        const int h = p->hanging[3 * i] - 1, x1 = p->hanging[3 * i + 1] - 1, x2 = p->hanging[3 * i + 2] - 1;
        for (int c = 0; c < 2; ++c) {
            const ld_t half = mfull[2 * h + c] / 2.0L;
            mfull[2 * x1 + c] = mfull[2 * x1 + c] + half;
            mfull[2 * x2 + c] = mfull[2 * x2 + c] + half;
        }
    }
    if (p->point_kind == 2)
        for (int q = 0; q < p->npoint; ++q) {
            const int d = 2 * (p->pointnodes[q] - 1) + 1;
            u[d] = orc_ricker(p->point_amp, p->point_sigma, 0.0 - p->point_start);
            up[d] = orc_ricker(p->point_amp, p->point_sigma, -dtd - p->point_start);
        }
    for (int i = 0; i <= p->nsteps; ++i) {
        if (i > 0) {
            for (int d = 0; d < nd; ++d) ro[d] = 0.0L;
            if (p->wave_kind != ORC_WAVE_NONE) {
                const double outt = (double)i * dtd - p->wt0;
                for (int s = 0; s < p->nseis; ++s) {
                    const int ele = p->seismic[2 * s] - 1, bck = p->seismic[2 * s + 1];
                    const int *c = corners[ele];
                    ld_t uu[8], aa[8], pp[8];
                    unsigned char emask[8];
                    for (int j = 0; j < 4; ++j) {
                        const double *x = &p->nodes[2 * (c[j] - 1)];
                        double up6[6];
                        orc_planewave(&w, outt, x[0] - p->worigin[0], x[1] - p->worigin[1], up6);
                        uu[2 * j] = up6[0];
                        uu[2 * j + 1] = up6[1];
                        aa[2 * j] = up6[4];
                        aa[2 * j + 1] = up6[5];
                    }
                    drm_mask(bck, emask);
                    multdrm_ld(&kmat[64 * (p->ematprop[ele] - 1)], uu, emask, pp);
                    for (int j = 0; j < 8; ++j) ro[2 * (c[j / 2] - 1) + (j % 2)] += pp[j];
                    multdrm_ld(mmat, aa, emask, pp);
                    for (int j = 0; j < 8; ++j) ro[2 * (c[j / 2] - 1) + (j % 2)] += pp[j] * mass[ele];
                }
            }
            /* element rows gathered per node in ascending element order */
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
            for (int n = 0; n < nn; ++n) {
                ld_t rx = ro[2 * n], rz = ro[2 * n + 1];
                for (int q = nptr[n]; q < nptr[n + 1]; ++q) {
                    const int e = nlist[q] / 4, jc = nlist[q] % 4;
                    const int *c = corners[e];
                    const ld_t *k = &kmat[64 * (p->ematprop[e] - 1)];
                    ld_t lu[8], lup[8];
                    for (int j = 0; j < 8; ++j) {
                        const int d = 2 * (c[j / 2] - 1) + (j % 2);
                        lu[j] = u[d];
                        lup[j] = up[d];
                    }
                    for (int comp = 0; comp < 2; ++comp) {
                        const int j = 2 * jc + comp;
                        ld_t r = (a3 * lu[j] - a1 * lup[j]) * mass[e];
                        r = r + a2 * damp[8 * e + j] * lup[j];
                        for (int col = 0; col < 8; ++col) r = r - lu[col] * k[j + 8 * col];
                        if (comp == 0) rx += r;
                        else rz += r;
                    }
                }
                ro[2 * n] = rx;
                ro[2 * n + 1] = rz;
            }
            if (p->point_kind == 1)
                for (int q = 0; q < p->npoint; ++q) {
                    const int d = 2 * (p->pointnodes[q] - 1) + 1;
                    ro[d] = ro[d] + orc_ricker(p->point_amp, p->point_sigma, (double)i * dtd - p->point_start);
                }
            for (int q = 0; q < p->nhang; ++q) {
                const int h = p->hanging[3 * q] - 1, x1 = p->hanging[3 * q + 1] - 1, x2 = p->hanging[3 * q + 2] - 1;
                for (int c = 0; c < 2; ++c) {
                    const ld_t half = ro[2 * h + c] / 2.0L;
                    ro[2 * x1 + c] = ro[2 * x1 + c] + half;
                    ro[2 * x2 + c] = ro[2 * x2 + c] + half;
                }
            }
            for (int d = 0; d < nd; ++d) {
                up[d] = u[d];
                u[d] = ro[d] / mfull[d];
            }
            if (p->point_kind == 2)
                for (int q = 0; q < p->npoint; ++q) {
                    const int d = 2 * (p->pointnodes[q] - 1) + 1;
                    u[d] = orc_ricker(p->point_amp, p->point_sigma, (double)i * dtd - p->point_start);
                }
            for (int q = 0; q < p->nhang; ++q) {
                const int h = p->hanging[3 * q] - 1, x1 = p->hanging[3 * q + 1] - 1, x2 = p->hanging[3 * q + 2] - 1;
                for (int c = 0; c < 2; ++c) u[2 * h + c] = (u[2 * x1 + c] + u[2 * x2 + c]) / 2.0L;
            }
        }
        if (recv_out && p->nrecv > 0 && p->recv_interval > 0 && i % p->recv_interval == 0) {
            const int k = i / p->recv_interval;
            const double time = (dtd * (double)k) * (double)p->recv_interval;
            for (int d = 0; d < nd; ++d) ud[d] = (double)u[d];
            for (int r = 0; r < p->nrecv; ++r)
                visu_field(p, &w, (const int(*)[4])corners, ud, p->recv_elem[r], &p->recv_xy[2 * r], time,
                           &recv_out[2 * ((size_t)nsamp * p->nrecv + r)]);
            ++nsamp;
        }
    }
    if (u_out)
        for (int d = 0; d < nd; ++d) u_out[d] = (double)u[d];
    if (up_out)
        for (int d = 0; d < nd; ++d) up_out[d] = (double)up[d];
    rc = nsamp;
done:
    free(kmat);
    free(corners);
    free(mass);
    free(damp);
    free(u);
    free(up);
    free(ro);
    free(mfull);
    free(ud);
    free(nptr);
    free(nlist);
    return rc;
}

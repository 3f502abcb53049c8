/*
 * oracle.h -- CPU restatement (FP64) of exfem's explicit elastic-wave step loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product (libdamian_hip.so) never links
 * or calls it.
 *
 * Conventions follow the reference Fortran: integer mesh arrays are 1-based
 * (eptr/eind values exactly as model.f90 holds them), matrices are
 * column-major 8x8 (k[i + 8*j] == K(i+1, j+1)), DOF 2n-1 = x, 2n = z with z
 * positive downwards (model.f90:126-133).
 *
 * Pinning status per function is listed in oracle.c's header comment.
 */
#ifndef DMP_ORACLE_H
#define DMP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Wave kinds (model.f90 m_prob 'WAVEP ' / 'WAVESV'). */
enum { ORC_WAVE_NONE = 0, ORC_WAVE_P = 1, ORC_WAVE_SV = 2 };

/* Free-field state filled by orc_initplanewaves (pwaves.f90:8-13 globals). */
typedef struct {
    int kind;
    double alpha, beta, freq;
    double sine, cose, sinf, cosf;
    double a, a1, b, b1;
    double ui[6]; /* f_wui(2,3) column-major */
} orc_wave;

typedef struct {
    int nelems, nnodes;
    const int *eptr;            /* nelems+1, 1-based */
    const int *eind;            /* eptr[nelems]-1 entries, 1-based node ids */
    const unsigned char *ediv;  /* same length, hanging flags (logical*1) */
    const double *nodes;        /* 2*nnodes: (x, z) */
    const int *ematprop;        /* nelems, 1-based material ids */
    int nummat;
    const double *matconst;     /* 5*nummat: mu, lambda, rho, vp, vs */
    const int *absorbing;       /* nnodes: 0 none, 1 ZDIR, 2 XDIR */
    int nhang;
    const int *hanging;         /* 3*nhang: (h, a1, a2), 1-based */
    int nseis;
    const int *seismic;         /* 2*nseis: (elem, kind), 1-based elem */
    int wave_kind;              /* ORC_WAVE_* */
    double worigin[2], wt0, wangle, wamp, wfreq, walpha, wbeta;
    double box_left, box_right, box_top; /* inside_soil box (model.f90:217-235) */
    double dt;
    int nsteps;
    int nrecv;
    const int *recv_elem;       /* nrecv, 1-based element ids (0 = not found) */
    const double *recv_xy;      /* 2*nrecv */
    int recv_interval;          /* sample every k steps, starting at step 0 */
    int hang_mode;              /* 2 = P2 (anchor += R_h/2, default), 1 = P1 literal */
    /* point source (model.f90:706-843): 0 none, 1 POINTN force, 2 POINTD displacement */
    int point_kind;
    double point_amp, point_sigma, point_start;
    int npoint;
    const int *pointnodes;      /* 1-based */
} orc_problem;

/* --- element kernels (formulation.f90) ------------------------------------ */
void orc_calck(double mu, double lam, double k[64]);
void orc_calcsqm(double m[64]);
double orc_calcm(double scl, double rho);
void orc_lysmer(double scl, double rho, double vp, double vs, int iszdir, double res[2]);
void orc_multdrm(const double k[64], const double u[8], const unsigned char e[8], double p[8]);
void orc_shapefuncs(double r, double s, double inter[4]);
void orc_recalcstep(double alpha, double length, double sperlen, double *step);
double orc_ricker(double amp, double sigma, double t);

/* --- free field (pwaves.f90) ----------------------------------------------- */
int orc_initplanewaves(orc_wave *w, int kind, double mag, double alpha, double beta,
                       double angle, double freq);
void orc_planewave(const orc_wave *w, double t, double x1, double x3, double u[6]);

/* --- setup arrays (model.f90) ------------------------------------------------ */
int orc_setup_hanging(int nelems, const int *eptr, const int *eind,
                      const unsigned char *ediv, int *hanging /* 3*count */);
void orc_create_absorbingbc(int nnodes, const double *nodes, double blength,
                            int totcols, int totrows, int *absorbing);
int orc_wave_elems(int nelems, const int *eptr, const int *eind, const unsigned char *ediv,
                   const double *nodes, double blength, int wcols, int wrows,
                   int totcols, int *seismic /* 2*count, may be NULL to count */);

/* --- the loop (explicit.f90:79-159) ---------------------------------------- */
/* Runs init_explicit + nsteps of run_explicit.  Outputs (each may be NULL):
 *   u_out, up_out, mfull_out: 2*nnodes
 *   recv_out: 2*nrecv per sampled step, samples at steps 0, k, 2k, ... <= nsteps
 * Returns number of receiver samples written, or <0 on error. */
int orc_run(const orc_problem *p, double *u_out, double *up_out, double *mfull_out,
            double *recv_out);

/* orc_run's loop in 80-bit long double (same formula and order, free field
 * and Ricker values in double): the exact solution of the discrete scheme to
 * about 1e-15, against which FP64 implementations are compared over long
 * runs, where the reference formula's own FP64 rounding accumulates.
 * Outputs as orc_run (no effective mass); nthreads OpenMP threads. */
int orc_run_ld(const orc_problem *p, double *u_out, double *up_out, double *recv_out, int nthreads);

/* The loop's receiver sampler (visu_field, visualization.f90:156-181): field
 * at point pt of element ele (1-based; 0 = not found -> zeros) from the state
 * u (2*nnodes, hanging nodes projected), plus the free field at `time` - t0
 * where inside_soil(pt).  Returns 0, or <0 on a bad element / wave. */
int orc_visu_field(const orc_problem *p, const double *u, int ele, const double pt[2],
                   double time, double field[2]);

/* orc_run with its element loop on nthreads OpenMP threads (a node-centric
 * gather adding each node's element rows in ascending element order):
 * bit-identical results.  Used as the all-cores CPU baseline. */
int orc_run_mt(const orc_problem *p, double *u_out, double *up_out, double *mfull_out,
               double *recv_out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif

/*
 * damian_step.h -- C-ABI of the MI355X (gfx950) explicit elastic-wave step
 * library: a drop-in for exfem's init_explicit()/run_explicit()
 * (explicit.f90:79-159) and the per-step routines they call.
 *
 * Conventions mirror the reference's Fortran<->C++ layer (io_utils.f90:14-286,
 * io.h:1-38): every argument is passed by pointer so that Fortran can bind it
 * with ISO_C_BINDING and default by-reference passing; integer mesh arrays are
 * the Fortran 1-based values, unchanged; logical*1 arrays are bool (1 byte);
 * inputs stay owned by the caller and are copied.  Unlike io.cpp (which
 * exit()s on error, io.cpp:29-35) every entry point returns an int status,
 * 0 on success, so the Fortran caller can route failures to error_exfem
 * (utilities.f90:1412-1418).  One host thread per handle; calls on a handle
 * are serialized; device work runs on the library's own stream.
 *
 * The whole mesh of an instance lives on one GPU: the single-rank halo calls
 * sum_neighbor_reactions / sum_neighbor_anchors / set_neighbor_displacements
 * (communicator.f90:1157-1344) are identities and have no counterpart here.
 */
#ifndef DAMIAN_STEP_H
#define DAMIAN_STEP_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define DMP_OK 0
#define DMP_EINVAL -1        /* bad argument or inconsistent input arrays    */
#define DMP_ENOMEM -2        /* host or device allocation failed             */
#define DMP_EHIP -3          /* HIP runtime error (no device, launch failure) */
#define DMP_EUNSUPPORTED -4  /* input outside what the library implements    */
#define DMP_ESTATE -5        /* call out of order (e.g. run before init)     */
#define DMP_EWAVE -6         /* grazing P / post-critical SV (pwaves.f90:32-35, 87-89) */

/* Precision modes for dmp_create. */
#define DMP_FP64 0           /* state, forces and update in FP64              */
#define DMP_FP32C 1          /* FP32 forces, displacement as compensated FP32 pair */

/* ---- lifetime ------------------------------------------------------------ */

/* New instance on HIP device *device (0-based).  Replaces the allocation part
 * of init_explicit (explicit.f90:86-90). */
int dmp_create(void **h, const int *precision, const int *device);
int dmp_destroy(void **h);
/* Copies the last error message of h (or of the library when h is NULL) into
 * buf, blank padded to *buflen like a Fortran character variable. */
int dmp_last_error(void *h, char *buf, const int *buflen);

/* ---- model state (model.f90:34-97) -------------------------------------- */

/* Mesh and materials.  eptr(nelems+1), eind/ediv(eptr(nelems+1)-1) are the
 * hanging-flag CSR of model.f90:57-60 (m_eptr, m_eind, m_ediv); nodes is
 * m_nodes(2,nnodes); ematprop m_ematprop(nelems); matconst m_matconst(5,nummat)
 * = (mu, lambda, rho, vp, vs).  elem_props, when not NULL, holds (mu, lambda,
 * rho, vp, vs) per element (5 x nelems) and overrides matconst (heterogeneous
 * media); ematprop is then ignored. */
int dmp_set_mesh(void *h, const int *nelems, const int *eptr, const int *eind,
                 const bool *ediv, const int *nnodes, const double *nodes,
                 const int *ematprop, const int *nummat, const double *matconst,
                 const double *elem_props);

/* Boundary arrays built by problem_setup (model.f90:958-980):
 * absorbing = m_absorbing(nnodes) in {0, AB_ZDIR=1, AB_XDIR=2} (model.f90:17,
 * 1080-1096); hanging = m_hanging(3,nhang) = (node, anchor1, anchor2)
 * (model.f90:654-675); seismic = m_seismic(2,nseis) = (element, kind) with
 * kind in {SHORI=1, SVERL=2, SVERR=4, CORNR=5, CORNL=6} (model.f90:19-20,
 * 1015-1045). */
int dmp_set_boundaries(void *h, const int *absorbing, const int *nhang, const int *hanging,
                       const int *nseis, const int *seismic);

/* Incident plane wave (init_wave, model.f90:982-1001 + initplanewaves,
 * pwaves.f90:16-98).  prob is m_prob (6 characters, 'WAVEP ' or 'WAVESV', no
 * terminator); origin = m_worigin (already snapped); t0 = m_wttoorigin;
 * angle in radians; vp, vs = m_wvelo; box = (left, right, top) of the
 * plane-wave box used by inside_soil (model.f90:217-235). */
int dmp_set_wave(void *h, const char *prob, const double *origin, const double *t0,
                 const double *angle, const double *amp, const double *freq,
                 const double *vp, const double *vs, const double *box);

/* Point source instead of a plane wave (setup_point_source, model.f90:1051-1078):
 * prob 'POINTN' adds amp * ricker(t - start, sigma) (formulation.f90:53-59) to
 * the z reaction of every listed node (point_reactions, model.f90:809-822);
 * 'POINTD' prescribes that value as the z displacement, also at steps 0 and -1
 * (point_displacements, model.f90:831-843; explicit.f90:93-94).  pointnodes
 * is m_pointnodes (1-based). */
int dmp_set_point_source(void *h, const char *prob, const double *amp, const double *sigma,
                         const double *start, const int *npoints, const int *pointnodes);

/* Time step (m_deltat after recalculate_deltat, communicator.f90:1629-1662),
 * the planned number of steps (m_nsteps; informative) and the receiver
 * sampling interval in steps (io_interval; io_utils.f90:390). */
int dmp_set_time(void *h, const double *dt, const int *nsteps, const int *io_interval);

/* Surface receivers (visualize SHEET: find_highest, visualization.f90:286-305;
 * visu_field, 156-181): element (1-based, 0 = not found -> zeros), point
 * (x, z), and whether the free field is added (inside_soil(point)). */
int dmp_set_receivers(void *h, const int *nrecv, const int *recv_elem, const double *recv_xy,
                      const bool *recv_add_freefield);

/* ---- the loop (explicit.f90:79-159) -------------------------------------- */

/* init_explicit: effective lumped mass a1*M + a2*C with the hanging-node
 * transfer (explicit.f90:96-107), u_0 = u_-1 = 0, device layout, first-step
 * plane-wave forces, receiver sample of step 0.  Also reserves the receiver
 * sample ring in pinned host memory: nsteps / io_interval + 2 samples of
 * 16 B per receiver (at least 64 samples, at most 256 MiB), used as a ring,
 * so any run length streams through it. */
int dmp_init(void *h);

/* Advance *nsteps_block steps of run_explicit (explicit.f90:125-154).
 * recv_out (may be NULL): receives 2 x nrecv doubles per receiver sample in
 * this block -- samples are taken at every step j with j % io_interval == 0
 * and j in (s, s+n] where s is the step reached before the call (the first
 * call also returns the step-0 sample first); *nsaved (may be NULL) gets the
 * number of samples written.  The kernels write the samples into the
 * pinned ring reserved by dmp_init; finished chunks are moved to recv_out
 * while the call's later steps run, and a chunk's slots are reused only
 * after it has been moved.  Returns when the call's steps are complete. */
int dmp_run(void *h, const int *nsteps_block, double *recv_out, int *nsaved);

/* dmp_run without allocating anything: no counterpart in the reference.  For
 * callers that cannot allocate device or pinned memory (tight serving loops,
 * harnesses that forbid allocation inside a run).  Same arguments, results
 * and state as dmp_run (bit for bit), but every step block is an eager
 * launch (no graph is instantiated), and the receiver samples go to a 16 MiB
 * device array of the library (one per GPU, for the life of the process)
 * and are copied to recv_out at the end of the call.  Limit: the call's
 * samples plus one must fit in it -- 16 MiB / (16 B x nrecv) samples;
 * beyond that it returns DMP_EINVAL and does nothing.  Calls on one GPU
 * share the array: calls that overlap in time (other handles, other host
 * threads) race. */
int dmp_run_noalloc(void *h, const int *nsteps_block, double *recv_out, int *nsaved);

/* Current state in the caller's node numbering: u = e_u, up = e_up (hanging
 * nodes projected as set_dispshanging does), mfull = e_mfull (what save_mass
 * writes, explicit.f90:108).  Any pointer may be NULL.  In DMP_FP32C mode u
 * and up are hi+lo rounded to double. */
int dmp_get_state(void *h, double *u, double *up, double *mfull);

/* Device time in milliseconds of the last dmp_run / dmp_batch_run call,
 * measured with HIP events on the library's stream from before the call's
 * first kernel to after its last one.  The reference's only timers are
 * whole-second time() counters (io_utils.f90:702-713). */
int dmp_last_run_ms(void *h, float *ms);
int dmp_batch_last_run_ms(void *b, float *ms);

/* Current step index (number of steps completed since dmp_init). */
int dmp_get_step(void *h, int *step);

/* Bit-exactness gate for the integer model arrays: rebuilds m_eptr, m_eind,
 * m_ediv, m_hanging and m_seismic from the device-resident reordered layout
 * and its inverse permutations.  Output sizes are those given to the set_*
 * calls. */
int dmp_export_mesh(void *h, int *eptr, int *eind, bool *ediv, int *hanging, int *seismic);

/* ---- tuning and batching -------------------------------------------------- */

/* Launch parameters (4 ints).  tuning[0] = nodes per tile on average (32..1024;
 * each workgroup owns one of ceil(nodes / tuning[0]) tiles; with several steps
 * per launch and at most two tiles per compute unit the tiles are cut at equal
 * work, so tiles holding hanging-node anchors are smaller), tuning[1] = XCD-aware tile remap (0/1), tuning[2] = steps
 * per captured hipGraph (0 = eager launches), tuning[3] = time steps one
 * launch advances (1..8; each tile then carries that many rings of
 * neighbours in LDS).  dmp_init starts from {256, 1, 60, 3} ({256, 1, 60, 1}
 * above 200,000 nodes), falling back to smaller shapes where the mesh needs
 * it; NULL restores that initial shape.  The shape's graph is instantiated
 * here, so that no dmp_run pays for it.  A shape the mesh cannot use (LDS,
 * descriptor range) returns DMP_EUNSUPPORTED and leaves the previous shape in
 * place. */
int dmp_set_tuning(void *h, const int *tuning);
int dmp_get_tuning(void *h, int *tuning);

/* Times the candidate launch shapes (nodes per tile 64/128/256 and the sizes
 * giving one or two tiles per compute unit, 1, 2, 3, 4, 6 or 8 steps per
 * launch -- 1 or 2 when the layout has more than 2 x 256 nodes per compute
 * unit --, XCD remap on/off, eager launches or graphs of min(*nsteps, 60)
 * steps in whole launches) on this instance for *nsteps steps each, once to
 * warm up, then timed until there are two runs and 2 ms of them (at most 15)
 * (state is restored afterwards), and keeps the shape with the shortest lower
 * median of its timed runs; the result is also written to tuning_out (4 ints)
 * when not NULL. */
int dmp_autotune(void *h, const int *nsteps, int *tuning_out);

/* Batched site-response: *n instances (same precision, device and
 * io_interval; sizes and time steps may differ) are packed into one device
 * layout and advanced together, one launch per step block for all of them.
 * Each instance needs its set_* calls done; dmp_init is not needed and is
 * best left out: the local layouts of instances not initialised are built
 * here, on host threads in parallel, and no per-instance device layout is
 * made.  Initialised instances keep their own state untouched (the batch
 * has its own copy).  The batch has its own receiver ring (as dmp_init, for
 * the longest instance's nsteps). */
int dmp_batch_create(void **b, const int *n, void *const *handles);
int dmp_batch_destroy(void **b);
/* Advance all instances by *nsteps steps.  recv_out (may be NULL) receives,
 * for each sample step, the receivers of instance 0, then instance 1, ... */
int dmp_batch_run(void *b, const int *nsteps, double *recv_out, int *nsaved);
int dmp_batch_get_state(void *b, const int *instance, double *u, double *up);
int dmp_batch_set_tuning(void *b, const int *tuning);
int dmp_batch_autotune(void *b, const int *nsteps, int *tuning_out);

/* ---- output formats ----------------------------------------------------- */

/* save_binstep-compatible snapshot record (io.cpp:454-482): mode 'n' creates
 * the file, anything else appends [int32 step][int32 ndofs][ndofs x float64].
 * nfname/fname are a Fortran character variable and its length. */
int dmp_write_step(const int *nfname, const char *fname, const char *mode, const int *step,
                   const int *ndofs, const double *u);
/* save_binmass-compatible file (io.cpp:409-429). */
int dmp_write_mass(const int *nfname, const char *fname, const int *ndofs, const double *mass);
/* save_binmesh-compatible file (io.cpp:333-360): [nelems][eptr(nelems+1)][eind][ediv]
 * [nnodes][nodes(2,nnodes)], with eptr[nelems] entries of eind and ediv like the
 * reference writer (one more than the mesh holds; written as 0 / false here instead
 * of reading past the arrays). */
int dmp_write_mesh(const int *nfname, const char *fname, const int *nelems, const int *eptr,
                   const int *eind, const bool *ediv, const int *nnodes, const double *nodes);
/* save_bintparams-compatible file (io.cpp:891-910): [float64 deltat][int32 iointerval]. */
int dmp_write_tparams(const int *nfname, const char *fname, const double *deltat,
                      const int *iointerval);
/* save_rank_file-compatible update of ranks.bin (io.cpp:230-267): appends *myrank
 * and increments the count in the header, creating the file when absent. */
int dmp_write_rank(const int *nfname, const char *fname, const int *myrank);

#ifdef __cplusplus
}
#endif
#endif /* DAMIAN_STEP_H */

"""Config 4 at full size: the 64 site-response instances of the batched
sweep (config-3 valleys scaled with their depth, von Karman valley Vs, 37k to
503k elements, 13.9M in total), set up and handed to dmp_batch_create
uninitialised, 10,000 steps in one batch per precision.

Gates: every instance's batched FP32C seismograms against its batched FP64
ones (relative L2 <= 1e-5); the smallest and the largest instance run on
their own in FP64 give the batched FP64 bits; for those two, batched FP64
seismograms and fields against the reference formula evaluated in 80-bit
precision (the oracle's orc_run_ld) over the first 3500 steps, <= 1e-12.
Over that many steps the reference formula's own FP64 rounding drifts by a
few 1e-12 (reported for the smallest instance), so the 80-bit evaluation is
the yardstick."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

NSTEPS = 10000
PREFIX = 3500   # the pulse enters the box by about 0.1 s + d / 1000 s (2900 steps at d = 400 m)


def _mk(k):
    from damian_hip import cases
    return cases.c4_instance(k, NSTEPS)


@pytest.fixture(scope="module")
def built():
    import os
    from multiprocessing import get_context
    with get_context("fork").Pool(min(16, len(os.sched_getaffinity(0)))) as pool:
        out = pool.map(_mk, range(64))
    print(f"config 4: {sum(c.nelems for c in out)} elements, "
          f"{min(c.nelems for c in out)}..{max(c.nelems for c in out)}", flush=True)
    return out


def split(rec, solvers):
    out, k = [], 0
    for s in solvers:
        out.append(rec[:, k:k + s.nrecv])
        k += s.nrecv
    return out


@pytest.mark.timeout(900)
def test_config4_batched_sweep(built):
    import oracle_ctypes as O
    from damian_hip import lib
    sizes = [c.nelems for c in built]
    lo, hi = int(np.argmin(sizes)), int(np.argmax(sizes))
    recs, prefix_state = {}, {}
    for prec in (1, 0):
        print(f"batch prec={prec}", flush=True)
        solvers = [lib.Solver(c, precision=prec, init=False) for c in built]
        b = lib.Batch(solvers)
        r0 = b.run(PREFIX)
        if prec == 0:
            prefix_state = {q: b.state(q) for q in (lo, hi)}
        rec = np.concatenate([r0, b.run(NSTEPS - PREFIX)], axis=0)
        b.close()
        recs[prec] = split(rec, solvers)
        for s in solvers:
            s.close()
    worst = 0.0
    for q, (a, b) in enumerate(zip(recs[1], recs[0])):
        assert a.shape == b.shape == (NSTEPS + 1, len(built[q].recv_elem), 2)
        assert np.abs(b).max() > 0.1, f"instance {q}: the wave must reach the receivers"
        e = rel_l2(a, b)
        worst = max(worst, e)
        assert e <= 1e-5, (q, built[q].name, e)
    print(f"fp32c vs fp64, worst instance rel-L2 {worst:.2e}", flush=True)
    for q in (lo, hi):
        case = built[q]
        print(f"instance {q} ({case.nelems} elements): single FP64 run and oracle prefix", flush=True)
        s = lib.Solver(case, precision=0)
        single = s.run(NSTEPS)
        s.close()
        assert np.array_equal(single, recs[0][q]), "batching must not change a bit"
        # the reference formula in 80-bit precision over the prefix
        u_l, up_l, rec_l = O.run_ld(case, nsteps=PREFIX, recv_interval=1, threads=16)
        assert np.abs(u_l).max() > 1e-3, "the wave must have entered the model"
        u, up = prefix_state[q]
        e_rec = rel_l2(recs[0][q][:PREFIX + 1], rec_l)
        e_u, e_up = rel_l2(u, u_l), rel_l2(up, up_l)
        print(f"instance {q}: FP64 vs the 80-bit oracle over {PREFIX} steps: seismograms "
              f"{e_rec:.2e}, u {e_u:.2e}, u_prev {e_up:.2e}", flush=True)
        assert e_rec <= 1e-12 and e_u <= 1e-12 and e_up <= 1e-12
        if q == lo:
            # the FP64 oracle (the reference formula in FP64) over the same
            # prefix: its own rounding drift, for the record
            u_o, up_o, mf_o, rec_o = O.run(case, nsteps=PREFIX, recv_interval=1, threads=16)
            d_rec, d_u = rel_l2(rec_o, rec_l), rel_l2(u_o, u_l)
            print(f"instance {q}: FP64 oracle vs the 80-bit oracle: seismograms {d_rec:.2e}, "
                  f"u {d_u:.2e}; FP64 path vs FP64 oracle: seismograms "
                  f"{rel_l2(recs[0][q][:PREFIX + 1], rec_o):.2e}", flush=True)
            assert d_rec <= 1e-10 and rel_l2(u, u_o) <= 1e-10

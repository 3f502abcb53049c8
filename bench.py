"""Benchmark of the explicit elastic-wave step loop on one MI355X.

Workload (default): semicircular canyon R = 250 m in a 4000 m x 1500 m
half-space (Vs 1000, Vp 2000, rho 2200), quadtree 20 m -> 2.5 m within 50 m of
the free surface (4 levels, hanging nodes), SV plane wave at 29 deg (30 deg is
the critical angle, rejected), Ricker f0 3 Hz, dt 2.5e-4 s, 201 surface
receivers between the nodes, sampled every step.  A "step" is one time step of
the whole mesh.

Prints ONE JSON line.  `value` is the FP32-compensated throughput in
element-steps/s (elements x steps / wall time of the timed block, receiver
device-to-host copy included); FP64 throughput and the FP32C-vs-FP64
seismogram error over the whole 20,000-step run (an untimed leg) are in the
same line.  `also` holds the uniform 2.5 m mesh of the same canyon (944,492
elements) and the batched sweep of 64 valley site-response instances
(config-3 valleys scaled with their depth, 37k to 503k elements, 13.9M in
total; P 0..60 deg, SV 0..28 deg): batched FP32C and FP64 against the same
instances run one after another in FP64, each autotuned.

    python bench.py [--steps K] [--warmup W] [--workload c2|c2u|c1]
                    [--cpu-seconds S] [--no-cpu] [--no-extra] [--sweep-steps N]
    python bench.py --config4           # only the batched sweep, 10,000 steps
    python bench.py --profile-steps K [--profile-calls N] [--tuning B,X,G,S] [--workload c2|c2u|c1|config4]
                                        # only the timed FP32C block (for rocprofv3)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "damian.par_amd" / "python"))

HBM_PEAK_GBS = 8000.0
METRIC = "element-steps/s (FP32-comp vs FP64, rel-L2<=1e-5); achieved HBM GB/s"


def algorithmic_bytes(case, solver_counts, precision):
    """Compulsory HBM bytes of one step (perfect on-chip reuse):
    FP64: 64 B per state-carrying node (read u_n, u_n-1, write u_n+1: 3 x 2 x 8,
    plus one 8 B coefficient per DOF); FP32C: 56 B per node (hi/lo pairs read
    and written, 4 B increment read and written, 4 B coefficient per DOF);
    20 B per element (4 x int32 corners + int32 material); 12 B per hanging
    node, 24 B per plane-wave layer element, 16 B per receiver sample."""
    na, nh, ns, nr = solver_counts
    per_node = 64 if precision == 0 else 56
    return per_node * na + 20 * case.nelems + 12 * nh + 24 * ns + 16 * nr


def measured_traffic(workload, precision, tuning):
    """HBM bytes per launch of the step kernel for this workload and launch
    shape, from the rocprofv3 FETCH_SIZE / WRITE_SIZE table committed under
    profiles/ (None when the shape was not profiled)."""
    path = ROOT / "profiles" / "step_kernel_traffic.json"
    if not path.exists():
        return None
    table = json.loads(path.read_text())
    for e in table["entries"]:
        if (e["workload"] == workload and e["precision"] == precision
                and e["nodes_per_tile"] == tuning[0] and e["steps_per_launch"] == tuning[3]
                and e.get("xcd_remap", tuning[1]) == tuning[1]):
            return e["bytes_per_launch"]
    return None


def counts(case):
    nh = len(case.hanging)
    return (case.mesh.nnodes - nh, nh, len(case.seismic), len(case.recv_elem))


def tf_at(trace, dt, f=1.0):
    """|FFT| * dt * 2 at the first bin >= f Hz, NFFT = 2^nextpow2(L)
    (sheets/obtainsingle.m:2-22, without the Ricker normalisation, which
    cancels in relative errors)."""
    L = trace.shape[0]
    nfft = 1 << int(math.ceil(math.log2(L)))
    Y = np.fft.rfft(trace, n=nfft, axis=0)
    freqs = (1.0 / dt) / 2 * np.linspace(0, 1, nfft // 2 + 1)
    n = int(np.nonzero(freqs >= f)[0][0])
    return dt * 2 * np.abs(Y[n])


ACCURACY_GATE = 1e-5        # seismogram relative L2, FP32C against FP64
SIGNAL_FLOOR = 1e-3         # receivers below this fraction of the largest trace norm carry no signal


def accuracy(rec32, rec64, dt):
    """FP32C against FP64 seismograms (nsamp, nrecv, 2): relative L2 over all
    receivers and both components (the gate), peak-amplitude and 1 Hz
    transfer-function errors relative to the largest FP64 value, and the
    worst per-receiver relative L2 over receivers whose FP64 trace norm is at
    least SIGNAL_FLOOR of the largest (a trace the wave has not reached holds
    rounding noise that FP32 does not represent)."""
    d = rec32 - rec64
    rel = float(np.linalg.norm(d) / np.linalg.norm(rec64))
    pk64 = np.abs(rec64).max(axis=0)
    pk32 = np.abs(rec32).max(axis=0)
    peak = float(np.max(np.abs(pk32 - pk64)) / pk64.max())
    tf64 = tf_at(rec64, dt)
    tf32 = tf_at(rec32, dt)
    tf = float(np.max(np.abs(tf32 - tf64)) / tf64.max())
    n64 = np.linalg.norm(rec64, axis=0)
    live = n64 >= SIGNAL_FLOOR * n64.max()
    per_recv = np.linalg.norm(d, axis=0)[live] / n64[live]
    return {"rel_l2": rel, "peak_amplitude_err": peak, "tf_1hz_err": tf,
            "max_receiver_rel_l2": float(per_recv.max()),
            "receiver_traces_compared": int(live.sum()), "receiver_traces": int(live.size),
            "gate": ACCURACY_GATE, "gate_ok": bool(rel <= ACCURACY_GATE)}


def accuracy_leg(lib, case, tuning, nsteps):
    """Untimed: the whole run (nsteps steps, every step recorded) in FP32C
    and in FP64 from the same inputs, compared."""
    recs = []
    for prec in (1, 0):
        s = lib.Solver(case, precision=prec, io_interval=1)
        if tuning:
            s.set_tuning(*tuning)
        recs.append(s.run(nsteps))
        s.close()
    out = accuracy(recs[0], recs[1], case.dt)
    out["steps"] = nsteps
    return out


def timed_run(lib, case, precision, warmup, steps, tuning=None):
    s = lib.Solver(case, precision=precision, io_interval=1)
    if tuning:
        s.set_tuning(*tuning)
    r0 = s.run(warmup) if warmup > 0 else None
    t0 = time.perf_counter()
    r1 = s.run(steps)
    wall = time.perf_counter() - t0
    dev_ms = s.last_run_ms()
    recs = [r for r in (r0, r1) if r is not None and len(r)]
    rec = np.concatenate(recs, axis=0)
    s.close()
    return rec, wall, dev_ms


def measure(lib, case, warmup, steps, tuning, label, accuracy_steps=None):
    """Timed FP64 and FP32C runs of `steps` steps after `warmup`; the
    accuracy of FP32C against FP64 over an untimed run of accuracy_steps
    steps (default: the timed runs' own warmup + steps)."""
    out = {}
    recs = {}
    for prec, name in ((0, "fp64"), (1, "fp32c")):
        rec, wall, dev_ms = timed_run(lib, case, prec, warmup, steps, tuning)
        recs[name] = rec
        B = algorithmic_bytes(case, counts(case), prec)
        per_launch_s = dev_ms * 1e-3 / steps
        achieved = B / per_launch_s / 1e9
        out[name] = {"element_steps_per_s": case.nelems * steps / wall,
                     "ms_per_step": wall * 1e3 / steps,
                     "device_ms_per_step": dev_ms / steps,
                     "bytes_per_step": B,
                     "achieved_GBs": achieved, "frac": achieved / HBM_PEAK_GBS}
        print(f"# {label} {name}: {out[name]['element_steps_per_s']:.3e} element-steps/s, "
              f"{out[name]['device_ms_per_step'] * 1e3:.2f} us/step, {achieved:.0f} GB/s",
              file=sys.stderr, flush=True)
    if accuracy_steps is None:
        out["accuracy_fp32c_vs_fp64"] = accuracy(recs["fp32c"], recs["fp64"], case.dt)
        out["accuracy_fp32c_vs_fp64"]["steps"] = warmup + steps
    else:
        out["accuracy_fp32c_vs_fp64"] = accuracy_leg(lib, case, tuning, accuracy_steps)
    a = out["accuracy_fp32c_vs_fp64"]
    print(f"# {label} accuracy over {a['steps']} steps: rel-L2 {a['rel_l2']:.2e} "
          f"(gate {a['gate']:.0e}: {'ok' if a['gate_ok'] else 'FAILED'})", file=sys.stderr, flush=True)
    return out


def _c4_instance(args):
    from damian_hip import cases
    k, nsteps = args
    return cases.c4_instance(k, nsteps)


def config4_instances(nsteps, workers=None):
    """The 64 site-response instances of the batched sweep (cases.c4_instance:
    config-3 valleys scaled with their depth, 37k to 503k elements), meshed
    in parallel processes (the mesh build is the driver's work, not the
    library's)."""
    workers = workers or min(16, len(os.sched_getaffinity(0)))
    if workers <= 1:
        return [_c4_instance((k, nsteps)) for k in range(64)]
    from multiprocessing import get_context
    with get_context("fork").Pool(workers) as pool:
        return pool.map(_c4_instance, [(k, nsteps) for k in range(64)])


def batched_sweep(lib, built, steps, warmup=50, tune_steps=60):
    """Batched site-response sweep: all instances advanced together, one
    launch per step block, FP32C and FP64, each batch set up from instances
    that are set up but not initialised (the library's set-up time is
    setup_s) and autotuned outside the timed region; against the same
    instances run one after another on the plain FP64 path, each autotuned
    outside its timed region.  Accuracy: every instance's seismograms,
    batched FP32C against batched FP64, over warmup + steps steps."""
    elems = [c.nelems for c in built]
    total = sum(elems)

    def batched(prec):
        t0 = time.perf_counter()
        solvers = [lib.Solver(c, precision=prec, io_interval=1, init=False) for c in built]
        b = lib.Batch(solvers)
        setup_s = time.perf_counter() - t0
        tun = b.autotune(tune_steps)
        r0 = b.run(warmup)
        t0 = time.perf_counter()
        r1 = b.run(steps)
        wall = time.perf_counter() - t0
        dev_s = b.last_run_ms() * 1e-3
        b.close()
        rec = np.concatenate([r0, r1], axis=0)
        out, k = [], 0
        for s in solvers:
            out.append(rec[:, k:k + s.nrecv])
            k += s.nrecv
            s.close()
        B = sum(algorithmic_bytes(c, counts(c), prec) for c in built)
        return {"element_steps_per_s": total * steps / wall, "run_s": wall, "setup_s": setup_s,
                "device_ms_per_step": dev_s * 1e3 / steps, "bytes_per_step": B,
                "achieved_GBs": B * steps / dev_s / 1e9, "frac": B * steps / dev_s / 1e9 / HBM_PEAK_GBS,
                "launch_shape": list(tun)}, out

    def sequential():
        wall = 0.0
        for c in built:
            s = lib.Solver(c, precision=0, io_interval=1)
            s.autotune(tune_steps // 2)
            s.run(warmup)
            t0 = time.perf_counter()
            s.run(steps)
            wall += time.perf_counter() - t0
            s.close()
        return total * steps / wall, wall

    m32, rec32 = batched(1)
    print(f"# sweep batched fp32c: {m32['element_steps_per_s']:.3e} element-steps/s, "
          f"{m32['achieved_GBs']:.0f} GB/s, setup {m32['setup_s']:.2f}s, run {m32['run_s']:.2f}s",
          file=sys.stderr, flush=True)
    m64, rec64 = batched(0)
    print(f"# sweep batched fp64: {m64['element_steps_per_s']:.3e} element-steps/s, "
          f"{m64['achieved_GBs']:.0f} GB/s", file=sys.stderr, flush=True)
    seq64, seq_wall = sequential()
    print(f"# sweep sequential fp64 (autotuned each): {seq64:.3e} element-steps/s", file=sys.stderr,
          flush=True)
    acc = [accuracy(a, b, c.dt) for a, b, c in zip(rec32, rec64, built)]
    worst = max(acc, key=lambda x: x["rel_l2"])
    return {"instances": len(built), "elements_total": total, "elements_min": min(elems),
            "elements_max": max(elems), "steps": steps, "warmup": warmup,
            "batched_fp32c": m32, "batched_fp64": m64,
            "sequential_fp64_element_steps_per_s": seq64, "sequential_fp64_run_s": seq_wall,
            "batched_fp32c_over_sequential_fp64": m32["element_steps_per_s"] / seq64,
            "batched_fp32c_over_batched_fp64": m32["element_steps_per_s"] / m64["element_steps_per_s"],
            "setup_s_over_run_s": m32["setup_s"] / m32["run_s"],
            "fp32c_vs_fp64_rel_l2_max": worst["rel_l2"],
            "fp32c_vs_fp64_peak_amplitude_err_max": max(x["peak_amplitude_err"] for x in acc),
            "fp32c_vs_fp64_max_receiver_rel_l2": max(x["max_receiver_rel_l2"] for x in acc),
            "fp32c_vs_fp64_gate_ok": all(x["gate_ok"] for x in acc)}


def host_threads():
    """Host threads for the all-cores baseline: OMP_NUM_THREADS when set
    (the GPU host's CPU share for one GPU), else every CPU this process may
    run on; and the number of CPUs the host shows."""
    visible = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env), visible
    return visible, visible


def cpu_baseline(case, seconds):
    """Oracle (FP64, reference summation order) on rank 0 over the first n
    steps of the same workload, n chosen to take about `seconds`: on one core,
    and on all host threads (element loop as a node-centric gather with the
    same summation order, bit-identical results).  The all-cores figure is
    the reported baseline."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_ctypes as O

    def timed(threads):
        n = 50
        t0 = time.perf_counter()
        O.run(case, nsteps=n, recv_interval=1, threads=threads)
        t1 = time.perf_counter() - t0
        n2 = min(case.nsteps, max(n, int(n * seconds / max(t1, 1e-6))))
        t0 = time.perf_counter()
        O.run(case, nsteps=n2, recv_interval=1, threads=threads)
        return n2, time.perf_counter() - t0

    n1, s1 = timed(1)
    nt, visible = host_threads()
    n2, s2 = timed(nt)
    return {"value": case.nelems * n2 / s2, "unit": "element-steps/s", "cores": nt,
            "kind": "port",
            "sample": f"oracle/oracle.c FP64 loop, first {n2} of the workload's steps "
                      f"({case.nelems} elements, {len(case.recv_elem)} receivers), {s2:.1f} s on "
                      f"{nt} host threads (the host shows {visible} CPUs; OMP_NUM_THREADS, when set, "
                      f"is this GPU's share); first {n1} steps on 1 core: {s1:.1f} s",
            "cores_visible": visible,
            "value_1core": case.nelems * n1 / s1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=19800)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="c2", choices=["c2", "c2u", "c1", "config4"],
                    help="config4: the batched sweep of 64 instances (with --profile-steps only)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--sweep-steps", type=int, default=10000,
                    help="steps of the config-4 batched sweep in the default run")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run the FP32C timed block of this many steps")
    ap.add_argument("--profile-calls", type=int, default=1,
                    help="with --profile-steps: that many run calls of --profile-steps steps, "
                         "after one untimed call of as many (short-call profiles)")
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--tuning", default="",
                    help="launch shape for --profile-steps: nodes per tile, XCD remap, "
                         "steps per graph, steps per launch (default: autotuned)")
    ap.add_argument("--config4", action="store_true",
                    help="only the config-4 batched sweep (64 instances, 10,000 steps)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="copy this many MiB with a plain device copy (counter calibration)")
    args = ap.parse_args()

    if args.copy_calibration:
        # plain device copy: the counter calibration, and the bandwidth a copy
        # reaches on this part (reference for the step kernel's traffic rate)
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        y.copy_(x)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(5):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        print(json.dumps({"copy_calibration_MiB": args.copy_calibration, "launches": 6,
                          "ms_per_copy": ms, "GBs_read_plus_write": 2 * n * 4 / ms / 1e6}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs one GPU only (the whole mesh lives on one device); "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)

    from damian_hip import cases, lib

    if args.config4:
        t0 = time.perf_counter()
        built = config4_instances(50 + 10000)
        mesh_s = time.perf_counter() - t0
        print(f"# config 4: {len(built)} instances, {sum(c.nelems for c in built)} elements, "
              f"meshes built in {mesh_s:.1f}s", file=sys.stderr, flush=True)
        out = batched_sweep(lib, built, 10000)
        out["mesh_build_s"] = mesh_s
        print(json.dumps({"config4": out}))
        return

    if args.workload == "config4":
        # one batched run of the sweep's instances (for profilers)
        if not args.profile_steps:
            sys.exit("bench.py: --workload config4 needs --profile-steps")
        # serial mesh build: no process is forked once a profiler may have
        # initialised the GPU in this one
        built = config4_instances(args.profile_steps, workers=1)
        solvers = [lib.Solver(c, precision=args.precision, io_interval=1, init=False) for c in built]
        b = lib.Batch(solvers)
        if args.tuning:
            b.set_tuning(*(int(x) for x in args.tuning.split(",")))
        b.run(args.profile_steps)
        print(json.dumps({"profile_steps": args.profile_steps, "device_ms": b.last_run_ms(),
                          "elements": sum(c.nelems for c in built), "tuning": args.tuning or "default",
                          "bytes_per_step": sum(algorithmic_bytes(c, counts(c), args.precision)
                                                for c in built)}))
        b.close()
        for x in solvers:
            x.close()
        return

    t0 = time.perf_counter()
    build = {"c2": cases.c2, "c2u": lambda: cases.c2(uniform=True), "c1": cases.c1}[args.workload]
    case = build()
    print(f"# workload {case.name}: {case.nelems} elements, {case.mesh.nnodes} nodes, "
          f"{len(case.hanging)} hanging, built in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)

    if args.profile_steps:
        if args.tuning:
            tuning = tuple(int(x) for x in args.tuning.split(","))
        else:
            tuner = lib.Solver(case, precision=args.precision, io_interval=1)
            tuning = tuner.autotune(60)
            tuner.close()
        s = lib.Solver(case, precision=args.precision, io_interval=1)
        s.set_tuning(*tuning)
        if args.profile_calls > 1:
            s.run(args.profile_steps)
            walls, devs = [], []
            for _ in range(args.profile_calls):
                t0 = time.perf_counter()
                s.run(args.profile_steps)
                walls.append((time.perf_counter() - t0) * 1e3)
                devs.append(s.last_run_ms())
            print(json.dumps({"profile_steps": args.profile_steps, "calls": args.profile_calls,
                              "wall_ms_median": float(np.median(walls)),
                              "device_ms_median": float(np.median(devs)), "tuning": list(tuning)}))
            return
        s.run(args.profile_steps)
        print(json.dumps({"profile_steps": args.profile_steps, "device_ms": s.last_run_ms(),
                          "tuning": list(tuning)}))
        return

    # Launch-shape autotuning on a throwaway instance, with run calls of the
    # timed call's length (at most 120 steps)
    tuner = lib.Solver(case, precision=1, io_interval=1)
    tuning = tuner.autotune(max(1, min(120, args.steps)))
    tuner.close()
    print(f"# autotuned launch shape {tuning}", file=sys.stderr, flush=True)

    main_m = measure(lib, case, args.warmup, args.steps, tuning, case.name,
                     accuracy_steps=case.nsteps)
    f32 = main_m["fp32c"]
    B = f32["bytes_per_step"]
    result = {
        "metric": METRIC,
        "value": f32["element_steps_per_s"],
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("c2: canyon R=250 m in 4000x1500 m, Vs 1000 Vp 2000 rho 2200, quadtree "
                         "20->2.5 m within 50 m of the free surface (4 levels), SV 29 deg, "
                         "f0 3 Hz, dt 2.5e-4 s, 201 receivers") if case.name == "c2" else case.name,
            "arithmetic": "FP32 compensated (hi/lo displacement pair, FP32 forces); FP64 plane-wave "
                          "phase and effective-mass inverse",
            "elements": case.nelems, "nodes": case.mesh.nnodes, "hanging_nodes": len(case.hanging),
            "dt": case.dt, "total_steps": args.warmup + args.steps,
            "launch_shape": {"nodes_per_tile": tuning[0], "xcd_remap": tuning[1],
                             "steps_per_graph": tuning[2], "steps_per_launch": tuning[3]},
        },
        "roofline": {"bound": "hbm", "achieved": f32["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": f32["frac"],
                     "traffic": measured_traffic(case.name, "fp32c", tuning),
                     "bytes_per_launch": B * tuning[3],
                     "steps_per_launch": tuning[3],
                     "launch_us": f32["device_ms_per_step"] * 1e3 * tuning[3]},
        "fp64": main_m["fp64"],
        "fp32c": f32,
        "accuracy_fp32c_vs_fp64": main_m["accuracy_fp32c_vs_fp64"],
        "fp32c_over_fp64": f32["element_steps_per_s"] / main_m["fp64"]["element_steps_per_s"],
    }
    if not args.no_extra:
        extra = {}
        dense = cases.c2(uniform=True)
        t_dense = lib.Solver(dense, precision=1, io_interval=1)
        dt_tuning = t_dense.autotune(120)
        t_dense.close()
        m = measure(lib, dense, 100, 4000, dt_tuning, "c2u")
        extra["c2u_dense_2.5m"] = {"elements": dense.nelems, "steps": 4000, "launch_shape": dt_tuning,
                                   **m}
        t0 = time.perf_counter()
        built = config4_instances(50 + args.sweep_steps)
        mesh_s = time.perf_counter() - t0
        print(f"# config 4: {len(built)} instances, {sum(c.nelems for c in built)} elements, "
              f"meshes built in {mesh_s:.1f}s", file=sys.stderr, flush=True)
        extra["config4_batched_sweep"] = batched_sweep(lib, built, args.sweep_steps)
        extra["config4_batched_sweep"]["mesh_build_s"] = mesh_s
        del built
        result["also"] = extra
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(case, args.cpu_seconds)
    print(json.dumps(result))


if __name__ == "__main__":
    main()

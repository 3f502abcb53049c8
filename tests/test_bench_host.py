"""Host-side logic of bench.py (no GPU): the one-GPU refusal, the algorithmic
byte model, the config-4 instance list and the CPU-baseline thread count."""
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402
from damian_hip import cases  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_algorithmic_bytes_model():
    """64 B per state-carrying node in FP64 (56 B in FP32C), 20 B per element,
    12 B per hanging node, 24 B per plane-wave layer element, 16 B per
    receiver sample, counted on a mesh with hanging nodes."""
    c = cases.f2()
    na, nh, ns, nr = bench.counts(c)
    assert na == c.mesh.nnodes - len(c.hanging) and nh == len(c.hanging) > 0
    assert ns == len(c.seismic) > 0 and nr == len(c.recv_elem) > 0
    b64 = bench.algorithmic_bytes(c, (na, nh, ns, nr), 0)
    b32 = bench.algorithmic_bytes(c, (na, nh, ns, nr), 1)
    assert b64 == 64 * na + 20 * c.nelems + 12 * nh + 24 * ns + 16 * nr
    assert b64 - b32 == 8 * na


def test_config4_instances():
    """The batched sweep's 64 sources and depth draws: P at 0..60 deg (two
    depths each), SV at 0..28 deg (four each), depth of instance k uniform on
    [100, 400] m with seed 4000 + k."""
    from damian_hip import cases
    specs = cases.c4_specs()
    assert len(specs) == 64
    p = sorted(a for kind, a, d in specs if kind == "P")
    sv = sorted(a for kind, a, d in specs if kind == "SV")
    assert p == sorted(list(range(0, 61, 4)) * 2) and sv == sorted(list(range(0, 29, 4)) * 4)
    for k, (kind, a, d) in enumerate(specs):
        assert d == np.random.default_rng(4000 + k).uniform(100.0, 400.0)


def test_config4_instance_is_config3_at_200m():
    """The scaled valley has config 3's mesh and materials at d = 200 m, and
    the smallest instance (d = 102 m) about a quarter of its size."""
    from damian_hip import cases
    a = cases.c4_valley(200.0, "P", 0.0, nsteps=10, field_seed=3003)
    b = cases.c3(nsteps=10)
    assert a.nelems == b.nelems and np.array_equal(a.mesh.eind, b.mesh.eind)
    assert np.array_equal(a.mesh.ematprop, b.mesh.ematprop)
    assert np.array_equal(a.elem_props, b.elem_props)
    k = min(range(64), key=lambda i: cases.c4_specs()[i][2])
    small = cases.c4_instance(k, nsteps=10)
    assert small.elem_props is not None and 20000 < small.nelems < 60000, small.nelems


def test_host_threads(monkeypatch):
    import os
    visible = len(os.sched_getaffinity(0))
    monkeypatch.setenv("OMP_NUM_THREADS", "5")
    assert bench.host_threads() == (5, visible)
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.host_threads() == (visible, visible)

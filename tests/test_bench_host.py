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


def test_config4_instances(monkeypatch):
    """64 instances: P at 0..60 deg (two depth draws each), SV at 0..28 deg
    (four each), depth of instance k from seed 4000 + k in [100, 400] m."""
    seen = []
    monkeypatch.setattr(cases, "valley", lambda **kw: seen.append(kw) or kw)
    out = bench.config4_instances(cases, 123)
    assert len(out) == 64
    p = sorted(kw["angle_deg"] for kw in seen if kw["kind"] == "P")
    sv = sorted(kw["angle_deg"] for kw in seen if kw["kind"] == "SV")
    assert p == sorted(list(range(0, 61, 4)) * 2) and sv == sorted(list(range(0, 29, 4)) * 4)
    for k, kw in enumerate(seen):
        assert kw["depth"] == np.random.default_rng(4000 + k).uniform(100.0, 400.0)
        assert 100.0 <= kw["depth"] <= 400.0 and kw["nsteps"] == 123
        assert kw["h0"] == 10.0 and kw["fine"] == 2


def test_host_threads(monkeypatch):
    import os
    visible = len(os.sched_getaffinity(0))
    monkeypatch.setenv("OMP_NUM_THREADS", "5")
    assert bench.host_threads() == (5, visible)
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.host_threads() == (visible, visible)

"""Extracts the digitized semicircular-canyon surface transfer-function
amplitudes (x/a, |TF| at 1 Hz) shipped with the reference
(sheets/canyonsvx.m, canyonsvy.m, canyonpx.m, canyonpy.m) into
canyon_tf.json.  The reference compares its canyon1 (SV) and canyon4 (P)
runs against these points (sheets/comparison.m:1-41).

    python tests/golden/make_canyon_fixture.py [/root/reference]
"""
import json
import re
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
out = {}
for name in ("canyonsvx", "canyonsvy", "canyonpx", "canyonpy"):
    text = (ref / "sheets" / f"{name}.m").read_text()
    body = text[text.index("[") + 1:text.index("]")]
    pts = [[float(v) for v in row.split()] for row in re.split(r"[\n;]", body) if row.strip()]
    out[name[6:].upper()] = pts
Path(__file__).with_name("canyon_tf.json").write_text(json.dumps(out, indent=0))
print({k: len(v) for k, v in out.items()})

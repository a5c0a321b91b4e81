"""Known-answer vectors for pcg3d / pcg3f and sincos_turn (written by the
oracle; the tests re-derive them independently in numpy).

Run from the repository root: python tests/golden/make_rng_kat.py
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle as O  # noqa: E402

words = np.array([[0, 0, 0], [1, 2, 3], [1920, 1080, 7], [0xFFFFFFFF, 0x80000000, 0x7FFFFFFF],
                  [12345, 67890, 0x9E3779B9]], np.uint32)
seeds = np.array([[0.0, 0.0, 0.0], [0.5, 0.25, 0.125], [0.9999999, 1e-30, 3.0]], np.float32)
u = np.array([0.0, 0.125, 0.25, 0.3, 0.5, 0.74, 0.875, 0.99999994], np.float32)
c, s = O.sincos_turn(u)
out = {
    "pcg3d_in": words.tolist(),
    "pcg3d_out": O.pcg3d(words).tolist(),
    "pcg3f_in": seeds.view(np.uint32).tolist(),
    "pcg3f_out": O.pcg3f(seeds).view(np.uint32).tolist(),
    "sincos_u": u.view(np.uint32).tolist(),
    "sincos_c": c.view(np.uint32).tolist(),
    "sincos_s": s.view(np.uint32).tolist(),
}
(Path(__file__).parent / "rng_kat.json").write_text(json.dumps(out) + "\n")
print("wrote rng_kat.json")

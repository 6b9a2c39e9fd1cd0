"""Digest of the oracle's 256^2 circular dam break after 2000 steps (float32 and float64).

The oracle (oracle/swe_oracle.py) runs the configured case (h = 10 m inside r < 40 cells, 1 m
outside, flat bed, walls, no friction, CFL 0.5, dry floor 1e-6 m) for 2000 CFL steps; the
SHA-256 of its interior state bytes (h, hu, hv as C-ordered (3, 256, 256) arrays), the time
and the dt sequence's digest go to c1_oracle_2000.json. The GPU test compares its own state
with this digest bit for bit, which saves it the oracle's minute of CPU per precision.
Usage: python tests/golden/make_c1_oracle_digest.py
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import swe_oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent / "c1_oracle_2000.json"


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    n = 256
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    h = np.where(np.hypot(i - (n - 1) / 2.0, j - (n - 1) / 2.0) < 40, 10.0, 1.0)
    out = {}
    for dtype in (np.float32, np.float64):
        case = O.DenseCase(h=h, hu=0 * h, hv=0 * h, bed=0 * h, dx=1.0, dy=1.0, bc_x="wall", bc_y="wall")
        o = O.DenseOracle(case, O.Physics(h_min=1e-6, h_cfl=1e-6, cfl=0.5), dtype)
        o.run_steps(2000)
        q = np.ascontiguousarray(o.interior)
        out[np.dtype(dtype).name] = {"steps": 2000, "t": o.t, "state_sha256": digest(q),
                                     "dts_sha256": digest(np.array(o.dts, np.float64)),
                                     "h_max": float(q[0].max()), "h_min": float(q[0].min())}
        print(np.dtype(dtype).name, out[np.dtype(dtype).name], flush=True)
    OUT.write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()

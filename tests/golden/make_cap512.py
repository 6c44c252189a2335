"""Writes tests/golden/cap_512_oracle.npz: the oracle's state after the
cap_512 workload (512 x 512, dx 1500 m, 5 levels, dt 10 a, 3 V-cycles per
step, 20 steps) from the seeded generator in glide_amd/synthetic.py.

The fields are stored on every 3rd row and column (about 1/9 of the values)
together with full-field L2 norms and checksums of the generated inputs.
Run: python tests/golden/make_cap512.py   (about 15 min on 8 cores)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from glide_amd import synthetic as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

STRIDE = 3


def main(n_steps=None):
    w = S.workload("cap_512")
    n_steps = n_steps or w["n_steps"]
    h = O.Hierarchy(w["ny"], w["nx"], w["dx"], n_levels=w["n_levels"])
    h.set_geometry(w["bed"], w["H"])
    h.set_parameters(B=w["B"], beta=w["beta"], smb=w["smb"])
    out = {}
    t0 = time.time()
    for step in range(n_steps):
        u, v, H = h.forward(w["dt"], n_vcycles=w["n_vcycles"])
        print(f"step {step} done at {time.time() - t0:.0f}s", flush=True)
        if step in (0, n_steps - 1):
            tag = f"s{step}"
            for name, a in (("u", u), ("v", v), ("H", H)):
                out[f"{tag}_{name}"] = a[::STRIDE, ::STRIDE].copy()
                out[f"{tag}_{name}_norm"] = np.linalg.norm(a)
    out["stride"] = STRIDE
    out["n_steps"] = n_steps
    out["input_checksums"] = np.array([w["bed"].sum(), w["beta"].sum(), w["H"].sum()])
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cap_512_oracle.npz"), **out)


if __name__ == "__main__":
    main()

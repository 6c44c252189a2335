"""Drop-in check at full size: the cap_512 workload (512 x 512, dx 1500 m,
5 levels, dt 10 a, 3 V-cycles per step, 20 steps) on the GPU against the
oracle's state stored in tests/golden/cap_512_oracle.npz (written by
tests/golden/make_cap512.py): relative L2 on u, v, H <= 1e-10 after step 1
and step 20, plus full-field norms."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import synthetic as S  # noqa: E402
from glide_amd.physics import IcePhysics  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cap_512_oracle.npz")


def test_cap512_twenty_steps_match_oracle():
    gold = np.load(GOLD)
    w = S.workload("cap_512")
    cs = np.array([w["bed"].sum(), w["beta"].sum(), w["H"].sum()])
    assert np.allclose(cs, gold["input_checksums"], rtol=1e-13, atol=0), "generator drifted"
    st = int(gold["stride"])
    ph = IcePhysics(w["ny"], w["nx"], w["dx"], n_levels=w["n_levels"])
    ph.set_geometry(w["bed"], w["H"])
    ph.set_parameters(B=w["B"], beta=w["beta"], smb=w["smb"])
    n = int(gold["n_steps"])
    for step in range(n):
        u, v, H = ph.forward(w["dt"], n_vcycles=w["n_vcycles"])
        if step in (0, n - 1):
            tag = f"s{step}"
            for name, a in (("u", u), ("v", v), ("H", H)):
                a = a.cpu().numpy()
                ref = gold[f"{tag}_{name}"]
                rel = np.linalg.norm(a[::st, ::st] - ref) / np.linalg.norm(ref)
                nrm = abs(np.linalg.norm(a) - float(gold[f"{tag}_{name}_norm"])) / float(gold[f"{tag}_{name}_norm"])
                print(f"step {step} {name}: rel L2 (sampled) {rel:.2e}, norm rel {nrm:.2e}", flush=True)
                assert rel <= 1e-10 and nrm <= 1e-10

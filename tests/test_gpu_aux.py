"""Solver-side auxiliaries: non-finite inputs are rejected, a non-finite
residual raises, a checkpoint continues the time stepping bit for bit, the
per-step JSON log, and determinism (the same step twice, single and batched,
gives the same bits)."""
import ctypes
import json

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import _lib  # noqa: E402
from glide_amd import synthetic as S  # noqa: E402
from glide_amd.ensemble import Ensemble  # noqa: E402
from glide_amd.physics import IcePhysics  # noqa: E402

DEV = torch.device("cuda")


def dome(n_levels=4):
    w = S.workload("dome_128")
    ph = IcePhysics(w["ny"], w["nx"], w["dx"], n_levels=n_levels)
    ph.set_geometry(w["bed"], w["H"])
    ph.set_parameters(B=w["B"], beta=w["beta"], smb=w["smb"])
    return w, ph


def state(ph):
    return tuple(t.clone() for t in (ph.grid.u, ph.grid.v, ph.grid.H, ph.grid.H_prev))


def same(a, b):
    return all(torch.equal(x, y) for x, y in zip(a, b))


def test_count_nonfinite_matches_numpy():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(100_003)
    x[rng.choice(x.size, 37, replace=False)] = np.nan
    x[rng.choice(x.size, 5, replace=False)] = np.inf
    x[7] = -np.inf
    t = torch.as_tensor(x, device=DEV)
    counter = torch.zeros(1, dtype=torch.int64, device=DEV)
    n = ctypes.c_ulonglong()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.call("glide_count_nonfinite", t.data_ptr(), t.numel(), ctypes.c_void_p(counter.data_ptr()), ctypes.byref(n), st)
    assert n.value == int(np.sum(~np.isfinite(x)))
    _lib.call("glide_count_nonfinite", t.data_ptr(), 0, ctypes.c_void_p(counter.data_ptr()), ctypes.byref(n), st)
    assert n.value == 0


def test_nonfinite_inputs_are_rejected_and_nothing_changes():
    w, ph = dome()
    before = (ph.grid.beta.clone(), ph.grid.bed.clone(), ph.grid.H.clone())
    bad = np.array(w["beta"], dtype=np.float64)
    bad[3, 4] = np.nan
    with pytest.raises(ValueError, match="beta"):
        ph.set_parameters(beta=bad)
    bed = np.array(w["bed"], dtype=np.float64)
    H = np.array(w["H"], dtype=np.float64)
    H[0, 0] = np.inf
    with pytest.raises(ValueError, match="thickness"):
        ph.set_geometry(bed, H)
    with pytest.raises(ValueError, match="smb"):
        ph.set_parameters(smb=float("nan"))
    assert same(before, (ph.grid.beta, ph.grid.bed, ph.grid.H))


def test_nonfinite_state_raises_in_a_tracked_step():
    w, ph = dome()
    ph.grid.u[5, 5] = float("nan")          # forced from the host: arithmetic on NaN, no fault
    with pytest.raises(FloatingPointError):
        ph.forward(w["dt"], tol=1e-8, max_vcycles=3)


def test_checkpoint_resume_is_bitwise(tmp_path):
    w, ph = dome()
    for _ in range(2):
        ph.forward(w["dt"], n_vcycles=3)
    path = tmp_path / "ckpt.pt"
    ph.save(path)
    for _ in range(2):
        ph.forward(w["dt"], n_vcycles=3)
    straight = state(ph)
    resumed = IcePhysics.load(path)
    assert resumed.steps == 2
    for _ in range(2):
        resumed.forward(w["dt"], n_vcycles=3)
    assert same(straight, state(resumed))
    assert resumed.steps == 4
    with pytest.raises(ValueError):
        IcePhysics(64, 64, w["dx"], n_levels=3).load_state_dict(torch.load(path, weights_only=True))


def test_step_log(tmp_path):
    w, ph = dome()
    log = tmp_path / "steps.jsonl"
    ph.forward(w["dt"], tol=1e-14, max_vcycles=4, log=log)     # tracked: runs all 4
    history = list(ph.last_history)
    ph.forward(w["dt"], n_vcycles=2, log=log)                   # untracked
    seen = []
    ph.forward(w["dt"], n_vcycles=1, log=seen.append)
    recs = [json.loads(line) for line in log.read_text().splitlines()]
    assert [r["step"] for r in recs] == [1, 2] and seen[0]["step"] == 3
    assert recs[0]["vcycles"] == 4 and recs[0]["rel_residual"] == history and recs[0]["r0"] > 0
    assert recs[1]["vcycles"] == 2 and recs[1]["rel_residual"] == [] and recs[1]["r0"] is None
    assert recs[1]["ms"] > 0
    assert recs[0]["grid"] == [128, 128] and recs[0]["n_levels"] == 4


def test_same_step_twice_gives_the_same_bits():
    w, a = dome()
    _, b = dome()
    a.forward(w["dt"], n_vcycles=3)
    b.forward(w["dt"], n_vcycles=3)
    assert same(state(a), state(b))
    runs = []
    for _ in range(2):
        mem = [dome()[1] for _ in range(3)]
        ens = Ensemble(mem)
        ens.forward(w["dt"], n_vcycles=2)
        torch.cuda.synchronize()
        runs.append([state(m) for m in mem])
        ens.close()
    for x, y in zip(*runs):
        assert same(x, y)

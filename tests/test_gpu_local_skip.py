"""Stopping a level's local passes at the first idle pass (glide_set_local_skip)
changes nothing but the work done: a forward solve of a seeded ice sheet,
alone and as a batched ensemble, ends in the same state with the switch on
and off, and the pass counters show that passes were skipped.  (Idle passes
can only flip -0.0 entries to +0.0, which compare equal.)"""
import ctypes

import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import _lib  # noqa: E402
from glide_amd.ensemble import Ensemble  # noqa: E402
from glide_amd.train_warmstart import TEST_SEED0, member  # noqa: E402


def counts():
    run, skipped = ctypes.c_long(), ctypes.c_long()
    _lib.call("glide_local_pass_counts", ctypes.byref(run), ctypes.byref(skipped), 1)
    return run.value, skipped.value


@pytest.fixture(autouse=True)
def _restore():
    yield
    _lib.call("glide_set_local_skip", 1)


def solve(skip, n=256, levels=4, cycles=6):
    _lib.call("glide_set_local_skip", int(skip))
    ph = member(TEST_SEED0 + 3, n, dx=1500.0, n_levels=levels)
    counts()
    ph.forward(1.0, n_vcycles=cycles, update_geometry=False)
    torch.cuda.synchronize()
    return (ph.grid.u.clone(), ph.grid.v.clone(), ph.grid.H.clone()), counts()


def test_single_domain_same_state_fewer_passes():
    off, (run_off, skip_off) = solve(False)
    on, (run_on, skip_on) = solve(True)
    print(f"local passes: without skipping {run_off} run; with skipping {run_on} run, {skip_on} skipped", flush=True)
    assert skip_off == 0 and run_off == 6 * 3 * 10
    assert run_on + skip_on == run_off and skip_on > 0
    assert all(torch.equal(a, b) for a, b in zip(off, on))


def test_batched_same_state():
    states = {}
    for skip in (False, True):
        _lib.call("glide_set_local_skip", int(skip))
        mem = [member(TEST_SEED0 + 3 + k, n, dx=1500.0, n_levels=lv)
               for k, (n, lv) in enumerate([(256, 4), (128, 3), (256, 4)])]
        ens = Ensemble(mem)
        ens.forward(1.0, n_vcycles=4, update_geometry=False)
        torch.cuda.synchronize()
        states[skip] = [(m.grid.u.clone(), m.grid.v.clone(), m.grid.H.clone()) for m in mem]
        ens.close()
    for a, b in zip(states[False], states[True]):
        assert all(torch.equal(x, y) for x, y in zip(a, b))

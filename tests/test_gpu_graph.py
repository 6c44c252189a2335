"""A FASCD V-cycle captured into a HIP graph.

glide_fascd_vcycle with the idle-local-pass check off launches kernels only:
no host synchronisation, no copies to the host, no allocation and no memset
nodes.  Captured once on one hierarchy and replayed, it must give the same
bits as the same V-cycles run eagerly on a second, identical hierarchy, on
every replay, with eager work and synchronisations between the replays.
"""
import ctypes

import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import _lib  # noqa: E402
from glide_amd.train_warmstart import TEST_SEED0, member  # noqa: E402

SCHED = dict(pre=4, post=4, coarse=10, local_passes=3, local_sweeps=2, boundary_weight=0.5, local_skip=False)


def _setup(seed, dt):
    ph = member(seed, 64, dx=1000.0, n_levels=3)
    ph.set_schedule(**SCHED)
    for lv in ph.grids:
        lv.dt = dt
    ph.grid.H.copy_(ph.grid.H_prev)
    _lib.call("glide_forward_setup", ph._levels(), ph.n_levels, dt, ph._stream())
    return ph


def test_vcycle_captured_in_a_graph_replays_bitwise():
    seed, dt = TEST_SEED0 + 60, 1.0
    a, b = _setup(seed, dt), _setup(seed, dt)
    # one eager V-cycle each first: the same state on both, and every launch
    # shape cached before the capture
    a.vcycle()
    b.vcycle()
    torch.cuda.synchronize()
    assert torch.equal(a.grid.U, b.grid.U)
    r0 = a.residual_norm()

    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        a.vcycle()          # recorded, not run
    torch.cuda.synchronize()
    assert torch.equal(a.grid.U, b.grid.U), "capturing must not execute the V-cycle"

    for rep in range(4):
        graph.replay()
        b.vcycle()
        torch.cuda.synchronize()
        assert torch.equal(a.grid.U, b.grid.U), f"replay {rep}"
        # eager work between replays: the residual norms, read back
        ra, rb = a.residual_norm(), b.residual_norm()
        assert ra == rb
        print(f"replay {rep}: ||r|| / ||r_1|| = {ra / r0:.3e}", flush=True)
    assert ra < r0
    del graph

"""Batched ensembles (libglide_hip.so glide_batch_*) against the same members
solved one at a time with the single-domain path (itself pinned to the oracle
bitwise): every member's state must be bitwise identical, for any tile shape
the autotuner may pick."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import synthetic as S  # noqa: E402
from glide_amd.autotune import tune  # noqa: E402
from glide_amd.ensemble import (BK_RESIDUAL, BK_RESIDUAL_CHILD, BK_VANKA, BK_VANKA_LOCAL,  # noqa: E402
                                Ensemble, stream_solve)
from glide_amd.physics import IcePhysics  # noqa: E402

# uneven sizes and level counts: 3, 2, 3, 2 and 1 level(s)
SMALL = [(64, 64), (48, 96), (96, 96), (32, 64), (16, 32)]
SCHED = dict(pre=2, post=2, coarse=20, local_passes=2, local_sweeps=2, n_newton=10)


def make(d, sched=True):
    ph = IcePhysics(d["ny"], d["nx"], d["dx"], n_levels=d["n_levels"])
    ph.set_geometry(d["bed"], d["H"])
    ph.set_parameters(B=d["B"], beta=d["beta"], smb=d["smb"])
    if sched:
        ph.set_schedule(**SCHED)
    return ph


def state(ph):
    return ph.grid.U.detach().cpu().numpy().copy()


@pytest.fixture(scope="module")
def members():
    return S.ensemble_members(SMALL, seed0=500, dx=2000.0)


def test_batched_vcycles_bitwise_equal_individual(members):
    print("members", [(d["ny"], d["nx"], d["n_levels"]) for d in members], flush=True)
    solo = [make(d) for d in members]
    for ph in solo:
        ph.forward(10.0, n_vcycles=2)
    batch = [make(d) for d in members]
    ens = Ensemble(batch)
    ens.set_schedule(**SCHED)
    assert ens.n_phases == 3 and ens.entries_per_phase == [2, 4, 5]
    ens.forward(10.0, n_vcycles=2)
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(zip(solo, batch)):
        sa, sb = state(a), state(b)
        assert np.array_equal(sa, sb), f"member {k}: max diff {np.max(np.abs(sa - sb))}"
        assert np.array_equal(a.grid.H_prev.cpu().numpy(), b.grid.H_prev.cpu().numpy())
    ens.close()


def test_batched_boundary_weight_bitwise_equal_individual(members):
    sched = dict(SCHED, boundary_weight=0.5)
    solo = [make(d, sched=False) for d in members]
    for ph in solo:
        ph.set_schedule(**sched)
        ph.forward(10.0, n_vcycles=2)
    batch = [make(d, sched=False) for d in members]
    ens = Ensemble(batch)
    ens.set_schedule(**sched)
    ens.forward(10.0, n_vcycles=2)
    torch.cuda.synchronize()
    ref = [make(d) for d in members]          # the reference restriction gives other states
    for ph in ref:
        ph.forward(10.0, n_vcycles=2)
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(zip(solo, batch)):
        assert np.array_equal(state(a), state(b)), f"member {k}"
    assert any(not np.array_equal(state(a), state(c)) for a, c in zip(solo, ref))
    ens.close()


def test_tile_shapes_do_not_change_results(members):
    ref = [make(d) for d in members]
    e1 = Ensemble(ref)
    e1.set_schedule(**SCHED)
    e1.forward(10.0, n_vcycles=1)
    alt = [make(d) for d in members]
    e2 = Ensemble(alt)
    e2.set_schedule(**SCHED)
    for p in range(e2.n_phases + 1):
        e2.set_tiles(BK_VANKA, p, 1)
        e2.set_tiles(BK_VANKA_LOCAL, p, 8)
        e2.set_tiles(BK_RESIDUAL, p, 16, 1)
        if p < e2.n_phases - 1:
            e2.set_tiles(BK_RESIDUAL_CHILD, p, 2, 3)
    e2.forward(10.0, n_vcycles=1)
    torch.cuda.synchronize()
    for a, b in zip(ref, alt):
        assert np.array_equal(state(a), state(b))
    e1.close()
    e2.close()


def test_batched_norms_and_residual_match_single(members):
    batch = [make(d) for d in members]
    ens = Ensemble(batch)
    ens.set_schedule(**SCHED)
    ens.forward(10.0, n_vcycles=1, update_geometry=False)
    norms = ens.residual_norms()
    for k, ph in enumerate(batch):
        r_batched = ph.grid.r.detach().cpu().numpy().copy()   # written by the batched residual
        single = ph.residual_norm()                            # single-domain residual + norm
        assert np.array_equal(r_batched, ph.grid.r.detach().cpu().numpy())
        assert abs(norms[k] - single) <= 1e-13 * single, (k, norms[k], single)
    ens.close()


def test_disabled_members_are_untouched(members):
    batch = [make(d) for d in members]
    ens = Ensemble(batch)
    ens.set_schedule(**SCHED)
    ens.forward_setup(10.0)
    before = [state(ph) for ph in batch]
    ens.set_enabled([True, False, True, False, True])
    ens.vcycle()
    torch.cuda.synchronize()
    after = [state(ph) for ph in batch]
    for k in (1, 3):
        assert np.array_equal(before[k], after[k])
    for k in (0, 2, 4):
        assert not np.array_equal(before[k], after[k])
    ens.close()


def test_restriction_timing_hook(members):
    """glide_batch_time_restriction counts 8 (N_total + N_total,child) per
    member with a child in the phase, returns a positive time, and leaves
    every member's state untouched (it writes the children's r only)."""
    batch = [make(d) for d in members]
    ens = Ensemble(batch)
    ens.set_schedule(**SCHED)
    ens.forward(10.0, n_vcycles=1)
    torch.cuda.synchronize()
    before = [state(ph) for ph in batch]

    def ntot(ny, nx):
        return ny * (nx + 1) + (ny + 1) * nx + ny * nx

    for p in range(ens.n_phases):
        ms, nb = ens.time_restriction(p, reps=3)
        # member m's level k sits in phase k + (L_max - L_m); it has a child if k < L_m - 1
        expect = 0
        for d in members:
            k = p - (ens.n_phases - d["n_levels"])
            if 0 <= k < d["n_levels"] - 1:
                ny, nx = d["ny"] >> k, d["nx"] >> k
                expect += 8 * (ntot(ny, nx) + ntot(ny // 2, nx // 2))
        assert nb == expect, (p, nb, expect)
        assert (ms > 0) == (expect > 0)
    torch.cuda.synchronize()
    for k, ph in enumerate(batch):
        assert np.array_equal(before[k], state(ph)), f"member {k}"
    ens.close()


def test_tolerance_mode_stops_converged_members():
    # ice-covered slabs converge; each member stops at its own V-cycle
    sizes = [(32, 32), (64, 64), (64, 32)]
    ds = []
    for k, (ny, nx) in enumerate(sizes):
        d = S.ice_slab(ny, nx, 2000.0, seed=40 + k)
        d.update(ny=ny, nx=nx, n_levels=S.max_levels(ny, nx))
        ds.append(d)
    solo = [make(d, sched=False) for d in ds]
    for ph in solo:
        ph.forward(10.0, tol=1e-10, max_vcycles=60, update_geometry=False)
    batch = [make(d, sched=False) for d in ds]
    ens = Ensemble(batch)
    st = ens.forward(10.0, tol=1e-10, max_vcycles=60, update_geometry=False)
    print("vcycles", st.vcycles, "solo", [len(p.last_history) for p in solo], flush=True)
    assert all(st.converged)
    for k, ph in enumerate(solo):
        assert st.vcycles[k] == len(ph.last_history)
        assert np.array_equal(state(ph), state(batch[k]))
    ens.close()


def test_autotuner_picks_shapes_and_keeps_results(members, tmp_path):
    batch = [make(d) for d in members]
    ens = Ensemble(batch)
    ens.set_schedule(**SCHED)
    rec = tune(ens, reps=2, path=str(tmp_path / "tune.json"))
    assert set(rec["choice"]["vanka"]) == {str(p) for p in range(ens.n_phases)}
    assert all(ty in (1, 2, 4, 8) for ty in rec["choice"]["vanka"].values())
    assert all(ty in (2, 4, 8, 16) for ty in rec["choice"]["residual"].values())
    rec2 = tune(ens, reps=2, path=str(tmp_path / "tune.json"))
    assert rec2["cached"] and rec2["choice"] == rec["choice"]
    ens.forward(10.0, n_vcycles=1)
    solo = [make(d) for d in members]
    for ph in solo:
        ph.forward(10.0, n_vcycles=1)
    torch.cuda.synchronize()
    for a, b in zip(solo, batch):
        assert np.array_equal(state(a), state(b))
    ens.close()


def test_full_size_members_one_vcycle():
    # one member of each ensemble size, full schedule, one V-cycle
    sizes = [(256, 256), (384, 512), (512, 512), (768, 768)]
    ds = S.ensemble_members(sizes, seed0=1000)
    batch = [make(d, sched=False) for d in ds]
    ens = Ensemble(batch)
    assert ens.n_phases == 6
    ens.forward(10.0, n_vcycles=1, update_geometry=False)
    for k, d in enumerate(ds):
        solo = make(d, sched=False)
        solo.forward(10.0, n_vcycles=1, update_geometry=False)
        torch.cuda.synchronize()
        assert np.array_equal(state(solo), state(batch[k])), f"member {d['ny']}x{d['nx']}"
        del solo
    ens.close()


def test_stream_solve_jobs_equal_their_solo_solves():
    """Continuous batching: 7 one-step jobs (64^2 sheets, 3 levels, seeds
    600-606) through 3 slots, each to 1e-8 of its cold residual.  Every job's
    state when it ends equals the same job solved alone for the same number
    of V-cycles, bit for bit, whatever the slot and whatever ran there
    before; slots are refilled as soon as a job ends."""
    from glide_amd.physics import TOLERANCE_SCHEDULE

    n, dx, n_jobs = 64, 1500.0, 7
    data = [S.ice_sheet(n, n, dx, 600 + k) for k in range(n_jobs)]

    def fresh(d):
        ph = IcePhysics(n, n, dx, n_levels=3)
        ph.set_geometry(d["bed"], d["H"])
        ph.set_parameters(B=d["B"], beta=d["beta"], smb=d["smb"])
        ph.set_schedule(**TOLERANCE_SCHEDULE)
        return ph

    def load(k, ph):
        d = data[k]
        ph.set_geometry(d["bed"], d["H"])
        ph.set_parameters(B=d["B"], beta=d["beta"], smb=d["smb"])

    slots = [fresh(data[0]) for _ in range(3)]
    ens = Ensemble(slots)
    ens.set_schedule(**TOLERANCE_SCHEDULE)
    done = {}
    st = stream_solve(ens, load, n_jobs, dt=1.0, tol=1e-8, max_vcycles=80,
                      on_done=lambda k, ph: done.__setitem__(k, state(ph)))
    ens.close()
    assert [j["job"] for j in st.jobs] == list(range(n_jobs)) and all(j["converged"] for j in st.jobs)
    assert len({j["slot"] for j in st.jobs}) == 3
    assert st.batched_vcycles < sum(j["vcycles"] for j in st.jobs)   # slots ran jobs side by side
    print("jobs", [(j["job"], j["slot"], j["vcycles"]) for j in st.jobs], "batched V-cycles", st.batched_vcycles,
          "occupancy", round(st.occupancy, 3), flush=True)
    for j in st.jobs:
        ph = fresh(data[j["job"]])
        ph.forward(1.0, n_vcycles=j["vcycles"], update_geometry=False)
        torch.cuda.synchronize()
        assert np.array_equal(state(ph), done[j["job"]]), f"job {j['job']}"

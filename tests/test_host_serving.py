"""Host-side logic of the server and of the evaluation measures (no GPU)."""
import numpy as np
import pytest

from miccai2016_amd import evaluation as ev
from miccai2016_amd import phantom as ph
from miccai2016_amd import serving


def test_fast_dice_for_matches_definition():
    rng = np.random.default_rng(0)
    seg = rng.uniform(size=(9, 8, 7))
    truth = (rng.uniform(size=seg.shape) < 0.3).astype(np.uint8)
    s = seg >= 0.5
    want = 2 * np.sum(s & (truth > 0)) / (s.sum() + truth.sum())
    assert ev.fast_dice_for(seg, truth) == pytest.approx(want, abs=0)
    assert ev.fast_dice_for(truth, truth) == 1.0
    assert ev.fast_dice_for(np.full(5, 0.5), np.zeros(5)) == 0.0
    # 0 / 0 is NaN in the reference's true division
    assert np.isnan(ev.fast_dice_for(np.zeros(5), np.zeros(5)))
    assert np.isnan(ph.fast_dice(np.zeros(5), np.zeros(5)))
    assert ph.dice_between(np.zeros(5), np.zeros(5)) == 1.0


def test_select_largest_and_dice_for():
    seg = np.zeros((6, 6, 6))
    seg[0, 0, 0:3] = 1        # 3 voxels, label 1
    seg[3, 3, 0:3] = 0.7      # 3 voxels, label 2 (tie: the first label wins)
    seg[5, 0:2, 5] = 0.6      # 2 voxels
    seg[2, 2, 2] = 0.4        # below the threshold
    comp, size = ev.select_largest(seg, return_size=True)
    assert size == 3 and comp[0, 0, 0] and not comp[3, 3, 0] and comp.sum() == 3
    # diagonal neighbours are separate components (face connectivity)
    d = np.zeros((3, 3, 3))
    d[0, 0, 0] = d[1, 1, 1] = 1
    assert ev.select_largest(d, return_size=True)[1] == 1
    truth = np.zeros_like(seg, dtype=np.uint8)
    truth[0, 0, 0:4] = 1
    assert ev.dice_for(seg, truth) == pytest.approx(2 * 3 / (3 + 4))
    assert ev.select_largest(np.zeros((2, 2, 2)), return_size=True) == (None, -1)
    with pytest.raises(ValueError):
        ev.dice_for(np.zeros((2, 2, 2)), np.ones((2, 2, 2)))


def test_synthetic_requests_follow_the_workload():
    rng = np.random.default_rng(5)
    for _ in range(200):
        r = serving.synthetic_request(rng)
        d = np.asarray(r.shape)
        assert d.max() / d.min() <= 2.0 + 1e-9
        assert 300 <= r.num_iter <= 1000
        assert 3 <= r.r_tube <= 8 and 20 <= r.pitch <= 80 and 0.1 <= r.sigma <= 1.5
        assert r.r_helix == 0.25 * d.min()
        assert r.weighting in ph.GRID_WEIGHTINGS and r.decay in ph.GRID_DECAYS
        # base side log-uniform in [64, 320]: the geometric mean of the sides stays within
        # sqrt(2) of that range
        g = np.exp(np.log(d).mean())
        assert 64 / np.sqrt(2) - 1 <= g <= 320 * np.sqrt(2) + 1


def test_footprints():
    shape = (10, 20, 30)
    n = 6000
    base = serving.job_footprint_bytes(shape, 500, 30)
    # fields (28 B per voxel) + state (q4 16, hp 12, pt 4, tmp 4 B per voxel and a few
    # 256-byte aligned control blocks)
    assert 64 * n <= base <= 64 * n + 16 * 1024
    assert serving.job_footprint_bytes(shape, 5000, 1) > base    # a longer criterion log
    z = np.zeros(shape, np.float32)
    cut = serving.CutRequest(z, np.zeros(shape + (3,), np.float32), z, z, z)
    assert serving.request_footprint_bytes(cut, 500, 30) == base
    vol = serving.VolumeRequest(z, 1.0, 0.1, 1.0)
    assert serving.request_footprint_bytes(vol, 500, 30) == serving.PREPARE_BYTES_PER_VOXEL * n
    ph_req = serving.PhantomCutRequest(shape, 3, 2, 20, 1, 0.1, 1.0)
    assert serving.request_footprint_bytes(ph_req, 500, 30) == base + 5 * n
    with pytest.raises(ValueError):
        serving.job_footprint_bytes((2000, 2000, 1000), 500, 30)


def test_ragged_set_follows_the_workload():
    specs, w, d = ph.ragged_set(64, seed=9)
    assert len(specs) == 64 and len(set(sp.seed for sp in specs)) == 64
    for sp in specs:
        assert all(48 <= v <= 192 for v in sp.shape)
        assert 3 <= sp.r_tube <= 8 and 20 <= sp.pitch <= 80 and 0.1 <= sp.sigma <= 1.5
        assert sp.r_helix == 0.25 * min(sp.shape)
    assert set(w) <= set(ph.GRID_WEIGHTINGS) and set(d) <= set(ph.GRID_DECAYS)
    again = ph.ragged_set(64, seed=9)[0]
    assert [sp.shape for sp in again] == [sp.shape for sp in specs]


def _bare_server():
    """A MaxFlowServer with its bookkeeping only (no device)."""
    import threading
    srv = object.__new__(serving.MaxFlowServer)
    srv._lock = threading.Lock()
    srv.reserved = 0
    srv._inflight = 0
    srv.completed = []
    srv._active = []
    srv._jobs_arr = None
    return srv


def test_failed_finish_releases_each_request_once():
    # two resident jobs finish in one step; the second one's finish raises:
    # the first stays finished, the second fails, and the budget and the
    # in-flight count return to zero exactly once each
    srv = _bare_server()
    jobs = []
    for k, need in enumerate((1000, 2500)):
        t = serving.Ticket(serving.PhantomCutRequest((4, 4, 4), 1, 1, 10, 0.5, 0.1, 1.0))
        job = serving._Job(t, (4, 4, 4), None, need)
        jobs.append(job)
        srv.reserved += need
        srv._inflight += 1
    srv._active = list(jobs)

    def finish(job):
        if job is jobs[1]:
            raise RuntimeError("download failed")
        with srv._lock:
            job.ticket._released = True
            srv.reserved -= job.reserved
            srv._inflight -= 1
        job.ticket._done.set()
        srv.completed.append(job.ticket)

    srv._finish = finish
    with pytest.raises(RuntimeError):
        srv._retire(jobs)
    assert srv._active == [jobs[1]]
    srv._fail_all([], RuntimeError("download failed"))
    assert srv.reserved == 0 and srv._inflight == 0 and srv._active == []
    assert jobs[0].ticket.error is None and jobs[0].ticket.done()
    assert isinstance(jobs[1].ticket.error, RuntimeError) and jobs[1].ticket.done()
    assert [t is jobs[0].ticket for t in srv.completed] == [True, False]
    # failing again (a later loop turn) changes nothing
    srv._fail(jobs[0].ticket, jobs[0].reserved, RuntimeError("again"))
    srv._fail(jobs[1].ticket, jobs[1].reserved, RuntimeError("again"))
    assert srv.reserved == 0 and srv._inflight == 0 and len(srv.completed) == 2


def test_full_queue_refuses_requests():
    import collections
    import threading
    from miccai2016_amd.processing.continuous_cut_anisotropic_gpu import CmfCutQAnisotropicGpu
    srv = _bare_server()
    srv.params = CmfCutQAnisotropicGpu.CutParameters()
    srv.resident_min = serving.RESIDENT_MIN_VOXELS
    srv.budget = 1 << 40
    srv.max_queue = 2
    srv.rejected = 0
    srv.peak_queue = 0
    srv._pending = collections.deque()
    srv._wake = threading.Condition(srv._lock)
    req = serving.PhantomCutRequest((8, 8, 8), 2, 2, 10, 0.5, 0.1, 1.0, num_iter=10)
    tickets = [srv.submit(req) for _ in range(4)]
    assert [t.done() for t in tickets] == [False, False, True, True]
    with pytest.raises(serving.OverloadError):
        tickets[2].wait(0)
    assert srv.rejected == 2 and srv.peak_queue == 2 and len(srv._pending) == 2

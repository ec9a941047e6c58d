"""GPU tests of the resident-job solver and the continuous-batching server:
a cut run in chunks beside other cuts of other shapes computes what it
computes alone, bit for bit (the same per-voxel arithmetic in the same
order, without contraction into fused multiply-adds), admission never
reserves more than the budget, and every request kind returns its result."""
import ctypes as C

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from miccai2016_amd import _lib
from miccai2016_amd import phantom as ph

pytestmark = pytest.mark.gpu


def _squeezed(rot):
    from oracle import ref32
    return ref32.qsqueeze(rot.as_quat()[..., [3, 0, 1, 2]].astype(np.float32))


def _case(seed, shape, sigma=0.8):
    spec = ph.HelixSpec(shape, 3, min(shape[1:]) / 4, 20, sigma, seed=seed)
    img, truth = ph.make_helix(spec)
    q = _squeezed(Rotation.random(int(np.prod(shape)), random_state=seed)).reshape(shape + (3,))
    cs, ct = ph.terminal_capacities(img, sigma)
    alpha = ph.nonterminal_weight(ph.grad_mag_sq(img), 10 ** -1.625, 10 ** 1.5)
    return (img.astype(np.float32), q, alpha.astype(np.float32), cs.astype(np.float32),
            ct.astype(np.float32), truth)


def _params(num_iter=200, update=30, error_bound=0.0):
    from miccai2016_amd.processing.continuous_cut_anisotropic_gpu import CmfCutQAnisotropicGpu
    p = CmfCutQAnisotropicGpu.CutParameters()
    p.num_iter, p.update, p.error_bound, p.steps = num_iter, update, error_bound, 0.16
    return p


def _alone(dev, case, eigs, num_iter, update=30, error_bound=0.0):
    from miccai2016_amd.processing.continuous_cut_anisotropic_gpu import CmfCutQAnisotropicGpu
    c = CmfCutQAnisotropicGpu(_params(num_iter, update, error_bound), dev)
    u = c.execute(case[0], case[1], eigs, case[2], case[3], case[4], verbose=False)
    return u, c.last_result


SHAPES = [(16, 18, 20), (9, 31, 12), (24, 8, 17), (5, 5, 70)]
EIGS = (1.0, 10 ** -1.5, 10 ** -1.5)


@pytest.mark.parametrize("resident_min", [1 << 40, 0, 5000])   # all packed, all resident, mixed
def test_chunked_jobs_equal_lone_solves(dev, resident_min):
    from miccai2016_amd.serving import CutRequest, MaxFlowServer
    cases = [_case(40 + k, s) for k, s in enumerate(SHAPES)]
    iters = [200, 77, 131, 260]
    srv = MaxFlowServer(_params(), chunk=37, resident_min_voxels=resident_min)
    tickets = [srv.submit(CutRequest(c[0], c[1], c[2], c[3], c[4], EIGS, num_iter=it))
               for c, it in zip(cases, iters)]
    srv.run_until_idle()
    assert srv.peak_resident == len(cases)
    for c, it, t in zip(cases, iters, tickets):
        u = t.wait(0)
        want, res = _alone(dev, c, EIGS, it)
        assert t.iterations == res.iterations == it and t.stopped == 0
        assert np.array_equal(u, want)


@pytest.mark.parametrize("resident_min", [1 << 40, 0])
def test_chunked_jobs_stop_like_lone_solves(dev, resident_min):
    from miccai2016_amd.serving import CutRequest, MaxFlowServer
    cases = [_case(60 + k, s) for k, s in enumerate(SHAPES[:3])]
    srv = MaxFlowServer(_params(num_iter=3000, update=7, error_bound=1e-3), chunk=50,
                        resident_min_voxels=resident_min)
    tickets = [srv.submit(CutRequest(c[0], c[1], c[2], c[3], c[4], (1, 1, 1))) for c in cases]
    # a NaN frame stops its job at the first criterion, as it stops a lone solve
    bad = list(cases[0])
    bad[1] = bad[1].copy()
    bad[1][3, 4, 5, 1] = np.nan
    tickets.append(srv.submit(CutRequest(bad[0], bad[1], bad[2], bad[3], bad[4], (1, 1, 1))))
    srv.run_until_idle()
    for c, t in zip(cases + [bad], tickets):
        u = t.wait(0)
        want, res = _alone(dev, c, (1, 1, 1), 3000, 7, 1e-3)
        assert (t.iterations, t.stopped) == (res.iterations, res.stopped)
        if res.stopped == 2:
            assert np.array_equal(u, want)
    assert tickets[-1].stopped == 1 and tickets[-1].iterations == 1


def test_admission_respects_the_budget(dev):
    from miccai2016_amd.serving import CutRequest, MaxFlowServer, request_footprint_bytes
    cases = [_case(80 + k, s) for k, s in enumerate(SHAPES)]
    reqs = [CutRequest(c[0], c[1], c[2], c[3], c[4], EIGS, num_iter=90) for c in cases]
    need = [request_footprint_bytes(r, 90, 30) for r in reqs]
    budget = max(need) + min(need) - 1         # never all four at once
    srv = MaxFlowServer(_params(), budget_bytes=budget, chunk=20)
    z = np.zeros((60, 60, 60), np.float32)
    too_big = srv.submit(CutRequest(z, np.zeros(z.shape + (3,), np.float32), z, z, z))
    with pytest.raises(MemoryError):
        too_big.wait(0)
    tickets = [srv.submit(r) for r in reqs]
    srv.run_until_idle()
    assert srv.peak_reserved <= budget and srv.peak_resident < len(reqs)
    assert srv.reserved == 0
    for c, t in zip(cases, tickets):
        want, _ = _alone(dev, c, EIGS, 90)
        assert np.array_equal(t.wait(0), want)


def test_phantom_requests_equal_lone_solves(dev):
    from miccai2016_amd.evaluation import fast_dice_for
    from miccai2016_amd.processing.continuous_cut_anisotropic_gpu import CmfCutQAnisotropicGpu
    from miccai2016_amd.serving import MaxFlowServer, PhantomCutRequest
    reqs = [PhantomCutRequest((30, 26, 34), 4.0, 6.5, 30.0, 0.5, 0.1, 10.0, seed=7, num_iter=150),
            PhantomCutRequest((20, 40, 22), 3.0, 5.0, 24.0, 0.9, 0.03, 3.0, seed=8, num_iter=90)]
    srv = MaxFlowServer(_params(), chunk=40)
    tickets = [srv.submit(r) for r in reqs]
    srv.run_until_idle()
    for r, t in zip(reqs, tickets):
        u = t.wait(0)
        n = int(np.prod(r.shape))
        L, h, s = dev.lib, dev.handle, dev.stream
        img, truth, tan = dev.empty(n), dev.empty(n, dtype=dev.torch.uint8), dev.empty(3 * n)
        dev.check(L.mh_helix_phantom(h, _lib.dims3(r.shape), r.r_tube, r.r_helix, r.pitch,
                                     r.sigma, C.c_uint64(r.seed), dev.ptr(img), dev.ptr(truth),
                                     dev.ptr(tan), s))
        dev.check(L.mh_complete_frames(h, dev.ptr(tan), dev.ptr(tan), n, s))
        x = dev.download(img).reshape(r.shape)
        cs, ct = ph.terminal_capacities(x.astype(np.float64), r.sigma)
        alpha = ph.nonterminal_weight(ph.grad_mag_sq(x.astype(np.float64)), r.weighting, r.decay)
        c = CmfCutQAnisotropicGpu(_params(r.num_iter), dev)
        want = c.execute(x, dev.download(tan).reshape(r.shape + (3,)), r.eigs,
                         alpha.astype(np.float32), cs.astype(np.float32), ct.astype(np.float32),
                         verbose=False)
        # the served fields are FP32 device formulas, the lone solve's float64 host ones
        assert np.abs(u - want).max() <= 1e-3
        assert ph.dice_between(u, want) >= 0.999
        assert t.dice == pytest.approx(fast_dice_for(u, dev.download(truth).reshape(r.shape)),
                                       abs=1e-12)


def test_volume_requests_run_the_pipeline(dev):
    from miccai2016_amd.pipeline import PipelineConfig, prepare, segment
    from miccai2016_amd.serving import MaxFlowServer, VolumeRequest
    spec = ph.HelixSpec((32, 36, 36), 3.0, 9.0, 20.0, 0.5, seed=21)
    img, truth = ph.make_helix(spec)
    cfg = PipelineConfig()
    cfg.gvf_iters, cfg.cut_iters = 40, 120
    srv = MaxFlowServer(_params(num_iter=120), chunk=30, prepare_cfg=cfg)
    t = srv.submit(VolumeRequest(img, spec.sigma, 10 ** -1.625, 10 ** 1.5, truth=truth))
    srv.run_until_idle()
    u = t.wait(0)
    host = segment(img, prepare(img, spec.sigma, cfg, dev), 10 ** -1.625, 10 ** 1.5, cfg, dev)
    assert ph.dice_between(u, host) >= 0.99
    assert t.dice == pytest.approx(ph.fast_dice(u, truth), abs=1e-12)
    srv.stop()


def test_prepared_phantom_requests_equal_the_device_pipeline(dev):
    """Phantom requests with the full preparation, prepared by two workers
    while other cuts run: each labelling equals the device pipeline run alone
    (phantom, prepare_device, weight, initial u, plain solve), bit for bit."""
    from miccai2016_amd.pipeline import PipelineConfig, prepare_device
    from miccai2016_amd.processing.continuous_cut_anisotropic_gpu import CmfCutQAnisotropicGpu
    from miccai2016_amd.serving import MaxFlowServer, PhantomCutRequest
    cfg = PipelineConfig()
    cfg.gvf_iters = 40
    reqs = [PhantomCutRequest((30, 26, 34), 4.0, 6.5, 30.0, 0.5, 0.1, 10.0, seed=7, num_iter=150,
                              prepare=True),
            PhantomCutRequest((20, 40, 22), 3.0, 5.0, 24.0, 0.9, 0.03, 3.0, seed=8, num_iter=90),
            PhantomCutRequest((36, 20, 28), 3.5, 5.0, 22.0, 1.2, 0.3, 30.0, seed=9, num_iter=120,
                              prepare=True),
            PhantomCutRequest((24, 24, 24), 3.0, 6.0, 25.0, 0.3, 0.1, 3.0, seed=10, num_iter=60,
                              prepare=True)]
    srv = MaxFlowServer(_params(), chunk=40, prepare_cfg=cfg, prepare_workers=2)
    tickets = [srv.submit(r) for r in reqs]
    srv.run_until_idle()
    srv.stop()
    for r, t in zip(reqs, tickets):
        u = t.wait(0)
        if not r.prepare:
            continue
        n = int(np.prod(r.shape))
        L, h, s = dev.lib, dev.handle, dev.stream
        img, truth = dev.empty(n), dev.empty(n, dtype=dev.torch.uint8)
        dev.check(L.mh_helix_phantom(h, _lib.dims3(r.shape), r.r_tube, r.r_helix, r.pitch,
                                     r.sigma, C.c_uint64(r.seed), dev.ptr(img), dev.ptr(truth),
                                     None, s))
        prep = prepare_device(img, r.shape, r.sigma, cfg, dev)
        a, u0 = dev.empty(n), dev.empty(n)
        dev.check(L.mh_nonterminal_weight(h, dev.ptr(prep.gms), n, r.weighting, r.decay,
                                          dev.ptr(a), s))
        work = dev.empty(int(L.mh_initial_u_work_floats(n)))
        dev.check(L.mh_initial_u(h, dev.ptr(img), n, dev.ptr(u0), dev.ptr(work), s))
        c = CmfCutQAnisotropicGpu(_params(r.num_iter), dev)
        c.execute_device(u0, prep.quats, a, prep.cs, prep.ct, const_sqevals=np.sqrt(r.eigs),
                         shape=r.shape)
        assert np.array_equal(u, dev.download(u0).reshape(r.shape))
        assert t.iterations == r.num_iter and t.dice is not None


def test_poisson_load_on_a_background_thread(dev):
    from miccai2016_amd.serving import MaxFlowServer, poisson_load, synthetic_request
    srv = MaxFlowServer(_params(), chunk=25).start()
    tickets = poisson_load(srv, lambda rng: synthetic_request(rng, side=(16, 40), iters=(30, 80)),
                           qps=200.0, n_requests=12, seed=3)
    srv.stop(drain=True, timeout=120)
    assert all(t.done() and t.error is None for t in tickets)
    assert all(t.iterations == t.request.num_iter for t in tickets)
    st = srv.stats()
    assert st["completed"] == 12 and st["p50_s"] <= st["p99_s"]
    assert st["voxel_iterations"] == sum(int(np.prod(t.request.shape)) * t.iterations
                                         for t in tickets)


@pytest.mark.parametrize("workers", [1, 3])
def test_ragged_pipeline_equals_per_volume_pipeline(dev, workers):
    """segment_ragged packs every volume's device-prepared fields and solves
    them in one ragged launch sequence: each labelling equals the volume's
    own preparation plus a plain solve, bit for bit, whether the volumes are
    prepared in turn or by concurrent workers."""
    from miccai2016_amd.pipeline import PipelineConfig, prepare_device, segment_ragged
    from miccai2016_amd.processing.continuous_cut_anisotropic_gpu import CmfCutQAnisotropicGpu
    specs, w, d = ph.ragged_set(3, seed=4, side=(20, 44))
    imgs = [ph.make_helix(sp)[0] for sp in specs]
    cfg = PipelineConfig()
    cfg.gvf_iters, cfg.cut_iters = 40, 90
    rr = segment_ragged(imgs, [sp.sigma for sp in specs], w, d, cfg, dev, workers=workers)
    assert [r[0] for r in rr.results] == [90] * 3
    for v, (img, sp) in enumerate(zip(imgs, specs)):
        x = img.astype(np.float32)
        n = x.size
        x_t = dev.upload(x).reshape(-1)
        prep = prepare_device(x_t, x.shape, sp.sigma, cfg, dev)
        a, u = dev.empty(n), dev.empty(n)
        dev.check(dev.lib.mh_nonterminal_weight(dev.handle, dev.ptr(prep.gms), n, float(w[v]),
                                                float(d[v]), dev.ptr(a), dev.stream))
        work = dev.empty(int(dev.lib.mh_initial_u_work_floats(n)))
        dev.check(dev.lib.mh_initial_u(dev.handle, dev.ptr(x_t), n, dev.ptr(u), dev.ptr(work),
                                       dev.stream))
        c = CmfCutQAnisotropicGpu(_params(90), dev)
        c.execute_device(u, prep.quats, a, prep.cs, prep.ct, const_sqevals=np.sqrt(EIGS),
                         shape=x.shape)
        assert np.array_equal(dev.download(rr.volume(v)), dev.download(u).reshape(x.shape))


def test_resident_sweep_in_serving_equals_lone_solve(dev):
    """A cut large enough for the two-level sweep, served as a resident solve
    in chunks of an odd size, equals the lone solve bit for bit."""
    from miccai2016_amd.serving import MaxFlowServer, PhantomCutRequest
    req = PhantomCutRequest((64, 96, 128), 5.0, 16.0, 40.0, 0.8, 0.05, 30.0, seed=5, num_iter=123)
    try:
        dev.check(dev.lib.mh_cmf_set_sweep(dev.handle, 2, 0, 0))
        srv = MaxFlowServer(_params(), chunk=17, resident_min_voxels=0)
        dev2 = srv.dev
        dev2.check(dev2.lib.mh_cmf_set_sweep(dev2.handle, 2, 0, 0))
        t = srv.submit(req)
        srv.run_until_idle()
        u = t.wait(0)
    finally:
        dev.check(dev.lib.mh_cmf_set_sweep(dev.handle, 0, 0, 0))
    n = int(np.prod(req.shape))
    L, h, s = dev.lib, dev.handle, dev.stream
    img, tan = dev.empty(n), dev.empty(3 * n)
    dev.check(L.mh_helix_phantom(h, _lib.dims3(req.shape), req.r_tube, req.r_helix, req.pitch,
                                 req.sigma, C.c_uint64(req.seed), dev.ptr(img), None, dev.ptr(tan), s))
    dev.check(L.mh_complete_frames(h, dev.ptr(tan), dev.ptr(tan), n, s))
    cs, ct, a, u0 = dev.empty(n), dev.empty(n), dev.empty(n), dev.empty(n)
    dev.check(L.mh_terminal_capacities(h, dev.ptr(img), n, req.sigma, 1e-9, dev.ptr(cs), dev.ptr(ct), s))
    dev.check(L.mh_grad_mag_sq(h, dev.ptr(img), _lib.dims3(req.shape), dev.ptr(a), s))
    dev.check(L.mh_nonterminal_weight(h, dev.ptr(a), n, req.weighting, req.decay, dev.ptr(a), s))
    wk = dev.empty(int(L.mh_initial_u_work_floats(n)))
    dev.check(L.mh_initial_u(h, dev.ptr(img), n, dev.ptr(u0), dev.ptr(wk), s))
    from miccai2016_amd.processing.continuous_cut_anisotropic_gpu import CmfCutQAnisotropicGpu
    c = CmfCutQAnisotropicGpu(_params(123), dev)
    dev.check(dev.lib.mh_cmf_set_sweep(dev.handle, 1, 0, 0))
    try:
        c.execute_device(u0, tan, a, cs, ct, const_sqevals=np.sqrt(req.eigs), shape=req.shape)
    finally:
        dev.check(dev.lib.mh_cmf_set_sweep(dev.handle, 0, 0, 0))
    assert t.iterations == 123
    assert np.array_equal(u, dev.download(u0).reshape(req.shape))

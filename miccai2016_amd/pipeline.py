"""The segmentation pipeline of the reference experiment, stage by stage
(phantom_experiment.py:319-368): Frangi vesselness -> min-max normalise ->
direction adjustment -> GVF of the response-weighted directions -> frame
completion -> capacities -> ATV max-flow cut(s).
"""
from dataclasses import dataclass, field

import numpy as np

from . import phantom
from .device import default_context
from .helpers.misc import normalize
from .processing.adjust_directions import AdvancedAdjuster, TransformationCompleter
from .processing.continuous_cut_anisotropic_gpu import CmfCutQAnisotropicGpu
from .processing.gradient_vector_flow import GvfGpu, GvfParameters
from .processing.qfrangi_gpu import FeatureFilter, VesselnessCalculation


@dataclass
class PipelineConfig:
    scales: tuple = (1.0, 2.0, 3)        # Frangi scales (start, stop, number), mm
    alpha: float = 0.5
    beta: float = 0.5
    c: float = None                      # None: half the max eigenvalue norm per scale
    bright: bool = True
    ridge_thresholds: tuple = (0.01, 0.3)
    clean_ridge: bool = True
    gvf_mu: float = 0.1
    gvf_iters: int = 200
    cut_iters: int = 500
    cut_update: int = 30
    cut_error_bound: float = 0.0         # 0: fixed iteration count (only a NaN stops)
    steps: float = 0.16
    cc: float = 0.2
    a_atv: float = phantom.A_ATV
    voxel_size: tuple = field(default=(1.0, 1.0, 1.0))


@dataclass
class Prepared:
    """Everything the cuts of one noisy volume share."""
    vesselness: np.ndarray   # normalised to [0, 1], float64
    quats: np.ndarray        # completed frames (D0, D1, D2, 3), float32
    cs: np.ndarray           # foreground capacity, float32
    ct: np.ndarray           # background capacity, float32
    gms: np.ndarray          # |grad I|^2, float64


def prepare(img, noise_level, cfg=None, loader=None, verbose=False):
    cfg = PipelineConfig() if cfg is None else cfg
    loader = default_context() if loader is None else loader
    vp = VesselnessCalculation.Parameters()
    vp.alpha, vp.beta, vp.c = cfg.alpha, cfg.beta, cfg.c
    vp.bright_objects = cfg.bright
    vp.scales = cfg.scales
    ff = FeatureFilter(vp, loader)
    ves, dirs_raw, _ = ff.execute_for(img, roi=None, voxel_size=cfg.voxel_size)
    ves = normalize(ves, dst_min=0.0, dst_max=1.0, perc_low=None, perc_up=None, clip=True)
    adj = AdvancedAdjuster(threshold=cfg.ridge_thresholds, qeigs=dirs_raw, response=ves,
                           voxel_size=cfg.voxel_size, ridge_detection_only=False,
                           clean_ridge=cfg.clean_ridge, loader=loader)
    vdirs = adj.execute(verbose=verbose)
    scaled = np.rollaxis(np.rollaxis(vdirs, -1) * ves, 0, vdirs.ndim)
    gp = GvfParameters()
    gp.mu, gp.tol, gp.maxiter = cfg.gvf_mu, 0, cfg.gvf_iters
    gp.callback, gp.update, gp.normalize = None, max(cfg.gvf_iters // 10, 1), False
    smoothed = GvfGpu(scaled, gp, cfg.voxel_size, cl_program_loader=loader).execute()
    quats = TransformationCompleter(smoothed, loader).complete()
    cs, ct = phantom.terminal_capacities(img, noise_level)
    return Prepared(ves, quats.astype(np.float32), cs.astype(np.float32), ct.astype(np.float32),
                    phantom.grad_mag_sq(img))


@dataclass
class PreparedDevice:
    """Device tensors (float32, flat) shared by the cuts of one volume."""
    quats: object    # completed frames, AoS x3 (3N)
    cs: object
    ct: object
    gms: object      # |grad I|^2
    ridge_voxels: int


def prepare_device(img_t, shape, noise_level, cfg=None, loader=None, verbose=False,
                   timings=None):
    """prepare() on a device tensor (float32, N): every stage runs on the GPU
    and only the ridge steps of the adjuster visit the host. With a dict as
    `timings`, the stream is synchronised after each stage and the seconds
    of 'frangi', 'adjust', 'gvf' and 'fields' are stored in it."""
    import time
    from . import _lib
    cfg = PipelineConfig() if cfg is None else cfg
    dev = default_context() if loader is None else loader
    clock = [time.perf_counter()]

    def lap(name):
        if timings is not None:
            dev.synchronize()
            now = time.perf_counter()
            timings[name] = now - clock[0]
            clock[0] = now
    shape = tuple(shape)
    n = int(np.prod(shape))
    img_t = img_t.reshape(-1)
    vp = VesselnessCalculation.Parameters()
    vp.alpha, vp.beta, vp.c = cfg.alpha, cfg.beta, cfg.c
    vp.bright_objects, vp.scales = cfg.bright, cfg.scales
    ff = FeatureFilter(vp, dev)
    val, quat = ff.execute_tensor(img_t, shape, cfg.voxel_size)
    ves = dev.empty(n)
    work = dev.empty(int(dev.lib.mh_normalize_work_floats(n)))
    dev.check(dev.lib.mh_minmax_normalize(dev.handle, dev.ptr(val), n, dev.ptr(ves), dev.ptr(work),
                                          dev.stream), "mh_minmax_normalize")
    lap("frangi")
    adj = AdvancedAdjuster(threshold=cfg.ridge_thresholds, qeigs=None, response=None,
                           voxel_size=cfg.voxel_size, clean_ridge=cfg.clean_ridge, loader=dev)
    dirs = adj.execute_tensor(quat, ves, shape, verbose=verbose)
    v3 = dev.empty(3 * n)
    dev.check(dev.lib.mh_weighted_dirs(dev.handle, dev.ptr(dirs), dev.ptr(ves), n, dev.ptr(v3),
                                       dev.stream), "mh_weighted_dirs")
    lap("adjust")
    del ff, val, quat
    gp = GvfParameters()
    gp.mu, gp.tol, gp.maxiter = cfg.gvf_mu, 0, cfg.gvf_iters
    gp.callback, gp.update, gp.normalize = None, max(cfg.gvf_iters // 10, 1), False
    gvf = GvfGpu(v3, gp, cfg.voxel_size, cl_program_loader=dev, shape=shape)
    g = gvf.execute_device()
    lap("gvf")
    quats = dev.empty(3 * n)
    dev.check(dev.lib.mh_soa_to_aos3(dev.handle, dev.ptr(g), n, dev.ptr(quats), dev.stream),
              "mh_soa_to_aos3")
    del gvf, g, v3
    TransformationCompleter(None, dev).complete_device(quats)
    cs, ct, gms = dev.empty(n), dev.empty(n), dev.empty(n)
    dev.check(dev.lib.mh_terminal_capacities(dev.handle, dev.ptr(img_t), n, float(noise_level),
                                             1e-9, dev.ptr(cs), dev.ptr(ct), dev.stream),
              "mh_terminal_capacities")
    dev.check(dev.lib.mh_grad_mag_sq(dev.handle, dev.ptr(img_t), _lib.dims3(shape), dev.ptr(gms),
                                     dev.stream), "mh_grad_mag_sq")
    lap("fields")
    return PreparedDevice(quats, cs, ct, gms, adj.ridge_voxels)


def cut_parameters(cfg):
    p = CmfCutQAnisotropicGpu.CutParameters()
    p.num_iter, p.update, p.error_bound = cfg.cut_iters, cfg.cut_update, cfg.cut_error_bound
    p.steps, p.cc = cfg.steps, cfg.cc
    return p


def segment(img, prep, weighting, decay, cfg=None, loader=None, atv=True):
    """One cut with C = w exp(-|grad I|^2 / decay^2); returns u (float32)."""
    cfg = PipelineConfig() if cfg is None else cfg
    cutter = CmfCutQAnisotropicGpu(cut_parameters(cfg), loader)
    a = cfg.a_atv if atv else 1.0
    ntw = phantom.nonterminal_weight(prep.gms, weighting, decay)
    return cutter.execute(img, prep.quats, (1, a, a), ntw, prep.cs, prep.ct, labels=None,
                          scaling=cfg.voxel_size, verbose=False)


@dataclass
class RaggedResult:
    """Labellings of a ragged set: `u` packed (float32 device tensor), volume
    v at element offset offsets[v] with shape shapes[v]."""
    u: object
    shapes: list
    offsets: np.ndarray
    results: list          # per-volume (iterations, stop flag)
    prepare_s: float
    solve_s: float

    def volume(self, v):
        o, s = int(self.offsets[v]), self.shapes[v]
        return self.u[o:o + int(np.prod(s))].reshape(s)


_worker_contexts = {}


def _worker_context(device, k):
    """The k-th preparation worker's own library context and stream on
    `device` (kept for the process: a context per thread, as the library
    requires of concurrent callers)."""
    key = (device, k)
    if key not in _worker_contexts:
        from .device import DeviceContext
        ctx = DeviceContext(device)
        _worker_contexts[key] = (ctx, ctx.torch.cuda.Stream(ctx.device))
    return _worker_contexts[key]


def _prepare_cut_fields(img, nl, w, d, cfg, dev):
    """One volume's packed-solve inputs: (shape, u0, frames, alpha, cs, ct)."""
    if isinstance(img, tuple):
        img_t, shape = img[0].reshape(-1), tuple(img[1])
    else:
        shape = np.shape(img)
        img_t = dev.upload(np.asarray(img, dtype=np.float32)).reshape(-1)
    n = int(np.prod(shape))
    prep = prepare_device(img_t, shape, nl, cfg, dev)
    a, u0 = dev.empty(n), dev.empty(n)
    dev.check(dev.lib.mh_nonterminal_weight(dev.handle, dev.ptr(prep.gms), n, float(w), float(d),
                                            dev.ptr(a), dev.stream), "mh_nonterminal_weight")
    work = dev.empty(int(dev.lib.mh_initial_u_work_floats(n)))
    dev.check(dev.lib.mh_initial_u(dev.handle, dev.ptr(img_t), n, dev.ptr(u0), dev.ptr(work),
                                   dev.stream), "mh_initial_u")
    return shape, u0, prep.quats, a, prep.cs, prep.ct


def segment_ragged(imgs, noise_levels, weightings, decays, cfg=None, loader=None, atv=True,
                   workers=4):
    """The experiment pipeline over volumes of different shapes, each with its
    own noise level and (weighting, decay): every volume is prepared on the
    device (prepare_device), its fields are packed at element offsets, and
    one ragged max-flow solve segments all of them together. imgs: numpy
    volumes or (device tensor, shape) pairs. Returns a RaggedResult.

    With workers > 1, volumes are prepared by that many threads, each with its
    own library context and stream, so one volume's host ridge steps overlap
    other volumes' kernels; results do not depend on it."""
    import time
    from concurrent.futures import ThreadPoolExecutor
    cfg = PipelineConfig() if cfg is None else cfg
    dev = default_context() if loader is None else loader
    torch = dev.torch
    t0 = time.perf_counter()
    jobs = list(zip(imgs, noise_levels, weightings, decays))
    workers = max(1, min(int(workers), len(jobs)))
    if workers == 1:
        prepared = [_prepare_cut_fields(*j, cfg, dev) for j in jobs]
    else:
        dev.synchronize()   # device inputs are complete before other streams read them
        slots = list(range(workers))

        def run(job):
            k = slots.pop()
            try:
                ctx, stream = _worker_context(dev.index, k)
                with torch.cuda.stream(stream):
                    out = _prepare_cut_fields(*job, cfg, ctx)
                    stream.synchronize()
                return out
            finally:
                slots.append(k)
        with ThreadPoolExecutor(workers) as pool:
            prepared = list(pool.map(run, jobs))
    shapes = [p[0] for p in prepared]
    packed = {k: torch.cat([p[i + 1] for p in prepared])
              for i, k in enumerate(("u", "q", "a", "cs", "ct"))}
    sizes = [int(np.prod(s)) for s in shapes]
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    dev.synchronize()
    del prepared
    t1 = time.perf_counter()
    a_atv = cfg.a_atv if atv else 1.0
    cutter = CmfCutQAnisotropicGpu(cut_parameters(cfg), dev)
    cutter.execute_ragged_device(packed["u"], packed["q"], packed["a"], packed["cs"],
                                 packed["ct"], shapes, offsets, (1, a_atv, a_atv),
                                 cfg.voxel_size)
    dev.synchronize()
    t2 = time.perf_counter()
    res = [(r.iterations, r.stopped) for r in cutter.last_result]
    return RaggedResult(packed["u"], shapes, offsets, res, t1 - t0, t2 - t1)

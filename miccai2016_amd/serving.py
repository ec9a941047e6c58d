"""Continuous-batching server for ATV max-flow cuts on one GPU.

Requests wait in a FIFO queue (bounded by ``max_queue`` when given: a
request that finds it full is refused at once with ``OverloadError``, which
keeps latency bounded beyond the server's capacity). A request is admitted when its device
footprint fits the unreserved part of an HBM budget. An admitted cut joins
the resident set at the next chunk boundary; every chunk, each cut advances
``chunk`` iterations: cuts of at least ``resident_min_voxels`` voxels as
resident solves (``mh_cmf_resident_step``: per-iteration kernels or the
multi-level sweep, whichever the library uses for that shape), the smaller ones
together in one packed launch sequence (``mh_cmf_jobs_step``, any mix of
shapes). A finished cut leaves the set and releases its memory. Every cut
computes exactly the iterations of a lone
``CmfCutQAnisotropicGpu.execute_device`` call.

Three request kinds:

* ``CutRequest``: the fields of one cut (img, frames, alpha, cs, ct), numpy.
* ``VolumeRequest``: a noisy volume; the preparation of the reference
  experiment (Frangi, direction adjustment, GVF, frames, capacities,
  phantom_experiment.py:319-368) runs on the server's preparation thread, on
  its own stream and library context, before the cut joins the set.
* ``PhantomCutRequest``: a synthetic helix generated on the device with its
  analytic centre-line frames, for load tests.

No counterpart in the reference: its experiment runs one solve at a time
(phantom_experiment.py:350-368).
"""
import collections
import ctypes as C
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .device import DeviceContext
from .processing.continuous_cut_anisotropic_gpu import derived_constants, initial_u

# device bytes per voxel held by a resident cut besides its state: u, frames
# (3 floats), alpha, cs, ct
FIELD_BYTES_PER_VOXEL = 28
# bytes per voxel reserved while a VolumeRequest is being prepared: the
# Frangi working set (image, Hessian AoS x6, response, frames), the GVF
# buffers (field, result and swap SoA x3, response) and the cut's fields
PREPARE_BYTES_PER_VOXEL = 160
# cuts of at least this many voxels run as resident solves, which may use the
# multi-level sweep; smaller ones share the packed job launches
RESIDENT_MIN_VOXELS = 1 << 20


@dataclass
class CutRequest:
    img: np.ndarray
    quats: np.ndarray            # squeezed frames (D0, D1, D2, 3)
    alpha: np.ndarray            # nonterminal weight (scalar or volume)
    cs: np.ndarray
    ct: np.ndarray
    eigs: tuple = (1.0, 10 ** -1.5, 10 ** -1.5)
    num_iter: int = None         # None: the server's CutParameters.num_iter


@dataclass
class VolumeRequest:
    img: np.ndarray              # noisy volume
    noise_level: float
    weighting: float
    decay: float
    eigs: tuple = (1.0, 10 ** -1.5, 10 ** -1.5)
    num_iter: int = None
    truth: np.ndarray = None     # optional mask: fast Dice reported with the result
    cfg: object = None           # pipeline.PipelineConfig


@dataclass
class PhantomCutRequest:
    shape: tuple
    r_tube: float
    r_helix: float
    pitch: float
    sigma: float
    weighting: float
    decay: float
    seed: int = 0
    eigs: tuple = (1.0, 10 ** -1.5, 10 ** -1.5)
    num_iter: int = None
    # False: the phantom's own centre-line frames; True: the full preparation
    # (Frangi, direction adjustment, GVF, frame completion) on a worker
    prepare: bool = False
    cfg: object = None           # pipeline.PipelineConfig when prepared


def _prepares(req):
    return isinstance(req, VolumeRequest) or (isinstance(req, PhantomCutRequest) and req.prepare)


class OverloadError(RuntimeError):
    """A request refused at submission because max_queue requests were
    already waiting for admission (load shedding beyond capacity)."""


class Ticket:
    """Handle of one submitted request."""

    def __init__(self, request):
        self.request = request
        self.t_submit = time.perf_counter()
        self.t_admit = None
        self.t_done = None
        self.u = None             # float32 volume (numpy), or device tensor with return_device
        self.iterations = None
        self.stopped = None
        self.dice = None          # fast Dice against the request's truth, when it has one
        self.error = None
        self._done = threading.Event()
        self._released = False    # its reservation was returned (set once, under the lock)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit

    def done(self):
        return self._done.is_set()

    def wait(self, timeout=None):
        if not self._done.wait(timeout):
            raise TimeoutError("request not finished")
        if self.error is not None:
            raise self.error
        return self.u


def job_footprint_bytes(shape, num_iter, update):
    """Device bytes a resident cut of this shape holds: its fields plus the
    solver state (mh_cmf_job_state_bytes)."""
    n = int(np.prod(shape))
    P = _lib.CmfParams()
    P.num_iter, P.update, P.const_evals = int(num_iter), int(update), 1
    st = int(_lib.lib().mh_cmf_job_state_bytes(n, C.byref(P)))
    if st < 0:
        raise ValueError(f"volume of shape {tuple(shape)} is not supported")
    return FIELD_BYTES_PER_VOXEL * n + st + 5 * 256


def resident_footprint_bytes(shape):
    """Device bytes a cut advanced as a resident solve holds: its fields plus
    the solve's workspace (both buffer sets of the multi-level sweep)."""
    n = int(np.prod(shape))
    w = int(_lib.lib().mh_cmf_workspace_bytes(n))
    if w < 0:
        raise ValueError(f"volume of shape {tuple(shape)} is not supported")
    return FIELD_BYTES_PER_VOXEL * n + w + 5 * 256


def request_footprint_bytes(req, num_iter, update, resident_min=None):
    """Reservation of a request: a cut of at least `resident_min` voxels runs
    as a resident solve (mh_cmf_resident_*), a smaller one as a packed job
    (mh_cmf_jobs_*); a volume to prepare needs the preparation's working set
    first."""
    resident_min = RESIDENT_MIN_VOXELS if resident_min is None else resident_min
    shape = req.shape if isinstance(req, PhantomCutRequest) else np.shape(req.img)
    n = int(np.prod(shape))
    b = (resident_footprint_bytes(shape) if n >= resident_min
         else job_footprint_bytes(shape, num_iter, update))
    if _prepares(req):
        b = max(b, PREPARE_BYTES_PER_VOXEL * n)
    if isinstance(req, PhantomCutRequest):
        b += 5 * n            # the generated image and mask while the job is set up
    return b


class _Job:
    def __init__(self, ticket, shape, P, reserved):
        self.ticket = ticket
        self.shape = tuple(int(v) for v in shape)
        self.n = int(np.prod(self.shape))
        self.P = P
        self.reserved = reserved
        self.t = {}               # device tensors kept alive while resident
        self.truth = None
        self.event = None         # producer stream event to wait on before use
        self.resident = False     # advanced as a resident solve (else a packed job)


class MaxFlowServer:
    """See the module docstring.

    cut_params: CmfCutQAnisotropicGpu.CutParameters (cc, steps, update,
    error_bound, default num_iter). budget_bytes: HBM the server may reserve
    for requests (default: 90 % of the device's free memory at start).
    chunk: iterations per step of the resident set. prepare_workers: threads
    preparing VolumeRequests (each with its own library context and stream).
    max_queue: most requests waiting for admission (None: unbounded); a
    request submitted to a full queue fails at once with OverloadError.
    """

    def __init__(self, cut_params, budget_bytes=None, chunk=50, device=0, return_device=False,
                 max_jobs=1024, prepare_cfg=None, resident_min_voxels=RESIDENT_MIN_VOXELS,
                 prepare_workers=2, max_queue=None):
        import torch
        self.resident_min = int(resident_min_voxels)
        self.torch = torch
        self.params = cut_params
        self.chunk = int(chunk)
        self.return_device = return_device
        self.max_jobs = min(int(max_jobs), 1024)
        self.prepare_cfg = prepare_cfg
        self.dev = DeviceContext(device)
        self.stream = torch.cuda.Stream(device=self.dev.device)
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.dev.device)
            budget_bytes = int(0.9 * free)
        self.budget = int(budget_bytes)
        self.max_queue = None if max_queue is None else int(max_queue)
        self.rejected = 0         # refused at submission: the queue was full
        self.peak_queue = 0       # most requests waiting for admission at once
        self.reserved = 0
        self.peak_reserved = 0
        self.peak_resident = 0
        self._pending = collections.deque()
        self._ready = queue.Queue()
        self._prep_q = queue.Queue()
        self._lock = threading.Lock()
        self._wake = threading.Condition(self._lock)
        self._active = []
        self._jobs_arr = None
        self._stop = False
        self._thread = None
        self.prepare_workers = max(1, int(prepare_workers))
        self._prep_threads = []
        self._preparing = 0
        self._inflight = 0        # admitted, not yet finished or failed
        if int(cut_params.update) < 1:
            raise ValueError("CutParameters.update must be >= 1")
        self.completed = []
        self.voxel_iterations = 0
        self.t_start = None

    # ------------------------------------------------------------- public --
    def submit(self, request):
        num_iter = int(request.num_iter or self.params.num_iter)
        t = Ticket(request)
        try:
            need = request_footprint_bytes(request, num_iter, int(self.params.update),
                                           self.resident_min)
        except Exception as e:        # unsupported shape
            t.error = e
            t._done.set()
            return t
        if need > self.budget:
            t.error = MemoryError(f"request needs {need} bytes, the budget is {self.budget}")
            t._done.set()
            return t
        with self._wake:
            if self.max_queue is not None and len(self._pending) >= self.max_queue:
                self.rejected += 1
                t.error = OverloadError(f"{len(self._pending)} requests are waiting for "
                                        f"admission (max_queue {self.max_queue})")
                t.t_done = time.perf_counter()
                t._done.set()
                return t
            self._pending.append((t, need))
            self.peak_queue = max(self.peak_queue, len(self._pending))
            self._wake.notify()
        return t

    def start(self):
        """Run the serving loop on a background thread."""
        if self._thread is None:
            self._thread = threading.Thread(target=self._loop, name="maxflow-serve", daemon=True)
            self._thread.start()
        return self

    def stop(self, drain=True, timeout=None):
        """Stop the loop; with drain, after every submitted request is done."""
        if drain:
            self.wait_idle(timeout)
        with self._wake:
            self._stop = True
            self._wake.notify()
        if self._thread is not None:
            self._thread.join(timeout)
        for _ in self._prep_threads:
            self._prep_q.put(None)
        for th in self._prep_threads:
            th.join(timeout)

    def wait_idle(self, timeout=None):
        t0 = time.perf_counter()
        while True:
            with self._lock:
                idle = not self._pending and self._inflight == 0
            if idle:
                return
            if timeout is not None and time.perf_counter() - t0 > timeout:
                raise TimeoutError("server not idle")
            time.sleep(0.002)

    def run_until_idle(self):
        """Serve in the calling thread until nothing is queued or running."""
        self._loop(until_idle=True)

    def stats(self):
        lat = np.array([t.latency for t in self.completed if t.error is None])
        wall = (max(t.t_done for t in self.completed) - self.t_start) if self.completed else 0.0
        out = dict(completed=len(lat), voxel_iterations=self.voxel_iterations, wall_s=wall,
                   rejected=self.rejected, peak_queue=self.peak_queue)
        if len(lat):
            waits = np.array([t.t_admit - t.t_submit for t in self.completed
                              if t.error is None and t.t_admit is not None])
            out.update(p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                       mean_s=float(lat.mean()),
                       queue_wait_p99_s=float(np.percentile(waits, 99)) if len(waits) else None,
                       voxel_iters_per_s=self.voxel_iterations / wall if wall > 0 else None)
        return out

    # ------------------------------------------------------------ serving --
    def _loop(self, until_idle=False):
        torch = self.torch
        if self.t_start is None:
            self.t_start = time.perf_counter()
        with torch.cuda.device(self.dev.device), torch.cuda.stream(self.stream):
            while True:
                new = self._admit()
                while True:
                    try:
                        job = self._ready.get_nowait()
                    except queue.Empty:
                        break
                    if isinstance(job, Exception):
                        continue
                    self.stream.wait_event(job.event)
                    for v in job.t.values():
                        v.record_stream(self.stream)
                    new.append(job)
                try:
                    if new:
                        self._init(new)
                    if self._active:
                        self._step()
                        continue
                except Exception as e:     # a failed launch fails every job it touched
                    self._fail_all(new, e)
                    continue
                with self._wake:
                    if self._stop:
                        return
                    if until_idle and not self._pending and self._inflight == 0:
                        return
                    # nothing resident: sleep until a submission or a finished
                    # preparation (or a short timeout) changes that
                    if self._ready.empty():
                        self._wake.wait(0.005)

    def _admit(self):
        """FIFO admission against the budget; returns jobs ready to initialise."""
        new = []
        while True:
            with self._lock:
                if not self._pending:
                    break
                t, need = self._pending[0]
                if self.reserved + need > self.budget or \
                        len(self._active) + len(new) + self._preparing >= self.max_jobs:
                    break
                self._pending.popleft()
                self._inflight += 1
                self.reserved += need
                self.peak_reserved = max(self.peak_reserved, self.reserved)
            t.t_admit = time.perf_counter()
            try:
                if _prepares(t.request):
                    with self._lock:
                        self._preparing += 1
                    self._ensure_prep_thread()
                    self._prep_q.put((t, need))
                elif isinstance(t.request, PhantomCutRequest):
                    new.append(self._phantom_job(t, need))
                else:
                    new.append(self._cut_job(t, need))
            except Exception as e:
                self._fail(t, need, e)
        return new

    def _fail(self, t, need, e):
        """Fail one request and return its reservation, unless it has already
        finished or failed (each request is released exactly once)."""
        with self._lock:
            if t._released:
                return
            t._released = True
            self.reserved -= need
            self._inflight -= 1
        t.error = e
        t.t_done = time.perf_counter()
        t._done.set()
        self.completed.append(t)

    def _fail_all(self, new, e):
        """After a failed launch or finish: fail the jobs being initialised and
        every resident job that has not finished, and empty the resident set."""
        for job in list(new) + [j for j in self._active if j not in new]:
            self._fail(job.ticket, job.reserved, e)
        self._active, self._jobs_arr = [], None

    def _params_for(self, req, shape):
        """The solver constants of one job (unit voxel size)."""
        p = self.params
        rcc, rs, ss = derived_constants(p, (1, 1, 1))
        sq = np.sqrt(np.asarray(req.eigs, np.float64))
        if sq.shape != (3,):
            raise ValueError("served cuts take 3 constant eigenvalues")
        P = _lib.CmfParams()
        P.rcc = float(rcc)
        for k in range(3):
            P.rs[k], P.ss[k], P.sqev[k] = float(rs[k]), float(ss[k]), float(sq[k])
        P.const_evals = 1
        P.num_iter = int(req.num_iter or p.num_iter)
        P.update = int(p.update)
        P.error_bound = float(p.error_bound)
        return P

    def _cut_job(self, t, need):
        r = t.request
        dev = self.dev
        img = np.require(r.img, np.float32, ["C", "A"])
        shape = img.shape
        job = _Job(t, shape, self._params_for(r, shape), need)
        alpha = r.alpha if isinstance(r.alpha, np.ndarray) else np.full(shape, r.alpha, np.float32)
        job.t = dict(u=dev.upload(initial_u(img)), eigs=dev.upload(r.quats),
                     alpha=dev.upload(alpha), cs=dev.upload(r.cs), ct=dev.upload(r.ct))
        return job

    def _phantom_job(self, t, need):
        r = t.request
        dev = self.dev
        shape = tuple(int(v) for v in r.shape)
        n = int(np.prod(shape))
        job = _Job(t, shape, self._params_for(r, shape), need)
        img, tangent = dev.empty(n), dev.empty(3 * n)
        truth = dev.empty(n, dtype=self.torch.uint8)
        L, h, s = dev.lib, dev.handle, dev.stream
        dev.check(L.mh_helix_phantom(h, _lib.dims3(shape), float(r.r_tube), float(r.r_helix),
                                     float(r.pitch), float(r.sigma), C.c_uint64(int(r.seed)),
                                     dev.ptr(img), dev.ptr(truth), dev.ptr(tangent), s),
                  "mh_helix_phantom")
        dev.check(L.mh_complete_frames(h, dev.ptr(tangent), dev.ptr(tangent), n, s),
                  "mh_complete_frames")
        cs, ct, alpha, u = dev.empty(n), dev.empty(n), dev.empty(n), dev.empty(n)
        dev.check(L.mh_terminal_capacities(h, dev.ptr(img), n, float(r.sigma), 1e-9, dev.ptr(cs),
                                           dev.ptr(ct), s), "mh_terminal_capacities")
        # |grad I|^2 into alpha, then alpha = w exp(-|grad I|^2 / decay^2) in place
        dev.check(L.mh_grad_mag_sq(h, dev.ptr(img), _lib.dims3(shape), dev.ptr(alpha), s),
                  "mh_grad_mag_sq")
        dev.check(L.mh_nonterminal_weight(h, dev.ptr(alpha), n, float(r.weighting),
                                          float(r.decay), dev.ptr(alpha), s),
                  "mh_nonterminal_weight")
        work = dev.empty(int(L.mh_initial_u_work_floats(n)))
        dev.check(L.mh_initial_u(h, dev.ptr(img), n, dev.ptr(u), dev.ptr(work), s), "mh_initial_u")
        job.t = dict(u=u, eigs=tangent, alpha=alpha, cs=cs, ct=ct)
        job.truth = truth
        return job

    # --------------------------------------------------------- preparation --
    def _ensure_prep_thread(self):
        if not self._prep_threads:
            for k in range(self.prepare_workers):
                th = threading.Thread(target=self._prep_loop, name=f"maxflow-prepare-{k}",
                                      daemon=True)
                th.start()
                self._prep_threads.append(th)

    def _prep_loop(self):
        from .pipeline import prepare_device
        torch = self.torch
        with torch.cuda.device(self.dev.device):
            pdev = DeviceContext(self.dev.index)
            pstream = torch.cuda.Stream(device=self.dev.device)
            with torch.cuda.stream(pstream):
                while True:
                    item = self._prep_q.get()
                    if item is None:
                        return
                    t, need = item
                    try:
                        job = self._volume_job(t, need, prepare_device, pdev, pstream)
                        self._ready.put(job)
                    except Exception as e:
                        self._fail(t, need, e)
                    finally:
                        with self._wake:
                            self._preparing -= 1
                            self._wake.notify()

    def _volume_job(self, t, need, prepare_device, dev, pstream):
        r = t.request
        truth_t = None
        if isinstance(r, PhantomCutRequest):
            shape = tuple(int(v) for v in r.shape)
            n = int(np.prod(shape))
            img_t, truth_t = dev.empty(n), dev.empty(n, dtype=self.torch.uint8)
            dev.check(dev.lib.mh_helix_phantom(dev.handle, _lib.dims3(shape), float(r.r_tube),
                                               float(r.r_helix), float(r.pitch), float(r.sigma),
                                               C.c_uint64(int(r.seed)), dev.ptr(img_t),
                                               dev.ptr(truth_t), None, dev.stream),
                      "mh_helix_phantom")
            noise = r.sigma
        else:
            img = np.require(r.img, np.float32, ["C", "A"])
            shape = img.shape
            n = img.size
            img_t = dev.upload(img).reshape(-1)
            noise = r.noise_level
        prep = prepare_device(img_t, shape, noise, r.cfg or self.prepare_cfg, dev)
        L, h, s = dev.lib, dev.handle, dev.stream
        alpha, u = dev.empty(n), dev.empty(n)
        dev.check(L.mh_nonterminal_weight(h, dev.ptr(prep.gms), n, float(r.weighting),
                                          float(r.decay), dev.ptr(alpha), s),
                  "mh_nonterminal_weight")
        work = dev.empty(int(L.mh_initial_u_work_floats(n)))
        dev.check(L.mh_initial_u(h, dev.ptr(img_t), n, dev.ptr(u), dev.ptr(work), s),
                  "mh_initial_u")
        job = _Job(t, shape, self._params_for(r, shape), need)
        job.t = dict(u=u, eigs=prep.quats, alpha=alpha, cs=prep.cs, ct=prep.ct)
        if truth_t is not None:
            job.truth = truth_t
            job.t["truth"] = job.truth
        elif r.truth is not None:
            job.truth = dev.upload(np.asarray(r.truth) != 0, dtype=np.uint8).reshape(-1)
            job.t["truth"] = job.truth
        job.event = self.torch.cuda.Event()
        job.event.record(pstream)
        # the inputs above are read on the serving stream from now on
        dev.synchronize()
        return job

    # ------------------------------------------------------ resident set --
    def _init(self, new):
        dev = self.dev
        jobs = [j for j in new if j.n < self.resident_min]
        for job in new:
            if job.n >= self.resident_min:
                self._init_resident(job)
        arr = (_lib.CmfJob * len(jobs))()
        for k, job in enumerate(jobs):
            job.state = dev.bytes(int(dev.lib.mh_cmf_job_state_bytes(job.n, C.byref(job.P))))
            a = arr[k]
            for i in range(3):
                a.dims[i] = job.shape[i]
            a.u, a.eigs, a.alpha = (job.t["u"].data_ptr(), job.t["eigs"].data_ptr(),
                                    job.t["alpha"].data_ptr())
            a.cs, a.ct, a.state = job.t["cs"].data_ptr(), job.t["ct"].data_ptr(), job.state.data_ptr()
            a.params = job.P
        for i in range(0, len(jobs), 1024):
            part = (_lib.CmfJob * min(1024, len(jobs) - i)).from_buffer(arr, i * C.sizeof(_lib.CmfJob))
            dev.check(dev.lib.mh_cmf_jobs_init(dev.handle, part, len(part), dev.stream),
                      "mh_cmf_jobs_init")
        for k, job in enumerate(jobs):
            job.handle = _lib.CmfJob.from_buffer_copy(arr[k])
        self._active.extend(new)
        self.peak_resident = max(self.peak_resident, len(self._active))
        self._jobs_arr = None

    def _init_resident(self, job):
        dev = self.dev
        job.resident = True
        job.state = dev.bytes(int(dev.lib.mh_cmf_workspace_bytes(job.n)))
        r = _lib.CmfResident()
        for i in range(3):
            r.dims[i] = job.shape[i]
        r.u, r.eigs, r.alpha = (job.t["u"].data_ptr(), job.t["eigs"].data_ptr(),
                                job.t["alpha"].data_ptr())
        r.cs, r.ct, r.work = job.t["cs"].data_ptr(), job.t["ct"].data_ptr(), job.state.data_ptr()
        r.params = job.P
        dev.check(dev.lib.mh_cmf_resident_init(dev.handle, C.byref(r), dev.stream),
                  "mh_cmf_resident_init")
        job.handle = r

    def _step(self):
        dev = self.dev
        residents = [j for j in self._active if j.resident]
        jobs = [j for j in self._active if not j.resident]
        # large cuts: each advances by a chunk through its own launches
        for job in residents:
            dev.check(dev.lib.mh_cmf_resident_step(dev.handle, C.byref(job.handle), self.chunk, 0,
                                                   dev.stream), "mh_cmf_resident_step")
        done = []
        if jobs:
            if self._jobs_arr is None:
                self._jobs_arr = (_lib.CmfJob * len(jobs))(*[j.handle for j in jobs])
            arr = self._jobs_arr
            dev.check(dev.lib.mh_cmf_jobs_step(dev.handle, arr, len(arr), self.chunk, dev.stream),
                      "mh_cmf_jobs_step")   # blocks: every launch above has run
            for k, job in enumerate(jobs):
                job.handle = arr[k]
                if arr[k].stopped or arr[k].iter >= job.P.num_iter:
                    done.append(job)
        for job in residents:
            if job.handle.iter >= job.P.num_iter:
                done.append(job)
                continue
            st = C.c_int(0)
            dev.check(dev.lib.mh_cmf_resident_status(dev.handle, C.byref(job.handle), C.byref(st),
                                                     dev.stream), "mh_cmf_resident_status")
            if st.value:
                done.append(job)
        self._retire(done)

    def _retire(self, done):
        """Finish the done jobs one by one, each leaving the resident set once
        it has finished. If a finish raises, that job and the ones after it
        stay resident and the exception reaches the loop, which fails them
        (_fail_all); the finished ones are not failed again."""
        for job in done:
            self._finish(job)
            self._active = [j for j in self._active if j is not job]
            if not job.resident:
                self._jobs_arr = None

    def _finish(self, job):
        dev = self.dev
        t = job.ticket
        h = job.handle
        if job.resident:
            res = _lib.CmfResult()
            rc = dev.lib.mh_cmf_resident_finish(dev.handle, C.byref(h), C.byref(res), None,
                                                dev.stream)
            dev.check(rc, "mh_cmf_resident_finish")
            t.iterations, t.stopped = int(res.iterations), int(res.stopped)
        else:
            t.iterations, t.stopped = int(h.iter), int(h.stopped)
        if job.truth is not None:
            from .evaluation import DeviceDice
            t.dice = float(DeviceDice(job.truth, dev)(job.t["u"], 1)[0])
        if self.return_device:
            t.u = job.t["u"].reshape(job.shape)
        else:
            t.u = dev.download(job.t["u"]).reshape(job.shape)
        self.voxel_iterations += job.n * t.iterations
        job.t = {}
        job.state = None
        with self._lock:
            if t._released:
                raise RuntimeError("request finished after it was released")
            t._released = True
            self.reserved -= job.reserved
            self._inflight -= 1
        t.t_done = time.perf_counter()
        self.completed.append(t)
        t._done.set()


# ------------------------------------------------------------- load tests --

def synthetic_request(rng, side=(64, 320), aspect=2.0, iters=(300, 1000), prepare=False):
    """A PhantomCutRequest drawn like the serving workload: base side
    log-uniform in `side`, each axis scaled by a factor within sqrt(aspect)
    of 1 (so no two axes differ by more than `aspect`), tube radius U[3, 8],
    helix radius a quarter of the smallest side, pitch U[20, 80], noise
    U[0.1, 1.5], (w, decay) from the 25 x 17 grid, iterations U[lo, hi].
    prepare: frames from the full preparation instead of the centre line."""
    from . import phantom as ph
    base = float(np.exp(rng.uniform(np.log(side[0]), np.log(side[1]))))
    f = np.exp(rng.uniform(-0.5, 0.5, size=3) * np.log(aspect))
    f = f / np.sqrt(f.max() * f.min())          # centre the spread so max/min <= aspect
    dims = tuple(int(max(16, round(base * v))) for v in f)
    w = float(rng.choice(ph.GRID_WEIGHTINGS))
    d = float(rng.choice(ph.GRID_DECAYS))
    return PhantomCutRequest(dims, float(rng.uniform(3, 8)), 0.25 * min(dims),
                             float(rng.uniform(20, 80)), float(rng.uniform(0.1, 1.5)), w, d,
                             seed=int(rng.integers(1 << 62)),
                             num_iter=int(rng.integers(iters[0], iters[1] + 1)), prepare=prepare)


def poisson_load(server, make_request, qps, n_requests, seed=0):
    """Submit n_requests at Poisson arrival times of rate qps from the calling
    thread; returns the tickets in submission order."""
    rng = np.random.default_rng(seed)
    tickets = []
    t_next = time.perf_counter()
    for _ in range(int(n_requests)):
        t_next += rng.exponential(1.0 / qps)
        req = make_request(rng)
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(server.submit(req))
    return tickets

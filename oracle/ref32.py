"""ctypes bindings to the FP32 C oracle (oracle/miccai_oracle.c).

Test infrastructure only (see oracle/__init__.py). Each wrapper takes and
returns numpy arrays in the reference's layouts: volumes C-order (D0, D1, D2),
Hessians AoS x6 [H00, H11, H22, H01, H02, H12], eigen output AoS x6
[l1, l2, l3, phi, theta, psi], squeezed quaternions AoS x3.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class CutParams(C.Structure):
    """Mirror of orc_cut_params (derived constants of
    continuous_cut_anisotropic_gpu.py:104-129)."""
    _fields_ = [("rcc", C.c_double), ("rs", C.c_double * 3), ("ss", C.c_double * 3),
                ("sqev", C.c_double * 3), ("const_evals", C.c_int), ("supervised", C.c_int)]


def build():
    """Compile liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def use_native_build():
    """Build the oracle for this host's CPU (-march=native, oracle/Makefile's
    `native` target) and load it instead of the portable build; returns the
    compiler command line. Both builds compute the same values (no
    contraction); the native one is the CPU baseline's timed build."""
    global _lib, _LIB_PATH
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "native"], check=True)
    flags = subprocess.run(["make", "-s", "-C", _HERE, "native-flags"], check=True,
                           capture_output=True, text=True).stdout.strip()
    _LIB_PATH = os.path.join(_HERE, "liboracle_native.so")
    _lib = None
    lib()
    return flags


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.orc_convolve_axis.argtypes = [_f32p, _f32p, _i64p, C.c_int, _f32p, C.c_int]
        L.orc_gauss_taps.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, vp]
        L.orc_gauss_taps.restype = C.c_int
        L.orc_hessian.argtypes = [_f32p, _i64p, _f64p, C.c_double, C.c_int, _f32p]
        L.orc_eigs.argtypes = [_f32p, _f32p, C.c_int64, C.c_int]
        L.orc_half_max_norm.argtypes = [_f32p, C.c_int64]
        L.orc_half_max_norm.restype = C.c_float
        L.orc_update_vesselness.argtypes = [_f32p, _f32p, _f32p, C.c_int64, C.c_float, C.c_float,
                                            C.c_float, C.c_int]
        L.orc_vesselness.argtypes = [_f32p, _i64p, _f64p, C.c_int, C.c_float, C.c_float, C.c_float,
                                     C.c_int, _f32p, _f32p]
        L.orc_detect_ridges.argtypes = [_f32p, _f32p, _f32p, _i64p, C.c_double]
        L.orc_ridge_margin.argtypes = [_f64p, _f32p, _f32p, _i64p]
        L.orc_main_dirs.argtypes = [_f32p, _f32p, C.c_int64]
        L.orc_adjust_nonridge_signs.argtypes = [_f32p, _i64p, _u16p, _f32p, C.c_int64, _f32p]
        L.orc_complete_frames.argtypes = [_f32p, _f32p, C.c_int64]
        L.orc_qexpand.argtypes = [_f32p, _f32p, C.c_int64]
        L.orc_qsqueeze.argtypes = [_f32p, _f32p, C.c_int64]
        L.orc_qrot.argtypes = [_f32p, _f32p, _f32p, C.c_int64]
        L.orc_quat2mat.argtypes = [_f32p, _f32p, C.c_int64]
        L.orc_mat2quat.argtypes = [_f32p, _f32p, C.c_int64]
        L.orc_sorter_table.argtypes = [_f32p]
        L.orc_sum_abs_reduce.argtypes = [_f32p, C.c_int64]
        L.orc_sum_abs_reduce.restype = C.c_float
        L.orc_gvf.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i64p, C.c_double,
                              C.c_double, _f64p, C.c_int, C.c_int, vp]
        L.orc_gvf_laplacian.argtypes = [_f32p, _i64p, _f64p, _f32p]
        L.orc_grad_fwd.argtypes = [_f32p, _i64p, _f64p, _f32p, _f32p, _f32p]
        L.orc_divergence.argtypes = [_f32p, _f32p, _f32p, _i64p, _f64p, _f32p]
        L.orc_update_p.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i64p,
                                   C.POINTER(CutParams)]
        L.orc_update_u.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, vp,
                                   _i64p, C.POINTER(CutParams)]
        L.orc_update_tmp.argtypes = [_f32p, _f32p, C.c_int64, C.c_double]
        L.orc_cmf_solve.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, vp, _i64p,
                                    C.POINTER(CutParams), C.c_int, C.c_int, C.c_double, vp,
                                    C.POINTER(C.c_int)]
        L.orc_cmf_solve.restype = C.c_int
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _dims(shape):
    return np.asarray(shape, dtype=np.int64)


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))


# ---------------------------------------------------------------- Frangi ----

def gauss_taps(sigma, order, scale=1.0, n_sigmas=4.0, round_up=False):
    L = lib()
    R = L.orc_gauss_taps(float(sigma), int(order), float(scale), float(n_sigmas), int(round_up), None)
    out = np.empty(2 * R + 1, np.float32)
    L.orc_gauss_taps(float(sigma), int(order), float(scale), float(n_sigmas), int(round_up),
                     out.ctypes.data)
    return out


def convolve_axis(img, axis, taps):
    img = _f32(img)
    out = np.empty_like(img)
    taps = _f32(taps)
    lib().orc_convolve_axis(out, img, _dims(img.shape), int(axis), taps, (len(taps) - 1) // 2)
    return out


def hessian(img, sigma, n_sigmas=4.0, round_up=False):
    img = _f32(img)
    sig = np.asarray(np.broadcast_to(np.asarray(sigma, np.float64), (3,)), np.float64).copy()
    out = np.empty(img.shape + (6,), np.float32)
    if lib().orc_hessian(img, _dims(img.shape), sig, float(n_sigmas), int(round_up), out):
        raise MemoryError("oracle hessian allocation failed")
    return out


def eigs(h6, maxiter=50):
    h6 = _f32(h6)
    out = np.empty_like(h6)
    lib().orc_eigs(h6, out, h6.size // 6, int(maxiter))
    return out


def half_max_norm(e6):
    e6 = _f32(e6)
    return float(lib().orc_half_max_norm(e6, e6.size // 6))


def update_vesselness(val, quat, e6, alpha, beta, c, bright=True):
    """In place on val (float32[N]) and quat (float32[N,3])."""
    lib().orc_update_vesselness(val, quat, _f32(e6), val.size, alpha, beta, c, int(bright))


def vesselness(img, scales_vx, alpha=0.5, beta=0.5, c=None, bright=True):
    """Multi-scale Frangi response; scales_vx is (M, 3) sigma per axis in voxels."""
    img = _f32(img)
    sc = np.ascontiguousarray(np.asarray(scales_vx, np.float64).reshape(-1, 3))
    val = np.empty(img.shape, np.float32)
    quat = np.empty(img.shape + (3,), np.float32)
    cf = float("nan") if c is None else float(c)
    if lib().orc_vesselness(img, _dims(img.shape), sc, sc.shape[0], alpha, beta, cf, int(bright),
                            val, quat):
        raise MemoryError("oracle vesselness allocation failed")
    return val, quat


# ------------------------------------------------------------ directions ----

def detect_ridges(resp, qeig3, threshold):
    resp = _f32(resp)
    out = np.empty_like(resp)
    lib().orc_detect_ridges(out, resp, _f32(qeig3), _dims(resp.shape), float(threshold))
    return out


def ridge_margin(resp, qeig3):
    """resp - the ridge test's sample maximum (float64), per voxel: a ridge
    decision is a near-tie where this is within rounding of zero."""
    resp = _f32(resp)
    out = np.empty(resp.shape, np.float64)
    lib().orc_ridge_margin(out, resp, _f32(qeig3), _dims(resp.shape))
    return out


def main_dirs(qeig3):
    q = _f32(qeig3)
    out = np.empty_like(q)
    lib().orc_main_dirs(q, out, q.size // 3)
    return out


def adjust_nonridge_signs(dirs, coords, rdirs, scaling=(1.0, 1.0, 1.0)):
    d = _f32(dirs).copy()
    co = np.ascontiguousarray(coords, dtype=np.uint16)
    rd = _f32(rdirs)
    lib().orc_adjust_nonridge_signs(d, _dims(d.shape[:3]), co, rd, len(co),
                                    np.asarray(scaling, np.float32))
    return d


def complete_frames(dirs):
    d = _f32(dirs)
    out = np.empty_like(d)
    lib().orc_complete_frames(d, out, d.size // 3)
    return out


def qexpand(v):
    v = _f32(v)
    out = np.empty(v.shape[:-1] + (4,), np.float32)
    lib().orc_qexpand(v, out, v.size // 3)
    return out


def qsqueeze(q):
    q = _f32(q)
    out = np.empty(q.shape[:-1] + (3,), np.float32)
    lib().orc_qsqueeze(q, out, q.size // 4)
    return out


def qrot(q, v):
    q = _f32(q)
    v = _f32(v)
    out = np.empty_like(v)
    lib().orc_qrot(q, v, out, v.size // 3)
    return out


def quat2mat(q):
    q = _f32(q)
    out = np.empty(q.shape[:-1] + (3, 3), np.float32)
    lib().orc_quat2mat(q, out, q.size // 4)
    return out


def mat2quat(m):
    m = _f32(m)
    out = np.empty(m.shape[:-2] + (4,), np.float32)
    lib().orc_mat2quat(m, out, m.size // 9)
    return out


def sorter_table():
    out = np.empty((6, 4), np.float32)
    lib().orc_sorter_table(out)
    return out


def sum_abs_reduce(x):
    x = _f32(x).ravel()
    return float(lib().orc_sum_abs_reduce(x, x.size))


# ------------------------------------------------------------------- GVF ----

def gvf_time_step(voxel_size, mu, ndim=3):
    """gradient_vector_flow.py:58-65."""
    return float(np.minimum(0.99, (np.min(voxel_size) ** ndim) / (2.01 ** ndim * mu)))


def gvf(vfield, mu, maxiter, voxel_size=(1.0, 1.0, 1.0), response=None, update=None):
    """Returns (g AoS (D0,D1,D2,3), mags or None)."""
    vfield = np.asarray(vfield)
    v = [_f32(vfield[..., k]) for k in range(3)]
    shape = v[0].shape
    vs = np.multiply(voxel_size, np.ones(3))
    dt = gvf_time_step(vs, mu, 3)
    rss = np.ascontiguousarray(1.0 / vs ** 2, dtype=np.float64)
    if response is None:
        # r = |v|^2 (gvf.cl:246-253), float32 per operation
        r = ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]).astype(np.float32)
    else:
        r = _f32(response)
    g = [np.empty(shape, np.float32) for _ in range(3)]
    mags = None
    mp = None
    if update:
        mags = np.empty(int(math.ceil(maxiter / update)), np.float32)
        mp = mags.ctypes.data
    lib().orc_gvf(g[0], g[1], g[2], v[0], v[1], v[2], r, _dims(shape), float(mu), float(dt), rss,
                  int(maxiter), int(update or 0), mp)
    return np.stack(g, axis=-1), mags


def gvf_laplacian(g, rss=(1.0, 1.0, 1.0)):
    """The mirrored, RSS-weighted 7-point Laplacian orc_gvf applies (gvf.cl:77-92)."""
    g = _f32(g)
    out = np.empty_like(g)
    lib().orc_gvf_laplacian(g, _dims(g.shape), np.ascontiguousarray(rss, np.float64), out)
    return out


# -------------------------------------------------------------- max-flow ----

def grad_fwd(x, w=(1.0, 1.0, 1.0)):
    """The weighted forward difference of update_p (update_qp.cl:179-181), (3, D0, D1, D2)."""
    x = _f32(x)
    g = [np.empty_like(x) for _ in range(3)]
    lib().orc_grad_fwd(x, _dims(x.shape), np.ascontiguousarray(w, np.float64), *g)
    return np.stack(g)


def divergence(y, w=(1.0, 1.0, 1.0)):
    """The weighted divergence of update_u (update_qp.cl:270-272); y is (3, D0, D1, D2)."""
    y = [_f32(c) for c in y]
    out = np.empty_like(y[0])
    lib().orc_divergence(*y, _dims(y[0].shape), np.ascontiguousarray(w, np.float64), out)
    return out


def cut_params(scaling=(1, 1, 1), steps=0.11, cc=0.2, eigs=(1.0, 1.0, 1.0), supervised=False):
    """Derived constants as in continuous_cut_anisotropic_gpu.py:104-129."""
    scaling = np.asarray(scaling, np.float64)
    rs = 1.0 / scaling
    mins = np.min(scaling)
    ss = steps / scaling * mins ** 2
    P = CutParams()
    P.rcc = 1.0 / (cc * mins ** 2)
    for k in range(3):
        P.rs[k] = rs[k]
        P.ss[k] = ss[k]
    ev = np.asarray(eigs)
    if ev.ndim == 1:
        sq = np.sqrt(ev.astype(np.float64))
        for k in range(3):
            P.sqev[k] = sq[k]
        P.const_evals = 1
    else:
        P.const_evals = 0
    P.supervised = int(bool(supervised))
    return P


def cmf_initial_u(img):
    """u0 = img / (max(img) + tiny) in float32 (continuous_cut_anisotropic_gpu.py:282-283)."""
    img = np.asarray(img, np.float32)
    return np.require(np.divide(img, np.max(img) + np.finfo(np.float32).tiny), np.float32, ["C", "A"])


def cmf_solve(img, quats, eigs, alpha, cs, ct, labels=None, scaling=None, steps=0.11, cc=0.2,
              num_iter=1000, update=1, error_bound=1e-4):
    """Reference-structured solve. Returns (u, iterations_done, criteria)."""
    img = np.asarray(img, np.float32)
    scaling = (1, 1, 1) if scaling is None else scaling
    ev = np.asarray(eigs)
    P = cut_params(scaling, steps, cc, ev, labels is not None)
    if ev.ndim == 1:
        e = _f32(quats)
    else:
        e = np.empty(img.shape + (6,), np.float32)
        e[..., :3] = np.sqrt(ev)
        e[..., 3:] = quats
    if not isinstance(alpha, np.ndarray):
        alpha = np.multiply(alpha, np.ones(img.shape), dtype=np.float32)
    u = cmf_initial_u(img).copy()
    lab = None
    labp = None
    if labels is not None:
        lab = np.ascontiguousarray(labels, dtype=np.uint8)
        labp = lab.ctypes.data
    ncheck = int(math.ceil(num_iter / update)) if update else 0
    crit = np.empty(max(ncheck, 1), np.float32)
    nc = C.c_int(0)
    done = lib().orc_cmf_solve(u, e, _f32(alpha), _f32(cs), _f32(ct), labp, _dims(img.shape),
                               C.byref(P), int(num_iter), int(update), float(error_bound),
                               crit.ctypes.data, C.byref(nc))
    return u, done, crit[:nc.value]


def update_p(hp, tmp, eigs_buf, alpha, P):
    """One update_p on copies; hp is (3, D0, D1, D2)."""
    hp = [_f32(h).copy() for h in hp]
    lib().orc_update_p(hp[0], hp[1], hp[2], _f32(tmp), _f32(eigs_buf), _f32(alpha),
                       _dims(hp[0].shape), C.byref(P))
    return np.stack(hp)


def update_u(hp, tmp, cs, ct, pt, u, P, labels=None):
    """One update_u on copies; returns (tmp, pt, u)."""
    tmp = _f32(tmp).copy()
    pt = _f32(pt).copy()
    u = _f32(u).copy()
    hp = [_f32(h) for h in hp]
    lab = None if labels is None else np.ascontiguousarray(labels, np.uint8)
    lib().orc_update_u(hp[0], hp[1], hp[2], tmp, _f32(cs), _f32(ct), pt, u,
                       None if lab is None else lab.ctypes.data, _dims(tmp.shape), C.byref(P))
    return tmp, pt, u

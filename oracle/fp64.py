"""Float64 numpy restatements of the hot-path stages.

Test infrastructure only (see oracle/__init__.py). These state the FP32 error
of each stage (GPU FP32 vs exact-ish FP64) and restate the reference's
host-side numpy/scipy steps. Each function cites the reference file:line it
follows (paths relative to the reference's src/).
"""
import numpy as np
import scipy.ndimage as ndi
from scipy.sparse.csgraph import minimum_spanning_tree
from scipy.spatial.distance import pdist, squareform


# ------------------------------------------------------------ quaternions ----

def qexpand(v):
    """quaternion_math.h:79-90 in float64; v (..., 3) -> q (..., 4) scalar first."""
    v = np.asarray(v, np.float64)
    pp = (v[..., 0] + v[..., 2]) * 0.5
    pm = (v[..., 0] - v[..., 2]) * 0.5
    ct, st = np.cos(v[..., 1] * 0.5), np.sin(v[..., 1] * 0.5)
    return np.stack([np.cos(pp) * ct, np.cos(pm) * st, np.sin(pm) * st, np.sin(pp) * ct], -1)


def quat2mat(q):
    """quaternion_math.h:222-250; (..., 4) -> (..., 3, 3)."""
    q = np.asarray(q, np.float64)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    m = np.empty(q.shape[:-1] + (3, 3))
    m[..., 0, 0] = 1 - 2 * (y * y + z * z); m[..., 0, 1] = 2 * (x * y - w * z); m[..., 0, 2] = 2 * (x * z + w * y)
    m[..., 1, 0] = 2 * (x * y + w * z); m[..., 1, 1] = 1 - 2 * (x * x + z * z); m[..., 1, 2] = 2 * (y * z - w * x)
    m[..., 2, 0] = 2 * (x * z - w * y); m[..., 2, 1] = 2 * (y * z + w * x); m[..., 2, 2] = 1 - 2 * (x * x + y * y)
    return m


def squeezed_to_mat(v):
    return quat2mat(qexpand(v))


# ----------------------------------------------------------------- Frangi ----

def gauss_taps(sigma, order, scale=1.0, n_sigmas=4.0, round_up=False):
    """qfrangi_gpu.py:158-185 in float64 (not rounded to float32)."""
    R = int(np.ceil(n_sigmas * sigma)) if round_up else int(n_sigmas * sigma + 0.5)
    x = np.arange(-R, R + 1, dtype=np.float64)
    s2 = sigma * sigma
    g = np.exp(-.5 * x ** 2 / s2)
    g = g / np.sum(g)
    poly = [np.ones_like(x), -x / s2, (x ** 2 - s2) / s2 ** 2, x * (3 * s2 - x ** 2) / s2 ** 3][order]
    return scale * g * poly


def mirror_index(idx, n):
    """Mirror extension with period 2(n-1) (convolve_naive.cl:38-56)."""
    p = 2 * (n - 1)
    t = np.mod(idx, p)
    return np.where(t < n, t, p - t)


def convolve_axis(img, axis, taps):
    """True convolution along one axis with mirror boundary, float64."""
    img = np.asarray(img, np.float64)
    R = (len(taps) - 1) // 2
    n = img.shape[axis]
    out = np.zeros_like(img)
    base = np.arange(n)
    for m in range(-R, R + 1):
        src = mirror_index(base + m, n)
        out += taps[R - m] * np.take(img, src, axis=axis)
    return out


def hessian(img, sigma, n_sigmas=4.0, round_up=False):
    """qfrangi_gpu.py:208-271 in float64; returns AoS x6 [H00,H11,H22,H01,H02,H12]."""
    sig = np.broadcast_to(np.asarray(sigma, np.float64), (3,))
    out = np.empty(np.shape(img) + (6,))
    for i in range(3):
        for j in range(i + 1):
            order = [0, 0, 0]
            order[i] = 1
            order[j] += 1
            comp = i if i == j else 2 + i + j
            a = np.asarray(img, np.float64)
            for ax in range(3):
                t = gauss_taps(sig[ax], order[ax], sig[ax] ** order[ax], n_sigmas, round_up)
                a = convolve_axis(a, ax, t)
            out[..., comp] = a
    return out


def hessian_scipy(img, sigma, truncate=4.0):
    """Library routine pinning the Hessian: scipy gaussian_filter with
    mode='mirror', scaled by sigma^order (scale-space gamma = 1)."""
    sig = np.broadcast_to(np.asarray(sigma, np.float64), (3,))
    out = np.empty(np.shape(img) + (6,))
    for i in range(3):
        for j in range(i + 1):
            order = [0, 0, 0]
            order[i] = 1
            order[j] += 1
            comp = i if i == j else 2 + i + j
            f = ndi.gaussian_filter(np.asarray(img, np.float64), sig, order=order, mode="mirror",
                                    truncate=truncate)
            out[..., comp] = f * np.prod(sig ** np.asarray(order))
    return out


def h6_to_mat(h6):
    h6 = np.asarray(h6, np.float64)
    m = np.empty(h6.shape[:-1] + (3, 3))
    m[..., 0, 0], m[..., 1, 1], m[..., 2, 2] = h6[..., 0], h6[..., 1], h6[..., 2]
    m[..., 0, 1] = m[..., 1, 0] = h6[..., 3]
    m[..., 0, 2] = m[..., 2, 0] = h6[..., 4]
    m[..., 1, 2] = m[..., 2, 1] = h6[..., 5]
    return m


def eig_sorted(h6):
    """numpy.linalg.eigh, eigenpairs sorted by ascending |lambda|
    (qfrangi_gpu.py:300-302). Returns (lam (...,3), vecs (...,3,3) columns)."""
    w, v = np.linalg.eigh(h6_to_mat(h6))
    order = np.argsort(np.abs(w), axis=-1, kind="stable")
    w = np.take_along_axis(w, order, -1)
    v = np.take_along_axis(v, order[..., None, :], -1)
    return w, v


def vesselness_from_eigs(lam, alpha, beta, c, bright=True):
    """frangi.cl:46-92 in float64."""
    lam = np.asarray(lam, np.float64)
    l1 = np.abs(lam[..., 0])
    l2 = lam[..., 1] if bright else -lam[..., 1]
    l3 = lam[..., 2] if bright else -lam[..., 2]
    valid = (l2 < 0) & (l3 < 0)
    l2, l3 = np.abs(l2), np.abs(l3)
    tiny = np.finfo(np.float32).tiny
    with np.errstate(invalid="ignore", divide="ignore"):
        rB = l1 / (np.sqrt(l2 * l3) + tiny)
        rA = l2 / (l3 + tiny)
        S2 = l1 ** 2 + l2 ** 2 + l3 ** 2
        return valid * ((1 - np.exp(-0.5 * (rA / alpha) ** 2)) * np.exp(-0.5 * (rB / beta) ** 2) *
                        (1 - np.exp(-0.5 * S2 / c ** 2)))


def scales_voxels(scales, voxel_size):
    """qfrangi_gpu.py:582-592: linspace(*scales) (mm) / voxel_size -> (M, 3)."""
    return np.dot(np.linspace(*scales).reshape(-1, 1), np.divide(1., voxel_size).reshape(1, -1))


def normalize(data, dst_min=0.0, dst_max=1.0):
    """misc.py:63-114 without percentiles (float64 result)."""
    src_min, src_max = np.min(data), np.max(data)
    if src_max - src_min == 0:
        return np.ones_like(data, dtype=np.float64) * dst_min
    result = (1. / (src_max - src_min)) * (data.astype(np.float64) - src_min)
    if dst_max != 1 and dst_min == 0:
        result = result * dst_max
    elif dst_min != 0:
        result = result * (dst_max - dst_min) + dst_min
    return result


# ------------------------------------------------------------- directions ----

def hysteresis(cand, sure):
    """ridge_detection.py:80-131 restated with scipy.ndimage.label (26-conn):
    the union of candidate components that contain a seed."""
    cand = np.asarray(cand, bool)
    sure = np.asarray(sure, bool) & cand
    lab, _ = ndi.label(cand, structure=np.ones((3, 3, 3), int))
    keep = np.unique(lab[sure])
    keep = keep[keep > 0]
    return np.isin(lab, keep)


def mst_signs(coords, dirs, voxel_size=(1, 1, 1)):
    """adjust_directions.py:365-417: scipy MST over the dense distance matrix,
    signs propagated from point 0 so that dot(child, parent) >= 0."""
    coords = np.asarray(coords)
    n = len(coords)
    out = np.full((n, 3), np.nan, np.float32)
    if n == 0:
        return out
    mst = minimum_spanning_tree(squareform(pdist(np.multiply(coords, voxel_size), "euclidean")),
                                overwrite=True).tocsr()
    sym = (mst + mst.T).tocsr()
    out[0] = dirs[0]
    stack = [0]
    while stack:
        i = stack.pop()
        for j in sym.indices[sym.indptr[i]:sym.indptr[i + 1]]:
            if np.isnan(out[j, 0]):
                d = dirs[j]
                out[j] = d if np.dot(d, out[i]) >= 0 else -d
                stack.append(j)
    return out


def nearest_ridge_signs(dirs, coords, rdirs, scaling=(1, 1, 1)):
    """adjust_directions.cl:182-246 restated with a brute-force argmin (numpy
    argmin picks the first index, as the reference's strict <)."""
    shape = dirs.shape[:3]
    grid = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    c = np.asarray(coords, np.float64)
    s = np.asarray(scaling, np.float64)
    out = np.array(dirs, np.float32).reshape(-1, 3)
    for start in range(0, len(grid), 4096):
        g = grid[start:start + 4096].astype(np.float64)
        d2 = (((g[:, None, :] - c[None, :, :]) * s) ** 2).sum(-1)
        j = np.argmin(d2, axis=1)
        dot = (out[start:start + 4096].astype(np.float64) * rdirs[j]).sum(-1)
        out[start:start + 4096][dot < 0] *= -1
    return out.reshape(dirs.shape)


# -------------------------------------------------------------------- GVF ----

def gvf(vfield, mu, maxiter, voxel_size=(1, 1, 1), response=None):
    """gvf.cl:44-173 with gradient_vector_flow.py:58-65, 179-203 in float64."""
    v = np.moveaxis(np.asarray(vfield, np.float64), -1, 0)
    vs = np.multiply(voxel_size, np.ones(3))
    dt = min(0.99, np.min(vs) ** 3 / (2.01 ** 3 * mu))
    rss = 1.0 / vs ** 2
    r = (v ** 2).sum(0) if response is None else np.asarray(response, np.float64)
    g = np.zeros_like(v)
    for _ in range(maxiter):
        lap = np.zeros_like(g)
        for ax in range(3):
            n = g.shape[ax + 1]
            idx = np.arange(n)
            bef = np.where(idx > 0, idx - 1, 1)
            aft = np.where(idx < n - 1, idx + 1, n - 2)
            lap += (np.take(g, bef, axis=ax + 1) + np.take(g, aft, axis=ax + 1) - 2 * g) * rss[ax]
        delta = r[None] * (g - v) - mu * lap
        g = g - dt * delta
    return np.moveaxis(g, 0, -1)


# --------------------------------------------------------------- max-flow ----

def cmf(img, quats, eigs, alpha, cs, ct, num_iter, steps=0.11, cc=0.2, scaling=(1, 1, 1),
        labels=None):
    """update_qp.cl:150-316 with continuous_cut_anisotropic_gpu.py:198-343 in
    float64, no early stop. Returns u after num_iter iterations."""
    img = np.asarray(img, np.float64)
    sc = np.asarray(scaling, np.float64)
    rs = 1 / sc
    mins = sc.min()
    ss = steps / sc * mins ** 2
    rcc = 1 / (cc * mins ** 2)
    ev = np.asarray(eigs, np.float64)
    sq = np.sqrt(ev) * np.ones(img.shape + (3,))
    R = squeezed_to_mat(quats)                    # (...,3,3), H = R diag(sq)
    alpha = np.broadcast_to(np.asarray(alpha, np.float64), img.shape)
    cs = np.asarray(cs, np.float64)
    ct = np.asarray(ct, np.float64)
    u = img / (img.max() + np.finfo(np.float32).tiny)
    hp = np.zeros((3,) + img.shape)
    pt = np.minimum(cs, ct)
    tmp = -u * rcc
    if labels is not None:
        uf = (np.asarray(labels) & 1).astype(np.float64)
        ub = (~(np.asarray(labels) & 2).astype(bool)).astype(np.float64)
    else:
        uf, ub = 0.0, 1.0
    for _ in range(num_iter):
        g = np.zeros((3,) + img.shape)
        g[0, :-1] = (tmp[1:] - tmp[:-1]) * ss[0]
        g[1, :, :-1] = (tmp[:, 1:] - tmp[:, :-1]) * ss[1]
        g[2, :, :, :-1] = (tmp[:, :, 1:] - tmp[:, :, :-1]) * ss[2]
        gt = np.einsum("...ji,j...->...i", R, g) * sq          # H^T g
        pin = np.einsum("...ji,j...->...i", R, hp) / sq        # H^-1 hp
        p = pin + gt
        nrm = np.sqrt((p ** 2).sum(-1))
        with np.errstate(invalid="ignore", divide="ignore"):
            scale = np.where(nrm > alpha, alpha / nrm, 1.0)
        p = p * scale[..., None]
        hp = np.moveaxis(np.einsum("...ij,...j->...i", R, p * sq), -1, 0)
        d0 = np.zeros(img.shape); d1 = np.zeros(img.shape); d2 = np.zeros(img.shape)
        d0[:-1] += hp[0, :-1]; d0[1:] -= hp[0, :-1]
        d1[:, :-1] += hp[1, :, :-1]; d1[:, 1:] -= hp[1, :, :-1]
        d2[:, :, :-1] += hp[2, :, :, :-1]; d2[:, :, 1:] -= hp[2, :, :, :-1]
        div = d0 * rs[0] + d1 * rs[1] + d2 * rs[2]
        ps =np.minimum(div + pt - rcc * (u - ub), cs)
        pt = np.minimum(-div + ps + rcc * (u - uf), ct)
        err = div - ps + pt
        u = u - err / rcc
        tmp = err - u * rcc
    return u


# ----------------------------------------------------- capacities / scores ----

def normpdf(x, mean, var):
    """phantom_experiment.py:29-35."""
    return np.exp((x - mean) ** 2 * (1 / (-2 * var))) * (1 / np.sqrt(2 * np.pi * var))


def capacities(image, noise_level, eps=1e-9):
    """phantom_experiment.py:280-301: returns (bgw, fgw) = (ct, cs)."""
    var = noise_level ** 2
    image = np.asarray(image, np.float64)
    bgp = normpdf(image, 0, var)
    fgp = normpdf(image, 1, var)
    p_hat = np.maximum(normpdf(0, 0, var), normpdf(1, 1, var))
    r = np.log((fgp + eps) / (bgp + eps))
    rq = 1 / np.log((p_hat + eps) / eps)
    return rq * np.maximum(-r, 0), rq * np.maximum(r, 0)


def grad_mag_sq(image):
    """phantom_experiment.py:345: |np.gradient(image)|^2."""
    return np.linalg.norm(np.gradient(np.asarray(image, np.float64)), axis=0) ** 2


def nonterminal_weights(gms, w, decay):
    """phantom_experiment.py:359."""
    return w * np.exp(((-1.) / (decay ** 2)) * gms)


def fast_dice(seg, truth):
    """phantom_experiment.py:38-49."""
    b = seg >= 0.5
    return 2 * np.sum(np.multiply(b, truth)) / (np.sum(b) + np.sum(truth))

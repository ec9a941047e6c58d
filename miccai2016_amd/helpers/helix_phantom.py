"""The reference experiment's "toric helix" phantom (reference:
helpers/helix_phantom.py:18-122), voxelised analytically.

The reference triangulates the tube's surface with tvtk and rasterises the
closed mesh (PolyDataToImageStencil). Here the same surface is tested point
by point: the tube's cross-section at curve parameter t is the disc of radius
r_t(t) around (r_h(t), z(t)) in the half-plane at azimuth 2 pi t (helix_phantom
.py:58-71), so a voxel centre at cylindrical (rho, phi, z) is inside when, for
some t = phi / (2 pi) + k in [0, n_windings], (rho - r_h(t))^2 + (z - z(t))^2
<= r_t(t)^2 (the flat end caps of the closed mesh are the cross-sections at
t = 0 and t = n_windings). The volume's extent and origin follow the
reference's: the bounds of the sampled surface points, plus the margins,
divided into voxels whose centres start half a voxel inside. Voxels on the
boundary can differ from the faceted mesh's rasterisation (100 points per
circle, 100 cross-sections per winding).
"""
import numpy as np


def _curve(r_helix, r_tube, slope, n_windings, factor_helix, factor_tube, factor_slope):
    """(r_h, r_t, z) as functions of the curve parameter t (helix_phantom.py:60-66)."""
    def r_h(t):
        return r_helix * ((factor_helix - 1) * t / n_windings + 1)

    def r_t(t):
        return r_tube * ((factor_tube - 1) * t / n_windings + 1)

    def z(t):
        # the ascent with slope and helix radius both changing linearly in t
        return 2 * np.pi * r_helix * slope * t * (
            1 + t * (factor_helix + factor_slope - 2) / (2 * n_windings)
            + t ** 2 * (factor_helix - 1) * (factor_slope - 1) / (3 * n_windings ** 2))
    return r_h, r_t, z


def _margins6(margins):
    m = np.asarray(margins, dtype=np.float64)
    if m.ndim == 0:
        return np.full(6, float(m))
    if m.size == 3:
        return np.concatenate([m, m])
    if m.size == 6:
        return m.astype(np.float64)
    raise ValueError("margins: a scalar, 3 values or 6 values expected")


def create_toric_helix(r_helix, r_tube, slope, n_windings, factor_helix=1.0, factor_tube=1.0,
                       factor_slope=1.0, points_per_circle=100, points_per_winding=100,
                       write_output=False, output_path=None, voxel_size=1.0, margins=0.0,
                       show=False):
    """A tube of radius r_tube (growing linearly by factor_tube over the
    course) winding n_windings times at radius r_helix (factor_helix) with
    the given slope (factor_slope) around axis 2, sampled at voxel_size
    (scalar or 3 values) with margins (scalar, 3 or 6 values: axis starts,
    then ends). Returns a uint8 volume indexed (axis 0, axis 1, axis 2) with
    100 inside the tube and 0 outside. points_per_circle and
    points_per_winding set the surface samples whose bounds define the
    volume's extent; write_output is accepted and ignored (a no-op in the
    reference too); show needs a display and is not supported."""
    if show:
        raise NotImplementedError("create_toric_helix: show=True needs a display (mayavi)")
    r_h, r_t, z = _curve(r_helix, r_tube, slope, n_windings, factor_helix, factor_tube,
                         factor_slope)
    # the surface samples of the reference's mesh (helix_phantom.py:56-71)
    t, theta = np.mgrid[0:n_windings:(n_windings * points_per_winding * 1j),
                        0:(2 * np.pi):((points_per_circle + 1) * 1j)]
    rho_s = r_h(t) + r_t(t) * np.cos(theta)
    surf = np.stack([rho_s * np.cos(2 * np.pi * t), rho_s * np.sin(2 * np.pi * t),
                     r_t(t) * np.sin(theta) + z(t)], -1).reshape(-1, 3)
    lo, hi = surf.min(0), surf.max(0)
    vs = np.multiply(voxel_size, np.ones(3))
    m = _margins6(margins)
    n_vox = np.ceil((hi - lo + m[:3] + m[3:]) / vs).astype(np.int64)
    origin = lo - m[:3] + 0.5 * vs
    c0 = origin[0] + vs[0] * np.arange(n_vox[0])
    c1 = origin[1] + vs[1] * np.arange(n_vox[1])
    c2 = origin[2] + vs[2] * np.arange(n_vox[2])
    x0, x1 = np.meshgrid(c0, c1, indexing="ij")
    rho = np.hypot(x0, x1)[..., None]                       # (n0, n1, 1)
    t0 = np.arctan2(x1, x0)[..., None] / (2 * np.pi)       # in (-1/2, 1/2]
    inside = np.zeros(tuple(n_vox), bool)
    zz = c2[None, None, :]
    for k in range(-1, int(np.ceil(n_windings)) + 2):
        tk = t0 + k
        ok = (tk >= 0) & (tk <= n_windings)
        if not ok.any():
            continue
        tc = np.clip(tk, 0, n_windings)
        d2 = (rho - r_h(tc)) ** 2 + (zz - z(tc)) ** 2
        inside |= ok & (d2 <= r_t(tc) ** 2)
    return np.where(inside, np.uint8(100), np.uint8(0))

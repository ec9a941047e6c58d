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

"""Differentiable (expected) film model on the GPU: forward and vector-Jacobian product.

No counterpart in the reference, whose path is not differentiable (integer latent
counts, a Boolean development mask).  The expected model replaces both by their
expectations (the C header states the model in full):

* expected dose per crystal: the mean latent count 5 * E * w * pa * sens of every hit
  (transport.rs:117-134) summed over packets that replay fg_expose's photon streams;
* development probability P[Poisson(dose) >= threshold] (development.rs:69-84);
* expected transmittance: the product over the crystals a scan ray crosses of
  1 - prob * (1 - exp(-sigma * chord)), averaged over the pixel's rays and mapped by the
  scan mode (transmittance.rs:122-152, image_pipeline.rs:1006-1011).

``ExpectedFilm`` is the VJP object (forward, vjp); ``ExpectedFilmFunction`` adapts it to
``torch.autograd.Function`` for callers that optimise with torch.  The backward pass
regenerates every packet from its Philox counters instead of storing paths.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, replace

import numpy as np

from . import _lib as L
from ._lib import check, ptr
from .api import (DevelopmentConfig, ExposureConfig, GelatinOptics, GpuEngine, ScanConfig,
                  VirtualEmulsionLayer, _engine)

__all__ = ["ExpectedModel", "ParamGrads", "ExpectedFilm", "ExpectedFilmFunction", "mse_loss"]


@dataclass
class ExpectedModel:
    exposure: ExposureConfig
    optics: GelatinOptics
    development: DevelopmentConfig
    scan: ScanConfig
    expose_seed: int = 1
    scan_seed: int = 2
    weight_cutoff: float = 1e-6

    def c(self) -> L.ExpectedModelC:
        return L.ExpectedModelC(self.exposure.c(), self.optics.c(), self.development.c(),
                                self.scan.c(), self.expose_seed & (2**64 - 1),
                                self.scan_seed & (2**64 - 1), self.weight_cutoff)

    def with_params(self, crystal_absorption_coeff=None, anisotropy_g=None,
                    antihalation_reflectance=None) -> "ExpectedModel":
        """A copy with the differentiable parameters replaced (None keeps a value)."""
        ex, op = self.exposure, self.optics
        if crystal_absorption_coeff is not None:
            ex = replace(ex, crystal_absorption_coeff=float(crystal_absorption_coeff))
        if anisotropy_g is not None:
            op = replace(op, anisotropy_g=float(anisotropy_g))
        if antihalation_reflectance is not None:
            op = replace(op, antihalation_reflectance=float(antihalation_reflectance))
        return replace(self, exposure=ex, optics=op)

    def params(self) -> np.ndarray:
        return np.array([self.exposure.crystal_absorption_coeff, self.optics.anisotropy_g,
                         self.optics.antihalation_reflectance], np.float64)


@dataclass
class ParamGrads:
    crystal_absorption_coeff: float = 0.0
    anisotropy_g: float = 0.0
    antihalation_reflectance: float = 0.0

    def array(self) -> np.ndarray:
        return np.array([self.crystal_absorption_coeff, self.anisotropy_g,
                         self.antihalation_reflectance], np.float64)


def _stats(st: L.AdjointStatsC) -> dict:
    return {k: getattr(st, k) for k, _ in L.AdjointStatsC._fields_ if k != "pad_"}


class ExpectedFilm:
    """Expected model of one emulsion tile: ``forward(E)`` and ``vjp(E, dL/dout)``."""

    def __init__(self, emulsion: VirtualEmulsionLayer, model: ExpectedModel,
                 engine: GpuEngine | None = None):
        self.emulsion = emulsion
        self.model = model
        self.engine = _engine(engine)
        self.last_stats: dict = {}

    def _map(self, exposure_map):
        em = np.ascontiguousarray(exposure_map, np.float32)
        if em.ndim != 2:
            raise ValueError("exposure map must be 2-D (rows, cols)")
        return em

    def forward(self, exposure_map, want_prob=False):
        """-> (out (h, w) f32, dose per crystal f32[, prob per crystal f32])."""
        em = self._map(exposure_map)
        th, tw = em.shape
        n = len(self.emulsion)
        s = self.model.scan
        out = np.zeros((s.height_px, s.width_px), np.float32)
        dose = np.zeros(n, np.float32)
        prob = np.zeros(n, np.float32) if want_prob else None
        cm = self.model.c()
        st = L.AdjointStatsC()
        check(L.lib().fg_expected_forward(self.engine._h, self.emulsion._h, C.byref(cm),
                                          ptr(em, L._f32p), tw, th, ptr(dose, L._f32p),
                                          ptr(prob, L._f32p), ptr(out, L._f32p), C.byref(st)))
        self.last_stats = _stats(st)
        return (out, dose, prob) if want_prob else (out, dose)

    def vjp(self, exposure_map, grad_out, dose=None, want_grad_dose=False):
        """-> (dL/dE (th, tw) f32, ParamGrads[, dL/ddose per crystal])."""
        em = self._map(exposure_map)
        th, tw = em.shape
        s = self.model.scan
        g = np.ascontiguousarray(grad_out, np.float32)
        if g.shape != (s.height_px, s.width_px):
            raise ValueError(f"grad_out shape {g.shape} != scan grid {(s.height_px, s.width_px)}")
        d = None if dose is None else np.ascontiguousarray(dose, np.float32)
        if d is not None and d.shape != (len(self.emulsion),):
            raise ValueError("dose must hold one value per crystal")
        gE = np.zeros((th, tw), np.float32)
        gd = np.zeros(len(self.emulsion), np.float32) if want_grad_dose else None
        pg = L.ParamGradsC()
        cm = self.model.c()
        st = L.AdjointStatsC()
        check(L.lib().fg_vjp(self.engine._h, self.emulsion._h, C.byref(cm), ptr(em, L._f32p), tw, th,
                             ptr(d, L._f32p), ptr(g, L._f32p), ptr(gE, L._f32p), C.byref(pg),
                             ptr(gd, L._f32p), C.byref(st)))
        self.last_stats = _stats(st)
        grads = ParamGrads(pg.crystal_absorption_coeff, pg.anisotropy_g, pg.antihalation_reflectance)
        return (gE, grads, gd) if want_grad_dose else (gE, grads)


def mse_loss(out, target):
    """L = mean((out - target)^2) and dL/dout."""
    diff = out.astype(np.float64) - target.astype(np.float64)
    return float(np.mean(diff * diff)), (2.0 * diff / diff.size).astype(np.float32)


def _torch_function():
    import torch

    class _Fn(torch.autograd.Function):
        """torch.autograd.Function over ExpectedFilm: inputs (exposure map, params
        [mu_abs, g, R]); the backward pass is fg_vjp (path replay)."""

        @staticmethod
        def forward(ctx, exposure, params, film: ExpectedFilm):
            p = params.detach().cpu().double().numpy()
            f = ExpectedFilm(film.emulsion, film.model.with_params(*p), film.engine)
            E = exposure.detach().cpu().float().numpy()
            out, dose = f.forward(E)
            ctx.film, ctx.E, ctx.dose = f, E, dose
            ctx.dev, ctx.pdev = exposure.device, params.device
            return torch.from_numpy(out).to(exposure.device)

        @staticmethod
        def backward(ctx, grad_out):
            gE, pg = ctx.film.vjp(ctx.E, grad_out.detach().cpu().float().numpy(), dose=ctx.dose)
            return (torch.from_numpy(gE).to(ctx.dev),
                    torch.from_numpy(pg.array()).to(ctx.pdev), None)

    return _Fn


class _LazyFunction:
    _fn = None

    def apply(self, exposure, params, film):
        if _LazyFunction._fn is None:
            _LazyFunction._fn = _torch_function()
        return _LazyFunction._fn.apply(exposure, params, film)


ExpectedFilmFunction = _LazyFunction()

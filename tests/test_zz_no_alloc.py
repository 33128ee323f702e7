"""fg_expose_noalloc (the allocation-free transport entry point) against fg_expose.

The caller's device memory (torch tensors: plumbing only) holds the emulsion's threaded
nodes, its records, the exposure map and the latent counts; the entry point must give
exactly the latent counts and statistics fg_expose gives for the same emulsion, seed and
configuration, on two calls in a row (the second with a smaller tile, so a scratch
block left over from the first call would show), and on the caller's own stream
without statistics (enqueue only).
"""
import ctypes as C

import numpy as np
import pytest
import torch

import fgrain_amd as F
from fgrain_amd import _lib as L
from helpers import trix_config

pytestmark = pytest.mark.gpu


def _ramp(w, h):
    return np.tile(np.linspace(0.05, 3.0, w, dtype=np.float32), (h, 1))


def _reference(engine, em, emap, x, seed):
    st = F.expose_emulsion(em, emap, x, seed, engine)
    return em.absorbed_photons().astype(np.int64), st


def _noalloc(engine, em, emap, x, seed, stream=None, with_stats=True):
    nodes, recs = em.device_layout()
    d_nodes = torch.from_numpy(nodes.view(np.int32).copy()).cuda()
    d_recs = torch.from_numpy(recs.view(np.int32).copy()).cuda()
    d_counts = torch.zeros(len(em), dtype=torch.int32, device="cuda")
    d_emap = torch.from_numpy(np.ascontiguousarray(emap)).cuda()
    torch.cuda.synchronize()
    tile = L.DeviceTileC(d_nodes.data_ptr(), d_recs.data_ptr(), d_counts.data_ptr(), d_emap.data_ptr(),
                         nodes.shape[0], len(em), emap.shape[1], emap.shape[0])
    st = L.ExposeStatsC()
    cfg = x.c()
    L.check(L.lib().fg_expose_noalloc(engine._h, em._h, C.byref(tile), C.byref(cfg), seed, stream,
                                      C.byref(st) if with_stats else None))
    torch.cuda.synchronize()
    leaf = d_counts.cpu().numpy().view(np.uint32).astype(np.int64)
    counts = np.zeros(len(em), np.int64)
    counts[em.bvh()["indices"]] = leaf          # leaf slot k holds crystal indices[k]
    return counts, {k: getattr(st, k) for k, _ in L.ExposeStatsC._fields_ if k != "pad_"}


STAT_KEYS = ("launched", "bounces", "nodes_visited", "prims_tested", "abs_events", "abs_total", "atomics",
             "packets_run")


def test_noalloc_transport_equals_fg_expose_on_two_calls(engine):
    x = F.ExposureConfig(0.0, 0.12, 0.9, 48, 16)
    sizes = [(80, 64, 3), (40, 24, 4)]           # second call: a smaller tile
    for w, h, seed in sizes:
        print(f"noalloc: tile {w}x{h} seed {seed}", flush=True)
        cfg = trix_config(w, h)
        emap = _ramp(w, h)
        em_ref = F.VirtualEmulsionLayer.synthesize(cfg, seed)
        ref, st_ref = _reference(engine, em_ref, emap, x, seed + 1)
        em = F.VirtualEmulsionLayer.synthesize(cfg, seed)
        got, st = _noalloc(engine, em, emap, x, seed + 1)
        assert ref.sum() > 1000
        assert np.array_equal(got, ref)
        for k in STAT_KEYS:
            assert st[k] == st_ref[k], k
        assert st["kernel_ms"] > 0


def test_noalloc_transport_on_the_callers_stream_without_stats(engine):
    x = F.ExposureConfig(0.0, 0.12, 0.9, 48, 8)
    cfg = trix_config(48, 40)
    emap = _ramp(48, 40)
    em_ref = F.VirtualEmulsionLayer.synthesize(cfg, 9)
    ref, _ = _reference(engine, em_ref, emap, x, 10)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 9)
    s = torch.cuda.Stream()
    print("noalloc: caller stream, no stats", flush=True)
    got, _ = _noalloc(engine, em, emap, x, 10, stream=C.c_void_p(s.cuda_stream), with_stats=False)
    assert np.array_equal(got, ref)


def test_noalloc_rejects_a_tile_that_does_not_match_the_emulsion(engine):
    cfg = trix_config(32, 32)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 1)
    nodes, _ = em.device_layout()
    tile = L.DeviceTileC(None, None, None, None, nodes.shape[0] + 1, len(em), 0, 0)
    cfg_x = F.ExposureConfig(1.0, 0.12, 0.9, 48, 0).c()
    st = L.lib().fg_expose_noalloc(engine._h, em._h, C.byref(tile), C.byref(cfg_x), 1, None, None)
    assert st == 7  # FG_ERR_INVALID_ARG
    assert b"nodes" in L.lib().fg_last_error()

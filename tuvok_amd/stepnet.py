"""Learned step proposal: the network, its training data and its training.

StepNet wraps the device network (libtvkhip.so: BF16 MFMA inference and
training kernels).  The learned march decides per chunk of 32 fixed steps of
a ray's segment through a brick.  Training data are chunk samples produced by
the fixed-step march itself (HIPRaycaster.GenerateSegments): for every
recorded chunk that is not fully transparent, the network input, the ray's
transmittance before the chunk and the chunk's RGBA marched at 1, 2, 4 and 8
times the fine step.  labels() turns them into targets: the largest
multiplier whose error against the fine march stays within
tau + rel * (the chunk's contribution), and whether the whole chunk
contributes at most tau_skip.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import TvkError

N_FEAT = 24


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class StepNet:
    def __init__(self, device=0, seed=0):
        self._L = _abi.load()
        h = C.c_void_p()
        rc = self._L.tvk_stepnet_create(C.byref(h), device, int(seed))
        if rc != 0:
            raise TvkError(rc, "tvk_stepnet_create failed (no GPU?)")
        self._h = h
        self.n_param = self._L.tvk_stepnet_param_count()
        self._n = 0

    def _chk(self, rc):
        if rc != 0:
            raise TvkError(rc, self._L.tvk_stepnet_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.tvk_stepnet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        th = np.zeros(self.n_param, np.float32)
        self._chk(self._L.tvk_stepnet_get_params(self._h, _fp(th), self.n_param))
        return th

    def set_params(self, th):
        th = np.ascontiguousarray(th, np.float32)
        self._chk(self._L.tvk_stepnet_set_params(self._h, _fp(th), th.size))

    def infer(self, feat):
        feat = np.ascontiguousarray(feat, np.float32).reshape(-1, N_FEAT)
        out = np.zeros((feat.shape[0], 2), np.float32)
        self._chk(self._L.tvk_stepnet_infer(self._h, _fp(feat), feat.shape[0], _fp(out)))
        return out

    def set_data(self, feat, tgt):
        feat = np.ascontiguousarray(feat, np.float32).reshape(-1, N_FEAT)
        tgt = np.ascontiguousarray(tgt, np.float32).reshape(-1, 3)
        self._chk(self._L.tvk_stepnet_set_data(self._h, _fp(feat), _fp(tgt), feat.shape[0]))
        self._n = feat.shape[0]

    def grad(self, n=None):
        """(gradient of the mean weighted loss over the first n loaded samples, summed loss)"""
        g = np.zeros(self.n_param + 1, np.float32)
        self._chk(self._L.tvk_stepnet_grad(self._h, int(n or self._n), _fp(g)))
        return g[:-1], float(g[-1])

    def train(self, epochs=10, batch=4096, lr=2e-3, weight_decay=0.0, seed=0, n=None):
        losses = np.zeros(int(epochs), np.float32)
        self._chk(self._L.tvk_stepnet_train(self._h, int(n or self._n), int(epochs), int(batch), float(lr),
                                            float(weight_decay), int(seed), _fp(losses)))
        return losses


def split_samples(rec):
    """Chunk records (n x 48) -> features, transmittance before, steps, RGBA[n, 4 multipliers, 4]."""
    rec = np.asarray(rec, np.float32)
    return rec[:, :24], rec[:, 24], rec[:, 25], rec[:, 26:42].reshape(-1, 4, 4)


def segment_errors(rec):
    """Error of each multiplier against the fine march, as the pixel sees it:
    transmittance before the chunk x the largest channel difference."""
    _, tb, _, c = split_samples(rec)
    d = np.abs(c - c[:, :1, :]).max(axis=2)          # n x 4
    return tb[:, None] * d


def labels(rec, tau, tau_skip, weight_pos=1.0, rel=0.0):
    """Targets (n x 3): t0 = log2(largest safe multiplier)/3, t1 = skip, weight.
    A multiplier is safe when it and every smaller one stay within
    tau + rel * (transmittance before x the chunk's alpha) of the fine march."""
    _, tb, _, c = split_samples(rec)
    e = segment_errors(rec)
    bound = tau + rel * tb * c[:, 0, 3]
    k = np.zeros(len(rec), np.int64)
    ok = np.ones(len(rec), bool)
    for m in (1, 2, 3):
        ok &= e[:, m] <= bound
        k = np.where(ok, m, k)
    contrib = tb * c[:, 0, :].max(axis=1)
    skip = (contrib <= tau_skip).astype(np.float32)
    w = np.where(skip > 0, 1.0, weight_pos).astype(np.float32)
    return np.stack([k / 3.0, skip, w], axis=1).astype(np.float32)


def decide(y, margin, skip_threshold, max_log2=3):
    """The kernel's decision rule on network outputs y (n x 2): (log2 multiplier, skip)."""
    k = np.clip(np.floor(3.0 * y[:, 0] + 0.5 - margin), 0, max_log2).astype(np.int64)
    p = 1.0 / (1.0 + np.exp(-y[:, 1].astype(np.float64)))
    return k, p > skip_threshold


def train_proposal(ren, rotations, cap=2_000_000, keep=1.0, tau=0.25 / 255, tau_skip=None, rel=0.0, epochs=20,
                   batch=4096, lr=2e-3, seed=0, net=None, device=0):
    """Chunk samples of the renderer's fixed-step march over the given views,
    their labels, and a network trained on them with the device kernels."""
    rec = ren.GenerateSegments(rotations, cap, keep, seed, fit=True)
    if len(rec) == 0:
        raise RuntimeError("no chunk samples (empty views?)")
    tgt = labels(rec, tau, tau / 2 if tau_skip is None else tau_skip, rel=rel)
    net = net or StepNet(device, seed=seed)
    net.set_data(split_samples(rec)[0], tgt)
    losses = net.train(epochs=epochs, batch=batch, lr=lr, seed=seed)
    info = dict(chunks=int(len(rec)), mean_safe_log2=float((tgt[:, 0] * 3).mean()),
                skip_fraction=float(tgt[:, 1].mean()), loss_first=float(losses[0]), loss_last=float(losses[-1]))
    return net, info


def calibrate(ren, net, rotations, bound=0.9 / 255, margins=(0.0, 0.5, 1.0, 1.5, 2.0, 3.0),
              skips=(0.999, 1.0), fallbacks=(0.0, 0.05, 0.02, 0.01, 0.005, 0.002, 0.001), max_log2=3,
              skip_bounds=(0.0,), fallback_tvs=(0.005, 0.002, 0.001), tv_margins=(0.0, 1.0)):
    """The most aggressive proposal setting (least device time) whose frames stay
    within `bound` of the fixed-step frames on the given views (fallback 0: no
    fallback).  skip_bounds: contribution bounds below which a chunk is skipped
    without the network (0: only chunks that cannot contribute).  fallback_tvs
    (with tv_margins, no skip, no plain fallback): the variation-risk fallback.
    The fixed march itself is a candidate: when no setting is faster within the
    bound the result is None (march the fixed step).  Returns (setting or None,
    [(setting, max error, ms)])."""
    def timed():
        # the faster of two launches: one noisy launch can hide the best setting
        ms = []
        for _ in range(2):
            ren.PaintViews(rotations, download=False)
            ms.append(float(ren.last_stats.ms))
        return min(ms)

    ren.SetStepProposal(None)
    refs = ren.PaintViews(rotations)
    t_fixed = timed()
    tried, best = [(None, 0.0, t_fixed)], (None, t_fixed)
    props = [dict(margin=m, skip_threshold=s, max_log2=max_log2, skip_bound=sb, fallback_bound=fb)
             for sb in skip_bounds for fb in fallbacks for m in margins for s in skips]
    props += [dict(margin=m, skip_threshold=1.0, max_log2=max_log2, skip_bound=skip_bounds[0], fallback_bound=0.0,
                   fallback_tv=tv) for tv in fallback_tvs for m in tv_margins]
    for prop in props:
        ren.SetStepProposal(net, **prop)
        imgs = ren.PaintViews(rotations)
        err = float(np.abs(imgs - refs).max())
        ms = timed() if err <= bound else float(ren.last_stats.ms)
        tried.append((prop, err, ms))
        if err <= bound and (best is None or ms < best[1]):
            best = (prop, ms)
    ren.SetStepProposal(None)
    return best[0], tried

"""Learned warm start of the surface solve (no reference counterpart).

The reference's GMRES starts from zero (gmres.rs:38-39).  A pointwise neural
operator predicts the rigid-body surface potential from the geometry, the
plane wave and ka; its guess seeds GMRES with FP64 refinement
(``solve.sweep(..., x0=...)``), and the GMRES residual judges it: the final
answer is the solver's, the network only saves iterations.

    sur = Surrogate(width=256, seed=0)
    data = sur.dataset([(predata, waves, frequencies, phi), ...])   # solver outputs
    sur.fit(data, steps=3000, batch=16384)
    x0 = sur.predict(predata, frequencies, wave)                   # (nfreq, n)
    phi, vn, stats = solve.sweep(predata, x0=x0)

Training and inference run on libccab's HIP kernels (BF16 MFMA GEMMs with
FP32 accumulation, FP32 Adam).
"""
from __future__ import annotations

import ctypes as ct
import math

import numpy as np
import torch

from . import capi
from .preprocess import IncidentWave, PreData

FEAT = 32


def _ptr(t) -> ct.c_void_p:
    return ct.c_void_p(t.data_ptr()) if t is not None else None


class Surrogate:
    def __init__(self, width: int = 256, seed: int = 0, lr: float = 1e-3, weight_decay: float = 0.0,
                 device: int = 0, ctx=None):
        self.ctx = ctx or capi.default_context(device)
        self.device = torch.device("cuda", device)
        cfg = capi.SurrogateCfg()
        cfg.width = int(width)
        cfg.seed = int(seed)
        cfg.lr = float(lr)
        cfg.beta1, cfg.beta2, cfg.eps = 0.9, 0.999, 1e-8
        cfg.weight_decay = float(weight_decay)
        h = ct.c_void_p()
        self.ctx.call("ccab_surrogate_create", ct.byref(cfg), ct.byref(h))
        self._h = h
        self.width = int(width)
        self.lr = float(lr)

    # ---- parameters -------------------------------------------------------
    @property
    def num_params(self) -> int:
        return int(capi.lib().ccab_surrogate_num_params(self._h))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.num_params, np.float32)
        self.ctx.call("ccab_surrogate_get_params", self._h, p.ctypes.data_as(ct.c_void_p))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params:
            raise ValueError("parameter count mismatch")
        self.ctx.call("ccab_surrogate_set_params", self._h, p.ctypes.data_as(ct.c_void_p))

    def set_lr(self, lr: float):
        rc = capi.lib().ccab_surrogate_set_lr(self._h, float(lr))
        if rc != capi.OK:
            raise capi.CcabError(rc, "bad learning rate")
        self.lr = float(lr)

    # ---- data ---------------------------------------------------------------
    def features(self, predata: PreData, frequencies, waves, phi: torch.Tensor | None = None):
        """Features (P, 32) uint16/bf16 and, with phi (nsys, n), targets (P, 2)
        for nsys systems on predata's mesh: frequencies[s] with waves[s]."""
        fs = list(frequencies)
        ws = list(waves) if isinstance(waves, (list, tuple)) else [waves] * len(fs)
        if len(ws) != len(fs):
            raise ValueError("one wave per frequency")
        nsys, n = len(fs), predata.get_num_eqn()
        ph = predata.physics_array(fs)
        warr = (capi.Wave * nsys)(*[w.to_c() for w in ws])
        feat = torch.empty((nsys * n, FEAT), dtype=torch.int16, device=self.device)
        tgt = None
        if phi is not None:
            phi = phi.to(torch.complex128).contiguous()
            if phi.shape != (nsys, n):
                raise ValueError("phi must have shape (nsys, n)")
            tgt = torch.empty((nsys * n, 2), dtype=torch.float32, device=self.device)
        self.ctx.call("ccab_surrogate_features", ph, warr, nsys, _ptr(predata.d_cpts), n,
                      _ptr(phi), _ptr(feat), _ptr(tgt))
        self.ctx.synchronize()
        return feat, tgt

    def dataset(self, items):
        """items: iterable of (predata, frequencies, waves, phi) -> (feat, target)."""
        fs, ts = [], []
        for pd, freqs, waves, phi in items:
            f, t = self.features(pd, freqs, waves, phi)
            fs.append(f)
            ts.append(t)
        return torch.cat(fs), torch.cat(ts)

    # ---- training -----------------------------------------------------------
    def loss_grad(self, feat: torch.Tensor, target: torch.Tensor):
        g = np.zeros(self.num_params, np.float32)
        loss = ct.c_double()
        self.ctx.call("ccab_surrogate_loss_grad", self._h, _ptr(feat.contiguous()),
                      _ptr(target.contiguous()), feat.shape[0], ct.byref(loss),
                      g.ctypes.data_as(ct.c_void_p))
        return loss.value, g

    def train_step(self, feat, target, index: torch.Tensor, want_loss=False):
        loss = ct.c_double()
        self.ctx.call("ccab_surrogate_train_step", self._h, _ptr(feat), _ptr(target), _ptr(index),
                      index.numel(), ct.byref(loss) if want_loss else None)
        return loss.value if want_loss else None

    def fit(self, data, steps: int = 2000, batch: int = 16384, seed: int = 0, lr: float | None = None,
            lr_min: float = 1e-5, log_every: int = 0):
        """Adam with cosine decay over random mini-batches; returns the losses
        sampled every log_every steps (and the last one)."""
        feat, target = data
        P = feat.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        lr0 = self.lr if lr is None else float(lr)
        losses = []
        perm = torch.randperm(P, generator=g)
        pos = 0
        for it in range(steps):
            if pos + batch > P:
                perm = torch.randperm(P, generator=g)
                pos = 0
            idx = perm[pos:pos + min(batch, P)].to(self.device)
            pos += batch
            self.set_lr(lr_min + 0.5 * (lr0 - lr_min) * (1 + math.cos(math.pi * it / steps)))
            want = (log_every and it % log_every == 0) or it == steps - 1
            loss = self.train_step(feat, target, idx, want_loss=bool(want))
            if want:
                losses.append((it, loss))
        self.ctx.synchronize()
        return losses

    # ---- inference ----------------------------------------------------------
    def forward(self, feat: torch.Tensor) -> torch.Tensor:
        out = torch.empty((feat.shape[0], 2), dtype=torch.float32, device=self.device)
        self.ctx.call("ccab_surrogate_forward", self._h, _ptr(feat.contiguous()), feat.shape[0],
                      _ptr(out))
        self.ctx.synchronize()
        return out

    def predict(self, predata: PreData, frequencies=None, waves=None) -> torch.Tensor:
        """Initial guesses (nsys, n) of the surface potential."""
        fs = predata.get_frequencies() if frequencies is None else list(frequencies)
        if waves is None:
            waves = predata.incident_wave
        ws = list(waves) if isinstance(waves, (list, tuple)) else [waves] * len(fs)
        nsys, n = len(fs), predata.get_num_eqn()
        ph = predata.physics_array(fs)
        warr = (capi.Wave * nsys)(*[w.to_c() for w in ws])
        x0 = torch.empty((nsys, n), dtype=torch.complex128, device=self.device)
        self.ctx.call("ccab_surrogate_predict", self._h, ph, warr, nsys, _ptr(predata.d_cpts), n,
                      _ptr(x0))
        self.ctx.synchronize()
        return x0

    def predict_bodies(self, bodies) -> list:
        """Initial guesses for bodies of uneven size (list of PreData, each at
        its current frequency with its own plane wave) in one pass; one (n_b,)
        tensor per body, as solve.solve_bodies takes them."""
        nb = len(bodies)
        ns = [pd.get_num_eqn() for pd in bodies]
        n_arr = (ct.c_int64 * nb)(*ns)
        cpts = torch.cat([pd.d_cpts for pd in bodies]).contiguous()
        ph = (capi.Physics * nb)(*[pd.physics() for pd in bodies])
        warr = (capi.Wave * nb)(*[pd.incident_wave.to_c() for pd in bodies])
        x0 = torch.empty(sum(ns), dtype=torch.complex128, device=self.device)
        self.ctx.call("ccab_surrogate_predict_bodies", self._h, nb, n_arr, _ptr(cpts), ph, warr,
                      _ptr(x0))
        self.ctx.synchronize()   # x0 is read on torch's stream next
        return list(x0.split(ns))

    def close(self):
        if getattr(self, "_h", None):
            capi.lib().ccab_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_direction(rng: np.random.Generator):
    d = rng.standard_normal(3)
    return tuple(d / np.linalg.norm(d))


def plane_waves(directions, amplitude=(1.0, 0.0)):
    return [IncidentWave(capi.WAVE_PLANE, tuple(d), amplitude) for d in directions]

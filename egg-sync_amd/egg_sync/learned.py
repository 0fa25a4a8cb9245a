"""Learned multi-iteration jump for the EGG-SynC solve (no counterpart in the reference).

A small MLP (4d inputs -> 128 -> 128 -> d, ReLU, FP32) predicts, for every
point, the remaining displacement to its attractor, in units of eps, from:
its position, the mean sin and mean cos of its own cell (the solver's cell
totals divided by the count) and its displacement over the last plain
iteration divided by eps.  It is trained on trajectories the plain solver
itself generates (`trajectory_samples`), with the project's HIP forward /
backward / Adam kernels (`egg_mlp_train_step_device`).

`solve_learned` runs `t0` plain iterations, applies the jump once, resumes the
plain iteration until EGG-SynC's own stop rule holds, and labels with the
plain rule.  The plain method judges the result: `solve_learned(verify=True)`
also runs the plain solve and, if the partitions differ, returns the plain
result with `fallback=True`.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _native as nat
from . import solver

HIDDEN = 128


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class JumpMLP:
    """Parameters, Adam state and workspace of the jump network on the GPU."""

    def __init__(self, d: int = 3, hidden: int = HIDDEN, seed: int = 0):
        self.d, self.n_in, self.h, self.n_out = d, 4 * d, hidden, d
        cnt = ctypes.c_int64()
        nat.check(nat.lib().egg_mlp_param_count(self.n_in, hidden, d, ctypes.byref(cnt)))
        self.n_params = cnt.value
        rng = np.random.default_rng(seed)
        parts = []
        for fan_in, fan_out in ((self.n_in, hidden), (hidden, hidden), (hidden, d)):
            parts.append(rng.normal(0.0, np.sqrt(2.0 / fan_in), (fan_out, fan_in)).ravel())
            parts.append(np.zeros(fan_out))
        p = np.concatenate(parts).astype(np.float32)
        p[-d:] = 0.0
        p[-d - d * hidden:-d] *= 0.1  # start close to the identity map of the plain step
        assert p.size == self.n_params
        self.params = torch.from_numpy(p).cuda()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.step = 0
        self._work = None
        self._work_b = 0

    def _workspace(self, batch):
        if self._work is None or self._work_b < batch:
            nf = nat.lib().egg_mlp_workspace_floats(self.n_in, self.h, self.n_out, batch)
            self._work = torch.empty(int(nf), dtype=torch.float32, device="cuda")
            self._work_b = batch
        return self._work

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        X = X.contiguous()
        Y = torch.empty((X.shape[0], self.n_out), dtype=torch.float32, device=X.device)
        nat.check(nat.lib().egg_mlp_forward_device(nat.ptr(self.params), self.n_in, self.h,
                                                   self.n_out, nat.ptr(X), X.shape[0], nat.ptr(Y),
                                                   _stream()), "egg_mlp_forward_device")
        return Y

    def grad_and_loss(self, X: torch.Tensor, T: torch.Tensor):
        """Gradient (flat, parameter layout) and loss on a batch, no update."""
        return self._step(X, T, 0.0)

    def train_step(self, X: torch.Tensor, T: torch.Tensor, lr: float):
        self.step += 1
        return self._step(X, T, lr)

    def _step(self, X, T, lr):
        X, T = X.contiguous(), T.contiguous()
        B = X.shape[0]
        work = self._workspace(B)
        loss = ctypes.c_double()
        nat.check(nat.lib().egg_mlp_train_step_device(
            nat.ptr(self.params), nat.ptr(self.m), nat.ptr(self.v), self.n_in, self.h, self.n_out,
            nat.ptr(X), nat.ptr(T), B, float(lr), max(self.step, 1), nat.ptr(work), work.numel(),
            nat.ptr(self.grad), ctypes.byref(loss), _stream()), "egg_mlp_train_step_device")
        return self.grad.clone(), loss.value

    def fit(self, X: torch.Tensor, T: torch.Tensor, epochs: int = 4, batch: int = 8192,
            lr: float = 1e-3, seed: int = 0, log=None):
        g = torch.Generator(device="cpu").manual_seed(seed)
        n = X.shape[0]
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).cuda()
            tot, k = 0.0, 0
            for s in range(0, n - batch + 1, batch):
                idx = perm[s:s + batch]
                _, l = self.train_step(X[idx], T[idx], lr)
                tot += l
                k += 1
            losses.append(tot / max(k, 1))
            if log:
                log(f"epoch {ep} loss {losses[-1]:.5f}")
        return losses

    def state(self):
        return self.params.cpu().numpy().copy()

    def load(self, p):
        self.params.copy_(torch.from_numpy(np.asarray(p, np.float32)).cuda())

    @classmethod
    def from_file(cls, path: str) -> "JumpMLP":
        """Network saved by `python -m egg_sync.train_jump` (a plain .npz, no pickles)."""
        with np.load(path, allow_pickle=False) as z:
            net = cls(int(z["n_out"]), hidden=int(z["hidden"]))
            net.load(z["params"])
            net.record = str(z["record"])
        return net


def features(X: torch.Tensor, prev: torch.Tensor, eps: float) -> torch.Tensor:
    n, d = X.shape
    F = torch.empty((n, 4 * d), dtype=torch.float32, device=X.device)
    nat.check(nat.lib().egg_sync_features_device(nat.ptr(X.contiguous()),
                                                 nat.ptr(None if prev is None else prev.contiguous()),
                                                 n, d, float(eps), nat.ptr(F), _stream()),
              "egg_sync_features_device")
    return F


def plain_steps(X: torch.Tensor, eps: float, t: int):
    """t plain iterations (the solver's own pass, in its visiting order);
    returns (positions after t-1 passes, after t passes)."""
    prev, cur, order = X, X, None
    for _ in range(t):
        nxt, order = solver.iteration_device(cur, eps, order)
        prev, cur = cur, nxt
    return prev, cur


def trajectory_samples(X: np.ndarray, eps: float, t0: int = 1):
    """Training pairs from one plain solve: features after t0 passes and the
    remaining displacement to the converged position, in units of eps."""
    D = solver._dev_f32(X)
    final = solver.solve(D, eps)["X"]
    prev, cur = plain_steps(D, eps, t0)
    F = features(cur, prev, eps)
    T = (torch.from_numpy(final).cuda() - cur) / eps
    return F, T


def solve_learned(X, eps: float, mlp: JumpMLP, t0: int = 1, verify: bool = False,
                  alpha: float = 1.0, clip: float = 0.0):
    """t0 plain passes, one learned jump, plain passes until EGG-SynC's stop rule.

    The jump moves each point by alpha * eps * (predicted displacement / eps),
    each component first clamped to [-clip, clip] (no clamp if clip <= 0)."""
    D = solver._dev_f32(X)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    prev, cur = plain_steps(D, eps, t0)
    F = features(cur, prev, eps)
    Y = mlp.forward(F)
    jumped = cur.clone()
    nat.check(nat.lib().egg_sync_jump_device(nat.ptr(jumped), nat.ptr(Y), jumped.shape[0],
                                             jumped.shape[1], float(alpha * eps), float(clip),
                                             _stream()),
              "egg_sync_jump_device")
    r = solver.solve(jumped, eps)
    torch.cuda.synchronize()
    t_learned = time.perf_counter() - t_start
    out = dict(labels=r["labels"], X=r["X"], iterations=t0 + r["stats"]["iterations"],
               seconds=t_learned, fallback=False)
    if verify:
        t1 = time.perf_counter()
        p = solver.solve(D, eps)
        torch.cuda.synchronize()
        out["plain_seconds"] = time.perf_counter() - t1
        out["plain_iterations"] = p["stats"]["iterations"]
        from .metrics import same_partition
        out["same_partition"] = same_partition(r["labels"], p["labels"])
        if not out["same_partition"]:
            out.update(labels=p["labels"], X=p["X"], fallback=True)
    return out

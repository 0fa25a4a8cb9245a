"""GPU tests of the learned jump: the HIP MLP kernels against the FP64 CPU
oracle (forward, gradients, Adam), training on the plain solver's own
trajectories, and the learned solve judged by the plain solve's labels."""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def L():
    assert torch.cuda.is_available()
    from egg_sync import learned
    return learned


def _rand_batch(rng, B, d=3):
    F = np.concatenate([rng.random((B, d)), rng.uniform(-1, 1, (B, 2 * d)),
                        rng.normal(0, 0.5, (B, d))], 1).astype(np.float32)
    T = rng.normal(0, 0.5, (B, d)).astype(np.float32)
    return torch.from_numpy(F).cuda(), torch.from_numpy(T).cuda()


def test_forward_matches_oracle(L):
    import mlp_oracle as M
    rng = np.random.default_rng(1)
    net = L.JumpMLP(3, seed=3)
    F, _ = _rand_batch(rng, 5000)
    got = net.forward(F).cpu().numpy()
    ref, _ = M.forward(net.state(), F.cpu().numpy(), 12, 128, 3)
    assert np.abs(got - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())


def test_gradients_match_oracle(L):
    import mlp_oracle as M
    rng = np.random.default_rng(2)
    net = L.JumpMLP(3, seed=4)
    F, T = _rand_batch(rng, 4096)
    g, loss = net.grad_and_loss(F, T)
    lr, gr = M.loss_and_grad(net.state(), F.cpu().numpy(), T.cpu().numpy(), 12, 128, 3)
    assert abs(loss - lr) <= 1e-5 * max(1.0, abs(lr))
    g = g.cpu().numpy()
    assert np.abs(g - gr).max() <= 2e-4 * np.abs(gr).max()
    # Adam step against the oracle's
    p0 = net.state()
    net.train_step(F, T, lr=1e-3)
    p1, _, _ = M.adam_step(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    assert np.abs(net.state() - p1).max() <= 1e-6


def test_training_on_plain_trajectories_beats_no_jump(L):
    from egg_sync import data
    torch.manual_seed(0)
    Fs, Ts = [], []
    for seed in range(8):
        X, eps = data.config1(seed=100 + seed, n=20000)
        F, T = L.trajectory_samples(X, eps)
        Fs.append(F)
        Ts.append(T)
    F, T = torch.cat(Fs[:-2]), torch.cat(Ts[:-2])
    Fv, Tv = torch.cat(Fs[-2:]), torch.cat(Ts[-2:])
    net = L.JumpMLP(3, seed=0)
    base = float((Tv ** 2).sum(1).mean())           # predicting no jump
    losses = net.fit(F, T, epochs=6, batch=4096, lr=2e-3, log=_say)
    val = float(((net.forward(Fv) - Tv) ** 2).sum(1).mean())
    _say(f"held-out squared jump error: no jump {base:.4f}, trained MLP {val:.4f}")
    assert losses[-1] < losses[0]
    assert val < 0.8 * base


def test_learned_solve_is_judged_by_plain_labels(L):
    from egg_sync import data, metrics
    net = L.JumpMLP(3, seed=0)
    Fs, Ts = [], []
    for seed in range(6):
        X, eps = data.config1(seed=200 + seed, n=20000)
        F, T = L.trajectory_samples(X, eps)
        Fs.append(F)
        Ts.append(T)
    net.fit(torch.cat(Fs), torch.cat(Ts), epochs=4, batch=4096, lr=2e-3)
    X, eps = data.config1(seed=300, n=20000)
    r = L.solve_learned(X, eps, net, verify=True)
    _say(f"learned: {r['iterations']} passes {r['seconds']*1e3:.1f} ms, plain: "
         f"{r['plain_iterations']} passes {r['plain_seconds']*1e3:.1f} ms, "
         f"same partition: {r['same_partition']}, fallback: {r['fallback']}")
    from egg_sync import solver
    p = solver.solve(X, eps)
    assert metrics.same_partition(r["labels"], p["labels"])

"""GPU tests of the learned jump: the HIP MLP kernels against the FP64 CPU
oracle (forward, gradients, Adam), training on the plain solver's own
trajectories, the learned solve on its own (no fallback) against the plain
solve, and the verified form's fallback rule."""
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


def _trained_net(L):
    net = L.JumpMLP(3, seed=0)
    Fs, Ts = [], []
    for seed in range(6):
        from egg_sync import data
        X, eps = data.config1(seed=200 + seed, n=20000)
        F, T = L.trajectory_samples(X, eps)
        Fs.append(F)
        Ts.append(T)
    net.fit(torch.cat(Fs), torch.cat(Ts), epochs=4, batch=4096, lr=2e-3)
    return net


@pytest.fixture(scope="module")
def net(L):
    return _trained_net(L)


def test_learned_solve_unverified(L, net, O):
    """The learned solve on its own (verify=False, no fallback): one plain pass,
    the jump, plain passes to EGG-SynC's stop rule.  Its labels are the cell
    rule applied to its own final positions, and at this size (the regime the
    network was trained in) its partition is within 1% of the plain solve's."""
    from egg_sync import data, metrics, solver
    X, eps = data.config1(seed=300, n=20000)
    r = L.solve_learned(X, eps, net, t0=1, alpha=0.5, clip=0.5, verify=False)
    assert r["fallback"] is False and "same_partition" not in r
    assert O.same_partition(r["labels"], O.egg_labels(r["X"], eps))
    p = solver.solve(X, eps)
    dis = metrics.disagreement(r["labels"], p["labels"])
    _say(f"learned: {r['iterations']} passes {r['seconds']*1e3:.1f} ms; plain "
         f"{p['stats']['iterations']} passes; disagreement {dis}")
    assert dis["unmatched"] + dis["noise_differs"] <= 0.01 * len(X)


def test_learned_solve_verify_falls_back_only_on_disagreement(L, net):
    """verify=True returns the learned labels when the partitions agree and the
    plain ones (fallback) when they do not; a zero-length jump keeps the plain
    trajectory's partition and does not fall back."""
    from egg_sync import data, metrics, solver
    X, eps = data.config1(seed=301, n=20000)
    u = L.solve_learned(X, eps, net, t0=1, alpha=0.5, clip=0.5, verify=False)
    v = L.solve_learned(X, eps, net, t0=1, alpha=0.5, clip=0.5, verify=True)
    p = solver.solve(X, eps)
    same = metrics.same_partition(u["labels"], p["labels"])
    _say(f"jump alpha=0.5: same partition {same}, fallback {v['fallback']}")
    assert v["same_partition"] == same and v["fallback"] == (not same)
    assert np.array_equal(v["labels"], u["labels"] if same else p["labels"])
    z = L.solve_learned(X, eps, net, t0=1, alpha=0.0, verify=True)
    assert z["same_partition"] and not z["fallback"]
    assert np.array_equal(z["labels"], L.solve_learned(X, eps, net, t0=1, alpha=0.0)["labels"])

"""egg_sync_solve_device_noalloc: the solve with its workspace in the
library's scratch array (no allocation), checked through the C ABI against
egg_sync_solve_device -- labels, final positions and statistics bit for bit,
on two calls in a row, the second one smaller -- and its size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


def _call(nat, fn, D, eps, pair_groups=0):
    n, d = D.shape
    opt = nat.default_options(pair_groups=pair_groups)
    labels = torch.full((n,), -7, dtype=torch.int32, device=D.device)
    final = torch.full_like(D, -1.0)
    st = nat.Stats()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(nat.ptr(D), n, d, float(eps), ctypes.byref(opt), nat.ptr(labels), nat.ptr(final),
            ctypes.byref(st), stream)
    torch.cuda.synchronize()
    return rc, labels.cpu().numpy(), final.cpu().numpy(), st


@pytest.fixture(scope="module")
def nat():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from egg_sync import _native
    return _native


def test_noalloc_equals_device_solve_two_calls(nat):
    from egg_sync import data
    L = nat.lib()
    cases = [data.config1(n=120_000), data.config1(seed=7, n=30_000),
             (data.config2(n=20_000)[0], 0.1)]
    for X, eps in cases:
        D = torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda()
        _say("noalloc vs device solve", X.shape, eps)
        rc_a, lab_a, fin_a, st_a = _call(nat, L.egg_sync_solve_device, D, eps)
        rc_b, lab_b, fin_b, st_b = _call(nat, L.egg_sync_solve_device_noalloc, D, eps)
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b)
        assert st_a.iterations == st_b.iterations and st_b.iterations > 1
        assert st_a.n_clusters == st_b.n_clusters and st_a.n_cells == st_b.n_cells
        assert np.array_equal(lab_a, lab_b)
        assert np.array_equal(fin_a, fin_b)
        assert 0 < st_b.workspace_bytes <= 64 << 20
        _say("workspace bytes", st_b.workspace_bytes, "per point", st_b.workspace_bytes / len(X))


def test_noalloc_rejects_too_large(nat):
    from egg_sync import data
    X, eps = data.config1(n=1_000_000)
    D = torch.from_numpy(X).cuda()
    _say("noalloc size limit")
    rc, lab, _, _ = _call(nat, nat.lib().egg_sync_solve_device_noalloc, D, eps)
    assert rc == nat.E_TOO_BIG
    assert "no-alloc" in nat.error_string(rc)
    assert (lab == -7).all()  # nothing was written
    # a call after the rejection still works
    X2, eps2 = data.config1(n=10_000)
    D2 = torch.from_numpy(X2).cuda()
    rc_a, lab_a, fin_a, _ = _call(nat, nat.lib().egg_sync_solve_device, D2, eps2)
    rc_b, lab_b, fin_b, _ = _call(nat, nat.lib().egg_sync_solve_device_noalloc, D2, eps2)
    assert rc_a == rc_b == 0
    assert np.array_equal(lab_a, lab_b) and np.array_equal(fin_a, fin_b)

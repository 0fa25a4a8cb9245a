"""Thin Python wrappers of the C ABI for device-resident data (torch tensors on the GPU).

These call libeggsync.so through ctypes; torch only provides device memory and
the current stream.  Used by the tests, the benchmark and callers that want
statistics or intermediate results of the solve.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as nat


def _dev_f32(X, device="cuda"):
    if isinstance(X, torch.Tensor):
        return X.to(device=device, dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).to(device)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def solve(X, eps: float, *, stop_rule=nat.STOP_EGG, label_rule=None, lam=1 - 1e-3, max_iter=0,
          compute_rc=False, timing=False, iters_per_sync=4, return_positions=True,
          itr_cap=16384, use_graph=False, queue_blocks=0, queue_threads=0, tile_group=0,
          pair_groups=0):
    """Solve on the GPU.  Returns dict(labels int32 numpy, X final numpy or None, stats).

    use_graph replays the loop body from a captured hipGraph; queue_blocks /
    queue_threads set the launch shape of the tile-queue kernels (0 = automatic);
    tile_group lets up to that many neighbouring cells share an update tile;
    pair_groups picks the cells updated point pair by point pair (0 automatic,
    1 sparse groups, 2 every group).  The first three do not change the result;
    tile_group and pair_groups change the order of FP32 additions only."""
    D = _dev_f32(X)
    n, d = D.shape
    if label_rule is None:
        label_rule = nat.LABELS_SYNC if stop_rule == nat.STOP_RC else nat.LABELS_CELLS
    opt = nat.default_options(stop_rule=stop_rule, label_rule=label_rule, lam=lam,
                              max_iter=max_iter, compute_rc=int(compute_rc), timing=int(timing),
                              iters_per_sync=iters_per_sync, use_graph=int(use_graph),
                              queue_blocks=queue_blocks, queue_threads=queue_threads,
                              tile_group=tile_group, pair_groups=pair_groups)
    labels = torch.empty(n, dtype=torch.int32, device=D.device)
    final = torch.empty_like(D) if return_positions else None
    st = nat.Stats()
    itr = np.zeros(itr_cap, np.float32)
    st.itr_ms = itr.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    st.itr_ms_cap = itr_cap
    cells_tr = np.zeros(4096, np.int32)
    rc_tr = np.zeros(4096, np.float64)
    st.trace_cap = 4096
    st.cells_trace = cells_tr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    st.rc_trace = rc_tr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = nat.lib().egg_sync_solve_device(nat.ptr(D), n, d, float(eps), ctypes.byref(opt),
                                         nat.ptr(labels), nat.ptr(final), ctypes.byref(st),
                                         _stream())
    nat.check(rc, "egg_sync_solve_device")
    stats = {k: getattr(st, k) for k in ("iterations", "status", "n_clusters", "n_cells", "rc",
                                         "total_ms", "update_kernel_ms", "update_launches",
                                         "workspace_bytes")}
    stats["stage_ms"] = list(st.stage_ms)
    stats["itr_ms"] = itr[: st.iterations].copy() if timing else None
    it = min(st.iterations, 4096)
    stats["cells_trace"] = cells_tr[:it].copy()
    stats["rc_trace"] = rc_tr[:it].copy()
    return dict(labels=labels.cpu().numpy(), X=None if final is None else final.cpu().numpy(),
                stats=stats, code=rc)


def pack_instances(datasets, eps_values=None):
    """Packs datasets (arrays or tensors, n_k x d_k) into one device buffer.

    With `eps_values` a list of lists, instance list = every dataset with each
    of its eps values, sharing the dataset's rows; otherwise `datasets` is a
    list of (X, eps) pairs.  Returns (D flat device tensor, [(offset, n, d, eps)])."""
    if eps_values is None:
        pairs = list(datasets)
        datasets = [x for x, _ in pairs]
        eps_values = [[e] for _, e in pairs]
    flat, specs, off = [], [], 0
    for X, es in zip(datasets, eps_values):
        T = _dev_f32(X)
        n, d = T.shape
        flat.append(T.reshape(-1))
        for e in es:
            specs.append((off, n, d, float(e)))
        off += n * d
    D = torch.cat(flat) if flat else torch.empty(0, dtype=torch.float32, device="cuda")
    return D, specs


def solve_batch(D: torch.Tensor, specs, *, stop_rule=nat.STOP_EGG, label_rule=None,
                lam=1 - 1e-3, max_iter=0, timing=False, iters_per_sync=4, use_graph=False,
                streams=4, queue_blocks=0, queue_threads=0, tile_group=0, pair_groups=0,
                return_positions=True):
    """Many independent solves in one call (egg_sync_solve_batch_device).

    `D` is a flat float32 device tensor, `specs` a list of (data_offset, n, d,
    eps).  Returns dict(labels=[int32 numpy per instance], X=[final positions or
    None], results=[dict per instance], stats)."""
    if label_rule is None:
        label_rule = nat.LABELS_SYNC if stop_rule == nat.STOP_RC else nat.LABELS_CELLS
    opt = nat.default_options(stop_rule=stop_rule, label_rule=label_rule, lam=lam,
                              max_iter=max_iter, timing=int(timing),
                              iters_per_sync=iters_per_sync, use_graph=int(use_graph),
                              queue_blocks=queue_blocks, queue_threads=queue_threads,
                              streams=streams, tile_group=tile_group, pair_groups=pair_groups)
    k = len(specs)
    inst = (nat.Instance * max(k, 1))()
    rows = fl = 0
    for i, (o, n, d, e) in enumerate(specs):
        inst[i].data_offset, inst[i].n, inst[i].d, inst[i].eps = int(o), int(n), int(d), float(e)
        rows += int(n)
        fl += int(n) * int(d)
    labels = torch.empty(max(rows, 1), dtype=torch.int32, device=D.device)
    final = torch.empty(max(fl, 1), dtype=torch.float32, device=D.device) \
        if return_positions else None
    res = (nat.InstanceResult * max(k, 1))()
    st = nat.Stats()
    rc = nat.lib().egg_sync_solve_batch_device(nat.ptr(D), inst, k, ctypes.byref(opt),
                                               nat.ptr(labels), nat.ptr(final), res,
                                               ctypes.byref(st), _stream())
    nat.check(rc, "egg_sync_solve_batch_device")
    lab = labels.cpu().numpy()
    fin = final.cpu().numpy() if final is not None else None
    out_l, out_x, out_r = [], [], []
    r0 = f0 = 0
    for i, (o, n, d, e) in enumerate(specs):
        out_l.append(lab[r0:r0 + n].copy())
        out_x.append(None if fin is None else fin[f0:f0 + n * d].reshape(n, d).copy())
        out_r.append(dict(iterations=res[i].iterations, status=res[i].status,
                          n_clusters=res[i].n_clusters, n_cells=res[i].n_cells))
        r0 += n
        f0 += n * d
    stats = {k_: getattr(st, k_) for k_ in ("iterations", "status", "n_clusters", "n_cells",
                                             "total_ms", "workspace_bytes")}
    stats["stage_ms"] = list(st.stage_ms)
    return dict(labels=out_l, X=out_x, results=out_r, stats=stats, code=rc)


def build_grid(X, eps: float, order_in=None):
    """Grid build on the GPU; same outputs as the CPU oracle's build_grid."""
    D = _dev_f32(X)
    n, d = D.shape
    g = nat.geometry(n, d, eps)
    dev = D.device
    oi = None if order_in is None else torch.as_tensor(np.asarray(order_in, np.int32)).to(dev)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    cell_of = torch.empty(n, dtype=torch.int32, device=dev)
    cell_start = torch.empty(n + 1, dtype=torch.int32, device=dev)
    cell_g = torch.empty(n * d, dtype=torch.int32, device=dev)
    ob = torch.empty(g["O"], dtype=torch.int32, device=dev)
    oe = torch.empty(g["O"], dtype=torch.int32, device=dev)
    C = ctypes.c_int32()
    rc = nat.lib().egg_sync_build_grid_device(nat.ptr(D), n, d, float(eps), nat.ptr(oi),
                                              nat.ptr(order), nat.ptr(cell_of),
                                              nat.ptr(cell_start), nat.ptr(cell_g), nat.ptr(ob),
                                              nat.ptr(oe), ctypes.byref(C), _stream())
    nat.check(rc, "egg_sync_build_grid_device")
    C = C.value
    return dict(order=order.cpu().numpy(), cell_of=cell_of.cpu().numpy(),
                cell_start=cell_start[: C + 1].cpu().numpy(),
                cell_g=cell_g[: C * d].cpu().numpy().reshape(C, d), ob=ob.cpu().numpy(),
                oe=oe.cpu().numpy(), C=C)


def iteration_device(D: torch.Tensor, eps: float, order_in: torch.Tensor = None):
    """One EGG iteration, device in / device out: (next positions, visiting order)."""
    D = D.contiguous()
    n, d = D.shape
    Xn = torch.empty_like(D)
    order = torch.empty(n, dtype=torch.int32, device=D.device)
    fl, ff = ctypes.c_int32(), ctypes.c_int32()
    nat.check(nat.lib().egg_sync_iteration_device(nat.ptr(D), n, d, float(eps), nat.ptr(order_in),
                                                  nat.ptr(Xn), nat.ptr(order), None, None, None,
                                                  None, ctypes.byref(fl), ctypes.byref(ff),
                                                  _stream()), "egg_sync_iteration_device")
    return Xn, order


def iteration(X, eps: float, order_in=None):
    """One EGG iteration on the GPU; same outputs as the CPU oracle's egg_iteration."""
    D = _dev_f32(X)
    n, d = D.shape
    dev = D.device
    oi = None if order_in is None else torch.as_tensor(np.asarray(order_in, np.int32)).to(dev)
    Xn = torch.empty_like(D)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    N, half, border, synced = (torch.empty(n, dtype=torch.int32, device=dev) for _ in range(4))
    fl, ff = ctypes.c_int32(), ctypes.c_int32()
    rc = nat.lib().egg_sync_iteration_device(nat.ptr(D), n, d, float(eps), nat.ptr(oi),
                                             nat.ptr(Xn), nat.ptr(order), nat.ptr(N),
                                             nat.ptr(half), nat.ptr(border), nat.ptr(synced),
                                             ctypes.byref(fl), ctypes.byref(ff), _stream())
    nat.check(rc, "egg_sync_iteration_device")
    return dict(X_next=Xn.cpu().numpy(), order=order.cpu().numpy(), N=N.cpu().numpy(),
                half=half.cpu().numpy(), border=border.cpu().numpy(),
                synced=synced.cpu().numpy(), flag_local=fl.value, flag_final=ff.value)


def simple_solve(X, eps: float, *, lam=1 - 1e-3, max_iter=100, itr_cap=4096):
    """Brute-force SynC on the GPU (simple_GPU_SynC.cu:105-237): every point scans
    all points each pass; stops once r >= lam or after max_iter passes (the
    reference's cap is 100).  Returns dict(labels int32 numpy -- smallest index
    among a point's eps-neighbours, numbered 0..k-1 --, X final numpy, stats)."""
    D = _dev_f32(X)
    n, d = D.shape
    labels = torch.empty(n, dtype=torch.int32, device=D.device)
    final = torch.empty_like(D)
    st = nat.Stats()
    itr = np.zeros(itr_cap, np.float32)
    st.itr_ms = itr.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    st.itr_ms_cap = itr_cap
    r_tr = np.zeros(itr_cap, np.float64)
    st.trace_cap = itr_cap
    st.rc_trace = r_tr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = nat.lib().egg_sync_simple_solve_device(nat.ptr(D), n, d, float(eps), float(lam),
                                                int(max_iter), nat.ptr(labels), nat.ptr(final),
                                                ctypes.byref(st), _stream())
    nat.check(rc, "egg_sync_simple_solve_device")
    it = st.iterations
    stats = {k: getattr(st, k) for k in ("iterations", "status", "n_clusters", "rc", "total_ms",
                                         "workspace_bytes")}
    stats["stage_ms"] = list(st.stage_ms)
    stats["itr_ms"] = itr[:min(it, itr_cap)].copy()
    stats["r_trace"] = r_tr[:min(it, itr_cap)].astype(np.float32)
    return dict(labels=labels.cpu().numpy(), X=final.cpu().numpy(), stats=stats, code=rc)


def simple_labels(X, eps: float):
    """Labels of GPU_synCluster + rename_cluster (simple_GPU_SynC.cu:70-102) on the GPU."""
    D = _dev_f32(X)
    n, d = D.shape
    labels = torch.empty(n, dtype=torch.int32, device=D.device)
    k = ctypes.c_int32()
    nat.check(nat.lib().egg_sync_simple_labels_device(nat.ptr(D), n, d, float(eps),
                                                      nat.ptr(labels), ctypes.byref(k),
                                                      _stream()), "egg_sync_simple_labels_device")
    return labels.cpu().numpy(), k.value


def rc(X, eps: float) -> float:
    D = _dev_f32(X)
    out = ctypes.c_double()
    nat.check(nat.lib().egg_sync_rc_device(nat.ptr(D), D.shape[0], D.shape[1], float(eps),
                                           ctypes.byref(out), _stream()), "egg_sync_rc_device")
    return out.value

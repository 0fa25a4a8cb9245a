"""C-ABI boundary checks that need no GPU: the library loads, exports every
function include/egg_sync.h declares, the ctypes struct layouts match the C
header, host-side geometry equals the oracle's, errors come back as codes."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "egg_sync.h")


def _declared_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(egg_(?:sync|mlp)_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    from egg_sync import _native as nat
    L = nat.lib()
    declared = _declared_functions()
    assert len(declared) >= 10
    assert sorted(declared) == sorted(nat.EXPORTED)
    for name in declared:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", nat.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}\b", out), name


def test_struct_layout_matches_header(tmp_path):
    from egg_sync import _native as nat
    import ctypes
    prog = tmp_path / "layout.c"
    fields_o = [f for f, _ in nat.Options._fields_]
    fields_s = [f for f, _ in nat.Stats._fields_]
    body = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void){',
            'printf("%zu %zu\\n", sizeof(egg_sync_options), sizeof(egg_sync_stats));']
    body += [f'printf("%zu\\n", offsetof(egg_sync_options, {f}));' for f in fields_o]
    body += [f'printf("%zu\\n", offsetof(egg_sync_stats, {f}));' for f in fields_s]
    body += ['return 0;}']
    prog.write_text("\n".join(body))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(prog), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(nat.Options) and vals[1] == ctypes.sizeof(nat.Stats)
    offs = vals[2:]
    exp = [getattr(nat.Options, f).offset for f in fields_o] + \
          [getattr(nat.Stats, f).offset for f in fields_s]
    assert offs == exp


@pytest.mark.parametrize("n,d,eps", [(2000, 2, 0.05), (1_000_000, 3, 0.02), (250_000, 8, 0.1),
                                     (4_000_000, 3, 0.01), (100, 3, 0.02), (1, 1, 0.5),
                                     (5000, 4, 0.06), (10, 2, 3.0), (200_000, 8, 0.01)])
def test_geometry_matches_oracle(O, n, d, eps):
    from egg_sync import _native as nat
    g = nat.geometry(n, d, eps)
    o = O.geometry(n, d, eps)
    assert (g["w"], g["od"], g["O"], g["R"]) == (o.w, o.od, o.O, o.R)
    assert np.float32(g["cs"]) == np.float32(o.cs)


def test_key_width_for_fine_high_dim_grid():
    from egg_sync import _native as nat
    g = nat.geometry(200_000, 8, 0.01)      # w = 565, od = 1: 10 + 7*10 bits
    assert g["key_bits"] == 80
    g = nat.geometry(1_000_000, 3, 0.02)    # 15 + 8 bits: a single 64-bit key word
    assert g["key_bits"] == 23


def test_errors_are_codes_not_exits():
    from egg_sync import _native as nat
    L = nat.lib()
    import ctypes
    out = (ctypes.c_int64 * 5)()
    cs = ctypes.c_float()
    assert L.egg_sync_geometry(10, 0, 0.1, out, ctypes.byref(cs)) == nat.E_ARG
    assert L.egg_sync_geometry(10, 21, 0.1, out, ctypes.byref(cs)) == nat.E_DIM
    assert L.egg_sync_geometry(10, 2, 0.0, out, ctypes.byref(cs)) == nat.E_ARG
    assert L.egg_sync_geometry(10, 2, float("nan"), out, ctypes.byref(cs)) == nat.E_ARG
    assert L.egg_sync_geometry(0, 2, 0.1, out, ctypes.byref(cs)) == nat.E_ARG
    # d=12, eps=1e-6: 12 coordinate fields of 23 bits = 276 bits > 256
    assert L.egg_sync_geometry(10, 12, 1e-6, out, ctypes.byref(cs)) == nat.E_KEY
    assert L.egg_sync_geometry(10, 8, 1e-6, out, ctypes.byref(cs)) == nat.OK
    assert out[4] == 8 * 23
    assert "[0, 1]" in nat.error_string(nat.E_RANGE)
    assert L.egg_sync_solve_ex(None, 5, 2, 0.1, None, None, None, None) == nat.E_ARG


def test_torch_extension_exposes_reference_names():
    from egg_sync import algorithms
    ext = algorithms._ext()
    for name in ("EGG_SynC_wrapper", "DynamicalClustering_wrapper",
                 "DynamicalClustering_parallel_wrapper", "GPU_DynamicalClustering_wrapper"):
        assert hasattr(ext, name)
    for name in ("EGG_SynC", "SynC", "SynC_parallel", "DynamicalClustering",
                 "DynamicalClustering_parallel", "GPU_DynamicalClustering", "simple_GPU_SynC"):
        assert callable(getattr(algorithms, name))


def test_no_oracle_in_product_path():
    pkg = os.path.join(ROOT, "egg-sync_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                txt = open(os.path.join(dirpath, f), errors="ignore").read()
                assert not re.search(r"^\s*(import|from)\s+oracle\b", txt, re.M), f
                assert "liboracle" not in txt and "egg_oracle" not in txt, f
                assert "sync_ref" not in txt and "_ref/" not in txt, f


def test_autotune_signatures_and_cache(tmp_path, monkeypatch):
    from egg_sync import autotune
    monkeypatch.setenv("EGG_SYNC_TUNE_CACHE", str(tmp_path / "t.json"))
    monkeypatch.setattr(autotune, "_device_key", lambda: "dev")
    assert autotune.plain_signature(1_000_000, 3, 0.02) == "plain|n2^20|d3|eps0.02"
    sig = autotune.batch_signature([(0, 1000, 2, 0.1), (0, 1000, 2, 0.05), (2000, 4000, 8, 0.1)])
    assert sig == "batch|k2^2|pts2^13|d2,8"
    assert autotune.cached(sig) is None
    autotune._store(sig, {"streams": 2}, [])
    assert autotune.cached(sig) == {"streams": 2}


def test_small_angle_sincos_error_bound(tmp_path):
    """sincos_small (egg_common.h), compiled for the host by hipcc, against FP64
    sin / cos: at most 1.5 ulp on [-1, 1] (and the library functions beyond)."""
    import shutil
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = tmp_path / "sc.cpp"
    src.write_text(r'''
#include <cmath>
#include <cstdio>
#include "%s"
int main() {
    double worst_s = 0, worst_c = 0;
    for (int i = -2000000; i <= 2000000; i++) {
        const float v = i * 5e-7f;
        float s, c;
        egg::sincos_small(v, s, c);
        const double es = std::fabs((double)s - std::sin((double)v)) /
                          std::ldexp(1.0, std::ilogb(std::fabs(std::sin((double)v)) + 1e-300) - 23);
        const double ec = std::fabs((double)c - std::cos((double)v)) /
                          std::ldexp(1.0, std::ilogb(std::cos((double)v)) - 23);
        if (v != 0.0f && es > worst_s) worst_s = es;
        if (ec > worst_c) worst_c = ec;
    }
    float s, c;
    egg::sincos_small(3.0f, s, c);
    std::printf("%%.3f %%.3f %%d\n", worst_s, worst_c,
                (int)(s == std::sin(3.0f) && c == std::cos(3.0f)));
    return 0;
}
''' % os.path.join(ROOT, "egg-sync_amd", "csrc", "egg_common.h"))
    exe = tmp_path / "sc"
    subprocess.run([hipcc, "-O2", "-ffp-contract=off", "-std=c++17", str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    ws, wc, big = subprocess.run([str(exe)], capture_output=True, text=True,
                                 check=True).stdout.split()
    assert float(ws) <= 1.5 and float(wc) <= 1.5, (ws, wc)
    assert big == "1"


def test_partition_metrics():
    from egg_sync.metrics import disagreement, same_partition
    a = np.array([0, 0, 1, 1, -1, 2, 2])
    assert same_partition(a, np.array([3, 3, 0, 0, -1, 9, 9]))
    assert not same_partition(a, np.array([3, 3, 0, 0, 4, 9, 9]))      # noise differs
    b = np.array([5, 5, 7, 7, -1, 7, 7])                                # two clusters merged
    assert not same_partition(a, b)
    assert disagreement(a, b) == {"noise_differs": 0, "unmatched": 2, "clusters_a": 3,
                                  "clusters_b": 2}
    assert disagreement(a, a)["unmatched"] == 0


def test_experiment_data_recipe():
    from egg_sync import experiments as E
    X = E.synthetic_gauss(5000, 3, 5, 5.0, seed=1)
    assert X.shape == (5000, 3) and X.dtype == np.float32
    assert X.min() == 0.0 and X.max() == 1.0
    assert np.array_equal(X, E.synthetic_gauss(5000, 3, 5, 5.0, seed=1))
    assert not np.array_equal(X, E.synthetic_gauss(5000, 3, 5, 5.0, seed=2))
    assert E.STAGES == ["Allocating", "Build structure", "Update", "Extra check", "Clustering",
                        "Free Memory"]
    assert set(E.SWEEPS) == {"n", "d", "std", "cl", "eps"}


def test_dataset_loaders_read_the_reference_files():
    """The loaders read the reference's own data files (python/data.py:46-97)."""
    from egg_sync import experiments as E
    data_dir = "/root/reference/data"
    if not os.path.isdir(data_dir):
        pytest.skip("reference data files not present")
    for name, (fname, _) in E.DATASETS.items():
        if not os.path.exists(os.path.join(data_dir, fname)):
            continue
        X = E.load_dataset(name, data_dir)
        assert X.ndim == 2 and X.shape[0] > 100 and 0.0 <= X.min() and X.max() <= 1.0, name


def test_real_fixture_inputs_are_the_loaders_output():
    """tests/golden/real_oracle.npz holds the loaders' output for the reference's files."""
    from egg_sync import experiments as E
    from conftest import golden
    data_dir = "/root/reference/data"
    if not os.path.isdir(data_dir):
        pytest.skip("reference data files not present")
    g = golden("real_oracle")
    for name in ("banknote", "Yeast", "Wilt", "CCPP", "pendigits"):
        assert np.array_equal(E.load_dataset(name, data_dir), g[f"{name}_X"]), name
    assert g["pendigits_X"].shape[1] == 17

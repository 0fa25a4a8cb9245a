"""The stop-rule pin catches the plausible mistakes in the oracle.

Each case plants one mistake in a copy of the CPU oracle's sources (built
into a temporary directory, never into oracle/), loads that copy and runs
the named checks of test_stop_rule.py against it: each must fail.  The
mistakes are the ones a restatement of EGG_SynC.cu:583-709 could plausibly
make -- a strict compare at eps/2, the own cell dropped from `half`, the
border test in float, eps/2-neighbours counted in partial cells, the
border branch of the synced rule dropped, a wrong border band.
"""
import importlib.util
import os
import shutil
import subprocess

import pytest

import test_stop_rule as T

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

MUTATIONS = {
    "strict_half_compare": (
        "egg_oracle.c", "#define LE_HALF(s) ((s) <= geo->thr_half)",
        "#define LE_HALF(s) ((s) < geo->thr_half)",
        ["test_hand_squared_distance_at_threshold"]),
    "own_cell_dropped": (
        "egg_core.inc", "                        half += s1 - s0;\n",
        "                        (void)0;\n",
        ["test_hand_exact_half_eps_on_border", "test_flags_golden_inputs"]),
    "border_in_float": (
        "egg_oracle.c", "int idp = (int)((xd + 0.00001) / (double)cs);",
        "int idp = (int)((x + 0.00001f) / cs);",
        ["test_hand_border_test_in_double"]),
    "half_in_partial_cells": (
        "egg_core.inc", "            } else if (included) {\n",
        "            } else if (included) {\n"
        "                if (border) for (int32_t j = s0; j < s1; j++)\n"
        "                    if (LE_HALF(F(dist2)(x, &X[(int64_t)G->order[j] * d], d))) half++;\n",
        ["test_hand_half_neighbour_in_partial_cell"]),
    "border_branch_dropped": (
        "egg_core.inc", "int synced = !((!border && N != own) || (border && N != half));",
        "int synced = !(N != own);",
        ["test_hand_exact_half_eps_on_border", "test_two_points_iteration_count_closed_form"]),
    "border_band_1e-4": (
        "egg_oracle.c", "0.00001", "0.0001",
        ["test_hand_border_band", "test_flags_golden_inputs"]),
}


def _mutated_oracle(tmp_path, fname, old, new):
    d = tmp_path / "oracle"
    d.mkdir()
    for f in ("egg_oracle.c", "egg_core.inc", "egg_oracle.h", "oracle.py"):
        shutil.copy(os.path.join(ORACLE, f), d / f)
    src = (d / fname).read_text()
    assert old in src, "mutation anchor not found: the oracle changed"
    (d / fname).write_text(src.replace(old, new))
    subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    "-std=c11", "-shared", "-o", str(d / "liboracle.so"), str(d / "egg_oracle.c"),
                    "-lm"], check=True, capture_output=True)
    spec = importlib.util.spec_from_file_location(f"oracle_mut_{tmp_path.name}", d / "oracle.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert mod.LIB_PATH == str(d / "liboracle.so")
    return mod


def _cases(fn):
    for m in getattr(fn, "pytestmark", []):
        if m.name == "parametrize":
            single = "," not in m.args[0]
            return [(c,) if single or not isinstance(c, tuple) else c for c in m.args[1]]
    return [()]


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistake_is_caught(tmp_path, name):
    fname, old, new, checks = MUTATIONS[name]
    M = _mutated_oracle(tmp_path, fname, old, new)
    for check in checks:
        fn = getattr(T, check)
        failed = False
        for case in _cases(fn):
            try:
                fn(M, *case)
            except AssertionError:
                failed = True
                break
        assert failed, f"{check} passes against the oracle with '{name}' planted"

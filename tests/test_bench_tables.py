"""bench.py's HBM-traffic and VALU tables against the rocprofv3 summaries
they are read from (profiles/*.md): every entry names its source file, and
that file's derived bytes per step and its per-launch SQ_INSTS_VALU of the
sweep kernels must equal the table's figures to the digits the table keeps.
"""
import importlib.util
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_tables", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _read(rel):
    with open(os.path.join(ROOT, rel)) as f:
        return f.read()


def _sweep_valu(text):
    """SQ_INSTS_VALU per launch of the sweep kernels in a summary's counter
    table: the two-direction k_sweep_xy, or else the x and y k_sweep launches
    of the step (the pair with the most wavefronts)."""
    rows = {}
    for ln in text.splitlines():
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cells) == 9 and cells[0].startswith("`k_sweep"):
            rows[cells[0].strip("`")] = (int(cells[6]), float(cells[7]))
    xy = [v for k, v in rows.items() if k.startswith("k_sweep_xy")]
    if xy:
        return xy[0][1]
    single = sorted(rows.values(), key=lambda v: -v[0])
    return single[0][1] + single[1][1] if len(single) == 2 else None


def test_traffic_table_matches_profiles(bench):
    assert bench.TRAFFIC
    for key, entry in bench.TRAFFIC.items():
        text = _read(entry["source"])
        m = re.search(r"= \*\*([0-9.]+) MB\*\*", text)
        assert m, entry["source"]
        assert abs(float(m.group(1)) * 1e6 - entry["bytes_per_step"]) <= 1e5, (key, m.group(1))
        # the summary was made on the tiling and in the layout the table names
        cx, cy, wx, wy = entry["tiling"]
        assert f"--chunks {cx} {cy} {wx} {wy}" in text, key
        assert ("--layout blocked" in text) == (key[4] == "blocked"), key


def test_valu_table_matches_profiles(bench):
    assert bench.VALU
    for key, entry in bench.VALU.items():
        text = _read(entry["source"])
        got = _sweep_valu(text)
        assert got is not None, entry["source"]
        assert abs(got - entry["wave_instr_per_launch"]) <= 5e-4 * got, (key, got, entry["wave_instr_per_launch"])
        cx, cy, wx, wy = entry["tiling"]
        assert f"--chunks {cx} {cy} {wx} {wy}" in text, key
        assert ("--layout blocked" in text) == (key[4] == "blocked"), key


def test_profile_commands_name_their_workload():
    """A summary's commands select the workload its title names (the bench
    default is the 4096^2 four-quadrant problem)."""
    for name, flag in (("kh8192", "--workload kh8192"), ("dmr1920x480", "--workload dmr")):
        for prec in ("fp64", "fp32split", "fp64_blocked"):
            text = _read(f"profiles/{name}_{prec}.md")
            cmds = [ln for ln in text.splitlines() if ln.startswith("cd /tmp") and "bench.py" in ln]
            assert cmds and all(flag in ln for ln in cmds), (name, prec)
            assert all(("--precision fp32split" in ln) == (prec == "fp32split") for ln in cmds), (name, prec)
            assert all(("--layout blocked" in ln) == prec.endswith("blocked") for ln in cmds), (name, prec)


def test_batch_table_matches_profile(bench):
    """The persistent batch's traffic and VALU per step equal its summary's."""
    assert bench.BATCH_TRAFFIC
    for structure, entry in bench.BATCH_TRAFFIC.items():
        chunk = entry["chunk"]
        text = _read(entry["source"])
        m = re.search(r"HBM traffic = \*\*([0-9.]+) MB\*\*", text)
        assert m, entry["source"]
        assert abs(float(m.group(1)) * 1e6 - entry["bytes_per_step"]) <= 1e5, (structure, chunk)
        m = re.search(r"VALU: ([0-9.e+]+) wave instructions per step", text)
        assert m and abs(float(m.group(1)) - entry["wave_instr_per_step"]) <= 1e-3 * entry["wave_instr_per_step"]
        assert f"--chunks {chunk} {chunk} 1 1" in text and "--workload batch512" in text
        # the numbers come from the persistent step's launches
        assert (structure == "persistent") == ("k_batch_persist" in text)
        r = {"value": 2.4e9, "cells": 24326144, "ms_per_step": 10.0, "chunk": 64,
             "step_structure": "persistent" if structure == "persistent" else "launch per phase"}
        roof = bench.batch_roofline(r)
        assert roof["traffic"] == entry["bytes_per_step"] and roof["traffic_chunk"] == chunk

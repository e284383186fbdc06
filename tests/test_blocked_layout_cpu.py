"""The blocked state layout's index map (blk_row / blk_col in
weno-cfd-solver_amd/csrc/wenox_sweep.h, with the plane geometry nb, Pb of
make_sweep_plan): every padded cell of an (nx+6) x (ny+6) plane gets its own
element below Pb, the 16 cells of a 4 x 4 block share one 128 B line, and a
row or column of 4 cells of a block lies in two different 32 B sectors, two
cells in each.  Checked by a host program built against the header with
hipcc (no GPU needed)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

PROGRAM = r'''
#include <cstdio>
#include <vector>
#include "wenox_sweep.h"
using namespace wenox;
int main() {
    const long sizes[][2] = {{3, 3}, {16, 16}, {45, 37}, {130, 96}, {1920, 480}, {257, 1029}};
    for (auto& sz : sizes) {
        const long nx = sz[0], ny = sz[1], M = nx + 6, N = ny + 6;
        const long nb = (ny + 6 + 3) / 4, Pb = (nx + 6 + 3) / 4 * nb * 16;
        std::vector<char> seen(Pb, 0);
        for (long i = 0; i < M; ++i)
            for (long j = 0; j < N; ++j) {
                const long e = blk_row(i, nb) + blk_col(j);
                if (e < 0 || e >= Pb) { std::printf("out of range %ld %ld\n", i, j); return 1; }
                if (seen[e]) { std::printf("collision %ld %ld\n", i, j); return 1; }
                seen[e] = 1;
                // the block's 16 cells in one 16-element (128 B) line
                const long base = blk_row(i & ~3L, nb) + blk_col(j & ~3L);
                if (base % 16 != 0 || e / 16 != base / 16) { std::printf("line %ld %ld\n", i, j); return 1; }
            }
        // 4 cells along either axis: two pairs of adjacent elements (16 B)
        // or two pairs one element apart, in two different 4-element sectors
        for (long a = 0; a < 8; a += 4)
            for (long b = 0; b < 8; ++b)
                for (int axis = 0; axis < 2; ++axis) {
                    long e[4];
                    for (int t = 0; t < 4; ++t) {
                        const long i = axis ? a + t : b, j = axis ? b : a + t;
                        e[t] = blk_row(i, nb) + blk_col(j);
                    }
                    if (e[0] / 4 != e[1] / 4 || e[2] / 4 != e[3] / 4 || e[0] / 4 == e[2] / 4) {
                        std::printf("sectors axis %d\n", axis);
                        return 1;
                    }
                }
    }
    std::printf("ok\n");
    return 0;
}
'''


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_blocked_index_map(tmp_path):
    src = tmp_path / "blk.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "blk"
    inc = os.path.join(ROOT, "weno-cfd-solver_amd", "csrc")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-I", inc, str(src), "-o", str(exe)], check=True,
                   capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout

"""Tile order for the 8 XCDs (vision_amd/csrc/xcd.h): xcd_order(b, n) must
be a bijection on [0, n) for every grid size, give workgroup b (dispatched
to XCD b mod 8) a tile in XCD (b mod 8)'s range, and give each XCD one
contiguous range walked in dispatch order; otherwise a kernel would skip or
repeat tiles.  Checked exhaustively for n < 4100 with the header compiled
on the host."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

CSRC = Path(__file__).resolve().parents[1] / "vision_amd" / "csrc"

PROG = r'''
#include <cstdio>
#include "xcd.h"
int main() {
    for (unsigned n = 1; n < 4100; ++n)
        for (unsigned b = 0; b < n; ++b) std::printf("%u\n", vs::xcd_order(b, n));
    return 0;
}
'''


@pytest.fixture(scope="module")
def orders(tmp_path_factory):
    d = tmp_path_factory.mktemp("xcd")
    src, exe = d / "xcd.cpp", d / "xcd"
    src.write_text(PROG)
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I", str(CSRC),
                    str(src), "-o", str(exe)], check=True, capture_output=True)
    out = np.array(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split(), np.int64)
    res, pos = {}, 0
    for n in range(1, 4100):
        res[n] = out[pos:pos + n]
        pos += n
    assert pos == out.size
    return res


def test_bijection_for_every_grid_size(orders):
    for n, t in orders.items():
        assert np.array_equal(np.sort(t), np.arange(n)), n


def test_each_xcd_walks_one_contiguous_range_in_dispatch_order(orders):
    for n in (1, 7, 8, 9, 63, 64, 65, 1000, 4099):
        t = orders[n]
        starts = []
        for x in range(8):
            mine = t[x::8]  # workgroups x, x + 8, ... run on XCD x, in this order
            if mine.size:
                assert np.array_equal(mine, mine[0] + np.arange(mine.size)), (n, x)
                starts.append((mine[0], mine.size))
        # ranges follow one another: XCD 0 first, ..., sizes differ by at most one
        edge = 0
        for s0, sz in starts:
            assert s0 == edge
            edge += sz
        sizes = [sz for _, sz in starts]
        assert max(sizes) - min(sizes) <= 1

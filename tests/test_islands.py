"""The island building's broadphase (beamlab_amd/csrc/bl_islands.h, host C++): the pairs of overlapping body boxes
from the 3-D grid against the plain all-pairs test, as sets, on random scenes - small boxes in clusters, a few large
ones, boxes touching at a face, and boxes that are not finite. Compiled with g++ on the host; no GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "beamlab_amd", "csrc")

HARNESS = r"""
#include "bl_islands.h"
#include <cstdio>
#include <vector>
int main() {
    size_t n;
    if (std::fscanf(stdin, "%zu", &n) != 1) return 2;
    std::vector<float> bx(6 * n);
    for (auto& v : bx) if (std::fscanf(stdin, "%f", &v) != 1) return 2;
    // (a first call on half the boxes leaves its arrays behind: the second, on all of them, reuses them)
    blg::PairScratch ws;
    blg::overlapping_pairs(bx.data(), n / 2, ws, [](uint32_t, uint32_t) {});
    blg::overlapping_pairs(bx.data(), n, ws, [](uint32_t i, uint32_t j) { std::printf("%u %u\n", i, j); });
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("islands")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    return str(exe)


def brute(bx):
    out = set()
    n = len(bx)
    for i in range(n):
        p = bx[i]
        for j in range(i + 1, n):
            q = bx[j]
            if p[0] <= q[3] and p[3] >= q[0] and p[1] <= q[4] and p[4] >= q[1] and p[2] <= q[5] and p[5] >= q[2]:
                out.add((i, j))
    return out


def run(exe, bx):
    txt = f"{len(bx)}\n" + "\n".join(" ".join(repr(float(v)) for v in b) for b in bx) + "\n"
    r = subprocess.run([exe], input=txt, capture_output=True, text=True, check=True)
    pairs = [tuple(map(int, l.split())) for l in r.stdout.split("\n") if l]
    assert len(pairs) == len(set(pairs)), "a pair reported twice"
    return set(pairs)


def scene(seed, n):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-20, 20, (n, 3)).astype(np.float32)
    c[: n // 2] = rng.normal(0, 2, (n // 2, 3)).astype(np.float32)  # a cluster
    h = rng.uniform(0.05, 1.0, (n, 3)).astype(np.float32)
    h[:3] = [30.0, 1.0, 4.0]  # large bodies (the odd path for one of them)
    h[3] = [200.0, 0.5, 200.0]
    bx = np.concatenate([c - h, c + h], 1).astype(np.float32)
    bx[5, 3] = bx[6, 0]  # touching at a face (closed intervals)
    bx[5, 1], bx[5, 4], bx[5, 2], bx[5, 5] = bx[6, 1], bx[6, 4], bx[6, 2], bx[6, 5]
    bx[7, 0] = np.nan
    bx[8, 3] = np.inf
    bx[9, 4] = -np.inf
    return bx


@pytest.mark.parametrize("seed,n", [(1, 40), (2, 300), (3, 700)])
def test_grid_pairs_equal_all_pairs(harness, seed, n):
    bx = scene(seed, n)
    got = run(harness, bx)
    want = brute(bx.astype(np.float64))
    assert got == want
    assert len(want) > n // 4


def test_sparse_wide_and_flat_scenes(harness):
    """A 2 km arena of small boxes (the grid's cells enlarged to bound their count) and a flat layer (one cell in y)."""
    rng = np.random.default_rng(5)
    c = rng.uniform(-1000, 1000, (400, 3)).astype(np.float32)
    c[:200] = rng.uniform(-3, 3, (200, 3)).astype(np.float32)
    h = rng.uniform(0.1, 0.8, (400, 3)).astype(np.float32)
    bx = np.concatenate([c - h, c + h], 1).astype(np.float32)
    assert run(harness, bx) == brute(bx.astype(np.float64))
    c[:, 1] = 0.0
    bx = np.concatenate([c - h, c + h], 1).astype(np.float32)
    assert run(harness, bx) == brute(bx.astype(np.float64))


def test_empty_and_single(harness):
    assert run(harness, np.zeros((0, 6), np.float32)) == set()
    assert run(harness, np.array([[0, 0, 0, 1, 1, 1]], np.float32)) == set()
    same = np.array([[0, 0, 0, 1, 1, 1]] * 5, np.float32)
    assert run(harness, same) == {(i, j) for i in range(5) for j in range(i + 1, 5)}

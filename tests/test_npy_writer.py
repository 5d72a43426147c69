"""The header-only .npy writer (include/ssim_npy.hpp) that the C++ front end uses for the per-step frames of the
reference drivers (deform_<step>.npy, tests/fem_cu/main.cu:86-92, loaded by tests/fem_time_step/visualize.py:9-12).

Pin: for every shape, including empty arrays and a header that needs no padding beyond its newline, the file is
byte-identical to numpy.save's output for the same float64 C-order array, and numpy.load returns the values
bit for bit."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "ssim_npy.hpp"
int main(int argc, char** argv) {
  const size_t rows = std::strtoull(argv[2], nullptr, 10), cols = std::strtoull(argv[3], nullptr, 10);
  std::vector<double> v(rows * cols);
  FILE* f = std::fopen(argv[4], "rb");
  if (!v.empty() && std::fread(v.data(), sizeof(double), v.size(), f) != v.size()) return 3;
  std::fclose(f);
  ssim::hip::io::write_npy(argv[1], v.data(), rows, cols);
  return 0;
}
"""


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    d = tmp_path_factory.mktemp("npy")
    src, exe = d / "w.cpp", d / "w"
    src.write_text(PROGRAM)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    return exe


@pytest.mark.parametrize("shape", [(0, 3), (1, 3), (5, 3), (1000, 3), (7, 11), (123456, 3), (10 ** 9, 0)])
def test_matches_numpy_save(writer, tmp_path, shape):
    rng = np.random.default_rng(sum(shape) % 1000)
    a = rng.standard_normal(shape) * 10.0 ** rng.integers(-300, 300, size=shape)
    a.reshape(-1)[:3] = [np.inf, -0.0, np.nan][:a.size]
    raw = tmp_path / "in.bin"
    a.tofile(raw)
    mine, ref = tmp_path / "mine.npy", tmp_path / "ref.npy"
    subprocess.run([str(writer), str(mine), str(shape[0]), str(shape[1]), str(raw)], check=True)
    np.save(ref, a)
    assert mine.read_bytes() == ref.read_bytes()
    b = np.load(mine)
    assert b.shape == shape and b.dtype == np.float64
    assert np.array_equal(b.view(np.uint64), a.view(np.uint64))


def test_header_lengths_around_the_alignment(writer, tmp_path):
    """Shapes whose digit counts move the dictionary across a 64-byte boundary."""
    for rows in (9, 99, 999, 9999, 10 ** 12):
        for cols in (0, 3, 10 ** 6):
            if rows * cols > 10 ** 6:
                continue
            a = np.zeros((rows, cols))
            raw, mine, ref = tmp_path / "in.bin", tmp_path / "mine.npy", tmp_path / "ref.npy"
            a.tofile(raw)
            subprocess.run([str(writer), str(mine), str(rows), str(cols), str(raw)], check=True)
            np.save(ref, a)
            assert mine.read_bytes() == ref.read_bytes(), (rows, cols)

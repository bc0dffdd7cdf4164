"""The scalar math of the kernels (lumina_amd/csrc/lh_math.h), compiled for
the host with g++ and checked against numpy/mpmath: the FP32 and FP64
sin/cos of the pair phase, and the contraction-free |d|^2 of the exclusion
test."""
import os
import subprocess

import mpmath as mp
import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <cstdio>
#include <cstdlib>
#include "lh_math.h"
int main(int argc, char** argv) {
  // read x values from stdin, print f32 sin cos, f64 sin cos
  double x;
  while (std::scanf("%lf", &x) == 1) {
    float s32, c32; double s64, c64;
    lh::sincos_f32((float)x, s32, c32);
    lh::sincos_f64(x, s64, c64);
    std::printf("%.9e %.9e %.17e %.17e\n", (double)s32, (double)c32, s64, c64);
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("math")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-Wno-unknown-pragmas", "-I",
                    os.path.join(ROOT, "lumina_amd", "csrc"), str(src), "-o", str(exe)], check=True)
    return exe


def run(exe, xs):
    out = subprocess.run([str(exe)], input="\n".join(repr(float(x)) for x in xs),
                         capture_output=True, text=True, check=True).stdout.split()
    return np.array(out, dtype=np.float64).reshape(-1, 4)


def test_sincos_accuracy(harness):
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.linspace(0, 12.0, 2001), rng.uniform(0, 600.0, 2000),
                         [0.0, 1e-8, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 2, 2 * np.pi]])
    r = run(harness, xs)
    mp.mp.dps = 30
    s_ref = np.array([float(mp.sin(mp.mpf(float(x)))) for x in xs])
    c_ref = np.array([float(mp.cos(mp.mpf(float(x)))) for x in xs])
    # FP32 (input rounded to FP32): absolute error, dominated by rounding of x
    xs32 = xs.astype(np.float32).astype(np.float64)
    s32_ref = np.sin(xs32)
    c32_ref = np.cos(xs32)
    assert np.max(np.abs(r[:, 0] - s32_ref)) < 4e-7
    assert np.max(np.abs(r[:, 1] - c32_ref)) < 4e-7
    # FP64: within a few ulp
    assert np.max(np.abs(r[:, 2] - s_ref)) < 5e-16 * max(1.0, xs.max() / 100)
    assert np.max(np.abs(r[:, 3] - c_ref)) < 5e-16 * max(1.0, xs.max() / 100)
    small = xs < 20
    assert np.max(np.abs(r[small, 2] - s_ref[small])) < 3e-16
    assert np.max(np.abs(r[small, 3] - c_ref[small])) < 3e-16


def test_sincos_quadrant_signs(harness):
    # every quadrant, both sides of each boundary
    xs = np.array([q * np.pi / 2 + d for q in range(0, 9) for d in (-0.3, 0.3)])
    xs = xs[xs >= 0]
    r = run(harness, xs)
    assert np.all(np.sign(r[:, 2]) == np.sign(np.sin(xs)))
    assert np.all(np.sign(r[:, 3]) == np.sign(np.cos(xs)))
    assert np.all(np.sign(r[:, 0]) == np.sign(np.sin(xs)))
    assert np.all(np.sign(r[:, 1]) == np.sign(np.cos(xs)))

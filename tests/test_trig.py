"""The kernels' FP64 sine and cosine (acoustic_reliefs_amd/csrc/trig.h, Cody-Waite reduction + Taylor
polynomials), compiled for the host from the same header, against the C library: within 1 ulp for
|x| <= pi/4 and within 2.3e-16 absolute up to |x| = 1e9 (the domain the entry points enforce through
the wavenumber limit), exact quadrant signs at multiples of pi/2."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "acoustic_reliefs_amd", "csrc")

SRC = r'''
#include "trig.h"
extern "C" void trig_batch(const double *x, int n, double *s, double *c) {
    for (int i = 0; i < n; i++) arh::trig_sincos(x[i], s + i, c + i);
}
'''


@pytest.fixture(scope="module")
def trig(tmp_path_factory):
    d = tmp_path_factory.mktemp("trig")
    src, so = d / "t.cpp", d / "t.so"
    src.write_text(SRC)
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-ffp-contract=off", "-I" + CSRC, str(src), "-o", str(so)],
                   check=True)
    lib = ctypes.CDLL(str(so))
    P = ctypes.POINTER(ctypes.c_double)

    def run(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        s, c = np.empty_like(x), np.empty_like(x)
        lib.trig_batch(x.ctypes.data_as(P), len(x), s.ctypes.data_as(P), c.ctypes.data_as(P))
        return s, c
    return run


def test_reduced_range_ulp(trig):
    x = np.random.default_rng(0).uniform(-np.pi / 4, np.pi / 4, 500000)
    s, c = trig(x)
    assert np.max(np.abs(s - np.sin(x)) / np.spacing(np.abs(np.sin(x)))) <= 1.0
    assert np.max(np.abs(c - np.cos(x)) / np.spacing(np.abs(np.cos(x)))) <= 1.0


@pytest.mark.parametrize("hi", [200.0, 1e4, 1e6, 1e9])
def test_absolute_error(trig, hi):
    x = np.random.default_rng(int(np.log10(hi))).uniform(-hi, hi, 400000)
    s, c = trig(x)
    assert np.max(np.abs(s - np.sin(x))) <= 2.3e-16
    assert np.max(np.abs(c - np.cos(x))) <= 2.3e-16


def test_quadrants_and_special_points(trig):
    k = np.arange(-8, 9)
    s, c = trig(k * (np.pi / 2))
    np.testing.assert_allclose(s, np.sin(k * np.pi / 2), rtol=0, atol=1e-15)
    np.testing.assert_allclose(c, np.cos(k * np.pi / 2), rtol=0, atol=1e-15)
    # the signs of the four quadrants at their midpoints
    m = (k + 0.5) * (np.pi / 2)
    s, c = trig(m)
    assert np.array_equal(np.sign(s), np.sign(np.sin(m))) and np.array_equal(np.sign(c), np.sign(np.cos(m)))
    s, c = trig(np.array([0.0, -0.0, 1e-300]))
    assert s[0] == 0.0 and c[0] == 1.0 and s[2] == 1e-300 and c[2] == 1.0
    # close to multiples of pi/2 the reduction keeps the small result accurate
    k = np.random.default_rng(5).integers(-10 ** 6, 10 ** 6, 20000)
    x = k * (np.pi / 2) + np.random.default_rng(6).uniform(-1e-9, 1e-9, len(k))
    s, c = trig(x)
    small_s, small_c = np.abs(np.sin(x)) < 1e-6, np.abs(np.cos(x)) < 1e-6
    assert small_s.any() and small_c.any()
    assert np.max(np.abs(s - np.sin(x))[small_s]) <= 1e-22
    assert np.max(np.abs(c - np.cos(x))[small_c]) <= 1e-22

"""FD coefficient tables and CPML profiles: bit-exact against the reference.

* Closed-form KATs of the float32 zero-sum second-derivative stencils follow
  the reference's test/test_laplacian_zero_sum.py:41-80 (c0 = 2 sum c exactly,
  order 4: 5/2, 4/3; order 6: 49/18, 3/2, -3/20; order 8: 205/72, 8/5, -1/5,
  8/315; the last weight closes the sum).
* tests/golden/tables.npz holds the reference's own tables and profiles
  (written by tests/golden/gen_acoustic2d.py from scalars.py / pml.py).
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

from conftest import GOLDEN

from sweep_amd import pml, scalars


@pytest.fixture(scope="module")
def tabs():
    return np.load(GOLDEN / "tables.npz")


@pytest.mark.parametrize("M", range(1, 9))
def test_second_derivative_zero_sum_exact(M):
    c0, c = scalars.second_derivative_coefficients_fp32(2 * M)
    assert Fraction(float(c0)) + 2 * sum(Fraction(float(v)) for v in c) == 0


def test_second_derivative_kats():
    f32 = np.float32
    c0, c = scalars.second_derivative_coefficients_fp32(4)
    assert c0 == f32(-5 / 2) and c[0] == f32(4 / 3)
    assert Fraction(float(c[1])) == -Fraction(float(c0)) / 2 - Fraction(float(c[0]))
    c0, c = scalars.second_derivative_coefficients_fp32(6)
    assert c0 == f32(-49 / 18) and c[0] == f32(3 / 2) and c[1] == f32(-3 / 20)
    assert abs(float(c[2]) - 1 / 90) < 6e-8
    c0, c = scalars.second_derivative_coefficients_fp32(8)
    assert c0 == f32(-205 / 72) and c[0] == f32(8 / 5) and c[1] == f32(-1 / 5) and c[2] == f32(8 / 315)
    assert abs(float(c[3]) + 1 / 560) < 6e-8


@pytest.mark.parametrize("M", range(1, 9))
def test_tables_match_reference(tabs, M):
    c0, c = scalars.second_derivative_coefficients_fp32(2 * M)
    assert np.float32(c0) == tabs[f"lap_c0_{M}"]
    assert np.array_equal(c, tabs[f"lap_c_{M}"])
    assert np.array_equal(scalars.fd_coefficients(1, 2 * M).astype(np.float32), tabs[f"grad_{M}"])
    assert np.array_equal(scalars.staggered_grid_coes(M), tabs[f"stag_{M}"])


@pytest.mark.parametrize("M", range(1, 5))
def test_static_constants_equal_tables(M):
    """The fixed-order forward weights the kernels hold equal the zero-sum
    tables (laplace.cuh:157-162 says so; checked, not assumed)."""
    c0, c = scalars.second_derivative_coefficients_fp32(2 * M)
    st = scalars.static_laplace_constants(M)
    assert np.array_equal(st, c)
    if M == 2:
        assert scalars.static_laplace_c0(M) == -c0


def test_staggered_closed_form():
    a = scalars.staggered_grid_coes(4)
    want = [1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168]
    assert np.allclose(a, want, rtol=0, atol=1e-7)


def test_first_derivative_closed_form():
    exact = {1: [1 / 2], 2: [2 / 3, -1 / 12], 3: [3 / 4, -3 / 20, 1 / 60], 4: [4 / 5, -1 / 5, 4 / 105, -1 / 280]}
    for M, want in exact.items():
        assert np.allclose(scalars.fd_coefficients(1, 2 * M), want, rtol=1e-12, atol=0)
        st = scalars.static_gradient_constants(M)
        assert np.allclose(st, np.float32(want), rtol=0, atol=1e-7)


GEO = {
    "r_384x1152": ("r", (384, 1152), 8.0, 8e-4, 4),
    "s_512x1024": ("s", (512, 1024), 5.0, 5e-4, 4),
    "r_256x512x512": ("r", (256, 512, 512), 10.0, 1e-3, 4),
    "r_96x128_o4": ("r", (96, 128), 10.0, 2e-3, 2),
}


@pytest.mark.parametrize("key", list(GEO))
def test_cpml_profiles_bit_exact(tabs, key):
    kind, shp, dh, dt, M = GEO[key]
    runtime = [s + 100 + 2 * M for s in shp]
    fn = pml.cpml_profiles_r if kind == "r" else pml.cpml_profiles_s
    prof = fn([50] * (2 * len(shp)), [M, M] * len(shp), dt, [dh] * len(shp), 4500.0, 25.0, tuple(runtime))
    for i, p in enumerate(prof):
        assert np.array_equal(np.asarray(p, dtype=np.float32), tabs[f"prof_{key}_{i}"]), (key, i)


def test_cpml_profile_invariants():
    """a = b = 0 off the band, 0 < a < 1 in it, |b| monotone into the pad,
    db non-zero one cell past the band (the leak cell)."""
    M, w = 4, 50
    n = 200 + 2 * w + 2 * M
    a, b, db = pml.cpml_profiles_r([w, w], [M, M], 1e-3, [10.0], 4500.0, 25.0, (n,))
    band = np.zeros(n, dtype=bool)
    band[:M + w] = True
    band[n - M - w:] = True
    assert np.all(a[~band] == 0) and np.all(b[~band] == 0)
    assert np.all((a[band] > 0) & (a[band] < 1))
    assert np.all(np.diff(np.abs(b[:M + w + 1])) <= 0)
    assert db[M + w] != 0 and np.all(db[M + w + 1:n - M - w - 1] == 0)

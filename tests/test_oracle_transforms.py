"""Pins of the oracle's transform layer against mathematics and numpy.

Each test fixes a routine of the oracle against something that does not come
from the oracle: numpy's FFT, numpy's Chebyshev-series algebra, closed forms
of Chebyshev polynomials, exact Clenshaw-Curtis quadrature.

Reference routines covered: fourier_fftw.F90:39-97 (forfft, invfft),
chebyshev_fftw.F90:88-427 (dcheb, chebtransform, chebinvtran, D1, D2),
integ.F90:36-76 (radInt), init.F90:120-140 (grid).
"""
from __future__ import annotations

import numpy as np
import numpy.polynomial.chebyshev as npc
import pytest

from conftest import rel

ETA = 0.35
RMIN = ETA / (1 - ETA)


# --------------------------------------------------------------------- DFT
@pytest.mark.parametrize("n", [8, 12, 30, 60, 64, 96, 128, 192, 250, 384, 768, 1024, 2304])
def test_fft_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(orc.fft(x, -1), np.fft.fft(x)) < 1e-14
    assert rel(orc.fft(x, +1), np.fft.ifft(x) * n) < 1e-14


# ------------------------------------------------------- azimuthal transforms
@pytest.mark.parametrize("Nm", [4, 16, 64, 128])
def test_forfft_closed_form(orc, Nm):
    """cos(k phi) -> f_k = 1/2, sin(k phi) -> f_k = -i/2, constant -> f_0."""
    Np = 3 * Nm
    phi = 2 * np.pi * np.arange(Np) / Np
    for k in (1, Nm // 2, Nm):
        f = orc.forfft(np.cos(k * phi) + 0.25, Nm)
        want = np.zeros(Nm + 1, complex)
        want[0] = 0.25
        want[k] += 0.5
        assert np.abs(f - want).max() < 1e-14
        f = orc.forfft(np.sin(k * phi), Nm)
        want = np.zeros(Nm + 1, complex)
        want[k] = -0.5j
        assert np.abs(f - want).max() < 1e-14


@pytest.mark.parametrize("Nm", [8, 32, 128])
def test_forfft_invfft_numpy_and_roundtrip(orc, Nm):
    Np = 3 * Nm
    rng = np.random.default_rng(Nm)
    x = rng.standard_normal(Np)
    f = orc.forfft(x, Nm)
    assert rel(f, np.fft.rfft(x)[: Nm + 1] / Np) < 1e-14
    # c2r of a band-limited spectrum: x_j = Re f0 + 2 Re sum f_m e^{+i m phi_j}
    g = rng.standard_normal(Nm + 1) + 1j * rng.standard_normal(Nm + 1)
    xs = orc.invfft(g, Np)
    phi = 2 * np.pi * np.arange(Np) / Np
    direct = g[0].real + 2 * np.real(np.exp(1j * np.outer(phi, np.arange(1, Nm + 1))) @ g[1:])
    assert rel(xs, direct) < 1e-13
    # Im(f_0) is ignored, as by an unnormalised complex-to-real transform
    g2 = g.copy()
    g2[0] = g[0].real + 5j
    assert np.array_equal(orc.invfft(g2, Np), xs)
    # round trip: forfft(invfft(g)) = g with Im(f_0) dropped
    g3 = g.copy()
    g3[0] = g3[0].real
    assert rel(orc.forfft(xs, Nm), g3) < 1e-13


# --------------------------------------------------------------- Chebyshev
def cheb_points(N):
    return np.cos((N - 1 - np.arange(N)) * np.pi / (N - 1))


@pytest.mark.parametrize("N", [5, 17, 33, 65, 97])
def test_dcheb_closed_forms(orc, N):
    t, D, D2, xp = orc.dcheb(N)
    x = cheb_points(N)
    assert np.abs(xp - x).max() < 1e-15
    j = np.arange(N)
    # T_j(x_i) = cos(j arccos x_i)
    assert np.abs(t - np.cos(np.outer(np.arccos(np.clip(x, -1, 1)), j))).max() < 1e-12
    # derivative columns against numpy's Chebyshev algebra (D = 2 d/dx, D2 = 4 d2/dx2)
    for jj in (0, 1, 2, N // 2, N - 1):
        c = np.zeros(N)
        c[jj] = 1.0
        d1 = 2 * npc.chebval(x, npc.chebder(c))
        d2 = 4 * npc.chebval(x, npc.chebder(c, 2))
        scale = max(1.0, np.abs(d2).max())
        assert np.abs(D[:, jj] - d1).max() <= 1e-12 * max(1.0, np.abs(d1).max())
        assert np.abs(D2[:, jj] - d2).max() <= 1e-11 * scale
    # T_n'(1) = n^2, T_n''(1) = n^2 (n^2 - 1) / 3
    n = j.astype(float)
    assert np.allclose(D[N - 1, :], 2 * n ** 2, rtol=1e-12, atol=1e-12)
    assert np.allclose(D2[N - 1, :], 4 * n ** 2 * (n ** 2 - 1) / 3, rtol=1e-11, atol=1e-9)


def series_to_ft(c):
    """numpy Chebyshev coefficients -> the reference's scaled coefficient
    vector: f(x) = C (ft_0/2 + sum_{k=1}^{N-2} ft_k T_k + ft_{N-1}/2 T_{N-1}),
    C = sqrt(2/(N-1))."""
    N = len(c)
    C = np.sqrt(2.0 / (N - 1))
    ft = np.asarray(c, dtype=complex) / C
    ft[0] *= 2
    ft[N - 1] *= 2
    return ft


@pytest.mark.parametrize("N", [5, 9, 17, 33, 65, 97, 129])
def test_chebtransform_known_series(orc, N):
    rng = np.random.default_rng(N)
    x = cheb_points(N)
    c = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    f = npc.chebval(x, c.real) + 1j * npc.chebval(x, c.imag)
    ft = orc.chebtransform(f)
    assert rel(ft, series_to_ft(c)) < 1e-13
    # the inverse transform, and the involution
    assert rel(orc.chebinvtran(series_to_ft(c)), f) < 1e-13
    assert rel(orc.chebinvtran(ft), f) < 1e-13
    # a single polynomial T_k sampled maps to a single coefficient
    k = N // 3
    e = np.zeros(N)
    e[k] = 1.0
    ftk = orc.chebtransform(npc.chebval(x, e).astype(complex))
    want = np.zeros(N, complex)
    want[k] = np.sqrt((N - 1) / 2.0)
    assert np.abs(ftk - want).max() < 1e-12


@pytest.mark.parametrize("N", [9, 17, 33, 65, 97])
def test_cheb_derivatives_exact_on_polynomials(orc, N):
    """D1/D2 of a degree N-1 polynomial, with d/dr = 2 d/dx."""
    rng = np.random.default_rng(100 + N)
    x = cheb_points(N)
    c = (rng.standard_normal(N) + 1j * rng.standard_normal(N)) / (1.0 + np.arange(N)) ** 2
    ft = series_to_ft(c)
    d1 = 2 * (npc.chebval(x, npc.chebder(c.real)) + 1j * npc.chebval(x, npc.chebder(c.imag)))
    d2 = 4 * (npc.chebval(x, npc.chebder(c.real, 2)) + 1j * npc.chebval(x, npc.chebder(c.imag, 2)))
    assert rel(orc.chebinvtranD1(ft), d1) < 1e-12
    assert rel(orc.chebinvtranD2(ft), d2) < 1e-11


@pytest.mark.parametrize("N", [9, 33, 65])
def test_grid(orc, N):
    r, rr, rr2 = orc.grid(N, ETA)
    assert abs(r[0] - RMIN) < 1e-15 and abs(r[-1] - (RMIN + 1)) < 1e-14
    assert np.allclose(r, RMIN + 0.5 * (1 + cheb_points(N)), rtol=0, atol=1e-15)
    assert np.allclose(rr, 1 / r, rtol=1e-15) and np.allclose(rr2, 1 / r ** 2, rtol=1e-15)


@pytest.mark.parametrize("N", [9, 17, 33, 65])
def test_radint_exact_clenshaw_curtis(orc, N):
    """radInt(f) = int_{rmin}^{rmax} f r dr, exact for f r a polynomial of degree < N."""
    r, _, _ = orc.grid(N, ETA)
    rng = np.random.default_rng(N)
    deg = N - 3
    a = rng.standard_normal(deg + 1)
    f = np.polynomial.polynomial.polyval(r, a)
    P = np.polynomial.Polynomial(np.polynomial.polynomial.polymul(a, [0.0, 1.0]))
    exact = P.integ()(RMIN + 1) - P.integ()(RMIN)
    assert abs(orc.radint(f, ETA) - exact) < 1e-12 * max(1.0, abs(exact))

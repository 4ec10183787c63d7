"""Pins of the oracle's nonlinear pass and of the seeded initial state.

Nonlinear pass (nonlin.F90:134-190, fourier_fftw.F90:39-97): the four
products u_phi T, u_r T, u_phi omega, u_r omega, computed here with numpy's
real FFTs on the same Np = 3 Nm points, so the aliasing of the k = 2 Nm term
onto mode Nm is reproduced rather than removed.  The mean <u_r omega>(r) read
by the mean-flow equation (rhs_create.F90:329-335) equals Re of the m = 0
product.

Seeded initial state (a build-defined extension, the reference has no RNG):
SplitMix64 against its published known-answer values, the conductive T_0(r)
of init.F90:207-211 (solution of T'' + T'/r = 0 with T(rmin) = 1,
T(rmax) = 0), zero walls for m >= 1 and the sin envelope.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel
from test_oracle_transforms import ETA

MASK = (1 << 64) - 1


def splitmix64(state):
    """Vigna's SplitMix64: returns (new_state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return state, z ^ (z >> 31)


def test_splitmix64_known_answers():
    # published first outputs of SplitMix64 seeded with 0
    s = 0
    outs = []
    for _ in range(3):
        s, z = splitmix64(s)
        outs.append(z)
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def seed_coeffs(seed, m):
    s = (seed * 0xD1B54A32D192ED03 + m) & MASK
    s, _ = splitmix64(s)
    s, za = splitmix64(s)
    s, zb = splitmix64(s)
    return (za >> 11) * 2.0 ** -53 * 2 - 1, (zb >> 11) * 2.0 ** -53 * 2 - 1


@pytest.mark.parametrize("Nr,Nm,seed", [(17, 8, 0), (33, 16, 7), (65, 128, 1000)])
def test_seed_state(orc, Nr, Nm, seed):
    amp = 1e-3
    t, o, u, w = orc.seed_state(Nr, Nm, ETA, seed, amp)
    r, _, _ = orc.grid(Nr, ETA)
    rmin, rmax = ETA / (1 - ETA), 1 / (1 - ETA)
    T0 = np.log(r / rmax) / np.log(rmin / rmax)
    assert np.abs(t[0] - T0).max() < 1e-14
    assert abs(t[0, 0] - 1) < 1e-15 and abs(t[0, -1]) < 1e-15
    env = np.sin(np.pi * (r - rmin))
    for m in (1, Nm // 2, Nm):
        a, b = seed_coeffs(seed, m)
        want = amp * (a + 1j * b) * env
        want[0] = want[-1] = 0.0
        assert np.abs(t[m] - want).max() < 1e-18
    for f in (o, u, w):
        assert not np.any(f)
    # a_m, b_m are in [-1, 1) and not degenerate
    ab = np.array([seed_coeffs(seed, m) for m in range(1, Nm + 1)])
    assert ab.min() >= -1 and ab.max() < 1 and np.std(ab) > 0.3


@pytest.mark.parametrize("Nr,Nm", [(9, 4), (17, 16), (33, 64)])
def test_nonlin_products_against_numpy(orc, Nr, Nm):
    Np = 3 * Nm
    rng = np.random.default_rng(Nr * Nm)

    def field():
        f = (rng.standard_normal((Nm + 1, Nr)) + 1j * rng.standard_normal((Nm + 1, Nr)))
        return f / (1.0 + np.arange(Nm + 1))[:, None]

    T, W, UP, UR = field(), field(), field(), field()
    upT, urT, upW, urW, mean = orc.nonlin(T, W, UP, UR)

    def phys(f):
        # unnormalised complex-to-real of the zero-padded half spectrum
        full = np.zeros((Np // 2 + 1, Nr), complex)
        full[: Nm + 1] = f
        full[0] = full[0].real
        return np.fft.irfft(full, n=Np, axis=0) * Np

    def spec(x):
        return np.fft.rfft(x, axis=0)[: Nm + 1] / Np

    t, w, up, ur = phys(T), phys(W), phys(UP), phys(UR)
    for got, want in ((upT, spec(up * t)), (urT, spec(ur * t)), (upW, spec(up * w)), (urW, spec(ur * w))):
        assert rel(got, want) < 1e-13
    assert np.abs(mean - (ur * w).mean(0)).max() < 1e-13 * np.abs(ur * w).max()
    assert np.abs(mean - urW[0].real).max() < 1e-13 * np.abs(ur * w).max()


def test_nonlin_aliasing_onto_top_mode(orc):
    """With Np = 3 Nm, cos(Nm phi)^2 = 1/2 + cos(2 Nm phi)/2 and 2 Nm aliases
    onto Nm (3 Nm - 2 Nm = Nm): the m = Nm product is not zero."""
    Nr, Nm = 5, 8
    f = np.zeros((Nm + 1, Nr), complex)
    f[Nm] = 0.5                    # u = cos(Nm phi)
    z = np.zeros_like(f)
    upT, _, _, _, _ = orc.nonlin(f, z, f, z)
    assert np.allclose(upT[0].real, 0.5) and np.allclose(upT[Nm].real, 0.25)

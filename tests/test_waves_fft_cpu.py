"""The FFT wave producer's oracle (oracle/waves_fft.py) pinned on the CPU, and
the library's host spectrum (sfh_wave_spectrum) against it.  No GPU."""
import math

import numpy as np
import pytest

import waves_fft as WF
from stonefish_amd import waves as W


@pytest.mark.parametrize("passes", [1, 3, 6])
def test_network_is_the_unnormalised_inverse_dft(passes):
    """With exact twiddles the reference's butterfly passes (x, then y) give
    sum_k Z(k) exp(+2 pi i k.x / n): n^2 ifft2."""
    n = 1 << passes
    rng = np.random.default_rng(passes)
    Z = rng.normal(size=(n, n, 2)) + 1j * rng.normal(size=(n, n, 2))
    R = WF.network(Z, WF.butterfly_table(passes, exact=True))
    ref = np.fft.ifft2(Z, axes=(0, 1)) * n * n
    np.testing.assert_allclose(R, ref, rtol=0, atol=1e-11 * np.abs(Z).sum())


def test_binary16_twiddles():
    """The table's twiddles are exp(2 pi i m / n) rounded to binary16: within
    2^-11 of the unit circle's points, the pairs exact negatives."""
    j1, j2, w = WF.butterfly_table(8)
    _, _, we = WF.butterfly_table(8, exact=True)
    assert np.all(np.abs(w - we) <= 2.0 ** -11 * np.sqrt(2))
    assert np.all(np.abs(np.abs(we) - 1) < 1e-15)
    assert not np.array_equal(w, we)


def test_single_line_is_a_travelling_cosine():
    """One spectral line h0 at texel (x0, 0) of grid 1: the height of grid 1 is
    2 sqrt2 |h0| cos(2 pi x0 x / n + w t + arg h0) with w = sqrt(g K (1 +
    (K / 370)^2)), K = 2 pi x0 / L; the other grids stay flat."""
    passes, x0, L = 5, 3, 40.0
    n = 1 << passes
    s12 = np.zeros((n, n, 4))
    s34 = np.zeros((n, n, 4))
    h0 = complex(0.5, -0.25)
    s12[0, x0, 0], s12[0, x0, 1] = h0.real, h0.imag
    grids = (L, 7.0, 3.0, 2.0)
    r2 = float(np.float32(1.414213562))
    for t in (0.0, 3.7):
        H, _ = WF.heights(s12, s34, passes, grids, t, half=False, exact=True)
        K = x0 / n * float(np.float32(2 * math.pi * n / L))
        w = math.sqrt(9.81 * K * (1 + (K / 370.0) ** 2))
        x = np.arange(n)
        expect = 2 * r2 * abs(h0) * np.cos(2 * math.pi * x0 * x / n + w * t + math.atan2(h0.imag, h0.real))
        np.testing.assert_allclose(H[:, :, 0], np.broadcast_to(expect, (n, n)), atol=1e-12)
        np.testing.assert_allclose(H[:, :, 1:], 0, atol=1e-12)


def test_lcg_is_the_c_standard_generator():
    """next = next * 1103515245 + 12345 (mod 2^31): from seed 1 the C
    standard's example rand() returns (next / 65536) % 32768 = 16838; the
    reference's uniforms are the top 24 of the 31 bits."""
    u = WF.lcg_uniforms(1, seed=1)[0]
    assert (int(u * 2 ** 24) >> 9) & 0x7FFF == 16838
    v = WF.lcg_uniforms(1000)
    assert np.all((v >= 0) & (v < 1)) and np.all(v * 2 ** 24 == np.floor(v * 2 ** 24))


@pytest.mark.parametrize("state", [0.5, 2.0])
def test_library_spectrum_matches_restatement(state):
    """sfh_wave_spectrum (float arithmetic, like the reference) against the FP64
    restatement: exact zeros for waves against +x (propagating spectrum) and
    in each grid's low-k box, values within FP32 rounding of the formula
    elsewhere.  Across the wind (kx = 0) the directional factor 1 + Delta
    cos(2 phi) nearly cancels (Delta ~ 1); formed in FP32 it is off by ~1e-7
    absolute, which moves h = sqrt(S / 2) dk there by up to 1e-5 of the
    spectrum's peak: the absolute tolerance."""
    passes = 5
    n = 1 << passes
    g12, g34 = W.reference_spectrum(state, passes)
    o12, o34 = WF.spectrum(state, passes)
    for g, o in ((g12, o12), (g34, o34)):
        scale = np.abs(o).max()
        np.testing.assert_allclose(g, o, rtol=2e-4, atol=1e-5 * scale)
    i = np.where(np.arange(n) >= n // 2, np.arange(n) - n, np.arange(n))
    I, J = np.meshgrid(i, i, indexing="xy")
    L = WF.DEFAULT_GRIDS
    kmin = [math.pi / L[0], math.pi * n / L[0], math.pi * n / L[1], math.pi * n / L[2]]
    for grid in range(4):
        s = (g12 if grid < 2 else g34)[..., 2 * (grid % 2):2 * (grid % 2) + 2]
        dk = 2 * math.pi / L[grid]
        box = (np.abs(I * dk) < kmin[grid]) & (np.abs(J * dk) < kmin[grid])
        assert not s[(I < 0) | box].any()


def test_spectrum_argument_checks():
    from stonefish_amd import _native as N
    with pytest.raises(N.SfhError):
        W.reference_spectrum(1.0, 12)
    with pytest.raises(N.SfhError):
        W.reference_spectrum(1.0, 4, grid_sizes=(893.0, 0.0, 21.0, 11.0))

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


# ---- the spectrum pinned by what the mathematics fixes -----------------------
# (no use of the restatement's formula: these follow from the structure of a
# directional wave spectrum S(k) (1 + Delta(k) cos 2 phi) / (2 pi), the
# propagating half plane, the long-wave cut-off and the energy of a fully
# developed sea)

def _grid_S(state, passes, grid):
    """S at the integer wavenumber indices (i, j) of one grid, from the
    library's samples h = sqrt(S / 2) dk (|sample| = h): S = 2 (h / dk)^2."""
    n = 1 << passes
    s12, s34 = W.reference_spectrum(state, passes)
    s = (s12 if grid < 2 else s34)[..., 2 * (grid % 2):2 * (grid % 2) + 2].astype(np.float64)
    h = np.hypot(s[..., 0], s[..., 1])
    dk = 2 * math.pi / WF.DEFAULT_GRIDS[grid]
    idx = np.where(np.arange(n) >= n // 2, np.arange(n) - n, np.arange(n))
    return 2.0 * (h / dk) ** 2, idx, dk


@pytest.mark.parametrize("state", [0.5, 2.0])
def test_spectrum_mirror_symmetric_across_the_wind(state):
    """cos 2 phi is even in phi: S(kx, ky) = S(kx, -ky) on the propagating half
    plane (the phases of the two samples are independent draws, their
    magnitudes are not)."""
    passes = 6
    n = 1 << passes
    for grid in range(4):
        S, idx, _ = _grid_S(state, passes, grid)
        for y in range(1, n // 2):
            ym = n - y
            np.testing.assert_allclose(S[y, :], S[ym, :], rtol=3e-6, atol=1e-12 * S.max())


@pytest.mark.parametrize("state", [0.5, 2.0])
def test_spectrum_angular_form(state):
    """On one ring |k| = const the spectrum varies as 1 + Delta(k) cos 2 phi:
    for (i, j) and (j, i) (the same |k|, phi and pi/2 - phi) the ratio
    r = S(i, j) / S(j, i) gives Delta(k) cos 2 phi = (r - 1) / (r + 1), so every
    pair on a ring must give the same Delta(k), with |Delta| <= 1 (a tanh).
    A wrong angular law (cos phi, a sign, a missing factor 2) fails it."""
    passes = 7
    n = 1 << passes
    S = max((_grid_S(state, passes, g)[0] for g in range(4)), key=lambda a: a.max())   # the grid holding the peak
    floor = 1e-6 * S.max()
    rings = {}
    for i in range(1, n // 2):
        for j in range(i + 1, n // 2):
            s1, s2 = S[j, i], S[i, j]          # rows y = ky index, columns x = kx index
            if s1 <= floor or s2 <= floor:
                continue
            r = s1 / s2
            c2 = (i * i - j * j) / (i * i + j * j)   # cos 2 phi of (kx, ky) = (i, j)
            rings.setdefault(i * i + j * j, []).append((r - 1) / (r + 1) / c2)
    multi = {k: v for k, v in rings.items() if len(v) > 1}
    assert len(multi) > 20
    for v in multi.values():
        assert max(abs(x) for x in v) <= 1.0 + 1e-5
        assert max(v) - min(v) <= 2e-4 * max(1.0, abs(np.mean(v)))


def test_spectrum_long_wave_cutoff_and_peak():
    """Wavelengths four times the peak's and longer carry almost no energy
    (the Pierson-Moskowitz factor exp(-5/4 (kp / k)^2) = 2e-9 at kp / 4
    against the k^-4 growth of 256), and the spectrum's
    variance peaks at |k| near kp = g (Omega / U10)^2."""
    state = 1.5
    wind = state * 5.0 + 2.0
    Om = 5.0 * math.exp(-state) + 0.2
    kp = 9.81 * (Om / wind) ** 2
    S, idx, dk = _grid_S(state, 7, 0)
    I, J = np.meshgrid(idx, idx, indexing="xy")
    k = np.hypot(I, J) * dk
    m = S > 0
    assert S[m & (k < kp / 4)].max() < 1e-4 * S.max()
    # variance per |k| bin (S k dk dphi summed over a ring) peaks within a factor 1.5 of kp
    kb = np.floor(k / dk).astype(int)
    var = np.bincount(kb[m], weights=S[m] * dk * dk)
    kpk = (np.argmax(var) + 0.5) * dk
    assert kp / 1.5 < kpk < kp * 1.5, (kpk, kp)


def test_fully_developed_sea_height():
    """At full development (inverse wave age Omega = 0.84) the sea is the
    Pierson-Moskowitz one with Hs = 0.21 U10^2 / g; Elfouhaily's long-wave
    part carries a JONSWAP peak enhancement (gamma = 1.7 for Omega < 1), which
    adds energy around the peak, so the sampled spectrum's Hs = 4 sqrt(m0),
    m0 = sum of S dk^2 over the four grids' samples, lies between 1.05 and
    1.40 times the PM height.  A propagating half plane without its doubling
    (Hs / sqrt 2), a missing 2 pi or a wrong power of k falls outside."""
    Om = 0.84
    state = -math.log((Om - 0.2) / 5.0)
    wind = state * 5.0 + 2.0
    s12, s34 = W.reference_spectrum(state, 7)
    m0 = 0.0
    for arr in (s12, s34):
        a = arr.astype(np.float64)
        m0 += 2.0 * float(np.sum(a ** 2))     # S dk^2 = 2 h^2 per sample
    ratio = 4.0 * math.sqrt(m0) / (0.21 * wind ** 2 / 9.81)
    assert 1.05 < ratio < 1.40, ratio

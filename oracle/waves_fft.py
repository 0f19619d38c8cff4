"""FP64 restatement of the reference's FFT wave producer -- TEST INFRASTRUCTURE.

Only tests/ (and the entry point's smoke check) use this file, as the checker
of sfh_wave_spectrum / sfh_wave_fft_*; the product never calls it.

  lrandom / frandom                  SystemUtil.hpp:207-217
  OpenGLOcean::omega, spectrum       OpenGLOcean.cpp:934-1001 (Elfouhaily et al. 1997)
  GetSpectrumSample                  OpenGLOcean.cpp:1003-1022
  GenerateWavesSpectrum              OpenGLOcean.cpp:1025-1046
  sea state                          OpenGLRealOcean.cpp:43-45, OpenGLOcean.cpp:63-71
  bitReverse, computeWeight          OpenGLOcean.cpp:856-877
  ComputeButterflyLookupTable        OpenGLOcean.cpp:879-925 (RGBA16F: twiddles as binary16)
  h(k, t) and the packing            oceanInit.frag:36-83 (spectra from RGBA16F textures)
  butterfly passes                   oceanFFTX.frag, oceanFFTY.frag; order OpenGLOcean.cpp:479-530

Pinned by tests/test_waves_fft_cpu.py: the butterfly network with exact
twiddles equals n^2 ifft2 (numpy), a one-line spectrum gives the closed-form
travelling cosine with the deep-water (capillary-corrected) dispersion, and
the LCG's 24-bit uniforms.  The Elfouhaily spectrum is pinned in its
structure and energy by what the mathematics fixes (the library's samples,
independent of this restatement's formula): its propagating half plane and
low-k boxes are exact zeros, it is mirror symmetric across the wind, on each
|k| ring it varies as 1 + Delta(k) cos 2 phi with one Delta per ring, it
carries no energy four peak wavelengths out and its variance peaks near
kp = g (Omega / U10)^2, and at full development its significant height
lies within the JONSWAP enhancement (1.05-1.40) of the Pierson-Moskowitz
0.21 U10^2 / g.  Its short-wave constants (alpha_m, cm, km inside B_h) are
checked only against this restatement: parity unpinned for those.
"""
from __future__ import annotations

import math

import numpy as np

DEFAULT_GRIDS = (893.0, 101.0, 21.0, 11.0)


def lcg_uniforms(count: int, seed: int = 1234) -> np.ndarray:
    """frandom: seed = (seed * 1103515245 + 12345) & 0x7FFFFFFF, value = (seed >> 7) / 2^24."""
    out = np.empty(count)
    s = seed
    for i in range(count):
        s = (s * 1103515245 + 12345) & 0x7FFFFFFF
        out[i] = (s >> 7) / float(1 << 24)
    return out


def _omega(k, km=370.0):
    return math.sqrt(9.81 * k * (1.0 + (k / km) ** 2))


def elfouhaily(kx, ky, wind, Om, A=1.0, km=370.0, cm=0.23, propagate=True):
    k = math.sqrt(kx * kx + ky * ky)
    c = _omega(k, km) / k
    kp = 9.81 * (Om / wind) ** 2
    cp = _omega(kp, km) / kp
    z0 = 3.7e-5 * wind ** 2 / 9.81 * (wind / cp) ** 0.9
    u_star = 0.41 * wind / math.log(10.0 / z0)
    Lpm = math.exp(-5.0 / 4.0 * (kp / k) ** 2)
    gamma = 1.7 if Om < 1.0 else 1.7 + 6.0 * math.log(Om)
    sigma = 0.08 * (1.0 + 4.0 / Om ** 3)
    Gamma = math.exp(-1.0 / (2.0 * sigma ** 2) * (math.sqrt(k / kp) - 1.0) ** 2)
    Fp = Lpm * gamma ** Gamma * math.exp(-Om / math.sqrt(10.0) * (math.sqrt(k / kp) - 1.0))
    Bl = 0.5 * 0.006 * math.sqrt(Om) * cp / c * Fp
    alpham = 0.01 * (1.0 + math.log(u_star / cm) if u_star < cm else 1.0 + 3.0 * math.log(u_star / cm))
    Bh = 0.5 * alpham * cm / c * math.exp(-0.25 * (k / km - 1.0) ** 2) * Lpm
    am = 0.13 * u_star / cm
    Delta = math.tanh(math.log(2.0) / 4.0 + 4.0 * (c / cp) ** 2.5 + am * (cm / c) ** 2.5)
    if propagate:
        if kx < 0.0:
            return 0.0
        Bl *= 2.0
        Bh *= 2.0
    return A * (Bl + Bh) * (1.0 + Delta * math.cos(2.0 * math.atan2(ky, kx))) / (2.0 * math.pi * k ** 4)


def spectrum(state: float, passes: int, grid_sizes=DEFAULT_GRIDS):
    """(spectrum12, spectrum34), each (n, n, 4), row y, column x."""
    n = 1 << passes
    wind = state * 5.0 + 2.0
    Om = 5.0 * math.exp(-state) + 0.2
    L = [float(g) for g in grid_sizes]
    kmins = [math.pi / L[0], math.pi * n / L[0], math.pi * n / L[1], math.pi * n / L[2]]
    s12 = np.zeros((n, n, 4))
    s34 = np.zeros((n, n, 4))
    seed = 1234
    for y in range(n):
        for x in range(n):
            i = x - n if x >= n // 2 else x
            j = y - n if y >= n // 2 else y
            for g in range(4):
                dk = 2.0 * math.pi / L[g]
                kx, ky = i * dk, j * dk
                dst = (s12 if g < 2 else s34)[y, x]
                o = 2 * (g % 2)
                if abs(kx) < kmins[g] and abs(ky) < kmins[g]:
                    continue
                S = elfouhaily(kx, ky, wind, Om)
                h = math.sqrt(S / 2.0) * dk
                seed = (seed * 1103515245 + 12345) & 0x7FFFFFFF
                phi = (seed >> 7) / float(1 << 24) * 2.0 * math.pi
                dst[o] = h * math.cos(phi)
                dst[o + 1] = h * math.sin(phi)
    return s12, s34


def bit_reverse(i: int, n: int) -> int:
    r, m, w = 0, n >> 1, 1
    while m:
        if i & m:
            r += w
        m >>= 1
        w <<= 1
    return r


def butterfly_table(passes: int, half: bool = True, exact: bool = False):
    """(j1, j2, w): per pass and output index the two inputs and the twiddle
    (complex): computeWeight's float, rounded to binary16 as in the
    reference's RGBA16F table (half), or exact (no rounding at all)."""
    n = 1 << passes
    j1 = np.zeros((passes, n), dtype=np.int64)
    j2 = np.zeros((passes, n), dtype=np.int64)
    w = np.zeros((passes, n), dtype=np.complex128)
    for p in range(passes):
        nblocks, nh = 1 << (passes - 1 - p), 1 << p
        for b in range(nblocks):
            for k in range(nh):
                i1 = b * nh * 2 + k
                i2 = i1 + nh
                a1 = bit_reverse(i1, n) if p == 0 else i1
                a2 = bit_reverse(i2, n) if p == 0 else i2
                arg = 2.0 * math.pi * k * nblocks / n
                wr, wi = math.cos(arg), math.sin(arg)
                if not exact:
                    wr, wi = float(np.float32(wr)), float(np.float32(wi))
                if half and not exact:
                    wr, wi = float(np.float16(wr)), float(np.float16(wi))
                j1[p, i1], j2[p, i1], w[p, i1] = a1, a2, complex(wr, wi)
                j1[p, i2], j2[p, i2], w[p, i2] = a1, a2, complex(-wr, -wi)
    return j1, j2, w


def network(Z: np.ndarray, table) -> np.ndarray:
    """The passes along x (axis 1) then y (axis 0) of a complex (n, n[, m]) array."""
    j1, j2, w = table
    Z = np.asarray(Z, dtype=np.complex128)
    ww = w.reshape(w.shape + (1,) * (Z.ndim - 2))
    for p in range(j1.shape[0]):
        Z = Z[:, j1[p]] + ww[p][None] * Z[:, j2[p]]
    for p in range(j1.shape[0]):
        Z = Z[j1[p]] + ww[p][:, None] * Z[j2[p]]
    return Z


def packed_spectrum(s12, s34, passes: int, grid_sizes, t: float, half: bool = True):
    """oceanInit.frag: (h1 + i h2, h3 + i h4) at every texel, shape (n, n, 2)."""
    n = 1 << passes
    r2 = float(np.float32(1.414213562))
    if half:
        s12 = np.asarray(s12, dtype=np.float32).astype(np.float16).astype(np.float64)
        s34 = np.asarray(s34, dtype=np.float32).astype(np.float16).astype(np.float64)
    s = np.concatenate([s12, s34], axis=2) * r2                 # (n, n, 8)
    idx = np.arange(n)
    neg = (n - idx) % n
    sc = s[neg][:, neg]                                          # texel of -k
    f = np.where(idx >= n // 2, idx - n, idx) / n
    FX, FY = np.meshgrid(f, f, indexing="xy")
    h = np.zeros((n, n, 4), dtype=np.complex128)
    for g in range(4):
        inv = float(np.float32(2.0 * math.pi * n / float(np.float32(grid_sizes[g]))))
        K = np.hypot(FX * inv, FY * inv)
        om = np.sqrt(9.81 * K * (1.0 + K * K / (370.0 * 370.0)))
        c, sn = np.cos(om * t), np.sin(om * t)
        px, py, qx, qy = s[..., 2 * g], s[..., 2 * g + 1], sc[..., 2 * g], sc[..., 2 * g + 1]
        h[..., g] = ((px + qx) * c - (py + qy) * sn) + 1j * ((px - qx) * sn + (py - qy) * c)
    A = h[..., 0] + 1j * h[..., 1]
    B = h[..., 2] + 1j * h[..., 3]
    return np.stack([A, B], axis=2)


def heights(s12, s34, passes: int, grid_sizes=DEFAULT_GRIDS, t: float = 0.0, half: bool = True,
            exact: bool = False):
    """(n, n, 4) heights of grids 1-4 (the reference's fftData layer 0) and
    the packed spectrum it came from."""
    Z = packed_spectrum(s12, s34, passes, grid_sizes, t, half)
    R = network(Z, butterfly_table(passes, half, exact))
    out = np.stack([R[..., 0].real, R[..., 0].imag, R[..., 1].real, R[..., 1].imag], axis=2)
    return out, Z

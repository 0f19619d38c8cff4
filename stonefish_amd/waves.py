"""Wave-field input generator: seeded sinusoids with Pierson-Moskowitz amplitudes.

The elevation grid (upward elevation eta, like the reference's fftData texels,
OpenGLRealOcean.cpp:333-356) is built on the GPU by sfh_wave_grid_build and
fed to Ocean::GetDepth's bilinear query.  Wavenumbers are snapped to integer
multiples of 2 pi / L so the field tiles the L x L grid periodically, which is
what the query's wrap-around assumes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

G = 9.81


def pm_components(n_waves: int = 64, wind_speed: float = 12.0, length: float = 1024.0, seed: int = 2,
                  w_lo: float = 0.5, w_hi: float = 3.0) -> np.ndarray:
    """Returns a structured array of sfh_wave_component fields (amplitude,
    phase, mx, my).  Pierson-Moskowitz spectrum S(w) = a g^2 w^-5
    exp(-b (w0/w)^4), a = 8.1e-3, b = 0.74, w0 = g / U; n_waves frequencies
    evenly spaced in [w_lo, w_hi] x the spectral peak, directions uniform,
    deep-water dispersion k = w^2 / g."""
    rng = np.random.default_rng(seed)
    w0 = G / wind_speed
    wp = (0.8 * 0.74) ** 0.25 * w0
    w = np.linspace(w_lo * wp, w_hi * wp, n_waves)
    dw = w[1] - w[0] if n_waves > 1 else wp
    S = 8.1e-3 * G * G / w ** 5 * np.exp(-0.74 * (w0 / w) ** 4)
    amp = np.sqrt(2.0 * S * dw)
    theta = rng.uniform(0.0, 2.0 * np.pi, n_waves)
    phase = rng.uniform(0.0, 2.0 * np.pi, n_waves)
    k = w * w / G
    mx = np.rint(k * np.cos(theta) * length / (2.0 * np.pi)).astype(np.int32)
    my = np.rint(k * np.sin(theta) * length / (2.0 * np.pi)).astype(np.int32)
    zero = (mx == 0) & (my == 0)
    mx[zero] = 1
    dt = np.dtype([("amplitude", np.float32), ("phase", np.float32), ("mx", np.int32), ("my", np.int32)])
    out = np.zeros(n_waves, dtype=dt)
    out["amplitude"], out["phase"], out["mx"], out["my"] = amp, phase, mx, my
    return out


def grid_reference(comps: np.ndarray, n: int) -> np.ndarray:
    """float64 evaluation of the same field (for checking the GPU builder)."""
    i = np.arange(n)
    I, J = np.meshgrid(i, i, indexing="xy")   # J row (y), I column (x)
    h = np.zeros((n, n))
    for c in comps:
        m = (np.int64(c["mx"]) * I + np.int64(c["my"]) * J) % n
        h += float(c["amplitude"]) * np.cos(2.0 * np.pi * m / n + float(c["phase"]))
    return h


def build_grid_device(comps: np.ndarray, n: int, d_grid: int, stream: int = 0) -> None:
    """Builds the n x n grid into device memory at d_grid (float32)."""
    arr = (N.sfh_wave_component * len(comps))()
    for k, c in enumerate(comps):
        arr[k].amplitude = float(c["amplitude"])
        arr[k].phase = float(c["phase"])
        arr[k].mx = int(c["mx"])
        arr[k].my = int(c["my"])
    N.check(N.load().sfh_wave_grid_build(arr, len(comps), n, C.c_void_p(d_grid), C.c_void_p(stream)))


# ---- the reference's FFT ocean, headless ----------------------------------------

REFERENCE_GRID_SIZES = (893.0, 101.0, 21.0, 11.0)


def reference_spectrum(state: float, passes: int = 8, grid_sizes=REFERENCE_GRID_SIZES):
    """sfh_wave_spectrum: (spectrum12, spectrum34), each (n, n, 4) float32,
    n = 2^passes (OpenGLOcean::GenerateWavesSpectrum for the sea state
    OpenGLRealOcean derives from `state`)."""
    n = 1 << passes
    s12 = np.zeros((n, n, 4), np.float32)
    s34 = np.zeros((n, n, 4), np.float32)
    g = np.ascontiguousarray(grid_sizes, dtype=np.float32)
    N.check(N.load().sfh_wave_spectrum(float(state), int(passes), g.ctypes.data_as(C.c_void_p),
                                       s12.ctypes.data_as(C.c_void_p), s34.ctypes.data_as(C.c_void_p)))
    return s12, s34


class WaveProducer:
    """sfh_wave_fft_*: the reference's wave heights (fftData layer 0) at time t
    computed on the device into an (n, n, 4) float32 buffer."""

    def __init__(self, spectrum12, spectrum34, passes: int = 8, grid_sizes=REFERENCE_GRID_SIZES):
        n = 1 << passes
        a = np.ascontiguousarray(spectrum12, dtype=np.float32).reshape(n, n, 4)
        b = np.ascontiguousarray(spectrum34, dtype=np.float32).reshape(n, n, 4)
        self.grid_sizes = np.ascontiguousarray(grid_sizes, dtype=np.float32)
        h = C.c_void_p()
        N.check(N.load().sfh_wave_fft_create(int(passes), self.grid_sizes.ctypes.data_as(C.c_void_p),
                                             a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                                             C.byref(h)))
        self.handle = h
        self.n = n
        self.passes = passes

    def step(self, t: float, d_heights: int, stream: int = 0) -> None:
        """Writes the heights at time t to device memory d_heights (n*n*4 floats)."""
        N.check(N.load().sfh_wave_fft_step(self.handle, float(t), C.c_void_p(d_heights), C.c_void_p(stream)))

    def close(self) -> None:
        if self.handle:
            N.check(N.load().sfh_wave_fft_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""GPU parity of the headless wave producer (sfh_wave_fft_*) against the FP64
restatement oracle/waves_fft.py, through the C ABI, on the reference's own
spectrum (sfh_wave_spectrum, sea state 1.5) and a random one.

Bound per height: the passes run in FP32; every intermediate value of a
line is bounded by sum |Z| of the packed spectrum (|w| <= 1 + 2^-11 for the
binary16 twiddles), each butterfly adds at most ~4 roundings of that size
and each output goes through 2 P butterflies after the FP32 init, so
|h - h64| <= (8 P + 4) 2^-24 sum |Z| (per packed complex signal)."""
import sys

import numpy as np
import pytest

import waves_fft as WF
from stonefish_amd import hydro as H
from stonefish_amd import waves as W

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(prod, t):
    import torch
    out = torch.zeros((prod.n, prod.n, 4), dtype=torch.float32, device="cuda:0")
    prod.step(t, out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("passes", [5, 8, 10])
def test_wave_producer_matches_oracle(gpu, passes):
    log(f"case: wave producer, passes={passes}")
    s12, s34 = W.reference_spectrum(1.5, passes)
    prod = W.WaveProducer(s12, s34, passes)
    for t in (0.0, 37.25):
        g = _run(prod, t)
        ref, Z = WF.heights(s12, s34, passes, W.REFERENCE_GRID_SIZES, t)
        for c, S in ((0, np.abs(Z[..., 0]).sum()), (1, np.abs(Z[..., 0]).sum()), (2, np.abs(Z[..., 1]).sum()),
                     (3, np.abs(Z[..., 1]).sum())):
            err = np.abs(g[..., c] - ref[..., c]).max()
            bound = (8 * passes + 4) * 2.0 ** -24 * S
            log(f"  t={t} grid {c + 1}: max |h| {np.abs(ref[..., c]).max():.3e}, err {err:.2e}, bound ratio {err / bound:.3f}")
            assert err <= bound
    prod.close()


def test_wave_producer_random_spectrum_and_ocean(gpu):
    """A random spectrum (every texel, including the Nyquist lines), and the
    heights fed to an ocean as the reference's two-grid layout."""
    log("case: wave producer, random spectrum, ocean hand-off")
    passes = 7
    n = 1 << passes
    rng = np.random.default_rng(9)
    s12 = rng.normal(size=(n, n, 4)).astype(np.float32) * 0.05
    s34 = rng.normal(size=(n, n, 4)).astype(np.float32) * 0.01
    grids = (200.0, 50.0, 10.0, 5.0)
    prod = W.WaveProducer(s12, s34, passes, grids)
    import torch
    out = torch.zeros((n, n, 4), dtype=torch.float32, device="cuda:0")
    prod.step(12.5, out.data_ptr())
    torch.cuda.synchronize()
    g = out.cpu().numpy().astype(np.float64)
    ref, Z = WF.heights(s12, s34, passes, grids, 12.5)
    for c in range(4):
        S = np.abs(Z[..., c // 2]).sum()
        assert np.abs(g[..., c] - ref[..., c]).max() <= (8 * passes + 4) * 2.0 ** -24 * S
    ocean = H.Ocean()
    ocean.set_waves(None, n, 4, list(grids[:2]), device_ptr=out.data_ptr())
    prod.close()


def test_wave_producer_rejects_bad_arguments(gpu):
    from stonefish_amd import _native as N
    s = np.zeros((4, 4, 4), np.float32)
    with pytest.raises(N.SfhError):
        W.WaveProducer(s, s, 2, (893.0, -1.0, 21.0, 11.0))
    with pytest.raises(N.SfhError):
        N.check(N.load().sfh_wave_fft_step(None, 0.0, None, None))

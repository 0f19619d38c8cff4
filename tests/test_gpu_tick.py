"""The tick driver (sfh_tick_*): recompute every fdPrescaler ticks over the
bodies the fluid overlaps, apply the held forces every tick
(SimulationManager.cpp:1583-1655, Ocean.cpp:219-268, SolidEntity.cpp:
1958-1968).  Checked exactly against the batch compute it drives (whose
parity with the oracle is tested elsewhere) and the reference's tick rules
restated in the test."""
import sys

import numpy as np
import pytest

from stonefish_amd import hydro as H
from stonefish_amd import meshes as M
from stonefish_amd import transforms as T
from stonefish_amd import waves as W

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _states(rng, n, depths):
    s = np.zeros((n, 22))
    q = T.random_unit_quaternions(rng, n)
    pos = np.stack([rng.uniform(0, 60, n), rng.uniform(0, 60, n), depths], axis=1)
    s[:, :16] = T.quat_to_gl(q, pos)
    s[:, 16:19] = rng.uniform(-2, 2, (n, 3))
    s[:, 19:22] = rng.uniform(-1, 1, (n, 3))
    return s


def _run_ticks(batch, fluid, n, prescaler, ticks, overlap_of, states_of, aero):
    import torch
    dev = torch.device("cuda", 0)
    drv = H.TickDriver(batch, prescaler)
    d_out = torch.zeros((n, 24), dtype=torch.float64, device=dev)
    d_app = torch.full((n, 6), 9.0, dtype=torch.float64, device=dev)
    ref_out = np.zeros((n, 24))
    for k in range(ticks):
        st = states_of(k)
        ov = overlap_of(k)
        d_st = torch.from_numpy(st).to(dev)
        d_ov = torch.from_numpy(ov.astype(np.uint8)).to(dev)
        rec = drv.step(fluid, d_st.data_ptr(), d_ov.data_ptr(), d_out.data_ptr(), d_app.data_ptr())
        torch.cuda.synchronize()
        assert rec == (k % prescaler == 0), k
        if rec:
            fresh = batch.compute_host(fluid, st, ref_out.copy())
            ref_out[ov] = fresh[ov]
        ref_app = np.zeros((n, 6))
        for i in np.nonzero(ov)[0]:
            o = ref_out[i]
            if aero:
                ref_app[i] = np.concatenate([o[6:9], o[9:12]])
            else:
                ref_app[i] = np.concatenate([(o[0:3] + o[6:9]) + o[12:15], (o[3:6] + o[9:12]) + o[15:18]])
        np.testing.assert_array_equal(d_out.cpu().numpy(), ref_out)
        np.testing.assert_array_equal(d_app.cpu().numpy(), ref_app)
    drv.close()


def test_tick_driver_hydro(gpu):
    """Six bodies on waves, prescaler 3 over 8 ticks; body 2 leaves the
    overlap on ticks 1-4, body 5 never overlaps (its row stays zero)."""
    log("case: tick driver, hydrodynamics, prescaler 3")
    rng = np.random.default_rng(3)
    shapes = [M.ico_ellipsoid(1.0, 0.4, 0.3, 3), M.box(0.5, 0.3, 0.2, 6), M.capsule(0.2, 1.5)]
    descs = []
    for i in range(6):
        tm = shapes[i % 3]
        descs.append(H.BodyDesc(H.Mesh(tm), volume=1.0, surface=2.0))
    batch = H.BodyBatch(descs)
    ocean = H.Ocean()
    comps = W.pm_components(16, 6.0, 64.0, seed=3)
    ocean.set_waves(W.grid_reference(comps, 64).astype(np.float32), 64, 1, [64.0])
    depths = np.array([-0.2, 0.1, 3.0, 0.0, -3.0, 0.3])
    states = [_states(rng, 6, depths + rng.uniform(-0.1, 0.1, 6)) for _ in range(8)]

    def overlap(k):
        ov = np.ones(6, bool)
        ov[5] = False
        if 1 <= k <= 4:
            ov[2] = False
        return ov
    _run_ticks(batch, ocean, 6, 3, 8, overlap, lambda k: states[k], aero=False)


def test_tick_driver_aero(gpu):
    """An aerodynamic batch under the driver applies Fda, Tda; prescaler 1
    recomputes every tick."""
    log("case: tick driver, aerodynamics, prescaler 1")
    rng = np.random.default_rng(4)
    descs = [H.AeroDesc(H.Mesh(M.box(0.4, 0.3, 0.2, 4)), approx_type=2, approx_params=[0.3, 0.8, 0.0]),
             H.AeroDesc(H.Mesh(M.ico_ellipsoid(0.5, 0.4, 0.3, 3)), approx_type=3, approx_params=[0.5, 0.4, 0.3]),
             H.AeroDesc(H.Mesh(M.capsule(0.1, 0.8)), aero=False)]
    batch = H.AeroBatch(descs)
    atm = H.Atmosphere(1.2)
    atm.add_uniform_current([3.0, 1.0, 0.0])
    states = [_states(rng, 3, np.array([-10.0, -12.0, -8.0])) for _ in range(3)]
    _run_ticks(batch, atm, 3, 1, 3, lambda k: np.ones(3, bool), lambda k: states[k], aero=True)


def test_tick_driver_arguments(gpu):
    from stonefish_amd import _native as N
    batch = H.BodyBatch([H.BodyDesc(H.Mesh(M.box(0.3, 0.2, 0.1, 2)))])
    with pytest.raises(N.SfhError):
        H.TickDriver(batch, 0)

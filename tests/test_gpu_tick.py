"""The tick driver (sfh_tick_*): recompute every fdPrescaler ticks over the
bodies the fluid overlaps, apply the held forces every tick
(SimulationManager.cpp:1583-1655, Ocean.cpp:219-268, SolidEntity.cpp:
1958-1968).  Checked against the FP64 oracle tick by tick (the held outputs
with the reference's stale-output rules, and the applied force and torque,
within the bound of tests/util.py) and exactly against the batch compute the
driver wraps and the reference's tick rules restated in the test."""
import sys

import numpy as np
import pytest

import oracle as O
import hull as HO
from stonefish_amd import hydro as H
from stonefish_amd import meshes as M
from stonefish_amd import transforms as T
from stonefish_amd import waves as W
from util import RTOL, STOL

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


def _applied(o, aero):
    if aero:
        return np.concatenate([o[6:9], o[9:12]])
    return np.concatenate([(o[0:3] + o[6:9]) + o[12:15], (o[3:6] + o[9:12]) + o[15:18]])


def _run_ticks(batch, fluid, n, prescaler, ticks, overlap_of, states_of, aero, oracle_eval):
    """oracle_eval(i, state, previous row) -> (FP64 oracle row, bound per
    output vector (6,)) of body i with the reference's untouched slots."""
    import torch
    dev = torch.device("cuda", 0)
    drv = H.TickDriver(batch, prescaler)
    d_out = torch.zeros((n, 24), dtype=torch.float64, device=dev)
    d_app = torch.full((n, 6), 9.0, dtype=torch.float64, device=dev)
    ref_out = np.zeros((n, 24))
    orc_out = np.zeros((n, 24))
    orc_bound = np.zeros((n, 6))
    worst = 0.0
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
            for i in np.nonzero(ov)[0]:
                row, bnd = oracle_eval(i, st[i], orc_out[i])
                wm = int(row[20])
                for j in range(6):
                    if (wm >> j) & 1:
                        orc_bound[i, j] = bnd[j]
                orc_out[i] = row
        # the held outputs and the applied force/torque against the oracle
        g = d_out.cpu().numpy()
        ga = d_app.cpu().numpy()
        for i in np.nonzero(ov)[0]:
            assert g[i, 20] == orc_out[i, 20] and g[i, 21] == orc_out[i, 21], (k, i)
            for j in range(6):
                e = np.linalg.norm(g[i, 3 * j:3 * j + 3] - orc_out[i, 3 * j:3 * j + 3])
                if orc_bound[i, j] > 0:
                    worst = max(worst, e / orc_bound[i, j])
                    assert e <= orc_bound[i, j], (k, i, j, e, orc_bound[i, j])
                else:
                    assert e == 0.0, (k, i, j)
            oa = _applied(orc_out[i], aero)
            bF = orc_bound[i, [2]].sum() if aero else orc_bound[i, [0, 2, 4]].sum()
            bT = orc_bound[i, [3]].sum() if aero else orc_bound[i, [1, 3, 5]].sum()
            assert np.linalg.norm(ga[i, :3] - oa[:3]) <= bF and np.linalg.norm(ga[i, 3:] - oa[3:]) <= bT, (k, i)
        ref_app = np.zeros((n, 6))
        for i in np.nonzero(ov)[0]:
            o = ref_out[i]
            if aero:
                ref_app[i] = np.concatenate([o[6:9], o[9:12]])
            else:
                ref_app[i] = np.concatenate([(o[0:3] + o[6:9]) + o[12:15], (o[3:6] + o[9:12]) + o[15:18]])
        np.testing.assert_array_equal(d_out.cpu().numpy(), ref_out)
        np.testing.assert_array_equal(d_app.cpu().numpy(), ref_app)
    log(f"  worst bound ratio against the oracle over {ticks} ticks: {worst:.2f}")
    drv.close()


def test_tick_driver_hydro(gpu):
    """Six bodies on waves, prescaler 3 over 8 ticks; body 2 leaves the
    overlap on ticks 1-4, body 5 never overlaps (its row stays zero)."""
    log("case: tick driver, hydrodynamics, prescaler 3")
    rng = np.random.default_rng(3)
    shapes = [M.ico_ellipsoid(1.0, 0.4, 0.3, 3), M.box(0.5, 0.3, 0.2, 6), M.capsule(0.2, 1.5)]
    descs, obodies = [], []
    for i in range(6):
        tm = shapes[i % 3]
        descs.append(H.BodyDesc(H.Mesh(tm), volume=1.0, surface=2.0))
        obodies.append(O.make_body(O.Mesh.from_tri(tm), volume=1.0, surface=2.0))
    batch = H.BodyBatch(descs)
    ocean = H.Ocean()
    comps = W.pm_components(16, 6.0, 64.0, seed=3)
    grid = W.grid_reference(comps, 64).astype(np.float32)
    ocean.set_waves(grid, 64, 1, [64.0])
    oocean = O.Ocean(wave_data=grid, wave_n=64, channels=1, lengths=[64.0])

    def oracle_eval(i, s, prev):
        row, ex = O.body_forces(64, obodies[i], oocean, s[:16], s[16:19], s[19:22], init=prev.copy(), full=True)
        bnd = np.zeros(6)
        for j in range(6):
            raw = row[3 * j:3 * j + 3] if j < 2 else ex[6 + 3 * (j - 2):9 + 3 * (j - 2)]
            b = RTOL * np.linalg.norm(raw) + STOL * ex[j]
            if j >= 2 and ex[18 + (j - 2)] != 1.0:   # through CorrectHydrodynamicForces
                b *= ex[18 + (j - 2)] + 2.0 * (0.5 if j < 4 else 1.0) * 1000.0 * np.linalg.norm(np.ones(3) if j < 4 else 0.1 * np.ones(3))
            bnd[j] = b
        return row, bnd
    depths = np.array([-0.2, 0.1, 3.0, 0.0, -3.0, 0.3])
    states = [_states(rng, 6, depths + rng.uniform(-0.1, 0.1, 6)) for _ in range(8)]

    def overlap(k):
        ov = np.ones(6, bool)
        ov[5] = False
        if 1 <= k <= 4:
            ov[2] = False
        return ov
    _run_ticks(batch, ocean, 6, 3, 8, overlap, lambda k: states[k], aero=False, oracle_eval=oracle_eval)


def test_tick_driver_aero(gpu):
    """An aerodynamic batch under the driver applies Fda, Tda; prescaler 1
    recomputes every tick."""
    log("case: tick driver, aerodynamics, prescaler 1")
    rng = np.random.default_rng(4)
    tms = [M.box(0.4, 0.3, 0.2, 4), M.ico_ellipsoid(0.5, 0.4, 0.3, 3), M.capsule(0.1, 0.8)]
    descs = [H.AeroDesc(H.Mesh(tms[0]), approx_type=2, approx_params=[0.3, 0.8, 0.0]),
             H.AeroDesc(H.Mesh(tms[1]), approx_type=3, approx_params=[0.5, 0.4, 0.3]),
             H.AeroDesc(H.Mesh(tms[2]), aero=False)]
    omeshes = [O.Mesh.from_tri(t) for t in tms]
    batch = H.AeroBatch(descs)
    atm = H.Atmosphere(1.2)
    atm.add_uniform_current([3.0, 1.0, 0.0])
    wind = [(0, [3.0, 1.0, 0.0])]
    states = [_states(rng, 3, np.array([-10.0, -12.0, -8.0])) for _ in range(3)]

    def oracle_eval(i, s, prev):
        d = descs[i]
        row = prev.copy()
        bnd = np.zeros(6)
        if not d.aero:          # not aerodynamic: the mode gate leaves the row
            row[20], row[21] = 0.0, -1.0
            return row, bnd
        F, Tq, (SF, ST) = O.aero(omeshes[i], 1.2, wind, s[:16], s[:16], s[16:19], s[19:22], sums=True)
        Fc, Tc = HO.correct_aero(d.approx_type, d.approx_params, d.T_CG2H, F, Tq)
        row[6:9], row[9:12] = Fc, Tc
        row[20], row[21] = 12.0, 1.0
        cor = np.linalg.norm(Fc) / max(np.linalg.norm(F), 1e-300)
        bF = RTOL * np.linalg.norm(F) + STOL * SF
        bT = RTOL * np.linalg.norm(Tq) + STOL * ST
        bnd[2] = (cor + 1.5) * bF
        bnd[3] = cor * bT + 1.5 * np.linalg.norm(Tq) * bF / max(np.linalg.norm(F), 1e-300)
        return row, bnd
    _run_ticks(batch, atm, 3, 1, 3, lambda k: np.ones(3, bool), lambda k: states[k], aero=True,
               oracle_eval=oracle_eval)


def test_tick_driver_arguments(gpu):
    from stonefish_amd import _native as N
    batch = H.BodyBatch([H.BodyDesc(H.Mesh(M.box(0.3, 0.2, 0.1, 2)))])
    with pytest.raises(N.SfhError):
        H.TickDriver(batch, 0)

"""The benchmark workloads at their configured sizes on the GPU.

* Fleet: 4,096 bodies (ellipsoids and boxes, half-axes U[0.1, 2] m, face
  counts log-uniform in [500, 50,000], 44.0 M faces), CoB depth U[-1, 3] m,
  on the 1024 x 1024 Pierson-Moskowitz wave grid (1 m spacing), one captured
  step.  Every body against the FP64 oracle (the error bound of tests/util.py;
  positions and write masks exact); two replays bit-identical.
* Training data: 8 hulls x 8,192 bodies x 31 steps = 2,031,616 exact-kernel
  samples on that sea; a sample of the stored targets against the oracle; a
  short training run of the 8 networks cuts the held-out error.
* Mixed rollout: 16,384 AUV compounds (8,192-triangle hull and four
  1,024-triangle fairings) on the fleet's wave field at 1 kHz with audits:
  the exact rows of the out-of-distribution bodies equal the full exact batch
  bit for bit, and audit rows match the FP64 oracle's compound evaluation on a
  sample.
"""
import sys

import numpy as np
import pytest

import oracle as O
from stonefish_amd import transforms as T
from stonefish_amd import workloads as WL
from util import bound_ratio, component_errors, format_component_table

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def log(msg):
    print(f"[configs] {msg}", flush=True)
    sys.stdout.flush()


def _fleet_ocean(w):
    from stonefish_amd import hydro as H
    from stonefish_amd import waves as W
    n = w.sea.wave_n
    grid = torch.empty((n, n), dtype=torch.float32, device="cuda")
    W.build_grid_device(w.sea.wave_components, n, grid.data_ptr())
    go = H.Ocean()
    go.set_waves(None, n, 1, [w.sea.wave_len], device_ptr=grid.data_ptr())
    torch.cuda.synchronize()
    g = grid.cpu().numpy()
    oo = O.Ocean(wave_data=g, wave_n=n, channels=1, lengths=[w.sea.wave_len])
    return go, oo, grid


def test_fleet_4096_bodies_every_body_against_oracle(gpu):
    from stonefish_amd import hydro as H
    log("case: fleet of 4,096 bodies on the 1024^2 wave grid (bench workload, seed 7)")
    w = WL.fleet(4096, seed=7)
    nf = sum(b.mesh.nf for b in w.bodies)
    assert 43_000_000 < nf < 45_000_000 and w.sea.wave_n == 1024
    go, oo, grid = _fleet_ocean(w)
    gms = [H.Mesh(b.mesh) for b in w.bodies]
    descs = [H.BodyDesc(mesh=g, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface) for g, b in zip(gms, w.bodies)]
    batch = H.BodyBatch(descs)
    st = WL.random_states(w, 1)[0]
    d_st = torch.tensor(st, dtype=torch.float64, device="cuda")
    outs = []
    s = torch.cuda.Stream()
    d_out = torch.zeros((4096, 24), dtype=torch.float64, device="cuda")
    with torch.cuda.stream(s):
        batch.capture(go, d_st.data_ptr(), d_out.data_ptr(), s.cuda_stream)
    torch.cuda.synchronize()
    for _ in range(2):
        d_out.fill_(-3.0)
        torch.cuda.synchronize()
        batch.replay(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        outs.append(d_out.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]), "two replays differ"
    g = outs[0]
    # every body through the FP64 oracle, with its condition scales
    omeshes = [O.Mesh.from_tri(b.mesh) for b in w.bodies]
    obodies = [O.make_body(m, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface) for m, b in zip(omeshes, w.bodies)]
    rr, xx = [], []
    for i in range(4096):
        a, e = O.body_forces(64, obodies[i], oo, st[i, :16], st[i, 16:19], st[i, 19:22], full=True)
        rr.append(a)
        xx.append(e)
    r, xx = np.stack(rr), np.stack(xx)
    pos = np.bincount(r[:, 21].astype(int) + 1, minlength=4)[1:]
    log(f"positions (outside, inside, crossing): {pos.tolist()}")
    assert pos[1] > 500 and pos[2] > 500
    np.testing.assert_array_equal(g[:, 21], r[:, 21])
    np.testing.assert_array_equal(g[:, 20], r[:, 20])
    # untouched slots keep the caller's contents
    for j in range(8):
        untouched = ((r[:, 20].astype(int) >> j) & 1) == 0
        sl = slice(3 * j, 3 * j + 3) if j < 6 else slice(18 + j - 6, 19 + j - 6)
        assert np.all(g[untouched, sl] == -3.0)
    Cd = np.stack([b.Cd for b in w.bodies])
    Cf = np.stack([b.Cf for b in w.bodies])
    per_body = []
    for i in range(4096):
        wr, _ = bound_ratio(g[i:i + 1], r[i:i + 1], xx[i:i + 1], Cd=Cd[i:i + 1], Cf=Cf[i:i + 1])
        per_body.append(max(wr.values()))
    order = np.argsort(per_body)[::-1]
    for i in order[:5]:
        b = w.bodies[i]
        log(f"  body {i}: {b.mesh.nf} faces, position {int(r[i, 21])}, Vsub {r[i, 19]:.3e}, "
            f"worst bound ratio {per_body[i]:.2f}")
    worst, plain = bound_ratio(g, r, xx, Cd=Cd, Cf=Cf)
    largest = int(np.argmax([b.mesh.nf for b in w.bodies]))
    log(f"all 4,096 bodies (largest: {w.bodies[largest].mesh.nf} faces): bound ratio "
        + ", ".join(f"{k} {v:.2f}" for k, v in worst.items())
        + " | plain norm-rel " + ", ".join(f"{k} {v:.1e}" for k, v in plain.items()))
    log(format_component_table(component_errors(g, r)))
    for k, v in worst.items():
        assert v <= 1.0, (k, v)
    # every body: Swet and Vsub relative 1e-6 where written
    wm = r[:, 20].astype(int)
    for j, col in ((6, 18), (7, 19)):
        m = ((wm >> j) & 1) == 1
        np.testing.assert_allclose(g[m, col], r[m, col], rtol=1e-6, atol=1e-9)
    del batch


def test_training_data_2m_samples_8_hulls(gpu):
    from stonefish_amd import surrogate as S
    log("case: 8 hulls x 8,192 bodies x 31 steps of exact-kernel samples (2,031,616)")
    w = S.hull_set(8, 17)
    ocean, grid = S.make_ocean(w.sea)
    batch, meshes = S.make_hull_batch(w.bodies, 8192)
    ds = S.HullDataset(batch, ocean, 8, 8192, 31, 3, seed=18, xy_range=(0.0, w.sea.wave_len), depth_range=(-1.0, 3.0))
    assert ds.states.shape[0] * ds.states.shape[1] == 2_031_616
    log(f"generated in {ds.gen_seconds:.2f}s")
    # a sample of the stored exact outputs against the oracle
    g = grid.cpu().numpy()
    oo = O.Ocean(wave_data=g, wave_n=w.sea.wave_n, channels=1, lengths=[w.sea.wave_len])
    omeshes = [O.Mesh.from_tri(b.mesh) for b in w.bodies]
    obodies = [O.make_body(m, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface) for m, b in zip(omeshes, w.bodies)]
    rng = np.random.default_rng(3)
    picks = [(int(s), int(h), int(i)) for s, h, i in zip(rng.integers(0, 31, 48), rng.integers(0, 8, 48),
                                                         rng.integers(0, 8192, 48))]
    rows = torch.stack([ds.states[s, h * 8192 + i] for s, h, i in picks]).cpu().numpy()
    got = torch.stack([ds.exact[s, h * 8192 + i] for s, h, i in picks]).cpu().numpy()
    rr, xx = [], []
    for (s, h, i), st in zip(picks, rows):
        a, e = O.body_forces(64, obodies[h], oo, st[:16], st[16:19], st[19:22], full=True)
        rr.append(a)
        xx.append(e)
    rr, xx = np.stack(rr), np.stack(xx)
    np.testing.assert_array_equal(got[:, 20:22], rr[:, 20:22])
    Cd = np.stack([w.bodies[h].Cd for _, h, _ in picks])
    Cf = np.stack([w.bodies[h].Cf for _, h, _ in picks])
    worst, _ = bound_ratio(got, rr, xx, Cd=Cd, Cf=Cf)
    log("sampled targets: bound ratio " + ", ".join(f"{k} {v:.2f}" for k, v in worst.items()))
    for k, v in worst.items():
        assert v <= 1.0, (k, v)
    # a short run of the 8 networks side by side cuts the held-out error
    norms = [ds.fit_norm(h) for h in range(8)]
    xs, ts = [], []
    for h in range(8):
        x, t, _ = ds.inputs(h, norms[h], range(ds.train_steps))
        xs.append(x)
        ts.append(t)
    nets = [S.Net(65536, seed=h) for h in range(8)]
    before = [S.evaluate(ds, h, nets[h], norms[h]) for h in range(8)]
    S.train_many(nets, xs, ts, S.TrainConfig(steps=400, batch=16384, lr=3e-3, warmup=40))
    after = [S.evaluate(ds, h, nets[h], norms[h]) for h in range(8)]
    fb = np.median([b["F_median"] for b in before])
    fa = np.median([a["F_median"] for a in after])
    log(f"held-out F median over hulls: untrained {fb:.3f}, after 400 steps {fa:.4f}")
    assert fa < 0.25 * fb
    for n in nets:
        n.close()
    del xs, ts, ds, batch


def test_rollout_16384_auv_compounds(gpu):
    from stonefish_amd import hydro as H
    from stonefish_amd import surrogate as S
    log("case: mixed rollout of 16,384 AUV compounds on the fleet's wave field")
    auv = WL.auv_compound()
    gm, om = {}, {}
    parts_g, parts_o = [], []
    for p in auv.parts:
        if id(p.mesh) not in gm:
            gm[id(p.mesh)] = H.Mesh(p.mesh)
            om[id(p.mesh)] = O.Mesh.from_tri(p.mesh)
        parts_g.append(H.PartDesc(mesh=gm[id(p.mesh)], origin=p.origin, Cd=p.Cd, Cf=p.Cf, external=p.external,
                                  buoyant=p.buoyant))
        solid = O.make_body(om[id(p.mesh)], Cd=p.Cd, Cf=p.Cf, volume=p.mesh.volume(), surface=p.mesh.area(),
                            buoyant=p.buoyant)
        parts_o.append((solid, p.origin, T.identity_gl(), p.external))
    assert sum(p.mesh.nf for p in auv.parts) == 12_288
    body = H.BodyDesc(mesh=None, aabb_min=auv.aabb_min, aabb_max=auv.aabb_max, volume=auv.volume,
                      surface=auv.surface, P_CB=auv.P_CB)
    n = 16384
    batch = H.CompoundBatch([H.CompoundDesc(body, parts_g)] * n)
    w = WL.fleet(1, seed=7)
    ocean, grid = S.make_ocean(w.sea)
    # a network trained briefly on exact samples of this compound
    ds = S.HullDataset(batch, ocean, 1, n, steps=3, holdout_steps=1, seed=5, xy_range=(0.0, w.sea.wave_len),
                       depth_range=(-1.0, 3.0))
    norm = ds.fit_norm(0)
    x, t, _ = ds.inputs(0, norm, range(ds.train_steps))
    net = S.Net(65536, seed=2)
    S.train(net, x, t, S.TrainConfig(steps=300, batch=8192, warmup=20))
    rng = np.random.default_rng(11)
    states = torch.from_numpy(S.sample_states(rng, n, xy_range=(0.0, w.sea.wave_len), depth_range=(-0.5, 2.5))).cuda()
    ro = S.MixedRollout(batch, ocean, net, norm, states, auv.mass, auv.inertia, dt=1e-3, audit_every=3)
    for _ in range(5):
        ro.step()
    torch.cuda.synchronize()
    assert torch.isfinite(states).all()
    summ = ro.audit_summary()
    log(f"audits: {summ}")
    assert summ["audits"] == 2
    # out-of-distribution rows of a masked exact pass equal the full exact batch
    plan_ft = torch.zeros((n, 6), dtype=torch.float64, device="cuda")
    ood = torch.zeros(n, dtype=torch.uint8, device="cuda")
    pos = torch.zeros(n, dtype=torch.int8, device="cuda")
    ro.plan.run(ocean, states, plan_ft, ood=ood, pos=pos)
    mask = ood.clone()
    mask[::97] = 1   # make sure some rows take the exact path
    full = torch.zeros((n, 24), dtype=torch.float64, device="cuda")
    part = torch.full((n, 24), -5.0, dtype=torch.float64, device="cuda")
    batch.compute(ocean, states.data_ptr(), full.data_ptr())
    batch.compute_masked(ocean, states.data_ptr(), mask.data_ptr(), part.data_ptr())
    torch.cuda.synchronize()
    m = mask.cpu().numpy().astype(bool)
    f, p = full.cpu().numpy(), part.cpu().numpy()
    log(f"out of distribution: {int(ood.sum())} of {n}; exact-path rows checked: {int(m.sum())}")
    assert np.array_equal(p[m, :22], f[m, :22]) and np.all(p[~m] == -5.0)
    # audit rows (the full exact batch of this state) against the FP64 oracle
    dummy = O.Mesh(np.zeros((3, 3), np.float32), np.zeros((1, 3), np.uint32))
    ocomp = O.make_body(dummy, aabb_min=list(auv.aabb_min), aabb_max=list(auv.aabb_max), volume=auv.volume,
                        surface=auv.surface, P_CB=list(auv.P_CB))
    g = grid.cpu().numpy()
    oo = O.Ocean(wave_data=g, wave_n=w.sea.wave_n, channels=1, lengths=[w.sea.wave_len])
    st = states.cpu().numpy()
    counts = np.zeros(3, int)
    L = 3.0
    for i in range(0, n, 512):
        r = O.compound_forces(ocomp, parts_o, oo, st[i, :16], st[i, 16:19], st[i, 19:22], init=np.zeros(24))
        assert f[i, 21] == r[21] and f[i, 20] == r[20], i
        counts[int(r[21])] += 1
        u = np.linalg.norm(st[i, 16:19]) + np.linalg.norm(st[i, 19:22]) * L
        SF = {"b": 1000.0 * 9.81 * auv.volume, "d": 1000.0 * u ** 3 * auv.surface}
        for j in range(6):
            if not (int(r[20]) >> j) & 1:
                continue
            sl = slice(3 * j, 3 * j + 3)
            S_ = SF["b" if j < 2 else "d"] * (L if j % 2 else 1.0)
            assert np.linalg.norm(f[i, sl] - r[sl]) <= 1e-6 * np.linalg.norm(r[sl]) + 1e-7 * S_, (i, j)
    log(f"oracle sample positions (outside, inside, crossing): {counts.tolist()}")
    assert counts[1] + counts[2] > 0
    net.close()

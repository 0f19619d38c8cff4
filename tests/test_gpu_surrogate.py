"""GPU tests of the surrogate: the MFMA layout, the forward / backward / Adam
kernels against the FP32 reference (tests/mlp_ref.py), the feature and decode
kernels against the oracle, the masked exact fallback, and a training run
that must beat the untrained network on held-out samples."""
import numpy as np
import pytest

import mlp_ref as R

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _net(max_batch=1024, seed=0):
    from stonefish_amd import surrogate as S
    return S.Net(max_batch, seed)


def _load(net, Ws, bs):
    for l in range(5):
        net.set_layer(l, Ws[l].numpy(), bs[l].numpy())


def test_mfma_operand_layout(gpu):
    from stonefish_amd import surrogate as S
    g = torch.Generator().manual_seed(1)
    a = torch.randn(32, 16, generator=g).to(torch.bfloat16)
    b = torch.randn(16, 32, generator=g).to(torch.bfloat16)
    c = torch.zeros(32, 32, dtype=torch.float32, device="cuda")
    da, db = a.cuda(), b.cuda()
    S.probe_mfma(da, db, c)
    ref = a.float() @ b.float()
    assert torch.allclose(c.cpu(), ref, rtol=1e-5, atol=1e-5)


def test_forward_matches_reference(gpu):
    net = _net(1024)
    Ws, bs = R.init_weights(11)
    _load(net, Ws, bs)
    x, _ = R.random_inputs(12, 992)   # a multiple of 32, not of 256
    y = torch.zeros(992, 6, dtype=torch.float32, device="cuda")
    net.forward(x.to(torch.bfloat16).cuda(), y)
    torch.cuda.synchronize()
    yr, _ = R.forward(Ws, bs, x, rounded=True)
    y0, _ = R.forward(Ws, bs, x, rounded=False)
    yc = y.cpu()
    scale = yr[:, :6].abs().max()
    # same BF16 rounding points: only FP32 summation order differs
    assert (yc - yr[:, :6]).abs().max() <= 3e-3 * scale
    # and the rounding itself stays small against the unrounded network
    assert (yc - y0[:, :6]).norm() <= 2e-2 * y0[:, :6].norm()
    net.close()


def test_fused_group_and_layer_paths_agree(gpu):
    from stonefish_amd import surrogate as S
    nets = [_net(1024), _net(1024)]
    ws = [R.init_weights(51), R.init_weights(52)]
    for n, (Ws, bs) in zip(nets, ws):
        _load(n, Ws, bs)
    counts = [96, 160]   # 32-row tiles (not multiples of 64)
    x, _ = R.random_inputs(53, sum(counts))
    dx = x.to(torch.bfloat16).cuda()
    y = torch.zeros(sum(counts), 6, dtype=torch.float32, device="cuda")
    S.forward_group(nets, counts, dx, y)
    yl = torch.zeros_like(y)
    nets[0].forward_layers(dx[:96], yl[:96])
    nets[1].forward_layers(dx[96:], yl[96:])
    torch.cuda.synchronize()
    yr = torch.cat([R.forward(*ws[0], x[:96])[0][:, :6], R.forward(*ws[1], x[96:])[0][:, :6]])
    scale = yr.abs().max()
    assert (y.cpu() - yr).abs().max() <= 3e-3 * scale
    assert (yl.cpu() - yr).abs().max() <= 3e-3 * scale
    for n in nets:
        n.close()


@pytest.mark.parametrize("batch", [512, 16384])   # 2 and 32/64 batch splits of the weight gradients
def test_gradients_match_reference(gpu, batch):
    net = _net(max(1024, batch))
    Ws, bs = R.init_weights(21)
    _load(net, Ws, bs)
    x, t = R.random_inputs(22, batch)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    net.train_step(x.to(torch.bfloat16).cuda(), t.cuda().contiguous(), loss, lr=0.0)
    torch.cuda.synchronize()
    y, cache = R.forward(Ws, bs, x, rounded=True)
    lref, dy = R.loss_and_grad_out(y, t)
    dWs, dbs = R.backward(Ws, cache, dy, rounded=True)
    y0, cache0 = R.forward(Ws, bs, x, rounded=False)
    _, dy0 = R.loss_and_grad_out(y0, t)
    dW0, db0 = R.backward(Ws, cache0, dy0, rounded=False)
    assert abs(float(loss.item()) - float(lref)) <= 1e-3 * float(lref)
    for l in range(5):
        gw, gb = net.gradients(l)
        gw, gb = gw.cpu(), gb.cpu()
        rw = float((gw - dWs[l]).norm() / dWs[l].norm())
        rb = float((gb - dbs[l]).norm() / dbs[l].norm())
        assert rw < 1e-2 and rb < 1e-2, (l, rw, rb)
        rw0 = float((gw - dW0[l]).norm() / dW0[l].norm())
        assert rw0 < 6e-2, (l, rw0)
    # padding stays out of the gradients
    gw0, _ = net.gradients(0)
    assert torch.all(gw0[:, 13:] == 0)
    gw4, gb4 = net.gradients(4)
    assert torch.all(gw4[6:] == 0) and torch.all(gb4[6:] == 0)
    net.close()


def test_adam_matches_torch(gpu):
    net = _net(512)
    Ws, bs = R.init_weights(31)
    _load(net, Ws, bs)
    params = [p.clone().requires_grad_() for pair in zip(Ws, bs) for p in pair]
    opt = torch.optim.Adam(params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    for step in range(3):
        x, t = R.random_inputs(40 + step, 256)
        net.train_step(x.to(torch.bfloat16).cuda(), t.cuda().contiguous(), loss, lr=1e-3)
        torch.cuda.synchronize()
        for l in range(5):
            gw, gb = net.gradients(l)
            params[2 * l].grad = gw.cpu()
            params[2 * l + 1].grad = gb.cpu()
        opt.step()
        for l in range(5):
            w, b = net.layer(l)
            assert np.allclose(w, params[2 * l].detach().numpy(), rtol=1e-5, atol=2e-7), (step, l)
            assert np.allclose(b, params[2 * l + 1].detach().numpy(), rtol=1e-5, atol=2e-7), (step, l)
    w0, _ = net.layer(0)
    w4, b4 = net.layer(4)
    assert np.all(w0[:, 13:] == 0) and np.all(w4[6:] == 0) and np.all(b4[6:] == 0)
    net.close()


def test_training_fits_a_smooth_map(gpu):
    from stonefish_amd import surrogate as S
    net = _net(4096, seed=5)
    g = torch.Generator().manual_seed(9)
    n = 16384
    xin = torch.rand(n, 13, generator=g) * 2 - 1
    tgt = torch.stack([torch.sin(2 * xin[:, 0]) + xin[:, 1] * xin[:, 2], xin[:, 3] ** 2 - 0.3,
                       torch.cos(xin[:, 4] + xin[:, 5]), xin[:, 6] * xin[:, 7] * xin[:, 8],
                       torch.tanh(3 * xin[:, 9]), xin[:, 10] - xin[:, 11] + xin[:, 12]], 1)
    x = torch.zeros(n, 32)
    x[:, :13] = xin
    dx = x.to(torch.bfloat16).cuda()
    dt = tgt.cuda().contiguous()
    y = S.predict(net, dx)
    before = float(((y - dt) ** 2).mean())
    S.train(net, dx, dt, S.TrainConfig(steps=1500, batch=1024, lr=2e-3, warmup=50))
    y = S.predict(net, dx)
    after = float(((y - dt) ** 2).mean())
    assert after < 0.02 * before and after < 5e-3, (before, after)
    net.close()


# ---- feature build, decode and masked fallback (libsfhydro) ------------------------

def _small_fleet(waves=True, n_hulls=2, per_hull=256):
    from stonefish_amd import hydro as H
    from stonefish_amd import meshes as M
    from stonefish_amd import waves as W
    from stonefish_amd import workloads as WL
    hulls = [M.ellipsoid_with_faces(1.0, 0.4, 0.3, 800), M.box_with_faces(0.6, 0.5, 0.3, 600)][:n_hulls]
    meshes = [H.Mesh(m) for m in hulls]
    descs = []
    for h, (m, dm) in enumerate(zip(hulls, meshes)):
        bd = WL.body_from_mesh(m, np.ptp(m.verts, axis=0) / 2)
        for _ in range(per_hull):
            descs.append(H.BodyDesc(mesh=dm, Cd=bd.Cd, Cf=bd.Cf, volume=bd.volume, surface=bd.surface,
                                    P_CB=np.array([0.05, -0.02, 0.01])))
    ocean = H.Ocean()
    grid = None
    comps = None
    if waves:
        comps = W.pm_components(16, 12.0, 256.0, seed=3)
        grid = torch.empty(256 * 256, dtype=torch.float32, device="cuda")
        W.build_grid_device(comps, 256, grid.data_ptr())
        torch.cuda.synchronize()
        ocean.set_waves(grid.cpu().numpy(), 256, 1, [256.0])
    ocean.add_uniform_current([0.3, -0.2, 0.05])
    ocean.enable_currents(True)
    batch = H.BodyBatch(descs)
    return batch, ocean, hulls, grid


def test_features_match_oracle(gpu):
    import oracle as O
    from stonefish_amd import surrogate as S
    batch, ocean, hulls, grid = _small_fleet()
    rng = np.random.default_rng(4)
    states = S.sample_states(rng, batch.n, xy_range=(0.0, 256.0))
    ds = torch.from_numpy(states).cuda()
    exact = torch.zeros(batch.n, 24, dtype=torch.float64, device="cuda")
    batch.compute(ocean, ds.data_ptr(), exact.data_ptr())
    raw = torch.zeros(batch.n, 13, dtype=torch.float32, device="cuda")
    rt = torch.zeros(batch.n, 6, dtype=torch.float32, device="cuda")
    S.features(batch, ocean, ds, 0, batch.n, raw=raw, exact=exact, raw_target=rt)
    torch.cuda.synchronize()
    raw, rt, ex = raw.cpu().numpy(), rt.cpu().numpy(), exact.cpu().numpy()
    g = grid.cpu().numpy()
    ocn = O.Ocean(wave_data=g, wave_n=256, channels=1, lengths=[256.0])
    lib = O.load()
    u = np.array([0.3, -0.2, 0.05])
    for i in range(0, batch.n, 7):
        s = states[i]
        Rm = np.array(s[:16]).reshape(4, 4).T[:3, :3]   # OpenGL column-major -> rows
        t = s[12:15]
        exp = np.zeros(13)
        exp[0:3] = Rm.T @ (s[16:19] - u)
        exp[3:6] = Rm.T @ s[19:22]
        exp[6:9] = Rm.T @ [0.0, 0.0, 1.0]          # gravity (NED, +z) in the body frame
        cb = Rm @ np.array([0.05, -0.02, 0.01]) + t
        # local surface: plane fitted over the footprint of the body's world AABB
        hx = lambda x, y: -lib.orc_wave_height_64(C_ocean(ocn), x, y)   # noqa: E731
        half = 0.5 * np.ptp(hulls[i // 256].verts.astype(np.float64), axis=0)
        ex, ey = max(np.abs(Rm[0]) @ half, 1.0), max(np.abs(Rm[1]) @ half, 1.0)
        wts = np.outer([0.25, 0.5, 0.25], [0.25, 0.5, 0.25])
        H = np.array([[hx(cb[0] + (a - 1) * ex, cb[1] + (c - 1) * ey) for c in range(3)] for a in range(3)])
        exp[9] = cb[2] + (wts * H).sum()
        gx = (wts * H * (np.arange(3) - 1)[:, None]).sum() / (0.5 * ex)
        gy = (wts * H * (np.arange(3) - 1)[None, :]).sum() / (0.5 * ey)
        nw = np.array([gx, gy, 1.0]) / np.sqrt(1 + gx * gx + gy * gy)   # grad(depth) direction
        exp[10:13] = Rm.T @ nw
        assert np.allclose(raw[i], exp, rtol=2e-6, atol=2e-6), (i, raw[i], exp)
    # the targets are checked by their round trip through the decode (test_masked_compute_and_apply)


def C_ocean(o):
    import ctypes as C
    return C.byref(o.s)


def test_normalised_inputs_and_ood(gpu):
    from stonefish_amd import surrogate as S
    batch, ocean, hulls, grid = _small_fleet(waves=False, n_hulls=1)
    rng = np.random.default_rng(8)
    states = S.sample_states(rng, batch.n, xy_range=(0.0, 10.0), depth_range=(0.0, 0.5))   # all in the water
    states[3, 16] = 50.0          # far outside any training range
    states[5, 17] = float("nan")  # non-finite velocity
    ds = torch.from_numpy(states).cuda()
    raw = torch.zeros(batch.n, 13, dtype=torch.float32, device="cuda")
    S.features(batch, ocean, ds, 0, batch.n, raw=raw)
    r = raw.cpu().numpy()
    ok = np.isfinite(r).all(1) & (np.abs(r[:, 0:3]).max(1) < 10)
    norm = S.Norm.fit(r[ok], np.ones((ok.sum(), 6)))
    rows = batch.n + 32
    x = torch.full((rows, 32), 7.0, dtype=torch.bfloat16, device="cuda")
    ood = torch.zeros(batch.n, dtype=torch.uint8, device="cuda")
    S.features(batch, ocean, ds, 0, batch.n, norm=norm, rows=rows, x=x, ood=ood)
    torch.cuda.synchronize()
    xo = x.float().cpu().numpy()
    o = ood.cpu().numpy()
    assert o[3] == 1 and o[5] == 1 and o[ok].sum() == 0
    expect = torch.from_numpy(((r[ok] - norm.in_shift) * norm.in_scale).astype(np.float32)).to(torch.bfloat16).float()
    assert np.array_equal(xo[:batch.n][ok][:, :13], expect.numpy())
    assert np.all(xo[:, 13:] == 0) and np.all(xo[batch.n:] == 0)
    # a jet makes every body in the water out of distribution (the features
    # carry the uniform current only)
    ocean.add_jet([5.0, 5.0, 0.2], [1.0, 0.0, 0.0], 0.5, 1.0)
    pos = torch.zeros(batch.n, dtype=torch.int8, device="cuda")
    S.features(batch, ocean, ds, 0, batch.n, norm=norm, rows=rows, x=x, ood=ood, pos=pos)
    torch.cuda.synchronize()
    wet = pos.cpu().numpy() > 0
    assert wet.sum() > 0 and np.all(ood.cpu().numpy()[wet] == 1)


def test_masked_compute_and_apply(gpu):
    from stonefish_amd import _native as N
    from stonefish_amd import surrogate as S
    batch, ocean, hulls, grid = _small_fleet()
    rng = np.random.default_rng(12)
    states = S.sample_states(rng, batch.n, xy_range=(0.0, 256.0))
    ds = torch.from_numpy(states).cuda()
    full = torch.full((batch.n, 24), 123.0, dtype=torch.float64, device="cuda")   # untouched outputs stay 123
    batch.compute(ocean, ds.data_ptr(), full.data_ptr())
    mask = torch.from_numpy((rng.random(batch.n) < 0.3).astype(np.uint8)).cuda()
    part = torch.full((batch.n, 24), 123.0, dtype=torch.float64, device="cuda")
    import ctypes as C
    N.check(N.load().sfh_batch_compute_masked(batch.handle, ocean.handle, C.c_void_p(ds.data_ptr()),
                                              C.c_void_p(mask.data_ptr()), C.c_void_p(part.data_ptr()),
                                              C.c_void_p(0)))
    torch.cuda.synchronize()
    m = mask.cpu().numpy().astype(bool)
    f, p = full.cpu().numpy(), part.cpu().numpy()
    assert np.array_equal(p[m, :22], f[m, :22])   # deterministic: identical rows (spare columns unwritten)
    assert np.all(p[~m] == 123.0)              # inactive rows untouched
    # decode: feed the exact normalised targets as network outputs
    raw = torch.zeros(batch.n, 13, dtype=torch.float32, device="cuda")
    rt = torch.zeros(batch.n, 6, dtype=torch.float32, device="cuda")
    S.features(batch, ocean, ds, 0, batch.n, raw=raw, exact=full, raw_target=rt)
    norm = S.Norm.fit(raw.cpu().numpy(), rt.cpu().numpy())
    y = torch.zeros(batch.n, 6, dtype=torch.float32, device="cuda")
    S.features(batch, ocean, ds, 0, batch.n, norm=norm, exact=full, target=y)
    ft = torch.zeros(batch.n, 6, dtype=torch.float64, device="cuda")
    err = torch.zeros(batch.n, 2, dtype=torch.float32, device="cuda")
    ood = mask
    S.apply(batch, ocean, ds, 0, batch.n, norm, y, ft, ood=ood, exact=full, err=err)
    torch.cuda.synchronize()
    e = err.cpu().numpy()
    fo = ft.cpu().numpy()
    w = f[:, 20].astype(np.int64)
    Fe = sum(np.where(((w >> (2 * k)) & 1)[:, None] == 1, f[:, 6 * k:6 * k + 3], 0.0) for k in range(3))
    Te = sum(np.where(((w >> (2 * k + 1)) & 1)[:, None] == 1, f[:, 6 * k + 3:6 * k + 6], 0.0) for k in range(3))
    assert np.array_equal(fo[m, :3], Fe[m]) and np.array_equal(fo[m, 3:], Te[m])   # fallback rows: exact
    nF = np.linalg.norm(Fe, axis=1)
    big = nF > 0.05 * nF.max()
    assert np.allclose(fo[~m & big, :3], Fe[~m & big], rtol=1e-5, atol=1e-5 * nF.max())
    assert np.all(e[big, 0] < 1e-5)


def test_surrogate_beats_untrained_on_held_out(gpu):
    from stonefish_amd import surrogate as S
    batch, ocean, hulls, grid = _small_fleet(waves=True, n_hulls=1, per_hull=2048)
    ds = S.HullDataset(batch, ocean, 1, 2048, steps=10, holdout_steps=2, seed=1, xy_range=(0.0, 256.0),
                       depth_range=(-0.5, 1.0))
    norm = ds.fit_norm(0)
    x, t, ood = ds.inputs(0, norm, range(ds.train_steps))
    net = S.Net(4096, seed=3)
    untrained = S.evaluate(ds, 0, net, norm)
    S.train(net, x, t, S.TrainConfig(steps=1500, batch=1024, lr=2e-3, warmup=50))
    trained = S.evaluate(ds, 0, net, norm)
    print(f"[surrogate] held-out F median {trained['F_median']:.4f} (untrained {untrained['F_median']:.4f}), "
          f"T median {trained['T_median']:.4f}", flush=True)
    assert trained["n"] >= 0.9 * (2 * 2048 - trained["n_out_of_water"] - trained["n_ood"]), trained
    # a small data set (16K samples) and a short run: the gain, not the final accuracy
    assert trained["F_median"] < 0.25 * untrained["F_median"], (trained, untrained)
    assert trained["T_median"] < 0.5 * untrained["T_median"], (trained, untrained)
    net.close()


def test_error_aware_gate(gpu):
    """gate_band flags exactly the bodies crossing the surface whose depth
    feature is below in_lo[9] + gate_band (FP32, as the kernel compares), in
    sfh_surrogate_features and in the one-launch plan; 0 turns it off."""
    from stonefish_amd import surrogate as S
    batch, ocean, hulls, grid = _small_fleet(waves=True, n_hulls=2, per_hull=256)
    rng = np.random.default_rng(41)
    st = torch.from_numpy(S.sample_states(rng, batch.n, xy_range=(0.0, 256.0), depth_range=(-0.5, 1.5))).cuda()
    raw = torch.zeros(batch.n, 13, dtype=torch.float32, device="cuda")
    pos = torch.zeros(batch.n, dtype=torch.int8, device="cuda")
    S.features(batch, ocean, st, 0, batch.n, raw=raw, pos=pos)
    torch.cuda.synchronize()
    r, p = raw.cpu().numpy(), pos.cpu().numpy()
    norm = S.Norm.fit(r, np.ones((batch.n, 6)))   # every body inside the training range
    cross = p == 2
    assert cross.sum() > 20 and (p == 1).sum() > 20
    norm.gate_band = float(np.median(r[cross, 9])) - float(norm.in_lo[9])
    expect = cross & (r[:, 9] < np.float32(norm.in_lo[9]) + np.float32(norm.gate_band))
    assert 0 < expect.sum() < cross.sum()
    rows = (batch.n + 31) // 32 * 32
    x = torch.zeros(rows, 32, dtype=torch.bfloat16, device="cuda")
    ood = torch.zeros(batch.n, dtype=torch.uint8, device="cuda")
    S.features(batch, ocean, st, 0, batch.n, norm=norm, rows=rows, x=x, ood=ood)
    torch.cuda.synchronize()
    assert np.array_equal(ood.cpu().numpy().astype(bool), expect)
    plan = S.Plan(batch, [S.Net(1024, seed=4)], [norm], [batch.n])
    ft = torch.zeros(batch.n, 6, dtype=torch.float64, device="cuda")
    o2 = torch.zeros(batch.n, dtype=torch.uint8, device="cuda")
    plan.run(ocean, st, ft, ood=o2)
    torch.cuda.synchronize()
    assert torch.equal(o2, ood)
    norm.gate_band = 0.0
    S.features(batch, ocean, st, 0, batch.n, norm=norm, rows=rows, x=x, ood=ood)
    torch.cuda.synchronize()
    assert int(ood.sum()) == 0


def test_one_launch_step_matches_separate_kernels(gpu):
    """sfh_surrogate_run (features + network + decode in one launch, two
    groups with their own networks and norms) against features, the fused
    forward and the decode run separately; merge against the exact rows."""
    from stonefish_amd import surrogate as S
    batch, ocean, hulls, grid = _small_fleet(waves=True, n_hulls=2, per_hull=256)
    rng = np.random.default_rng(31)
    st = torch.from_numpy(S.sample_states(rng, batch.n, xy_range=(0.0, 256.0), depth_range=(-0.5, 1.5))).cuda()
    exact = torch.full((batch.n, 24), 0.0, dtype=torch.float64, device="cuda")
    batch.compute(ocean, st.data_ptr(), exact.data_ptr())
    raw = torch.zeros(batch.n, 13, dtype=torch.float32, device="cuda")
    rt = torch.zeros(batch.n, 6, dtype=torch.float32, device="cuda")
    S.features(batch, ocean, st, 0, batch.n, raw=raw, exact=exact, raw_target=rt)
    norms = [S.Norm.fit(raw[:256].cpu().numpy(), rt[:256].cpu().numpy()),
             S.Norm.fit(raw[256:].cpu().numpy(), rt[256:].cpu().numpy())]
    # tighten the training range so that some units are out of distribution
    for nm in norms:
        nm.in_hi[0] = 1.0
    nets = [S.Net(1024, seed=4), S.Net(1024, seed=5)]
    counts = [224, 288]   # groups need not match the hulls or the tile size
    plan = S.Plan(batch, nets, norms_for(norms, counts), counts)
    ft = torch.zeros(batch.n, 6, dtype=torch.float64, device="cuda")
    ood = torch.zeros(batch.n, dtype=torch.uint8, device="cuda")
    pos = torch.zeros(batch.n, dtype=torch.int8, device="cuda")
    plan.run(ocean, st, ft, ood=ood, pos=pos)
    # separate path
    ref = torch.zeros_like(ft)
    off = 0
    for g, c in enumerate(counts):
        nm = norms_for(norms, counts)[g]
        rows = (c + 31) // 32 * 32
        x = torch.zeros(rows, 32, dtype=torch.bfloat16, device="cuda")
        o2 = torch.zeros(c, dtype=torch.uint8, device="cuda")
        p2 = torch.zeros(c, dtype=torch.int8, device="cuda")
        S.features(batch, ocean, st, off, c, norm=nm, rows=rows, x=x, ood=o2, pos=p2)
        y = torch.zeros(rows, 6, dtype=torch.float32, device="cuda")
        nets[g].forward(x, y)
        S.apply(batch, ocean, st, off, c, nm, y[:c], ref[off:off + c])
        assert torch.equal(o2, ood[off:off + c]) and torch.equal(p2, pos[off:off + c])
        off += c
    torch.cuda.synchronize()
    f, r = ft.cpu().numpy(), ref.cpu().numpy()
    scale = np.abs(r).max(0) + 1e-9
    assert np.all(np.abs(f - r) <= 1e-5 * scale + 1e-6 * np.abs(r)), np.abs(f - r).max(0)
    assert int(ood.sum()) > 0
    # merge: exact totals on the flagged units, errors against the exact rows
    err = torch.zeros(batch.n, 2, dtype=torch.float32, device="cuda")
    sur = ft.clone()
    plan.merge(exact, ft, ood=ood, err=err)
    torch.cuda.synchronize()
    o = ood.cpu().numpy().astype(bool)
    ex = exact.cpu().numpy()
    w = ex[:, 20].astype(np.int64)
    Fe = sum(np.where(((w >> (2 * k)) & 1)[:, None] == 1, ex[:, 6 * k:6 * k + 3], 0.0) for k in range(3))
    fm = ft.cpu().numpy()
    assert np.array_equal(fm[o, :3], Fe[o]) and np.array_equal(fm[~o], sur.cpu().numpy()[~o])
    e = err.cpu().numpy()
    nF = np.linalg.norm(Fe, axis=1)
    m = nF > 0
    d = np.linalg.norm(sur.cpu().numpy()[:, :3] - Fe, axis=1)
    np.testing.assert_allclose(e[m, 0], d[m] / nF[m], rtol=1e-5)
    plan.close()
    for n in nets:
        n.close()


def norms_for(norms, counts):
    return [norms[0], norms[1]][:len(counts)]


def test_hydrostatic_table_reproduces_static_buoyancy(gpu):
    """A hull at rest on a flat sea: with its hydrostatic table the
    surrogate's baseline is the exact static buoyancy up to the table's
    interpolation error, so the network's target (the residual in units of
    the body's scales) is small; without the table it is not."""
    from stonefish_amd import surrogate as S
    batch, ocean, hulls, grid = _small_fleet(waves=False, n_hulls=1, per_hull=256)
    rng = np.random.default_rng(41)
    st = S.sample_states(rng, batch.n, xy_range=(0.0, 10.0), depth_range=(-0.2, 0.2), v_max=0.0, w_max=0.0)
    ds = torch.from_numpy(st).cuda()
    exact = torch.zeros(batch.n, 24, dtype=torch.float64, device="cuda")
    batch.compute(ocean, ds.data_ptr(), exact.data_ptr())
    def residual():
        rt = torch.zeros(batch.n, 6, dtype=torch.float32, device="cuda")
        pos = torch.zeros(batch.n, dtype=torch.int8, device="cuda")
        S.features(batch, ocean, ds, 0, batch.n, pos=pos, exact=exact, raw_target=rt)
        p = pos.cpu().numpy()
        return rt.cpu().numpy()[p == 2]
    before = np.abs(residual()[:, :3]).max()
    tables = S.attach_tables(batch, ocean, [(0, batch.n)], ns=33, nth=17, nph=32)
    r = residual()
    assert r.shape[0] > 50
    after = np.abs(r[:, :3]).max()
    print(f"[surrogate] static residual max {before:.4f} without table, {after:.5f} with", flush=True)
    assert after < 0.02 and after < 0.2 * before
    del tables

"""Rollout harness on the GPU: the rigid-body step against closed forms
(free fall, constant force, torque-free rotation of a symmetric body, angular
momentum of an asymmetric one) and a short mixed surrogate/exact rollout of
AUV compounds with audits."""
import numpy as np
import pytest

from stonefish_amd import transforms as T

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _gl_R(s):
    return s[:16].reshape(4, 4).T[:3, :3]


def _skew(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def _expm(w, t):
    th = np.linalg.norm(w) * t
    if th == 0:
        return np.eye(3)
    K = _skew(w / np.linalg.norm(w))
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


def test_rigid_step_closed_forms(gpu):
    from stonefish_amd import hydro as H
    from stonefish_amd import surrogate as S
    ocean = H.Ocean()
    n, dt, steps, m = 4, 1e-3, 500, 2.0
    rng = np.random.default_rng(1)
    q = T.random_unit_quaternions(rng, n)
    st = np.zeros((n, 22))
    st[:, :16] = T.quat_to_gl(q, np.zeros((n, 3)))
    st[1, 16:19] = [0.5, -0.2, 0.1]
    st[2, 19:22] = [0.3, -0.7, 0.4]
    st[3, 19:22] = [0.2, 0.1, 0.05]
    R0 = [_gl_R(st[i]).copy() for i in range(n)]
    ds = torch.from_numpy(st.copy()).cuda()
    ft = torch.zeros(n, 6, dtype=torch.float64, device="cuda")
    ft[1, 0] = 3.0   # constant force along x on body 1
    # bodies 0-2 share a symmetric inertia; body 3 gets its own asymmetric one
    for _ in range(steps):
        S.integrate(ocean, ds[:3], 3, ft[:3], m, (1.0, 1.0, 1.0), dt)
        S.integrate(ocean, ds[3:], 1, ft[3:], m, (1.0, 2.0, 3.0), dt)
    out = ds.cpu().numpy()
    t = steps * dt
    g = 9.81
    # free fall (body 0): v_z = g t exactly in steps, z = g dt^2 (1 + ... + steps)
    np.testing.assert_allclose(out[0, 18], g * t, rtol=1e-12)
    np.testing.assert_allclose(out[0, 14], g * dt * dt * steps * (steps + 1) / 2, rtol=1e-12)
    # constant force (body 1)
    np.testing.assert_allclose(out[1, 16], 0.5 + 3.0 / m * t, rtol=1e-12)
    # torque-free symmetric body (body 2): omega constant, R(t) = exp([w] t) R0
    np.testing.assert_allclose(out[2, 19:22], [0.3, -0.7, 0.4], rtol=1e-10)
    np.testing.assert_allclose(_gl_R(out[2]), _expm(np.array([0.3, -0.7, 0.4]), t) @ R0[2], atol=1e-10)
    # asymmetric body: |L| and L (world) conserved to O(dt)
    for i in range(n):
        R = _gl_R(out[i])
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
    I = np.diag([1.0, 2.0, 3.0])
    L0 = R0[3] @ I @ R0[3].T @ st[3, 19:22]
    R = _gl_R(out[3])
    L1 = R @ I @ R.T @ out[3, 19:22]
    assert np.linalg.norm(L1 - L0) <= 2e-3 * np.linalg.norm(L0)


def test_mixed_rollout_runs_with_audits(gpu):
    from stonefish_amd import hydro as H
    from stonefish_amd import surrogate as S
    from stonefish_amd import workloads as WL
    auv = WL.auv_compound()
    meshes = {}
    parts = []
    for p in auv.parts:
        if id(p.mesh) not in meshes:
            meshes[id(p.mesh)] = H.Mesh(p.mesh)
        parts.append(H.PartDesc(mesh=meshes[id(p.mesh)], origin=p.origin, Cd=p.Cd, Cf=p.Cf))
    body = H.BodyDesc(mesh=None, aabb_min=auv.aabb_min, aabb_max=auv.aabb_max, volume=auv.volume,
                      surface=auv.surface, P_CB=auv.P_CB)
    n = 256
    batch = H.CompoundBatch([H.CompoundDesc(body, parts)] * n)
    ocean = H.Ocean()
    ds = S.HullDataset(batch, ocean, 1, n, steps=6, holdout_steps=1, seed=3, xy_range=(0.0, 100.0),
                       depth_range=(-1.0, 3.0))
    norm = ds.fit_norm(0)
    x, t, _ = ds.inputs(0, norm, range(ds.train_steps))
    net = S.Net(1024, seed=1)
    S.train(net, x, t, S.TrainConfig(steps=100, batch=256, warmup=10))
    rng = np.random.default_rng(9)
    init = torch.from_numpy(S.sample_states(rng, n, xy_range=(0.0, 100.0), depth_range=(0.5, 2.0))).cuda()
    # the same rollout eagerly and with the ordinary step replayed from a HIP graph
    runs = []
    for graph in (False, True):
        states = init.clone()
        ro = S.MixedRollout(batch, ocean, net, norm, states, auv.mass, auv.inertia, dt=1e-3, audit_every=10)
        if graph:
            ro.step()   # audit step 0 eagerly, then capture (the capture itself does not run)
            ro.capture()
            for _ in range(24):
                ro.step()
        else:
            for _ in range(25):
                ro.step()
        torch.cuda.synchronize()
        assert torch.isfinite(states).all()
        summ = ro.audit_summary()
        assert summ["audits"] == 3 and summ["n"] > 0
        runs.append(states.cpu().numpy())
    assert np.array_equal(runs[0], runs[1])
    assert not np.array_equal(runs[0], init.cpu().numpy())
    net.close()


def test_integrate_takes_the_exact_totals_of_flagged_bodies(gpu):
    """sfh_rollout_integrate with d_ood / d_exact: the flagged bodies' exact
    totals (fields with their written bit) replace their forces before the
    step, the others keep theirs; the result equals merging first."""
    from stonefish_amd import hydro as H
    from stonefish_amd import surrogate as S
    ocean = H.Ocean()
    n = 64
    rng = np.random.default_rng(5)
    st = np.zeros((n, 22))
    st[:, :16] = T.quat_to_gl(T.random_unit_quaternions(rng, n), rng.uniform(-1, 1, (n, 3)))
    st[:, 16:22] = rng.uniform(-1, 1, (n, 6))
    ft = rng.uniform(-100, 100, (n, 6))
    exact = np.zeros((n, 24))
    exact[:, :18] = rng.uniform(-100, 100, (n, 18))
    exact[:, 20] = rng.integers(0, 64, n)        # written bits of Fb Tb Fdq Tdq Fdf Tdf
    ood = (rng.random(n) < 0.3).astype(np.uint8)
    merged = ft.copy()
    for i in np.nonzero(ood)[0]:
        wm = int(exact[i, 20])
        F, Tq = np.zeros(3), np.zeros(3)
        for q in range(3):
            if wm & (1 << (2 * q)):
                F += exact[i, 6 * q:6 * q + 3]
            if wm & (1 << (2 * q + 1)):
                Tq += exact[i, 6 * q + 3:6 * q + 6]
        merged[i] = np.concatenate([F, Tq])
    a, b = torch.from_numpy(st.copy()).cuda(), torch.from_numpy(st.copy()).cuda()
    fa, fb = torch.from_numpy(ft.copy()).cuda(), torch.from_numpy(merged).cuda()
    S.integrate(ocean, a, n, fa, 2.0, (1.0, 2.0, 3.0), 1e-3, ood=torch.from_numpy(ood).cuda(),
                exact=torch.from_numpy(exact).cuda())
    S.integrate(ocean, b, n, fb, 2.0, (1.0, 2.0, 3.0), 1e-3)
    torch.cuda.synchronize()
    assert np.array_equal(fa.cpu().numpy(), merged)
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())

"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Bound per F/T vector (tests/util.py): |x - x64| <= 1e-6 |x64| + 2e-7 S with S
the sum of the norms of the per-face contributions, carried through
CorrectHydrodynamicForces for body-level outputs; Swet and Vsub relative 1e-6;
write masks and body/fluid positions exact.
"""
import sys

import numpy as np
import pytest

import oracle as O
from stonefish_amd import hydro as H
from stonefish_amd import meshes as M
from stonefish_amd import transforms as T
from stonefish_amd import waves as W
from stonefish_amd import workloads as WL
from util import NAMES, bound_ratio, component_errors, format_component_table

pytestmark = pytest.mark.gpu
RTOL = 1e-6


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def assert_bound(gpu, ref, extras, what="", names=NAMES, Cd=None, Cf=None):
    worst, plain = bound_ratio(gpu, ref, extras, names=names, Cd=Cd, Cf=Cf)
    log(f"{what}: bound ratio " + ", ".join(f"{k} {v:.2f}" for k, v in worst.items())
        + " | plain norm-rel " + ", ".join(f"{k} {v:.1e}" for k, v in plain.items()))
    log(format_component_table(component_errors(gpu, ref, names=names)))
    for k, v in worst.items():
        assert v <= 1.0, f"{what}: {k} error exceeds the bound by {v:.2f}x"


def test_submerged_capsule(gpu):
    log("case: submerged capsule, 32 random states")
    w = WL.capsule_submerged()
    tm = w.bodies[0].mesh
    gm, om = H.Mesh(tm), O.Mesh.from_tri(tm)
    go, oo = H.Ocean(), O.Ocean()
    st = WL.random_states(w, 32)[:, 0, :]
    g, r, x = np.zeros((32, 24)), np.zeros((32, 24)), np.zeros((32, 22))
    for s in range(32):
        Fdq, Tdq, Fdf, Tdf = (np.zeros(3) for _ in range(4))
        H.ComputeHydrodynamicForcesSubmerged(gm, go, st[s, :16], st[s, :16], st[s, 16:19], st[s, 19:22],
                                             Fdq, Tdq, Fdf, Tdf)
        g[s, 6:9], g[s, 9:12], g[s, 12:15], g[s, 15:18] = Fdq, Tdq, Fdf, Tdf
        r[s], x[s] = O.submerged(64, om, oo, st[s, :16], st[s, :16], st[s, 16:19], st[s, 19:22], full=True)
        r[s, 20] = 60   # drag vectors written
    assert_bound(g, r, x, "submerged capsule", names=("Fdq", "Tdq", "Fdf", "Tdf"))


@pytest.mark.parametrize("waves", [False, True])
def test_surface_ellipsoid(gpu, waves):
    log(f"case: surface ellipsoid ico5, waves={waves}")
    w = WL.ellipsoid_surface()
    tm = w.bodies[0].mesh
    gm, om = H.Mesh(tm), O.Mesh.from_tri(tm)
    go = H.Ocean()
    if waves:
        comps = W.pm_components(16, 6.0, 64.0, seed=3)
        grid = W.grid_reference(comps, 64).astype(np.float32)
        go.set_waves(grid, 64, 1, [64.0])
        oo = O.Ocean(wave_data=grid, wave_n=64, channels=1, lengths=[64.0])
    else:
        oo = O.Ocean()
    st = WL.random_states(w, 16)[:, 0, :]
    g, r, x = np.zeros((16, 24)), np.zeros((16, 24)), np.zeros((16, 22))
    settings = H.HydrodynamicsSettings(True, True)
    for s in range(16):
        outs = [np.zeros(3) for _ in range(6)]
        sv = np.zeros(2)
        H.ComputeHydrodynamicForcesSurface(settings, gm, go, st[s, :16], st[s, :16], st[s, 16:19], st[s, 19:22],
                                           *outs, sv)
        g[s, :18] = np.concatenate(outs)
        g[s, 18:20] = sv
        r[s], x[s] = O.surface(64, 1, 1, om, oo, st[s, :16], st[s, :16], st[s, 16:19], st[s, 19:22], full=True)
        assert r[s, 20] == 255
    assert_bound(g, r, x, f"surface ellipsoid waves={waves}")
    np.testing.assert_allclose(g[:, 18], r[:, 18], rtol=RTOL)
    np.testing.assert_allclose(g[:, 19], r[:, 19], rtol=RTOL)


def build_fleet(n, f_hi=4000, waves=True, seed=11):
    w = WL.fleet(n, seed=seed, f_lo=200, f_hi=f_hi, waves=waves, wave_n=128, wave_len=128.0)
    gms = [H.Mesh(b.mesh) for b in w.bodies]
    oms = [O.Mesh.from_tri(b.mesh) for b in w.bodies]
    go = H.Ocean()
    grid = None
    if waves:
        grid = W.grid_reference(w.sea.wave_components, w.sea.wave_n).astype(np.float32)
        go.set_waves(grid, w.sea.wave_n, 1, [w.sea.wave_len])
        oo = O.Ocean(wave_data=grid, wave_n=w.sea.wave_n, channels=1, lengths=[w.sea.wave_len])
    else:
        oo = O.Ocean()
    descs, obodies = [], []
    for b, gm, om in zip(w.bodies, gms, oms):
        descs.append(H.BodyDesc(mesh=gm, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface))
        obodies.append(O.make_body(om, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface))
    return w, go, oo, descs, obodies


@pytest.mark.parametrize("waves", [False, True])
def test_batch_fleet(gpu, waves):
    log(f"case: batch fleet 48 bodies, waves={waves}")
    w, go, oo, descs, obodies = build_fleet(48, waves=waves)
    batch = H.BodyBatch(descs)
    st = WL.random_states(w, 2)
    Cd = np.stack([b.Cd for b in w.bodies])
    Cf = np.stack([b.Cf for b in w.bodies])
    for s in range(2):
        g = batch.compute_host(go, st[s])
        rr = [O.body_forces(64, obodies[i], oo, st[s, i, :16], st[s, i, 16:19], st[s, i, 19:22], full=True)
              for i in range(len(obodies))]
        r = np.stack([a for a, _ in rr])
        x = np.stack([b for _, b in rr])
        np.testing.assert_array_equal(g[:, 21], r[:, 21])   # position
        np.testing.assert_array_equal(g[:, 20], r[:, 20])   # written mask
        log(f"positions (outside, inside, crossing): {np.bincount(r[:, 21].astype(int), minlength=3)}")
        assert_bound(g, r, x, f"fleet waves={waves} step {s}", Cd=Cd, Cf=Cf)
        np.testing.assert_allclose(g[:, 18], r[:, 18], rtol=RTOL)
        m = r[:, 20].astype(int) & 128 > 0
        np.testing.assert_allclose(g[m, 19], r[m, 19], rtol=RTOL)


def _jet_pipe_oceans(centre, waves):
    """The same jets, pipe and uniform current on the device ocean and the
    oracle's, laid across a body at `centre`; optionally waves.  The pipe's
    exponent is 1.5: with an exponent below 1 the profile (1 - d/r)^gamma has
    an unbounded slope at the wall, where a 1e-7 m change of a face centroid
    (the reference's own FP32 centroids) moves the velocity by ~1e-3 m/s, so
    no FP32 evaluation, the reference's included, meets the FP64 bound there."""
    c = np.asarray(centre, dtype=float)
    jet = (c + [-2.0, 0.1, 0.05], [1.0, 0.05, 0.0], 0.4, 3.0)
    jet2 = (c + [0.2, -2.5, 0.1], [0.0, 1.0, -0.1], 0.3, 2.0)
    pipe = (c + [0.3, -3.0, -0.2], c + [0.3, 3.0, 0.2], 0.8, 0.5, 1.5, 1.5)
    go = H.Ocean()
    go.add_uniform_current([0.2, -0.1, 0.0])
    go.add_jet(*jet)
    go.add_pipe(*pipe)
    go.add_jet(*jet2)
    go.enable_currents(True)
    cur = [(0, 1, [0.2, -0.1, 0.0])]
    for pt, d, r, v in (jet, jet2):
        d = np.asarray(d) / np.linalg.norm(d)
        cur.append((1, 1, list(pt) + list(d) + [r, v]))
    p1, p2, r1, r2, vin, gam = pipe
    ln = float(np.linalg.norm(p2 - p1))
    cur.insert(2, (2, 1, list(p1) + list((p2 - p1) / ln) + [ln, r1, r2, vin, gam]))
    kw = {}
    if waves:
        comps = W.pm_components(16, 6.0, 64.0, seed=3)
        grid = W.grid_reference(comps, 64).astype(np.float32)
        go.set_waves(grid, 64, 1, [64.0])
        kw = dict(wave_data=grid, wave_n=64, channels=1, lengths=[64.0])
    oo = O.Ocean(currents=cur, currents_enabled=True, **kw)
    return go, oo


def test_jets_and_pipes_submerged(gpu):
    """Jets and a pipe (evaluated per face centroid) over a submerged capsule."""
    log("case: submerged capsule in two jets, a pipe and a uniform current")
    w = WL.capsule_submerged()
    tm = w.bodies[0].mesh
    gm, om = H.Mesh(tm), O.Mesh.from_tri(tm)
    go, oo = _jet_pipe_oceans([0.0, 0.0, 5.0], False)
    rng = np.random.default_rng(17)
    n = 16
    g, r, x = np.zeros((n, 24)), np.zeros((n, 24)), np.zeros((n, 22))
    for s in range(n):
        q = T.random_unit_quaternions(rng, 1)[0]
        Tcg = T.quat_to_gl(q, [rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), 5.0 + rng.uniform(-0.3, 0.3)])
        v, om_ = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        Fdq, Tdq, Fdf, Tdf = (np.zeros(3) for _ in range(4))
        H.ComputeHydrodynamicForcesSubmerged(gm, go, Tcg, Tcg, v, om_, Fdq, Tdq, Fdf, Tdf)
        g[s, 6:9], g[s, 9:12], g[s, 12:15], g[s, 15:18] = Fdq, Tdq, Fdf, Tdf
        r[s], x[s] = O.submerged(64, om, oo, Tcg, Tcg, v, om_, full=True)
        r[s, 20] = 60
    # the currents change the result (they are not ignored)
    go0, oo0 = H.Ocean(), O.Ocean()
    r0, _ = O.submerged(64, om, oo0, Tcg, Tcg, v, om_, full=True)
    assert np.linalg.norm(r0[6:9] - r[n - 1, 6:9]) > 1e-3 * np.linalg.norm(r[n - 1, 6:9])
    assert_bound(g, r, x, "submerged capsule, jets and pipe", names=("Fdq", "Tdq", "Fdf", "Tdf"))


@pytest.mark.parametrize("waves", [False, True])
def test_jets_and_pipes_surface(gpu, waves):
    """Jets and a pipe over an ellipsoid crossing the surface (clipped faces
    take the current at their wet centroid)."""
    log(f"case: surface ellipsoid in two jets, a pipe and a uniform current, waves={waves}")
    w = WL.ellipsoid_surface()
    tm = w.bodies[0].mesh
    gm, om = H.Mesh(tm), O.Mesh.from_tri(tm)
    go, oo = _jet_pipe_oceans([0.0, 0.0, 0.1], waves)
    rng = np.random.default_rng(23 + int(waves))
    n = 12
    g, r, x = np.zeros((n, 24)), np.zeros((n, 24)), np.zeros((n, 22))
    settings = H.HydrodynamicsSettings(True, True)
    for s in range(n):
        q = T.random_unit_quaternions(rng, 1)[0]
        Tcg = T.quat_to_gl(q, [rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(-0.2, 0.2)])
        v, om_ = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        outs = [np.zeros(3) for _ in range(6)]
        sv = np.zeros(2)
        H.ComputeHydrodynamicForcesSurface(settings, gm, go, Tcg, Tcg, v, om_, *outs, sv)
        g[s, :18] = np.concatenate(outs)
        g[s, 18:20] = sv
        r[s], x[s] = O.surface(64, 1, 1, om, oo, Tcg, Tcg, v, om_, full=True)
    assert_bound(g, r, x, f"surface ellipsoid, jets and pipe, waves={waves}")
    np.testing.assert_allclose(g[:, 18], r[:, 18], rtol=RTOL)


def test_reference_wave_layout(gpu):
    """Two-layer RGBA texel layout of the reference's fftData (channels 0 and 1,
    sizes 893 m and 101 m, OpenGLOcean.cpp:70)."""
    log("case: reference RGBA 2-layer wave layout")
    rng = np.random.default_rng(5)
    n = 64
    data = (0.3 * rng.standard_normal((n, n, 4))).astype(np.float32)
    lengths = [893.0, 101.0]
    tm = M.ico_ellipsoid(1.0, 0.3, 0.25, 3)
    gm, om = H.Mesh(tm), O.Mesh.from_tri(tm)
    go = H.Ocean()
    go.set_waves(data, n, 4, lengths)
    oo = O.Ocean(wave_data=data, wave_n=n, channels=4, lengths=lengths)
    g, r, x = np.zeros((8, 24)), np.zeros((8, 24)), np.zeros((8, 22))
    for s in range(8):
        q = T.random_unit_quaternions(rng, 1)[0]
        Tcg = T.quat_to_gl(q, [rng.uniform(-900, 900), rng.uniform(-900, 900), rng.uniform(-0.2, 0.2)])
        v, om_ = rng.uniform(-2, 2, 3), rng.uniform(-1, 1, 3)
        outs = [np.zeros(3) for _ in range(6)]
        sv = np.zeros(2)
        H.ComputeHydrodynamicForcesSurface(H.HydrodynamicsSettings(), gm, go, Tcg, Tcg, v, om_, *outs, sv)
        g[s, :18] = np.concatenate(outs)
        g[s, 18:20] = sv
        r[s], x[s] = O.surface(64, 1, 1, om, oo, Tcg, Tcg, v, om_, full=True)
    assert_bound(g, r, x, "reference wave layout")


def test_untouched_outputs(gpu):
    """Buoyancy outputs stay untouched when the submerged volume is <= 1e-9
    (SolidEntity.cpp:1690); drag outputs stay untouched without damping."""
    log("case: untouched outputs")
    tm = M.icosphere(2)
    gm = H.Mesh(tm)
    go = H.Ocean()
    Tcg = T.translation_gl([0.0, 0.0, -5.0])   # fully dry: 5 m above the surface
    sentinel = [np.full(3, 7.0) for _ in range(6)]
    sv = np.array([3.0, 4.0])
    H.ComputeHydrodynamicForcesSurface(H.HydrodynamicsSettings(False, True), gm, go, Tcg, Tcg, [1, 0, 0],
                                       [0, 0, 0], *sentinel, sv)
    for k in range(6):
        assert np.all(sentinel[k] == 7.0), k
    assert sv[0] == 0.0 and sv[1] == 4.0


def test_batch_determinism_and_graph(gpu):
    import torch
    log("case: determinism + graph replay")
    w, go, oo, descs, obodies = build_fleet(24, waves=True, seed=3)
    batch = H.BodyBatch(descs)
    st = WL.random_states(w, 3)
    dev = torch.device("cuda:0")
    ds = torch.tensor(st[0], dtype=torch.float64, device=dev).contiguous()
    out1 = torch.zeros((24, 24), dtype=torch.float64, device=dev)
    out2 = torch.zeros_like(out1)
    stream = torch.cuda.current_stream().cuda_stream
    batch.compute(go, ds.data_ptr(), out1.data_ptr(), stream)
    batch.compute(go, ds.data_ptr(), out2.data_ptr(), stream)
    torch.cuda.synchronize()
    assert torch.equal(out1, out2)
    s = torch.cuda.Stream()
    gout = torch.zeros_like(out1)
    with torch.cuda.stream(s):
        batch.capture(go, ds.data_ptr(), gout.data_ptr(), s.cuda_stream)
    for rep in range(3):
        ds.copy_(torch.tensor(st[rep], dtype=torch.float64))
        eager = torch.zeros_like(out1)
        batch.compute(go, ds.data_ptr(), eager.data_ptr(), stream)
        torch.cuda.synchronize()
        gout.zero_()
        torch.cuda.synchronize()
        batch.replay(s.cuda_stream)
        s.synchronize()
        assert torch.equal(gout, eager), rep


def test_rows_independent_of_launch_shape(gpu):
    """The face launch's wavefront count (and with it the epilogue's, whose
    wavefronts finish several bodies at once or one at a time) and one or two
    face launches change the speed, not the bits (the faces per lane decide
    which faces a lane sums in FP32, so they are held fixed)."""
    log("case: launch shapes, 64 bodies up to 20,000 faces")
    w, go, oo, descs, obodies = build_fleet(64, f_hi=20000, waves=True, seed=5)
    batch = H.BodyBatch(descs, faces_per_lane=16)
    st = WL.random_states(w, 1)[0]
    ref = None
    for waves, split in ((64, 0), (1024, 0), (16384, 1), (4096, 1), (8, 0)):
        batch.set_opts(16, waves, split)
        g = batch.compute_host(go, st)
        if ref is None:
            ref = g
            assert np.bincount(g[:, 21].astype(int) + 1, minlength=4)[2:].min() > 0   # inside and crossing bodies
        else:
            assert np.array_equal(g.view(np.uint64), ref.view(np.uint64)), (waves, split)


def test_wave_grid_builder(gpu):
    import torch
    comps = W.pm_components(64, 12.0, 1024.0, seed=8)
    n = 256
    g = torch.empty((n, n), dtype=torch.float32, device="cuda:0")
    W.build_grid_device(comps, n, g.data_ptr())
    ref = W.grid_reference(comps, n)
    err = np.abs(g.cpu().numpy() - ref).max()
    log(f"wave grid max abs error {err:.2e}, rms height {ref.std():.3f}")
    assert err < 2e-5


def test_mesh_lifetime(gpu):
    """Regression: a mesh destroyed and a new one created (possibly at the same
    host address) must not reuse device pointers of the freed one; a batch keeps
    its meshes alive after sfh_mesh_destroy."""
    import gc
    log("case: mesh lifetime")
    oo, go = O.Ocean(), H.Ocean()
    Tcg = T.quat_to_gl([0.1, 0.2, 0.3, 0.9], [1.0, 2.0, 0.05])
    for sub in (1, 2, 3):
        tm = M.ico_ellipsoid(0.5 + 0.1 * sub, 0.3, 0.2, sub)
        gm = H.Mesh(tm)
        outs = [np.zeros(3) for _ in range(6)]
        sv = np.zeros(2)
        H.ComputeHydrodynamicForcesSurface(H.HydrodynamicsSettings(), gm, go, Tcg, Tcg, [1, 0.5, 0], [0, 0, 1],
                                           *outs, sv)
        r = O.surface(64, 1, 1, O.Mesh.from_tri(tm), oo, Tcg, Tcg, [1, 0.5, 0], [0, 0, 1])
        assert abs(sv[1] / r[19] - 1) < 1e-6
        del gm
        gc.collect()
    # a batch outlives the Python-side handle of its mesh
    tm = M.ico_ellipsoid(0.7, 0.3, 0.2, 2)
    gm = H.Mesh(tm)
    b = H.BodyBatch([H.BodyDesc(mesh=gm, volume=tm.volume(), surface=tm.area())])
    gm.close()
    st = np.concatenate([Tcg, [0.3, 0, 0], [0, 0, 0.2]])[None]
    g = b.compute_host(go, st)
    ob = O.make_body(O.Mesh.from_tri(tm), volume=tm.volume(), surface=tm.area())
    r = O.body_forces(64, ob, oo, Tcg, [0.3, 0, 0], [0, 0, 0.2])
    assert abs(g[0, 19] / r[19] - 1) < 1e-6


def test_unstaged_chunks(gpu):
    """Faces in random order: vertex windows exceed the LDS stage, so the
    chunks take the global-gather path; results must not change."""
    log("case: scrambled face order (unstaged chunks)")
    rng = np.random.default_rng(9)
    tm = M.ico_ellipsoid(1.0, 0.4, 0.3, 5)
    perm = rng.permutation(tm.nf)
    tms = M.TriMesh(tm.verts, tm.faces[perm])
    oo, go = O.Ocean(), H.Ocean()
    g, r, x = np.zeros((6, 24)), np.zeros((6, 24)), np.zeros((6, 22))
    for s in range(6):
        gm = H.Mesh(tms if s % 2 else tm)
        om = O.Mesh.from_tri(tms if s % 2 else tm)
        Tcg = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], [rng.uniform(-5, 5), 2.0, rng.uniform(-0.2, 0.2)])
        v, w_ = rng.uniform(-2, 2, 3), rng.uniform(-1, 1, 3)
        outs = [np.zeros(3) for _ in range(6)]
        sv = np.zeros(2)
        H.ComputeHydrodynamicForcesSurface(H.HydrodynamicsSettings(), gm, go, Tcg, Tcg, v, w_, *outs, sv)
        g[s, :18] = np.concatenate(outs)
        g[s, 18:20] = sv
        r[s], x[s] = O.surface(64, 1, 1, om, oo, Tcg, Tcg, v, w_, full=True)
    assert_bound(g, r, x, "unstaged chunks")


def test_statics_from_many_threads_with_mesh_switching(gpu):
    """The reference calls its statics from pool threads, one body per task,
    with any mesh.  Calls from 4 host threads, each alternating between three
    meshes of different sizes, equal the same calls made serially bit for bit
    (each call takes its own workspace; a mesh switch rebuilds nothing)."""
    import threading
    log("case: per-body statics from 4 threads, 3 meshes")
    ws = [WL.ellipsoid_surface(), WL.capsule_submerged()]
    meshes = [H.Mesh(ws[0].bodies[0].mesh), H.Mesh(ws[1].bodies[0].mesh), H.Mesh(M.box(0.5, 0.3, 0.2, 3))]
    go = H.Ocean()
    rng = np.random.default_rng(4)
    settings = H.HydrodynamicsSettings(True, True)
    jobs = []
    for t in range(4):
        for k in range(9):
            q = T.random_unit_quaternions(rng, 1)[0]
            Tm = T.quat_to_gl(q, np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-0.2, 0.2)]))
            jobs.append((t, k % 3, np.ascontiguousarray(Tm, dtype=np.float64).reshape(-1), rng.uniform(-2, 2, 3),
                         rng.uniform(-1, 1, 3)))

    def run(job):
        _, mi, Tm, v, w = job
        outs = [np.zeros(3) for _ in range(6)]
        sv = np.zeros(2)
        H.ComputeHydrodynamicForcesSurface(settings, meshes[mi], go, Tm, Tm, v, w, *outs, sv)
        return np.concatenate(outs + [sv])

    serial = [run(j) for j in jobs]
    par = [None] * len(jobs)

    def worker(t):
        for i, j in enumerate(jobs):
            if j[0] == t:
                par[i] = run(j)
    th = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for a, b in zip(serial, par):
        np.testing.assert_array_equal(a, b)
    assert max(np.abs(a).max() for a in serial) > 0

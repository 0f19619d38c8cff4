"""Benchmark: hydrodynamic body-steps/s of the batched exact kernel on one MI355X.

Workload (synthetic, seeded): 4,096 bodies, each an ellipsoid or a box with
half-axes U[0.1, 2]^3 m and a face count log-uniform in [500, 50,000] (about
44 M faces in total), CoB depth U[-1, 3] m, on a 1024 x 1024 wave grid (1 m
spacing) built from 64 Pierson-Moskowitz sinusoids (U = 12 m/s); per step
random velocities, angular velocities and orientations.  One step evaluates
every body: classification, per-face loop, FP64 reduction, buoyancy and
CorrectHydrodynamicForces.  Inputs are resident in HBM when timing starts.

Prints one JSON line.  `value` is body-steps/s of the whole job.  `roofline`
prices the per-face kernel (the dominant one) against HBM bandwidth with the
algorithmic bytes 12 F + 12 V per body (indices and vertices) plus the wave
grid, over its average duration measured with HIP events on its stream.
`cpu_baseline` is the FP32 CPU restatement of the reference path (oracle,
one task per body on a thread pool) timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "hydro body-steps/s, exact vs surrogate at ≤1% F/T error; % of HBM peak"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_fleet(n_bodies: int, seed: int):
    from stonefish_amd import workloads as WL
    t0 = time.time()
    w = WL.fleet(n_bodies, seed=seed)
    log(f"generated {n_bodies} meshes in {time.time() - t0:.1f}s")
    return w


def cpu_baseline(w, states: np.ndarray, ocean_grid: np.ndarray, budget_s: float = 12.0) -> dict:
    """FP32 restatement of the reference path on the host cores (oracle)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    oo = O.Ocean(wave_data=ocean_grid, wave_n=w.sea.wave_n, channels=1, lengths=[w.sea.wave_len])
    meshes = [O.Mesh.from_tri(b.mesh) for b in w.bodies]
    bodies = [O.make_body(m, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface) for m, b in zip(meshes, w.bodies)]
    # calibrate on a small slice, then size the sample to the budget
    n_cal = min(256, len(bodies))
    _, secs = O.batch_forces(32, bodies[:n_cal], oo, states[:n_cal], steps=1, nthreads=threads)
    per_body = max(secs / n_cal, 1e-7)
    n = len(bodies)
    steps = max(1, int(budget_s / (per_body * n)))
    _, secs = O.batch_forces(32, bodies, oo, states, steps=steps, nthreads=threads)
    rate = n * steps / secs
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": rate, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": f"all {n} bodies x {steps} step(s) of the same workload, FP32 restatement of "
                      f"SolidEntity::ComputeHydrodynamicForces (oracle ref32), one task per body on "
                      f"{threads} threads ({cpu_model})",
            "seconds": secs}


def parity_sample(w, states, grid, out_gpu: np.ndarray, n: int = 96) -> dict:
    """Error of the GPU outputs against the FP64 oracle on the first n bodies."""
    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle as O
    from util import bound_ratio
    oo = O.Ocean(wave_data=grid, wave_n=w.sea.wave_n, channels=1, lengths=[w.sea.wave_len])
    r, x = [], []
    for i in range(n):
        b = w.bodies[i]
        om = O.Mesh.from_tri(b.mesh)
        ob = O.make_body(om, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface)
        a, e = O.body_forces(64, ob, oo, states[i, :16], states[i, 16:19], states[i, 19:22], full=True)
        r.append(a)
        x.append(e)
    r, x = np.stack(r), np.stack(x)
    Cd = np.stack([b.Cd for b in w.bodies[:n]])
    Cf = np.stack([b.Cf for b in w.bodies[:n]])
    worst, plain = bound_ratio(out_gpu[:n], r, x, Cd=Cd, Cf=Cf)
    pos_ok = bool(np.array_equal(out_gpu[:n, 21], r[:, 21]) and np.array_equal(out_gpu[:n, 20], r[:, 20]))
    return {"bodies": n, "positions_and_masks_equal": pos_ok,
            "bound_ratio_max": {k: round(float(v), 4) for k, v in worst.items()},
            "norm_rel_err_max": {k: float(f"{v:.3e}") for k, v in plain.items()}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--bodies", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--faces-per-lane", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0, help="resident wavefronts (0 = library default)")
    ap.add_argument("--graph", type=int, default=1, help="replay a captured HIP graph per step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-parity", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        return 2

    import torch
    from stonefish_amd import hydro as H
    from stonefish_amd import waves as W
    from stonefish_amd import workloads as WL

    assert torch.cuda.is_available(), "bench needs a GPU"
    H.init(0)
    dev = torch.device("cuda:0")
    w = build_fleet(args.bodies, args.seed)
    nf = np.array([b.mesh.nf for b in w.bodies], dtype=np.int64)
    nv = np.array([b.mesh.nv for b in w.bodies], dtype=np.int64)
    t0 = time.time()
    gms = [H.Mesh(b.mesh) for b in w.bodies]
    descs = [H.BodyDesc(mesh=g, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface) for g, b in zip(gms, w.bodies)]
    batch = H.BodyBatch(descs, faces_per_lane=args.faces_per_lane, waves=args.waves)
    log(f"uploaded {nf.sum() / 1e6:.1f} M faces / {nv.sum() / 1e6:.1f} M vertices in {time.time() - t0:.1f}s")

    # wave field built on the GPU, fed to the depth query like the reference's fftData
    n = w.sea.wave_n
    grid = torch.empty((n, n), dtype=torch.float32, device=dev)
    W.build_grid_device(w.sea.wave_components, n, grid.data_ptr())
    go = H.Ocean()
    go.set_waves(None, n, 1, [w.sea.wave_len], device_ptr=grid.data_ptr())
    torch.cuda.synchronize()

    nsets = 8
    states_h = WL.random_states(w, nsets)
    states = [torch.tensor(states_h[k], dtype=torch.float64, device=dev).contiguous() for k in range(nsets)]
    out = torch.zeros((args.bodies, 24), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream

    graphs = None
    if args.graph:
        # one captured graph per state set (kernel arguments are baked into a graph)
        gstream = torch.cuda.Stream()
        graphs = []
        for k in range(nsets):
            b = H.BodyBatch(descs, faces_per_lane=args.faces_per_lane, waves=args.waves)
            with torch.cuda.stream(gstream):
                b.capture(go, states[k].data_ptr(), out.data_ptr(), gstream.cuda_stream)
            graphs.append(b)
        torch.cuda.synchronize()

    def step(k: int) -> None:
        if graphs is not None:
            graphs[k % nsets].replay(stream)
        else:
            batch.compute(go, states[k % nsets].data_ptr(), out.data_ptr(), stream)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = args.bodies * args.steps / elapsed

    # per-kernel timing with HIP events on the launch stream (eager launches)
    batch.timing(True)
    for k in range(max(10, min(args.steps, 50))):
        batch.compute(go, states[k % nsets].data_ptr(), out.data_ptr(), stream)
        if k % 4 == 3:
            batch.kernel_ms()   # harvest
    torch.cuda.synchronize()
    face_ms, total_ms, count = batch.kernel_ms()
    batch.timing(False)

    face_bytes = int((12 * nf + 12 * nv).sum()) + n * n * 16   # indices + vertices + wave quad texture
    achieved = face_bytes / (face_ms * 1e-3) / 1e9
    nb, nchunks, nfaces = batch.info()
    res = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 per face, f64 reduction",
        "data": "synthetic",
        "config": {"workload": f"fleet of {args.bodies} bodies (ellipsoids/boxes, half-axes U[0.1,2] m, "
                               f"faces log-U[500,50000]), depth U[-1,3] m, 1024x1024 PM wave grid",
                   "bodies": args.bodies, "faces": int(nfaces), "vertices": int(nv.sum()), "chunks": nchunks,
                   "graph": bool(args.graph), "faces_per_lane": args.faces_per_lane or 8},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_faces", "kernel_ms": round(face_ms, 5), "step_kernels_ms": round(total_ms, 5),
                     "algorithmic_bytes": face_bytes, "timed_launches": count},
    }
    if not args.no_parity:
        t0 = time.time()
        out_h = out.cpu().numpy()
        # the last eager launch used state set (count-1) % nsets
        kk = (max(10, min(args.steps, 50)) - 1) % nsets
        res["parity"] = parity_sample(w, states_h[kk], grid.cpu().numpy(), out_h)
        log(f"parity sample in {time.time() - t0:.1f}s")
    if not args.no_cpu:
        t0 = time.time()
        res["cpu_baseline"] = cpu_baseline(w, states_h[0], grid.cpu().numpy())
        log(f"cpu baseline in {time.time() - t0:.1f}s")
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

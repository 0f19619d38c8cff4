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


BOX_CPU_SHARE = 16   # host threads one GPU's job may use on the measuring machine


def host_cpu_info() -> dict:
    """CPU model, logical CPUs this process may run on, the physical cores
    among them (unique (package, core) pairs of /proc/cpuinfo, the count the
    reference's pool is sized to: SimulationApp.cpp:47, GetPhysicalCores),
    and the cgroup CPU quota."""
    info = {"model": "", "logical_cpus": 0, "physical_cores": 0, "host_physical_cores": 0, "cgroup_cpus": None}
    try:
        aff = sorted(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = list(range(os.cpu_count() or 1))
    info["logical_cpus"] = len(aff)
    cores, host_cores, cpu = set(), set(), {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "processor":
                cpu = {"id": int(v)}
            elif k == "model name" and not info["model"]:
                info["model"] = v
            elif k in ("physical id", "core id"):
                cpu[k] = v
                if "physical id" in cpu and "core id" in cpu:
                    key = (cpu["physical id"], cpu["core id"])
                    host_cores.add(key)
                    if cpu["id"] in aff:
                        cores.add(key)
    except OSError:
        pass
    info["physical_cores"] = len(cores) or len(aff)
    info["host_physical_cores"] = len(host_cores) or info["physical_cores"]
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            info["cgroup_cpus"] = round(int(q) / int(per), 2)
    except (OSError, ValueError):
        pass
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                m = line.split(":", 1)[1].strip()
                if len(m) > len(info["model"]):
                    info["model"] = m
    except Exception:
        pass
    return info


def cpu_baseline(w, states: np.ndarray, ocean_grid: np.ndarray, budget_s: float = 12.0, budget1_s: float = 6.0) -> dict:
    """FP32 restatement of the reference path (oracle ref32) on the host cores:
    one task per body on a pool of N threads, N the physical cores this job
    may use (the reference's pool size, SimulationApp.cpp:47), capped at the
    measuring machine's CPU share of one GPU's job; and on one core."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    info = host_cpu_info()
    threads = max(1, min(info["physical_cores"], BOX_CPU_SHARE,
                         int(info["cgroup_cpus"]) if info["cgroup_cpus"] else 1 << 30))
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
    # one core: the first n1 bodies (in order, so the face-count mix is the
    # generator's) for about budget1_s
    n1 = int(min(n, max(64, budget1_s / (per_body * threads))))
    _, secs1 = O.batch_forces(32, bodies[:n1], oo, states[:n1], steps=1, nthreads=1)
    rate1 = n1 / secs1
    nf1 = sum(b.mesh.nf for b in w.bodies[:n1])
    nf = sum(b.mesh.nf for b in w.bodies)
    return {"value": rate, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": f"all {n} bodies x {steps} step(s) of the same workload, FP32 restatement of "
                      f"SolidEntity::ComputeHydrodynamicForces (oracle ref32), one task per body on a pool of "
                      f"{threads} threads",
            "seconds": round(secs, 3),
            "one_core": {"value": round(rate1, 1), "unit": "body-steps/s", "cores": 1,
                         "sample": f"first {n1} bodies x 1 step ({nf1} faces; the whole fleet has {nf})",
                         "seconds": round(secs1, 3),
                         "faces_per_s": round(nf1 / secs1, 1)},
            "faces_per_s": round(nf * steps / secs, 1),
            "host": info,
            "pool_rule": f"physical cores available to this job ({info['physical_cores']}) capped at the "
                         f"{BOX_CPU_SHARE}-thread CPU share one GPU's job has on the measuring machine "
                         f"(host: {info['host_physical_cores']} physical cores, {info['logical_cpus']} logical "
                         f"CPUs visible)"}


def parity_sample(w, states, grid, out_gpu: np.ndarray, n: int = 96) -> dict:
    """Error of the GPU outputs against the FP64 oracle on the first n bodies."""
    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle as O
    from util import bound_ratio, component_errors
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
            "norm_rel_err_max": {k: float(f"{v:.3e}") for k, v in plain.items()},
            "component_err_max": component_errors(out_gpu[:n], r)}


def measured_traffic():
    """HBM bytes per step of the face passes from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (profiles/), or None."""
    p = ROOT / "profiles" / "k_faces_fleet4096.txt"
    try:
        for line in p.read_text().splitlines():
            if line.startswith("traffic_bytes_per_step"):
                return int(line.split()[1])
    except (OSError, ValueError, IndexError):
        pass
    return None


def _section(name: str, fn):
    """Runs an optional section of the report; a Python-level failure in it is
    recorded in its place (and logged) instead of costing the headline line."""
    try:
        return fn()
    except Exception as e:   # noqa: BLE001 - reported in the JSON line, not swallowed
        log(f"{name} section failed: {e!r}")
        return {"error": repr(e)}


def _time_replays(fn, k: int) -> float:
    """Seconds per call of fn over k calls (after 3 untimed), bracketed by syncs."""
    import torch
    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / k


def measure_peaks(dev) -> dict:
    """Peaks of this GPU measured in the same run: HBM bandwidth of a
    STREAM-style device copy (2 GiB read, 2 GiB written per pass, best of 5)
    and the device properties the FP32 peak is priced from."""
    import torch
    n = 512 * 1024 * 1024   # floats: 2 GiB
    a = torch.ones(n, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    del a, b
    torch.cuda.empty_cache()
    props = torch.cuda.get_device_properties(dev)
    return {"hbm_copy_GBs": round(2 * 4 * n / best / 1e9, 1), "hbm_copy_bytes": 2 * 4 * n,
            "hbm_peak_spec_GBs": HBM_PEAK_GBS,
            "compute_units": int(props.multi_processor_count), "name": props.name,
            "gcn_arch": getattr(props, "gcnArchName", ""),
            "memory_GiB": round(props.total_memory / 2 ** 30, 1)}


def single_body_section(args) -> dict:
    """The two single-body workloads: a 2,048-triangle capsule fully submerged
    at 5 m (1,000 steps) and a 20,480-triangle ellipsoid crossing a flat
    surface (10,000 steps), random velocities and orientations per step.

    Each is timed three ways: through the per-body statics of the drop-in
    boundary (one synchronous call per step: ComputeHydrodynamicForcesSubmerged
    for the capsule, ComputeHydrodynamicForcesSurface for the ellipsoid, host
    buffers in and out, as SolidEntity calls them); as one batch of all the
    steps' states of the one body (SolidEntity::ComputeHydrodynamicForces,
    classification, buoyancy and CorrectHydrodynamicForces included, states
    resident in HBM, one captured graph per replay); and on the host with the
    FP32 restatement (oracle ref32) on the N-thread pool and on one core.
    The error of the batch against the FP64 oracle is stated per vector and
    per component on a sample of steps."""
    import torch
    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle as O
    from stonefish_amd import hydro as H
    from stonefish_amd import workloads as WL
    from util import bound_ratio, component_errors
    dev = torch.device("cuda:0")
    info = host_cpu_info()
    threads = max(1, min(info["physical_cores"], BOX_CPU_SHARE,
                         int(info["cgroup_cpus"]) if info["cgroup_cpus"] else 1 << 30))
    res = {}
    for key, w, nsteps in (("capsule_submerged_2048", WL.capsule_submerged(), 1000),
                           ("ellipsoid_surface_20480", WL.ellipsoid_surface(), 10000)):
        b = w.bodies[0]
        gm = H.Mesh(b.mesh)
        go = H.Ocean()
        st = WL.random_states(w, nsteps)[:, 0, :]            # (steps, 22)
        submerged = key.startswith("capsule")
        # (1) per-body statics, one synchronous call per step
        outs = [np.zeros(3) for _ in range(6)]
        sv = np.zeros(2)
        settings = H.HydrodynamicsSettings(True, True)

        def call(i, outs=outs, sv=sv):
            if submerged:
                H.ComputeHydrodynamicForcesSubmerged(gm, go, st[i, :16], st[i, :16], st[i, 16:19], st[i, 19:22],
                                                     *outs[2:])
            else:
                H.ComputeHydrodynamicForcesSurface(settings, gm, go, st[i, :16], st[i, :16], st[i, 16:19],
                                                   st[i, 19:22], *outs, sv)
        for i in range(20):
            call(i)
        t0 = time.perf_counter()
        for i in range(nsteps):
            call(i)
        statics_rate = nsteps / (time.perf_counter() - t0)
        # the same calls from a pool of host threads, one task per body-step,
        # as the reference's pool threads call the statics (each thread takes
        # a workspace with its own stream from the library's per-device pool)
        from concurrent.futures import ThreadPoolExecutor
        nthr = BOX_CPU_SHARE

        def worker(tid):
            o, s2 = [np.zeros(3) for _ in range(6)], np.zeros(2)
            for i in range(tid, nsteps, nthr):
                call(i, o, s2)

        with ThreadPoolExecutor(nthr) as ex:
            list(ex.map(worker, range(nthr)))   # every thread's first call sets up its workspace
            t0 = time.perf_counter()
            list(ex.map(worker, range(nthr)))
            statics_pool_rate = nsteps / (time.perf_counter() - t0)
        # (2) all steps as one batch of the body's states (body level)
        desc = H.BodyDesc(mesh=gm, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface)
        batch = H.BodyBatch([desc] * nsteps)
        d_st = torch.tensor(st, dtype=torch.float64, device=dev).contiguous()
        d_out = torch.zeros((nsteps, 24), dtype=torch.float64, device=dev)
        stream = torch.cuda.current_stream().cuda_stream
        if not args.no_autotune:
            batch.autotune(go, d_st.data_ptr(), d_out.data_ptr(), stream)
            torch.cuda.synchronize()
        gstream = torch.cuda.Stream()
        with torch.cuda.stream(gstream):
            batch.capture(go, d_st.data_ptr(), d_out.data_ptr(), gstream.cuda_stream)
        torch.cuda.synchronize()
        k = 20
        # graph replay and eager launches, the faster
        sec = min(_time_replays(lambda: batch.replay(stream), k),
                  _time_replays(lambda: batch.compute(go, d_st.data_ptr(), d_out.data_ptr(), stream), k))
        batch_rate = nsteps / sec
        batch.timing(True)
        for _ in range(10):
            batch.compute(go, d_st.data_ptr(), d_out.data_ptr(), stream)
            batch.kernel_ms()
        torch.cuda.synchronize()
        face_ms, total_ms, _ = batch.kernel_ms()
        batch.timing(False)
        out = d_out.cpu().numpy()
        nf, nv = b.mesh.nf, b.mesh.nv
        alg = (12 * nf + 12 * nv) * nsteps
        # (3) host: FP32 restatement on N threads and on one core, bounded sample
        om = O.Mesh.from_tri(b.mesh)
        ob = O.make_body(om, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface)
        oo = O.Ocean()
        ns_cal = min(nsteps, 64)
        _, c = O.batch_forces(32, [ob] * ns_cal, oo, st[:ns_cal], steps=1, nthreads=1)
        per = max(c / ns_cal, 1e-7)
        n1 = int(min(nsteps, max(32, 4.0 / per)))
        _, c1 = O.batch_forces(32, [ob] * n1, oo, st[:n1], steps=1, nthreads=1)
        nN = int(min(nsteps, max(64, 4.0 * threads / per)))
        reps = max(1, int(4.0 * threads / (per * nN)))
        _, cN = O.batch_forces(32, [ob] * nN, oo, st[:nN], steps=reps, nthreads=threads)
        # error of the batch against ref64 on a sample of steps
        ns = 64
        r, x = [], []
        for i in range(ns):
            a, e = O.body_forces(64, ob, oo, st[i, :16], st[i, 16:19], st[i, 19:22], full=True)
            r.append(a)
            x.append(e)
        r, x = np.stack(r), np.stack(x)
        Cd = np.repeat(np.asarray(b.Cd)[None], ns, 0)
        Cf = np.repeat(np.asarray(b.Cf)[None], ns, 0)
        worst, plain = bound_ratio(out[:ns], r, x, Cd=Cd, Cf=Cf)
        pos = np.bincount(out[:, 21].astype(int) + 1, minlength=4)[1:].tolist()
        res[key] = {
            "faces": nf, "vertices": nv, "steps": nsteps,
            "positions_outside_inside_crossing": pos,
            "gpu_per_body_statics": {"value": round(statics_rate, 1), "unit": "body-steps/s",
                                     "call": "ComputeHydrodynamicForcesSubmerged" if submerged
                                     else "ComputeHydrodynamicForcesSurface",
                                     "note": "one synchronous call per step from one host thread, host buffers",
                                     "pool_threads": nthr,
                                     "pool_value": round(statics_pool_rate, 1)},
            "gpu_batch": {"value": round(batch_rate, 1), "unit": "body-steps/s", "ms_per_step": round(sec * 1e3, 4),
                          "opts": batch.get_opts(),
                          "face_pass_ms": round(face_ms, 5), "step_kernels_ms": round(total_ms, 5),
                          "roofline": {"bound": "hbm", "achieved": round(alg / (face_ms * 1e-3) / 1e9, 1),
                                       "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": round(alg / (face_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                       "algorithmic_bytes": alg,
                                       "note": "12 F + 12 V per body-step; the steps share one mesh, so the "
                                               "face passes read it from cache, not HBM"}},
            "cpu_n_cores": {"value": round(nN * reps / cN, 1), "unit": "body-steps/s", "cores": threads,
                            "sample": f"{nN} steps x {reps}"},
            "cpu_one_core": {"value": round(n1 / c1, 1), "unit": "body-steps/s", "cores": 1,
                             "sample": f"{n1} steps"},
            "parity": {"steps": ns, "positions_and_masks_equal": bool(np.array_equal(out[:ns, 20:22], r[:, 20:22])),
                       "bound_ratio_max": {k2: round(float(v), 4) for k2, v in worst.items()},
                       "norm_rel_err_max": {k2: float(f"{v:.3e}") for k2, v in plain.items()},
                       "component_err_max": component_errors(out[:ns], r)},
        }
        log(f"{key}: statics {statics_rate:.0f}/s, batch {batch_rate:.0f}/s, cpu N {nN * reps / cN:.0f}/s, "
            f"cpu 1 {n1 / c1:.0f}/s")
        del batch
    return res


def surrogate_section(args) -> dict:
    """Training workload (8 hulls of the fleet distribution, samples from the
    exact kernel), held-out error, and exact vs surrogate body-steps/s on the
    same hull set; MFMA rate of the fused inference and of training."""
    import torch
    from stonefish_amd import surrogate as S
    t0 = time.time()
    cfg = S.TrainConfig(steps=args.sur_train_steps, batch=args.sur_batch, lr=args.sur_lr, warmup=200,
                        t_weight=args.sur_t_weight)
    gate_tau = args.sur_gate_tau if args.sur_gate_tau > 0 else None
    r = S.run_training_workload(8, args.sur_per_hull, args.sur_steps, args.sur_holdout, cfg, seed=17, log=log,
                                gate_tau=gate_tau)
    hulls, nets, norms, w, ocean = r["hulls"], r["nets"], r["norms"], r["workload"], r["ocean"]
    ds = r["dataset"]
    log(f"surrogate training workload in {time.time() - t0:.1f}s")
    keys = ("F_median", "F_p90", "F_p99", "T_median", "T_p90", "T_p99", "F_within_1pct", "T_within_1pct")
    n_tot = sum(h["n"] for h in hulls)
    agg = {k: round(float(sum(h[k] * h["n"] for h in hulls) / max(1, n_tot)), 5) for k in keys}
    # exact vs surrogate on 8 x 512 bodies of the same hulls
    per = 512
    batch, meshes = S.make_hull_batch(w.bodies, per)
    S.attach_tables(batch, ocean, [(h * per, per) for h in range(8)], tables=r["tables"])
    nb = 8 * per
    rng = np.random.default_rng(99)
    states = torch.from_numpy(S.sample_states(rng, nb, (0.0, w.sea.wave_len), (-1.0, 3.0))).cuda()
    out = torch.zeros((nb, 24), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    # the exact step in the configuration the autotuner picks for it, eager
    # launches and graph replay (the faster)
    tune = {}
    if not args.no_autotune:
        batch.autotune(ocean, states.data_ptr(), out.data_ptr(), stream)
        tune["exact"] = batch.get_opts()
    gs = torch.cuda.Stream()
    with torch.cuda.stream(gs):
        batch.capture(ocean, states.data_ptr(), out.data_ptr(), gs.cuda_stream)
    torch.cuda.synchronize()
    t_exact = min(_time_replays(lambda: batch.replay(stream), 30),
                  _time_replays(lambda: batch.compute(ocean, states.data_ptr(), out.data_ptr(), stream), 30))
    step = S.SurrogateStep(batch, ocean, nets, norms, [per] * 8, states)
    # the surrogate alone is one kernel per step: a plain launch costs less
    # than a one-node graph replay, so both are timed
    t_eager = _time_replays(step.run, 200)
    step.capture()
    t_graph = _time_replays(step.replay, 200)
    t_sur_only = min(t_eager, t_graph)
    # the step that delivers outputs: surrogate, exact kernel on the bodies it
    # flags out of distribution, merge (graph replay and eager, the faster);
    # the masked exact step is tuned on this step's flags
    if not args.no_autotune:
        step.run()
        torch.cuda.synchronize()
        mask = step.ood.clone()
        batch.autotune(ocean, states.data_ptr(), out.data_ptr(), stream, d_active=mask.data_ptr())
        tune["fallback"] = batch.get_opts()
        torch.cuda.synchronize()
    t_fb_eager = _time_replays(step.run_with_fallback, 100)
    step.capture(with_fallback=True)
    t_fb_graph = _time_replays(step.replay, 100)
    t_sur = min(t_fb_eager, t_fb_graph)
    n_fallback = int(step.ood.sum())
    # the surrogate's error on this sample against the exact outputs
    err = torch.zeros((nb, 2), dtype=torch.float32, device="cuda")
    step.run()
    step.plan.merge(out, step.ft.clone(), err=err)
    torch.cuda.synchronize()
    sample_err = S.error_summary(err.cpu().numpy(), step.ood.cpu().numpy().astype(bool), step.pos.cpu().numpy())
    rnd = lambda d: {k: (round(v, 5) if isinstance(v, float) else v) for k, v in d.items()}   # noqa: E731
    keys_d = ("F_median", "F_p90", "F_p99", "T_median", "T_p90", "T_p99", "F_within_1pct", "T_within_1pct")

    def gated_point(tau: float, full: bool) -> dict:
        """The step with the error-aware gate at threshold tau (fitted per hull
        on the training errors): the barely wet crossing bodies whose training
        error exceeded tau go to the exact kernel too; held-out errors of the
        network and delivered, the flagged share, and the step's time with its
        larger fallback (the masked exact step tuned on its flags)."""
        gnorms = [S.with_gate(norms[h], S.fit_gate(*r["gate_data"][h], float(norms[h].in_lo[9]), tau))
                  for h in range(8)]
        ev = [S.evaluate(ds, h, nets[h], gnorms[h], delivered=True) for h in range(8)]
        n_d = sum(e["delivered"]["n"] for e in ev)
        n_n = sum(e["n"] for e in ev)
        gstep = S.SurrogateStep(batch, ocean, nets, gnorms, [per] * 8, states)
        if not args.no_autotune:
            gstep.run()
            torch.cuda.synchronize()
            gmask = gstep.ood.clone()
            batch.autotune(ocean, states.data_ptr(), out.data_ptr(), stream, d_active=gmask.data_ptr())
            torch.cuda.synchronize()
        tg_eager = _time_replays(gstep.run_with_fallback, 100)
        gstep.capture(with_fallback=True)
        tg_graph = _time_replays(gstep.replay, 100)
        tg = min(tg_eager, tg_graph)
        pt = {
            "tau": tau,
            "held_out_flagged_fraction": round(sum(e["n_ood"] for e in ev) / max(1, n_d), 4),
            "held_out_delivered_error": {k: round(float(sum(e["delivered"][k] * e["delivered"]["n"] for e in ev)
                                                        / max(1, n_d)), 5) for k in keys_d},
            "speedup": round(t_exact / tg, 2),
        }
        if full:
            gerr = torch.zeros((nb, 2), dtype=torch.float32, device="cuda")
            gstep.run()
            gstep.plan.merge(out, gstep.ft.clone(), err=gerr)
            torch.cuda.synchronize()
            ge, go, gp = gerr.cpu().numpy(), gstep.ood.cpu().numpy().astype(bool), gstep.pos.cpu().numpy()
            pt.update({
                "depth_below_m": [round(float(g.in_lo[9]) + float(g.gate_band), 3) for g in gnorms],
                "held_out_network_error": {k: round(float(sum(e[k] * e["n"] for e in ev) / max(1, n_n)), 5)
                                           for k in keys_d},
                "fallback_bodies": int(go.sum()),
                "autotune": batch.get_opts(),
                "us_per_step": {"with_fallback_launch": round(tg_eager * 1e6, 2),
                                "with_fallback_graph": round(tg_graph * 1e6, 2)},
                "throughput_sample_network_error": rnd(S.error_summary(ge, go, gp)),
                "throughput_sample_delivered_error": rnd(S.delivered_summary(ge, go, gp)),
            })
        return pt

    gated, sweep, at_1pct = None, [], None
    if gate_tau is not None:
        gated = gated_point(gate_tau, True)
        # lower thresholds trade speed for accuracy: the largest speed-up whose
        # delivered held-out F and T medians are both within 1 %
        for tau in [float(t) for t in args.sur_gate_sweep.split(",") if t.strip()]:
            sweep.append(gated_point(tau, False))
        ok = [p for p in [gated] + sweep if p["held_out_delivered_error"]["F_median"] <= 0.01
              and p["held_out_delivered_error"]["T_median"] <= 0.01]
        if ok:
            best = max(ok, key=lambda p: p["speedup"])
            at_1pct = {"tau": best["tau"], "speedup": best["speedup"],
                       "held_out_flagged_fraction": best["held_out_flagged_fraction"]}
    # MFMA rate of the fused inference at a large batch (all hulls' held-out inputs, one launch)
    xs = [ds.inputs(h, norms[h], range(ds.train_steps, ds.steps))[0] for h in range(8)]
    x = torch.cat(xs).contiguous()
    y = torch.empty((x.shape[0], 6), dtype=torch.float32, device="cuda")
    counts = [xx.shape[0] for xx in xs]
    nets[0].timing(True)
    for _ in range(20):
        S.forward_group(nets, counts, x, y)
    ms, flops, cnt = nets[0].kernel_ms()
    nets[0].timing(False)
    inf_tflops = flops / (ms * 1e-3) / 1e12
    train_tflops = float(r["train_tflops"])
    res = {
        "workload": f"8 hulls x {args.sur_per_hull} bodies x {args.sur_steps} steps of exact-kernel samples "
                    f"({8 * args.sur_per_hull * args.sur_steps} total, {args.sur_holdout} steps held out), "
                    f"{args.sur_train_steps} Adam steps of {args.sur_batch} per hull",
        "held_out_error": agg,
        "per_hull": [{k: h[k] for k in ("hull", "faces", "train_samples", "F_median", "F_p90", "T_median",
                                        "T_p90", "n", "n_ood", "n_out_of_water", "final_loss")}
                     for h in hulls],
        "train_seconds": round(float(r["train_seconds"]), 1),
        "throughput_bodies": nb,
        "exact_body_steps_per_s": round(nb / t_exact, 1),
        "surrogate_body_steps_per_s": round(nb / t_sur, 1),
        "surrogate_us_per_step": {"with_fallback_launch": round(t_fb_eager * 1e6, 2),
                                  "with_fallback_graph": round(t_fb_graph * 1e6, 2),
                                  "surrogate_only_launch": round(t_eager * 1e6, 2),
                                  "surrogate_only_graph": round(t_graph * 1e6, 2)},
        "fallback_bodies": n_fallback,
        "autotune": tune,
        "speedup": round(t_exact / t_sur, 2),
        "speedup_surrogate_only": round(t_exact / t_sur_only, 2),
        "throughput_sample_error": {k: (round(v, 5) if isinstance(v, float) else v) for k, v in sample_err.items()},
        "gated": gated,
        "gate_sweep": sweep,
        "speedup_at_1pct_median_F_and_T": at_1pct,
        "mfma": {"inference_tflops": round(inf_tflops, 1), "inference_rows": int(x.shape[0]),
                 "inference_kernel_ms": round(ms, 4), "training_tflops": round(train_tflops, 1),
                 "peak_tflops_bf16_dense": 2500.0, "inference_frac": round(inf_tflops / 2500.0, 4),
                 "training_frac": round(train_tflops / 2500.0, 4)},
        "seconds": round(time.time() - t0, 1),
    }
    return res


def rollout_section(args, sea) -> dict:
    """Mixed rollout: AUV compounds (8,192-triangle hull + four 1,024-triangle
    fairings) at 1 kHz on the wave field, surrogate for every body, exact
    fallback outside the training range, exact audit every 100 steps."""
    import torch
    from stonefish_amd import hydro as H
    from stonefish_amd import surrogate as S
    from stonefish_amd import workloads as WL
    t0 = time.time()
    n = args.rollout_bodies
    auv = WL.auv_compound()
    meshes = {}
    parts = []
    for p in auv.parts:
        if id(p.mesh) not in meshes:
            meshes[id(p.mesh)] = H.Mesh(p.mesh)
        parts.append(H.PartDesc(mesh=meshes[id(p.mesh)], origin=p.origin, Cd=p.Cd, Cf=p.Cf))
    body = H.BodyDesc(mesh=None, aabb_min=auv.aabb_min, aabb_max=auv.aabb_max, volume=auv.volume,
                      surface=auv.surface, P_CB=auv.P_CB)
    batch = H.CompoundBatch([H.CompoundDesc(body, parts)] * n, faces_per_lane=args.rollout_faces_per_lane or 8)
    _, _, nfaces = batch.info()
    ocean, grid = S.make_ocean(sea)
    # trained on twice the rollout's range of velocities (its compounds start
    # with |v_i| <= 1 m/s, |w_i| <= 0.5 rad/s): with a quarter of margin
    # instead, the audits' F median fell from 0.98 % to 0.60 % and T from
    # 9.3 % to 7.0 %, but the trajectories left the training range six times
    # as often (3.3 % of the compounds at the audits against 0.6 %) and the
    # exact fallback took the rollout from 9.6x to 5.7x
    ds = S.HullDataset(batch, ocean, 1, n, steps=args.rollout_data_steps, holdout_steps=2, seed=31,
                       xy_range=(0.0, sea.wave_len), depth_range=(-1.0, 3.0), v_max=2.0, w_max=1.0)
    tables = S.attach_tables(batch, ocean, [(0, n)])
    norm = ds.fit_norm(0)
    x, t, _ = ds.inputs(0, norm, range(ds.train_steps))
    net = S.Net(65536, seed=7)
    hist = S.train(net, x, t, S.TrainConfig(steps=args.sur_train_steps, batch=args.sur_batch, lr=args.sur_lr,
                                            warmup=200))
    held = S.evaluate(ds, 0, net, norm)
    log(f"rollout surrogate trained ({ds.gen_seconds:.1f}s data, {hist['seconds']:.1f}s train): "
        f"F median {held['F_median']:.4f}")
    rng = np.random.default_rng(77)
    states = torch.from_numpy(S.sample_states(rng, n, (0.0, sea.wave_len), (0.0, 2.5), v_max=1.0,
                                              w_max=0.5)).cuda()
    states0 = states.clone()   # the rollout integrates in place; the gated rollout starts from the same states
    # exact kernel alone on the same bodies (graph replay of one evaluation),
    # in the configuration the autotuner picks for the whole batch; the
    # rollout's batch evaluates only the few out-of-distribution bodies per
    # step and is tuned on that masked step: the rollout's states after a few
    # steps with 1% of the compounds active (a seeded draw; the audits find
    # 0.5-1% out of distribution, and the first steps' states are all inside)
    exact = H.CompoundBatch([H.CompoundDesc(body, parts)] * n)
    out = torch.zeros((n, 24), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    tune = {}
    if not args.no_autotune:
        pts = exact.autotune(ocean, states.data_ptr(), out.data_ptr(), stream)
        tune["exact"] = {"chosen": exact.get_opts(), "points": pts}
    gs = torch.cuda.Stream()
    with torch.cuda.stream(gs):
        exact.capture(ocean, states.data_ptr(), out.data_ptr(), gs.cuda_stream)
    torch.cuda.synchronize()
    t_exact = _time_replays(lambda: exact.replay(stream), 10)
    ro = S.MixedRollout(batch, ocean, net, norm, states, auv.mass, auv.inertia, dt=1e-3, audit_every=100,
                        audit_batch=exact)
    for _ in range(3):
        ro.step()
    torch.cuda.synchronize()
    if not args.no_autotune:
        mask = torch.from_numpy((np.random.default_rng(5).random(n) < 0.01).astype(np.uint8)).cuda()
        st_now = ro.states.clone()
        pts = batch.autotune(ocean, st_now.data_ptr(), out.data_ptr(), stream, d_active=mask.data_ptr())
        torch.cuda.synchronize()
        tune["rollout_masked"] = {"chosen": batch.get_opts(), "points": pts, "active_bodies": int(mask.sum())}
    steps = args.rollout_steps

    def timed_steps(r, start) -> float:
        """steps of rollout r from the states `start` (copied in place: a
        captured graph keeps the states' address), audits every 100."""
        r.states.copy_(start)
        r.audits.clear()
        r.k = 0
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            r.step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t1) / steps

    # eager launches, then the ordinary step replayed from a HIP graph, both
    # from the same states (the faster is the rollout's rate; the audits are
    # those of the second run, the same trajectory)
    start = ro.states.clone()
    t_ro_eager = timed_steps(ro, start)
    ro.capture()
    t_ro_graph = timed_steps(ro, start)
    t_ro = min(t_ro_eager, t_ro_graph)
    summ = ro.audit_summary()
    ood_frac = float(torch.cat([a[2] for a in ro.audits]).float().mean()) if ro.audits else None
    rnd = lambda d: {k: (round(v, 5) if isinstance(v, float) else v) for k, v in d.items()}   # noqa: E731
    gated = None
    if args.sur_gate_tau > 0:
        # the same rollout with the error-aware gate fitted on the training
        # steps: the flagged compounds (out of distribution, or barely wet
        # where the training error exceeded tau) take the exact kernel
        d, e = S.gate_errors(ds, 0, net, norm, range(min(4, ds.train_steps)))
        band = S.fit_gate(d, e, float(norm.in_lo[9]), args.sur_gate_tau)
        gnorm = S.Norm(**{k: getattr(norm, k) for k in ("in_shift", "in_scale", "in_lo", "in_hi", "out_shift",
                                                        "out_scale")}, gate_band=band)
        held_g = S.evaluate(ds, 0, net, gnorm, delivered=True)
        rg = S.MixedRollout(batch, ocean, net, gnorm, states0, auv.mass, auv.inertia, dt=1e-3, audit_every=100,
                            audit_batch=exact)
        for _ in range(3):
            rg.step()
        torch.cuda.synchronize()
        if not args.no_autotune:
            gmask = rg.ood.clone()
            st_now = rg.states.clone()
            pts = batch.autotune(ocean, st_now.data_ptr(), out.data_ptr(), stream, d_active=gmask.data_ptr())
            torch.cuda.synchronize()
            tune["rollout_masked_gated"] = {"chosen": batch.get_opts(), "points": pts,
                                            "active_bodies": int(gmask.sum())}
        gstart = rg.states.clone()
        t_rg_eager = timed_steps(rg, gstart)
        rg.capture()
        t_rg_graph = timed_steps(rg, gstart)
        t_rg = min(t_rg_eager, t_rg_graph)
        sg = rg.audit_summary()
        gated = {
            "tau": args.sur_gate_tau,
            "depth_below_m": round(float(norm.in_lo[9]) + band, 3),
            "ms_per_step": round(1e3 * t_rg, 4),
            "ms_per_step_eager_graph": [round(1e3 * t_rg_eager, 4), round(1e3 * t_rg_graph, 4)],
            "speedup_vs_exact": round(t_exact / t_rg, 2),
            "flagged_fraction_at_audits": (float(torch.cat([a[2] for a in rg.audits]).float().mean())
                                           if rg.audits else None),
            "audit_network_error": rnd({k: v for k, v in sg.items() if k != "delivered"}),
            "audit_delivered_error": rnd(sg.get("delivered", {})),
            "held_out_network_error": rnd({k: v for k, v in held_g.items() if k != "delivered"}),
            "held_out_delivered_error": rnd(held_g["delivered"]),
        }
    return {
        "workload": f"{n} AUV compounds (8,192 + 4 x 1,024 triangles, {nfaces} faces), 1 kHz, "
                    f"1024 x 1024 PM wave grid, audit every 100 steps, {steps} steps",
        "body_steps_per_s": round(n / t_ro, 1),
        "ms_per_step": round(1e3 * t_ro, 4),
        "ms_per_step_eager_graph": [round(1e3 * t_ro_eager, 4), round(1e3 * t_ro_graph, 4)],
        "exact_body_steps_per_s": round(n / t_exact, 1),
        "speedup_vs_exact": round(t_exact / t_ro, 2),
        "faces_per_lane": {"rollout": batch.get_opts()[0], "exact": exact.get_opts()[0]},
        "waves": {"rollout": batch.get_opts()[1], "exact": exact.get_opts()[1]},
        "autotune": tune,
        "audit_error": {k: (round(v, 5) if isinstance(v, float) else v) for k, v in summ.items() if k != "delivered"},
        "audit_delivered_error": {k: (round(v, 5) if isinstance(v, float) else v)
                                  for k, v in summ.get("delivered", {}).items()},
        "ood_fraction_at_audits": ood_frac,
        "held_out_error": {k: (round(v, 5) if isinstance(v, float) else v) for k, v in held.items()},
        "gated": gated,
        "seconds": round(time.time() - t0, 1),
    }


def compound_exact_section(args, sea) -> dict:
    """The rollout's exact step of the AUV compounds alone, for counter runs:
    `args.rollout_bodies` compounds on the fleet's wave field in the
    autotuner's configuration, `args.steps` eager steps after the warm-up,
    then 10 steps timed with HIP events on the launch stream."""
    import torch
    from stonefish_amd import hydro as H
    from stonefish_amd import surrogate as S
    from stonefish_amd import workloads as WL
    n = args.rollout_bodies
    auv = WL.auv_compound()
    meshes, parts = {}, []
    for p in auv.parts:
        if id(p.mesh) not in meshes:
            meshes[id(p.mesh)] = H.Mesh(p.mesh)
        parts.append(H.PartDesc(mesh=meshes[id(p.mesh)], origin=p.origin, Cd=p.Cd, Cf=p.Cf))
    body = H.BodyDesc(mesh=None, aabb_min=auv.aabb_min, aabb_max=auv.aabb_max, volume=auv.volume,
                      surface=auv.surface, P_CB=auv.P_CB)
    batch = H.CompoundBatch([H.CompoundDesc(body, parts)] * n)
    _, nchunks, nfaces = batch.info()
    ocean, grid = S.make_ocean(sea)
    rng = np.random.default_rng(77)
    states = torch.from_numpy(S.sample_states(rng, n, (0.0, sea.wave_len), (0.0, 2.5), v_max=1.0,
                                              w_max=0.5)).cuda()
    out = torch.zeros((n, 24), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    if not args.no_autotune:
        batch.autotune(ocean, states.data_ptr(), out.data_ptr(), stream)
    t = _time_replays(lambda: batch.compute(ocean, states.data_ptr(), out.data_ptr(), stream), max(1, args.steps))
    batch.timing(True)
    for k in range(10):
        batch.compute(ocean, states.data_ptr(), out.data_ptr(), stream)
        batch.kernel_ms()
    torch.cuda.synchronize()
    face_ms, total_ms, _ = batch.kernel_ms()
    batch.timing(False)
    return {"workload": f"{n} AUV compounds ({nfaces} faces, {nchunks} chunks), exact step",
            "ms_per_step": round(1e3 * t, 4), "face_ms": round(face_ms, 5), "step_kernels_ms": round(total_ms, 5),
            "faces": int(nfaces), "opts": batch.get_opts()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--bodies", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--faces-per-lane", type=int, default=0)
    ap.add_argument("--rollout-faces-per-lane", type=int, default=0,
                    help="faces per lane of the rollout's batch before tuning (0: 8)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the given (or default) faces per lane and waves instead of tuning them")
    ap.add_argument("--waves", type=int, default=0, help="resident wavefronts (0 = library default)")
    ap.add_argument("--split-passes", type=int, default=0,
                    help="1: submerged and crossing chunks in face launches of their own")
    ap.add_argument("--graph", type=int, default=1,
                    help="also time replays of captured HIP graphs of the step (reported beside the eager rate)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-single", action="store_true", help="skip the single-body workloads")
    ap.add_argument("--single-only", action="store_true", help="run only the single-body workloads")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate and rollout sections")
    ap.add_argument("--rollout-only", action="store_true", help="run only the mixed rollout section (profiling)")
    ap.add_argument("--compound-exact-only", action="store_true",
                    help="run only the exact step of the rollout's compounds (profiling)")
    ap.add_argument("--surrogate-only", action="store_true", help="run only the surrogate section (profiling)")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a plain 256 MiB device copy (reference dispatch for HBM byte counters)")
    ap.add_argument("--sur-per-hull", type=int, default=8192)
    ap.add_argument("--sur-steps", type=int, default=31)
    ap.add_argument("--sur-holdout", type=int, default=3)
    ap.add_argument("--sur-train-steps", type=int, default=60000)
    ap.add_argument("--sur-lr", type=float, default=5e-3, help="peak Adam learning rate (cosine schedule)")
    ap.add_argument("--sur-batch", type=int, default=16384)
    ap.add_argument("--sur-t-weight", type=float, default=1.0, help="torque outputs' weight in the training loss")
    ap.add_argument("--sur-gate-sweep", default="0.05,0.03,0.02,0.01",
                    help="further gate thresholds measured (held-out delivered error and speed-up)")
    ap.add_argument("--sur-gate-tau", type=float, default=0.1,
                    help="error-aware gate: crossing bodies in the depth band whose training error exceeded this "
                         "fall back to the exact kernel (0: off)")
    ap.add_argument("--rollout-bodies", type=int, default=16384)
    ap.add_argument("--rollout-steps", type=int, default=300)
    ap.add_argument("--rollout-data-steps", type=int, default=14)
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
    if args.calibrate:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.rollout_only:
        print(json.dumps({"rollout": rollout_section(args, WL.fleet(1, seed=args.seed).sea)}), flush=True)
        return 0
    if args.compound_exact_only:
        print(json.dumps({"compound_exact": compound_exact_section(args, WL.fleet(1, seed=args.seed).sea)}),
              flush=True)
        return 0
    if args.single_only:
        print(json.dumps({"single_body": single_body_section(args)}), flush=True)
        return 0
    if args.surrogate_only:
        print(json.dumps({"surrogate": surrogate_section(args)}), flush=True)
        return 0
    w = build_fleet(args.bodies, args.seed)
    nf = np.array([b.mesh.nf for b in w.bodies], dtype=np.int64)
    nv = np.array([b.mesh.nv for b in w.bodies], dtype=np.int64)
    t0 = time.time()
    gms = [H.Mesh(b.mesh) for b in w.bodies]
    descs = [H.BodyDesc(mesh=g, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface) for g, b in zip(gms, w.bodies)]
    batch = H.BodyBatch(descs, faces_per_lane=args.faces_per_lane, waves=args.waves, split_passes=args.split_passes)
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
    sweep = None
    if not args.no_autotune and not args.faces_per_lane and not args.waves and not args.split_passes:
        t0 = time.time()
        sweep = batch.autotune(go, states[0].data_ptr(), out.data_ptr(), stream)
        torch.cuda.synchronize()
        args.faces_per_lane, args.waves, args.split_passes = batch.get_opts()
        log(f"autotune: {len(sweep)} points in {time.time() - t0:.1f}s, chose faces_per_lane "
            f"{args.faces_per_lane}, {args.waves} waves, split_passes {args.split_passes}")

    def timed(step) -> float:
        for k in range(args.warmup):
            step(k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            step(k)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # the timed steps: eager launches of the step (prologue, face launch,
    # epilogue) on one stream, the state set changing every step
    elapsed = timed(lambda k: batch.compute(go, states[k % nsets].data_ptr(), out.data_ptr(), stream))
    ms_per_step = 1e3 * elapsed / args.steps
    value = args.bodies * args.steps / elapsed

    graph_rate = None
    if args.graph:
        # the same steps as replays of captured HIP graphs, one per state set
        # (kernel arguments are baked into a graph), reported beside the rate
        gstream = torch.cuda.Stream()
        graphs = []
        for k in range(nsets):
            b = H.BodyBatch(descs, faces_per_lane=args.faces_per_lane, waves=args.waves,
                            split_passes=args.split_passes)
            with torch.cuda.stream(gstream):
                b.capture(go, states[k].data_ptr(), out.data_ptr(), gstream.cuda_stream)
            graphs.append(b)
        torch.cuda.synchronize()
        g_el = timed(lambda k: graphs[k % nsets].replay(stream))
        graph_rate = {"value": round(args.bodies * args.steps / g_el, 1), "ms_per_step": round(1e3 * g_el / args.steps, 4)}
        del graphs

    # per-kernel timing with HIP events on the launch stream (eager launches)
    batch.timing(True)
    for k in range(max(10, min(args.steps, 50))):
        batch.compute(go, states[k % nsets].data_ptr(), out.data_ptr(), stream)
        if k % 4 == 3:
            batch.kernel_ms()   # harvest
    torch.cuda.synchronize()
    face_ms, total_ms, count = batch.kernel_ms()
    batch.timing(False)

    # the host-buffer boundary (sfh_batch_compute_host: states in and outputs
    # out over PCIe, synchronous), the rate a caller with host arrays sees
    st_h = states_h[0]
    out_h0 = np.zeros((args.bodies, 24))
    batch.compute_host(go, st_h, out_h0)
    t0 = time.perf_counter()
    nh = 5
    for _ in range(nh):
        batch.compute_host(go, st_h, out_h0)
    host_io = args.bodies * nh / (time.perf_counter() - t0)

    face_bytes = int((12 * nf + 12 * nv).sum()) + n * n * 16   # indices + vertices + wave quad texture
    traffic = measured_traffic()
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
                   "launch": "eager", "faces_per_lane": batch.get_opts()[0],
                   "waves": batch.get_opts()[1], "split_passes": batch.get_opts()[2],
                   "autotune": None if sweep is None else {
                       "chosen": {"faces_per_lane": batch.get_opts()[0], "waves": batch.get_opts()[1],
                                  "split_passes": batch.get_opts()[2]},
                       "points": sweep}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_faces", "kernel_ms": round(face_ms, 5), "step_kernels_ms": round(total_ms, 5),
                     "algorithmic_bytes": face_bytes, "timed_launches": count},
        "host_buffers_body_steps_per_s": round(host_io, 1),
        "graph_replay": graph_rate,
    }
    res["roofline"]["peaks_measured"] = measure_peaks(dev)
    if not args.no_parity:
        t0 = time.time()
        out_h = out.cpu().numpy()
        # the last eager launch used state set (count-1) % nsets
        kk = (max(10, min(args.steps, 50)) - 1) % nsets
        res["parity"] = _section("parity", lambda: parity_sample(w, states_h[kk], grid.cpu().numpy(), out_h))
        log(f"parity sample in {time.time() - t0:.1f}s")
    if not args.no_cpu:
        t0 = time.time()
        res["cpu_baseline"] = _section("cpu_baseline", lambda: cpu_baseline(w, states_h[0], grid.cpu().numpy()))
        log(f"cpu baseline in {time.time() - t0:.1f}s")
    if not args.no_single:
        t0 = time.time()
        graphs = None
        res["single_body"] = _section("single_body", lambda: single_body_section(args))
        log(f"single-body workloads in {time.time() - t0:.1f}s")
    if not args.no_surrogate:
        graphs = None
        res["surrogate"] = _section("surrogate", lambda: surrogate_section(args))
        if "surrogate_body_steps_per_s" in res["surrogate"]:
            res["surrogate"]["speedup_vs_fleet_exact"] = round(res["surrogate"]["surrogate_body_steps_per_s"] / value,
                                                               2)
        res["rollout"] = _section("rollout", lambda: rollout_section(args, w.sea))
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

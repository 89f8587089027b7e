"""Benchmark: one fused fwd+bwd+optimiser iteration of the differentiable renderer on
config 2 (icosphere subdivision 5: 10242 vertices / 20480 triangles, Lambert with one
directional light and constant albedo 0.8, 1024x1024, 4 spp, 2^20 edge samples, FP32).

A step = re-commit from the device-resident parameters (LBVH build or refit), render + L2
loss + adjoint, d_render (interior texture adjoint + edge sampling), Adam update of the
vertex positions and the albedo -- all on the handle's stream, inputs resident in HBM, replayed
from a hipGraph.  Beside it: the same step with the albedo frozen (the interior pass then has
no output and is left out), the rays the step casts (counted by the kernels), the plain
one-job-at-a-time loop through the host interface, the served forms (continuous batching) with
their p50/p99 per-iteration latency, the multi-view and serving-stream workloads, and the CPU
baselines (the oracle's CPU port at config 2, the reference's own code at config 1).  Prints
one JSON line.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu] [--no-serve]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The served legs run up to 16 jobs, each on its own stream: give the HIP runtime as many
# hardware queues (read once, at its initialisation; the default is 4) so their steps run
# side by side.  The single-stream legs are unaffected.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(kernel, F, V, E, W, H, spp_sq, N, P, tiles):
    """Compulsory HBM bytes of one launch: every stream read or written once."""
    scene = 36 * F + 32 * (2 * F - 1) + 24 * V          # SoA triangles + LBVH nodes + vertex attributes
    px = W * H
    if kernel == "k_render":
        return scene + 12 * px + 12 * px + 12 * px + 8 * tiles   # image write, target read, adjoint write
    if kernel == "k_interior":
        return scene + 12 * px + 4 * P                            # adjoint read, gradient
    if kernel == "k_edge_trace":   # rng, endpoints, adjoint, colour differences written; edges + cdf
        return scene + N * (8 + 32 + 12 + 16) + 12 * E
    if kernel == "k_edge_grad":    # colour differences, rng, endpoints, adjoint, positions; pmf; gradient
        return N * (16 + 8 + 32 + 12 + 24) + 4 * E + 4 * P
    return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (profiles/traffic.json, copy-calibrated), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f)["bytes_per_launch"].get(kernel)
    except (OSError, KeyError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def fused_rate(drh, r, cams, views, d_target, N, d_loss, steps, warmup):
    """Warm up, then time `steps` fused steps (graph replays) between two synchronisations."""
    for _ in range(warmup):
        r.opt_step_dev(cams, views, d_target.ptr, N, d_loss.ptr)
    r.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.opt_step_dev(cams, views, d_target.ptr, N, d_loss.ptr)
    r.sync()
    return time.perf_counter() - t0


JITTER_SEED = 2024   # pixel-sample jitter pattern of the config-2 legs


def plain_leg(drh, meshes, cam, target, N, spp, lam, iters):
    """The plain one-job-at-a-time GPU path: the reference's optimiser loop
    (src/optimizer_simple.cpp:272-309) driving the drop-in through its host interface --
    per iteration commit(scene), d_render_and_compare with host target and host gradient
    (src/drender.h:169-191), then Adam on the host and the scene update.  Returns iters/s."""
    r = drh.DiffRender(mode=drh.LAMBERT, spp=spp, light=lam)
    m = meshes[0].copy()
    V = m.V
    r.commit([m])
    r.set_pixel_jitter(True, JITTER_SEED)
    P = r.grad_size()
    mom, vel = np.zeros(P, np.float32), np.zeros(P, np.float32)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8

    def it(t):
        r.commit([m])
        _, g, _ = r.d_render_and_compare([cam], [target], N, images=False)
        mom[:] = b1 * mom + (1 - b1) * g
        vel[:] = b2 * vel + (1 - b2) * g * g
        step = lr * (mom / (1 - b1 ** t)) / (np.sqrt(vel / (1 - b2 ** t)) + eps)
        m.pos = (m.pos.reshape(-1) - step[:3 * V]).reshape(-1, 3).astype(np.float32)

    it(1)
    t0 = time.perf_counter()
    for t in range(2, iters + 2):
        it(t)
    dt = time.perf_counter() - t0
    r.close()
    return iters / dt


def plain_fit(drh, meshes, cams, targets, mode, spp, N, lam, iters):
    """The plain one-job-at-a-time path for one job: create + commit, then per iteration
    commit, d_render_and_compare with host targets / host gradient, host Adam (lr 1e-3) on
    positions (and colours in vertex-colour mode) -- the reference's optimiser loop over the
    drop-in.  Returns the parameters after `iters` iterations."""
    r = drh.DiffRender(mode=mode, spp=spp, light=lam)
    m = meshes[0].copy()
    V = m.V
    r.commit([m])
    P = r.grad_size()
    mom, vel = np.zeros(P, np.float32), np.zeros(P, np.float32)
    b1, b2, eps, lr = 0.9, 0.999, 1e-8, 1e-3
    nparam = 6 * V if mode == drh.VERTEX_COLOR else 3 * V
    for t in range(1, iters + 1):
        r.commit([m])
        _, g, _ = r.d_render_and_compare(cams, targets, N, images=False)
        mom[:] = b1 * mom + (1 - b1) * g
        vel[:] = b2 * vel + (1 - b2) * g * g
        step = lr * (mom / (1 - b1 ** t)) / (np.sqrt(vel / (1 - b2 ** t)) + eps)
        m.pos = (m.pos.reshape(-1) - step[:3 * V]).reshape(-1, 3).astype(np.float32)
        if nparam > 3 * V:
            m.col = (m.col.reshape(-1) - step[3 * V:6 * V]).reshape(-1, 3).astype(np.float32)
    r.close()
    return m


def sequential_fused(drh, jobs, targets, lam, iters):
    """Each job's fused device step (graph replays of four steps, as the server launches them; no
    host round trips) run job after job on one handle at a time: handle, commit, Adam set-up,
    `iters` steps, synchronise, tear down.  Returns the wall seconds of the whole batch."""
    t0 = time.perf_counter()
    for j, tg in zip(jobs, targets):
        r = drh.DiffRender(mode=j["mode"], spp=j["spp"], light=lam)
        r.commit(j["meshes"])
        r.opt_init(drh.OPT_ADAM, 1e-3, 1e-3)
        r.set_bvh_rebuild(iters + 1 if iters <= 64 else 8)
        dt = r.device_array(np.concatenate([x.reshape(-1) for x in tg]))
        dl = r.device_array(np.zeros(1, np.float32))
        cs = drh.cameras_array(j["cams"])
        done = 0
        while done < iters:   # as the server steps: the first step alone, then graphs of 4 steps
            k = 1 if done == 0 else min(4, iters - done)
            r.opt_steps_dev(cs, len(j["cams"]), dt.ptr, j["edge_samples"], dl.ptr, k)
            done += k
        r.sync()
        dt.free()
        dl.free()
        r.close()
    return time.perf_counter() - t0


def mix_legs(drh, S, lam, njobs, iters, max_active):
    """Config 3's uneven batch (64 jobs by default): the plain path fits the jobs one after
    another through the host interface; the single-job fused device step runs them one after
    another; the server takes them all at once (continuous batching, `max_active` slots).  Wall
    time of each includes per-job setup.  Returns a dict for the bench line."""
    jobs = S.config3_jobs(njobs)
    targets = []
    for j in jobs:
        rt = drh.DiffRender(mode=j["mode"], spp=j["spp"], light=lam)
        rt.commit(j["target"])
        targets.append(rt.render(j["cams"]))
        rt.close()
    t0 = time.perf_counter()
    for j, tg in zip(jobs, targets):
        plain_fit(drh, j["meshes"], j["cams"], tg, j["mode"], j["spp"], j["edge_samples"], lam, iters)
    plain_s = time.perf_counter() - t0
    seq_s = sequential_fused(drh, jobs, targets, lam, iters)
    srv = drh.Server(hbm_budget=64 << 30, max_active=max_active)
    t0 = time.perf_counter()
    ids = [srv.submit(j["meshes"], j["cams"], tg, iters, j["edge_samples"], mode=j["mode"], spp=j["spp"], light=lam)
           for j, tg in zip(jobs, targets)]
    res = [srv.wait(i) for i in ids]
    served_s = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    assert all(x["status"] == 0 for x in res), [x["status"] for x in res]
    total = njobs * iters
    tris = [int(j["meshes"][0].F) for j in jobs]
    return {"workload": f"config3: {njobs} uneven jobs (F log-uniform 12..80k as cube-spheres, {min(tris)}..{max(tris)} "
                        f"triangles drawn; 128-512 px, VCOL/Lambert, N = 4 px), {iters} Adam iterations each",
            "plain_iters_per_s": round(total / plain_s, 2),
            "sequential_fused_iters_per_s": round(total / seq_s, 2),
            "served_iters_per_s": round(total / served_s, 2),
            "served_over_plain": round(plain_s / served_s, 3),
            "served_over_sequential_fused": round(seq_s / served_s, 3),
            "p50_ms": round(st["p50_ms"], 4), "p99_ms": round(st["p99_ms"], 4), "peak_active": st["peak_active"],
            "mean_active": round(st["mean_active"], 3), "served_wall_s": round(served_s, 4),
            "server_thread_s": {"admit": round(st["admit_s"], 4), "step": round(st["step_s"], 4),
                                "finish": round(st["finish_s"], 4)},
            "admitters": int(os.environ.get("DRH_SERVER_ADMITTERS", "1")), "triangles": tris}


def poisson_leg(drh, S, lam, njobs, rate, max_active, iter_scale):
    """Config 5's serving stream: `njobs` mesh-fitting jobs submitted at Poisson arrival times
    (rate jobs/s, seed 7) to one server under a 288 GB HBM budget; reports per-iteration
    p50 / p99 latency, jobs/s and iters/s over the whole stream.  `iter_scale` scales the
    50-500 iteration counts (the same scale for every job)."""
    jobs = S.config5_jobs(njobs)
    targets = []
    for j in jobs:
        rt = drh.DiffRender(mode=j["mode"], spp=j["spp"], light=lam)
        rt.commit(j["target"])
        targets.append(rt.render(j["cams"]))
        rt.close()
    gaps = np.random.default_rng(7).exponential(1.0 / rate, size=njobs)
    srv = drh.Server(hbm_budget=288 << 30, max_active=max_active)
    t0 = time.perf_counter()
    ids, iters_total, due = [], 0, 0.0
    for j, tg, g in zip(jobs, targets, gaps):
        due += g
        while time.perf_counter() - t0 < due:
            time.sleep(0.0005)
        it = max(1, int(round(j["iterations"] * iter_scale)))
        iters_total += it
        ids.append(srv.submit(j["meshes"], j["cams"], tg, it, j["edge_samples"], mode=j["mode"], spp=j["spp"], light=lam,
                              flags=drh.JOB_FULL_JACOBIAN))
    res = [srv.wait(i) for i in ids]
    dt = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    assert all(x["status"] == 0 for x in res), [x["status"] for x in res]
    tris = [int(j["meshes"][0].F) for j in jobs]
    return {"workload": f"config5: {njobs} Lambert jobs at Poisson rate {rate}/s (F log-uniform 12..200k as cube-spheres, "
                        f"{min(tris)}..{max(tris)} drawn; 128-1024 px, 1-8 ring views, 50-500 iterations x {iter_scale}), "
                        f"full screen Jacobian, 288 GB budget, {max_active} slots",
            "jobs_per_s": round(njobs / dt, 3), "iters_per_s": round(iters_total / dt, 2),
            "p50_ms": round(st["p50_ms"], 4), "p99_ms": round(st["p99_ms"], 4), "peak_active": st["peak_active"],
            "mean_active": round(st["mean_active"], 3), "wall_s": round(dt, 3)}


def multiview_leg(drh, S, lam, steps, N):
    """Config 4's multi-view adjoint: a displaced sphere of 199,808 triangles (radial noise
    0.02, seed 6), per-vertex colours, 16 cameras on a ring (radius 3, 20 deg), 512x512 each,
    spp 1, Lambert with the vertex colour as albedo; a fused step fits positions and colours
    (full screen Jacobian).  Times the step with the reference's Hammersley interior samples
    (recomputed), and with render-aligned samples recomputed, stored, and chosen by budget."""
    base = S.uv_sphere(224, 448)
    m = S.with_random_colors(S.radial_displace(base, 0.02, seed=6, smooth=False), 6)
    t = S.with_random_colors(S.radial_displace(base, 0.02, seed=16, smooth=False), 16)
    cams = S.cameras_ring(16, 512, 512)
    mode = drh.LAMBERT_VERTEX_COLOR
    rt = drh.DiffRender(mode=mode, spp=1, light=lam)
    rt.commit([t])
    targets = rt.render(cams)
    rt.close()
    hit_bytes = 16 * 16 * 512 * 512
    out = {"workload": f"config4: {m.F} triangles / {m.V} vertices, 16 views x 512x512, spp 1, Lambert with vertex-colour "
                       f"albedo, {N} edge samples per view, Adam on positions + colours",
           "hit_buffer_bytes": hit_bytes}
    for name, aligned, store, limit in (("hammersley_recompute", 0, 0, None), ("aligned_recompute", 1, 0, None),
                                        ("aligned_store", 1, 1, None), ("aligned_auto_small_budget", 1, 2, hit_bytes // 2)):
        r = drh.DiffRender(mode=mode, spp=1, light=lam)
        r.commit([m])
        r.set_edge_grad_mode(drh.EDGE_GRAD_FULL_JACOBIAN)
        r.set_interior_samples(aligned, store)
        if limit is not None:
            r.set_hit_buffer_limit(limit)
        r.opt_init(drh.OPT_ADAM, 1e-3, 1e-3)
        dt = r.device_array(np.concatenate([x.reshape(-1) for x in targets]))
        dl = r.device_array(np.zeros(1, np.float32))
        cs = drh.cameras_array(cams)
        sec = fused_rate(drh, r, cs, 16, dt, N, dl, steps, 2)
        out[name + "_ms_per_step"] = round(1e3 * sec / steps, 3)
        if name == "aligned_auto_small_budget":
            out[name + "_hit_bytes_held"] = r.hit_buffer_bytes()
        r.close()
    return out


def served_leg(drh, S, meshes, cam, N, spp, lam, jobs, iters):
    """The served form: `jobs` config-2 mesh-fitting jobs (distinct targets) submitted to the
    continuous-batching server at once; each keeps its scene, LBVH, Adam moments and target
    on the device and steps by graph replay, the jobs' steps overlapping on the GPU.  Wall
    time from the first submit to the last result (admission and downloads included)."""
    targets = []
    for k in range(jobs):
        t = S.with_albedo(S.radial_displace(S.icosphere(5, 1.0), 0.1, seed=100 + k), 0.8)
        rt = drh.DiffRender(mode=drh.LAMBERT, spp=spp, light=lam)
        rt.commit([t])
        targets.append(rt.render([cam]))
        rt.close()
    srv = drh.Server(hbm_budget=64 << 30, max_active=jobs)
    t0 = time.perf_counter()
    ids = [srv.submit(meshes, [cam], targets[k], iters, N, mode=drh.LAMBERT, spp=spp, light=lam)
           for k in range(jobs)]
    res = [srv.wait(i) for i in ids]
    dt = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    assert all(x["status"] == 0 for x in res), [x["status"] for x in res]
    return {"iters_per_s": round(jobs * iters / dt, 3), "jobs": jobs, "iterations_per_job": iters,
            "p50_ms": round(st["p50_ms"], 4), "p99_ms": round(st["p99_ms"], 4),
            "peak_active": st["peak_active"], "mean_active": round(st["mean_active"], 3),
            "wall_s": round(dt, 4)}


def config1_leg(drh, S, steps, with_reference):
    """Config 1 (cube, vertex colours, 256x256, spp 1, 2^16 edge samples) with the reference's
    optimiser (OptSimple, GD_Adam): the GPU's fused step (eager: its learning-rate decay is host
    state) and, where oracle/_ref holds the reference compiled from its own sources, the
    reference's own OptSimple::Run with its brute-force tracer on the host cores."""
    c = S.config1()
    cam = c["cams"][0]
    rt = drh.DiffRender(mode=drh.VERTEX_COLOR, spp=1)
    rt.commit(c["target"])
    tg = rt.render([cam])
    rt.close()
    r = drh.DiffRender(mode=drh.VERTEX_COLOR, spp=1)
    r.commit(c["meshes"])
    r.opt_init(drh.OPT_SIMPLE, 0.1, 0.1)
    dt = r.device_array(tg[0].reshape(-1))
    dl = r.device_array(np.zeros(1, np.float32))
    cs = drh.cameras_array([cam])
    sec = fused_rate(drh, r, cs, 1, dt, c["edge_samples"], dl, steps, 5)
    r.close()
    out = {"workload": "config1: cube 12 triangles, vertex colours, 256x256, spp 1, 2^16 edge samples, OptSimple GD_Adam",
           "gpu_iters_per_s": round(steps / sec, 2)}
    if with_reference:
        from oracle import ref_run
        if ref_run.available():
            iters = 200
            res = ref_run.run(c["meshes"], [cam], 2, 1, ref_run.OP_OPTIMIZE, edge_samples=c["edge_samples"], images=tg,
                              opt_iters=iters)
            cpu = iters / float(res["opt_seconds"][0])
            out["cpu_reference"] = {"value": round(cpu, 3), "unit": "iters/s", "cores": int(res["omp_threads"][0]),
                                    "kind": "reference", "cpu_model": cpu_model(),
                                    "sample": f"{iters} iterations of the reference's own OptSimple::Run "
                                              "(commit + d_render_and_compare + OptStep) with its brute-force tracer, "
                                              "compiled from its sources (Embree binary absent)"}
            out["gpu_over_cpu_reference"] = round(out["gpu_iters_per_s"] / cpu, 2)
    return out


def cpu_port_leg(meshes, target, cam, lam, N, spp, seconds):
    """The oracle's C++ port with its CPU BVH (the Embree stand-in) on config 2, the same work
    per iteration as the GPU step: commit (scene flattening + BVH build), render, loss, d_render
    (interior + edges), Adam on positions and albedo.  OpenMP over the host cores."""
    import oracle as O
    m = meshes[0].copy()
    V = m.V
    mom = vel = None
    iters = 0
    c0 = time.perf_counter()
    while True:
        ref = O.Scene([m], O.TRACER_BVH)
        ref.set_jitter(True, JITTER_SEED + iters + 1)   # the GPU step's pattern of step t
        img = ref.render(4, spp, [cam], lam=lam)[0]
        _, adj = O.loss(img, target)
        g = ref.d_render(4, spp, [cam], [adj], N, lam=lam)
        if mom is None:
            mom, vel = np.zeros_like(g), np.zeros_like(g)
        t = iters + 1
        mom = 0.9 * mom + 0.1 * g
        vel = 0.999 * vel + 0.001 * g * g
        step = 1e-3 * (mom / (1 - 0.9 ** t)) / (np.sqrt(vel / (1 - 0.999 ** t)) + 1e-8)
        m.pos = (m.pos.reshape(-1) - step[:3 * V]).reshape(-1, 3).astype(np.float32)
        m.tex = (m.tex.reshape(-1) - step[3 * V:3 * V + 3]).reshape(m.tex.shape).astype(np.float32)
        iters += 1
        if time.perf_counter() - c0 > seconds or iters >= 5:
            break
    cdt = time.perf_counter() - c0
    cores = host_cores()
    omp = int(os.environ.get("OMP_NUM_THREADS", cores))
    return {"value": round(iters / cdt, 5), "unit": "iters/s", "cores": min(omp, cores), "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{iters} full config-2 iterations of the oracle (its C++ port of the reference loop, -O3, "
                      f"OpenMP) with its CPU BVH tracer as the Embree stand-in (Embree binary absent): commit with "
                      f"BVH build, render (4 spp jittered), loss, d_render (interior + edges), Adam on positions "
                      f"and albedo"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--serve-jobs", type=int, default=4, help="concurrent config-2 jobs in the served leg (0: skip the plain and served legs)")
    ap.add_argument("--serve-iters", type=int, default=50, help="Adam iterations per served job")
    ap.add_argument("--plain-iters", type=int, default=20, help="iterations of the plain host-driven loop")
    ap.add_argument("--no-serve", action="store_true", help="skip the plain, served, multi-view and stream legs")
    ap.add_argument("--mix-jobs", type=int, default=64, help="jobs in the config-3 uneven mix (0: skip)")
    ap.add_argument("--mix-iters", type=int, default=100, help="Adam iterations per mix job (config 3: 100 steps)")
    ap.add_argument("--mix-active", type=int, default=16, help="server slots for the mix")
    ap.add_argument("--multiview-steps", type=int, default=5, help="timed config-4 steps per interior mode (0: skip)")
    ap.add_argument("--stream-jobs", type=int, default=24, help="jobs in the config-5 Poisson stream (0: skip)")
    ap.add_argument("--stream-rate", type=float, default=20.0, help="Poisson arrival rate, jobs/s")
    ap.add_argument("--stream-iter-scale", type=float, default=0.1, help="scale of the 50-500 iteration counts")
    ap.add_argument("--config1-steps", type=int, default=200, help="timed config-1 steps (0: skip)")
    ap.add_argument("--eager", action="store_true",
                    help="config-2 steps without hipGraph replay or second-stream overlap (for kernel traces)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy 3 times (k_copy_f32), the known-size reference "
                         "for FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_calibration:
        from diffrender_tutorials_amd import drh
        r = drh.DiffRender(mode=drh.VERTEX_COLOR, spp=1)
        n = 64 * 1024 * 1024
        a, b = drh.DeviceBuffer(r, 4 * n), drh.DeviceBuffer(r, 4 * n)
        for _ in range(3):
            r._check(drh.load().drh_debug_copy(r.h, b.ptr, a.ptr, n))
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this path runs one job "
              "per GPU; jobs spread over GPUs without communication)", file=sys.stderr)
        sys.exit(2)

    from diffrender_tutorials_amd import drh, scenes as S

    c = S.config2()
    meshes, cam = c["meshes"], c["cams"][0]
    lam = S.lambert_config2()
    N = c["edge_samples"]
    spp = c["spp"]
    W, H = cam.width, cam.height
    F, V = meshes[0].F, meshes[0].V

    # target image, rendered on the GPU from the displaced mesh
    rt = drh.DiffRender(mode=drh.LAMBERT, spp=spp, light=lam)
    rt.commit(c["target"])
    target = rt.render([cam])[0]
    rt.close()

    def handle(tex_lr):
        r = drh.DiffRender(mode=drh.LAMBERT, spp=spp, light=lam)
        r.commit(meshes)
        r.set_pixel_jitter(True, JITTER_SEED)   # 4 spp jittered: a fresh stratified pattern per step
        if args.eager:
            # every step launched kernel by kernel, one stream: a kernel trace's durations are then
            # those the per-kernel HIP events measure
            drh.load().drh_opt_use_graph(r.h, 0)
            r.set_device_sharing(2)
        r.opt_init(drh.OPT_ADAM, 1e-3, 1e-3)
        r.opt_set_texture_lr(tex_lr)
        return r

    # headline: the full-gradient step (albedo learning rate 1e-3: the interior texture adjoint
    # runs every step, as in the reference's d_render), captured once, replayed as a hipGraph
    r = handle(1e-3)
    P = r.grad_size()
    d_target = r.device_array(target.reshape(-1))
    d_loss = r.device_array(np.zeros(1, np.float32))
    cams = drh.cameras_array([cam])
    for _ in range(args.warmup):
        r.opt_step_dev(cams, 1, d_target.ptr, N, d_loss.ptr)
    r.sync()
    loss0 = float(d_loss.download(1)[0])
    r.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.opt_step_dev(cams, 1, d_target.ptr, N, d_loss.ptr)
    r.sync()
    dt = time.perf_counter() - t0
    loss1 = float(d_loss.download(1)[0])
    ms = 1e3 * dt / args.steps

    # the same steps again, eager with a HIP event pair around every kernel (per-kernel times
    # on the library's stream; the events keep the step out of the graph)
    r.profile(True)
    r.sync()
    g0 = time.perf_counter()
    for _ in range(args.steps):
        r.opt_step_dev(cams, 1, d_target.ptr, N, d_loss.ptr)
    r.sync()
    events_ms = 1e3 * (time.perf_counter() - g0) / args.steps
    kernels = ["k_render", "k_interior", "k_edge_trace", "k_edge_grad", "k_cdf_par", "k_edge_bin",
               "k_suspects", "k_radix_sort", "k_refit", "k_karras",
               "k_tri_setup", "k_morton", "k_pack_nodes", "k_prepare", "k_project", "k_sus_bin", "k_light_bin",
               "k_loss_final", "k_fill_f32", "k_fill_u32", "k_adam"]
    prof = {}
    for k in kernels:
        tot, cnt = r.profile_read(k)
        if cnt:
            prof[k] = {"ms_per_launch": tot / cnt, "launches_per_step": cnt / args.steps}
    r.profile(False)

    # rays the step casts, counted by the kernels' counting copies over one eager step
    r.count_rays(True)
    r.opt_step_dev(cams, 1, d_target.ptr, N, d_loss.ptr)
    r.sync()
    rc = r.ray_counts()
    r.count_rays(False)
    r.close()
    rays = {"render_camera": rc["render_camera"], "render_shadow": rc["render_shadow"],
            "interior": rc["interior_closest"] + rc["interior_any"], "edge_side": rc["edge_side"],
            "edge_shadow": rc["edge_shadow"]}
    rays["total"] = sum(rays.values())

    dom = max(("k_render", "k_interior", "k_edge_trace", "k_edge_grad"), key=lambda k: prof.get(k, {}).get("ms_per_launch", 0))
    tiles = ((W + 7) // 8) * ((H + 7) // 8)
    E = 3 * F // 2
    sq = int(np.sqrt(spp))
    bytes_dom = algorithmic_bytes(dom, F, V, E, W, H, sq * sq, N, P, tiles)
    achieved = bytes_dom / (prof[dom]["ms_per_launch"] * 1e-3) / 1e9
    value = args.steps / dt

    # the same step with the albedo frozen: the interior pass (whose only output in Lambert is
    # the texture gradient) is left out
    rf = handle(0.0)
    d_tf = rf.device_array(target.reshape(-1))
    d_lf = rf.device_array(np.zeros(1, np.float32))
    fdt = fused_rate(drh, rf, cams, 1, d_tf, N, d_lf, args.steps, args.warmup)
    rf.close()

    result = {
        "metric": "fwd+bwd iters/s",
        "value": round(value, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "config2: icosphere-5 (10242 V / 20480 F), Lambert, 1024x1024, 4 spp jittered, 2^20 edge samples,"
                               " fused commit (LBVH build / refit) + render + loss + d_render (interior texture adjoint"
                               " + edge sampling) + Adam on the vertex positions and the albedo per step",
                   "triangles": F, "vertices": V, "resolution": [W, H], "spp": spp, "edge_samples": N,
                   "pixel_samples": f"stratified jitter, seed {JITTER_SEED} + step",
                   "parallelism": "single GPU"},
        "rays_per_step": rays,
        "rays_per_s": round(rays["total"] * value, 1),
        "edge_samples_per_s": round(N * value, 1),
        "edge_samples_kept_per_step": rays["edge_side"] // 2,
        "loss_first_last": [loss0, loss1],
        "frozen_albedo_step": {"iters_per_s": round(args.steps / fdt, 3), "ms_per_step": round(1e3 * fdt / args.steps, 4),
                               "note": "interior pass left out (its only output is the frozen albedo's gradient)"},
        "ms_per_step_with_kernel_events": round(events_ms, 4),
        "kernels": {k: round(v["ms_per_launch"], 5) for k, v in prof.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(dom),
                     "algorithmic_bytes": bytes_dom},
        "cpu_baseline": None,
    }

    # The serving legs run under a watchdog: should one not finish within its time limit, the
    # line is printed with what was measured (that leg marked as timed out) and the process
    # ends, so a stuck leg cannot cost the headline.
    def guarded(name, limit_s, fn):
        def expire():
            result[name] = {"error": f"timed out after {limit_s} s"}
            print(json.dumps(result), flush=True)
            os._exit(0)
        timer = threading.Timer(limit_s, expire)
        timer.daemon = True
        timer.start()
        try:
            return fn()
        finally:
            timer.cancel()

    if not args.no_serve:
        if args.serve_jobs > 0:
            plain = guarded("plain_one_job_iters_per_s", 300,
                            lambda: plain_leg(drh, meshes, cam, target, N, spp, lam, args.plain_iters))
            served = guarded("served", 300, lambda: served_leg(drh, S, meshes, cam, N, spp, lam, args.serve_jobs, args.serve_iters))
            result["plain_one_job_iters_per_s"] = round(plain, 3)
            result["served"] = served
            result["served_over_plain"] = round(served["iters_per_s"] / plain, 3)
            result["served_over_single_fused_step"] = round(served["iters_per_s"] / result["frozen_albedo_step"]["iters_per_s"], 3)
            result["p50_ms"], result["p99_ms"] = served["p50_ms"], served["p99_ms"]
        if args.mix_jobs > 0:
            result["mix"] = guarded("mix", 300, lambda: mix_legs(drh, S, lam, args.mix_jobs, args.mix_iters, args.mix_active))
        if args.multiview_steps > 0:
            result["multiview"] = guarded("multiview", 300, lambda: multiview_leg(drh, S, lam, args.multiview_steps, 1 << 18))
        if args.stream_jobs > 0:
            result["stream"] = guarded("stream", 300, lambda: poisson_leg(drh, S, lam, args.stream_jobs, args.stream_rate, 8,
                                                                          args.stream_iter_scale))
    if args.config1_steps > 0:
        result["config1"] = config1_leg(drh, S, args.config1_steps, not args.no_cpu)

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_port_leg(meshes, target, cam, lam, N, spp, args.cpu_seconds)
        result["gpu_over_cpu"] = round(result["value"] / result["cpu_baseline"]["value"], 2)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

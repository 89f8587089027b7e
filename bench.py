"""Benchmark: one fused fwd+bwd+optimiser iteration of the differentiable renderer on
config 2 (icosphere subdivision 5: 10242 vertices / 20480 triangles, Lambert with one
directional light and constant albedo 0.8, 1024x1024, 4 spp, 2^20 edge samples, FP32).

A step = re-commit from the device-resident vertices (LBVH rebuild), render + L2 loss +
adjoint, interior adjoint, edge sampling, Adam update of the vertex positions -- all on the
handle's stream, inputs resident in HBM.  Beside it: the plain one-job-at-a-time loop
through the host interface, and the served form (several config-2 jobs on the
continuous-batching server) with its p50/p99 per-iteration latency.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu] [--no-serve]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
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


def plain_leg(drh, meshes, cam, target, N, spp, lam, iters):
    """The plain one-job-at-a-time GPU path: the reference's optimiser loop
    (src/optimizer_simple.cpp:272-309) driving the drop-in through its host interface --
    per iteration commit(scene), d_render_and_compare with host target and host gradient
    (src/drender.h:169-191), then Adam on the host and the scene update.  Returns iters/s."""
    r = drh.DiffRender(mode=drh.LAMBERT, spp=spp, light=lam)
    m = meshes[0].copy()
    V = m.V
    r.commit([m])
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


def mix_legs(drh, S, lam, njobs, iters, max_active):
    """Config 3's uneven batch (64 jobs by default): the plain path fits the jobs one after
    another; the server takes them all at once (continuous batching, `max_active` slots).
    Wall time of each includes per-job setup.  Returns a dict for the bench line."""
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
    return {"workload": f"config3: {njobs} uneven jobs (F log-uniform 12..80k, 128-512 px, VCOL/Lambert, N = 4 px),"
                        f" {iters} Adam iterations each",
            "plain_iters_per_s": round(total / plain_s, 2), "served_iters_per_s": round(total / served_s, 2),
            "served_over_plain": round(plain_s / served_s, 3), "p50_ms": round(st["p50_ms"], 4),
            "p99_ms": round(st["p99_ms"], 4), "peak_active": st["peak_active"],
            "mean_active": round(st["mean_active"], 3), "triangles": [int(j["meshes"][0].F) for j in jobs]}


def poisson_leg(drh, S, lam, njobs, rate, max_active, iter_scale):
    """Config 5's serving stream: `njobs` mesh-fitting jobs submitted at Poisson arrival times
    (rate jobs/s, seed 7) to one server under a 288 GB HBM budget; reports per-iteration
    p50 / p99 latency, jobs/s and iters/s over the whole stream.  `iter_scale` shrinks the
    50-500 iteration counts to keep the run short (the same scale for every job)."""
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
        ids.append(srv.submit(j["meshes"], j["cams"], tg, it, j["edge_samples"], mode=j["mode"], spp=j["spp"], light=lam))
    res = [srv.wait(i) for i in ids]
    dt = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    assert all(x["status"] == 0 for x in res), [x["status"] for x in res]
    return {"workload": f"config5: {njobs} Lambert jobs at Poisson rate {rate}/s (F log-uniform 12..200k, 128-1024 px, "
                        f"1-8 views, 50-500 iterations x {iter_scale}), 288 GB budget, {max_active} slots",
            "jobs_per_s": round(njobs / dt, 3), "iters_per_s": round(iters_total / dt, 2),
            "p50_ms": round(st["p50_ms"], 4), "p99_ms": round(st["p99_ms"], 4), "peak_active": st["peak_active"],
            "mean_active": round(st["mean_active"], 3), "wall_s": round(dt, 3)}


def multiview_leg(drh, S, steps, N):
    """Config 4's multi-view adjoint: a displaced sphere of 199,808 triangles (radial noise
    0.02, seed 6), per-vertex colours, 16 cameras on a ring (radius 3, 20 deg), 512x512 each,
    spp 4; a fused step fits positions and colours (vertex-colour shading, so the interior
    pass produces the colour gradient).  Times the step with the reference's Hammersley
    interior samples (recomputed), and with aligned samples recomputed vs stored."""
    base = S.uv_sphere(224, 448)
    m = S.with_random_colors(S.radial_displace(base, 0.02, seed=6, smooth=False), 6)
    t = S.with_random_colors(S.radial_displace(base, 0.02, seed=16, smooth=False), 16)
    cams = S.cameras_ring(16, 512, 512)
    rt = drh.DiffRender(mode=drh.VERTEX_COLOR, spp=4)
    rt.commit([t])
    targets = rt.render(cams)
    rt.close()
    out = {"workload": f"config4: {m.F} triangles / {m.V} vertices, 16 views x 512x512, spp 4, vertex colours, "
                       f"{N} edge samples per view, Adam on positions + colours",
           "hit_buffer_bytes": 16 * 16 * 512 * 512 * 4}
    for name, aligned, store in (("hammersley_recompute", 0, 0), ("aligned_recompute", 1, 0), ("aligned_store", 1, 1)):
        r = drh.DiffRender(mode=drh.VERTEX_COLOR, spp=4)
        r.commit([m])
        r.set_interior_samples(aligned, store)
        r.opt_init(drh.OPT_ADAM, 1e-3, 1e-3)
        dt = r.device_array(np.concatenate([x.reshape(-1) for x in targets]))
        dl = r.device_array(np.zeros(1, np.float32))
        cs = drh.cameras_array(cams)
        for _ in range(2):
            r.opt_step_dev(cs, 16, dt.ptr, N, dl.ptr)
        r.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            r.opt_step_dev(cs, 16, dt.ptr, N, dl.ptr)
        r.sync()
        out[name + "_ms_per_step"] = round(1e3 * (time.perf_counter() - t0) / steps, 3)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--serve-jobs", type=int, default=4, help="concurrent config-2 jobs in the served leg (0: skip the plain and served legs)")
    ap.add_argument("--serve-iters", type=int, default=50, help="Adam iterations per served job")
    ap.add_argument("--plain-iters", type=int, default=20, help="iterations of the plain host-driven loop")
    ap.add_argument("--no-serve", action="store_true", help="skip the plain and served legs")
    ap.add_argument("--mix-jobs", type=int, default=64, help="jobs in the config-3 uneven mix (0: skip)")
    ap.add_argument("--mix-iters", type=int, default=30, help="Adam iterations per mix job")
    ap.add_argument("--mix-active", type=int, default=16, help="server slots for the mix")
    ap.add_argument("--multiview-steps", type=int, default=5, help="timed config-4 steps per interior mode (0: skip)")
    ap.add_argument("--stream-jobs", type=int, default=24, help="jobs in the config-5 Poisson stream (0: skip)")
    ap.add_argument("--stream-rate", type=float, default=20.0, help="Poisson arrival rate, jobs/s")
    ap.add_argument("--stream-iter-scale", type=float, default=0.1, help="scale of the 50-500 iteration counts")
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

    r = drh.DiffRender(mode=drh.LAMBERT, spp=spp, light=lam)
    r.commit(meshes)
    P = r.grad_size()
    r.opt_init(drh.OPT_ADAM, 1e-3, 1e-3)
    d_target = r.device_array(target.reshape(-1))
    d_loss = r.device_array(np.zeros(1, np.float32))
    cams = drh.cameras_array([cam])

    def step():
        r.opt_step_dev(cams, 1, d_target.ptr, N, d_loss.ptr)

    for _ in range(args.warmup):
        step()
    r.sync()
    loss0 = float(d_loss.download(1)[0])

    # timed steps: the fused step as callers run it (captured once, replayed as a hipGraph)
    r.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    r.sync()
    t1 = time.perf_counter()
    loss1 = float(d_loss.download(1)[0])
    dt = t1 - t0
    ms = 1e3 * dt / args.steps

    # the same steps again, eager with a HIP event pair around every kernel (per-kernel times
    # on the library's stream; the events keep the step out of the graph)
    r.profile(True)
    r.sync()
    g0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    r.sync()
    events_ms = 1e3 * (time.perf_counter() - g0) / args.steps
    kernels = ["k_render", "k_interior", "k_edge_trace", "k_edge_grad", "k_cdf_par", "k_edge_bin",
               "k_suspects", "k_radix_sort", "k_refit", "k_karras",
               "k_tri_setup", "k_morton", "k_reorder", "k_pack_nodes", "k_prepare", "k_project", "k_edge_len",
               "k_loss_final", "k_fill_f32", "k_fill_u32", "k_adam"]
    prof = {}
    for k in kernels:
        tot, cnt = r.profile_read(k)
        if cnt:
            prof[k] = {"ms_per_launch": tot / cnt, "launches_per_step": cnt / args.steps}
    r.profile(False)
    dom = max(("k_render", "k_interior", "k_edge_trace", "k_edge_grad"), key=lambda k: prof.get(k, {}).get("ms_per_launch", 0))
    tiles = ((W + 7) // 8) * ((H + 7) // 8)
    E = len(r.edges()[0])
    sq = int(np.sqrt(spp))
    bytes_dom = algorithmic_bytes(dom, F, V, E, W, H, sq * sq, N, P, tiles)
    achieved = bytes_dom / (prof[dom]["ms_per_launch"] * 1e-3) / 1e9
    rays_per_step = W * H * sq * sq * 2 + 2 * N     # primary ray queries (render, interior, two per edge sample)

    result = {
        "metric": "fwd+bwd iters/s",
        "value": round(args.steps / dt, 3),
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
        "config": {"workload": "config2: icosphere-5 (10242 V / 20480 F), Lambert, 1024x1024, spp 4, 2^20 edge samples,"
                               " fused commit(LBVH)+render+loss+d_render+Adam per step on the vertex positions"
                               " (texture frozen, so the interior pass -- texture adjoint only in Lambert -- is left"
                               " out; the plain leg computes the full gradient)",
                   "triangles": F, "vertices": V, "resolution": [W, H], "spp": spp, "edge_samples": N,
                   "parallelism": "single GPU"},
        "primary_rays_per_s": round(rays_per_step * args.steps / dt, 1),
        "edge_samples_per_s": round(N * args.steps / dt, 1),
        "loss_first_last": [loss0, loss1],
        "ms_per_step_with_kernel_events": round(events_ms, 4),
        "kernels": {k: round(v["ms_per_launch"], 5) for k, v in prof.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(dom),
                     "algorithmic_bytes": bytes_dom},
        "cpu_baseline": None,
    }

    if not args.no_serve:
        if args.serve_jobs > 0:
            plain = plain_leg(drh, meshes, cam, target, N, spp, lam, args.plain_iters)
            served = served_leg(drh, S, meshes, cam, N, spp, lam, args.serve_jobs, args.serve_iters)
            result["plain_one_job_iters_per_s"] = round(plain, 3)
            result["served"] = served
            result["served_over_plain"] = round(served["iters_per_s"] / plain, 3)
            result["p50_ms"], result["p99_ms"] = served["p50_ms"], served["p99_ms"]
        if args.mix_jobs > 0:
            result["mix"] = mix_legs(drh, S, lam, args.mix_jobs, args.mix_iters, args.mix_active)
        if args.multiview_steps > 0:
            result["multiview"] = multiview_leg(drh, S, args.multiview_steps, 1 << 18)
        if args.stream_jobs > 0:
            result["stream"] = poisson_leg(drh, S, lam, args.stream_jobs, args.stream_rate, 8, args.stream_iter_scale)

    if not args.no_cpu:
        import oracle as O
        ref = O.Scene(meshes, O.TRACER_BVH)
        tgt = target
        cores = os.cpu_count()
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            pass
        iters = 0
        c0 = time.perf_counter()
        while True:
            img = ref.render(4, spp, [cam], lam=lam)[0]
            _, adj = O.loss(img, tgt)
            ref.d_render(4, spp, [cam], [adj], N, lam=lam)
            iters += 1
            if time.perf_counter() - c0 > args.cpu_seconds or iters >= 5:
                break
        cdt = time.perf_counter() - c0
        omp = int(os.environ.get("OMP_NUM_THREADS", cores))
        result["cpu_baseline"] = {
            "value": round(iters / cdt, 5), "unit": "iters/s", "cores": min(omp, cores), "kind": "port",
            "sample": f"{iters} full config-2 fwd+bwd iterations (render+loss+d_render) of the oracle with its CPU "
                      f"BVH tracer, OpenMP over host cores (Embree stand-in; Embree binary absent)",
        }
        result["gpu_over_cpu"] = round(result["value"] / result["cpu_baseline"]["value"], 2)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

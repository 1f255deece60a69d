"""Benchmark: primary rays/s at 4K on one MI355X, full pass sequence.

Workload (one "step" = one frame): the 16x4x16-chunk seeded world with a
5-level clipmap, frame f of the 120-frame seeded camera flight, rendered at
3840x2160 (render scale 1.0) through primary -> local TAA -> godray ->
composite -> composite TAA.  Inputs are resident in HBM before timing starts;
each step updates the clipmap for the frame's camera (Clipmap::update with the
new texels generated on the device, as the reference's frame begins), writes
the 464-byte camera and launches the frame.

Prints one JSON line.  value = primary rays per second of the whole frame
(rays per frame / wall time per frame).  Also reports the primary kernel's
own rate, per-pass milliseconds, p50/p99 per-frame latency, the primary
kernel's roofline position (algorithmic bytes from the counting variant of
the same kernel / its event-timed duration, against 8 TB/s) and a CPU
baseline: the oracle (C++ restatement of the reference passes) on a bounded
sample of the same frame on the host cores.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "voxel-engine-v0_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--res", default="3840x2160")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-views", type=int, default=64,
                    help="views of the mixed-resolution batching workload (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-1080p", action="store_true", help="skip the 1080p per-view latency run")
    ap.add_argument("--query-batches", type=int, default=256,
                    help="ray batches of the incoherent ray-query workload (0: skip it)")
    ap.add_argument("--serve-target-ms", type=float, default=100.0,
                    help="the serving leg's latency target (admission and launch size); 0 runs without one")
    ap.add_argument("--serve-seconds", type=float, default=6.0,
                    help="duration of the serving workload's arrival schedule (0: skip it)")
    ap.add_argument("--single-batch", action="store_true",
                    help="render the 4K view through a one-view batch (work-stealing launches) instead of vx_render")
    ap.add_argument("--no-chunk-builds", action="store_true", help="skip the chunk builder measurement")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed views + ray batches workload")
    ap.add_argument("--only", choices=["mixed", "chunk_builds", "ray_queries", "serving", "serving_device_builds",
                                       "serving_1080p", "config0"],
                    help="run one workload alone and print its result (not the benchmark line)")
    ap.add_argument("--static-clipmap", action="store_true",
                    help="keep frame 0's clipmap instead of updating it every frame")
    ap.add_argument("--no-count", action="store_true", help="skip the counting run (roofline)")
    ap.add_argument("--cpu-rows", type=int, default=0, help="8-pixel tile rows of the CPU sample (0: the whole frame)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device three times (the plain copy kernel that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


TRAFFIC_FILE = REPO / "profiles" / "k_primary_batch_traffic.md"
PROFILE_FILE = REPO / "profiles" / "k_primary_batch_4k_flight.md"
QUERY_PROFILE_FILE = REPO / "profiles" / "k_query_32x4x32.md"
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9  # CUs x SIMDs x lanes x clock: 39.3 T non-packed lane-ops/s


def kernel_build() -> str:
    """A digest of the kernel sources, the public header and the build flags: a
    profile's figures describe this build only if it carries the same digest."""
    h = hashlib.sha256()
    pkg = REPO / "voxel-engine-v0_amd"
    for path in sorted((pkg / "csrc").iterdir()) + [pkg / "build.py", REPO / "include" / "voxhip.h"]:
        h.update(path.name.encode())
        h.update(path.read_bytes())
    return h.hexdigest()[:16]


def profile_keys(path: Path) -> dict:
    """The `key: value` lines of a committed profile summary."""
    out = {}
    if path.exists():
        for line in path.read_text().splitlines():
            key, sep, val = line.partition(":")
            if sep and key and " " not in key.strip() and not key.startswith(("|", "#", "*", "`")):
                out[key.strip()] = val.strip()
    return out


def profile_of_this_build(path: Path):
    """The profile's keys if it was made with this kernel build, else (None, reason)."""
    keys = profile_keys(path)
    if not keys:
        return None, f"{path.relative_to(REPO)} is missing"
    if keys.get("kernel_build") != kernel_build():
        return None, (f"{path.relative_to(REPO)} describes kernel build {keys.get('kernel_build')}, "
                      f"not this one ({kernel_build()})")
    return keys, None


def measured_traffic(W, H):
    """HBM bytes per k_primary_batch launch at this resolution, from the committed
    counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of this kernel
    build, or (None, reason)."""
    keys, why = profile_of_this_build(TRAFFIC_FILE)
    if keys is None:
        return None, why
    val, _, res = keys.get("traffic_bytes_per_launch", "").partition("@")
    if res.strip() != f"{W}x{H}":
        return None, f"no traffic figure at {W}x{H}"
    return int(val.strip()), None


def profiled_valu(path: Path = PROFILE_FILE):
    """SQ/TCC counter figures of a kernel from a committed rocprofv3 summary of
    the same workload and this kernel build (by default k_primary_batch at 4K),
    or (None, reason)."""
    keys, why = profile_of_this_build(path)
    if keys is None:
        return None, why
    try:
        return {"valu_insts_per_launch": float(keys["valu_insts_per_launch"]),
                "valu_busy": float(keys["valu_busy"]), "valu_lane_utilisation": float(keys["valu_lane_utilisation"]),
                "l2_hit_rate": float(keys["l2_hit_rate"]), "source": str(path.relative_to(REPO))}, None
    except (KeyError, ValueError) as e:
        return None, f"{path.relative_to(REPO)}: {e}"


def query_traffic(n):
    """HBM bytes of the fused k_query launch over n rays, from the committed
    counter summary of this kernel build (FETCH_SIZE + WRITE_SIZE,
    copy-calibrated), or (None, reason)."""
    keys, why = profile_of_this_build(QUERY_PROFILE_FILE)
    if keys is None:
        return None, why
    val, _, rays = keys.get("traffic_bytes_per_launch", "").partition("@")
    if rays.strip() != f"{n} rays":
        return None, f"no traffic figure for {n} rays"
    return int(val.strip()), None


def copy_calibration(mib):
    """Three device-to-device copies of MIB MiB through the HIP runtime's own
    copy kernel: the reference workload for the FETCH_SIZE / WRITE_SIZE scale."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(n))
    for _ in range(3):
        hip.hipMemcpy(dst, src, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              "(the path shards by request: one server per GPU, each with a full replica of the world)", file=sys.stderr)
        sys.exit(2)
    import voxhip as vx
    from voxhip import scenes

    if a.only:
        # One workload alone, its JSON object on one line (for profiling a leg).
        legs = {"mixed": lambda: mixed(vx, scenes), "chunk_builds": lambda: chunk_builds(vx, scenes),
                "config0": lambda: config0(vx, scenes),
                "ray_queries": lambda: ray_queries(vx, scenes, a.query_batches),
                "serving": lambda: serve(vx, scenes, a.serve_seconds, target_ms=a.serve_target_ms),
                "serving_device_builds": lambda: serve(vx, scenes, a.serve_seconds, target_ms=a.serve_target_ms,
                                                       device_build=True),
                "serving_1080p": lambda: serve_1080p(vx, scenes, 4.0)}
        print(json.dumps({a.only: legs[a.only]()}), flush=True)
        if a.copy_calib:
            copy_calibration(a.copy_calib)
        return

    W, H = (int(v) for v in a.res.lower().split("x"))
    t0 = time.time()
    sc = scenes.c1(resolution=(W, H))
    hw = vx.HostWorld(sc.params, sc.grid_origin, sc.grid_dims)
    clip, heights = scenes.scene_clipmap(sc, 0)
    ctx = vx.Context(node_capacity=len(hw.nodes), chunk_capacity=hw.chunk_count, grid_capacity=len(hw.grid),
                     clip_levels=5, clip_res=256)
    ctx.upload_world(hw)
    ctx.write_clipmap(clip, heights)
    view = vx.View(ctx, (W, H), (W, H))
    setup_s = time.time() - t0
    # Every frame starts with the clipmap update (the reference's frame order:
    # Clipmap::update, then its uploads, then the passes), heights generated on
    # the device; frame 0's first touch rebuilds every level.
    clip_state = None if a.static_clipmap else vx.ClipmapState(levels=5, res=256)

    def clip_update(f):
        if clip_state is not None:
            pose, t_s = sc.poses[f % len(sc.poses)]
            c, rects = clip_state.update(float(pose.pos[0]), float(pose.pos[2]), t_s)
            ctx.write_clipmap_rects(c, sc.params, rects)

    cams = []
    prev = None
    for f in range(a.warmup + a.steps):
        cam = scenes.frame_camera(sc, hw, f, prev_view_proj=prev, jitter=True)
        prev = np.array(cam.view_proj, np.float32)
        cams.append(cam)

    render = view.render
    if a.single_batch:
        one = vx.Batch(ctx, [view])
        render = one.render
    for f in range(a.warmup):
        clip_update(f)
        view.write_camera(cams[f])
        render(vx.PASS_FRAME)
    ctx.sync()

    # Timed region: K frames back to back.
    lat = []
    pass_ms = []
    t_start = time.perf_counter()
    for k in range(a.steps):
        f = a.warmup + k
        t_f = time.perf_counter()
        clip_update(f)
        view.write_camera(cams[f])
        render(vx.PASS_FRAME)
        if a.single_batch:
            ms = [one.timing(), 0.0, 0.0, 0.0, 0.0]  # the fused sequence is timed as a whole
        else:
            ms = view.timings()  # waits for the frame: per-frame latency is arrival -> result ready
        lat.append((time.perf_counter() - t_f) * 1e3)
        pass_ms.append(ms)
    ctx.sync()
    total_s = time.perf_counter() - t_start

    rays = W * H
    ms_per_step = total_s * 1e3 / a.steps
    pass_ms = np.array(pass_ms)
    prim_ms = float(np.mean(pass_ms[:, 0]))
    value = rays * a.steps / total_s

    roofline = roofline_valu = None
    counts = None
    passes = None
    if not a.no_count:
        # The counting forms of the passes over the first three timed frames:
        # primary (the reference algorithm's loads, then the loads the product
        # kernel executes), then godray.
        view.reset()
        tot = {n: 0 for n in vx.COUNT_NAMES}
        tot_x = {n: 0 for n in vx.COUNT_NAMES}
        tot_g = {n: 0 for n in vx.COUNT_NAMES}
        parts = {n: 0 for n in vx.COUNT_PARTS}
        parts_x = {n: 0 for n in vx.COUNT_PARTS}
        ncount = min(3, a.steps)
        ray_b = []
        view.counters(reset=True)
        for k in range(ncount):
            view.write_camera(cams[a.warmup + k])
            view.render(vx.PASS_PRIMARY | vx.PASS_COUNT)
            pr = view.count_parts()
            c = view.counters(reset=True)
            ray_b.append(view.read(vx.OUT_RAY_BYTES).ravel())
            view.render(vx.PASS_PRIMARY | vx.PASS_COUNT | vx.PASS_COUNT_EXECUTED)
            px = view.count_parts()
            cx = view.counters(reset=True)
            view.render(vx.PASS_GODRAY | vx.PASS_COUNT | vx.PASS_END_FRAME)
            cg = view.counters(reset=True)
            for n in tot:
                tot[n] += c[n]
                tot_x[n] += cx[n]
                tot_g[n] += cg[n]
            for n in parts:
                parts[n] += pr[n]
                parts_x[n] += px[n]
        counts = {n: v / ncount for n, v in tot.items()}
        bytes_per_launch = vx.algorithmic_bytes(counts)
        exec_per_launch = vx.algorithmic_bytes({n: v / ncount for n, v in tot_x.items()})
        rb = np.concatenate(ray_b).astype(np.float64)
        achieved = bytes_per_launch / (prim_ms * 1e-3) / 1e9
        achieved_x = exec_per_launch / (prim_ms * 1e-3) / 1e9
        traffic, traffic_note = measured_traffic(W, H)
        prof, prof_note = profiled_valu()
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                    "kernel": "k_primary_batch", "kernel_build": kernel_build(),
                    "algorithmic_bytes_per_launch": int(bytes_per_launch),
                    "algorithmic_bytes_per_ray": {"mean": round(bytes_per_launch / rays, 1),
                                                  "p50": round(float(np.percentile(rb, 50)), 1),
                                                  "p99": round(float(np.percentile(rb, 99)), 1),
                                                  "max": round(float(rb.max()), 1),
                                                  "sum_check": round(float(rb.sum()) / ncount / max(bytes_per_launch, 1), 6)},
                    "bytes_per_ray_by_part": {n: round(v / ncount / rays, 1) for n, v in parts.items()},
                    "executed": {"note": "the loads the product kernel makes (its exact work skips on)",
                                 "bytes_per_launch": int(exec_per_launch),
                                 "bytes_per_ray": round(exec_per_launch / rays, 1),
                                 "bytes_per_ray_by_part": {n: round(v / ncount / rays, 1) for n, v in parts_x.items()},
                                 "achieved": round(achieved_x, 2), "frac": round(achieved_x / HBM_PEAK_GBS, 5)},
                    "profiled": prof}
        if traffic is None:
            roofline["traffic_note"] = traffic_note
        if prof is None:
            roofline["profiled_note"] = prof_note
        else:
            # The VALU roof: lane-operations the kernel executes (VALU instructions
            # x 64 lanes x the measured lane utilisation) per second of this run's
            # primary pass, against 39.3 T non-packed lane-ops/s.
            lane_ops = prof["valu_insts_per_launch"] * 64.0 * prof["valu_lane_utilisation"]
            issue_ops = prof["valu_insts_per_launch"] * 64.0
            roofline_valu = {"bound": "valu", "achieved": round(lane_ops / (prim_ms * 1e-3) / 1e12, 3),
                             "peak": round(VALU_PEAK_LANE_OPS / 1e12, 3), "unit": "T lane-op/s",
                             "frac": round(lane_ops / (prim_ms * 1e-3) / VALU_PEAK_LANE_OPS, 4),
                             "issue_frac": round(issue_ops / (prim_ms * 1e-3) / VALU_PEAK_LANE_OPS, 4),
                             "lane_ops_per_ray": round(lane_ops / rays, 1),
                             "source": prof["source"]}
        # The other passes against the HBM roof: godray's counted loads, and the
        # streaming passes' fixed bytes per pixel (local TAA 48, composite 212,
        # composite TAA 64).
        godray_b = vx.algorithmic_bytes({n: v / ncount for n, v in tot_g.items()})
        pm = pass_ms.mean(axis=0)
        passes = {}
        for name, bts, ms in (("godray", godray_b, pm[2]), ("local_taa", 48 * rays, pm[1]),
                              ("composite", 212 * rays, pm[3]), ("composite_taa", 64 * rays, pm[4])):
            gbs = bts / (ms * 1e-3) / 1e9
            passes[name] = {"bytes_per_launch": int(bts), "ms": round(float(ms), 4), "achieved_gbs": round(gbs, 1),
                            "frac": round(gbs / HBM_PEAK_GBS, 4)}

    def leg(name, fn):
        # One progress line per workload, on stderr (stdout carries only the JSON line).
        t = time.time()
        r = fn()
        print(f"bench.py: {name} done in {time.time() - t:.1f} s", file=sys.stderr, flush=True)
        return r

    print(f"bench.py: {W}x{H} flight frames done ({total_s:.1f} s timed)", file=sys.stderr, flush=True)
    lat1080 = None if a.no_1080p else leg("1080p latency", lambda: latency_1080p(
        vx, scenes, sc, hw, ctx, max(a.steps, 20), a.warmup, not a.static_clipmap))
    clipmap = None if a.static_clipmap else leg("clipmap update", lambda: clipmap_costs(vx, sc, ctx))
    batched = None if a.batch_views <= 0 else leg("batching", lambda: batching(
        vx, scenes, sc, hw, ctx, a.batch_views, a.batch_steps))
    queries = None if a.query_batches <= 0 else leg("ray queries", lambda: ray_queries(vx, scenes, a.query_batches))
    mix = None if a.no_mixed else leg("mixed", lambda: mixed(vx, scenes))
    serving = None if a.serve_seconds <= 0 else leg("serving", lambda: serve(
        vx, scenes, a.serve_seconds, target_ms=a.serve_target_ms))
    serving_untargeted = serving_device = None
    if a.serve_seconds > 0 and a.serve_target_ms > 0:
        serving_untargeted = leg("serving, no target", lambda: serve(vx, scenes, a.serve_seconds, target_ms=0.0))
    if a.serve_seconds > 0:
        serving_device = leg("serving, device builds", lambda: serve(
            vx, scenes, a.serve_seconds, target_ms=a.serve_target_ms, device_build=True))
    serving1080 = None if a.serve_seconds <= 0 else leg("serving 1080p", lambda: serve_1080p(vx, scenes, 4.0))
    builds = None if a.no_chunk_builds else leg("chunk builds", lambda: chunk_builds(vx, scenes))

    cpu = None
    if not a.no_cpu_baseline:
        cpu = leg("cpu baseline", lambda: cpu_baseline(sc, hw, clip, heights, cams[a.warmup], W, H, a.cpu_rows,
                                                       frame=a.warmup))
    c0 = None if a.no_cpu_baseline else leg("config 0", lambda: config0(vx, scenes))

    out = {
        "metric": "primary rays/s at 4K + p50/p99 per-view latency, served batched vs plain (1 GPU)",
        "value": round(value, 1),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seed-42 world, seeded camera flight)",
        "config": {"workload": f"plain one-view-per-launch full pipeline at {W}x{H}, 16x4x16-chunk world, "
                               f"5-level clipmap ({'static' if a.static_clipmap else 'updated every frame on the device'}), "
                               f"camera flight frames {a.warmup}..{a.warmup + a.steps - 1}",
                   "resolution": [W, H], "chunks": hw.chunk_count, "nodes": int(len(hw.nodes)),
                   "path": "plain", "primary_kernels": "persistent work-stealing over the view's 8x8 tiles"},
        "primary_kernel_rays_per_s": round(rays / (prim_ms * 1e-3), 1),
        "pass_ms_mean": {n: round(float(v), 4) for n, v in
                         zip(("primary", "local_taa", "godray", "composite", "composite_taa"), pass_ms.mean(axis=0))},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "roofline": roofline,
        "roofline_valu": roofline_valu,
        "pass_roofline_hbm": passes,
        "cpu_baseline": cpu,
        "config0_primary_640x360": c0,
        "latency_1080p_ms": lat1080,
        "clipmap_update": clipmap,
        "batched": batched,
        "ray_queries": queries,
        "mixed_batched_vs_plain": mix,
        "serving": serving,
        "serving_without_latency_target": serving_untargeted,
        "serving_device_builds": serving_device,
        "serving_1080p": serving1080,
        "chunk_builds": builds,
        "setup_s": round(setup_s, 2),
    }
    if counts is not None:
        out["primary_counts_per_launch"] = {k: int(v) for k, v in counts.items()}
    print(json.dumps(out), flush=True)
    if a.copy_calib:
        copy_calibration(a.copy_calib)


def latency_1080p(vx, scenes, sc, hw, ctx, steps, warmup, clip_every_frame=True):
    """Per-view frame latency of one 1920x1080 view along the same flight
    (clipmap update, camera write -> all five passes done), p50/p99 in ms."""
    res = (1920, 1080)
    view = vx.View(ctx, res, res)
    st = vx.ClipmapState(levels=5, res=256) if clip_every_frame else None
    lat, prev = [], None
    for f in range(warmup + steps):
        cam = scenes.frame_camera(sc, hw, f, prev_view_proj=prev, jitter=True, render=res)
        prev = np.array(cam.view_proj, np.float32)
        t = time.perf_counter()
        if st is not None:
            pose, t_s = sc.poses[f % len(sc.poses)]
            c, rects = st.update(float(pose.pos[0]), float(pose.pos[2]), t_s)
            ctx.write_clipmap_rects(c, sc.params, rects)
        view.write_camera(cam)
        view.render(vx.PASS_FRAME)
        view.timings()
        if f >= warmup:
            lat.append((time.perf_counter() - t) * 1e3)
    view.close()
    return {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
            "frames": steps}


def clipmap_costs(vx, sc, ctx, rebuilds=5):
    """Clipmap::update and its uploads per frame along the 120-frame flight, and
    a whole-clipmap rebuild (first touch: 5 levels x 256^2 texels), two ways:
    heights on the host then patch uploads (the reference's path), and heights
    generated on the device by the fill kernel.  Wall milliseconds per call."""
    def run(device, frames):
        st = vx.ClipmapState(levels=5, res=256)
        per, texels = [], 0
        for x, z, t in frames:
            t0 = time.perf_counter()
            c, rects = st.update(x, z, t)
            if device:
                ctx.write_clipmap_rects(c, sc.params, rects)
            else:
                ctx.write_clipmap_patches(c, rects, vx.clipmap_fill_rects(sc.params, rects))
            per.append((time.perf_counter() - t0) * 1e3)
            texels += sum(r.w * r.h for r in rects)
        return per, texels

    flight = [(float(p.pos[0]), float(p.pos[2]), t) for p, t in sc.poses]
    out = {}
    for name, device in (("host_fill", False), ("device_fill", True)):
        full = []
        for _ in range(rebuilds):
            per, _ = run(device, flight[:1])
            full.append(per[0])
        per, texels = run(device, flight)
        steady = per[1:]
        out[name] = {"rebuild_ms": round(float(np.median(full)), 3),
                     "frame_ms_mean": round(float(np.mean(steady)), 4),
                     "frame_ms_p99": round(float(np.percentile(steady, 99)), 4),
                     "texels_per_frame": round((texels - 5 * 256 * 256) / max(len(steady), 1), 1)}
    out["rebuild_texels"] = 5 * 256 * 256
    out["rebuild_speedup"] = round(out["host_fill"]["rebuild_ms"] / max(out["device_fill"]["rebuild_ms"], 1e-9), 2)
    return out


def batching(vx, scenes, sc, hw, ctx, n_views, steps):
    """The mixed-resolution multi-view workload: n_views views with seeded
    resolutions from {320x180 .. 3840x2160} and seeded poses over the same
    world, each step one full frame of every view.  Plain = one view per
    launch sequence (vx_render per view); batched = one fused launch per pass
    (vx_batch_render), after the autotuner picked its configuration."""
    res, frames = scenes.batch_views(n_views, len(sc.poses))
    rays = int(sum(w * h for w, h in res))

    def cams(step):
        return [scenes.frame_camera(sc, hw, int(frames[i]) + step, jitter=True, render=res[i]) for i in range(n_views)]

    def run(views, render_step):
        times = []
        for step in range(steps + 1):
            cs = cams(step)
            ctx.sync()
            t = time.perf_counter()
            render_step(views, cs)
            ctx.sync()
            if step > 0:  # step 0 warms up
                times.append(time.perf_counter() - t)
        return float(np.mean(times))

    def plain_step(views, cs):
        for v, c in zip(views, cs):
            v.write_camera(c)
            v.render(vx.PASS_FRAME)

    views = [vx.View(ctx, r, r) for r in res]
    t_plain = run(views, plain_step)
    for v in views:
        v.close()
    views = [vx.View(ctx, r, r) for r in res]
    batch = vx.Batch(ctx, views)
    for v, c in zip(views, cams(0)):
        v.write_camera(c)
    tuned = batch.tune(vx.PASS_FRAME, iters=1).as_dict()

    def batched_step(vs, cs):
        for v, c in zip(vs, cs):
            v.write_camera(c)
        batch.render(vx.PASS_FRAME)

    t_batch = run(views, batched_step)
    batch.close()
    for v in views:
        v.close()
    return {"workload": f"{n_views} views, resolutions uniform over 320x180..3840x2160 (seed 7), full pipeline, "
                        f"same world and clipmap", "rays_per_step": rays,
            "plain_rays_per_s": round(rays / t_plain, 1), "batched_rays_per_s": round(rays / t_batch, 1),
            "speedup": round(t_plain / t_batch, 3), "plain_ms_per_step": round(t_plain * 1e3, 2),
            "batched_ms_per_step": round(t_batch * 1e3, 2), "tuned": tuned}


def mixed(vx, scenes, n_views=64, n_batches=256, steps=2):
    """The mixed workload of the target: the batching leg's 64 views (the same
    seeded resolutions from {320x180 .. 3840x2160}) and the ray-query leg's 256
    batches (the same seeded sizes, log-uniform 1e3..1e6 rays), over one world,
    the 32x4x32-chunk seed-42 world with views along seeded flights.  Plain: one
    launch sequence per view (vx_render) and one query launch per ray batch,
    all on the context stream.  Batched: the views as one fused batch
    (vx_batch_render, autotuned), with all ray batches as one set of rays
    traced inside the views' persistent primary launch (vx_batch_set_rays),
    so ray blocks fill the time the last view tiles take; the batch and the
    query configuration are autotuned."""
    sc = scenes.c4(streams=n_views, frames=200)
    res, _ = scenes.batch_views(n_views, len(sc.poses))
    sizes_q, rays = scenes.query_batches(n_batches, sc.grid_origin, sc.grid_dims)
    n_rays = int(sizes_q.sum())
    hw = vx.HostWorld(sc.params, sc.grid_origin, sc.grid_dims)
    clip, heights = scenes.scene_clipmap(sc, 0)
    ctx = vx.Context(node_capacity=len(hw.nodes), chunk_capacity=hw.chunk_count, grid_capacity=len(hw.grid),
                     clip_levels=5, clip_res=256)
    ctx.upload_world(hw)
    ctx.write_clipmap(clip, heights)
    qcam = vx.camera_build(vx.make_pose((1.0, 1.0, 1.0), 0.0, 0.0), (64, 64), (64, 64), 0, grid_origin=sc.grid_origin,
                           grid_dims=sc.grid_dims, chunk_count=hw.chunk_count)
    d_rays = vx.DeviceArray.from_numpy(rays)
    d_hits = vx.DeviceArray(n_rays * vx.HIT_DTYPE.itemsize)
    offs = np.concatenate([[0], np.cumsum(sizes_q)])
    stride = vx.HIT_DTYPE.itemsize
    view_rays = int(sum(w * h for w, h in res))

    def cams(step):
        out = []
        for i, (w, h) in enumerate(res):
            pose, t_s = sc.stream_poses[i][step % len(sc.stream_poses[i])]
            out.append(vx.camera_build(pose, (w, h), (w, h), step, time_s=t_s, grid_origin=sc.grid_origin,
                                       grid_dims=sc.grid_dims, chunk_count=hw.chunk_count, jitter=True))
        return out

    def run(step_fn):
        times = []
        for step in range(steps + 1):
            cs = cams(step)
            ctx.sync()
            t = time.perf_counter()
            step_fn(cs)
            if step > 0:  # step 0 warms up
                times.append(time.perf_counter() - t)
        return float(np.mean(times))

    views = [vx.View(ctx, r, r) for r in res]

    def plain_step(cs):
        for v, c in zip(views, cs):
            v.write_camera(c)
            v.render(vx.PASS_FRAME)
        for i in range(n_batches):
            k = int(sizes_q[i])
            vx.query_rays_device(ctx, qcam, d_rays.view(int(offs[i]) * 32, k * 32),
                                 d_hits.view(int(offs[i]) * stride, k * stride), k)
        ctx.sync()

    t_plain = run(plain_step)
    for v in views:
        v.close()
    views = [vx.View(ctx, r, r) for r in res]
    batch = vx.Batch(ctx, views)
    for v, c in zip(views, cams(0)):
        v.write_camera(c)
    # The ray batches ride in the views' persistent primary launch
    # (vx_batch_set_rays): the query kernel's configuration is tuned alone,
    # then the fused launch with the rays attached.
    tuned_rays = vx.query_tune(ctx, qcam, d_rays, d_hits, n_rays).as_dict()
    batch.set_rays(qcam, d_rays, d_hits, n_rays)
    tuned_views = batch.tune(vx.PASS_FRAME, iters=2).as_dict()

    def batched_step(cs):
        for v, c in zip(views, cs):
            v.write_camera(c)
        batch.render(vx.PASS_FRAME)
        ctx.sync()

    t_batch = run(batched_step)
    batch.set_rays(None, None, None, 0)
    batch.close()
    for v in views:
        v.close()
    d_rays.close()
    d_hits.close()
    ctx.close()
    total = view_rays + n_rays
    return {"workload": f"{n_views} views (320x180..3840x2160, seed 7) + {n_batches} ray batches (1e3..1e6 rays, "
                        f"seed 11) per step, 32x4x32-chunk world + 5-level clipmap, full pipeline for the views",
            "view_rays": view_rays, "query_rays": n_rays,
            "plain_rays_per_s": round(total / t_plain, 1), "batched_rays_per_s": round(total / t_batch, 1),
            "speedup": round(t_plain / t_batch, 3), "plain_ms_per_step": round(t_plain * 1e3, 2),
            "batched_ms_per_step": round(t_batch * 1e3, 2), "tuned_views": tuned_views, "tuned_rays": tuned_rays}


def ray_queries(vx, scenes, n_batches):
    """The incoherent ray-query workload: n_batches batches of log-uniform
    sizes in [1e3, 1e6] rays (seed 11), origins uniform in the 32x4x32-chunk
    world's AABB, directions uniform on the sphere, max distance 512 voxels.
    Fused = all batches in one vx_query_rays launch (after the autotuner);
    plain = one launch per batch."""
    dims, origin = (32, 4, 32), (0, 0, 0)
    sizes, rays = scenes.query_batches(n_batches, origin, dims)
    n = int(sizes.sum())
    t0 = time.time()
    hw = vx.HostWorld(vx.world_params(42), origin, dims)
    build_s = time.time() - t0
    ctx = vx.Context(node_capacity=len(hw.nodes), chunk_capacity=hw.chunk_count, grid_capacity=len(hw.grid),
                     clip_levels=1, clip_res=256)
    ctx.upload_world(hw)
    cam = vx.camera_build(vx.make_pose((1.0, 1.0, 1.0), 0.0, 0.0), (64, 64), (64, 64), 0, grid_origin=origin,
                          grid_dims=dims, chunk_count=hw.chunk_count)
    d_rays = vx.DeviceArray.from_numpy(rays)
    d_hits = vx.DeviceArray(n * vx.HIT_DTYPE.itemsize)
    tuned = vx.query_tune(ctx, cam, d_rays, d_hits, n).as_dict()
    vx.query_rays_device(ctx, cam, d_rays, d_hits, n)  # warm
    ctx.sync()
    t = time.perf_counter()
    vx.query_rays_device(ctx, cam, d_rays, d_hits, n)
    ctx.sync()
    t_fused = time.perf_counter() - t
    gpu_fused_ms = vx.query_timing(ctx)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    stride = vx.HIT_DTYPE.itemsize
    t = time.perf_counter()
    for i in range(n_batches):
        k = int(sizes[i])
        vx.query_rays_device(ctx, cam, d_rays.view(int(offs[i]) * 32, k * 32), d_hits.view(int(offs[i]) * stride, k * stride), k)
    ctx.sync()
    t_plain = time.perf_counter() - t
    hits = d_hits.download(np.empty(n, vx.HIT_DTYPE))
    # The query's roofline: the counting forms of the same rays (one ray per
    # lane, per-lane load counters) give the reference algorithm's bytes and the
    # bytes k_query loads; both over the fused launch's HIP-event time.
    c_ref, p_ref = vx.query_count(ctx, cam, d_rays, d_hits, n)
    c_exe, _ = vx.query_count(ctx, cam, d_rays, d_hits, n, executed=True)
    d_rays.close()
    d_hits.close()
    b_ref, b_exe = vx.algorithmic_bytes(c_ref), vx.algorithmic_bytes(c_exe)
    achieved = b_ref / (gpu_fused_ms * 1e-3) / 1e9
    achieved_x = b_exe / (gpu_fused_ms * 1e-3) / 1e9
    traffic, traffic_note = query_traffic(n)
    prof, prof_note = profiled_valu(QUERY_PROFILE_FILE)
    roofline = {"bound": "hbm", "kernel": "k_query", "kernel_build": kernel_build(), "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "algorithmic_bytes_per_launch": int(b_ref), "algorithmic_bytes_per_ray": round(b_ref / n, 1),
                "loads_per_ray": {k: round(v / n, 3) for k, v in c_ref.items()},
                "executed": {"bytes_per_launch": int(b_exe), "bytes_per_ray": round(b_exe / n, 1),
                             "achieved": round(achieved_x, 2), "frac": round(achieved_x / HBM_PEAK_GBS, 5)},
                "profiled": prof}
    if traffic is None:
        roofline["traffic_note"] = traffic_note
    if prof is None:
        roofline["profiled_note"] = prof_note
    else:
        # the VALU roof as for the primary pass: lane-operations executed per second
        lane_ops = prof["valu_insts_per_launch"] * 64.0 * prof["valu_lane_utilisation"]
        issue_ops = prof["valu_insts_per_launch"] * 64.0
        roofline["valu"] = {"bound": "valu", "achieved": round(lane_ops / (gpu_fused_ms * 1e-3) / 1e12, 3),
                            "peak": round(VALU_PEAK_LANE_OPS / 1e12, 3), "unit": "T lane-op/s",
                            "frac": round(lane_ops / (gpu_fused_ms * 1e-3) / VALU_PEAK_LANE_OPS, 4),
                            "issue_frac": round(issue_ops / (gpu_fused_ms * 1e-3) / VALU_PEAK_LANE_OPS, 4),
                            "lane_ops_per_ray": round(lane_ops / n, 1)}
    return {"workload": f"{n_batches} batches, log-uniform 1e3..1e6 rays (seed 11), 32x4x32-chunk world, origins uniform "
                        f"in its AABB, directions uniform, tmax 51.2 m", "rays": n, "chunks": hw.chunk_count,
            "world_build_s": round(build_s, 2), "fused_rays_per_s": round(n / t_fused, 1),
            "fused_gpu_ms": round(gpu_fused_ms, 3), "plain_rays_per_s": round(n / t_plain, 1),
            "speedup": round(t_plain / t_fused, 3), "hit_fraction": round(float(np.mean(hits["kind"] == 1)), 4),
            "tuned": tuned, "roofline": roofline}


def serve(vx, scenes, seconds, rate=400.0, edit_rate=1000.0, n_streams=32, budget=64 << 30, target_ms=100.0,
          device_build=False):
    """The serving workload: Poisson arrivals (rate req/s, seed 5) for `seconds`,
    half render frames of n_streams client view streams (resolutions drawn
    as in the batching workload, each stream on its own seeded flight, at most
    one frame pending per stream) and half ray-query batches (log-uniform
    1e3..1e6 rays), against the 32x4x32-chunk world with a 5-level clipmap,
    with a concurrent stream of edit_rate random single-voxel edits/s, and
    admission against a `budget`-byte device memory budget and, when
    target_ms > 0, against that latency target (the server also bounds each
    launch to a third of the target by its launch-cost model).  device_build: edited chunks
    are rebuilt by the device chunk builder from HBM-resident material instead
    of on host worker threads."""
    import threading

    rng = np.random.default_rng(5)
    sc = scenes.c4(streams=n_streams, frames=2000)
    hw = vx.HostWorld(sc.params, sc.grid_origin, sc.grid_dims)
    clip, heights = scenes.scene_clipmap(sc, 0)
    ctx = vx.Context(node_capacity=int(1.5 * len(hw.nodes)), chunk_capacity=hw.chunk_count, grid_capacity=len(hw.grid),
                     clip_levels=5, clip_res=256)
    ctx.upload_world(hw)
    ctx.write_clipmap(clip, heights)
    sizes = [(320, 180), (640, 360), (1280, 720), (1920, 1080), (3840, 2160)]
    res = [sizes[i] for i in rng.integers(0, len(sizes), n_streams)]
    qcam = vx.camera_build(vx.make_pose((1.0, 1.0, 1.0), 0.0, 0.0), (64, 64), (64, 64), 0, grid_origin=sc.grid_origin,
                           grid_dims=sc.grid_dims, chunk_count=hw.chunk_count)
    srv = vx.Server(ctx, hw, memory_budget=budget, max_queued_frames=1, latency_target_ms=target_ms,
                    device_build=device_build)
    pool = scenes.query_rays(rng, 2_000_000, sc.grid_origin, sc.grid_dims, tmax_m=51.2)  # batches are slices of it
    # arrival schedule
    t_arr = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.5)))
    t_arr = t_arr[t_arr < seconds]
    kinds = rng.random(len(t_arr)) < 0.5  # True: render frame
    frame_of = [0] * n_streams
    stop = threading.Event()
    ext = sc.grid_dims[0] * 64, sc.grid_dims[2] * 64

    def edits():
        erng = np.random.default_rng(6)
        period = 0.02
        per = max(1, int(edit_rate * period))
        while not stop.is_set():
            x = erng.integers(0, ext[0], per)
            z = erng.integers(0, ext[1], per)
            y = np.array([vx.ground_height(sc.params, int(a), int(b)) for a, b in zip(x, z)]) + erng.integers(-3, 4, per)
            srv.edit(np.stack([x, y, z], 1), erng.integers(0, 6, per))
            stop.wait(period)

    editor = threading.Thread(target=edits, daemon=True)
    t0 = time.perf_counter()
    cpu0 = time.process_time()
    editor.start()
    submitted = rejected = 0
    view_rays = ray_rays = 0
    done = []
    for i, ta in enumerate(t_arr):
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        if kinds[i]:
            sid = int(rng.integers(0, n_streams))
            pose, t_s = sc.stream_poses[sid][frame_of[sid] % len(sc.stream_poses[sid])]
            w, h = res[sid]
            cam = vx.camera_build(pose, (w, h), (w, h), frame_of[sid], time_s=t_s, grid_origin=sc.grid_origin,
                                  grid_dims=sc.grid_dims, chunk_count=hw.chunk_count, jitter=True)
            ok = srv.submit_view(sid, cam, w, h, tag=i)
            if ok:
                frame_of[sid] += 1
                view_rays += w * h
        else:
            k = int(np.exp(rng.uniform(np.log(1e3), np.log(1e6))))
            off = int(rng.integers(0, len(pool) - k))
            ok = srv.submit_rays(qcam, pool[off:off + k], tag=i, out=np.empty(k, vx.HIT_DTYPE))
            if ok:
                ray_rays += k
        submitted += 1
        rejected += 0 if ok else 1
        while (r := srv.poll()) is not None:  # results release their hit buffers
            done.append((r.submit_ms, r.start_ms, r.done_ms))
    stop.set()
    editor.join()
    srv.drain()
    wall = time.perf_counter() - t0
    cpu_s = time.process_time() - cpu0
    while (r := srv.poll()) is not None:
        done.append((r.submit_ms, r.start_ms, r.done_ms))
    d = np.array(done) if done else np.zeros((1, 3))
    lat = d[:, 2] - d[:, 0]
    slow = d[lat >= np.percentile(lat, 99)]
    srv.wait_edits()
    st = srv.stats()
    srv.close()
    return {"workload": f"Poisson {rate:g} req/s for {seconds:g} s (seed 5), half frames of {n_streams} view streams "
                        f"(320x180..3840x2160, at most 1 pending frame each), half ray batches (1e3..1e6 rays, hits returned), "
                        f"{edit_rate:g} single-voxel edits/s, 32x4x32 chunks + 5-level clipmap, {budget >> 30} GB budget, "
                        + (f"{target_ms:g} ms latency target" if target_ms > 0 else "no latency target")
                        + (", edited chunks rebuilt on the device" if device_build else
                           ", edited chunks rebuilt on 8 host threads"),
            "requests": submitted, "rejected": rejected, "rejected_rate": round(rejected / max(1, submitted), 4),
            "completed": int(st["completed"]), "launches": int(st["launches"]),
            "rays_per_s": round((view_rays + ray_rays) / wall, 1), "view_rays": view_rays, "query_rays": ray_rays,
            "view_latency_ms": {"p50": round(st["view_p50_ms"], 2), "p99": round(st["view_p99_ms"], 2)},
            "rays_latency_ms": {"p50": round(st["rays_p50_ms"], 2), "p99": round(st["rays_p99_ms"], 2)},
            "edits_applied_chunks": int(st["edits_applied"]), "gpu_busy_frac": round(st["busy_ms"] / (wall * 1e3), 3),
            "host_cpu_s": round(cpu_s, 2), "wall_s": round(wall, 2),
            "slowest_1pct_ms": {"wait_for_launch": round(float(np.mean(slow[:, 1] - slow[:, 0])), 2),
                                "in_launch": round(float(np.mean(slow[:, 2] - slow[:, 1])), 2),
                                "launch_ms_p50": round(float(np.percentile(d[:, 2] - d[:, 1], 50)), 2),
                                "launch_ms_max": round(float(np.max(d[:, 2] - d[:, 1])), 2)}}


def serve_1080p(vx, scenes, seconds, streams=(2, 3, 4), fps=30.0, target_ms=16.0):
    """Interactive serving at 1080p: `n` client view streams of 1920x1080 on the
    16x4x16-chunk world with a 5-level clipmap, each submitting a frame every
    1/fps s (phases spread evenly over the period) along its own stretch of the
    seeded flight, at most one frame pending per stream (a frame whose
    predecessor is still pending is rejected: the client drops it), under a
    target_ms latency target; for each n in `streams`, p50/p99 of submit -> done
    over the served frames."""
    sc = scenes.c1()
    hw = vx.HostWorld(sc.params, sc.grid_origin, sc.grid_dims)
    clip, heights = scenes.scene_clipmap(sc, 0)
    res = (1920, 1080)
    out = []
    for n in streams:
        ctx = vx.Context(node_capacity=len(hw.nodes), chunk_capacity=hw.chunk_count, grid_capacity=len(hw.grid),
                         clip_levels=5, clip_res=256)
        ctx.upload_world(hw)
        ctx.write_clipmap(clip, heights)
        srv = vx.Server(ctx, hw, memory_budget=64 << 30, max_queued_frames=1, latency_target_ms=target_ms)
        # warm-up: one frame per stream (creates the resident views), not counted
        for k in range(n):
            pose, t_s = sc.poses[(37 * k) % len(sc.poses)]
            srv.submit_view(k, vx.camera_build(pose, res, res, 0, time_s=t_s, grid_origin=sc.grid_origin,
                                               grid_dims=sc.grid_dims, chunk_count=hw.chunk_count, jitter=True),
                            *res, tag=k)
        srv.drain()
        while srv.poll() is not None:
            pass
        base = srv.stats()
        ticks = sorted((j / fps + k / (fps * n), k, j) for k in range(n) for j in range(int(seconds * fps)))
        frame_of = [1] * n
        offered = dropped = 0
        done = []
        t0 = time.perf_counter()
        for ta, k, _ in ticks:
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            f = frame_of[k]
            pose, t_s = sc.poses[(37 * k + f) % len(sc.poses)]
            cam = vx.camera_build(pose, res, res, f, time_s=t_s, grid_origin=sc.grid_origin, grid_dims=sc.grid_dims,
                                  chunk_count=hw.chunk_count, jitter=True)
            offered += 1
            if srv.submit_view(k, cam, *res, tag=n + offered):
                frame_of[k] += 1
            else:
                dropped += 1
            while (r := srv.poll()) is not None:
                done.append(r.done_ms - r.submit_ms)
        srv.drain()
        wall = time.perf_counter() - t0
        while (r := srv.poll()) is not None:
            done.append(r.done_ms - r.submit_ms)
        st = srv.stats()
        srv.close()
        ctx.close()
        lat = np.array(done) if done else np.zeros(1)
        served = int(st["completed"] - base["completed"])
        out.append({"streams": n, "offered_fps": round(n * fps, 1), "served_fps": round(served / wall, 1),
                    "dropped_frames": dropped, "served_frames": served,
                    "rays_per_s": round(served * res[0] * res[1] / wall, 1),
                    "launches": int(st["launches"] - base["launches"]),
                    "view_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                                        "p99": round(float(np.percentile(lat, 99)), 2),
                                        "max": round(float(np.max(lat)), 2)}})
    return {"workload": f"1920x1080 view streams at {fps:g} fps each (evenly phased), 16x4x16 chunks + 5-level "
                        f"clipmap, at most 1 pending frame per stream (a frame whose predecessor is pending is "
                        f"dropped), {target_ms:g} ms latency target, {seconds:g} s per load; latency = submit -> done",
            "loads": out}


def chunk_builds(vx, scenes, repeats=3, cpu_sample=24):
    """Chunk builds of the 16x4x16-chunk world (1024 chunks, materials filled
    on the host beforehand and resident in HBM): the device builder's count +
    emit for the whole batch, against the host builder (vx_build_chunk) on one
    core over a sample of the same chunks."""
    import ctypes as C

    sc = scenes.c1()
    p = sc.params
    gx, gy, gz = sc.grid_dims
    coords = [(x, y, z) for z in range(gz) for y in range(gy) for x in range(gx)]
    mats, grounds = [], []
    for cx, cy, cz in coords:
        m, g = vx.fill_chunk(p, cx, cy, cz)
        mats.append(m)
        grounds.append(g)
    n = len(coords)
    oy = np.array([cy * 64 for _, cy, _ in coords], np.int32)
    d_mat = vx.DeviceArray.from_numpy(np.concatenate(mats))
    d_ground = vx.DeviceArray.from_numpy(np.concatenate(grounds))
    probe = vx.Context(node_capacity=16, chunk_capacity=1, grid_capacity=1)
    b = vx.ChunkBuilder(probe, max_chunks=n)
    counts = b.count(d_mat, d_ground, oy, p.dirt_depth)
    b.close()
    probe.close()
    total = int(counts.sum())
    ctx = vx.Context(node_capacity=total, chunk_capacity=n, grid_capacity=1)
    b = vx.ChunkBuilder(ctx, max_chunks=n)
    base = np.concatenate([[0], np.cumsum(counts[:-1].astype(np.uint64))]).astype(np.uint64)
    slots = np.arange(n, dtype=np.uint32)
    times = []
    for _ in range(repeats):
        t = time.perf_counter()
        b.count(d_mat, d_ground, oy, p.dirt_depth)
        b.emit(base, slots * 16, slots * 2048)
        times.append(time.perf_counter() - t)
    gpu_s = float(np.median(times))
    b.close()
    ctx.close()
    # host builder, one core
    cap = vx.MAX_CHUNK_NODES
    nodes = np.zeros(cap, vx.NODE_DTYPE)
    ropes = np.zeros(cap, vx.ROPES_DTYPE)
    macro = np.zeros(16, np.uint32)
    colinfo = np.zeros(2048, np.uint32)
    cnt = C.c_uint32()
    pick = np.linspace(0, n - 1, cpu_sample).astype(int)
    t = time.perf_counter()
    for i in pick:
        vx.LIB.vx_build_chunk(mats[i].ctypes.data, grounds[i].ctypes.data, int(oy[i]), p.dirt_depth, nodes.ctypes.data,
                              ropes.ctypes.data, cap, C.byref(cnt), macro.ctypes.data, colinfo.ctypes.data)
    cpu_s = (time.perf_counter() - t) / len(pick)
    return {"workload": f"{n} chunks of the seed-42 16x4x16 world ({total} nodes), materials resident in HBM",
            "device_ms": round(gpu_s * 1e3, 3), "device_chunks_per_s": round(n / gpu_s, 1),
            "device_nodes_per_s": round(total / gpu_s, 1),
            "host_1core_ms_per_chunk": round(cpu_s * 1e3, 3), "host_1core_chunks_per_s": round(1.0 / cpu_s, 1),
            "host_sample_chunks": int(len(pick))}


def config0(vx, scenes, runs=7, cpu_runs=5):
    """The reference's CPU-runnable case: the plain primary pass at 640x360 over
    the 8x4x8-chunk seed-42 world, one camera at the fixed seeded pose, frame 1
    (no history, the view reset before each run), 5-level clipmap.  GPU: median
    of the primary pass's event time over `runs`; CPU: the oracle's main_primary
    on the host cores, median of `cpu_runs` whole frames."""
    sys.path.insert(0, str(REPO))
    from oracle import oracle as O

    sc = scenes.c0()
    hw = vx.HostWorld(sc.params, sc.grid_origin, sc.grid_dims)
    clip, heights = scenes.scene_clipmap(sc)
    cam = scenes.frame_camera(sc, hw, 0)
    ctx = vx.Context(node_capacity=len(hw.nodes), chunk_capacity=hw.chunk_count, grid_capacity=len(hw.grid),
                     clip_levels=5, clip_res=256)
    ctx.upload_world(hw)
    ctx.write_clipmap(clip, heights)
    view = vx.View(ctx, sc.present, sc.render)
    ms = []
    for _ in range(runs + 1):
        view.reset()
        view.write_camera(cam)
        view.render(vx.PASS_PRIMARY | vx.PASS_END_FRAME)
        ms.append(view.timings()[0])
    view.close()
    ctx.close()
    gpu_ms = float(np.median(ms[1:]))
    threads = max(1, min(16, os.cpu_count() or 1))
    w = O.World.from_host(hw, clip, heights)
    cpu = []
    for _ in range(cpu_runs):
        fr = O.Frame(*sc.render)
        t = time.perf_counter()
        O.primary(w, cam, fr, threads=threads)
        cpu.append(time.perf_counter() - t)
    cpu_s = float(np.median(cpu))
    rays = sc.render[0] * sc.render[1]
    return {"workload": "plain primary pass, 640x360, 8x4x8-chunk seed-42 world, fixed seeded pose, frame 1, "
                        "5-level clipmap",
            "gpu_primary_ms_median": round(gpu_ms, 4), "gpu_rays_per_s": round(rays / (gpu_ms * 1e-3), 1),
            "cpu_ms_median": round(cpu_s * 1e3, 2), "cpu_rays_per_s": round(rays / cpu_s, 1), "cpu_cores": threads,
            "cpu_kind": "port", "gpu_over_cpu": round(cpu_s / (gpu_ms * 1e-3), 1)}


def cpu_baseline(sc, hw, clip, heights, cam, W, H, rows, frame=0):
    """The oracle's main_primary on a band of tile rows through the middle of
    the same 4K frame (by default all of it), timed on the host cores."""
    sys.path.insert(0, str(REPO))
    from oracle import oracle as O

    threads = max(1, min(16, os.cpu_count() or 1))
    tx = (W + 7) // 8
    ty = (H + 7) // 8
    if rows <= 0:
        rows = ty  # the whole frame (about 5 s on 16 host threads)
    r0 = max(0, ty // 2 - rows // 2)
    w = O.World.from_host(hw, clip, heights)
    fr = O.Frame(W, H)
    t = time.perf_counter()
    O.primary(w, cam, fr, tiles=(r0 * tx, (r0 + rows) * tx), threads=threads)
    dt = time.perf_counter() - t
    n = rows * 8 * W
    return {"value": round(n / dt, 1), "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"main_primary over tile rows {r0}..{r0 + rows - 1} ({n} rays) of flight frame {frame} at "
                      f"{W}x{H}, {dt:.1f} s"}


if __name__ == "__main__":
    main()

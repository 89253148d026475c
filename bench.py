"""Benchmark: closest-hit ray casting on one MI355X.

Workload (one "step" = one closest-hit pass over one batch resident in HBM):
2,097,152-triangle fBm heightfield (1024 x 1024 cells), octree split 16 /
depth 12, batch B = 4,194,304 incoherent rays (origins uniform in the scene
AABB, directions uniform on S^2), clipBackfaces = false.  The coherent 1080p
primary batch (A) is timed beside it.

Prints ONE JSON line (rank 0) with the metric, a roofline object for the
traversal kernel (HIP events on the launch stream) and the oracle CPU path
timed on this host as cpu_baseline.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0                      # MI355X HBM3E
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # lane-ops/s: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz
# Work of the reference algorithm per ray (counted by the oracle on the same rays):
# bytes = 24 ray in + 8 (id, t) out + 28 per box (6 fp32 bounds + link) + 4 per leaf-list entry
#         + 36 per triangle test (9 fp32 vertices); ops = 6 + 24 per box + 155 per triangle test
B_RAY, B_OUT, B_BOX, B_REF, B_TRI = 24, 8, 28, 4, 36
OPS_RAY, OPS_BOX, OPS_TRI = 6, 24, 155
# Work of the kernel's traversal per ray (counted by dr_cast_stats on the same rays): a group
# expansion reads 196 B (8 child boxes + the non-empty mask) and pops read 8 B (link, count); a leaf
# reads its 24-B box; a plane test reads the 4-B record index of the leaf entry and 32 B of the
# record, an inside test 32 B more.
# ops: 6 (1/d) + 12 per octant test (the children of an expanded group with a non-empty leaf below)
#      + 24 per slab test (those not behind the origin) + 25 per plane test
#      + 140 per inside test (6 edge vectors, 3 dots, 3 square roots, 3 divisions).
K_RAY, K_OUT, K_GROUP, K_POP, K_LEAF, K_PLANE, K_INSIDE = 24, 8, 196, 8, 24, 36, 32
KOPS_RAY, KOPS_OCTANT, KOPS_SLAB, KOPS_PLANE, KOPS_INSIDE = 6, 12, 24, 25, 140


# host threads of the CPU legs: the cores this process may run on (the lease's share of the host)
HOST_THREADS = len(os.sched_getaffinity(0))
# octree knobs of the path-trace and batch scenes: the reference's defaults, leaf split above
# 8 triangles and depth limit 24 (Octree.cuh:11,13); the cast workload keeps 16 / 12
REF_KNOBS = (8, 24)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(path):
    """Measured HBM bytes per launch from a committed rocprofv3 --pmc summary (or None)."""
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


# Counter passes this run launches (rocprofv3 --pmc, one child process each, on the profile-only
# mode of this script: the timed incoherent casts and a 256 MiB calibration copy).  Counter
# budget per pass: FETCH_SIZE = 3 TCC counters, WRITE_SIZE = 2, TCC_HIT/MISS 1 each, VALUBusy and
# VALUUtilization = 2 SQ + 1 GRBM counters.
PMC_PASSES = (("fetch", "FETCH_SIZE VALUBusy VALUUtilization"), ("write", "WRITE_SIZE TCC_HIT_sum TCC_MISS_sum"))
CALIB_BYTES = 268435456
CAST_KERNEL = "k_cast<false>"


def _read_counters(outdir):
    """{kernel name: {counter: [(dispatch id, value)]}} from rocprofv3 counter_collection.csv files."""
    import csv
    import glob
    res = {}
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                name, ctr, val = row.get("Kernel_Name", ""), row.get("Counter_Name", ""), row.get("Counter_Value", "")
                try:
                    res.setdefault(name, {}).setdefault(ctr, []).append((int(row.get("Dispatch_Id", 0)), float(val)))
                except ValueError:
                    pass
    return res


def measure_counters(steps=3, keep_dir=None):
    """Fabric bytes, L2 hit rate, VALU busy and VALU lane utilization of the cast kernel on the
    incoherent batch, from rocprofv3 counter passes run by this invocation on this GPU; None
    when rocprofv3 is missing or a pass fails.  With keep_dir the passes' CSV files stay there
    (one sub-directory per pass); otherwise they go to a temporary directory that is removed."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    base = os.path.abspath(keep_dir) if keep_dir else tempfile.mkdtemp(prefix="bench_pmc_", dir="/tmp")
    os.makedirs(base, exist_ok=True)
    env = dict(os.environ, TMPDIR="/tmp")
    per = {}
    for tag, counters in PMC_PASSES:
        out = os.path.join(base, tag)
        cmd = ["timeout", "-s", "KILL", "120", prof, "--pmc", *counters.split(), "--output-format", "csv", "-d", out,
               "-o", tag, "--", sys.executable, os.path.abspath(__file__), "--profile", "--calib-copy", "--steps",
               str(steps), "--warmup", "1"]
        log(f"counter pass {tag}: {counters}")
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        data = _read_counters(out)
        if not data:
            log(f"counter pass {tag} gave nothing (rc {r.returncode}): {r.stderr[-400:]}")
            return None
        for k, v in data.items():
            per.setdefault(k, {}).update(v)
    if not keep_dir:
        shutil.rmtree(base, ignore_errors=True)
    cast = next((v for k, v in per.items() if CAST_KERNEL in k), None)
    if not cast:
        return None
    # the calibration copy is the last copy kernel of the profile run (uploads of the workload
    # also run as copy kernels, earlier)
    copies = {}
    for k, v in per.items():
        if "copyBuffer" in k:
            for ctr, xs in v.items():
                copies.setdefault(ctr, []).extend(xs)
    if "FETCH_SIZE" not in copies or "WRITE_SIZE" not in copies:
        return None
    last = lambda xs: max(xs)[1]
    mean = lambda xs: float(np.mean([x for _, x in xs]))
    f_fac = CALIB_BYTES / (last(copies["FETCH_SIZE"]) * 1024)
    w_fac = CALIB_BYTES / (last(copies["WRITE_SIZE"]) * 1024)
    fetch = mean(cast["FETCH_SIZE"]) * 1024 * f_fac
    write = mean(cast["WRITE_SIZE"]) * 1024 * w_fac
    hit, miss = mean(cast["TCC_HIT_sum"]), mean(cast["TCC_MISS_sum"])
    return {"fetch_bytes": fetch, "write_bytes": write, "l2_hit_rate": hit / (hit + miss),
            "valu_busy": mean(cast["VALUBusy"]) / 100.0, "valu_lane_utilization": mean(cast["VALUUtilization"]) / 100.0,
            "calibration": {"copy_bytes": CALIB_BYTES, "fetch_factor": round(f_fac, 4), "write_factor": round(w_fac, 4)},
            "dispatches": len(cast["FETCH_SIZE"])}


def _event_ms(fn, k, stream):
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(k):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / k


def _device_scene(D, tris, split=16, depth=12):
    t = D.Octree(split, depth)
    t.push(tris)
    t.build()
    return t.upload()


def render_roofline(cnt, ms, pixels):
    """Roofline of a render from the work its kernel counted (dr_render_stats): the traversal
    lane-ops of every cast (as for the cast kernel, plus 6 ops of ray setup per cast) against
    the VALU peak, and the bytes those casts read (groups, leaf boxes and links, 32-B halves of
    triangle records; leaf-ordered records, so no record index) plus the 16-B framebuffer store
    of every pixel, against HBM.  Shading arithmetic (BSDF sampling, light sampling, Philox) is
    not counted, so the VALU fraction understates the kernel's instruction use."""
    seg, ge, lv, pl, ins = (cnt[k] for k in ("segments", "group_expansions", "leaves", "plane_tests", "inside_tests"))
    ops = (KOPS_RAY * seg + KOPS_OCTANT * cnt["live_children"] + KOPS_SLAB * cnt["slab_tests"] + KOPS_PLANE * pl
           + KOPS_INSIDE * ins)
    pops = max(ge - seg, 0) + lv
    byts = K_GROUP * ge + K_POP * pops + K_LEAF * lv + (K_PLANE - 4) * pl + K_INSIDE * ins + 16 * pixels
    t = ms / 1e3
    return {"bound": "valu", "achieved": round(ops / t / 1e12, 3), "peak": round(VALU_PEAK_TOPS, 2),
            "unit": "T lane-ops/s", "frac": round(ops / t / 1e12 / VALU_PEAK_TOPS, 4),
            "hbm_logical": {"achieved": round(byts / t / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(byts / t / 1e9 / HBM_PEAK_GBS, 4)},
            "per_frame": {"casts": seg, "photon_casts": cnt["photons"], "group_expansions": ge, "leaves": lv,
                          "plane_tests": pl, "inside_tests": ins, "live_children": cnt["live_children"],
                          "slab_tests": cnt["slab_tests"], "paths": cnt["paths"], "traversal_ops": ops,
                          "bytes": byts}}


def bench_path_trace(D, S, stream, cpu_seconds, with_cpu):
    """1920x1080, 16 spp (4x4), max depth 8, Russian roulette from depth 3 on three displaced
    level-7 icospheres (983,040 tris) + ground quad, Lambert / mirror / Schlick metal, two area lights."""
    import torch
    sc = S.icosphere_scene(level=7, seed=3)
    ds = _device_scene(D, sc.tris, *REF_KNOBS)
    W, H, spp, depth = 1920, 1080, (4, 4), 8
    g = D.DumbRenderer(ds, sc.mat_id, sc.materials, sc.lights, sc.camera, W, H, spp, depth, rr_start=3,
                       rr_survive=0.75)
    frame = D.BlockBasedFrameBuffer(W, H)
    ms = _event_ms(lambda: g.render(frame), 3, stream)
    out = {"workload": f"path trace {W}x{H}, 16 spp, depth 8, RR from 3 (survive 0.75), {len(sc.tris)} tris, "
                       "Lambert/mirror/Schlick metal, 2 area lights, octree split 8 / depth 24",
           "ms_per_frame": round(ms, 3), "mpix_spp_s": round(W * H * 16 / ms / 1e3, 2)}
    log(f"path trace 1080p16: {ms:.2f} ms/frame, {out['mpix_spp_s']} Mpix*spp/s")
    _, cnt = g.render_stats()
    out["roofline"] = render_roofline(cnt, ms, W * H)
    log(f"path trace counters: {cnt}; roofline {out['roofline']}")
    if with_cpu:
        import oracle
        ot = oracle.Octree(sc.tris, *REF_KNOBS)
        threads = HOST_THREADS
        o = oracle.Renderer(ot, sc.mat_id, sc.materials, sc.lights, sc.camera.as_array(), W, H, spp, depth,
                            rr_start=3, rr_survive=0.75)
        # a band of rows through the middle of the frame, grown until ~cpu_seconds/4
        rows = 8
        while True:
            r0 = H // 2 - rows // 2
            t0 = time.perf_counter()
            o.render(nthreads=threads, rows=(r0, r0 + rows))
            spent = time.perf_counter() - t0
            if spent >= cpu_seconds / 4 or rows >= 256:
                break
            rows *= 2
        out["cpu_oracle"] = {"value": round(W * rows * 16 / spent / 1e6, 4), "unit": "Mpix*spp/s", "cores": threads,
                             "kind": "port", "sample": f"rows {r0}..{r0 + rows - 1} of the {W}x{H} frame at 16 spp "
                                                       f"through the oracle estimator, {spent:.1f} s"}
        log(f"path trace cpu oracle: {out['cpu_oracle']}")
    return out


def bench_scene_file(D, stream, with_cpu, cpu_seconds):
    """The reference's example scene (tests/golden/example_scene_0: cargo.obj, 1670 tris, 9
    reference-shader materials, 6 lights) loaded through the scene-file loader and rendered
    at 1920x1080 with its own settings (1x1 spp, 2 bounces), then converted to RGBA8 bytes."""
    import torch
    path = os.path.join(ROOT, "tests", "golden", "example_scene_0", "0.dumb")
    t0 = time.perf_counter()
    ctx = D.DumbRenderContext.fromFile(path)
    load_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    tree = ctx.octree()
    build_ms = (time.perf_counter() - t0) * 1e3
    W, H = 1920, 1080
    g = ctx.renderer(W, H, tree=tree)
    frame = D.BlockBasedFrameBuffer(W, H)
    ms = _event_ms(lambda: g.render(frame), 10, stream)
    spp = ctx.samples_per_pixel[0] * ctx.samples_per_pixel[1]
    out8 = torch.empty((H, W, 4), dtype=torch.uint8, device=frame.data.device)

    def conv():
        D.lib().dr_frame_to_rgba8(ctypes.c_void_p(frame.data.data_ptr()), W, H, ctypes.c_void_p(out8.data_ptr()),
                                  ctypes.c_void_p(stream.cuda_stream))
    conv_ms = _event_ms(conv, 50, stream)
    conv_bytes = frame.data.numel() * 4 + out8.numel()
    out = {"workload": f"example_scene_0 (cargo.obj, {ctx.triangle_count} tris, reference plugins) {W}x{H}, "
                       f"{spp} spp, {ctx.max_bounces} bounces",
           "load_ms": round(load_ms, 2), "octree_build_ms": round(build_ms, 2), "ms_per_frame": round(ms, 3),
           "mpix_spp_s": round(W * H * spp / ms / 1e3, 2),
           "frame_to_rgba8": {"ms": round(conv_ms, 4), "bytes": conv_bytes,
                              "gb_s": round(conv_bytes / (conv_ms / 1e3) / 1e9, 1)}}
    log(f"example scene 1080p: {ms:.3f} ms/frame, {out['mpix_spp_s']} Mpix*spp/s; rgba8 {conv_ms:.4f} ms")
    if with_cpu:
        import oracle
        from oracle import scene_file as SF
        sf = SF.load(path)
        o = oracle.Renderer(oracle.Octree(sf.tris, 8, 24), sf.mat_id, sf.materials, sf.lights, sf.cam, W, H, sf.fsaa,
                            sf.max_bounces, ignore_backfaces=sf.ignore_backfaces, normals=sf.normals,
                            boxing=sf.boxing)
        threads = HOST_THREADS
        rows = 16
        while True:
            r0 = H // 2 - rows // 2
            t0 = time.perf_counter()
            o.render_ref(nthreads=threads, rows=(r0, r0 + rows))
            spent = time.perf_counter() - t0
            if spent >= cpu_seconds / 8 or rows >= 1024:
                break
            rows *= 2
        out["cpu_oracle"] = {"value": round(W * rows * spp / spent / 1e6, 4), "unit": "Mpix*spp/s", "cores": threads,
                             "kind": "port", "sample": f"rows {r0}..{r0 + rows - 1} of the {W}x{H} frame through the "
                                                       f"oracle's reference estimator, {spent:.2f} s"}
    return out


def bench_rmse(D, S, stream):
    """The 256x256 config-0 scene (depth 4) rendered on the GPU at 4 spp (its own setting) and at
    16 spp, against the FP64 oracle image at 4096 spp committed in tests/golden (made by
    tests/golden/make_fp64_reference.py with independent Philox keys)."""
    import torch
    ref = np.load(os.path.join(ROOT, "tests", "golden", "spheres8_256_fp64_4096spp.npy")).astype(np.float64)
    sc = S.spheres_scene(8, seed=0)
    ds = _device_scene(D, sc.tris)
    out = {"reference": "FP64 oracle estimator, 4096 spp, 256x256, depth 4", "mean_radiance": round(float(ref.mean()), 6)}
    for spp in ((2, 2), (4, 4)):
        g = D.DumbRenderer(ds, sc.mat_id, sc.materials, sc.lights, sc.camera, 256, 256, spp, 4, seed=11)
        f = g.render()
        torch.cuda.synchronize()
        d = f.to_image()[..., :3].cpu().numpy().astype(np.float64) - ref
        n = spp[0] * spp[1]
        out[f"{n}spp"] = {"rmse": float(f"{np.sqrt(np.mean(d ** 2)):.5g}"), "mean_difference": float(f"{d.mean():.3g}"),
                          "mean_difference_se": float(f"{d.std() / np.sqrt(d.size):.3g}")}
    log(f"rmse vs FP64 4096 spp: {out}")
    return out


def bench_small(D, S, stream):
    """256x256, 4 spp (2x2), depth 4, Lambert only: ground quad + 8 UV spheres, one 1x1 area light."""
    sc = S.spheres_scene(8, seed=0)
    g = D.DumbRenderer(_device_scene(D, sc.tris), sc.mat_id, sc.materials, sc.lights, sc.camera, 256, 256, (2, 2), 4)
    frame = D.BlockBasedFrameBuffer(256, 256)
    ms = _event_ms(lambda: g.render(frame), 10, stream)
    return {"workload": "path trace 256x256, 4 spp, depth 4, 2050 tris", "ms_per_frame": round(ms, 4),
            "mpix_spp_s": round(256 * 256 * 4 / ms / 1e3, 2)}


def bench_adjoint(D, S, stream, adam_steps):
    """512x512, 16 spp, depth 4; 64 UV spheres (64 Lambert materials) + ground, 4 area lights.
    Forward render vs dr_backward time, then Adam (lr 0.01) on albedo + emission from a
    random init toward a target rendered from the hidden seeded parameters."""
    import torch
    from dumbray_amd.autograd import render
    sc = S.spheres_scene(64, seed=5, n_lights=4, ground_material=False)
    W, H = 512, 512
    g = D.DumbRenderer(_device_scene(D, sc.tris), sc.mat_id, sc.materials, sc.lights, sc.camera, W, H, (4, 4), 4)
    frame = D.BlockBasedFrameBuffer(W, H)
    fwd = _event_ms(lambda: g.render(frame), 5, stream)
    dl = torch.randn_like(frame.data)
    bwd = _event_ms(lambda: g.backward(dl), 5, stream)
    out = {"workload": f"adjoint {W}x{H}, 16 spp, depth 4, {len(sc.tris)} tris, 64 materials, 4 lights",
           "forward_ms": round(fwd, 3), "backward_ms": round(bwd, 3), "ratio": round(bwd / fwd, 3)}
    log(f"adjoint: forward {fwd:.2f} ms, backward {bwd:.2f} ms, ratio {bwd / fwd:.3f}")
    if adam_steps > 0:
        hidden_a = g.albedo.clone()
        hidden_e = g.emission.clone()
        target = render(hidden_a, hidden_e, g).detach()

        def adam(em_scale):
            """Adam (lr 0.01) on albedo and on emission / em_scale from the same random init."""
            rng = torch.Generator(device="cpu").manual_seed(1234)
            alb0 = torch.rand(hidden_a.shape, generator=rng).mul(0.6).add(0.2).cuda()
            em0 = torch.rand(hidden_e.shape, generator=rng).mul(10.0).add(5.0).cuda()
            alb = alb0.clone().requires_grad_(True)
            em_p = (em0 / em_scale).requires_grad_(True)
            opt = torch.optim.Adam([alb, em_p], lr=0.01)
            losses = []
            t0 = time.perf_counter()
            for step in range(adam_steps):
                g.seed = 0x5EED + 1 + step                 # fresh Philox key per step
                img = render(alb, em_p * em_scale, g)
                loss = ((img[:, :3] - target[:, :3]) ** 2).sum()
                opt.zero_grad()
                loss.backward()
                opt.step()
                with torch.no_grad():
                    alb.clamp_(0.0, 1.0)
                    em_p.clamp_(min=0.0)
                losses.append(float(loss.detach()))
            torch.cuda.synchronize()
            secs = time.perf_counter() - t0
            g.seed = 0x5EED
            rmse = lambda a, b: float(torch.sqrt(((a - b) ** 2).mean()))
            return {"adam_steps": adam_steps, "lr": 0.01, "emission_parameter": f"emission / {em_scale:g}",
                    "seconds": round(secs, 2), "loss_first": round(losses[0], 6), "loss_last": round(losses[-1], 6),
                    "albedo_rmse_first": round(rmse(alb0, hidden_a), 5),
                    "albedo_rmse_last": round(rmse(alb.detach(), hidden_a), 5),
                    "emission_rmse_first": round(rmse(em0, hidden_e), 5),
                    "emission_rmse_last": round(rmse(em_p.detach() * em_scale, hidden_e), 5)}

        # as configured (lr 0.01 on both tables: an emission of 5-15 moves at most 2 in 200 steps), and
        # with the emission parameter scaled to the albedo's range
        out["inverse_rendering"] = adam(1.0)
        log(f"inverse rendering: {out['inverse_rendering']}")
        out["inverse_rendering_emission_scaled"] = adam(10.0)
        log(f"inverse rendering, emission / 10: {out['inverse_rendering_emission_scaled']}")
    return out


def bench_batch(D, S, stream):
    """64 renders in one persistent launch (hipGraph replay) over 8 seeded scenes of 3 displaced
    icospheres (levels 5-7: 61k-983k tris) + ground; resolution from {128^2..1024^2}, spp from
    {1, 4, 16, 64}, depth from {2, 4, 8}; launch shape chosen by the autotuner.  Compared with
    rendering the same requests one dr_render_blocks call at a time."""
    import torch
    specs, qs = S.batch_workload()
    scenes = [(sc, _device_scene(D, sc.tris, *REF_KNOBS)) for sc in specs]
    reqs = []
    total = 0
    for q in qs:
        sc, ds = scenes[q["scene"]]
        r = D.DumbRenderer(ds, sc.mat_id, sc.materials, sc.lights, sc.camera, q["width"], q["height"], q["fsaa"],
                           q["depth"], rr_start=3, rr_survive=0.75, seed=q["seed"])
        reqs.append((r, D.BlockBasedFrameBuffer(q["width"], q["height"]), 1, q["request"]))
        total += q["width"] * q["height"] * q["fsaa"][0] * q["fsaa"][1]
    t0 = time.perf_counter()
    batch = D.RenderBatch(reqs)
    tune_s = time.perf_counter() - t0
    ms = _event_ms(lambda: batch.run(stream), 3, stream)

    def sequential():
        for r, fr, it, rid in reqs:
            r.renderBlocksGPU(fr, 0, fr.getBlockCount(), it, stream)
    ms_seq = _event_ms(sequential, 2, stream)
    out = {"workload": f"64 requests over 8 icosphere scenes ({sum(len(sc.tris) for sc, _ in scenes)} tris), "
                       f"{total / 1e6:.1f} M pixel-samples, octree split 8 / depth 24", "ms_per_batch": round(ms, 2),
           "mpix_spp_s": round(total / ms / 1e3, 2), "sequential_ms": round(ms_seq, 2),
           "speedup_vs_sequential": round(ms_seq / ms, 3), "tuning": batch.tuning, "create_and_tune_s": round(tune_s, 2)}
    log(f"batch: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--rays", type=int, default=1 << 22)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true", help="also run a 256 MiB device copy (pmc calibration)")
    ap.add_argument("--profile", action="store_true",
                    help="only the timed incoherent casts (+ calibration copy): for rocprofv3 runs")
    ap.add_argument("--no-render", action="store_true", help="skip the path-tracing and adjoint legs")
    ap.add_argument("--adam-steps", type=int, default=200)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-requests leg")
    ap.add_argument("--legs", default="path_trace,rmse,small,adjoint,batch,example",
                    help="render legs to run (comma list of path_trace, rmse, small, adjoint, batch, example)")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 counter passes of the cast kernel")
    ap.add_argument("--pmc-dir", default=None, help="keep the counter passes' CSV files in this directory")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one MI355X only; runs on several GPUs (rays, tiles or requests "
              "split across them) are not implemented.", file=sys.stderr)
        sys.exit(2)

    import torch
    import dumbray_amd as D
    from dumbray_amd import scenes as S

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log("generating workload")
    tris, coherent, incoherent, cam = S.raycast_workload(cells=args.cells, n_incoherent=args.rays)
    t0 = time.perf_counter()
    tree = D.Octree(16, 12)
    tree.push(tris)
    tree.build()
    build_s = time.perf_counter() - t0
    sizes = tree.device_layout_sizes()
    scene = tree.upload()
    log(f"tris={len(tris)} nodes={sizes['nodes']} groups={sizes['groups']} records={sizes['records']} "
        f"build={build_s:.2f}s device_bytes={scene.device_bytes()}")
    rB = torch.from_numpy(incoherent).to(dev)
    rA = torch.from_numpy(coherent).to(dev)
    stream = torch.cuda.current_stream(dev)

    def step(r):
        return scene.cast(r, clip_backfaces=False, stream=stream)

    for _ in range(args.warmup):
        step(rB)
        if not args.profile:
            step(rA)
    torch.cuda.synchronize()

    def timed(r, k):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        e0.record(stream)
        for _ in range(k):
            out = step(r)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        return e0.elapsed_time(e1) / k, wall, out

    kB_ms, wallB, outB = timed(rB, args.steps)
    log(f"batch B: {kB_ms:.3f} ms/launch ({len(incoherent) / kB_ms / 1e3:.1f} Mrays/s)")
    if args.profile:
        if args.calib_copy:
            a = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
            b = torch.empty_like(a)
            b.copy_(a)
            torch.cuda.synchronize()
            log("calibration copy: 268435456 bytes read + written")
        print(json.dumps({"profile": True, "kernel_ms": kB_ms}), flush=True)
        return
    st = scene.cast_stats(rB, clip_backfaces=False, stream=stream)
    torch.cuda.synchronize()
    cnt = st["counts"].double().mean(dim=0).cpu().numpy()
    assert torch.equal(st["id"], outB["id"]), "dr_cast_stats disagrees with dr_cast"
    ge, lv, pl, ins, live, slab = (float(x) for x in cnt)
    kernel_work = {"group_expansions": ge, "leaves": lv, "plane_tests": pl, "inside_tests": ins,
                   "live_children": live, "slab_tests": slab,
                   "bytes": K_RAY + K_OUT + K_GROUP * ge + K_POP * (ge - 1 + lv) + K_LEAF * lv + K_PLANE * pl
                   + K_INSIDE * ins,
                   "ops": KOPS_RAY + KOPS_OCTANT * live + KOPS_SLAB * slab + KOPS_PLANE * pl + KOPS_INSIDE * ins}
    log(f"kernel work per ray: {kernel_work}")
    kA_ms, wallA, outA = timed(rA, args.steps)
    log(f"batch A: {kA_ms:.3f} ms/launch ({len(coherent) / kA_ms / 1e3:.1f} Mrays/s)")
    ms_per_step = wallB * 1e3 / args.steps
    value = len(incoherent) / (ms_per_step / 1e3) / 1e6

    # ---------------------------------------------------------------- oracle leg
    per_ray = None
    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        oracle.build()
        log("oracle: building the reference octree on the host")
        ot = oracle.Octree(tris, 16, 12)
        threads = HOST_THREADS
        rng = np.random.default_rng(11)
        idx = np.sort(rng.choice(len(incoherent), 65536, replace=False))
        ores = ot.cast(incoherent[idx], counters=True, nthreads=threads)
        gid = outB["id"].cpu().numpy()[idx]
        parity = int(np.sum(gid != ores["id"]))
        nb, nr, nt = (float(ores[k].mean()) for k in ("n_box", "n_leafref", "n_tri"))
        per_ray = {"n_box": nb, "n_leafref": nr, "n_tri": nt,
                   "bytes": B_RAY + B_OUT + B_BOX * nb + B_REF * nr + B_TRI * nt,
                   "ops": OPS_RAY + OPS_BOX * nb + OPS_TRI * nt}
        log(f"oracle sample parity mismatches={parity}; per ray boxes={nb:.1f} refs={nr:.1f} tris={nt:.1f}")
        # CPU baseline: the oracle's Octree::cast over the whole incoherent batch, repeated
        # until ~cpu_seconds have passed (at least 3 runs), median rate
        rates, spent = [], 0.0
        while len(rates) < 3 or (spent < args.cpu_seconds and len(rates) < 50):
            t0 = time.perf_counter()
            ot.cast(incoherent, nthreads=threads)
            dt = time.perf_counter() - t0
            spent += dt
            rates.append(len(incoherent) / dt / 1e6)
        cpu = {"value": round(float(np.median(rates)), 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"all {len(incoherent)} incoherent rays through the oracle's Octree::cast on {threads} "
                         f"threads, median of {len(rates)} runs ({spent:.1f} s); octree build excluded"}
        log(f"cpu baseline: {cpu['value']} Mrays/s on {threads} threads")

    # ---------------------------------------------------------------- roofline of the cast kernel
    # Work the kernel did (dr_cast_stats on the same rays) against the VALU peak, the resource
    # that binds it; its fabric bytes, L2 hit rate, VALU busy and lane utilization come from the
    # counter passes this run launches (or from the committed profile when they cannot run).
    n = len(incoherent)
    t = kB_ms / 1e3
    ops = kernel_work["ops"] * n
    logical = kernel_work["bytes"] * n
    pmc = None if args.no_pmc else measure_counters(keep_dir=args.pmc_dir)
    if pmc is not None:
        traffic, src = pmc["fetch_bytes"] + pmc["write_bytes"], "rocprofv3 --pmc passes launched by this bench run"
    else:
        traffic = read_traffic(os.path.join(ROOT, "profiles", "cast_incoherent_hf2M_counters.txt"))
        src = "profiles/cast_incoherent_hf2M_counters.txt (committed; the counter passes did not run)"
    roofline = {"bound": "valu", "achieved": round(ops / t / 1e12, 3), "peak": round(VALU_PEAK_TOPS, 2),
                "unit": "T lane-ops/s", "frac": round(ops / t / 1e12 / VALU_PEAK_TOPS, 4),
                "traffic": None if traffic is None else int(traffic),
                "kernel": "dr::k_cast<false>", "kernel_ms": round(kB_ms, 4), "ops_per_launch": int(ops),
                "binding": "instruction issue, then load latency: the useful lane-ops above are the kernel's box, "
                           "plane and inside tests; the rest of the issue slots go to idle lanes and control, and "
                           "the time the VALU is idle (1 - valu_busy) to waits on dependent loads",
                "counters": None if pmc is None else {k: (round(v, 4) if isinstance(v, float) else v)
                                                      for k, v in pmc.items()},
                "counters_source": src,
                "hbm": None if traffic is None else {
                    "achieved": round(traffic / t / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(traffic / t / 1e9 / HBM_PEAK_GBS, 4),
                    "note": "FETCH_SIZE + WRITE_SIZE scaled by a copy of known size; they count every L2 miss, "
                            "including those the last-level (Infinity) cache serves, so this bounds HBM from above"},
                "logical_bytes": {"per_ray": round(kernel_work["bytes"], 1), "achieved": round(logical / t / 1e9, 1),
                                  "unit": "GB/s", "note": "bytes the kernel's loads name; most are L2 hits"}}
    reference_equivalent = None
    if per_ray is not None:
        reference_equivalent = {"bytes_per_ray": round(per_ray["bytes"], 1), "ops_per_ray": round(per_ray["ops"], 1),
                                "bytes_rate": round(per_ray["bytes"] * n / t / 1e9, 1), "bytes_unit": "GB/s",
                                "ops_rate": round(per_ray["ops"] * n / t / 1e12, 3), "ops_unit": "T lane-ops/s",
                                "note": "the reference algorithm's work (boxes, leaf refs and triangle tests it "
                                        "runs on these rays) at this kernel's ray rate; the kernel prunes most of "
                                        "it, so these rates can exceed the hardware peaks"}

    # clipBackfaces = true (the reference's default and the renderer's setting), same batch
    kC_ms = _event_ms(lambda: scene.cast(rB, clip_backfaces=True, stream=stream), args.steps, stream)
    clip_true = {"kernel_ms": round(kC_ms, 4), "mrays_s": round(n / kC_ms / 1e3, 2), "kernel": "dr::k_cast<true>"}
    log(f"batch B, clipBackfaces=true: {kC_ms:.3f} ms/launch ({clip_true['mrays_s']} Mrays/s)")

    render = None
    if not args.no_render:
        legs = set(args.legs.split(","))
        if args.no_batch:
            legs.discard("batch")
        run = lambda leg, fn: fn() if leg in legs else None
        render = {"path_trace_1080p": run("path_trace", lambda: bench_path_trace(D, S, stream, args.cpu_seconds,
                                                                                 not args.no_cpu)),
                  "rmse_vs_fp64_4096spp": run("rmse", lambda: bench_rmse(D, S, stream)),
                  "path_trace_256": run("small", lambda: bench_small(D, S, stream)),
                  "adjoint_512": run("adjoint", lambda: bench_adjoint(D, S, stream, args.adam_steps)),
                  "batch_64_requests": run("batch", lambda: bench_batch(D, S, stream)),
                  "scene_file_example_0": run("example", lambda: bench_scene_file(D, stream, not args.no_cpu,
                                                                                  args.cpu_seconds))}

    pt = render["path_trace_1080p"] if render else None
    adj = render["adjoint_512"] if render else None
    line = {
        "metric": "closest-hit Mrays/s & Mpix·spp/s at 1080p; adjoint/forward time ratio",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"closest-hit cast, {len(incoherent)} incoherent rays (batch B) on a "
                               f"{args.cells}x{args.cells} fBm heightfield ({len(tris)} tris)",
                   "octree": "split 16 / depth 12", "clip_backfaces": False, "parallelism": "single GPU"},
        "coherent_1080p_mrays_s": round(len(coherent) / kA_ms / 1e3, 2),
        "incoherent_kernel_mrays_s": round(len(incoherent) / kB_ms / 1e3, 2),
        "octree_build_s": round(build_s, 3),
        "roofline": roofline,
        "clip_backfaces_true": clip_true,
        "kernel_work_per_ray": kernel_work,
        "reference_work_per_ray": per_ray,
        "reference_equivalent": reference_equivalent,
        "host": {"cores_used": HOST_THREADS, "nproc": os.cpu_count()},
        "oracle_sample_mismatches": parity,
        "cpu_baseline": cpu,
        "mpix_spp_s_1080p": pt["mpix_spp_s"] if pt else None,
        "adjoint_forward_ratio": adj["ratio"] if adj else None,
        "render": render,
    }
    if cpu:
        line["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

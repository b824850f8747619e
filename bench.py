#!/usr/bin/env python3
"""Benchmark of the MI355X TSDF fusion path on one GPU.

Workload (one "step"): the 512^3 float32 TSDF + weight volume (4 mm voxels,
truncation 5 voxels, weight cap 128) is reset and 300 ray-cast 640x480 depth
frames of a 40-box scene are integrated one frame per call (the drop-in
TSDFVolume::integrate path), with a marching-cubes extraction after every
50 frames.  Depth frames are resident in HBM before timing starts.

Printed value: swept voxel-updates per second over the whole step
(512^3 * 300 / step time), in GVoxel-updates/s.  The roofline object
covers one integrate call (its three launches: depth_bounds_kernel,
cull_kernel, integrate_list_kernel): algorithmic bytes per call
(16 B per updated voxel + 4 B per depth pixel, the update count measured
exactly by the kernel's counting mode outside the timed region) divided by
the call's average duration from HIP events on the launch stream (one event
pair around each run of consecutive calls between extractions, in the timed
steps; a record per call would add its own ~10 us per call); traffic
is the HBM bytes per call from the committed rocprofv3 summary
(profiles/integrate_plain_large_512.md).

Side legs in the same JSON line: meshing (soup vs welded), colour, sparse
bricks, the small 128^3 x 10-frame workload in full on the GPU and on the
CPU (1 thread and all host cores), point-cloud render, batched volumes, the
differentiable forward + adjoint, batched pose refinement.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cu-reconstruction_amd"))

METRIC = "GVoxel-updates/s + % HBM roofline; batched vs looped; fwd+bwd/fwd time ratio"
HBM_PEAK_GBS = 8000.0
PROFILE = ROOT / "profiles" / "integrate_plain_large_512.md"
TUNE_CACHE = ROOT / "cu-reconstruction_amd" / "tuning" / "batch_mi355x.json"


def profiled_traffic():
    """HBM bytes per integrate call measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (copy-kernel scaled) and committed with the commands that
    produced them; None when no such summary is present."""
    import re
    if not PROFILE.exists():
        return None
    m = re.search(r"^traffic_bytes_per_call: (\d+)", PROFILE.read_text(), re.M)
    return int(m.group(1)) if m else None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str | None:
    """The host CPU's model name (/proc/cpuinfo), None where it is not readable."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(w, depth, seconds: float) -> dict:
    """Times the CPU restatement (oracle) of the per-frame integrate on the
    host cores over the same workload's frames in order, all of them unless
    `seconds` have elapsed first."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_ffi as O

    threads = min(os.cpu_count() or 1, 16)
    vol = O.volume_from_config(w.config())
    D, W = O.init(vol)
    t0 = time.perf_counter()
    frames = 0
    while frames < w.num_frames:
        O.integrate_f32(vol, D, W, depth[frames], w.poses[frames], nthreads=threads)
        frames += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": round(w.num_voxels * frames / dt / 1e9, 4), "unit": "GVoxel-updates/s",
            "cores": threads, "host_cpus": os.cpu_count(), "cpu_model": cpu_model(), "kind": "port",
            "sample": f"oracle float32 integrate of the first {frames} frames of the "
                      f"{w.name} workload ({dt:.1f} s, OpenMP over voxel rows)"}


def bench_small(steps: int, cpu: bool = True) -> dict:
    """The small CPU-runnable workload (128^3 volume, 8 mm voxels, 10 frames
    of 640x480) in full: the GPU per-frame integrate and the fused ten-frame
    integrate (HIP events on the launch stream, `steps` passes each) beside
    the oracle's float32 integrate of all 10 frames on one host thread and on
    up to 16 threads (best of five passes each)."""
    import torch

    import recon
    from recon import scenes

    w = scenes.plain_small()
    depth = w.render(nthreads=min(os.cpu_count() or 1, 16))
    dev = torch.from_numpy(depth).cuda()
    vol = recon.TSDFVolume(w.config())
    stream = torch.cuda.current_stream()
    F = w.num_frames

    def one_pass():
        vol.reset(stream)
        for f in range(F):
            vol.integrate(dev[f], w.poses[f], stream)

    one_pass()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        one_pass()
    e1.record(stream)
    torch.cuda.synchronize()
    pass_ms = e0.elapsed_time(e1) / steps

    def fused_pass():
        vol.reset(stream)
        vol.integrate_frames(dev, w.poses, stream)

    fused_pass()
    e0.record(stream)
    for _ in range(steps):
        fused_pass()
    e1.record(stream)
    torch.cuda.synchronize()
    fused_ms = e0.elapsed_time(e1) / steps
    vol.close()
    out = {"workload": w.name, "grid": list(w.grid), "frames": F, "image": list(w.image),
           "gpu_ms_per_pass": round(pass_ms, 4),
           "gpu_gvu_s": round(w.num_voxels * F / (pass_ms / 1000.0) / 1e9, 3),
           "gpu_fused_ms_per_pass": round(fused_ms, 4),
           "gpu_fused_gvu_s": round(w.num_voxels * F / (fused_ms / 1000.0) / 1e9, 3)}
    del dev
    if cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle_ffi as O

        ov = O.volume_from_config(w.config())
        for threads in (1, min(os.cpu_count() or 1, 16)):
            best = float("inf")
            for _ in range(5):  # a pass is 0.01-0.2 s: best of five against host noise
                Dc, Wc = O.init(ov)
                t0 = time.perf_counter()
                for f in range(F):
                    O.integrate_f32(ov, Dc, Wc, depth[f], w.poses[f], nthreads=threads)
                best = min(best, time.perf_counter() - t0)
            out[f"cpu_{threads}t_ms_per_pass"] = round(1000.0 * best, 1)
            out[f"cpu_{threads}t_gvu_s"] = round(w.num_voxels * F / best / 1e9, 4)
    return out


def bench_batched(steps: int, autotune_to: str | None = None) -> dict:
    """64 uneven volumes: one graph launch (recon_batch_run) against the loop
    of the single-volume path over volumes and frames (reset, one integrate
    call per frame, mesh extraction), with identical results."""
    import numpy as np
    import torch

    import recon
    from recon import scenes

    ws = scenes.batched_uneven()
    t0 = time.perf_counter()
    depths = [torch.from_numpy(w.render(nthreads=min(os.cpu_count() or 1, 16))).cuda() for w in ws]
    torch.cuda.synchronize()
    log(f"batched: rendered {sum(w.num_frames for w in ws)} frames for {len(ws)} volumes in "
        f"{time.perf_counter() - t0:.1f} s")
    spec = [(w.config(), d, w.poses) for w, d in zip(ws, depths)]
    if autotune_to:
        Path(autotune_to).parent.mkdir(parents=True, exist_ok=True)
        recon.Batch.autotune(spec, autotune_to, iterations=3)
        log(f"batched: autotuned into {autotune_to}")
    vols = [recon.TSDFVolume(w.config()) for w in ws]
    stream = torch.cuda.current_stream()

    def looped():
        counts = []
        for v, w, d in zip(vols, ws, depths):
            v.reset(stream)
            for f in range(w.num_frames):
                v.integrate(d[f], w.poses[f], stream)
            verts, _ = v.extract_mesh(stream=stream)
            counts.append(verts.shape[0])
        return counts

    counts = looped()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        looped()
    torch.cuda.synchronize()
    looped_ms = 1000.0 * (time.perf_counter() - t0) / steps
    cache = str(TUNE_CACHE) if TUNE_CACHE.exists() else None
    batch = recon.Batch(spec, mesh_capacity=int(sum(counts) * 1.05) + 1024, tune_cache=cache)
    batch.run(stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run(stream)
    torch.cuda.synchronize()
    batched_ms = 1000.0 * (time.perf_counter() - t0) / steps
    # identical results: every mesh size, and D / W of a sample of volumes
    ranges = batch.mesh_ranges()
    same = [int(r) for r in ranges[:, 1]] == counts
    for i in range(0, len(ws), 16):
        D, W = batch.download(i)
        D0, W0 = vols[i].download()
        same = same and np.array_equal(D.view(np.uint32), D0.view(np.uint32)) and \
            np.array_equal(W.view(np.uint32), W0.view(np.uint32))
    upd = sum(w.num_voxels * w.num_frames for w in ws)
    batch.close()
    for v in vols:
        v.close()
    return {"volumes": len(ws), "voxel_frames": upd, "batched_ms": round(batched_ms, 3),
            "looped_ms": round(looped_ms, 3), "batched_vs_looped": round(looped_ms / batched_ms, 2),
            "batched_gvu_s_swept": round(upd / (batched_ms / 1000.0) / 1e9, 2),
            "identical_results": bool(same), "tuned": cache is not None}


def bench_differentiable(steps: int) -> dict:
    """The differentiable workload: 256^3 volume (6 mm voxels), 20 frames of
    640x480 of the sphere-and-plane scene, loss at 100k surface points.
    forward = reset + recon_tsdf_diff_integrate (20 frames, bilinear) +
    recon_tsdf_surface_loss; backward = recon_tsdf_diff_adjoint (gradients
    for the 20 poses and the 20 depth maps).  HIP events on the launch
    stream bracket each part."""
    import numpy as np
    import torch

    import recon
    from recon import scenes

    w = scenes.differentiable()
    depth = torch.from_numpy(w.render(nthreads=min(os.cpu_count() or 1, 16))).cuda()
    rt = torch.from_numpy(recon.pose_to_matrix(w.poses)).cuda()
    pts = torch.from_numpy(scenes.surface_points(w, 100_000).astype(np.float32)).cuda()
    vol = recon.TSDFVolume(w.config())
    stream = torch.cuda.current_stream()

    def forward():
        vol.reset(stream)
        vol.diff_integrate(depth, rt, stream)
        return vol.surface_loss(pts, stream=stream)

    loss, _, ids, lam = forward()
    gp, gd = vol.diff_adjoint(depth, rt, ids, lam, stream=stream)
    torch.cuda.synchronize()
    # one event pair around a loop of forward passes, one around a loop of
    # forward + backward passes: the backward time is their difference
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(stream)
    for s in range(steps):
        loss, _, ids, lam = forward()
    ev[1].record(stream)
    ev[2].record(stream)
    for s in range(steps):
        loss, _, ids, lam = forward()
        gp, gd = vol.diff_adjoint(depth, rt, ids, lam, stream=stream)
    ev[3].record(stream)
    torch.cuda.synchronize()
    fwd = np.array([ev[0].elapsed_time(ev[1]) / steps])
    bwd = np.array([ev[2].elapsed_time(ev[3]) / steps]) - fwd
    out = {"workload": w.name, "grid": list(w.grid), "frames": w.num_frames,
           "points": int(pts.shape[0]), "fwd_ms": round(float(fwd.mean()), 4),
           "bwd_ms": round(float(bwd.mean()), 4),
           "fwd_bwd_over_fwd": round(float((fwd + bwd).mean() / fwd.mean()), 3),
           "fwd_gvu_s_swept": round(w.num_voxels * w.num_frames / (fwd.mean() / 1000.0) / 1e9, 2),
           "loss": float(loss.item()), "grad_pose_norm": float(gp.norm().item())}
    vol.close()
    return out


def bench_render(w, depth, steps: int, frames: int = 30, cpu: bool = True) -> dict:
    """Point-cloud z-buffer render (recon_points_render, the device version of
    cudaPointCloud::project_to_color_depth): frames 0..frames-1 of the
    workload back-projected on the device into one cloud (valid mask and
    colours kept), rendered at the pose of the next frame into a depth map
    and an rgb image.  HIP events on the launch stream bracket each call
    (clear + splat + resolve).  The CPU leg runs the oracle's restatement of
    the reference's host loop on the same cloud, one thread, as the
    reference does."""
    import numpy as np
    import torch

    import recon
    from recon import scenes

    def inverse(pose):
        R, t = scenes.pose_matrix(pose)
        return np.r_[pose[0], -np.asarray(pose[1:4]), -(R.T @ t)]

    cols, rows = w.image
    n1 = rows * cols
    xyz = torch.empty((frames * n1, 3), dtype=torch.float32, device="cuda")
    valid = torch.empty((frames * n1,), dtype=torch.uint8, device="cuda")
    colors = torch.empty((frames * n1, 3), dtype=torch.uint8, device="cuda")
    rng = np.random.default_rng(7)
    for f in range(frames):
        rgb = torch.from_numpy(rng.integers(0, 256, (rows, cols, 3), dtype=np.uint8)).cuda()
        p, v, c = recon.depth_to_points(torch.from_numpy(depth[f]).cuda(), w.intrinsics,
                                        inverse(w.poses[f]), rgb=rgb, with_colors=True)
        xyz[f * n1:(f + 1) * n1] = p
        valid[f * n1:(f + 1) * n1] = v
        colors[f * n1:(f + 1) * n1] = c
    pose = w.poses[frames]
    stream = torch.cuda.current_stream()
    dout = torch.empty((rows, cols), dtype=torch.float32, device="cuda")
    cout = torch.empty((rows, cols, 3), dtype=torch.uint8, device="cuda")
    zb = torch.empty((n1,), dtype=torch.int64, device="cuda")

    def call():
        recon.render_points(xyz, pose, w.intrinsics, w.image, valid=valid, colors=colors,
                            depth_out=dout, color_out=cout, zbuffer=zb, stream=stream)

    call()
    torch.cuda.synchronize()
    reps = max(steps, 1) * 10
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for r in range(reps):
        call()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = np.array([e0.elapsed_time(e1) / reps])
    n = xyz.shape[0]
    hits = int((dout > 0).sum().item())
    # algorithmic bytes per call: 12 B xyz + 1 B valid per point; per pixel
    # 8 B z-buffer clear, 8 B z-buffer read, 4 B depth, 3 B colour written;
    # 3 B colour gathered per covered pixel
    alg = 13 * n + 23 * n1 + 3 * hits
    out = {"workload": f"{w.name} frames 0-{frames - 1} as points, rendered at frame {frames}",
           "points": int(n), "valid_points": int(valid.sum().item()), "image": [cols, rows],
           "covered_pixels": hits, "ms_per_render": round(float(ms.mean()), 4),
           "gpoints_s": round(n / (ms.mean() / 1000.0) / 1e9, 2),
           "algorithmic_bytes": int(alg),
           "achieved_gb_s": round(alg / (ms.mean() / 1000.0) / 1e9, 1)}
    if cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle_ffi as O
        hx, hv, hc = xyz.cpu().numpy(), valid.cpu().numpy(), colors.cpu().numpy()
        t0 = time.perf_counter()
        od, oc = O.render_points(hx, pose, w.intrinsics, w.image, valid=hv, colors=hc)
        dt = time.perf_counter() - t0
        same = bool(np.array_equal(od.view(np.uint32), dout.cpu().numpy().view(np.uint32)) and
                    np.array_equal(oc, cout.cpu().numpy()))
        out["cpu_baseline"] = {"value": round(n / dt / 1e9, 4), "unit": "Gpoints/s", "cores": 1,
                               "kind": "port", "ms": round(dt * 1000.0, 1),
                               "sample": f"oracle host z-buffer loop over the same {n} points"}
        out["identical_to_oracle"] = same
    return out


def bench_pose_refinement(steps: int = 20, fragments: int = 32, grid: int = 128,
                          frames: int = 10, points: int = 20000, frame_weight: float = 100.0,
                          frame_trunc: float = 0.010, trunc_ratio: float = 0.3,
                          gn_step: float = 0.7, damping: float = 1e-3) -> dict:
    """Batched pose refinement: `fragments` fragments (128^3, 8 mm, 10 frames
    of 640x480 of a distinct seeded scene of boxes and spheres on a floor),
    initial poses perturbed by up to 1 degree and 1 cm, `steps` on-device
    damped Gauss-Newton steps (one graph launch per step: forward, fused-TSDF
    surface loss plus the model-to-frame term with weight `frame_weight` and
    its point-to-plane normal matrices, adjoint, a 6x6 solve per frame) over
    all fragments; the model-to-frame truncation shrinks from `frame_trunc`
    by `trunc_ratio` over the first half of the steps.  Reports the time per
    step, per-frame pose errors, the loss (at `frame_trunc`) and, as the
    quality measure, the RMS distance to the analytic surface of the plain
    integrate's mesh with initial, refined and ground-truth poses."""
    import numpy as np
    import torch

    import recon
    from recon import scenes

    ws = scenes.pose_refinement(num_fragments=fragments, frames=frames, grid=grid,
                                voxel_size=1.024 / grid)
    spec, depths_np = [], []
    npairs = 0
    t0 = time.perf_counter()
    for i, w in enumerate(ws):
        depth = w.render(nthreads=min(os.cpu_count() or 1, 16))
        pts, wts, nrm = scenes.refinement_points(w, depth, points, seed=1000 + i, with_normals=True)
        pairs = scenes.frame_pairs(w, depth, pts, nrm)
        npairs += len(pairs)
        depths_np.append(depth)
        spec.append((w.config(), torch.from_numpy(depth).cuda(),
                     torch.from_numpy(pts.astype(np.float32)).cuda(),
                     torch.from_numpy(wts).cuda(), torch.from_numpy(pairs).cuda(), frame_weight,
                     frame_trunc, torch.from_numpy(nrm.astype(np.float32)).cuda()))
    log(f"pose refinement: {fragments} fragments prepared in {time.perf_counter() - t0:.1f} s")
    init = np.concatenate([recon.pose_to_matrix(w.meta["init"]) for w in ws])
    gt = np.concatenate([recon.pose_to_matrix(w.poses) for w in ws])
    b = recon.DiffBatch(spec)
    loss_gt = b.forward(torch.from_numpy(gt).cuda()).sum().item()
    loss_init = b.forward(torch.from_numpy(init).cuda()).sum().item()
    rt = torch.from_numpy(init.copy()).cuda()
    b.refine_gn(rt, 1)  # warm-up: capture and first replay
    rt = torch.from_numpy(init.copy()).cuda()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    b.refine_gn(rt, steps, step=gn_step, damping=damping, trunc_ratio=trunc_ratio)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    refined = rt.cpu().numpy()
    loss_final = b.forward(rt).sum().item()
    b.close()
    a0, c0 = scenes.pose_errors(init, gt)
    a1, c1 = scenes.pose_errors(refined, gt)

    def mesh_error(rts):
        """Mean over fragments of the RMS distance to the analytic surface of
        the plain integrate's mesh vertices in observed space."""
        errs = []
        f0 = 0
        for w, depth in zip(ws, depths_np):
            F = w.num_frames
            v = recon.TSDFVolume(w.config())
            dd = torch.from_numpy(depth).cuda()
            for f in range(F):
                R = rts[f0 + f, :9].reshape(3, 3).astype(np.float64)
                pose = np.r_[scenes.quat_from_matrix(R), rts[f0 + f, 9:]]
                v.integrate(dd[f], pose)
            verts, _ = v.extract_mesh()
            _, W = v.download()
            vv = verts.cpu().numpy().reshape(-1, 3)
            errs.append(scenes.mesh_rms(w, vv[scenes.observed_vertices(w, W, vv)]))
            v.close()
            f0 += F
        return float(np.mean(errs))

    def common_mode(rts):
        """Mean over fragments of |mean over its frames of the world-frame
        error twist|: (translation mm, rotation deg)."""
        tw = scenes.world_error_twists(rts, gt)
        f0, tr, rot = 0, [], []
        for w in ws:
            m = tw[f0:f0 + w.num_frames].mean(axis=0)
            tr.append(np.linalg.norm(m[:3]) * 1e3)
            rot.append(np.degrees(np.linalg.norm(m[3:])))
            f0 += w.num_frames
        return round(float(np.mean(tr)), 3), round(float(np.mean(rot)), 4)

    return {"fragments": fragments, "grid": grid, "frames": frames, "points_per_fragment": points,
            "frame_pairs": int(npairs), "frame_weight": frame_weight,
            "optimizer": {"method": "damped Gauss-Newton, point-to-plane normal matrix",
                          "step": gn_step, "damping": damping, "frame_trunc_m": frame_trunc,
                          "trunc_ratio": round(trunc_ratio, 6)},
            "steps": steps, "ms_per_step": round(1000.0 * dt / max(steps, 1), 3),
            "loss_initial": float(loss_init), "loss_final": float(loss_final),
            "loss_ground_truth": float(loss_gt),
            "rotation_error_deg": {"initial_mean": round(float(a0.mean()), 4),
                                   "final_mean": round(float(a1.mean()), 4),
                                   "initial_max": round(float(a0.max()), 4),
                                   "final_p99": round(float(np.percentile(a1, 99)), 4),
                                   "final_max": round(float(a1.max()), 4),
                                   "frames_above_initial_max": int((a1 > a0.max()).sum())},
            "centre_error_mm": {"initial_mean": round(float(c0.mean()) * 1e3, 3),
                                "final_mean": round(float(c1.mean()) * 1e3, 3),
                                "initial_max": round(float(c0.max()) * 1e3, 3),
                                "final_p99": round(float(np.percentile(c1, 99)) * 1e3, 3),
                                "final_max": round(float(c1.max()) * 1e3, 3),
                                "frames_above_initial_max": int((c1 > c0.max()).sum())},
            "common_mode_error": {"initial_mm_deg": common_mode(init),
                                  "final_mm_deg": common_mode(refined)},
            "mesh_rms_mm": {"initial": round(mesh_error(init) * 1e3, 4),
                            "refined": round(mesh_error(refined) * 1e3, 4),
                            "ground_truth": round(mesh_error(gt) * 1e3, 4)}}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="cap on the CPU baseline's sample (the 300-frame workload takes ~18 s on 16 threads)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-volume measurement")
    ap.add_argument("--autotune", metavar="JSON",
                    help="run the batch autotuner on the batched workload and write its table")
    ap.add_argument("--no-diff", action="store_true",
                    help="skip the differentiable (forward + adjoint) measurement")
    ap.add_argument("--no-refine", action="store_true",
                    help="skip the batched pose refinement measurement")
    ap.add_argument("--no-render", action="store_true",
                    help="skip the point-cloud z-buffer render measurement")
    ap.add_argument("--only", choices=["integrate", "copy", "differentiable", "refine", "render"],
                    help="profiling aids: 'integrate' runs --steps passes of the per-frame "
                         "integrate only; 'copy' copies a 256 MiB buffer 20 times (the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE); "
                         "'differentiable' runs the forward + adjoint leg only; 'refine' the "
                         "batched pose refinement only; 'render' the point-cloud render only")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the TSDF path here is single-GPU "
              "(--gpus 1); multi-GPU sharding is a design note, not a run.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import recon
    from recon import scenes

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 1
    torch.cuda.set_device(0)
    if args.only == "copy":
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(20):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 20 x 256 MiB")
        return 0
    if args.only == "differentiable":
        log(json.dumps(bench_differentiable(args.steps)))
        return 0
    if args.only == "refine":
        log(json.dumps(bench_pose_refinement()))
        return 0
    w = scenes.plain_large(num_frames=args.frames)
    t0 = time.perf_counter()
    depth = w.render(nthreads=min(os.cpu_count() or 1, 16))
    log(f"rendered {w.num_frames} frames {w.image} in {time.perf_counter() - t0:.1f} s")
    dev = torch.from_numpy(depth).cuda()
    torch.cuda.synchronize()
    vol = recon.TSDFVolume(w.config())
    stream = torch.cuda.current_stream()
    F = w.num_frames
    if args.only == "render":
        log(json.dumps(bench_render(w, depth, args.steps, cpu=False)))
        return 0
    if args.only == "integrate":
        for s in range(args.steps):
            vol.reset(stream)
            for f in range(F):
                vol.integrate(dev[f], w.poses[f], stream)
        torch.cuda.synchronize()
        log(f"integrated {args.steps} x {F} frames")
        return 0
    mesh_at = [f for f in range(F) if w.mesh_every and (f + 1) % w.mesh_every == 0]

    # exact update counts U_f (counting mode, outside the timed region)
    counters = torch.zeros(F, dtype=torch.int64, device="cuda")
    vol.reset()
    for f in range(F):
        vol.set_update_counter(counters[f:f + 1])
        vol.integrate(dev[f], w.poses[f])
    vol.set_update_counter(None)
    torch.cuda.synchronize()
    U = counters.cpu().numpy().astype(np.int64)
    log(f"updates per frame: mean {U.mean():.3e} min {U.min()} max {U.max()}")

    # mesh buffer capacity from the counts at the extraction points
    vol.reset()
    cap = 1
    for f in range(F):
        vol.integrate(dev[f], w.poses[f])
        if f in mesh_at:
            cap = max(cap, vol.count_triangles())
    cap = int(cap * 1.05) + 1024
    verts = torch.empty((cap, 3, 3), dtype=torch.float32, device="cuda")
    tris = torch.empty((cap, 3), dtype=torch.int32, device="cuda")
    log(f"mesh capacity {cap} triangles")

    # runs of consecutive integrate calls between extractions (an event
    # record costs ~10 us of GPU time, so the timed steps bracket runs, not
    # single calls)
    runs = []
    a = 0
    for f in range(F):
        if f in mesh_at or f == F - 1:
            runs.append((a, f + 1))
            a = f + 1

    def step(events=None):
        vol.reset(stream)
        ntri = 0
        for k, (a, b) in enumerate(runs):
            if events is not None:
                events[k][0].record(stream)
            for f in range(a, b):
                vol.integrate(dev[f], w.poses[f], stream)
            if events is not None:
                events[k][1].record(stream)
            if b - 1 in mesh_at:
                v, _ = vol.extract_mesh(verts, tris, stream)
                ntri = v.shape[0]
        return ntri

    for i in range(args.warmup):
        step()
        log(f"warmup step {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()

    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in runs] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    ntri = 0
    for s in range(args.steps):
        ntri = step(ev[s])
        log(f"timed step {s + 1}/{args.steps} enqueued")
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    ms_per_step = 1000.0 * t_total / args.steps
    run_ms = np.array([[e0.elapsed_time(e1) for (e0, e1) in ev[s]] for s in range(args.steps)])
    avg_launch_ms = float(run_ms.sum()) / (args.steps * F)  # ms per integrate call

    # fused-frame path: same frames, integrate_frames between extractions
    seg = [0] + [f + 1 for f in mesh_at]
    if seg[-1] != F:
        seg.append(F)

    def step_fused(events=None):
        vol.reset(stream)
        for k, (a, b) in enumerate(zip(seg[:-1], seg[1:])):
            if events is not None:
                events[k][0].record(stream)
            vol.integrate_frames(dev[a:b], w.poses[a:b], stream)
            if events is not None:
                events[k][1].record(stream)
            if b - 1 in mesh_at:
                vol.extract_mesh(verts, tris, stream)

    # voxels each fused sweep (16 frames) reads and writes: the touch counter
    # outside the timed region
    touched = torch.zeros(1, dtype=torch.int64, device="cuda")
    vol.set_touch_counter(touched)
    step_fused()
    vol.set_touch_counter(None)
    torch.cuda.synchronize()
    touched_per_step = int(touched.item())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step_fused()
    torch.cuda.synchronize()
    fused_ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    fev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            for _ in seg[:-1]] for _ in range(args.steps)]
    for s_ in range(args.steps):
        step_fused(fev[s_])
    torch.cuda.synchronize()
    fused_integrate_ms = float(sum(e0.elapsed_time(e1) for st in fev for (e0, e1) in st)) / args.steps

    swept = w.num_voxels * F
    value = swept / (ms_per_step / 1000.0) / 1e9
    pix = w.image[0] * w.image[1]
    alg_bytes = 16.0 * U + 4.0 * pix           # per launch (frame)
    achieved = float(alg_bytes.mean()) / (avg_launch_ms / 1000.0) / 1e9
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GVoxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "grid": list(w.grid), "voxel_m": w.voxel_size,
                   "frames": F, "image": list(w.image), "mesh_every": w.mesh_every,
                   "integrate": "one launch per frame (TSDFVolume::integrate)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(),
                     "traffic_source": str(PROFILE.relative_to(ROOT)) if PROFILE.exists() else None,
                     "algorithmic_bytes": int(alg_bytes.mean())},
        "integrate_ms_per_launch": round(avg_launch_ms, 4),
        "integrate_gvu_s_swept": round(w.num_voxels / (avg_launch_ms / 1000.0) / 1e9, 2),
        "integrate_gvu_s_touched": round(float(U.mean()) / (avg_launch_ms / 1000.0) / 1e9, 2),
        "updates_per_frame_mean": int(U.mean()),
        "triangles_last_mesh": int(ntri),
        "fused_frames_ms_per_step": round(fused_ms, 3),
        "fused_frames_gvu_s": round(swept / (fused_ms / 1000.0) / 1e9, 3),
    }
    # fused path (integrate_frames, 16 frames per sweep): algorithmic bytes
    # 16 B per voxel a sweep updates at least once + 4 B per depth pixel per
    # frame, over the integrate calls' own time (HIP events)
    fused_bytes = 16.0 * touched_per_step + 4.0 * pix * F
    fused_achieved = fused_bytes / (fused_integrate_ms / 1000.0) / 1e9
    result["roofline_fused"] = {
        "bound": "hbm", "achieved": round(fused_achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(fused_achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "algorithmic_bytes_per_step": int(fused_bytes),
        "touched_voxels_per_step": touched_per_step,
        "integrate_ms_per_step": round(fused_integrate_ms, 3),
        "sweeps_per_step": sum((b - a + 15) // 16 for a, b in zip(seg[:-1], seg[1:]))}
    # meshing of the final volume: the triangle soup (count + emit, the
    # drop-in generate_triangular_surface path) against the welded mesh with
    # normals (count + vertices/normals + indexed triangles); both are
    # synchronous calls, timed on the host clock
    def timed(fn, reps=5):
        out = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return 1000.0 * (time.perf_counter() - t0) / reps, out

    log("timing the meshing of the final volume")
    soup_ms, _ = timed(lambda: vol.extract_mesh(verts, tris, stream))
    nv, nt = vol.count_welded(stream)
    wbuf = (torch.empty((nv, 3), dtype=torch.float32, device="cuda"),
            torch.empty((nv, 3), dtype=torch.float32, device="cuda"),
            torch.empty((nt, 3), dtype=torch.int32, device="cuda"))
    weld_ms, (wv, _, wt) = timed(lambda: vol.extract_welded_mesh(out=wbuf, stream=stream))
    result["meshing"] = {
        "grid": list(w.grid), "triangles": int(wt.shape[0]),
        "soup_ms": round(soup_ms, 3), "soup_vertices": 3 * int(wt.shape[0]),
        "welded_ms": round(weld_ms, 3), "welded_vertices": int(wv.shape[0]),
        "welded_bytes_vs_soup": round((24.0 * wv.shape[0] + 12.0 * wt.shape[0]) /
                                      (36.0 * wt.shape[0] + 12.0 * wt.shape[0]), 3)}
    del wv, wt, wbuf
    # colour integrate: the first 30 frames, per-frame calls, with random
    # colour images, against the same calls without colour (HIP events)
    ncol = min(30, F)
    rng = np.random.default_rng(11)
    rgbs = torch.from_numpy(rng.integers(0, 256, (ncol, w.image[1], w.image[0], 3),
                                         dtype=np.uint8)).cuda()

    def frames_ms(colour):
        vol.reset(stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for f in range(ncol):
            vol.integrate(dev[f], w.poses[f], stream, rgb=rgbs[f] if colour else None)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / ncol

    log("timing the colour integrate")
    vol.enable_color(True, stream)
    frames_ms(True)
    colour_ms = frames_ms(True)
    plain_ms = frames_ms(False)
    vol.enable_color(False, stream)
    del rgbs
    # sparse brick volume on the same lattice and frames: per-frame calls
    # (allocation + integrate), bricks and memory against the dense volume;
    # the mesh of the exported grid box against the dense mesh
    log("timing the sparse volume")
    sv = recon.SparseVolume(w.config(), 1 << 20)
    sv.set_bounds((0, 0, 0), w.grid)        # the dense grid's region
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for f in range(F):
        sv.integrate(dev[f], w.poses[f], stream)
    e1.record(stream)
    torch.cuda.synchronize()
    sparse_ms = e0.elapsed_time(e1) / F
    nb, flags = sv.stats(stream)
    Ds, _ = sv.export((0, 0, 0), w.grid, stream)
    sverts, _ = recon.mc_extract(Ds, w.grid, list(vol.info.voxel_size), list(vol.info.global_offset),
                                 stream=stream)
    torch.cuda.synchronize()
    result["sparse"] = {"frames": F, "ms_per_call": round(sparse_ms, 4),
                        "dense_ms_per_call": round(avg_launch_ms, 4), "bricks": nb,
                        "overflow_flags": flags, "brick_bytes": nb * 4096,
                        "dense_bytes": 8 * w.num_voxels,
                        "mesh_triangles_grid_box": int(sverts.shape[0])}
    del Ds, sverts
    sv.close()
    result["colour"] = {"frames": ncol, "ms_per_call_colour": round(colour_ms, 4),
                        "ms_per_call_plain": round(plain_ms, 4),
                        "colour_over_plain": round(colour_ms / plain_ms, 3),
                        "colour_volume_bytes": 16 * w.num_voxels}
    vol.close()
    del dev
    torch.cuda.empty_cache()
    log("timing the small workload (128^3 x 10 frames) on the GPU and the CPU")
    result["small"] = bench_small(max(args.steps, 5), cpu=not args.no_cpu)
    if not args.no_render:
        log("timing the point-cloud render")
        result["render"] = bench_render(w, depth, args.steps, cpu=not args.no_cpu)
        torch.cuda.empty_cache()
    if not args.no_batched:
        log("timing the batched volumes")
        result["batched"] = bench_batched(max(args.steps, 2), args.autotune)
    if not args.no_diff:
        log("timing the differentiable forward + adjoint")
        result["differentiable"] = bench_differentiable(max(args.steps, 3))
    if not args.no_refine:
        log("batched pose refinement")
        result["pose_refinement"] = bench_pose_refinement()
    if not args.no_cpu:
        log("timing the CPU baseline")
        result["cpu_baseline"] = cpu_baseline(w, depth, args.cpu_seconds)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

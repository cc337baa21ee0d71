"""Benchmark of the path-tracing hot path on one MI355X.

Workload (default): 64 displaced subdivision-6 icospheres (81,920 triangles
each; 5.24M unique, 65.5M instanced triangles), 800 instances with random
rigid transforms in a 100^3 box, 32 emissive instances; 1920x1080, 1 sample
per pixel per step, 4 bounces (max_depth 4) with next event estimation
(GGX sampling + Lambert NEE).  One step = one frame: generate, 4x(extend,
shade, shadow), accumulate, with the scene already resident in HBM.

value = incoherent Mrays/s: extension rays at depth >= 1 plus shadow rays,
divided by the wall time of those stages (hipEvents on the library's two
streams: the shadow rays of bounce k trace beside bounce k+1's extension
rays, and the time counted is from each shade stage's end until both
launches it feeds are done; exact counts from the queue lengths).

Prints one JSON line.  --gpus N > 1 is not supported on this build.
"""
from __future__ import annotations

import argparse
import ctypes as C
import csv
import json
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "incoherent Mrays/s + fwd+adjoint ms/iter vs plain fwd, 1 MI355X, 1080p"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of the traversal kernels, per unit of work.  Node pairs are
# fetched together (both children of an interior node: 64 B); the TLAS root is
# read once per ray and each entered BLAS reads its root (32 B), the instance
# index (4 B) and the instance fields used (world_object 64 B, object_world
# 64 B, node/index offsets 8 B); a triangle test reads the traversal-ordered
# triangle record (48 B).
B_NODE_PAIR = 64
B_ROOT = 32
B_INSTANCE = 32 + 4 + 64 + 64 + 8
B_TRIANGLE = 48
B_RAY_EXTEND = 32 + 20     # ray in, hit out
B_RAY_SHADOW = 48 + 16     # ray + threshold + colour in, colour out
# The 4-wide traversal (loop 2): a node visit reads one 64 B quantised node
# (the roots included); an entered instance reads the two 64 B lines of its
# instance record (world_object rows 0-2 and the BLAS root on entry;
# object_world rows 0-2, index offset and instance when the ray leaves it).
B_WIDE_NODE = 64
B_INSTANCE_WIDE = 64 + 64


def bytes_per_ray(work: dict, io: int) -> float:
    r = max(work["rays"], 1)
    return (B_ROOT * r + B_NODE_PAIR * (work["tlas_interior"] + work["blas_interior"])
            + B_INSTANCE * work["instances_entered"] + B_TRIANGLE * work["tri_tests"]) / r + io


def bytes_per_ray_wide(work, io: int) -> float:
    """work = [rays, wide node visits, triangle tests, instances entered]
    from the kernels' own census (PULSE_FLAG_COUNT_WORK)."""
    r = max(int(work[0]), 1)
    return (B_WIDE_NODE * work[1] + B_INSTANCE_WIDE * work[3] + B_TRIANGLE * work[2]) / r + io


def cpu_baseline(scene, cam, prm, seconds: float):
    """The f32 oracle (scalar C++ restatement) timed on the host cores on a
    bounded sample of rows of the same frame; also counts the traversal work
    of the kernels' algorithm for the bytes model."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import OracleScene

    # one thread per host core this process may run on (the affinity mask:
    # a container's CPU share, not the machine's core count)
    try:
        threads = max(1, len(os.sched_getaffinity(0)))
    except (AttributeError, OSError):
        threads = max(1, os.cpu_count() or 1)
    orc = OracleScene(scene)
    acc = np.zeros((cam.height, cam.width, 4), np.float32)
    rows, ms_total = 0, 0.0
    tot = {"rays": 0, "primary": 0, "shadow": 0}
    step = max(1, threads // 2)
    row = 0
    while row < cam.height and ms_total < seconds * 1000.0:
        r1 = min(cam.height, row + step)
        _, c, ms = orc.render(cam, prm, accum=acc, threads=threads, rows=(row, r1))
        ms_total += ms
        tot["rays"] += c["rays"]
        tot["primary"] += c["primary_rays"]
        tot["shadow"] += c["shadow_rays"]
        rows += r1 - row
        row = r1
        step = min(step * 2, 64)
    incoherent = tot["rays"] - tot["primary"]
    # work counters of the kernels' traversal on a few rows of the sample
    orc.set_cull_world(True)
    r0 = cam.height // 2
    _, cw, _ = orc.render(cam, prm, accum=np.zeros_like(acc), threads=threads, rows=(r0, min(cam.height, r0 + 16)))
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    base = {
        "value": incoherent / (ms_total / 1e3) / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model,
        "kind": "port",
        "sample": f"{rows} of {cam.height} rows of the same frame ({rows * cam.width} paths, {tot['rays']} rays, "
                  f"{incoherent} incoherent) in {ms_total / 1e3:.1f} s, one thread per core of the "
                  f"affinity mask ({threads} of nproc {os.cpu_count()}); "
                  f"value = incoherent rays / whole render time",
        "all_rays_mrays_per_s": tot["rays"] / (ms_total / 1e3) / 1e6,
    }
    return base, cw["by_category"]


def bench_adjoint(scene, camf, iters: int = 5) -> dict:
    """Differentiable form at 512x512, 16 spp, 3 bounces (GGX + Lambert NEE,
    no Russian roulette) on the same geometry with 64 materials: the target
    is rendered at 64 spp from hidden parameters (albedo +-0.3, roughness
    +-0.3 from the start values); timed: plain forward vs forward + loss +
    adjoint (pulse_forward_adjoint: its forward pass records the paths'
    closest hits and shadow visibilities, which the adjoint reads instead of
    tracing the paths again)."""
    from pulse_amd import _native as N
    from pulse_amd.path_tracer import Device, PulseAdjoint, apply_material_params, make_params, material_params

    dev = Device(0)
    dev.upload_scene(scene)
    cam = camf(512, 512)
    mats = scene.buffers["materials"]
    nm = len(mats)
    adj = PulseAdjoint(dev, 512, 512, nm)
    rng = np.random.default_rng(4)
    start = material_params(mats)
    hidden = start.copy()
    hidden[:, 0:3] = np.clip(start[:, 0:3] + rng.uniform(-0.3, 0.3, (nm, 3)), 0.02, 0.98)
    hidden[:, 3] = np.clip(start[:, 3] + rng.uniform(-0.3, 0.3, nm), 0.05, 1.0)
    adj.set_materials(apply_material_params(mats, hidden))
    clamp = N.FLAG_HORIZON_CLAMP
    adj.set_target(adj.forward(cam, make_params(spp=64, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=3,
                                                flags=clamp)))
    adj.set_materials(mats)
    prm = make_params(spp=16, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=3, flags=clamp)
    adj.forward(cam, prm)
    adj.forward_backward(cam, prm)  # warm-up
    # both legs end with the image on the device: the forward is pulse_trace
    # and a synchronisation, the other pulse_forward_adjoint (forward, loss,
    # adjoint) and the gradient's copy to the host
    lib = N.lib()
    t0 = time.perf_counter()
    for _ in range(iters):
        N.check(lib.pulse_trace(dev.handle, C.byref(cam), C.byref(prm), adj.image.ptr))
        dev.sync()
    fwd_ms = (time.perf_counter() - t0) * 1e3 / iters
    t0 = time.perf_counter()
    for _ in range(iters):
        loss, g = adj.forward_backward(cam, prm)
    fa_ms = (time.perf_counter() - t0) * 1e3 / iters
    dev.close()
    return {"forward_ms": round(fwd_ms, 3), "forward_adjoint_ms": round(fa_ms, 3),
            "ratio": round(fa_ms / fwd_ms, 3), "loss": loss, "grad_norm": float(np.linalg.norm(g)),
            "config": "512x512, 16 spp, max_depth 3, GGX + Lambert NEE, no RR, horizon clamp, "
                      "65 materials x 7 parameters; the forward pass of pulse_forward_adjoint records its paths "
                      "(closest hits, shadow visibilities) and the adjoint reads them"}


def bench_adjoint_views(iters: int = 3) -> dict:
    """Batched adjoint on the large workload's geometry: 256 displaced
    subdivision-6 icosphere BLAS (21M unique triangles), 1024 instances (256
    emissive), one material per instance (4096 per-instance albedo and
    roughness parameters, 7175 gradient entries with the emission: the
    gradient lives in device memory); 8 views of 512x512 on a circle around
    the field, 4 spp, 6 bounces, GGX + Lambert NEE, no RR, horizon clamp.
    Targets rendered at 16 spp from hidden parameters (albedo and roughness
    +-0.3).  Timed (best of `iters`, each ending in a synchronisation): the 8
    forward renders; pulse_adjoint_views over the 8 views in one launch (the
    default: each wave's lanes summed per material before their atomic);
    the same with per-lane atomics; 8 pulse_adjoint calls accumulating into
    one gradient."""
    from pulse_amd import _native as N
    from pulse_amd import procedural as P
    from pulse_amd.path_tracer import Device, MultiViewAdjoint, apply_material_params, make_params, material_params

    try:
        threads = max(1, len(os.sched_getaffinity(0)))
    except (AttributeError, OSError):
        threads = max(1, os.cpu_count() or 1)
    t0 = time.perf_counter()
    scene, _ = P.icosphere_field(n_meshes=256, subdiv=6, n_instances=1024, n_emissive=256, box=110.0, seed=5,
                                 threads=min(threads, 16), roughness=(0.3, 1.0), per_instance_materials=True)
    build_s = time.perf_counter() - t0
    dev = Device(0)
    lib = N.lib()
    try:
        dev.upload_scene(scene)
        mats = scene.buffers["materials"]
        nm = len(mats)
        cams = [P.make_camera(512, 512, eye=(140.0 * np.cos(a), 25.0, 140.0 * np.sin(a)), target=(0.0, 0.0, 0.0))
                for a in np.linspace(0.0, 2.0 * np.pi, 8, endpoint=False)]
        mv = MultiViewAdjoint(dev, cams, nm)
        rng = np.random.default_rng(5)
        start = material_params(mats)
        hidden = start.copy()
        hidden[:, 0:3] = np.clip(start[:, 0:3] + rng.uniform(-0.3, 0.3, (nm, 3)), 0.02, 0.98)
        hidden[:, 3] = np.clip(start[:, 3] + rng.uniform(-0.3, 0.3, nm), 0.05, 1.0)
        clamp = N.FLAG_HORIZON_CLAMP
        mh = np.ascontiguousarray(apply_material_params(mats, hidden), N.MATERIAL_DTYPE)
        N.check(lib.pulse_scene_update_materials(dev.handle, mh.ctypes.data, nm))
        mv.set_targets(mv.forward(make_params(spp=16, max_depth=6, mode=N.MODE_GGX_NEE, rr_start_depth=6,
                                              flags=clamp)))
        m0 = np.ascontiguousarray(mats, N.MATERIAL_DTYPE)
        N.check(lib.pulse_scene_update_materials(dev.handle, m0.ctypes.data, nm))
        prm = make_params(spp=4, max_depth=6, mode=N.MODE_GGX_NEE, rr_start_depth=6, flags=clamp)
        acc = N.Params.from_buffer_copy(prm)
        acc.flags |= N.FLAG_ADJOINT_ACCUMULATE
        mv.forward(prm)
        singles = [(N.AdjointView * 1)(mv.views[v]) for v in range(len(cams))]

        def best(fn):
            fn()
            dev.sync()
            t = 1e30
            for _ in range(iters):
                t1 = time.perf_counter()
                fn()
                dev.sync()
                t = min(t, time.perf_counter() - t1)
            return t * 1e3

        def forward():
            for cam, img in zip(mv.cameras, mv.images):
                N.check(lib.pulse_trace(dev.handle, C.byref(cam), C.byref(prm), img.ptr))

        def batched():
            N.check(lib.pulse_adjoint_views(dev.handle, mv.views, len(cams), C.byref(prm), mv.grad.ptr))

        def sequential():
            for v, one in enumerate(singles):
                N.check(lib.pulse_adjoint_views(dev.handle, one, 1, C.byref(prm if v == 0 else acc), mv.grad.ptr))

        fwd_ms = best(forward)
        batched_ms = best(batched)
        g = mv.grad.copy_to(np.zeros(nm * 7, np.float32))
        seq_ms = best(sequential)
        dev.set_option(N.OPT_ADJOINT_REDUCE_ROUNDS, 0)
        per_lane_ms = best(batched)
        dev.set_option(N.OPT_ADJOINT_REDUCE_ROUNDS, -1)
        # the views' forward passes keep their paths (one record per camera)
        # and the batched adjoint reads them instead of tracing again
        rec = N.Params.from_buffer_copy(prm)
        rec.flags |= N.FLAG_RECORD_PATHS

        def forward_recorded():
            for cam, img in zip(mv.cameras, mv.images):
                N.check(lib.pulse_trace(dev.handle, C.byref(cam), C.byref(rec), img.ptr))

        fwd_rec_ms = best(forward_recorded)
        batched_rec_ms = best(batched)
        read_record = dev.get_option(N.OPT_ADJOINT_READ_RECORD)
        g_rec = mv.grad.copy_to(np.zeros(nm * 7, np.float32))
        losses = mv.losses()
    finally:
        dev.close()
    return {"forward_8_views_ms": round(fwd_ms, 3), "adjoint_batched_ms": round(batched_ms, 3),
            "adjoint_sequential_ms": round(seq_ms, 3), "adjoint_batched_per_lane_atomics_ms": round(per_lane_ms, 3),
            "batched_speedup": round(seq_ms / batched_ms, 3),
            "forward_plus_batched_adjoint_ratio": round((fwd_ms + batched_ms) / fwd_ms, 3),
            "recorded_paths": {"forward_8_views_ms": round(fwd_rec_ms, 3),
                               "adjoint_batched_ms": round(batched_rec_ms, 3), "read_record": int(read_record),
                               "forward_plus_batched_adjoint_ratio": round((fwd_rec_ms + batched_rec_ms) / fwd_ms, 3),
                               "grad_rel_diff": float(np.abs(g_rec - g).max() / max(np.abs(g).max(), 1e-30))},
            "loss_sum": float(sum(losses)), "grad_norm": float(np.linalg.norm(g)),
            "scene_build_s": round(build_s, 1),
            "config": "256 BLAS x 1024 instances (21M unique triangles), 1025 materials x 7 parameters, "
                      "8 views 512x512, 4 spp, max_depth 6, GGX + Lambert NEE, no RR, horizon clamp"}


def bench_instance_update(dev, scene, full_upload_s: float) -> dict:
    """Per-frame scene change as the reference does it every frame (new
    instance transforms -> instance records + TLAS rebuilt on the host,
    src/scene/mod.rs:561-697): 80 of the 800 instances moved, host rebuild
    (pulse_sb_update_instances) and device update of the three instance
    buffers (pulse_scene_update_instances), against a full 13-buffer upload."""
    sb = scene.builder
    u = np.random.default_rng(5).uniform(-3.0, 3.0, (80, 3))
    inst = scene.buffers["instances"]
    for k, i in enumerate(range(0, 800, 10)):
        m = np.asarray(inst[i]["object_world"], np.float64).reshape(4, 4).T.copy()  # column-major -> row-major
        m[:3, 3] += u[k]
        sb.set_instance_transform(i, m)
    t0 = time.perf_counter()
    moved = sb.update_instances(scene)
    host_ms = (time.perf_counter() - t0) * 1e3
    dev.sync()
    t0 = time.perf_counter()
    dev.update_instances(moved)
    dev.sync()
    device_ms = (time.perf_counter() - t0) * 1e3
    nbytes = sum(moved.buffers[k].nbytes for k in ("instances", "tlas_nodes", "instance_indices"))
    # the same motion without the host: the 800 matrices go up, instance
    # records, world boxes and the 4-wide TLAS refit are computed on the GPU
    # (pulse_scene_refit_instances; the first call sizes its scratch)
    mats = moved.buffers["instances"]["object_world"]
    dev.refit_instances(mats)
    t0 = time.perf_counter()
    dev.refit_instances(mats)
    refit_ms = (time.perf_counter() - t0) * 1e3
    # the same with a new 4-wide TLAS built on the GPU (synchronous)
    dev.rebuild_instances(mats)
    t0 = time.perf_counter()
    dev.rebuild_instances(mats)
    rebuild_ms = (time.perf_counter() - t0) * 1e3
    return {"moved_instances": 80, "host_rebuild_ms": round(host_ms, 2),
            "builder_instances_tlas_ms": round(moved.info["ms_instances_tlas"], 2),
            "device_update_ms": round(device_ms, 2), "device_update_bytes": int(nbytes),
            "device_refit_ms": round(refit_ms, 3), "device_refit_bytes": int(mats.nbytes),
            "device_rebuild_ms": round(rebuild_ms, 3),
            "full_upload_ms": round(full_upload_s * 1e3, 1)}


def _pmc_run(counter: str, extra_args, timeout_s: float = 150.0):
    """Runs this script under `rocprofv3 --pmc <counter>` (a child process,
    its own process group, killed at the time limit) and returns the rows of
    the counter CSV in dispatch order, or None."""
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    out = tempfile.mkdtemp(prefix="pulse_pmc_", dir="/tmp")
    cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "pmc", "--",
           sys.executable, str(Path(__file__).resolve())] + list(extra_args)
    env = dict(os.environ, TMPDIR="/tmp")
    try:
        p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                             start_new_session=True)
        try:
            p.wait(timeout=timeout_s)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            return None
        if p.returncode != 0:
            return None
        files = list(Path(out).rglob("*counter_collection.csv"))
        if not files:
            return None
        rows = list(csv.DictReader(open(files[0])))
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        return rows
    except OSError:
        return None
    finally:
        shutil.rmtree(out, ignore_errors=True)


def pmc_traffic(max_depth: int, loop: int) -> dict | None:
    """HBM bytes per dispatch of the traversal launches, from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch, one counter per run) over one frame of this
    workload, each scaled by what the same counter reads for a plain 256 MiB
    device copy (bytes moved / counter reading).  Dispatches are told apart
    by their order within the frame: extend (trace index 0 = primary, then
    bounces) and shadow launches alternate with shading."""
    frame_args = ["--steps", "1", "--warmup", "0", "--no-cpu", "--no-adjoint", "--no-batch", "--no-pmc", "--no-loop1",
                  "--no-subdiv5", "--max-depth", str(max_depth), "--loop", str(loop)]
    scale = {}
    per_kind = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = _pmc_run(counter, ["--copy-calibration", "256"])
        if not cal:
            return None
        copies = [float(r["Counter_Value"]) for r in cal if "copy" in r["Kernel_Name"].lower()]
        if not copies:
            return None
        scale[counter] = (256 * 1024) / (sum(copies) / len(copies))  # KiB moved per KiB counted
        rows = _pmc_run(counter, frame_args)
        if not rows:
            return None
        ext = [float(r["Counter_Value"]) for r in rows if "k_trace_queue" in r["Kernel_Name"] and "false" in r["Kernel_Name"]]
        sh = [float(r["Counter_Value"]) for r in rows if "k_trace_queue" in r["Kernel_Name"] and "true" in r["Kernel_Name"]]
        if not ext or not sh:
            return None
        # the measured frame is the first max_depth extend / shadow dispatches
        # (a work-census frame may follow it)
        ext, sh = ext[:max_depth], sh[:max_depth]
        per_kind.setdefault("extend_primary", {})[counter] = ext[0]
        per_kind.setdefault("extend_bounce", {})[counter] = sum(ext[1:]) / max(len(ext) - 1, 1)
        per_kind.setdefault("shadow", {})[counter] = sum(sh) / len(sh)
        per_kind.setdefault("extend", {})[counter] = sum(ext) / len(ext)
    out = {"scale": {k: round(v, 3) for k, v in scale.items()}}
    for kind, v in per_kind.items():
        out[kind] = (v["FETCH_SIZE"] * scale["FETCH_SIZE"] + v["WRITE_SIZE"] * scale["WRITE_SIZE"]) * 1024.0
    return out


def bench_loop1(dev, node, cam, args, frame: int, frames: int = 4) -> dict:
    """The same frames with traversal loop 1 (the reference's binary BVHs and
    visiting order: bit-exact with the oracle), timed the same way."""
    from pulse_amd import _native as N
    from pulse_amd.path_tracer import make_params

    dev.set_option(N.OPT_TRAVERSAL_LOOP, 1)
    flags = N.FLAG_STAGE_TIMING | N.FLAG_COLLECT_STATS
    try:
        node.run(cam, make_params(accumulation_count=frame, max_depth=args.max_depth, mode=N.MODE_GGX_NEE, flags=flags))
        rays = ms = 0.0
        for i in range(frames):
            node.run(cam, make_params(accumulation_count=frame + 1 + i, max_depth=args.max_depth,
                                      mode=N.MODE_GGX_NEE, flags=flags))
            st = dev.stats()
            rays += st.incoherent_rays
            ms += st.ms_incoherent
    finally:
        dev.set_option(N.OPT_TRAVERSAL_LOOP, args.loop)
    return {"value": round(rays / (ms / 1e3) / 1e6, 2), "unit": "Mrays/s", "frames": frames}


def bench_subdiv5(args, frames: int = 10, warmup: int = 2) -> dict:
    """The same workload with subdivision-5 meshes (20,480 triangles each:
    1.31M unique / 16.4M instanced, about 1M unique as the workload's
    description reads it), timed the same way in its own context (same
    traversal loop and launch options)."""
    from pulse_amd import _native as N
    from pulse_amd import procedural as P
    from pulse_amd.path_tracer import Device, PulsePathTracerNode, make_params

    scene, camf = P.icosphere_field(subdiv=5, threads=min(16, os.cpu_count() or 1))
    cam = camf(args.width, args.height)
    dev = Device(0)
    try:
        dev.set_option(N.OPT_TRAVERSAL_LOOP, args.loop)
        if args.concurrent is not None:
            dev.set_option(N.OPT_CONCURRENT_SHADOW, args.concurrent)
        dev.upload_scene(scene)
        node = PulsePathTracerNode(dev, args.width, args.height)
        flags = N.FLAG_STAGE_TIMING | N.FLAG_COLLECT_STATS
        rays = ms = ms_total = 0.0
        for f in range(warmup + frames):
            node.run(cam, make_params(accumulation_count=f, max_depth=args.max_depth, mode=N.MODE_GGX_NEE, flags=flags))
            st = dev.stats()
            if st.error_flags:
                raise RuntimeError(f"device error flags {st.error_flags}")
            if f >= warmup:
                rays += st.incoherent_rays
                ms += st.ms_incoherent
                ms_total += st.ms_total
        dev.sync()
    finally:
        dev.close()
    return {"value": round(rays / (ms / 1e3) / 1e6, 2), "unit": "Mrays/s", "frames": frames,
            "ms_per_frame": round(ms_total / frames, 3), "unique_triangles": scene.info["unique_triangles"],
            "instanced_triangles": scene.info["instanced_triangles"]}


def bench_batch(frames: int = 2, check_views: int = 4) -> dict:
    """Batched form: 32 views, resolutions from {512^2, 720p, 1080p}, icosphere
    fields of 10k-2M unique triangles (log-uniform), 2 spp, 4 bounces, GGX +
    Lambert NEE with Russian roulette.  Timed: 32 sequential plain launches
    (bind scene, pulse_trace) per frame against one batched frame (autotuned
    launch shape, hipGraph replay); both with scenes resident."""
    from pulse_amd import _native as N
    from pulse_amd import procedural as P
    from pulse_amd.path_tracer import Batch, Device, make_params

    t0 = time.perf_counter()
    work = P.batch_workload(threads=min(16, os.cpu_count() or 1))
    build_s = time.perf_counter() - t0
    dev = Device(0)
    scenes = [dev.create_scene(s) for s, _, _ in work]
    cams = [c for _, c, _ in work]
    prm = make_params(spp=2, max_depth=4, mode=N.MODE_GGX_NEE)
    seq_acc = [dev.alloc(c.width * c.height * 16) for c in cams]

    def sequential(acc):
        p = N.Params.from_buffer_copy(prm)
        p.accumulation_count = acc
        for sc, cam, buf in zip(scenes, cams, seq_acc):
            dev.bind_scene(sc)
            N.check(N.lib().pulse_trace(dev.handle, cam, p, buf.ptr))

    sequential(0)
    dev.sync()
    t0 = time.perf_counter()
    for f in range(frames):
        sequential(2 * f + 2)
    dev.sync()
    seq_ms = (time.perf_counter() - t0) * 1e3 / frames
    print(f"bench: batch workload built in {build_s:.1f} s; sequential {seq_ms:.1f} ms/frame", file=sys.stderr,
          flush=True)

    batch = Batch(dev, list(zip(scenes, cams)), prm)
    batch.render(0)
    dev.sync()
    t0 = time.perf_counter()
    batch.render(0)
    dev.sync()
    default_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    cfg, tuned_ms = batch.autotune(reps=1)
    tune_s = time.perf_counter() - t0
    batch.render(0)
    batch.stats()
    t0 = time.perf_counter()
    for f in range(frames):
        batch.render(2 * f + 2)
    dev.sync()
    batch_ms = (time.perf_counter() - t0) * 1e3 / frames
    st = batch.stats()
    # bit-exact check of the smallest views against the sequential images (same frames)
    order = sorted(range(len(cams)), key=lambda i: cams[i].width * cams[i].height)[:check_views]
    imgs = batch.images()
    exact = 0
    for i in order:
        ref = seq_acc[i].copy_to(np.zeros((cams[i].height, cams[i].width, 4), np.float32))
        exact += int(np.array_equal(ref.view(np.uint32), imgs[i].view(np.uint32)))
    rays = (st.extension_rays + st.shadow_rays) / frames
    out = {"views": len(cams), "pixels": int(sum(c.width * c.height for c in cams)),
           "unique_triangles": int(sum(i["unique_triangles"] for _, _, i in work)),
           "spp": 2, "max_depth": 4, "host_build_s": round(build_s, 1),
           "sequential_ms": round(seq_ms, 2), "batched_ms": round(batch_ms, 2),
           "batched_default_shape_ms": round(default_ms, 2), "speedup": round(seq_ms / batch_ms, 3),
           "rays_per_frame": rays, "batched_mrays_per_s": round(rays / batch_ms / 1e3, 1),
           "autotune_s": round(tune_s, 1), "autotuned_shape": cfg.as_dict(),
           "bit_exact_views": f"{exact}/{len(order)}"}
    batch.close()
    for s in scenes:
        s.close()
    dev.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--subdiv", type=int, default=6)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--max-depth", type=int, default=4)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--loop", type=int, default=2,
                    help="traversal loop (0 nested, 1 single loop: the reference's order; 2: 4-wide BVHs)")
    ap.add_argument("--persistent", type=int, default=None, help="persistent traversal waves per CU (0: off)")
    ap.add_argument("--refill", type=int, default=None, help="idle lanes that trigger a persistent refill")
    ap.add_argument("--gather", type=int, default=None, help="lanes gathered before an instance change runs")
    ap.add_argument("--period", type=int, default=None, help="persistent: instance changes run at least every n steps")
    ap.add_argument("--pair-waves", type=int, default=None,
                    help="persistent: waves per CU of each of two concurrent launches (0: half)")
    ap.add_argument("--pair-waves-shadow", type=int, default=None,
                    help="persistent: waves per CU of the shadow launch of a concurrent pair (0: --pair-waves)")
    ap.add_argument("--lds", type=int, default=None, help="persistent: traversal stack entries kept in LDS")
    ap.add_argument("--tps", type=int, default=None, help="persistent: leaf triangles tested per step (1 or 2)")
    ap.add_argument("--concurrent", type=int, default=None, choices=(0, 1),
                    help="persistent: shadow rays of bounce k beside bounce k+1's extension rays (default 1)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE runs")
    ap.add_argument("--no-loop1", action="store_true", help="skip the reference-order (loop 1) comparison frames")
    ap.add_argument("--no-subdiv5", action="store_true", help="skip the subdivision-5 variant of the workload")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (reference kernel for HBM counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build renders on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard pixel rows or views across GPUs, see the design notes).", file=sys.stderr)
        return 2

    from pulse_amd import _native as N
    from pulse_amd import procedural as P
    from pulse_amd.path_tracer import Device, PulsePathTracerNode, make_params

    t0 = time.perf_counter()
    scene, camf = P.icosphere_field(subdiv=args.subdiv, threads=min(16, os.cpu_count() or 1), keep_builder=True)
    host_build_s = time.perf_counter() - t0
    cam = camf(args.width, args.height)
    print(f"bench: scene built in {host_build_s:.1f} s: {scene.info['unique_triangles']} unique / "
          f"{scene.info['instanced_triangles']} instanced triangles", file=sys.stderr, flush=True)

    dev = Device(0)
    dev.set_option(N.OPT_TRAVERSAL_LOOP, args.loop)
    if args.persistent is not None:
        dev.set_option(N.OPT_PERSISTENT_WAVES, args.persistent)
    if args.refill is not None:
        dev.set_option(N.OPT_REFILL_LANES, args.refill)
    if args.gather is not None:
        dev.set_option(N.OPT_GATHER_LANES, args.gather)
    if args.period is not None:
        dev.set_option(N.OPT_ADVANCE_PERIOD, args.period)
    if args.pair_waves is not None:
        dev.set_option(N.OPT_PAIR_WAVES, args.pair_waves)
    if args.pair_waves_shadow is not None:
        dev.set_option(N.OPT_PAIR_WAVES_SHADOW, args.pair_waves_shadow)
    if args.lds is not None:
        dev.set_option(N.OPT_LDS_STACK_ENTRIES, args.lds)
    if args.tps is not None:
        dev.set_option(N.OPT_TRIS_PER_STEP, args.tps)
    if args.concurrent is not None:
        dev.set_option(N.OPT_CONCURRENT_SHADOW, args.concurrent)
    t0 = time.perf_counter()
    dev.upload_scene(scene)
    dev.sync()
    upload_s = time.perf_counter() - t0
    node = PulsePathTracerNode(dev, args.width, args.height)
    flags = N.FLAG_STAGE_TIMING | N.FLAG_COLLECT_STATS
    frame = 0
    for _ in range(args.warmup):
        node.run(cam, make_params(accumulation_count=frame, max_depth=args.max_depth, mode=N.MODE_GGX_NEE, flags=flags))
        frame += 1
    dev.sync()
    print("bench: warmup done", file=sys.stderr, flush=True)

    agg = {k: 0.0 for k in ["incoherent_rays", "extension_rays", "primary_rays", "shadow_rays", "ms_total",
                            "ms_incoherent", "ms_primary_extend", "ms_extend_bounce", "ms_shadow", "ms_shade",
                            "extend_launches", "shadow_launches"]}
    dev.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        node.run(cam, make_params(accumulation_count=frame, max_depth=args.max_depth, mode=N.MODE_GGX_NEE, flags=flags))
        frame += 1
        st = dev.stats()
        if st.error_flags:
            raise RuntimeError(f"device error flags {st.error_flags}")
        for k in agg:
            agg[k] += getattr(st, k)
    dev.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = agg["incoherent_rays"] / (agg["ms_incoherent"] / 1e3) / 1e6
    all_rays = agg["extension_rays"] + agg["shadow_rays"]
    img = node.readback()
    census = None
    if args.loop == 2:
        # one more frame (untimed) with the kernels' work census for the bytes model
        node.run(cam, make_params(accumulation_count=frame, max_depth=args.max_depth, mode=N.MODE_GGX_NEE,
                                  flags=N.FLAG_COUNT_WORK))
        w = dev.stats().work
        census = {c: [int(w[i][j]) for j in range(4)] for i, c in enumerate(("primary", "bounce", "shadow"))}
    loop1 = None
    if args.loop == 2 and not args.no_loop1:
        loop1 = bench_loop1(dev, node, cam, args, frame + 1)
    instance_update = bench_instance_update(dev, scene, upload_s)
    print(f"bench: {args.steps} frames, {ms_per_step:.2f} ms/frame, {value:.1f} incoherent Mrays/s", file=sys.stderr,
          flush=True)

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural displaced icospheres, seeded)",
        "config": {
            "workload": f"icosphere_field: 64 BLAS x subdiv-{args.subdiv} displaced icospheres, 800 instances, "
                        f"32 emissive, {args.width}x{args.height}, 1 spp/step, max_depth {args.max_depth}, "
                        "GGX sampling + Lambert NEE, Russian roulette",
            "unique_triangles": scene.info["unique_triangles"],
            "instanced_triangles": scene.info["instanced_triangles"],
            "width": args.width, "height": args.height, "spp_per_step": 1, "max_depth": args.max_depth,
        },
        "rays_per_step": {k: agg[k] / args.steps for k in ["primary_rays", "extension_rays", "shadow_rays",
                                                            "incoherent_rays"]},
        "stage_ms_per_step": {k: round(agg[k] / args.steps, 3) for k in ["ms_total", "ms_primary_extend",
                                                                         "ms_incoherent", "ms_extend_bounce",
                                                                         "ms_shadow", "ms_shade"]},
        "all_rays_mrays_per_s": round(all_rays / wall / 1e6, 2),
        "host_build_s": round(host_build_s, 2),
        "upload_s": round(upload_s, 2),
        "traversal_loop": args.loop,
        "persistent_waves_per_cu": dev.get_option(N.OPT_PERSISTENT_WAVES),
        "refill_lanes": dev.get_option(N.OPT_REFILL_LANES),
        "concurrent_shadow": dev.get_option(N.OPT_CONCURRENT_SHADOW),
        "gather_lanes": dev.get_option(N.OPT_GATHER_LANES),
        "stack_entries": dev.get_option(N.OPT_STACK_ENTRIES),
        "fast_division": dev.get_option(N.OPT_FAST_DIVISION),
        "image_mean": float(np.mean(img[..., :3])),
        "instance_update": instance_update,
    }
    if loop1 is not None:
        result["reference_order_loop1"] = loop1
    if census is not None:
        result["work_census"] = census
    # The legs below each open a context of their own; close this one first.
    # HIP serves a process's streams from a few hardware queues per device
    # (GPU_MAX_HW_QUEUES, 4), and with this context still open the next
    # one's two streams were placed on one queue, which runs its concurrent
    # extension and shadow launches one after the other (the 512x512 forward
    # of the adjoint leg: 35.5 ms alone, 47.5 ms beside an open context).
    del node
    dev.close()
    if not args.no_subdiv5 and args.subdiv == 6:
        result["subdiv5"] = bench_subdiv5(args)
        print(f"bench: subdivision-5 meshes {result['subdiv5']['value']} incoherent Mrays/s", file=sys.stderr,
              flush=True)

    if not args.no_adjoint:
        result["fwd_adjoint"] = bench_adjoint(scene, camf)
        print(f"bench: forward+adjoint {result['fwd_adjoint']['ratio']}x forward", file=sys.stderr, flush=True)
        result["adjoint_views"] = bench_adjoint_views()
        print(f"bench: batched adjoint over 8 views {result['adjoint_views']['batched_speedup']}x sequential",
              file=sys.stderr, flush=True)

    if not args.no_batch:
        result["batched"] = bench_batch()
        print(f"bench: batched {result['batched']['speedup']}x sequential", file=sys.stderr, flush=True)

    if not args.no_cpu:
        prm = make_params(accumulation_count=0, max_depth=args.max_depth, mode=N.MODE_GGX_NEE)
        base, work = cpu_baseline(scene, cam, prm, args.cpu_seconds)
        result["cpu_baseline"] = base
        if census is not None:
            bpr = {"primary": bytes_per_ray_wide(census["primary"], B_RAY_EXTEND),
                   "bounce": bytes_per_ray_wide(census["bounce"], B_RAY_EXTEND),
                   "shadow": bytes_per_ray_wide(census["shadow"], B_RAY_SHADOW)}
        else:
            bpr = {"primary": bytes_per_ray(work["primary"], B_RAY_EXTEND),
                   "bounce": bytes_per_ray(work["bounce"], B_RAY_EXTEND),
                   "shadow": bytes_per_ray(work["shadow"], B_RAY_SHADOW)}
        # the two traversal kernels over all their launches: the extension
        # kernel (primary and bounce rays: one kernel, the largest share of
        # GPU time in the kernel trace) and the shadow kernel.  roofline
        # names the extension kernel; each launch's own duration (hipEvents
        # on the stream it runs on) includes the share of the GPU a launch on
        # the other stream takes while both run
        persistent = result["persistent_waves_per_cu"] > 0 and args.loop != 0
        ext_name = "k_trace_queue<..., false>" if persistent else "k_extend"
        sh_name = "k_trace_queue<..., true>" if persistent else "k_shadow"
        bounce_rays = agg["extension_rays"] - agg["primary_rays"]
        kernels = {
            "extend": (f"{ext_name} (extension rays, primary and bounce)",
                       agg["ms_primary_extend"] + agg["ms_extend_bounce"],
                       agg["primary_rays"] * bpr["primary"] + bounce_rays * bpr["bounce"], agg["extend_launches"]),
            "shadow": (f"{sh_name} (shadow rays)", agg["ms_shadow"], agg["shadow_rays"] * bpr["shadow"],
                       agg["shadow_launches"]),
        }
        pmc = None
        if not args.no_pmc and persistent:
            pmc = pmc_traffic(args.max_depth, args.loop)
        per_kernel = {}
        for kind, (kname, ms, nbytes, launches) in kernels.items():
            rate = nbytes / (ms / 1e3) / 1e9
            per_kernel[kind] = {"kernel": kname, "achieved": round(rate, 1), "frac": round(rate / HBM_PEAK_GBS, 4),
                                "launches_per_step": round(launches / args.steps, 2),
                                "avg_launch_ms": round(ms / max(launches, 1), 4),
                                "algorithmic_bytes_per_launch": round(nbytes / max(launches, 1)),
                                "traffic": round(pmc[kind]) if pmc is not None else None}
        dom = per_kernel["extend"]
        result["roofline"] = {"bound": "hbm", "achieved": dom["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": dom["frac"], "traffic": dom["traffic"], "kernel": dom["kernel"],
                              "avg_launch_ms": dom["avg_launch_ms"],
                              "algorithmic_bytes_per_launch": dom["algorithmic_bytes_per_launch"],
                              "bytes_per_ray": {k: round(v, 1) for k, v in bpr.items()},
                              "per_kernel": per_kernel}
        if pmc is not None:
            result["roofline"]["hbm_bytes_per_launch"] = {k: round(v) for k, v in pmc.items() if k != "scale"}
            result["roofline"]["pmc_scale"] = pmc["scale"]
        # the bounce extension and shadow launches share the GPU (concurrent
        # streams): their algorithmic bytes over the wall time they take together
        inc_bytes = ((agg["extension_rays"] - agg["primary_rays"]) * bpr["bounce"] + agg["shadow_rays"] * bpr["shadow"])
        inc_rate = inc_bytes / (agg["ms_incoherent"] / 1e3) / 1e9
        result["roofline"]["incoherent_stages"] = {"achieved": round(inc_rate, 1), "unit": "GB/s",
                                                   "frac": round(inc_rate / HBM_PEAK_GBS, 4),
                                                   "ms_per_step": round(agg["ms_incoherent"] / args.steps, 3)}
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())

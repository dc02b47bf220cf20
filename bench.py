"""Benchmark of the MI355X surface builder + span rasteriser (Square-Wheel's per-frame pixel path).

Default workload (one GPU): the plain drop-in path at 1920x1080, Color32, directional lightmaps,
on a seeded 32x32x4 grid map (p_solid = 0.35, 32 point lights + sun, 256x256 textures with normal
maps and mips, a third of the materials specular). A step renders one frame of a 64-pose seeded
camera path: surface build + draw setup + binning + raster, with the frame's command buffers already
resident in HBM (the CPU front-end that produces them is not timed, as on the CPU side).

Also measured in the same run:
  - p50 / p99 frame latency of the plain path,
  - the batched form: 256 views of log-uniform size 160x120 .. 1920x1080 (seeded poses) rendered by
    one captured HIP graph per batch, reported as aggregate Gpixel/s,
  - per-kernel durations with HIP events on the context's stream, for the roofline of the dominant
    kernel (algorithmic bytes / average launch duration vs 8 TB/s HBM peak),
  - the CPU baseline: the C++ restatement of the reference's CPU renderer (oracle/), multithreaded
    with the reference's decomposition, timed on this host's cores on a bounded sample.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "square-wheel_amd"))

import numpy as np  # noqa: E402

import sw_abi  # noqa: E402
import sw_host  # noqa: E402

METRIC = "1080p frames/s + batched Gpixel/s vs plain (HBM GB/s); p99 frame latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(stats, color_bits, e_bytes):
    """Algorithmic bytes the path must move per frame, split per kernel: a surface texel reads its
    8-byte texture element and writes a c-byte surface texel, each lightmap element of the rows a
    surface touches is read once (40 B directional, 12 B simple), an emissive texel reads 4 B and
    read-modify-writes the surface; a polygon pixel reads a c-byte surface texel and writes the
    framebuffer, a triangle pixel reads a 4-byte lite texel and writes, a blending write also reads
    the destination (c = 4 for Color32, 8 for Color64)."""
    c = 8 if color_bits == 64 else 4
    surf = stats["surface_texels"] * (8 + c) + stats["lightmap_elements"] * e_bytes + stats["emissive_texels"] * (4 + c)
    raster = stats["polygon_pixels"] * (c + c) + stats["triangle_pixels"] * (4 + c) + stats["blend_pixels"] * c
    return {"build_surfaces": surf, "raster_tiles": raster, "setup_and_surfaces": surf}


def read_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` in `workload` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by its reading for a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json.txt")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            data = json.load(f)
        v = data.get(workload, {}).get(kernel)
        return None if v is None else int(v["traffic_bytes"])
    except (OSError, ValueError, KeyError, TypeError):
        return None


def cpu_model_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(m, batches, seconds, threads):
    """The C++ restatement of the reference's CPU renderer (oracle/), timed frame by frame on this
    host's cores with the reference's decomposition (surfaces in parallel, the screen split into
    `threads` rects). Returns (frames/s, frames, seconds, per-frame latencies in s)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the CPU restatement; timed as the baseline only

    oracle.load()
    lat = []
    t0 = time.perf_counter()
    while True:
        b = batches[len(lat) % len(batches)]
        ts = time.perf_counter()
        oracle.render(m.data, b, mode=oracle.MODE_SSE, num_threads=threads)
        lat.append(time.perf_counter() - ts)
        el = time.perf_counter() - t0
        if el >= seconds or len(lat) >= 4096:
            break
    return len(lat) / el, len(lat), el, np.array(lat)


def dropin_leg(ctx, m, mh, frames, args):
    """The reference's own call shape, Renderer::draw_frame writing into the caller's &mut [ColorT]
    (renderer.rs:106-124): per frame sw_frame_begin, sw_build_surfaces and sw_raster_draws with the
    command buffers in host memory (copied to the device by the calls), then sw_frame_end into a
    pinned host framebuffer. Serial: one context, each frame finished before the next starts.
    Overlapped: two contexts (two HIP streams) take alternate frames, so frame N's copy back to the
    host overlaps frame N+1's uploads and kernels; each frame is still complete in host memory when
    its sw_frame_end returns."""
    import ctypes as C
    W, H = args.width, args.height
    lib = ctx.lib
    ctx2 = sw_abi.Context(0)
    mh2 = ctx2.upload_map(m.data)
    ctxs, mhs = [ctx, ctx2], [mh, mh2]
    fb_bytes = W * H * 4
    pinned = []
    for c in ctxs:
        p = C.c_void_p()
        c.check(lib.sw_host_alloc(c.h, fb_bytes, C.byref(p)), "sw_host_alloc")
        pinned.append(p)
    prep = []
    for f in frames:
        v = f["view"]
        arrs = [np.ascontiguousarray(f[k]) for k in ("surfaces", "lights", "light_indices", "draws", "vertex_data")]
        cam = (C.c_float * 3)(*[float(x) for x in np.asarray(v["cam_pos"]).reshape(-1)[:3]])
        prep.append((cam, int(np.asarray(v["clear_background"]).item()), int(np.asarray(v["background_color"]).item()), arrs))

    def draw(c, i):
        cam, clear, bgc, (s, li, lx, d, vd) = prep[i % len(prep)]
        h = ctxs[c].h
        ctxs[c].check(lib.sw_frame_begin(h, mhs[c], 32, W, H, cam, clear, bgc), "sw_frame_begin")
        ctxs[c].check(lib.sw_build_surfaces(h, C.c_void_p(s.ctypes.data), len(s), C.c_void_p(li.ctypes.data), len(li),
                                            C.c_void_p(lx.ctypes.data), len(lx)), "sw_build_surfaces")
        ctxs[c].check(lib.sw_raster_draws(h, C.c_void_p(d.ctypes.data), len(d), C.c_void_p(vd.ctypes.data), len(vd)),
                      "sw_raster_draws")

    def end(c):
        ctxs[c].check(lib.sw_frame_end(ctxs[c].h, pinned[c], W * 4, 0), "sw_frame_end")

    n = max(args.steps, len(frames))
    out = {}
    for i in range(len(frames)):
        draw(0, i)
        end(0)
    lat = np.zeros(n)
    t0 = time.perf_counter()
    for i in range(n):
        ts = time.perf_counter()
        draw(0, i)
        end(0)
        lat[i] = time.perf_counter() - ts
    wall = time.perf_counter() - t0
    out["serial"] = {"frames_s": round(n / wall, 2), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 4),
                     "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 4)}
    # overlapped: frame i is drawn on context i % 2 while frame i - 1 finishes on the other one
    lat = np.zeros(n)
    starts = np.zeros(n)
    t0 = time.perf_counter()
    for i in range(n):
        starts[i] = time.perf_counter()
        draw(i % 2, i)
        if i > 0:
            end((i - 1) % 2)
            lat[i - 1] = time.perf_counter() - starts[i - 1]
    end((n - 1) % 2)
    lat[n - 1] = time.perf_counter() - starts[n - 1]
    wall = time.perf_counter() - t0
    out["overlapped"] = {"frames_s": round(n / wall, 2), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 4),
                         "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 4), "contexts": 2}
    out["workload"] = (f"{W}x{H} Color32 frames of the plain path through sw_frame_begin / sw_build_surfaces / "
                       "sw_raster_draws (host-resident command buffers) / sw_frame_end into pinned host memory")
    out["frame_bytes_to_host"] = fb_bytes
    for c, p in zip(ctxs, pinned):
        lib.sw_host_free(c.h, p)
    ctx2.release_map(mh2)
    ctx2.close()
    log(f"drop-in: {out}")
    return out


def hdr_leg(ctx, args):
    """The HDR + models frame: 2560x1440 Color64 with 16 dynamic lights, 64 animated triangle
    models, 256 decals and 10% alpha-blended materials, rendered from HBM-resident command buffers
    and tonemapped on the device into a Color32 frame (sw_render_batch + sw_tonemap_device)."""
    import ctypes as C
    W, H = 2560, 1440
    m = sw_host.large_map_alpha(args.seed)
    mh = ctx.upload_map(m.data)
    fbld = sw_host.FrameBuilder()
    frames = sw_host.hdr_models_frames(m, args.seed, args.hdr_poses, W, H, fbld)
    batches = [sw_host.batch_from_frames([f]) for f in frames]
    dev = [sw_abi.DeviceCommandBuffers(ctx, b) for b in batches]
    hdr = ctx.alloc(W * H * 8)
    ldr = ctx.alloc(W * H * 4)
    cbs = [b.ctypes_batch(inputs_on_device=True, device_ptrs=d.ptrs()) for b, d in zip(batches, dev)]
    mod = (C.c_float * 3)(1.0, 1.0, 1.0)
    sums = (C.c_float * 3)()
    lib = ctx.lib

    def step(i):
        ctx.check(lib.sw_render_batch(ctx.h, mh, C.byref(cbs[i % len(cbs)]), hdr.ptr), "sw_render_batch")
        ctx.check(lib.sw_tonemap_device(ctx.h, hdr.ptr, W, H, C.c_float(1.0), mod, ldr.ptr, W * 4, sums), "sw_tonemap_device")

    tris = [int(np.sum(f["draws"]["kind"] == sw_abi.DRAW_TRIANGLE)) for f in frames]
    rows, tune = ctx.autotune_many(mh, batches[::4], dev[::4], hdr)
    log(f"autotune (hdr): {rows} rows, scalar spans {ctx.scalar_spans()}; ms per frame by rows {tune}")
    for i in range(4 * len(cbs)):
        step(i)
    ctx.profile_read(reset=True)
    lat = np.zeros(args.hdr_steps)
    t0 = time.perf_counter()
    for i in range(args.hdr_steps):
        sample = i % 8 == 0
        if sample:
            ctx.profile_enable(True)
        ts = time.perf_counter()
        step(i)
        lat[i] = time.perf_counter() - ts
        if sample:
            ctx.profile_enable(False)
    wall = time.perf_counter() - t0
    prof = ctx.profile_read(reset=True)

    # the same frames postprocessed with bloom (sw_postprocess_device): the reference's default
    # settings (sigma 1, half-resolution bloom buffer, scale 0.25, nearest upsampling) and a wide
    # linear-filtered bloom (sigma 8)
    bloom_out = {}
    avg = (C.c_float * 3)()
    for name, bp in (("default", sw_abi.BloomParams(1.0, 1, 0.25, 0)), ("sigma8_linear", sw_abi.BloomParams(8.0, 1, 0.25, 1))):
        def bstep(i):
            ctx.check(lib.sw_render_batch(ctx.h, mh, C.byref(cbs[i % len(cbs)]), hdr.ptr), "sw_render_batch")
            ctx.check(lib.sw_postprocess_device(ctx.h, hdr.ptr, W, H, C.c_float(1.0), mod, C.byref(bp), ldr.ptr, W * 4, avg),
                      "sw_postprocess_device")
        for i in range(4 * len(cbs)):
            bstep(i)
        ctx.profile_read(reset=True)
        blat = np.zeros(args.hdr_steps)
        tb = time.perf_counter()
        for i in range(args.hdr_steps):
            sample = i % 8 == 0
            if sample:
                ctx.profile_enable(True)
            ts = time.perf_counter()
            bstep(i)
            blat[i] = time.perf_counter() - ts
            if sample:
                ctx.profile_enable(False)
        bwall = time.perf_counter() - tb
        bprof = ctx.profile_read(reset=True)
        bus = bprof["bloom"][0] / max(bprof["bloom"][1], 1) * 1e3
        # the postprocess's algorithmic bytes: the HDR frame read, the Color32 frame written
        bloom_out[name] = {"frames_s": round(args.hdr_steps / bwall, 2),
                           "p50_ms": round(float(np.percentile(blat, 50)) * 1e3, 4),
                           "p99_ms": round(float(np.percentile(blat, 99)) * 1e3, 4),
                           "postprocess_us": round(bus, 3),
                           "postprocess_gb_s": round(W * H * 12 / (bus * 1e-6) / 1e9, 1)}
    for d in dev:
        d.free()
    hdr.free()
    ldr.free()
    ctx.release_map(mh)
    out = {"workload": f"{W}x{H} Color64 + tonemap, 16 dynamic lights, 64 models, 256 decals, 10% alpha materials, "
                       f"{len(frames)}-pose path",
           "frames_s": round(args.hdr_steps / wall, 2), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 4),
           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 4),
           "triangle_draws_per_frame": round(float(np.mean(tris)), 1),
           "kernel_us_per_launch": {k: round(v[0] / max(v[1], 1) * 1e3, 3) for k, v in prof.items() if v[1]},
           "tile_rows": rows, "scalar_spans": ctx.scalar_spans(), "with_bloom": bloom_out}
    log(f"hdr + models: {out}")
    return out


def shadows_leg(ctx, m, mh, args):
    """Shadowed frames: 1920x1080 Color32 with 8 dynamic lights, each with a 256x256 cube shadow map
    drawn every frame from the front-end's draw lists (swh_shadow_add_cube) by sw_render_depth, then
    the frame rendered with the maps sampled by the surface builder (sw_set_light_shadows +
    sw_render_batch). Draw lists and command buffers are built before the timed region (front-end
    CPU work); the timed step is the depth pass plus the frame."""
    import ctypes as C
    W, H, size, nl = args.width, args.height, 256, 8
    fbld = sw_host.FrameBuilder()
    poses = m.camera_path(args.seed + 7, 8)
    steps = []
    for i, p in enumerate(poses):
        f = fbld.build(m, p, width=W, height=H, color_bits=32, use_directional_lightmaps=1, num_dynamic_lights=nl,
                       seed=args.seed + 40 + i, clear_background=1, force_clear=1)
        b = sw_host.batch_from_frames([f])
        targets, draws, vdata, n, sh = sw_host.cube_shadows_for_lights(m, b.lights, size)
        steps.append((b, sw_abi.DeviceCommandBuffers(ctx, b), targets, draws, vdata, n, sh))
    n_max = max(st[5] for st in steps)
    depth = ctx.alloc(max(n_max, 1) * 4)
    out = ctx.alloc(W * H * 4)
    cbs = [st[0].ctypes_batch(inputs_on_device=True, device_ptrs=st[1].ptrs()) for st in steps]
    lib = ctx.lib

    def step(i):
        b, _, tg, dr, vd, n, sh = steps[i % len(steps)]
        ctx.check(lib.sw_render_depth(ctx.h, C.c_void_p(tg.ctypes.data), len(tg), C.c_void_p(dr.ctypes.data), len(dr),
                                      C.c_void_p(vd.ctypes.data), len(vd), depth.ptr, n), "sw_render_depth")
        ctx.check(lib.sw_set_light_shadows(ctx.h, C.c_void_p(sh.ctypes.data), len(sh), depth.ptr, n), "sw_set_light_shadows")
        ctx.check(lib.sw_render_batch(ctx.h, mh, C.byref(cbs[i % len(cbs)]), out.ptr), "sw_render_batch")

    # warm-up: eight passes of the path, so one-time runtime work (pools growing on first use) stays
    # out of the timed steps
    for i in range(8 * len(steps)):
        step(i)
    ctx.check(lib.sw_sync(ctx.h), "sw_sync")
    reps = 256
    ctx.profile_read(reset=True)
    lat = np.zeros(reps)
    t0 = time.perf_counter()
    for i in range(reps):
        sample = i % 8 == 0
        if sample:
            ctx.profile_enable(True)
        ts = time.perf_counter()
        step(i)
        lat[i] = time.perf_counter() - ts
        if sample:
            ctx.profile_enable(False)
    ctx.check(lib.sw_sync(ctx.h), "sw_sync")
    wall = time.perf_counter() - t0
    prof = ctx.profile_read(reset=True)
    ctx.check(lib.sw_set_light_shadows(ctx.h, None, 0, None, 0), "sw_set_light_shadows")
    dus = prof["depth"][0] / max(prof["depth"][1], 1) * 1e3
    texels = float(np.mean([st[5] for st in steps]))
    draws = float(np.mean([len(st[3]) for st in steps]))
    for st in steps:
        st[1].free()
    depth.free()
    out.free()
    res = {"workload": f"{W}x{H} Color32, {nl} dynamic lights with {size}x{size} cube shadow maps drawn per frame, "
                       f"{len(steps)}-pose path",
           "frames_s": round(reps / wall, 2), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 4),
           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 4),
           "depth_draws_per_frame": round(draws, 1), "depth_texels_per_frame": int(texels),
           "depth_pass_us": round(dus, 2),
           # algorithmic bytes of the depth pass: one 4-byte write per depth texel (a closed map covers them all)
           "depth_gtexel_s": round(texels / (dus * 1e-6) / 1e9, 2),
           "depth_gb_s": round(texels * 4 / (dus * 1e-6) / 1e9, 1),
           "kernel_us_per_launch": {k: round(v[0] / max(v[1], 1) * 1e3, 3) for k, v in prof.items() if v[1]}}
    log(f"shadows: {res}")
    return res


def water_leg(ctx, args):
    """Water textures: a 1024x1024 field (the largest the effect allows is 2^21 elements) stepped
    four times per frame with 16 wave-source points per step (sw_water_step), then the wavy texture
    with source-colour deformation and its four mips (sw_water_make_texture), into a map texture."""
    import ctypes as C
    lw = lh = 10
    n = 1 << (lw + lh)
    m = sw_host.Map(seed=args.seed + 9, grid=(10, 10, 4), texture_size=1024, num_materials=2, num_decal_textures=1,
                    num_model_textures=1)
    lv = m.data.levels
    lv["flags"][0:5] |= 1
    mh = ctx.upload_map(m.data)
    rng = np.random.default_rng(args.seed + 11)
    a = ctx.to_device(rng.normal(size=n).astype(np.float32) * np.float32(0.1))
    b = ctx.to_device(np.zeros(n, np.float32))
    pts = np.zeros(16, dtype=sw_abi.WAVE_POINT_DTYPE)
    pts["x"], pts["y"], pts["value"] = rng.integers(0, 1024, 16), rng.integers(0, 1024, 16), 1.0
    params = sw_abi.WaterTextureParams(0, 5, sw_abi.WATER_COLOR_DEFORMED, 0, 0.4, 0)
    lib = ctx.lib
    att = C.c_float(1.0 - 1.0 / 100.0)

    def frame():
        nonlocal a, b
        for _ in range(4):
            ctx.check(lib.sw_water_step(ctx.h, a.ptr, b.ptr, lw, lh, att, C.c_void_p(pts.ctypes.data), len(pts)), "sw_water_step")
            a, b = b, a
        ctx.check(lib.sw_water_make_texture(ctx.h, mh, a.ptr, lw, lh, C.byref(params)), "sw_water_make_texture")

    for _ in range(8):
        frame()
    ctx.check(lib.sw_sync(ctx.h), "sw_sync")
    reps = 100
    ctx.profile_enable(True)
    ctx.profile_read(reset=True)
    t0 = time.perf_counter()
    for _ in range(reps):
        frame()
    ctx.check(lib.sw_sync(ctx.h), "sw_sync")
    wall = (time.perf_counter() - t0) / reps
    ctx.profile_enable(False)
    prof = ctx.profile_read(reset=True)
    launches = prof["textures"][1]
    kus = prof["textures"][0] / max(reps, 1) * 1e3  # event-timed device time per frame (4 steps + texture)
    a.free()
    b.free()
    ctx.release_map(mh)
    fire = fire_leg(ctx, args)
    # algorithmic bytes per frame: each step reads and writes the two fields (4 + 4 + 4 B per element),
    # the texture reads the field and the colour level and writes 8 B texels, the mips 1/3 more
    step_bytes = 12 * n
    tex_bytes = (4 + 8 + 8) * n + (8 + 8) * n // 3
    res = {"workload": "1024x1024 water field, 4 steps x 16 source points + wavy texture (deformed colours) + 4 mips per frame",
           "us_per_frame": round(wall * 1e6, 2), "device_us_per_frame": round(kus, 2), "event_launches": int(launches),
           "gb_s": round((4 * step_bytes + tex_bytes) / (kus * 1e-6) / 1e9, 1), "fire": fire}
    log(f"water: {res}")
    return res


def fire_leg(ctx, args):
    """Fire textures: a 64x64 heat map (bytes; the map's emissive textures are 64 wide), 4 steps per
    frame with 64 set_heat points each, then
    the palette texture and its four lite mips (sw_fire_step, sw_fire_make_texture), both update
    variants."""
    import ctypes as C
    lw = lh = 6
    n = 1 << (lw + lh)
    m = sw_host.Map(seed=args.seed + 13, grid=(10, 10, 4), texture_size=1024, num_materials=2, num_decal_textures=1,
                    num_model_textures=1)
    lv = m.data.levels
    dst = next(i for i in range(len(lv) - 4) if (lv[i]["flags"] & 256) and lv[i]["width"] == 64 and lv[i]["height"] == 64)
    mh = ctx.upload_map(m.data)
    rng = np.random.default_rng(args.seed + 15)
    a = ctx.to_device(rng.integers(0, 256, n).astype(np.uint8))
    b = ctx.alloc(n)
    pts = np.zeros(64, dtype=sw_abi.HEAT_POINT_DTYPE)
    pts["x"], pts["y"], pts["heat"] = rng.integers(0, 64, 64), 62, 255
    palette = (np.arange(256, dtype=np.uint32) * 0x010101).astype(np.uint32)
    lib = ctx.lib
    out = {}
    for slow in (0, 1):
        def frame():
            nonlocal a, b
            for _ in range(4):
                ctx.check(lib.sw_fire_step(ctx.h, a.ptr, b.ptr, lw, lh, C.c_float(0.99), slow, C.c_void_p(pts.ctypes.data), len(pts)),
                          "sw_fire_step")
                a, b = b, a
            ctx.check(lib.sw_fire_make_texture(ctx.h, mh, a.ptr, lw, lh, dst, C.c_void_p(palette.ctypes.data)), "sw_fire_make_texture")

        for _ in range(8):
            frame()
        ctx.check(lib.sw_sync(ctx.h), "sw_sync")
        reps = 100
        ctx.profile_enable(True)
        ctx.profile_read(reset=True)
        t0 = time.perf_counter()
        for _ in range(reps):
            frame()
        ctx.check(lib.sw_sync(ctx.h), "sw_sync")
        wall = (time.perf_counter() - t0) / reps
        ctx.profile_enable(False)
        prof = ctx.profile_read(reset=True)
        out["slow" if slow else "fast"] = {"us_per_frame": round(wall * 1e6, 2),
                                           "device_us_per_frame": round(prof["textures"][0] / reps * 1e3, 2)}
    a.free()
    b.free()
    ctx.release_map(mh)
    out["workload"] = "64x64 heat map, 4 steps x 64 points + palette texture + 4 lite mips per frame"
    return out


def models_leg(ctx, args):
    """Triangle-model preparation (sw_prepare_models): 64 instances of a 1280-triangle, 642-vertex
    vertex-animated mesh (8 frames), meshes resident in HBM, outputs in HBM; the kernel's duration
    from HIP events, the oracle's single-thread CPU time on the same instances beside it."""
    import ctypes as C
    import sw_model_scenes as MS
    rng = np.random.default_rng(args.seed + 31)
    meshes = [MS.vertex_animated_mesh(rng, subdiv=3, frames=8)]
    inst = np.ascontiguousarray(MS.instances(rng, meshes, 64))
    arr, bufs = ctx.upload_meshes(meshes)
    nv = len(inst) * meshes[0].num_vertices
    nt = len(inst) * len(meshes[0].triangles)
    vbuf, tbuf, cbuf = ctx.alloc(nv * 32), ctx.alloc(nt * 6), ctx.alloc(len(inst) * 4)
    lib = ctx.lib

    def call():
        ctx.check(lib.sw_prepare_models(ctx.h, arr, len(meshes), C.c_void_p(inst.ctypes.data), len(inst), vbuf.ptr, nv,
                                        tbuf.ptr, nt, cbuf.ptr), "sw_prepare_models")

    for _ in range(5):
        call()
    ctx.check(lib.sw_sync(ctx.h), "sw_sync")
    reps = 200
    ctx.profile_enable(True)
    ctx.profile_read(reset=True)
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    ctx.check(lib.sw_sync(ctx.h), "sw_sync")
    wall = (time.perf_counter() - t0) / reps
    ctx.profile_enable(False)
    prof = ctx.profile_read(reset=True)
    kus = prof["models"][0] / max(prof["models"][1], 1) * 1e3
    visible = int(cbuf.read(np.uint32, len(inst)).sum())
    for b in bufs + [vbuf, tbuf, cbuf]:
        b.free()
    out = {"workload": f"{len(inst)} instances x {len(meshes[0].triangles)}-triangle vertex-animated mesh (lerped frames, "
                       "per-vertex light cube + 2 directional lights, back faces rejected, back-to-front sort)",
           "kernel_us": round(kus, 2), "us_per_call": round(wall * 1e6, 2), "visible_triangles": visible,
           "mtriangles_s": round(nt / (kus * 1e-6) / 1e6, 1)}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 2.0:
            oracle.prepare_models(meshes, inst)
            n += 1
        cms = (time.perf_counter() - t0) / n * 1e3
        out.update(cpu_oracle_ms=round(cms, 3), cpu_cores=1, speedup_vs_cpu=round(cms * 1e3 / kus, 1))
    log(f"models: {out}")
    return out


def device_models_leg(ctx, args):
    """The HDR + models frames with their models prepared on the device (swh_frame_config.device_models):
    the front-end's CPU time per frame with and without its own model preparation, and the frame
    through sw_prepare_models + sw_expand_mesh_runs (which waits for the fan sizes) + the draw list
    completed on the host + sw_render_batch, against the same frames with host-prepared models
    (both with host-resident command buffers, uploaded per frame)."""
    import ctypes as C
    W, H = 2560, 1440
    m = sw_host.large_map_alpha(args.seed)
    mh = ctx.upload_map(m.data)
    fbld = sw_host.FrameBuilder()
    poses = m.camera_path(args.seed + 3, 16)
    kw = dict(width=W, height=H, color_bits=64, use_directional_lightmaps=1, num_dynamic_lights=16, num_models=64,
              num_decals=256, seed=args.seed + 5, clear_background=1, force_clear=1)
    fe = {}
    frames = {}
    for dev in (0, 1):
        t0 = time.perf_counter()
        frames[dev] = [fbld.build(m, p, time=0.1 * i, device_models=dev, **kw) for i, p in enumerate(poses)]
        fe[dev] = (time.perf_counter() - t0) / len(poses) * 1e3
    dm = sw_host.DeviceModels(ctx, m)
    hdr = ctx.alloc(W * H * 8)
    host_batches = [sw_host.batch_from_frames([f]) for f in frames[0]]
    lib = ctx.lib
    cbs = [b.ctypes_batch() for b in host_batches]
    for i in range(len(poses)):
        dm.render(mh, frames[1][i], out=hdr)
        ctx.check(lib.sw_render_batch(ctx.h, mh, C.byref(cbs[i]), hdr.ptr), "sw_render_batch")
    reps = 4 * len(poses)
    ctx.profile_read(reset=True)
    ctx.profile_enable(True)
    t0 = time.perf_counter()
    for i in range(reps):
        dm.render(mh, frames[1][i % len(poses)], out=hdr)
    dev_ms = (time.perf_counter() - t0) / reps * 1e3
    ctx.profile_enable(False)
    prof = ctx.profile_read(reset=True)
    t0 = time.perf_counter()
    for i in range(reps):
        ctx.check(lib.sw_render_batch(ctx.h, mh, C.byref(cbs[i % len(poses)]), hdr.ptr), "sw_render_batch")
    host_ms = (time.perf_counter() - t0) / reps * 1e3
    runs = float(np.mean([len(f["models"]["runs"]) for f in frames[1]]))
    tris = float(np.mean([int(np.sum(f["draws"]["kind"] == sw_abi.DRAW_TRIANGLE)) for f in frames[0]]))
    dm.close()
    hdr.free()
    ctx.release_map(mh)
    out = {"workload": f"{W}x{H} Color64, 64 models (1280-triangle vertex-animated mesh), 256 decals, 16 lights, "
                       f"{len(poses)}-pose path; command buffers uploaded per frame",
           "front_end_ms_host_models": round(fe[0], 3), "front_end_ms_device_models": round(fe[1], 3),
           "frame_ms_device_models": round(dev_ms, 4), "frame_ms_host_models": round(host_ms, 4),
           "prepare_and_expand_us_per_frame": round(prof["models"][0] / reps * 1e3, 2),
           "mesh_runs_per_frame": round(runs, 1), "model_and_decal_triangle_draws_per_frame": round(tris, 1)}
    log(f"device models: {out}")
    return out


def serving_leg(args):
    """Open-loop serving: four resident maps (16x16x4 .. 48x48x4 grids), Poisson arrivals at fixed
    rates, each request a random map, pose, resolution (log-uniform 320x200 .. 2560x1440, 16:10) and
    colour depth (Color32, or Color64 tonemapped on the device and returned as the tonemapped image
    only). The server batches requests on the fly onto three concurrent slots under an HBM budget;
    results are copied by the device into pinned host buffers; latency is arrival -> result in host
    memory."""
    grids = sw_host.SERVING_GRIDS
    maps = sw_host.serving_maps(args.seed)
    pool = sw_host.serving_request_pool(maps, args.seed, per_map=48)
    rng = np.random.default_rng(args.seed + 78)  # arrivals and picks
    mean_px = float(np.mean([int(np.asarray(f["view"]["width"]).item()) * int(np.asarray(f["view"]["height"]).item()) for _, f in pool]))
    srv = sw_abi.Server(slots=3, hbm_budget=16 << 30, max_views=32)
    ids = [srv.add_map(m.data) for m in maps]
    out = {"workload": f"4 maps {grids}, {len(pool)} distinct requests, mean {mean_px / 1e6:.2f} Mpixel, "
                       "log-uniform 320x200..2560x1440, half Color64 + tonemap (tonemapped image returned), "
                       "results in pinned host memory", "rates": []}
    # warm-up: every request once
    def submit(mi, f):
        return srv.submit(ids[mi], f, tonemap_exposure=1.0 if f["color_bits"] == 64 else None, hdr=False, pinned=True)

    for mi, f in pool:
        submit(mi, f)
    for c in srv.drain():
        srv.release(c)
    for rate in (200, 500, 1000, 2000, 5000, 8000):
        n = max(int(rate * args.serve_seconds), 50)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
        picks = rng.integers(0, len(pool), size=n)
        lat = []
        views = []
        submitted = {}
        t0 = time.perf_counter()
        k = 0
        while k < n or srv.pending():
            now = time.perf_counter() - t0
            while k < n and arrivals[k] <= now:
                mi, f = pool[int(picks[k])]
                t = submit(mi, f)
                submitted[t] = arrivals[k]
                k += 1
            for c in srv.poll():
                done_at = time.perf_counter() - t0
                lat.append((done_at - submitted.pop(c["ticket"])) * 1e3)
                views.append(c["batch_views"])
                if c["status"] != 0:
                    raise RuntimeError(f"request failed: {c}")
                srv.release(c)
        wall = time.perf_counter() - t0
        lat = np.array(lat)
        row = {"offered_req_s": rate, "requests": n, "achieved_req_s": round(n / wall, 1),
               "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
               "mean_batch_views": round(float(np.mean(views)), 2),
               "gpixel_s": round(n * mean_px / wall / 1e9, 3)}
        out["rates"].append(row)
        log(f"serving {row}")
    # closed loop: a fixed number of requests in flight, each completion replaced by a new request;
    # the completion rate is the server's sustained throughput
    for depth in (16, 96):
        n = max(int(8000 * args.serve_seconds), 200)
        picks = rng.integers(0, len(pool), size=n)
        lat, views = [], []
        submitted = {}
        t0 = time.perf_counter()
        k = 0
        while k < n or srv.pending():
            while k < n and srv.pending() < depth:
                mi, f = pool[int(picks[k])]
                submitted[submit(mi, f)] = time.perf_counter()
                k += 1
            for c in srv.poll():
                lat.append((time.perf_counter() - submitted.pop(c["ticket"])) * 1e3)
                views.append(c["batch_views"])
                if c["status"] != 0:
                    raise RuntimeError(f"request failed: {c}")
                srv.release(c)
        wall = time.perf_counter() - t0
        row = {"in_flight": depth, "requests": n, "achieved_req_s": round(n / wall, 1),
               "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
               "mean_batch_views": round(float(np.mean(views)), 2), "gpixel_s": round(n * mean_px / wall / 1e9, 3)}
        out.setdefault("closed_loop", []).append(row)
        log(f"serving closed loop {row}")
    out["server_stats"] = srv.stats()
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=384)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--poses", type=int, default=64)
    ap.add_argument("--batch-views", type=int, default=256)
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--no-hdr", action="store_true")
    ap.add_argument("--no-models", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-shadows", action="store_true")
    ap.add_argument("--no-water", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--hdr-poses", type=int, default=16)
    ap.add_argument("--hdr-steps", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run device-to-device copies of this size (HBM counter calibration) and exit")
    args = ap.parse_args()

    if args.copy_calibration_mb:
        # A plain copy kernel of known size: rocprofv3's FETCH_SIZE / WRITE_SIZE of it against its
        # byte count gives the scale of those counters on this device.
        ctx = sw_abi.Context(0)
        n = args.copy_calibration_mb << 20
        a, b = ctx.alloc(n), ctx.alloc(n)
        for _ in range(3):
            ctx.copy_device(b, a, n)
        a.free()
        b.free()
        ctx.close()
        print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the renderer runs on a single GPU; "
              "views and requests would shard across GPUs with no collective", flush=True)
        sys.exit(2)

    color_bits = 32
    t = time.perf_counter()
    m = sw_host.large_map(args.seed)
    log(f"map built in {time.perf_counter() - t:.1f}s: {m.info}")
    frames = sw_host.plain_1080p_frames(m, args.seed, args.poses, args.width, args.height)
    batches = [sw_host.batch_from_frames([f]) for f in frames]
    log(f"{len(frames)} frames built; draws/frame {np.mean([len(f['draws']) for f in frames]):.0f}, "
        f"surface texels/frame {np.mean([f['arena_texels'] for f in frames]):.0f}")

    ctx = sw_abi.Context(0)
    mh = ctx.upload_map(m.data)
    dev = [sw_abi.DeviceCommandBuffers(ctx, b) for b in batches]
    out = ctx.alloc(args.width * args.height * 4)
    lib = ctx.lib
    import ctypes as C
    cbs = [b.ctypes_batch(inputs_on_device=True, device_ptrs=d.ptrs()) for b, d in zip(batches, dev)]

    def render(i):
        r = lib.sw_render_batch(ctx.h, mh, C.byref(cbs[i % len(cbs)]), out.ptr)
        if r != 0:
            raise RuntimeError(lib.sw_last_error(ctx.h).decode())

    # per-frame work counters (identical every pass of the path)
    per_frame_stats = []
    for i in range(len(frames)):
        render(i)
        st = ctx.stats()
        st["surface_texels"] = int(frames[i]["arena_texels"])
        s = frames[i]["surfaces"]
        st["lightmap_elements"] = int(np.sum(2 * s["lightmap_size"][:, 0].astype(np.int64) * s["size"][:, 1] *
                                             (s["lightmap_offset"] != 0)))
        st["emissive_texels"] = 0
        per_frame_stats.append(st)

    # tile height and span mode for this frame size, chosen by the autotuner on every fourth pose
    plain_rows, plain_tune = ctx.autotune_many(mh, batches[::4], dev[::4], out)
    log(f"autotune (plain): {plain_rows} rows, scalar spans {ctx.scalar_spans()}; ms per frame by rows {plain_tune}")
    for i in range(args.warmup):
        render(i)
    # Timed region: K frames of the path, each call returns with its frame finished.
    t0 = time.perf_counter()
    for i in range(args.steps):
        render(i)
    wall = time.perf_counter() - t0
    fps = args.steps / wall
    # Frame latency over a fixed sample (four passes of the path), whatever --steps is.
    lat_frames = 4 * len(frames)
    lat = np.zeros(lat_frames)
    for i in range(lat_frames):
        ts = time.perf_counter()
        render(i)
        lat[i] = time.perf_counter() - ts
    # Frames in flight: a second context (its own HIP stream, map copy, command buffers and
    # framebuffer) takes every other frame; frame i is launched before frame i - 1 is waited for, so
    # one frame's raster overlaps the next frame's setup and surfaces. Each frame is complete in HBM
    # when its wait returns.
    ctx2 = sw_abi.Context(0)
    mh2 = ctx2.upload_map(m.data)
    dev2 = [sw_abi.DeviceCommandBuffers(ctx2, b) for b in batches]
    out2 = ctx2.alloc(args.width * args.height * 4)
    cbs2 = [b.ctypes_batch(inputs_on_device=True, device_ptrs=d.ptrs()) for b, d in zip(batches, dev2)]
    ctx2.autotune_many(mh2, batches[::4], dev2[::4], out2)
    pipe = [(ctx, mh, cbs, out), (ctx2, mh2, cbs2, out2)]

    def launch(i):
        c, h, cb, o = pipe[i % 2]
        r = lib.sw_render_batch_async(c.h, h, C.byref(cb[i % len(cb)]), o.ptr)
        if r != 0:
            raise RuntimeError(lib.sw_last_error(c.h).decode())

    def wait(i):
        c = pipe[i % 2][0]
        r = lib.sw_batch_wait(c.h)
        if r != 0:
            raise RuntimeError(lib.sw_last_error(c.h).decode())

    for i in range(args.warmup):
        launch(i)
        if i > 0:
            wait(i - 1)
    wait(args.warmup - 1)
    starts = np.zeros(args.steps)
    plat = np.zeros(args.steps)
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i] = time.perf_counter()
        launch(i)
        if i > 0:
            wait(i - 1)
            plat[i - 1] = time.perf_counter() - starts[i - 1]
    wait(args.steps - 1)
    plat[-1] = time.perf_counter() - starts[-1]
    pwall = time.perf_counter() - t0
    pipelined_fps = args.steps / pwall
    log(f"plain, two frames in flight: {pipelined_fps:.1f} frames/s")
    for d in dev2:
        d.free()
    out2.free()
    ctx2.release_map(mh2)
    ctx2.close()

    # Per-kernel durations for the roofline: HIP events around the launches on the context's stream,
    # over one pass of the path (recording events costs host time, so the timed frames go without).
    ctx.profile_read(reset=True)
    ctx.profile_enable(True)
    for i in range(len(frames)):
        render(i)
    ctx.profile_enable(False)
    prof = ctx.profile_read(reset=True)
    log(f"plain: {fps:.1f} frames/s, p50 {np.percentile(lat, 50) * 1e3:.3f} ms, p99 {np.percentile(lat, 99) * 1e3:.3f} ms")
    log(f"kernels: { {k: (round(v[0] / max(v[1], 1) * 1e3, 2), v[1]) for k, v in prof.items()} } (us/launch, launches)")

    # roofline of the dominant kernel over the event-timed pass
    steps_stats = [per_frame_stats[i] for i in range(len(frames))]
    bytes_sum = {"build_surfaces": 0, "raster_tiles": 0, "setup_and_surfaces": 0}
    for st in steps_stats:
        for k, v in algorithmic_bytes(st, color_bits, 40).items():
            bytes_sum[k] += v
    kernel_ms = {k: prof[k][0] for k in ("build_surfaces", "raster_tiles", "setup_and_surfaces")}
    dominant = max(kernel_ms, key=kernel_ms.get)
    launches = max(prof[dominant][1], 1)
    avg_s = prof[dominant][0] / 1e3 / launches
    achieved = bytes_sum[dominant] / launches / avg_s / 1e9
    raster_key = f"k_raster_tiles_rows{plain_rows}" + ("_sspan" if ctx.scalar_spans() else "")
    traffic_key = {"raster_tiles": raster_key, "setup_and_surfaces": "k_setup_and_surfaces_256",
                   "build_surfaces": "k_build_surfaces"}.get(dominant)
    traffic = read_traffic("plain1080p", traffic_key) if traffic_key else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_source": "profiles/traffic.json.txt: rocprofv3 FETCH_SIZE + WRITE_SIZE runs of this workload, "
                                  "scaled by a copy kernel's reading" if traffic is not None else None,
                "kernel": dominant,
                "algorithmic_bytes_per_launch": int(bytes_sum[dominant] / launches),
                "avg_launch_us": round(avg_s * 1e6, 3)}

    result = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 surfaces / int32 spans, Color32 out",
        "data": "synthetic (seeded grid map, value-noise textures, generated lightmaps)",
        "config": {"workload": f"plain drop-in frame: {args.width}x{args.height} Color32, directional lightmaps, "
                               f"32x32x4 grid p=0.35, 32 lights + sun, 256^2 textures + normal maps, "
                               f"{args.poses}-pose camera path, command buffers resident in HBM"},
        "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 4),
        "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 4),
        "latency_frames": lat_frames,
        "plain_gpixel_s": round(fps * args.width * args.height / 1e9, 4),
        "two_frames_in_flight": {"frames_s": round(pipelined_fps, 2),
                                 "p50_ms": round(float(np.percentile(plat, 50)) * 1e3, 4),
                                 "p99_ms": round(float(np.percentile(plat, 99)) * 1e3, 4), "contexts": 2,
                                 "note": "frame i launched (sw_render_batch_async) before frame i - 1 is waited for, "
                                         "alternate frames on two contexts; value above is one frame at a time"},
        "kernel_us_per_launch": {k: round(v[0] / max(v[1], 1) * 1e3, 3) for k, v in prof.items()},
        "tile_rows": plain_rows,
        "event_sampled_frames": len(frames),
        "roofline": roofline,
    }

    if not args.no_dropin:
        result["dropin"] = dropin_leg(ctx, m, mh, frames, args)

    if not args.no_batch:
        bframes = sw_host.batch_view_frames(m, args.seed, args.batch_views, color_bits=color_bits)
        total_px = sum(int(np.asarray(f["view"]["width"]).item()) * int(np.asarray(f["view"]["height"]).item()) for f in bframes)
        bb = sw_host.batch_from_frames(bframes)
        bdev = sw_abi.DeviceCommandBuffers(ctx, bb)
        bout = ctx.alloc(bb.fb_texels * 4)
        bcb = bb.ctypes_batch(inputs_on_device=True, device_ptrs=bdev.ptrs())
        batch_rows, batch_tune = ctx.autotune(mh, bb, bdev, bout)
        log(f"autotune (batch): {batch_rows} rows; ms per batch by rows {batch_tune}")
        ctx.check(lib.sw_render_batch_graph(ctx.h, mh, C.byref(bcb), bout.ptr), "sw_render_batch_graph")
        ctx.profile_enable(True)
        ctx.profile_read(reset=True)
        tb = time.perf_counter()
        for _ in range(args.batch_reps):
            ctx.check(lib.sw_render_batch_graph(ctx.h, mh, C.byref(bcb), bout.ptr), "sw_render_batch_graph")
        bwall = (time.perf_counter() - tb) / args.batch_reps
        # per-kernel durations of the batch: HIP events around the launches of eager runs (a
        # captured graph's launches carry no events)
        ctx.profile_read(reset=True)
        eager_runs = 2
        for _ in range(eager_runs):
            ctx.check(lib.sw_render_batch(ctx.h, mh, C.byref(bcb), bout.ptr), "sw_render_batch")
        ctx.profile_enable(False)
        bprof = ctx.profile_read(reset=True)
        bst = ctx.stats()
        surf_texels = int(bb.arena_texels)
        s = bb.surfaces
        lm = int(np.sum(2 * s["lightmap_size"][:, 0].astype(np.int64) * s["size"][:, 1] * (s["lightmap_offset"] != 0)))
        bst.update(surface_texels=surf_texels, lightmap_elements=lm, emissive_texels=0)
        bbytes = sum(v for k, v in algorithmic_bytes(bst, color_bits, 40).items() if k != "setup_and_surfaces")
        result["batched"] = {"views": args.batch_views, "pixels": total_px, "ms_per_batch": round(bwall * 1e3, 3),
                             "gpixel_s": round(total_px / bwall / 1e9, 4),
                             "vs_plain_gpixel_s": round((total_px / bwall) / (fps * args.width * args.height), 3),
                             "achieved_gb_s": round(bbytes / bwall / 1e9, 2),
                             "hbm_frac": round(bbytes / bwall / 1e9 / HBM_PEAK_GBS, 5),
                             "surfaces": int(len(bb.surfaces)), "draws": int(len(bb.draws)), "tile_rows": batch_rows,
                             "kernel_launches_per_batch": {k: v[1] // eager_runs for k, v in bprof.items() if v[1]},
                             # summed kernel durations (a large batch's surface and raster chunks overlap
                             # on two streams, so these need not add up to the batch's time)
                             "kernel_us_per_batch": {k: round(v[0] / eager_runs * 1e3, 2) for k, v in bprof.items()}}
        log(f"batched: {result['batched']}")
        bdev.free()
        bout.free()

    if not args.no_shadows:
        result["shadows"] = shadows_leg(ctx, m, mh, args)

    if not args.no_hdr:
        result["hdr_models"] = hdr_leg(ctx, args)

    if not args.no_models:
        result["models"] = models_leg(ctx, args)
        result["device_models"] = device_models_leg(ctx, args)

    if not args.no_water:
        result["water"] = water_leg(ctx, args)

    if not args.no_serve:
        result["serving"] = serving_leg(args)

    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        threads = max(1, min(threads, 64))
        cfps, n, el, clat = cpu_baseline(m, batches, args.cpu_seconds, threads)
        result["cpu_baseline"] = {"value": round(cfps, 3), "unit": "frames/s", "cores": threads, "kind": "port",
                                  "p50_ms": round(float(np.percentile(clat, 50)) * 1e3, 3),
                                  "p99_ms": round(float(np.percentile(clat, 99)) * 1e3, 3),
                                  "cpu": cpu_model_name(),
                                  "sample": f"{n} frames of the same {args.poses}-pose 1080p path in {el:.1f}s "
                                            f"(surface build + raster, reference decomposition over {threads} threads, "
                                            "x86 rsqrtss/rcpss numerics, per-thread work counters)"}
        result["speedup_vs_cpu"] = round(fps / cfps, 2)
        if "dropin" in result:
            for k in ("serial", "overlapped"):
                result["dropin"][k]["speedup_vs_cpu"] = round(result["dropin"][k]["frames_s"] / cfps, 2)
        log(f"cpu baseline: {cfps:.2f} frames/s on {threads} threads")

    for d in dev:
        d.free()
    out.free()
    ctx.release_map(mh)
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

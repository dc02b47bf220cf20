"""GPU parity at the benchmark workloads' full sizes: the same seeded maps, poses, resolutions and
command buffers bench.py times, checked against the oracle (bit-exact against its exact mode,
within one LSB per channel of its x86 rsqrtss / rcpss mode) and, for the batched and serving
forms, against the plain path view by view.

  - the plain 1080p frame (32x32x4 map, directional lightmaps, normal maps, specular);
  - the HDR + models frame (2560x1440 Color64 with its tonemap, 16 lights, 64 models, 256 decals,
    10% alpha-blended materials);
  - the 256-view batch (160x90 .. 1920x1080), eager and replayed from its captured graph;
  - the serving workload (four maps of 16x16x4 .. 48x48x4, requests up to 2560x1440 in both
    colour depths), every result against its solo render.
The oracle runs single-threaded per view (its N-rect split changes the reference's output), views
in parallel host threads.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
import sw_abi
import sw_host

pytestmark = pytest.mark.gpu

POOL = 8


def _say(msg):
    print(msg, flush=True)


def _max_lsb(a, b, color_bits):
    ch = np.uint16 if color_bits == 64 else np.uint8
    return int(np.abs(a.view(ch).astype(np.int64) - b.view(ch).astype(np.int64)).max())


@pytest.fixture(scope="module")
def map_1080p():
    return sw_host.large_map(0)


def test_plain_1080p_frames(gpu_ctx, map_1080p):
    frames = sw_host.plain_1080p_frames(map_1080p, 0, 64)
    mh = gpu_ctx.upload_map(map_1080p.data)
    try:
        for k in (0, 21, 63):
            _say(f"case 1080p pose {k}")
            b = sw_host.batch_from_frames([frames[k]])
            img = gpu_ctx.render_plain(mh, b).reshape(-1)
            ref, _, st, oob = oracle.render(map_1080p.data, b)
            assert oob == 0 and st["polygon_pixels"] > 1920 * 1080
            assert np.array_equal(img, ref), f"pose {k}: {np.count_nonzero(img != ref)} pixels differ"
            sse, _, _, _ = oracle.render(map_1080p.data, b, mode=oracle.MODE_SSE)
            assert _max_lsb(img, sse, 32) <= 1
            # the same frame through the batched entry point with HBM-resident command buffers
            dev = sw_abi.DeviceCommandBuffers(gpu_ctx, b)
            try:
                out = gpu_ctx.render_batch(mh, b, device_inputs=dev)
            finally:
                dev.free()
            assert np.array_equal(out, ref)
    finally:
        gpu_ctx.release_map(mh)


def test_hdr_models_frames(gpu_ctx):
    m = sw_host.large_map_alpha(0)
    frames = sw_host.hdr_models_frames(m, 0, 16)
    mh = gpu_ctx.upload_map(m.data)
    try:
        for k in (0, 9):
            _say(f"case hdr+models 2560x1440 pose {k}")
            b = sw_host.batch_from_frames([frames[k]])
            img, out, sums = gpu_ctx.render_plain(mh, b, tonemap=(1.0, (1.0, 1.0, 1.0)))
            ref, _, st, oob = oracle.render(m.data, b)
            assert oob == 0 and st["triangle_pixels"] > 0 and st["blend_pixels"] > 0
            img = img.reshape(-1)
            assert np.array_equal(img, ref), f"pose {k}: {np.count_nonzero(img != ref)} pixels differ"
            tref, tsums = oracle.tonemap(ref.reshape(1440, 2560), 1.0, (1.0, 1.0, 1.0))
            assert np.array_equal(out, tref)
            sse, _, _, _ = oracle.render(m.data, b, mode=oracle.MODE_SSE)
            assert _max_lsb(img, sse, 64) <= 1
    finally:
        gpu_ctx.release_map(mh)


def _oracle_views(m_data, frames, mode=oracle.MODE_EXACT):
    def one(f):
        ref, _, _, oob = oracle.render(m_data, sw_host.batch_from_frames([f]), mode=mode)
        assert oob == 0
        return ref
    with ThreadPoolExecutor(POOL) as ex:
        return list(ex.map(one, frames))


def test_batch_256_views(gpu_ctx, map_1080p):
    frames = sw_host.batch_view_frames(map_1080p, 0, 256)
    b = sw_host.batch_from_frames(frames)
    px = sum(int(v["width"]) * int(v["height"]) for v in b.views)
    _say(f"case batch 256 views, {px} pixels, {len(b.surfaces)} surfaces, {len(b.draws)} draws")
    assert px > 10 ** 8
    refs = _oracle_views(map_1080p.data, frames)
    mh = gpu_ctx.upload_map(map_1080p.data)
    dev = sw_abi.DeviceCommandBuffers(gpu_ctx, b)
    try:
        eager = gpu_ctx.render_batch(mh, b, device_inputs=dev)
        graph = [gpu_ctx.render_batch(mh, b, graph=True, device_inputs=dev) for _ in range(3)]
        for vi, v in enumerate(b.views):
            off, n = int(v["fb_offset"]), int(v["width"]) * int(v["height"])
            for name, out in [("eager", eager)] + [(f"graph replay {r}", g) for r, g in enumerate(graph)]:
                assert np.array_equal(out[off:off + n], refs[vi]), f"view {vi} ({name}) differs from the oracle"
        # view by view through the plain path (the drop-in call sequence) on a sample of the views
        for vi in range(0, 256, 16):
            _say(f"case batch view {vi} plain")
            one = sw_host.batch_from_frames([frames[vi]])
            plain = gpu_ctx.render_plain(mh, one).reshape(-1)
            off, n = int(b.views[vi]["fb_offset"]), plain.size
            assert np.array_equal(plain, eager[off:off + n]), f"view {vi}: batched differs from plain"
        # the x86-mode oracle within one LSB on a sample of the views
        sse = _oracle_views(map_1080p.data, frames[::32], mode=oracle.MODE_SSE)
        for j, vi in enumerate(range(0, 256, 32)):
            off, n = int(b.views[vi]["fb_offset"]), sse[j].size
            assert _max_lsb(eager[off:off + n], sse[j], 32) <= 1
    finally:
        dev.free()
        gpu_ctx.release_map(mh)


def test_serving_workload(gpu_ctx):
    maps = sw_host.serving_maps(0)
    pool = sw_host.serving_request_pool(maps, 0)
    _say(f"case serving: {len(pool)} requests over maps {sw_host.SERVING_GRIDS}")
    srv = sw_abi.Server(slots=3, hbm_budget=16 << 30, max_views=32)
    handles = []
    try:
        ids = [srv.add_map(m.data) for m in maps]
        order = np.random.default_rng(5).permutation(len(pool))
        tickets = {}
        for k in order:
            mi, f = pool[int(k)]
            t = srv.submit(ids[mi], f, tonemap_exposure=1.0 if f["color_bits"] == 64 else None, pinned=True, hdr=True)
            tickets[t] = int(k)
        done = srv.drain()
        assert len(done) == len(pool)
        assert srv.stats()["batches"] < len(pool)
        handles = [gpu_ctx.upload_map(m.data) for m in maps]
        checked = 0
        for c in done:
            assert c["status"] == 0, c
            mi, f = pool[tickets[c["ticket"]]]
            b = sw_host.batch_from_frames([f])
            alone = gpu_ctx.render_plain(handles[mi], b)
            assert np.array_equal(c["framebuffer"], alone), f"request {tickets[c['ticket']]} differs from its solo render"
            if f["color_bits"] == 64:
                tref, _ = oracle.tonemap(alone, 1.0, (1.0, 1.0, 1.0))
                assert np.array_equal(c["tonemapped"], tref)
            if tickets[c["ticket"]] % 6 == 0:
                ref, _, _, oob = oracle.render(maps[mi].data, b)
                assert oob == 0 and np.array_equal(c["framebuffer"].reshape(-1), ref)
                checked += 1
            srv.release(c)
        assert checked >= 6
    finally:
        for h in handles:
            gpu_ctx.release_map(h)
        srv.close()

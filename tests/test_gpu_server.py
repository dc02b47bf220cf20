"""The frame-serving scheduler (include/sw_server.h) on the GPU: requests of several maps, sizes and
colour depths are batched on the fly onto concurrent slots; every result equals the same frame
rendered alone through the plain path (and the oracle), the HBM budget gates admission."""
import numpy as np
import pytest

import oracle
import sw_abi
import sw_host

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def maps():
    return [sw_host.large_map(21, grid=(10, 10, 4)), sw_host.large_map_alpha(22, grid=(12, 10, 4))]


def _requests(maps, n, seed):
    rng = np.random.default_rng(seed)
    fb = sw_host.FrameBuilder()
    reqs = []
    paths = [m.camera_path(seed + i, 16) for i, m in enumerate(maps)]
    for i in range(n):
        mi = int(rng.integers(0, len(maps)))
        pose = paths[mi][int(rng.integers(0, 16))]
        w = int(rng.integers(96, 480))
        h = int(rng.integers(64, 320))
        bits = 64 if rng.uniform() < 0.4 else 32
        f = fb.build(maps[mi], pose, width=w, height=h, color_bits=bits, use_directional_lightmaps=1,
                     num_dynamic_lights=int(rng.integers(0, 8)), num_models=int(rng.integers(0, 12)),
                     num_decals=int(rng.integers(0, 40)), seed=int(rng.integers(0, 1000)), time=float(rng.uniform(0, 3)))
        reqs.append((mi, f))
    return reqs


@pytest.mark.parametrize("pinned", [False, True])
def test_mixed_requests_equal_alone(gpu_ctx, maps, pinned):
    """pinned: result buffers in pinned host memory (copied into directly by the device), and every
    third 64-bit request returns only its tonemapped image."""
    _say(f"case server mix pinned={pinned}")
    reqs = _requests(maps, 28, 5)
    srv = sw_abi.Server(slots=3, hbm_budget=2 << 30, max_views=6)
    try:
        ids = [srv.add_map(m.data) for m in maps]
        tickets = {}
        for k, (mi, f) in enumerate(reqs):
            t = srv.submit(ids[mi], f, tonemap_exposure=1.0 if f["color_bits"] == 64 else None,
                           pinned=pinned, hdr=not (pinned and k % 3 == 0))
            tickets[t] = k
            if k % 5 == 4:
                # requests keep arriving while earlier batches run
                for c in srv.poll():
                    tickets.setdefault(("done", c["ticket"]), c)
        done = srv.drain()
        for key in list(tickets):
            if isinstance(key, tuple):
                done.append(tickets.pop(key))
        st = srv.stats()
        assert len(done) == len(reqs), (len(done), len(reqs))
        assert st["batches"] < len(reqs), st  # requests were batched
        # pinned results stay valid until released: compare before closing the server
        handles = [gpu_ctx.upload_map(m.data) for m in maps]
        try:
            checked = 0
            for c in done:
                assert c["status"] == 0, c
                mi, f = reqs[tickets[c["ticket"]]]
                b = sw_host.batch_from_frames([f])
                alone = gpu_ctx.render_plain(handles[mi], b)
                if c["framebuffer"] is not None:
                    assert np.array_equal(c["framebuffer"], alone), f"request {tickets[c['ticket']]} differs from its solo render"
                else:
                    assert f["color_bits"] == 64
                if f["color_bits"] == 64:
                    ref, _ = oracle.tonemap(alone, 1.0, (1.0, 1.0, 1.0))
                    assert np.array_equal(c["tonemapped"], ref)
                # a few against the oracle as well
                if checked < 4 and c["framebuffer"] is not None:
                    ref, _, _, _ = oracle.render(maps[mi].data, b)
                    assert np.array_equal(c["framebuffer"].reshape(-1), ref)
                    checked += 1
                srv.release(c)
        finally:
            for h in handles:
                gpu_ctx.release_map(h)
    finally:
        srv.close()


def test_budget_gates_admission(maps):
    _say("case server budget")
    reqs = _requests(maps[:1], 6, 9)
    fb = sw_host.FrameBuilder()
    big = fb.build(maps[0], maps[0].camera_path(3, 2)[0], width=1920, height=1080, color_bits=64, use_directional_lightmaps=1)
    # budget: any single small request fits (<= 6.5 MB), the 1080p 64-bit one (> 16 MB) does not
    srv = sw_abi.Server(slots=2, hbm_budget=12 << 20, max_views=8)
    try:
        mid = srv.add_map(maps[0].data)
        tickets = [srv.submit(mid, f) for _, f in reqs]
        tb = srv.submit(mid, big)
        done = srv.drain()
        st = srv.stats()
    finally:
        srv.close()
    by = {c["ticket"]: c for c in done}
    assert by[tb]["status"] == sw_abi.SW_ERR_BUDGET
    for t in tickets:
        assert by[t]["status"] == 0, by[t]
    assert st["peak_bytes_in_flight"] <= 12 << 20
    assert st["rejected"] == 1

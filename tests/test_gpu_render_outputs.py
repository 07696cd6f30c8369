"""drmlt_render: the loop with the reference's intermediate outputs
(delayed.cpp:225-242, 610-655) -- partial images of the indirect buffer
scaled by pixels / mutations, their times, the statistics file -- and the
final film through drmlt_write_exr."""
import numpy as np
import pytest

from exr_read import read_exr

pytestmark = pytest.mark.gpu


def test_render_writes_partials_times_and_stats(cfg1, tmp_path):
    from drmlt import capi

    opt = capi.options_from_dpt(cfg1["options"], num_init_samples=8192, num_chains=1024, seed=21)
    ctx = capi.Context(cfg1, opt)
    ctx.prepass()
    ctx.direct_lighting()
    part = tmp_path / "drmlt_partial"
    print("render cfg1: 96 iterations, reports every 1 ms of wall time", flush=True)
    ctx.render(96, capi.BATCHED_GRAPH, interval_seconds=0.001, partial_dir=part, base_name="drmlt", chunk_iters=16)
    imgs = sorted(part.glob("drmlt_*.exr"), key=lambda p: int(p.stem.split("_")[1]))
    times = [float(x.rstrip(",")) for x in (part / "drmlt_time.csv").read_text().splitlines()]
    assert 1 <= len(imgs) <= 6 and len(times) == len(imgs)
    assert [p.name for p in imgs] == [f"drmlt_{k}.exr" for k in range(1, len(imgs) + 1)]
    assert all(b >= a for a, b in zip(times, times[1:]))
    # the last chunk's report: the indirect buffer x pixels / mutations so far
    W, H = ctx.W, ctx.H
    st = ctx.stats()
    assert st["large_step"] + st["stage1"] == 96 * 1024
    if len(imgs) == 6:
        # one report per chunk: the last image follows the last chunk
        ch, attrs = read_exr(imgs[-1])
        want = (ctx.film(1) * np.float32(W * H / (96 * 1024))).astype(np.float16)
        got = np.stack([ch["R"], ch["G"], ch["B"]], -1)
        assert np.abs(got.astype(np.float32) - want.astype(np.float32)).max() <= 1e-3 * max(float(np.abs(want).max()), 1e-6)
    assert (part / "drmlt_stats.txt").read_text() == capi.stats_text(st, reference_labels=True)
    ctx.write_stats(tmp_path / "s.txt", reference_labels=False)
    assert (tmp_path / "s.txt").read_text() == capi.stats_text(st, reference_labels=False)
    film = ctx.film(0)
    capi.write_exr(tmp_path / "image.exr", film)
    ch, _ = read_exr(tmp_path / "image.exr")
    assert np.array_equal(np.stack([ch["R"], ch["G"], ch["B"]], -1).view(np.uint16), film.astype(np.float16).view(np.uint16))
    ctx.close()


def test_render_without_reports_equals_run(cfg1):
    """No partial directory: drmlt_render advances the chains exactly as
    drmlt_run does (same counts, same film)."""
    from drmlt import capi

    res = []
    for use_render in (False, True):
        opt = capi.options_from_dpt(cfg1["options"], num_init_samples=8192, num_chains=1024, seed=22)
        ctx = capi.Context(cfg1, opt)
        ctx.prepass()
        if use_render:
            ctx.render(40, capi.BATCHED, chunk_iters=7)
        else:
            ctx.run(40, capi.BATCHED)
        res.append((ctx.stats(), ctx.film(1)))
        ctx.close()
    assert res[0][0] == res[1][0]
    np.testing.assert_allclose(res[0][1], res[1][1], rtol=1e-4, atol=1e-6)


def test_scene_file_renders_like_the_procedural_scene(cfg1, tmp_path):
    """The config-1 room written as XML + OBJ files and read back by
    drmlt_scene_load renders as the procedural scene does: the arrays agree
    up to the rounding of normalised normals and the recovered look-at, so
    the chains' statistics and the film agree closely."""
    import sys
    import os
    from drmlt import capi

    sys.path.insert(0, os.path.dirname(__file__))
    from scene_files import export_scene

    ld = capi.load_scene(export_scene(cfg1, tmp_path))
    res = []
    for sc in (cfg1, ld):
        opt = capi.options_from_dpt(cfg1["options"], num_init_samples=8192, num_chains=2048, seed=23)
        ctx = capi.Context(sc, opt)
        ctx.prepass()
        ctx.direct_lighting()
        ctx.run(16, capi.BATCHED_GRAPH)
        res.append((ctx.stats(), ctx.film(0)))
        ctx.close()
    (s0, f0), (s1, f1) = res
    print("procedural", s0, "\nfrom files", s1, flush=True)
    assert s0["large_step"] + s0["stage1"] == s1["large_step"] + s1["stage1"]
    for k in ("stage1_accepted", "stage2_iso", "stage2_h2mc"):
        assert abs(s0[k] - s1[k]) <= 0.02 * max(s0[k], 1), k
    assert abs(f0.mean() - f1.mean()) <= 0.02 * f0.mean()


def test_mcmc_render_writes_the_mcmc_stats_text(cfg1, tmp_path):
    from drmlt import capi

    opt = capi.options_from_dpt(dict(cfg1["options"], integrator="mcmc"), num_init_samples=8192, num_chains=1024, seed=24)
    ctx = capi.Context(cfg1, opt)
    ctx.prepass()
    ctx.render(8, capi.BATCHED_GRAPH, partial_dir=tmp_path / "mcmc_partial", base_name="mcmc")
    st = ctx.stats()
    assert st["stage2_iso"] == st["stage2_h2mc"] == 0 and st["large_step"] + st["stage1"] == 8 * 1024
    assert (tmp_path / "mcmc_partial" / "mcmc_stats.txt").read_text() == capi.stats_text(st, integrator="mcmc")
    ctx.close()


def test_h2mc_render_writes_its_stats_text(cfg1, tmp_path):
    from drmlt import capi

    opt = capi.options_from_dpt(dict(cfg1["options"], integrator="h2mc", lensperturbprob=0.0), num_init_samples=8192,
                                num_chains=1024, seed=25)
    ctx = capi.Context(cfg1, opt)
    ctx.prepass()
    ctx.render(8, capi.BATCHED_GRAPH, partial_dir=tmp_path / "h2mc_partial", base_name="h2mc")
    st = ctx.stats()
    assert st["large_step"] + st["stage1"] + st["stage2_h2mc"] == 8 * 1024
    assert (tmp_path / "h2mc_partial" / "h2mc_stats.txt").read_text() == capi.stats_text(st, integrator="h2mc")
    ctx.close()


def test_acceptance_map(cfg1, tmp_path):
    """acceptanceMap (delayed.cpp:246-275): radiance splats are replaced by a
    (1, 0, 0) splat per accepted first-stage small step and (0, 1, 0) per
    accepted second stage, the image the reference's tools/heatmap.py reads.
    The film holds integer counts, so the plain and graph-replayed batched
    forms (identical chains) give identical maps. The splat goes to the state
    being left; seed states carry no splat list, and a chain leaves its seed at
    most once, so red and green fall short of the accepted counts by at most
    one per chain in total. The map survives the EXR round trip."""
    from drmlt import capi
    from exr_read import read_exr

    films = {}
    for mode in (capi.PLAIN, capi.BATCHED_GRAPH):
        opt = capi.options_from_dpt(dict(cfg1["options"], acceptanceMap=True), num_init_samples=8192, num_chains=2048, seed=26)
        ctx = capi.Context(cfg1, opt)
        ctx.prepass()
        print(f"acceptance map, form {mode}: 12 iterations", flush=True)
        ctx.run(12, mode)
        films[mode] = (ctx.film(1), ctx.stats())
        ctx.close()
    (f, st), (fb, stb) = films[capi.PLAIN], films[capi.BATCHED_GRAPH]
    assert st == stb and np.array_equal(f, fb)
    assert np.array_equal(f, np.round(f)) and f[..., 2].sum() == 0
    red, green = f[..., 0].sum(), f[..., 1].sum()
    acc2 = st["stage2_iso_accepted"] + st["stage2_h2mc_accepted"]
    print("red", red, "stage-1 small accepted", st["stage1_accepted"], "green", green, "stage-2 accepted", acc2, flush=True)
    assert red <= st["stage1_accepted"] and green <= acc2
    assert (st["stage1_accepted"] - red) + (acc2 - green) <= 2048
    capi.write_exr(tmp_path / "map.exr", f, capi.EXR_NONE)
    ch, _ = read_exr(tmp_path / "map.exr")
    assert np.array_equal(ch["R"].astype(np.float32), f[..., 0]) and np.array_equal(ch["G"].astype(np.float32), f[..., 1])
    # heatmap.py's stage map, green / (red + green + eps), lies in [0, 1)
    ratio = f[..., 1] / (f[..., 0] + f[..., 1] + 1e-2)
    assert ratio.min() >= 0 and ratio.max() < 1

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

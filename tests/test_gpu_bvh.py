"""The device BVH build (bvh_device.hip) against the host builder
(bvh_build.cpp), both replacing the reference's Embree build
(src/scene.cpp:42-54): the same binned-SAH tree (node count, 4-wide depth,
stack bound), bit-identical answers to closest-hit and any-hit queries, and a
DR loop that advances identically over either."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _ctx(sc, host_bvh, **kw):
    from drmlt import capi

    opt = capi.options_from_dpt(sc["options"], host_bvh=host_bvh, **kw)
    return capi.Context(sc, opt)


def _rays(sc, n, seed):
    rng = np.random.default_rng(seed)
    p = sc["tri_p"].reshape(-1, 3)
    lo, hi = p.min(0), p.max(0)
    org = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # half the rays aimed at triangle corners (grazing and shared-edge cases)
    k = n // 2
    tgt = p[rng.integers(0, len(p), k)]
    dd = tgt - org[:k]
    d[:k] = dd / np.maximum(np.linalg.norm(dd, axis=1, keepdims=True), 1e-12)
    tmax = np.full(n, np.inf, np.float32)
    tmax[: n // 4] = rng.uniform(0.1, 2.0, n // 4).astype(np.float32)
    return org, d.astype(np.float32), np.full(n, 1e-4, np.float32), tmax


@pytest.mark.parametrize("which", ["cfg1", "cfg2_small", "cfg4_small", "cfg5_small"])
def test_device_bvh_answers_like_host_bvh(which, request):
    sc = request.getfixturevalue(which)
    dev = _ctx(sc, 0, num_chains=1024, num_init_samples=4096, seed=3)
    host = _ctx(sc, 1, num_chains=1024, num_init_samples=4096, seed=3)
    bd, bh = dev.bvh_info(), host.bvh_info()
    say(which, "device", bd, "host", bh)
    assert bd["on_device"] and not bh["on_device"]
    assert (bd["nodes"], bd["depth4"], bd["stack_need"]) == (bh["nodes"], bh["depth4"], bh["stack_need"])
    org, d, tmin, tmax = _rays(sc, 200000, 7)
    for any_hit in (False, True):
        pd, td = dev.trace(org, d, tmin, tmax, any_hit)
        ph, th = host.trace(org, d, tmin, tmax, any_hit)
        hits = int((ph >= 0).sum())
        say(f"{which} any_hit={any_hit}: {hits} hits of {len(org)}")
        assert hits > len(org) // 10
        if any_hit:
            np.testing.assert_array_equal(pd >= 0, ph >= 0)
        else:
            np.testing.assert_array_equal(pd, ph)
            np.testing.assert_array_equal(td[ph >= 0], th[ph >= 0])
    dev.close()
    host.close()


def test_loop_over_device_bvh_matches_host_bvh(cfg2_small):
    """Two batched iterations from the same seeds over either BVH: identical
    acceptance counts, films equal up to float summation order."""
    from drmlt import capi

    res = []
    for host_bvh in (1, 0):
        ctx = _ctx(cfg2_small, host_bvh, num_chains=4096, num_init_samples=8192, seed=5)
        ctx.prepass()
        ctx.run(2, capi.BATCHED)
        res.append((ctx.stats(), ctx.film(1).sum()))
        ctx.close()
    say(res)
    assert res[0][0] == res[1][0]
    assert np.isclose(res[0][1], res[1][1], rtol=1e-4)

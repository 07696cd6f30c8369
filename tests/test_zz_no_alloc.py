"""drmlt_grad_hess_noalloc, the derivative call that allocates no device
memory, against drmlt_grad_hess through the C ABI: the same kernels on the
same records give identical gradients, Hessians and flags, on two calls in a
row (the second smaller, so stale contents of the shared device array would
show), in FP32 and FP64; a call beyond its 4096-state limit is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def states(cfg2_small):
    """Valid path states of the DR loop on the reduced config-2 scene
    (pre-pass seeds plus two iterations, read back)."""
    from drmlt import capi

    opt = capi.options_from_dpt(cfg2_small["options"], num_init_samples=20000, num_chains=8192, seed=13)
    ctx = capi.Context(cfg2_small, opt)
    say("no-alloc states: pre-pass + 2 iterations, 8192 chains")
    ctx.prepass()
    ctx.run(2, capi.BATCHED_GRAPH)
    recs, c8 = ctx.read_chains()
    recs = recs[(c8[:, 0] > 0) & (recs["c"] + recs["l"] - 1 <= 12)]
    yield ctx, recs
    ctx.close()


@pytest.mark.parametrize("fp64", [False, True])
def test_noalloc_matches_grad_hess(states, fp64):
    from drmlt import capi

    ctx, recs = states
    assert len(recs) >= 4096 + 700
    big, small = recs[:4096], recs[4096:4096 + 700]
    for name, rs in (("first call, 4096 states", big), ("second call, 700 states", small)):
        say(f"grad_hess_noalloc fp64={fp64}: {name}")
        g1, h1, ok1 = ctx.grad_hess_noalloc(rs, fp64=fp64)
        say(f"grad_hess fp64={fp64}: {name}")
        g0, h0, ok0 = ctx.grad_hess(rs, fp64=fp64)
        assert ok0.sum() > len(rs) // 2
        assert np.array_equal(ok1, ok0)
        assert np.array_equal(g1, g0)
        assert np.array_equal(h1, h0)
        D = 2 * np.maximum(rs["c"] + rs["l"] - 1, 2)
        for i in range(0, len(rs), 97):  # entries beyond D are zero, the Hessian symmetric
            d = D[i]
            assert not g1[i, d:].any() and not h1[i, d:, :].any() and not h1[i, :, d:].any()
            assert np.array_equal(h1[i], h1[i].T)
    say("grad_hess_noalloc: 4097 states is refused")
    with pytest.raises(capi.DrmltError, match="at most 4096"):
        ctx.grad_hess_noalloc(np.concatenate([big, big[:1]]), fp64=fp64)

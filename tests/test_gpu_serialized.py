"""drmlt_path_func_derv: the reference's derivative call (PathFuncDerv with
its Serialize layout, src/path.h:145-149, src/path.cpp:2303-2391) on states
the loop realised, serialized on the CPU by the oracle's restatement, against
drmlt_grad_hess on the same records."""
import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_path_func_derv_equals_grad_hess(cfg2_small):
    """Same FP32 jets through either entry point: identical gradients and
    Hessians for every (c, l) the loop's states have; a scene layout that
    differs from the context's is rejected."""
    from drmlt import capi

    sc = cfg2_small
    opt = capi.options_from_dpt(dict(sc["options"], maxdepth=8), num_init_samples=8192, num_chains=8192, seed=17)
    ctx = capi.Context(sc, opt)
    ctx.prepass()
    ctx.run(2, capi.BATCHED)
    recs, c8 = ctx.read_chains()
    keep = (c8[:, 0] > 0) & (c8[:, 5] > 1e-15)
    recs = recs[keep]
    O = orc.Scene(sc)
    sp = O.serialize_scene(use_light_coord=bool(opt.use_light_coordinate_sampling))
    pairs = sorted(set(zip(recs["c"].tolist(), recs["l"].tolist())))
    say(f"{len(recs)} states, (c, l) pairs {pairs}")
    assert len(pairs) >= 6
    checked = 0
    for c, l in pairs:
        st = recs[(recs["c"] == c) & (recs["l"] == l)][:400]
        D = 2 * max(c + l - 1, 2)
        primary, vert = O.serialize(st)
        say(f"(c, l) = ({c}, {l}): {len(st)} states, D = {D}")
        g, h, ok = ctx.path_func_derv(c, l, st["screen"], primary, sp, vert)
        g32, h32, ok32 = ctx.grad_hess(st, fp64=False)
        fin = (ok32 == 1) & np.isfinite(g32[:, :D]).all(1) & np.isfinite(h32[:, :D, :D]).all((1, 2))
        np.testing.assert_array_equal(ok == 1, fin)
        np.testing.assert_array_equal(g[fin], g32[fin, :D].astype(np.float32))
        np.testing.assert_array_equal(h[fin], h32[fin, :D, :D].astype(np.float32))
        assert (g[~fin] == 0).all() and (h[~fin] == 0).all()
        checked += int(fin.sum())
    assert checked > 500
    bad = sp.copy()
    bad[33] *= 1.01  # camera dist
    st = recs[:4]
    c, l = int(st["c"][0]), int(st["l"][0])
    st = recs[(recs["c"] == c) & (recs["l"] == l)][:4]
    primary, vert = O.serialize(st)
    with pytest.raises(capi.DrmltError, match="dist"):
        ctx.path_func_derv(c, l, st["screen"], primary, bad, vert)
    bad = sp.copy()
    bad[0] = 1.0 - bad[0]
    with pytest.raises(capi.DrmltError, match="useLightCoordinateSampling"):
        ctx.path_func_derv(c, l, st["screen"], primary, bad, vert)
    ctx.close()

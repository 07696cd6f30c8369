"""DRMLT mutation-loop benchmark on one MI355X.

A step is one DR iteration (delayed.cpp:396-656) of every chain. The
workload is config 4, the configuration the equal-RMSE target is quoted on:
a seeded procedural scene of 512 icospheres and extruded boxes on a
displaced floor (1,836,036 triangles), Phong n log-uniform in [100, 1000] on
60% of the surfaces, a 0.2 x 0.2 m light of radiance 50, 1024 x 1024 film,
1,048,576 chains, maximum depth 10; FP32 path evaluation, FP64 Gaussians and
acceptance ratios. Chain states, scene and BVH are resident in HBM before
the timed region; the pre-pass, the direct-lighting pass and the autotuner
run before it.

Prints one JSON line: mutations/s of the batched + adjoint form (hipGraph
replay of the per-stage persistent kernels), the plain one-kernel-per-step
form beside it, the roofline of the dominant kernel, the CPU baseline (the
oracle's C++ restatement of the reference loop on the host cores), the
equal-RMSE study on the same configuration (the time the batched form needs
to reach the plain form's 60 s RMSE), config 2 at its maximum depth 9, the
mcmc and h2mc integrators and the derivative kernels alone (config 3).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "drmlt-dpt_amd"))

HBM_PEAK_GBS = 8000.0
STAGES = ["stage1", "deriv_cur", "gauss_cur", "stage2_propose", "deriv_prop", "gauss_prop", "stage2_accept"]
STAGE_KERNELS = ["k_b_stage1", "k_b_deriv", "k_b_gauss", "k_b_stage2_propose", "k_b_deriv", "k_b_gauss", "k_b_stage2_accept"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(sc, threads, chains, iters, seed=11):
    """The oracle's CPU loop (FP32 path, like the reference as shipped) on a
    bounded sample: `chains` chains x `iters` iterations after a pre-pass."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc

    s = orc.Scene(sc)
    maxd = int(sc["options"]["maxdepth"])
    b, lf, seeds, c8, _ = s.prepass(20000, maxd, chains, seed=seed, threads=threads)
    _, stats, sec = s.render(seeds, c8, lf, b, iters, threads=threads, seed=seed)
    acc = int(stats[1]) + int(stats[3]) + int(stats[5]) + int(stats[7])
    return chains * iters / sec, sec, acc / sec


def derivative_workload(n=262144, seed=4321):
    """The derivative kernels alone: gradient and full Hessian of log f at n
    path states of the config-2 scene, in FP32 with an FP64 check. The states
    are realised by the loop itself (pre-pass seeds plus two DR iterations at
    maximum depth 9, so D = 2 * path length reaches 18), read back with
    drmlt_read_chains and kept where the reference computes derivatives
    (drsmallstep.cpp:135-139: f > 1e-15, c + l >= 3, path length <= 9)."""
    import numpy as np
    from drmlt import capi, scenes

    sc = scenes.config2()
    opt = capi.options_from_dpt(dict(sc["options"], maxdepth=9), num_chains=n, seed=seed)
    ctx = capi.Context(sc, opt)
    ctx.prepass()
    ctx.run(2, capi.BATCHED_GRAPH)
    recs, c8 = ctx.read_chains()
    P = recs["c"] + recs["l"] - 1
    keep = (c8[:, 0] > 0) & (c8[:, 5] > 1e-15) & (recs["c"] + recs["l"] >= 3) & (P <= int(opt.max_derivatives_depth))
    recs = recs[keep]
    D = 2 * np.maximum(recs["c"] + recs["l"] - 1, 2)
    log(f"derivative workload: {len(recs)} of {n} chain states, D histogram {np.bincount(D).tolist()}")
    g32, h32, ok32 = ctx.grad_hess(recs, fp64=False)
    ms32 = float(ctx.timing()[0])
    g64, h64, ok64 = ctx.grad_hess(recs, fp64=True)
    ms64 = float(ctx.timing()[0])
    _, okl = ctx.eval_logf(recs, fp64=False)
    msl = float(ctx.timing()[0])
    ctx.close()
    both = (ok32 == 1) & (ok64 == 1)
    eg, eh = [], []
    for i in np.nonzero(both)[0][:20000]:
        d = D[i]
        eg.append(np.linalg.norm(g32[i, :d] - g64[i, :d]) / max(np.linalg.norm(g64[i, :d]), 1e-30))
        eh.append(np.linalg.norm(h32[i, :d, :d] - h64[i, :d, :d]) / max(np.linalg.norm(h64[i, :d, :d]), 1e-30))
    return {"workload": f"config3: gradient + Hessian of log f at {len(recs)} states of the config-2 scene "
                        f"(D {int(D.min())}-{int(D.max())}), states from the DR loop",
            "states": int(len(recs)), "d_histogram": np.bincount(D).tolist(),
            "fp32_ms": round(ms32, 3), "fp32_states_per_s": round(len(recs) / (ms32 / 1000.0), 1),
            "fp64_ms": round(ms64, 3), "fp64_states_per_s": round(len(recs) / (ms64 / 1000.0), 1),
            "logf_fp32_ms": round(msl, 3),
            "fp32_vs_fp64_rel_err": {"grad_median": float(np.median(eg)), "grad_p99": float(np.percentile(eg, 99)),
                                     "hess_median": float(np.median(eh)), "hess_p99": float(np.percentile(eh, 99)),
                                     "states_compared": len(eg)},
            "ok_fraction": float(both.mean())}


def rmse_study(config, budget_s, ref_steps, sched, cpu_threads=0, seed=1234, plain_chunk=4, batched_chunk=4):
    """Equal-time image RMSE against a long reference render.

    The reference is a batched render from independent seeds with 256 direct
    samples per pixel. The plain form, the batched form (under the tuned
    schedule `sched`) and, if cpu_threads > 0, the oracle's CPU loop on the same
    number of chains each render for `budget_s` seconds of their own time
    (device time for the GPU forms); the RMSE of the full image and of the DR
    (indirect) part are taken over the film against the reference's. The
    batched form's curve gives the time it needs to reach the plain form's
    indirect RMSE at the budget."""
    import numpy as np
    from drmlt import capi, scenes

    sc = scenes.CONFIGS[config]()
    dpt = dict(sc["options"])
    chains = int(dpt["numchains"])
    W, H = int(sc["cam"][10]), int(sc["cam"][11])

    def make(seed_, b_common=None, **kw):
        ctx = capi.Context(sc, capi.options_from_dpt(dpt, num_chains=chains, seed=seed_, **kw))
        ctx.set_schedule(**sched)
        b, lf = ctx.prepass()
        if b_common is not None:
            # the chains' own seed states, the reference's normalisation: the
            # pre-pass estimate of b scales the whole DR image, and its
            # sampling error would otherwise put a floor under every RMSE
            st, c8 = ctx.read_chains()
            ctx.seed_chains(st, c8, b_common, lf)
        ctx.direct_lighting()
        return ctx, b

    def images(ctx):
        full = ctx.film(0)
        return full, full - ctx.film(2) / float(ctx.opt.direct_spp)

    rmse = lambda a, b: float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)))
    # relative MSE (per pixel and channel, against the reference squared plus 0.01)
    relmse = lambda a, b: float(np.mean((a.astype(np.float64) - b) ** 2 / (b.astype(np.float64) ** 2 + 1e-2)))
    t0 = time.time()
    ref, b_ref = make(seed + 7777, direct_spp=256)
    if not sched:
        # a schedule for this workload, tuned on the reference's chains (its
        # tuning steps only add to the reference render)
        sched, _ = ref.autotune(steps=2)
        log(f"rmse study: schedule {sched}")
    chunk = 50
    for k in range(0, ref_steps, chunk):
        ref.run(min(chunk, ref_steps - k), capi.BATCHED_GRAPH)
        if time.time() - t0 > 10:
            log(f"rmse study: reference {k + chunk}/{ref_steps} steps")
            t0 = time.time()
    rfull, rind = images(ref)
    ref.close()

    def curve(mode, step_chunk, target=None):
        """RMSE points every step_chunk steps until budget_s of device time;
        with a target, stop once the curve has stayed at or below it for
        30% (at least 1 s) beyond its first crossing."""
        ctx, _ = make(seed, b_common=b_ref)
        t, steps, pts, crossed = 0.0, 0, [], None
        t1 = time.time()
        while t < budget_s * 1000.0:
            ctx.run(step_chunk, mode)
            t += float(ctx.timing()[0])
            steps += step_chunk
            full, ind = images(ctx)
            pts.append((t / 1000.0, steps, rmse(full, rfull), rmse(ind, rind), relmse(ind, rind),
                        float(ind.mean() / max(rind.mean(), 1e-30))))
            if target is not None:
                if pts[-1][3] > target:
                    crossed = None
                elif crossed is None:
                    crossed = pts[-1][0]
                if crossed is not None and pts[-1][0] >= max(1.3 * crossed, crossed + 1.0):
                    break
            if time.time() - t1 > 10:
                log(f"rmse study: mode {mode} {t / 1000.0:.1f}/{budget_s} s")
                t1 = time.time()
        ctx.close()
        return pts

    plain = curve(capi.PLAIN, plain_chunk)
    batched = curve(capi.BATCHED_GRAPH, batched_chunk, target=plain[-1][3])
    # the batched form's time to the plain form's indirect RMSE at the budget:
    # the first point from which its curve stays at or below it (an almost
    # empty early film can lie below it before the estimate has settled)
    target = plain[-1][3]
    reach = None
    for i in range(len(batched) - 1, -1, -1):
        if batched[i][3] > target:
            break
        reach = batched[i]
    pick = lambda pts: {"time_s": round(pts[-1][0], 3), "steps": pts[-1][1], "rmse": pts[-1][2], "rmse_indirect": pts[-1][3],
                        "relmse_indirect": pts[-1][4], "mean_ratio_indirect": round(pts[-1][5], 4),
                        "curve_s_rmse_indirect": [(round(p[0], 3), p[3]) for p in pts[:: max(1, len(pts) // 12)]]}
    out = {"workload": f"config{config}: {len(sc['tri_p'])} tris, {chains} chains, {W}x{H}",
           "budget_s": budget_s, "schedule": sched,
           "criterion": "indirect (DR) RMSE: the direct pass is the same 16-spp pass in every form",
           "batched_curve": "stops once it has stayed at or below the plain form's final RMSE for 30% (>= 1 s) past its crossing",
           "reference": {"steps": ref_steps, "direct_spp": 256, "seed": "independent", "b": b_ref,
                         "note": "test renders keep their own seed states and take the reference's normalisation b"},
           "plain": pick(plain), "batched": pick(batched),
           "batched_time_to_plain_rmse_s": round(reach[0], 4) if reach else None,
           "batched_steps_to_plain_rmse": reach[1] if reach else None,
           "time_ratio_plain_over_batched_at_equal_rmse": round(plain[-1][0] / reach[0], 2) if reach else None}
    if cpu_threads > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import orc

        O = orc.Scene(sc)
        b, lf, seeds, c8, _ = O.prepass(int(dpt["numinitsamples"]), int(dpt["maxdepth"]), chains, seed=seed, threads=cpu_threads)
        _, _, sec1 = O.render(seeds, c8, lf, b, 1, threads=cpu_threads, seed=seed)
        iters = max(1, int(budget_s / max(sec1, 1e-6)))
        film, _, sec = O.render(seeds, c8, lf, b_ref, iters, threads=cpu_threads, seed=seed)
        ind = film / (iters * chains / float(W * H))
        dctx, _ = make(seed)  # the same direct pass as the GPU forms' films
        direct = dctx.film(2) / float(dctx.opt.direct_spp)
        dctx.close()
        out["cpu"] = {"time_s": round(sec, 3), "steps": iters, "threads": cpu_threads, "kind": "port",
                      "rmse": rmse(ind + direct, rfull), "rmse_indirect": rmse(ind, rind), "relmse_indirect": relmse(ind, rind),
                      "mean_ratio_indirect": round(float(ind.mean() / max(rind.mean(), 1e-30)), 4)}
    return out


def cpu_info():
    """The host CPU's model name and count as the OS reports them."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def config2_depth9(steps, warmup, tune_steps):
    """Config 2 (65,536 chains, 457,732 triangles) at maximum depth 9, the
    path lengths 3..9 (D 6..18) its definition names: graph-replayed step time
    under its own autotuned schedule, the plain form beside it."""
    from drmlt import capi, scenes

    sc = scenes.config2()
    dpt = dict(sc["options"], maxdepth=9)
    chains = int(dpt["numchains"])
    ctx = capi.Context(sc, capi.options_from_dpt(dpt, num_chains=chains, seed=1234))
    ctx.prepass()
    ctx.direct_lighting()
    ctx.run(2, capi.PLAIN)
    plain_ms = ctx.timing()[0] / 2
    sched, _ = ctx.autotune(steps=tune_steps)
    ctx.run(warmup, capi.BATCHED_GRAPH)
    s0 = ctx.stats()
    ctx.run(steps, capi.BATCHED_GRAPH)
    ms = ctx.timing()[0] / steps
    acc = accepted(ctx.stats()) - accepted(s0)
    ctx.close()
    return {"workload": f"config2: {len(sc['tri_p'])} tris, {chains} chains, maxdepth 9, 512x512", "schedule": sched,
            "ms_per_step": round(ms, 4), "mutations_per_s": round(chains / (ms / 1000.0), 1),
            "accepted_mutations_per_s": round(acc / (steps * ms / 1000.0), 1),
            "plain_ms_per_step": round(plain_ms, 3), "plain_mutations_per_s": round(chains / (plain_ms / 1000.0), 1)}


def accepted(st):
    """Accepted mutations in a stats dict: accept.first or accept.second of an
    iteration (delayed.cpp:565), counted per stage."""
    return (st["large_step_accepted"] + st["stage1_accepted"] + st["stage2_iso_accepted"] + st["stage2_h2mc_accepted"] +
            st.get("lens_step_accepted", 0))


def copy_calibration(mb):
    """One device-to-device hipMemcpy of `mb` MB (read mb, write mb)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {mb} MB device to device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--chains", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16, help="threads of the CPU baseline (the host CPU share of one GPU)")
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--tune-steps", type=int, default=8, help="timed steps per candidate schedule of the autotuner")
    ap.add_argument("--no-deriv", action="store_true", help="skip the derivative-kernel workload (config 3)")
    ap.add_argument("--no-mcmc", action="store_true", help="skip the step times of the mcmc and h2mc integrators")
    ap.add_argument("--no-config2", action="store_true", help="skip config 2 at its maximum depth 9")
    ap.add_argument("--rmse-config", type=int, default=-1,
                    help="scene configuration of the equal-RMSE study (-1: the bench configuration, 0: none)")
    ap.add_argument("--rmse-budget", type=float, default=60.0, help="seconds the plain form renders in the RMSE study")
    ap.add_argument("--rmse-ref-steps", type=int, default=2000, help="batched steps of the RMSE study's reference render")
    ap.add_argument("--rmse-cpu", action="store_true", help="also render with the CPU loop in the RMSE study")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch schedule (profiled runs: no candidate schedules among the launches)")
    ap.add_argument("--schedule", type=str, default="",
                    help='a fixed launch schedule as JSON (e.g. \'{"stage_lanes": 32}\') instead of the autotuner\'s; '
                         "profiled runs use the schedule the autotuner picks for the workload")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="first copy MB megabytes device to device (hipMemcpy), the known-size kernel that scales "
                         "rocprof FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the DR chains would shard by chain "
              "over GPUs with a film all-reduce; no multi-GPU run is part of this build)", file=sys.stderr)
        sys.exit(2)

    from drmlt import capi, scenes

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    t0 = time.time()
    sc = scenes.CONFIGS[args.config]()
    dpt = dict(sc["options"])
    chains = args.chains or int(dpt["numchains"])
    opt = capi.options_from_dpt(dpt, num_chains=chains, seed=1234)
    log(f"scene cfg{args.config}: {len(sc['tri_p'])} triangles, {chains} chains, built in {time.time() - t0:.1f}s")
    # the BVH: the context's device build, and the host build of the same
    # tree for comparison (a context that holds only the scene)
    hctx = capi.Context(sc, capi.options_from_dpt(dpt, num_chains=1024, seed=1234, host_bvh=1))
    bvh_host = hctx.bvh_info()
    hctx.close()
    ctx = capi.Context(sc, opt)
    bvh = ctx.bvh_info()
    log(f"BVH {bvh['nodes']} 4-wide nodes, depth {bvh['depth4']}, stack {bvh['stack_need']}: device build {bvh['build_ms']:.1f} ms, "
        f"host build {bvh_host['build_ms']:.1f} ms")
    t1 = time.time()
    b, lf = ctx.prepass()
    ctx.direct_lighting()
    log(f"pre-pass b={b:.5f} + direct pass {time.time() - t1:.1f}s")

    # plain one-kernel-per-step form, same chains
    ctx.run(max(1, args.plain_steps // 2), capi.PLAIN)
    s0 = ctx.stats()
    ctx.run(args.plain_steps, capi.PLAIN)
    plain_ms = ctx.timing()[0] / args.plain_steps
    plain_acc = accepted(ctx.stats()) - accepted(s0)
    log(f"plain form {plain_ms:.2f} ms/step")

    # launch schedule of the batched form chosen by the autotuner on these chains
    if args.schedule:
        ctx.set_schedule(**json.loads(args.schedule))
        sched = ctx.schedule()
        log(f"fixed schedule {sched}")
    elif args.no_tune:
        sched = ctx.schedule()
        log(f"default schedule {sched}")
    else:
        sched, tuned_ms = ctx.autotune(steps=args.tune_steps)
        log(f"autotuned schedule {sched}: {tuned_ms:.3f} ms/step")

    # batched form with per-stage timing (events between kernels)
    ctx.run(3, capi.BATCHED)
    stage_ms = ctx.timing()[1:8] / 3
    log("batched per-stage ms: " + ", ".join(f"{n}={v:.2f}" for n, v in zip(STAGES, stage_ms)))

    # work counters for one batched step
    work = ctx.count_work()
    log(f"work counters: {work}")

    # batched + graph: the measured form
    ctx.run(args.warmup, capi.BATCHED_GRAPH)
    s0 = ctx.stats()
    ctx.run(args.steps, capi.BATCHED_GRAPH)
    total_ms = ctx.timing()[0]
    ms_per_step = total_ms / args.steps
    value = chains * args.steps / (total_ms / 1000.0)
    st = ctx.stats()
    batched_acc = accepted(st) - accepted(s0)
    log(f"batched+graph {ms_per_step:.3f} ms/step, {value:.4g} mutations/s; stats {st}")

    # roofline of the dominant stage kernel
    dom = int(max(range(len(STAGES)), key=lambda k: stage_ms[k]))
    nodes, tris, rays, state_bytes = work[dom]
    # 128 B per 4-wide BVH node visited (one cache line), 48 B per triangle tested
    alg_bytes = nodes * 128 + tris * 48 + state_bytes
    achieved = alg_bytes / (stage_ms[dom] / 1000.0) / 1e9
    traffic, traffic_src, valu, lanes = None, None, None, None
    k = STAGE_KERNELS[dom]
    if sched.get("stage_waves") == 3 and k in ("k_b_stage1", "k_b_stage2_propose", "k_b_stage2_accept"):
        k += "_w3"  # the path-stage kernels bounded to three waves per SIMD
    prof = os.path.join(ROOT, "profiles", f"dr_step_config{args.config}_hbm_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            pt = json.load(f)
        if k in pt.get("kernels", {}):
            traffic = int(pt["kernels"][k]["hbm_bytes_per_launch"])
            valu = pt["kernels"][k].get("valu_util")
            lanes = pt["kernels"][k].get("lanes_per_valu")
            traffic_src = (f"rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of {k} at commit {pt['commit']}, scaled by a "
                           f"256 MiB device copy (x{pt['fetch_scale']:.2f}, x{pt['write_scale']:.2f}); "
                           f"profiles/dr_step_config{args.config}_hbm_traffic.md")
    roof = {"kernel": STAGES[dom], "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": int(alg_bytes), "launch_ms": round(float(stage_ms[dom]), 4),
            "rays_per_launch": int(rays)}
    if traffic_src:
        roof["traffic_source"] = traffic_src
    if valu is not None:
        roof["valu_util"] = valu
        roof["valu_util_source"] = (f"4 x SQ_ACTIVE_INST_VALU / (1024 SIMDs x GRBM_GUI_ACTIVE / 8) of {k}, "
                                    f"profiles/dr_step_config{args.config}_valu.md")
    if lanes is not None:
        roof["active_lanes_per_valu_instr"] = lanes  # SQ_THREAD_CYCLES_VALU / SQ_ACTIVE_INST_VALU, of 64

    ctx.close()

    # the mcmc integrator (the first stage alone) and the h2mc integrator
    # (lens mutations at the reference's lensperturbprob 0.3) on the same
    # scene and chains
    others = {}
    if not args.no_mcmc:
        for name, extra in (("mcmc", {}), ("h2mc", {"lensperturbprob": 0.3})):
            mctx = capi.Context(sc, capi.options_from_dpt(dict(dpt, integrator=name, **extra), num_chains=chains, seed=1234))
            mctx.set_schedule(**sched)  # the path-stage scheduling tuned for these chains
            mctx.prepass()
            mctx.run(args.warmup, capi.BATCHED_GRAPH)
            m0 = mctx.stats()
            mctx.run(args.steps, capi.BATCHED_GRAPH)
            mms = mctx.timing()[0] / args.steps
            macc = accepted(mctx.stats()) - accepted(m0)
            mctx.close()
            others[name] = {"ms_per_step": round(mms, 4), "mutations_per_s": round(chains / (mms / 1000.0), 1),
                            "accepted_mutations_per_s": round(macc / (args.steps * mms / 1000.0), 1)}
            log(f"{name} integrator {mms:.3f} ms/step")
    cfg2 = None
    if not args.no_config2:
        cfg2 = config2_depth9(args.steps, args.warmup, args.tune_steps)
        log(f"config 2 at maxdepth 9: {cfg2}")
    study = None
    rmse_cfg = args.config if args.rmse_config < 0 else args.rmse_config
    if rmse_cfg:
        study = rmse_study(rmse_cfg, args.rmse_budget, args.rmse_ref_steps, sched if rmse_cfg == args.config else None,
                           cpu_threads=min(args.cpu_threads, os.cpu_count() or 1) if args.rmse_cpu else 0)
        log(f"rmse study: {study}")
    deriv = None
    if not args.no_deriv:
        deriv = derivative_workload()
        log(f"derivative workload: {deriv}")

    cpu = None
    if not args.no_cpu:
        ncpu = min(args.cpu_threads, os.cpu_count() or 1)
        # about 10-20 s of chain loop on 16 threads
        cchains, citers = {1: (1024, 2048), 2: (16384, 512)}.get(args.config, (16384, 768))
        model, host_cpus = cpu_info()
        log(f"cpu baseline: the oracle loop, {cchains} chains x {citers} iterations on {ncpu} threads ({model}, {host_cpus} CPUs)")
        cval, csec, cacc = cpu_baseline(sc, ncpu, cchains, citers)
        cpu = {"value": round(cval, 1), "unit": "mutations/s", "cores": ncpu, "kind": "port",
               "accepted_mutations_per_s": round(cacc, 1),
               "sample": f"{cchains} chains x {citers} iterations of the same scene after a 20000-sample pre-pass "
                         f"({csec:.1f} s of chain loop)",
               "host_cpu_model": model, "host_cpu_count": host_cpus,
               "threads_note": "one std::thread per core over chains; the host grants one GPU's share of 16 CPU threads"}
        log(f"cpu baseline {cval:.4g} mutations/s on {ncpu} threads")

    out = {
        "metric": "DRMLT mutations/s + sec to equal RMSE: batched+adjoint vs plain GPU vs CPU",
        "value": round(value, 1), "unit": "mutations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "seconds_to_equal_rmse": None if not study else {
            "batched_s": study["batched_time_to_plain_rmse_s"], "plain_s": study["plain"]["time_s"],
            "ratio_plain_over_batched": study["time_ratio_plain_over_batched_at_equal_rmse"], "target": "plain form's RMSE at its budget"},
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 path / f64 gaussian+ratio", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": f"config{args.config}: {len(sc['tri_p'])} tris, {chains} chains, maxdepth {dpt['maxdepth']}, "
                               f"{int(sc['cam'][10])}x{int(sc['cam'][11])} film, type={dpt['type']}",
                   "chains": chains, "form": "batched+graph", "schedule": sched},
        "accepted_mutations_per_s": round(batched_acc / (total_ms / 1000.0), 1),
        "plain_form": {"ms_per_step": round(plain_ms, 3), "mutations_per_s": round(chains / (plain_ms / 1000.0), 1),
                       "accepted_mutations_per_s": round(plain_acc / (args.plain_steps * plain_ms / 1000.0), 1)},
        # the forms advance identical chains (identical acceptance statistics, tests/test_gpu_loop.py), so
        # equal RMSE is reached after equal iteration counts and the time ratio is the step-time ratio
        "equal_rmse_time_ratio_plain_over_batched": round(plain_ms / ms_per_step, 3),
        "mcmc_integrator": others.get("mcmc"),
        "h2mc_integrator": others.get("h2mc"),
        "stage_ms": {n: round(float(v), 4) for n, v in zip(STAGES, stage_ms)},
        "acceptance": st,
        "roofline": roof,
        "cpu_baseline": cpu,
        "derivatives": deriv,
        "equal_time_rmse": study,
        "config2_maxdepth9": cfg2,
        "bvh": {"builder": "device" if bvh["on_device"] else "host", "build_ms": round(bvh["build_ms"], 2),
                "host_build_ms": round(bvh_host["build_ms"], 2), "nodes_4wide": bvh["nodes"], "depth4": bvh["depth4"],
                "stack_need": bvh["stack_need"], "same_tree": (bvh["nodes"], bvh["depth4"], bvh["stack_need"]) ==
                (bvh_host["nodes"], bvh_host["depth4"], bvh_host["stack_need"])},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

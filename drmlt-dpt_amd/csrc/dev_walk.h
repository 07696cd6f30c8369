// Path walks with one trace site per round.
//
// The path code of dev_path.h traces rays from many call sites (the light and
// camera subpath loops, each kind of connection). When the lanes of a wave are
// at different points of their paths -- different strategies (c, l), a large
// step beside small steps -- the wave runs each site's traversal in turn, each
// time for the few lanes that reached it. A walk instead advances one lane's
// path evaluation (PerturbPathBidir, path.cpp:1480-1628, or GenerateSubpath,
// path.cpp:1123-1269) up to its next ray query and stops there; then every
// lane of the wave with a query traces it together at one site, and each walk
// resumes with its result.
//
// The arithmetic and the order of the random draws are those of
// perturb_eval<float, false> and generate_subpath (the plain form runs those,
// and both forms produce the same chains). Two things are reordered without
// changing any result:
//  * a visibility test is traced after the contribution it gates has been
//    computed (a contribution that comes out zero needs no test);
//  * the visibility test of a light-coordinate sampled edge (path.cpp:
//    1430-1478) is answered by the closest-hit query along the same ray that
//    the next vertex needs: the segment (eps, (1 - eps) d) is occluded exactly
//    when that closest hit lies before its end.
#pragma once
#include "drmlt_kernels.h"

namespace drmlt {

// Where a walk's primary samples come from: generated from uniforms
// (GenerateSubpath), or the state's own u plus an offset -- N(0, sigma)
// draws, the chain's offset array, or the difference of two states (Green's
// y*, mutation.cpp:210-237).
enum : int { WK_GEN = 0, WK_OFF_NORMAL = 1, WK_OFF_ARRAY = 2, WK_OFF_DIFF = 3 };
// What a walk waits for: a light-subpath hit, a camera-subpath hit, or the
// visibility of a finished contribution.
enum : int { WP_LHIT = 1, WP_CHIT = 2, WP_VIS = 3 };

// The per-lane state of a walk. Chain state is named by slot index (p, and
// px / py for the y* offset) against the chain arrays, which every lane of a
// kernel shares, so no per-lane pointers are held; the pending query is the
// walk's own ray r (for a visibility test, the shadow segment), and a
// contribution awaiting its visibility sits in the walk's output already.
struct Walk {
    int p;                     // the state slot evaluated (and, for WK_GEN, written) in place
    int mode, c, l;
    Philox *g;                 // uniforms (WK_GEN); offset and discrete-dimension normals (perturbation)
    float sigma;               // WK_OFF_NORMAL sigma
    int offChain;              // WK_OFF_ARRAY: offset k of chain offChain in the chains' offset array
    int px, py;                // WK_OFF_DIFF: offset k = u_k of slot px - u_k of slot py
    int k;                     // next primary-sample slot
    int phase, d;              // what the walk waits for; vertex index of the subpath being built
    RayT<float> r;             // the ray of the pending query
    bool qany;                 // the pending query is a visibility test
    BState<float> ls, cs;
    int lastLightPrim;
    float sx, sy;
    bool lcp;                  // light-coordinate sampling of the last camera edge
    float lcpEnd;              // end of the light-coordinate segment awaiting its answer (< 0: none)
    bool visScreen;            // the contribution awaiting visibility also sets the state's screen position
};

DEV float walk_next_u(const Chains &ch, Walk &w, const SRef &ref) {  // PerturbU::next with the walk's offset
    float off;
    if (w.mode == WK_OFF_NORMAL) off = float(w.sigma * w.g->normal());
    else if (w.mode == WK_OFF_ARRAY) off = chain_off(ch, w.offChain)[w.k];
    else off = slot_u(ch, w.px)[w.k] - slot_u(ch, w.py)[w.k];
    float *u = ref.u + w.k * ref.s;
    const float v = modulo1(*u + off);
    *u = v;
    w.k++;
    return v;
}
DEV float walk_disc(Walk &w, float dsigma) { return float(dsigma * w.g->normal()); }  // DiscNormal
DEV void walk_put(Walk &w, const SRef &ref, float v) { ref.u[w.k * ref.s] = v; w.k++; }  // UWriter::put

// The next query is the closest hit along w.r.
DEV void walk_closest(Walk &w) { w.qany = false; }
// occluded() of dev_scene.h as a query: the shadow segment replaces the ray
// (the walk ends with the answer); the contribution is already in out.
DEV void walk_visibility(Walk &w, const ShadowQ &q, bool screen) {
    w.r.org = q.o; w.r.dir = q.d;
    w.r.minT = kIsectEps; w.r.maxT = (1.0f - kShadowEps) * q.dist;
    w.qany = true;
    w.visScreen = screen;
    w.phase = WP_VIS;
}

DEV bool walk_camera_start(const DevScene &S, const Chains &ch, Walk &w, const SRef &ref) {
    if (w.mode == WK_GEN) {
        const float sx = w.g->uni(), sy = w.g->uni();
        walk_put(w, ref, sx); walk_put(w, ref, sy);
        ref.screen[0] = sx; ref.screen[ref.s] = sy;
        w.sx = sx; w.sy = sy;
    } else {
        w.sx = walk_next_u(ch, w, ref);
        w.sy = walk_next_u(ch, w, ref);
    }
    emit_from_camera(S, w.sx, w.sy, w.r, w.cs);
    w.lcp = false;
    w.lcpEnd = -1.0f;
    w.d = 0;
    if (w.mode != WK_GEN && w.c < 2) return false;
    w.phase = WP_CHIT;
    walk_closest(w);
    return true;
}

// Begins the walk; true when a query is pending, false when it has ended
// (out holds its contribution).
DEV bool walk_start(const DevScene &S, const Opts &o, const Chains &ch, Walk &w, POut<float> &out) {
    out.valid = false;
    const SRef ref = slot_ref(ch, w.p);
    const int s = ref.s, vs = ref.vs;
    w.k = 0;
    w.lastLightPrim = -1;
    w.lcpEnd = -1.0f;
    if (w.mode == WK_GEN) {
        *ref.c = w.c; *ref.l = w.l;
        *ref.light = -1; *ref.lprim = -1;
        for (int k = 0; k < kVMax; k++) {
            ref.cprim[k * vs] = -1; ref.lgprim[k * vs] = -1; ref.crr[k * vs] = 1.0f; ref.lrr[k * vs] = 1.0f;
            ref.cdisc[k * vs] = 0.0f; ref.ldisc[k * vs] = 0.0f;
        }
        for (int k = 0; k < kDMax; k++) ref.u[k * s] = 0.0f;
        // as generate_subpath: no path without two camera vertices or a light
        // subpath that ends in the camera connection
        if (w.c < 2 && w.l < 2) return false;
    }
    if (w.l > 1) {
        float p0, p1, d0, d1;
        int li, lp;
        if (w.mode == WK_GEN) {
            p0 = w.g->uni(); p1 = w.g->uni(); d0 = w.g->uni(); d1 = w.g->uni();
            li = pick_light(S, w.g->uni());
            lp = pick_light_tri(S, li, w.g->uni());
            *ref.light = li; *ref.lprim = lp;
            walk_put(w, ref, p0); walk_put(w, ref, p1); walk_put(w, ref, d0); walk_put(w, ref, d1);
        } else {
            li = *ref.light; lp = *ref.lprim;
            p0 = walk_next_u(ch, w, ref); p1 = walk_next_u(ch, w, ref); d0 = walk_next_u(ch, w, ref); d1 = walk_next_u(ch, w, ref);
        }
        emit_from_light(S, o, li, lp, p0, p1, d0, d1, w.r, w.ls);
        w.r.minT = kIsectEps;
        w.r.maxT = __builtin_inff();
        w.d = 0;
        w.phase = WP_LHIT;
        walk_closest(w);
        return true;
    }
    return walk_camera_start(S, ch, w, ref);
}

// Resumes the walk with the answer h to its query; true when the next query
// is pending, false when the walk has ended.
DEV bool walk_resume(const DevScene &S, const Opts &o, const Chains &ch, float dsigma, Walk &w, const TraceHit &h, POut<float> &out) {
    const SRef ref = slot_ref(ch, w.p);
    const int s = ref.s, vs = ref.vs;
    const bool gen = w.mode == WK_GEN;
    if (w.phase == WP_VIS) {
        if (h.prim >= 0) out.valid = false;
        else if (w.visScreen) { ref.screen[0] = out.screen[0]; ref.screen[s] = out.screen[1]; }
        return false;
    }
    if (w.phase == WP_LHIT) {
        if (h.prim < 0 || !re_intersect(S, h.prim, w.r, w.ls.isect)) return false;
        const int prim = w.ls.isect.prim;
        const int d = w.d;
        ref.lgprim[d * vs] = prim;
        float dv;
        if (gen) {
            dv = w.g->uni();
            ref.ldisc[d * vs] = dv;
            w.ls.wi = -w.r.dir;
        } else {
            w.ls.wi = -w.r.dir;
            dv = modulo1(ref.ldisc[d * vs] + walk_disc(w, dsigma));
            ref.ldisc[d * vs] = dv;
        }
        convert_mis(w.r, w.ls);
        w.lastLightPrim = prim;
        if (d == w.l - 2) {
            if (w.c == 1) {
                ShadowQ q;
                connect_to_camera<float, false>(S, o, d, w.ls, prim, dv, w.r.org, out, &q);
                if (!out.valid) return false;
                walk_visibility(w, q, true);
                return true;
            }
            return walk_camera_start(S, ch, w, ref);
        }
        float r0, r1;
        if (gen) { r0 = w.g->uni(); r1 = w.g->uni(); }
        else { r0 = walk_next_u(ch, w, ref); r1 = walk_next_u(ch, w, ref); }
        V3<float> dir, bc;
        if (!bsdf_sampling(S, o, true, !gen, w.ls, prim, dv, r0, r1, w.ls, dir, bc)) return false;
        if (gen) { walk_put(w, ref, r0); walk_put(w, ref, r1); }
        else w.ls.thr = w.ls.thr * ref.lrr[d * vs];
        w.r.org = w.ls.isect.pos;
        w.r.dir = dir;
        w.d = d + 1;
        walk_closest(w);
        return true;
    }
    // WP_CHIT: camera-subpath vertex d
    const int d = w.d;
    if (w.lcpEnd >= 0.0f) {
        // the light-coordinate edge that led here is occluded
        if (h.prim >= 0 && h.t < w.lcpEnd) return false;
        w.lcpEnd = -1.0f;
    }
    const bool hit = h.prim >= 0 && re_intersect(S, h.prim, w.r, w.cs.isect);
    const int prim = hit ? w.cs.isect.prim : -1;
    if (hit) ref.cprim[d * vs] = prim;
    w.cs.wi = -w.r.dir;
    if (hit) convert_mis(w.r, w.cs);
    if (gen) {
        if (d + 2 >= w.c && w.l == 0) {
            if (hit && S.tlight[prim] >= 0) {
                const int li = S.tlight[prim];
                if (d > 1) {
                    float u0, u1;
                    get_sample_param(S, prim, w.cs.isect.pos, u0, u1);
                    ref.u[(w.k - 2) * s] = u0;
                    ref.u[(w.k - 1) * s] = u1;
                    V3<float> toPrev = w.cs.isect.pos - w.r.org;
                    const float d2 = length2(toPrev), id2 = 1.0f / d2, id = sqrtf(id2);
                    toPrev = toPrev * id;
                    w.cs.ssJ *= fabsf(dot(toPrev, w.cs.isect.sn) * id2) * (w.cs.lcJ * (1.0f / S.lightArea[li]));
                }
                handle_hit_light(S, d, li, w.r, w.sx, w.sy, w.cs, out);
            }
            return false;
        }
        if (!hit) return false;
        if (d == 1 && length2(w.cs.isect.pos - w.r.org) <= 0.0f) return false;
        const float dv = w.g->uni();
        ref.cdisc[d * vs] = dv;
        if (d + 2 == w.c) {
            ShadowQ q;
            if (w.l == 1) {
                const int li = pick_light(S, w.g->uni());
                const float r0 = w.g->uni(), r1 = w.g->uni();
                const int lp = pick_light_tri(S, li, w.g->uni());
                *ref.light = li; *ref.lprim = lp;
                walk_put(w, ref, r0); walk_put(w, ref, r1);
                direct_lighting<float, false>(S, o, d, w.cs, w.sx, w.sy, li, lp, r0, r1, prim, out, &q);
            } else {
                connect_vertex<float, false>(S, d, w.l - 2, w.ls, w.lastLightPrim, w.cs, prim, w.sx, w.sy, out, &q);
            }
            if (!out.valid) return false;
            walk_visibility(w, q, false);
            return true;
        }
        float r0 = w.g->uni(), r1 = w.g->uni();
        V3<float> dir, bc;
        if (!bsdf_sampling(S, o, false, false, w.cs, prim, dv, r0, r1, w.cs, dir, bc)) return false;
        walk_put(w, ref, r0); walk_put(w, ref, r1);
        w.r.org = w.cs.isect.pos;
        w.r.dir = dir;
        w.r.minT = kIsectEps;
        w.r.maxT = __builtin_inff();
        w.d = d + 1;
        walk_closest(w);
        return true;
    }
    const int nc = w.c - 1, nl = w.l - 1;
    if (d == nc - 1 && w.l == 0) {
        if (hit) {
            const int li = S.tlight[prim];
            if (li >= 0) handle_hit_light(S, d, li, w.r, w.sx, w.sy, w.cs, out);
        }
        return false;
    }
    if (!hit) return false;
    const float dv = modulo1(ref.cdisc[d * vs] + walk_disc(w, dsigma));
    ref.cdisc[d * vs] = dv;
    if (d == 1 && length2(w.cs.isect.pos - w.r.org) <= 0.0f) return false;
    if (d == nc - 1) {
        ShadowQ q;
        if (w.l == 1) {
            const float r0 = walk_next_u(ch, w, ref), r1 = walk_next_u(ch, w, ref);
            direct_lighting<float, false>(S, o, d, w.cs, w.sx, w.sy, *ref.light, *ref.lprim, r0, r1, prim, out, &q);
        } else {
            connect_vertex<float, false>(S, d, nl - 1, w.ls, w.lastLightPrim, w.cs, prim, w.sx, w.sy, out, &q);
        }
        if (!out.valid) return false;
        walk_visibility(w, q, false);
        return true;
    }
    const float r0 = walk_next_u(ch, w, ref), r1 = walk_next_u(ch, w, ref);
    if (o.lightCoord && d == nc - 2 && w.l == 0) {
        const int np = ref.cprim[(nc - 1) * vs];
        if (np >= 0 && S.tlight[np] >= 0) w.lcp = true;
    }
    V3<float> dir, bc;
    if (w.lcp) {
        ShadowQ q;
        if (!light_coordinate_sampling<float, false>(S, prim, r0, r1, ref.cprim[(nc - 1) * vs], w.cs, dir, &q)) return false;
        w.lcpEnd = (1.0f - kShadowEps) * q.dist;
    } else {
        float a0 = r0, a1 = r1;
        if (!bsdf_sampling(S, o, false, true, w.cs, prim, dv, a0, a1, w.cs, dir, bc)) return false;
    }
    w.cs.thr = w.cs.thr * ref.crr[d * vs];
    w.r.org = w.cs.isect.pos;
    w.r.dir = dir;
    w.r.minT = kIsectEps;
    w.r.maxT = __builtin_inff();
    w.d = d + 1;
    walk_closest(w);
    return true;
}

// Runs the walks of the lanes with `active` set to their end, tracing every
// round's queries at one site (the traversal inlined there: the stage
// kernels trace nowhere else). dsigma: the discrete-dimension sigma.
DEV void walk_run(const DevScene &S, const Opts &o, const Chains &ch, float dsigma, Walk &w, bool active, POut<float> &out) {
    const BvhRef B = bvh_of(S);
    bool q = active && walk_start(S, o, ch, w, out);
    while (__any(q)) {
        if (q) {
            const float org[3] = {w.r.org.x, w.r.org.y, w.r.org.z}, dir[3] = {w.r.dir.x, w.r.dir.y, w.r.dir.z};
            TraceHit h;
            h.prim = trace_impl(B, org, dir, w.r.minT, w.r.maxT, w.qany, h.t);
            q = walk_resume(S, o, ch, dsigma, w, h, out);
        }
    }
}

}  // namespace drmlt

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
#include "dev_path.h"

namespace drmlt {

// Where a walk's primary samples come from: generated from uniforms
// (GenerateSubpath), or the state's own u plus an offset -- N(0, sigma)
// draws, an offset array, or the difference of two states (Green's y*,
// mutation.cpp:210-237).
enum : int { WK_GEN = 0, WK_OFF_NORMAL = 1, WK_OFF_ARRAY = 2, WK_OFF_DIFF = 3 };
// What a walk waits for: a light-subpath hit, a camera-subpath hit, or the
// visibility of a finished contribution.
enum : int { WP_LHIT = 1, WP_CHIT = 2, WP_VIS = 3 };

struct Walk {
    SRef ref;                  // the state evaluated (and, for WK_GEN, written) in place
    int mode, c, l;
    Philox *g;                 // uniforms (WK_GEN); offset and discrete-dimension normals (perturbation)
    float sigma, dsigma;       // WK_OFF_NORMAL sigma; discrete-dimension sigma
    const float *oa;           // WK_OFF_ARRAY: offset k at oa[k * os]
    int os;
    const float *ux, *uy;      // WK_OFF_DIFF: offset k = ux[k * us] - uy[k * us]
    int us;
    int k;                     // next primary-sample slot
    int phase, d;              // what the walk waits for; vertex index of the subpath being built
    RayT<float> r;
    BState<float> ls, cs;
    int lastLightPrim;
    float sx, sy;
    bool lcp;                  // light-coordinate sampling of the last camera edge
    float lcpEnd;              // end of the light-coordinate segment awaiting its answer (< 0: none)
    bool visScreen;            // the deferred contribution also sets the state's screen position
    POut<float> pend;          // contribution awaiting its visibility
    V3<float> qo, qd;          // the pending query
    float qmin, qmax;
    int qany;
};

DEV float walk_next_u(Walk &w) {  // PerturbU::next with the walk's offset
    float off;
    if (w.mode == WK_OFF_NORMAL) off = float(w.sigma * w.g->normal());
    else if (w.mode == WK_OFF_ARRAY) off = w.oa[w.k * w.os];
    else off = w.ux[w.k * w.us] - w.uy[w.k * w.us];
    float *u = w.ref.u + w.k * w.ref.s;
    const float v = modulo1(*u + off);
    *u = v;
    w.k++;
    return v;
}
DEV float walk_disc(Walk &w) { return float(w.dsigma * w.g->normal()); }  // DiscNormal
DEV void walk_put(Walk &w, float v) { w.ref.u[w.k * w.ref.s] = v; w.k++; }  // UWriter::put

DEV void walk_closest(Walk &w) {
    w.qo = w.r.org; w.qd = w.r.dir;
    w.qmin = w.r.minT; w.qmax = w.r.maxT;
    w.qany = 0;
}
// occluded() of dev_scene.h as a query
DEV void walk_visibility(Walk &w, const ShadowQ &q, const POut<float> &pend, bool screen) {
    w.qo = q.o; w.qd = q.d;
    w.qmin = kIsectEps; w.qmax = (1.0f - kShadowEps) * q.dist;
    w.qany = 1;
    w.pend = pend;
    w.visScreen = screen;
    w.phase = WP_VIS;
}

DEV bool walk_camera_start(const DevScene &S, Walk &w) {
    if (w.mode == WK_GEN) {
        const float sx = w.g->uni(), sy = w.g->uni();
        walk_put(w, sx); walk_put(w, sy);
        w.ref.screen[0] = sx; w.ref.screen[w.ref.s] = sy;
        w.sx = sx; w.sy = sy;
    } else {
        w.sx = walk_next_u(w);
        w.sy = walk_next_u(w);
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
DEV bool walk_start(const DevScene &S, const Opts &o, Walk &w, POut<float> &out) {
    out.valid = false;
    const SRef &ref = w.ref;
    const int s = ref.s;
    w.k = 0;
    w.lastLightPrim = -1;
    w.lcpEnd = -1.0f;
    if (w.mode == WK_GEN) {
        *ref.c = w.c; *ref.l = w.l;
        *ref.light = -1; *ref.lprim = -1;
        for (int k = 0; k < kVMax; k++) {
            ref.cprim[k * s] = -1; ref.lgprim[k * s] = -1; ref.crr[k * s] = 1.0f; ref.lrr[k * s] = 1.0f;
            ref.cdisc[k * s] = 0.0f; ref.ldisc[k * s] = 0.0f;
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
            walk_put(w, p0); walk_put(w, p1); walk_put(w, d0); walk_put(w, d1);
        } else {
            li = *ref.light; lp = *ref.lprim;
            p0 = walk_next_u(w); p1 = walk_next_u(w); d0 = walk_next_u(w); d1 = walk_next_u(w);
        }
        emit_from_light(S, o, li, lp, p0, p1, d0, d1, w.r, w.ls);
        w.r.minT = kIsectEps;
        w.r.maxT = __builtin_inff();
        w.d = 0;
        w.phase = WP_LHIT;
        walk_closest(w);
        return true;
    }
    return walk_camera_start(S, w);
}

// Resumes the walk with the answer h to its query; true when the next query
// is pending, false when the walk has ended.
DEV bool walk_resume(const DevScene &S, const Opts &o, Walk &w, const TraceHit &h, POut<float> &out) {
    const SRef &ref = w.ref;
    const int s = ref.s;
    const bool gen = w.mode == WK_GEN;
    if (w.phase == WP_VIS) {
        if (h.prim < 0) {
            out = w.pend;
            if (w.visScreen) { ref.screen[0] = out.screen[0]; ref.screen[s] = out.screen[1]; }
        }
        return false;
    }
    if (w.phase == WP_LHIT) {
        if (h.prim < 0 || !re_intersect(S, h.prim, w.r, w.ls.isect)) return false;
        const int prim = w.ls.isect.prim;
        const int d = w.d;
        ref.lgprim[d * s] = prim;
        float dv;
        if (gen) {
            dv = w.g->uni();
            ref.ldisc[d * s] = dv;
            w.ls.wi = -w.r.dir;
        } else {
            w.ls.wi = -w.r.dir;
            dv = modulo1(ref.ldisc[d * s] + walk_disc(w));
            ref.ldisc[d * s] = dv;
        }
        convert_mis(w.r, w.ls);
        w.lastLightPrim = prim;
        if (d == w.l - 2) {
            if (w.c == 1) {
                ShadowQ q;
                POut<float> pend;
                pend.valid = false;
                connect_to_camera<float, false>(S, o, d, w.ls, prim, dv, w.r.org, pend, &q);
                if (!pend.valid) return false;
                walk_visibility(w, q, pend, true);
                return true;
            }
            return walk_camera_start(S, w);
        }
        float r0, r1;
        if (gen) { r0 = w.g->uni(); r1 = w.g->uni(); }
        else { r0 = walk_next_u(w); r1 = walk_next_u(w); }
        V3<float> dir, bc;
        if (!bsdf_sampling(S, o, true, !gen, w.ls, prim, dv, r0, r1, w.ls, dir, bc)) return false;
        if (gen) { walk_put(w, r0); walk_put(w, r1); }
        else w.ls.thr = w.ls.thr * ref.lrr[d * s];
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
    if (hit) ref.cprim[d * s] = prim;
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
        ref.cdisc[d * s] = dv;
        if (d + 2 == w.c) {
            ShadowQ q;
            POut<float> pend;
            pend.valid = false;
            if (w.l == 1) {
                const int li = pick_light(S, w.g->uni());
                const float r0 = w.g->uni(), r1 = w.g->uni();
                const int lp = pick_light_tri(S, li, w.g->uni());
                *ref.light = li; *ref.lprim = lp;
                walk_put(w, r0); walk_put(w, r1);
                direct_lighting<float, false>(S, o, d, w.cs, w.sx, w.sy, li, lp, r0, r1, prim, pend, &q);
            } else {
                connect_vertex<float, false>(S, d, w.l - 2, w.ls, w.lastLightPrim, w.cs, prim, w.sx, w.sy, pend, &q);
            }
            if (!pend.valid) return false;
            walk_visibility(w, q, pend, false);
            return true;
        }
        float r0 = w.g->uni(), r1 = w.g->uni();
        V3<float> dir, bc;
        if (!bsdf_sampling(S, o, false, false, w.cs, prim, dv, r0, r1, w.cs, dir, bc)) return false;
        walk_put(w, r0); walk_put(w, r1);
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
    const float dv = modulo1(ref.cdisc[d * s] + walk_disc(w));
    ref.cdisc[d * s] = dv;
    if (d == 1 && length2(w.cs.isect.pos - w.r.org) <= 0.0f) return false;
    if (d == nc - 1) {
        ShadowQ q;
        POut<float> pend;
        pend.valid = false;
        if (w.l == 1) {
            const float r0 = walk_next_u(w), r1 = walk_next_u(w);
            direct_lighting<float, false>(S, o, d, w.cs, w.sx, w.sy, *ref.light, *ref.lprim, r0, r1, prim, pend, &q);
        } else {
            connect_vertex<float, false>(S, d, nl - 1, w.ls, w.lastLightPrim, w.cs, prim, w.sx, w.sy, pend, &q);
        }
        if (!pend.valid) return false;
        walk_visibility(w, q, pend, false);
        return true;
    }
    const float r0 = walk_next_u(w), r1 = walk_next_u(w);
    if (o.lightCoord && d == nc - 2 && w.l == 0) {
        const int np = ref.cprim[(nc - 1) * s];
        if (np >= 0 && S.tlight[np] >= 0) w.lcp = true;
    }
    V3<float> dir, bc;
    if (w.lcp) {
        ShadowQ q;
        if (!light_coordinate_sampling<float, false>(S, prim, r0, r1, ref.cprim[(nc - 1) * s], w.cs, dir, &q)) return false;
        w.lcpEnd = (1.0f - kShadowEps) * q.dist;
    } else {
        float a0 = r0, a1 = r1;
        if (!bsdf_sampling(S, o, false, true, w.cs, prim, dv, a0, a1, w.cs, dir, bc)) return false;
    }
    w.cs.thr = w.cs.thr * ref.crr[d * s];
    w.r.org = w.cs.isect.pos;
    w.r.dir = dir;
    w.r.minT = kIsectEps;
    w.r.maxT = __builtin_inff();
    w.d = d + 1;
    walk_closest(w);
    return true;
}

// Runs the walks of the lanes with `active` set to their end, tracing every
// round's queries at one call site.
DEV void walk_run(const DevScene &S, const Opts &o, Walk &w, bool active, POut<float> &out) {
    const BvhRef B = bvh_of(S);
    bool q = active && walk_start(S, o, w, out);
    while (__any(q)) {
        if (q) {
            const TraceHit h = trace_call(B, w.qo.x, w.qo.y, w.qo.z, w.qd.x, w.qd.y, w.qd.z, w.qmin, w.qmax, w.qany);
            q = walk_resume(S, o, w, h, out);
        }
    }
}

}  // namespace drmlt

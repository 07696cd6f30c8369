"""ctypes binding of the oracle (oracle/liborc.so), the CPU restatement the
GPU results are checked against. Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liborc.so")

PATH_REC = np.dtype([
    ("c", np.int32), ("l", np.int32), ("light", np.int32), ("lprim", np.int32),
    ("screen", np.float32, 2), ("u", np.float32, 32),
    ("cprim", np.int32, 16), ("lgprim", np.int32, 16),
    ("cdisc", np.float32, 16), ("ldisc", np.float32, 16),
    ("crr", np.float32, 16), ("lrr", np.float32, 16),
])
assert PATH_REC.itemsize == 536


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.orc_scene_new.restype = C.c_void_p
        _lib.orc_gaussian_logpdf.restype = C.c_double
        _lib.orc_roughness.restype = C.c_double
        assert _lib.orc_path_rec_size() == PATH_REC.itemsize
    return _lib


def P(a):
    return a.ctypes.data_as(C.c_void_p)


class Scene:
    def __init__(self, sc, use_bvh=True):
        self.sc = sc
        L = lib()
        self._keep = [np.ascontiguousarray(sc[k]) for k in
                      ("tri_p", "tri_n", "tri_mat", "tri_light", "mats", "light_tris", "light_emit", "cam")]
        tp, tn, tm, tl, m, lt, le, cam = self._keep
        self.h = L.orc_scene_new(len(tp), P(tp), P(tn), P(tm), P(tl), len(m), P(m), len(lt), P(lt), P(le), P(cam),
                                 1 if use_bvh else 0)
        self.W, self.H = int(cam[10]), int(cam[11])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_free(C.c_void_p(self.h))
            self.h = None

    def trace(self, org, dirs, tmin, tmax):
        n = len(org)
        org = np.ascontiguousarray(org, np.float32); dirs = np.ascontiguousarray(dirs, np.float32)
        tmin = np.ascontiguousarray(tmin, np.float32); tmax = np.ascontiguousarray(tmax, np.float32)
        prim = np.zeros(n, np.int32); t = np.zeros(n, np.float32)
        lib().orc_trace(C.c_void_p(self.h), n, P(org), P(dirs), P(tmin), P(tmax), P(prim), P(t))
        return prim, t

    def occluded(self, org, dirs, tmin, tmax):
        n = len(org)
        org = np.ascontiguousarray(org, np.float32); dirs = np.ascontiguousarray(dirs, np.float32)
        tmin = np.ascontiguousarray(tmin, np.float32); tmax = np.ascontiguousarray(tmax, np.float32)
        occ = np.zeros(n, np.int32)
        lib().orc_occluded(C.c_void_p(self.h), n, P(org), P(dirs), P(tmin), P(tmax), P(occ))
        return occ

    def perturb(self, recs, off, disc, dbl=True, occl=True, fixed=False, chad_quirk=False):
        n = len(recs)
        recs = np.ascontiguousarray(recs, PATH_REC)
        off = np.ascontiguousarray(off, np.float32).reshape(n, 32)
        disc = np.ascontiguousarray(disc, np.float32).reshape(n, 32)
        out = np.zeros(n, PATH_REC); c8 = np.zeros((n, 8), np.float32)
        lib().orc_perturb(C.c_void_p(self.h), int(dbl), n, P(recs), P(off), P(disc), int(occl), int(fixed), int(chad_quirk),
                          P(out), P(c8))
        return out, c8

    def generate_subpath(self, cl, uni, dbl=True):
        n = len(cl)
        cl = np.ascontiguousarray(cl, np.int32).reshape(n, 2)
        uni = np.ascontiguousarray(uni, np.float32).reshape(n, -1)
        out = np.zeros(n, PATH_REC); c8 = np.zeros((n, 8), np.float32)
        lib().orc_generate_subpath(C.c_void_p(self.h), int(dbl), n, P(cl), P(uni), uni.shape[1], P(out), P(c8))
        return out, c8

    def serialize(self, recs, max_depth=9):
        """The reference's serialized form of path records (Serialize,
        path.cpp:2303-2391): primary n x GetPrimaryParamSize(max, max) and
        vertParams n x GetVertParamSize(max, max), as the reference sizes them."""
        recs = np.ascontiguousarray(recs, PATH_REC)
        n = len(recs)
        ps = 2 * max(2 * max_depth - 1, 2) + 1
        vs = 2 * max_depth * (46 + 10 + 3) + 56 + 5
        primary = np.zeros((n, ps), np.float32); vert = np.zeros((n, vs), np.float32)
        lib().orc_serialize(C.c_void_p(self.h), n, P(recs), ps, P(primary), vs, P(vert))
        return primary, vert

    def serialize_scene(self, use_light_coord=True):
        """Serialize(scene) (scene.cpp:126-131): 38 floats."""
        sp = np.zeros(38, np.float32)
        lib().orc_serialize_scene(C.c_void_p(self.h), int(use_light_coord), P(sp))
        return sp

    def generate_bidir(self, uni, min_depth, max_depth, dbl=True, maxc=64):
        uni = np.ascontiguousarray(uni, np.float32)
        n = len(uni)
        cs = np.zeros((n, maxc, 10), np.float32); nc = np.zeros(n, np.int32)
        lib().orc_generate_bidir(C.c_void_p(self.h), int(dbl), n, min_depth, max_depth, P(uni), uni.shape[1], maxc, P(cs), P(nc))
        return cs, nc

    def grad_hess_fd(self, recs, h=1e-4, chad_quirk=False):
        n = len(recs)
        recs = np.ascontiguousarray(recs, PATH_REC)
        logf = np.zeros(n); g = np.zeros((n, 32)); H = np.zeros((n, 32, 32))
        lib().orc_grad_hess_fd(C.c_void_p(self.h), n, P(recs), C.c_double(h), int(chad_quirk), P(logf), P(g), P(H))
        return logf, g, H

    def logf_fixed(self, recs, off, chad_quirk=False):
        n = len(recs)
        recs = np.ascontiguousarray(recs, PATH_REC)
        off = np.ascontiguousarray(off, np.float64).reshape(n, 32)
        out = np.zeros(n)
        lib().orc_logf_fixed(C.c_void_p(self.h), n, P(recs), P(off), int(chad_quirk), P(out))
        return out

    def lens_score(self, recs, dbl=True):
        """lensScore and lensVertexPos of each state's fixed re-evaluation: n x 4."""
        recs = np.ascontiguousarray(recs, PATH_REC)
        out = np.zeros((len(recs), 4), np.float32)
        lib().orc_lens_score(C.c_void_p(self.h), int(dbl), len(recs), P(recs), P(out))
        return out

    def generate_subpath_lens(self, cl, uni, dbl=True):
        """GenerateSubpath's own (tracked) lensScore per sample."""
        cl = np.ascontiguousarray(cl, np.int32); uni = np.ascontiguousarray(uni, np.float32)
        out = np.zeros(len(cl), np.float32)
        lib().orc_generate_subpath_lens(C.c_void_p(self.h), int(dbl), len(cl), P(cl), P(uni), uni.shape[1], P(out))
        return out

    def perturb_lens(self, recs, screen, dbl=True, occl=True, fixed=False):
        """PerturbLensBidir to screen positions (n x 2): states, c9 = contribution 8 + lensScore."""
        recs = np.ascontiguousarray(recs, PATH_REC); n = len(recs)
        screen = np.ascontiguousarray(screen, np.float32).reshape(n, 2)
        out = np.zeros(n, PATH_REC); c9 = np.zeros((n, 9), np.float32)
        lib().orc_perturb_lens(C.c_void_p(self.h), int(dbl), n, P(recs), P(screen), int(occl), int(fixed), P(out), P(c9))
        return out, c9

    def lens_grad_hess_fd(self, recs, screen, h=1e-4):
        recs = np.ascontiguousarray(recs, PATH_REC); n = len(recs)
        screen = np.ascontiguousarray(screen, np.float32).reshape(n, 2)
        logl = np.zeros(n); g = np.zeros((n, 2)); H = np.zeros((n, 2, 2))
        lib().orc_lens_grad_hess_fd(C.c_void_p(self.h), n, P(recs), P(screen), C.c_double(h), P(logl), P(g), P(H))
        return logl, g, H

    def lens_accept(self, recs, c8, off, sigma, fd_step=2e-6, max_derv=9):
        """The oracle loop's lens-mutation acceptance (LensMutateAt) of each
        state for the screen offset off (n x 2), in double."""
        recs = np.ascontiguousarray(recs, PATH_REC); n = len(recs)
        c8 = np.ascontiguousarray(c8, np.float32).reshape(n, 8)
        off = np.ascontiguousarray(off, np.float32).reshape(n, 2)
        a = np.zeros(n)
        lib().orc_lens_accept(C.c_void_p(self.h), n, P(recs), P(c8), P(off), C.c_double(sigma), C.c_double(fd_step), int(max_derv), P(a))
        return a

    def direct_pass(self, spp, seed=1, threads=8):
        film = np.zeros((self.H, self.W, 3), np.float32)
        lib().orc_direct_pass(C.c_void_p(self.h), spp, C.c_uint64(seed), threads, P(film))
        return film

    def prepass(self, num_init, max_depth, num_chains, seed=1, threads=8):
        b = np.zeros(1); nlen = max_depth + 2
        lf = np.zeros(nlen, np.float32)
        seeds = np.zeros(num_chains, PATH_REC); c8 = np.zeros((num_chains, 8), np.float32)
        ncand = lib().orc_prepass(C.c_void_p(self.h), num_init, max_depth, num_chains, C.c_uint64(seed), threads, P(b), P(lf),
                                  nlen, P(seeds), P(c8))
        return float(b[0]), lf, seeds, c8, ncand

    def render(self, seeds, c8, length_func, b, iters, opts=None, threads=8, seed=7):
        o = {"type": 0, "largestepprob": 0.3, "perturbstddev": 0.01, "anisoperturbprob": 0.5, "anisostddev": 0.0025,
             "discretestddev": 0.01, "roughthr": 0.03, "maxdervdepth": 9, "acceptancemap": 0, "chadquirk": 1, "fdstep": 2e-6,
             "secondstage": 1, "h2mc": 0, "uniformmixing": 0.2, "lensperturbprob": 0.0, "lensperturbstddev": 0.01}
        o.update(opts or {})
        ov = np.array([o["type"], o["largestepprob"], o["perturbstddev"], o["anisoperturbprob"], o["anisostddev"],
                       o["discretestddev"], o["roughthr"], o["maxdervdepth"], o["acceptancemap"], o["chadquirk"],
                       o["fdstep"], o["secondstage"], o["h2mc"], o["uniformmixing"], o["lensperturbprob"],
                       o["lensperturbstddev"]],
                      np.float64)
        seeds = np.ascontiguousarray(seeds, PATH_REC); c8 = np.ascontiguousarray(c8, np.float32)
        lf = np.ascontiguousarray(length_func, np.float32)
        film = np.zeros((self.H, self.W, 3), np.float32); stats = np.zeros(10, np.uint64); sec = np.zeros(1)
        lib().orc_render(C.c_void_p(self.h), P(ov), P(seeds), P(c8), len(seeds), P(lf), len(lf), C.c_float(b),
                         C.c_int64(iters), threads, C.c_uint64(seed), P(film), P(stats), P(sec))
        return film, stats, float(sec[0])


def bsdf_eval(mat, wi, n, wo, adjoint=False):
    out = np.zeros(6)
    mat = np.ascontiguousarray(mat, np.float32)
    wi, n, wo = (np.ascontiguousarray(x, np.float64) for x in (wi, n, wo))
    lib().orc_bsdf_eval(P(mat), int(adjoint), P(wi), P(n), P(wo), P(out))
    return out


def bsdf_eval_batch(mat, wi, n, wo, adjoint=False):
    """bsdf_eval over rows of wi and wo (m x 3) with one normal: m x 6."""
    wi = np.ascontiguousarray(np.broadcast_to(wi, np.shape(wo)), np.float64)
    wo = np.ascontiguousarray(wo, np.float64)
    out = np.zeros((len(wo), 6))
    lib().orc_bsdf_eval_batch(P(np.ascontiguousarray(mat, np.float32)), int(adjoint), len(wo), P(wi),
                              P(np.ascontiguousarray(n, np.float64)), P(wo), P(out))
    return out


class planted:
    """Context manager: the oracle's BSDF code with mistake `which` planted
    (orc_core.h, g_plant), removed on exit."""

    def __init__(self, which):
        self.which = which

    def __enter__(self):
        lib().orc_plant(int(self.which))
        return self

    def __exit__(self, *exc):
        lib().orc_plant(0)
        return False


def bsdf_sample(mat, wi, n, r0, r1, disc, adjoint=False):
    out = np.zeros(6); wo = np.zeros(3)
    mat = np.ascontiguousarray(mat, np.float32)
    wi, n = (np.ascontiguousarray(x, np.float64) for x in (wi, n))
    ok = lib().orc_bsdf_sample(P(mat), int(adjoint), P(wi), P(n), C.c_double(r0), C.c_double(r1), C.c_double(disc), P(wo), P(out))
    return bool(ok), wo, out


def compute_gaussian(Ds, sc, sigma, grad, hess):
    n = len(Ds)
    Ds = np.ascontiguousarray(Ds, np.int32); sc = np.ascontiguousarray(sc, np.float64)
    grad = np.ascontiguousarray(grad, np.float64).reshape(n, 32); hess = np.ascontiguousarray(hess, np.float64).reshape(n, 32, 32)
    mean = np.zeros((n, 32)); covL = np.zeros((n, 32, 32)); invCov = np.zeros((n, 32, 32)); logDet = np.zeros(n)
    lib().orc_compute_gaussian(n, P(Ds), P(sc), C.c_double(sigma), P(grad), P(hess), P(mean), P(covL), P(invCov), P(logDet))
    return mean, covL, invCov, logDet


def gaussian_logpdf(D, x, mean, invCov32, logDet):
    x = np.ascontiguousarray(x, np.float64); mean = np.ascontiguousarray(mean, np.float64)
    invCov32 = np.ascontiguousarray(invCov32, np.float64)
    return lib().orc_gaussian_logpdf(D, P(x), P(mean), P(invCov32), C.c_double(logDet))


def gaussian_sample(D, nrm, mean, covL32):
    nrm = np.ascontiguousarray(nrm, np.float64); mean = np.ascontiguousarray(mean, np.float64)
    covL32 = np.ascontiguousarray(covL32, np.float64)
    x = np.zeros(D)
    lib().orc_gaussian_sample(D, P(nrm), P(mean), P(covL32), P(x))
    return x


def jacobi_eigen(A):
    A = np.ascontiguousarray(A, np.float64); D = len(A)
    lam = np.zeros(D); V = np.zeros((D, D))
    lib().orc_jacobi_eigen(D, P(A), P(lam), P(V))
    return lam, V

"""ctypes binding of libdrmlt_hip.so (include/drmlt_amd.h).

The library is the product: every call here goes to the HIP implementation.
Loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libdrmlt_hip.so")

EXPORTS = [
    "drmlt_last_error", "drmlt_options_default", "drmlt_ctx_create", "drmlt_ctx_destroy", "drmlt_prepass",
    "drmlt_seed_chains", "drmlt_direct_lighting", "drmlt_run", "drmlt_read_film", "drmlt_read_stats", "drmlt_read_timing",
    "drmlt_count_work", "drmlt_trace", "drmlt_perturb", "drmlt_generate_subpath", "drmlt_generate_bidir",
    "drmlt_grad_hess", "drmlt_grad_hess_noalloc", "drmlt_lens_grad_hess", "drmlt_compute_gaussian_f64", "drmlt_gaussian_logpdf_f64", "drmlt_dr_alpha2_f64",
    "drmlt_get_schedule", "drmlt_set_schedule", "drmlt_autotune", "drmlt_eval_logf", "drmlt_read_chains", "drmlt_read_step", "drmlt_read_lens", "drmlt_gaussian_eval_f64",
    "drmlt_path_func_derv", "drmlt_serialized_to_rec", "drmlt_material_from_serialized", "drmlt_bvh_info",
    "drmlt_write_exr", "drmlt_stats_text", "drmlt_write_stats", "drmlt_render",
    "drmlt_scene_load", "drmlt_scene_get", "drmlt_scene_output", "drmlt_scene_free",
]

PLAIN, BATCHED, BATCHED_GRAPH = 0, 1, 2
INTEGRATORS = {"drmlt": 0, "mcmc": 1, "h2mc": 2}  # the <dpt> integrator values the device path runs
EXR_NONE, EXR_ZIP = 0, 3

PATH_REC = np.dtype([
    ("c", np.int32), ("l", np.int32), ("light", np.int32), ("lprim", np.int32),
    ("screen", np.float32, 2), ("u", np.float32, 32),
    ("cprim", np.int32, 16), ("lgprim", np.int32, 16),
    ("cdisc", np.float32, 16), ("ldisc", np.float32, 16),
    ("crr", np.float32, 16), ("lrr", np.float32, 16),
])


class SceneDesc(C.Structure):
    _fields_ = [("num_triangles", C.c_int32), ("tri_p", C.c_void_p), ("tri_n", C.c_void_p), ("tri_mat", C.c_void_p),
                ("tri_light", C.c_void_p), ("num_materials", C.c_int32), ("materials", C.c_void_p),
                ("num_lights", C.c_int32), ("light_tris", C.c_void_p), ("light_emit", C.c_void_p), ("camera", C.c_void_p)]


class Options(C.Structure):
    _fields_ = [("spp", C.c_int32), ("num_init_samples", C.c_int32), ("min_depth", C.c_int32), ("max_depth", C.c_int32),
                ("direct_spp", C.c_int32), ("large_step_prob", C.c_float), ("perturb_std_dev", C.c_float),
                ("roughness_threshold", C.c_float), ("num_chains", C.c_int32), ("seed_offset", C.c_int32),
                ("use_light_coordinate_sampling", C.c_int32), ("large_step_multiplexed", C.c_int32),
                ("discrete_std_dev", C.c_float), ("type", C.c_int32), ("aniso_perturb_prob", C.c_float),
                ("aniso_perturb_std_dev", C.c_float), ("acceptance_map", C.c_int32),
                ("max_derivatives_depth", C.c_int32), ("quirk_ignore_aniso_std_dev", C.c_int32),
                ("quirk_chad_mis", C.c_int32), ("integrator", C.c_int32), ("uniform_mixing_prob", C.c_float),
                ("lens_perturb_prob", C.c_float), ("seed", C.c_uint64), ("host_bvh", C.c_int32),
                ("lens_perturb_std_dev", C.c_float)]


class Schedule(C.Structure):
    _fields_ = [("bin_stages", C.c_int32), ("stage_blocks", C.c_int32), ("deriv_blocks", C.c_int32),
                ("gauss_blocks", C.c_int32), ("stage_lanes", C.c_int32),
                ("stage_split", C.c_int32), ("stage_waves", C.c_int32)]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("large_step", "large_step_accepted", "stage1", "stage1_accepted", "stage2_iso",
                                          "stage2_iso_accepted", "stage2_h2mc", "stage2_h2mc_accepted", "lens_step",
                                          "lens_step_accepted")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class Report(C.Structure):
    _fields_ = [("interval_seconds", C.c_double), ("partial_dir", C.c_char_p), ("base_name", C.c_char_p),
                ("compression", C.c_int32), ("reference_labels", C.c_int32), ("chunk_iters", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _lib.drmlt_last_error.restype = C.c_char_p
        _lib.drmlt_ctx_create.argtypes = [C.POINTER(SceneDesc), C.POINTER(Options), C.POINTER(C.c_void_p)]
    return _lib


class DrmltError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise DrmltError(f"drmlt error {rc}: {lib().drmlt_last_error().decode()}")


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def default_options():
    o = Options()
    check(lib().drmlt_options_default(C.byref(o)))
    return o


# <dpt> keys of the reference (src/parsescene.cpp:505-564) -> Options fields
_KEYMAP = {
    "spp": ("spp", int), "numinitsamples": ("num_init_samples", int), "mindepth": ("min_depth", int),
    "maxdepth": ("max_depth", int), "directspp": ("direct_spp", int), "largestepprob": ("large_step_prob", float),
    "perturbstddev": ("perturb_std_dev", float), "roughnessthreshold": ("roughness_threshold", float),
    "numchains": ("num_chains", int), "seedoffset": ("seed_offset", int),
    "uselightcoordinatesampling": ("use_light_coordinate_sampling", int),
    "largestepmultiplexed": ("large_step_multiplexed", int), "anisoperturbprob": ("aniso_perturb_prob", float),
    "anisoperturbstddev": ("aniso_perturb_std_dev", float), "acceptanceMap": ("acceptance_map", int),
    "uniformmixingprobability": ("uniform_mixing_prob", float), "lensperturbprob": ("lens_perturb_prob", float),
    "lensperturbstddev": ("lens_perturb_std_dev", float),
}


def options_from_dpt(dpt: dict, **overrides):
    """Options from a dict of the reference's <dpt> keys (types as the XML
    parser converts them); 'type' is "green" or "mira"."""
    o = default_options()
    for k, v in dpt.items():
        if k == "type":
            o.type = 1 if v == "mira" else 0
        elif k == "integrator":
            o.integrator = INTEGRATORS.get(v, -1)
        elif k in _KEYMAP:
            f, cast = _KEYMAP[k]
            setattr(o, f, cast(bool(v) if isinstance(v, bool) else v))
            if k == "anisoperturbstddev":  # the reference's parser writes lensPerturbStdDev (parsescene.cpp:553-554)
                o.lens_perturb_std_dev = float(v)
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


class Context:
    """One drmlt_ctx: the scene on the GPU plus the chain state."""

    def __init__(self, scene: dict, options: Options):
        self._keep = [np.ascontiguousarray(scene[k]) for k in
                      ("tri_p", "tri_n", "tri_mat", "tri_light", "mats", "light_tris", "light_emit", "cam")]
        tp, tn, tm, tl, m, lt, le, cam = self._keep
        assert tp.dtype == np.float32 and tm.dtype == np.int32 and le.dtype == np.float32
        d = SceneDesc(len(tp), P(tp), P(tn), P(tm), P(tl), len(m), P(m), len(lt), P(lt), P(le), P(cam))
        self.opt = options
        h = C.c_void_p()
        check(lib().drmlt_ctx_create(C.byref(d), C.byref(options), C.byref(h)))
        self.h = h
        self.W, self.H = int(cam[10]), int(cam[11])

    def bvh_info(self):
        """The context's BVH: built on the device, build ms, nodes, 4-wide depth, stack bound."""
        dev, n, d4, st = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        ms = C.c_double()
        check(lib().drmlt_bvh_info(self.h, C.byref(dev), C.byref(ms), C.byref(n), C.byref(d4), C.byref(st)))
        return {"on_device": bool(dev.value), "build_ms": ms.value, "nodes": n.value, "depth4": d4.value, "stack_need": st.value}

    def close(self):
        if getattr(self, "h", None):
            lib().drmlt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- loop
    def prepass(self, n_len=16):
        b = C.c_double(0)
        lf = np.zeros(n_len, np.float32)
        check(lib().drmlt_prepass(self.h, C.byref(b), P(lf), n_len))
        return b.value, lf

    def seed_chains(self, states, contrib8, b, length_func):
        states = np.ascontiguousarray(states, PATH_REC)
        contrib8 = np.ascontiguousarray(contrib8, np.float32)
        lf = np.ascontiguousarray(length_func, np.float32)
        check(lib().drmlt_seed_chains(self.h, P(states), P(contrib8), C.c_int64(len(states)), C.c_float(b), P(lf), len(lf)))

    def direct_lighting(self):
        check(lib().drmlt_direct_lighting(self.h))

    def run(self, iters, mode=BATCHED_GRAPH):
        check(lib().drmlt_run(self.h, C.c_int64(iters), mode))

    def film(self, which=0):
        out = np.zeros((self.H, self.W, 3), np.float32)
        check(lib().drmlt_read_film(self.h, P(out), self.W, self.H, which))
        return out

    def stats(self):
        s = Stats()
        check(lib().drmlt_read_stats(self.h, C.byref(s)))
        return s.as_dict()

    def write_stats(self, path, reference_labels=True):
        """The reference's statistics text (Stats::output) of the counters so far."""
        check(lib().drmlt_write_stats(self.h, os.fsencode(path), int(reference_labels)))

    def render(self, iters, mode=BATCHED_GRAPH, interval_seconds=0.0, partial_dir=None, base_name="drmlt",
               compression=EXR_ZIP, reference_labels=True, chunk_iters=0):
        """run() with the reference's intermediate outputs in partial_dir:
        <base>_<k>.exr, <base>_time.csv and <base>_stats.txt."""
        rep = Report(float(interval_seconds), os.fsencode(partial_dir) if partial_dir else None,
                     base_name.encode() if base_name else None, int(compression), int(reference_labels), int(chunk_iters))
        check(lib().drmlt_render(self.h, C.c_int64(iters), mode, C.byref(rep)))

    def count_work(self):
        """One batched step with work counters: per stage (nodes, triangles, rays, state bytes)."""
        out = np.zeros((7, 4), np.uint64)
        check(lib().drmlt_count_work(self.h, P(out)))
        return out.astype(np.int64)

    def schedule(self):
        sc = Schedule()
        check(lib().drmlt_get_schedule(self.h, C.byref(sc)))
        return sc.as_dict()

    def set_schedule(self, **fields):
        sc = Schedule()
        check(lib().drmlt_get_schedule(self.h, C.byref(sc)))
        for k, v in fields.items():
            setattr(sc, k, int(v))
        check(lib().drmlt_set_schedule(self.h, C.byref(sc)))

    def autotune(self, steps=3):
        """Times candidate launch schedules on this context's chains and keeps
        the fastest; returns (schedule dict, ms per step)."""
        sc = Schedule()
        ms = C.c_double(0)
        check(lib().drmlt_autotune(self.h, int(steps), C.byref(sc), C.byref(ms)))
        return sc.as_dict(), ms.value

    def timing(self):
        out = np.zeros(8)
        check(lib().drmlt_read_timing(self.h, P(out), 8))
        return out

    # -- function level
    def trace(self, org, dirs, tmin, tmax, any_hit=False):
        n = len(org)
        org = np.ascontiguousarray(org, np.float32); dirs = np.ascontiguousarray(dirs, np.float32)
        tmin = np.ascontiguousarray(tmin, np.float32); tmax = np.ascontiguousarray(tmax, np.float32)
        prim = np.zeros(n, np.int32); t = np.zeros(n, np.float32)
        check(lib().drmlt_trace(self.h, C.c_int64(n), P(org), P(dirs), P(tmin), P(tmax), int(any_hit), P(prim), P(t)))
        return prim, t

    def perturb(self, states, off, disc):
        states = np.array(states, PATH_REC)
        n = len(states)
        off = np.ascontiguousarray(off, np.float32).reshape(n, 32)
        disc = np.ascontiguousarray(disc, np.float32).reshape(n, 32)
        c8 = np.zeros((n, 8), np.float32)
        check(lib().drmlt_perturb(self.h, C.c_int64(n), P(states), P(off), P(disc), P(c8)))
        return states, c8

    def generate_subpath(self, cl, uni):
        cl = np.ascontiguousarray(cl, np.int32)
        n = len(cl)
        uni = np.ascontiguousarray(uni, np.float32).reshape(n, -1)
        st = np.zeros(n, PATH_REC); c8 = np.zeros((n, 8), np.float32)
        check(lib().drmlt_generate_subpath(self.h, C.c_int64(n), P(cl), P(uni), uni.shape[1], P(st), P(c8)))
        return st, c8

    def generate_bidir(self, uni, min_depth, max_depth, maxc=64):
        uni = np.ascontiguousarray(uni, np.float32)
        n = len(uni)
        cs = np.zeros((n, maxc, 10), np.float32); nc = np.zeros(n, np.int32)
        check(lib().drmlt_generate_bidir(self.h, C.c_int64(n), min_depth, max_depth, P(uni), uni.shape[1], maxc, P(cs), P(nc)))
        return cs, nc

    def read_chains(self, n=None):
        """Current states of chains 0..n-1 as records, with contrib8 rows."""
        n = int(self.opt.num_chains if n is None else n)
        st = np.zeros(n, PATH_REC)
        c8 = np.zeros((n, 8), np.float32)
        check(lib().drmlt_read_chains(self.h, C.c_int64(n), P(st), P(c8)))
        return st, c8

    def read_step(self, n=None):
        """What the last iteration left in chains 0..n-1: a dict with the four
        physical slots (n x 4 records and n x 4 x 8 contrib8 rows), perm,
        flags, a1, alpha2 and the H2MC offset (n x 32; after a lens
        mutation, flag bit 7, its screen offset in the first two entries)."""
        n = int(self.opt.num_chains if n is None else n)
        slots = np.zeros((n, 4), PATH_REC)
        c8 = np.zeros((n, 4, 8), np.float32)
        perm = np.zeros(n, np.int32); flags = np.zeros(n, np.int32); a1 = np.zeros(n, np.float32)
        a2 = np.zeros(n); off = np.zeros((n, 32), np.float32)
        check(lib().drmlt_read_step(self.h, C.c_int64(n), P(slots), P(c8), P(perm), P(flags), P(a1), P(a2), P(off)))
        return {"slots": slots, "contrib8": c8, "perm": perm, "flags": flags, "a1": a1, "alpha2": a2, "off": off}

    def read_lens(self, n=None):
        """The lens score the loop holds for each of the four physical slots
        of chains 0..n-1, its lens flags (bit 0 score known, bit 1 lens
        Gaussian initialised) and the lensVertexPos held with it:
        (n x 4 floats, n x 4 ints, n x 4 x 3 floats)."""
        n = int(self.opt.num_chains if n is None else n)
        ls = np.zeros((n, 4), np.float32)
        lf = np.zeros((n, 4), np.int32)
        lp = np.zeros((n, 4, 3), np.float32)
        check(lib().drmlt_read_lens(self.h, C.c_int64(n), P(ls), P(lf), P(lp)))
        return ls, lf, lp

    def eval_logf(self, states, fp64=False):
        """log f at each state with its triangles fixed (the function
        grad_hess differentiates); returns (logf, ok)."""
        states = np.ascontiguousarray(states, PATH_REC)
        n = len(states)
        lf = np.zeros(n)
        ok = np.zeros(n, np.int32)
        check(lib().drmlt_eval_logf(self.h, C.c_int64(n), P(states), int(fp64), P(lf), P(ok)))
        return lf, ok

    def lens_grad_hess(self, states, screen, lens_pos=None):
        """Lens score and lensVertexPos of each state, and the gradient and
        Hessian of log lensScore in the screen position at `screen` (n x 2),
        about the lensVertexPos `lens_pos` (n x 3; None: the re-evaluated
        one): returns (n x 9 floats, ok)."""
        states = np.ascontiguousarray(states, PATH_REC)
        n = len(states)
        screen = np.ascontiguousarray(screen, np.float32).reshape(n, 2)
        lp = None if lens_pos is None else np.ascontiguousarray(lens_pos, np.float32).reshape(n, 3)
        out = np.zeros((n, 9), np.float32)
        ok = np.zeros(n, np.int32)
        check(lib().drmlt_lens_grad_hess(self.h, C.c_int64(n), P(states), P(screen), None if lp is None else P(lp), P(out), P(ok)))
        return out, ok

    def grad_hess(self, states, fp64=False):
        states = np.ascontiguousarray(states, PATH_REC)
        n = len(states)
        g = np.zeros((n, 32)); h = np.zeros((n, 32, 32)); ok = np.zeros(n, np.int32)
        check(lib().drmlt_grad_hess(self.h, C.c_int64(n), P(states), int(fp64), P(g), P(h), P(ok)))
        return g, h, ok

    def grad_hess_noalloc(self, states, fp64=False):
        """grad_hess through the entry point that allocates no device memory
        (at most 4096 states per call)."""
        states = np.ascontiguousarray(states, PATH_REC)
        n = len(states)
        g = np.zeros((n, 32)); h = np.zeros((n, 32, 32)); ok = np.zeros(n, np.int32)
        check(lib().drmlt_grad_hess_noalloc(self.h, C.c_int64(n), P(states), int(fp64), P(g), P(h), P(ok)))
        return g, h, ok

    def path_func_derv(self, c, l, lens, primary, scene_params, vert_params):
        """The reference's PathFuncDerv call, batched over the rows of lens
        (n x 2), primary (n x >= 2 max(c + l - 1, 2) + 1) and vert_params
        (n x V), for one (c, l): grad n x D, hess n x D x D (float32) and ok."""
        lens = np.ascontiguousarray(lens, np.float32).reshape(-1, 2)
        primary = np.ascontiguousarray(primary, np.float32)
        vert = np.ascontiguousarray(vert_params, np.float32)
        sp = np.ascontiguousarray(scene_params, np.float32)
        n = len(lens)
        D = 2 * max(c + l - 1, 2)
        g = np.zeros((n, D), np.float32); h = np.zeros((n, D, D), np.float32); ok = np.zeros(n, np.int32)
        check(lib().drmlt_path_func_derv(self.h, int(c), int(l), C.c_int64(n), P(lens), P(primary), C.c_int64(primary.shape[1]), P(sp),
                                         P(vert), C.c_int64(vert.shape[1]), P(g), P(h), P(ok)))
        return g, h, ok


def serialized_to_rec(scene, c, l, lens, primary, vert_params):
    """Host only: the reference's serialized states (the drmlt_path_func_derv
    layout) as path records of `scene` (a scene dict)."""
    keep = [np.ascontiguousarray(scene[k]) for k in ("tri_p", "tri_n", "tri_mat", "tri_light", "mats", "light_tris", "light_emit", "cam")]
    tp, tn, tm, tl, m, lt, le, cam = keep
    d = SceneDesc(len(tp), P(tp), P(tn), P(tm), P(tl), len(m), P(m), len(lt), P(lt), P(le), P(cam))
    lens = np.ascontiguousarray(lens, np.float32).reshape(-1, 2)
    primary = np.ascontiguousarray(primary, np.float32)
    vert = np.ascontiguousarray(vert_params, np.float32)
    n = len(lens)
    out = np.zeros(n, PATH_REC)
    check(lib().drmlt_serialized_to_rec(C.byref(d), int(c), int(l), C.c_int64(n), P(lens), P(primary), C.c_int64(primary.shape[1]),
                                        P(vert), C.c_int64(vert.shape[1]), P(out)))
    return out


def material_from_serialized(bsdf, two_sided=False):
    """A serialized BSDF record (10 floats) as a 16-float material."""
    b = np.ascontiguousarray(bsdf, np.float32)
    m = np.zeros(16, np.float32)
    check(lib().drmlt_material_from_serialized(P(b), int(two_sided), P(m)))
    return m


def compute_gaussian(D, sc, sigma, grad, hess):
    D = np.ascontiguousarray(D, np.int32)
    n = len(D)
    sc = np.ascontiguousarray(sc, np.float64)
    grad = np.ascontiguousarray(grad, np.float64).reshape(n, 32)
    hess = np.ascontiguousarray(hess, np.float64).reshape(n, 32, 32)
    mean = np.zeros((n, 32)); covL = np.zeros((n, 24, 24)); inv = np.zeros((n, 24, 24)); ld = np.zeros(n)
    check(lib().drmlt_compute_gaussian_f64(C.c_int64(n), P(D), P(sc), C.c_double(sigma), P(grad), P(hess), P(mean), P(covL),
                                           P(inv), P(ld)))
    return mean, covL, inv, ld


def gaussian_logpdf(D, x, sign, mean, covL, inv, ld):
    D = np.ascontiguousarray(D, np.int32)
    n = len(D)
    x = np.ascontiguousarray(x, np.float64).reshape(n, 32)
    out = np.zeros(n)
    check(lib().drmlt_gaussian_logpdf_f64(C.c_int64(n), P(D), P(x), C.c_double(sign), P(np.ascontiguousarray(mean)),
                                          P(np.ascontiguousarray(covL)), P(np.ascontiguousarray(inv)),
                                          P(np.ascontiguousarray(ld)), P(out)))
    return out


def gaussian_eval(D, sc, sigma, grad, hess, x, sign, nrm):
    """The loop's own ComputeGaussian -> GaussianLogPdf(sign x) ->
    GenerateSample(nrm) chain on the device: (logpdf n, sample n x 32)."""
    D = np.ascontiguousarray(D, np.int32)
    n = len(D)
    arr = lambda a, shape: np.ascontiguousarray(a, np.float64).reshape(shape)
    sc, grad, hess = arr(sc, n), arr(grad, (n, 32)), arr(hess, (n, 32, 32))
    x, nrm = arr(x, (n, 32)), arr(nrm, (n, 32))
    lp = np.zeros(n); sm = np.zeros((n, 32))
    check(lib().drmlt_gaussian_eval_f64(C.c_int64(n), P(D), P(sc), C.c_double(sigma), P(grad), P(hess), P(x), C.c_double(sign), P(nrm),
                                        P(lp), P(sm)))
    return lp, sm


def dr_alpha2(rows):
    rows = np.ascontiguousarray(rows, np.float64).reshape(-1, 6)
    out = np.zeros(len(rows))
    check(lib().drmlt_dr_alpha2_f64(C.c_int64(len(rows)), P(rows), P(out)))
    return out


def write_exr(path, rgb, compression=EXR_ZIP):
    """An H x W x 3 float image as a HALF RGB OpenEXR file (the reference's WriteImage)."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    if rgb.ndim != 3 or rgb.shape[2] != 3:
        raise ValueError("expected an H x W x 3 image")
    check(lib().drmlt_write_exr(os.fsencode(path), P(rgb), rgb.shape[1], rgb.shape[0], int(compression)))


def stats_text(stats: dict, reference_labels=True, integrator="drmlt"):
    """Stats::output of a stats dict (Context.stats()) as the reference prints
    it: the DR integrator's text (with its swapped labels, or corrected ones),
    or the mcmc / h2mc integrators' (small steps: first-stage plus H2MC moves)."""
    s = Stats(*[int(stats.get(n, 0)) for n, _ in Stats._fields_])
    fmt = 2 if integrator in ("mcmc", "h2mc") else int(bool(reference_labels))
    n = lib().drmlt_stats_text(C.byref(s), fmt, None, 0)
    if n < 0:
        check(n)
    buf = C.create_string_buffer(n + 1)
    lib().drmlt_stats_text(C.byref(s), fmt, buf, n + 1)
    return buf.value.decode()


def load_scene(xml_path):
    """A scene file in the reference's XML format (drmlt_scene_load): the
    scene dict Context takes (arrays copied out of the loaded scene), with its
    <dpt> options as an Options struct under "opt", the output image name,
    the integrator name and reportIntervalSeconds."""
    L = lib()
    h = C.c_void_p()
    check(L.drmlt_scene_load(os.fsencode(xml_path), C.byref(h)))
    try:
        d, o = SceneDesc(), Options()
        check(L.drmlt_scene_get(h, C.byref(d), C.byref(o)))
        n = L.drmlt_scene_output(h, None, 0, None, 0, None)
        if n < 0:
            check(n)
        out, integ, rep = C.create_string_buffer(n + 1), C.create_string_buffer(n + 1), C.c_int32()
        check(min(0, L.drmlt_scene_output(h, out, n + 1, integ, n + 1, C.byref(rep))))

        def arr(ptr, count, dtype, cols):
            if count == 0:
                return np.zeros((0, cols) if cols > 1 else 0, dtype)
            buf = (C.c_char * (count * cols * np.dtype(dtype).itemsize)).from_address(ptr)
            a = np.frombuffer(buf, dtype).copy()
            return a.reshape(count, cols) if cols > 1 else a

        nt, nm, nl = d.num_triangles, d.num_materials, d.num_lights
        sc = {"tri_p": arr(d.tri_p, nt, np.float32, 9), "tri_n": arr(d.tri_n, nt, np.float32, 9),
              "tri_mat": arr(d.tri_mat, nt, np.int32, 1), "tri_light": arr(d.tri_light, nt, np.int32, 1),
              "mats": arr(d.materials, nm, np.float32, 16), "light_tris": arr(d.light_tris, nl, np.int32, 2),
              "light_emit": arr(d.light_emit, nl, np.float32, 4), "cam": arr(d.camera, 14, np.float32, 1),
              "opt": o, "output": out.value.decode(), "integrator": integ.value.decode(),
              "report_interval_seconds": int(rep.value)}
        return sc
    finally:
        L.drmlt_scene_free(h)

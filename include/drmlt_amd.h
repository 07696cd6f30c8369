/*
 * drmlt_amd.h -- C ABI of libdrmlt_hip.so, the MI355X (gfx950) implementation
 * of the DRMLT data-parallel mutation loop of joeylitalien/drmlt-dpt.
 *
 * Two tiers:
 *  1. Loop level: replaces the chain loop of DRIntegrator::Render
 *     (src/integrators/delayed/delayed.cpp:334-673) together with its
 *     pre-pass (delayed.cpp:14-173) and direct-lighting pass
 *     (src/integrator.h:33-74). A host integrator creates a context from a
 *     scene and the reference's DptOptions (src/dptoptions.h:8-37), seeds
 *     the chains, runs iterations and reads back the film and statistics.
 *  2. Function level: batched versions of the functions the loop is made of,
 *     for parity tests and for callers that drive their own loop.
 * Around them, on the host: scene files in the reference's XML format
 * (drmlt_scene_*) and the output files of a render (drmlt_write_exr,
 * drmlt_render, drmlt_write_stats).
 *
 * Conventions: every function returns an int status (DRMLT_OK = 0, negative
 * on error; drmlt_last_error() describes the last failure). Buffers are
 * caller-owned host memory unless a name says otherwise; functions copy in and
 * out. A context owns one HIP stream on the device current at creation and is
 * not thread-safe (one host thread drives it). No function allocates on behalf
 * of the caller beyond the context's own device state and the host object of
 * a loaded scene (released by drmlt_scene_free).
 */
#ifndef DRMLT_AMD_H
#define DRMLT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRMLT_OK 0
#define DRMLT_ERR_ARG (-1)
#define DRMLT_ERR_HIP (-2)
#define DRMLT_ERR_STATE (-3)

typedef struct drmlt_ctx drmlt_ctx;

/* Scene as flat arrays (the reference parses Mitsuba-style XML into
 * TriangleMesh/BSDF/AreaLight objects, src/parsescene.cpp:136-503).
 * tri_p, tri_n: 9 floats per triangle (three positions / three shading
 * normals). materials: 16 floats each -- type (0 Lambertian, 1 Phong,
 * 2 rough dielectric), Kd|Ks (3), Ks|Kt (3), Phong exponent, dielectric
 * alpha, eta = intIOR/extIOR, 1/eta, Phong KsWeight, twoSided.
 * light_tris: (first triangle, triangle count) per area light; light_emit:
 * radiance RGB and sampling weight per light. camera: position, target, up,
 * horizontal fov in degrees, width, height, near clip, far clip. */
typedef struct {
    int32_t num_triangles;
    const float *tri_p;
    const float *tri_n;
    const int32_t *tri_mat;
    const int32_t *tri_light; /* area-light index of an emissive triangle, else -1 */
    int32_t num_materials;
    const float *materials;
    int32_t num_lights;
    const int32_t *light_tris;
    const float *light_emit;
    const float *camera; /* 14 floats */
} drmlt_scene_desc;

/* The reference's DptOptions (src/dptoptions.h:8-37) plus the command-line
 * knobs that reach the loop (src/main.cpp:28-42). */
typedef struct {
    int32_t spp;                 /* mutations per pixel */
    int32_t num_init_samples;    /* pre-pass samples */
    int32_t min_depth, max_depth;
    int32_t direct_spp;
    float large_step_prob;
    float perturb_std_dev;       /* stage-1 sigma and H2MC prior sigma */
    float roughness_threshold;
    int32_t num_chains;
    int32_t seed_offset;
    int32_t use_light_coordinate_sampling;
    int32_t large_step_multiplexed;  /* only multiplexed large steps are implemented */
    float discrete_std_dev;
    int32_t type;                /* 0 = "green", 1 = "mira" */
    float aniso_perturb_prob;
    float aniso_perturb_std_dev; /* the stage-2 isotropic sigma */
    int32_t acceptance_map;
    int32_t max_derivatives_depth;   /* --max-derivatives-depth (src/main.cpp:37, default 9) */
    /* The reference stores the XML key anisoperturbstddev into lensPerturbStdDev
     * (src/parsescene.cpp:553-554), so the loop always sees 0.0025. 1 keeps
     * that behaviour (aniso_perturb_std_dev is ignored), 0 honours the option. */
    int32_t quirk_ignore_aniso_std_dev;
    /* The reference's derivative code weights area-light emission MIS as for a
     * point light (src/path.cpp:2598-2604). 1 differentiates that function
     * (what the reference's H2MC sees), 0 differentiates the primal log f. */
    int32_t quirk_chad_mis;
    /* The <dpt> integrator: 0 "drmlt" (DRIntegrator, delayed.cpp), 1 "mcmc"
     * (MCMCIntegrator, src/integrators/mcmc.cpp:11-244: the loop's first
     * stage alone -- a large or isotropic small step, splats weighted 1 - a
     * and a, acceptance with probability a; no acceptance map), 2 "h2mc"
     * (H2MCIntegrator, src/integrators/h2mc.cpp:196-600: a small step is
     * isotropic with probability uniform_mixing_prob; otherwise, when the
     * current state has a lens score, a lens mutation (PerturbLensBidir,
     * path.cpp:1914-2067) with probability lens_perturb_prob, else one H2MC
     * move; each accepted with min(1, ratio)). */
    int32_t integrator;
    float uniform_mixing_prob;  /* uniformmixingprobability (dptoptions.h:29, h2mc only) */
    float lens_perturb_prob;    /* lensperturbprob (dptoptions.h:20, h2mc only) */
    uint64_t seed;
    /* 0 (default): the context builds its BVH on the device; 1: on the host.
     * Both run the same binned-SAH build and answer every ray query alike
     * (the reference builds with Embree, src/scene.cpp:42-54). */
    int32_t host_bvh;
    /* lensperturbstddev (dptoptions.h:21, h2mc only): sigma of the lens
     * Gaussian. The reference's XML key anisoperturbstddev also writes it
     * (parsescene.cpp:553-554); the scene loader does the same. */
    float lens_perturb_std_dev;
} drmlt_options;

/* One path state (the reference's Path + SubpathContrib, src/path.h:12-107):
 * camera / light subpath vertex counts c, l (SubpathContrib camDepth /
 * lightDepth); the emitting light and triangle (l >= 2) or the direct-light
 * light and triangle (l == 1); the contribution's screen position; the
 * primary sample u in GetPathPos order (src/path.cpp:2393-2439); per surface
 * vertex the hit triangle, bsdfDiscrete and Russian-roulette weight. 536 bytes. */
typedef struct {
    int32_t c, l, light, lprim;
    float screen[2];
    float u[32];
    int32_t cprim[16], lgprim[16];
    float cdisc[16], ldisc[16];
    float crr[16], lrr[16];
} drmlt_path_rec;

/* delayed.cpp:287-329 (Stats), counted per iteration */
typedef struct {
    uint64_t large_step, large_step_accepted;
    uint64_t stage1, stage1_accepted;
    uint64_t stage2_iso, stage2_iso_accepted;
    uint64_t stage2_h2mc, stage2_h2mc_accepted;
    /* h2mc integrator: lens mutations and their acceptances (h2mc.cpp:109-115) */
    uint64_t lens_step, lens_step_accepted;
} drmlt_stats;

/* Launch schedule of the batched form. The chains advance identically under
 * every schedule (the same random streams, the same arithmetic); a schedule
 * only changes speed. The reference has no counterpart: its chains run as
 * CPU threads (src/parallel.cpp). Zero fields take the defaults. */
typedef struct {
    int32_t bin_stages;    /* 1: the path-stage kernels take their chains from path-length bins
                              (waves of equal-length paths); 2: only the second-stage kernels do
                              (compacted lists of the chains with a second stage), stage 1 in
                              chain order; 0 (default): all in chain order */
    int32_t stage_blocks;  /* 64-lane workgroups of the path-stage kernels (default: one per 64 chains, at most 16 per CU) */
    int32_t deriv_blocks;  /* workgroups of the derivative kernel (default 16 per CU) */
    int32_t gauss_blocks;  /* one-wave workgroups of the Gaussian kernel (default 16 per CU) */
    int32_t stage_lanes;   /* chains per 64-lane workgroup of the path-stage kernels, 1..64 (default
                              64); fewer helps when the chains fill few waves per SIMD */
    int32_t stage_split;   /* 2..4: the graph-replayed step runs that many contiguous ranges of the
                              chains on as many streams that fork and join around it (one range's
                              kernel tails overlap the others' work); 0 or 1 (default 0): one
                              stream. The other forms ignore it */
    int32_t stage_waves;   /* resident waves per SIMD the path-stage kernels are built for: 3 (168
                              VGPRs, some spills, more traversal latency hidden) or 2 (0, default:
                              256 VGPRs); 3 pays when the chains fill more than two waves per SIMD */
} drmlt_schedule;

/* Execution forms of drmlt_run */
#define DRMLT_PLAIN 0          /* one kernel launch per step, one thread per chain iteration */
#define DRMLT_BATCHED 1        /* one persistent kernel per DR stage, under the context's schedule */
#define DRMLT_BATCHED_GRAPH 2  /* DRMLT_BATCHED with each step replayed from a captured hipGraph */

const char *drmlt_last_error(void);
int drmlt_options_default(drmlt_options *opt);

/* ---------------- tier 1: the loop ---------------- */

/* Replaces Scene::Scene + the integrator construction (src/main.cpp:101-128,
 * src/scene.cpp:10-54): uploads the scene and builds the BVH. */
int drmlt_ctx_create(const drmlt_scene_desc *scene, const drmlt_options *opt, drmlt_ctx **out);
int drmlt_ctx_destroy(drmlt_ctx *ctx);

/* The context's BVH: whether it was built on the device (1) or the host (0),
 * its build time in milliseconds (wall clock, triangle arrays already on the
 * host), its 4-wide node count, 4-wide depth and the largest traversal stack
 * a ray can need. No reference counterpart (Embree builds internally). */
int drmlt_bvh_info(drmlt_ctx *ctx, int32_t *on_device, double *build_ms, int32_t *nodes, int32_t *depth4, int32_t *stack_need);

/* DRIntegrator::PrePass (delayed.cpp:14-173) on the device: num_init_samples
 * bidirectional paths, the normalisation b, the path-length distribution and
 * num_chains equal-spaced seed states. */
int drmlt_prepass(drmlt_ctx *ctx, double *b_out, float *length_func_out, int32_t n_len);

/* Seeds the chains from caller-provided states instead (for example the
 * states a reference pre-pass produced). contrib8 per state: valid, rgb,
 * lsScore, ssScore, screen x, y. */
int drmlt_seed_chains(drmlt_ctx *ctx, const drmlt_path_rec *states, const float *contrib8, int64_t n, float b,
                      const float *length_func, int32_t n_len);

/* DirectLighting (src/integrator.h:33-74) into the context's direct buffer. */
int drmlt_direct_lighting(drmlt_ctx *ctx);

/* Runs iters_per_chain DR iterations on every chain (delayed.cpp:396-656). */
int drmlt_run(drmlt_ctx *ctx, int64_t iters_per_chain, int32_t mode);

/* which: 0 merged film (direct/directSpp + indirect/spp with spp the
 * mutations per pixel run so far, delayed.cpp:682-685), 1 raw indirect
 * buffer, 2 raw direct buffer. rgb holds w*h*3 floats. */
int drmlt_read_film(drmlt_ctx *ctx, float *rgb, int32_t w, int32_t h, int32_t which);
int drmlt_read_stats(drmlt_ctx *ctx, drmlt_stats *stats);
/* ---------------- scene files (host) ---------------- */

/* A scene loaded from the reference's XML format; owns the arrays its
 * drmlt_scene_desc points to. */
typedef struct drmlt_scene drmlt_scene;

/* Replaces ParseScene (src/parsescene.cpp:589-613) with its OBJ and
 * .serialized loaders (src/parseobj.cpp, src/loadserialized.cpp) for what
 * the device path renders: <sensor type="perspective"> with a look-at
 * toWorld and <film> width/height/filename; <bsdf> diffuse, phong,
 * roughdielectric (constant parameters) and twosided, by id or inline;
 * <shape> obj and serialized (shapeIndex) with toWorld transforms
 * (scale, translate, rotate, lookat, matrix) and <emitter type="area">
 * radiance; <dpt> options. Relative file names resolve against the XML
 * file's directory. Textures, point and environment emitters and moving
 * shapes are rejected (DRMLT_ERR_ARG, drmlt_last_error names them). */
int drmlt_scene_load(const char *xml_path, drmlt_scene **out);
/* The scene's arrays (valid until drmlt_scene_free) and its <dpt> options
 * applied over drmlt_options_default. Either output may be NULL. */
int drmlt_scene_get(const drmlt_scene *scene, drmlt_scene_desc *desc, drmlt_options *opt);
/* The output image name (the film's filename, else <xml dir>/image.exr),
 * the integrator name and reportIntervalSeconds; strings are NUL-terminated
 * and truncated to cap - 1; returns the longer string's full length. */
int drmlt_scene_output(const drmlt_scene *scene, char *output_name, int32_t cap, char *integrator, int32_t icap,
                       int32_t *report_interval_seconds);
int drmlt_scene_free(drmlt_scene *scene);

/* ---------------- output files (host) ---------------- */

/* OpenEXR compression codes accepted by the writers */
#define DRMLT_EXR_NONE 0
#define DRMLT_EXR_ZIP 3

/* Replaces WriteImage (src/image.cpp:45-58, OpenImageIO): writes a w x h RGB
 * float image (row-major, top row first) as a scanline OpenEXR file with
 * three HALF channels. Host only; no context needed. */
int drmlt_write_exr(const char *path, const float *rgb, int32_t w, int32_t h, int32_t compression);

/* Stats::output (delayed.cpp:311-324) as text into buf (NUL-terminated,
 * truncated to cap - 1 bytes); returns the full length. format 1 keeps the
 * reference's swapped second-stage labels (delayed.cpp:320-321), 0 names
 * each line's own component, 2 is the mcmc and h2mc integrators' text
 * (mcmc.cpp:76-86, h2mc.cpp:140-150: large steps, small steps -- first-stage
 * plus H2MC moves -- and lens steps). */
int drmlt_stats_text(const drmlt_stats *stats, int32_t format, char *buf, int32_t cap);
/* The same text of the context's counters into a file (truncated); a context
 * of the mcmc or h2mc integrator writes format 2 whatever reference_labels says. */
int drmlt_write_stats(drmlt_ctx *ctx, const char *path, int32_t reference_labels);

/* Intermediate outputs of drmlt_render (delayed.cpp:225-242, 610-655). */
typedef struct {
    double interval_seconds;   /* DptOptions reportIntervalSeconds; <= 0: no partial images */
    const char *partial_dir;   /* <output dir>/<integrator or output name>_partial; NULL: no files */
    const char *base_name;     /* the integrator name ("drmlt") or the output name */
    int32_t compression;       /* DRMLT_EXR_NONE or DRMLT_EXR_ZIP */
    int32_t reference_labels;  /* stats file labels, as drmlt_stats_text */
    int32_t chunk_iters;       /* iterations per chain between report checks (0: 16) */
} drmlt_report;

/* drmlt_run with the reference's intermediate outputs: partial images of
 * the indirect buffer (scaled by pixels / mutations so far) as
 * <dir>/<base>_<k>.exr once interval_seconds x k have passed, their times in
 * <dir>/<base>_time.csv, and the statistics in <dir>/<base>_stats.txt. */
int drmlt_render(drmlt_ctx *ctx, int64_t iters_per_chain, int32_t mode, const drmlt_report *report);

/* Device time of the last drmlt_run in milliseconds (after a function-level
 * call that says so: of its kernel, in out[0]), total and per stage
 * kernel (hipEvents on the context stream): out[0] total, out[1..7] stage 1,
 * derivatives (current), Gaussian (current), stage-2 proposal, derivatives
 * (proposal), Gaussian (proposal), stage-2 acceptance. n_out <= 8. */
int drmlt_read_timing(drmlt_ctx *ctx, double *out, int32_t n_out);
/* Runs one batched DR step (advancing every chain by one iteration) with work
 * counters on: out is 7 x 4 words, one row per stage kernel in the order of
 * drmlt_read_timing, holding BVH nodes visited, triangles tested, rays traced
 * and algorithmic chain-state bytes. */
int drmlt_count_work(drmlt_ctx *ctx, uint64_t *out);

/* The batched form's launch schedule. Setting one drops the captured step
 * graph (it is captured again at the next DRMLT_BATCHED_GRAPH run). */
int drmlt_get_schedule(drmlt_ctx *ctx, drmlt_schedule *out);
int drmlt_set_schedule(drmlt_ctx *ctx, const drmlt_schedule *sched);
/* Autotuner: times candidate schedules on the context's own chains (each
 * candidate one untimed and `steps` timed batched steps, hipEvents on the
 * context stream; stage_split by the best of two rounds of max(steps, 10)
 * graph-replayed steps per choice, then the derivative and Gaussian
 * workgroup counts once more against their defaults under that split) and keeps the fastest. The steps are ordinary DR iterations:
 * they advance the chains, count in the statistics and splat into the film.
 * best and best_ms (both optional) receive the schedule and its time per step. */
int drmlt_autotune(drmlt_ctx *ctx, int32_t steps, drmlt_schedule *best, double *best_ms);

/* The current state of chains 0..n-1 as records (the Markov states the
 * loop holds; delayed.cpp:343-345 keeps them per chain), contrib8 as in
 * drmlt_seed_chains (valid flag, rgb, lsScore, ssScore, screen x, y). */
int drmlt_read_chains(drmlt_ctx *ctx, int64_t n, drmlt_path_rec *states, float *contrib8);

/* What the last iteration left in chains 0..n-1, for per-state checks of the
 * loop (the reference's per-chain MarkovStates and DR variables,
 * delayed.cpp:343-371): the four physical state slots of each chain as
 * records (slots n x 4; slot 3 holds Green's y*) with their contrib8 rows
 * (n x 4 x 8, [0] the slot's valid flag); perm (2 bits each: the slot of the
 * current state, of the first- and of the second-stage proposal); flags
 * (bit 0 first stage accepted, 1 second stage run, 2 H2MC second stage,
 * 3 large step, 4 second-stage proposal valid, 5 first-stage proposal
 * valid, 6 y* traced and valid, 7 lens mutation, h2mc.cpp:463-530); a1;
 * alpha2, the second-stage acceptance (0 where no second stage ran); off, the
 * H2MC offset (n x 32), or after a lens mutation its screen offset in
 * off[0..1].
 * Read before an iteration, perm names the roles the slots take in it. */
int drmlt_read_step(drmlt_ctx *ctx, int64_t n, drmlt_path_rec *slots, float *contrib8, int32_t *perm, int32_t *flags, float *a1,
                    double *alpha2, float *off);

/* The h2mc integrator's per-state lens bookkeeping of the four physical slots
 * of chains 0..n-1 (the lensScore of the reference's Contrib and the
 * MarkovState's lens Gaussian flag, h2mc.cpp:463-530): lscore (n x 4), the
 * lens score the loop holds for the slot's state, and lflags (n x 4), bit 0
 * set where that score is known (computed or carried over from a lens
 * mutation), bit 1 where the state's lens Gaussian is initialised; and
 * lens_pos (n x 4 x 3), the lensVertexPos held with that score (zeros for
 * slot 3, which holds none, and where no lens mutation is enabled). */
int drmlt_read_lens(drmlt_ctx *ctx, int64_t n, float *lscore, int32_t *lflags, float *lens_pos);

/* ---------------- tier 2: batched functions ---------------- */

/* Scene::Intersect / Occluded (src/scene.cpp:80-111). any = 0: closest hit,
 * prim[i] = scene triangle index or -1 and t[i]; any = 1: prim[i] = 1 if any
 * triangle is hit in (tmin, tmax). */
int drmlt_trace(drmlt_ctx *ctx, int64_t n, const float *org, const float *dir, const float *tmin, const float *tmax, int32_t any,
                int32_t *prim, float *t);

/* PerturbPathBidir (src/path.cpp:1480-1628) on n states: off (n x 32) are the
 * primary-sample offsets in GetPathPos order, disc (n x 32) the discrete
 * offsets (camera vertex d at d, light vertex d at 16 + d). states are
 * updated in place; contrib8 as in drmlt_seed_chains. */
int drmlt_perturb(drmlt_ctx *ctx, int64_t n, drmlt_path_rec *states, const float *off, const float *disc, float *contrib8);

/* GenerateSubpath (src/path.cpp:1123-1269) for strategies cl (n x 2: c, l)
 * drawing from the uniforms uni (n x nuni) in the documented order.
 * Currently switched off: returns DRMLT_ERR_STATE after argument checks (its
 * batched kernel made a GPU memory fault whose cause is not established).
 * The same device routine runs inside drmlt_run's large steps. */
int drmlt_generate_subpath(drmlt_ctx *ctx, int64_t n, const int32_t *cl, const float *uni, int32_t nuni, drmlt_path_rec *states,
                           float *contrib8);

/* GeneratePathBidir (src/path.cpp:978-1121): up to maxc contributions per
 * sample, 10 floats each (c, l, rgb, lsScore, ssScore, screen x, y, MIS
 * weight); ncontrib[i] is the number produced. */
int drmlt_generate_bidir(drmlt_ctx *ctx, int64_t n, int32_t min_depth, int32_t max_depth, const float *uni, int32_t nuni, int32_t maxc,
                         float *contribs, int32_t *ncontrib);

/* log f at each state with its triangles fixed: the primal of the
 * reference's chad library, evaluate_path_bidir_{c}_{l}_static
 * (src/chad.cpp:183-238, logLumValue of src/path.cpp:3159), i.e. the
 * function drmlt_grad_hess differentiates. ok[i] = 0 where the state has no
 * contribution. fp64 = 1 evaluates in double. Afterwards drmlt_read_timing's
 * out[0] is the device time of the kernel. */
int drmlt_eval_logf(drmlt_ctx *ctx, int64_t n, const drmlt_path_rec *states, int32_t fp64, double *logf, int32_t *ok);

/* Gradient and Hessian of log f (the function of evaluate_path_bidir_{c}_{l}
 * _static_derv, src/path.cpp:2987-3175 / src/chad.cpp:494-551) at each state,
 * triangles fixed. grad n x 32, hess n x 32 x 32 (row-major, leading
 * dimension 32); ok[i] = 0 where the fixed-topology evaluation fails.
 * fp64 = 1 evaluates in double. Afterwards drmlt_read_timing's out[0] is the
 * device time of the derivative kernel. */
int drmlt_grad_hess(drmlt_ctx *ctx, int64_t n, const drmlt_path_rec *states, int32_t fp64, double *grad, double *hess, int32_t *ok);

/* drmlt_grad_hess for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It has no counterpart in the reference. Same arguments and
 * results as drmlt_grad_hess, at most 4096 states per call (more:
 * DRMLT_ERR_ARG); it calls no hipMalloc or hipFree and stages the records,
 * the pass table and the results in one fixed device array of 51.6 MiB per
 * GPU, held from the library's load until the process ends. Calls that
 * overlap on one GPU share that array and race. */
int drmlt_grad_hess_noalloc(drmlt_ctx *ctx, int64_t n, const drmlt_path_rec *states, int32_t fp64, double *grad, double *hess, int32_t *ok);

/* The h2mc integrator's lens derivative functions (chad's
 * evaluate_path_bidir_{c}_{l}_lens _derv, PathFuncMode::Lens, path.cpp:
 * 2987-3175; call site h2mc.cpp:473-481), batched: for each record its lens
 * score and lensVertexPos (the fixed-topology re-evaluation that the lens
 * mutation reads) and the gradient and Hessian of log lensScore in the
 * screen position at screen[2i], screen[2i + 1], FP32, about the
 * lensVertexPos lens_pos[3i..3i+2] (n x 3; NULL: the re-evaluated one).
 * out9[9i..] = lensScore, lensVertexPos (3, re-evaluated), d/ds0, d/ds1,
 * H00, H01, H11; ok[i] = 0 where the lens function has no contribution.
 * Every vertex must name a triangle. */
int drmlt_lens_grad_hess(drmlt_ctx *ctx, int64_t n, const drmlt_path_rec *states, const float *screen, const float *lens_pos, float *out9,
                         int32_t *ok);

/* The reference's derivative call with its own argument set, batched:
 * PathFuncDerv(lens, primary, sceneParams, vertParams, grad, hess)
 * (src/path.h:145-149) as DRMLT's H2MC step calls it
 * (src/integrators/delayed/drsmallstep.cpp:140-148), for n states of one
 * (camera, light) vertex count pair (the key of the reference's
 * staticFuncDervMap). State i reads lens + 2 i (the contribution's screen
 * position), primary + primary_stride i (the time, then the primary sample in
 * GetPathPos order; primary_stride >= GetPrimaryParamSize(c, l) =
 * 2 max(c + l - 1, 2) + 1) and vert_params + vert_stride i, written by the
 * reference's Serialize(scene, path, subPath) (src/path.cpp:2303-2391; the
 * reference sizes it GetVertParamSize(maxDervDepth, maxDervDepth),
 * path.cpp:2291-2301). scene_params is Serialize(scene)
 * (src/scene.cpp:126-131, 38 floats), checked against the context's scene
 * and options. The serialized triangles are matched to the context's by the
 * exact bits of their (p0, p1 - p0, p2 - p0); the BSDF and light records
 * are checked against the triangles' materials and area lights, and the
 * values are the context's own (the scene the context was created from).
 * Out: grad n x D, hess n x D x D (row-major), D = 2 max(c + l - 1, 2), the
 * derivatives of log luminance with respect to primary[1..D] (the static
 * function differentiates all primary entries but the time), computed in
 * FP32 like the reference's Float; a non-finite result is written as zeros,
 * as the reference's caller does (drsmallstep.cpp:149-155), and ok[i] = 0
 * there. Needs c + l - 1 <= 12. */
int drmlt_path_func_derv(drmlt_ctx *ctx, int32_t cam_depth, int32_t lgt_depth, int64_t n, const float *lens, const float *primary,
                         int64_t primary_stride, const float *scene_params, const float *vert_params, int64_t vert_stride, float *grad,
                         float *hess, int32_t *ok);

/* Host only: the serialized states drmlt_path_func_derv reads, as path
 * records of the scene `scene` (the arrays a context would be created from),
 * with the same matching and checks. Fields the serialized layout does not
 * carry take the record defaults: bsdfDiscrete of the last camera vertex 0,
 * the Russian-roulette weight of each subpath's last vertex 1. */
int drmlt_serialized_to_rec(const drmlt_scene_desc *scene, int32_t cam_depth, int32_t lgt_depth, int64_t n, const float *lens,
                            const float *primary, int64_t primary_stride, const float *vert_params, int64_t vert_stride,
                            drmlt_path_rec *states);

/* Host only: a serialized BSDF record (BSDF::Serialize: type, then Lambertian
 * Kd; Phong Kd, Ks, exponent, KsWeight; RoughDielectric Ks, Kt, eta, invEta,
 * alpha -- src/lambertian.cpp:10-13, src/phong.cpp:13-19,
 * src/roughdielectric.cpp:13-20) as a 16-float material of
 * drmlt_scene_desc, with the given twoSided flag (the reference keeps it out
 * of the record). For scene flattening on the reference side
 * (include/drmlt_reference_flatten.hpp). */
int drmlt_material_from_serialized(const float *bsdf, int32_t two_sided, float *material);

/* PH2MCSmallStep::ComputeGaussian (drsmallstep.cpp:18-93, 219-266) in FP64
 * on the current device: D[i] dimensions, sc[i] the state's ssScore, sigma the
 * prior, grad n x 32, hess n x 32 x 32. Outputs mean n x 32, covL and invCov
 * n x 24 x 24 (leading dimension 24), logDet n. */
int drmlt_compute_gaussian_f64(int64_t n, const int32_t *D, const double *sc, double sigma, const double *grad, const double *hess,
                               double *mean, double *covL, double *invCov, double *logDet);

/* GaussianLogPdf (gaussian.cpp:18-26) of sign * x[i] under the Gaussians
 * laid out as drmlt_compute_gaussian_f64 returns them. */
int drmlt_gaussian_logpdf_f64(int64_t n, const int32_t *D, const double *x, double sign, const double *mean, const double *covL,
                              const double *invCov, const double *logDet, double *out);

/* The loop's own Gaussian code on n inputs (the device functions its H2MC
 * kernels run; dimensions <= 8 in groups of 8 lanes, as there):
 * ComputeGaussian (drsmallstep.cpp:18-93, 219-266) from D, sc, sigma, grad
 * n x 32 and hess n x 32 x 32 as drmlt_compute_gaussian_f64 takes them, then
 * GaussianLogPdf (gaussian.cpp:18-26) of sign * x[i] (x n x 32) into logpdf
 * and GenerateSample (gaussian.cpp:36-47) from the standard normals nrm
 * (n x 32) into sample (n x 32). */
int drmlt_gaussian_eval_f64(int64_t n, const int32_t *D, const double *sc, double sigma, const double *grad, const double *hess,
                            const double *x, double sign, const double *nrm, double *logpdf, double *sample);

/* Second-stage acceptance of delayed.cpp:466-541 with the loop's own
 * function: per row (type, a1, a2, aStar, transRatio, secondValid) ->
 * alpha2. */
int drmlt_dr_alpha2_f64(int64_t n, const double *in6, double *alpha2);

#ifdef __cplusplus
}
#endif

#endif /* DRMLT_AMD_H */

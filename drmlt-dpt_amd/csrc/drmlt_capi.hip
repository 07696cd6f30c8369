// C ABI of libdrmlt_hip.so (declared in include/drmlt_amd.h): context and
// scene upload, pre-pass / direct pass / DR loop orchestration on one HIP
// stream, hipGraph capture of the batched step, and the batched
// function-level parity surface.
#include "../../include/drmlt_amd.h"
#include "bvh_build.h"
#include "host_io.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "kernels_decl.h"
#include "serialized.h"
#include "bvh_device.h"

using namespace drmlt;

static thread_local std::string g_err;
static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
// for the host-side loaders in other translation units
int drmlt_set_error(int code, const char *msg) { return fail(code, msg); }
#define HIPCHK(x)                                                                                          \
    do {                                                                                                   \
        hipError_t e_ = (x);                                                                               \
        if (e_ != hipSuccess) return fail(DRMLT_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T> static hipError_t dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void **)p, n * sizeof(T));
}
template <class T> static hipError_t upload(T **p, const T *src, size_t n, hipStream_t s) {
    hipError_t e = dalloc(p, n);
    if (e != hipSuccess) return e;
    if (n) e = hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, s);
    return e;
}
// device buffer freed when it goes out of scope (also on an early error return)
template <class T> struct DBuf {
    T *p = nullptr;
    ~DBuf() { if (p) (void)hipFree(p); }
};

// PiecewiseConstant1D cdf (distribution.h:8-33) in float
static void build_cdf(const float *f, int n, std::vector<float> &cdf, float &funcInt) {
    cdf.assign(n + 1, 0.0f);
    for (int i = 1; i < n + 1; i++) cdf[i] = cdf[i - 1] + f[i - 1] / float(n);
    funcInt = cdf[n];
    if (funcInt == 0.0f) {
        for (int i = 1; i < n + 1; i++) cdf[i] = float(i) / float(n);
    } else {
        for (int i = 1; i < n + 1; i++) cdf[i] /= funcInt;
    }
}

// 4x4 helpers for the camera (camera.cpp:11-28, transform.cpp:4-90)
static void m4mul(const double a[16], const double b[16], double r[16]) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += a[4 * i + k] * b[4 * k + j];
            r[4 * i + j] = s;
        }
}
static void m4inv(const double m[16], double o[16]) {
    double a[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++) a[i][j] = j < 4 ? m[4 * i + j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        for (int j = 0; j < 8; j++) std::swap(a[c][j], a[p][j]);
        const double d = a[c][c];
        for (int j = 0; j < 8; j++) a[c][j] /= d;
        for (int r = 0; r < 4; r++)
            if (r != c) {
                const double f = a[r][c];
                for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
            }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) o[4 * i + j] = a[i][j + 4];
}
static DevCamera make_camera(const float *cam) {
    const double pos[3] = {cam[0], cam[1], cam[2]}, tgt[3] = {cam[3], cam[4], cam[5]}, up[3] = {cam[6], cam[7], cam[8]};
    const double fov = cam[9];
    const int W = int(cam[10]), H = int(cam[11]);
    const double nearC = cam[12], farC = cam[13];
    const double aspect = double(W) / double(H);
    const double S[16] = {-0.5, 0, 0, 0, 0, -0.5 * aspect, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const double T[16] = {1, 0, 0, -1.0, 0, 1, 0, -1.0 / aspect, 0, 0, 1, 0, 0, 0, 0, 1};
    const double recip = 1.0 / (farC - nearC);
    const double cot = 1.0 / std::tan((M_PI / 180.0) * (fov / 2.0));
    const double Pm[16] = {cot, 0, 0, 0, 0, cot, 0, 0, 0, 0, farC * recip, -nearC * farC * recip, 0, 0, 1, 0};
    double ST[16], c2s[16], s2c[16];
    m4mul(S, T, ST);
    m4mul(ST, Pm, c2s);
    m4inv(c2s, s2c);
    double d[3] = {tgt[0] - pos[0], tgt[1] - pos[1], tgt[2] - pos[2]};
    const double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (double &x : d) x /= dl;
    const double ul = std::sqrt(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
    const double un[3] = {up[0] / ul, up[1] / ul, up[2] / ul};
    double left[3] = {un[1] * d[2] - un[2] * d[1], un[2] * d[0] - un[0] * d[2], un[0] * d[1] - un[1] * d[0]};
    const double ll = std::sqrt(left[0] * left[0] + left[1] * left[1] + left[2] * left[2]);
    for (double &x : left) x /= ll;
    const double nu[3] = {d[1] * left[2] - d[2] * left[1], d[2] * left[0] - d[0] * left[2], d[0] * left[1] - d[1] * left[0]};
    const double c2w[16] = {left[0], nu[0], d[0], pos[0], left[1], nu[1], d[1], pos[1], left[2], nu[2], d[2], pos[2], 0, 0, 0, 1};
    double w2c[16];
    m4inv(c2w, w2c);
    DevCamera c;
    for (int i = 0; i < 16; i++) {
        c.s2c[i] = float(s2c[i]);
        c.c2s[i] = float(c2s[i]);
        c.c2w[i] = float(c2w[i]);
        c.w2c[i] = float(w2c[i]);
    }
    c.dist = float(W / (2.0 * std::tan((fov / 2.0) * (M_PI / 180.0))));
    c.nearClip = float(nearC);
    c.farClip = float(farC);
    c.W = W;
    c.H = H;
    return c;
}

// ---------------------------------------------------------------- no-alloc scratch
// Device memory of drmlt_grad_hess_noalloc: one array per GPU, held from the
// library's load until the process ends, shared by every call on that GPU.
// On the host the array's name stands for a host-side placeholder, not the
// device copy, so host code never hands it to a kernel or a copy: it uses the
// device address hipGetSymbolAddress gives, looked up once per device when
// the first context is created on it.
constexpr int64_t kNoAllocStates = 4096;  // most states per drmlt_grad_hess_noalloc call
constexpr size_t na_align(size_t x) { return (x + 255) / 256 * 256; }
// layout: records | pass offsets | ok flags | FP32 gradient | FP32 Hessian | FP64 gradient | FP64 Hessian
constexpr size_t kNaRec = 0;
constexpr size_t kNaPass = na_align(kNaRec + size_t(kNoAllocStates) * sizeof(drmlt_path_rec));
constexpr size_t kNaOk = na_align(kNaPass + size_t(kNoAllocStates + 1) * 4);
constexpr size_t kNaG32 = na_align(kNaOk + size_t(kNoAllocStates) * 4);
constexpr size_t kNaH32 = na_align(kNaG32 + size_t(kNoAllocStates) * 32 * 4);
constexpr size_t kNaG64 = na_align(kNaH32 + size_t(kNoAllocStates) * 1024 * 4);
constexpr size_t kNaH64 = na_align(kNaG64 + size_t(kNoAllocStates) * 32 * 8);
constexpr size_t kScratchBytes = na_align(kNaH64 + size_t(kNoAllocStates) * 1024 * 8);  // 51.6 MiB
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static char *g_scratch_ptr[kMaxDevices];
// The device address of `scratch` on `device`, which must be current.
static int scratch_address(int device, char **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return fail(DRMLT_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(DRMLT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<char *>(p);
    }
    *out = g_scratch_ptr[device];
    return DRMLT_OK;
}

// Fills n words with v (a kernel, not a memset).
__global__ void __launch_bounds__(256) k_na_fill(uint32_t *p, size_t n, uint32_t v) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}
// The derivative kernels' output (gradient, upper Hessian triangle) as FP64
// gradient rows and full, mirrored Hessians: FP32 results are converted,
// FP64 results get their lower triangle filled in place.
__global__ void __launch_bounds__(256) k_na_finish(int n, int fp64, const float *g32, const float *h32, double *g64, double *h64) {
    const size_t total = size_t(n) * (32 + 1024);
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += size_t(gridDim.x) * blockDim.x) {
        const size_t r = t / (32 + 1024);
        const int e = int(t % (32 + 1024));
        if (e < 32) {
            if (!fp64) g64[r * 32 + e] = g32[r * 32 + e];
            continue;
        }
        const int a = (e - 32) / 32, b = (e - 32) % 32;
        if (fp64) {
            if (b < a) h64[(r * 32 + a) * 32 + b] = h64[(r * 32 + b) * 32 + a];
        } else {
            h64[(r * 32 + a) * 32 + b] = h32[(r * 32 + (a < b ? a : b)) * 32 + (a < b ? b : a)];
        }
    }
}
static unsigned na_grid(size_t n) { return unsigned(std::max<size_t>(1, std::min<size_t>(2048, (n + 255) / 256))); }

struct drmlt_ctx {
    int device = 0;
    int cus = 256;
    char *scratch = nullptr;  // this device's address of `scratch`
    hipStream_t stream = nullptr;
    drmlt_options opt;
    DevScene S;
    std::vector<void *> sceneAllocs;
    int W = 0, H = 0, nt = 0;
    int bvhDepth = 0;
    // chains
    bool haveChains = false;
    Chains ch;
    std::vector<void *> chainAllocs;
    std::vector<size_t> chainBytes;
    Lists L;
    uint32_t *iterDev = nullptr;
    uint32_t iterHost = 0;
    int64_t itersDone = 0;
    // length distribution
    float *lenFunc = nullptr, *lenCdf = nullptr;
    float lenFuncInt = 0;
    int nLen = 0;
    float b = 0;
    // films and stats
    float *filmInd = nullptr, *filmDir = nullptr;
    unsigned long long *stats = nullptr;
    bool directDone = false;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double timing[8] = {0};
    hipEvent_t ev[9];
    bool evInit = false;
    // parts 1.. of the split step: their streams (default priority), join
    // events and work lists (part 0 runs on `stream` with `L`)
    hipStream_t streamX[kMaxParts - 1] = {};
    hipEvent_t evFork = nullptr, evJoin[kMaxParts - 1] = {};
    Lists LX[kMaxParts - 1];
    // launch schedule of the batched step (drmlt_set_schedule / drmlt_autotune)
    drmlt_schedule sched = {0, 0, 0, 0, 0, 0, 0};
    // the scene's triangles by serialized key, their material types and
    // area lights (drmlt_path_func_derv; built on first use)
    int nmat = 0;
    // the BVH: built on the device (1) or the host (0), its build time
    // (milliseconds, wall clock), 4-wide nodes and traversal stack bound
    int bvhOnDevice = 0, bvhNodes = 0, bvhStack = 0;
    double bvhMs = 0;
    TriKeyMap triMap;
    std::vector<int> triMatType, triLight;
};

static DrParams make_params(const drmlt_ctx *c) {
    DrParams P;
    const drmlt_options &o = c->opt;
    P.largeStepProb = o.large_step_prob;
    P.perturbStdDev = o.perturb_std_dev;
    P.anisoPerturbProb = o.aniso_perturb_prob;
    P.anisoStdDev = o.quirk_ignore_aniso_std_dev ? 0.0025f : o.aniso_perturb_std_dev;
    P.discreteStdDev = o.discrete_std_dev;
    P.type = o.type;
    P.maxDervDepth = o.max_derivatives_depth;
    P.integrator = o.integrator;
    P.uniformMixing = o.uniform_mixing_prob;
    P.lensProb = o.integrator == 2 ? o.lens_perturb_prob : 0.0f;
    P.h2lens = make_h2param(o.lens_perturb_std_dev);
    P.acceptanceMap = o.integrator == 0 ? o.acceptance_map : 0;  // a DR option (delayed.cpp:255-275)
    P.b = c->b;
    P.nLen = c->nLen;
    P.lenFunc = c->lenFunc;
    P.lenCdf = c->lenCdf;
    P.lenFuncInt = c->lenFuncInt;
    P.seed = o.seed + uint64_t(uint32_t(o.seed_offset)) * 0x9E3779B97F4A7C15ull;
    P.h2 = make_h2param(o.perturb_std_dev);
    P.o.roughThr = o.roughness_threshold;
    P.o.lightCoord = o.use_light_coordinate_sampling;
    P.o.chadMisQuirk = 0;
    P.oDeriv = P.o;
    P.oDeriv.chadMisQuirk = o.quirk_chad_mis;
    P.binStages = c->sched.bin_stages;
    P.stageLanes = c->sched.stage_lanes > 0 ? c->sched.stage_lanes : 64;
    P.chainBase = 0;
    P.chainCount = c->ch.C;
    return P;
}

extern "C" const char *drmlt_last_error(void) { return g_err.c_str(); }

extern "C" int drmlt_options_default(drmlt_options *o) {
    if (!o) return fail(DRMLT_ERR_ARG, "null options");
    std::memset(o, 0, sizeof(*o));
    o->spp = 16;                    // dptoptions.h:12
    o->num_init_samples = 100000;   // dptoptions.h:13
    o->min_depth = -1;
    o->max_depth = -1;
    o->direct_spp = 16;
    o->large_step_prob = 0.3f;
    o->perturb_std_dev = 0.01f;
    o->roughness_threshold = 0.03f;
    o->num_chains = 1024;
    o->seed_offset = 0;
    o->use_light_coordinate_sampling = 1;
    o->large_step_multiplexed = 1;
    o->discrete_std_dev = 0.01f;
    o->type = 0;
    o->aniso_perturb_prob = 0.5f;
    o->aniso_perturb_std_dev = 0.0025f;
    o->acceptance_map = 0;
    o->max_derivatives_depth = 9;   // main.cpp:76
    o->quirk_ignore_aniso_std_dev = 1;
    o->quirk_chad_mis = 1;
    o->integrator = 0;
    o->uniform_mixing_prob = 0.2f;  // dptoptions.h:29
    o->lens_perturb_prob = 0.3f;    // dptoptions.h:20
    o->lens_perturb_std_dev = 0.01f;  // dptoptions.h:21
    o->seed = 1;
    return DRMLT_OK;
}

static void free_list(std::vector<void *> &v) {
    for (void *p : v)
        if (p) (void)hipFree(p);
    v.clear();
}

extern "C" int drmlt_ctx_destroy(drmlt_ctx *c) {
    if (!c) return DRMLT_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    free_list(c->sceneAllocs);
    free_list(c->chainAllocs);
    for (void *p : {(void *)c->lenFunc, (void *)c->lenCdf, (void *)c->filmInd, (void *)c->filmDir, (void *)c->stats, (void *)c->iterDev})
        if (p) (void)hipFree(p);
    for (hipStream_t x : c->streamX)
        if (x) (void)hipStreamSynchronize(x);
    if (c->evInit)
        for (auto &e : c->ev) (void)hipEventDestroy(e);
    if (c->evFork) (void)hipEventDestroy(c->evFork);
    for (hipEvent_t e : c->evJoin)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t x : c->streamX)
        if (x) (void)hipStreamDestroy(x);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DRMLT_OK;
}

extern "C" int drmlt_bvh_info(drmlt_ctx *c, int32_t *onDevice, double *ms, int32_t *nodes, int32_t *depth4, int32_t *stack) {
    if (!c || !onDevice || !ms || !nodes || !depth4 || !stack) return fail(DRMLT_ERR_ARG, "bad argument");
    *onDevice = c->bvhOnDevice;
    *ms = c->bvhMs;
    *nodes = c->bvhNodes;
    *depth4 = c->bvhDepth;
    *stack = c->bvhStack;
    return DRMLT_OK;
}

extern "C" int drmlt_ctx_create(const drmlt_scene_desc *sc, const drmlt_options *opt, drmlt_ctx **out) {
    if (!sc || !opt || !out) return fail(DRMLT_ERR_ARG, "null argument");
    if (sc->num_triangles <= 0 || sc->num_lights <= 0 || sc->num_materials <= 0) return fail(DRMLT_ERR_ARG, "empty scene");
    if (opt->max_depth > 12 || opt->max_depth < 3) return fail(DRMLT_ERR_ARG, "max_depth must be in [3, 12]");
    if (opt->max_derivatives_depth > 12) return fail(DRMLT_ERR_ARG, "max_derivatives_depth must be <= 12");
    if (!opt->large_step_multiplexed) return fail(DRMLT_ERR_ARG, "only multiplexed large steps are implemented");
    if (opt->integrator < 0 || opt->integrator > 2) return fail(DRMLT_ERR_ARG, "integrator must be 0 (drmlt), 1 (mcmc) or 2 (h2mc)");
    if (opt->integrator == 2 && !(opt->lens_perturb_prob >= 0.0f && opt->lens_perturb_prob <= 1.0f))
        return fail(DRMLT_ERR_ARG, "lensperturbprob must be in [0, 1]");
    if (opt->integrator == 2 && opt->lens_perturb_prob > 0.0f && !(opt->lens_perturb_std_dev > 0.0f))
        return fail(DRMLT_ERR_ARG, "lensperturbstddev must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(DRMLT_ERR_HIP, "no HIP device");
    drmlt_ctx *c = new drmlt_ctx();
    c->opt = *opt;
    *out = nullptr;
    auto bail = [&](int code) { drmlt_ctx_destroy(c); return code; };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(DRMLT_ERR_HIP, "hipGetDevice"));
    if (int e = scratch_address(c->device, &c->scratch)) return bail(e);
    (void)hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->evFork, hipEventDisableTiming) != hipSuccess)
        return bail(fail(DRMLT_ERR_HIP, "stream"));
    for (int k = 0; k < kMaxParts - 1; k++)
        if (hipStreamCreateWithFlags(&c->streamX[k], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->evJoin[k], hipEventDisableTiming) != hipSuccess)
            return bail(fail(DRMLT_ERR_HIP, "stream"));
    const int nt = sc->num_triangles;
    c->nt = nt;
    c->nmat = sc->num_materials;
    std::vector<float4> s0(nt), s1(nt), s2(nt);
    for (int i = 0; i < nt; i++) {
        const float *p = sc->tri_p + 9 * size_t(i);
        s0[i] = make_float4(p[0], p[1], p[2], 0.0f);
        s1[i] = make_float4(p[3] - p[0], p[4] - p[1], p[5] - p[2], 0.0f);
        s2[i] = make_float4(p[6] - p[0], p[7] - p[1], p[8] - p[2], 0.0f);
    }
    hipStream_t s = c->stream;
    auto up = [&](auto **p, const auto *src, size_t n) {
        hipError_t e = upload(p, src, n, s);
        c->sceneAllocs.push_back((void *)*p);
        return e;
    };
    float4 *d_tri = nullptr, *d_nodes = nullptr, *d_s0, *d_s1, *d_s2;
    if (up(&d_s0, s0.data(), nt) != hipSuccess || up(&d_s1, s1.data(), nt) != hipSuccess || up(&d_s2, s2.data(), nt) != hipSuccess)
        return bail(fail(DRMLT_ERR_HIP, "scene upload"));
    // BVH over the triangles and the leaf-order triangle copy for traversal,
    // built on the device (bvh_device.hip) unless the options ask for the
    // host builder (or the scene is a single leaf); the binary depth cap is
    // lowered until no ray can need more than the traversal stack
    const auto tb0 = std::chrono::steady_clock::now();
    bool hostBuild = c->opt.host_bvh != 0 || nt <= 4;
    DeviceBvh db;
    if (!hostBuild) {
        float *d_triP = nullptr;
        if (upload(&d_triP, sc->tri_p, size_t(nt) * 9, s) != hipSuccess) return bail(fail(DRMLT_ERR_HIP, "scene upload"));
        hipError_t e = hipSuccess;
        for (int cap = 48;; cap -= 4) {
            e = BuildBvhDevice(nt, d_triP, d_s0, d_s1, d_s2, cap, s, db);
            if (e != hipSuccess || db.stackNeed <= kBvhStack || cap <= 8) break;
            (void)hipFree(db.nodes);
            (void)hipFree(db.tri);
        }
        (void)hipFree(d_triP);
        if (e == hipErrorInvalidValue) hostBuild = true;  // a single-leaf scene
        else if (e != hipSuccess) return bail(fail(DRMLT_ERR_HIP, std::string("device BVH build: ") + hipGetErrorString(e)));
    }
    if (!hostBuild) {
        c->sceneAllocs.push_back(db.nodes);
        c->sceneAllocs.push_back(db.tri);
        if (db.stackNeed > kBvhStack) return bail(fail(DRMLT_ERR_ARG, "BVH does not fit the traversal stack"));
        d_nodes = db.nodes;
        d_tri = db.tri;
        c->bvhDepth = db.depth4;
        c->bvhNodes = db.numNodes;
        c->bvhStack = db.stackNeed;
        c->bvhOnDevice = 1;
    } else {
        BvhOut bvh;
        for (int cap = 48;; cap -= 4) {
            BuildBvh(nt, sc->tri_p, cap, bvh);
            if (bvh.stackNeed <= kBvhStack) break;
            if (cap <= 8) return bail(fail(DRMLT_ERR_ARG, "BVH does not fit the traversal stack"));
        }
        std::vector<float4> tri(3 * size_t(nt));
        for (int k = 0; k < nt; k++) {
            const int i = bvh.order[k];
            int bits;
            std::memcpy(&bits, &i, 4);
            float w;
            std::memcpy(&w, &bits, 4);
            tri[3 * size_t(k)] = make_float4(s0[i].x, s0[i].y, s0[i].z, w);
            tri[3 * size_t(k) + 1] = s1[i];
            tri[3 * size_t(k) + 2] = s2[i];
        }
        if (up(&d_tri, tri.data(), 3 * size_t(nt)) != hipSuccess ||
            up(&d_nodes, (const float4 *)bvh.nodes.data(), bvh.nodes.size() / 4) != hipSuccess)
            return bail(fail(DRMLT_ERR_HIP, "scene upload"));
        if (hipStreamSynchronize(s) != hipSuccess) return bail(fail(DRMLT_ERR_HIP, "scene upload"));
        c->bvhDepth = bvh.depth4;
        c->bvhNodes = int(bvh.nodes.size() / 32);
        c->bvhStack = bvh.stackNeed;
        c->bvhOnDevice = 0;
    }
    c->bvhMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tb0).count();
    // lights: area cdf per light (TriangleMesh::SetAreaLight, trianglemesh.cpp:274-290)
    const int nl = sc->num_lights;
    std::vector<float> area(nl), cdfAll, funcAll;
    std::vector<int> cdfOff(nl);
    std::vector<float> weights(nl);
    float wsum = 0;
    for (int li = 0; li < nl; li++) {
        const int f = sc->light_tris[2 * li], n = sc->light_tris[2 * li + 1];
        if (f < 0 || n <= 0 || f + n > nt) return bail(fail(DRMLT_ERR_ARG, "light triangle range"));
        std::vector<float> a(n);
        float tot = 0;
        for (int k = 0; k < n; k++) {
            const float4 e1 = s1[f + k], e2 = s2[f + k];
            const float cx = e1.y * e2.z - e1.z * e2.y, cy = e1.z * e2.x - e1.x * e2.z, cz = e1.x * e2.y - e1.y * e2.x;
            a[k] = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
            tot += a[k];
        }
        area[li] = tot;
        std::vector<float> cdf;
        float fi;
        build_cdf(a.data(), n, cdf, fi);
        cdfOff[li] = int(cdfAll.size());
        cdfAll.insert(cdfAll.end(), cdf.begin(), cdf.end());
        funcAll.insert(funcAll.end(), a.begin(), a.end());
        weights[li] = sc->light_emit[4 * li + 3];
        wsum += weights[li];
    }
    std::vector<float> pickCdf;
    float pfi;
    build_cdf(weights.data(), nl, pickCdf, pfi);
    DevScene &S = c->S;
    float *d_sn, *d_mats, *d_le, *d_area, *d_cdf, *d_func, *d_pick;
    int *d_mat, *d_light, *d_lt, *d_cdfOff;
    bool ok = up(&d_sn, sc->tri_n, size_t(nt) * 9) == hipSuccess && up(&d_mat, sc->tri_mat, nt) == hipSuccess &&
              up(&d_light, sc->tri_light, nt) == hipSuccess && up(&d_mats, sc->materials, size_t(sc->num_materials) * 16) == hipSuccess &&
              up(&d_lt, sc->light_tris, size_t(nl) * 2) == hipSuccess && up(&d_le, sc->light_emit, size_t(nl) * 4) == hipSuccess &&
              up(&d_area, area.data(), nl) == hipSuccess && up(&d_cdf, cdfAll.data(), cdfAll.size()) == hipSuccess &&
              up(&d_func, funcAll.data(), funcAll.size()) == hipSuccess && up(&d_cdfOff, cdfOff.data(), nl) == hipSuccess &&
              up(&d_pick, pickCdf.data(), pickCdf.size()) == hipSuccess;
    if (!ok) return bail(fail(DRMLT_ERR_HIP, "scene upload"));
    for (int i = 0; i < nt; i++)
        if (sc->tri_mat[i] < 0 || sc->tri_mat[i] >= sc->num_materials || sc->tri_light[i] >= nl)
            return bail(fail(DRMLT_ERR_ARG, "triangle material / light index out of range"));
    S.tri = d_tri; S.nodes = d_nodes;
    S.sp0 = d_s0; S.se1 = d_s1; S.se2 = d_s2; S.sn = d_sn;
    S.tmat = d_mat; S.tlight = d_light; S.mats = d_mats;
    S.nt = nt; S.nlights = nl;
    S.lightTris = d_lt; S.lightEmit = d_le; S.lightArea = d_area;
    S.lightCdf = d_cdf; S.lightCdfOff = d_cdfOff; S.lightFunc = d_func; S.pickCdf = d_pick;
    S.lightWeightSum = wsum;
    S.cam = make_camera(sc->camera);
    c->W = S.cam.W;
    c->H = S.cam.H;
    const size_t filmN = size_t(c->W) * c->H * 3;
    if (dalloc(&c->filmInd, filmN) != hipSuccess || dalloc(&c->filmDir, filmN) != hipSuccess || dalloc(&c->stats, kStatCounters) != hipSuccess ||
        dalloc(&c->iterDev, 1) != hipSuccess)
        return bail(fail(DRMLT_ERR_HIP, "film allocation"));
    (void)hipMemsetAsync(c->filmInd, 0, filmN * 4, s);
    (void)hipMemsetAsync(c->filmDir, 0, filmN * 4, s);
    (void)hipMemsetAsync(c->stats, 0, kStatCounters * 8, s);
    (void)hipMemsetAsync(c->iterDev, 0, 4, s);
    for (auto &e : c->ev) (void)hipEventCreate(&e);
    c->evInit = true;
    if (hipStreamSynchronize(s) != hipSuccess) return bail(fail(DRMLT_ERR_HIP, "scene upload sync"));
    *out = c;
    return DRMLT_OK;
}

static int alloc_chains(drmlt_ctx *c, int C) {
    if (c->haveChains) {
        if (c->ch.C == C) return DRMLT_OK;
        free_list(c->chainAllocs);
        c->haveChains = false;
    }
    Chains &ch = c->ch;
    ch.C = C;
    ch.NS = 4 * C;
    ch.GS = 3 * C;
    const size_t NS = ch.NS, GS = ch.GS;
    c->chainBytes.clear();
    auto A = [&](auto **p, size_t n) {
        hipError_t e = dalloc(p, n);
        c->chainAllocs.push_back((void *)*p);
        c->chainBytes.push_back(std::max<size_t>(n, 1) * sizeof(**p));
        return e == hipSuccess;
    };
    bool ok = A(&ch.slots, size_t(kSlotWords) * NS) && A(&ch.perm, size_t(C)) && A(&ch.gauss, size_t(kGSlot) * GS) &&
              A(&ch.off, 32 * size_t(C)) && A(&ch.grad, size_t(kGMax) * C) && A(&ch.hess, size_t(kGMax * kGMax) * C) &&
              A(&ch.a1, size_t(C)) && A(&ch.a2, size_t(C)) && A(&ch.a2d, size_t(C)) && A(&ch.flags, size_t(C)) &&
              A(&c->L.count, kListKinds * kMaxBins) && A(&c->L.items, size_t(kListKinds) * kMaxBins * C);
    // lens Gaussians and lensVertexPos: the h2mc integrator with lens mutations only
    if (c->opt.integrator == 2 && c->opt.lens_perturb_prob > 0.0f) ok = ok && A(&ch.lensG, size_t(kLSlot) * GS);
    else ch.lensG = nullptr;
    for (Lists &lx : c->LX) ok = ok && A(&lx.count, kListKinds * kMaxBins) && A(&lx.items, size_t(kListKinds) * kMaxBins * C);
    if (!ok) return fail(DRMLT_ERR_HIP, "chain state allocation");
    bind_slot_fields(ch);
    hipStream_t s = c->stream;
    // every state field starts zeroed: c = l = 0 is an empty path
    for (size_t k = 0; k < c->chainAllocs.size(); k++)
        if (c->chainAllocs[k]) HIPCHK(hipMemsetAsync(c->chainAllocs[k], 0, c->chainBytes[k], s));
    (void)hipMemsetAsync(c->L.count, 0, kListKinds * kMaxBins * 4, s);
    for (Lists &lx : c->LX) (void)hipMemsetAsync(lx.count, 0, kListKinds * kMaxBins * 4, s);
    c->haveChains = true;
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    return DRMLT_OK;
}

static bool recs_ok(const drmlt_ctx *c, const drmlt_path_rec *r, int64_t n);
static bool recs_fixed_ok(const drmlt_path_rec *r, int64_t n);

static void drop_graph(drmlt_ctx *c) {
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
}

static int set_length_dist(drmlt_ctx *c, const float *lf, int n) {
    drop_graph(c);
    std::vector<float> cdf;
    float fi;
    build_cdf(lf, n, cdf, fi);
    if (c->lenFunc) (void)hipFree(c->lenFunc);
    if (c->lenCdf) (void)hipFree(c->lenCdf);
    c->lenFunc = c->lenCdf = nullptr;
    HIPCHK(upload(&c->lenFunc, lf, n, c->stream));
    HIPCHK(upload(&c->lenCdf, cdf.data(), cdf.size(), c->stream));
    c->lenFuncInt = fi;
    c->nLen = n;
    return DRMLT_OK;
}

// Host images of one slot record (the layout of Chains::slots).
template <class T> static void wput(uint32_t *w, int i, T v) {
    static_assert(sizeof(T) == 4, "32-bit field");
    std::memcpy(w + i, &v, 4);
}
template <class T> static T wget(const uint32_t *w, int i) {
    static_assert(sizeof(T) == 4, "32-bit field");
    T v;
    std::memcpy(&v, w + i, 4);
    return v;
}
// An empty state: c = l = 0, no light, every vertex triangle -1 and
// Russian-roulette factor 1, everything else 0.
static void slot_default(uint32_t *w) {
    std::memset(w, 0, kSlotWords * 4);
    wput<int32_t>(w, SO_LIGHT, -1);
    wput<int32_t>(w, SO_LPRIM, -1);
    for (int k = 0; k < kVMax; k++) {
        wput<int32_t>(w, SO_CV + 3 * k, -1); wput<float>(w, SO_CV + 3 * k + 2, 1.0f);
        wput<int32_t>(w, SO_LV + 3 * k, -1); wput<float>(w, SO_LV + 3 * k + 2, 1.0f);
    }
}
// A record and its contrib8 row (valid, rgb, lsScore, ssScore, screen) into a slot.
static void slot_pack(uint32_t *w, const drmlt_path_rec &R, const float *c8) {
    slot_default(w);
    wput<int32_t>(w, SO_C, R.c); wput<int32_t>(w, SO_L, R.l);
    wput<int32_t>(w, SO_LIGHT, R.light); wput<int32_t>(w, SO_LPRIM, R.lprim);
    wput<int32_t>(w, SO_VALID, c8[0] != 0.0f ? 1 : 0);
    for (int k = 0; k < 3; k++) wput<float>(w, SO_RGB + k, c8[1 + k]);
    wput<float>(w, SO_LS, c8[4]); wput<float>(w, SO_SS, c8[5]);
    wput<float>(w, SO_SCREEN, c8[6]); wput<float>(w, SO_SCREEN + 1, c8[7]);
    for (int k = 0; k < kDMax; k++) wput<float>(w, SO_U + k, R.u[k]);
    for (int k = 0; k < kVMax; k++) {
        wput<int32_t>(w, SO_CV + 3 * k, R.cprim[k]); wput<float>(w, SO_CV + 3 * k + 1, R.cdisc[k]); wput<float>(w, SO_CV + 3 * k + 2, R.crr[k]);
        wput<int32_t>(w, SO_LV + 3 * k, R.lgprim[k]); wput<float>(w, SO_LV + 3 * k + 1, R.ldisc[k]); wput<float>(w, SO_LV + 3 * k + 2, R.lrr[k]);
    }
}
// A slot as a record and its contrib8 row.
static void slot_unpack(const uint32_t *w, drmlt_path_rec &R, float *o) {
    R.c = wget<int32_t>(w, SO_C); R.l = wget<int32_t>(w, SO_L);
    R.light = wget<int32_t>(w, SO_LIGHT); R.lprim = wget<int32_t>(w, SO_LPRIM);
    R.screen[0] = wget<float>(w, SO_SCREEN); R.screen[1] = wget<float>(w, SO_SCREEN + 1);
    for (int k = 0; k < kDMax; k++) R.u[k] = wget<float>(w, SO_U + k);
    for (int k = 0; k < kVMax; k++) {
        R.cprim[k] = wget<int32_t>(w, SO_CV + 3 * k); R.cdisc[k] = wget<float>(w, SO_CV + 3 * k + 1); R.crr[k] = wget<float>(w, SO_CV + 3 * k + 2);
        R.lgprim[k] = wget<int32_t>(w, SO_LV + 3 * k); R.ldisc[k] = wget<float>(w, SO_LV + 3 * k + 1); R.lrr[k] = wget<float>(w, SO_LV + 3 * k + 2);
    }
    o[0] = wget<int32_t>(w, SO_VALID) ? 1.0f : 0.0f;
    for (int k = 0; k < 3; k++) o[1 + k] = wget<float>(w, SO_RGB + k);
    o[4] = wget<float>(w, SO_LS); o[5] = wget<float>(w, SO_SS);
    o[6] = R.screen[0]; o[7] = R.screen[1];
}

// Host-side mirror of the seed records into chain slot 0; slots 1-3 get the
// seed's vertex counts and are otherwise empty.
extern "C" int drmlt_seed_chains(drmlt_ctx *c, const drmlt_path_rec *st, const float *c8, int64_t n, float b, const float *lf, int32_t nLen) {
    if (!c || !st || !c8 || !lf || n <= 0 || nLen <= 0 || nLen > kMaxBins) return fail(DRMLT_ERR_ARG, "bad argument");
    // records flagged invalid (contrib8[0] == 0) seed an empty invalid state,
    // which the first iteration replaces by a large step
    for (int64_t i = 0; i < n; i++) {
        if (c8[8 * i] == 0.0f) continue;
        if (!recs_ok(c, st + i, 1) || !recs_fixed_ok(st + i, 1)) return fail(DRMLT_ERR_ARG, "seed record out of range");
        if (st[i].c + st[i].l - 1 > c->opt.max_depth) return fail(DRMLT_ERR_ARG, "seed path longer than max_depth");
    }
    int r = alloc_chains(c, int(n));
    if (r) return r;
    Chains &ch = c->ch;
    const size_t C = size_t(n), W = kSlotWords;
    std::vector<uint32_t> recs(C * W);
    std::vector<int> perm(C, 0 | (1 << 2) | (2 << 4));
    hipStream_t s = c->stream;
    // slot 0: the seeds (a record flagged invalid, contrib8[0] == 0, stays empty)
    for (size_t i = 0; i < C; i++) {
        if (c8[8 * i] == 0.0f) slot_default(&recs[i * W]);
        else slot_pack(&recs[i * W], st[i], c8 + 8 * i);
    }
    HIPCHK(hipMemcpyAsync(ch.slots, recs.data(), C * W * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    // slots 1-3: empty states with the seed's c and l
    for (size_t i = 0; i < C; i++) {
        uint32_t *w = &recs[i * W];
        const bool seeded = c8[8 * i] != 0.0f;
        slot_default(w);
        wput<int32_t>(w, SO_C, seeded ? st[i].c : 0);
        wput<int32_t>(w, SO_L, seeded ? st[i].l : 0);
    }
    for (size_t k = 1; k < 4; k++) HIPCHK(hipMemcpyAsync(ch.slots + k * C * W, recs.data(), C * W * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ch.perm, perm.data(), C * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    c->b = b;
    r = set_length_dist(c, lf, nLen);
    if (r) return r;
    HIPCHK(hipMemsetAsync(c->filmInd, 0, size_t(c->W) * c->H * 3 * 4, s));
    HIPCHK(hipMemsetAsync(c->stats, 0, kStatCounters * 8, s));
    c->itersDone = 0;
    HIPCHK(hipStreamSynchronize(s));
    return DRMLT_OK;
}

extern "C" int drmlt_prepass(drmlt_ctx *c, double *bOut, float *lfOut, int32_t nLenOut) {
    if (!c) return fail(DRMLT_ERR_ARG, "null context");
    const int N = c->opt.num_init_samples, C = c->opt.num_chains, maxDepth = c->opt.max_depth;
    if (N <= 0 || C <= 0) return fail(DRMLT_ERR_ARG, "num_init_samples and num_chains must be positive");
    hipStream_t s = c->stream;
    DBuf<int> bCnt, bCl;
    DBuf<float> bLs;
    DBuf<double> bLenSum, bTotal;
    HIPCHK(dalloc(&bCnt.p, N)); HIPCHK(dalloc(&bCl.p, size_t(N) * kPreMaxC)); HIPCHK(dalloc(&bLs.p, size_t(N) * kPreMaxC));
    HIPCHK(dalloc(&bLenSum.p, kMaxBins)); HIPCHK(dalloc(&bTotal.p, 1));
    int *cnt = bCnt.p, *cl = bCl.p;
    float *ls = bLs.p;
    double *lenSum = bLenSum.p, *total = bTotal.p;
    HIPCHK(hipMemsetAsync(lenSum, 0, kMaxBins * 8, s));
    HIPCHK(hipMemsetAsync(total, 0, 8, s));
    DrParams P = make_params(c);
    hipLaunchKernelGGL(k_prepass_gen, dim3((N + 63) / 64), dim3(64), 0, s, c->S, P.o, N, maxDepth, P.seed ^ 0x5151515151ull, cnt, cl, ls,
                       lenSum, total);
    HIPCHK(hipGetLastError());
    std::vector<int> hcnt(N), hcl(size_t(N) * kPreMaxC);
    std::vector<float> hls(size_t(N) * kPreMaxC);
    std::vector<double> hlen(kMaxBins);
    double htot = 0;
    HIPCHK(hipMemcpyAsync(hcnt.data(), cnt, N * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hcl.data(), cl, hcl.size() * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hls.data(), ls, hls.size() * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hlen.data(), lenSum, kMaxBins * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&htot, total, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // candidates in sample order, equal-spaced seeding (Veach's thesis p. 340, delayed.cpp:123-167)
    std::vector<int> candSample, candCl;
    std::vector<double> cdf(1, 0.0);
    for (int i = 0; i < N; i++)
        for (int k = 0; k < hcnt[i]; k++) {
            candSample.push_back(i);
            candCl.push_back(hcl[size_t(i) * kPreMaxC + k]);
            cdf.push_back(cdf.back() + double(hls[size_t(i) * kPreMaxC + k]));
        }
    const int ncand = int(candSample.size());
    if (ncand < 2) return fail(DRMLT_ERR_STATE, "pre-pass found fewer than two contributing paths");
    const int nLen = maxDepth + 1;
    std::vector<float> lf(nLen, 0.0f);
    for (int k = 0; k < nLen && k < kMaxBins; k++) lf[k] = float(hlen[k]);
    const double interval = cdf.back() / double(C);
    std::mt19937_64 rng{uint64_t(ncand)};
    double pos = std::uniform_real_distribution<double>(0.0, interval)(rng);
    int cdfPos = 0;
    std::vector<int> selS(C), selCl(C);
    for (int i = 0; i < C; i++) {
        // the cursor is capped at ncand - 1 as in delayed.cpp:138-140, where a
        // position inside the last candidate's interval never ends the walk;
        // here the walk stops at the cap and takes the candidate it names
        while (cdfPos < ncand - 1 && pos > cdf[cdfPos]) cdfPos++;
        const int k = std::max(cdfPos - 1, 0);
        selS[i] = candSample[k];
        selCl[i] = candCl[k];
        pos += interval;
    }
    int r = alloc_chains(c, C);
    if (r) return r;
    DBuf<int> dS, dCl;
    HIPCHK(upload(&dS.p, selS.data(), C, s));
    HIPCHK(upload(&dCl.p, selCl.data(), C, s));
    hipLaunchKernelGGL(k_prepass_seed, dim3((C + 63) / 64), dim3(64), 0, s, c->S, P.o, c->ch, maxDepth, P.seed ^ 0x5151515151ull, dS.p, dCl.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    c->b = float(htot / double(N));
    r = set_length_dist(c, lf.data(), nLen);
    if (r) return r;
    HIPCHK(hipMemsetAsync(c->filmInd, 0, size_t(c->W) * c->H * 3 * 4, s));
    HIPCHK(hipMemsetAsync(c->stats, 0, kStatCounters * 8, s));
    HIPCHK(hipStreamSynchronize(s));
    c->itersDone = 0;
    if (bOut) *bOut = htot / double(N);
    if (lfOut)
        for (int k = 0; k < nLenOut; k++) lfOut[k] = k < nLen ? lf[k] : 0.0f;
    return DRMLT_OK;
}

extern "C" int drmlt_direct_lighting(drmlt_ctx *c) {
    if (!c) return fail(DRMLT_ERR_ARG, "null context");
    const int spp = c->opt.direct_spp;
    hipStream_t s = c->stream;
    HIPCHK(hipMemsetAsync(c->filmDir, 0, size_t(c->W) * c->H * 3 * 4, s));
    if (spp > 0 && c->opt.min_depth <= 2) {
        const long n = long(c->W) * c->H * spp;
        DrParams P = make_params(c);
        hipLaunchKernelGGL(k_direct, dim3(unsigned((n + 63) / 64)), dim3(64), 0, s, c->S, P.o, spp, P.seed ^ 0xD1EC7ull, c->filmDir);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));
    c->directDone = true;
    return DRMLT_OK;
}

static int persistent_blocks(const drmlt_ctx *c) { return c->cus * 16; }
// Workgroup counts of the batched step's kernels under the context's schedule.
struct StepGrid {
    int stage, deriv, gauss;
};
// (for `C` chains)
static StepGrid step_grid(const drmlt_ctx *c, int C) {
    const int pb = persistent_blocks(c);
    StepGrid g;
    const int lanes = c->sched.stage_lanes > 0 ? c->sched.stage_lanes : 64;
    const int perChain = (C + lanes - 1) / lanes;
    g.stage = c->sched.stage_blocks > 0 ? std::min(c->sched.stage_blocks, perChain) : std::min(perChain, pb);
    g.deriv = c->sched.deriv_blocks > 0 ? c->sched.deriv_blocks : pb;
    g.gauss = c->sched.gauss_blocks > 0 ? c->sched.gauss_blocks : pb;
    return g;
}

// One batched DR step: the stage-1 binning kernel, seven stage kernels and
// the list reset (the binning kernel is timed with stage 1); without a
// second stage (the mcmc integrator) only stage 1 and the reset run, the
// events of the other stages still recorded. With
// `counters` (7 x 4 words) the counting forms of the stage kernels run
// instead, each accumulating its traversal work and state bytes into its own
// row.
// The step's kernels for the chains of P on stream s with lists L (the
// split step's halves).
// The path-stage kernels built for the schedule's resident waves per SIMD.
struct PathStageKernels {
    decltype(&k_b_stage1) stage1;
    decltype(&k_b_stage2_propose) propose;
    decltype(&k_b_stage2_accept) accept;
};
static PathStageKernels path_stage_kernels(const drmlt_ctx *c) {
    if (c->sched.stage_waves == 3) return {k_b_stage1_w3, k_b_stage2_propose_w3, k_b_stage2_accept_w3};
    return {k_b_stage1, k_b_stage2_propose, k_b_stage2_accept};
}

// h2mc with lens mutations: the lens scores before stage 1 (its lens coin
// reads them) and the lens mutations of list 5 after it.
static void launch_lens_scores(drmlt_ctx *c, const DrParams &P, hipStream_t s) {
    if (P.lensProb > 0.0f)
        hipLaunchKernelGGL(k_b_lens_score, dim3(std::max(1, std::min((P.chainCount + 63) / 64, persistent_blocks(c)))), dim3(64), 0, s, c->S, c->ch,
                           P);
}
static void launch_lens_moves(drmlt_ctx *c, const DrParams &P, hipStream_t s, const Lists &L, const StepGrid &G) {
    if (P.lensProb > 0.0f)
        hipLaunchKernelGGL(k_b_lens, dim3(G.stage), dim3(64), 0, s, c->S, c->ch, P, c->iterDev, c->filmInd, c->stats, L);
}

static void launch_step_kernels(drmlt_ctx *c, const DrParams &P, hipStream_t s, const Lists &L, const StepGrid &G) {
    const PathStageKernels K = path_stage_kernels(c);
    launch_lens_scores(c, P, s);
    if (P.binStages == 1) hipLaunchKernelGGL(k_b_stage1_bin, dim3(G.stage), dim3(64), 0, s, c->ch, P, c->iterDev, L);
    hipLaunchKernelGGL(K.stage1, dim3(G.stage), dim3(64), 0, s, c->S, c->ch, P, c->iterDev, c->filmInd, c->stats, L);
    launch_lens_moves(c, P, s, L, G);
    if (P.integrator == 1) return;
    hipLaunchKernelGGL(k_b_deriv, dim3(G.deriv), dim3(64), 0, s, c->S, c->ch, P, L, 4, 0);
    hipLaunchKernelGGL(k_b_gauss, dim3(G.gauss), dim3(64), 0, s, c->S, c->ch, P, L, 0, 0, 0, c->iterDev);
    hipLaunchKernelGGL(K.propose, dim3(G.stage), dim3(64), 0, s, c->S, c->ch, P, c->iterDev, L);
    hipLaunchKernelGGL(k_b_deriv, dim3(G.deriv), dim3(64), 0, s, c->S, c->ch, P, L, 2, 1);
    hipLaunchKernelGGL(k_b_gauss, dim3(G.gauss), dim3(64), 0, s, c->S, c->ch, P, L, 2, 1, 1, c->iterDev);
    hipLaunchKernelGGL(K.accept, dim3(G.stage), dim3(64), 0, s, c->S, c->ch, P, c->iterDev, c->filmInd, c->stats, L);
}

static void launch_batched_step(drmlt_ctx *c, const DrParams &P, bool timed, unsigned long long *counters = nullptr) {
    hipStream_t s = c->stream;
    const StepGrid G = step_grid(c, c->ch.C);
    if (counters) {
        DevSceneCounted Sk[7];
        for (int k = 0; k < 7; k++) {
            static_cast<DevSceneData &>(Sk[k]) = c->S;
            Sk[k].counters = counters + 4 * k;
        }
        launch_lens_scores(c, P, s);  // the lens kernels have no counting form: their work is not counted
        if (P.binStages == 1) hipLaunchKernelGGL(k_b_stage1_bin_counted, dim3(G.stage), dim3(64), 0, s, c->ch, P, c->iterDev, c->L);
        hipLaunchKernelGGL(k_b_stage1_counted, dim3(G.stage), dim3(64), 0, s, Sk[0], c->ch, P, c->iterDev, c->filmInd, c->stats, c->L);
        launch_lens_moves(c, P, s, c->L, G);
        if (P.integrator != 1) {
            hipLaunchKernelGGL(k_b_deriv_counted, dim3(G.deriv), dim3(64), 0, s, Sk[1], c->ch, P, c->L, 4, 0);
            hipLaunchKernelGGL(k_b_gauss_counted, dim3(G.gauss), dim3(64), 0, s, Sk[2], c->ch, P, c->L, 0, 0, 0, c->iterDev);
            hipLaunchKernelGGL(k_b_stage2_propose_counted, dim3(G.stage), dim3(64), 0, s, Sk[3], c->ch, P, c->iterDev, c->L);
            hipLaunchKernelGGL(k_b_deriv_counted, dim3(G.deriv), dim3(64), 0, s, Sk[4], c->ch, P, c->L, 2, 1);
            hipLaunchKernelGGL(k_b_gauss_counted, dim3(G.gauss), dim3(64), 0, s, Sk[5], c->ch, P, c->L, 2, 1, 1, c->iterDev);
            hipLaunchKernelGGL(k_b_stage2_accept_counted, dim3(G.stage), dim3(64), 0, s, Sk[6], c->ch, P, c->iterDev, c->filmInd, c->stats, c->L);
        }
        hipLaunchKernelGGL(k_b_reset, dim3(1), dim3(64), 0, s, c->L, c->iterDev, 1);
        return;
    }
    const int parts = std::min(c->sched.stage_split, c->ch.C);
    if (!timed && parts > 1) {
        // the chains in `parts` contiguous ranges, each with its own lists,
        // on streams that fork and join around the step; the iteration
        // counter advances after the join
        const int C = c->ch.C;
        (void)hipEventRecord(c->evFork, s);
        for (int k = 0; k < parts; k++) {
            DrParams Pk = P;
            Pk.chainBase = int(int64_t(C) * k / parts);
            Pk.chainCount = int(int64_t(C) * (k + 1) / parts) - Pk.chainBase;
            hipStream_t sk = k == 0 ? s : c->streamX[k - 1];
            const Lists &Lk = k == 0 ? c->L : c->LX[k - 1];
            if (k > 0) (void)hipStreamWaitEvent(sk, c->evFork, 0);
            launch_step_kernels(c, Pk, sk, Lk, step_grid(c, Pk.chainCount));
            if (k > 0) {
                hipLaunchKernelGGL(k_b_reset, dim3(1), dim3(64), 0, sk, Lk, c->iterDev, 0);
                (void)hipEventRecord(c->evJoin[k - 1], sk);
            }
        }
        for (int k = 1; k < parts; k++) (void)hipStreamWaitEvent(s, c->evJoin[k - 1], 0);
        hipLaunchKernelGGL(k_b_reset, dim3(1), dim3(64), 0, s, c->L, c->iterDev, 1);
        return;
    }
    const PathStageKernels K = path_stage_kernels(c);
    if (timed) (void)hipEventRecord(c->ev[0], s);
    launch_lens_scores(c, P, s);  // timed with stage 1
    if (P.binStages == 1) hipLaunchKernelGGL(k_b_stage1_bin, dim3(G.stage), dim3(64), 0, s, c->ch, P, c->iterDev, c->L);
    hipLaunchKernelGGL(K.stage1, dim3(G.stage), dim3(64), 0, s, c->S, c->ch, P, c->iterDev, c->filmInd, c->stats, c->L);
    launch_lens_moves(c, P, s, c->L, G);
    if (timed) (void)hipEventRecord(c->ev[1], s);
    if (P.integrator != 1) hipLaunchKernelGGL(k_b_deriv, dim3(G.deriv), dim3(64), 0, s, c->S, c->ch, P, c->L, 4, 0);
    if (timed) (void)hipEventRecord(c->ev[2], s);
    if (P.integrator != 1) hipLaunchKernelGGL(k_b_gauss, dim3(G.gauss), dim3(64), 0, s, c->S, c->ch, P, c->L, 0, 0, 0, c->iterDev);
    if (timed) (void)hipEventRecord(c->ev[3], s);
    if (P.integrator != 1) hipLaunchKernelGGL(K.propose, dim3(G.stage), dim3(64), 0, s, c->S, c->ch, P, c->iterDev, c->L);
    if (timed) (void)hipEventRecord(c->ev[4], s);
    if (P.integrator != 1) hipLaunchKernelGGL(k_b_deriv, dim3(G.deriv), dim3(64), 0, s, c->S, c->ch, P, c->L, 2, 1);
    if (timed) (void)hipEventRecord(c->ev[5], s);
    if (P.integrator != 1) hipLaunchKernelGGL(k_b_gauss, dim3(G.gauss), dim3(64), 0, s, c->S, c->ch, P, c->L, 2, 1, 1, c->iterDev);
    if (timed) (void)hipEventRecord(c->ev[6], s);
    if (P.integrator != 1) hipLaunchKernelGGL(K.accept, dim3(G.stage), dim3(64), 0, s, c->S, c->ch, P, c->iterDev, c->filmInd, c->stats, c->L);
    if (timed) (void)hipEventRecord(c->ev[7], s);
    hipLaunchKernelGGL(k_b_reset, dim3(1), dim3(64), 0, s, c->L, c->iterDev, 1);
    if (timed) (void)hipEventRecord(c->ev[8], s);
}

extern "C" int drmlt_run(drmlt_ctx *c, int64_t iters, int32_t mode) {
    if (!c || iters < 0) return fail(DRMLT_ERR_ARG, "bad argument");
    if (!c->haveChains || !c->lenCdf) return fail(DRMLT_ERR_STATE, "chains are not seeded (call drmlt_prepass or drmlt_seed_chains)");
    if (mode < 0 || mode > 2) return fail(DRMLT_ERR_ARG, "unknown mode");
    hipStream_t s = c->stream;
    const DrParams P = make_params(c);
    const int C = c->ch.C;
    for (double &t : c->timing) t = 0;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // the device iteration counter follows the host one (plain form reads the host one)
    HIPCHK(hipMemcpyAsync(c->iterDev, &c->iterHost, 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(c->L.count, 0, kListKinds * kMaxBins * 4, s));
    for (Lists &lx : c->LX) HIPCHK(hipMemsetAsync(lx.count, 0, kListKinds * kMaxBins * 4, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipEventRecord(e0, s));
    if (mode == DRMLT_PLAIN) {
        for (int64_t it = 0; it < iters; it++) {
            hipLaunchKernelGGL(k_dr_plain, dim3((C + 63) / 64), dim3(64), 0, s, c->S, c->ch, P, c->iterHost + uint32_t(it), c->filmInd,
                               c->stats);
        }
        HIPCHK(hipGetLastError());
    } else if (mode == DRMLT_BATCHED) {
        for (int64_t it = 0; it < iters; it++) {
            launch_batched_step(c, P, true);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventSynchronize(c->ev[8]));
            for (int k = 0; k < 7; k++) {
                float ms = 0;
                (void)hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]);
                c->timing[1 + k] += ms;
            }
        }
    } else {
        if (!c->gexec) {
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            launch_batched_step(c, P, false);
            HIPCHK(hipStreamEndCapture(s, &c->graph));
            HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        }
        for (int64_t it = 0; it < iters; it++) HIPCHK(hipGraphLaunch(c->gexec, s));
    }
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    c->timing[0] = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipGetLastError());
    c->iterHost += uint32_t(iters);
    c->itersDone += iters;
    return DRMLT_OK;
}

extern "C" int drmlt_read_timing(drmlt_ctx *c, double *out, int32_t n) {
    if (!c || !out) return fail(DRMLT_ERR_ARG, "bad argument");
    for (int k = 0; k < n && k < 8; k++) out[k] = c->timing[k];
    return DRMLT_OK;
}

extern "C" int drmlt_get_schedule(drmlt_ctx *c, drmlt_schedule *out) {
    if (!c || !out) return fail(DRMLT_ERR_ARG, "bad argument");
    *out = c->sched;
    return DRMLT_OK;
}

extern "C" int drmlt_set_schedule(drmlt_ctx *c, const drmlt_schedule *sc) {
    if (!c || !sc) return fail(DRMLT_ERR_ARG, "bad argument");
    if (sc->bin_stages < 0 || sc->bin_stages > 2 || sc->stage_blocks < 0 || sc->deriv_blocks < 0 || sc->gauss_blocks < 0 ||
        sc->stage_lanes < 0 || sc->stage_lanes > 64 || sc->stage_split < 0 || sc->stage_split > kMaxParts ||
        !(sc->stage_waves == 0 || sc->stage_waves == 2 || sc->stage_waves == 3))
        return fail(DRMLT_ERR_ARG, "schedule out of range");
    c->sched = *sc;
    drop_graph(c);  // the captured step bakes in the launch geometry
    return DRMLT_OK;
}

// Coordinate search over the schedule: first the path-stage scheduling
// (chains per workgroup, binned or chain order, the stage kernels'
// workgroup count) by the whole step's time, then the derivative and Gaussian kernels' workgroup
// counts by their own kernel times. Every candidate runs one untimed and
// `steps` timed batched steps of the context's own chains; those are ordinary
// DR iterations (they advance the chains and splat into the film).
extern "C" int drmlt_autotune(drmlt_ctx *c, int32_t steps, drmlt_schedule *best, double *bestMs) {
    if (!c || steps <= 0) return fail(DRMLT_ERR_ARG, "bad argument");
    if (!c->haveChains || !c->lenCdf) return fail(DRMLT_ERR_STATE, "chains are not seeded (call drmlt_prepass or drmlt_seed_chains)");
    const int C = c->ch.C, cus = c->cus;
    // per-step milliseconds of stage kernels k0..k1 (timing slots) under schedule sc
    auto measure = [&](const drmlt_schedule &sc, int k0, int k1, double &ms) -> int {
        c->sched = sc;
        int r = drmlt_run(c, 1, DRMLT_BATCHED);
        if (r) return r;
        r = drmlt_run(c, steps, DRMLT_BATCHED);
        if (r) return r;
        ms = 0;
        for (int k = k0; k <= k1; k++) ms += c->timing[k] / steps;
        return DRMLT_OK;
    };
    drmlt_schedule cur = c->sched, cand;
    double curMs = 0, ms = 0;
    int r = measure(cur, 1, 7, curMs);
    if (r) return r;
    // chains per workgroup: full waves, half waves, and the count that fills
    // exactly two waves per SIMD (the path-stage kernels' occupancy), which
    // measured best when the chains fill no more than that
    const int fill = std::min(64, std::max(1, (C + 2 * 4 * cus - 1) / (2 * 4 * cus)));
    std::vector<int> laneOpts = {64, 32};
    if (fill != 64 && fill != 32) laneOpts.push_back(fill);
    // and the path-stage kernels' register bound (three resident waves per
    // SIMD against two), searched together: the best workgroup count depends
    // on it (with 1M chains the three-wave kernels do best at one workgroup
    // per 64 chains, the two-wave ones at a quarter of that)
    for (int lanes : laneOpts)
        for (int bin = 0; bin <= 2; bin++) {
            const int perChain = (C + lanes - 1) / lanes;
            for (int sb : {cus * 4, cus * 8, cus * 16, perChain})
                for (int waves : {2, 3}) {
                    cand = cur;
                    cand.stage_lanes = lanes;
                    cand.bin_stages = bin;
                    cand.stage_blocks = std::min(sb, perChain);
                    cand.stage_waves = waves;
                    if ((r = measure(cand, 1, 7, ms))) return r;
                    if (ms < curMs) { cur = cand; curMs = ms; }
                }
        }
    const int kernelSlots[2][2] = {{2, 5}, {3, 6}};  // derivative and Gaussian kernels (current + proposal)
    for (int which = 0; which < 2; which++) {
        double kMs = 0;
        if ((r = measure(cur, kernelSlots[which][0], kernelSlots[which][0], kMs))) return r;
        kMs += c->timing[kernelSlots[which][1]] / steps;
        for (int mult : {4, 8, 16, 32}) {
            cand = cur;
            (which == 0 ? cand.deriv_blocks : cand.gauss_blocks) = cus * mult;
            if ((r = measure(cand, kernelSlots[which][0], kernelSlots[which][0], ms))) return r;
            ms += c->timing[kernelSlots[which][1]] / steps;
            if (ms < kMs) { cur = cand; kMs = ms; }
        }
    }
    // the split step (graph-replayed form only), by the captured step's whole
    // time over at least 10 steps (the ranges differ by a few percent)
    const int gsteps = std::max(steps, 10);
    auto graphMs = [&](const drmlt_schedule &sc, double &ms) -> int {
        c->sched = sc;
        drop_graph(c);
        int rr = drmlt_run(c, 1, DRMLT_BATCHED_GRAPH);
        if (rr) return rr;
        rr = drmlt_run(c, gsteps, DRMLT_BATCHED_GRAPH);
        if (rr) return rr;
        ms = c->timing[0] / gsteps;
        return DRMLT_OK;
    };
    // (the best of two interleaved rounds per choice: one stream, 2..kMaxParts ranges)
    double splitMs[kMaxParts + 1];
    for (double &m : splitMs) m = 1e30;
    for (int round = 0; round < 2; round++)
        for (int parts = 1; parts <= kMaxParts; parts++) {
            cand = cur;
            cand.stage_split = parts == 1 ? 0 : parts;
            if ((r = graphMs(cand, ms))) return r;
            splitMs[parts] = std::min(splitMs[parts], ms);
        }
    int bestParts = 1;
    for (int parts = 2; parts <= kMaxParts; parts++)
        if (splitMs[parts] < splitMs[bestParts]) bestParts = parts;
    cur.stage_split = bestParts == 1 ? 0 : bestParts;
    // the derivative and Gaussian workgroup counts were timed on one stream:
    // against their defaults under the chosen split, by the captured step's time
    if (cur.deriv_blocks != 0 || cur.gauss_blocks != 0) {
        cand = cur;
        cand.deriv_blocks = cand.gauss_blocks = 0;
        double tunedMs = 1e30, defMs = 1e30;
        for (int round = 0; round < 2; round++) {
            if ((r = graphMs(cur, ms))) return r;
            tunedMs = std::min(tunedMs, ms);
            if ((r = graphMs(cand, ms))) return r;
            defMs = std::min(defMs, ms);
        }
        if (defMs < tunedMs) cur = cand;
    }
    if ((r = measure(cur, 1, 7, curMs))) return r;
    c->sched = cur;
    drop_graph(c);
    if (best) *best = cur;
    if (bestMs) *bestMs = curMs;
    return DRMLT_OK;
}

extern "C" int drmlt_read_film(drmlt_ctx *c, float *rgb, int32_t w, int32_t h, int32_t which) {
    if (!c || !rgb || w != c->W || h != c->H) return fail(DRMLT_ERR_ARG, "film size mismatch");
    const size_t n = size_t(w) * h * 3;
    std::vector<float> ind(n), dir(n);
    HIPCHK(hipMemcpyAsync(ind.data(), c->filmInd, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(dir.data(), c->filmDir, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (which == 1) { std::memcpy(rgb, ind.data(), n * 4); return DRMLT_OK; }
    if (which == 2) { std::memcpy(rgb, dir.data(), n * 4); return DRMLT_OK; }
    const double sppDone = c->haveChains ? double(c->itersDone) * double(c->ch.C) / (double(w) * h) : 0.0;
    const float dw = c->opt.direct_spp > 0 ? 1.0f / float(c->opt.direct_spp) : 0.0f;
    const float iw = sppDone > 0 ? float(1.0 / sppDone) : 0.0f;
    for (size_t i = 0; i < n; i++) rgb[i] = dw * dir[i] + iw * ind[i];
    return DRMLT_OK;
}

extern "C" int drmlt_read_stats(drmlt_ctx *c, drmlt_stats *st) {
    if (!c || !st) return fail(DRMLT_ERR_ARG, "bad argument");
    unsigned long long v[kStatCounters];
    HIPCHK(hipMemcpyAsync(v, c->stats, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    st->large_step = v[0]; st->large_step_accepted = v[1];
    st->stage1 = v[2]; st->stage1_accepted = v[3];
    st->stage2_iso = v[4]; st->stage2_iso_accepted = v[5];
    st->stage2_h2mc = v[6]; st->stage2_h2mc_accepted = v[7];
    st->lens_step = v[8]; st->lens_step_accepted = v[9];
    return DRMLT_OK;
}

// ---------------------------------------------------------------- output files

extern "C" int drmlt_write_exr(const char *path, const float *rgb, int32_t w, int32_t h, int32_t compression) {
    std::string err;
    if (write_exr(path, rgb, w, h, compression, err) != 0) return fail(DRMLT_ERR_ARG, err.c_str());
    return DRMLT_OK;
}

extern "C" int drmlt_stats_text(const drmlt_stats *st, int32_t format, char *buf, int32_t cap) {
    if (!st || cap < 0 || (cap > 0 && !buf) || format < 0 || format > 2) return fail(DRMLT_ERR_ARG, "bad argument");
    const uint64_t v[10] = {st->large_step, st->large_step_accepted, st->stage1, st->stage1_accepted,
                            st->stage2_iso, st->stage2_iso_accepted, st->stage2_h2mc, st->stage2_h2mc_accepted,
                            st->lens_step, st->lens_step_accepted};
    const std::string t = stats_text(v, format);
    if (cap > 0) {
        const size_t n = std::min(t.size(), size_t(cap - 1));
        std::memcpy(buf, t.data(), n);
        buf[n] = '\0';
    }
    return int(t.size());
}

static int write_text(const std::string &path, const std::string &text, bool append) {
    FILE *fp = std::fopen(path.c_str(), append ? "ab" : "wb");
    if (!fp) return fail(DRMLT_ERR_ARG, ("cannot open " + path).c_str());
    const size_t wr = std::fwrite(text.data(), 1, text.size(), fp);
    if (std::fclose(fp) != 0 || wr != text.size()) return fail(DRMLT_ERR_ARG, ("write failed: " + path).c_str());
    return DRMLT_OK;
}

extern "C" int drmlt_write_stats(drmlt_ctx *c, const char *path, int32_t reference_labels) {
    if (!c || !path) return fail(DRMLT_ERR_ARG, "bad argument");
    drmlt_stats st;
    const int rc = drmlt_read_stats(c, &st);
    if (rc != DRMLT_OK) return rc;
    const uint64_t v[10] = {st.large_step, st.large_step_accepted, st.stage1, st.stage1_accepted,
                            st.stage2_iso, st.stage2_iso_accepted, st.stage2_h2mc, st.stage2_h2mc_accepted,
                            st.lens_step, st.lens_step_accepted};
    return write_text(path, stats_text(v, c->opt.integrator != 0 ? 2 : (reference_labels != 0 ? 1 : 0)), false);
}

// The chain loop with the reference's intermediate outputs (delayed.cpp:
// 234-242, 610-655): the partial directory is created if missing; whenever
// the time since the call began reaches interval_seconds x k (k = 1, 2, ...),
// the indirect buffer scaled by pixels / mutations so far is written to
// <dir>/<base>_<k>.exr and the elapsed seconds to <dir>/<base>_time.csv
// (truncated at k = 1, appended after); <dir>/<base>_stats.txt is rewritten
// with every partial image and at the end. The loop runs in chunks of
// chunk_iters iterations per chain, the reports fall between chunks.
extern "C" int drmlt_render(drmlt_ctx *c, int64_t iters_per_chain, int32_t mode, const drmlt_report *rep) {
    if (!c || iters_per_chain < 0 || !rep) return fail(DRMLT_ERR_ARG, "bad argument");
    if (mode < 0 || mode > 2) return fail(DRMLT_ERR_ARG, "unknown mode");
    if (rep->compression != kExrNone && rep->compression != kExrZip) return fail(DRMLT_ERR_ARG, "unsupported EXR compression");
    const bool files = rep->partial_dir && rep->partial_dir[0];
    if (files && (!rep->base_name || !rep->base_name[0])) return fail(DRMLT_ERR_ARG, "partial outputs need a base name");
    if (!c->haveChains || !c->lenCdf) return fail(DRMLT_ERR_STATE, "chains are not seeded (call drmlt_prepass or drmlt_seed_chains)");
    std::string dir, base;
    if (files) {
        dir = rep->partial_dir;
        base = dir + "/" + rep->base_name;
        struct stat sb;
        if (stat(dir.c_str(), &sb) != 0 && mkdir(dir.c_str(), 0755) != 0) return fail(DRMLT_ERR_ARG, ("cannot create " + dir).c_str());
    }
    const int64_t chunk = rep->chunk_iters > 0 ? rep->chunk_iters : 16;
    const auto t0 = std::chrono::system_clock::now();
    const size_t np = size_t(c->W) * c->H;
    int imgId = 1;
    for (int64_t done = 0; done < iters_per_chain;) {
        const int64_t k = std::min(chunk, iters_per_chain - done);
        int rc = drmlt_run(c, k, mode);
        if (rc != DRMLT_OK) return rc;
        done += k;
        if (!files || !(rep->interval_seconds > 0)) continue;
        const long ms = long(std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::system_clock::now() - t0).count());
        const float elapsed = float(ms) / 1000.f;
        if (!(elapsed >= float(rep->interval_seconds) * float(imgId))) continue;
        std::vector<float> ind(np * 3);
        if ((rc = drmlt_read_film(c, ind.data(), c->W, c->H, 1)) != DRMLT_OK) return rc;
        const float scale = float(np) / float(double(c->itersDone) * double(c->ch.C));
        for (float &v : ind) v *= scale;
        std::string err;
        if (write_exr((base + "_" + std::to_string(imgId) + ".exr").c_str(), ind.data(), c->W, c->H, rep->compression, err) != 0)
            return fail(DRMLT_ERR_ARG, err.c_str());
        if ((rc = write_text(base + "_time.csv", float_text(elapsed) + ",\n", imgId > 1)) != DRMLT_OK) return rc;
        if ((rc = drmlt_write_stats(c, (base + "_stats.txt").c_str(), rep->reference_labels)) != DRMLT_OK) return rc;
        imgId++;
    }
    if (files) return drmlt_write_stats(c, (base + "_stats.txt").c_str(), rep->reference_labels);
    return DRMLT_OK;
}

extern "C" int drmlt_count_work(drmlt_ctx *c, uint64_t *out) {
    if (!c || !out) return fail(DRMLT_ERR_ARG, "bad argument");
    if (!c->haveChains || !c->lenCdf) return fail(DRMLT_ERR_STATE, "chains are not seeded");
    hipStream_t s = c->stream;
    DBuf<unsigned long long> buf;
    HIPCHK(dalloc(&buf.p, 28));
    unsigned long long *d = buf.p;
    HIPCHK(hipMemsetAsync(d, 0, 28 * 8, s));
    HIPCHK(hipMemcpyAsync(c->iterDev, &c->iterHost, 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(c->L.count, 0, kListKinds * kMaxBins * 4, s));
    const DrParams P = make_params(c);
    launch_batched_step(c, P, false, d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, d, 28 * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    c->iterHost += 1;
    c->itersDone += 1;
    return DRMLT_OK;
}

// ---------------------------------------------------------------- tier 2
#define DUP(buf, src, n) HIPCHK(upload(&buf.p, src, size_t(n), s))
#define DNEW(buf, n) HIPCHK(dalloc(&buf.p, size_t(n)))
#define DGET(dst, buf, n) HIPCHK(hipMemcpyAsync(dst, buf.p, size_t(n) * sizeof(*buf.p), hipMemcpyDeviceToHost, s))

extern "C" int drmlt_trace(drmlt_ctx *c, int64_t n, const float *org, const float *dir, const float *tmin, const float *tmax, int32_t any,
                           int32_t *prim, float *t) {
    if (!c || n < 0 || !org || !dir || !tmin || !tmax || !prim || (!any && !t)) return fail(DRMLT_ERR_ARG, "bad argument");
    if (n == 0) return DRMLT_OK;
    hipStream_t s = c->stream;
    DBuf<float> o, d, a, b, tt;
    DBuf<int> pr;
    DUP(o, org, 3 * n); DUP(d, dir, 3 * n); DUP(a, tmin, n); DUP(b, tmax, n); DNEW(pr, n); DNEW(tt, n);
    hipLaunchKernelGGL(k_trace, dim3(unsigned((n + 63) / 64)), dim3(64), 0, s, c->S, int(n), o.p, d.p, a.p, b.p, any, pr.p, tt.p);
    HIPCHK(hipGetLastError());
    DGET(prim, pr, n);
    if (!any) DGET(t, tt, n);
    HIPCHK(hipStreamSynchronize(s));
    return DRMLT_OK;
}

// Validates path records before any kernel reads them: vertex counts within
// the fixed slots, light and triangle indices within the scene.
static bool recs_ok(const drmlt_ctx *c, const drmlt_path_rec *r, int64_t n) {
    for (int64_t i = 0; i < n; i++) {
        const drmlt_path_rec &R = r[i];
        const int cc = R.c, ll = R.l;
        if (cc < 1 || ll < 0 || cc - 1 > 16 || ll - 1 > 16 || cc + ll - 1 < 2 || 2 * std::max(cc + ll - 1, 2) > 32) return false;
        if (cc == 1 && ll < 2) return false;
        if (ll >= 1 && (R.light < 0 || R.light >= c->S.nlights || R.lprim < 0 || R.lprim >= c->nt)) return false;
        for (int k = 0; k < cc - 1; k++)
            if (R.cprim[k] < -1 || R.cprim[k] >= c->nt) return false;
        for (int k = 0; k < ll - 1; k++)
            if (R.lgprim[k] < -1 || R.lgprim[k] >= c->nt) return false;
    }
    return true;
}
// For fixed-topology evaluation every vertex must name a triangle.
static bool recs_fixed_ok(const drmlt_path_rec *r, int64_t n) {
    for (int64_t i = 0; i < n; i++) {
        for (int k = 0; k < r[i].c - 1; k++)
            if (r[i].cprim[k] < 0) return false;
        for (int k = 0; k < r[i].l - 1; k++)
            if (r[i].lgprim[k] < 0) return false;
    }
    return true;
}

extern "C" int drmlt_perturb(drmlt_ctx *c, int64_t n, drmlt_path_rec *st, const float *off, const float *disc, float *c8) {
    if (!c || n < 0 || !st || !off || !disc || !c8) return fail(DRMLT_ERR_ARG, "bad argument");
    if (!recs_ok(c, st, n)) return fail(DRMLT_ERR_ARG, "path record out of range");
    // The batched PerturbPathBidir kernel made a GPU memory fault whose cause
    // is not established; it stays switched off until it is. Perturbation
    // itself runs inside the loop's small steps.
    return fail(DRMLT_ERR_STATE, "drmlt_perturb is switched off in this build (unexplained GPU memory fault of its kernel)");
}

extern "C" int drmlt_generate_subpath(drmlt_ctx *c, int64_t n, const int32_t *cl, const float *uni, int32_t nuni, drmlt_path_rec *st,
                                      float *c8) {
    if (!c || n < 0 || !cl || !uni || nuni <= 0 || !st || !c8) return fail(DRMLT_ERR_ARG, "bad argument");
    // The batched GenerateSubpath kernel made a GPU memory fault whose cause
    // is not established; it stays switched off until it is. Subpath
    // generation itself runs inside the loop's large steps.
    return fail(DRMLT_ERR_STATE, "drmlt_generate_subpath is switched off in this build (unexplained GPU memory fault of its kernel)");
}

extern "C" int drmlt_generate_bidir(drmlt_ctx *c, int64_t n, int32_t minDepth, int32_t maxDepth, const float *uni, int32_t nuni, int32_t maxc,
                                    float *cs, int32_t *nc) {
    if (!c || n < 0 || !uni || nuni <= 0 || maxc <= 0 || !cs || !nc) return fail(DRMLT_ERR_ARG, "bad argument");
    if (maxDepth > 15 || maxDepth < 1) return fail(DRMLT_ERR_ARG, "max_depth out of range");
    if (n == 0) return DRMLT_OK;
    hipStream_t s = c->stream;
    DBuf<float> u, o;
    DBuf<int> k;
    DUP(u, uni, size_t(n) * nuni); DNEW(o, size_t(n) * maxc * 10); DNEW(k, n);
    const DrParams P = make_params(c);
    hipLaunchKernelGGL(k_genbidir_rec, dim3(unsigned((n + 63) / 64)), dim3(64), 0, s, c->S, P.o, int(n), minDepth, maxDepth, u.p, nuni, maxc,
                       o.p, k.p);
    HIPCHK(hipGetLastError());
    DGET(cs, o, size_t(n) * maxc * 10);
    DGET(nc, k, n);
    HIPCHK(hipStreamSynchronize(s));
    return DRMLT_OK;
}

static int jet_passes_host(int D) {
    const int nch = (D + kJetK - 1) / kJetK;
    int np = 0;
    for (int i = 0; i < D; i++) np += nch - i / kJetK;
    return np;
}

// Device time of the function-level kernel bracketed by ev[0] and ev[1], kept
// as timing[0] (drmlt_read_timing).
static void kernel_time(drmlt_ctx *c) {
    float ms = 0;
    if (hipEventSynchronize(c->ev[1]) == hipSuccess && hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) {
        for (double &t : c->timing) t = 0;
        c->timing[0] = ms;
    }
}

extern "C" int drmlt_eval_logf(drmlt_ctx *c, int64_t n, const drmlt_path_rec *st, int32_t fp64, double *logf, int32_t *okOut) {
    if (!c || n < 0 || !st || !logf || !okOut) return fail(DRMLT_ERR_ARG, "bad argument");
    if (!recs_ok(c, st, n) || !recs_fixed_ok(st, n)) return fail(DRMLT_ERR_ARG, "path record out of range");
    if (n == 0) return DRMLT_OK;
    hipStream_t s = c->stream;
    DBuf<char> r;
    DBuf<double> lf;
    DBuf<int> ok;
    DUP(r, (const char *)st, n * sizeof(drmlt_path_rec));
    DNEW(lf, n); DNEW(ok, n);
    const DrParams P = make_params(c);
    HIPCHK(hipEventRecord(c->ev[0], s));
    if (fp64) hipLaunchKernelGGL(k_logf_rec_f64, dim3(unsigned((n + 63) / 64)), dim3(64), 0, s, c->S, P.oDeriv, int(n), r.p, lf.p, ok.p);
    else hipLaunchKernelGGL(k_logf_rec_f32, dim3(unsigned((n + 63) / 64)), dim3(64), 0, s, c->S, P.oDeriv, int(n), r.p, lf.p, ok.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[1], s));
    DGET(logf, lf, n);
    DGET(okOut, ok, n);
    HIPCHK(hipStreamSynchronize(s));
    kernel_time(c);
    return DRMLT_OK;
}

extern "C" int drmlt_lens_grad_hess(drmlt_ctx *c, int64_t n, const drmlt_path_rec *st, const float *screen, const float *lensPos, float *out9,
                                    int32_t *okOut) {
    if (!c || n < 0 || !st || !screen || !out9 || !okOut) return fail(DRMLT_ERR_ARG, "bad argument");
    if (!recs_ok(c, st, n) || !recs_fixed_ok(st, n)) return fail(DRMLT_ERR_ARG, "path record out of range");
    if (n == 0) return DRMLT_OK;
    hipStream_t s = c->stream;
    DBuf<char> r;
    DBuf<float> sc, lp, o9;
    DBuf<int> ok;
    DUP(r, (const char *)st, n * sizeof(drmlt_path_rec));
    DUP(sc, screen, 2 * n);
    if (lensPos) DUP(lp, lensPos, 3 * n);
    DNEW(o9, 9 * n); DNEW(ok, n);
    const DrParams P = make_params(c);
    hipLaunchKernelGGL(k_lens_rec, dim3(unsigned((n + 63) / 64)), dim3(64), 0, s, c->S, P.o, int(n), r.p, sc.p, lensPos ? lp.p : nullptr, o9.p,
                       ok.p);
    HIPCHK(hipGetLastError());
    DGET(out9, o9, 9 * n);
    DGET(okOut, ok, n);
    HIPCHK(hipStreamSynchronize(s));
    return DRMLT_OK;
}

// The state of chains 0..n-1 in host memory: their four slots (slot k of
// chain i at slots[(k * n + i) * kSlotWords]) and the per-chain scratch of
// the last iteration.
struct HostChains {
    size_t n = 0;
    std::vector<int> perm, flags;
    std::vector<uint32_t> slots;
    std::vector<float> a1, off;
    std::vector<double> a2;
    const uint32_t *slot(size_t chain, int k) const { return &slots[(size_t(k) * n + chain) * kSlotWords]; }
};
static int fetch_chains(drmlt_ctx *c, HostChains &h, size_t n, bool scratch) {
    const Chains &ch = c->ch;
    const size_t C = size_t(ch.C), W = kSlotWords;
    hipStream_t s = c->stream;
    auto get = [&](const auto *dev, size_t count, auto &host) {
        host.resize(count);
        return hipMemcpyAsync(host.data(), dev, count * sizeof(host[0]), hipMemcpyDeviceToHost, s);
    };
    h.n = n;
    h.slots.resize(4 * n * W);
    HIPCHK(get(ch.perm, n, h.perm));
    for (size_t k = 0; k < 4; k++)
        HIPCHK(hipMemcpyAsync(&h.slots[k * n * W], ch.slots + k * C * W, n * W * 4, hipMemcpyDeviceToHost, s));
    if (scratch) {
        HIPCHK(get(ch.flags, n, h.flags)); HIPCHK(get(ch.a1, n, h.a1)); HIPCHK(get(ch.a2d, n, h.a2)); HIPCHK(get(ch.off, 32 * n, h.off));
    }
    HIPCHK(hipStreamSynchronize(s));
    return DRMLT_OK;
}

extern "C" int drmlt_read_chains(drmlt_ctx *c, int64_t n, drmlt_path_rec *st, float *c8) {
    if (!c || n < 0 || !st || !c8) return fail(DRMLT_ERR_ARG, "bad argument");
    if (!c->haveChains) return fail(DRMLT_ERR_STATE, "chains are not seeded (call drmlt_prepass or drmlt_seed_chains)");
    if (n > c->ch.C) return fail(DRMLT_ERR_ARG, "more states asked for than there are chains");
    if (n == 0) return DRMLT_OK;
    HostChains h;
    if (int r = fetch_chains(c, h, size_t(n), false)) return r;
    for (int64_t i = 0; i < n; i++) slot_unpack(h.slot(size_t(i), h.perm[i] & 3), st[i], c8 + 8 * size_t(i));
    return DRMLT_OK;
}

extern "C" int drmlt_read_step(drmlt_ctx *c, int64_t n, drmlt_path_rec *slots, float *c8, int32_t *perm, int32_t *flags, float *a1,
                               double *alpha2, float *off) {
    if (!c || n < 0 || !slots || !c8 || !perm || !flags || !a1 || !alpha2 || !off) return fail(DRMLT_ERR_ARG, "bad argument");
    if (!c->haveChains) return fail(DRMLT_ERR_STATE, "chains are not seeded (call drmlt_prepass or drmlt_seed_chains)");
    if (n > c->ch.C) return fail(DRMLT_ERR_ARG, "more chains asked for than there are");
    if (n == 0) return DRMLT_OK;
    HostChains h;
    if (int r = fetch_chains(c, h, size_t(n), true)) return r;
    for (int64_t i = 0; i < n; i++) {
        for (int k = 0; k < 4; k++) slot_unpack(h.slot(size_t(i), k), slots[4 * i + k], c8 + 8 * (4 * size_t(i) + k));
        perm[i] = h.perm[i];
        flags[i] = h.flags[i];
        a1[i] = h.a1[i];
        alpha2[i] = (h.flags[i] & 2) ? h.a2[i] : 0.0;
        for (int k = 0; k < 32; k++) off[32 * i + k] = h.off[32 * size_t(i) + k];
    }
    return DRMLT_OK;
}

extern "C" int drmlt_read_lens(drmlt_ctx *c, int64_t n, float *lscore, int32_t *lflags, float *lensPos) {
    if (!c || n < 0 || !lscore || !lflags || !lensPos) return fail(DRMLT_ERR_ARG, "bad argument");
    if (!c->haveChains) return fail(DRMLT_ERR_STATE, "chains are not seeded (call drmlt_prepass or drmlt_seed_chains)");
    if (n > c->ch.C) return fail(DRMLT_ERR_ARG, "more chains asked for than there are");
    if (n == 0) return DRMLT_OK;
    HostChains h;
    if (int r = fetch_chains(c, h, size_t(n), false)) return r;
    const size_t C = size_t(c->ch.C);
    std::vector<double> lg;
    if (c->ch.lensG) {
        // lens slots of physical slots 0..2: slot k of chain i at k * C + i
        lg.resize(3 * C * kLSlot);
        HIPCHK(hipMemcpy(lg.data(), c->ch.lensG, lg.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) {
            const uint32_t *w = h.slot(size_t(i), k);
            lscore[4 * i + k] = wget<float>(w, SO_LENS);
            lflags[4 * i + k] = wget<int32_t>(w, SO_LFLAGS);
            for (int d = 0; d < 3; d++)
                lensPos[12 * i + 3 * k + d] = (k < 3 && !lg.empty()) ? float(lg[(size_t(k) * C + size_t(i)) * kLSlot + 9 + d]) : 0.0f;
        }
    return DRMLT_OK;
}

extern "C" int drmlt_grad_hess(drmlt_ctx *c, int64_t n, const drmlt_path_rec *st, int32_t fp64, double *grad, double *hess, int32_t *okOut) {
    if (!c || n < 0 || !st || !grad || !hess || !okOut) return fail(DRMLT_ERR_ARG, "bad argument");
    if (!recs_ok(c, st, n) || !recs_fixed_ok(st, n)) return fail(DRMLT_ERR_ARG, "path record out of range");
    if (n == 0) return DRMLT_OK;
    std::vector<int> passOff(n + 1, 0);
    for (int64_t i = 0; i < n; i++) {
        const int D = 2 * std::max(st[i].c + st[i].l - 1, 2);
        if (D > kGMax) return fail(DRMLT_ERR_ARG, "derivatives are provided up to D = 24");
        passOff[i + 1] = passOff[i] + jet_passes_host(D);
    }
    const int total = passOff[n];
    hipStream_t s = c->stream;
    DBuf<char> r;
    DBuf<int> po, ok;
    DUP(r, (const char *)st, n * sizeof(drmlt_path_rec)); DUP(po, passOff.data(), n + 1);
    std::vector<int> ones(n, 1);
    DUP(ok, ones.data(), n);
    const DrParams P = make_params(c);
    if (fp64) {
        DBuf<double> g, h;
        DNEW(g, 32 * n); DNEW(h, 1024 * n);
        HIPCHK(hipMemsetAsync(g.p, 0, 32 * n * 8, s));
        HIPCHK(hipMemsetAsync(h.p, 0, 1024 * n * 8, s));
        HIPCHK(hipEventRecord(c->ev[0], s));
        hipLaunchKernelGGL(k_gradhess_rec_f64, dim3(unsigned((total + 63) / 64)), dim3(64), 0, s, c->S, P.oDeriv, int(n), r.p, po.p,
                           total, g.p, h.p, ok.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[1], s));
        DGET(grad, g, 32 * n);
        DGET(hess, h, 1024 * n);
    } else {
        DBuf<float> g, h;
        DNEW(g, 32 * n); DNEW(h, 1024 * n);
        HIPCHK(hipMemsetAsync(g.p, 0, 32 * n * 4, s));
        HIPCHK(hipMemsetAsync(h.p, 0, 1024 * n * 4, s));
        HIPCHK(hipEventRecord(c->ev[0], s));
        hipLaunchKernelGGL(k_gradhess_rec_f32, dim3(unsigned((total + 63) / 64)), dim3(64), 0, s, c->S, P.oDeriv, int(n), r.p, po.p,
                           total, g.p, h.p, ok.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[1], s));
        std::vector<float> hg(32 * n), hh(1024 * n);
        DGET(hg.data(), g, 32 * n);
        DGET(hh.data(), h, 1024 * n);
        HIPCHK(hipStreamSynchronize(s));
        for (size_t i = 0; i < hg.size(); i++) grad[i] = hg[i];
        for (size_t i = 0; i < hh.size(); i++) hess[i] = hh[i];
    }
    DGET(okOut, ok, n);
    HIPCHK(hipStreamSynchronize(s));
    kernel_time(c);
    // the passes write the upper triangle (column >= row); mirror it
    for (int64_t i = 0; i < n; i++) {
        double *H = hess + size_t(i) * 1024;
        for (int a = 0; a < 32; a++)
            for (int b = 0; b < a; b++) H[a * 32 + b] = H[b * 32 + a];
    }
    return DRMLT_OK;
}

// drmlt_grad_hess without allocating: the records, the pass table and the
// results go through this device's `scratch` array (at most kNoAllocStates
// states per call), cleared by fill kernels. Overlapping calls on one GPU
// share the array and race.
// ------------------------------------------------ the reference's serialized form
// Material types and area lights per triangle from host scene arrays.
static void scene_facts_host(const drmlt_scene_desc *sd, std::vector<int> &matType, std::vector<int> &triLight) {
    const int nt = sd->num_triangles;
    matType.resize(nt);
    triLight.resize(nt);
    for (int i = 0; i < nt; i++) {
        const int m = sd->tri_mat[i];
        matType[i] = (m >= 0 && m < sd->num_materials) ? int(sd->materials[16 * size_t(m)]) : -1;
        triLight[i] = sd->tri_light[i];
    }
}
// (p0, p1 - p0, p2 - p0) per triangle in float, as the context's scene holds them.
static std::vector<float> tri_keys_host(const drmlt_scene_desc *sd) {
    const int nt = sd->num_triangles;
    std::vector<float> k(9 * size_t(nt));
    for (int i = 0; i < nt; i++) {
        const float *p = sd->tri_p + 9 * size_t(i);
        float *o = &k[9 * size_t(i)];
        for (int a = 0; a < 3; a++) {
            o[a] = p[a];
            o[3 + a] = p[3 + a] - p[a];
            o[6 + a] = p[6 + a] - p[a];
        }
    }
    return k;
}
static int serialized_args_ok(int32_t cam, int32_t lgt, int64_t n, const float *lens, const float *primary, int64_t pstride,
                              const float *vert, int64_t vstride) {
    if (n < 0 || cam < 1 || lgt < 0 || cam + lgt < 3 || cam + lgt - 1 > 12 || (cam == 1 && lgt < 2))
        return fail(DRMLT_ERR_ARG, "vertex counts out of range (c >= 1, l >= 0, 3 <= c + l, path length <= 12)");
    if (n > 0 && (!lens || !primary || !vert)) return fail(DRMLT_ERR_ARG, "bad argument");
    if (pstride < ser_primary_size(cam, lgt)) return fail(DRMLT_ERR_ARG, "primary stride shorter than GetPrimaryParamSize(c, l)");
    if (vstride <= 3) return fail(DRMLT_ERR_ARG, "vertParams stride too short");
    return DRMLT_OK;
}
static int decode_all(const TriKeyMap &tm, const SceneFacts &sf, int32_t cam, int32_t lgt, int64_t n, const float *lens,
                      const float *primary, int64_t pstride, const float *vert, int64_t vstride, drmlt_path_rec *out) {
    std::string err;
    for (int64_t i = 0; i < n; i++)
        if (decode_subpath(tm, sf, cam, lgt, lens + 2 * i, primary + pstride * i, vert + vstride * i, vstride, out[i], err))
            return fail(DRMLT_ERR_ARG, "serialized state " + std::to_string(i) + ": " + err);
    return DRMLT_OK;
}

extern "C" int drmlt_material_from_serialized(const float *bsdf, int32_t two_sided, float *material) {
    if (!bsdf || !material) return fail(DRMLT_ERR_ARG, "bad argument");
    if (!material_from_serialized(bsdf, two_sided != 0, material)) return fail(DRMLT_ERR_ARG, "unknown serialized BSDF type");
    return DRMLT_OK;
}

extern "C" int drmlt_serialized_to_rec(const drmlt_scene_desc *sd, int32_t cam_depth, int32_t lgt_depth, int64_t n, const float *lens,
                                       const float *primary, int64_t primary_stride, const float *vert_params, int64_t vert_stride,
                                       drmlt_path_rec *states) {
    if (!sd || (n > 0 && !states) || sd->num_triangles <= 0 || !sd->tri_p || !sd->tri_mat || !sd->tri_light || !sd->materials)
        return fail(DRMLT_ERR_ARG, "bad argument");
    if (int r = serialized_args_ok(cam_depth, lgt_depth, n, lens, primary, primary_stride, vert_params, vert_stride)) return r;
    if (n == 0) return DRMLT_OK;
    TriKeyMap tm;
    tm.build(sd->num_triangles, tri_keys_host(sd).data());
    std::vector<int> matType, triLight;
    scene_facts_host(sd, matType, triLight);
    return decode_all(tm, SceneFacts{matType.data(), triLight.data()}, cam_depth, lgt_depth, n, lens, primary, primary_stride, vert_params,
                      vert_stride, states);
}

// The context's triangle keys, material types and area lights, read back
// from the device scene once.
static int ensure_tri_map(drmlt_ctx *c) {
    if (!c->triMap.empty()) return DRMLT_OK;
    const size_t nt = size_t(c->nt);
    std::vector<float4> s0(nt), s1(nt), s2(nt);
    std::vector<int> tmat(nt);
    std::vector<float> mats(16 * size_t(c->nmat));
    c->triLight.resize(nt);
    HIPCHK(hipMemcpy(s0.data(), c->S.sp0, nt * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s1.data(), c->S.se1, nt * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s2.data(), c->S.se2, nt * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(tmat.data(), c->S.tmat, nt * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(c->triLight.data(), c->S.tlight, nt * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(mats.data(), c->S.mats, mats.size() * 4, hipMemcpyDeviceToHost));
    std::vector<float> keys(9 * nt);
    for (size_t i = 0; i < nt; i++) {
        const float4 *v[3] = {&s0[i], &s1[i], &s2[i]};
        for (int a = 0; a < 3; a++) {
            keys[9 * i + 3 * a] = v[a]->x;
            keys[9 * i + 3 * a + 1] = v[a]->y;
            keys[9 * i + 3 * a + 2] = v[a]->z;
        }
    }
    c->triMatType.resize(nt);
    for (size_t i = 0; i < nt; i++) c->triMatType[i] = int(mats[16 * size_t(tmat[i])]);
    c->triMap.build(int(nt), keys.data());
    return DRMLT_OK;
}

// The scene layout (scene.cpp:126-131) against the context's scene: the
// light-coordinate option, sampleToCam (column-major), the camera position,
// the pixel count and dist.
static int scene_params_ok(const drmlt_ctx *c, const float *sp) {
    const DevCamera &K = c->S.cam;
    auto near = [](float a, float b) { return std::fabs(a - b) <= 1e-4f * (1.0f + std::fabs(b)); };
    if ((sp[0] != 0.0f) != (c->opt.use_light_coordinate_sampling != 0))
        return fail(DRMLT_ERR_ARG, "sceneParams: useLightCoordinateSampling differs from the context's option");
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            if (!near(sp[1 + 4 * i + j], K.s2c[4 * j + i])) return fail(DRMLT_ERR_ARG, "sceneParams: sampleToCam differs from the context's camera");
    // camToWorld: isMoving, translate[0] (the camera position), ...
    for (int a = 0; a < 3; a++)
        if (!near(sp[18 + a], K.c2w[4 * a + 3])) return fail(DRMLT_ERR_ARG, "sceneParams: camera position differs from the context's camera");
    if (sp[32] != float(K.W) * float(K.H)) return fail(DRMLT_ERR_ARG, "sceneParams: pixel count differs from the context's film");
    if (!near(sp[33], K.dist)) return fail(DRMLT_ERR_ARG, "sceneParams: camera dist differs from the context's camera");
    return DRMLT_OK;
}

extern "C" int drmlt_path_func_derv(drmlt_ctx *c, int32_t cam_depth, int32_t lgt_depth, int64_t n, const float *lens, const float *primary,
                                    int64_t primary_stride, const float *scene_params, const float *vert_params, int64_t vert_stride,
                                    float *grad, float *hess, int32_t *okOut) {
    if (!c || !scene_params || (n > 0 && (!grad || !hess || !okOut))) return fail(DRMLT_ERR_ARG, "bad argument");
    if (int r = serialized_args_ok(cam_depth, lgt_depth, n, lens, primary, primary_stride, vert_params, vert_stride)) return r;
    if (int r = scene_params_ok(c, scene_params)) return r;
    if (n == 0) return DRMLT_OK;
    if (int r = ensure_tri_map(c)) return r;
    std::vector<drmlt_path_rec> recs(n);
    if (int r = decode_all(c->triMap, SceneFacts{c->triMatType.data(), c->triLight.data()}, cam_depth, lgt_depth, n, lens, primary,
                           primary_stride, vert_params, vert_stride, recs.data()))
        return r;
    std::vector<double> g(size_t(n) * 32), h(size_t(n) * 32 * 32);
    std::vector<int32_t> ok(n);
    if (int r = drmlt_grad_hess(c, n, recs.data(), 0, g.data(), h.data(), ok.data())) return r;
    const int D = 2 * std::max(cam_depth + lgt_depth - 1, 2);
    for (int64_t i = 0; i < n; i++) {
        float *gi = grad + size_t(i) * D, *hi = hess + size_t(i) * D * D;
        bool finite = ok[i] != 0;
        for (int r = 0; r < D && finite; r++) {
            gi[r] = float(g[size_t(i) * 32 + r]);
            finite = std::isfinite(gi[r]);
            for (int k = 0; k < D && finite; k++) {
                hi[r * D + k] = float(h[(size_t(i) * 32 + r) * 32 + k]);
                finite = std::isfinite(hi[r * D + k]);
            }
        }
        if (!finite) {
            // as the reference's caller does with a non-finite result (drsmallstep.cpp:149-155)
            std::fill(gi, gi + D, 0.0f);
            std::fill(hi, hi + size_t(D) * D, 0.0f);
        }
        okOut[i] = finite ? 1 : 0;
    }
    return DRMLT_OK;
}

extern "C" int drmlt_grad_hess_noalloc(drmlt_ctx *c, int64_t n, const drmlt_path_rec *st, int32_t fp64, double *grad, double *hess,
                                       int32_t *okOut) {
    if (!c || n < 0 || !st || !grad || !hess || !okOut) return fail(DRMLT_ERR_ARG, "bad argument");
    if (n > kNoAllocStates) return fail(DRMLT_ERR_ARG, "drmlt_grad_hess_noalloc takes at most 4096 states per call");
    if (!c->scratch) return fail(DRMLT_ERR_STATE, "no scratch address for this context's device");
    if (!recs_ok(c, st, n) || !recs_fixed_ok(st, n)) return fail(DRMLT_ERR_ARG, "path record out of range");
    if (n == 0) return DRMLT_OK;
    int passOff[kNoAllocStates + 1];
    passOff[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        const int D = 2 * std::max(st[i].c + st[i].l - 1, 2);
        if (D > kGMax) return fail(DRMLT_ERR_ARG, "derivatives are provided up to D = 24");
        passOff[i + 1] = passOff[i] + jet_passes_host(D);
    }
    const int total = passOff[n];
    hipStream_t s = c->stream;
    char *base = c->scratch;
    char *dRec = base + kNaRec;
    int *dPass = reinterpret_cast<int *>(base + kNaPass), *dOk = reinterpret_cast<int *>(base + kNaOk);
    float *g32 = reinterpret_cast<float *>(base + kNaG32), *h32 = reinterpret_cast<float *>(base + kNaH32);
    double *g64 = reinterpret_cast<double *>(base + kNaG64), *h64 = reinterpret_cast<double *>(base + kNaH64);
    HIPCHK(hipMemcpyAsync(dRec, st, size_t(n) * sizeof(drmlt_path_rec), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dPass, passOff, size_t(n + 1) * 4, hipMemcpyHostToDevice, s));
    // results start at 0 (entries beyond D stay 0), the ok flags at 1
    const size_t wOk = size_t(n), wG = size_t(n) * 32 * (fp64 ? 2 : 1), wH = size_t(n) * 1024 * (fp64 ? 2 : 1);
    uint32_t *zg = reinterpret_cast<uint32_t *>(fp64 ? static_cast<void *>(g64) : static_cast<void *>(g32));
    uint32_t *zh = reinterpret_cast<uint32_t *>(fp64 ? static_cast<void *>(h64) : static_cast<void *>(h32));
    hipLaunchKernelGGL(k_na_fill, dim3(na_grid(wOk)), dim3(256), 0, s, reinterpret_cast<uint32_t *>(dOk), wOk, 1u);
    hipLaunchKernelGGL(k_na_fill, dim3(na_grid(wG)), dim3(256), 0, s, zg, wG, 0u);
    hipLaunchKernelGGL(k_na_fill, dim3(na_grid(wH)), dim3(256), 0, s, zh, wH, 0u);
    HIPCHK(hipGetLastError());
    const DrParams P = make_params(c);
    HIPCHK(hipEventRecord(c->ev[0], s));
    if (fp64)
        hipLaunchKernelGGL(k_gradhess_rec_f64, dim3(unsigned((total + 63) / 64)), dim3(64), 0, s, c->S, P.oDeriv, int(n), dRec, dPass, total,
                           g64, h64, dOk);
    else
        hipLaunchKernelGGL(k_gradhess_rec_f32, dim3(unsigned((total + 63) / 64)), dim3(64), 0, s, c->S, P.oDeriv, int(n), dRec, dPass, total,
                           g32, h32, dOk);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[1], s));
    const size_t wFin = size_t(n) * (32 + 1024);
    hipLaunchKernelGGL(k_na_finish, dim3(na_grid(wFin)), dim3(256), 0, s, int(n), int(fp64 != 0), g32, h32, g64, h64);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(grad, g64, size_t(n) * 32 * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hess, h64, size_t(n) * 1024 * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(okOut, dOk, size_t(n) * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    kernel_time(c);
    return DRMLT_OK;
}

extern "C" int drmlt_compute_gaussian_f64(int64_t n, const int32_t *D, const double *sc, double sigma, const double *grad, const double *hess,
                                          double *mean, double *covL, double *invCov, double *logDet) {
    if (n < 0 || !D || !sc || !grad || !hess || !mean || !covL || !invCov || !logDet) return fail(DRMLT_ERR_ARG, "bad argument");
    for (int64_t i = 0; i < n; i++)
        if (D[i] < 4 || D[i] > kGMax || D[i] % 2) return fail(DRMLT_ERR_ARG, "D must be even, 4 <= D <= 24");
    if (n == 0) return DRMLT_OK;
    hipStream_t s = 0;
    DBuf<int> d;
    DBuf<double> a, g, h, sl, m, L, Q, ld;
    DUP(d, D, n); DUP(a, sc, n); DUP(g, grad, 32 * n); DUP(h, hess, 1024 * n);
    DNEW(sl, size_t(kGSlot) * n); DNEW(m, 32 * size_t(n)); DNEW(L, size_t(kGMax * kGMax) * n); DNEW(Q, size_t(kGMax * kGMax) * n);
    DNEW(ld, n);
    const unsigned blocks = unsigned(std::min<int64_t>(n, 65535));
    hipLaunchKernelGGL(k_gauss_rec, dim3(blocks), dim3(64), 0, s, int(n), make_h2param(sigma), d.p, a.p, g.p, h.p, sl.p);
    HIPCHK(hipGetLastError());
    const int64_t ne = n * kGMax * kGMax;
    hipLaunchKernelGGL(k_gauss_expand, dim3(unsigned((ne + 63) / 64)), dim3(64), 0, s, int(n), d.p, sl.p, m.p, L.p, Q.p, ld.p);
    HIPCHK(hipGetLastError());
    DGET(mean, m, 32 * n); DGET(covL, L, size_t(kGMax * kGMax) * n); DGET(invCov, Q, size_t(kGMax * kGMax) * n); DGET(logDet, ld, n);
    HIPCHK(hipStreamSynchronize(s));
    return DRMLT_OK;
}

extern "C" int drmlt_gaussian_logpdf_f64(int64_t n, const int32_t *D, const double *x, double sign, const double *mean, const double *covL,
                                         const double *invCov, const double *logDet, double *out) {
    if (n < 0 || !D || !x || !mean || !covL || !invCov || !logDet || !out) return fail(DRMLT_ERR_ARG, "bad argument");
    if (n == 0) return DRMLT_OK;
    for (int64_t i = 0; i < n; i++)
        if (D[i] < 1 || D[i] > kGMax) return fail(DRMLT_ERR_ARG, "D out of range");
    hipStream_t s = 0;
    DBuf<int> d;
    DBuf<double> xx, dm, dQ, dl, o;
    DUP(d, D, n); DUP(xx, x, 32 * n); DUP(dm, mean, 32 * n); DUP(dQ, invCov, size_t(kGMax * kGMax) * n); DUP(dl, logDet, n); DNEW(o, n);
    hipLaunchKernelGGL(k_logpdf_rec, dim3(unsigned((n + 63) / 64)), dim3(64), 0, s, int(n), d.p, xx.p, sign, dm.p, dQ.p, dl.p, o.p);
    HIPCHK(hipGetLastError());
    DGET(out, o, n);
    HIPCHK(hipStreamSynchronize(s));
    return DRMLT_OK;
}

extern "C" int drmlt_gaussian_eval_f64(int64_t n, const int32_t *D, const double *sc, double sigma, const double *grad, const double *hess,
                                       const double *x, double sign, const double *nrm, double *logpdf, double *sample) {
    if (n < 0 || !D || !sc || !grad || !hess || !x || !nrm || !logpdf || !sample) return fail(DRMLT_ERR_ARG, "bad argument");
    for (int64_t i = 0; i < n; i++)
        if (D[i] < 4 || D[i] > kGMax || D[i] % 2) return fail(DRMLT_ERR_ARG, "D must be even, 4 <= D <= 24");
    if (n == 0) return DRMLT_OK;
    std::vector<int> small, big;
    for (int64_t i = 0; i < n; i++) (D[i] <= kGSmallD ? small : big).push_back(int(i));
    hipStream_t s = 0;
    DBuf<int> d, is, ib;
    DBuf<double> a, g, h, xx, nn, sl, lp, sm;
    DUP(d, D, n); DUP(a, sc, n); DUP(g, grad, 32 * n); DUP(h, hess, 1024 * n); DUP(xx, x, 32 * n); DUP(nn, nrm, 32 * n);
    DUP(is, small.data(), small.size()); DUP(ib, big.data(), big.size());
    DNEW(sl, size_t(kGSlot) * n); DNEW(lp, n); DNEW(sm, 32 * size_t(n));
    HIPCHK(hipMemsetAsync(sm.p, 0, 32 * size_t(n) * 8, s));
    const H2Param h2 = make_h2param(sigma);
    constexpr int NG = 64 / DRMLT_GAUSS_GROUP;
    if (!small.empty())
        hipLaunchKernelGGL(k_gauss_eval_rec, dim3(unsigned(std::min<size_t>((small.size() + NG - 1) / NG, 65535))), dim3(64), 0, s,
                           int(small.size()), is.p, 1, h2, d.p, a.p, g.p, h.p, xx.p, sign, nn.p, sl.p, lp.p, sm.p);
    if (!big.empty())
        hipLaunchKernelGGL(k_gauss_eval_rec, dim3(unsigned(std::min<size_t>(big.size(), 65535))), dim3(64), 0, s, int(big.size()), ib.p, 0,
                           h2, d.p, a.p, g.p, h.p, xx.p, sign, nn.p, sl.p, lp.p, sm.p);
    HIPCHK(hipGetLastError());
    DGET(logpdf, lp, n);
    DGET(sample, sm, 32 * n);
    HIPCHK(hipStreamSynchronize(s));
    return DRMLT_OK;
}

extern "C" int drmlt_dr_alpha2_f64(int64_t n, const double *in6, double *a2) {
    if (n < 0 || !in6 || !a2) return fail(DRMLT_ERR_ARG, "bad argument");
    if (n == 0) return DRMLT_OK;
    hipStream_t s = 0;
    DBuf<double> i, o;
    DUP(i, in6, 6 * n); DNEW(o, n);
    hipLaunchKernelGGL(k_alpha2, dim3(unsigned((n + 63) / 64)), dim3(64), 0, s, int(n), i.p, o.p);
    HIPCHK(hipGetLastError());
    DGET(a2, o, n);
    HIPCHK(hipStreamSynchronize(s));
    return DRMLT_OK;
}

// Kernel declarations for the host translation unit (definitions in k_*.hip).
#pragma once
#include "drmlt_kernels.h"

namespace drmlt {
constexpr int kPreMaxC = 64;  // pre-pass contributions kept per sample (k_loop.hip)
__global__ void k_dr_plain(DevScene, Chains, DrParams, uint32_t, float *, unsigned long long *);
__global__ void k_b_stage1_bin(Chains, DrParams, const uint32_t *, Lists);
__global__ void k_b_stage1(DevScene, Chains, DrParams, const uint32_t *, float *, unsigned long long *, Lists);
__global__ void k_b_deriv(DevScene, Chains, DrParams, Lists, int, int);
__global__ void k_b_gauss(DevScene, Chains, DrParams, Lists, int, int, int, const uint32_t *);
__global__ void k_b_stage2_propose(DevScene, Chains, DrParams, const uint32_t *, Lists);
__global__ void k_b_stage2_accept(DevScene, Chains, DrParams, uint32_t *, float *, unsigned long long *, Lists);
// the same path-stage kernels bounded to three waves per SIMD
__global__ void k_b_stage1_w3(DevScene, Chains, DrParams, const uint32_t *, float *, unsigned long long *, Lists);
__global__ void k_b_stage2_propose_w3(DevScene, Chains, DrParams, const uint32_t *, Lists);
__global__ void k_b_stage2_accept_w3(DevScene, Chains, DrParams, uint32_t *, float *, unsigned long long *, Lists);
__global__ void k_b_reset(Lists, uint32_t *, int);
// h2mc lens mutations: lens scores of the current states, then the mutations of list 5
__global__ void k_b_lens_score(DevScene, Chains, DrParams);
__global__ void k_b_lens(DevScene, Chains, DrParams, const uint32_t *, float *, unsigned long long *, Lists);
// counting forms of the stage kernels (k_count.hip)
__global__ void k_b_stage1_bin_counted(Chains, DrParams, const uint32_t *, Lists);
__global__ void k_b_stage1_counted(DevSceneCounted, Chains, DrParams, const uint32_t *, float *, unsigned long long *, Lists);
__global__ void k_b_deriv_counted(DevSceneCounted, Chains, DrParams, Lists, int, int);
__global__ void k_b_gauss_counted(DevSceneCounted, Chains, DrParams, Lists, int, int, int, const uint32_t *);
__global__ void k_b_stage2_propose_counted(DevSceneCounted, Chains, DrParams, const uint32_t *, Lists);
__global__ void k_b_stage2_accept_counted(DevSceneCounted, Chains, DrParams, uint32_t *, float *, unsigned long long *, Lists);
__global__ void k_prepass_gen(DevScene, Opts, int, int, uint64_t, int *, int *, float *, double *, double *);
__global__ void k_prepass_seed(DevScene, Opts, Chains, int, uint64_t, const int *, const int *);
__global__ void k_direct(DevScene, Opts, int, uint64_t, float *);
__global__ void k_trace(DevScene, int, const float *, const float *, const float *, const float *, int, int *, float *);
__global__ void k_genbidir_rec(DevScene, Opts, int, int, int, const float *, int, int, float *, int *);
__global__ void k_lens_rec(DevScene, Opts, int, char *, const float *, const float *, float *, int *);
__global__ void k_logf_rec_f32(DevScene, Opts, int, const char *, double *, int *);
__global__ void k_logf_rec_f64(DevScene, Opts, int, const char *, double *, int *);
__global__ void k_gradhess_rec_f32(DevScene, Opts, int, char *, const int *, int, float *, float *, int *);
__global__ void k_gradhess_rec_f64(DevScene, Opts, int, char *, const int *, int, double *, double *, int *);
__global__ void k_gauss_rec(int, H2Param, const int *, const double *, const double *, const double *, double *);
__global__ void k_gauss_expand(int, const int *, const double *, double *, double *, double *, double *);
__global__ void k_logpdf_rec(int, const int *, const double *, double, const double *, const double *, const double *, double *);
__global__ void k_alpha2(int, const double *, double *);
__global__ void k_gauss_eval_rec(int, const int *, int, H2Param, const int *, const double *, const double *, const double *, const double *,
                                 double, const double *, double *, double *, double *);
}  // namespace drmlt

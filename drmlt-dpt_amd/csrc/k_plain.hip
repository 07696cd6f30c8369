// Plain DR form: one thread runs one whole DR iteration of one chain (delayed.cpp:396-656).
#include "dr_stages.h"

namespace drmlt {

// ------------------------------------------------------------ plain form
__global__ void __launch_bounds__(64) k_dr_plain(DevScene Sc, Chains ch, DrParams P, uint32_t iter, float *film, unsigned long long *stats,
                                                 float *scratch) {
    const int chain = blockIdx.x * blockDim.x + threadIdx.x;
    if (chain >= ch.C) return;
    const int flags = dr_stage1(Sc, ch, P, chain, iter, film, stats);
    if (!(flags & 2)) return;
    float *gbuf = scratch + size_t(chain) * (kGMax + kGMax * kGMax);
    float *hbuf = gbuf + kGMax;
    const int perm = ch.perm[chain];
    const int pc = chain + ch.C * perm_cur(perm), ps = chain + ch.C * perm_second(perm);
    if (flags & 4) {
        if (!ch.gInit[pc]) init_gaussian_thread(Sc, ch, P, pc, gbuf, hbuf);
    }
    dr_stage2_propose(Sc, ch, P, chain, iter);
    if ((flags & 4) && (ch.flags[chain] & 16)) {
        init_gaussian_thread(Sc, ch, P, ps, gbuf, hbuf);
        dr_stage2_ratio(ch, chain);
    }
    dr_stage2_accept(Sc, ch, P, chain, iter, film, stats);
}

}  // namespace drmlt

// Plain DR form: one thread runs one whole DR iteration of one chain (delayed.cpp:396-656).
#include "dr_stages.h"

namespace drmlt {

// ------------------------------------------------------------ plain form
__global__ void __launch_bounds__(64) k_dr_plain(DevScene Sc, Chains ch, DrParams P, uint32_t iter, float *film, unsigned long long *stats) {
    const int chain = blockIdx.x * blockDim.x + threadIdx.x;
    if (chain >= ch.C) return;
    const int flags = dr_stage1(Sc, ch, P, chain, iter, film, stats);
    if (!(flags & 2)) return;
    const int perm = ch.perm[chain];
    const int pc = chain + ch.C * perm_cur(perm), ps = chain + ch.C * perm_second(perm);
    const SeqExec x;
    GWork w;
    if (flags & 4) {
        if (!ch.gInit[pc]) {
            derivs_thread(Sc, ch, P, chain, pc);
            gaussian_from_derivs(x, ch, P, chain, pc, w);
        }
        h2mc_offset(x, ch, P, chain, iter, w);
    }
    dr_stage2_propose(Sc, ch, P, chain, iter);
    if ((flags & 4) && (ch.flags[chain] & 16)) {
        derivs_thread(Sc, ch, P, chain, ps);
        gaussian_from_derivs(x, ch, P, chain, ps, w);
        h2mc_ratio(x, ch, chain, w);
    }
    dr_stage2_accept(Sc, ch, P, chain, iter, film, stats);
}

}  // namespace drmlt

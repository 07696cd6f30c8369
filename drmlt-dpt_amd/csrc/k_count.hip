// Counting kernels of the batched DR step: the stage kernels compiled with
// every traversal, path-state access and Gaussian counted into the launch's
// counters (drmlt_count_work). The product kernels contain none of this code.
#define DRMLT_COUNT_WORK 1
#include "k_stages.h"

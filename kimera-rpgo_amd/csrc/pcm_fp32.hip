// FP32 PCM pairwise kernel with FP64 near-threshold recheck (placeholder
// until the anchor-blocked formulation lands).
#include "rpgo_internal.h"

namespace rpgo {

int pcm_pairs_fp32(rpgo_ctx* ctx, double lc_thr, double margin, bool want_dist,
                   int64_t* n_recheck, int64_t* n_flip, float* ms_pairs, float* ms_recheck) {
  (void)ctx; (void)lc_thr; (void)margin; (void)want_dist;
  (void)n_recheck; (void)n_flip; (void)ms_pairs; (void)ms_recheck;
  set_error("rpgo_pcm_run: RPGO_PCM_FP32_RECHECK is not available in this build");
  return RPGO_ERR_ARG;
}

}  // namespace rpgo

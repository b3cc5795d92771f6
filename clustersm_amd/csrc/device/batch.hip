// Batched matching: persistent kernel over a device work queue.
#include "common.hpp"

extern "C" csm_status csm_batch_match(csm_ctx *ctx, const csm_graph *g, csm_query *const *qs, uint32_t n,
                                      const csm_batch_opts *opts, csm_result *out) {
  (void)g; (void)qs; (void)n; (void)opts; (void)out;
  return csm::set_error(ctx, CSM_E_ARG, "batched matching is not built yet");
}

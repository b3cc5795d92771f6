// Serving loop: continuous admission of queries against a resident graph.
#include "common.hpp"

namespace csm {
class Server {};
}  // namespace csm

extern "C" {
csm_status csm_serve_start(csm_ctx *ctx, const csm_graph *, const csm_serve_opts *) {
  return csm::set_error(ctx, CSM_E_ARG, "serving is not built yet");
}
csm_status csm_serve_submit(csm_ctx *ctx, const csm_query_plan *, uint64_t *) {
  return csm::set_error(ctx, CSM_E_ARG, "serving is not built yet");
}
csm_status csm_serve_poll(csm_ctx *ctx, uint64_t, int, csm_result *) {
  return csm::set_error(ctx, CSM_E_ARG, "serving is not built yet");
}
csm_status csm_serve_stop(csm_ctx *ctx) { return ctx ? CSM_OK : CSM_E_ARG; }
}

// Adjoint of the Chebyshev propagation (no reference counterpart).
#include "lightne_internal.h"

extern "C" int lightne_propagate_backward(lightne_ctx* ctx, const lightne_graph* g, const float* a, int64_t dim,
                                          const lightne_prop_params* p, int32_t with_aggregate,
                                          const float* grad_out, float* grad_a, double* grad_mu_host,
                                          double* grad_theta_host, int32_t checkpoint_every) {
  (void)g; (void)a; (void)dim; (void)p; (void)with_aggregate; (void)grad_out; (void)grad_a;
  (void)grad_mu_host; (void)grad_theta_host; (void)checkpoint_every;
  return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_propagate_backward: not built yet");
}

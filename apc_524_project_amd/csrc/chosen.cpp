// chosen.cpp -- FP32 chosen form (cell index + in-cell offset SoA, fused
// gather/push/deposit); not available in this build yet.
#include "emo_internal.hpp"

namespace emo {
int fp32_create(emo_ctx *c) {
  c->err = "the FP32 chosen form is not built into this library yet";
  return EMO_ERR_UNSUPPORTED;
}
void fp32_destroy(emo_ctx *) {}
int fp32_load(emo_ctx *, int64_t, const double *, const double *, const double *, const int32_t *,
              const int64_t *) {
  return EMO_ERR_UNSUPPORTED;
}
int fp32_get_particles(emo_ctx *, double *, double *, double *, int32_t *, int64_t *) { return EMO_ERR_UNSUPPORTED; }
int fp32_get_cells(emo_ctx *, int32_t *) { return EMO_ERR_UNSUPPORTED; }
int fp32_get_field(emo_ctx *, int, double *) { return EMO_ERR_UNSUPPORTED; }
int fp32_set_field(emo_ctx *, int, const double *) { return EMO_ERR_UNSUPPORTED; }
int fp32_init_constant(emo_ctx *) { return EMO_ERR_UNSUPPORTED; }
int fp32_init_poisson(emo_ctx *, int32_t, double, int32_t *, double *) { return EMO_ERR_UNSUPPORTED; }
int fp32_step(emo_ctx *, int32_t, double, double) { return EMO_ERR_UNSUPPORTED; }
int fp32_diag(emo_ctx *, double *) { return EMO_ERR_UNSUPPORTED; }
}  // namespace emo

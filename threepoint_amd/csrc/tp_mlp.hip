// tp_mlp.hip -- learned B_kappa surrogate (placeholder until the MLP kernels land).
#include "tp_internal.h"

void tp_surrogate_free(tp_ctx *ctx) { (void)ctx; }

int tp_surrogate_round(tp_ctx *ctx, const tp_cosmo *c, size_t nreg, const TpRegion *d_reg,
                       const TpIntegralDesc *d_desc, double *d_fvals, TpRegionOut *d_out, int *d_flag, int *fell_back)
{
    (void)ctx; (void)c; (void)nreg; (void)d_reg; (void)d_desc; (void)d_fvals; (void)d_out; (void)d_flag; (void)fell_back;
    return tp_set_error(TP_EINVAL, "surrogate not available in this build");
}
extern "C" int tp_surrogate_train(tp_ctx *, const tp_train_config *, tp_train_report *) { return tp_set_error(TP_EINVAL, "surrogate not available"); }
extern "C" int tp_surrogate_save(tp_ctx *, const char *) { return tp_set_error(TP_EINVAL, "surrogate not available"); }
extern "C" int tp_surrogate_load(tp_ctx *, const char *) { return tp_set_error(TP_EINVAL, "surrogate not available"); }
extern "C" int tp_surrogate_config(tp_ctx *, double, double, uint64_t) { return tp_set_error(TP_EINVAL, "surrogate not available"); }
extern "C" int tp_surrogate_predict(tp_ctx *, size_t, const double *, double *) { return tp_set_error(TP_EINVAL, "surrogate not available"); }
extern "C" int tp_mlp_num_params(size_t *) { return tp_set_error(TP_EINVAL, "surrogate not available"); }
extern "C" int tp_mlp_get_params(tp_ctx *, float *) { return tp_set_error(TP_EINVAL, "surrogate not available"); }
extern "C" int tp_mlp_set_params(tp_ctx *, const float *) { return tp_set_error(TP_EINVAL, "surrogate not available"); }
extern "C" int tp_mlp_loss_grad(tp_ctx *, size_t, const float *, const float *, double *, float *) { return tp_set_error(TP_EINVAL, "surrogate not available"); }

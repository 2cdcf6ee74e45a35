// dg_stage_p2.hip -- kernels and launch sequences for degree 2 (one TU per degree).
#include "dg_internal.hpp"

namespace oceano {
static __constant__ Basis<double> c_tab[4];
static __constant__ Basis<float> c_tabf[4];
}

#include "dg_launch.hpp"

namespace oceano {

int upload_tables_d2(const DegreeTables *t, const Basis<float> *tf)
{
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_tab), t, sizeof(DegreeTables) * 4));
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_tabf), tf, sizeof(Basis<float>) * 4));
  return OC_OK;
}

int stage_d2(oc_dg *op, const OpArgs &O, uint32_t stage, const double *factor,
             const double *ri_h, const double *ri_q, const double *ri_v, double *const *ki_h,
             double *const *ki_q, uint32_t n_ki, double *w_h, double *w_q, double *r_h,
             double *r_q, int write_scratch, hipStream_t s)
{
  if (op->precision == OC_PRECISION_MIXED)
    return launch_stage<2, float>(op, O, stage, factor, ri_h, ri_q, ri_v, ki_h, ki_q, n_ki, w_h,
                                 w_q, r_h, r_q, write_scratch, s);
  return launch_stage<2, double>(op, O, stage, factor, ri_h, ri_q, ri_v, ki_h, ki_q, n_ki, w_h, w_q,
                         r_h, r_q, write_scratch, s);
}

int rhs_apply_d2(oc_dg *op, const OpArgs &O, bool invert, const double *src_h,
                 const double *src_q, const double *src_v, double *dst_h, double *dst_q,
                 hipStream_t s)
{
  if (op->precision == OC_PRECISION_MIXED)
    return launch_rhs_apply<2, float>(op, O, invert, src_h, src_q, src_v, dst_h, dst_q, s);
  return launch_rhs_apply<2, double>(op, O, invert, src_h, src_q, src_v, dst_h, dst_q, s);
}

int post_d2(oc_dg *op, const OpArgs &O, int what, const double *zeta, const double *q,
             double *out, hipStream_t s)
{
  return launch_post<2>(op, O, what, zeta, q, out, s);
}

int stage_ark_d2(oc_dg *op, const OpArgs &O, uint32_t stage, uint32_t n_stages,
                  const double *factor, const double *factor_tilde, double fm,
                  const double *ri_h, const double *ri_q, const double *ri_v,
                  double *const *ki_h, double *const *ki_q, double *w_h, double *w_q,
                  double *r_h, double *r_q, hipStream_t s)
{
  if (op->precision == OC_PRECISION_MIXED)
    return launch_stage_ark<2, float>(op, O, stage, n_stages, factor, factor_tilde, fm, ri_h,
                                       ri_q, ri_v, ki_h, ki_q, w_h, w_q, r_h, r_q, s);
  return launch_stage_ark<2, double>(op, O, stage, n_stages, factor, factor_tilde, fm, ri_h,
                                      ri_q, ri_v, ki_h, ki_q, w_h, w_q, r_h, r_q, s);
}

int closure_d2(oc_dg *op, const OpArgs &O, const double *zeta, const double *q,
                float *features, hipStream_t s)
{
  return launch_closure<2>(op, O, zeta, q, features, s);
}

}  // namespace oceano

// dg_stage_degree.hip -- kernels and launch sequences of one polynomial degree.
//
// Compiled once per degree with -DOC_DEGREE=1..4 (oceano_amd/build.py): every degree is its
// own translation unit, with its own __constant__ basis tables and kernel instances, and the
// four compile in parallel.  The entry points are named <name>_d<degree>, declared in
// dg_internal.hpp.
#include "dg_internal.hpp"

#ifndef OC_DEGREE
#error "compile with -DOC_DEGREE=<1..4>"
#endif

namespace oceano {
static __constant__ Basis<double> c_tab[4];
static __constant__ Basis<float> c_tabf[4];
}

#include "dg_launch.hpp"

#define OC_CAT2(a, b) a##b
#define OC_CAT(a, b) OC_CAT2(a, b)
#define OC_NAME(name) OC_CAT(name, OC_CAT(_d, OC_DEGREE))

namespace oceano {

namespace {
constexpr int P = OC_DEGREE;
}

int OC_NAME(upload_tables)(const DegreeTables *t, const Basis<float> *tf)
{
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_tab), t, sizeof(DegreeTables) * 4));
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_tabf), tf, sizeof(Basis<float>) * 4));
  return OC_OK;
}

int OC_NAME(stage)(oc_dg *op, const OpArgs &O, uint32_t stage, const double *factor,
             const double *ri_h, const double *ri_q, const double *ri_v, double *const *ki_h,
             double *const *ki_q, uint32_t n_ki, double *w_h, double *w_q, double *r_h,
             double *r_q, int write_scratch, hipStream_t s)
{
  if (op->precision == OC_PRECISION_MIXED)
    return launch_stage<P, float>(op, O, stage, factor, ri_h, ri_q, ri_v, ki_h, ki_q, n_ki, w_h,
                                 w_q, r_h, r_q, write_scratch, s);
  return launch_stage<P, double>(op, O, stage, factor, ri_h, ri_q, ri_v, ki_h, ki_q, n_ki, w_h, w_q,
                         r_h, r_q, write_scratch, s);
}

int OC_NAME(rhs_apply)(oc_dg *op, const OpArgs &O, bool invert, const double *src_h,
                 const double *src_q, const double *src_v, double *dst_h, double *dst_q,
                 hipStream_t s)
{
  if (op->precision == OC_PRECISION_MIXED)
    return launch_rhs_apply<P, float>(op, O, invert, src_h, src_q, src_v, dst_h, dst_q, s);
  return launch_rhs_apply<P, double>(op, O, invert, src_h, src_q, src_v, dst_h, dst_q, s);
}

int OC_NAME(post)(oc_dg *op, const OpArgs &O, int what, const double *zeta, const double *q,
             double *out, hipStream_t s)
{
  return launch_post<P>(op, O, what, zeta, q, out, s);
}

int OC_NAME(stage_ark)(oc_dg *op, const OpArgs &O, uint32_t stage, uint32_t n_stages,
                  const double *factor, const double *factor_tilde, double fm,
                  const double *ri_h, const double *ri_q, const double *ri_v,
                  double *const *ki_h, double *const *ki_q, double *w_h, double *w_q,
                  double *r_h, double *r_q, hipStream_t s)
{
  if (op->precision == OC_PRECISION_MIXED)
    return launch_stage_ark<P, float>(op, O, stage, n_stages, factor, factor_tilde, fm, ri_h,
                                       ri_q, ri_v, ki_h, ki_q, w_h, w_q, r_h, r_q, s);
  return launch_stage_ark<P, double>(op, O, stage, n_stages, factor, factor_tilde, fm, ri_h,
                                      ri_q, ri_v, ki_h, ki_q, w_h, w_q, r_h, r_q, s);
}

int OC_NAME(closure)(oc_dg *op, const OpArgs &O, const double *zeta, const double *q,
                float *features, hipStream_t s)
{
  return launch_closure<P>(op, O, zeta, q, features, s);
}

}  // namespace oceano

// C-ABI of the ns2d tendency and the fused ns2d Runge-Kutta stages (single
// runs and batches of runs).
#include "fsx_internal.h"

extern "C" {

int fsx_ns2d_tendency(fsx_plan p, const void* rot_fft, void* out_fft,
                      double beta, int want_maxima, void* stream) {
  PlanScope scope_(p);
  if (!p || !rot_fft || !out_fft) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2) return fail(FSX_ERR_ARG, "ns2d needs a 2D plan");
  if (p->nf_max < 4) return fail(FSX_ERR_ARG, "ns2d needs nf_max >= 4");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? do_ns2d<float>(p, rot_fft, out_fft, beta, want_maxima, s)
                      : do_ns2d<double>(p, rot_fft, out_fft, beta, want_maxima, s);
}

int fsx_ns2d_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                      void* t1, double beta, const double* visc, double cfl,
                      double dt_max, void* stream) {
  PlanScope scope_(p);
  if (!p || !S || !t1 || !visc) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2 || p->nf_max < 4) return fail(FSX_ERR_ARG, "ns2d needs a 2D plan with nf_max >= 4");
  if (scheme != 4 && scheme != 2) return fail(FSX_ERR_ARG, "scheme must be 2 or 4");
  if (stage < 0 || stage >= scheme) return fail(FSX_ERR_ARG, "bad stage");
  if (scheme == 4 && !tmp) return fail(FSX_ERR_ARG, "RK4 needs tmp");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? do_ns2d_rk<float>(p, scheme, stage, S, tmp, t1, beta, visc,
                                          cfl, dt_max, s)
                      : do_ns2d_rk<double>(p, scheme, stage, S, tmp, t1, beta,
                                           visc, cfl, dt_max, s);
}

int fsx_ns2d_rk_stage_ext(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                          void* t1, double beta, const double* visc, double cfl,
                          double dt_max, const void* ext_fft, void* stream) {
  PlanScope scope_(p);
  if (!p || !S || !t1 || !visc) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2 || p->nf_max < 4) return fail(FSX_ERR_ARG, "ns2d needs a 2D plan with nf_max >= 4");
  if (scheme != 4 && scheme != 2) return fail(FSX_ERR_ARG, "scheme must be 2 or 4");
  if (stage < 0 || stage >= scheme) return fail(FSX_ERR_ARG, "bad stage");
  if (scheme == 4 && !tmp) return fail(FSX_ERR_ARG, "RK4 needs tmp");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? do_ns2d_rk<float>(p, scheme, stage, S, tmp, t1, beta, visc,
                                          cfl, dt_max, s, ext_fft)
                      : do_ns2d_rk<double>(p, scheme, stage, S, tmp, t1, beta,
                                           visc, cfl, dt_max, s, ext_fft);
}

int fsx_ns2d_batch_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp, void* t1,
                            double beta, void* stream) {
  PlanScope scope_(p);
  if (!p || !S || !t1 || (scheme == 4 && !tmp)) return fail(FSX_ERR_ARG, "bad argument");
  if (p->nb < 1 || !p->d_bpar) return fail(FSX_ERR_ARG, "not a batch plan");
  if (!((scheme == 4 && stage >= 0 && stage < 4) || (scheme == 2 && stage >= 0 && stage < 2)))
    return fail(FSX_ERR_ARG, "scheme must be 4 (stages 0-3) or 2 (stages 0-1)");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? do_ns2d_batch_rk<float>(p, scheme, stage, S, tmp, t1, beta, s)
                      : do_ns2d_batch_rk<double>(p, scheme, stage, S, tmp, t1, beta, s);
}

}  // extern "C"


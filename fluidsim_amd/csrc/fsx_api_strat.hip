// C-ABI of the ns3d.strat tendency and the fused ns3d.strat Runge-Kutta stage.
#include "fsx_internal.h"

extern "C" {

int fsx_strat_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                       void* t1, int has_f, double f, double N, int project,
                       int no_vz_kz0, const double* visc, double cfl,
                       double dt_max, void* stream) {
  PlanScope scope_(p);
  if (!p || !S || !t1 || !visc) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 3 || p->nf_max < 7)
    return fail(FSX_ERR_ARG, "ns3d.strat needs a 3D plan with nf_max >= 7");
  if (scheme != 4 && scheme != 2) return fail(FSX_ERR_ARG, "scheme must be 2 or 4");
  if (stage < 0 || stage >= scheme) return fail(FSX_ERR_ARG, "bad stage");
  if (scheme == 4 && !tmp) return fail(FSX_ERR_ARG, "RK4 needs tmp");
  if (project < 0 || project > 3) return fail(FSX_ERR_ARG, "bad projection kind");
  hipStream_t s = (hipStream_t)stream;
  void* src = stage == 0 ? S : t1;
  return p->prec == 4
             ? do_strat<float>(p, scheme, stage, S, tmp, t1, nullptr, src, has_f, f,
                               N * N, project, no_vz_kz0, visc, cfl, dt_max, false, s)
             : do_strat<double>(p, scheme, stage, S, tmp, t1, nullptr, src, has_f,
                                f, N * N, project, no_vz_kz0, visc, cfl, dt_max,
                                false, s);
}

int fsx_strat_tendency(fsx_plan p, const void* state, void* out_fft, int has_f,
                       double f, double N, int project, int no_vz_kz0,
                       int want_maxima, void* stream) {
  PlanScope scope_(p);
  if (!p || !state || !out_fft) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 3 || p->nf_max < 7)
    return fail(FSX_ERR_ARG, "ns3d.strat needs a 3D plan with nf_max >= 7");
  if (project < 0 || project > 3) return fail(FSX_ERR_ARG, "bad projection kind");
  hipStream_t s = (hipStream_t)stream;
  const double visc[4] = {0, 0, 0, 0};
  void* st = const_cast<void*>(state);
  // stage -1: the fused pass stores the tendency instead of a stage update
  return p->prec == 4
             ? do_strat<float>(p, 4, -1, st, nullptr, st, out_fft, state, has_f, f,
                               N * N, project, no_vz_kz0, visc, 0.0, 0.0,
                               want_maxima != 0, s)
             : do_strat<double>(p, 4, -1, st, nullptr, st, out_fft, state, has_f,
                                f, N * N, project, no_vz_kz0, visc, 0.0, 0.0,
                                want_maxima != 0, s);
}

}  // extern "C"


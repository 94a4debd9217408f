// C-ABI of the ns3d tendency and the fused ns3d Runge-Kutta stage.
#include "fsx_internal.h"

extern "C" {

int fsx_ns3d_tendency(fsx_plan p, const void* v_fft, void* out_fft, int has_f,
                      double f, int want_maxima, void* stream) {
  PlanScope scope_(p);
  if (!p || !v_fft || !out_fft) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 3) return fail(FSX_ERR_ARG, "ns3d needs a 3D plan");
  if (p->nf_max < 6) return fail(FSX_ERR_ARG, "ns3d needs nf_max >= 6");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4
             ? do_ns3d<float>(p, v_fft, out_fft, has_f, f, want_maxima, s)
             : do_ns3d<double>(p, v_fft, out_fft, has_f, f, want_maxima, s);
}

int fsx_ns3d_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                      void* t1, int has_f, double f, int project, int no_vz_kz0,
                      const double* visc, double cfl, double dt_max,
                      void* stream) {
  PlanScope scope_(p);
  if (!p || !S || !t1 || !visc) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 3 || p->nf_max < 6) return fail(FSX_ERR_ARG, "ns3d needs a 3D plan with nf_max >= 6");
  if (scheme != 4 && scheme != 2) return fail(FSX_ERR_ARG, "scheme must be 2 or 4");
  if (stage < 0 || stage >= scheme) return fail(FSX_ERR_ARG, "bad stage");
  if (scheme == 4 && !tmp) return fail(FSX_ERR_ARG, "RK4 needs tmp");
  if (project < 0 || project > 3) return fail(FSX_ERR_ARG, "bad projection kind");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4
             ? do_ns3d_rk<float>(p, scheme, stage, S, tmp, t1, has_f, f, project,
                                 no_vz_kz0, visc, cfl, dt_max, s)
             : do_ns3d_rk<double>(p, scheme, stage, S, tmp, t1, has_f, f, project,
                                  no_vz_kz0, visc, cfl, dt_max, s);
}

}  // extern "C"


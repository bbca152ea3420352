// nh_ocv.h -- device arrays and launchers for the steps of OCVTOT_NH around the solve (nh_ocv.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace onh {

// Device copies, in the caller's Fortran layout with halos, of the velocity-side metrics
// (attached once) and of the velocity state of the current step.
struct OcvArrays {
  int IE, JE, KE, i_start, j_start;
  int cyc_x, cyc_y, open_w, open_e, open_n, open_s;
  const double *dlxp, *dlyp, *dtdxpo, *dtdyo, *zo;  // (IE,JE)
  const double *ddpo, *weto;                        // (IE,JE,KE)
  const double *amsuo, *dduo;                       // (I_start:IE,JE,KE)
  const double *dtdxuo, *dlyu;                      // (I_start:IE,JE)
  const double *amsue, *ddue;                       // (IE,J_start:JE,KE)
  const double *dpye, *dlxv;                        // (IE,J_start:JE)
  const double *uoo, *uoo_1, *voe, *voe_1, *wno, *wno_1;  // velocity state (divergence inputs)
  double dt, dti;
};

// DIVG on the interior cells (ocvtot_nh.f90:11-62) into the (IE,JE,KE) array divg.
hipError_t launch_divergence(hipStream_t st, const OcvArrays& a, double* divg);
// bounds_exch (mo_parallel.f90:551-668) for one process on a(i0:IE, j0:JE, 1:nk).
hipError_t launch_exchange(hipStream_t st, double* arr, int IE, int JE, int nk, int i0, int j0, int cyc_x, int cyc_y);
// ocvtot_nh.f90:151-236: PNH (interior from the padded solution p64, halo by bounds_exch and the
// open-boundary copies), U/V pressure correction in place, WNO by continuity.
hipError_t launch_project(hipStream_t st, const OcvArrays& a, const double* p64, int pitch, long long plane,
                          double* pnh, double* uoo, double* voe, double* wno);

}  // namespace onh

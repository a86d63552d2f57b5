// setup.h - host-side start-up chain of the PCOM time stepper: the static
// fields every kernel reads and the initial state, computed once from the
// raw inputs the caller hands over (main.f90:252-704).
#pragma once

#include <string>
#include <vector>

#include "../../include/pcom_gpu.h"

namespace pcom {

struct Calendar {
  int y = 1, m = 1, d = 1, h = 0, mi = 0, s = 0, mm = 31;
};
int days_in_month(int y, int m);       // mod_type.f90:403-438
Calendar time_plus(Calendar t, int dt);  // mod_type.f90:593-635
// int_bnd month pair and weight (mod_int.f90:656-676); f is float32.
void bnd_weights(const Calendar& t, int* m1, int* m2, double* f);

struct HostSetup {
  int glo_ni = 0, glo_nj = 0, ni = 0, nj = 0, nk = 0, nkp = 0;
  size_t n2 = 0, n3 = 0, n3p = 0;
  // horizontal staggers g1j..g4j: index 1..4
  std::vector<double> rh[5], tn[5], dx1[5], dx2[5];
  std::vector<double> pr1, dz1, dpr1, z1, pr2, dz2, dpr2, z2;  // gi1 / gi2
  std::vector<int> lev_t, lev_u;
  std::vector<double> phih_t, prh_t, phih_u, prh_u;
  std::vector<double> cv1, cv2, cor, graphihx, graphihy;
  std::vector<double> km, kh;        // (ni,nj,nkp)
  std::vector<double> frc[6];        // taux tauy t s pa fw: (ni,nj,12)
  std::vector<double> tc, sc;        // initial T, S (ni,nj,nk)
};

// Returns "" on success or an error message.
std::string build_setup(const pcom_inputs& in, HostSetup& hs);

}  // namespace pcom

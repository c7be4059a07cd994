// Host-side driver of the adjoint's per-contact map (pe_amd/csrc/adjoint.hpp) for the CPU tests:
// reads contacts from stdin, one per line
//   model rad1 rad2 im1 im2 I1[9] I2[9] plane pn[3] pd xp[3] erp dtinv z[22]
// and prints for each the map's 15 outputs and its 22 x 15 Jacobian (one input direction a row),
// evaluated in dual numbers exactly as the device's k_adj_jacobian does.  No GPU is used.
#include <cstdio>

#include "adjoint.hpp"

using namespace pe_amd::ad;

template <int MODEL>
static void run(const ContactConst& K, const double* zv) {
  Dual z[kAdjIn], out[kAdjOut];
  for (int j = 0; j < kAdjIn; ++j) z[j] = mkd(zv[j]);
  contact_map<Dual, MODEL>(K, z, out);
  for (int k = 0; k < kAdjOut; ++k) printf("%.17g%c", out[k].v, k + 1 < kAdjOut ? ' ' : '\n');
  for (int j = 0; j < kAdjIn; ++j) {
    for (int k = 0; k < kAdjIn; ++k) z[k].d = k == j ? 1.0 : 0.0;
    contact_map<Dual, MODEL>(K, z, out);
    for (int k = 0; k < kAdjOut; ++k) printf("%.17g%c", out[k].d, k + 1 < kAdjOut ? ' ' : '\n');
  }
}

int main() {
  int model, plane;
  while (scanf("%d", &model) == 1) {
    ContactConst K{};
    double z[kAdjIn];
    if (scanf("%lf %lf %lf %lf", &K.rad1, &K.rad2, &K.im1, &K.im2) != 4) return 1;
    for (int k = 0; k < 9; ++k)
      if (scanf("%lf", &K.I1[k]) != 1) return 1;
    for (int k = 0; k < 9; ++k)
      if (scanf("%lf", &K.I2[k]) != 1) return 1;
    if (scanf("%d %lf %lf %lf %lf %lf %lf %lf %lf %lf", &plane, &K.pn[0], &K.pn[1], &K.pn[2], &K.pd, &K.xp[0],
              &K.xp[1], &K.xp[2], &K.erp, &K.dtinv) != 10)
      return 1;
    K.plane = plane;
    for (int j = 0; j < kAdjIn; ++j)
      if (scanf("%lf", &z[j]) != 1) return 1;
    if (model == 0) run<0>(K, z);
    else if (model == 1) run<1>(K, z);
    else if (model == 3) run<3>(K, z);
    else return 2;
  }
  return 0;
}

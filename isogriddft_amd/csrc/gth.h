// Local GTH pseudopotential radial function shared by the field, gradient and patch
// kernels (pseudo/local.py:138-180).
#pragma once

#include <hip/hip_runtime.h>

// One atom as the host passes it: position, valence charge, r_loc and C1..C4.
struct GthAtom {
  double x, y, z, zion, rloc, c[4];
};

// V_loc(r) = -Z erf(r / (sqrt2 r_loc)) / r + exp(-x^2/2) (C1 + C2 x^2 + C3 x^4 + C4 x^6),
// x = r / r_loc; r = 0 limit -Z sqrt(2/pi) / r_loc.
__device__ __forceinline__ double gth_vloc(double r, const GthAtom& a) {
  const double s = r / (1.4142135623730951 * a.rloc);
  const double coul = (r > 0.0) ? -a.zion * erf(s) / r : -a.zion * sqrt(2.0 / 3.14159265358979323846) / a.rloc;
  const double x2 = (r / a.rloc) * (r / a.rloc);
  const double poly = a.c[0] + x2 * (a.c[1] + x2 * (a.c[2] + x2 * a.c[3]));
  return coul + exp(-0.5 * x2) * poly;
}

// d V_loc / d r
__device__ __forceinline__ double gth_dvloc(double r, const GthAtom& a) {
  constexpr double PI_ = 3.14159265358979323846;
  const double al = 1.0 / (1.4142135623730951 * a.rloc);
  double dcoul = 0.0;
  if (r > 0.0) dcoul = a.zion * (erf(al * r) / (r * r) - 2.0 * al / sqrt(PI_) * exp(-(al * r) * (al * r)) / r);
  const double x = r / a.rloc;
  const double x2 = x * x;
  const double poly = a.c[0] + x2 * (a.c[1] + x2 * (a.c[2] + x2 * a.c[3]));
  const double dpoly = x * (2.0 * a.c[1] + x2 * (4.0 * a.c[2] + x2 * 6.0 * a.c[3]));
  return dcoul + exp(-0.5 * x2) * (dpoly - x * poly) / a.rloc;
}

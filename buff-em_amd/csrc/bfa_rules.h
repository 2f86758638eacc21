// bfa_rules.h -- fixed cubature rules shared by host and device code.
//
// TetCR4/16/33 are the symmetric tetrahedron rules of the reference
// (src/libs/libbuff/TetCR.cc:35-95: degree 2, 4 and 7).  They are built here
// from their orbits under vertex permutation, {u1,u2,u3,w} per point with
// x = V0 + u1 (V1-V0) + u2 (V2-V0) + u3 (V3-V0) and sum(w) = 1/6.
#pragma once
#include <cmath>

namespace bfa {

inline int tet_orbit(double *tab, int n, const double b[4], double w)
{
  for (int a = 0; a < 4; a++)
    for (int c = 0; c < 4; c++)
      for (int d = 0; d < 4; d++)
        for (int e = 0; e < 4; e++) {
          if (a == c || a == d || a == e || c == d || c == e || d == e) continue;
          double v1 = b[c], v2 = b[d], v3 = b[e];
          bool dup = false;
          for (int q = 0; q < n && !dup; q++)
            dup = tab[4 * q] == v1 && tab[4 * q + 1] == v2 && tab[4 * q + 2] == v3;
          if (dup) continue;
          tab[4 * n] = v1; tab[4 * n + 1] = v2; tab[4 * n + 2] = v3; tab[4 * n + 3] = w;
          n++;
        }
  return n;
}

// Fills tab with npts*4 doubles; returns npts (0 if unsupported).
inline int tet_rule(int npts, double *tab)
{
  if (npts == 4) {
    const double a = 0.13819660112501051, b[4] = {1.0 - 3 * a, a, a, a};
    return tet_orbit(tab, 0, b, 0.041666666666666666);
  }
  if (npts == 16) {
    const double b1[4] = {0.771642902067237, 0.076119032644254315, 0.076119032644254315, 0.076119032644254315};
    const double b2[4] = {0.07183164526766925, 0.11970052779780202, 0.40423391346726440, 0.40423391346726440};
    int n = tet_orbit(tab, 0, b1, 8.3956323500204695e-3);
    return tet_orbit(tab, n, b2, 0.011090344772215399);
  }
  if (npts == 33) {
    const double c[4] = {0.25, 0.25, 0.25, 0.25};
    const double s1[4] = {0.05777714911165610, 0.31407428362944793, 0.31407428362944793, 0.31407428362944793};
    const double s2[4] = {0.81235958507558810, 0.06254680497480394, 0.06254680497480394, 0.06254680497480394};
    const double s3[4] = {0.61738099120939840, 0.04120213879937276, 0.04120213879937276, 0.30021473119185615};
    const double s4[4] = {0.44311853122966216, 0.05292032110257111, 0.25198057383388334, 0.25198057383388334};
    int n = tet_orbit(tab, 0, c, 0.017783308938202055);
    n = tet_orbit(tab, n, s1, -0.019099406118263309);
    n = tet_orbit(tab, n, s2, 3.8252276536630236e-3);
    n = tet_orbit(tab, n, s3, 0.0035475499144103266);
    return tet_orbit(tab, n, s4, 0.013950789384495153);
  }
  return 0;
}

// Gauss-Legendre nodes/weights on [0,1] (Newton iteration on P_n).
inline void gauss01(int n, double *x, double *w)
{
  for (int i = 0; i < n; i++) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
    for (int it = 0; it < 100; it++) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        double p3 = p2; p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-16) break;
    }
    x[i] = 0.5 * (1.0 - z);
    w[i] = 1.0 / ((1.0 - z * z) * pp * pp);
  }
}

}  // namespace bfa

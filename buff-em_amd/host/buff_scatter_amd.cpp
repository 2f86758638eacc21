// buff_scatter_amd.cpp -- scattering driver on the drop-in C++ API.
//
// Reproduces the frequency loop of the reference's buff-scatter
// (src/applications/buff-scatter/buff-scatter.cc:236-363): per frequency
// assemble the VIE matrix, LU-factor it once, then per incident field
// assemble the RHS, solve in place and report the EMT absorbed/scattered
// power (the --PFTFile output, OutputModules.cc:149-211).
//
//   buff_scatter_amd (--geometry FILE.buffgeo | --vmsh FILE | --kuhn n R)
//                    [--eps RE IM]  (material of --vmsh / --kuhn)
//                    --omega W [--omega W ...]
//                    --pw EX EY EZ NX NY NZ [--pw ...]
//                    [--rytov T_EXPR T_ENV]
// Output: one line per (omega, incident field): omega  ifield  P_abs  P_scat.
// With --rytov, per omega also buff-neq's flux of this object: the EMT power
// of the dressed Rytov matrix (GetFlux.cc:309-399) for the temperature
// expression T_EXPR (SVTensor syntax, e.g. "300 + 40*x") and environment
// temperature T_ENV, as the line  omega  rytov  P_abs  P_scat.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "libbuff_amd.h"

using namespace buffamd;

int main(int argc, char **argv)
{
  const char *vmsh = nullptr, *geo = nullptr;
  int kuhn_n = 0;
  double R = 1.0, er = 4.0, ei = 0.1;
  std::vector<double> omegas;
  std::vector<PlaneWave> fields;
  const char *texpr = nullptr;
  double tenv = 0.0;
  for (int i = 1; i < argc; i++) {
    auto need = [&](int k) {
      if (i + k >= argc) { fprintf(stderr, "missing value for %s\n", argv[i]); exit(2); }
    };
    if (!strcmp(argv[i], "--vmsh")) { need(1); vmsh = argv[++i]; }
    else if (!strcmp(argv[i], "--geometry")) { need(1); geo = argv[++i]; }
    else if (!strcmp(argv[i], "--kuhn")) { need(2); kuhn_n = atoi(argv[++i]); R = atof(argv[++i]); }
    else if (!strcmp(argv[i], "--eps")) { need(2); er = atof(argv[++i]); ei = atof(argv[++i]); }
    else if (!strcmp(argv[i], "--omega")) { need(1); omegas.push_back(atof(argv[++i])); }
    else if (!strcmp(argv[i], "--pw")) {
      need(6);
      PlaneWave p;
      for (int c = 0; c < 3; c++) p.E0[c] = atof(argv[++i]);
      for (int c = 0; c < 3; c++) p.nHat[c] = atof(argv[++i]);
      fields.push_back(p);
    } else if (!strcmp(argv[i], "--rytov")) {
      need(2);
      texpr = argv[++i];
      tenv = atof(argv[++i]);
    } else { fprintf(stderr, "unknown option %s\n", argv[i]); return 2; }
  }
  if ((!vmsh && !geo && kuhn_n <= 0) || omegas.empty() || (fields.empty() && !texpr)) {
    fprintf(stderr, "usage: %s (--geometry FILE.buffgeo | --vmsh FILE | --kuhn n R) [--eps RE IM] --omega W "
                    "--pw EX EY EZ NX NY NZ [--rytov T_EXPR T_ENV]\n", argv[0]);
    return 2;
  }
  try {
    bfa_material mat{0, {er, ei, 0, 0, 0, 0}};
    SWGGeometry *G = geo ? new SWGGeometry(geo)
                   : vmsh ? new SWGGeometry(vmsh, mat) : new SWGGeometry(kuhn_n, R, 0.05, 0x5EED, mat);
    HMatrix *M = G->AllocateVIEMatrix();
    HVector *J = G->AllocateRHSVector();
    bfa_material T{};
    if (texpr) check(bfa_material_expr(texpr, &T), "temperature expression");
    HMatrix *DR = nullptr;
    for (double w : omegas) {
      if (texpr) {
        DR = G->ComputeDressedRytovMatrix(w, &T, tenv, DR);
        std::vector<double> P = G->GetPFTMatrix(DR, w);
        printf("%.6e rytov %.12e %.12e\n", w, P[0], P[1]);
      }
      if (fields.empty()) continue;
      M = G->AssembleVIEMatrix(w, M);
      int info = M->LUFactorize();
      if (info) { fprintf(stderr, "singular VIE matrix at omega=%g (info %d)\n", w, info); return 3; }
      for (size_t f = 0; f < fields.size(); f++) {
        std::vector<PlaneWave> IF(1, fields[f]);
        G->AssembleRHSVector(w, IF, J);
        M->LUSolve(J);
        std::vector<double> P = G->GetPFTMatrix(J, w, IF);
        printf("%.6e %zu %.12e %.12e\n", w, f, P[0], P[1]);
      }
    }
    delete DR;
    delete J;
    delete M;
    delete G;
    if (texpr) bfa_material_release(&T);
  } catch (const Error &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}

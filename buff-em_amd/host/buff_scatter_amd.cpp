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
// Output: one line per (omega, incident field): omega  ifield  P_abs  P_scat
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
    } else { fprintf(stderr, "unknown option %s\n", argv[i]); return 2; }
  }
  if ((!vmsh && !geo && kuhn_n <= 0) || omegas.empty() || fields.empty()) {
    fprintf(stderr, "usage: %s (--geometry FILE.buffgeo | --vmsh FILE | --kuhn n R) [--eps RE IM] --omega W "
                    "--pw EX EY EZ NX NY NZ\n", argv[0]);
    return 2;
  }
  try {
    bfa_material mat{0, {er, ei, 0, 0, 0, 0}};
    SWGGeometry *G = geo ? new SWGGeometry(geo)
                   : vmsh ? new SWGGeometry(vmsh, mat) : new SWGGeometry(kuhn_n, R, 0.05, 0x5EED, mat);
    HMatrix *M = G->AllocateVIEMatrix();
    HVector *J = G->AllocateRHSVector();
    for (double w : omegas) {
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
    delete J;
    delete M;
    delete G;
  } catch (const Error &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}

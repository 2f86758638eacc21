// lffield.cc -- low-frequency dielectric sphere through the source-compatible
// drop-in (libbuff.h): the calls of the reference's LFField unit test
// (src/unitTests/unit-test-LFField.cc:50-69) -- geometry file, VIE matrix,
// plane-wave RHS, LU factor and solve, scattered and incident fields at the
// points of an evaluation-point file.
//
//   lffield GEOFILE EPFILE OMEGA OUTFILE
// OUTFILE: per point, x y z, then the 6 complex scattered-field components
// (GetFields with IF = 0), then the 6 incident ones (GetFields with J = 0).
#include <cstdlib>

#include "libbuff.h"

using namespace scuff;
using namespace buff;

int main(int argc, char *argv[])
{
  if (argc != 5) {
    fprintf(stderr, "usage: %s GEOFILE EPFILE OMEGA OUTFILE\n", argv[0]);
    return 2;
  }
  SetLogFileName("lffield.log");
  Log("lffield running on %s", GetHostName());

  SWGGeometry *G = new SWGGeometry(argv[1]);
  HMatrix *M = G->AllocateVIEMatrix();
  HVector *J = G->AllocateRHSVector();
  cdouble Omega = atof(argv[3]);

  HMatrix *XMatrix = new HMatrix(argv[2], LHM_TEXT, "-ncol 3");
  if (XMatrix->ErrMsg) ErrExit(XMatrix->ErrMsg);

  cdouble E0[3] = {1.0, 0.0, 0.0};
  double nHat[3] = {0.0, 0.0, 1.0};
  PlaneWave *PW = new PlaneWave(E0, nHat);

  G->AssembleVIEMatrix(Omega, M);
  G->AssembleRHSVector(Omega, PW, J);
  M->LUFactorize();
  M->LUSolve(J);
  HMatrix *Scattered = G->GetFields(0, J, Omega, XMatrix);
  HMatrix *Incident = G->GetFields(PW, 0, Omega, XMatrix);

  FILE *f = fopen(argv[4], "w");
  if (!f) ErrExit("cannot write %s", argv[4]);
  for (int nx = 0; nx < XMatrix->NR; nx++) {
    double X[3];
    cdouble FS[6], FI[6];
    XMatrix->GetEntriesD(nx, "0:2", X);
    Scattered->GetEntries(nx, "0:5", FS);
    Incident->GetEntries(nx, "0:5", FI);
    fprintVec(f, X, 3);
    fprintVec(f, FS, 6);
    fprintVecCR(f, FI, 6);
  }
  fclose(f);
  delete Scattered;
  delete Incident;
  delete PW;
  delete XMatrix;
  delete J;
  delete M;
  delete G;
  return 0;
}

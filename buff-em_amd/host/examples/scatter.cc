// scatter.cc -- the scattering driver's frequency loop through the
// source-compatible drop-in (libbuff.h), in the structure of the reference's
// buff-scatter (src/applications/buff-scatter/buff-scatter.cc:182-314 and
// WritePFTFile, OutputModules.cc:149-211): incident plane waves chained
// into one IncField, a PFTOptions for the EMT method, and per frequency
// AssembleVIEMatrix, LUFactorize, then per incident-field chain
// AssembleRHSVector, a saved copy of the RHS, LUSolve and GetPFTMatrix.
//
//   scatter GEOFILE PFTFILE --omega W [--omega W ...]
//           --pw PX PY PZ DX DY DZ [--pw ...]
// PFTFILE gets one line per frequency and object: omega, label, then the
// NUMPFT quantities (absorbed, scattered power; force/torque columns NaN).
#include <cstdlib>
#include <vector>

#include "libbuff.h"

using namespace scuff;
using namespace buff;

static void write_pft(SWGGeometry *G, HVector *J, HVector *RHS, IncField *IF, cdouble Omega, PFTOptions *Options,
                      const char *PFTFile)
{
  Options->IF = IF;
  Options->RHSVector = RHS;
  Options->PFTMethod = SCUFF_PFT_EMT;
  HMatrix *PFTMatrix = G->GetPFTMatrix(J, Omega, Options);
  FILE *f = fopen(PFTFile, "a");
  if (!f) ErrExit("cannot append to %s", PFTFile);
  for (int no = 0; no < G->NumObjects; no++) {
    fprintf(f, "%e %s ", real(Omega), G->Objects[no]->Label);
    for (int nq = 0; nq < NUMPFT; nq++) fprintf(f, "%.15e ", PFTMatrix->GetEntryD(no, nq));
    fprintf(f, "\n");
  }
  fclose(f);
  delete PFTMatrix;
}

int main(int argc, char *argv[])
{
  if (argc < 3) {
    fprintf(stderr, "usage: %s GEOFILE PFTFILE --omega W ... --pw PX PY PZ DX DY DZ ...\n", argv[0]);
    return 2;
  }
  const char *GeoFile = argv[1], *PFTFile = argv[2];
  std::vector<double> omegas;
  std::vector<double> pol, dir;
  for (int i = 3; i < argc; i++) {
    if (!strcmp(argv[i], "--omega") && i + 1 < argc) omegas.push_back(atof(argv[++i]));
    else if (!strcmp(argv[i], "--pw") && i + 6 < argc) {
      for (int c = 0; c < 3; c++) pol.push_back(atof(argv[++i]));
      for (int c = 0; c < 3; c++) dir.push_back(atof(argv[++i]));
    } else ErrExit("unknown or incomplete option %s", argv[i]);
  }
  if (omegas.empty() || pol.empty()) ErrExit("need at least one --omega and one --pw");
  HVector *OmegaList = new HVector((int)omegas.size(), LHM_COMPLEX);
  for (size_t n = 0; n < omegas.size(); n++) OmegaList->SetEntry((int)n, cdouble(omegas[n], 0.0));

  // plane waves chained into one incident field, as buff-scatter builds it
  IncField *IF = 0;
  for (size_t npw = 0; npw < pol.size() / 3; npw++) {
    cdouble E0[3] = {pol[3 * npw], pol[3 * npw + 1], pol[3 * npw + 2]};
    PlaneWave *PW = new PlaneWave(E0, &dir[3 * npw]);
    PW->Next = IF;
    IF = PW;
  }
  IncFieldList *IFList = AddIncFieldToList(IF, const_cast<char *>("Default"));

  PFTOptions MyPFTOptions, *pftOptions = &MyPFTOptions;
  BUFF_InitPFTOptions(pftOptions);

  SWGGeometry *G = new SWGGeometry(GeoFile);
  HMatrix *M = G->AllocateVIEMatrix();
  HVector *RHS = G->AllocateRHSVector();
  HVector *J = G->AllocateRHSVector();

  char OmegaStr[MAXSTR];
  for (int nFreq = 0; nFreq < OmegaList->N; nFreq++) {
    cdouble Omega = OmegaList->GetEntry(nFreq);
    z2s(Omega, OmegaStr);
    Log("Working at frequency %s...", OmegaStr);
    G->AssembleVIEMatrix(Omega, M);
    Log("  LU-factorizing VIE matrix...");
    if (M->LUFactorize()) ErrExit("singular VIE matrix at omega=%s", OmegaStr);
    for (int nIF = 0; nIF < IFList->NumIFs; nIF++) {
      IncField *F = IFList->IFs[nIF];
      G->AssembleRHSVector(Omega, F, J);
      RHS->Copy(J);
      M->LUSolve(J);
      write_pft(G, J, RHS, F, Omega, pftOptions, PFTFile);
    }
  }
  delete J;
  delete RHS;
  delete M;
  delete G;
  delete OmegaList;
  return 0;
}

// libbuff.h -- source-compatible drop-in for buff-em's scattering API on MI355X.
//
// A caller of the reference library (src/libs/libbuff/libbuff.h:200-283)
// and of the SCUFF-EM pieces it uses on the dense VIE path -- HMatrix /
// HVector (libhmat), IncField / PlaneWave (libIncField), PFTOptions, and the
// libhrutil helpers ErrExit / Log -- compiles against this header unchanged
// apart from its include path, with everything heavy on the GPU through the
// C ABI of include/buffem_amd.h:
//
//   SWGGeometry *G = new SWGGeometry("E10Sphere_533.buffgeo");
//   HMatrix *M = G->AllocateVIEMatrix();           libbuff.h:230
//   HVector *J = G->AllocateRHSVector();           libbuff.h:232
//   G->AssembleVIEMatrix(Omega, M);                libbuff.h:231
//   G->AssembleRHSVector(Omega, IF, J);            libbuff.h:233 (IF: IncField chain)
//   M->LUFactorize();  M->LUSolve(J);              buff-scatter.cc:285,314
//   HMatrix *F = G->GetFields(IF, J, Omega, X);    libbuff.h:235-236
//   HMatrix *PFT = G->GetPFTMatrix(J, Omega, &Opt);libbuff.h:237-238
//
// Storage: HMatrix entries live in HBM (column-major, ld = NR; complex128 or
// real64), GetEntry/SetEntry/GetEntries move single entries or rows.  As in
// the reference: AssembleVIEMatrix reallocates a wrong-size matrix and
// returns the new one; LUFactorize overwrites the matrix with its factors;
// LUSolve works in place; GetPFTMatrix returns NumObjects x NUMPFT (row 0,
// columns PFT_PABS = 0 and PFT_PSCAT = 1; the force and torque columns
// 2..7, not computed here, are NaN); ErrExit prints and exits.
//
// Scope: one object per geometry (the reference's multi-object geometries
// are out of scope), plane-wave incident fields, the EMT power method
// (PFTOptions::PFTMethod = SCUFF_PFT_EMT, with or without a DRMatrix).
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <strings.h>
#include <limits>
#include <string>
#include <vector>
#include <unistd.h>

#include "../../include/buffem_amd.h"

typedef std::complex<double> cdouble;

namespace scuff {

#define LHM_REAL    0
#define LHM_COMPLEX 1
#define LHM_NORMAL  0
#define LHM_TEXT    1

#ifndef ZVAC
#define ZVAC 376.73031346177
#endif

// ---- libhrutil helpers the callers use ----------------------------------
inline FILE *&log_file()
{
  static FILE *f = nullptr;
  return f;
}
inline void SetLogFileName(const char *name)
{
  if (log_file()) fclose(log_file());
  log_file() = fopen(name, "a");
}
inline void Log(const char *fmt, ...)
{
  FILE *f = log_file() ? log_file() : stderr;
  va_list ap;
  va_start(ap, fmt);
  vfprintf(f, fmt, ap);
  va_end(ap);
  fputc('\n', f);
  fflush(f);
}
[[noreturn]] inline void ErrExit(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "error: ");
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fputc('\n', stderr);
  exit(1);
}
inline const char *GetHostName()
{
  static char name[256] = "";
  if (!name[0] && gethostname(name, sizeof(name) - 1) != 0) strcpy(name, "unknown");
  return name;
}
// full double precision, so that written results can be compared exactly
inline void fprintVec(FILE *f, const double *v, int n)
{
  for (int i = 0; i < n; i++) fprintf(f, "%+.16e ", v[i]);
}
inline void fprintVec(FILE *f, const cdouble *v, int n)
{
  for (int i = 0; i < n; i++) fprintf(f, "%+.16e %+.16e ", v[i].real(), v[i].imag());
}
inline void fprintVecCR(FILE *f, const double *v, int n) { fprintVec(f, v, n); fputc('\n', f); }
inline void fprintVecCR(FILE *f, const cdouble *v, int n) { fprintVec(f, v, n); fputc('\n', f); }
inline char *z2s(cdouble z, char *buf)
{
  if (z.imag() == 0.0) snprintf(buf, 64, "%g", z.real());
  else snprintf(buf, 64, "%g+%gi", z.real(), z.imag());
  return buf;
}

inline void check(int rc, const char *what)
{
  if (rc != BFA_OK) ErrExit("%s: %s", what, bfa_last_error());
}

// column ranges "a:b" (inclusive), ":" (all), "a"
inline void parse_range(const char *s, int n, int *lo, int *hi)
{
  const char *c = strchr(s, ':');
  if (!c) { *lo = *hi = atoi(s); }
  else {
    *lo = (c == s) ? 0 : atoi(s);
    *hi = c[1] ? atoi(c + 1) : n - 1;
  }
  if (*lo < 0 || *hi >= n || *lo > *hi) ErrExit("bad index range \"%s\" for %i columns", s, n);
}

// ---- HMatrix / HVector (libhmat): dense, device-resident ----------------
class HMatrix {
 public:
  int NR = 0, NC = 0, RealComplex = LHM_COMPLEX;
  char *ErrMsg = nullptr;

  HMatrix(int nr, int nc, int realcomplex = LHM_COMPLEX, int storage = LHM_NORMAL)
      : NR(nr), NC(nc), RealComplex(realcomplex)
  {
    (void)storage;
    alloc();
    Zero();
  }
  // text file: whitespace-separated numbers, one row per line; options
  // "-ncol N" keeps the first N columns (the reference's EPFile reader)
  HMatrix(const char *FileName, int FileType = LHM_TEXT, const char *Options = nullptr)
  {
    RealComplex = LHM_REAL;
    if (FileType != LHM_TEXT) { set_error("unsupported file type for %s", FileName); return; }
    int ncol = -1;
    if (Options && sscanf(Options, " -ncol %d", &ncol) != 1) ncol = -1;
    FILE *f = fopen(FileName, "r");
    if (!f) { set_error("could not open file %s", FileName); return; }
    std::vector<std::vector<double>> rows;
    char line[4096];
    while (fgets(line, sizeof line, f)) {
      std::vector<double> r;
      char *p = line, *end = nullptr;
      if (*p == '#') continue;
      for (double v = strtod(p, &end); end != p; v = strtod(p, &end)) {
        r.push_back(v);
        p = end;
      }
      if (r.empty()) continue;
      if (ncol > 0) {
        if ((int)r.size() < ncol) { fclose(f); set_error("%s: line with fewer than %i columns", FileName, ncol); return; }
        r.resize(ncol);
      }
      if (!rows.empty() && r.size() != rows[0].size()) {
        fclose(f);
        set_error("%s: rows of unequal length", FileName);
        return;
      }
      rows.push_back(r);
    }
    fclose(f);
    if (rows.empty()) { set_error("%s: no data", FileName); return; }
    NR = (int)rows.size();
    NC = (int)rows[0].size();
    alloc();
    std::vector<double> h((size_t)NR * NC);
    for (int r = 0; r < NR; r++)
      for (int c = 0; c < NC; c++) h[(size_t)c * NR + r] = rows[r][c];
    check(bfa_memcpy_htod(d_, h.data(), h.size() * 8, nullptr), "HMatrix upload");
  }
  virtual ~HMatrix()
  {
    if (d_) bfa_device_free(d_);
    if (ipiv_) bfa_device_free(ipiv_);
    free(ErrMsg);
  }
  HMatrix(const HMatrix &) = delete;
  HMatrix &operator=(const HMatrix &) = delete;

  void *DevicePtr() { return d_; }
  const void *DevicePtr() const { return d_; }
  size_t ElemBytes() const { return RealComplex == LHM_COMPLEX ? 16 : 8; }

  cdouble GetEntry(int r, int c) const
  {
    bounds(r, c);
    double v[2] = {0.0, 0.0};
    check(bfa_memcpy_dtoh(v, at(r, c), ElemBytes(), nullptr), "HMatrix::GetEntry");
    return cdouble(v[0], v[1]);
  }
  double GetEntryD(int r, int c) const { return GetEntry(r, c).real(); }
  void SetEntry(int r, int c, cdouble z)
  {
    bounds(r, c);
    double v[2] = {z.real(), z.imag()};
    if (RealComplex == LHM_REAL && z.imag() != 0.0) ErrExit("HMatrix::SetEntry: complex value in a real matrix");
    check(bfa_memcpy_htod(at(r, c), v, ElemBytes(), nullptr), "HMatrix::SetEntry");
    factored_ = false;
  }
  void SetEntry(int r, int c, double x) { SetEntry(r, c, cdouble(x, 0.0)); }
  void AddEntry(int r, int c, cdouble z) { SetEntry(r, c, GetEntry(r, c) + z); }
  void AddEntry(int r, int c, double x) { AddEntry(r, c, cdouble(x, 0.0)); }

  void GetEntries(int r, const char *cols, cdouble *out) const
  {
    int lo, hi;
    parse_range(cols, NC, &lo, &hi);
    for (int c = lo; c <= hi; c++) out[c - lo] = GetEntry(r, c);
  }
  void GetEntriesD(int r, const char *cols, double *out) const
  {
    int lo, hi;
    parse_range(cols, NC, &lo, &hi);
    for (int c = lo; c <= hi; c++) out[c - lo] = GetEntryD(r, c);
  }
  void SetEntries(int r, const char *cols, const cdouble *in)
  {
    int lo, hi;
    parse_range(cols, NC, &lo, &hi);
    for (int c = lo; c <= hi; c++) SetEntry(r, c, in[c - lo]);
  }
  void SetEntriesD(int r, const char *cols, const double *in)
  {
    int lo, hi;
    parse_range(cols, NC, &lo, &hi);
    for (int c = lo; c <= hi; c++) SetEntry(r, c, in[c - lo]);
  }

  void Zero()
  {
    std::vector<char> z((size_t)NR * NC * ElemBytes(), 0);
    if (!z.empty()) check(bfa_memcpy_htod(d_, z.data(), z.size(), nullptr), "HMatrix::Zero");
    factored_ = false;
  }
  // entries of a matrix of the same size and type (the reference copies
  // between equal shapes, e.g. BSD->RHS->Copy(J), buff-scatter.cc:301)
  void Copy(const HMatrix *src)
  {
    if (src->NR != NR || src->NC != NC || src->RealComplex != RealComplex)
      ErrExit("HMatrix::Copy: size or type mismatch (%ix%i <- %ix%i)", NR, NC, src->NR, src->NC);
    std::vector<char> h((size_t)NR * NC * ElemBytes());
    check(bfa_memcpy_dtoh(h.data(), src->d_, h.size(), nullptr), "HMatrix::Copy");
    check(bfa_memcpy_htod(d_, h.data(), h.size(), nullptr), "HMatrix::Copy");
    factored_ = false;
  }

  // HMatrix::LUFactorize: zgetrf semantics in place (buff-scatter.cc:285);
  // returns 0 or the 1-based column of the first exactly zero pivot
  int LUFactorize()
  {
    if (NR != NC || RealComplex != LHM_COMPLEX) ErrExit("LUFactorize: needs a square complex matrix");
    if (!ipiv_) check(bfa_device_malloc((void **)&ipiv_, (size_t)NR * 4), "pivot allocation");
    int info = 0;
    check(bfa_lu_factor_z(d_, NR, NR, ipiv_, &info, nullptr), "LUFactorize");
    factored_ = true;
    return info;
  }
  // HMatrix::LUSolve(HVector *) / LUSolve(HMatrix *): in place, every column
  // of B a right-hand side (buff-scatter.cc:314, GetFlux.cc:265)
  int LUSolve(HMatrix *B)
  {
    if (!factored_) ErrExit("LUSolve: LUFactorize was not called");
    if (B->NR != NR || B->RealComplex != LHM_COMPLEX) ErrExit("LUSolve: right-hand side of the wrong size or type");
    check(bfa_lu_solve_z(d_, NR, NR, ipiv_, B->d_, B->NC, B->NR, nullptr), "LUSolve");
    return 0;
  }

 protected:
  HMatrix() = default;
  void alloc()
  {
    const size_t bytes = std::max<size_t>((size_t)NR * NC * ElemBytes(), 16);
    check(bfa_device_malloc(&d_, bytes), "HMatrix allocation");
  }
  void *at(int r, int c) const { return (char *)d_ + ((size_t)c * NR + r) * ElemBytes(); }
  void bounds(int r, int c) const
  {
    if (r < 0 || r >= NR || c < 0 || c >= NC) ErrExit("HMatrix: entry (%i,%i) outside %ix%i", r, c, NR, NC);
  }
  void set_error(const char *fmt, ...)
  {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    free(ErrMsg);
    ErrMsg = strdup(buf);
  }
  void *d_ = nullptr;
  int *ipiv_ = nullptr;
  bool factored_ = false;
};

class HVector : public HMatrix {
 public:
  int N = 0;
  explicit HVector(int n, int realcomplex = LHM_COMPLEX) : HMatrix(n, 1, realcomplex), N(n) {}
  cdouble GetEntry(int i) const { return HMatrix::GetEntry(i, 0); }
  double GetEntryD(int i) const { return HMatrix::GetEntryD(i, 0); }
  void SetEntry(int i, cdouble z) { HMatrix::SetEntry(i, 0, z); }
  void SetEntry(int i, double x) { HMatrix::SetEntry(i, 0, x); }
  void AddEntry(int i, cdouble z) { HMatrix::AddEntry(i, 0, z); }
  using HMatrix::GetEntry;
  using HMatrix::SetEntry;
};

// ---- incident fields (libIncField) --------------------------------------
class IncField {
 public:
  IncField *Next = nullptr;
  cdouble Omega = 0.0;
  virtual ~IncField() {}
  void SetFrequency(cdouble W, bool Traverse = true)
  {
    Omega = W;
    if (Traverse)
      for (IncField *f = Next; f; f = f->Next) f->Omega = W;
  }
  // E and H of this element alone at X (vacuum, k = Omega)
  virtual void GetFields(const double X[3], cdouble EH[6]) = 0;
  // the element as a plane wave for the device path (false: not one)
  virtual bool AsPlaneWave(bfa_planewave *pw) const = 0;
};

// E = E0 exp(i Omega nHat.x), H = nHat x E / Z0
class PlaneWave : public IncField {
 public:
  cdouble E0[3];
  double nHat[3];
  PlaneWave(const cdouble pE0[3], const double pnHat[3])
  {
    for (int i = 0; i < 3; i++) { E0[i] = pE0[i]; nHat[i] = pnHat[i]; }
  }
  void GetFields(const double X[3], cdouble EH[6]) override
  {
    const cdouble ph = std::exp(cdouble(0.0, 1.0) * Omega * (nHat[0] * X[0] + nHat[1] * X[1] + nHat[2] * X[2]));
    for (int i = 0; i < 3; i++) EH[i] = E0[i] * ph;
    EH[3] = (nHat[1] * EH[2] - nHat[2] * EH[1]) / ZVAC;
    EH[4] = (nHat[2] * EH[0] - nHat[0] * EH[2]) / ZVAC;
    EH[5] = (nHat[0] * EH[1] - nHat[1] * EH[0]) / ZVAC;
  }
  bool AsPlaneWave(bfa_planewave *pw) const override
  {
    for (int c = 0; c < 3; c++) {
      pw->E0[2 * c] = E0[c].real();
      pw->E0[2 * c + 1] = E0[c].imag();
      pw->dir[c] = nHat[c];
    }
    return true;
  }
};

// the elements of an IncField chain as plane waves
inline std::vector<bfa_planewave> plane_waves(const IncField *IF, const char *who)
{
  std::vector<bfa_planewave> v;
  for (const IncField *f = IF; f; f = f->Next) {
    bfa_planewave pw;
    if (!f->AsPlaneWave(&pw)) ErrExit("%s: only plane-wave incident fields are supported", who);
    v.push_back(pw);
  }
  return v;
}

// a list of labelled incident-field chains (buff-scatter's --IFFile / default)
struct IncFieldList {
  int NumIFs = 0;
  IncField **IFs = nullptr;
  char **Labels = nullptr;
};
inline IncFieldList *AddIncFieldToList(IncField *IF, char *Label, IncFieldList *List = nullptr)
{
  if (!List) List = new IncFieldList;
  List->IFs = (IncField **)realloc(List->IFs, (List->NumIFs + 1) * sizeof(IncField *));
  List->Labels = (char **)realloc(List->Labels, (List->NumIFs + 1) * sizeof(char *));
  List->IFs[List->NumIFs] = IF;
  List->Labels[List->NumIFs] = Label ? strdup(Label) : nullptr;
  List->NumIFs++;
  return List;
}

// ---- power, force, torque options (PFTOptions.h) ------------------------
#define SCUFF_PFT_OVERLAP 0
#define SCUFF_PFT_EMT     1
#define SCUFF_PFT_MOMENTS 2
#define SCUFF_PFT_DSI     3
#define PFT_PABS  0
#define PFT_PSCAT 1

struct PFTOptions {
  int PFTMethod = SCUFF_PFT_EMT;
  IncField *IF = nullptr;
  HVector *RHSVector = nullptr;
  HMatrix *DRMatrix = nullptr;
  char *DSIMesh = nullptr;
  double DSIRadius = 5.0;
  int DSIPoints = 110;
};
inline PFTOptions *InitPFTOptions(PFTOptions *o = nullptr)
{
  if (!o) o = new PFTOptions;
  *o = PFTOptions();
  return o;
}

}  // namespace scuff

namespace buff {

using namespace scuff;

#define NUMPFT 8
#define MAXSTR 200

// the object of a one-object geometry: its label (.buffgeo OBJECT name),
// mesh file and number of basis functions
struct SWGVolume {
  char *Label = nullptr;
  char *MeshFileName = nullptr;
  int NumInteriorFaces = 0;
  ~SWGVolume() { free(Label); free(MeshFileName); }
};

inline PFTOptions *BUFF_InitPFTOptions(PFTOptions *o = nullptr) { return InitPFTOptions(o); }

class SWGGeometry {
 public:
  int NumObjects = 1, TotalBFs = 0, LogLevel = 0;
  SWGVolume **Objects = nullptr;
  char *GeoFileName = nullptr;

  // one-object .buffgeo file (SWGGeometry.cc:214-383): MESHFILE, MATERIAL
  // or SVTENSOR, DISPLACED / ROTATED; the static (FIBBI) records are built
  // on the device once per geometry
  explicit SWGGeometry(const char *GeoFile)
  {
    int nv = 0, nt = 0;
    check(bfa_buffgeo_load(GeoFile, &nv, &nt, nullptr, nullptr, nullptr), "SWGGeometry");
    std::vector<double> V(3 * (size_t)nv);
    std::vector<int> T(4 * (size_t)nt);
    check(bfa_buffgeo_load(GeoFile, &nv, &nt, V.data(), T.data(), &mat_), "SWGGeometry");
    check(bfa_geometry_create(nv, V.data(), nt, T.data(), &g_), "SWGGeometry");
    check(bfa_geometry_info(g_, nullptr, nullptr, &TotalBFs, nullptr), "SWGGeometry");
    check(bfa_geometry_upload(g_, nullptr), "SWGGeometry upload");
    check(bfa_static_build(g_, 0, nullptr), "static (FIBBI) records");
    GeoFileName = strdup(GeoFile);
    Objects = new SWGVolume *[1];
    Objects[0] = new SWGVolume;
    Objects[0]->NumInteriorFaces = TotalBFs;
    read_labels(GeoFileName, Objects[0]);
  }
  ~SWGGeometry()
  {
    if (Objects) { delete Objects[0]; delete[] Objects; }
    free(GeoFileName);
    if (g_) bfa_geometry_destroy(g_);
    bfa_material_release(&mat_);
  }
  SWGGeometry(const SWGGeometry &) = delete;
  SWGGeometry &operator=(const SWGGeometry &) = delete;

  HMatrix *AllocateVIEMatrix(bool PureImagFreq = false)
  {
    (void)PureImagFreq;
    return new HMatrix(TotalBFs, TotalBFs, LHM_COMPLEX);
  }
  // a wrong-size M is deleted and a new one returned (VIEMatrix.cc:433-439)
  HMatrix *AssembleVIEMatrix(cdouble Omega, HMatrix *M)
  {
    if (M && (M->NR != TotalBFs || M->NC != TotalBFs || M->RealComplex != LHM_COMPLEX)) { delete M; M = nullptr; }
    if (!M) M = AllocateVIEMatrix();
    check(bfa_assemble_vie(g_, Omega.real(), Omega.imag(), &mat_, M->DevicePtr(), TotalBFs, nullptr),
          "AssembleVIEMatrix");
    check(bfa_stream_synchronize(nullptr), "AssembleVIEMatrix");
    return M;
  }
  HVector *AllocateRHSVector() { return new HVector(TotalBFs, LHM_COMPLEX); }
  // RHS_a = -1/(i Omega Z0) <b_a, E_inc>, E_inc summed over the IF chain
  // (RHSVector.cc:62-124)
  HVector *AssembleRHSVector(cdouble Omega, IncField *IF, HVector *RHS)
  {
    if (!RHS || RHS->N != TotalBFs) { delete RHS; RHS = AllocateRHSVector(); }
    if (Omega.imag() != 0.0) ErrExit("AssembleRHSVector: real frequencies only");
    std::vector<bfa_planewave> pw = plane_waves(IF, "AssembleRHSVector");
    if (pw.empty()) ErrExit("AssembleRHSVector: no incident field");
    if (IF) IF->SetFrequency(Omega, true);
    HMatrix cols(TotalBFs, (int)pw.size(), LHM_COMPLEX);
    check(bfa_assemble_rhs(g_, Omega.real(), pw.data(), (int)pw.size(), cols.DevicePtr(), TotalBFs, nullptr),
          "AssembleRHSVector");
    std::vector<cdouble> h((size_t)TotalBFs * pw.size()), s(TotalBFs, 0.0);
    check(bfa_memcpy_dtoh(h.data(), cols.DevicePtr(), h.size() * 16, nullptr), "AssembleRHSVector");
    for (size_t j = 0; j < pw.size(); j++)
      for (int a = 0; a < TotalBFs; a++) s[a] += h[j * TotalBFs + a];
    check(bfa_memcpy_htod(RHS->DevicePtr(), s.data(), s.size() * 16, nullptr), "AssembleRHSVector");
    return RHS;
  }

  // total E and H at the rows of XMatrix (NR x >= 3): the IF chain's fields
  // plus those of the currents J (GetFields.cc:288-382); either may be null.
  // FMatrix (NR x 6, complex) is reallocated when missing or of the wrong size.
  HMatrix *GetFields(IncField *IF, HVector *J, cdouble Omega, HMatrix *XMatrix, HMatrix *FMatrix = nullptr)
  {
    if (!XMatrix || XMatrix->NC < 3 || XMatrix->NR == 0)
      ErrExit("wrong-size XMatrix (%ix%i) passed to GetFields", XMatrix ? XMatrix->NR : 0, XMatrix ? XMatrix->NC : 0);
    if (FMatrix && (FMatrix->NR != XMatrix->NR || FMatrix->NC != 6)) FMatrix = nullptr;
    if (!FMatrix) FMatrix = new HMatrix(XMatrix->NR, 6, LHM_COMPLEX);
    const int n = XMatrix->NR;
    std::vector<double> X(3 * (size_t)n);
    for (int r = 0; r < n; r++) XMatrix->GetEntriesD(r, "0:2", &X[3 * (size_t)r]);
    std::vector<cdouble> F = fields(IF, J, Omega, n, X.data());
    std::vector<cdouble> colmajor((size_t)n * 6);
    for (int r = 0; r < n; r++)
      for (int c = 0; c < 6; c++) colmajor[(size_t)c * n + r] = F[(size_t)r * 6 + c];
    check(bfa_memcpy_htod(FMatrix->DevicePtr(), colmajor.data(), colmajor.size() * 16, nullptr), "GetFields");
    return FMatrix;
  }
  void GetFields(IncField *IF, HVector *J, cdouble Omega, double *X, cdouble *EH)
  {
    std::vector<cdouble> F = fields(IF, J, Omega, 1, X);
    for (int c = 0; c < 6; c++) EH[c] = F[c];
  }

  // EMT absorbed and scattered power (GetPFT.cc:133-189, EMTPFT.cc:324-546)
  // of J under the incident chain Options->IF, or of Options->DRMatrix
  HMatrix *GetPFTMatrix(HVector *JVector, cdouble Omega, PFTOptions *Options = nullptr, HMatrix *PFTMatrix = nullptr)
  {
    if (PFTMatrix && (PFTMatrix->NR != NumObjects || PFTMatrix->NC != NUMPFT)) { delete PFTMatrix; PFTMatrix = nullptr; }
    if (!PFTMatrix) PFTMatrix = new HMatrix(NumObjects, NUMPFT, LHM_REAL);
    PFTOptions defaults;
    if (!Options) Options = InitPFTOptions(&defaults);
    if (Options->PFTMethod != SCUFF_PFT_EMT) ErrExit("GetPFTMatrix: only the EMT method is available");
    double pabs = 0.0, pscat = 0.0;
    if (Options->DRMatrix) {
      double out[2];
      check(bfa_emt_power_dr(g_, Omega.real(), Options->DRMatrix->DevicePtr(), TotalBFs, out, nullptr),
            "GetPFTMatrix");
      pabs = out[0];
      pscat = out[1];
    } else {
      if (!JVector) ErrExit("GetPFTMatrix: no current vector");
      std::vector<bfa_planewave> pw = plane_waves(Options->IF, "GetPFTMatrix");
      if (pw.empty()) ErrExit("GetPFTMatrix: the EMT method needs the incident field (PFTOptions::IF)");
      // P_scat depends on J alone; P_ext is linear in the incident field,
      // so over a chain P_abs = sum_k (P_abs,k + P_scat) - P_scat
      double pext = 0.0;
      for (const bfa_planewave &p : pw) {
        double out[2];
        check(bfa_emt_power(g_, Omega.real(), &p, 1, JVector->DevicePtr(), TotalBFs, out, nullptr), "GetPFTMatrix");
        pext += out[0] + out[1];
        pscat = out[1];
      }
      pabs = pext - pscat;
    }
    const double nan = std::numeric_limits<double>::quiet_NaN();
    double row[NUMPFT] = {pabs, pscat, nan, nan, nan, nan, nan, nan};
    PFTMatrix->SetEntriesD(0, ":", row);
    return PFTMatrix;
  }

 private:
  // OBJECT label and MESHFILE of the .buffgeo file
  static void read_labels(const char *path, SWGVolume *O)
  {
    FILE *f = fopen(path, "r");
    if (!f) return;
    char line[1024], key[256], val[768];
    while (fgets(line, sizeof line, f)) {
      if (sscanf(line, " %255s %767s", key, val) != 2) continue;
      if (!strcasecmp(key, "OBJECT") && !O->Label) O->Label = strdup(val);
      else if (!strcasecmp(key, "MESHFILE") && !O->MeshFileName) O->MeshFileName = strdup(val);
    }
    fclose(f);
    if (!O->Label) O->Label = strdup(O->MeshFileName ? O->MeshFileName : "Object");
  }
  std::vector<cdouble> fields(IncField *IF, HVector *J, cdouble Omega, int n, const double *X)
  {
    std::vector<bfa_planewave> pw = IF ? plane_waves(IF, "GetFields") : std::vector<bfa_planewave>();
    if (IF) IF->SetFrequency(Omega, true);
    std::vector<cdouble> F((size_t)n * 6);
    check(bfa_get_fields(g_, Omega.real(), pw.empty() ? nullptr : pw.data(), (int)pw.size(),
                         J ? J->DevicePtr() : nullptr, n, X, (double *)F.data(), nullptr),
          "GetFields");
    return F;
  }
  bfa_geometry *g_ = nullptr;
  bfa_material mat_{};
};

}  // namespace buff

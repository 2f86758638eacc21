// libbuff_amd.h -- C++ drop-in for the reference's scattering API on MI355X.
//
// Mirrors the subset of HomerReid/buff-em's public C++ interface that the
// dense VIE hot path uses (src/libs/libbuff/libbuff.h:200-283 and SCUFF-EM's
// HMatrix/HVector as called by src/applications/buff-scatter/buff-scatter.cc:
// 236-314), with device-resident storage underneath:
//
//   SWGGeometry G(meshfile | Kuhn sphere);        SWGGeometry::SWGGeometry
//   HMatrix *M = G.AllocateVIEMatrix();           libbuff.h:230
//   M = G.AssembleVIEMatrix(Omega, M);            libbuff.h:231 (VIEMatrix.cc:431-460)
//   HVector *J = G.AllocateRHSVector();           libbuff.h:232
//   G.AssembleRHSVector(Omega, IF, J);            libbuff.h:233 (RHSVector.cc:85-124)
//   M->LUFactorize();  M->LUSolve(J);             buff-scatter.cc:285,314
//   HMatrix *PFT = G.GetPFTMatrix(J, Omega, IF);  libbuff.h:237 (EMT method)
//
// Conventions follow the reference: AssembleVIEMatrix reallocates and
// returns a new matrix when handed one of the wrong size, the matrix is
// overwritten in place by its LU factors, LUSolve works in place.  Errors
// throw buffamd::Error (the reference calls ErrExit).  Matrices and vectors
// live in HBM; GetEntry/SetEntry move single entries for inspection.
#pragma once
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/buffem_amd.h"

namespace buffamd {

typedef std::complex<double> cdouble;

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void check(int rc, const char *what)
{
  if (rc != BFA_OK) throw Error(rc, std::string(what) + ": " + bfa_last_error());
}

// Dense complex128 matrix in HBM (column-major, ld = NR); after
// LUFactorize() it holds the LU factors and the pivot vector.
class HMatrix {
 public:
  int NR, NC;
  HMatrix(int nr, int nc) : NR(nr), NC(nc)
  {
    check(bfa_device_malloc(&d_, (size_t)nr * nc * 16), "HMatrix alloc");
  }
  ~HMatrix()
  {
    if (d_) bfa_device_free(d_);
    if (ipiv_) bfa_device_free(ipiv_);
  }
  HMatrix(const HMatrix &) = delete;
  HMatrix &operator=(const HMatrix &) = delete;

  void *data() { return d_; }
  const void *data() const { return d_; }
  bool factored() const { return factored_; }

  cdouble GetEntry(int r, int c) const
  {
    double v[2];
    check(bfa_memcpy_dtoh(v, (const char *)d_ + ((size_t)c * NR + r) * 16, 16, nullptr), "GetEntry");
    return cdouble(v[0], v[1]);
  }
  void SetEntry(int r, int c, cdouble z)
  {
    double v[2] = {z.real(), z.imag()};
    check(bfa_memcpy_htod((char *)d_ + ((size_t)c * NR + r) * 16, v, 16, nullptr), "SetEntry");
    factored_ = false;
  }

  // HMatrix::LUFactorize -> zgetrf semantics, in place
  int LUFactorize()
  {
    if (NR != NC) throw Error(BFA_ERR_ARG, "LUFactorize: matrix not square");
    if (!ipiv_) check(bfa_device_malloc((void **)&ipiv_, (size_t)NR * 4), "pivot alloc");
    int info = 0;
    check(bfa_lu_factor_z(d_, NR, NR, ipiv_, &info, nullptr), "LUFactorize");
    factored_ = true;
    return info;
  }

  // HMatrix::LUSolve(HMatrix *) : in-place multi-RHS solve (GetFlux.cc:265)
  void LUSolve(HMatrix *B) const
  {
    if (!factored_) throw Error(BFA_ERR_STATE, "LUSolve: LUFactorize not called");
    if (B->NR != NR) throw Error(BFA_ERR_ARG, "LUSolve: size mismatch");
    check(bfa_lu_solve_z(d_, NR, NR, ipiv_, B->d_, B->NC, B->NR, nullptr), "LUSolve");
  }

 protected:
  void *d_ = nullptr;
  int *ipiv_ = nullptr;
  bool factored_ = false;
};

// Column vector: an N x 1 HMatrix (HVector in the reference).
class HVector : public HMatrix {
 public:
  int N;
  explicit HVector(int n) : HMatrix(n, 1), N(n) {}
  cdouble GetEntry(int i) const { return HMatrix::GetEntry(i, 0); }
  void SetEntry(int i, cdouble z) { HMatrix::SetEntry(i, 0, z); }
  std::vector<cdouble> Download() const
  {
    std::vector<cdouble> h(N);
    check(bfa_memcpy_dtoh(h.data(), d_, (size_t)N * 16, nullptr), "HVector download");
    return h;
  }
};

// HMatrix::LUSolve(HVector *) : in place
inline void LUSolve(const HMatrix &M, HVector *J) { M.LUSolve(J); }

// Incident plane wave list (PlaneWave chained IncField, buff-scatter.cc:186)
struct PlaneWave {
  cdouble E0[3];
  double nHat[3];
};

class SWGGeometry {
 public:
  int TotalBFs = 0;

  // GMSH .msh/.vmsh mesh, one object, the given material.  As the
  // reference's per-mesh FIBBI cache (FIBBICache.cc:108-143, VIEMatrix.cc:
  // 382-424): the static records are preloaded from
  // ${BUFF_CACHE_PATH}/<mesh base>.cache, else ./<mesh base>.cache, records
  // missing there are computed, and the cache is written back after the
  // next AssembleVIEMatrix when anything was computed.
  SWGGeometry(const char *MeshFile, const bfa_material &mat, int td_order = 0) : mat_(mat), mesh_(MeshFile)
  {
    int nv = 0, nt = 0;
    check(bfa_mesh_read_vmsh(MeshFile, &nv, &nt, nullptr, nullptr), "mesh size");
    std::vector<double> V(3 * (size_t)nv);
    std::vector<int> T(4 * (size_t)nt);
    check(bfa_mesh_read_vmsh(MeshFile, &nv, &nt, V.data(), T.data()), "mesh read");
    init(nv, V.data(), nt, T.data(), td_order);
  }
  // .buffgeo geometry file (SWGGeometry::SWGGeometry(const char *GeoFileName),
  // SWGGeometry.cc:214-383): one object, its MESHFILE, MATERIAL or SVTENSOR
  // and DISPLACED/ROTATED transformations; the FIBBI cache is keyed by the
  // mesh file as in the reference
  explicit SWGGeometry(const char *GeoFile, int td_order = 0)
  {
    int nv = 0, nt = 0;
    check(bfa_buffgeo_load(GeoFile, &nv, &nt, nullptr, nullptr, nullptr), "geometry file");
    std::vector<double> V(3 * (size_t)nv);
    std::vector<int> T(4 * (size_t)nt);
    check(bfa_buffgeo_load(GeoFile, &nv, &nt, V.data(), T.data(), &mat_), "geometry file");
    owns_mat_ = true;
    init(nv, V.data(), nt, T.data(), td_order);
  }
  // synthetic Kuhn-sphere mesh
  SWGGeometry(int n, double radius, double jitter, uint64_t seed, const bfa_material &mat, int td_order = 0)
      : mat_(mat)
  {
    int nv = 0, nt = 0;
    check(bfa_mesh_kuhn_sphere(n, radius, jitter, seed, &nv, &nt, nullptr, nullptr), "mesh size");
    std::vector<double> V(3 * (size_t)nv);
    std::vector<int> T(4 * (size_t)nt);
    check(bfa_mesh_kuhn_sphere(n, radius, jitter, seed, &nv, &nt, V.data(), T.data()), "mesh build");
    init(nv, V.data(), nt, T.data(), td_order);
  }
  ~SWGGeometry()
  {
    delete sinv_;
    if (work_) bfa_device_free(work_);
    if (g_) bfa_geometry_destroy(g_);
    if (owns_mat_) bfa_material_release(&mat_);
  }
  SWGGeometry(const SWGGeometry &) = delete;
  SWGGeometry &operator=(const SWGGeometry &) = delete;

  HMatrix *AllocateVIEMatrix() { return new HMatrix(TotalBFs, TotalBFs); }
  HVector *AllocateRHSVector() { return new HVector(TotalBFs); }

  // wrong-size M is deleted and a new one returned (VIEMatrix.cc:433-439)
  HMatrix *AssembleVIEMatrix(cdouble Omega, HMatrix *M)
  {
    if (M && (M->NR != TotalBFs || M->NC != TotalBFs)) { delete M; M = nullptr; }
    if (!M) M = AllocateVIEMatrix();
    check(bfa_assemble_vie(g_, Omega.real(), Omega.imag(), &mat_, M->data(), TotalBFs, nullptr),
          "AssembleVIEMatrix");
    check(bfa_stream_synchronize(nullptr), "AssembleVIEMatrix sync");
    store_cache();   // GCache->Store(OA->MeshFileName), VIEMatrix.cc:420-424
    return M;
  }

  // The incident fields of one IncField chain are summed (RHSVector.cc:
  // 62-80); pass one PlaneWave per chain element.
  HVector *AssembleRHSVector(cdouble Omega, const std::vector<PlaneWave> &IF, HVector *RHS)
  {
    if (!RHS) RHS = AllocateRHSVector();
    if (IF.empty()) throw Error(BFA_ERR_ARG, "AssembleRHSVector: no incident field");
    std::vector<bfa_planewave> pw(IF.size());
    for (size_t i = 0; i < IF.size(); i++)
      for (int c = 0; c < 3; c++) {
        pw[i].E0[2 * c] = IF[i].E0[c].real();
        pw[i].E0[2 * c + 1] = IF[i].E0[c].imag();
        pw[i].dir[c] = IF[i].nHat[c];
      }
    // one column per plane wave, then summed into RHS
    HMatrix cols(TotalBFs, (int)IF.size());
    check(bfa_assemble_rhs(g_, Omega.real(), pw.data(), (int)pw.size(), cols.data(), TotalBFs, nullptr),
          "AssembleRHSVector");
    std::vector<cdouble> h((size_t)TotalBFs * IF.size()), s(TotalBFs, 0.0);
    check(bfa_memcpy_dtoh(h.data(), cols.data(), h.size() * 16, nullptr), "RHS download");
    for (size_t j = 0; j < IF.size(); j++)
      for (int a = 0; a < TotalBFs; a++) s[a] += h[j * TotalBFs + a];
    check(bfa_memcpy_htod(RHS->data(), s.data(), s.size() * 16, nullptr), "RHS upload");
    return RHS;
  }

  // SWGGeometry::GetFields (GetFields.cc:288-382): total E and H at the rows
  // of X (NR x 3, row-major) from the incident chain IF and the currents J
  // (either may be empty / null); returns NR x 6 (Ex Ey Ez Hx Hy Hz).
  std::vector<cdouble> GetFields(const std::vector<PlaneWave> &IF, const HVector *J, cdouble Omega,
                                 const std::vector<double> &X)
  {
    if (X.size() % 3) throw Error(BFA_ERR_ARG, "GetFields: X must hold 3 coordinates per point");
    std::vector<bfa_planewave> pw(IF.size());
    for (size_t i = 0; i < IF.size(); i++)
      for (int c = 0; c < 3; c++) {
        pw[i].E0[2 * c] = IF[i].E0[c].real();
        pw[i].E0[2 * c + 1] = IF[i].E0[c].imag();
        pw[i].dir[c] = IF[i].nHat[c];
      }
    const int npts = (int)(X.size() / 3);
    std::vector<cdouble> F((size_t)npts * 6);
    check(bfa_get_fields(g_, Omega.real(), pw.empty() ? nullptr : pw.data(), (int)pw.size(), J ? J->data() : nullptr,
                         npts, X.data(), (double *)F.data(), nullptr),
          "GetFields");
    return F;
  }

  // GetPFTMatrix with the EMT method: returns {P_abs, P_scat} for the
  // object (columns 0 and 1 of the reference's NumObjects x 8 matrix).
  std::vector<double> GetPFTMatrix(const HVector *J, cdouble Omega, const std::vector<PlaneWave> &IF)
  {
    if (IF.size() != 1)
      throw Error(BFA_ERR_ARG, "GetPFTMatrix: pass the single plane wave the RHS was built from");
    bfa_planewave pw;
    for (int c = 0; c < 3; c++) {
      pw.E0[2 * c] = IF[0].E0[c].real();
      pw.E0[2 * c + 1] = IF[0].E0[c].imag();
      pw.dir[c] = IF[0].nHat[c];
    }
    double out[2];
    check(bfa_emt_power(g_, Omega.real(), &pw, 1, J->data(), TotalBFs, out, nullptr), "GetPFTMatrix");
    return {out[0], out[1]};
  }

  // ---- buff-neq (one object) ----------------------------------------------
  // AssembleGBlock(0, 0, Omega, G) (VIEMatrix.cc:365-426)
  HMatrix *AssembleGBlock(int noa, int nob, cdouble Omega, HMatrix *G)
  {
    if (noa != 0 || nob != 0) throw Error(BFA_ERR_ARG, "AssembleGBlock: one object");
    if (!G) G = AllocateVIEMatrix();
    check(bfa_assemble_g(g_, Omega.real(), Omega.imag(), G->data(), TotalBFs, nullptr), "AssembleGBlock");
    return G;
  }

  // AssembleOverlapBlocks(0, Omega, TSVT, TAvg, ThetaEnv, V, VInv, Rytov)
  // (VIEMatrix.cc:321-360): rows of <= 7 entries; the temperature is a
  // compiled expression (nullptr: none) and T_avg is taken from it as
  // GetAverageTemperature does.
  struct OverlapRows {
    std::vector<int> cnt, cols;             // N, N x 7
    std::vector<cdouble> V, VInv;           // N x 7
    std::vector<double> Rytov;              // N x 7
  };
  OverlapRows AssembleOverlapBlocks(cdouble Omega, const bfa_material *Temperature = nullptr,
                                    double TEnvironment = 0.0)
  {
    const size_t n = (size_t)TotalBFs, m = n * 7;
    OverlapRows r{std::vector<int>(n), std::vector<int>(m), std::vector<cdouble>(m), std::vector<cdouble>(m),
                  std::vector<double>(m)};
    check(bfa_overlap_blocks(g_, Omega.real(), &mat_, Temperature, TEnvironment, r.cnt.data(), r.cols.data(),
                             (double *)r.V.data(), (double *)r.VInv.data(), r.Rytov.data()),
          "AssembleOverlapBlocks");
    return r;
  }

  // SInverse of CreateBNEQData (CreateBNEQData.cc:119-142): the inverse SWG
  // overlap matrix, computed on first use and kept
  const HMatrix *GetOverlapInverse()
  {
    if (!sinv_) {
      sinv_ = new HMatrix(TotalBFs, TotalBFs);
      check(bfa_gram_inverse(g_, sinv_->data(), TotalBFs, nullptr), "overlap matrix inverse");
    }
    return sinv_;
  }

  // ComputeDressedRytovMatrix (buff-neq GetFlux.cc:155-262) for this
  // object: DR = W (S^-1 Rytov S^-1)^H, W = (1 + S^-1 V S^-1 G)^-1.  DR is
  // reallocated when missing or of the wrong size; the work matrices are kept.
  HMatrix *ComputeDressedRytovMatrix(cdouble Omega, const bfa_material *Temperature, double TEnvironment,
                                     HMatrix *DR)
  {
    if (DR && (DR->NR != TotalBFs || DR->NC != TotalBFs)) { delete DR; DR = nullptr; }
    if (!DR) DR = AllocateVIEMatrix();
    const HMatrix *S = GetOverlapInverse();
    if (!work_) check(bfa_device_malloc(&work_, bfa_dressed_rytov_work_bytes(TotalBFs, TotalBFs)), "work alloc");
    check(bfa_dressed_rytov(g_, Omega.real(), &mat_, Temperature, TEnvironment, S->data(), DR->data(), TotalBFs,
                            work_, nullptr),
          "ComputeDressedRytovMatrix");
    return DR;
  }

  // GetPFTMatrix(IF = 0, Omega, {EMT, DRMatrix}) as GetFlux calls it:
  // {P_abs, P_scat} of the dressed Rytov matrix (EMTPFT.cc:389-476)
  std::vector<double> GetPFTMatrix(const HMatrix *DRMatrix, cdouble Omega)
  {
    if (DRMatrix->NR != TotalBFs || DRMatrix->NC != TotalBFs) throw Error(BFA_ERR_ARG, "GetPFTMatrix: DR size");
    double out[2];
    check(bfa_emt_power_dr(g_, Omega.real(), DRMatrix->data(), TotalBFs, out, nullptr), "GetPFTMatrix(DR)");
    return {out[0], out[1]};
  }

 private:
  void init(int nv, const double *V, int nt, const int *T, int td_order)
  {
    check(bfa_geometry_create(nv, V, nt, T, &g_), "SWGGeometry");
    int nbf = 0;
    check(bfa_geometry_info(g_, nullptr, nullptr, &nbf, nullptr), "SWGGeometry info");
    TotalBFs = nbf;
    check(bfa_geometry_upload(g_, nullptr), "geometry upload");
    // the frequency-independent static table plays the role of the
    // reference's per-mesh FIBBI cache (FIBBICache.cc:108-143)
    if (!mesh_.empty()) {
      for (const std::string &path : cache_candidates()) {
        FILE *f = fopen(path.c_str(), "rb");
        if (!f) continue;
        fclose(f);
        int64_t hits = 0, misses = 0;
        if (bfa_fibbi_preload(g_, path.c_str(), td_order, &hits, &misses, nullptr) == BFA_OK) {
          fprintf(stderr, "FIBBI cache: preloaded %s (%lld hits, %lld computed)\n", path.c_str(), (long long)hits,
                  (long long)misses);
          cache_dirty_ = misses > 0;
          return;
        }
        fprintf(stderr, "FIBBI cache: %s: %s (skipping cache preload)\n", path.c_str(), bfa_last_error());
      }
      cache_dirty_ = true;
    }
    check(bfa_static_build(g_, td_order, nullptr), "static table");
  }
  // FIBBICache::Store's file name: ${BUFF_CACHE_PATH}/<base>.cache or
  // ./<base>.cache, <base> = mesh file name without directory and extension
  std::string cache_base() const
  {
    std::string b = mesh_.substr(mesh_.find_last_of('/') == std::string::npos ? 0 : mesh_.find_last_of('/') + 1);
    const size_t dot = b.find_last_of('.');
    if (dot != std::string::npos && dot > 0) b.resize(dot);
    return b + ".cache";
  }
  std::vector<std::string> cache_candidates() const
  {
    std::vector<std::string> v;
    if (const char *dir = getenv("BUFF_CACHE_PATH")) v.push_back(std::string(dir) + "/" + cache_base());
    v.push_back(cache_base());
    return v;
  }
  void store_cache()
  {
    if (mesh_.empty() || !cache_dirty_) return;
    const std::string path = cache_candidates().front();
    int64_t n = 0;
    if (bfa_fibbi_store(g_, path.c_str(), &n) == BFA_OK) {
      fprintf(stderr, "FIBBI cache: wrote %lld records to %s\n", (long long)n, path.c_str());
      cache_dirty_ = false;
    } else {
      fprintf(stderr, "FIBBI cache: %s (aborting cache dump)\n", bfa_last_error());
    }
  }
  bfa_geometry *g_ = nullptr;
  bfa_material mat_;
  std::string mesh_;
  bool cache_dirty_ = false;
  bool owns_mat_ = false;
  HMatrix *sinv_ = nullptr;     // buff-neq: S^-1, built on first use
  void *work_ = nullptr;        // buff-neq: dressed-Rytov work matrices
};

}  // namespace buffamd

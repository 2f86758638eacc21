// bfa_internal.h -- internal types shared by the host code and the HIP
// kernels of the dense VIE hot path.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/buffem_amd.h"

namespace bfa {

// Thread-local error message, set by fail().
int fail(int status, const char *fmt, ...);

// BF tile edge of the tetrahedron-pair far-field kernel and the stride of
// its per-tile table: [2*FT2] distinct tetrahedra (-1 padded), [2*FT2] slot
// of BF (local l, side s) at 2l+s in that list, then the count.
#define FT2 32
#define FTAB (4 * FT2 + 4)

// Device-side view of the geometry (all pointers are device pointers).
// Cubature tables hold, per tetrahedron and point, {x, y, z, w} with
// w = 6 * Vol_signed * w_rule (src/libs/libbuff/Cubature.cc:201-218).
struct DevGeom {
  int nv, nt, nbf;
  int64_t nnz;            // touching BF pairs (a <= b)
  const double *V;        // [3*nv]
  const int *tetVI;       // [4*nt]
  const int *tetFI;       // [4*nt]
  const double *tetVol;   // [nt] signed
  const double *q4;       // [nt][4][4]
  const double *q16;      // [nt][16][4]
  const double *q33;      // [nt][33][4]
  const int *bfV;         // [nbf][5]  QP, V1, V2, V3, QM
  const int *bfT;         // [nbf][2]  PTet, MTet
  const double *bfPF;     // [nbf][2]  +Area/(3 V_P), -Area/(3 V_M) (signed V)
  const double *bfQ;      // [nbf][6]  QP xyz, QM xyz
  const double *bfRad;    // [nbf]
  const int *rowptr;      // [nbf+1]   touching pairs CSR
  const int *cols;        // [nnz]
  const int *pairRow;     // [nnz]     row of each CSR entry
  const int *ftTab;       // [ceil(nbf/FT2)][FTAB] far-field tile tetrahedra
};

// Device buffers owned by a geometry handle.
struct DevBuffers {
  void *V = nullptr, *tetVI = nullptr, *tetFI = nullptr, *tetVol = nullptr;
  void *q4 = nullptr, *q16 = nullptr, *q33 = nullptr;
  void *bfV = nullptr, *bfT = nullptr, *bfPF = nullptr, *bfQ = nullptr, *bfRad = nullptr;
  void *rowptr = nullptr, *cols = nullptr, *pairRow = nullptr, *ftTab = nullptr;
  // static (cold) data
  void *tdPairs = nullptr;     // [ntd] int4 {tA, tB, ncv, 0}
  void *tdVerts = nullptr;     // [ntd][24] double: VA(4x3), VB(4x3) ordered
  void *tdMom = nullptr;       // [ntd][3][8] moments
  void *pairTD = nullptr;      // [nnz][4] int: TD index*2+swap, or -1 for 16-pt
  void *table = nullptr;       // [nnz][6] FIBBI records
  void *tpTab = nullptr;       // far field: 16 moments per tetrahedron pair S <= T (per frequency)
  size_t tpTabBytes = 0;
};

}  // namespace bfa

struct bfa_geometry;
namespace bfa {
// touching-pair pass of the assembly alone (near field + static + Gram)
int assemble_near_only(bfa_geometry *g, double omega_re, double omega_im, const bfa_material *mat, void *d_M,
                       int64_t ld, void *stream);
}  // namespace bfa

struct bfa_geometry {
  int nv = 0, nt = 0, nbf = 0, nftot = 0;
  std::vector<double> V;                      // 3*nv
  std::vector<int> tetVI, tetFI;              // 4*nt
  std::vector<double> tetVol, tetCen;         // nt, 3*nt
  std::vector<int> fV, fQP, fQM, fPT, fMT, fPI, fMI;  // face tables (nftot)
  std::vector<double> fArea, fCen, fRad;
  std::vector<int> rowptr, cols;              // touching pairs, a <= b
  // Taylor-Duffy tet pairs (ncv >= 2) used by touching BF pairs
  std::vector<int> tdA, tdB, tdNcv;           // canonical tA <= tB
  std::vector<int> tdOV;                      // [ntd][8] ordered vertices (A then B)
  std::vector<int> pairTD;                    // [nnz][4]
  // device state
  int device = -1;
  bool uploaded = false, static_built = false;
  int td_order = 0;
  bfa::DevBuffers d;
  bfa::DevGeom dg{};
};

// bfa_noalloc.hip -- an LU solve that allocates no device memory.
//
// bfa_lu_solve_z allocates its scratch (row permutation, timeout flag,
// inverted 64 x 64 diagonal blocks of L and U, the permuted right-hand
// sides, the forward and backward solutions and the streaming solves'
// tickets) with hipMalloc on every call.  Callers that cannot allocate --
// code under HIP graph capture, tight serving loops, harnesses that forbid
// it -- use bfa_lu_solve_z_noalloc instead: the same kernels, with the
// scratch in one statically sized device array per GPU, sized for
// n <= BFA_NOALLOC_MAX_N and nrhs <= BFA_NOALLOC_MAX_NRHS (47.6 MiB, held by
// every device the library runs on until the process ends).
//
// Every call on a device uses that device's one array, so calls that
// overlap on one device (different host threads or streams without an
// ordering between them) race on it.  The array keeps the last call's
// contents; the solve's kernels clear what must start at zero (the timeout
// flag, the tickets and the tag-filled solution vectors) before use.
#include <hip/hip_runtime.h>

#include "bfa_internal.h"
#include "bfa_lu.h"

namespace {

constexpr size_t al256(size_t x) { return (x + 255) / 256 * 256; }

// the layout of lu_solve's scratch for the streaming form (nrhs < 64):
// [perm n ints][flag 256 B][2 x nblk inverted 64 x 64 blocks][P b, Y, X][tickets]
constexpr size_t noalloc_bytes(size_t n, size_t nrhs)
{
  return al256(n * 4) + 256 + al256(2 * ((n + 63) / 64) * 64 * 64 * 16) + 3 * al256(n * nrhs * 16) +
         al256(2 * ((nrhs + 7) / 8) * 4);
}

constexpr size_t SCRATCH_BYTES = noalloc_bytes(BFA_NOALLOC_MAX_N, BFA_NOALLOC_MAX_NRHS);

// the scratch array; host code never uses its name as an address
__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// device addresses of `scratch`, one per device, from hipGetSymbolAddress
void *g_scratch[64] = {};

int scratch_ptr(void **out)
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return bfa::fail(BFA_ERR_HIP, "bfa_lu_solve_z_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= 64) return bfa::fail(BFA_ERR_ARG, "bfa_lu_solve_z_noalloc: device %d unsupported", dev);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return bfa::fail(BFA_ERR_HIP, "bfa_lu_solve_z_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return BFA_OK;
}

}  // namespace

extern "C" int bfa_lu_solve_z_noalloc(const void *d_A, int n, int64_t lda, const int *d_ipiv, void *d_B, int nrhs,
                                      int64_t ldb, void *stream)
{
  if (!d_A || !d_ipiv || !d_B || n <= 0 || nrhs <= 0 || lda < n || ldb < n)
    return bfa::fail(BFA_ERR_ARG, "bfa_lu_solve_z_noalloc: bad argument");
  if (n > BFA_NOALLOC_MAX_N || nrhs > BFA_NOALLOC_MAX_NRHS)
    return bfa::fail(BFA_ERR_ARG, "bfa_lu_solve_z_noalloc: n = %d, nrhs = %d exceed the static scratch "
                     "(n <= %d, nrhs <= %d); use bfa_lu_solve_z", n, nrhs, BFA_NOALLOC_MAX_N, BFA_NOALLOC_MAX_NRHS);
  static_assert(BFA_NOALLOC_MAX_NRHS < 64, "the streaming solve form serves fewer than 64 right-hand sides");
  if (bfa::lu_solve_scratch_bytes(n, nrhs, 16) > SCRATCH_BYTES)
    return bfa::fail(BFA_ERR_ARG, "bfa_lu_solve_z_noalloc: scratch layout larger than the static array");
  void *s = nullptr;
  int rc = scratch_ptr(&s);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  rc = bfa::lu_solve_z(d_A, n, lda, d_ipiv, d_B, nrhs, ldb, st, s);
  if (rc == BFA_OK) rc = bfa::lu_solve_check(s, n, st);
  return rc;
}

// bfa_fibbi.cpp -- FIBBI cache files: the reference's on-disk format for the
// frequency-independent static records of touching BF pairs.
//
// Format (src/libs/libbuff/FIBBICache.cc:247-463, non-portable w.r.t.
// endianness exactly like the reference): the 13 bytes "FIBBI_GCACHE\0",
// then records of a 27-float search key followed by the 6-double record
// (FIBBIDATALEN, FIBBICache.h).  The key of the pair (A, B) is the nine
// vectors QM_A, V1_A, V2_A, V3_A, QP_B, QM_B, V1_B, V2_B, V3_B minus QP_A,
// each component rounded to float after the double subtraction
// (GetFIBBICacheKey, FIBBICache.cc:171-203): translation-independent, not
// rotation-independent.  Keys compare bytewise (KeyCmp, memcmp); a key that
// occurs twice keeps its first record (unordered_map::insert).
//
// The record is what this build's static table holds per touching pair
// a <= b (the reference's AssembleGBlock calls GetGMatrixElement(nfa, nfb)
// with nfa <= nfb, VIEMatrix.cc:404-414), so a cache written here can be
// preloaded by the reference and vice versa.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <unordered_map>
#include <vector>

#include "bfa_internal.h"

using namespace bfa;

namespace {

const char kSignature[] = "FIBBI_GCACHE";     // written with its NUL: 13 bytes
constexpr size_t kSigSize = sizeof(kSignature);
constexpr int kKeyLen = 27, kDataLen = 6;
constexpr size_t kKeySize = kKeyLen * sizeof(float), kDataSize = kDataLen * sizeof(double);

void pair_key(const bfa_geometry *g, int a, int b, float *k)
{
  const double *V = g->V.data();
  const double *qpa = V + 3 * (size_t)g->fQP[a];
  const int va[4] = {g->fQM[a], g->fV[3 * a], g->fV[3 * a + 1], g->fV[3 * a + 2]};
  const int vb[5] = {g->fQP[b], g->fQM[b], g->fV[3 * b], g->fV[3 * b + 1], g->fV[3 * b + 2]};
  int o = 0;
  for (int i = 0; i < 4; i++, o += 3)
    for (int d = 0; d < 3; d++) k[o + d] = (float)(V[3 * (size_t)va[i] + d] - qpa[d]);
  for (int i = 0; i < 5; i++, o += 3)
    for (int d = 0; d < 3; d++) k[o + d] = (float)(V[3 * (size_t)vb[i] + d] - qpa[d]);
}

std::vector<float> all_keys(const bfa_geometry *g)
{
  std::vector<float> keys((size_t)g->cols.size() * kKeyLen);
  for (int a = 0; a < g->nbf; a++)
    for (int j = g->rowptr[a]; j < g->rowptr[a + 1]; j++) pair_key(g, a, g->cols[j], &keys[(size_t)j * kKeyLen]);
  return keys;
}

std::string key_bytes(const float *k) { return std::string((const char *)k, kKeySize); }

// whole file, validated as FIBBICache::PreLoad does (signature, size)
int read_file(const char *path, std::vector<float> &keys, std::vector<double> &data)
{
  FILE *f = fopen(path, "rb");
  if (!f) return fail(BFA_ERR_IO, "FIBBI cache %s: cannot open", path);
  struct stat st;
  if (fstat(fileno(f), &st) != 0 || st.st_size < (off_t)kSigSize) {
    fclose(f);
    return fail(BFA_ERR_IO, "FIBBI cache %s: invalid cache file", path);
  }
  char sig[kSigSize];
  if (fread(sig, kSigSize, 1, f) != 1 || memcmp(sig, kSignature, kSigSize) != 0) {
    fclose(f);
    return fail(BFA_ERR_IO, "FIBBI cache %s: invalid cache file", path);
  }
  const size_t body = (size_t)st.st_size - kSigSize, rec = kKeySize + kDataSize;
  if (body % rec != 0) {
    fclose(f);
    return fail(BFA_ERR_IO, "FIBBI cache %s: cache file has incorrect size", path);
  }
  const size_t n = body / rec;
  keys.resize(n * kKeyLen);
  data.resize(n * kDataLen);
  for (size_t i = 0; i < n; i++) {
    if (fread(&keys[i * kKeyLen], kKeySize, 1, f) != 1 || fread(&data[i * kDataLen], kDataSize, 1, f) != 1) {
      fclose(f);
      return fail(BFA_ERR_IO, "FIBBI cache %s: read only %zu/%zu records", path, i, n);
    }
  }
  fclose(f);
  return BFA_OK;
}

int write_file(const char *path, size_t n, const float *keys, const double *data, const std::vector<char> *keep,
               int64_t *written)
{
  FILE *f = fopen(path, "wb");
  if (!f) return fail(BFA_ERR_IO, "FIBBI cache %s: cannot open for writing", path);
  int64_t w = 0;
  bool ok = fwrite(kSignature, kSigSize, 1, f) == 1;
  for (size_t i = 0; ok && i < n; i++) {
    if (keep && !(*keep)[i]) continue;
    ok = fwrite(keys + i * kKeyLen, kKeySize, 1, f) == 1 && fwrite(data + i * kDataLen, kDataSize, 1, f) == 1;
    w += ok;
  }
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(BFA_ERR_IO, "FIBBI cache %s: write failed", path);
  if (written) *written = w;
  return BFA_OK;
}

}  // namespace

extern "C" int bfa_fibbi_keys(const bfa_geometry *g, float *keys)
{
  if (!g || !keys) return fail(BFA_ERR_ARG, "bfa_fibbi_keys: null argument");
  const std::vector<float> k = all_keys(g);
  memcpy(keys, k.data(), k.size() * sizeof(float));
  return BFA_OK;
}

extern "C" int bfa_fibbi_read(const char *path, int64_t *nrec, float *keys, double *data)
{
  if (!path || !nrec) return fail(BFA_ERR_ARG, "bfa_fibbi_read: null argument");
  std::vector<float> k;
  std::vector<double> d;
  int rc = read_file(path, k, d);
  if (rc) return rc;
  const int64_t n = (int64_t)(k.size() / kKeyLen);
  if ((keys || data) && *nrec < n) return fail(BFA_ERR_ARG, "bfa_fibbi_read: room for %lld of %lld records",
                                               (long long)*nrec, (long long)n);
  *nrec = n;
  if (keys) memcpy(keys, k.data(), k.size() * sizeof(float));
  if (data) memcpy(data, d.data(), d.size() * sizeof(double));
  return BFA_OK;
}

extern "C" int bfa_fibbi_write(const char *path, int64_t nrec, const float *keys, const double *data)
{
  if (!path || nrec < 0 || (nrec > 0 && (!keys || !data))) return fail(BFA_ERR_ARG, "bfa_fibbi_write: bad argument");
  return write_file(path, (size_t)nrec, keys, data, nullptr, nullptr);
}

// FIBBICache::Store: every distinct key of the geometry's touching pairs
// once, with the first pair's record.
extern "C" int bfa_fibbi_store(const bfa_geometry *g, const char *path, int64_t *written)
{
  if (!g || !path) return fail(BFA_ERR_ARG, "bfa_fibbi_store: null argument");
  if (!g->static_built) return fail(BFA_ERR_STATE, "bfa_fibbi_store: static table not built");
  const size_t n = g->cols.size();
  std::vector<double> table(n * kDataLen);
  int rc = bfa_static_download(g, table.data());
  if (rc) return rc;
  const std::vector<float> keys = all_keys(g);
  std::vector<char> keep(n, 0);
  std::unordered_map<std::string, size_t> seen;
  seen.reserve(n * 2);
  for (size_t j = 0; j < n; j++) keep[j] = seen.emplace(key_bytes(&keys[j * kKeyLen]), j).second;
  return write_file(path, n, keys.data(), table.data(), &keep, written);
}

// FIBBICache::PreLoad + GetFIBBIData over every touching pair: records whose
// key is in the file are taken from it (hits); if any pair misses, the
// static table is built on the device (td_order) and only the misses keep
// the computed record.  With no misses the cold build is skipped.
extern "C" int bfa_fibbi_preload(bfa_geometry *g, const char *path, int td_order, int64_t *hits, int64_t *misses,
                                 void *stream)
{
  if (!g || !path) return fail(BFA_ERR_ARG, "bfa_fibbi_preload: null argument");
  if (!g->uploaded) return fail(BFA_ERR_STATE, "bfa_fibbi_preload: geometry not uploaded");
  std::vector<float> fk;
  std::vector<double> fd;
  int rc = read_file(path, fk, fd);
  if (rc) return rc;
  const size_t nf = fk.size() / kKeyLen;
  std::unordered_map<std::string, size_t> map;
  map.reserve(nf * 2);
  for (size_t i = 0; i < nf; i++) map.emplace(key_bytes(&fk[i * kKeyLen]), i);
  const size_t n = g->cols.size();
  const std::vector<float> keys = all_keys(g);
  std::vector<double> table(n * kDataLen, 0.0);
  std::vector<char> hit(n, 0);
  int64_t nh = 0;
  for (size_t j = 0; j < n; j++) {
    auto it = map.find(key_bytes(&keys[j * kKeyLen]));
    if (it == map.end()) continue;
    memcpy(&table[j * kDataLen], &fd[it->second * kDataLen], kDataSize);
    hit[j] = 1;
    nh++;
  }
  if (nh < (int64_t)n) {
    if ((rc = bfa_static_build(g, td_order, stream))) return rc;
    std::vector<double> computed(n * kDataLen);
    if ((rc = bfa_static_download(g, computed.data()))) return rc;
    for (size_t j = 0; j < n; j++)
      if (!hit[j]) memcpy(&table[j * kDataLen], &computed[j * kDataLen], kDataSize);
  }
  if ((rc = bfa_static_upload(g, table.data()))) return rc;
  if (hits) *hits = nh;
  if (misses) *misses = (int64_t)n - nh;
  return BFA_OK;
}

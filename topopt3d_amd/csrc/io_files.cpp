// On-disk formats of the reference, written and read on the host.
//
// PETSc binary (PetscViewerBinaryOpen + VecView/MatView, TopOpt.cpp:312-347, 442-508):
// big-endian; a Vec is  classid 1211214, n, then n doubles; a sequential AIJ Mat is
// classid 1211216, rows, cols, nnz, the row lengths, the column indices, then the values.
// Integers are 32-bit, or 64-bit for a PETSc built with 64-bit indices (index_bytes = 8).
//
// Raw mesh files (TopOpt::MeshOut, TopOpt.cpp:97-360; read back by TopOpt::LoadMesh,
// RecMesh.cpp:34-412): native-endian PetscInt / double / bool arrays, row-major.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_setup.hpp"
#include "tofe.h"

namespace {

constexpr int64_t kVecClassId = 1211214;
constexpr int64_t kMatClassId = 1211216;

struct File {
  FILE *f = nullptr;
  File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
  ~File()
  {
    if (f) std::fclose(f);
  }
};

bool put_int(FILE *f, int64_t v, int bytes)
{
  if (bytes == 4) {
    const uint32_t b = __builtin_bswap32((uint32_t)(int32_t)v);
    return std::fwrite(&b, 4, 1, f) == 1;
  }
  const uint64_t b = __builtin_bswap64((uint64_t)v);
  return std::fwrite(&b, 8, 1, f) == 1;
}

bool get_int(FILE *f, int64_t &v, int bytes)
{
  if (bytes == 4) {
    uint32_t b;
    if (std::fread(&b, 4, 1, f) != 1) return false;
    v = (int32_t)__builtin_bswap32(b);
    return true;
  }
  uint64_t b;
  if (std::fread(&b, 8, 1, f) != 1) return false;
  v = (int64_t)__builtin_bswap64(b);
  return true;
}

bool put_doubles(FILE *f, const double *v, int64_t n)
{
  std::vector<uint64_t> buf(4096);
  for (int64_t i = 0; i < n; i += 4096) {
    const int64_t m = std::min<int64_t>(4096, n - i);
    for (int64_t j = 0; j < m; ++j) {
      uint64_t u;
      std::memcpy(&u, v + i + j, 8);
      buf[j] = __builtin_bswap64(u);
    }
    if (std::fwrite(buf.data(), 8, m, f) != (size_t)m) return false;
  }
  return true;
}

bool get_doubles(FILE *f, double *v, int64_t n)
{
  std::vector<uint64_t> buf(4096);
  for (int64_t i = 0; i < n; i += 4096) {
    const int64_t m = std::min<int64_t>(4096, n - i);
    if (std::fread(buf.data(), 8, m, f) != (size_t)m) return false;
    for (int64_t j = 0; j < m; ++j) {
      const uint64_t u = __builtin_bswap64(buf[j]);
      std::memcpy(v + i + j, &u, 8);
    }
  }
  return true;
}

bool bad_bytes(int b) { return b != 4 && b != 8; }

template <typename T>
int write_raw(const std::string &path, const T *v, size_t n)
{
  File f(path.c_str(), "wb");
  if (!f.f) return TOFE_ERR_ARG;
  if (n && std::fwrite(v, sizeof(T), n, f.f) != n) return TOFE_ERR_HIP;
  return TOFE_OK;
}

// Raw native-endian PetscInt array of 4 or 8 bytes per entry.
int write_raw_int(const std::string &path, const std::vector<int64_t> &v, int bytes)
{
  if (bytes == 8) return write_raw(path, v.data(), v.size());
  std::vector<int32_t> w(v.begin(), v.end());
  return write_raw(path, w.data(), w.size());
}

}  // namespace

extern "C" {

int tofe_petsc_vec_write(const char *path, const double *v, int64_t n, int index_bytes)
{
  if (!path || (n > 0 && !v) || n < 0 || bad_bytes(index_bytes)) return TOFE_ERR_ARG;
  File f(path, "wb");
  if (!f.f) return TOFE_ERR_ARG;
  if (!put_int(f.f, kVecClassId, index_bytes) || !put_int(f.f, n, index_bytes) ||
      !put_doubles(f.f, v, n))
    return TOFE_ERR_HIP;
  return TOFE_OK;
}

int tofe_petsc_vec_read(const char *path, double *v, int64_t *n, int index_bytes)
{
  if (!path || !n || bad_bytes(index_bytes)) return TOFE_ERR_ARG;
  File f(path, "rb");
  if (!f.f) return TOFE_ERR_ARG;
  int64_t cid, len;
  if (!get_int(f.f, cid, index_bytes) || cid != kVecClassId || !get_int(f.f, len, index_bytes) ||
      len < 0)
    return TOFE_ERR_ARG;
  if (v) {
    if (*n < len) return TOFE_ERR_ARG;
    if (!get_doubles(f.f, v, len)) return TOFE_ERR_ARG;
  }
  *n = len;
  return TOFE_OK;
}

int tofe_petsc_aij_write(const char *path, int64_t m, int64_t ncols, const int64_t *rowptr,
                         const int32_t *col, const double *val, int index_bytes)
{
  if (!path || m < 0 || ncols < 0 || !rowptr || bad_bytes(index_bytes)) return TOFE_ERR_ARG;
  const int64_t nnz = rowptr[m] - rowptr[0];
  if (nnz > 0 && (!col || !val)) return TOFE_ERR_ARG;
  File f(path, "wb");
  if (!f.f) return TOFE_ERR_ARG;
  bool ok = put_int(f.f, kMatClassId, index_bytes) && put_int(f.f, m, index_bytes) &&
            put_int(f.f, ncols, index_bytes) && put_int(f.f, nnz, index_bytes);
  for (int64_t i = 0; ok && i < m; ++i) ok = put_int(f.f, rowptr[i + 1] - rowptr[i], index_bytes);
  for (int64_t p = rowptr[0]; ok && p < rowptr[m]; ++p) ok = put_int(f.f, col[p], index_bytes);
  ok = ok && put_doubles(f.f, val + rowptr[0], nnz);
  return ok ? TOFE_OK : TOFE_ERR_HIP;
}

int tofe_petsc_aij_read(const char *path, int64_t *m, int64_t *ncols, int64_t *nnz,
                        int64_t *rowptr, int32_t *col, double *val, int index_bytes)
{
  if (!path || !m || !ncols || !nnz || bad_bytes(index_bytes)) return TOFE_ERR_ARG;
  File f(path, "rb");
  if (!f.f) return TOFE_ERR_ARG;
  int64_t cid, M, N, Z;
  if (!get_int(f.f, cid, index_bytes) || cid != kMatClassId || !get_int(f.f, M, index_bytes) ||
      !get_int(f.f, N, index_bytes) || !get_int(f.f, Z, index_bytes) || M < 0 || N < 0 || Z < 0)
    return TOFE_ERR_ARG;
  if (rowptr) {
    if (*m < M || *nnz < Z) return TOFE_ERR_ARG;
    rowptr[0] = 0;
    for (int64_t i = 0; i < M; ++i) {
      int64_t len;
      if (!get_int(f.f, len, index_bytes) || len < 0) return TOFE_ERR_ARG;
      rowptr[i + 1] = rowptr[i] + len;
    }
    if (rowptr[M] != Z) return TOFE_ERR_ARG;
    if (col) {
      for (int64_t p = 0; p < Z; ++p) {
        int64_t c;
        if (!get_int(f.f, c, index_bytes) || c < 0 || c >= N) return TOFE_ERR_ARG;
        col[p] = (int32_t)c;
      }
      if (val && !get_doubles(f.f, val, Z)) return TOFE_ERR_ARG;
    }
  }
  *m = M;
  *ncols = N;
  *nnz = Z;
  return TOFE_OK;
}

}  // extern "C"

namespace tofe {

int write_mesh_files(const std::string &dir, int nx, int ny, int nz, double hx, double hy,
                     double hz, const double origin[3], int index_bytes)
{
  if (bad_bytes(index_bytes)) return TOFE_ERR_ARG;
  const int64_t ne = (int64_t)nx * ny * nz, Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  const int64_t nn = Nx * Ny * Nz;
  int rc;
  // one rank: the distributions are [0, nElem] and [0, nNode]
  if ((rc = write_raw_int(dir + "/Element_Distribution.bin", {0, ne}, index_bytes))) return rc;
  if ((rc = write_raw_int(dir + "/Node_Distribution.bin", {0, nn}, index_bytes))) return rc;
  std::vector<int64_t> el((size_t)ne * 8);
  static const int LN[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                               {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int64_t e = 0; e < ne; ++e) {
    const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / ((int64_t)nx * ny);
    for (int l = 0; l < 8; ++l)
      el[8 * e + l] = (ex + LN[l][0]) + Nx * ((ey + LN[l][1]) + Ny * (ez + LN[l][2]));
  }
  if ((rc = write_raw_int(dir + "/elements.bin", el, index_bytes))) return rc;
  std::vector<double> xyz((size_t)nn * 3);
  for (int64_t n = 0; n < nn; ++n) {
    xyz[3 * n] = origin[0] + hx * (double)(n % Nx);
    xyz[3 * n + 1] = origin[1] + hy * (double)((n / Nx) % Ny);
    xyz[3 * n + 2] = origin[2] + hz * (double)(n / (Nx * Ny));
  }
  if ((rc = write_raw(dir + "/nodes.bin", xyz.data(), xyz.size()))) return rc;
  std::vector<uint8_t> active((size_t)ne, 1);
  static_assert(sizeof(bool) == 1, "bool is one byte");
  return write_raw(dir + "/active.bin", active.data(), active.size());
}

}  // namespace tofe

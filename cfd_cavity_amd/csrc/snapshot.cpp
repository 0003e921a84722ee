// snapshot.cpp — the CAVSNAP1 checkpoint format and the bilinear regrid used
// for restarts (bbrangeo/cfd_cavity src/uniform/snapshot.rs:10-131).
//
// Layout: 8-byte magic "CAVSNAP1", nx and ny as u64 little-endian, then rho,
// ux, uy, T as f64 little-endian, each nx*ny values in k = iy*nx + ix order.
// Host-side file I/O; nothing here touches the GPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cavlbm.h"

namespace {

thread_local std::string g_io_error;

const char kMagic[8] = {'C', 'A', 'V', 'S', 'N', 'A', 'P', '1'};

int io_fail(const std::string& m) {
  g_io_error = m;
  return CAVLBM_E_IO;
}

void put_u64(std::vector<unsigned char>& b, uint64_t v) {
  for (int i = 0; i < 8; ++i) b.push_back((unsigned char)(v >> (8 * i)));
}

uint64_t get_u64(const unsigned char* p) {
  uint64_t v = 0;
  for (int i = 7; i >= 0; --i) v = (v << 8) | p[i];
  return v;
}

void put_f64(std::vector<unsigned char>& b, double x) {
  uint64_t v;
  std::memcpy(&v, &x, 8);
  put_u64(b, v);
}

double get_f64(const unsigned char* p) {
  const uint64_t v = get_u64(p);
  double x;
  std::memcpy(&x, &v, 8);
  return x;
}

// Nodes of a grid read from a header and the bytes of its four f64 fields,
// with `room` bytes left in a uint64 size; false for an empty grid or when
// nx*ny or 32*nx*ny does not fit (a corrupt header).
bool grid_payload(uint64_t nx, uint64_t ny, uint64_t room, uint64_t* n, uint64_t* bytes) {
  if (nx == 0 || ny == 0) return false;
  const uint64_t nn = nx * ny;
  if (nn / nx != ny || nn > room / 32) return false;
  *n = nn;
  *bytes = 32 * nn;
  return true;
}

}  // namespace

extern "C" {

// write_snapshot (snapshot.rs:14-31)
int cavlbm_snapshot_write(const char* path, uint64_t nx, uint64_t ny, const double* rho,
                          const double* ux, const double* uy, const double* t) {
  if (!path || !rho || !ux || !uy || !t) return io_fail("null argument");
  const uint64_t n = nx * ny;
  std::vector<unsigned char> b;
  b.reserve(24 + 32 * n);
  b.insert(b.end(), kMagic, kMagic + 8);
  put_u64(b, nx);
  put_u64(b, ny);
  for (const double* a : {rho, ux, uy, t})
    for (uint64_t k = 0; k < n; ++k) put_f64(b, a[k]);
  FILE* f = std::fopen(path, "wb");
  if (!f) return io_fail(std::string("cannot create ") + path);
  const size_t w = std::fwrite(b.data(), 1, b.size(), f);
  const int c = std::fclose(f);
  if (w != b.size() || c != 0) return io_fail(std::string("write failed: ") + path);
  return 0;
}

// write_snapshot_atomic (snapshot.rs:37-58): temporary file in the same
// directory, then rename over the destination.
int cavlbm_snapshot_write_atomic(const char* path, uint64_t nx, uint64_t ny, const double* rho,
                                 const double* ux, const double* uy, const double* t) {
  if (!path) return io_fail("null path");
  const std::string tmp = std::string(path) + ".tmp";
  int rc = cavlbm_snapshot_write(tmp.c_str(), nx, ny, rho, ux, uy, t);
  if (rc) {
    std::remove(tmp.c_str());
    return rc;
  }
  if (std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return io_fail(std::string("rename failed: ") + path);
  }
  return 0;
}

// read_snapshot (snapshot.rs:61-95).  With null field pointers only nx, ny
// are returned; otherwise each buffer must hold `capacity` >= nx*ny values.
int cavlbm_snapshot_read(const char* path, uint64_t* nx, uint64_t* ny, double* rho, double* ux,
                         double* uy, double* t, uint64_t capacity) {
  if (!path || !nx || !ny) return io_fail("null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return io_fail(std::string("cannot open ") + path);
  std::vector<unsigned char> data;
  unsigned char buf[1 << 16];
  size_t r;
  while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) data.insert(data.end(), buf, buf + r);
  std::fclose(f);
  if (data.size() < 24 || std::memcmp(data.data(), kMagic, 8) != 0)
    return io_fail("snapshot: bad magic or file too short");
  const uint64_t sx = get_u64(data.data() + 8), sy = get_u64(data.data() + 16);
  uint64_t n = 0, payload = 0;
  if (!grid_payload(sx, sy, UINT64_MAX - 24, &n, &payload))
    return io_fail("snapshot: invalid grid " + std::to_string(sx) + "x" + std::to_string(sy));
  const uint64_t expect = 24 + payload;
  if (data.size() < expect)
    return io_fail("snapshot: expected " + std::to_string(expect) + " bytes, read " +
                   std::to_string(data.size()));
  *nx = sx;
  *ny = sy;
  if (!rho && !ux && !uy && !t) return 0;
  if (!rho || !ux || !uy || !t || capacity < n) return io_fail("snapshot: output buffers too small");
  const unsigned char* p = data.data() + 24;
  for (double* a : {rho, ux, uy, t})
    for (uint64_t k = 0; k < n; ++k, p += 8) a[k] = get_f64(p);
  return 0;
}

// interpolate_field (snapshot.rs:100-131): bilinear in normalised
// coordinates x, y in [0, 1]; identity when the grids are equal.
int cavlbm_interpolate_field(const double* src, uint64_t nx0, uint64_t ny0, double* dst,
                             uint64_t nx1, uint64_t ny1) {
  if (!src || !dst || nx0 == 0 || ny0 == 0) return io_fail("null or empty field");
  if (nx0 == nx1 && ny0 == ny1) {
    std::memcpy(dst, src, nx0 * ny0 * sizeof(double));
    return 0;
  }
  for (uint64_t iy = 0; iy < ny1; ++iy) {
    const double vy = ny1 > 1 ? (double)iy / (double)(ny1 - 1) : 0.0;
    const double sy = vy * (double)(ny0 - 1);
    const uint64_t y0 = (uint64_t)std::floor(sy);
    const uint64_t y1 = std::min<uint64_t>(y0 + 1, ny0 - 1);
    const double wy = sy - (double)y0;
    for (uint64_t ix = 0; ix < nx1; ++ix) {
      const double vx = nx1 > 1 ? (double)ix / (double)(nx1 - 1) : 0.0;
      const double sx = vx * (double)(nx0 - 1);
      const uint64_t x0 = (uint64_t)std::floor(sx);
      const uint64_t x1 = std::min<uint64_t>(x0 + 1, nx0 - 1);
      const double wx = sx - (double)x0;
      const double v00 = src[y0 * nx0 + x0], v10 = src[y0 * nx0 + x1];
      const double v01 = src[y1 * nx0 + x0], v11 = src[y1 * nx0 + x1];
      dst[iy * nx1 + ix] = v00 * (1.0 - wx) * (1.0 - wy) + v10 * wx * (1.0 - wy) +
                           v01 * (1.0 - wx) * wy + v11 * wx * wy;
    }
  }
  return 0;
}

// ---- OCTSNAP1: multi-level snapshots of the octree solver
// (src/octree/snapshot.rs:1-170).  Header: magic "OCTSNAP1", version u32 = 1,
// levels u32, step u64, n_grids u32, then nx, ny (u64) of every grid; then
// for every grid in the same order its rho, ux, uy, T (f64), all
// little-endian.  The field arrays here are the grids' fields concatenated in
// grid order.

namespace {

const char kOctMagic[8] = {'O', 'C', 'T', 'S', 'N', 'A', 'P', '1'};

void put_u32(std::vector<unsigned char>& b, uint32_t v) {
  for (int i = 0; i < 4; ++i) b.push_back((unsigned char)(v >> (8 * i)));
}

uint32_t get_u32(const unsigned char* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

int write_file(const std::string& path, const std::vector<unsigned char>& b) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return io_fail("cannot create " + path);
  const size_t w = std::fwrite(b.data(), 1, b.size(), f);
  const int fl = std::fflush(f);
  const int c = std::fclose(f);
  if (w != b.size() || fl != 0 || c != 0) return io_fail("write failed: " + path);
  return 0;
}

}  // namespace

// encode (snapshot.rs:57-79) + write_snapshot_atomic (snapshot.rs:140-161)
int cavlbm_octsnap_write(const char* path, uint32_t levels, uint64_t step, uint32_t n_grids,
                         const uint64_t* dims, const double* rho, const double* ux,
                         const double* uy, const double* t, int32_t atomic) {
  if (!path || !dims || !rho || !ux || !uy || !t) return io_fail("null argument");
  std::vector<unsigned char> b;
  b.insert(b.end(), kOctMagic, kOctMagic + 8);
  put_u32(b, 1);
  put_u32(b, levels);
  put_u64(b, step);
  put_u32(b, n_grids);
  for (uint32_t g = 0; g < n_grids; ++g) {
    put_u64(b, dims[2 * g]);
    put_u64(b, dims[2 * g + 1]);
  }
  uint64_t off = 0;
  for (uint32_t g = 0; g < n_grids; ++g) {
    const uint64_t n = dims[2 * g] * dims[2 * g + 1];
    for (const double* a : {rho, ux, uy, t})
      for (uint64_t k = 0; k < n; ++k) put_f64(b, a[off + k]);
    off += n;
  }
  if (!atomic) return write_file(path, b);
  const std::string tmp = std::string(path) + ".tmp";
  int rc = write_file(tmp, b);
  if (rc) {
    std::remove(tmp.c_str());
    return rc;
  }
  if (std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return io_fail(std::string("rename failed: ") + path);
  }
  return 0;
}

// read_snapshot + decode (snapshot.rs:83-137, :164-170): the exact length is
// checked.  With null field pointers only the header (levels, step,
// n_grids, dims: room for 2*5 values) is returned; otherwise each field
// buffer must hold `capacity` >= the sum of nx*ny values.
int cavlbm_octsnap_read(const char* path, uint32_t* levels, uint64_t* step, uint32_t* n_grids,
                        uint64_t* dims, double* rho, double* ux, double* uy, double* t,
                        uint64_t capacity) {
  if (!path || !levels || !step || !n_grids || !dims) return io_fail("null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return io_fail(std::string("cannot open ") + path);
  std::vector<unsigned char> data;
  unsigned char buf[1 << 16];
  size_t r;
  while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) data.insert(data.end(), buf, buf + r);
  std::fclose(f);
  if (data.size() < 28 || std::memcmp(data.data(), kOctMagic, 8) != 0)
    return io_fail("snapshot octree : magic invalide ou fichier trop court");
  const uint32_t version = get_u32(data.data() + 8);
  if (version != 1)
    return io_fail("snapshot octree : version " + std::to_string(version) +
                   " non supportée (attendu 1)");
  const uint32_t lv = get_u32(data.data() + 12);
  const uint64_t st = get_u64(data.data() + 16);
  const uint32_t ng = get_u32(data.data() + 24);
  if (!(lv == 2 || lv == 3) || ng != (lv == 2 ? 3u : 5u))
    return io_fail("snapshot octree : incohérence levels=" + std::to_string(lv) +
                   " / n_grids=" + std::to_string(ng) + " (attendu 3/5)");
  const size_t dims_end = 28 + (size_t)ng * 16;
  if (data.size() < dims_end) return io_fail("snapshot octree : header dims tronqué");
  uint64_t expect = dims_end, total = 0;
  uint64_t d[10];
  for (uint32_t g = 0; g < ng; ++g) {
    d[2 * g] = get_u64(data.data() + 28 + 16 * g);
    d[2 * g + 1] = get_u64(data.data() + 36 + 16 * g);
    uint64_t n = 0, payload = 0;
    if (!grid_payload(d[2 * g], d[2 * g + 1], UINT64_MAX - expect, &n, &payload))
      return io_fail("snapshot octree : grille " + std::to_string(g) + " invalide " +
                     std::to_string(d[2 * g]) + "x" + std::to_string(d[2 * g + 1]));
    expect += payload;
    total += n;
  }
  if (data.size() != expect)
    return io_fail("snapshot octree : attendu " + std::to_string(expect) + " octets, lu " +
                   std::to_string(data.size()));
  *levels = lv;
  *step = st;
  *n_grids = ng;
  std::memcpy(dims, d, sizeof(uint64_t) * 2 * ng);
  if (!rho && !ux && !uy && !t) return 0;
  if (!rho || !ux || !uy || !t || capacity < total)
    return io_fail("snapshot octree : output buffers too small");
  const unsigned char* p = data.data() + dims_end;
  uint64_t off = 0;
  for (uint32_t g = 0; g < ng; ++g) {
    const uint64_t n = d[2 * g] * d[2 * g + 1];
    for (double* a : {rho, ux, uy, t})
      for (uint64_t k = 0; k < n; ++k, p += 8) a[off + k] = get_f64(p);
    off += n;
  }
  return 0;
}

const char* cavlbm_snapshot_last_error(void) { return g_io_error.c_str(); }

}  // extern "C"

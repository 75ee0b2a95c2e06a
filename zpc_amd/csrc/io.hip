// Particle layout and IO adapters: zpc's TileVector particle layout to and
// from the per-attribute arrays the substep runs on, TileVector tile
// reordering, binary checkpoints (particles + HashTable) and Houdini bgeo
// output of positions with one optional attribute.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/zsmpm.h"
#include "zsmpm_internal.hpp"

namespace zsmpm {

namespace {

constexpr int kIoThreads = 256;

int groups_for_n(int64_t n) {
  const int64_t g = (n + kIoThreads - 1) / kIoThreads;
  return (int)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

// TileVector<f32, L> element (chn, i) at (i / L * nch + chn) * L + i % L
// (container/TileVector.hpp:108,397,911-913). A workgroup takes 256
// consecutive elements and stages each attribute in LDS: on the tile side
// thread t moves element b0 + t (a wave's 64 lanes read or write 64 / L whole
// L-float channel stripes per channel), on the array side the workgroup's
// rows are one contiguous span of 256 * width floats, moved as 16-byte
// vectors when aligned.
template <bool kToTiles>
__global__ __launch_bounds__(kIoThreads) void tile_convert_kernel(TileAttrs a, float* __restrict__ tiles) {
  __shared__ float stage[kIoThreads * 9];
  const int shift = a.lane_shift;
  const int64_t mask = ((int64_t)1 << shift) - 1;
  const int t = threadIdx.x;
  for (int64_t b0 = (int64_t)blockIdx.x * kIoThreads; b0 < a.n; b0 += (int64_t)gridDim.x * kIoThreads) {
    const int cnt = (int)(a.n - b0 < kIoThreads ? a.n - b0 : kIoThreads);
    const int64_t i = b0 + t;
    const int64_t base = ((i >> shift) * a.nch << shift) + (i & mask);
    for (int k = 0; k < a.count; ++k) {
      const TileAttr at = a.attr[k];
      const int w = at.width;
      float* rows = at.soa + b0 * w;  // the workgroup's span of rows
      const int span = cnt * w;
      const bool vec = ((uintptr_t)rows % 16) == 0;
      if (kToTiles) {
        if (vec) {
          for (int q = t; q < span / 4; q += kIoThreads)
            *reinterpret_cast<zs_f4*>(stage + 4 * q) = *reinterpret_cast<const zs_f4*>(rows + 4 * q);
          for (int q = (span / 4) * 4 + t; q < span; q += kIoThreads) stage[q] = rows[q];
        } else {
          for (int q = t; q < span; q += kIoThreads) stage[q] = rows[q];
        }
        __syncthreads();
        if (t < cnt)
          for (int e = 0; e < w; ++e) {
            // row-major channels of a 3x3 (vec<f32,3,3>::val order) against
            // the column-major rows: channel 3r + c holds row entry 3c + r
            const int col = at.transpose ? (e % 3) * 3 + e / 3 : e;
            tiles[base + ((int64_t)(at.chn + e) << shift)] = stage[t * w + col];
          }
        __syncthreads();
      } else {
        if (t < cnt)
          for (int e = 0; e < w; ++e) {
            const int col = at.transpose ? (e % 3) * 3 + e / 3 : e;
            stage[t * w + col] = tiles[base + ((int64_t)(at.chn + e) << shift)];
          }
        __syncthreads();
        if (vec) {
          for (int q = t; q < span / 4; q += kIoThreads)
            *reinterpret_cast<zs_f4*>(rows + 4 * q) = *reinterpret_cast<const zs_f4*>(stage + 4 * q);
          for (int q = (span / 4) * 4 + t; q < span; q += kIoThreads) rows[q] = stage[q];
        } else {
          for (int q = t; q < span; q += kIoThreads) rows[q] = stage[q];
        }
        __syncthreads();
      }
    }
  }
}

// TileVectorTileReorder (container/TileVector.hpp:644-680): gather moves tile
// map[t] to t, scatter moves tile t to map[t]; whole tiles of L * nch floats.
template <bool kScatter>
__global__ __launch_bounds__(kIoThreads) void tile_reorder_kernel(const zs_f4* __restrict__ src, zs_f4* __restrict__ dst,
                                                                  const int32_t* __restrict__ map, int64_t ntiles,
                                                                  int64_t tile_vec4, uint32_t* err) {
  const int64_t total = ntiles * tile_vec4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += stride) {
    const int64_t t = e / tile_vec4, o = e - t * tile_vec4;
    const int64_t m = map[t];
    if (m < 0 || m >= ntiles) {
      atomicOr(err, kErrCapacity);
      continue;
    }
    if (kScatter)
      __builtin_nontemporal_store(__builtin_nontemporal_load(&src[e]), &dst[m * tile_vec4 + o]);
    else
      __builtin_nontemporal_store(__builtin_nontemporal_load(&src[m * tile_vec4 + o]), &dst[e]);
  }
}

template <bool kScatter>
__global__ __launch_bounds__(kIoThreads) void tile_reorder_scalar_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                                         const int32_t* __restrict__ map, int64_t ntiles,
                                                                         int64_t tile_floats, uint32_t* err) {
  const int64_t total = ntiles * tile_floats;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += stride) {
    const int64_t t = e / tile_floats, o = e - t * tile_floats;
    const int64_t m = map[t];
    if (m < 0 || m >= ntiles) {
      atomicOr(err, kErrCapacity);
      continue;
    }
    if (kScatter)
      dst[m * tile_floats + o] = src[e];
    else
      dst[e] = src[m * tile_floats + o];
  }
}

}  // namespace

hipError_t tiles_convert(const TileAttrs& a, float* tiles, bool to_tiles, hipStream_t s) {
  if (a.n == 0 || a.count == 0) return hipSuccess;
  if (to_tiles)
    tile_convert_kernel<true><<<groups_for_n(a.n), kIoThreads, 0, s>>>(a, tiles);
  else
    tile_convert_kernel<false><<<groups_for_n(a.n), kIoThreads, 0, s>>>(a, tiles);
  return hipGetLastError();
}

hipError_t tiles_reorder(const float* src, float* dst, const int32_t* map, int64_t ntiles, int64_t tile_floats,
                         bool scatter, uint32_t* err, hipStream_t s) {
  if (ntiles == 0) return hipSuccess;
  const bool vec4 = (tile_floats % 4) == 0 && ((uintptr_t)src % 16) == 0 && ((uintptr_t)dst % 16) == 0;
  if (vec4) {
    const int64_t tv = tile_floats / 4;
    const int g = groups_for_n(ntiles * tv);
    if (scatter)
      tile_reorder_kernel<true><<<g, kIoThreads, 0, s>>>((const zs_f4*)src, (zs_f4*)dst, map, ntiles, tv, err);
    else
      tile_reorder_kernel<false><<<g, kIoThreads, 0, s>>>((const zs_f4*)src, (zs_f4*)dst, map, ntiles, tv, err);
  } else {
    const int g = groups_for_n(ntiles * tile_floats);
    if (scatter)
      tile_reorder_scalar_kernel<true><<<g, kIoThreads, 0, s>>>(src, dst, map, ntiles, tile_floats, err);
    else
      tile_reorder_scalar_kernel<false><<<g, kIoThreads, 0, s>>>(src, dst, map, ntiles, tile_floats, err);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host-side file formats
// ---------------------------------------------------------------------------
namespace {

// 64-bit FNV-1a over little-endian 64-bit words (offset basis and prime of
// FNV-1a 64), the tail bytes one by one: a checksum of the checkpoint body.
struct Fnv64 {
  uint64_t h = 0xcbf29ce484222325ull;
  static constexpr uint64_t kPrime = 0x100000001b3ull;
  void add(const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
      uint64_t w;
      memcpy(&w, b + i, 8);
      h = (h ^ w) * kPrime;
    }
    for (; i < n; ++i) h = (h ^ b[i]) * kPrime;
  }
};

// Pinned staging of device arrays in fixed chunks.
struct Staging {
  static constexpr size_t kChunk = size_t(64) << 20;
  unsigned char* buf = nullptr;
  hipError_t init() { return hipHostMalloc((void**)&buf, kChunk, hipHostMallocDefault); }
  ~Staging() {
    if (buf) (void)hipHostFree(buf);
  }
};

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

constexpr char kMagic[8] = {'Z', 'S', 'M', 'P', 'M', 'C', 'K', 'P'};
constexpr uint32_t kVersion = 1;

// header: magic[8], u32 version, u32 fields, i64 n, i32 table size,
// i32 reserved, f64 time, i64 step (48 bytes, little-endian)
struct CkHeader {
  char magic[8];
  uint32_t version;
  uint32_t fields;
  int64_t n;
  int32_t table_size;
  int32_t reserved;
  double time;
  int64_t step;
};
static_assert(sizeof(CkHeader) == 48, "checkpoint header layout");

// device array -> file (checksummed), through the staging buffer
int32_t put_device(File& f, Fnv64& h, Staging& st, const void* dev, size_t bytes, hipStream_t s) {
  const unsigned char* d = (const unsigned char*)dev;
  for (size_t off = 0; off < bytes; off += Staging::kChunk) {
    const size_t len = bytes - off < Staging::kChunk ? bytes - off : Staging::kChunk;
    if (hipMemcpyAsync(st.buf, d + off, len, hipMemcpyDeviceToHost, s) != hipSuccess) return ZSMPM_DEVICE_ERROR;
    if (hipStreamSynchronize(s) != hipSuccess) return ZSMPM_DEVICE_ERROR;
    h.add(st.buf, len);
    if (fwrite(st.buf, 1, len, f.f) != len) return ZSMPM_IO_ERROR;
  }
  return ZSMPM_OK;
}

// file -> device array (checksummed)
int32_t get_device(File& f, Fnv64& h, Staging& st, void* dev, size_t bytes, hipStream_t s) {
  unsigned char* d = (unsigned char*)dev;
  for (size_t off = 0; off < bytes; off += Staging::kChunk) {
    const size_t len = bytes - off < Staging::kChunk ? bytes - off : Staging::kChunk;
    if (fread(st.buf, 1, len, f.f) != len) return ZSMPM_IO_ERROR;
    h.add(st.buf, len);
    if (hipMemcpyAsync(d + off, st.buf, len, hipMemcpyHostToDevice, s) != hipSuccess) return ZSMPM_DEVICE_ERROR;
    if (hipStreamSynchronize(s) != hipSuccess) return ZSMPM_DEVICE_ERROR;
  }
  return ZSMPM_OK;
}

struct Field {
  uint32_t bit;
  void* ptr;
  size_t bytes;
};

std::vector<Field> particle_fields(const ParticlesView& p) {
  const size_t n = (size_t)p.n;
  std::vector<Field> f = {{ZSMPM_CK_X, p.x, 12 * n},  {ZSMPM_CK_V, p.v, 12 * n}, {ZSMPM_CK_C, p.C, 36 * n},
                          {ZSMPM_CK_F, p.F, 36 * n},  {ZSMPM_CK_M, p.m, 4 * n},  {ZSMPM_CK_LOGJP, p.logJp, 4 * n},
                          {ZSMPM_CK_ID, p.id, 4 * n}};
  std::vector<Field> out;
  for (const Field& x : f)
    if (x.ptr) out.push_back(x);
  return out;
}

int32_t read_header(File& f, CkHeader& hd) {
  if (fread(&hd, sizeof hd, 1, f.f) != 1) return ZSMPM_IO_ERROR;
  if (memcmp(hd.magic, kMagic, 8) != 0 || hd.version != kVersion || hd.n < 0 || hd.table_size < 0)
    return ZSMPM_IO_ERROR;
  return ZSMPM_OK;
}

template <typename T>
void put_be(std::vector<unsigned char>& b, T v) {
  unsigned char raw[sizeof(T)];
  memcpy(raw, &v, sizeof(T));
  for (size_t i = 0; i < sizeof(T); ++i) b.push_back(raw[sizeof(T) - 1 - i]);
}

}  // namespace

int32_t checkpoint_save(const char* path, const ParticlesView& p, const TableView* t, double time, int64_t step,
                        hipStream_t s) {
  File f;
  f.f = fopen(path, "wb");
  if (!f.f) return ZSMPM_IO_ERROR;
  Staging st;
  if (st.init() != hipSuccess) return ZSMPM_DEVICE_ERROR;
  CkHeader hd{};
  memcpy(hd.magic, kMagic, 8);
  hd.version = kVersion;
  hd.n = p.n;
  hd.table_size = t ? t->T : 0;
  hd.time = time;
  hd.step = step;
  const std::vector<Field> fields = particle_fields(p);
  for (const Field& x : fields) hd.fields |= x.bit;
  Fnv64 h;
  h.add(&hd, sizeof hd);
  if (fwrite(&hd, sizeof hd, 1, f.f) != 1) return ZSMPM_IO_ERROR;
  for (const Field& x : fields)
    if (int32_t r = put_device(f, h, st, x.ptr, x.bytes, s); r != ZSMPM_OK) return r;
  if (t) {
    // HashTable serialisation order (container/HashTable.hpp:576-590): keys,
    // indices, status, table size, cnt, active keys
    const size_t T = (size_t)t->T;
    int32_t r = put_device(f, h, st, t->keys, 12 * T, s);
    if (r == ZSMPM_OK) r = put_device(f, h, st, t->indices, 4 * T, s);
    if (r == ZSMPM_OK) r = put_device(f, h, st, t->status, 4 * T, s);
    if (r != ZSMPM_OK) return r;
    const int32_t ts = t->T;
    h.add(&ts, 4);
    if (fwrite(&ts, 4, 1, f.f) != 1) return ZSMPM_IO_ERROR;
    r = put_device(f, h, st, t->cnt, 4, s);
    if (r == ZSMPM_OK) r = put_device(f, h, st, t->active, 12 * T, s);
    if (r != ZSMPM_OK) return r;
  }
  if (fwrite(&h.h, 8, 1, f.f) != 1) return ZSMPM_IO_ERROR;
  if (fflush(f.f) != 0) return ZSMPM_IO_ERROR;
  return ZSMPM_OK;
}

int32_t checkpoint_info(const char* path, zsmpm_checkpoint_info* out) {
  File f;
  f.f = fopen(path, "rb");
  if (!f.f) return ZSMPM_IO_ERROR;
  CkHeader hd;
  if (int32_t r = read_header(f, hd); r != ZSMPM_OK) return r;
  out->n = hd.n;
  out->table_size = hd.table_size;
  out->fields = (int32_t)hd.fields;
  out->time = hd.time;
  out->step = hd.step;
  return ZSMPM_OK;
}

int32_t checkpoint_load(const char* path, const ParticlesView& p, const TableView* t, zsmpm_checkpoint_info* out,
                        hipStream_t s) {
  File f;
  f.f = fopen(path, "rb");
  if (!f.f) return ZSMPM_IO_ERROR;
  CkHeader hd;
  if (int32_t r = read_header(f, hd); r != ZSMPM_OK) return r;
  const std::vector<Field> fields = particle_fields(p);
  uint32_t have = 0;
  for (const Field& x : fields) have |= x.bit;
  // the destination names exactly the file's arrays, at its count and table size
  if (hd.n != p.n || have != hd.fields || (t ? t->T : 0) != hd.table_size) return ZSMPM_INVALID_ARGUMENT;
  Staging st;
  if (st.init() != hipSuccess) return ZSMPM_DEVICE_ERROR;
  Fnv64 h;
  h.add(&hd, sizeof hd);
  for (const Field& x : fields)
    if (int32_t r = get_device(f, h, st, x.ptr, x.bytes, s); r != ZSMPM_OK) return r;
  if (t) {
    const size_t T = (size_t)t->T;
    int32_t r = get_device(f, h, st, t->keys, 12 * T, s);
    if (r == ZSMPM_OK) r = get_device(f, h, st, t->indices, 4 * T, s);
    if (r == ZSMPM_OK) r = get_device(f, h, st, t->status, 4 * T, s);
    if (r != ZSMPM_OK) return r;
    int32_t ts = 0;
    if (fread(&ts, 4, 1, f.f) != 1) return ZSMPM_IO_ERROR;
    h.add(&ts, 4);
    if (ts != t->T) return ZSMPM_IO_ERROR;
    r = get_device(f, h, st, t->cnt, 4, s);
    if (r == ZSMPM_OK) r = get_device(f, h, st, t->active, 12 * T, s);
    if (r != ZSMPM_OK) return r;
  }
  uint64_t sum = 0;
  if (fread(&sum, 8, 1, f.f) != 1 || sum != h.h) return ZSMPM_IO_ERROR;
  if (out) {
    out->n = hd.n;
    out->table_size = hd.table_size;
    out->fields = (int32_t)hd.fields;
    out->time = hd.time;
    out->step = hd.step;
  }
  return ZSMPM_OK;
}

// Houdini bgeo (version 5) as partio's BGEO writer lays it out, big-endian:
// magic "Bgeo", 'V', version 5, point / primitive / group / attribute counts;
// one definition per point attribute other than the position (name length,
// name, component count, Houdini type: 0 float, 5 vector; zero defaults);
// per point the position as (x, y, z, 1) followed by its attribute values;
// then the bytes 0x00 0xff.
int32_t write_bgeo(const char* path, const float* pos, int64_t n, const float* attr, int32_t width,
                   const char* name, hipStream_t s) {
  if (n > INT32_MAX) return ZSMPM_INVALID_ARGUMENT;
  std::vector<float> hp((size_t)n * 3), ha(attr ? (size_t)n * width : 0);
  if (n > 0) {
    if (hipMemcpyAsync(hp.data(), pos, hp.size() * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
      return ZSMPM_DEVICE_ERROR;
    if (attr && hipMemcpyAsync(ha.data(), attr, ha.size() * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
      return ZSMPM_DEVICE_ERROR;
  }
  if (hipStreamSynchronize(s) != hipSuccess) return ZSMPM_DEVICE_ERROR;
  std::vector<unsigned char> b;
  b.reserve(64 + (size_t)n * (16 + 4 * (size_t)(attr ? width : 0)));
  put_be<int32_t>(b, ((('B' * 256 + 'g') * 256 + 'e') * 256) + 'o');
  b.push_back('V');
  put_be<int32_t>(b, 5);
  put_be<int32_t>(b, (int32_t)n);  // points
  put_be<int32_t>(b, 0);           // primitives
  put_be<int32_t>(b, 0);           // point groups
  put_be<int32_t>(b, 0);           // primitive groups
  put_be<int32_t>(b, attr ? 1 : 0);  // point attributes besides the position
  put_be<int32_t>(b, 0);           // vertex attributes
  put_be<int32_t>(b, 0);           // primitive attributes
  put_be<int32_t>(b, 0);           // detail attributes
  if (attr) {
    const size_t len = strlen(name);
    put_be<uint16_t>(b, (uint16_t)len);
    b.insert(b.end(), name, name + len);
    put_be<uint16_t>(b, (uint16_t)width);
    put_be<int32_t>(b, width == 1 ? 0 : 5);
    for (int32_t k = 0; k < width; ++k) put_be<int32_t>(b, 0);
  }
  for (int64_t i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) put_be<float>(b, hp[3 * i + d]);
    put_be<float>(b, 1.f);
    for (int32_t k = 0; attr && k < width; ++k) put_be<float>(b, ha[(size_t)i * width + k]);
  }
  b.push_back(0x00);
  b.push_back(0xff);
  File f;
  f.f = fopen(path, "wb");
  if (!f.f) return ZSMPM_IO_ERROR;
  if (fwrite(b.data(), 1, b.size(), f.f) != b.size()) return ZSMPM_IO_ERROR;
  if (fflush(f.f) != 0) return ZSMPM_IO_ERROR;
  return ZSMPM_OK;
}

}  // namespace zsmpm

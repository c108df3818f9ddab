// mpsa_ckpt.cpp -- versioned on-disk checkpoint of block-sparse tensors (an
// MPS, a stack of environments), in place of the reference's Boost binary
// archives (MPSBase::save/load src/MPSTypes.hpp:377-426,
// Container::write/read src/container.hpp:87-146).
//
// File layout (little-endian):
//   "MPSACKPT" | u32 version (1) | u32 flags (bit 0: complex128, planar) |
//   u64 tensor count | u64 metadata bytes | metadata (caller text, e.g. JSON)
//   per tensor: u32 ns | u32 0 | u64 blocks | q (2 u64 per block) |
//               p (ns u32 per block) | rows, cols, offsets (i64 per block) |
//               u64 doubles | data (doubles; planar: all real parts, then all
//               imaginary parts) | u64 FNV-1a 64 of the tensor's bytes above
//   "MPSAEND\0" | u64 FNV-1a 64 of every byte before it
// A reader checks the magic, the version, every tensor's hash, the trailer
// and the file length, and refuses a file that fails any of them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpstools_amd.h"
#include "mpsa_internal.h"

namespace mpsa {
namespace {

struct CErr : std::runtime_error {
  int code;
  CErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

constexpr uint32_t kVersion = 1;
constexpr char kMagic[8] = {'M', 'P', 'S', 'A', 'C', 'K', 'P', 'T'};
constexpr char kEnd[8] = {'M', 'P', 'S', 'A', 'E', 'N', 'D', '\0'};

// FNV-1a 64 (Fowler, Noll, Vo): offset basis and prime of the published
// 64-bit variant
struct Fnv {
  uint64_t h = 14695981039346656037ull;
  void add(const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) {
      h ^= b[i];
      h *= 1099511628211ull;
    }
  }
};

struct Writer {
  FILE* f = nullptr;
  Fnv all, part;
  void put(const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f) != n) throw CErr(MPSA_EINVAL, "checkpoint: write failed");
    all.add(p, n);
    part.add(p, n);
  }
  template <typename T>
  void put(const T& v) {
    put(&v, sizeof(T));
  }
};

struct Tensor {
  int32_t ns = 1;
  std::vector<uint64_t> q;
  std::vector<uint32_t> p;
  std::vector<int64_t> rows, cols, off;
  std::vector<double> data;
};

int64_t extent_of(const mpsa_blocks* b) {
  int64_t e = 0;
  for (int64_t i = 0; i < b->n; ++i) e = std::max(e, b->offset[i] + b->rows[i] * b->cols[i]);
  return e;
}

}  // namespace
}  // namespace mpsa

using namespace mpsa;

struct mpsa_ckpt {
  uint32_t version = 0, flags = 0;
  std::string meta;
  std::vector<Tensor> t;
};

extern "C" {

int mpsa_ckpt_write(const char* path, int ntensors, const mpsa_blocks* tables, const double* const* data,
                    int data_on_device, const char* meta, mpsa_stream stream) {
  Writer w;
  try {
    if (!path || ntensors < 0 || (ntensors > 0 && (!tables || !data)))
      throw CErr(MPSA_EINVAL, "checkpoint_write: bad arguments");
    int cplx = -1;
    for (int i = 0; i < ntensors; ++i) {
      const mpsa_blocks& b = tables[i];
      if ((b.ns != 1 && b.ns != 2) || b.n < 0 || (b.is_complex != 0 && b.is_complex != 1) ||
          (b.n > 0 && (!b.q || !b.p || !b.rows || !b.cols || !b.offset)))
        throw CErr(MPSA_EINVAL, "checkpoint_write: bad table " + std::to_string(i));
      if (cplx >= 0 && cplx != b.is_complex) throw CErr(MPSA_EINVAL, "checkpoint_write: mixed real and complex");
      cplx = b.is_complex;
      if (extent_of(&b) > 0 && !data[i]) throw CErr(MPSA_EINVAL, "checkpoint_write: null data");
    }
    if (data_on_device) {
      const hipError_t e = hipStreamSynchronize((hipStream_t)stream);
      if (e != hipSuccess) throw CErr(MPSA_EHIP, std::string("checkpoint_write: ") + hipGetErrorString(e));
    }
    w.f = std::fopen(path, "wb");
    if (!w.f) throw CErr(MPSA_EINVAL, std::string("checkpoint_write: cannot open ") + path);
    const std::string m = meta ? meta : "";
    w.put(kMagic, 8);
    w.put(kVersion);
    w.put((uint32_t)(cplx == 1 ? 1 : 0));
    w.put((uint64_t)ntensors);
    w.put((uint64_t)m.size());
    w.put(m.data(), m.size());
    for (int i = 0; i < ntensors; ++i) {
      const mpsa_blocks& b = tables[i];
      w.part = Fnv();
      w.put((uint32_t)b.ns);
      w.put((uint32_t)0);
      w.put((uint64_t)b.n);
      w.put(b.q, sizeof(uint64_t) * 2 * (size_t)b.n);
      w.put(b.p, sizeof(uint32_t) * (size_t)b.ns * (size_t)b.n);
      w.put(b.rows, sizeof(int64_t) * (size_t)b.n);
      w.put(b.cols, sizeof(int64_t) * (size_t)b.n);
      w.put(b.offset, sizeof(int64_t) * (size_t)b.n);
      const uint64_t nd = (uint64_t)extent_of(&b) * (b.is_complex ? 2 : 1);
      w.put(nd);
      std::vector<double> host;
      const double* src = data[i];
      if (data_on_device && nd) {
        host.resize((size_t)nd);
        const hipError_t e = hipMemcpy(host.data(), data[i], sizeof(double) * nd, hipMemcpyDeviceToHost);
        if (e != hipSuccess) throw CErr(MPSA_EHIP, std::string("checkpoint_write: ") + hipGetErrorString(e));
        src = host.data();
      }
      w.put(src, sizeof(double) * (size_t)nd);
      const uint64_t h = w.part.h;
      w.put(h);
    }
    w.put(kEnd, 8);
    const uint64_t h = w.all.h;
    w.put(h);
    if (std::fclose(w.f) != 0) {
      w.f = nullptr;
      throw CErr(MPSA_EINVAL, "checkpoint_write: close failed");
    }
    return MPSA_OK;
  } catch (const CErr& e) {
    if (w.f) std::fclose(w.f);
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    if (w.f) std::fclose(w.f);
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

int mpsa_ckpt_open(const char* path, mpsa_ckpt** out, int* ntensors, int* is_complex, int64_t* meta_bytes) {
  FILE* f = nullptr;
  try {
    if (!path || !out) throw CErr(MPSA_EINVAL, "checkpoint_open: bad arguments");
    *out = nullptr;
    f = std::fopen(path, "rb");
    if (!f) throw CErr(MPSA_EINVAL, std::string("checkpoint_open: cannot open ") + path);
    std::fseek(f, 0, SEEK_END);
    const long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (len < 0) throw CErr(MPSA_EINVAL, "checkpoint_open: cannot size the file");
    std::vector<unsigned char> buf((size_t)len);
    if (len && std::fread(buf.data(), 1, (size_t)len, f) != (size_t)len)
      throw CErr(MPSA_EINVAL, "checkpoint_open: read failed");
    std::fclose(f);
    f = nullptr;
    size_t pos = 0;
    auto need = [&](size_t n) {
      if (n > buf.size() || pos > buf.size() - n) throw CErr(MPSA_EINVAL, "checkpoint_open: file is truncated");
    };
    auto get = [&](void* dst, size_t n) {
      need(n);
      if (n) std::memcpy(dst, buf.data() + pos, n);
      pos += n;
    };
    auto get_u64 = [&]() {
      uint64_t v;
      get(&v, 8);
      return v;
    };
    auto get_u32 = [&]() {
      uint32_t v;
      get(&v, 4);
      return v;
    };
    char magic[8];
    get(magic, 8);
    if (std::memcmp(magic, kMagic, 8) != 0) throw CErr(MPSA_EINVAL, "checkpoint_open: not a checkpoint file");
    std::unique_ptr<mpsa_ckpt> c(new mpsa_ckpt);
    c->version = get_u32();
    if (c->version != kVersion)
      throw CErr(MPSA_EINVAL, "checkpoint_open: format version " + std::to_string(c->version) +
                                  " (this library reads version " + std::to_string(kVersion) + ")");
    c->flags = get_u32();
    if (c->flags & ~1u) throw CErr(MPSA_EINVAL, "checkpoint_open: unknown flags");
    const uint64_t nt = get_u64();
    const uint64_t ml = get_u64();
    need((size_t)ml);
    c->meta.assign((const char*)buf.data() + pos, (size_t)ml);
    pos += (size_t)ml;
    if (nt > (uint64_t)1 << 24) throw CErr(MPSA_EINVAL, "checkpoint_open: implausible tensor count");
    for (uint64_t i = 0; i < nt; ++i) {
      const size_t start = pos;
      Tensor T;
      T.ns = (int32_t)get_u32();
      (void)get_u32();
      if (T.ns != 1 && T.ns != 2) throw CErr(MPSA_EINVAL, "checkpoint_open: bad tensor rank");
      const uint64_t n = get_u64();
      if (n > (buf.size() - pos) / 32) throw CErr(MPSA_EINVAL, "checkpoint_open: file is truncated");
      T.q.resize((size_t)2 * n);
      T.p.resize((size_t)T.ns * n);
      T.rows.resize((size_t)n);
      T.cols.resize((size_t)n);
      T.off.resize((size_t)n);
      get(T.q.data(), 16 * (size_t)n);
      get(T.p.data(), 4 * (size_t)T.ns * (size_t)n);
      get(T.rows.data(), 8 * (size_t)n);
      get(T.cols.data(), 8 * (size_t)n);
      get(T.off.data(), 8 * (size_t)n);
      int64_t ext = 0;
      for (size_t b = 0; b < (size_t)n; ++b) {
        if (T.rows[b] < 0 || T.cols[b] < 0 || T.off[b] < 0) throw CErr(MPSA_EINVAL, "checkpoint_open: bad table");
        ext = std::max(ext, T.off[b] + T.rows[b] * T.cols[b]);
      }
      const uint64_t nd = get_u64();
      if (nd != (uint64_t)ext * ((c->flags & 1u) ? 2 : 1))
        throw CErr(MPSA_EINVAL, "checkpoint_open: data length does not match the table");
      if (nd > (buf.size() - pos) / 8) throw CErr(MPSA_EINVAL, "checkpoint_open: file is truncated");
      T.data.resize((size_t)nd);
      get(T.data.data(), 8 * (size_t)nd);
      Fnv h;
      h.add(buf.data() + start, pos - start);
      if (get_u64() != h.h) throw CErr(MPSA_EINVAL, "checkpoint_open: tensor " + std::to_string(i) + " is corrupt");
      c->t.push_back(std::move(T));
    }
    char end[8];
    get(end, 8);
    if (std::memcmp(end, kEnd, 8) != 0) throw CErr(MPSA_EINVAL, "checkpoint_open: missing trailer");
    Fnv h;
    h.add(buf.data(), pos);
    if (get_u64() != h.h) throw CErr(MPSA_EINVAL, "checkpoint_open: file hash mismatch");
    if (pos != buf.size()) throw CErr(MPSA_EINVAL, "checkpoint_open: trailing bytes after the checkpoint");
    if (ntensors) *ntensors = (int)c->t.size();
    if (is_complex) *is_complex = (int)(c->flags & 1u);
    if (meta_bytes) *meta_bytes = (int64_t)c->meta.size();
    *out = c.release();
    return MPSA_OK;
  } catch (const CErr& e) {
    if (f) std::fclose(f);
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    if (f) std::fclose(f);
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

int mpsa_ckpt_meta(const mpsa_ckpt* c, char* dst, int64_t cap) {
  if (!c || (cap > 0 && !dst)) {
    set_last_error("checkpoint_meta: bad arguments");
    return MPSA_EINVAL;
  }
  const size_t n = std::min<size_t>(c->meta.size(), (size_t)std::max<int64_t>(cap, 0));
  if (n) std::memcpy(dst, c->meta.data(), n);
  return MPSA_OK;
}

int mpsa_ckpt_table(const mpsa_ckpt* c, int i, int* ns, int64_t* n_blocks, int64_t* len, uint64_t* q, uint32_t* p,
                    int64_t* rows, int64_t* cols, int64_t* offsets) {
  if (!c || i < 0 || i >= (int)c->t.size() || !n_blocks) {
    set_last_error("checkpoint_table: bad arguments");
    return MPSA_EINVAL;
  }
  const Tensor& T = c->t[(size_t)i];
  if (ns) *ns = T.ns;
  *n_blocks = (int64_t)T.rows.size();
  if (len) *len = (int64_t)T.data.size();
  const size_t n = T.rows.size();
  if (q) std::memcpy(q, T.q.data(), 16 * n);
  if (p) std::memcpy(p, T.p.data(), 4 * (size_t)T.ns * n);
  if (rows) std::memcpy(rows, T.rows.data(), 8 * n);
  if (cols) std::memcpy(cols, T.cols.data(), 8 * n);
  if (offsets) std::memcpy(offsets, T.off.data(), 8 * n);
  return MPSA_OK;
}

int mpsa_ckpt_read(const mpsa_ckpt* c, int i, double* dst, int dst_on_device, mpsa_stream stream) {
  if (!c || i < 0 || i >= (int)c->t.size() || (!dst && !c->t[(size_t)i].data.empty())) {
    set_last_error("checkpoint_read: bad arguments");
    return MPSA_EINVAL;
  }
  const auto& d = c->t[(size_t)i].data;
  if (d.empty()) return MPSA_OK;
  if (!dst_on_device) {
    std::memcpy(dst, d.data(), sizeof(double) * d.size());
    return MPSA_OK;
  }
  hipError_t e = hipMemcpyAsync(dst, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);  // the host copy lives in the handle
  if (e != hipSuccess) {
    set_last_error((std::string("checkpoint_read: ") + hipGetErrorString(e)).c_str());
    return MPSA_EHIP;
  }
  return MPSA_OK;
}

int mpsa_ckpt_close(mpsa_ckpt* c) {
  delete c;
  return MPSA_OK;
}

}  // extern "C"

// olamgpu.hip -- MI355X implementation of OLAM's RK3 dynamical core behind the C-ABI
// declared in include/olamgpu.h.
//
// Host side: grid renumbering with a checked round trip, device buffers, the RK
// drivers of prog_wrtv_rk (omodel/prog_wrtv_rk.f90:58-488) and scalar_transport_rk
// (omodel/scalar_transport_rk.F90:30-385), the tracer adjoint, and the C entry
// points.  Kernels live in kernels_tracer.h and kernels_dyn.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/olamgpu.h"
#include "common.h"
#ifndef OLAMGPU_V_ITEMS
#define OLAMGPU_V_ITEMS 2
#endif
#ifndef OLAMGPU_VEL_ITEMS
#define OLAMGPU_VEL_ITEMS 8
#endif
#ifndef OLAMGPU_SETUPV_ITEMS
#define OLAMGPU_SETUPV_ITEMS 2
#endif
#ifndef OLAMGPU_DAMP_ITEMS
#define OLAMGPU_DAMP_ITEMS 4
#endif
#include "kernels_dyn.h"
#include "kernels_tracer.h"
#include "kernels_limiter.h"

namespace {

struct Err : std::runtime_error {
  int code;
  Err(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                               \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess)                                                                       \
      throw Err(e_ == hipErrorOutOfMemory ? OLAMGPU_ENOMEM : OLAMGPU_EHIP,                     \
                std::string(#x) + ": " + hipGetErrorString(e_));                                \
  } while (0)

// ------------------------------------------------------------------ device memory
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { if (p) (void)hipFree(p); }
  void alloc(size_t b) {
    if (p) { (void)hipFree(p); p = nullptr; }
    bytes = b;
    if (b) HIPCHK(hipMalloc(&p, b));
  }
  template <class T> T *as() const { return (T *)p; }
};

// ------------------------------------------------------------------ host transfers
// A host array in the reference's Fortran layout (mza, n + 1), dummy column first, moves
// as raw bytes between host memory and a device staging buffer through two pinned
// bounce buffers (the host-side copy of one chunk overlaps the DMA of the other, and
// the host-side copy is split over threads); kernels then convert it to or from the
// device layout, renumbered columns of ldk levels (levels mza..ldk-1 zero).
template <class D, class H>
__global__ __launch_bounds__(256) void k_unpack(D *dst, const H *src, const int *perm, int n, int mza, int ldk) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * ldk) return;
  const int c = (int)(i / ldk), l = (int)(i % ldk);
  dst[i] = l < mza ? (D)src[(size_t)(perm[c] + 1) * mza + l] : D(0);
}
template <class D, class H>
__global__ __launch_bounds__(256) void k_pack(H *dst, const D *src, const int *inv, int n, int mza, int ldk) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)(n + 1) * mza) return;
  const int j = (int)(i / mza), l = (int)(i % mza);
  dst[i] = j == 0 ? H(0) : (H)src[(size_t)inv[j - 1] * ldk + l];
}

struct HostXfer {
  static constexpr size_t CHUNK = size_t(64) << 20;
  void *pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  DevBuf stage;
  int nthr = 1;
  ~HostXfer() {
    for (int k = 0; k < 2; ++k) {
      if (pin[k]) (void)hipHostFree(pin[k]);
      if (ev[k]) (void)hipEventDestroy(ev[k]);
    }
  }
  void ready(size_t bytes, hipStream_t s) {
    if (!pin[0]) {
      for (int k = 0; k < 2; ++k) {
        HIPCHK(hipHostMalloc(&pin[k], CHUNK, hipHostMallocDefault));
        HIPCHK(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
        HIPCHK(hipEventRecord(ev[k], s));
      }
      nthr = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    }
    if (stage.bytes < bytes) stage.alloc(bytes);
  }
  void copy(void *dst, const void *src, size_t nb) const {
    const size_t part = (nb + nthr - 1) / nthr;
    if (nthr == 1 || nb < (size_t(4) << 20)) { std::memcpy(dst, src, nb); return; }
    std::vector<std::thread> th;
    for (int k = 0; k < nthr; ++k) {
      const size_t o = (size_t)k * part;
      if (o >= nb) break;
      th.emplace_back([=] { std::memcpy((char *)dst + o, (const char *)src + o, std::min(part, nb - o)); });
    }
    for (auto &x : th) x.join();
  }
  // host -> stage
  void up(const void *h, size_t bytes, hipStream_t s) {
    ready(bytes, s);
    for (size_t o = 0, i = 0; o < bytes; o += CHUNK, ++i) {
      const int k = (int)(i & 1);
      const size_t nb = std::min(CHUNK, bytes - o);
      HIPCHK(hipEventSynchronize(ev[k]));          // the DMA that last read pin[k] is done
      copy(pin[k], (const char *)h + o, nb);
      HIPCHK(hipMemcpyAsync((char *)stage.p + o, pin[k], nb, hipMemcpyHostToDevice, s));
      HIPCHK(hipEventRecord(ev[k], s));
    }
  }
  // stage -> host (the stream's earlier work, the packing kernel, comes first)
  void down(void *h, size_t bytes, hipStream_t s) {
    ready(bytes, s);
    size_t prev_o = 0, prev_nb = 0;
    int prev_k = -1;
    for (size_t o = 0, i = 0; o < bytes; o += CHUNK, ++i) {
      const int k = (int)(i & 1);
      const size_t nb = std::min(CHUNK, bytes - o);
      HIPCHK(hipMemcpyAsync(pin[k], (const char *)stage.p + o, nb, hipMemcpyDeviceToHost, s));
      HIPCHK(hipEventRecord(ev[k], s));
      if (prev_k >= 0) {
        HIPCHK(hipEventSynchronize(ev[prev_k]));
        copy((char *)h + prev_o, pin[prev_k], prev_nb);
      }
      prev_k = k; prev_o = o; prev_nb = nb;
    }
    if (prev_k >= 0) {
      HIPCHK(hipEventSynchronize(ev[prev_k]));
      copy((char *)h + prev_o, pin[prev_k], prev_nb);
    }
  }
};

// ------------------------------------------------------------------ kernel timing
// Optional per-kernel timing with HIP events recorded on the context's stream
// around each launch (olamgpu_set_profiling).  Elapsed times are collected when the
// API call has synchronised its stream.
struct Prof {
  bool on = false;
  struct Rec { const char *name; hipEvent_t a, b; };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::vector<std::string> names;
  std::vector<long> count;
  std::vector<double> ms;
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
  void flush() {
    for (auto &r : pending) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
      size_t i = 0;
      while (i < names.size() && names[i] != r.name) ++i;
      if (i == names.size()) { names.push_back(r.name); count.push_back(0); ms.push_back(0.0); }
      count[i] += 1;
      ms[i] += t;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }
  void reset() { pending.clear(); names.clear(); count.clear(); ms.clear(); }
  ~Prof() { for (auto e : pool) (void)hipEventDestroy(e); }
};

#define LAUNCH_SH(NAME, KERNEL, GRID, BLOCK, SHMEM, ...)                       \
  do {                                                                          \
    hipEvent_t e0_ = nullptr, e1_ = nullptr;                                    \
    if (prof && prof->on) {                                                     \
      e0_ = prof->get(); e1_ = prof->get();                                     \
      HIPCHK(hipEventRecord(e0_, s));                                           \
    }                                                                           \
    hipLaunchKernelGGL(KERNEL, GRID, BLOCK, SHMEM, s, __VA_ARGS__);             \
    if (e0_) {                                                                  \
      HIPCHK(hipEventRecord(e1_, s));                                           \
      prof->pending.push_back({NAME, e0_, e1_});                                \
    }                                                                           \
  } while (0)
#define LAUNCH(NAME, KERNEL, GRID, BLOCK, ...) LAUNCH_SH(NAME, KERNEL, GRID, BLOCK, 0, __VA_ARGS__)

// ------------------------------------------------------------------ renumbering
// 3-D Hilbert index of a point on the unit sphere (Skilling's transpose method,
// 21 bits per axis).  Gives each XCD's contiguous block range a compact patch.
// Locality key of a point on the unit sphere: the cube face it projects to (largest
// component), then the 2-D Hilbert index of its equal-angle face coordinates.  Runs of
// consecutive keys form compact patches on the sphere, which keeps the halo of a block
// of consecutive columns small.
static uint64_t hilbert2(uint32_t x, uint32_t y, int bits) {
  const uint32_t n = 1u << bits;
  uint64_t d = 0;
  for (uint32_t s = n >> 1; s > 0; s >>= 1) {
    uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += (uint64_t)s * s * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) { x = n - 1 - x; y = n - 1 - y; }
      std::swap(x, y);
    }
  }
  return d;
}
static uint64_t sphere_key(double x, double y, double z) {
  const double c[3] = {x, y, z};
  int ax = 0;
  for (int i = 1; i < 3; ++i) if (std::fabs(c[i]) > std::fabs(c[ax])) ax = i;
  const int face = ax * 2 + (c[ax] > 0.0 ? 1 : 0);
  const double den = std::fabs(c[ax]) > 0.0 ? std::fabs(c[ax]) : 1.0;
  const int bits = 20;
  auto coord = [&](double v) {
    double a = (std::atan(v / den) / (0.25 * M_PI) + 1.0) * 0.5;
    a = std::min(std::max(a, 0.0), 1.0);
    return (uint32_t)std::min((double)((1u << bits) - 1), a * (double)(1u << bits));
  };
  const uint32_t u = coord(c[(ax + 1) % 3]), v = coord(c[(ax + 2) % 3]);
  return ((uint64_t)face << 41) | hilbert2(u, v, bits);
}

// Column patches for the LDS-staged tracer kernels.  Patch p owns the consecutive
// columns [p*P, p*P + n_own); its cell list is the owned columns, then ring 1 (their
// neighbours outside the patch), then ring 2 (neighbours of ring 1 outside both).  For
// every owned and ring-1 cell, nbr holds the patch-local index of each of its NSLOT
// neighbour slots (padded slots point at the cell itself, like w_nbr).
struct Patches {
  int P = 0, np = 0, pmax = 0, rmax = 0;
  std::vector<int> cells;        // np * pmax global column ids (unused entries: 0)
  std::vector<int> cnt;          // np * 4: n_own, n_r1, n_tot, 0
  std::vector<uint16_t> nbr;     // np * rmax * NSLOT
};

struct Topo {
  int nw, nv, nm, mza, ldk;
  std::vector<int> perm_w, inv_w, perm_v, inv_v, perm_m, inv_m;   // perm: new -> old (0-based)
  std::vector<int> w_npoly, w_nbr, w_face, w_back, v_iw, v_im, m_iv, m_im;
};

// Column stride: one wave.  Lane l of a column's wave always addresses element l of
// that column, which the column loads rely on (see ld in common.h).
// column stride: one wave of levels, or two for tall columns (64 < mza <= 128)
static int ld_k(int mza) { return mza <= WAVE ? WAVE : 2 * WAVE; }

// Fortran accessors on the input descriptor (1-based, dummy index 1)
struct In {
  const olamgpu_mesh *m;
  int iw(int j, int w) const { return m->w_iw[(size_t)(w + 1) * 7 + j] - 2; }   // w: 0-based real
  int iv(int j, int w) const { return m->w_iv[(size_t)(w + 1) * 7 + j] - 2; }
  int np(int w) const { return m->w_npoly[w + 1]; }
  int viw(int s, int v) const { return m->v_iw[(size_t)(v + 1) * 2 + s] - 2; }
  int vim(int s, int v) const { return m->v_im[(size_t)(v + 1) * 2 + s] - 2; }
  int miv(int j, int i) const { return m->m_iv[(size_t)(i + 1) * 3 + j] - 2; }
  int mim(int j, int i) const { return m->m_im[(size_t)(i + 1) * 3 + j] - 2; }
};

static Topo build_topo(const olamgpu_mesh *m) {
  Topo t;
  t.nw = m->mwa - 1; t.nv = m->mva - 1; t.nm = m->mma - 1; t.mza = m->mza; t.ldk = ld_k(m->mza);
  In in{m};
  if (t.nw < 4 || t.nv < 6 || t.nm < 4) throw Err(OLAMGPU_EINVAL, "mesh too small");
  if (m->mza < 4 || m->mza > 2 * WAVE) throw Err(OLAMGPU_EINVAL, "mza must be in [4, 128]");
  // a tall column takes two waves of a block; whole blocks of whole items keep every
  // wave of a block on the same sequence of LDS exchanges
  if (m->mza > WAVE && ((t.nw | t.nv | t.nm) & 1))
    throw Err(OLAMGPU_EINVAL, "more than 64 levels need even numbers of W, V and M points");
  for (int i = 1; i < m->mwa; ++i)
    if (m->lpw[i] != 2) throw Err(OLAMGPU_EINVAL, "only flat terrain (lpw = 2) is supported");
  for (int i = 1; i < m->mva; ++i)
    if (m->lpv[i] != 2) throw Err(OLAMGPU_EINVAL, "only flat terrain (lpv = 2) is supported");
  for (int i = 1; i < m->mma; ++i)
    if (m->lpm[i] != 2) throw Err(OLAMGPU_EINVAL, "only flat terrain (lpm = 2) is supported");
  for (int w = 0; w < t.nw; ++w) {
    int np = in.np(w);
    if (np < 3 || np > 7) throw Err(OLAMGPU_EINVAL, "npoly out of range");
    for (int j = 0; j < np; ++j) {
      int a = in.iw(j, w), v = in.iv(j, w);
      if (a < 0 || a >= t.nw || v < 0 || v >= t.nv) throw Err(OLAMGPU_EINVAL, "itab_w index out of range");
      int p = in.viw(0, v), q = in.viw(1, v);
      if (!((p == w && q == a) || (p == a && q == w))) throw Err(OLAMGPU_EINVAL, "itab_w/itab_v mismatch");
    }
  }
  for (int v = 0; v < t.nv; ++v)
    for (int s = 0; s < 2; ++s) {
      int a = in.viw(s, v), b = in.vim(s, v);
      if (a < 0 || a >= t.nw || b < 0 || b >= t.nm) throw Err(OLAMGPU_EINVAL, "itab_v index out of range");
    }
  for (int i = 0; i < t.nm; ++i)
    for (int j = 0; j < 3; ++j) {
      int v = in.miv(j, i), o = in.mim(j, i);
      if (v < 0 || v >= t.nv || o < 0 || o >= t.nm) throw Err(OLAMGPU_EINVAL, "itab_m index out of range");
    }
  // W: locality order of the unit position vectors (wnx, wny, wnz)
  std::vector<uint64_t> key(t.nw);
  for (int w = 0; w < t.nw; ++w) key[w] = sphere_key(m->wnx[w + 1], m->wny[w + 1], m->wnz[w + 1]);
  t.perm_w.resize(t.nw);
  std::iota(t.perm_w.begin(), t.perm_w.end(), 0);
  std::stable_sort(t.perm_w.begin(), t.perm_w.end(), [&](int a, int b) { return key[a] < key[b]; });
  t.inv_w.resize(t.nw);
  for (int c = 0; c < t.nw; ++c) t.inv_w[t.perm_w[c]] = c;
  // V: by owning (lower-numbered) W, then the other W
  std::vector<uint64_t> kv(t.nv);
  for (int v = 0; v < t.nv; ++v) {
    uint64_t a = t.inv_w[in.viw(0, v)], b = t.inv_w[in.viw(1, v)];
    kv[v] = (std::min(a, b) << 32) | std::max(a, b);
  }
  t.perm_v.resize(t.nv);
  std::iota(t.perm_v.begin(), t.perm_v.end(), 0);
  std::stable_sort(t.perm_v.begin(), t.perm_v.end(), [&](int a, int b) { return kv[a] < kv[b]; });
  t.inv_v.resize(t.nv);
  for (int c = 0; c < t.nv; ++c) t.inv_v[t.perm_v[c]] = c;
  // M: by the lowest new index of its three V points' W neighbours
  std::vector<uint64_t> km(t.nm);
  for (int i = 0; i < t.nm; ++i) {
    uint64_t lo = UINT64_MAX;
    for (int j = 0; j < 3; ++j) {
      int v = in.miv(j, i);
      lo = std::min<uint64_t>(lo, t.inv_v[v]);
    }
    km[i] = lo;
  }
  t.perm_m.resize(t.nm);
  std::iota(t.perm_m.begin(), t.perm_m.end(), 0);
  std::stable_sort(t.perm_m.begin(), t.perm_m.end(), [&](int a, int b) { return km[a] < km[b]; });
  t.inv_m.resize(t.nm);
  for (int c = 0; c < t.nm; ++c) t.inv_m[t.perm_m[c]] = c;
  // renumbered tables
  t.w_npoly.assign(t.nw, 0);
  t.w_nbr.assign((size_t)t.nw * NSLOT, 0);
  t.w_face.assign((size_t)t.nw * NSLOT, 0);
  t.w_back.assign((size_t)t.nw * NSLOT, 0);
  for (int c = 0; c < t.nw; ++c) {
    int o = t.perm_w[c], np = in.np(o);
    t.w_npoly[c] = np;
    for (int j = 0; j < np; ++j) {
      t.w_nbr[(size_t)c * NSLOT + j] = t.inv_w[in.iw(j, o)];
      t.w_face[(size_t)c * NSLOT + j] = t.inv_v[in.iv(j, o)];
    }
    for (int j = np; j < NSLOT; ++j) { t.w_nbr[(size_t)c * NSLOT + j] = c; t.w_face[(size_t)c * NSLOT + j] = 0; }
  }
  for (int c = 0; c < t.nw; ++c)
    for (int j = 0; j < t.w_npoly[c]; ++j) {
      int n = t.w_nbr[(size_t)c * NSLOT + j], b = -1;
      for (int i = 0; i < t.w_npoly[n]; ++i) if (t.w_nbr[(size_t)n * NSLOT + i] == c) b = i;
      if (b < 0) throw Err(OLAMGPU_EINVAL, "W adjacency is not symmetric");
      t.w_back[(size_t)c * NSLOT + j] = b;
    }
  t.v_iw.resize((size_t)t.nv * 2); t.v_im.resize((size_t)t.nv * 2);
  for (int e = 0; e < t.nv; ++e) {
    int o = t.perm_v[e];
    for (int s = 0; s < 2; ++s) { t.v_iw[e * 2 + s] = t.inv_w[in.viw(s, o)]; t.v_im[e * 2 + s] = t.inv_m[in.vim(s, o)]; }
  }
  t.m_iv.assign((size_t)t.nm * 4, 0); t.m_im.assign((size_t)t.nm * 4, 0);
  for (int i = 0; i < t.nm; ++i) {
    int o = t.perm_m[i];
    for (int j = 0; j < 3; ++j) { t.m_iv[i * 4 + j] = t.inv_v[in.miv(j, o)]; t.m_im[i * 4 + j] = t.inv_m[in.mim(j, o)]; }
    t.m_iv[i * 4 + 3] = t.m_iv[i * 4]; t.m_im[i * 4 + 3] = t.m_im[i * 4];
  }
  // round trip: map every renumbered table back and compare with the input exactly
  for (int c = 0; c < t.nw; ++c) {
    int o = t.perm_w[c];
    if (t.w_npoly[c] != in.np(o)) throw Err(OLAMGPU_EINVAL, "renumbering check failed (npoly)");
    for (int j = 0; j < t.w_npoly[c]; ++j) {
      if (t.perm_w[t.w_nbr[(size_t)c * NSLOT + j]] + 2 != m->w_iw[(size_t)(o + 1) * 7 + j] ||
          t.perm_v[t.w_face[(size_t)c * NSLOT + j]] + 2 != m->w_iv[(size_t)(o + 1) * 7 + j])
        throw Err(OLAMGPU_EINVAL, "renumbering check failed (itab_w)");
    }
  }
  for (int e = 0; e < t.nv; ++e) {
    int o = t.perm_v[e];
    for (int s = 0; s < 2; ++s)
      if (t.perm_w[t.v_iw[e * 2 + s]] + 2 != m->v_iw[(size_t)(o + 1) * 2 + s] ||
          t.perm_m[t.v_im[e * 2 + s]] + 2 != m->v_im[(size_t)(o + 1) * 2 + s])
        throw Err(OLAMGPU_EINVAL, "renumbering check failed (itab_v)");
  }
  for (int i = 0; i < t.nm; ++i) {
    int o = t.perm_m[i];
    for (int j = 0; j < 3; ++j)
      if (t.perm_v[t.m_iv[i * 4 + j]] + 2 != m->m_iv[(size_t)(o + 1) * 3 + j] ||
          t.perm_m[t.m_im[i * 4 + j]] + 2 != m->m_im[(size_t)(o + 1) * 3 + j])
        throw Err(OLAMGPU_EINVAL, "renumbering check failed (itab_m)");
  }
  return t;
}

static Patches build_patches(const Topo &t, int P) {
  Patches pt;
  pt.P = P;
  pt.np = (t.nw + P - 1) / P;
  std::vector<int> stamp(t.nw, -1), loc(t.nw, 0);
  std::vector<std::vector<int>> lists(pt.np);
  for (int p = 0; p < pt.np; ++p) {
    std::vector<int> &L = lists[p];
    const int c0 = p * P, c1 = std::min(t.nw, c0 + P);
    for (int c = c0; c < c1; ++c) { stamp[c] = p; loc[c] = (int)L.size(); L.push_back(c); }
    const int n_own = (int)L.size();
    for (int i = 0; i < n_own; ++i)
      for (int j = 0; j < t.w_npoly[L[i]]; ++j) {
        int n = t.w_nbr[(size_t)L[i] * NSLOT + j];
        if (stamp[n] != p) { stamp[n] = p; loc[n] = (int)L.size(); L.push_back(n); }
      }
    const int n_r1 = (int)L.size() - n_own;
    for (int i = n_own; i < n_own + n_r1; ++i)
      for (int j = 0; j < t.w_npoly[L[i]]; ++j) {
        int n = t.w_nbr[(size_t)L[i] * NSLOT + j];
        if (stamp[n] != p) { stamp[n] = p; loc[n] = (int)L.size(); L.push_back(n); }
      }
    pt.pmax = std::max(pt.pmax, (int)L.size());
    pt.rmax = std::max(pt.rmax, n_own + n_r1);
    pt.cnt.push_back(n_own); pt.cnt.push_back(n_r1); pt.cnt.push_back((int)L.size()); pt.cnt.push_back(0);
  }
  if (pt.pmax > 65535) throw Err(OLAMGPU_EINVAL, "patch too large");
  pt.cells.assign((size_t)pt.np * pt.pmax, 0);
  pt.nbr.assign((size_t)pt.np * pt.rmax * NSLOT, 0);
  for (int p = 0; p < pt.np; ++p) {
    const std::vector<int> &L = lists[p];
    for (size_t i = 0; i < L.size(); ++i) { stamp[L[i]] = p; loc[L[i]] = (int)i; pt.cells[(size_t)p * pt.pmax + i] = L[i]; }
    const int nr = pt.cnt[p * 4] + pt.cnt[p * 4 + 1];
    for (int i = 0; i < nr; ++i)
      for (int j = 0; j < NSLOT; ++j) {
        int n = t.w_nbr[(size_t)L[i] * NSLOT + j];
        if (stamp[n] != p) throw Err(OLAMGPU_EINVAL, "patch construction failed");
        pt.nbr[((size_t)p * pt.rmax + i) * NSLOT + j] = (uint16_t)loc[n];
      }
  }
  return pt;
}

// ------------------------------------------------------------------ implementation
// ---------------------------------------------------------------- allocation-free mass sums
// olamgpu_mass_sums_noalloc forms the sums of olamgpu_mass_sums with its per-block partial
// sums in this file-scope device array instead of a buffer the library allocates: a
// fixed grid of MASS_NB blocks strides over the columns, each block keeping MASS_NV
// partials (dry mass and up to 64 tracers).  Each GPU holds its own copy of the array
// until the process ends; every call on a GPU shares it, so calls that overlap on one GPU
// race.  A call writes every partial it reads, so nothing needs clearing between calls.
constexpr int MASS_NB = 1024, MASS_NV = 65;
static __device__ __align__(256) double scratch[MASS_NB * MASS_NV];     // 520 KiB per GPU

// Device address of `scratch` on each device, taken once per device with that device
// current (olamgpu_create): host code uses this, never the array's name.
static std::mutex g_scratch_mu;
static double *g_scratch[64] = {};
static void scratch_setup(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= 64) throw Err(OLAMGPU_EINVAL, "device index beyond the scratch table");
  if (g_scratch[dev]) return;
  void *p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[dev] = static_cast<double *>(p);
}
static double *scratch_ptr(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return (dev >= 0 && dev < 64) ? g_scratch[dev] : nullptr;
}

// Partial sums of rho volt and rho volt q_n (FP64) per block over the columns
// blockIdx.x * WPB + wave, stepping by gridDim.x * WPB; each wave keeps its own row of
// LDS partials, summed in a fixed order, so the result is deterministic.
template <class R, class RP>
__global__ __launch_bounds__(256) void k_mass_partial_gs(DMesh<R> m, const RP *rho, const R *q, int ntr,
                                                         long tstride, double *part) {
  __shared__ double sh[WPB][MASS_NV];
  const int w = threadIdx.x >> 6, lane = lane_id();
  for (int v = lane; v <= ntr; v += WAVE) sh[w][v] = 0.0;
  __syncthreads();
  const int nitem = m.tall ? 2 * m.nw : m.nw;                // a tall column: two waves
  for (int g = blockIdx.x * WPB + w; g < nitem; g += gridDim.x * WPB) {
    const int c = m.tall ? g >> 1 : g;
    const int l = (m.tall ? (g & 1) * WAVE : 0) + lane, k = l + 1;
    const bool lev = k >= 2 && k <= m.mza;
    const long ic = (long)c * m.ldk + l;
    const double mass = lev ? (double)rho[ic] * m.volt[ic] : 0.0;
    for (int n = -1; n < ntr; ++n) {
      double v = n < 0 ? mass : (lev ? mass * (double)q[n * tstride + ic] : 0.0);
      for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, WAVE);
      if (lane == 0) sh[w][n + 1] += v;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x <= ntr) {
    double s = 0.0;
    for (int i = 0; i < WPB; ++i) s += sh[i][threadIdx.x];
    part[(long)blockIdx.x * MASS_NV + threadIdx.x] = s;
  }
}

struct ImplBase {
  virtual ~ImplBase() {}
  virtual void upload_state(const olamgpu_state *) = 0;
  virtual void download_state(olamgpu_state *) = 0;
  virtual void upload_scalar(int, const float *, const float *) = 0;
  virtual void download_scalar(int, float *) = 0;
  virtual void upload_adj(int, const float *) = 0;
  virtual void download_adj(int, float *) = 0;
  virtual void init_wrtv_rk() = 0;
  virtual void tend0() = 0;
  virtual void prog_wrtv_rk() = 0;
  virtual void timeavg_momsc() = 0;
  virtual void comp_alpha_press(const float *, const float *, const float *) = 0;
  virtual void set_turbulence(const float *, const float *, const int32_t *, const int32_t *, const float *) = 0;
  virtual void set_nudging(const float *, const float *, const double *) = 0;
  virtual void scalar_transport_rk(const float *) = 0;
  virtual void scalar_transport_rk_adj() = 0;
  virtual void mass_sums(double *) = 0;
  virtual void cfl_report(double *) = 0;
  virtual void copy_field(int, int, void *, int, int *) = 0;
  virtual size_t checkpoint_bytes(int) = 0;
  virtual void checkpoint(int, void *, int) = 0;
  virtual void set_pdef(int, int) = 0;
  virtual void set_sgsmix(int, int) = 0;
  virtual void tracer_maxmin(float *) = 0;
  virtual void mass_sums_noalloc(double *, double *) = 0;
};

// x**n for an integer n >= 0 in the type of x, one rounding per factor (Fortran's
// real**integer, e.g. vort_div_damp.f90:80,318)
template <class T> static inline T ipow(T x, int n) { T r = T(1); for (int i = 0; i < n; ++i) r *= x; return r; }

// FP32 x**y of a real exponent: FP64 pow rounded to FP32
static inline float powr(float x, float y) { return (float)std::pow((double)x, (double)y); }

// Rayleigh friction profiles of rayf_init (mem_rayf.f90:41-156), the W factor b20 of
// init_wrtv_rk (prog_wrtv_rk.f90:46-52) and the divergence profile zf of divh_damp
// (vort_div_damp.f90:60-66), in the reference's FP32.  Arrays are per level [k-1].
struct Rayf {
  bool dorayf = false, dorayfw = false, dorayfdiv = false, dorayfmix = false;
  int krayf_bot, krayfw_bot, krayfdiv_bot, krayfmix_bot;
  std::vector<float> cof, cofw, cofmix, zf, b20;
};
static Rayf rayf_profiles(const olamgpu_params &p, int mza, const float *zm, const float *zt) {
  Rayf r;
  r.krayf_bot = r.krayfw_bot = r.krayfdiv_bot = r.krayfmix_bot = mza + 1;
  r.cof.assign(mza, 0.f); r.cofw.assign(mza, 0.f); r.cofmix.assign(mza, 0.f); r.zf.assign(mza, 0.f);
  r.b20.assign(mza, 1.f);
  const bool any = p.rayf_fact > 1.e-7f || p.rayfw_fact > 1.e-7f || p.rayfdiv_fact > 1.e-7f ||
                   p.rayfmix_fact > 1.e-7f;
  if (!any) return r;
  if (!zm || !zt) throw Err(OLAMGPU_EINVAL, "Rayleigh friction needs the zm and zt level heights");
  auto Z = [&](const float *a, int k) { return a[k - 1]; };
  const float dti = 1.0f / (float)p.dtsm;
  const float ztop = Z(zm, mza);
  if (p.rayf_fact > 1.e-7f && p.initial == 1) {
    for (int k = 2; k <= mza; ++k) if (Z(zt, k) > p.rayf_zmin) { r.krayf_bot = k; break; }
    if (r.krayf_bot <= mza) {
      r.dorayf = true;
      const float distimi = p.rayf_fact * dti;
      for (int k = r.krayf_bot; k <= mza; ++k)
        r.cof[k - 1] = distimi * powr((Z(zt, k) - p.rayf_zmin) / (ztop - p.rayf_zmin), p.rayf_expon);
    }
  }
  if (p.rayfw_fact > 1.e-7f) {
    for (int k = 2; k <= mza - 1; ++k) if (Z(zm, k) > p.rayfw_zmin) { r.krayfw_bot = k; break; }
    if (r.krayfw_bot < mza) {
      r.dorayfw = true;
      const float distimi = p.rayfw_fact * dti;
      for (int k = r.krayfw_bot; k <= mza - 1; ++k) {
        r.cofw[k - 1] = distimi * powr((Z(zm, k) - p.rayfw_zmin) / (ztop - p.rayfw_zmin), p.rayfw_expon);
        r.b20[k - 1] = 1.0f + (float)p.dtsm * r.cofw[k - 1];
      }
    }
  }
  if (p.rayfdiv_fact > 1.e-7f) {
    for (int k = 2; k <= mza; ++k) if (Z(zt, k) > p.rayfdiv_zmin) { r.krayfdiv_bot = k; break; }
    if (r.krayfdiv_bot <= mza) {
      r.dorayfdiv = true;
      for (int k = r.krayfdiv_bot; k <= mza; ++k)
        r.zf[k - 1] = powr((Z(zt, k) - p.rayfdiv_zmin) / (ztop - p.rayfdiv_zmin), p.rayfdiv_expon);
    }
  }
  if (p.rayfmix_fact > 1.e-7f) {
    for (int k = 2; k <= mza - 1; ++k) if (Z(zm, k) > p.rayfmix_zmin) { r.krayfmix_bot = k; break; }
    if (r.krayfmix_bot < mza) {
      r.dorayfmix = true;
      // the profile is normalised by zm(mza) - rayfw_zmin, as in mem_rayf.f90:150
      for (int k = r.krayfmix_bot; k <= mza - 1; ++k)
        r.cofmix[k - 1] = 0.5f * p.rayfmix_fact *
                          powr((Z(zm, k) - p.rayfmix_zmin) / (ztop - p.rayfw_zmin), p.rayfmix_expon);
    }
  }
  return r;
}

template <class R, class RP>
struct Impl : ImplBase {
  hipStream_t s;
  Prof *prof = nullptr;
  olamgpu_params prm;
  Topo t;
  int nw, nv, nm, mza, ldk;
  size_t wsz, vsz, msz;        // elements per W/V/M field
  DMesh<R> dm;
  // mesh storage
  DevBuf b_npoly, b_nbr, b_face, b_back, b_viw, b_vim, b_miv, b_mim, b_invw, b_invv;
  DevBuf b_dirv, b_gxc, b_gyc, b_dxs, b_dys, b_dxn, b_dyn, b_ecx, b_ecy, b_ecz, b_xxyy, b_wn[3], b_arw0, b_zfacit, b_dniv;
  DevBuf b_vn[3], b_vfact, b_xxyym, b_lev[11], b_lev8[4];
  DevBuf b_arw, b_volti, b_volwi, b_volt, b_arv, b_volvi;
  DevBuf b_q[2][5], b_os[2][3], b_on[2][3];   // order-3 reconstruction (DMesh::w_q, w_os, w_on)
  // column patches of the LDS-staged tracer kernels (forward and adjoint)
  struct PatchSet {
    Patches h;
    DevBuf cells, cnt, nbr;
    PatchTab tab{};
    int cpw = 1;
    size_t lds = 0;
  } pf, pa;
  bool use_patch = false;   // OLAMGPU_TRACER=patch selects the LDS-staged tracer kernels
  // Order-2 tracer stages (outside a limited final stage) run as a gradient pass and an
  // update pass with the gradient pairs stored side by side (default), as the same two
  // passes with gx and gy in separate arrays (OLAMGPU_TRACER=twopass), or as one pass
  // that forms every gradient where it is used (OLAMGPU_TRACER=fused; measured 2x slower
  // at the bench size: it gathers the two-ring stencil of every upwind neighbour).
  bool two_pass = false, fused_fwd = false;
  bool fuse_adj = true;     // one pass for the column and face parts of the adjoint stage
  // state
  DevBuf thil[2], wmc, wc, vxe, vye, vze, vmc, vc, vmasc, wmasc, alpha, pwfac, pvfac, thilt, vmxet, vmyet, vmzet, vmt;
  DevBuf rho, press, rho_old;
  int cur_thil = 0;
  // prog_wrtv_rk scratch
  DevBuf rho0, rth0, wmc0, vmc0, thilt_s, vmxet_s, vmyet_s, vmzet_s, vmt_s, vmxet_rk, vmyet_rk, vmzet_rk, grads[8];
  DevBuf gradq[4][3];          // order 3: quadratic terms of thil, vxe, vye, vze
  // fluxes_wrtv_monot (ithil_monot / iwind_monot): RK-start copies, scpt, scales
  DevBuf mq0[4], mscpt[4], ms1[4], ms2[4];
  int mono_on[4] = {0, 0, 0, 0};
  DevBuf div2d[3], vort2d[2], fdivl, fdivh, fvorth;
  int div_ih = 0, vort_ih = 0, div_l = 0, kdivl = 2;
  bool do_div = false, do_vort = false, do_divl = false, do_divh = false;
  // Rayleigh friction: profiles, the wind vc03d the friction relaxes toward, and
  // whether init_wrtv_rk has taken vc03d
  Rayf rf;
  DevBuf b_rayf, vc03d;
  bool rayf_ready = false;
  // turbulence of the host's PBL scheme and comp_horiz_k's face coefficients
  DevBuf tb_vkh, tb_vkm, tb_akm, tb_dzitb, tb_khtop, tb_kmtop, tb_akhodx, tb_akmodx, tb_khtopv, tb_kmtopv;
  DTurb<R> dturb{};
  // nudging fields of mem_nudge (zero until olamgpu_set_nudging)
  DevBuf nud_w, nud_v, nud_r;
  // tracers
  int ntr;
  long tstride;
  // trG holds the gradients of the stage input: gx then gy (ntr fields each), or at order
  // 2 outside a limited stage the (gx, gy) pairs side by side (measured at NXP=320, L=60,
  // 10 tracers: grad + update 6.47 ms per stage against 6.66 with separate arrays; the
  // adjoint's pair layout was slower, 7.64 against 7.20, and is not used); lgG holds the
  // adjoint's gradient adjoints, lgx then lgy
  DevBuf trS, trX, trY, trG, trsct;
  DevBuf trgq[3];              // order 3: quadratic reconstruction terms
  bool have_sct = false;
  // final-stage limiters (iscal_monot = 1 monotone, 2 positive definite)
  std::vector<int> pdef;            // scalar_tab(n)%pdef, default .true. (var_tables.f90:58)
  std::vector<int> sgsmix;          // scalar_tab(n)%do_sgsmix, default .true. (var_tables.f90:59)
  DevBuf sgs_dev, rhof, lsct;       // do_sgsmix per tracer; nudged tracer density; adjoint of sct
  DevBuf lim_mode, lim_pre, lim_s1, lim_s2, lim_tsv, lim_tsw;
  bool any_lim = false, any_pre = false, taped_lim = false;
  DevBuf ffimp, cflmax, nimp, anyimp, masspart, massout, mmkeys, mmout;
  // adjoint
  DevBuf adjA, adjB, lam0, nu, lpart, lgG, lgq[3];
  int cur_adj = 0;
  bool have_fwd = false;
  // ffimp (and whether any column is implicit) is current for the present vmasc, wmasc
  // and rho_old: set by check_cfls, cleared by every call that can change those fields
  bool cfl_fresh = false, cfl_any = false;
  double cfl_out[4] = {0, 0, 0, 0};

  template <class T> T *dp(DevBuf &b) { return b.as<T>(); }

  void up(DevBuf &b, const void *h, size_t bytes) {
    b.alloc(bytes);
    HIPCHK(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  template <class T> void upv(DevBuf &b, const std::vector<T> &v) { up(b, v.data(), v.size() * sizeof(T)); }

  // waves per point: 2 for tall columns (each item's halves in one block)
  int halves = 1;
  unsigned nblk(int n) const { return (unsigned)(((long)n * halves + WPB - 1) / WPB); }
  // blocks for q items per wave (per wave pair for tall columns)
  unsigned nblk_per(int n, int q) const { return (unsigned)((((long)n + q - 1) / q * halves + WPB - 1) / WPB); }
  // tracers per wave in the tracer kernels: all of them up to 16 (measured at NXP=320,
  // L=60, 10 tracers: 1 -> 151, 2 -> 121, 5 -> 107, 10 -> 104 ms per step, because the
  // per-column gathers of winds and metrics are shared); OLAMGPU_TPW overrides
  int tpw = 16;
  unsigned nblk_tr() const { int ng = (std::max(ntr, 1) + tpw - 1) / tpw; return nblk(nw * ng); }

  Impl(const olamgpu_mesh *m, const olamgpu_params *p, hipStream_t st, Prof *pr, Topo &&topo)
      : s(st), prof(pr), prm(*p), t(std::move(topo)) {
    nw = t.nw; nv = t.nv; nm = t.nm; mza = t.mza; ldk = t.ldk;
    wsz = (size_t)nw * ldk; vsz = (size_t)nv * ldk; msz = (size_t)nm * ldk;
    if (prm.nrk_wrtv < 2 || prm.nrk_wrtv > 3 || prm.nrk_scal < 2 || prm.nrk_scal > 3)
      throw Err(OLAMGPU_EINVAL, "RK levels must be 2 or 3");
    for (int o : {prm.thil_horiz_adv_order, prm.wind_horiz_adv_order, prm.scal_horiz_adv_order}) {
      if (o < 1 || o > 3) throw Err(OLAMGPU_EINVAL, "horizontal advection order must be 1, 2 or 3");
      if (o == 3 && !(m->xy_h && m->xy_hu && m->xx0_v && m->xy0_v && m->yy0_v && m->xx0_vu && m->xy0_vu &&
                      m->yy0_vu))
        throw Err(OLAMGPU_EINVAL, "order 3 needs xy_h, xy_hu and the xx0_v/xx0_vu face offsets");
    }
    if (prm.iscal_monot < 0 || prm.iscal_monot > 2) throw Err(OLAMGPU_EINVAL, "iscal_monot must be 0, 1 or 2");
    if (prm.num_scalar < 0 || prm.num_scalar > 64) throw Err(OLAMGPU_EINVAL, "num_scalar must be in [0, 64]");
    if (mza > WAVE) halves = 2;   // tall columns: two waves per column
    build_mesh(m);
    alloc_state();
    if (prm.nudflag > 0) alloc_nudging();
    if (const char *e = std::getenv("OLAMGPU_TPW")) tpw = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("OLAMGPU_TRACER")) {
      use_patch = std::strcmp(e, "patch") == 0;
      two_pass = std::strcmp(e, "twopass") == 0;
      fused_fwd = std::strcmp(e, "fused") == 0;
      fuse_adj = std::strcmp(e, "split") != 0;   // OLAMGPU_TRACER=split: separate adjoint passes
    }
    if (halves == 2) {
      // one tracer group (every item of a block runs the same tracer loop) and the
      // default kernel arrangement
      tpw = std::max(ntr, 1);
      use_patch = two_pass = fused_fwd = false;
      fuse_adj = true;
    }
    build_patch_tables();
  }

  // Patches of 8 * cpw columns (one block of 8 waves, cpw owned columns per wave),
  // halved from the largest size until the LDS of the kernel fits the device's limit
  // per workgroup.  Forward: the stencil records of owned + ring-1 cells and one staged
  // tracer row per patch cell.  Adjoint: in addition the gradient adjoints (Gx, Gy) of
  // owned + ring-1 cells.
  size_t lds_limit = 65536;
  // staged rows per wave that the forward kernel prefetches into registers
  static constexpr int rows_per_wave(int cpw) { return cpw == 4 ? 20 : (cpw == 2 ? 16 : 12); }
  size_t lds_fwd(const Patches &h) const { return (size_t)h.rmax * sizeof(PCell<R>) + (size_t)h.pmax * WAVE * sizeof(R); }
  size_t lds_adj(const Patches &h) const { return lds_fwd(h) + (size_t)h.rmax * 2 * WAVE * sizeof(R); }
  void build_patch_set(PatchSet &ps, bool adj) {
    int cpw = 4;
    // OLAMGPU_PATCH_CPW (forward) / OLAMGPU_ADJ_CPW (adjoint): 1, 2 or 4, for tuning
    if (const char *e = std::getenv(adj ? "OLAMGPU_ADJ_CPW" : "OLAMGPU_PATCH_CPW")) {
      const int v = std::atoi(e);
      if (v == 1 || v == 2 || v == 4) cpw = v;
    }
    for (;;) {
      ps.h = build_patches(t, 8 * cpw);
      ps.lds = adj ? lds_adj(ps.h) : lds_fwd(ps.h);
      if ((ps.lds <= lds_limit && ps.h.pmax <= 8 * rows_per_wave(cpw)) || cpw == 1) break;
      cpw /= 2;
    }
    if (ps.lds > lds_limit || ps.h.pmax > 8 * rows_per_wave(cpw))
      throw Err(OLAMGPU_EINVAL, "mesh neighbourhoods too large for the LDS patches");
    ps.cpw = cpw;
    upv(ps.cells, ps.h.cells); upv(ps.cnt, ps.h.cnt);
    ps.nbr.alloc(ps.h.nbr.size() * sizeof(uint16_t));
    HIPCHK(hipMemcpyAsync(ps.nbr.p, ps.h.nbr.data(), ps.nbr.bytes, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    ps.tab.cells = (const int *)ps.cells.p; ps.tab.cnt = (const int *)ps.cnt.p; ps.tab.nbr = (const uint16_t *)ps.nbr.p;
    ps.tab.pmax = ps.h.pmax; ps.tab.rmax = ps.h.rmax;
  }
  void build_patch_tables() {
    int dev = 0, v = 0;
    HIPCHK(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      lds_limit = (size_t)v;
    build_patch_set(pf, false);
    build_patch_set(pa, true);
  }

  // -------------------------------------------------------------- mesh upload
  void build_mesh(const olamgpu_mesh *m) {
    In in{m};
    const bool f64 = sizeof(R) == 8;
    std::vector<R> dirv(nw * NSLOT, 0), gxc(nw * NSLOT, 0), gyc(nw * NSLOT, 0), dxs(nw * NSLOT, 0),
        dys(nw * NSLOT, 0), dxn(nw * NSLOT, 0), dyn(nw * NSLOT, 0), ecx(nw * NSLOT, 0), ecy(nw * NSLOT, 0),
        ecz(nw * NSLOT, 0), xxyy(nw * NSLOT, 0), wn[3] = {std::vector<R>(nw), std::vector<R>(nw), std::vector<R>(nw)}, aw0(nw);
    for (int c = 0; c < nw; ++c) {
      int o = t.perm_w[c];
      size_t fo = (size_t)(o + 1) * 7;
      for (int j = 0; j < t.w_npoly[c]; ++j) {
        size_t i = (size_t)c * NSLOT + j;
        dirv[i] = m->w_dirv[fo + j];
        gxc[i] = m->gxps_coef[fo + j]; gyc[i] = m->gyps_coef[fo + j];
        ecx[i] = m->w_ecvec_vx[fo + j]; ecy[i] = m->w_ecvec_vy[fo + j]; ecz[i] = m->w_ecvec_vz[fo + j];
        xxyy[i] = m->xx_yy[fo + j];
        int v = in.iv(j, o);
        int side = (m->w_dirv[fo + j] < 0.f) ? 0 : 1;   // this cell is iw1 (dirv < 0) or iw2
        dxs[i] = m->v_dxps[(size_t)(v + 1) * 2 + side]; dys[i] = m->v_dyps[(size_t)(v + 1) * 2 + side];
        dxn[i] = m->v_dxps[(size_t)(v + 1) * 2 + 1 - side]; dyn[i] = m->v_dyps[(size_t)(v + 1) * 2 + 1 - side];
      }
      wn[0][c] = m->wnx[o + 1]; wn[1][c] = m->wny[o + 1]; wn[2][c] = m->wnz[o + 1];
      aw0[c] = m->arw0[o + 1];
    }
    // order-3 terms: xy_h(t, n, iw) at [(iw - 1) * 35 + (n - 1) * 5 + t - 1]; face offsets
    // xx0_v(side, iv) at [(iv - 1) * 2 + side - 1]
    std::vector<R> q3[2][5], os3[2][3], on3[2][3];
    const bool have_q3 = m->xy_h && m->xy_hu && m->xx0_v && m->xy0_v && m->yy0_v && m->xx0_vu && m->xy0_vu &&
                         m->yy0_vu;
    for (int b = 0; b < 2; ++b) {
      for (auto &v : q3[b]) v.assign(nw * NSLOT, R(0));
      for (auto &v : os3[b]) v.assign(nw * NSLOT, R(0));
      for (auto &v : on3[b]) v.assign(nw * NSLOT, R(0));
    }
    if (have_q3) {
      const float *xy[2] = {m->xy_h, m->xy_hu};
      const float *off[2][3] = {{m->xx0_v, m->xy0_v, m->yy0_v}, {m->xx0_vu, m->xy0_vu, m->yy0_vu}};
      for (int c = 0; c < nw; ++c) {
        const int o = t.perm_w[c];
        const size_t fo = (size_t)(o + 1) * 7;
        for (int j = 0; j < t.w_npoly[c]; ++j) {
          const size_t i = (size_t)c * NSLOT + j;
          const int v = in.iv(j, o);
          const int side = (m->w_dirv[fo + j] < 0.f) ? 0 : 1;
          for (int b = 0; b < 2; ++b) {
            for (int q = 0; q < 5; ++q) q3[b][q][i] = xy[b][(size_t)(o + 1) * 35 + (size_t)j * 5 + q];
            for (int q = 0; q < 3; ++q) {
              os3[b][q][i] = off[b][q][(size_t)(v + 1) * 2 + side];
              on3[b][q][i] = off[b][q][(size_t)(v + 1) * 2 + 1 - side];
            }
          }
        }
      }
    }
    // coefficient round trip (bitwise against the FP32 inputs)
    for (int c = 0; c < nw; ++c) {
      int o = t.perm_w[c];
      size_t fo = (size_t)(o + 1) * 7;
      for (int j = 0; j < t.w_npoly[c]; ++j) {
        size_t i = (size_t)c * NSLOT + j;
        if ((float)gxc[i] != m->gxps_coef[fo + j] || (float)gyc[i] != m->gyps_coef[fo + j] ||
            (float)dirv[i] != m->w_dirv[fo + j])
          throw Err(OLAMGPU_EINVAL, "renumbering check failed (coefficients)");
      }
    }
    std::vector<R> vn[3] = {std::vector<R>(nv), std::vector<R>(nv), std::vector<R>(nv)};
    for (int e = 0; e < nv; ++e) {
      int o = t.perm_v[e] + 1;
      vn[0][e] = m->vnx[o]; vn[1][e] = m->vny[o]; vn[2][e] = m->vnz[o];
    }
    std::vector<R> vfact(nm * 4, 0), xxyym(nm * 4, 0);
    for (int i = 0; i < nm; ++i) {
      int o = t.perm_m[i];
      R arm0 = m->arm0[o + 1];
      for (int j = 0; j < 3; ++j) {
        int vo = in.miv(j, o);
        R dnv = m->dnv[vo + 1];
        // vort_div_damp.f90:282-291
        vfact[i * 4 + j] = (in.vim(1, vo) == o) ? dnv / arm0 : -dnv / arm0;
        xxyym[i * 4 + j] = m->xx_yy_m[(size_t)(o + 1) * 3 + j];
      }
    }
    // per-level arrays (ldk entries: every lane may read its level)
    std::vector<R> lev[11];
    for (auto &v : lev) v.assign(ldk, R(0));
    std::vector<double> lev8[4];
    for (auto &v : lev8) v.assign(ldk, 0.0);
    for (int k = 1; k <= mza; ++k) {
      int l = k - 1;
      lev[0][l] = m->dzim[l];
      lev[1][l] = (R)m->gdz_wgtm8[l];     // real(gdz_wgtm8), mem_grid.f90:408-410
      lev[2][l] = (R)m->gdz_wgtp8[l];
      lev[3][l] = m->a_w[l + mza * (0 + 2 * 0)];    // a_w(k,1,1)
      lev[4][l] = m->a_w[l + mza * (1 + 2 * 0)];    // a_w(k,2,1)
      lev[5][l] = m->a_w[l + mza * (0 + 2 * 1)];    // a_w(k,1,2)
      lev[6][l] = m->a_w[l + mza * (1 + 2 * 1)];    // a_w(k,2,2)
      lev[7][l] = m->a_wu[l + mza * (0 + 2 * 0)];   // a_wu(k,1,1) (unbounded)
      lev[8][l] = m->a_wu[l + mza * (1 + 2 * 0)];   // a_wu(k,2,1)
      lev[9][l] = m->a_wu[l + mza * (0 + 2 * 1)];   // a_wu(k,1,2)
      lev[10][l] = m->a_wu[l + mza * (1 + 2 * 1)];  // a_wu(k,2,2)
      lev8[0][l] = m->zwgt_top8[l]; lev8[1][l] = m->zwgt_bot8[l];
      lev8[2][l] = m->gdz_wgtm8[l]; lev8[3][l] = m->gdz_wgtp8[l];
    }
    // control volumes
    std::vector<R> arw(wsz, 0), volti(wsz, 0), volwi(wsz, 0), arv(vsz, 0), volvi(vsz, 0);
    std::vector<double> volt(wsz, 0);
    for (int c = 0; c < nw; ++c) {
      size_t fo = (size_t)(t.perm_w[c] + 1) * mza;
      for (int l = 0; l < mza; ++l) {
        size_t i = (size_t)c * ldk + l;
        arw[i] = m->arw[fo + l];
        volt[i] = m->volt[fo + l];
        if (!f64) { volti[i] = m->volti[fo + l]; volwi[i] = m->volwi[fo + l]; }
      }
      if (f64) {   // inverse volumes rebuilt from volt in double (see olamgpu.h)
        for (int k = 2; k <= mza; ++k) volti[(size_t)c * ldk + k - 1] = (R)(1.0 / volt[(size_t)c * ldk + k - 1]);
        for (int k = 2; k <= mza - 1; ++k)
          volwi[(size_t)c * ldk + k - 1] = (R)(1.0 / (volt[(size_t)c * ldk + k - 1] + volt[(size_t)c * ldk + k]));
      }
    }
    for (int e = 0; e < nv; ++e) {
      size_t fo = (size_t)(t.perm_v[e] + 1) * mza;
      int w1 = t.v_iw[e * 2], w2 = t.v_iw[e * 2 + 1];
      for (int l = 0; l < mza; ++l) {
        size_t i = (size_t)e * ldk + l;
        arv[i] = m->arv[fo + l];
        if (!f64) volvi[i] = m->volvi[fo + l];
        else if (l >= 1) volvi[i] = (R)(1.0 / (volt[(size_t)w1 * ldk + l] + volt[(size_t)w2 * ldk + l]));
      }
    }
    upv(b_npoly, t.w_npoly); upv(b_nbr, t.w_nbr); upv(b_face, t.w_face); upv(b_back, t.w_back);
    upv(b_viw, t.v_iw); upv(b_vim, t.v_im); upv(b_miv, t.m_iv); upv(b_mim, t.m_im); upv(b_invw, t.inv_w); upv(b_invv, t.inv_v);
    upv(b_dirv, dirv); upv(b_gxc, gxc); upv(b_gyc, gyc); upv(b_dxs, dxs); upv(b_dys, dys);
    upv(b_dxn, dxn); upv(b_dyn, dyn); upv(b_ecx, ecx); upv(b_ecy, ecy); upv(b_ecz, ecz); upv(b_xxyy, xxyy);
    for (int i = 0; i < 3; ++i) { upv(b_wn[i], wn[i]); upv(b_vn[i], vn[i]); }
    upv(b_arw0, aw0);
    {
      std::vector<R> zf(ldk, R(0)), dn(nv, R(0));
      for (int k = 1; k <= mza; ++k) zf[k - 1] = (R)m->zfacit[k - 1];
      for (int e = 0; e < nv; ++e) dn[e] = (R)m->dniv[t.perm_v[e] + 1];
      upv(b_zfacit, zf);
      upv(b_dniv, dn);
    }
    if (m->dzit_bot) {
      std::vector<R> db(ldk, R(0));
      for (int k = 1; k <= mza; ++k) db[k - 1] = (R)m->dzit_bot[k - 1];
      upv(tb_dzitb, db);
    }
    upv(b_vfact, vfact); upv(b_xxyym, xxyym);
    for (int i = 0; i < 11; ++i) upv(b_lev[i], lev[i]);
    for (int i = 0; i < 4; ++i) upv(b_lev8[i], lev8[i]);
    upv(b_arw, arw); upv(b_volti, volti); upv(b_volwi, volwi); upv(b_volt, volt); upv(b_arv, arv); upv(b_volvi, volvi);
    for (int b = 0; b < 2; ++b) {
      for (int q = 0; q < 5; ++q) upv(b_q[b][q], q3[b][q]);
      for (int q = 0; q < 3; ++q) { upv(b_os[b][q], os3[b][q]); upv(b_on[b][q], on3[b][q]); }
    }
    // device copies of the integer tables must equal the host tables
    verify_dev(b_nbr, t.w_nbr); verify_dev(b_face, t.w_face); verify_dev(b_viw, t.v_iw); verify_dev(b_miv, t.m_iv);

    dm.nw = nw; dm.nv = nv; dm.nm = nm; dm.mza = mza; dm.ldk = ldk;
    dm.tall = mza > WAVE ? 1 : 0;
    dm.w_npoly = dp<int>(b_npoly); dm.w_nbr = dp<int>(b_nbr); dm.w_face = dp<int>(b_face); dm.w_back = dp<int>(b_back);
    dm.w_dirv = dp<R>(b_dirv); dm.w_gxc = dp<R>(b_gxc); dm.w_gyc = dp<R>(b_gyc);
    dm.w_dxs = dp<R>(b_dxs); dm.w_dys = dp<R>(b_dys); dm.w_dxn = dp<R>(b_dxn); dm.w_dyn = dp<R>(b_dyn);
    dm.w_ecx = dp<R>(b_ecx); dm.w_ecy = dp<R>(b_ecy); dm.w_ecz = dp<R>(b_ecz); dm.w_xxyy = dp<R>(b_xxyy);
    for (int b = 0; b < 2; ++b) {
      for (int q = 0; q < 5; ++q) dm.w_q[b][q] = dp<R>(b_q[b][q]);
      for (int q = 0; q < 3; ++q) { dm.w_os[b][q] = dp<R>(b_os[b][q]); dm.w_on[b][q] = dp<R>(b_on[b][q]); }
    }
    dm.wnx = dp<R>(b_wn[0]); dm.wny = dp<R>(b_wn[1]); dm.wnz = dp<R>(b_wn[2]); dm.arw0 = dp<R>(b_arw0);
    dm.zfacit = dp<R>(b_zfacit); dm.dniv = dp<R>(b_dniv);
    dm.v_iw = dp<int>(b_viw); dm.v_im = dp<int>(b_vim);
    dm.vnx = dp<R>(b_vn[0]); dm.vny = dp<R>(b_vn[1]); dm.vnz = dp<R>(b_vn[2]);
    dm.m_iv = dp<int>(b_miv); dm.m_im = dp<int>(b_mim); dm.m_vfact = dp<R>(b_vfact); dm.m_xxyym = dp<R>(b_xxyym);
    dm.dzim = dp<R>(b_lev[0]); dm.gdz_wgtm = dp<R>(b_lev[1]); dm.gdz_wgtp = dp<R>(b_lev[2]);
    dm.aw11 = dp<R>(b_lev[3]); dm.aw21 = dp<R>(b_lev[4]); dm.aw12 = dp<R>(b_lev[5]); dm.aw22 = dp<R>(b_lev[6]);
    dm.awu11 = dp<R>(b_lev[7]); dm.awu21 = dp<R>(b_lev[8]); dm.awu12 = dp<R>(b_lev[9]); dm.awu22 = dp<R>(b_lev[10]);
    dm.zwgt_top8 = dp<double>(b_lev8[0]); dm.zwgt_bot8 = dp<double>(b_lev8[1]);
    dm.gdz_wgtm8 = dp<double>(b_lev8[2]); dm.gdz_wgtp8 = dp<double>(b_lev8[3]);
    dm.arw = dp<R>(b_arw); dm.volti = dp<R>(b_volti); dm.volwi = dp<R>(b_volwi); dm.arv = dp<R>(b_arv);
    dm.volvi = dp<R>(b_volvi); dm.volt = dp<double>(b_volt);

    // Rayleigh friction profiles and b20 (mem_rayf.f90:41-156, prog_wrtv_rk.f90:46-52)
    rf = rayf_profiles(prm, mza, m->zm, m->zt);
    if (rf.dorayf && !(m->th01d && m->dn01d))
      throw Err(OLAMGPU_EINVAL, "rayf_fact > 0 with initial == 1 needs the th01d/dn01d reference sounding");
    {
      std::vector<R> lv((size_t)5 * ldk, R(0));
      for (int k = 1; k <= mza; ++k) {
        lv[0 * ldk + k - 1] = (R)rf.cof[k - 1];
        lv[1 * ldk + k - 1] = (R)rf.cofmix[k - 1];
        lv[2 * ldk + k - 1] = rf.dorayf ? (R)m->th01d[k - 1] : R(0);
        lv[3 * ldk + k - 1] = rf.dorayf ? (R)m->dn01d[k - 1] : R(0);
        lv[4 * ldk + k - 1] = (R)rf.b20[k - 1];
      }
      upv(b_rayf, lv);
      R *base = dp<R>(b_rayf);
      dm.rayf_cof = rf.dorayf ? base : nullptr;
      dm.rayf_cofmix = rf.dorayfmix ? base + ldk : nullptr;
      dm.th01d = base + 2 * ldk;
      dm.dn01d = base + 3 * ldk;
      dm.b20 = rf.dorayfw ? base + 4 * ldk : nullptr;
      dm.krayf_bot = rf.krayf_bot;
      dm.krayfmix_bot = rf.krayfmix_bot;
    }

    // damping coefficients (first-call SAVEd factors of vort_div_damp.f90:39-119, 271-328)
    const double dtm = prm.dtsm;
    const int dl = prm.divh_damp_level, vl = prm.vort_damp_level;
    do_div = prm.divh_damp_fact > 1.e-7f && dl > 0;
    do_vort = prm.vort_damp_fact >= 1.e-7f && vl > 0;
    // second-order (fdivl: level 1 and/or the Rayleigh profile) and higher-order (fdivh)
    // divergence damping factors, vort_div_damp.f90:68-119
    do_divl = rf.dorayfdiv || (do_div && dl == 1);
    do_divh = do_div && dl > 1;
    kdivl = (do_div && dl == 1) ? 2 : rf.krayfdiv_bot;
    if (do_divl || do_divh) {
      std::vector<R> fl(do_divl ? vsz : 0, R(0)), fh(do_divh ? vsz : 0, R(0));
      const double facth = do_divh ? -(double)ipow((R)(-prm.divh_damp_fact), dl) / dtm : 0.0;
      const R fact2 = (do_div && dl == 1) ? (R)(prm.divh_damp_fact / (R)dtm) : R(0);
      R factr = R(0);
      if (rf.dorayfdiv && do_div) factr = (R)(std::max(prm.divh_damp_fact, prm.rayfdiv_fact) / (R)dtm);
      else if (rf.dorayfdiv) factr = (R)(prm.rayfdiv_fact / (R)dtm);
      for (int e = 0; e < nv; ++e) {
        int o = t.perm_v[e] + 1;
        int w1o = in.viw(0, o - 1) + 1, w2o = in.viw(1, o - 1) + 1;
        const double factb = do_divh ? facth * (double)ipow((R)std::min(m->arw0[w1o], m->arw0[w2o]), dl - 1) : 0.0;
        for (int k = 2; k <= mza; ++k) {
          const double vmin = std::min(m->volt[(size_t)w1o * mza + k - 1], m->volt[(size_t)w2o * mza + k - 1]);
          const R zf = (R)rf.zf[k - 1];
          if (do_divl && k >= kdivl)
            fl[(size_t)e * ldk + k - 1] = (R)((double)((fact2 * (R(1) - zf) + factr * zf) * (R)m->dniv[o] *
                                                       (R)m->zfacit[k - 1] * (R)m->dzit[k - 1]) * vmin);
          if (do_divh)
            fh[(size_t)e * ldk + k - 1] = (R)(factb * m->dniv[o] * m->zfacit[k - 1] * m->dzit[k - 1] * vmin *
                                              (double)(R(1) - zf));
        }
      }
      if (do_divl) upv(fdivl, fl);
      if (do_divh) upv(fdivh, fh);
      // the divergence and its Laplacians; a third buffer keeps the divergence for
      // fdivl when two Laplacians follow it (level 3 with the Rayleigh profile)
      const int nb = (do_divl && dl == 3 && do_div) ? 3 : 2;
      for (int i = 0; i < nb; ++i) {
        div2d[i].alloc(wsz * sizeof(R));
        HIPCHK(hipMemsetAsync(div2d[i].p, 0, div2d[i].bytes, s));
      }
    }
    if (do_vort) {
      std::vector<R> f(nv, R(0));
      for (int e = 0; e < nv; ++e) {
        int o = t.perm_v[e] + 1;
        int m1 = in.vim(0, o - 1) + 1, m2 = in.vim(1, o - 1) + 1;
        R a = std::min((R)m->arm0[m1], (R)m->arm0[m2]);
        f[e] = R(0.5) * (R)m->dniu[o] / (R)dtm * ipow(-(R)prm.vort_damp_fact * a, vl);
      }
      upv(fvorth, f);
      vort2d[0].alloc(msz * sizeof(R)); vort2d[1].alloc(msz * sizeof(R));
      HIPCHK(hipMemsetAsync(vort2d[0].p, 0, vort2d[0].bytes, s));
      HIPCHK(hipMemsetAsync(vort2d[1].p, 0, vort2d[1].bytes, s));
    }
    HIPCHK(hipStreamSynchronize(s));
  }

  template <class T> void verify_dev(DevBuf &b, const std::vector<T> &h) {
    std::vector<T> back(h.size());
    HIPCHK(hipMemcpyAsync(back.data(), b.p, h.size() * sizeof(T), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (std::memcmp(back.data(), h.data(), h.size() * sizeof(T)) != 0)
      throw Err(OLAMGPU_EHIP, "device copy of a neighbour table differs from the host table");
  }

  void zero(DevBuf &b) { HIPCHK(hipMemsetAsync(b.p, 0, b.bytes, s)); }

  void alloc_state() {
    for (DevBuf *b : {&thil[0], &thil[1], &wmc, &wc, &vxe, &vye, &vze, &wmasc, &alpha, &pwfac, &thilt, &vmxet,
                      &vmyet, &vmzet, &rho_old, &wmc0, &thilt_s, &vmxet_s, &vmyet_s, &vmzet_s, &vmxet_rk,
                      &vmyet_rk, &vmzet_rk}) { b->alloc(wsz * sizeof(R)); zero(*b); }
    for (DevBuf *b : {&vmc, &vc, &vmasc, &pvfac, &vmt, &vmc0, &vmt_s}) { b->alloc(vsz * sizeof(R)); zero(*b); }
    for (DevBuf *b : {&rho, &press, &rho0, &rth0}) { b->alloc(wsz * sizeof(RP)); zero(*b); }
    for (auto &g : grads) { g.alloc(wsz * sizeof(R)); zero(g); }
    if (prm.thil_horiz_adv_order == 3 || prm.wind_horiz_adv_order == 3)
      for (auto &gv : gradq)
        for (auto &g : gv) { g.alloc(wsz * sizeof(R)); zero(g); }
    mono_on[0] = prm.ithil_monot > 0;
    mono_on[1] = mono_on[2] = mono_on[3] = prm.iwind_monot > 0;
    for (int q = 0; q < 4; ++q)
      if (mono_on[q])
        for (DevBuf *b : {&mq0[q], &mscpt[q], &ms1[q], &ms2[q]}) { b->alloc(wsz * sizeof(R)); zero(*b); }
    ntr = prm.num_scalar;
    tstride = (long)wsz;
    size_t tb = std::max<size_t>(1, (size_t)ntr) * wsz * sizeof(R);
    for (DevBuf *b : {&trS, &trX, &trY}) { b->alloc(tb); zero(*b); }
    trG.alloc(2 * tb); zero(trG);
    if (prm.scal_horiz_adv_order == 3)
      for (DevBuf &b : trgq) { b.alloc(tb); zero(b); }
    pdef.assign(std::max(ntr, 1), 1);
    sgsmix.assign(std::max(ntr, 1), 1);
    sgs_dev.alloc(std::max(ntr, 1) * sizeof(int));
    if (prm.iscal_monot > 0) {
      for (DevBuf *b : {&lim_s1, &lim_s2}) { b->alloc(tb); zero(*b); }
      lim_mode.alloc(std::max(ntr, 1) * sizeof(int)); lim_pre.alloc(std::max(ntr, 1) * sizeof(int));
      if (prm.adjoint_tape > 0) {
        lim_tsv.alloc(std::max<size_t>(1, (size_t)ntr) * vsz * sizeof(R)); zero(lim_tsv);
        lim_tsw.alloc(tb); zero(lim_tsw);
      }
    }
    ffimp.alloc((size_t)3 * nw * sizeof(R)); zero(ffimp);
    cflmax.alloc(4 * sizeof(unsigned)); nimp.alloc(sizeof(int)); anyimp.alloc(sizeof(int));
    masspart.alloc((size_t)nblk(nw) * 65 * sizeof(double));
    massout.alloc(65 * sizeof(double));
    mmkeys.alloc(2 * 64 * sizeof(int)); mmout.alloc(2 * 64 * sizeof(float));
    HIPCHK(hipStreamSynchronize(s));
  }

  void alloc_adj() {
    if (adjA.p) return;
    size_t tb = std::max<size_t>(1, (size_t)ntr) * wsz * sizeof(R);
    for (DevBuf *b : {&adjA, &adjB, &lam0, &nu, &lpart}) { b->alloc(tb); zero(*b); }
    lgG.alloc(2 * tb); zero(lgG);
    if (prm.scal_horiz_adv_order == 3)
      for (DevBuf &b : lgq) { b.alloc(tb); zero(b); }
    HIPCHK(hipStreamSynchronize(s));
  }

  // -------------------------------------------------------------- transfers
  // Fortran (mza, n + 1) host array with dummy column <-> device (n, ldk) renumbered,
  // repacked on the device
  HostXfer xfer;
  DevBuf dperm_w, dperm_v, dinv_w, dinv_v;
  const int *dev_perm(const std::vector<int> &perm, bool inverse) {
    const bool w = &perm == &t.perm_w;
    if (!w && &perm != &t.perm_v) throw Err(OLAMGPU_EINVAL, "transfer with an unknown numbering");
    DevBuf &p = w ? dperm_w : dperm_v, &q = w ? dinv_w : dinv_v;
    if (!p.p) { upv(p, perm); upv(q, w ? t.inv_w : t.inv_v); }
    return (inverse ? q : p).as<int>();
  }
  template <class D, class H> void h2d_at(void *dst, const H *h, const std::vector<int> &perm) {
    const int n = (int)perm.size();
    xfer.up(h, (size_t)(n + 1) * mza * sizeof(H), s);
    const size_t cnt = (size_t)n * ldk;
    hipLaunchKernelGGL((k_unpack<D, H>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
                       (D *)dst, xfer.stage.as<const H>(), dev_perm(perm, false), n, mza, ldk);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
  }
  template <class D, class H> void h2d(DevBuf &b, const H *h, const std::vector<int> &perm) {
    h2d_at<D>(b.p, h, perm);
  }
  template <class D, class H> void h2d_off(DevBuf &b, size_t off, const H *h, const std::vector<int> &perm) {
    h2d_at<D>((char *)b.p + off * sizeof(D), h, perm);
  }
  template <class D, class H> void d2h(const DevBuf &b, H *h, const std::vector<int> &perm, size_t off = 0) {
    const int n = (int)perm.size();
    const size_t cnt = (size_t)(n + 1) * mza;
    xfer.ready(cnt * sizeof(H), s);
    hipLaunchKernelGGL((k_pack<D, H>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
                       xfer.stage.as<H>(), (const D *)((const char *)b.p + off * sizeof(D)),
                       dev_perm(perm, true), n, mza, ldk);
    HIPCHK(hipGetLastError());
    xfer.down(h, cnt * sizeof(H), s);
  }

  DevBuf &cur_thil_buf() { return thil[cur_thil]; }

  void upload_state(const olamgpu_state *h) override {
    cfl_fresh = false;
    auto W = [&](DevBuf &b, const float *p) { if (p) h2d<R>(b, p, t.perm_w); };
    auto V = [&](DevBuf &b, const float *p) { if (p) h2d<R>(b, p, t.perm_v); };
    W(cur_thil_buf(), h->thil); W(wmc, h->wmc); W(wc, h->wc); W(vxe, h->vxe); W(vye, h->vye); W(vze, h->vze);
    if (h->rho) h2d<RP>(rho, h->rho, t.perm_w);
    if (h->press) h2d<RP>(press, h->press, t.perm_w);
    V(vmc, h->vmc); V(vc, h->vc); W(wmasc, h->wmasc); V(vmasc, h->vmasc);
    W(alpha, h->alpha_press); W(pwfac, h->pwfac); V(pvfac, h->pvfac);
    W(thilt, h->thilt); W(vmxet, h->vmxet); W(vmyet, h->vmyet); W(vmzet, h->vmzet); V(vmt, h->vmt);
  }
  void download_state(olamgpu_state *h) override {
    auto W = [&](DevBuf &b, float *p) { if (p) d2h<R>(b, p, t.perm_w); };
    auto V = [&](DevBuf &b, float *p) { if (p) d2h<R>(b, p, t.perm_v); };
    W(cur_thil_buf(), h->thil); W(wmc, h->wmc); W(wc, h->wc); W(vxe, h->vxe); W(vye, h->vye); W(vze, h->vze);
    if (h->rho) d2h<RP>(rho, h->rho, t.perm_w);
    if (h->press) d2h<RP>(press, h->press, t.perm_w);
    V(vmc, h->vmc); V(vc, h->vc); W(wmasc, h->wmasc); V(vmasc, h->vmasc);
    W(alpha, h->alpha_press); W(pwfac, h->pwfac); V(pvfac, h->pvfac);
    W(thilt, h->thilt); W(vmxet, h->vmxet); W(vmyet, h->vmyet); W(vmzet, h->vmzet); V(vmt, h->vmt);
  }
  void check_n(int n) { if (n < 1 || n > ntr) throw Err(OLAMGPU_EINVAL, "scalar index out of range"); }
  void upload_scalar(int n, const float *p, const float *tt) override {
    check_n(n);
    if (p) h2d_off<R>(trS, (size_t)(n - 1) * tstride, p, t.perm_w);
    if (tt) {
      if (!trsct.p) { trsct.alloc(std::max<size_t>(1, ntr) * wsz * sizeof(R)); zero(trsct); }
      h2d_off<R>(trsct, (size_t)(n - 1) * tstride, tt, t.perm_w);
      have_sct = true;
    } else if (have_sct) {
      HIPCHK(hipMemsetAsync((char *)trsct.p + (size_t)(n - 1) * tstride * sizeof(R), 0, wsz * sizeof(R), s));
    }
  }
  void download_scalar(int n, float *p) override { check_n(n); d2h<R>(trS, p, t.perm_w, (size_t)(n - 1) * tstride); }
  DevBuf &cur_adj_buf() { return cur_adj ? adjB : adjA; }
  void upload_adj(int n, const float *p) override {
    check_n(n); alloc_adj();
    h2d_off<R>(cur_adj_buf(), (size_t)(n - 1) * tstride, p, t.perm_w);
  }
  void download_adj(int n, float *p) override {
    check_n(n); alloc_adj();
    d2h<R>(cur_adj_buf(), p, t.perm_w, (size_t)(n - 1) * tstride);
  }

  // -------------------------------------------------------------- tend0 etc.
  void tend0() override {
    cfl_fresh = false;
    HIPCHK(hipMemsetAsync(vmasc.p, 0, vmasc.bytes, s));
    HIPCHK(hipMemsetAsync(wmasc.p, 0, wmasc.bytes, s));
    if (have_sct) HIPCHK(hipMemsetAsync(trsct.p, 0, trsct.bytes, s));
    copy_rho_old();
  }
  void copy_rho_old();
  void timeavg_momsc() override;

  // -------------------------------------------------------------- prog_wrtv_rk
  // mem_nudge's wmanud, vmanud, rhot_nud (host Fortran layout; NULL keeps a field)
  void set_nudging(const float *wm, const float *vm, const double *rt) override {
    if (!nud_w.p) alloc_nudging();
    if (wm) h2d<R>(nud_w, wm, t.perm_w);
    if (vm) h2d<R>(nud_v, vm, t.perm_v);
    if (rt) h2d<RP>(nud_r, rt, t.perm_w);
  }
  void alloc_nudging() {
    nud_w.alloc(wsz * sizeof(R)); nud_v.alloc(vsz * sizeof(R)); nud_r.alloc(wsz * sizeof(RP));
    zero(nud_w); zero(nud_v); zero(nud_r);
    HIPCHK(hipStreamSynchronize(s));
  }

  // the PBL scheme's mem_turb fields, then comp_horiz_k (pbl_driver.f90:136-238)
  void set_turbulence(const float *vkh, const float *vkm, const int32_t *khtop, const int32_t *kmtop,
                      const float *akm_sfc) override {
    if (!vkh || !vkm || !khtop || !kmtop || !akm_sfc)
      throw Err(OLAMGPU_EINVAL, "set_turbulence needs vkh, vkm, khtop, kmtop and akm_sfc");
    if (!tb_dzitb.p) throw Err(OLAMGPU_EINVAL, "set_turbulence needs dzit_bot in the mesh descriptor");
    std::vector<int> kh(nw), km(nw);
    std::vector<R> as(nw);
    for (int c = 0; c < nw; ++c) {
      const int o = t.perm_w[c] + 1;
      kh[c] = khtop[o]; km[c] = kmtop[o]; as[c] = (R)akm_sfc[o];
      if (kh[c] < 0 || kh[c] > mza - 1 || km[c] < 0 || km[c] > mza - 1)
        throw Err(OLAMGPU_EINVAL, "khtop and kmtop must lie in [0, mza - 1]");
    }
    if (!tb_vkh.p) {
      tb_vkh.alloc(wsz * sizeof(R)); tb_vkm.alloc(wsz * sizeof(R));
      tb_akhodx.alloc(vsz * sizeof(R)); tb_akmodx.alloc(vsz * sizeof(R));
      tb_khtopv.alloc(nv * sizeof(int)); tb_kmtopv.alloc(nv * sizeof(int));
    }
    h2d<R>(tb_vkh, vkh, t.perm_w);
    h2d<R>(tb_vkm, vkm, t.perm_w);
    upv(tb_khtop, kh); upv(tb_kmtop, km); upv(tb_akm, as);
    DTurb<R> d{};
    d.vkh = dp<R>(tb_vkh); d.vkm = dp<R>(tb_vkm); d.akm_sfc = dp<R>(tb_akm); d.dzit_bot = dp<R>(tb_dzitb);
    d.khtop = dp<int>(tb_khtop); d.kmtop = dp<int>(tb_kmtop);
    d.akhodx = dp<R>(tb_akhodx); d.akmodx = dp<R>(tb_akmodx);
    d.khtopv = dp<int>(tb_khtopv); d.kmtopv = dp<int>(tb_kmtopv);
    d.on = 1;
    LAUNCH("k_horiz_k", (k_horiz_k<R, RP>), dim3(nblk(nv)), dim3(256), dm, d, dp<RP>(rho), (R)prm.dtlm,
           (R)prm.akmin);
    HIPCHK(hipStreamSynchronize(s));
    dturb = d;
  }

  // comp_alpha_press (timestep.f90:500-559) from host moisture fields
  DevBuf moist[3];
  void comp_alpha_press(const float *rr_v, const float *rr_w, const float *theta) override {
    if (!rr_v || !rr_w || !theta) throw Err(OLAMGPU_EINVAL, "comp_alpha_press needs rr_v, rr_w and theta");
    const float *h[3] = {rr_v, rr_w, theta};
    for (int i = 0; i < 3; ++i) {
      if (!moist[i].p) moist[i].alloc(wsz * sizeof(R));
      h2d<R>(moist[i], h[i], t.perm_w);
    }
    // pc1 = p00i**rocv (consts_coms.f90:76)
    const R pc1 = sizeof(R) == 4 ? (R)powr(1.0f / 1.0e5f, 287.0f / 717.0f) : (R)std::pow(1.0 / 1.0e5, 287.0 / 717.0);
    LAUNCH("k_alpha_w", (k_alpha_w<R>), dim3(nblk(nw)), dim3(256), dm, dp<R>(moist[0]), dp<R>(moist[1]),
           dp<R>(moist[2]), dp<R>(cur_thil_buf()), dp<R>(alpha), dp<R>(pwfac), pc1);
    LAUNCH("k_alpha_v", (k_alpha_v<R>), dim3(nblk(nv)), dim3(256), dm, dp<R>(moist[1]), dp<R>(pvfac));
    HIPCHK(hipStreamSynchronize(s));
  }

  // k_dyn_w instance for the stage: monotone fluxes, nudging, order-3 fields
  // (tall columns: the instances whose vertical moves cross the two waves of a column)
  template <bool MONO, bool NUD, bool Q3> void launch_dyn_w_inst(const DynBuf<R, RP> &b, unsigned gw) {
    if (dm.tall) LAUNCH("k_dyn_w", (k_dyn_w<R, RP, MONO, NUD, Q3, true>), dim3(gw), dim3(256), dm, b);
    else LAUNCH("k_dyn_w", (k_dyn_w<R, RP, MONO, NUD, Q3>), dim3(gw), dim3(256), dm, b);
  }
  template <bool MONO> void launch_dyn_w(const DynBuf<R, RP> &b, bool nud, unsigned gw) {
    const bool q3 = b.thil_order == 3 || b.wind_order == 3;
    if (nud) {
      if (q3) launch_dyn_w_inst<MONO, true, true>(b, gw);
      else launch_dyn_w_inst<MONO, true, false>(b, gw);
    } else {
      if (q3) launch_dyn_w_inst<MONO, false, true>(b, gw);
      else launch_dyn_w_inst<MONO, false, false>(b, gw);
    }
  }

  // rayf_init's copy of the initial wind (mem_rayf.f90:158-171); the profiles and b20
  // were built with the mesh
  void init_wrtv_rk() override {
    if (rf.dorayf) {
      if (!vc03d.p) vc03d.alloc(vsz * sizeof(R));
      HIPCHK(hipMemcpyAsync(vc03d.p, vc.p, vsz * sizeof(R), hipMemcpyDeviceToDevice, s));
      HIPCHK(hipStreamSynchronize(s));
    }
    rayf_ready = true;
  }

  void prog_wrtv_rk() override {
    if (rf.dorayf && !rayf_ready)
      throw Err(OLAMGPU_EINVAL, "Rayleigh friction on thil/vmc: call olamgpu_init_wrtv_rk after the initial state upload");
    cfl_fresh = false;
    const unsigned gw = nblk(nw), gv = nblk(nv), gm = nblk(nm);
    LAUNCH("k_dyn_setup_w", (k_dyn_setup_w<R, RP>), dim3(gw), dim3(256), dm, dturb, (R)prm.dtsm,
                       dp<R>(cur_thil_buf()), dp<RP>(rho),
                       dp<R>(wmc), dp<R>(thilt), dp<R>(vmxet), dp<R>(vmyet), dp<R>(vmzet), dp<RP>(rth0),
                       dp<RP>(rho0), dp<R>(wmc0), dp<R>(thilt_s), dp<R>(vmxet_s), dp<R>(vmyet_s), dp<R>(vmzet_s),
                       dp<R>(vxe), dp<R>(vye), dp<R>(vze));
    // RK-start copies of the monotone-flux fields (prog_wrtv_rk.f90:188-199)
    {
      DevBuf *src[4] = {&cur_thil_buf(), &vxe, &vye, &vze};
      for (int q = 0; q < 4; ++q)
        if (mono_on[q]) HIPCHK(hipMemcpyAsync(mq0[q].p, src[q]->p, wsz * sizeof(R), hipMemcpyDeviceToDevice, s));
    }
    // short-step damping: divergence at W and vorticity at M, then their Laplacians.
    // div2d[0] plays the reference's SAVEd div2d(:,:,1), onto which each call adds the
    // new divergence; with three buffers the divergence itself goes to div2d[2] so that
    // fdivl still sees it after the two Laplacians have overwritten div2d[0].
    if (do_divl || do_divh) {
      const int dst = div2d[2].p ? 2 : 0;
      if (OLAMGPU_DAMP_ITEMS > 1)
        LAUNCH("k_div", (k_divn<R, OLAMGPU_DAMP_ITEMS>), dim3(nblk_per(nw, OLAMGPU_DAMP_ITEMS)), dim3(256), dm,
               dp<R>(vmc), dp<R>(div2d[0]), dp<R>(div2d[dst]));
      else LAUNCH("k_div", (k_div<R>), dim3(gw), dim3(256), dm, dp<R>(vmc), dp<R>(div2d[0]), dp<R>(div2d[dst]));
      div_l = dst;
      int ih = dst;
      if (do_divh) {
        // Laplacians ping-pong between the two other buffers and end in div2d[0] for
        // level 3 (vort_div_damp.f90:196-212 with il/ih = 1,2 then 2,1)
        const int order[3] = {dst, 1, 0};
        for (int nn = 1; nn <= prm.divh_damp_level - 1; ++nn) {
          const int il = order[nn - 1], io = order[nn];
          if (OLAMGPU_DAMP_ITEMS > 1)
            LAUNCH("k_lap_w", (k_lap_wn<R, OLAMGPU_DAMP_ITEMS>), dim3(nblk_per(nw, OLAMGPU_DAMP_ITEMS)), dim3(256), dm,
                   dp<R>(div2d[il]), dp<R>(div2d[io]));
          else LAUNCH("k_lap_w", (k_lap_w<R>), dim3(gw), dim3(256), dm, dp<R>(div2d[il]), dp<R>(div2d[io]));
          ih = io;
        }
      }
      div_ih = ih;
    }
    if (do_vort) {
      if (OLAMGPU_DAMP_ITEMS > 1)
        LAUNCH("k_vort", (k_vortn<R, OLAMGPU_DAMP_ITEMS>), dim3(nblk_per(nm, OLAMGPU_DAMP_ITEMS)), dim3(256), dm,
               dp<R>(vc), dp<R>(vort2d[0]));
      else LAUNCH("k_vort", (k_vort<R>), dim3(gm), dim3(256), dm, dp<R>(vc), dp<R>(vort2d[0]));
      int ih = 0;
      for (int nn = 1; nn <= prm.vort_damp_level - 1; ++nn) {
        int il = (nn - 1) % 2;
        ih = nn % 2;
        if (OLAMGPU_DAMP_ITEMS > 1)
          LAUNCH("k_lap_m", (k_lap_mn<R, OLAMGPU_DAMP_ITEMS>), dim3(nblk_per(nm, OLAMGPU_DAMP_ITEMS)), dim3(256), dm,
                 dp<R>(vort2d[il]), dp<R>(vort2d[ih]));
        else LAUNCH("k_lap_m", (k_lap_m<R>), dim3(gm), dim3(256), dm, dp<R>(vort2d[il]), dp<R>(vort2d[ih]));
      }
      vort_ih = ih;
    }
    {
      R *dl = do_divl ? dp<R>(div2d[div_l]) : (R *)nullptr, *dh = do_divh ? dp<R>(div2d[div_ih]) : (R *)nullptr;
      R *vo = do_vort ? dp<R>(vort2d[vort_ih]) : (R *)nullptr;
      // several faces per wave, their loads issued together
      LAUNCH("k_dyn_setup_v", (k_dyn_setup_vn<R, RP, OLAMGPU_SETUPV_ITEMS>), dim3(nblk_per(nv, OLAMGPU_SETUPV_ITEMS)),
             dim3(256), dm, dp<R>(vmc), dp<R>(vmt), dp<R>(vmc0), dp<R>(vmt_s), dl, dp<R>(fdivl), kdivl, dh,
             dp<R>(fdivh), vo, dp<R>(fvorth), dp<RP>(rho), dp<R>(vc), dp<R>(vc03d));
    }
    const int nrk = prm.nrk_wrtv;
    for (int istage = 1; istage <= nrk; ++istage) {
      double frac = (nrk <= 2) ? (istage == 1 ? 0.5 : 1.0)
                               : (istage == 1 ? 1.0 / 3.0 : (istage == 2 ? 0.5 : 1.0));
      DynBuf<R, RP> b;
      b.thil_in = dp<R>(thil[cur_thil]);
      b.thil_out = dp<R>(thil[1 - cur_thil]);
      b.wmc = dp<R>(wmc); b.wc = dp<R>(wc); b.vxe = dp<R>(vxe); b.vye = dp<R>(vye); b.vze = dp<R>(vze);
      b.vmc = dp<R>(vmc); b.vc = dp<R>(vc); b.vmasc = dp<R>(vmasc); b.wmasc = dp<R>(wmasc);
      b.rho = dp<RP>(rho); b.press = dp<RP>(press);
      b.alpha_press = dp<R>(alpha); b.pwfac = dp<R>(pwfac); b.pvfac = dp<R>(pvfac);
      b.rho0 = dp<RP>(rho0); b.rth0 = dp<RP>(rth0); b.wmc0 = dp<R>(wmc0); b.vmc0 = dp<R>(vmc0);
      b.thilt_s = dp<R>(thilt_s); b.vmxet_s = dp<R>(vmxet_s); b.vmyet_s = dp<R>(vmyet_s); b.vmzet_s = dp<R>(vmzet_s);
      b.vmt_s = dp<R>(vmt_s);
      b.vmxet_rk = dp<R>(vmxet_rk); b.vmyet_rk = dp<R>(vmyet_rk); b.vmzet_rk = dp<R>(vmzet_rk);
      for (int i = 0; i < 8; ++i) b.g[i] = dp<R>(grads[i]);
      for (int q = 0; q < 4; ++q)
        for (int t = 0; t < 3; ++t) b.gq[q][t] = dp<R>(gradq[q][t]);
      b.thil_order = prm.thil_horiz_adv_order; b.wind_order = prm.wind_horiz_adv_order;
      b.istage = istage; b.nrk = nrk;
      b.dt8 = frac * prm.dtsm;
      b.fcor = (R)prm.fcoriol;
      b.wmanud = dp<R>(nud_w); b.vmanud = dp<R>(nud_v); b.rhot_nud = dp<RP>(nud_r);
      b.nud_vm = prm.nudflag > 0 && prm.nud_preserve_total_mass;
      b.nud_w = prm.nudflag == 1 && prm.nud_preserve_total_mass;
      b.nud_rho = prm.nudflag > 0 && !prm.nud_preserve_total_mass;
      const bool nud = b.nud_vm || b.nud_w || b.nud_rho;
      int dth = prm.thil_horiz_adv_order == 2, dwn = prm.wind_horiz_adv_order == 2;
      if (dth || dwn)
        LAUNCH("k_dyn_grad", (k_dyn_grad<R>), dim3(gw), dim3(256), dm, b.thil_in, b.vxe, b.vye, b.vze,
                           dp<R>(grads[0]), dp<R>(grads[1]), dp<R>(grads[2]), dp<R>(grads[3]), dp<R>(grads[4]),
                           dp<R>(grads[5]), dp<R>(grads[6]), dp<R>(grads[7]), dth, dwn);
      // the final stage runs fluxes_wrtv_monot for the monotone fields (:370-399)
      const bool last = istage == nrk;
      int any_mono = 0, bmask = 0;
      for (int q = 0; q < 4; ++q) {
        b.mono[q] = last && mono_on[q];
        any_mono |= b.mono[q];
        bmask |= b.mono[q] << q;
        b.scpt[q] = dp<R>(mscpt[q]); b.ms1[q] = dp<R>(ms1[q]); b.ms2[q] = dp<R>(ms2[q]);
      }
      const int qth = prm.thil_horiz_adv_order == 3, qwn = prm.wind_horiz_adv_order == 3;
      if (qth || qwn) {
        DynGrad3<R> o;
        for (int q = 0; q < 4; ++q) {
          o.g[q][0] = dp<R>(grads[2 * q]); o.g[q][1] = dp<R>(grads[2 * q + 1]);
          for (int t = 0; t < 3; ++t) o.g[q][2 + t] = dp<R>(gradq[q][t]);
        }
        LAUNCH("k_dyn_grad3", (k_dyn_grad3<R>), dim3(gw), dim3(256), dm, b.thil_in, b.vxe, b.vye, b.vze, o, qth,
               qwn, bmask);
      }
      if (any_mono) {
        DynMono<R, RP> mo;
        const R *sts[4] = {dp<R>(thilt_s), dp<R>(vmxet_s), dp<R>(vmyet_s), dp<R>(vmzet_s)};
        for (int q = 0; q < 4; ++q) {
          mo.q0[q] = dp<R>(mq0[q]); mo.st[q] = sts[q]; mo.scpt[q] = dp<R>(mscpt[q]);
          mo.s1[q] = dp<R>(ms1[q]); mo.s2[q] = dp<R>(ms2[q]); mo.on[q] = b.mono[q];
        }
        LAUNCH("k_dyn_monot_pre", (k_dyn_monot_pre<R, RP>), dim3(gw), dim3(256), dm, mo, dp<RP>(rho0), (R)b.dt8);
        LAUNCH("k_dyn_monot_scales", (k_dyn_monot_scales<R, RP>), dim3(gw), dim3(256), dm, b, mo);
        launch_dyn_w<true>(b, nud, gw);
      } else {
        launch_dyn_w<false>(b, nud, gw);
      }
      cur_thil = 1 - cur_thil;
      // several faces or columns per wave, their loads issued together
      constexpr int NV = OLAMGPU_V_ITEMS, NVEL = OLAMGPU_VEL_ITEMS;
      LAUNCH("k_dyn_v", (k_dyn_vn<R, RP, NV>), dim3(nblk_per(nv, NV)), dim3(256), dm, b, istage == nrk - 1 ? 1 : 0);
      LAUNCH("k_dyn_vel", (k_dyn_veln<R, RP, NVEL>), dim3(nblk_per(nw, NVEL)), dim3(256), dm, b);
    }
    HIPCHK(hipGetLastError());
  }

  // -------------------------------------------------------------- scalar_transport_rk
  void stage_dts(R *dtlr, int nrk) const {
    R dtl = (R)prm.dtlm;
    const R one3 = R(1.) / R(3.);
    if (nrk == 3) { dtlr[0] = dtl * one3; dtlr[1] = dtl * R(0.5); dtlr[2] = dtl * R(1.0); }
    else { dtlr[0] = dtl * R(0.5); dtlr[1] = dtl * R(1.0); }
  }

  // check_cfls: implicit fractions (ffimp) for this step; returns whether any column
  // needs the implicit branch (one small device-to-host copy).
  bool run_cfl(bool report) {
    const int nrk = prm.nrk_scal;
    HIPCHK(hipMemsetAsync(anyimp.p, 0, anyimp.bytes, s));
    if (report) {
      HIPCHK(hipMemsetAsync(cflmax.p, 0, cflmax.bytes, s));
      HIPCHK(hipMemsetAsync(nimp.p, 0, nimp.bytes, s));
    }
    LAUNCH("k_tr_cfl", (k_tr_cfl<R>), dim3(nblk(nw)), dim3(256), dm, dp<R>(vmasc), dp<R>(wmasc),
           dp<R>(rho_old), dp<R>(wc), (R)prm.dtlm, nrk, dp<R>(ffimp), cflmax.as<unsigned>(), nimp.as<int>(),
           anyimp.as<int>(), report ? 1 : 0);
    int any = 0;
    HIPCHK(hipMemcpyAsync(&any, anyimp.p, sizeof(int), hipMemcpyDeviceToHost, s));
    if (report) {
      unsigned mb[4];
      int ni = 0;
      HIPCHK(hipMemcpyAsync(mb, cflmax.p, 3 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
      HIPCHK(hipMemcpyAsync(&ni, nimp.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      for (int i = 0; i < 3; ++i) { float f; std::memcpy(&f, &mb[i], 4); cfl_out[i] = f; }
      cfl_out[3] = ni;
    }
    HIPCHK(hipStreamSynchronize(s));
    return any != 0;
  }

  void set_pdef(int n, int v) override { check_n(n); pdef[n - 1] = v != 0; }
  void set_sgsmix(int n, int v) override { check_n(n); sgsmix[n - 1] = v != 0; }

  // whether the tracer step diffuses some tracer horizontally (scalar_transport_rk.F90:
  // 244-274: do_sgsmix tracers once the PBL's diffusivities are set); uploads the flags
  bool sgs_setup() {
    bool any = false;
    for (int n = 0; n < ntr; ++n) any = any || sgsmix[n];
    if (!dturb.on || !any) return false;
    HIPCHK(hipMemcpyAsync(sgs_dev.p, sgsmix.data(), ntr * sizeof(int), hipMemcpyHostToDevice, s));
    return true;
  }
  // the density the tracer step divides by: rho, or with nudging that preserves mixing
  // ratios rho - dtlm * rhot_nud (scalar_transport_rk.F90:120-133); oname_check.f90:218
  // turns nud_preserve_mix_ratio off when nud_preserve_total_mass is on
  RP *tracer_rho() {
    if (!(prm.nudflag > 0 && prm.nud_preserve_mix_ratio && !prm.nud_preserve_total_mass && nud_r.p))
      return dp<RP>(rho);
    if (!rhof.p) rhof.alloc(wsz * sizeof(RP));
    const long n = (long)wsz;
    LAUNCH("k_tr_rhof", (k_tr_rhof<RP>), dim3((unsigned)((n + 255) / 256)), dim3(256), dp<RP>(rho), dp<RP>(nud_r),
           prm.dtlm, dp<RP>(rhof), n);
    return dp<RP>(rhof);
  }

  // Per-tracer final-stage treatment (scalar_transport_rk.F90:297-337): mode 1 with
  // iscal_monot = 1, mode 2 with iscal_monot = 2 for pdef tracers, else 0 (regular
  // update); scp0 gets the low-order tendency update when iscal_monot > 0 and pdef.
  void lim_setup() {
    any_lim = any_pre = false;
    if (prm.iscal_monot == 0 || ntr == 0) return;
    std::vector<int> mode(ntr), pre(ntr);
    for (int n = 0; n < ntr; ++n) {
      mode[n] = prm.iscal_monot == 1 ? 1 : (pdef[n] ? 2 : 0);
      pre[n] = pdef[n] ? 1 : 0;
      any_lim = any_lim || mode[n] != 0;
      any_pre = any_pre || (pre[n] && have_sct);
    }
    HIPCHK(hipMemcpyAsync(lim_mode.p, mode.data(), ntr * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(lim_pre.p, pre.data(), ntr * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  TrLim<R> lim_args(bool tape) {
    TrLim<R> L;
    L.tmode = lim_mode.as<int>();
    L.s1 = dp<R>(lim_s1); L.s2 = dp<R>(lim_s2);
    L.tsv = tape && lim_tsv.p ? dp<R>(lim_tsv) : nullptr;
    L.tsw = tape && lim_tsw.p ? dp<R>(lim_tsw) : nullptr;
    L.vstride = (long)vsz;
    return L;
  }

  void scalar_transport_rk(const float *rho_old_host) override {
    if (rho_old_host) h2d<R>(rho_old, rho_old_host, t.perm_w);
    const int nrk = prm.nrk_scal;
    R dtlr[3];
    stage_dts(dtlr, nrk);
    const bool any_imp = run_cfl(false);
    cfl_fresh = true;
    cfl_any = any_imp;
    if (ntr == 0) { have_fwd = true; return; }
    lim_setup();
    taped_lim = false;
    if (sgs_setup()) {
      if (!trsct.p) { trsct.alloc(std::max<size_t>(1, ntr) * wsz * sizeof(R)); zero(trsct); }
      have_sct = true;
      LAUNCH("k_tr_sgsmix", (k_tr_sgsmix<R>), dim3(nblk(nw)), dim3(256), dm, dp<R>(trS), dp<R>(trsct),
             dp<R>(tb_akhodx), dp<int>(tb_khtopv), sgs_dev.as<int>(), ntr, tstride);
    }
    RP *rho_tr = tracer_rho();
    // RK3: S -> X -> Y -> X;  RK2: S -> X -> Y
    DevBuf *in = &trS, *outs[3] = {&trX, &trY, &trX};
    if (nrk == 2) { outs[0] = &trX; outs[1] = &trY; }
    for (int istage = 1; istage <= nrk; ++istage) {
      DevBuf *out = outs[istage - 1];
      TrStage<R, RP> a;
      a.q_in = dp<R>(*in); a.q0 = dp<R>(trS); a.sct = have_sct ? dp<R>(trsct) : nullptr;
      a.gx = dp<R>(trG); a.gy = dp<R>(trG) + (size_t)std::max(ntr, 1) * tstride; a.q_out = dp<R>(*out);
      a.gxy = trG.as<P2<R>>();
      for (int t = 0; t < 3; ++t) a.gq[t] = dp<R>(trgq[t]);
      a.vmasc = dp<R>(vmasc); a.wmasc = dp<R>(wmasc); a.rho_old = dp<R>(rho_old); a.ffimp = dp<R>(ffimp);
      a.rho = rho_tr;
      a.ntr = ntr; a.nrk = nrk; a.istage = istage; a.order = prm.scal_horiz_adv_order; a.tstride = tstride;
      a.dtr = dtlr[istage - 1];
      a.tpw = tpw;
      a.tmode = lim_mode.as<int>();
      a.lim = (istage == nrk && any_lim) ? 1 : 0;
      if (istage == nrk && any_pre)
        LAUNCH("k_tr_lim_pre", (k_tr_lim_pre<R>), dim3(nblk(nw)), dim3(256), dm, dp<R>(trS), dp<R>(trsct),
               dp<R>(rho_old), lim_pre.as<int>(), ntr, tstride, a.dtr);
      if (a.lim) {
        // limited final stage: gradients, regular update of the other tracers, then the
        // two limiter passes
        const TrLim<R> L = lim_args(prm.adjoint_tape > 0);
        tracer_update(a);
        LAUNCH("k_tr_lim_scales", (k_tr_lim_scales<R, RP>), dim3(nblk_tr()), dim3(256), dm, a, L);
        LAUNCH("k_tr_lim_update", (k_tr_lim_update<R, RP>), dim3(nblk_tr()), dim3(256), dm, a, L);
        taped_lim = prm.adjoint_tape > 0;
      } else if (use_patch && a.order < 3) {
        const dim3 g(pf.h.np), b(512);
        if (pf.cpw == 4) LAUNCH_SH("k_tr_patch", (k_tr_patch<R, RP, 4, rows_per_wave(4)>), g, b, pf.lds, dm, a, pf.tab);
        else if (pf.cpw == 2) LAUNCH_SH("k_tr_patch", (k_tr_patch<R, RP, 2, rows_per_wave(2)>), g, b, pf.lds, dm, a, pf.tab);
        else LAUNCH_SH("k_tr_patch", (k_tr_patch<R, RP, 1, rows_per_wave(1)>), g, b, pf.lds, dm, a, pf.tab);
      } else {
        tracer_update(a);
      }
      if (any_imp) LAUNCH("k_tr_implicit", (k_tr_implicit<R, RP>), dim3(nblk(nw)), dim3(256), dm, a);
      in = out;
    }
    // the new state is in `in`; make it the state buffer
    std::swap(trS.p, in->p);
    std::swap(trS.bytes, in->bytes);
    HIPCHK(hipGetLastError());
    have_fwd = true;
  }

  // reconstruction coefficients of the stage input, then the update of the unlimited
  // tracers (grad_t2d or grad_t2d_quadratic, scalar_v_column + scalar_w_column)
  void tracer_update(const TrStage<R, RP> &a) {
    // a limited final stage keeps the gradients: the limiter passes read them
    if (a.order < 3 && fused_fwd && !a.lim) {
      LAUNCH("k_tr_stage", (k_tr_stage<R, RP>), dim3(nblk_tr()), dim3(256), dm, a);
      return;
    }
    if (a.order == 2 && !two_pass && !a.lim) {
      LAUNCH("k_tr_grad", (k_tr_grad<R, true>), dim3(nblk_tr()), dim3(256), dm, a.q_in, nullptr, nullptr,
             (P2<R> *)a.gxy, ntr, tstride, tpw);
      if (dm.tall) LAUNCH("k_tr_update", (k_tr_update<R, RP, false, true, true>), dim3(nblk_tr()), dim3(256), dm, a);
      else LAUNCH("k_tr_update", (k_tr_update<R, RP, false, true>), dim3(nblk_tr()), dim3(256), dm, a);
      return;
    }
    if (a.order == 2)
      LAUNCH("k_tr_grad", (k_tr_grad<R, false>), dim3(nblk_tr()), dim3(256), dm, a.q_in, (R *)a.gx, (R *)a.gy, nullptr,
             ntr, tstride, tpw);
    if (a.order == 3) {
      const Q5<R> g{{(R *)a.gx, (R *)a.gy, dp<R>(trgq[0]), dp<R>(trgq[1]), dp<R>(trgq[2])}};
      LAUNCH("k_tr_grad3", (k_tr_grad3<R>), dim3(nblk_tr()), dim3(256), dm, a.q_in, g, ntr, tstride, tpw, a.tmode,
             a.lim);
      if (dm.tall) LAUNCH("k_tr_update", (k_tr_update<R, RP, true, false, true>), dim3(nblk_tr()), dim3(256), dm, a);
      else LAUNCH("k_tr_update", (k_tr_update<R, RP, true>), dim3(nblk_tr()), dim3(256), dm, a);
    } else {
      if (dm.tall) LAUNCH("k_tr_update", (k_tr_update<R, RP, false, false, true>), dim3(nblk_tr()), dim3(256), dm, a);
      else LAUNCH("k_tr_update", (k_tr_update<R, RP, false>), dim3(nblk_tr()), dim3(256), dm, a);
    }
  }

  // -------------------------------------------------------------- adjoint
  void scalar_transport_rk_adj() override {
    if (!have_fwd) throw Err(OLAMGPU_ESTATE, "adjoint called before a forward tracer step");
    alloc_adj();
    if (ntr == 0) return;
    lim_setup();
    if (any_lim && !taped_lim)
      throw Err(OLAMGPU_ESTATE, "limited step has no taped scales (adjoint_tape = 0 at create)");
    const int nrk = prm.nrk_scal;
    R dtlr[3];
    stage_dts(dtlr, nrk);
    // the forward step's implicit fractions are reused when nothing changed the winds
    const bool any_imp = cfl_fresh ? cfl_any : run_cfl(false);
    cfl_fresh = true;
    cfl_any = any_imp;
    const bool sgs = sgs_setup();
    if (sgs) {
      for (int n = 0; n < ntr; ++n)
        if (sgsmix[n] && any_lim && (prm.iscal_monot == 1 || pdef[n]))
          throw Err(OLAMGPU_EINVAL, "adjoint of tracer horizontal diffusion with a final-stage limiter is not built");
      if (!lsct.p) lsct.alloc(std::max<size_t>(1, ntr) * wsz * sizeof(R));
    }
    RP *rho_tr = tracer_rho();
    // lam0 is not cleared: the first adjoint stage (a.first) starts it from zero
    for (int istage = nrk; istage >= 1; --istage) {
      TrAdj<R, RP> a;
      DevBuf &muB = cur_adj_buf();
      DevBuf &lqB = cur_adj ? adjA : adjB;
      a.mu = dp<R>(muB); a.lam0 = dp<R>(lam0); a.nu = dp<R>(nu); a.lpart = dp<R>(lpart);
      a.lgx = dp<R>(lgG); a.lgy = dp<R>(lgG) + (size_t)std::max(ntr, 1) * tstride; a.lq = dp<R>(lqB);
      for (int t = 0; t < 3; ++t) a.lgq[t] = dp<R>(lgq[t]);
      a.vmasc = dp<R>(vmasc); a.wmasc = dp<R>(wmasc); a.rho_old = dp<R>(rho_old); a.ffimp = dp<R>(ffimp);
      a.rho = rho_tr;
      a.ntr = ntr; a.nrk = nrk; a.istage = istage; a.order = prm.scal_horiz_adv_order; a.last = istage == 1;
      a.first = istage == nrk;
      a.tstride = tstride; a.dtr = dtlr[istage - 1];
      a.tpw = tpw;
      a.tmode = lim_mode.as<int>();
      a.lim = (istage == nrk && any_lim) ? 1 : 0;
      if (any_imp) LAUNCH("k_tr_adj_implicit", (k_tr_adj_implicit<R, RP>), dim3(nblk(nw)), dim3(256), dm, a);
      if (sgs)
        LAUNCH("k_tr_adj_sct", (k_tr_adj_sct<R, RP>), dim3(nblk(nw)), dim3(256), dm, a, dp<R>(lsct),
               sgs_dev.as<int>());
      if (a.lim) {
        const TrLim<R> L = lim_args(true);
        adjoint_faces(a);
        LAUNCH("k_tr_lim_adj0", (k_tr_lim_adj0<R, RP>), dim3(nblk_tr()), dim3(256), dm, a, L);
        LAUNCH("k_tr_lim_adj1", (k_tr_lim_adj1<R, RP>), dim3(nblk_tr()), dim3(256), dm, a, L);
        adjoint_coefs(a);
      } else if (use_patch && a.order < 3) {
        const dim3 g(pa.h.np), b(512);
        if (pa.cpw == 4) LAUNCH_SH("k_tr_adj_patch", (k_tr_adj_patch<R, RP, 4>), g, b, pa.lds, dm, a, pa.tab);
        else if (pa.cpw == 2) LAUNCH_SH("k_tr_adj_patch", (k_tr_adj_patch<R, RP, 2>), g, b, pa.lds, dm, a, pa.tab);
        else LAUNCH_SH("k_tr_adj_patch", (k_tr_adj_patch<R, RP, 1>), g, b, pa.lds, dm, a, pa.tab);
      } else {
        adjoint_faces(a);
        adjoint_coefs(a);
      }
      cur_adj = 1 - cur_adj;
    }
    // the diffusion tendency was built from the tracers before the step
    if (sgs)
      LAUNCH("k_tr_sgsmix_adj", (k_tr_sgsmix_adj<R>), dim3(nblk(nw)), dim3(256), dm, dp<R>(lsct),
             dp<R>(cur_adj_buf()), dp<R>(tb_akhodx), dp<int>(tb_khtopv), sgs_dev.as<int>(), ntr, tstride);
    HIPCHK(hipGetLastError());
  }
  // column and face parts of an adjoint stage for the unlimited tracers: fused in one
  // pass (default) or as k_tr_adj0 + k_tr_adj1 (OLAMGPU_TRACER=split)
  void adjoint_faces(const TrAdj<R, RP> &a) {
    const bool q3 = a.order == 3;
    if (fuse_adj) {
      if (q3 && dm.tall) LAUNCH("k_tr_adj01", (k_tr_adj01<R, RP, true, true>), dim3(nblk_tr()), dim3(256), dm, a);
      else if (q3) LAUNCH("k_tr_adj01", (k_tr_adj01<R, RP, true>), dim3(nblk_tr()), dim3(256), dm, a);
      else if (dm.tall) LAUNCH("k_tr_adj01", (k_tr_adj01<R, RP, false, true>), dim3(nblk_tr()), dim3(256), dm, a);
      else LAUNCH("k_tr_adj01", (k_tr_adj01<R, RP, false>), dim3(nblk_tr()), dim3(256), dm, a);
    } else {
      LAUNCH("k_tr_adj0", (k_tr_adj0<R, RP>), dim3(nblk_tr()), dim3(256), dm, a);
      if (q3) LAUNCH("k_tr_adj1", (k_tr_adj1<R, RP, true>), dim3(nblk_tr()), dim3(256), dm, a);
      else LAUNCH("k_tr_adj1", (k_tr_adj1<R, RP, false>), dim3(nblk_tr()), dim3(256), dm, a);
    }
  }
  // transpose of the reconstruction coefficients into the stage input's adjoint
  void adjoint_coefs(const TrAdj<R, RP> &a) {
    if (a.order == 3) LAUNCH("k_tr_adj2q", (k_tr_adj2q<R, RP>), dim3(nblk_tr()), dim3(256), dm, a);
    else LAUNCH("k_tr_adj2", (k_tr_adj2<R, RP>), dim3(nblk_tr()), dim3(256), dm, a);
  }

  // -------------------------------------------------------------- diagnostics
  void mass_sums(double *out) override {
    const unsigned nb = nblk(nw);
    LAUNCH("k_mass_partial", (k_mass_partial<R, RP>), dim3(nb), dim3(256), dm, dp<RP>(rho), dp<R>(trS), ntr, tstride,
                       masspart.as<double>());
    LAUNCH("k_mass_final", k_mass_final, dim3(1), dim3(256), masspart.as<double>(), (int)nb, 1 + ntr,
                       massout.as<double>());
    HIPCHK(hipMemcpyAsync(out, massout.p, (1 + ntr) * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  // compute_mass_sums into a device array, with the partial sums in the static scratch
  // (no allocation, no host copy)
  void mass_sums_noalloc(double *out_dev, double *scr) override {
    if (1 + ntr > MASS_NV) throw Err(OLAMGPU_EINVAL, "mass_sums_noalloc carries at most 64 tracers");
    if (!scr) throw Err(OLAMGPU_ESTATE, "scratch address not set up on this device");
    hipLaunchKernelGGL((k_mass_partial_gs<R, RP>), dim3(MASS_NB), dim3(256), 0, s, dm, dp<RP>(rho), dp<R>(trS), ntr,
                       tstride, scr);
    hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(256), 0, s, (const double *)scr, MASS_NB, 1 + ntr, out_dev);
    HIPCHK(hipGetLastError());
  }
  // tracer_maxmin (check_nans.F90:164-203): out[2n] = max, out[2n+1] = min of tracer n+1
  void tracer_maxmin(float *out) override {
    if (ntr == 0) return;
    LAUNCH("k_maxmin_init", k_maxmin_init, dim3(1), dim3(128), mmkeys.as<int>(), 2 * ntr);
    LAUNCH("k_tr_maxmin", (k_tr_maxmin<R>), dim3(nblk(nw)), dim3(256), dm, dp<R>(trS), ntr, tstride,
           mmkeys.as<int>());
    LAUNCH("k_maxmin_final", k_maxmin_final, dim3(1), dim3(128), mmkeys.as<int>(), 2 * ntr, mmout.as<float>());
    HIPCHK(hipMemcpyAsync(out, mmout.p, 2 * ntr * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  void cfl_report(double *o) override {
    run_cfl(true);
    for (int i = 0; i < 4; ++i) o[i] = cfl_out[i];
  }

  void copy_field(int id, int index, void *buf, int dir, int *eb) override;

  // -------------------------------------------------------------- checkpoints
  // STATE: every field that the library's own calls modify and read back later
  // (prognostic fields, accumulated mass fluxes, rho_old), so a restored context
  // continues bit-identically as long as the host-set inputs (alpha_press, pwfac,
  // pvfac, long-step tendencies, tracer tendencies) are the same.  TAPE: what a tracer
  // step and its adjoint read besides the tracers (vmasc, wmasc, rho, rho_old).
  std::vector<DevBuf *> ck_list(int kind) {
    // With a final-stage limiter the tracers themselves are part of both: the scales
    // depend on them, so a recomputation needs them and the adjoint needs the scales.
    const bool lim = prm.iscal_monot > 0;
    if (kind == OLAMGPU_CKPT_STATE) {
      std::vector<DevBuf *> v = {&cur_thil_buf(), &wmc, &wc, &vxe, &vye, &vze, &vmc, &vc, &rho, &press, &vmasc,
                                 &wmasc, &rho_old};
      if (lim) v.push_back(&trS);
      return v;
    }
    if (kind == OLAMGPU_CKPT_TAPE) {
      std::vector<DevBuf *> v = {&vmasc, &wmasc, &rho, &rho_old};
      if (lim && lim_tsv.p) { v.push_back(&lim_tsv); v.push_back(&lim_tsw); }
      return v;
    }
    throw Err(OLAMGPU_EINVAL, "unknown checkpoint kind");
  }
  static size_t ck_round(size_t b) { return (b + 255) & ~(size_t)255; }
  size_t checkpoint_bytes(int kind) override {
    size_t tot = 0;
    for (DevBuf *b : ck_list(kind)) tot += ck_round(b->bytes);
    return tot;
  }
  void checkpoint(int kind, void *buf, int restore) override {
    if (restore) cfl_fresh = false;
    size_t off = 0;
    for (DevBuf *b : ck_list(kind)) {
      char *d = (char *)buf + off;
      if (restore) HIPCHK(hipMemcpyAsync(b->p, d, b->bytes, hipMemcpyDeviceToDevice, s));
      else HIPCHK(hipMemcpyAsync(d, b->p, b->bytes, hipMemcpyDeviceToDevice, s));
      off += ck_round(b->bytes);
    }
    // the adjoint of a tracer step depends only on the taped winds and densities
    if (restore && kind == OLAMGPU_CKPT_TAPE) {
      have_fwd = true;
      taped_lim = prm.iscal_monot > 0 && lim_tsv.p;
    }
  }
};

// ---------------------------------------------------------------- small kernels
template <class R, class RP>
__global__ void k_cast_rho(const RP *rho, R *ro, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ro[i] = (R)rho[i];
}
template <class R>
__global__ void k_scale(R *a, long n, R f, int mza, int ldk, int klo, int khi) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int k = (int)(i % ldk) + 1;
  if (k >= klo && k <= khi) a[i] = a[i] / f;
}
// dense (n, mza) original order <-> device (n, ldk) renumbered
template <class T>
__global__ void k_perm_copy(const int *inv, T *dev, T *dense, int n, int mza, int ldk, int dir) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n * mza) return;
  long o = i / mza;
  int l = (int)(i % mza);
  long d = (long)inv[o] * ldk + l;
  if (dir == 0) dense[i] = dev[d];
  else dev[d] = dense[i];
}

template <class R, class RP>
void Impl<R, RP>::copy_rho_old() {
  long n = (long)wsz;
  LAUNCH("k_cast_rho", (k_cast_rho<R, RP>), dim3((unsigned)((n + 255) / 256)), dim3(256), dp<RP>(rho),
                     dp<R>(rho_old), n);
  HIPCHK(hipGetLastError());
}

template <class R, class RP>
void Impl<R, RP>::timeavg_momsc() {
  cfl_fresh = false;
  if (prm.nacoust <= 1) return;
  R f = (R)prm.nacoust;
  LAUNCH("k_scale", (k_scale<R>), dim3((unsigned)((vsz + 255) / 256)), dim3(256), dp<R>(vmasc), (long)vsz, f,
                     mza, ldk, 2, mza);
  LAUNCH("k_scale", (k_scale<R>), dim3((unsigned)((wsz + 255) / 256)), dim3(256), dp<R>(wmasc), (long)wsz, f,
                     mza, ldk, 2, mza - 1);
  HIPCHK(hipGetLastError());
}

template <class R, class RP>
void Impl<R, RP>::copy_field(int id, int index, void *buf, int dir, int *eb) {
  if (dir) cfl_fresh = false;
  DevBuf *b = nullptr;
  size_t off = 0;
  bool isrp = false, isv = false;
  switch (id) {
    case OLAMGPU_F_SCALAR: check_n(index); b = &trS; off = (size_t)(index - 1) * tstride; break;
    case OLAMGPU_F_SCALAR_ADJ: check_n(index); alloc_adj(); b = &cur_adj_buf(); off = (size_t)(index - 1) * tstride; break;
    case OLAMGPU_F_RHO: b = &rho; isrp = true; break;
    case OLAMGPU_F_PRESS: b = &press; isrp = true; break;
    case OLAMGPU_F_RHO_OLD: b = &rho_old; break;
    case OLAMGPU_F_VMASC: b = &vmasc; isv = true; break;
    case OLAMGPU_F_WMASC: b = &wmasc; break;
    case OLAMGPU_F_THIL: b = &thil[cur_thil]; break;
    case OLAMGPU_F_VMC: b = &vmc; isv = true; break;
    case OLAMGPU_F_WMC: b = &wmc; break;
    default: throw Err(OLAMGPU_EINVAL, "unknown field id");
  }
  int n = isv ? nv : nw;
  const int *dinv = isv ? b_invv.as<int>() : b_invw.as<int>();
  long tot = (long)n * mza;
  unsigned g = (unsigned)((tot + 255) / 256);
  if (isrp) {
    if (eb) *eb = sizeof(RP);
    LAUNCH("k_perm_copy", k_perm_copy<RP>, dim3(g), dim3(256), dinv, b->as<RP>() + off, (RP *)buf, n, mza, ldk, dir);
  } else {
    if (eb) *eb = sizeof(R);
    LAUNCH("k_perm_copy", k_perm_copy<R>, dim3(g), dim3(256), dinv, b->as<R>() + off, (R *)buf, n, mza, ldk, dir);
  }
  HIPCHK(hipGetLastError());
}

}  // namespace

// ------------------------------------------------------------------ C-ABI
struct olamgpu_ctx {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  Prof prof;
  std::unique_ptr<ImplBase> impl;
};

static std::string g_err;

template <class F>
static int guarded(olamgpu_ctx *c, F &&f) {
  if (!c) { g_err = "null context"; return OLAMGPU_EINVAL; }
  try {
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    f();
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->prof.flush();
    return OLAMGPU_OK;
  } catch (const Err &e) {
    c->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    c->err = e.what();
    return OLAMGPU_EHIP;
  }
}

extern "C" {

int olamgpu_create(const olamgpu_mesh *mesh, const olamgpu_params *params, int32_t device, olamgpu_ctx **out) {
  if (!mesh || !params || !out) { g_err = "null argument"; return OLAMGPU_EINVAL; }
  *out = nullptr;
  auto *c = new olamgpu_ctx;
  try {
    int dev = device;
    if (dev < 0) HIPCHK(hipGetDevice(&dev));
    c->device = dev;
    HIPCHK(hipSetDevice(dev));
    scratch_setup(dev);
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    Topo t = build_topo(mesh);
    if (params->precision == OLAMGPU_PREC_F64)
      c->impl.reset(new Impl<double, double>(mesh, params, c->stream, &c->prof, std::move(t)));
    else if (params->precision == OLAMGPU_PREC_MIXED)
      c->impl.reset(new Impl<float, double>(mesh, params, c->stream, &c->prof, std::move(t)));
    else if (params->precision == OLAMGPU_PREC_F32)
      c->impl.reset(new Impl<float, float>(mesh, params, c->stream, &c->prof, std::move(t)));
    else
      throw Err(OLAMGPU_EINVAL, "unknown precision mode");
  } catch (const Err &e) {
    g_err = e.what();
    olamgpu_destroy(c);
    return e.code;
  } catch (const std::exception &e) {
    g_err = e.what();
    olamgpu_destroy(c);
    return OLAMGPU_EHIP;
  }
  *out = c;
  return OLAMGPU_OK;
}

int olamgpu_destroy(olamgpu_ctx *c) {
  if (!c) return OLAMGPU_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->impl.reset();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return OLAMGPU_OK;
}

const char *olamgpu_last_error(const olamgpu_ctx *c) { return c ? c->err.c_str() : g_err.c_str(); }

int olamgpu_upload_state(olamgpu_ctx *c, const olamgpu_state *h) {
  return guarded(c, [&] { if (!h) throw Err(OLAMGPU_EINVAL, "null state"); c->impl->upload_state(h); });
}
int olamgpu_download_state(olamgpu_ctx *c, olamgpu_state *h) {
  return guarded(c, [&] { if (!h) throw Err(OLAMGPU_EINVAL, "null state"); c->impl->download_state(h); });
}
int olamgpu_upload_scalar(olamgpu_ctx *c, int32_t n, const float *p, const float *t) {
  return guarded(c, [&] { c->impl->upload_scalar(n, p, t); });
}
int olamgpu_download_scalar(olamgpu_ctx *c, int32_t n, float *p) {
  return guarded(c, [&] { if (!p) throw Err(OLAMGPU_EINVAL, "null buffer"); c->impl->download_scalar(n, p); });
}
int olamgpu_set_nudging(olamgpu_ctx *c, const float *wmanud, const float *vmanud, const double *rhot_nud) {
  return guarded(c, [&] { c->impl->set_nudging(wmanud, vmanud, rhot_nud); });
}
int olamgpu_set_turbulence(olamgpu_ctx *c, const float *vkh, const float *vkm, const int32_t *khtop,
                           const int32_t *kmtop, const float *akm_sfc) {
  return guarded(c, [&] { c->impl->set_turbulence(vkh, vkm, khtop, kmtop, akm_sfc); });
}
int olamgpu_comp_alpha_press(olamgpu_ctx *c, const float *rr_v, const float *rr_w, const float *theta) {
  return guarded(c, [&] { c->impl->comp_alpha_press(rr_v, rr_w, theta); });
}
int olamgpu_init_wrtv_rk(olamgpu_ctx *c) { return guarded(c, [&] { c->impl->init_wrtv_rk(); }); }
int olamgpu_tend0(olamgpu_ctx *c) { return guarded(c, [&] { c->impl->tend0(); }); }
int olamgpu_prog_wrtv_rk(olamgpu_ctx *c) { return guarded(c, [&] { c->impl->prog_wrtv_rk(); }); }
int olamgpu_timeavg_momsc(olamgpu_ctx *c) { return guarded(c, [&] { c->impl->timeavg_momsc(); }); }
int olamgpu_scalar_transport_rk(olamgpu_ctx *c, const float *rho_old) {
  return guarded(c, [&] { c->impl->scalar_transport_rk(rho_old); });
}
int olamgpu_scalar_transport_rk_adj(olamgpu_ctx *c) { return guarded(c, [&] { c->impl->scalar_transport_rk_adj(); }); }
int olamgpu_upload_scalar_adj(olamgpu_ctx *c, int32_t n, const float *p) {
  return guarded(c, [&] { if (!p) throw Err(OLAMGPU_EINVAL, "null buffer"); c->impl->upload_adj(n, p); });
}
int olamgpu_download_scalar_adj(olamgpu_ctx *c, int32_t n, float *p) {
  return guarded(c, [&] { if (!p) throw Err(OLAMGPU_EINVAL, "null buffer"); c->impl->download_adj(n, p); });
}
int olamgpu_mass_sums(olamgpu_ctx *c, double *out) {
  return guarded(c, [&] { if (!out) throw Err(OLAMGPU_EINVAL, "null buffer"); c->impl->mass_sums(out); });
}
int olamgpu_mass_sums_noalloc(olamgpu_ctx *c, double *out_dev) {
  return guarded(c, [&] {
    if (!out_dev) throw Err(OLAMGPU_EINVAL, "null buffer");
    c->impl->mass_sums_noalloc(out_dev, scratch_ptr(c->device));
  });
}
int olamgpu_cfl_report(olamgpu_ctx *c, double *out) {
  return guarded(c, [&] { if (!out) throw Err(OLAMGPU_EINVAL, "null buffer"); c->impl->cfl_report(out); });
}
int olamgpu_copy_field(olamgpu_ctx *c, int32_t field_id, int32_t index, void *dev_buf, int32_t direction,
                       int32_t *elem_bytes) {
  return guarded(c, [&] {
    if (!dev_buf || (direction != 0 && direction != 1)) throw Err(OLAMGPU_EINVAL, "bad copy arguments");
    c->impl->copy_field(field_id, index, dev_buf, direction, elem_bytes);
  });
}
int olamgpu_tracer_maxmin(olamgpu_ctx *c, float *out) {
  return guarded(c, [&] {
    if (!out) throw Err(OLAMGPU_EINVAL, "tracer_maxmin needs an output array");
    c->impl->tracer_maxmin(out);
  });
}
int olamgpu_set_scalar_sgsmix(olamgpu_ctx *c, int32_t n, int32_t v) {
  return guarded(c, [&] { c->impl->set_sgsmix(n, v); });
}
int olamgpu_set_scalar_pdef(olamgpu_ctx *c, int32_t n, int32_t pdef) {
  return guarded(c, [&] { c->impl->set_pdef(n, pdef); });
}
int olamgpu_checkpoint_bytes(olamgpu_ctx *c, int32_t kind, int64_t *bytes) {
  return guarded(c, [&] {
    if (!bytes) throw Err(OLAMGPU_EINVAL, "null pointer");
    *bytes = (int64_t)c->impl->checkpoint_bytes(kind);
  });
}
int olamgpu_checkpoint_save(olamgpu_ctx *c, int32_t kind, void *dev_buf) {
  return guarded(c, [&] {
    if (!dev_buf) throw Err(OLAMGPU_EINVAL, "null buffer");
    c->impl->checkpoint(kind, dev_buf, 0);
  });
}
int olamgpu_checkpoint_restore(olamgpu_ctx *c, int32_t kind, const void *dev_buf) {
  return guarded(c, [&] {
    if (!dev_buf) throw Err(OLAMGPU_EINVAL, "null buffer");
    c->impl->checkpoint(kind, const_cast<void *>(dev_buf), 1);
  });
}
int olamgpu_set_profiling(olamgpu_ctx *c, int32_t on) {
  if (!c) return OLAMGPU_EINVAL;
  c->prof.reset();
  c->prof.on = on != 0;
  return OLAMGPU_OK;
}
int olamgpu_kernel_profile(olamgpu_ctx *c, char *text, int32_t len) {
  if (!c || !text || len <= 0) return OLAMGPU_EINVAL;
  std::string out;
  char line[256];
  for (size_t i = 0; i < c->prof.names.size(); ++i) {
    std::snprintf(line, sizeof line, "%s %ld %.6f\n", c->prof.names[i].c_str(), c->prof.count[i], c->prof.ms[i]);
    out += line;
  }
  std::snprintf(text, (size_t)len, "%s", out.c_str());
  return (int)out.size() < len ? OLAMGPU_OK : OLAMGPU_EINVAL;
}
int olamgpu_last_call_ms(olamgpu_ctx *c, double *ms) {
  if (!c || !ms) return OLAMGPU_EINVAL;
  *ms = c->last_ms;
  return OLAMGPU_OK;
}

}  // extern "C"

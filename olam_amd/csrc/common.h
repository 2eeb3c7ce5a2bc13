// common.h -- device-side layout and wave helpers shared by the HIP kernels.
//
// Layout in HBM.  Every 3-D field is stored column by column exactly as the
// reference's Fortran arrays (vertical index contiguous), with the dummy point
// removed, points renumbered for locality, and the column stride padded to LDK = 64
// elements (one wave; mza <= 64), levels above mza zero.  Fortran level k sits at offset k-1 of its
// column.  One 64-lane wavefront owns one column: lane l holds level k = l + 1, so
// every column load/store is a single coalesced 4*mza-byte access and neighbour
// columns are gathered whole.  Tall columns (64 < mza <= 128) have LDK = 128 and two
// consecutive waves of a block per column, the first holding levels 1..64 and the second
// 65..128; the vertical moves between them go through LDS (vshdn, vshup, vlane0, vpcr).  Per-point neighbour records are padded to 8 slots so
// a wave reads them with scalar loads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Expressions are evaluated as written, without fusing a*b+c into an FMA, in the
// kernels whose results must track the oracle (compiled with -ffp-contract=off) to the
// last bit: the dynamics, where small differences of large terms amplify rounding, and
// the limiters, whose scale factors switch on comparisons of such differences.
#define NOCONTRACT _Pragma("clang fp contract(off)")

#define WAVE 64
#define WPB 4          // waves per 256-thread block
#define NSLOT 8        // padded neighbour slots per W point

template <class R>
struct DMesh {
  int nw, nv, nm, mza, ldk;
  // W records (nw x NSLOT)
  const int *w_npoly, *w_nbr, *w_face, *w_back;
  const R *w_dirv, *w_gxc, *w_gyc, *w_dxs, *w_dys, *w_dxn, *w_dyn;
  const R *w_ecx, *w_ecy, *w_ecz, *w_xxyy;
  // order-3 reconstruction per slot: the least-squares terms xy_h ([0]) / xy_hu ([1])
  // of neighbour j, and the quadratic face offsets (xx, xy, yy) of this cell's side
  // (w_os) and of the neighbour's side (w_on), bounded ([0], xx0_v) / unbounded ([1])
  const R *w_q[2][5], *w_os[2][3], *w_on[2][3];
  const R *wnx, *wny, *wnz, *arw0;
  // V records (nv x 2)
  const int *v_iw, *v_im;
  const R *vnx, *vny, *vnz;
  // M records (nm x 4)
  const int *m_iv, *m_im;
  const R *m_vfact, *m_xxyym;
  // per level (ldk)
  const R *dzim, *gdz_wgtm, *gdz_wgtp, *aw11, *aw21, *aw12, *aw22;
  const R *awu11, *awu21, *awu12, *awu22;      // unbounded vertical quadratic (a_wu)
  const double *zwgt_top8, *zwgt_bot8, *gdz_wgtm8, *gdz_wgtp8;
  const R *zfacit;                             // per level
  const R *dniv;                               // per V face
  // control volumes (n x ldk)
  const R *arw, *volti, *volwi, *arv, *volvi;
  const double *volt;
  // Rayleigh friction near the model top (mem_rayf.f90), per level (ldk): rayf_cof is
  // null unless thil/vmc friction is on, rayf_cofmix unless the vertical wind mixing is
  // on, b20 (the W factor of the implicit solve) unless W friction is on
  const R *rayf_cof, *rayf_cofmix, *th01d, *dn01d, *b20;
  int krayf_bot, krayfmix_bot;
  int tall;      // 1: two waves per column (64 < mza <= 128)
};

// Blocks are handed to the 8 XCDs round-robin; give each XCD a contiguous range of
// logical blocks so neighbouring columns (which share gathers) hit the same L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  int per = nb >> 3, rem = nb & 7;
  int x = b & 7, idx = b >> 3;
  return x * per + (x < rem ? x : rem) + idx;
}

__device__ __forceinline__ int wave_item() {
  int b = xcd_block(blockIdx.x, gridDim.x);
  return __builtin_amdgcn_readfirstlane(b * WPB + (threadIdx.x >> 6));
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Work item (column, V face or M point) of a wave and the level index l (0-based, so
// Fortran level k = l + 1) of its lane.  With tall columns the two waves of an item are
// consecutive in the block: half h holds levels 64 h + 1 .. 64 h + 64.
struct WPos { int item, l, h; };
__device__ __forceinline__ WPos wave_pos(int tall) {
  const int w = wave_item();
  WPos p;
  p.h = tall ? (w & 1) : 0;
  p.item = tall ? (w >> 1) : w;
  p.l = p.h * WAVE + lane_id();
  return p;
}

// Work item of a tracer kernel: column c and the tracer range [n0, n1) handled by this
// wave.  The tpw tracers of one group share a wave; the groups of one column are
// consecutive items, so they run in the same block and share the column's gathers.
struct TrItem { int c, n0, n1, l; };
__device__ __forceinline__ TrItem tr_item(int nw, int ntr, int tpw, int tall) {
  const int ng = (ntr + tpw - 1) / tpw;
  const WPos p = wave_pos(tall);
  const int it = p.item;
  TrItem t;
  t.l = p.l;
  t.c = it / ng;
  t.n0 = (it - t.c * ng) * tpw;
  t.n1 = min(ntr, t.n0 + tpw);
  if (t.c >= nw) t.c = nw;
  return t;
}

// Cross-lane moves.  Call them with every lane of the wave active (never inside a
// per-lane branch or a per-lane ?: operand): a lane that does not take part supplies 0.
// One-lane shifts go through the DPP wave_shl:1 / wave_shr:1 operand modifiers (one
// VALU move per 32-bit word, no LDS round trip); a lane without a source keeps its own
// value, as __shfl_down / __shfl_up do.  Longer shifts use __shfl_down / __shfl_up.
template <int CTRL, class T>
__device__ __forceinline__ T dpp_shift1(T v) {
  static_assert(sizeof(T) % 4 == 0 || sizeof(T) < 4, "32-bit words");
  if constexpr (sizeof(T) < 4) {
    const int w = (int)v;
    return (T)__builtin_amdgcn_update_dpp(w, w, CTRL, 0xF, 0xF, false);
  } else {
    constexpr int NW = sizeof(T) / 4;
    int w[NW];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = __builtin_amdgcn_update_dpp(w[i], w[i], CTRL, 0xF, 0xF, false);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
  }
}
// The value of lane 0 (level 1 of the column) in every lane: one scalar read per
// 32-bit word.  Every lane must be active, so lane 0 is the first active lane.
template <class T>
__device__ __forceinline__ T lane0(T v) {
  constexpr int NW = sizeof(T) / 4;
  int w[NW];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < NW; ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}
template <class T>
__device__ __forceinline__ T shdn(T v, int d) {
  if (d == 1) return dpp_shift1<0x130>(v);     // wave_shl:1, lane l reads lane l + 1
  return __shfl_down(v, d, WAVE);
}
template <class T>
__device__ __forceinline__ T shup(T v, int d) {
  if (d == 1) return dpp_shift1<0x138>(v);     // wave_shr:1, lane l reads lane l - 1
  return __shfl_up(v, d, WAVE);
}

// Vertical moves that also work for tall columns.  The two waves of a tall column swap
// the words the other half needs through LDS (one 64-bit word per lane, two barriers),
// so every wave of the block must make the same calls: only in code that all waves of
// the block run alike, never inside a branch on the item's own data.
__device__ __forceinline__ unsigned long long *xbuf() {
  __shared__ unsigned long long x[WPB * WAVE];
  return x;
}
template <class T>
__device__ __forceinline__ unsigned long long to_u64(T v) {
  unsigned long long u = 0;
  __builtin_memcpy(&u, &v, sizeof(T));
  return u;
}
template <class T>
__device__ __forceinline__ T from_u64(unsigned long long u) {
  T v;
  __builtin_memcpy(&v, &u, sizeof(T));
  return v;
}
// level l + d (a level without one keeps its own value, as shdn)
template <class T>
__device__ __forceinline__ T vshdn(int tall, T v, int d) {
  T r = shdn(v, d);
  if (tall) {
    unsigned long long *x = xbuf();
    const int w = threadIdx.x >> 6, lane = lane_id();
    x[w * WAVE + lane] = to_u64(v);
    __syncthreads();
    if (!(w & 1) && lane + d >= WAVE) r = from_u64<T>(x[(w + 1) * WAVE + lane + d - WAVE]);
    __syncthreads();
  }
  return r;
}
// level l - d
template <class T>
__device__ __forceinline__ T vshup(int tall, T v, int d) {
  T r = shup(v, d);
  if (tall) {
    unsigned long long *x = xbuf();
    const int w = threadIdx.x >> 6, lane = lane_id();
    x[w * WAVE + lane] = to_u64(v);
    __syncthreads();
    if ((w & 1) && lane < d) r = from_u64<T>(x[(w - 1) * WAVE + WAVE - d + lane]);
    __syncthreads();
  }
  return r;
}
// the value at level 1 in every lane of the column
template <class T>
__device__ __forceinline__ T vlane0(int tall, T v) {
  T r = lane0(v);
  if (tall) {
    unsigned long long *x = xbuf();
    const int w = threadIdx.x >> 6, lane = lane_id();
    if (!(w & 1) && lane == 0) x[w * WAVE] = to_u64(v);
    __syncthreads();
    if (w & 1) r = from_u64<T>(x[(w - 1) * WAVE]);
    __syncthreads();
  }
  return r;
}

// Parallel cyclic reduction of a tridiagonal system held one row per lane:
// a x(l-1) + b x(l) + c x(l+1) = d.  Rows outside the system must be identity rows
// (a = c = d = 0, b = 1).  Returns x for this lane.  FP64 throughout.
__device__ __forceinline__ void pcr_step(int tall, int l, int nrow, int s, double &a, double &b, double &c,
                                         double &d) {
  double rb = 1.0 / b;
  double am = vshup(tall, a, s), rbm = vshup(tall, rb, s), cm = vshup(tall, c, s), dm = vshup(tall, d, s);
  double ap = vshdn(tall, a, s), rbp = vshdn(tall, rb, s), cp = vshdn(tall, c, s), dp = vshdn(tall, d, s);
  if (l < s) { am = 0.0; rbm = 1.0; cm = 0.0; dm = 0.0; }
  if (l + s >= nrow) { ap = 0.0; rbp = 1.0; cp = 0.0; dp = 0.0; }
  double alpha = -a * rbm;
  double gamma = -c * rbp;
  double na = alpha * am;
  double nc = gamma * cp;
  double nb = b + alpha * cm + gamma * ap;
  double nd = d + alpha * dm + gamma * dp;
  a = na; b = nb; c = nc; d = nd;
}
// Tall columns: the 128 rows of a column lie in two waves.  Each step publishes every
// row's (a, 1/b, c, d) in an LDS row array of the column and reads rows l - s and l + s
// from it, with the arithmetic of pcr_step.  Two arrays used in turn need one barrier
// per step: a wave writes the array of step k + 2 only after the barrier of step k + 1,
// which every wave reaches after its reads of step k; the barrier at the end does the
// same for the next call.  Every wave of the block must make the same calls.
struct PcrRow {
  double a, rb, c, d;
};
__device__ __forceinline__ double vpcr_tall(int l, double a, double b, double c, double d) {
  __shared__ PcrRow rows[2][WPB / 2][2 * WAVE];
  const int col = threadIdx.x >> 7;   // the block's two waves of one column
  const int nrow = 2 * WAVE;
  int k = 0;
#pragma unroll
  for (int s = 1; s < nrow; s <<= 1, ++k) {
    PcrRow *r = rows[k & 1][col];
    const double rb = 1.0 / b;
    r[l] = PcrRow{a, rb, c, d};
    __syncthreads();
    double am = 0.0, rbm = 1.0, cm = 0.0, dm = 0.0;
    double ap = 0.0, rbp = 1.0, cp = 0.0, dp = 0.0;
    if (l >= s) { const PcrRow q = r[l - s]; am = q.a; rbm = q.rb; cm = q.c; dm = q.d; }
    if (l + s < nrow) { const PcrRow q = r[l + s]; ap = q.a; rbp = q.rb; cp = q.c; dp = q.d; }
    double alpha = -a * rbm;
    double gamma = -c * rbp;
    double na = alpha * am;
    double nc = gamma * cp;
    double nb = b + alpha * cm + gamma * ap;
    double nd = d + alpha * dm + gamma * dp;
    a = na; b = nb; c = nc; d = nd;
  }
  __syncthreads();
  return d / b;
}
// l: the row's level index; tall columns have 128 rows over the two waves
__device__ __forceinline__ double vpcr(int tall, int l, double a, double b, double c, double d) {
  if (tall) return vpcr_tall(l, a, b, c, d);
#pragma unroll
  for (int s = 1; s < WAVE; s <<= 1) pcr_step(0, l, WAVE, s, a, b, c, d);
  return d / b;
}
__device__ __forceinline__ double pcr_solve(double a, double b, double c, double d) {
  return vpcr(0, lane_id(), a, b, c, d);
}

// pcr_solve for three right-hand sides of the same matrix (the eliminations of each d are
// those of pcr_solve, so each solution equals pcr_solve's for that d).
__device__ __forceinline__ void pcr_solve3(double a, double b, double c, double d0, double d1, double d2,
                                           double &x0, double &x1, double &x2) {
  const int l = lane_id();
#pragma unroll
  for (int s = 1; s < WAVE; s <<= 1) {
    double rb = 1.0 / b;
    double am = shup(a, s), rbm = shup(rb, s), cm = shup(c, s);
    double ap = shdn(a, s), rbp = shdn(rb, s), cp = shdn(c, s);
    double d0m = shup(d0, s), d1m = shup(d1, s), d2m = shup(d2, s);
    double d0p = shdn(d0, s), d1p = shdn(d1, s), d2p = shdn(d2, s);
    if (l < s) { am = 0.0; rbm = 1.0; cm = 0.0; d0m = d1m = d2m = 0.0; }
    if (l + s >= WAVE) { ap = 0.0; rbp = 1.0; cp = 0.0; d0p = d1p = d2p = 0.0; }
    double alpha = -a * rbm;
    double gamma = -c * rbp;
    double na = alpha * am;
    double nc = gamma * cp;
    double nb = b + alpha * cm + gamma * ap;
    d0 = d0 + alpha * d0m + gamma * d0p;
    d1 = d1 + alpha * d1m + gamma * d1p;
    d2 = d2 + alpha * d2m + gamma * d2p;
    a = na; b = nb; c = nc;
  }
  x0 = d0 / b; x1 = d1 / b; x2 = d2 / b;
}
// pcr_solve3 for a column that may be tall (l: the row's level index).  A tall column
// solves each right-hand side in turn: the eliminations of each d are those of the
// joint solve, and the column's LDS row array holds one right-hand side.
__device__ __forceinline__ void vpcr3(int tall, int l, double a, double b, double c, double d0, double d1,
                                      double d2, double &x0, double &x1, double &x2) {
  if (tall) {
    x0 = vpcr_tall(l, a, b, c, d0);
    x1 = vpcr_tall(l, a, b, c, d1);
    x2 = vpcr_tall(l, a, b, c, d2);
    return;
  }
  pcr_solve3(a, b, c, d0, d1, d2, x0, x1, x2);
}

// Column loads.  ld predicates the load per lane (lanes outside the column's levels
// read nothing); its result may stay in flight past later loads, which the pipelined
// tracer kernels rely on.  ldv loads for every lane and selects the default after the
// load; it is safe because the column stride is at least one wave, so every lane
// addresses an element of its own column.
template <class T>
__device__ __forceinline__ T ld(const T *p, long i, bool ok) { return ok ? p[i] : T(0); }
// Streaming accesses: data a wave reads or writes once and no other wave gathers
// (own-column inputs, outputs).  The non-temporal hint keeps them from displacing the
// gathered neighbour columns in the caches.
template <class T>
__device__ __forceinline__ T ld_nt(const T *p, long i, bool ok) { return ok ? __builtin_nontemporal_load(p + i) : T(0); }
template <class T>
__device__ __forceinline__ void st_nt(T *p, long i, T v) { __builtin_nontemporal_store(v, p + i); }
// Two values of one cell and level stored side by side (a gradient pair), so one
// 8- or 16-byte gather per lane fetches both.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <class R> struct Pair2;
template <> struct Pair2<float> { using T = f32x2; };
template <> struct Pair2<double> { using T = f64x2; };
template <class R> using P2 = typename Pair2<R>::T;
template <class R>
__device__ __forceinline__ P2<R> ld2(const P2<R> *p, long i, bool ok) {
  P2<R> z;
  z.x = R(0); z.y = R(0);
  return ok ? p[i] : z;
}
template <class R>
__device__ __forceinline__ P2<R> mk2(R x, R y) {
  P2<R> v;
  v.x = x; v.y = y;
  return v;
}
// Unpredicated column load: a wave-uniform column base and the lane index, so the load
// takes a scalar base and a 32-bit lane offset (no per-lane 64-bit address arithmetic,
// no exec-mask branch).  Every lane reads an element of its own column (the column
// stride is at least one wave); lanes above the column's levels read its padding, so
// use it only where those lanes' results are discarded.
template <class T>
__device__ __forceinline__ T ldl(const T *col, unsigned lane) { return col[lane]; }
template <class T>
__device__ __forceinline__ T ldl_nt(const T *col, unsigned lane) { return __builtin_nontemporal_load(col + lane); }
template <class T>
__device__ __forceinline__ T ldv(const T *p, long i, bool ok, T dflt) {
  const T v = p[i];
  return ok ? v : dflt;
}

// Neighbour slots: slots 0..5 are always processed, so every gather of a column is
// issued without a branch in between (pentagons pad slot 5 with the column itself and
// zero weights, which adds exact zeros); slot 6 exists only for heptagons.
template <class F>
__device__ __forceinline__ void for_slots(int np, F &&f) {
#pragma unroll
  for (int j = 0; j < 6; ++j) f(j);
  if (np > 6) f(6);
}

// ---------------------------------------------------------------- flux limiters
// Shared by the tracer limiters (update_scalar_monot/pd) and the monotone dynamics
// fluxes (fluxes_wrtv_monot): the limiter constants of scalar_transport_rk.F90:14-17
// and prog_wrtv_rk.f90:20-23, and the vertical quadratic of the column.
template <class R>
struct LimConst {
  static constexpr R eps0 = R(1.e-32), eps1 = R(3.e-07);
};

// grad_z_quadratic (grad_lib.f90:166-229) of the column held one level per lane:
// monotone form (mode 1) unbounded (a_wu) with constant-gradient ends, PD form
// (mode 2) bounded (a_w) with zero-gradient ends.
template <class R>
__device__ __forceinline__ void lim_vrec(const DMesh<R> &m, R qc, int k, int mza, int mode, R &sb, R &st) {
  NOCONTRACT
  const int l = k - 1;
  const R qp1 = vshdn(m.tall, qc, 1);
  const bool inner = k >= 2 && k <= mza - 1;
  R ds = inner ? qp1 - qc : R(0);
  const R ds_p1 = vshdn(m.tall, ds, 1), ds_m1 = vshup(m.tall, ds, 1);
  if (mode == 1) {
    if (k == 1) ds = ds_p1;
    if (k == mza) ds = ds_m1;
  }
  const R dsm1 = vshup(m.tall, ds, 1);
  if (mode == 1) {
    sb = qc + ds * m.awu11[l] + dsm1 * m.awu21[l];
    st = qc + ds * m.awu12[l] + dsm1 * m.awu22[l];
  } else {
    sb = qc + ds * m.aw11[l] + dsm1 * m.aw21[l];
    st = qc + ds * m.aw12[l] + dsm1 * m.aw22[l];
  }
}

// kernels_tracer.h -- scalar_transport_rk (forward, adjoint) and its diagnostics.
//
// Reference: omodel/scalar_transport_rk.F90.  One wavefront per W column, lane l
// holds level k = l+1.  The horizontal reconstruction is materialised per stage as
// the gradient pair (gx, gy) of grad_t2d (grad_lib.f90:236-269), or at order 3 the five
// terms of grad_t2d_quadratic (:278-341: gx, gy and the quadratic terms gq[0..2]); the
// update kernel then rebuilds each V-face value from the upwind cell's coefficients,
// exactly the scalar_v_column formula (:474-533), so no V-face array is ever written.  Columns
// whose vertical CFL needs the implicit branch (ff_implic > 0.001, rare) are finished
// by separate kernels that run only when such columns exist.
#pragma once
#include "common.h"

// Register budgets of the hottest tracer kernels, as the minimum number of 256-thread
// blocks per CU (= waves per SIMD) the compiler must allow: 4 gives 128 VGPRs, 5 gives
// 96, 6 gives 80.  Measured choices; a build may override them (-D...) to re-tune.
#ifndef OLAMGPU_BLOCKS_UPDATE64
#define OLAMGPU_BLOCKS_UPDATE64 3
#endif
#ifndef OLAMGPU_BLOCKS_UPDATE
#define OLAMGPU_BLOCKS_UPDATE 4
#endif
#ifndef OLAMGPU_BLOCKS_ADJ01
#define OLAMGPU_BLOCKS_ADJ01 1
#endif

template <class R, class RP>
struct TrStage {
  const R *q_in, *q0, *sct, *gx, *gy;
  const P2<R> *gxy;  // order 2, unlimited stage: the gradient pairs (gx, gy) side by side
  const R *gq[3];    // order 3: the quadratic terms (xx, xy, yy) of the reconstruction
  R *q_out;
  const R *vmasc, *wmasc, *rho_old, *ffimp;
  const RP *rho;
  int ntr, nrk, istage, order, tpw;
  long tstride;
  R dtr;
  const int *tmode;  // final stage: tracers with tmode[n] != 0 go through the limiter
  int lim;           // this stage is a limited final stage
  __device__ __forceinline__ bool limited(int n) const { return lim && tmode[n] != 0; }
};

// ---------------------------------------------------------------- grad_t2d
template <class R, bool PK>
__global__ __launch_bounds__(256) void k_tr_grad(DMesh<R> m, const R *__restrict__ q, R *__restrict__ gx,
                                                 R *__restrict__ gy, P2<R> *__restrict__ gxy, int ntr, long tstride,
                                                 int tpw) {
  const TrItem ti = tr_item(m.nw, ntr, tpw, m.tall);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1;
  const bool act = k <= m.mza, body = act && k >= 2;
  const int np = m.w_npoly[c];
  int nb[7];
  R cx[7], cy[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    nb[j] = m.w_nbr[c * NSLOT + j];
    cx[j] = m.w_gxc[c * NSLOT + j];
    cy[j] = m.w_gyc[c * NSLOT + j];
  }
  const long ldk = m.ldk;
  const unsigned ul = (unsigned)l;
  for (int n = ti.n0; n < ti.n1; ++n) {
    // unpredicated column loads: lanes above mza only feed their own (unstored) results
    const R *qn = q + n * tstride;
    const R qc = ldl(qn + (long)c * ldk, ul);
    R ax = R(0), ay = R(0);
    for_slots(np, [&](int j) {
      const R d = ldl(qn + (long)nb[j] * ldk, ul) - qc;
      ax = ax + cx[j] * d;
      ay = ay + cy[j] * d;
    });
    if (!body) { ax = R(0); ay = R(0); }
    if (act) {
      if constexpr (PK) {
        st_nt(gxy, n * tstride + c * ldk + l, mk2(ax, ay));
      } else {
        st_nt(gx, n * tstride + c * ldk + l, ax);
        st_nt(gy, n * tstride + c * ldk + l, ay);
      }
    }
  }
}

// Output pointers of the order-3 reconstruction (five terms per tracer).
template <class R>
struct Q5 { R *p[5]; };

// ---------------------------------------------------------------- grad_t2d_quadratic
// grad_lib.f90:278-341 for all tracers: the five least-squares terms (x, y, xx, xy,
// yy) from the neighbour differences.  The bounded set (xy_h) serves the unlimited
// stages and the positive-definite final stage (scalar_transport_rk.F90:429-430,
// 1129-1130), the unbounded set (xy_hu) the monotone final stage (:664-665).
template <class R>
__global__ __launch_bounds__(256) void k_tr_grad3(DMesh<R> m, const R *__restrict__ q, Q5<R> g, int ntr,
                                                  long tstride, int tpw, const int *tmode, int lim) {
  const TrItem ti = tr_item(m.nw, ntr, tpw, m.tall);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1;
  const bool act = k <= m.mza, body = act && k >= 2;
  const int np = m.w_npoly[c];
  int nb[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) nb[j] = m.w_nbr[c * NSLOT + j];
  const long ldk = m.ldk;
  for (int n = ti.n0; n < ti.n1; ++n) {
    const int set = (lim && tmode[n] == 1) ? 1 : 0;
    const R *qn = q + n * tstride;
    const R qc = ld(qn, c * ldk + l, act);
    R acc[5] = {R(0), R(0), R(0), R(0), R(0)};
    for_slots(np, [&](int j) {
      const R d = ld(qn, nb[j] * ldk + l, act) - qc;
#pragma unroll
      for (int t = 0; t < 5; ++t) acc[t] = acc[t] + d * m.w_q[set][t][c * NSLOT + j];
    });
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      if (!body) acc[t] = R(0);
      if (act) g.p[t][n * tstride + c * ldk + l] = acc[t];
    }
  }
}

// stage density used by the implicit branch, scalar_transport_rk.F90:139-150
template <class R>
__device__ __forceinline__ R stage_rho(R ro, R rf, int istage, int nrk) {
  if (nrk == 3) {
    if (istage == 1) return R(2.) / R(3.) * ro + R(1.) / R(3.) * rf;
    if (istage == 2) return R(0.5) * (ro + rf);
    return rf;
  }
  return istage == 1 ? R(0.5) * (ro + rf) : rf;
}

// ---------------------------------------------------------------- update_scalar
// One explicit RK stage of all tracers: scalar_v_column + scalar_w_column
// (:474-611), with the explicit weight fexpl = 1 - ff_implic of :578.
template <class R, class RP, bool Q3, bool PK = false, bool TALL = false>
// FP64 instances get the register room of fewer waves (held to the FP32 bounds they
// spilled 38-88 registers per lane; the packed one ran 2.2x faster without)
__global__ __launch_bounds__(256, sizeof(R) == 8 ? (Q3 ? 1 : (PK ? OLAMGPU_BLOCKS_UPDATE64 : 2))
                                                : (Q3 ? 2 : (PK ? OLAMGPU_BLOCKS_UPDATE : 4)))
void k_tr_update(DMesh<R> m, TrStage<R, RP> a) {
  const TrItem ti = tr_item(m.nw, a.ntr, a.tpw, TALL);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;      // W levels with a flux
  const long ldk = m.ldk;
  const long ic = c * ldk + l;
  const int np = m.w_npoly[c];
  constexpr int NQ = Q3 ? 3 : 1;
  int nb[7];
  R dvm[7], dxs[7], dys[7], dxn[7], dyn[7], oqs[7][NQ], oqn[7][NQ], cgx[7], cgy[7];
  bool sel[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    nb[j] = m.w_nbr[c * NSLOT + j];
    if constexpr (PK) { cgx[j] = m.w_gxc[c * NSLOT + j]; cgy[j] = m.w_gyc[c * NSLOT + j]; }
    const int iv = m.w_face[c * NSLOT + j];
    const R dv = m.w_dirv[c * NSLOT + j];
    dxs[j] = m.w_dxs[c * NSLOT + j]; dys[j] = m.w_dys[c * NSLOT + j];
    dxn[j] = m.w_dxn[c * NSLOT + j]; dyn[j] = m.w_dyn[c * NSLOT + j];
    if constexpr (Q3) {
#pragma unroll
      for (int t = 0; t < 3; ++t) { oqs[j][t] = m.w_os[0][t][c * NSLOT + j]; oqn[j][t] = m.w_on[0][t][c * NSLOT + j]; }
    }
    const R vm = (j < 6 || np > 6) ? ld(a.vmasc, (long)iv * ldk + l, act) : R(0);
    dvm[j] = dv * vm;
    // neighbour is upwind: this cell is iw1 (dirv < 0) and vm < 0, or iw2 and vm >= 0
    sel[j] = (dv < R(0)) ? (vm < R(0)) : !(vm < R(0));
  }
  const R wm = ld(a.wmasc, ic, act);
  const R wm_m1 = vshup(TALL, wm, 1);
  const R fexpl = R(1) - a.ffimp[(long)(a.istage - 1) * m.nw + c];
  const R vt = ld(m.volti, ic, act);
  const R rf = (R)ldv(a.rho, ic, act, RP(1));
  const R rinv = R(1) / (body ? rf : R(1));
  const R a11 = m.aw11[l], a21 = m.aw21[l], a12 = m.aw12[l], a22 = m.aw22[l];
  // outside the packed path a neighbour is read only if some level takes that face's
  // value from it
  unsigned need = 0;
#pragma unroll
  for (int j = 0; j < 7; ++j) need |= (__ballot(sel[j] && act) != 0 ? 1u : 0u) << j;
  // the next tracer's loads are issued before the current one is computed
  struct Ld { R qc, q0, st, gxc, gyc, qn[7], gxn[7], gyn[7], gqc[NQ], gqn[7][NQ]; };
  auto load = [&](int n, Ld &L) {
    const long off = n * a.tstride;
    const R *qi = a.q_in + off;
    if constexpr (PK) {   // unpredicated: the vertical terms at mza select away lane mza
      const long cb = off + (long)c * ldk;
      L.qc = ldl(a.q_in + cb, (unsigned)l);
      L.q0 = ldl_nt(a.q0 + cb, (unsigned)l);
      L.st = a.sct ? ldl_nt(a.sct + cb, (unsigned)l) : R(0);
    } else {
      L.qc = ld(qi, ic, act);
      L.q0 = ld_nt(a.q0 + off, ic, act);
      L.st = a.sct ? ld_nt(a.sct + off, ic, act) : R(0);
    }
    L.gxc = R(0); L.gyc = R(0);
    if constexpr (PK) {
      // the column's own gradient is rebuilt from the neighbour values below (every
      // neighbour is read) instead of being streamed from the gradient array
    } else if (a.order > 1) {
      L.gxc = ld(a.gx + off, ic, act); L.gyc = ld(a.gy + off, ic, act);
    }
    if constexpr (Q3) {
#pragma unroll
      for (int t = 0; t < 3; ++t) L.gqc[t] = ld(a.gq[t] + off, ic, act);
    }
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      L.qn[j] = R(0); L.gxn[j] = R(0); L.gyn[j] = R(0);
      if constexpr (Q3) {
#pragma unroll
        for (int t = 0; t < 3; ++t) L.gqn[j][t] = R(0);
      }
      if constexpr (PK) {   // values of lanes above mza only reach discarded lanes
        if (j < 6 || np > 6) L.qn[j] = ldl(qi + (long)nb[j] * ldk, (unsigned)l);
      }
      // the packed path gathers every neighbour's gradient pair: skipping the faces no
      // level takes from the neighbour costs more in branches and held masks than the
      // cache reads it saves (4 % of the kernel)
      if ((j < 6 || np > 6) && (PK || ((need >> j) & 1u))) {
        const long in = (long)nb[j] * ldk + l;
        if constexpr (!PK) L.qn[j] = ld(qi, in, act);
        if constexpr (PK) {
          const P2<R> g = ldl(a.gxy + off + (long)nb[j] * ldk, (unsigned)l);
          L.gxn[j] = g.x; L.gyn[j] = g.y;
        } else if (a.order > 1) {
          L.gxn[j] = ld(a.gx + off, in, act);
          L.gyn[j] = ld(a.gy + off, in, act);
          if constexpr (Q3) {
#pragma unroll
            for (int t = 0; t < 3; ++t) L.gqn[j][t] = ld(a.gq[t] + off, in, act);
          }
        }
      }
    }
  };
  Ld nx;
  if (ti.n0 < ti.n1) load(ti.n0, nx);
  for (int n = ti.n0; n < ti.n1; ++n) {
    const Ld L = nx;
    if (n + 1 < ti.n1) load(n + 1, nx);
    if (a.limited(n)) continue;
    const long off = n * a.tstride;
    const R qc = L.qc, q0 = L.q0;
    R gxc = L.gxc, gyc = L.gyc;
    if constexpr (PK) {   // grad_t2d of the column, the arithmetic and order of k_tr_grad
      R ax = R(0), ay = R(0);
      for_slots(np, [&](int j) {
        const R d = L.qn[j] - qc;
        ax = ax + cgx[j] * d;
        ay = ay + cgy[j] * d;
      });
      gxc = body ? ax : R(0);
      gyc = body ? ay : R(0);
    }
    R F = R(0);
    for_slots(np, [&](int j) {
      R fs = qc, fn = L.qn[j];
      if (a.order > 1) {
        fs = qc + dxs[j] * gxc + dys[j] * gyc;
        fn = L.qn[j] + dxn[j] * L.gxn[j] + dyn[j] * L.gyn[j];
      }
      if constexpr (Q3) {   // bounded offsets xx0_v (:515-524)
        fs = fs + oqs[j][0] * L.gqc[0] + oqs[j][1] * L.gqc[1] + oqs[j][2] * L.gqc[2];
        fn = fn + oqn[j][0] * L.gqn[j][0] + oqn[j][1] * L.gqn[j][1] + oqn[j][2] * L.gqn[j][2];
      }
      const R f = sel[j] ? fn : fs;
      F = F + dvm[j] * (f - q0);
    });
    // vertical: grad_z_quadratic (bounded, zero-gradient ends) and upwinding
    const R qp1 = vshdn(TALL, qc, 1);
    const R ds = inner ? qp1 - qc : R(0);
    const R dsm1 = vshup(TALL, ds, 1);
    const R scpb = qc + ds * a11 + dsm1 * a21;
    const R scpt = qc + ds * a12 + dsm1 * a22;
    const R scpb_p1 = vshdn(TALL, scpb, 1);
    R upw = (wm < R(0)) ? scpb_p1 : scpt;
    if (k == 1) upw = qp1;
    if (k == mza) upw = qc;
    const R upw_m1 = vshup(TALL, upw, 1);
    F = F + fexpl * (wm_m1 * (upw_m1 - q0) - wm * (upw - q0));
    R qn_new = q0 + a.dtr * (L.st + vt * F) * rinv;
    const R lev1 = vshdn(TALL, qn_new, 1);             // level 1 copies level 2
    if (k == 1) qn_new = lev1;
    if (act) st_nt(a.q_out, off + ic, qn_new);
  }
}

// ---------------------------------------------------------------- fused stage
// One explicit RK stage of all tracers at order 1 or 2 with the reconstruction computed
// where it is used instead of materialised: grad_t2d (grad_lib.f90:236-269) of the
// column from its neighbours, and for every face whose value some level takes from the
// neighbour (a wave ballot), grad_t2d of that neighbour from the neighbour's own
// stencil (its record read with scalar loads, its neighbours' columns gathered).  The
// gradients are formed with the expressions of k_tr_grad and the update with those of
// k_tr_update, so the results match the two-pass arrangement; no gradient array is
// written or read back.  Per tracer the HBM traffic is the stage input, scp0, sct and
// the output.
template <class R, class RP>
__global__ __launch_bounds__(256, 4) void k_tr_stage(DMesh<R> m, TrStage<R, RP> a) {
  const TrItem ti = tr_item(m.nw, a.ntr, a.tpw, m.tall);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk;
  const long ic = c * ldk + l;
  const int np = m.w_npoly[c];
  const bool o2 = a.order > 1;
  int nb[7];
  R dvm[7];
  bool sel[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    nb[j] = m.w_nbr[c * NSLOT + j];
    const int iv = m.w_face[c * NSLOT + j];
    const R dv = m.w_dirv[c * NSLOT + j];
    const R vm = (j < 6 || np > 6) ? ld(a.vmasc, (long)iv * ldk + l, act) : R(0);
    dvm[j] = dv * vm;
    // neighbour is upwind: this cell is iw1 (dirv < 0) and vm < 0, or iw2 and vm >= 0
    sel[j] = (dv < R(0)) ? (vm < R(0)) : !(vm < R(0));
  }
  const R wm = ld(a.wmasc, ic, act);
  const R wm_m1 = vshup(m.tall, wm, 1);
  const R fexpl = R(1) - a.ffimp[(long)(a.istage - 1) * m.nw + c];
  const R vt = ld(m.volti, ic, act);
  const R rf = (R)ldv(a.rho, ic, act, RP(1));
  const R rinv = R(1) / (body ? rf : R(1));
  const R a11 = m.aw11[l], a21 = m.aw21[l], a12 = m.aw12[l], a22 = m.aw22[l];
  unsigned need = 0;
#pragma unroll
  for (int j = 0; j < 7; ++j) need |= (__ballot(sel[j] && act) != 0 ? 1u : 0u) << j;
  if (!o2) need = 0;
  // the next tracer's column and neighbour values are loaded before the current one is
  // computed
  struct Ld { R qc, q0, st, qn[7]; };
  auto load = [&](int n, Ld &L) {
    const long off = n * a.tstride;
    const R *qi = a.q_in + off;
    L.qc = ld(qi, ic, act);
    L.q0 = ld_nt(a.q0 + off, ic, act);
    L.st = a.sct ? ld_nt(a.sct + off, ic, act) : R(0);
#pragma unroll
    for (int j = 0; j < 7; ++j) L.qn[j] = (j < 6 || np > 6) ? ld(qi, (long)nb[j] * ldk + l, act) : R(0);
  };
  Ld nx;
  if (ti.n0 < ti.n1) load(ti.n0, nx);
  for (int n = ti.n0; n < ti.n1; ++n) {
    const Ld L = nx;
    if (n + 1 < ti.n1) load(n + 1, nx);
    if (a.limited(n)) continue;
    const long off = n * a.tstride;
    const R *qi = a.q_in + off;
    const R qc = L.qc, q0 = L.q0;
    R gxc = R(0), gyc = R(0);
    if (o2) {   // grad_t2d of the column (k_tr_grad's expressions)
      R ax = R(0), ay = R(0);
      for_slots(np, [&](int j) {
        const R d = L.qn[j] - qc;
        ax = ax + m.w_gxc[c * NSLOT + j] * d;
        ay = ay + m.w_gyc[c * NSLOT + j] * d;
      });
      if (body) { gxc = ax; gyc = ay; }
    }
    R F = R(0);
    for_slots(np, [&](int j) {
      R fs = qc, fn = L.qn[j];
      if (o2) {
        fs = qc + m.w_dxs[c * NSLOT + j] * gxc + m.w_dys[c * NSLOT + j] * gyc;
        if ((need >> j) & 1u) {   // wave-uniform: grad_t2d of the upwind neighbour
          const int u = nb[j];
          const int npu = m.w_npoly[u];
          R bx = R(0), by = R(0);
          for_slots(npu, [&](int i) {
            const R d = ld(qi, (long)m.w_nbr[u * NSLOT + i] * ldk + l, act) - fn;
            bx = bx + m.w_gxc[u * NSLOT + i] * d;
            by = by + m.w_gyc[u * NSLOT + i] * d;
          });
          if (!body) { bx = R(0); by = R(0); }
          fn = fn + m.w_dxn[c * NSLOT + j] * bx + m.w_dyn[c * NSLOT + j] * by;
        }
      }
      const R f = sel[j] ? fn : fs;
      F = F + dvm[j] * (f - q0);
    });
    // vertical: grad_z_quadratic (bounded, zero-gradient ends) and upwinding
    const R qp1 = vshdn(m.tall, qc, 1);
    const R ds = inner ? qp1 - qc : R(0);
    const R dsm1 = vshup(m.tall, ds, 1);
    const R scpb = qc + ds * a11 + dsm1 * a21;
    const R scpt = qc + ds * a12 + dsm1 * a22;
    const R scpb_p1 = vshdn(m.tall, scpb, 1);
    R upw = (wm < R(0)) ? scpb_p1 : scpt;
    if (k == 1) upw = qp1;
    if (k == mza) upw = qc;
    const R upw_m1 = vshup(m.tall, upw, 1);
    F = F + fexpl * (wm_m1 * (upw_m1 - q0) - wm * (upw - q0));
    R qn_new = q0 + a.dtr * (L.st + vt * F) * rinv;
    const R lev1 = vshdn(m.tall, qn_new, 1);             // level 1 copies level 2
    if (k == 1) qn_new = lev1;
    if (act) st_nt(a.q_out, off + ic, qn_new);
  }
}

// ---------------------------------------------------------------- LDS-staged stage
// Patch tables (built on the host): see Patches in olamgpu.hip.
struct PatchTab {
  const int *cells, *cnt;
  const uint16_t *nbr;
  int pmax, rmax;
};

// Stencil record of one patch cell, staged in LDS once per block: the patch-local
// indices of its NSLOT neighbour slots, its grad_t2d weights and npoly.  Laid out so
// that each group is one or two 128-bit LDS reads.
template <class R>
struct __align__(16) PCell {
  uint16_t nl[NSLOT];
  R cx[NSLOT], cy[NSLOT];
  int np, pad_[3];
};

// Fill the stencil records of the owned and ring-1 cells of patch p.
template <class R>
__device__ __forceinline__ void stage_pcells(const DMesh<R> &m, const PatchTab &pt, int p, int nr, PCell<R> *sc) {
  const int *cells = pt.cells + (long)p * pt.pmax;
  const uint16_t *lnb = pt.nbr + (long)p * pt.rmax * NSLOT;
  for (int e = threadIdx.x; e < nr * NSLOT; e += blockDim.x) {
    const int i = e / NSLOT, j = e - i * NSLOT;
    const int g = cells[i];
    sc[i].nl[j] = lnb[e];
    sc[i].cx[j] = m.w_gxc[g * NSLOT + j];
    sc[i].cy[j] = m.w_gyc[g * NSLOT + j];
    if (j == 0) sc[i].np = m.w_npoly[g];
  }
}

// Register copy of a stencil record (wide LDS reads of a wave-uniform address).
template <class R>
struct PRec {
  int nl[NSLOT];
  R cx[NSLOT], cy[NSLOT];
  int np;
  __device__ __forceinline__ void load(const PCell<R> &c) {
    const uint4 v = *reinterpret_cast<const uint4 *>(c.nl);
    nl[0] = v.x & 0xffff; nl[1] = v.x >> 16; nl[2] = v.y & 0xffff; nl[3] = v.y >> 16;
    nl[4] = v.z & 0xffff; nl[5] = v.z >> 16; nl[6] = v.w & 0xffff; nl[7] = v.w >> 16;
#pragma unroll
    for (int j = 0; j < NSLOT; ++j) { cx[j] = c.cx[j]; cy[j] = c.cy[j]; }
    np = c.np;
  }
};

// One explicit RK stage of all tracers for one patch of columns, with the gradients
// computed in the block instead of materialised.  The block stages the stencil
// records of its owned and ring-1 cells in LDS once, and each wave keeps the
// tracer-independent data of its CPW owned columns (face mass fluxes, upwind choices,
// wmasc, volti, 1/rho) in registers.  Per tracer, the stage-input column of every
// patch cell (owned, ring 1, ring 2) is staged in LDS, one row of 64 levels per cell;
// each face value is then rebuilt from LDS: the column's own grad_t2d for outflow
// faces and the upwind neighbour's grad_t2d (from its neighbours, ring 2 at the patch
// edge) for inflow faces, skipped when no level of the face takes it.  The next
// tracer's rows and scp0 are loaded into registers while the current one is computed.
// The arithmetic and its order are those of k_tr_grad + k_tr_update.
template <class R, class RP, int CPW, int RPW>
__global__ __launch_bounds__(512) void k_tr_patch(DMesh<R> m, TrStage<R, RP> a, PatchTab pt) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int p = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x));
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int NWV = 8;
  const int l = lane_id(), k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk;
  const int n_own = pt.cnt[p * 4], n_r1 = pt.cnt[p * 4 + 1], n_tot = pt.cnt[p * 4 + 2];
  const int *cells = pt.cells + (long)p * pt.pmax;
  PCell<R> *sc = reinterpret_cast<PCell<R> *>(smem_raw);
  R *sq = reinterpret_cast<R *>(smem_raw + (size_t)pt.rmax * sizeof(PCell<R>));
  stage_pcells(m, pt, p, n_own + n_r1, sc);
  const R a11 = m.aw11[l], a21 = m.aw21[l], a12 = m.aw12[l], a22 = m.aw22[l];
  // tracer-independent data of this wave's owned columns
  R dvm[CPW][7], wmr[CPW], vtr[CPW], rinvr[CPW], fexr[CPW];
  unsigned selm[CPW];
  long icr[CPW];
#pragma unroll
  for (int t = 0; t < CPW; ++t) {
    const int i = w + t * NWV;
    selm[t] = 0; wmr[t] = R(0); vtr[t] = R(0); rinvr[t] = R(1); fexr[t] = R(0); icr[t] = l;
#pragma unroll
    for (int j = 0; j < 7; ++j) dvm[t][j] = R(0);
    if (i < n_own) {
      const int c = cells[i];
      const long ic = (long)c * ldk + l;
      icr[t] = ic;
      const int np = m.w_npoly[c];
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        const int iv = m.w_face[c * NSLOT + j];
        const R dv = m.w_dirv[c * NSLOT + j];
        const R vm = (j < 6 || np > 6) ? ld(a.vmasc, (long)iv * ldk + l, act) : R(0);
        dvm[t][j] = dv * vm;
        const bool sel = (dv < R(0)) ? (vm < R(0)) : !(vm < R(0));
        selm[t] |= (sel ? 1u : 0u) << j;
      }
      wmr[t] = ld(a.wmasc, ic, act);
      fexr[t] = R(1) - a.ffimp[(long)(a.istage - 1) * m.nw + c];
      vtr[t] = ld(m.volti, ic, act);
      const R rf = (R)ldv(a.rho, ic, act, RP(1));
      rinvr[t] = R(1) / (body ? rf : R(1));
    }
  }
  // prefetch of tracer 0
  R rq[RPW], q0r[CPW], str[CPW];
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int i = w + r * NWV;
    rq[r] = (i < n_tot) ? ld(a.q_in, (long)cells[i] * ldk + l, act) : R(0);
  }
#pragma unroll
  for (int t = 0; t < CPW; ++t) {
    const bool own = w + t * NWV < n_own;
    q0r[t] = own ? ld(a.q0, icr[t], act) : R(0);
    str[t] = (own && a.sct) ? ld(a.sct, icr[t], act) : R(0);
  }
  for (int n = 0; n < a.ntr; ++n) {
    const long off = n * a.tstride;
    __syncthreads();                                  // records staged / previous tracer done
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const int i = w + r * NWV;
      if (i < n_tot) sq[i * WAVE + l] = rq[r];
    }
    __syncthreads();
    R q0c[CPW], stc[CPW];
#pragma unroll
    for (int t = 0; t < CPW; ++t) { q0c[t] = q0r[t]; stc[t] = str[t]; }
    if (n + 1 < a.ntr) {                              // next tracer, in flight during compute
      const long offn = off + a.tstride;
#pragma unroll
      for (int r = 0; r < RPW; ++r) {
        const int i = w + r * NWV;
        rq[r] = (i < n_tot) ? ld(a.q_in + offn, (long)cells[i] * ldk + l, act) : R(0);
      }
#pragma unroll
      for (int t = 0; t < CPW; ++t) {
        const bool own = w + t * NWV < n_own;
        q0r[t] = own ? ld(a.q0 + offn, icr[t], act) : R(0);
        str[t] = (own && a.sct) ? ld(a.sct + offn, icr[t], act) : R(0);
      }
    }
#pragma unroll
    for (int t = 0; t < CPW; ++t) {
      const int i = w + t * NWV;
      if (i >= n_own) continue;
      const int c = cells[i];
      PRec<R> me;
      me.load(sc[i]);
      const int np = me.np;
      const R qc = sq[i * WAVE + l];
      const R q0 = q0c[t];
      R gxc = R(0), gyc = R(0);
      if (a.order > 1) {
        R ax = R(0), ay = R(0);
        for_slots(np, [&](int j) {
          const R d = sq[me.nl[j] * WAVE + l] - qc;
          ax = ax + me.cx[j] * d;
          ay = ay + me.cy[j] * d;
        });
        if (body) { gxc = ax; gyc = ay; }
      }
      R F = R(0);
      for_slots(np, [&](int j) {
        const int u = me.nl[j];
        const R qn = sq[u * WAVE + l];
        const bool sel = (selm[t] >> j) & 1u;
        R fs = qc, fn = qn;
        if (a.order > 1) {
          fs = qc + m.w_dxs[c * NSLOT + j] * gxc + m.w_dys[c * NSLOT + j] * gyc;
          if (__ballot(sel && act) != 0) {
            PRec<R> nb;
            nb.load(sc[u]);
            R bx = R(0), by = R(0);
            for_slots(nb.np, [&](int i2) {
              const R d = sq[nb.nl[i2] * WAVE + l] - qn;
              bx = bx + nb.cx[i2] * d;
              by = by + nb.cy[i2] * d;
            });
            if (!body) { bx = R(0); by = R(0); }
            fn = qn + m.w_dxn[c * NSLOT + j] * bx + m.w_dyn[c * NSLOT + j] * by;
          }
        }
        const R f = sel ? fn : fs;
        F = F + dvm[t][j] * (f - q0);
      });
      const R wm = wmr[t];
      const R wm_m1 = vshup(m.tall, wm, 1);
      const R qp1 = vshdn(m.tall, qc, 1);
      const R ds = inner ? qp1 - qc : R(0);
      const R dsm1 = vshup(m.tall, ds, 1);
      const R scpb = qc + ds * a11 + dsm1 * a21;
      const R scpt = qc + ds * a12 + dsm1 * a22;
      const R scpb_p1 = vshdn(m.tall, scpb, 1);
      R upw = (wm < R(0)) ? scpb_p1 : scpt;
      if (k == 1) upw = qp1;
      if (k == mza) upw = qc;
      const R upw_m1 = vshup(m.tall, upw, 1);
      F = F + fexr[t] * (wm_m1 * (upw_m1 - q0) - wm * (upw - q0));
      R qn_new = q0 + a.dtr * (stc[t] + vtr[t] * F) * rinvr[t];
      const R lev1 = vshdn(m.tall, qn_new, 1);
      if (k == 1) qn_new = lev1;
      if (act) a.q_out[off + icr[t]] = qn_new;
    }
  }
}

// Implicit vertical transport of one stage for columns with ff_implic > 0.001:
// the tridif_prep/tridif_fini pair (tridiag_lib.f90:321-364) applied after the
// explicit update (scalar_transport_rk.F90:186-217, :347-375), solved in FP64 by
// PCR.  Level 1 keeps the explicit value, as in the reference.
template <class R, class RP>
__global__ __launch_bounds__(256) void k_tr_implicit(DMesh<R> m, TrStage<R, RP> a) {
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const R fimpl = a.ffimp[(long)(a.istage - 1) * m.nw + c];
  if (!(fimpl > R(0.001))) return;
  const int l = wp_.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk, ic = c * ldk + l;
  const R wm = ld(a.wmasc, ic, act);
  const R wm_m1 = vshup(m.tall, wm, 1);
  const R vt = ld(m.volti, ic, act);
  const R rf = (R)ldv(a.rho, ic, act, RP(1));
  const R ro = ld(a.rho_old, ic, act);
  const R rs = stage_rho(ro, rf, a.istage, a.nrk);
  const R wup = inner ? (wm > R(0) ? wm : R(0)) : R(0);
  const R wdn = inner ? (wm < R(0) ? wm : R(0)) : R(0);
  const R wup_m1 = vshup(m.tall, wup, 1), wdn_m1 = vshup(m.tall, wdn, 1);
  const R c2 = fimpl * a.dtr * vt / (body ? rs : R(1));
  double ib1 = 0.0, ib2 = 1.0, ib3 = 0.0;
  if (body) {
    ib1 = (double)(-c2 * wup_m1);
    ib2 = (double)(R(1) + c2 * (wup - wdn_m1));
    ib3 = (double)(c2 * wdn);
  }
  for (int n = 0; n < a.ntr; ++n) {
    const long off = n * a.tstride;
    const R q0 = ld(a.q0 + off, ic, act);
    const R qe = ld(a.q_out + off, ic, act);
    const R q0p1 = vshdn(m.tall, q0, 1);
    const R dq0 = inner ? q0p1 - q0 : R(0);
    const R dq0_m1 = vshup(m.tall, dq0, 1);
    const R delex = qe - q0 - c2 * ((wm < R(0) ? wm : R(0)) * dq0 + (wm_m1 > R(0) ? wm_m1 : R(0)) * dq0_m1);
    const double del = vpcr(m.tall, l, ib1, ib2, ib3, body ? (double)delex : 0.0);
    if (body) a.q_out[off + ic] = q0 + (R)del;
  }
}

// ---------------------------------------------------------------- check_cfls
// scalar_transport_rk.F90:1476-1712: implicit fractions per stage; with report != 0
// also the maxima of total CFL, horizontal CFL and |wc| (unsigned bit patterns of
// non-negative floats, one atomic per block) and the number of implicit columns.
// any_imp is set (plain store) when some column needs the implicit branch.
template <class R>
__global__ __launch_bounds__(256) void k_tr_cfl(DMesh<R> m, const R *vmasc, const R *wmasc, const R *rho_o,
                                                const R *wc, R dt, int nrk, R *ffimp, unsigned *maxbits,
                                                int *nimp, int *any_imp, int report) {
  __shared__ float sh[WPB][4];
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  const int w = threadIdx.x >> 6;
  const int l = wp_.l, k = l + 1, mza = m.mza;
  float v0 = 0.f, v1 = 0.f, wmx = 0.f;
  bool any = false;
  if (c < m.nw) {
    const bool act = k <= mza, body = act && k >= 2;
    const long ldk = m.ldk, ic = c * ldk + l;
    const int np = m.w_npoly[c];
    R cflh = R(0);
    for_slots(np, [&](int j) {
      const int iv = m.w_face[c * NSLOT + j];
      const R dv = m.w_dirv[c * NSLOT + j];
      const R x = dv * ld(vmasc, (long)iv * ldk + l, act);
      cflh = cflh - (x < R(0) ? x : R(0));
    });
    const R wm = ld(wmasc, ic, act), wm_m1 = vshup(m.tall, wm, 1);
    R fimp[3] = {R(0), R(0), R(0)};
    R cfl = R(0);
    if (body) {
      const R dovr = dt * m.volti[ic] / rho_o[ic];
      cflh = cflh * dovr;
      R cflw = ((wm > R(0) ? wm : R(0)) - (wm_m1 < R(0) ? wm_m1 : R(0))) * dovr;
      cfl = cflh + cflw;
      cflw = cflw > R(0.1) ? cflw : R(0.1);
      fimp[nrk - 1] = (cfl - R(0.98)) / cflw;
      fimp[nrk - 2] = (cfl - R(1.98)) / cflw;
      if (nrk == 3) fimp[0] = (cfl - R(2.98)) / cflw;
    } else {
      cflh = R(0);
      fimp[0] = fimp[1] = fimp[2] = R(-1e30);
    }
    for (int n = 0; n < nrk; ++n) {
      R v = fimp[n];
      for (int s = 32; s >= 1; s >>= 1) { R o = __shfl_xor(v, s, WAVE); v = v > o ? v : o; }
      if (m.tall) {                    // the maximum over both halves of the column
        const R o = vshdn(m.tall, v, WAVE), p = vshup(m.tall, v, WAVE);
        v = v > o ? v : o;
        v = v > p ? v : p;
      }
      v = v > R(0) ? v : R(0);         // fimp starts at 0 (:1558)
      R f = v < R(1) ? v : R(1);
      if (f < R(0.02)) f = R(0);
      if (l == 0) ffimp[(long)n * m.nw + c] = f;
      any = any || f > R(0.001);
    }
    if (report) {
      wmx = body ? fabsf((float)wc[ic]) : 0.f;
      v0 = (float)cfl;
      v1 = (float)cflh;
      for (int s = 32; s >= 1; s >>= 1) {
        v0 = fmaxf(v0, __shfl_xor(v0, s, WAVE));
        v1 = fmaxf(v1, __shfl_xor(v1, s, WAVE));
        wmx = fmaxf(wmx, __shfl_xor(wmx, s, WAVE));
      }
    }
  }
  if (any && l == 0) *any_imp = 1;
  if (!report) return;
  // each column counted once (by the wave holding level 1)
  if (lane_id() == 0) { sh[w][0] = v0; sh[w][1] = v1; sh[w][2] = wmx; sh[w][3] = (any && l == 0) ? 1.f : 0.f; }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    int cnt = 0;
    for (int i = 0; i < WPB; ++i) {
      a0 = fmaxf(a0, sh[i][0]); a1 = fmaxf(a1, sh[i][1]); a2 = fmaxf(a2, sh[i][2]);
      cnt += sh[i][3] > 0.f;
    }
    atomicMax(&maxbits[0], __float_as_uint(fmaxf(a0, 0.f)));
    atomicMax(&maxbits[1], __float_as_uint(fmaxf(a1, 0.f)));
    atomicMax(&maxbits[2], __float_as_uint(a2));
    if (cnt) atomicAdd(nimp, cnt);
  }
}

// ---------------------------------------------------------------- adjoint
// Transpose of k_tr_update (+ k_tr_implicit) for fixed winds (upwind choices
// frozen).  Per stage, in reverse order:
//   k_tr_adj_implicit (implicit columns only): mu -> T^-T mu, with the scp0 terms of
//      the implicit branch into lam0;
//   A0 (per column): nu = D mu, the scp0 terms into lam0 and the transpose of the
//      vertical reconstruction -> lpart;
//   A1 (per column): face adjoint fluxes phi = vm (nu(iw2) - nu(iw1)) routed to the
//      upwind side: lpart += phi, (Gx, Gy) = sum (dxps, dyps) phi;
//   A2 (per column): transpose of grad_t2d: lq = lpart + sum over neighbours of
//      their (Gx, Gy) weighted by the coefficient that points back to this column.
template <class R, class RP>
struct TrAdj {
  R *mu;             // adjoint of the stage output (overwritten by the implicit pass)
  R *lam0;           // accumulates the adjoint of scp0
  R *nu, *lpart, *lgx, *lgy;
  R *lgq[3];         // order 3: adjoints of the quadratic terms
  R *lq;             // adjoint of the stage input
  const R *vmasc, *wmasc, *rho_old, *ffimp;
  const RP *rho;
  int ntr, nrk, istage, order, last, tpw;
  int first;         // the first adjoint stage (istage = nrk): lam0 starts at zero
  long tstride;
  R dtr;
  const int *tmode;  // as in TrStage
  int lim;
  __device__ __forceinline__ bool limited(int n) const { return lim && tmode[n] != 0; }
};

template <class R, class RP>
__global__ __launch_bounds__(256) void k_tr_adj_implicit(DMesh<R> m, TrAdj<R, RP> a) {
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const R fimpl = a.ffimp[(long)(a.istage - 1) * m.nw + c];
  if (!(fimpl > R(0.001))) return;
  const int l = wp_.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk, ic = c * ldk + l;
  const R wm = ld(a.wmasc, ic, act), wm_m1 = vshup(m.tall, wm, 1);
  const R vt = ld(m.volti, ic, act);
  const R rf = (R)ldv(a.rho, ic, act, RP(1));
  const R ro = ld(a.rho_old, ic, act);
  const R rs = stage_rho(ro, rf, a.istage, a.nrk);
  const R wup = inner ? (wm > R(0) ? wm : R(0)) : R(0);
  const R wdn = inner ? (wm < R(0) ? wm : R(0)) : R(0);
  const R wup_m1 = vshup(m.tall, wup, 1), wdn_m1 = vshup(m.tall, wdn, 1);
  const R c2 = fimpl * a.dtr * vt / (body ? rs : R(1));
  const R b1 = body ? -c2 * wup_m1 : R(0);
  const R b2 = body ? R(1) + c2 * (wup - wdn_m1) : R(1);
  const R b3 = body ? c2 * wdn : R(0);
  // transposed system: sub-diagonal b3(k-1), diagonal b2(k), super-diagonal b1(k+1)
  const R b3_m1 = vshup(m.tall, b3, 1), b1_p1 = vshdn(m.tall, b1, 1);
  double tb1 = 0.0, tb2 = 1.0, tb3 = 0.0;
  if (body) { tb1 = (double)b3_m1; tb2 = (double)b2; tb3 = (double)(k < mza ? b1_p1 : R(0)); }
  for (int n = 0; n < a.ntr; ++n) {
    const long off = n * a.tstride;
    const bool limn = a.limited(n);
    R mu = ld(a.mu + off, ic, act);
    const R mu1 = vlane0(m.tall, mu);
    if (k == 1) mu = R(0);
    R l0 = a.first ? R(0) : ld(a.lam0 + off, ic, act);
    const double le = vpcr(m.tall, l, tb1, tb2, tb3, body ? (double)mu : 0.0);
    R lef = body ? (R)le : R(0);
    l0 = l0 + mu - lef;
    // -c2 (min(wm(k),0) dq0(k) + max(wm(k-1),0) dq0(k-1)),  dq0(k) = q0(k+1) - q0(k)
    const R av = -c2 * lef;
    const R t1 = inner ? av * (wm < R(0) ? wm : R(0)) : R(0);
    const R t2 = (body && k - 1 >= 2) ? av * (wm_m1 > R(0) ? wm_m1 : R(0)) : R(0);
    const R t1_m1 = vshup(m.tall, t1, 1), t2_p1 = vshdn(m.tall, t2, 1);
    if (body) l0 = l0 - t1 + (k - 1 >= 2 ? t1_m1 : R(0)) + t2 - (k + 1 <= mza ? t2_p1 : R(0));
    // level 1 kept the explicit level-2 value, so its adjoint joins that of the explicit
    // result at level 2 (on a limited stage level 1 is the stage input: its adjoint
    // stays in place for the limited explicit pass)
    if (k == 2 && !limn) lef = lef + mu1;
    if (k == 1 && limn) lef = mu1;
    if (act) {
      a.lam0[off + ic] = l0;
      a.mu[off + ic] = lef;                   // level 1: 0, or mu(1) on a limited stage
    }
  }
}

template <class R, class RP>
__global__ __launch_bounds__(256) void k_tr_adj0(DMesh<R> m, TrAdj<R, RP> a) {
  const TrItem ti = tr_item(m.nw, a.ntr, a.tpw, m.tall);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk, ic = c * ldk + l;
  const int np = m.w_npoly[c];
  R csum = R(0);
  for_slots(np, [&](int j) {
    const int iv = m.w_face[c * NSLOT + j];
    const R dv = m.w_dirv[c * NSLOT + j];
    csum = csum + dv * ld(a.vmasc, (long)iv * ldk + l, act);
  });
  const R wm = ld(a.wmasc, ic, act), wm_m1 = vshup(m.tall, wm, 1);
  const R fexpl = R(1) - a.ffimp[(long)(a.istage - 1) * m.nw + c];
  const R vt = ld(m.volti, ic, act);
  const R rf = (R)ldv(a.rho, ic, act, RP(1));
  const R rinv = R(1) / (body ? rf : R(1));
  const R cc = csum + fexpl * (wm_m1 - wm);
  const R a11 = m.aw11[l], a12 = m.aw12[l];
  const R a21_p1 = vshdn(m.tall, m.aw21[l], 1), a22_p1 = vshdn(m.tall, m.aw22[l], 1);
  const bool neg = wm < R(0);
  const bool neg_m1 = vshup(m.tall, (int)neg, 1) != 0, neg_p1 = vshdn(m.tall, (int)neg, 1) != 0;
  // the next tracer's loads are issued before the current one is computed (all of a
  // tracer's loads are in the prefetch, so no later load waits on them)
  R nmu = R(0), nl0 = R(0);
  // on the first adjoint stage lam0 holds only what the implicit pass of this column wrote
  const bool rd0 = !a.first || a.ffimp[(long)(a.istage - 1) * m.nw + c] > R(0.001);
  if (ti.n0 < ti.n1) { nmu = ld(a.mu + ti.n0 * a.tstride, ic, act); nl0 = ld(a.lam0 + ti.n0 * a.tstride, ic, act && rd0); }
  for (int n = ti.n0; n < ti.n1; ++n) {
    R mu = nmu, l0 = nl0;
    if (n + 1 < ti.n1) {
      nmu = ld(a.mu + (n + 1) * a.tstride, ic, act);
      nl0 = ld(a.lam0 + (n + 1) * a.tstride, ic, act && rd0);
    }
    if (a.limited(n)) continue;
    const long off = n * a.tstride;
    const R mu1 = vlane0(m.tall, mu);
    if (k == 1) mu = R(0);
    if (k == 2) mu = mu + mu1;                 // level 1 copied level 2 (explicit)
    const R nu = body ? mu * a.dtr * vt * rinv : R(0);
    if (body) l0 = l0 + mu - nu * cc;
    // vertical transpose: upw(j) feeds F(j+1) with +fexpl wm(j), F(j) with -fexpl wm(j)
    const R nu_p1 = vshdn(m.tall, nu, 1);
    const R lupw = inner ? fexpl * wm * ((k + 1 <= mza ? nu_p1 : R(0)) - nu) : R(0);
    const R lupw_m1 = vshup(m.tall, lupw, 1), lupw_p1 = vshdn(m.tall, lupw, 1);
    R lqv = R(0), lds = R(0);
    if (body) {
      if (inner && !neg) lqv = lqv + lupw;
      if (k - 1 >= 2 && neg_m1) lqv = lqv + lupw_m1;
      // lds(k): from W level k-1 (neg) via aw11(k), W level k (neg) via aw21(k+1),
      // W level k (not neg) via aw12(k), W level k+1 (not neg) via aw22(k+1)
      if (k - 1 >= 2 && neg_m1) lds = lds + lupw_m1 * a11;
      if (inner && neg) lds = lds + lupw * a21_p1;
      if (inner && !neg) lds = lds + lupw * a12;
      if (k + 1 <= mza - 1 && !neg_p1) lds = lds + lupw_p1 * a22_p1;
    }
    if (!inner) lds = R(0);
    const R lds_m1 = vshup(m.tall, lds, 1);
    if (body) lqv = lqv + (k - 1 >= 2 ? lds_m1 : R(0)) - lds;
    if (act) {
      a.lam0[off + ic] = l0;
      a.nu[off + ic] = nu;
      a.lpart[off + ic] = lqv;
    }
  }
}

template <class R, class RP, bool Q3>
__global__ __launch_bounds__(256) void k_tr_adj1(DMesh<R> m, TrAdj<R, RP> a) {
  const TrItem ti = tr_item(m.nw, a.ntr, a.tpw, m.tall);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const long ldk = m.ldk, ic = c * ldk + l;
  const int np = m.w_npoly[c];
  constexpr int NQ = Q3 ? 3 : 1;
  int nb[7];
  R vmj[7], dxs[7], dys[7], oqs[7][NQ];
  bool iw1[7], up[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    nb[j] = m.w_nbr[c * NSLOT + j];
    const int iv = m.w_face[c * NSLOT + j];
    const R dv = m.w_dirv[c * NSLOT + j];
    dxs[j] = m.w_dxs[c * NSLOT + j]; dys[j] = m.w_dys[c * NSLOT + j];
    if constexpr (Q3) {
#pragma unroll
      for (int t = 0; t < 3; ++t) oqs[j][t] = m.w_os[0][t][c * NSLOT + j];
    }
    const R vm = (j < 6 || np > 6) ? ld(a.vmasc, (long)iv * ldk + l, act) : R(0);
    vmj[j] = vm * dv * dv;                       // zero on padded slots
    iw1[j] = dv < R(0);
    up[j] = iw1[j] ? !(vm < R(0)) : (vm < R(0));  // this column is the upwind side
  }
  // the next tracer's loads (nu of the column and its neighbours, lpart) are issued
  // before the current one is reduced; all of a tracer's loads are in the prefetch
  struct Ld { R nuc, lp, nun[7]; };
  auto load = [&](int n, Ld &L) {
    const long off = n * a.tstride;
    L.nuc = ld(a.nu + off, ic, act);
    L.lp = ld(a.lpart + off, ic, act);
#pragma unroll
    for (int j = 0; j < 7; ++j) L.nun[j] = (j < 6 || np > 6) ? ld(a.nu + off, (long)nb[j] * ldk + l, act) : R(0);
  };
  Ld nx;
  if (ti.n0 < ti.n1) load(ti.n0, nx);
  for (int n = ti.n0; n < ti.n1; ++n) {
    const Ld L = nx;
    if (n + 1 < ti.n1) load(n + 1, nx);
    if (a.limited(n)) continue;
    const long off = n * a.tstride;
    const R nuc = L.nuc;
    R P = R(0), Gx = R(0), Gy = R(0), Gq[NQ];
#pragma unroll
    for (int t = 0; t < NQ; ++t) Gq[t] = R(0);
    for_slots(np, [&](int j) {
      const R nun = L.nun[j];
      const R phi = iw1[j] ? vmj[j] * (nun - nuc) : vmj[j] * (nuc - nun);
      const R ph = up[j] ? phi : R(0);
      P = P + ph;
      Gx = Gx + dxs[j] * ph;
      Gy = Gy + dys[j] * ph;
      if constexpr (Q3) {
#pragma unroll
        for (int t = 0; t < 3; ++t) Gq[t] = Gq[t] + oqs[j][t] * ph;
      }
    });
    if (!body) { P = R(0); Gx = R(0); Gy = R(0); }
    if (act) {
      a.lpart[off + ic] = L.lp + P;
      if (a.order > 1) { a.lgx[off + ic] = Gx; a.lgy[off + ic] = Gy; }
      if constexpr (Q3) {
#pragma unroll
        for (int t = 0; t < 3; ++t) a.lgq[t][off + ic] = body ? Gq[t] : R(0);
      }
    }
  }
}

// k_tr_adj0 and k_tr_adj1 in one pass: the face adjoints need nu of the neighbours,
// which is rebuilt here from their mu (nu = mu dtr volti / rho, the same expression as
// in k_tr_adj0) instead of being written and read back.  Per tracer it reads mu of the
// column and its neighbours and lam0, and writes lam0, lpart, lgx, lgy.
template <class R, class RP, bool Q3, bool TALL = false>
__global__ __launch_bounds__(256, Q3 ? 1 : OLAMGPU_BLOCKS_ADJ01) void k_tr_adj01(DMesh<R> m, TrAdj<R, RP> a) {
  const TrItem ti = tr_item(m.nw, a.ntr, a.tpw, TALL);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk, ic = c * ldk + l;
  const int np = m.w_npoly[c];
  constexpr int NQ = Q3 ? 3 : 1;
  int nb[7];
  R vmj[7], dxs[7], dys[7], vtn[7], rinvn[7], oqs[7][NQ];
  bool iw1[7], up[7];
  unsigned need = 0;
  R csum = R(0);
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    nb[j] = m.w_nbr[c * NSLOT + j];
    const int iv = m.w_face[c * NSLOT + j];
    const R dv = m.w_dirv[c * NSLOT + j];
    dxs[j] = m.w_dxs[c * NSLOT + j]; dys[j] = m.w_dys[c * NSLOT + j];
    if constexpr (Q3) {
#pragma unroll
      for (int t = 0; t < 3; ++t) oqs[j][t] = m.w_os[0][t][c * NSLOT + j];
    }
    const R vm = (j < 6 || np > 6) ? ld(a.vmasc, (long)iv * ldk + l, act) : R(0);
    if (j < 6 || np > 6) csum = csum + dv * vm;
    vmj[j] = vm * dv * dv;
    iw1[j] = dv < R(0);
    up[j] = iw1[j] ? !(vm < R(0)) : (vm < R(0));
  }
  // a neighbour's volume and density are read only if this column is upwind of that face
  // at some level (its mu, per tracer, is read for every face: gating those loads costs
  // more than the reads)
#pragma unroll
  for (int j = 0; j < 7; ++j) need |= (__ballot(up[j] && act) != 0 ? 1u : 0u) << j;
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    const bool take = (need >> j) & 1u;      // wave-uniform
    const long in = (long)nb[j] * ldk + l;
    vtn[j] = take ? ld(m.volti, in, act) : R(0);
    const R rfn = take ? (R)ldv(a.rho, in, act, RP(1)) : R(1);
    rinvn[j] = R(1) / (body ? rfn : R(1));
  }
  const R wm = ld(a.wmasc, ic, act), wm_m1 = vshup(TALL, wm, 1);
  const R fexpl = R(1) - a.ffimp[(long)(a.istage - 1) * m.nw + c];
  // on the first adjoint stage lam0 holds only what the implicit pass of this column wrote
  const bool rd0 = !a.first || a.ffimp[(long)(a.istage - 1) * m.nw + c] > R(0.001);
  const R vt = ld(m.volti, ic, act);
  const R rf = (R)ldv(a.rho, ic, act, RP(1));
  const R rinv = R(1) / (body ? rf : R(1));
  const R cc = csum + fexpl * (wm_m1 - wm);
  const R a11 = m.aw11[l], a12 = m.aw12[l];
  const R a21_p1 = vshdn(TALL, m.aw21[l], 1), a22_p1 = vshdn(TALL, m.aw22[l], 1);
  const bool neg = wm < R(0);
  const bool neg_m1 = vshup(TALL, (int)neg, 1) != 0, neg_p1 = vshdn(TALL, (int)neg, 1) != 0;
  // mu of the stage output with level 1 folded into level 2 (the explicit copy)
  auto fold = [&](R v) {
    const R v1 = vlane0(TALL, v);
    if (k == 1) v = R(0);
    if (k == 2) v = v + v1;
    return v;
  };
  // the next tracer's loads are issued before the current one is computed
  struct Ld { R mu, l0, mun[7]; };
  auto load = [&](int n, Ld &L) {
    const long off = n * a.tstride;
    // unpredicated column loads: lanes above mza only feed their own (unstored) results,
    // and nu is zero there
    L.mu = ldl(a.mu + off + (long)c * ldk, (unsigned)l);
    L.l0 = rd0 ? ldl_nt(a.lam0 + off + (long)c * ldk, (unsigned)l) : R(0);
#pragma unroll
    for (int j = 0; j < 7; ++j)
      L.mun[j] = (j < 6 || np > 6) ? ldl(a.mu + off + (long)nb[j] * ldk, (unsigned)l) : R(0);
  };
  Ld nx;
  if (ti.n0 < ti.n1) load(ti.n0, nx);
  for (int n = ti.n0; n < ti.n1; ++n) {
    const Ld L = nx;
    if (n + 1 < ti.n1) load(n + 1, nx);
    if (a.limited(n)) continue;
    const long off = n * a.tstride;
    const R mu = fold(L.mu);
    R l0 = L.l0;
    const R nu = body ? mu * a.dtr * vt * rinv : R(0);
    if (body) l0 = l0 + mu - nu * cc;
    // vertical transpose (as k_tr_adj0)
    const R nu_p1 = vshdn(TALL, nu, 1);
    const R lupw = inner ? fexpl * wm * ((k + 1 <= mza ? nu_p1 : R(0)) - nu) : R(0);
    const R lupw_m1 = vshup(TALL, lupw, 1), lupw_p1 = vshdn(TALL, lupw, 1);
    R lqv = R(0), lds = R(0);
    if (body) {
      if (inner && !neg) lqv = lqv + lupw;
      if (k - 1 >= 2 && neg_m1) lqv = lqv + lupw_m1;
      if (k - 1 >= 2 && neg_m1) lds = lds + lupw_m1 * a11;
      if (inner && neg) lds = lds + lupw * a21_p1;
      if (inner && !neg) lds = lds + lupw * a12;
      if (k + 1 <= mza - 1 && !neg_p1) lds = lds + lupw_p1 * a22_p1;
    }
    if (!inner) lds = R(0);
    const R lds_m1 = vshup(TALL, lds, 1);
    if (body) lqv = lqv + (k - 1 >= 2 ? lds_m1 : R(0)) - lds;
    // face adjoints to the upwind side (as k_tr_adj1)
    R P = R(0), Gx = R(0), Gy = R(0), Gq[NQ];
#pragma unroll
    for (int t = 0; t < NQ; ++t) Gq[t] = R(0);
    for_slots(np, [&](int j) {
      const R mun = fold(L.mun[j]);
      const R nun = body ? mun * a.dtr * vtn[j] * rinvn[j] : R(0);
      const R phi = iw1[j] ? vmj[j] * (nun - nu) : vmj[j] * (nu - nun);
      const R ph = up[j] ? phi : R(0);
      P = P + ph;
      Gx = Gx + dxs[j] * ph;
      Gy = Gy + dys[j] * ph;
      if constexpr (Q3) {
#pragma unroll
        for (int t = 0; t < 3; ++t) Gq[t] = Gq[t] + oqs[j][t] * ph;
      }
    });
    if (!body) { P = R(0); Gx = R(0); Gy = R(0); }
    if (act) {
      st_nt(a.lam0, off + ic, l0);
      st_nt(a.lpart, off + ic, lqv + P);
      if (a.order > 1) { st_nt(a.lgx, off + ic, Gx); st_nt(a.lgy, off + ic, Gy); }
      if constexpr (Q3) {
#pragma unroll
        for (int t = 0; t < 3; ++t) a.lgq[t][off + ic] = body ? Gq[t] : R(0);
      }
    }
  }
}

template <class R, class RP>
__global__ __launch_bounds__(256) void k_tr_adj2(DMesh<R> m, TrAdj<R, RP> a) {
  const TrItem ti = tr_item(m.nw, a.ntr, a.tpw, m.tall);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const long ldk = m.ldk, ic = c * ldk + l;
  const int np = m.w_npoly[c];
  int nb[7];
  R bx[7], by[7], cx[7], cy[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    nb[j] = m.w_nbr[c * NSLOT + j];
    const int b = m.w_back[c * NSLOT + j];
    cx[j] = m.w_gxc[c * NSLOT + j]; cy[j] = m.w_gyc[c * NSLOT + j];
    bx[j] = (j < np) ? m.w_gxc[nb[j] * NSLOT + b] : R(0);
    by[j] = (j < np) ? m.w_gyc[nb[j] * NSLOT + b] : R(0);
  }
  // the next tracer's loads are issued before the current one is reduced
  struct Ld { R lp, l0, gxs, gys, gxn[7], gyn[7]; };
  auto load = [&](int n, Ld &L) {
    const long off = n * a.tstride;
    // unpredicated column loads: every result is per level and stored for act lanes only
    const unsigned ul = (unsigned)l;
    const long cb = off + (long)c * ldk;
    L.lp = ldl_nt(a.lpart + cb, ul);
    L.l0 = a.last ? ldl_nt(a.lam0 + cb, ul) : R(0);
    L.gxs = R(0); L.gys = R(0);
#pragma unroll
    for (int j = 0; j < 7; ++j) { L.gxn[j] = R(0); L.gyn[j] = R(0); }
    if (a.order > 1) {
      L.gxs = ldl(a.lgx + cb, ul); L.gys = ldl(a.lgy + cb, ul);
#pragma unroll
      for (int j = 0; j < 7; ++j)
        if (j < 6 || np > 6) {
          const long nbb = off + (long)nb[j] * ldk;
          L.gxn[j] = ldl(a.lgx + nbb, ul);
          L.gyn[j] = ldl(a.lgy + nbb, ul);
        }
    }
  };
  Ld nx;
  if (ti.n0 < ti.n1) load(ti.n0, nx);
  for (int n = ti.n0; n < ti.n1; ++n) {
    const Ld L = nx;
    if (n + 1 < ti.n1) load(n + 1, nx);
    const long off = n * a.tstride;
    R lq = L.lp;
    const R lev1 = lq;                                  // limited stage: mu(1) passes through
    if (a.order > 1) {
      for_slots(np, [&](int j) {
        lq = lq + (bx[j] * L.gxn[j] + by[j] * L.gyn[j]);
        lq = lq - (cx[j] * L.gxs + cy[j] * L.gys);
      });
    }
    if (!body) lq = a.limited(n) ? lev1 : R(0);
    if (a.last) lq = lq + L.l0;                        // stage 1 read q = scp0
    if (act) st_nt(a.lq, off + ic, lq);
  }
}

// k_tr_adj2 at order 3: the transpose of grad_t2d_quadratic, with each tracer's
// coefficient set (unbounded xy_hu for a monotone-limited final stage, bounded xy_h
// otherwise).  Padded slots carry no coefficients (their back slot is not a real one).
template <class R, class RP>
__global__ __launch_bounds__(256) void k_tr_adj2q(DMesh<R> m, TrAdj<R, RP> a) {
  const TrItem ti = tr_item(m.nw, a.ntr, a.tpw, m.tall);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1;
  const bool act = k <= m.mza, body = act && k >= 2;
  const long ldk = m.ldk, ic = c * ldk + l;
  const int np = m.w_npoly[c];
  int nb[7], bk[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) { nb[j] = m.w_nbr[c * NSLOT + j]; bk[j] = m.w_back[c * NSLOT + j]; }
  for (int n = ti.n0; n < ti.n1; ++n) {
    const long off = n * a.tstride;
    const int set = (a.limited(n) && a.tmode[n] == 1) ? 1 : 0;
    const R *lg[5] = {a.lgx + off, a.lgy + off, a.lgq[0] + off, a.lgq[1] + off, a.lgq[2] + off};
    R lq = ld(a.lpart + off, ic, act);
    const R lev1 = lq;                                  // limited stage: mu(1) passes through
    R gs[5];
#pragma unroll
    for (int t = 0; t < 5; ++t) gs[t] = ld(lg[t], ic, act);
    for_slots(np, [&](int j) {
      const long in = (long)nb[j] * ldk + l;
      R gin = R(0), gout = R(0);
#pragma unroll
      for (int t = 0; t < 5; ++t) {
        const R bq = (j < np) ? m.w_q[set][t][nb[j] * NSLOT + bk[j]] : R(0);
        gin = gin + bq * ld(lg[t], in, act);
        gout = gout + m.w_q[set][t][c * NSLOT + j] * gs[t];
      }
      lq = lq + gin;
      lq = lq - gout;
    });
    if (!body) lq = a.limited(n) ? lev1 : R(0);
    if (a.last) lq = lq + ld(a.lam0 + off, ic, act);   // stage 1 read q = scp0
    if (act) a.lq[off + ic] = lq;
  }
}

// ---------------------------------------------------------------- LDS-staged adjoint
// One adjoint RK stage of all tracers for one patch: k_tr_adj0, k_tr_adj1 and
// k_tr_adj2 in one pass, with nu, the face-flux sums and the gradient adjoints kept
// in LDS instead of global memory.  Per tracer:
//   A. nu = mu dtr volti / rho of every patch cell (owned, rings 1 and 2) -> LDS;
//   B. for owned and ring-1 cells, the face adjoints phi of the faces where the cell
//      is upwind, summed into the gradient adjoints (Gx, Gy) -> LDS; for owned cells
//      also the scp0 terms, the vertical transpose and the face sum P (registers);
//   C. for owned cells, the grad_t2d transpose from the (Gx, Gy) of the column and of
//      its neighbours, and the stage-input adjoint.
// The arithmetic and its order are those of the three separate kernels.
template <class R, class RP, int CPW>
__global__ __launch_bounds__(512) void k_tr_adj_patch(DMesh<R> m, TrAdj<R, RP> a, PatchTab pt) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int p = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x));
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nwv = blockDim.x >> 6;
  const int l = lane_id(), k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk;
  const int n_own = pt.cnt[p * 4], n_r1 = pt.cnt[p * 4 + 1], n_tot = pt.cnt[p * 4 + 2];
  const int nr = n_own + n_r1;
  const int *cells = pt.cells + (long)p * pt.pmax;
  PCell<R> *sc = reinterpret_cast<PCell<R> *>(smem_raw);
  R *snu = reinterpret_cast<R *>(smem_raw + (size_t)pt.rmax * sizeof(PCell<R>));
  R *sgx = snu + (size_t)pt.pmax * WAVE;
  R *sgy = sgx + (size_t)pt.rmax * WAVE;
  stage_pcells(m, pt, p, nr, sc);
  const R a11 = m.aw11[l], a12 = m.aw12[l];
  const R a21_p1 = vshdn(m.tall, m.aw21[l], 1), a22_p1 = vshdn(m.tall, m.aw22[l], 1);
  // tracer-independent data of this wave's owned columns
  R vmj[CPW][7], cc[CPW], wmr[CPW], fexr[CPW];
  unsigned upm[CPW], iw1m[CPW];
#pragma unroll
  for (int t = 0; t < CPW; ++t) {
    const int i = w + t * nwv;
    upm[t] = 0; iw1m[t] = 0; cc[t] = R(0); wmr[t] = R(0); fexr[t] = R(0);
#pragma unroll
    for (int j = 0; j < 7; ++j) vmj[t][j] = R(0);
    if (i < n_own) {
      const int c = cells[i];
      const long ic = (long)c * ldk + l;
      const int np = sc[i].np;
      R csum = R(0);
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        const int iv = m.w_face[c * NSLOT + j];
        const R dv = m.w_dirv[c * NSLOT + j];
        const R vmf = ld(a.vmasc, (long)iv * ldk + l, act);
        if (j < 6 || np > 6) csum = csum + dv * vmf;
        const R vm = (j < 6 || np > 6) ? vmf : R(0);
        vmj[t][j] = vm * dv * dv;
        const bool iw1 = dv < R(0);
        const bool up = iw1 ? !(vm < R(0)) : (vm < R(0));
        iw1m[t] |= (iw1 ? 1u : 0u) << j;
        upm[t] |= (up ? 1u : 0u) << j;
      }
      wmr[t] = ld(a.wmasc, ic, act);
      fexr[t] = R(1) - a.ffimp[(long)(a.istage - 1) * m.nw + c];
      cc[t] = csum + fexr[t] * (vshup(m.tall, wmr[t], 1) - wmr[t]);
    }
  }
  for (int n = 0; n < a.ntr; ++n) {
    const long off = n * a.tstride;
    __syncthreads();                                  // records staged / previous tracer done
    // A: nu of every patch cell
    for (int i = w; i < n_tot; i += nwv) {
      const long ix = (long)cells[i] * ldk + l;
      R mu = ld(a.mu + off, ix, act);
      const R mu1 = vlane0(m.tall, mu);
      if (k == 1) mu = R(0);
      if (k == 2) mu = mu + mu1;
      const R vt = ld(m.volti, ix, act);
      const R rf = (R)ldv(a.rho, ix, act, RP(1));
      const R rinv = R(1) / (body ? rf : R(1));
      snu[i * WAVE + l] = body ? mu * a.dtr * vt * rinv : R(0);
    }
    __syncthreads();
    // B: owned columns
    R lpart[CPW], l0r[CPW];
#pragma unroll
    for (int t = 0; t < CPW; ++t) {
      lpart[t] = R(0); l0r[t] = R(0);
      const int i = w + t * nwv;
      if (i >= n_own) continue;
      const int c = cells[i];
      const long ic = (long)c * ldk + l;
      const PCell<R> &me = sc[i];
      R mu = ld(a.mu + off, ic, act);
      const R mu1 = vlane0(m.tall, mu);
      if (k == 1) mu = R(0);
      if (k == 2) mu = mu + mu1;
      // on the first adjoint stage lam0 holds only what the implicit pass wrote
      const bool rd0 = !a.first || a.ffimp[(long)(a.istage - 1) * m.nw + c] > R(0.001);
      R l0 = rd0 ? ld(a.lam0 + off, ic, act) : R(0);
      const R nu = snu[i * WAVE + l];
      if (body) l0 = l0 + mu - nu * cc[t];
      const R wm = wmr[t], fexpl = fexr[t];
      const bool neg = wm < R(0);
      const bool neg_m1 = vshup(m.tall, (int)neg, 1) != 0, neg_p1 = vshdn(m.tall, (int)neg, 1) != 0;
      const R nu_p1 = vshdn(m.tall, nu, 1);
      const R lupw = inner ? fexpl * wm * ((k + 1 <= mza ? nu_p1 : R(0)) - nu) : R(0);
      const R lupw_m1 = vshup(m.tall, lupw, 1), lupw_p1 = vshdn(m.tall, lupw, 1);
      R lqv = R(0), lds = R(0);
      if (body) {
        if (inner && !neg) lqv = lqv + lupw;
        if (k - 1 >= 2 && neg_m1) lqv = lqv + lupw_m1;
        if (k - 1 >= 2 && neg_m1) lds = lds + lupw_m1 * a11;
        if (inner && neg) lds = lds + lupw * a21_p1;
        if (inner && !neg) lds = lds + lupw * a12;
        if (k + 1 <= mza - 1 && !neg_p1) lds = lds + lupw_p1 * a22_p1;
      }
      if (!inner) lds = R(0);
      const R lds_m1 = vshup(m.tall, lds, 1);
      if (body) lqv = lqv + (k - 1 >= 2 ? lds_m1 : R(0)) - lds;
      R P = R(0), Gx = R(0), Gy = R(0);
      for_slots(me.np, [&](int j) {
        const R nun = snu[me.nl[j] * WAVE + l];
        const R phi = ((iw1m[t] >> j) & 1u) ? vmj[t][j] * (nun - nu) : vmj[t][j] * (nu - nun);
        const R ph = ((upm[t] >> j) & 1u) ? phi : R(0);
        P = P + ph;
        Gx = Gx + m.w_dxs[c * NSLOT + j] * ph;
        Gy = Gy + m.w_dys[c * NSLOT + j] * ph;
      });
      if (!body) { P = R(0); Gx = R(0); Gy = R(0); }
      lpart[t] = lqv + P;
      l0r[t] = l0;
      if (act) a.lam0[off + ic] = l0;
      sgx[i * WAVE + l] = Gx;
      sgy[i * WAVE + l] = Gy;
    }
    // B: ring-1 cells (gradient adjoints only)
    if (a.order > 1) {
      for (int i = n_own + w; i < nr; i += nwv) {
        const int x = cells[i];
        const PCell<R> &me = sc[i];
        const R nu = snu[i * WAVE + l];
        R Gx = R(0), Gy = R(0);
        for_slots(me.np, [&](int j) {
          const int iv = m.w_face[x * NSLOT + j];
          const R dv = m.w_dirv[x * NSLOT + j];
          const R vm = ld(a.vmasc, (long)iv * ldk + l, act);
          const R vj = vm * dv * dv;
          const bool iw1 = dv < R(0);
          const bool up = iw1 ? !(vm < R(0)) : (vm < R(0));
          const R nun = snu[me.nl[j] * WAVE + l];
          const R phi = iw1 ? vj * (nun - nu) : vj * (nu - nun);
          const R ph = up ? phi : R(0);
          Gx = Gx + m.w_dxs[x * NSLOT + j] * ph;
          Gy = Gy + m.w_dys[x * NSLOT + j] * ph;
        });
        if (!body) { Gx = R(0); Gy = R(0); }
        sgx[i * WAVE + l] = Gx;
        sgy[i * WAVE + l] = Gy;
      }
    }
    __syncthreads();
    // C: grad_t2d transpose and the stage-input adjoint of owned columns
#pragma unroll
    for (int t = 0; t < CPW; ++t) {
      const int i = w + t * nwv;
      if (i >= n_own) continue;
      const int c = cells[i];
      const long ic = (long)c * ldk + l;
      const PCell<R> &me = sc[i];
      R lq = lpart[t];
      if (a.order > 1) {
        const R gxs = sgx[i * WAVE + l], gys = sgy[i * WAVE + l];
        for_slots(me.np, [&](int j) {
          const int u = me.nl[j];
          const int b = m.w_back[c * NSLOT + j];
          const R bx = (j < me.np) ? sc[u].cx[b] : R(0);
          const R by = (j < me.np) ? sc[u].cy[b] : R(0);
          lq = lq + (bx * sgx[u * WAVE + l] + by * sgy[u * WAVE + l]);
          lq = lq - (me.cx[j] * gxs + me.cy[j] * gys);
        });
      }
      if (!body) lq = R(0);
      if (a.last) lq = lq + l0r[t];
      if (act) a.lq[off + ic] = lq;
    }
  }
}

// ---------------------------------------------------------------- nudged density, sgs mixing
// rhof of scalar_transport_rk.F90:120-133 when nudging preserves the mixing ratio: the
// density less the long step's nudging density tendency, formed in FP64.
template <class RP>
__global__ void k_tr_rhof(const RP *rho, const RP *rhot_nud, double dtl, RP *out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (RP)((double)rho[i] - dtl * (double)rhot_nud[i]);
}

// Horizontal diffusion of the tracers into their tendencies (scalar_transport_rk.F90:
// 244-274): sct += volti sum_faces akhodx (q_nbr - q) on levels 2..khtopv of each face.
// Only columns with some face diffusing (kwtop >= 2, :160) change; sgs[n] is the
// tracer's do_sgsmix.  One wave per column, all tracers.
template <class R>
__global__ __launch_bounds__(256) void k_tr_sgsmix(DMesh<R> m, const R *__restrict__ q, R *__restrict__ sct,
                                                   const R *akhodx, const int *khtopv, const int *sgs, int ntr,
                                                   long tstride) {
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1;
  const bool act = k <= m.mza;
  const long ldk = m.ldk, ic = c * ldk + l;
  const int np = m.w_npoly[c];
  int nb[7], kwtop = 0;
  R ak[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    nb[j] = m.w_nbr[c * NSLOT + j];
    const int iv = m.w_face[c * NSLOT + j];
    const int kt = j < np ? khtopv[iv] : 0;
    kwtop = max(kwtop, kt);
    ak[j] = (act && k >= 2 && k <= kt) ? akhodx[(long)iv * ldk + l] : R(0);
  }
  if (kwtop < 2) return;
  const R vt = ld(m.volti, ic, act);
  const bool upd = act && k >= 2 && k <= kwtop;
  for (int n = 0; n < ntr; ++n) {
    if (!sgs[n]) continue;
    const R *qn = q + n * tstride;
    const R qc = ld(qn, ic, act);
    R fd = R(0);
    for_slots(np, [&](int j) { fd = fd + ak[j] * (ld(qn, (long)nb[j] * ldk + l, act) - qc); });
    if (upd) sct[n * tstride + ic] = sct[n * tstride + ic] + fd * vt;
  }
}

// Adjoint of the tendency of the diffused tracers: per stage, lsct += dtr mu / rhos on
// the prognostic levels, with mu of level 1 folded into level 2 (the explicit copy; the
// implicit pass has already moved it there).  The first adjoint stage starts from zero.
template <class R, class RP>
__global__ __launch_bounds__(256) void k_tr_adj_sct(DMesh<R> m, TrAdj<R, RP> a, R *lsct, const int *sgs) {
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1;
  const bool act = k <= m.mza, body = act && k >= 2;
  const long ldk = m.ldk, ic = c * ldk + l;
  const R rf = (R)ldv(a.rho, ic, act, RP(1));
  const R rinv = R(1) / (body ? rf : R(1));
  for (int n = 0; n < a.ntr; ++n) {
    if (!sgs[n]) continue;
    const long off = n * a.tstride;
    R mu = ld(a.mu + off, ic, act);
    const R mu1 = vlane0(m.tall, mu);
    if (k == 1) mu = R(0);
    if (k == 2) mu = mu + mu1;
    const R v = body ? mu * a.dtr * rinv : R(0);
    const R prev = a.first ? R(0) : ld(lsct + off, ic, act);
    if (act) lsct[off + ic] = prev + v;
  }
}

// lq += D^T lsct for the diffused tracers: the transpose of k_tr_sgsmix,
// lq(c) += sum_faces akhodx (volti(nbr) lsct(nbr) - volti(c) lsct(c)).
template <class R>
__global__ __launch_bounds__(256) void k_tr_sgsmix_adj(DMesh<R> m, const R *__restrict__ lsct, R *__restrict__ lq,
                                                       const R *akhodx, const int *khtopv, const int *sgs, int ntr,
                                                       long tstride) {
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1;
  const bool act = k <= m.mza;
  const long ldk = m.ldk, ic = c * ldk + l;
  const int np = m.w_npoly[c];
  int nb[7];
  R ak[7], vtn[7];
  bool any = false;
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    nb[j] = m.w_nbr[c * NSLOT + j];
    const int iv = m.w_face[c * NSLOT + j];
    const int kt = j < np ? khtopv[iv] : 0;
    any = any || kt >= 2;
    ak[j] = (act && k >= 2 && k <= kt) ? akhodx[(long)iv * ldk + l] : R(0);
    vtn[j] = ld(m.volti, (long)nb[j] * ldk + l, act);
  }
  if (!any) return;
  const R vt = ld(m.volti, ic, act);
  for (int n = 0; n < ntr; ++n) {
    if (!sgs[n]) continue;
    const long off = n * tstride;
    const R lc = ld(lsct + off, ic, act);
    R acc = ld(lq + off, ic, act);
    for_slots(np, [&](int j) { acc = acc + ak[j] * (vtn[j] * ld(lsct + off, (long)nb[j] * ldk + l, act) - vt * lc); });
    if (act) lq[off + ic] = acc;
  }
}

// ---------------------------------------------------------------- tracer_maxmin
// check_nans.F90:164-203: per tracer the maximum and minimum of var_p over the prognostic
// levels lpw..mza of every column.  Floats are ordered through their bit patterns (an int
// key that sorts as the float does) so each wave's reduction is folded in with one integer
// atomic max per tracer and bound; max and min are exact, in any order.  NaN values are
// skipped (fmax / fmin keep the other operand).
__device__ __forceinline__ int f2key(float f) {
  const int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float key2f(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff); }

__global__ void k_maxmin_init(int *keys, int n) {
  const int i = threadIdx.x;
  if (i < n) keys[i] = f2key(-3.402823466e+38f);   // -huge(1.)
}

template <class R>
__global__ __launch_bounds__(256) void k_tr_maxmin(DMesh<R> m, const R *q, int ntr, long tstride, int *keys) {
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1;
  const bool body = k >= 2 && k <= m.mza;
  const long ic = (long)c * m.ldk + l;
  for (int n = 0; n < ntr; ++n) {
    const float v = (float)ld(q + n * tstride, ic, body);
    float hi = body ? v : -3.402823466e+38f, lo = body ? -v : -3.402823466e+38f;
    for (int s = 32; s >= 1; s >>= 1) {
      hi = fmaxf(hi, __shfl_xor(hi, s, WAVE));
      lo = fmaxf(lo, __shfl_xor(lo, s, WAVE));
    }
    if (lane_id() == 0) {
      atomicMax(keys + 2 * n, f2key(hi));
      atomicMax(keys + 2 * n + 1, f2key(lo));
    }
  }
}

__global__ void k_maxmin_final(const int *keys, int n, float *out) {
  const int i = threadIdx.x;
  if (i < n) out[i] = (i & 1) ? -key2f(keys[i]) : key2f(keys[i]);
}

// ---------------------------------------------------------------- mass sums
// compute_mass_sums (check_nans.F90:224-262): per-block FP64 partial sums of
// rho*volt and rho*volt*q_n, reduced in a fixed order (deterministic).
template <class R, class RP>
__global__ __launch_bounds__(256) void k_mass_partial(DMesh<R> m, const RP *rho, const R *q, int ntr, long tstride,
                                                      double *part) {
  __shared__ double sh[WPB][65];
  const int b = blockIdx.x;
  const int w = threadIdx.x >> 6, gw = b * WPB + w;
  const int c = m.tall ? gw >> 1 : gw;                       // a tall column: two waves
  const int l = (m.tall ? (gw & 1) * WAVE : 0) + lane_id(), k = l + 1;
  const bool ok = c < m.nw && k >= 2 && k <= m.mza;
  const long ic = (long)c * m.ldk + l;
  const double mass = ok ? (double)rho[ic] * m.volt[ic] : 0.0;
  for (int n = -1; n < ntr && n < 64; ++n) {
    double v = n < 0 ? mass : (ok ? mass * (double)q[n * tstride + ic] : 0.0);
    for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, WAVE);
    if (lane_id() == 0) sh[w][n + 1] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x <= ntr && threadIdx.x < 65) {
    double s = 0.0;
    for (int i = 0; i < WPB; ++i) s += sh[i][threadIdx.x];
    part[(long)b * 65 + threadIdx.x] = s;
  }
}

__global__ void k_mass_final(const double *part, int nblk, int nval, double *out) {
  __shared__ double sh[256];
  for (int v = 0; v < nval; ++v) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 256) s += part[(long)i * 65 + v];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st >= 1; st >>= 1) {
      if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[v] = sh[0];
    __syncthreads();
  }
}

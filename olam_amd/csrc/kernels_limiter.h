// kernels_limiter.h -- final-stage flux limiters of scalar_transport_rk and their
// adjoint with the renormalisation scales frozen.
//
// Reference: update_scalar_monot (omodel/scalar_transport_rk.F90:619-1081,
// iscal_monot = 1) and update_scalar_pd (:1085-1472, iscal_monot = 2), applied on the
// final RK stage.  Each face flux is the low-order upwind value from scp0 plus the
// high-order correction (reconstruction from the stage input minus that value) times a
// renormalisation factor built from per-column in/out flux sums.  Two passes, one wave
// per column (lane l = level k = l + 1):
//   k_tr_lim_scales  the factors scale_inout(k, iw, 1:2) (monotone) or scale_out (PD);
//   k_tr_lim_update  the renormalised face and vertical fluxes and the forward step,
//                    taping the applied scales when the adjoint is enabled.
// Gradients of the stage input come from k_tr_grad.  Tracers that the final stage
// does not limit (tmode = 0) are left to k_tr_update.  No FMA contraction: where the
// high-order correction vanishes the scale is arbitrary for the forward step but not
// for the frozen adjoint, so the scales must come out exactly as in the oracle.
#pragma once
#include "kernels_tracer.h"

template <class R>
struct TrLim {
  const int *tmode;   // per tracer: 0 regular update, 1 monotone, 2 positive definite
  R *s1, *s2;         // scale_inout(:,:,1:2) or scale_out (s1), tstride per tracer
  R *tsv, *tsw;       // taped face (vstride per tracer) and vertical scales, or null
  long vstride;
};

// Column context shared by the two forward passes.
template <class R>
struct LimCol {
  int c, np, nb[7], iv[7];
  R dv[7], vm[7], dxs[7], dys[7], dxn[7], dyn[7];
  R wm, wm_m1, fimpl, fexpl, vt, rs;
};

template <class R, class RP>
__device__ __forceinline__ void lim_col(const DMesh<R> &m, const TrStage<R, RP> &a, int c, int l, bool act,
                                        bool body, LimCol<R> &L) {
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  L.c = c;
  L.np = m.w_npoly[c];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    L.nb[j] = m.w_nbr[c * NSLOT + j];
    L.iv[j] = m.w_face[c * NSLOT + j];
    L.dv[j] = m.w_dirv[c * NSLOT + j];
    L.dxs[j] = m.w_dxs[c * NSLOT + j]; L.dys[j] = m.w_dys[c * NSLOT + j];
    L.dxn[j] = m.w_dxn[c * NSLOT + j]; L.dyn[j] = m.w_dyn[c * NSLOT + j];
    L.vm[j] = (j < 6 || L.np > 6) ? ld(a.vmasc, (long)L.iv[j] * ldk + l, act) : R(0);
  }
  L.wm = ld(a.wmasc, ic, act);
  L.wm_m1 = vshup(m.tall, L.wm, 1);
  L.fimpl = a.ffimp[(long)(a.istage - 1) * m.nw + c];
  L.fexpl = R(1) - L.fimpl;
  L.vt = ld(m.volti, ic, act);
  L.rs = (R)ldv(a.rho, ic, act, RP(1));
  if (!body) L.rs = R(1);
}

// Face j of column c for tracer data at offset off: the low-order value (scp0 of the
// upwind cell), the high-order correction and its flux (scalar_vflux_monot :748-827,
// scalar_vflux_pd :1227-1290).  up_c: this column is the upwind one.  At order 3 the
// quadratic terms gqc (this column's) are added with the unbounded face offsets for the
// monotone form (xx0_vu, :793-807) and the bounded ones for the positive-definite form
// (xx0_v, :1252-1266).
template <class R, class RP>
__device__ __forceinline__ void lim_face(const DMesh<R> &m, const TrStage<R, RP> &a, const LimCol<R> &L, int j,
                                         long off, int l, bool act, int mode, R qc, R q0c, R gxc, R gyc,
                                         const R *gqc, R &low, R &hiv, R &sf, R &q0n) {
  NOCONTRACT
  const long ldk = m.ldk, in = (long)L.nb[j] * ldk + l;
  const R vm = L.vm[j];
  const bool iw1side = L.dv[j] < R(0);
  const bool up_c = iw1side ? !(vm < R(0)) : (vm < R(0));
  q0n = ld(a.q0 + off, in, act);
  low = up_c ? q0c : q0n;
  const bool hi_c = (mode == 2 && a.order == 1) ? iw1side : up_c;
  const R qn = ld(a.q_in + off, in, act);
  R high;
  if (a.order == 1) {
    high = hi_c ? qc : qn;
  } else if (hi_c) {
    high = qc + L.dxs[j] * gxc + L.dys[j] * gyc;
    if (a.order == 3) {
      const int set = mode == 1 ? 1 : 0;
      const long e = (long)L.c * NSLOT + j;
      high = high + m.w_os[set][0][e] * gqc[0] + m.w_os[set][1][e] * gqc[1] + m.w_os[set][2][e] * gqc[2];
    }
  } else {
    high = qn + L.dxn[j] * ld(a.gx + off, in, act) + L.dyn[j] * ld(a.gy + off, in, act);
    if (a.order == 3) {
      const int set = mode == 1 ? 1 : 0;
      const long e = (long)L.c * NSLOT + j;
      high = high + m.w_on[set][0][e] * ld(a.gq[0] + off, in, act) + m.w_on[set][1][e] * ld(a.gq[1] + off, in, act) +
             m.w_on[set][2][e] * ld(a.gq[2] + off, in, act);
    }
  }
  hiv = high - low;
  sf = vm * hiv;
}

template <class R, class RP>
__global__ __launch_bounds__(256) void k_tr_lim_scales(DMesh<R> m, TrStage<R, RP> a, TrLim<R> T) {
  NOCONTRACT
  const TrItem ti = tr_item(m.nw, a.ntr, a.tpw, m.tall);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  LimCol<R> L;
  lim_col(m, a, c, l, act, body, L);
  const R eps0 = LimConst<R>::eps0, eps1 = LimConst<R>::eps1;
  const R onep = R(1) + eps1, onem = R(1) - eps1;
  for (int n = ti.n0; n < ti.n1; ++n) {
    const int mode = T.tmode[n];
    if (mode == 0) continue;
    const long off = n * a.tstride;
    const R qc = ld(a.q_in + off, ic, act), q0c = ld(a.q0 + off, ic, act);
    R gxc = R(0), gyc = R(0), gqc[3] = {R(0), R(0), R(0)};
    if (a.order > 1) { gxc = ld(a.gx + off, ic, act); gyc = ld(a.gy + off, ic, act); }
    if (a.order == 3)
      for (int t = 0; t < 3; ++t) gqc[t] = ld(a.gq[t] + off, ic, act);
    // vertical: low-order upwind values from scp0, high-order from the stage input
    const R q0p1 = vshdn(m.tall, q0c, 1), qp1 = vshdn(m.tall, qc, 1);
    R upwl = (L.wm < R(0)) ? q0p1 : q0c;
    if (k == 1) upwl = (mode == 1) ? q0p1 : qp1;
    if (k == mza) upwl = (mode == 1) ? q0c : qc;
    R sb, st;
    lim_vrec(m, qc, k, mza, mode, sb, st);
    const R sb_p1 = vshdn(m.tall, sb, 1);
    const R hiw = inner ? ((L.wm < R(0)) ? sb_p1 : st) - upwl : R(0);
    const R sfw = inner ? L.fexpl * L.wm * hiw : R(0);
    const R sfw_m1 = vshup(m.tall, sfw, 1);
    const R upwl_m1 = vshup(m.tall, upwl, 1);
    R smax = q0c, smin = q0c;
    if (mode == 1) {
      const R q0m1 = vshup(m.tall, q0c, 1);
      if (k >= 3 && L.wm_m1 > R(0)) { smax = fmax(smax, q0m1); smin = fmin(smin, q0m1); }
      if (inner && L.wm < R(0)) { smax = fmax(smax, q0p1); smin = fmin(smin, q0p1); }
    }
    R fl = R(0), fin = R(0), fout = R(0);
    for_slots(L.np, [&](int j) {
      R low, hiv, sf, q0n;
      lim_face(m, a, L, j, off, l, act, mode, qc, q0c, gxc, gyc, gqc, low, hiv, sf, q0n);
      const R d = L.dv[j];
      fl = fl + d * L.vm[j] * (low - q0c);
      if (mode == 1) fin = fin + fmax(d * sf, R(0));
      fout = fout + fmin(d * sf, R(0));
      if (mode == 1 && d * L.vm[j] > R(0)) { smax = fmax(smax, q0n); smin = fmin(smin, q0n); }
    });
    if (L.fimpl > R(.001)) {
      if (mode == 1) {
        const R dwi = L.fimpl * (L.wm - L.wm_m1);
        fl = fl + L.fexpl * (L.wm_m1 * (upwl_m1 - q0c) - L.wm * (upwl - q0c)) + q0c * dwi;
        fin = fin + fmax(sfw_m1, R(0)) - fmin(sfw, R(0));
        fout = fout + fmin(sfw_m1, R(0)) - fmax(sfw, R(0));
        const R imp = L.rs + a.dtr * L.vt * dwi;
        smax = imp * smax;
        smin = imp * smin;
      } else {
        fl = fl + L.fexpl * (L.wm_m1 * (upwl_m1 - q0c) - L.wm * (upwl - q0c)) + L.fimpl * q0c * (L.wm - L.wm_m1);
        fout = fout + fmin(sfw_m1, R(0)) - fmax(sfw, R(0));
      }
    } else {
      fl = fl + L.wm_m1 * (upwl_m1 - q0c) - L.wm * (upwl - q0c);
      if (mode == 1) fin = fin + fmax(sfw_m1, R(0)) - fmin(sfw, R(0));
      fout = fout + fmin(sfw_m1, R(0)) - fmax(sfw, R(0));
      if (mode == 1) { smax = L.rs * smax; smin = L.rs * smin; }
    }
    const R rlow = q0c * L.rs + a.dtr * L.vt * fl;
    R s1 = R(1), s2 = R(1);
    if (mode == 1) {
      const R dtrp = onep * a.dtr;
      const R flux_in = fmax(dtrp * L.vt * fin, eps0);
      const R flux_out = fmin(dtrp * L.vt * fout, -eps0);
      const R epsr = eps1 * fabs(rlow);
      const R sc_in = fmax(smax - (rlow + epsr) - eps0, R(0));
      const R sc_out = fmin(smin - (rlow - epsr) + eps0, R(0));
      s1 = fmin(R(1), sc_in / fmax(flux_in, R(0.01) * sc_in));
      s2 = fmin(R(1), sc_out / fmin(flux_out, R(0.01) * sc_out));
    } else {
      const R mdtrp = -onep * a.dtr;
      const R flux_out = fmax(mdtrp * L.vt * fout, eps0);
      const R sc_out = fmax(onem * rlow - eps0, R(0));
      s1 = fmin(R(1), sc_out / fmax(flux_out, R(0.01) * sc_out));
    }
    if (body) {
      T.s1[off + ic] = s1;
      if (mode == 1) T.s2[off + ic] = s2;
    }
  }
}

template <class R, class RP>
__global__ __launch_bounds__(256) void k_tr_lim_update(DMesh<R> m, TrStage<R, RP> a, TrLim<R> T) {
  NOCONTRACT
  const TrItem ti = tr_item(m.nw, a.ntr, a.tpw, m.tall);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  LimCol<R> L;
  lim_col(m, a, c, l, act, body, L);
  for (int n = ti.n0; n < ti.n1; ++n) {
    const int mode = T.tmode[n];
    if (mode == 0) continue;
    const long off = n * a.tstride;
    const R qc = ld(a.q_in + off, ic, act), q0c = ld(a.q0 + off, ic, act);
    R gxc = R(0), gyc = R(0), gqc[3] = {R(0), R(0), R(0)};
    if (a.order > 1) { gxc = ld(a.gx + off, ic, act); gyc = ld(a.gy + off, ic, act); }
    if (a.order == 3)
      for (int t = 0; t < 3; ++t) gqc[t] = ld(a.gq[t] + off, ic, act);
    const R s1c = ldv(T.s1, off + ic, body, R(1));
    const R s2c = (mode == 1) ? ldv(T.s2, off + ic, body, R(1)) : R(1);
    // horizontal: renormalised face values (:704-716, :1183-1193)
    R fd = R(0);
    for_slots(L.np, [&](int j) {
      R low, hiv, sf, q0n;
      lim_face(m, a, L, j, off, l, act, mode, qc, q0c, gxc, gyc, gqc, low, hiv, sf, q0n);
      const long in = (long)L.nb[j] * ldk + l;
      const bool iw1side = L.dv[j] < R(0);
      const R s1n = ldv(T.s1, off + in, body, R(1));
      const R s2n = (mode == 1) ? ldv(T.s2, off + in, body, R(1)) : R(1);
      const R s1_1 = iw1side ? s1c : s1n, s1_2 = iw1side ? s1n : s1c;
      const R s2_1 = iw1side ? s2c : s2n, s2_2 = iw1side ? s2n : s2c;
      R scale;
      if (mode == 1) scale = (sf < R(0)) ? fmin(s1_1, s2_2) : fmin(s1_2, s2_1);
      else scale = (sf < R(0)) ? s1_2 : s1_1;
      const R upv = low + hiv * scale;
      fd = fd + L.dv[j] * L.vm[j] * (upv - q0c);
      if (T.tsv && iw1side && body) T.tsv[n * T.vstride + (long)L.iv[j] * ldk + l] = scale;
    });
    // vertical: renormalised fluxes (:1003-1008, :1396-1401)
    const R q0p1 = vshdn(m.tall, q0c, 1), qp1 = vshdn(m.tall, qc, 1);
    R upwl = (L.wm < R(0)) ? q0p1 : q0c;
    if (k == 1) upwl = (mode == 1) ? q0p1 : qp1;
    if (k == mza) upwl = (mode == 1) ? q0c : qc;
    R sb, st;
    lim_vrec(m, qc, k, mza, mode, sb, st);
    const R sb_p1 = vshdn(m.tall, sb, 1);
    const R hiw = inner ? ((L.wm < R(0)) ? sb_p1 : st) - upwl : R(0);
    const R sfw = inner ? L.fexpl * L.wm * hiw : R(0);
    const R s1p = vshdn(m.tall, s1c, 1), s2p = vshdn(m.tall, s2c, 1);
    R vs = R(0);
    if (mode == 1) vs = (sfw < R(0)) ? fmin(s1c, s2p) : fmin(s1p, s2c);
    else vs = (sfw < R(0)) ? s1p : s1c;
    R upw = upwl;
    if (inner) upw = upwl + hiw * vs;
    if (T.tsw && inner) T.tsw[off + ic] = vs;
    const R upw_m1 = vshup(m.tall, upw, 1);
    if (L.fimpl > R(.001)) fd = fd + L.fexpl * (L.wm_m1 * (upw_m1 - q0c) - L.wm * (upw - q0c));
    else fd = fd + L.wm_m1 * (upw_m1 - q0c) - L.wm * (upw - q0c);
    // scalar_tend_monot / scalar_tend_pd: level 1 keeps the stage input
    const R qnew = body ? q0c + a.dtr * L.vt * fd / L.rs : qc;
    if (act) a.q_out[off + ic] = qnew;
  }
}

// ---------------------------------------------------------------- frozen adjoint
// Adjoint of k_tr_lim_update with the taped scales: per face the value is
// (1 - s) scp0(upwind) + s * reconstruction(stage input), so its adjoint phi goes as
// (1 - s) phi to scp0 and s phi to the stage input (through grad_t2d for horizontal
// faces, through the vertical quadratic for vertical ones).  Level 1 of the stage
// output is the stage input's level 1.  Two passes, followed by k_tr_adj2 (grad_t2d
// transpose), like the unlimited adjoint.
template <class R, class RP>
__global__ __launch_bounds__(256) void k_tr_lim_adj0(DMesh<R> m, TrAdj<R, RP> a, TrLim<R> T) {
  NOCONTRACT
  const TrItem ti = tr_item(m.nw, a.ntr, a.tpw, m.tall);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  const int np = m.w_npoly[c];
  R gsum = R(0);
  for_slots(np, [&](int j) {
    const int iv = m.w_face[c * NSLOT + j];
    gsum = gsum + m.w_dirv[c * NSLOT + j] * ld(a.vmasc, (long)iv * ldk + l, act);
  });
  const R wm = ld(a.wmasc, ic, act), wm_m1 = vshup(m.tall, wm, 1);
  const R ff = a.ffimp[(long)(a.istage - 1) * m.nw + c];
  const R fexp = (ff > R(0.001)) ? R(1) - ff : R(1);
  const R vt = ld(m.volti, ic, act);
  const R rf = (R)ldv(a.rho, ic, body, RP(1));
  const bool neg = wm < R(0);
  for (int n = ti.n0; n < ti.n1; ++n) {
    const int mode = T.tmode[n];
    if (mode == 0) continue;
    const long off = n * a.tstride;
    R mu = ld(a.mu + off, ic, act);
    const R mu1 = vlane0(m.tall, mu);
    if (k == 1) mu = R(0);
    // on the first adjoint stage lam0 holds only what the implicit pass wrote
    R l0 = (!a.first || ff > R(0.001)) ? ld(a.lam0 + off, ic, act) : R(0);
    const R nu = body ? mu * a.dtr * vt / rf : R(0);
    if (body) l0 = l0 + mu - nu * (gsum + fexp * (wm_m1 - wm));
    // every lane takes part in a shuffle: a lane that skipped it would hand 0 to its
    // neighbour (here lane mza - 1 would lose nu(mza))
    const R nu_dn = vshdn(m.tall, nu, 1);
    const R nu_p1 = (k + 1 <= mza) ? nu_dn : R(0);
    const R psi = act ? fexp * wm * (nu_p1 - nu) : R(0);          // adjoint of upw(k)
    const R psi_m1 = vshup(m.tall, psi, 1);
    const R s = inner ? ld(T.tsw + off, ic, act) : R(0);
    const R low = inner ? (R(1) - s) * psi : R(0);
    const R lowB = neg ? low : R(0), lowA = neg ? R(0) : low;
    const R h = inner ? s * psi : R(0);
    const R hneg = neg ? h : R(0), hpos = neg ? R(0) : h;
    const R lowB_m1 = vshup(m.tall, lowB, 1);
    if (body) {
      if (mode == 1) {
        if (k == 2) l0 = l0 + psi_m1;          // upw(1) = scp0(2)
        if (k == mza) l0 = l0 + psi;           // upw(mza) = scp0(mza)
      }
      l0 = l0 + lowB_m1 + lowA;
    }
    // vertical quadratic transpose
    const R A11 = mode == 1 ? m.awu11[l] : m.aw11[l];
    const R A21 = mode == 1 ? m.awu21[l] : m.aw21[l];
    const R A12 = mode == 1 ? m.awu12[l] : m.aw12[l];
    const R A22 = mode == 1 ? m.awu22[l] : m.aw22[l];
    const R A21_p1 = vshdn(m.tall, A21, 1);
    const R hneg_m1 = vshup(m.tall, hneg, 1);
    R lds = hneg_m1 * A11 + hneg * A21_p1 + hpos * A12 + vshdn(m.tall, hpos * A22, 1);
    if (k == 1 && mode != 1) lds = R(0);
    if (mode == 1) {                                // ds(1) = ds(2), ds(mza) = ds(mza-1)
      const R l_m1 = vshup(m.tall, lds, 1), l_p1 = vshdn(m.tall, lds, 1);
      if (k == 2) lds = lds + l_m1;
      if (k == mza - 1) lds = lds + l_p1;
    }
    if (!inner) lds = R(0);
    const R lds_m1 = vshup(m.tall, lds, 1);
    R lqv = hpos + hneg_m1 + (k - 1 >= 2 ? lds_m1 : R(0)) - lds;
    if (mode == 2) {
      if (k == 2) lqv = lqv + psi_m1;          // upw(1) = scp(2)
      if (k == mza) lqv = lqv + psi;           // upw(mza) = scp(mza)
    }
    if (!body) lqv = R(0);
    if (k == 1) lqv = mu1;                     // level 1 passes through
    if (act) {
      a.lam0[off + ic] = l0;
      a.nu[off + ic] = nu;
      a.lpart[off + ic] = lqv;
    }
  }
}

template <class R, class RP>
__global__ __launch_bounds__(256) void k_tr_lim_adj1(DMesh<R> m, TrAdj<R, RP> a, TrLim<R> T) {
  NOCONTRACT
  const TrItem ti = tr_item(m.nw, a.ntr, a.tpw, m.tall);
  const int c = ti.c;
  if (c >= m.nw) return;
  const int l = ti.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  const int np = m.w_npoly[c];
  int nb[7], iv[7];
  R vmj[7], dxs[7], dys[7];
  bool iw1[7], up[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    nb[j] = m.w_nbr[c * NSLOT + j];
    iv[j] = m.w_face[c * NSLOT + j];
    const R dv = m.w_dirv[c * NSLOT + j];
    dxs[j] = m.w_dxs[c * NSLOT + j]; dys[j] = m.w_dys[c * NSLOT + j];
    const R vm = (j < 6 || np > 6) ? ld(a.vmasc, (long)iv[j] * ldk + l, act) : R(0);
    vmj[j] = vm * dv * dv;
    iw1[j] = dv < R(0);
    up[j] = iw1[j] ? !(vm < R(0)) : (vm < R(0));
  }
  for (int n = ti.n0; n < ti.n1; ++n) {
    const int mode = T.tmode[n];
    if (mode == 0) continue;
    const long off = n * a.tstride;
    const R nuc = ld(a.nu + off, ic, act);
    const int set = mode == 1 ? 1 : 0;
    R L0 = R(0), P = R(0), Gx = R(0), Gy = R(0), Gq[3] = {R(0), R(0), R(0)};
    for_slots(np, [&](int j) {
      const R nun = ld(a.nu + off, (long)nb[j] * ldk + l, act);
      const R phi = iw1[j] ? vmj[j] * (nun - nuc) : vmj[j] * (nuc - nun);
      const R s = ld(T.tsv + n * T.vstride, (long)iv[j] * ldk + l, act);
      const bool hi_c = (mode == 2 && a.order == 1) ? iw1[j] : up[j];
      if (up[j]) L0 = L0 + (R(1) - s) * phi;
      const R hh = hi_c ? s * phi : R(0);
      P = P + hh;
      Gx = Gx + dxs[j] * hh;
      Gy = Gy + dys[j] * hh;
      if (a.order == 3)
        for (int t = 0; t < 3; ++t) Gq[t] = Gq[t] + m.w_os[set][t][c * NSLOT + j] * hh;
    });
    if (!body) { L0 = R(0); P = R(0); Gx = R(0); Gy = R(0); Gq[0] = Gq[1] = Gq[2] = R(0); }
    if (act) {
      a.lam0[off + ic] = a.lam0[off + ic] + L0;
      a.lpart[off + ic] = a.lpart[off + ic] + P;
      if (a.order > 1) { a.lgx[off + ic] = Gx; a.lgy[off + ic] = Gy; }
      if (a.order == 3)
        for (int t = 0; t < 3; ++t) a.lgq[t][off + ic] = Gq[t];
    }
  }
}

// scp0 += dtr sct / rho_old before the limited final stage (scalar_transport_rk.F90:
// 297-313), for tracers with pre[n] != 0 (its copy to levels 2..lpw-1 is empty here).
template <class R>
__global__ __launch_bounds__(256) void k_tr_lim_pre(DMesh<R> m, R *q0, const R *sct, const R *rho_old,
                                                    const int *pre, int ntr, long tstride, R dtr) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1;
  const bool act = k <= m.mza, body = act && k >= 2;
  const long ic = (long)c * m.ldk + l;
  const R ro = ldv(rho_old, ic, body, R(1));
  for (int n = 0; n < ntr; ++n) {
    if (!pre[n]) continue;
    const long off = n * tstride;
    if (body) q0[off + ic] = q0[off + ic] + dtr * sct[off + ic] / ro;
  }
}

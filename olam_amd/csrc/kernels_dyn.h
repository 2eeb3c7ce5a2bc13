// kernels_dyn.h -- prog_wrtv_rk: the acoustic RK step of W/V momentum, rho, rho*thil
// and press, with divergence/vorticity damping and vel_t3d.
//
// Reference: omodel/prog_wrtv_rk.f90, omodel/vort_div_damp.f90, omodel/vel_t3d.f90,
// omodel/grad_lib.f90.  Per RK stage four launches replace the reference's ten
// OpenMP loops: k_dyn_grad (grad_t2d of thil, vxe, vye, vze), k_dyn_w (fluxes_wrtv's
// face and vertical upwinding fused with prog_wrt_begs, including the FP64
// tridiagonal W solve by parallel cyclic reduction inside the wavefront),
// k_dyn_vn (prog_v_begs) and k_dyn_veln (vel_t3d_hex).
#pragma once
#ifndef OLAMGPU_BLOCKS_DYNW
#define OLAMGPU_BLOCKS_DYNW 8
#endif
#include "common.h"

// All dynamics kernels evaluate their expressions exactly as written (no FMA
// contraction): the W-momentum balance subtracts nearly equal pressure-gradient and
// buoyancy terms, so every ulp upstream of it shows in w.  The oracle does the same.

template <class R, class RP>
struct DynBuf {
  // prognostic state
  const R *thil_in;
  R *thil_out;
  R *wmc, *wc, *vxe, *vye, *vze, *vmc, *vc, *vmasc, *wmasc;
  RP *rho, *press;
  const R *alpha_press, *pwfac, *pvfac;
  // RK start values and volume-weighted short-step tendencies
  const RP *rho0, *rth0;
  const R *wmc0, *vmc0, *thilt_s, *vmxet_s, *vmyet_s, *vmzet_s, *vmt_s;
  R *vmxet_rk, *vmyet_rk, *vmzet_rk;
  // horizontal gradients: [thil, vxe, vye, vze] x [gx, gy]; order 3 adds the quadratic
  // terms [thil, vxe, vye, vze] x [xx, xy, yy]
  const R *g[8];
  const R *gq[4][3];
  // fluxes_wrtv_monot on the final stage for the fields with mono[q] != 0 (thil: q = 0,
  // winds: 1..3): low-order values scpt and the scales of scalar_monot_limiter
  const R *scpt[4], *ms1[4], *ms2[4];
  int mono[4];
  // nudging (mem_nudge): mass fluxes wmanud/vmanud added to the W and face mass fluxes
  // (nud_vm), the vertical nudging fluxes of nudflag == 1 (nud_w), the density
  // tendency rhot_nud (nud_rho)
  const R *wmanud, *vmanud;
  const RP *rhot_nud;
  int nud_vm, nud_w, nud_rho;
  int thil_order, wind_order, istage, nrk;
  double dt8;
  R fcor;
};

// R-precision power evaluated through an FP64 pow and rounded once, so the FP32
// result does not depend on which math library computes it (the oracle does the same).
__device__ __forceinline__ float pow_r(float x, float y) { return (float)pow((double)x, (double)y); }
__device__ __forceinline__ double pow_r(double x, double y) { return pow(x, y); }

// ---------------------------------------------------------------- turbulence
// The turbulence fields of the PBL scheme (mem_turb) and the per-face horizontal
// diffusion coefficients of comp_horiz_k; on = 0 until the host has set them.
template <class R>
struct DTurb {
  const R *vkh, *vkm;            // eddy diffusivities at W levels (nw x ldk)
  const R *akm_sfc;              // surface drag coefficient per column
  const R *dzit_bot;             // per level
  const int *khtop, *kmtop;      // top levels of heat / momentum diffusion per column
  R *akhodx, *akmodx;            // arv * K / dx per face and level (nv x ldk)
  int *khtopv, *kmtopv;          // per face
  int on;
};

// ---------------------------------------------------------------- setup loop
// prog_wrtv_rk.f90:165-251: RK start values and volume-weighted tendencies, then in the
// reference's order the implicit vertical eddy diffusion of heat and momentum
// (solve_eddy_diff_heat / _vxe, pbl_driver.f90:519-742, flat terrain: one surface
// level), the Rayleigh terms (vertical mixing of vxe/vye/vze near the top,
// rayf_mix_top_vxe mem_rayf.f90:177-220, and the relaxation of thil toward th01d), and
// the horizontal diffusion across the faces.  The column solves are FP64 PCR within
// the wave (the reference's tridiffo / tridv are FP32 Thomas).
template <class R, class RP>
__global__ __launch_bounds__(256) void k_dyn_setup_w(DMesh<R> m, DTurb<R> tb, R dts, const R *thil,
                                                     const RP *rho, const R *wmc, const R *thilt,
                                                     const R *vmxet, const R *vmyet, const R *vmzet, RP *rth0,
                                                     RP *rho0, R *wmc0, R *thilt_s, R *vmxet_s, R *vmyet_s,
                                                     R *vmzet_s, const R *vxe, const R *vye, const R *vze) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1, mza = m.mza;
  const bool act = k >= 2 && k <= mza;
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  // own-column values; lanes outside 2..mza read the padding of this column
  const R *v[3] = {vxe, vye, vze};
  const RP r = rho[ic];
  const R th = thil[ic];
  // the own winds are read only by the diffusion and the Rayleigh top mixing
  R vv[3] = {R(0), R(0), R(0)};
  if (tb.on || m.rayf_cofmix)
    for (int q = 0; q < 3; ++q) vv[q] = v[q][ic];
  const R v4 = (R)__builtin_nontemporal_load(m.volt + ic);
  R tt = __builtin_nontemporal_load(thilt + ic) * v4;
  R vm[3] = {__builtin_nontemporal_load(vmxet + ic) * v4, __builtin_nontemporal_load(vmyet + ic) * v4,
             __builtin_nontemporal_load(vmzet + ic) * v4};
  if (act) {
    st_nt(rth0, ic, (RP)th * r);
    st_nt(rho0, ic, r);
    st_nt(wmc0, ic, __builtin_nontemporal_load(wmc + ic));
  }
  if (tb.on) {
    const R forw_imp = R(1.1);
    const R vti = m.volti[ic], arw = m.arw[ic], dzim = m.dzim[l];
    // solve_eddy_diff_heat, pbl_driver.f90:519-605: levels ka..kmax = khtop + 1, fluxes at
    // W levels ka..khtop from the implicit system vctr5 x(k-1) + vctr6 x(k) + vctr7 x(k+1)
    const int khtop = tb.khtop[c];
    if (khtop >= 2) {
      const int kmax = khtop + 1;
      const R dtom = dts * vti / (R)r;
      const R dtomass = forw_imp * dtom;
      const R varp = th + dtom * tt;
      const R varp1 = vshdn(m.tall, varp, 1), dtomass1 = vshdn(m.tall, dtomass, 1);
      const bool row = k >= 2 && k <= khtop;
      double a = 0.0, b = 1.0, cc = 0.0, d = 0.0;
      if (row) {
        const R akodz = arw * tb.vkh[ic] * dzim;
        const R rhs = akodz * (varp - varp1);
        const R v5 = -akodz * dtomass;
        const R v7 = -akodz * dtomass1;
        const R v6 = R(1) - v5 - v7;
        a = (k > 2) ? (double)v5 : 0.0;
        b = (double)v6;
        cc = (k < khtop) ? (double)v7 : 0.0;
        d = (double)rhs;
      }
      const double x = vpcr(m.tall, l, a, b, cc, d);
      const R flux = row ? (R)x : R(0);
      const R flux_m1 = vshup(m.tall, flux, 1);
      if (k >= 2 && k <= kmax) tt = tt + (R)((double)flux_m1 - (double)flux);
    }
    // solve_eddy_diff_vxe, pbl_driver.f90:609-742: levels ka..kmax = max(kmtop + 1, ka),
    // surface drag on level ka
    {
      const int kmtop = tb.kmtop[c];
      const int kmax = max(kmtop + 1, 2);
      const R akodz = (k >= 2 && k <= kmtop) ? arw * tb.vkm[ic] * dzim : R(0);
      const R akodz_m1 = vshup(m.tall, akodz, 1);
      const bool row = k >= 2 && k <= kmax;
      const R dtom = (R)((RP)(dts * vti) / r);
      const R dtomass = forw_imp * dtom;
      R v5 = -dtomass * akodz_m1;
      const R v7 = -dtomass * akodz;
      R v6 = R(1) - v5 - v7;
      R vctr3 = R(0);
      if (k == 2) {
        vctr3 = tb.akm_sfc[c] * tb.dzit_bot[l];
        v6 = v6 + dtomass * vctr3;
      }
      double a = 0.0, b = 1.0, cc = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
      if (row) {
        a = (k > 2) ? (double)v5 : 0.0;
        b = (double)v6;
        cc = (k < kmax) ? (double)v7 : 0.0;
        d0 = (double)vv[0]; d1 = (double)vv[1]; d2 = (double)vv[2];
      }
      double x[3];
      vpcr3(m.tall, l, a, b, cc, d0, d1, d2, x[0], x[1], x[2]);
      for (int q = 0; q < 3; ++q) {
        const R so = row ? (R)x[q] : R(0);
        const R so1 = vshdn(m.tall, so, 1);
        const R v2 = (k >= 2 && k <= kmax - 1) ? akodz * (so - so1) : R(0);
        const R v2m = vshup(m.tall, v2, 1);
        if (k == 2) vm[q] = vm[q] + (v2m - v2 - vctr3 * so);
        else if (k >= 3 && k <= kmax) vm[q] = vm[q] + v2m - v2;
      }
    }
  }
  // rayf_mix_top_vxe: vflux(k) = arw0 * rayf_cofmix(k) * real(rho(k+1) + rho(k)) *
  // (v(k) - v(k+1)) on ka..mza-1, zero at ka-1 and mza; the tendency of level k gains
  // vflux(k-1) - vflux(k) on ka..mza.  Every lane takes part in the shifts.
  if (m.rayf_cofmix) {
    const bool fl = k >= m.krayfmix_bot && k <= mza - 1;
    const RP r1 = vshdn(m.tall, r, 1);
    const R fact = m.arw0[c] * m.rayf_cofmix[l] * (R)(r1 + r);
    for (int q = 0; q < 3; ++q) {
      const R a1 = vshdn(m.tall, vv[q], 1);
      const R fk = fl ? fact * (vv[q] - a1) : R(0);
      const R fkm1 = vshup(m.tall, fk, 1);
      if (act && k >= m.krayfmix_bot) vm[q] = vm[q] + fkm1 - fk;
    }
  }
  if (!act) return;
  if (m.rayf_cof && k >= m.krayf_bot) tt = tt + m.rayf_cof[l] * m.dn01d[l] * (m.th01d[l] - th) * v4;
  // horizontal diffusion, prog_wrtv_rk.f90:235-251
  if (tb.on) {
    const int np = m.w_npoly[c];
    for_slots(np, [&](int j) {
      if (j >= np) return;            // a pentagon's padded slot adds nothing in the reference
      const int iv = m.w_face[c * NSLOT + j];
      const long in = (long)m.w_nbr[c * NSLOT + j] * ldk + l, ie = (long)iv * ldk + l;
      if (k <= tb.khtopv[iv]) tt = tt + tb.akhodx[ie] * (thil[in] - th);
      if (k <= tb.kmtopv[iv]) {
        const R am = tb.akmodx[ie];
        for (int q = 0; q < 3; ++q) vm[q] = vm[q] + am * (v[q][in] - vv[q]);
      }
    });
  }
  st_nt(thilt_s, ic, tt);
  st_nt(vmxet_s, ic, vm[0]);
  st_nt(vmyet_s, ic, vm[1]);
  st_nt(vmzet_s, ic, vm[2]);
}

// comp_horiz_k_column, pbl_driver.f90:154-238 (flat terrain: lpv = 2; one resolution
// level, so akmin(mrl) is akmin): the face's diffusion tops, the stability bound
// akstab and arv * K / dx for momentum (akmodx) and heat (akhodx).
template <class R, class RP>
__global__ __launch_bounds__(256) void k_horiz_k(DMesh<R> m, DTurb<R> tb, const RP *rho, R dtl, R akmin) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int iv = wp_.item;
  if (iv >= m.nv) return;
  const int l = wp_.l, k = l + 1;
  const long ldk = m.ldk, ie = (long)iv * ldk + l;
  const int iw1 = m.v_iw[iv * 2], iw2 = m.v_iw[iv * 2 + 1];
  const long i1 = (long)iw1 * ldk + l, i2 = (long)iw2 * ldk + l;
  const R fact1 = R(0.95) / (dtl * (R)m.w_npoly[iw1]);
  const R fact2 = R(0.95) / (dtl * (R)m.w_npoly[iw2]);
  // every lane loads and shifts (k-1 levels of the W-level diffusivities)
  const R m1 = tb.vkm[i1], m2 = tb.vkm[i2], h1 = tb.vkh[i1], h2 = tb.vkh[i2];
  const R m1m = vshup(m.tall, m1, 1), m2m = vshup(m.tall, m2, 1), h1m = vshup(m.tall, h1, 1), h2m = vshup(m.tall, h2, 1);
  const RP r1 = rho[i1], r2 = rho[i2];
  const bool amin = akmin > R(1.e-7);
  int khv, kmv;
  R hk0 = R(0);
  if (amin) {
    const R bkmin = akmin * R(0.075) * pow_r(min(m.arw0[iw1], m.arw0[iw2]), (R)0.66666666f);
    khv = kmv = m.mza;
    hk0 = R(0.5) * (R)(r1 + r2) * bkmin;
  } else {
    khv = max(tb.khtop[iw1], tb.khtop[iw2]);
    khv = (khv >= 2) ? khv + 1 : 1;
    kmv = max(tb.kmtop[iw1], tb.kmtop[iw2]);
    kmv = (kmv >= 2) ? kmv + 1 : 1;
  }
  if (l == 0) { tb.khtopv[iv] = khv; tb.kmtopv[iv] = kmv; }
  if (k < 2 || k > m.mza) return;
  R akstab = R(0);
  if (k <= max(kmv, khv)) {
    const R stab1 = (R)(r1 * (RP)(R)m.volt[i1] * (RP)fact1);
    const R stab2 = (R)(r2 * (RP)(R)m.volt[i2] * (RP)fact2);
    akstab = min(stab1, stab2);
  }
  R am = R(0), ah = R(0);
  if (k <= kmv) {
    R hkm = R(0.25) * (m1 + m2 + m1m + m2m);
    if (amin) hkm = max(hkm, hk0);
    am = min(m.dniv[iv] * m.arv[ie] * hkm, akstab);
  }
  if (k <= khv) {
    R hkh = R(0.25) * (h1 + h2 + h1m + h2m);
    if (amin) hkh = max(hkh, hk0);
    ah = min(m.dniv[iv] * m.arv[ie] * hkh, akstab);
  }
  tb.akmodx[ie] = am;
  tb.akhodx[ie] = ah;
}


// ---------------------------------------------------------------- moist factors
// comp_alpha_press, timestep.f90:515-536 (W): alpha_press = pc1 * ((rdry + rvap * rr_v)
// * theta / thil)**cpocv on lpw..mza and pwfac = dzim * (1 - rw + rw**2) with rw the
// W-level interpolation of rr_w on lpw..mza-1.
template <class R>
__global__ __launch_bounds__(256) void k_alpha_w(DMesh<R> m, const R *rr_v, const R *rr_w, const R *theta,
                                                 const R *thil, R *alpha, R *pwfac, R pc1) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1;
  const long ic = (long)c * m.ldk + l;
  const R w0 = rr_w[ic], w1 = vshdn(m.tall, w0, 1);    // rr_w(k), rr_w(k+1); every lane shifts
  if (k < 2 || k > m.mza) return;
  const R rdry = R(287), rvap = R(461), cpocv = (R)(1004.0 / 717.0);
  alpha[ic] = pc1 * pow_r((rdry + rvap * rr_v[ic]) * theta[ic] / thil[ic], cpocv);
  if (k <= m.mza - 1) {
    const R rw = (R)m.zwgt_top8[l] * w1 + (R)m.zwgt_bot8[l] * w0;
    pwfac[ic] = m.dzim[l] * (R(1) - rw + rw * rw);
  }
}

// comp_alpha_press, timestep.f90:540-553 (V): pvfac = dniv * zfacit * (1 - rw + rw**2)
// with rw the mean rr_w of the face's two cells, on lpv..mza
template <class R>
__global__ __launch_bounds__(256) void k_alpha_v(DMesh<R> m, const R *rr_w, R *pvfac) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int iv = wp_.item;
  if (iv >= m.nv) return;
  const int l = wp_.l, k = l + 1;
  if (k < 2 || k > m.mza) return;
  const long ldk = m.ldk;
  const int iw1 = m.v_iw[iv * 2], iw2 = m.v_iw[iv * 2 + 1];
  const R rw = R(0.5) * (rr_w[(long)iw1 * ldk + l] + rr_w[(long)iw2 * ldk + l]);
  pvfac[(long)iv * ldk + l] = m.dniv[iv] * m.zfacit[l] * (R(1) - rw + rw * rw);
}

// ---------------------------------------------------------------- damping
// divh_damp divergence, vort_div_damp.f90:140-156.  div2d(:,:,1) is SAVEd by the
// reference and never reset, so the divergence accumulates onto what the previous
// call left there (the last Laplacian for level 3); reproduced as is.
template <class R>
__global__ __launch_bounds__(256) void k_div(DMesh<R> m, const R *vmc, const R *div_in, R *div_out) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1;
  if (k < 2 || k > m.mza) return;
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  const int np = m.w_npoly[c];
  R d = div_in[ic];
  for_slots(np, [&](int j) {
    const long iv = (long)m.w_face[c * NSLOT + j] * ldk + l;
    d = d - m.arv[iv] * m.w_dirv[c * NSLOT + j] * vmc[iv];
  });
  div_out[ic] = d * m.volti[ic];
}

// laplacian2d, grad_lib.f90:494-524
template <class R>
__global__ __launch_bounds__(256) void k_lap_w(DMesh<R> m, const R *src, R *dst) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1;
  if (k > m.mza) return;
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  const int np = m.w_npoly[c];
  R d = R(0);
  if (k >= 2) {
    const R s0 = src[ic];
    for_slots(np, [&](int j) { d = d + (src[(long)m.w_nbr[c * NSLOT + j] * ldk + l] - s0) * m.w_xxyy[c * NSLOT + j]; });
  }
  dst[ic] = d;
}

// vort_damp vorticity at M points, vort_div_damp.f90:304-336 (flat terrain)
template <class R>
__global__ __launch_bounds__(256) void k_vort(DMesh<R> m, const R *vc, R *vort) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int im = wp_.item;
  if (im >= m.nm) return;
  const int l = wp_.l, k = l + 1;
  if (k > m.mza) return;
  const long ldk = m.ldk;
  R v = R(0);
  if (k >= 2)
    for (int j = 0; j < 3; ++j) v = v + vc[(long)m.m_iv[im * 4 + j] * ldk + l] * m.m_vfact[im * 4 + j];
  vort[(long)im * ldk + l] = v;
}

// laplacian_m, grad_lib.f90:531-559
template <class R>
__global__ __launch_bounds__(256) void k_lap_m(DMesh<R> m, const R *src, R *dst) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int im = wp_.item;
  if (im >= m.nm) return;
  const int l = wp_.l, k = l + 1;
  if (k > m.mza) return;
  const long ldk = m.ldk, ic = (long)im * ldk + l;
  R d = R(0);
  if (k >= 2) {
    const R s0 = src[ic];
    for (int j = 0; j < 3; ++j)
      d = d + (src[(long)m.m_im[im * 4 + j] * ldk + l] - s0) * m.m_xxyym[im * 4 + j];
  }
  dst[ic] = d;
}

// k_div, k_lap_w, k_vort and k_lap_m with N points per wave (per wave pair for tall
// columns, no move crosses the two waves): the
// gathers of all N points are issued before any is reduced, with each kernel's own
// arithmetic.  Points past the last one read the wave's first point and store nothing.
template <class R, int N>
__global__ __launch_bounds__(256) void k_divn(DMesh<R> m, const R *vmc, const R *div_in, R *div_out) {
  NOCONTRACT
  const WPos p = wave_pos(m.tall);   // tall columns: the half from the wave index
  const int w = p.item;
  const int l = p.l, k = l + 1;
  if (N * w >= m.nw || k < 2 || k > m.mza) return;
  const long ldk = m.ldk;
  R d0[N], av[N][7], vv[N][7];
  int cs[N], nps[N];
  bool ok[N];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    ok[r] = N * w + r < m.nw;
    const int c = ok[r] ? N * w + r : N * w;
    cs[r] = c;
    nps[r] = m.w_npoly[c];
    d0[r] = div_in[(long)c * ldk + l];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      const long iv = (long)m.w_face[c * NSLOT + j] * ldk + l;
      const bool use = j < 6 || nps[r] > 6;
      av[r][j] = use ? m.arv[iv] : R(0);
      vv[r][j] = use ? vmc[iv] : R(0);
    }
  }
#pragma unroll
  for (int r = 0; r < N; ++r) {
    if (!ok[r]) continue;
    const int c = cs[r];
    const long ic = (long)c * ldk + l;
    R d = d0[r];
    for_slots(nps[r], [&](int j) { d = d - av[r][j] * m.w_dirv[c * NSLOT + j] * vv[r][j]; });
    div_out[ic] = d * m.volti[ic];
  }
}

template <class R, int N>
__global__ __launch_bounds__(256) void k_lap_wn(DMesh<R> m, const R *src, R *dst) {
  NOCONTRACT
  const WPos p = wave_pos(m.tall);   // tall columns: the half from the wave index
  const int w = p.item;
  const int l = p.l, k = l + 1;
  if (N * w >= m.nw || k > m.mza) return;
  const long ldk = m.ldk;
  R s0[N], sn[N][7];
  int cs[N], nps[N];
  bool ok[N];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    ok[r] = N * w + r < m.nw;
    const int c = ok[r] ? N * w + r : N * w;
    cs[r] = c;
    nps[r] = m.w_npoly[c];
    s0[r] = src[(long)c * ldk + l];
#pragma unroll
    for (int j = 0; j < 7; ++j)
      sn[r][j] = (j < 6 || nps[r] > 6) ? src[(long)m.w_nbr[c * NSLOT + j] * ldk + l] : R(0);
  }
#pragma unroll
  for (int r = 0; r < N; ++r) {
    if (!ok[r]) continue;
    const int c = cs[r];
    R d = R(0);
    if (k >= 2) for_slots(nps[r], [&](int j) { d = d + (sn[r][j] - s0[r]) * m.w_xxyy[c * NSLOT + j]; });
    dst[(long)c * ldk + l] = d;
  }
}

template <class R, int N>
__global__ __launch_bounds__(256) void k_vortn(DMesh<R> m, const R *vc, R *vort) {
  NOCONTRACT
  const WPos p = wave_pos(m.tall);   // tall columns: the half from the wave index
  const int w = p.item;
  const int l = p.l, k = l + 1;
  if (N * w >= m.nm || k > m.mza) return;
  const long ldk = m.ldk;
  R vj[N][3];
  int ims[N];
  bool ok[N];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    ok[r] = N * w + r < m.nm;
    const int im = ok[r] ? N * w + r : N * w;
    ims[r] = im;
#pragma unroll
    for (int j = 0; j < 3; ++j) vj[r][j] = vc[(long)m.m_iv[im * 4 + j] * ldk + l];
  }
#pragma unroll
  for (int r = 0; r < N; ++r) {
    if (!ok[r]) continue;
    const int im = ims[r];
    R v = R(0);
    if (k >= 2)
      for (int j = 0; j < 3; ++j) v = v + vj[r][j] * m.m_vfact[im * 4 + j];
    vort[(long)im * ldk + l] = v;
  }
}

template <class R, int N>
__global__ __launch_bounds__(256) void k_lap_mn(DMesh<R> m, const R *src, R *dst) {
  NOCONTRACT
  const WPos p = wave_pos(m.tall);   // tall columns: the half from the wave index
  const int w = p.item;
  const int l = p.l, k = l + 1;
  if (N * w >= m.nm || k > m.mza) return;
  const long ldk = m.ldk;
  R s0[N], sn[N][3];
  int ims[N];
  bool ok[N];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    ok[r] = N * w + r < m.nm;
    const int im = ok[r] ? N * w + r : N * w;
    ims[r] = im;
    s0[r] = src[(long)im * ldk + l];
#pragma unroll
    for (int j = 0; j < 3; ++j) sn[r][j] = src[(long)m.m_im[im * 4 + j] * ldk + l];
  }
#pragma unroll
  for (int r = 0; r < N; ++r) {
    if (!ok[r]) continue;
    const int im = ims[r];
    R d = R(0);
    if (k >= 2)
      for (int j = 0; j < 3; ++j) d = d + (sn[r][j] - s0[r]) * m.m_xxyym[im * 4 + j];
    dst[(long)im * ldk + l] = d;
  }
}


// V setup (prog_wrtv_rk.f90:256-262), then the Rayleigh friction of vmc toward vc03d
// (:268-284, global domain) and the damping tendencies (vort_div_damp.f90:169-185,
// 216-228 and 424-441), added to vmt_short in the reference's order.
// N faces per wave (a wave pair for tall columns): the loads of all N faces are issued
// before any tendency is formed.  Faces past the last one read the wave's first face
// and store nothing.
template <class R, class RP, int N>
__global__ __launch_bounds__(256) void k_dyn_setup_vn(DMesh<R> m, const R *vmc, const R *vmt, R *vmc0, R *vmt_s,
                                                      const R *divl, const R *fdivl, int kdivl, const R *divh,
                                                      const R *fdivh, const R *vort, const R *fvorth,
                                                      const RP *rho, const R *vc, const R *vc03d) {
  NOCONTRACT
  const WPos p = wave_pos(m.tall);   // tall columns: the half from the wave index
  const int w = p.item;
  const int l = p.l, k = l + 1;
  if (N * w >= m.nv || k > m.mza) return;
  const long ldk = m.ldk;
  struct Face {
    long ie;
    R mc, mt, c03, cc, dl1, dl2, fl, dh1, dh2, fh, vo1, vo2, fv;
    RP r1, r2;
  };
  Face f[N];
  bool ok[N];
  const bool rayf = m.rayf_cof && k >= m.krayf_bot;
#pragma unroll
  for (int r = 0; r < N; ++r) {
    ok[r] = N * w + r < m.nv;
    const int iv = ok[r] ? N * w + r : N * w;
    Face &F = f[r];
    F.ie = (long)iv * ldk + l;
    F.mc = vmc[F.ie];
    F.mt = vmt[F.ie];
    if (k >= 2) {
      const int iw1 = m.v_iw[iv * 2], iw2 = m.v_iw[iv * 2 + 1];
      const long i1 = (long)iw1 * ldk + l, i2 = (long)iw2 * ldk + l;
      if (rayf || vort) { F.r1 = rho[i1]; F.r2 = rho[i2]; }
      if (rayf) { F.c03 = vc03d[F.ie]; F.cc = vc[F.ie]; }
      if (divl && k >= kdivl) { F.fl = fdivl[F.ie]; F.dl1 = divl[i1]; F.dl2 = divl[i2]; }
      if (divh) { F.fh = fdivh[F.ie]; F.dh1 = divh[i1]; F.dh2 = divh[i2]; }
      if (vort) {
        const int im1 = m.v_im[iv * 2], im2 = m.v_im[iv * 2 + 1];
        F.fv = fvorth[iv];
        F.vo1 = vort[(long)im1 * ldk + l]; F.vo2 = vort[(long)im2 * ldk + l];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < N; ++r) {
    if (!ok[r]) continue;
    const Face &F = f[r];
    vmc0[F.ie] = F.mc;
    R t = F.mt;
    if (k >= 2) {
      if (rayf) t = (R)((RP)t + (RP)(m.rayf_cof[l] * R(0.5)) * (F.r1 + F.r2) * (RP)(F.c03 - F.cc));
      if (divl && k >= kdivl) t = t + F.fl * (F.dl2 - F.dl1);
      if (divh) t = t + F.fh * (F.dh2 - F.dh1);
      if (vort) t = t + F.fv * (F.vo2 - F.vo1) * (R)(F.r1 + F.r2);
    }
    vmt_s[F.ie] = t;
  }
}

// ---------------------------------------------------------------- gradients
// grad_t2d of thil, vxe, vye, vze (fluxes_wrtv, prog_wrtv_rk.f90:976-987)
template <class R>
__global__ __launch_bounds__(256) void k_dyn_grad(DMesh<R> m, const R *thil, const R *vxe, const R *vye,
                                                  const R *vze, R *g0, R *g1, R *g2, R *g3, R *g4, R *g5, R *g6,
                                                  R *g7, int do_thil, int do_wind) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1;
  const bool act = k <= m.mza, body = act && k >= 2;
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  const int np = m.w_npoly[c];
  const R *f[4] = {thil, vxe, vye, vze};
  R *gx[4] = {g0, g2, g4, g6};
  R *gy[4] = {g1, g3, g5, g7};
  for (int q = 0; q < 4; ++q) {
    if ((q == 0 && !do_thil) || (q > 0 && !do_wind)) continue;
    const R s0 = ld(f[q], ic, act);
    R ax = R(0), ay = R(0);
    for_slots(np, [&](int j) {
      const R d = ld(f[q], (long)m.w_nbr[c * NSLOT + j] * ldk + l, act) - s0;
      ax = ax + m.w_gxc[c * NSLOT + j] * d;
      ay = ay + m.w_gyc[c * NSLOT + j] * d;
    });
    if (!body) { ax = R(0); ay = R(0); }
    if (act) { st_nt(gx[q], ic, ax); st_nt(gy[q], ic, ay); }
  }
}

// grad_t2d_quadratic with bounded=.false. (prog_wrtv_rk.f90:983-984) of the order-3
// fields: the five least-squares terms per level, written to (gx, gy) and gq
template <class R>
struct DynGrad3 { R *g[4][5]; };

template <class R>
__global__ __launch_bounds__(256) void k_dyn_grad3(DMesh<R> m, const R *thil, const R *vxe, const R *vye,
                                                   const R *vze, DynGrad3<R> out, int do_thil, int do_wind,
                                                   int bounded_mask) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1;
  const bool act = k <= m.mza, body = act && k >= 2;
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  const int np = m.w_npoly[c];
  const R *f[4] = {thil, vxe, vye, vze};
  for (int q = 0; q < 4; ++q) {
    if ((q == 0 && !do_thil) || (q > 0 && !do_wind)) continue;
    const R s0 = ld(f[q], ic, act);
    // unbounded fit (xy_hu) except for a monotone-flux field on the final stage, whose
    // fluxes_wrtv_monot fits the bounded quadratic (:1191-1192)
    const int set = ((bounded_mask >> q) & 1) ? 0 : 1;
    R acc[5] = {R(0), R(0), R(0), R(0), R(0)};
    for_slots(np, [&](int j) {
      const R sc = ld(f[q], (long)m.w_nbr[c * NSLOT + j] * ldk + l, act) - s0;
#pragma unroll
      for (int t = 0; t < 5; ++t) acc[t] = acc[t] + sc * m.w_q[set][t][c * NSLOT + j];
    });
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      if (!body) acc[t] = R(0);
      if (act) out.g[q][t][ic] = acc[t];
    }
  }
}

// vertical upwinded face value at W level k of a column held one level per lane
// (fluxes_wrtv scalar_w_column :1085-1126 with grad_z_quadratic, bounded)
template <class R>
__device__ __forceinline__ R vert_upw(int tall, R q, R wsign, int k, int mza, R a11, R a21, R a12, R a22) {
  NOCONTRACT
  const bool inner = k >= 2 && k <= mza - 1;
  const R qp1 = vshdn(tall, q, 1);
  const R ds = inner ? qp1 - q : R(0);
  const R dsm1 = vshup(tall, ds, 1);
  const R scpb = q + ds * a11 + dsm1 * a21;
  const R scpt = q + ds * a12 + dsm1 * a22;
  const R scpb_p1 = vshdn(tall, scpb, 1);
  R u = (wsign < R(0)) ? scpb_p1 : scpt;
  if (k == 1) u = qp1;
  if (k == mza) u = q;
  return u;
}

// fluxes_wrtv_monot's vertical face values (scalar_wflux_monot :1380-1445 and the
// renormalisation :1235-1246): scpt upwinded by wmc plus the correction from the
// unbounded constant-gradient quadratic, scaled by the column's in/out factors.
template <class R>
__device__ __forceinline__ R vert_upw_monot(const DMesh<R> &m, R q, const R *scpt, const R *s1, const R *s2,
                                            long ic, bool act, R wmcc, R arw, int k, int mza) {
  NOCONTRACT
  const bool inner = k >= 2 && k <= mza - 1;
  R sb, st;
  lim_vrec(m, q, k, mza, 1, sb, st);
  const R sc = ld(scpt, ic, act), sc_p1 = vshdn(m.tall, sc, 1);
  const R low = (wmcc < R(0)) ? sc_p1 : sc;
  const R sb_p1 = vshdn(m.tall, sb, 1);
  const R hiw = inner ? ((wmcc < R(0)) ? sb_p1 : st) - low : R(0);
  const R sfw = inner ? arw * wmcc * hiw : R(0);
  const R s1c = ld(s1, ic, act), s2c = ld(s2, ic, act);
  const R s1p = vshdn(m.tall, s1c, 1), s2p = vshdn(m.tall, s2c, 1);
  const R scale = (sfw < R(0)) ? fmin(s1c, s2p) : fmin(s1p, s2c);
  const R qp1 = vshdn(m.tall, q, 1);
  R u = low + hiw * scale;
  if (k == 1) u = qp1;
  if (k == mza) u = q;
  return u;
}

// ---------------------------------------------------------------- prog_wrt_begs
// prog_wrtv_rk.f90:494-884 (global domain) fused with the face
// values of fluxes_wrtv's scalar_v_column (:1022-1081) and scalar_w_column.

// The FP32 default instance is held to 64 VGPRs (8 waves per SIMD; a few spills, 3 %
// faster at NXP=320, L=60); the other instances would spill far more under that bound.
// Q3: some field has horizontal order 3 (the quadratic terms are only kept then).
template <class R, class RP, bool MONO, bool NUD, bool Q3, bool TALL = false>
__global__ __launch_bounds__(256, (sizeof(R) == 4 && sizeof(RP) == 4 && !MONO && !NUD && !Q3) ? OLAMGPU_BLOCKS_DYNW : 1)
void k_dyn_w(DMesh<R> m, DynBuf<R, RP> b) {
  NOCONTRACT
  const WPos wp_ = wave_pos(TALL);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  const int np = m.w_npoly[c];
  const R cpocv = (R)(1004.0 / 717.0), rocv = (R)(287.0 / 717.0);
  const RP fr = RP(0.70), fp2 = RP(0.70), fp3 = RP(0.70);
  const RP pc2 = fp2 * (RP)cpocv, pc3 = fp3 * (RP)cpocv;
  const double dt8 = b.dt8;
  const R dts = (R)dt8;

  // own column
  const R thc = ld(b.thil_in, ic, act), vxc = ld(b.vxe, ic, act), vyc = ld(b.vye, ic, act),
          vzc = ld(b.vze, ic, act);
  R gxs[4] = {R(0), R(0), R(0), R(0)}, gys[4] = {R(0), R(0), R(0), R(0)};
  R gqs[4][3] = {};
  for (int q = 0; q < 4; ++q) {
    const int ord = q == 0 ? b.thil_order : b.wind_order;
    if (ord > 1) { gxs[q] = ld(b.g[2 * q], ic, act); gys[q] = ld(b.g[2 * q + 1], ic, act); }
    if (Q3 && ord == 3)
      for (int t = 0; t < 3; ++t) gqs[q][t] = ld(b.gq[q][t], ic, act);
  }
  const R qc[4] = {thc, vxc, vyc, vzc};
  // horizontal advective fluxes (:632-661)
  RP hflx = RP(0);
  R tt = ld_nt(b.thilt_s, ic, act), tx = ld_nt(b.vmxet_s, ic, act), ty = ld_nt(b.vmyet_s, ic, act),
    tz = ld_nt(b.vmzet_s, ic, act);
  // nudging mass flux wmanud carrying the vertically upwinded values (:605-621,
  // nudflag == 1 with nud_preserve_total_mass), before the face sums
  if (NUD && b.nud_w) {
    const R wn = ldv(b.wmanud, ic, act, R(0)), wn_m1 = vshup(TALL, wn, 1);
    const R wmc_n = ld(b.wmc, ic, act);
    R up[4];
    for (int q = 0; q < 4; ++q) {
      up[q] = vert_upw(TALL, qc[q], wmc_n, k, mza, m.aw11[l], m.aw21[l], m.aw12[l], m.aw22[l]);
      if constexpr (MONO)
        if (b.mono[q]) up[q] = vert_upw_monot(m, qc[q], b.scpt[q], b.ms1[q], b.ms2[q], ic, act, wmc_n, ld(m.arw, ic, act), k, mza);
    }
    const RP vt = inner ? (RP)(wn * up[0]) : RP(0);
    const R vx = inner ? wn * up[1] : R(0), vy = inner ? wn * up[2] : R(0), vz = inner ? wn * up[3] : R(0);
    const RP vt_m1 = vshup(TALL, vt, 1);
    const R vx_m1 = vshup(TALL, vx, 1), vy_m1 = vshup(TALL, vy, 1), vz_m1 = vshup(TALL, vz, 1);
    if (body) {
      hflx = (RP)(wn_m1 - wn);
      tt = (R)((RP)tt + vt_m1 - vt);
      tx = tx + vx_m1 - vx;
      ty = ty + vy_m1 - vy;
      tz = tz + vz_m1 - vz;
    }
  }
  // face mass fluxes first: a neighbour's values and gradients are gathered only if some
  // level takes that face's value from the neighbour (upwind side)
  R vmca_j[7], vmc_j[7], arv_j[7];
  bool sel_j[7];
  unsigned need = 0;
#pragma unroll
  for (int j = 0; j < 7; ++j) {       // all face loads first, then the ballots
    const long ie = (long)m.w_face[c * NSLOT + j] * ldk + l;
    vmc_j[j] = (j < 6 || np > 6) ? ldv(b.vmc, ie, act, R(0)) : R(0);
    arv_j[j] = (j < 6 || np > 6) ? ldv(m.arv, ie, act, R(0)) : R(0);
  }
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    const R dv = m.w_dirv[c * NSLOT + j];
    vmca_j[j] = (j < np) ? vmc_j[j] * arv_j[j] : R(0);
    if (NUD && b.nud_vm && j < np)   // vmca = vmc * arv + vmanud (:347-355)
      vmca_j[j] = vmca_j[j] + ldv(b.vmanud, (long)m.w_face[c * NSLOT + j] * ldk + l, act, R(0));
    const bool selj = (dv < R(0)) ? (vmc_j[j] < R(0)) : !(vmc_j[j] < R(0));
    if constexpr (NUD) sel_j[j] = selj;
    need |= (__ballot(selj && act) != 0 ? 1u : 0u) << j;
  }
  for_slots(np, [&](int j) {
    const int nbr = m.w_nbr[c * NSLOT + j];
    const R dv = m.w_dirv[c * NSLOT + j];
    const R dxs = m.w_dxs[c * NSLOT + j], dys = m.w_dys[c * NSLOT + j];
    const R dxn = m.w_dxn[c * NSLOT + j], dyn = m.w_dyn[c * NSLOT + j];
    const long in = (long)nbr * ldk + l;
    const R vmca = vmca_j[j];
    // upwind side from the sign of vmc; without nudging vmca = vmc * arv has that sign on
    // every level 2..mza (arv > 0 there), so it is taken from vmca instead of being kept
    // per face (fewer live registers)
    bool sel;
    if constexpr (NUD) sel = sel_j[j];
    else sel = (dv < R(0)) ? (vmca < R(0)) : !(vmca < R(0));
    // wave-uniform: every neighbour is gathered in the instances without order-3 or
    // monotone terms (the branches and held masks of the gating cost more than the
    // cache reads it saves); the others read only faces some level takes
    const bool take = (!Q3 && !MONO) || ((need >> j) & 1u);
    const bool nact = act && take;
    const long nbb = (long)nbr * ldk;
    R f[4];
    for (int q = 0; q < 4; ++q) {
      const int ord = q == 0 ? b.thil_order : b.wind_order;
      const R *src = q == 0 ? b.thil_in : (q == 1 ? b.vxe : (q == 2 ? b.vye : b.vze));
      // unpredicated neighbour loads: a lane above mza only feeds its own face sums,
      // which the vertical terms at mza select away
      const R qn = take ? ldl(src + nbb, (unsigned)l) : R(0);
      R fs = qc[q], fn = qn;
      if (ord > 1) {
        fs = qc[q] + dxs * gxs[q] + dys * gys[q];
        const R gxn = take ? ldl(b.g[2 * q] + nbb, (unsigned)l) : R(0);
        const R gyn = take ? ldl(b.g[2 * q + 1] + nbb, (unsigned)l) : R(0);
        fn = qn + dxn * gxn + dyn * gyn;
      }
      if (Q3 && ord == 3) {   // the unbounded fit at the bounded offsets xx0_v (:1057-1071)
        const long e = (long)c * NSLOT + j;
        fs = fs + m.w_os[0][0][e] * gqs[q][0] + m.w_os[0][1][e] * gqs[q][1] + m.w_os[0][2][e] * gqs[q][2];
        fn = fn + m.w_on[0][0][e] * ld(b.gq[q][0], in, nact) + m.w_on[0][1][e] * ld(b.gq[q][1], in, nact) +
             m.w_on[0][2][e] * ld(b.gq[q][2], in, nact);
      }
      f[q] = sel ? fn : fs;
      if constexpr (MONO) {
        if (b.mono[q]) {
          // fluxes_wrtv_monot: scpt of the upwind cell plus the renormalised correction
          // (:1326-1374, :1251-1266); at order 3 the bounded fit with xx0_vu offsets
          R hs = fs, hn = fn;
          if (Q3 && ord == 3) {
            const long e = (long)c * NSLOT + j;
            hs = qc[q] + dxs * gxs[q] + dys * gys[q];
            hs = hs + m.w_os[1][0][e] * gqs[q][0] + m.w_os[1][1][e] * gqs[q][1] + m.w_os[1][2][e] * gqs[q][2];
            hn = qn + dxn * ld(b.g[2 * q], in, nact) + dyn * ld(b.g[2 * q + 1], in, nact);
            hn = hn + m.w_on[1][0][e] * ld(b.gq[q][0], in, nact) + m.w_on[1][1][e] * ld(b.gq[q][1], in, nact) +
                 m.w_on[1][2][e] * ld(b.gq[q][2], in, nact);
          }
          const R low = sel ? ld(b.scpt[q], in, act) : ld(b.scpt[q], ic, act);
          const R hiv = (sel ? hn : hs) - low;
          const R sfv = vmca * hiv;
          const bool iw1side = dv < R(0);
          const R s1c = ld(b.ms1[q], ic, act), s2c = ld(b.ms2[q], ic, act);
          const R s1n = ld(b.ms1[q], in, act), s2n = ld(b.ms2[q], in, act);
          const R s1_1 = iw1side ? s1c : s1n, s1_2 = iw1side ? s1n : s1c;
          const R s2_1 = iw1side ? s2c : s2n, s2_2 = iw1side ? s2n : s2c;
          const R scale = (sfv < R(0)) ? fmin(s1_1, s2_2) : fmin(s1_2, s2_1);
          f[q] = low + hiv * scale;
        }
      }
    }
    if (dv > R(0)) {
      hflx = hflx + (RP)vmca;
      tt = tt + vmca * f[0]; tx = tx + vmca * f[1]; ty = ty + vmca * f[2]; tz = tz + vmca * f[3];
    } else {
      hflx = hflx - (RP)vmca;
      tt = tt - vmca * f[0]; tx = tx - vmca * f[1]; ty = ty - vmca * f[2]; tz = tz - vmca * f[3];
    }
  });
  // Coriolis, global (:675-681)
  const RP rho_c = act ? b.rho[ic] : RP(0);
  const double vol = act ? m.volt[ic] : 0.0;
  {
    const R mass = (R)(rho_c * (RP)vol);
    tx = tx + mass * b.fcor * vyc;
    ty = ty - mass * b.fcor * vxc;
  }
  // vertical upwinded values (own column, wmc sign)
  const R wmcc = ld(b.wmc, ic, act);
  const R a11 = m.aw11[l], a21 = m.aw21[l], a12 = m.aw12[l], a22 = m.aw22[l];
  R upws[4];
  for (int q = 0; q < 4; ++q) {
    upws[q] = vert_upw(TALL, qc[q], wmcc, k, mza, a11, a21, a12, a22);
    if constexpr (MONO)
      if (b.mono[q]) upws[q] = vert_upw_monot(m, qc[q], b.scpt[q], b.ms1[q], b.ms2[q], ic, act, wmcc, ld(m.arw, ic, act), k, mza);
  }
  const R tupw = upws[0], xupw = upws[1], yupw = upws[2], zupw = upws[3];
  // explicit density and theta tendencies (:687-723)
  const R vti = ld(m.volti, ic, act);
  const R alp = ld_nt(b.alpha_press, ic, act);
  const RP rhothil = rho_c * (RP)thc;
  const RP b10 = (RP)dt8 * (RP)vti;
  const RP b5 = body ? (RP)(alp * pow_r((R)rhothil, rocv)) : RP(0);
  const R b15 = (R)(b10 * b5);
  RP delex_rho = b10 * hflx;
  RP delex_rt = b10 * (RP)tt;
  if (NUD && b.nud_rho)   // :703-706, in FP64 as the reference's real(r8) terms
    delex_rho = (RP)((double)delex_rho + dt8 * (double)ld(b.rhot_nud, ic, act));
  if (b.istage > 1) {
    delex_rho = delex_rho + ld(b.rho0, ic, act) - rho_c;
    delex_rt = delex_rt + ld_nt(b.rth0, ic, act) - rhothil;
  }
  const RP press_ex = b5 * (rhothil + pc2 * delex_rt);
  const RP rho_ex = rho_c + fr * delex_rho;
  // explicit W momentum (:727-745), W levels ka..mza-1
  const RP pex_p1 = vshdn(TALL, press_ex, 1), rex_p1 = vshdn(TALL, rho_ex, 1);
  const R tx_p1 = vshdn(TALL, tx, 1), ty_p1 = vshdn(TALL, ty, 1), tz_p1 = vshdn(TALL, tz, 1);
  const R pwf = ld_nt(b.pwfac, ic, act), vwi = ld(m.volwi, ic, act);
  R delex_wm = R(0);
  if (inner)
    delex_wm = ld_nt(b.wmc0, ic, act) +
               dts * ((R)(press_ex - pex_p1) * pwf -
                      (R)((RP)m.gdz_wgtm8[l] * rho_ex + (RP)m.gdz_wgtp8[l] * rex_p1) +
                      (m.wnx[c] * (tx + tx_p1) + m.wny[c] * (ty + ty_p1) + m.wnz[c] * (tz + tz_p1)) * vwi);
  // implicit W matrix (:749-769); b20 = 1 + dtsm * rayf_cofw (:46-52), 1 without
  // Rayleigh friction on W
  const R b20 = m.b20 ? m.b20[l] : R(1);
  const R c8 = dts * (R)pc2;
  const R c9 = -dts * dts * (R)fr;
  const R wcc = ld(b.wc, ic, act);
  const R wc_m1 = vshup(TALL, wcc, 1), wc_p1 = vshdn(TALL, wcc, 1);
  const R b15_p1 = vshdn(TALL, b15, 1), vti_p1 = vshdn(TALL, vti, 1);
  const R arw = ld(m.arw, ic, act), arw_m1 = vshup(TALL, arw, 1), arw_p1 = vshdn(TALL, arw, 1);
  const R tupw_m1 = vshup(TALL, tupw, 1), tupw_p1 = vshdn(TALL, tupw, 1);
  double ma = 0.0, mb = 1.0, mc = 0.0, md = 0.0;
  if (inner) {
    const R b7 = dts * vwi;
    const R b8 = c8 * pwf;
    const R b21 = b7 * wc_m1;
    const R b22 = b7 * wc_p1;
    const R b23 = b8 * b15;
    const R b24 = b8 * b15_p1;
    const R b25 = c9 * vti * m.gdz_wgtm[l];
    const R b26 = c9 * vti_p1 * m.gdz_wgtp[l];
    ma = (double)(-arw_m1 * (b21 + b25 + b23 * tupw_m1));
    mb = (double)(b20 + arw * (b25 - b26 + (b23 + b24) * tupw));
    mc = (double)(arw_p1 * (b22 + b26 - b24 * tupw_p1));
    md = (double)delex_wm;
  }
  // tridiffo (tridiag_lib.f90:10-36) replaced by FP64 PCR within the wavefront
  const double wmf = vpcr(TALL, l, ma, mb, mc, md);
  const R wmff = (R)wmf;
  R wma = inner ? wmff * arw : R(0);
  if (NUD && b.nud_vm && inner) wma = wma + b.wmanud[ic];                                 // :782-788
  const RP vflux_thil = (RP)(wma * tupw);
  const R vfx = wma * xupw, vfy = wma * yupw, vfz = wma * zupw;
  if (b.istage == b.nrk && inner) b.wmasc[ic] = b.wmasc[ic] + wma;
  // T-level updates (:818-844)
  const R wma_m1 = vshup(TALL, wma, 1);
  const RP vft_m1 = vshup(TALL, vflux_thil, 1);
  const R vfx_m1 = vshup(TALL, vfx, 1), vfy_m1 = vshup(TALL, vfy, 1), vfz_m1 = vshup(TALL, vfz, 1);
  RP rho_new = rho_c;
  R thil_new = thc;
  if (body) {
    rho_new = ld(b.rho0, ic, act) + (RP)dt8 * (RP)vti * (hflx + (RP)wma_m1 - (RP)wma);
    delex_rt = delex_rt + b10 * (vft_m1 - vflux_thil);
    st_nt(b.press, ic, b5 * (rhothil + pc3 * delex_rt));
    thil_new = (R)((rhothil + delex_rt) / rho_new);
    b.rho[ic] = rho_new;
    st_nt(b.vmxet_rk, ic, tx + vfx_m1 - vfx);
    st_nt(b.vmyet_rk, ic, ty + vfy_m1 - vfy);
    st_nt(b.vmzet_rk, ic, tz + vfz_m1 - vfz);
  }
  // W-level updates (:869-882)
  const RP rho_new_p1 = vshdn(TALL, rho_new, 1);
  R wc_new = wcc;
  if (inner) {
    st_nt(b.wmc, ic, wmff);
    wc_new = wmff / (R)((RP)m.zwgt_bot8[l] * rho_new + (RP)m.zwgt_top8[l] * rho_new_p1);
  }
  const R wc_new_p1 = vshdn(TALL, wc_new, 1);
  const R thil_new_p1 = vshdn(TALL, thil_new, 1);
  if (k == 1) { wc_new = wc_new_p1; thil_new = thil_new_p1; }
  if (act && k <= mza - 1) st_nt(b.wc, ic, wc_new);
  if (act) st_nt(b.thil_out, ic, thil_new);
}

// ---------------------------------------------------------------- fluxes_wrtv_monot
// prog_wrtv_rk.f90:1133-1554 on the final stage, for the fields with mono[q] != 0.
// The RK-start copy of the field advanced by its long-step tendency (:1174-1181).
template <class R, class RP>
struct DynMono {
  const R *q0[4], *st[4];
  R *scpt[4], *s1[4], *s2[4];
  int on[4];
};

template <class R, class RP>
__global__ __launch_bounds__(256) void k_dyn_monot_pre(DMesh<R> m, DynMono<R, RP> mo, const RP *rho0, R dts) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1;
  if (k < 2 || k > m.mza) return;
  const long ic = (long)c * m.ldk + l;
  const R vti = m.volti[ic], r0 = (R)rho0[ic];
  for (int q = 0; q < 4; ++q)
    if (mo.on[q]) mo.scpt[q][ic] = mo.q0[q][ic] + dts * vti * mo.st[q][ic] / r0;
}

// scalar_vflux_monot / scalar_wflux_monot / scalar_monot_limiter (:1295-1550): the
// in/out renormalisation scales of each column and level.  Face values come from the
// upwind side's reconstruction (bounded fit, xx0_vu offsets at order 3), the vertical
// ones from the unbounded constant-gradient quadratic; rho is the density before this
// stage's column update.
template <class R, class RP>
__global__ __launch_bounds__(256) void k_dyn_monot_scales(DMesh<R> m, DynBuf<R, RP> b, DynMono<R, RP> mo) {
  NOCONTRACT
  const WPos wp_ = wave_pos(m.tall);
  const int c = wp_.item;
  if (c >= m.nw) return;
  const int l = wp_.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza, body = act && k >= 2;
  const bool inner = k >= 2 && k <= mza - 1;
  const long ldk = m.ldk, ic = (long)c * ldk + l;
  const int np = m.w_npoly[c];
  const R eps0 = LimConst<R>::eps0, eps1 = LimConst<R>::eps1;
  const R onep = R(1) + eps1;
  const R dtr = (R)b.dt8, dtrp = onep * dtr;
  R vmca_j[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    const long ie = (long)m.w_face[c * NSLOT + j] * ldk + l;
    vmca_j[j] = (j < np) ? ld(b.vmc, ie, act) * ld(m.arv, ie, act) : R(0);
    if (b.nud_vm && j < np) vmca_j[j] = vmca_j[j] + ld(b.vmanud, ie, act);
  }
  const R wmcc = ld(b.wmc, ic, act), arw = ld(m.arw, ic, act);
  const R wmca = inner ? wmcc * arw : R(0);
  const R wmca_m1 = vshup(m.tall, wmca, 1);
  const R vti = ld(m.volti, ic, act);
  const R rhos = act ? (R)b.rho[ic] : R(1);
  for (int q = 0; q < 4; ++q) {
    if (!mo.on[q]) continue;
    const int ord = q == 0 ? b.thil_order : b.wind_order;
    const R *src = q == 0 ? b.thil_in : (q == 1 ? b.vxe : (q == 2 ? b.vye : b.vze));
    const R qc = ld(src, ic, act);
    const R sc = ld(mo.scpt[q], ic, act);
    R gx = R(0), gy = R(0), gq[3] = {R(0), R(0), R(0)};
    if (ord > 1) { gx = ld(b.g[2 * q], ic, act); gy = ld(b.g[2 * q + 1], ic, act); }
    if (ord == 3)
      for (int t = 0; t < 3; ++t) gq[t] = ld(b.gq[q][t], ic, act);
    // vertical low-order and high-order values and fluxes (scalar_wflux_monot)
    R sb, st;
    lim_vrec(m, qc, k, mza, 1, sb, st);
    const R sc_p1 = vshdn(m.tall, sc, 1), sc_m1 = vshup(m.tall, sc, 1), sb_p1 = vshdn(m.tall, sb, 1), qp1 = vshdn(m.tall, qc, 1);
    R upwl = (wmcc < R(0)) ? sc_p1 : sc;
    if (k == 1) upwl = qp1;
    if (k == mza) upwl = qc;
    const R hiw = inner ? ((wmcc < R(0)) ? sb_p1 : st) - upwl : R(0);
    const R sfw = inner ? arw * wmcc * hiw : R(0);
    const R sfw_m1 = vshup(m.tall, sfw, 1), upwl_m1 = vshup(m.tall, upwl, 1);
    // upwind-biased extrema (vertical)
    R smax = sc, smin = sc;
    if (k >= 3 && wmca_m1 > R(0)) { smax = fmax(smax, sc_m1); smin = fmin(smin, sc_m1); }
    if (inner && wmca < R(0)) { smax = fmax(smax, sc_p1); smin = fmin(smin, sc_p1); }
    // horizontal faces
    R fl = R(0), fin = R(0), fout = R(0);
    for_slots(np, [&](int j) {
      const int nbr = m.w_nbr[c * NSLOT + j];
      const long in = (long)nbr * ldk + l;
      const R d = m.w_dirv[c * NSLOT + j];
      const R vm = vmca_j[j];
      const bool iw1side = d < R(0);
      const bool up_c = iw1side ? !(vm < R(0)) : (vm < R(0));
      const R scn = ld(mo.scpt[q], in, act);
      const R low = up_c ? sc : scn;
      R high;
      const long e = (long)c * NSLOT + j;
      if (up_c) {
        high = qc;
        if (ord > 1) high = qc + m.w_dxs[e] * gx + m.w_dys[e] * gy;
        if (ord == 3) high = high + m.w_os[1][0][e] * gq[0] + m.w_os[1][1][e] * gq[1] + m.w_os[1][2][e] * gq[2];
      } else {
        const R qn = ld(src, in, act);
        high = qn;
        if (ord > 1) high = qn + m.w_dxn[e] * ld(b.g[2 * q], in, act) + m.w_dyn[e] * ld(b.g[2 * q + 1], in, act);
        if (ord == 3)
          high = high + m.w_on[1][0][e] * ld(b.gq[q][0], in, act) + m.w_on[1][1][e] * ld(b.gq[q][1], in, act) +
                 m.w_on[1][2][e] * ld(b.gq[q][2], in, act);
      }
      const R sfv = vm * (high - low);
      fl = fl + d * vm * (low - sc);
      fin = fin + fmax(d * sfv, R(0));
      fout = fout + fmin(d * sfv, R(0));
      if (d * vm > R(0)) { smax = fmax(smax, scn); smin = fmin(smin, scn); }
    });
    // vertical flux divergence and the scales (scalar_monot_limiter :1531-1548)
    fl = fl + wmca_m1 * (upwl_m1 - sc) - wmca * (upwl - sc);
    fin = fin + fmax(sfw_m1, R(0)) - fmin(sfw, R(0));
    fout = fout + fmin(sfw_m1, R(0)) - fmax(sfw, R(0));
    const R rlow = sc * rhos + dtr * vti * fl;
    const R flux_in = fmax(dtrp * vti * fin, eps0);
    const R flux_out = fmin(dtrp * vti * fout, -eps0);
    const R epsr = eps1 * fabs(rlow);
    const R sc_in = fmax(rhos * smax - (rlow + epsr) - eps0, R(0));
    const R sc_out = fmin(rhos * smin - (rlow - epsr) + eps0, R(0));
    if (body) {
      mo.s1[q][ic] = fmin(R(1), sc_in / fmax(flux_in, R(0.01) * sc_in));
      mo.s2[q][ic] = fmin(R(1), sc_out / fmin(flux_out, R(0.01) * sc_out));
    }
  }
}

// ---------------------------------------------------------------- prog_v_begs
// prog_wrtv_rk.f90:888-946.  At stage nrk-1 the new vmc*arv is what the last stage
// uses as vmca, so it is added to vmasc here (:356-365).
// N faces per wave (a wave pair for tall columns): the loads of all N faces are issued
// before any is computed, so a wave keeps N times the loads in flight.  Faces past the
// last one read the wave's first face and store nothing.
template <class R, class RP, int N>
__global__ __launch_bounds__(256) void k_dyn_vn(DMesh<R> m, DynBuf<R, RP> b, int acc_vmasc) {
  NOCONTRACT
  const WPos p = wave_pos(m.tall);   // tall columns: the half from the wave index
  const int w = p.item;
  const int l = p.l, k = l + 1;
  if (N * w >= m.nv || k < 2 || k > m.mza) return;
  const long ldk = m.ldk;
  const R dts = (R)b.dt8;
  struct Face {
    long ie, i1, i2;
    RP p1, p2, r1, r2;
    R pv, v0, vt, x1, x2, y1, y2, z1, z2, vvi, nx, ny, nz;
  };
  Face f[N];
  bool ok[N];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    ok[r] = N * w + r < m.nv;
    const int iv = ok[r] ? N * w + r : N * w;
    Face &F = f[r];
    F.ie = (long)iv * ldk + l;
    const int iw1 = m.v_iw[iv * 2], iw2 = m.v_iw[iv * 2 + 1];
    F.i1 = (long)iw1 * ldk + l; F.i2 = (long)iw2 * ldk + l;
    F.p1 = b.press[F.i1]; F.p2 = b.press[F.i2];
    F.pv = __builtin_nontemporal_load(b.pvfac + F.ie);
    F.v0 = __builtin_nontemporal_load(b.vmc0 + F.ie);
    F.vt = __builtin_nontemporal_load(b.vmt_s + F.ie);
    F.x1 = b.vmxet_rk[F.i1]; F.x2 = b.vmxet_rk[F.i2];
    F.y1 = b.vmyet_rk[F.i1]; F.y2 = b.vmyet_rk[F.i2];
    F.z1 = b.vmzet_rk[F.i1]; F.z2 = b.vmzet_rk[F.i2];
    F.vvi = m.volvi[F.ie];
    F.r1 = b.rho[F.i1]; F.r2 = b.rho[F.i2];
    F.nx = m.vnx[iv]; F.ny = m.vny[iv]; F.nz = m.vnz[iv];
  }
#pragma unroll
  for (int r = 0; r < N; ++r) {
    if (!ok[r]) continue;
    const Face &F = f[r];
    const R pgf = (R)(F.p1 - F.p2) * F.pv;
    const R v = F.v0 + dts * (F.vt + pgf + (F.nx * (F.x1 + F.x2) + F.ny * (F.y1 + F.y2) + F.nz * (F.z1 + F.z2)) * F.vvi);
    st_nt(b.vmc, F.ie, v);
    st_nt(b.vc, F.ie, R(2) * v / (R)(F.r1 + F.r2));
    if (acc_vmasc) b.vmasc[F.ie] = b.vmasc[F.ie] + (b.nud_vm ? v * m.arv[F.ie] + b.vmanud[F.ie] : v * m.arv[F.ie]);
  }
}

// ---------------------------------------------------------------- vel_t3d_hex
// vel_t3d.f90:39-108 (no closed faces on flat terrain).
// N columns per wave (a wave pair for tall columns): the loads of all N columns are
// issued before any is reduced.  Columns past the last one read the wave's first column
// and store nothing.
template <class R, class RP, int N>
__global__ __launch_bounds__(256) void k_dyn_veln(DMesh<R> m, DynBuf<R, RP> b) {
  NOCONTRACT
  const WPos p = wave_pos(m.tall);   // tall columns: the half from the wave index
  const int w = p.item;
  if (N * w >= m.nw) return;
  const int l = p.l, k = l + 1, mza = m.mza;
  const bool act = k <= mza;
  const long ldk = m.ldk;
  R wcc[N], vj[N][7];
  int cs[N], nps[N];
  bool ok[N];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    ok[r] = N * w + r < m.nw;
    const int c = ok[r] ? N * w + r : N * w;
    cs[r] = c;
    nps[r] = m.w_npoly[c];
    wcc[r] = ld(b.wc, (long)c * ldk + l, act);
#pragma unroll
    for (int j = 0; j < 7; ++j)
      vj[r][j] = (j < 6 || nps[r] > 6) ? ld(b.vc, (long)m.w_face[c * NSLOT + j] * ldk + l, act) : R(0);
  }
#pragma unroll
  for (int r = 0; r < N; ++r) {
    const int c = cs[r];
    const long ic = (long)c * ldk + l;
    const R wst = vshup(m.tall, wcc[r], 1) + wcc[r];   // every item shifts, in both waves
    R x = wst * (m.wnx[c] * R(0.5)), y = wst * (m.wny[c] * R(0.5)), z = wst * (m.wnz[c] * R(0.5));
    for_slots(nps[r], [&](int j) {
      const R v = vj[r][j];
      x = x + m.w_ecx[c * NSLOT + j] * v;
      y = y + m.w_ecy[c * NSLOT + j] * v;
      z = z + m.w_ecz[c * NSLOT + j] * v;
    });
    const R x2 = vshdn(m.tall, x, 1), y2 = vshdn(m.tall, y, 1), z2 = vshdn(m.tall, z, 1);
    if (k == 1) { x = x2; y = y2; z = z2; }
    if (act && ok[r]) { st_nt(b.vxe, ic, x); st_nt(b.vye, ic, y); st_nt(b.vze, ic, z); }
  }
}

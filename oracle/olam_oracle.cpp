// olam_oracle.cpp -- CPU restatement of OLAM's RK3 dynamical core (TEST INFRASTRUCTURE).
//
// This file is the oracle: a from-scratch C++ restatement of the reference
// algorithm, written loop for loop after the Fortran so that each function can be
// checked against the file:line it follows.  It exists only to check the MI355X
// implementation (tests/, __graft_entry__.smoke()) and to time the CPU baseline
// (bench.py's cpu_baseline leg).  The product library never links or calls it.
//
// Precision is a template policy <R, RP, RT>: R = bulk stencil type, RP = type of
// rho/press and the prog_wrt_begs temporaries the reference keeps in real(r8),
// RT = type of the vertical tridiagonal solve.  Modes: F32 <float,float,double>,
// MIXED <float,double,double>, F64 <double,double,double>, and REF
// <float,double,float> which is the reference's own type map (prog_wrtv_rk.f90:557-580,
// tridiag_lib.f90:16-21).
//
// Parity pins (which test checks what) are listed in the header of
// tests/test_oracle.py.  Arrays use the reference's Fortran layout: (mza, n)
// column-major, 1-based, dummy index 1.
//
// Restricted to flat terrain (lpw = lpv = lpm = 2), horizontal orders 1, 2 and 3,
// vertical order 3 (iorderv = 3), the global domain (mdomain /= 4), no nudging or
// eddy diffusion (their switches are off in every configuration exercised here).

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <omp.h>

#include "../include/olamgpu.h"

namespace {

template <class T>
struct F2 {  // Fortran (m1, n) array, 1-based
  std::vector<T> v;
  int m1 = 0, n = 0;
  void alloc(int m1_, int n_, T init = T(0)) { m1 = m1_; n = n_; v.assign((size_t)m1 * n, init); }
  inline T &operator()(int k, int i) { return v[(size_t)(i - 1) * m1 + (k - 1)]; }
  inline const T &operator()(int k, int i) const { return v[(size_t)(i - 1) * m1 + (k - 1)]; }
  T *col(int i) { return &v[(size_t)(i - 1) * m1]; }
};

template <class T, class S>
void load2(F2<T> &a, const S *src, int m1, int n) {
  a.alloc(m1, n);
  if (!src) return;
  for (size_t i = 0; i < (size_t)m1 * n; ++i) a.v[i] = (T)src[i];
}
template <class T, class S>
void load1(std::vector<T> &a, const S *src, int n) {
  a.assign(n + 1, T(0));  // 1-based
  if (!src) return;
  for (int i = 0; i < n; ++i) a[i + 1] = (T)src[i];
}

template <class R_, class RP_, class RT_>
struct Prec { using R = R_; using RP = RP_; using RT = RT_; };

// FP32 power through an FP64 pow, rounded once (the reference's real ** real,
// prog_wrtv_rk.f90:691, evaluated without a math-library dependent last bit)
inline float pow_r(float x, float y) { return (float)std::pow((double)x, (double)y); }
inline double pow_r(double x, double y) { return std::pow(x, y); }
// Fortran real**integer: repeated products in the type of x
template <class T> T powi(T x, int n) { T r = T(1); for (int i = 0; i < n; ++i) r = r * x; return r; }

// tridiffo, tridiag_lib.f90:10-36: Thomas algorithm on rows ka..kz (arrays indexed by
// row), cim1 = sub-diagonal, ci = diagonal, cip1 = super-diagonal.
template <class RT>
void tridiffo_t(int ka, int kz, const RT *cim1, const RT *ci, const RT *cip1, const RT *rhs, RT *soln) {
  std::vector<RT> scr1(kz + 2);
  RT cji = RT(1) / ci[ka];
  soln[ka] = rhs[ka] * cji;
  for (int k = ka + 1; k <= kz; ++k) {
    scr1[k] = cji * cip1[k - 1];
    cji = RT(1) / (ci[k] - cim1[k] * scr1[k]);
    soln[k] = cji * (rhs[k] - cim1[k] * soln[k - 1]);
  }
  for (int k = kz - 1; k >= ka; --k) soln[k] = soln[k] - scr1[k + 1] * soln[k + 1];
}

struct Base {
  virtual ~Base() {}
  virtual int upload_state(const olamgpu_state *) = 0;
  virtual int download_state(olamgpu_state *) = 0;
  virtual int upload_scalar(int, const float *, const float *) = 0;
  virtual int download_scalar(int, float *) = 0;
  virtual int upload_adj(int, const float *) = 0;
  virtual int download_adj(int, float *) = 0;
  virtual void init_wrtv_rk() = 0;
  virtual void comp_alpha_press(const float *, const float *, const float *) = 0;
  virtual int set_turbulence(const float *, const float *, const int32_t *, const int32_t *, const float *) = 0;
  virtual int set_nudging(const float *, const float *, const double *) = 0;
  virtual void tend0() = 0;
  virtual void prog_wrtv_rk() = 0;
  virtual void timeavg_momsc() = 0;
  virtual void scalar_transport_rk(const float *rho_old) = 0;
  virtual void scalar_transport_rk_adj() = 0;
  virtual void mass_sums(double *) = 0;
  virtual void cfl_report(double *) = 0;
  virtual void tracer_maxmin(float *) = 0;
  virtual int set_pdef(int, int) = 0;
  virtual int set_sgsmix(int, int) = 0;
  virtual void set_frozen(int) = 0;
  virtual int tape_scales(int, int, double *) = 0;
  virtual int recon(const float *, int, int, int, double *, double *) = 0;
  virtual double rhothil_sum() = 0;
  virtual void rayf_profiles(float *, int32_t *) = 0;
  virtual int setup_tendencies(double *) = 0;
  std::string err;
  // Test hook: a deliberate mistake planted in one term, so a pin can show that it
  // detects that mistake (0 = none; the codes are listed at oracle_plant below).
  int plant = 0;
};

template <class P>
struct Model : Base {
  using R = typename P::R;
  using RP = typename P::RP;
  using RT = typename P::RT;

  int mza, mwa, mva, mma;
  olamgpu_params prm;

  // ---- tables (mem_ijtabs.f90:45-114) ----
  std::vector<int> npoly;
  F2<int> w_iv, w_iw, w_im, v_iw, v_im, m_iv, m_iw, m_im;
  F2<R> dirv, ecx, ecy, ecz, gxc, gyc, xx_yy, dxps, dyps, xx_yy_m;
  // order-3 reconstruction (mem_adv.f90:11-17): xy_h / xy_hu (5, 7, mwa) as (35, mwa)
  F2<R> xyh, xyhu, xx0v, xy0v, yy0v, xx0vu, xy0vu, yy0vu;
  // ---- metrics ----
  std::vector<R> wnx, wny, wnz, wnxo2, wnyo2, wnzo2, arw0, vnx, vny, vnz, dnu, dniu, dnv, dniv, arm0;
  std::vector<R> dzim, dzit, zfacit, gdz_wgtm, gdz_wgtp;
  std::vector<double> zwgt_top8, zwgt_bot8, gdz_wgtm8, gdz_wgtp8;
  std::vector<R> aw;   // a_w(k,i,j) at aw[((j-1)*2 + (i-1))*mza + k-1]
  std::vector<R> awu;  // a_wu, same layout (unbounded vertical quadratic)
  F2<R> arw, volti, volwi, arv, volvi;
  F2<double> volt;
  // ---- state (mem_basic, mem_tend) ----
  F2<R> thil, wmc, wc, vxe, vye, vze, vmc, vc, vmasc, wmasc, alpha_press, pwfac, pvfac;
  F2<R> thilt, vmxet, vmyet, vmzet, vmt, rho_old;
  F2<RP> rho, press;
  std::vector<F2<R>> scp, sct, lam;
  std::vector<int> pdef;                    // scalar_tab(n)%pdef (var_tables.f90:58)
  std::vector<int> sgsmix;                  // scalar_tab(n)%do_sgsmix (var_tables.f90:59)
  // frozen limiter scales of the last forward step's final stage, per tracer
  std::vector<F2<R>> tape_sv, tape_sw;      // per V face / per W level
  std::vector<char> tape_lim;               // 0 none, 1 monotone, 2 positive definite
  bool frozen_lim = false;                  // test hook: replay the taped scales
  // ---- Rayleigh friction (mem_rayf.f90:7-37, prog_wrtv_rk.f90:7), FP32 as there ----
  bool dorayf = false, dorayfw = false, dorayfdiv = false, dorayfmix = false;
  int krayf_bot = 0, krayfw_bot = 0, krayfdiv_bot = 0, krayfmix_bot = 0;
  std::vector<float> zm, zt, th01d, dn01d, rayf_cof, rayf_cofw, rayf_cofmix, b20;
  std::vector<R> dzit_bot;                  // mem_grid.f90:374-378, for the surface drag
  F2<R> vc03d;
  // ---- SAVEd damping state (vort_div_damp.f90:23-33,255-262) ----
  bool div_first = true, vort_first = true;
  F2<R> div2d[2], vort2d[2], fdivh, fdivl;
  std::vector<R> fvorth, fvortl;
  F2<R> vfact;
  // ---- last tracer step (for the adjoint) ----
  bool have_tape = false;
  F2<R> tape_vmasc, tape_wmasc, tape_rhos[3];
  std::vector<R> tape_ffimp;  // (nrk, mwa)
  double cfl_out[4] = {0, 0, 0, 0};

  R aw_(int k, int i, int j) const { return aw[((j - 1) * 2 + (i - 1)) * mza + (k - 1)]; }
  R awu_(int k, int i, int j) const { return awu[((j - 1) * 2 + (i - 1)) * mza + (k - 1)]; }

  Model(const olamgpu_mesh *m, const olamgpu_params *p) {
    prm = *p;
    mza = m->mza; mwa = m->mwa; mva = m->mva; mma = m->mma;
    npoly.assign(mwa + 1, 0);
    for (int i = 0; i < mwa; ++i) npoly[i + 1] = m->w_npoly[i];
    load2(w_iv, m->w_iv, 7, mwa); load2(w_iw, m->w_iw, 7, mwa); load2(w_im, m->w_im, 7, mwa);
    load2(dirv, m->w_dirv, 7, mwa);
    load2(ecx, m->w_ecvec_vx, 7, mwa); load2(ecy, m->w_ecvec_vy, 7, mwa); load2(ecz, m->w_ecvec_vz, 7, mwa);
    load2(gxc, m->gxps_coef, 7, mwa); load2(gyc, m->gyps_coef, 7, mwa); load2(xx_yy, m->xx_yy, 7, mwa);
    load2(v_iw, m->v_iw, 2, mva); load2(v_im, m->v_im, 2, mva);
    load2(dxps, m->v_dxps, 2, mva); load2(dyps, m->v_dyps, 2, mva);
    if (m->xy_h && m->xy_hu && m->xx0_v && m->xy0_v && m->yy0_v && m->xx0_vu && m->xy0_vu && m->yy0_vu) {
      load2(xyh, m->xy_h, 35, mwa); load2(xyhu, m->xy_hu, 35, mwa);
      load2(xx0v, m->xx0_v, 2, mva); load2(xy0v, m->xy0_v, 2, mva); load2(yy0v, m->yy0_v, 2, mva);
      load2(xx0vu, m->xx0_vu, 2, mva); load2(xy0vu, m->xy0_vu, 2, mva); load2(yy0vu, m->yy0_vu, 2, mva);
    }
    load2(m_iv, m->m_iv, 3, mma); load2(m_iw, m->m_iw, 3, mma); load2(m_im, m->m_im, 3, mma);
    load2(xx_yy_m, m->xx_yy_m, 3, mma);
    load1(wnx, m->wnx, mwa); load1(wny, m->wny, mwa); load1(wnz, m->wnz, mwa); load1(arw0, m->arw0, mwa);
    wnxo2.assign(mwa + 1, 0); wnyo2.assign(mwa + 1, 0); wnzo2.assign(mwa + 1, 0);
    for (int i = 1; i <= mwa; ++i) {  // mem_grid.f90:541-543
      wnxo2[i] = wnx[i] * R(0.5); wnyo2[i] = wny[i] * R(0.5); wnzo2[i] = wnz[i] * R(0.5);
    }
    load1(vnx, m->vnx, mva); load1(vny, m->vny, mva); load1(vnz, m->vnz, mva);
    load1(dnu, m->dnu, mva); load1(dniu, m->dniu, mva); load1(dnv, m->dnv, mva); load1(dniv, m->dniv, mva);
    load1(arm0, m->arm0, mma);
    load1(dzim, m->dzim, mza); load1(dzit, m->dzit, mza); load1(zfacit, m->zfacit, mza);
    load1(zwgt_top8, m->zwgt_top8, mza); load1(zwgt_bot8, m->zwgt_bot8, mza);
    load1(gdz_wgtm8, m->gdz_wgtm8, mza); load1(gdz_wgtp8, m->gdz_wgtp8, mza);
    gdz_wgtm.assign(mza + 1, 0); gdz_wgtp.assign(mza + 1, 0);
    for (int k = 1; k <= mza; ++k) { gdz_wgtm[k] = (R)gdz_wgtm8[k]; gdz_wgtp[k] = (R)gdz_wgtp8[k]; }
    aw.assign(4 * mza, 0);
    for (int i = 0; i < 4 * mza; ++i) aw[i] = (R)m->a_w[i];
    awu.assign(4 * mza, 0);
    for (int i = 0; i < 4 * mza; ++i) awu[i] = (R)m->a_wu[i];
    load2(arw, m->arw, mza, mwa); load2(arv, m->arv, mza, mva);
    load2(volt, m->volt, mza, mwa);
    if (sizeof(R) == 8) {
      // FP64 mode: inverse volumes rebuilt from volt in double (mem_grid.f90:523-556
      // computes them in r8 and rounds to real; the FP64 oracle skips the rounding).
      volti.alloc(mza, mwa); volwi.alloc(mza, mwa); volvi.alloc(mza, mva);
      for (int iw = 2; iw <= mwa; ++iw) {
        for (int k = 2; k <= mza; ++k) volti(k, iw) = (R)(1.0 / volt(k, iw));
        for (int k = 2; k <= mza - 1; ++k) volwi(k, iw) = (R)(1.0 / (volt(k, iw) + volt(k + 1, iw)));
      }
      for (int iv = 2; iv <= mva; ++iv) {
        int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
        for (int k = 2; k <= mza; ++k) volvi(k, iv) = (R)(1.0 / (volt(k, iw1) + volt(k, iw2)));
      }
    } else {
      load2(volti, m->volti, mza, mwa); load2(volwi, m->volwi, mza, mwa); load2(volvi, m->volvi, mza, mva);
    }
    // state
    thil.alloc(mza, mwa); wmc.alloc(mza, mwa); wc.alloc(mza, mwa); vxe.alloc(mza, mwa);
    vye.alloc(mza, mwa); vze.alloc(mza, mwa); vmc.alloc(mza, mva); vc.alloc(mza, mva);
    vmasc.alloc(mza, mva); wmasc.alloc(mza, mwa); alpha_press.alloc(mza, mwa); pwfac.alloc(mza, mwa);
    pvfac.alloc(mza, mva); thilt.alloc(mza, mwa); vmxet.alloc(mza, mwa); vmyet.alloc(mza, mwa);
    vmzet.alloc(mza, mwa); vmt.alloc(mza, mva); rho_old.alloc(mza, mwa);
    rho.alloc(mza, mwa); press.alloc(mza, mwa);
    int ns = std::max(prm.num_scalar, 0);
    scp.resize(ns); sct.resize(ns); lam.resize(ns);
    for (int n = 0; n < ns; ++n) { scp[n].alloc(mza, mwa); sct[n].alloc(mza, mwa); lam[n].alloc(mza, mwa); }
    pdef.assign(ns, 1);
    sgsmix.assign(ns, 1);
    tape_sv.resize(ns); tape_sw.resize(ns); tape_lim.assign(ns, 0);
    if (m->dzit_bot) load1(dzit_bot, m->dzit_bot, mza);
    if (prm.nudflag > 0) { wmanud.alloc(mza, mwa); vmanud.alloc(mza, mva); rhot_nud.alloc(mza, mwa); }
    if (m->zm && m->zt) {   // oracle_create checks that Rayleigh friction has them
      load1(zm, m->zm, mza);
      load1(zt, m->zt, mza);
      if (m->th01d) load1(th01d, m->th01d, mza);
      if (m->dn01d) load1(dn01d, m->dn01d, mza);
      rayf_init();
    } else {
      b20.assign(mza + 1, 1.f);
    }
  }

  // rayf_init, mem_rayf.f90:41-156: friction profiles from the namelist values, and
  // init_wrtv_rk's b20 = 1 + dtsm * rayf_cofw (prog_wrtv_rk.f90:46-52)
  void rayf_init() {
    const float dti = 1.0f / (float)prm.dtsm;
    rayf_cof.assign(mza + 1, 0.f); rayf_cofw.assign(mza + 1, 0.f); rayf_cofmix.assign(mza + 1, 0.f);
    b20.assign(mza + 1, 1.f);
    krayf_bot = krayfw_bot = krayfdiv_bot = krayfmix_bot = mza + 1;
    if (prm.rayf_fact > 1.e-7f && prm.initial == 1) {
      for (int k = 2; k <= mza; ++k)
        if (zt[k] > prm.rayf_zmin) { krayf_bot = k; break; }
      if (krayf_bot <= mza) {
        dorayf = true;
        float distimi = prm.rayf_fact * dti;
        for (int k = krayf_bot; k <= mza; ++k)
          rayf_cof[k] = distimi * pow_r((zt[k] - prm.rayf_zmin) / (zm[mza] - prm.rayf_zmin), prm.rayf_expon);
      }
    }
    if (prm.rayfw_fact > 1.e-7f) {
      for (int k = 2; k <= mza - 1; ++k)
        if (zm[k] > prm.rayfw_zmin) { krayfw_bot = k; break; }
      if (krayfw_bot < mza) {
        dorayfw = true;
        float distimi = prm.rayfw_fact * dti;
        for (int k = krayfw_bot; k <= mza - 1; ++k)
          rayf_cofw[k] = distimi * pow_r((zm[k] - prm.rayfw_zmin) / (zm[mza] - prm.rayfw_zmin), prm.rayfw_expon);
      }
    }
    if (prm.rayfdiv_fact > 1.e-7f) {
      for (int k = 2; k <= mza; ++k)
        if (zt[k] > prm.rayfdiv_zmin) { krayfdiv_bot = k; break; }
      if (krayfdiv_bot <= mza) dorayfdiv = true;
    }
    if (prm.rayfmix_fact > 1.e-7f) {
      for (int k = 2; k <= mza - 1; ++k)
        if (zm[k] > prm.rayfmix_zmin) { krayfmix_bot = k; break; }
      if (krayfmix_bot < mza) {
        dorayfmix = true;
        for (int k = krayfmix_bot; k <= mza - 1; ++k)   // normalised by zm(mza) - rayfw_zmin
          rayf_cofmix[k] = 0.5f * prm.rayfmix_fact *
                           pow_r((zm[k] - prm.rayfmix_zmin) / (zm[mza] - prm.rayfw_zmin), prm.rayfmix_expon);
      }
    }
    if (dorayfw)
      for (int k = krayfw_bot; k <= mza - 1; ++k) b20[k] = 1.0f + (float)prm.dtsm * rayf_cofw[k];
  }

  // ---- nudging fields (oisan/mem_nudge.f90:46,90-91), zero until set ----
  F2<R> wmanud, vmanud;
  F2<RP> rhot_nud;
  int set_nudging(const float *wm, const float *vm, const double *rt) override {
    if (!wmanud.n) { wmanud.alloc(mza, mwa); vmanud.alloc(mza, mva); rhot_nud.alloc(mza, mwa); }
    if (wm) load2(wmanud, wm, mza, mwa);
    if (vm) load2(vmanud, vm, mza, mva);
    if (rt) load2(rhot_nud, rt, mza, mwa);
    return 0;
  }

  // ---- turbulence of the host's PBL scheme (mem_turb) and comp_horiz_k's results ----
  bool turb_on = false;
  F2<R> vkh, vkm, akhodx, akmodx;
  std::vector<int> khtop, kmtop, khtopv, kmtopv;
  std::vector<R> akm_sfc;

  int set_turbulence(const float *vkh_h, const float *vkm_h, const int32_t *kh, const int32_t *km,
                     const float *as) override {
    if (dzit_bot.empty() || !vkh_h || !vkm_h || !kh || !km || !as) return OLAMGPU_EINVAL;
    khtop.assign(mwa + 1, 0); kmtop.assign(mwa + 1, 0); akm_sfc.assign(mwa + 1, R(0));
    for (int iw = 2; iw <= mwa; ++iw) {
      khtop[iw] = kh[iw - 1]; kmtop[iw] = km[iw - 1]; akm_sfc[iw] = (R)as[iw - 1];
      if (khtop[iw] < 0 || khtop[iw] > mza - 1 || kmtop[iw] < 0 || kmtop[iw] > mza - 1) return OLAMGPU_EINVAL;
    }
    load2(vkh, vkh_h, mza, mwa); load2(vkm, vkm_h, mza, mwa);
    comp_horiz_k();
    turb_on = true;
    return 0;
  }

  // comp_horiz_k_column, pbl_driver.f90:154-238 (lpv = 2, akmin(mrl) = akmin)
  void comp_horiz_k() {
    akhodx.alloc(mza, mva); akmodx.alloc(mza, mva);
    khtopv.assign(mva + 1, 0); kmtopv.assign(mva + 1, 0);
    const R dtl = (R)prm.dtlm, akmin = (R)prm.akmin;
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv) {
      int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
      R fact1 = R(0.95) / (dtl * (R)npoly[iw1]);
      R fact2 = R(0.95) / (dtl * (R)npoly[iw2]);
      R hk0[130], akstab[130];
      if (akmin > R(1.e-7)) {
        R bkmin = akmin * R(0.075) * pow_r(std::min(arw0[iw1], arw0[iw2]), (R)0.66666666f);
        khtopv[iv] = mza;
        kmtopv[iv] = mza;
        for (int k = 2; k <= mza; ++k) hk0[k] = R(0.5) * (R)(rho(k, iw1) + rho(k, iw2)) * bkmin;
      } else {
        khtopv[iv] = std::max(khtop[iw1], khtop[iw2]);
        khtopv[iv] = (khtopv[iv] >= 2) ? khtopv[iv] + 1 : 1;
        kmtopv[iv] = std::max(kmtop[iw1], kmtop[iw2]);
        kmtopv[iv] = (kmtopv[iv] >= 2) ? kmtopv[iv] + 1 : 1;
      }
      for (int k = 2; k <= std::max(kmtopv[iv], khtopv[iv]); ++k) {
        R stab1 = (R)(rho(k, iw1) * (RP)(R)volt(k, iw1) * (RP)fact1);
        R stab2 = (R)(rho(k, iw2) * (RP)(R)volt(k, iw2) * (RP)fact2);
        akstab[k] = std::min(stab1, stab2);
      }
      for (int k = 2; k <= kmtopv[iv]; ++k) {
        R hkm = R(0.25) * (vkm(k, iw1) + vkm(k, iw2) + vkm(k - 1, iw1) + vkm(k - 1, iw2));
        if (akmin > R(1.e-7)) hkm = std::max(hkm, hk0[k]);
        akmodx(k, iv) = std::min(dniv[iv] * arv(k, iv) * hkm, akstab[k]);
      }
      for (int k = 2; k <= khtopv[iv]; ++k) {
        R hkh = R(0.25) * (vkh(k, iw1) + vkh(k, iw2) + vkh(k - 1, iw1) + vkh(k - 1, iw2));
        if (akmin > R(1.e-7)) hkh = std::max(hkh, hk0[k]);
        akhodx(k, iv) = std::min(dniv[iv] * arv(k, iv) * hkh, akstab[k]);
      }
    }
  }

  // solve_eddy_diff_heat, pbl_driver.f90:519-605 (idiffk /= 1: no non-local term)
  void solve_eddy_diff_heat(int iw, R *thiltc) const {
    if (khtop[iw] < 2) return;
    const int ka = 2, kz = khtop[iw], kmax = kz + 1;
    const R dts = (R)prm.dtsm, forw_imp = R(1.1);
    R dtomass[130], varp[130];
    RT v5[130], v6[130], v7[130], rhs[130], soln[130];
    for (int k = ka; k <= kmax; ++k) {
      R dtom = dts * volti(k, iw) / (R)rho(k, iw);
      dtomass[k] = forw_imp * dtom;
      varp[k] = thil(k, iw) + dtom * thiltc[k];
    }
    for (int k = ka; k <= kz; ++k) {
      R akodz = arw(k, iw) * vkh(k, iw) * dzim[k];
      rhs[k] = (RT)(akodz * (varp[k] - varp[k + 1]));
      R a5 = -akodz * dtomass[k], a7 = -akodz * dtomass[k + 1];
      v5[k] = (RT)a5; v7[k] = (RT)a7; v6[k] = (RT)(R(1) - a5 - a7);
    }
    tridiffo(ka, kz, v5, v6, v7, rhs, soln);
    double flux[130];
    flux[ka - 1] = 0.0;
    flux[kmax] = 0.0;
    for (int k = ka; k <= kmax - 1; ++k) flux[k] = (double)(R)soln[k];
    for (int k = ka; k <= kmax; ++k) thiltc[k] = thiltc[k] + (R)(flux[k - 1] - flux[k]);
  }

  // solve_eddy_diff_vxe, pbl_driver.f90:609-742 (lsw = 1 on flat terrain); tridv is
  // tridiffo's elimination applied to each right-hand side
  void solve_eddy_diff_vxe(int iw, R *vmx, R *vmy, R *vmz) const {
    const int ka = 2, kmax = std::max(kmtop[iw] + 1, ka);
    const R dts = (R)prm.dtsm, forw_imp = R(1.1);
    R akodz[130], dtomass[130], a5[130], a6[130], a7[130];
    for (int k = 0; k <= kmax; ++k) akodz[k] = R(0);
    for (int k = ka; k <= kmtop[iw]; ++k) akodz[k] = arw(k, iw) * vkm(k, iw) * dzim[k];
    for (int k = ka; k <= kmax; ++k) {
      R dtom = (R)((RP)(dts * volti(k, iw)) / rho(k, iw));
      dtomass[k] = forw_imp * dtom;
      a5[k] = -dtomass[k] * akodz[k - 1];
      a7[k] = -dtomass[k] * akodz[k];
      a6[k] = R(1) - a5[k] - a7[k];
    }
    const R vctr3 = akm_sfc[iw] * dzit_bot[ka];
    a6[ka] = a6[ka] + dtomass[ka] * vctr3;
    RT v5[130], v6[130], v7[130], rhs[130], soln[130];
    for (int k = ka; k <= kmax; ++k) { v5[k] = (RT)a5[k]; v6[k] = (RT)a6[k]; v7[k] = (RT)a7[k]; }
    const F2<R> *vel[3] = {&vxe, &vye, &vze};
    R *ten[3] = {vmx, vmy, vmz};
    for (int q = 0; q < 3; ++q) {
      for (int k = ka; k <= kmax; ++k) rhs[k] = (RT)(*vel[q])(k, iw);
      tridiffo(ka, kmax, v5, v6, v7, rhs, soln);
      R so[130], v2[130];
      for (int k = ka; k <= kmax; ++k) so[k] = (R)soln[k];
      v2[ka - 1] = R(0);
      v2[kmax] = R(0);
      for (int k = ka; k <= kmax - 1; ++k) v2[k] = akodz[k] * (so[k] - so[k + 1]);
      ten[q][ka] = ten[q][ka] + (v2[ka - 1] - v2[ka] - vctr3 * so[ka]);
      for (int k = ka + 1; k <= kmax; ++k) ten[q][k] = ten[q][k] + v2[k - 1] - v2[k];
    }
  }

  // comp_alpha_press, timestep.f90:500-559 (called when miclevel > 0).  rr_v, rr_w and
  // theta are host (mza, mwa) FP32 arrays; zwgt_top/zwgt_bot are the FP32 roundings of
  // zwgt_top8/zwgt_bot8 (mem_grid.f90:432-433).
  void comp_alpha_press(const float *rr_v_h, const float *rr_w_h, const float *theta_h) override {
    F2<R> rr_v, rr_w, theta;
    load2(rr_v, rr_v_h, mza, mwa); load2(rr_w, rr_w_h, mza, mwa); load2(theta, theta_h, mza, mwa);
    const R pc1 = sizeof(R) == 4 ? (R)pow_r(1.0f / 1.0e5f, 287.0f / 717.0f) : (R)std::pow(1.0 / 1.0e5, 287.0 / 717.0);
    const R rdry = R(287), rvap = R(461), cpocv = (R)(1004.0 / 717.0);
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw) {
      for (int k = 2; k <= mza; ++k)
        alpha_press(k, iw) = pc1 * pow_r((rdry + rvap * rr_v(k, iw)) * theta(k, iw) / thil(k, iw), cpocv);
      for (int k = 2; k <= mza - 1; ++k) {
        R rw = (R)zwgt_top8[k] * rr_w(k + 1, iw) + (R)zwgt_bot8[k] * rr_w(k, iw);
        pwfac(k, iw) = dzim[k] * (R(1) - rw + rw * rw);
      }
    }
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv) {
      int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
      for (int k = 2; k <= mza; ++k) {
        R rw = R(0.5) * (rr_w(k, iw1) + rr_w(k, iw2));
        pvfac(k, iv) = dniv[iv] * zfacit[k] * (R(1) - rw + rw * rw);
      }
    }
  }

  // rayf_init's copy of the initial wind (mem_rayf.f90:158-171), taken where
  // olam_run.f90:555 calls it: after the initial state exists
  void init_wrtv_rk() override {
    if (!dorayf) return;
    vc03d.alloc(mza, mva);
    for (int iv = 2; iv <= mva; ++iv)
      for (int k = 2; k <= mza; ++k) vc03d(k, iv) = vc(k, iv);
  }

  // rayf_mix_top_vxe, mem_rayf.f90:177-220 (vmxet.. are the volume-weighted short-step
  // tendencies of column iw)
  void rayf_mix_top_vxe(int iw, R *vmxet_c, R *vmyet_c, R *vmzet_c) const {
    R vxflux[130], vyflux[130], vzflux[130];
    int ka = std::max(krayfmix_bot, 2);
    vxflux[mza] = vyflux[mza] = vzflux[mza] = 0;
    vxflux[ka - 1] = vyflux[ka - 1] = vzflux[ka - 1] = 0;
    for (int k = ka; k <= mza - 1; ++k) {
      R fact = arw0[iw] * (R)rayf_cofmix[k] * (R)(rho(k + 1, iw) + rho(k, iw));
      vxflux[k] = fact * (vxe(k, iw) - vxe(k + 1, iw));
      vyflux[k] = fact * (vye(k, iw) - vye(k + 1, iw));
      vzflux[k] = fact * (vze(k, iw) - vze(k + 1, iw));
    }
    for (int k = ka; k <= mza; ++k) {
      vmxet_c[k] = vmxet_c[k] + vxflux[k - 1] - vxflux[k];
      vmyet_c[k] = vmyet_c[k] + vyflux[k - 1] - vyflux[k];
      vmzet_c[k] = vmzet_c[k] + vzflux[k - 1] - vzflux[k];
    }
  }
  void set_frozen(int v) override { frozen_lim = v != 0; }
  // test hook: sum of rho * thil * volt over the prognostic cells at full precision
  double rhothil_sum() override {
    double t = 0;
    for (int iw = 2; iw <= mwa; ++iw)
      for (int k = 2; k <= mza; ++k) t += (double)rho(k, iw) * (double)thil(k, iw) * volt(k, iw);
    return t;
  }
  // test hook: the Rayleigh friction profiles per level k = 1..mza (rayf_cof, rayf_cofw,
  // rayf_cofmix, b20) and krayf_bot, krayfw_bot, krayfdiv_bot, krayfmix_bot
  void rayf_profiles(float *out, int32_t *kb) override {
    const std::vector<float> *a[4] = {&rayf_cof, &rayf_cofw, &rayf_cofmix, &b20};
    for (int i = 0; i < 4; ++i)
      for (int k = 1; k <= mza; ++k) out[(size_t)i * mza + k - 1] = (*a[i]).empty() ? (i == 3 ? 1.f : 0.f) : (*a[i])[k];
    kb[0] = krayf_bot; kb[1] = krayfw_bot; kb[2] = krayfdiv_bot; kb[3] = krayfmix_bot;
  }
  // test hook: run the setup part of prog_wrtv_rk (which advances the SAVEd divergence
  // of the damping) and return thilt_short, vmxet_short, vmyet_short, vmzet_short
  // (mza, mwa each) and vmt_short (mza, mva) in Fortran layout
  int setup_tendencies(double *out) override {
    if (dorayf && !vc03d.n) return OLAMGPU_EINVAL;
    setup_short();
    size_t o = 0;
    for (const F2<R> *a : {&thilt_s, &vmxet_s, &vmyet_s, &vmzet_s, &vmt_s})
      for (size_t i = 0; i < a->v.size(); ++i) out[o++] = (double)a->v[i];
    return 0;
  }
  // test hook: reconstruction coefficients of a field (5, mwa, mza) and its face values
  // on both sides of every face (2, mva, mza), for one order and coefficient/offset set
  int recon(const float *field, int iorder, int bounded_coef, int bounded_off, double *coefs,
            double *faces) override {
    if (iorder < 1 || iorder > 3 || (iorder == 3 && !xyh.n)) { err = "bad order"; return OLAMGPU_EINVAL; }
    F2<R> s, g[5];
    s.alloc(mza, mwa);
    for (auto &x : g) x.alloc(mza, mwa);
    for (size_t i = 0; i < s.v.size(); ++i) s.v[i] = (R)field[i];
    recon_coefs(s, g, iorder, bounded_coef != 0);
    for (int t = 0; t < 5; ++t)
      for (size_t i = 0; i < g[t].v.size(); ++i) coefs[(size_t)t * g[t].v.size() + i] = (double)g[t].v[i];
    const size_t fv = (size_t)mza * mva;
    for (int side = 1; side <= 2; ++side)
      for (int iv = 2; iv <= mva; ++iv)
        for (int k = 2; k <= mza; ++k)
          faces[(size_t)(side - 1) * fv + (size_t)(iv - 1) * mza + (k - 1)] =
              (double)face_value(s, g, iorder, side, iv, k, bounded_off != 0);
    return 0;
  }
  // test hook: the taped limiter scales of tracer n (which 0: faces (mza, mva), 1: W levels (mza, mwa))
  int tape_scales(int n, int which, double *out) override {
    if (n < 1 || n > (int)tape_sv.size() || !tape_lim[n - 1]) { err = "no taped scales"; return OLAMGPU_ESTATE; }
    const F2<R> &t = which ? tape_sw[n - 1] : tape_sv[n - 1];
    for (size_t i = 0; i < t.v.size(); ++i) out[i] = (double)t.v[i];
    return 0;
  }
  int set_pdef(int n, int v) override {
    if (n < 1 || n > (int)pdef.size()) { err = "bad scalar index"; return OLAMGPU_EINVAL; }
    pdef[n - 1] = v != 0;
    return 0;
  }
  int set_sgsmix(int n, int v) override {
    if (n < 1 || n > (int)sgsmix.size()) { err = "bad scalar index"; return OLAMGPU_EINVAL; }
    sgsmix[n - 1] = v != 0;
    return 0;
  }

  // ---------------------------------------------------------------- transfers
  template <class T, class S>
  static void cp_in(F2<T> &a, const S *src) { if (src) for (size_t i = 0; i < a.v.size(); ++i) a.v[i] = (T)src[i]; }
  template <class T, class S>
  static void cp_out(const F2<T> &a, S *dst) { if (dst) for (size_t i = 0; i < a.v.size(); ++i) dst[i] = (S)a.v[i]; }

  int upload_state(const olamgpu_state *s) override {
    cp_in(thil, s->thil); cp_in(wmc, s->wmc); cp_in(wc, s->wc); cp_in(vxe, s->vxe); cp_in(vye, s->vye);
    cp_in(vze, s->vze); cp_in(rho, s->rho); cp_in(press, s->press); cp_in(vmc, s->vmc); cp_in(vc, s->vc);
    cp_in(wmasc, s->wmasc); cp_in(vmasc, s->vmasc); cp_in(alpha_press, s->alpha_press);
    cp_in(pwfac, s->pwfac); cp_in(pvfac, s->pvfac); cp_in(thilt, s->thilt); cp_in(vmxet, s->vmxet);
    cp_in(vmyet, s->vmyet); cp_in(vmzet, s->vmzet); cp_in(vmt, s->vmt);
    return 0;
  }
  int download_state(olamgpu_state *s) override {
    cp_out(thil, s->thil); cp_out(wmc, s->wmc); cp_out(wc, s->wc); cp_out(vxe, s->vxe); cp_out(vye, s->vye);
    cp_out(vze, s->vze); cp_out(rho, s->rho); cp_out(press, s->press); cp_out(vmc, s->vmc); cp_out(vc, s->vc);
    cp_out(wmasc, s->wmasc); cp_out(vmasc, s->vmasc); cp_out(alpha_press, s->alpha_press);
    cp_out(pwfac, s->pwfac); cp_out(pvfac, s->pvfac); cp_out(thilt, s->thilt); cp_out(vmxet, s->vmxet);
    cp_out(vmyet, s->vmyet); cp_out(vmzet, s->vmzet); cp_out(vmt, s->vmt);
    return 0;
  }
  int upload_scalar(int n, const float *p, const float *t) override {
    if (n < 1 || n > (int)scp.size()) { err = "bad scalar index"; return OLAMGPU_EINVAL; }
    cp_in(scp[n - 1], p);
    if (t) cp_in(sct[n - 1], t); else std::fill(sct[n - 1].v.begin(), sct[n - 1].v.end(), R(0));
    return 0;
  }
  int download_scalar(int n, float *p) override {
    if (n < 1 || n > (int)scp.size()) { err = "bad scalar index"; return OLAMGPU_EINVAL; }
    cp_out(scp[n - 1], p); return 0;
  }
  int upload_adj(int n, const float *p) override {
    if (n < 1 || n > (int)lam.size()) { err = "bad scalar index"; return OLAMGPU_EINVAL; }
    cp_in(lam[n - 1], p); return 0;
  }
  int download_adj(int n, float *p) override {
    if (n < 1 || n > (int)lam.size()) { err = "bad scalar index"; return OLAMGPU_EINVAL; }
    cp_out(lam[n - 1], p); return 0;
  }

  // ---------------------------------------------------------------- tend0
  // timestep.f90:430-496 (tendencies of thil/momentum are physics inputs here and
  // are left as uploaded; vmasc/wmasc/var_t are zeroed, rho_old saved).
  void tend0() override {
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw)
      for (int k = 2; k <= mza; ++k) {
        rho_old(k, iw) = (R)rho(k, iw);
        wmasc(k, iw) = 0;
        for (auto &t : sct) t(k, iw) = 0;
      }
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv)
      for (int k = 2; k <= mza; ++k) vmasc(k, iv) = 0;
  }

  // timestep.f90:563-612
  void timeavg_momsc() override {
    if (prm.nacoust <= 1) return;
    R f = (R)prm.nacoust;
    for (int iv = 2; iv <= mva; ++iv) for (int k = 2; k <= mza; ++k) vmasc(k, iv) = vmasc(k, iv) / f;
    for (int iw = 2; iw <= mwa; ++iw) for (int k = 2; k <= mza - 1; ++k) wmasc(k, iw) = wmasc(k, iw) / f;
  }

  // ---------------------------------------------------------------- grad_lib
  // grad_t2d, grad_lib.f90:236-269
  void grad_t2d(int iw, const F2<R> &s, R *gx, R *gy) const {
    for (int k = 1; k <= mza; ++k) { gx[k] = 0; gy[k] = 0; }
    for (int n = 1; n <= npoly[iw]; ++n) {
      int iwn = w_iw(n, iw);
      for (int k = 2; k <= mza; ++k) {
        R d = s(k, iwn) - s(k, iw);
        gx[k] = gx[k] + gxc(n, iw) * d;
        gy[k] = gy[k] + gyc(n, iw) * d;
      }
    }
  }
  // grad_t2d_quadratic, grad_lib.f90:278-341: the five least-squares terms (x, y, xx,
  // xy, yy) from the neighbour differences, bounded (xy_h) or unbounded (xy_hu)
  void grad_t2d_quadratic(int iw, const F2<R> &s, F2<R> *g, bool bounded) const {
    const F2<R> &c = bounded ? xyh : xyhu;
    for (int t = 0; t < 5; ++t) for (int k = 1; k <= mza; ++k) g[t](k, iw) = 0;
    for (int n = 1; n <= npoly[iw]; ++n) {
      int iwn = w_iw(n, iw);
      for (int k = 2; k <= mza; ++k) {
        R sc = s(k, iwn) - s(k, iw);
        for (int t = 0; t < 5; ++t) g[t](k, iw) = g[t](k, iw) + sc * c(t + 1 + 5 * (n - 1), iw);
      }
    }
  }
  // reconstruction coefficients at every W: order 2 grad_t2d (g[0], g[1]), order 3
  // grad_t2d_quadratic (g[0..4])
  void recon_coefs(const F2<R> &s, F2<R> *g, int iorder, bool bounded) const {
    if (iorder == 2) {
#pragma omp parallel for
      for (int iw = 2; iw <= mwa; ++iw) grad_t2d(iw, s, g[0].col(iw) - 1, g[1].col(iw) - 1);
    } else if (iorder == 3) {
#pragma omp parallel for
      for (int iw = 2; iw <= mwa; ++iw) grad_t2d_quadratic(iw, s, g, bounded);
    }
  }
  // the reconstruction of cell v_iw(side, iv) at face iv (scalar_v_column forms,
  // e.g. scalar_transport_rk.F90:490-525): order 3 adds the quadratic terms with the
  // bounded (xx0_v) or unbounded (xx0_vu) face offsets
  R face_value(const F2<R> &s, const F2<R> *g, int iorder, int side, int iv, int k, bool bounded) const {
    const int iw = v_iw(side, iv);
    if (iorder == 1) return s(k, iw);
    const int os = plant == 4 ? 3 - side : side;   // plant 4: the other side's offsets
    R v = s(k, iw) + dxps(os, iv) * g[0](k, iw) + dyps(os, iv) * g[1](k, iw);
    if (iorder == 3) {
      const F2<R> &ox = bounded ? xx0v : xx0vu, &oxy = bounded ? xy0v : xy0vu, &oy = bounded ? yy0v : yy0vu;
      v = v + ox(side, iv) * g[2](k, iw) + oxy(side, iv) * g[3](k, iw) + oy(side, iv) * g[4](k, iw);
    }
    return v;
  }
  // transposes: h, the adjoint of a face value, into the adjoints of the coefficients
  void face_value_adj(F2<R> *lg, int iorder, int side, int iv, int k, R h, bool bounded) const {
    const int iw = v_iw(side, iv);
    if (iorder == 1) return;
    lg[0](k, iw) += dxps(side, iv) * h;
    lg[1](k, iw) += dyps(side, iv) * h;
    if (iorder == 3) {
      const F2<R> &ox = bounded ? xx0v : xx0vu, &oxy = bounded ? xy0v : xy0vu, &oy = bounded ? yy0v : yy0vu;
      lg[2](k, iw) += ox(side, iv) * h;
      lg[3](k, iw) += oxy(side, iv) * h;
      lg[4](k, iw) += oy(side, iv) * h;
    }
  }
  // coefficient t of neighbour slot n of cell iw
  R recon_c(int iorder, bool bounded, int t, int n, int iw) const {
    if (iorder == 2) return t == 0 ? gxc(n, iw) : gyc(n, iw);
    return (bounded ? xyh : xyhu)(t + 1 + 5 * (n - 1), iw);
  }
  // lq += transpose of recon_coefs applied to lg: g_t(c) = sum_n c_t(n, c) (s(n) - s(c))
  void recon_coefs_adj(const F2<R> *lg, F2<R> &lq, int iorder, bool bounded) const {
    if (iorder == 1) return;
    const int nt = iorder == 2 ? 2 : 5;
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw) {
      for (int n = 1; n <= npoly[iw]; ++n) {
        int iwn = w_iw(n, iw);
        int back = 0;
        for (int m = 1; m <= npoly[iwn]; ++m) if (w_iw(m, iwn) == iw) back = m;
        for (int k = 2; k <= mza; ++k) {
          R in = R(0), out = R(0);
          for (int t = 0; t < nt; ++t) {
            in = in + recon_c(iorder, bounded, t, back, iwn) * lg[t](k, iwn);
            out = out + recon_c(iorder, bounded, t, n, iw) * lg[t](k, iw);
          }
          lq(k, iw) += in;
          lq(k, iw) -= out;
        }
      }
    }
  }
  // grad_z_quadratic bounded, itopbc = ibotbc = 0, grad_lib.f90:166-229
  void grad_z_quadratic(const R *s, R *sb, R *st) const {
    std::vector<R> ds(mza + 2, R(0));
    for (int k = 2; k <= mza - 1; ++k) ds[k] = s[k + 1] - s[k];
    ds[1] = 0; ds[mza] = 0;
    for (int k = 2; k <= mza; ++k) {
      sb[k] = s[k] + ds[k] * aw_(k, 1, 1) + ds[k - 1] * aw_(k, 2, 1);
      st[k] = s[k] + ds[k] * aw_(k, 1, 2) + ds[k - 1] * aw_(k, 2, 2);
    }
  }
  // laplacian2d, grad_lib.f90:494-524
  void laplacian2d(int iw, const F2<R> &s, R *d) const {
    for (int k = 1; k <= mza; ++k) d[k] = 0;
    for (int n = 1; n <= npoly[iw]; ++n) {
      int iwn = w_iw(n, iw);
      for (int k = 2; k <= mza; ++k) d[k] = d[k] + (s(k, iwn) - s(k, iw)) * xx_yy(n, iw);
    }
  }
  // laplacian_m, grad_lib.f90:531-559
  void laplacian_m(int im, const F2<R> &s, R *d) const {
    for (int k = 1; k <= mza; ++k) d[k] = 0;
    for (int j = 1; j <= 3; ++j) {
      int imn = m_im(j, im);
      for (int k = 2; k <= mza; ++k) d[k] = d[k] + (s(k, imn) - s(k, im)) * xx_yy_m(j, im);
    }
  }

  // tridiffo, tridiag_lib.f90:10-36 (RT precision)
  void tridiffo(int ka, int kz, const RT *cim1, const RT *ci, const RT *cip1, const RT *rhs, RT *soln) const {
    tridiffo_t<RT>(ka, kz, cim1, ci, cip1, rhs, soln);
  }

  // ---------------------------------------------------------------- fluxes_wrtv
  // prog_wrtv_rk.f90:950-1129 (iorderv = 3).  Order 3 fits the unbounded quadratic
  // (grad_t2d_quadratic bounded=.false., :983-984) and evaluates it at the faces with
  // the bounded offsets xx0_v (:1057-1071), as the reference does.
  void fluxes_wrtv(const F2<R> &s, F2<R> &upw, F2<R> &upv, int iorderh, F2<R> *g) {
    recon_coefs(s, g, iorderh, false);
#pragma omp parallel
    {
      std::vector<R> sb(mza + 2), st(mza + 2);
#pragma omp for
      for (int iw = 2; iw <= mwa; ++iw) {  // scalar_w_column :1085-1126
        int kb = 2;
        upw(kb - 1, iw) = s(kb, iw);
        upw(mza, iw) = s(mza, iw);
        grad_z_quadratic(&s(1, iw) - 1, sb.data(), st.data());
        for (int k = kb; k <= mza - 1; ++k) {
          upw(k, iw) = st[k];
          if (wmc(k, iw) < R(0)) upw(k, iw) = sb[k + 1];
        }
      }
    }
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv) {  // scalar_v_column :1022-1081
      int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
      (void)iw1; (void)iw2;
      for (int k = 2; k <= mza; ++k) {
        const R s1 = face_value(s, g, iorderh, 1, iv, k, true);
        const R s2 = face_value(s, g, iorderh, 2, iv, k, true);
        upv(k, iv) = s1;
        if (vmc(k, iv) < R(0)) upv(k, iv) = s2;
      }
    }
  }

  // fluxes_wrtv_monot, prog_wrtv_rk.f90:1133-1554 (iorderv = 3, centered_monot =
  // .false., :26-27): low-order upwind face values from the RK-start field advanced by
  // its long-step tendency (scpt, :1174-1181), high-order corrections from the stage
  // field (bounded horizontal fit with the unbounded face offsets xx0_vu, :1186-1196 and
  // :1341-1354; unbounded constant-gradient vertical quadratic, :1418-1420), the
  // in/out renormalisation scales of scalar_monot_limiter (:1452-1550) and the
  // renormalised face values (:1235-1266).  vmca is the stage's vmc * arv, rho the
  // density before prog_wrt_begs of this stage.
  void fluxes_wrtv_monot(const F2<R> &s, const F2<R> &s0, const F2<R> &st, F2<R> &upw, F2<R> &upv, R dtr,
                         int iorderh, F2<R> *g) {
    const R eps0 = R(1.e-32), eps1 = R(3.e-07), onep = R(1) + eps1;
    F2<R> &scpt = m_scpt, &hiv = m_hiv, &sfv = m_sfv, &hiw = m_hiw, &sfw = m_sfw, &s1 = m_s1, &s2 = m_s2;
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw)
      for (int k = 2; k <= mza; ++k) scpt(k, iw) = s0(k, iw) + dtr * volti(k, iw) * st(k, iw) / (R)rho0(k, iw);
    recon_coefs(s, g, iorderh, true);
    // scalar_wflux_monot (:1380-1445)
#pragma omp parallel
    {
      R sb[130], stp[130];
#pragma omp for
      for (int iw = 2; iw <= mwa; ++iw) {
        const int kb = 2;
        upw(kb - 1, iw) = s(kb, iw);
        upw(mza, iw) = s(mza, iw);
        for (int k = kb; k <= mza - 1; ++k) {
          upw(k, iw) = scpt(k, iw);
          if (wmc(k, iw) < R(0)) upw(k, iw) = scpt(k + 1, iw);
        }
        grad_z_q(&s(1, iw) - 1, sb, stp, 1, 1, false);
        sfw(kb - 1, iw) = 0;
        sfw(mza, iw) = 0;
        for (int k = kb; k <= mza - 1; ++k) {
          R h = stp[k];
          if (wmc(k, iw) < R(0)) h = sb[k + 1];
          hiw(k, iw) = h - upw(k, iw);
          sfw(k, iw) = arw(k, iw) * wmc(k, iw) * hiw(k, iw);
        }
      }
    }
    // scalar_vflux_monot (:1295-1377)
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv) {
      const int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
      for (int k = 2; k <= mza; ++k) {
        upv(k, iv) = scpt(k, iw1);
        if (vmca(k, iv) < R(0)) upv(k, iv) = scpt(k, iw2);
        const R h1 = face_value(s, g, iorderh, 1, iv, k, false);
        const R h2 = face_value(s, g, iorderh, 2, iv, k, false);
        R h = h1;
        if (vmca(k, iv) < R(0)) h = h2;
        hiv(k, iv) = h - upv(k, iv);
        sfv(k, iv) = vmca(k, iv) * hiv(k, iv);
      }
    }
    // scalar_monot_limiter (:1450-1550)
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw) {
      const int kb = 2;
      const R dtrp = onep * dtr;
      R wmca[130], smax[130], smin[130], rhos[130], fl[130], fin[130], fout[130];
      wmca[kb - 1] = 0;
      for (int k = kb; k <= mza - 1; ++k) wmca[k] = wmc(k, iw) * arw(k, iw);
      wmca[mza] = 0;
      for (int k = kb; k <= mza; ++k) { smax[k] = scpt(k, iw); smin[k] = scpt(k, iw); rhos[k] = (R)rho(k, iw); }
      for (int k = kb + 1; k <= mza; ++k)
        if (wmca[k - 1] > R(0)) { smax[k] = std::max(smax[k], scpt(k - 1, iw)); smin[k] = std::min(smin[k], scpt(k - 1, iw)); }
      for (int k = kb; k <= mza - 1; ++k)
        if (wmca[k] < R(0)) { smax[k] = std::max(smax[k], scpt(k + 1, iw)); smin[k] = std::min(smin[k], scpt(k + 1, iw)); }
      for (int k = 1; k <= mza; ++k) { fl[k] = 0; fin[k] = 0; fout[k] = 0; }
      for (int jv = 1; jv <= npoly[iw]; ++jv) {
        const int iv = w_iv(jv, iw), iwn = w_iw(jv, iw);
        const R d = dirv(jv, iw);
        for (int k = 2; k <= mza; ++k) {
          fl[k] = fl[k] + d * vmca(k, iv) * (upv(k, iv) - scpt(k, iw));
          fin[k] = fin[k] + std::max(d * sfv(k, iv), R(0));
          fout[k] = fout[k] + std::min(d * sfv(k, iv), R(0));
          if (d * vmca(k, iv) > R(0)) { smax[k] = std::max(smax[k], scpt(k, iwn)); smin[k] = std::min(smin[k], scpt(k, iwn)); }
        }
      }
      for (int k = kb; k <= mza; ++k) {
        fl[k] = fl[k] + wmca[k - 1] * (upw(k - 1, iw) - scpt(k, iw)) - wmca[k] * (upw(k, iw) - scpt(k, iw));
        fin[k] = fin[k] + std::max(sfw(k - 1, iw), R(0)) - std::min(sfw(k, iw), R(0));
        fout[k] = fout[k] + std::min(sfw(k - 1, iw), R(0)) - std::max(sfw(k, iw), R(0));
        const R rlow = scpt(k, iw) * rhos[k] + dtr * volti(k, iw) * fl[k];
        const R flux_in = std::max(dtrp * volti(k, iw) * fin[k], eps0);
        const R flux_out = std::min(dtrp * volti(k, iw) * fout[k], -eps0);
        const R epsr = eps1 * std::fabs(rlow);
        const R sc_in = std::max(rhos[k] * smax[k] - (rlow + epsr) - eps0, R(0));
        const R sc_out = std::min(rhos[k] * smin[k] - (rlow - epsr) + eps0, R(0));
        s1(k, iw) = std::min(R(1), sc_in / std::max(flux_in, R(0.01) * sc_in));
        s2(k, iw) = std::min(R(1), sc_out / std::min(flux_out, R(0.01) * sc_out));
      }
    }
    // renormalised vertical (:1235-1246) and horizontal (:1251-1266) face values
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw)
      for (int k = 2; k <= mza - 1; ++k) {
        R sc = std::min(s1(k + 1, iw), s2(k, iw));
        if (sfw(k, iw) < R(0)) sc = std::min(s1(k, iw), s2(k + 1, iw));
        upw(k, iw) = upw(k, iw) + hiw(k, iw) * sc;
      }
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv) {
      const int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
      for (int k = 2; k <= mza; ++k) {
        R sc = std::min(s1(k, iw2), s2(k, iw1));
        if (sfv(k, iv) < R(0)) sc = std::min(s1(k, iw1), s2(k, iw2));
        upv(k, iv) = upv(k, iv) + hiv(k, iv) * sc;
      }
    }
  }

  // ---------------------------------------------------------------- damping
  // divh_damp, vort_div_damp.f90:1-230
  void divh_damp(F2<R> &vmtx, double dtm) {
    const int lev = prm.divh_damp_level;
    const bool dodiv = (prm.divh_damp_fact > 1.e-7f) && lev > 0;
    if (!dorayfdiv && !dodiv) return;
    const bool dol = dorayfdiv || (dodiv && lev == 1), doh = dodiv && lev > 1;
    if (div_first) {
      div_first = false;
      div2d[0].alloc(mza, mwa); div2d[1].alloc(mza, mwa);
      // Rayleigh friction divergence damping profile, :60-66
      std::vector<R> zf(mza + 1, R(0));
      if (dorayfdiv)
        for (int k = krayfdiv_bot; k <= mza; ++k)
          zf[k] = (R)pow_r((zt[k] - prm.rayfdiv_zmin) / (zm[mza] - prm.rayfdiv_zmin), prm.rayfdiv_expon);
      if (dol) fdivl.alloc(mza, mva);
      if (doh) fdivh.alloc(mza, mva);
      double facth = doh ? -(double)powi(-(R)prm.divh_damp_fact, lev) / dtm : 0.0;
      if (plant == 5) facth = -facth;
      if (plant == 6 && doh) facth = -(double)powi(-(R)prm.divh_damp_fact, lev - 1) / dtm;
      const R fact2 = (dodiv && lev == 1) ? (R)(prm.divh_damp_fact / (R)dtm) : R(0);
      R factr = R(0);
      if (dorayfdiv && dodiv) factr = (R)(std::max(prm.divh_damp_fact, prm.rayfdiv_fact) / (R)dtm);
      else if (dorayfdiv) factr = (R)(prm.rayfdiv_fact / (R)dtm);
      for (int iv = 2; iv <= mva; ++iv) {
        int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
        if (dol) {
          int ka = (dodiv && lev == 1) ? 2 : krayfdiv_bot;
          for (int k = ka; k <= mza; ++k)
            fdivl(k, iv) = (R)((double)((fact2 * (R(1) - zf[k]) + factr * zf[k]) * dniv[iv] * zfacit[k] * dzit[k]) *
                               std::min(volt(k, iw1), volt(k, iw2)));
        }
        if (doh) {
          double factb = facth * (double)powi(std::min(arw0[iw1], arw0[iw2]), lev - 1);
          for (int k = 2; k <= mza; ++k)
            fdivh(k, iv) = (R)(factb * (double)dniv[iv] * (double)zfacit[k] * (double)dzit[k] *
                               std::min(volt(k, iw1), volt(k, iw2)) * (double)(R(1) - zf[k]));
        }
      }
    }
    // horizontal divergence (accumulates onto the saved div2d(:,:,1), :140-156)
    const int kd = dodiv ? 2 : std::max(2, krayfdiv_bot);
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw) {
      for (int j = 1; j <= npoly[iw]; ++j) {
        int iv = w_iv(j, iw);
        for (int k = kd; k <= mza; ++k)
          div2d[0](k, iw) = div2d[0](k, iw) - arv(k, iv) * dirv(j, iw) * vmc(k, iv);
      }
      for (int k = kd; k <= mza; ++k) div2d[0](k, iw) = div2d[0](k, iw) * volti(k, iw);
    }
    // second-order (and Rayleigh) damping, :168-185
    if (dol) {
      const int ka = (dodiv && lev == 1) ? 2 : krayfdiv_bot;
#pragma omp parallel for
      for (int iv = 2; iv <= mva; ++iv) {
        int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
        for (int k = ka; k <= mza; ++k)
          vmtx(k, iv) = vmtx(k, iv) + fdivl(k, iv) * (div2d[0](k, iw2) - div2d[0](k, iw1));
      }
    }
    if (!doh) return;
    // Laplacians of the divergence, :189-207, and the higher-order tendency, :211-226
    int ih = 0;
    for (int nn = 1; nn <= lev - 1; ++nn) {
      int il = (nn - 1) % 2;
      ih = nn % 2;
#pragma omp parallel for
      for (int iw = 2; iw <= mwa; ++iw) laplacian2d(iw, div2d[il], div2d[ih].col(iw) - 1);
    }
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv) {
      int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
      for (int k = 2; k <= mza; ++k)
        vmtx(k, iv) = vmtx(k, iv) + fdivh(k, iv) * (div2d[ih](k, iw2) - div2d[ih](k, iw1));
    }
  }

  // vort_damp, vort_div_damp.f90:234-445 (flat terrain: lpmmin = lpm = lpv)
  void vort_damp(F2<R> &vmtx, double dtm) {
    int lev = prm.vort_damp_level;
    if (prm.vort_damp_fact < 1.e-7f || lev <= 0) return;
    if (vort_first) {
      vort_first = false;
      fvorth.assign(mva + 1, 0); fvortl.assign(mva + 1, 0);
      vfact.alloc(3, mma);
      vort2d[0].alloc(mza, mma); vort2d[1].alloc(mza, mma);
      for (int im = 2; im <= mma; ++im)
        for (int jv = 1; jv <= 3; ++jv) {
          int iv = m_iv(jv, im);
          vfact(jv, im) = (v_im(2, iv) == im) ? dnv[iv] / arm0[im] : -dnv[iv] / arm0[im];
        }
      for (int iv = 2; iv <= mva; ++iv) {
        int im1 = v_im(1, iv), im2 = v_im(2, iv);
        R a = std::min(arm0[im1], arm0[im2]);
        fvorth[iv] = R(0.5) * dniu[iv] / (R)dtm * powi(-(R)prm.vort_damp_fact * a, lev);
        if (plant == 7) fvorth[iv] = -fvorth[iv];
        if (lev > 2)
          fvortl[iv] = R(0.5) * dniu[iv] / (R)dtm * powi(-(R)prm.vort_damp_fact * a, lev - 1);
      }
    }
#pragma omp parallel for
    for (int im = 2; im <= mma; ++im) {
      for (int k = 1; k <= mza; ++k) vort2d[0](k, im) = 0;
      for (int jv = 1; jv <= 3; ++jv) {
        int iv = m_iv(jv, im);
        for (int k = 2; k <= mza; ++k) vort2d[0](k, im) = vort2d[0](k, im) + vc(k, iv) * vfact(jv, im);
      }
    }
    int ih = 0;
    if (lev > 1) {
      for (int nn = 1; nn <= lev - 1; ++nn) {
        int il = (nn - 1) % 2;
        ih = nn % 2;
#pragma omp parallel for
        for (int im = 2; im <= mma; ++im) laplacian_m(im, vort2d[il], vort2d[ih].col(im) - 1);
      }
    }
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv) {
      int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv), im1 = v_im(1, iv), im2 = v_im(2, iv);
      for (int k = 2; k <= mza; ++k)
        vmtx(k, iv) = vmtx(k, iv) + fvorth[iv] * (vort2d[ih](k, im2) - vort2d[ih](k, im1)) *
                                        (R)(rho(k, iw1) + rho(k, iw2));
    }
  }

  // ---------------------------------------------------------------- prog_wrtv_rk
  // Scratch of prog_wrtv_rk (automatic arrays, prog_wrtv_rk.f90:97-137)
  F2<R> thil_upv, vxe_upv, vye_upv, vze_upv, thil_upw, vxe_upw, vye_upw, vze_upw;
  F2<R> vmxet_rk, vmyet_rk, vmzet_rk, thilt_rk, wmc0, vmc0, thilt_s, vmxet_s, vmyet_s, vmzet_s, vmt_s, vmca;
  F2<RP> rho0, rth0;
  F2<R> gs[5];
  // RK start values of the monotone-flux fields (prog_wrtv_rk.f90:117-120,188-199)
  F2<R> thl0, vxe0, vye0, vze0;
  F2<R> m_scpt, m_hiv, m_sfv, m_hiw, m_sfw, m_s1, m_s2;

  void alloc_wrtv_scratch() {
    if (thil_upv.n) return;
    for (auto *a : {&thil_upv, &vxe_upv, &vye_upv, &vze_upv, &vmc0, &vmt_s, &vmca}) a->alloc(mza, mva);
    for (auto *a : {&thil_upw, &vxe_upw, &vye_upw, &vze_upw, &vmxet_rk, &vmyet_rk, &vmzet_rk, &thilt_rk,
                    &wmc0, &thilt_s, &vmxet_s, &vmyet_s, &vmzet_s, &gs[0], &gs[1], &gs[2], &gs[3], &gs[4]})
      a->alloc(mza, mwa);
    rho0.alloc(mza, mwa); rth0.alloc(mza, mwa);
    if (prm.ithil_monot > 0 || prm.iwind_monot > 0) {
      for (auto *a : {&thl0, &vxe0, &vye0, &vze0, &m_scpt, &m_hiw, &m_sfw, &m_s1, &m_s2}) a->alloc(mza, mwa);
      m_hiv.alloc(mza, mva); m_sfv.alloc(mza, mva);
    }
  }

  // the part of prog_wrtv_rk before the RK stages (:165-308): RK start values, the
  // volume-weighted short-step tendencies with the Rayleigh terms, vmc0/vmt_short and
  // the damping tendencies
  void setup_short() {
    alloc_wrtv_scratch();
    const double dtsm = prm.dtsm;
    // setup loop, prog_wrtv_rk.f90:165-253 (no eddy diffusion)
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw) {
      thil_upw(mza, iw) = 0; thil_upw(1, iw) = 0;
      for (int k = 2; k <= mza; ++k) {
        R v4 = (R)volt(k, iw);
        rth0(k, iw) = (RP)thil(k, iw) * rho(k, iw);
        rho0(k, iw) = rho(k, iw);
        wmc0(k, iw) = wmc(k, iw);
        thilt_s(k, iw) = thilt(k, iw) * v4;
        vmxet_s(k, iw) = vmxet(k, iw) * v4;
        vmyet_s(k, iw) = vmyet(k, iw) * v4;
        vmzet_s(k, iw) = vmzet(k, iw) * v4;
      }
      // vertical eddy diffusion, :213-217
      if (turb_on) {
        solve_eddy_diff_heat(iw, thilt_s.col(iw) - 1);
        solve_eddy_diff_vxe(iw, vmxet_s.col(iw) - 1, vmyet_s.col(iw) - 1, vmzet_s.col(iw) - 1);
      }
      // Rayleigh friction on the velocity gradient at the top and on thil, :219-233
      if (dorayfmix) rayf_mix_top_vxe(iw, vmxet_s.col(iw) - 1, vmyet_s.col(iw) - 1, vmzet_s.col(iw) - 1);
      if (dorayf && prm.initial == 1)
        for (int k = krayf_bot; k <= mza; ++k) {
          R v4 = (R)volt(k, iw);
          thilt_s(k, iw) = thilt_s(k, iw) + (R)rayf_cof[k] * (R)dn01d[k] * ((R)th01d[k] - thil(k, iw)) * v4;
        }
      // horizontal diffusion, :235-251
      if (turb_on)
        for (int jv = 1; jv <= npoly[iw]; ++jv) {
          int iv = w_iv(jv, iw), iwn = w_iw(jv, iw);
          for (int k = 2; k <= khtopv[iv]; ++k)
            thilt_s(k, iw) = thilt_s(k, iw) + akhodx(k, iv) * (thil(k, iwn) - thil(k, iw));
          for (int k = 2; k <= kmtopv[iv]; ++k) {
            vmxet_s(k, iw) = vmxet_s(k, iw) + akmodx(k, iv) * (vxe(k, iwn) - vxe(k, iw));
            vmyet_s(k, iw) = vmyet_s(k, iw) + akmodx(k, iv) * (vye(k, iwn) - vye(k, iw));
            vmzet_s(k, iw) = vmzet_s(k, iw) + akmodx(k, iv) * (vze(k, iwn) - vze(k, iw));
          }
        }
      if (prm.ithil_monot > 0)
        for (int k = 2; k <= mza; ++k) thl0(k, iw) = thil(k, iw);
      if (prm.iwind_monot > 0)
        for (int k = 2; k <= mza; ++k) { vxe0(k, iw) = vxe(k, iw); vye0(k, iw) = vye(k, iw); vze0(k, iw) = vze(k, iw); }
    }
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv)
      for (int k = 1; k <= mza; ++k) { vmc0(k, iv) = vmc(k, iv); vmt_s(k, iv) = vmt(k, iv); }
    // Rayleigh friction on vmc toward vc03d, :268-284 (mdomain /= 4)
    if (dorayf) {
#pragma omp parallel for
      for (int iv = 2; iv <= mva; ++iv) {
        int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
        for (int k = krayf_bot; k <= mza; ++k)
          vmt_s(k, iv) = (R)((RP)vmt_s(k, iv) + (RP)((R)rayf_cof[k] * R(0.5)) * (rho(k, iw1) + rho(k, iw2)) *
                                                    (RP)(vc03d(k, iv) - vc(k, iv)));
      }
    }
    // short-step damping, :307-308
    divh_damp(vmt_s, dtsm);
    vort_damp(vmt_s, dtsm);
  }

  void prog_wrtv_rk() override {
    if (dorayf && !vc03d.n) { err = "Rayleigh friction on thil/vmc needs init_wrtv_rk"; return; }
    setup_short();
    const double dtsm = prm.dtsm;
    const double onethird = 1.0 / 3.0;
    for (int istage = 1; istage <= prm.nrk_wrtv; ++istage) {
      double dt8 = (prm.nrk_wrtv <= 2) ? (istage == 1 ? 0.5 : 1.0) * dtsm
                                       : (istage == 1 ? onethird : (istage == 2 ? 0.5 : 1.0)) * dtsm;
      if (plant == 8 && istage == 1) dt8 = 0.5 * dtsm;   // plant 8: a wrong first-stage weight
      R dts = (R)dt8;
#pragma omp parallel for
      for (int iw = 2; iw <= mwa; ++iw)
        for (int k = 2; k <= mza; ++k) {
          vmxet_rk(k, iw) = vmxet_s(k, iw); vmyet_rk(k, iw) = vmyet_s(k, iw);
          vmzet_rk(k, iw) = vmzet_s(k, iw); thilt_rk(k, iw) = thilt_s(k, iw);
        }
#pragma omp parallel for
      for (int iv = 2; iv <= mva; ++iv) {
        if (prm.nudflag > 0 && prm.nud_preserve_total_mass)  // :347-355
          for (int k = 2; k <= mza; ++k) vmca(k, iv) = vmc(k, iv) * arv(k, iv) + vmanud(k, iv);
        else
          for (int k = 2; k <= mza; ++k) vmca(k, iv) = vmc(k, iv) * arv(k, iv);
        if (istage == prm.nrk_wrtv)
          for (int k = 2; k <= mza; ++k) vmasc(k, iv) = vmasc(k, iv) + vmca(k, iv);
      }
      const bool last = istage == prm.nrk_wrtv;
      if (last && prm.ithil_monot > 0)   // :370-378
        fluxes_wrtv_monot(thil, thl0, thilt_s, thil_upw, thil_upv, dts, prm.thil_horiz_adv_order, gs);
      else
        fluxes_wrtv(thil, thil_upw, thil_upv, prm.thil_horiz_adv_order, gs);
      if (last && prm.iwind_monot > 0) {  // :380-399
        fluxes_wrtv_monot(vxe, vxe0, vmxet_s, vxe_upw, vxe_upv, dts, prm.wind_horiz_adv_order, gs);
        fluxes_wrtv_monot(vye, vye0, vmyet_s, vye_upw, vye_upv, dts, prm.wind_horiz_adv_order, gs);
        fluxes_wrtv_monot(vze, vze0, vmzet_s, vze_upw, vze_upv, dts, prm.wind_horiz_adv_order, gs);
      } else {
        fluxes_wrtv(vxe, vxe_upw, vxe_upv, prm.wind_horiz_adv_order, gs);
        fluxes_wrtv(vye, vye_upw, vye_upv, prm.wind_horiz_adv_order, gs);
        fluxes_wrtv(vze, vze_upw, vze_upv, prm.wind_horiz_adv_order, gs);
      }
#pragma omp parallel for
      for (int iw = 2; iw <= mwa; ++iw) prog_wrt_begs(iw, istage, dt8);
#pragma omp parallel for
      for (int iv = 2; iv <= mva; ++iv) prog_v_begs(iv, dts);
#pragma omp parallel for
      for (int iw = 2; iw <= mwa; ++iw) vel_t3d_hex(iw);
    }
  }

  // prog_wrt_begs, prog_wrtv_rk.f90:494-884 (global domain, no nudging)
  void prog_wrt_begs(int iw, int istage, double dt8) {
    const RP fr = RP(0.70), fp2 = RP(0.70), fp3 = RP(0.70);
    const R cpocv = (R)(1004.0 / 717.0), rocv = (R)(287.0 / 717.0);
    const RP pc2 = fp2 * (RP)cpocv, pc3 = fp3 * (RP)cpocv;
    const R fcor = plant == 1 ? -(R)prm.fcoriol : (R)prm.fcoriol;
    R dts = (R)dt8;
    int ka = 2;
    RP hflx_rho[130], delex_rho[130], delex_rhothil[130], rhothil[130], vflux_thil[130];
    RP press_ex[130], rho_ex[130], b5[130], b10[130];
    R delex_wm[130], wma[130], vflux_vxe[130], vflux_vye[130], vflux_vze[130];
    R b15[130];
    RT b31[130], b32[130], b33[130], rhs[130], wmf[130];
    vflux_thil[ka - 1] = 0; vflux_thil[mza] = 0;
    vflux_vxe[ka - 1] = 0; vflux_vxe[mza] = 0;
    vflux_vye[ka - 1] = 0; vflux_vye[mza] = 0;
    vflux_vze[ka - 1] = 0; vflux_vze[mza] = 0;
    if (prm.nudflag == 1 && prm.nud_preserve_total_mass) {  // nudging mass fluxes, :605-621
      for (int k = ka; k <= mza - 1; ++k) {
        vflux_thil[k] = (RP)(wmanud(k, iw) * thil_upw(k, iw));
        vflux_vxe[k] = wmanud(k, iw) * vxe_upw(k, iw);
        vflux_vye[k] = wmanud(k, iw) * vye_upw(k, iw);
        vflux_vze[k] = wmanud(k, iw) * vze_upw(k, iw);
      }
      for (int k = ka; k <= mza; ++k) {
        hflx_rho[k] = (RP)(wmanud(k - 1, iw) - wmanud(k, iw));
        thilt_rk(k, iw) = (R)((RP)thilt_rk(k, iw) + vflux_thil[k - 1] - vflux_thil[k]);
        vmxet_rk(k, iw) = vmxet_rk(k, iw) + vflux_vxe[k - 1] - vflux_vxe[k];
        vmyet_rk(k, iw) = vmyet_rk(k, iw) + vflux_vye[k - 1] - vflux_vye[k];
        vmzet_rk(k, iw) = vmzet_rk(k, iw) + vflux_vze[k - 1] - vflux_vze[k];
      }
    } else {
      for (int k = ka; k <= mza; ++k) hflx_rho[k] = 0;
    }
    for (int jv = 1; jv <= npoly[iw]; ++jv) {  // :635-661
      int iv = w_iv(jv, iw);
      if ((dirv(jv, iw) > R(0)) != (plant == 3)) {
        // plant 3 flips the sign of the face fluxes of the momentum (and thil) only
        for (int k = 2; k <= mza; ++k) {
          hflx_rho[k] = hflx_rho[k] + (plant == 3 ? -(RP)vmca(k, iv) : (RP)vmca(k, iv));
          thilt_rk(k, iw) = thilt_rk(k, iw) + vmca(k, iv) * thil_upv(k, iv);
          vmxet_rk(k, iw) = vmxet_rk(k, iw) + vmca(k, iv) * vxe_upv(k, iv);
          vmyet_rk(k, iw) = vmyet_rk(k, iw) + vmca(k, iv) * vye_upv(k, iv);
          vmzet_rk(k, iw) = vmzet_rk(k, iw) + vmca(k, iv) * vze_upv(k, iv);
        }
      } else {
        for (int k = 2; k <= mza; ++k) {
          hflx_rho[k] = hflx_rho[k] - (plant == 3 ? -(RP)vmca(k, iv) : (RP)vmca(k, iv));
          thilt_rk(k, iw) = thilt_rk(k, iw) - vmca(k, iv) * thil_upv(k, iv);
          vmxet_rk(k, iw) = vmxet_rk(k, iw) - vmca(k, iv) * vxe_upv(k, iv);
          vmyet_rk(k, iw) = vmyet_rk(k, iw) - vmca(k, iv) * vye_upv(k, iv);
          vmzet_rk(k, iw) = vmzet_rk(k, iw) - vmca(k, iv) * vze_upv(k, iv);
        }
      }
    }
    for (int k = ka; k <= mza; ++k) {  // Coriolis, global, :675-681
      R mass = (R)(rho(k, iw) * (RP)volt(k, iw));
      vmxet_rk(k, iw) = vmxet_rk(k, iw) + mass * fcor * vye(k, iw);
      vmyet_rk(k, iw) = vmyet_rk(k, iw) - mass * fcor * vxe(k, iw);
    }
    for (int k = ka; k <= mza; ++k) {  // :687-700
      rhothil[k] = rho(k, iw) * (RP)thil(k, iw);
      b10[k] = (RP)dt8 * (RP)volti(k, iw);
      b5[k] = (RP)(alpha_press(k, iw) * pow_r((R)rhothil[k], rocv));
      b15[k] = (R)(b10[k] * b5[k]);
      delex_rho[k] = b10[k] * hflx_rho[k];
      delex_rhothil[k] = b10[k] * (RP)thilt_rk(k, iw);
    }
    if (prm.nudflag > 0 && !prm.nud_preserve_total_mass)  // :703-706
      for (int k = ka; k <= mza; ++k) delex_rho[k] = delex_rho[k] + dt8 * rhot_nud(k, iw);
    if (istage > 1)
      for (int k = ka; k <= mza; ++k) {  // :710-715
        delex_rho[k] = delex_rho[k] + rho0(k, iw) - rho(k, iw);
        delex_rhothil[k] = delex_rhothil[k] + rth0(k, iw) - rhothil[k];
      }
    for (int k = ka; k <= mza; ++k) {
      press_ex[k] = b5[k] * (rhothil[k] + pc2 * delex_rhothil[k]);
      rho_ex[k] = rho(k, iw) + fr * delex_rho[k];
    }
    for (int k = ka; k <= mza - 1; ++k)  // :727-745
      delex_wm[k] = wmc0(k, iw) +
          dts * ((R)(press_ex[k] - press_ex[k + 1]) * pwfac(k, iw) -
                 (R)((RP)gdz_wgtm8[k] * rho_ex[k] + (RP)gdz_wgtp8[k] * rho_ex[k + 1]) +
                 (wnx[iw] * (vmxet_rk(k, iw) + vmxet_rk(k + 1, iw)) +
                  wny[iw] * (vmyet_rk(k, iw) + vmyet_rk(k + 1, iw)) +
                  wnz[iw] * (vmzet_rk(k, iw) + vmzet_rk(k + 1, iw))) * volwi(k, iw));
    R c8 = dts * (R)pc2;
    R c9 = -dts * dts * (R)fr;
    for (int k = ka; k <= mza - 1; ++k) {  // :752-769
      R b7 = dts * volwi(k, iw);
      R b8 = c8 * pwfac(k, iw);
      R b21 = b7 * wc(k - 1, iw);
      R b22 = b7 * wc(k + 1, iw);
      R b23 = b8 * b15[k];
      R b24 = b8 * b15[k + 1];
      R b25 = c9 * volti(k, iw) * gdz_wgtm[k];
      R b26 = c9 * volti(k + 1, iw) * gdz_wgtp[k];
      b31[k] = (RT)(-arw(k - 1, iw) * (b21 + b25 + b23 * thil_upw(k - 1, iw)));
      b32[k] = (RT)((R)b20[k] + arw(k, iw) * (b25 - b26 + (b23 + b24) * thil_upw(k, iw)));
      b33[k] = (RT)(arw(k + 1, iw) * (b22 + b26 - b24 * thil_upw(k + 1, iw)));
      rhs[k] = (RT)delex_wm[k];
    }
    if (ka <= mza - 1) tridiffo(ka, mza - 1, b31, b32, b33, rhs, wmf);
    wma[ka - 1] = 0; wma[mza] = 0;
    if (prm.nudflag > 0 && prm.nud_preserve_total_mass)  // :782-788
      for (int k = ka; k <= mza - 1; ++k) wma[k] = (R)wmf[k] * arw(k, iw) + wmanud(k, iw);
    else
      for (int k = ka; k <= mza - 1; ++k) wma[k] = (R)wmf[k] * arw(k, iw);
    for (int k = ka; k <= mza - 1; ++k) {
      vflux_thil[k] = (RP)(wma[k] * thil_upw(k, iw));
      vflux_vxe[k] = wma[k] * vxe_upw(k, iw);
      vflux_vye[k] = wma[k] * vye_upw(k, iw);
      vflux_vze[k] = wma[k] * vze_upw(k, iw);
    }
    if (istage == prm.nrk_wrtv)
      for (int k = ka; k <= mza - 1; ++k) wmasc(k, iw) = wmasc(k, iw) + wma[k];
    for (int k = ka; k <= mza; ++k) {  // :818-844
      rho(k, iw) = rho0(k, iw) + (RP)dt8 * (RP)volti(k, iw) * (hflx_rho[k] + (RP)wma[k - 1] - (RP)wma[k]);
      delex_rhothil[k] = delex_rhothil[k] + b10[k] * (vflux_thil[k - 1] - vflux_thil[k]);
      press(k, iw) = b5[k] * (rhothil[k] + pc3 * delex_rhothil[k]);
      thil(k, iw) = (R)((rhothil[k] + delex_rhothil[k]) / rho(k, iw));
      vmxet_rk(k, iw) = vmxet_rk(k, iw) + vflux_vxe[k - 1] - vflux_vxe[k];
      vmyet_rk(k, iw) = vmyet_rk(k, iw) + vflux_vye[k - 1] - vflux_vye[k];
      vmzet_rk(k, iw) = vmzet_rk(k, iw) + vflux_vze[k - 1] - vflux_vze[k];
    }
    for (int k = ka; k <= mza - 1; ++k) {  // :869-872
      wmc(k, iw) = (R)wmf[k];
      wc(k, iw) = (R)wmf[k] / (R)((RP)zwgt_bot8[k] * rho(k, iw) + (RP)zwgt_top8[k] * rho(k + 1, iw));
    }
    wc(ka - 1, iw) = wc(ka, iw);
    for (int k = 1; k <= ka - 1; ++k) thil(k, iw) = thil(ka, iw);
  }

  // prog_v_begs, prog_wrtv_rk.f90:888-946
  void prog_v_begs(int iv, R dts) {
    int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
    for (int k = 2; k <= mza; ++k) {
      R pgf = (R)(plant == 2 ? press(k, iw2) - press(k, iw1) : press(k, iw1) - press(k, iw2)) * pvfac(k, iv);
      vmc(k, iv) = vmc0(k, iv) + dts * (vmt_s(k, iv) + pgf +
                                        (vnx[iv] * (vmxet_rk(k, iw1) + vmxet_rk(k, iw2)) +
                                         vny[iv] * (vmyet_rk(k, iw1) + vmyet_rk(k, iw2)) +
                                         vnz[iv] * (vmzet_rk(k, iw1) + vmzet_rk(k, iw2))) * volvi(k, iv));
      vc(k, iv) = R(2) * vmc(k, iv) / (R)(rho(k, iw1) + rho(k, iw2));
    }
  }

  // vel_t3d_hex, vel_t3d.f90:39-108 (lve2 = 0)
  void vel_t3d_hex(int iw) {
    int ka = 2;
    for (int k = ka; k <= mza; ++k) {
      R wst = wc(k - 1, iw) + wc(k, iw);
      vxe(k, iw) = wst * wnxo2[iw];
      vye(k, iw) = wst * wnyo2[iw];
      vze(k, iw) = wst * wnzo2[iw];
    }
    for (int jv = 1; jv <= npoly[iw]; ++jv) {
      int iv = w_iv(jv, iw);
      for (int k = 2; k <= mza; ++k) {
        vxe(k, iw) = vxe(k, iw) + ecx(jv, iw) * vc(k, iv);
        vye(k, iw) = vye(k, iw) + ecy(jv, iw) * vc(k, iv);
        vze(k, iw) = vze(k, iw) + ecz(jv, iw) * vc(k, iv);
      }
    }
    vxe(1, iw) = vxe(ka, iw); vye(1, iw) = vye(ka, iw); vze(1, iw) = vze(ka, iw);
  }

  // ---------------------------------------------------------------- scalar transport
  F2<R> scp0, gt[5], upv;
  std::vector<R> ffimp;   // ff_implic(n, iw) at (iw-1)*nrk + n-1
  F2<R> rhos[3];

  // check_cfls, scalar_transport_rk.F90:1476-1712 (diagnostic maxima kept for the report)
  void check_cfls(const F2<R> &rho_o, int nrk) {
    const R dt = (R)prm.dtlm;
    double ctot = 0, chor = 0, wmax = 0;
    ffimp.assign((size_t)mwa * nrk, R(0));
#pragma omp parallel for reduction(max : ctot, chor, wmax)
    for (int iw = 2; iw <= mwa; ++iw) {
      R cflh[130], cfl[130], fimp[3] = {0, 0, 0};
      for (int k = 1; k <= mza; ++k) cflh[k] = 0;
      for (int n = 1; n <= npoly[iw]; ++n) {
        int iv = w_iv(n, iw);
        for (int k = 2; k <= mza; ++k) cflh[k] = cflh[k] - std::min(dirv(n, iw) * vmasc(k, iv), R(0));
      }
      for (int k = 2; k <= mza; ++k) {
        R dovr = dt * volti(k, iw) / rho_o(k, iw);
        cflh[k] = cflh[k] * dovr;
        R cflw = (std::max(wmasc(k, iw), R(0)) - std::min(wmasc(k - 1, iw), R(0))) * dovr;
        cfl[k] = cflh[k] + cflw;
        cflw = std::max(cflw, R(0.1));
        fimp[nrk - 1] = std::max(fimp[nrk - 1], (cfl[k] - R(0.98)) / cflw);
        fimp[nrk - 2] = std::max(fimp[nrk - 2], (cfl[k] - R(1.98)) / cflw);
        if (nrk == 3) fimp[0] = std::max(fimp[0], (cfl[k] - R(2.98)) / cflw);
      }
      for (int n = 0; n < nrk; ++n) {
        R f = std::min(R(1), fimp[n]);
        if (f < R(0.02)) f = 0;
        ffimp[(size_t)(iw - 1) * nrk + n] = f;
      }
      for (int k = 2; k <= mza; ++k) {
        ctot = std::max(ctot, (double)cfl[k]);
        chor = std::max(chor, (double)cflh[k]);
        wmax = std::max(wmax, (double)std::fabs(wc(k, iw)));
      }
    }
    cfl_out[0] = ctot; cfl_out[1] = chor; cfl_out[2] = wmax;
  }

  // update_scalar, scalar_transport_rk.F90:389-615 (iorderv = 3).  q is updated
  // from its own old values; rhosf is the final-stage density (:328-337).
  void update_scalar(F2<R> &q, const F2<R> &q0, const F2<R> &st, const F2<R> &rhosf, R dtr,
                     int iorder, int istage, int nrk) {
    recon_coefs(q, gt, iorder, true);   // bounded quadratic (:429-430)
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv) {  // scalar_v_column :474-533
      for (int k = 2; k <= mza; ++k) {
        const R s1 = face_value(q, gt, iorder, 1, iv, k, true);
        const R s2 = face_value(q, gt, iorder, 2, iv, k, true);
        upv(k, iv) = s1;
        if (vmasc(k, iv) < R(0)) upv(k, iv) = s2;
      }
    }
#pragma omp parallel
    {
      std::vector<R> sb(mza + 2), stp(mza + 2), upw(mza + 2), fd(mza + 2);
#pragma omp for
      for (int iw = 2; iw <= mwa; ++iw) {  // scalar_w_column :537-611
        int kb = 2;
        R fimpl = ffimp[(size_t)(iw - 1) * nrk + istage - 1];
        R fexpl = R(1) - fimpl;
        upw[kb - 1] = q(kb, iw);
        upw[mza] = q(mza, iw);
        grad_z_quadratic(&q(1, iw) - 1, sb.data(), stp.data());
        for (int k = kb; k <= mza - 1; ++k) {
          upw[k] = stp[k];
          if (wmasc(k, iw) < R(0)) upw[k] = sb[k + 1];
        }
        for (int k = 1; k <= mza; ++k) fd[k] = 0;
        for (int jv = 1; jv <= npoly[iw]; ++jv) {
          int iv = w_iv(jv, iw);
          for (int k = 2; k <= mza; ++k)
            fd[k] = fd[k] + dirv(jv, iw) * vmasc(k, iv) * (upv(k, iv) - q0(k, iw));
        }
        for (int k = kb; k <= mza; ++k) {
          fd[k] = fd[k] + fexpl * (wmasc(k - 1, iw) * (upw[k - 1] - q0(k, iw)) -
                                   wmasc(k, iw) * (upw[k] - q0(k, iw)));
          q(k, iw) = q0(k, iw) + dtr * (st(k, iw) + volti(k, iw) * fd[k]) / rhosf(k, iw);
        }
        for (int k = 1; k <= kb - 1; ++k) q(k, iw) = q(kb, iw);
      }
    }
  }

  // implicit vertical coefficients, scalar_transport_rk.F90:186-217 with
  // tridif_prep (tridiag_lib.f90:321-340); solve with tridif_fini (:344-364).
  void implicit_column(F2<R> &q, const F2<R> &q0, int iw, int istage, R dtr, const R *dtlr, int nrk) {
    R ff = ffimp[(size_t)(iw - 1) * nrk + istage - 1];
    if (!(ff > R(0.001))) return;
    int kb = 2;
    R wup[130], wdn[130], b1[130], b2a[130], b3a[130], scr1[130], cji[130], dscp0[130], delex[130], del[130];
    wup[mza] = 0; wdn[mza] = 0; wup[kb - 1] = 0; wdn[kb - 1] = 0;
    for (int k = kb; k <= mza - 1; ++k) { wup[k] = std::max(wmasc(k, iw), R(0)); wdn[k] = std::min(wmasc(k, iw), R(0)); }
    R c1 = ff * dtlr[istage - 1];
    for (int k = kb; k <= mza; ++k) {
      R c2 = c1 * volti(k, iw) / rhos[istage - 1](k, iw);
      b1[k] = -c2 * wup[k - 1];
      b2a[k] = R(1) + c2 * (wup[k] - wdn[k - 1]);
      b3a[k] = c2 * wdn[k];
    }
    cji[kb] = R(1) / b2a[kb];
    for (int k = kb + 1; k <= mza; ++k) {
      scr1[k] = cji[k - 1] * b3a[k - 1];
      cji[k] = R(1) / (b2a[k] - b1[k] * scr1[k]);
    }
    dscp0[mza] = 0; dscp0[kb - 1] = 0;
    for (int k = kb; k <= mza - 1; ++k) dscp0[k] = q0(k + 1, iw) - q0(k, iw);
    for (int k = kb; k <= mza; ++k) {
      R cc = ff * dtr * volti(k, iw) / rhos[istage - 1](k, iw);
      delex[k] = q(k, iw) - q0(k, iw) - cc * (std::min(wmasc(k, iw), R(0)) * dscp0[k] +
                                              std::max(wmasc(k - 1, iw), R(0)) * dscp0[k - 1]);
    }
    del[kb] = delex[kb] * cji[kb];
    for (int k = kb + 1; k <= mza; ++k) del[k] = cji[k] * (delex[k] - b1[k] * del[k - 1]);
    for (int k = mza - 1; k >= kb; --k) del[k] = del[k] - scr1[k + 1] * del[k + 1];
    for (int k = kb; k <= mza; ++k) q(k, iw) = q0(k, iw) + del[k];
  }

  // rhof of scalar_transport_rk.F90:120-133: the density minus the long step's nudging
  // density tendency when nudging preserves the mixing ratio (global grid: every column
  // is within max_nud_mrl), the density otherwise; FP64 arithmetic rounded to FP32.
  // oname_check.f90:218 turns nud_preserve_mix_ratio off when nud_preserve_total_mass
  // is on (rhot_nud is then the potential of the nudging mass fluxes).
  bool nudged_rhof() const {
    return prm.nudflag > 0 && prm.nud_preserve_mix_ratio && !prm.nud_preserve_total_mass && rhot_nud.n;
  }
  R rhof_at(int k, int iw) const {
    if (nudged_rhof()) return (R)((double)rho(k, iw) - prm.dtlm * (double)rhot_nud(k, iw));
    return (R)rho(k, iw);
  }

  // horizontal diffusion of a tracer into its tendency, scalar_transport_rk.F90:244-274
  // (kwtop(iw) = max khtopv over the cell's faces, :160; a face contributes up to its
  // own khtopv)
  void tracer_sgsmix(const F2<R> &q, F2<R> &st) const {
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw) {
      R fd[130];
      int kwtop = 0;
      for (int jv = 1; jv <= npoly[iw]; ++jv) kwtop = std::max(kwtop, khtopv[w_iv(jv, iw)]);
      if (kwtop < 2) continue;
      for (int k = 2; k <= kwtop; ++k) fd[k] = 0;
      for (int jv = 1; jv <= npoly[iw]; ++jv) {
        int iv = w_iv(jv, iw), iwn = w_iw(jv, iw);
        for (int k = 2; k <= khtopv[iv]; ++k) fd[k] = fd[k] + akhodx(k, iv) * (q(k, iwn) - q(k, iw));
      }
      for (int k = 2; k <= kwtop; ++k) st(k, iw) = st(k, iw) + fd[k] * volti(k, iw);
    }
  }
  // transpose of tracer_sgsmix: lq += D^T ls, D q = volti sum_faces akhodx (q_n - q_c)
  void tracer_sgsmix_adj(const F2<R> &ls, F2<R> &lq) const {
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw)
      for (int jv = 1; jv <= npoly[iw]; ++jv) {
        int iv = w_iv(jv, iw), iwn = w_iw(jv, iw);
        for (int k = 2; k <= khtopv[iv]; ++k)
          lq(k, iw) = lq(k, iw) + akhodx(k, iv) * (volti(k, iwn) * ls(k, iwn) - volti(k, iw) * ls(k, iw));
      }
  }

  void stage_densities(const F2<R> &rho_o, int nrk) {
    for (int n = 0; n < nrk; ++n) if (!rhos[n].n) rhos[n].alloc(mza, mwa);
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw)
      for (int k = 2; k <= mza; ++k) {  // scalar_transport_rk.F90:120-150
        R rf = rhof_at(k, iw);
        if (nrk == 3) {
          rhos[0](k, iw) = R(2.) / R(3.) * rho_o(k, iw) + R(1.) / R(3.) * rf;
          rhos[1](k, iw) = R(0.5) * (rho_o(k, iw) + rf);
          rhos[2](k, iw) = rf;
        } else {
          rhos[0](k, iw) = R(0.5) * (rho_o(k, iw) + rf);
          rhos[1](k, iw) = rf;
        }
      }
  }

  void stage_dts(R *dtlr, int nrk) const {
    R dtl = (R)prm.dtlm;
    const R one3 = R(1.) / R(3.);
    if (nrk == 3) { dtlr[0] = dtl * one3; dtlr[1] = dtl * R(0.5); dtlr[2] = dtl * R(1.0); }
    else { dtlr[0] = dtl * R(0.5); dtlr[1] = dtl * R(1.0); }
    if (plant == 9 && nrk == 3) dtlr[0] = dtl * R(0.5);   // plant 9: a wrong first-stage weight
  }

  // scalar_transport_rk, scalar_transport_rk.F90:30-385 (limiters in update_scalar_lim)
  void scalar_transport_rk(const float *rho_old_host) override {
    if (!scp0.n) {
      scp0.alloc(mza, mwa); upv.alloc(mza, mva);
      for (auto &g : gt) g.alloc(mza, mwa);
    }
    F2<R> rho_o;
    const F2<R> *ro = &rho_old;
    if (rho_old_host) { load2(rho_o, rho_old_host, mza, mwa); ro = &rho_o; }
    const int nrk = prm.nrk_scal;
    R dtlr[3];
    stage_dts(dtlr, nrk);
    check_cfls(*ro, nrk);
    stage_densities(*ro, nrk);
    int nimp = 0;
    for (int iw = 2; iw <= mwa; ++iw) {
      bool any = false;
      for (int n = 0; n < nrk; ++n) any = any || ffimp[(size_t)(iw - 1) * nrk + n] > R(0.001);
      nimp += any;
    }
    cfl_out[3] = nimp;
    for (size_t n = 0; n < scp.size(); ++n) {
      F2<R> &q = scp[n];
      scp0 = q;
      if (turb_on && sgsmix[n]) tracer_sgsmix(q, sct[n]);
      const int imonot = prm.iscal_monot;
      tape_lim[n] = 0;
      for (int istage = 1; istage <= nrk; ++istage) {
        const R dtr = dtlr[istage - 1];
        // final stage of the monotone / positive-definite options: low-order update of
        // scp0 by the tendency first (scalar_transport_rk.F90:297-313)
        if (istage == nrk && imonot > 0 && pdef[n]) {
#pragma omp parallel for
          for (int iw = 2; iw <= mwa; ++iw)      // (the copy scp0(2:lpw-1) is empty for lpw = 2)
            for (int k = 2; k <= mza; ++k) scp0(k, iw) = scp0(k, iw) + dtr * sct[n](k, iw) / (*ro)(k, iw);
        }
        if (istage == nrk && imonot == 1)
          update_scalar_lim(n, q, scp0, rhos[nrk - 1], dtr, prm.scal_horiz_adv_order, istage, nrk, 1);
        else if (istage == nrk && imonot == 2 && pdef[n])
          update_scalar_lim(n, q, scp0, rhos[nrk - 1], dtr, prm.scal_horiz_adv_order, istage, nrk, 2);
        else
          update_scalar(q, scp0, sct[n], rhos[nrk - 1], dtr, prm.scal_horiz_adv_order, istage, nrk);
        if (nimp) {
#pragma omp parallel for
          for (int iw = 2; iw <= mwa; ++iw) implicit_column(q, scp0, iw, istage, dtlr[istage - 1], dtlr, nrk);
        }
      }
    }
    // tape for the adjoint: winds and stage quantities of this step
    tape_vmasc = vmasc; tape_wmasc = wmasc; tape_ffimp = ffimp;
    for (int n = 0; n < nrk; ++n) tape_rhos[n] = rhos[n];
    have_tape = true;
  }

  // ---------------------------------------------------------------- limiters
  // grad_z_quadratic (grad_lib.f90:166-229) with its options: itop/ibot = 1 extend the
  // end differences (constant gradient), bounded selects a_w or the unbounded a_wu.
  void grad_z_q(const R *s, R *sb, R *st, int itop, int ibot, bool bounded) const {
    R ds[130];
    const int ka = 2;
    for (int k = ka; k <= mza - 1; ++k) ds[k] = s[k + 1] - s[k];
    ds[ka - 1] = ibot ? ds[ka] : R(0);
    ds[mza] = itop ? ds[mza - 1] : R(0);
    for (int k = ka; k <= mza; ++k) {
      if (bounded) {
        sb[k] = s[k] + ds[k] * aw_(k, 1, 1) + ds[k - 1] * aw_(k, 2, 1);
        st[k] = s[k] + ds[k] * aw_(k, 1, 2) + ds[k - 1] * aw_(k, 2, 2);
      } else {
        sb[k] = s[k] + ds[k] * awu_(k, 1, 1) + ds[k - 1] * awu_(k, 2, 1);
        st[k] = s[k] + ds[k] * awu_(k, 1, 2) + ds[k - 1] * awu_(k, 2, 2);
      }
    }
  }

  F2<R> l_sfv, l_upv, l_hiv, l_s1, l_s2, l_upw, l_hiw, l_sfw;
  // update_scalar_monot (mode 1, scalar_transport_rk.F90:619-1081) and
  // update_scalar_pd (mode 2, :1085-1472): the final-stage flux renormalisation of
  // the high-order corrections (scp_hiv, scp_hiw) to the low-order upwind fluxes from
  // scp0, with iorderv = 3 and centered_monot = .false. (:20-21).  The applied face and
  // vertical scales are kept in tape_sv / tape_sw for the frozen-limiter adjoint.
  void update_scalar_lim(int n, F2<R> &q, const F2<R> &q0, const F2<R> &rhosf, R dtr, int iorder, int istage,
                         int nrk, int mode) {
    const R eps0 = R(1.e-32), eps1 = R(3.e-07), onep = R(1) + eps1, onem = R(1) - eps1;
    if (!l_sfv.n) {
      l_sfv.alloc(mza, mva); l_upv.alloc(mza, mva); l_hiv.alloc(mza, mva);
      l_s1.alloc(mza, mwa); l_s2.alloc(mza, mwa); l_upw.alloc(mza + 1, mwa); l_hiw.alloc(mza + 1, mwa);
      l_sfw.alloc(mza + 1, mwa);
    }
    if (!tape_sv[n].n) { tape_sv[n].alloc(mza, mva); tape_sw[n].alloc(mza, mwa); }
    tape_lim[n] = (char)mode;
    const bool frozen = frozen_lim;   // the linear map with the last forward's scales
    // order 3: unbounded quadratic with xx0_vu for the monotone form (:665, :793-807),
    // bounded with xx0_v for the positive-definite form (:1130, :1252-1266)
    const bool bnd = mode == 2;
    recon_coefs(q, gt, iorder, bnd);
    // scalar_vflux_monot (:748-827) / scalar_vflux_pd (:1227-1290)
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv) {
      int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
      for (int k = 2; k <= mza; ++k) {
        R s1, s2;
        if (iorder == 1) { s1 = q(k, iw1); s2 = (mode == 2) ? q(k, iw1) : q(k, iw2); }
        else {
          s1 = face_value(q, gt, iorder, 1, iv, k, bnd);
          s2 = face_value(q, gt, iorder, 2, iv, k, bnd);
        }
        R up = q0(k, iw1);
        if (vmasc(k, iv) < R(0)) up = q0(k, iw2);
        R hi = s1;
        if (vmasc(k, iv) < R(0)) hi = s2;
        l_upv(k, iv) = up;
        l_hiv(k, iv) = hi - up;
        l_sfv(k, iv) = vmasc(k, iv) * l_hiv(k, iv);
      }
    }
    // scalar_wflux_and_scales_monot (:830-1010) / scalar_wflux_and_scale_pd (:1293-1402)
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw) {
      const int kb = 2;
      const R fimpl = ffimp[(size_t)(iw - 1) * nrk + istage - 1];
      const R fexpl = R(1) - fimpl;
      R sb[130], stp[130], smax[130], smin[130], fl[130], fin[130], fout[130];
      R *upw = &l_upw(1, iw) - 1, *hiw = &l_hiw(1, iw) - 1, *sfw = &l_sfw(1, iw) - 1;
      if (mode == 1) { upw[kb - 1] = q0(kb, iw); upw[mza] = q0(mza, iw); }
      else { upw[kb - 1] = q(kb, iw); upw[mza] = q(mza, iw); }
      for (int k = kb; k <= mza - 1; ++k) {
        upw[k] = q0(k, iw);
        if (wmasc(k, iw) < R(0)) upw[k] = q0(k + 1, iw);
      }
      if (mode == 1) grad_z_q(&q(1, iw) - 1, sb, stp, 1, 1, false);
      else grad_z_q(&q(1, iw) - 1, sb, stp, 0, 0, true);
      sfw[kb - 1] = 0; sfw[mza] = 0; hiw[kb - 1] = 0; hiw[mza] = 0;
      for (int k = kb; k <= mza - 1; ++k) {
        hiw[k] = stp[k];
        if (wmasc(k, iw) < R(0)) hiw[k] = sb[k + 1];
        hiw[k] = hiw[k] - upw[k];
        sfw[k] = fexpl * wmasc(k, iw) * hiw[k];
      }
      for (int k = 1; k <= mza; ++k) { fl[k] = 0; fin[k] = 0; fout[k] = 0; }
      if (mode == 1) {
        for (int k = kb; k <= mza; ++k) { smax[k] = q0(k, iw); smin[k] = q0(k, iw); }
        for (int k = kb + 1; k <= mza; ++k)
          if (wmasc(k - 1, iw) > R(0)) {
            smax[k] = std::max(smax[k], q0(k - 1, iw));
            smin[k] = std::min(smin[k], q0(k - 1, iw));
          }
        for (int k = kb; k <= mza - 1; ++k)
          if (wmasc(k, iw) < R(0)) {
            smax[k] = std::max(smax[k], q0(k + 1, iw));
            smin[k] = std::min(smin[k], q0(k + 1, iw));
          }
      }
      for (int jv = 1; jv <= npoly[iw]; ++jv) {
        int iv = w_iv(jv, iw), iwn = w_iw(jv, iw);
        for (int k = 2; k <= mza; ++k) {
          fl[k] = fl[k] + dirv(jv, iw) * vmasc(k, iv) * (l_upv(k, iv) - q0(k, iw));
          if (mode == 1) fin[k] = fin[k] + std::max(dirv(jv, iw) * l_sfv(k, iv), R(0));
          fout[k] = fout[k] + std::min(dirv(jv, iw) * l_sfv(k, iv), R(0));
          if (mode == 1 && dirv(jv, iw) * vmasc(k, iv) > R(0)) {
            smax[k] = std::max(smax[k], q0(k, iwn));
            smin[k] = std::min(smin[k], q0(k, iwn));
          }
        }
      }
      if (fimpl > R(.001)) {
        for (int k = kb; k <= mza; ++k) {
          if (mode == 1) {
            R dwi = fimpl * (wmasc(k, iw) - wmasc(k - 1, iw));
            fl[k] = fl[k] + fexpl * (wmasc(k - 1, iw) * (upw[k - 1] - q0(k, iw)) - wmasc(k, iw) * (upw[k] - q0(k, iw)))
                    + q0(k, iw) * dwi;
            fin[k] = fin[k] + std::max(sfw[k - 1], R(0)) - std::min(sfw[k], R(0));
            fout[k] = fout[k] + std::min(sfw[k - 1], R(0)) - std::max(sfw[k], R(0));
            R imp = rhosf(k, iw) + dtr * volti(k, iw) * dwi;
            smax[k] = imp * smax[k];
            smin[k] = imp * smin[k];
          } else {
            fl[k] = fl[k] + fexpl * (wmasc(k - 1, iw) * (upw[k - 1] - q0(k, iw)) - wmasc(k, iw) * (upw[k] - q0(k, iw)))
                    + fimpl * q0(k, iw) * (wmasc(k, iw) - wmasc(k - 1, iw));
            fout[k] = fout[k] + std::min(sfw[k - 1], R(0)) - std::max(sfw[k], R(0));
          }
        }
      } else {
        for (int k = kb; k <= mza; ++k) {
          fl[k] = fl[k] + wmasc(k - 1, iw) * (upw[k - 1] - q0(k, iw)) - wmasc(k, iw) * (upw[k] - q0(k, iw));
          if (mode == 1) fin[k] = fin[k] + std::max(sfw[k - 1], R(0)) - std::min(sfw[k], R(0));
          fout[k] = fout[k] + std::min(sfw[k - 1], R(0)) - std::max(sfw[k], R(0));
          if (mode == 1) { smax[k] = rhosf(k, iw) * smax[k]; smin[k] = rhosf(k, iw) * smin[k]; }
        }
      }
      for (int k = kb; k <= mza; ++k) {
        R rlow = q0(k, iw) * rhosf(k, iw) + dtr * volti(k, iw) * fl[k];
        if (mode == 1) {
          R dtrp = onep * dtr;
          R flux_in = std::max(dtrp * volti(k, iw) * fin[k], eps0);
          R flux_out = std::min(dtrp * volti(k, iw) * fout[k], -eps0);
          R epsr = eps1 * std::fabs(rlow);
          R sc_in = std::max(smax[k] - (rlow + epsr) - eps0, R(0));
          R sc_out = std::min(smin[k] - (rlow - epsr) + eps0, R(0));
          l_s1(k, iw) = std::min(R(1), sc_in / std::max(flux_in, R(0.01) * sc_in));
          l_s2(k, iw) = std::min(R(1), sc_out / std::min(flux_out, R(0.01) * sc_out));
        } else {
          R mdtrp = -onep * dtr;
          R flux_out = std::max(mdtrp * volti(k, iw) * fout[k], eps0);
          R sc_out = std::max(onem * rlow - eps0, R(0));
          l_s1(k, iw) = std::min(R(1), sc_out / std::max(flux_out, R(0.01) * sc_out));
        }
      }
      // vertical renormalisation (:1003-1008, :1396-1401)
      for (int k = kb; k <= mza - 1; ++k) {
        R scale;
        if (mode == 1) {
          scale = std::min(l_s1(k + 1, iw), l_s2(k, iw));
          if (sfw[k] < R(0)) scale = std::min(l_s1(k, iw), l_s2(k + 1, iw));
        } else {
          scale = l_s1(k, iw);
          if (sfw[k] < R(0)) scale = l_s1(k + 1, iw);
        }
        if (frozen) scale = tape_sw[n](k, iw);
        else tape_sw[n](k, iw) = scale;
        upw[k] = upw[k] + hiw[k] * scale;
      }
    }
    // horizontal renormalisation (:704-716, :1183-1193)
#pragma omp parallel for
    for (int iv = 2; iv <= mva; ++iv) {
      int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
      for (int k = 2; k <= mza; ++k) {
        R scale;
        if (mode == 1) {
          scale = std::min(l_s1(k, iw2), l_s2(k, iw1));
          if (l_sfv(k, iv) < R(0)) scale = std::min(l_s1(k, iw1), l_s2(k, iw2));
        } else {
          scale = l_s1(k, iw1);
          if (l_sfv(k, iv) < R(0)) scale = l_s1(k, iw2);
        }
        if (frozen) scale = tape_sv[n](k, iv);
        else tape_sv[n](k, iv) = scale;
        l_upv(k, iv) = l_upv(k, iv) + l_hiv(k, iv) * scale;
      }
    }
    // scalar_tend_monot (:1013-1078) / scalar_tend_pd (:1405-1469); level 1 is left
    // as it was (the copy scp(2:kb-1) is empty for kb = 2)
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw) {
      const int kb = 2;
      const R fimpl = ffimp[(size_t)(iw - 1) * nrk + istage - 1];
      R fd[130];
      const R *upw = &l_upw(1, iw) - 1;
      for (int k = 1; k <= mza; ++k) fd[k] = 0;
      for (int jv = 1; jv <= npoly[iw]; ++jv) {
        int iv = w_iv(jv, iw);
        for (int k = 2; k <= mza; ++k) fd[k] = fd[k] + dirv(jv, iw) * vmasc(k, iv) * (l_upv(k, iv) - q0(k, iw));
      }
      if (fimpl > R(.001)) {
        R fexpl = R(1) - fimpl;
        for (int k = kb; k <= mza; ++k) {
          fd[k] = fd[k] + fexpl * (wmasc(k - 1, iw) * (upw[k - 1] - q0(k, iw)) - wmasc(k, iw) * (upw[k] - q0(k, iw)));
          q(k, iw) = q0(k, iw) + dtr * volti(k, iw) * fd[k] / rhosf(k, iw);
        }
      } else {
        for (int k = kb; k <= mza; ++k) {
          fd[k] = fd[k] + wmasc(k - 1, iw) * (upw[k - 1] - q0(k, iw)) - wmasc(k, iw) * (upw[k] - q0(k, iw));
          q(k, iw) = q0(k, iw) + dtr * volti(k, iw) * fd[k] / rhosf(k, iw);
        }
      }
    }
  }

  // Adjoint of update_scalar_lim with the renormalisation scales frozen at the values
  // the forward step applied (tape_sv, tape_sw).  Per face, the flux value is
  // (1 - s) q0(upwind) + s * reconstruction(q), so the face adjoint phi is split into
  // (1 - s) phi to the adjoint of scp0 and s phi to that of q (through grad_t2d or the
  // vertical quadratic), exactly as in the forward selections.
  void update_scalar_lim_adj(int n, const F2<R> &mu_in, F2<R> &lq, F2<R> &lq0, const F2<R> &rhosf, R dtr,
                             int iorder, int istage, int nrk, const R *dtlr) {
    const F2<R> &vm = tape_vmasc, &wm = tape_wmasc;
    const int mode = tape_lim[n];
    const bool bnd = mode == 2;     // the forward's coefficient and offset sets
    alloc_adj_scratch();
    F2<R> &nu = a_nu, *lg = a_lg;
    for (auto *a : {&nu, &lg[0], &lg[1], &lg[2], &lg[3], &lg[4]}) std::fill(a->v.begin(), a->v.end(), R(0));
    for (int iw = 2; iw <= mwa; ++iw) {
      const int kb = 2;
      R mu[130];
      for (int k = 1; k <= mza; ++k) mu[k] = mu_in(k, iw);
      const R ff = tape_ffimp[(size_t)(iw - 1) * nrk + istage - 1];
      const bool imp = ff > R(0.001);
      const R fexp = imp ? R(1) - ff : R(1);
      lq(1, iw) += mu[1];                      // level 1 passes through unchanged
      mu[1] = 0;
      if (imp) {
        R wup[130], wdn[130], b1[130], b2a[130], b3a[130], le[130], c[130], scr[130], cj[130];
        wup[mza] = 0; wdn[mza] = 0; wup[kb - 1] = 0; wdn[kb - 1] = 0;
        for (int k = kb; k <= mza - 1; ++k) { wup[k] = std::max(wm(k, iw), R(0)); wdn[k] = std::min(wm(k, iw), R(0)); }
        R c1 = ff * dtlr[istage - 1];
        for (int k = kb; k <= mza; ++k) {
          R c2 = c1 * volti(k, iw) / tape_rhos[istage - 1](k, iw);
          b1[k] = -c2 * wup[k - 1]; b2a[k] = R(1) + c2 * (wup[k] - wdn[k - 1]); b3a[k] = c2 * wdn[k];
          c[k] = ff * dtr * volti(k, iw) / tape_rhos[istage - 1](k, iw);
        }
        cj[kb] = R(1) / b2a[kb];
        le[kb] = mu[kb] * cj[kb];
        for (int k = kb + 1; k <= mza; ++k) {
          scr[k] = cj[k - 1] * b1[k];
          cj[k] = R(1) / (b2a[k] - b3a[k - 1] * scr[k]);
          le[k] = cj[k] * (mu[k] - b3a[k - 1] * le[k - 1]);
        }
        for (int k = mza - 1; k >= kb; --k) le[k] = le[k] - scr[k + 1] * le[k + 1];
        for (int k = kb; k <= mza; ++k) lq0(k, iw) = lq0(k, iw) + mu[k] - le[k];
        for (int k = kb; k <= mza; ++k) {
          R a = -c[k] * le[k];
          if (k <= mza - 1) { R t = a * std::min(wm(k, iw), R(0)); lq0(k + 1, iw) += t; lq0(k, iw) -= t; }
          if (k - 1 >= kb) { R t = a * std::max(wm(k - 1, iw), R(0)); lq0(k, iw) += t; lq0(k - 1, iw) -= t; }
        }
        for (int k = kb; k <= mza; ++k) mu[k] = le[k];
      }
      R gsum[130];
      for (int k = 1; k <= mza; ++k) gsum[k] = 0;
      for (int jv = 1; jv <= npoly[iw]; ++jv) {
        int iv = w_iv(jv, iw);
        for (int k = 2; k <= mza; ++k) gsum[k] = gsum[k] + dirv(jv, iw) * vm(k, iv);
      }
      for (int k = kb; k <= mza; ++k) {
        R v = mu[k] * dtr * volti(k, iw) / rhosf(k, iw);
        nu(k, iw) = v;
        lq0(k, iw) = lq0(k, iw) + mu[k] - v * (gsum[k] + fexp * (wm(k - 1, iw) - wm(k, iw)));
      }
      // vertical faces: upw(k) feeds fluxdiv(k+1) with +fexp wm(k), fluxdiv(k) with -fexp wm(k)
      R psi[131], lds[131];
      for (int k = 0; k <= mza; ++k) { psi[k] = 0; lds[k] = 0; }
      for (int k = kb - 1; k <= mza; ++k) {
        R nup = (k + 1 <= mza) ? nu(k + 1, iw) : R(0);
        R nuk = (k >= kb) ? nu(k, iw) : R(0);
        psi[k] = fexp * wm(k, iw) * (nup - nuk);
      }
      if (mode == 1) { lq0(kb, iw) += psi[kb - 1]; lq0(mza, iw) += psi[mza]; }
      else { lq(kb, iw) += psi[kb - 1]; lq(mza, iw) += psi[mza]; }
      for (int k = kb; k <= mza - 1; ++k) {
        const R s = tape_sw[n](k, iw);
        const bool neg = wm(k, iw) < R(0);
        lq0(neg ? k + 1 : k, iw) += (R(1) - s) * psi[k];
        const R h = s * psi[k];
        if (neg) {        // scpb(k+1) = q(k+1) + ds(k+1) A(k+1,1,1) + ds(k) A(k+1,2,1)
          lq(k + 1, iw) += h;
          lds[k + 1] += h * (mode == 1 ? awu_(k + 1, 1, 1) : aw_(k + 1, 1, 1));
          lds[k] += h * (mode == 1 ? awu_(k + 1, 2, 1) : aw_(k + 1, 2, 1));
        } else {          // scpt(k) = q(k) + ds(k) A(k,1,2) + ds(k-1) A(k,2,2)
          lq(k, iw) += h;
          lds[k] += h * (mode == 1 ? awu_(k, 1, 2) : aw_(k, 1, 2));
          lds[k - 1] += h * (mode == 1 ? awu_(k, 2, 2) : aw_(k, 2, 2));
        }
      }
      // ds(kb-1) and ds(mza): constant-gradient ends (monotone) copy ds(kb), ds(mza-1)
      if (mode == 1) { lds[kb] += lds[kb - 1]; lds[mza - 1] += lds[mza]; }
      for (int k = kb; k <= mza - 1; ++k) { lq(k + 1, iw) += lds[k]; lq(k, iw) -= lds[k]; }
    }
    // horizontal faces
    for (int iv = 2; iv <= mva; ++iv) {
      int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
      for (int k = 2; k <= mza; ++k) {
        const R phi = vm(k, iv) * (nu(k, iw2) - nu(k, iw1));
        const R s = tape_sv[n](k, iv);
        const bool neg = vm(k, iv) < R(0);
        lq0(k, neg ? iw2 : iw1) += (R(1) - s) * phi;
        const R h = s * phi;
        int side = neg ? 2 : 1;
        if (iorder == 1 && mode == 2) side = 1;       // scalar_vflux_pd order-1 form
        const int iwh = side == 1 ? iw1 : iw2;
        lq(k, iwh) += h;
        face_value_adj(lg, iorder, side, iv, k, h, bnd);
      }
    }
    recon_coefs_adj(lg, lq, iorder, bnd);
  }

  // ---------------------------------------------------------------- adjoint
  F2<R> a_nu, a_lqpart, a_lg[5];
  F2<R> a_lsct;                 // adjoint of the tracer tendency (with horizontal diffusion)
  bool a_lsct_on = false;
  void alloc_adj_scratch() {
    if (a_nu.n) return;
    a_nu.alloc(mza, mwa); a_lqpart.alloc(mza, mwa);
    a_lsct.alloc(mza, mwa);
    for (auto &g : a_lg) g.alloc(mza, mwa);
  }
  // Transpose of update_scalar (+ implicit column) for fixed winds.  Derivation:
  // q_new = q0 + D (sct + volti F), F linear in (q, q0) with upwind choices frozen;
  // see the formulae in the comments below.  lq = adjoint of the stage input q,
  // lq0 accumulates the adjoint of scp0.
  void update_scalar_adj(const F2<R> &mu_in, F2<R> &lq, F2<R> &lq0, const F2<R> &rhosf, R dtr,
                         int iorder, int istage, int nrk, const R *dtlr) {
    const F2<R> &vm = tape_vmasc, &wm = tape_wmasc;
    alloc_adj_scratch();
    F2<R> &nu = a_nu, &lqpart = a_lqpart, *lg = a_lg;
#pragma omp parallel for
    for (int iw = 1; iw <= mwa; ++iw)
      for (int k = 1; k <= mza; ++k) {
        nu(k, iw) = 0; lqpart(k, iw) = 0;
        for (int t = 0; t < 5; ++t) lg[t](k, iw) = 0;
      }
    // (1) implicit columns: mu -> lambda of the explicit result; (2) nu = D volti mu_e
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw) {
      int kb = 2;
      R mu[130];
      for (int k = 1; k <= mza; ++k) mu[k] = mu_in(k, iw);
      R ff = tape_ffimp[(size_t)(iw - 1) * nrk + istage - 1];
      bool imp = ff > R(0.001);
      R fexpl = R(1) - ff;
      // level 1 copy: explicit q(1) = q(kb); implicit branch leaves q(1) at the explicit value
      R mu1 = mu[1];
      mu[1] = 0;
      if (imp) {
        // u = q0 + T^-1 (e - q0 - c M q0); adjoint: le = T^-T mu, lq0 += mu - le - M^T(c le)
        R wup[130], wdn[130], b1[130], b2a[130], b3a[130], le[130], c[130];
        wup[mza] = 0; wdn[mza] = 0; wup[kb - 1] = 0; wdn[kb - 1] = 0;
        for (int k = kb; k <= mza - 1; ++k) { wup[k] = std::max(wm(k, iw), R(0)); wdn[k] = std::min(wm(k, iw), R(0)); }
        R c1 = ff * dtlr[istage - 1];
        for (int k = kb; k <= mza; ++k) {
          R c2 = c1 * volti(k, iw) / tape_rhos[istage - 1](k, iw);
          b1[k] = -c2 * wup[k - 1]; b2a[k] = R(1) + c2 * (wup[k] - wdn[k - 1]); b3a[k] = c2 * wdn[k];
          c[k] = ff * dtr * volti(k, iw) / tape_rhos[istage - 1](k, iw);
        }
        // transpose system: sub-diagonal of T^T is b3a(k-1), super-diagonal b1(k+1)
        R scr[130], cj[130];
        cj[kb] = R(1) / b2a[kb];
        le[kb] = mu[kb] * cj[kb];
        for (int k = kb + 1; k <= mza; ++k) {
          scr[k] = cj[k - 1] * b1[k];
          cj[k] = R(1) / (b2a[k] - b3a[k - 1] * scr[k]);
          le[k] = cj[k] * (mu[k] - b3a[k - 1] * le[k - 1]);
        }
        for (int k = mza - 1; k >= kb; --k) le[k] = le[k] - scr[k + 1] * le[k + 1];
        for (int k = kb; k <= mza; ++k) lq0(k, iw) = lq0(k, iw) + mu[k] - le[k];
        // - c(k) (min(wm(k),0) dq0(k) + max(wm(k-1),0) dq0(k-1)), dq0(k) = q0(k+1) - q0(k)
        for (int k = kb; k <= mza; ++k) {
          R a = -c[k] * le[k];
          if (k <= mza - 1) { R t = a * std::min(wm(k, iw), R(0)); lq0(k + 1, iw) += t; lq0(k, iw) -= t; }
          if (k - 1 >= kb) { R t = a * std::max(wm(k - 1, iw), R(0)); lq0(k, iw) += t; lq0(k - 1, iw) -= t; }
        }
        le[kb] = le[kb] + mu1;  // q(1) kept the explicit value q_e(kb)
        for (int k = kb; k <= mza; ++k) mu[k] = le[k];
      } else {
        mu[kb] = mu[kb] + mu1;
      }
      // explicit: q = q0 + dtr (sct + volti F)/rhos
      R gsum[130];
      for (int k = 1; k <= mza; ++k) gsum[k] = 0;
      for (int jv = 1; jv <= npoly[iw]; ++jv) {
        int iv = w_iv(jv, iw);
        for (int k = 2; k <= mza; ++k) gsum[k] = gsum[k] + dirv(jv, iw) * vm(k, iv);
      }
      // the tendency sct enters as dtr sct / rhos
      if (a_lsct_on)
        for (int k = kb; k <= mza; ++k) a_lsct(k, iw) = a_lsct(k, iw) + mu[k] * dtr / rhosf(k, iw);
      for (int k = kb; k <= mza; ++k) {
        R v = mu[k] * dtr * volti(k, iw) / rhosf(k, iw);
        nu(k, iw) = v;
        R cc = gsum[k] + fexpl * (wm(k - 1, iw) - wm(k, iw));
        lq0(k, iw) = lq0(k, iw) + mu[k] - v * cc;
      }
      // vertical part: F += fexpl (wm(k-1) upw(k-1) - wm(k) upw(k))
      R lupw[130], lds[130];
      for (int k = 0; k <= mza + 1 && k < 130; ++k) { lupw[k] = 0; lds[k] = 0; }
      for (int k = kb; k <= mza; ++k) {
        lupw[k - 1] = lupw[k - 1] + nu(k, iw) * fexpl * wm(k - 1, iw);
        lupw[k] = lupw[k] - nu(k, iw) * fexpl * wm(k, iw);
      }
      R lqv[130];
      for (int k = 1; k <= mza; ++k) lqv[k] = 0;
      lqv[kb] = lqv[kb] + lupw[kb - 1];
      lqv[mza] = lqv[mza] + lupw[mza];
      for (int k = kb; k <= mza - 1; ++k) {
        if (wm(k, iw) < R(0)) {  // upw(k) = scpb(k+1)
          R l = lupw[k];
          lqv[k + 1] += l;
          lds[k + 1] += l * aw_(k + 1, 1, 1);
          lds[k] += l * aw_(k + 1, 2, 1);
        } else {                 // upw(k) = scpt(k)
          R l = lupw[k];
          lqv[k] += l;
          lds[k] += l * aw_(k, 1, 2);
          lds[k - 1] += l * aw_(k, 2, 2);
        }
      }
      for (int k = kb; k <= mza - 1; ++k) { lqv[k + 1] += lds[k]; lqv[k] -= lds[k]; }
      for (int k = 1; k <= mza; ++k) lqpart(k, iw) = lqv[k];
    }
    // (3) face adjoint fluxes phi = vm (nu(w2) - nu(w1)) routed to the upwind side
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw) {
      for (int jv = 1; jv <= npoly[iw]; ++jv) {
        int iv = w_iv(jv, iw);
        int iw1 = v_iw(1, iv), iw2 = v_iw(2, iv);
        int side = (iw == iw1) ? 1 : 2;
        for (int k = 2; k <= mza; ++k) {
          R phi = vm(k, iv) * (nu(k, iw2) - nu(k, iw1));
          bool up2 = vm(k, iv) < R(0);
          if ((side == 2) == up2) {
            lqpart(k, iw) += phi;
            face_value_adj(lg, iorder, side, iv, k, phi, true);
          }
        }
      }
    }
    // (4) transpose of the reconstruction coefficients (bounded for update_scalar)
#pragma omp parallel for
    for (int iw = 2; iw <= mwa; ++iw)
      for (int k = 1; k <= mza; ++k) lq(k, iw) = lqpart(k, iw);
    recon_coefs_adj(lg, lq, iorder, true);
  }

  void scalar_transport_rk_adj() override {
    if (!have_tape) { err = "adjoint called before a forward tracer step"; throw 1; }
    const int nrk = prm.nrk_scal;
    R dtlr[3];
    stage_dts(dtlr, nrk);
    F2<R> lq, lq0, mu;
    lq.alloc(mza, mwa); lq0.alloc(mza, mwa);
    for (size_t n = 0; n < lam.size(); ++n) {
      const bool sgs = turb_on && sgsmix[n];
      if (sgs && tape_lim[n]) {
        err = "adjoint of tracer horizontal diffusion with a final-stage limiter is not built";
        throw 1;
      }
      alloc_adj_scratch();
      a_lsct_on = sgs;
      if (sgs) std::fill(a_lsct.v.begin(), a_lsct.v.end(), R(0));
      mu = lam[n];
      std::fill(lq0.v.begin(), lq0.v.end(), R(0));
      for (int istage = nrk; istage >= 1; --istage) {
        std::fill(lq.v.begin(), lq.v.end(), R(0));
        if (istage == nrk && tape_lim[n])
          update_scalar_lim_adj(n, mu, lq, lq0, tape_rhos[nrk - 1], dtlr[istage - 1], prm.scal_horiz_adv_order,
                                istage, nrk, dtlr);
        else
          update_scalar_adj(mu, lq, lq0, tape_rhos[nrk - 1], dtlr[istage - 1], prm.scal_horiz_adv_order,
                            istage, nrk, dtlr);
        mu = lq;
      }
      // stage 1 read q = q0: both adjoints land on the initial tracer
      for (size_t i = 0; i < lq0.v.size(); ++i) lam[n].v[i] = lq0.v[i] + mu.v[i];
      // the diffusion tendency was built from the initial tracer
      if (sgs) tracer_sgsmix_adj(a_lsct, lam[n]);
    }
  }

  // ---------------------------------------------------------------- mass sums
  // compute_mass_sums, check_nans.F90:207-262 (FP64 accumulation)
  void mass_sums(double *out) override {
    double dry = 0;
    std::vector<double> tr(scp.size(), 0.0);
    for (int iw = 2; iw <= mwa; ++iw)
      for (int k = 2; k <= mza; ++k) {
        double m = (double)rho(k, iw) * volt(k, iw);
        dry += m;
        for (size_t n = 0; n < scp.size(); ++n) tr[n] += m * (double)scp[n](k, iw);
      }
    out[0] = dry;
    for (size_t n = 0; n < scp.size(); ++n) out[1 + n] = tr[n];
  }
  void cfl_report(double *o) override { for (int i = 0; i < 4; ++i) o[i] = cfl_out[i]; }
  // tracer_maxmin, check_nans.F90:164-203: smax(1) = max over the prognostic levels,
  // smax(2) = max of -var_p, from -huge(1.); out[2n] = smax(1), out[2n+1] = -smax(2)
  void tracer_maxmin(float *out) override {
    for (size_t n = 0; n < scp.size(); ++n) {
      float s1 = -FLT_MAX, s2 = -FLT_MAX;
      for (int iw = 2; iw <= mwa; ++iw)
        for (int k = 2; k <= mza; ++k) {
          const float v = (float)scp[n](k, iw);
          s1 = std::max(s1, v);
          s2 = std::max(s2, -v);
        }
      out[2 * n] = s1;
      out[2 * n + 1] = -s2;
    }
  }
};

}  // namespace

// ---------------------------------------------------------------------- C API
extern "C" {

struct oracle_ctx { Base *m; std::string err; };

int oracle_create(const olamgpu_mesh *mesh, const olamgpu_params *p, int32_t ref_types, void **out) {
  auto *c = new oracle_ctx;
  if (mesh->mza > 128) { delete c; return OLAMGPU_EINVAL; }
  const bool q3 = p->thil_horiz_adv_order == 3 || p->wind_horiz_adv_order == 3 || p->scal_horiz_adv_order == 3;
  if (q3 && !(mesh->xy_h && mesh->xy_hu && mesh->xx0_v && mesh->xy0_v && mesh->yy0_v && mesh->xx0_vu &&
              mesh->xy0_vu && mesh->yy0_vu)) { delete c; return OLAMGPU_EINVAL; }
  const bool rayf = p->rayf_fact > 1.e-7f || p->rayfw_fact > 1.e-7f || p->rayfdiv_fact > 1.e-7f ||
                    p->rayfmix_fact > 1.e-7f;
  if (rayf && !(mesh->zm && mesh->zt)) { delete c; return OLAMGPU_EINVAL; }
  if (p->rayf_fact > 1.e-7f && p->initial == 1 && !(mesh->th01d && mesh->dn01d)) { delete c; return OLAMGPU_EINVAL; }
  if (ref_types) c->m = new Model<Prec<float, double, float>>(mesh, p);
  else if (p->precision == OLAMGPU_PREC_F64) c->m = new Model<Prec<double, double, double>>(mesh, p);
  else if (p->precision == OLAMGPU_PREC_MIXED) c->m = new Model<Prec<float, double, double>>(mesh, p);
  else c->m = new Model<Prec<float, float, double>>(mesh, p);
  *out = c;
  return 0;
}
int oracle_destroy(void *h) { auto *c = (oracle_ctx *)h; delete c->m; delete c; return 0; }
int oracle_upload_state(void *h, const olamgpu_state *s) { return ((oracle_ctx *)h)->m->upload_state(s); }
int oracle_download_state(void *h, olamgpu_state *s) { return ((oracle_ctx *)h)->m->download_state(s); }
int oracle_upload_scalar(void *h, int32_t n, const float *p, const float *t) { return ((oracle_ctx *)h)->m->upload_scalar(n, p, t); }
int oracle_download_scalar(void *h, int32_t n, float *p) { return ((oracle_ctx *)h)->m->download_scalar(n, p); }
int oracle_upload_scalar_adj(void *h, int32_t n, const float *p) { return ((oracle_ctx *)h)->m->upload_adj(n, p); }
int oracle_download_scalar_adj(void *h, int32_t n, float *p) { return ((oracle_ctx *)h)->m->download_adj(n, p); }
int oracle_tend0(void *h) { ((oracle_ctx *)h)->m->tend0(); return 0; }
int oracle_rayf_profiles(void *h, float *out, int32_t *kb) { ((oracle_ctx *)h)->m->rayf_profiles(out, kb); return 0; }
int oracle_setup_tendencies(void *h, double *out) { return ((oracle_ctx *)h)->m->setup_tendencies(out); }
int oracle_set_nudging(void *h, const float *wmanud, const float *vmanud, const double *rhot_nud) {
  return ((oracle_ctx *)h)->m->set_nudging(wmanud, vmanud, rhot_nud);
}
int oracle_set_turbulence(void *h, const float *vkh, const float *vkm, const int32_t *khtop, const int32_t *kmtop,
                          const float *akm_sfc) {
  return ((oracle_ctx *)h)->m->set_turbulence(vkh, vkm, khtop, kmtop, akm_sfc);
}
int oracle_comp_alpha_press(void *h, const float *rr_v, const float *rr_w, const float *theta) {
  ((oracle_ctx *)h)->m->comp_alpha_press(rr_v, rr_w, theta);
  return 0;
}
int oracle_init_wrtv_rk(void *h) { ((oracle_ctx *)h)->m->init_wrtv_rk(); return 0; }
int oracle_prog_wrtv_rk(void *h) {
  Base *m = ((oracle_ctx *)h)->m;
  m->err.clear();
  m->prog_wrtv_rk();
  return m->err.empty() ? 0 : OLAMGPU_EINVAL;
}
int oracle_timeavg_momsc(void *h) { ((oracle_ctx *)h)->m->timeavg_momsc(); return 0; }
int oracle_scalar_transport_rk(void *h, const float *rho_old) { ((oracle_ctx *)h)->m->scalar_transport_rk(rho_old); return 0; }
int oracle_scalar_transport_rk_adj(void *h) {
  try { ((oracle_ctx *)h)->m->scalar_transport_rk_adj(); } catch (...) { return OLAMGPU_ESTATE; }
  return 0;
}
int oracle_mass_sums(void *h, double *out) { ((oracle_ctx *)h)->m->mass_sums(out); return 0; }
int oracle_tracer_maxmin(void *h, float *out) { ((oracle_ctx *)h)->m->tracer_maxmin(out); return 0; }
int oracle_cfl_report(void *h, double *out) { ((oracle_ctx *)h)->m->cfl_report(out); return 0; }
int oracle_set_scalar_pdef(void *h, int32_t n, int32_t pdef) { return ((oracle_ctx *)h)->m->set_pdef(n, pdef); }
int oracle_set_scalar_sgsmix(void *h, int32_t n, int32_t v) { return ((oracle_ctx *)h)->m->set_sgsmix(n, v); }
int oracle_set_frozen_limiter(void *h, int32_t on) { ((oracle_ctx *)h)->m->set_frozen(on); return 0; }
int oracle_tape_scales(void *h, int32_t n, int32_t which, double *out) { return ((oracle_ctx *)h)->m->tape_scales(n, which, out); }
int oracle_rhothil_sum(void *h, double *out) { *out = ((oracle_ctx *)h)->m->rhothil_sum(); return 0; }
int oracle_recon(void *h, const float *field, int32_t order, int32_t bounded_coef, int32_t bounded_off, double *coefs,
                 double *faces) {
  return ((oracle_ctx *)h)->m->recon(field, order, bounded_coef, bounded_off, coefs, faces);
}
// Test hook (pins only): plant a deliberate mistake.  1 Coriolis sign; 2 sign of the
// pressure difference of prog_v_begs; 3 sign of the face fluxes of prog_wrt_begs; 4
// order-2 face values with the other side's offsets; 5 sign of the divergence damping;
// 6 divergence damping coefficient with exponent level-1; 7 sign of the vorticity
// damping; 8 first RK stage of prog_wrtv_rk with weight 1/2; 9 first RK stage of
// scalar_transport_rk with weight 1/2.
int oracle_plant(void *h, int32_t code) { ((oracle_ctx *)h)->m->plant = code; return 0; }
int oracle_set_threads(int32_t n) { if (n > 0) omp_set_num_threads(n); return omp_get_max_threads(); }

// tridiffo (tridiag_lib.f90:10-36) on arrays indexed 1..n, for the unit test of the
// column solve: a = sub-diagonal, b = diagonal, c = super-diagonal.
int oracle_tridiffo(int32_t n, const double *a, const double *b, const double *c, const double *d, double *x) {
  std::vector<double> A(n + 2), B(n + 2), Cc(n + 2), D(n + 2), X(n + 2);
  for (int i = 0; i < n; ++i) { A[i + 1] = a[i]; B[i + 1] = b[i]; Cc[i + 1] = c[i]; D[i + 1] = d[i]; }
  tridiffo_t<double>(1, n, A.data(), B.data(), Cc.data(), D.data(), X.data());
  for (int i = 0; i < n; ++i) x[i] = X[i + 1];
  return 0;
}

}  // extern "C"

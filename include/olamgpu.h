/*
 * olamgpu.h -- C-ABI of the MI355X implementation of OLAM's RK3 dynamical core.
 *
 * The reference exposes this path as two Fortran module procedures with no explicit
 * data arguments (state lives in module globals):
 *
 *   module wrtv_rk            omodel/prog_wrtv_rk.f90:1
 *     init_wrtv_rk()          omodel/prog_wrtv_rk.f90:36   (caller olam_run.f90:566)
 *     prog_wrtv_rk()          omodel/prog_wrtv_rk.f90:58   (caller timestep.f90:205)
 *   module scalar_transport   omodel/scalar_transport_rk.F90:1
 *     scalar_transport_rk(rho_old)  scalar_transport_rk.F90:30 (caller timestep.f90:254)
 *
 * plus the glue those calls rely on: tend0 (timestep.f90:430), timeavg_momsc
 * (timestep.f90:563) and compute_mass_sums (outils/check_nans.F90:207).  Every entry
 * point below names the reference routine it replaces.  A Fortran host reaches them
 * through the ISO_C_BINDING interfaces in olam_amd/fortran/olamgpu_iface.f90.
 *
 * Conventions
 *  - All host arrays use the reference's Fortran layout: column-major (mza, n) with the
 *    vertical index k contiguous, horizontal index 1 a dummy (so n = mwa, mva or mma),
 *    and neighbour tables holding 1-based Fortran indices.  The library copies what it
 *    needs to device memory in create/upload and never keeps a host pointer past a call.
 *  - Every call returns 0 on success or a negative OLAMGPU_E* code; the message of the
 *    last failure is returned by olamgpu_last_error().  Nothing throws or aborts across
 *    the ABI.
 *  - Calls are ordered on the context's HIP stream and return after the stream is
 *    synchronised.  One host thread per context.
 *  - Only flat terrain (lpw = lpv = lpm = 2 everywhere) and 4 <= mza <= 128 are
 *    supported (columns of more than 64 levels run as two wavefronts per column, with
 *    every option); horizontal advection orders 1, 2 and 3 are built (order 3 needs the
 *    mem_adv least-squares arrays in the mesh descriptor).  create() rejects anything
 *    else with OLAMGPU_EINVAL.
 */
#ifndef OLAMGPU_H
#define OLAMGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define OLAMGPU_OK        0
#define OLAMGPU_EINVAL   -1   /* invalid argument or unsupported configuration */
#define OLAMGPU_EHIP     -2   /* a HIP runtime call failed */
#define OLAMGPU_ENOMEM   -3   /* device allocation failed */
#define OLAMGPU_ESTATE   -4   /* call out of order (e.g. adjoint without a forward tape) */

/* arithmetic modes (create-time) */
#define OLAMGPU_PREC_F32    0  /* FP32 stencils and state; FP64 column solve and mass sums */
#define OLAMGPU_PREC_MIXED  1  /* as F32, but rho/press kept in FP64 as the reference does */
#define OLAMGPU_PREC_F64    2  /* everything FP64 (inverse volumes rebuilt from volt)     */

typedef struct olamgpu_ctx olamgpu_ctx;

/* Grid description: the fields of itab_w/itab_v/itab_m (modules/mem_ijtabs.f90:45-114)
 * and of mem_grid/mem_adv that the path reads, in Fortran layout. */
typedef struct olamgpu_mesh {
  int32_t mza, mwa, mva, mma;   /* Fortran sizes including the dummy index 1 */
  /* itab_w: (mwa) and (7, mwa) */
  const int32_t *w_npoly;
  const int32_t *w_iv, *w_iw, *w_im;
  const float   *w_dirv;
  const float   *w_ecvec_vx, *w_ecvec_vy, *w_ecvec_vz;
  const float   *gxps_coef, *gyps_coef;      /* mem_grid.f90:111-113 */
  const float   *xx_yy;                      /* mem_adv.f90:121-131   */
  /* itab_v: (2, mva) */
  const int32_t *v_iw, *v_im;
  const float   *v_dxps, *v_dyps;
  /* itab_m: (3, mma) */
  const int32_t *m_iv, *m_iw, *m_im;
  const float   *xx_yy_m;                    /* (3, mma) mem_adv.f90:137-147 */
  /* lowest prognostic levels (must all be 2) */
  const int32_t *lpw, *lpv, *lpm;
  /* per-point metrics */
  const float *wnx, *wny, *wnz, *arw0;       /* (mwa) */
  const float *vnx, *vny, *vnz;              /* (mva) */
  const float *dnu, *dniu, *dnv, *dniv;      /* (mva) */
  const float *arm0;                         /* (mma) */
  /* per-level (mza) */
  const float  *dzim, *dzit, *zfacit;
  const double *zwgt_top8, *zwgt_bot8, *gdz_wgtm8, *gdz_wgtp8;
  const float  *a_w, *a_wu;                  /* (mza, 2, 2) mem_adv.f90:58-117 */
  /* control volumes (mza, n) */
  const float  *arw, *volti, *volwi;         /* (mza, mwa) */
  const double *volt;                        /* (mza, mwa) FP64 as mem_grid.f90:66 */
  const float  *arv, *volvi;                 /* (mza, mva) */
  /* horizontal quadratic reconstruction (order 3), mem_adv.f90:11-17,161-354.  Read
   * only when an *_horiz_adv_order is 3; may be NULL otherwise. */
  const float  *xy_h, *xy_hu;                /* (5, 7, mwa) bounded / unbounded LSQ   */
  const float  *xx0_v, *xy0_v, *yy0_v;       /* (2, mva) face offsets, bounded        */
  const float  *xx0_vu, *xy0_vu, *yy0_vu;    /* (2, mva) face offsets, unbounded      */
  /* Rayleigh friction near the model top (mem_rayf.f90:41-173).  Read only when one of
   * the rayf*_fact parameters exceeds 1e-7; may be NULL otherwise. */
  const float  *zm, *zt;                     /* (mza) heights of M and T levels, mem_grid */
  const float  *th01d, *dn01d;               /* (mza) reference sounding, misc_coms.f90:112-113;
                                                read when rayf_fact > 0 and initial == 1 */
  const float  *dzit_bot;                    /* (mza) 1 / (zt(k) - zm(k-1)), mem_grid.f90:374-378;
                                                read by olamgpu_set_turbulence (surface drag) */
} olamgpu_mesh;

/* Namelist options and time steps the path reads (oname_coms.f90:190-238,
 * misc_coms.f90:66-107, consts_coms.f90:159-163). */
typedef struct olamgpu_params {
  int32_t precision;                 /* OLAMGPU_PREC_* */
  int32_t num_scalar;                /* tracers carried by scalar_transport_rk */
  int32_t nrk_wrtv;                  /* acoust_timestep_level: 2 or 3 */
  int32_t nrk_scal;                  /* scalar_timestep_level: 2 or 3 */
  int32_t thil_horiz_adv_order;      /* 1, 2 or 3 */
  int32_t wind_horiz_adv_order;      /* 1, 2 or 3 */
  int32_t scal_horiz_adv_order;      /* 1, 2 or 3 */
  int32_t iscal_monot;               /* 0 none, 1 monotone, 2 positive definite (final stage) */
  int32_t divh_damp_level;           /* 0 disables; 1..3 */
  int32_t vort_damp_level;           /* 0 disables; 1..3 */
  float   divh_damp_fact;
  float   vort_damp_fact;
  double  dtsm;                      /* acoustic (short) step [s] */
  double  dtlm;                      /* long step [s] */
  int32_t nacoust;                   /* acoustic steps per long step */
  int32_t adjoint_tape;              /* >0: record forward tracer tapes for the adjoint */
  double  fcoriol;                   /* 2 Omega on the global grid */
  int32_t ithil_monot;               /* >0: monotone thil fluxes on the final acoustic RK stage */
  int32_t iwind_monot;               /* >0: monotone wind fluxes on the final acoustic RK stage */
  /* Rayleigh friction (oname_coms.f90:224-238, mem_rayf.f90:41-173): each kind is on
   * when its fact exceeds 1e-7 and some level lies above its zmin.  rayf_* damps thil
   * and vmc toward the initial state and needs initial == 1 (horizontally homogeneous
   * initialization); rayfw_* damps W through the implicit solve (prog_wrtv_rk.f90:46-52);
   * rayfdiv_* adds a second-order divergence damping profile (vort_div_damp.f90:37-121);
   * rayfmix_* mixes vxe/vye/vze vertically (rayf_mix_top_vxe, mem_rayf.f90:177-220). */
  int32_t initial;                   /* misc_coms initial: 1 = horizontally homogeneous */
  float   rayf_zmin, rayf_fact, rayf_expon;
  float   rayfw_zmin, rayfw_fact, rayfw_expon;
  float   rayfdiv_zmin, rayfdiv_fact, rayfdiv_expon;
  float   rayfmix_zmin, rayfmix_fact, rayfmix_expon;
  float   akmin;                     /* akmin(1), oname_coms.f90:266: minimum horizontal eddy
                                        diffusivity ratio of comp_horiz_k */
  /* nudging (oname_coms.f90:134-140): with nudflag > 0 and nud_preserve_total_mass the
   * nudging mass fluxes vmanud/wmanud join the face and vertical mass fluxes
   * (prog_wrtv_rk.f90:347-355,782-788; with nudflag == 1 also the vertical nudging
   * fluxes of :605-621), otherwise the density tendency rhot_nud enters the explicit
   * density (:703-706).  The fields are set by olamgpu_set_nudging (zero until then). */
  int32_t nudflag;
  int32_t nud_preserve_total_mass;
  /* nud_preserve_mix_ratio (oname_coms.f90:139, default true): with nudflag > 0 the
   * tracer step's final-stage density is rho - dtlm * rhot_nud (scalar_transport_rk.F90:
   * 120-133), so the nudging of the density does not change the mixing ratios.  As
   * oname_check.f90:218 does, it is ignored while nud_preserve_total_mass is on. */
  int32_t nud_preserve_mix_ratio;
} olamgpu_params;

/* Prognostic and diagnostic fields of mem_basic/mem_tend (mem_basic.f90:8-34,
 * mem_tend.f90:5-11) in Fortran layout.  Any pointer may be NULL to skip it. */
typedef struct olamgpu_state {
  float  *thil, *wmc, *wc, *vxe, *vye, *vze;   /* (mza, mwa) */
  double *rho, *press;                          /* (mza, mwa) real(r8) */
  float  *vmc, *vc;                             /* (mza, mva) */
  float  *wmasc;                                /* (mza, mwa) */
  float  *vmasc;                                /* (mza, mva) */
  float  *alpha_press, *pwfac;                  /* (mza, mwa) */
  float  *pvfac;                                /* (mza, mva) */
  float  *thilt, *vmxet, *vmyet, *vmzet;        /* (mza, mwa) long-step tendencies */
  float  *vmt;                                  /* (mza, mva) */
} olamgpu_state;

/* ---- lifecycle ------------------------------------------------------------ */

/* Copies the grid to device memory, renumbers W/V/M points for locality and
 * checks the renumbered tables against the originals element by element.
 * Replaces the allocation side of alloc_basic/alloc_grid_other/alloc_adv for the
 * fields the path reads (modules/mem_basic.f90:53, modules/mem_grid.f90:464,
 * modules/mem_adv.f90:23).  device < 0 uses the current device. */
int olamgpu_create(const olamgpu_mesh *mesh, const olamgpu_params *params,
                   int32_t device, olamgpu_ctx **out);
int olamgpu_destroy(olamgpu_ctx *ctx);
const char *olamgpu_last_error(const olamgpu_ctx *ctx);   /* ctx may be NULL */

/* ---- state transfer ------------------------------------------------------- */

/* Host <-> device copies of mem_basic/mem_tend fields (NULL members skipped). */
int olamgpu_upload_state(olamgpu_ctx *ctx, const olamgpu_state *host);
int olamgpu_download_state(olamgpu_ctx *ctx, olamgpu_state *host);
/* scalar_tab(n)%var_p and %var_t (var_tables.f90:53-69), n = 1..num_scalar.
 * var_t may be NULL (treated as zero tendency). */
int olamgpu_upload_scalar(olamgpu_ctx *ctx, int32_t n, const float *var_p, const float *var_t);
int olamgpu_download_scalar(olamgpu_ctx *ctx, int32_t n, float *var_p);
/* scalar_tab(n)%pdef (var_tables.f90:58, default true): with iscal_monot = 2 only pdef
 * tracers are limited, and with iscal_monot > 0 only pdef tracers get the low-order
 * tendency update of scp0 before the final stage (scalar_transport_rk.F90:297-337). */
int olamgpu_set_scalar_pdef(olamgpu_ctx *ctx, int32_t n, int32_t pdef);
/* scalar_tab(n)%do_sgsmix (var_tables.f90:59, default true): once olamgpu_set_turbulence
 * has supplied the horizontal diffusivities, each such tracer's tendency gets the
 * horizontal diffusion of scalar_transport_rk.F90:244-274 at the start of every tracer
 * step (built from the tracer before the step, added to var_t).  The adjoint includes
 * its transpose; the adjoint of a diffused tracer that also has a final-stage limiter
 * is not built and fails with OLAMGPU_EINVAL. */
int olamgpu_set_scalar_sgsmix(olamgpu_ctx *ctx, int32_t n, int32_t do_sgsmix);

/* ---- the hot path ---------------------------------------------------------- */

/* rayf_init (mem_rayf.f90:158-171) and init_wrtv_rk (prog_wrtv_rk.f90:36-54), called
 * where olam_run.f90:555,566 call them: after the initial state has been uploaded and
 * before the first prog_wrtv_rk.  With rayf_fact > 0 and initial == 1 it keeps the
 * current vc as the wind vc03d that the friction relaxes toward; prog_wrtv_rk then
 * fails with OLAMGPU_EINVAL until it has been called.  The friction profiles and the
 * W factor b20 are built by olamgpu_create from the params. */
int olamgpu_init_wrtv_rk(olamgpu_ctx *ctx);

/* tend0 (timestep.f90:430-496): zero vmasc/wmasc and the tracer tendencies, save
 * rho_old = real(rho).  The momentum/thil tendencies are left to the host (physics). */
int olamgpu_tend0(olamgpu_ctx *ctx);

/* prog_wrtv_rk (prog_wrtv_rk.f90:58-488): one acoustic RK step of W/V momentum,
 * rho, rho*thil and press, with divergence/vorticity damping (vort_div_damp.f90)
 * and vel_t3d_hex (vel_t3d.f90:39-108).  Accumulates vmasc/wmasc. */
int olamgpu_prog_wrtv_rk(olamgpu_ctx *ctx);

/* timeavg_momsc (timestep.f90:563-612): vmasc, wmasc /= nacoust when nacoust > 1. */
int olamgpu_timeavg_momsc(olamgpu_ctx *ctx);

/* comp_alpha_press (timestep.f90:500-559), for a moist atmosphere (OLAM calls it only
 * when miclevel > 0): from the host's vapor and total water mixing ratios rr_v, rr_w
 * and potential temperature theta ((mza, mwa) FP32 each, Fortran layout) and the
 * context's thil, sets alpha_press and pwfac at W and pvfac at V on the device, the
 * factors prog_wrtv_rk reads (olamgpu_upload_state can set them instead). */
int olamgpu_comp_alpha_press(olamgpu_ctx *ctx, const float *rr_v, const float *rr_w, const float *theta);

/* The turbulence fields of the host's PBL scheme (mem_turb: eddy diffusivities vkh,
 * vkm (mza, mwa) at W levels, the top levels khtop, kmtop (mwa) of heat and momentum
 * diffusion, and akm_sfc (mwa), the surface drag coefficient of the one surface level
 * of flat terrain), followed by comp_horiz_k (pbl_driver.f90:136-238) on the device.
 * Each later prog_wrtv_rk then applies, in its setup loop, the implicit vertical eddy
 * diffusion of solve_eddy_diff_heat and solve_eddy_diff_vxe (pbl_driver.f90:519-742)
 * and the horizontal diffusion of prog_wrtv_rk.f90:235-251.  OLAM's pbl_driver sets
 * these once per long step; call this where it does.  Before the first call the
 * diffusion is off.  khtop and kmtop must lie in [0, mza - 1]; the non-local PBL term
 * of idiffk = 1 is not part of this entry. */
int olamgpu_set_turbulence(olamgpu_ctx *ctx, const float *vkh, const float *vkm, const int32_t *khtop,
                           const int32_t *kmtop, const float *akm_sfc);

/* The nudging fields of mem_nudge (oisan/mem_nudge.f90:46,90-91): the mass fluxes
 * wmanud (mza, mwa) and vmanud (mza, mva) FP32, and the density tendency rhot_nud
 * (mza, mwa) FP64.  Each may be NULL (left as it was; all start at zero, as the
 * reference allocates them).  Set them where OLAM's nudging updates them, once per
 * long step before the acoustic steps. */
int olamgpu_set_nudging(olamgpu_ctx *ctx, const float *wmanud, const float *vmanud, const double *rhot_nud);

/* scalar_transport_rk(rho_old) (scalar_transport_rk.F90:30-385): one long step of
 * all registered tracers.  rho_old = NULL uses the copy saved by olamgpu_tend0;
 * otherwise a host (mza, mwa) FP32 array. */
int olamgpu_scalar_transport_rk(olamgpu_ctx *ctx, const float *rho_old);

/* ---- adjoint (no counterpart in the reference) ------------------------------ */

/* Adjoint of the last olamgpu_scalar_transport_rk call with the upwind choices
 * and limiter scales frozen from that forward pass (params.adjoint_tape > 0 records
 * the scales of a limited final stage).
 * On entry the device adjoint buffers hold dJ/dq(after the step); on return they
 * hold dJ/dq(before the step).  Buffers are set/read with the calls below. */
int olamgpu_scalar_transport_rk_adj(olamgpu_ctx *ctx);
int olamgpu_upload_scalar_adj(olamgpu_ctx *ctx, int32_t n, const float *lambda);
int olamgpu_download_scalar_adj(olamgpu_ctx *ctx, int32_t n, float *lambda);

/* ---- diagnostics ------------------------------------------------------------ */

/* compute_mass_sums (check_nans.F90:207-332): out[0] = sum rho*volt (dry mass),
 * out[1+n] = sum rho*volt*q_n for n = 0..num_scalar-1, all FP64 accumulations. */
int olamgpu_mass_sums(olamgpu_ctx *ctx, double *out);

/* No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The sums of olamgpu_mass_sums, written to out_dev, a DEVICE array of 1 + num_scalar
 * doubles, with no allocation and no host copy: the partial sums live in a static
 * device array of 1024 x 65 doubles (520 KiB per GPU, held until the process ends), so
 * num_scalar may be at most 64.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (from several threads or contexts) race.  Returns after the
 * context's stream is synchronised, like every entry point. */
int olamgpu_mass_sums_noalloc(olamgpu_ctx *ctx, double *out_dev);

/* tracer_maxmin (check_nans.F90:164-203): out[2n] and out[2n+1] = the maximum and the
 * minimum of tracer n+1 (n = 0..num_scalar-1) over levels lpw..mza of every column, FP32
 * as var_p; NaN values are skipped.  The reference prints them; here they are returned. */
int olamgpu_tracer_maxmin(olamgpu_ctx *ctx, float *out);

/* check_cfls diagnostics (scalar_transport_rk.F90:1476-1712) of the last tracer
 * step: out[0] = max total CFL, out[1] = max horizontal CFL, out[2] = max |wc|,
 * out[3] = number of implicit columns (any stage). */
int olamgpu_cfl_report(olamgpu_ctx *ctx, double *out);

/* ---- device-resident access (for harnesses that keep data on the GPU) ------- */

/* Field identifiers for olamgpu_copy_field. */
#define OLAMGPU_F_SCALAR      1   /* tracer n (index = n) */
#define OLAMGPU_F_SCALAR_ADJ  2   /* tracer adjoint n */
#define OLAMGPU_F_RHO         3
#define OLAMGPU_F_RHO_OLD     4
#define OLAMGPU_F_VMASC       5
#define OLAMGPU_F_WMASC       6
#define OLAMGPU_F_THIL        7
#define OLAMGPU_F_VMC         8
#define OLAMGPU_F_WMC         9
#define OLAMGPU_F_PRESS      10

/* Device-to-device copy between a library field and a caller buffer laid out as a
 * dense (n, mza) row-major array over real points only (no dummy index, original
 * point order), element type float (or double when the field is FP64 in the
 * context's precision mode; *elem_bytes reports it).  direction: 0 = field -> buf,
 * 1 = buf -> field.  buf is a device pointer. */
int olamgpu_copy_field(olamgpu_ctx *ctx, int32_t field_id, int32_t index,
                       void *dev_buf, int32_t direction, int32_t *elem_bytes);

/* ---- checkpoints for multi-step adjoint windows (no counterpart in the reference) --- */

#define OLAMGPU_CKPT_STATE  1   /* fields the hot-path calls modify (prognostic state,
                                   vmasc, wmasc, rho_old); host-set inputs excluded */
#define OLAMGPU_CKPT_TAPE   2   /* vmasc, wmasc, rho, rho_old: what a tracer step and its
                                   adjoint read besides the tracers (with iscal_monot > 0
                                   also the taped limiter scales; STATE then also holds
                                   the tracers) */

/* Size in bytes of a checkpoint of the given kind. */
int olamgpu_checkpoint_bytes(olamgpu_ctx *ctx, int32_t kind, int64_t *bytes);
/* Device-to-device copy of the context's state into / out of a caller-owned device
 * buffer of checkpoint_bytes bytes.  Restoring a STATE checkpoint makes the following
 * calls reproduce bit-identically what followed the save (given the same host-set
 * inputs and tracers); restoring a TAPE makes
 * olamgpu_scalar_transport_rk_adj valid for the tracer step that the tape was saved
 * after (the adjoint depends on the tracers only linearly). */
int olamgpu_checkpoint_save(olamgpu_ctx *ctx, int32_t kind, void *dev_buf);
int olamgpu_checkpoint_restore(olamgpu_ctx *ctx, int32_t kind, const void *dev_buf);

/* Synchronise the context's stream and report the time of the last call's
 * kernels in milliseconds (HIP events around the call). */
int olamgpu_last_call_ms(olamgpu_ctx *ctx, double *ms);

/* ---- kernel timing (no counterpart in the reference) ------------------------ */

/* on != 0: record HIP events around every kernel launch on the context's stream
 * and accumulate per-kernel counts and times; resets the counters either way. */
int olamgpu_set_profiling(olamgpu_ctx *ctx, int32_t on);
/* Writes one line "kernel_name launches total_ms" per kernel into text[len]. */
int olamgpu_kernel_profile(olamgpu_ctx *ctx, char *text, int32_t len);

#ifdef __cplusplus
}
#endif
#endif /* OLAMGPU_H */

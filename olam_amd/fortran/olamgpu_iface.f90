! olamgpu_iface.f90 -- ISO_C_BINDING interfaces to libolamgpu (include/olamgpu.h).
!
! A Fortran host (OLAM's model driver) uses this module to call the MI355X dynamical
! core in place of the module procedures it replaces:
!   prog_wrtv_rk()              omodel/prog_wrtv_rk.f90:58  (called at timestep.f90:205)
!   scalar_transport_rk(rho_old) omodel/scalar_transport_rk.F90:30 (timestep.f90:254)
!   tend0 / timeavg_momsc       omodel/timestep.f90:430 / :563
!   compute_mass_sums           outils/check_nans.F90:207
! The derived types below mirror the C structs member by member (bind(C) gives them
! the C layout).  Pointer members are filled with c_loc() of the host's arrays, which
! must be contiguous and keep the reference's (mza, n) layout with the dummy index 1.

module olamgpu_iface
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: OLAMGPU_OK = 0, OLAMGPU_EINVAL = -1, &
       OLAMGPU_EHIP = -2, OLAMGPU_ENOMEM = -3, OLAMGPU_ESTATE = -4
  integer(c_int32_t), parameter, public :: OLAMGPU_PREC_F32 = 0, OLAMGPU_PREC_MIXED = 1, &
       OLAMGPU_PREC_F64 = 2
  integer(c_int32_t), parameter, public :: OLAMGPU_CKPT_STATE = 1, OLAMGPU_CKPT_TAPE = 2

  type, bind(C), public :: olamgpu_mesh_t
     integer(c_int32_t) :: mza = 0, mwa = 0, mva = 0, mma = 0
     type(c_ptr) :: w_npoly = c_null_ptr, w_iv = c_null_ptr, w_iw = c_null_ptr, w_im = c_null_ptr
     type(c_ptr) :: w_dirv = c_null_ptr
     type(c_ptr) :: w_ecvec_vx = c_null_ptr, w_ecvec_vy = c_null_ptr, w_ecvec_vz = c_null_ptr
     type(c_ptr) :: gxps_coef = c_null_ptr, gyps_coef = c_null_ptr, xx_yy = c_null_ptr
     type(c_ptr) :: v_iw = c_null_ptr, v_im = c_null_ptr, v_dxps = c_null_ptr, v_dyps = c_null_ptr
     type(c_ptr) :: m_iv = c_null_ptr, m_iw = c_null_ptr, m_im = c_null_ptr, xx_yy_m = c_null_ptr
     type(c_ptr) :: lpw = c_null_ptr, lpv = c_null_ptr, lpm = c_null_ptr
     type(c_ptr) :: wnx = c_null_ptr, wny = c_null_ptr, wnz = c_null_ptr, arw0 = c_null_ptr
     type(c_ptr) :: vnx = c_null_ptr, vny = c_null_ptr, vnz = c_null_ptr
     type(c_ptr) :: dnu = c_null_ptr, dniu = c_null_ptr, dnv = c_null_ptr, dniv = c_null_ptr
     type(c_ptr) :: arm0 = c_null_ptr
     type(c_ptr) :: dzim = c_null_ptr, dzit = c_null_ptr, zfacit = c_null_ptr
     type(c_ptr) :: zwgt_top8 = c_null_ptr, zwgt_bot8 = c_null_ptr
     type(c_ptr) :: gdz_wgtm8 = c_null_ptr, gdz_wgtp8 = c_null_ptr
     type(c_ptr) :: a_w = c_null_ptr, a_wu = c_null_ptr
     type(c_ptr) :: arw = c_null_ptr, volti = c_null_ptr, volwi = c_null_ptr, volt = c_null_ptr
     type(c_ptr) :: arv = c_null_ptr, volvi = c_null_ptr
     ! order-3 reconstruction (mem_adv.f90:11-17): xy_h, xy_hu (5,7,mwa); face offsets (2,mva)
     type(c_ptr) :: xy_h = c_null_ptr, xy_hu = c_null_ptr
     type(c_ptr) :: xx0_v = c_null_ptr, xy0_v = c_null_ptr, yy0_v = c_null_ptr
     type(c_ptr) :: xx0_vu = c_null_ptr, xy0_vu = c_null_ptr, yy0_vu = c_null_ptr
     ! Rayleigh friction (mem_rayf.f90:41-173): zm, zt (mza) of mem_grid; the reference
     ! sounding th01d, dn01d (mza) of misc_coms
     type(c_ptr) :: zm = c_null_ptr, zt = c_null_ptr, th01d = c_null_ptr, dn01d = c_null_ptr
     ! surface drag of the eddy diffusion: dzit_bot (mza) of mem_grid
     type(c_ptr) :: dzit_bot = c_null_ptr
  end type olamgpu_mesh_t

  type, bind(C), public :: olamgpu_params_t
     integer(c_int32_t) :: precision = OLAMGPU_PREC_F32
     integer(c_int32_t) :: num_scalar = 0
     integer(c_int32_t) :: nrk_wrtv = 3, nrk_scal = 3
     integer(c_int32_t) :: thil_horiz_adv_order = 2, wind_horiz_adv_order = 2
     integer(c_int32_t) :: scal_horiz_adv_order = 2
     integer(c_int32_t) :: iscal_monot = 0
     integer(c_int32_t) :: divh_damp_level = 3, vort_damp_level = 3
     real(c_float)      :: divh_damp_fact = 0.1, vort_damp_fact = 0.075
     real(c_double)     :: dtsm = 0.0d0, dtlm = 0.0d0
     integer(c_int32_t) :: nacoust = 1
     integer(c_int32_t) :: adjoint_tape = 0
     real(c_double)     :: fcoriol = 0.0d0
     integer(c_int32_t) :: ithil_monot = 0, iwind_monot = 0
     ! Rayleigh friction namelist values (oname_coms.f90:224-238) and misc_coms initial
     integer(c_int32_t) :: initial = 1
     real(c_float)      :: rayf_zmin = 30000.0, rayf_fact = 0.0, rayf_expon = 1.0
     real(c_float)      :: rayfw_zmin = 30000.0, rayfw_fact = 0.0, rayfw_expon = 1.0
     real(c_float)      :: rayfdiv_zmin = 30000.0, rayfdiv_fact = 0.0, rayfdiv_expon = 1.0
     real(c_float)      :: rayfmix_zmin = 30000.0, rayfmix_fact = 0.0, rayfmix_expon = 1.0
     real(c_float)      :: akmin = 0.0     ! nl%akmin(1), comp_horiz_k
     ! nudging: nl%nudflag, nl%nud_preserve_total_mass, nl%nud_preserve_mix_ratio (1 = .true.)
     integer(c_int32_t) :: nudflag = 0, nud_preserve_total_mass = 1
     integer(c_int32_t) :: nud_preserve_mix_ratio = 1
  end type olamgpu_params_t

  type, bind(C), public :: olamgpu_state_t
     type(c_ptr) :: thil = c_null_ptr, wmc = c_null_ptr, wc = c_null_ptr
     type(c_ptr) :: vxe = c_null_ptr, vye = c_null_ptr, vze = c_null_ptr
     type(c_ptr) :: rho = c_null_ptr, press = c_null_ptr
     type(c_ptr) :: vmc = c_null_ptr, vc = c_null_ptr
     type(c_ptr) :: wmasc = c_null_ptr, vmasc = c_null_ptr
     type(c_ptr) :: alpha_press = c_null_ptr, pwfac = c_null_ptr, pvfac = c_null_ptr
     type(c_ptr) :: thilt = c_null_ptr, vmxet = c_null_ptr, vmyet = c_null_ptr, vmzet = c_null_ptr
     type(c_ptr) :: vmt = c_null_ptr
  end type olamgpu_state_t

  public :: olamgpu_create, olamgpu_destroy, olamgpu_last_error, olamgpu_upload_state, &
       olamgpu_download_state, olamgpu_upload_scalar, olamgpu_download_scalar, &
       olamgpu_init_wrtv_rk, olamgpu_tend0, olamgpu_prog_wrtv_rk, olamgpu_timeavg_momsc, &
       olamgpu_comp_alpha_press, olamgpu_set_turbulence, olamgpu_set_nudging, &
       olamgpu_scalar_transport_rk, olamgpu_scalar_transport_rk_adj, &
       olamgpu_upload_scalar_adj, olamgpu_download_scalar_adj, olamgpu_mass_sums, &
       olamgpu_cfl_report, olamgpu_copy_field, olamgpu_last_call_ms, &
       olamgpu_set_profiling, olamgpu_kernel_profile, olamgpu_checkpoint_bytes, &
       olamgpu_checkpoint_save, olamgpu_checkpoint_restore, olamgpu_set_scalar_pdef, &
       olamgpu_set_scalar_sgsmix, olamgpu_tracer_maxmin, olamgpu_mass_sums_noalloc, &
       olamgpu_error_message

  interface
     integer(c_int) function olamgpu_create(mesh, params, device, ctx) bind(C, name="olamgpu_create")
       import :: c_int, c_int32_t, c_ptr, olamgpu_mesh_t, olamgpu_params_t
       type(olamgpu_mesh_t), intent(in) :: mesh
       type(olamgpu_params_t), intent(in) :: params
       integer(c_int32_t), value :: device
       type(c_ptr), intent(out) :: ctx
     end function olamgpu_create

     integer(c_int) function olamgpu_destroy(ctx) bind(C, name="olamgpu_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function olamgpu_destroy

     type(c_ptr) function olamgpu_last_error(ctx) bind(C, name="olamgpu_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function olamgpu_last_error

     integer(c_int) function olamgpu_upload_state(ctx, host) bind(C, name="olamgpu_upload_state")
       import :: c_int, c_ptr, olamgpu_state_t
       type(c_ptr), value :: ctx
       type(olamgpu_state_t), intent(in) :: host
     end function olamgpu_upload_state

     integer(c_int) function olamgpu_download_state(ctx, host) bind(C, name="olamgpu_download_state")
       import :: c_int, c_ptr, olamgpu_state_t
       type(c_ptr), value :: ctx
       type(olamgpu_state_t), intent(in) :: host
     end function olamgpu_download_state

     integer(c_int) function olamgpu_upload_scalar(ctx, n, var_p, var_t) bind(C, name="olamgpu_upload_scalar")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       type(c_ptr), value :: var_p, var_t
     end function olamgpu_upload_scalar

     integer(c_int) function olamgpu_download_scalar(ctx, n, var_p) bind(C, name="olamgpu_download_scalar")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       type(c_ptr), value :: var_p
     end function olamgpu_download_scalar

     ! scalar_tab(n)%pdef of the limiter options (default true)
     integer(c_int) function olamgpu_set_scalar_pdef(ctx, n, pdef) bind(C, name="olamgpu_set_scalar_pdef")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n, pdef
     end function olamgpu_set_scalar_pdef

     ! no counterpart in the reference: for callers that cannot allocate device memory;
     ! out_dev is a device array of 1 + num_scalar doubles
     integer(c_int) function olamgpu_mass_sums_noalloc(ctx, out_dev) bind(C, name="olamgpu_mass_sums_noalloc")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, out_dev
     end function olamgpu_mass_sums_noalloc

     integer(c_int) function olamgpu_tracer_maxmin(ctx, out) bind(C, name="olamgpu_tracer_maxmin")
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(out) :: out(*)
     end function olamgpu_tracer_maxmin

     integer(c_int) function olamgpu_set_scalar_sgsmix(ctx, n, do_sgsmix) bind(C, name="olamgpu_set_scalar_sgsmix")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n, do_sgsmix
     end function olamgpu_set_scalar_sgsmix

     integer(c_int) function olamgpu_init_wrtv_rk(ctx) bind(C, name="olamgpu_init_wrtv_rk")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function olamgpu_init_wrtv_rk

     integer(c_int) function olamgpu_tend0(ctx) bind(C, name="olamgpu_tend0")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function olamgpu_tend0

     integer(c_int) function olamgpu_prog_wrtv_rk(ctx) bind(C, name="olamgpu_prog_wrtv_rk")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function olamgpu_prog_wrtv_rk

     integer(c_int) function olamgpu_timeavg_momsc(ctx) bind(C, name="olamgpu_timeavg_momsc")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function olamgpu_timeavg_momsc

     ! comp_alpha_press (timestep.f90:500-559): c_loc of real(4) (mza, mwa) rr_v, rr_w, theta
     integer(c_int) function olamgpu_comp_alpha_press(ctx, rr_v, rr_w, theta) &
          bind(C, name="olamgpu_comp_alpha_press")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, rr_v, rr_w, theta
     end function olamgpu_comp_alpha_press

     ! mem_turb fields of the PBL scheme: vkh, vkm real(4) (mza, mwa); khtop, kmtop
     ! integer(4) (mwa); akm_sfc real(4) (mwa); then comp_horiz_k (pbl_driver.f90:136-238)
     integer(c_int) function olamgpu_set_turbulence(ctx, vkh, vkm, khtop, kmtop, akm_sfc) &
          bind(C, name="olamgpu_set_turbulence")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, vkh, vkm, khtop, kmtop, akm_sfc
     end function olamgpu_set_turbulence

     ! mem_nudge fields: wmanud real(4) (mza, mwa), vmanud real(4) (mza, mva), rhot_nud
     ! real(8) (mza, mwa); c_null_ptr leaves a field as it was
     integer(c_int) function olamgpu_set_nudging(ctx, wmanud, vmanud, rhot_nud) &
          bind(C, name="olamgpu_set_nudging")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, wmanud, vmanud, rhot_nud
     end function olamgpu_set_nudging

     ! rho_old: c_loc of a real(4) (mza, mwa) array, or c_null_ptr for the copy saved by tend0
     integer(c_int) function olamgpu_scalar_transport_rk(ctx, rho_old) &
          bind(C, name="olamgpu_scalar_transport_rk")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, rho_old
     end function olamgpu_scalar_transport_rk

     integer(c_int) function olamgpu_scalar_transport_rk_adj(ctx) bind(C, name="olamgpu_scalar_transport_rk_adj")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function olamgpu_scalar_transport_rk_adj

     integer(c_int) function olamgpu_upload_scalar_adj(ctx, n, lambda) bind(C, name="olamgpu_upload_scalar_adj")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       type(c_ptr), value :: lambda
     end function olamgpu_upload_scalar_adj

     integer(c_int) function olamgpu_download_scalar_adj(ctx, n, lambda) &
          bind(C, name="olamgpu_download_scalar_adj")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       type(c_ptr), value :: lambda
     end function olamgpu_download_scalar_adj

     integer(c_int) function olamgpu_mass_sums(ctx, out) bind(C, name="olamgpu_mass_sums")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(*)
     end function olamgpu_mass_sums

     integer(c_int) function olamgpu_cfl_report(ctx, out) bind(C, name="olamgpu_cfl_report")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(4)
     end function olamgpu_cfl_report

     integer(c_int) function olamgpu_copy_field(ctx, field_id, index, dev_buf, direction, elem_bytes) &
          bind(C, name="olamgpu_copy_field")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: field_id, index
       type(c_ptr), value :: dev_buf
       integer(c_int32_t), value :: direction
       integer(c_int32_t), intent(out) :: elem_bytes
     end function olamgpu_copy_field

     integer(c_int) function olamgpu_last_call_ms(ctx, ms) bind(C, name="olamgpu_last_call_ms")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ms
     end function olamgpu_last_call_ms

     integer(c_int) function olamgpu_set_profiling(ctx, on) bind(C, name="olamgpu_set_profiling")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: on
     end function olamgpu_set_profiling

     integer(c_int) function olamgpu_kernel_profile(ctx, text, len) bind(C, name="olamgpu_kernel_profile")
       import :: c_int, c_int32_t, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(out) :: text(*)
       integer(c_int32_t), value :: len
     end function olamgpu_kernel_profile

     integer(c_int) function olamgpu_checkpoint_bytes(ctx, kind, bytes) bind(C, name="olamgpu_checkpoint_bytes")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: kind
       integer(c_int64_t), intent(out) :: bytes
     end function olamgpu_checkpoint_bytes

     ! dev_buf: a device pointer (e.g. from hipMalloc) of checkpoint_bytes bytes
     integer(c_int) function olamgpu_checkpoint_save(ctx, kind, dev_buf) bind(C, name="olamgpu_checkpoint_save")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: kind
       type(c_ptr), value :: dev_buf
     end function olamgpu_checkpoint_save

     integer(c_int) function olamgpu_checkpoint_restore(ctx, kind, dev_buf) &
          bind(C, name="olamgpu_checkpoint_restore")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: kind
       type(c_ptr), value :: dev_buf
     end function olamgpu_checkpoint_restore
  end interface

contains

  ! The message of olamgpu_last_error as a Fortran string (blank when there is none).
  function olamgpu_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=512) :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    p = olamgpu_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [512])
    do i = 1, 512
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    end do
  end function olamgpu_error_message

end module olamgpu_iface

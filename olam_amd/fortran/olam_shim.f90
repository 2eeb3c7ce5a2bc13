! olam_shim.f90 -- the drop-in for OLAM's RK3 dynamical core: modules wrtv_rk (public
! init_wrtv_rk, prog_wrtv_rk; omodel/prog_wrtv_rk.f90:1-58) and scalar_transport (public
! scalar_transport_rk; omodel/scalar_transport_rk.F90:1-51) with the reference's
! procedure names and arguments, running the path on the GPU through the C-ABI
! (module olamgpu_iface, include/olamgpu.h).  With this file in place of those two in
! the build, timestep.f90 and olam_run.f90 compile unchanged.
!
! The prognostic state lives on the GPU from init_wrtv_rk on.  Per long step:
!   * before the first acoustic step (prog_wrtv_rk): the device side of tend0 and the
!     long-step tendencies of the host's physics (thilt, vmxet, vmyet, vmzet, vmt) go up;
!   * scalar_transport_rk: the device side of timeavg_momsc, each tracer's tendency
!     var_t goes up, the tracers are transported, then the fields the host's physics
!     reads (thil, rho, press, vxe, vye, vze, wc) and the tracers come down.
! olamgpu_shim_download_all brings the whole state down (history output).  The host's
! own tend0 and timeavg_momsc keep running on the host arrays; the device copies are the
! ones that count.
!
! Host modules read: mem_grid, mem_ijtabs, mem_adv, mem_basic, mem_tend, misc_coms,
! consts_coms, oname_coms, var_tables, by the names they have in OLAM.  Arrays of those
! modules are passed by address; the library copies what it needs during each call and
! keeps no host pointer.

module olamgpu_host
  use iso_c_binding
  use olamgpu_iface
  implicit none
  private
  public :: ctx, olamgpu_precision, olamgpu_adjoint_tape, st_all, st_tend, st_phys
  public :: olamgpu_host_init, olamgpu_shim_download_all, shim_check

  type(c_ptr), save :: ctx = c_null_ptr
  ! OLAMGPU_PREC_MIXED is the reference's own precision map (FP64 rho, press); a host may
  ! select OLAMGPU_PREC_F32 (FP32 bulk) or OLAMGPU_PREC_F64 before init_wrtv_rk
  integer(c_int32_t), save :: olamgpu_precision = OLAMGPU_PREC_MIXED
  integer(c_int32_t), save :: olamgpu_adjoint_tape = 0
  type(olamgpu_state_t), save :: st_all, st_tend, st_phys

  ! contiguous copies of the itab_w / itab_v / itab_m components the path reads
  integer(c_int32_t), allocatable, target, save :: w_npoly(:), w_iv(:,:), w_iw(:,:), w_im(:,:)
  real(c_float),      allocatable, target, save :: w_dirv(:,:), w_ecx(:,:), w_ecy(:,:), w_ecz(:,:)
  integer(c_int32_t), allocatable, target, save :: v_iw(:,:), v_im(:,:), m_iv(:,:), m_iw(:,:), m_im(:,:)
  real(c_float),      allocatable, target, save :: v_dxps(:,:), v_dyps(:,:)
  ! the host's lpw/lpv/lpm as 32-bit integers
  integer(c_int32_t), allocatable, target, save :: lpw32(:), lpv32(:), lpm32(:)

  ! Address of a contiguous host array: module arrays of the host need not carry the
  ! TARGET attribute, and a contiguous actual argument is passed by reference, so the
  ! dummy's address is the array's.
  interface addr
     module procedure addr_r1, addr_r2, addr_r3, addr_d1, addr_d2
  end interface addr

contains

  function addr_r1(a) result(p)
    real(c_float), target, contiguous, intent(in) :: a(:)
    type(c_ptr) :: p
    p = c_loc(a)
  end function addr_r1
  function addr_r2(a) result(p)
    real(c_float), target, contiguous, intent(in) :: a(:,:)
    type(c_ptr) :: p
    p = c_loc(a)
  end function addr_r2
  function addr_r3(a) result(p)
    real(c_float), target, contiguous, intent(in) :: a(:,:,:)
    type(c_ptr) :: p
    p = c_loc(a)
  end function addr_r3
  function addr_d1(a) result(p)
    real(c_double), target, contiguous, intent(in) :: a(:)
    type(c_ptr) :: p
    p = c_loc(a)
  end function addr_d1
  function addr_d2(a) result(p)
    real(c_double), target, contiguous, intent(in) :: a(:,:)
    type(c_ptr) :: p
    p = c_loc(a)
  end function addr_d2

  subroutine shim_check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= OLAMGPU_OK) then
       print '(a,a,a,i0,a,a)', 'olamgpu ', what, ' failed (', code, '): ', trim(olamgpu_error_message(ctx))
       error stop 'olamgpu'
    end if
  end subroutine shim_check

  ! Create the context from the host's grid, namelist and tables, and point the three
  ! state descriptors at the host's mem_basic / mem_tend arrays.
  subroutine olamgpu_host_init()
    use mem_grid,    only: mza, mwa, mva, mma, lpw, lpv, lpm, arw, arv, volt, volti, volwi, volvi, &
                           dzim, dzit, zfacit, zwgt_top8, zwgt_bot8, gdz_wgtm8, gdz_wgtp8, &
                           wnx, wny, wnz, vnx, vny, vnz, dnu, dniu, dnv, dniv, arw0, arm0, &
                           gxps_coef, gyps_coef, zm, zt, dzit_bot
    use mem_ijtabs,  only: itab_w, itab_v, itab_m
    use mem_adv,     only: a_w, a_wu, xx_yy, xx_yy_m, xy_h, xy_hu, xx0_v, xy0_v, yy0_v, &
                           xx0_vu, xy0_vu, yy0_vu
    use mem_basic,   only: thil, wmc, wc, vxe, vye, vze, rho, press, vmc, vc, wmasc, vmasc, &
                           alpha_press, pwfac, pvfac
    use mem_tend,    only: thilt, vmxet, vmyet, vmzet, vmt
    use misc_coms,   only: dtsm, dtlm, nacoust, nrk_wrtv, nrk_scal, initial, th01d, dn01d
    use consts_coms, only: fcoriol
    use oname_coms,  only: nl
    use var_tables,  only: num_scalar, scalar_tab
    type(olamgpu_mesh_t)   :: m
    type(olamgpu_params_t) :: p
    integer :: iw, iv, im, n

    allocate(w_npoly(mwa), w_iv(7, mwa), w_iw(7, mwa), w_im(7, mwa))
    allocate(w_dirv(7, mwa), w_ecx(7, mwa), w_ecy(7, mwa), w_ecz(7, mwa))
    allocate(v_iw(2, mva), v_im(2, mva), v_dxps(2, mva), v_dyps(2, mva))
    allocate(m_iv(3, mma), m_iw(3, mma), m_im(3, mma))
    allocate(lpw32(mwa), lpv32(mva), lpm32(mma))
    do iw = 1, mwa
       w_npoly(iw) = itab_w(iw)%npoly
       w_iv(:, iw) = itab_w(iw)%iv;   w_iw(:, iw) = itab_w(iw)%iw;   w_im(:, iw) = itab_w(iw)%im
       w_dirv(:, iw) = itab_w(iw)%dirv
       w_ecx(:, iw) = itab_w(iw)%ecvec_vx
       w_ecy(:, iw) = itab_w(iw)%ecvec_vy
       w_ecz(:, iw) = itab_w(iw)%ecvec_vz
    end do
    do iv = 1, mva
       v_iw(:, iv) = itab_v(iv)%iw;   v_im(:, iv) = itab_v(iv)%im
       v_dxps(:, iv) = itab_v(iv)%dxps; v_dyps(:, iv) = itab_v(iv)%dyps
    end do
    do im = 1, mma
       m_iv(:, im) = itab_m(im)%iv;   m_iw(:, im) = itab_m(im)%iw;   m_im(:, im) = itab_m(im)%im
    end do
    lpw32 = lpw; lpv32 = lpv; lpm32 = lpm

    m%mza = mza; m%mwa = mwa; m%mva = mva; m%mma = mma
    m%w_npoly = c_loc(w_npoly); m%w_iv = c_loc(w_iv); m%w_iw = c_loc(w_iw); m%w_im = c_loc(w_im)
    m%w_dirv = c_loc(w_dirv)
    m%w_ecvec_vx = c_loc(w_ecx); m%w_ecvec_vy = c_loc(w_ecy); m%w_ecvec_vz = c_loc(w_ecz)
    m%gxps_coef = addr(gxps_coef); m%gyps_coef = addr(gyps_coef); m%xx_yy = addr(xx_yy)
    m%v_iw = c_loc(v_iw); m%v_im = c_loc(v_im); m%v_dxps = c_loc(v_dxps); m%v_dyps = c_loc(v_dyps)
    m%m_iv = c_loc(m_iv); m%m_iw = c_loc(m_iw); m%m_im = c_loc(m_im); m%xx_yy_m = addr(xx_yy_m)
    m%lpw = c_loc(lpw32); m%lpv = c_loc(lpv32); m%lpm = c_loc(lpm32)
    m%wnx = addr(wnx); m%wny = addr(wny); m%wnz = addr(wnz); m%arw0 = addr(arw0)
    m%vnx = addr(vnx); m%vny = addr(vny); m%vnz = addr(vnz)
    m%dnu = addr(dnu); m%dniu = addr(dniu); m%dnv = addr(dnv); m%dniv = addr(dniv)
    m%arm0 = addr(arm0)
    m%dzim = addr(dzim); m%dzit = addr(dzit); m%zfacit = addr(zfacit)
    m%zwgt_top8 = addr(zwgt_top8); m%zwgt_bot8 = addr(zwgt_bot8)
    m%gdz_wgtm8 = addr(gdz_wgtm8); m%gdz_wgtp8 = addr(gdz_wgtp8)
    m%a_w = addr(a_w); m%a_wu = addr(a_wu)
    m%arw = addr(arw); m%volti = addr(volti); m%volwi = addr(volwi); m%volt = addr(volt)
    m%arv = addr(arv); m%volvi = addr(volvi)
    if (allocated(xy_h)) m%xy_h = addr(xy_h)
    if (allocated(xy_hu)) m%xy_hu = addr(xy_hu)
    if (allocated(xx0_v)) then
       m%xx0_v = addr(xx0_v); m%xy0_v = addr(xy0_v); m%yy0_v = addr(yy0_v)
    end if
    if (allocated(xx0_vu)) then
       m%xx0_vu = addr(xx0_vu); m%xy0_vu = addr(xy0_vu); m%yy0_vu = addr(yy0_vu)
    end if
    m%zm = addr(zm); m%zt = addr(zt); m%dzit_bot = addr(dzit_bot)
    if (allocated(th01d)) m%th01d = addr(th01d)
    if (allocated(dn01d)) m%dn01d = addr(dn01d)

    p%precision = olamgpu_precision
    p%num_scalar = num_scalar
    p%nrk_wrtv = nrk_wrtv; p%nrk_scal = nrk_scal
    p%thil_horiz_adv_order = nl%thil_horiz_adv_order
    p%wind_horiz_adv_order = nl%wind_horiz_adv_order
    p%scal_horiz_adv_order = nl%scal_horiz_adv_order
    p%iscal_monot = nl%iscal_monot
    p%ithil_monot = nl%ithil_monot; p%iwind_monot = nl%iwind_monot
    p%divh_damp_level = nl%divh_damp_level; p%divh_damp_fact = nl%divh_damp_fact
    p%vort_damp_level = nl%vort_damp_level; p%vort_damp_fact = nl%vort_damp_fact
    p%dtsm = dtsm; p%dtlm = dtlm; p%nacoust = nacoust
    p%adjoint_tape = olamgpu_adjoint_tape
    p%fcoriol = fcoriol
    p%initial = initial
    p%rayf_zmin = nl%rayf_zmin; p%rayf_fact = nl%rayf_fact; p%rayf_expon = nl%rayf_expon
    p%rayfw_zmin = nl%rayfw_zmin; p%rayfw_fact = nl%rayfw_fact; p%rayfw_expon = nl%rayfw_expon
    p%rayfdiv_zmin = nl%rayfdiv_zmin; p%rayfdiv_fact = nl%rayfdiv_fact; p%rayfdiv_expon = nl%rayfdiv_expon
    p%rayfmix_zmin = nl%rayfmix_zmin; p%rayfmix_fact = nl%rayfmix_fact; p%rayfmix_expon = nl%rayfmix_expon
    p%akmin = nl%akmin(1)
    p%nudflag = nl%nudflag
    p%nud_preserve_total_mass = merge(1_c_int32_t, 0_c_int32_t, nl%nud_preserve_total_mass)
    p%nud_preserve_mix_ratio = merge(1_c_int32_t, 0_c_int32_t, nl%nud_preserve_mix_ratio)
    call shim_check(olamgpu_create(m, p, -1_c_int32_t, ctx), 'create')
    do n = 1, num_scalar
       call shim_check(olamgpu_set_scalar_pdef(ctx, int(n, c_int32_t), &
            merge(1_c_int32_t, 0_c_int32_t, scalar_tab(n)%pdef)), 'set_scalar_pdef')
       call shim_check(olamgpu_set_scalar_sgsmix(ctx, int(n, c_int32_t), &
            merge(1_c_int32_t, 0_c_int32_t, scalar_tab(n)%do_sgsmix)), 'set_scalar_sgsmix')
    end do

    st_all%thil = addr(thil); st_all%wmc = addr(wmc); st_all%wc = addr(wc)
    st_all%vxe = addr(vxe); st_all%vye = addr(vye); st_all%vze = addr(vze)
    st_all%rho = addr(rho); st_all%press = addr(press)
    st_all%vmc = addr(vmc); st_all%vc = addr(vc)
    st_all%wmasc = addr(wmasc); st_all%vmasc = addr(vmasc)
    st_all%alpha_press = addr(alpha_press); st_all%pwfac = addr(pwfac); st_all%pvfac = addr(pvfac)
    st_all%thilt = addr(thilt); st_all%vmxet = addr(vmxet); st_all%vmyet = addr(vmyet)
    st_all%vmzet = addr(vmzet); st_all%vmt = addr(vmt)
    st_tend%thilt = st_all%thilt; st_tend%vmxet = st_all%vmxet; st_tend%vmyet = st_all%vmyet
    st_tend%vmzet = st_all%vmzet; st_tend%vmt = st_all%vmt
    st_phys%thil = st_all%thil; st_phys%rho = st_all%rho; st_phys%press = st_all%press
    st_phys%vxe = st_all%vxe; st_phys%vye = st_all%vye; st_phys%vze = st_all%vze
    st_phys%wc = st_all%wc
  end subroutine olamgpu_host_init

  ! The whole state and the tracers, device -> host (history and restart output).
  subroutine olamgpu_shim_download_all()
    use var_tables, only: num_scalar, scalar_tab
    integer :: n
    call shim_check(olamgpu_download_state(ctx, st_all), 'download_state')
    do n = 1, num_scalar
       call shim_check(olamgpu_download_scalar(ctx, int(n, c_int32_t), c_loc(scalar_tab(n)%var_p)), &
            'download_scalar')
    end do
  end subroutine olamgpu_shim_download_all

end module olamgpu_host


module wrtv_rk
  use iso_c_binding
  use olamgpu_iface
  use olamgpu_host
  implicit none
  private
  public :: init_wrtv_rk, prog_wrtv_rk
  integer, save :: nacoust_done = 0

contains

  ! olam_run.f90:566, after rayf_init and the initial state: create the context, upload
  ! the state and the tracers once (init_wrtv_rk keeps the state's vc as the wind that
  ! Rayleigh friction relaxes toward, prog_wrtv_rk.f90:36-54).
  subroutine init_wrtv_rk()
    use var_tables, only: num_scalar, scalar_tab
    integer :: n
    call olamgpu_host_init()
    call shim_check(olamgpu_upload_state(ctx, st_all), 'upload_state')
    call shim_check(olamgpu_init_wrtv_rk(ctx), 'init_wrtv_rk')
    do n = 1, num_scalar
       call shim_check(olamgpu_upload_scalar(ctx, int(n, c_int32_t), c_loc(scalar_tab(n)%var_p), &
            c_null_ptr), 'upload_scalar')
    end do
    nacoust_done = 0
  end subroutine init_wrtv_rk

  ! timestep.f90:205, nacoust times per long step.
  subroutine prog_wrtv_rk()
    use misc_coms, only: nacoust
    if (nacoust_done == 0) then
       call shim_check(olamgpu_tend0(ctx), 'tend0')
       call shim_check(olamgpu_upload_state(ctx, st_tend), 'upload tendencies')
    end if
    call shim_check(olamgpu_prog_wrtv_rk(ctx), 'prog_wrtv_rk')
    nacoust_done = mod(nacoust_done + 1, max(nacoust, 1))
  end subroutine prog_wrtv_rk

end module wrtv_rk


module scalar_transport
  use iso_c_binding
  use olamgpu_iface
  use olamgpu_host
  implicit none
  private
  public :: scalar_transport_rk

contains

  ! timestep.f90:254.  rho_old is the density olamgpu_tend0 kept on the device at the
  ! start of the long step (the host's copy is the same field), so it is not sent.
  subroutine scalar_transport_rk(rho_old)
    use mem_grid,   only: mza, mwa
    use var_tables, only: num_scalar, scalar_tab
    real, intent(in) :: rho_old(mza, mwa)
    integer :: n
    call shim_check(olamgpu_timeavg_momsc(ctx), 'timeavg_momsc')
    do n = 1, num_scalar
       if (associated(scalar_tab(n)%var_t)) &
            call shim_check(olamgpu_upload_scalar(ctx, int(n, c_int32_t), c_null_ptr, &
            c_loc(scalar_tab(n)%var_t)), 'upload tracer tendency')
    end do
    call shim_check(olamgpu_scalar_transport_rk(ctx, c_null_ptr), 'scalar_transport_rk')
    call shim_check(olamgpu_download_state(ctx, st_phys), 'download_state')
    do n = 1, num_scalar
       call shim_check(olamgpu_download_scalar(ctx, int(n, c_int32_t), c_loc(scalar_tab(n)%var_p)), &
            'download_scalar')
    end do
  end subroutine scalar_transport_rk

end module scalar_transport

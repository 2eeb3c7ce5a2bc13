! shim_driver.f90 -- test host for olam_amd/fortran/olam_shim.f90: fills the host
! modules (tests/fortran/host_modules.f90) from a case directory written by
! tests/helpers.py:write_fortran_case (params.txt, fields.txt, one raw .bin per array),
! then calls the shim as OLAM's olam_run / timestep do: init_wrtv_rk once, and per long
! step nacoust x prog_wrtv_rk and scalar_transport_rk(rho_old).  The long-step
! tendencies thilt, vmxet, vmyet, vmzet, vmt of the case stand for the host physics'
! (the same every long step); tracer tendencies come from the case's 'scalart' arrays,
! zero where it has none.
!
! Writes <dir>/shim_phys_<name>.bin for the fields the shim brings back each long step,
! <dir>/shim_<name>.bin for every state member after olamgpu_shim_download_all, and
! <dir>/shim_scalar<n>.bin for the tracers.
program shim_driver
  use iso_c_binding
  use olamgpu_iface, only: OLAMGPU_PREC_F32
  use olamgpu_host,  only: olamgpu_precision, olamgpu_adjoint_tape, olamgpu_shim_download_all
  use wrtv_rk,       only: init_wrtv_rk, prog_wrtv_rk
  use scalar_transport, only: scalar_transport_rk
  use consts_coms
  use misc_coms
  use oname_coms
  use mem_grid
  use mem_ijtabs
  use mem_adv
  use mem_basic
  use mem_tend
  use var_tables
  implicit none

  character(len=1024) :: dir
  character(len=32) :: group, name
  integer(c_int64_t) :: nbytes
  integer :: u, ios, istep, iac, nsteps, n, iw, iv, im, prec, tape, nuds(3)
  integer :: sizes(4)
  integer, allocatable :: i2(:,:)
  real, allocatable :: r2(:,:)
  real, allocatable :: rho_old(:,:)
  real, allocatable :: thilt0(:,:), vmxet0(:,:), vmyet0(:,:), vmzet0(:,:), vmt0(:,:)
  integer :: nt

  call get_command_argument(1, dir)
  if (len_trim(dir) == 0) stop 'usage: olam_shim_driver <dir>'

  open(newunit=u, file=trim(dir)//'/params.txt', status='old', action='read')
  read(u, *) prec, num_scalar, nrk_wrtv, nrk_scal, &
       nl%thil_horiz_adv_order, nl%wind_horiz_adv_order, nl%scal_horiz_adv_order, &
       nl%iscal_monot, nl%divh_damp_level, nl%vort_damp_level, nl%divh_damp_fact, &
       nl%vort_damp_fact, dtsm, dtlm, nacoust, tape, fcoriol, nsteps, &
       nl%ithil_monot, nl%iwind_monot, initial, &
       nl%rayf_zmin, nl%rayf_fact, nl%rayf_expon, nl%rayfw_zmin, nl%rayfw_fact, nl%rayfw_expon, &
       nl%rayfdiv_zmin, nl%rayfdiv_fact, nl%rayfdiv_expon, &
       nl%rayfmix_zmin, nl%rayfmix_fact, nl%rayfmix_expon, nl%akmin(1), nuds
  read(u, *) sizes
  close(u)
  nl%nudflag = nuds(1)
  nl%nud_preserve_total_mass = nuds(2) /= 0
  nl%nud_preserve_mix_ratio = nuds(3) /= 0
  olamgpu_precision = int(prec, c_int32_t)
  olamgpu_adjoint_tape = int(tape, c_int32_t)
  mza = sizes(1); mwa = sizes(2); mva = sizes(3); mma = sizes(4)
  allocate(itab_w(mwa), itab_v(mva), itab_m(mma), scalar_tab(num_scalar))

  n = 0
  nt = 0
  open(newunit=u, file=trim(dir)//'/fields.txt', status='old', action='read')
  do
     read(u, *, iostat=ios) group, name, nbytes
     if (ios /= 0) exit
     select case (trim(group))
     case ('mesh')
        call read_mesh(trim(name))
     case ('state')
        call read_state(trim(name))
     case ('scalar')
        n = n + 1
        allocate(scalar_tab(n)%var_p(mza, mwa), scalar_tab(n)%var_t(mza, mwa))
        call rd_r2(trim(name), scalar_tab(n)%var_p)
        scalar_tab(n)%var_t = 0.
     case ('scalart')
        nt = nt + 1
        call rd_r2(trim(name), scalar_tab(nt)%var_t)
     end select
  end do
  close(u)
  allocate(rho_old(mza, mwa))
  thilt0 = thilt; vmxet0 = vmxet; vmyet0 = vmyet; vmzet0 = vmzet; vmt0 = vmt

  call init_wrtv_rk()
  do istep = 1, nsteps
     ! the host's tend0 and physics
     thilt = thilt0; vmxet = vmxet0; vmyet = vmyet0; vmzet = vmzet0; vmt = vmt0
     rho_old = real(rho)
     do iac = 1, nacoust
        call prog_wrtv_rk()
     end do
     call scalar_transport_rk(rho_old)
  end do

  call wr_r2('shim_phys_thil', thil); call wr_d2('shim_phys_rho', rho); call wr_d2('shim_phys_press', press)
  call wr_r2('shim_phys_vxe', vxe); call wr_r2('shim_phys_vye', vye); call wr_r2('shim_phys_vze', vze)
  call wr_r2('shim_phys_wc', wc)
  call olamgpu_shim_download_all()
  call wr_r2('shim_thil', thil); call wr_r2('shim_wmc', wmc); call wr_r2('shim_wc', wc)
  call wr_r2('shim_vxe', vxe); call wr_r2('shim_vye', vye); call wr_r2('shim_vze', vze)
  call wr_d2('shim_rho', rho); call wr_d2('shim_press', press)
  call wr_r2('shim_vmc', vmc); call wr_r2('shim_vc', vc)
  call wr_r2('shim_wmasc', wmasc); call wr_r2('shim_vmasc', vmasc)
  do n = 1, num_scalar
     call wr_r2('shim_scalar'//char(ichar('0') + n), scalar_tab(n)%var_p)
  end do
  print '(a,i0,a)', 'olam_shim_driver: ', nsteps, ' long steps done'

contains

  subroutine rd_r1(nm, a)
    character(len=*), intent(in) :: nm
    real, intent(out) :: a(:)
    integer :: v
    open(newunit=v, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='old')
    read(v) a
    close(v)
  end subroutine rd_r1
  subroutine rd_r2(nm, a)
    character(len=*), intent(in) :: nm
    real, intent(out) :: a(:,:)
    integer :: v
    open(newunit=v, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='old')
    read(v) a
    close(v)
  end subroutine rd_r2
  subroutine rd_r3(nm, a)
    character(len=*), intent(in) :: nm
    real, intent(out) :: a(:,:,:)
    integer :: v
    open(newunit=v, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='old')
    read(v) a
    close(v)
  end subroutine rd_r3
  subroutine rd_d1(nm, a)
    character(len=*), intent(in) :: nm
    real(kind(1.d0)), intent(out) :: a(:)
    integer :: v
    open(newunit=v, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='old')
    read(v) a
    close(v)
  end subroutine rd_d1
  subroutine rd_d2(nm, a)
    character(len=*), intent(in) :: nm
    real(kind(1.d0)), intent(out) :: a(:,:)
    integer :: v
    open(newunit=v, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='old')
    read(v) a
    close(v)
  end subroutine rd_d2
  subroutine rd_i1(nm, a)
    character(len=*), intent(in) :: nm
    integer, intent(out) :: a(:)
    integer :: v
    open(newunit=v, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='old')
    read(v) a
    close(v)
  end subroutine rd_i1
  subroutine rd_i2(nm, a)
    character(len=*), intent(in) :: nm
    integer, intent(out) :: a(:,:)
    integer :: v
    open(newunit=v, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='old')
    read(v) a
    close(v)
  end subroutine rd_i2
  subroutine wr_r2(nm, a)
    character(len=*), intent(in) :: nm
    real, intent(in) :: a(:,:)
    integer :: v
    open(newunit=v, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='replace')
    write(v) a
    close(v)
  end subroutine wr_r2
  subroutine wr_d2(nm, a)
    character(len=*), intent(in) :: nm
    real(kind(1.d0)), intent(in) :: a(:,:)
    integer :: v
    open(newunit=v, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='replace')
    write(v) a
    close(v)
  end subroutine wr_d2

  ! one array of the case into its host module, by the C-ABI descriptor's name
  subroutine read_mesh(nm)
    character(len=*), intent(in) :: nm
    select case (nm)
    case ('w_npoly')
       allocate(i2(1, mwa)); call rd_i2(nm, i2)
       do iw = 1, mwa; itab_w(iw)%npoly = i2(1, iw); end do
       deallocate(i2)
    case ('w_iv', 'w_iw', 'w_im')
       allocate(i2(7, mwa)); call rd_i2(nm, i2)
       do iw = 1, mwa
          if (nm == 'w_iv') itab_w(iw)%iv = i2(:, iw)
          if (nm == 'w_iw') itab_w(iw)%iw = i2(:, iw)
          if (nm == 'w_im') itab_w(iw)%im = i2(:, iw)
       end do
       deallocate(i2)
    case ('w_dirv', 'w_ecvec_vx', 'w_ecvec_vy', 'w_ecvec_vz')
       allocate(r2(7, mwa)); call rd_r2(nm, r2)
       do iw = 1, mwa
          if (nm == 'w_dirv') itab_w(iw)%dirv = r2(:, iw)
          if (nm == 'w_ecvec_vx') itab_w(iw)%ecvec_vx = r2(:, iw)
          if (nm == 'w_ecvec_vy') itab_w(iw)%ecvec_vy = r2(:, iw)
          if (nm == 'w_ecvec_vz') itab_w(iw)%ecvec_vz = r2(:, iw)
       end do
       deallocate(r2)
    case ('v_iw', 'v_im')
       allocate(i2(2, mva)); call rd_i2(nm, i2)
       do iv = 1, mva
          if (nm == 'v_iw') itab_v(iv)%iw = i2(:, iv)
          if (nm == 'v_im') itab_v(iv)%im = i2(:, iv)
       end do
       deallocate(i2)
    case ('v_dxps', 'v_dyps')
       allocate(r2(2, mva)); call rd_r2(nm, r2)
       do iv = 1, mva
          if (nm == 'v_dxps') itab_v(iv)%dxps = r2(:, iv)
          if (nm == 'v_dyps') itab_v(iv)%dyps = r2(:, iv)
       end do
       deallocate(r2)
    case ('m_iv', 'm_iw', 'm_im')
       allocate(i2(3, mma)); call rd_i2(nm, i2)
       do im = 1, mma
          if (nm == 'm_iv') itab_m(im)%iv = i2(:, im)
          if (nm == 'm_iw') itab_m(im)%iw = i2(:, im)
          if (nm == 'm_im') itab_m(im)%im = i2(:, im)
       end do
       deallocate(i2)
    case ('gxps_coef'); allocate(gxps_coef(7, mwa)); call rd_r2(nm, gxps_coef)
    case ('gyps_coef'); allocate(gyps_coef(7, mwa)); call rd_r2(nm, gyps_coef)
    case ('xx_yy'); allocate(xx_yy(7, mwa)); call rd_r2(nm, xx_yy)
    case ('xx_yy_m'); allocate(xx_yy_m(3, mma)); call rd_r2(nm, xx_yy_m)
    case ('lpw'); allocate(lpw(mwa)); call rd_i1(nm, lpw)
    case ('lpv'); allocate(lpv(mva)); call rd_i1(nm, lpv)
    case ('lpm'); allocate(lpm(mma)); call rd_i1(nm, lpm)
    case ('wnx'); allocate(wnx(mwa)); call rd_r1(nm, wnx)
    case ('wny'); allocate(wny(mwa)); call rd_r1(nm, wny)
    case ('wnz'); allocate(wnz(mwa)); call rd_r1(nm, wnz)
    case ('arw0'); allocate(arw0(mwa)); call rd_r1(nm, arw0)
    case ('vnx'); allocate(vnx(mva)); call rd_r1(nm, vnx)
    case ('vny'); allocate(vny(mva)); call rd_r1(nm, vny)
    case ('vnz'); allocate(vnz(mva)); call rd_r1(nm, vnz)
    case ('dnu'); allocate(dnu(mva)); call rd_r1(nm, dnu)
    case ('dniu'); allocate(dniu(mva)); call rd_r1(nm, dniu)
    case ('dnv'); allocate(dnv(mva)); call rd_r1(nm, dnv)
    case ('dniv'); allocate(dniv(mva)); call rd_r1(nm, dniv)
    case ('arm0'); allocate(arm0(mma)); call rd_r1(nm, arm0)
    case ('dzim'); allocate(dzim(mza)); call rd_r1(nm, dzim)
    case ('dzit'); allocate(dzit(mza)); call rd_r1(nm, dzit)
    case ('zfacit'); allocate(zfacit(mza)); call rd_r1(nm, zfacit)
    case ('zm'); allocate(zm(mza)); call rd_r1(nm, zm)
    case ('zt'); allocate(zt(mza)); call rd_r1(nm, zt)
    case ('dzit_bot'); allocate(dzit_bot(mza)); call rd_r1(nm, dzit_bot)
    case ('th01d'); allocate(th01d(mza)); call rd_r1(nm, th01d)
    case ('dn01d'); allocate(dn01d(mza)); call rd_r1(nm, dn01d)
    case ('zwgt_top8'); allocate(zwgt_top8(mza)); call rd_d1(nm, zwgt_top8)
    case ('zwgt_bot8'); allocate(zwgt_bot8(mza)); call rd_d1(nm, zwgt_bot8)
    case ('gdz_wgtm8'); allocate(gdz_wgtm8(mza)); call rd_d1(nm, gdz_wgtm8)
    case ('gdz_wgtp8'); allocate(gdz_wgtp8(mza)); call rd_d1(nm, gdz_wgtp8)
    case ('a_w'); allocate(a_w(mza, 2, 2)); call rd_r3(nm, a_w)
    case ('a_wu'); allocate(a_wu(mza, 2, 2)); call rd_r3(nm, a_wu)
    case ('arw'); allocate(arw(mza, mwa)); call rd_r2(nm, arw)
    case ('volti'); allocate(volti(mza, mwa)); call rd_r2(nm, volti)
    case ('volwi'); allocate(volwi(mza, mwa)); call rd_r2(nm, volwi)
    case ('volt'); allocate(volt(mza, mwa)); call rd_d2(nm, volt)
    case ('arv'); allocate(arv(mza, mva)); call rd_r2(nm, arv)
    case ('volvi'); allocate(volvi(mza, mva)); call rd_r2(nm, volvi)
    case ('xy_h'); allocate(xy_h(5, 7, mwa)); call rd_r3(nm, xy_h)
    case ('xy_hu'); allocate(xy_hu(5, 7, mwa)); call rd_r3(nm, xy_hu)
    case ('xx0_v'); allocate(xx0_v(2, mva)); call rd_r2(nm, xx0_v)
    case ('xy0_v'); allocate(xy0_v(2, mva)); call rd_r2(nm, xy0_v)
    case ('yy0_v'); allocate(yy0_v(2, mva)); call rd_r2(nm, yy0_v)
    case ('xx0_vu'); allocate(xx0_vu(2, mva)); call rd_r2(nm, xx0_vu)
    case ('xy0_vu'); allocate(xy0_vu(2, mva)); call rd_r2(nm, xy0_vu)
    case ('yy0_vu'); allocate(yy0_vu(2, mva)); call rd_r2(nm, yy0_vu)
    case default
       print '(a,a)', 'unknown mesh array ', nm
       error stop 2
    end select
  end subroutine read_mesh

  subroutine read_state(nm)
    character(len=*), intent(in) :: nm
    select case (nm)
    case ('thil'); allocate(thil(mza, mwa)); call rd_r2(nm, thil)
    case ('wmc'); allocate(wmc(mza, mwa)); call rd_r2(nm, wmc)
    case ('wc'); allocate(wc(mza, mwa)); call rd_r2(nm, wc)
    case ('vxe'); allocate(vxe(mza, mwa)); call rd_r2(nm, vxe)
    case ('vye'); allocate(vye(mza, mwa)); call rd_r2(nm, vye)
    case ('vze'); allocate(vze(mza, mwa)); call rd_r2(nm, vze)
    case ('rho'); allocate(rho(mza, mwa)); call rd_d2(nm, rho)
    case ('press'); allocate(press(mza, mwa)); call rd_d2(nm, press)
    case ('vmc'); allocate(vmc(mza, mva)); call rd_r2(nm, vmc)
    case ('vc'); allocate(vc(mza, mva)); call rd_r2(nm, vc)
    case ('wmasc'); allocate(wmasc(mza, mwa)); call rd_r2(nm, wmasc)
    case ('vmasc'); allocate(vmasc(mza, mva)); call rd_r2(nm, vmasc)
    case ('alpha_press'); allocate(alpha_press(mza, mwa)); call rd_r2(nm, alpha_press)
    case ('pwfac'); allocate(pwfac(mza, mwa)); call rd_r2(nm, pwfac)
    case ('pvfac'); allocate(pvfac(mza, mva)); call rd_r2(nm, pvfac)
    case ('thilt'); allocate(thilt(mza, mwa)); call rd_r2(nm, thilt)
    case ('vmxet'); allocate(vmxet(mza, mwa)); call rd_r2(nm, vmxet)
    case ('vmyet'); allocate(vmyet(mza, mwa)); call rd_r2(nm, vmyet)
    case ('vmzet'); allocate(vmzet(mza, mwa)); call rd_r2(nm, vmzet)
    case ('vmt'); allocate(vmt(mza, mva)); call rd_r2(nm, vmt)
    case default
       print '(a,a)', 'unknown state array ', nm
       error stop 2
    end select
  end subroutine read_state

end program shim_driver

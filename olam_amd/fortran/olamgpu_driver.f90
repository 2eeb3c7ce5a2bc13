! olamgpu_driver.f90 -- a Fortran host that runs long steps of the dynamical core
! through the ISO_C_BINDING interfaces, in the call order of OLAM's timestep
! (omodel/timestep.f90:190-260):  tend0, nacoust x prog_wrtv_rk, timeavg_momsc,
! scalar_transport_rk.
!
! Usage: olamgpu_fdriver <dir>
!   <dir>/params.txt  one line: precision num_scalar nrk_wrtv nrk_scal thil_order
!                     wind_order scal_order iscal_monot divh_level vort_level
!                     divh_fact vort_fact dtsm dtlm nacoust adjoint_tape fcoriol nsteps
!                     ithil_monot iwind_monot initial, then (zmin fact expon) of
!                     rayf, rayfw, rayfdiv, rayfmix, then akmin nudflag
!                     nud_preserve_total_mass
!   <dir>/fields.txt  one line per array: group name bytes, group = mesh | state | scalar
!                     (mesh arrays also carry mza mwa mva mma in params.txt's mesh line)
!   <dir>/<name>.bin  raw bytes of each array in Fortran layout
! Writes <dir>/out_<name>.bin for every state field and scalar, and out_mass.bin.

program olamgpu_fdriver
  use iso_c_binding
  use olamgpu_iface
  implicit none

  type buf_t
     character(len=32) :: group, name
     integer(c_int64_t) :: nbytes
     integer(c_int8_t), pointer :: b(:) => null()
  end type buf_t

  character(len=1024) :: dir
  type(olamgpu_mesh_t) :: mesh
  type(olamgpu_params_t) :: prm
  type(olamgpu_state_t) :: st
  type(c_ptr) :: ctx
  type(buf_t) :: bufs(128)
  integer :: nbuf, u, ios, i, istep, nsteps, iac, n
  integer :: sizes(4)
  character(len=32) :: group, name
  integer(c_int64_t) :: nbytes
  real(c_double), allocatable :: mass(:)
  integer(c_int32_t) :: rc

  call get_command_argument(1, dir)
  if (len_trim(dir) == 0) stop 'usage: olamgpu_fdriver <dir>'

  open(newunit=u, file=trim(dir)//'/params.txt', status='old', action='read')
  read(u, *) prm%precision, prm%num_scalar, prm%nrk_wrtv, prm%nrk_scal, &
       prm%thil_horiz_adv_order, prm%wind_horiz_adv_order, prm%scal_horiz_adv_order, &
       prm%iscal_monot, prm%divh_damp_level, prm%vort_damp_level, prm%divh_damp_fact, &
       prm%vort_damp_fact, prm%dtsm, prm%dtlm, prm%nacoust, prm%adjoint_tape, prm%fcoriol, nsteps, &
       prm%ithil_monot, prm%iwind_monot, prm%initial, &
       prm%rayf_zmin, prm%rayf_fact, prm%rayf_expon, prm%rayfw_zmin, prm%rayfw_fact, prm%rayfw_expon, &
       prm%rayfdiv_zmin, prm%rayfdiv_fact, prm%rayfdiv_expon, &
       prm%rayfmix_zmin, prm%rayfmix_fact, prm%rayfmix_expon, prm%akmin, &
       prm%nudflag, prm%nud_preserve_total_mass, prm%nud_preserve_mix_ratio
  read(u, *) sizes
  close(u)
  mesh%mza = sizes(1); mesh%mwa = sizes(2); mesh%mva = sizes(3); mesh%mma = sizes(4)

  nbuf = 0
  open(newunit=u, file=trim(dir)//'/fields.txt', status='old', action='read')
  do
     read(u, *, iostat=ios) group, name, nbytes
     if (ios /= 0) exit
     nbuf = nbuf + 1
     if (nbuf > size(bufs)) stop 'too many arrays'
     bufs(nbuf)%group = group
     bufs(nbuf)%name = name
     bufs(nbuf)%nbytes = nbytes
     call read_raw(trim(dir)//'/'//trim(name)//'.bin', bufs(nbuf))
     select case (trim(group))
     case ('mesh')
        call set_mesh(mesh, trim(name), c_loc(bufs(nbuf)%b))
     case ('state')
        call set_state(st, trim(name), c_loc(bufs(nbuf)%b))
     end select
  end do
  close(u)

  call check(olamgpu_create(mesh, prm, 0_c_int32_t, ctx), 'create')
  ! the initial state first: init_wrtv_rk keeps its vc for Rayleigh friction, as
  ! rayf_init does after initialization (olam_run.f90:555,566)
  call check(olamgpu_upload_state(ctx, st), 'upload_state')
  call check(olamgpu_init_wrtv_rk(ctx), 'init_wrtv_rk')
  n = 0
  do i = 1, nbuf
     if (trim(bufs(i)%group) == 'scalar') then
        n = n + 1
        call check(olamgpu_upload_scalar(ctx, int(n, c_int32_t), c_loc(bufs(i)%b), c_null_ptr), 'upload_scalar')
     end if
  end do

  do istep = 1, nsteps
     call check(olamgpu_tend0(ctx), 'tend0')
     do iac = 1, prm%nacoust
        call check(olamgpu_prog_wrtv_rk(ctx), 'prog_wrtv_rk')
     end do
     call check(olamgpu_timeavg_momsc(ctx), 'timeavg_momsc')
     call check(olamgpu_scalar_transport_rk(ctx, c_null_ptr), 'scalar_transport_rk')
  end do

  allocate(mass(1 + prm%num_scalar))
  call check(olamgpu_mass_sums(ctx, mass), 'mass_sums')
  call check(olamgpu_download_state(ctx, st), 'download_state')
  n = 0
  do i = 1, nbuf
     if (trim(bufs(i)%group) == 'scalar') then
        n = n + 1
        call check(olamgpu_download_scalar(ctx, int(n, c_int32_t), c_loc(bufs(i)%b)), 'download_scalar')
     end if
     if (trim(bufs(i)%group) /= 'mesh') call write_raw(trim(dir)//'/out_'//trim(bufs(i)%name)//'.bin', bufs(i))
  end do
  open(newunit=u, file=trim(dir)//'/out_mass.bin', access='stream', form='unformatted', status='replace')
  write(u) mass
  close(u)
  call check(olamgpu_destroy(ctx), 'destroy')
  print '(a,i0,a)', 'olamgpu_fdriver: ', nsteps, ' steps done'

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= OLAMGPU_OK) then
       print '(a,a,a,i0,a,a)', 'olamgpu_', what, ' failed (', code, '): ', trim(olamgpu_error_message(ctx))
       error stop 1
    end if
  end subroutine check

  subroutine read_raw(path, buf)
    character(len=*), intent(in) :: path
    type(buf_t), intent(inout) :: buf
    integer :: v
    allocate(buf%b(buf%nbytes))
    open(newunit=v, file=path, access='stream', form='unformatted', status='old', action='read')
    read(v) buf%b
    close(v)
  end subroutine read_raw

  subroutine write_raw(path, buf)
    character(len=*), intent(in) :: path
    type(buf_t), intent(in) :: buf
    integer :: v
    open(newunit=v, file=path, access='stream', form='unformatted', status='replace')
    write(v) buf%b
    close(v)
  end subroutine write_raw

  subroutine set_mesh(m, nm, p)
    type(olamgpu_mesh_t), intent(inout) :: m
    character(len=*), intent(in) :: nm
    type(c_ptr), intent(in) :: p
    select case (nm)
    case ('w_npoly'); m%w_npoly = p
    case ('w_iv'); m%w_iv = p
    case ('w_iw'); m%w_iw = p
    case ('w_im'); m%w_im = p
    case ('w_dirv'); m%w_dirv = p
    case ('w_ecvec_vx'); m%w_ecvec_vx = p
    case ('w_ecvec_vy'); m%w_ecvec_vy = p
    case ('w_ecvec_vz'); m%w_ecvec_vz = p
    case ('gxps_coef'); m%gxps_coef = p
    case ('gyps_coef'); m%gyps_coef = p
    case ('xx_yy'); m%xx_yy = p
    case ('v_iw'); m%v_iw = p
    case ('v_im'); m%v_im = p
    case ('v_dxps'); m%v_dxps = p
    case ('v_dyps'); m%v_dyps = p
    case ('m_iv'); m%m_iv = p
    case ('m_iw'); m%m_iw = p
    case ('m_im'); m%m_im = p
    case ('xx_yy_m'); m%xx_yy_m = p
    case ('lpw'); m%lpw = p
    case ('lpv'); m%lpv = p
    case ('lpm'); m%lpm = p
    case ('wnx'); m%wnx = p
    case ('wny'); m%wny = p
    case ('wnz'); m%wnz = p
    case ('arw0'); m%arw0 = p
    case ('vnx'); m%vnx = p
    case ('vny'); m%vny = p
    case ('vnz'); m%vnz = p
    case ('dnu'); m%dnu = p
    case ('dniu'); m%dniu = p
    case ('dnv'); m%dnv = p
    case ('dniv'); m%dniv = p
    case ('arm0'); m%arm0 = p
    case ('dzim'); m%dzim = p
    case ('dzit'); m%dzit = p
    case ('zfacit'); m%zfacit = p
    case ('zwgt_top8'); m%zwgt_top8 = p
    case ('zwgt_bot8'); m%zwgt_bot8 = p
    case ('gdz_wgtm8'); m%gdz_wgtm8 = p
    case ('gdz_wgtp8'); m%gdz_wgtp8 = p
    case ('a_w'); m%a_w = p
    case ('a_wu'); m%a_wu = p
    case ('arw'); m%arw = p
    case ('volti'); m%volti = p
    case ('volwi'); m%volwi = p
    case ('volt'); m%volt = p
    case ('arv'); m%arv = p
    case ('volvi'); m%volvi = p
    case ('xy_h'); m%xy_h = p
    case ('xy_hu'); m%xy_hu = p
    case ('xx0_v'); m%xx0_v = p
    case ('xy0_v'); m%xy0_v = p
    case ('yy0_v'); m%yy0_v = p
    case ('xx0_vu'); m%xx0_vu = p
    case ('xy0_vu'); m%xy0_vu = p
    case ('yy0_vu'); m%yy0_vu = p
    case ('zm'); m%zm = p
    case ('zt'); m%zt = p
    case ('th01d'); m%th01d = p
    case ('dn01d'); m%dn01d = p
    case ('dzit_bot'); m%dzit_bot = p
    case default
       print '(a,a)', 'unknown mesh array ', nm
       error stop 2
    end select
  end subroutine set_mesh

  subroutine set_state(s, nm, p)
    type(olamgpu_state_t), intent(inout) :: s
    character(len=*), intent(in) :: nm
    type(c_ptr), intent(in) :: p
    select case (nm)
    case ('thil'); s%thil = p
    case ('wmc'); s%wmc = p
    case ('wc'); s%wc = p
    case ('vxe'); s%vxe = p
    case ('vye'); s%vye = p
    case ('vze'); s%vze = p
    case ('rho'); s%rho = p
    case ('press'); s%press = p
    case ('vmc'); s%vmc = p
    case ('vc'); s%vc = p
    case ('wmasc'); s%wmasc = p
    case ('vmasc'); s%vmasc = p
    case ('alpha_press'); s%alpha_press = p
    case ('pwfac'); s%pwfac = p
    case ('pvfac'); s%pvfac = p
    case ('thilt'); s%thilt = p
    case ('vmxet'); s%vmxet = p
    case ('vmyet'); s%vmyet = p
    case ('vmzet'); s%vmzet = p
    case ('vmt'); s%vmt = p
    case default
       print '(a,a)', 'unknown state array ', nm
       error stop 2
    end select
  end subroutine set_state

end program olamgpu_fdriver

! host_modules.f90 -- minimal stand-ins for the OLAM host modules that
! olam_amd/fortran/olam_shim.f90 reads (test infrastructure).  They declare, under the
! names OLAM uses, only the variables the shim reads, so the shim can be compiled and
! run here; in OLAM the real modules replace them.  Array shapes follow the reference's
! Fortran layout: (mza, n) fields, (7, mwa) / (2, mva) / (3, mma) tables.

module consts_coms
  implicit none
  real :: fcoriol = 0.
end module consts_coms

module misc_coms
  implicit none
  integer, parameter :: r8 = selected_real_kind(13, 300)
  real(r8) :: dtsm = 0., dtlm = 0.
  integer :: nacoust = 1, nrk_wrtv = 3, nrk_scal = 3, initial = 1
  real, allocatable :: th01d(:), dn01d(:)
end module misc_coms

module oname_coms
  implicit none
  type oname_vars
     integer :: thil_horiz_adv_order = 2, wind_horiz_adv_order = 2, scal_horiz_adv_order = 2
     integer :: iscal_monot = 0, ithil_monot = 0, iwind_monot = 0
     real    :: divh_damp_fact = 0.1
     integer :: divh_damp_level = 3
     real    :: vort_damp_fact = 0.075
     integer :: vort_damp_level = 3
     real    :: rayf_zmin = 30000., rayf_fact = 0., rayf_expon = 1.
     real    :: rayfw_zmin = 30000., rayfw_fact = 0., rayfw_expon = 1.
     real    :: rayfdiv_zmin = 30000., rayfdiv_fact = 0., rayfdiv_expon = 1.
     real    :: rayfmix_zmin = 30000., rayfmix_fact = 0., rayfmix_expon = 1.
     real    :: akmin(1) = 0.
     integer :: nudflag = 0
     logical :: nud_preserve_mix_ratio = .true., nud_preserve_total_mass = .true.
  end type oname_vars
  type(oname_vars) :: nl
end module oname_coms

module mem_grid
  implicit none
  integer :: mza = 0, mwa = 0, mva = 0, mma = 0
  integer, allocatable :: lpw(:), lpv(:), lpm(:)
  real, allocatable :: arw(:,:), arv(:,:), volti(:,:), volwi(:,:), volvi(:,:)
  real(kind(1.d0)), allocatable :: volt(:,:)
  real, allocatable :: dzim(:), dzit(:), zfacit(:), zm(:), zt(:), dzit_bot(:)
  real(kind(1.d0)), allocatable :: zwgt_top8(:), zwgt_bot8(:), gdz_wgtm8(:), gdz_wgtp8(:)
  real, allocatable :: wnx(:), wny(:), wnz(:), arw0(:), vnx(:), vny(:), vnz(:)
  real, allocatable :: dnu(:), dniu(:), dnv(:), dniv(:), arm0(:)
  real, allocatable :: gxps_coef(:,:), gyps_coef(:,:)
end module mem_grid

module mem_ijtabs
  implicit none
  type itab_w_vars
     integer :: npoly = 0
     integer :: im(7) = 1, iv(7) = 1, iw(7) = 1
     real    :: dirv(7) = 0., ecvec_vx(7) = 0., ecvec_vy(7) = 0., ecvec_vz(7) = 0.
  end type itab_w_vars
  type itab_v_vars
     integer :: im(2) = 1, iw(2) = 1
     real    :: dxps(2) = 0., dyps(2) = 0.
  end type itab_v_vars
  type itab_m_vars
     integer :: iv(3) = 1, im(3) = 1, iw(3) = 1
  end type itab_m_vars
  type(itab_w_vars), allocatable, target :: itab_w(:)
  type(itab_v_vars), allocatable, target :: itab_v(:)
  type(itab_m_vars), allocatable, target :: itab_m(:)
end module mem_ijtabs

module mem_adv
  implicit none
  real, allocatable :: a_w(:,:,:), a_wu(:,:,:), xx_yy(:,:), xx_yy_m(:,:)
  real, allocatable :: xy_h(:,:,:), xy_hu(:,:,:)
  real, allocatable :: xx0_v(:,:), xy0_v(:,:), yy0_v(:,:), xx0_vu(:,:), xy0_vu(:,:), yy0_vu(:,:)
end module mem_adv

module mem_basic
  implicit none
  real, allocatable :: vmc(:,:), vc(:,:), wmc(:,:), wc(:,:), thil(:,:)
  real, allocatable :: vxe(:,:), vye(:,:), vze(:,:), wmasc(:,:), vmasc(:,:)
  real, allocatable :: alpha_press(:,:), pwfac(:,:), pvfac(:,:)
  real(kind(1.d0)), allocatable :: rho(:,:), press(:,:)
end module mem_basic

module mem_tend
  implicit none
  real, allocatable :: thilt(:,:), vmxet(:,:), vmyet(:,:), vmzet(:,:), vmt(:,:)
end module mem_tend

module var_tables
  implicit none
  type scalar_table
     real, pointer, contiguous :: var_p(:,:) => null()
     real, pointer, contiguous :: var_t(:,:) => null()
     logical :: pdef = .true.
     logical :: do_sgsmix = .true.
  end type scalar_table
  type(scalar_table), allocatable :: scalar_tab(:)
  integer :: num_scalar = 0
end module var_tables

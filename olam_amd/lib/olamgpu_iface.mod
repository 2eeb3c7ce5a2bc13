﻿!mod$ v1 sum:26c4e1f26ba30752
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module olamgpu_iface
use,intrinsic::__fortran_builtins,only:__builtin_c_ptr
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::__builtin_c_ptr
private::c_associated
private::c_funloc
private::c_funptr
private::c_f_pointer
private::c_loc
private::c_null_funptr
private::c_null_ptr
private::c_ptr
private::c_sizeof
private::operator(==)
private::operator(/=)
private::c_int8_t
private::c_int16_t
private::c_int32_t
private::c_int64_t
private::c_int128_t
private::c_int
private::c_short
private::c_long
private::c_long_long
private::c_signed_char
private::c_size_t
private::c_intmax_t
private::c_intptr_t
private::c_ptrdiff_t
private::c_int_least8_t
private::c_int_fast8_t
private::c_int_least16_t
private::c_int_fast16_t
private::c_int_least32_t
private::c_int_fast32_t
private::c_int_least64_t
private::c_int_fast64_t
private::c_int_least128_t
private::c_int_fast128_t
private::c_float
private::c_double
private::c_long_double
private::c_float_complex
private::c_double_complex
private::c_long_double_complex
private::c_bool
private::c_char
private::c_null_char
private::c_alert
private::c_backspace
private::c_form_feed
private::c_new_line
private::c_carriage_return
private::c_horizontal_tab
private::c_vertical_tab
private::c_float128
private::c_float128_complex
private::c_uint8_t
private::c_uint16_t
private::c_uint32_t
private::c_uint64_t
private::c_uint128_t
private::c_unsigned_char
private::c_unsigned_short
private::c_unsigned
private::c_unsigned_long
private::c_unsigned_long_long
private::c_uintmax_t
private::c_uint_fast8_t
private::c_uint_fast16_t
private::c_uint_fast32_t
private::c_uint_fast64_t
private::c_uint_fast128_t
private::c_uint_least8_t
private::c_uint_least16_t
private::c_uint_least32_t
private::c_uint_least64_t
private::c_uint_least128_t
private::c_f_procpointer
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
integer(4),parameter::olamgpu_ok=0_4
integer(4),parameter::olamgpu_einval=-1_4
integer(4),parameter::olamgpu_ehip=-2_4
integer(4),parameter::olamgpu_enomem=-3_4
integer(4),parameter::olamgpu_estate=-4_4
integer(4),parameter::olamgpu_prec_f32=0_4
integer(4),parameter::olamgpu_prec_mixed=1_4
integer(4),parameter::olamgpu_prec_f64=2_4
type,bind(c)::olamgpu_mesh_t
integer(4)::mza=0_4
integer(4)::mwa=0_4
integer(4)::mva=0_4
integer(4)::mma=0_4
type(c_ptr)::w_npoly=__builtin_c_ptr(__address=0_8)
type(c_ptr)::w_iv=__builtin_c_ptr(__address=0_8)
type(c_ptr)::w_iw=__builtin_c_ptr(__address=0_8)
type(c_ptr)::w_im=__builtin_c_ptr(__address=0_8)
type(c_ptr)::w_dirv=__builtin_c_ptr(__address=0_8)
type(c_ptr)::w_ecvec_vx=__builtin_c_ptr(__address=0_8)
type(c_ptr)::w_ecvec_vy=__builtin_c_ptr(__address=0_8)
type(c_ptr)::w_ecvec_vz=__builtin_c_ptr(__address=0_8)
type(c_ptr)::gxps_coef=__builtin_c_ptr(__address=0_8)
type(c_ptr)::gyps_coef=__builtin_c_ptr(__address=0_8)
type(c_ptr)::xx_yy=__builtin_c_ptr(__address=0_8)
type(c_ptr)::v_iw=__builtin_c_ptr(__address=0_8)
type(c_ptr)::v_im=__builtin_c_ptr(__address=0_8)
type(c_ptr)::v_dxps=__builtin_c_ptr(__address=0_8)
type(c_ptr)::v_dyps=__builtin_c_ptr(__address=0_8)
type(c_ptr)::m_iv=__builtin_c_ptr(__address=0_8)
type(c_ptr)::m_iw=__builtin_c_ptr(__address=0_8)
type(c_ptr)::m_im=__builtin_c_ptr(__address=0_8)
type(c_ptr)::xx_yy_m=__builtin_c_ptr(__address=0_8)
type(c_ptr)::lpw=__builtin_c_ptr(__address=0_8)
type(c_ptr)::lpv=__builtin_c_ptr(__address=0_8)
type(c_ptr)::lpm=__builtin_c_ptr(__address=0_8)
type(c_ptr)::wnx=__builtin_c_ptr(__address=0_8)
type(c_ptr)::wny=__builtin_c_ptr(__address=0_8)
type(c_ptr)::wnz=__builtin_c_ptr(__address=0_8)
type(c_ptr)::arw0=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vnx=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vny=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vnz=__builtin_c_ptr(__address=0_8)
type(c_ptr)::dnu=__builtin_c_ptr(__address=0_8)
type(c_ptr)::dniu=__builtin_c_ptr(__address=0_8)
type(c_ptr)::dnv=__builtin_c_ptr(__address=0_8)
type(c_ptr)::dniv=__builtin_c_ptr(__address=0_8)
type(c_ptr)::arm0=__builtin_c_ptr(__address=0_8)
type(c_ptr)::dzim=__builtin_c_ptr(__address=0_8)
type(c_ptr)::dzit=__builtin_c_ptr(__address=0_8)
type(c_ptr)::zfacit=__builtin_c_ptr(__address=0_8)
type(c_ptr)::zwgt_top8=__builtin_c_ptr(__address=0_8)
type(c_ptr)::zwgt_bot8=__builtin_c_ptr(__address=0_8)
type(c_ptr)::gdz_wgtm8=__builtin_c_ptr(__address=0_8)
type(c_ptr)::gdz_wgtp8=__builtin_c_ptr(__address=0_8)
type(c_ptr)::a_w=__builtin_c_ptr(__address=0_8)
type(c_ptr)::a_wu=__builtin_c_ptr(__address=0_8)
type(c_ptr)::arw=__builtin_c_ptr(__address=0_8)
type(c_ptr)::volti=__builtin_c_ptr(__address=0_8)
type(c_ptr)::volwi=__builtin_c_ptr(__address=0_8)
type(c_ptr)::volt=__builtin_c_ptr(__address=0_8)
type(c_ptr)::arv=__builtin_c_ptr(__address=0_8)
type(c_ptr)::volvi=__builtin_c_ptr(__address=0_8)
end type
type,bind(c)::olamgpu_params_t
integer(4)::precision=0_4
integer(4)::num_scalar=0_4
integer(4)::nrk_wrtv=3_4
integer(4)::nrk_scal=3_4
integer(4)::thil_horiz_adv_order=2_4
integer(4)::wind_horiz_adv_order=2_4
integer(4)::scal_horiz_adv_order=2_4
integer(4)::iscal_monot=0_4
integer(4)::divh_damp_level=3_4
integer(4)::vort_damp_level=3_4
real(4)::divh_damp_fact=1.00000001490116119384765625e-1_4
real(4)::vort_damp_fact=7.500000298023223876953125e-2_4
real(8)::dtsm=0._8
real(8)::dtlm=0._8
integer(4)::nacoust=1_4
integer(4)::adjoint_tape=0_4
real(8)::fcoriol=0._8
end type
type,bind(c)::olamgpu_state_t
type(c_ptr)::thil=__builtin_c_ptr(__address=0_8)
type(c_ptr)::wmc=__builtin_c_ptr(__address=0_8)
type(c_ptr)::wc=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vxe=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vye=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vze=__builtin_c_ptr(__address=0_8)
type(c_ptr)::rho=__builtin_c_ptr(__address=0_8)
type(c_ptr)::press=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vmc=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vc=__builtin_c_ptr(__address=0_8)
type(c_ptr)::wmasc=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vmasc=__builtin_c_ptr(__address=0_8)
type(c_ptr)::alpha_press=__builtin_c_ptr(__address=0_8)
type(c_ptr)::pwfac=__builtin_c_ptr(__address=0_8)
type(c_ptr)::pvfac=__builtin_c_ptr(__address=0_8)
type(c_ptr)::thilt=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vmxet=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vmyet=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vmzet=__builtin_c_ptr(__address=0_8)
type(c_ptr)::vmt=__builtin_c_ptr(__address=0_8)
end type
interface
function olamgpu_create(mesh,params,device,ctx) bind(c,name="olamgpu_create")
import::c_ptr
import::olamgpu_mesh_t
import::olamgpu_params_t
type(olamgpu_mesh_t),intent(in)::mesh
type(olamgpu_params_t),intent(in)::params
integer(4),value::device
type(c_ptr),intent(out)::ctx
integer(4)::olamgpu_create
end
end interface
interface
function olamgpu_destroy(ctx) bind(c,name="olamgpu_destroy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::olamgpu_destroy
end
end interface
interface
function olamgpu_last_error(ctx) bind(c,name="olamgpu_last_error")
import::c_ptr
type(c_ptr),value::ctx
type(c_ptr)::olamgpu_last_error
end
end interface
interface
function olamgpu_upload_state(ctx,host) bind(c,name="olamgpu_upload_state")
import::c_ptr
import::olamgpu_state_t
type(c_ptr),value::ctx
type(olamgpu_state_t),intent(in)::host
integer(4)::olamgpu_upload_state
end
end interface
interface
function olamgpu_download_state(ctx,host) bind(c,name="olamgpu_download_state")
import::c_ptr
import::olamgpu_state_t
type(c_ptr),value::ctx
type(olamgpu_state_t),intent(in)::host
integer(4)::olamgpu_download_state
end
end interface
interface
function olamgpu_upload_scalar(ctx,n,var_p,var_t) bind(c,name="olamgpu_upload_scalar")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::n
type(c_ptr),value::var_p
type(c_ptr),value::var_t
integer(4)::olamgpu_upload_scalar
end
end interface
interface
function olamgpu_download_scalar(ctx,n,var_p) bind(c,name="olamgpu_download_scalar")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::n
type(c_ptr),value::var_p
integer(4)::olamgpu_download_scalar
end
end interface
interface
function olamgpu_init_wrtv_rk(ctx) bind(c,name="olamgpu_init_wrtv_rk")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::olamgpu_init_wrtv_rk
end
end interface
interface
function olamgpu_tend0(ctx) bind(c,name="olamgpu_tend0")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::olamgpu_tend0
end
end interface
interface
function olamgpu_prog_wrtv_rk(ctx) bind(c,name="olamgpu_prog_wrtv_rk")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::olamgpu_prog_wrtv_rk
end
end interface
interface
function olamgpu_timeavg_momsc(ctx) bind(c,name="olamgpu_timeavg_momsc")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::olamgpu_timeavg_momsc
end
end interface
interface
function olamgpu_scalar_transport_rk(ctx,rho_old) bind(c,name="olamgpu_scalar_transport_rk")
import::c_ptr
type(c_ptr),value::ctx
type(c_ptr),value::rho_old
integer(4)::olamgpu_scalar_transport_rk
end
end interface
interface
function olamgpu_scalar_transport_rk_adj(ctx) bind(c,name="olamgpu_scalar_transport_rk_adj")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::olamgpu_scalar_transport_rk_adj
end
end interface
interface
function olamgpu_upload_scalar_adj(ctx,n,lambda) bind(c,name="olamgpu_upload_scalar_adj")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::n
type(c_ptr),value::lambda
integer(4)::olamgpu_upload_scalar_adj
end
end interface
interface
function olamgpu_download_scalar_adj(ctx,n,lambda) bind(c,name="olamgpu_download_scalar_adj")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::n
type(c_ptr),value::lambda
integer(4)::olamgpu_download_scalar_adj
end
end interface
interface
function olamgpu_mass_sums(ctx,out) bind(c,name="olamgpu_mass_sums")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::out(1_8:*)
integer(4)::olamgpu_mass_sums
end
end interface
interface
function olamgpu_cfl_report(ctx,out) bind(c,name="olamgpu_cfl_report")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::out(1_8:4_8)
integer(4)::olamgpu_cfl_report
end
end interface
interface
function olamgpu_copy_field(ctx,field_id,index,dev_buf,direction,elem_bytes) bind(c,name="olamgpu_copy_field")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::field_id
integer(4),value::index
type(c_ptr),value::dev_buf
integer(4),value::direction
integer(4),intent(out)::elem_bytes
integer(4)::olamgpu_copy_field
end
end interface
interface
function olamgpu_last_call_ms(ctx,ms) bind(c,name="olamgpu_last_call_ms")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::ms
integer(4)::olamgpu_last_call_ms
end
end interface
interface
function olamgpu_set_profiling(ctx,on) bind(c,name="olamgpu_set_profiling")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::on
integer(4)::olamgpu_set_profiling
end
end interface
interface
function olamgpu_kernel_profile(ctx,text,len) bind(c,name="olamgpu_kernel_profile")
import::c_ptr
type(c_ptr),value::ctx
character(1_8,1),intent(out)::text(1_8:*)
integer(4),value::len
integer(4)::olamgpu_kernel_profile
end
end interface
contains
function olamgpu_error_message(ctx) result(msg)
type(c_ptr),intent(in)::ctx
character(512_4,1)::msg
end
end

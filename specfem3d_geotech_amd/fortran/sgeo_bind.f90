! Fortran bindings (iso_c_binding) of the libsgeo C ABI (include/sgeo.h)
! that the two drop-in seams call, and the one context a process uses.
!
! The context is created on first use on device SGEO_DEVICE (environment,
! default 0).  SGEO_REDUCTION=sequential makes every PCG dot product sum in
! the reference's order (bit-for-bit runs); the default is the fast
! blocked-tree order.
module sgeo_bind
use iso_c_binding
implicit none
private
public :: sgeo_context, sgeo_fail
public :: sgeo_pcg_solver, sgeo_mesh_create, sgeo_mesh_destroy,               &
          sgeo_mesh_build_stiffness, sgeo_slope_create, sgeo_slope_destroy,   &
          sgeo_slope_begin_srf, sgeo_slope_plastic_iter, sgeo_slope_download
integer(c_int),parameter,public :: SGEO_K_F64=0
type(c_ptr),save :: ctx=c_null_ptr

interface
  integer(c_int) function sgeo_init(device,ctx) bind(C,name='sgeo_init')
    import :: c_int,c_ptr
    integer(c_int),value :: device
    type(c_ptr) :: ctx
  end function sgeo_init
  integer(c_int) function sgeo_set_reduction(ctx,mode)                        &
    bind(C,name='sgeo_set_reduction')
    import :: c_int,c_ptr
    type(c_ptr),value :: ctx
    integer(c_int),value :: mode
  end function sgeo_set_reduction
  type(c_ptr) function sgeo_last_error() bind(C,name='sgeo_last_error')
    import :: c_ptr
  end function sgeo_last_error
  ! src/solver.f90:14-27, with the module globals cg_tol and cg_maxiter
  ! passed explicitly
  integer(c_int) function sgeo_pcg_solver(ctx,neq,nelmt,k,u,f,dprecon,gdof,   &
    tol,maxiter,cg_iter,errcode,errtag) bind(C,name='sgeo_pcg_solver')
    import :: c_int,c_int32_t,c_double,c_ptr,c_char
    type(c_ptr),value :: ctx
    integer(c_int32_t),value :: neq,nelmt
    real(c_double),intent(in) :: k(*),f(*),dprecon(*)
    real(c_double),intent(inout) :: u(*)
    integer(c_int32_t),intent(in) :: gdof(*)
    real(c_double),value :: tol
    integer(c_int32_t),value :: maxiter
    integer(c_int32_t),intent(out) :: cg_iter,errcode
    character(kind=c_char),intent(out) :: errtag(*)
  end function sgeo_pcg_solver
  integer(c_int) function sgeo_mesh_create(ctx,nelmt,neq,corners,gdof,mat_id, &
    nmat,mesh) bind(C,name='sgeo_mesh_create')
    import :: c_int,c_int32_t,c_double,c_ptr
    type(c_ptr),value :: ctx
    integer(c_int32_t),value :: nelmt,neq,nmat
    real(c_double),intent(in) :: corners(*)
    integer(c_int32_t),intent(in) :: gdof(*),mat_id(*)
    type(c_ptr) :: mesh
  end function sgeo_mesh_create
  integer(c_int) function sgeo_mesh_destroy(mesh) bind(C,name='sgeo_mesh_destroy')
    import :: c_int,c_ptr
    type(c_ptr),value :: mesh
  end function sgeo_mesh_destroy
  ! src/preprocess.f90:74-148 on the device; extload and dprecon_out may be
  ! c_null_ptr (pass c_loc of a contiguous array otherwise)
  integer(c_int) function sgeo_mesh_build_stiffness(mesh,gam,ym,nu,kprec,      &
    extload,dprecon_out) bind(C,name='sgeo_mesh_build_stiffness')
    import :: c_int,c_double,c_ptr
    type(c_ptr),value :: mesh
    real(c_double),intent(in) :: gam(*),ym(*),nu(*)
    integer(c_int),value :: kprec
    type(c_ptr),value :: extload,dprecon_out
  end function sgeo_mesh_build_stiffness
  integer(c_int) function sgeo_slope_create(mesh,extload,slope)                &
    bind(C,name='sgeo_slope_create')
    import :: c_int,c_double,c_ptr
    type(c_ptr),value :: mesh
    real(c_double),intent(in) :: extload(*)
    type(c_ptr) :: slope
  end function sgeo_slope_create
  integer(c_int) function sgeo_slope_destroy(slope) bind(C,name='sgeo_slope_destroy')
    import :: c_int,c_ptr
    type(c_ptr),value :: slope
  end function sgeo_slope_destroy
  integer(c_int) function sgeo_slope_begin_srf(slope,ym,nuf,phif,cohf,psif,dt) &
    bind(C,name='sgeo_slope_begin_srf')
    import :: c_int,c_double,c_ptr
    type(c_ptr),value :: slope
    real(c_double),intent(in) :: ym(*),nuf(*),phif(*),cohf(*),psif(*)
    real(c_double),value :: dt
  end function sgeo_slope_begin_srf
  ! one pass of src/slope.F90:286-391
  integer(c_int) function sgeo_slope_plastic_iter(slope,cg_tol,cg_maxiter,     &
    nl_tol,is_last,cg_iter,errcode,uerr,nl_isconv,fmax)                        &
    bind(C,name='sgeo_slope_plastic_iter')
    import :: c_int,c_int32_t,c_double,c_ptr
    type(c_ptr),value :: slope
    real(c_double),value :: cg_tol,nl_tol
    integer(c_int32_t),value :: cg_maxiter
    integer(c_int),value :: is_last
    integer(c_int32_t),intent(out) :: cg_iter,errcode,nl_isconv
    real(c_double),intent(out) :: uerr,fmax
  end function sgeo_slope_plastic_iter
  ! host copies of the trial state; any pointer may be c_null_ptr
  integer(c_int) function sgeo_slope_download(slope,x,bodyload,evpt,stress,    &
    flags) bind(C,name='sgeo_slope_download')
    import :: c_int,c_ptr
    type(c_ptr),value :: slope,x,bodyload,evpt,stress,flags
  end function sgeo_slope_download
end interface

contains

! The process's context: created at the first call.
function sgeo_context() result(c)
type(c_ptr) :: c
character(len=32) :: val
integer :: device,ios,n
integer(c_int) :: rc
if(.not.c_associated(ctx))then
  device=0
  call get_environment_variable('SGEO_DEVICE',val,length=n)
  if(n>0)then
    read(val,*,iostat=ios)device
    if(ios/=0)device=0
  endif
  rc=sgeo_init(int(device,c_int),ctx)
  if(rc/=0)call sgeo_fail(rc,'sgeo_init')
  call get_environment_variable('SGEO_REDUCTION',val,length=n)
  if(n>0 .and. val(1:1)=='s')then
    rc=sgeo_set_reduction(ctx,1_c_int)
    if(rc/=0)call sgeo_fail(rc,'sgeo_set_reduction')
  endif
endif
c=ctx
end function sgeo_context

! Stops with the library's message after a failed call.
subroutine sgeo_fail(rc,what)
integer(c_int),intent(in) :: rc
character(len=*),intent(in) :: what
character(kind=c_char),pointer :: msg(:)
character(len=512) :: text
integer :: i
type(c_ptr) :: p
text=''
p=sgeo_last_error()
if(c_associated(p))then
  call c_f_pointer(p,msg,[512])
  do i=1,512
    if(msg(i)==c_null_char)exit
    text(i:i)=msg(i)
  enddo
endif
write(*,'(a,i0,a)')'ERROR: '//what//' failed (',rc,'): '//trim(text)
stop 1
end subroutine sgeo_fail

end module sgeo_bind

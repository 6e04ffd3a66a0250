! Seam 1: a module `solver` that replaces src/solver.f90 at link time.
! pcg_solver keeps the reference's name, arguments and semantics
! (src/solver.f90:14-79) and solves on the GPU through sgeo_pcg_solver, so
! its callers (src/slope.F90:293, src/excavation.F90:231,525) compile and
! link unchanged.  The library caches the device copies of k, dprecon and
! gdof_elmt between calls (keyed on the arrays and their content).
module solver
use iso_c_binding
use set_precision
use global, only : cg_maxiter,cg_tol,nedof
use sgeo_bind
implicit none
private
public :: pcg_solver

contains

subroutine pcg_solver(neq,nelmt,k,u,f,dprecon,gdof_elmt,cg_iter,errcode,errtag)
integer,intent(in) :: neq,nelmt
real(kind=kreal),dimension(nedof,nedof,nelmt),intent(in) :: k
real(kind=kreal),dimension(0:neq),intent(inout) :: u
real(kind=kreal),dimension(0:neq),intent(in) :: f,dprecon
integer,dimension(:,:),intent(in) :: gdof_elmt
integer,intent(out) :: cg_iter
integer,intent(out) :: errcode
character(len=250),intent(out) :: errtag
integer(c_int) :: rc
integer(c_int32_t) :: it,ec
integer,allocatable :: g(:,:)

if(is_contiguous(gdof_elmt))then
  ! slope3d passes the whole array
  rc=sgeo_pcg_solver(sgeo_context(),int(neq,c_int32_t),int(nelmt,c_int32_t),  &
  k,u,f,dprecon,gdof_elmt,real(cg_tol,c_double),int(cg_maxiter,c_int32_t),it, &
  ec,errtag)
else
  ! excavation3d passes gdof_elmt(:,elmt_intact): a contiguous copy
  g=gdof_elmt
  rc=sgeo_pcg_solver(sgeo_context(),int(neq,c_int32_t),int(nelmt,c_int32_t),  &
  k,u,f,dprecon,g,real(cg_tol,c_double),int(cg_maxiter,c_int32_t),it,ec,errtag)
endif
if(rc/=0)call sgeo_fail(rc,'sgeo_pcg_solver')
cg_iter=it; errcode=ec
end subroutine pcg_solver

end module solver

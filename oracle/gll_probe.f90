! Prints the GLL tables the reference computes at run time for NGLL=3
! (zwgljd, gll_quadrature, dshape_function_hex8 from the reference's own
! gll_library / shape_library objects) in hexadecimal-exact form, one value
! per line, so that the oracle's constant tables can be checked bit for bit.
! Test infrastructure only (built into oracle/_ref/ by oracle/Makefile).
program gll_probe
use set_precision
use gll_library
use shape_library
implicit none
integer,parameter :: ng=3, ngll=27
real(kind=kreal) :: xi(ng),w(ng)
real(kind=kreal) :: gll_weights(ngll),gll_points(ndim,ngll)
real(kind=kreal) :: lagrange_gll(ngll,ngll),dlagrange_gll(ndim,ngll,ngll)
real(kind=kreal) :: dshape_hex8(ndim,8,ngll)
integer :: i,j,k
call zwgljd(xi,w,ng,0.0_kreal,0.0_kreal)
write(*,'(a)')'# zwgljd points, weights'
do i=1,ng
  write(*,'(es25.17e3,1x,es25.17e3)')xi(i),w(i)
enddo
call dshape_function_hex8(8,ng,ng,ng,xi,xi,xi,dshape_hex8)
call gll_quadrature(ndim,ng,ng,ng,ngll,gll_points,gll_weights,lagrange_gll,    &
dlagrange_gll)
write(*,'(a)')'# gll_weights(27)'
do i=1,ngll
  write(*,'(es25.17e3)')gll_weights(i)
enddo
write(*,'(a)')'# dlagrange_gll(3,27,27) fortran order'
do k=1,ngll
  do j=1,ngll
    do i=1,ndim
      write(*,'(es25.17e3)')dlagrange_gll(i,j,k)
    enddo
  enddo
enddo
write(*,'(a)')'# dshape_hex8(3,8,27) fortran order'
do k=1,ngll
  do j=1,8
    do i=1,ndim
      write(*,'(es25.17e3)')dshape_hex8(i,j,k)
    enddo
  enddo
enddo
end program gll_probe

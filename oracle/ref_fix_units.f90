! Closes Fortran unit 11 if it is still connected.  The reference's
! read_input (src/read_input.f90:557-600) leaves the material list open on
! unit 11 and the driver then re-opens unit 11 with status='replace'
! (src/semgeotech.F90:143); gfortran and ifort close the old connection
! implicitly, flang's runtime reports an error instead.  Called by the
! read_input wrapper in oracle/ref_tap.c right after the real read_input,
! which is what the implicit close would have done.  Test infrastructure.
subroutine sgeo_ref_close_unit11() bind(C, name='sgeo_ref_close_unit11')
implicit none
integer :: ios
logical :: isopen
inquire(unit=11, opened=isopen)
if(isopen)close(11, iostat=ios)
end subroutine sgeo_ref_close_unit11

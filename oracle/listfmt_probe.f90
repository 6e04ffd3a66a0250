! List-directed output of the compiler the reference is built with
! (amdflang), for the summary-file writer's formatting tests (test
! infrastructure, built into oracle/_ref/).  The reference writes its
! summary rows with list-directed WRITE (src/slope.F90:217-219,415;
! src/excavation.F90:270-279,692).
!   usage: listfmt_probe <in.bin>
!   in.bin: int32 n, then n doubles, then int32 m, then m records of
!           (double srf, int32 cg, int32 nl, double uxmax, umax, fmax)
!   stdout: one list-directed line per double, then one list-directed
!           write of every record (the summary row form, which the
!           runtime wraps at its record length)
program listfmt_probe
implicit none
character(len=512) :: fin
integer :: n,m,i,cg,nl
real(kind=8) :: v,srf,ux,um,fm
call get_command_argument(1,fin)
open(21,file=trim(fin),access='stream',form='unformatted',status='old')
read(21)n
do i=1,n
  read(21)v
  write(*,*)v
enddo
read(21)m
do i=1,m
  read(21)srf,cg,nl,ux,um,fm
  write(*,*)srf,cg,nl,ux,um,fm
enddo
close(21)
end program listfmt_probe

! Runs the reference's own compiled constitutive routines on inputs read
! from a file, so each routine of the C restatement can be compared with it
! value by value (test infrastructure, built into oracle/_ref/).
!   usage: fn_probe <in.bin> <out.bin>
!   in.bin : int32 n, then n records of 9 doubles
!            (stress(6), phi, c, psi)             phi/psi in degrees
!   out.bin: n records of 115 doubles
!            sigm, dsbar, theta, f, dq1, dq2, dq3, m1(36), m2(36), m3(36)
!            (stress_invariant src/math_library.f90:144, mohcouf
!            src/plastic_library.f90:95, mohcouq :116, formm :150)
!            then, once, for each of the first min(n,16) records used as
!            (srf=1+phi/45, nu=0.1+c/100, ym=1e5, coh=c, phi, psi):
!            cohf, nuf, phif, psif, istat(phinu on), dt_viscoplas
program fn_probe
use set_precision
use math_library
use plastic_library
implicit none
character(len=512) :: fin,fout
integer :: n,i,k,istat
real(kind=kreal) :: rec(9),s(6),sigm,dsbar,theta,f,dq1,dq2,dq3
real(kind=kreal) :: m1(6,6),m2(6,6),m3(6,6)
real(kind=kreal) :: coh(1),nu(1),phi(1),psi(1),cohf(1),nuf(1),phif(1),psif(1),ym(1)
real(kind=kreal) :: srf,dt
call get_command_argument(1,fin)
call get_command_argument(2,fout)
open(21,file=trim(fin),access='stream',form='unformatted',status='old')
open(22,file=trim(fout),access='stream',form='unformatted',status='replace')
read(21)n
do i=1,n
  read(21)rec
  s=rec(1:6)
  call stress_invariant(s,sigm,dsbar,theta)
  call mohcouf(rec(7),rec(8),sigm,dsbar,theta,f)
  call mohcouq(rec(9),dsbar,theta,dq1,dq2,dq3)
  call formm(s,m1,m2,m3)
  write(22)sigm,dsbar,theta,f,dq1,dq2,dq3,m1,m2,m3
enddo
close(21)
open(21,file=trim(fin),access='stream',form='unformatted',status='old')
read(21)n
do i=1,min(n,16)
  read(21)rec
  srf=1.0_kreal+rec(7)/45.0_kreal
  coh(1)=rec(8); nu(1)=0.1_kreal+rec(8)/100.0_kreal; phi(1)=rec(7); psi(1)=rec(9)
  ym(1)=1.0e5_kreal
  call strength_reduction(srf,.true.,1,coh,nu,phi,psi,cohf,nuf,phif,psif,istat)
  dt=dt_viscoplas(1,nuf,phif,ym)
  write(22)cohf(1),nuf(1),phif(1),psif(1),real(istat,kreal),dt
enddo
close(21)
close(22)
end program fn_probe

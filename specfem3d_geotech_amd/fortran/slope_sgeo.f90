! Seam 2: a module `slope` that replaces src/slope.F90 at link time.
!
! slope3d keeps the reference's entry point (src/slope.F90:13,38-39, called
! from src/semgeotech.F90:329) and its host-side order: boundary conditions
! and equation numbering, the summary file, the strength-reduction loop with
! its failure rule, the nodal displacement, the summary row and save_data
! (src/slope.F90:244-443).  What differs is where the work happens: the
! mesh, the elastic stiffness, and the trial's x, bodyload, evpt and
! stress_elmt live on the GPU for the whole sweep; the stiffness and the
! gravity load are built there (sgeo_mesh_build_stiffness), and every
! plastic iteration -- load, PCG, uerr, the element x GLL body and the body
! load -- is one sgeo_slope_plastic_iter call.  The host downloads x and
! stress_elmt once per SRF, for the summary row and the result files.
!
! Not on this path (each stops with a message): traction and water-table
! loads, pseudo-static loads and all-elastic models.
module slope
contains

subroutine slope3d(sum_file,format_str)
use iso_c_binding
use set_precision
use global
use math_constants
use element,only:hex8_gnode
use plastic_library,only:dt_viscoplas,strength_reduction
use serial_library,only:error_stop
use math_library_serial,only:maxvec
use postprocess,only:save_data
use sgeo_bind
implicit none
character(len=250),intent(in) :: sum_file
character(len=20),intent(in) :: format_str

integer :: i,j,neq,istat,i_elmt,i_node,i_srf,nl_iter,nl_tot,cg_tot,errcode,ios
integer(c_int) :: rc
integer(c_int32_t) :: cg_iter,ec,isconv
real(kind=kreal) :: dt,fmax,uerr,uxmax,umax
real(c_double) :: fmax_c,uerr_c
character(len=250) :: errtag
type(c_ptr) :: mesh=c_null_ptr,trial=c_null_ptr
integer,allocatable :: gdof_elmt(:,:),node_valency(:)
real(kind=kreal),allocatable :: corners(:,:,:)
real(kind=kreal),allocatable,target :: extload(:),dprecon(:),x(:),            &
stress_elmt(:,:,:)
real(kind=kreal),allocatable :: nodalu(:,:),stress_global(:,:),scf(:),vmeps(:)
real(kind=kreal),allocatable :: cohf_blk(:),nuf_blk(:),phif_blk(:),psif_blk(:)

errtag=""; errcode=-1
if(istraction .or. iswater .or. iseqload .or. allelastic)then
  call error_stop('ERROR: traction, water, pseudo-static loads and all-elastic '//&
  'models are not on the GPU slope path!')
endif

! displacement boundary conditions and equation numbers (src/apply_bc.f90)
if(myrank==0)write(stdout,'(a)',advance='no')'applying BC...'
allocate(gdof(nndof,nnode),stat=istat)
if(istat/=0)call error_stop('ERROR: cannot allocate memory!')
gdof=1
call apply_bc(neq,errcode,errtag)
if(errcode/=0)call error_stop(errtag)
if(myrank==0)write(stdout,*)'complete!'
if(myrank==0)write(stdout,*)'degrees of freedoms => total:',neq

! element equations and corner coordinates: what the device mesh holds
allocate(gdof_elmt(nedof,nelmt),corners(ndim,8,nelmt),stat=istat)
if(istat/=0)call error_stop('ERROR: cannot allocate memory!')
do i_elmt=1,nelmt
  gdof_elmt(:,i_elmt)=reshape(gdof(:,g_num(:,i_elmt)),(/nedof/))
  corners(:,:,i_elmt)=g_coord(:,g_num(hex8_gnode,i_elmt))
enddo

! resident mesh; stiffness, inverted diagonal and gravity load on the GPU
if(myrank==0)write(stdout,'(a)',advance='no')'preprocessing on the GPU...'
rc=sgeo_mesh_create(sgeo_context(),int(nelmt,c_int32_t),int(neq,c_int32_t),   &
corners,gdof_elmt,mat_id,int(nmatblk,c_int32_t),mesh)
if(rc/=0)call sgeo_fail(rc,'sgeo_mesh_create')
allocate(extload(0:neq),dprecon(0:neq),x(0:neq),stress_elmt(nst,ngll,nelmt),  &
stat=istat)
if(istat/=0)call error_stop('ERROR: cannot allocate memory!')
rc=sgeo_mesh_build_stiffness(mesh,gam_blk,ym_blk,nu_blk,SGEO_K_F64,           &
c_loc(extload),c_loc(dprecon))
if(rc/=0)call sgeo_fail(rc,'sgeo_mesh_build_stiffness')
rc=sgeo_slope_create(mesh,extload,trial)
if(rc/=0)call sgeo_fail(rc,'sgeo_slope_create')
if(myrank==0)write(stdout,*)'complete!'
if(myrank==0)write(stdout,'(a)')'--------------------------------------------'

allocate(nodalu(nndof,nnode),stress_global(nst,nnode),scf(nnode),vmeps(nnode), &
node_valency(nnode))
node_valency=0
do i_elmt=1,nelmt
  node_valency(g_num(:,i_elmt))=node_valency(g_num(:,i_elmt))+1
enddo

! summary header (src/slope.F90:212-219)
open(unit=10,file=trim(sum_file),status='old',position='append',action='write',&
iostat=ios)
write(10,*)'CG_MAXITER, CG_TOL, NL_MAXITER, NL_TOL'
write(10,*)cg_maxiter,cg_tol,nl_maxiter,nl_tol
write(10,*)'Number of SRFs'
write(10,*)nsrf
write(10,*)'SRF, CGITER, NLITER, UXMAX, UMAX, fmax'
close(10)

allocate(cohf_blk(nmatblk),nuf_blk(nmatblk),phif_blk(nmatblk),psif_blk(nmatblk))

! strength-reduction loop (src/slope.F90:244-443)
srf_loop: do i_srf=1,nsrf
  if(myrank==0)write(stdout,'(/,a,f7.4)')'SRF:',srf(i_srf)
  nodalu=zero; vmeps=zero; scf=inftol
  call strength_reduction(srf(i_srf),phinu,nmatblk,coh_blk,nu_blk,phi_blk,     &
  psi_blk,cohf_blk,nuf_blk,phif_blk,psif_blk,istat)
  dt=dt_viscoplas(nmatblk,nuf_blk,phif_blk,ym_blk)
  if(istat==1)then
    ! nu changed with the SRF: the device stiffness and diagonal follow
    rc=sgeo_mesh_build_stiffness(mesh,gam_blk,ym_blk,nuf_blk,SGEO_K_F64,      &
    c_null_ptr,c_null_ptr)
    if(rc/=0)call sgeo_fail(rc,'sgeo_mesh_build_stiffness')
  endif
  ! the trial restarts from zero state on the device
  rc=sgeo_slope_begin_srf(trial,ym_blk,nuf_blk,phif_blk,cohf_blk,psif_blk,dt)
  if(rc/=0)call sgeo_fail(rc,'sgeo_slope_begin_srf')
  cg_tot=0; fmax=zero; uerr=zero
  plastic: do nl_iter=1,nl_maxiter
    rc=sgeo_slope_plastic_iter(trial,real(cg_tol,c_double),                    &
    int(cg_maxiter,c_int32_t),real(nl_tol,c_double),                           &
    merge(1_c_int,0_c_int,nl_iter==nl_maxiter),cg_iter,ec,uerr_c,isconv,fmax_c)
    if(rc/=0)call sgeo_fail(rc,'sgeo_slope_plastic_iter')
    if(ec/=0)call error_stop('ERROR: PCG solver doesn''t converge!')
    cg_tot=cg_tot+cg_iter
    fmax=fmax_c; uerr=uerr_c
    if(myrank==0)then
      write(stdout,'(a,i4,a,f12.6,a,f12.6)')' nl_iter:',nl_iter,' f_max:',fmax,&
      ' uerr:',uerr
    endif
    if(isconv/=0 .or. nl_iter==nl_maxiter)exit
  end do plastic
  if(nl_iter>=nl_maxiter .and. isconv==0)then
    write(stdout,*)'WARNING: nonconvergence in nonlinear iterations!'
    write(stdout,*)'desired tolerance:',nl_tol,' achieved tolerance:',uerr
  endif
  nl_tot=nl_iter

  ! the trial's displacement and stress, once per SRF
  rc=sgeo_slope_download(trial,c_loc(x),c_null_ptr,c_null_ptr,                 &
  c_loc(stress_elmt),c_null_ptr)
  if(rc/=0)call sgeo_fail(rc,'sgeo_slope_download')
  do i=1,nndof
    do j=1,nnode
      if(gdof(i,j)/=0)nodalu(i,j)=nodalu(i,j)+x(gdof(i,j))
    enddo
  enddo
  uxmax=maxvec(abs(reshape(nodalu,(/nndof*nnode/))))
  umax=maxvec(sqrt(nodalu(1,:)*nodalu(1,:)+nodalu(2,:)*nodalu(2,:)+           &
  nodalu(3,:)*nodalu(3,:)))
  open(10,file=trim(sum_file),status='old',position='append',action='write')
  write(10,*)srf(i_srf),cg_tot,nl_tot,uxmax,umax,fmax
  close(10)

  ! nodal averages for the result files (src/slope.F90:418-437)
  do i_node=1,nnode
    vmeps(i_node)=vmeps(i_node)/real(node_valency(i_node),kreal)
  enddo
  stress_global=zero
  do i_elmt=1,nelmt
    stress_global(:,g_num(:,i_elmt))=stress_global(:,g_num(:,i_elmt))+         &
    stress_elmt(:,:,i_elmt)
  enddo
  do i_node=1,nnode
    stress_global(:,i_node)=stress_global(:,i_node)/                           &
    real(node_valency(i_node),kreal)
  enddo
  call save_data(format_str,i_srf,nnode,nodalu,scf,vmeps,stress_global)

  if(nl_iter==nl_maxiter)exit
enddo srf_loop

rc=sgeo_slope_destroy(trial)
rc=sgeo_mesh_destroy(mesh)
deallocate(mat_id,gam_blk,ym_blk,coh_blk,nu_blk,phi_blk,psi_blk,srf)
deallocate(g_coord,g_num)
deallocate(extload,dprecon,x,stress_elmt,gdof_elmt,corners)
return
end subroutine slope3d

end module slope

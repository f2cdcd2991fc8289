! example_driver.f90 -- the reference's frame loop (src/main/trj_analysis.f90:173-246)
! over the trj_hip Fortran modules.
!
! Input (raw little-endian stream): int32 ndim, nsp, natoms, nframes; real32
! deltar, rcrdf, qmax, qmin; int32 periodic(3); real32 bsc(nsp);
! then per frame: real32 sidel(3), int32 itype(natoms), real32 r(ndim,natoms).
! Output (raw stream): int32 lsmax, nit, nqmax; int64 histomix(lsmax,nsp,nsp);
! real64 sqf(0:nqmax), sqfp(0:nqmax,nsp); int64 nq(0:nqmax).
program example_driver
  use iso_c_binding
  use trj_hip
  use trj_hip_rdf
  use trj_hip_sq
  implicit none
  character(len=512) :: fin, fout
  integer(c_int32_t) :: ndim, nsp, natoms, nframes, per(3)
  real(c_float) :: deltar, rcrdf, qmax, qmin, sidel(3)
  real(c_float), allocatable :: bsc(:), r(:, :)
  integer(c_int32_t), allocatable :: itype(:)
  integer :: i, u
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) ndim, nsp, natoms, nframes
  read (u) deltar, rcrdf, qmax, qmin
  read (u) per
  allocate (bsc(nsp), itype(natoms), r(ndim, natoms))
  read (u) bsc
  do i = 1, nframes
    read (u) sidel, itype, r
    if (i == 1) then
      call RDF_init(nsp, ndim, deltar, rcrdf, sidel, per /= 0)
      call sq_init(natoms, nsp, ndim, qmax, qmin, bsc, sidel)
    end if
    call RDFcomp(natoms, r, itype, sidel)
    call SQcalc(natoms, r, itype, sidel)
  end do
  close (u)
  call rdf_histomix()
  call sq_transfer_gpu_cpu()
  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) int(lsmax, c_int32_t), int(nit, c_int32_t), int(nqmax, c_int32_t)
  write (u) histomix
  write (u) sqf, sqfp, nq
  close (u)
  call rdf_clear()
  call sq_clear()
  print '(a,i0,a)', " example_driver: ", nframes, " frames done"
end program example_driver

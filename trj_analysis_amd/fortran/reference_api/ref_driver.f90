! ref_driver -- the reference's frame loop on the reference-signature modules:
! per frame the cell setup of reformat_input_conf (netcdf.f90:600-622), then
! trj_analysis.f90:193-217 (init_modules on the first frame, transfer_cpu_gpu,
! RDFcomp, SQcalc); after the loop sq_transfer_gpu_cpu (print_results,
! util.f90:269) and clean_memory's sq_clear / rdf_clear (util.f90:203-210).
!
! Input (raw little-endian stream): int32 ndim, nsp, natoms, nframes; real32
! deltar, rcrdf, qmax, qmin (qmax <= 0: no S(Q)); int32 periodic(3); real32
! bsc(nsp); then per frame: real32 sidel(3), int32 itype(natoms), real32
! r(ndim,natoms).
! Output (raw stream): int32 lsmax, nit, nqmax, nqmin; real32
! histomix(lsmax,nsp,nsp), gtot(0:lsmax), rmingr, qmin, dq; then, with S(Q),
! real32 sqf(0:nqmax), sqfp(0:nqmax,nsp) and int32 nq(0:nqmax).
program ref_driver
  use mod_precision
  use mod_common
  use mod_input
  use mod_rdf
  use mod_sq
  implicit none
  character(len=512) :: fin, fout
  integer :: nframes, per(3), i, k, u
  logical :: first_configuration
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) ndim, nsp, natms, nframes
  read (u) deltar, rcrdf, qmax, qmin
  read (u) per
  periodic(:) = per(:) /= 0
  allocate (bsc(nsp))
  read (u) bsc
  Nsites = natms
  nbcuda = Nsites/nthread
  if (nbcuda*nthread /= Nsites) nbcuda = nbcuda + 1
  run_rdf = .true.
  run_sq = qmax > 0.0
  allocate (r(ndim, Nsites), itype(Nsites))
  first_configuration = .true.
  nconf = 0
  do i = 1, nframes
    read (u) sidel, itype, r
    ! reformat_input_conf: cell lengths, the cutoff kept within half the
    ! smallest side, side2 and the volume
    do k = 1, ndim
      if (.not. periodic(k)) then
        sidelv(k) = 2.5*sidel(k)
      else
        sidelv(k) = sidel(k)
      end if
    end do
    side = minval(sidel(1:ndim))
    rcrdf = min(rcrdf, side/2)
    if (rcrdf > 0.0) then
      side2 = rcrdf**2
    else
      side2 = (side/2)**2
    end if
    volumen = 1.0
    do k = 1, ndim
      volumen = volumen*sidel(k)
    end do
    ! trj_analysis.f90:193-217
    if (first_configuration) then
      if (run_rdf) call RDF_init(nsp)
      if (run_sq) call sq_init(Nsites, nsp, nbcuda)
    end if
    first_configuration = .false.
    call transfer_cpu_gpu(ndim)
    if (run_rdf) call RDFcomp(Nsites, i, nbcuda, nthread, 'lj')
    if (run_sq) call SQcalc()
    nconf = nconf + 1
  end do
  close (u)
  if (run_sq) call sq_transfer_gpu_cpu()
  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) lsmax, nit, merge(nqmax, 0, run_sq), merge(nqmin, 0, run_sq)
  write (u) histomix, gtot, rmingr, qmin, dq
  if (run_sq) write (u) sqf, sqfp, nq
  close (u)
  if (run_sq) call sq_clear()
  call rdf_clear()
  print '(a,i0,a)', " ref_driver: ", nconf, " frames done"
end program ref_driver

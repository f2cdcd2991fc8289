! mod_common -- the part of the reference's mod_common (src/core/common.cuf:
! 40-110, 368-378) that the RDF and S(Q) path reads and writes, with the same
! names, kinds and shapes: the frame (r(ndim,Nsites), itype, sidel, side2,
! volumen), the sizes (natms, Nsites, nbcuda, lsmax, nit, nqmax, nqmin) and
! the results (histomix, gtot, sqf, sqfp, nq, rmingr).  transfer_cpu_gpu
! hands the frame to the trj_hip library, which runs the frame's RDF and
! S(Q) work for the modules initialised; RDFcomp and SQcalc collect it.
!
! Not carried: the device arrays (r_d, itype_d, histomix_d, sqf_d: the
! library holds its own), slices (twoDstruc_3D), charges (ex_qc) and number
! fluctuations (nrandom > 0), which the trj_hip_rdf / trj_hip_sq modules and
! the C ABI provide; mod_rdf / mod_sq stop when they are set.
module mod_common
  use iso_c_binding
  use mod_precision
  use trj_hip
  implicit none
  integer, parameter :: nspmax = 10
  integer, parameter :: nitmax = nspmax*(nspmax + 1)/2
  real(myprec), parameter :: pi = 3.1415926535898D0
  integer :: nbcuda, natms, Nsites = 0, nqmax, lsmax, nqmin, ndr, nit, nconf, nrandom = 0
  integer, allocatable, dimension(:) :: itype, nq
  real(myprec) :: volumen, side, side2, dq, rmingr
  real(myprec), dimension(3) :: sidel, sidelv
  real, dimension(3) :: fk
  real(myprec), allocatable, dimension(:) :: gtot
  real(myprec), allocatable, target, dimension(:, :) :: r
  real, allocatable, dimension(:) :: sqf
  real, allocatable, dimension(:, :) :: sqfp
  real, allocatable, dimension(:, :, :) :: histomix
  logical :: run_rdf = .false., run_sq = .false., ex_qc = .false., periodic(3) = .true., &
             twoDstruc_3D = .false.
  ! The library context shared by mod_rdf and mod_sq, and which of the two
  ! were initialised (no counterpart in the reference).
  type(c_ptr), save :: trj_ref_ctx = c_null_ptr
  logical, save :: trj_rdf_on = .false., trj_sq_on = .false.
contains
  ! Creates the library context on device 0 the first time it is needed.
  subroutine trj_ref_context()
    if (.not. c_associated(trj_ref_ctx)) call trj_ref_check(trj_create(trj_ref_ctx, 0_c_int), "trj_create")
  end subroutine trj_ref_context

  ! Stops with the library's message when a call fails (the reference stops on
  ! kernel errors, rdf.cuf:239).
  subroutine trj_ref_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    call trj_check(trj_ref_ctx, rc, what)
  end subroutine trj_ref_check

  ! transfer_cpu_gpu (common.cuf:368-378): r(1:ndim,1:Nsites), itype and the
  ! cell lengths of the current frame to the device.  The frame's RDF and
  ! S(Q) work for the initialised modules starts here (asynchronously; the
  ! host arrays may change once this returns).
  subroutine transfer_cpu_gpu(ndim)
    integer, intent(in) :: ndim
    integer(c_int32_t) :: what
    real(c_double) :: box(3)
    real(c_float), allocatable, target :: rc(:, :)
    what = 0
    if (trj_rdf_on) what = ior(what, TRJ_WANT_RDF)
    if (trj_sq_on) what = ior(what, TRJ_WANT_SQ)
    if (what == 0) return
    box = real(sidel, c_double)
    if (size(r, 1) == ndim) then
      call trj_ref_check(trj_frame(trj_ref_ctx, c_loc(r), TRJ_F32, TRJ_AOS, itype, int(Nsites, c_int64_t), box, &
                                   what), "trj_frame")
    else
      rc = r(1:ndim, 1:Nsites)
      call trj_ref_check(trj_frame(trj_ref_ctx, c_loc(rc), TRJ_F32, TRJ_AOS, itype, int(Nsites, c_int64_t), box, &
                                   what), "trj_frame")
    end if
  end subroutine transfer_cpu_gpu

  ! Destroys the library context once neither module uses it.
  subroutine trj_ref_release()
    if (.not. trj_rdf_on .and. .not. trj_sq_on .and. c_associated(trj_ref_ctx)) then
      call trj_destroy(trj_ref_ctx)
      trj_ref_ctx = c_null_ptr
    end if
  end subroutine trj_ref_release
end module mod_common

! mod_sq -- the reference's S(Q) module (src/modules/sq.cuf) with its own
! procedure names and argument lists over the trj_hip library:
!   sq_init(Nsites, nsp, nbcuda)   sq.cuf:150-370
!   SQcalc()                       sq.cuf:58-126
!   sq_transfer_gpu_cpu()          sq.cuf:128-148
!   sq_clear()                     sq.cuf:774
! sq_init sets the module variables the reference sets (dq, nqmax, nqmin,
! fk, and qmin moved to the edge of the fully sampled region) from the
! library's plan of the same q-set, and allocates sqf(0:nqmax),
! sqfp(0:nqmax,nsp) and nq(0:nqmax).  The contraction of each frame runs on
! the device when transfer_cpu_gpu (mod_common) hands the frame over, so
! SQcalc has nothing left to launch; sq_transfer_gpu_cpu copies the
! accumulated shell sums (FP64 in the library) into sqf and sqfp.
module mod_sq
  use iso_c_binding
  use mod_precision
  use mod_common
  use mod_input
  use trj_hip
  implicit none
contains

  subroutine sq_init(Nsites, nsp, nbcuda)
    implicit none
    integer, intent(IN) :: Nsites, nsp, nbcuda
    type(trj_sq_params) :: p
    type(trj_sq_info) :: info
    real(c_double) :: box(3)
    real(c_double), allocatable :: s(:), sp(:, :)
    integer(c_int64_t), allocatable :: n64(:)
    if (twoDstruc_3D .or. ex_qc) then
      print *, " *** mod_sq: slices and charges go through the trj_hip C ABI"
      stop 1
    end if
    if (nsp > nspmax) then
      print *, " *** mod_sq: nsp larger than nspmax"
      stop 1
    end if
    call trj_ref_context()
    p%ndim = ndim
    p%nsp = nsp
    p%qmax = real(qmax, c_double)
    p%qmin = real(qmin, c_double)
    p%bsc(:) = 0.0d0
    p%bsc(1:nsp) = real(bsc(1:nsp), c_double)
    p%qset = TRJ_QSET_REF
    p%grid(:) = 0
    p%precision = TRJ_PREC_FP64
    p%flags = 0
    box = real(sidel, c_double)
    call trj_ref_check(trj_sq_init(trj_ref_ctx, p, box, info), "trj_sq_init")
    dq = real(info%dq, myprec)
    nqmax = info%nqmax
    nqmin = info%nqmin
    fk(:) = real(info%fk(:))
    qmin = real(info%qmin_eff, myprec)
    if (allocated(sqf)) deallocate (sqf)
    if (allocated(sqfp)) deallocate (sqfp)
    if (allocated(nq)) deallocate (nq)
    allocate (sqf(0:nqmax), sqfp(0:nqmax, nsp), nq(0:nqmax))
    sqf(:) = 0.0
    sqfp(:, :) = 0
    ! nq(:): vectors per shell, as sq_init counts them
    allocate (s(0:nqmax), sp(0:nqmax, nsp), n64(0:nqmax))
    call trj_ref_check(trj_sq_shells(trj_ref_ctx, s, sp, n64), "trj_sq_shells")
    nq(:) = int(n64(:))
    trj_sq_on = .true.
  end subroutine sq_init

  subroutine SQcalc()
    implicit none
    if (.not. trj_sq_on) then
      print *, " *** SQcalc: sq_init has not been called"
      stop 1
    end if
  end subroutine SQcalc

  subroutine sq_transfer_gpu_cpu()
    implicit none
    real(c_double), allocatable :: s(:), sp(:, :)
    integer(c_int64_t), allocatable :: n64(:)
    integer :: nspq
    if (.not. trj_sq_on) return
    nspq = size(sqfp, 2)
    allocate (s(0:nqmax), sp(0:nqmax, nspq), n64(0:nqmax))
    call trj_ref_check(trj_sq_shells(trj_ref_ctx, s, sp, n64), "trj_sq_shells")
    sqfp(0:nqmax, 1:nspq) = real(sp(0:nqmax, 1:nspq))
    sqf(0:nqmax) = real(s(0:nqmax))
  end subroutine sq_transfer_gpu_cpu

  subroutine sq_clear()
    implicit none
    if (allocated(sqf)) deallocate (sqf)
    if (allocated(sqfp)) deallocate (sqfp)
    if (allocated(nq)) deallocate (nq)
    trj_sq_on = .false.
    call trj_ref_release()
  end subroutine sq_clear
end module mod_sq

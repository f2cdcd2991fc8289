! trj_hip_modules.f90 -- the reference's mod_rdf / mod_sq procedures over the
! trj_hip C ABI.  Same procedure names and argument meaning as
! src/modules/rdf.cuf (RDF_init 51, RDFcomp 184, rdf_clear 124) and
! src/modules/sq.cuf (sq_init 150, SQcalc 58, sq_transfer_gpu_cpu 128,
! sq_clear 774); the module globals the reference reads from mod_common /
! mod_input (r, itype, sidel, deltar, rcrdf, qmax, qmin, bsc, periodic) are
! passed explicitly.  Results land in module arrays with the reference's
! names and layouts: histomix(lsmax,nsp,nsp) (exact int64 counts, symmetric),
! sqf(0:nqmax), sqfp(0:nqmax,nsp), nq(0:nqmax).
module trj_hip_rdf
  use iso_c_binding
  use trj_hip
  implicit none
  type(c_ptr), save :: rdf_ctx = c_null_ptr
  integer, save :: lsmax = 0, nit = 0, nsp_rdf = 0
  integer(c_int64_t), allocatable, save :: histomix(:, :, :)
contains
  subroutine RDF_init(nsp, ndim, deltar, rcrdf, sidel, periodic, device)
    integer, intent(in) :: nsp, ndim
    real(c_float), intent(in) :: deltar, rcrdf, sidel(3)
    logical, intent(in) :: periodic(3)
    integer, intent(in), optional :: device
    type(trj_rdf_params) :: p
    real(c_double) :: box(3)
    integer(c_int32_t) :: ls, ni
    integer :: k, dev
    dev = 0
    if (present(device)) dev = device
    if (.not. c_associated(rdf_ctx)) call trj_check(rdf_ctx, trj_create(rdf_ctx, int(dev, c_int)), "trj_create")
    p%ndim = ndim
    p%nsp = nsp
    p%deltar = real(deltar, c_double)
    p%rcrdf = real(rcrdf, c_double)
    do k = 1, 3
      p%periodic(k) = merge(1, 0, periodic(k))
    end do
    box = real(sidel, c_double)
    call trj_check(rdf_ctx, trj_rdf_init(rdf_ctx, p, box, ls, ni), "trj_rdf_init")
    lsmax = ls
    nit = ni
    nsp_rdf = nsp
    if (allocated(histomix)) deallocate (histomix)
    allocate (histomix(lsmax, nsp, nsp))
    histomix = 0
  end subroutine

  ! One frame: r(ndim,Nsites) as in the reference (common.cuf:52), itype 1..nsp.
  subroutine RDFcomp(Nsites, r, itype, sidel)
    integer, intent(in) :: Nsites
    real(c_float), intent(in), target :: r(:, :)
    integer(c_int32_t), intent(in) :: itype(:)
    real(c_float), intent(in) :: sidel(3)
    real(c_double) :: box(3)
    box = real(sidel, c_double)
    call trj_check(rdf_ctx, trj_frame(rdf_ctx, c_loc(r), TRJ_F32, TRJ_AOS, itype, int(Nsites, c_int64_t), &
                                      box, TRJ_WANT_RDF), "trj_frame")
  end subroutine

  ! The host reduction of RDFcomp (rdf.cuf:301-310): fills histomix(:,k,l).
  subroutine rdf_histomix()
    integer(c_int64_t), allocatable :: counts(:)
    integer :: k, l, ind
    allocate (counts(lsmax * nit))
    call trj_check(rdf_ctx, trj_rdf_counts(rdf_ctx, counts), "trj_rdf_counts")
    ind = 0
    do k = 1, nsp_rdf
      do l = k, nsp_rdf
        histomix(:, k, l) = counts(ind * lsmax + 1:(ind + 1) * lsmax)
        histomix(:, l, k) = histomix(:, k, l)
        ind = ind + 1
      end do
    end do
  end subroutine

  subroutine rdf_clear()
    if (c_associated(rdf_ctx)) call trj_check(rdf_ctx, trj_reset(rdf_ctx), "trj_reset")
    if (allocated(histomix)) deallocate (histomix)
  end subroutine
end module trj_hip_rdf

module trj_hip_sq
  use iso_c_binding
  use trj_hip
  use trj_hip_rdf, only: rdf_ctx
  implicit none
  integer, save :: nqmax = 0, nqmin = 0, nsp_sq = 0
  real(c_double), save :: dq = 0.0d0
  real(c_double), allocatable, save :: sqf(:), sqfp(:, :)
  integer(c_int64_t), allocatable, save :: nq(:)
contains
  subroutine sq_init(Nsites, nsp, ndim, qmax, qmin, bsc, sidel)
    integer, intent(in) :: Nsites, nsp, ndim
    real(c_float), intent(in) :: qmax, qmin, bsc(:), sidel(3)
    type(trj_sq_params) :: p
    type(trj_sq_info) :: info
    real(c_double) :: box(3)
    if (.not. c_associated(rdf_ctx)) call trj_check(rdf_ctx, trj_create(rdf_ctx, 0_c_int), "trj_create")
    p%ndim = ndim
    p%nsp = nsp
    p%qmax = real(qmax, c_double)
    p%qmin = real(qmin, c_double)
    p%bsc = 1.0d0
    p%bsc(1:nsp) = real(bsc(1:nsp), c_double)
    p%qset = TRJ_QSET_REF
    p%grid = 0
    p%precision = TRJ_PREC_FP64
    p%flags = 0
    box = real(sidel, c_double)
    call trj_check(rdf_ctx, trj_sq_init(rdf_ctx, p, box, info), "trj_sq_init")
    nqmax = info%nqmax
    nqmin = info%nqmin
    dq = info%dq
    nsp_sq = nsp
    if (allocated(sqf)) deallocate (sqf, sqfp, nq)
    allocate (sqf(0:nqmax), sqfp(0:nqmax, nsp), nq(0:nqmax))
    sqf = 0
    sqfp = 0
    nq = 0
    if (Nsites < 1) stop "sq_init: Nsites < 1"
  end subroutine

  subroutine SQcalc(Nsites, r, itype, sidel)
    integer, intent(in) :: Nsites
    real(c_float), intent(in), target :: r(:, :)
    integer(c_int32_t), intent(in) :: itype(:)
    real(c_float), intent(in) :: sidel(3)
    real(c_double) :: box(3)
    box = real(sidel, c_double)
    call trj_check(rdf_ctx, trj_frame(rdf_ctx, c_loc(r), TRJ_F32, TRJ_AOS, itype, int(Nsites, c_int64_t), &
                                      box, TRJ_WANT_SQ), "trj_frame")
  end subroutine

  subroutine sq_transfer_gpu_cpu()
    real(c_double), allocatable :: fp(:)
    integer :: s
    allocate (fp((nqmax + 1) * nsp_sq))
    call trj_check(rdf_ctx, trj_sq_shells(rdf_ctx, sqf, fp, nq), "trj_sq_shells")
    do s = 1, nsp_sq
      sqfp(0:nqmax, s) = fp((s - 1) * (nqmax + 1) + 1:s * (nqmax + 1))
    end do
  end subroutine

  subroutine sq_clear()
    if (c_associated(rdf_ctx)) call trj_check(rdf_ctx, trj_reset(rdf_ctx), "trj_reset")
    if (allocated(sqf)) deallocate (sqf, sqfp, nq)
  end subroutine
end module trj_hip_sq

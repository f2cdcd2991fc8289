! mod_rdf -- the reference's RDF module (src/modules/rdf.cuf) with its own
! procedure names and argument lists over the trj_hip library:
!   RDF_init(nsp)                                 rdf.cuf:51-122
!   RDFcomp(Nsites, Iconf, nbcuda, nthread, units) rdf.cuf:184-353
!   rdf_clear()                                   rdf.cuf:124-182
! The pair histogram of each frame is computed on the device when
! transfer_cpu_gpu (mod_common) hands the frame over; RDFcomp then does the
! host part of the reference's RDFcomp on the module variables: histomix
! (lsmax, nsp, nsp), filled for l >= k as rdf.cuf:301-310 does, gtot(1:lsmax)
! (rdf.cuf:311-341) and rmingr (find_first_minimum_interp, rdf.cuf:343).
! histomix holds the accumulated counts (exact int64 in the library, rounded
! once to real here rather than summed frame by frame in real).
! nbcuda and nthread size the reference's launch and are not needed here.
module mod_rdf
  use iso_c_binding
  use mod_precision
  use mod_common
  use mod_input
  use trj_hip
  implicit none
contains

  subroutine RDF_init(nsp)
    implicit none
    integer, intent(in) :: nsp
    type(trj_rdf_params) :: p
    real(c_double) :: box(3)
    integer(c_int32_t) :: ls, ni
    integer :: k
    if (twoDstruc_3D .or. ex_qc .or. nrandom > 0) then
      print *, " *** mod_rdf: slices, charges and number fluctuations go through the trj_hip C ABI"
      stop 1
    end if
    call trj_ref_context()
    p%ndim = ndim
    p%nsp = nsp
    p%deltar = real(deltar, c_double)
    p%rcrdf = real(rcrdf, c_double)
    do k = 1, 3
      p%periodic(k) = merge(1, 0, periodic(k))
    end do
    p%flags = 0
    box = real(sidel, c_double)
    ! lsmax = nint(rcrdf/deltar) + 1, or from half the smallest box side when
    ! rcrdf <= 0 (rdf.cuf:54-59); the library's histogram needs no
    ! shared-memory check (rdf.cuf:61-66): it tiles its LDS.
    call trj_ref_check(trj_rdf_init(trj_ref_ctx, p, box, ls, ni), "trj_rdf_init")
    lsmax = ls
    ndr = lsmax
    nit = ni
    if (allocated(histomix)) deallocate (histomix)
    if (allocated(gtot)) deallocate (gtot)
    allocate (histomix(lsmax, nsp, nsp), gtot(0:lsmax))
    histomix(:, :, :) = 0
    gtot(:) = 0
    trj_rdf_on = .true.
  end subroutine RDF_init

  subroutine RDFcomp(Nsites, Iconf, nbcuda, nthread, units)
    implicit none
    integer, intent(IN) :: Nsites, Iconf, nbcuda, nthread
    character(len=*), intent(IN) :: units
    integer(c_int64_t), allocatable :: counts(:, :)
    real(c_double) :: rmin
    real :: deltaV, ri
    integer :: i, j, k, l, ind, nspr
    if (.not. trj_rdf_on) then
      print *, " *** RDFcomp: RDF_init has not been called"
      stop 1
    end if
    nspr = size(histomix, 2)
    ! Accumulated counts, counts(1:lsmax, ind) with ind running over k <= l.
    allocate (counts(lsmax, nit))
    call trj_ref_check(trj_rdf_counts(trj_ref_ctx, counts), "trj_rdf_counts")
    ind = 1
    do k = 1, nspr
      do l = k, nspr
        histomix(1:lsmax, k, l) = real(counts(:, ind))
        ind = ind + 1
      end do
    end do
    ! Total g(r) of the accumulated histograms (rdf.cuf:311-341, 3D and 2D).
    do i = 1, lsmax
      ri = real(i)*deltar
      if (ndim == 3) then
        gtot(i) = 0
        deltaV = 4*pi*((ri + deltar/2)**3 - (ri - deltar/2)**3)/3.0
      else
        gtot(i) = 0
        deltaV = pi*((ri + deltar/2)**2 - (ri - deltar/2)**2)
      end if
      do j = 1, nspr
        do l = j, nspr
          gtot(i) = gtot(i) + histomix(i, j, l)
        end do
      end do
      gtot(i) = 2*volumen*gtot(i)/(deltaV*natms*natms)
    end do
    call trj_ref_check(trj_rdf_first_minimum(trj_ref_ctx, rmin), "trj_rdf_first_minimum")
    rmingr = real(rmin, myprec)
    if (Iconf == 1) then
      if (units == 'lj') then
        write (*, "(/' *** Suggested rcl (NN and/or connectivity distance): ',f10.4,' σ')") rmingr
      else
        write (*, "(/' *** Suggested rcl (NN and/or connectivity distance): ',f10.4,' Å')") rmingr
      end if
    end if
  end subroutine RDFcomp

  subroutine rdf_clear()
    implicit none
    if (allocated(histomix)) deallocate (histomix)
    if (allocated(gtot)) deallocate (gtot)
    trj_rdf_on = .false.
    call trj_ref_release()
  end subroutine rdf_clear
end module mod_rdf

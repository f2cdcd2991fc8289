! mod_input -- the namelist variables of the reference's mod_input
! (src/io/input.f90:51-70) that the RDF and S(Q) modules read: /INPUT/
! ndim, nsp, nthread, idir; /INPUT_RDF/ deltar, rcrdf; /INPUT_SQ/ qmax, qmin,
! bsc.  Same names, kinds and defaults; filled by the caller (the reference
! reads them with read_input).
module mod_input
  use mod_precision
  implicit none
  integer :: nthread = 64, ndim, idir = 3, nsp
  real(myprec) :: deltar, qmin, qmax, rcrdf
  real(myprec), allocatable, dimension(:) :: bsc
end module mod_input

! mod_precision -- the kinds of the reference's mod_precision
! (src/core/precision.f90:1-8): single precision is the working kind.
module mod_precision
  use, intrinsic :: iso_fortran_env, only: sp => real32, dp => real64
  implicit none
  integer, parameter :: myprec = sp
  integer, parameter :: float = sp
  integer, parameter :: double = dp
end module mod_precision

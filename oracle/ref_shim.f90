! C entry points over routines compiled from the reference's own sources
! (test infrastructure only: the tests compare the oracle with them).
!
! The Makefile target `ref` compiles, where the reference tree is present,
! common/puffin_kinds.f90, common/puffin_constants.f90 and
! utilities/gamma_to_p2.f90 of the reference with this file into
! oracle/_ref/libpuffin_ref.so.  These are the only modules of the path that
! compile on their own (everything else needs MPI and FFTW3-MPI).
module ref_shim
  use iso_c_binding
  use puffin_kinds, only: wp
  use puffin_constants, only: pi, c, q_e
  use gtop2, only: getP2, getGamma
  implicit none
contains

  ! getP2 (gamma_to_p2.f90:47-87) over n macroparticles
  subroutine ref_get_p2(n, gam, px, py, eta, gamma0, aw, p2) bind(C, name="ref_get_p2")
    integer(c_int64_t), value :: n
    real(c_double), intent(in) :: gam(n), px(n), py(n)
    real(c_double), value :: eta, gamma0, aw
    real(c_double), intent(out) :: p2(n)
    call getP2(p2, gam, px, py, eta, gamma0, aw)
  end subroutine ref_get_p2

  ! getGamma (gamma_to_p2.f90:22-43) over n macroparticles
  subroutine ref_get_gamma(n, p2, px, py, eta, gamma0, aw, gam) bind(C, name="ref_get_gamma")
    integer(c_int64_t), value :: n
    real(c_double), intent(in) :: p2(n), px(n), py(n)
    real(c_double), value :: eta, gamma0, aw
    real(c_double), intent(out) :: gam(n)
    call getGamma(gam, p2, px, py, eta, gamma0, aw)
  end subroutine ref_get_gamma

  ! the constants of puffin_constants.f90 the path uses: pi, c, q_e
  subroutine ref_constants(out) bind(C, name="ref_constants")
    real(c_double), intent(out) :: out(3)
    out(1) = pi
    out(2) = c
    out(3) = q_e
  end subroutine ref_constants

end module ref_shim

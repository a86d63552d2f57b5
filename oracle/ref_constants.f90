! Prints the physical constants and tunables of the reference's mod_con.f90
! (compiled from the reference's own mod_kind.f90 and mod_con.f90, without
! -r8, as its active Makefile builds them) so the oracle's constant table
! can be checked against what the reference's compiler makes of them.
! Test infrastructure only.
program ref_constants
  use mod_kind, only: wp
  use mod_con
  implicit none
  call put('rho0', rho0)
  call put('a0', a0)
  call put('g', g)
  call put('a', a)
  call put('pi', pi)
  call put('omega', omega)
  call put('torad', torad)
  call put('gamma_t', gamma_t)
  call put('gamma_s', gamma_s)
  call put('gamma_b', gamma_b)
  call put('deltat', deltat)
  call put('deltas', deltas)
  call put('cdbot', cdbot)
  call put('tice', tice)
  call put('km_c', km_c)
  call put('am_c', am_c)
  call put('smag_c', smag_c)
  call put('ah_c', ah_c)
contains
  subroutine put(name, v)
    character(len=*), intent(in) :: name
    real(kind=wp), intent(in) :: v
    write(*, '(a, 1x, z16.16)') name, v
  end subroutine put
end program ref_constants

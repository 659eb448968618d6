!! Fortran host driver: a perfect 3x3x3 diamond Si crystal through
!! init / bind_to / energy_and_forces of the atomistica_gpu module.
!! Prints "E/atom <value>" and the largest force component, then the same
!! for the screened variant (TersoffScr).
program tersoff_check
  use, intrinsic :: iso_c_binding
  use atomistica_gpu
  implicit none
  integer, parameter :: n = 3, nat = 8 * n**3
  real(c_double), parameter :: a = 5.431d0
  real(c_double) :: basis(3, 8), r(3, nat), cell(3, 3), f(3, nat), wpot(3, 3), epot
  real(c_double) :: epa(nat)
  integer :: Z(nat), ix, iy, iz, b, k, ierr
  type(particles_gpu_t) :: p
  type(neighbors_gpu_t) :: nl
  type(potential_gpu_t) :: pot, pscr
  type(neighbors_gpu_t) :: nl2

  basis = reshape([0d0, 0d0, 0d0, 0d0, .5d0, .5d0, .5d0, 0d0, .5d0, .5d0, .5d0, 0d0, &
                   .25d0, .25d0, .25d0, .25d0, .75d0, .75d0, .75d0, .25d0, .75d0, .75d0, .75d0, .25d0], [3, 8])
  k = 0
  do ix = 0, n - 1
     do iy = 0, n - 1
        do iz = 0, n - 1
           do b = 1, 8
              k = k + 1
              r(:, k) = a * ([real(ix, c_double), real(iy, c_double), real(iz, c_double)] + basis(:, b))
           end do
        end do
     end do
  end do
  Z = 14
  cell = 0d0
  cell(1, 1) = n * a
  cell(2, 2) = n * a
  cell(3, 3) = n * a

  call particles_init(p, nat, ierr); call check(ierr)
  call particles_set_cell(p, cell, [.true., .true., .true.], ierr); call check(ierr)
  call particles_set_Z(p, Z, ierr); call check(ierr)
  call particles_set_positions(p, r, ierr); call check(ierr)
  call neighbors_init(nl, 100, ierr); call check(ierr)
  call tersoff_init(pot, ierr); call check(ierr)
  call bind_to(pot, p, nl, ierr); call check(ierr)
  f = 0d0
  wpot = 0d0
  epa = 0d0
  call energy_and_forces(pot, p, nl, epot, f, wpot, ierr, epot_per_at=epa); call check(ierr)
  write (*, '(A,F20.12)') 'E/atom ', epot / nat
  write (*, '(A,ES12.4)') 'max|f| ', maxval(abs(f))
  write (*, '(A,ES12.4)') 'sum(epot_per_at)-epot ', sum(epa) - epot
  ! the screened variant: in a perfect crystal every second neighbour is
  ! screened and the bonds are equal, so the energy is the same
  call neighbors_init(nl2, 200, ierr); call check(ierr)
  call tersoff_scr_init(pscr, ierr); call check(ierr)
  call bind_to(pscr, p, nl2, ierr); call check(ierr)
  f = 0d0
  wpot = 0d0
  call energy_and_forces(pscr, p, nl2, epot, f, wpot, ierr); call check(ierr)
  write (*, '(A,F20.12)') 'E/atom TersoffScr ', epot / nat
  write (*, '(A,ES12.4)') 'max|f| TersoffScr ', maxval(abs(f))
  call del(pscr)
  call del(nl2)
  call del(pot)
  call del(nl)
  call del(p)

contains

  subroutine check(ierr)
    integer, intent(in) :: ierr
    if (ierr /= AGPU_OK) then
       write (*, '(A,I3,2A)') 'error ', ierr, ': ', trim(last_error())
       stop 1
    end if
  end subroutine check

end program tersoff_check

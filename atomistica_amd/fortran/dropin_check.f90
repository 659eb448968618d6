!! Fortran host driver of the reference-shaped types (atomistica_gpu_dropin):
!! reads a structure (nat; Abox column by column; pbc; Z x y z per atom)
!! from the file named on the command line, fills particles_t itself, and
!! evaluates Tersoff through bind_to / energy_and_forces with per-atom and
!! per-bond arrays; then moves atom 1 (I_changed_positions) and evaluates
!! again.  Prints the results for the caller to compare.
program dropin_check
  use, intrinsic :: iso_c_binding
  use atomistica_gpu, only: tersoff_init
  use atomistica_gpu_dropin
  implicit none
  integer, parameter :: DP = c_double
  type(particles_t) :: p
  type(neighbors_t) :: nl
  type(potential_t) :: pot
  character(len=512) :: fn
  integer :: nat, i, ierr, pbc(3), nb
  real(DP) :: Abox(3, 3), epot, wpot(3, 3)
  real(DP), allocatable :: f(:, :), epa(:), wpa(:, :, :), epb(:), fpb(:, :), wpb(:, :, :)

  call get_command_argument(1, fn)
  open(10, file=trim(fn), status='old', action='read')
  read(10, *) nat
  read(10, *) Abox
  read(10, *) pbc
  call init(p)
  call allocate(p, nat)
  do i = 1, nat
     read(10, *) p%Z(i), p%r_non_cyc(:, i)
  end do
  close(10)
  call set_cell(p, Abox, pbc /= 0)
  call update_elements(p)
  call I_changed_positions(p)

  call init(nl, avgn=100)
  call tersoff_init(pot%gpu, ierr)
  call check(ierr)
  call bind_to(pot, p, nl, ierr)
  call check(ierr)
  nb = nl%neighbors_size
  allocate(f(3, nat), epa(nat), wpa(3, 3, nat), epb(nb), fpb(3, nb), wpb(3, 3, nb))

  call evaluate('A')
  ! move atom 1 the way a host moves atoms: in place, then signal
  p%r_non_cyc(:, 1) = p%r_non_cyc(:, 1) + [0.013d0, -0.021d0, 0.017d0]
  call I_changed_positions(p)
  call evaluate('B')
  call del(p)
  call del(nl)

contains

  subroutine evaluate(tag)
    character(len=1), intent(in) :: tag
    integer :: k
    f = 0
    wpot = 0
    epa = 0
    wpa = 0
    epb = 0
    fpb = 0
    wpb = 0
    call energy_and_forces(pot, p, nl, epot, f, wpot, epot_per_at=epa, epot_per_bond=epb, &
         f_per_bond=fpb, wpot_per_at=wpa, wpot_per_bond=wpb, ierror=ierr)
    call check(ierr)
    write(*, '(a, 1x, a, 1x, es25.17)') tag, 'E', epot
    write(*, '(a, 1x, a, 9(1x, es25.17))') tag, 'W', wpot
    write(*, '(a, 1x, a, 1x, es25.17)') tag, 'sum_epb', sum(epb)
    write(*, '(a, 1x, a, 1x, es25.17)') tag, 'sum_epa', sum(epa)
    write(*, '(a, 1x, a, 9(1x, es25.17))') tag, 'sum_wpa', sum(wpa, dim=3)
    write(*, '(a, 1x, a, 9(1x, es25.17))') tag, 'sum_wpb', sum(wpb, dim=3)
    do k = 1, nat
       write(*, '(a, 1x, a, 1x, i6, 4(1x, es25.17))') tag, 'F', k, f(:, k), epa(k)
    end do
  end subroutine evaluate

  subroutine check(rc)
    integer, intent(in) :: rc
    if (rc /= 0) then
       write(*, *) 'error', rc
       stop 1
    end if
  end subroutine check

end program dropin_check

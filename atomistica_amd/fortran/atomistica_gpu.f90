!! Fortran host interface of libatomistica_amd (gfx950).
!!
!! Keeps the class interface of the reference's potentials --
!! init / bind_to / energy_and_forces (e.g. tersoff_module.f90:30-78,
!! default_bind_to_func.f90:25-146, default_compute_func.f90:25-100) -- and
!! reaches the HIP kernels through the flat C-ABI of include/atomistica_amd.h
!! with iso_c_binding.  Arrays are the reference's Fortran shapes:
!! r(3,nat), f(3,nat), cell(3,3) with cell(:,k) = lattice vector k,
!! wpot(3,3), epot_per_at(nat), wpot_per_at(3,3,nat).  Host arrays are staged
!! through HBM by the library.
module atomistica_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter, public :: AGPU_OK = 0
  integer(c_int), parameter :: AGPU_HOST = 1
  integer, parameter, public :: AGPU_PREC_FP64 = 0, AGPU_PREC_MIXED = 1

  type, public :: particles_gpu_t
     type(c_ptr) :: h = c_null_ptr
     integer :: nat = 0
  end type particles_gpu_t

  type, public :: neighbors_gpu_t
     type(c_ptr) :: h = c_null_ptr
  end type neighbors_gpu_t

  type, public :: potential_gpu_t
     type(c_ptr) :: h = c_null_ptr
  end type potential_gpu_t

  !> Tersoff parameter database (BOP_DB_TYPE of tersoff_params.f90:32-78)
  type, bind(C), public :: tersoff_params_t
     integer(c_int) :: nel
     integer(c_int) :: Z(3)
     real(c_double) :: A(6), B(6), xi(6), lambda(6), mu(6), omega(6), mubo(6)
     integer(c_int) :: m(6)
     real(c_double) :: beta(3), n(3), c(3), d(3), h(3)
     real(c_double) :: r1(6), r2(6)
  end type tersoff_params_t

  !> Kumagai parameters (BOP_DB_TYPE of kumagai_params.f90:34-94, one element)
  type, bind(C), public :: kumagai_params_t
     integer(c_int) :: Z
     real(c_double) :: A, B, lambda1, lambda2, eta, delta, alpha
     integer(c_int) :: beta
     real(c_double) :: c1, c2, c3, c4, c5, h, r1, r2
  end type kumagai_params_t

  !> Brenner parameters (BOP_DB_TYPE of brenner_params.f90:34-71, no screening)
  type, bind(C), public :: brenner_params_t
     integer(c_int) :: nel
     integer(c_int) :: Z(3)
     real(c_double) :: D0(6), r0(6), S(6), beta(6), gamma(6), c(6), d(6), h(6), mu(6), n(6)
     integer(c_int) :: m(6)
     real(c_double) :: r1(6), r2(6)
  end type brenner_params_t

  !> Juslin parameters (BOP_DB_TYPE of juslin_params.f90:26-72): ordered
  !> element pairs j + i nel, ordered triplets k + nel (j + nel i)
  type, bind(C), public :: juslin_params_t
     integer(c_int) :: nel
     integer(c_int) :: Z(3)
     real(c_double) :: D0(9), r0(9), S(9), beta(9), gamma(9), c(9), d(9), h(9), n(9)
     real(c_double) :: alpha(27), omega(27)
     integer(c_int) :: m(27)
     real(c_double) :: r1(9), r2(9)
  end type juslin_params_t

  !> Screening parameters of the screened variants (tersoff_params.f90:82-89):
  !> outer and bond-order cutoffs and Cmin, Cmax per element pair
  type, bind(C), public :: screening_params_t
     ! per element pair: symmetric index (6 used), ordered for Juslin (9)
     real(c_double) :: or1(9), or2(9), bor1(9), bor2(9), Cmin(9), Cmax(9)
  end type screening_params_t

  public :: particles_init, particles_set_cell, particles_set_Z, particles_set_positions
  public :: neighbors_init, tersoff_init, eam_alloy_init, ewald_init, rebo2_init
  public :: kumagai_init, brenner_init
  public :: tersoff_scr_init, kumagai_scr_init, brenner_scr_init, pme_init, juslin_init, juslin_scr_init
  public :: set_precision, bind_to, energy_and_forces, coulomb_energy_and_forces
  public :: last_error, del

  interface del
     module procedure particles_del, neighbors_del, potential_del
  end interface del

  interface
     integer(c_int) function agpu_get_error(buf, len) bind(C, name="agpu_get_error")
       import :: c_int, c_char
       character(kind=c_char) :: buf(*)
       integer(c_int), value :: len
     end function
     integer(c_int) function agpu_particles_create(p) bind(C, name="agpu_particles_create")
       import :: c_int, c_ptr
       type(c_ptr) :: p
     end function
     integer(c_int) function agpu_particles_destroy(p) bind(C, name="agpu_particles_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: p
     end function
     integer(c_int) function agpu_particles_allocate(p, nat) bind(C, name="agpu_particles_allocate")
       import :: c_int, c_ptr
       type(c_ptr), value :: p
       integer(c_int), value :: nat
     end function
     integer(c_int) function agpu_particles_set_cell(p, cell, pbc) bind(C, name="agpu_particles_set_cell")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: p
       real(c_double) :: cell(3, 3)
       integer(c_int) :: pbc(3)
     end function
     integer(c_int) function agpu_particles_set_atomic_numbers(p, Z, flags) &
          bind(C, name="agpu_particles_set_atomic_numbers")
       import :: c_int, c_ptr
       type(c_ptr), value :: p
       integer(c_int) :: Z(*)
       integer(c_int), value :: flags
     end function
     integer(c_int) function agpu_particles_set_positions(p, r, flags) &
          bind(C, name="agpu_particles_set_positions")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: p
       real(c_double) :: r(3, *)
       integer(c_int), value :: flags
     end function
     integer(c_int) function agpu_neighbors_create(nl, avgn) bind(C, name="agpu_neighbors_create")
       import :: c_int, c_ptr
       type(c_ptr) :: nl
       integer(c_int), value :: avgn
     end function
     integer(c_int) function agpu_neighbors_destroy(nl) bind(C, name="agpu_neighbors_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: nl
     end function
     integer(c_int) function agpu_tersoff_create(pot, params) bind(C, name="agpu_tersoff_create")
       import :: c_int, c_ptr
       type(c_ptr) :: pot
       type(c_ptr), value :: params
     end function
     integer(c_int) function agpu_kumagai_create(pot, params) bind(C, name="agpu_kumagai_create")
       import :: c_int, c_ptr
       type(c_ptr) :: pot
       type(c_ptr), value :: params
     end function
     integer(c_int) function agpu_brenner_create(pot, params) bind(C, name="agpu_brenner_create")
       import :: c_int, c_ptr
       type(c_ptr) :: pot
       type(c_ptr), value :: params
     end function
     integer(c_int) function agpu_juslin_create(pot, params) bind(C, name="agpu_juslin_create")
       import :: c_int, c_ptr
       type(c_ptr) :: pot
       type(c_ptr), value :: params
     end function
     integer(c_int) function agpu_tersoff_scr_create(pot, params, scr) bind(C, name="agpu_tersoff_scr_create")
       import :: c_int, c_ptr
       type(c_ptr) :: pot
       type(c_ptr), value :: params, scr
     end function
     integer(c_int) function agpu_kumagai_scr_create(pot, params, scr) bind(C, name="agpu_kumagai_scr_create")
       import :: c_int, c_ptr
       type(c_ptr) :: pot
       type(c_ptr), value :: params, scr
     end function
     integer(c_int) function agpu_brenner_scr_create(pot, params, scr) bind(C, name="agpu_brenner_scr_create")
       import :: c_int, c_ptr
       type(c_ptr) :: pot
       type(c_ptr), value :: params, scr
     end function
     integer(c_int) function agpu_juslin_scr_create(pot, params, scr) bind(C, name="agpu_juslin_scr_create")
       import :: c_int, c_ptr
       type(c_ptr) :: pot
       type(c_ptr), value :: params, scr
     end function
     integer(c_int) function agpu_eam_alloy_create(pot, fn, elements) bind(C, name="agpu_eam_alloy_create")
       import :: c_int, c_ptr, c_char
       type(c_ptr) :: pot
       character(kind=c_char) :: fn(*), elements(*)
     end function
     integer(c_int) function agpu_ewald_create(pot, cutoff, beta, kmax) bind(C, name="agpu_ewald_create")
       import :: c_int, c_ptr, c_double
       type(c_ptr) :: pot
       real(c_double), value :: cutoff, beta, kmax
     end function
     integer(c_int) function agpu_pme_create(pot, cutoff, order, grid) bind(C, name="agpu_pme_create")
       import :: c_int, c_ptr, c_double
       type(c_ptr) :: pot
       real(c_double), value :: cutoff
       integer(c_int), value :: order
       integer(c_int), intent(in) :: grid(3)
     end function
     integer(c_int) function agpu_rebo2_create(pot, flags) bind(C, name="agpu_rebo2_create")
       import :: c_int, c_ptr
       type(c_ptr) :: pot
       integer(c_int), value :: flags
     end function
     integer(c_int) function agpu_potential_destroy(pot) bind(C, name="agpu_potential_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: pot
     end function
     integer(c_int) function agpu_potential_set_precision(pot, prec) bind(C, name="agpu_potential_set_precision")
       import :: c_int, c_ptr
       type(c_ptr), value :: pot
       integer(c_int), value :: prec
     end function
     integer(c_int) function agpu_potential_bind_to(pot, p, nl) bind(C, name="agpu_potential_bind_to")
       import :: c_int, c_ptr
       type(c_ptr), value :: pot, p, nl
     end function
     integer(c_int) function agpu_potential_energy_and_forces(pot, p, nl, q, epot, f, wpot, mask, &
          epot_per_at, epot_per_bond, f_per_bond, wpot_per_at, wpot_per_bond, flags) &
          bind(C, name="agpu_potential_energy_and_forces")
       import :: c_int, c_ptr
       type(c_ptr), value :: pot, p, nl, q, epot, f, wpot, mask, epot_per_at, epot_per_bond, &
            f_per_bond, wpot_per_at, wpot_per_bond
       integer(c_int), value :: flags
     end function
  end interface

contains

  !> Message of the last error raised on this thread
  function last_error() result(msg)
    character(len=1024) :: msg
    character(kind=c_char) :: buf(1024)
    integer :: i, rc
    rc = agpu_get_error(buf, 1024_c_int)
    msg = ' '
    do i = 1, 1024
       if (buf(i) == c_null_char) exit
       msg(i:i) = buf(i)
    end do
  end function last_error

  subroutine particles_init(this, nat, ierror)
    type(particles_gpu_t), intent(inout) :: this
    integer, intent(in) :: nat
    integer, intent(out) :: ierror
    ierror = agpu_particles_create(this%h)
    if (ierror /= AGPU_OK) return
    ierror = agpu_particles_allocate(this%h, int(nat, c_int))
    this%nat = nat
  end subroutine particles_init

  subroutine particles_set_cell(this, cell, pbc, ierror)
    type(particles_gpu_t), intent(inout) :: this
    real(c_double), intent(in) :: cell(3, 3)
    logical, intent(in) :: pbc(3)
    integer, intent(out) :: ierror
    real(c_double) :: c(3, 3)
    integer(c_int) :: ip(3)
    c = cell
    ip = merge(1, 0, pbc)
    ierror = agpu_particles_set_cell(this%h, c, ip)
  end subroutine particles_set_cell

  subroutine particles_set_Z(this, Z, ierror)
    type(particles_gpu_t), intent(inout) :: this
    integer, intent(in) :: Z(:)
    integer, intent(out) :: ierror
    integer(c_int) :: zz(size(Z))
    zz = Z
    ierror = agpu_particles_set_atomic_numbers(this%h, zz, AGPU_HOST)
  end subroutine particles_set_Z

  subroutine particles_set_positions(this, r, ierror)
    type(particles_gpu_t), intent(inout) :: this
    real(c_double), intent(in) :: r(:, :)
    integer, intent(out) :: ierror
    real(c_double) :: rr(3, size(r, 2))
    rr = r
    ierror = agpu_particles_set_positions(this%h, rr, AGPU_HOST)
  end subroutine particles_set_positions

  subroutine neighbors_init(this, avgn, ierror)
    type(neighbors_gpu_t), intent(inout) :: this
    integer, intent(in) :: avgn
    integer, intent(out) :: ierror
    ierror = agpu_neighbors_create(this%h, int(avgn, c_int))
  end subroutine neighbors_init

  !> tersoff_init: db absent selects Tersoff_PRB_39_5566_SiC
  subroutine tersoff_init(this, ierror, db)
    type(potential_gpu_t), intent(inout) :: this
    integer, intent(out) :: ierror
    type(tersoff_params_t), optional, target, intent(in) :: db
    if (present(db)) then
       ierror = agpu_tersoff_create(this%h, c_loc(db))
    else
       ierror = agpu_tersoff_create(this%h, c_null_ptr)
    end if
  end subroutine tersoff_init

  !> kumagai_init (kumagai_module.f90:30-78): db absent selects
  !> Kumagai_CompMaterSci_39_457_Si
  subroutine kumagai_init(this, ierror, db)
    type(potential_gpu_t), intent(inout) :: this
    integer, intent(out) :: ierror
    type(kumagai_params_t), optional, target, intent(in) :: db
    if (present(db)) then
       ierror = agpu_kumagai_create(this%h, c_loc(db))
    else
       ierror = agpu_kumagai_create(this%h, c_null_ptr)
    end if
  end subroutine kumagai_init

  !> brenner_init (brenner_module.f90:27-152): db absent selects
  !> Erhart_PRB_71_035211_SiC
  subroutine brenner_init(this, ierror, db)
    type(potential_gpu_t), intent(inout) :: this
    integer, intent(out) :: ierror
    type(brenner_params_t), optional, target, intent(in) :: db
    if (present(db)) then
       ierror = agpu_brenner_create(this%h, c_loc(db))
    else
       ierror = agpu_brenner_create(this%h, c_null_ptr)
    end if
  end subroutine brenner_init

  !> juslin_init (juslin_module.f90:27-150): db absent selects
  !> Juslin_J_Appl_Phys_98_123520_WCH
  subroutine juslin_init(this, ierror, db)
    type(potential_gpu_t), intent(inout) :: this
    integer, intent(out) :: ierror
    type(juslin_params_t), optional, target, intent(in) :: db
    if (present(db)) then
       ierror = agpu_juslin_create(this%h, c_loc(db))
    else
       ierror = agpu_juslin_create(this%h, c_null_ptr)
    end if
  end subroutine juslin_init

  !> tersoff_scr_init (tersoff_scr.f90 with tersoff_module.f90:30-83): db and
  !> scr absent select the screened Tersoff_PRB_39_5566_SiC; both or neither
  subroutine tersoff_scr_init(this, ierror, db, scr)
    type(potential_gpu_t), intent(inout) :: this
    integer, intent(out) :: ierror
    type(tersoff_params_t), optional, target, intent(in) :: db
    type(screening_params_t), optional, target, intent(in) :: scr
    if (present(db) .and. present(scr)) then
       ierror = agpu_tersoff_scr_create(this%h, c_loc(db), c_loc(scr))
    else
       ierror = agpu_tersoff_scr_create(this%h, c_null_ptr, c_null_ptr)
    end if
  end subroutine tersoff_scr_init

  !> kumagai_scr_init (kumagai_scr.f90): db and scr absent select the
  !> screened Kumagai_CompMaterSci_39_457_Si
  subroutine kumagai_scr_init(this, ierror, db, scr)
    type(potential_gpu_t), intent(inout) :: this
    integer, intent(out) :: ierror
    type(kumagai_params_t), optional, target, intent(in) :: db
    type(screening_params_t), optional, target, intent(in) :: scr
    if (present(db) .and. present(scr)) then
       ierror = agpu_kumagai_scr_create(this%h, c_loc(db), c_loc(scr))
    else
       ierror = agpu_kumagai_scr_create(this%h, c_null_ptr, c_null_ptr)
    end if
  end subroutine kumagai_scr_init

  !> brenner_scr_init (brenner_scr.f90): db and scr absent select the
  !> screened Erhart_PRB_71_035211_SiC
  subroutine brenner_scr_init(this, ierror, db, scr)
    type(potential_gpu_t), intent(inout) :: this
    integer, intent(out) :: ierror
    type(brenner_params_t), optional, target, intent(in) :: db
    type(screening_params_t), optional, target, intent(in) :: scr
    if (present(db) .and. present(scr)) then
       ierror = agpu_brenner_scr_create(this%h, c_loc(db), c_loc(scr))
    else
       ierror = agpu_brenner_scr_create(this%h, c_null_ptr, c_null_ptr)
    end if
  end subroutine brenner_scr_init

  !> juslin_scr_init (juslin_scr.f90): db and scr absent select the screened
  !> Juslin W-C-H set; scr holds one entry per ordered element pair
  subroutine juslin_scr_init(this, ierror, db, scr)
    type(potential_gpu_t), intent(inout) :: this
    integer, intent(out) :: ierror
    type(juslin_params_t), optional, target, intent(in) :: db
    type(screening_params_t), optional, target, intent(in) :: scr
    if (present(db) .and. present(scr)) then
       ierror = agpu_juslin_scr_create(this%h, c_loc(db), c_loc(scr))
    else
       ierror = agpu_juslin_scr_create(this%h, c_null_ptr, c_null_ptr)
    end if
  end subroutine juslin_scr_init

  subroutine eam_alloy_init(this, fn, elements, ierror)
    type(potential_gpu_t), intent(inout) :: this
    character(len=*), intent(in) :: fn, elements
    integer, intent(out) :: ierror
    ierror = agpu_eam_alloy_create(this%h, trim(fn) // c_null_char, trim(elements) // c_null_char)
  end subroutine eam_alloy_init

  subroutine ewald_init(this, cutoff, beta, kmax, ierror)
    type(potential_gpu_t), intent(inout) :: this
    real(c_double), intent(in) :: cutoff, beta, kmax
    integer, intent(out) :: ierror
    ierror = agpu_ewald_create(this%h, cutoff, beta, kmax)
  end subroutine ewald_init

  !> pme_init (pme.f90:129-147): cutoff, order and grid as there; absent
  !> arguments keep the defaults 10, 8, 64^3
  subroutine pme_init(this, ierror, cutoff, order, grid)
    type(potential_gpu_t), intent(inout) :: this
    integer, intent(out) :: ierror
    real(c_double), intent(in), optional :: cutoff
    integer, intent(in), optional :: order
    integer, intent(in), optional :: grid(3)
    real(c_double) :: c
    integer(c_int) :: o, g(3)
    c = 10.0d0
    o = 8
    g = 64
    if (present(cutoff)) c = cutoff
    if (present(order)) o = order
    if (present(grid)) g = grid
    ierror = agpu_pme_create(this%h, c, o, g)
  end subroutine pme_init

  !> REBO2 (rebo2_init, rebo2_module.f90:30-42): dihedral = with_dihedral,
  !> zero_tables as the reference's switch of the same name
  subroutine rebo2_init(this, ierror, dihedral, zero_tables)
    type(potential_gpu_t), intent(inout) :: this
    integer, intent(out) :: ierror
    logical, intent(in), optional :: dihedral, zero_tables
    integer(c_int) :: flags
    flags = 0
    if (present(zero_tables)) then
       if (zero_tables) flags = ior(flags, 1_c_int)
    end if
    if (present(dihedral)) then
       if (dihedral) flags = ior(flags, 2_c_int)
    end if
    ierror = agpu_rebo2_create(this%h, flags)
  end subroutine rebo2_init

  subroutine set_precision(this, prec, ierror)
    type(potential_gpu_t), intent(inout) :: this
    integer, intent(in) :: prec
    integer, intent(out) :: ierror
    ierror = agpu_potential_set_precision(this%h, int(prec, c_int))
  end subroutine set_precision

  subroutine bind_to(this, p, nl, ierror)
    type(potential_gpu_t), intent(inout) :: this
    type(particles_gpu_t), intent(inout) :: p
    type(neighbors_gpu_t), intent(inout) :: nl
    integer, intent(out) :: ierror
    ierror = agpu_potential_bind_to(this%h, p%h, nl%h)
  end subroutine bind_to

  !> energy_and_forces (default_compute_func.f90:25-100 argument meaning):
  !> epot is set, f and wpot (and per-atom arrays) are accumulated.
  subroutine energy_and_forces(this, p, nl, epot, f, wpot, ierror, mask, epot_per_at, wpot_per_at)
    type(potential_gpu_t), intent(inout) :: this
    type(particles_gpu_t), intent(inout) :: p
    type(neighbors_gpu_t), intent(inout) :: nl
    real(c_double), target, intent(inout) :: epot
    real(c_double), target, intent(inout) :: f(3, p%nat)
    real(c_double), target, intent(inout) :: wpot(3, 3)
    integer, intent(out) :: ierror
    integer(c_int), target, optional, intent(in) :: mask(p%nat)
    real(c_double), target, optional, intent(inout) :: epot_per_at(p%nat)
    real(c_double), target, optional, intent(inout) :: wpot_per_at(3, 3, p%nat)
    type(c_ptr) :: pm, pe, pw
    pm = c_null_ptr
    pe = c_null_ptr
    pw = c_null_ptr
    if (present(mask)) pm = c_loc(mask)
    if (present(epot_per_at)) pe = c_loc(epot_per_at)
    if (present(wpot_per_at)) pw = c_loc(wpot_per_at)
    ierror = agpu_potential_energy_and_forces(this%h, p%h, nl%h, c_null_ptr, c_loc(epot), c_loc(f), &
         c_loc(wpot), pm, pe, c_null_ptr, c_null_ptr, pw, c_null_ptr, AGPU_HOST)
  end subroutine energy_and_forces

  !> Coulomb flavour: energy_and_forces(p, nl, q, epot, f, wpot)
  subroutine coulomb_energy_and_forces(this, p, nl, q, epot, f, wpot, ierror)
    type(potential_gpu_t), intent(inout) :: this
    type(particles_gpu_t), intent(inout) :: p
    type(neighbors_gpu_t), intent(inout) :: nl
    real(c_double), target, intent(in) :: q(p%nat)
    real(c_double), target, intent(inout) :: epot
    real(c_double), target, intent(inout) :: f(3, p%nat)
    real(c_double), target, intent(inout) :: wpot(3, 3)
    integer, intent(out) :: ierror
    ierror = agpu_potential_energy_and_forces(this%h, p%h, nl%h, c_loc(q), c_loc(epot), c_loc(f), &
         c_loc(wpot), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, AGPU_HOST)
  end subroutine coulomb_energy_and_forces

  subroutine particles_del(this)
    type(particles_gpu_t), intent(inout) :: this
    integer :: rc
    if (c_associated(this%h)) rc = agpu_particles_destroy(this%h)
    this%h = c_null_ptr
  end subroutine particles_del

  subroutine neighbors_del(this)
    type(neighbors_gpu_t), intent(inout) :: this
    integer :: rc
    if (c_associated(this%h)) rc = agpu_neighbors_destroy(this%h)
    this%h = c_null_ptr
  end subroutine neighbors_del

  subroutine potential_del(this)
    type(potential_gpu_t), intent(inout) :: this
    integer :: rc
    if (c_associated(this%h)) rc = agpu_potential_destroy(this%h)
    this%h = c_null_ptr
  end subroutine potential_del

end module atomistica_gpu

!! Reference-shaped host types for Fortran callers of libatomistica_amd.
!!
!! particles_t and neighbors_t carry the public components and routines of
!! the reference's interface modules (python_particles.f90:84-170,
!! python_neighbors.f90:51-139): the host fills Abox, pbc, Z and r_non_cyc
!! itself and signals changes with I_changed_positions / I_changed_other /
!! set_cell, as it does for the reference.  potential_t wraps a potential
!! handle of module atomistica_gpu (created there by tersoff_init, rebo2_init,
!! ...), and its bind_to and energy_and_forces take the reference's argument
!! list (default_compute_func.f90:25-45, per-bond arrays included).  Before
!! each call the components whose revision counters moved are pushed to the
!! device handles; host arrays are staged through HBM by the library.
module atomistica_gpu_dropin
  use, intrinsic :: iso_c_binding
  use atomistica_gpu, only: potential_gpu_t
  implicit none
  private

  integer, parameter :: DP = c_double
  integer, parameter :: MAX_Z = 116
  integer(c_int), parameter :: AGPU_HOST = 1

  !> particles_t (python_particles.f90:84-170): the components a host sets
  !> and reads; the device handle and the revisions last pushed are private
  type, public :: particles_t
     logical :: initialized = .false.
     integer :: pos_rev = 0
     integer :: cell_rev = 0
     integer :: other_rev = 0
     real(DP), pointer :: Abox(:, :) => null()
     real(DP) :: Bbox(3, 3) = 0.0_DP
     integer, pointer :: pbc(:) => null()
     integer :: nat = 0
     integer :: natloc = 0
     integer :: maxnatloc = 0
     integer :: totnat = 0
     integer, pointer :: Z(:) => null()
     integer, pointer :: el(:) => null()
     real(DP), pointer :: r_non_cyc(:, :) => null()
     integer :: nZ(MAX_Z) = 0
     integer :: nel = 0
     integer :: el2Z(MAX_Z) = 0
     integer :: Z2el(MAX_Z) = 0
     type(c_ptr), private :: h = c_null_ptr
     integer, private :: dev_nat = -1
     integer, private :: dev_pos_rev = -1
     integer, private :: dev_cell_rev = -1
     integer, private :: dev_other_rev = -1
  end type particles_t

  !> neighbors_t (python_neighbors.f90:51-139): avgn, interaction range and
  !> the list capacity neighbors_size (nat * avgn), the length per-bond
  !> arrays are allocated with
  type, public :: neighbors_t
     logical :: initialized = .false.
     integer :: avgn = 0
     real(DP) :: interaction_range = 0.0_DP
     real(DP) :: cutoff = 0.0_DP
     real(DP) :: requested_bin_size = 0.0_DP
     integer :: neighbors_size = 0
     integer :: it = 0
     integer :: nupdate = 0
     type(c_ptr), private :: h = c_null_ptr
  end type neighbors_t

  !> a potential: gpu is the handle created by atomistica_gpu's *_init
  type, public :: potential_t
     type(potential_gpu_t) :: gpu
  end type potential_t

  public :: init, allocate, set_cell, I_changed_positions, I_changed_other, update_elements
  public :: request_interaction_range, update, bind_to, energy_and_forces, del

  interface init
     module procedure particles_init, neighbors_init
  end interface init
  interface allocate
     module procedure particles_allocate
  end interface allocate
  interface set_cell
     module procedure particles_set_cell
  end interface set_cell
  interface I_changed_positions
     module procedure particles_I_changed_positions
  end interface I_changed_positions
  interface I_changed_other
     module procedure particles_I_changed_other
  end interface I_changed_other
  interface update_elements
     module procedure particles_update_elements
  end interface update_elements
  interface request_interaction_range
     module procedure neighbors_request_interaction_range
  end interface request_interaction_range
  interface update
     module procedure neighbors_update
  end interface update
  interface bind_to
     module procedure potential_bind_to
  end interface bind_to
  interface energy_and_forces
     module procedure potential_energy_and_forces
  end interface energy_and_forces
  interface del
     module procedure particles_del, neighbors_del
  end interface del

  interface
     integer(c_int) function c_particles_create(p) bind(C, name="agpu_particles_create")
       import :: c_int, c_ptr
       type(c_ptr), intent(out) :: p
     end function
     integer(c_int) function c_particles_destroy(p) bind(C, name="agpu_particles_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: p
     end function
     integer(c_int) function c_particles_allocate(p, nat) bind(C, name="agpu_particles_allocate")
       import :: c_int, c_ptr
       type(c_ptr), value :: p
       integer(c_int), value :: nat
     end function
     integer(c_int) function c_particles_set_cell(p, cell, pbc) bind(C, name="agpu_particles_set_cell")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: p
       real(c_double), intent(in) :: cell(9)
       integer(c_int), intent(in) :: pbc(3)
     end function
     integer(c_int) function c_particles_set_Z(p, Z, flags) bind(C, name="agpu_particles_set_atomic_numbers")
       import :: c_int, c_ptr
       type(c_ptr), value :: p
       integer(c_int), intent(in) :: Z(*)
       integer(c_int), value :: flags
     end function
     integer(c_int) function c_particles_set_positions(p, r, flags) bind(C, name="agpu_particles_set_positions")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: p
       real(c_double), intent(in) :: r(*)
       integer(c_int), value :: flags
     end function
     integer(c_int) function c_neighbors_create(nl, avgn) bind(C, name="agpu_neighbors_create")
       import :: c_int, c_ptr
       type(c_ptr), intent(out) :: nl
       integer(c_int), value :: avgn
     end function
     integer(c_int) function c_neighbors_destroy(nl) bind(C, name="agpu_neighbors_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: nl
     end function
     integer(c_int) function c_neighbors_request(nl, cutoff) &
          bind(C, name="agpu_neighbors_request_interaction_range")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: nl
       real(c_double), value :: cutoff
     end function
     integer(c_int) function c_neighbors_set_bin_size(nl, bin_size) bind(C, name="agpu_neighbors_set_bin_size")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: nl
       real(c_double), value :: bin_size
     end function
     integer(c_int) function c_neighbors_update(nl, p) bind(C, name="agpu_neighbors_update")
       import :: c_int, c_ptr
       type(c_ptr), value :: nl, p
     end function
     integer(c_int) function c_potential_bind_to(pot, p, nl) bind(C, name="agpu_potential_bind_to")
       import :: c_int, c_ptr
       type(c_ptr), value :: pot, p, nl
     end function
     integer(c_int) function c_energy_and_forces(pot, p, nl, q, epot, f, wpot, mask, epot_per_at, &
          epot_per_bond, f_per_bond, wpot_per_at, wpot_per_bond, flags) &
          bind(C, name="agpu_potential_energy_and_forces")
       import :: c_int, c_ptr
       type(c_ptr), value :: pot, p, nl, q, epot, f, wpot, mask, epot_per_at, epot_per_bond, f_per_bond, &
            wpot_per_at, wpot_per_bond
       integer(c_int), value :: flags
     end function
  end interface

contains

  !> particles_init (python_particles.f90:433-471): a 3x3 box, periodic
  subroutine particles_init(this)
    type(particles_t), intent(inout) :: this
    integer :: rc
    if (.not. c_associated(this%h)) rc = c_particles_create(this%h)
    if (.not. associated(this%Abox)) allocate(this%Abox(3, 3))
    if (.not. associated(this%pbc)) allocate(this%pbc(3))
    this%Abox = 0.0_DP
    this%pbc = 1
    this%initialized = .true.
  end subroutine particles_init

  !> particles_allocate (python_particles.f90:524-555): storage for nat atoms
  subroutine particles_allocate(this, nat, totnat, allow_def, error)
    type(particles_t), intent(inout) :: this
    integer, intent(in) :: nat
    integer, optional, intent(in) :: totnat
    logical, optional, intent(in) :: allow_def
    integer, optional, intent(inout) :: error
    if (.not. this%initialized) call particles_init(this)
    if (associated(this%Z)) deallocate(this%Z)
    if (associated(this%el)) deallocate(this%el)
    if (associated(this%r_non_cyc)) deallocate(this%r_non_cyc)
    allocate(this%Z(nat), this%el(nat), this%r_non_cyc(3, nat))
    this%Z = 0
    this%el = 0
    this%r_non_cyc = 0.0_DP
    this%nat = nat
    this%natloc = nat
    this%maxnatloc = nat
    this%totnat = nat
    if (present(totnat)) this%totnat = totnat
    this%other_rev = this%other_rev + 1
    this%pos_rev = this%pos_rev + 1
    if (present(error)) error = 0
  end subroutine particles_allocate

  !> particles_set_cell (python_particles.f90:286-346): Abox(:, k) is lattice
  !> vector k; Bbox its inverse
  subroutine particles_set_cell(this, Abox, pbc, scale_atoms, error)
    type(particles_t), intent(inout) :: this
    real(DP), intent(in) :: Abox(3, 3)
    logical, optional, intent(in) :: pbc(3)
    logical, optional, intent(in) :: scale_atoms
    integer, optional, intent(inout) :: error
    real(DP) :: det
    if (.not. this%initialized) call particles_init(this)
    if (present(scale_atoms)) then
       if (scale_atoms .and. this%nat > 0) then
          ! keep the fractional coordinates: r -> Abox_new Bbox_old r
          this%r_non_cyc = matmul(Abox, matmul(this%Bbox, this%r_non_cyc))
          this%pos_rev = this%pos_rev + 1
       end if
    end if
    this%Abox = Abox
    if (present(pbc)) then
       this%pbc = merge(1, 0, pbc)
    end if
    det = Abox(1, 1) * (Abox(2, 2) * Abox(3, 3) - Abox(2, 3) * Abox(3, 2)) &
         - Abox(1, 2) * (Abox(2, 1) * Abox(3, 3) - Abox(2, 3) * Abox(3, 1)) &
         + Abox(1, 3) * (Abox(2, 1) * Abox(3, 2) - Abox(2, 2) * Abox(3, 1))
    this%Bbox(1, 1) = (Abox(2, 2) * Abox(3, 3) - Abox(2, 3) * Abox(3, 2)) / det
    this%Bbox(1, 2) = (Abox(1, 3) * Abox(3, 2) - Abox(1, 2) * Abox(3, 3)) / det
    this%Bbox(1, 3) = (Abox(1, 2) * Abox(2, 3) - Abox(1, 3) * Abox(2, 2)) / det
    this%Bbox(2, 1) = (Abox(2, 3) * Abox(3, 1) - Abox(2, 1) * Abox(3, 3)) / det
    this%Bbox(2, 2) = (Abox(1, 1) * Abox(3, 3) - Abox(1, 3) * Abox(3, 1)) / det
    this%Bbox(2, 3) = (Abox(1, 3) * Abox(2, 1) - Abox(1, 1) * Abox(2, 3)) / det
    this%Bbox(3, 1) = (Abox(2, 1) * Abox(3, 2) - Abox(2, 2) * Abox(3, 1)) / det
    this%Bbox(3, 2) = (Abox(1, 2) * Abox(3, 1) - Abox(1, 1) * Abox(3, 2)) / det
    this%Bbox(3, 3) = (Abox(1, 1) * Abox(2, 2) - Abox(1, 2) * Abox(2, 1)) / det
    this%cell_rev = this%cell_rev + 1
    if (present(error)) error = 0
  end subroutine particles_set_cell

  !> I_changed_positions (python_particles.f90:768-777)
  subroutine particles_I_changed_positions(this)
    type(particles_t), intent(inout) :: this
    this%pos_rev = this%pos_rev + 1
  end subroutine particles_I_changed_positions

  !> I_changed_other: atomic numbers (or anything but positions and cell)
  subroutine particles_I_changed_other(this)
    type(particles_t), intent(inout) :: this
    this%other_rev = this%other_rev + 1
  end subroutine particles_I_changed_other

  !> update_elements (python_particles.f90:617-658): element statistics
  !> from Z, el = index of Z among the elements present (1-based)
  subroutine particles_update_elements(this)
    type(particles_t), intent(inout) :: this
    integer :: i, z
    this%nZ = 0
    this%el2Z = 0
    this%Z2el = 0
    this%nel = 0
    do i = 1, this%nat
       z = this%Z(i)
       if (z >= 1 .and. z <= MAX_Z) this%nZ(z) = this%nZ(z) + 1
    end do
    do z = 1, MAX_Z
       if (this%nZ(z) > 0) then
          this%nel = this%nel + 1
          this%el2Z(this%nel) = z
          this%Z2el(z) = this%nel
       end if
    end do
    do i = 1, this%nat
       z = this%Z(i)
       this%el(i) = 0
       if (z >= 1 .and. z <= MAX_Z) this%el(i) = this%Z2el(z)
    end do
    this%other_rev = this%other_rev + 1
  end subroutine particles_update_elements

  !> push what changed since the last call to the device handle
  subroutine particles_sync(this, ierror)
    type(particles_t), intent(inout) :: this
    integer, intent(out) :: ierror
    real(c_double) :: cell(9)
    integer(c_int) :: pbc(3)
    ierror = 0
    if (.not. this%initialized) call particles_init(this)
    if (this%dev_nat /= this%nat) then
       ierror = c_particles_allocate(this%h, int(this%nat, c_int))
       if (ierror /= 0) return
       this%dev_nat = this%nat
       this%dev_pos_rev = -1
       this%dev_other_rev = -1
    end if
    if (this%dev_cell_rev /= this%cell_rev) then
       cell = reshape(this%Abox, (/ 9 /))
       pbc = int(this%pbc(1:3), c_int)
       ierror = c_particles_set_cell(this%h, cell, pbc)
       if (ierror /= 0) return
       this%dev_cell_rev = this%cell_rev
    end if
    if (this%dev_other_rev /= this%other_rev .and. this%nat > 0) then
       ierror = c_particles_set_Z(this%h, int(this%Z, c_int), AGPU_HOST)
       if (ierror /= 0) return
       this%dev_other_rev = this%other_rev
    end if
    if (this%dev_pos_rev /= this%pos_rev .and. this%nat > 0) then
       ierror = c_particles_set_positions(this%h, reshape(this%r_non_cyc, (/ 3 * this%nat /)), AGPU_HOST)
       if (ierror /= 0) return
       this%dev_pos_rev = this%pos_rev
    end if
  end subroutine particles_sync

  subroutine particles_del(this)
    type(particles_t), intent(inout) :: this
    integer :: rc
    if (c_associated(this%h)) rc = c_particles_destroy(this%h)
    this%h = c_null_ptr
    if (associated(this%Abox)) deallocate(this%Abox)
    if (associated(this%pbc)) deallocate(this%pbc)
    if (associated(this%Z)) deallocate(this%Z)
    if (associated(this%el)) deallocate(this%el)
    if (associated(this%r_non_cyc)) deallocate(this%r_non_cyc)
    this%initialized = .false.
    this%dev_nat = -1
  end subroutine particles_del

  !> neighbors_init (python_neighbors.f90:224-272): avgn neighbours per atom
  !> (capacity nat * avgn), an optional range and bin size; no Verlet shell
  subroutine neighbors_init(this, avgn, cutoff, verlet_shell, bin_size, error)
    type(neighbors_t), intent(inout) :: this
    integer, optional, intent(in) :: avgn
    real(DP), optional, intent(in) :: cutoff
    real(DP), optional, intent(in) :: verlet_shell
    real(DP), optional, intent(in) :: bin_size
    integer, optional, intent(out) :: error
    integer :: rc
    this%avgn = 100
    if (present(avgn)) this%avgn = avgn
    rc = c_neighbors_create(this%h, int(this%avgn, c_int))
    if (rc == 0 .and. present(cutoff)) call neighbors_request_interaction_range(this, cutoff)
    if (rc == 0 .and. present(bin_size)) then
       this%requested_bin_size = bin_size
       rc = c_neighbors_set_bin_size(this%h, bin_size)
    end if
    this%initialized = rc == 0
    if (present(error)) error = rc
  end subroutine neighbors_init

  !> request_interaction_range (python_neighbors.f90:381-403): keeps the max
  subroutine neighbors_request_interaction_range(this, cutoff)
    type(neighbors_t), intent(inout) :: this
    real(DP), intent(in) :: cutoff
    integer :: rc
    this%interaction_range = max(this%interaction_range, cutoff)
    this%cutoff = this%interaction_range
    rc = c_neighbors_request(this%h, cutoff)
  end subroutine neighbors_request_interaction_range

  !> neighbors_update (python_neighbors.f90:430-456): rebuild when positions
  !> or the cell changed
  subroutine neighbors_update(this, p, error)
    type(neighbors_t), intent(inout) :: this
    type(particles_t), intent(inout) :: p
    integer, optional, intent(out) :: error
    integer :: rc
    call particles_sync(p, rc)
    if (rc == 0) rc = c_neighbors_update(this%h, p%h)
    if (rc == 0) then
       this%neighbors_size = p%nat * this%avgn
       this%nupdate = this%nupdate + 1
    end if
    if (present(error)) error = rc
  end subroutine neighbors_update

  subroutine neighbors_del(this)
    type(neighbors_t), intent(inout) :: this
    integer :: rc
    if (c_associated(this%h)) rc = c_neighbors_destroy(this%h)
    this%h = c_null_ptr
    this%initialized = .false.
  end subroutine neighbors_del

  !> bind_to (default_bind_to_func.f90:25-146): element map, range request
  subroutine potential_bind_to(this, p, nl, ierror)
    type(potential_t), intent(inout) :: this
    type(particles_t), intent(inout) :: p
    type(neighbors_t), intent(inout) :: nl
    integer, optional, intent(out) :: ierror
    integer :: rc
    call particles_sync(p, rc)
    if (rc == 0) rc = c_potential_bind_to(this%gpu%h, p%h, nl%h)
    if (rc == 0) nl%neighbors_size = p%nat * nl%avgn
    if (present(ierror)) ierror = rc
  end subroutine potential_bind_to

  !> energy_and_forces with the argument list of default_compute_func.f90:
  !> 25-45: epot is set; f, wpot and the per-atom and per-bond arrays are
  !> accumulated.  Per-bond arrays are indexed by list entry (length
  !> nl%neighbors_size, the first entries used).
  subroutine potential_energy_and_forces(this, p, nl, epot, f, wpot, mask, epot_per_at, epot_per_bond, &
       f_per_bond, wpot_per_at, wpot_per_bond, ierror)
    type(potential_t), intent(inout) :: this
    type(particles_t), intent(inout) :: p
    type(neighbors_t), intent(inout) :: nl
    real(DP), target, intent(inout) :: epot
    real(DP), target, intent(inout) :: f(3, p%maxnatloc)
    real(DP), target, intent(inout) :: wpot(3, 3)
    integer, target, optional, intent(in) :: mask(p%maxnatloc)
    real(DP), target, optional, intent(inout) :: epot_per_at(p%maxnatloc)
    real(DP), target, optional, intent(inout) :: epot_per_bond(nl%neighbors_size)
    real(DP), target, optional, intent(inout) :: f_per_bond(3, nl%neighbors_size)
    real(DP), target, optional, intent(inout) :: wpot_per_at(3, 3, p%maxnatloc)
    real(DP), target, optional, intent(inout) :: wpot_per_bond(3, 3, nl%neighbors_size)
    integer, optional, intent(out) :: ierror
    integer(c_int), allocatable, target :: m(:)
    type(c_ptr) :: pm, pea, peb, pfb, pwa, pwb
    integer :: rc
    call particles_sync(p, rc)
    if (rc == 0) then
       pm = c_null_ptr
       pea = c_null_ptr
       peb = c_null_ptr
       pfb = c_null_ptr
       pwa = c_null_ptr
       pwb = c_null_ptr
       if (present(mask)) then
          allocate(m(p%nat))
          m = int(mask(1:p%nat), c_int)
          pm = c_loc(m)
       end if
       if (present(epot_per_at)) pea = c_loc(epot_per_at)
       if (present(epot_per_bond)) peb = c_loc(epot_per_bond)
       if (present(f_per_bond)) pfb = c_loc(f_per_bond)
       if (present(wpot_per_at)) pwa = c_loc(wpot_per_at)
       if (present(wpot_per_bond)) pwb = c_loc(wpot_per_bond)
       rc = c_energy_and_forces(this%gpu%h, p%h, nl%h, c_null_ptr, c_loc(epot), c_loc(f), c_loc(wpot), pm, &
            pea, peb, pfb, pwa, pwb, AGPU_HOST)
       nl%it = nl%it + 1
    end if
    if (present(ierror)) ierror = rc
  end subroutine potential_energy_and_forces

end module atomistica_gpu_dropin

!> Writer of the .model file: the same sections, keys and order as
!> write_gpr (src/model/gpr/gpr.f90:329-398) with the kernel, mean, data and
!> weights blocks of src/kernels/*.f90, src/mean/*.f90,
!> src/data/data_set_type.f90:60-79 and src/model/gpr/likelihood.f90:255-264.
module fb_model_writer
  use iso_c_binding, only: c_double, c_int
  use ferebus_hip, only: FB_LEAF_RBF, FB_LEAF_PERIODIC, FB_LEAF_RBF_CYCLIC, FB_LEAF_LINEAR, &
                         FB_LEAF_CONSTANT, FB_OP_SUM, FB_LIKELIHOOD_CONCENTRATED
  use fb_gpr, only: GaussianProcessRegressor
  implicit none
  private
  integer, parameter :: wp = c_double

  type, public :: ModelMeta
    character(len=:), allocatable :: system_name, atom_name, property
    integer :: natoms = 0
    integer :: alf(3) = 0
    character(len=256), allocatable :: note_keys(:), note_values(:)
  end type ModelMeta

  public :: write_model, model_file_name, input_unit_name, output_unit_name

contains

  !> SYSTEM_property_ATOM.model (gpr.f90:339-340)
  function model_file_name(meta) result(fname)
    type(ModelMeta), intent(in) :: meta
    character(len=:), allocatable :: fname
    fname = meta%system_name // "_" // meta%property // "_" // meta%atom_name // ".model"
  end function model_file_name

  !> units.f90:55-64: angles at features 3, 6, 9, ... and 5, 8, 11, ...
  function input_unit_name(i) result(u)
    integer, intent(in) :: i
    character(len=:), allocatable :: u
    if (modulo(i, 3) == 0 .or. (i > 3 .and. modulo(i + 1, 3) == 0)) then
      u = "radians"
    else
      u = "bohr"
    end if
  end function input_unit_name

  !> units.f90:66-78
  function output_unit_name(label) result(u)
    character(len=*), intent(in) :: label
    character(len=:), allocatable :: u
    select case (trim(label))
    case ("iqa", "IQA"); u = "Ha"
    case ("q00", "Q00"); u = "e"
    case default; u = "unknown"
    end select
  end function output_unit_name

  subroutine write_model(gp, theta, meta, unit, rc)
    type(GaussianProcessRegressor), intent(inout) :: gp
    real(wp), intent(in) :: theta(:)
    type(ModelMeta), intent(in) :: meta
    integer, intent(in) :: unit
    integer, intent(out) :: rc
    real(wp), allocatable :: alpha(:)
    real(wp) :: out4(4), nugget
    integer(c_int) :: info
    integer :: i, j, k, nleaf, off, doff
    character(len=64) :: fmt
    allocate(alpha(gp%n))
    call gp%weights(theta, alpha, out4, info, rc)
    if (rc /= 0) return
    nugget = gp%noise
    if (gp%optimise_noise) nugget = theta(gp%nparams)
    write(unit, '(A)') "# [metadata]"
    write(unit, '(A)') "# program ferebus"
    write(unit, '(A)') "# version 7.2.0"
    write(unit, '(A, 1x, G0)') "# nugget", nugget
    write(unit, '(A, 1x, G0)') "# likelihood", out4(1)
    if (allocated(meta%note_keys)) then
      do i = 1, size(meta%note_keys)
        write(unit, '(A, 1x, A, 1x, A, 1x, A)') "#", trim(meta%note_keys(i)), "=", trim(meta%note_values(i))
      end do
    end if
    write(unit, '(A)') ""
    write(unit, '(A)') "[system]"
    write(unit, '(A, 1x, A)') "name", meta%system_name
    write(unit, '(A, 1x, A)') "atom", meta%atom_name
    write(unit, '(A, 1x, A)') "property", meta%property
    write(unit, '(A, 1X, I0, 1X, I0, 1X, I0)') "ALF", meta%alf(1), meta%alf(2), meta%alf(3)
    write(unit, '(A)')
    write(unit, '(A)') "[dimensions]"
    write(unit, '(A, 1x, I0)') "number_of_atoms", meta%natoms
    write(unit, '(A, 1x, I0)') "number_of_features", gp%d
    write(unit, '(A, 1x, I0)') "number_of_training_points", gp%n
    write(unit, '(A)')
    if (gp%likelihood == FB_LIKELIHOOD_CONCENTRATED) then
      ! gpr.f90:374-381: the concentrated likelihood's GLS constant
      write(unit, '(A)') "[mean]"
      write(unit, '(A, 1x, A)') "type", "constant"
      write(unit, '(A, 1x, G0)') "value", out4(3)
    else
      call gp%mean%write(unit)
    end if
    write(unit, '(A)')
    nleaf = gp%prog%n_leaves()
    write(unit, '(A)') "[kernels]"
    write(unit, '(A, 1x, I0)') "number_of_kernels", nleaf
    write(unit, '(A, 1x, A)') "composition", gp%prog%composition
    write(unit, '(A)')
    off = 0
    doff = 0
    j = 0
    do k = 1, gp%prog%n_nodes
      if (gp%prog%node_type(k) >= FB_OP_SUM) cycle
      j = j + 1
      call write_leaf(unit, gp, k, theta, off, doff)
      if (j < nleaf) write(unit, '(A)')
    end do
    write(unit, '(A)')
    ! data_set_type.f90:60-79
    write(fmt, '(a,i0,a)') "(a,1x,(", gp%d, "(a,1x)))"
    write(unit, "(a)") "[training_data]"
    write(unit, trim(fmt)) "units.x", (input_unit_name(i), i = 1, gp%d)
    write(unit, "(a,1x,a)") "units.y", output_unit_name(meta%property)
    write(unit, "(a)") "scaling.x none"
    write(unit, "(a)") "scaling.y none"
    write(unit, "(a)")
    write(fmt, '(a,i0,a)') "(", gp%d, "(g0,1x))"
    write(unit, "(A)") "[training_data.x]"
    do i = 1, gp%n
      write(unit, trim(fmt)) (gp%x(i, j), j = 1, gp%d)
    end do
    write(unit, "(a)")
    write(unit, "(A)") "[training_data.y]"
    do i = 1, gp%n
      write(unit, "(g0)") gp%y(i)
    end do
    write(unit, '(A)')
    write(unit, "(A)") "[weights]"                 ! likelihood.f90:255-264
    do i = 1, gp%n
      write(unit, "(g0)") alpha(i)
    end do
    write(unit, '(A)')
    rc = int(info)
  end subroutine write_model

  !> One [kernel.<name>] block: rbf_kernel.f90:127-149, periodic_kernel.f90:126-148,
  !> rbf_cyclic_kernel.f90:123-145, linear_kernel.f90:108-126, constant.f90:98-109.
  subroutine write_leaf(unit, gp, k, theta, off, doff)
    integer, intent(in) :: unit, k
    type(GaussianProcessRegressor), intent(in) :: gp
    real(wp), intent(in) :: theta(:)
    integer, intent(inout) :: off, doff
    integer :: nd, t
    character(len=64) :: fmt, ifmt
    real(wp), allocatable :: l(:)
    t = gp%prog%node_type(k)
    nd = gp%prog%node_ndim(k)
    write(unit, "(A, A, A)") "[kernel.", trim(gp%prog%leaf_name(k)), "]"
    if (t == FB_LEAF_CONSTANT) then
      write(unit, "(A, 1x, A)") "type", "constant"
      write(unit, "(A, 1x, I0)") "number_of_dimensions", 1
      write(unit, "(A, 1x, A)") "active_dimensions", "<TODO>"
      write(unit, "(A, 1x, G0)") "value", theta(off + 1)
      off = off + 1
      return
    end if
    write(fmt, '(a,i0,a)') "(a,1x,(", nd, "(g0,1x)))"
    write(ifmt, '(a,i0,a)') "(a,1x,(", nd, "(i0,1x)))"
    select case (t)
    case (FB_LEAF_RBF); write(unit, "(A, 1x, A)") "type", "rbf"
    case (FB_LEAF_PERIODIC); write(unit, "(A, 1x, A)") "type", "periodic"
    case (FB_LEAF_RBF_CYCLIC); write(unit, "(A, 1x, A)") "type", "rbf-cyclic"
    case (FB_LEAF_LINEAR); write(unit, "(A, 1x, A)") "type", "linear"
    end select
    write(unit, "(A, 1x, I0)") "number_of_dimensions", nd
    write(unit, trim(ifmt)) "active_dimensions", gp%prog%active_dims(doff + 1:doff + nd)
    select case (t)
    case (FB_LEAF_RBF, FB_LEAF_RBF_CYCLIC)
      l = exp(theta(off + 1:off + nd))
      write(unit, trim(fmt)) "theta", 0.5_wp * l            ! DL_POLY theta = l/2
    case (FB_LEAF_PERIODIC)
      l = 1.0_wp / exp(theta(off + 1:off + nd))
      l = l * l
      write(unit, trim(fmt)) "theta", 0.5_wp * l
    case (FB_LEAF_LINEAR)
      write(unit, "(A, 1x, G0)") "order", 0.0_wp
      write(unit, trim(fmt)) "c", exp(-theta(off + 1:off + nd))
    end select
    off = off + nd
    doff = doff + nd
  end subroutine write_leaf
end module fb_model_writer

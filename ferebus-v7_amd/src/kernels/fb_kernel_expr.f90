!> Kernel expressions of ferebus.toml -> postfix kernel programs (fb_kernel_spec).
!>
!> Grammar of the reference interpreter (src/interpreter/parser.f90:102-197):
!>   expr   := term (('+' | '-') term)*      left associative
!>   term   := factor (('*' | '/') term)?    right recursive
!>   factor := '+' factor | '-' factor | NUMBER | '(' expr ')' | IDENT
!> Visitor (src/interpreter/ast.f90:173-271): '+' sum, '*' product,
!> identifiers name [kernels.<name>] tables, numbers become constant kernels
!> whose value is one theta entry.  '-' and '/' are rejected ("not yet
!> implemented" in the reference).  Leaves are emitted left to right, the
!> order of their parameters in theta (src/kernels/kernel_module.f90:227-245).
module fb_kernel_expr
  use iso_c_binding, only: c_int
  use ferebus_hip, only: FB_LEAF_RBF, FB_LEAF_PERIODIC, FB_LEAF_RBF_CYCLIC, FB_LEAF_LINEAR, &
                         FB_LEAF_CONSTANT, FB_OP_SUM, FB_OP_PROD
  implicit none
  private

  integer, parameter :: MAX_NODES = 64

  !> One [kernels.<name>] table.
  type, public :: KernelConfigEntry
    character(len=64) :: name = ""
    character(len=32) :: ktype = "rbf"
    integer, allocatable :: active_dims(:)
  end type KernelConfigEntry

  !> Compiled postfix program plus the leaves' names (for the .model writer).
  type, public :: KernelProgram
    integer :: n_nodes = 0
    integer(c_int) :: node_type(MAX_NODES) = 0
    integer(c_int) :: node_ndim(MAX_NODES) = 0
    integer(c_int), allocatable :: active_dims(:)
    character(len=64) :: leaf_name(MAX_NODES) = ""
    character(len=:), allocatable :: composition
    character(len=:), allocatable :: error
  contains
    procedure :: n_params => program_n_params
    procedure :: n_leaves => program_n_leaves
  end type KernelProgram

  type :: Parser
    character(len=:), allocatable :: s
    integer :: pos = 1
    type(KernelConfigEntry), pointer :: cfg(:) => null()
    integer :: nfeats = 0
    type(KernelProgram), pointer :: prog => null()
    logical :: failed = .false.
  end type Parser

  public :: compile_kernel

contains

  integer function program_n_params(self) result(n)
    class(KernelProgram), intent(in) :: self
    integer :: k
    n = 0
    do k = 1, self%n_nodes
      if (self%node_type(k) == FB_LEAF_CONSTANT) then
        n = n + 1
      else if (self%node_type(k) < FB_OP_SUM) then
        n = n + self%node_ndim(k)
      end if
    end do
  end function program_n_params

  integer function program_n_leaves(self) result(n)
    class(KernelProgram), intent(in) :: self
    n = count(self%node_type(1:self%n_nodes) < FB_OP_SUM)
  end function program_n_leaves

  !> Compiles `text` against the configured kernels for `nfeats` features.
  !> On error prog%error is allocated with the reason.
  subroutine compile_kernel(text, cfg, nfeats, prog)
    character(len=*), intent(in) :: text
    type(KernelConfigEntry), target, intent(in) :: cfg(:)
    integer, intent(in) :: nfeats
    type(KernelProgram), target, intent(out) :: prog
    type(Parser) :: p
    character(len=:), allocatable :: name
    allocate(prog%active_dims(0))
    p%s = trim(text)
    p%pos = 1
    p%cfg => cfg
    p%nfeats = nfeats
    p%prog => prog
    call expr(p, name)
    call skip_space(p)
    if (.not. p%failed .and. p%pos <= len(p%s)) call fail(p, "trailing input in kernel expression")
    if (.not. p%failed) prog%composition = name
  end subroutine compile_kernel

  subroutine fail(p, msg)
    type(Parser), intent(inout) :: p
    character(len=*), intent(in) :: msg
    if (.not. p%failed) p%prog%error = msg
    p%failed = .true.
  end subroutine fail

  subroutine skip_space(p)
    type(Parser), intent(inout) :: p
    do while (p%pos <= len(p%s))
      if (p%s(p%pos:p%pos) /= " ") exit
      p%pos = p%pos + 1
    end do
  end subroutine skip_space

  character function peek(p)
    type(Parser), intent(inout) :: p
    call skip_space(p)
    if (p%pos > len(p%s)) then
      peek = achar(0)
    else
      peek = p%s(p%pos:p%pos)
    end if
  end function peek

  subroutine emit(p, t, nd, name)
    type(Parser), intent(inout) :: p
    integer, intent(in) :: t, nd
    character(len=*), intent(in) :: name
    if (p%prog%n_nodes >= MAX_NODES) then
      call fail(p, "kernel expression too long")
      return
    end if
    p%prog%n_nodes = p%prog%n_nodes + 1
    p%prog%node_type(p%prog%n_nodes) = t
    p%prog%node_ndim(p%prog%n_nodes) = nd
    p%prog%leaf_name(p%prog%n_nodes) = name
  end subroutine emit

  recursive subroutine expr(p, name)
    type(Parser), intent(inout) :: p
    character(len=:), allocatable, intent(out) :: name
    character(len=:), allocatable :: rhs
    character :: c
    call term(p, name)
    do while (.not. p%failed)
      c = peek(p)
      if (c /= "+" .and. c /= "-") exit
      p%pos = p%pos + 1
      if (c == "-") then
        call fail(p, "Kernel minus not yet implemented")
        return
      end if
      call term(p, rhs)
      call emit(p, FB_OP_SUM, 0, "")
      name = "(" // name // "+" // rhs // ")"
    end do
  end subroutine expr

  recursive subroutine term(p, name)
    type(Parser), intent(inout) :: p
    character(len=:), allocatable, intent(out) :: name
    character(len=:), allocatable :: rhs
    character :: c
    call factor(p, name)
    if (p%failed) return
    c = peek(p)
    if (c == "*" .or. c == "/") then
      p%pos = p%pos + 1
      if (c == "/") then
        call fail(p, "Kernel divide not yet implemented")
        return
      end if
      call term(p, rhs)
      call emit(p, FB_OP_PROD, 0, "")
      name = "(" // name // "*" // rhs // ")"
    end if
  end subroutine term

  recursive subroutine factor(p, name)
    type(Parser), intent(inout) :: p
    character(len=:), allocatable, intent(out) :: name
    character :: c
    integer :: start, k, kind
    name = ""
    c = peek(p)
    if (c == "+") then
      p%pos = p%pos + 1
      call factor(p, name)
    else if (c == "-") then
      call fail(p, "Kernel minus not yet implemented")
    else if (c == "(") then
      p%pos = p%pos + 1
      call expr(p, name)
      if (peek(p) /= ")") then
        call fail(p, "missing ) in kernel expression")
        return
      end if
      p%pos = p%pos + 1
    else if (index("0123456789", c) > 0 .and. c /= achar(0)) then
      ! lexer.f90:95-130: digits, optional fraction, optional exponent
      call skip_digits(p)
      if (p%pos <= len(p%s)) then
        if (p%s(p%pos:p%pos) == ".") then
          p%pos = p%pos + 1
          call skip_digits(p)
        end if
      end if
      if (p%pos <= len(p%s)) then
        if (p%s(p%pos:p%pos) == "e" .or. p%s(p%pos:p%pos) == "E") then
          p%pos = p%pos + 1
          if (p%pos <= len(p%s)) then
            if (p%s(p%pos:p%pos) == "+" .or. p%s(p%pos:p%pos) == "-") p%pos = p%pos + 1
          end if
          call skip_digits(p)
        end if
      end if
      call emit(p, FB_LEAF_CONSTANT, 0, "c")
      name = "c"
    else if (is_alpha(c)) then
      start = p%pos
      do while (p%pos <= len(p%s))
        if (.not. (is_alpha(p%s(p%pos:p%pos)) .or. index("0123456789", p%s(p%pos:p%pos)) > 0)) exit
        p%pos = p%pos + 1
      end do
      name = p%s(start:p%pos-1)
      do k = 1, size(p%cfg)
        if (trim(p%cfg(k)%name) == name) exit
      end do
      if (k > size(p%cfg)) then
        call fail(p, "kernel " // name // " is not configured under [kernels]")
        return
      end if
      select case (trim(p%cfg(k)%ktype))
      case ("rbf"); kind = FB_LEAF_RBF
      case ("rbf-cyclic"); kind = FB_LEAF_RBF_CYCLIC
      case ("periodic"); kind = FB_LEAF_PERIODIC
      case ("linear"); kind = FB_LEAF_LINEAR
      case default
        call fail(p, "unknown kernel type " // trim(p%cfg(k)%ktype))
        return
      end select
      call add_leaf(p, kind, p%cfg(k), name)
    else
      call fail(p, "unexpected character in kernel expression")
    end if
  end subroutine factor

  subroutine add_leaf(p, kind, cfg, name)
    type(Parser), intent(inout) :: p
    integer, intent(in) :: kind
    type(KernelConfigEntry), intent(in) :: cfg
    character(len=*), intent(in) :: name
    integer, allocatable :: dims(:)
    integer :: f
    ! src/config/kernel/kernel_type_config.f90:54-73: default dims 1..nfeats
    if (allocated(cfg%active_dims)) then
      dims = cfg%active_dims
    else
      dims = [(f, f = 1, p%nfeats)]
    end if
    if (size(dims) < 1 .or. any(dims < 1) .or. any(dims > p%nfeats)) then
      call fail(p, "kernel " // name // ": active_dimensions outside 1..nfeats")
      return
    end if
    if (kind == FB_LEAF_RBF_CYCLIC) then
      if (size(dims) /= p%nfeats) then
        call fail(p, "rbf-cyclic sums over every feature")
        return
      end if
    end if
    call emit(p, kind, size(dims), name)
    p%prog%active_dims = [p%prog%active_dims, int(dims, c_int)]
  end subroutine add_leaf

  !> str.f90:174-180
  logical function is_alpha(c)
    character, intent(in) :: c
    is_alpha = (c >= "a" .and. c <= "z") .or. (c >= "A" .and. c <= "Z")
  end function is_alpha

  subroutine skip_digits(p)
    type(Parser), intent(inout) :: p
    do while (p%pos <= len(p%s))
      if (index("0123456789", p%s(p%pos:p%pos)) == 0) exit
      p%pos = p%pos + 1
    end do
  end subroutine skip_digits
end module fb_kernel_expr

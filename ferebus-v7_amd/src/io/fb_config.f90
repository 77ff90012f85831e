!> ferebus.toml reader (the subset of TOML the reference's examples use) and
!> training-set CSV reader.
!>
!> Keys and defaults follow the reference's config surface:
!>   [system] name, natoms, atoms = [{name, alf}], properties, nfeats
!>     (src/config/system/system.f90:52-122; nfeats = 3*natoms - 6)
!>   [model] mean="constant", optimiser="pso", kernel, noise=1e-10,
!>     optimise_noise=false, likelihood="marginal", standardise=false
!>     (src/config/model_config.f90:47-116); noise_max=1e-6, the upper end of
!>     the optimised nugget's search box (hard-coded in the reference,
!>     src/model/gpr/gpr.f90:229,242; a key of this build)
!>   [optimiser] search_min=0, search_max=1 (src/config/optimiser/optimiser.f90:35-38)
!>   [optimiser.pso] swarm_size=50, iterations=1000, inertia_weight=0.789,
!>     cognitive/social_learning_rate=1.494, stopping_criteria
!>     (src/config/optimiser/pso/pso.f90:48-56)
!>   [optimiser.pso.relative_change] tolerance=1e-6, stall_iterations=20
!>     (src/config/optimiser/pso/stopping_criteria.f90:60-62)
!>   [optimiser.pso] swarm_updater="global" ("default"; "forced" with
!>     [optimiser.pso.forced] delta=1e-7), parameter_updater="constant"
!>     ("none"; "unified_adaptive"/"uapso"/"ua", "individual_adaptive"/
!>     "iapso"/"ia" with their table -- or [optimiser.pso.adaptive], or the
!>     other adaptive variant's -- holding min/max/n of the three
!>     coefficients, reward/penalty step lengths and success_threshold;
!>     "random_unified"/"random"/"ru", "random_individual"/"ri")
!>     (src/config/optimiser/pso/pso.f90:92-157, parameter_updater.f90:62-75,
!>     swarm_updater.f90:58).  The three names are sanitised as the reference
!>     does (spaces to underscores, lower case; src/utils/str.f90:85-98), and
!>     an unknown name -- or a stopping criterion the reference parses but
!>     never implements (potential_based, sign_test; pso.f90:360-366) -- is
!>     an error, not a silent default.
!>   [kernels.<name>] type="rbf", active_dimensions (src/config/config_module.f90:165-190)
!>   [notes] key = value
!> The CSV reader takes the header convention of src/csv.f90:37-177: an empty
!> header is an index column, f<digits> a feature, anything else a property.
module fb_config
  use iso_c_binding, only: c_double
  use fb_kernel_expr, only: KernelConfigEntry
  implicit none
  private
  integer, parameter :: wp = c_double

  type :: Entry
    character(len=:), allocatable :: key, value
  end type Entry

  type, public :: TomlDoc
    type(Entry), allocatable :: e(:)
    integer :: n = 0
    character(len=:), allocatable :: error
  contains
    procedure :: load => toml_load
    procedure :: has => toml_has
    procedure :: raw => toml_raw
    procedure :: get_str => toml_get_str
    procedure :: get_int => toml_get_int
    procedure :: get_real => toml_get_real
    procedure :: get_bool => toml_get_bool
    procedure :: get_int_array => toml_get_int_array
    procedure :: keys_under => toml_keys_under
  end type TomlDoc

  type, public :: AtomConfig
    character(len=64) :: name = ""
    integer :: alf(3) = 0
  end type AtomConfig

  type, public :: FerebusConfig
    character(len=:), allocatable :: system_name
    integer :: natoms = 0, nfeats = 0
    type(AtomConfig), allocatable :: atoms(:)
    character(len=64), allocatable :: properties(:)      ! empty = all
    character(len=:), allocatable :: mean, kernel, likelihood
    real(wp) :: noise = 1e-10_wp
    logical :: optimise_noise = .false.
    real(wp) :: noise_max = 1e-6_wp
    real(wp) :: search_min = 0.0_wp, search_max = 1.0_wp
    integer :: swarm_size = 50, iterations = 1000
    real(wp) :: inertia_weight = 0.789_wp, cognitive = 1.494_wp, social = 1.494_wp
    character(len=:), allocatable :: stopping
    real(wp) :: tolerance = 1e-6_wp
    integer :: stall_iterations = 20
    character(len=:), allocatable :: swarm_updater, parameter_updater
    real(wp) :: delta = 1e-7_wp
    real(wp) :: min_w = 0.0_wp, max_w = 1.0_wp, min_c1 = 0.0_wp, max_c1 = 4.0_wp, min_c2 = 0.0_wp, max_c2 = 4.0_wp
    integer :: n_w = 10, n_c1 = 10, n_c2 = 10
    real(wp) :: reward_step = 0.01_wp, penalty_step = 0.01_wp, success_threshold = 0.5_wp
    integer :: lbfgs_every = 0, lbfgs_steps = 5
    type(KernelConfigEntry), allocatable :: kernels(:)
    character(len=256), allocatable :: note_keys(:), note_values(:)
  end type FerebusConfig

  public :: read_config, read_training_set, sanitise

contains

  !> Spaces to underscores and lower case (src/utils/str.f90:85-98).
  function sanitise(s) result(r)
    character(len=*), intent(in) :: s
    character(len=:), allocatable :: r
    integer :: i, c
    r = s
    do i = 1, len(r)
      if (r(i:i) == " ") then
        r(i:i) = "_"
      else
        c = iachar(r(i:i))
        if (c >= iachar("A") .and. c <= iachar("Z")) r(i:i) = achar(c + 32)
      end if
    end do
  end function sanitise

  function strip(s) result(r)
    character(len=*), intent(in) :: s
    character(len=:), allocatable :: r
    r = trim(adjustl(s))
  end function strip

  function unquote(s) result(r)
    character(len=*), intent(in) :: s
    character(len=:), allocatable :: r
    r = strip(s)
    if (len(r) >= 2) then
      if ((r(1:1) == '"' .and. r(len(r):len(r)) == '"') .or. &
          (r(1:1) == "'" .and. r(len(r):len(r)) == "'")) r = r(2:len(r) - 1)
    end if
  end function unquote

  !> Drops a # comment that is not inside a string.
  function drop_comment(s) result(r)
    character(len=*), intent(in) :: s
    character(len=:), allocatable :: r
    integer :: i
    logical :: instr
    character :: q
    instr = .false.
    q = " "
    do i = 1, len(s)
      if (instr) then
        if (s(i:i) == q) instr = .false.
      else if (s(i:i) == '"' .or. s(i:i) == "'") then
        instr = .true.
        q = s(i:i)
      else if (s(i:i) == "#") then
        r = s(1:i - 1)
        return
      end if
    end do
    r = s
  end function drop_comment

  integer function bracket_balance(s) result(b)
    character(len=*), intent(in) :: s
    integer :: i
    logical :: instr
    b = 0
    instr = .false.
    do i = 1, len(s)
      if (s(i:i) == '"') instr = .not. instr
      if (instr) cycle
      if (s(i:i) == "[" .or. s(i:i) == "{") b = b + 1
      if (s(i:i) == "]" .or. s(i:i) == "}") b = b - 1
    end do
  end function bracket_balance

  subroutine push(self, key, value)
    class(TomlDoc), intent(inout) :: self
    character(len=*), intent(in) :: key, value
    type(Entry), allocatable :: tmp(:)
    if (.not. allocated(self%e)) allocate(self%e(16))
    if (self%n == size(self%e)) then
      allocate(tmp(2 * self%n))
      tmp(1:self%n) = self%e(1:self%n)
      call move_alloc(tmp, self%e)
    end if
    self%n = self%n + 1
    self%e(self%n)%key = key
    self%e(self%n)%value = value
  end subroutine push

  subroutine toml_load(self, path, ok)
    class(TomlDoc), intent(inout) :: self
    character(len=*), intent(in) :: path
    logical, intent(out) :: ok
    integer :: u, ios, eq
    character(len=4096) :: line
    character(len=:), allocatable :: l, table, key, val
    ok = .false.
    open(newunit=u, file=path, status="old", action="read", iostat=ios)
    if (ios /= 0) then
      self%error = "cannot open " // path
      return
    end if
    table = ""
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      l = strip(drop_comment(line))
      if (len(l) == 0) cycle
      if (l(1:1) == "[") then
        table = strip(l(2:index(l, "]", back=.true.) - 1))
        cycle
      end if
      eq = index(l, "=")
      if (eq == 0) cycle
      key = unquote(l(1:eq - 1))
      val = strip(l(eq + 1:))
      do while (bracket_balance(val) > 0)        ! multi-line arrays
        read(u, '(A)', iostat=ios) line
        if (ios /= 0) exit
        val = val // " " // strip(drop_comment(line))
      end do
      if (len(table) > 0) key = table // "." // key
      call push(self, key, val)
    end do
    close(u)
    ok = .true.
  end subroutine toml_load

  integer function find(self, key)
    class(TomlDoc), intent(in) :: self
    character(len=*), intent(in) :: key
    do find = 1, self%n
      if (self%e(find)%key == key) return
    end do
    find = 0
  end function find

  logical function toml_has(self, key)
    class(TomlDoc), intent(in) :: self
    character(len=*), intent(in) :: key
    toml_has = find(self, key) > 0
  end function toml_has

  function toml_raw(self, key) result(v)
    class(TomlDoc), intent(in) :: self
    character(len=*), intent(in) :: key
    character(len=:), allocatable :: v
    integer :: k
    k = find(self, key)
    v = ""
    if (k > 0) v = self%e(k)%value
  end function toml_raw

  function toml_get_str(self, key, default) result(v)
    class(TomlDoc), intent(in) :: self
    character(len=*), intent(in) :: key, default
    character(len=:), allocatable :: v
    if (self%has(key)) then
      v = unquote(self%raw(key))
    else
      v = default
    end if
  end function toml_get_str

  integer function toml_get_int(self, key, default) result(v)
    class(TomlDoc), intent(in) :: self
    character(len=*), intent(in) :: key
    integer, intent(in) :: default
    integer :: ios
    v = default
    if (self%has(key)) then
      read(self%raw(key), *, iostat=ios) v
      if (ios /= 0) v = default
    end if
  end function toml_get_int

  real(wp) function toml_get_real(self, key, default) result(v)
    class(TomlDoc), intent(in) :: self
    character(len=*), intent(in) :: key
    real(wp), intent(in) :: default
    integer :: ios
    v = default
    if (self%has(key)) then
      read(self%raw(key), *, iostat=ios) v
      if (ios /= 0) v = default
    end if
  end function toml_get_real

  logical function toml_get_bool(self, key, default) result(v)
    class(TomlDoc), intent(in) :: self
    character(len=*), intent(in) :: key
    logical, intent(in) :: default
    character(len=:), allocatable :: s
    v = default
    if (self%has(key)) then
      s = strip(self%raw(key))
      v = s == "true"
    end if
  end function toml_get_bool

  !> Integers of a flat array text "[1, 2, 3]".
  function parse_int_array(text) result(a)
    character(len=*), intent(in) :: text
    integer, allocatable :: a(:)
    character(len=:), allocatable :: t
    integer :: i, start, v, ios
    allocate(a(0))
    t = strip(text)
    if (len(t) < 2) return
    t = t(2:len(t) - 1) // ","
    start = 1
    do i = 1, len(t)
      if (t(i:i) == ",") then
        if (len(strip(t(start:i - 1))) > 0) then
          read(t(start:i - 1), *, iostat=ios) v
          if (ios == 0) a = [a, v]
        end if
        start = i + 1
      end if
    end do
  end function parse_int_array

  function toml_get_int_array(self, key) result(a)
    class(TomlDoc), intent(in) :: self
    character(len=*), intent(in) :: key
    integer, allocatable :: a(:)
    a = parse_int_array(self%raw(key))
  end function toml_get_int_array

  !> Distinct next components of keys starting with prefix // "."
  function toml_keys_under(self, prefix) result(names)
    class(TomlDoc), intent(in) :: self
    character(len=*), intent(in) :: prefix
    character(len=64), allocatable :: names(:)
    character(len=:), allocatable :: rest
    integer :: k, dot, j
    logical :: seen
    allocate(names(0))
    do k = 1, self%n
      if (len(self%e(k)%key) <= len(prefix) + 1) cycle
      if (self%e(k)%key(1:len(prefix) + 1) /= prefix // ".") cycle
      rest = self%e(k)%key(len(prefix) + 2:)
      dot = index(rest, ".")
      if (dot > 0) rest = rest(1:dot - 1)
      seen = .false.
      do j = 1, size(names)
        if (trim(names(j)) == rest) seen = .true.
      end do
      if (.not. seen) names = [names, [character(len=64) :: rest]]
    end do
  end function toml_keys_under

  !> Reads ferebus.toml into the typed configuration.
  subroutine read_config(path, cfg, ok, msg)
    character(len=*), intent(in) :: path
    type(FerebusConfig), intent(out) :: cfg
    logical, intent(out) :: ok
    character(len=:), allocatable, intent(out) :: msg
    type(TomlDoc) :: doc
    character(len=:), allocatable :: atoms, item, props, ptable, pre
    character(len=64), allocatable :: knames(:)
    integer :: k, p0, p1, a0, b0
    call doc%load(path, ok)
    if (.not. ok) then
      msg = doc%error
      return
    end if
    ok = .false.            ! until every key has been read
    cfg%system_name = doc%get_str("system.name", "")
    cfg%natoms = doc%get_int("system.natoms", 0)
    cfg%nfeats = doc%get_int("system.nfeats", 3 * cfg%natoms - 6)
    ! atoms = [{name="O1", alf=[1, 2, 3]}, ...]
    allocate(cfg%atoms(0))
    atoms = doc%raw("system.atoms")
    p0 = index(atoms, "{")
    do while (p0 > 0)
      p1 = index(atoms(p0:), "}") + p0 - 1
      item = atoms(p0 + 1:p1 - 1)
      cfg%atoms = [cfg%atoms, AtomConfig()]
      a0 = index(item, "name")
      if (a0 > 0) then
        b0 = index(item(a0:), '"') + a0
        cfg%atoms(size(cfg%atoms))%name = item(b0:index(item(b0:), '"') + b0 - 2)
      end if
      a0 = index(item, "alf")
      if (a0 > 0) cfg%atoms(size(cfg%atoms))%alf = parse_int_array(item(index(item(a0:), "[") + a0 - 1: &
                                                                          index(item(a0:), "]") + a0 - 1))
      atoms = atoms(p1 + 1:)
      p0 = index(atoms, "{")
    end do
    props = strip(doc%raw("system.properties"))
    allocate(cfg%properties(0))
    if (len(props) > 0 .and. props /= '"all"') then
      props = props(2:len(props) - 1) // ","
      p0 = 1
      do k = 1, len(props)
        if (props(k:k) == ",") then
          if (len(unquote(props(p0:k - 1))) > 0) &
            cfg%properties = [cfg%properties, [character(len=64) :: unquote(props(p0:k - 1))]]
          p0 = k + 1
        end if
      end do
    end if
    cfg%mean = doc%get_str("model.mean", "constant")
    cfg%kernel = doc%get_str("model.kernel", "")
    cfg%likelihood = doc%get_str("model.likelihood", "marginal")
    cfg%noise = doc%get_real("model.noise", 1e-10_wp)
    cfg%optimise_noise = doc%get_bool("model.optimise_noise", .false.)
    cfg%noise_max = doc%get_real("model.noise_max", 1e-6_wp)
    cfg%search_min = doc%get_real("optimiser.search_min", 0.0_wp)
    cfg%search_max = doc%get_real("optimiser.search_max", 1.0_wp)
    cfg%swarm_size = doc%get_int("optimiser.pso.swarm_size", 50)
    cfg%iterations = doc%get_int("optimiser.pso.iterations", 1000)
    cfg%inertia_weight = doc%get_real("optimiser.pso.inertia_weight", 0.789_wp)
    cfg%cognitive = doc%get_real("optimiser.pso.cognitive_learning_rate", 1.494_wp)
    cfg%social = doc%get_real("optimiser.pso.social_learning_rate", 1.494_wp)
    cfg%stopping = sanitise(doc%get_str("optimiser.pso.stopping_criteria", "none"))
    select case (cfg%stopping)
    case ("none", "relative_change")
    case ("potential_based", "sign_test")
      msg = "stopping_criteria " // cfg%stopping // " has no implementation (the reference parses it but " // &
            "never implements it)"
      return
    case default
      msg = "unknown stopping_criteria: " // cfg%stopping
      return
    end select
    cfg%tolerance = doc%get_real("optimiser.pso." // cfg%stopping // ".tolerance", 1e-6_wp)
    cfg%stall_iterations = doc%get_int("optimiser.pso." // cfg%stopping // ".stall_iterations", 20)
    cfg%swarm_updater = sanitise(doc%get_str("optimiser.pso.swarm_updater", "global"))
    select case (cfg%swarm_updater)
    case ("global", "default")
      cfg%swarm_updater = "global"
    case ("forced")
      cfg%delta = doc%get_real("optimiser.pso.forced.delta", 1e-7_wp)
    case default
      msg = "unknown swarm_updater: " // cfg%swarm_updater
      return
    end select
    ptable = sanitise(doc%get_str("optimiser.pso.parameter_updater", "constant"))
    select case (ptable)
    case ("constant", "none")
      cfg%parameter_updater = "constant"
    case ("unified_adaptive", "uapso", "ua")
      cfg%parameter_updater = "unified_adaptive"
    case ("individual_adaptive", "iapso", "ia")
      cfg%parameter_updater = "individual_adaptive"
    case ("random_unified", "random", "ru")
      cfg%parameter_updater = "random_unified"
    case ("random_individual", "ri")
      cfg%parameter_updater = "random_individual"
    case default
      msg = "unknown parameter_updater: " // ptable
      return
    end select
    if (cfg%parameter_updater == "unified_adaptive" .or. cfg%parameter_updater == "individual_adaptive") then
      ! the table named as given, else [adaptive], else the other variant's
      ! (src/config/optimiser/pso/pso.f90:124-139)
      if (size(doc%keys_under("optimiser.pso." // ptable)) == 0) then
        if (size(doc%keys_under("optimiser.pso.adaptive")) > 0) then
          ptable = "adaptive"
        else if (cfg%parameter_updater == "unified_adaptive") then
          ptable = "individual_adaptive"
        else
          ptable = "unified_adaptive"
        end if
      end if
      pre = "optimiser.pso." // ptable // "."
      cfg%min_w = doc%get_real(pre // "min_inertia_weight", 0.0_wp)
      cfg%max_w = doc%get_real(pre // "max_inertia_weight", 1.0_wp)
      cfg%n_w = doc%get_int(pre // "n_inertia_weight", 10)
      cfg%min_c1 = doc%get_real(pre // "min_cognitive_learning_rate", 0.0_wp)
      cfg%max_c1 = doc%get_real(pre // "max_cognitive_learning_rate", 4.0_wp)
      cfg%n_c1 = doc%get_int(pre // "n_cognitive_learning_rate", 10)
      cfg%min_c2 = doc%get_real(pre // "min_social_learning_rate", 0.0_wp)
      cfg%max_c2 = doc%get_real(pre // "max_social_learning_rate", 4.0_wp)
      cfg%n_c2 = doc%get_int(pre // "n_social_learning_rate", 10)
      cfg%reward_step = doc%get_real(pre // "reward_step_length", 0.01_wp)
      cfg%penalty_step = doc%get_real(pre // "penalty_step_length", 0.01_wp)
      cfg%success_threshold = doc%get_real(pre // "success_threshold", 0.5_wp)
      if (min(cfg%n_w, cfg%n_c1, cfg%n_c2) < 2) then
        msg = "adaptive parameter updater: every n_* must be at least 2"
        return
      end if
    end if
    ! build-side extension (no reference key): L-BFGS polishing of gbest
    cfg%lbfgs_every = doc%get_int("optimiser.pso.lbfgs_every", 0)
    cfg%lbfgs_steps = doc%get_int("optimiser.pso.lbfgs_steps", 5)
    knames = doc%keys_under("kernels")
    allocate(cfg%kernels(size(knames)))
    do k = 1, size(knames)
      cfg%kernels(k)%name = knames(k)
      cfg%kernels(k)%ktype = doc%get_str("kernels." // trim(knames(k)) // ".type", "rbf")
      if (doc%has("kernels." // trim(knames(k)) // ".active_dimensions")) &
        cfg%kernels(k)%active_dims = doc%get_int_array("kernels." // trim(knames(k)) // ".active_dimensions")
    end do
    knames = doc%keys_under("notes")
    allocate(cfg%note_keys(size(knames)), cfg%note_values(size(knames)))
    do k = 1, size(knames)
      cfg%note_keys(k) = knames(k)
      cfg%note_values(k) = doc%get_str("notes." // trim(knames(k)), "")
    end do
    ok = .true.
    msg = ""
  end subroutine read_config

  !> Reads a training-set CSV: x (ntrain, nfeats) and the property columns.
  subroutine read_training_set(path, nfeats, x, y, labels, ok, msg)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nfeats
    real(wp), allocatable, intent(out) :: x(:, :), y(:, :)
    character(len=64), allocatable, intent(out) :: labels(:)
    logical, intent(out) :: ok
    character(len=:), allocatable, intent(out) :: msg
    integer :: u, ios, nrow, ncol, i, c, nf, p0, k, ny
    character(len=65536) :: line
    character(len=64), allocatable :: head(:)
    integer, allocatable :: kind(:)          ! 0 index, 1 feature, 2 property
    real(wp), allocatable :: vals(:)
    ok = .false.
    open(newunit=u, file=path, status="old", action="read", iostat=ios)
    if (ios /= 0) then
      msg = "cannot open " // path
      return
    end if
    read(u, '(A)') line
    allocate(head(0))
    p0 = 1
    do k = 1, len_trim(line) + 1
      if (k > len_trim(line) .or. line(k:k) == ",") then
        head = [head, [character(len=64) :: strip(line(p0:k - 1))]]
        p0 = k + 1
      end if
    end do
    ncol = size(head)
    allocate(kind(ncol))
    nf = 0
    do c = 1, ncol
      if (len_trim(head(c)) == 0) then
        kind(c) = 0
      else if (head(c)(1:1) == "f" .and. verify(trim(head(c)(2:)), "0123456789") == 0 &
               .and. len_trim(head(c)) > 1) then
        kind(c) = 1
        nf = nf + 1
        if (nf > nfeats) kind(c) = 3     ! beyond system.nfeats: ignored
      else
        kind(c) = 2
      end if
    end do
    nrow = 0
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      if (len_trim(line) > 0) nrow = nrow + 1
    end do
    rewind(u)
    read(u, '(A)') line
    ny = count(kind == 2)
    allocate(x(nrow, min(nf, nfeats)), y(nrow, ny), vals(ncol), labels(ny))
    labels = pack(head, kind == 2)
    do i = 1, nrow
      read(u, *, iostat=ios) vals
      if (ios /= 0) then
        msg = "malformed row in " // path
        close(u)
        return
      end if
      x(i, :) = pack(vals, kind == 1)
      y(i, :) = pack(vals, kind == 2)
    end do
    close(u)
    ok = .true.
    msg = ""
  end subroutine read_training_set
end module fb_config

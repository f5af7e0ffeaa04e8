!> Qvort-style driver with the velocity stage on the GPU.
!>
!> Keeps the reference's run.in (label value lines, read as cdata.f90:420-807
!> does, unknown labels ignored) and its f(:) arrays in the host layout
!> x(3,pcount), infront, behind (cdata.f90:20-41).  The time loop follows
!> run.f90:36-169 for the hot path: velocity (ghostp + construct_tree +
!> calc_velocity), the AB ladder of pmotion (or pmotion_RK3 with
!> runge_kutta T, run.f90:50-54), the periodic wrap, t = t + dt,
!> and every `shots` steps the restart file data/var.dat (data_dump,
!> output.f90:272-285), the ts.log line and the var%04d.log snapshot
!> (output.f90:55-143).  With tree_print T, the leaves of the step's tree go
!> to data/tree_mesh%03d.log as construct_tree prints them (tree.f90:116-123).
!> A data/var.dat present at start-up restarts the run from it as
!> init_setup does (initial.f90:168-169, data_restore initial.f90:500-529):
!> the loop continues at the dumped itime + 1 (at 1 with restart_rewind T).
!> With line_ops T (this driver's own label, default F) every step runs the
!> line operations of run.f90:60-72 after the motion: pinsert, and when
!> recon_shots divides itime the neighbour search on the device tree,
!> precon (recon_type 'original') and premove; they run on the host through
!> the library (qv_line_*), the state going down and back up around them.
!> With mesh_size > 0 the velocity on setup_mesh's grid (initial.f90:541-574)
!> is summed after the motion when mesh_shots divides itime and itime >
!> mesh_print_delay (run.f90:56-58, mesh_velocity timestep.f90:374-452) by
!> qv_probe_velocity, and data/mesh%03d.dat is written as print_mesh does
!> (output.f90:305-321) when shots and mesh_shots divide itime, numbered
!> (itime - mesh_print_delay)/mesh_shots (output.f90:32-41).  u_norm is 0:
!> the driver has no normal fluid (normal_velocity zero).  In Tree mode the
!> mesh sums carry Gamma/(4 pi); mesh_reference_scale T (this driver's own
!> label) keeps the factors of tree_walk_general (tree.f90:460,479) so that
!> the files match the reference's.
!>
!> run.in labels used: nsteps shots dt delta box_size velocity boundary
!> quant_circ corea tree_theta tree_extra_correction deriv_order runge_kutta initf pcount
!> tree_print restart_rewind recon_shots mesh_size mesh_shots mesh_print_delay
!> normal_velocity remove_mean_pos switch_off_recon, and gpu_precision
!> (fp64 | fp32c), line_ops (T | F) and mesh_reference_scale (T | F), this
!> driver's own.  Labels that change the motion and are not carried here
!> (recon_type other than original, delta_adapt, phonon_emission,
!> hyperviscosity, alpha, force, sticky_z_boundary, kelvin_wave_casc,
!> svistunov_hamiltonian, killing_sphere_on, moving_box,
!> multiple_initial_loop, particles, vortex injection) stop the run when
!> set to anything but their defaults.
!>
!> var.dat holds f(:) as the reference's type qvort (cdata.f90:20-38); the
!> components this driver does not carry (x_interp, acc, u_mf, u_s_LI,
!> u_s_BS, the ghosts, f_mf, stretch, curv, closest*, pinned*, delta,
!> t_recon) are written as zero, u_sup as u; a restart reads x, u, u1..u3,
!> infront and behind.  With runge_kutta T the reference prints the leaves of
!> all three stage trees into one tree_mesh file; this driver prints the first
!> stage's tree (the one of the step's start).
!> initf: single_loop (initial_loop.f90:11-39) or state_file (state.bin:
!> int32 version, pcount, nsteps, flags, nsample; real64 x(3,pcount);
!> int32 infront(pcount), behind(pcount); real64 u1, u2, u3 (3,pcount)).
program qv_run
  use iso_c_binding
  use qv_gpu
  implicit none
  real(c_double), parameter :: pi = 3.14159265358979324_c_double
  ! f(:) arrays
  ! arrays hold capacity >= pcount slots; the first pcount are f(:)
  ! (pinned host memory from qv_host_alloc where the runtime gives it,
  ! ordinary memory otherwise: host_pinned)
  real(c_double), pointer, contiguous :: x(:,:) => null(), u(:,:) => null(), &
    u1(:,:) => null(), u2(:,:) => null(), u3(:,:) => null()
  integer(c_int32_t), pointer, contiguous :: infront(:) => null(), behind(:) => null()
  ! ghost zones and neighbour-search results, for the line operations
  real(c_double), pointer, contiguous :: gi(:,:) => null(), gb(:,:) => null(), &
    gii(:,:) => null(), gbb(:,:) => null(), closestd(:) => null()
  integer(c_int32_t), pointer, contiguous :: closest(:) => null()
  logical :: host_pinned = .true., host_decided = .false.
  integer(c_int32_t) :: pcount
  ! run.in
  integer :: nsteps = 1, shots = 1, init_pcount = 100
  real(c_double) :: dt = 1e-6_c_double, delta = 5e-4_c_double, box_size = 0._c_double
  real(c_double) :: quant_circ = 9.97e-4_c_double, corea = 8.244023e-9_c_double
  real(c_double) :: tree_theta = 0._c_double
  logical :: tree_extra_correction = .true., runge_kutta = .false.
  character(len=30) :: velocity = 'Tree', boundary = 'periodic', deriv_order = 'second'
  character(len=30) :: initf = 'single_loop', gpu_precision = 'fp64'
  logical :: tree_print = .false., restart_rewind = .false., line_ops = .false.
  integer :: recon_shots = 1
  integer :: mesh_size = 0, mesh_shots = 100, mesh_print_delay = 0
  logical :: mesh_reference_scale = .false.
  character(len=30) :: normal_velocity = 'zero'
  logical :: remove_mean_pos = .false., switch_off_recon = .false.
  integer :: inject_skip = 10000000
  type(qv_line_counts) :: counts
  ! the mesh points (3, mesh_size**3) in the order of mesh(k,j,i) and their
  ! superfluid velocity
  real(c_double), allocatable, target :: mesh_x(:,:), mesh_u(:,:)
  ! type qvort (cdata.f90:20-38), component for component: the record layout
  ! of f(:) in data/var.dat
  type :: qvort_record
    real(c_double) :: x(3), x_interp(3)
    real(c_double) :: u(3), u1(3), u2(3), u3(3)
    real(c_double) :: acc(3)
    real(c_double) :: u_sup(3), u_mf(3)
    real(c_double) :: u_s_LI(3), u_s_BS(3)
    real(c_double) :: ghosti(3), ghostb(3)
    real(c_double) :: ghostii(3), ghostbb(3)
    real(c_double) :: f_mf(3), stretch, curv
    integer(c_int32_t) :: infront, behind
    integer(c_int32_t) :: closest
    real(c_double) :: closestd
    real(c_double) :: closestd_loop
    integer(c_int32_t) :: pinnedi  ! logical(4) in type qvort: its 4 bytes (0 = .false.)
    integer(c_int32_t) :: pinnedb
    real(c_double) :: delta
    real(c_double) :: t_recon(2)
  end type
  integer :: nstart = 1
  ! bytes of one f(:) element in data/var.dat
  integer(c_int64_t), parameter :: kRecordBytes = 460
  ! state
  type(qv_params) :: p
  type(c_ptr) :: ctx
  type(qv_timings) :: tm
  integer(c_int64_t), target :: evals
  integer :: itime, i
  real(c_double) :: t = 0._c_double
  logical :: have_hist = .false.

  call read_run_in()
  call init_filament()
  call setup_mesh()

  p%quant_circ = quant_circ
  p%corea = corea
  p%box_size = box_size
  p%delta = delta
  p%tree_theta = tree_theta
  p%dt = dt
  select case (velocity)
    case ('BS')
      p%velocity = QV_VEL_BS
    case ('Tree')
      p%velocity = QV_VEL_TREE
    case default
      call stop_with('init_setup', 'velocity must be BS or Tree on the GPU')
  end select
  select case (boundary)
    case ('periodic')
      p%boundary = QV_BC_PERIODIC
    case ('openx')
      p%boundary = QV_BC_OPENX
    case ('openxy')
      p%boundary = QV_BC_OPENXY
    case ('open')
      p%boundary = QV_BC_OPEN
    case default
      call stop_with('init_setup:', 'incorrect boundary parameter')
  end select
  p%tree_extra_correction = merge(1, 0, tree_extra_correction)
  p%deriv_order = merge(4, 2, deriv_order == 'fourth')
  p%precision = merge(QV_FP32C, QV_FP64, gpu_precision == 'fp32c')
  p%flags = merge(QV_FLAG_REMOVE_MEAN_POS, 0, remove_mean_pos)

  call qv_check(qv_create(p, ctx), 'qv_create')
  call qv_check(qv_upload(ctx, pcount, x, infront, behind), 'qv_upload')
  if (have_hist) call qv_check(qv_upload_history(ctx, c_loc(u1), c_loc(u2), c_loc(u3)), &
                               'qv_upload_history')
  call execute_command_line('mkdir -p data')
  write(*,'(a,i0,a,a,a,a)') ' qv_run: ', count(infront(1:pcount) > 0), ' points, velocity ', &
        trim(velocity), ', precision ', trim(gpu_precision)

  do itime = nstart, nsteps
    if (runge_kutta) then
      if (tree_print .and. tree_theta > 0._c_double .and. mod(itime, shots) == 0) then
        ! the step's first tree: the velocity stage at the step's start
        call qv_check(qv_velocity(ctx, c_null_ptr, c_null_ptr, c_null_ptr, c_loc(evals)), &
                      'qv_velocity')
        call print_tree_mesh(itime / shots)
      end if
      call qv_check(qv_rk3_step(ctx), 'qv_rk3_step')
    else
      call qv_check(qv_velocity(ctx, c_null_ptr, c_null_ptr, c_null_ptr, c_loc(evals)), &
                    'qv_velocity')
      if (tree_print .and. tree_theta > 0._c_double .and. mod(itime, shots) == 0) &
        call print_tree_mesh(itime / shots)
      call qv_check(qv_ab_step(ctx), 'qv_ab_step')
    end if
    if (mesh_size > 0) then
      if (mod(itime, mesh_shots) == 0 .and. itime > mesh_print_delay) call mesh_velocity()
    end if
    if (line_ops) call line_step()
    call qv_check(qv_enforce_periodic(ctx), 'qv_enforce_periodic')
    ! in openx / openxy boxes the boundary step removes loops that leave
    ! through an open face: the host's links follow
    if (p%boundary == QV_BC_OPENX .or. p%boundary == QV_BC_OPENXY) &
      call qv_check(qv_download_links(ctx, infront, behind), 'qv_download_links')
    t = t + dt
    if (mod(itime, shots) == 0) then
      ! the device holds the state, the line operations' changes included
      call qv_check(qv_download(ctx, c_loc(x), c_loc(u), c_loc(u1), c_loc(u2), c_loc(u3)), &
                    'qv_download')
      call data_dump()
      call print_info(itime / shots)
      call printf(itime / shots)
      if (mesh_size > 0 .and. mod(itime, mesh_shots) == 0) then
        if (itime > mesh_print_delay) call print_mesh((itime - mesh_print_delay) / mesh_shots)
      end if
    end if
  end do
  call qv_check(qv_get_timings(ctx, tm), 'qv_get_timings')
  write(*,'(a,f10.3,a,f10.3,a,f10.3,a)') ' last stage: build ', tm%build_ms, ' ms, walk ', &
        tm%walk_ms, ' ms, total ', tm%total_ms, ' ms'
  call qv_check(qv_destroy(ctx), 'qv_destroy')

contains

  subroutine stop_with(location, message)
    character(len=*), intent(in) :: location, message
    write (*,*) '-------------------------FATAL ERROR-------------------------'
    write (*,*) trim(location), ": ", trim(message)
    write (*,*) '-------------------------------------------------------------'
    stop 1
  end subroutine

  !> label value lines; the first blank ends the label (cdata.f90:437-441)
  subroutine read_run_in()
    character(len=100) :: buffer, label
    integer :: pos, ios, fh
    open(newunit=fh, file='run.in', status='old', action='read', iostat=ios)
    if (ios /= 0) call stop_with('read_run_in', 'run.in not found')
    do
      read(fh, '(A)', iostat=ios) buffer
      if (ios /= 0) exit
      pos = scan(buffer, ' ')
      label = buffer(1:pos)
      buffer = buffer(pos+1:)
      select case (label)
        case ('nsteps')
          read(buffer, *, iostat=ios) nsteps
        case ('shots')
          read(buffer, *, iostat=ios) shots
        case ('dt')
          read(buffer, *, iostat=ios) dt
        case ('delta')
          read(buffer, *, iostat=ios) delta
        case ('box_size')
          read(buffer, *, iostat=ios) box_size
        case ('velocity')
          read(buffer, *, iostat=ios) velocity
        case ('boundary')
          read(buffer, *, iostat=ios) boundary
        case ('quant_circ')
          read(buffer, *, iostat=ios) quant_circ
        case ('corea')
          read(buffer, *, iostat=ios) corea
        case ('tree_theta')
          read(buffer, *, iostat=ios) tree_theta
        case ('tree_extra_correction')
          read(buffer, *, iostat=ios) tree_extra_correction
        case ('deriv_order')
          read(buffer, *, iostat=ios) deriv_order
        case ('initf')
          read(buffer, *, iostat=ios) initf
        case ('pcount')
          read(buffer, *, iostat=ios) init_pcount
        case ('gpu_precision')
          read(buffer, *, iostat=ios) gpu_precision
        case ('runge_kutta')
          read(buffer, *, iostat=ios) runge_kutta
        case ('tree_print')
          read(buffer, *, iostat=ios) tree_print
        case ('restart_rewind')
          read(buffer, *, iostat=ios) restart_rewind
        case ('recon_shots')
          read(buffer, *, iostat=ios) recon_shots
        case ('line_ops')
          read(buffer, *, iostat=ios) line_ops
        case ('mesh_size')
          read(buffer, *, iostat=ios) mesh_size
        case ('mesh_shots')
          read(buffer, *, iostat=ios) mesh_shots
        case ('mesh_print_delay')
          read(buffer, *, iostat=ios) mesh_print_delay
        case ('mesh_reference_scale')
          read(buffer, *, iostat=ios) mesh_reference_scale
        case ('normal_velocity')
          read(buffer, *, iostat=ios) normal_velocity
        case ('remove_mean_pos')
          read(buffer, *, iostat=ios) remove_mean_pos
        case ('switch_off_recon')
          read(buffer, *, iostat=ios) switch_off_recon
        case ('inject_skip')
          read(buffer, *, iostat=ios) inject_skip
        case default
          call refuse_unsupported(label, buffer)
      end select
      ios = 0
    end do
    close(fh)
    ! vortex_inject runs when inject_skip divides itime (run.f90:38-40)
    if (inject_skip <= nsteps) call unsupported('inject_skip', 'vortex injection')
    if (.not. box_size > 0._c_double) call stop_with('init_setup:', 'box size is less than zero')
    ! init_setup's checks of the velocity choice (initial.f90:355-363)
    if (velocity == 'Tree') then
      if (tree_theta < epsilon(0._c_double)) call stop_with('init.mod:init_setup', &
        'runnning with tree velocity but tree_theta is zero')
      if (tree_theta > 0.6_c_double) call warning('init_setup', &
        'tree_theta>0.6, the vortices may be unstable')
    end if
  end subroutine

  !> run.in switches of the reference that change the motion of the points
  !> and that this driver does not carry: a run.in that sets one of them to
  !> anything but its default (cdata.f90:196-416) stops here rather than run
  !> different physics.  Labels of output, diagnostics and initial
  !> conditions the driver does not read are ignored, as before.
  subroutine refuse_unsupported(label, value)
    character(len=*), intent(in) :: label, value
    character(len=30) :: word
    logical :: flag
    real(c_double) :: a(2)
    integer :: n, ios
    select case (label)
      case ('recon_type')
        read(value, *, iostat=ios) word
        if (ios == 0 .and. word /= 'original') call unsupported(label, word)
      case ('force')
        read(value, *, iostat=ios) word
        if (ios == 0 .and. word /= 'off') call unsupported(label, word)
      case ('delta_adapt', 'phonon_emission', 'hyperviscosity', 'sticky_z_boundary', &
            'kelvin_wave_casc', 'svistunov_hamiltonian', 'killing_sphere_on', 'moving_box', &
            'multiple_initial_loop')
        read(value, *, iostat=ios) flag
        if (ios == 0 .and. flag) call unsupported(label, 'T')
      case ('alpha')
        a = 0._c_double
        read(value, *, iostat=ios) a
        if (any(a /= 0._c_double)) call unsupported(label, 'mutual friction')
      case ('quasi_pcount', 'part_count')
        read(value, *, iostat=ios) n
        if (ios == 0 .and. n > 0) call unsupported(label, 'particles')
    end select
  end subroutine

  subroutine unsupported(label, value)
    character(len=*), intent(in) :: label, value
    call stop_with('qv_run:read_run_in', trim(label)//' '//trim(value)// &
                   ' changes the motion and is not implemented by this driver')
  end subroutine

  subroutine init_filament()
    integer :: fh, version, nst, flags, nsample
    logical :: restart
    inquire(file='./data/var.dat', exist=restart)
    if (restart) then
      call data_restore()
      return
    end if
    select case (initf)
      case ('single_loop')
        ! setup_single_loop, initial_loop.f90:11-39
        pcount = init_pcount
        call alloc_f()
        block
          real(c_double) :: radius
          radius = (0.75_c_double * pcount * delta) / (2 * pi)
          do i = 1, pcount
            x(1, i) = radius * sin(pi * real(2 * i - 1, c_double) / pcount)
            x(2, i) = radius * cos(pi * real(2 * i - 1, c_double) / pcount)
            x(3, i) = 0._c_double
            infront(i) = i + 1
            behind(i) = i - 1
          end do
          infront(pcount) = 1
          behind(1) = pcount
        end block
      case ('state_file')
        open(newunit=fh, file='state.bin', access='stream', form='unformatted', status='old', &
             action='read')
        read(fh) version, pcount, nst, flags, nsample
        call alloc_f()
        read(fh) x(:, 1:pcount)
        read(fh) infront(1:pcount)
        read(fh) behind(1:pcount)
        read(fh) u1(:, 1:pcount), u2(:, 1:pcount), u3(:, 1:pcount)
        close(fh)
        have_hist = any(u1 /= 0._c_double) .or. any(u2 /= 0._c_double) .or. &
                    any(u3 /= 0._c_double)
      case default
        call stop_with('init_setup', 'initf must be single_loop or state_file')
    end select
  end subroutine

  !> f(:) in the file is its elements' components back to back, 460 bytes
  !> each (the Fortran runtime transfers a derived type component by
  !> component, without the alignment padding of the memory layout); these
  !> copy whole runs of components between the two layouts, where the
  !> runtime's item-by-item transfer takes microseconds per component
  subroutine record_runs(rec, nrun, mo, fo, ln)
    type(qvort_record), target, intent(in) :: rec
    integer, intent(out) :: nrun
    integer, intent(out) :: mo(27), fo(27), ln(27)
    integer(c_intptr_t) :: base, off(27)
    integer :: sz(27), k, f
    base = transfer(c_loc(rec), base)
    off = [transfer(c_loc(rec%x), base), transfer(c_loc(rec%x_interp), base), &
           transfer(c_loc(rec%u), base), transfer(c_loc(rec%u1), base), &
           transfer(c_loc(rec%u2), base), transfer(c_loc(rec%u3), base), &
           transfer(c_loc(rec%acc), base), transfer(c_loc(rec%u_sup), base), &
           transfer(c_loc(rec%u_mf), base), transfer(c_loc(rec%u_s_LI), base), &
           transfer(c_loc(rec%u_s_BS), base), transfer(c_loc(rec%ghosti), base), &
           transfer(c_loc(rec%ghostb), base), transfer(c_loc(rec%ghostii), base), &
           transfer(c_loc(rec%ghostbb), base), transfer(c_loc(rec%f_mf), base), &
           transfer(c_loc(rec%stretch), base), transfer(c_loc(rec%curv), base), &
           transfer(c_loc(rec%infront), base), transfer(c_loc(rec%behind), base), &
           transfer(c_loc(rec%closest), base), transfer(c_loc(rec%closestd), base), &
           transfer(c_loc(rec%closestd_loop), base), transfer(c_loc(rec%pinnedi), base), &
           transfer(c_loc(rec%pinnedb), base), transfer(c_loc(rec%delta), base), &
           transfer(c_loc(rec%t_recon), base)] - base
    sz = [(24, k = 1, 16), 8, 8, 4, 4, 4, 8, 8, 4, 4, 8, 16]
    nrun = 0
    f = 0
    do k = 1, 27
      if (nrun > 0) then
        if (int(off(k)) - f == mo(nrun) - fo(nrun)) then
          ln(nrun) = ln(nrun) + sz(k)
          f = f + sz(k)
          cycle
        end if
      end if
      nrun = nrun + 1
      mo(nrun) = int(off(k)); fo(nrun) = f; ln(nrun) = sz(k)
      f = f + sz(k)
    end do
  end subroutine

  subroutine pack_records(rec, bytes)
    type(qvort_record), target, intent(in) :: rec(:)
    integer(c_int8_t), allocatable, intent(out) :: bytes(:)
    integer(c_int8_t), pointer :: mem(:)
    integer :: nrun, mo(27), fo(27), ln(27), r
    integer(c_int64_t) :: i, m, f, stride
    allocate(bytes(size(rec, kind=c_int64_t) * kRecordBytes))
    if (size(rec) == 0) return
    stride = storage_size(rec(1)) / 8
    call record_runs(rec(1), nrun, mo, fo, ln)
    call c_f_pointer(c_loc(rec), mem, [size(rec, kind=c_int64_t) * stride])
    do i = 0, size(rec, kind=c_int64_t) - 1
      m = i * stride
      f = i * kRecordBytes
      do r = 1, nrun
        bytes(f + fo(r) + 1:f + fo(r) + ln(r)) = mem(m + mo(r) + 1:m + mo(r) + ln(r))
      end do
    end do
  end subroutine

  subroutine unpack_records(bytes, rec)
    integer(c_int8_t), intent(in) :: bytes(:)
    type(qvort_record), target, intent(inout) :: rec(:)
    integer(c_int8_t), pointer :: mem(:)
    integer :: nrun, mo(27), fo(27), ln(27), r
    integer(c_int64_t) :: i, m, f, stride
    if (size(rec) == 0) return
    stride = storage_size(rec(1)) / 8
    call record_runs(rec(1), nrun, mo, fo, ln)
    call c_f_pointer(c_loc(rec), mem, [size(rec, kind=c_int64_t) * stride])
    do i = 0, size(rec, kind=c_int64_t) - 1
      m = i * stride
      f = i * kRecordBytes
      do r = 1, nrun
        mem(m + mo(r) + 1:m + mo(r) + ln(r)) = bytes(f + fo(r) + 1:f + fo(r) + ln(r))
      end do
    end do
  end subroutine

  !> data_restore's reads (initial.f90:506-525); quasi particles and
  !> particles are not part of this driver and must be absent
  subroutine data_restore()
    integer :: fh, dummy_itime, quasi_pcount, part_count, recon_count
    type(qvort_record), allocatable, target :: rec(:)
    integer(c_int8_t), allocatable :: bytes(:)
    open(newunit=fh, file='./data/var.dat', form='unformatted', status='old', action='read')
    read(fh) pcount
    read(fh) recon_count
    counts%recon_count = recon_count
    read(fh) dummy_itime
    read(fh) t
    allocate(rec(pcount), bytes(int(pcount, c_int64_t) * kRecordBytes))
    read(fh) bytes
    call unpack_records(bytes, rec)
    deallocate(bytes)
    read(fh) quasi_pcount
    if (quasi_pcount > 0) call stop_with('data_restore', 'quasi particles are not supported')
    read(fh) part_count
    if (part_count > 0) call stop_with('data_restore', 'particles are not supported')
    close(fh)
    call alloc_f()
    do i = 1, pcount
      x(:, i) = rec(i)%x
      u(:, i) = rec(i)%u
      u1(:, i) = rec(i)%u1
      u2(:, i) = rec(i)%u2
      u3(:, i) = rec(i)%u3
      infront(i) = rec(i)%infront
      behind(i) = rec(i)%behind
    end do
    have_hist = any(u1 /= 0._c_double) .or. any(u2 /= 0._c_double) .or. &
                any(u3 /= 0._c_double)
    if (restart_rewind) then
      nstart = 1
    else
      nstart = dummy_itime + 1
    end if
    write(*,*) 'data read in from dump file at t=', t
  end subroutine

  !> data_dump's writes (output.f90:274-283)
  subroutine data_dump()
    integer :: fh
    type(qvort_record), allocatable, target :: rec(:)
    integer(c_int8_t), allocatable :: bytes(:)
    allocate(rec(pcount))
    do i = 1, pcount
      rec(i)%x = x(:, i)
      rec(i)%x_interp = 0._c_double
      rec(i)%u = u(:, i)
      rec(i)%u1 = u1(:, i)
      rec(i)%u2 = u2(:, i)
      rec(i)%u3 = u3(:, i)
      rec(i)%acc = 0._c_double
      rec(i)%u_sup = u(:, i)
      rec(i)%u_mf = 0._c_double
      rec(i)%u_s_LI = 0._c_double
      rec(i)%u_s_BS = 0._c_double
      rec(i)%ghosti = 0._c_double
      rec(i)%ghostb = 0._c_double
      rec(i)%ghostii = 0._c_double
      rec(i)%ghostbb = 0._c_double
      rec(i)%f_mf = 0._c_double
      rec(i)%stretch = 0._c_double
      rec(i)%curv = 0._c_double
      rec(i)%infront = infront(i)
      rec(i)%behind = behind(i)
      rec(i)%closest = 0
      rec(i)%closestd = 0._c_double
      rec(i)%closestd_loop = 0._c_double
      rec(i)%pinnedi = 0
      rec(i)%pinnedb = 0
      rec(i)%delta = 0._c_double
      rec(i)%t_recon = 0._c_double
    end do
    open(newunit=fh, file='./data/var.dat', form='unformatted', status='replace')
    write(fh) pcount
    write(fh) int(counts%recon_count, c_int32_t)
    write(fh) itime
    write(fh) t
    call pack_records(rec, bytes)
    write(fh) bytes
    deallocate(bytes)
    write(fh) 0_c_int32_t
    write(fh) 0_c_int32_t
    close(fh)
  end subroutine

  !> construct_tree's leaf print (tree.f90:116-123): every cell that is not
  !> split, in the order the reference creates them (DFS), appended to
  !> data/tree_mesh%03d.log as list-directed lines
  subroutine print_tree_mesh(filenumber)
    integer, intent(in) :: filenumber
    character(len=40) :: print_file
    integer :: fh
    integer(c_int64_t) :: nn, k
    real(c_double), allocatable, target :: nodes(:,:)
    call qv_check(qv_tree_dump(ctx, c_null_ptr, 0_c_int64_t, nn), 'qv_tree_dump')
    allocate(nodes(22, nn))
    call qv_check(qv_tree_dump(ctx, c_loc(nodes), nn, nn), 'qv_tree_dump')
    write(print_file, '(a,i3.3,a)') './data/tree_mesh', filenumber, '.log'
    open(newunit=fh, file=print_file, action='write', position='append')
    do k = 1, nn
      if (nodes(2, k) > 1._c_double) cycle
      write(fh, *) nodes(4, k), nodes(4, k) + nodes(7, k), nodes(5, k), &
                   nodes(5, k) + nodes(7, k), nodes(6, k), nodes(6, k) + nodes(7, k), &
                   int(nodes(2, k))
    end do
    close(fh)
  end subroutine

  subroutine alloc_f()
    integer :: cap
    cap = pcount
    if (line_ops) cap = 2 * pcount + 64
    call new3(x, cap); call new3(u, cap); call new3(u1, cap); call new3(u2, cap)
    call new3(u3, cap)
    call new3(gi, cap); call new3(gb, cap); call new3(gii, cap); call new3(gbb, cap)
    call newi(infront, cap); call newi(behind, cap); call newi(closest, cap)
    call new1(closestd, cap)
  end subroutine

  !> pinned host memory for an f(:) array.  The first request decides for
  !> the run: without pinned memory (no device, say) every array is ordinary
  !> memory; a later failure is fatal, so that all arrays stay of one kind
  function host_mem(bytes) result(p)
    integer(c_size_t), intent(in) :: bytes
    type(c_ptr) :: p
    p = c_null_ptr
    if (.not. host_pinned) return
    if (qv_host_alloc(bytes, p) == QV_OK) then
      host_decided = .true.
      return
    end if
    if (host_decided) call stop_with('qv_host_alloc', 'out of pinned host memory')
    host_decided = .true.
    host_pinned = .false.
    p = c_null_ptr
  end function

  !> zeroed (3, cap) / (cap) arrays
  subroutine new3(a, cap)
    real(c_double), pointer, contiguous, intent(out) :: a(:,:)
    integer, intent(in) :: cap
    type(c_ptr) :: p
    p = host_mem(int(24, c_size_t) * cap)
    if (c_associated(p)) then
      call c_f_pointer(p, a, [3, cap])
    else
      allocate(a(3, cap))
    end if
    a = 0._c_double
  end subroutine

  subroutine new1(a, cap)
    real(c_double), pointer, contiguous, intent(out) :: a(:)
    integer, intent(in) :: cap
    type(c_ptr) :: p
    p = host_mem(int(8, c_size_t) * cap)
    if (c_associated(p)) then
      call c_f_pointer(p, a, [cap])
    else
      allocate(a(cap))
    end if
    a = 0._c_double
  end subroutine

  subroutine newi(a, cap)
    integer(c_int32_t), pointer, contiguous, intent(out) :: a(:)
    integer, intent(in) :: cap
    type(c_ptr) :: p
    p = host_mem(int(4, c_size_t) * cap)
    if (c_associated(p)) then
      call c_f_pointer(p, a, [cap])
    else
      allocate(a(cap))
    end if
    a = 0
  end subroutine

  subroutine free3(a)
    real(c_double), pointer, contiguous, intent(inout) :: a(:,:)
    if (host_pinned) then
      call qv_check(qv_host_free(c_loc(a)), 'qv_host_free')
    else
      deallocate(a)
    end if
    nullify(a)
  end subroutine

  subroutine free1(a)
    real(c_double), pointer, contiguous, intent(inout) :: a(:)
    if (host_pinned) then
      call qv_check(qv_host_free(c_loc(a)), 'qv_host_free')
    else
      deallocate(a)
    end if
    nullify(a)
  end subroutine

  subroutine freei(a)
    integer(c_int32_t), pointer, contiguous, intent(inout) :: a(:)
    if (host_pinned) then
      call qv_check(qv_host_free(c_loc(a)), 'qv_host_free')
    else
      deallocate(a)
    end if
    nullify(a)
  end subroutine

  !> room for pinsert: one new point per slot at most, so 2 pcount slots
  subroutine ensure_capacity()
    integer :: cap, old
    old = size(infront)
    if (old >= 2 * pcount) return
    cap = 2 * pcount + 64
    call grow3(x); call grow3(u); call grow3(u1); call grow3(u2); call grow3(u3)
    call grow3(gi); call grow3(gb); call grow3(gii); call grow3(gbb)
    call growi(infront); call growi(behind); call growi(closest)
    call grow1(closestd)
  contains
    subroutine grow3(a)
      real(c_double), pointer, contiguous, intent(inout) :: a(:,:)
      real(c_double), pointer, contiguous :: t(:,:)
      call new3(t, cap)
      t(:, 1:old) = a(:, 1:old)
      call free3(a)
      a => t
    end subroutine
    subroutine grow1(a)
      real(c_double), pointer, contiguous, intent(inout) :: a(:)
      real(c_double), pointer, contiguous :: t(:)
      call new1(t, cap)
      t(1:old) = a(1:old)
      call free1(a)
      a => t
    end subroutine
    subroutine growi(a)
      integer(c_int32_t), pointer, contiguous, intent(inout) :: a(:)
      integer(c_int32_t), pointer, contiguous :: t(:)
      call newi(t, cap)
      t(1:old) = a(1:old)
      call freei(a)
      a => t
    end subroutine
  end subroutine

  function line_state() result(ls)
    type(qv_line_state) :: ls
    ls%pcount = pcount
    ls%capacity = size(infront)
    ls%x = c_loc(x); ls%u = c_loc(u); ls%u1 = c_loc(u1); ls%u2 = c_loc(u2); ls%u3 = c_loc(u3)
    ls%ghosti = c_loc(gi); ls%ghostb = c_loc(gb); ls%ghostii = c_loc(gii)
    ls%ghostbb = c_loc(gbb)
    ls%infront = c_loc(infront); ls%behind = c_loc(behind)
    ls%closest = c_loc(closest); ls%closestd = c_loc(closestd)
  end function

  !> run.f90:60-72 after the motion: pinsert, then every recon_shots steps
  !> pclose_tree / pclose (device), precon and premove (host); the state
  !> goes back to the device with its AB history
  subroutine line_step()
    type(qv_line_state) :: ls
    call ensure_capacity()
    ! the positions and the ghosts the operations read; the velocities and
    ! the AB history stay on the device, which gives new and cleared slots
    ! their values (qv_upload_topology)
    call qv_check(qv_download(ctx, c_loc(x), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), &
                  'qv_download')
    if (p%deriv_order == 4) then
      call qv_check(qv_download_ghosts(ctx, c_loc(gi), c_loc(gb), c_loc(gii), c_loc(gbb)), &
                    'qv_download_ghosts')
    else
      call qv_check(qv_download_ghosts(ctx, c_loc(gi), c_loc(gb), c_null_ptr, c_null_ptr), &
                    'qv_download_ghosts')
    end if
    ls = line_state()
    call qv_check(qv_line_pinsert(p, ls, counts), 'line.mod:pinsert')
    pcount = ls%pcount
    call qv_check(qv_upload_topology(ctx, pcount, c_loc(x), infront, behind), &
                  'qv_upload_topology')
    if (mod(itime, recon_shots) == 0) then
      call qv_check(qv_closest(ctx, closest, closestd), 'qv_closest')
      ls = line_state()
      if (.not. switch_off_recon) call qv_check(qv_line_precon(p, ls, counts), 'line.mod:precon')
      call qv_check(qv_line_premove(p, ls, counts), 'line.mod:premove')
      ! precon and premove only relink and clear: the device keeps its
      ! positions and zeroes the cleared slots'
      call qv_check(qv_upload_topology(ctx, pcount, c_null_ptr, infront, behind), &
                    'qv_upload_topology')
    end if
  end subroutine

  !> setup_mesh (initial.f90:541-574) with its checks and messages, and the
  !> warning of initial.f90:285-286
  subroutine setup_mesh()
    integer :: i, j, k, q
    real(c_double) :: mesh_delta
    if (mesh_size <= 0) then
      write(*,*) 'velocity fields not being stored on a mesh - (no spectra etc.)'
      return
    end if
    if (mod(mesh_size, 2) /= 0) call stop_with('init.mod:setup_mesh', &
      'mesh size must be a multiple of 2')
    if (mesh_size < 16) call warning('setup_mesh', 'warning mesh size is small')
    if (mesh_size > 1290) call stop_with('setup_mesh', &
      'mesh_size**3 points must fit a 32-bit count')
    if (normal_velocity /= 'zero') call stop_with('setup_mesh', &
      'the driver has no normal fluid: normal_velocity must be zero')
    mesh_delta = box_size / mesh_size
    write(*,'(a,i3.2,a,f7.6)') ' creating an n^3 mesh, n=', mesh_size, ' resolution=', mesh_delta
    write(*,'(a,i5.2,a)') ' mesh information will be printed every ', mesh_shots, ' time-steps'
    if (mod(mesh_print_delay, mesh_shots) == 0) then
      if (mesh_print_delay > 0) write(*,'(a,i5.2,a)') ' will not start printing until itime= ', &
        mesh_print_delay
    else
      call stop_with('init.mod:setup_mesh', 'mesh_print_delay must be a multiply of mesh_shots')
    end if
    if (mesh_shots < shots) call warning('init.mod', &
      'mesh shots < shots which will create output anomolies')
    allocate(mesh_x(3, mesh_size**3), mesh_u(3, mesh_size**3))
    ! mesh(k,j,i) at (x_i, y_j, z_k), k fastest as print_mesh writes the arrays
    q = 0
    do i = 1, mesh_size
      do j = 1, mesh_size
        do k = 1, mesh_size
          q = q + 1
          mesh_x(1, q) = mesh_delta * real(2*i-1, c_double) / 2._c_double - (box_size / 2._c_double)
          mesh_x(2, q) = mesh_delta * real(2*j-1, c_double) / 2._c_double - (box_size / 2._c_double)
          mesh_x(3, q) = mesh_delta * real(2*k-1, c_double) / 2._c_double - (box_size / 2._c_double)
        end do
      end do
    end do
    mesh_u = 0._c_double
  end subroutine

  subroutine warning(location, message)
    character(len=*), intent(in) :: location, message
    write (*,*) '-------------------------WARNING----------------------------'
    write (*,*) trim(location), ": ", trim(message)
    write (*,*) '------------------------------------------------------------'
  end subroutine

  !> mesh_velocity (timestep.f90:374-452): the superfluid velocity at the
  !> mesh points from the step's positions, ghosts and tree
  subroutine mesh_velocity()
    integer(c_int32_t) :: flags
    flags = 0
    if (velocity == 'Tree' .and. mesh_reference_scale) flags = QV_PROBE_REFERENCE_SCALE
    call qv_check(qv_probe_velocity(ctx, int(mesh_size**3, c_int32_t), mesh_x, mesh_u, &
                                    c_null_ptr, flags), 'qv_probe_velocity')
  end subroutine

  !> data/mesh%03d.dat as print_mesh writes it (output.f90:311-321); u_norm = 0
  subroutine print_mesh(filenumber)
    integer, intent(in) :: filenumber
    character(len=40) :: print_file
    integer :: fh, i, m
    m = mesh_size
    write(print_file, '(a,i3.3,a)') './data/mesh', filenumber, '.dat'
    open(newunit=fh, file=print_file, form='unformatted', status='replace', access='stream')
    write(fh) t
    ! mesh(m/2, m/2, 1:m)%x(1)
    write(fh) (mesh_x(1, m/2 + m*(m/2 - 1) + m*m*(i - 1)), i = 1, m)
    write(fh) (0._c_double, i = 1, 3 * m**3)
    write(fh) mesh_u(1, :)
    write(fh) mesh_u(2, :)
    write(fh) mesh_u(3, :)
    close(fh)
  end subroutine

  !> the ts.log line of print_info (output.f90:55-77): recon and removed are
  !> the line operations' counters; maxdu and curv, not computed here, are 0
  subroutine print_info(filenumber)
    integer, intent(in) :: filenumber
    integer :: fh, nact, j, k
    real(c_double) :: length, maxu, d(3)
    nact = count(infront(1:pcount) > 0)
    length = 0._c_double
    maxu = 0._c_double
    do j = 1, pcount
      if (infront(j) == 0) cycle
      d = x(:, infront(j)) - x(:, j)
      if (boundary /= 'open') then
        do k = 1, 3
          if (boundary == 'openx' .and. k == 1) cycle
          if (boundary == 'openxy' .and. k < 3) cycle
          if (d(k) > box_size / 2) d(k) = d(k) - box_size
          if (d(k) < -box_size / 2) d(k) = d(k) + box_size
        end do
      end if
      length = length + sqrt(sum(d**2))
      maxu = max(maxu, sqrt(sum(u(:, j)**2)))
    end do
    open(newunit=fh, file='data/ts.log', position='append')
    if (filenumber == 1) write(fh, *) '%--var--------t-------pcount-------recon----avg_d-----length&
        &--------maxu---------maxdu-----num eval----curv------removed'
    write(fh, '(i6.4,f13.7,i10.1,i13.1,f7.4,f13.6,f13.5,f13.5,f10.2,f10.2,i13.1)') &
      filenumber, t, nact, int(counts%recon_count), length / nact / delta, length, maxu, &
      0._c_double, real(evals, c_double) / nact, 0._c_double, int(counts%remove_count)
    close(fh)
  end subroutine

  !> data/var%04d.log as printf writes it (output.f90:116-143); u_mf, curv,
  !> stretch and f_mf are not computed here and written as 0
  subroutine printf(filenumber)
    integer, intent(in) :: filenumber
    character(len=40) :: print_file
    integer :: fh
    real(c_double), allocatable :: z(:)
    allocate(z(pcount))
    z = 0._c_double
    write(print_file, '(a,i4.4,a)') './data/var', filenumber, '.log'
    open(newunit=fh, file=print_file, status='replace', form='unformatted', access='stream')
    write(fh) t
    write(fh) pcount
    write(fh) x(1, 1:pcount)
    write(fh) x(2, 1:pcount)
    write(fh) x(3, 1:pcount)
    write(fh) infront(1:pcount)
    write(fh) u(1, 1:pcount)
    write(fh) u(2, 1:pcount)
    write(fh) u(3, 1:pcount)
    write(fh) z, z, z, z, z, z, z, z
    close(fh)
  end subroutine
end program

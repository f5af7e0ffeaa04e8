!> Test-infrastructure driver around the reference Qvort hot path.
!>
!> This program is NOT product code.  It is linked (by oracle/Makefile) against
!> objects compiled from the reference sources under /root/reference/src and is
!> used only (a) to write golden vectors under tests/golden/ and (b) as the
!> "reference" CPU baseline timed by bench.py.  It calls the reference's own
!> routines unchanged:
!>   ghostp          periodic.f90:8-17
!>   construct_tree  tree.f90:44-80
!>   pmotion         timestep.f90:19-93 (calc_velocity timestep.f90:180-370)
!>   pmotion_RK3     timestep.f90:95-176  (run.in runge_kutta T, as run.f90:50-54)
!>   enforce_periodic* periodic.f90:188-303
!>   empty_tree      tree.f90:189-203
!>   pinsert         line.f90:18-108      (flags bit 3)
!>   pclose_tree     tree.f90:219-278     (flags bit 2, tree_theta > 0)
!>   pclose          line.f90:205-234     (flags bit 2, tree_theta = 0)
!>   biot_savart_general(_shift) timestep.f90:516-563, tree_walk_general
!>                   tree.f90:425-488     (flags bit 4, probe points)
!>   the restart file data/var.dat: written after the run with the statements
!>                   of data_dump (output.f90:272-285, flags bit 6), read before
!>                   it with those of data_restore (initial.f90:500-529, flags
!>                   bit 7; the run then continues at the restored itime + 1
!>                   and t).  output.f90 and initial.f90 do not build here
!>                   (itime clash, stiff_solver), so the statements are
!>                   restated; the record layout is that of the reference's
!>                   own type(qvort) under this compiler.
!>   construct_tree's leaf print data/tree_mesh%03d.log (tree.f90:116-123)
!>                   when run.in sets tree_print T and shots divides itime
!>   mesh_velocity   timestep.f90:374-452, when run.in sets mesh_size > 0: the
!>                   grid of setup_mesh (initial.f90:541-574, restated: the
!>                   module does not build here), mesh_velocity after the
!>                   motion when mesh_shots divides itime and itime >
!>                   mesh_print_delay (run.f90:56-58), and data/mesh%03d.dat
!>                   with print_mesh's writes (output.f90:305-321, restated)
!>                   numbered as mesh_shot_print_calls does (output.f90:32-41)
!>                   when shots and mesh_shots divide itime (run.f90:111-124)
!>   the line operations of run.f90:60-72 (flags bit 8): pinsert
!>                   (line.f90:18-108), and when recon_shots divides itime
!>                   pclose_tree / pclose, precon (line.f90:238-252, run.in
!>                   recon_type, default precon_original reconnection.f90:145-256)
!>                   and premove (line.f90:114-161, loop_killer
!>                   reconnection.f90:421-461), between pmotion and the final
!>                   ghostp + wrap of every step
!> and mirrors the order of the main loop in run.f90:36-169 with the line
!> operations left out, except in the neighbour-search mode (flags bit 2):
!> one step of ghostp, construct_tree, pmotion, [pinsert], then the search,
!> as run.f90:41-66 orders them (the tree is the pre-motion one).
!>
!> Files (stream access, little endian), in the working directory:
!>   run.in     read by the reference's read_run_file (cdata.f90:420-807)
!>   state.bin  int32 version, pcount, nsteps, flags, nsample
!>              real64 x(3,pcount); int32 infront(pcount), behind(pcount)
!>              real64 u1(3,pcount), u2(3,pcount), u3(3,pcount)
!>   out.bin    int32 pcount; int64 eval_counter (first step)
!>              real64 u, u_s_LI, u_s_BS, ghosti, ghostb   (3,pcount each, first step)
!>              real64 x, u1, u2, u3                        (3,pcount each, after nsteps)
!>              (flags bit 9) int32 infront, behind (pcount each, after nsteps);
!>              int32 boundary_loop_remove_count; real64 boundary_loop_remove_length
!>   tree.bin   (flags bit 0) every node of the first step's tree in DFS order
!>   timing.txt (flags bit 1) build/walk wall times of a target sample
!>   closest.bin (flags bit 2) int32 pcount after pinsert;
!>              real64 x(3,pcount); int32 infront, behind, closest (before the
!>              search); real64 closestd (before); int32 closest (after);
!>              real64 closestd (after)
!>   line.bin   (flags bit 8) the state after pmotion, pinsert, the search,
!>              precon and premove of the first step (each: int32 pcount;
!>              real64 x, u, u1, u2, ghosti, ghostb, ghostii, ghostbb
!>              (3,pcount each); int32 infront, behind, closest; real64
!>              closestd), the counters after the first step, then after
!>              nsteps: int32 pcount; real64 x, u1, u2 (3,pcount);
!>              int32 infront, behind; counters.  Counters: int32
!>              recon_count, remove_count, small_loop_remove_count; real64
!>              total_length, avg_sep, small_loop_remove_length.  out.bin is
!>              not written in this mode.
!>   probes.bin (flags bit 4, read) int32 nprobe; real64 y(3,nprobe)
!>   probe.bin  (flags bit 4) int32 pcount; real64 x(3,pcount), ghosti(3,pcount)
!>              as the probe sums saw them; real64 u(3,nprobe).  The sums
!>              follow mesh_velocity (timestep.f90:374-452): ghostp,
!>              construct_tree, [pmotion if flags bit 5], then per probe the
!>              central shell and the image shells.
program qv_ref_driver
  use cdata
  use general
  use periodic
  use tree
  use timestep
  use line
  use omp_lib
  implicit none
  integer :: version, nsteps_run, flags, nsample
  integer :: i, k, stride, nact, nwalk
  integer(8) :: eval_first
  real, allocatable :: u_first(:,:), uli_first(:,:), ubs_first(:,:), gi_first(:,:), gb_first(:,:)
  real :: uvel(3)
  real(8) :: t0, t1, t2
  integer :: nodes_written, dummy_itime

  call read_run_file()
  call set_boundary_flags()
  if (mesh_size > 0) call setup_mesh_grid()

  open(unit=41, file='state.bin', access='stream', form='unformatted', status='old', action='read')
  read(41) version, pcount, nsteps_run, flags, nsample
  allocate(f(pcount))
  do i = 1, pcount
    f(i)%x = 0.; f(i)%u = 0.; f(i)%u1 = 0.; f(i)%u2 = 0.; f(i)%u3 = 0.
    f(i)%ghosti = 0.; f(i)%ghostb = 0.; f(i)%ghostii = 0.; f(i)%ghostbb = 0.
    f(i)%u_s_LI = 0.; f(i)%u_s_BS = 0.; f(i)%u_sup = 0.; f(i)%u_mf = 0.
    f(i)%closest = 0; f(i)%closestd = 0.
    f(i)%x_interp = 0.; f(i)%acc = 0.; f(i)%f_mf = 0.; f(i)%stretch = 0.; f(i)%curv = 0.
    f(i)%closestd_loop = 0.; f(i)%delta = 0.; f(i)%t_recon = 0.
  end do
  do i = 1, pcount
    read(41) f(i)%x
  end do
  do i = 1, pcount
    read(41) f(i)%infront
  end do
  do i = 1, pcount
    read(41) f(i)%behind
  end do
  do i = 1, pcount
    read(41) f(i)%u1
  end do
  do i = 1, pcount
    read(41) f(i)%u2
  end do
  do i = 1, pcount
    read(41) f(i)%u3
  end do
  close(41)

  nstart = 1
  t = 0.
  if (btest(flags, 7)) then
    ! data_restore's reads (initial.f90:506-520); no quasi particles or
    ! particles are written by this driver
    deallocate(f)
    open(unit=63, file='./data/var.dat', form='unformatted', status='old', action='read')
    read(63) pcount
    read(63) recon_count
    read(63) dummy_itime
    read(63) t
    allocate(f(pcount))
    read(63) f
    read(63) quasi_pcount
    read(63) part_count
    close(63)
    nstart = dummy_itime + 1
  end if

  if (btest(flags, 1)) then
    call timing_run()
    stop
  end if
  if (btest(flags, 2)) then
    call closest_run()
    stop
  end if
  if (btest(flags, 4)) then
    call probe_run()
    stop
  end if

  allocate(u_first(3,pcount), uli_first(3,pcount), ubs_first(3,pcount))
  allocate(gi_first(3,pcount), gb_first(3,pcount))
  eval_first = 0
  do itime = nstart, nstart + nsteps_run - 1
    call ghostp
    if (tree_theta > 0) call construct_tree
    if (itime == nstart .and. btest(flags, 0) .and. tree_theta > 0) then
      open(unit=42, file='tree.bin', access='stream', form='unformatted', status='replace')
      nodes_written = 0
      call dump_node(vtree, 0)
      close(42)
    end if
    if (runge_kutta) then
      call pmotion_RK3
    else
      call pmotion
    end if
    if (mesh_size > 0) then
      if ((mod(itime,mesh_shots)==0).and.(itime>mesh_print_delay)) call mesh_velocity
    end if
    if (itime == nstart) then
      do i = 1, pcount
        u_first(:,i) = f(i)%u
        uli_first(:,i) = f(i)%u_s_LI
        ubs_first(:,i) = f(i)%u_s_BS
        gi_first(:,i) = f(i)%ghosti
        gb_first(:,i) = f(i)%ghostb
      end do
      if (tree_theta > 0) eval_first = eval_counter
    end if
    if (btest(flags, 8)) then
      if (itime == nstart) then
        open(unit=48, file='line.bin', access='stream', form='unformatted', status='replace')
        call dump_line_state()
      end if
      call pinsert
      if (itime == nstart) call dump_line_state()
      if (mod(itime, recon_shots) == 0) then
        if (tree_theta > 0) then
          call pclose_tree
        else
          call pclose
        end if
        if (itime == nstart) call dump_line_state()
        call precon
        if (itime == nstart) call dump_line_state()
        call premove
        if (itime == nstart) call dump_line_state()
      end if
      if (itime == nstart) call dump_line_counters()
    end if
    call ghostp
    select case(boundary)
      case('periodic')
        call enforce_periodic
      case('openx')
        call enforce_periodic_yz
      case('openxy')
        call enforce_periodic_z
    end select
    t = t + dt
    if (mesh_size > 0 .and. mod(itime, shots) == 0 .and. mod(itime, mesh_shots) == 0) then
      if (itime > mesh_print_delay) call write_mesh((itime-mesh_print_delay)/mesh_shots)
    end if
    if (tree_theta > 0) then
      call empty_tree(vtree)
      deallocate(vtree%parray) ; deallocate(vtree)
      nullify(vtree)
    end if
  end do
  itime = nstart + nsteps_run - 1

  if (btest(flags, 6)) then
    ! data_dump's writes (output.f90:274-283)
    quasi_pcount = 0
    part_count = 0
    open(unit=53, file='./data/var.dat', form='unformatted', status='replace')
    write(53) pcount
    write(53) recon_count
    write(53) itime
    write(53) t
    write(53) f
    write(53) quasi_pcount
    write(53) part_count
    close(53)
  end if

  if (btest(flags, 8)) then
    write(48) pcount
    do i = 1, pcount
      write(48) f(i)%x
    end do
    do i = 1, pcount
      write(48) f(i)%u1
    end do
    do i = 1, pcount
      write(48) f(i)%u2
    end do
    write(48) (f(i)%infront, i = 1, pcount)
    write(48) (f(i)%behind, i = 1, pcount)
    call dump_line_counters()
    close(48)
    stop
  end if

  open(unit=43, file='out.bin', access='stream', form='unformatted', status='replace')
  write(43) pcount
  write(43) eval_first
  write(43) u_first, uli_first, ubs_first, gi_first, gb_first
  do i = 1, pcount
    write(43) f(i)%x
  end do
  do i = 1, pcount
    write(43) f(i)%u1
  end do
  do i = 1, pcount
    write(43) f(i)%u2
  end do
  do i = 1, pcount
    write(43) f(i)%u3
  end do
  if (btest(flags, 9)) then
    ! the links after the run and the counters of boundary_loop_remove
    ! (periodic.f90:306-327), for runs whose loops leave through an open face
    write(43) (f(i)%infront, i = 1, pcount)
    write(43) (f(i)%behind, i = 1, pcount)
    write(43) boundary_loop_remove_count
    write(43) boundary_loop_remove_length
  end if
  close(43)

contains

  !> boundary -> periodic flags, as init_setup does (initial.f90:104-136)
  subroutine set_boundary_flags()
    select case(boundary)
      case('periodic')
        periodic_bc = .true.
      case('openx')
        periodic_bc_notx = .true.
      case('openxy')
        periodic_bc_notxy = .true.
    end select
  end subroutine

  !> setup_mesh's grid (initial.f90:550-571): mesh(k,j,i)%x = (x_i, y_j, z_k)
  subroutine setup_mesh_grid()
    integer :: ii, jj, kk
    real :: xx, yy, zz
    mesh_delta=real(box_size)/mesh_size
    allocate(mesh(mesh_size,mesh_size,mesh_size))
    do kk=1, mesh_size
      do jj=1, mesh_size
        do ii=1, mesh_size
          xx=mesh_delta*real(2*ii-1)/2.-(box_size/2.)
          yy=mesh_delta*real(2*jj-1)/2.-(box_size/2.)
          zz=mesh_delta*real(2*kk-1)/2.-(box_size/2.)
          mesh(kk,jj,ii)%x(1)=xx ; mesh(kk,jj,ii)%x(2)=yy ; mesh(kk,jj,ii)%x(3)=zz
          mesh(kk,jj,ii)%u_sup=0. ; mesh(kk,jj,ii)%u_norm=0.
        end do
      end do
    end do
  end subroutine

  !> print_mesh's writes (output.f90:311-321)
  subroutine write_mesh(filenumber)
    integer, intent(in) :: filenumber
    character(len=40) :: print_file
    write(unit=print_file,fmt="(a,i3.3,a)")"./data/mesh",filenumber,".dat"
    open(unit=92,file=print_file,form='unformatted',status='replace',access='stream')
      write(92) t
      write(92) mesh(mesh_size/2,mesh_size/2,1:mesh_size)%x(1)
      write(92) mesh(1:mesh_size,1:mesh_size,1:mesh_size)%u_norm(1)
      write(92) mesh(1:mesh_size,1:mesh_size,1:mesh_size)%u_norm(2)
      write(92) mesh(1:mesh_size,1:mesh_size,1:mesh_size)%u_norm(3)
      write(92) mesh(1:mesh_size,1:mesh_size,1:mesh_size)%u_sup(1)
      write(92) mesh(1:mesh_size,1:mesh_size,1:mesh_size)%u_sup(2)
      write(92) mesh(1:mesh_size,1:mesh_size,1:mesh_size)%u_sup(3)
    close(92)
  end subroutine

  !> the f(:) fields the line operations read or write, to line.bin
  subroutine dump_line_state()
    integer :: q
    write(48) pcount
    do q = 1, pcount
      write(48) f(q)%x
    end do
    do q = 1, pcount
      write(48) f(q)%u
    end do
    do q = 1, pcount
      write(48) f(q)%u1
    end do
    do q = 1, pcount
      write(48) f(q)%u2
    end do
    do q = 1, pcount
      write(48) f(q)%ghosti
    end do
    do q = 1, pcount
      write(48) f(q)%ghostb
    end do
    do q = 1, pcount
      write(48) f(q)%ghostii
    end do
    do q = 1, pcount
      write(48) f(q)%ghostbb
    end do
    write(48) (f(q)%infront, q = 1, pcount)
    write(48) (f(q)%behind, q = 1, pcount)
    write(48) (f(q)%closest, q = 1, pcount)
    write(48) (f(q)%closestd, q = 1, pcount)
  end subroutine

  subroutine dump_line_counters()
    write(48) recon_count, remove_count, small_loop_remove_count
    write(48) total_length, avg_sep, small_loop_remove_length
  end subroutine

  !> every node (empty children included) in the DFS order of tree_walk
  recursive subroutine dump_node(nd, level)
    type(node), pointer :: nd
    integer, intent(in) :: level
    integer :: leafj
    leafj = 0
    if (nd%pcount == 1) then
      if (nd%parray(1)%infront > 0) leafj = f(nd%parray(1)%infront)%behind
    end if
    write(42) level, nd%pcount, leafj
    write(42) nd%posx, nd%posy, nd%posz, nd%width
    write(42) nd%centx, nd%centy, nd%centz
    write(42) nd%circ
    write(42) nd%Cij
    nodes_written = nodes_written + 1
    if (nd%pcount > 1) then
      call dump_node(nd%fbl, level+1) ; call dump_node(nd%fbr, level+1)
      call dump_node(nd%ftl, level+1) ; call dump_node(nd%ftr, level+1)
      call dump_node(nd%bbl, level+1) ; call dump_node(nd%bbr, level+1)
      call dump_node(nd%btl, level+1) ; call dump_node(nd%btr, level+1)
    end if
  end subroutine

  !> one step up to the neighbour search, in the order of run.f90:41-66
  subroutine closest_run()
    itime = 1
    t = 0.
    call ghostp
    if (tree_theta > 0) call construct_tree
    call pmotion
    if (btest(flags, 3)) call pinsert
    open(unit=45, file='closest.bin', access='stream', form='unformatted', status='replace')
    write(45) pcount
    do k = 1, pcount
      write(45) f(k)%x
    end do
    write(45) (f(k)%infront, k = 1, pcount)
    write(45) (f(k)%behind, k = 1, pcount)
    write(45) (f(k)%closest, k = 1, pcount)
    write(45) (f(k)%closestd, k = 1, pcount)
    if (tree_theta > 0) then
      call pclose_tree
    else
      call pclose
    end if
    write(45) (f(k)%closest, k = 1, pcount)
    write(45) (f(k)%closestd, k = 1, pcount)
    close(45)
  end subroutine

  !> velocity at probe points in the way mesh_velocity sums it
  subroutine probe_run()
    integer :: np, q, peri, perj, perk
    real, allocatable :: y(:,:), up(:,:)
    open(unit=46, file='probes.bin', access='stream', form='unformatted', status='old', &
         action='read')
    read(46) np
    allocate(y(3,np), up(3,np))
    read(46) y
    close(46)
    itime = 1
    t = 0.
    call ghostp
    if (tree_theta > 0) call construct_tree
    if (btest(flags, 5)) call pmotion
    do q = 1, np
      up(:,q) = 0.
      select case(velocity)
        case('BS')
          call biot_savart_general(y(:,q), up(:,q))
          if (periodic_bc) then
            do peri=-1,1 ; do perj=-1,1 ; do perk=-1,1
              if (peri==0.and.perj==0.and.perk==0) cycle
              call biot_savart_general_shift(y(:,q), up(:,q), &
                   (/peri*box_size,perj*box_size,perk*box_size/))
            end do ; end do ; end do
          else if (periodic_bc_notx) then
            do perj=-1,1 ; do perk=-1,1
              if (perj==0.and.perk==0) cycle
              call biot_savart_general_shift(y(:,q), up(:,q), (/0.,perj*box_size,perk*box_size/))
            end do ; end do
          else if (periodic_bc_notxy) then
            do perk=-1,1
              if (perk==0) cycle
              call biot_savart_general_shift(y(:,q), up(:,q), (/0.,0.,perk*box_size/))
            end do
          end if
        case('Tree')
          call tree_walk_general(y(:,q), vtree, (/0.,0.,0./), up(:,q))
          if (periodic_bc) then
            do peri=-1,1 ; do perj=-1,1 ; do perk=-1,1
              if (peri==0.and.perj==0.and.perk==0) cycle
              call tree_walk_general(y(:,q), vtree, (/peri*box_size,perj*box_size,perk*box_size/), &
                   up(:,q))
            end do ; end do ; end do
          else if (periodic_bc_notx) then
            do perj=-1,1 ; do perk=-1,1
              if (perj==0.and.perk==0) cycle
              call tree_walk_general(y(:,q), vtree, (/0.,perj*box_size,perk*box_size/), up(:,q))
            end do ; end do
          else if (periodic_bc_notxy) then
            do perk=-1,1
              if (perk==0) cycle
              call tree_walk_general(y(:,q), vtree, (/0.,0.,perk*box_size/), up(:,q))
            end do
          end if
      end select
    end do
    open(unit=47, file='probe.bin', access='stream', form='unformatted', status='replace')
    write(47) pcount
    do k = 1, pcount
      write(47) f(k)%x
    end do
    do k = 1, pcount
      write(47) f(k)%ghosti
    end do
    write(47) up
    close(47)
  end subroutine

  !> time construct_tree and calc_velocity over a strided sample of targets
  subroutine timing_run()
    integer :: ii
    itime = 1
    t = 0.
    nact = count(mask=f(:)%infront > 0)
    if (nsample <= 0 .or. nsample > nact) nsample = nact
    stride = max(1, nact / nsample)
    call ghostp
    t0 = omp_get_wtime()
    if (tree_theta > 0) call construct_tree
    t1 = omp_get_wtime()
    nwalk = 0
    !$omp parallel do private(ii, uvel) reduction(+:nwalk) schedule(dynamic, 16)
    do ii = 1, pcount, stride
      if (f(ii)%infront == 0) cycle
      call calc_velocity(uvel, ii)
      f(ii)%u = uvel
      nwalk = nwalk + 1
    end do
    !$omp end parallel do
    t2 = omp_get_wtime()
    open(unit=44, file='timing.txt', status='replace')
    write(44, '(a,i0)') 'n_active ', nact
    write(44, '(a,i0)') 'n_sample ', nwalk
    write(44, '(a,i0)') 'threads ', omp_get_max_threads()
    write(44, '(a,es24.16)') 'build_s ', t1 - t0
    write(44, '(a,es24.16)') 'walk_s ', t2 - t1
    close(44)
    k = 0
  end subroutine
end program

!> iso_c_binding interface to libqvort_gpu.so (include/qvort_gpu.h).
!>
!> A Fortran driver keeps Qvort's f(:) arrays on the host in the layout
!> x(3,pcount), infront(pcount), behind(pcount) (cdata.f90:20-41) and hands
!> them to the device with qv_upload; qv_velocity then replaces
!>   ghostp (periodic.f90:8-17) + construct_tree (tree.f90:44-80) + the
!>   velocity loop of pmotion (timestep.f90:41-47),
!> qv_ab_step the Adams-Bashforth ladder of pmotion (timestep.f90:48-71),
!> qv_enforce_periodic the boundary step of run.f90:82-90 (enforce_periodic,
!> or enforce_periodic_yz / _z with their loop removal at open faces,
!> periodic.f90:188-327),
!> qv_closest the neighbour search pclose_tree / pclose (tree.f90:219-278,
!> line.f90:205-234) that precon reads, qv_line_pinsert / qv_line_precon /
!> qv_line_premove the host line operations of run.f90:60-72 (pinsert,
!> precon_original, premove with loop_killer), and qv_destroy the empty_tree /
!> deallocate(vtree) of run.f90:132-138.  qv_vjp / qv_jvp (no counterpart in
!> the reference) give the adjoint and tangent of the direct-sum velocity.
!> Default reals of a -fdefault-real-8 build are real(c_double); default
!> integers are integer(c_int32_t).
module qv_gpu
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: QV_VEL_BS = 1, QV_VEL_TREE = 2
  integer(c_int32_t), parameter, public :: QV_BC_OPEN = 0, QV_BC_PERIODIC = 1, &
                                           QV_BC_OPENX = 2, QV_BC_OPENXY = 3
  integer(c_int32_t), parameter, public :: QV_FP64 = 0, QV_FP32C = 1
  integer(c_int32_t), parameter, public :: QV_PROBE_GHOSTP = 1, QV_PROBE_REFERENCE_SCALE = 2, &
                                           QV_PROBE_EMPTY_CORE = 4
  integer(c_int32_t), parameter, public :: QV_FLAG_REMOVE_MEAN_POS = 1
  integer(c_int), parameter, public :: QV_OK = 0, QV_ERR_ARG = 1, QV_ERR_HIP = 2, &
                                       QV_ERR_SINGULAR = 3, QV_ERR_TREE_DEPTH = 4, &
                                       QV_ERR_STATE = 5, QV_ERR_NO_DEVICE = 6, &
                                       QV_ERR_LINE_EMPTY = 7, QV_ERR_CAPACITY = 8, &
                                       QV_ERR_SCRATCH = 9

  !> run.in parameters of the hot path (cdata.f90:196-416)
  type, bind(C), public :: qv_params
    real(c_double) :: quant_circ, corea, box_size, delta, tree_theta, dt
    integer(c_int32_t) :: velocity, boundary, tree_extra_correction, deriv_order
    integer(c_int32_t) :: precision, flags
  end type

  type, bind(C), public :: qv_timings
    real(c_float) :: ghost_lia_ms, build_ms, walk_ms, total_ms
    integer(c_int64_t) :: nodes, evals, visits, leaf_evals, cell_evals, wave_steps
  end type

  !> host arrays of the line operations: c_loc of x(3,capacity) ... , the
  !> first pcount slots being f(:)
  type, bind(C), public :: qv_line_state
    integer(c_int32_t) :: pcount, capacity
    type(c_ptr) :: x, u, u1, u2, u3, ghosti, ghostb, ghostii, ghostbb
    type(c_ptr) :: infront, behind, closest, closestd
  end type

  !> recon_count, remove_count, small_loop_remove_count, total_length,
  !> avg_sep, small_loop_remove_length of cdata.f90:146-154, 314-315
  type, bind(C), public :: qv_line_counts
    integer(c_int64_t) :: inserted = 0, recon_count = 0, remove_count = 0
    integer(c_int64_t) :: small_loop_remove_count = 0
    real(c_double) :: total_length = 0, avg_sep = 0, small_loop_remove_length = 0
  end type

  public :: qv_create, qv_destroy, qv_set_params, qv_upload, qv_upload_history, &
            qv_velocity, qv_ab_step, qv_rk3_step, qv_enforce_periodic, qv_step, qv_download, &
            qv_download_links, qv_boundary_removed, qv_upload_topology, &
            qv_tree_dump, qv_closest, qv_probe_velocity, qv_vjp, qv_jvp, qv_get_timings, &
            qv_download_ghosts, qv_line_pinsert, qv_line_precon, qv_line_premove, &
            qv_host_alloc, qv_host_free, &
            qv_status_string, qv_check

  interface
    integer(c_int) function qv_create(params, ctx) bind(C, name='qv_create')
      import :: c_int, c_ptr, qv_params
      type(qv_params), intent(in) :: params
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function qv_destroy(ctx) bind(C, name='qv_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function qv_set_params(ctx, params) bind(C, name='qv_set_params')
      import :: c_int, c_ptr, qv_params
      type(c_ptr), value :: ctx
      type(qv_params), intent(in) :: params
    end function
    integer(c_int) function qv_upload(ctx, pcount, x, infront, behind) bind(C, name='qv_upload')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: pcount
      real(c_double), intent(in) :: x(3, *)
      integer(c_int32_t), intent(in) :: infront(*), behind(*)
    end function
    integer(c_int) function qv_upload_history(ctx, u1, u2, u3) bind(C, name='qv_upload_history')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u1, u2, u3
    end function
    integer(c_int) function qv_velocity(ctx, u, u_s_li, u_s_bs, evals) bind(C, name='qv_velocity')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u, u_s_li, u_s_bs, evals
    end function
    integer(c_int) function qv_ab_step(ctx) bind(C, name='qv_ab_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function qv_rk3_step(ctx) bind(C, name='qv_rk3_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function qv_enforce_periodic(ctx) bind(C, name='qv_enforce_periodic')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function qv_step(ctx, nsteps) bind(C, name='qv_step')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function qv_download(ctx, x, u, u1, u2, u3) bind(C, name='qv_download')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, x, u, u1, u2, u3
    end function
    !> positions and links after the host line operations, keeping the
    !> device's velocities and AB history (new points and cleared slots get
    !> pinsert's and clear_particle's values on the device)
    integer(c_int) function qv_upload_topology(ctx, pcount, x, infront, behind) &
        bind(C, name='qv_upload_topology')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: pcount
      type(c_ptr), value :: x
      integer(c_int32_t), intent(in) :: infront(*), behind(*)
    end function
    !> the device links; they change on the device only when the boundary
    !> step removes loops at an open face
    integer(c_int) function qv_download_links(ctx, infront, behind) &
        bind(C, name='qv_download_links')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: infront(*), behind(*)
    end function
    !> boundary_loop_remove_count / _length of the context's boundary steps
    integer(c_int) function qv_boundary_removed(ctx, count, length) &
        bind(C, name='qv_boundary_removed')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: count
      real(c_double), intent(out) :: length
    end function
    integer(c_int) function qv_tree_dump(ctx, out, max_nodes, nnodes) bind(C, name='qv_tree_dump')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, out
      integer(c_int64_t), value :: max_nodes
      integer(c_int64_t), intent(out) :: nnodes
    end function
    integer(c_int) function qv_closest(ctx, closest, closestd) bind(C, name='qv_closest')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(inout) :: closest(*)
      real(c_double), intent(inout) :: closestd(*)
    end function
    !> u (and, if grad is not c_null_ptr, du_k/dy_m in grad(3,3,n) with
    !> grad(m,k,q) = du_k/dy_m) at probe points y(3,n): mesh_velocity's sums,
    !> and the gradient velocity_quasip takes by central differences
    integer(c_int) function qv_probe_velocity(ctx, nprobe, y, u, grad, flags) &
        bind(C, name='qv_probe_velocity')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nprobe
      real(c_double), intent(in) :: y(3, *)
      real(c_double), intent(out) :: u(3, *)
      type(c_ptr), value :: grad
      integer(c_int32_t), value :: flags
    end function
    integer(c_int) function qv_vjp(ctx, w, xbar) bind(C, name='qv_vjp')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: w(3, *)
      real(c_double), intent(out) :: xbar(3, *)
    end function
    integer(c_int) function qv_jvp(ctx, v, jv) bind(C, name='qv_jvp')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: v(3, *)
      real(c_double), intent(out) :: jv(3, *)
    end function
    !> pinned host memory for the f(:) arrays (QV_ERR_HIP: use ordinary memory)
    integer(c_int) function qv_host_alloc(bytes, ptr) bind(C, name='qv_host_alloc')
      import :: c_int, c_size_t, c_ptr
      integer(c_size_t), value :: bytes
      type(c_ptr), intent(out) :: ptr
    end function
    integer(c_int) function qv_host_free(ptr) bind(C, name='qv_host_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: ptr
    end function
    !> ghost zones of the last ghostp, each x(3,pcount)-shaped (c_null_ptr skips)
    integer(c_int) function qv_download_ghosts(ctx, gi, gb, gii, gbb) &
        bind(C, name='qv_download_ghosts')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, gi, gb, gii, gbb
    end function
    integer(c_int) function qv_line_pinsert(params, s, c) bind(C, name='qv_line_pinsert')
      import :: c_int, qv_params, qv_line_state, qv_line_counts
      type(qv_params), intent(in) :: params
      type(qv_line_state), intent(inout) :: s
      type(qv_line_counts), intent(inout) :: c
    end function
    integer(c_int) function qv_line_precon(params, s, c) bind(C, name='qv_line_precon')
      import :: c_int, qv_params, qv_line_state, qv_line_counts
      type(qv_params), intent(in) :: params
      type(qv_line_state), intent(inout) :: s
      type(qv_line_counts), intent(inout) :: c
    end function
    integer(c_int) function qv_line_premove(params, s, c) bind(C, name='qv_line_premove')
      import :: c_int, qv_params, qv_line_state, qv_line_counts
      type(qv_params), intent(in) :: params
      type(qv_line_state), intent(inout) :: s
      type(qv_line_counts), intent(inout) :: c
    end function
    integer(c_int) function qv_get_timings(ctx, t) bind(C, name='qv_get_timings')
      import :: c_int, c_ptr, qv_timings
      type(c_ptr), value :: ctx
      type(qv_timings), intent(out) :: t
    end function
    type(c_ptr) function qv_status_string_c(status) bind(C, name='qv_status_string')
      import :: c_ptr, c_int
      integer(c_int), value :: status
    end function
  end interface

contains

  !> the library's message for a status code
  function qv_status_string(status) result(msg)
    integer(c_int), intent(in) :: status
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    p = qv_status_string_c(status)
    call c_f_pointer(p, s, [512])
    n = 0
    do while (n < 512)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

  !> stop with the reference's fatal_error layout (cdata.f90:955-970) when a
  !> call did not return QV_OK
  subroutine qv_check(status, location)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: location
    if (status == QV_OK) return
    write (*,*) '-------------------------FATAL ERROR-------------------------'
    write (*,*) trim(location), ": ", qv_status_string(status)
    write (*,*) '-------------------------------------------------------------'
    stop 1
  end subroutine
end module

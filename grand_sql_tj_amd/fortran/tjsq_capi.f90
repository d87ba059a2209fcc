! tjsq_capi.f90 -- ISO_C_BINDING interface to libtjsq_hip.so
! (declarations in include/tjsq_capi.h).  Standalone: it depends on no
! module of the reference code.
module tjsq_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: TJSQ_OK = 0
  integer(c_int), parameter :: TJSQ_FP64 = 0, TJSQ_FP32 = 1, TJSQ_BF16X3 = 2

  type, bind(C) :: tjsq_lattice
    integer(c_int) :: ny, nx_tables
    real(c_double) :: jt(2), jd(2), jn(2), esite
    integer(c_int) :: up_bias
  end type tjsq_lattice

  type, bind(C) :: tjsq_opset
    integer(c_int) :: length, nops
    type(c_ptr) :: hdr, sub, off, data
    integer(c_int64_t) :: ndata
    integer(c_int) :: data_on_device
  end type tjsq_opset

  type, bind(C) :: tjsq_basis
    integer(c_int) :: length, nsec
    type(c_ptr) :: sec, sub
  end type tjsq_basis

  type, bind(C) :: tjsq_lanczos_opts
    real(c_double) :: tol
    integer(c_int) :: max_iter, precision
    real(c_double) :: refine_tol
    integer(c_int) :: max_refine
    real(c_double) :: energy_tol, rtol
  end type tjsq_lanczos_opts

  type, bind(C) :: tjsq_lanczos_result
    real(c_double) :: e0, e_ritz, residual
    integer(c_int) :: iterations, applies, refinements, breakdown, refine_iterations
    real(c_double) :: gap, energy_err_est, seconds
  end type tjsq_lanczos_result

  type, bind(C) :: tjsq_trunc_info
    integer(c_int) :: num_state, dim
    real(c_double) :: trunc_err, entropy, entropy1, entropy2
  end type tjsq_trunc_info

  interface
    integer(c_int) function tjsq_density_create(side, len, nw, wsec, levels, prog) bind(C, name="tjsq_density_create")
      import :: c_int, c_ptr
      integer(c_int), value :: side, len, nw, levels
      integer(c_int), intent(in) :: wsec(*)
      type(c_ptr), intent(out) :: prog
    end function

    integer(c_int) function tjsq_dia_transport_create(ops, site, basis, trun, prog) &
        bind(C, name="tjsq_dia_transport_create")
      import :: c_int, c_ptr, tjsq_opset, tjsq_basis
      type(tjsq_opset), intent(in) :: ops, site, trun
      type(tjsq_basis), intent(in) :: basis
      type(c_ptr), intent(out) :: prog
    end function

    integer(c_int) function tjsq_measure_create(side, ops, site, basis, nw, wsec, prog) &
        bind(C, name="tjsq_measure_create")
      import :: c_int, c_ptr, tjsq_opset, tjsq_basis
      integer(c_int), value :: side, nw
      type(tjsq_opset), intent(in) :: ops, site
      type(tjsq_basis), intent(in) :: basis
      integer(c_int), intent(in) :: wsec(*)
      type(c_ptr), intent(out) :: prog
    end function

    integer(c_int) function tjsq_block_update_create(lat, side, sys_len, env_len, pre, basis, trun, prog) &
        bind(C, name="tjsq_block_update_create")
      import :: c_int, c_ptr, tjsq_lattice, tjsq_opset, tjsq_basis
      type(tjsq_lattice), intent(in) :: lat
      integer(c_int), value :: side, sys_len, env_len
      type(tjsq_opset), intent(in) :: pre, trun
      type(tjsq_basis), intent(in) :: basis
      type(c_ptr), intent(out) :: prog
    end function

    integer(c_int) function tjsq_wave_transform_create(dir, sys_bs, env_bs, sys_trun, env_trun, nw_old, wsec_old, &
        nw_new, wsec_new, prog) bind(C, name="tjsq_wave_transform_create")
      import :: c_int, c_ptr, tjsq_opset, tjsq_basis
      integer(c_int), value :: dir, nw_old, nw_new
      type(tjsq_basis), intent(in) :: sys_bs, env_bs
      type(tjsq_opset), intent(in) :: sys_trun, env_trun
      integer(c_int), intent(in) :: wsec_old(*), wsec_new(*)
      type(c_ptr), intent(out) :: prog
    end function

    integer(c_int) function tjsq_program_sizes(prog, sizes) bind(C, name="tjsq_program_sizes")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: prog
      integer(c_int64_t), intent(out) :: sizes(14)
    end function

    integer(c_int) function tjsq_program_layout(prog, hdr, sub, off) bind(C, name="tjsq_program_layout")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: prog
      integer(c_int), intent(out) :: hdr(*), sub(*)
      integer(c_int64_t), intent(out) :: off(*)
    end function

    integer(c_int) function tjsq_program_run_host(prog, ops, x, y) bind(C, name="tjsq_program_run_host")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: prog
      real(c_double), intent(in) :: ops(*), x(*)
      real(c_double), intent(out) :: y(*)
    end function

    integer(c_int) function tjsq_program_destroy(prog) bind(C, name="tjsq_program_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: prog
    end function

    integer(c_int) function tjsq_truncation_operator_host(nsec, den_sub, den_off, den, kept_min, kept_max, trun_err, &
        trun_sub, trun_off, trun_num, spectrum, info) bind(C, name="tjsq_truncation_operator_host")
      import :: c_int, c_ptr, c_double, c_int64_t, tjsq_trunc_info
      integer(c_int), value :: nsec, kept_min, kept_max
      integer(c_int), intent(in) :: den_sub(*)
      integer(c_int64_t), intent(in) :: den_off(*)
      real(c_double), intent(inout) :: den(*)
      real(c_double), value :: trun_err
      integer(c_int), intent(out) :: trun_sub(*)
      integer(c_int64_t), intent(out) :: trun_off(*)
      integer(c_int), intent(out) :: trun_num
      type(c_ptr), value :: spectrum
      type(tjsq_trunc_info), intent(out) :: info
    end function

    integer(c_int) function tjsq_plan_create(lat, sys_len, env_len, sys, env, sys_bs, env_bs, nw, wsec, &
        precision_mask, stream, plan) bind(C, name="tjsq_plan_create")
      import :: c_int, c_ptr, tjsq_lattice, tjsq_opset, tjsq_basis
      type(tjsq_lattice), intent(in) :: lat
      integer(c_int), value :: sys_len, env_len, nw, precision_mask
      type(tjsq_opset), intent(in) :: sys, env
      type(tjsq_basis), intent(in) :: sys_bs, env_bs
      integer(c_int), intent(in) :: wsec(*)
      type(c_ptr), value :: stream
      type(c_ptr), intent(out) :: plan
    end function

    integer(c_int) function tjsq_plan_set_ops(plan, sys, env, stream) bind(C, name="tjsq_plan_set_ops")
      import :: c_int, c_ptr, tjsq_opset
      type(c_ptr), value :: plan
      type(tjsq_opset), intent(in) :: sys, env
      type(c_ptr), value :: stream
    end function

    integer(c_int) function tjsq_plan_destroy(plan) bind(C, name="tjsq_plan_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
    end function

    integer(c_int) function tjsq_heff_apply_host(plan, x, y) bind(C, name="tjsq_heff_apply_host")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
    end function

    integer(c_int) function tjsq_lanczos_solve_host(plan, psi, opts, res) bind(C, name="tjsq_lanczos_solve_host")
      import :: c_int, c_ptr, c_double, tjsq_lanczos_opts, tjsq_lanczos_result
      type(c_ptr), value :: plan
      real(c_double), intent(inout) :: psi(*)
      type(tjsq_lanczos_opts), intent(in) :: opts
      type(tjsq_lanczos_result), intent(out) :: res
    end function

    type(c_ptr) function tjsq_last_error_c() bind(C, name="tjsq_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! message of the last failed call, as a Fortran string
  function tjsq_last_error() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    call c_f_pointer(tjsq_last_error_c(), p, [4096])
    n = 0
    do while (n < 4096)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function tjsq_last_error
end module tjsq_capi

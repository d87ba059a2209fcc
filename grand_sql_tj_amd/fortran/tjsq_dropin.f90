! tjsq_dropin.f90 -- drop-in GPU replacements for the Lanczos step of the
! reference tjsq code, with the reference's own routine signatures:
!
!   gpu_lanczos_memo_new(sys,env,sys_bs,env_bs,super,eigvec,levels,num_idx)
!        replaces lanczos_memo_new, code/tjsq/eigen.f90:2-10
!   gpu_model_wave(sys,env,sys_bs,env_bs,invec,outvec)
!        replaces model_wave (with hmerge), code/tjsq/get_sys.f90:4127-4134
!
! It is compiled inside the reference's build (it uses module pubdata,
! code/tjsq/common.f90) and calls libtjsq_hip.so through module tjsq_capi.
! The derived types are flattened into the C-ABI layout; the lattice
! couplings are read from pubdata (uniform per bond class, as
! code/tjsq/main.f90:35-47 sets them).  A nonzero status from the library
! ends the program with `error stop`, where the reference would `stop`.
module tjsq_dropin
  use iso_c_binding
  use pubdata
  use tjsq_capi
  implicit none
  private
  public :: gpu_lanczos_memo_new, gpu_model_wave, tjsq_dropin_precision

  ! contraction precision of the drop-in solve (TJSQ_FP64 by default)
  integer(c_int), save :: tjsq_dropin_precision = TJSQ_FP64

  type flat_model
    integer(c_int), allocatable :: hdr(:, :), sub(:, :)
    integer(c_int64_t), allocatable :: off(:)
    real(c_double), allocatable :: data(:)
  end type flat_model

  type flat_basis
    integer(c_int), allocatable :: sec(:, :), sub(:, :)
  end type flat_basis

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= TJSQ_OK) then
      write(*, *) 'tjsq: ', what, ' failed: ', tjsq_last_error()
      error stop 1
    end if
  end subroutine check

  integer function count_subs(b)
    type(Total_Block), intent(in) :: b
    count_subs = 0
    if (allocated(b%sub)) count_subs = b%num
  end function count_subs

  ! Total_Model -> flat operator set: ham, then (elec_up, elec_down, spin_sd,
  ! num_sn) of each allocated active site
  subroutine flatten_model(m, f, nops)
    type(Total_Model), intent(in), target :: m
    type(flat_model), intent(out) :: f
    integer, intent(out) :: nops
    integer :: nact, i, q, nsub, ndata, k, s, r, c
    type(Total_Block), pointer :: b
    nact = 0
    do i = 1, nxx
      if (.not. allocated(m%sub(i)%elec_up%sub)) exit
      nact = i
    end do
    nops = 1 + 4 * nact
    nsub = count_subs(m%ham)
    ndata = 0
    do k = 1, nops
      b => op_of(m, k)
      do s = 1, count_subs(b)
        ndata = ndata + b%sub(s)%row_dim * b%sub(s)%sdim
      end do
      if (k > 1) nsub = nsub + count_subs(b)
    end do
    allocate(f%hdr(4, nops), f%sub(5, max(1, nsub)), f%off(max(1, nsub)), f%data(max(1, ndata)))
    q = 0
    ndata = 0
    do k = 1, nops
      b => op_of(m, k)
      f%hdr(:, k) = [m%len, b%up_dif, b%down_dif, count_subs(b)]
      do s = 1, count_subs(b)
        q = q + 1
        f%sub(:, q) = [b%sub(s)%num_up, b%sub(s)%num_down, b%sub(s)%down_dif, b%sub(s)%row_dim, b%sub(s)%sdim]
        f%off(q) = ndata
        do c = 1, b%sub(s)%sdim
          do r = 1, b%sub(s)%row_dim
            f%data(ndata + (c - 1) * b%sub(s)%row_dim + r) = b%sub(s)%mat(r, c)
          end do
        end do
        ndata = ndata + b%sub(s)%row_dim * b%sub(s)%sdim
      end do
    end do
  end subroutine flatten_model

  function op_of(m, k) result(b)
    type(Total_Model), intent(in), target :: m
    integer, intent(in) :: k
    type(Total_Block), pointer :: b
    integer :: site
    if (k == 1) then
      b => m%ham
      return
    end if
    site = (k - 2) / 4 + 1
    select case (mod(k - 2, 4))
    case (0)
      b => m%sub(site)%elec_up
    case (1)
      b => m%sub(site)%elec_down
    case (2)
      b => m%sub(site)%spin_sd
    case default
      b => m%sub(site)%num_sn
    end select
  end function op_of

  subroutine flatten_basis(bs, f)
    type(Total_Basis), intent(in) :: bs
    type(flat_basis), intent(out) :: f
    integer :: i, j, q, n
    n = 0
    do i = 1, bs%num
      n = n + bs%sub(i)%num
    end do
    allocate(f%sec(4, max(1, bs%num)), f%sub(6, max(1, n)))
    q = 0
    do i = 1, bs%num
      f%sec(:, i) = [bs%sub(i)%new_num_up, bs%sub(i)%new_num_down, bs%sub(i)%dim, bs%sub(i)%num]
      do j = 1, bs%sub(i)%num
        q = q + 1
        f%sub(:, q) = [bs%sub(i)%sub(j)%bl_num_up, bs%sub(i)%sub(j)%bl_num_down, bs%sub(i)%sub(j)%st_num_up, &
                       bs%sub(i)%sub(j)%st_num_down, bs%sub(i)%sub(j)%spos, bs%sub(i)%sub(j)%sdim]
      end do
    end do
  end subroutine flatten_basis

  ! couplings of pubdata (main.f90:35-47): bond class 1 = directions 1-4,
  ! class 2 = directions 5-8
  function lattice_from_pubdata() result(lat)
    type(tjsq_lattice) :: lat
    lat%ny = Ny
    lat%nx_tables = Nx
    lat%jt = [jt(1, 1), jt(5, 1)]
    lat%jd = [jd(1, 1), jd(5, 1)]
    lat%jn = [jn(1), jn(5)]
    lat%esite = esite(1)
    ! the particle label the site operators carry (dmrg_sub.f90:380-381):
    ! the apply kernels see up_bias only through the operators' up_dif
    lat%up_bias = st_elec_up%up_dif
  end function lattice_from_pubdata

  subroutine make_plan(sys, env, sys_bs, env_bs, wsec, nw, plan, fs, fe, bs, be)
    type(Total_Model), intent(in) :: sys, env
    type(Total_Basis), intent(in) :: sys_bs, env_bs
    integer(c_int), intent(in), target :: wsec(:, :)
    integer, intent(in) :: nw
    type(c_ptr), intent(out) :: plan
    type(flat_model), intent(inout), target :: fs, fe
    type(flat_basis), intent(inout), target :: bs, be
    type(tjsq_lattice) :: lat
    type(tjsq_opset) :: os, oe
    type(tjsq_basis) :: cbs, cbe
    integer :: ns, ne
    call flatten_model(sys, fs, ns)
    call flatten_model(env, fe, ne)
    call flatten_basis(sys_bs, bs)
    call flatten_basis(env_bs, be)
    lat = lattice_from_pubdata()
    os = tjsq_opset(sys%len, ns, c_loc(fs%hdr), c_loc(fs%sub), c_loc(fs%off), c_loc(fs%data), &
                    int(size(fs%data), c_int64_t), 0)
    oe = tjsq_opset(env%len, ne, c_loc(fe%hdr), c_loc(fe%sub), c_loc(fe%off), c_loc(fe%data), &
                    int(size(fe%data), c_int64_t), 0)
    cbs = tjsq_basis(sys_bs%len, sys_bs%num, c_loc(bs%sec), c_loc(bs%sub))
    cbe = tjsq_basis(env_bs%len, env_bs%num, c_loc(be%sec), c_loc(be%sub))
    call check(tjsq_plan_create(lat, sys%len, env%len, os, oe, cbs, cbe, int(nw, c_int), wsec, &
                                int(7, c_int), c_null_ptr, plan), 'tjsq_plan_create')
  end subroutine make_plan

  subroutine pack_wave(w, x)
    type(Wavefunction), intent(in) :: w
    real(c_double), intent(out) :: x(:)
    integer :: i, n, o
    o = 0
    do i = 1, w%num
      n = w%sub(i)%sys_dim * w%sub(i)%env_dim
      x(o + 1:o + n) = reshape(w%sub(i)%vec, [n])
      o = o + n
    end do
  end subroutine pack_wave

  subroutine unpack_wave(x, w)
    real(c_double), intent(in) :: x(:)
    type(Wavefunction), intent(inout) :: w
    integer :: i, n, o
    o = 0
    do i = 1, w%num
      n = w%sub(i)%sys_dim * w%sub(i)%env_dim
      w%sub(i)%vec = reshape(x(o + 1:o + n), [w%sub(i)%sys_dim, w%sub(i)%env_dim])
      o = o + n
    end do
  end subroutine unpack_wave

  subroutine wave_sectors(w, wsec)
    type(Wavefunction), intent(in) :: w
    integer(c_int), allocatable, intent(out) :: wsec(:, :)
    integer :: i
    allocate(wsec(6, w%num))
    do i = 1, w%num
      wsec(:, i) = [w%sub(i)%sys_num_up, w%sub(i)%sys_num_down, w%sub(i)%env_num_up, w%sub(i)%env_num_down, &
                    w%sub(i)%sys_dim, w%sub(i)%env_dim]
    end do
  end subroutine wave_sectors

  ! lanczos_memo_new (eigen.f90:2-175) for levels = 1: start vector in,
  ! normalised ground state out, E0 into Energys(1)
  subroutine gpu_lanczos_memo_new(sys, env, sys_bs, env_bs, super, eigvec, levels, num_idx)
    type(Total_Model), intent(in) :: sys, env
    type(Total_Basis), intent(in) :: sys_bs, env_bs
    type(Super_Basis), intent(in) :: super
    integer, intent(in) :: levels, num_idx
    type(Wavefunction), intent(inout) :: eigvec(levels)
    type(flat_model) :: fs, fe
    type(flat_basis) :: bs, be
    integer(c_int), allocatable, target :: wsec(:, :)
    real(c_double), allocatable :: psi(:)
    type(c_ptr) :: plan
    type(tjsq_lanczos_opts) :: opts
    type(tjsq_lanczos_result) :: res
    if (levels /= 1) then
      write(*, *) 'tjsq: gpu_lanczos_memo_new solves the ground state only (levels = 1)'
      error stop 1
    end if
    if (num_idx < 0 .or. super%num /= eigvec(1)%num) then
      write(*, *) 'tjsq: wavefunction and super basis disagree'
      error stop 1
    end if
    call wave_sectors(eigvec(1), wsec)
    call make_plan(sys, env, sys_bs, env_bs, wsec, eigvec(1)%num, plan, fs, fe, bs, be)
    allocate(psi(eigvec(1)%dim))
    call pack_wave(eigvec(1), psi)
    opts%tol = lanzero
    opts%max_iter = 110
    opts%precision = tjsq_dropin_precision
    opts%refine_tol = 1.0d-6
    opts%max_refine = 1
    opts%energy_tol = 0.0d0
    opts%rtol = 0.0d0
    call check(tjsq_lanczos_solve_host(plan, psi, opts, res), 'tjsq_lanczos_solve_host')
    call unpack_wave(psi, eigvec(1))
    Energys(1) = res%e0
    call check(tjsq_plan_destroy(plan), 'tjsq_plan_destroy')
  end subroutine gpu_lanczos_memo_new

  ! hmerge + model_wave (get_sys.f90:4383, 4127): outvec = H_eff invec
  subroutine gpu_model_wave(sys, env, sys_bs, env_bs, invec, outvec)
    type(Total_Model), intent(in) :: sys, env
    type(Total_Basis), intent(in) :: sys_bs, env_bs
    type(Wavefunction), intent(in) :: invec
    type(Wavefunction), intent(inout) :: outvec
    type(flat_model) :: fs, fe
    type(flat_basis) :: bs, be
    integer(c_int), allocatable, target :: wsec(:, :)
    real(c_double), allocatable :: x(:), y(:)
    type(c_ptr) :: plan
    call wave_sectors(invec, wsec)
    call make_plan(sys, env, sys_bs, env_bs, wsec, invec%num, plan, fs, fe, bs, be)
    allocate(x(invec%dim), y(invec%dim))
    call pack_wave(invec, x)
    call check(tjsq_heff_apply_host(plan, x, y), 'tjsq_heff_apply_host')
    call unpack_wave(y, outvec)
    call check(tjsq_plan_destroy(plan), 'tjsq_plan_destroy')
  end subroutine gpu_model_wave
end module tjsq_dropin

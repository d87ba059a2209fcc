! tjsq_dropin.f90 -- drop-in GPU replacements for the Lanczos step of the
! reference tjsq code, with the reference's own routine signatures:
!
!   gpu_lanczos_memo_new(sys,env,sys_bs,env_bs,super,eigvec,levels,num_idx)
!        replaces lanczos_memo_new, code/tjsq/eigen.f90:2-10
!   gpu_model_wave(sys,env,sys_bs,env_bs,invec,outvec)
!        replaces model_wave (with hmerge), code/tjsq/get_sys.f90:4127-4134
!
! and the step after a solve (code/tjsq/dmrg_sub.f90:656-693, 778, 960):
!
!   gpu_Get_density_matrix_sys_new(sys_den,wave,levels) / _env_new
!        replace Get_density_matrix_sys_new / _env_new, den_mat.f90:84-134
!   gpu_Get_truncation_operator(den_mat,trun_op,trun_err,num_state)
!        replaces Get_truncation_operator, den_mat.f90:141-383 (kept_min,
!        kept_max from pubdata; sets Entropys; appends the reference's log
!        lines)
!   gpu_Get_Hamiltonian_Sys(pre,next,basis,trun) / _Env
!        replace Get_Hamiltonian_Sys / _Env, get_sys.f90:4-478: pre becomes
!        the projected model one site longer, next is left untouched (the
!        reference uses it as scratch).  Operators spin_sz, which nothing on
!        the solve path reads, are not carried.
!   gpu_wave_trans_LR / gpu_wave_trans_RL(sys_bs,env_bs,sys_trun,env_trun,
!        old_wave,new_wave) replace wave_trans_LR / _RL, dmrg_sub.f90:
!        1093-1384; new_wave comes allocated with the new sectors.
!
! The block-update routines work at the bond of the last gpu_lanczos_memo_new
! or gpu_model_wave call (the reference calls dmrg_setup(sys%len, env%len)
! and lanczos_memo_new before each bond's update), or the one gpu_set_bond
! sets.  Data stay on the host between calls, as in the reference: each call
! copies its inputs to the device and its outputs back.
!
! Plan cache: gpu_lanczos_memo_new and gpu_model_wave keep the last device
! plan, keyed by the block lengths, the couplings and the layouts of the two
! block models, their fused bases and the wavefunction's sectors.  A call
! with the same key uploads the models' data into that plan
! (tjsq_plan_set_ops: no host compile); any other key compiles a new plan
! and frees the old one.  A reference build that swaps only model_wave, which
! lanczos_memo_new calls once per iteration, so compiles once per bond.
! gpu_release_plan frees the cached plan; tjsq_dropin_plan_builds counts the
! compiles.
!
! Run logs as the reference writes them: gpu_lanczos_memo_new appends its
! Energy_All.dat line (and prints it, eigen.f90:166-171; the time is the
! solve's wall time), gpu_Get_truncation_operator appends to Entropy.dat,
! entropy1.dat and Num_State.dat (den_mat.f90:299-317).
!
! It is compiled inside the reference's build (it uses module pubdata,
! code/tjsq/common.f90) and calls libtjsq_hip.so through module tjsq_capi.
! The derived types are flattened into the C-ABI layout; the lattice
! couplings are read from pubdata, where they must be uniform per bond class
! (directions 1-4 and 5-8, directions 9-12 zero, esite the same on every
! site), as code/tjsq/main.f90:35-47 sets them: other values end the program
! with a message (pind / pinz / pint, which no routine on the path reads,
! are not checked).  A nonzero status from the library ends the program with
! `error stop`, where the reference would `stop`.
module tjsq_dropin
  use iso_c_binding
  use pubdata
  use tjsq_capi
  implicit none
  private
  public :: gpu_lanczos_memo_new, gpu_model_wave, tjsq_dropin_precision, gpu_set_bond
  public :: gpu_Get_density_matrix_sys_new, gpu_Get_density_matrix_env_new, gpu_Get_truncation_operator
  public :: gpu_Get_Hamiltonian_Sys, gpu_Get_Hamiltonian_Env, gpu_wave_trans_LR, gpu_wave_trans_RL
  public :: gpu_measure_operator_dia, gpu_release_plan, tjsq_dropin_plan_builds

  ! contraction precision of the drop-in solve (TJSQ_FP64 by default)
  integer(c_int), save :: tjsq_dropin_precision = TJSQ_FP64

  ! plan cache (one plan) and its key
  type(c_ptr), save :: cached_plan = c_null_ptr
  integer, save :: tjsq_dropin_plan_builds = 0
  integer, save :: key_len(2) = -1
  real(c_double), save :: key_coup(7) = 0.0d0
  integer(c_int), allocatable, save :: key_shdr(:, :), key_ssub(:, :), key_ehdr(:, :), key_esub(:, :)
  integer(c_int64_t), allocatable, save :: key_soff(:), key_eoff(:)
  integer(c_int), allocatable, save :: key_bsec(:, :), key_bsub(:, :), key_besec(:, :), key_besub(:, :)
  integer(c_int), allocatable, save :: key_wsec(:, :)

  ! block lengths of the current bond
  integer, save :: bond_sys_len = -1, bond_env_len = -1

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

  ! the couplings the library takes: uniform per bond class on every site
  ! (common.f90:36-39 holds them per direction and site)
  subroutine check_uniform_couplings()
    integer :: i, k
    logical :: ok
    ok = .true.
    do i = 1, Num_site
      do k = 1, neibt
        if (k <= 4) then
          ok = ok .and. jt(k, i) == jt(1, 1) .and. jd(k, i) == jd(1, 1)
        else if (k <= 8) then
          ok = ok .and. jt(k, i) == jt(5, 1) .and. jd(k, i) == jd(5, 1)
        else
          ok = ok .and. jt(k, i) == 0.0d0 .and. jd(k, i) == 0.0d0
        end if
      end do
      ok = ok .and. esite(i) == esite(1)
    end do
    do k = 1, neibt
      if (k <= 4) then
        ok = ok .and. jn(k) == jn(1)
      else if (k <= 8) then
        ok = ok .and. jn(k) == jn(5)
      else
        ok = ok .and. jn(k) == 0.0d0
      end if
    end do
    if (.not. ok) then
      write(*, *) 'tjsq: the couplings jt / jd / jn / esite are not uniform per bond class; ', &
                  'the GPU path takes directions 1-4 and 5-8 with one value each, 9-12 zero, one esite'
      error stop 1
    end if
  end subroutine check_uniform_couplings

  ! couplings of pubdata (main.f90:35-47): bond class 1 = directions 1-4,
  ! class 2 = directions 5-8
  function lattice_from_pubdata() result(lat)
    type(tjsq_lattice) :: lat
    call check_uniform_couplings()
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

  logical function same_i2(a, b)
    integer(c_int), allocatable, intent(in) :: a(:, :)
    integer(c_int), intent(in) :: b(:, :)
    same_i2 = .false.
    if (.not. allocated(a)) return
    if (any(shape(a) /= shape(b))) return
    same_i2 = all(a == b)
  end function same_i2

  logical function same_i8(a, b)
    integer(c_int64_t), allocatable, intent(in) :: a(:)
    integer(c_int64_t), intent(in) :: b(:)
    same_i8 = .false.
    if (.not. allocated(a)) return
    if (size(a) /= size(b)) return
    same_i8 = all(a == b)
  end function same_i8

  ! the device plan of a bond: the cached one when the key matches (with the
  ! models' data uploaded into it), else a new one that replaces it
  subroutine get_plan(sys, env, sys_bs, env_bs, wsec, nw, plan, fs, fe, bs, be)
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
    real(c_double) :: coup(7)
    integer :: ns, ne
    logical :: hit
    call flatten_model(sys, fs, ns)
    call flatten_model(env, fe, ne)
    call flatten_basis(sys_bs, bs)
    call flatten_basis(env_bs, be)
    lat = lattice_from_pubdata()
    coup = [lat%jt(1), lat%jt(2), lat%jd(1), lat%jd(2), lat%jn(1), lat%jn(2), lat%esite]
    os = tjsq_opset(sys%len, ns, c_loc(fs%hdr), c_loc(fs%sub), c_loc(fs%off), c_loc(fs%data), &
                    int(size(fs%data), c_int64_t), 0)
    oe = tjsq_opset(env%len, ne, c_loc(fe%hdr), c_loc(fe%sub), c_loc(fe%off), c_loc(fe%data), &
                    int(size(fe%data), c_int64_t), 0)
    hit = c_associated(cached_plan) .and. key_len(1) == sys%len .and. key_len(2) == env%len
    if (hit) hit = all(key_coup == coup) .and. same_i2(key_shdr, fs%hdr) .and. same_i2(key_ssub, fs%sub)
    if (hit) hit = same_i8(key_soff, fs%off) .and. same_i2(key_ehdr, fe%hdr) .and. same_i2(key_esub, fe%sub)
    if (hit) hit = same_i8(key_eoff, fe%off) .and. same_i2(key_bsec, bs%sec) .and. same_i2(key_bsub, bs%sub)
    if (hit) hit = same_i2(key_besec, be%sec) .and. same_i2(key_besub, be%sub) .and. same_i2(key_wsec, wsec(:, 1:nw))
    if (hit) then
      call check(tjsq_plan_set_ops(cached_plan, os, oe, c_null_ptr), 'tjsq_plan_set_ops')
      plan = cached_plan
      return
    end if
    call gpu_release_plan()
    cbs = tjsq_basis(sys_bs%len, sys_bs%num, c_loc(bs%sec), c_loc(bs%sub))
    cbe = tjsq_basis(env_bs%len, env_bs%num, c_loc(be%sec), c_loc(be%sub))
    call check(tjsq_plan_create(lat, sys%len, env%len, os, oe, cbs, cbe, int(nw, c_int), wsec, &
                                int(7, c_int), c_null_ptr, plan), 'tjsq_plan_create')
    tjsq_dropin_plan_builds = tjsq_dropin_plan_builds + 1
    cached_plan = plan
    key_len = [sys%len, env%len]
    key_coup = coup
    key_shdr = fs%hdr
    key_ssub = fs%sub
    key_soff = fs%off
    key_ehdr = fe%hdr
    key_esub = fe%sub
    key_eoff = fe%off
    key_bsec = bs%sec
    key_bsub = bs%sub
    key_besec = be%sec
    key_besub = be%sub
    key_wsec = wsec(:, 1:nw)
  end subroutine get_plan

  ! frees the cached plan (the next solve or apply compiles a new one)
  subroutine gpu_release_plan()
    if (c_associated(cached_plan)) call check(tjsq_plan_destroy(cached_plan), 'tjsq_plan_destroy')
    cached_plan = c_null_ptr
    key_len = -1
  end subroutine gpu_release_plan

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
    call gpu_set_bond(sys%len, env%len)
    call get_plan(sys, env, sys_bs, env_bs, wsec, eigvec(1)%num, plan, fs, fe, bs, be)
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
    ! the reference's output line (eigen.f90:166-171)
    open(11, file="Energy_All.dat", position='append')
    write(11, 111) "Sys_len=", eigvec(1)%sys_len, "E0=", Energys(1), "Iter=", res%iterations, "time=", res%seconds
    write(*, 111) "Sys_len=", eigvec(1)%sys_len, "E0=", Energys(1), "Iter=", res%iterations, "time=", res%seconds
    close(11)
111 format(A8,I3,2X,A3,2X,F18.12,2X,A5,I4,2X,A5,2X,F12.6)
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
    call gpu_set_bond(sys%len, env%len)
    call get_plan(sys, env, sys_bs, env_bs, wsec, invec%num, plan, fs, fe, bs, be)
    allocate(x(invec%dim), y(invec%dim))
    call pack_wave(invec, x)
    call check(tjsq_heff_apply_host(plan, x, y), 'tjsq_heff_apply_host')
    call unpack_wave(y, outvec)
  end subroutine gpu_model_wave

  subroutine gpu_set_bond(sys_len, env_len)
    integer, intent(in) :: sys_len, env_len
    bond_sys_len = sys_len
    bond_env_len = env_len
  end subroutine gpu_set_bond

  ! one operator as a one-operator flat set
  subroutine flatten_block(b, f)
    type(Total_Block), intent(in) :: b
    type(flat_model), intent(out) :: f
    integer :: s, r, c, n, nd
    n = count_subs(b)
    nd = 0
    do s = 1, n
      nd = nd + b%sub(s)%row_dim * b%sub(s)%sdim
    end do
    allocate(f%hdr(4, 1), f%sub(5, max(1, n)), f%off(max(1, n)), f%data(max(1, nd)))
    f%hdr(:, 1) = [b%len, b%up_dif, b%down_dif, n]
    nd = 0
    do s = 1, n
      f%sub(:, s) = [b%sub(s)%num_up, b%sub(s)%num_down, b%sub(s)%down_dif, b%sub(s)%row_dim, b%sub(s)%sdim]
      f%off(s) = nd
      do c = 1, b%sub(s)%sdim
        do r = 1, b%sub(s)%row_dim
          f%data(nd + (c - 1) * b%sub(s)%row_dim + r) = b%sub(s)%mat(r, c)
        end do
      end do
      nd = nd + b%sub(s)%row_dim * b%sub(s)%sdim
    end do
  end subroutine flatten_block

  function opset_of(f, length, nops) result(o)
    type(flat_model), intent(in), target :: f
    integer, intent(in) :: length, nops
    type(tjsq_opset) :: o
    o = tjsq_opset(length, nops, c_loc(f%hdr), c_loc(f%sub), c_loc(f%off), c_loc(f%data), &
                   int(size(f%data), c_int64_t), 0)
  end function opset_of

  function basis_of(bs, f) result(o)
    type(Total_Basis), intent(in) :: bs
    type(flat_basis), intent(in), target :: f
    type(tjsq_basis) :: o
    o = tjsq_basis(bs%len, bs%num, c_loc(f%sec), c_loc(f%sub))
  end function basis_of

  ! one operator of a flat layout (header hdr, sub-blocks s0 + 1 .. s0 + hdr(4)
  ! of sub / off into data) into a Total_Block of total dimension dim
  subroutine block_from_flat(b, hdr, sub, off, data, s0, dim)
    type(Total_Block), intent(inout) :: b
    integer(c_int), intent(in) :: hdr(4), sub(:, :)
    integer(c_int64_t), intent(in) :: off(:)
    real(c_double), intent(in) :: data(:)
    integer, intent(in) :: s0, dim
    integer :: s, r, c, rows, cols
    integer(c_int64_t) :: o
    if (allocated(b%sub)) deallocate(b%sub)
    b%len = hdr(1)
    b%up_dif = hdr(2)
    b%down_dif = hdr(3)
    b%num = hdr(4)
    b%dim = dim
    allocate(b%sub(b%num))
    do s = 1, b%num
      rows = sub(4, s0 + s)
      cols = sub(5, s0 + s)
      b%sub(s)%num_up = sub(1, s0 + s)
      b%sub(s)%num_down = sub(2, s0 + s)
      b%sub(s)%down_dif = sub(3, s0 + s)
      b%sub(s)%row_dim = rows
      b%sub(s)%sdim = cols
      allocate(b%sub(s)%mat(rows, cols))
      o = off(s0 + s)
      do c = 1, cols
        do r = 1, rows
          b%sub(s)%mat(r, c) = data(o + (c - 1) * rows + r)
        end do
      end do
    end do
  end subroutine block_from_flat

  ! a program's output layout and data on the host
  subroutine program_output(prog, ops, x, hdr, sub, off, y)
    type(c_ptr), intent(in) :: prog
    real(c_double), intent(in) :: ops(:), x(:)
    integer(c_int), allocatable, intent(out) :: hdr(:, :), sub(:, :)
    integer(c_int64_t), allocatable, intent(out) :: off(:)
    real(c_double), allocatable, intent(out) :: y(:)
    integer(c_int64_t) :: sz(14)
    call check(tjsq_program_sizes(prog, sz), 'tjsq_program_sizes')
    allocate(hdr(4, max(1_c_int64_t, sz(1))), sub(5, max(1_c_int64_t, sz(2))), off(max(1_c_int64_t, sz(2))), &
             y(max(1_c_int64_t, sz(3))))
    call check(tjsq_program_layout(prog, hdr, sub, off), 'tjsq_program_layout')
    call check(tjsq_program_run_host(prog, ops, x, y), 'tjsq_program_run_host')
  end subroutine program_output

  ! Get_density_matrix_sys_new / _env_new (den_mat.f90:84-134): the levels
  ! states summed and normalised, rho = psi psi^T per sector (side 0) or
  ! psi^T psi (side 1)
  subroutine den_matrix(side, den, wave, levels)
    integer, intent(in) :: side, levels
    type(Total_Block), intent(inout) :: den
    type(Wavefunction), intent(in) :: wave(levels)
    integer(c_int), allocatable :: wsec(:, :), hdr(:, :), sub(:, :)
    integer(c_int64_t), allocatable :: off(:)
    real(c_double), allocatable :: psi(:), y(:)
    real(c_double) :: none(1)
    type(c_ptr) :: prog
    integer :: l, n, s, dim, length
    call wave_sectors(wave(1), wsec)
    n = wave(1)%dim
    allocate(psi(max(1, n * levels)))
    do l = 1, levels
      call pack_wave(wave(l), psi((l - 1) * n + 1:l * n))
    end do
    length = wave(1)%sys_len
    if (side == 1) length = wave(1)%env_len
    call check(tjsq_density_create(int(side, c_int), int(length, c_int), int(wave(1)%num, c_int), wsec, &
                                   int(levels, c_int), prog), 'tjsq_density_create')
    none = 0.0d0
    call program_output(prog, none, psi, hdr, sub, off, y)
    dim = 0
    do s = 1, hdr(4, 1)
      dim = dim + sub(5, s)
    end do
    call block_from_flat(den, hdr(:, 1), sub, off, y, 0, dim)
    call check(tjsq_program_destroy(prog), 'tjsq_program_destroy')
  end subroutine den_matrix

  subroutine gpu_Get_density_matrix_sys_new(sys_den, wave, levels)
    integer, intent(in) :: levels
    type(Wavefunction), intent(in) :: wave(levels)
    type(Total_Block), intent(inout) :: sys_den
    call den_matrix(0, sys_den, wave, levels)
  end subroutine gpu_Get_density_matrix_sys_new

  subroutine gpu_Get_density_matrix_env_new(env_den, wave, levels)
    integer, intent(in) :: levels
    type(Wavefunction), intent(in) :: wave(levels)
    type(Total_Block), intent(inout) :: env_den
    call den_matrix(1, env_den, wave, levels)
  end subroutine gpu_Get_density_matrix_env_new

  ! Get_truncation_operator (den_mat.f90:141-383): every sector diagonalised
  ! on the device, the kept states chosen as the reference chooses them
  ! (kept_min, kept_max of pubdata, trun_err), their eigenvectors copied out
  subroutine gpu_Get_truncation_operator(den_mat, trun_op, trun_err, num_state)
    real(8), intent(in) :: trun_err
    type(Total_Block), intent(in) :: den_mat
    type(Total_Block), intent(inout) :: trun_op
    integer, intent(inout) :: num_state
    type(flat_model) :: f
    integer(c_int), allocatable :: tsub(:, :)
    integer(c_int64_t), allocatable :: toff(:)
    integer(c_int) :: nt, hdr(4)
    type(tjsq_trunc_info) :: info
    integer :: n, s, dim, slen
    call flatten_block(den_mat, f)
    n = max(1, count_subs(den_mat))
    allocate(tsub(5, n), toff(n))
    call check(tjsq_truncation_operator_host(int(count_subs(den_mat), c_int), f%sub, f%off, f%data, &
                                             int(kept_min, c_int), int(kept_max, c_int), real(trun_err, c_double), &
                                             tsub, toff, nt, c_null_ptr, info), 'tjsq_truncation_operator_host')
    dim = 0
    do s = 1, nt
      dim = dim + tsub(5, s)
    end do
    hdr = [int(den_mat%len, c_int), int(den_mat%up_dif, c_int), int(den_mat%down_dif, c_int), nt]
    call block_from_flat(trun_op, hdr, tsub, toff, f%data, 0, dim)
    num_state = info%num_state
    Entropys = info%entropy
    ! the reference's logs (den_mat.f90:299-317)
    slen = den_mat%len
    if (.not. system_block) slen = Num_site - den_mat%len
    open(60, file="Entropy.dat", position='append')
    write(60, *) "t=", jt(1, 1), jt(7, 1), "J=", jz(1, 1), jz(7, 1), "N=", Num_site
    close(60)
    open(60, file="entropy1.dat", position='append')
    write(60, 22) slen, info%entropy1, info%entropy2, info%entropy, Num_site, tot_num_up, tot_num_down
    close(60)
22  format(i8, 3f18.8, 3i9)
    open(60, file="Num_State.dat", position='append')
    write(60, *) "Sys_len=", slen, "Num_State=", info%num_state, "Err=", info%trunc_err
    close(60)
  end subroutine gpu_Get_truncation_operator

  ! Get_Hamiltonian_Sys (side 0) / _Env (side 1), get_sys.f90:4-478
  subroutine enlarge_block(side, pre, basis, trun)
    integer, intent(in) :: side
    type(Total_Model), intent(inout), target :: pre
    type(Total_Basis), intent(in) :: basis
    type(Total_Block), intent(in) :: trun
    type(flat_model), target :: fp, ft
    type(flat_basis), target :: fb
    type(tjsq_lattice) :: lat
    type(c_ptr) :: prog
    integer(c_int), allocatable :: hdr(:, :), sub(:, :)
    integer(c_int64_t), allocatable :: off(:)
    real(c_double), allocatable :: y(:)
    integer :: nops, k, s0, len0
    len0 = bond_sys_len
    if (side == 1) len0 = bond_env_len
    if (bond_sys_len < 0 .or. pre%len /= len0) then
      write(*, *) 'tjsq: the block does not belong to the current bond (gpu_set_bond)'
      error stop 1
    end if
    call flatten_model(pre, fp, nops)
    call flatten_block(trun, ft)
    call flatten_basis(basis, fb)
    lat = lattice_from_pubdata()
    call check(tjsq_block_update_create(lat, int(side, c_int), int(bond_sys_len, c_int), int(bond_env_len, c_int), &
                                        opset_of(fp, pre%len, nops), basis_of(basis, fb), opset_of(ft, trun%len, 1), &
                                        prog), 'tjsq_block_update_create')
    call program_output(prog, ft%data, fp%data, hdr, sub, off, y)
    call check(tjsq_program_destroy(prog), 'tjsq_program_destroy')
    ! the new model replaces pre (ham, then elec_up, elec_down, spin_sd,
    ! num_sn of each active site; the other sites' operators are released)
    do k = 1, nxx
      call drop_block(pre%sub(k)%elec_up)
      call drop_block(pre%sub(k)%elec_down)
      call drop_block(pre%sub(k)%spin_sd)
      call drop_block(pre%sub(k)%spin_sz)
      call drop_block(pre%sub(k)%num_sn)
    end do
    if ((size(hdr, 2) - 1) / 4 > nxx) then
      write(*, *) 'tjsq: the new block has more active sites than nxx'
      error stop 1
    end if
    s0 = 0
    do k = 1, size(hdr, 2)
      call set_op(pre, k, hdr(:, k), sub, off, y, s0, trun_dim(trun))
      s0 = s0 + hdr(4, k)
    end do
    pre%len = pre%len + 1
  end subroutine enlarge_block

  ! operator k of a model (the order of flatten_model) from a flat layout
  subroutine set_op(m, k, hdr, sub, off, data, s0, dim)
    type(Total_Model), intent(inout) :: m
    integer, intent(in) :: k, s0, dim
    integer(c_int), intent(in) :: hdr(4), sub(:, :)
    integer(c_int64_t), intent(in) :: off(:)
    real(c_double), intent(in) :: data(:)
    integer :: site
    if (k == 1) then
      call block_from_flat(m%ham, hdr, sub, off, data, s0, dim)
      return
    end if
    site = (k - 2) / 4 + 1
    select case (mod(k - 2, 4))
    case (0)
      call block_from_flat(m%sub(site)%elec_up, hdr, sub, off, data, s0, dim)
    case (1)
      call block_from_flat(m%sub(site)%elec_down, hdr, sub, off, data, s0, dim)
    case (2)
      call block_from_flat(m%sub(site)%spin_sd, hdr, sub, off, data, s0, dim)
    case default
      call block_from_flat(m%sub(site)%num_sn, hdr, sub, off, data, s0, dim)
    end select
  end subroutine set_op

  integer function trun_dim(t)
    type(Total_Block), intent(in) :: t
    integer :: s
    trun_dim = 0
    do s = 1, count_subs(t)
      trun_dim = trun_dim + t%sub(s)%sdim
    end do
  end function trun_dim

  subroutine drop_block(b)
    type(Total_Block), intent(inout) :: b
    if (allocated(b%sub)) deallocate(b%sub)
    b%num = 0
    b%dim = 0
  end subroutine drop_block

  subroutine gpu_Get_Hamiltonian_Sys(pre, next, basis, trun)
    type(Total_Basis), intent(in) :: basis
    type(Total_Model), intent(inout), target :: pre
    type(Total_Model), intent(inout) :: next
    type(Total_Block), intent(in) :: trun
    call enlarge_block(0, pre, basis, trun)
  end subroutine gpu_Get_Hamiltonian_Sys

  subroutine gpu_Get_Hamiltonian_Env(pre, next, basis, trun)
    type(Total_Basis), intent(in) :: basis
    type(Total_Model), intent(inout), target :: pre
    type(Total_Model), intent(inout) :: next
    type(Total_Block), intent(in) :: trun
    call enlarge_block(1, pre, basis, trun)
  end subroutine gpu_Get_Hamiltonian_Env

  ! wave_trans_LR (dir 0) / wave_trans_RL (dir 1), dmrg_sub.f90:1093-1384,
  ! with wave_normalize of the result
  subroutine move_wave(dir, sys_bs, env_bs, sys_trun, env_trun, old_wave, new_wave)
    integer, intent(in) :: dir
    type(Total_Basis), intent(in) :: sys_bs, env_bs
    type(Total_Block), intent(in) :: sys_trun, env_trun
    type(Wavefunction), intent(in) :: old_wave
    type(Wavefunction), intent(inout) :: new_wave
    type(flat_basis), target :: fs, fe
    type(flat_model), target :: ts, te
    integer(c_int), allocatable :: wold(:, :), wnew(:, :), hdr(:, :), sub(:, :)
    integer(c_int64_t), allocatable :: off(:)
    real(c_double), allocatable :: x(:), y(:), ops(:)
    type(c_ptr) :: prog
    call flatten_basis(sys_bs, fs)
    call flatten_basis(env_bs, fe)
    call flatten_block(sys_trun, ts)
    call flatten_block(env_trun, te)
    call wave_sectors(old_wave, wold)
    call wave_sectors(new_wave, wnew)
    call check(tjsq_wave_transform_create(int(dir, c_int), basis_of(sys_bs, fs), basis_of(env_bs, fe), &
                                          opset_of(ts, sys_trun%len, 1), opset_of(te, env_trun%len, 1), &
                                          int(old_wave%num, c_int), wold, int(new_wave%num, c_int), wnew, prog), &
               'tjsq_wave_transform_create')
    allocate(x(max(1, old_wave%dim)), ops(size(ts%data) + size(te%data)))
    call pack_wave(old_wave, x)
    ops(1:size(ts%data)) = ts%data
    ops(size(ts%data) + 1:) = te%data
    call program_output(prog, ops, x, hdr, sub, off, y)
    call check(tjsq_program_destroy(prog), 'tjsq_program_destroy')
    call unpack_wave(y, new_wave)
  end subroutine move_wave

  subroutine gpu_wave_trans_LR(sys_bs, env_bs, sys_trun, env_trun, old_wave, new_wave)
    type(Total_Basis), intent(in) :: sys_bs, env_bs
    type(Total_Block), intent(in) :: sys_trun, env_trun
    type(Wavefunction), intent(in) :: old_wave
    type(Wavefunction), intent(inout) :: new_wave
    call move_wave(0, sys_bs, env_bs, sys_trun, env_trun, old_wave, new_wave)
  end subroutine gpu_wave_trans_LR

  subroutine gpu_wave_trans_RL(sys_bs, env_bs, sys_trun, env_trun, old_wave, new_wave)
    type(Total_Basis), intent(in) :: sys_bs, env_bs
    type(Total_Block), intent(in) :: sys_trun, env_trun
    type(Wavefunction), intent(in) :: old_wave
    type(Wavefunction), intent(inout) :: new_wave
    call move_wave(1, sys_bs, env_bs, sys_trun, env_trun, old_wave, new_wave)
  end subroutine gpu_wave_trans_RL
  ! an identity truncation over the sectors of a fused basis (the reference
  ! skips update_trun_dia at lengths up to trun_idx)
  subroutine identity_flat(bs, f)
    type(Total_Basis), intent(in) :: bs
    type(flat_model), intent(out) :: f
    integer :: k, r, n, nd
    nd = 0
    do k = 1, bs%num
      nd = nd + bs%sub(k)%dim * bs%sub(k)%dim
    end do
    allocate(f%hdr(4, 1), f%sub(5, max(1, bs%num)), f%off(max(1, bs%num)), f%data(max(1, nd)))
    f%hdr(:, 1) = [bs%len, 0, 0, bs%num]
    f%data = 0.0d0
    nd = 0
    do k = 1, bs%num
      n = bs%sub(k)%dim
      f%sub(:, k) = [bs%sub(k)%new_num_up, bs%sub(k)%new_num_down, 0, n, n]
      f%off(k) = nd
      do r = 1, n
        f%data(nd + (r - 1) * n + r) = 1.0d0
      end do
      nd = nd + n * n
    end do
  end subroutine identity_flat

  ! <st_oper> on the sites of one side of the wavefunction: every site's
  ! operator carried along the block (Get_sys_oper_dia / Get_env_oper_dia,
  ! get_sys.f90:764-891), all of them in one device program per block
  ! length, then contracted with the wavefunction (measure_sys_block /
  ! _site, measure_env_block / _site, measure.f90:1241-1470).  v(i) is site
  ! i counted from the side's outer end; v(length) the bond site.
  subroutine side_measure(side, st, wave, trun_idx, v)
    integer, intent(in) :: side, trun_idx
    type(Total_Block), intent(in) :: st
    type(Wavefunction), intent(in) :: wave
    real(8), allocatable, intent(out) :: v(:)
    type(flat_model), target :: cur, fs, ft
    type(flat_basis), target :: fb
    integer(c_int), allocatable :: hdr(:, :), sub(:, :), wsec(:, :)
    integer(c_int64_t), allocatable :: off(:)
    real(c_double), allocatable :: y(:), psi(:)
    type(c_ptr) :: prog
    integer :: length, L, nops, j, D
    length = wave%sys_len
    if (side == 1) length = wave%env_len
    call flatten_block(st, fs)
    call flatten_block(st, cur)
    nops = 1
    do L = 2, length - 1
      if (side == 0) then
        call flatten_basis(sys_bsm(L), fb)
        if (L > trun_idx) then
          call flatten_block(systruns(L), ft)
        else
          call identity_flat(sys_bsm(L), ft)
        end if
        call check(tjsq_dia_transport_create(opset_of(cur, L - 1, nops), opset_of(fs, 1, 1), &
                                             basis_of(sys_bsm(L), fb), opset_of(ft, L, 1), prog), &
                   'tjsq_dia_transport_create')
      else
        call flatten_basis(env_bsm(L), fb)
        if (L > trun_idx) then
          call flatten_block(envtruns(L), ft)
        else
          call identity_flat(env_bsm(L), ft)
        end if
        call check(tjsq_dia_transport_create(opset_of(cur, L - 1, nops), opset_of(fs, 1, 1), &
                                             basis_of(env_bsm(L), fb), opset_of(ft, L, 1), prog), &
                   'tjsq_dia_transport_create')
      end if
      call program_output(prog, ft%data, cur%data, hdr, sub, off, y)
      call check(tjsq_program_destroy(prog), 'tjsq_program_destroy')
      nops = nops + 1
      call move_alloc(hdr, cur%hdr)
      call move_alloc(sub, cur%sub)
      call move_alloc(off, cur%off)
      call move_alloc(y, cur%data)
    end do
    call wave_sectors(wave, wsec)
    D = wave%dim
    allocate(psi(max(1, D)))
    call pack_wave(wave, psi)
    if (side == 0) then
      call flatten_basis(sys_bsm(length), fb)
      call check(tjsq_measure_create(0_c_int, opset_of(cur, length - 1, nops), opset_of(fs, 1, 1), &
                                     basis_of(sys_bsm(length), fb), int(wave%num, c_int), wsec, prog), &
                 'tjsq_measure_create')
    else
      call flatten_basis(env_bsm(length), fb)
      call check(tjsq_measure_create(1_c_int, opset_of(cur, length - 1, nops), opset_of(fs, 1, 1), &
                                     basis_of(env_bsm(length), fb), int(wave%num, c_int), wsec, prog), &
                 'tjsq_measure_create')
    end if
    call program_output(prog, cur%data, psi, hdr, sub, off, y)
    call check(tjsq_program_destroy(prog), 'tjsq_program_destroy')
    allocate(v(length))
    do j = 1, length
      v(j) = dot_product(y((j - 1) * D + 1:j * D), psi(1:D))
    end do
  end subroutine side_measure

  ! measure_operator_dia (measure.f90:3-57): <st_oper> on every site of the
  ! lattice in the wavefunction, appended to Filename as the reference
  ! writes it.  Only trun_idx = 1 is taken, the value the reference's driver
  ! passes (main.f90:80,250): with a larger trun_idx the reference still
  ! projects with systruns(idx_two) at the last length (get_sys.f90:817-822),
  ! a case this routine does not restate, so it stops instead.
  subroutine gpu_measure_operator_dia(ave_dia, st_oper, wave, trun_idx, Filename)
    character(len=30) :: Filename
    integer, intent(in) :: trun_idx
    type(Total_Block), intent(in) :: st_oper
    type(Wavefunction), intent(in) :: wave
    real(8), intent(inout) :: ave_dia(Num_site)
    real(8), allocatable :: v(:)
    integer :: i
    if (trun_idx /= 1) then
      write(*, *) 'tjsq: gpu_measure_operator_dia takes trun_idx = 1 only, got ', trun_idx
      error stop 1
    end if
    ave_dia = 0.0d0
    call side_measure(0, st_oper, wave, trun_idx, v)
    ave_dia(1:wave%sys_len) = v(1:wave%sys_len)
    call side_measure(1, st_oper, wave, trun_idx, v)
    do i = 1, wave%env_len
      ave_dia(Num_site - i + 1) = v(i)
    end do
    open(10, file=Filename, position='append')
    write(10, *) "Nx=", Nx, "Ny=", Ny, "N_up=", tot_num_up, "N_down=", tot_num_down, "M1=", kept_min, &
      "M2=", kept_max
    do i = 1, Num_site
      write(10, 111) Lattice(1, i), Lattice(2, i), ave_dia(i)
    end do
    write(10, *)
    close(10)
111 format(I4,1X,I2,1X,F21.15)
  end subroutine gpu_measure_operator_dia

end module tjsq_dropin

! ref_bridge.f90 -- C-callable bridge into the reference tjsq Fortran routines.
!
! TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
! reference's own unmodified sources (read where they lie under the reference
! checkout) into oracle/_ref/libtjsq_ref.so by oracle/Makefile.  It lets the
! tests run the reference's hmerge + model_wave (code/tjsq/get_sys.f90:4383,
! :4127) and lanczos_memo_new (code/tjsq/eigen.f90:2) on the same flat inputs
! the oracle and the HIP path receive.  The flat layout is the one described
! in oracle/tjsq_oracle.h.
!
! The reference is compiled with its shipped parameters (common.f90:10-12:
! Nx=24, Ny=6, up_bias=0); the bridge builds the site operators itself so a
! run may use (N, 2S) quantum numbers (up_bias = 1), which the apply kernels
! handle through the operators' up_dif alone.
module ref_bridge_util
  use pubdata
  implicit none
  ! operators returned by the last refb_* call that produces operators
  type(Total_Block), allocatable, save :: res(:)
  ! basis and wavefunction read by the last refb_basis_from_disk / refb_wave_from_disk
  type(Total_Basis), save :: bres
  type(Wavefunction), save :: wres
contains
  subroutine set_result(n)
    integer, intent(in) :: n
    integer :: i
    if (allocated(res)) then
      do i = 1, size(res)
        call free_block(res(i))
      end do
      deallocate(res)
    end if
    allocate(res(n))
  end subroutine set_result

  subroutine put_block(b, len, hdr, sub, off, data)
    type(Total_Block), intent(inout) :: b
    integer, intent(in) :: len, hdr(4), sub(5, *)
    integer(8), intent(in) :: off(*)
    real(8), intent(in) :: data(*)
    integer :: i, r, c
    integer(8) :: o
    b%len = len
    b%up_dif = hdr(2)
    b%down_dif = hdr(3)
    b%num = hdr(4)
    b%dim = 0
    if (allocated(b%sub)) deallocate(b%sub)
    allocate(b%sub(b%num))
    do i = 1, b%num
      b%sub(i)%num_up = sub(1, i)
      b%sub(i)%num_down = sub(2, i)
      b%sub(i)%down_dif = sub(3, i)
      b%sub(i)%row_dim = sub(4, i)
      b%sub(i)%sdim = sub(5, i)
      b%dim = b%dim + sub(5, i)
      allocate(b%sub(i)%mat(sub(4, i), sub(5, i)))
      o = off(i)
      do c = 1, sub(5, i)
        do r = 1, sub(4, i)
          b%sub(i)%mat(r, c) = data(o + (c - 1) * sub(4, i) + r)
        end do
      end do
    end do
  end subroutine put_block

  subroutine put_model(m, len, nops, hdr, sub, off, data)
    type(Total_Model), intent(inout) :: m
    integer, intent(in) :: len, nops, hdr(4, *), sub(5, *)
    integer(8), intent(in) :: off(*)
    real(8), intent(in) :: data(*)
    integer :: op, first, site
    m%len = len
    first = 1
    call put_block(m%ham, len, hdr(:, 1), sub(1, first), off(first), data)
    first = first + hdr(4, 1)
    do op = 2, nops
      site = (op - 2) / 4 + 1
      select case (mod(op - 2, 4))
      case (0)
        call put_block(m%sub(site)%elec_up, len, hdr(:, op), sub(1, first), off(first), data)
      case (1)
        call put_block(m%sub(site)%elec_down, len, hdr(:, op), sub(1, first), off(first), data)
      case (2)
        call put_block(m%sub(site)%spin_sd, len, hdr(:, op), sub(1, first), off(first), data)
      case (3)
        call put_block(m%sub(site)%num_sn, len, hdr(:, op), sub(1, first), off(first), data)
      end select
      first = first + hdr(4, op)
    end do
  end subroutine put_model

  subroutine put_basis(bs, len, nsec, bsec, bsub)
    type(Total_Basis), intent(inout) :: bs
    integer, intent(in) :: len, nsec, bsec(4, *), bsub(6, *)
    integer :: i, j, q
    bs%len = len
    bs%num = nsec
    bs%dim = 0
    allocate(bs%sub(nsec))
    q = 0
    do i = 1, nsec
      bs%sub(i)%new_num_up = bsec(1, i)
      bs%sub(i)%new_num_down = bsec(2, i)
      bs%sub(i)%dim = bsec(3, i)
      bs%sub(i)%num = bsec(4, i)
      bs%sub(i)%idx = bsec(4, i)
      bs%dim = bs%dim + bsec(3, i)
      allocate(bs%sub(i)%sub(bsec(4, i)))
      do j = 1, bsec(4, i)
        q = q + 1
        bs%sub(i)%sub(j)%bl_num_up = bsub(1, q)
        bs%sub(i)%sub(j)%bl_num_down = bsub(2, q)
        bs%sub(i)%sub(j)%st_num_up = bsub(3, q)
        bs%sub(i)%sub(j)%st_num_down = bsub(4, q)
        bs%sub(i)%sub(j)%spos = bsub(5, q)
        bs%sub(i)%sub(j)%sdim = bsub(6, q)
      end do
    end do
  end subroutine put_basis

  subroutine put_wave(w, sys_len, env_len, nw, wsec, x)
    type(Wavefunction), intent(inout) :: w
    integer, intent(in) :: sys_len, env_len, nw, wsec(6, *)
    real(8), intent(in) :: x(*)
    integer :: i, r, c
    integer(8) :: o
    w%sys_len = sys_len
    w%env_len = env_len
    w%num = nw
    w%dim = 0
    allocate(w%sub(nw))
    o = 0
    do i = 1, nw
      w%sub(i)%sys_num_up = wsec(1, i)
      w%sub(i)%sys_num_down = wsec(2, i)
      w%sub(i)%env_num_up = wsec(3, i)
      w%sub(i)%env_num_down = wsec(4, i)
      w%sub(i)%sys_dim = wsec(5, i)
      w%sub(i)%env_dim = wsec(6, i)
      allocate(w%sub(i)%vec(wsec(5, i), wsec(6, i)))
      do c = 1, wsec(6, i)
        do r = 1, wsec(5, i)
          w%sub(i)%vec(r, c) = x(o + (c - 1) * wsec(5, i) + r)
        end do
      end do
      o = o + int(wsec(5, i), 8) * wsec(6, i)
      w%dim = w%dim + wsec(5, i) * wsec(6, i)
    end do
  end subroutine put_wave

  subroutine get_wave(w, y)
    type(Wavefunction), intent(in) :: w
    real(8), intent(inout) :: y(*)
    integer :: i, r, c
    integer(8) :: o
    o = 0
    do i = 1, w%num
      do c = 1, w%sub(i)%env_dim
        do r = 1, w%sub(i)%sys_dim
          y(o + (c - 1) * w%sub(i)%sys_dim + r) = w%sub(i)%vec(r, c)
        end do
      end do
      o = o + int(w%sub(i)%sys_dim, 8) * w%sub(i)%env_dim
    end do
  end subroutine get_wave

  subroutine free_block(b)
    type(Total_Block), intent(inout) :: b
    integer :: i
    if (.not. allocated(b%sub)) return
    do i = 1, size(b%sub)
      if (allocated(b%sub(i)%mat)) deallocate(b%sub(i)%mat)
    end do
    deallocate(b%sub)
  end subroutine free_block

  subroutine free_set(s)
    type(Block_Set), intent(inout) :: s
    call free_block(s%elec_up)
    call free_block(s%elec_down)
    call free_block(s%spin_sd)
    call free_block(s%spin_sz)
    call free_block(s%num_sn)
  end subroutine free_set

  subroutine free_model(m)
    type(Total_Model), intent(inout) :: m
    integer :: i
    call free_block(m%ham)
    do i = 1, nxx
      call free_set(m%sub(i))
      call free_set(m%sub_s(i))
    end do
    do i = 1, 2
      call free_set(m%snc(i))
      call free_set(m%sdc(i))
    end do
  end subroutine free_model

  ! site operators of Get_site_operator_su2_U0 (dmrg_sub.f90:370-473) with the
  ! particle-number label ub taking the place of the compile-time up_bias
  subroutine make_site_ops(ub)
    integer, intent(in) :: ub
    call free_block(st_elec_up)
    call free_block(st_elec_down)
    call free_block(num_elec)
    call free_block(st_sd)
    call free_block(st_si)
    st_elec_up%len = 1; st_elec_up%num = 1; st_elec_up%dim = 1
    st_elec_up%up_dif = ub; st_elec_up%down_dif = 1
    allocate(st_elec_up%sub(1))
    st_elec_up%sub(1)%num_up = 0; st_elec_up%sub(1)%num_down = 0
    st_elec_up%sub(1)%down_dif = 1; st_elec_up%sub(1)%row_dim = 1; st_elec_up%sub(1)%sdim = 1
    allocate(st_elec_up%sub(1)%mat(1, 1)); st_elec_up%sub(1)%mat(1, 1) = -dsqrt(2.0d0)

    st_elec_down%len = 1; st_elec_down%num = 1; st_elec_down%dim = 1
    st_elec_down%up_dif = -ub; st_elec_down%down_dif = 1
    allocate(st_elec_down%sub(1))
    st_elec_down%sub(1)%num_up = ub; st_elec_down%sub(1)%num_down = 1
    st_elec_down%sub(1)%down_dif = -1; st_elec_down%sub(1)%row_dim = 1; st_elec_down%sub(1)%sdim = 1
    allocate(st_elec_down%sub(1)%mat(1, 1)); st_elec_down%sub(1)%mat(1, 1) = -dsqrt(2.0d0)

    num_elec%len = 1; num_elec%num = 2; num_elec%dim = 2; num_elec%up_dif = 0; num_elec%down_dif = 0
    allocate(num_elec%sub(2))
    num_elec%sub(1)%num_up = 0; num_elec%sub(1)%num_down = 0; num_elec%sub(1)%down_dif = 0
    num_elec%sub(1)%row_dim = 1; num_elec%sub(1)%sdim = 1
    allocate(num_elec%sub(1)%mat(1, 1)); num_elec%sub(1)%mat(1, 1) = 0.0d0
    num_elec%sub(2)%num_up = ub; num_elec%sub(2)%num_down = 1; num_elec%sub(2)%down_dif = 0
    num_elec%sub(2)%row_dim = 1; num_elec%sub(2)%sdim = 1
    allocate(num_elec%sub(2)%mat(1, 1)); num_elec%sub(2)%mat(1, 1) = dsqrt(2.0d0)

    st_sd%len = 1; st_sd%num = 1; st_sd%dim = 1; st_sd%up_dif = 0; st_sd%down_dif = su
    allocate(st_sd%sub(1))
    st_sd%sub(1)%num_up = ub; st_sd%sub(1)%num_down = 1; st_sd%sub(1)%down_dif = 0
    st_sd%sub(1)%row_dim = 1; st_sd%sub(1)%sdim = 1
    allocate(st_sd%sub(1)%mat(1, 1)); st_sd%sub(1)%mat(1, 1) = dsqrt(1.50d0)

    st_si%len = 1; st_si%num = 2; st_si%dim = 2; st_si%up_dif = 0; st_si%down_dif = 0
    allocate(st_si%sub(2))
    st_si%sub(1)%num_up = 0; st_si%sub(1)%num_down = 0; st_si%sub(1)%down_dif = 0
    st_si%sub(1)%row_dim = 1; st_si%sub(1)%sdim = 1
    allocate(st_si%sub(1)%mat(1, 1)); st_si%sub(1)%mat(1, 1) = 1.0d0
    st_si%sub(2)%num_up = ub; st_si%sub(2)%num_down = 1; st_si%sub(2)%down_dif = 0
    st_si%sub(2)%row_dim = 1; st_si%sub(2)%sdim = 1
    allocate(st_si%sub(2)%mat(1, 1)); st_si%sub(2)%mat(1, 1) = dsqrt(2.0d0)
  end subroutine make_site_ops
end module ref_bridge_util

! Set couplings (main.f90:35-47 convention, uniform per bond class), lattice
! tables (main.f90:50-52), site operators and the 6j machinery.
integer(c_int) function refb_init(ny_, ub, t_nn, t_nnn, d_nn, d_nnn, n_nn, n_nnn, esite_, lanzero_) &
    bind(C, name="refb_init")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: ny_, ub
  real(c_double), value :: t_nn, t_nnn, d_nn, d_nnn, n_nn, n_nnn, esite_, lanzero_
  logical, save :: caches_ready = .false.
  refb_init = 0
  if (ny_ /= Ny) then
    refb_init = -1
    return
  end if
  jz = 0; jd = 0; jt = 0; jn = 0
  jt(1:4, 1:num_site) = t_nn
  jt(5:8, 1:num_site) = t_nnn
  jd(1:4, 1:num_site) = d_nn
  jd(5:8, 1:num_site) = d_nnn
  jn(1:4) = n_nn
  jn(5:8) = n_nnn
  jz(1:4, 1:num_site) = -4.0d0 * n_nn
  jz(5:8, 1:num_site) = -4.0d0 * n_nnn
  esite = esite_
  lanzero = lanzero_
  call dmrg_setup0()
  call dmrg_setup(num_site/2 - 1, num_site/2 - 1)
  call dmrg_setup0()
  call Get_Lattice
  call make_site_ops(ub)
  ! the 6j memo tables (common.f90:16-21) hold values of a pure function of
  ! their keys, so they are cleared once and kept across calls
  if (.not. caches_ready) then
    iw6j1 = 0; w6j1 = 0.d0; iw6j2 = 0; w6j2 = 0.d0; iw6j3 = 0; w6j3 = 0.d0
    call factrl
    caches_ready = .true.
  end if
  cone = 1.0d0
  czero = 0.0d0
end function refb_init

! y = H_eff x through the reference's hmerge + model_wave.
integer(c_int) function refb_model_wave(sys_len, env_len, sys_nops, sys_hdr, sys_sub, sys_off, sys_data, &
    env_nops, env_hdr, env_sub, env_off, env_data, sys_nsec, sys_bsec, sys_bsub, env_nsec, env_bsec, env_bsub, &
    nw, wsec, x, y) bind(C, name="refb_model_wave")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: sys_len, env_len, sys_nops, env_nops, sys_nsec, env_nsec, nw
  integer(c_int), intent(in) :: sys_hdr(4, *), sys_sub(5, *), env_hdr(4, *), env_sub(5, *)
  integer(c_int64_t), intent(in) :: sys_off(*), env_off(*)
  real(c_double), intent(in) :: sys_data(*), env_data(*), x(*)
  integer(c_int), intent(in) :: sys_bsec(4, *), sys_bsub(6, *), env_bsec(4, *), env_bsub(6, *), wsec(6, *)
  real(c_double), intent(inout) :: y(*)
  type(Total_Model) :: sys, env
  type(Total_Basis) :: sys_bs, env_bs
  type(Wavefunction) :: invec, outvec
  call put_model(sys, sys_len, sys_nops, sys_hdr, sys_sub, sys_off, sys_data)
  call put_model(env, env_len, env_nops, env_hdr, env_sub, env_off, env_data)
  call put_basis(sys_bs, sys_len + 1, sys_nsec, sys_bsec, sys_bsub)
  call put_basis(env_bs, env_len + 1, env_nsec, env_bsec, env_bsub)
  call put_wave(invec, sys_len + 1, env_len + 1, nw, wsec, x)
  call put_wave(outvec, sys_len + 1, env_len + 1, nw, wsec, x)
  call dmrg_setup(sys_len, env_len)
  call hmerge(sys, env, sys_bs, env_bs)
  call model_wave(sys, env, sys_bs, env_bs, invec, outvec)
  call get_wave(outvec, y)
  call free_model(sys)
  call free_model(env)
  refb_model_wave = 0
end function refb_model_wave

! Ground state through the reference's lanczos_memo_new with an injected start
! vector (eigen.f90:2-175); returns E0 and overwrites x with the ground state.
integer(c_int) function refb_lanczos(sys_len, env_len, sys_nops, sys_hdr, sys_sub, sys_off, sys_data, &
    env_nops, env_hdr, env_sub, env_off, env_data, sys_nsec, sys_bsec, sys_bsub, env_nsec, env_bsec, env_bsub, &
    nw, wsec, x, e0) bind(C, name="refb_lanczos")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: sys_len, env_len, sys_nops, env_nops, sys_nsec, env_nsec, nw
  integer(c_int), intent(in) :: sys_hdr(4, *), sys_sub(5, *), env_hdr(4, *), env_sub(5, *)
  integer(c_int64_t), intent(in) :: sys_off(*), env_off(*)
  real(c_double), intent(in) :: sys_data(*), env_data(*)
  real(c_double), intent(inout) :: x(*)
  integer(c_int), intent(in) :: sys_bsec(4, *), sys_bsub(6, *), env_bsec(4, *), env_bsub(6, *), wsec(6, *)
  real(c_double), intent(out) :: e0
  type(Total_Model) :: sys, env
  type(Total_Basis) :: sys_bs, env_bs
  type(Super_Basis) :: super
  type(Wavefunction) :: eigvec(1)
  integer :: i
  call put_model(sys, sys_len, sys_nops, sys_hdr, sys_sub, sys_off, sys_data)
  call put_model(env, env_len, env_nops, env_hdr, env_sub, env_off, env_data)
  call put_basis(sys_bs, sys_len + 1, sys_nsec, sys_bsec, sys_bsub)
  call put_basis(env_bs, env_len + 1, env_nsec, env_bsec, env_bsub)
  call put_wave(eigvec(1), sys_len + 1, env_len + 1, nw, wsec, x)
  super%sys_len = sys_len + 1
  super%env_len = env_len + 1
  super%num = nw
  super%dim = eigvec(1)%dim
  super%idx = nw
  allocate(super%sub(nw))
  do i = 1, nw
    super%sub(i)%sys_num_up = wsec(1, i)
    super%sub(i)%sys_num_down = wsec(2, i)
    super%sub(i)%env_num_up = wsec(3, i)
    super%sub(i)%env_num_down = wsec(4, i)
    super%sub(i)%sys_dim = wsec(5, i)
    super%sub(i)%env_dim = wsec(6, i)
  end do
  call dmrg_setup(sys_len, env_len)
  call lanczos_memo_new(sys, env, sys_bs, env_bs, super, eigvec, 1, 1)
  e0 = Energys(1)
  call get_wave(eigvec(1), x)
  call free_model(sys)
  call free_model(env)
  refb_lanczos = 0
end function refb_lanczos

! W6JS of the reference (block.f90:1564-1615); refb_init must have run FACTRL.
real(c_double) function refb_w6js(j1, j2, j3, l1, l2, l3) bind(C, name="refb_w6js")
  use iso_c_binding
  implicit none
  integer(c_int), value :: j1, j2, j3, l1, l2, l3
  real(8), external :: w6js
  refb_w6js = w6js(j1, j2, j3, l1, l2, l3)
end function refb_w6js

! Get_basis of the reference (block.f90:9-104) for a block given by its
! sector list blk(3, nb) = (num_up, num_down, dim) and the site st_sz-like
! list site(3, ns).  Returns the counts; sec(4, *) and sub(6, *) receive the
! tables in the flat layout.
integer(c_int) function refb_get_basis(blen, nb, blk, ns, site, sec, sub, nsec, nsub) bind(C, name="refb_get_basis")
  use iso_c_binding
  use pubdata
  implicit none
  integer(c_int), value :: blen, nb, ns
  integer(c_int), intent(in) :: blk(3, *), site(3, *)
  integer(c_int), intent(inout) :: sec(4, *), sub(6, *)
  integer(c_int), intent(out) :: nsec, nsub
  type(Total_Block) :: block, st
  type(Total_Basis) :: basis
  integer :: i, j, q
  block%len = blen; block%num = nb; block%up_dif = 0; block%down_dif = 0
  allocate(block%sub(nb))
  do i = 1, nb
    block%sub(i)%num_up = blk(1, i); block%sub(i)%num_down = blk(2, i)
    block%sub(i)%down_dif = 0; block%sub(i)%row_dim = blk(3, i); block%sub(i)%sdim = blk(3, i)
  end do
  st%len = 1; st%num = ns; st%up_dif = 0; st%down_dif = 0
  allocate(st%sub(ns))
  do i = 1, ns
    st%sub(i)%num_up = site(1, i); st%sub(i)%num_down = site(2, i)
    st%sub(i)%down_dif = 0; st%sub(i)%row_dim = site(3, i); st%sub(i)%sdim = site(3, i)
  end do
  call Get_basis(block, st, basis)
  nsec = basis%num
  q = 0
  do i = 1, basis%num
    sec(1, i) = basis%sub(i)%new_num_up; sec(2, i) = basis%sub(i)%new_num_down
    sec(3, i) = basis%sub(i)%dim; sec(4, i) = basis%sub(i)%num
    do j = 1, basis%sub(i)%num
      q = q + 1
      sub(1, q) = basis%sub(i)%sub(j)%bl_num_up; sub(2, q) = basis%sub(i)%sub(j)%bl_num_down
      sub(3, q) = basis%sub(i)%sub(j)%st_num_up; sub(4, q) = basis%sub(i)%sub(j)%st_num_down
      sub(5, q) = basis%sub(i)%sub(j)%spos; sub(6, q) = basis%sub(i)%sub(j)%sdim
    end do
  end do
  nsub = q
  refb_get_basis = 0
end function refb_get_basis

! Get_super_basis of the reference (block.f90:110-161) over two fused bases
! given as flat (sec, sub) tables; the sector list comes back as rows of
! (sys_up, sys_down, env_up, env_down, sys_dim, env_dim) in the routine's order.
integer(c_int) function refb_get_super_basis(slen, snsec, ssec, ssub, elen, ensec, esec, esub, total_up, &
    total_down, wsec, nw, dim) bind(C, name="refb_get_super_basis")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: slen, snsec, elen, ensec, total_up, total_down
  integer(c_int), intent(in) :: ssec(4, *), ssub(6, *), esec(4, *), esub(6, *)
  integer(c_int), intent(inout) :: wsec(6, *)
  integer(c_int), intent(out) :: nw, dim
  type(Total_Basis) :: sbs, ebs
  type(Super_Basis) :: super
  integer :: i
  call put_basis(sbs, slen, snsec, ssec, ssub)
  call put_basis(ebs, elen, ensec, esec, esub)
  call Get_super_basis(sbs, ebs, super, total_up, total_down)
  nw = super%num
  dim = super%dim
  do i = 1, super%num
    wsec(1, i) = super%sub(i)%sys_num_up; wsec(2, i) = super%sub(i)%sys_num_down
    wsec(3, i) = super%sub(i)%env_num_up; wsec(4, i) = super%sub(i)%env_num_down
    wsec(5, i) = super%sub(i)%sys_dim; wsec(6, i) = super%sub(i)%env_dim
  end do
  refb_get_super_basis = 0
end function refb_get_super_basis

! Sizes of the operators held from the last producing call.
integer(c_int) function refb_result_sizes(nops, nsub, ndata) bind(C, name="refb_result_sizes")
  use iso_c_binding
  use ref_bridge_util
  implicit none
  integer(c_int), intent(out) :: nops, nsub
  integer(c_int64_t), intent(out) :: ndata
  integer :: o, i
  nops = 0; nsub = 0; ndata = 0
  if (allocated(res)) then
    nops = size(res)
    do o = 1, size(res)
      if (.not. allocated(res(o)%sub)) cycle
      do i = 1, res(o)%num
        nsub = nsub + 1
        ndata = ndata + int(res(o)%sub(i)%row_dim, 8) * res(o)%sub(i)%sdim
      end do
    end do
  end if
  refb_result_sizes = 0
end function refb_result_sizes

! The held operators in the flat layout.
integer(c_int) function refb_result_copy(hdr, sub, off, data) bind(C, name="refb_result_copy")
  use iso_c_binding
  use ref_bridge_util
  implicit none
  integer(c_int), intent(inout) :: hdr(4, *), sub(5, *)
  integer(c_int64_t), intent(inout) :: off(*)
  real(c_double), intent(inout) :: data(*)
  integer :: o, i, r, c, s
  integer(8) :: nd
  s = 0; nd = 0
  do o = 1, size(res)
    hdr(1, o) = res(o)%len; hdr(2, o) = res(o)%up_dif; hdr(3, o) = res(o)%down_dif; hdr(4, o) = 0
    if (.not. allocated(res(o)%sub)) cycle
    hdr(4, o) = res(o)%num
    do i = 1, res(o)%num
      s = s + 1
      sub(1, s) = res(o)%sub(i)%num_up; sub(2, s) = res(o)%sub(i)%num_down
      sub(3, s) = res(o)%sub(i)%down_dif; sub(4, s) = res(o)%sub(i)%row_dim; sub(5, s) = res(o)%sub(i)%sdim
      off(s) = nd
      do c = 1, res(o)%sub(i)%sdim
        do r = 1, res(o)%sub(i)%row_dim
          data(nd + (c - 1) * res(o)%sub(i)%row_dim + r) = res(o)%sub(i)%mat(r, c)
        end do
      end do
      nd = nd + int(res(o)%sub(i)%row_dim, 8) * res(o)%sub(i)%sdim
    end do
  end do
  refb_result_copy = 0
end function refb_result_copy

! Get_Hamiltonian_Sys (side 0) / Get_Hamiltonian_Env (side 1) on a block model
! of length sys_len (or env_len) at the superblock (sys_len, env_len); mshift
! as the sweep sets it (dmrg_sub.f90:692, 989).  Result: the new model's
! Hamiltonian and nleg11(len+1) x {elec_up, elec_down, spin_sd, num_sn}.
integer(c_int) function refb_get_hamiltonian(side, sys_len, env_len, pre_nops, pre_hdr, pre_sub, pre_off, &
    pre_data, blen, nsec, bsec, bsub, trun_hdr, trun_sub, trun_off, trun_data) bind(C, name="refb_get_hamiltonian")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: side, sys_len, env_len, pre_nops, blen, nsec
  integer(c_int), intent(in) :: pre_hdr(4, *), pre_sub(5, *), bsec(4, *), bsub(6, *), trun_hdr(4), trun_sub(5, *)
  integer(c_int64_t), intent(in) :: pre_off(*), trun_off(*)
  real(c_double), intent(in) :: pre_data(*), trun_data(*)
  type(Total_Model) :: pre, nxt
  type(Total_Basis) :: basis
  type(Total_Block) :: trun
  integer :: plen, n1, ii
  plen = sys_len
  if (side == 1) plen = env_len
  call put_model(pre, plen, pre_nops, pre_hdr, pre_sub, pre_off, pre_data)
  call put_basis(basis, blen, nsec, bsec, bsub)
  call put_block(trun, blen, trun_hdr, trun_sub, trun_off, trun_data)
  call dmrg_setup(sys_len, env_len)
  mshift = sys_len + 1
  if (side == 0) then
    call Get_Hamiltonian_Sys(pre, nxt, basis, trun)
  else
    call Get_Hamiltonian_Env(pre, nxt, basis, trun)
  end if
  n1 = nleg11(plen + 1)
  call set_result(1 + 4 * n1)
  call block_transfer(pre%ham, res(1))
  do ii = 1, n1
    call block_transfer(pre%sub(ii)%elec_up, res(2 + 4 * (ii - 1)))
    call block_transfer(pre%sub(ii)%elec_down, res(3 + 4 * (ii - 1)))
    call block_transfer(pre%sub(ii)%spin_sd, res(4 + 4 * (ii - 1)))
    call block_transfer(pre%sub(ii)%num_sn, res(5 + 4 * (ii - 1)))
  end do
  call free_model(pre)
  call free_model(nxt)
  call free_block(trun)
  refb_get_hamiltonian = 0
end function refb_get_hamiltonian

! update_trun_dia (ndia(o) = 0) / update_trun_ndia (ndia(o) = 1) of each operator.
integer(c_int) function refb_update_trun(nops, hdr, sub, off, data, ndia, trun_hdr, trun_sub, trun_off, &
    trun_data) bind(C, name="refb_update_trun")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: nops
  integer(c_int), intent(in) :: hdr(4, *), sub(5, *), ndia(*), trun_hdr(4), trun_sub(5, *)
  integer(c_int64_t), intent(in) :: off(*), trun_off(*)
  real(c_double), intent(in) :: data(*), trun_data(*)
  type(Total_Block) :: ori, trun
  integer :: o, first
  call put_block(trun, trun_hdr(1), trun_hdr, trun_sub, trun_off, trun_data)
  call set_result(nops)
  first = 1
  do o = 1, nops
    call put_block(ori, hdr(1, o), hdr(:, o), sub(1, first), off(first), data)
    if (ndia(o) /= 0) then
      call update_trun_ndia(ori, res(o), trun)
    else
      call update_trun_dia(ori, res(o), trun)
    end if
    call free_block(ori)
    first = first + hdr(4, o)
  end do
  call free_block(trun)
  refb_update_trun = 0
end function refb_update_trun

! Get_density_matrix_sys_new (side 0) / _env_new (side 1) of `levels` states
! stored back to back in psi.
integer(c_int) function refb_density(side, sys_len, env_len, nw, wsec, levels, psi) bind(C, name="refb_density")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: side, sys_len, env_len, nw, levels
  integer(c_int), intent(in) :: wsec(6, *)
  real(c_double), intent(in) :: psi(*)
  type(Wavefunction), allocatable :: w(:)
  integer :: l, i
  integer(8) :: d
  d = 0
  do i = 1, nw
    d = d + int(wsec(5, i), 8) * wsec(6, i)
  end do
  allocate(w(levels))
  do l = 1, levels
    call put_wave(w(l), sys_len, env_len, nw, wsec, psi((l - 1) * d + 1))
  end do
  call set_result(1)
  if (side == 0) then
    call Get_density_matrix_sys_new(res(1), w, levels)
  else
    call Get_density_matrix_env_new(res(1), w, levels)
  end if
  do l = 1, levels
    call deallocate_wave(w(l))
  end do
  refb_density = 0
end function refb_density

! Get_truncation_operator (den_mat.f90:141-383) with kept_min / kept_max and
! the error bound given; writes the reference's log files into the current
! directory.  Result: the truncation operator; num_state and Entropys out.
integer(c_int) function refb_truncation(den_hdr, den_sub, den_off, den_data, kmin, kmax, terr, sysblk, &
    num_state, entropy) bind(C, name="refb_truncation")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), intent(in) :: den_hdr(4), den_sub(5, *)
  integer(c_int64_t), intent(in) :: den_off(*)
  real(c_double), intent(in) :: den_data(*)
  integer(c_int), value :: kmin, kmax, sysblk
  real(c_double), value :: terr
  integer(c_int), intent(out) :: num_state
  real(c_double), intent(out) :: entropy
  type(Total_Block) :: den
  integer :: ns
  call put_block(den, den_hdr(1), den_hdr, den_sub, den_off, den_data)
  kept_min = kmin
  kept_max = kmax
  system_block = sysblk /= 0
  call set_result(1)
  ns = 0
  call Get_truncation_operator(den, res(1), terr, ns)
  num_state = ns
  entropy = Entropys
  call free_block(den)
  refb_truncation = 0
end function refb_truncation

! QuickSort_Double (block.f90:288-332) over arr(1:n), idx permuted alike.
integer(c_int) function refb_quicksort(n, arr, idx) bind(C, name="refb_quicksort")
  use iso_c_binding
  implicit none
  integer(c_int), value :: n
  real(c_double), intent(inout) :: arr(*)
  integer(c_int), intent(inout) :: idx(*)
  if (n > 0) call QuickSort_Double(arr, idx, 1, n, n)
  refb_quicksort = 0
end function refb_quicksort

! Flushes the reference's standard output (unit 6) so a caller can route what
! the reference prints (FACTRL's last factorial, the Lanczos log lines) away
! from its own standard output.
integer(c_int) function refb_flush() bind(C, name="refb_flush")
  use iso_c_binding
  implicit none
  flush(6)
  refb_flush = 0
end function refb_flush

! wave_trans_LR (dir 0) / wave_trans_RL (dir 1) (dmrg_sub.f90:1093-1384) from
! the old wavefunction x into y (sectors wsec_new, normalised by the routine).
integer(c_int) function refb_wave_trans(dir, sys_nsec, sys_bsec, sys_bsub, env_nsec, env_bsec, env_bsub, &
    st_hdr, st_sub, st_off, st_data, et_hdr, et_sub, et_off, et_data, nw_old, wsec_old, x, nw_new, wsec_new, y) &
    bind(C, name="refb_wave_trans")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: dir, sys_nsec, env_nsec, nw_old, nw_new
  integer(c_int), intent(in) :: sys_bsec(4, *), sys_bsub(6, *), env_bsec(4, *), env_bsub(6, *)
  integer(c_int), intent(in) :: st_hdr(4), st_sub(5, *), et_hdr(4), et_sub(5, *), wsec_old(6, *), wsec_new(6, *)
  integer(c_int64_t), intent(in) :: st_off(*), et_off(*)
  real(c_double), intent(in) :: st_data(*), et_data(*), x(*)
  real(c_double), intent(inout) :: y(*)
  type(Total_Basis) :: sbs, ebs
  type(Total_Block) :: st, et
  type(Wavefunction) :: ow, nw
  real(8), allocatable :: zero(:)
  integer(8) :: dn
  integer :: i
  call put_basis(sbs, 0, sys_nsec, sys_bsec, sys_bsub)
  call put_basis(ebs, 0, env_nsec, env_bsec, env_bsub)
  call put_block(st, st_hdr(1), st_hdr, st_sub, st_off, st_data)
  call put_block(et, et_hdr(1), et_hdr, et_sub, et_off, et_data)
  call put_wave(ow, 0, 0, nw_old, wsec_old, x)
  dn = 0
  do i = 1, nw_new
    dn = dn + int(wsec_new(5, i), 8) * wsec_new(6, i)
  end do
  allocate(zero(max(dn, 1_8)))
  zero = 0.0d0
  call put_wave(nw, 0, 0, nw_new, wsec_new, zero)
  if (dir == 0) then
    call wave_trans_LR(sbs, ebs, st, et, ow, nw)
  else
    call wave_trans_RL(sbs, ebs, st, et, ow, nw)
  end if
  call get_wave(nw, y)
  call free_block(st)
  call free_block(et)
  deallocate(sbs%sub, ebs%sub, ow%sub, nw%sub, zero)
  refb_wave_trans = 0
end function refb_wave_trans

! ---------------------------------------------------------------------------
! Checkpoint files (block.f90:2760-3097), written into and read from the
! current directory under the names of Get_model_name (dmrg_sub.f90:294).
! flag /= 0 selects the system names, 0 the environment names.
! ---------------------------------------------------------------------------

! The `dim` field of each operator held from the last producing call.
integer(c_int) function refb_result_dims(dims) bind(C, name="refb_result_dims")
  use iso_c_binding
  use ref_bridge_util
  implicit none
  integer(c_int), intent(inout) :: dims(*)
  integer :: o
  do o = 1, size(res)
    dims(o) = res(o)%dim
  end do
  refb_result_dims = 0
end function refb_result_dims

! model_to_disk of a flat model of length len with the given dim fields;
! nops must be 1 + 4 nleg11(len).
integer(c_int) function refb_model_to_disk(len, nops, hdr, sub, off, data, dims, idx, flag) &
    bind(C, name="refb_model_to_disk")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: len, nops, idx, flag
  integer(c_int), intent(in) :: hdr(4, *), sub(5, *), dims(*)
  integer(c_int64_t), intent(in) :: off(*)
  real(c_double), intent(in) :: data(*)
  type(Total_Model) :: m
  integer :: ii
  if (1 + 4 * nleg11(len) /= nops) then
    refb_model_to_disk = -1
    return
  end if
  call Get_model_name(sysname, envname)
  call put_model(m, len, nops, hdr, sub, off, data)
  m%ham%dim = dims(1)
  do ii = 1, nleg11(len)
    m%sub(ii)%elec_up%dim = dims(2 + 4 * (ii - 1))
    m%sub(ii)%elec_down%dim = dims(3 + 4 * (ii - 1))
    m%sub(ii)%spin_sd%dim = dims(4 + 4 * (ii - 1))
    m%sub(ii)%num_sn%dim = dims(5 + 4 * (ii - 1))
  end do
  call model_to_disk(m, 1001, idx, flag /= 0)
  call free_model(m)
  refb_model_to_disk = 0
end function refb_model_to_disk

! model_from_disk; the model is held as the result operators.
integer(c_int) function refb_model_from_disk(idx, flag) bind(C, name="refb_model_from_disk")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: idx, flag
  type(Total_Model) :: m
  integer :: n1, ii
  call Get_model_name(sysname, envname)
  call model_from_disk(m, 1001, idx, flag /= 0)
  n1 = nleg11(m%len)
  call set_result(1 + 4 * n1)
  call block_transfer(m%ham, res(1))
  do ii = 1, n1
    call block_transfer(m%sub(ii)%elec_up, res(2 + 4 * (ii - 1)))
    call block_transfer(m%sub(ii)%elec_down, res(3 + 4 * (ii - 1)))
    call block_transfer(m%sub(ii)%spin_sd, res(4 + 4 * (ii - 1)))
    call block_transfer(m%sub(ii)%num_sn, res(5 + 4 * (ii - 1)))
  end do
  call free_model(m)
  refb_model_from_disk = m%len
end function refb_model_from_disk

! truns_to_disk of one flat block with the given dim field.
integer(c_int) function refb_truns_to_disk(hdr, sub, off, data, dim, idx, flag) bind(C, name="refb_truns_to_disk")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), intent(in) :: hdr(4), sub(5, *)
  integer(c_int64_t), intent(in) :: off(*)
  real(c_double), intent(in) :: data(*)
  integer(c_int), value :: dim, idx, flag
  type(Total_Block) :: b
  call Get_model_name(sysname, envname)
  call put_block(b, hdr(1), hdr, sub, off, data)
  b%dim = dim
  call truns_to_disk(b, 1001, idx, flag /= 0)
  call free_block(b)
  refb_truns_to_disk = 0
end function refb_truns_to_disk

! truns_from_disk; the block is held as the single result operator.
integer(c_int) function refb_truns_from_disk(idx, flag) bind(C, name="refb_truns_from_disk")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: idx, flag
  call Get_model_name(sysname, envname)
  call set_result(1)
  call truns_from_disk(res(1), 1001, idx, flag /= 0)
  refb_truns_from_disk = 0
end function refb_truns_from_disk

! basis_to_disk of a flat basis.
integer(c_int) function refb_basis_to_disk(len, nsec, bsec, bsub, idx, flag) bind(C, name="refb_basis_to_disk")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: len, nsec, idx, flag
  integer(c_int), intent(in) :: bsec(4, *), bsub(6, *)
  type(Total_Basis) :: bs
  call Get_model_name(sysname, envname)
  call put_basis(bs, len, nsec, bsec, bsub)
  call basis_to_disk(bs, 1001, idx, flag /= 0)
  call deallocate_basis(bs)
  refb_basis_to_disk = 0
end function refb_basis_to_disk

! basis_from_disk; returns the length and the table sizes, the tables follow
! from refb_basis_copy.
integer(c_int) function refb_basis_from_disk(idx, flag, len, nsec, nsub) bind(C, name="refb_basis_from_disk")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: idx, flag
  integer(c_int), intent(out) :: len, nsec, nsub
  integer :: i
  call Get_model_name(sysname, envname)
  if (allocated(bres%sub)) call deallocate_basis(bres)
  call basis_from_disk(bres, 1001, idx, flag /= 0)
  len = bres%len
  nsec = bres%num
  nsub = 0
  do i = 1, bres%num
    nsub = nsub + bres%sub(i)%num
  end do
  refb_basis_from_disk = 0
end function refb_basis_from_disk

integer(c_int) function refb_basis_copy(bsec, bsub, bidx) bind(C, name="refb_basis_copy")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), intent(inout) :: bsec(4, *), bsub(6, *), bidx(*)
  integer :: i, j, q
  q = 0
  do i = 1, bres%num
    bsec(1, i) = bres%sub(i)%new_num_up; bsec(2, i) = bres%sub(i)%new_num_down
    bsec(3, i) = bres%sub(i)%dim; bsec(4, i) = bres%sub(i)%num
    bidx(i) = bres%sub(i)%idx
    do j = 1, bres%sub(i)%num
      q = q + 1
      bsub(1, q) = bres%sub(i)%sub(j)%bl_num_up; bsub(2, q) = bres%sub(i)%sub(j)%bl_num_down
      bsub(3, q) = bres%sub(i)%sub(j)%st_num_up; bsub(4, q) = bres%sub(i)%sub(j)%st_num_down
      bsub(5, q) = bres%sub(i)%sub(j)%spos; bsub(6, q) = bres%sub(i)%sub(j)%sdim
    end do
  end do
  refb_basis_copy = 0
end function refb_basis_copy

! wave_to_disk of a flat wavefunction into the file `name`.
integer(c_int) function refb_wave_to_disk(sys_len, env_len, nw, wsec, x, name, name_len) &
    bind(C, name="refb_wave_to_disk")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: sys_len, env_len, nw, name_len
  integer(c_int), intent(in) :: wsec(6, *)
  real(c_double), intent(in) :: x(*)
  character(kind=c_char), intent(in) :: name(*)
  character(len=name_len) :: fname
  type(Wavefunction) :: w
  integer :: i
  do i = 1, name_len
    fname(i:i) = name(i)
  end do
  call put_wave(w, sys_len, env_len, nw, wsec, x)
  call wave_to_disk(w, 1001, fname, name_len)
  call deallocate_wave(w)
  refb_wave_to_disk = 0
end function refb_wave_to_disk

! wave_from_disk of the file `name`; returns sizes, the data follow from
! refb_wave_copy.
integer(c_int) function refb_wave_from_disk(name, name_len, sys_len, env_len, nw, dim) &
    bind(C, name="refb_wave_from_disk")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: name_len
  character(kind=c_char), intent(in) :: name(*)
  integer(c_int), intent(out) :: sys_len, env_len, nw, dim
  character(len=name_len) :: fname
  integer :: i
  do i = 1, name_len
    fname(i:i) = name(i)
  end do
  if (allocated(wres%sub)) call deallocate_wave(wres)
  call wave_from_disk(wres, 1001, fname, name_len)
  sys_len = wres%sys_len
  env_len = wres%env_len
  nw = wres%num
  dim = wres%dim
  refb_wave_from_disk = 0
end function refb_wave_from_disk

integer(c_int) function refb_wave_copy(wsec, y) bind(C, name="refb_wave_copy")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), intent(inout) :: wsec(6, *)
  real(c_double), intent(inout) :: y(*)
  integer :: i
  do i = 1, wres%num
    wsec(1, i) = wres%sub(i)%sys_num_up; wsec(2, i) = wres%sub(i)%sys_num_down
    wsec(3, i) = wres%sub(i)%env_num_up; wsec(4, i) = wres%sub(i)%env_num_down
    wsec(5, i) = wres%sub(i)%sys_dim; wsec(6, i) = wres%sub(i)%env_dim
  end do
  call get_wave(wres, y)
  refb_wave_copy = 0
end function refb_wave_copy

! ---------------------------------------------------------------------------
! Measurements: measure_operator_dia (measure.f90:3-57) of the site operator
! num_elec, over bases and truncation operators stored into the reference's
! globals sys_bsm / env_bsm / systruns / envtruns (common.f90:159-160), as
! Measurement reads them from disk (main.f90:220-238).
! ---------------------------------------------------------------------------
integer(c_int) function refb_set_bsm(side, idx, len, nsec, bsec, bsub) bind(C, name="refb_set_bsm")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: side, idx, len, nsec
  integer(c_int), intent(in) :: bsec(4, *), bsub(6, *)
  if (side == 0) then
    if (allocated(sys_bsm(idx)%sub)) call deallocate_basis(sys_bsm(idx))
    call put_basis(sys_bsm(idx), len, nsec, bsec, bsub)
  else
    if (allocated(env_bsm(idx)%sub)) call deallocate_basis(env_bsm(idx))
    call put_basis(env_bsm(idx), len, nsec, bsec, bsub)
  end if
  refb_set_bsm = 0
end function refb_set_bsm

integer(c_int) function refb_set_truns(side, idx, hdr, sub, off, data) bind(C, name="refb_set_truns")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: side, idx
  integer(c_int), intent(in) :: hdr(4), sub(5, *)
  integer(c_int64_t), intent(in) :: off(*)
  real(c_double), intent(in) :: data(*)
  if (side == 0) then
    call put_block(systruns(idx), hdr(1), hdr, sub, off, data)
  else
    call put_block(envtruns(idx), hdr(1), hdr, sub, off, data)
  end if
  refb_set_truns = 0
end function refb_set_truns

! <n_i> of every lattice site in the wavefunction (sys_len, env_len are the
! wavefunction's lengths, block + site); ave(Num_site) receives ave_dia.
integer(c_int) function refb_measure_density(sys_len, env_len, nw, wsec, x, trun_idx, ave) &
    bind(C, name="refb_measure_density")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: sys_len, env_len, nw, trun_idx
  integer(c_int), intent(in) :: wsec(6, *)
  real(c_double), intent(in) :: x(*)
  real(c_double), intent(inout) :: ave(*)
  type(Wavefunction) :: w
  real(8) :: ave_dia(Num_site)
  character(len=30) :: fname
  fname = "num_elec.dat"
  call put_wave(w, sys_len, env_len, nw, wsec, x)
  call measure_operator_dia(ave_dia, num_elec, w, trun_idx, fname)
  ave(1:Num_site) = ave_dia
  call deallocate_wave(w)
  refb_measure_density = Num_site
end function refb_measure_density

! The across-the-bond pairs of Get_oper_cor_dia (measure.f90:468-634) for
! num_elec, with the reference's own routines: every system-side site i
! (1..sys_len, the last the bond site) with every environment-side site j
! (1..env_len, the last the bond site); cor(sys_len, env_len).
integer(c_int) function refb_cross_cor(sys_len, env_len, nw, wsec, x, cor) bind(C, name="refb_cross_cor")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: sys_len, env_len, nw
  integer(c_int), intent(in) :: wsec(6, *)
  real(c_double), intent(in) :: x(*)
  real(c_double), intent(inout) :: cor(sys_len, env_len)
  type(Wavefunction) :: w
  type(Total_Block) :: so, eo
  integer :: i, j
  real(8) :: v
  call put_wave(w, sys_len, env_len, nw, wsec, x)
  do i = 1, sys_len - 1
    call Get_sys_oper_dia(num_elec, i, so, sys_len - 1, 1, .true.)
    do j = 1, env_len - 1
      call Get_env_oper_dia(num_elec, j, eo, env_len - 1, 1, .true.)
      call sys_block_env_block_cor_dia(v, so, eo, sys_bsm(sys_len), env_bsm(env_len), w)
      cor(i, j) = v
      call deallocate_block(eo)
    end do
    call sys_block_env_site_cor_dia(v, so, num_elec, sys_bsm(sys_len), env_bsm(env_len), w)
    cor(i, env_len) = v
    call deallocate_block(so)
  end do
  do j = 1, env_len - 1
    call Get_env_oper_dia(num_elec, j, eo, env_len - 1, 1, .true.)
    call sys_site_env_block_cor_dia(v, num_elec, eo, sys_bsm(sys_len), env_bsm(env_len), w)
    cor(sys_len, j) = v
    call deallocate_block(eo)
  end do
  call sys_site_env_site_cor_dia(v, num_elec, num_elec, sys_bsm(sys_len), env_bsm(env_len), w)
  cor(sys_len, env_len) = v
  call deallocate_wave(w)
  refb_cross_cor = 0
end function refb_cross_cor

! Get_oper_cor_dia (measure.f90:468-634) of num_elec in full: cor(Num_site,
! Num_site) receives oper_dia (it appends to oper_dia_tmp.dat and the
! given file in the current directory).
integer(c_int) function refb_oper_cor_dia(sys_len, env_len, nw, wsec, x, cor) bind(C, name="refb_oper_cor_dia")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: sys_len, env_len, nw
  integer(c_int), intent(in) :: wsec(6, *)
  real(c_double), intent(in) :: x(*)
  real(c_double), intent(inout) :: cor(Num_site, Num_site)
  type(Wavefunction) :: w
  real(8), allocatable :: od(:, :)
  character(len=30) :: fname
  fname = "num_elec_cor.dat"
  allocate(od(Num_site, Num_site))
  call put_wave(w, sys_len, env_len, nw, wsec, x)
  call Get_oper_cor_dia(od, num_elec, w, 1, fname)
  cor = od
  call deallocate_wave(w)
  deallocate(od)
  refb_oper_cor_dia = Num_site
end function refb_oper_cor_dia

! W3JS of the reference (block.f90:1522-1560); refb_init must have run FACTRL.
real(c_double) function refb_w3js(j1, j2, j3, m1, m2, m3) bind(C, name="refb_w3js")
  use iso_c_binding
  implicit none
  integer(c_int), value :: j1, j2, j3, m1, m2, m3
  real(8), external :: w3js
  refb_w3js = w3js(j1, j2, j3, m1, m2, m3)
end function refb_w3js

! The pairs of Get_density_cor (measure.f90:3110-3223) of system-side site
! sid1 with every later site of a chain of sys_len + env_len sites, with the
! reference's own routines in the order that routine calls them (trun_idx
! 1): cor(sid2) for the same-side sites sid2 = sid1+1 .. sys_len-1
! (Get_sys2b + measure_sys_block), cor(sys_len) (sys_block_site_cor_ndia),
! cor(sys_len+1) (sys_block_env_site_cor_ndia) and cor(sid3) for
! sid3 >= sys_len+2, environment site n - sid3 + 1
! (sys_block_env_block_cor_ndia), each scaled as Get_density_cor scales it.
! kind 0: num_elec x num_elec ('B'), 1: st_sd x st_sd ('B'),
! 2: st_elec_up x st_elec_down ('F').  The reference's routine walks sid3 to
! its compile-time Num_site and picks the sites in line with sid1; this one
! walks the chain's own sites and returns them all.
integer(c_int) function refb_density_cor(kind, sid1, sys_len, env_len, nw, wsec, x, cor) &
    bind(C, name="refb_density_cor")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: kind, sid1, sys_len, env_len, nw
  integer(c_int), intent(in) :: wsec(6, *)
  real(c_double), intent(in) :: x(*)
  real(c_double), intent(inout) :: cor(*)
  type(Wavefunction) :: w
  type(Total_Block) :: o1, o2, so, eo, nw1, nw2, tmp1
  character(len=1) :: flags
  integer :: n, sid2, sid3, eid1, trun_idx
  real(8) :: value
  refb_density_cor = -1
  if (sid1 < 1 .or. sid1 > sys_len - 1 .or. env_len < 2) return
  select case (kind)
  case (0)
    call block_transfer(num_elec, o1); call block_transfer(num_elec, o2); flags = 'B'
  case (1)
    call block_transfer(st_sd, o1); call block_transfer(st_sd, o2); flags = 'B'
  case (2)
    call block_transfer(st_elec_up, o1); call block_transfer(st_elec_down, o2); flags = 'F'
  case default
    return
  end select
  n = sys_len + env_len
  trun_idx = 1
  call put_wave(w, sys_len, env_len, nw, wsec, x)
  do sid2 = sid1 + 1, sys_len - 1
    call Get_sys2b(sid1, sid2, o1, o2, so, trun_idx, w, flags)
    call measure_sys_block(value, so, sys_bsm(sys_len), w)
    value = value * dsqrt(1.0d0 + o1%down_dif)
    if (o1%down_dif == 2) value = -value
    cor(sid2) = value
    call deallocate_block(so)
  end do
  call Get_sys_oper_ndia(o1, sid1, nw1, sid1, trun_idx, .false., flags)
  if (sid1 <= trun_idx) then
    call block_transfer(nw1, nw2)
  else
    call update_trun_ndia(nw1, nw2, systruns(sid1))
  end if
  call deallocate_block(nw1)
  call Change_sys_oper_ndia(nw2, sid1, so, sys_len - 1, trun_idx, .true., flags)
  call deallocate_block(nw2)
  do sid3 = sys_len, n
    value = 0.0d0
    if (sid3 == sys_len) then
      if (o2%down_dif == 1) then
        call block_transfer_trans(o2, tmp1)
        call sys_block_site_cor_ndia(value, so, tmp1, sys_bsm(sys_len), w, flags)
        call deallocate_block(tmp1)
      else
        call sys_block_site_cor_ndia(value, so, o2, sys_bsm(sys_len), w, flags)
      end if
    else if (sid3 == sys_len + 1) then
      if (o2%down_dif == 1) then
        call block_transfer_trans(o2, tmp1)
        call sys_block_env_site_cor_ndia(value, so, tmp1, sys_bsm(sys_len), env_bsm(env_len), w, flags)
        call deallocate_block(tmp1)
      else
        call sys_block_env_site_cor_ndia(value, so, o2, sys_bsm(sys_len), env_bsm(env_len), w, flags)
      end if
    else
      eid1 = n - sid3 + 1
      call Get_env_oper_ndia(o2, eid1, nw1, eid1, trun_idx, .false., flags)
      if (eid1 <= trun_idx) then
        call block_transfer(nw1, nw2)
      else
        call update_trun_ndia(nw1, nw2, envtruns(eid1))
      end if
      call deallocate_block(nw1)
      call Change_env_oper_ndia(nw2, eid1, eo, env_len - 1, trun_idx, .true., flags)
      call deallocate_block(nw2)
      if (o2%down_dif /= 0) then
        call block_transfer_trans(eo, tmp1)
        call sys_block_env_block_cor_ndia(value, so, tmp1, sys_bsm(sys_len), env_bsm(env_len), w, flags)
        call deallocate_block(tmp1)
      else
        call sys_block_env_block_cor_ndia(value, so, eo, sys_bsm(sys_len), env_bsm(env_len), w, flags)
      end if
      call deallocate_block(eo)
    end if
    value = value * dsqrt(1.0d0 + o1%down_dif)
    if (o1%down_dif == 2) value = -value
    cor(sid3) = value
  end do
  call deallocate_block(so)
  call deallocate_block(o1)
  call deallocate_block(o2)
  call deallocate_wave(w)
  refb_density_cor = n
end function refb_density_cor

! Get_SC_Order (measure.f90:3226-3376) over a chain of nb = sys_len + env_len
! sites with the neighbour table nbt (column k of site i = neib(i, k)), with
! the reference's own routines in the order and branches of that routine
! (trun_idx 1).  The routine addresses environment sites through its
! compile-time Num_site; here n = nb takes its place.  out(2k-1), out(2k) are
! val1(1), val1(2) of the k-th sid11 = xc1, xc1 + ny_, ...; the routine never
! assigns its val1(3:4).
integer(c_int) function refb_sc_order(xc1, xc2, ny_, sys_len, env_len, nb, nbt, nw, wsec, x, out) &
    bind(C, name="refb_sc_order")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: xc1, xc2, ny_, sys_len, env_len, nb, nw
  integer(c_int), intent(in) :: nbt(12, 0:nb), wsec(6, *)
  real(c_double), intent(in) :: x(*)
  real(c_double), intent(inout) :: out(*)
  type(Wavefunction) :: w
  type(Total_Basis) :: sys_bs, env_bs
  type(Total_Block) :: sys_oper, sys_oper1, sys_oper2, env_oper, new_oper, new_oper1, new_oper2, tmp1
  integer :: sid11, sid1, sid2, sid3, sid4, eid1, eid2, ij, k, n, trun_idx, lo, hi
  real(8) :: value, val1(2)
  refb_sc_order = -1
  n = nb
  trun_idx = 1
  call put_wave(w, sys_len, env_len, nw, wsec, x)
  call basis_transfer(sys_bsm(sys_len), sys_bs)
  call basis_transfer(env_bsm(env_len), env_bs)
  val1 = 0.0d0
  k = 0
  do sid11 = xc1, xc2, ny_
    sid1 = sid11
    do ij = 1, 3, 2
      sid2 = nbt(ij, sid1)
      lo = min(sid1, sid2)
      hi = max(sid1, sid2)
      value = 0.0d0
      if (hi < sys_len) then
        call Get_sys_two(lo, hi, sys_oper2, trun_idx, w)
        call Change_sys_oper_ndia(sys_oper2, hi, sys_oper, sys_len - 1, trun_idx, .true., 'B')
        call deallocate_block(sys_oper2)
        call measure_sys_block(value, sys_oper, sys_bsm(sys_len), w)
        call deallocate_block(sys_oper)
        val1((ij + 1) / 2) = value
      end if
      if (lo == sys_len) then
        if (hi == sys_len + 1) then
          call sys_site_env_site_cor_ndia(value, st_elec_up, st_elec_down, sys_bs, env_bs, w, 'F')
          val1((ij + 1) / 2) = value
        end if
        if (hi > sys_len + 1) then
          eid1 = n - hi + 1
          call env_cdag(eid1, env_oper)
          call block_transfer_trans(env_oper, tmp1)
          call sys_site_env_block_cor_ndia(value, st_elec_up, tmp1, sys_bs, env_bs, w, 'F')
          call deallocate_block(tmp1)
          call deallocate_block(env_oper)
        end if
        val1((ij + 1) / 2) = value
      end if
      if (lo == sys_len + 1) then
        if (hi > sys_len + 1) then
          eid1 = n - hi + 1
          call env_cdag(eid1, env_oper)
          call block_transfer_trans(env_oper, tmp1)
          call env_block_site_cor_ndia(value, tmp1, st_elec_up, env_bs, w, 'F')
          value = -value
          call deallocate_block(tmp1)
          call deallocate_block(env_oper)
        end if
        val1((ij + 1) / 2) = value
      end if
      if (lo < sys_len .and. hi >= sys_len) then
        call Get_sys_oper_ndia(st_elec_up, lo, new_oper, lo, trun_idx, .false., 'F')
        if (lo <= trun_idx) then
          call block_transfer(new_oper, new_oper1)
        else
          call update_trun_ndia(new_oper, new_oper1, systruns(lo))
        end if
        call deallocate_block(new_oper)
        call Change_sys_oper_ndia(new_oper1, lo, sys_oper1, sys_len - 1, trun_idx, .true., 'F')
        call deallocate_block(new_oper1)
        if (hi == sys_len) then
          call sys_block_site_cor_ndia(value, sys_oper1, st_elec_down, sys_bs, w, 'F')
        else if (hi == sys_len + 1) then
          call sys_block_env_site_cor_ndia(value, sys_oper1, st_elec_down, sys_bs, env_bs, w, 'F')
        else
          eid1 = n - hi + 1
          call env_cdag(eid1, env_oper)
          call block_transfer_trans(env_oper, tmp1)
          call sys_block_env_block_cor_ndia(value, sys_oper1, tmp1, sys_bs, env_bs, w, 'F')
          call deallocate_block(tmp1)
          call deallocate_block(env_oper)
        end if
        call deallocate_block(sys_oper1)
        val1((ij + 1) / 2) = value
      end if
      sid3 = sid1
      eid1 = n - sid3 + 1
      sid4 = nbt(ij, sid3)
      eid2 = n - sid4 + 1
      if (max(eid1, eid2) <= env_len - 1) then
        call Get_SCOP_Operator_Env(min(eid1, eid2), max(eid1, eid2), env_oper, env_len - 1, trun_idx, .true., 'S0')
        call measure_env_block(value, env_oper, env_bsm(env_len), w)
        call deallocate_block(env_oper)
        val1((ij + 1) / 2) = value
      end if
    end do
    out(2 * k + 1) = val1(1)
    out(2 * k + 2) = val1(2)
    k = k + 1
  end do
  call deallocate_wave(w)
  refb_sc_order = k
contains
  ! c^+ of environment site e at env_len - 1, as Get_SC_Order forms it
  subroutine env_cdag(e, op)
    integer, intent(in) :: e
    type(Total_Block), intent(inout) :: op
    type(Total_Block) :: a, b
    call Get_env_oper_ndia(st_elec_up, e, a, e, trun_idx, .false., 'F')
    if (e <= trun_idx) then
      call block_transfer(a, b)
    else
      call update_trun_ndia(a, b, envtruns(e))
    end if
    call deallocate_block(a)
    call Change_env_oper_ndia(b, e, op, env_len - 1, trun_idx, .true., 'F')
    call deallocate_block(b)
  end subroutine env_cdag
end function refb_sc_order

! Get_mSCOP_Operator_Cor0 (measure.f90:3377-3507) with xc1 = xc2 over a chain
! of nb = sys_len + env_len sites with the neighbour table nbt, with the
! reference's own routines in that routine's order (trun_idx 1, Flags 'S0',
! n = nb in place of Num_site).  Row r of out (7 doubles) holds sid1, sid3,
! mod(sid3 - sid1 + n, ny), (sid3-1)/ny - (sid1-1)/ny and val1(1:3) of each
! line the routine writes (its val2 columns come from the never-filled
! cor11 table).  Returns the number of rows, at most maxrows.
integer(c_int) function refb_sc_pair_cor0(xc1, ij, ny_, sys_len, env_len, nb, nbt, nw, wsec, x, maxrows, out) &
    bind(C, name="refb_sc_pair_cor0")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  implicit none
  integer(c_int), value :: xc1, ij, ny_, sys_len, env_len, nb, nw, maxrows
  integer(c_int), intent(in) :: nbt(12, 0:nb), wsec(6, *)
  real(c_double), intent(in) :: x(*)
  real(c_double), intent(inout) :: out(7, *)
  type(Wavefunction) :: w
  type(Total_Basis) :: sys_bs, env_bs
  type(Total_Block) :: sys_oper, sys_oper1, sys_oper2, env_oper
  integer :: sid11, sid1, sid2, sid3, sid4, eid1, eid2, i, n, trun_idx, nrow
  real(8) :: value, val1(3)
  refb_sc_pair_cor0 = -1
  n = nb
  trun_idx = 1
  nrow = 0
  call put_wave(w, sys_len, env_len, nw, wsec, x)
  call basis_transfer(sys_bsm(sys_len), sys_bs)
  call basis_transfer(env_bsm(env_len), env_bs)
  sid11 = xc1
  sid1 = sid11
  sid2 = nbt(ij, sid1)
  call Get_sys_two(min(sid1, sid2), max(sid1, sid2), sys_oper2, trun_idx, w)
  do sid3 = sid1 + ny_, sys_len - 3, ny_
    val1 = 0.0d0
    do i = 1, 5, 2
      sid4 = nbt(i, sid3)
      if (min(sid4, sid3) <= max(sid1, sid2)) cycle
      if (max(sid3, sid4) <= sys_len - 1) then
        call Get_sys_two1(max(sid1, sid2), min(sid3, sid4), max(sid3, sid4), sys_oper2, sys_oper, trun_idx, w)
        call measure_sys_block(value, sys_oper, sys_bsm(sys_len), w)
        val1((i + 1) / 2) = value
        call deallocate_block(sys_oper)
      end if
    end do
    call add_row(sid1, sid3)
  end do
  if (ny_ >= 6) then
    sid1 = sid11 + 4
    sid2 = nbt(ij, sid1)
    call Get_sys_two(min(sid1, sid2), max(sid1, sid2), sys_oper1, trun_idx, w)
    call Change_sys_oper_ndia(sys_oper1, sid2, sys_oper, sys_len - 1, trun_idx, .true., 'B')
    call deallocate_block(sys_oper1)
    do sid3 = sys_len + 2, sys_len + ny_ + 1
      if (mod(sid3 - sid1, ny_) /= 0) cycle
      call env_part(sid3)
      call add_row(sid1, sid3)
    end do
    call deallocate_block(sys_oper)
  end if
  sid1 = sid11
  sid2 = nbt(ij, sid1)
  call Change_sys_oper_ndia(sys_oper2, sid2, sys_oper, sys_len - 1, trun_idx, .true., 'B')
  call deallocate_block(sys_oper2)
  do sid3 = sys_len + 2, n - ny_
    if (mod(sid3 - sid1, ny_) /= 0) cycle
    call env_part(sid3)
    call add_row(sid1, sid3)
  end do
  call deallocate_block(sys_oper)
  call deallocate_wave(w)
  refb_sc_pair_cor0 = nrow
contains
  subroutine env_part(s3)
    integer, intent(in) :: s3
    integer :: k, s4, e1, e2
    val1 = 0.0d0
    e1 = n - s3 + 1
    do k = 1, 5, 2
      s4 = nbt(k, s3)
      e2 = n - s4 + 1
      if (max(e1, e2) <= env_len - 1) then
        call Get_SCOP_Operator_Env(min(e1, e2), max(e1, e2), env_oper, env_len - 1, trun_idx, .true., 'S0')
        call sys_block_env_block_cor_ndia(value, sys_oper, env_oper, sys_bs, env_bs, w, 'B')
        call deallocate_block(env_oper)
        val1((k + 1) / 2) = value
      end if
    end do
  end subroutine env_part
  subroutine add_row(s1, s3)
    integer, intent(in) :: s1, s3
    nrow = nrow + 1
    if (nrow > maxrows) return
    out(1, nrow) = s1
    out(2, nrow) = s3
    out(3, nrow) = mod(s3 - s1 + n, ny_)
    out(4, nrow) = (s3 - 1) / ny_ - (s1 - 1) / ny_
    out(5:7, nrow) = val1
  end subroutine add_row
end function refb_sc_pair_cor0

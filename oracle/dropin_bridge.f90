! dropin_bridge.f90 -- test entry points for the Fortran drop-in shim.
!
! TEST INFRASTRUCTURE ONLY.  Builds the reference's derived types from the
! flat test inputs (ref_bridge_util) and calls the drop-in routines of
! grand_sql_tj_amd/fortran/tjsq_dropin.f90 exactly as the reference's DMRG
! driver would call lanczos_memo_new / model_wave.
integer(c_int) function dropin_model_wave(sys_len, env_len, sys_nops, sys_hdr, sys_sub, sys_off, sys_data, &
    env_nops, env_hdr, env_sub, env_off, env_data, sys_nsec, sys_bsec, sys_bsub, env_nsec, env_bsec, env_bsub, &
    nw, wsec, x, y) bind(C, name="dropin_model_wave")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  use tjsq_dropin
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
  call gpu_model_wave(sys, env, sys_bs, env_bs, invec, outvec)
  call get_wave(outvec, y)
  call free_model(sys)
  call free_model(env)
  dropin_model_wave = 0
end function dropin_model_wave

integer(c_int) function dropin_lanczos(sys_len, env_len, sys_nops, sys_hdr, sys_sub, sys_off, sys_data, &
    env_nops, env_hdr, env_sub, env_off, env_data, sys_nsec, sys_bsec, sys_bsub, env_nsec, env_bsec, env_bsub, &
    nw, wsec, x, e0) bind(C, name="dropin_lanczos")
  use iso_c_binding
  use pubdata
  use ref_bridge_util
  use tjsq_dropin
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
  allocate(super%sub(nw))
  do i = 1, nw
    super%sub(i)%sys_num_up = wsec(1, i)
    super%sub(i)%sys_num_down = wsec(2, i)
    super%sub(i)%env_num_up = wsec(3, i)
    super%sub(i)%env_num_down = wsec(4, i)
    super%sub(i)%sys_dim = wsec(5, i)
    super%sub(i)%env_dim = wsec(6, i)
  end do
  call gpu_lanczos_memo_new(sys, env, sys_bs, env_bs, super, eigvec, 1, 1)
  e0 = Energys(1)
  call get_wave(eigvec(1), x)
  call free_model(sys)
  call free_model(env)
  dropin_lanczos = 0
end function dropin_lanczos

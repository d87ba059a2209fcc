! c1_driver.f90 -- the reference's own C1 run, for the fixture c1_reference.json.
!
! TEST INFRASTRUCTURE ONLY (built and run by make_c1.py in a scratch build
! directory; never part of the product).  A main program in place of the
! reference's code/tjsq/main.f90: it sets the parameters of main.f90's
! "change parameters here" block (main.f90:13-47) from c1.in instead of
! constants, does main.f90's set-up calls (main.f90:50-80), and runs the
! schedule of Get_Ground_State (main.f90:96-187; restart1 = restart2 = 0,
! iter1 = 0) with the reference's own warm_up_point, left_to_right and
! right_to_left (code/tjsq/dmrg_sub.f90:556-1087).  Measurement is not run.
! Each lanczos_memo_new appends its "Sys_len= ... E0= ..." line to
! Energy_All.dat in the working directory (eigen.f90:166-171).
program c1_driver
  use pubdata
  implicit none
  real(8) :: t_nn, t_nnn, d_nn, d_nnn, n_nn, n_nnn, e_site, lanerr, trun_err, tmp_err
  integer :: kmax, kmaxf, nsweep, nup, trun_idx, levels, iter, idx1

  open(21, file='c1.in', status='old')
  read(21, *) t_nn, t_nnn, d_nn, d_nnn, n_nn, n_nnn, e_site
  read(21, *) lanerr, trun_err
  read(21, *) kmax, kmaxf, nsweep, nup
  close(21)

  ! main.f90:13-47 with the C1 values
  restart1 = 0
  restart2 = 0
  sys_len_start = 0
  iter1 = 0
  kept_max = kmax
  sweep_num = nsweep
  kept_min = min(3 * kept_max / 4, 3600)
  kept_maxf = kmaxf
  kept_min = min(1000, kept_max * 3 / 4)
  kept_max = min(1200, kept_max)
  tot_num_up = nup
  jd = 0
  jz = 0
  jn = 0
  jt = 0
  jt(1:4, 1:num_site) = t_nn
  jt(5:8, 1:num_site) = t_nnn
  jd(1:4, 1:num_site) = d_nn
  jd(5:8, 1:num_site) = d_nnn
  jn(1:4) = n_nn
  jn(5:8) = n_nnn
  jz(1:4, 1:num_site) = -4.0d0 * n_nn
  jz(5:8, 1:num_site) = -4.0d0 * n_nnn
  esite = e_site

  ! main.f90:50-80
  call dmrg_setup0()
  call dmrg_setup(num_site / 2 - 1, num_site / 2 - 1)
  call dmrg_setup0()
  call Get_Lattice
  call Get_site_operator_su2_U0
  call Get_single_site_basis
  call Get_model_name(sysname, envname)
  call Get_wave_name
  iw6j1 = 0
  w6j1 = 0.d0
  iw6j2 = 0
  w6j2 = 0.d0
  iw6j3 = 0
  w6j3 = 0.d0
  call factrl
  infi_del = 0
  cone = 1.0d0
  czero = 0.0d0
  openy = 0
  pinz = 1
  pind = 1
  trun_idx = 1
  tot_num_down = 0
  levels = 1

  ! Get_Ground_State, main.f90:96-187 (restart1 = restart2 = 0, iter1 = 0)
  iter = iter1
  spectrum = .false.
  lanzero = lanerr * (10 ** 3)
  tmp_err = trun_err * 200
  call warm_up_point(trun_idx, Num_site / 2, tmp_err, levels)
  if (iter == 0) then
    lanzero = lanerr * 30
    tmp_err = trun_err * 15
  end if
  kept_max = kept_max * 2 / 3 + kept_maxf / 3
  kept_min = min(3 * kept_max / 4, 3600)
  idx1 = trun_idx
  sys_len_start = num_site / 2 - 1
  restart1 = restart2
  call left_to_right(sys_len_start, Num_site - idx1 - 1, tmp_err, levels, .true.)
  restart1 = 1
  restart2 = 1
  kept_min = min(3 * kept_max / 4, 3600)
  sys_len_start = num_site - idx1
  restart1 = restart2
  call right_to_left(sys_len_start - 1, idx1, tmp_err, levels, .true.)
  restart1 = 1
  kept_max = kept_maxf
  kept_min = min(3 * kept_max / 4, 3600)
  iter = iter + 1
  do while (iter <= sweep_num)
    lanzero = lanerr
    tmp_err = trun_err
    call left_to_right(idx1, Num_site - idx1 - 1, tmp_err, levels, .false.)
    call right_to_left(Num_site - idx1 - 1, idx1, tmp_err, levels, .false.)
    iter = iter + 1
  end do
  spectrum = .true.
  lanzero = lanerr
  tmp_err = trun_err
  call left_to_right(idx1, Num_site0 / 2 + 3, tmp_err, levels, .false.)
end program c1_driver

! coag_driver -- a Fortran host that advances one flat-weighted box with the
! GPU coagulation step through the pmc_coag_gpu module, the way
! run_part_timestep calls mc_coag (src/run_part.F90:919-920).
!
! usage: coag_driver DIR N_SPEC N_PART N_STEPS KERNEL DT SEED TEMP PRES BETA
!                    WMAG N_BIN R_MIN R_MAX MODE [N_PART_IDEAL]
! Reads DIR/dens.bin (n_spec doubles), DIR/vol.bin (vol(n_spec, n_part)),
! DIR/ids.bin (n_part int64); writes DIR/counters.txt (step, tot_n_samp,
! tot_n_coag per line), DIR/out_n.txt, DIR/out_vol.bin and DIR/out_ids.bin.
! With N_PART_IDEAL > 0 each step is followed by aero_state_rebalance with
! doubling and halving allowed (src/run_part.F90:982-984), and the final
! weight magnitude is written to DIR/out_wmag.txt.
program coag_driver

  use iso_c_binding
  use pmc_coag_gpu
  implicit none

  character(len=1024) :: dir, arg
  integer :: n_spec, n_part, n_steps, kernel, mode, n_bin, step, u, i
  integer :: tot_n_samp, tot_n_coag, n_out
  integer(c_int64_t) :: seed, next_id
  real(c_double) :: dt, temp, pres, beta, wmag, r_min, r_max, n_ideal
  real(c_double), allocatable :: dens(:), vol(:, :), vol_out(:, :)
  integer(c_int64_t), allocatable :: ids(:), ids_out(:)
  integer(c_int), allocatable :: group(:), klass(:), nprim(:)
  integer(c_int) :: wtype(1), n_got, ng, ncl
  real(c_double) :: wm(1), we(1)
  type(c_ptr) :: h

  call get_command_argument(1, dir)
  call int_arg(2, n_spec)
  call int_arg(3, n_part)
  call int_arg(4, n_steps)
  call int_arg(5, kernel)
  call real_arg(6, dt)
  call get_command_argument(7, arg)
  read(arg, *) seed
  call real_arg(8, temp)
  call real_arg(9, pres)
  call real_arg(10, beta)
  call real_arg(11, wmag)
  call int_arg(12, n_bin)
  call real_arg(13, r_min)
  call real_arg(14, r_max)
  call int_arg(15, mode)
  n_ideal = 0d0
  if (command_argument_count() >= 16) call real_arg(16, n_ideal)

  allocate(dens(n_spec), vol(n_spec, n_part), ids(n_part))
  allocate(group(n_part), klass(n_part), nprim(n_part))
  open(newunit=u, file=trim(dir)//'/dens.bin', access='stream', form='unformatted')
  read(u) dens
  close(u)
  open(newunit=u, file=trim(dir)//'/vol.bin', access='stream', form='unformatted')
  read(u) vol
  close(u)
  open(newunit=u, file=trim(dir)//'/ids.bin', access='stream', form='unformatted')
  read(u) ids
  close(u)
  group = 1
  klass = 1
  nprim = 1

  call coag_gpu_check(pmc_gpu_create(h, 0_c_int, 1_c_int, int(n_spec, c_int), &
       dens, 0_c_int))
  call coag_gpu_check(pmc_gpu_set_env(h, 0_c_int, temp, pres, beta))
  wtype(1) = 1
  wm(1) = wmag
  we(1) = 0d0
  call coag_gpu_check(pmc_gpu_set_weights(h, 0_c_int, 1_c_int, 1_c_int, wtype, &
       wm, we))
  call coag_gpu_check(pmc_gpu_set_rng(h, seed, 0_c_int32_t))
  call coag_gpu_check(pmc_gpu_upload(h, 0_c_int, int(n_part, c_int), &
       int(2 * n_part + 64, c_int), vol, ids, group, klass, nprim, c_null_ptr, &
       c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr))
  next_id = int(n_part, c_int64_t) + 1
  call coag_gpu_check(pmc_gpu_set_next_id(h, 0_c_int, next_id, 1_c_int64_t, &
       0_c_int64_t))
  if (n_bin > 0) then
     call coag_gpu_check(pmc_gpu_set_bin_grid(h, 0_c_int, int(n_bin, c_int), &
          r_min, r_max))
  end if
  call coag_gpu_check(pmc_gpu_set_mode(h, int(mode, c_int), PMC_PREC_FP64, 1d-5))
  if (n_ideal > 0d0) call coag_gpu_check(pmc_gpu_set_n_part_ideal(h, 0_c_int, n_ideal))

  open(newunit=u, file=trim(dir)//'/counters.txt', form='formatted')
  do step = 1, n_steps
     call coag_gpu_mc_coag(h, kernel, dt, tot_n_samp, tot_n_coag)
     if (n_ideal > 0d0) call coag_gpu_check(pmc_gpu_rebalance(h, 1_c_int, 1_c_int))
     write(u, '(i0,1x,i0,1x,i0)') step - 1, tot_n_samp, tot_n_coag
  end do
  close(u)

  call coag_gpu_check(pmc_gpu_get_n_part(h, 0_c_int, n_got))
  n_out = max(int(n_got), 1)
  allocate(vol_out(n_spec, n_out), ids_out(n_out))
  deallocate(group, klass, nprim)
  allocate(group(n_out), klass(n_out), nprim(n_out))
  call coag_gpu_check(pmc_gpu_download(h, 0_c_int, int(n_out, c_int), n_got, &
       vol_out, ids_out, group, klass, nprim, c_null_ptr, c_null_ptr, &
       c_null_ptr, c_null_ptr, c_null_ptr))
  open(newunit=u, file=trim(dir)//'/out_n.txt', form='formatted')
  write(u, '(i0)') n_got
  close(u)
  open(newunit=u, file=trim(dir)//'/out_vol.bin', access='stream', &
       form='unformatted', status='replace')
  write(u) (vol_out(:, i), i = 1, n_got)
  close(u)
  open(newunit=u, file=trim(dir)//'/out_ids.bin', access='stream', &
       form='unformatted', status='replace')
  write(u) ids_out(1:n_got)
  close(u)
  if (n_ideal > 0d0) then
     call coag_gpu_check(pmc_gpu_get_weights(h, 0_c_int, 1_c_int, ng, ncl, &
          wtype, wm, we))
     open(newunit=u, file=trim(dir)//'/out_wmag.txt', form='formatted')
     write(u, '(es25.17)') wm(1)
     close(u)
  end if
  call coag_gpu_check(pmc_gpu_destroy(h))

contains

  subroutine int_arg(k, v)
    integer, intent(in) :: k
    integer, intent(out) :: v
    character(len=256) :: a
    call get_command_argument(k, a)
    read(a, *) v
  end subroutine int_arg

  subroutine real_arg(k, v)
    integer, intent(in) :: k
    real(c_double), intent(out) :: v
    character(len=256) :: a
    call get_command_argument(k, a)
    read(a, *) v
  end subroutine real_arg

end program coag_driver

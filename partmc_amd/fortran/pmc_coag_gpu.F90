! pmc_coag_gpu -- Fortran ISO_C_BINDING interface to the gfx950 coagulation
! library (include/partmc_coag.h).
!
! The reference calls mc_coag(coag_kernel_type, env_state, aero_data,
! aero_state, del_t, tot_n_samp, tot_n_coag) (src/coagulation.F90:43-59)
! from run_part_timestep (src/run_part.F90:919-920).  This module gives a
! Fortran program the same step on the GPU: set up a handle once, mirror
! the particles (vol(n_spec, n_part) column-major, as aero_particle_t%vol),
! then call coag_gpu_mc_coag each timestep.  The C entry points return 0 or
! an error code; coag_gpu_check stops with the code and the library's text,
! as die_msg does (src/util.F90:134-143).
module pmc_coag_gpu

  use iso_c_binding
  implicit none

  integer, parameter :: PMC_MODE_EXACT = 0, PMC_MODE_PARALLEL = 1
  integer, parameter :: PMC_PREC_FP64 = 0, PMC_PREC_FP32_RECHECK = 1

  interface
     integer(c_int) function pmc_gpu_create(h, device, n_box, n_spec, density, &
          max_bin) bind(c, name="pmc_gpu_create")
       import :: c_int, c_ptr, c_double
       type(c_ptr), intent(out) :: h
       integer(c_int), value :: device, n_box, n_spec, max_bin
       real(c_double), intent(in) :: density(*)
     end function pmc_gpu_create

     integer(c_int) function pmc_gpu_destroy(h) bind(c, name="pmc_gpu_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function pmc_gpu_destroy

     type(c_ptr) function pmc_gpu_last_error() bind(c, name="pmc_gpu_last_error")
       import :: c_ptr
     end function pmc_gpu_last_error

     integer(c_int) function pmc_gpu_set_env(h, box, temp, pressure, beta) &
          bind(c, name="pmc_gpu_set_env")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: box
       real(c_double), value :: temp, pressure, beta
     end function pmc_gpu_set_env

     integer(c_int) function pmc_gpu_set_weights(h, box, n_group, n_class, &
          wtype, magnitude, exponent) bind(c, name="pmc_gpu_set_weights")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: box, n_group, n_class
       integer(c_int), intent(in) :: wtype(*)
       real(c_double), intent(in) :: magnitude(*), exponent(*)
     end function pmc_gpu_set_weights

     integer(c_int) function pmc_gpu_set_rng(h, seed, step) &
          bind(c, name="pmc_gpu_set_rng")
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: h
       integer(c_int64_t), value :: seed
       integer(c_int32_t), value :: step
     end function pmc_gpu_set_rng

     integer(c_int) function pmc_gpu_set_next_id(h, box, next_id, stride, &
          offset) bind(c, name="pmc_gpu_set_next_id")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: h
       integer(c_int), value :: box
       integer(c_int64_t), value :: next_id, stride, offset
     end function pmc_gpu_set_next_id

     integer(c_int) function pmc_gpu_get_next_id(h, box, next_id) &
          bind(c, name="pmc_gpu_get_next_id")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: h
       integer(c_int), value :: box
       integer(c_int64_t), intent(out) :: next_id
     end function pmc_gpu_get_next_id

     integer(c_int) function pmc_gpu_upload(h, box, n_part, capacity, vol, id, &
          group, klass, n_primary, n_comp, comp_source, comp_time, least, &
          greatest) bind(c, name="pmc_gpu_upload")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: h
       integer(c_int), value :: box, n_part, capacity
       real(c_double), intent(in) :: vol(*)
       integer(c_int64_t), intent(in) :: id(*)
       integer(c_int), intent(in) :: group(*), klass(*), n_primary(*)
       type(c_ptr), value :: n_comp, comp_source, comp_time, least, greatest
     end function pmc_gpu_upload

     integer(c_int) function pmc_gpu_set_bin_grid(h, box, n_bin, r_min, r_max) &
          bind(c, name="pmc_gpu_set_bin_grid")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: box, n_bin
       real(c_double), value :: r_min, r_max
     end function pmc_gpu_set_bin_grid

     integer(c_int) function pmc_gpu_get_n_part(h, box, n_part) &
          bind(c, name="pmc_gpu_get_n_part")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: box
       integer(c_int), intent(out) :: n_part
     end function pmc_gpu_get_n_part

     integer(c_int) function pmc_gpu_download(h, box, max_part, n_part, vol, &
          id, group, klass, n_primary, n_comp, comp_source, comp_time, least, &
          greatest) bind(c, name="pmc_gpu_download")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: h
       integer(c_int), value :: box, max_part
       integer(c_int), intent(out) :: n_part
       real(c_double), intent(out) :: vol(*)
       integer(c_int64_t), intent(out) :: id(*)
       integer(c_int), intent(out) :: group(*), klass(*), n_primary(*)
       type(c_ptr), value :: n_comp, comp_source, comp_time, least, greatest
     end function pmc_gpu_download

     integer(c_int) function pmc_gpu_set_mode(h, mode, precision, band) &
          bind(c, name="pmc_gpu_set_mode")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: mode, precision
       real(c_double), value :: band
     end function pmc_gpu_set_mode

     integer(c_int) function pmc_gpu_mc_coag(h, kernel_type, del_t, n_samp, &
          n_coag) bind(c, name="pmc_gpu_mc_coag")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: h
       integer(c_int), value :: kernel_type
       real(c_double), value :: del_t
       integer(c_int64_t), intent(out) :: n_samp(*), n_coag(*)
     end function pmc_gpu_mc_coag

     integer(c_int) function pmc_gpu_set_n_part_ideal(h, box, n_part) &
          bind(c, name="pmc_gpu_set_n_part_ideal")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: box
       real(c_double), value :: n_part
     end function pmc_gpu_set_n_part_ideal

     integer(c_int) function pmc_gpu_rebalance(h, allow_doubling, allow_halving) &
          bind(c, name="pmc_gpu_rebalance")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: allow_doubling, allow_halving
     end function pmc_gpu_rebalance

     integer(c_int) function pmc_gpu_get_weights(h, box, max_n, n_group, &
          n_class, type, magnitude, exponent) bind(c, name="pmc_gpu_get_weights")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: box, max_n
       integer(c_int), intent(out) :: n_group, n_class, type(*)
       real(c_double), intent(out) :: magnitude(*), exponent(*)
     end function pmc_gpu_get_weights

     integer(c_int) function pmc_gpu_size_dist(h, box, quantity, n_bin, d_min, &
          d_max, hist) bind(c, name="pmc_gpu_size_dist")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: box, quantity, n_bin
       real(c_double), value :: d_min, d_max
       real(c_double), intent(out) :: hist(*)
     end function pmc_gpu_size_dist

     integer(c_int) function pmc_gpu_dilute(h, sample_prob) &
          bind(c, name="pmc_gpu_dilute")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: sample_prob(*)
     end function pmc_gpu_dilute

     integer(c_int) function pmc_gpu_particle_loss(h, loss_function, del_t, &
          height) bind(c, name="pmc_gpu_particle_loss")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: loss_function
       real(c_double), value :: del_t
       real(c_double), intent(in) :: height(*)
     end function pmc_gpu_particle_loss

     integer(c_int) function pmc_gpu_scale_weights(h, box, factor) &
          bind(c, name="pmc_gpu_scale_weights")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: box
       real(c_double), value :: factor
     end function pmc_gpu_scale_weights

     integer(c_int) function pmc_gpu_append(h, box, n, vol, id, group, klass, &
          n_primary, n_comp, comp_source, comp_time, least, greatest) &
          bind(c, name="pmc_gpu_append")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: box, n
       real(c_double), intent(in) :: vol(*)
       type(c_ptr), value :: id
       integer(c_int), intent(in) :: group(*), klass(*)
       type(c_ptr), value :: n_primary, n_comp, comp_source, comp_time, least, greatest
     end function pmc_gpu_append

     integer(c_int) function pmc_gpu_clear_removed(h, box) &
          bind(c, name="pmc_gpu_clear_removed")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: box
     end function pmc_gpu_clear_removed
  end interface

contains

  !> Stop with the library's message if a call failed (cf. die_msg).
  subroutine coag_gpu_check(code)
    integer(c_int), intent(in) :: code
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (code == 0) return
    call c_f_pointer(pmc_gpu_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    write(0, '(a,i0,a,a)') 'ERROR (pmc_coag_gpu-', code, '): ', trim(text)
    stop 3
  end subroutine coag_gpu_check

  !> One coagulation timestep of a single-box handle, with mc_coag's
  !> argument meaning (src/coagulation.F90:43-59): kernel type code, del_t,
  !> and the two counters out.
  subroutine coag_gpu_mc_coag(h, coag_kernel_type, del_t, tot_n_samp, &
       tot_n_coag)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: coag_kernel_type
    real(kind=c_double), intent(in) :: del_t
    integer, intent(out) :: tot_n_samp, tot_n_coag
    integer(c_int64_t) :: ns(1), nc(1)
    call coag_gpu_check(pmc_gpu_mc_coag(h, int(coag_kernel_type, c_int), &
         del_t, ns, nc))
    tot_n_samp = int(ns(1))
    tot_n_coag = int(nc(1))
  end subroutine coag_gpu_mc_coag

end module pmc_coag_gpu

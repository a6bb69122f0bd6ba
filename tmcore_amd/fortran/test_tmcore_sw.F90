! Namelist-driven shallow-water driver, the counterpart of the reference's
! test_tmcore_swm program (reference src/core/test_tmcore_swm.F90:1-24).
! Usage: test_tmcore_sw [namelist]   (default ./namelist.tmcore_sw)
program test_tmcore_sw

  use tm_swm_gpu_mod

  implicit none

  character(256) :: path
  integer :: start_clock, end_clock, rate

  call system_clock(start_clock, rate)
  if (command_argument_count() >= 1) then
    call get_command_argument(1, path)
  else
    path = './namelist.tmcore_sw'
  end if

  call tm_swm_init(trim(path))
  call tm_swm_run()
  call tm_swm_final()

  call system_clock(end_clock)
  write(*, '(A,F10.3,A)') '[Notice]: wall time ', real(end_clock - start_clock) / real(rate), ' s'

end program test_tmcore_sw

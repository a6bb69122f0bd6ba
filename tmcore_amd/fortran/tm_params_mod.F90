! Namelist parameters of the test_tmcore_sw driver.  The four groups
! time_control, io_control, tmcore and test_case keep the names and meaning of
! the reference's params_mod (reference src/core/params_mod.F90:15-65).  The
! group gpu_control is new: the mesh is generated (the reference reads an
! MPAS grid file, mesh_mod.F90:79-167), and the device options and seeded
! initial state are chosen here.
module tm_params_mod

  use iso_c_binding, only: c_int64_t

  implicit none

  integer, parameter :: r8 = 8
  integer, parameter :: max_file_path_len = 256
  integer, parameter :: max_name_len = 30

  ! const_mod.F90:11-13
  real(r8), parameter :: radius = 6371220d0
  real(r8), parameter :: g      = 9.80616d0
  real(r8), parameter :: omega  = 7.292d-5

  integer :: run_days = 0
  integer :: run_hours = 0
  integer :: run_minutes = 0
  integer :: run_seconds = 0
  integer :: start_time(5) = [0, 0, 0, 0, 0]
  integer :: end_time  (5) = [0, 0, 0, 0, 0]
  real(r8) :: dt = 600

  namelist /time_control/ run_days, run_hours, run_minutes, run_seconds, start_time, end_time, dt

  character(max_name_len) :: history_interval(1) = 'N/A'
  character(max_name_len) :: time_units = 'days'
  character(max_file_path_len) :: output_file_prefix = 'N/A'
  character(max_name_len) :: frames_per_file = 'N/A'

  namelist /io_control/ history_interval, time_units, output_file_prefix, frames_per_file

  integer :: time_scheme = 2
  integer :: energy_scheme = 1
  integer :: pv_scheme = 3
  logical :: conserve_energy = .true.
  logical :: flux_4th_order_correct = .false.
  real(r8) :: apvm_weight = 0.5
  real(r8) :: clust_weight = 0.25
  integer :: adv_order = 2
  logical :: adv_mono = .false.
  integer :: interp_pv_order = 3
  character(max_file_path_len) :: mesh_file_path = 'N/A'

  namelist /tmcore/ time_scheme, energy_scheme, pv_scheme, conserve_energy, flux_4th_order_correct, &
                    apvm_weight, clust_weight, adv_order, adv_mono, interp_pv_order, mesh_file_path

  character(max_name_len) :: case_name = 'seeded'
  character(max_name_len) :: subcase_name = 'N/A'

  namelist /test_case/ case_name, subcase_name

  integer :: mesh_level = 5          ! icosahedral bisection level (10*4**L+2 cells)
  integer :: precision = 64          ! 64: plain FP64; 32: FP32 state with FP64 geometry/weights
  integer :: ntracers = 0            ! passive tracers advected with the stage winds
  integer(c_int64_t) :: seed = 1     ! seeded initial state
  integer(c_int64_t) :: tracer_seed0 = 101  ! tracer k uses tracer_seed0 + k - 1
  integer :: device = 0
  integer :: renumber = 1
  integer :: print_every = 1         ! steps between diagnostic lines
  character(max_file_path_len) :: closure_file = 'N/A'  ! trained closure: raw little-endian float32
                                     ! [38530 parameters, 40 input shifts, 40 input scales, 2 output scales]

  integer :: closure_interval = 1    ! steps between MLP evaluations
  integer :: closure_lumped = 0      ! 0: correction after every step; 1: once per interval
  real(8) :: closure_gain = 1.0d0    ! factor on the applied correction
  logical :: use_gpu = .true.        ! .false.: the steps run on the host CPU in the Fortran FP64 CPU path
                                     ! (tm_swm_cpu_mod: CRK4 with beta, pv_scheme 3 at interp_pv_order
                                     ! 2-4, adv_order 2, precision 64, no closure, no history)

  namelist /gpu_control/ mesh_level, precision, ntracers, seed, tracer_seed0, device, renumber, print_every, &
                         closure_file, closure_interval, closure_lumped, &
                         closure_gain, use_gpu

contains

  ! params_parse_namelist (params_mod.F90:69-80) plus the gpu_control group.
  subroutine params_parse_namelist(file_path)

    character(*), intent(in) :: file_path
    integer :: ios

    open(10, file=file_path, status='old', iostat=ios)
    if (ios /= 0) then
      write(*, '(A)') '[Error]: cannot open namelist file ' // trim(file_path)
      stop 1
    end if
    read(10, nml=time_control)
    rewind(10)
    read(10, nml=io_control, iostat=ios)
    rewind(10)
    read(10, nml=tmcore)
    rewind(10)
    read(10, nml=test_case, iostat=ios)
    rewind(10)
    read(10, nml=gpu_control, iostat=ios)
    close(10)

  end subroutine params_parse_namelist

  ! Number of steps of the run: the reference loops until curr_time >= end_time
  ! (time_mod.F90:138-142), i.e. ceiling(run length / dt).
  integer function run_steps() result(n)

    real(r8) :: seconds

    seconds = run_days * 86400.0_r8 + run_hours * 3600.0_r8 + run_minutes * 60.0_r8 + run_seconds
    n = ceiling(seconds / dt - 1.0e-9_r8)

  end function run_steps

end module tm_params_mod

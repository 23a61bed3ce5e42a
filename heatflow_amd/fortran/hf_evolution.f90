!! hf_evolution.f90 -- drop-in replacement of HeatFlow's module `evolution`
!! (src/heatflow/mod_evolve.f90) that runs the time step on an MI355X through the C-ABI of
!! libheatflow_hip.so (include/heatflow_hip.h).
!!
!! Same module name, same single public procedure `simulate(itime)` with the same kind and
!! intent (mod_evolve.f90:36, 45-46).  Build it in place of mod_evolve.f90 together with the
!! other HeatFlow modules and link libheatflow_hip.so; the driver (src/heatflow.f90), the input
!! files, setup, material handling and output stay as they are.
!!
!! On the first call the module gathers the per-cell fields of grid(:,:,:) into
!! structure-of-arrays, creates a device context, uploads and assembles H on the device.  Each
!! call runs one step on the device.  The state stays there: it comes back to Temp_p/Temp_pp
!! only on the steps whose output reads it -- data_write's log records and steady-state check
!! (src/heatflow/mod_output.f90:69-111, 117-131) and final_print at the last step or, with
!! IVERB >= 3, at every step (heatflow.f90:86-87, mod_output.f90:139-142, 207-238) -- so those
!! see the same arrays as with the CPU solver, and the other steps copy nothing.  The host
!! history goes to the device on the first call and at step 1 (after initial_evolve, which may
!! read a restart); the driver does not change it between steps otherwise.  With TempDepProp = 1
!! the reference rebuilds H after every step (ChangeProp, mod_tempdep.f90:38-65, 64); here H is
!! re-assembled on the device from the fields already there.  heated_volume is set as heater
!! sets it (mod_heating.f90:137-141).  The reference's fatal NaN guards are reproduced: the
!! message of mod_evolve.f90:70-158 and `stop 1`.
module evolution
  use iso_c_binding
  use constants, only: real12, int12
  use inputs, only: NA, nx, ny, nz, icattaneo, isteady, time_step, grid, power_in, freq, &
       heated_steps, kappaBoundx1, kappaBoundy1, kappaBoundz1, kappaBoundNx, kappaBoundNy, &
       kappaBoundNz, T_Bathx1, T_Bathx2, T_Bathy1, T_Bathy2, T_Bathz1, T_Bathz2, T_Bath, &
       T_BathCG, BR, Periodicx, Periodicy, Periodicz, CG_x_m, CG_x_p, CG_y_m, CG_y_p, &
       CG_z_m, CG_z_p, TempDepProp, ntime, write_every, WriteToTxt, Test_Run, &
       Check_Steady_State, IVERB
  use globe_data, only: Temp_p, Temp_pp, heated_volume
  implicit none
  private
  public :: simulate

  !! struct hf_params (include/heatflow_hip.h)
  type, bind(C) :: hf_params_t
     integer(c_int64_t) :: nx, ny, nz
     real(c_double) :: time_step
     integer(c_int32_t) :: icattaneo, isteady
     integer(c_int32_t) :: periodic(3)
     integer(c_int32_t) :: cg(6)
     real(c_double) :: kappa_bound(6), t_bath(6)
     real(c_double) :: br, t_bath_cg, power_in, freq, t_bath_rad
     integer(c_int64_t) :: heated_steps
     real(c_double) :: tol
     integer(c_int64_t) :: itmax
     integer(c_int32_t) :: precision, inner_itmax
     real(c_double) :: inner_tol
  end type hf_params_t

  interface
     subroutine hf_params_default(p) bind(C, name="hf_params_default")
       import :: hf_params_t
       type(hf_params_t), intent(out) :: p
     end subroutine
     integer(c_int) function hf_create(ctx, p) bind(C, name="hf_create")
       import :: c_ptr, c_int, hf_params_t
       type(c_ptr), intent(out) :: ctx
       type(hf_params_t), intent(in) :: p
     end function
     integer(c_int) function hf_upload_cells(ctx, kappa, rho, hc, tau, em, lx, ly, lz, vol, &
          iheater, qextra) bind(C, name="hf_upload_cells")
       import :: c_ptr, c_int, c_double, c_int64_t
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: kappa(*), rho(*), hc(*), tau(*), em(*), lx(*), ly(*), &
            lz(*), vol(*)
       integer(c_int64_t), intent(in) :: iheater(*)
       type(c_ptr), value :: qextra
     end function
     integer(c_int) function hf_assemble(ctx) bind(C, name="hf_assemble")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function hf_set_state(ctx, tp, tpp) bind(C, name="hf_set_state")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: tp(*), tpp(*)
     end function
     integer(c_int) function hf_get_state(ctx, tp, tpp) bind(C, name="hf_get_state")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: tp(*)
       real(c_double), intent(out), optional :: tpp(*)     ! absent: not copied
     end function
     integer(c_int) function hf_step(ctx, itime, iter, err) bind(C, name="hf_step")
       import :: c_ptr, c_int, c_int64_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: itime
       integer(c_int64_t), intent(out) :: iter
       real(c_double), intent(out) :: err
     end function
     type(c_ptr) function hf_last_error(ctx) bind(C, name="hf_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function
  end interface

  type(c_ptr), save :: ctx = c_null_ptr
  !! the device holds Temp_p / Temp_pp as the host last received them
  logical, save :: device_current = .false.
  !! heater's heated_volume: the volume of the cells with iheater > 0, in loop order
  real(real12), save :: volume_heated = 0.0_real12

contains

  !! evolution::simulate(itime) (src/heatflow/mod_evolve.f90:45-177)
  subroutine simulate(itime)
    integer(int12), intent(in) :: itime
    integer(c_int) :: rc
    integer(c_int64_t) :: iter
    real(c_double) :: err

    if (.not. c_associated(ctx)) then
       call device_setup()
       device_current = .false.
    end if
    ! The driver changes the history only through initial_evolve, before step 1
    ! (heatflow.f90:79-80, mod_init_evolve.f90:34-51); otherwise the device copy is current.
    if (itime == 1) device_current = .false.
    if (.not. device_current) then
       rc = hf_set_state(ctx, Temp_p, Temp_pp)
       if (rc /= 0) call fail(rc)
    end if
    rc = hf_step(ctx, int(itime, c_int64_t), iter, err)
    if (rc /= 0) call fail(rc)
    device_current = .true.
    heated_volume = volume_heated
    ! Temp_pp = Temp_p; Temp_p = x (mod_evolve.f90:169-170), on the host when output reads them
    if (host_reads_state(itime)) then
       rc = hf_get_state(ctx, Temp_p, Temp_pp)
       if (rc /= 0) call fail(rc)
    end if
    ! ChangeProp (mod_tempdep.f90:38-65) rebuilds H from the grid after every step
    ! (mod_evolve.f90:172-174)
    if (TempDepProp .eq. 1) then
       rc = hf_assemble(ctx)
       if (rc /= 0) call fail(rc)
    end if
  end subroutine simulate

  !! whether the driver reads Temp_p / Temp_pp after step itime: data_write builds its log
  !! record from Temp_pp at step 1 and every write_every steps when writing to text, and
  !! compares it in the steady-state check (mod_output.f90:69-131); final_print writes both at
  !! the last step, and at every step with IVERB >= 3 (heatflow.f90:86-87)
  logical function host_reads_state(itime)
    integer(int12), intent(in) :: itime
    host_reads_state = itime == ntime .or. IVERB >= 3 .or. &
         (Check_Steady_State .and. itime > 1)
    if (.not. Test_Run) then
       if (itime == 1) host_reads_state = .true.
       if (WriteToTxt .and. write_every > 0) then
          if (mod(itime, write_every) == 0) host_reads_state = .true.
       end if
    end if
  end function host_reads_state

  subroutine device_setup()
    type(hf_params_t) :: p
    integer(c_int) :: rc
    call hf_params_default(p)       ! tol = 1e-32, itmax = 50000 as mod_evolve.f90:142-144
    call solver_from_environment(p)
    p%nx = nx; p%ny = ny; p%nz = nz
    p%time_step = time_step
    p%icattaneo = int(icattaneo, c_int32_t)
    p%isteady = int(isteady, c_int32_t)
    p%periodic = [merge(1, 0, Periodicx), merge(1, 0, Periodicy), merge(1, 0, Periodicz)]
    p%cg = [merge(1, 0, CG_x_m), merge(1, 0, CG_x_p), merge(1, 0, CG_y_m), &
         merge(1, 0, CG_y_p), merge(1, 0, CG_z_m), merge(1, 0, CG_z_p)]
    p%kappa_bound = [kappaBoundx1, kappaBoundNx, kappaBoundy1, kappaBoundNy, kappaBoundz1, &
         kappaBoundNz]
    p%t_bath = [T_Bathx1, T_Bathx2, T_Bathy1, T_Bathy2, T_Bathz1, T_Bathz2]
    p%br = BR
    p%t_bath_cg = T_BathCG
    p%power_in = power_in
    p%freq = freq
    p%t_bath_rad = T_Bath
    p%heated_steps = heated_steps
    rc = hf_create(ctx, p)
    if (rc /= 0) call fail(rc)
    call upload_cells()
  end subroutine device_setup

  !! Optional solver controls from the environment; unset, the reference's stand (FP64, tol
  !! 1e-32, hard-coded in mod_evolve.f90:142-144).  HF_PRECISION = fp64 | fp32_ir (FP32 inner CG
  !! inside FP64 refinement), HF_TOL = the scaled-residual tolerance.  Read once, at set-up.
  subroutine solver_from_environment(p)
    type(hf_params_t), intent(inout) :: p
    character(len=64) :: v
    integer :: n, st
    real(c_double) :: t
    call get_environment_variable("HF_PRECISION", v, length=n, status=st)
    if (st == 0 .and. n > 0) then
       select case (trim(v))
       case ("fp64")
          p%precision = 0_c_int32_t
       case ("fp32_ir")
          p%precision = 1_c_int32_t
       case default
          write(0,*) "HF_PRECISION must be fp64 or fp32_ir, not ", trim(v)
          stop 1
       end select
    end if
    call get_environment_variable("HF_TOL", v, length=n, status=st)
    if (st == 0 .and. n > 0) then
       read(v, *, iostat=st) t
       if (st /= 0 .or. .not. (t > 0.0_c_double)) then
          write(0,*) "HF_TOL must be a positive number, not ", trim(v)
          stop 1
       end if
       p%tol = t
    end if
  end subroutine solver_from_environment

  !! grid(:,:,:) (mod_constructions.f90:25-31) as structure-of-arrays, x fastest
  subroutine upload_cells()
    real(c_double), allocatable :: kappa(:), rho(:), hc(:), tau(:), em(:), lx(:), ly(:), lz(:), &
         vol(:)
    integer(c_int64_t), allocatable :: ih(:)
    integer(c_int) :: rc
    integer(int12) :: i
    kappa = reshape(grid%kappa, [NA])
    rho = reshape(grid%rho, [NA])
    hc = reshape(grid%heat_capacity, [NA])
    tau = reshape(grid%tau, [NA])
    em = reshape(grid%em, [NA])
    lx = reshape(grid%Length(1), [NA])
    ly = reshape(grid%Length(2), [NA])
    lz = reshape(grid%Length(3), [NA])
    vol = reshape(grid%volume, [NA])
    ih = reshape(grid%iheater, [NA])
    rc = hf_upload_cells(ctx, kappa, rho, hc, tau, em, lx, ly, lz, vol, ih, c_null_ptr)
    if (rc /= 0) call fail(rc)
    rc = hf_assemble(ctx)
    if (rc /= 0) call fail(rc)
    volume_heated = 0.0_real12
    do i = 1, NA
       if (ih(i) > 0) volume_heated = volume_heated + vol(i)
    end do
  end subroutine upload_cells

  !! the reference's fatal guards (mod_evolve.f90:70-158): message, then stop 1
  subroutine fail(rc)
    integer(c_int), intent(in) :: rc
    select case (rc)
    case (1)
       write(0,*) "fatal error: NAN in B vector"
    case (2)
       write(0,*) "fatal error: NAN in Q vector"
    case (3)
       write(0,*) "fatal error: NAN in Qdens vector"
    case (4)
       write(*,*) "fatal error: NAN in S_CAT vector"
    case (5)
       write(0,*) "fatal error: NAN in S vector"
    case (6)
       write(0,*) "fatal error: NAN in x tempurature vector"
    case default
       write(0,*) "fatal error: heatflow_hip error ", rc, ": ", trim(c_message(hf_last_error(ctx)))
    end select
    stop 1
  end subroutine fail

  function c_message(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=512) :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: i
    s = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, chars, [512])
    do i = 1, 512
       if (chars(i) == c_null_char) exit
       s(i:i) = chars(i)
    end do
  end function c_message

end module evolution

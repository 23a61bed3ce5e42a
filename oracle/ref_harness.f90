!! ref_harness.f90 -- C-callable driver over the reference HeatFlow modules (TEST INFRASTRUCTURE).
!!
!! It is compiled together with the unmodified reference sources (read in place from the
!! HeatFlow checkout) into oracle/_ref/libhfref.so by oracle/Makefile.  Like the reference's own
!! unit tests (test/test_mod_evolve.f90:48-101) it sets the module globals of `inputs` and
!! `globe_data` directly, so per-cell material fields can be given without the material table.
!! The per-cell setup below restates the non-material part of set_global_variables
!! (src/heatflow/mod_setup.f90:38-45, 59-64).  Everything on the time-step path -- sparse_Hmatrix,
!! boundary, heater, S_catS, linbcg, simulate -- is the reference's own code.
!! Used only by tests and by the CPU-baseline leg of the benchmark.
module hf_ref_harness
  use iso_c_binding
  use constants, only: real12, int12
  use inputs
  use globe_data, only: Temp_cur, Temp_p, Temp_pp, Q_P, inverse_time, heat, lin_rhoc, ra
  use setup, only: sparse_Hmatrix
  use evolution, only: simulate
  use boundary_vector, only: boundary
  use heating, only: heater
  use cattaneo, only: S_catS
  use solver, only: linbcg
  use materials, only: ref_material => material
  implicit none

contains

  subroutine hfr_reset() bind(C, name="hfr_reset")
    if (allocated(grid)) deallocate(grid)
    if (allocated(Temp_cur)) deallocate(Temp_cur)
    if (allocated(Temp_p)) deallocate(Temp_p)
    if (allocated(Temp_pp)) deallocate(Temp_pp)
    if (allocated(lin_rhoc)) deallocate(lin_rhoc)
    if (allocated(Q_P)) deallocate(Q_P)
    if (associated(ra%val)) then
       deallocate(ra%val, ra%irow, ra%jcol)
       nullify(ra%val, ra%irow, ra%jcol)
    end if
  end subroutine hfr_reset

  !! iparam: [nx, ny, nz, icattaneo, isteady, px, py, pz, cg_xm, cg_xp, cg_ym, cg_yp, cg_zm, cg_zp,
  !!          heated_steps]
  !! dparam: [time_step, kb_x1, kb_Nx, kb_y1, kb_Ny, kb_z1, kb_Nz, Tb_x1, Tb_x2, Tb_y1, Tb_y2,
  !!          Tb_z1, Tb_z2, BR, T_BathCG, power_in, freq, T_Bath]
  !! Per-cell arrays are x-fastest; tau is the physical relaxation time (scaled here by 1/dt^2).
  subroutine hfr_setup(iparam, dparam, kappa_c, rho_c, hc_c, tau_c, em_c, lx_c, ly_c, lz_c, &
       iheater_c) bind(C, name="hfr_setup")
    integer(c_int64_t), intent(in) :: iparam(15)
    real(c_double), intent(in) :: dparam(18)
    real(c_double), intent(in) :: kappa_c(*), rho_c(*), hc_c(*), tau_c(*), em_c(*)
    real(c_double), intent(in) :: lx_c(*), ly_c(*), lz_c(*)
    integer(c_int64_t), intent(in) :: iheater_c(*)
    integer(int12) :: ix, iy, iz, idx

    call hfr_reset()
    nx = iparam(1); ny = iparam(2); nz = iparam(3)
    NA = nx*ny*nz
    icattaneo = iparam(4); isteady = iparam(5)
    Periodicx = iparam(6) /= 0; Periodicy = iparam(7) /= 0; Periodicz = iparam(8) /= 0
    CG_x_m = iparam(9) /= 0; CG_x_p = iparam(10) /= 0
    CG_y_m = iparam(11) /= 0; CG_y_p = iparam(12) /= 0
    CG_z_m = iparam(13) /= 0; CG_z_p = iparam(14) /= 0
    heated_steps = iparam(15)
    time_step = dparam(1)
    kappaBoundx1 = dparam(2); kappaBoundNx = dparam(3)
    kappaBoundy1 = dparam(4); kappaBoundNy = dparam(5)
    kappaBoundz1 = dparam(6); kappaBoundNz = dparam(7)
    T_Bathx1 = dparam(8); T_Bathx2 = dparam(9)
    T_Bathy1 = dparam(10); T_Bathy2 = dparam(11)
    T_Bathz1 = dparam(12); T_Bathz2 = dparam(13)
    BR = dparam(14); T_BathCG = dparam(15)
    power_in = dparam(16); freq = dparam(17); T_Bath = dparam(18)
    IVERB = 0
    TempDepProp = 0

    allocate(grid(nx, ny, nz))
    allocate(Temp_cur(nx, ny, nz), Temp_p(NA), Temp_pp(NA), lin_rhoc(NA), Q_P(NA))
    Q_P(:) = 0.0_real12
    heat = 0.0_real12
    inverse_time = 1.0_real12/time_step
    idx = 0
    do iz = 1, nz
       do iy = 1, ny
          do ix = 1, nx
             idx = idx + 1
             grid(ix,iy,iz)%imaterial_type = 1
             grid(ix,iy,iz)%Length(1) = lx_c(idx)
             grid(ix,iy,iz)%Length(2) = ly_c(idx)
             grid(ix,iy,iz)%Length(3) = lz_c(idx)
             grid(ix,iy,iz)%volume = lx_c(idx)*ly_c(idx)*lz_c(idx)
             grid(ix,iy,iz)%iheater = iheater_c(idx)
             grid(ix,iy,iz)%kappa = kappa_c(idx)
             grid(ix,iy,iz)%rho = rho_c(idx)
             grid(ix,iy,iz)%heat_capacity = hc_c(idx)
             grid(ix,iy,iz)%tau = tau_c(idx)*inverse_time*inverse_time
             grid(ix,iy,iz)%em = em_c(idx)
             lin_rhoc(idx) = rho_c(idx)*hc_c(idx)
          end do
       end do
    end do
    call sparse_Hmatrix()
  end subroutine hfr_setup

  subroutine hfr_coo_sizes(n, len) bind(C, name="hfr_coo_sizes")
    integer(c_int64_t), intent(out) :: n, len
    n = ra%n
    len = ra%len
  end subroutine hfr_coo_sizes

  subroutine hfr_coo(val, irow, jcol) bind(C, name="hfr_coo")
    real(c_double), intent(out) :: val(*)
    integer(c_int64_t), intent(out) :: irow(*), jcol(*)
    val(1:ra%len) = ra%val(1:ra%len)
    irow(1:ra%len) = ra%irow(1:ra%len)
    jcol(1:ra%len) = ra%jcol(1:ra%len)
  end subroutine hfr_coo

  subroutine hfr_set_state(tp, tpp) bind(C, name="hfr_set_state")
    real(c_double), intent(in) :: tp(*), tpp(*)
    Temp_p(1:NA) = tp(1:NA)
    Temp_pp(1:NA) = tpp(1:NA)
  end subroutine hfr_set_state

  subroutine hfr_get_state(tp, tpp) bind(C, name="hfr_get_state")
    real(c_double), intent(out) :: tp(*), tpp(*)
    tp(1:NA) = Temp_p(1:NA)
    tpp(1:NA) = Temp_pp(1:NA)
  end subroutine hfr_get_state

  !! the reference time step, unchanged (src/heatflow/mod_evolve.f90:45-177)
  subroutine hfr_simulate(itime) bind(C, name="hfr_simulate")
    integer(c_int64_t), value :: itime
    integer(int12) :: it
    it = itime
    call simulate(it)
  end subroutine hfr_simulate

  subroutine hfr_boundary(b) bind(C, name="hfr_boundary")
    real(c_double), intent(out) :: b(*)
    real(real12), allocatable :: bb(:)
    allocate(bb(NA))
    bb = 0.0_real12
    call boundary(bb)
    b(1:NA) = bb
  end subroutine hfr_boundary

  subroutine hfr_heater(itime, q, qdens, heated_vol) bind(C, name="hfr_heater")
    integer(c_int64_t), value :: itime
    real(c_double), intent(out) :: q(*), qdens(*)
    real(c_double), intent(out) :: heated_vol
    real(real12), allocatable :: qq(:), qd(:)
    integer(int12) :: it
    allocate(qq(NA), qd(NA))
    qq = 0.0_real12
    qd = 0.0_real12
    it = itime
    call heater(it, qq, qd)
    q(1:NA) = qq
    qdens(1:NA) = qd
    heated_vol = heated_volume_value()
  end subroutine hfr_heater

  function heated_volume_value() result(v)
    use globe_data, only: heated_volume
    real(real12) :: v
    v = heated_volume
  end function heated_volume_value

  subroutine hfr_scat(sc) bind(C, name="hfr_scat")
    real(c_double), intent(out) :: sc(*)
    real(real12), allocatable :: s(:)
    allocate(s(NA))
    s = 0.0_real12
    call S_catS(s)
    sc(1:NA) = s
  end subroutine hfr_scat

  !! linbcg on the assembled reference matrix, with caller-chosen tolerance
  subroutine hfr_linbcg(b, x, itol, tol, itmax, iter, err) bind(C, name="hfr_linbcg")
    real(c_double), intent(in) :: b(*)
    real(c_double), intent(inout) :: x(*)
    integer(c_int64_t), value :: itol, itmax
    real(c_double), value :: tol
    integer(c_int64_t), intent(out) :: iter
    real(c_double), intent(out) :: err
    real(real12), allocatable :: bb(:), xx(:)
    integer(int12) :: it
    allocate(bb(NA), xx(NA))
    bb = b(1:NA)
    xx = x(1:NA)
    it = 0
    call linbcg(bb, xx, itol, tol, itmax, it, err)
    x(1:NA) = xx
    iter = it
  end subroutine hfr_linbcg

  !! One time step composed from the reference's own pieces (boundary, heater, S_catS, linbcg)
  !! in simulate's order (src/heatflow/mod_evolve.f90:55-170), with one superset input the
  !! reference cannot express: a per-cell volumetric power density added to Qdens.  Used for
  !! workloads with non-uniform heating.  Returns linbcg's iteration count.
  subroutine hfr_simulate_qextra(itime, qextra, tol, iter) bind(C, name="hfr_simulate_qextra")
    use constants, only: TINY
    integer(c_int64_t), value :: itime
    real(c_double), intent(in) :: qextra(*)
    real(c_double), value :: tol
    integer(c_int64_t), intent(out) :: iter
    real(real12), allocatable :: S(:), x(:), Q(:), Qdens(:), S_CAT(:), B(:)
    integer(int12) :: it, itol, itmax
    real(real12) :: err
    allocate(S(NA), x(NA), Q(NA), Qdens(NA), S_CAT(NA), B(NA))
    B = 0.0_real12; Q = 0.0_real12; Qdens = 0.0_real12; S_CAT = 0.0_real12; S = 0.0_real12
    call boundary(B)
    if (any(grid%iheater .gt. 0)) call heater(itime, Q, Qdens)
    Qdens = Qdens + qextra(1:NA)
    if (icattaneo .eq. 1) call S_catS(S_CAT)
    if (isteady .eq. 0) then
       S = - inverse_time * Temp_p * lin_rhoc - Qdens - B
       if (icattaneo .eq. 1) S = S + S_CAT
    else
       S = -Qdens - B
    end if
    x = Temp_p + (Temp_p - Temp_pp)
    if (any(x - Temp_p .lt. TINY)) x = x + TINY
    itol = 1
    itmax = 50000
    it = 0
    call linbcg(S, x, itol, tol, itmax, it, err)
    iter = it
    Temp_pp = Temp_p
    Temp_p = x
  end subroutine hfr_simulate_qextra

  !! The reference's material() (src/heatflow/mod_material.f90:54-267) over n cells in order,
  !! as set_global_variables calls it (src/heatflow/mod_setup.f90:53-68): the output locals
  !! persist from cell to cell, so a built-in material that does not assign tau or em keeps the
  !! previous cell's value.  The locals start at 0 here (uninitialised in the reference).
  !! File materials (index <= 99) come from the nfile entries given (input_materials, read
  !! from mat.in in the reference).  Every id must be known (material() calls exit otherwise).
  subroutine hfr_materials(n, ids, nfile, findex, fkappa, frho, fhc, ftau, fem, &
       kappa_o, rho_o, hc_o, tau_o, em_o) bind(C, name="hfr_materials")
    integer(c_int64_t), value :: n, nfile
    integer(c_int64_t), intent(in) :: ids(*), findex(*)
    real(c_double), intent(in) :: fkappa(*), frho(*), fhc(*), ftau(*), fem(*)
    real(c_double), intent(out) :: kappa_o(*), rho_o(*), hc_o(*), tau_o(*), em_o(*)
    real(real12) :: kappa, kappa3D, h_conv, heat_capacity, rho, sound_speed, tau, em
    integer(int12) :: i
    if (allocated(input_materials)) deallocate(input_materials)
    allocate(input_materials(max(nfile, 1_c_int64_t)))
    do i = 1, nfile
       input_materials(i)%index = findex(i)
       input_materials(i)%heat_capacity = fhc(i)
       input_materials(i)%h_conv = 0.0_real12
       input_materials(i)%kappa = fkappa(i)
       input_materials(i)%kappa3D = fkappa(i)
       input_materials(i)%rho = frho(i)
       input_materials(i)%sound_speed = 0.0_real12
       input_materials(i)%tau = ftau(i)
       input_materials(i)%em = fem(i)
    end do
    if (nfile == 0) input_materials(1)%index = -1
    kappa = 0; kappa3D = 0; h_conv = 0; heat_capacity = 0; rho = 0; sound_speed = 0
    tau = 0; em = 0
    do i = 1, n
       call ref_material(ids(i), kappa, kappa3D, h_conv, heat_capacity, rho, sound_speed, tau, em)
       kappa_o(i) = kappa
       rho_o(i) = rho
       hc_o(i) = heat_capacity
       tau_o(i) = tau
       em_o(i) = em
    end do
  end subroutine hfr_materials

end module hf_ref_harness

!------------------------------------------------------------------------------
! elas3d_gpu: the main program of src/elas3dxtal_pcg.f90 (program elas3d,
! f90:186-552) with its solver routines replaced by the gfx950 library through
! the C ABI (module e3x_bindings).
!
! The reference's compile-time parameters (f90:77-148) are run-time inputs here,
! read from a namelist file (first command-line argument, default 'elas3d.nml'),
! with the reference's defaults.  Inputs and outputs keep the reference's files:
!   input_structure_poly.h5     (/pix, /orientation)           read  (f90:2879)
!   output_SS316L_polycrystal.txt, cg_history_polycrystal.txt,
!   cgitr_SS316L_polycrystal.txt                               logs  (f90:278, 466, 1458)
!   fullfield_poly.h5           (/pix, /vm, /stress_tensor)     write (f90:2840)
!   top_peaks.txt               the MATLAB post-processing's hotspots
!                               (applications/sec3_2/A_2_0/elas3dxtal_postprocessing.m
!                               :86-243), when hot_threshold_mpa > 0
! The grid extents come from the /pix 'dimensions' attribute when present.
!------------------------------------------------------------------------------
      program elas3d_gpu
      use, intrinsic :: iso_fortran_env, dp => real64
      use, intrinsic :: iso_c_binding
      use e3x_bindings
      implicit none

      ! ---- run-time parameters (defaults of f90:77-148) -------------------------
      integer :: nx = 101, ny = 101, nz = 101, n_grains = 910
      real(dp) :: c11_local = 206.00d9, c12_local = 133.00d9, c44_local = 119.00d9
      real(dp) :: e1 = 0.0d0, nu1 = 0.0d0          ! pore phase (f90:114-115)
      real(dp) :: k_matrix = 0.0d0, g_matrix = 0.0d0   ! flag_m = 0 matrix moduli
      integer :: flag_m = 1
      real(dp) :: aml = real(1.0e-3, dp)            ! single-precision literal (f90:125)
      real(dp) :: aml_exx = 0.0d0, aml_eyy = 0.0d0, aml_ezz = -1.0d0
      real(dp) :: aml_exz = 0.0d0, aml_eyz = 0.0d0, aml_exy = 0.0d0
      integer :: kmax = 1, ldemb = 10000
      real :: tol = 1.0d-8                          ! `real` (f90:148)
      integer :: precision = E3X_PREC_FP64, replace_every = 50, device = 0
      integer :: stop_rule = E3X_STOP_REFERENCE, pcg_variant = E3X_PCG_REFERENCE
      ! hotspots (top_peaks.txt): off unless hot_threshold_mpa > 0; sigma, peaks and the
      ! voxel coordinates x = hot_x0 + i * hot_dx (i 0-based) as the script's xvec
      real(dp) :: hot_threshold_mpa = -1.0d0, hot_sigma = 0.4d0, hot_dx = 1.0d0, hot_x0 = 0.0d0
      integer :: hot_npeaks = 3
      character(len=512) :: input_file = 'input_structure_poly.h5'
      character(len=512) :: output_dir = '.'
      namelist /elas3d/ nx, ny, nz, n_grains, c11_local, c12_local, c44_local, e1, nu1, &
                        k_matrix, g_matrix, flag_m, aml, aml_exx, aml_eyy, aml_ezz, aml_exz, &
                        aml_eyz, aml_exy, kmax, ldemb, tol, precision, replace_every, device, &
                        stop_rule, pcg_variant, hot_threshold_mpa, hot_sigma, hot_npeaks, hot_dx, &
                        hot_x0, input_file, output_dir

      ! ---- state ------------------------------------------------------------------
      type(c_ptr) :: h = c_null_ptr
      integer :: nphase, i, kkk, ltot, lhist, ios, nargs, npk
      integer(c_int64_t) :: hidx(64), nfound, m
      real(dp) :: hval(64)
      integer(int64) :: ns
      integer(c_int64_t) :: npix, nori
      integer(c_int32_t) :: dims(3)
      integer(c_int32_t), allocatable :: pix(:)
      real(dp), allocatable :: orientation(:), prob(:), vm(:), stress_field(:,:), history(:)
      real(dp), allocatable :: phasemod(:,:)
      real(dp) :: strain6(6), pm(4), cval(1), utot(1), gg(1), gg0(1), ggo(1)
      real(dp) :: stress6(6), strain_avg(6), youngs, rel_res, elapsed_time, pre_ms, pcg_ms
      integer(c_int) :: lstep(1), exitf(1), resets(1)
      integer :: t1, t2, tc
      character(len=512) :: nml_file
      character(len=:), allocatable :: odir
      real(dp), allocatable :: energies(:), ggs(:)
      integer, allocatable :: cgiter(:)

      ! ---- parameters ------------------------------------------------------------
      nargs = command_argument_count()
      nml_file = 'elas3d.nml'
      if (nargs >= 1) call get_command_argument(1, nml_file)
      open(unit=11, file=trim(nml_file), status='old', action='read', iostat=ios)
      if (ios == 0) then
        read(11, nml=elas3d, iostat=ios)
        if (ios /= 0) then
          print *, 'ERROR: cannot parse namelist &elas3d in ', trim(nml_file)
          stop 1
        end if
        close(11)
      end if
      if (aml_ezz < 0.0d0) aml_ezz = aml                 ! default load: ezz = aml (f90:128)
      odir = trim(output_dir)//'/'

      ! grid and grain count from the input file when it carries them
      call e3x_check(e3x_h5_probe_input(trim(input_file)//c_null_char, npix, nori, dims), &
                     'e3x_h5_probe_input', io=.true.)
      if (dims(1) > 0 .and. int(dims(1), int64)*dims(2)*dims(3) == npix) then
        nx = dims(1); ny = dims(2); nz = dims(3)
      end if
      if (flag_m == 1) n_grains = int(nori / 3)
      nphase = n_grains + 1
      if (flag_m == 0) nphase = 2
      ns = int(nx, int64) * ny * nz

      allocate(pix(ns), orientation(max(3*(nphase-1), 1)), prob(nphase), vm(ns), stress_field(ns, 6))
      allocate(phasemod(nphase, 2), history(max(ldemb, 1)))
      allocate(energies(kmax+1), ggs(kmax+1), cgiter(kmax+1))

      open(unit=7, file=odir//'output_SS316L_polycrystal.txt', status='replace', iostat=ios)
      call system_clock(t1, tc)
      write(7, '(a,i4,a,i4,a,i4,a,i12)') 'nx = ', nx, ', ny =', ny, ', nz =', nz, ', ns =', ns
      print '(a,i4,a,i4,a,i4,a,i12)', 'nx = ', nx, ', ny =', ny, ', nz =', nz, ', ns =', ns

      ! ---- phase moduli (f90:295-317): E, nu -> K, G ------------------------------
      phasemod = 0.0d0
      if (flag_m == 0) then
        phasemod(1, 1) = k_matrix; phasemod(1, 2) = g_matrix
        youngs = e1
        phasemod(2, 1) = e1 / 3.d0 / (1.d0 - 2.d0*nu1)
        phasemod(2, 2) = youngs / 2.d0 / (1.d0 + nu1)
        pm = [phasemod(1, 1), phasemod(1, 2), phasemod(2, 1), phasemod(2, 2)]
      else
        youngs = e1
        phasemod(nphase, 1) = e1 / 3.d0 / (1.d0 - 2.d0*nu1)
        phasemod(nphase, 2) = youngs / 2.d0 / (1.d0 + nu1)
        pm = [phasemod(nphase, 1), phasemod(nphase, 2), 0.0d0, 0.0d0]
      end if

      ! ---- microstructure (ppixel_hdf5 f90:380, assig f90:388) -------------------
      call e3x_check(e3x_h5_read_input(trim(input_file)//c_null_char, ns, nphase, pix, orientation), &
                     'ppixel_hdf5', io=.true.)
      call e3x_check(e3x_create(nx, ny, nz, nphase, 1, device, h), 'e3x_create')
      call e3x_check(e3x_set_stop_rule(h, stop_rule), 'e3x_set_stop_rule')
      call e3x_check(e3x_set_pcg_variant(h, pcg_variant), 'e3x_set_pcg_variant')
      call e3x_check(e3x_set_microstructure(h, pix, orientation), 'e3x_set_microstructure')
      print *, 'Unique phases in pix:', minval(pix), maxval(pix)
      do i = 1, nphase
        write(7, '(A,I7,A,F20.6,A,F20.6)') 'Phase ', i, ' bulk=', phasemod(i, 1), ' shear=', phasemod(i, 2)
      end do
      call e3x_check(e3x_assig(h, prob), 'assig')
      do i = 1, nphase
        write(7, '(A,I7,A,F20.6)') 'Volume fraction of phase ', i, ' is ', prob(i)
      end do

      ! ---- applied strain (f90:394-406) -----------------------------------------
      strain6 = [aml_exx, aml_eyy, aml_ezz, aml_exz, aml_eyz, aml_exy]
      write(7, *) 'Applied engineering strains'
      write(7, *) ' exx eyy ezz exz eyz exy'
      write(7, *) strain6(1), strain6(2), strain6(3), 2.d0*strain6(4), 2.d0*strain6(5), 2.d0*strain6(6)
      print *, 'Applied engineering strains: exx eyy ezz exz eyz exy'
      print '(6E15.5)', strain6(1), strain6(2), strain6(3), 2.d0*strain6(4), 2.d0*strain6(5), 2.d0*strain6(6)

      ! ---- femat + affine guess (f90:417-438) -----------------------------------
      call e3x_check(e3x_femat(h, c11_local, c12_local, c44_local, pm, flag_m, strain6, cval), 'femat')
      call e3x_check(e3x_affine_guess(h), 'affine guess')

      ! ---- PCG relaxation (f90:445-511) ----------------------------------------
      ltot = 0
      lhist = 1
      call e3x_check(e3x_energy(h, utot, gg), 'energy')
      gg0 = gg
      write(7, *) 'Initial energy=', utot(1), 'gg=', gg(1)
      write(7, *) '   '
      call flush(7)
      print '(A,E12.5,A,E12.5)', 'Initial energy=', utot(1), ', Initial gg=', gg(1)
      print *, '   '
      energies(lhist) = utot(1)
      ggs(lhist) = sqrt(gg(1)/gg0(1))
      cgiter(lhist) = 0
      open(unit=99, file=odir//'cg_history_polycrystal.txt', status='replace')
      write(99, '(I8,1X,1PE18.10,1X,1PE18.10)') cgiter(lhist), energies(lhist), ggs(lhist)

      do kkk = 1, kmax
        call e3x_check(e3x_dembx_ex(h, precision, replace_every, gg0, tol, ldemb, kkk, lstep, ggo, &
                                    history, exitf, resets), 'dembx')
        gg = ggo
        if (exitf(1) == E3X_PCG_HAH_BREAK) print *, 'WARNING: hAh is nearly zero, breaking CG step.'
        if (resets(1) > 0) print *, '  Direction reset count:', resets(1)
        ! cgitr file (f90:1458, 2019-2021)
        open(unit=100, file=odir//'cgitr_SS316L_polycrystal.txt', status='replace', iostat=ios)
        do i = 1, lstep(1)
          if (history(i) == history(i)) write(100, '(I8, 4X, E20.8)') i, history(i)
        end do
        close(100)
        ltot = ltot + lstep(1)
        lhist = lhist + 1
        call e3x_check(e3x_energy(h, utot, ggo), 'energy')
        write(7, *) 'Energy=', utot(1), 'gg=', gg(1)
        write(7, *) 'Number of conjugate steps=', ltot
        call flush(7)
        print '(A,E12.5,A,E12.5,A,E12.5)', 'Energy=', utot(1), ', gg=', gg(1), &
              ', rel. residual norm=', sqrt(gg(1)/gg0(1))
        print *, 'Number of conjugate steps=', ltot
        energies(lhist) = utot(1)
        ggs(lhist) = sqrt(gg(1)/gg0(1))
        cgiter(lhist) = ltot
        write(99, '(I8,1X,1PE18.10,1X,1PE18.10)') cgiter(lhist), energies(lhist), ggs(lhist)
        rel_res = sqrt(gg(1)/gg0(1))
        if (rel_res < tol) exit
        call e3x_check(e3x_stress(h, stress6, strain_avg), 'stress')
        write(7, *) 'stresses: xx, yy, zz, xz, yz, xy'
        write(7, '(6E15.6)') stress6
        write(7, *) 'strains: xx, yy, zz, xz, yz, xy'
        write(7, '(6E15.6)') strain_avg
        write(7, *) '     '
        call flush(7)
      end do
      close(99)

      call e3x_check(e3x_stress(h, stress6, strain_avg), 'stress')
      write(7, *) 'stresses: xx, yy, zz, xz, yz, xy'
      write(7, '(6E15.6)') stress6
      write(7, *) 'strains: xx, yy, zz, xz, yz, xy'
      write(7, '(6E15.6)') strain_avg
      print *, 'End of CG'

      call system_clock(t2)
      elapsed_time = dble(t2 - t1) / dble(tc)
      call e3x_check(e3x_last_timing(h, pre_ms, pcg_ms), 'timing')
      print *, ''
      print *, 'Calculation complete!'
      print *, 'Total calculation time: ', elapsed_time, ' seconds'
      print '(A,F12.3,A,F12.3,A)', ' GPU preconditioner build: ', pre_ms, ' ms, PCG loop: ', pcg_ms, ' ms'
      write(7, *) ''
      write(7, *) 'Calculation complete!'
      write(7, *) 'Total calculation time: ', elapsed_time, ' seconds'

      ! ---- full field (f90:540, 2639-2877) -----------------------------------------
      call e3x_check(e3x_stress_fullfield(h, 0, stress_field, vm), 'stress_fullfield')
      call e3x_check(e3x_h5_write_fullfield(odir//'fullfield_poly.h5'//c_null_char, ns, pix, vm, &
                                            stress_field), 'fullfield HDF5 write', io=.true.)

      ! ---- hotspots of the MATLAB post-processing (postprocessing.m:86-243) ---------
      ! solid = every label up to max(pix), as the script sets it (.m:67); the first
      ! min(3, found) peaks in the script's format (.m:232-243)
      if (hot_threshold_mpa > 0.0d0) then
        npk = max(1, min(hot_npeaks, 64))
        call e3x_check(e3x_hotspots(h, 0, c_null_ptr, int(maxval(pix), c_int), hot_sigma, &
                                    hot_threshold_mpa, int(npk, c_int), hidx, hval, nfound), 'hotspots')
        if (nfound == 0) then
          print *, 'No hotspots found exceeding the threshold!'
        else
          npk = int(min(3_c_int64_t, min(int(npk, c_int64_t), nfound)))
          open(unit=101, file=odir//'top_peaks.txt', status='replace')
          write(101, '(A,I0,A)') 'Top ', npk, ' Local Maxima (Von Mises Stress):'
          do i = 1, npk
            m = hidx(i)
            write(101, '(A,I0,A,A,A,A,A,A,A,A,A)') 'Peak #', i, ': ', &
                fmt_f(real(real(hval(i), c_float) / 1.0e6_c_float, dp), 3), ' MPa at (', &
                fmt_f(hot_x0 + real(mod(m, int(nx, c_int64_t)), dp) * hot_dx, 2), ', ', &
                fmt_f(hot_x0 + real(mod(m / nx, int(ny, c_int64_t)), dp) * hot_dx, 2), ', ', &
                fmt_f(hot_x0 + real(m / (int(nx, c_int64_t) * ny), dp) * hot_dx, 2), ')'
          end do
          close(101)
        end if
      end if
      call e3x_check(e3x_destroy(h), 'e3x_destroy')
      close(7)
      contains

      ! x with d decimals and no padding, as C's %.<d>f prints it
      function fmt_f(x, d) result(s)
        real(dp), intent(in) :: x
        integer, intent(in) :: d
        character(len=:), allocatable :: s
        character(len=48) :: buf, f
        write(f, '(A,I0,A)') '(F48.', d, ')'
        write(buf, f) x
        s = trim(adjustl(buf))
        if (s(1:1) == '.') s = '0'//s
        if (len(s) > 1) then
          if (s(1:2) == '-.') s = '-0'//s(2:)
        end if
      end function fmt_f
      end program elas3d_gpu

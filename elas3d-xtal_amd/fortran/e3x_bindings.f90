!------------------------------------------------------------------------------
! e3x_bindings: iso_c_binding interfaces of the gfx950 solver's C ABI
! (include/e3x.h, include/e3x_io.h).  Each interface names the routine of
! src/elas3dxtal_pcg.f90 it replaces.  Arrays are passed by address in their
! Fortran layout (u(ns,3), stress_field(ns,6)); optional outputs take c_null_ptr.
!------------------------------------------------------------------------------
      module e3x_bindings
      use, intrinsic :: iso_c_binding
      implicit none

      integer(c_int), parameter :: E3X_OK = 0, E3X_EINVAL = 1, E3X_EDEVICE = 2, &
                                   E3X_ENOMEM = 3, E3X_EPHASE = 4, E3X_ESTATE = 5, E3X_EIO = 6
      integer(c_int), parameter :: E3X_PCG_CONVERGED = 1, E3X_PCG_HAH_BREAK = 2, E3X_PCG_MAXITER = 3
      integer(c_int), parameter :: E3X_FIELD_U = 0, E3X_FIELD_GB = 1, E3X_FIELD_B = 2, E3X_FIELD_H = 3
      integer(c_int), parameter :: E3X_PREC_FP64 = 0, E3X_PREC_MIXED = 1
      integer(c_int), parameter :: E3X_STOP_REFERENCE = 0, E3X_STOP_NIST_ABS = 1
      integer(c_int), parameter :: E3X_PCG_REFERENCE = 0, E3X_PCG_SWEEP = 1, E3X_PCG_PLAIN_CG = 2

      interface
        function e3x_last_error() bind(C, name="e3x_last_error") result(p)
          import :: c_ptr
          type(c_ptr) :: p
        end function

        function e3x_io_last_error() bind(C, name="e3x_io_last_error") result(p)
          import :: c_ptr
          type(c_ptr) :: p
        end function

        ! allocations of program elas3d (f90:242-256)
        integer(c_int) function e3x_create(nx, ny, nz, nphase, nrhs, device, h) &
            bind(C, name="e3x_create")
          import :: c_int, c_ptr
          integer(c_int), value :: nx, ny, nz, nphase, nrhs, device
          type(c_ptr) :: h
        end function

        integer(c_int) function e3x_destroy(h) bind(C, name="e3x_destroy")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        ! data part of ppixel_hdf5 (f90:2879-2991)
        integer(c_int) function e3x_set_microstructure(h, pix, orientation) &
            bind(C, name="e3x_set_microstructure")
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: h
          integer(c_int32_t), intent(in) :: pix(*)
          real(c_double), intent(in) :: orientation(*)
        end function

        ! assig (f90:2410-2468)
        integer(c_int) function e3x_assig(h, prob) bind(C, name="e3x_assig")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double) :: prob(*)
        end function

        ! femat_OpenMP (f90:555-1104)
        integer(c_int) function e3x_femat(h, c11, c12, c44, phasemod, flag_m, strain6, c_out) &
            bind(C, name="e3x_femat")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), value :: c11, c12, c44
          real(c_double), intent(in) :: phasemod(*), strain6(*)
          integer(c_int), value :: flag_m
          real(c_double) :: c_out(*)
        end function

        ! affine initial guess (f90:423-438)
        integer(c_int) function e3x_affine_guess(h) bind(C, name="e3x_affine_guess")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        ! energy_OpenMP (f90:1107-1353) and gg = sum(gb**2) (f90:452)
        integer(c_int) function e3x_energy(h, utot, gg) bind(C, name="e3x_energy")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double) :: utot(*), gg(*)
        end function

        ! stop test of the following dembx calls: the reference's rel_res < tol
        ! (f90:1996-2004, default) or NIST ELAS3D's sum(gb**2) < 1e-8*ns
        integer(c_int) function e3x_set_stop_rule(h, rule) bind(C, name="e3x_set_stop_rule")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
          integer(c_int), value :: rule
        end function

        ! loop of the following dembx calls: src (f90:1356), the anisotropy-sweep
        ! programs' (A_2_0:1045) or the plain-CG programs' (cg:915) dembx_OpenMP
        integer(c_int) function e3x_set_pcg_variant(h, variant) bind(C, name="e3x_set_pcg_variant")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
          integer(c_int), value :: variant
        end function

        ! dembx_OpenMP (f90:1356-2043), with a precision mode
        integer(c_int) function e3x_dembx_ex(h, precision, replace_every, gg0, tol, ldemb, kkk, &
                                             lstep, gg, history, exit_flag, reset_count) &
            bind(C, name="e3x_dembx_ex")
          import :: c_int, c_ptr, c_double, c_float
          type(c_ptr), value :: h
          integer(c_int), value :: precision, replace_every, ldemb, kkk
          real(c_double), intent(in) :: gg0(*)
          real(c_float), value :: tol
          integer(c_int) :: lstep(*), exit_flag(*), reset_count(*)
          real(c_double) :: gg(*), history(*)
        end function

        ! the MATLAB post-processing's hotspots for top_peaks.txt
        ! (applications/sec3_2/A_2_0/elas3dxtal_postprocessing.m:86-243); vm may be
        ! c_null_ptr for the handle's own field of right-hand side rhs
        integer(c_int) function e3x_hotspots(h, rhs, vm, nsolid, sigma, threshold_mpa, npeaks, &
                                             idx, val, found) bind(C, name="e3x_hotspots")
          import :: c_int, c_ptr, c_double, c_int64_t
          type(c_ptr), value :: h, vm
          integer(c_int), value :: rhs, nsolid, npeaks
          real(c_double), value :: sigma, threshold_mpa
          integer(c_int64_t) :: idx(*), found
          real(c_double) :: val(*)
        end function

        ! stress_OpenMP (f90:2149-2408)
        integer(c_int) function e3x_stress(h, stress6, strain6) bind(C, name="e3x_stress")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double) :: stress6(*), strain6(*)
        end function

        ! six unit load cases as six right-hand sides -> 6x6 effective stiffness
        ! (no counterpart in the reference; the handle needs nrhs = 6)
        integer(c_int) function e3x_effective_stiffness(h, c11, c12, c44, phasemod, flag_m, magnitude, &
            precision, replace_every, tol, ldemb, c_eff, lstep, exit_flag) &
            bind(C, name="e3x_effective_stiffness")
          import :: c_int, c_ptr, c_double, c_float
          type(c_ptr), value :: h
          real(c_double), value :: c11, c12, c44, magnitude
          real(c_double), intent(in) :: phasemod(*)
          integer(c_int), value :: flag_m, precision, replace_every, ldemb
          real(c_float), value :: tol
          real(c_double) :: c_eff(*)
          integer(c_int) :: lstep(*), exit_flag(*)
        end function

        ! stress_fullfield_OpenMP, computation (f90:2639-2833)
        integer(c_int) function e3x_stress_fullfield(h, rhs, stress, vm) &
            bind(C, name="e3x_stress_fullfield")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          integer(c_int), value :: rhs
          real(c_double) :: stress(*), vm(*)
        end function

        integer(c_int) function e3x_last_timing(h, precond_ms, pcg_ms) bind(C, name="e3x_last_timing")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double) :: precond_ms, pcg_ms
        end function

        ! HDF5 probe / ppixel_hdf5 file part / full-field writer (f90:2840-2874)
        integer(c_int) function e3x_h5_probe_input(path, npix, norient, dims) &
            bind(C, name="e3x_h5_probe_input")
          import :: c_int, c_char, c_int64_t, c_int32_t
          character(kind=c_char), intent(in) :: path(*)
          integer(c_int64_t) :: npix, norient
          integer(c_int32_t) :: dims(3)
        end function

        integer(c_int) function e3x_h5_read_input(path, ns, nphase, pix, orientation) &
            bind(C, name="e3x_h5_read_input")
          import :: c_int, c_char, c_int64_t, c_int32_t, c_double
          character(kind=c_char), intent(in) :: path(*)
          integer(c_int64_t), value :: ns
          integer(c_int), value :: nphase
          integer(c_int32_t) :: pix(*)
          real(c_double) :: orientation(*)
        end function

        integer(c_int) function e3x_h5_write_fullfield(path, ns, pix, vm, stress) &
            bind(C, name="e3x_h5_write_fullfield")
          import :: c_int, c_char, c_int64_t, c_int32_t, c_double
          character(kind=c_char), intent(in) :: path(*)
          integer(c_int64_t), value :: ns
          integer(c_int32_t), intent(in) :: pix(*)
          real(c_double), intent(in) :: vm(*), stress(*)
        end function

        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
          import :: c_size_t, c_ptr
          type(c_ptr), value :: s
        end function
      end interface

      contains

      ! C string (char*) -> Fortran string
      function c_to_f(p) result(s)
        type(c_ptr), intent(in) :: p
        character(len=:), allocatable :: s
        character(kind=c_char), pointer :: a(:)
        integer :: n, i
        if (.not. c_associated(p)) then
          s = ''
          return
        end if
        n = int(c_strlen(p))
        call c_f_pointer(p, a, [n])
        allocate(character(len=n) :: s)
        do i = 1, n
          s(i:i) = a(i)
        end do
      end function

      ! Stop with the library's message on a nonzero status (the reference `stop`s on
      ! input errors, f90:2924, 2941, 2968, 2988).
      subroutine e3x_check(rc, what, io)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        logical, intent(in), optional :: io
        logical :: from_io
        if (rc == E3X_OK) return
        from_io = .false.
        if (present(io)) from_io = io
        if (from_io) then
          print *, 'ERROR in ', what, ' (', rc, '): ', c_to_f(e3x_io_last_error())
        else
          print *, 'ERROR in ', what, ' (', rc, '): ', c_to_f(e3x_last_error())
        end if
        stop 1
      end subroutine

      end module e3x_bindings

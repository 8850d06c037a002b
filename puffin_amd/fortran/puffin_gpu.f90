! Fortran binding of the gfx950 integrator's C ABI (include/puffin_gpu.h)
! through iso_c_binding, and drop-in replacements of the reference's
! rk4par / diffractIM / DiffractionStep that keep its call signatures and its
! module-global, scaled-variable arrays:
!
!   rk4par(sZ, h, qD)                  interaction/puffin_mpi_RK4.f90:38-59
!   diffractIM(sStep, qDiffrctd, qOK)  diffraction/diffraction.f90:46-52
!   DiffractionStep(h, sAr, sAi, qOK)  diffraction/diffraction.f90:209-220
!   allact_rk4_arrs / deallact_rk4_arrs  interaction/puffin_mpi_RK4.f90:374-460
!   matchIn(sZ) / matchOut(sZ)          acc_lattice.f90:743-828 / 638-725
!   correctTrans()                     acc_lattice.f90:601-623
! and the bodies of the other lattice elements with their values passed in
! (the reference indexes its lattice arrays):
!   gpu_drift(del_dr_z)                driftSection, acc_lattice.f90:478-518
!   gpu_disperse(disp, slip)           disperse, acc_lattice.f90:420-463
!   gpu_quad(qfx, qfy)                 Quad, acc_lattice.f90:531-561
!   gpu_bmodulation(wavenum, mag)      bModulation, acc_lattice.f90:575-585
! and the output reductions on the device particles:
!   getSliceTwiss(nslices, slicetrim, aveX, ..., sdata)  io/av_write.f90:405-773
!   getCurr(sam_len, Iarray)                            io/av_write.f90:297-368
!
! The particle arrays sElX_G .. sElGam_G, s_chi_bar_G (deriv_globals.f90:94,
! 128-133) and the field ac_rfield_in / ac_ifield_in (puffin_mpi_RK4.f90:24,
! whole mesh here, x fastest) stay resident on the device between
! allact_rk4_arrs and deallact_rk4_arrs; the flags qPArrOK_G / qInnerXYOK_G
! (deriv_globals.f90:407-409) are set from the status code of each step.
! A lattice element called while no device state is live uploads the module
! globals first and leaves the device state live: from then on the device
! copy is the current one until deallact_rk4_arrs or gpu_sync_to_host, and a
! following allact_rk4_arrs keeps it (updating only the module's undulator
! and taper).

module puffin_gpu_capi

  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: PUFF_FP64 = 0, PUFF_MIXED = 1
  integer(c_int32_t), parameter :: PUFF_MEM_HOST = 0, PUFF_MEM_DEVICE = 1
  integer(c_int), parameter :: PUFF_OK = 0, PUFF_OUT_TRANSVERSE = 1, &
                               PUFF_OUT_Z2_BUFFER = 2, PUFF_OUT_Z2_MESH = 3

  ! puff_params (include/puffin_gpu.h)
  type, bind(C) :: puff_params_t
    integer(c_int32_t) :: nx, ny, nz2, periodic, one_d, und_type, und_ends, &
                          focusing, electrons_evolve, field_evolve, coupling, &
                          filter, nbx, nby, nbz2, reserved
    real(c_double) :: dx, dy, dz2, rho, eta, kappa, gamma_r, aw, fx, fy, &
                      szfs, szfe, kbeta_x, kbeta_y, kx_und, ky_und, beta, &
                      filt_cutoff, pi_ref
  end type puff_params_t

  interface

    function puff_device_count() bind(C, name="puff_device_count") result(r)
      import :: c_int
      integer(c_int) :: r
    end function

    function puff_last_error() bind(C, name="puff_last_error") result(r)
      import :: c_ptr
      type(c_ptr) :: r
    end function

    function puff_create(ctx, params, n, precision) bind(C, name="puff_create") result(r)
      import :: c_ptr, c_int, c_int64_t, c_int32_t, puff_params_t
      type(c_ptr), intent(out) :: ctx
      type(puff_params_t), intent(in) :: params
      integer(c_int64_t), value :: n
      integer(c_int32_t), value :: precision
      integer(c_int) :: r
    end function

    function puff_destroy(ctx) bind(C, name="puff_destroy") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: r
    end function

    function puff_upload_particles(ctx, kind, x, y, z2, px, py, gam, chi) &
        bind(C, name="puff_upload_particles") result(r)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kind
      real(c_double), intent(in) :: x(*), y(*), z2(*), px(*), py(*), gam(*), chi(*)
      integer(c_int) :: r
    end function

    function puff_download_particles(ctx, kind, x, y, z2, px, py, gam) &
        bind(C, name="puff_download_particles") result(r)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kind
      real(c_double), intent(out) :: x(*), y(*), z2(*), px(*), py(*), gam(*)
      integer(c_int) :: r
    end function

    function puff_upload_field(ctx, kind, ar, ai) bind(C, name="puff_upload_field") result(r)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kind
      real(c_double), intent(in) :: ar(*), ai(*)
      integer(c_int) :: r
    end function

    function puff_download_field(ctx, kind, ar, ai) bind(C, name="puff_download_field") result(r)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kind
      real(c_double), intent(out) :: ar(*), ai(*)
      integer(c_int) :: r
    end function

    function puff_set_taper(ctx, n2col, n2col0, undgrad) bind(C, name="puff_set_taper") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: n2col, n2col0, undgrad
      integer(c_int) :: r
    end function

    function puff_get_taper(ctx, n2col, n2col0) bind(C, name="puff_get_taper") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: n2col, n2col0
      integer(c_int) :: r
    end function

    function puff_rk4par(ctx, sz, h) bind(C, name="puff_rk4par") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: sz, h
      integer(c_int) :: r
    end function

    function puff_diffraction_step(ctx, h) bind(C, name="puff_diffraction_step") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: h
      integer(c_int) :: r
    end function

    function puff_wrap_z2(ctx) bind(C, name="puff_wrap_z2") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: r
    end function

    function puff_power(ctx, power) bind(C, name="puff_power") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: power(*)
      integer(c_int) :: r
    end function

    function puff_bunching(ctx, harmonic, nslices, slice_len, b) bind(C, name="puff_bunching") result(r)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: harmonic, nslices
      real(c_double), value :: slice_len
      real(c_double), intent(out) :: b(*)
      integer(c_int) :: r
    end function

    function puff_slice_stats(ctx, nslices, slice_len, out) bind(C, name="puff_slice_stats") result(r)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nslices
      real(c_double), value :: slice_len
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: r
    end function

    function puff_current(ctx, npts, sam_len, out) bind(C, name="puff_current") result(r)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: npts
      real(c_double), value :: sam_len
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: r
    end function

    function puff_drift(ctx, dz) bind(C, name="puff_drift") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dz
      integer(c_int) :: r
    end function

    function puff_chicane(ctx, disp, slip) bind(C, name="puff_chicane") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: disp, slip
      integer(c_int) :: r
    end function

    function puff_quad(ctx, fx, fy) bind(C, name="puff_quad") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: fx, fy
      integer(c_int) :: r
    end function

    function puff_modulation(ctx, wavenum, mag) bind(C, name="puff_modulation") result(r)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: wavenum, mag
      integer(c_int) :: r
    end function

    function puff_correct_trans(ctx) bind(C, name="puff_correct_trans") result(r)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: r
    end function

    function puff_match(ctx, sz, n2col, sign) bind(C, name="puff_match") result(r)
      import :: c_ptr, c_int, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), value :: sz, n2col
      integer(c_int32_t), value :: sign
      integer(c_int) :: r
    end function

    function puff_set_undulator(ctx, und_type, fx, fy, kbeta_x, kbeta_y, szfs, szfe) &
        bind(C, name="puff_set_undulator") result(r)
      import :: c_ptr, c_int, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: und_type
      real(c_double), value :: fx, fy, kbeta_x, kbeta_y, szfs, szfe
      integer(c_int) :: r
    end function

  end interface

end module puffin_gpu_capi


module puffin_gpu

  use iso_c_binding
  use puffin_gpu_capi
  implicit none

  integer, parameter :: wp = c_double

  ! module constants of the current undulator module (filled by the caller as
  ! setup / initUndulator leave them)
  type(puff_params_t), save :: tParams_G
  integer(c_int32_t), save :: iPrecision_G = PUFF_FP64

  ! macroparticles (deriv_globals.f90:94,128-133)
  real(wp), allocatable, save :: sElX_G(:), sElY_G(:), sElZ2_G(:), &
                                 sElPX_G(:), sElPY_G(:), sElGam_G(:), s_chi_bar_G(:)
  ! field on the whole mesh (puffin_mpi_RK4.f90:24)
  real(wp), allocatable, save :: ac_rfield_in(:), ac_ifield_in(:)
  ! taper state (wiggler_taper.f90)
  real(wp), save :: n2col = 1.0_wp, n2col0 = 1.0_wp, undgrad = 0.0_wp
  ! parallel-layout flags (deriv_globals.f90:407-409)
  logical, save :: qPArrOK_G = .true., qInnerXYOK_G = .true.
  ! fatal temporal-mesh overflow (the reference STOPs)
  logical, save :: qZ2MeshOK_G = .true.

  type(c_ptr), save :: ctx_G = c_null_ptr

  ! scalings of the output routines (deriv_globals.f90:102-106, 199), filled
  ! by the caller as setup leaves them
  real(wp), save :: sRho_G = 0.01_wp, npk_bar_G = 1.0_wp, ata_G = 1.0_wp, lc_G = 1.0_wp
  logical, save :: qOneD_G = .false.

contains

  subroutine gpu_check(r, where)
    integer(c_int), intent(in) :: r
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    if (r < 0) then
      call c_f_pointer(puff_last_error(), msg, [512])
      write(*, '(a)', advance='no') 'puffin_gpu: '//where//' failed: '
      do i = 1, 512
        if (msg(i) == c_null_char) exit
        write(*, '(a)', advance='no') msg(i)
      end do
      write(*, *)
      error stop 2
    end if
  end subroutine gpu_check

  ! Upload the module globals to the device (replaces allact_rk4_arrs,
  ! puffin_mpi_RK4.f90:374-418, and its outer2Inner copy).
  subroutine allact_rk4_arrs()
    integer(c_int64_t) :: n
    n = int(size(sElX_G), c_int64_t)
    if (c_associated(ctx_G)) then
      ! device state left live by a lattice element: keep it, take the
      ! module's undulator and taper
      call gpu_check(puff_set_undulator(ctx_G, tParams_G%und_type, tParams_G%fx, tParams_G%fy, &
                     tParams_G%kbeta_x, tParams_G%kbeta_y, tParams_G%szfs, tParams_G%szfe), &
                     'puff_set_undulator')
      call gpu_check(puff_set_taper(ctx_G, n2col, n2col0, undgrad), 'puff_set_taper')
      qInnerXYOK_G = .true.
      qPArrOK_G = .true.
      return
    end if
    call gpu_check(puff_create(ctx_G, tParams_G, n, iPrecision_G), 'puff_create')
    call gpu_check(puff_upload_particles(ctx_G, PUFF_MEM_HOST, sElX_G, sElY_G, sElZ2_G, &
                   sElPX_G, sElPY_G, sElGam_G, s_chi_bar_G), 'puff_upload_particles')
    call gpu_check(puff_upload_field(ctx_G, PUFF_MEM_HOST, ac_rfield_in, ac_ifield_in), &
                   'puff_upload_field')
    call gpu_check(puff_set_taper(ctx_G, n2col, n2col0, undgrad), 'puff_set_taper')
    qInnerXYOK_G = .true.
    qPArrOK_G = .true.
  end subroutine allact_rk4_arrs

  ! Copy the device state back into the module globals.
  subroutine gpu_sync_to_host()
    call gpu_check(puff_download_particles(ctx_G, PUFF_MEM_HOST, sElX_G, sElY_G, sElZ2_G, &
                   sElPX_G, sElPY_G, sElGam_G), 'puff_download_particles')
    call gpu_check(puff_download_field(ctx_G, PUFF_MEM_HOST, ac_rfield_in, ac_ifield_in), &
                   'puff_download_field')
    call gpu_check(puff_get_taper(ctx_G, n2col, n2col0), 'puff_get_taper')
  end subroutine gpu_sync_to_host

  ! Download and release (replaces deallact_rk4_arrs, puffin_mpi_RK4.f90:423-460).
  subroutine deallact_rk4_arrs()
    integer(c_int) :: r
    if (.not. c_associated(ctx_G)) return
    call gpu_sync_to_host()
    r = puff_destroy(ctx_G)
    ctx_G = c_null_ptr
  end subroutine deallact_rk4_arrs

  ! One RK4 step of particles and field (rk4par, puffin_mpi_RK4.f90:38-366).
  ! qD is vestigial in the reference (only reset, :154-166).
  subroutine rk4par(sZ, h, qD)
    real(wp), intent(in) :: sZ
    real(wp), intent(in) :: h
    logical, intent(inout) :: qD
    integer(c_int) :: r
    qD = .false.
    r = puff_rk4par(ctx_G, sZ, h)
    call gpu_check(r, 'puff_rk4par')
    select case (r)
    case (PUFF_OUT_TRANSVERSE)
      qInnerXYOK_G = .false.
      qPArrOK_G = .false.
    case (PUFF_OUT_Z2_BUFFER)
      qPArrOK_G = .false.
    case (PUFF_OUT_Z2_MESH)
      qZ2MeshOK_G = .false.
      qPArrOK_G = .false.
    end select
  end subroutine rk4par

  ! Diffract the device field over sStep (diffractIM, diffraction.f90:46-91).
  subroutine diffractIM(sStep, qDiffrctd, qOK)
    real(wp), intent(in) :: sStep
    logical, intent(out) :: qDiffrctd, qOK
    call gpu_check(puff_diffraction_step(ctx_G, sStep), 'puff_diffraction_step')
    qDiffrctd = .true.
    qOK = .true.
  end subroutine diffractIM

  ! Diffract a host field held in split real / imaginary arrays
  ! (DiffractionStep, diffraction.f90:209-348); uses the module's context.
  subroutine DiffractionStep(h, sAr, sAi, qOK)
    real(wp), intent(in) :: h
    real(wp), intent(inout) :: sAr(:), sAi(:)
    logical, intent(out) :: qOK
    call gpu_check(puff_upload_field(ctx_G, PUFF_MEM_HOST, sAr, sAi), 'puff_upload_field')
    call gpu_check(puff_diffraction_step(ctx_G, h), 'puff_diffraction_step')
    call gpu_check(puff_download_field(ctx_G, PUFF_MEM_HOST, sAr, sAi), 'puff_download_field')
    qOK = .true.
  end subroutine DiffractionStep

  ! Periodic z2 wrap done by getLocalFieldIndices (para_field.f90:212-215).
  subroutine wrap_z2()
    call gpu_check(puff_wrap_z2(ctx_G), 'puff_wrap_z2')
  end subroutine wrap_z2

  ! Power per z2 node of the device field (gPowerP, av_write.f90:39-64).
  subroutine gPowerP(power)
    real(wp), intent(out) :: power(:)
    call gpu_check(puff_power(ctx_G, power), 'puff_power')
  end subroutine gPowerP

  ! ---- lattice elements (acc_lattice.f90) on the device macroparticles

  subroutine ensure_device()
    if (.not. c_associated(ctx_G)) call allact_rk4_arrs()
  end subroutine ensure_device

  ! matchIn (acc_lattice.f90:743-828): undulator entry offsets at zbar sZ
  ! with the current n2col, for tParams_G's undulator type.
  subroutine matchIn(sZ)
    real(wp), intent(in) :: sZ
    call ensure_device()
    call gpu_check(puff_match(ctx_G, sZ, n2col, 1_c_int32_t), 'puff_match')
  end subroutine matchIn

  ! matchOut (acc_lattice.f90:638-725): the exit offsets taken off again.
  subroutine matchOut(sZ)
    real(wp), intent(in) :: sZ
    call ensure_device()
    call gpu_check(puff_match(ctx_G, sZ, n2col, -1_c_int32_t), 'puff_match')
  end subroutine matchOut

  ! correctTrans (acc_lattice.f90:601-623): centre x, y, px, py on their
  ! chi-weighted means.
  subroutine correctTrans()
    call ensure_device()
    call gpu_check(puff_correct_trans(ctx_G), 'puff_correct_trans')
  end subroutine correctTrans

  ! driftSection's particle part (acc_lattice.f90:490-508); the caller
  ! diffracts over del_dr_z (diffractIM) and advances sZ as :512-515 do.
  subroutine gpu_drift(del_dr_z)
    real(wp), intent(in) :: del_dr_z
    call ensure_device()
    call gpu_check(puff_drift(ctx_G, del_dr_z), 'puff_drift')
  end subroutine gpu_drift

  ! disperse's particle part with the periodic wrap (acc_lattice.f90:436-461).
  subroutine gpu_disperse(disp, slip)
    real(wp), intent(in) :: disp, slip
    call ensure_device()
    call gpu_check(puff_chicane(ctx_G, disp, slip), 'puff_chicane')
  end subroutine gpu_disperse

  ! Quad (acc_lattice.f90:544-554); no kick on a 1-D mesh.
  subroutine gpu_quad(qfx, qfy)
    real(wp), intent(in) :: qfx, qfy
    call ensure_device()
    call gpu_check(puff_quad(ctx_G, qfx, qfy), 'puff_quad')
  end subroutine gpu_quad

  ! bModulation (acc_lattice.f90:579-580).
  subroutine gpu_bmodulation(wavenum, mag)
    real(wp), intent(in) :: wavenum, mag
    call ensure_device()
    call gpu_check(puff_modulation(ctx_G, wavenum, mag), 'puff_modulation')
  end subroutine gpu_bmodulation

  ! ---- output reductions (io/av_write.f90) on the device particles

  ! getSliceTwiss (av_write.f90:405-773): slices of 4 pi rho.  slicetrim is
  ! unused there too; the Twiss alpha and beta outputs are zero, as the
  ! reference leaves them.
  subroutine getSliceTwiss(nslices, slicetrim, aveX, aveY, avePX, avePY, sdX, sdY, sdpx, sdpy, &
                           eX, eY, ax, ay, bx, by, aveGamma, aveDgamma, b1, b2, b3, b4, b5, sdata)
    integer, intent(in) :: nslices
    real(wp), intent(in) :: slicetrim
    real(wp), dimension(nslices), intent(out) :: aveX, aveY, avePX, avePY, sdX, sdY, sdpx, sdpy, &
                                                 eX, eY, ax, ay, bx, by, aveGamma, aveDgamma, &
                                                 b1, b2, b3, b4, b5, sdata
    real(wp), allocatable :: o(:)
    real(wp) :: sl
    integer :: n
    n = nslices
    sl = 4.0_wp * pi_ref() * sRho_G
    call ensure_device()
    allocate(o(13 * n))
    call gpu_check(puff_slice_stats(ctx_G, int(n, c_int32_t), sl, o), 'puff_slice_stats')
    sdata = o(1:n);            aveX = o(n+1:2*n);       aveY = o(2*n+1:3*n)
    avePX = o(3*n+1:4*n);      avePY = o(4*n+1:5*n);    sdX = o(5*n+1:6*n)
    sdY = o(6*n+1:7*n);        sdpx = o(7*n+1:8*n);     sdpy = o(8*n+1:9*n)
    eX = o(9*n+1:10*n);        eY = o(10*n+1:11*n);     aveGamma = o(11*n+1:12*n)
    aveDgamma = o(12*n+1:13*n)
    ax = 0.0_wp; ay = 0.0_wp; bx = 0.0_wp; by = 0.0_wp
    call gpu_check(puff_bunching(ctx_G, 1_c_int32_t, int(n, c_int32_t), sl, b1), 'puff_bunching')
    call gpu_check(puff_bunching(ctx_G, 2_c_int32_t, int(n, c_int32_t), sl, b2), 'puff_bunching')
    call gpu_check(puff_bunching(ctx_G, 3_c_int32_t, int(n, c_int32_t), sl, b3), 'puff_bunching')
    call gpu_check(puff_bunching(ctx_G, 4_c_int32_t, int(n, c_int32_t), sl, b4), 'puff_bunching')
    call gpu_check(puff_bunching(ctx_G, 5_c_int32_t, int(n, c_int32_t), sl, b5), 'puff_bunching')
    deallocate(o)
  end subroutine getSliceTwiss

  ! getCurr (av_write.f90:297-368): charge on size(Iarray) nodes sam_len
  ! apart, in amperes, with the reference's single-precision constants.
  subroutine getCurr(sam_len, Iarray)
    real(wp), intent(in) :: sam_len
    real(wp), intent(inout) :: Iarray(:)
    real(wp), parameter :: q_e = 1.60217653e-19, c = 2.99792458e8
    call ensure_device()
    call gpu_check(puff_current(ctx_G, int(size(Iarray), c_int32_t), sam_len, Iarray), 'puff_current')
    Iarray = Iarray * npk_bar_G
    if (qOneD_G) Iarray = Iarray * ata_G
    Iarray = Iarray * q_e / sam_len
    Iarray = Iarray * c / lc_G
  end subroutine getCurr

  ! pi as the reference's default-real literal, widened
  pure function pi_ref() result(p)
    real(wp) :: p
    p = real(3.14159265358979323846, wp)
  end function pi_ref

end module puffin_gpu

! One undulator module driven from Fortran through the drop-in routines of
! module puffin_gpu: the split-step schedule of UndSection
! (undulator.f90:103-333) with rk4par, diffractIM, the periodic re-layout and
! the power writes.  With ``elements`` set, the module is entered with
! matchIn before the arrays go to the device and left as UndSection and the
! element chain leave it (undulator.f90:86, 337-340; puffin_module.f90:58-90):
! matchOut, correctTrans, then a drift (with its diffraction), a quad, a
! chicane and an energy modulation, all on the device.
!
! usage: und_section <input file> <output file>
!
! Input (stream, native byte order): the puff_params record, then
!   int64  n, nsteps, steps4diff, redist_every, diffraction, periodic, write_every, elements
!   real64 h, n2col, n2col0, undgrad
!   real64 drift, quad_fx, quad_fy, chic_disp, chic_slip, enmod_wavenum, enmod_mag
!   real64 x(n), y(n), z2(n), px(n), py(n), gam(n), chi(n), ar(nmesh), ai(nmesh)
! Output: int64 nwrites, then real64 x, y, z2, px, py, gam (n each), ar, ai
! (nmesh each), power(nz2, nwrites); with ``elements``, int64 nslices and
! real64 sdata, aveGamma, aveDgamma, avePX, sdpx, b1 (nslices each) from
! getSliceTwiss and Iarray (nslices + 1) from getCurr with unit npk_bar and
! Lc; and the int64 final status.

program und_section

  use iso_c_binding
  use puffin_gpu_capi
  use puffin_gpu
  implicit none

  character(len=4096) :: fin, fout
  integer(c_int64_t) :: n, nsteps, steps4diff, redist_every, idiff, iper, write_every, ielem
  real(wp) :: el(7)
  integer :: nsl
  real(wp), allocatable :: tw(:,:), cur(:)
  integer(c_int64_t) :: nmesh, istep, iCsteps, nw, status
  real(wp) :: h, szl, dzdS, dzdF, dzd
  real(wp), allocatable :: powers(:,:), pw(:)
  logical :: qD, qDiffrctd, qOK, qWrite, qDone
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  if (puff_device_count() < 1) then
    write(*, *) 'und_section: no HIP device'
    error stop 1
  end if

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) tParams_G
  read(u) n, nsteps, steps4diff, redist_every, idiff, iper, write_every, ielem
  read(u) h, n2col, n2col0, undgrad
  read(u) el
  nmesh = int(tParams_G%nx, c_int64_t) * tParams_G%ny * tParams_G%nz2
  allocate(sElX_G(n), sElY_G(n), sElZ2_G(n), sElPX_G(n), sElPY_G(n), sElGam_G(n), s_chi_bar_G(n))
  allocate(ac_rfield_in(nmesh), ac_ifield_in(nmesh))
  read(u) sElX_G, sElY_G, sElZ2_G, sElPX_G, sElPY_G, sElGam_G, s_chi_bar_G
  read(u) ac_rfield_in, ac_ifield_in
  close(u)

  allocate(powers(tParams_G%nz2, max(1_c_int64_t, nsteps)), pw(tParams_G%nz2))
  nw = 0
  status = 0
  iCsteps = 0
  szl = 0.0_wp
  qD = .true.

  if (ielem /= 0) call matchIn(szl)                   ! undulator.f90:86, before the arrays
  call allact_rk4_arrs()
  if (iper /= 0) call wrap_z2()                       ! undulator.f90:103
  dzdS = 0.0_wp
  if (idiff /= 0) then                                 ! undulator.f90:156-164
    dzdS = real(steps4diff, wp) * h / 2.0_wp
    call diffractIM(dzdS, qDiffrctd, qOK)
  end if

  istep = 0
  do
    istep = istep + 1
    if (istep > nsteps) exit
    iCsteps = iCsteps + 1
    call rk4par(szl, h, qD)                           ! undulator.f90:199
    if (.not. qPArrOK_G) then
      status = 2
      if (.not. qInnerXYOK_G) status = 1
      if (.not. qZ2MeshOK_G) status = 3
      exit
    end if
    szl = szl + h                                      ! undulator.f90:222
    qWrite = .false.                                   ! (.and. does not short-circuit)
    if (write_every > 0) qWrite = (mod(iCsteps, write_every) == 0) .or. (iCsteps == nsteps)
    qDone = .false.
    if (idiff /= 0) then
      if ((mod(istep, steps4diff) == 0) .or. (istep == nsteps)) then   ! :229-289
        dzdF = dzdS
        if (istep == nsteps) then
          dzdS = 0.0_wp
        else if ((istep + steps4diff) <= nsteps) then
          dzdS = real(steps4diff, wp) * h / 2
        else
          dzdS = real(nsteps - istep, wp) * h / 2
        end if
        if (.not. qWrite) then
          dzd = dzdF + dzdS
          call diffractIM(dzd, qDiffrctd, qOK)
        else
          call diffractIM(dzdF, qDiffrctd, qOK)
          nw = nw + 1
          call gPowerP(pw)
          powers(:, nw) = pw
          qDone = .true.
          if (dzdS > 0.0_wp) call diffractIM(dzdS, qDiffrctd, qOK)
        end if
      end if
    end if
    if (qWrite .and. .not. qDone) then
      nw = nw + 1
      call gPowerP(pw)
      powers(:, nw) = pw
    end if
    if (iper /= 0 .and. redist_every > 0) then        ! :325-331
      if (mod(iCsteps, redist_every) == 0) call wrap_z2()
    end if
  end do
  call deallact_rk4_arrs()
  if (ielem /= 0 .and. status == 0) then
    call matchOut(szl)                                 ! undulator.f90:337
    call correctTrans()                                ! :340
    call gpu_drift(el(1))                              ! driftSection
    call diffractIM(el(1), qDiffrctd, qOK)
    call gpu_quad(el(2), el(3))                        ! Quad
    call gpu_disperse(el(4), el(5))                    ! disperse (no zbar: no diffraction)
    call gpu_bmodulation(el(6), el(7))                 ! bModulation
    ! output reductions of the final state (slices of 4 pi rho)
    sRho_G = tParams_G%rho
    qOneD_G = tParams_G%one_d /= 0
    nsl = ceiling(tParams_G%dz2 * real(tParams_G%nz2 - 1, wp) / (4.0_wp * pi_ref() * sRho_G))
    allocate(tw(nsl, 22), cur(nsl + 1))
    call getSliceTwiss(nsl, 0.0_wp, tw(:,1), tw(:,2), tw(:,3), tw(:,4), tw(:,5), tw(:,6), tw(:,7), &
                       tw(:,8), tw(:,9), tw(:,10), tw(:,11), tw(:,12), tw(:,13), tw(:,14), tw(:,15), &
                       tw(:,16), tw(:,17), tw(:,18), tw(:,19), tw(:,20), tw(:,21), tw(:,22))
    call getCurr(4.0_wp * pi_ref() * sRho_G, cur)
    call deallact_rk4_arrs()
  end if

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) nw
  write(u) sElX_G, sElY_G, sElZ2_G, sElPX_G, sElPY_G, sElGam_G
  write(u) ac_rfield_in, ac_ifield_in
  if (nw > 0) write(u) powers(:, 1:nw)
  if (allocated(tw)) then
    write(u) int(nsl, c_int64_t)
    write(u) tw(:,22), tw(:,15), tw(:,16), tw(:,3), tw(:,7), tw(:,17)
    write(u) cur
  end if
  write(u) status
  close(u)

end program und_section

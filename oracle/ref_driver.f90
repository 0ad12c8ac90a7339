! Test infrastructure only (never shipped, never on the product's path): a
! trace recorder that drives the *reference* FastMC modules, compiled
! from the read-only reference sources by oracle/Makefile, through the GCMC
! insert / delete / displace / translate / rotate sequence and writes every
! trial's outcome to the file TRACE in the working directory.
!
! The start-up restates gcmc.f:207-548 (the main program is not a module, so
! its internal procedures cannot be called); single_point and
! excluding_ewald_charges restate gcmc.f:2549-2761; the move drivers restate
! mc_insert / mc_delete / mc_displace / displace_guest (gcmc.f:1693-1935,
! 2816-2901) and the per-step move choice and step-size adaptation of
! gcmc.f:805-911.  Everything numerical is done by the reference routines
! themselves (insertion, deletion, guest_energy, accept_move, reject_move,
! energy_eval, random_ins, random_disp, random_rot, translate, rotation, ewald1,
! ewald2, ewald3, srfrce, lrcorrect, parlst, erfcgen, duni).
!
! usage: ref_driver.x NSTEPS [SKEVERY]
program ref_driver
  use utility_pack
  use readinputs
  use ewald_module
  use vdw_module
  use mc_moves
  implicit none

  logical lprob, lwidom, lspe, ljob, lrestart, laccsample, lnumg, lfuga
  logical lwanglandau, lblock, loverlap, lnewsurf, linitsurf, accepted
  logical insert, delete, displace, tran, rota
  integer idnode, mxnode, imcon, keyfce, kmax1, kmax2, kmax3, ntpguest, ntprob
  integer ntpsite, ntpvdw, maxmls, mxatm, mxatyp, nwind, nwindsteps, maxvdw
  integer ntpatm, ntpmls, ntpfram, totatm, levcfg, mcsteps, eqsteps, nnumg
  integer nhis, swap_max, nwidstep, visittol, maxn, minn, ebins, numblocks
  integer newld, gridsize, i, j, ik, mol, iguest, nmols, natms, randchoice
  integer nsteps, istep, skevery, movekind, nargs, gcmccount, k
  integer, dimension(3) :: gridfactor
  character*1 cfgname(80)
  character(len=64) :: arg
  real(8) volm, rcut, eps, alpha, initdelr, rvdw, griddim, dlrpot, engunit, sumchg
  real(8) statvolm, det, temp, overlap, surftol, wlprec, flatcoeff, prectol
  real(8) init, rotangle, jumpangle, beta, epsq, drewd, rota_rotangle
  real(8) spenergy, vdwsum, ecoul, dlpeng, delrc, estep, rande, randmov
  real(8) chgtmp, engsictmp, a, b, c, q1, q2, q3, q4, sk1, sk2, sk3, sk4
  real(8), dimension(10) :: celprp
  real(8) :: xyzrec(3, 64)
  integer :: nrec, drew_rande

  idnode = 0
  mxnode = 1
  nargs = command_argument_count()
  nsteps = 1000
  skevery = 100
  if (nargs >= 1) then
    call get_command_argument(1, arg)
    read(arg, *) nsteps
  end if
  if (nargs >= 2) then
    call get_command_argument(2, arg)
    read(arg, *) skevery
  end if

  ! defaults of gcmc.f:140-205
  overlap = 1.5d0
  surftol = 0.d0
  ntpatm = 0
  mxatm = 0
  mxegrd = 0
  nnumg = 1
  griddim = 0.1d0
  gridfactor = (/1, 1, 1/)
  newld = 0
  mcsteps = 1
  eqsteps = 0
  delrc = 0.d0
  flatcoeff = 0.7d0
  visittol = 10000
  wlprec = dexp(1.d0)
  prectol = 1d-8
  maxn = 200
  minn = 0
  ebins = 100
  nwidstep = 20
  lprob = .false.
  lwidom = .false.
  lspe = .false.
  ljob = .false.
  lrestart = .false.
  laccsample = .false.
  lnumg = .false.
  lfuga = .false.
  lwanglandau = .false.
  lblock = .false.
  numblocks = 0

  open(nrite, file='OUTPUT')
  call initscan(idnode, imcon, volm, keyfce, rcut, eps, alpha, kmax1, kmax2, kmax3, lprob, &
       initdelr, rvdw, ntpguest, ntprob, ntpsite, ntpvdw, maxmls, mxatm, mxatyp, &
       griddim, gridfactor, nwind, nwindsteps, lwidom)
  maxvdw = max(ntpvdw, (mxatyp*(mxatyp+1))/2)
  call alloc_config_arrays(idnode, mxnode, maxmls, mxatm, mxatyp, volm, ntpguest, rcut, rvdw, &
       nwind, initdelr)
  delr(:) = initdelr
  disp_ratio(:) = 0.5d0
  tran_ratio(:) = 0.5d0
  rota_ratio(:) = 0.5d0
  call alloc_vdw_arrays(idnode, maxvdw, maxmls, mxatyp)
  call readfield(idnode, ntpvdw, maxvdw, ntpatm, ntpmls, ntpguest, ntpfram, totatm, rvdw, &
       dlrpot, engunit, sumchg)
  call readconfig(idnode, mxnode, imcon, cfgname, levcfg, ntpmls, maxmls, totatm, volm, rcut, celprp)
  call alloc_ewald_arrays(idnode, maxmls, kmax1, kmax2, kmax3, rvdw, totatm, maxguest)
  statvolm = volm*1d-30
  call invert(cell, rcell, det)
  gridsize = 1
  ntprob = 1
  call alloc_prob_arrays(idnode, ntpguest, ntpsite, ntprob, gridsize)
  call readcontrol(idnode, lspe, temp, ljob, mcsteps, eqsteps, ntpguest, lrestart, laccsample, &
       lnumg, nnumg, nhis, swap_max, lfuga, overlap, surftol, nwidstep, lwanglandau, wlprec, &
       flatcoeff, visittol, prectol, maxn, minn, ebins, lblock)
  overlap = overlap**2
  call fugacity(idnode, lfuga, temp, ntpguest)
  ! move-frequency normalisation in REAL(4), gcmc.f:341-366
  do i = 1, ntpguest
    mcmvnorm(i) = mcinsf(i)+mcdelf(i)+mcdisf(i)+mcjmpf(i)+mcswpf(i)+mctraf(i)+mcrotf(i)+mcswif(i)
    if (mcmvnorm(i) .eq. 0.d0) then
      mcmvnorm(i) = 1
      mcinsf(i) = 1.d0/3.d0
      mcdelf(i) = 1.d0/3.d0
      mcdisf(i) = 1.d0/3.d0
      mcjmpf(i) = 0.d0
      mcswpf(i) = 0.d0
      mctraf(i) = 0.d0
      mcrotf(i) = 0.d0
      mcswif(i) = 0.d0
    end if
    mcinsf(i) = mcinsf(i)/mcmvnorm(i)
    mcdelf(i) = mcinsf(i) + (mcdelf(i)/mcmvnorm(i))
    mcdisf(i) = mcdelf(i) + (mcdisf(i)/mcmvnorm(i))
    mcjmpf(i) = mcdisf(i) + (mcjmpf(i)/mcmvnorm(i))
    mcswpf(i) = mcjmpf(i) + (mcswpf(i)/mcmvnorm(i))
    mctraf(i) = mcswpf(i) + (mctraf(i)/mcmvnorm(i))
    mcrotf(i) = mctraf(i) + (mcrotf(i)/mcmvnorm(i))
    mcswif(i) = mcrotf(i) + (mcswif(i)/mcmvnorm(i))
  end do
  init = duni(idnode)
  rotangle = pi/3.d0
  jumpangle = 2.d0*pi
  call guest_exclude(ntpguest)
  call condense(totatm, ntpfram, ntpguest)
  beta = 1.d0/(kboltz*temp)
  epsq = 1.d0
  call erfcgen(keyfce, alpha, rcut, drewd)
  call my_single_point()
  ! initial per-guest energies, gcmc.f:554-582
  do iguest = 1, ntpguest
    mol = locguest(iguest)
    nmols = nummols(mol)
    spenergy = 0.d0
    do j = 1, ntpmls
      ik = loc2(mol, j)
      spenergy = spenergy + (ewald1en(ik) + ewald2en(ik) + elrc_mol(ik) + vdwen(ik))/engunit
    end do
    if (nmols .gt. 0) spenergy = spenergy - nmols*ewald3en(mol)/engunit
    energy(mol) = spenergy
  end do

  open(91, file='TRACE')
  call dump_system()

  do i = 1, ntpguest
    delrdisp(i) = delr(i)
    tran_delr(i) = delr(i)
  end do
  rota_rotangle = rotangle
  gcmccount = 0
  do istep = 1, nsteps
    gcmccount = gcmccount + 1
    insert = .false.
    delete = .false.
    displace = .false.
    tran = .false.
    rota = .false.
    if (ntpguest .gt. 1) then
      iguest = floor(duni(idnode)*ntpguest) + 1
    else
      iguest = 1
    end if
    mol = locguest(iguest)
    natms = numatoms(mol)
    nmols = nummols(mol)
    randmov = duni(idnode)
    if (randmov .lt. mcinsf(iguest)) then
      insert = .true.
    else if (randmov .lt. mcdelf(iguest)) then
      delete = .true.
    else if (randmov .lt. mcdisf(iguest)) then
      displace = .true.
    else if (randmov .lt. mcjmpf(iguest)) then
      stop 'jump move is outside the traced path'
    else if (randmov .lt. mcswpf(iguest)) then
      stop 'swap move is outside the traced path'
    else if (randmov .lt. mctraf(iguest)) then
      displace = .true.
      tran = .true.
    else if (randmov .lt. mcrotf(iguest)) then
      displace = .true.
      rota = .true.
    else if (randmov .lt. mcswif(iguest)) then
      stop 'switch move is outside the traced path'
    else
      displace = .true.
    end if
    ! step-size adaptation, gcmc.f:859-900
    if (nmols .ge. 1 .and. disp_count(iguest) .ge. 1) then
      if (mod(disp_count(iguest), 100) .eq. 0) then
        if ((accept_disp(iguest)/dble(disp_count(iguest))) .gt. disp_ratio(iguest)) then
          delrdisp(iguest) = delrdisp(iguest)*1.05d0
        else
          if (delrdisp(iguest) .gt. delrmin) delrdisp(iguest) = delrdisp(iguest)*0.95d0
        end if
      else if (mod(disp_count(iguest), 50) .eq. 0) then
        if ((accept_disp(iguest)/dble(disp_count(iguest))) .gt. disp_ratio(iguest)) then
          rotangle = rotangle*1.05d0
        else
          if (rotangle .gt. minangle) rotangle = rotangle*0.95d0
        end if
      end if
    end if
    if ((nmols .ge. 1) .and. (tran_count(iguest) .ge. 1) .and. mod(tran_count(iguest), 100) .eq. 0) then
      if ((accept_tran(iguest)/dble(tran_count(iguest))) .gt. tran_ratio(iguest)) then
        tran_delr(iguest) = tran_delr(iguest)*1.05d0
      else
        if (tran_delr(iguest) .gt. delrmin) tran_delr(iguest) = tran_delr(iguest)*0.95d0
      end if
    end if
    if ((nmols .ge. 1) .and. (rota_count(iguest) .ge. 1) .and. mod(rota_count(iguest), 100) .eq. 0) then
      if ((accept_rota(iguest)/dble(rota_count(iguest))) .gt. rota_ratio(iguest)) then
        rota_rotangle = rota_rotangle*1.05d0
      else
        if (rota_rotangle .gt. minangle) rota_rotangle = rota_rotangle*0.95d0
      end if
    end if
    if (nmols .eq. 0) then
      insert = .true.
      delete = .false.
      displace = .false.
      tran = .false.
      rota = .false.
    end if

    rande = -1.d0
    drew_rande = 0
    randchoice = 0
    loverlap = .false.
    accepted = .false.
    estep = 0.d0
    if (insert) then
      movekind = 1
      call do_insert()
    else if (delete) then
      movekind = 2
      call do_delete()
    else
      if (tran) then
        movekind = 4
      else if (rota) then
        movekind = 5
      else
        movekind = 3
      end if
      call do_displace()
    end if
    write(91, '(a,i9,4i4,i6,2i2,es26.17e3,es26.17e3,i2)') 'S ', istep, iguest, movekind, &
         randchoice, nmols, natms, merge(1, 0, loverlap), merge(1, 0, accepted), estep, rande, drew_rande
    write(91, '(a,64es26.17e3)') 'X ', (xyzrec(1, k), xyzrec(2, k), xyzrec(3, k), k = 1, natms)
    write(91, '(a,16es26.17e3)') 'E ', (energy(k), k = 1, maxmls), (delE(k), k = 1, maxmls)
    write(91, '(a,3es26.17e3)') 'A ', delrdisp(iguest), rotangle, rota_rotangle
    if (mod(istep, skevery) .eq. 0) then
      sk1 = 0.d0; sk2 = 0.d0; sk3 = 0.d0; sk4 = 0.d0
      do k = 1, newld
        sk1 = sk1 + ckcsum(maxmls+1, k)
        sk2 = sk2 + ckssum(maxmls+1, k)
        sk3 = sk3 + abs(ckcsum(maxmls+1, k))*dble(mod(k, 7) + 1)
        sk4 = sk4 + abs(ckssum(maxmls+1, k))*dble(mod(k, 5) + 1)
      end do
      write(91, '(a,i9,4es26.17e3)') 'K ', istep, sk1, sk2, sk3, sk4
      write(91, '(a,i9,4i6)') 'N ', istep, (nummols(locguest(k)), k = 1, ntpguest)
    end if
  end do
  close(91)

contains

  subroutine record_new(n)
    integer n, kk
    do kk = 1, n
      xyzrec(1, kk) = newx(kk)
      xyzrec(2, kk) = newy(kk)
      xyzrec(3, kk) = newz(kk)
    end do
  end subroutine record_new

  ! gcmc.f:1693-1752
  subroutine do_insert()
    real(8) gpress
    integer rc
    mol = locguest(iguest)
    natms = numatoms(mol)
    lnewsurf = .false.
    engsicorig = engsic
    ins(iguest) = 1
    ins_count(iguest) = ins_count(iguest) + 1
    call random_ins(idnode, natms, iguest, rcut, delr(iguest))
    call record_new(natms)
    estep = 0.d0
    call insertion(imcon, iguest, keyfce, alpha, rcut, delr(iguest), drewd, totatm, volm, &
         kmax1, kmax2, kmax3, epsq, dlrpot, ntpatm, maxvdw, engunit, delrc, estep, sumchg, &
         chgtmp, engsictmp, maxmls, loverlap, lnewsurf, surftol, overlap, newld, lblock, numblocks)
    accepted = .false.
    if (.not. loverlap) then
      gpress = gstfuga(iguest)
      rande = duni(idnode)
      drew_rande = 1
      call energy_eval(estep, rande, statvolm, iguest, 0, temp, beta, &
           .false., .true., .false., .false., accepted)
    end if
    if (accepted) then
      accept_ins(iguest) = accept_ins(iguest) + 1
      rc = 0
      call accept_move(iguest, .true., .false., .false., lnewsurf, delrc, totatm, rc, ntpfram, &
           ntpguest, maxmls, sumchg, engsictmp, chgtmp, newld)
    else
      call reject_move(iguest, 0, .true., .false., .false., .false.)
    end if
  end subroutine do_insert

  ! gcmc.f:1754-1810
  subroutine do_delete()
    integer nmol
    mol = locguest(iguest)
    nmol = nummols(mol)
    engsicorig = engsic
    linitsurf = .false.
    del(iguest) = 1
    del_count(iguest) = del_count(iguest) + 1
    randchoice = floor(duni(idnode)*nmol) + 1
    estep = 0.d0
    call deletion(imcon, keyfce, iguest, randchoice, alpha, rcut, delr(iguest), drewd, maxmls, &
         totatm, volm, kmax1, kmax2, kmax3, epsq, dlrpot, ntpatm, maxvdw, engunit, delrc, estep, &
         linitsurf, surftol, sumchg, engsictmp, chgtmp, overlap, newld)
    call record_new(numatoms(mol))
    accepted = .false.
    rande = duni(idnode)
    drew_rande = 1
    call energy_eval(-estep, rande, statvolm, iguest, 0, temp, beta, &
         .false., .false., .true., .false., accepted)
    if (accepted) then
      accept_del(iguest) = accept_del(iguest) + 1
      call accept_move(iguest, .false., .true., .false., linitsurf, delrc, totatm, randchoice, &
           ntpfram, ntpguest, maxmls, sumchg, engsictmp, chgtmp, newld)
    else
      call reject_move(iguest, 0, .false., .true., .false., .false.)
    end if
  end subroutine do_delete

  ! gcmc.f:1812-1935
  subroutine do_displace()
    integer nmol
    mol = locguest(iguest)
    nmol = nummols(mol)
    a = 0.d0; b = 0.d0; c = 0.d0; q1 = 0.d0; q2 = 0.d0; q3 = 0.d0; q4 = 0.d0
    if (tran) then
      dis(iguest) = 2
      tran_count(iguest) = tran_count(iguest) + 1
      call random_disp(idnode, tran_delr(iguest), a, b, c)
    else if (rota) then
      dis(iguest) = 3
      rota_count(iguest) = rota_count(iguest) + 1
      call random_rot(idnode, rota_rotangle, q1, q2, q3, q4)
    else
      dis(iguest) = 1
      disp_count(iguest) = disp_count(iguest) + 1
      call random_disp(idnode, delrdisp(iguest), a, b, c)
      call random_rot(idnode, rotangle, q1, q2, q3, q4)
    end if
    do ik = 1, newld
      ckcsorig(mol, ik) = ckcsum(mol, ik)
      ckssorig(mol, ik) = ckssum(mol, ik)
      ckcsorig(maxmls+1, ik) = ckcsum(maxmls+1, ik)
      ckssorig(maxmls+1, ik) = ckssum(maxmls+1, ik)
    end do
    randchoice = floor(duni(idnode)*nmol) + 1
    call my_displace_guest(randchoice)
    accepted = .false.
    if (.not. loverlap) then
      if (estep .lt. 0.d0) then
        accepted = .true.
      else
        rande = duni(idnode)
        drew_rande = 1
        call energy_eval(estep, rande, statvolm, iguest, 0, temp, beta, &
             .true., .false., .false., .false., accepted)
      end if
    end if
    if (accepted) then
      call accept_move(iguest, .false., .false., .true., lnewsurf, delrc, totatm, randchoice, &
           ntpfram, ntpguest, maxmls, sumchg, engsictmp, chgtmp, newld)
      if (tran) then
        accept_tran(iguest) = accept_tran(iguest) + 1
      else if (rota) then
        accept_rota(iguest) = accept_rota(iguest) + 1
      else
        accept_disp(iguest) = accept_disp(iguest) + 1
      end if
    else
      do ik = 1, newld
        ckcsum(mol, ik) = ckcsorig(mol, ik)
        ckssum(mol, ik) = ckssorig(mol, ik)
        ckcsum(maxmls+1, ik) = ckcsorig(maxmls+1, ik)
        ckssum(maxmls+1, ik) = ckssorig(maxmls+1, ik)
      end do
      call reject_move(iguest, 0, .false., .false., .true., .false.)
    end if
  end subroutine do_displace

  ! gcmc.f:2816-2901
  subroutine my_displace_guest(imol)
    integer imol, nmolsl, natmsl, mxcmls, ii
    real(8) comx, comy, comz, enginit, ewld3sum, ewld1eng, ewld2sum, vdwsum2
    call get_guest(iguest, imol, mol, natmsl, nmolsl)
    mxcmls = maxmls*(maxmls-1)/2 + maxmls
    mol = locguest(iguest)
    call guest_energy(imcon, iguest, alpha, rcut, delr(iguest), drewd, totatm, maxmls, volm, &
         kmax1, kmax2, kmax3, epsq, dlrpot, engunit, vdwsum2, ewld2sum, ewld1eng, linitsurf, &
         newld, surftol, sumchg, chgtmp, engsictmp, loverlap, overlap, estep, .true., lblock, &
         numblocks)
    ewld3sum = ewald3en(mol)
    enginit = estep
    do ii = 1, mxcmls
      ewald1entmp(ii) = ewald1en(ii)
      ewald2entmp(ii) = ewald2en(ii)
      vdwentmp(ii) = vdwen(ii)
    end do
    do ii = 1, newld
      ckcsum(mol, ii) = ckcsum(mol, ii) - ckcsnew(mol, ii)
      ckssum(mol, ii) = ckssum(mol, ii) - ckssnew(mol, ii)
      ckcsum(maxmls+1, ii) = ckcsum(maxmls+1, ii) - ckcsnew(maxmls+1, ii)
      ckssum(maxmls+1, ii) = ckssum(maxmls+1, ii) - ckssnew(maxmls+1, ii)
    end do
    call translate(natmsl, mol, newx, newy, newz, cell, rcell, comx, comy, comz, a, b, c)
    call rotation(newx, newy, newz, comx, comy, comz, natmsl, q1, q2, q3, q4)
    call record_new(natmsl)
    call guest_energy(imcon, iguest, alpha, rcut, delr(iguest), drewd, totatm, maxmls, volm, &
         kmax1, kmax2, kmax3, epsq, dlrpot, engunit, vdwsum2, ewld2sum, ewld1eng, lnewsurf, &
         newld, surftol, sumchg, chgtmp, engsictmp, loverlap, overlap, estep, .false., lblock, &
         numblocks)
    estep = estep - enginit
    do ii = 1, maxmls
      ik = loc2(mol, ii)
      if (ii .ne. mol) delE(ii) = delE(ii) + (ewald1en(ik) - ewald1entmp(ik) + ewald2en(ik) - &
           ewald2entmp(ik) + vdwen(ik) - vdwentmp(ik))/engunit
    end do
    delE(mol) = delE(mol) + estep
  end subroutine my_displace_guest

  ! gcmc.f:2549-2661
  subroutine my_single_point()
    logical latmsurf, lmolsurf, lov
    integer ii, jj, iatm, imol, jatm, sittyp, ikk
    real(8) ewald1sum, ewald2sum, vdwsum1, ewald2eng, vdweng, ewald3eng, chg, engsic1
    ewald2sum = 0.d0
    ewald3eng = 0.d0
    vdwsum1 = 0.d0
    do ikk = 1, maxmls
      vdwen(ikk) = 0.d0
      ewald2en(ikk) = 0.d0
      ewald3en(ikk) = 0.d0
      ewald1en(ikk) = 0.d0
      delrc_mol(ikk) = 0.d0
      elrc_mol(ikk) = 0.d0
    end do
    call lrcorrect(imcon, keyfce, totatm, ntpatm, maxvdw, rcut, volm, maxmls, ntpguest)
    call condense(totatm, ntpfram, ntpguest)
    call parlst(imcon, totatm, rcut, initdelr)
    call ewald1(imcon, ewald1sum, engsic1, totatm, volm, alpha, sumchg, kmax1, kmax2, kmax3, &
         epsq, newld, maxmls)
    call my_excluding_ewald_charges(ewald3eng)
    ii = 0
    lov = .false.
    do mol = 1, ntpmls
      if (nummols(mol) .gt. 0) then
        do imol = 1, nummols(mol)
          lmolsurf = .false.
          do iatm = 1, numatoms(mol)
            ii = ii + 1
            ikk = 0
            if (lentry(ii) .gt. 0) then
              do jj = 1, lentry(ii)
                ikk = ikk + 1
                jatm = list(ii, jj)
                ilist(jj) = jatm
                moldf(jj) = moltype(jatm)
                xdf(ikk) = xxx(ii) - xxx(jatm)
                ydf(ikk) = yyy(ii) - yyy(jatm)
                zdf(ikk) = zzz(ii) - zzz(jatm)
              end do
              call images(imcon, ikk, cell, xdf, ydf, zdf)
              do jj = 1, lentry(ii)
                rsqdf(jj) = xdf(jj)**2 + ydf(jj)**2 + zdf(jj)**2
                call surface_check(ii, jj, mol, surftol, overlap, lov, latmsurf)
                if (latmsurf) lmolsurf = .true.
              end do
              chg = atmcharge(ii)
              sittyp = ltype(ii)
              call ewald2(chg, lentry(ii), ewald2eng, mol, drewd, rcut, epsq)
              ewald2sum = ewald2sum + ewald2eng
              call srfrce(sittyp, lentry(ii), mol, vdweng, rcut, dlrpot)
              vdwsum1 = vdwsum1 + vdweng
            end if
          end do
          if (lmolsurf) surfacemols(mol) = surfacemols(mol) + 1
        end do
      end if
    end do
    dlpeng = (ewald1sum + ewald2sum + vdwsum1 - ewald3eng + elrc)
    ecoul = ewald2sum + ewald1sum - ewald3eng
    vdwsum = vdwsum1
    write(nrite, '(a,es26.17e3)') 'DRIVER Electrostatic energy ', ecoul/engunit
    write(nrite, '(a,es26.17e3)') 'DRIVER recip energy ', ewald1sum/engunit
    write(nrite, '(a,es26.17e3)') 'DRIVER vdw energy ', (vdwsum1 + elrc)/engunit
  end subroutine my_single_point

  ! gcmc.f:2662-2761 (the framework branch only feeds ewald3en of the framework)
  subroutine my_excluding_ewald_charges(ewald3sum)
    logical lguest, lfrzi, lfrzj
    integer nm, mm, imol, iatm, itgst, gstmol, katm, jatm, jmol, ii, jj, ikk, na
    real(8) ewald3mol, ewald3sum, engcpe, chg
    ewald3sum = 0.d0
    ewald3en(:) = 0.d0
    do mm = 1, ntpmls
      lguest = .false.
      ewald3mol = 0.d0
      nm = nummols(mm)
      na = numatoms(mm)
      do itgst = 1, ntpguest
        gstmol = locguest(itgst)
        if (gstmol .eq. mm) then
          lguest = .true.
          do iatm = 1, na - 1
            ikk = 0
            do jatm = iatm + 1, na
              ikk = ikk + 1
              jlist(ikk) = jatm
              xdf(ikk) = guestx(itgst, iatm) - guestx(itgst, jatm)
              ydf(ikk) = guesty(itgst, iatm) - guesty(itgst, jatm)
              zdf(ikk) = guestz(itgst, iatm) - guestz(itgst, jatm)
            end do
            call images(imcon, ikk, cell, xdf, ydf, zdf)
            chg = atmchg(mm, iatm)
            call ewald3(chg, mm, ikk, alpha, engcpe, epsq)
            ewald3mol = ewald3mol + engcpe
            ewald3en(mm) = ewald3en(mm) + engcpe
          end do
          ewald3sum = ewald3sum + nm*ewald3mol
        end if
      end do
      iatm = 0
      if (.not. lguest) then
        do imol = 1, nm
          do ii = 1, na
            iatm = iatm + 1
            jatm = iatm
            ikk = 0
            lfrzi = (lfzsite(mm, ii) .ne. 0)
            do jmol = imol, nm
              katm = 1
              if (imol .eq. jmol) katm = iatm + 1
              do jj = katm, na
                jatm = jatm + 1
                lfrzj = (lfzsite(mm, jj) .ne. 0)
                if ((lfrzi) .and. (lfrzj)) then
                  ikk = ikk + 1
                  jlist(ikk) = jatm
                  xdf(ikk) = molxxx(mm, iatm) - molxxx(mm, jatm)
                  ydf(ikk) = molyyy(mm, iatm) - molyyy(mm, jatm)
                  zdf(ikk) = molzzz(mm, iatm) - molzzz(mm, jatm)
                end if
              end do
            end do
            call images(imcon, ikk, cell, xdf, ydf, zdf)
            chg = atmchg(mm, iatm)
            call ewald3(chg, mm, ikk, alpha, engcpe, epsq)
            ewald3en(mm) = ewald3en(mm) + engcpe
          end do
        end do
      end if
    end do
  end subroutine my_excluding_ewald_charges

  subroutine dump_system()
    integer kk, ll, nvv
    write(91, '(a,10i8)') 'I ', imcon, kmax1, kmax2, kmax3, newld, mxegrd, maxmls, ntpguest, ntpatm, maxvdw
    write(91, '(a,i8,2i8)') 'J ', totatm, ntpvdw, ntpfram
    write(91, '(a,12es26.17e3)') 'R ', volm, alpha, rcut, delr(1), drewd, dlrpot, engunit, sumchg, &
         temp, overlap, beta, statvolm
    write(91, '(a,9es26.17e3)') 'C ', (cell(kk), kk = 1, 9)
    write(91, '(a,9es26.17e3)') 'Q ', (rcell(kk), kk = 1, 9)
    write(91, '(a,8es26.17e3)') 'M ', mcinsf(1), mcdelf(1), mcdisf(1), mcjmpf(1), mcswpf(1), &
         mctraf(1), mcrotf(1), mcswif(1)
    write(91, '(a,4es26.17e3)') 'G ', (gstfuga(kk), kk = 1, ntpguest)
    do kk = 1, mxegrd
      if (mod(kk, 37) .eq. 1 .or. kk .le. 3 .or. kk .ge. mxegrd - 3) then
        write(91, '(a,i6,es26.17e3)') 'T ', kk, erc(kk)
      end if
    end do
    nvv = ntpvdw
    do ll = 1, nvv
      write(91, '(a,2i6,3es26.17e3)') 'P ', ll, ltpvdw(ll), prmvdw(ll, 1), prmvdw(ll, 2), steadd(ll)
      do kk = 1, mxegrd
        if (mod(kk, 97) .eq. 5 .or. kk .le. 2 .or. kk .ge. mxegrd - 2) then
          write(91, '(a,2i6,es26.17e3)') 'V ', ll, kk, vvv(kk, ll)
        end if
      end do
    end do
    do kk = 1, (ntpatm*(ntpatm+1))/2
      write(91, '(a,2i6)') 'L ', kk, lstvdw(kk)
    end do
    do kk = 1, maxmls
      write(91, '(a,i4,6es26.17e3)') 'W ', kk, energy(kk), ewald3en(kk), chgsum_mol(kk), &
           engsic(loc2(kk, kk))
    end do
    write(91, '(a,3es26.17e3)') 'Z ', elrc, (elrc_mol(kk), kk = 1, 2)
    do kk = 1, newld
      if (mod(kk, 23) .eq. 1 .or. kk .eq. newld) then
        write(91, '(a,i7,4es26.17e3)') 'K0 ', kk, ckcsum(maxmls+1, kk), ckssum(maxmls+1, kk), &
             ckcsum(maxmls, kk), ckssum(maxmls, kk)
      end if
    end do
    do kk = 1, ntpguest
      write(91, '(a,3i4,3i4)') 'U ', kk, locguest(kk), numatoms(locguest(kk)), &
           (ltpsit(locguest(kk), ll), ll = 1, min(3, numatoms(locguest(kk))))
      do ll = 1, numatoms(locguest(kk))
        write(91, '(a,2i4,5es26.17e3)') 'UG ', kk, ll, guestx(kk, ll), guesty(kk, ll), guestz(kk, ll), &
             atmchg(locguest(kk), ll), atmwght(locguest(kk), ll)
      end do
    end do
    do kk = 1, totatm
      if (mod(kk, 11) .eq. 1 .or. kk .eq. totatm) then
        write(91, '(a,2i7,4es26.17e3)') 'F ', kk, ltype(kk), xxx(kk), yyy(kk), zzz(kk), atmcharge(kk)
      end if
    end do
  end subroutine dump_system

end program ref_driver

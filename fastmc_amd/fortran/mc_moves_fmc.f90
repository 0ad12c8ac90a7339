! Drop-in replacement for FastMC's module mc_moves (mc_moves.f): the same
! module name, public procedures and dummy argument lists, so gcmc.f,
! readinputs.f and the CONFIG/FIELD/CONTROL inputs stay as they are.  The
! trial energies of insertion, deletion and guest_energy are evaluated by
! libfastmc_hip.so (include/fastmc.h) on the GPU; this module keeps the host
! bookkeeping that gcmc.f reads (coordinates, counters, energies, tails).
!
! Build: compile against the reference's module files in place of
! mc_moves.f and link libfastmc_hip.so (see oracle/Makefile, target dropin).
!
! The device handle is created at the first trial from the module globals the
! reference's readers filled (readinputs.f, utility_pack.f, vdw_module.f,
! ewald_module.f) and follows the chain through accept_move / reject_move.
! Environment:
!   FMC_SHIM_PRECISION=mixed   FP32 pair/phase arithmetic (default FP64)
!   FMC_SHIM_TRACE=<file>      one line per trial and per decision
!
! Not mirrored to the host: the trial structure factors ckcsnew/ckssnew and
! the running ckcsum/ckssum (S(k) lives on the device).  gcmc.f also writes
! the host arrays itself: it backs up and restores S(k) around displacements
! (mc_displace gcmc.f:1863-1930; the device trial leaves S(k) unchanged, so
! nothing is needed), and after a rejected swap it restores the coordinates,
! nummols and S(k) of both guest kinds (gcmc.f:2416-2437) after the deletion
! and the insertion were accepted through this module.  The module therefore
! keeps a shadow of the molecules the device holds; before each trial it
! compares the host's molecule lists with it and reloads a guest kind whose
! list changed behind its back (fmc_load_guests, which re-sums S(k) in FP64).
! An accept of an overlapping insertion (mc_swap and mc_switch accept it and
! then reject the whole move) leaves the device as it is.  mc_switch with
! three or more guest kinds restores host arrays between its own trials, with
! the stale-S(k) energies of the reference in between: such runs are refused
! at start-up.  With one or two guest kinds mc_switch never moves
! (gcmc.f:2098).
! The surface-atom flags (surface_check, mc_moves.f:10-44) come from the
! device trial; for the molecule in place of a displacement that takes one
! extra zero-displacement trial, whose flag bit 1 is the position's.
module mc_moves
  use, intrinsic :: iso_c_binding
  use readinputs
  use ewald_module
  use vdw_module
  use utility_pack
  implicit none

  ! ---- bind(C) mirrors of the structs of include/fastmc.h
  type, bind(C) :: fmc_species_c
    integer(c_int) :: is_guest, natoms, nmols
    type(c_ptr) :: site_charge, site_mass, site_type, site_frozen, template_xyz, xyz
  end type fmc_species_c

  type, bind(C) :: fmc_system_c
    integer(c_int) :: imcon
    real(c_double) :: cell(9)
    integer(c_int) :: nspecies
    type(c_ptr) :: species
    integer(c_int) :: ntypes, npairs
    type(c_ptr) :: pair_i, pair_j, pair_kind, pair_eps, pair_sig
    real(c_double) :: rcut, rvdw, delr, alpha
    integer(c_int) :: kmax(3)
    real(c_double) :: engunit, epsq, overlap, surftol, sumchg
  end type fmc_system_c

  type, bind(C) :: fmc_walker_c
    real(c_double) :: temperature
    real(c_double) :: fugacity(8)
    real(c_float) :: move_cum(8, 8)
    real(c_double) :: delr(8)
    integer(c_int) :: seed(4)
    type(c_ptr) :: framework_xyz
  end type fmc_walker_c

  type, bind(C) :: fmc_trial_c
    real(c_double) :: estep
    real(c_double) :: dele(8)
    real(c_double) :: recip, creal, vdw, tail
    integer(c_int) :: overlap, surface
  end type fmc_trial_c

  interface
    integer(c_int) function fmc_create(sys, nwalkers, walkers, precision, out) bind(C, name='fmc_create')
      import :: c_int, c_ptr, fmc_system_c, fmc_walker_c
      type(fmc_system_c), intent(in) :: sys
      integer(c_int), value :: nwalkers
      type(fmc_walker_c), intent(in) :: walkers
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: out
    end function fmc_create
    integer(c_int) function fmc_trial_insert(h, w, iguest, xyz, out) bind(C, name='fmc_trial_insert')
      import :: c_int, c_ptr, c_double, fmc_trial_c
      type(c_ptr), value :: h
      integer(c_int), value :: w, iguest
      real(c_double), intent(in) :: xyz(*)
      type(fmc_trial_c), intent(out) :: out
    end function fmc_trial_insert
    integer(c_int) function fmc_trial_delete(h, w, iguest, choice, out) bind(C, name='fmc_trial_delete')
      import :: c_int, c_ptr, fmc_trial_c
      type(c_ptr), value :: h
      integer(c_int), value :: w, iguest, choice
      type(fmc_trial_c), intent(out) :: out
    end function fmc_trial_delete
    integer(c_int) function fmc_trial_displace(h, w, iguest, choice, xyz, out) bind(C, name='fmc_trial_displace')
      import :: c_int, c_ptr, c_double, fmc_trial_c
      type(c_ptr), value :: h
      integer(c_int), value :: w, iguest, choice
      real(c_double), intent(in) :: xyz(*)
      type(fmc_trial_c), intent(out) :: out
    end function fmc_trial_displace
    integer(c_int) function fmc_accept(h, w) bind(C, name='fmc_accept')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: w
    end function fmc_accept
    integer(c_int) function fmc_reject(h, w) bind(C, name='fmc_reject')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: w
    end function fmc_reject
    integer(c_int) function fmc_load_guests(h, w, iguest, xyz, nmols, sumchg) bind(C, name='fmc_load_guests')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: w, iguest, nmols
      real(c_double), intent(in) :: xyz(*)
      real(c_double), intent(in) :: sumchg
    end function fmc_load_guests
  end interface

  type(c_ptr), save :: fmc_h = c_null_ptr
  integer, save :: fmc_pending = 0          ! 1 while a device trial awaits accept/reject
  logical, save :: fmc_pend_overlap = .false. ! the pending trial is an overlapping insertion
  ! the molecules the device holds, per species (guests only): count and sites
  integer, allocatable, save :: fmc_shn(:)
  real(8), allocatable, save :: fmc_shx(:, :), fmc_shy(:, :), fmc_shz(:, :)
  integer, save :: fmc_disp_iguest = 0      ! guest_energy(lexisting) seen: molecule to displace
  integer, save :: fmc_disp_choice = 0
  integer, save :: fmc_trace_unit = 0
  logical, save :: fmc_tracing = .false.
  ! storage the device handle was built from
  type(fmc_species_c), allocatable, target, save :: fmc_sp(:)
  real(c_double), allocatable, target, save :: fmc_q(:, :), fmc_m(:, :), fmc_tm(:, :), fmc_xyz(:, :)
  integer(c_int), allocatable, target, save :: fmc_t(:, :), fmc_f(:, :)
  integer(c_int), allocatable, target, save :: fmc_pi(:), fmc_pj(:), fmc_pk(:)
  real(c_double), allocatable, target, save :: fmc_pe(:), fmc_ps(:)

contains

  ! ---------------------------------------------------------------- setup --
  subroutine fmc_fail(where, rc)
    character(*), intent(in) :: where
    integer(c_int), intent(in) :: rc
    write(*, '(a,a,a,i6)') 'fastmc: ', where, ' returned error code ', rc
    call error(0, int(rc))
  end subroutine fmc_fail

  subroutine fmc_shim_init(imcon, maxmls, alpha, kmax1, kmax2, kmax3, rcut, ddelr, epsq, engunit, &
                           overlap, surftol, sumchg, dlrpot)
    integer, intent(in) :: imcon, maxmls, kmax1, kmax2, kmax3
    real(8), intent(in) :: alpha, rcut, ddelr, epsq, engunit, overlap, surftol, sumchg, dlrpot
    type(fmc_system_c) :: sys
    type(fmc_walker_c) :: wk
    integer :: ntpguest, m, g, i, j, k, ntypes, nsite, nat, np, isg(maxmls)
    integer(c_int) :: rc, prec
    character(len=256) :: env
    integer :: envlen, envst

    ntpguest = size(guestx, 1)
    isg = 0
    do g = 1, ntpguest
      isg(locguest(g)) = g
    enddo
    ntypes = 0
    nsite = 1
    do m = 1, maxmls
      nsite = max(nsite, numatoms(m) * max(nummols(m), 1))
      do i = 1, numatoms(m)
        ntypes = max(ntypes, ltpsit(m, i))
      enddo
    enddo
    allocate(fmc_sp(maxmls))
    allocate(fmc_q(nsite, maxmls), fmc_m(nsite, maxmls), fmc_t(nsite, maxmls), fmc_f(nsite, maxmls))
    allocate(fmc_tm(3 * nsite, maxmls), fmc_xyz(3 * nsite, maxmls))
    fmc_tm = 0.d0
    fmc_xyz = 0.d0
    do m = 1, maxmls
      nat = numatoms(m)
      do i = 1, nat
        fmc_q(i, m) = atmchg(m, i)
        fmc_m(i, m) = atmwght(m, i)
        fmc_t(i, m) = ltpsit(m, i)
        fmc_f(i, m) = lfzsite(m, i)
      enddo
      g = isg(m)
      if (g > 0) then
        do i = 1, nat
          fmc_tm(3 * i - 2, m) = guestx(g, i)
          fmc_tm(3 * i - 1, m) = guesty(g, i)
          fmc_tm(3 * i, m) = guestz(g, i)
        enddo
      endif
      do i = 1, nat * nummols(m)
        fmc_xyz(3 * i - 2, m) = molxxx(m, i)
        fmc_xyz(3 * i - 1, m) = molyyy(m, i)
        fmc_xyz(3 * i, m) = molzzz(m, i)
      enddo
      fmc_sp(m)%is_guest = merge(1, 0, g > 0)
      fmc_sp(m)%natoms = nat
      fmc_sp(m)%nmols = nummols(m)
      fmc_sp(m)%site_charge = c_loc(fmc_q(1, m))
      fmc_sp(m)%site_mass = c_loc(fmc_m(1, m))
      fmc_sp(m)%site_type = c_loc(fmc_t(1, m))
      fmc_sp(m)%site_frozen = c_loc(fmc_f(1, m))
      fmc_sp(m)%template_xyz = c_loc(fmc_tm(1, m))
      fmc_sp(m)%xyz = c_loc(fmc_xyz(1, m))
    enddo
    ! one LJ row per type pair that has a 12-6 entry (lstvdw / ltpvdw / prmvdw, vdw_terms.f:1-179)
    allocate(fmc_pi(ntypes * (ntypes + 1) / 2), fmc_pj(ntypes * (ntypes + 1) / 2))
    allocate(fmc_pk(ntypes * (ntypes + 1) / 2))
    allocate(fmc_pe(ntypes * (ntypes + 1) / 2), fmc_ps(ntypes * (ntypes + 1) / 2))
    np = 0
    do i = 1, ntypes
      do j = 1, i
        k = lstvdw(loc2(i, j))
        if (k < 1 .or. k > size(ltpvdw)) cycle
        if (ltpvdw(k) /= 2) cycle
        np = np + 1
        fmc_pi(np) = i
        fmc_pj(np) = j
        fmc_pk(np) = 2
        fmc_pe(np) = prmvdw(k, 1)
        fmc_ps(np) = prmvdw(k, 2)
      enddo
    enddo
    sys%imcon = imcon
    sys%cell = cell
    sys%nspecies = maxmls
    sys%species = c_loc(fmc_sp(1))
    sys%ntypes = ntypes
    sys%npairs = np
    sys%pair_i = c_loc(fmc_pi(1))
    sys%pair_j = c_loc(fmc_pj(1))
    sys%pair_kind = c_loc(fmc_pk(1))
    sys%pair_eps = c_loc(fmc_pe(1))
    sys%pair_sig = c_loc(fmc_ps(1))
    sys%rcut = rcut
    ! the table grid: dlrpot = rvdw / (mxegrd - 4), vdw_terms.f:202
    sys%rvdw = dlrpot * dble(mxegrd - 4)
    sys%delr = ddelr
    sys%alpha = alpha
    sys%kmax = (/ kmax1, kmax2, kmax3 /)
    sys%engunit = engunit
    sys%epsq = epsq
    sys%overlap = sqrt(overlap)   ! gcmc.f:325 squared it
    sys%surftol = surftol
    sys%sumchg = sumchg
    wk%temperature = 300.d0       ! the seam never draws: temperature, moves and seeds are unused
    wk%fugacity = 0.d0
    wk%move_cum = 0.0
    wk%delr = ddelr
    wk%seed = (/ 12, 34, 56, 78 /)
    wk%framework_xyz = c_null_ptr
    prec = 0
    call get_environment_variable('FMC_SHIM_PRECISION', env, envlen, envst)
    if (envst == 0 .and. envlen > 0) then
      if (env(1:5) == 'mixed') prec = 1
    endif
    ! mc_switch moves molecules only with three or more guest kinds, and then
    ! restores the host arrays between its own trials (see the header)
    if (ntpguest >= 3) then
      do g = 1, ntpguest
        if (mcswif(g) > mcrotf(g)) then
          write(*, '(a)') 'fastmc: switch moves with three or more guest types are not supported'
          call fmc_fail('fmc_shim_init', 9001_c_int)
        endif
      enddo
    endif
    rc = fmc_create(sys, 1_c_int, wk, prec, fmc_h)
    if (rc /= 0) call fmc_fail('fmc_create', rc)
    allocate(fmc_shn(maxmls))
    allocate(fmc_shx(size(molxxx, 2), maxmls), fmc_shy(size(molxxx, 2), maxmls), fmc_shz(size(molxxx, 2), maxmls))
    fmc_shn = 0
    do g = 1, ntpguest
      call fmc_shadow_take(locguest(g))
    enddo
    call get_environment_variable('FMC_SHIM_TRACE', env, envlen, envst)
    if (envst == 0 .and. envlen > 0) then
      open(newunit=fmc_trace_unit, file=env(1:envlen), status='replace', action='write')
      fmc_tracing = .true.
    endif
  end subroutine fmc_shim_init

  subroutine fmc_pack(natms, xyz)
    integer, intent(in) :: natms
    real(c_double), intent(out) :: xyz(3 * natms)
    integer :: i
    do i = 1, natms
      xyz(3 * i - 2) = newx(i)
      xyz(3 * i - 1) = newy(i)
      xyz(3 * i) = newz(i)
    enddo
  end subroutine fmc_pack

  subroutine fmc_log_trial(kind, iguest, choice, natms, xyz, t)
    integer, intent(in) :: kind, iguest, choice, natms
    real(c_double), intent(in) :: xyz(3 * natms)
    type(fmc_trial_c), intent(in) :: t
    if (.not. fmc_tracing) return
    write(fmc_trace_unit, '(a,4i8,1x,es24.16e3,i3,*(1x,es24.16e3))') 'T', kind, iguest, choice, natms, &
      t%estep, t%overlap, xyz
  end subroutine fmc_log_trial

  ! sum of site charges and the self-interaction term of ewald1_guest
  ! (ewald.f:96-105, 322-325)
  subroutine fmc_selfterms(mol, natms, alpha, epsq, lexisting, chgtmp, engsictmp)
    integer, intent(in) :: mol, natms
    real(8), intent(in) :: alpha, epsq
    logical, intent(in) :: lexisting
    real(8), intent(out) :: chgtmp, engsictmp
    integer :: n
    real(8) :: q
    chgtmp = 0.d0
    engsictmp = 0.d0
    do n = 1, natms
      q = atmchg(mol, n)
      chgtmp = chgtmp + q
      engsictmp = engsictmp + q * q
    enddo
    engsictmp = -r4pie0 / epsq * alpha * engsictmp / sqrpi
    if (lexisting) engsictmp = -engsictmp
  end subroutine fmc_selfterms

  subroutine fmc_zero_terms(maxmls)
    integer, intent(in) :: maxmls
    integer :: mxcmls
    mxcmls = maxmls * (maxmls - 1) / 2 + maxmls
    delE(1:mxcmls) = 0.d0
    ewald1en(1:mxcmls) = 0.d0
    ewald2en(1:mxcmls) = 0.d0
    vdwen(1:mxcmls) = 0.d0
    delrc_mol(1:mxcmls) = 0.d0
  end subroutine fmc_zero_terms

  ! ------------------------------------------------------------- checks --
  ! Overlap and surface-atom tests of one neighbour pair (rsqdf(j), ilist(j)),
  ! as used by single_point; semantics of mc_moves.f:10-44.
  subroutine surface_check(iatm, j, mol, surftol, overlap, loverlap, latmsurf)
    integer :: iatm, j, mol
    real(8) :: surftol, overlap
    logical :: loverlap, latmsurf
    integer :: ta, tb, jatm, ivdw
    real(8) :: reach
    latmsurf = .false.
    jatm = ilist(j)
    ta = ltpsit(mol, iatm)
    tb = ltype(jatm)
    ivdw = lstvdw(loc2(ta, tb))
    reach = (surftol + prmvdw(ivdw, 2))**2
    if (rsqdf(j) < reach .and. lfreezesite(jatm) == 1) latmsurf = .true.
    if (rsqdf(j) < overlap) loverlap = .true.
  end subroutine surface_check

  ! Pore-blocking test of site iatm (mc_moves.f:46-82); the reference reads
  ! the squared distance from xdf/ydf/zdf into an integer, and so does this.
  subroutine poreblock_check(iatm, imcon, numblocks, loverlap)
    integer :: iatm, imcon, numblocks
    logical :: loverlap
    integer :: b, rsq
    do b = 1, numblocks
      bxdf(b) = newx(iatm) - bxxx(b)
      bydf(b) = newy(iatm) - byyy(b)
      bzdf(b) = newz(iatm) - bzzz(b)
    enddo
    call images(imcon, numblocks, cell, bxdf, bydf, bzdf)
    do b = 1, numblocks
      rsq = int(xdf(b)**2 + ydf(b)**2 + zdf(b)**2)
      if (rsq - bradii(b) <= 0.d0) then
        loverlap = .true.
        exit
      endif
    enddo
  end subroutine poreblock_check

  ! the device now holds the host's molecule list of species mol
  subroutine fmc_shadow_take(mol)
    integer, intent(in) :: mol
    integer :: n
    n = numatoms(mol) * nummols(mol)
    fmc_shn(mol) = nummols(mol)
    fmc_shx(1:n, mol) = molxxx(mol, 1:n)
    fmc_shy(1:n, mol) = molyyy(mol, 1:n)
    fmc_shz(1:n, mol) = molzzz(mol, 1:n)
  end subroutine fmc_shadow_take

  ! Before a trial: reload every guest kind whose host molecule list is not
  ! the one the device holds (gcmc.f restored it, e.g. after a rejected swap).
  subroutine fmc_sync(sumchg)
    real(8), intent(in) :: sumchg
    integer :: g, mol, n, a
    integer(c_int) :: rc
    logical :: same
    real(c_double), allocatable :: xyz(:)
    do g = 1, size(guestx, 1)
      mol = locguest(g)
      n = numatoms(mol) * nummols(mol)
      same = nummols(mol) == fmc_shn(mol)
      if (same) same = all(molxxx(mol, 1:n) == fmc_shx(1:n, mol)) .and. &
                       all(molyyy(mol, 1:n) == fmc_shy(1:n, mol)) .and. &
                       all(molzzz(mol, 1:n) == fmc_shz(1:n, mol))
      if (same) cycle
      allocate(xyz(3 * max(n, 1)))
      do a = 1, n
        xyz(3 * a - 2) = molxxx(mol, a)
        xyz(3 * a - 1) = molyyy(mol, a)
        xyz(3 * a) = molzzz(mol, a)
      enddo
      rc = fmc_load_guests(fmc_h, 0_c_int, int(g, c_int), xyz, int(nummols(mol), c_int), sumchg)
      deallocate(xyz)
      if (rc /= 0) call fmc_fail('fmc_load_guests', rc)
      call fmc_shadow_take(mol)
      if (fmc_tracing) write(fmc_trace_unit, '(a,2i8)') 'L', g, nummols(mol)
    enddo
  end subroutine fmc_sync

  ! -------------------------------------------------------------- trials --
  ! Initial loading to nguests molecules with random non-overlapping
  ! insertions, accepted regardless of energy (mc_moves.f:84-149).
  subroutine insert_guests(idnode, imcon, totatm, ntpguest, ntpfram, iguest, nguests, rcut, ddelr, &
                           sumchg, surftol, overlap, keyfce, alpha, drewd, volm, newld, kmax1, kmax2, &
                           kmax3, epsq, dlrpot, ntpatm, maxvdw, engunit, delrc, maxmls, iter, lblock, &
                           numblocks)
    integer :: idnode, imcon, totatm, ntpguest, ntpfram, iguest, nguests, keyfce, newld
    integer :: kmax1, kmax2, kmax3, ntpatm, maxvdw, maxmls, iter, numblocks
    real(8) :: rcut, ddelr, sumchg, surftol, overlap, alpha, drewd, volm, epsq, dlrpot, engunit, delrc
    logical :: lblock
    logical :: loverlap, lnewsurf
    integer :: mol, natms, maxiter, idum
    real(8) :: estep, chgtmp, engsictmp
    mol = locguest(iguest)
    natms = numatoms(mol)
    maxiter = 1000 * nguests
    iter = 0
    do while (nummols(mol) < nguests)
      if (iter >= maxiter) call error(idnode, 2320)
      call random_ins(idnode, natms, iguest, rcut, ddelr)
      estep = 0.d0
      engsicorig = engsic
      call insertion(imcon, iguest, keyfce, alpha, rcut, ddelr, drewd, totatm, volm, kmax1, kmax2, &
                     kmax3, epsq, dlrpot, ntpatm, maxvdw, engunit, delrc, estep, sumchg, chgtmp, &
                     engsictmp, maxmls, loverlap, lnewsurf, surftol, overlap, newld, lblock, numblocks)
      if (.not. loverlap) then
        call accept_move(iguest, .true., .false., .false., lnewsurf, delrc, totatm, idum, ntpfram, &
                         ntpguest, maxmls, sumchg, engsictmp, chgtmp, newld)
      else
        call reject_move(iguest, 0, .true., .false., .false., .false.)
        engsic = engsicorig
      endif
      delE(:) = 0.d0
      iter = iter + 1
    enddo
  end subroutine insert_guests

  ! Energy of one more guest at newx/newy/newz (mc_moves.f:150-276).
  subroutine insertion(imcon, iguest, keyfce, alpha, rcut, ddelr, drewd, totatm, volm, kmax1, kmax2, &
                       kmax3, epsq, dlrpot, ntpatm, maxvdw, engunit, delrc, estep, sumchg, chgtmp, &
                       engsictmp, maxmls, loverlap, lnewsurf, surftol, overlap, newld, lblock, numblocks)
    integer :: imcon, iguest, keyfce, totatm, kmax1, kmax2, kmax3, ntpatm, maxvdw, maxmls, newld
    integer :: numblocks
    real(8) :: alpha, rcut, ddelr, drewd, volm, epsq, dlrpot, engunit, delrc, estep, sumchg, chgtmp
    real(8) :: engsictmp, surftol, overlap
    logical :: loverlap, lnewsurf, lblock
    type(fmc_trial_c) :: t
    real(c_double) :: xyz(3 * mxguestsite)
    integer :: mol, natms, i
    integer(c_int) :: rc
    if (lblock) call error(0, 9001)   ! pore blocking is not on the device path
    if (.not. c_associated(fmc_h)) call fmc_shim_init(imcon, maxmls, alpha, kmax1, kmax2, kmax3, rcut, &
                                                      ddelr, epsq, engunit, overlap, surftol, sumchg, dlrpot)
    mol = locguest(iguest)
    natms = numatoms(mol)
    lnewsurf = .false.
    loverlap = .false.
    call fmc_sync(sumchg)
    call fmc_zero_terms(maxmls)
    call gstlrcorrect(imcon, iguest, keyfce, ntpatm, maxvdw, delrc, volm, maxmls, .false.)
    call fmc_pack(natms, xyz)
    rc = fmc_trial_insert(fmc_h, 0_c_int, int(iguest, c_int), xyz, t)
    if (rc /= 0) call fmc_fail('fmc_trial_insert', rc)
    fmc_pending = 1
    fmc_pend_overlap = t%overlap /= 0
    call fmc_log_trial(1, iguest, 0, natms, xyz, t)
    loverlap = t%overlap /= 0
    lnewsurf = btest(t%surface, 0)
    if (loverlap) return
    call fmc_selfterms(mol, natms, alpha, epsq, .false., chgtmp, engsictmp)
    estep = estep + t%estep
    do i = 1, maxmls
      if (i == mol) cycle
      ewald1en(loc2(mol, i)) = t%dele(i) * engunit - delrc_mol(loc2(mol, i))
      delE(i) = delE(i) + t%dele(i)
    enddo
    delE(mol) = delE(mol) + estep
  end subroutine insertion

  ! Energy of guest molecule `choice` in place, with (lexisting) or without
  ! it in S(k) (mc_moves.f:277-374).  displace_guest (gcmc.f:2816-2901) calls
  ! it for the molecule in place, then at its displaced position: the first
  ! call only records the molecule (from ind(), filled by get_guest); the
  ! second evaluates the whole displacement on the device and returns
  ! E_new - E_old, so that displace_guest's estep - enginit is the same dU.
  subroutine guest_energy(imcon, iguest, alpha, rcut, ddelr, drewd, totatm, maxmls, volm, kmax1, kmax2, &
                          kmax3, epsq, dlrpot, engunit, vdwsum, ewld2sum, ewld1eng, lsurf, newld, &
                          surftol, sumchg, chgtmp, engsictmp, loverlap, overlap, estep, lexisting, &
                          lblock, numblocks)
    integer :: imcon, iguest, totatm, maxmls, kmax1, kmax2, kmax3, newld, numblocks
    real(8) :: alpha, rcut, ddelr, drewd, volm, epsq, dlrpot, engunit, vdwsum, ewld2sum, ewld1eng
    real(8) :: surftol, sumchg, chgtmp, engsictmp, overlap, estep
    logical :: lsurf, loverlap, lexisting, lblock
    type(fmc_trial_c) :: t
    real(c_double) :: xyz(3 * mxguestsite)
    integer :: mol, natms, i, g, atmadd
    integer(c_int) :: rc
    if (lblock .and. .not. lexisting) call error(0, 9001)
    if (.not. c_associated(fmc_h)) call fmc_shim_init(imcon, maxmls, alpha, kmax1, kmax2, kmax3, rcut, &
                                                      ddelr, epsq, engunit, overlap, surftol, sumchg, dlrpot)
    mol = locguest(iguest)
    natms = numatoms(mol)
    lsurf = .false.
    loverlap = .false.
    call fmc_zero_terms(maxmls)
    estep = 0.d0
    vdwsum = 0.d0
    ewld2sum = 0.d0
    ewld1eng = 0.d0
    call fmc_selfterms(mol, natms, alpha, epsq, lexisting, chgtmp, engsictmp)
    if (lexisting) then
      call fmc_sync(sumchg)
      atmadd = 0
      do g = 1, iguest - 1
        atmadd = atmadd + numatoms(locguest(g)) * nummols(locguest(g))
      enddo
      fmc_disp_iguest = iguest
      fmc_disp_choice = (ind(1) - atmadd - 1) / natms + 1
      ! surface flag of the position in place: a zero displacement, rejected
      call fmc_pack(natms, xyz)
      rc = fmc_trial_displace(fmc_h, 0_c_int, int(iguest, c_int), int(fmc_disp_choice, c_int), xyz, t)
      if (rc /= 0) call fmc_fail('fmc_trial_displace', rc)
      lsurf = btest(t%surface, 1)
      rc = fmc_reject(fmc_h, 0_c_int)
      if (rc /= 0) call fmc_fail('fmc_reject', rc)
      return
    endif
    if (fmc_disp_iguest /= iguest) call fmc_fail('guest_energy without a molecule in place', 9002_c_int)
    call fmc_pack(natms, xyz)
    rc = fmc_trial_displace(fmc_h, 0_c_int, int(iguest, c_int), int(fmc_disp_choice, c_int), xyz, t)
    if (rc /= 0) call fmc_fail('fmc_trial_displace', rc)
    fmc_pending = 1
    fmc_pend_overlap = .false.
    fmc_disp_iguest = 0
    call fmc_log_trial(3, iguest, fmc_disp_choice, natms, xyz, t)
    loverlap = t%overlap /= 0
    lsurf = btest(t%surface, 0)
    estep = t%estep
    ewld1eng = t%recip * engunit
    ewld2sum = t%creal * engunit
    vdwsum = t%vdw * engunit
    do i = 1, maxmls
      if (i == mol) cycle
      ewald1en(loc2(mol, i)) = t%dele(i) * engunit
    enddo
  end subroutine guest_energy

  ! Energy released by removing guest molecule `choice` (mc_moves.f:375-484).
  subroutine deletion(imcon, keyfce, iguest, choice, alpha, rcut, ddelr, drewd, maxmls, totatm, volm, &
                      kmax1, kmax2, kmax3, epsq, dlrpot, ntpatm, maxvdw, engunit, delrc, estep, &
                      linitsurf, surftol, sumchg, engsictmp, chgtmp, overlap, newld)
    integer :: imcon, keyfce, iguest, choice, maxmls, totatm, kmax1, kmax2, kmax3, ntpatm, maxvdw, newld
    real(8) :: alpha, rcut, ddelr, drewd, volm, epsq, dlrpot, engunit, delrc, estep, surftol, sumchg
    real(8) :: engsictmp, chgtmp, overlap
    logical :: linitsurf
    type(fmc_trial_c) :: t
    real(c_double) :: xyz(3 * mxguestsite)
    integer :: mol, natms, i, a
    integer(c_int) :: rc
    if (.not. c_associated(fmc_h)) call fmc_shim_init(imcon, maxmls, alpha, kmax1, kmax2, kmax3, rcut, &
                                                      ddelr, epsq, engunit, overlap, surftol, sumchg, dlrpot)
    mol = locguest(iguest)
    natms = numatoms(mol)
    linitsurf = .false.
    call fmc_sync(sumchg)
    call fmc_zero_terms(maxmls)
    call gstlrcorrect(imcon, iguest, keyfce, ntpatm, maxvdw, delrc, volm, maxmls, .true.)
    do i = 1, natms
      a = (choice - 1) * natms + i
      xyz(3 * i - 2) = molxxx(mol, a)
      xyz(3 * i - 1) = molyyy(mol, a)
      xyz(3 * i) = molzzz(mol, a)
    enddo
    rc = fmc_trial_delete(fmc_h, 0_c_int, int(iguest, c_int), int(choice, c_int), t)
    if (rc /= 0) call fmc_fail('fmc_trial_delete', rc)
    fmc_pending = 1
    fmc_pend_overlap = .false.
    call fmc_log_trial(2, iguest, choice, natms, xyz, t)
    linitsurf = btest(t%surface, 1)
    call fmc_selfterms(mol, natms, alpha, epsq, .true., chgtmp, engsictmp)
    estep = estep + t%estep
    do i = 1, maxmls
      if (i == mol) cycle
      ewald1en(loc2(mol, i)) = -t%dele(i) * engunit - delrc_mol(loc2(mol, i))
      delE(i) = delE(i) + t%dele(i)
    enddo
    delE(mol) = delE(mol) - estep
  end subroutine deletion

  ! ------------------------------------------------------------ decisions --
  ! Commits the last trial: on the device, and in the host arrays gcmc.f
  ! reads (semantics of mc_moves.f:485-621; S(k) stays on the device).
  subroutine accept_move(iguest, insert, delete, displace, lsurf, delrc, totatm, choice, ntpfram, &
                         ntpguest, maxmls, sumchg, engsictmp, chgtmp, newld)
    integer :: iguest, totatm, choice, ntpfram, ntpguest, maxmls, newld
    logical :: insert, delete, displace, lsurf
    real(8) :: delrc, sumchg, engsictmp, chgtmp
    integer :: mol, natms, kk, first, last, i, s
    integer(c_int) :: rc
    logical :: on_device
    on_device = .false.
    if (fmc_pending /= 0) then
      if (fmc_pend_overlap) then
        ! an overlapping insertion accepted on the way to rejecting a swap or
        ! switch: the device keeps its state, the host list is reloaded if
        ! the move's restores leave it different
        rc = fmc_reject(fmc_h, 0_c_int)
        if (rc /= 0) call fmc_fail('fmc_reject', rc)
      else
        rc = fmc_accept(fmc_h, 0_c_int)
        if (rc /= 0) call fmc_fail('fmc_accept', rc)
        on_device = .true.
      endif
      fmc_pending = 0
      fmc_pend_overlap = .false.
      if (fmc_tracing) write(fmc_trace_unit, '(a)') 'D 1'
    endif
    mol = locguest(iguest)
    natms = numatoms(mol)
    kk = loc2(mol, mol)
    energy(1:maxmls) = energy(1:maxmls) + delE(1:maxmls)
    if (insert) then
      engsic(kk) = engsic(kk) + engsictmp
      chgsum_mol(mol) = chgsum_mol(mol) + chgtmp
      sumchg = sumchg + chgtmp
      if (lsurf) surfacemols(mol) = surfacemols(mol) + 1
      first = natms * nummols(mol)
      do s = 1, natms
        molxxx(mol, first + s) = newx(s)
        molyyy(mol, first + s) = newy(s)
        molzzz(mol, first + s) = newz(s)
        call fmc_count_site(mol, s, 1)
      enddo
      elrc = elrc + delrc
      elrc_mol(:) = elrc_mol(:) + delrc_mol(:)
      nummols(mol) = nummols(mol) + 1
      totatm = totatm + natms
      call condense(totatm, ntpfram, ntpguest)
      choice = nummols(mol)
    elseif (delete) then
      engsic(kk) = engsic(kk) - engsictmp
      chgsum_mol(mol) = chgsum_mol(mol) - chgtmp
      sumchg = sumchg - chgtmp
      if (lsurf) surfacemols(mol) = surfacemols(mol) - 1
      do s = 1, natms
        call fmc_count_site(mol, s, -1)
      enddo
      ! molecules after `choice` move down one slot, in order
      first = (choice - 1) * natms + 1
      last = natms * (nummols(mol) - 1)
      do i = first, last
        molxxx(mol, i) = molxxx(mol, i + natms)
        molyyy(mol, i) = molyyy(mol, i + natms)
        molzzz(mol, i) = molzzz(mol, i + natms)
      enddo
      elrc = elrc - delrc
      elrc_mol(:) = elrc_mol(:) - delrc_mol(:)
      nummols(mol) = nummols(mol) - 1
      totatm = totatm - natms
      call condense(totatm, ntpfram, ntpguest)
    elseif (displace) then
      first = (choice - 1) * natms
      do s = 1, natms
        molxxx(mol, first + s) = newx(s)
        molyyy(mol, first + s) = newy(s)
        molzzz(mol, first + s) = newz(s)
      enddo
      call condense(totatm, ntpfram, ntpguest)
    endif
    if (on_device) call fmc_shadow_take(mol)
  end subroutine accept_move

  subroutine fmc_count_site(mol, s, dn)
    integer, intent(in) :: mol, s, dn
    integer :: ka
    ka = ltpsit(mol, s)
    numtyp(ka) = numtyp(ka) + dn
    numtyp_mol(mol, ka) = numtyp_mol(mol, ka) + dn
    if (lfzsite(mol, s) /= 0) then
      numfrz(ka) = numfrz(ka) + dn
      numfrz_mol(mol, ka) = numfrz_mol(mol, ka) + dn
    endif
  end subroutine fmc_count_site

  ! Discards the last trial (semantics of mc_moves.f:623-673).
  subroutine reject_move(iguest, jguest, insert, delete, displace, swap)
    integer :: iguest, jguest
    logical :: insert, delete, displace, swap
    integer :: mol, jmol, s
    integer(c_int) :: rc
    if (fmc_pending /= 0) then
      rc = fmc_reject(fmc_h, 0_c_int)
      if (rc /= 0) call fmc_fail('fmc_reject', rc)
      fmc_pending = 0
      fmc_pend_overlap = .false.
      if (fmc_tracing) write(fmc_trace_unit, '(a)') 'D 0'
    endif
    if (swap .and. fmc_tracing) write(fmc_trace_unit, '(a)') 'R'
    delE(:) = 0.d0
    if (insert .or. delete) then
      engsic = engsicorig
    elseif (swap) then
      mol = locguest(iguest)
      jmol = locguest(jguest)
      engsic = engsicorig
      elrc_mol = origelrc_mol
      do s = 1, numatoms(mol)
        call fmc_count_site(mol, s, 1)
      enddo
      do s = 1, numatoms(jmol)
        call fmc_count_site(jmol, s, -1)
      enddo
    endif
    if (displace) fmc_disp_iguest = 0
  end subroutine reject_move

end module mc_moves

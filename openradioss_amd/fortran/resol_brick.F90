!  resol_brick.F90 -- a minimal RESOL-style explicit driver over the C ABI.
!
!  usage: resol_brick <deck.bin> <ncycle> <group|device> <out.bin> [mixed]
!
!  group : the reference's call surface.  Each cycle follows RESOL
!          (engine/source/engine/resol.F:2586-9595): DT1=DT2, DT2=EP06; the
!          FORINT group loop (elements/forint.F:325-347) calls SFORC3's drop-in
!          orad_sforc3_async per NVSIZ=128 group and orad_group_wait after the
!          loop (the DT2T merge); ASSPAR4's drop-in orad_asspar4 returns A and
!          STIFN; then DT2 selection, ACCELE, BCS10, VELOCITY and DEPLA run here on
!          the host in FP64, and orad_update_xv hands X, V back to the device.
!  device: the whole cycle on the GPU, orad_run_cycles.
!
!  deck.bin (written by openradioss_amd.engine.write_deck): int32 magic, version,
!  NUMNOD, NUMELS; X(3,NUMNOD), V(3,NUMNOD), MS(NUMNOD) real(8); ICODT(NUMNOD),
!  IXS(11,NUMELS) int32; VOL0(NUMELS), law-2 parameters(11) real(8); ICC, VP,
!  ISRATE, pad int32; qa, qb, h, DTFAC real(8).
program resol_brick
  use iso_c_binding
  use orad_bind
  implicit none
  integer, parameter :: NVSIZ = 128
  character(len=512) :: deckf, outf, arg
  integer(c_int) :: hdr(4), numnod, numels, ncycle, rc, prec, ipar(4)
  integer :: ncyc, nft, nel, n, u
  real(c_double), allocatable :: x(:,:), v(:,:), d(:,:), a(:,:), ms(:), stifn(:), vol0(:)
  integer(c_int), allocatable :: icodt(:), ixs(:,:)
  real(c_double) :: mpar(11), prop(4), tt, dt1, dt2, dt12, dt2old, dtmx, dt2t, rtmp, vdt
  type(c_ptr) :: ctx
  type(orad_time_t) :: t
  logical :: group_mode

  call get_command_argument(1, deckf)
  call get_command_argument(2, arg); read(arg, *) ncycle
  call get_command_argument(3, arg); group_mode = (trim(arg) == 'group')
  call get_command_argument(4, outf)
  prec = ORAD_FP64
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg)
    if (trim(arg) == 'mixed') prec = ORAD_MIXED
  end if

  open(newunit=u, file=trim(deckf), access='stream', form='unformatted', status='old')
  read(u) hdr
  if (hdr(1) /= int(z'4F524144', c_int) .or. hdr(2) /= 1) stop 'bad deck file'
  numnod = hdr(3); numels = hdr(4)
  allocate(x(3,numnod), v(3,numnod), d(3,numnod), a(3,numnod), ms(numnod), stifn(numnod))
  allocate(icodt(numnod), ixs(11,numels), vol0(numels))
  read(u) x, v, ms, icodt, ixs, vol0, mpar, ipar, prop
  close(u)

  call orad_check(ctx, orad_create(ctx, 0_c_int, prec), 'create')
  call orad_check(ctx, orad_set_nodes(ctx, numnod, x, v, ms, icodt), 'set_nodes')
  call orad_check(ctx, orad_set_law2(ctx, mpar, ipar), 'set_law2')
  call orad_check(ctx, orad_set_solid_prop(ctx, prop(1), prop(2), prop(3), prop(4)), 'set_solid_prop')
  call orad_check(ctx, orad_set_bricks(ctx, numels, ixs, vol0), 'set_bricks')
  d = 0.0d0

  if (.not. group_mode) then
    call orad_check(ctx, orad_run_cycles(ctx, ncycle, t), 'run_cycles')
    call orad_check(ctx, orad_get_nodes(ctx, x, v, d, a, stifn), 'get_nodes')
  else
    ! starter values (contrl.F:1108-1113) and engine DTMX default
    tt = 0.0d0; dt1 = 0.0d0; dt2 = 0.0d0; dt2old = 1.0d6/1.1d0; dtmx = 1.0d20
    do ncyc = 1, ncycle
      dt1 = dt2
      dt2 = 1.0d6
      dt2t = 1.0d30
      do nft = 0, numels - 1, NVSIZ                  ! FORINT group loop, SFORC3 per group
        nel = min(NVSIZ, numels - nft)
        call orad_check(ctx, orad_sforc3_async(ctx, nft, nel, dt1), 'sforc3')
      end do
      call orad_check(ctx, orad_group_wait(ctx, dt2t), 'group_wait')
      call orad_check(ctx, orad_asspar4(ctx, a, stifn), 'asspar4')
      if (dt2t < dt2) dt2 = dt2t
      dt2 = min(dt2, 1.1*dt2old, dtmx)
      dt2old = dt2
      dt12 = 0.5d0*(dt1 + dt2)
      do n = 1, numnod
        if (ms(n) > 0.0d0) then                       ! ACCELE
          rtmp = 1.0d0/ms(n)
          a(1,n) = a(1,n)*rtmp; a(2,n) = a(2,n)*rtmp; a(3,n) = a(3,n)*rtmp
        else
          a(:,n) = 0.0d0
        end if
        if (iand(icodt(n), 4) /= 0) then; v(1,n) = 0.0d0; a(1,n) = 0.0d0; end if   ! BCS10
        if (iand(icodt(n), 2) /= 0) then; v(2,n) = 0.0d0; a(2,n) = 0.0d0; end if
        if (iand(icodt(n), 1) /= 0) then; v(3,n) = 0.0d0; a(3,n) = 0.0d0; end if
        v(1,n) = v(1,n) + dt12*a(1,n)                ! VELOCITY
        v(2,n) = v(2,n) + dt12*a(2,n)
        v(3,n) = v(3,n) + dt12*a(3,n)
        vdt = dt2*v(1,n); d(1,n) = d(1,n) + vdt; x(1,n) = x(1,n) + vdt     ! DEPLA
        vdt = dt2*v(2,n); d(2,n) = d(2,n) + vdt; x(2,n) = x(2,n) + vdt
        vdt = dt2*v(3,n); d(3,n) = d(3,n) + vdt; x(3,n) = x(3,n) + vdt
      end do
      tt = tt + dt2
      call orad_check(ctx, orad_update_xv(ctx, x, v), 'update_xv')   ! no VR for bricks
    end do
    write(*,'(a,i0,a,es22.15,a,es22.15)') 'cycles ', ncycle, ' tt ', tt, ' dt2 ', dt2
  end if

  open(newunit=u, file=trim(outf), access='stream', form='unformatted', status='replace')
  write(u) x, v, d
  close(u)
  rc = orad_destroy(ctx)
end program resol_brick

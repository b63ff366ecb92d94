!  orad_bind.F90 -- ISO_C_BINDING interface to the C ABI in include/orad.h.
!
!  Style follows the reference's coupling adapters (engine/source/coupling/
!  coupling_adapter.F90:83-131): type(c_ptr), value handles, scalars by value,
!  arrays by reference in the engine's own layouts (X(3,NUMNOD), IXS(NIXS,NUMELS),
!  IXC(NIXC,NUMELC)).
module orad_bind
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ORAD_OK = 0, ORAD_EINVAL = 1, ORAD_EHIP = 2, &
                               ORAD_ESTATE = 3, ORAD_ENEGVOL = 4
  integer(c_int), parameter :: ORAD_FP64 = 0, ORAD_MIXED = 1

  type, bind(c) :: orad_time_t
    real(c_double) :: tt, dt1, dt2, dt12, dt2old, dtmx
    integer(c_int) :: ncycle, mstop
  end type orad_time_t

  type, bind(c) :: orad_balance_t
    real(c_double) :: encin, enrot, enint, ehour, wfext, entot, entot0, err, tt
    integer(c_int) :: ncycle, pad
  end type orad_balance_t

  type, bind(c) :: orad_launch_t
    integer(c_int) :: persistent, lds_kb, xcd_remap, grid
    integer(c_int) :: block, waves, order_rows, node_tile
    integer(c_int) :: adj_vjp
  end type orad_launch_t

  interface
    integer(c_int) function orad_create(ctx, device, precision) bind(c, name='orad_create')
      import :: c_ptr, c_int
      type(c_ptr) :: ctx
      integer(c_int), value :: device, precision
    end function
    integer(c_int) function orad_destroy(ctx) bind(c, name='orad_destroy')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function orad_last_error(ctx) bind(c, name='orad_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function orad_set_nodes(ctx, numnod, x, v, ms, icodt) bind(c, name='orad_set_nodes')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: numnod
      real(c_double) :: x(3,*), v(3,*), ms(*)
      integer(c_int) :: icodt(*)
    end function
    integer(c_int) function orad_set_bricks(ctx, numels, ixs, vol0) bind(c, name='orad_set_bricks')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: numels
      integer(c_int) :: ixs(11,*)
      real(c_double) :: vol0(*)
    end function
    integer(c_int) function orad_set_law2(ctx, p, ip) bind(c, name='orad_set_law2')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: p(*)
      integer(c_int) :: ip(*)
    end function
    integer(c_int) function orad_set_solid_prop(ctx, qa, qb, hvis, dtfac) bind(c, name='orad_set_solid_prop')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: qa, qb, hvis, dtfac
    end function
    integer(c_int) function orad_get_pon(ctx, adsky, iads) bind(c, name='orad_get_pon')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: adsky(*), iads(8,*)
    end function
    integer(c_int) function orad_set_time(ctx, t) bind(c, name='orad_set_time')
      import :: c_ptr, c_int, orad_time_t
      type(c_ptr), value :: ctx
      type(orad_time_t) :: t
    end function
    integer(c_int) function orad_get_time(ctx, t) bind(c, name='orad_get_time')
      import :: c_ptr, c_int, orad_time_t
      type(c_ptr), value :: ctx
      type(orad_time_t) :: t
    end function
    integer(c_int) function orad_sforc3(ctx, nft, nel, dt1, dt2t) bind(c, name='orad_sforc3')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nft, nel
      real(c_double), value :: dt1
      real(c_double) :: dt2t
    end function
    integer(c_int) function orad_sforc3_async(ctx, nft, nel, dt1) bind(c, name='orad_sforc3_async')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nft, nel
      real(c_double), value :: dt1
    end function
    integer(c_int) function orad_cforc3_async(ctx, nft, nel, dt1) bind(c, name='orad_cforc3_async')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nft, nel
      real(c_double), value :: dt1
    end function
    integer(c_int) function orad_group_wait(ctx, dt2t) bind(c, name='orad_group_wait')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: dt2t
    end function
    integer(c_int) function orad_update_xv(ctx, x, v, vr) bind(c, name='orad_update_xv')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: x(3,*), v(3,*)
      real(c_double), optional :: vr(3,*)
    end function
    integer(c_int) function orad_asspar4(ctx, a, stifn) bind(c, name='orad_asspar4')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: a(3,*), stifn(*)
    end function
    integer(c_int) function orad_run_cycles(ctx, ncycle, t) bind(c, name='orad_run_cycles')
      import :: c_ptr, c_int, orad_time_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: ncycle
      type(orad_time_t) :: t
    end function
    integer(c_int) function orad_get_nodes(ctx, x, v, d, a, stifn) bind(c, name='orad_get_nodes')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: x(3,*), v(3,*), d(3,*), a(3,*), stifn(*)
    end function
    integer(c_int) function orad_energies(ctx, eint, ekin, ehour) bind(c, name='orad_energies')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: eint, ekin, ehour
    end function
    integer(c_int) function orad_energy_balance(ctx, b) bind(c, name='orad_energy_balance')
      import :: c_ptr, c_int, orad_balance_t
      type(c_ptr), value :: ctx
      type(orad_balance_t) :: b
    end function
    integer(c_int) function orad_energies_noalloc(ctx, eint, ekin, ehour) &
        bind(c, name='orad_energies_noalloc')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: eint, ekin, ehour
    end function
    integer(c_int) function orad_set_fixvel(ctx, n, nodes, dir, vel) bind(c, name='orad_set_fixvel')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      integer(c_int) :: nodes(*), dir(*)
      real(c_double) :: vel(*)
    end function
    ! ---- shells (CFORC3 path) ----
    integer(c_int) function orad_set_nodes_rot(ctx, vr, in, icodr) bind(c, name='orad_set_nodes_rot')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: vr(3,*), in(*)
      integer(c_int) :: icodr(*)
    end function
    integer(c_int) function orad_set_shells(ctx, numelc, ixc, thke) bind(c, name='orad_set_shells')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: numelc
      integer(c_int) :: ixc(7,*)
      real(c_double) :: thke(*)
    end function
    integer(c_int) function orad_set_shell_law2(ctx, p, ip) bind(c, name='orad_set_shell_law2')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: p(*)
      integer(c_int) :: ip(*)
    end function
    integer(c_int) function orad_set_shell_prop(ctx, g, ish, npt) bind(c, name='orad_set_shell_prop')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: g(*)
      integer(c_int), value :: ish, npt
    end function
    integer(c_int) function orad_set_shell_ishell(ctx, ishell, dn) bind(c, name='orad_set_shell_ishell')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ishell
      real(c_double), value :: dn
    end function
    integer(c_int) function orad_cforc3(ctx, nft, nel, dt1, dt2t) bind(c, name='orad_cforc3')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nft, nel
      real(c_double), value :: dt1
      real(c_double) :: dt2t
    end function
    integer(c_int) function orad_asspar4_rot(ctx, a, ar, stifn, stifr) bind(c, name='orad_asspar4_rot')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: a(3,*), ar(3,*), stifn(*), stifr(*)
    end function
    integer(c_int) function orad_get_nodes_rot(ctx, vr, ar) bind(c, name='orad_get_nodes_rot')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: vr(3,*), ar(3,*)
    end function
    ! rows: 53 values per shell, 60 with QEPH (orad_set_shell_ishell)
    integer(c_int) function orad_get_shell_state(ctx, rows) bind(c, name='orad_get_shell_state')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: rows(*)
    end function
    ! rows: 53 values per shell, 60 with QEPH (orad_set_shell_ishell)
    integer(c_int) function orad_set_shell_state(ctx, rows) bind(c, name='orad_set_shell_state')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: rows(*)
    end function
    ! p and ip: c_loc of the per-model law-2 records, or c_null_ptr
    integer(c_int) function orad_set_shell_models(ctx, nmodel, node_off, elem_off, p, ip) &
        bind(c, name='orad_set_shell_models')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nmodel
      integer(c_int) :: node_off(*), elem_off(*)
      type(c_ptr), value :: p, ip
    end function
    integer(c_int) function orad_get_model_time(ctx, model, t) bind(c, name='orad_get_model_time')
      import :: c_ptr, c_int, orad_time_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: model
      type(orad_time_t) :: t
    end function
    integer(c_int) function orad_model_energies(ctx, model, eint, ekin, ehour) &
        bind(c, name='orad_model_energies')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: model
      real(c_double) :: eint, ekin, ehour
    end function
    integer(c_int) function orad_set_launch(ctx, l) bind(c, name='orad_set_launch')
      import :: c_ptr, c_int, orad_launch_t
      type(c_ptr), value :: ctx
      type(orad_launch_t) :: l
    end function
    integer(c_int) function orad_get_launch(ctx, l) bind(c, name='orad_get_launch')
      import :: c_ptr, c_int, orad_launch_t
      type(c_ptr), value :: ctx
      type(orad_launch_t) :: l
    end function
    integer(c_int) function orad_autotune(ctx, cycles, best, best_ms) bind(c, name='orad_autotune')
      import :: c_ptr, c_int, c_double, orad_launch_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: cycles
      type(orad_launch_t) :: best
      real(c_double) :: best_ms
    end function
    integer(c_int) function orad_set_contact7(ctx, nsn, nsv, nrtm, irect, stfm, stfns, params, &
        cand_cap) bind(c, name='orad_set_contact7')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsn, nrtm, cand_cap
      integer(c_int) :: nsv(*), irect(4, *)
      real(c_double) :: stfm(*), stfns(*), params(4)
    end function
    integer(c_int) function orad_contact7(ctx, dt1, dt2t) bind(c, name='orad_contact7')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt1
      real(c_double) :: dt2t
    end function
    integer(c_int) function orad_get_contact7(ctx, energies, counts, stfns) &
        bind(c, name='orad_get_contact7')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: energies(3), stfns(*)
      integer(c_int) :: counts(4)
    end function
    integer(c_int) function orad_get_contact7_sky(ctx, cap, fsky, isky, nsky) &
        bind(c, name='orad_get_contact7_sky')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: cap
      real(c_double) :: fsky(4, *)
      integer(c_int) :: isky(*), nsky
    end function
    integer(c_int) function orad_write_restart(ctx, path) bind(c, name='orad_write_restart')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*) :: path
    end function
    integer(c_int) function orad_read_restart(ctx, path) bind(c, name='orad_read_restart')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*) :: path
    end function
    integer(c_int) function orad_adjoint_forward(ctx, ncycle, every) &
        bind(c, name='orad_adjoint_forward')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: ncycle, every
    end function
    ! every argument: c_loc of an array, or c_null_ptr
    integer(c_int) function orad_adjoint_reverse(ctx, lam_x, lam_v, lam_vr, lam_d, lam_state, &
        g_jc, g_thk, g_ms, g_in, lam_x0, lam_v0, lam_vr0) bind(c, name='orad_adjoint_reverse')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      type(c_ptr), value :: lam_x, lam_v, lam_vr, lam_d, lam_state
      type(c_ptr), value :: g_jc, g_thk, g_ms, g_in, lam_x0, lam_v0, lam_vr0
    end function
  end interface

contains

  ! Status check in place of the reference's ANCMSG/ARRET exit.
  subroutine orad_check(ctx, rc, what)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == ORAD_OK) return
    call c_f_pointer(orad_last_error(ctx), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n+1) == c_null_char) exit
      n = n + 1
    end do
    write(*,'(a,a,a,i0,a,512a)') 'orad: ', what, ' failed, status ', rc, ': ', msg(1:n)
    stop 2
  end subroutine orad_check

end module orad_bind

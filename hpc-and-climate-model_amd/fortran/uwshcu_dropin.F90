! uwshcu_dropin.F90 -- Fortran declarations for the MI355X UW shallow-cumulus library.
!
! The physics driver keeps its call of bxcompute_uwshcu_inv exactly as the reference
! makes it (convect_shallow.F90.error:605-619): an implicit interface, every argument
! by reference, external name bxcompute_uwshcu_inv_, which the library exports
! (include/uwshcu.h).  This module adds what the driver needs around that call:
!
!   uwshcu_params_t            the tunables (rpen, rle, rkm), layout of uwshcu_params
!   uwshcu_set_dropin_params   pass the uwshcu_nl namelist's rpen once after
!                              uwshcu_readnl (uwshcu.F90:58-98); until then the
!                              drop-in computes nothing (rpen has no default:
!                              init_uwshcu stops the model, uwshcu.F90:241-242)
!   uwshcu_dropin_last_status  0, or the negative code of the last chunk's fatal
!                              condition (UWSHCU_ERR_* in uwshcu_types.h) and its
!                              column, for a driver that wants to call endrun
!
! Module uwshcu_dropin_explicit declares the same entry point with an explicit
! interface (bxcompute_uwshcu_inv_explicit), for drivers that prefer one.  It lives
! in a module of its own because a scope that sees a BIND(C) interface for
! bxcompute_uwshcu_inv_ may not also make the implicit call.
!
! Usage:
!   use uwshcu_dropin
!   rc = uwshcu_set_dropin_params(uwshcu_params_t(uwshcu_rpen, 0.1_c_double, 14.0_c_double))
!   call bxcompute_uwshcu_inv(pcols, pver, ncol, pcnst, ztodt, ...)   ! unchanged
!   if (uwshcu_dropin_last_status(bad_col) /= 0) call endrun('UW shallow convection')
module uwshcu_dropin
   use iso_c_binding, only: c_int, c_double
   implicit none
   private
   public :: uwshcu_params_t, uwshcu_set_dropin_params, uwshcu_dropin_last_status, UWSHCU_ERR_PARAMS

   integer(c_int), parameter :: UWSHCU_ERR_PARAMS = -16

   type, bind(C) :: uwshcu_params_t
      real(c_double) :: rpen, rle, rkm
   end type uwshcu_params_t

   interface
      integer(c_int) function uwshcu_set_dropin_params(par) bind(C, name="uwshcu_set_dropin_params")
         import :: c_int, uwshcu_params_t
         type(uwshcu_params_t), intent(in) :: par
      end function uwshcu_set_dropin_params

      integer(c_int) function uwshcu_dropin_last_status(bad_col) bind(C, name="uwshcu_dropin_last_status")
         import :: c_int
         integer(c_int), intent(out) :: bad_col
      end function uwshcu_dropin_last_status
   end interface
end module uwshcu_dropin

module uwshcu_dropin_explicit
   use iso_c_binding, only: c_int, c_double
   implicit none
   interface
      subroutine bxcompute_uwshcu_inv_explicit(pcols, pver, ncol, pcnst, ztodt,                  &
           ps0, zs0, p0, z0, dp0, u0, v0, qv0, ql0, qi0, t0, s0, tr0, tke, cldfrc, concldfrc,     &
           pblh, cush, umf, slflx, qtflx, flxprc1, flxsnow1, qvten, qlten, qiten, sten, uten,     &
           vten, trten, qrten, qsten, precip, snow, evapc, cufrc, qcu, qlu, qiu, cbmf, qc,        &
           rliq, cnt, cnb, lchnk, dpdry0, wu) bind(C, name="bxcompute_uwshcu_inv_")
         import :: c_int, c_double
         integer(c_int), intent(in) :: pcols, pver, ncol, pcnst, lchnk
         real(c_double), intent(in) :: ztodt
         real(c_double), dimension(pcols, pver+1), intent(in) :: ps0, zs0, tke
         real(c_double), dimension(pcols, pver), intent(in) :: p0, z0, dp0, u0, v0, qv0, ql0, qi0, t0, s0, &
              cldfrc, concldfrc, dpdry0
         real(c_double), dimension(pcols, pver, pcnst), intent(in) :: tr0
         real(c_double), dimension(pcols), intent(in) :: pblh
         real(c_double), dimension(pcols), intent(inout) :: cush
         ! outputs: only columns 1..ncol are written, the rest keep the caller's values
         real(c_double), dimension(pcols, pver+1), intent(inout) :: umf, slflx, qtflx, flxprc1, flxsnow1, wu
         real(c_double), dimension(pcols, pver), intent(inout) :: qvten, qlten, qiten, sten, uten, vten, &
              qrten, qsten, evapc, cufrc, qcu, qlu, qiu, qc
         real(c_double), dimension(pcols, pver, pcnst), intent(inout) :: trten
         real(c_double), dimension(pcols), intent(inout) :: precip, snow, cbmf, rliq, cnt, cnb
      end subroutine bxcompute_uwshcu_inv_explicit
   end interface
end module uwshcu_dropin_explicit

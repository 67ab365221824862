! explicit_call.F90 -- TEST HARNESS: a driver that prefers the explicit interface
! (module uwshcu_dropin_explicit); compiled for syntax only by the CPU tests.
subroutine shallow_with_explicit_interface(pcols, pver, ncol, pcnst, ztodt, ps0, zs0, p0, z0, dp0, u0, v0, qv0, &
     ql0, qi0, t0, s0, tr0, tke, cld, concld, pblh, cush, umf, slflx, qtflx, flxprc1, flxsnow1, qvten, qlten,  &
     qiten, sten, uten, vten, trten, qrten, qsten, precip, snow, evapc, cufrc, qcu, qlu, qiu, cbmf, qc, rliq,  &
     cnt, cnb, lchnk, dpdry0, wu, uwshcu_rpen)
   use iso_c_binding, only: c_int, c_double
   use uwshcu_dropin
   use uwshcu_dropin_explicit
   implicit none
   integer(c_int), intent(in) :: pcols, pver, ncol, pcnst, lchnk
   real(c_double), intent(in) :: ztodt, uwshcu_rpen
   real(c_double), dimension(pcols, pver+1) :: ps0, zs0, tke, umf, slflx, qtflx, flxprc1, flxsnow1, wu
   real(c_double), dimension(pcols, pver) :: p0, z0, dp0, u0, v0, qv0, ql0, qi0, t0, s0, cld, concld, dpdry0, &
        qvten, qlten, qiten, sten, uten, vten, qrten, qsten, evapc, cufrc, qcu, qlu, qiu, qc
   real(c_double), dimension(pcols, pver, pcnst) :: tr0, trten
   real(c_double), dimension(pcols) :: pblh, cush, precip, snow, cbmf, rliq, cnt, cnb
   integer(c_int) :: rc, bad_col
   rc = uwshcu_set_dropin_params(uwshcu_params_t(uwshcu_rpen, 0.1_c_double, 14.0_c_double))
   call bxcompute_uwshcu_inv_explicit(pcols, pver, ncol, pcnst, ztodt, ps0, zs0, p0, z0, dp0, u0, v0, qv0, ql0, &
        qi0, t0, s0, tr0, tke, cld, concld, pblh, cush, umf, slflx, qtflx, flxprc1, flxsnow1, qvten, qlten,     &
        qiten, sten, uten, vten, trten, qrten, qsten, precip, snow, evapc, cufrc, qcu, qlu, qiu, cbmf, qc, rliq, &
        cnt, cnb, lchnk, dpdry0, wu)
   if (uwshcu_dropin_last_status(bad_col) /= 0) stop 1
end subroutine shallow_with_explicit_interface

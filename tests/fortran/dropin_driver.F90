! dropin_driver.F90 -- TEST HARNESS: a Fortran physics driver that calls the library
! the way convect_shallow_tend calls the reference's bxcompute_uwshcu_inv
! (convect_shallow.F90.error:605-619): implicit interface, the physics state's arrays
! and constituent slices passed directly, rpen passed once beforehand as
! uwshcu_readnl would have read it (uwshcu.F90:58-98).
!
!   dropin_driver <in.bin> <out.bin>
!
! in.bin (stream, native endian): int32 pcols, pver, ncol, pcnst, lchnk; float64 ztodt,
! rpen; then the state fields pint zi pmid zm pdel u v t s q(pcols,pver,pcnst) pdeldry
! tke cld concld pblht cush, each in Fortran order.  out.bin: int32 status of the call
! made before rpen is set, int32 status and bad column of the real call, then the
! outputs in argument order.
program dropin_driver
   use iso_c_binding, only: c_int, c_double
   use uwshcu_dropin
   implicit none
   type physics_state
      real(c_double), allocatable :: pint(:,:), zi(:,:), pmid(:,:), zm(:,:), pdel(:,:), u(:,:), v(:,:), &
           t(:,:), s(:,:), q(:,:,:), pdeldry(:,:)
   end type
   type(physics_state) :: state
   integer(c_int) :: pcols, pver, pverp, ncol, pcnst, lchnk, ixcldliq, ixcldice, rc0, rc, badc
   real(c_double) :: ztodt, rpen
   real(c_double), allocatable :: tke(:,:), cld(:,:), concld(:,:), pblht(:), cush1(:)
   real(c_double), allocatable :: cmfmc2(:,:), slflx(:,:), qtflx(:,:), flxprec(:,:), flxsnow(:,:), wu_out(:,:)
   real(c_double), allocatable :: qv_t(:,:), ql_t(:,:), qi_t(:,:), s_t(:,:), u_t(:,:), v_t(:,:), ptend_tracer(:,:,:)
   real(c_double), allocatable :: rprdsh(:,:), cmfdqs(:,:), evapcsh(:,:), shfrc(:,:), iccmr_UW(:,:), &
        icwmr_UW(:,:), icimr_UW(:,:), qc2(:,:)
   real(c_double), allocatable :: precc(:), snow(:), cbmf(:), rliq2(:), cnt2(:), cnb2(:)
   character(len=512) :: fin, fout
   integer :: u_in, u_out

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
   read(u_in) pcols, pver, ncol, pcnst, lchnk
   read(u_in) ztodt, rpen
   pverp = pver + 1
   ixcldliq = 2
   ixcldice = 3
   allocate(state%pint(pcols, pverp), state%zi(pcols, pverp), state%pmid(pcols, pver), state%zm(pcols, pver), &
        state%pdel(pcols, pver), state%u(pcols, pver), state%v(pcols, pver), state%t(pcols, pver), &
        state%s(pcols, pver), state%q(pcols, pver, pcnst), state%pdeldry(pcols, pver))
   allocate(tke(pcols, pverp), cld(pcols, pver), concld(pcols, pver), pblht(pcols), cush1(pcols))
   read(u_in) state%pint, state%zi, state%pmid, state%zm, state%pdel, state%u, state%v, state%t, state%s, &
        state%q, state%pdeldry, tke, cld, concld, pblht, cush1
   close(u_in)

   allocate(cmfmc2(pcols, pverp), slflx(pcols, pverp), qtflx(pcols, pverp), flxprec(pcols, pverp), &
        flxsnow(pcols, pverp), wu_out(pcols, pverp))
   allocate(qv_t(pcols, pver), ql_t(pcols, pver), qi_t(pcols, pver), s_t(pcols, pver), u_t(pcols, pver), &
        v_t(pcols, pver), ptend_tracer(pcols, pver, pcnst))
   allocate(rprdsh(pcols, pver), cmfdqs(pcols, pver), evapcsh(pcols, pver), shfrc(pcols, pver), &
        iccmr_UW(pcols, pver), icwmr_UW(pcols, pver), icimr_UW(pcols, pver), qc2(pcols, pver))
   allocate(precc(pcols), snow(pcols), cbmf(pcols), rliq2(pcols), cnt2(pcols), cnb2(pcols))
   sten_zero: block
      cmfmc2 = 0; slflx = 0; qtflx = 0; flxprec = 0; flxsnow = 0; wu_out = 0
      qv_t = 0; ql_t = 0; qi_t = 0; s_t = 0; u_t = 0; v_t = 0; ptend_tracer = 0
      rprdsh = 0; cmfdqs = 0; evapcsh = 0; shfrc = 0; iccmr_UW = 0; icwmr_UW = 0; icimr_UW = 0; qc2 = 0
      precc = 0; snow = 0; cbmf = 0; rliq2 = 0; cnt2 = 0; cnb2 = 0
   end block sten_zero

   ! before uwshcu_readnl has supplied rpen: the chunk must be refused
   rc0 = uwshcu_set_dropin_params_unset()
   call bxcompute_uwshcu_inv( pcols     , pver    , ncol           , pcnst         , ztodt         ,                   &
                              state%pint, state%zi, state%pmid     , state%zm      , state%pdel    ,                   &
                              state%u   , state%v , state%q(:,:,1) , state%q(:,:,ixcldliq), state%q(:,:,ixcldice),     &
                              state%t   , state%s , state%q(:,:,:) ,                                                   &
                              tke       , cld     , concld         , pblht         , cush1          ,                  &
                              cmfmc2    , slflx   , qtflx          ,                                                   &
                              flxprec, flxsnow,                                                                        &
                              qv_t      , ql_t    , qi_t           ,                                                   &
                              s_t       , u_t     , v_t            , ptend_tracer  ,                                   &
                              rprdsh              , cmfdqs         , precc         , snow          ,                   &
                              evapcsh             , shfrc          , iccmr_UW      , icwmr_UW      ,                   &
                              icimr_UW            , cbmf           , qc2           , rliq2         ,                   &
                              cnt2                , cnb2                , lchnk         , state%pdeldry, wu_out )
   rc0 = uwshcu_dropin_last_status(badc)

   ! uwshcu_readnl has read uwshcu_rpen: pass it once, then the driver's call, unchanged
   rc = uwshcu_set_dropin_params(uwshcu_params_t(rpen, 0.1_c_double, 14.0_c_double))
   call bxcompute_uwshcu_inv( pcols     , pver    , ncol           , pcnst         , ztodt         ,                   &
                              state%pint, state%zi, state%pmid     , state%zm      , state%pdel    ,                   &
                              state%u   , state%v , state%q(:,:,1) , state%q(:,:,ixcldliq), state%q(:,:,ixcldice),     &
                              state%t   , state%s , state%q(:,:,:) ,                                                   &
                              tke       , cld     , concld         , pblht         , cush1          ,                  &
                              cmfmc2    , slflx   , qtflx          ,                                                   &
                              flxprec, flxsnow,                                                                        &
                              qv_t      , ql_t    , qi_t           ,                                                   &
                              s_t       , u_t     , v_t            , ptend_tracer  ,                                   &
                              rprdsh              , cmfdqs         , precc         , snow          ,                   &
                              evapcsh             , shfrc          , iccmr_UW      , icwmr_UW      ,                   &
                              icimr_UW            , cbmf           , qc2           , rliq2         ,                   &
                              cnt2                , cnb2                , lchnk         , state%pdeldry, wu_out )
   rc = uwshcu_dropin_last_status(badc)

   open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(u_out) rc0, rc, badc
   write(u_out) cmfmc2, slflx, qtflx, flxprec, flxsnow, qv_t, ql_t, qi_t, s_t, u_t, v_t, ptend_tracer, &
        rprdsh, cmfdqs, precc, snow, evapcsh, shfrc, iccmr_UW, icwmr_UW, icimr_UW, cbmf, qc2, rliq2, &
        cnt2, cnb2, wu_out, cush1
   close(u_out)

contains

   ! uwshcu_set_dropin_params(NULL): back to the unset state
   integer(c_int) function uwshcu_set_dropin_params_unset()
      use iso_c_binding, only: c_ptr, c_null_ptr
      interface
         integer(c_int) function set_raw(par) bind(C, name="uwshcu_set_dropin_params")
            import :: c_int, c_ptr
            type(c_ptr), value :: par
         end function set_raw
      end interface
      uwshcu_set_dropin_params_unset = set_raw(c_null_ptr)
   end function uwshcu_set_dropin_params_unset
end program dropin_driver

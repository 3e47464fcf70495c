! neorx_shim.f90 -- drop-in replacements for the reference's geminal
! integral and Fock-contraction subroutines, forwarding to libneorx_hip.so
! through iso_c_binding.
!
! Every subroutine keeps the name, argument order and meaning of the
! reference routine it replaces (file:line given above each one), so the
! reference's SCF drivers link against this object instead of
! gam2_IC1_OMP.o / gam3_IC1_OMP.o / gam4_ICR_OMP.o / gam_1_OMP.o / gam_ee.o.
! Integers are 64-bit as in the reference build (-integer-size 64,
! source/comp:7); the shim declares them as integer(c_int64_t) explicitly.
! Failures print the library's message on unit 0 and stop the program: the
! reference's callers have no status argument to check.
module neorx_capi
  use iso_c_binding
  implicit none
  interface
    function neorx_init(dev) bind(C, name="neorx_init")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: dev
      integer(c_int) :: neorx_init
    end function
    function neorx_last_error() bind(C, name="neorx_last_error")
      import :: c_ptr
      type(c_ptr) :: neorx_last_error
    end function
    function neorx_system_create(nebf, npebf, npbf, nat, ngtg, pmass, cat, zan, bcoef, gamma, kpestr, kpeend, &
                                 ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc, handle) &
                                 bind(C, name="neorx_system_create")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: nebf, npebf, npbf, nat, ngtg
      real(c_double), value :: pmass
      real(c_double) :: cat(*), zan(*), bcoef(*), gamma(*), agebfcc(*), agnbfcc(*), elcex(*), nucex(*)
      real(c_double) :: elcbfc(*), nucbfc(*)
      integer(c_int64_t) :: kpestr(*), kpeend(*), ampeb2c(*), elcam(*), nucam(*)
      integer(c_int64_t) :: handle
      integer(c_int) :: neorx_system_create
    end function
    function neorx_system_destroy(handle) bind(C, name="neorx_system_destroy")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: handle
      integer(c_int) :: neorx_system_destroy
    end function
    function neorx_gam1(handle, g, gs) bind(C, name="neorx_gam1")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: handle
      real(c_double) :: g(*), gs(*)
      integer(c_int) :: neorx_gam1
    end function
    function neorx_gam2_ic1(handle, g, gs) bind(C, name="neorx_gam2_ic1")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: handle
      real(c_double) :: g(*), gs(*)
      integer(c_int) :: neorx_gam2_ic1
    end function
    function neorx_gam3_ic1(handle, g) bind(C, name="neorx_gam3_ic1")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: handle
      real(c_double) :: g(*)
      integer(c_int) :: neorx_gam3_ic1
    end function
    function neorx_gam_ee(handle, g) bind(C, name="neorx_gam_ee")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: handle
      real(c_double) :: g(*)
      integer(c_int) :: neorx_gam_ee
    end function
    function neorx_gam4_icr(ne, np, g2s, gee, g4) bind(C, name="neorx_gam4_icr")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: g2s(*), gee(*), g4(*)
      integer(c_int) :: neorx_gam4_icr
    end function
    function neorx_eg2ic1(ne, np, de, dp, g, gs, fe, se, fp, sp, e, s) bind(C, name="neorx_eg2ic1")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: de(*), dp(*), g(*), gs(*), fe(*), se(*), fp(*), sp(*), e, s
      integer(c_int) :: neorx_eg2ic1
    end function
    function neorx_eg3ic1(ne, np, de, dp, g, fe, fp, e) bind(C, name="neorx_eg3ic1")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: de(*), dp(*), g(*), fe(*), fp(*), e
      integer(c_int) :: neorx_eg3ic1
    end function
    function neorx_eg4icr(ne, np, de, dp, g, fe, fp, e) bind(C, name="neorx_eg4icr")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: de(*), dp(*), g(*), fe(*), fp(*), e
      integer(c_int) :: neorx_eg4icr
    end function
    function neorx_gam2_dscf(handle, de, dp, fe, fp, se, sp, e, s) bind(C, name="neorx_gam2_dscf")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: handle
      real(c_double) :: de(*), dp(*), fe(*), fp(*), se(*), sp(*), e, s
      integer(c_int) :: neorx_gam2_dscf
    end function
    function neorx_gam3_dscf(handle, de, dp, fe, fp, e) bind(C, name="neorx_gam3_dscf")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: handle
      real(c_double) :: de(*), dp(*), fe(*), fp(*), e
      integer(c_int) :: neorx_gam3_dscf
    end function
    function neorx_gam4_dscf(ne, np, g2s, gee, de, dp, fe, fp, e) bind(C, name="neorx_gam4_dscf")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: g2s(*), gee(*), de(*), dp(*), fe(*), fp(*), e
      integer(c_int) :: neorx_gam4_dscf
    end function
    function neorx_fae_omg2(nae, nbe, ne, np, da, db, dp, g, gs, f, s2, e, s) bind(C, name="neorx_fae_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: nae, nbe, ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), gs(*), f(*), s2(*), e, s
      integer(c_int) :: neorx_fae_omg2
    end function
    function neorx_ufp_omg2(nae, nbe, ne, np, da, db, dp, g, gs, f, s2, e, s) bind(C, name="neorx_ufp_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: nae, nbe, ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), gs(*), f(*), s2(*), e, s
      integer(c_int) :: neorx_ufp_omg2
    end function
    function neorx_fae_omg3(nae, nbe, ne, np, da, db, dp, g, f, e) bind(C, name="neorx_fae_omg3")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: nae, nbe, ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_fae_omg3
    end function
    function neorx_ufp_omg3(nae, nbe, ne, np, da, db, dp, g, f, e) bind(C, name="neorx_ufp_omg3")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: nae, nbe, ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_ufp_omg3
    end function
    function neorx_fae_omg4(nae, nbe, ne, np, da, db, dp, g, f, e) bind(C, name="neorx_fae_omg4")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: nae, nbe, ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_fae_omg4
    end function
    function neorx_ufp_omg4(nae, nbe, ne, np, da, db, dp, g, f, e) bind(C, name="neorx_ufp_omg4")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: nae, nbe, ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_ufp_omg4
    end function
    function neorx_omg4_icr(ne, np, g2s, gee, g4) bind(C, name="neorx_omg4_icr")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: g2s(*), gee(*), g4(*)
      integer(c_int) :: neorx_omg4_icr
    end function
    function neorx_rxchf_gam2_ic1(h, g1, g2, gs) bind(C, name="neorx_rxchf_gam2_ic1")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      real(c_double) :: g1(*), g2(*), gs(*)
      integer(c_int) :: neorx_rxchf_gam2_ic1
    end function
    function neorx_rxchf_fae_omg2(ne, np, da, db, dp, g1, g2, gs, f, sf, e, sv) bind(C, name="neorx_rxchf_fae_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g1(*), g2(*), gs(*), f(*), sf(*), e, sv
      integer(c_int) :: neorx_rxchf_fae_omg2
    end function
    function neorx_rxchf_fp_omg2(ne, np, da, db, dp, g1, g2, gs, f, sf, e, sv) bind(C, name="neorx_rxchf_fp_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g1(*), g2(*), gs(*), f(*), sf(*), e, sv
      integer(c_int) :: neorx_rxchf_fp_omg2
    end function
    function neorx_rxchf_fbe_omg2(ne, np, da, db, dp, g1, g2, gs, f, sf, e, sv) bind(C, name="neorx_rxchf_fbe_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g1(*), g2(*), gs(*), f(*), sf(*), e, sv
      integer(c_int) :: neorx_rxchf_fbe_omg2
    end function
    function neorx_rxchf_gam3_ic1(h, g1, g2) bind(C, name="neorx_rxchf_gam3_ic1")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      real(c_double) :: g1(*), g2(*)
      integer(c_int) :: neorx_rxchf_gam3_ic1
    end function
    function neorx_rxchf_fae_omg3(ne, np, da, db, dp, g1, g2, f, e) bind(C, name="neorx_rxchf_fae_omg3")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g1(*), g2(*), f(*), e
      integer(c_int) :: neorx_rxchf_fae_omg3
    end function
    function neorx_rxchf_fp_omg3(ne, np, da, db, dp, g1, g2, f, e) bind(C, name="neorx_rxchf_fp_omg3")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g1(*), g2(*), f(*), e
      integer(c_int) :: neorx_rxchf_fp_omg3
    end function
    function neorx_rxchf_fbe_omg3(ne, np, da, db, dp, g1, g2, f, e) bind(C, name="neorx_rxchf_fbe_omg3")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g1(*), g2(*), f(*), e
      integer(c_int) :: neorx_rxchf_fbe_omg3
    end function
    function neorx_rxchf_fp_omg1(ne, np, db, dp, g, gs, f, sf, e, sv) bind(C, name="neorx_rxchf_fp_omg1")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: db(*), dp(*), g(*), gs(*), f(*), sf(*), e, sv
      integer(c_int) :: neorx_rxchf_fp_omg1
    end function
    function neorx_rxchf_fbe_omg1(ne, np, db, dp, g, gs, f, sf, e, sv) bind(C, name="neorx_rxchf_fbe_omg1")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: db(*), dp(*), g(*), gs(*), f(*), sf(*), e, sv
      integer(c_int) :: neorx_rxchf_fbe_omg1
    end function
    function neorx_rxcuhf_fe_omg2(ne, np, dl, dbt, dt, db, dp, g1, g2, gs, fa, fb, sa, e, sv) &
        bind(C, name="neorx_rxcuhf_fe_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: dl(*), dbt(*), dt(*), db(*), dp(*), g1(*), g2(*), gs(*), fa(*), fb(*), sa(*), e, sv
      integer(c_int) :: neorx_rxcuhf_fe_omg2
    end function
    function neorx_rxcuhf_fp_omg2(ne, np, dl, dt, db, dp, g1, g2, gs, f, sf, e, sv) bind(C, name="neorx_rxcuhf_fp_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: dl(*), dt(*), db(*), dp(*), g1(*), g2(*), gs(*), f(*), sf(*), e, sv
      integer(c_int) :: neorx_rxcuhf_fp_omg2
    end function
    function neorx_rxcuhf_fbe_omg2(ne, np, dl, dt, db, dp, g1, g2, gs, f, sf, e, sv) &
        bind(C, name="neorx_rxcuhf_fbe_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: dl(*), dt(*), db(*), dp(*), g1(*), g2(*), gs(*), f(*), sf(*), e, sv
      integer(c_int) :: neorx_rxcuhf_fbe_omg2
    end function
    function neorx_rxcuhf_fe_omg3(ne, np, dl, dbt, dt, db, dp, g1, g2, fa, fb, e) bind(C, name="neorx_rxcuhf_fe_omg3")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: dl(*), dbt(*), dt(*), db(*), dp(*), g1(*), g2(*), fa(*), fb(*), e
      integer(c_int) :: neorx_rxcuhf_fe_omg3
    end function
    function neorx_rxcuhf_fp_omg3(ne, np, dl, dbt, dt, db, dp, g1, g2, f, e) bind(C, name="neorx_rxcuhf_fp_omg3")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: dl(*), dbt(*), dt(*), db(*), dp(*), g1(*), g2(*), f(*), e
      integer(c_int) :: neorx_rxcuhf_fp_omg3
    end function
    function neorx_rxcuhf_fbe_omg3(ne, np, dl, dbt, dt, db, dp, g1, g2, f, e) bind(C, name="neorx_rxcuhf_fbe_omg3")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: dl(*), dbt(*), dt(*), db(*), dp(*), g1(*), g2(*), f(*), e
      integer(c_int) :: neorx_rxcuhf_fbe_omg3
    end function
    function neorx_rxchfne_gam2_ic1(h, g, gx) bind(C, name="neorx_rxchfne_gam2_ic1")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      real(c_double) :: g(*), gx(*)
      integer(c_int) :: neorx_rxchfne_gam2_ic1
    end function
    function neorx_rxchfne_fae_omg2(ne, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfne_fae_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfne_fae_omg2
    end function
    function neorx_rxchfne_fp_omg2(ne, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfne_fp_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfne_fp_omg2
    end function
    function neorx_rxchfne_fbe_omg2(ne, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfne_fbe_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfne_fbe_omg2
    end function
    function neorx_rxcuhfne_fe_omg2(ne, np, daa, dat, db, dp, g, gx, fa, fb, e) bind(C, name="neorx_rxcuhfne_fe_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: daa(*), dat(*), db(*), dp(*), g(*), gx(*), fa(*), fb(*), e
      integer(c_int) :: neorx_rxcuhfne_fe_omg2
    end function
    function neorx_rxcuhfne_fp_omg2(ne, np, daa, dat, db, dp, g, gx, f, e) bind(C, name="neorx_rxcuhfne_fp_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: daa(*), dat(*), db(*), dp(*), g(*), gx(*), f(*), e
      integer(c_int) :: neorx_rxcuhfne_fp_omg2
    end function
    function neorx_rxcuhfne_fbe_omg2(ne, np, daa, dat, db, dp, g, gx, f, e) bind(C, name="neorx_rxcuhfne_fbe_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: daa(*), dat(*), db(*), dp(*), g(*), gx(*), f(*), e
      integer(c_int) :: neorx_rxcuhfne_fbe_omg2
    end function
    function neorx_rxchfne_fae_gamee(ne, gee, da, f, e) bind(C, name="neorx_rxchfne_fae_gamee")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne
      real(c_double) :: gee(*), da(*), f(*), e
      integer(c_int) :: neorx_rxchfne_fae_gamee
    end function
    function neorx_mult_system_create(nebf, npebf, nebfbe, npebfbe, npbf, nat, ngtg, pmass, cat, zan, bcoef, gamma, &
                                      kpestr, kpeend, ampeb2c, agebfcc, elcex, elcam, elcbfc, kpestr_be, kpeend_be, &
                                      ampeb2c_be, agebfcc_be, elcex_be, elcam_be, elcbfc_be, agnbfcc, nucex, nucam, &
                                      nucbfc, handle) bind(C, name="neorx_mult_system_create")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: nebf, npebf, nebfbe, npebfbe, npbf, nat, ngtg
      real(c_double), value :: pmass
      real(c_double) :: cat(*), zan(*), bcoef(*), gamma(*), agebfcc(*), elcex(*), elcbfc(*)
      real(c_double) :: agebfcc_be(*), elcex_be(*), elcbfc_be(*), agnbfcc(*), nucex(*), nucbfc(*)
      integer(c_int64_t) :: kpestr(*), kpeend(*), ampeb2c(*), elcam(*), kpestr_be(*), kpeend_be(*)
      integer(c_int64_t) :: ampeb2c_be(*), elcam_be(*), nucam(*), handle
      integer(c_int) :: neorx_mult_system_create
    end function
    function neorx_rxchfmult_gam2(h, gi, gx, gs) bind(C, name="neorx_rxchfmult_gam2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      real(c_double) :: gi(*), gx(*), gs(*)
      integer(c_int) :: neorx_rxchfmult_gam2
    end function
    function neorx_rxchfmult_gam2_ic1ex(h, g1, g2, g3, gs) bind(C, name="neorx_rxchfmult_gam2_ic1ex")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      real(c_double) :: g1(*), g2(*), g3(*), gs(*)
      integer(c_int) :: neorx_rxchfmult_gam2_ic1ex
    end function
    function neorx_rxchfmult_gam3_ic1ex(h, g2, g3, g4) bind(C, name="neorx_rxchfmult_gam3_ic1ex")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      real(c_double) :: g2(*), g3(*), g4(*)
      integer(c_int) :: neorx_rxchfmult_gam3_ic1ex
    end function
    function neorx_rxchfmult_fae_omg2ex(ne, np, da, db, dp, g, gx, f, e) bind(C, name="neorx_rxchfmult_fae_omg2ex")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), gx(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fae_omg2ex
    end function
    function neorx_rxchfmult_fae_omg3ex(ne, np, da, db, dp, g, gx1, gx2, f, e) bind(C, name="neorx_rxchfmult_fae_omg3ex")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), gx1(*), gx2(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fae_omg3ex
    end function
    function neorx_rxchfmult_fp_omg2ex(ne, np, da, db, dp, g, gx, f, e) bind(C, name="neorx_rxchfmult_fp_omg2ex")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), gx(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fp_omg2ex
    end function
    function neorx_rxchfmult_fp_omg3ex(ne, np, da, db, dp, g, gx1, gx2, f, e) bind(C, name="neorx_rxchfmult_fp_omg3ex")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), gx1(*), gx2(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fp_omg3ex
    end function
    function neorx_rxchfmult_fbe_omg2ex(ne, np, da, db, dp, g, gx, f, e) bind(C, name="neorx_rxchfmult_fbe_omg2ex")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), gx(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fbe_omg2ex
    end function
    function neorx_rxchfmult_fbe_omg3ex(ne, np, da, db, dp, g, gx1, gx2, f, e) bind(C, name="neorx_rxchfmult_fbe_omg3ex")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: ne, np
      real(c_double) :: da(*), db(*), dp(*), g(*), gx1(*), gx2(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fbe_omg3ex
    end function
    function neorx_rxchfmult_gam1(h, g, gs) bind(C, name="neorx_rxchfmult_gam1")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      real(c_double) :: g(*), gs(*)
      integer(c_int) :: neorx_rxchfmult_gam1
    end function
    function neorx_rxchfmult_gam3_xchf(h, g) bind(C, name="neorx_rxchfmult_gam3_xchf")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      real(c_double) :: g(*)
      integer(c_int) :: neorx_rxchfmult_gam3_xchf
    end function
    function neorx_rxchfmult_gam3_int(h, g) bind(C, name="neorx_rxchfmult_gam3_int")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      real(c_double) :: g(*)
      integer(c_int) :: neorx_rxchfmult_gam3_int
    end function
    function neorx_rxchfmult_gam4_int(na, nb, np, g2s, gee, g4) bind(C, name="neorx_rxchfmult_gam4_int")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: na, nb, np
      real(c_double) :: g2s(*), gee(*), g4(*)
      integer(c_int) :: neorx_rxchfmult_gam4_int
    end function
    function neorx_rxchfmult_fae_omg2(na, nb, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfmult_fae_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: na, nb, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fae_omg2
    end function
    function neorx_rxchfmult_fp_omg2(na, nb, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfmult_fp_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: na, nb, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fp_omg2
    end function
    function neorx_rxchfmult_fbe_omg2(na, nb, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfmult_fbe_omg2")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: na, nb, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fbe_omg2
    end function
    function neorx_rxchfmult_fae_omg3(na, nb, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfmult_fae_omg3")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: na, nb, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fae_omg3
    end function
    function neorx_rxchfmult_fp_omg3(na, nb, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfmult_fp_omg3")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: na, nb, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fp_omg3
    end function
    function neorx_rxchfmult_fbe_omg3(na, nb, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfmult_fbe_omg3")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: na, nb, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fbe_omg3
    end function
    function neorx_rxchfmult_fae_omg4(na, nb, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfmult_fae_omg4")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: na, nb, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fae_omg4
    end function
    function neorx_rxchfmult_fp_omg4(na, nb, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfmult_fp_omg4")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: na, nb, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fp_omg4
    end function
    function neorx_rxchfmult_fbe_omg4(na, nb, np, da, db, dp, g, f, e) bind(C, name="neorx_rxchfmult_fbe_omg4")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: na, nb, np
      real(c_double) :: da(*), db(*), dp(*), g(*), f(*), e
      integer(c_int) :: neorx_rxchfmult_fbe_omg4
    end function
  end interface
contains
  subroutine neorx_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    type(c_ptr) :: p
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    p = neorx_last_error()
    n = 0
    call c_f_pointer(p, msg, [512])
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(0, '(a,a,i0,a,512a)') where, ': neorx error ', rc, ': ', msg(1:n)
    error stop 1
  end subroutine

  ! contracted-function ranges from the primitive -> contracted map
  ! (the primitives of one contracted function are contiguous, make_KPE
  ! source/xcneo_HYBRID.f:1269-1307)
  subroutine neorx_kpe(nebf, npebf, ampeb2c, kpestr, kpeend)
    integer(c_int64_t), intent(in) :: nebf, npebf, ampeb2c(npebf)
    integer(c_int64_t), intent(out) :: kpestr(nebf), kpeend(nebf)
    integer(c_int64_t) :: i, c
    kpestr = 0
    kpeend = -1
    do i = 1, npebf
      c = ampeb2c(i)
      if (kpestr(c) == 0) kpestr(c) = i
      kpeend(c) = i
    end do
  end subroutine
  ! Record length of the .ufm direct-access files, in bytes.  The reference
  ! opens them with RECL=8 (e.g. source/gam_ee.f:143-144): flang and gfortran
  ! count RECL in bytes (8-byte records), ifort in 4-byte words (32-byte
  ! records, each double followed by 24 zero bytes; the reference's own
  ! tests/TEST_XCRHF/TEST_06/novlap.ufm is laid out so).  NEORX_UFM_RECL=32
  ! makes these files match an ifort-built driver; 8 is the default.
  integer(c_int64_t) function ufm_recl()
    character(len=32) :: v
    integer :: n, st
    ufm_recl = 8
    call get_environment_variable('NEORX_UFM_RECL', v, n, st)
    if (st == 0 .and. n > 0) read(v, *) ufm_recl
    if (ufm_recl /= 8 .and. ufm_recl /= 32) then
      write(0, '(a,a)') 'neorx: NEORX_UFM_RECL must be 8 or 32, got ', trim(v)
      error stop 1
    end if
  end function

  ! write a(1:n) to fname, one double per record
  subroutine ufm_write(u, fname, n, a)
    integer, intent(in) :: u
    character(len=*), intent(in) :: fname
    integer(c_int64_t), intent(in) :: n
    real(c_double), intent(in) :: a(n)
    integer(c_int64_t) :: i, rl
    rl = ufm_recl()
    open(u, file=fname, form='unformatted', status='unknown', access='direct', recl=rl)
    do i = 1, n
      if (rl == 8) then
        write(u, rec=i) a(i)
      else
        write(u, rec=i) a(i), 0.0d0, 0.0d0, 0.0d0
      end if
    end do
    close(u)
  end subroutine

  ! read a(1:n) from fname, one double per record
  subroutine ufm_read(u, fname, n, a)
    integer, intent(in) :: u
    character(len=*), intent(in) :: fname
    integer(c_int64_t), intent(in) :: n
    real(c_double), intent(out) :: a(n)
    integer(c_int64_t) :: i
    open(u, file=fname, form='unformatted', status='old', access='direct', recl=ufm_recl())
    do i = 1, n
      read(u, rec=i) a(i)
    end do
    close(u)
  end subroutine
end module neorx_capi

! GAM1_OMP_MD, source/gam_1_OMP.f:2-397: e-p integrals written to
! GAM_1.ufm / GAM_1s.ufm (direct access, one double per record) in the
! pack_4D(nebf,nebf,npbf,jec1,iec1,jp,ip) order.
subroutine GAM1_OMP_MD(nebf, npebf, npbf, ng1, ng1prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                       AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: nebf, npebf, npbf, ng1, ng1prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  integer(c_int64_t) :: h, kpestr(nebf), kpeend(nebf)
  real(c_double), allocatable :: g(:), gs(:)
  call neorx_kpe(nebf, npebf, AMPEB2C, kpestr, kpeend)
  allocate(g(ng1), gs(ng1))
  call neorx_check(neorx_system_create(nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, kpestr, &
                   kpeend, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, h), 'GAM1_OMP_MD')
  call neorx_check(neorx_gam1(h, g, gs), 'GAM1_OMP_MD')
  call neorx_check(neorx_system_destroy(h), 'GAM1_OMP_MD')
  call ufm_write(801, 'GAM_1.ufm', ng1, g)
  call ufm_write(802, 'GAM_1s.ufm', ng1, gs)
  deallocate(g, gs)
end subroutine

! calc_GAM_ee, source/gam_ee.f:2-168: contracted ERIs written to GAM_ee.ufm
subroutine calc_GAM_ee(nebf, npebf, ngee, AMPEB2C, AGEBFCC, ELCEX, ELCAM, ELCBFC)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: nebf, npebf, ngee
  integer(c_int64_t) :: AMPEB2C(npebf), ELCAM(npebf, 3)
  real(c_double) :: AGEBFCC(npebf), ELCEX(npebf), ELCBFC(npebf, 3)
  integer(c_int64_t) :: h, kpestr(nebf), kpeend(nebf), nucam(3)
  real(c_double) :: one(1), zero3(3), g(ngee)
  call neorx_kpe(nebf, npebf, AMPEB2C, kpestr, kpeend)
  one = 1.0d0
  zero3 = 0.0d0
  nucam = 0
  call neorx_check(neorx_system_create(nebf, npebf, 1_c_int64_t, 0_c_int64_t, 1_c_int64_t, 1.0d0, zero3, one, one, &
                   one, kpestr, kpeend, AMPEB2C, AGEBFCC, one, ELCEX, one, ELCAM, nucam, ELCBFC, zero3, h), &
                   'calc_GAM_ee')
  call neorx_check(neorx_gam_ee(h, g), 'calc_GAM_ee')
  call neorx_check(neorx_system_destroy(h), 'calc_GAM_ee')
  call ufm_write(810, 'GAM_ee.ufm', ngee, g)
end subroutine

! GAM2_IC1, source/gam2_IC1_OMP.f:2-212
subroutine GAM2_IC1(Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                    KPESTR, KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, &
                    GM2ICR, GM2sICR)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  real(c_double) :: GM2ICR(ng2), GM2sICR(ng2)
  integer(c_int64_t) :: h
  call neorx_check(neorx_system_create(nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, KPESTR, &
                   KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, h), 'GAM2_IC1')
  call neorx_check(neorx_gam2_ic1(h, GM2ICR, GM2sICR), 'GAM2_IC1')
  call neorx_check(neorx_system_destroy(h), 'GAM2_IC1')
end subroutine

! GAM3_IC1, source/gam3_IC1_OMP.f:2-298
subroutine GAM3_IC1(Nchunks, nebf, npebf, npbf, ng3, ng3prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                    KPESTR, KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, GM3ICR)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, npbf, ng3, ng3prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  real(c_double) :: GM3ICR(ng3)
  integer(c_int64_t) :: h
  call neorx_check(neorx_system_create(nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, KPESTR, &
                   KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, h), 'GAM3_IC1')
  call neorx_check(neorx_gam3_ic1(h, GM3ICR), 'GAM3_IC1')
  call neorx_check(neorx_system_destroy(h), 'GAM3_IC1')
end subroutine

! GAM4_ICR, source/gam4_ICR_OMP.f:2-138: reads GAM_ee.ufm from the working
! directory like read_GAM_ee (source/xchf1_independent/GAM_IO.f:28)
subroutine GAM4_ICR(Nchunks, ne, np, ngee, ng2, ng4, GAM_2s, GAM_4)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, ne, np, ngee, ng2, ng4
  real(c_double) :: GAM_2s(ng2), GAM_4(ng4)
  real(c_double), allocatable :: gee(:)
  allocate(gee(ngee))
  call ufm_read(810, 'GAM_ee.ufm', ngee, gee)
  call neorx_check(neorx_gam4_icr(ne, np, GAM_2s, gee, GAM_4), 'GAM4_ICR')
  deallocate(gee)
end subroutine

! EG2IC1, source/gam2_IC1_OMP.f:339-443
subroutine EG2IC1(Nchunks, nebf, npbf, ng2, DE, DP, GM2ICR, GM2SICR, FE, SE, FP, SP, E_gam2, S_gam2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2
  real(c_double) :: DE(nebf, nebf), DP(npbf, npbf), GM2ICR(ng2), GM2SICR(ng2)
  real(c_double) :: FE(nebf, nebf), SE(nebf, nebf), FP(npbf, npbf), SP(npbf, npbf), E_gam2, S_gam2
  call neorx_check(neorx_eg2ic1(nebf, npbf, DE, DP, GM2ICR, GM2SICR, FE, SE, FP, SP, E_gam2, S_gam2), 'EG2IC1')
end subroutine

! EG3IC1, source/gam3_IC1_OMP.f:433-534
subroutine EG3IC1(Nchunks, nebf, npbf, ng3, DE, DP, GM3ICR, focke, fockp, E_gam3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3
  real(c_double) :: DE(nebf, nebf), DP(npbf, npbf), GM3ICR(ng3), focke(nebf, nebf), fockp(npbf, npbf), E_gam3
  call neorx_check(neorx_eg3ic1(nebf, npbf, DE, DP, GM3ICR, focke, fockp, E_gam3), 'EG3IC1')
end subroutine

! EG4ICR, source/gam4_ICR_OMP.f:273-376
subroutine EG4ICR(Nchunks, nebf, npbf, ng4, DE, DP, GM4ICR, focke, fockp, E_gam4)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng4
  real(c_double) :: DE(nebf, nebf), DP(npbf, npbf), GM4ICR(ng4), focke(nebf, nebf), fockp(npbf, npbf), E_gam4
  call neorx_check(neorx_eg4icr(nebf, npbf, DE, DP, GM4ICR, focke, fockp, E_gam4), 'EG4ICR')
end subroutine

! GAM2_DSCF, source/gam2_DSCF_OMP.f:2-139
subroutine GAM2_DSCF(Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                     KPESTR, KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, &
                     DE, DP, focke, fockp, se2, sp2, E_gam2, S_gam2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  real(c_double) :: DE(nebf, nebf), DP(npbf, npbf), focke(nebf, nebf), fockp(npbf, npbf)
  real(c_double) :: se2(nebf, nebf), sp2(npbf, npbf), E_gam2, S_gam2
  integer(c_int64_t) :: h
  call neorx_check(neorx_system_create(nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, KPESTR, &
                   KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, h), 'GAM2_DSCF')
  call neorx_check(neorx_gam2_dscf(h, DE, DP, focke, fockp, se2, sp2, E_gam2, S_gam2), 'GAM2_DSCF')
  call neorx_check(neorx_system_destroy(h), 'GAM2_DSCF')
end subroutine

! GAM3_DSCF, source/gam3_DSCF_OMP.f:2-138
subroutine GAM3_DSCF(Nchunks, nebf, npebf, npbf, ng3, ng3prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                     KPESTR, KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, &
                     DE, DP, focke, fockp, E_gam3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, npbf, ng3, ng3prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  real(c_double) :: DE(nebf, nebf), DP(npbf, npbf), focke(nebf, nebf), fockp(npbf, npbf), E_gam3
  integer(c_int64_t) :: h
  call neorx_check(neorx_system_create(nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, KPESTR, &
                   KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, h), 'GAM3_DSCF')
  call neorx_check(neorx_gam3_dscf(h, DE, DP, focke, fockp, E_gam3), 'GAM3_DSCF')
  call neorx_check(neorx_system_destroy(h), 'GAM3_DSCF')
end subroutine

! GAM4_DSCF, source/gam4_DSCF_OMP.f:2-104: reads GAM_ee.ufm and GAM_2s.ufm
! from the working directory like read_GAM_ee / read_GAM_2s
! (source/xchf1_independent/GAM_IO.f:28, source/gam_4_OMP.f:1195)
subroutine GAM4_DSCF(NG4CHK, ne, np, ngee, ng2, ng4, DE, DP, focke, fockp, E_gam4)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: NG4CHK, ne, np, ngee, ng2, ng4
  real(c_double) :: DE(ne, ne), DP(np, np), focke(ne, ne), fockp(np, np), E_gam4
  real(c_double), allocatable :: gee(:), g2s(:)
  allocate(gee(ngee), g2s(ng2))
  call ufm_read(810, 'GAM_ee.ufm', ngee, gee)
  call ufm_read(804, 'GAM_2s.ufm', ng2, g2s)
  call neorx_check(neorx_gam4_dscf(ne, np, g2s, gee, DE, DP, focke, fockp, E_gam4), 'GAM4_DSCF')
  deallocate(gee, g2s)
end subroutine

! ---------------------------------------------------------------------
! XCUHF contractions (source/xcuhf/xcuhf_omg{2,3,4}.f)
! ---------------------------------------------------------------------
! FAE_OMG2, source/xcuhf/xcuhf_omg2.f:2-136 (DETOT is not used by it)
subroutine FAE_OMG2(Nchunks, NAE, NBE, nebf, npbf, ng2, DAE, DBE, DETOT, DP, GM2ICR, GM2SICR, FAE, SAE2, &
                    E_AE_OMG2, S_AE_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, NAE, NBE, nebf, npbf, ng2
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DETOT(nebf, nebf), DP(npbf, npbf), GM2ICR(ng2), GM2SICR(ng2)
  real(c_double) :: FAE(nebf, nebf), SAE2(nebf, nebf), E_AE_OMG2, S_AE_OMG2
  call neorx_check(neorx_fae_omg2(NAE, NBE, nebf, npbf, DAE, DBE, DP, GM2ICR, GM2SICR, FAE, SAE2, E_AE_OMG2, &
                   S_AE_OMG2), 'FAE_OMG2')
end subroutine

! UFP_OMG2, source/xcuhf/xcuhf_omg2.f:302-440
subroutine UFP_OMG2(Nchunks, nebf, npbf, ng2, NAE, NBE, DAE, DBE, DETOT, DP, GM2ICR, GM2SICR, FP, SP2, &
                    E_P_OMG2, S_P_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2, NAE, NBE
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DETOT(nebf, nebf), DP(npbf, npbf), GM2ICR(ng2), GM2SICR(ng2)
  real(c_double) :: FP(npbf, npbf), SP2(npbf, npbf), E_P_OMG2, S_P_OMG2
  call neorx_check(neorx_ufp_omg2(NAE, NBE, nebf, npbf, DAE, DBE, DP, GM2ICR, GM2SICR, FP, SP2, E_P_OMG2, &
                   S_P_OMG2), 'UFP_OMG2')
end subroutine

! FAE_OMG3, source/xcuhf/xcuhf_omg3.f:2-128
subroutine FAE_OMG3(Nchunks, nebf, npbf, ng3, NAE, NBE, DAE, DBE, DP, GM3ICR, FAE, E_AE_OMG3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3, NAE, NBE
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM3ICR(ng3), FAE(nebf, nebf), E_AE_OMG3
  call neorx_check(neorx_fae_omg3(NAE, NBE, nebf, npbf, DAE, DBE, DP, GM3ICR, FAE, E_AE_OMG3), 'FAE_OMG3')
end subroutine

! UFP_OMG3, source/xcuhf/xcuhf_omg3.f:284-407
subroutine UFP_OMG3(Nchunks, nebf, npbf, ng3, NAE, NBE, DAE, DBE, DP, GM3ICR, FP, E_P_OMG3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3, NAE, NBE
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM3ICR(ng3), FP(npbf, npbf), E_P_OMG3
  call neorx_check(neorx_ufp_omg3(NAE, NBE, nebf, npbf, DAE, DBE, DP, GM3ICR, FP, E_P_OMG3), 'UFP_OMG3')
end subroutine

! FAE_OMG4, source/xcuhf/xcuhf_omg4.f:2-124
subroutine FAE_OMG4(Nchunks, nebf, npbf, ng4, NAE, NBE, DAE, DBE, DP, GM4ICR, FAE, E_AE_OMG4)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng4, NAE, NBE
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM4ICR(ng4), FAE(nebf, nebf), E_AE_OMG4
  call neorx_check(neorx_fae_omg4(NAE, NBE, nebf, npbf, DAE, DBE, DP, GM4ICR, FAE, E_AE_OMG4), 'FAE_OMG4')
end subroutine

! UFP_OMG4, source/xcuhf/xcuhf_omg4.f:512-636
subroutine UFP_OMG4(Nchunks, nebf, npbf, ng4, NAE, NBE, DAE, DBE, DP, GM4ICR, FP, E_P_OMG4)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng4, NAE, NBE
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM4ICR(ng4), FP(npbf, npbf), E_P_OMG4
  call neorx_check(neorx_ufp_omg4(NAE, NBE, nebf, npbf, DAE, DBE, DP, GM4ICR, FP, E_P_OMG4), 'UFP_OMG4')
end subroutine

! OMG4_ICR, source/xcuhf/xcuhf_omg4.f:1080-1210: reads GAM_ee.ufm from the
! working directory like read_GAM_ee
subroutine OMG4_ICR(Nchunks, ne, np, ngee, ng2, ng4, GAM_2s, GAM_4)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, ne, np, ngee, ng2, ng4
  real(c_double) :: GAM_2s(ng2), GAM_4(ng4)
  real(c_double), allocatable :: gee(:)
  allocate(gee(ngee))
  call ufm_read(810, 'GAM_ee.ufm', ngee, gee)
  call neorx_check(neorx_omg4_icr(ne, np, GAM_2s, gee, GAM_4), 'OMG4_ICR')
  deallocate(gee)
end subroutine

! ---------------------------------------------------------------------
! RXCHF-ne (source/rxchfne/)
! ---------------------------------------------------------------------
! RXCHFne_GAM2_IC1, source/rxchfne/rxchfne_gam2_IC1_OMP.f:2-243
subroutine RXCHFne_GAM2_IC1(Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                            KPESTR, KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, &
                            GM2ICR, GM2exICR)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  real(c_double) :: GM2ICR(ng2), GM2exICR(ng2)
  integer(c_int64_t) :: h
  call neorx_check(neorx_system_create(nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, KPESTR, &
                   KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, h), 'RXCHFne_GAM2_IC1')
  call neorx_check(neorx_rxchfne_gam2_ic1(h, GM2ICR, GM2exICR), 'RXCHFne_GAM2_IC1')
  call neorx_check(neorx_system_destroy(h), 'RXCHFne_GAM2_IC1')
end subroutine

! RXCHF_GAM2_IC1, source/rxchf/rxchf_gam2_IC1_OMP.f:2-250
subroutine RXCHF_GAM2_IC1(Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                          KPESTR, KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, &
                          GM2_1ICR, GM2_2ICR, GM2sICR)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  real(c_double) :: GM2_1ICR(ng2), GM2_2ICR(ng2), GM2sICR(ng2)
  integer(c_int64_t) :: h
  call neorx_check(neorx_system_create(nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, KPESTR, &
                   KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, h), 'RXCHF_GAM2_IC1')
  call neorx_check(neorx_rxchf_gam2_ic1(h, GM2_1ICR, GM2_2ICR, GM2sICR), 'RXCHF_GAM2_IC1')
  call neorx_check(neorx_system_destroy(h), 'RXCHF_GAM2_IC1')
end subroutine

! RXCHF_FAE_OMG2, source/rxchf/rxchf_omg2.f:2-110
subroutine RXCHF_FAE_OMG2(Nchunks, nebf, npbf, ng2, DAE, DBE, DP, GM2_1ICR, GM2_2ICR, GM2sICR, FAE, SAE, &
                          E_AE_OMG2, S_AE_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM2_1ICR(ng2), GM2_2ICR(ng2), GM2sICR(ng2)
  real(c_double) :: FAE(nebf, nebf), SAE(nebf, nebf), E_AE_OMG2, S_AE_OMG2
  call neorx_check(neorx_rxchf_fae_omg2(nebf, npbf, DAE, DBE, DP, GM2_1ICR, GM2_2ICR, GM2sICR, FAE, SAE, E_AE_OMG2, &
                   S_AE_OMG2), 'RXCHF_FAE_OMG2')
end subroutine

! RXCHF_FP_OMG2, source/rxchf/rxchf_omg2.f:231-330
subroutine RXCHF_FP_OMG2(Nchunks, nebf, npbf, ng2, DAE, DBE, DP, GM2_1ICR, GM2_2ICR, GM2sICR, FP, SP, E_P_OMG2, &
                         S_P_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM2_1ICR(ng2), GM2_2ICR(ng2), GM2sICR(ng2)
  real(c_double) :: FP(npbf, npbf), SP(npbf, npbf), E_P_OMG2, S_P_OMG2
  call neorx_check(neorx_rxchf_fp_omg2(nebf, npbf, DAE, DBE, DP, GM2_1ICR, GM2_2ICR, GM2sICR, FP, SP, E_P_OMG2, &
                   S_P_OMG2), 'RXCHF_FP_OMG2')
end subroutine

! RXCHF_FBE_OMG2, source/rxchf/rxchf_omg2.f:449-548
subroutine RXCHF_FBE_OMG2(Nchunks, nebf, npbf, ng2, DAE, DBE, DP, GM2_1ICR, GM2_2ICR, GM2sICR, FBE, SBE, &
                          E_BE_OMG2, S_BE_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM2_1ICR(ng2), GM2_2ICR(ng2), GM2sICR(ng2)
  real(c_double) :: FBE(nebf, nebf), SBE(nebf, nebf), E_BE_OMG2, S_BE_OMG2
  call neorx_check(neorx_rxchf_fbe_omg2(nebf, npbf, DAE, DBE, DP, GM2_1ICR, GM2_2ICR, GM2sICR, FBE, SBE, E_BE_OMG2, &
                   S_BE_OMG2), 'RXCHF_FBE_OMG2')
end subroutine

! RXCHF_GAM1_OMP_MD, source/rxchf/rxchf_gam_1_OMP.f:2-397: the same
! computation as RXCHFmult_GAM1_MD (GAM_1.ufm / GAM_1s.ufm)
subroutine RXCHF_GAM1_OMP_MD(nebf, npebf, npbf, ng1, ng1prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                             AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: nebf, npebf, npbf, ng1, ng1prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  call RXCHFmult_GAM1_MD(nebf, npebf, npbf, ng1, ng1prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                         AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC)
end subroutine

! RXCHF_GAM3_IC1, source/rxchf/rxchf_gam3_IC1_OMP.f:2-283
subroutine RXCHF_GAM3_IC1(Nchunks, nebf, npebf, npbf, ng3, ng3prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                          KPESTR, KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, &
                          GM3_1ICR, GM3_2ICR)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, npbf, ng3, ng3prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  real(c_double) :: GM3_1ICR(ng3), GM3_2ICR(ng3)
  integer(c_int64_t) :: h
  call neorx_check(neorx_system_create(nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, KPESTR, &
                   KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, h), 'RXCHF_GAM3_IC1')
  call neorx_check(neorx_rxchf_gam3_ic1(h, GM3_1ICR, GM3_2ICR), 'RXCHF_GAM3_IC1')
  call neorx_check(neorx_system_destroy(h), 'RXCHF_GAM3_IC1')
end subroutine

! RXCHF_FAE_OMG3, source/rxchf/rxchf_omg3.f:2-105
subroutine RXCHF_FAE_OMG3(Nchunks, nebf, npbf, ng3, DAE, DBE, DP, GM3_1ICR, GM3_2ICR, FAE, E_AE_OMG3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM3_1ICR(ng3), GM3_2ICR(ng3)
  real(c_double) :: FAE(nebf, nebf), E_AE_OMG3
  call neorx_check(neorx_rxchf_fae_omg3(nebf, npbf, DAE, DBE, DP, GM3_1ICR, GM3_2ICR, FAE, E_AE_OMG3), &
                   'RXCHF_FAE_OMG3')
end subroutine

! RXCHF_FP_OMG3, source/rxchf/rxchf_omg3.f:263-366
subroutine RXCHF_FP_OMG3(Nchunks, nebf, npbf, ng3, DAE, DBE, DP, GM3_1ICR, GM3_2ICR, FP, E_P_OMG3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM3_1ICR(ng3), GM3_2ICR(ng3)
  real(c_double) :: FP(npbf, npbf), E_P_OMG3
  call neorx_check(neorx_rxchf_fp_omg3(nebf, npbf, DAE, DBE, DP, GM3_1ICR, GM3_2ICR, FP, E_P_OMG3), 'RXCHF_FP_OMG3')
end subroutine

! RXCHF_FBE_OMG3, source/rxchf/rxchf_omg3.f:486-589
subroutine RXCHF_FBE_OMG3(Nchunks, nebf, npbf, ng3, DAE, DBE, DP, GM3_1ICR, GM3_2ICR, FBE, E_BE_OMG3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM3_1ICR(ng3), GM3_2ICR(ng3)
  real(c_double) :: FBE(nebf, nebf), E_BE_OMG3
  call neorx_check(neorx_rxchf_fbe_omg3(nebf, npbf, DAE, DBE, DP, GM3_1ICR, GM3_2ICR, FBE, E_BE_OMG3), &
                   'RXCHF_FBE_OMG3')
end subroutine

! RXCHF_FP_OMG1, source/rxchf/rxchf_omg1.f:2-82: reads GAM_1.ufm / GAM_1s.ufm
subroutine RXCHF_FP_OMG1(nebf, npbf, ng1, DBE, DP, FP, SP, E_P_OMG1, S_P_OMG1)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: nebf, npbf, ng1
  real(c_double) :: DBE(nebf, nebf), DP(npbf, npbf), FP(npbf, npbf), SP(npbf, npbf), E_P_OMG1, S_P_OMG1
  real(c_double), allocatable :: g(:), gs(:)
  allocate(g(ng1), gs(ng1))
  call ufm_read(801, 'GAM_1.ufm', ng1, g)
  call ufm_read(802, 'GAM_1s.ufm', ng1, gs)
  call neorx_check(neorx_rxchf_fp_omg1(nebf, npbf, DBE, DP, g, gs, FP, SP, E_P_OMG1, S_P_OMG1), 'RXCHF_FP_OMG1')
  deallocate(g, gs)
end subroutine

! RXCHF_FBE_OMG1, source/rxchf/rxchf_omg1.f:84-165: reads GAM_1.ufm / GAM_1s.ufm
subroutine RXCHF_FBE_OMG1(nebf, npbf, ng1, DBE, DP, FBE, SBE, E_BE_OMG1, S_BE_OMG1)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: nebf, npbf, ng1
  real(c_double) :: DBE(nebf, nebf), DP(npbf, npbf), FBE(nebf, nebf), SBE(nebf, nebf), E_BE_OMG1, S_BE_OMG1
  real(c_double), allocatable :: g(:), gs(:)
  allocate(g(ng1), gs(ng1))
  call ufm_read(801, 'GAM_1.ufm', ng1, g)
  call ufm_read(802, 'GAM_1s.ufm', ng1, gs)
  call neorx_check(neorx_rxchf_fbe_omg1(nebf, npbf, DBE, DP, g, gs, FBE, SBE, E_BE_OMG1, S_BE_OMG1), &
                   'RXCHF_FBE_OMG1')
  deallocate(g, gs)
end subroutine

! RXCUHF_FE_OMG2, source/rxchf/rxcuhf_omg2.f:2-126
subroutine RXCUHF_FE_OMG2(Nchunks, nebf, npbf, ng2, DAalpE, DAbetE, DAtotE, DBE, DP, GM2_1ICR, GM2_2ICR, GM2sICR, &
                          FAalpE, FAbetE, SAalpE, E_AE_OMG2, S_AE_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2
  real(c_double) :: DAalpE(nebf, nebf), DAbetE(nebf, nebf), DAtotE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf)
  real(c_double) :: GM2_1ICR(ng2), GM2_2ICR(ng2), GM2sICR(ng2)
  real(c_double) :: FAalpE(nebf, nebf), FAbetE(nebf, nebf), SAalpE(nebf, nebf), E_AE_OMG2, S_AE_OMG2
  call neorx_check(neorx_rxcuhf_fe_omg2(nebf, npbf, DAalpE, DAbetE, DAtotE, DBE, DP, GM2_1ICR, GM2_2ICR, GM2sICR, &
                   FAalpE, FAbetE, SAalpE, E_AE_OMG2, S_AE_OMG2), 'RXCUHF_FE_OMG2')
end subroutine

! RXCUHF_FP_OMG2, source/rxchf/rxcuhf_omg2.f:253-352
subroutine RXCUHF_FP_OMG2(Nchunks, nebf, npbf, ng2, NAalpE, NBE, DAalpE, DAtotE, DBE, DP, GM2_1ICR, GM2_2ICR, &
                          GM2sICR, FP, SP, E_P_OMG2, S_P_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2, NAalpE, NBE
  real(c_double) :: DAalpE(nebf, nebf), DAtotE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf)
  real(c_double) :: GM2_1ICR(ng2), GM2_2ICR(ng2), GM2sICR(ng2), FP(npbf, npbf), SP(npbf, npbf), E_P_OMG2, S_P_OMG2
  call neorx_check(neorx_rxcuhf_fp_omg2(nebf, npbf, DAalpE, DAtotE, DBE, DP, GM2_1ICR, GM2_2ICR, GM2sICR, FP, SP, &
                   E_P_OMG2, S_P_OMG2), 'RXCUHF_FP_OMG2')
end subroutine

! RXCUHF_FBE_OMG2, source/rxchf/rxcuhf_omg2.f:469-576
subroutine RXCUHF_FBE_OMG2(Nchunks, nebf, npbf, ng2, NAalpE, NBE, DAalpE, DAtotE, DBE, DP, GM2_1ICR, GM2_2ICR, &
                           GM2sICR, FBE, SBE, E_BE_OMG2, S_BE_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2, NAalpE, NBE
  real(c_double) :: DAalpE(nebf, nebf), DAtotE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf)
  real(c_double) :: GM2_1ICR(ng2), GM2_2ICR(ng2), GM2sICR(ng2), FBE(nebf, nebf), SBE(nebf, nebf), E_BE_OMG2, S_BE_OMG2
  call neorx_check(neorx_rxcuhf_fbe_omg2(nebf, npbf, DAalpE, DAtotE, DBE, DP, GM2_1ICR, GM2_2ICR, GM2sICR, FBE, SBE, &
                   E_BE_OMG2, S_BE_OMG2), 'RXCUHF_FBE_OMG2')
end subroutine

! RXCUHF_FE_OMG3, source/rxchf/rxcuhf_omg3.f:2-107
subroutine RXCUHF_FE_OMG3(Nchunks, nebf, npbf, ng3, DAalpE, DAbetE, DAtotE, DBE, DP, GM3_1ICR, GM3_2ICR, FAalpE, &
                          FAbetE, E_AE_OMG3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3
  real(c_double) :: DAalpE(nebf, nebf), DAbetE(nebf, nebf), DAtotE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf)
  real(c_double) :: GM3_1ICR(ng3), GM3_2ICR(ng3), FAalpE(nebf, nebf), FAbetE(nebf, nebf), E_AE_OMG3
  call neorx_check(neorx_rxcuhf_fe_omg3(nebf, npbf, DAalpE, DAbetE, DAtotE, DBE, DP, GM3_1ICR, GM3_2ICR, FAalpE, &
                   FAbetE, E_AE_OMG3), 'RXCUHF_FE_OMG3')
end subroutine

! RXCUHF_FP_OMG3, source/rxchf/rxcuhf_omg3.f:270-371
subroutine RXCUHF_FP_OMG3(Nchunks, nebf, npbf, ng3, DAalpE, DAbetE, DAtotE, DBE, DP, GM3_1ICR, GM3_2ICR, FP, E_P_OMG3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3
  real(c_double) :: DAalpE(nebf, nebf), DAbetE(nebf, nebf), DAtotE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf)
  real(c_double) :: GM3_1ICR(ng3), GM3_2ICR(ng3), FP(npbf, npbf), E_P_OMG3
  call neorx_check(neorx_rxcuhf_fp_omg3(nebf, npbf, DAalpE, DAbetE, DAtotE, DBE, DP, GM3_1ICR, GM3_2ICR, FP, &
                   E_P_OMG3), 'RXCUHF_FP_OMG3')
end subroutine

! RXCUHF_FBE_OMG3, source/rxchf/rxcuhf_omg3.f:499-600
subroutine RXCUHF_FBE_OMG3(Nchunks, nebf, npbf, ng3, DAalpE, DAbetE, DAtotE, DBE, DP, GM3_1ICR, GM3_2ICR, FBE, &
                           E_BE_OMG3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3
  real(c_double) :: DAalpE(nebf, nebf), DAbetE(nebf, nebf), DAtotE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf)
  real(c_double) :: GM3_1ICR(ng3), GM3_2ICR(ng3), FBE(nebf, nebf), E_BE_OMG3
  call neorx_check(neorx_rxcuhf_fbe_omg3(nebf, npbf, DAalpE, DAbetE, DAtotE, DBE, DP, GM3_1ICR, GM3_2ICR, FBE, &
                   E_BE_OMG3), 'RXCUHF_FBE_OMG3')
end subroutine

! RXCHFne_FAE_OMG2, source/rxchfne/rxchfne_omg2.f:2-116
subroutine RXCHFne_FAE_OMG2(Nchunks, nebf, npbf, ng2, DAE, DBE, DP, GM2ICR, FAE, E_AE_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM2ICR(ng2), FAE(nebf, nebf), E_AE_OMG2
  call neorx_check(neorx_rxchfne_fae_omg2(nebf, npbf, DAE, DBE, DP, GM2ICR, FAE, E_AE_OMG2), 'RXCHFne_FAE_OMG2')
end subroutine

! RXCHFne_FP_OMG2, source/rxchfne/rxchfne_omg2.f:212-314
subroutine RXCHFne_FP_OMG2(Nchunks, nebf, npbf, ng2, NAE, NBE, DAE, DBE, DP, GM2ICR, FP, E_P_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2, NAE, NBE
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM2ICR(ng2), FP(npbf, npbf), E_P_OMG2
  call neorx_check(neorx_rxchfne_fp_omg2(nebf, npbf, DAE, DBE, DP, GM2ICR, FP, E_P_OMG2), 'RXCHFne_FP_OMG2')
end subroutine

! RXCHFne_FBE_OMG2, source/rxchfne/rxchfne_omg2.f:409-511
subroutine RXCHFne_FBE_OMG2(Nchunks, nebf, npbf, ng2, NAE, NBE, DAE, DBE, DP, GM2ICR, FBE, E_BE_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2, NAE, NBE
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM2ICR(ng2), FBE(nebf, nebf), E_BE_OMG2
  call neorx_check(neorx_rxchfne_fbe_omg2(nebf, npbf, DAE, DBE, DP, GM2ICR, FBE, E_BE_OMG2), 'RXCHFne_FBE_OMG2')
end subroutine

! RXCUHFne_FE_OMG2, source/rxchfne/rxcuhfne_omg2.f:2-129 (DAbetE is not used by it)
subroutine RXCUHFne_FE_OMG2(Nchunks, nebf, npbf, ng2, DAalpE, DAbetE, DAtotE, DBE, DP, GM2ICR, GM2exICR, &
                            FAalpE, FAbetE, E_AE_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2
  real(c_double) :: DAalpE(nebf, nebf), DAbetE(nebf, nebf), DAtotE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf)
  real(c_double) :: GM2ICR(ng2), GM2exICR(ng2), FAalpE(nebf, nebf), FAbetE(nebf, nebf), E_AE_OMG2
  call neorx_check(neorx_rxcuhfne_fe_omg2(nebf, npbf, DAalpE, DAtotE, DBE, DP, GM2ICR, GM2exICR, FAalpE, FAbetE, &
                   E_AE_OMG2), 'RXCUHFne_FE_OMG2')
end subroutine

! RXCUHFne_FP_OMG2, source/rxchfne/rxcuhfne_omg2.f:239-345
subroutine RXCUHFne_FP_OMG2(Nchunks, nebf, npbf, ng2, NAalpE, NBE, DAalpE, DAtotE, DBE, DP, GM2ICR, GM2exICR, &
                            FP, E_P_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2, NAalpE, NBE
  real(c_double) :: DAalpE(nebf, nebf), DAtotE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf)
  real(c_double) :: GM2ICR(ng2), GM2exICR(ng2), FP(npbf, npbf), E_P_OMG2
  call neorx_check(neorx_rxcuhfne_fp_omg2(nebf, npbf, DAalpE, DAtotE, DBE, DP, GM2ICR, GM2exICR, FP, E_P_OMG2), &
                   'RXCUHFne_FP_OMG2')
end subroutine

! RXCUHFne_FBE_OMG2, source/rxchfne/rxcuhfne_omg2.f:445-551
subroutine RXCUHFne_FBE_OMG2(Nchunks, nebf, npbf, ng2, NAalpE, NBE, DAalpE, DAtotE, DBE, DP, GM2ICR, GM2exICR, &
                             FBE, E_BE_OMG2)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2, NAalpE, NBE
  real(c_double) :: DAalpE(nebf, nebf), DAtotE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf)
  real(c_double) :: GM2ICR(ng2), GM2exICR(ng2), FBE(nebf, nebf), E_BE_OMG2
  call neorx_check(neorx_rxcuhfne_fbe_omg2(nebf, npbf, DAalpE, DAtotE, DBE, DP, GM2ICR, GM2exICR, FBE, E_BE_OMG2), &
                   'RXCUHFne_FBE_OMG2')
end subroutine

! RXCHFne_FAE_GAMee, source/rxchfne/rxchfne_hfterms.f:2-60
subroutine RXCHFne_FAE_GAMee(nebf, ngee, GAM_ee, DAE, FAE, E_AE_GAMee)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: nebf, ngee
  real(c_double) :: GAM_ee(ngee), DAE(nebf, nebf), FAE(nebf, nebf), E_AE_GAMee
  call neorx_check(neorx_rxchfne_fae_gamee(nebf, GAM_ee, DAE, FAE, E_AE_GAMee), 'RXCHFne_FAE_GAMee')
end subroutine

! ---- RXCHFmult (source/rxchfmult/) ------------------------------------------

! RXCHFmult_GAM2, source/rxchfmult/rxchfmult_gam2.f:2-265
subroutine RXCHFmult_GAM2(Nchunks, nebf, npebf, nebfBE, npebfBE, npbf, dimINT2, dimXCHF2, nat, ngtg1, pmass, cat, &
                          zan, bcoef1, gamma1, KPESTR, KPEEND, AMPEB2C, AGEBFCC, ELCEX, ELCAM, ELCBFC, KPESTR_be, &
                          KPEEND_be, AMPEB2C_be, AGEBFCC_be, ELCEX_be, ELCAM_be, ELCBFC_be, AGNBFCC, NUCEX, NUCAM, &
                          NUCBFC, INT_GAM2, XCHF_GAM2, XCHF_GAM2s)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, nebfBE, npebfBE, npbf, dimINT2, dimXCHF2, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3)
  integer(c_int64_t) :: KPESTR_be(nebfBE), KPEEND_be(nebfBE), AMPEB2C_be(npebfBE), ELCAM_be(npebfBE, 3)
  integer(c_int64_t) :: NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), ELCEX(npebf), ELCBFC(npebf, 3)
  real(c_double) :: AGEBFCC_be(npebfBE), ELCEX_be(npebfBE), ELCBFC_be(npebfBE, 3)
  real(c_double) :: AGNBFCC(npbf), NUCEX(npbf), NUCBFC(npbf, 3)
  real(c_double) :: INT_GAM2(dimINT2), XCHF_GAM2(dimXCHF2), XCHF_GAM2s(dimXCHF2)
  integer(c_int64_t) :: h
  call neorx_check(neorx_mult_system_create(nebf, npebf, nebfBE, npebfBE, npbf, nat, ngtg1, pmass, cat, zan, &
                   bcoef1, gamma1, KPESTR, KPEEND, AMPEB2C, AGEBFCC, ELCEX, ELCAM, ELCBFC, KPESTR_be, KPEEND_be, &
                   AMPEB2C_be, AGEBFCC_be, ELCEX_be, ELCAM_be, ELCBFC_be, AGNBFCC, NUCEX, NUCAM, NUCBFC, h), &
                   'RXCHFmult_GAM2')
  call neorx_check(neorx_rxchfmult_gam2(h, INT_GAM2, XCHF_GAM2, XCHF_GAM2s), 'RXCHFmult_GAM2')
  call neorx_check(neorx_system_destroy(h), 'RXCHFmult_GAM2')
end subroutine

! RXCHFmult_GAM1_MD, source/rxchfmult/rxchfmult_gam1.f:2-398: as GAM1_OMP_MD
! (GAM_1.ufm / GAM_1s.ufm), without the one-geminal terms
subroutine RXCHFmult_GAM1_MD(nebf, npebf, npbf, ng1, ng1prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                             AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: nebf, npebf, npbf, ng1, ng1prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  integer(c_int64_t) :: h, kpestr(nebf), kpeend(nebf)
  real(c_double), allocatable :: g(:), gs(:)
  call neorx_kpe(nebf, npebf, AMPEB2C, kpestr, kpeend)
  allocate(g(ng1), gs(ng1))
  call neorx_check(neorx_system_create(nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, kpestr, &
                   kpeend, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, h), &
                   'RXCHFmult_GAM1_MD')
  call neorx_check(neorx_rxchfmult_gam1(h, g, gs), 'RXCHFmult_GAM1_MD')
  call neorx_check(neorx_system_destroy(h), 'RXCHFmult_GAM1_MD')
  call ufm_write(801, 'GAM_1.ufm', ng1, g)
  call ufm_write(802, 'GAM_1s.ufm', ng1, gs)
  deallocate(g, gs)
end subroutine

! RXCHFmult_GAM3_XCHF, source/rxchfmult/rxchfmult_gam3-xchf.f:2-268 (the
! driver passes the special-electron basis, rxchfmult_driver.f:548)
subroutine RXCHFmult_GAM3_XCHF(Nchunks, nebf, npebf, npbf, ng3, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                               KPESTR, KPEEND, AMPEB2C, AGEBFCC, ELCEX, ELCAM, ELCBFC, AGNBFCC, NUCEX, NUCAM, &
                               NUCBFC, XCHF_GAM3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, npbf, ng3, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  real(c_double) :: XCHF_GAM3(ng3)
  integer(c_int64_t) :: h
  call neorx_check(neorx_system_create(nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, KPESTR, &
                   KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, h), &
                   'RXCHFmult_GAM3_XCHF')
  call neorx_check(neorx_rxchfmult_gam3_xchf(h, XCHF_GAM3), 'RXCHFmult_GAM3_XCHF')
  call neorx_check(neorx_system_destroy(h), 'RXCHFmult_GAM3_XCHF')
end subroutine

! RXCHFmult_GAM4_XCHF, source/rxchfmult/rxchfmult_gam4-xchf.f:2-157: GAM4_ICR
! over the special basis with the leading nebfBE^4 block of GAM_ee.ufm
subroutine RXCHFmult_GAM4_XCHF(Nchunks, nebf, nebfBE, npbf, ngee, ng2, ng4, GAM_2s, GAM_4)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, nebfBE, npbf, ngee, ng2, ng4
  real(c_double) :: GAM_2s(ng2), GAM_4(ng4)
  real(c_double), allocatable :: gee(:,:,:,:), geebe(:,:,:,:)
  allocate(gee(nebf, nebf, nebf, nebf), geebe(nebfBE, nebfBE, nebfBE, nebfBE))
  ! record j2 + nebf (i2-1 + nebf (j1-1 + nebf (i1-1))) holds gee(j2, i2, j1, i1)
  call ufm_read(812, 'GAM_ee.ufm', nebf ** 4, gee)
  geebe = gee(1:nebfBE, 1:nebfBE, 1:nebfBE, 1:nebfBE)
  call neorx_check(neorx_gam4_icr(nebfBE, npbf, GAM_2s, geebe, GAM_4), 'RXCHFmult_GAM4_XCHF')
  deallocate(gee, geebe)
end subroutine

! RXCHFmult_GAM3_INT, source/rxchfmult/rxchfmult_gam3-int.f:2-254
subroutine RXCHFmult_GAM3_INT(Nchunks, nebf, npebf, nebfBE, npebfBE, npbf, ng3, nat, ngtg1, pmass, cat, zan, &
                              bcoef1, gamma1, KPESTR, KPEEND, AMPEB2C, AGEBFCC, ELCEX, ELCAM, ELCBFC, KPESTR_be, &
                              KPEEND_be, AMPEB2C_be, AGEBFCC_be, ELCEX_be, ELCAM_be, ELCBFC_be, AGNBFCC, NUCEX, &
                              NUCAM, NUCBFC, INT_GAM3)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, nebfBE, npebfBE, npbf, ng3, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3)
  integer(c_int64_t) :: KPESTR_be(nebfBE), KPEEND_be(nebfBE), AMPEB2C_be(npebfBE), ELCAM_be(npebfBE, 3)
  integer(c_int64_t) :: NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), ELCEX(npebf), ELCBFC(npebf, 3)
  real(c_double) :: AGEBFCC_be(npebfBE), ELCEX_be(npebfBE), ELCBFC_be(npebfBE, 3)
  real(c_double) :: AGNBFCC(npbf), NUCEX(npbf), NUCBFC(npbf, 3)
  real(c_double) :: INT_GAM3(ng3)
  integer(c_int64_t) :: h
  call neorx_check(neorx_mult_system_create(nebf, npebf, nebfBE, npebfBE, npbf, nat, ngtg1, pmass, cat, zan, &
                   bcoef1, gamma1, KPESTR, KPEEND, AMPEB2C, AGEBFCC, ELCEX, ELCAM, ELCBFC, KPESTR_be, KPEEND_be, &
                   AMPEB2C_be, AGEBFCC_be, ELCEX_be, ELCAM_be, ELCBFC_be, AGNBFCC, NUCEX, NUCAM, NUCBFC, h), &
                   'RXCHFmult_GAM3_INT')
  call neorx_check(neorx_rxchfmult_gam3_int(h, INT_GAM3), 'RXCHFmult_GAM3_INT')
  call neorx_check(neorx_system_destroy(h), 'RXCHFmult_GAM3_INT')
end subroutine

! RXCHFmult_GAM4_INT, source/rxchfmult/rxchfmult_gam4-int.f:2-153: reads
! GAM_ee.ufm (nebf^4, regular basis) like read_GAM_ee
subroutine RXCHFmult_GAM4_INT(Nchunks, nebf, nebfBE, npbf, ngee, ng2, ng4, GAM_2s, GAM_4)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, nebfBE, npbf, ngee, ng2, ng4
  real(c_double) :: GAM_2s(ng2), GAM_4(ng4)
  real(c_double), allocatable :: gee(:)
  allocate(gee(ngee))
  call ufm_read(811, 'GAM_ee.ufm', ngee, gee)
  call neorx_check(neorx_rxchfmult_gam4_int(nebf, nebfBE, npbf, GAM_2s, gee, GAM_4), 'RXCHFmult_GAM4_INT')
  deallocate(gee)
end subroutine

! RXCHFmult_FAE_OMG2, source/rxchfmult/rxchfmult_omg2.f:2-90: FAE accumulates, E overwritten
subroutine RXCHFmult_FAE_OMG2(Nchunks, nebf, nebfBE, npbf, ng, DAE, DBE, DP, GM, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, nebfBE, npbf, ng
  real(c_double) :: DAE(nebf, nebf), DBE(nebfBE, nebfBE), DP(npbf, npbf), GM(ng), F(nebf, nebf), E
  call neorx_check(neorx_rxchfmult_fae_omg2(nebf, nebfBE, npbf, DAE, DBE, DP, GM, F, E), &
                   'RXCHFmult_FAE_OMG2')
end subroutine

! RXCHFmult_FP_OMG2, source/rxchfmult/rxchfmult_omg2.f:175-263: FP accumulates, E overwritten
subroutine RXCHFmult_FP_OMG2(Nchunks, nebf, nebfBE, npbf, ng, DAE, DBE, DP, GM, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, nebfBE, npbf, ng
  real(c_double) :: DAE(nebf, nebf), DBE(nebfBE, nebfBE), DP(npbf, npbf), GM(ng), F(npbf, npbf), E
  call neorx_check(neorx_rxchfmult_fp_omg2(nebf, nebfBE, npbf, DAE, DBE, DP, GM, F, E), &
                   'RXCHFmult_FP_OMG2')
end subroutine

! RXCHFmult_FBE_OMG2, source/rxchfmult/rxchfmult_omg2.f:348-436: FBE accumulates, E overwritten
subroutine RXCHFmult_FBE_OMG2(Nchunks, nebf, nebfBE, npbf, ng, DAE, DBE, DP, GM, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, nebfBE, npbf, ng
  real(c_double) :: DAE(nebf, nebf), DBE(nebfBE, nebfBE), DP(npbf, npbf), GM(ng), F(nebfBE, nebfBE), E
  call neorx_check(neorx_rxchfmult_fbe_omg2(nebf, nebfBE, npbf, DAE, DBE, DP, GM, F, E), &
                   'RXCHFmult_FBE_OMG2')
end subroutine

! RXCHFmult_FAE_OMG3, source/rxchfmult/rxchfmult_omg3.f:2-96: FAE accumulates, E overwritten
subroutine RXCHFmult_FAE_OMG3(Nchunks, nebf, nebfBE, npbf, ng, DAE, DBE, DP, GM, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, nebfBE, npbf, ng
  real(c_double) :: DAE(nebf, nebf), DBE(nebfBE, nebfBE), DP(npbf, npbf), GM(ng), F(nebf, nebf), E
  call neorx_check(neorx_rxchfmult_fae_omg3(nebf, nebfBE, npbf, DAE, DBE, DP, GM, F, E), &
                   'RXCHFmult_FAE_OMG3')
end subroutine

! RXCHFmult_FP_OMG3, source/rxchfmult/rxchfmult_omg3.f:201-295: FP accumulates, E overwritten
subroutine RXCHFmult_FP_OMG3(Nchunks, nebf, nebfBE, npbf, ng, DAE, DBE, DP, GM, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, nebfBE, npbf, ng
  real(c_double) :: DAE(nebf, nebf), DBE(nebfBE, nebfBE), DP(npbf, npbf), GM(ng), F(npbf, npbf), E
  call neorx_check(neorx_rxchfmult_fp_omg3(nebf, nebfBE, npbf, DAE, DBE, DP, GM, F, E), &
                   'RXCHFmult_FP_OMG3')
end subroutine

! RXCHFmult_FBE_OMG3, source/rxchfmult/rxchfmult_omg3.f:400-494: FBE accumulates, E overwritten
subroutine RXCHFmult_FBE_OMG3(Nchunks, nebf, nebfBE, npbf, ng, DAE, DBE, DP, GM, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, nebfBE, npbf, ng
  real(c_double) :: DAE(nebf, nebf), DBE(nebfBE, nebfBE), DP(npbf, npbf), GM(ng), F(nebfBE, nebfBE), E
  call neorx_check(neorx_rxchfmult_fbe_omg3(nebf, nebfBE, npbf, DAE, DBE, DP, GM, F, E), &
                   'RXCHFmult_FBE_OMG3')
end subroutine

! RXCHFmult_FAE_OMG4, source/rxchfmult/rxchfmult_omg4.f:2-108: FAE accumulates, E overwritten
subroutine RXCHFmult_FAE_OMG4(Nchunks, nebf, nebfBE, npbf, ng, DAE, DBE, DP, GM, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, nebfBE, npbf, ng
  real(c_double) :: DAE(nebf, nebf), DBE(nebfBE, nebfBE), DP(npbf, npbf), GM(ng), F(nebf, nebf), E
  call neorx_check(neorx_rxchfmult_fae_omg4(nebf, nebfBE, npbf, DAE, DBE, DP, GM, F, E), &
                   'RXCHFmult_FAE_OMG4')
end subroutine

! RXCHFmult_FP_OMG4, source/rxchfmult/rxchfmult_omg4.f:207-313: FP accumulates, E overwritten
subroutine RXCHFmult_FP_OMG4(Nchunks, nebf, nebfBE, npbf, ng, DAE, DBE, DP, GM, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, nebfBE, npbf, ng
  real(c_double) :: DAE(nebf, nebf), DBE(nebfBE, nebfBE), DP(npbf, npbf), GM(ng), F(npbf, npbf), E
  call neorx_check(neorx_rxchfmult_fp_omg4(nebf, nebfBE, npbf, DAE, DBE, DP, GM, F, E), &
                   'RXCHFmult_FP_OMG4')
end subroutine

! RXCHFmult_FBE_OMG4, source/rxchfmult/rxchfmult_omg4.f:412-518: FBE accumulates, E overwritten
subroutine RXCHFmult_FBE_OMG4(Nchunks, nebf, nebfBE, npbf, ng, DAE, DBE, DP, GM, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, nebfBE, npbf, ng
  real(c_double) :: DAE(nebf, nebf), DBE(nebfBE, nebfBE), DP(npbf, npbf), GM(ng), F(nebfBE, nebfBE), E
  call neorx_check(neorx_rxchfmult_fbe_omg4(nebf, nebfBE, npbf, DAE, DBE, DP, GM, F, E), &
                   'RXCHFmult_FBE_OMG4')
end subroutine

! RXCHFmult_GAM2_IC1ex, source/rxchfmult/rxchfmult_gam2ex.f:2-229 (one basis:
! the handle takes the regular basis for both electrons)
subroutine RXCHFmult_GAM2_IC1ex(Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1, pmass, cat, zan, bcoef1, &
                                gamma1, KPESTR, KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, &
                                ELCBFC, NUCBFC, GM2_1ICR, GM2_2ICR, GM2_3ICR, GM2sICR)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  real(c_double) :: GM2_1ICR(ng2), GM2_2ICR(ng2), GM2_3ICR(ng2), GM2sICR(ng2)
  integer(c_int64_t) :: h
  call neorx_check(neorx_mult_system_create(nebf, npebf, nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, &
                   gamma1, KPESTR, KPEEND, AMPEB2C, AGEBFCC, ELCEX, ELCAM, ELCBFC, KPESTR, KPEEND, AMPEB2C, &
                   AGEBFCC, ELCEX, ELCAM, ELCBFC, AGNBFCC, NUCEX, NUCAM, NUCBFC, h), 'RXCHFmult_GAM2_IC1ex')
  call neorx_check(neorx_rxchfmult_gam2_ic1ex(h, GM2_1ICR, GM2_2ICR, GM2_3ICR, GM2sICR), 'RXCHFmult_GAM2_IC1ex')
  call neorx_check(neorx_system_destroy(h), 'RXCHFmult_GAM2_IC1ex')
end subroutine

! RXCHFmult_GAM3_IC1ex, source/rxchfmult/rxchfmult_gam3ex.f:2-316
subroutine RXCHFmult_GAM3_IC1ex(Nchunks, nebf, npebf, npbf, ng3, ng3prm, nat, ngtg1, pmass, cat, zan, bcoef1, &
                                gamma1, KPESTR, KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, &
                                ELCBFC, NUCBFC, GM3_2ICR, GM3_3ICR, GM3_4ICR)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npebf, npbf, ng3, ng3prm, nat, ngtg1
  real(c_double) :: pmass, cat(3, nat), zan(nat), bcoef1(ngtg1), gamma1(ngtg1)
  integer(c_int64_t) :: KPESTR(nebf), KPEEND(nebf), AMPEB2C(npebf), ELCAM(npebf, 3), NUCAM(npbf, 3)
  real(c_double) :: AGEBFCC(npebf), AGNBFCC(npbf), ELCEX(npebf), NUCEX(npbf), ELCBFC(npebf, 3), NUCBFC(npbf, 3)
  real(c_double) :: GM3_2ICR(ng3), GM3_3ICR(ng3), GM3_4ICR(ng3)
  integer(c_int64_t) :: h
  call neorx_check(neorx_mult_system_create(nebf, npebf, nebf, npebf, npbf, nat, ngtg1, pmass, cat, zan, bcoef1, &
                   gamma1, KPESTR, KPEEND, AMPEB2C, AGEBFCC, ELCEX, ELCAM, ELCBFC, KPESTR, KPEEND, AMPEB2C, &
                   AGEBFCC, ELCEX, ELCAM, ELCBFC, AGNBFCC, NUCEX, NUCAM, NUCBFC, h), 'RXCHFmult_GAM3_IC1ex')
  call neorx_check(neorx_rxchfmult_gam3_ic1ex(h, GM3_2ICR, GM3_3ICR, GM3_4ICR), 'RXCHFmult_GAM3_IC1ex')
  call neorx_check(neorx_system_destroy(h), 'RXCHFmult_GAM3_IC1ex')
end subroutine

! RXCHFmult_FAE_OMG2ex, source/rxchfmult/rxchfmult_omg2ex.f:2-93
subroutine RXCHFmult_FAE_OMG2ex(Nchunks, nebf, npbf, ng2, DAE, DBE, DP, GM2ICR, GM2exICR, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM2ICR(ng2), GM2exICR(ng2), F(nebf, nebf), E
  call neorx_check(neorx_rxchfmult_fae_omg2ex(nebf, npbf, DAE, DBE, DP, GM2ICR, GM2exICR, F, E), &
                   'RXCHFmult_FAE_OMG2ex')
end subroutine

! RXCHFmult_FAE_OMG3ex, source/rxchfmult/rxchfmult_omg3ex.f:2-100
subroutine RXCHFmult_FAE_OMG3ex(Nchunks, nebf, npbf, ng3, DAE, DBE, DP, GM3ICR, GM3ex1ICR, GM3ex2ICR, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM3ICR(ng3), GM3ex1ICR(ng3), GM3ex2ICR(ng3)
  real(c_double) :: F(nebf, nebf), E
  call neorx_check(neorx_rxchfmult_fae_omg3ex(nebf, npbf, DAE, DBE, DP, GM3ICR, GM3ex1ICR, GM3ex2ICR, F, E), &
                   'RXCHFmult_FAE_OMG3ex')
end subroutine

! RXCHFmult_FP_OMG2ex, source/rxchfmult/rxchfmult_omg2ex.f:181-272
subroutine RXCHFmult_FP_OMG2ex(Nchunks, nebf, npbf, ng2, DAE, DBE, DP, GM2ICR, GM2exICR, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM2ICR(ng2), GM2exICR(ng2), F(npbf, npbf), E
  call neorx_check(neorx_rxchfmult_fp_omg2ex(nebf, npbf, DAE, DBE, DP, GM2ICR, GM2exICR, F, E), &
                   'RXCHFmult_FP_OMG2ex')
end subroutine

! RXCHFmult_FP_OMG3ex, source/rxchfmult/rxchfmult_omg3ex.f:217-315
subroutine RXCHFmult_FP_OMG3ex(Nchunks, nebf, npbf, ng3, DAE, DBE, DP, GM3ICR, GM3ex1ICR, GM3ex2ICR, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM3ICR(ng3), GM3ex1ICR(ng3), GM3ex2ICR(ng3)
  real(c_double) :: F(npbf, npbf), E
  call neorx_check(neorx_rxchfmult_fp_omg3ex(nebf, npbf, DAE, DBE, DP, GM3ICR, GM3ex1ICR, GM3ex2ICR, F, E), &
                   'RXCHFmult_FP_OMG3ex')
end subroutine

! RXCHFmult_FBE_OMG2ex, source/rxchfmult/rxchfmult_omg2ex.f:360-451
subroutine RXCHFmult_FBE_OMG2ex(Nchunks, nebf, npbf, ng2, DAE, DBE, DP, GM2ICR, GM2exICR, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng2
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM2ICR(ng2), GM2exICR(ng2), F(nebf, nebf), E
  call neorx_check(neorx_rxchfmult_fbe_omg2ex(nebf, npbf, DAE, DBE, DP, GM2ICR, GM2exICR, F, E), &
                   'RXCHFmult_FBE_OMG2ex')
end subroutine

! RXCHFmult_FBE_OMG3ex, source/rxchfmult/rxchfmult_omg3ex.f:432-530
subroutine RXCHFmult_FBE_OMG3ex(Nchunks, nebf, npbf, ng3, DAE, DBE, DP, GM3ICR, GM3ex1ICR, GM3ex2ICR, F, E)
  use neorx_capi
  implicit none
  integer(c_int64_t) :: Nchunks, nebf, npbf, ng3
  real(c_double) :: DAE(nebf, nebf), DBE(nebf, nebf), DP(npbf, npbf), GM3ICR(ng3), GM3ex1ICR(ng3), GM3ex2ICR(ng3)
  real(c_double) :: F(nebf, nebf), E
  call neorx_check(neorx_rxchfmult_fbe_omg3ex(nebf, npbf, DAE, DBE, DP, GM3ICR, GM3ex1ICR, GM3ex2ICR, F, E), &
                   'RXCHFmult_FBE_OMG3ex')
end subroutine

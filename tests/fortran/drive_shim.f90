! drive_shim.f90 -- test driver that calls the drop-in subroutines exactly as
! the reference's integral phase and Fock build do (xcneo_HYBRID.f:761-1020,
! xchf1_fock.f:246-299), linked against libneorx_fshim.so.
! Input and output are raw stream files written/read by the GPU tests.
program drive_shim
  use iso_c_binding
  use neorx_capi, only: ufm_read, ufm_write
  implicit none
  interface
    subroutine GAM2_IC1(Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1, pmass, cat, zan, bcoef1, gamma1, &
                        KPESTR, KPEEND, AMPEB2C, AGEBFCC, AGNBFCC, ELCEX, NUCEX, ELCAM, NUCAM, ELCBFC, NUCBFC, &
                        GM2ICR, GM2sICR)
      import :: c_int64_t, c_double
      integer(c_int64_t) :: Nchunks, nebf, npebf, npbf, ng2, ng2prm, nat, ngtg1
      real(c_double) :: pmass, cat(*), zan(*), bcoef1(*), gamma1(*)
      integer(c_int64_t) :: KPESTR(*), KPEEND(*), AMPEB2C(*), ELCAM(*), NUCAM(*)
      real(c_double) :: AGEBFCC(*), AGNBFCC(*), ELCEX(*), NUCEX(*), ELCBFC(*), NUCBFC(*), GM2ICR(*), GM2sICR(*)
    end subroutine
  end interface
  integer(c_int64_t) :: nebf, npebf, npbf, nat, ngtg, ng2, ng3, ng4, ngee, ng1, one, zero, ia, nae, nbe
  real(c_double) :: pmass, e2, s2, e3, e4
  real(c_double), allocatable :: cat(:), zan(:), bcoef(:), gamma(:), agebfcc(:), agnbfcc(:), elcex(:), nucex(:)
  real(c_double), allocatable :: elcbfc(:), nucbfc(:), de(:,:), dp(:,:), dae(:,:), dbe(:,:)
  integer(c_int64_t), allocatable :: kpestr(:), kpeend(:), ampeb2c(:), elcam(:), nucam(:)
  real(c_double), allocatable :: gm2(:), gm2s(:), gm3(:), gm4(:), fe(:,:), se(:,:), fp(:,:), sp(:,:), g1(:)
  real(c_double), allocatable :: gx1(:), gx2(:), gx3(:), fb(:,:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nebf, npebf, npbf, nat, ngtg, pmass
  allocate(cat(3*nat), zan(nat), bcoef(ngtg), gamma(ngtg), kpestr(nebf), kpeend(nebf), ampeb2c(npebf))
  allocate(agebfcc(npebf), agnbfcc(npbf), elcex(npebf), nucex(npbf), elcam(3*npebf), nucam(3*npbf))
  allocate(elcbfc(3*npebf), nucbfc(3*npbf), de(nebf, nebf), dp(npbf, npbf))
  read(10) cat, zan, bcoef, gamma, kpestr, kpeend, ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, &
           elcbfc, nucbfc, de, dp
  close(10)
  one = 1
  zero = 0
  ng1 = nebf**2 * npbf**2
  ng2 = nebf**4 * npbf**2
  ng3 = nebf**6 * npbf**2
  ng4 = nebf**8 * npbf**2
  ngee = nebf**4
  allocate(gm2(ng2), gm2s(ng2), gm3(ng3), gm4(ng4), fe(nebf, nebf), se(nebf, nebf), fp(npbf, npbf), sp(npbf, npbf))

  call GAM1_OMP_MD(nebf, npebf, npbf, ng1, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, &
                   ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc)
  allocate(g1(ng1))
  call read_ufm('GAM_1.ufm', ng1, g1)
  call calc_GAM_ee(nebf, npebf, ngee, ampeb2c, agebfcc, elcex, elcam, elcbfc)
  call GAM2_IC1(one, nebf, npebf, npbf, ng2, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, kpestr, kpeend, &
                ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc, gm2, gm2s)
  call GAM3_IC1(one, nebf, npebf, npbf, ng3, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, kpestr, kpeend, &
                ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc, gm3)
  call GAM4_ICR(one, nebf, npbf, ngee, ng2, ng4, gm2s, gm4)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  fe = 0.0d0
  fp = 0.0d0
  call EG2IC1(one, nebf, npbf, ng2, de, dp, gm2, gm2s, fe, se, fp, sp, e2, s2)
  write(11) gm2, gm2s, fe, se, fp, sp, e2, s2
  fe = 0.0d0
  fp = 0.0d0
  call EG3IC1(one, nebf, npbf, ng3, de, dp, gm3, fe, fp, e3)
  write(11) gm3, fe, fp, e3
  fe = 0.0d0
  fp = 0.0d0
  call EG4ICR(one, nebf, npbf, ng4, de, dp, gm4, fe, fp, e4)
  write(11) gm4, fe, fp, e4
  ! direct-SCF routines (xchf1_fock.f:247, :273, :290); GAM4_DSCF reads
  ! GAM_2s.ufm, which the reference's integral phase writes
  ! (the driver's own .ufm I/O follows NEORX_UFM_RECL as the library does,
  ! like a driver and library built with the same compiler)
  call ufm_write(804, 'GAM_2s.ufm', ng2, gm2s)
  fe = 0.0d0
  fp = 0.0d0
  se = 0.0d0
  sp = 0.0d0
  call GAM2_DSCF(one, nebf, npebf, npbf, ng2, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, kpestr, kpeend, &
                 ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc, de, dp, fe, fp, se, sp, e2, s2)
  write(11) fe, fp, se, sp, e2, s2
  fe = 0.0d0
  fp = 0.0d0
  call GAM3_DSCF(one, nebf, npebf, npbf, ng3, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, kpestr, kpeend, &
                 ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc, de, dp, fe, fp, e3)
  write(11) fe, fp, e3
  fe = 0.0d0
  fp = 0.0d0
  call GAM4_DSCF(one, nebf, npbf, ngee, ng2, ng4, de, dp, fe, fp, e4)
  write(11) fe, fp, e4
  ! XCUHF Fock pieces (xcuhf_fock.f:896-916, :1014-1034) with NAE=2, NBE=1,
  ! DAE = 2/3 DE, DBE = 1/3 DE; alpha and beta FAE share one call shape
  nae = 2
  nbe = 1
  allocate(dae(nebf, nebf), dbe(nebf, nebf))
  dae = de * (2.0d0 / 3.0d0)
  dbe = de * (1.0d0 / 3.0d0)
  call OMG4_ICR(one, nebf, npbf, ngee, ng2, ng4, gm2s, gm4)
  write(11) gm4
  fe = 0.0d0
  call FAE_OMG2(one, nae, nbe, nebf, npbf, ng2, dae, dbe, de, dp, gm2, gm2s, fe, se, e2, s2)
  write(11) fe, se, e2, s2
  fp = 0.0d0
  call UFP_OMG2(one, nebf, npbf, ng2, nae, nbe, dae, dbe, de, dp, gm2, gm2s, fp, sp, e2, s2)
  write(11) fp, sp, e2, s2
  fe = 0.0d0
  call FAE_OMG3(one, nebf, npbf, ng3, nae, nbe, dae, dbe, dp, gm3, fe, e3)
  write(11) fe, e3
  fp = 0.0d0
  call UFP_OMG3(one, nebf, npbf, ng3, nae, nbe, dae, dbe, dp, gm3, fp, e3)
  write(11) fp, e3
  fe = 0.0d0
  call FAE_OMG4(one, nebf, npbf, ng4, nae, nbe, dae, dbe, dp, gm4, fe, e4)
  write(11) fe, e4
  fp = 0.0d0
  call UFP_OMG4(one, nebf, npbf, ng4, nae, nbe, dae, dbe, dp, gm4, fp, e4)
  write(11) fp, e4
  ! RXCHF-ne (rxchfne_fock.f:391, :446, :501; rxcuhfne_fock.f:376): DBE here
  ! is the special-electron density (dbe), DAE the regular one (dae)
  call RXCHFne_GAM2_IC1(one, nebf, npebf, npbf, ng2, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, kpestr, &
                        kpeend, ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc, gm2, gm2s)
  write(11) gm2, gm2s
  fe = 0.0d0
  call RXCHFne_FAE_OMG2(one, nebf, npbf, ng2, dae, dbe, dp, gm2, fe, e2)
  write(11) fe, e2
  fp = 0.0d0
  call RXCHFne_FP_OMG2(one, nebf, npbf, ng2, nae, nbe, dae, dbe, dp, gm2, fp, e2)
  write(11) fp, e2
  fe = 0.0d0
  call RXCHFne_FBE_OMG2(one, nebf, npbf, ng2, nae, nbe, dae, dbe, dp, gm2, fe, e2)
  write(11) fe, e2
  fe = 0.0d0
  se = 0.0d0
  call RXCUHFne_FE_OMG2(one, nebf, npbf, ng2, dbe, dae, de, dbe, dp, gm2, gm2s, fe, se, e2)
  write(11) fe, se, e2
  ! RXCHFmult (rxchfmult_driver.f:398-633, rxchfmult_fock-int.f:366-577) with
  ! the special-electron basis equal to the regular one (nebfBE = nebf):
  ! INT_GAM2 in gm2, XCHF_GAM2 in gm2s, XCHF_GAM2s in gm3(1:ng2)
  call RXCHFmult_GAM2(one, nebf, npebf, nebf, npebf, npbf, ng2, ng2, nat, ngtg, pmass, cat, zan, bcoef, gamma, &
                      kpestr, kpeend, ampeb2c, agebfcc, elcex, elcam, elcbfc, kpestr, kpeend, ampeb2c, agebfcc, &
                      elcex, elcam, elcbfc, agnbfcc, nucex, nucam, nucbfc, gm2, gm2s, gm3)
  write(11) gm2, gm2s, gm3(1:ng2)
  gm2s(1:ng2) = gm3(1:ng2)
  call RXCHFmult_GAM4_INT(one, nebf, nebf, npbf, ngee, ng2, ng4, gm2s, gm4)
  call RXCHFmult_GAM3_INT(one, nebf, npebf, nebf, npebf, npbf, ng3, nat, ngtg, pmass, cat, zan, bcoef, gamma, &
                          kpestr, kpeend, ampeb2c, agebfcc, elcex, elcam, elcbfc, kpestr, kpeend, ampeb2c, agebfcc, &
                          elcex, elcam, elcbfc, agnbfcc, nucex, nucam, nucbfc, gm3)
  write(11) gm3, gm4
  fe = 0.0d0
  call RXCHFmult_FAE_OMG2(one, nebf, nebf, npbf, ng2, dae, dbe, dp, gm2, fe, e2)
  write(11) fe, e2
  fp = 0.0d0
  call RXCHFmult_FP_OMG2(one, nebf, nebf, npbf, ng2, dae, dbe, dp, gm2, fp, e2)
  write(11) fp, e2
  fe = 0.0d0
  call RXCHFmult_FBE_OMG2(one, nebf, nebf, npbf, ng2, dae, dbe, dp, gm2, fe, e2)
  write(11) fe, e2
  fe = 0.0d0
  call RXCHFmult_FAE_OMG3(one, nebf, nebf, npbf, ng3, dae, dbe, dp, gm3, fe, e3)
  write(11) fe, e3
  fp = 0.0d0
  call RXCHFmult_FP_OMG3(one, nebf, nebf, npbf, ng3, dae, dbe, dp, gm3, fp, e3)
  write(11) fp, e3
  fe = 0.0d0
  call RXCHFmult_FBE_OMG3(one, nebf, nebf, npbf, ng3, dae, dbe, dp, gm3, fe, e3)
  write(11) fe, e3
  fe = 0.0d0
  call RXCHFmult_FAE_OMG4(one, nebf, nebf, npbf, ng4, dae, dbe, dp, gm4, fe, e4)
  write(11) fe, e4
  fp = 0.0d0
  call RXCHFmult_FP_OMG4(one, nebf, nebf, npbf, ng4, dae, dbe, dp, gm4, fp, e4)
  write(11) fp, e4
  fe = 0.0d0
  call RXCHFmult_FBE_OMG4(one, nebf, nebf, npbf, ng4, dae, dbe, dp, gm4, fe, e4)
  write(11) fe, e4
  ! RXCHFmult's XCHF-side builders over the special basis (= the regular one
  ! here): GAM1_MD overwrites GAM_1.ufm, so the GAM1_OMP_MD file goes first
  write(11) g1
  call RXCHFmult_GAM1_MD(nebf, npebf, npbf, ng1, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, &
                         ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc)
  call read_ufm('GAM_1.ufm', ng1, g1)
  write(11) g1
  call RXCHFmult_GAM3_XCHF(one, nebf, npebf, npbf, ng3, nat, ngtg, pmass, cat, zan, bcoef, gamma, kpestr, &
                           kpeend, ampeb2c, agebfcc, elcex, elcam, elcbfc, agnbfcc, nucex, nucam, nucbfc, gm3)
  call RXCHFmult_GAM4_XCHF(one, nebf, nebf, npbf, ngee, ng2, ng4, gm2s, gm4)
  write(11) gm3, gm4
  ! RXCHFmult with exchange (LADDEXCH, rxchfmult_driver.f:384, :470;
  ! rxchfmult_fock-int.f:366, :379, :459, :472, :552, :565): one basis
  allocate(gx1(ng2), gx2(ng2))
  call RXCHFmult_GAM2_IC1ex(one, nebf, npebf, npbf, ng2, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, kpestr, &
                            kpeend, ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc, gm2, &
                            gm2s, gx1, gx2)
  write(11) gm2, gm2s, gx1, gx2
  fe = 0.0d0
  call RXCHFmult_FAE_OMG2ex(one, nebf, npbf, ng2, dae, dbe, dp, gm2s, gx1, fe, e2)
  write(11) fe, e2
  fp = 0.0d0
  call RXCHFmult_FP_OMG2ex(one, nebf, npbf, ng2, dae, dbe, dp, gm2s, gx1, fp, e2)
  write(11) fp, e2
  fe = 0.0d0
  call RXCHFmult_FBE_OMG2ex(one, nebf, npbf, ng2, dae, dbe, dp, gm2s, gx1, fe, e2)
  write(11) fe, e2
  deallocate(gm4)
  allocate(gm4(ng3), gx3(ng3))
  call RXCHFmult_GAM3_IC1ex(one, nebf, npebf, npbf, ng3, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, kpestr, &
                            kpeend, ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc, gm3, &
                            gm4, gx3)
  write(11) gm3, gm4, gx3
  fe = 0.0d0
  call RXCHFmult_FAE_OMG3ex(one, nebf, npbf, ng3, dae, dbe, dp, gm3, gm4, gx3, fe, e3)
  write(11) fe, e3
  fp = 0.0d0
  call RXCHFmult_FP_OMG3ex(one, nebf, npbf, ng3, dae, dbe, dp, gm3, gm4, gx3, fp, e3)
  write(11) fp, e3
  fe = 0.0d0
  call RXCHFmult_FBE_OMG3ex(one, nebf, npbf, ng3, dae, dbe, dp, gm3, gm4, gx3, fe, e3)
  write(11) fe, e3
  ! RXCHF with one special electron (rxchf_fock.f:495, :580, :665): DAE the
  ! regular density, DBE the special electron's
  call RXCHF_GAM2_IC1(one, nebf, npebf, npbf, ng2, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, kpestr, &
                      kpeend, ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc, gm2, gx1, gx2)
  write(11) gm2, gx1, gx2
  fe = 0.0d0
  se = 0.0d0
  call RXCHF_FAE_OMG2(one, nebf, npbf, ng2, dae, dbe, dp, gm2, gx1, gx2, fe, se, e2, s2)
  write(11) fe, se, e2, s2
  fp = 0.0d0
  sp = 0.0d0
  call RXCHF_FP_OMG2(one, nebf, npbf, ng2, dae, dbe, dp, gm2, gx1, gx2, fp, sp, e2, s2)
  write(11) fp, sp, e2, s2
  fe = 0.0d0
  se = 0.0d0
  call RXCHF_FBE_OMG2(one, nebf, npbf, ng2, dae, dbe, dp, gm2, gx1, gx2, fe, se, e2, s2)
  write(11) fe, se, e2, s2
  ! RXCHF e-e-e-p and GAM_1 pieces (xcneo_HYBRID.f:756, :907; rxchf_fock.f:501,
  ! :586, :671, :577, :662)
  call RXCHF_GAM3_IC1(one, nebf, npebf, npbf, ng3, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, kpestr, &
                      kpeend, ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc, gm3, gx3)
  write(11) gm3, gx3
  fe = 0.0d0
  call RXCHF_FAE_OMG3(one, nebf, npbf, ng3, dae, dbe, dp, gm3, gx3, fe, e3)
  write(11) fe, e3
  fp = 0.0d0
  call RXCHF_FP_OMG3(one, nebf, npbf, ng3, dae, dbe, dp, gm3, gx3, fp, e3)
  write(11) fp, e3
  fe = 0.0d0
  call RXCHF_FBE_OMG3(one, nebf, npbf, ng3, dae, dbe, dp, gm3, gx3, fe, e3)
  write(11) fe, e3
  call RXCHF_GAM1_OMP_MD(nebf, npebf, npbf, ng1, zero, nat, ngtg, pmass, cat, zan, bcoef, gamma, &
                         ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc)
  fp = 0.0d0
  sp = 0.0d0
  call RXCHF_FP_OMG1(nebf, npbf, ng1, dbe, dp, fp, sp, e2, s2)
  write(11) fp, sp, e2, s2
  fe = 0.0d0
  se = 0.0d0
  call RXCHF_FBE_OMG1(nebf, npbf, ng1, dbe, dp, fe, se, e2, s2)
  write(11) fe, se, e2, s2
  ! RXCUHF (rxcuhf_fock.f:368, :463, :555, :376, :470, :562) over the RXCHF
  ! tensors: alpha = dae, beta = dbe, total = de, special-electron density dbe
  allocate(fb(nebf, nebf))
  fe = 0.0d0
  fb = 0.0d0
  se = 0.0d0
  call RXCUHF_FE_OMG2(one, nebf, npbf, ng2, dae, dbe, de, dbe, dp, gm2, gx1, gx2, fe, fb, se, e2, s2)
  write(11) fe, fb, se, e2, s2
  fp = 0.0d0
  sp = 0.0d0
  call RXCUHF_FP_OMG2(one, nebf, npbf, ng2, nae, nbe, dae, de, dbe, dp, gm2, gx1, gx2, fp, sp, e2, s2)
  write(11) fp, sp, e2, s2
  fe = 0.0d0
  se = 0.0d0
  call RXCUHF_FBE_OMG2(one, nebf, npbf, ng2, nae, nbe, dae, de, dbe, dp, gm2, gx1, gx2, fe, se, e2, s2)
  write(11) fe, se, e2, s2
  fe = 0.0d0
  fb = 0.0d0
  call RXCUHF_FE_OMG3(one, nebf, npbf, ng3, dae, dbe, de, dbe, dp, gm3, gx3, fe, fb, e3)
  write(11) fe, fb, e3
  fp = 0.0d0
  call RXCUHF_FP_OMG3(one, nebf, npbf, ng3, dae, dbe, de, dbe, dp, gm3, gx3, fp, e3)
  write(11) fp, e3
  fe = 0.0d0
  call RXCUHF_FBE_OMG3(one, nebf, npbf, ng3, dae, dbe, de, dbe, dp, gm3, gx3, fe, e3)
  write(11) fe, e3
  close(11)
contains
  subroutine read_ufm(name, n, a)
    character(len=*), intent(in) :: name
    integer(c_int64_t), intent(in) :: n
    real(c_double), intent(out) :: a(n)
    call ufm_read(820, name, n, a)
  end subroutine
end program

C     Test infrastructure only (checker side).
C     Copies the CKLIB / TRANLIB pointer common blocks into plain integer
C     arrays so ref_harness can write them out next to the work arrays.
C     Layouts follow cklib.f CKINIT (cklib.f:3972-4150) and tranlib.f
C     MCINIT (tranlib.f:93-326); both blocks are all-integer apart from
C     the three leading reals of /MCMCMC/.
      SUBROUTINE RHCKPT (IOUT)
      INTEGER IOUT(85), CK(85), I
      COMMON /CKSTRT/ CK
      DO 10 I = 1, 85
         IOUT(I) = CK(I)
   10 CONTINUE
      RETURN
      END

      SUBROUTINE RHMCPT (DOUT, IOUT)
      DOUBLE PRECISION DOUT(3), MCD(3)
      INTEGER IOUT(32), MCI(32), I
      COMMON /MCMCMC/ MCD, MCI
      DO 10 I = 1, 3
         DOUT(I) = MCD(I)
   10 CONTINUE
      DO 20 I = 1, 32
         IOUT(I) = MCI(I)
   20 CONTINUE
      RETURN
      END

      SUBROUTINE RHMACH (DOUT)
      DOUBLE PRECISION DOUT(3), SMALL, BIG, EXPARG
      COMMON /MACH/ SMALL, BIG, EXPARG
      DOUT(1) = SMALL
      DOUT(2) = BIG
      DOUT(3) = EXPARG
      RETURN
      END

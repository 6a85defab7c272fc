/*
 * redi_oracle.c -- CPU restatement of the reference Redi neutral-diffusion
 * step, FP64.  TEST INFRASTRUCTURE ONLY: it checks the results of the HIP
 * product and serves as the CPU baseline; it is never part of the product path.
 *
 * Each function restates one reference routine, cited as file:line under
 * redi/ in the reference tree, in its own code but with the reference's
 * floating-point evaluation order (left-to-right within a parenthesised
 * group, no contraction, no reassociation), so that on baseline x86-64 it
 * rounds exactly like the reference's amdflang DATA=none build.  That
 * bit-identity is checked against the compiled reference (oracle/_ref) and
 * the committed golden fixtures under tests/golden/.
 */
#include "redi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GRAVITY 9.80665   /* constants.F90:19 */
#define H_DIV_EPS 1.0e-20 /* constants.F90:21 */
#define H_NEGLECT 1.0e-30 /* ocean_redi.F90:242 default */

/* Fortran SIGN(a,b): |a| carrying the sign bit of b. */
static inline double f_sign(double a, double b) { return copysign(fabs(a), b); }
static inline double f_min(double a, double b) { return a < b ? a : b; }
static inline double f_max(double a, double b) { return a > b ? a : b; }

/* ---------------------------------------------------------------------- */
/* PPM / PLM column reconstruction                                         */
/* ---------------------------------------------------------------------- */

/* redi_fv_diff, ocean_redi.F90:135-150 */
double oracle_fv_diff(double hkm1, double hk, double hkp1, double skm1,
                      double sk, double skp1) {
    double h_sum = (hkm1 + hkp1) + hk;
    if (h_sum != 0.0) h_sum = 1.0 / h_sum;
    double hm = hkm1 + hk;
    if (hm != 0.0) hm = 1.0 / hm;
    double hp = hkp1 + hk;
    if (hp != 0.0) hp = 1.0 / hp;
    double a = ((2.0 * hkm1 + hk) * hp) * (skp1 - sk);
    double b = ((2.0 * hkp1 + hk) * hm) * (sk - skm1);
    return (hk * h_sum) * (a + b);
}

/* redi_signum, ocean_redi.F90:153-160 (with a = 1) */
static inline double signum1(double x) {
    double s = f_sign(1.0, x);
    if (x == 0.0) s = 0.0;
    return s;
}

/* redi_plm_diff, ocean_redi.F90:165-190 */
void oracle_plm_diff(int nk, const double *h, const double *s, double *diff) {
    for (int k = 1; k < nk - 1; ++k) {
        if ((h[k + 1] + h[k]) * (h[k - 1] + h[k]) > 0.0) {
            double dc = oracle_fv_diff(h[k - 1], h[k], h[k + 1], s[k - 1], s[k],
                                       s[k + 1]);
            double dl = 2.0 * (s[k] - s[k - 1]);
            double dr = 2.0 * (s[k + 1] - s[k]);
            if (signum1(dl) * signum1(dr) <= 0.0) {
                diff[k] = 0.0;
            } else {
                double m = f_min(f_min(fabs(dl), fabs(dc)), fabs(dr));
                diff[k] = f_sign(m, dc);
            }
        } else {
            diff[k] = 0.0;
        }
    }
    diff[0] = 0.0;
    diff[nk - 1] = 0.0;
}

/* redi_ppm_edge, ocean_redi.F90:194-225 */
double oracle_ppm_edge(double hkm1, double hk, double hkp1, double hkp2,
                       double ak, double akp1, double pk, double pkp1,
                       double h_neglect) {
    double r = hk + hkp1;
    if (r <= 0.0) return 0.5 * (ak + akp1);
    r = 1.0 / r;
    double e;
    if (hk < hkp1)
        e = ak + (hk * r) * (akp1 - ak);
    else
        e = akp1 + (hkp1 * r) * (ak - akp1);
    double r2a = 1.0 / ((2.0 * hk + hkp1) + h_neglect);
    double r2b = 1.0 / ((hk + 2.0 * hkp1) + h_neglect);
    double f1 = 1.0 / ((hk + hkp1) + (hkm1 + hkp2));
    double f2 = ((2.0 * (hkp1 * hk)) * r) * (((hkm1 + hk) * r2a) - ((hkp2 + hkp1) * r2b));
    double f3 = (hk * (hkm1 + hk)) * r2a;
    double f4 = (hkp1 * (hkp1 + hkp2)) * r2b;
    return e + f1 * (f2 * (akp1 - ak) - (f3 * pkp1 - f4 * pk));
}

/* redi_interface_scalar, ocean_redi.F90:231-254.  edge has nk+1 entries. */
void oracle_interface_scalar(int nk, const double *h, const double *tr,
                             double *edge, double h_neglect) {
    double *diff = (double *)malloc(sizeof(double) * (size_t)(nk > 0 ? nk : 1));
    oracle_plm_diff(nk, h, tr, diff);
    edge[0] = tr[0] - 0.5 * diff[0];
    for (int k = 1; k < nk; ++k) { /* reference k = 2..nk */
        int km2 = k - 2 < 0 ? 0 : k - 2;
        int kp1 = k + 1 > nk - 1 ? nk - 1 : k + 1;
        edge[k] = oracle_ppm_edge(h[km2], h[k - 1], h[k], h[kp1], tr[k - 1],
                                  tr[k], diff[k - 1], diff[k], h_neglect);
    }
    edge[nk] = tr[nk - 1] + 0.5 * diff[nk - 1];
    free(diff);
}

/* ---------------------------------------------------------------------- */
/* Neutral-surface sweep                                                   */
/* ---------------------------------------------------------------------- */

/* redi_interpolate_position, ocean_redi.F90:260-285 */
double oracle_interpolate_position(double dRhoNeg, double Pneg, double dRhoPos,
                                   double Ppos) {
    if ((Ppos > Pneg) && (dRhoPos - dRhoNeg >= 0.0)) {
        if (dRhoPos - dRhoNeg > 0.0)
            return f_min(1.0, f_max(0.0, -dRhoNeg / (dRhoPos - dRhoNeg)));
        if (dRhoNeg > 0.0) return 0.0;
        if (dRhoNeg < 0.0) return 1.0;
        return 0.5;
    }
    return 0.5; /* Ppos == Pneg, or the clamped MOM6 error branch */
}

/* redi_absolute_position, ocean_redi.F90:289-298 (K is 1-based) */
static inline double abs_pos(const double *P, int K, double frac) {
    return P[K - 1] + frac * (P[K] - P[K - 1]);
}

/* secant density difference between interface a of column X and interface
 * b of column Y: 0.5*((dRdT_X+dRdT_Y)*(T_X-T_Y) + (dRdS_X+dRdS_Y)*(S_X-S_Y))
 * (ocean_redi.F90:344-345, 367-371, 406-410); a, b are 1-based. */
static inline double sec_drho(const double *dTx, const double *Tx,
                              const double *dSx, const double *Sx, int a,
                              const double *dTy, const double *Ty,
                              const double *dSy, const double *Sy, int b) {
    return 0.5 * ((dTx[a - 1] + dTy[b - 1]) * (Tx[a - 1] - Ty[b - 1]) +
                  (dSx[a - 1] + dSy[b - 1]) * (Sx[a - 1] - Sy[b - 1]));
}

/* redi_neutral_positions_continuous, ocean_redi.F90:307-464.  Column arrays
 * have nk+1 interface entries (top-down); outputs have ns = 2nk+2 entries
 * (hEff ns-1, in the position unit of P). */
void oracle_neutral_positions_continuous(
    int nk, const double *Pl, const double *Tl, const double *Sl,
    const double *dRdTl, const double *dRdSl, const double *Pr,
    const double *Tr, const double *Sr, const double *dRdTr,
    const double *dRdSr, double *PoL, double *PoR, int *KoL, int *KoR,
    double *hEff) {
    const int ns = 2 * nk + 2;
    int kr = 1, kl = 1, lastK_right = 1, lastK_left = 1;
    double lastP_right = 0.0, lastP_left = 0.0;
    int reached_bottom = 0, search_l = 0, search_r = 0;

    for (int ks = 1; ks <= ns; ++ks) {
        const int s = ks - 1;
        int klm1 = kl - 1 > 1 ? kl - 1 : 1;
        int krm1 = kr - 1 > 1 ? kr - 1 : 1;
        double dRho = sec_drho(dRdTr, Tr, dRdSr, Sr, kr, dRdTl, Tl, dRdSl, Sl, kl);
        if (!reached_bottom) {
            if (dRho < 0.0) {
                search_l = 1; search_r = 0;
            } else if (dRho > 0.0) {
                search_r = 1; search_l = 0;
            } else if (kl + kr == 2) {
                search_l = 1; search_r = 0;
            } else {
                search_l = !search_l; search_r = !search_r;
            }
        }
        if (search_l) {
            double top = sec_drho(dRdTl, Tl, dRdSl, Sl, klm1, dRdTr, Tr, dRdSr, Sr, kr);
            double bot = sec_drho(dRdTl, Tl, dRdSl, Sl, klm1 + 1, dRdTr, Tr, dRdSr, Sr, kr);
            if (top > 0.0 || kr + kl == 2)
                PoL[s] = 0.0;
            else if (top >= bot)
                PoL[s] = 1.0;
            else
                PoL[s] = oracle_interpolate_position(top, Pl[klm1 - 1], bot, Pl[klm1]);
            if (PoL[s] >= 1.0 && klm1 < nk) {
                klm1 += 1;
                PoL[s] = PoL[s] - 1.0;
            }
            if ((double)(klm1 - lastK_left) + (PoL[s] - lastP_left) < 0.0) {
                PoL[s] = lastP_left;
                klm1 = lastK_left;
            }
            KoL[s] = klm1;
            if (kr <= nk) {
                PoR[s] = 0.0; KoR[s] = kr;
            } else {
                PoR[s] = 1.0; KoR[s] = nk;
            }
            if (kr <= nk) {
                kr += 1;
            } else {
                reached_bottom = 1; search_r = 1; search_l = 0;
            }
        } else if (search_r) {
            double top = sec_drho(dRdTr, Tr, dRdSr, Sr, krm1, dRdTl, Tl, dRdSl, Sl, kl);
            double bot = sec_drho(dRdTr, Tr, dRdSr, Sr, krm1 + 1, dRdTl, Tl, dRdSl, Sl, kl);
            if (top >= 0.0 || kr + kl == 2)
                PoR[s] = 0.0;
            else if (top >= bot)
                PoR[s] = 1.0;
            else
                PoR[s] = oracle_interpolate_position(top, Pr[krm1 - 1], bot, Pr[krm1]);
            if (PoR[s] >= 1.0 && krm1 < nk) {
                krm1 += 1;
                PoR[s] = PoR[s] - 1.0;
            }
            if ((double)(krm1 - lastK_right) + (PoR[s] - lastP_right) < 0.0) {
                PoR[s] = lastP_right;
                krm1 = lastK_right;
            }
            KoR[s] = krm1;
            if (kl <= nk) {
                PoL[s] = 0.0; KoL[s] = kl;
            } else {
                PoL[s] = 1.0; KoL[s] = nk;
            }
            if (kl <= nk) {
                kl += 1;
            } else {
                reached_bottom = 1; search_r = 0; search_l = 1;
            }
        }
        lastK_left = KoL[s];
        lastP_left = PoL[s];
        lastK_right = KoR[s];
        lastP_right = PoR[s];
        if (ks > 1) {
            double hL = abs_pos(Pl, KoL[s], PoL[s]) - abs_pos(Pl, KoL[s - 1], PoL[s - 1]);
            double hR = abs_pos(Pr, KoR[s], PoR[s]) - abs_pos(Pr, KoR[s - 1], PoR[s - 1]);
            if (hL + hR > 0.0)
                hEff[s - 1] = ((2.0 * hL) * hR) / (hL + hR);
            else
                hEff[s - 1] = 0.0;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Equation of state (ocean_eos.F90)                                       */
/* ---------------------------------------------------------------------- */

/* Wright (1997) Table A1 (ocean_eos.F90:83-97). */
#define W_A0 7.057924e-4
#define W_A1 3.480336e-7
#define W_A2 -1.112733e-7
#define W_B0 5.790749e8
#define W_B1 3.516535e6
#define W_B2 -4.002714e4
#define W_B3 2.084372e2
#define W_B4 5.944068e5
#define W_B5 -9.643486e3
#define W_C0 1.704853e5
#define W_C1 7.904722e2
#define W_C2 -7.984422e0
#define W_C3 5.140652e-2
#define W_C4 -2.302158e2
#define W_C5 -3.079464e0

/* Roquet et al. (2015) SpV polynomial, published coefficients with the
 * reference's scalings folded in (ocean_eos.F90:116-238).  Integer powers are
 * formed the way the reference compiler folds `x**n` (right-to-left binary
 * exponentiation: x^5 = x*x^4, x^6 = x^2*x^4). */
#define PW2(x) ((x) * (x))
#define PW3(x) (PW2(x) * (x))
#define PW4(x) (PW2(x) * PW2(x))
#define PW5(x) ((x) * PW4(x))
#define PW6(x) (PW2(x) * PW4(x))
#define Q_PA2KB 1.0e-8
#define Q_RDELTAS 24.0
#define Q_R1_S0 (0.875 / 35.16504)
#define Q_I_TS 0.025
#define Q_SRF (35.16504 / 35.0)
#define Q_V00 (-4.4015007269e-05 * Q_PA2KB)
#define Q_V01 (6.9232335784e-06 * PW2(Q_PA2KB))
#define Q_V02 (-7.5004675975e-07 * PW3(Q_PA2KB))
#define Q_V03 (1.7009109288e-08 * PW4(Q_PA2KB))
#define Q_V04 (-1.6884162004e-08 * PW5(Q_PA2KB))
#define Q_V05 (1.9613503930e-09 * PW6(Q_PA2KB))
#define S000 1.0772899069e-03
#define S100 (-3.1263658781e-04)
#define S200 6.7615860683e-04
#define S300 (-8.6127884515e-04)
#define S400 5.9010812596e-04
#define S500 (-2.1503943538e-04)
#define S600 3.2678954455e-05
#define S010 (-1.4949652640e-05 * Q_I_TS)
#define S110 (3.1866349188e-05 * Q_I_TS)
#define S210 (-3.8070687610e-05 * Q_I_TS)
#define S310 (2.9818473563e-05 * Q_I_TS)
#define S410 (-1.0011321965e-05 * Q_I_TS)
#define S510 (1.0751931163e-06 * Q_I_TS)
#define S020 (2.7546851539e-05 * PW2(Q_I_TS))
#define S120 (-3.6597334199e-05 * PW2(Q_I_TS))
#define S220 (3.4489154625e-05 * PW2(Q_I_TS))
#define S320 (-1.7663254122e-05 * PW2(Q_I_TS))
#define S420 (3.5965131935e-06 * PW2(Q_I_TS))
#define S030 (-1.6506828994e-05 * PW3(Q_I_TS))
#define S130 (2.4412359055e-05 * PW3(Q_I_TS))
#define S230 (-1.4606740723e-05 * PW3(Q_I_TS))
#define S330 (2.3293406656e-06 * PW3(Q_I_TS))
#define S040 (6.7896174634e-06 * PW4(Q_I_TS))
#define S140 (-8.7951832993e-06 * PW4(Q_I_TS))
#define S240 (4.4249040774e-06 * PW4(Q_I_TS))
#define S050 (-7.2535743349e-07 * PW5(Q_I_TS))
#define S150 (-3.4680559205e-07 * PW5(Q_I_TS))
#define S060 (1.9041365570e-07 * PW6(Q_I_TS))
#define S001 (-1.6889436589e-05 * Q_PA2KB)
#define S101 (2.1106556158e-05 * Q_PA2KB)
#define S201 (-2.1322804368e-05 * Q_PA2KB)
#define S301 (1.7347655458e-05 * Q_PA2KB)
#define S401 (-4.3209400767e-06 * Q_PA2KB)
#define S011 (1.5355844621e-05 * (Q_I_TS * Q_PA2KB))
#define S111 (2.0914122241e-06 * (Q_I_TS * Q_PA2KB))
#define S211 (-5.7751479725e-06 * (Q_I_TS * Q_PA2KB))
#define S311 (1.0767234341e-06 * (Q_I_TS * Q_PA2KB))
#define S021 (-9.6659393016e-06 * (PW2(Q_I_TS) * Q_PA2KB))
#define S121 (-7.0686982208e-07 * (PW2(Q_I_TS) * Q_PA2KB))
#define S221 (1.4488066593e-06 * (PW2(Q_I_TS) * Q_PA2KB))
#define S031 (3.1134283336e-06 * (PW3(Q_I_TS) * Q_PA2KB))
#define S131 (7.9562529879e-08 * (PW3(Q_I_TS) * Q_PA2KB))
#define S041 (-5.6590253863e-07 * (PW4(Q_I_TS) * Q_PA2KB))
#define S002 (1.0500241168e-06 * PW2(Q_PA2KB))
#define S102 (1.9600661704e-06 * PW2(Q_PA2KB))
#define S202 (-2.1666693382e-06 * PW2(Q_PA2KB))
#define S012 (-3.8541359685e-06 * (Q_I_TS * PW2(Q_PA2KB)))
#define S112 (1.0157632247e-06 * (Q_I_TS * PW2(Q_PA2KB)))
#define S022 (1.7178343158e-06 * (PW2(Q_I_TS) * PW2(Q_PA2KB)))
#define S003 (-4.1503454190e-07 * PW3(Q_PA2KB))
#define S103 (3.5627020989e-07 * PW3(Q_PA2KB))
#define S013 (-1.1293871415e-07 * (Q_I_TS * PW3(Q_PA2KB)))
/* dSV/dCT table: d/dzt of the SV table. */
#define A000 S010
#define A100 S110
#define A200 S210
#define A300 S310
#define A400 S410
#define A500 S510
#define A010 (2.0 * S020)
#define A110 (2.0 * S120)
#define A210 (2.0 * S220)
#define A310 (2.0 * S320)
#define A410 (2.0 * S420)
#define A020 (3.0 * S030)
#define A120 (3.0 * S130)
#define A220 (3.0 * S230)
#define A320 (3.0 * S330)
#define A030 (4.0 * S040)
#define A130 (4.0 * S140)
#define A230 (4.0 * S240)
#define A040 (5.0 * S050)
#define A140 (5.0 * S150)
#define A050 (6.0 * S060)
#define A001 S011
#define A101 S111
#define A201 S211
#define A301 S311
#define A011 (2.0 * S021)
#define A111 (2.0 * S121)
#define A211 (2.0 * S221)
#define A021 (3.0 * S031)
#define A121 (3.0 * S131)
#define A031 (4.0 * S041)
#define A002 S012
#define A102 S112
#define A012 (2.0 * S022)
#define A003 S013
/* dSV/dSA table: d/dzs with the 0.5*r1_S0 chain factor folded per term. */
#define B000 ((0.5 * S100) * Q_R1_S0)
#define B100 (S200 * Q_R1_S0)
#define B200 ((1.5 * S300) * Q_R1_S0)
#define B300 ((2.0 * S400) * Q_R1_S0)
#define B400 ((2.5 * S500) * Q_R1_S0)
#define B500 ((3.0 * S600) * Q_R1_S0)
#define B010 ((0.5 * S110) * Q_R1_S0)
#define B110 (S210 * Q_R1_S0)
#define B210 ((1.5 * S310) * Q_R1_S0)
#define B310 ((2.0 * S410) * Q_R1_S0)
#define B410 ((2.5 * S510) * Q_R1_S0)
#define B020 ((0.5 * S120) * Q_R1_S0)
#define B120 (S220 * Q_R1_S0)
#define B220 ((1.5 * S320) * Q_R1_S0)
#define B320 ((2.0 * S420) * Q_R1_S0)
#define B030 ((0.5 * S130) * Q_R1_S0)
#define B130 (S230 * Q_R1_S0)
#define B230 ((1.5 * S330) * Q_R1_S0)
#define B040 ((0.5 * S140) * Q_R1_S0)
#define B140 (S240 * Q_R1_S0)
#define B050 ((0.5 * S150) * Q_R1_S0)
#define B001 ((0.5 * S101) * Q_R1_S0)
#define B101 (S201 * Q_R1_S0)
#define B201 ((1.5 * S301) * Q_R1_S0)
#define B301 ((2.0 * S401) * Q_R1_S0)
#define B011 ((0.5 * S111) * Q_R1_S0)
#define B111 (S211 * Q_R1_S0)
#define B211 ((1.5 * S311) * Q_R1_S0)
#define B021 ((0.5 * S121) * Q_R1_S0)
#define B121 (S221 * Q_R1_S0)
#define B031 ((0.5 * S131) * Q_R1_S0)
#define B002 ((0.5 * S102) * Q_R1_S0)
#define B102 (S202 * Q_R1_S0)
#define B012 ((0.5 * S112) * Q_R1_S0)
#define B003 ((0.5 * S103) * Q_R1_S0)
/* 7-term PT->CT surface polynomial (gsw_CT_from_pt). */
#define Q_CP0 3991.86795711963
#define Q_CT_SFAC 0.0248826675584615

/* roquet_spv_point, ocean_eos.F90:282-448 */
void oracle_roquet_spv_point(double T_pt, double S_sp, double p, double *sv,
                             double *dsv_dt, double *dsv_ds) {
    const double SR = S_sp * Q_SRF;
    const double x2 = f_max(Q_CT_SFAC * SR, 1.0e-20);
    const double xx = sqrt(x2);
    const double yy = T_pt * 0.025;
    double c0 = 61.01362420681071 +
                x2 * (268.5520265845071 +
                      xx * (937.2099110620707 +
                            xx * (-1687.914374187449 + xx * 246.9598888781377)));
    double c1 = 168776.46138048015 +
                x2 * (-12019.028203559312 +
                      xx * (588.1802812170108 +
                            xx * (936.3206544460336 + xx * 123.59576582457964)));
    double c2 = -2735.2785605119625 +
                x2 * (3734.858026725145 +
                      xx * (248.39476522971285 +
                            xx * (-942.7827304544439 + xx * (-48.5891069025409))));
    double c3 = 2574.2164453821433 +
                x2 * (-2046.7671145057618 +
                      xx * (-3.871557904936333 + xx * 369.4389437509002));
    double c4 = -1536.6644434977543 +
                x2 * (465.28655623126450 +
                      xx * (-2.6268019854268356 + xx * (-33.83664947895248)));
    double c5 = 545.7340497931629 +
                x2 * (-0.6370820302831379 + xx * (-9.987880382780322));
    double c6 = -50.91091728474331 + x2 * (-10.650848542359153);
    double c7 = -18.30489878927802;
    double d0 = xx * ((2.0 * 268.5520265845071) +
                      xx * ((3.0 * 937.2099110620707) +
                            xx * ((4.0 * (-1687.914374187449)) +
                                  (xx * 5.0) * 246.9598888781377)));
    double d1 = xx * ((2.0 * (-12019.028203559312)) +
                      xx * ((3.0 * 588.1802812170108) +
                            xx * ((4.0 * 936.3206544460336) +
                                  (xx * 5.0) * 123.59576582457964)));
    double d2 = xx * ((2.0 * 3734.858026725145) +
                      xx * ((3.0 * 248.39476522971285) +
                            xx * ((4.0 * (-942.7827304544439)) +
                                  (xx * 5.0) * (-48.5891069025409))));
    double d3 = xx * ((2.0 * (-2046.7671145057618)) +
                      xx * ((3.0 * (-3.871557904936333)) +
                            (xx * 4.0) * 369.4389437509002));
    double d4 = xx * ((2.0 * 465.28655623126450) +
                      xx * ((3.0 * (-2.6268019854268356)) +
                            (xx * 4.0) * (-33.83664947895248)));
    double d5 = xx * ((2.0 * (-0.6370820302831379)) +
                      (xx * 3.0) * (-9.987880382780322));
    double d6 = (xx * 2.0) * (-10.650848542359153);
    double d7 = 0.0;
    double hh = c0 + yy * (c1 + yy * (c2 + yy * (c3 + yy * (c4 + yy * (c5 + yy * (c6 + yy * c7))))));
    double dh_dy =
        c1 + yy * ((2.0 * c2) +
                   yy * ((3.0 * c3) +
                         yy * ((4.0 * c4) +
                               yy * ((5.0 * c5) +
                                     yy * ((6.0 * c6) + (yy * 7.0) * c7)))));
    double dh_dx = d0 + yy * (d1 + yy * (d2 + yy * (d3 + yy * (d4 + yy * (d5 + yy * (d6 + yy * d7))))));
    const double CT = hh / Q_CP0;
    const double dct_dpt = (0.025 * dh_dy) / Q_CP0;
    const double dct_dsr = (dh_dx * (Q_CT_SFAC / (2.0 * xx))) / Q_CP0;

    const double zt = CT;
    const double zs = sqrt(fabs(S_sp * Q_SRF + Q_RDELTAS) * Q_R1_S0);
    const double zp = p;

    double ts3 = S003 + (zs * S103 + zt * S013);
    double ts2 = S002 + (zs * (S102 + zs * S202) + zt * (S012 + (zs * S112 + zt * S022)));
    double ts1 =
        S001 + (zs * (S101 + zs * (S201 + zs * (S301 + zs * S401))) +
                zt * (S011 + (zs * (S111 + zs * (S211 + zs * S311)) +
                              zt * (S021 + (zs * (S121 + zs * S221) +
                                            zt * (S031 + (zs * S131 + zt * S041)))))));
    double ts0 =
        zt * (S010 +
              (zs * (S110 + zs * (S210 + zs * (S310 + zs * (S410 + zs * S510)))) +
               zt * (S020 +
                     (zs * (S120 + zs * (S220 + zs * (S320 + zs * S420))) +
                      zt * (S030 +
                            (zs * (S130 + zs * (S230 + zs * S330)) +
                             zt * (S040 + (zs * (S140 + zs * S240) +
                                           zt * (S050 + (zs * S150 + zt * S060))))))))));
    double s0s0 = S000 + zs * (S100 + zs * (S200 + zs * (S300 + zs * (S400 + zs * (S500 + zs * S600)))));
    double s00p = zp * (Q_V00 + zp * (Q_V01 + zp * (Q_V02 + zp * (Q_V03 + zp * (Q_V04 + zp * Q_V05)))));
    *sv = ((ts0 + s0s0) + zp * (ts1 + zp * (ts2 + zp * ts3))) + s00p;

    double t3 = A003;
    double t2 = A002 + (zs * A102 + zt * A012);
    double t1 = A001 + (zs * (A101 + zs * (A201 + zs * A301)) +
                        zt * (A011 + (zs * (A111 + zs * A211) +
                                      zt * (A021 + (zs * A121 + zt * A031)))));
    double t0 =
        A000 + (zs * (A100 + zs * (A200 + zs * (A300 + zs * (A400 + zs * A500)))) +
                zt * (A010 + (zs * (A110 + zs * (A210 + zs * (A310 + zs * A410))) +
                              zt * (A020 + (zs * (A120 + zs * (A220 + zs * A320)) +
                                            zt * (A030 + (zt * (A040 + (zs * A140 + zt * A050)) +
                                                          zs * (A130 + zs * A230))))))));
    const double dsv_dct = t0 + zp * (t1 + zp * (t2 + zp * t3));

    double u3 = B003;
    double u2 = B002 + (zs * B102 + zt * B012);
    double u1 = B001 + (zs * (B101 + zs * (B201 + zs * B301)) +
                        zt * (B011 + (zs * (B111 + zs * B211) +
                                      zt * (B021 + (zs * B121 + zt * B031)))));
    double u0 =
        B000 + (zs * (B100 + zs * (B200 + zs * (B300 + zs * (B400 + zs * B500)))) +
                zt * (B010 + (zs * (B110 + zs * (B210 + zs * (B310 + zs * B410))) +
                              zt * (B020 + (zs * (B120 + zs * (B220 + zs * B320)) +
                                            zt * (B030 + (zt * (B040 + (zs * B140 + zt * B050)) +
                                                          zs * (B130 + zs * B230))))))));
    const double dsv_dsa = (u0 + zp * (u1 + zp * (u2 + zp * u3))) / zs;

    *dsv_dt = dsv_dct * dct_dpt;
    *dsv_ds = (dsv_dsa + dsv_dct * dct_dsr) * Q_SRF;
}

/* eos_specvol_derivs, ocean_eos.F90:450-514 */
void oracle_eos_specvol_derivs(const oracle_eos_t *eos, double T, double S,
                               double p, double *dsv_dt, double *dsv_ds) {
    if (eos->variant == 4) {
        double sv;
        oracle_roquet_spv_point(T, S, p, &sv, dsv_dt, dsv_ds);
    } else if (eos->variant == 2) {
        double T_sq = T * T;
        double p_0 = ((((W_B0 + W_B1 * T) + W_B2 * T_sq) + (W_B3 * T_sq) * T) + W_B4 * S) + (W_B5 * S) * T;
        double lam = ((((W_C0 + W_C1 * T) + W_C2 * T_sq) + (W_C3 * T_sq) * T) + W_C4 * S) + (W_C5 * S) * T;
        double dp0_dt = ((W_B1 + (2.0 * W_B2) * T) + (3.0 * W_B3) * T_sq) + W_B5 * S;
        double dlam_dt = ((W_C1 + (2.0 * W_C2) * T) + (3.0 * W_C3) * T_sq) + W_C5 * S;
        double dp0_ds = W_B4 + W_B5 * T;
        double dlam_ds = W_C4 + W_C5 * T;
        double inv_p = 1.0 / (p + p_0);
        double inv_p2 = inv_p * inv_p;
        *dsv_dt = (W_A1 + dlam_dt * inv_p) - (lam * dp0_dt) * inv_p2;
        *dsv_ds = (W_A2 + dlam_ds * inv_p) - (lam * dp0_ds) * inv_p2;
    } else {
        *dsv_dt = eos->alpha_T / (eos->rho0 * eos->rho0);
        *dsv_ds = -(eos->beta_S / (eos->rho0 * eos->rho0));
    }
}

/* eos_density_point, ocean_eos.F90:516-546 */
double oracle_eos_density_point(const oracle_eos_t *eos, double T, double S,
                                double p) {
    if (eos->variant == 4) {
        double sv, d1, d2;
        oracle_roquet_spv_point(T, S, p, &sv, &d1, &d2);
        return 1.0 / sv;
    } else if (eos->variant == 2) {
        double T_sq = T * T;
        double a0 = (W_A0 + W_A1 * T) + W_A2 * S;
        double p_0 = ((((W_B0 + W_B1 * T) + W_B2 * T_sq) + (W_B3 * T_sq) * T) + W_B4 * S) + (W_B5 * S) * T;
        double lam = ((((W_C0 + W_C1 * T) + W_C2 * T_sq) + (W_C3 * T_sq) * T) + W_C4 * S) + (W_C5 * S) * T;
        double pp = p + p_0;
        return pp / (lam + a0 * pp);
    }
    return (eos->rho0 + eos->beta_S * (S - eos->S_ref)) - eos->alpha_T * (T - eos->T_ref);
}

/* ---------------------------------------------------------------------- */
/* Phase A                                                                 */
/* ---------------------------------------------------------------------- */

/* redi_build_column, ocean_redi.F90:609-654.  h/thtr/shtr are native
 * bottom-up columns of nz; outputs are top-down with nz+1 interfaces. */
void oracle_build_column(int nz, const double *h_col, const double *thtr_col,
                         const double *shtr_col, const oracle_eos_t *eos,
                         double *Pint, double *Tint, double *Sint,
                         double *dRdT, double *dRdS) {
    double *htd = (double *)malloc(sizeof(double) * 3 * (size_t)nz);
    double *ttd = htd + nz, *std_ = htd + 2 * nz;
    for (int k = 0; k < nz; ++k) {
        int kf = nz - 1 - k;
        double he = f_max(h_col[k], H_DIV_EPS);
        htd[kf] = h_col[k];
        ttd[kf] = thtr_col[k] / he;
        std_[kf] = shtr_col[k] / he;
    }
    oracle_interface_scalar(nz, htd, ttd, Tint, H_NEGLECT);
    oracle_interface_scalar(nz, htd, std_, Sint, H_NEGLECT);
    Pint[0] = 0.0;
    for (int k = 0; k < nz; ++k) Pint[k + 1] = Pint[k] + (GRAVITY * eos->rho0) * htd[k];
    for (int k = 0; k <= nz; ++k) {
        double rho = oracle_eos_density_point(eos, Tint[k], Sint[k], Pint[k]);
        double dt, ds;
        oracle_eos_specvol_derivs(eos, Tint[k], Sint[k], Pint[k], &dt, &ds);
        dRdT[k] = -((rho * rho) * dt);
        dRdS[k] = -((rho * rho) * ds);
    }
    free(htd);
}

/* Opt-in neutral-slope taper -- NOT part of the reference, which tapers
 * nothing; the plain restatement never calls it (slope_max = 0).  Sublayer s
 * lies between surfaces s and s+1; its mid-depth in each column is the mean
 * of the absolute positions (redi_absolute_position, ocean_redi.F90:289-298)
 * of its bounding surfaces, converted Pa -> m; slope S = (zR - zL) * inv_dx.
 * Where |S| > slope_max, hEff (m) is scaled by (slope_max/|S|)^2, the
 * quadratic taper of Gerdes, Koberle and Willebrand (1991). */
static void slope_taper(int nz, const double *Pl, const double *Pr,
                        const double *PoL, const double *PoR, const int *KoL,
                        const int *KoR, double pa_to_h, double inv_dx,
                        double slope_max, double *hEff) {
    const int ns = 2 * nz + 2;
    for (int s = 0; s < ns - 1; ++s) {
        const double hm = hEff[s];
        if (hm == 0.0) continue;
        const double topL = abs_pos(Pl, KoL[s], PoL[s]), botL = abs_pos(Pl, KoL[s + 1], PoL[s + 1]);
        const double topR = abs_pos(Pr, KoR[s], PoR[s]), botR = abs_pos(Pr, KoR[s + 1], PoR[s + 1]);
        const double dz = ((0.5 * (topR + botR)) - (0.5 * (topL + botL))) * pa_to_h;
        const double aS = fabs(dz * inv_dx);
        if (aS > slope_max) {
            const double q = slope_max / aS;
            hEff[s] = hm * (q * q);
        }
    }
}

static void face_coeffs(int nz, const double *hL, const double *tL,
                        const double *sL, const double *hR, const double *tR,
                        const double *sR, const oracle_eos_t *eos,
                        double inv_dx, double slope_max,
                        double *PoL, double *PoR, int *KoL, int *KoR,
                        double *hEff);

/* redi_face_coeffs, ocean_redi.F90:656-692 (hEff converted Pa -> m). */
void oracle_face_coeffs(int nz, const double *hL, const double *tL,
                        const double *sL, const double *hR, const double *tR,
                        const double *sR, const oracle_eos_t *eos,
                        double *PoL, double *PoR, int *KoL, int *KoR,
                        double *hEff) {
    face_coeffs(nz, hL, tL, sL, hR, tR, sR, eos, 0.0, 0.0, PoL, PoR, KoL, KoR, hEff);
}

/* redi_face_coeffs followed by the opt-in slope taper (slope_max > 0). */
void oracle_face_coeffs_taper(int nz, const double *hL, const double *tL,
                              const double *sL, const double *hR, const double *tR,
                              const double *sR, const oracle_eos_t *eos,
                              double inv_dx, double slope_max,
                              double *PoL, double *PoR, int *KoL, int *KoR,
                              double *hEff) {
    face_coeffs(nz, hL, tL, sL, hR, tR, sR, eos, inv_dx, slope_max, PoL, PoR, KoL, KoR, hEff);
}

static void face_coeffs(int nz, const double *hL, const double *tL,
                        const double *sL, const double *hR, const double *tR,
                        const double *sR, const oracle_eos_t *eos,
                        double inv_dx, double slope_max,
                        double *PoL, double *PoR, int *KoL, int *KoR,
                        double *hEff) {
    const int n1 = nz + 1, ns = 2 * nz + 2;
    double *w = (double *)malloc(sizeof(double) * 10 * (size_t)n1);
    double *Pl = w, *Tl = w + n1, *Sl = w + 2 * n1, *dTl = w + 3 * n1, *dSl = w + 4 * n1;
    double *Pr = w + 5 * n1, *Tr = w + 6 * n1, *Sr = w + 7 * n1, *dTr = w + 8 * n1, *dSr = w + 9 * n1;
    oracle_build_column(nz, hL, tL, sL, eos, Pl, Tl, Sl, dTl, dSl);
    oracle_build_column(nz, hR, tR, sR, eos, Pr, Tr, Sr, dTr, dSr);
    oracle_neutral_positions_continuous(nz, Pl, Tl, Sl, dTl, dSl, Pr, Tr, Sr,
                                        dTr, dSr, PoL, PoR, KoL, KoR, hEff);
    const double pa_to_h = 1.0 / (GRAVITY * eos->rho0);
    for (int s = 0; s < ns - 1; ++s) hEff[s] = hEff[s] * pa_to_h;
    if (slope_max > 0.0)
        slope_taper(nz, Pl, Pr, PoL, PoR, KoL, KoR, pa_to_h, inv_dx, slope_max, hEff);
    free(w);
}

static void calc_faces(int nx, int ny, int nz, const oracle_eos_t *eos,
                       const double *h, const double *t, const double *sa,
                       const double *wet, int nfa, int nfb, int is_u,
                       const double *inv_dx, double slope_max,
                       double *PoL, double *PoR, int *KoL, int *KoR,
                       double *hEff) {
    const int ns = 2 * nz + 2;
    const size_t fstride = (size_t)nfa * nfb, cstride = (size_t)nx * ny;
    /* u: faces i=2..nx, j=1..ny (ocean_redi.F90:720); v: j=2..ny, i=1..nx
     * (:779).  Loop indices here are 0-based face coordinates. */
    const int ia = is_u ? 1 : 0, ib = is_u ? nx - 1 : nx - 1;
    const int ja = is_u ? 0 : 1, jb = ny - 1;
    const long nfi = (long)(ib - ia + 1), nfj = (long)(jb - ja + 1);
#pragma omp parallel
    {
        double *col = (double *)malloc(sizeof(double) * 6 * (size_t)nz);
        double *po = (double *)malloc(sizeof(double) * (3 * (size_t)ns));
        int *ko = (int *)malloc(sizeof(int) * 2 * (size_t)ns);
#pragma omp for schedule(dynamic, 64)
        for (long f = 0; f < nfi * nfj; ++f) {
            const int i = ia + (int)(f % nfi), j = ja + (int)(f / nfi);
            const size_t fo = (size_t)i + (size_t)nfa * j;
            for (int s = 0; s < ns; ++s) {
                PoL[fo + s * fstride] = 0.0;
                PoR[fo + s * fstride] = 0.0;
                KoL[fo + s * fstride] = 1;
                KoR[fo + s * fstride] = 1;
            }
            for (int s = 0; s < ns - 1; ++s) hEff[fo + s * fstride] = 0.0;
            if (!(wet[fo] > 0.0)) continue;
            const size_t cl = is_u ? (size_t)(i - 1) + (size_t)nx * j : (size_t)i + (size_t)nx * (j - 1);
            const size_t cr = (size_t)i + (size_t)nx * j;
            double *hL = col, *tL = col + nz, *sL = col + 2 * nz;
            double *hR = col + 3 * nz, *tR = col + 4 * nz, *sR = col + 5 * nz;
            for (int k = 0; k < nz; ++k) {
                hL[k] = h[cl + k * cstride]; tL[k] = t[cl + k * cstride]; sL[k] = sa[cl + k * cstride];
                hR[k] = h[cr + k * cstride]; tR[k] = t[cr + k * cstride]; sR[k] = sa[cr + k * cstride];
            }
            face_coeffs(nz, hL, tL, sL, hR, tR, sR, eos, slope_max > 0.0 ? inv_dx[fo] : 0.0,
                        slope_max, po, po + ns, ko, ko + ns, po + 2 * ns);
            for (int s = 0; s < ns; ++s) {
                PoL[fo + s * fstride] = po[s];
                PoR[fo + s * fstride] = po[ns + s];
                KoL[fo + s * fstride] = ko[s];
                KoR[fo + s * fstride] = ko[ns + s];
            }
            for (int s = 0; s < ns - 1; ++s) hEff[fo + s * fstride] = po[2 * ns + s];
        }
        free(col); free(po); free(ko);
    }
}

/* redi_calc_coeffs_x / _y, ocean_redi.F90:694-812 */
void oracle_calc_coeffs(int nx, int ny, int nz, const oracle_eos_t *eos,
                        const double *h_layer, const double *t_htr,
                        const double *s_htr, const double *wet_u,
                        const double *wet_v, double *uPoL, double *uPoR,
                        int *uKoL, int *uKoR, double *uhEff, double *vPoL,
                        double *vPoR, int *vKoL, int *vKoR, double *vhEff) {
    oracle_calc_coeffs_taper(nx, ny, nz, eos, h_layer, t_htr, s_htr, wet_u, wet_v,
                             NULL, NULL, 0.0, uPoL, uPoR, uKoL, uKoR, uhEff,
                             vPoL, vPoR, vKoL, vKoR, vhEff);
}

/* oracle_calc_coeffs with the opt-in slope taper (idxCu / idyCv give the
 * face spacings; slope_max = 0 is the plain restatement). */
void oracle_calc_coeffs_taper(int nx, int ny, int nz, const oracle_eos_t *eos,
                              const double *h_layer, const double *t_htr,
                              const double *s_htr, const double *wet_u,
                              const double *wet_v, const double *idxCu,
                              const double *idyCv, double slope_max,
                              double *uPoL, double *uPoR, int *uKoL, int *uKoR,
                              double *uhEff, double *vPoL, double *vPoR, int *vKoL,
                              int *vKoR, double *vhEff) {
    calc_faces(nx, ny, nz, eos, h_layer, t_htr, s_htr, wet_u, nx + 1, ny, 1,
               idxCu, slope_max, uPoL, uPoR, uKoL, uKoR, uhEff);
    calc_faces(nx, ny, nz, eos, h_layer, t_htr, s_htr, wet_v, nx, ny + 1, 0,
               idyCv, slope_max, vPoL, vPoR, vKoL, vKoR, vhEff);
}

/* ---------------------------------------------------------------------- */
/* Phase B                                                                 */
/* ---------------------------------------------------------------------- */

/* redi_ppm_ave, ocean_redi.F90:833-850 */
double oracle_ppm_ave(double xL, double xR, double aL, double aR, double aMean) {
    double dx = xR - xL;
    double xave = 0.5 * (xR + xL);
    double a6o3 = 2.0 * aMean - (aL + aR);
    double a6 = 3.0 * a6o3;
    if (dx > 0.0 && dx <= 1.0)
        return (aL + xave * ((aR - aL) + a6)) - a6o3 * ((xR * xR + xR * xL) + xL * xL);
    return (aL + (aR - aL) * xR) + (a6 * xR) * (1.0 - xR);
}

/* redi_tracer_column, ocean_redi.F90:852-891 */
void oracle_tracer_column(int nz, const double *h_col, const double *htr_col,
                          double *Tlay, double *Tint, double *aLe,
                          double *aRe) {
    double *htd = (double *)malloc(sizeof(double) * (size_t)nz);
    for (int k = 0; k < nz; ++k) {
        int kf = nz - 1 - k;
        double he = f_max(h_col[k], H_DIV_EPS);
        htd[kf] = h_col[k];
        Tlay[kf] = htr_col[k] / he;
    }
    oracle_interface_scalar(nz, htd, Tlay, Tint, H_NEGLECT);
    for (int k = 0; k < nz; ++k) {
        double alk = Tint[k], ark = Tint[k + 1], tlk = Tlay[k];
        if (signum1(ark - tlk) * signum1(tlk - alk) <= 0.0) {
            alk = tlk;
            ark = tlk;
        } else if (f_sign(3.0, ark - alk) * ((tlk - alk) + (tlk - ark)) > fabs(ark - alk)) {
            alk = tlk + 2.0 * (tlk - ark);
        } else if (f_sign(3.0, ark - alk) * ((tlk - alk) + (tlk - ark)) < -fabs(ark - alk)) {
            ark = tlk + 2.0 * (tlk - alk);
        }
        aLe[k] = alk;
        aRe[k] = ark;
    }
    free(htd);
}

/* redi_sublayer_dT, ocean_redi.F90:893-925 (1-based layer indices) */
double oracle_sublayer_dT(int klt, int klb, int krt, int krb, double PoLt,
                          double PoLb, double PoRt, double PoRb,
                          const double *TlL, const double *TiL,
                          const double *aLL, const double *aRL,
                          const double *TlR, const double *TiR,
                          const double *aLR, const double *aRR) {
    double tlt = (1.0 - PoLt) * TiL[klt - 1] + PoLt * TiL[klt];
    double tlb = (1.0 - PoLb) * TiL[klb - 1] + PoLb * TiL[klb];
    double trt = (1.0 - PoRt) * TiR[krt - 1] + PoRt * TiR[krt];
    double trb = (1.0 - PoRb) * TiR[krb - 1] + PoRb * TiR[krb];
    double tlay = oracle_ppm_ave(PoLt, PoLb + (double)(klb - klt), aLL[klt - 1],
                                 aRL[klt - 1], TlL[klt - 1]);
    double trlay = oracle_ppm_ave(PoRt, PoRb + (double)(krb - krt), aLR[krt - 1],
                                  aRR[krt - 1], TlR[krt - 1]);
    double dT_top = trt - tlt;
    double dT_bot = trb - tlb;
    double dT_ave = 0.5 * (dT_top + dT_bot);
    double dT_layer = trlay - tlay;
    if (signum1(dT_top) * signum1(dT_bot) <= 0.0 ||
        signum1(dT_ave) * signum1(dT_layer) <= 0.0)
        return 0.0;
    return dT_layer;
}

typedef struct {
    double *Tl, *Ti, *aL, *aR;
} tcol_t;

static void build_tcol(int nz, const double *h, const double *hTr_in,
                       size_t c, size_t cstride, double *tmp, tcol_t *o) {
    double *hc = tmp, *tc = tmp + nz;
    for (int k = 0; k < nz; ++k) {
        hc[k] = h[c + k * cstride];
        tc[k] = hTr_in[c + k * cstride];
    }
    oracle_tracer_column(nz, hc, tc, o->Tl, o->Ti, o->aL, o->aR);
}

/* redi_face_flux, ocean_redi.F90:1037-1092 (columns pre-built) */
static void face_flux(int nz, size_t fo, size_t fstride, const tcol_t *L,
                      const tcol_t *R, const double *PoL, const double *PoR,
                      const int *KoL, const int *KoR, const double *hEff,
                      double coef, int is_left, double *dTr) {
    const int ns = 2 * nz + 2;
    for (int s = 0; s < ns - 1; ++s) {
        const double he = hEff[fo + s * fstride];
        if (he != 0.0) {
            const size_t a = fo + s * fstride, b = fo + (s + 1) * fstride;
            double d = oracle_sublayer_dT(KoL[a], KoL[b], KoR[a], KoR[b], PoL[a],
                                          PoL[b], PoR[a], PoR[b], L->Tl, L->Ti,
                                          L->aL, L->aR, R->Tl, R->Ti, R->aL, R->aR);
            double flx = (d * he) * coef;
            if (is_left)
                dTr[nz - KoL[a]] += flx;
            else
                dTr[nz - KoR[a]] -= flx;
        }
    }
}

/* redi_apply_flux_impl, ocean_redi.F90:1094-1181 */
void oracle_apply_flux_tracer(
    int nx, int ny, int nz, double dt, int nghost, int nxp, int nyp,
    int wall_w, int wall_e, int wall_s, int wall_n, const double *khtr_u,
    const double *khtr_v, const double *dy_cu, const double *dx_cv,
    const double *idxCu, const double *idyCv, const double *areaT,
    const double *h_layer, const double *hTr_in, double *hTr,
    const double *uPoL, const double *uPoR, const int *uKoL, const int *uKoR,
    const double *uhEff, const double *vPoL, const double *vPoR,
    const int *vKoL, const int *vKoR, const double *vhEff) {
    /* 1-based wall face indices (:1139-1142) */
    const int wuf_w = nghost + 1, wuf_e = nghost + nxp + 1;
    const int wvf_s = nghost + 1, wvf_n = nghost + nyp + 1;
    const size_t cstride = (size_t)nx * ny;
    const size_t ufs = (size_t)(nx + 1) * ny, vfs = (size_t)nx * (ny + 1);
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * (size_t)(2 * nz + 2 * 4 * (nz + 1) + nz));
        double *tmp = buf, *dTr = buf + 2 * nz;
        double *cb = dTr + nz;
        tcol_t C = {cb, cb + nz, cb + 2 * nz + 1, cb + 3 * nz + 1};
        double *nb = cb + 4 * (nz + 1);
        tcol_t N = {nb, nb + nz, nb + 2 * nz + 1, nb + 3 * nz + 1};
#pragma omp for schedule(static)
        for (long c = 0; c < (long)cstride; ++c) {
            const int i = (int)(c % nx) + 1, j = (int)(c / nx) + 1; /* 1-based */
            for (int k = 0; k < nz; ++k) dTr[k] = 0.0;
            /* The centre column is rebuilt per face in the reference; the
             * reconstruction is a pure function of the column, so building
             * it once gives the same bits (ocean_redi.F90:1193-1199). */
            build_tcol(nz, h_layer, hTr_in, (size_t)c, cstride, tmp, &C);
            if (i >= 2 && !((wall_w && i == wuf_w) || (wall_e && i == wuf_e))) {
                size_t fo = (size_t)(i - 1) + (size_t)(nx + 1) * (j - 1);
                build_tcol(nz, h_layer, hTr_in, (size_t)c - 1, cstride, tmp, &N);
                double coef = ((dt * khtr_u[fo]) * dy_cu[fo]) * idxCu[fo];
                face_flux(nz, fo, ufs, &N, &C, uPoL, uPoR, uKoL, uKoR, uhEff, coef, 0, dTr);
            }
            if (i <= nx - 1 && !((wall_w && i + 1 == wuf_w) || (wall_e && i + 1 == wuf_e))) {
                size_t fo = (size_t)i + (size_t)(nx + 1) * (j - 1);
                build_tcol(nz, h_layer, hTr_in, (size_t)c + 1, cstride, tmp, &N);
                double coef = ((dt * khtr_u[fo]) * dy_cu[fo]) * idxCu[fo];
                face_flux(nz, fo, ufs, &C, &N, uPoL, uPoR, uKoL, uKoR, uhEff, coef, 1, dTr);
            }
            if (j >= 2 && !((wall_s && j == wvf_s) || (wall_n && j == wvf_n))) {
                size_t fo = (size_t)(i - 1) + (size_t)nx * (j - 1);
                build_tcol(nz, h_layer, hTr_in, (size_t)c - nx, cstride, tmp, &N);
                double coef = ((dt * khtr_v[fo]) * dx_cv[fo]) * idyCv[fo];
                face_flux(nz, fo, vfs, &N, &C, vPoL, vPoR, vKoL, vKoR, vhEff, coef, 0, dTr);
            }
            if (j <= ny - 1 && !((wall_s && j + 1 == wvf_s) || (wall_n && j + 1 == wvf_n))) {
                size_t fo = (size_t)(i - 1) + (size_t)nx * j;
                build_tcol(nz, h_layer, hTr_in, (size_t)c + nx, cstride, tmp, &N);
                double coef = ((dt * khtr_v[fo]) * dx_cv[fo]) * idyCv[fo];
                face_flux(nz, fo, vfs, &C, &N, vPoL, vPoR, vKoL, vKoR, vhEff, coef, 1, dTr);
            }
            const double iaij = 1.0 / areaT[c];
            for (int k = 0; k < nz; ++k) hTr[(size_t)c + k * cstride] = hTr[(size_t)c + k * cstride] + dTr[k] * iaij;
        }
        free(buf);
    }
}

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* fcfd_oracle.c -- CPU restatement of the reference's cell and interface update.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, see fcfd_oracle.h).  Built with -ffp-contract=off so that
 * every +,-,*,/ rounds exactly as the reference's Fortran does; expression trees follow the Fortran
 * source's left-to-right evaluation.  Paths are relative to the reference root; CKLIB / TRANLIB
 * lines refer to WORK/KineticModels/Chem_bin/GRI3.0_CH4/cklib.f and tranlib.f.
 */
#include "fcfd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Transcendental functions of the kinetics, transport and step-size code.  Default build: the
 * library's own fcfd_math.h, the same operations the GPU kernels execute, so the restatement and the
 * kernels round identically.  With -DFO_LIBM_MATH (libfcfd_oracle_libm.so): the C library's exp /
 * log / log10 / pow, which is what the reference Fortran calls, so that build reproduces the
 * reference's own outputs bit for bit. */
#ifdef FO_LIBM_MATH
#define FO_EXP exp
#define FO_LOG log
#define FO_LOG10 log10
#define FO_POW pow
#define FO_EXP10(x) pow(10.0, (x))
#define FO_TANH tanh
#define FO_EXPM1 expm1
#else
#include "../fortrancfd_amd/csrc/fcfd_math.h"
#define FO_EXP fm_exp
#define FO_LOG fm_log
#define FO_LOG10 fm_log10
#define FO_POW fm_pow
#define FO_EXP10 fm_exp10
#define FO_TANH fm_tanh
#define FO_EXPM1 fm_expm1
#endif

/* ============================================================================================ */
/* Mechanism tables                                                                              */
/* ============================================================================================ */
struct fo_mech {
    int kk, ii, mm, mxsp, mxtb, mxtp, ncp, npar, nfar;
    int nfal, nthb;
    double ru, pa, small, big, exparg;
    double *wt;            /* [kk]                                   */
    int *nt;               /* [kk]  number of NASA temperatures      */
    double *tt;            /* [kk][mxtp]                             */
    double *aa;            /* [kk][mxtp-1][ncp+2]                    */
    int *ns, *nr;          /* [ii]                                   */
    double *par;           /* [ii][npar+1]   A, b, E/R, 1.0          */
    int *nu, *nk;          /* [ii][mxsp]                             */
    int *ifl, *ifo, *kfl;  /* [nfal]                                 */
    double *fpar;          /* [nfal][nfar]                           */
    int *itb, *ntbs, *nktb;/* [nthb], [nthb], [nthb][mxtb]           */
    double *aik;           /* [nthb][mxtb]                           */
    double *h298;          /* [kk] CKHMS(298.15) erg/g               */
    int *ncf;              /* [kk][mm] element counts                */
    /* transport */
    int has_tran, no;
    double patmos;
    double *twt, *cofeta, *coflam, *cofd; /* [kk], [kk][no], [kk][no], [kk][kk][no] */
};

typedef struct {
    unsigned char *buf;
    long len, pos;
} recfile;

static int rf_open(recfile *r, const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    r->len = ftell(f);
    fseek(f, 0, SEEK_SET);
    r->buf = (unsigned char *)malloc((size_t)r->len);
    r->pos = 0;
    size_t got = fread(r->buf, 1, (size_t)r->len, f);
    fclose(f);
    return got == (size_t)r->len ? 0 : -1;
}

/* Fortran sequential unformatted record: int32 length, payload, int32 length. */
static const unsigned char *rf_next(recfile *r, int32_t *n) {
    if (r->pos + 4 > r->len) return NULL;
    int32_t a, b;
    memcpy(&a, r->buf + r->pos, 4);
    if (a < 0 || r->pos + 8 + a > r->len) return NULL;
    memcpy(&b, r->buf + r->pos + 4 + a, 4);
    if (a != b) return NULL;
    const unsigned char *p = r->buf + r->pos + 4;
    r->pos += 8 + a;
    *n = a;
    return p;
}

static int geti(const unsigned char **p) { int32_t v; memcpy(&v, *p, 4); *p += 4; return v; }
static double getd(const unsigned char **p) { double v; memcpy(&v, *p, 8); *p += 8; return v; }

fo_mech *fo_mech_load(const char *chem_bin, const char *tran_bin) {
    recfile rf;
    if (rf_open(&rf, chem_bin)) { fprintf(stderr, "fo: cannot read %s\n", chem_bin); return NULL; }
    fo_mech *m = (fo_mech *)calloc(1, sizeof(fo_mech));
    int32_t n;
    const unsigned char *p = rf_next(&rf, &n);                     /* VERS, PREC, KERR */
    p = rf_next(&rf, &n);                                          /* sizes            */
    int hdr[29];
    for (int i = 0; i < 29; ++i) hdr[i] = geti(&p);
    m->mm = hdr[3]; m->kk = hdr[4]; m->ii = hdr[5]; m->mxsp = hdr[6]; m->mxtb = hdr[7];
    m->mxtp = hdr[8]; m->ncp = hdr[9]; m->npar = hdr[10]; m->nfar = hdr[12];
    int nrv = hdr[13], nfl = hdr[14], ntb = hdr[15], nlt = hdr[16], nrl = hdr[17], nw = hdr[18];
    int nei = hdr[20], nijan = hdr[22], nif1 = hdr[24], nex = hdr[25], nsto = hdr[26], nor = hdr[27];
    if (nrv || nlt || nrl || nw || nei || nijan || nif1 || nex || nsto || nor) {
        fprintf(stderr, "fo: mechanism uses CHEMKIN features outside this restatement\n");
        free(m);
        free(rf.buf);
        return NULL;
    }
    m->nfal = nfl; m->nthb = ntb;
    /* CKINIT constants (cklib.f:4050-4057) */
    m->ru = 8.314510e7; m->pa = 1.01325e6;
    m->small = pow(10.0, -300); m->big = pow(10.0, 300); m->exparg = log(m->big);
    const int kk = m->kk, ii = m->ii, mxtp = m->mxtp, ncp2 = m->ncp + 2, mm = m->mm;
    p = rf_next(&rf, &n);                                          /* elements         */
    p = rf_next(&rf, &n);                                          /* species          */
    m->wt = calloc(kk, sizeof(double)); m->nt = calloc(kk, sizeof(int));
    m->tt = calloc((size_t)kk * mxtp, sizeof(double));
    m->aa = calloc((size_t)kk * (mxtp - 1) * ncp2, sizeof(double));
    m->ncf = calloc((size_t)kk * mm, sizeof(int));
    for (int k = 0; k < kk; ++k) {
        p += 16;
        for (int e = 0; e < mm; ++e) m->ncf[k * mm + e] = geti(&p);
        (void)geti(&p); (void)geti(&p);
        m->wt[k] = getd(&p);
        m->nt[k] = geti(&p);
        for (int l = 0; l < mxtp; ++l) m->tt[k * mxtp + l] = getd(&p);
        for (int l = 0; l < mxtp - 1; ++l)
            for (int c = 0; c < ncp2; ++c) m->aa[(k * (mxtp - 1) + l) * ncp2 + c] = getd(&p);
    }
    p = rf_next(&rf, &n);                                          /* reactions        */
    const int np1 = m->npar + 1, mxsp = m->mxsp;
    m->ns = calloc(ii, sizeof(int)); m->nr = calloc(ii, sizeof(int));
    m->par = calloc((size_t)ii * np1, sizeof(double));
    m->nu = calloc((size_t)ii * mxsp, sizeof(int)); m->nk = calloc((size_t)ii * mxsp, sizeof(int));
    for (int i = 0; i < ii; ++i) {
        m->ns[i] = geti(&p); m->nr[i] = geti(&p);
        for (int c = 0; c < m->npar; ++c) m->par[i * np1 + c] = getd(&p);
        m->par[i * np1 + m->npar] = 1.0;
        for (int s = 0; s < mxsp; ++s) { m->nu[i * mxsp + s] = geti(&p); m->nk[i * mxsp + s] = geti(&p); }
    }
    if (nfl > 0) {
        p = rf_next(&rf, &n);
        m->ifl = calloc(nfl, sizeof(int)); m->ifo = calloc(nfl, sizeof(int));
        m->kfl = calloc(nfl, sizeof(int)); m->fpar = calloc((size_t)nfl * m->nfar, sizeof(double));
        for (int f = 0; f < nfl; ++f) {
            m->ifl[f] = geti(&p); m->ifo[f] = geti(&p); m->kfl[f] = geti(&p);
            for (int c = 0; c < m->nfar; ++c) m->fpar[f * m->nfar + c] = getd(&p);
        }
    }
    if (ntb > 0) {
        p = rf_next(&rf, &n);
        m->itb = calloc(ntb, sizeof(int)); m->ntbs = calloc(ntb, sizeof(int));
        m->nktb = calloc((size_t)ntb * m->mxtb, sizeof(int));
        m->aik = calloc((size_t)ntb * m->mxtb, sizeof(double));
        for (int t = 0; t < ntb; ++t) {
            m->itb[t] = geti(&p); m->ntbs[t] = geti(&p);
            for (int l = 0; l < m->mxtb; ++l) {
                m->nktb[t * m->mxtb + l] = geti(&p);
                m->aik[t * m->mxtb + l] = getd(&p);
            }
        }
    }
    free(rf.buf);
    m->h298 = calloc(kk, sizeof(double));
    fo_ckhms(m, 298.15, m->h298);

    if (tran_bin) {
        recfile tf;
        if (rf_open(&tf, tran_bin)) { fprintf(stderr, "fo: cannot read %s\n", tran_bin); fo_mech_free(m); return NULL; }
        p = rf_next(&tf, &n);
        p = rf_next(&tf, &n);
        (void)geti(&p); (void)geti(&p);
        int no = geti(&p), tk = geti(&p), nlite = geti(&p);
        if (tk != kk) { fprintf(stderr, "fo: tran.bin species count mismatch\n"); fo_mech_free(m); free(tf.buf); return NULL; }
        p = rf_next(&tf, &n);
        m->no = no;
        m->patmos = getd(&p);
        m->twt = calloc(kk, sizeof(double));
        for (int k = 0; k < kk; ++k) {
            m->twt[k] = getd(&p);
            for (int c = 0; c < 5; ++c) (void)getd(&p);
            (void)geti(&p);
        }
        m->coflam = calloc((size_t)kk * no, sizeof(double));
        m->cofeta = calloc((size_t)kk * no, sizeof(double));
        m->cofd = calloc((size_t)kk * kk * no, sizeof(double));
        for (int i = 0; i < kk * no; ++i) m->coflam[i] = getd(&p);
        for (int i = 0; i < kk * no; ++i) m->cofeta[i] = getd(&p);
        for (int i = 0; i < kk * kk * no; ++i) m->cofd[i] = getd(&p);
        (void)nlite;
        m->has_tran = 1;
        free(tf.buf);
    }
    return m;
}

void fo_mech_free(fo_mech *m) {
    if (!m) return;
    void *ptrs[] = {m->wt, m->nt, m->tt, m->aa, m->ns, m->nr, m->par, m->nu, m->nk, m->ifl, m->ifo,
                    m->kfl, m->fpar, m->itb, m->ntbs, m->nktb, m->aik, m->h298, m->twt, m->cofeta,
                    m->coflam, m->cofd, m->ncf};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(m);
}

int fo_mech_kk(const fo_mech *m) { return m->kk; }
int fo_mech_mm(const fo_mech *m) { return m->mm; }
void fo_mech_ncf(const fo_mech *m, int *ncf) { memcpy(ncf, m->ncf, sizeof(int) * m->kk * m->mm); }
int fo_mech_ii(const fo_mech *m) { return m->ii; }
void fo_mech_weights(const fo_mech *m, double *wt) { memcpy(wt, m->wt, sizeof(double) * m->kk); }
void fo_mech_nasa(const fo_mech *m, double *tt, double *aa) {
    memcpy(tt, m->tt, sizeof(double) * m->kk * m->mxtp);
    memcpy(aa, m->aa, sizeof(double) * m->kk * (m->mxtp - 1) * (m->ncp + 2));
}
void fo_mech_arrhenius(const fo_mech *m, double *par, int *nu, int *nk, int *ns, int *nr) {
    memcpy(par, m->par, sizeof(double) * m->ii * (m->npar + 1));
    memcpy(nu, m->nu, sizeof(int) * m->ii * m->mxsp);
    memcpy(nk, m->nk, sizeof(int) * m->ii * m->mxsp);
    memcpy(ns, m->ns, sizeof(int) * m->ii);
    memcpy(nr, m->nr, sizeof(int) * m->ii);
}

/* ============================================================================================ */
/* CKLIB thermo                                                                                  */
/* ============================================================================================ */
/* T**n for small integer n the way the Fortran integer power is evaluated (square-and-multiply). */
static double ipow(double x, int n) {
    if (n < 0) return 1.0 / ipow(x, -n);
    double r = 1.0;
    int first = 1;
    while (n) {
        if (n & 1) { r = first ? x : r * x; first = 0; }
        n >>= 1;
        if (n) x = x * x;
    }
    return r;
}

/* NASA range index: L increments while T > TT(K,N), N = 2..NT-1 (cklib.f:3544-3548). */
static const double *nasa_coef(const fo_mech *m, int k, double T) {
    int l = 0;
    for (int n = 1; n < m->nt[k] - 1; ++n)
        if (T > m->tt[k * m->mxtp + n]) ++l;
    return m->aa + (k * (m->mxtp - 1) + l) * (m->ncp + 2);
}

void fo_ckhml(const fo_mech *m, double T, double *hml) {          /* cklib.f:3537-3556 */
    double rut = T * m->ru, tn[10];
    tn[0] = 1.0;
    for (int n = 2; n <= m->ncp; ++n) tn[n - 1] = ipow(T, n - 1) / n;
    for (int k = 0; k < m->kk; ++k) {
        const double *a = nasa_coef(m, k, T);
        double sum = 0.0;
        for (int n = 0; n < m->ncp; ++n) sum = sum + tn[n] * a[n];
        hml[k] = rut * (sum + a[m->ncp] / T);
    }
}

void fo_ckuml(const fo_mech *m, double T, double *uml) {          /* cklib.f:8446-8450 */
    fo_ckhml(m, T, uml);
    double rut = T * m->ru;
    for (int k = 0; k < m->kk; ++k) uml[k] = uml[k] - rut;
}

void fo_ckcpml(const fo_mech *m, double T, double *cpml) {        /* cklib.f:1849-1865 */
    double tn[10];
    tn[0] = 1.0;
    for (int n = 2; n <= m->ncp; ++n) tn[n - 1] = ipow(T, n - 1);
    for (int k = 0; k < m->kk; ++k) {
        const double *a = nasa_coef(m, k, T);
        double c = 0.0;
        for (int n = 0; n < m->ncp; ++n) c = c + m->ru * tn[n] * a[n];
        cpml[k] = c;
    }
}

void fo_cksmh(const fo_mech *m, double T, double *smh) {          /* cklib.f:7540-7560 */
    double tn[10];
    tn[0] = FO_LOG(T) - 1.0;
    for (int n = 2; n <= m->ncp; ++n) tn[n - 1] = ipow(T, n - 1) / ((n - 1) * n);
    for (int k = 0; k < m->kk; ++k) {
        const double *a = nasa_coef(m, k, T);
        double sum = 0.0;
        for (int n = 0; n < m->ncp; ++n) sum = sum + tn[n] * a[n];
        smh[k] = sum + a[m->ncp + 1] - a[m->ncp] / T;
    }
}

void fo_ckhms(const fo_mech *m, double T, double *hms) {          /* cklib.f:3598-3618 */
    fo_ckhml(m, T, hms);
    for (int k = 0; k < m->kk; ++k) hms[k] = hms[k] / m->wt[k];
}

void fo_ckytx(const fo_mech *m, const double *Y, double *X) {     /* cklib.f:9626-9632 */
    double s = 0.0;
    for (int k = 0; k < m->kk; ++k) s = s + Y[k] / m->wt[k];
    for (int k = 0; k < m->kk; ++k) X[k] = Y[k] / (s * m->wt[k]);
}

void fo_ckxty(const fo_mech *m, const double *X, double *Y) {     /* cklib.f:9461-9468 */
    double s = 0.0;
    for (int k = 0; k < m->kk; ++k) s = s + X[k] * m->wt[k];
    for (int k = 0; k < m->kk; ++k) Y[k] = X[k] * m->wt[k] / s;
}

void fo_adjust_sum(int n, double *v) {                            /* SRC/thermo.f90:136-149 */
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        v[i] = fmax(v[i], 0.0);
        s = s + v[i];
    }
    for (int i = 0; i < n; ++i) v[i] = v[i] / s;
}

static double mix_w(const fo_mech *m, const double *X) {
    double s = 0.0;
    for (int k = 0; k < m->kk; ++k) s = s + X[k] * m->wt[k];
    return s;
}

static double ckubml(const fo_mech *m, double T, const double *X) { /* cklib.f:8346-8351 */
    double u[64], s = 0.0;
    fo_ckuml(m, T, u);
    for (int k = 0; k < m->kk; ++k) s = s + X[k] * u[k];
    return s;
}

static double ckcpbl(const fo_mech *m, double T, const double *X) { /* cklib.f:1746-1751 */
    double c[64], s = 0.0;
    fo_ckcpml(m, T, c);
    for (int k = 0; k < m->kk; ++k) s = s + X[k] * c[k];
    return s;
}

static double ckcvbl(const fo_mech *m, double T, const double *X) { /* cklib.f:2553-2558 */
    double c[64], s = 0.0;
    fo_ckcpml(m, T, c);
    for (int k = 0; k < m->kk; ++k) c[k] = c[k] - m->ru;
    for (int k = 0; k < m->kk; ++k) s = s + X[k] * c[k];
    return s;
}

/* ============================================================================================ */
/* thermo.f90 wrappers                                                                           */
/* ============================================================================================ */
double fo_pressure(const fo_mech *m, double rho, double T, const double *X) {
    double rc = rho * 1.e-3;                                        /* thermo.f90:176-178 */
    double p = rc * m->ru * T / mix_w(m, X);                        /* CKPX cklib.f:5752-5756 */
    return p * 0.1;
}

double fo_density(const fo_mech *m, double p, double T, const double *X) {
    double pc = p * 10.0;                                           /* thermo.f90:161-163 */
    double r = mix_w(m, X) * pc / (m->ru * T);                      /* CKRHOX cklib.f:7197-7202 */
    return r * 1.e3;
}

double fo_specific_energy(const fo_mech *m, double T, const double *X) { /* thermo.f90:369-374 */
    return ckubml(m, T, X) * 1.e-4 / mix_w(m, X);
}

double fo_cp(const fo_mech *m, double T, const double *X) {       /* thermo.f90:335-341 */
    return ckcpbl(m, T, X) * 1.e-4 / mix_w(m, X);
}

double fo_cv(const fo_mech *m, double T, const double *X) {       /* thermo.f90:315-321 */
    return ckcvbl(m, T, X) * 1.e-4 / mix_w(m, X);
}

double fo_sonic_speed(const fo_mech *m, double rho, double p, double T, const double *X) {
    double g = fo_cp(m, T, X) / fo_cv(m, T, X);                     /* thermo.f90:352-353 */
    return sqrt(g * p / rho);                                       /* thermo.f90:418,429 */
}

double fo_temperature_con(const fo_mech *m, double e, const double *X) { /* thermo.f90:211-257 */
    double wtm = mix_w(m, X);
    double spe = e * 1.e4 * wtm;
    double T = 1000.0, f = 1.e30, dT = 1.e30;
    int it = 0;
    while (fabs(f) >= fabs(spe) * 1.e-9 && fabs(dT) >= T * 1.e-6 && it < 10) {
        double ub = ckubml(m, T, X);
        double cv = ckcvbl(m, T, X);
        f = ub - spe;
        dT = -f / cv;
        T = T + dT;
        T = fmax(20.0, fmin(6000.0, T));
        ++it;
    }
    return T;
}

double fo_heat_release_rate(const fo_mech *m, const double *Y0, const double *Y, double rho,
                            double dt) {                            /* thermo.f90:98-101 */
    double rc = rho * 1.e-3, s = 0.0;
    for (int k = 0; k < m->kk; ++k) s = s + rc * m->h298[k] * (Y[k] - Y0[k]) / dt;
    return -s * 1.e-1;
}

double fo_heat_release_rate_omegadot(const fo_mech *m, const double *Y, double rho, double T) {
    double cd[64], dd[64], s = 0.0;                                 /* thermo.f90:128-132 */
    fo_ckcdyr(m, rho * 1.e-3, T, Y, cd, dd);
    for (int k = 0; k < m->kk; ++k) s = s + m->wt[k] * m->h298[k] * (cd[k] - dd[k]);
    return -s * 1.e-1;
}

/* ============================================================================================ */
/* eos.f90                                                                                       */
/* ============================================================================================ */
void fo_qf_to_qc(const fo_mech *m, const double *qf, double *Qc) {
    const int kk = m->kk;
    double y[64], x[64];
    Qc[0] = qf[0];
    Qc[1] = qf[0] * qf[1];
    memcpy(y, qf + 3, sizeof(double) * kk);
    fo_adjust_sum(kk, y);
    for (int k = 0; k < kk; ++k) Qc[3 + k] = qf[0] * y[k];
    fo_ckytx(m, qf + 3, x);                                         /* raw Y, eos.f90:30 */
    fo_adjust_sum(kk, x);
    double e = fo_specific_energy(m, qf[2], x);
    Qc[2] = qf[0] * (e + 0.5 * (qf[1] * qf[1]));
}

void fo_qc_to_qf(const fo_mech *m, const double *Qc, double *qf) {
    const int kk = m->kk;
    double x[64];
    qf[0] = Qc[0];
    qf[1] = Qc[1] / Qc[0];
    for (int k = 0; k < kk; ++k) qf[3 + k] = Qc[3 + k] / Qc[0];
    fo_adjust_sum(kk, qf + 3);
    fo_ckytx(m, qf + 3, x);
    fo_adjust_sum(kk, x);
    double e = Qc[2] / Qc[0] - 0.5 * (qf[1] * qf[1]);
    qf[2] = fo_temperature_con(m, e, x);
}

/* ============================================================================================ */
/* Reconstruction                                                                                */
/* ============================================================================================ */
void fo_weno3z(const double *s, double *sl, double *sr) {         /* SRC/weno.f90:11-65 */
    const double eps = 1.0e-6;
    double b2 = (s[0] - s[1]) * (s[0] - s[1]);
    double b1 = (s[1] - s[2]) * (s[1] - s[2]);
    double tau = fabs(b2 - b1);
    b2 = 1.0 + (tau / (eps + b2)) * (tau / (eps + b2));
    b1 = 1.0 + (tau / (eps + b1)) * (tau / (eps + b1));
    double a2 = 2.0 * b2, a1 = 1.0 * b1, ai = 1.0 / (a2 + a1);
    double v2 = -s[0] + 3.0 * s[1], v1 = s[1] + s[2];
    *sl = 0.5 * ai * (a2 * v2 + a1 * v1);
    b2 = (s[3] - s[2]) * (s[3] - s[2]);
    b1 = (s[2] - s[1]) * (s[2] - s[1]);
    tau = fabs(b2 - b1);
    b2 = 1.0 + (tau / (eps + b2)) * (tau / (eps + b2));
    b1 = 1.0 + (tau / (eps + b1)) * (tau / (eps + b1));
    a2 = 2.0 * b2; a1 = 1.0 * b1; ai = 1.0 / (a2 + a1);
    v2 = 3.0 * s[2] - s[3]; v1 = s[1] + s[2];
    *sr = 0.5 * ai * (a2 * v2 + a1 * v1);
}

/* Fortran x**2 squares first: c*(e)**2 == c*(e*e). */
static double sq(double x) { return x * x; }

static void w5_weights(double *b, int flag) {
    const double eps = 1.0e-6;
    if (flag == 2) {                                                /* weno.f90:97-102 */
        for (int j = 0; j < 3; ++j) b[j] = 1.0 / ((eps + b[j]) * (eps + b[j]));
    } else if (flag == 3) {                                         /* weno.f90:104-111 */
        double tau = fabs(b[2] - b[0]);
        for (int j = 2; j >= 0; --j) b[j] = 1.0 + (tau / (eps + b[j])) * (tau / (eps + b[j]));
    }
}

void fo_weno5(const double *s, double *sl, double *sr, int flag) { /* SRC/weno.f90:68-166 */
    const double c13 = 13.0 / 12.0;
    double b[3];
    b[2] = c13 * sq(s[0] - 2.0 * s[1] + s[2]) + 0.25 * sq(s[0] - 4.0 * s[1] + 3.0 * s[2]);
    b[1] = c13 * sq(s[1] - 2.0 * s[2] + s[3]) + 0.25 * (s[1] - s[3]) * (s[1] - s[3]);
    b[0] = c13 * sq(s[2] - 2.0 * s[3] + s[4]) + 0.25 * sq(3.0 * s[2] - 4.0 * s[3] + s[4]);
    w5_weights(b, flag);
    double a3 = b[2], a2 = 6.0 * b[1], a1 = 3.0 * b[0];
    double ai = 1.0 / (a3 + a2 + a1);
    double v3 = 2.0 * s[0] - 7.0 * s[1] + 11.0 * s[2];
    double v2 = -s[1] + 5.0 * s[2] + 2.0 * s[3];
    double v1 = 2.0 * s[2] + 5.0 * s[3] - s[4];
    *sl = (1.0 / 6.0) * ai * (a3 * v3 + a2 * v2 + a1 * v1);
    b[2] = c13 * sq(s[5] - 2.0 * s[4] + s[3]) + 0.25 * sq(s[5] - 4.0 * s[4] + 3.0 * s[3]);
    b[1] = c13 * sq(s[4] - 2.0 * s[3] + s[2]) + 0.25 * (s[4] - s[2]) * (s[4] - s[2]);
    b[0] = c13 * sq(s[3] - 2.0 * s[2] + s[1]) + 0.25 * sq(3.0 * s[3] - 4.0 * s[2] + s[1]);
    w5_weights(b, flag);
    a3 = b[2]; a2 = 6.0 * b[1]; a1 = 3.0 * b[0];
    ai = 1.0 / (a3 + a2 + a1);
    v3 = 11.0 * s[3] - 7.0 * s[4] + 2.0 * s[5];
    v2 = -s[4] + 5.0 * s[3] + 2.0 * s[2];
    v1 = 2.0 * s[3] + 5.0 * s[2] - s[1];
    *sr = (1.0 / 6.0) * ai * (a3 * v3 + a2 * v2 + a1 * v1);
}

static void w7_side(const double *q, double *out) {
    /* q[0..6] in the orientation of the left-biased formula (weno.f90:184-219). */
    const double eps = 1.0e-6;
    const double d[4] = {1.0 / 35.0, 12.0 / 35.0, 18.0 / 35.0, 4.0 / 35.0};
    double b4 = q[0] * (547.0 * q[0] - 3882.0 * q[1] + 4642.0 * q[2] - 1854.0 * q[3]) +
                q[1] * (7043.0 * q[1] - 17246.0 * q[2] + 7042.0 * q[3]) +
                q[2] * (11003.0 * q[2] - 9402.0 * q[3]) + 2107.0 * (q[3] * q[3]);
    double b3 = q[1] * (267.0 * q[1] - 1642.0 * q[2] + 1602.0 * q[3] - 494.0 * q[4]) +
                q[2] * (2843.0 * q[2] - 5966.0 * q[3] + 1922.0 * q[4]) +
                q[3] * (3443.0 * q[3] - 2522.0 * q[4]) + 547.0 * (q[4] * q[4]);
    double b2 = q[2] * (547.0 * q[2] - 2522.0 * q[3] + 1922.0 * q[4] - 494.0 * q[5]) +
                q[3] * (3443.0 * q[3] - 5966.0 * q[4] + 1602.0 * q[5]) +
                q[4] * (2843.0 * q[4] - 1642.0 * q[5]) + 267.0 * (q[5] * q[5]);
    double b1 = q[3] * (2107.0 * q[3] - 9402.0 * q[4] + 7042.0 * q[5] - 1854.0 * q[6]) +
                q[4] * (11003.0 * q[4] - 17246.0 * q[5] + 4642.0 * q[6]) +
                q[5] * (7043.0 * q[5] - 3882.0 * q[6]) + 547.0 * (q[6] * q[6]);
    double tau = fabs(b4 - b1);
    b4 = 1.0 + (tau / (eps + b4)) * (tau / (eps + b4));
    b3 = 1.0 + (tau / (eps + b3)) * (tau / (eps + b3));
    b2 = 1.0 + (tau / (eps + b2)) * (tau / (eps + b2));
    b1 = 1.0 + (tau / (eps + b1)) * (tau / (eps + b1));
    double a4 = d[0] * b4, a3 = d[1] * b3, a2 = d[2] * b2, a1 = d[3] * b1;
    double ai = 1.0 / (a4 + a3 + a2 + a1);
    double v4 = -3.0 * q[0] + 13.0 * q[1] - 23.0 * q[2] + 25.0 * q[3];
    double v3 = 1.0 * q[1] - 5.0 * q[2] + 13.0 * q[3] + 3.0 * q[4];
    double v2 = -1.0 * q[2] + 7.0 * q[3] + 7.0 * q[4] - 1.0 * q[5];
    double v1 = 3.0 * q[3] + 13.0 * q[4] - 5.0 * q[5] + 1.0 * q[6];
    *out = (1.0 / 12.0) * ai * (a4 * v4 + a3 * v3 + a2 * v2 + a1 * v1);
}

void fo_weno7z(const double *s, double *sl, double *sr) {         /* SRC/weno.f90:169-267 */
    w7_side(s, sl);
    double r[7] = {s[7], s[6], s[5], s[4], s[3], s[2], s[1]};      /* mirrored (weno.f90:222-258) */
    w7_side(r, sr);
}

static double minmod(double x, double y) {                        /* SRC/muscl.f90:79-93 */
    double sx = copysign(1.0, x);
    return sx * fmax(0.0, fmin(fabs(x), sx * y));
}
static double superbee(double x, double y) {                      /* SRC/muscl.f90:96-119 */
    return fmax(0.0, fmax(fmin(x, 2.0 * y), fmin(2.0 * x, y)));
}
static double mc(double x, double y) {                            /* SRC/muscl.f90:122-140 */
    return fmax(0.0, fmin(fmin(2.0 * x, 2.0 * y), 0.5 * (x + y)));
}

void fo_muscl3(const double *s, double *sl, double *sr, int flag) { /* SRC/muscl.f90:9-76 */
    const double k = 1.0 / 3.0;
    const double b = (3.0 - k) / (1.0 - k);
    double dmi = s[1] - s[0], dpi = s[2] - s[1], dmip = s[2] - s[1], dpip = s[3] - s[2];
    if (flag == 1) {
        *sl = s[1] + (1.0 / 4.0) * ((1.0 - k) * dmi + (1.0 + k) * dpi);
        *sr = s[2] - (1.0 / 4.0) * ((1.0 - k) * dpip + (1.0 + k) * dmip);
        return;
    }
    double (*lim)(double, double) = flag == 2 ? minmod : (flag == 3 ? superbee : mc);
    double lpi = lim(dpi, b * dmi), lmi = lim(dmi, b * dpi);
    double lpip = lim(dpip, b * dmip), lmip = lim(dmip, b * dpip);
    *sl = s[1] + (1.0 / 4.0) * ((1.0 - k) * lmi + (1.0 + k) * lpi);
    *sr = s[2] - (1.0 / 4.0) * ((1.0 - k) * lpip + (1.0 + k) * lmip);
}

int fo_lvir(const fo_opts *o) {                                   /* SRC/params.f90:161-173 */
    if (o->flag_acu == 0) return 1;
    if (o->flag_acu == 1) return 2;
    if (o->flag_weno == 1) return 2;
    if (o->flag_weno == 4) return 4;
    return 3;
}

/* ============================================================================================ */
/* Interface flux                                                                                */
/* ============================================================================================ */
/* One interface between cells c and c+1; qf points at cell c's row, rows nequ apart. */
static void interface_flux(const fo_mech *m, const fo_opts *o, const double *qfc, double dx_block,
                           double *F, double *god) {
    const int kk = m->kk, nequ = 3 + kk;
    double qL[67], qR[67], QL[67], QR[67], xL[64], xR[64], fL[67], fR[67], fx[67], vh[67];
#define QF(n, off) qfc[(off) * nequ + (n)]
    for (int n = 0; n < nequ; ++n) {                               /* flux.f90:180-238 */
        if (o->flag_acu == 1) {
            double s[4] = {QF(n, -1), QF(n, 0), QF(n, 1), QF(n, 2)};
            fo_muscl3(s, &qL[n], &qR[n], o->flag_muscl);
        } else if (o->flag_acu == 2 && o->flag_weno == 1) {
            double s[4] = {QF(n, -1), QF(n, 0), QF(n, 1), QF(n, 2)};
            fo_weno3z(s, &qL[n], &qR[n]);
        } else if (o->flag_acu == 2 && (o->flag_weno == 2 || o->flag_weno == 3)) {
            double s[6] = {QF(n, -2), QF(n, -1), QF(n, 0), QF(n, 1), QF(n, 2), QF(n, 3)};
            fo_weno5(s, &qL[n], &qR[n], o->flag_weno);
        } else if (o->flag_acu == 2 && o->flag_weno == 4) {
            double s[8] = {QF(n, -3), QF(n, -2), QF(n, -1), QF(n, 0), QF(n, 1), QF(n, 2), QF(n, 3),
                           QF(n, 3)};                               /* s(8) = qf(i+3), flux.f90:232 */
            fo_weno7z(s, &qL[n], &qR[n]);
        }
    }
    if (o->flag_acu == 0 || o->dx_comp != dx_block) {              /* flux.f90:240-245 */
        for (int n = 0; n < nequ; ++n) { qL[n] = QF(n, 0); qR[n] = QF(n, 1); }
    }
#undef QF
    fo_qf_to_qc(m, qL, QL);
    fo_qf_to_qc(m, qR, QR);
    fo_ckytx(m, qL + 3, xL);
    fo_ckytx(m, qR + 3, xR);
    fo_adjust_sum(kk, xL);
    fo_adjust_sum(kk, xR);
    double rhoL = qL[0], rhoR = qR[0], uL = qL[1], uR = qR[1], tL = qL[2], tR = qR[2];
    double pL = fo_pressure(m, rhoL, tL, xL), pR = fo_pressure(m, rhoR, tR, xR);
    double aL = fo_sonic_speed(m, rhoL, pL, tL, xL), aR = fo_sonic_speed(m, rhoR, pR, tR, xR);
    double eL = QL[2], eR = QR[2];
    fL[0] = rhoL * uL; fR[0] = rhoR * uR;                           /* flux.f90:270-284 */
    if (o->is_spherical == 1) { fL[1] = rhoL * uL * uL; fR[1] = rhoR * uR * uR; }
    else { fL[1] = rhoL * uL * uL + pL; fR[1] = rhoR * uR * uR + pR; }
    fL[2] = (eL + pL) * uL; fR[2] = (eR + pR) * uR;
    for (int k = 0; k < kk; ++k) { fL[3 + k] = rhoL * qL[3 + k] * uL; fR[3 + k] = rhoR * qR[3 + k] * uR; }

    if (o->select_riemann == 0) {                                   /* flux.f90:287-294 */
        double lL = fmax(fmax(fabs(uL - aL), fabs(uL)), fabs(uL + aL));
        double lR = fmax(fmax(fabs(uR - aR), fabs(uR)), fabs(uR + aR));
        double al = fmax(lL, lR);
        for (int n = 0; n < nequ; ++n) fx[n] = 0.5 * (fL[n] + fR[n] - al * (QR[n] - QL[n]));
    } else {
        double rt = sqrt(rhoR / rhoL);
        double ur = (uL + rt * uR) / (1.0 + rt);
        double ar = (aL + rt * aR) / (1.0 + rt);
        double sL = fmin(uL - aL, ur - ar), sR = fmax(uR + aR, ur + ar);
        if (o->select_riemann == 1) {                               /* flux.f90:296-315 */
            if (0.0 <= sL) memcpy(fx, fL, sizeof(double) * nequ);
            else if (sL <= 0.0 && 0.0 <= sR) {
                for (int n = 0; n < nequ; ++n)
                    fx[n] = (sR * fL[n] - sL * fR[n] + sL * sR * (QR[n] - QL[n])) / (sR - sL);
            } else memcpy(fx, fR, sizeof(double) * nequ);
        } else {                                                    /* flux.f90:317-354 */
            double sM = (pL - pR + rhoR * uR * (sR - uR) - rhoL * uL * (sL - uL)) /
                        (rhoR * (sR - uR) - rhoL * (sL - uL));
            if (0.0 <= sL) memcpy(fx, fL, sizeof(double) * nequ);
            else if (sL <= 0.0 && 0.0 <= sM) {
                vh[0] = 1.0; vh[1] = sM;
                vh[2] = eL / rhoL + (sM - uL) * (sM + pL / (rhoL * (sL - uL)));
                for (int k = 0; k < kk; ++k) vh[3 + k] = QL[3 + k] / rhoL;
                double c = rhoL * (sL - uL) / (sL - sM);
                for (int n = 0; n < nequ; ++n) fx[n] = fL[n] + sL * (c * vh[n] - QL[n]);
            } else if (sM <= 0.0 && 0.0 <= sR) {
                vh[0] = 1.0; vh[1] = sM;
                vh[2] = eR / rhoR + (sM - uR) * (sM + pR / (rhoR * (sR - uR)));
                for (int k = 0; k < kk; ++k) vh[3 + k] = QR[3 + k] / rhoR;
                double c = rhoR * (sR - uR) / (sR - sM);
                for (int n = 0; n < nequ; ++n) fx[n] = fR[n] + sR * (c * vh[n] - QR[n]);
            } else memcpy(fx, fR, sizeof(double) * nequ);
        }
    }
    for (int n = 0; n < nequ; ++n) F[n] = F[n] + fx[n];            /* flux.f90:358 */
    double rt = sqrt(rhoR / rhoL);                                  /* flux.f90:360-367 */
    god[0] = sqrt(rhoR * rhoL);
    god[1] = (uL + rt * uR) / (1.0 + rt);
    god[2] = (pL + rt * pR) / (1.0 + rt);
}

void fo_flux_advection(const fo_mech *m, const fo_opts *o, int n, int lv, double dx_block,
                       const double *qf, double *Fplus, double *godunov) {
    const int nequ = 3 + m->kk;
#pragma omp parallel for schedule(static)
    for (int i = 0; i <= n; ++i)  /* interface ista-1+i, between cells (ista-1+i) and (ista+i) */
        interface_flux(m, o, qf + (size_t)(lv - 1 + i) * nequ, dx_block, Fplus + (size_t)i * nequ,
                       godunov + (size_t)i * 3);
}

/* ============================================================================================ */
/* TRANLIB + transport flux                                                                      */
/* ============================================================================================ */
static double horner(const double *c, int no, double x) {
    if (no == 4) return ((c[3] * x + c[2]) * x + c[1]) * x + c[0]; /* MCEVAL4 tranlib.f:1388 */
    double v = c[no - 1];                                           /* MCEVAL  tranlib.f:1335 */
    for (int i = 1; i < no; ++i) v = v * x + c[no - 1 - i];
    return v;
}

double fo_mcavis(const fo_mech *m, double T, const double *X) {   /* tranlib.f:734-762 */
    const int kk = m->kk;
    double vis[64], lt = FO_LOG(T);
    for (int k = 0; k < kk; ++k) vis[k] = FO_EXP(horner(m->cofeta + k * m->no, m->no, lt));
    double so = 0.0;
    for (int k = 0; k < kk; ++k) {
        double si = 0.0;
        for (int j = 0; j < kk; ++j) {
            double a = 1.0 + sqrt(vis[k] / vis[j] * sqrt(m->twt[j] / m->twt[k]));
            si = si + X[j] * (a * a) / sqrt(1.0 + m->twt[k] / m->twt[j]);
        }
        so = so + X[k] * vis[k] / si;
    }
    return so * sqrt(8.0);
}

double fo_mcacon(const fo_mech *m, double T, const double *X) {   /* tranlib.f:876-891 */
    double lt = FO_LOG(T), s = 0.0, sr = 0.0;
    for (int k = 0; k < m->kk; ++k) {
        double c = FO_EXP(horner(m->coflam + k * m->no, m->no, lt));
        s = s + X[k] * c;
        sr = sr + X[k] / c;
    }
    return 0.5 * (s + 1.0 / sr);
}

void fo_mcadif(const fo_mech *m, double p, double T, const double *X, double *D) {
    const int kk = m->kk, no = m->no;                               /* MCEDIF tranlib.f:1095-1148 */
    double lt = FO_LOG(T), xx[64], wtm = 0.0;
    static _Thread_local double djk[64 * 64];
    for (int k = 1; k < kk; ++k)
        for (int j = 0; j < k; ++j) djk[j + 64 * k] = horner(m->cofd + (k * kk + j) * no, no, lt);
    for (int k = 0; k < kk; ++k) {
        for (int j = 0; j < k; ++j) {
            djk[j + 64 * k] = FO_EXP(-djk[j + 64 * k]);
            djk[k + 64 * j] = djk[j + 64 * k];
        }
        djk[k + 64 * k] = 0.0;
    }
    for (int k = 0; k < kk; ++k) {
        wtm = wtm + m->twt[k] * X[k];
        xx[k] = fmax(X[k], 1.0e-50);
    }
    for (int k = 0; k < kk; ++k) {
        double sxw = -xx[k] * m->twt[k], sxod = 0.0;
        for (int j = 0; j < kk; ++j) {
            sxw = sxw + xx[j] * m->twt[j];
            sxod = sxod + xx[j] * djk[j + 64 * k];
        }
        D[k] = sxw / (wtm * sxod);
    }
    for (int k = 0; k < kk; ++k) D[k] = D[k] * m->patmos / p;      /* MCADIF tranlib.f:1022-1024 */
}

void fo_flux_tran(const fo_mech *m, int n, int lv, double dx, const double *Qc, double *Fplus) {
    const int kk = m->kk, nequ = 3 + kk;                            /* SRC/flux.f90:18-127 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i <= n; ++i) {
        const double *qa = Qc + (size_t)(lv - 1 + i) * nequ, *qb = qa + nequ;
        double y[2][64], x[2][64], T[2], p[2], mu[2], kap[2], dm[2][64], dif[2][64], u[2], rho[2];
        for (int s = 0; s < 2; ++s) {
            const double *q = s ? qb : qa;
            rho[s] = q[0];
            u[s] = q[1] / rho[s];
            for (int k = 0; k < kk; ++k) y[s][k] = q[3 + k] / rho[s];
            fo_ckytx(m, y[s], x[s]);
            double et = q[2] / rho[s], ek = 0.5 * (u[s] * u[s]);
            T[s] = fo_temperature_con(m, et - ek, x[s]);
            p[s] = fo_pressure(m, rho[s], T[s], x[s]);
            mu[s] = fo_mcavis(m, T[s], x[s]) * 0.1;                 /* trans.f90:57-59 */
            kap[s] = fo_mcacon(m, T[s], x[s]) * 1.e-5;              /* trans.f90:74-76 */
            fo_mcadif(m, p[s] * 10.0, T[s], x[s], dm[s]);           /* trans.f90:95-99 */
            for (int k = 0; k < kk; ++k) dm[s][k] = dm[s][k] * 1.e-4;
            for (int k = 0; k < kk; ++k) {
                dif[s][k] = (1.0 - x[s][k]) / (1.0 - y[s][k]) * dm[s][k];
                if (y[s][k] == 1.0) dif[s][k] = 0.0;
            }
        }
        double rho_a = 0.5 * (rho[0] + rho[1]), u_a = 0.5 * (u[0] + u[1]);
        double T_a = 0.5 * (T[0] + T[1]), mu_a = 0.5 * (mu[0] + mu[1]);
        double kap_a = 0.5 * (kap[0] + kap[1]);
        double tau = -(2.0 / 3.0 * mu_a * (2.0 * ((u[1] - u[0]) / dx)));
        double hs[64], hsum = 0.0;
        fo_ckhms(m, T_a, hs);
        for (int k = 0; k < kk; ++k) hs[k] = hs[k] * 1.e-4;
        for (int k = 0; k < kk; ++k) {
            double da = 0.5 * (dif[0][k] + dif[1][k]);
            hsum = hsum - hs[k] * da * ((y[1][k] - y[0][k]) / dx);
        }
        double *F = Fplus + (size_t)i * nequ;
        F[0] = F[0] + 0.0;
        F[1] = F[1] + tau;
        F[2] = F[2] + (tau * u_a + rho_a * hsum - kap_a * ((T[1] - T[0]) / dx));
        for (int k = 0; k < kk; ++k) {
            double da = 0.5 * (dif[0][k] + dif[1][k]);
            F[3 + k] = F[3 + k] - rho_a * da * ((y[1][k] - y[0][k]) / dx);
        }
    }
}

/* ============================================================================================ */
/* Domain: boundary, halo, time integration, CFL                                                 */
/* ============================================================================================ */
static double *cellp(fo_domain *d, double *base, int ib, int i) {
    /* i: local cell index, -lv .. ncell-1+lv */
    return base + ((size_t)ib * (d->ncell + 2 * d->lv) + (i + d->lv)) * d->nequ;
}

static void lbound_inlet(fo_domain *d, double *gh, const double *in) { /* SRC/boundary.f90:70-99 */
    const fo_aux *a = &d->aux;
    const double rho_in = in[0];
    const double p_in = fo_pressure(d->m, rho_in, a->T_in, a->x_in);
    (void)p_in;                                  /* specific_energy's p argument, unused by it */
    gh[0] = rho_in;
    gh[1] = a->rho_u_in;
    const double u_in = a->rho_u_in / rho_in;
    const double e_in = fo_specific_energy(d->m, a->T_in, a->x_in);
    gh[2] = rho_in * (e_in + 0.5 * (u_in * u_in));
    for (int k = 0; k < d->m->kk; ++k) gh[3 + k] = rho_in * a->y_in[k];
}

static void rbound_outlet(fo_domain *d, double *gh, const double *in) { /* SRC/boundary.f90:102-149 */
    const int kk = d->m->kk;
    double y[64] = {0.0}, x[64];
    const double u = in[1] / in[0];
    const double e = in[2] / in[0] - 0.5 * (u * u);
    for (int k = 0; k < kk; ++k) y[k] = in[3 + k] / in[0];
    fo_ckytx(d->m, y, x);
    const double T = fo_temperature_con(d->m, e, x);
    const double rho = fo_density(d->m, d->aux.p_out, T, x);
    gh[0] = rho;
    gh[1] = rho * u;
    const double e2 = fo_specific_energy(d->m, T, x);
    gh[2] = rho * (e2 + 0.5 * (u * u));
    for (int k = 0; k < kk; ++k) gh[3 + k] = rho * y[k];
}

static void update_boundary(fo_domain *d) {                       /* SRC/boundary.f90:12-67 */
    const int lv = d->lv, nequ = d->nequ, nl = d->nblk - 1, nc = d->ncell;
    for (int g = 1; g <= lv; ++g) {                                /* lbound_wall :152-174 */
        double *gh = cellp(d, d->Qc, 0, -g), *in = cellp(d, d->Qc, 0, g - 1);
        if (d->aux.lbound == 1) { lbound_inlet(d, gh, in); continue; }
        memcpy(gh, in, sizeof(double) * nequ);
        gh[1] = -in[1];
    }
    for (int g = 1; g <= lv; ++g) fo_qc_to_qf(d->m, cellp(d, d->Qc, 0, -g), cellp(d, d->qf, 0, -g));
    for (int g = 1; g <= lv; ++g) {                                /* rbound_wall :177-199 */
        double *gh = cellp(d, d->Qc, nl, nc - 1 + g), *in = cellp(d, d->Qc, nl, nc - g);
        if (d->aux.rbound == 1) { rbound_outlet(d, gh, in); continue; }
        memcpy(gh, in, sizeof(double) * nequ);
        gh[1] = -in[1];
    }
    for (int g = 1; g <= lv; ++g)
        fo_qc_to_qf(d->m, cellp(d, d->Qc, nl, nc - 1 + g), cellp(d, d->qf, nl, nc - 1 + g));
}

/* minmod slope of the coarse/fine fill (SRC/block.f90:522-525) */
static double smr_slope(double m, double p) {
    return 0.5 * (copysign(1.0, m) + copysign(1.0, p)) * fmin(fabs(m), fabs(p));
}

static void interpolation_block(fo_domain *d) {                   /* SRC/block.f90:475-565 */
    const int lv = d->lv, nc = d->ncell, ne = d->nequ;
    for (int ib = 0; ib < d->nblk - 1; ++ib) {
        const double dl = d->dxb[ib], dr = d->dxb[ib + 1];
        if (dl == dr) {                                            /* :490-497 */
            for (int g = 1; g <= lv; ++g) {
                memcpy(cellp(d, d->Qc, ib + 1, -g), cellp(d, d->Qc, ib, nc - g), sizeof(double) * ne);
                memcpy(cellp(d, d->Qc, ib, nc - 1 + g), cellp(d, d->Qc, ib + 1, g - 1), sizeof(double) * ne);
            }
        } else if (dl > dr) {                                      /* coarse | fine, :499-531 */
            for (int g = 1; g <= lv; ++g) {
                double *gh = cellp(d, d->Qc, ib, nc - 1 + g);
                const double *a = cellp(d, d->Qc, ib + 1, 2 * (g - 1)), *b = cellp(d, d->Qc, ib + 1, 2 * (g - 1) + 1);
                for (int n = 0; n < ne; ++n) gh[n] = (a[n] + b[n]) / 2.0;
            }
            for (int g = 1; g <= lv; ++g) {
                const int ic = -((g - 1) / 2);
                const double *c0 = cellp(d, d->Qc, ib, nc - 1 + ic), *cm = cellp(d, d->Qc, ib, nc - 2 + ic),
                             *cp = cellp(d, d->Qc, ib, nc + ic);
                double *gh = cellp(d, d->Qc, ib + 1, -g);
                for (int n = 0; n < ne; ++n) {
                    const double s = smr_slope(c0[n] - cm[n], cp[n] - c0[n]);
                    gh[n] = (g % 2 == 1) ? c0[n] + 0.25 * s : c0[n] - 0.25 * s;
                }
            }
        } else {                                                   /* fine | coarse, :533-563 */
            for (int g = 1; g <= lv; ++g) {
                double *gh = cellp(d, d->Qc, ib + 1, -g);
                const double *a = cellp(d, d->Qc, ib, nc - 1 - 2 * (g - 1)), *b = cellp(d, d->Qc, ib, nc - 1 - (2 * (g - 1) + 1));
                for (int n = 0; n < ne; ++n) gh[n] = (a[n] + b[n]) / 2.0;
            }
            for (int g = 1; g <= lv; ++g) {
                const int ic = (g - 1) / 2;
                const double *c0 = cellp(d, d->Qc, ib + 1, ic), *cm = cellp(d, d->Qc, ib + 1, ic - 1),
                             *cp = cellp(d, d->Qc, ib + 1, ic + 1);
                double *gh = cellp(d, d->Qc, ib, nc - 1 + g);
                for (int n = 0; n < ne; ++n) {
                    const double s = smr_slope(c0[n] - cm[n], cp[n] - c0[n]);
                    gh[n] = (g % 2 == 1) ? c0[n] - 0.25 * s : c0[n] + 0.25 * s;
                }
            }
        }
    }
}

static void qf_all(fo_domain *d) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int ib = 0; ib < d->nblk; ++ib)
        for (int i = -d->lv; i < d->ncell + d->lv; ++i)
            fo_qc_to_qf(d->m, cellp(d, d->Qc, ib, i), cellp(d, d->qf, ib, i));
}

void fo_refresh(fo_domain *d) {                                   /* SRC/main.f90:149-163 */
    update_boundary(d);
    interpolation_block(d);
    update_boundary(d);
    qf_all(d);
}

/* start-up ghost fill (SRC/main.f90:96-114): boundaries, block halos, QcToqf on the ghosts */
static void init_ghosts(fo_domain *d) {
    update_boundary(d);
    interpolation_block(d);
    update_boundary(d);
    for (int ib = 0; ib < d->nblk; ++ib)
        for (int g = 1; g <= d->lv; ++g) {
            fo_qc_to_qf(d->m, cellp(d, d->Qc, ib, -g), cellp(d, d->qf, ib, -g));
            fo_qc_to_qf(d->m, cellp(d, d->Qc, ib, d->ncell - 1 + g), cellp(d, d->qf, ib, d->ncell - 1 + g));
        }
}

fo_domain *fo_domain_create(const fo_mech *m, const fo_opts *o, int nblk, int ncell, double dx,
                            const double *cells) {
    fo_domain *d = (fo_domain *)calloc(1, sizeof(fo_domain));
    const int kk = m->kk;
    d->m = m; d->o = *o; d->nblk = nblk; d->ncell = ncell; d->lv = fo_lvir(o); d->nequ = 3 + kk;
    d->dx = dx;
    size_t ntot = (size_t)nblk * (ncell + 2 * d->lv) * d->nequ;
    size_t nint = (size_t)nblk * ncell * d->nequ;
    d->Qc = calloc(ntot, sizeof(double)); d->qf = calloc(ntot, sizeof(double));
    d->Fplus = calloc((size_t)nblk * (ncell + 1) * d->nequ, sizeof(double));
    d->godunov = calloc((size_t)nblk * (ncell + 1) * 3, sizeof(double));
    d->Qc0 = calloc(nint, sizeof(double));
    d->k1 = calloc(nint, sizeof(double)); d->k2 = calloc(nint, sizeof(double));
    d->k3 = calloc(nint, sizeof(double)); d->k4 = calloc(nint, sizeof(double));
    d->hmem = calloc((size_t)nblk * ncell, sizeof(double));
    d->cstats = calloc((size_t)nblk * ncell * 4, sizeof(int));
    d->dxb = calloc((size_t)nblk, sizeof(double));
    for (int ib = 0; ib < nblk; ++ib) d->dxb[ib] = dx;
    for (int ib = 0; ib < nblk; ++ib)
        for (int i = 0; i < ncell; ++i) {
            const double *c = cells + ((size_t)ib * ncell + i) * (3 + kk);
            double x[64], y[64];
            memcpy(x, c + 3, sizeof(double) * kk);
            fo_adjust_sum(kk, x);                                   /* params.f90: adjust_sum */
            fo_ckxty(m, x, y);
            double rho = fo_density(m, c[1], c[0], x);
            double e = fo_specific_energy(m, c[0], x);
            double *q = cellp(d, d->qf, ib, i), *Q = cellp(d, d->Qc, ib, i);
            q[0] = rho; q[1] = c[2]; q[2] = c[0];
            Q[0] = rho; Q[1] = rho * c[2]; Q[2] = rho * (e + 0.5 * (c[2] * c[2]));
            for (int k = 0; k < kk; ++k) { q[3 + k] = y[k]; Q[3 + k] = rho * y[k]; }
        }
    /* cell centres of the uniform block layout (SRC/block.f90:264-305 without refinement): a
     * running sum of dx from 0, shifted by the half cell length_smr = -dx/2 */
    d->xm = calloc((size_t)nblk * ncell, sizeof(double));
    d->qrad = calloc((size_t)nblk * ncell, sizeof(double));
    d->qw = calloc((size_t)nblk * ncell, sizeof(double));
    double xl = 0.0;
    const double length_smr = 0.0 - dx / 2.0;
    for (size_t g = 0; g < (size_t)nblk * ncell; ++g) {
        if (g > 0) xl = xl + dx;
        d->xm[g] = xl - length_smr;
    }
    init_ghosts(d);
    return d;
}

fo_domain *fo_domain_create_smr(const fo_mech *m, const fo_opts *o, int nblk_comp, int ncell,
                                double dx_comp, int lev_smr_0, int num_smr_0, int lev_smr_n,
                                int num_smr_n, const double *cells) {
    if (lev_smr_0 < 1 || lev_smr_n < 1 || num_smr_0 < 0 || num_smr_n < 0 || nblk_comp < 1) return NULL;
    const int n0 = (lev_smr_0 - 1) * num_smr_0, nn = (lev_smr_n - 1) * num_smr_n;
    const int nblk = nblk_comp + n0 + nn;                          /* SRC/block.f90:136-146 */
    fo_domain *d = fo_domain_create(m, o, nblk, ncell, dx_comp, cells);
    /* block spacings (:160-218): the left refinement blocks coarsest first, the right ones
     * coarsest last, each level num_smr blocks, doubling per level */
    double length_smr_0 = 0.0;
    int b = 0;
    for (int lev = 2; lev <= lev_smr_0; ++lev)
        for (int k = 0; k < num_smr_0; ++k, ++b) {
            d->dxb[b] = dx_comp * (double)(1 << (lev_smr_0 - lev + 1));
            length_smr_0 = length_smr_0 + d->dxb[b] * ncell;
        }
    b = 0;
    for (int lev = 2; lev <= lev_smr_n; ++lev)
        for (int k = 0; k < num_smr_n; ++k, ++b)
            d->dxb[nblk - 1 - b] = dx_comp * (double)(1 << (lev_smr_n - lev + 1));
    /* cell centres (:264-305): a running position that steps by each cell's dx, half a coarse
     * cell further on a coarse -> fine step and a quarter back on fine -> coarse, then shifted so
     * that the left refinement region ends at x = 0 */
    double xl = 0.0, dxl = 0.0;
    for (int ib = 0; ib < nblk; ++ib)
        for (int i = 0; i < ncell; ++i) {
            const double h = d->dxb[ib];
            double *x = d->xm + (size_t)ib * ncell + i;
            if (ib == 0 && i == 0) {
                *x = 0.0; xl = 0.0; dxl = 0.0;
            } else {
                *x = xl + h;
                xl = xl + h;
                if (dxl > h) { *x = *x + h / 2.0; xl = xl + h / 2.0; }
                if (dxl < h) { *x = *x - h / 4.0; xl = xl - h / 4.0; }
            }
            dxl = h;
        }
    const double length_smr = length_smr_0 - d->dxb[0] / 2.0;
    for (size_t g = 0; g < (size_t)nblk * ncell; ++g) d->xm[g] = d->xm[g] - length_smr;
    init_ghosts(d);
    return d;
}

void fo_domain_free(fo_domain *d) {
    if (!d) return;
    free(d->Qc); free(d->qf); free(d->Fplus); free(d->godunov);
    free(d->Qc0); free(d->k1); free(d->k2); free(d->k3); free(d->k4); free(d->hmem); free(d->cstats);
    free(d->xm); free(d->qrad); free(d->qw); free(d->dxb);
    free(d);
}

double fo_calc_dt(fo_domain *d, double cfl) {                     /* SRC/update.f90:13-59 */
    double cvel = 0.0;
#pragma omp parallel for collapse(2) reduction(max : cvel) schedule(static)
    for (int ib = 0; ib < d->nblk; ++ib)
        for (int i = 0; i < d->ncell; ++i) {
            double x[64];
            const double *q = cellp(d, d->qf, ib, i);
            fo_ckytx(d->m, q + 3, x);
            double p = fo_pressure(d->m, q[0], q[2], x);
            double a = fo_sonic_speed(d->m, q[0], p, q[2], x);
            double au = fmax(fmax(fabs(q[1] - a), fabs(q[1])), fabs(q[1] + a));
            if (au > cvel) cvel = au;
        }
    return (cfl * d->dx) / cvel;
}

static void compute_fluxes(fo_domain *d) {
    const int nequ = d->nequ;
    memset(d->Fplus, 0, sizeof(double) * d->nblk * (d->ncell + 1) * nequ);
    for (int ib = 0; ib < d->nblk; ++ib) {
        double *F = d->Fplus + (size_t)ib * (d->ncell + 1) * nequ;
        if (d->o.do_tran == 1)
            fo_flux_tran(d->m, d->ncell, d->lv, d->dxb[ib], cellp(d, d->Qc, ib, -d->lv), F);
        if (d->o.do_hydro == 1)
            fo_flux_advection(d->m, &d->o, d->ncell, d->lv, d->dxb[ib], cellp(d, d->qf, ib, -d->lv), F,
                              d->godunov + (size_t)ib * (d->ncell + 1) * 3);
    }
}

static double *fint(fo_domain *d, int ib, int j) {   /* interface ista-1+j of block ib */
    return d->Fplus + ((size_t)ib * (d->ncell + 1) + j) * d->nequ;
}

static void stage_sync(fo_domain *d) {
    update_boundary(d);
    interpolation_block(d);
    update_boundary(d);
    qf_all(d);
}

/* Spherical update of one interior cell (SRC/timeintegration.f90:58-72 and 177-196): face areas
 * (r -+ dx/2)^2 and volume r^2 dx at the cell centre r = x_m; the pressure gradient from the Roe
 * pressures of the two interfaces (godunov_state(iqpres)), which the spherical momentum flux
 * leaves out (SRC/flux.f90:272-278), subtracted after the combination with its full dt/dx. */
static void sph_update(fo_domain *d, int ib, int i, double dt, double h1, double h2, const double *Q0,
                       double *Q) {
    const int nequ = d->nequ, nc = d->ncell;
    const double dx = d->dxb[ib];
    const double rm = d->xm[(size_t)ib * nc + i];
    const double Al = (rm - (dx / 2.0)) * (rm - (dx / 2.0));
    const double Ar = (rm + (dx / 2.0)) * (rm + (dx / 2.0));
    const double dV = rm * rm * dx;
    const double *Fr = fint(d, ib, i + 1), *Fl = fint(d, ib, i);
    const double *Gr = d->godunov + ((size_t)ib * (nc + 1) + i + 1) * 3;
    const double *Gl = d->godunov + ((size_t)ib * (nc + 1) + i) * 3;
    for (int n = 0; n < nequ; ++n) {
        const double v = Q[n] - (dt / dV) * (Ar * Fr[n] - Al * Fl[n]);
        Q[n] = Q0 ? h1 * Q0[n] + h2 * v : v;
    }
    Q[1] = Q[1] - (dt / dx) * (Gr[2] - Gl[2]);
}

void fo_solve_fluid(fo_domain *d, double dt) {
    const int nequ = d->nequ, nc = d->ncell;
    const int sph = d->o.is_spherical == 1;
    if (d->o.select_time == 0) {                                   /* explicit_euler :12-96 */
        compute_fluxes(d);
        for (int ib = 0; ib < d->nblk; ++ib)
            for (int i = 0; i < nc; ++i) {
                double *Q = cellp(d, d->Qc, ib, i), *Fr = fint(d, ib, i + 1), *Fl = fint(d, ib, i);
                const double dx = d->dxb[ib];
                if (sph) { sph_update(d, ib, i, dt, 1.0, 1.0, NULL, Q); continue; }
                for (int n = 0; n < nequ; ++n) Q[n] = Q[n] - (dt / dx) * (Fr[n] - Fl[n]);
            }
        stage_sync(d);
    } else if (d->o.select_time == 1) {                            /* TVDRK3 :99-219 */
        for (int ib = 0; ib < d->nblk; ++ib)
            for (int i = 0; i < nc; ++i)
                memcpy(d->Qc0 + ((size_t)ib * nc + i) * nequ, cellp(d, d->Qc, ib, i), sizeof(double) * nequ);
        for (int st = 1; st <= 3; ++st) {
            compute_fluxes(d);
            double h1, h2;
            if (st == 1) {
                h1 = 1.0; h2 = 1.0;
                memset(d->Qc, 0, sizeof(double) * d->nblk * (nc + 2 * d->lv) * nequ);
            } else if (st == 2) { h1 = 3.0 / 4.0; h2 = 1.0 / 4.0; }
            else { h1 = 1.0 / 3.0; h2 = 2.0 / 3.0; }
#pragma omp parallel for collapse(2) schedule(static)
            for (int ib = 0; ib < d->nblk; ++ib)
                for (int i = 0; i < nc; ++i) {
                    double *Q = cellp(d, d->Qc, ib, i), *Fr = fint(d, ib, i + 1), *Fl = fint(d, ib, i);
                    const double *Q0 = d->Qc0 + ((size_t)ib * nc + i) * nequ;
                    const double dx = d->dxb[ib];
                    if (sph) { sph_update(d, ib, i, dt, h1, h2, Q0, Q); continue; }
                    for (int n = 0; n < nequ; ++n)
                        Q[n] = h1 * Q0[n] + h2 * (Q[n] - (dt / dx) * (Fr[n] - Fl[n]));
                }
            stage_sync(d);
        }
    } else {                                                       /* RK4 :222-356 */
        double *ks[4] = {d->k1, d->k2, d->k3, d->k4};
        for (int ib = 0; ib < d->nblk; ++ib)
            for (int i = 0; i < nc; ++i)
                memcpy(d->Qc0 + ((size_t)ib * nc + i) * nequ, cellp(d, d->Qc, ib, i), sizeof(double) * nequ);
        for (int st = 0; st < 4; ++st) {
            compute_fluxes(d);
            for (int ib = 0; ib < d->nblk; ++ib)
                for (int i = 0; i < nc; ++i) {
                    double *Q = cellp(d, d->Qc, ib, i), *Fr = fint(d, ib, i + 1), *Fl = fint(d, ib, i);
                    double *k = ks[st] + ((size_t)ib * nc + i) * nequ;
                    const double dx = d->dxb[ib];
                    for (int n = 0; n < nequ; ++n) {
                        k[n] = -(dt / dx) * (Fr[n] - Fl[n]);
                        if (st < 2) Q[n] = Q[n] - 0.5 * (0.5 * dt / dx) * (Fr[n] - Fl[n]);
                        else if (st == 2) Q[n] = Q[n] - (dt / dx) * (Fr[n] - Fl[n]);
                    }
                }
            if (st < 3) update_boundary(d);
        }
        for (int ib = 0; ib < d->nblk; ++ib)
            for (int i = 0; i < nc; ++i) {
                double *Q = cellp(d, d->Qc, ib, i);
                size_t o = ((size_t)ib * nc + i) * nequ;
                for (int n = 0; n < nequ; ++n)
                    Q[n] = d->Qc0[o + n] +
                           (1.0 / 6.0) * (d->k1[o + n] + 2.0 * d->k2[o + n] + 2.0 * d->k3[o + n] + d->k4[o + n]);
            }
        update_boundary(d);
        qf_all(d);
    }
}

/* ============================================================================================ */
/* CKLIB kinetics                                                                                */
/* ============================================================================================ */
/* CKRATT + CKYTCR + CKRATX (cklib.f:6630-6993, 9533-9583): forward/reverse rates of progress. */
static void rop_k(const fo_mech *m, double rho, double T, const double *Y, double *qf, double *qr,
                  double *kf_eff, double *kr_eff) {
    const int kk = m->kk, ii = m->ii, mxsp = m->mxsp, np1 = m->npar + 1;
    double rkft[512], rkrt[512], eqk[512], smh[64], c[64], ctb[512];
    double alogt = FO_LOG(T);
    for (int i = 0; i < ii; ++i) {                                 /* :6664-6666 */
        const double *p = m->par + i * np1;
        rkft[i] = p[0] * FO_EXP(p[1] * alogt - p[2] / T);
    }
    fo_cksmh(m, T, smh);
    for (int i = 0; i < ii; ++i) {                                 /* :6677-6689 */
        double s = 0.0;
        for (int n = 0; n < mxsp; ++n) {
            int k = m->nk[i * mxsp + n];
            if (k != 0) s = s + (double)m->nu[i * mxsp + n] * smh[k - 1];
        }
        /* CKRATT leaves EQK untouched when the sum is exactly zero; no GRI-3.0 reaction has that. */
        eqk[i] = s != 0.0 ? FO_EXP(fmin(s, m->exparg)) : 1.0;
    }
    double pfac = m->pa / (m->ru * T);                              /* :6700-6704 */
    for (int i = 0; i < ii; ++i) {
        int nus = 0;
        for (int n = 0; n < 6; ++n) nus += m->nu[i * mxsp + n];
        eqk[i] = eqk[i] * ipow(pfac, nus);
    }
    for (int i = 0; i < ii; ++i) {                                 /* :6712-6718 */
        rkrt[i] = 0.0;
        if (m->ns[i] > 0) rkrt[i] = rkft[i] / fmax(eqk[i], m->small);
    }
    for (int k = 0; k < kk; ++k) c[k] = rho * Y[k] / m->wt[k];     /* CKYTCR */
    for (int i = 0; i < ii; ++i) ctb[i] = 1.0;                     /* CKRATX :6851-6855 */
    if (m->nthb > 0) {                                              /* :6859-6869 */
        double ctot = 0.0;
        for (int k = 0; k < kk; ++k) ctot = ctot + c[k];
        for (int t = 0; t < m->nthb; ++t) {
            int i = m->itb[t] - 1;
            ctb[i] = ctot;
            for (int l = 0; l < m->ntbs[t]; ++l)
                ctb[i] = ctb[i] + (m->aik[t * m->mxtb + l] - 1.0) * c[m->nktb[t * m->mxtb + l] - 1];
        }
    }
    for (int f = 0; f < m->nfal; ++f) {                            /* :6873-6925 */
        const double *fp = m->fpar + f * m->nfar;
        int i = m->ifl[f] - 1;
        double rklow = fp[0] * FO_EXP(fp[1] * alogt - fp[2] / T);
        double pr;
        if (m->kfl[f] == 0) { pr = rklow * ctb[i] / rkft[i]; ctb[i] = 1.0; }
        else pr = rklow * c[m->kfl[f] - 1] / rkft[i];
        double pcor = pr / (1.0 + pr);
        if (m->ifo[f] > 1) {
            double prlog = FO_LOG10(fmax(pr, m->small)), fc;
            if (m->ifo[f] == 2) {
                double xp = 1.0 / (1.0 + prlog * prlog);
                fc = FO_POW(fp[3] * FO_EXP(-fp[4] / T) + FO_EXP(-T / fp[5]), xp) * fp[6] * FO_POW(T, fp[7]);
            } else {
                double fcent = (1.0 - fp[3]) * FO_EXP(-T / fp[4]) + fp[3] * FO_EXP(-T / fp[5]);
                if (m->ifo[f] == 4) fcent = fcent + FO_EXP(-fp[6] / T);
                double fclog = FO_LOG10(fmax(fcent, m->small));
                /* single-precision literals of cklib.f:6915-6917 */
                double xn = 0.75 - (double)1.27f * fclog;
                double cprlog = prlog - ((double)0.4f + (double)0.67f * fclog);
                double q = cprlog / (xn - (double)0.14f * cprlog);
                double flog = fclog / (1.0 + q * q);
                fc = FO_EXP10(flog);                                /* 10**flog */
            }
            pcor = fc * pcor;
        }
        rkft[i] = rkft[i] * pcor;
        rkrt[i] = rkrt[i] * pcor;
    }
    for (int i = 0; i < ii; ++i) {                                 /* :6931-6948 */
        const int *nu = m->nu + i * mxsp, *nk = m->nk + i * mxsp;
        rkft[i] = rkft[i] * ctb[i] * m->par[i * np1 + m->npar];
        rkrt[i] = rkrt[i] * ctb[i] * m->par[i * np1 + m->npar];
        double f = 0.0, r = 0.0;
        if (nu[0] != 0) {
            f = rkft[i] * ipow(c[nk[0] - 1], abs(nu[0]));
            r = rkrt[i] * ipow(c[nk[3] - 1], nu[3]);
            if (nk[1] != 0) {
                f = f * ipow(c[nk[1] - 1], abs(nu[1]));
                if (nk[2] != 0) f = f * ipow(c[nk[2] - 1], abs(nu[2]));
            }
            if (nk[4] != 0) {
                r = r * ipow(c[nk[4] - 1], nu[4]);
                if (nk[5] != 0) r = r * ipow(c[nk[5] - 1], nu[5]);
            }
        }
        qf[i] = f;
        qr[i] = r;
        if (kf_eff) { kf_eff[i] = rkft[i]; kr_eff[i] = rkrt[i]; }
    }
}

static void rop(const fo_mech *m, double rho, double T, const double *Y, double *qf, double *qr) {
    rop_k(m, rho, T, Y, qf, qr, NULL, NULL);
}

void fo_ckwyr(const fo_mech *m, double rho, double T, const double *Y, double *wdot, double *qf_,
              double *qr_) {
    double qf[512], qr[512];
    rop(m, rho, T, Y, qf, qr);
    for (int k = 0; k < m->kk; ++k) wdot[k] = 0.0;
    for (int n = 0; n < m->mxsp; ++n)                               /* :9179-9192 */
        for (int i = 0; i < m->ii; ++i) {
            int k = m->nk[i * m->mxsp + n];
            if (k != 0) wdot[k - 1] = wdot[k - 1] + (qf[i] - qr[i]) * (double)m->nu[i * m->mxsp + n];
        }
    if (qf_) memcpy(qf_, qf, sizeof(double) * m->ii);
    if (qr_) memcpy(qr_, qr, sizeof(double) * m->ii);
}

void fo_ckcdyr(const fo_mech *m, double rho, double T, const double *Y, double *cdot, double *ddot) {
    double qf[512], qr[512];
    rop(m, rho, T, Y, qf, qr);
    for (int k = 0; k < m->kk; ++k) { cdot[k] = 0.0; ddot[k] = 0.0; }
    for (int i = 0; i < m->ii; ++i)                                /* :1501-1517 */
        for (int n = 0; n < 3; ++n) {
            int kr = m->nk[i * m->mxsp + n], kp = m->nk[i * m->mxsp + n + 3];
            if (kr != 0) {
                double nur = (double)abs(m->nu[i * m->mxsp + n]);
                cdot[kr - 1] = cdot[kr - 1] + qr[i] * nur;
                ddot[kr - 1] = ddot[kr - 1] + qf[i] * nur;
            }
            if (kp != 0) {
                double nup = (double)m->nu[i * m->mxsp + n + 3];
                cdot[kp - 1] = cdot[kp - 1] + qf[i] * nup;
                ddot[kp - 1] = ddot[kp - 1] + qr[i] * nup;
            }
        }
}

/* ============================================================================================ */
/* Chemistry source: constant-volume adiabatic stiff integrator (this project's own design; the  */
/* reference calls macks(dt, nspe+1, yiT, rho) from module ODEs, SRC/source.f90:58, whose source */
/* is not part of the reference).  Same algorithm as the HIP kernel k_chem:                      */
/*   state y = (Y_1..Y_KK, T) at fixed rho; f from CKWYR rates;                                  */
/*   3-stage L-stable SDIRK of order 3 (Alexander 1977, gamma = 0.43586652150845900) with an     */
/*   embedded order-2 solution (b^ from the first two stages), simplified Newton on every stage    */
/*   with M = I - h*gamma*J and an LU with partial pivoting, error = M^{-1} h*sum (b-b^)_i k_i;    */
/*   analytic mass-action Jacobian                                                               */
/*   with frozen third-body / fall-off factors and a finite-difference temperature column;       */
/*   T finally corrected so that sum Y_k u_k(T)/W_k keeps its initial value (energy exact).      */
/*   Before any implicit step an explicit Heun step over dt is tried and accepted when its        */
/*   embedded Euler error estimate is within tolerance (quiescent cells skip the Jacobian).       */
/* ============================================================================================ */
#define CH_GAMMA 0.435866521508459
#define CH_A21 0.2820667392457705
#define CH_B1 1.20849664917601
#define CH_B2 -0.6443631706844692
#define CH_E1 0.43586652150845895
#define CH_E2 -0.8717330430169181
#define CH_E3 0.435866521508459
#define CH_KAPPA 0.03
#define CH_MAXIT 7
#define CH_MAXSTEPS 5000

void fo_chem_rhs(const fo_mech *m, double rho, const double *y, double *f) {
    const int kk = m->kk;
    const double T = y[kk];
    double w[64], u[64], cp[64];
    fo_ckwyr(m, rho, T, y, w, NULL, NULL);
    fo_ckuml(m, T, u);
    fo_ckcpml(m, T, cp);
    double cvm = 0.0, su = 0.0;
    for (int k = 0; k < kk; ++k) {
        cvm = cvm + y[k] * (cp[k] - m->ru) / m->wt[k];
        su = su + u[k] * w[k];
        f[k] = w[k] * m->wt[k] / rho;
    }
    f[kk] = -su / (rho * cvm);
}

void fo_chem_jac(const fo_mech *m, double rho, const double *y, double *J) {
    const int kk = m->kk, n = kk + 1, mxsp = m->mxsp;
    const double T = y[kk];
    double qf[512], qr[512], kf[512], kr[512], c[64], u[64], cp[64], f[65], fp[65], yp[65];
    static _Thread_local double JC[64 * 64];
    rop_k(m, rho, T, y, qf, qr, kf, kr);
    for (int k = 0; k < kk; ++k) c[k] = rho * y[k] / m->wt[k];
    memset(JC, 0, sizeof(JC));
    for (int i = 0; i < m->ii; ++i) {
        const int *nu = m->nu + i * mxsp, *nk = m->nk + i * mxsp;
        if (nu[0] == 0) continue;
        double dq[6] = {0, 0, 0, 0, 0, 0};
        for (int s = 0; s < 6; ++s) {
            if (nk[s] == 0) continue;
            const int lo = s < 3 ? 0 : 3;
            const int a = abs(nu[s]);
            double d = (double)a * ipow(c[nk[s] - 1], a - 1);
            for (int t = lo; t < lo + 3; ++t)
                if (t != s && nk[t] != 0) d = d * ipow(c[nk[t] - 1], abs(nu[t]));
            dq[s] = s < 3 ? kf[i] * d : -(kr[i] * d);
        }
        for (int s = 0; s < 6; ++s) {
            if (nk[s] == 0) continue;
            const int col = nk[s] - 1;
            for (int t = 0; t < 6; ++t)
                if (nk[t] != 0) JC[(nk[t] - 1) * 64 + col] += (double)nu[t] * dq[s];
        }
    }
    fo_ckuml(m, T, u);
    fo_ckcpml(m, T, cp);
    fo_chem_rhs(m, rho, y, f);
    double cvm = 0.0;
    for (int k = 0; k < kk; ++k) cvm = cvm + y[k] * (cp[k] - m->ru) / m->wt[k];
    for (int r = 0; r < kk; ++r)
        for (int k = 0; k < kk; ++k) J[r * n + k] = JC[r * 64 + k] * m->wt[r] / m->wt[k];
    for (int k = 0; k < kk; ++k) {
        double s = 0.0;
        for (int j = 0; j < kk; ++j) s = s + u[j] * JC[j * 64 + k];
        s = s * rho / m->wt[k];
        J[kk * n + k] = -s / (rho * cvm) - f[kk] * ((cp[k] - m->ru) / m->wt[k]) / cvm;
    }
    const double dT = T * 1.0e-6;
    memcpy(yp, y, sizeof(double) * n);
    yp[kk] = T + dT;
    fo_chem_rhs(m, rho, yp, fp);
    for (int r = 0; r < n; ++r) J[r * n + kk] = (fp[r] - f[r]) / dT;
}

/* LU of A (n x n, row-major) with partial pivoting, rows not swapped: rank[r] is the elimination
 * step at which row r was the pivot, perm[k] the pivot row of step k (ties go to the lower row),
 * rdiag[k] = 1 / pivot of step k.  The trailing update is one fused multiply-add per entry. */
static int ch_lu(double *A, int n, int *perm, int *rank, double *rdiag) {
    for (int r = 0; r < n; ++r) rank[r] = n;
    for (int k = 0; k < n; ++k) {
        int p = -1;
        double best = -1.0;
        for (int r = 0; r < n; ++r)
            if (rank[r] == n && fabs(A[r * n + k]) > best) { best = fabs(A[r * n + k]); p = r; }
        if (!(best > 0.0)) return -1;
        rank[p] = k;
        perm[k] = p;
        const double piv = A[p * n + k];
        rdiag[k] = 1.0 / piv;
        for (int r = 0; r < n; ++r) {
            if (rank[r] != n) continue;
            const double l = A[r * n + k] / piv;
            A[r * n + k] = l;
            for (int j = k + 1; j < n; ++j) A[r * n + j] = fma(-l, A[p * n + j], A[r * n + j]);
        }
    }
    return 0;
}

/* Solves A x = b in place (b -> x, x in natural order): fused multiply-adds, and the back
 * substitution multiplies by the stored reciprocal pivots. */
static void ch_solve(const double *A, int n, const int *perm, const int *rank, const double *rdiag, double *b) {
    double x[65];
    for (int k = 0; k < n; ++k) {
        const double bp = b[perm[k]];
        for (int r = 0; r < n; ++r)
            if (rank[r] > k) b[r] = fma(-A[r * n + k], bp, b[r]);
    }
    for (int k = n - 1; k >= 0; --k) {
        const int p = perm[k];
        x[k] = b[p] * rdiag[k];
        for (int r = 0; r < n; ++r)
            if (rank[r] < k) b[r] = fma(-A[r * n + k], x[k], b[r]);
    }
    memcpy(b, x, sizeof(double) * n);
}

static double ch_norm(const double *v, const double *w, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = s + (v[i] * w[i]) * (v[i] * w[i]);
    return sqrt(s / n);
}

/* Mixture internal energy per unit mass (erg/g) and cv (erg/(g K)) at T for composition y. */
static void ch_energy(const fo_mech *m, const double *y, double T, double *e, double *cv) {
    double u[64], cp[64];
    fo_ckuml(m, T, u);
    fo_ckcpml(m, T, cp);
    double s = 0.0, c = 0.0;
    for (int k = 0; k < m->kk; ++k) {
        s = s + y[k] * u[k] / m->wt[k];
        c = c + y[k] * (cp[k] - m->ru) / m->wt[k];
    }
    *e = s;
    *cv = c;
}

/* One integration over dt; h_io (may be NULL) carries the step-size proposal from one call to the
 * next for the same cell: in, the first SDIRK step size if positive and below dt (else dt); out, the
 * controller's proposal after the last step, or 0 when an explicit step finished the cell. */
static int chem_step_h(const fo_mech *m, double dt, double rho_si, double *yT, double rtol, double atol,
                       double *h_io, fo_chem_stats *st){
    const int kk = m->kk, n = kk + 1;
    const double rho = rho_si * 1.e-3, g = CH_GAMMA;
    double y[65], w[65], f[65], z[65], base[65], k[3][65], r[65], e[65], ya[64];
    static _Thread_local double J[65 * 65], M[65 * 65];
    int perm[65], rank[65];
    double rdiag[65];
    fo_chem_stats s = {0, 0, 0, 0, 0};
    memcpy(y, yT, sizeof(double) * n);
    double e0, cv0;
    memcpy(ya, y, sizeof(double) * kk);
    fo_adjust_sum(kk, ya);
    ch_energy(m, ya, y[kk], &e0, &cv0);
    double t = 0.0, h = dt;
    if (h_io && *h_io > 0.0 && *h_io < dt) h = *h_io;
    int expl = 0;
    {   /* explicit pre-step: Heun over dt, accepted when the embedded Euler error is in tolerance */
        double f0[65], f1[65], y1[65];
        for (int i = 0; i < kk; ++i) w[i] = 1.0 / (atol + rtol * fabs(y[i]));
        w[kk] = 1.0 / (rtol * fabs(y[kk]));
        fo_chem_rhs(m, rho, y, f0);
        for (int i = 0; i < n; ++i) y1[i] = y[i] + dt * f0[i];
        fo_chem_rhs(m, rho, y1, f1);
        for (int i = 0; i < n; ++i) e[i] = (dt * 0.5) * (f1[i] - f0[i]);
        if (ch_norm(e, w, n) <= 1.0) {
            for (int i = 0; i < n; ++i) y[i] = y[i] + (dt * 0.5) * (f0[i] + f1[i]);
            t = dt;
            s.steps = 1;
            expl = 1;
        }
    }
    if (t < dt) {   /* explicit retry: two Heun steps of dt/2, each within tolerance on its own */
        const double hs = dt * 0.5;
        double yh[65], fa[65], fb[65], yb[65];
        memcpy(yh, y, sizeof(double) * n);
        int ok = 1;
        for (int sub = 0; sub < 2 && ok; ++sub) {
            for (int i = 0; i < kk; ++i) w[i] = 1.0 / (atol + rtol * fabs(yh[i]));
            w[kk] = 1.0 / (rtol * fabs(yh[kk]));
            fo_chem_rhs(m, rho, yh, fa);
            for (int i = 0; i < n; ++i) yb[i] = yh[i] + hs * fa[i];
            fo_chem_rhs(m, rho, yb, fb);
            for (int i = 0; i < n; ++i) e[i] = (hs * 0.5) * (fb[i] - fa[i]);
            if (ch_norm(e, w, n) <= 1.0) {
                for (int i = 0; i < n; ++i) yh[i] = yh[i] + (hs * 0.5) * (fa[i] + fb[i]);
            } else {
                ok = 0;
            }
        }
        if (ok) {
            memcpy(y, yh, sizeof(double) * n);
            t = dt;
            s.steps = 2;
            expl = 1;
        }
    }
    while (t < dt && s.steps + s.rejects < CH_MAXSTEPS) {
        if (h > dt - t) h = dt - t;
        for (int i = 0; i < kk; ++i) w[i] = 1.0 / (atol + rtol * fabs(y[i]));
        w[kk] = 1.0 / (rtol * fabs(y[kk]));
        fo_chem_jac(m, rho, y, J);
        s.jacobians++;
        for (;;) {
            const double hg = h * g;
            for (int i = 0; i < n * n; ++i) M[i] = -hg * J[i];
            for (int i = 0; i < n; ++i) M[i * n + i] = M[i * n + i] + 1.0;
            int ok = ch_lu(M, n, perm, rank, rdiag) == 0;
            /* stage s: z = y + h sum_{j<s} a_sj k_j + h g f(z); predictor z = base + h g k_{s-1} */
            for (int st = 0; st < 3 && ok; ++st) {
                for (int i = 0; i < n; ++i) {
                    double b = y[i];
                    if (st == 1) b = b + h * CH_A21 * k[0][i];
                    if (st == 2) b = b + h * CH_B1 * k[0][i] + h * CH_B2 * k[1][i];
                    base[i] = b;
                    z[i] = st == 0 ? b : b + hg * k[st - 1][i];
                }
                double prev = 0.0;
                int conv = 0;
                for (int it = 0; it < CH_MAXIT; ++it) {
                    fo_chem_rhs(m, rho, z, f);
                    for (int i = 0; i < n; ++i) r[i] = -(z[i] - base[i] - hg * f[i]);
                    ch_solve(M, n, perm, rank, rdiag, r);
                    for (int i = 0; i < n; ++i) z[i] = z[i] + r[i];
                    s.newton++;
                    const double nrm = ch_norm(r, w, n);
                    if (!(nrm == nrm) || (it > 0 && nrm > 0.9 * prev)) break;
                    if (nrm <= CH_KAPPA) { conv = 1; break; }
                    prev = nrm;
                }
                if (!conv) { ok = 0; break; }
                for (int i = 0; i < n; ++i) k[st][i] = (z[i] - base[i]) / hg;
            }
            if (!ok) {
                s.rejects++;
                h = h * 0.25;
                if (s.steps + s.rejects >= CH_MAXSTEPS) break;
                continue;
            }
            for (int i = 0; i < n; ++i) e[i] = h * (CH_E1 * k[0][i] + CH_E2 * k[1][i] + CH_E3 * k[2][i]);
            ch_solve(M, n, perm, rank, rdiag, e);
            const double err = ch_norm(e, w, n);
            const double fac = 0.9 * FO_POW(fmax(err, 1.0e-10), -1.0 / 3.0);
            if (err <= 1.0) {
                memcpy(y, z, sizeof(double) * n);
                t = t + h;
                s.steps++;
                h = h * fmin(4.0, fmax(0.2, fac));
                break;
            }
            s.rejects++;
            h = h * fmax(0.2, fac);
            if (s.steps + s.rejects >= CH_MAXSTEPS) break;
        }
    }
    if (t < dt) s.failed = 1;
    if (h_io) *h_io = expl ? 0.0 : h;
    /* energy correction: sum Y u(T)/W with the renormalised composition equals its initial value */
    memcpy(ya, y, sizeof(double) * kk);
    fo_adjust_sum(kk, ya);
    double T = y[kk];
    for (int it = 0; it < 20; ++it) {
        double ek, cvk;
        ch_energy(m, ya, T, &ek, &cvk);
        const double d = (e0 - ek) / cvk;
        T = T + d;
        if (fabs(d) <= 1.0e-12 * T) break;
    }
    y[kk] = T;
    memcpy(yT, y, sizeof(double) * n);
    if (st) *st = s;
    return s.failed ? -1 : 0;
}

/* reaction_macks (SRC/source.f90:17-84) on every interior cell of a domain; hrr / hrr_dot hold
 * one value per interior cell (block-major), as eos_obj(i)%HRR / %HRR_dot. */
int fo_chem_step(const fo_mech *m, double dt, double rho_si, double *yT, double rtol, double atol,
                 fo_chem_stats *st) {
    return chem_step_h(m, dt, rho_si, yT, rtol, atol, NULL, st);
}

void fo_solve_source(fo_domain *d, double dt, double rtol, double atol, double *hrr,
                     double *hrr_dot, int *nfail) {
    const fo_mech *m = d->m;
    const int kk = m->kk;
    int fails = 0;
#pragma omp parallel for collapse(2) reduction(+ : fails) schedule(dynamic, 16)
    for (int ib = 0; ib < d->nblk; ++ib)
        for (int i = 0; i < d->ncell; ++i) {
            double *q = cellp(d, d->qf, ib, i);
            double y0[64], yT[65];
            memcpy(y0, q + 3, sizeof(double) * kk);
            memcpy(yT, q + 3, sizeof(double) * kk);
            fo_adjust_sum(kk, yT);
            yT[kk] = q[2];
            fo_chem_stats cs;
            if (chem_step_h(m, dt, q[0], yT, rtol, atol, d->hmem + (size_t)ib * d->ncell + i, &cs) != 0) ++fails;
            int *cst = d->cstats + ((size_t)ib * d->ncell + i) * 4;
            cst[0] = cs.steps; cst[1] = cs.rejects; cst[2] = cs.newton; cst[3] = cs.failed;
            fo_adjust_sum(kk, yT);
            q[2] = yT[kk];
            memcpy(q + 3, yT, sizeof(double) * kk);
            const size_t c = (size_t)ib * d->ncell + i;
            if (hrr) hrr[c] = fo_heat_release_rate(m, y0, q + 3, q[0], dt);
            if (hrr_dot) hrr_dot[c] = fo_heat_release_rate_omegadot(m, q + 3, q[0], q[2]);
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int ib = 0; ib < d->nblk; ++ib)
        for (int i = 0; i < d->ncell; ++i) fo_qf_to_qc(m, cellp(d, d->qf, ib, i), cellp(d, d->Qc, ib, i));
    if (nfail) *nfail = fails;
}

/* ============================================================================================ */
/* Auxiliary per-cell sources and open-boundary parameters                                        */
/* ============================================================================================ */
#define FO_PI_IG 3.14159265359             /* SRC/ignitionsource.f90:27 */
#define FO_PI_IG_15 0x1.645f7c63f2ed6p+2   /* FO_PI_IG**1.5, correctly rounded */

double fo_radiation_heat_loss(const fo_aux *a, double p, double T, const double *X) {
    const double sigma = 5.6703744191844294e-08;                   /* SRC/radiation.f90:96 */
    const double T1 = T, T2 = T1 * T1, T3 = T2 * T1, T4 = T3 * T1, T5 = T4 * T1;
    const double T0 = 300.0, T04 = T0 * T0 * T0 * T0;
    double Kp = 0.0;
    for (int r = 0; r < a->nrad; ++r) {                             /* :110-116 */
        const double *c = a->pmac[r];
        const double kpi = c[0] + c[1] * T1 + c[2] * T2 + c[3] * T3 + c[4] * T4 + c[5] * T5;
        Kp = Kp + kpi * X[a->rad_idx[r]] * (p / 101325.0);
    }
    return -(4.0 * Kp * sigma * (T4 - T04));                       /* :119 */
}

double fo_wall_temperature(const fo_aux *a, double x) {          /* SRC/mfr.f90:63-78 */
    const double A = 0.5 * (a->Tw_max - a->Tw_min);
    const double B = 0.5 * (a->Tw_max + a->Tw_min);
    return A * FO_TANH(a->alpha * (x - a->beta)) + B;
}

double fo_wall_heat(const fo_mech *m, const fo_aux *a, double T, const double *X, double x_m) {
    const double lambda = fo_mcacon(m, T, X) * 1.e-5;              /* SRC/trans.f90:74-76 */
    return -(4.0 * lambda * a->Nu * (T - fo_wall_temperature(a, x_m)) / (a->d_m * a->d_m));  /* mfr.f90:57 */
}

double fo_ignition_power(const fo_aux *a, double t, double x_m) { /* SRC/ignitionsource.f90:150-264 */
    double Eadd = 0.0;              /* energy_source zeroes it; a later active source overwrites */
    for (int s = 0; s < a->n_ig; ++s) {
        if (!(a->t0_s[s] < t && t < a->t0_s[s] + a->t_s[s])) continue;
        const double q = x_m / a->r_m[s];
        if (a->ig_type == 1) {                                      /* A. Frendi: :165-167 */
            const double coef = a->E_j[s] / (a->vol[s] * a->t_s[s]);
            const double q3 = q * q * q;
            Eadd = coef * FO_EXP(-(FO_PI_IG / 4.0) * (q3 * q3));
        } else {                                                    /* W. Zhang: :218-219 */
            const double r = a->r_m[s];
            const double coef = a->E_j[s] / (FO_PI_IG_15 * (r * r * r) * a->t_s[s]);
            Eadd = coef * FO_EXP(-(q * q));
        }
    }
    return Eadd;
}

int fo_domain_set_flow_boundary(fo_domain *d, int flag_lbound, int flag_rbound, double T_in,
                                double u_in_cm_s, const double *x_in, double p_out_atm) {
    if ((flag_lbound != 0 && flag_lbound != 1) || (flag_rbound != 0 && flag_rbound != 1)) return -1;
    fo_aux *a = &d->aux;
    const int kk = d->m->kk;
    a->lbound = flag_lbound;
    a->rbound = flag_rbound;
    for (int k = 0; k < kk; ++k) a->x_in[k] = x_in ? x_in[k] : 0.0;
    a->T_in = T_in;
    a->u_in = u_in_cm_s * 1.e-2;                                    /* SRC/flow_bound.f90:50-51 */
    a->p_out = p_out_atm * 101325.0;
    if (flag_lbound == 1) {
        fo_ckxty(d->m, a->x_in, a->y_in);                           /* :53-54 */
        a->rho_u_in = fo_density(d->m, a->p_out, a->T_in, a->x_in) * a->u_in;
    }
    init_ghosts(d);
    return 0;
}

int fo_domain_set_radiation(fo_domain *d, int n, const int *species, const double *pmac) {
    if (n < 0 || n > FO_MAX_RAD) return -1;
    for (int r = 0; r < n; ++r) {
        if (species[r] < 0 || species[r] >= d->m->kk) return -1;
        d->aux.rad_idx[r] = species[r];
        for (int j = 0; j < 6; ++j) d->aux.pmac[r][j] = pmac[6 * r + j];
    }
    d->aux.nrad = n;
    return 0;
}

void fo_domain_set_wall_heat(fo_domain *d, int on, double Nu, double d_mm, double Tw_min,
                             double Tw_max, double alpha, double beta) {
    fo_aux *a = &d->aux;
    a->wall_heat = on ? 1 : 0;
    a->Nu = Nu; a->d_m = d_mm * 1.0e-3;                             /* SRC/mfr.f90:36 */
    a->Tw_min = Tw_min; a->Tw_max = Tw_max; a->alpha = alpha; a->beta = beta;
}

int fo_domain_set_ignition(fo_domain *d, int type, int n, const double *E_mJ, const double *r_um,
                           const double *t_us, const double *t0_us) {
    if (type < 0 || type > 2 || n < 0 || n > FO_MAX_IG || (type != 0 && n == 0)) return -1;
    fo_aux *a = &d->aux;
    for (int s = 0; s < n; ++s) {                                   /* SRC/ignitionsource.f90:57-61 */
        a->E_j[s] = E_mJ[s] * 1.e-3;
        a->r_m[s] = r_um[s] * 1.e-6;
        a->t_s[s] = t_us[s] * 1.e-6;
        a->t0_s[s] = t0_us[s] * 1.e-6;
        a->vol[s] = 4.0 * FO_PI_IG * (a->r_m[s] * a->r_m[s] * a->r_m[s]) / 3.0;
    }
    a->ig_type = type;
    a->n_ig = type ? n : 0;
    return 0;
}

void fo_domain_set_xm(fo_domain *d, const double *xm) {
    memcpy(d->xm, xm, sizeof(double) * d->nblk * d->ncell);
}

static int aux_on(const fo_domain *d) {
    return d->aux.nrad > 0 || d->aux.wall_heat || d->aux.ig_type != 0;
}

void fo_aux_sources(fo_domain *d, double dt, double time) {
    const fo_aux *a = &d->aux;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ib = 0; ib < d->nblk; ++ib)
        for (int i = 0; i < d->ncell; ++i) {
            const size_t n = (size_t)ib * d->ncell + i;
            double *q = cellp(d, d->qf, ib, i), *Q = cellp(d, d->Qc, ib, i);
            double x[64];
            if (a->nrad > 0) {                                      /* SRC/source.f90:85-132 */
                fo_ckytx(d->m, q + 3, x);
                const double p = fo_pressure(d->m, q[0], q[2], x);
                const double Qrad = fo_radiation_heat_loss(a, p, q[2], x);
                d->qrad[n] = Qrad;
                Q[2] = Q[2] + dt * Qrad;
                fo_qc_to_qf(d->m, Q, q);
            }
            if (a->wall_heat) {                                     /* :135-182 */
                fo_ckytx(d->m, q + 3, x);
                const double Qw = fo_wall_heat(d->m, a, q[2], x, d->xm[n]);
                d->qw[n] = Qw;
                Q[2] = Q[2] + dt * Qw;
                fo_qc_to_qf(d->m, Q, q);
            }
            if (a->ig_type != 0) {                                  /* :185-236 */
                const double Eadd = fo_ignition_power(a, time, d->xm[n]);
                Q[2] = Q[2] + dt * Eadd;
                fo_qc_to_qf(d->m, Q, q);
            }
        }
}

void fo_eos_fields(const fo_domain *d, const double *hrr, const double *hrr_dot, int idx0, double *out) {
    const size_t n = (size_t)d->nblk * d->ncell;
    const int kk = d->m->kk;
    for (int ib = 0; ib < d->nblk; ++ib)
        for (int i = 0; i < d->ncell; ++i) {
            const size_t g = (size_t)ib * d->ncell + i;
            const double *q = ((const double *)d->qf) + ((size_t)ib * (d->ncell + 2 * d->lv) + (i + d->lv)) * d->nequ;
            const double *Q = ((const double *)d->Qc) + ((size_t)ib * (d->ncell + 2 * d->lv) + (i + d->lv)) * d->nequ;
            double x[64];
            fo_ckytx(d->m, q + 3, x);                              /* SRC/block.f90:458-466 */
            const double rho = q[0], u = q[1], T = q[2];
            out[0 * n + g] = rho;
            out[1 * n + g] = u;
            out[2 * n + g] = T;
            out[3 * n + g] = fo_pressure(d->m, rho, T, x);
            out[4 * n + g] = Q[2] / rho;
            out[5 * n + g] = 0.5 * (u * u);
            out[6 * n + g] = fo_specific_energy(d->m, T, x);
            out[7 * n + g] = hrr ? hrr[g] : 0.0;
            out[8 * n + g] = hrr_dot ? hrr_dot[g] : 0.0;
            out[9 * n + g] = d->qrad[g];
            out[10 * n + g] = d->aux.wall_heat ? fo_wall_temperature(&d->aux, d->xm[g]) : 0.0;  /* params.f90:495-498 */
            out[11 * n + g] = d->qw[g];
            out[12 * n + g] = d->xm[g];
            out[13 * n + g] = d->dxb[ib];
            out[14 * n + g] = (double)(idx0 + (int)g);
            for (int k = 0; k < kk; ++k) out[(15 + k) * n + g] = x[k];
        }
}

void fo_domain_restart(fo_domain *d, const double *f) {
    const size_t n = (size_t)d->nblk * d->ncell;
    const int kk = d->m->kk;
    for (int ib = 0; ib < d->nblk; ++ib)
        for (int i = 0; i < d->ncell; ++i) {
            const size_t g = (size_t)ib * d->ncell + i;
            double x[64], y[64];
            for (int k = 0; k < kk; ++k) x[k] = f[(4 + k) * n + g];
            fo_ckxty(d->m, x, y);                                   /* SRC/read_hdf5.f90:415-417 */
            const double rho = f[g], u = f[n + g], T = f[2 * n + g], et = f[3 * n + g];
            double *q = cellp(d, d->qf, ib, i), *Q = cellp(d, d->Qc, ib, i);
            q[0] = rho; q[1] = u; q[2] = T;                         /* SRC/block.f90:415-424 */
            Q[0] = rho; Q[1] = rho * u; Q[2] = rho * et;
            for (int k = 0; k < kk; ++k) { q[3 + k] = y[k]; Q[3 + k] = rho * y[k]; }
        }
    init_ghosts(d);
}

/* One time step of the driver loop (SRC/main.f90:130-163) with the chemistry split either the way
 * the reference driver does it (strang = 0, Lie: calc_dt, solve_fluid(dt), solve_source(dt),
 * refresh) or symmetrically (strang = 1: calc_dt, solve_source(dt/2), refresh, solve_fluid(dt),
 * solve_source(dt/2), refresh; the refresh after the first half step fills the ghost cells the fluid
 * stage reads).  react = 0: calc_dt, solve_fluid.  hrr / hrr_dot (may be NULL) receive the
 * diagnostics of the last source call.  Returns dt. */
double fo_step(fo_domain *d, double cfl, int react, int strang, double rtol, double atol, double *hrr,
               double *hrr_dot, int *nfail) {
    const double dt = fo_calc_dt(d, cfl);
    d->time = d->time + dt;                                         /* SRC/main.f90:139 */
    int nf = 0, nf2 = 0;
    if (react && strang) {
        fo_solve_source(d, dt * 0.5, rtol, atol, NULL, NULL, &nf);
        fo_refresh(d);
    }
    fo_solve_fluid(d, dt);
    if (react) fo_solve_source(d, strang ? dt * 0.5 : dt, rtol, atol, hrr, hrr_dot, &nf2);
    if (aux_on(d)) fo_aux_sources(d, dt, d->time);                 /* SRC/update.f90:127-147 */
    if (react || aux_on(d)) fo_refresh(d);
    if (nfail) *nfail = nf + nf2;
    return dt;
}

/* Threads the parallel loops of this library use (OpenMP; 1 without it). */
int fo_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* The shared transcendental functions on n arguments (which: 0 exp, 1 log, 2 log10, 3 pow(x, y), 4 10^x),
 * for the tests that compare them with the C library and with extended-precision values.  In the
 * FO_LIBM_MATH build these are the C library's functions. */
void fo_math_eval(int which, int n, const double *x, const double *y, double *out) {
    for (int i = 0; i < n; ++i) {
        switch (which) {
        case 0: out[i] = FO_EXP(x[i]); break;
        case 1: out[i] = FO_LOG(x[i]); break;
        case 2: out[i] = FO_LOG10(x[i]); break;
        case 4: out[i] = FO_EXP10(x[i]); break;
        case 5: out[i] = FO_TANH(x[i]); break;
        case 6: out[i] = FO_EXPM1(x[i]); break;
        default: out[i] = FO_POW(x[i], y[i]); break;
        }
    }
}

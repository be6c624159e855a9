/*
 * ORACLE (test infrastructure) — the reference's GPU chemistry step on the
 * CPU, restated in C with every float/double rounding where the CUDA source
 * has it:
 *   reaction coefficients   src/reaction_rate.cu:33-212 (case B)
 *   cooling functions       src/cooling_rate.cu:31-137 (case B)
 *   heat - cool             calc_heatcool_rate_dev, src/heatcool_rate.cu:10-74
 *   energy update           advance_heatcool_dev, src/step_heatcool.cu:19-181
 *   safe step               calc_dtime_dev, src/step_reaction.cu:18-196
 *   network + energy        advance_reaction_and_heatcool_dev,
 *                           src/step_reaction.cu:402-733
 *   per-cell kernels        src/chemistry.cu:224-433
 * Flag set: __HELIUM__ on, __HYDROGEN_MOL__ off, __HEATCOOL__ and
 * __CASE_B__ on.  XHYDROGEN/YHELIUM are chemistry.h's 0.755/0.245; the
 * helium number fraction is a parameter.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <omp.h>
#include "oracle.h"

#define KTOEV 8.61623e-5
#define XH 0.755
#define YHE 0.245
#define TINYC 1.0e-31

static double mx_d(double a, double b) { return a > b ? a : b; }
static double mn_d(double a, double b) { return a > b ? b : a; }
static float mx_f(float a, float b) { return a > b ? a : b; }

/* WMOL(c), chemistry.h:69 */
static double wmol_of(const orc_chem *c) { return 1.0 / ((1.0 + c->felec) * XH + 0.25 * YHE); }

/* ------------------------------------------------------------- fits ---- */

static double poly_exp8(const double *a, double lt)
{
    /* a[0] + a[1] lt + a[2] lt^2 + ... + a[8] lt^8, powers by repeated
     * multiplication, summed left to right */
    double pw[9];
    pw[1] = lt;
    for (int i = 2; i <= 8; i++) pw[i] = pw[i - 1] * lt;
    double s = a[0];
    for (int i = 1; i <= 8; i++) s += a[i] * pw[i];
    return exp(s);
}

static double hg_fit(double a, double lam, double p, double b, double q, double r)
{
    return a * pow(lam, p) / pow(1.0 + pow(lam / b, q), r);
}

void orc_chem_k(double t, double *k)
{
    static const double c01[9] = {-32.71396786375, 13.53655609057, -5.739328757388,
                                  1.563154982022, -0.2877056004391, 0.03482559773736999,
                                  -0.00263197617559, 0.0001119543953861, -2.039149852002e-6};
    static const double c03[9] = {-44.09864886561001, 23.91596563469, -10.75323019821,
                                  3.058038757198, -0.5685118909884001, 0.06795391233790001,
                                  -0.005009056101857001, 0.0002067236157507, -3.649161410833e-6};
    static const double c05[9] = {-68.71040990212001, 43.93347632635, -18.48066993568,
                                  4.701626486759002, -0.7692466334492, 0.08113042097303,
                                  -0.005324020628287001, 0.0001975705312221, -3.165581065665e-6};
    double tev = t * KTOEV;
    double lt = log(tev);
    k[0] = poly_exp8(c01, lt);
    k[1] = hg_fit(2.753e-14, 2.0 * 157807.0 / t, 1.5, 2.740, 0.407, 2.242);
    k[2] = tev > 0.8 ? poly_exp8(c03, lt) : 0.0;
    if (t > 500.3)
        k[3] = 1.26e-14 * pow(2.0 * 285335.0 / t, 0.75);
    else
        k[3] = (1.12e-10 - 2.1e-11 * log10(t)) / sqrt(t);
    k[4] = tev > 0.8 ? poly_exp8(c05, lt) : 0.0;
    k[5] = hg_fit(1.3765e-14, 2.0 * 631515.0 / t, 1.5, 2.740, 0.407, 2.242);
}

/* creh2 crehe2 crexihe2 crehe3 cioh1 ciohe1 ciohe2 cexh1 cexhe1 cexhe2 brems compt */
void orc_chem_cooling(double t, float zred, double *o)
{
    double t5 = t * 1.e-5;
    double den = 1.0 + sqrt(t5);
    o[0] = hg_fit(3.435e-30 * t, 2.0 * 157807.0 / t, 1.97, 2.25, 0.376, 3.72);
    o[1] = t < 980.0 ? 1.7e-26 * pow(t, 0.35) : 1.9e-25 * pow(t / 1.0e3, 0.15);
    o[2] = 1.24e-13 / pow(t, 1.5) * exp(-4.7 / t5) * (1.0 + 0.3 * exp(-0.94 / t5));
    o[3] = hg_fit(3.435e-30 * t, 2.0 * (631515.0 / t), 1.97, 2.25, 0.376, 3.72);
    o[4] = 1.27e-21 * sqrt(t) / den * exp(-1.578091 / t5);
    o[5] = 9.38e-22 * sqrt(t) / den * exp(-2.853354 / t5);
    o[6] = 4.95e-22 * sqrt(t) / den * exp(-6.31515 / t5);
    o[7] = 7.5e-19 / den * exp(-1.18348 / t5);
    o[8] = 9.1e-27 * pow(t, -0.1687) / den * exp(-0.13179 / t5);
    o[9] = 5.54e-17 * pow(t, -0.397) / den * exp(-4.73638 / t5);
    double u = 5.5 - log10(t);
    o[10] = 1.42e-27 * (1.1 + 0.34 * exp(-u * u / 3.0)) * sqrt(t);
    double z1 = 1.0 + zred;
    o[11] = 5.406e-36 * (z1 * z1 * z1 * z1) * (t - 2.73 * z1);
}

double orc_chem_heatcool(const orc_chem *c, float zred, double nH, double T, double hf)
{
    double cl[12];
    orc_chem_cooling(T, zred, cl);
    double nHI = nH * c->fHI, nHII = nH * c->fHII;
    double nHe = nH * hf;
    double nHeI = nHe * c->fHeI, nHeII = nHe * c->fHeII, nHeIII = nHe * c->fHeIII;
    double ne = nH * c->felec;
    double heat = (double)c->HeatHI * nHI;
    heat += (double)c->HeatHeI * nHeI + (double)c->HeatHeII * nHeII;
    double cool = (cl[0] * nHII + cl[4] * nHI + cl[7] * nHI) * ne + cl[10] * ne * nHII + cl[11] * ne;
    cool += ((cl[1] + cl[2]) * nHeII + cl[3] * nHeIII) * ne + (cl[5] * nHeI + cl[6] * nHeII) * ne +
            (cl[8] * ne * nHeI + cl[9] * nHeII) * ne + cl[10] * ne * (nHeII + nHeIII);
    return heat - cool;
}

/* ---------------------------------------------------- energy update ---- */

int orc_chem_advance_heatcool(float dens_cell, float *uene, const orc_chem *c,
                              const orc_chem_params *p, float dtime)
{
    double eunit = (p->lunit * p->lunit) / (p->tunit * p->tunit);
    double convfact = p->tunit / eunit * (p->lunit * p->lunit * p->lunit) / p->munit;
    float wmin = 4.0 / (5.0 * XH + 3.0);
    float wmax = 1.0 / (XH + 0.25 * YHE);
    float umin = 10.0 / p->uenetok / wmax;   /* pow(10, 0.5 + 0.5) */
    float umax = 1.0e8 / p->uenetok / wmin;  /* pow(10, 8.5 - 0.5) */
    float a3i = 1.0 / (p->anow * p->anow * p->anow);
    double dens = dens_cell * a3i;
    double uold = *uene, uup = uold, ulo = uold;
    if (uold < umin) {
        uup = ulo = uold = umin;
    } else if (uold > umax) {
        uup = ulo = uold = umax;
    }
    float wmol = wmol_of(c);
    double nH = dens * p->denstonh;
#define HC(TT) (orc_chem_heatcool(c, p->znow, nH, (TT), p->helium_fact) / dens * convfact)
    double hc = HC((*uene) * p->uenetok * wmol);
    if (hc > 0.0) {
        uup = mn_d(uup * 1.0488088481, umax);
        ulo = mx_d(ulo / 1.0488088481, umin);
        hc = HC(uup * p->uenetok * wmol);
        while (uup - uold - hc * dtime < 0.0) {
            uup *= 1.1;
            ulo *= 1.1;
            if (uup > umax) break;
            hc = HC(uup * p->uenetok * wmol);
        }
    } else {
        uup = mn_d(uup * 1.0488088481, umax);
        ulo = mx_d(ulo / 1.0488088481, umin);
        hc = HC(ulo * p->uenetok * wmol);
        while (ulo - uold - hc * dtime > 0.0) {
            uup /= 1.1;
            ulo /= 1.1;
            if (ulo < umin) break;
            hc = HC(ulo * p->uenetok * wmol);
        }
    }
    int it = 0;
    double unew;
    do {
        unew = 0.5 * (uup + ulo);
        hc = HC(0.5 * (uold + unew) * p->uenetok * wmol);
        if (unew - uold - hc * dtime > 0.0)
            uup = unew;
        else
            ulo = unew;
        it++;
    } while ((uup - ulo) > 1.0e-5 * unew && it < 32);
#undef HC
    *uene = unew;
    return it;
}

/* ------------------------------------------------------- network ---- */

typedef struct {
    float nH, nHe, HI, HII, HeI, HeII, HeIII, e;
    float k[6];
} ndens;

static void fill_ndens(ndens *n, float dens, const orc_chem *c, const orc_chem_params *p)
{
    float a3i = 1.0f / (p->anow * p->anow * p->anow);
    n->nH = dens * p->denstonh * a3i;
    n->HI = n->nH * c->fHI;
    n->HII = n->nH * c->fHII;
    n->nHe = n->nH * p->helium_fact;
    n->HeI = n->nHe * c->fHeI;
    n->HeII = n->nHe * c->fHeII;
    n->HeIII = n->nHe * c->fHeIII;
    n->e = n->nH * c->felec;
}

static void fill_k(float *k, float T)
{
    double kd[6];
    orc_chem_k(T, kd);
    for (int i = 0; i < 6; i++) k[i] = kd[i];
}

static float dtime_of(const ndens *n, const orc_chem *c, const float *k)
{
    float gHI = c->GammaHI;
    double gHeI = c->GammaHeI, gHeII = c->GammaHeII;
    double ep = k[0] * n->HI * n->e + gHI * n->HI;
    double em = k[1] * n->HII * n->e;
    ep += k[2] * n->HeI * n->e + k[4] * n->HeII * n->e + gHeI * n->HeI + gHeII * n->HeII;
    em += k[3] * n->HeII * n->e + k[5] * n->HeIII * n->e;
    double hp = k[1] * n->HII * n->e;
    double hm = k[0] * n->HI * n->e + gHI * n->HI;
    double de = n->e / (fabs(ep - em) + 1.e-20f);
    double dh = n->HI / (fabs(hp - hm) + 1.e-20f);
    double dt = 0.1 * de + 0.00001 * dh;
    return mx_d(dt, 1.0);
}

float orc_chem_dtime(float dens, float mesh_uene, const orc_chem *c, const orc_chem_params *p)
{
    ndens n;
    fill_ndens(&n, dens, c, p);
    float T = mesh_uene * p->uenetok * wmol_of(c);
    float k[6];
    fill_k(k, T);
    return dtime_of(&n, c, k);
}

static float backward_euler(float n, float s, float a, double dt)
{
    return (s * dt + n) / (1.0f + a * dt);
}

void orc_chem_advance(float dens, float mesh_uene, float *uene, orc_chem *c,
                      const orc_chem_params *p, float dtime, int *nsub, int *niter)
{
    double tp = dtime * p->tunit;
    ndens n;
    fill_ndens(&n, dens, c, p);
    float nH = n.nH, nHe = n.nHe;
    float gHI = c->GammaHI;
    double gHeI = c->GammaHeI, gHeII = c->GammaHeII;
    float k[6];
    fill_k(k, (float)(mesh_uene * p->uenetok * wmol_of(c)));
    double dts = 0.1 * orc_chem_dtime(dens, mesh_uene, c, p);
    if (tp < dts) dts = tp;
    double el = 0.0;
    int ns = 0, ni = 0;
    float HI = n.HI, HII = n.HII, HeI = n.HeI, HeII = n.HeII, HeIII = n.HeIII, e = n.e;
    /* the GPU's guards: stop on a NaN state or after 1e6 sub-steps */
    while (el < tp && ns < 1000000 && c->felec == c->felec && *uene == *uene) {
        if (el + dts > tp) dts = tp - el;
        el += dts;
        ns++;
        float s, a;
        s = k[1] * HII * e;
        a = k[0] * e + gHI;
        HI = backward_euler(HI, s, a, dts);
        s = k[0] * HI * e + gHI * HI;
        a = k[1] * e;
        HII = backward_euler(HII, s, a, dts);
        s = k[3] * HeII * e;
        a = k[2] * e + gHeI;
        HeI = backward_euler(HeI, s, a, dts);
        s = (k[2] * HeI + k[5] * HeIII) * e + gHeI * HeI;
        a = (k[3] + k[4]) * e + gHeII;
        HeII = backward_euler(HeII, s, a, dts);
        s = k[4] * HeII * e + gHeII * HeII;
        a = k[5] * e;
        HeIII = backward_euler(HeIII, s, a, dts);
        s = k[0] * HI * e + gHI * HI;
        a = k[1] * HII;
        s += (k[2] * HeI * e + k[4] * HeII * e + gHeI * HeI + gHeII * HeII);
        a += (k[3] * HeII + k[5] * HeIII);
        e = backward_euler(e, s, a, dts);

        c->fHI = HI / nH;
        c->fHII = HII / nH;
        float cf = 1.0f / (c->fHI + c->fHII);
        c->fHI *= cf;
        c->fHII *= cf;
        c->fHeI = HeI / (nHe + TINYC);
        c->fHeII = HeII / (nHe + TINYC);
        c->fHeIII = HeIII / (nHe + TINYC);
        cf = 1.0f / (c->fHeI + c->fHeII + c->fHeIII + TINYC);
        c->fHeI *= cf;
        c->fHeII *= cf;
        c->fHeIII *= cf;
        c->felec = c->fHII;
        c->felec += (p->helium_fact * (c->fHeII + 2.0 * c->fHeIII) + TINYC);

        dts /= p->tunit;
        ni += orc_chem_advance_heatcool(dens, uene, c, p, dts);

        fill_ndens(&n, dens, c, p);
        HI = n.HI;
        HII = n.HII;
        HeI = n.HeI;
        HeII = n.HeII;
        HeIII = n.HeIII;
        e = n.e;
        fill_k(k, (float)((*uene) * p->uenetok * wmol_of(c)));
        dts = 0.1 * orc_chem_dtime(dens, mesh_uene, c, p);
    }
    if (nsub) *nsub = ns;
    if (niter) *niter = ni;
}

/* --------------------------------------------------- mesh kernels ---- */

#define MESH_STRIDE 156
#define OFF_DENS 0
#define OFF_ENEG 4
#define OFF_MOM 8
#define OFF_UENE 20
#define OFF_DUENE 24
#define OFF_CHEM 44
#define OFF_PCHEM 96
#define OFF_PUENE 148

static float getf(const unsigned char *m, int off)
{
    float v;
    memcpy(&v, m + off, 4);
    return v;
}
static void putf(unsigned char *m, int off, float v) { memcpy(m + off, &v, 4); }

void orc_chem_step_mesh(unsigned char *mesh, int64_t ncell, const orc_chem_params *p,
                        float dtime, float *diff_chem, float *diff_uene, uint64_t *nsub,
                        uint64_t *niter)
{
    float dc = 0.0f, du = 0.0f;
    uint64_t ns = 0, ni = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : dc, du) reduction(+ : ns, ni)
    for (int64_t i = 0; i < ncell; i++) {
        unsigned char *m = mesh + i * MESH_STRIDE;
        float dens = getf(m, OFF_DENS), mu = getf(m, OFF_UENE), u = getf(m, OFF_PUENE);
        orc_chem c, old;
        memcpy(&c, m + OFF_PCHEM, sizeof c);
        memcpy(&old, m + OFF_CHEM, sizeof old);
        putf(m, OFF_DUENE, 0.0f);
        int a = 0, b = 0;
        orc_chem_advance(dens, mu, &u, &c, p, dtime, &a, &b);
        float dH = fabsf(c.GammaHI - old.GammaHI) / (c.GammaHI + 1.0e-30);
        float de = fabsf(c.felec - old.felec) / (c.felec + 1.0e-30);
        float d1 = mx_f(dH, de);
        float d2 = fabsf(u - mu) / (mx_f(u, mu) + 1.0e-30);
        dc = d1 > dc ? d1 : dc;
        du = d2 > du ? d2 : du;
        putf(m, OFF_UENE, u);
        memcpy(m + OFF_CHEM, &c, sizeof c);
        ns += (uint64_t)a;
        ni += (uint64_t)b;
    }
    *diff_chem = dc;
    *diff_uene = du;
    if (nsub) *nsub = ns;
    if (niter) *niter = ni;
}

float orc_chem_dtime_mesh(const unsigned char *mesh, int64_t ncell, const orc_chem_params *p)
{
    float dmin = 3.402823466e38f;
#pragma omp parallel for reduction(min : dmin)
    for (int64_t i = 0; i < ncell; i++) {
        const unsigned char *m = mesh + i * MESH_STRIDE;
        orc_chem c;
        memcpy(&c, m + OFF_PCHEM, sizeof c);
        float d = orc_chem_dtime(getf(m, OFF_DENS), getf(m, OFF_UENE), &c, p);
        dmin = d < dmin ? d : dmin;
    }
    return dmin;
}

void orc_chem_update_mesh(unsigned char *mesh, int64_t ncell)
{
    for (int64_t i = 0; i < ncell; i++) {
        unsigned char *m = mesh + i * MESH_STRIDE;
        memmove(m + OFF_PCHEM, m + OFF_CHEM, sizeof(orc_chem));
        float u = getf(m, OFF_UENE), d = getf(m, OFF_DENS);
        putf(m, OFF_PUENE, u);
        float mx = getf(m, OFF_MOM), my = getf(m, OFF_MOM + 4), mz = getf(m, OFF_MOM + 8);
        putf(m, OFF_ENEG, u * d + 0.5 * (mx * mx + my * my + mz * mz) / d);
    }
}

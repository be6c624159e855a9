/*
 * ORACLE (test infrastructure) — flat entry points into the reference's own
 * CPU chemistry (src/prim_chem/reaction_rate.c, cooling_rate.c,
 * heatcool_rate.c, step_heatcool.c, step_reaction.c), compiled unmodified
 * from where they lie into oracle/_ref/libref_chem.so together with this
 * file.  The reference's structs come from its own prim_chem headers; these
 * wrappers only fill them from flat arguments.
 *
 * Compiled with -D__HELIUM__ -D__HEATCOOL__ -D__CASE_B__, so HELIUM_FACT,
 * XHYDROGEN and YHELIUM are chemistry.h's (Y = 0.245).
 */
#include <string.h>
#include <stdio.h>
#include "run_param.h"
#include "fluid.h"
#include "chemistry.h"

double k01(double), k02(double), k03(double), k04(double), k05(double), k06(double);
double creh2(double), crehe2(double), crexihe2(double), crehe3(double);
double cioh1(double), ciohe1(double), ciohe2(double), cexh1(double), cexhe1(double);
double cexhe2(double), brems(double), compt(double, float);
int advance_heatcool(struct fluid_mesh *, float *, struct prim_chem *, struct run_param *, float,
                     int *, int *);
double calc_dtime(struct fluid_mesh *, struct prim_chem *, struct run_param *);
void advance_reaction_and_heatcool(struct fluid_mesh *, float *, struct prim_chem *,
                                   struct run_param *, float, int *, int *);

/* units[5] = lunit, munit, tunit, uenetok, denstonh */
static void fill_run(struct run_param *r, const double *units, float anow, float znow)
{
    memset(r, 0, sizeof *r);
    r->lunit = units[0];
    r->munit = units[1];
    r->tunit = units[2];
    r->uenetok = units[3];
    r->denstonh = units[4];
    r->anow = anow;
    r->znow = znow;
}

double ref_chem_helium_fact(void) { return HELIUM_FACT; }

void ref_chem_k(double t, double *k)
{
    k[0] = k01(t);
    k[1] = k02(t);
    k[2] = k03(t);
    k[3] = k04(t);
    k[4] = k05(t);
    k[5] = k06(t);
}

void ref_chem_cooling(double t, float zred, double *o)
{
    o[0] = creh2(t);
    o[1] = crehe2(t);
    o[2] = crexihe2(t);
    o[3] = crehe3(t);
    o[4] = cioh1(t);
    o[5] = ciohe1(t);
    o[6] = ciohe2(t);
    o[7] = cexh1(t);
    o[8] = cexhe1(t);
    o[9] = cexhe2(t);
    o[10] = brems(t);
    o[11] = compt(t, zred);
}

double ref_chem_heatcool(const float *chem13, float zred, double nH, double T)
{
    struct prim_chem c;
    memcpy(&c, chem13, sizeof c);
    return calc_heatcool_rate(&c, zred, nH, T);
}

int ref_chem_advance_heatcool(float dens, float mesh_uene, float *uene, const float *chem13,
                              const double *units, float anow, float znow, float dtime)
{
    struct run_param r;
    struct fluid_mesh m;
    struct prim_chem c;
    fill_run(&r, units, anow, znow);
    memset(&m, 0, sizeof m);
    m.dens = dens;
    m.uene = mesh_uene;
    memcpy(&c, chem13, sizeof c);
    int nrec = 0, niter = 0;
    advance_heatcool(&m, uene, &c, &r, dtime, &nrec, &niter);
    return niter;
}

double ref_chem_dtime(float dens, float mesh_uene, const float *chem13, const double *units,
                      float anow, float znow)
{
    struct run_param r;
    struct fluid_mesh m;
    struct prim_chem c;
    fill_run(&r, units, anow, znow);
    memset(&m, 0, sizeof m);
    m.dens = dens;
    m.uene = mesh_uene;
    memcpy(&c, chem13, sizeof c);
    return calc_dtime(&m, &c, &r);
}

/* The reference's CPU network step (double-precision densities): chem13 in
 * (prev_chem) and out, *uene in (prev_uene) and out. */
void ref_chem_advance(float dens, float mesh_uene, float *uene, float *chem13,
                      const double *units, float anow, float znow, float dtime)
{
    struct run_param r;
    struct fluid_mesh m;
    struct prim_chem c;
    fill_run(&r, units, anow, znow);
    memset(&m, 0, sizeof m);
    m.dens = dens;
    m.uene = mesh_uene;
    memcpy(&c, chem13, sizeof c);
    int nrec = 0, niter = 0;
    advance_reaction_and_heatcool(&m, uene, &c, &r, dtime, &nrec, &niter);
    memcpy(chem13, &c, sizeof c);
}

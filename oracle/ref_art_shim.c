/*
 * ORACLE (test infrastructure) — the reference's own recombination fits of
 * the diffuse-photon source function (src/diffuse_photon/diffuse_chemistry.c,
 * case-A and case-B rates k02, k04, k06), compiled unmodified from where it
 * lies into oracle/_ref/libref_art.so.  The reference includes that file
 * into calc_rmesh_data.c (:7) after its headers; this file does the same
 * with the one macro it needs (K_to_eV, src/prim_chem/chemistry.h) and
 * exports the six fits through one flat function.
 */
#include <math.h>
#include "chemistry.h"
#include "diffuse_chemistry.c"

/* out = k02_A, k02_B, k04_A, k04_B, k06_A, k06_B at temperature t [K] */
void ref_art_recomb(double t, double *out)
{
    out[0] = k02_A(t);
    out[1] = k02_B(t);
    out[2] = k04_A(t);
    out[3] = k04_B(t);
    out[4] = k06_A(t);
    out[5] = k06_B(t);
}

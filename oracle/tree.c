/*
 * ORACLE (test infrastructure) — the source oct-tree and the per-target
 * ray lists, restating src/setup_light_ray_tree.c.
 *
 * The reference stores nodes in a hashed cell list (clist, key_to_adr);
 * that is storage only and does not change which rays are produced or
 * their order, so the oracle keeps the nodes in a plain array with child
 * links.  What is kept exactly: the Morton keys (:488-515), the unstable
 * quicksort (:53-81, ties decide summation order), the node sums in the
 * reference's order (:177-229), the root record (:529-569), the node
 * lengths (:136), the opening test r^2 theta^2 > l^2 in FP32 (:259, :354)
 * and the child visiting order (:366-380).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef struct {
    float cm[3];
    float length;
    double L;
    int n, ifirst;
    int child[8];
    int nchild;
    int spec_row;
} onode;

struct orc_tree {
    int ngrid, nsrc, nnode, cap;
    onode *node;
    double *spec;            /* node spectra, nnode rows */
    int *index;              /* sorted source order */
    uint64_t *key;           /* sorted keys */
    float *src_pos;          /* [3*nsrc] original order */
    double *src_spec;        /* [nsrc*ngrid] original order */
    float theta_sq;
    float base_length;
};

/* kawaiqsort, src/setup_light_ray_tree.c:53-81: the same partition steps,
 * so equal keys end up in the same (unstable) order. */
static void key_sort(int *a, uint64_t *key, int lo, int up)
{
    while (up > lo) {
        int i = lo, j = up;
        int pivot_a = a[lo];
        uint64_t pivot_k = key[lo];
        while (i < j) {
            while (key[j] > pivot_k) j--;
            a[i] = a[j];
            key[i] = key[j];
            while (i < j && key[i] <= pivot_k) i++;
            a[j] = a[i];
            key[j] = key[i];
        }
        a[i] = pivot_a;
        key[i] = pivot_k;
        if (i - lo < up - i) {
            key_sort(a, key, lo, i - 1);
            lo = i + 1;
        } else {
            key_sort(a, key, i + 1, up);
            up = i - 1;
        }
    }
}

static int new_node(orc_tree *t)
{
    if (t->nnode == t->cap) {
        t->cap *= 2;
        t->node = realloc(t->node, sizeof(onode) * t->cap);
        t->spec = realloc(t->spec, sizeof(double) * t->ngrid * (size_t)t->cap);
    }
    int k = t->nnode++;
    memset(&t->node[k], 0, sizeof(onode));
    for (int c = 0; c < 8; c++) t->node[k].child[c] = -1;
    t->node[k].spec_row = k;
    return k;
}

static double spec_sum(const double *s, int ng)
{
    double L = 0.0;
    for (int k = 0; k < ng; k++) L += s[k];
    return L;
}

/* plant_tree, src/setup_light_ray_tree.c:100-239.  Builds the children of
 * `parent` (whose sources are sorted positions [ifirst, ifirst+n)) and
 * returns, through the out-arguments, the subtree's summed luminosity,
 * summed spectrum and sum of L*position, accumulated child by child. */
static void plant(orc_tree *t, int parent, int ifirst, int n, int key_level,
                  double *total_L, double *total_spec, double *cL)
{
    const int ng = t->ngrid;
    key_level -= 3;
    int ncell[8] = {0}, ifirstc[8] = {0};
    int prev = -1;
    for (int i = ifirst; i < ifirst + n; i++) {
        int oct = (int)((t->key[i] >> key_level) & 0x7);
        ncell[oct]++;
        if (oct != prev) ifirstc[oct] = i;
        prev = oct;
    }
    int kid[8];
    for (int c = 0; c < 8; c++) {
        kid[c] = -1;
        if (!ncell[c]) continue;
        int k = new_node(t);
        kid[c] = k;
        t->node[k].n = ncell[c];
        t->node[k].ifirst = ifirstc[c];
        t->node[k].length = t->base_length * pow(2.0, -(63.0 - key_level) / 3.0);
        t->node[parent].child[c] = k;
        t->node[parent].nchild++;
    }

    double cm_acc[3] = {0.0, 0.0, 0.0};
    *total_L = 0.0;
    for (int k = 0; k < ng; k++) total_spec[k] = 0.0;
    double *sub_spec = malloc(sizeof(double) * ng);

    for (int c = 0; c < 8; c++) {
        if (!ncell[c]) continue;
        int k = kid[c];
        if (ncell[c] > 1 && key_level > 0) {           /* nleaf = 1 */
            double sub_L, sub_cm[3];
            plant(t, k, ifirstc[c], ncell[c], key_level, &sub_L, sub_spec, sub_cm);
            *total_L += sub_L;
            for (int q = 0; q < ng; q++) total_spec[q] += sub_spec[q];
            for (int d = 0; d < 3; d++) cm_acc[d] += sub_cm[d];
            t->node[k].L = sub_L;
            memcpy(t->spec + (size_t)k * ng, sub_spec, sizeof(double) * ng);
            for (int d = 0; d < 3; d++) t->node[k].cm[d] = sub_cm[d] / sub_L;
        } else {
            double m_tot = 0.0, m_pos[3] = {0.0, 0.0, 0.0};
            for (int q = 0; q < ng; q++) sub_spec[q] = 0.0;
            for (int j = ifirstc[c]; j < ifirstc[c] + ncell[c]; j++) {
                int s = t->index[j];
                const double *ps = t->src_spec + (size_t)s * ng;
                double m = spec_sum(ps, ng);
                for (int q = 0; q < ng; q++) sub_spec[q] += ps[q];
                m_tot += m;
                for (int d = 0; d < 3; d++) m_pos[d] += m * t->src_pos[3 * s + d];
            }
            memcpy(t->spec + (size_t)k * ng, sub_spec, sizeof(double) * ng);
            t->node[k].L = spec_sum(sub_spec, ng);
            *total_L += m_tot;
            for (int q = 0; q < ng; q++) total_spec[q] += sub_spec[q];
            for (int d = 0; d < 3; d++) cm_acc[d] += m_pos[d];
            double inv = 1.0 / m_tot;
            for (int d = 0; d < 3; d++) t->node[k].cm[d] = m_pos[d] * inv;
        }
    }
    for (int d = 0; d < 3; d++) cL[d] = cm_acc[d];
    free(sub_spec);
}

/* Interleave the low 21 bits of x, y, z: bit b of x at 3b+2, of y at 3b+1,
 * of z at 3b; bit 63 set (src/setup_light_ray_tree.c:492-514). */
static uint64_t morton_key(uint64_t x, uint64_t y, uint64_t z)
{
    uint64_t k = ((uint64_t)1) << 63;
    for (int b = 0; b < 21; b++) {
        uint64_t bit = ((uint64_t)1) << b;
        uint64_t tri = ((x & bit) << 2) | ((y & bit) << 1) | (z & bit);
        k |= tri << (2 * b);
    }
    return k;
}

orc_tree *orc_tree_build(const orc_grid *g, int nsrc, const float *pos,
                         const double *spec)
{
    const int ng = g->ngrid;
    orc_tree *t = calloc(1, sizeof(orc_tree));
    t->ngrid = ng;
    t->nsrc = nsrc;
    t->cap = 64;
    t->node = malloc(sizeof(onode) * t->cap);
    t->spec = malloc(sizeof(double) * ng * (size_t)t->cap);
    t->index = malloc(sizeof(int) * (nsrc > 0 ? nsrc : 1));
    t->key = malloc(sizeof(uint64_t) * (nsrc > 0 ? nsrc : 1));
    t->src_pos = malloc(sizeof(float) * 3 * (nsrc > 0 ? nsrc : 1));
    t->src_spec = malloc(sizeof(double) * ng * (size_t)(nsrc > 0 ? nsrc : 1));
    memcpy(t->src_pos, pos, sizeof(float) * 3 * nsrc);
    memcpy(t->src_spec, spec, sizeof(double) * ng * (size_t)nsrc);
    /* init_tree_param, :573-585 */
    float th = g->theta_crit;
    t->theta_sq = th * th;

    /* construct_tree, :465-571 */
    float xl = g->xmax - g->xmin, yl = g->ymax - g->ymin, zl = g->zmax - g->zmin;
    float bl = yl > zl ? yl : zl;
    t->base_length = xl > bl ? xl : bl;
    double scale = (double)(1 << 21);
    for (int i = 0; i < nsrc; i++) {
        t->index[i] = i;
        uint64_t ip[3];
        ip[0] = (int)((pos[3 * i + 0] - g->xmin) / t->base_length * scale);
        ip[1] = (int)((pos[3 * i + 1] - g->ymin) / t->base_length * scale);
        ip[2] = (int)((pos[3 * i + 2] - g->zmin) / t->base_length * scale);
        t->key[i] = morton_key(ip[0], ip[1], ip[2]);
    }
    if (nsrc > 0) key_sort(t->index, t->key, 0, nsrc - 1);

    int root = new_node(t);
    double xc = 0.0, yc = 0.0, zc = 0.0, lum = 0.0;
    double *all = calloc(ng, sizeof(double));
    for (int s = 0; s < nsrc; s++) {
        const double *ps = spec + (size_t)s * ng;
        xc += pos[3 * s + 0] * ps[0];
        yc += pos[3 * s + 1] * ps[0];
        zc += pos[3 * s + 2] * ps[0];
        lum += ps[0];
        for (int q = 0; q < ng; q++) all[q] += ps[q];
    }
    xc /= lum; yc /= lum; zc /= lum;
    t->node[root].n = nsrc;
    t->node[root].ifirst = 0;
    t->node[root].length = t->base_length;
    t->node[root].cm[0] = xc;
    t->node[root].cm[1] = yc;
    t->node[root].cm[2] = zc;
    t->node[root].L = spec_sum(all, ng);
    memcpy(t->spec + (size_t)root * ng, all, sizeof(double) * ng);
    free(all);
    if (nsrc > 0) {
        double tl, cl[3];
        double *ts = malloc(sizeof(double) * ng);
        plant(t, root, 0, nsrc, 63, &tl, ts, cl);
        free(ts);
    }
    return t;
}

void orc_tree_free(orc_tree *t)
{
    if (!t) return;
    free(t->node); free(t->spec); free(t->index); free(t->key);
    free(t->src_pos); free(t->src_spec); free(t);
}

int orc_tree_nnode(const orc_tree *t) { return t->nnode; }

/* The sources below a node (original indices, in sorted-key order):
 * index[ifirst .. ifirst + n) of the sorted source list, the range the
 * reference's clist record names (clist[].ifirst, clist[].n). */
int orc_tree_node_members(const orc_tree *t, int node, int *out)
{
    const onode *p = &t->node[node];
    for (int j = 0; j < p->n; j++) out[j] = t->index[p->ifirst + j];
    return p->n;
}

void orc_tree_node(const orc_tree *t, int node, float *cm, float *length,
                   int *n, int *nchild, const double **spec)
{
    const onode *p = &t->node[node];
    for (int d = 0; d < 3; d++) cm[d] = p->cm[d];
    *length = p->length;
    *n = p->n;
    *nchild = p->nchild;
    *spec = t->spec + (size_t)p->spec_row * t->ngrid;
}

const double *orc_ray_spec(const orc_tree *t, int kind, int id)
{
    if (kind == 0) return t->spec + (size_t)t->node[id].spec_row * t->ngrid;
    return t->src_spec + (size_t)id * t->ngrid;
}

/* Target cell centre, src/setup_light_ray_tree.c:341-343: the sum is
 * formed in double ((float)i + 0.5 is double) and rounded to float. */
static void cell_centre(const orc_grid *g, int ix, int iy, int iz, float *p)
{
    p[0] = g->xmin + ((float)ix + 0.5) * g->dx;
    p[1] = g->ymin + ((float)iy + 0.5) * g->dy;
    p[2] = g->zmin + ((float)iz + 0.5) * g->dz;
}

typedef struct {
    uint64_t count, cap;
    int *kind, *id;
    float *pos;
} ray_out;

static void emit(ray_out *o, int kind, int id, const float *p)
{
    if (o->count < o->cap) {
        if (o->kind) o->kind[o->count] = kind;
        if (o->id) o->id[o->count] = id;
        if (o->pos) {
            o->pos[3 * o->count + 0] = p[0];
            o->pos[3 * o->count + 1] = p[1];
            o->pos[3 * o->count + 2] = p[2];
        }
    }
    o->count++;
}

/* setup_light_ray_to / count_src, src/setup_light_ray_tree.c:241-283 and
 * :321-395: accept the node when r^2 theta^2 > l^2 (FP32), else open it
 * (n > nleaf: visit the non-empty children in octant order; a node with
 * n > 1 but no children, i.e. coincident keys at the deepest level, then
 * yields nothing) or emit its n <= 1 sources. */
static void visit(const orc_tree *t, int k, const float *tp, ray_out *o)
{
    const onode *p = &t->node[k];
    float dx = tp[0] - p->cm[0];
    float dy = tp[1] - p->cm[1];
    float dz = tp[2] - p->cm[2];
    float r2 = dx * dx + dy * dy + dz * dz;
    if (r2 * t->theta_sq > p->length * p->length) {
        emit(o, 0, k, p->cm);
        return;
    }
    if (p->n > 1) {
        for (int c = 0; c < 8; c++)
            if (p->child[c] >= 0) visit(t, p->child[c], tp, o);
    } else {
        for (int i = p->ifirst; i < p->ifirst + p->n; i++) {
            int s = t->index[i];
            emit(o, 1, s, t->src_pos + 3 * s);
        }
    }
}

uint64_t orc_count_ray_to(const orc_tree *t, const orc_grid *g,
                          int ix, int iy, int iz)
{
    float tp[3];
    cell_centre(g, ix, iy, iz, tp);
    ray_out o = {0, 0, NULL, NULL, NULL};
    visit(t, 0, tp, &o);
    return o.count;
}

uint64_t orc_rays_to(const orc_tree *t, const orc_grid *g, int ix, int iy,
                     int iz, uint64_t cap, int *kind, int *id, float *pos)
{
    float tp[3];
    cell_centre(g, ix, iy, iz, tp);
    ray_out o = {0, cap, kind, id, pos};
    visit(t, 0, tp, &o);
    return o.count;
}

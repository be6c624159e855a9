// Host-side source oct-tree for the fused step (product code).
//
// Builds the tree the reference builds in construct_tree
// (src/setup_light_ray_tree.c:465-571) and lays it out for the GPU in
// depth-first pre-order, children in octant order, which is the order
// setup_light_ray_to (:321-395) visits it.  Numerics follow the reference:
// Morton keys (:488-515), the same unstable quicksort (:53-81, whose tie
// order fixes the summation order), node sums accumulated child by child
// (:177-229), node length base*2^-(63-level)/3 (:136), the root's
// photon_rate[0]-weighted centre (:529-556).  The reference's hashed cell
// list (clist) is storage only and is replaced by the flat arrays.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "tree.h"

namespace argot {
namespace {

struct Build {
  int ngrid;
  int nsrc;
  const float* pos;         // [3*nsrc]
  const double* spec;       // source s's spectrum at spec + s*stride
  size_t stride;
  std::vector<uint64_t> key;
  std::vector<int> order;   // sorted source order
  std::vector<double> scratch;  // [ngrid]: the root's children sum (unused)
  std::vector<double> lum;      // [nsrc]: each source's sum over the bins
  float base_length;
  HostTree* out;
};

// Unstable quicksort of (order, key) by key, with the reference's exact
// partition steps.
void sort_keys(int* a, uint64_t* key, int lo, int up) {
  while (up > lo) {
    int i = lo, j = up;
    const int pa = a[lo];
    const uint64_t pk = key[lo];
    while (i < j) {
      while (key[j] > pk) j--;
      a[i] = a[j];
      key[i] = key[j];
      while (i < j && key[i] <= pk) i++;
      a[j] = a[i];
      key[j] = key[i];
    }
    a[i] = pa;
    key[i] = pk;
    if (i - lo < up - i) {
      sort_keys(a, key, lo, i - 1);
      lo = i + 1;
    } else {
      sort_keys(a, key, i + 1, up);
      up = i - 1;
    }
  }
}

uint64_t interleave(uint64_t x, uint64_t y, uint64_t z) {
  uint64_t k = 1ull << 63;
  for (int b = 0; b < 21; b++) {
    const uint64_t m = 1ull << b;
    k |= (((x & m) << 2) | ((y & m) << 1) | (z & m)) << (2 * b);
  }
  return k;
}

int add_node(HostTree* t) {
  const int id = (int)t->link.size();
  t->link.push_back(NodeLink{-1, 0, -1, -1});
  t->cm.insert(t->cm.end(), {0.0f, 0.0f, 0.0f, 0.0f});
  t->spec64.resize(t->spec64.size() + t->ngrid, 0.0);
  return id;
}

// Nodes below a node holding the sorted keys [first, first + n) at `level`:
// the children plant() creates, counted the same way (so the arrays can be
// reserved once).
int count_nodes(const uint64_t* key, int first, int n, int level) {
  level -= 3;
  int total = 0;
  for (int i = first; i < first + n;) {
    const uint64_t oct = (key[i] >> level) & 7;
    int j = i + 1;
    while (j < first + n && ((key[j] >> level) & 7) == oct) j++;
    total++;
    if (j - i > 1 && level > 0) total += count_nodes(key, i, j - i, level);
    i = j;
  }
  return total;
}

// Spectrum row of node u (u >= 0) or the scratch row (u < 0).  Rows are
// addressed by index: add_node may move the array.
double* spec_row(Build& b, int u) {
  return u >= 0 ? &b.out->spec64[(size_t)u * b.ngrid] : b.scratch.data();
}

struct Sums {
  double L;
  double m[3];
};

// Pre-order: the caller has created node `self`; this creates and fills its
// children (octant order), recursing into each child before creating the
// next, and returns the subtree's luminosity and L-weighted position sums;
// the sum of the children's spectra goes to row `out` (zero on entry).
Sums plant(Build& b, int self, int first, int n, int level, int out) {
  HostTree* t = b.out;
  const int ng = b.ngrid;
  level -= 3;
  int cnt[8] = {0}, beg[8] = {0};
  int prev = -1;
  for (int i = first; i < first + n; i++) {
    const int oct = (int)((b.key[i] >> level) & 7);
    cnt[oct]++;
    if (oct != prev) beg[oct] = i;
    prev = oct;
  }
  Sums tot{0.0, {0.0, 0.0, 0.0}};
  bool first_child = true;
  for (int c = 0; c < 8; c++) {
    if (!cnt[c]) continue;
    const int id = add_node(t);
    if (first_child) {
      t->link[self].first = id;
      first_child = false;
    }
    t->link[id].n = cnt[c];
    const float len = (float)(b.base_length * pow(2.0, -(63.0 - level) / 3.0));
    t->cm[4 * id + 3] = len * len;
    if (cnt[c] > 1 && level > 0) {
      const Sums s = plant(b, id, beg[c], cnt[c], level, id);
      tot.L += s.L;
      const double* sub = spec_row(b, id);
      double* so = spec_row(b, out);
      for (int k = 0; k < ng; k++) so[k] += sub[k];
      for (int d = 0; d < 3; d++) tot.m[d] += s.m[d];
      for (int d = 0; d < 3; d++) t->cm[4 * id + d] = (float)(s.m[d] / s.L);
    } else {
      double mt = 0.0, mp[3] = {0.0, 0.0, 0.0};
      double* sub = spec_row(b, id);
      for (int j = beg[c]; j < beg[c] + cnt[c]; j++) {
        const int s = b.order[j];
        const double* ps = b.spec + (size_t)s * b.stride;
        const double m = b.lum[s];
        for (int k = 0; k < ng; k++) sub[k] += ps[k];
        mt += m;
        for (int d = 0; d < 3; d++) mp[d] += m * b.pos[3 * s + d];
      }
      tot.L += mt;
      double* so = spec_row(b, out);
      for (int k = 0; k < ng; k++) so[k] += sub[k];
      for (int d = 0; d < 3; d++) tot.m[d] += mp[d];
      const double inv = 1.0 / mt;
      for (int d = 0; d < 3; d++) t->cm[4 * id + d] = (float)(mp[d] * inv);
      if (cnt[c] == 1) t->link[id].src = b.order[beg[c]];
    }
    t->link[id].skip = (int)t->link.size();
  }
  return tot;
}

}  // namespace

int build_tree(const float* pos, const double* spec, size_t spec_stride, int nsrc, int ngrid,
               const TreeBox& box, float theta_crit, HostTree* t) {
  if (nsrc < 0 || ngrid <= 0 || ngrid > kNgridMax || spec_stride < (size_t)ngrid) return -1;
  t->ngrid = ngrid;
  t->nsrc = nsrc;
  t->theta_sq = theta_crit * theta_crit;
  t->link.clear();
  t->cm.clear();
  t->spec64.clear();
  t->src_pos.assign(4 * (size_t)std::max(nsrc, 1), 0.0f);
  for (int s = 0; s < nsrc; s++)
    for (int d = 0; d < 3; d++) t->src_pos[4 * s + d] = pos[3 * s + d];

  Build b;
  b.ngrid = ngrid;
  b.nsrc = nsrc;
  b.pos = pos;
  b.spec = spec;
  b.stride = spec_stride;
  b.scratch.assign(ngrid, 0.0);
  b.out = t;
  const float xl = box.xmax - box.xmin, yl = box.ymax - box.ymin, zl = box.zmax - box.zmin;
  const float byz = yl > zl ? yl : zl;
  b.base_length = xl > byz ? xl : byz;
  b.key.resize(std::max(nsrc, 1));
  b.order.resize(std::max(nsrc, 1));
  const double scale = (double)(1 << 21);
  for (int i = 0; i < nsrc; i++) {
    b.order[i] = i;
    const uint64_t ix = (uint64_t)(int)((pos[3 * i + 0] - box.xmin) / b.base_length * scale);
    const uint64_t iy = (uint64_t)(int)((pos[3 * i + 1] - box.ymin) / b.base_length * scale);
    const uint64_t iz = (uint64_t)(int)((pos[3 * i + 2] - box.zmin) / b.base_length * scale);
    b.key[i] = interleave(ix, iy, iz);
  }
  if (nsrc > 0) sort_keys(b.order.data(), b.key.data(), 0, nsrc - 1);
  // Each source's luminosity, its bins summed in order; four sources at a
  // time, so that four independent chains of additions overlap.
  b.lum.assign(std::max(nsrc, 1), 0.0);
  {
    int s = 0;
    for (; s + 4 <= nsrc; s += 4) {
      const double* p0 = spec + (size_t)s * spec_stride;
      const double* p1 = p0 + spec_stride;
      const double* p2 = p1 + spec_stride;
      const double* p3 = p2 + spec_stride;
      double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
      for (int k = 0; k < ngrid; k++) {
        l0 += p0[k];
        l1 += p1[k];
        l2 += p2[k];
        l3 += p3[k];
      }
      b.lum[s] = l0;
      b.lum[s + 1] = l1;
      b.lum[s + 2] = l2;
      b.lum[s + 3] = l3;
    }
    for (; s < nsrc; s++) {
      const double* ps = spec + (size_t)s * spec_stride;
      double l = 0.0;
      for (int k = 0; k < ngrid; k++) l += ps[k];
      b.lum[s] = l;
    }
  }
  const size_t nnode = 1 + (nsrc > 0 ? (size_t)count_nodes(b.key.data(), 0, nsrc, 63) : 0);
  t->link.reserve(nnode);
  t->cm.reserve(4 * nnode);
  t->spec64.reserve(nnode * ngrid);

  // Root (construct_tree, :529-569).
  const int root = add_node(t);
  double xc = 0.0, yc = 0.0, zc = 0.0, lum = 0.0;
  double* all = &t->spec64[0];
  for (int s = 0; s < nsrc; s++) {
    const double* ps = spec + (size_t)s * spec_stride;
    xc += pos[3 * s + 0] * ps[0];
    yc += pos[3 * s + 1] * ps[0];
    zc += pos[3 * s + 2] * ps[0];
    lum += ps[0];
    for (int k = 0; k < ngrid; k++) all[k] += ps[k];
  }
  xc /= lum;
  yc /= lum;
  zc /= lum;
  t->cm[0] = (float)xc;
  t->cm[1] = (float)yc;
  t->cm[2] = (float)zc;
  t->cm[3] = b.base_length * b.base_length;
  t->link[root].n = nsrc;
  if (nsrc == 1) t->link[root].src = 0;
  if (nsrc > 0) plant(b, root, 0, nsrc, 63, -1);
  t->link[root].skip = (int)t->link.size();
  // A node with n > 1 and no children (coincident keys at the deepest
  // level) keeps first = -1: opening it yields no ray, as in the reference.
  const int nn = (int)t->link.size();
  t->spec.assign((size_t)nn * ngrid, 0.0f);
  t->spec_scale.assign(nn, 0.0);
  for (int u = 0; u < nn; u++) {
    const double* row = &t->spec64[(size_t)u * ngrid];
    // four independent running maxima: max is exact, so splitting the
    // comparisons leaves the result unchanged and breaks the serial chain
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
    int k = 0;
    for (; k + 4 <= ngrid; k += 4) {
      m0 = std::max(m0, row[k]);
      m1 = std::max(m1, row[k + 1]);
      m2 = std::max(m2, row[k + 2]);
      m3 = std::max(m3, row[k + 3]);
    }
    for (; k < ngrid; k++) m0 = std::max(m0, row[k]);
    const double mx = std::max(std::max(m0, m1), std::max(m2, m3));
    t->spec_scale[u] = mx;
    const double inv = mx > 0.0 ? 1.0 / mx : 0.0;
    for (int k = 0; k < ngrid; k++) t->spec[(size_t)u * ngrid + k] = (float)(row[k] * inv);
  }
  return 0;
}

// Host replica of the device traversal (setup_light_ray_to, FP32 test).
namespace {
void visit(const HostTree& t, int u, const float* tp, std::vector<HostRay>* out, uint64_t* n) {
  const float dx = tp[0] - t.cm[4 * u + 0], dy = tp[1] - t.cm[4 * u + 1],
              dz = tp[2] - t.cm[4 * u + 2];
  const float r2 = dx * dx + dy * dy + dz * dz;
  const NodeLink& l = t.link[u];
  if (r2 * t.theta_sq > t.cm[4 * u + 3]) {
    if (out) out->push_back(HostRay{0, u, {t.cm[4 * u], t.cm[4 * u + 1], t.cm[4 * u + 2]}});
    (*n)++;
    return;
  }
  if (l.n > 1) {
    for (int c = l.first; c >= 0 && c < l.skip; c = t.link[c].skip) visit(t, c, tp, out, n);
  } else if (l.n == 1) {
    const int s = l.src;
    if (out)
      out->push_back(HostRay{1, s, {t.src_pos[4 * s], t.src_pos[4 * s + 1], t.src_pos[4 * s + 2]}});
    (*n)++;
  }
}
}  // namespace

uint64_t tree_rays_to(const HostTree& t, const float* target, std::vector<HostRay>* out) {
  uint64_t n = 0;
  if (!t.link.empty()) visit(t, 0, target, out, &n);
  return n;
}

}  // namespace argot

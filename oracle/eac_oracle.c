/*
 * eac_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded C restatement of the reference's CPU path for the EAC
 * hot path (diogo-aos/masters_final, Python 2 + Numba, read as source text;
 * numba is not installed and the reference is never run). Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, as a
 * checker or a timed baseline. The product (masters_final_amd/) never calls
 * it.
 *
 * Each function cites the reference lines it follows. Where the reference's
 * result depends on an unstable sort or on thread timing, the function
 * documents the canonical choice made instead ("canonical:" notes).
 *
 * Weights are carried as double: every integer weight of the reference's
 * graphs (uint8/uint16/int32) is exact in double and comparisons are the
 * reference's < and ==.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_ERR_VALUE (-7)
#define ORC_ERR_NOT_FOUND (-5)
#define ORC_ERR_ARG (-1)

/* ------------------------------------------------------------------ scans */

/* MyML/MyML/utils/scan.py:41-53 exprefixsumNumba: out[i] = carry; carry +=
 * in[i]; returns the final carry (init + total). */
int64_t orc_exprefixsum_i64(const int64_t* in, int64_t* out, int64_t n, int64_t init) {
  int64_t carry = init;
  for (int64_t i = 0; i < n; ++i) {
    out[i] = carry;
    carry += in[i];
  }
  return carry;
}

/* MyML/MyML/utils/scan.py:56-72 exprefixsumNumbaSingle: in place, keeper
 * holds the value being overwritten. Reads in[0] even for n == 0 in the
 * reference; here n == 0 returns init. */
int64_t orc_exprefixsum_single_i64(int64_t* a, int64_t n, int64_t init) {
  if (n <= 0) return init;
  int64_t carry = init;
  int64_t keeper = a[0];
  a[0] = init;
  for (int64_t i = 1; i < n; ++i) {
    carry += keeper;
    keeper = a[i];
    a[i] = carry;
  }
  carry += keeper;
  return carry;
}

/* ------------------------------------------------- Boruvka (reference CPU) */

/* mst.py:249-279 findMinEdgeNumba: strict < on weight, ties to the smaller
 * (contracted) destination; the first CSR entry wins among equal pairs. */
static void find_min_edge(int64_t* minedge, const double* w, const int64_t* fe,
                          const int64_t* od, const int64_t* dest, int64_t nv) {
  for (int64_t v = 0; v < nv; ++v) {
    int64_t deg = od[v];
    if (deg == 0) { minedge[v] = -1; continue; }
    int64_t start = fe[v], end = start + deg;
    int64_t me = start;
    double mw = w[start];
    int64_t md = dest[start];
    for (int64_t e = start + 1; e < end; ++e) {
      double ew = w[e];
      int64_t ed = dest[e];
      if (ew < mw) { me = e; mw = ew; md = ed; }
      else if (ew == mw && ed < md) { me = e; mw = ew; md = ed; }
    }
    minedge[v] = me;
  }
}

/* mst.py:282-305 removeMirroredNumba (sequential, in place). */
static void remove_mirrored(int64_t* minedge, const int64_t* dest, int64_t nv) {
  for (int64_t v = 0; v < nv; ++v) {
    int64_t my = minedge[v];
    if (my == -1) continue;
    int64_t succ = dest[my];
    int64_t se = minedge[succ];
    if (se == -1) continue;
    int64_t ss = dest[se];
    if (v == ss && v < succ) minedge[v] = -1;
  }
}

/* mst.py:320-351 initColorsNumba + propagateColorsNumba until converged. */
static void colors_to_roots(const int64_t* minedge, const int64_t* dest, int64_t* colors,
                            int64_t nv) {
  for (int64_t v = 0; v < nv; ++v) colors[v] = minedge[v] != -1 ? dest[minedge[v]] : v;
  int converged = 0;
  while (!converged) {
    converged = 1;
    for (int64_t v = 0; v < nv; ++v) {
      int64_t c = colors[v];
      if (c != v) {
        int64_t nc = colors[c];
        if (nc != c) converged = 0;
        colors[v] = nc;
      }
    }
  }
}

/* mst.py:143-232 boruvka_minho_seq with helpers mst.py:235-407.
 * Inputs: CSR (dest, w, fe, od) with nv vertices and ne entries.
 * Output: mst[0:n_mst] = original edge ids (CSR positions) in the order the
 * reference appends them (round by round, vertex order). Returns n_mst.
 * components (nullable, [nv]) receives the final supervertex id of every
 * vertex, following connected_components.py:88-94,162-173. */
int64_t orc_boruvka_minho_seq(const int64_t* dest_in, const double* w_in, const int64_t* fe_in,
                              const int64_t* od_in, int64_t nv, int64_t ne, int64_t* mst,
                              int64_t* components) {
  if (nv <= 0) return 0;
  int64_t* dest = malloc(sizeof(int64_t) * (ne > 0 ? ne : 1));
  double* w = malloc(sizeof(double) * (ne > 0 ? ne : 1));
  int64_t* eid = malloc(sizeof(int64_t) * (ne > 0 ? ne : 1));
  int64_t* fe = malloc(sizeof(int64_t) * nv);
  int64_t* od = malloc(sizeof(int64_t) * nv);
  memcpy(dest, dest_in, sizeof(int64_t) * ne);
  memcpy(w, w_in, sizeof(double) * ne);
  memcpy(fe, fe_in, sizeof(int64_t) * nv);
  memcpy(od, od_in, sizeof(int64_t) * nv);
  for (int64_t e = 0; e < ne; ++e) eid[e] = e;
  int64_t* labels = NULL;
  if (components) {
    labels = components;
    for (int64_t v = 0; v < nv; ++v) labels[v] = v;
  }
  int64_t ncomp = nv, n_mst = 0;
  for (;;) {
    int64_t* minedge = malloc(sizeof(int64_t) * ncomp);
    int64_t* colors = malloc(sizeof(int64_t) * ncomp);
    find_min_edge(minedge, w, fe, od, dest, ncomp);
    remove_mirrored(minedge, dest, ncomp);
    for (int64_t v = 0; v < ncomp; ++v)  /* mst.py:308-317 */
      if (minedge[v] != -1) mst[n_mst++] = eid[minedge[v]];
    colors_to_roots(minedge, dest, colors, ncomp);
    /* mst.py:235-243 buildFlag + exclusive scan (scan.py:56-72) */
    int64_t* newv = minedge;
    for (int64_t v = 0; v < ncomp; ++v) newv[v] = (colors[v] == v);
    int64_t nnew = orc_exprefixsum_single_i64(newv, ncomp, 0);
    if (labels)  /* connected_components.py:162-173 */
      for (int64_t v = 0; v < nv; ++v) labels[v] = newv[colors[labels[v]]];
    if (nnew == 1) { free(minedge); free(colors); break; }
    /* mst.py:354-372 countNewEdgesNumba */
    int64_t* nod = calloc(nnew, sizeof(int64_t));
    for (int64_t v = 0; v < ncomp; ++v) {
      int64_t c = colors[v], cid = newv[c];
      for (int64_t e = fe[v]; e < fe[v] + od[v]; ++e)
        if (colors[dest[e]] != c) nod[cid] += 1;
    }
    int64_t* nfe = malloc(sizeof(int64_t) * nnew);
    int64_t nne = orc_exprefixsum_i64(nod, nfe, nnew, 0);
    if (nne == 0) { free(minedge); free(colors); free(nod); free(nfe); break; }
    /* mst.py:378-407 assignInsertNumba */
    int64_t* ndest = malloc(sizeof(int64_t) * nne);
    double* nw = malloc(sizeof(double) * nne);
    int64_t* neid = malloc(sizeof(int64_t) * nne);
    int64_t* top = malloc(sizeof(int64_t) * nnew);
    memcpy(top, nfe, sizeof(int64_t) * nnew);
    for (int64_t v = 0; v < ncomp; ++v) {
      int64_t c = colors[v];
      for (int64_t e = fe[v]; e < fe[v] + od[v]; ++e) {
        int64_t sc = colors[dest[e]];
        if (c != sc) {
          int64_t sv = newv[c], ssv = newv[sc];
          int64_t p = top[sv]++;
          ndest[p] = ssv;
          nw[p] = w[e];
          neid[p] = eid[e];
        }
      }
    }
    free(dest); free(w); free(eid); free(fe); free(od); free(top);
    free(minedge); free(colors);
    dest = ndest; w = nw; eid = neid; fe = nfe; od = nod;
    ncomp = nnew;
  }
  free(dest); free(w); free(eid); free(fe); free(od);
  return n_mst;
}

/* connected_components.py:46-137 connected_comps_seq: same contraction,
 * labels are final supervertex ids (not first-occurrence numbered). */
void orc_connected_comps_seq(const int64_t* dest, const double* w, const int64_t* fe,
                             const int64_t* od, int64_t nv, int64_t ne, int64_t* labels) {
  int64_t* mst = malloc(sizeof(int64_t) * (nv > 1 ? nv - 1 : 1));
  orc_boruvka_minho_seq(dest, w, fe, od, nv, ne, mst, labels);
  free(mst);
}

/* --------------------------------------------- canonical Kruskal (checker)
 * Unique minimum spanning forest under the strict order (w, min, max).
 * Edges given as undirected triples; ties of identical triples keep the
 * first. Output: in_forest[i] flags; parent[] receives min-id roots. */
typedef struct { double w; int64_t lo, hi, idx; } orc_edge;

static int cmp_edge(const void* a, const void* b) {
  const orc_edge* x = a;
  const orc_edge* y = b;
  if (x->w < y->w) return -1;
  if (x->w > y->w) return 1;
  if (x->lo != y->lo) return x->lo < y->lo ? -1 : 1;
  if (x->hi != y->hi) return x->hi < y->hi ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

static int64_t uf_find(int64_t* p, int64_t x) {
  while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
  return x;
}
/* union by min id: the root of every set is its smallest vertex */
static int uf_union(int64_t* p, int64_t a, int64_t b) {
  a = uf_find(p, a); b = uf_find(p, b);
  if (a == b) return 0;
  if (a < b) p[b] = a; else p[a] = b;
  return 1;
}

int64_t orc_kruskal_canonical(const int64_t* u, const int64_t* v, const double* w, int64_t ne,
                              int64_t nv, uint8_t* in_forest, int64_t* order_out) {
  orc_edge* E = malloc(sizeof(orc_edge) * (ne > 0 ? ne : 1));
  for (int64_t i = 0; i < ne; ++i) {
    E[i].w = w[i];
    E[i].lo = u[i] < v[i] ? u[i] : v[i];
    E[i].hi = u[i] < v[i] ? v[i] : u[i];
    E[i].idx = i;
  }
  qsort(E, ne, sizeof(orc_edge), cmp_edge);
  int64_t* p = malloc(sizeof(int64_t) * nv);
  for (int64_t i = 0; i < nv; ++i) p[i] = i;
  memset(in_forest, 0, ne);
  int64_t nf = 0;
  for (int64_t i = 0; i < ne; ++i) {
    if (E[i].lo == E[i].hi) continue;
    if (uf_union(p, E[i].lo, E[i].hi)) {
      in_forest[E[i].idx] = 1;
      if (order_out) order_out[nf] = E[i].idx;
      ++nf;
    }
  }
  free(E); free(p);
  return nf;
}

/* All-core CPU baseline of the same forest: Borůvka on OpenMP threads over
 * the symmetric CSR (indptr[n+1], dest, int32 weights), keys (w - wmin,
 * min id, max id) so the forest is the canonical one. Each round: every
 * vertex's lightest key to another component, a compare-and-swap minimum
 * per component, hooks (a mutual choice keeps the smaller root), then every
 * vertex walks to its root. Writes the forest's (lo, hi) pairs and returns
 * their count (-1 when the key does not fit 64 bits). Not the reference's
 * algorithm (mst.py is single-threaded): it is the parallel CPU line the
 * measurements report beside it. */
int64_t orc_boruvka_omp(const int64_t* indptr, const int32_t* dest, const int32_t* w,
                        int64_t n, int64_t* lo_out, int64_t* hi_out, double* weight_out) {
  int b = 0;
  while (b < 63 && (((uint64_t)(n > 1 ? n - 1 : 1)) >> b) != 0) ++b;
  int64_t wmin = INT64_MAX, wmax = INT64_MIN;
  const int64_t m = indptr[n];
#pragma omp parallel for reduction(min : wmin) reduction(max : wmax)
  for (int64_t e = 0; e < m; ++e) {
    wmin = w[e] < wmin ? w[e] : wmin;
    wmax = w[e] > wmax ? w[e] : wmax;
  }
  if (m == 0) { wmin = 0; wmax = 0; }
  int wb = 0;
  while (wb < 63 && (((uint64_t)(wmax - wmin)) >> wb) != 0) ++wb;
  if (wb + 2 * b > 64) return -1;
  const uint64_t idmask = (b >= 64) ? ~0ull : ((1ull << b) - 1);
  int64_t* comp = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  int64_t* parent = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  uint64_t* best = malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
  uint64_t* tree = malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
#pragma omp parallel for
  for (int64_t v = 0; v < n; ++v) { comp[v] = v; parent[v] = v; tree[v] = ~0ull; }
  for (;;) {
    int any = 0;
#pragma omp parallel for
    for (int64_t v = 0; v < n; ++v) best[v] = ~0ull;
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : any)
    for (int64_t u = 0; u < n; ++u) {
      const int64_t cu = comp[u];
      uint64_t k = ~0ull;
      for (int64_t e = indptr[u]; e < indptr[u + 1]; ++e) {
        const int64_t v = dest[e];
        if (comp[v] == cu) continue;
        const uint64_t a = (uint64_t)(u < v ? u : v), c = (uint64_t)(u < v ? v : u);
        const uint64_t key = ((uint64_t)(w[e] - wmin) << (2 * b)) | (a << b) | c;
        k = key < k ? key : k;
      }
      if (k == ~0ull) continue;
      any = 1;
      uint64_t cur = __atomic_load_n(&best[cu], __ATOMIC_RELAXED);
      while (k < cur && !__atomic_compare_exchange_n(&best[cu], &cur, k, 1, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED)) {
      }
    }
    if (!any) break;
#pragma omp parallel for
    for (int64_t r = 0; r < n; ++r) {
      if (comp[r] != r || best[r] == ~0ull) continue;
      const uint64_t k = best[r];
      const int64_t a = (int64_t)((k >> b) & idmask), c = (int64_t)(k & idmask);
      const int64_t other = comp[a] == r ? comp[c] : comp[a];
      if (best[other] == k && r < other) continue;
      parent[r] = other;
      tree[r] = k;
    }
#pragma omp parallel for
    for (int64_t v = 0; v < n; ++v) {
      int64_t x = v;
      while (parent[x] != x) x = parent[x];
      comp[v] = x;
    }
  }
  int64_t nf = 0;
  double tw = 0.0;
  for (int64_t r = 0; r < n; ++r) {
    if (tree[r] == ~0ull) continue;
    const uint64_t k = tree[r];
    lo_out[nf] = (int64_t)((k >> b) & idmask);
    hi_out[nf] = (int64_t)(k & idmask);
    tw += (double)((int64_t)(k >> (2 * b)) + wmin);
    ++nf;
  }
  *weight_out = tw;
  free(comp); free(parent); free(best); free(tree);
  return nf;
}

/* All-core CPU baseline of the batched cut: orc_sp_sl_lifetime (eac.py rule,
 * max_dist None) on every graph of a disjoint union, graphs spread over
 * OpenMP threads (largest first). Graph g owns vertices [goff[g],
 * goff[g+1]) and their rows of the symmetric CSR (union ids in dest).
 * Writes per graph the component count (0 where the reference returns empty
 * labels, -1 where it raises) and the first-occurrence labels in
 * labels[goff[g]...]. */
int orc_sp_sl_lifetime(const int64_t* u, const int64_t* v, const double* w, int64_t ne,
                       int64_t nv, double max_dist, int64_t n_clusters, int64_t* labels,
                       int64_t* n_comps, int64_t* labels_empty, double* th_out,
                       int64_t* nc_stable_out);

void orc_sl_batched_omp(const int64_t* indptr, const int32_t* dest, const int32_t* weight,
                        const int64_t* goff, int64_t G, const int64_t* order, int64_t* nc_out,
                        int64_t* labels) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t t = 0; t < G; ++t) {
    const int64_t g = order ? order[t] : t;
    const int64_t a = goff[g], z = goff[g + 1], nv = z - a;
    int64_t ne = 0;
    for (int64_t r = a; r < z; ++r)
      for (int64_t e = indptr[r]; e < indptr[r + 1]; ++e) ne += dest[e] > r;
    int64_t* u = malloc(sizeof(int64_t) * (ne > 0 ? ne : 1));
    int64_t* v = malloc(sizeof(int64_t) * (ne > 0 ? ne : 1));
    double* w = malloc(sizeof(double) * (ne > 0 ? ne : 1));
    int64_t k = 0;
    for (int64_t r = a; r < z; ++r)
      for (int64_t e = indptr[r]; e < indptr[r + 1]; ++e)
        if (dest[e] > r) { u[k] = r - a; v[k] = dest[e] - a; w[k] = (double)weight[e]; ++k; }
    int64_t nc = 0, empty = 0, ncs = 0;
    double th = 0.0;
    const int rc = orc_sp_sl_lifetime(u, v, w, ne, nv, -1.0, 0, labels + a, &nc, &empty, &th, &ncs);
    nc_out[g] = rc != ORC_OK ? -1 : (empty ? 0 : nc);
    free(u); free(v); free(w);
  }
}

int orc_omp_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Components of an undirected edge list, numbered by first occurrence
 * (the numbering scipy.sparse.csgraph.connected_components returns, used by
 * eac.py:262). Returns the number of components. */
int64_t orc_components(const int64_t* u, const int64_t* v, int64_t ne, int64_t nv,
                       int64_t* labels) {
  int64_t* p = malloc(sizeof(int64_t) * (nv > 0 ? nv : 1));
  for (int64_t i = 0; i < nv; ++i) p[i] = i;
  for (int64_t i = 0; i < ne; ++i) uf_union(p, u[i], v[i]);
  int64_t* rank = malloc(sizeof(int64_t) * (nv > 0 ? nv : 1));
  int64_t nc = 0;
  for (int64_t i = 0; i < nv; ++i) {
    int64_t r = uf_find(p, i);
    if (r == i) rank[i] = nc++;
    labels[i] = rank[r];  /* r <= i, so rank[r] is already set */
  }
  free(p); free(rank);
  return nc;
}

/* First-occurrence renumbering of arbitrary labels in [0, nvals). */
int64_t orc_canonical_relabel(const int64_t* in, int64_t n, int64_t nvals, int64_t* out) {
  int64_t* map = malloc(sizeof(int64_t) * (nvals > 0 ? nvals : 1));
  for (int64_t i = 0; i < nvals; ++i) map[i] = -1;
  int64_t c = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (map[in[i]] < 0) map[in[i]] = c++;
    out[i] = map[in[i]];
  }
  free(map);
  return c;
}

/* ------------------------------------------------------- graph from edges */

/* build.py:18-45 binaryOriginVertexSearch (integer division of Python 2). */
static int64_t origin_search(int64_t key, const int64_t* fe, const int64_t* od, int64_t nv) {
  int64_t imin = 0, imax = nv;
  while (imin < imax) {
    int64_t imid = (imax + imin) / 2;
    int64_t f = fe[imid];
    if (key < f) imax = imid;
    else if (key > f + od[imid] - 1) imin = imid + 1;
    else return imid;
  }
  return -1;
}

/* build.py:49-93 getGraphFromEdges_seq. nod must be zeroed by the caller
 * (the reference's sl_mst_lifetime_seq passes np.zeros_like, singlelink.py:55).
 * Returns 0, or -1 when an edge id has no origin row. */
int64_t orc_graph_from_edges(const int64_t* dest, const double* w, const int64_t* fe,
                             const int64_t* od, int64_t nv, const int64_t* edges, int64_t nedges,
                             int64_t* nod, int64_t* nfe, int64_t* ndest, double* nw) {
  for (int64_t e = 0; e < nedges; ++e) {
    int64_t ed = edges[e];
    int64_t ov = dest[ed];
    int64_t iv = origin_search(ed, fe, od, nv);
    if (iv == -1) return -1;
    nod[ov] += 1;
    nod[iv] += 1;
  }
  orc_exprefixsum_i64(nod, nfe, nv, 0);
  int64_t* top = malloc(sizeof(int64_t) * (nv > 0 ? nv : 1));
  memcpy(top, nfe, sizeof(int64_t) * nv);
  for (int64_t e = 0; e < nedges; ++e) {
    int64_t ed = edges[e];
    int64_t ov = dest[ed];
    int64_t iv = origin_search(ed, fe, od, nv);
    if (iv == -1) { free(top); return -1; }
    int64_t ip = top[iv], op = top[ov];
    ndest[ip] = ov;
    ndest[op] = iv;
    nw[ip] = w[ed];
    nw[op] = w[ed];
    top[iv] += 1;
    top[ov] += 1;
  }
  free(top);
  return 0;
}

/* ------------------------------------------------------- lifetime cuts */

typedef struct { double w; int64_t lo, hi; } orc_fedge;
static int cmp_fedge(const void* a, const void* b) {
  const orc_fedge* x = a;
  const orc_fedge* y = b;
  if (x->w < y->w) return -1;
  if (x->w > y->w) return 1;
  if (x->lo != y->lo) return x->lo < y->lo ? -1 : 1;
  return (x->hi > y->hi) - (x->hi < y->hi);
}

/* eac.py:366-391 lifetime_n_clusters on ascending weights.
 * Returns ORC_ERR_VALUE where numpy raises (argmax of an empty array, or
 * np.min over no close-to-zero lifetime). */
static int lifetime_n_clusters(const double* sw, int64_t m, int64_t* cont_out, double* th_out,
                               int64_t* a_out) {
  if (m < 2) return ORC_ERR_VALUE;
  int64_t a = 0;
  double th = sw[1] - sw[0];
  for (int64_t i = 1; i < m - 1; ++i) {
    double lt = sw[i + 1] - sw[i];
    if (lt > th) { th = lt; a = i; }
  }
  int64_t above = 0;
  for (int64_t i = 0; i < m; ++i) above += sw[i] > sw[a];
  int64_t cont = above == 0 ? 1 : above + 1;
  /* np.isclose(lt, 0): |lt| <= 1e-8 (atol) + 1e-5 * 0 */
  int found = 0;
  double mn = 0.0;
  for (int64_t i = 0; i < m - 1; ++i) {
    double lt = sw[i + 1] - sw[i];
    if (fabs(lt) <= 1e-8) {
      if (!found || lt < mn) mn = lt;
      found = 1;
    }
  }
  if (!found) return ORC_ERR_VALUE;
  if (th < 2 * mn) cont = 1;
  *cont_out = cont;
  *th_out = th;
  *a_out = a;
  return ORC_OK;
}

/* eac.py:199-267 sp_sl_lifetime on an undirected graph given as edge
 * triples (any orientation; each undirected edge once).
 * max_dist < 0 means None (max weight + 1, eac.py:220-221); otherwise
 * max_dist + 1 (eac.py:223). n_clusters == 0 uses the lifetime rule.
 * Writes labels (first-occurrence numbering, scipy's order) and returns the
 * number of components in *n_comps; *labels_empty = 1 when the reference
 * returns an empty label array (nc_stable <= 1, eac.py:263-265).
 * canonical: Kruskal's unstable argsort (_min_spanning_tree.py:91,
 * eac.py:232) is replaced by the (w, min, max) order; under the lifetime
 * rule the partition does not depend on it (cuts are all edges with
 * w > w[a]); with an explicit n_clusters the tied edges at the cut are
 * taken largest (w, min, max) first. */
int orc_sp_sl_lifetime(const int64_t* u, const int64_t* v, const double* w, int64_t ne,
                       int64_t nv, double max_dist, int64_t n_clusters, int64_t* labels,
                       int64_t* n_comps, int64_t* labels_empty, double* th_out,
                       int64_t* nc_stable_out) {
  uint8_t* inf = malloc(ne > 0 ? ne : 1);
  int64_t* order = malloc(sizeof(int64_t) * (ne > 0 ? ne : 1));
  double maxw = -INFINITY;
  for (int64_t i = 0; i < ne; ++i) maxw = w[i] > maxw ? w[i] : maxw;
  if (max_dist < 0) max_dist = maxw + 1; else max_dist += 1;
  /* Kruskal drops zero weights through eliminate_zeros
     (_min_spanning_tree.py:163-175); only positive weights count. */
  int64_t nf = orc_kruskal_canonical(u, v, w, ne, nv, inf, order);
  orc_fedge* F = malloc(sizeof(orc_fedge) * (nf > 0 ? nf : 1));
  int64_t k = 0;
  for (int64_t i = 0; i < nf; ++i) {
    int64_t e = order[i];
    if (w[e] == 0.0) continue;
    F[k].w = w[e];
    F[k].lo = u[e] < v[e] ? u[e] : v[e];
    F[k].hi = u[e] < v[e] ? v[e] : u[e];
    ++k;
  }
  nf = k;
  qsort(F, nf, sizeof(orc_fedge), cmp_fedge);
  int64_t n_disc = nv - nf;  /* eac.py:229 mst.shape[0] - mst.nnz */
  double* sw = malloc(sizeof(double) * (nf > 0 ? nf : 1));
  for (int64_t i = 0; i < nf; ++i) sw[i] = F[i].w;
  int64_t nc_stable;
  double th = 0.0;
  if (n_clusters == 0) {
    int64_t cont, a;
    int rc = lifetime_n_clusters(sw, nf, &cont, &th, &a);
    if (rc != ORC_OK) { free(inf); free(order); free(F); free(sw); return rc; }
    if (n_disc > 1) {
      double disconnect_lifetime = max_dist + 1 - sw[nf - 1];
      if (th > disconnect_lifetime) cont += n_disc - 1;
      else cont = n_disc;
    }
    nc_stable = cont;
  } else {
    nc_stable = n_clusters;
  }
  int64_t keep = nf;
  if (nc_stable > n_disc) {
    int64_t n_cuts = nc_stable - n_disc;
    keep = nf - n_cuts;
    if (keep < 0) keep = 0;
  }
  *th_out = th;
  *nc_stable_out = nc_stable;
  if (nc_stable > 1) {
    int64_t* cu = malloc(sizeof(int64_t) * (keep > 0 ? keep : 1));
    int64_t* cv = malloc(sizeof(int64_t) * (keep > 0 ? keep : 1));
    for (int64_t i = 0; i < keep; ++i) { cu[i] = F[i].lo; cv[i] = F[i].hi; }
    *n_comps = orc_components(cu, cv, keep, nv, labels);
    *labels_empty = 0;
    free(cu); free(cv);
  } else {
    *n_comps = 1;
    *labels_empty = 1;
  }
  free(inf); free(order); free(F); free(sw);
  return ORC_OK;
}

/* singlelink.py:14-65 sl_mst_lifetime_seq on the forest of an undirected
 * graph (edge triples). disconnect_weight < 0 means None (max graph weight,
 * singlelink.py:16-17). Keeps the a smallest forest edges when the largest
 * lifetime is >= disconnect_weight - max forest weight (the reference drops
 * m - a edges, singlelink.py:43-48), then labels the kept forest's
 * components (first occurrence). canonical: forest edges ordered by
 * (w, min, max) instead of numpy's unstable argsort (singlelink.py:27).
 * Returns ORC_ERR_VALUE for fewer than 2 forest edges (argmax of empty). */
int orc_sl_mst_lifetime(const int64_t* u, const int64_t* v, const double* w, int64_t ne,
                        int64_t nv, double disconnect_weight, int64_t* labels,
                        int64_t* n_comps, int64_t* n_kept) {
  uint8_t* inf = malloc(ne > 0 ? ne : 1);
  int64_t* order = malloc(sizeof(int64_t) * (ne > 0 ? ne : 1));
  double maxw = -INFINITY;
  for (int64_t i = 0; i < ne; ++i) maxw = w[i] > maxw ? w[i] : maxw;
  if (disconnect_weight < 0) disconnect_weight = maxw;
  int64_t nf = orc_kruskal_canonical(u, v, w, ne, nv, inf, order);
  orc_fedge* F = malloc(sizeof(orc_fedge) * (nf > 0 ? nf : 1));
  for (int64_t i = 0; i < nf; ++i) {
    int64_t e = order[i];
    F[i].w = w[e];
    F[i].lo = u[e] < v[e] ? u[e] : v[e];
    F[i].hi = u[e] < v[e] ? v[e] : u[e];
  }
  qsort(F, nf, sizeof(orc_fedge), cmp_fedge);
  if (nf < 2) { free(inf); free(order); free(F); return ORC_ERR_VALUE; }
  int64_t a = 0;
  double mlt = F[1].w - F[0].w;
  for (int64_t i = 1; i < nf - 1; ++i) {
    double lt = F[i + 1].w - F[i].w;
    if (lt > mlt) { mlt = lt; a = i; }
  }
  double lt_threshold = disconnect_weight - F[nf - 1].w;
  int64_t keep = nf;
  if (mlt >= lt_threshold) keep = a;  /* n_discarded = (nf-1) - a + 1 */
  int64_t* cu = malloc(sizeof(int64_t) * (keep > 0 ? keep : 1));
  int64_t* cv = malloc(sizeof(int64_t) * (keep > 0 ? keep : 1));
  for (int64_t i = 0; i < keep; ++i) { cu[i] = F[i].lo; cv[i] = F[i].hi; }
  *n_comps = orc_components(cu, cv, keep, nv, labels);
  *n_kept = keep;
  free(cu); free(cv); free(inf); free(order); free(F);
  return ORC_OK;
}

/* ------------------------------------------------------- co-association */

/* Co-association counts of a set of partitions, as the condensed sparse
 * builder of MyML/MyML/EAC/sparse.py accumulates them with no capacity cap
 * (so nothing is discarded): the first partition processed is the one with
 * the fewest clusters (sparse.py:154-179, np.argmin -> first minimum); its
 * clusters write count 1 for every pair j > i of ascending members
 * (sparse.py:445-465); every other partition increments an existing
 * (row, col) found by binary search in the sorted row, or inserts it keeping
 * the row sorted (sparse.py:956-1040 with binary_search_interval
 * sparse.py:1069-1099).
 * labels: P x n int64 (row-major), values in [0, k_p). Output: rows of the
 * upper triangle (col > row) sorted by col, as CSR: indptr[n+1] (int64),
 * indices, counts. cap_total bounds indices/counts; returns nnz or -1 if
 * cap_total is too small. */
int64_t orc_coassoc_condensed(const int64_t* labels, const int64_t* kp, int64_t P, int64_t n,
                              int64_t* indptr, int64_t* indices, int64_t* counts,
                              int64_t cap_total) {
  /* per-row growable arrays */
  int64_t** rc = calloc(n, sizeof(int64_t*));
  int64_t** rv = calloc(n, sizeof(int64_t*));
  int64_t* deg = calloc(n, sizeof(int64_t));
  int64_t* cap = calloc(n, sizeof(int64_t));
  int64_t first = 0;
  for (int64_t p = 1; p < P; ++p) if (kp[p] < kp[first]) first = p;
  int64_t kmax = 0;
  for (int64_t p = 0; p < P; ++p) kmax = kp[p] > kmax ? kp[p] : kmax;
  /* members of each cluster, ascending (partition.py:73-86 order) */
  int64_t* cnt = malloc(sizeof(int64_t) * (kmax + 1));
  int64_t* off = malloc(sizeof(int64_t) * (kmax + 1));
  int64_t* mem = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  for (int64_t step = 0; step < P; ++step) {
    int64_t p = step == 0 ? first : (step <= first ? step - 1 : step);
    const int64_t* lab = labels + p * n;
    /* a negative label: the point is in no cluster of this partition */
    memset(cnt, 0, sizeof(int64_t) * (kmax + 1));
    for (int64_t i = 0; i < n; ++i)
      if (lab[i] >= 0) cnt[lab[i]]++;
    off[0] = 0;
    for (int64_t c = 0; c < kmax; ++c) off[c + 1] = off[c] + cnt[c];
    memset(cnt, 0, sizeof(int64_t) * (kmax + 1));
    for (int64_t i = 0; i < n; ++i)
      if (lab[i] >= 0) mem[off[lab[i]] + cnt[lab[i]]++] = i;
    for (int64_t c = 0; c < kp[p]; ++c) {
      const int64_t* S = mem + off[c];
      int64_t s = off[c + 1] - off[c];
      for (int64_t a = 0; a < s; ++a) {
        int64_t r = S[a];
        for (int64_t b = a + 1; b < s; ++b) {
          int64_t col = S[b];
          /* binary search for col in row r */
          int64_t lo = 0, hi = deg[r];
          while (lo < hi) {
            int64_t mid = (lo + hi) / 2;
            if (rc[r][mid] < col) lo = mid + 1; else hi = mid;
          }
          if (lo < deg[r] && rc[r][lo] == col) { rv[r][lo] += 1; continue; }
          if (deg[r] == cap[r]) {
            cap[r] = cap[r] ? cap[r] * 2 : 8;
            rc[r] = realloc(rc[r], sizeof(int64_t) * cap[r]);
            rv[r] = realloc(rv[r], sizeof(int64_t) * cap[r]);
          }
          memmove(rc[r] + lo + 1, rc[r] + lo, sizeof(int64_t) * (deg[r] - lo));
          memmove(rv[r] + lo + 1, rv[r] + lo, sizeof(int64_t) * (deg[r] - lo));
          rc[r][lo] = col;
          rv[r][lo] = 1;
          deg[r] += 1;
        }
      }
    }
  }
  /* sparse.py:336-349 condense_eac_csr: rows packed in order */
  int64_t nnz = 0;
  indptr[0] = 0;
  for (int64_t r = 0; r < n; ++r) {
    if (nnz + deg[r] > cap_total) { nnz = -1; break; }
    for (int64_t j = 0; j < deg[r]; ++j) {
      indices[nnz + j] = rc[r][j];
      counts[nnz + j] = rv[r][j];
    }
    nnz += deg[r];
    indptr[r + 1] = nnz;
  }
  for (int64_t r = 0; r < n; ++r) { free(rc[r]); free(rv[r]); }
  free(rc); free(rv); free(deg); free(cap); free(cnt); free(off); free(mem);
  return nnz;
}

/* binary_search_interval, MyML/MyML/EAC/sparse.py:1069-1099: index of key in
 * the sorted ary[start:end), else -(insertion point) - 1. */
static int64_t orc_bsearch_interval(int64_t key, const int64_t* ary, int64_t start, int64_t end) {
  int64_t imin = start, imax = end;
  while (imin < imax) {
    int64_t imid = (imax + imin) / 2;
    if (key < ary[imid]) imax = imid;
    else if (key > ary[imid]) imin = imid + 1;
    else return imid;
  }
  return -imin - 1;
}

/* The reference's capacity-capped co-association, EAC_CSR with its default
 * sort_mode "surgical" (MyML/MyML/EAC/sparse.py):
 *  - storage: one flat indices/data pair with row r owning
 *    [cap_off[r], cap_off[r+1]) (indptr_constant :368-369 / indptr_linear
 *    :371-413, computed by the caller) and degree[r] entries in use;
 *  - update_ensemble (:154-166): the partition with the fewest clusters
 *    (np.argmin of the cluster-list lengths, first minimum; clusters are the
 *    distinct labels present, partition.py:73-86) goes first, then every
 *    other partition in order;
 *  - first partition: update_cluster_fp_condensed (:445-465) writes count 1
 *    for every later member of the cluster (update_cluster_fp :425-443 in
 *    full mode: every other member), with no capacity check;
 *  - other partitions: update_cluster_sorted_surgical_condensed (:956-1040)
 *    / update_cluster_sorted_surgical (:866-950, full mode): for each row, a
 *    co-member found in the row (binary search over the entries the row had
 *    before this cluster) gets +1; a new one is queued while the row's
 *    degree is below its capacity (cap_off[r+1] - cap_off[r] condensed,
 *    max_assocs in full mode) and otherwise discarded and counted; the
 *    queued ones are merged into the row keeping it sorted. In full mode a
 *    row id above 405109 returns from the cluster (:877-878);
 *  - _condense (:212-225, condense_eac_csr :336-349) packs the rows.
 * labels: P x n (-1 = not listed in that partition). Outputs: indptr[n+1],
 * indices/counts (room for cap_off[n] entries), *nnz, *discarded (the sum of
 * the per-cluster discard counts the surgical update returns, i.e. every
 * discarded co-occurrence). Returns ORC_ERR_VALUE, with *bad_row set, when
 * the first partition gives a row more entries than its capacity: the
 * reference then writes past the row into the next row's storage, which is
 * not restated. */
int orc_coassoc_capped(const int64_t* labels, int64_t P, int64_t n, int full_rows,
                       const int64_t* cap_off, int64_t max_assocs, int64_t* indptr,
                       int64_t* indices, int64_t* counts, int64_t* nnz_out,
                       int64_t* discarded_out, int64_t* bad_row) {
  const int64_t total = cap_off[n];
  int64_t* ind = malloc(sizeof(int64_t) * (total > 0 ? total : 1));
  int64_t* dat = malloc(sizeof(int64_t) * (total > 0 ? total : 1));
  int64_t* deg = calloc(n > 0 ? n : 1, sizeof(int64_t));
  int64_t kmax = 0;
  for (int64_t t = 0; t < P * n; ++t) kmax = labels[t] + 1 > kmax ? labels[t] + 1 : kmax;
  int64_t* cnt = malloc(sizeof(int64_t) * (kmax + 1));
  int64_t* off = malloc(sizeof(int64_t) * (kmax + 1));
  int64_t* mem = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  int64_t* q_col = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  int64_t* q_at = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  int64_t* tmp_i = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  int64_t* tmp_d = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  /* first partition: fewest distinct labels present, first minimum */
  int64_t first = 0, best = -1;
  for (int64_t p = 0; p < P; ++p) {
    memset(cnt, 0, sizeof(int64_t) * (kmax + 1));
    int64_t distinct = 0;
    for (int64_t i = 0; i < n; ++i) {
      int64_t l = labels[p * n + i];
      if (l >= 0 && cnt[l]++ == 0) ++distinct;
    }
    if (best < 0 || distinct < best) { best = distinct; first = p; }
  }
  int rc = ORC_OK;
  int64_t discarded = 0;
  *bad_row = -1;
  for (int64_t step = 0; step < P && rc == ORC_OK; ++step) {
    int64_t p = step == 0 ? first : (step <= first ? step - 1 : step);
    const int64_t* lab = labels + p * n;
    memset(cnt, 0, sizeof(int64_t) * (kmax + 1));
    for (int64_t i = 0; i < n; ++i) if (lab[i] >= 0) cnt[lab[i]]++;
    off[0] = 0;
    for (int64_t c = 0; c < kmax; ++c) off[c + 1] = off[c] + cnt[c];
    memset(cnt, 0, sizeof(int64_t) * (kmax + 1));
    for (int64_t i = 0; i < n; ++i) if (lab[i] >= 0) mem[off[lab[i]] + cnt[lab[i]]++] = i;
    for (int64_t c = 0; c < kmax && rc == ORC_OK; ++c) {
      const int64_t* S = mem + off[c];
      const int64_t s = off[c + 1] - off[c];
      if (s == 0) continue;  /* absent label: no cluster in the list format */
      if (step == 0) {
        for (int64_t a = 0; a < s; ++a) {
          const int64_t r = S[a], fa = cap_off[r];
          int64_t add = 0;
          for (int64_t b = full_rows ? 0 : a + 1; b < s; ++b) {
            if (b == a) continue;
            if (fa + add >= cap_off[r + 1]) { rc = ORC_ERR_VALUE; *bad_row = r; break; }
            ind[fa + add] = S[b];
            dat[fa + add] = 1;
            ++add;
          }
          if (rc != ORC_OK) break;
          deg[r] = add;
        }
        continue;
      }
      for (int64_t a = 0; a < (full_rows ? s : s - 1); ++a) {
        const int64_t r = S[a];
        if (full_rows && r > 405109) break;
        const int64_t fa = cap_off[r], n_degree = deg[r], la = fa + n_degree;
        const int64_t cap = full_rows ? max_assocs : cap_off[r + 1] - fa;
        int64_t new_deg = n_degree, nq = 0;
        for (int64_t b = full_rows ? 0 : a + 1; b < s; ++b) {
          if (b == a) continue;
          const int64_t col = S[b];
          const int64_t at = orc_bsearch_interval(col, ind, fa, la);
          if (at >= 0) { dat[at] += 1; continue; }
          if (new_deg >= cap) { ++discarded; continue; }
          q_col[nq] = col;
          q_at[nq] = -(at + 1);
          ++nq;
          ++new_deg;
        }
        if (nq == 0) continue;
        /* merge the queued columns into the row (sorted), as the shifting
         * loop of :1015-1038 does */
        int64_t o = 0, e = fa;
        for (int64_t t = 0; t < nq; ++t) {
          while (e < q_at[t]) { tmp_i[o] = ind[e]; tmp_d[o] = dat[e]; ++o; ++e; }
          tmp_i[o] = q_col[t]; tmp_d[o] = 1; ++o;
        }
        while (e < la) { tmp_i[o] = ind[e]; tmp_d[o] = dat[e]; ++o; ++e; }
        memcpy(ind + fa, tmp_i, sizeof(int64_t) * o);
        memcpy(dat + fa, tmp_d, sizeof(int64_t) * o);
        deg[r] = new_deg;
      }
    }
  }
  if (rc == ORC_OK) {
    int64_t nnz = 0;
    indptr[0] = 0;
    for (int64_t r = 0; r < n; ++r) {
      memcpy(indices + nnz, ind + cap_off[r], sizeof(int64_t) * deg[r]);
      memcpy(counts + nnz, dat + cap_off[r], sizeof(int64_t) * deg[r]);
      nnz += deg[r];
      indptr[r + 1] = nnz;
    }
    *nnz_out = nnz;
    *discarded_out = discarded;
  }
  free(ind); free(dat); free(deg); free(cnt); free(off); free(mem);
  free(q_col); free(q_at); free(tmp_i); free(tmp_d);
  return rc;
}

/* ------------------------------------------------ dense EAC + single link */

/* condensed_index, MyML/MyML/EAC/full.py:168-181 (SciPy's): (i, j), i != j. */
static int64_t orc_cidx(int64_t n, int64_t i, int64_t j) {
  if (i > j) { int64_t t = i; i = j; j = t; }
  return n * i - (i * (i + 1)) / 2 + (j - i - 1);
}

/* EAC_FULL (MyML/MyML/EAC/full.py:8-131) update_ensemble: every partition,
 * every cluster (labels present, ascending), every pair of members i < j
 * (positions) gets +1: condensed (numba_update_condensed_coassoc_with_cluster
 * :155-166) or both (i, j) and (j, i) of the n x n matrix
 * (numba_update_coassoc_with_cluster :141-153). labels: P x n (-1: not
 * listed). out: n(n-1)/2 or n*n counts, zeroed by the caller. */
void orc_coassoc_full(const int64_t* labels, int64_t P, int64_t n, int condensed,
                      int64_t* out) {
  for (int64_t p = 0; p < P; ++p) {
    const int64_t* lab = labels + p * n;
    for (int64_t i = 0; i < n; ++i) {
      if (lab[i] < 0) continue;
      for (int64_t j = i + 1; j < n; ++j) {
        if (lab[j] != lab[i]) continue;
        if (condensed) out[orc_cidx(n, i, j)] += 1;
        else { out[i * n + j] += 1; out[j * n + i] += 1; }
      }
    }
  }
}

/* argsort1D, scipy_numba/scipy_numba/cluster/_argsort.py:6-64: quicksort
 * with the first element as pivot (partition :6-29) and an explicit stack of
 * (left, right) ranges, pushing the left part, then the right part, popping
 * the last pushed. Its order under ties is a property of this procedure, so
 * it is restated step by step. */
static int64_t orc_qs_partition(const double* values, int64_t* idxs, int64_t left,
                                int64_t right) {
  double piv = values[idxs[left]];
  int64_t i = left + 1, j = right;
  for (;;) {
    while (i <= j && values[idxs[i]] <= piv) ++i;
    while (j >= i && values[idxs[j]] >= piv) --j;
    if (j <= i) break;
    int64_t t = idxs[i]; idxs[i] = idxs[j]; idxs[j] = t;
  }
  int64_t t = idxs[left]; idxs[left] = idxs[j]; idxs[j] = t;
  return j;
}

void orc_argsort1d(const double* values, int64_t n, int64_t* idxs) {
  for (int64_t i = 0; i < n; ++i) idxs[i] = i;
  if (n < 2) return;
  int64_t* stk = malloc(sizeof(int64_t) * 2 * (n + 1));
  int64_t nd = 0;
  stk[0] = 0; stk[1] = n - 1; nd = 1;
  while (nd > 0) {
    --nd;
    int64_t left = stk[2 * nd], right = stk[2 * nd + 1];
    int64_t piv = orc_qs_partition(values, idxs, left, right);
    if (piv - 1 > left) { stk[2 * nd] = left; stk[2 * nd + 1] = piv - 1; ++nd; }
    if (piv + 1 < right) { stk[2 * nd] = piv + 1; stk[2 * nd + 1] = right; ++nd; }
  }
  free(stk);
}

/* The SLINK algorithm (R. Sibson, 1973) as
 * scipy_numba/scipy_numba/cluster/_hierarchy_eac.py:884-927 writes it, then
 * from_pointer_representation (:842-881, argsort1D order) and
 * calculate_cluster_sizes (:329-356). dists: condensed, carried in double
 * (the reference's arrays take the weights' dtype, where np.inf of
 * Lambda[0] does not fit an integer type; every integer and float32 weight
 * is exact in double). Z: (n-1) x 4 (left, right, height, size). */
void orc_slink(const double* dists, int64_t n, double* Z) {
  double* M = malloc(sizeof(double) * (n > 0 ? n : 1));
  double* Lambda = malloc(sizeof(double) * (n > 0 ? n : 1));
  int64_t* Pi = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  Pi[0] = 0;
  Lambda[0] = INFINITY;
  for (int64_t i = 1; i < n; ++i) {
    Pi[i] = i;
    Lambda[i] = INFINITY;
    for (int64_t j = 0; j < i; ++j) M[j] = dists[orc_cidx(n, i, j)];
    for (int64_t j = 0; j < i; ++j) {
      if (Lambda[j] >= M[j]) {
        M[Pi[j]] = M[Pi[j]] < Lambda[j] ? M[Pi[j]] : Lambda[j];
        Lambda[j] = M[j];
        Pi[j] = i;
      } else {
        M[Pi[j]] = M[Pi[j]] < M[j] ? M[Pi[j]] : M[j];
      }
    }
    for (int64_t j = 0; j < i; ++j)
      if (Lambda[j] >= Lambda[Pi[j]]) Pi[j] = i;
  }
  int64_t* sorted_idx = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  int64_t* node_ids = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  orc_argsort1d(Lambda, n, sorted_idx);
  for (int64_t i = 0; i < n; ++i) node_ids[i] = i;
  for (int64_t i = 0; i + 1 < n; ++i) {
    int64_t leaf = sorted_idx[i], pi = Pi[leaf];
    if (node_ids[leaf] < node_ids[pi]) {
      Z[4 * i] = (double)node_ids[leaf];
      Z[4 * i + 1] = (double)node_ids[pi];
    } else {
      Z[4 * i] = (double)node_ids[pi];
      Z[4 * i + 1] = (double)node_ids[leaf];
    }
    Z[4 * i + 2] = Lambda[leaf];
    node_ids[pi] = n + i;
  }
  for (int64_t i = 0; i + 1 < n; ++i) {
    int64_t l = (int64_t)Z[4 * i], r = (int64_t)Z[4 * i + 1];
    double cs = 0.0;
    cs += l >= n ? Z[4 * (l - n) + 3] : 1.0;
    cs += r >= n ? Z[4 * (r - n) + 3] : 1.0;
    Z[4 * i + 3] = cs;
  }
  free(M); free(Lambda); free(Pi); free(sorted_idx); free(node_ids);
}

/* labels_from_Z_numba, MyML/MyML/cluster/linkage.py:303-357 (the function
 * eac.py imports as labels_from_Z): every sample starts as its own cluster,
 * the first n - n_clusters rows of Z merge their two clusters into cluster
 * n + row, and the final clusters are numbered in order of first
 * occurrence over the samples. */
void orc_labels_from_Z(const double* Z, int64_t n, int64_t n_clusters, int64_t* labels) {
  int64_t* track = labels;
  for (int64_t i = 0; i < n; ++i) track[i] = i;
  for (int64_t zp = 0; zp != n - n_clusters && zp + 1 < n; ++zp) {
    int64_t c0 = (int64_t)Z[4 * zp], c1 = (int64_t)Z[4 * zp + 1];
    for (int64_t i = 0; i < n; ++i)
      if (track[i] == c0 || track[i] == c1) track[i] = n + zp;
  }
  int64_t* key = malloc(sizeof(int64_t) * (n_clusters > 0 ? n_clusters : 1));
  for (int64_t k = 0; k < n_clusters; ++k) key[k] = -1;
  for (int64_t l = 0; l < n; ++l) {
    int64_t c = track[l], k = 0;
    for (; k < n_clusters; ++k)
      if (key[k] == c || key[k] == -1) break;
    if (k < n_clusters && key[k] == -1) key[k] = c;
    track[l] = k;
  }
  free(key);
}

/* full_sl_lifetime, MyML/MyML/EAC/eac.py:337-364: SLINK on the condensed
 * dissimilarities, the lifetime rule (lifetime_n_clusters :366-391) on the
 * float32 heights Z[:, 2] unless n_clusters is given, labels from Z when
 * more than one cluster remains (empty labels otherwise). Returns
 * ORC_ERR_VALUE where numpy raises (fewer than two heights; no lifetime
 * close to zero). */
int orc_full_sl_lifetime(const double* dists, int64_t n, int64_t n_clusters, double* Z,
                         int64_t* labels, int64_t* nc_out, int64_t* labels_empty) {
  orc_slink(dists, n, Z);
  int64_t nc = n_clusters;
  if (n_clusters == 0) {
    if (n < 3) return ORC_ERR_VALUE;  /* np.argmax of an empty array */
    float best = 0.0f;
    int64_t m = 0;
    int have_zero = 0;
    double minimum = 0.0;
    for (int64_t i = 0; i + 2 < n; ++i) {
      float lt = (float)Z[4 * (i + 1) + 2] - (float)Z[4 * i + 2];
      if (i == 0 || lt > best) { best = lt; m = i; }
      if (fabs((double)lt) <= 1e-8) {
        if (!have_zero || (double)lt < minimum) minimum = (double)lt;
        have_zero = 1;
      }
    }
    if (!have_zero) return ORC_ERR_VALUE;  /* np.min of an empty selection */
    float wm = (float)Z[4 * m + 2];
    int64_t above = 0;
    for (int64_t i = 0; i + 1 < n; ++i) above += (float)Z[4 * i + 2] > wm;
    nc = above == 0 ? 1 : above + 1;
    if ((double)best < 2.0 * minimum) nc = 1;
  }
  *nc_out = nc;
  *labels_empty = nc > 1 ? 0 : 1;
  if (nc > 1) orc_labels_from_Z(Z, n, nc, labels);
  return ORC_OK;
}

/* ------------------------------------------------------------- K-Means */

/* MyML/MyML/cluster/KMeans.py:453-490 py_label / numba_label as the FP64
 * oracle: dist = sum over d ascending of (x_d - c_kd)^2 in double, strict <
 * so the lowest k wins exact ties. data (n, d) and centroids (k, d) are
 * float32 row-major; dists are returned in double. */
void orc_kmeans_label_f64(const float* X, const float* Cc, int64_t n, int64_t d, int64_t k,
                          int32_t* labels, double* dists) {
  for (int64_t i = 0; i < n; ++i) {
    const float* x = X + i * d;
    double best = 0.0;
    int32_t bl = 0;
    for (int64_t c = 0; c < k; ++c) {
      const float* cc = Cc + c * d;
      double dist = 0.0;
      for (int64_t j = 0; j < d; ++j) {
        double diff = (double)x[j] - (double)cc[j];
        dist += diff * diff;
      }
      if (c == 0 || dist < best) { best = dist; bl = (int32_t)c; }
    }
    labels[i] = bl;
    if (dists) dists[i] = best;
  }
}

/* MyML/MyML/utils/sorting.py:8-46 arg_k_select (quickselect over argument
 * indices; Python 2 integer division at :21). out[0:k] = the arguments of
 * the k smallest values in the order the partitioning leaves them. */
void orc_arg_k_select(const float* ary, int64_t size, int64_t k, int64_t* out) {
  int64_t* args = malloc(sizeof(int64_t) * (size > 0 ? size : 1));
  for (int64_t i = 0; i < size; ++i) args[i] = i;
  int64_t fro = 0, to = size - 1;
  while (fro < to) {
    int64_t r = fro, w = to;
    int64_t mid_arg = args[(r + w) / 2];
    float mid = ary[mid_arg];
    while (r < w) {
      if (ary[args[r]] >= mid) {
        int64_t tmp = args[w];
        args[w] = args[r];
        args[r] = tmp;
        w -= 1;
      } else {
        r += 1;
      }
    }
    if (ary[args[r]] > mid) r -= 1;
    if (k <= r) to = r; else fro = r + 1;
  }
  for (int64_t i = 0; i < k; ++i) out[i] = args[i];
  free(args);
}

/* KMeans.py:709-754 py_recompute_centroids / numba_recompute_centroids:
 * float32 sums in point order, then division by the count in double
 * (float32 / int32 in Numba), stored as float32. Empty clusters take, in
 * ascending cluster order, the points returned by arg_k_select(dists, e)
 * (the e SMALLEST dists, despite the comment at KMeans.py:727-728).
 * reseed_mode 0: that quickselect order; 1 (canonical): ascending
 * (dist, index). dists are the float32 distances of the last labeling. */
void orc_recompute_centroids(const float* X, const int32_t* labels, const float* dists,
                             int64_t n, int64_t d, int64_t k, int reseed_mode, float* out) {
  float* acc = calloc(k * d, sizeof(float));
  int64_t* cnt = calloc(k, sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i) cnt[labels[i]]++;
  int64_t ne = 0;
  for (int64_t c = 0; c < k; ++c) ne += cnt[c] == 0;
  int64_t* far = malloc(sizeof(int64_t) * (ne > 0 ? ne : 1));
  if (ne > 0) {
    if (reseed_mode == 0) {
      orc_arg_k_select(dists, n, ne, far);
    } else {
      /* canonical: the ne smallest (dist, index) ascending, by selection */
      uint8_t* used = calloc(n, 1);
      for (int64_t t = 0; t < ne; ++t) {
        int64_t bi = -1;
        for (int64_t i = 0; i < n; ++i) {
          if (used[i]) continue;
          if (bi < 0 || dists[i] < dists[bi]) bi = i;
        }
        far[t] = bi;
        used[bi] = 1;
      }
      free(used);
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    float* a = acc + labels[i] * d;
    const float* x = X + i * d;
    for (int64_t j = 0; j < d; ++j) a[j] = a[j] + x[j];
  }
  int64_t t = 0;
  for (int64_t c = 0; c < k; ++c) {
    if (cnt[c] != 0) {
      for (int64_t j = 0; j < d; ++j)
        out[c * d + j] = (float)((double)acc[c * d + j] / (double)cnt[c]);
    } else {
      int64_t src = far[t++];
      for (int64_t j = 0; j < d; ++j) out[c * d + j] = X[src * d + j];
    }
  }
  free(acc); free(cnt); free(far);
}

/* numpy's float32 summation, as `dists.sum()` computes the inertia at
 * KMeans.py:164 (numpy, the reference's array library, version unpinned).
 * The add reduction starts from the identity 0 and visits the contiguous
 * array in buffer-sized chunks (numpy's default buffer, 8192 elements),
 * adding each chunk's pairwise sum to the running result in order; the
 * pairwise sum is numpy/_core/src/umath/loops_utils.h.src
 * @TYPE@_pairwise_sum with PW_BLOCKSIZE 128 (numpy >= 1.9). Checked
 * bit-for-bit against numpy 2.2's float32 sum. Every addition is a float32
 * addition:
 *   n < 8:    sequential sum from 0;
 *   n <= 128: eight partial sums r[j] over a[j], a[j+8], ... up to the last
 *             multiple of 8, combined as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)),
 *             then the remaining elements added in order;
 *   else:     split at n2 = n/2 rounded down to a multiple of 8, sum of the
 *             two halves. */
static float pairwise_f32(const float* a, int64_t n) {
  if (n < 8) {
    float res = 0.0f;
    for (int64_t i = 0; i < n; ++i) res = res + a[i];
    return res;
  }
  if (n <= 128) {
    float r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    int64_t i;
    for (i = 8; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] = r[j] + a[i + j];
    float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res = res + a[i];
    return res;
  }
  int64_t n2 = n / 2;
  n2 -= n2 % 8;
  return pairwise_f32(a, n2) + pairwise_f32(a + n2, n - n2);
}

#define ORC_NUMPY_BUFSIZE 8192
float orc_sum_f32(const float* a, int64_t n) {
  float res = 0.0f;
  for (int64_t i = 0; i < n; i += ORC_NUMPY_BUFSIZE)
    res = res + pairwise_f32(a + i, n - i < ORC_NUMPY_BUFSIZE ? n - i : ORC_NUMPY_BUFSIZE);
  return res;
}

/* KMeans.py:117-189 K_Means.fit with given initial centroids: label, count
 * the iteration, then (converge mode, KMeans.py:160-175) new_inertia =
 * dists.sum() in float32 (orc_sum_f32), error = |new_inertia - inertia| in
 * float32 (inertia starts at +inf, so the first error is +inf), stop when
 * error <= tol (a float32 comparison: numpy compares a float32 scalar with a
 * Python float in float32) or iters >= max_iter, else recompute centroids.
 * A negative tol never stops early (converge=False). The labels/dists
 * of the last labeling, the final centroids and the last inertia are
 * returned; the iteration count is the return value. */
int64_t orc_kmeans_fit(const float* X, const float* init, int64_t n, int64_t d, int64_t k,
                       int64_t max_iter, double tol, int reseed_mode, int32_t* labels,
                       float* dists, float* centroids, float* inertia_out) {
  float* C = malloc(sizeof(float) * k * d);
  float* Cn = malloc(sizeof(float) * k * d);
  double* dd = malloc(sizeof(double) * (n > 0 ? n : 1));
  memcpy(C, init, sizeof(float) * k * d);
  const float tol32 = (float)tol;
  float inertia = INFINITY;
  int64_t iters = 0;
  for (;;) {
    orc_kmeans_label_f64(X, C, n, d, k, labels, dd);
    for (int64_t i = 0; i < n; ++i) dists[i] = (float)dd[i];
    iters += 1;
    int stop = 0;
    const float s = orc_sum_f32(dists, n);
    const float err = fabsf(s - inertia);
    inertia = s;
    if (err <= tol32) stop = 1;
    if (iters >= max_iter) stop = 1;
    if (stop) break;
    orc_recompute_centroids(X, labels, dists, n, d, k, reseed_mode, Cn);
    float* t = C; C = Cn; Cn = t;
  }
  memcpy(centroids, C, sizeof(float) * k * d);
  if (inertia_out) *inertia_out = inertia;
  free(C); free(Cn); free(dd);
  return iters;
}

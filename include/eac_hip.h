/*
 * eac_hip.h -- C ABI of libeac_hip.so, the MI355X (gfx950) implementation of
 * the evidence-accumulation-clustering (EAC) hot path of diogo-aos/masters_final.
 *
 * Conventions (every entry point):
 *   - returns int status: EAC_OK (0) or a negative EAC_ERR_* code; the message
 *     is available from eac_last_error() (thread-local, valid until the next
 *     failing call on the same host thread). No C++ exception crosses the ABI.
 *   - all array arguments are DEVICE pointers owned by the caller, unless the
 *     parameter name starts with h_ (host pointer).
 *   - scratch memory is a caller-provided device workspace of at least the
 *     size returned by the matching *_workspace_bytes() query (no allocation
 *     happens inside the library).
 *   - work is enqueued on `stream` (a hipStream_t; NULL = default stream).
 *     Entry points that return host scalars (h_*) synchronise that stream.
 *   - vertex ids are int32, CSR offsets int64 (the reference uses int32
 *     offsets; its Python shims widen them).
 *
 * The Python reference passes numpy / Numba device arrays; the shims in
 * masters_final_amd/ keep its function names and signatures and call these.
 */
#ifndef EAC_HIP_H
#define EAC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define EAC_OK 0
#define EAC_ERR_ARG (-1)        /* bad argument (size, null pointer, dtype)   */
#define EAC_ERR_HIP (-2)        /* HIP runtime error                          */
#define EAC_ERR_WORKSPACE (-3)  /* workspace smaller than the query returned  */
#define EAC_ERR_KEY_WIDTH (-4)  /* weight range + vertex ids exceed 64 bits   */
#define EAC_ERR_NOT_FOUND (-5)  /* an edge id is not inside the input graph   */
#define EAC_ERR_UNSUPPORTED (-6)
#define EAC_ERR_VALUE (-7)      /* the reference would raise ValueError       */

/* ---- weight dtypes ----------------------------------------------------- */
#define EAC_W_I32 0
#define EAC_W_F32 1
#define EAC_W_U16 2
#define EAC_W_U8 3
#define EAC_W_I64 4
#define EAC_W_I16 5
#define EAC_W_U32 6

const char* eac_last_error(void);
int eac_abi_version(void);

/* ---- kernel profiling (no counterpart in the reference) -----------------
 * When enabled, the dominant kernels are bracketed by HIP events on the
 * stream they run on; eac_profile_read returns their summed duration, the
 * number of launches and the algorithmic bytes those launches had to move
 * (the per-unit byte figures are documented with each slot). Enabling
 * resets the counters. The switch, the records and the totals belong to the
 * calling host thread: other threads' calls are neither timed nor counted. */
#define EAC_PROF_SLOTS 16
#define EAC_PROF_MST_ROUND0 0 /* CSR stream: 8 B/entry + 16 B/vertex + 8 B/edge out */
#define EAC_PROF_MST_ROUND 1  /* list pass: 16 B/key in (key + 2 root lookups) + 8 B/kept key */
#define EAC_PROF_MST_HOOK 2   /* slot scan + pick-tree walks: 8 B per vertex */
#define EAC_PROF_MST_JUMP 3   /* unused (links and root lookups need no jump pass) */
#define EAC_PROF_MST_FINAL 4  /* forest output + labels */
#define EAC_PROF_CC 5         /* union-find over the CSR: 4 B/entry + 8 B/vertex */
#define EAC_PROF_KM_LABEL 6   /* labeling pass; "bytes" = 2 n d sum(k) FLOPs     */
#define EAC_PROF_KM_RESCORE 7 /* FP64 re-scoring; "bytes" = 3 d k FLOPs/point   */
#define EAC_PROF_CO_PLAN 8    /* member lists: 16 B per (partition, point)       */
#define EAC_PROF_CO_COUNT 9   /* co-association count pass                       */
#define EAC_PROF_CO_FILL 10   /* co-association fill: 8 B/nnz + 8 B/row written  */
#define EAC_PROF_CO_STAGE 11  /* one-pass build: staged pass                     */
#define EAC_PROF_CO_COMPACT 12 /* one-pass build: compaction                     */
int eac_profile_enable(int32_t on);
int eac_profile_read(int32_t slot, double* total_ms, int64_t* launches, int64_t* alg_bytes);

/* ---- tuning knobs (no counterpart in the reference) ---------------------
 * Launch and algorithm parameters that change speed, never results; set by
 * the autotuner in masters_final_amd/tune.py (or by hand) for the calls
 * that follow on the calling host thread (each thread has its own
 * settings, so threads never race on them). 0 selects the built-in choice.
 *   EAC_TUNE_LIGHT_X10      light/heavy split of Borůvka: light undirected
 *                           edges per vertex x 10 (default 15; -1 = no split)
 *   EAC_TUNE_FIRST_BATCH    rounds enqueued before the first read-back
 *                           (default 3/4 of bits(n-1), at least 6)
 *   EAC_TUNE_FILTER_BPCU    256-thread blocks per compute unit of the
 *                           edge-list filter (default: as many as are
 *                           resident); -1: one 1024-thread block per unit
 *   EAC_TUNE_CO_WINDOW      co-association LDS window in columns (even,
 *                           <= 32768; default: the row range when n is
 *                           smaller, else 32768)
 *   EAC_TUNE_LEVEL_STEP     growth of the light edges per vertex from one
 *                           weight level of the split to the next (default
 *                           4, up to 4 levels; 1 or -1 = one split only)
 *   EAC_TUNE_CO_UNROLL      co-association member loads in flight per lane
 *                           while counting (4 by default, or 8)
 *   EAC_TUNE_CO_RANGES      co-association row ranges with a row counter
 *                           each (1..8, default 8: one per XCD; 1 below
 *                           65536 rows)
 *   EAC_TUNE_KM_WAVES       waves per SIMD the split-BF16 labeling kernel
 *                           is compiled for (2, 3 or 4; default 3 for
 *                           d <= 64, else unconstrained) */
#define EAC_TUNE_LIGHT_X10 0
#define EAC_TUNE_FIRST_BATCH 1
#define EAC_TUNE_FILTER_BPCU 2
#define EAC_TUNE_CO_WINDOW 3
#define EAC_TUNE_LEVEL_STEP 4
#define EAC_TUNE_CO_UNROLL 5
#define EAC_TUNE_CO_RANGES 6
#define EAC_TUNE_KM_WAVES 7
#define EAC_TUNE_COUNT 8
int eac_tune_set(int32_t knob, int32_t value);
int eac_tune_get(int32_t knob, int32_t* value);

/* ---- exclusive prefix sum ----------------------------------------------
 * out[i] = init + sum(in[0:i]); *d_total (device, may be NULL) = init + sum(in).
 * in == out is allowed (in place).
 * Replaces MyML/MyML/utils/scan.py:144-249 scan_gpu (in-place Blelloch scan
 * returning the device total) and the CPU scans scan.py:41-53
 * exprefixsumNumba / scan.py:56-72 exprefixsumNumbaSingle.
 */
size_t eac_scan_workspace_bytes(int64_t n, int32_t elem_bytes);
int eac_scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n,
                           int32_t init, int32_t* d_total, void* ws,
                           size_t ws_bytes, void* stream);
int eac_scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n,
                           int64_t init, int64_t* d_total, void* ws,
                           size_t ws_bytes, void* stream);
/* int32 counts -> int64 offsets (row offsets of CSRs with nnz > 2^31). */
int eac_scan_exclusive_i32_i64(const int32_t* in, int64_t* out, int64_t n,
                               int64_t init, int64_t* d_total, void* ws,
                               size_t ws_bytes, void* stream);

/* ---- Boruvka minimum spanning forest -----------------------------------
 * Graph in the reference's CSR form: for vertex v its entries are
 * [first_edge[v], first_edge[v] + out_degree[v]) of dest/weight
 * (MyML/MyML/graph/mst.py:143 boruvka_minho_seq(dest, weight, firstEdge,
 * outDegree)). first_edge must be non-decreasing. out_degree may be NULL, in
 * which case first_edge holds n+1 offsets.
 * symmetric = 1: every undirected edge {u,v} is stored in both rows (the
 *   reference's Boruvka input); only the entries with dest > row are read
 *   as edges.
 * symmetric = 0: every entry u != v is an undirected edge (e.g. the
 *   condensed upper-triangular co-association of MyML/MyML/EAC/sparse.py).
 * Ties are broken by the strict total order (weight, min(u,v), max(u,v)), so
 * the forest is unique.
 * Outputs (device, each may be NULL):
 *   mst_edge_ids[n-1]: CSR position of each forest edge, taken in the row of
 *     its smaller endpoint (replaces the `mst` array of mst.py:143-232 and
 *     mst.py:417-564 boruvka_minho_gpu);
 *   mst_weight[n-1]: weight of each forest edge, in the input dtype;
 *   labels[n]: connected-component labels, numbered in order of first
 *     occurrence (vertex 0 gets label 0);
 *   h_out[4] (host): n_mst, n_components, boruvka rounds, max_weight_key_bits.
 */
size_t eac_boruvka_workspace_bytes(int32_t n, int64_t m);
int eac_boruvka_mst(const int32_t* dest, const void* weight,
                    int32_t weight_dtype, const int64_t* first_edge,
                    const int32_t* out_degree, int32_t n, int64_t m,
                    int32_t symmetric, int64_t* mst_edge_ids,
                    void* mst_weight, int32_t* labels, int64_t* h_out,
                    void* ws, size_t ws_bytes, void* stream);

/* ---- connected components ----------------------------------------------
 * labels[v] = component of v numbered in first-occurrence order. Replaces
 * MyML/MyML/graph/connected_components.py:46-137 connected_comps_seq and
 * :176-314 connected_comps_gpu (whose labels are supervertex ids; they equal
 * ours after first-occurrence relabelling). Weights are not needed.
 * h_ncomp (host, may be NULL) receives the number of components.
 */
size_t eac_cc_workspace_bytes(int32_t n, int64_t m);
int eac_connected_components(const int32_t* dest, const int64_t* first_edge,
                             const int32_t* out_degree, int32_t n, int64_t m,
                             int32_t* labels, int64_t* h_ncomp, void* ws,
                             size_t ws_bytes, void* stream);

/* ---- connected components with no allocation ----------------------------
 * No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops). The same labels as
 * eac_connected_components, from the kernels' scratch in a fixed device
 * array of the library (eac_connected_components_noalloc_max_bytes() bytes,
 * 64 MiB; every GPU the library runs on holds its own copy until the
 * process ends). Graphs whose eac_cc_workspace_bytes(n, m) exceeds it are
 * rejected (EAC_ERR_UNSUPPORTED). Nothing is allocated and nothing is
 * synchronised: d_out (device int32[2]) receives the number of components
 * and a non-zero flag when dest holds ids outside [0, n). Calls on one GPU
 * share that GPU's scratch: calls that overlap on one GPU (two streams, or
 * two host threads) race; calls in one stream are ordered. */
size_t eac_connected_components_noalloc_max_bytes(void);
int eac_connected_components_noalloc(const int32_t* dest, const int64_t* first_edge,
                                     const int32_t* out_degree, int32_t n, int64_t m,
                                     int32_t* labels, int32_t* d_out, void* stream);

/* ---- single-link lifetime cut -----------------------------------------
 * Minimum spanning forest (as eac_boruvka_mst), its edges in ascending
 * (weight, min, max) order, the largest-lifetime cut, and the components of
 * the kept forest edges as first-occurrence labels.
 * rule = EAC_CUT_EAC: MyML/MyML/EAC/eac.py:199-267 sp_sl_lifetime with
 *   lifetime_n_clusters (eac.py:366-391); max_dist < 0 means None; n_clusters
 *   = 0 uses the lifetime rule. Forest edges of weight 0 leave the forest
 *   (Kruskal's eliminate_zeros). When the reference returns an empty label
 *   array (one cluster), h_out[1] = 1 and labels are not written.
 * rule = EAC_CUT_SINGLELINK: MyML/MyML/graph/singlelink.py:14-65
 *   sl_mst_lifetime_seq (and the intended GPU twin :69-153); max_dist is
 *   disconnect_weight (< 0: None = largest graph weight); the forest keeps
 *   its a smallest edges when the largest lifetime (first index a) is >=
 *   disconnect_weight - largest forest weight (the reference drops m - a).
 * Ties at the cut follow the (weight, min, max) order (the reference uses
 * numpy's unstable argsort). Returns EAC_ERR_VALUE where the reference raises
 * ValueError. h_out[6] (host): n_components, labels_empty, nc_stable,
 * kept forest edges, forest edges, Boruvka rounds.
 */
#define EAC_CUT_EAC 0
#define EAC_CUT_SINGLELINK 1
size_t eac_sl_lifetime_workspace_bytes(int32_t n, int64_t m);
int eac_sl_lifetime(const int32_t* dest, const void* weight, int32_t weight_dtype,
                    const int64_t* first_edge, const int32_t* out_degree, int32_t n,
                    int64_t m, int32_t symmetric, int32_t rule, double max_dist,
                    int32_t n_clusters, int32_t* labels, int64_t* h_out, void* ws,
                    size_t ws_bytes, void* stream);

/* ---- batched single-link over many independent graphs -----------------
 * G graphs as one disjoint union: graph g owns vertices [h_goff[g],
 * h_goff[g+1]) (h_goff HOST, h_goff[0] = 0, h_goff[G] = N) and their rows of
 * the union CSR (dest holds union ids). Per graph: its own canonical forest,
 * its own lifetime cut (rule, max_dist, n_clusters as eac_sl_lifetime) and
 * labels numbered by first occurrence inside the graph (labels[N]).
 * gout[4*g + {0,1,2,3}] (device int64): components, labels_empty, nc_stable,
 * kept forest edges (-1 where the reference raises ValueError for that
 * graph). h_out[4] (host): total components, forest edges, Boruvka rounds,
 * kept edges. One launch sequence serves the whole batch. */
size_t eac_sl_lifetime_batched_workspace_bytes(int32_t N, int64_t M, int32_t G);
int eac_sl_lifetime_batched(const int32_t* dest, const void* weight, int32_t weight_dtype,
                            const int64_t* first_edge, const int32_t* out_degree, int32_t N,
                            int64_t M, int32_t symmetric, const int64_t* h_goff, int32_t G,
                            int32_t rule, double max_dist, int32_t n_clusters, int32_t* labels,
                            int64_t* gout, int64_t* h_out, void* ws, size_t ws_bytes,
                            void* stream);

/* ---- the batched cut captured once into a hipGraph ----------------------
 * No reference counterpart as an API: the reference runs its GPU MST one
 * graph per host loop iteration (eac.py:199-267 via boruvka_gpu,
 * graph/mst.py:24-176); this is the same batched cut with every host
 * decision moved onto the device, recorded once and replayed per batch.
 * The graph is bound to the buffers given here (dest, weight, first_edge,
 * out_degree, goff, labels, gout, summary, ws): refill them in place (same
 * N, M, G) and call _launch for each batch. goff is a DEVICE array of G+1
 * offsets. summary (device int64[5]) receives: total components, forest
 * edges, Boruvka rounds, kept edges, error flags (1 edge list overflow,
 * 2 dest ids outside [0, N), 4 (unused), 8 key
 * wider than 64 bits, 16 rounds not converged within max_rounds, 32 bad
 * goff); results are valid only when the flags are 0. max_rounds <= 0
 * takes the bound L * (bits(N-1) + 3) that always suffices (L: weight
 * levels of the Borůvka split, 1 to 4 by edges per vertex); a smaller value
 * shortens replays and sets flag 16 when a batch needs more. Launches on
 * one graph must not overlap (they share ws). */
typedef struct eac_sl_graph eac_sl_graph;
int eac_sl_lifetime_batched_graph_create(const int32_t* dest, const void* weight,
                                         int32_t weight_dtype, const int64_t* first_edge,
                                         const int32_t* out_degree, int32_t N, int64_t M,
                                         int32_t symmetric, const int64_t* goff, int32_t G,
                                         int32_t rule, double max_dist, int32_t n_clusters,
                                         int32_t max_rounds, int32_t* labels, int64_t* gout,
                                         int64_t* summary, void* ws, size_t ws_bytes,
                                         eac_sl_graph** out);
int eac_sl_lifetime_batched_graph_launch(eac_sl_graph* graph, void* stream);
/* Borůvka rounds that always suffice for a graph of n vertices and m CSR
 * entries under the current tuning knobs (the max_rounds <= 0 choice of
 * _graph_create); 0 for bad sizes. No counterpart in the reference. */
int eac_mst_round_bound(int32_t n, int64_t m, int32_t symmetric);
int eac_sl_lifetime_batched_graph_rounds(const eac_sl_graph* graph);
int eac_sl_lifetime_batched_graph_destroy(eac_sl_graph* graph);

/* ---- stable LSD radix sort of 64-bit keys (optional 32-bit values) ------
 * Sorts by bits [begin_bit, end_bit). keys_out/vals_out must not alias the
 * inputs. No reference counterpart as an API: it replaces the numpy argsorts
 * of the MST weights (eac.py:232, singlelink.py:27) inside the cuts. */
size_t eac_sort_workspace_bytes(int64_t n, int32_t key_bytes, int32_t val_bytes);
int eac_radix_sort_u64(const uint64_t* keys_in, uint64_t* keys_out,
                       const uint32_t* vals_in, uint32_t* vals_out, int64_t n,
                       int32_t begin_bit, int32_t end_bit, void* ws, size_t ws_bytes,
                       void* stream);

/* ---- K-Means labeling and centroid update ------------------------------
 * X: n x d float32 row-major (the reference transposes to D x N for its
 * kernel, KMeans.py:321; the row-major layout feeds the MFMA fragments
 * directly). P partitions share X; partition p has centroids
 * C[h_koff[p] : h_koff[p+1]] (k_p x d float32, row-major), h_koff a HOST
 * array of P+1 offsets with h_koff[0] = 0. labels/dists are [P][n].
 * Modes:
 *   EAC_KM_F32     plain FP32 direct difference (the plain GPU baseline);
 *   EAC_KM_F64     FP64 direct difference, ascending dimensions, strict < ->
 *                  exactly MyML/MyML/cluster/KMeans.py:453-490 numba_label
 *                  evaluated in FP64 (the oracle);
 *   EAC_KM_BF16X3  split-BF16 MFMA contraction with fp32 accumulation, a
 *                  rigorous per-point error bound, FP64 re-scoring of every
 *                  point whose top-2 margin is within it -> labels and
 *                  distances equal EAC_KM_F64's (d < 16 runs EAC_KM_F64);
 *   EAC_KM_F32_NUMPY numpy's float32 arithmetic of KMeans.py:493-522
 *                  np_label: ((x - c) ** 2).sum(axis=1) with a float32
 *                  difference and square and numpy's pairwise row sum,
 *                  strict < (lowest k on ties); d <= 128.
 * eac_kmeans_label replaces KMeans.py:296-412 _cu_label (kernel
 * KMeans.py:586-629); h_out[0] = points re-scored in FP64.
 * eac_kmeans_update replaces KMeans.py:709-754 numba_recompute_centroids
 * (fp32 sums in point order, / count in double); empty clusters take the
 * points of smallest (dist, index) in ascending cluster order (the
 * reference's quickselect, sorting.py:8-46, returns the same set in its own
 * order); h_out[0] = clusters re-seeded.
 * eac_kmeans_fit replaces KMeans.py:117-189 K_Means.fit with given initial
 * centroids (C_inout), batched: each partition stops on its own when
 * |sum(dists) - previous| <= tol (numpy's float32 sum and float32
 * arithmetic, as eac_sum_f32_rows) or after max_iter labelings. */
#define EAC_KM_F32 0
#define EAC_KM_F64 1
#define EAC_KM_BF16X3 2
#define EAC_KM_F32_NUMPY 3
size_t eac_kmeans_workspace_bytes(int64_t n, int32_t d, int64_t k_total, int32_t P);
int eac_kmeans_label(const float* X, int64_t n, int32_t d, const float* C,
                     const int64_t* h_koff, int32_t P, int32_t mode, int32_t* labels,
                     float* dists, int64_t* h_out, void* ws, size_t ws_bytes, void* stream);
int eac_kmeans_update(const float* X, int64_t n, int32_t d, const int32_t* labels,
                      const float* dists, const int64_t* h_koff, int32_t P, float* C_new,
                      int64_t* h_out, void* ws, size_t ws_bytes, void* stream);
int eac_kmeans_fit(const float* X, int64_t n, int32_t d, float* C_inout,
                   const int64_t* h_koff, int32_t P, int32_t max_iter, double tol,
                   int32_t mode, int32_t* labels, float* dists, int32_t* h_iters,
                   int64_t* h_out, void* ws, size_t ws_bytes, void* stream);

/* eac_kmeans_fit for capture in a HIP graph (the ensemble of
 * MyML/MyML/utils/partition.py:117-122 + KMeans.py:117-189 with the
 * partitions batched and no host round trip): no host read, no host copy,
 * no synchronisation. d_koff = h_koff already on the device. Exactly
 * max_iter labelings run; a partition that meets the stopping rule keeps
 * its centroids (its later labelings repeat its labels) and its count in
 * d_iters (device int32[P]) stops, so labels, dists, centroids and
 * iteration counts equal eac_kmeans_fit's. Flags are re-scored from a
 * device count; empty clusters are reseeded on the device (the same rule).
 * d_info (device int64[2]): flags re-scored, clusters reseeded; d_err
 * (device uint32): 1 = flag list overflow, 2 = more than 2048 empty
 * clusters in one partition (results then invalid: run eac_kmeans_fit). */
size_t eac_kmeans_fit_graph_workspace_bytes(int64_t n, int32_t d, int64_t k_total, int32_t P);
int eac_kmeans_fit_graph(const float* X, int64_t n, int32_t d, float* C_inout,
                         const int64_t* h_koff, const int64_t* d_koff, int32_t P,
                         int32_t max_iter, double tol, int32_t mode, int32_t* labels,
                         float* dists, int32_t* d_iters, int64_t* d_info, uint32_t* d_err,
                         void* ws, size_t ws_bytes, void* stream);

/* numpy's float32 sum of each row of x (rows x n, row-major): out[r] equals
 * x[r].sum() of a contiguous float32 numpy array to the bit (numpy's add
 * reduction: chunks of its 8192-element buffer, each summed pairwise with
 * PW_BLOCKSIZE 128, chunk sums added in order from 0). It is the inertia of
 * MyML/MyML/cluster/KMeans.py:164 (new_inertia = self._dists.sum()), which
 * eac_kmeans_fit's stopping rule computes on the device: err =
 * |new_inertia - previous| in float32, stop when err <= (float)tol. */
size_t eac_sum_f32_workspace_bytes(int64_t n, int64_t rows);
int eac_sum_f32_rows(const float* x, int64_t n, int64_t rows, float* out, void* ws,
                     size_t ws_bytes, void* stream);

/* ---- graph from a list of edge ids -------------------------------------
 * Replaces MyML/MyML/graph/build.py:49-93 getGraphFromEdges_seq and
 * build.py:96-151 getGraphFromEdges_gpu: for each id in edges (CSR
 * positions of (dest, first_edge, out_degree); -1 entries are skipped as in
 * build.py:170-172) the origin is found by binary search (build.py:18-45),
 * both directions are inserted, rows keep list order (origin side first).
 * Outputs: nod[n] int32 degrees, nfe[n] int64 offsets, ndest[2*n_edges],
 * nweight[2*n_edges] (weight elements of weight_bytes each, nullable);
 * h_nitems = entries written. EAC_ERR_NOT_FOUND when an id has no origin
 * (the CPU function returns -1). */
size_t eac_graph_from_edges_workspace_bytes(int32_t n, int64_t n_edges);
int eac_graph_from_edges(const int32_t* dest, const void* weight, int32_t weight_bytes,
                         const int64_t* first_edge, const int32_t* out_degree, int32_t n,
                         int64_t m, const int64_t* edges, int64_t n_edges, int32_t* nod,
                         int64_t* nfe, int32_t* ndest, void* nweight, int64_t* h_nitems,
                         void* ws, size_t ws_bytes, void* stream);

/* ---- sparse co-association (EAC) --------------------------------------
 * labels: [P][n] int32, labels[p][i] in [0, k_p) or < 0 for "not in any
 * cluster of partition p"; h_koff (HOST) [P+1] offsets of the per-partition
 * cluster counts. Replaces MyML/MyML/EAC/sparse.py EAC_CSR.update_ensemble
 * (:154-195) + _condense (:212-225) + make_diassoc (:295-311), exact (no
 * max_assocs cap, int64 offsets).
 * full_rows = 0: condensed upper triangle (col > row, the reference's
 *   condensed=True default); 1: every co-member j != i (symmetric CSR).
 * Two passes sharing one workspace (pass the same ws, unchanged):
 *   eac_coassoc_count -> indptr[n+1] (int64), nnz, largest count;
 *   eac_coassoc_fill  -> indices[nnz] (columns ascending per row) and
 *                        data[nnz] = count, or w_offset - count when
 *                        w_offset != 0 (make_diassoc with max + 1). */
size_t eac_coassoc_workspace_bytes(int64_t n, int32_t P, int64_t k_total);
int eac_coassoc_count(const int32_t* labels, int64_t n, const int64_t* h_koff,
                      int32_t P, int32_t full_rows, int64_t* indptr, int64_t* h_nnz,
                      int32_t* h_max_count, void* ws, size_t ws_bytes, void* stream);
int eac_coassoc_fill(const int32_t* labels, int64_t n, const int64_t* h_koff,
                     int32_t P, int32_t full_rows, const int64_t* indptr, int32_t w_offset,
                     int32_t* indices, int32_t* data, void* ws, size_t ws_bytes,
                     void* stream);
/* eac_coassoc_fill (w_offset 0) that also writes rank[nnz]: for each entry,
 * the position, in the reference's processing order, of the first partition
 * in which the pair co-occurs. The order is the one of
 * MyML/MyML/EAC/sparse.py:154-166 update_ensemble: first the partition with
 * the fewest non-empty clusters (the first such), then the others by index
 * (rank 0 = that first partition). Same workspace as the count pass. */
int eac_coassoc_fill_ranked(const int32_t* labels, int64_t n, const int64_t* h_koff,
                            int32_t P, int32_t full_rows, const int64_t* indptr,
                            int32_t* indices, int32_t* data, int32_t* rank, void* ws,
                            size_t ws_bytes, void* stream);
/* One-pass build of the same matrix (the same reference functions): the
 * pairs are visited once instead of twice.
 *   eac_coassoc_estimate: plan + count pass over rows 0, s, 2s, ...
 *     (s = sample_stride >= 1): h_estimate = estimated nnz (exact for s = 1);
 *   the caller allocates a staging buffer of
 *     eac_coassoc_stage_bytes(n, full_rows, P, pool_entries) bytes: room
 *     for pool_entries staged entries (the estimate plus a margin) beyond
 *     the workgroups' partly used chunks;
 *   eac_coassoc_stage (same ws, unchanged): indptr[n+1], h_nnz, largest
 *     count; h_nnz = -1 when the pool was too small (then build with
 *     eac_coassoc_count + eac_coassoc_fill);
 *   eac_coassoc_compact: indices[nnz], data[nnz] as eac_coassoc_fill writes
 *     them, from the staging buffer (its header holds the layout). */
int eac_coassoc_estimate(const int32_t* labels, int64_t n, const int64_t* h_koff,
                         int32_t P, int32_t full_rows, int64_t sample_stride,
                         int64_t* h_estimate, void* ws, size_t ws_bytes, void* stream);
size_t eac_coassoc_stage_bytes(int64_t n, int32_t full_rows, int32_t P, int64_t pool_entries);
int eac_coassoc_stage(const int32_t* labels, int64_t n, const int64_t* h_koff, int32_t P,
                      int32_t full_rows, int64_t* indptr, void* stage, size_t stage_bytes,
                      int64_t* h_nnz, int32_t* h_max_count, void* ws, size_t ws_bytes,
                      void* stream);
int eac_coassoc_compact(int64_t n, const int64_t* indptr, const void* stage,
                        size_t stage_bytes, int32_t w_offset, int32_t* indices,
                        int32_t* data, void* stream);
/* The one-pass build with no host read, copy or synchronisation, for
 * capture in a HIP graph (the same matrix): d_koff = h_koff on the device
 * (h_koff still gives the sizes); the staging buffer is sized by the caller
 * (eac_coassoc_stage_bytes; no estimate pass); indices/data hold up to
 * cap_nnz entries; w_offset as in eac_coassoc_fill, or -1 for (largest
 * count + 1) taken on the device. d_out (device int64[4]): nnz, largest
 * count, error bits (1: pool too small, 2: nnz > cap_nnz, 4: a label >= k
 * of its partition; then no entry is written), 0. */
int eac_coassoc_graph(const int32_t* labels, int64_t n, const int64_t* h_koff,
                      const int64_t* d_koff, int32_t P, int32_t full_rows, int32_t w_offset,
                      int64_t* indptr, int32_t* indices, int32_t* data, int64_t cap_nnz,
                      void* stage, size_t stage_bytes, int64_t* d_out, void* ws,
                      size_t ws_bytes, void* stream);

/* ---- capacity-capped co-association (the reference's lossy matrix) ------
 * From the exact matrix (indptr, indices, count) and its ranks (above),
 * the matrix the reference's EAC_CSR builds with per-row capacities:
 * cap_off[n+1] are its capacity offsets (indptr_constant sparse.py:368-369
 * or indptr_linear :371-413; condensed rows, full_rows = 0) or, with
 * full_rows = 1, every row holds max_assocs (cap_off unused). A row keeps
 * the first `capacity` columns in the order (rank, column) at their full
 * counts, as the surgical updates (sparse.py:956-1040 condensed, :866-950
 * full) admit new columns only while the row has room and discard every
 * later co-occurrence of the others (counted, nnz_discarded); with full rows
 * a row above id 405109 keeps only the first partition's columns at count 1
 * (sparse.py:877-878). Replaces the capped update_ensemble + _condense.
 * Output: out_indptr[n+1] (int64), out_indices/out_data (room for
 * out_capacity entries; the reference's own allocation, cap_off[n], always
 * suffices), h_out[4] (host): kept entries, discarded co-occurrences,
 * largest kept count, -1. EAC_ERR_VALUE when the first partition alone
 * gives a row more entries than its capacity (the reference then writes
 * past the row into the next one). */
size_t eac_coassoc_cap_workspace_bytes(int64_t n);
int eac_coassoc_cap(const int64_t* indptr, const int32_t* indices, const int32_t* count,
                    const int32_t* rank, int64_t n, int32_t P, int32_t full_rows,
                    const int64_t* cap_off, int64_t max_assocs, int64_t out_capacity,
                    int64_t* out_indptr, int32_t* out_indices, int32_t* out_data,
                    int64_t* h_out, void* ws, size_t ws_bytes, void* stream);

/* ---- dense co-association and single linkage (small n) ------------------
 * Matrix dtypes: EAC_W_U8, EAC_W_U16, EAC_W_I32, EAC_W_F32. condensed = 1:
 * the n(n-1)/2 upper-triangle vector in SciPy's order (pdist/squareform);
 * 0: the n x n matrix (diagonal 0).
 * eac_coassoc_dense replaces MyML/MyML/EAC/full.py:8-166 EAC_FULL
 * update_ensemble (numba_update_coassoc_with_cluster :141-153,
 * numba_update_condensed_coassoc_with_cluster :155-166): out[i, j] = number
 * of partitions with labels[p][i] == labels[p][j] >= 0, in the matrix dtype
 * (uint8 wraps as the reference's += 1 does); *h_max (host, nullable)
 * receives the largest stored value. ws: 256 bytes.
 * eac_dense_diassoc replaces full.py:98-112 make_diassoc (_1d / _2d):
 * mat[i] = val - mat[i] over len elements (integer types wrap).
 * eac_dense_degree replaces full.py:83-95 get_degree: degree[i] = non-zero
 * entries of row i (condensed: its columns j > i only), *h_nnz the total.
 * eac_full_sl_lifetime replaces MyML/MyML/EAC/eac.py:337-364
 * full_sl_lifetime on the dissimilarities: single linkage (the minimum
 * spanning tree of the complete graph), the lifetime rule of eac.py:366-391
 * on the float32 heights unless n_clusters > 0, labels of the clusters
 * numbered by first occurrence (labels_from_Z_numba,
 * MyML/MyML/cluster/linkage.py:303-357); h_out[2]: clusters, labels_empty
 * (one cluster: labels not written). Z ((n-1) x 4 float32, nullable): the
 * linkage matrix. EAC_ERR_VALUE where numpy raises (no lifetime within
 * 1e-8 of zero, or fewer than two heights).
 * eac_slink replaces MyML/MyML/cluster/linkage.py:144-149 (SL: SLINK,
 * scipy_numba/.../_hierarchy_eac.py:884-927, and from_pointer_representation
 * :842-881): Z rows (left, right, height, size), heights ascending, equal
 * heights in (height, min id, max id) order of their tree edges.
 * eac_labels_from_Z replaces linkage.py:303-357 labels_from_Z_numba for any
 * valid Z: the first n - n_clusters merges, clusters numbered by first
 * occurrence. n <= 65536 for the single-linkage calls. */
int eac_coassoc_dense(const int32_t* labels, int64_t n, int32_t P, int32_t condensed,
                      int32_t dtype, void* out, int64_t* h_max, void* ws, size_t ws_bytes,
                      void* stream);
int eac_dense_diassoc(void* mat, int64_t len, int32_t dtype, double val, void* stream);
int eac_dense_degree(const void* mat, int32_t dtype, int64_t n, int32_t condensed,
                     int32_t* degree, int64_t* h_nnz, void* ws, size_t ws_bytes, void* stream);
size_t eac_dense_sl_workspace_bytes(int64_t n, int32_t dtype, int32_t condensed);
int eac_full_sl_lifetime(const void* dists, int32_t dtype, int64_t n, int32_t condensed,
                         int32_t n_clusters, float* Z, int32_t* labels, int64_t* h_out,
                         void* ws, size_t ws_bytes, void* stream);
int eac_slink(const void* dists, int32_t dtype, int64_t n, int32_t condensed, float* Z,
              void* ws, size_t ws_bytes, void* stream);
size_t eac_labels_from_Z_workspace_bytes(int64_t n);
int eac_labels_from_Z(const float* Z, int64_t n, int32_t n_clusters, int32_t* labels, void* ws,
                      size_t ws_bytes, void* stream);

/* ---- first-occurrence relabelling --------------------------------------
 * out[v] = rank of labels_in[v] among distinct values ordered by first
 * occurrence. Requires 0 <= labels_in[v] < n_values.
 */
size_t eac_relabel_workspace_bytes(int32_t n, int32_t n_values);
int eac_canonical_relabel(const int32_t* labels_in, int32_t n,
                          int32_t n_values, int32_t* out, int64_t* h_count,
                          void* ws, size_t ws_bytes, void* stream);

/* ---- binary cluster-by-sample matrix of a partition ---------------------
 * out[k][N] (uint8, 16-byte aligned) is cleared, then out[rows[t]][cols[t]]
 * = 1 for t < m (cols NULL: cols[t] = t, m <= N). Replaces
 * MyML/MyML/utils/partition.py:17-35 convertIndexToBin (rows = the cluster
 * of each listed index, cols = the indices, N = the number listed) and
 * partition.py:37-70 convertClusterStringToBin (rows = label + offset, one
 * column per sample). Entries outside [0, k) x [0, N) are skipped and
 * counted into *d_bad (device int64, nullable), where numpy raises
 * IndexError. */
int eac_partition_to_bin(const int32_t* rows, const int64_t* cols, int64_t m, int32_t k,
                         int64_t N, uint8_t* out, int64_t* d_bad, void* stream);

/* ---- contingency table of two partitions (accuracy indices) -------------
 * out[(a[s] - a0) * kb + (b[s] - b0)] = number of samples s with those two
 * labels; out[ka * kb] = samples whose labels fall outside
 * [a0, a0 + ka) x [b0, b0 + kb) (the caller rejects those). out has
 * ka * kb + 1 entries and is cleared by the call. For partitions this is the
 * product of the binary cluster-by-sample matrices that
 * MyML/MyML/validation/consistency.py:70-125 (ConsistencyIndex._match_: A,
 * and B, C as its row and column sums) and
 * MyML/MyML/validation/hindex.py:44 (HungarianIndex: match_matrix) compute
 * with dense dot products. */
int eac_contingency(const int32_t* a, const int32_t* b, int64_t n, int32_t a0, int32_t ka,
                    int32_t b0, int32_t kb, int64_t* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* EAC_HIP_H */

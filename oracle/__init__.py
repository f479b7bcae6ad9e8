"""CPU oracle of the EAC hot path -- TEST INFRASTRUCTURE ONLY.

`eac_oracle.c` restates the reference's CPU path (diogo-aos/masters_final,
read as source text) in plain C; `oracle.py` wraps it with numpy. Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, as the checker or as the timed CPU baseline; the product in
masters_final_amd/ never does.

Parity pinning of each oracle function (what independent check fixes it):
  scan .............. closed form (sequential sum), reference test patterns
  boruvka_minho_seq . MST weight/edge count vs scipy minimum_spanning_tree on
                      4elt (2,964,801 / 15,605) and random graphs; brute force
                      over all spanning trees of tiny graphs
  kruskal_canonical . scipy MST weight; brute force on tiny graphs (unique
                      forest under the strict (w, min, max) order)
  components ........ scipy.sparse.csgraph.connected_components labels
  graph_from_edges .. closed form (row multisets), 4elt round trip
  sp_sl_lifetime .... closed form on hand cases; 4elt known answer (6
                      clusters, sizes 15600/2/1/1/1/1); CC-of-threshold
                      invariant vs scipy
  sl_mst_lifetime ... hand cases derived from singlelink.py:14-65 text
  coassoc ........... dense brute force (np.add.at over cluster pairs)
  kmeans_label_f64 .. numpy float64 direct difference
  recompute_centroids closed form (numpy float32 cumulative sums)
  arg_k_select ...... set of returned indices == argsort smallest k
"""

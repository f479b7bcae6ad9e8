// Graph rebuilt from a list of edge ids of a CSR graph, on gfx950.
//
// Reference: MyML/MyML/graph/build.py:49-93 getGraphFromEdges_seq (origin of
// each edge by binaryOriginVertexSearch build.py:18-45, degree count,
// exclusive scan, both directions inserted in list order) and its GPU twin
// build.py:96-151 (countEdges :155-179 / addEdges :181-210 with atomic slot
// order; ids of -1 are skipped, :170-172).
//
// Here the within-row order is the sequential one of the CPU function for
// every input: each edge contributes two items (origin side first), a stable
// radix sort by row keeps list order inside rows, and the sorted position is
// the output slot. Degrees come from a histogram and offsets from a scan.
// For a self-loop both items land in the vertex's row (the reference writes
// one slot twice and leaves the next one unset).
#include <cstring>

#include "prims.hpp"
#include "sort.hpp"

namespace eac {
namespace {

__device__ __forceinline__ int32_t origin_search(int64_t key, const int64_t* fe,
                                                 const int32_t* od, int32_t n) {
  int32_t imin = 0, imax = n;
  while (imin < imax) {
    int32_t imid = (int32_t)(((int64_t)imax + imin) / 2);
    int64_t f = fe[imid];
    int64_t end = od ? f + od[imid] : fe[imid + 1];
    if (key < f) imax = imid;
    else if (key > end - 1) imin = imid + 1;
    else return imid;
  }
  return -1;
}

__global__ void k_gfe_items(const int64_t* __restrict__ edges, int64_t ne,
                            const int32_t* __restrict__ dest, const int64_t* __restrict__ fe,
                            const int32_t* __restrict__ od, int32_t n, int64_t m,
                            uint32_t* __restrict__ key, uint32_t* __restrict__ item,
                            int32_t* __restrict__ org, int32_t* __restrict__ dst,
                            int32_t* __restrict__ deg, unsigned int* __restrict__ bad) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t id = edges[e];
    uint32_t k0 = (uint32_t)n, k1 = (uint32_t)n;
    if (id != -1) {
      int32_t o = (id >= 0 && id < m) ? origin_search(id, fe, od, n) : -1;
      int32_t v = o >= 0 ? dest[id] : -1;
      if (o < 0 || v < 0 || v >= n) {
        atomicOr(bad, 1u);
      } else {
        org[e] = o;
        dst[e] = v;
        k0 = (uint32_t)o;
        k1 = (uint32_t)v;
        atomicAdd(&deg[o], 1);
        atomicAdd(&deg[v], 1);
      }
    }
    key[2 * e] = k0;
    key[2 * e + 1] = k1;
    item[2 * e] = (uint32_t)(2 * e);
    item[2 * e + 1] = (uint32_t)(2 * e + 1);
  }
}

template <typename T>
__global__ void k_gfe_scatter(const uint32_t* __restrict__ skey,
                              const uint32_t* __restrict__ sitem, int64_t nitems, int32_t n,
                              const int64_t* __restrict__ edges,
                              const int32_t* __restrict__ org, const int32_t* __restrict__ dst,
                              const T* __restrict__ weight, int32_t* __restrict__ ndest,
                              T* __restrict__ nweight) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nitems;
       p += (int64_t)gridDim.x * blockDim.x) {
    if (skey[p] == (uint32_t)n) continue;  // skipped or invalid edges sort last
    uint32_t it = sitem[p];
    int64_t e = it >> 1;
    ndest[p] = (it & 1u) ? org[e] : dst[e];
    if (nweight) nweight[p] = weight[edges[e]];
  }
}

int d2h(void* host, const void* dev, size_t bytes, hipStream_t s) {
  void* pin = pinned_scratch(bytes);
  if (!pin) return fail(EAC_ERR_HIP, "pinned host allocation failed");
  EAC_HIP_CHECK(hipMemcpyAsync(pin, dev, bytes, hipMemcpyDeviceToHost, s));
  EAC_HIP_CHECK(hipStreamSynchronize(s));
  memcpy(host, pin, bytes);
  return EAC_OK;
}

}  // namespace
}  // namespace eac

using namespace eac;

extern "C" size_t eac_graph_from_edges_workspace_bytes(int32_t n, int64_t n_edges) {
  if (n < 1) n = 1;
  if (n_edges < 1) n_edges = 1;
  Arena ar(nullptr, 0, true);
  int64_t items = 2 * n_edges;
  ar.take<uint32_t>(items);
  ar.take<uint32_t>(items);
  ar.take<uint32_t>(items);
  ar.take<uint32_t>(items);
  ar.take<int32_t>(n_edges);
  ar.take<int32_t>(n_edges);
  ar.take<unsigned int>(4);
  size_t a = radix_sort_ws_bytes(items, 4, 4);
  size_t b = scan_ws_bytes(n + 1, 8);
  return ar.used + (a > b ? a : b) + 4096;
}

extern "C" int eac_graph_from_edges(const int32_t* dest, const void* weight,
                                    int32_t weight_bytes, const int64_t* first_edge,
                                    const int32_t* out_degree, int32_t n, int64_t m,
                                    const int64_t* edges, int64_t n_edges, int32_t* nod,
                                    int64_t* nfe, int32_t* ndest, void* nweight,
                                    int64_t* h_nitems, void* ws, size_t ws_bytes,
                                    void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n < 1 || m < 0 || n_edges < 0) return fail(EAC_ERR_ARG, "graph_from_edges: bad sizes");
  if (!first_edge || !nod || !nfe || (n_edges > 0 && (!edges || !dest || !ndest)))
    return fail(EAC_ERR_ARG, "graph_from_edges: null array");
  if (nweight && weight_bytes != 1 && weight_bytes != 2 && weight_bytes != 4 && weight_bytes != 8)
    return fail(EAC_ERR_ARG, "graph_from_edges: weight element size must be 1, 2, 4 or 8");
  if (2 * n_edges >= (int64_t)UINT32_MAX) return fail(EAC_ERR_UNSUPPORTED, "too many edges");
  size_t need = eac_graph_from_edges_workspace_bytes(n, n_edges);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "graph_from_edges: workspace %zu < %zu", ws_bytes, need);
  Arena ar(ws, ws_bytes);
  const int64_t items = 2 * n_edges;
  uint32_t* key = ar.take<uint32_t>(items);
  uint32_t* item = ar.take<uint32_t>(items);
  uint32_t* skey = ar.take<uint32_t>(items);
  uint32_t* sitem = ar.take<uint32_t>(items);
  int32_t* org = ar.take<int32_t>(n_edges);
  int32_t* dst = ar.take<int32_t>(n_edges);
  unsigned int* bad = ar.take<unsigned int>(4);
  const size_t mark = ar.used;
  EAC_TRY(fill<int32_t>(nod, n, 0, s));
  EAC_TRY(fill<int32_t>((int32_t*)bad, 4, 0, s));
  if (n_edges > 0) {
    k_gfe_items<<<grid_for(n_edges, 256, 8192), 256, 0, s>>>(edges, n_edges, dest, first_edge,
                                                             out_degree, n, m, key, item, org,
                                                             dst, nod, bad);
    EAC_LAUNCH_CHECK();
    int bits = bits_for((uint64_t)n);
    EAC_TRY((radix_sort<uint32_t, uint32_t>(key, skey, item, sitem, items, 0, bits < 1 ? 1 : bits,
                                            ar, s)));
    ar.used = mark;
  }
  int64_t* d_tot = reinterpret_cast<int64_t*>(bad + 2);
  EAC_TRY((scan_excl<int32_t, int64_t>(nod, nfe, n, 0, d_tot, ar, s)));
  if (n_edges > 0) {
    int g = grid_for(items, 256, 8192);
    switch (nweight ? weight_bytes : 0) {
      case 1: k_gfe_scatter<uint8_t><<<g, 256, 0, s>>>(skey, sitem, items, n, edges, org, dst, (const uint8_t*)weight, ndest, (uint8_t*)nweight); break;
      case 2: k_gfe_scatter<uint16_t><<<g, 256, 0, s>>>(skey, sitem, items, n, edges, org, dst, (const uint16_t*)weight, ndest, (uint16_t*)nweight); break;
      case 8: k_gfe_scatter<uint64_t><<<g, 256, 0, s>>>(skey, sitem, items, n, edges, org, dst, (const uint64_t*)weight, ndest, (uint64_t*)nweight); break;
      default: k_gfe_scatter<uint32_t><<<g, 256, 0, s>>>(skey, sitem, items, n, edges, org, dst, (const uint32_t*)weight, ndest, (uint32_t*)nweight); break;
    }
    EAC_LAUNCH_CHECK();
  }
  unsigned int hb[4];
  EAC_TRY(d2h(hb, bad, sizeof(hb), s));
  if (hb[0]) return fail(EAC_ERR_NOT_FOUND, "graph_from_edges: an edge id has no origin vertex");
  int64_t tot;
  memcpy(&tot, &hb[2], sizeof(tot));
  if (h_nitems) *h_nitems = tot;
  return EAC_OK;
}

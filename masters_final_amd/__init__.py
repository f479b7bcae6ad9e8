"""MI355X-native EAC hot path (K-Means labeling, CSR co-association, Boruvka MST,
connected components, single-link lifetime cut) behind the reference's Python API.

Sub-packages mirror the reference's module layout (MyML/MyML/...):
  utils.scan, graph.mst, graph.connected_components, graph.build,
  graph.singlelink, EAC.eac, EAC.sparse, cluster.KMeans.
All compute runs in libeac_hip.so (HIP, gfx950) through a ctypes C ABI.
"""
__version__ = "0.1.0"

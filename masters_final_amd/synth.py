"""Seeded synthetic inputs for the EAC hot-path workloads (numpy PCG64).

  eac_small()        n=10,000 2-D points from a 6-component Gaussian mixture
  knn_graph()        symmetrized 3-D k-NN graph with int32 weights (the
                     1,000,000-vertex MST/CC workload and its smaller cases)
  gmm_points()       Gaussian-mixture points for the K-Means workloads
  random_graph()     small random symmetric CSR graphs for parity tests
  graph_batch()      many independent graphs as one disjoint-union CSR (host)
  knn_batch_device() the 512-graph batched single-link workload: k-NN graphs
                     found on the host, symmetrized and laid out on the GPU
"""
from __future__ import annotations

import numpy as np


def csr_from_undirected(n, lo, hi, w):
    """Symmetric CSR (both directions, rows sorted by column) from unique
    undirected edges lo < hi. Returns (indptr int64[n+1], dest int32, weight)."""
    rows = np.concatenate([lo, hi]).astype(np.int64)
    cols = np.concatenate([hi, lo]).astype(np.int64)
    ww = np.concatenate([w, w])
    order = np.argsort(rows * n + cols, kind="stable")
    rows, cols, ww = rows[order], cols[order], ww[order]
    indptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=indptr[1:])
    return indptr, cols.astype(np.int32), ww


def knn_graph(n=1_000_000, k=20, seed=1, frac_islands=0.10, n_islands=4, wmin=1, wmax=100,
              dtype=np.int32):
    """3-D uniform points, exact k-NN (k neighbours besides the point itself),
    symmetrized as a union; `frac_islands` of the points form `n_islands`
    separate clusters translated by +10 per island along x, so their k-NN
    stays inside. Weights uniform integers in [wmin, wmax] per undirected
    edge. Returns dict(indptr, dest, weight, n, n_undirected)."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    pts = rng.random((n, 3))
    n_isl = int(round(n * frac_islands))
    if n_islands > 0 and n_isl >= n_islands * (k + 1):
        owner = rng.permutation(n)[:n_isl]
        grp = np.arange(n_isl) % n_islands
        pts[owner, 0] += 10.0 * (grp + 1)
    kk = min(k + 1, n)
    tree = cKDTree(pts)
    _, nb = tree.query(pts, k=kk, workers=-1)
    nb = nb[:, 1:].astype(np.int64)
    src = np.repeat(np.arange(n, dtype=np.int64), nb.shape[1])
    dst = nb.ravel()
    lo = np.minimum(src, dst)
    hi = np.maximum(src, dst)
    key = np.unique(lo * n + hi)
    lo, hi = key // n, key % n
    w = rng.integers(wmin, wmax + 1, size=lo.size).astype(dtype)
    indptr, dest, weight = csr_from_undirected(n, lo, hi, w)
    return {"indptr": indptr, "dest": dest, "weight": weight, "n": n,
            "n_undirected": int(lo.size)}


def random_graph(n, avg_deg, seed, wmax=10, n_isolated=0, n_parts=1, dtype=np.int32):
    """Small random symmetric graph: Erdos-Renyi edges inside `n_parts`
    vertex blocks (so the graph has at least n_parts components), a few
    isolated vertices, integer weights in [1, wmax] (many ties)."""
    rng = np.random.default_rng(seed)
    m_target = max(int(n * avg_deg / 2), 1)
    part = rng.integers(0, n_parts, size=n)
    iso = rng.permutation(n)[:n_isolated] if n_isolated else np.zeros(0, np.int64)
    u = rng.integers(0, n, size=3 * m_target)
    v = rng.integers(0, n, size=3 * m_target)
    ok = (u != v) & (part[u] == part[v])
    if iso.size:
        bad = np.zeros(n, bool)
        bad[iso] = True
        ok &= ~bad[u] & ~bad[v]
    u, v = u[ok], v[ok]
    lo, hi = np.minimum(u, v), np.maximum(u, v)
    key = np.unique(lo * n + hi)[:m_target]
    lo, hi = key // n, key % n
    w = rng.integers(1, wmax + 1, size=lo.size).astype(dtype)
    indptr, dest, weight = csr_from_undirected(n, lo, hi, w)
    return {"indptr": indptr, "dest": dest, "weight": weight, "n": n,
            "n_undirected": int(lo.size)}


def gmm_points(n, d, n_comp, seed, mean_scale=None, mean_std=None, std=1.0, anisotropic=None):
    """Gaussian mixture, equal weights. Means uniform in [-mean_scale, mean_scale]^d
    or N(0, mean_std^2 I); isotropic `std`, or anisotropic covariances
    Q diag(lam) Q^T with lam ~ U[a, b] when anisotropic=(a, b). float32."""
    rng = np.random.default_rng(seed)
    if mean_std is not None:
        means = rng.normal(0.0, mean_std, size=(n_comp, d))
    else:
        s = 10.0 if mean_scale is None else mean_scale
        means = rng.uniform(-s, s, size=(n_comp, d))
    comp = rng.integers(0, n_comp, size=n)
    X = np.empty((n, d), np.float64)
    if anisotropic is None:
        X[:] = means[comp] + std * rng.standard_normal((n, d))
    else:
        a, b = anisotropic
        for c in range(n_comp):
            idx = np.nonzero(comp == c)[0]
            q, r = np.linalg.qr(rng.standard_normal((d, d)))
            q = q * np.sign(np.diag(r))
            lam = rng.uniform(a, b, size=d)
            L = q * np.sqrt(lam)
            X[idx] = means[c] + rng.standard_normal((idx.size, d)) @ L.T
    return X.astype(np.float32), comp


def eac_small(seed=0):
    """n=10,000 points in 2-D, 6-component GMM, means U[-10,10]^2, unit std."""
    X, y = gmm_points(10_000, 2, 6, seed, mean_scale=10.0, std=1.0)
    return X, y


def graph_batch(n_graphs, v_range, deg_range, seed, wmax=256, kind="knn", dtype=np.int32):
    """Independent graphs as one disjoint union CSR: vertex counts log-uniform
    in v_range, mean degree uniform in deg_range (k-NN with k = 0.75 * target,
    or Erdos-Renyi for kind="er"), integer weights U[1, wmax]. Returns
    dict(indptr, dest, weight, goff, n, targets)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(v_range[0]), np.log(v_range[1])
    nv = np.exp(rng.uniform(lo, hi, n_graphs)).astype(np.int64)
    deg = rng.uniform(deg_range[0], deg_range[1], n_graphs)
    parts = []
    for g in range(n_graphs):
        sub_seed = int(rng.integers(0, 2**31 - 1))
        if kind == "knn":
            k = max(1, int(round(0.75 * deg[g])))
            gg = knn_graph(n=int(nv[g]), k=min(k, int(nv[g]) - 1), seed=sub_seed, frac_islands=0.0,
                           n_islands=0, wmin=1, wmax=wmax, dtype=dtype)
        else:
            gg = random_graph(int(nv[g]), deg[g], sub_seed, wmax=wmax, dtype=dtype)
        parts.append(gg)
    goff = np.zeros(n_graphs + 1, np.int64)
    np.cumsum([p["n"] for p in parts], out=goff[1:])
    eoff = np.zeros(n_graphs + 1, np.int64)
    np.cumsum([p["dest"].size for p in parts], out=eoff[1:])
    dest = np.concatenate([p["dest"].astype(np.int64) + goff[i] for i, p in enumerate(parts)])
    weight = np.concatenate([p["weight"] for p in parts])
    indptr = np.concatenate([p["indptr"][:-1] + eoff[i] for i, p in enumerate(parts)] +
                            [np.array([eoff[-1]])])
    return {"indptr": indptr.astype(np.int64), "dest": dest.astype(np.int32),
            "weight": weight, "goff": goff, "n": int(goff[-1]), "targets": deg}


def graph_batch_device(n_graphs, v_range, deg_range, seed, wmax=256, device="cuda"):
    """Large disjoint-union batches generated on the GPU (torch): graph g has
    a vertex count log-uniform in v_range and n_g * deg_g / 2 undirected
    Erdos-Renyi edges (deg_g uniform in deg_range; self loops and repeats
    dropped), int32 weights U[1, wmax] per undirected edge, stored in both
    rows of a symmetric CSR sorted by (row, column). Sizes come from numpy
    PCG64(seed), edges from a torch CUDA generator seeded with `seed`.
    Returns dict(indptr int64, dest int32, weight int32 (device tensors),
    goff (host int64), n, targets)."""
    import torch

    rng = np.random.default_rng(seed)
    lo, hi = np.log(v_range[0]), np.log(v_range[1])
    nv = np.exp(rng.uniform(lo, hi, n_graphs)).astype(np.int64)
    deg = rng.uniform(deg_range[0], deg_range[1], n_graphs)
    me = np.maximum((nv * deg / 2).astype(np.int64), 1)
    goff = np.zeros(n_graphs + 1, np.int64)
    np.cumsum(nv, out=goff[1:])
    N = int(goff[-1])
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    dev = torch.device(device)
    g_of = torch.repeat_interleave(torch.arange(n_graphs, device=dev),
                                   torch.from_numpy(me).to(dev))
    base = torch.from_numpy(goff[:-1]).to(dev)[g_of]
    size = torch.from_numpy(nv).to(dev)[g_of]
    u = base + (torch.rand(g_of.numel(), generator=gen, device=dev, dtype=torch.float64)
                * size).to(torch.int64).clamp_max(size - 1)
    v = base + (torch.rand(g_of.numel(), generator=gen, device=dev, dtype=torch.float64)
                * size).to(torch.int64).clamp_max(size - 1)
    del g_of, base, size
    keep = u != v
    a = torch.minimum(u, v)[keep]
    b = torch.maximum(u, v)[keep]
    del u, v, keep
    key = torch.unique(a * N + b)  # sorted
    del a, b
    a, b = key // N, key % N
    del key
    w = torch.randint(1, wmax + 1, (a.numel(),), generator=gen, device=dev, dtype=torch.int32)
    rows = torch.cat([a, b])
    cols = torch.cat([b, a])
    ww = torch.cat([w, w])
    del a, b, w
    order = torch.sort(rows * N + cols).indices
    rows, cols, ww = rows[order], cols[order], ww[order]
    del order
    indptr = torch.zeros(N + 1, dtype=torch.int64, device=dev)
    indptr[1:] = torch.cumsum(torch.bincount(rows, minlength=N), 0)
    del rows
    return {"indptr": indptr, "dest": cols.to(torch.int32), "weight": ww, "goff": goff,
            "n": N, "targets": deg}


def _batch_sizes(n_graphs, v_range, deg_range, rng):
    lo, hi = np.log(v_range[0]), np.log(v_range[1])
    nv = np.exp(rng.uniform(lo, hi, n_graphs)).astype(np.int64)
    deg = rng.uniform(deg_range[0], deg_range[1], n_graphs)
    return nv, deg


def _csr_on_device(a, b, w, N, torch):
    """Symmetric CSR sorted by (row, column) from unique undirected device
    edges a < b with weights w (both directions stored)."""
    rows = torch.cat([a, b])
    cols = torch.cat([b, a])
    ww = torch.cat([w, w])
    del a, b, w
    order = torch.sort(rows * N + cols).indices
    rows, cols, ww = rows[order], cols[order], ww[order]
    del order
    indptr = torch.zeros(N + 1, dtype=torch.int64, device=rows.device)
    indptr[1:] = torch.cumsum(torch.bincount(rows, minlength=N), 0)
    del rows
    return indptr, cols.to(torch.int32), ww


def knn_batch_device(n_graphs, v_range, deg_range, seed, wmax=256, device="cuda"):
    """Independent 3-D k-NN graphs as one disjoint-union CSR on the GPU (the
    batched single-link workload): graph g has V_g points uniform in [0,1)^3
    (V_g log-uniform in v_range), k_g = round(0.75 * deg_g) exact nearest
    neighbours (deg_g uniform in deg_range; cKDTree on the host, as
    knn_graph), symmetrized as a union, int32 weights U[1, wmax] per
    undirected edge from a torch CUDA generator seeded with `seed`. Sizes and
    points come from numpy PCG64(seed). Returns dict(indptr int64, dest
    int32, weight int32 (device tensors), goff (host int64), n, targets,
    knn_k)."""
    import torch
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    nv, deg = _batch_sizes(n_graphs, v_range, deg_range, rng)
    goff = np.zeros(n_graphs + 1, np.int64)
    np.cumsum(nv, out=goff[1:])
    N = int(goff[-1])
    dev = torch.device(device)
    ks = np.zeros(n_graphs, np.int64)
    keys = []
    for g in range(n_graphs):
        V = int(nv[g])
        pts = np.random.default_rng(int(rng.integers(0, 2**31 - 1))).random((V, 3))
        k = min(max(1, int(round(0.75 * deg[g]))), V - 1)
        ks[g] = k
        _, nb = cKDTree(pts).query(pts, k=k + 1, workers=-1)
        dst = torch.from_numpy(nb[:, 1:].astype(np.int32)).to(dev).reshape(-1).to(torch.int64)
        src = torch.arange(V, device=dev, dtype=torch.int64).repeat_interleave(k)
        lo = torch.minimum(src, dst) + int(goff[g])
        hi = torch.maximum(src, dst) + int(goff[g])
        keys.append(lo * N + hi)
        del src, dst, lo, hi
    key = torch.unique(torch.cat(keys))  # sorted
    del keys
    a, b = key // N, key % N
    del key
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    w = torch.randint(1, wmax + 1, (a.numel(),), generator=gen, device=dev, dtype=torch.int32)
    indptr, dest, weight = _csr_on_device(a, b, w, N, torch)
    return {"indptr": indptr, "dest": dest, "weight": weight, "goff": goff, "n": N,
            "targets": deg, "knn_k": ks}

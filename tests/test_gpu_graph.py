"""GPU parity of the scan, Boruvka forest, connected components and relabel
kernels against the CPU oracle (bit-exact: integer/index work)."""
import numpy as np
import pytest

from masters_final_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def _edges_of_ids(indptr_or_fe, od, dest, ids):
    fe = np.asarray(indptr_or_fe)[: len(od)]
    rows = np.repeat(np.arange(len(od)), od)
    pos_row = rows
    return {(min(pos_row[e], dest[e]), max(pos_row[e], dest[e])) for e in ids}


# -------------------------------------------------------------------- scan
@pytest.mark.parametrize("n", [1, 2, 1023, 1024, 4095, 4096, 4097, 100_000, 2_000_000,
                               17_000_001])
@pytest.mark.parametrize("dt", ["int32", "int64"])
def test_scan_matches_closed_form(torch, n, dt):
    from masters_final_amd.utils.scan import scan_gpu
    rng = np.random.default_rng(n)
    hi = 10 if dt == "int32" and n > 1_000_000 else 10000
    a = rng.integers(0, hi, n).astype(dt)
    d = torch.from_numpy(a).cuda()
    tot = scan_gpu(d)
    want, wtot = O.exprefixsum(a)
    assert int(tot.item()) == wtot
    assert np.array_equal(d.cpu().numpy().astype(np.int64), want)


def test_scan_host_array_and_init(torch):
    from masters_final_amd.utils.scan import exprefixsumNumba, exprefixsumNumbaSingle, scan_gpu
    a = np.arange(5000, dtype=np.int32)
    b = a.copy()
    tot = scan_gpu(b)  # host input: written back in place
    want, wt = O.exprefixsum(a)
    assert tot == wt and np.array_equal(b, want)
    out = np.empty_like(a)
    assert exprefixsumNumba(a, out, init=11) == wt + 11
    assert np.array_equal(out, want + 11)
    c = a.copy()
    assert exprefixsumNumbaSingle(c, init=3) == wt + 3 and np.array_equal(c, want + 3)
    with pytest.raises(TypeError):
        scan_gpu(np.zeros(4, np.float32))


def test_scan_widening_i32_to_i64(torch):
    from masters_final_amd import _lib as L
    lib = L.load_library()
    n = 3_000_000
    a = np.full(n, 1000, np.int32)  # total 3e9 > 2^31
    d = torch.from_numpy(a).cuda()
    out = torch.empty(n, dtype=torch.int64, device="cuda")
    tot = torch.empty(1, dtype=torch.int64, device="cuda")
    ws = L.workspace(lib.eac_scan_workspace_bytes(n, 8))
    L.check(lib.eac_scan_exclusive_i32_i64(L.ptr(d), L.ptr(out), n, 0, L.ptr(tot), L.ptr(ws),
                                           ws.numel(), L.stream_ptr()))
    assert int(tot.item()) == 3_000_000_000
    assert np.array_equal(out.cpu().numpy(), np.arange(n, dtype=np.int64) * 1000)


# ------------------------------------------------------------------ Boruvka
def _check_forest(g, r, symmetric=True):
    n = g["n"]
    fe = g["indptr"][:-1]
    od = np.diff(g["indptr"])
    u, v, w, pos = O.csr_edges(g["dest"], g["weight"], fe, od, symmetric=symmetric)
    inf, _ = O.kruskal_canonical(u, v, w, n)
    want = {(a, b) for a, b, f in zip(u, v, inf) if f}
    nm = r["n_mst"]
    ids = r["ids"][:nm].cpu().numpy()
    rows = np.repeat(np.arange(n), od)
    got = [(min(rows[e], g["dest"][e]), max(rows[e], g["dest"][e])) for e in ids]
    assert len(got) == len(set(got)) == len(want)
    assert set(got) == want
    # each id is the entry in the smaller endpoint's row
    assert np.all(rows[ids] < g["dest"][ids]) if symmetric else True
    if r["weights"] is not None:
        assert np.array_equal(r["weights"][:nm].cpu().numpy(), g["weight"][ids])
    nc, lab = O.components(u, v, n)
    assert r["n_components"] == nc
    if r["labels"] is not None:
        assert np.array_equal(r["labels"].cpu().numpy(), lab)


def test_boruvka_4elt_golden(torch, elt4):
    from masters_final_amd.graph.mst import boruvka_minho_gpu, boruvka_minho_seq
    g = elt4
    mst, nm = boruvka_minho_seq(g["dest"], g["weight"], g["fe"], g["od"])
    assert nm == 15605  # test_boruvka.py:62-99
    assert int(g["weight"][mst[:nm]].astype(np.int64).sum()) == 2964801
    mst2, nm2 = boruvka_minho_gpu(g["dest"], g["weight"], g["fe"], g["od"], MAX_TPB=256)
    assert nm2 == nm and np.array_equal(mst, mst2)  # deterministic


def test_boruvka_4elt_edge_set_and_labels(torch, elt4):
    from masters_final_amd.graph.mst import boruvka_mst
    g = dict(elt4)
    r = boruvka_mst(g["dest"], g["weight"], g["fe"], g["od"], want_weights=True,
                    want_labels=True)
    _check_forest({"n": g["n"], "indptr": g["indptr"], "dest": g["dest"],
                   "weight": g["weight"]}, r)
    assert r["rounds"] >= 2


@pytest.mark.parametrize("seed,n,deg,wmax,iso,parts", [
    (0, 50, 3, 2, 3, 1), (1, 1000, 4, 5, 10, 3), (2, 5000, 8, 100, 0, 1),
    (3, 20000, 6, 3, 50, 7), (4, 3000, 1.5, 10, 100, 20)])
def test_boruvka_random_graphs(torch, seed, n, deg, wmax, iso, parts):
    from masters_final_amd.graph.mst import boruvka_mst
    g = synth.random_graph(n, deg, seed, wmax=wmax, n_isolated=iso, n_parts=parts)
    r = boruvka_mst(g["dest"], g["weight"], g["indptr"][:-1], np.diff(g["indptr"]),
                    want_weights=True, want_labels=True)
    _check_forest(g, r)


def test_boruvka_offsets_without_outdegree_and_float_weights(torch):
    from masters_final_amd.graph.mst import boruvka_mst
    g = synth.random_graph(4000, 5, 9, wmax=7)
    gf = dict(g, weight=(g["weight"].astype(np.float32) * 0.25 - 1.0))  # negatives too
    r = boruvka_mst(gf["dest"], gf["weight"], gf["indptr"], None, want_weights=True,
                    want_labels=True)
    _check_forest(gf, r)


def test_boruvka_condensed_upper_triangle(torch):
    """symmetric=0: every entry is an edge (the condensed co-association form)."""
    from masters_final_amd.graph.mst import boruvka_mst
    g = synth.random_graph(3000, 6, 21, wmax=4, n_isolated=7, n_parts=2)
    n = g["n"]
    rows = np.repeat(np.arange(n), np.diff(g["indptr"]))
    keep = g["dest"] > rows
    indptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows[keep], minlength=n), out=indptr[1:])
    gu = {"n": n, "indptr": indptr, "dest": g["dest"][keep], "weight": g["weight"][keep]}
    r = boruvka_mst(gu["dest"], gu["weight"], indptr, None, symmetric=False,
                    want_weights=True, want_labels=True)
    _check_forest(gu, r, symmetric=False)


def test_boruvka_edge_cases(torch):
    from masters_final_amd.graph.mst import boruvka_mst
    # single vertex, no edges
    r = boruvka_mst(np.zeros(0, np.int32), np.zeros(0, np.int32), np.array([0, 0]), None,
                    want_labels=True)
    assert r["n_mst"] == 0 and r["n_components"] == 1
    assert r["labels"].cpu().tolist() == [0]
    # isolated vertices only
    r = boruvka_mst(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(6, np.int64), None,
                    want_labels=True)
    assert r["n_components"] == 5 and r["labels"].cpu().tolist() == [0, 1, 2, 3, 4]
    # two vertices, one edge, plus a self loop
    dest = np.array([0, 1, 0], np.int32)
    w = np.array([9, 4, 4], np.int32)
    r = boruvka_mst(dest, w, np.array([0, 2, 3]), None, want_weights=True, want_labels=True)
    assert r["n_mst"] == 1 and r["ids"][0].item() == 1 and r["weights"][0].item() == 4


@pytest.mark.parametrize("n,direction,extra_deg", [(30_000, -1, 0), (30_000, 1, 0),
                                                   (20_000, -1, 8)])
def test_boruvka_long_pick_chains(torch, n, direction, extra_deg):
    """A path whose weights fall (or rise) along it: every vertex picks the
    same neighbour side, so one round's picks form a single tree of depth
    n - 2 (the deepest walk a hook can take). With extra heavier random
    edges the path is the light level of the split."""
    from masters_final_amd.graph.mst import boruvka_mst
    i = np.arange(n - 1, dtype=np.int64)
    lo, hi = i, i + 1
    w = (n - i) if direction < 0 else (i + 1)
    if extra_deg:
        rng = np.random.default_rng(7)
        m = n * extra_deg // 2
        a = rng.integers(0, n, m)
        b = rng.integers(0, n, m)
        keep = np.abs(a - b) > 1
        a, b = np.minimum(a[keep], b[keep]), np.maximum(a[keep], b[keep])
        key = np.unique(a * n + b)
        a, b = key // n, key % n
        lo, hi = np.concatenate([lo, a]), np.concatenate([hi, b])
        w = np.concatenate([w, rng.integers(n, 4 * n, a.size)])
    indptr, dest, weight = synth.csr_from_undirected(n, lo, hi, w.astype(np.int32))
    g = {"n": n, "indptr": indptr, "dest": dest, "weight": weight}
    r = boruvka_mst(dest, weight, indptr, None, want_weights=True, want_labels=True)
    _check_forest(g, r)


def test_boruvka_asymmetric_input_passed_as_symmetric_is_reported(torch):
    """Rows that list an edge only from one side can make a round's picks a
    cycle with no mirror pair (a picks (a,b), b picks (b,c), c picks (c,a)):
    the pick-tree walk stops after n steps and the call reports the misuse
    instead of returning a forest."""
    from masters_final_amd._lib import EACError
    from masters_final_amd.graph.mst import boruvka_mst
    # row 0: (0,1); row 1: (1,2); row 2: (2,0), all weight 1
    dest = np.array([1, 2, 0], np.int32)
    w = np.array([1, 1, 1], np.int32)
    with pytest.raises(EACError, match="not symmetric"):
        boruvka_mst(dest, w, np.array([0, 1, 2, 3]), None, symmetric=True, want_labels=True)
    # the triangle in the condensed form (symmetric=False: upper triangle,
    # every entry an edge) gives its forest
    r = boruvka_mst(np.array([1, 2, 2], np.int32), w, np.array([0, 2, 3, 3]), None,
                    symmetric=False, want_labels=True)
    assert r["n_mst"] == 2 and r["n_components"] == 1


def test_boruvka_knn_200k(torch):
    from masters_final_amd.graph.mst import boruvka_mst
    g = synth.knn_graph(n=200_000, k=20, seed=5)
    r = boruvka_mst(g["dest"], g["weight"], g["indptr"], None, want_weights=True,
                    want_labels=True)
    _check_forest(g, r)
    assert r["n_components"] == 5
    r2 = boruvka_mst(g["dest"], g["weight"], g["indptr"], None, want_weights=True,
                     want_labels=True)
    nm = r["n_mst"]
    assert r2["n_mst"] == nm
    assert np.array_equal(r["ids"][:nm].cpu().numpy(), r2["ids"][:nm].cpu().numpy())


# ------------------------------------------------------ connected components
@pytest.mark.parametrize("seed", range(3))
def test_connected_components(torch, seed):
    from masters_final_amd.graph.connected_components import (connected_comps_gpu,
                                                               connected_comps_seq)
    g = synth.random_graph(20000, 1.1 + seed, seed, n_isolated=30, n_parts=5)
    fe, od = g["indptr"][:-1], np.diff(g["indptr"]).astype(np.int32)
    u, v, _, _ = O.csr_edges(g["dest"], g["weight"], fe, od)
    nc, lab = O.components(u, v, g["n"])
    got = connected_comps_seq(g["dest"], g["weight"], fe, od)
    assert np.array_equal(got, lab)
    got2 = connected_comps_gpu(g["dest"], g["weight"], fe, od)
    assert np.array_equal(got2, lab)
    # the reference's supervertex ids give the same partition after relabel
    ref_lab = O.connected_comps_seq(g["dest"], g["weight"], fe, od)
    assert np.array_equal(O.canonical_relabel(ref_lab)[0], lab)


def test_canonical_relabel(torch):
    from masters_final_amd.graph.connected_components import canonical_relabel
    rng = np.random.default_rng(7)
    lab = rng.integers(0, 500, 100_000)
    out, cnt = canonical_relabel(lab.astype(np.int32), 500)
    want, wc = O.canonical_relabel(lab)
    assert cnt == wc and np.array_equal(out.cpu().numpy(), want)


@pytest.mark.parametrize("seed,n,deg,wmax,parts,fdt", [
    (11, 20000, 16, 100, 1, np.int32), (12, 30000, 20, 3, 4, np.int32),
    (13, 15000, 24, 1000, 2, np.float32), (14, 40000, 12, 255, 1, np.uint8)])
def test_boruvka_dense_graphs_light_heavy_split(torch, seed, n, deg, wmax, parts, fdt):
    """Graphs with >= 6 entries per vertex take the light/heavy split path."""
    from masters_final_amd.graph.mst import boruvka_mst
    g = synth.random_graph(n, deg, seed, wmax=wmax, n_parts=parts, dtype=fdt)
    if fdt == np.float32:
        g["weight"] = (g["weight"] * 0.37).astype(np.float32)
    assert g["dest"].size >= 6 * n
    r = boruvka_mst(g["dest"], g["weight"], g["indptr"], None, want_weights=True,
                    want_labels=True)
    _check_forest(g, r)


def test_graph_from_edges_matches_oracle(torch, elt4):
    from masters_final_amd.graph.build import getGraphFromEdges_gpu, getGraphFromEdges_seq
    from masters_final_amd.graph.mst import boruvka_minho_seq
    g = elt4
    mst, nm = boruvka_minho_seq(g["dest"], g["weight"], g["fe"], g["od"])
    edges = mst[:nm].astype(np.int64)
    want = O.graph_from_edges(g["dest"], g["weight"], g["fe"], g["od"], edges)
    n = g["n"]
    nod = np.zeros(n, np.int32)
    nfe = np.zeros(n, np.int32)
    ndest = np.zeros(2 * nm, np.int32)
    nw = np.zeros(2 * nm, g["weight"].dtype)
    assert getGraphFromEdges_seq(g["dest"], g["weight"], g["fe"], g["od"], edges, nod, nfe,
                                 ndest, nw) == 0
    assert np.array_equal(nod, want[0]) and np.array_equal(nfe, want[1])
    assert np.array_equal(ndest, want[2]) and np.array_equal(nw, want[3])
    # -1 ids are skipped by the GPU variant; bad ids are reported
    e2 = np.concatenate([edges[:10], [-1], edges[10:20]])
    nd, w2, fe2, od2 = getGraphFromEdges_gpu(g["dest"], g["weight"], g["fe"], g["od"], e2)
    want2 = O.graph_from_edges(g["dest"], g["weight"], g["fe"], g["od"], edges[:20])
    assert np.array_equal(nd.cpu().numpy(), want2[2]) and np.array_equal(od2.cpu().numpy(), want2[0])
    assert getGraphFromEdges_seq(g["dest"], g["weight"], g["fe"], g["od"],
                                 np.array([len(g["dest"]) + 3]), nod, nfe, ndest, nw) == -1


@pytest.mark.parametrize("light_x10,first_batch,bpcu,step,wmax", [
    (-1, 0, 0, 0, 60), (5, 2, 1, 2, 60), (40, 3, 2, 0, 60), (15, 30, 0, -1, 60),
    (10, 0, 0, 2, 3), (5, 4, 0, 3, 100000),
    (10, 0, -1, 8, 60), (-1, 0, -1, 0, 60), (5, 2, -1, 3, 100000)])
def test_boruvka_tuning_knobs_keep_the_forest(torch, light_x10, first_batch, bpcu, step, wmax):
    """Every tuning knob setting gives the canonical forest (knobs change
    speed only): no split, one split, up to four weight levels (also with
    fewer distinct weights than levels), tiny and large light fractions, a
    first round batch far too short (rounds continue after the read-back)
    or long; 256-thread filter workgroups or one 1024-thread workgroup
    per compute unit."""
    from masters_final_amd import tune
    from masters_final_amd.graph.mst import boruvka_mst
    g = synth.random_graph(30000, 16, 21, wmax=wmax, n_parts=3)
    before = tune.get_knobs()
    try:
        tune.set_knobs(light_x10=light_x10, first_batch=first_batch, filter_bpcu=bpcu,
                       level_step=step)
        r = boruvka_mst(g["dest"], g["weight"], g["indptr"], None, want_weights=True,
                        want_labels=True)
        _check_forest(g, r)
    finally:
        tune.set_knobs(**before)


def test_autotune_mst_applies_a_candidate(torch, tmp_path, monkeypatch):
    from masters_final_amd import tune
    monkeypatch.setenv("EAC_TUNE_CACHE", str(tmp_path / "tune.json"))
    g = synth.random_graph(20000, 16, 22, wmax=50)
    d = {k: torch.from_numpy(g[k]).cuda() for k in ("dest", "weight", "indptr")}
    before = tune.get_knobs()
    try:
        res = tune.autotune_mst(d["dest"], d["weight"], d["indptr"], reps=1,
                                grid={"light_x10": (-1, 15), "first_batch": (0,),
                                      "filter_bpcu": (0,)})
        assert not res["cached"] and len(res["candidates"]) == 2
        assert tune.get_knobs()["light_x10"] == res["knobs"]["light_x10"]
        again = tune.autotune_mst(d["dest"], d["weight"], d["indptr"])
        assert again["cached"] and again["knobs"] == res["knobs"]
    finally:
        tune.set_knobs(**before)

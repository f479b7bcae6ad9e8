"""GPU parity of the lifetime cuts and the radix sort against the oracle."""
import numpy as np
import pytest
import scipy.sparse as sp

from masters_final_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


@pytest.mark.parametrize("n,bits", [(1, 64), (1000, 20), (4096, 64), (300_000, 48),
                                    (1_000_003, 33)])
def test_radix_sort_u64(torch, n, bits):
    from masters_final_amd import _lib as L
    lib = L.load_library()
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 2**63, n, dtype=np.uint64) & np.uint64((1 << bits) - 1 if bits < 64
                                                                  else 2**64 - 1)
    keys[: n // 3] = keys[0]  # many duplicates: stability is visible through the values
    vals = np.arange(n, dtype=np.uint32)
    dk = torch.from_numpy(keys.view(np.int64)).cuda()
    dv = torch.from_numpy(vals.view(np.int32)).cuda()
    ok = torch.empty_like(dk)
    ov = torch.empty_like(dv)
    ws = L.workspace(lib.eac_sort_workspace_bytes(n, 8, 4))
    L.check(lib.eac_radix_sort_u64(L.ptr(dk), L.ptr(ok), L.ptr(dv), L.ptr(ov), n, 0, bits,
                                   L.ptr(ws), ws.numel(), L.stream_ptr()))
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ok.cpu().numpy().view(np.uint64), keys[order])
    assert np.array_equal(ov.cpu().numpy().view(np.uint32), vals[order])


def _graph_edges(g):
    fe, od = g["indptr"][:-1], np.diff(g["indptr"])
    return O.csr_edges(g["dest"], g["weight"], fe, od)


@pytest.mark.parametrize("seed,n,deg,wmax,iso,parts", [
    (0, 200, 3, 4, 2, 1), (1, 2000, 5, 30, 10, 3), (2, 20000, 6, 100, 40, 5),
    (3, 5000, 2, 7, 200, 10)])
def test_sl_mst_lifetime_vs_oracle(torch, seed, n, deg, wmax, iso, parts):
    from masters_final_amd.graph.singlelink import lifetime_cut, sl_mst_lifetime_seq
    g = synth.random_graph(n, deg, seed, wmax=wmax, n_isolated=iso, n_parts=parts)
    u, v, w, _ = _graph_edges(g)
    for dw in (None, float(wmax) * 3):
        want, nc, kept = O.sl_mst_lifetime(u, v, w, n, disconnect_weight=dw)
        lab, info = lifetime_cut(g["dest"], g["weight"], g["indptr"], None, rule=1, max_dist=dw)
        assert info["kept_edges"] == kept and info["n_components"] == nc
        assert np.array_equal(lab.cpu().numpy(), want)
    got = sl_mst_lifetime_seq(g["dest"], g["weight"], g["indptr"][:-1],
                              np.diff(g["indptr"]).astype(np.int32))
    assert np.array_equal(got, O.sl_mst_lifetime(u, v, w, n)[0])


@pytest.mark.parametrize("seed,n,deg,wmax,iso,parts", [
    (5, 300, 4, 6, 3, 1), (6, 3000, 5, 40, 20, 4), (7, 30000, 8, 100, 0, 1),
    (8, 8000, 3, 9, 100, 6)])
def test_sp_sl_lifetime_vs_oracle(torch, seed, n, deg, wmax, iso, parts):
    from masters_final_amd.EAC.eac import sp_sl_lifetime
    g = synth.random_graph(n, deg, seed, wmax=wmax, n_isolated=iso, n_parts=parts)
    u, v, w, _ = _graph_edges(g)
    # the driver passes the condensed upper triangle
    up = sp.csr_matrix((w.astype(np.int32), (u, v)), shape=(n, n))
    up.sort_indices()
    for md, ncl in ((None, 0), (wmax + 5, 0), (None, 7)):
        try:
            want_nc, want = O.sp_sl_lifetime(u, v, w, n, max_dist=md, n_clusters=ncl)
        except ValueError:
            with pytest.raises(ValueError):
                sp_sl_lifetime(up, md, n_clusters=ncl)
            continue
        nc, lab = sp_sl_lifetime(up, md, n_clusters=ncl)
        assert nc == want_nc
        assert np.array_equal(lab, want)


def test_sp_sl_lifetime_4elt_known_answer(torch, elt4):
    from masters_final_amd.EAC.eac import sp_sl_lifetime
    g = elt4
    fe, od = g["fe"], g["od"]
    u, v, w, _ = O.csr_edges(g["dest"], g["weight"], fe, od)
    up = sp.csr_matrix((w.astype(np.uint16), (u, v)), shape=(g["n"], g["n"]))
    nc, lab = sp_sl_lifetime(up)
    assert nc == 6
    assert sorted(np.bincount(lab).tolist(), reverse=True) == [15600, 2, 1, 1, 1, 1]
    assert np.array_equal(lab, O.sp_sl_lifetime(u, v, w, g["n"])[1])


def test_sp_sl_lifetime_hand_cases(torch):
    from masters_final_amd.EAC.eac import sp_sl_lifetime

    def mat(u, v, w, n):
        return sp.csr_matrix((np.asarray(w, np.int32), (u, v)), shape=(n, n))
    nc, lab = sp_sl_lifetime(mat([0, 1, 2], [1, 2, 3], [1, 1, 5], 4))
    assert nc == 2 and lab.tolist() == [0, 0, 0, 1]
    nc, lab = sp_sl_lifetime(mat([0, 1, 2], [1, 2, 3], [2, 2, 2], 4))
    assert nc == 1 and lab.size == 0
    with pytest.raises(ValueError):
        sp_sl_lifetime(mat([0, 1, 2], [1, 2, 3], [1, 2, 4], 4))
    m2 = mat([0, 1, 3, 4], [1, 2, 4, 5], [1, 1, 1, 5], 6)
    assert sp_sl_lifetime(m2)[1].tolist() == [0, 0, 0, 1, 1, 2]
    assert sp_sl_lifetime(m2, 10)[1].tolist() == [0, 0, 0, 1, 1, 1]
    assert sp_sl_lifetime(m2, n_clusters=4)[1].tolist() == [0, 0, 0, 1, 2, 3]


@pytest.mark.parametrize("rule,kind", [(0, "er"), (1, "er"), (0, "knn")])
def test_batched_lifetime_cut_matches_per_graph_oracle(torch, rule, kind):
    from masters_final_amd.graph.singlelink import lifetime_cut_batched
    b = synth.graph_batch(12, (50, 3000), (3, 12), seed=7 + rule, wmax=40, kind=kind)
    labels, gout, info = lifetime_cut_batched(b["dest"], b["weight"], b["indptr"], b["goff"],
                                              rule=rule)
    labels = labels.cpu().numpy()
    gout = gout.cpu().numpy()
    fe = b["indptr"][:-1]
    od = np.diff(b["indptr"])
    u, v, w, _ = O.csr_edges(b["dest"], b["weight"], fe, od)
    for g in range(len(b["goff"]) - 1):
        a, z = b["goff"][g], b["goff"][g + 1]
        sel = (u >= a) & (u < z)
        try:
            if rule == 0:
                nc, want = O.sp_sl_lifetime(u[sel] - a, v[sel] - a, w[sel], z - a)
            else:
                want, nc, kept = O.sl_mst_lifetime(u[sel] - a, v[sel] - a, w[sel], z - a)
        except ValueError:
            assert gout[g, 3] == -1, g
            continue
        if rule == 0 and want.size == 0:
            assert gout[g, 1] == 1
            continue
        assert gout[g, 0] == nc, g
        assert np.array_equal(labels[a:z], want), g


@pytest.mark.parametrize("rule", [0, 1])
def test_batched_graph_replays_match_eager(torch, rule):
    """hipGraph replays of the batched cut: three batches of different sizes
    through one recording, eager work and host reads between replays, every
    replay equal to the eager batched call on the same batch."""
    from masters_final_amd.graph.singlelink import LifetimeCutGraph, lifetime_cut_batched
    batches = [synth.graph_batch(10, (40, 2500), (3, 10), seed=31 + rule, wmax=30, kind="knn"),
               synth.graph_batch(6, (30, 900), (2, 8), seed=41 + rule, wmax=5, kind="er"),
               synth.graph_batch(14, (20, 1500), (3, 12), seed=51 + rule, wmax=200, kind="knn")]
    # zero weights (symmetric: chosen per undirected edge) leave the forest under rule 0
    b1 = batches[1]
    rows = np.repeat(np.arange(b1["n"]), np.diff(b1["indptr"]))
    b1["weight"][(rows + b1["dest"]) % 7 == 0] = 0
    n_cap = max(b["n"] for b in batches) + 5
    m_cap = max(b["dest"].size for b in batches) + 11
    g_cap = max(len(b["goff"]) - 1 for b in batches) + 2
    cg = LifetimeCutGraph(n_cap, m_cap, g_cap, torch.int32, rule=rule)
    for rep in range(2):
        for i, b in enumerate(batches):
            print(f"replay {rep} batch {i}: N={b['n']} M={b['dest'].size}", flush=True)
            lab, gout, info = cg.run(b["dest"], b["weight"], b["indptr"], b["goff"])
            lab_h, gout_h = lab.cpu().numpy().copy(), gout.cpu().numpy().copy()
            want_lab, want_gout, want_info = lifetime_cut_batched(
                b["dest"], b["weight"], b["indptr"], b["goff"], rule=rule)
            assert np.array_equal(gout_h, want_gout.cpu().numpy()), (rep, i)
            assert np.array_equal(lab_h, want_lab.cpu().numpy()), (rep, i)
            assert info["kept_edges"] == want_info["kept_edges"]
            assert info["forest_edges"] == want_info["forest_edges"]
    cg.close()


def test_batched_graph_tuned_rounds_fall_back(torch):
    """A recording with too few rounds is detected on the device and the
    batch is run again at the bound; tuning records the observed rounds."""
    from masters_final_amd.graph.singlelink import LifetimeCutGraph, lifetime_cut_batched
    b = synth.graph_batch(8, (200, 3000), (4, 10), seed=77, wmax=50, kind="er")
    cg = LifetimeCutGraph(b["n"] + 1, b["dest"].size + 1, 10, torch.int32, rule=0, max_rounds=2)
    assert cg.rounds == 2
    lab, gout, info = cg.run(b["dest"], b["weight"], b["indptr"], b["goff"])
    assert cg.rounds == cg.bound_rounds
    want_lab, want_gout, _ = lifetime_cut_batched(b["dest"], b["weight"], b["indptr"], b["goff"])
    assert np.array_equal(lab.cpu().numpy(), want_lab.cpu().numpy())
    tuned = LifetimeCutGraph(b["n"] + 1, b["dest"].size + 1, 10, torch.int32, rule=0, tune=True)
    lab, gout, info = tuned.run(b["dest"], b["weight"], b["indptr"], b["goff"])
    assert tuned.rounds == min(info["rounds"] + tuned.margin, tuned.bound_rounds)
    lab, gout, info = tuned.run(b["dest"], b["weight"], b["indptr"], b["goff"])
    assert np.array_equal(lab.cpu().numpy(), want_lab.cpu().numpy())
    assert np.array_equal(gout.cpu().numpy(), want_gout.cpu().numpy())


def test_device_batch_generator_and_batched_cut(torch):
    """The GPU batch generator of the large batched workload gives symmetric
    CSR graphs (rows sorted, weights mirrored, no self loops); the batched
    cut over it equals the per-graph oracle."""
    from masters_final_amd.graph.singlelink import lifetime_cut_batched
    b = synth.graph_batch_device(9, (50, 4000), (4, 20), seed=5, wmax=256)
    ip = b["indptr"].cpu().numpy()
    dest = b["dest"].cpu().numpy().astype(np.int64)
    w = b["weight"].cpu().numpy()
    N = b["n"]
    rows = np.repeat(np.arange(N), np.diff(ip))
    assert np.all(rows != dest)
    key = rows * N + dest
    assert np.all(np.diff(key) > 0)  # sorted by (row, column), no repeats
    mirror = np.searchsorted(key, dest * N + rows)
    assert np.array_equal(key[mirror], dest * N + rows)
    assert np.array_equal(w[mirror], w)
    goff = b["goff"]
    gid = np.searchsorted(goff, rows, side="right") - 1
    assert np.all((dest >= goff[gid]) & (dest < goff[gid + 1]))
    labels, gout, _ = lifetime_cut_batched(b["dest"], b["weight"], b["indptr"], goff, rule=0)
    labels, gout = labels.cpu().numpy(), gout.cpu().numpy()
    u, v, ww, _ = O.csr_edges(dest, w, ip[:-1], np.diff(ip))
    for g in range(len(goff) - 1):
        a, z = goff[g], goff[g + 1]
        sel = (u >= a) & (u < z)
        try:
            nc, want = O.sp_sl_lifetime(u[sel] - a, v[sel] - a, ww[sel], z - a)
        except ValueError:
            assert gout[g, 3] == -1
            continue
        if want.size == 0:
            assert gout[g, 1] == 1
            continue
        assert gout[g, 0] == nc
        assert np.array_equal(labels[a:z], want), g

"""The learned guess's pair features, targets and density assembly (vectorized
over atom pairs) against their definition written pair by pair: the S block,
the local core Hamiltonian T + V(A) (+ V(B) for B != A, each nucleus's own
attraction matrix V symmetric) built from the own-nucleus matrix O, element flags,
A == B, R, exp(-R), 1/(1+R), zero-padded to NMAX functions per atom."""
import numpy as np

from quiqbox_amd import learned_guess as lg
from quiqbox_amd.particles import NuclearCluster


def _case(seed=0):
    rng = np.random.default_rng(seed)
    syms = ["O", "H", "H", "O", "H", "H", "H"]
    nuc = NuclearCluster(syms, [tuple(rng.normal(size=3) * 3) for _ in syms], pairwiseSort=False)
    sizes = [14 if s == "O" else 5 for s in syms]
    perm = rng.permutation(sum(sizes))
    blocks, k = [], 0
    for z in sizes:
        blocks.append(np.sort(perm[k:k + z]))
        k += z
    n = sum(sizes)
    mats = [rng.normal(size=(n, n)) for _ in range(2 + len(syms))]
    S, T, Vn = mats[0], mats[1], [0.5 * (m + m.T) for m in mats[2:]]
    # own-nucleus attraction: row i takes the attraction of function i's atom
    owner = np.empty(n, dtype=int)
    for a, b in enumerate(blocks):
        owner[b] = a
    O = np.stack([Vn[owner[i]][i] for i in range(n)])
    return nuc, blocks, n, S, T, Vn, O


def test_pair_features_follow_the_pairwise_definition():
    nuc, blocks, n, S, T, Vn, O = _case()
    X = lg.pair_features(nuc, blocks, S, (T, O))
    na = len(blocks)
    R = nuc.coord_array
    o = 2 * lg.NMAX * lg.NMAX
    for a in range(na):
        for b in range(na):
            ia, ib = blocks[a], blocks[b]
            row = X[a * na + b]
            s = row[:lg.NMAX * lg.NMAX].reshape(lg.NMAX, lg.NMAX)
            h = row[lg.NMAX * lg.NMAX:o].reshape(lg.NMAX, lg.NMAX)
            hl = T[np.ix_(ia, ib)] + Vn[a][np.ix_(ia, ib)] + (Vn[b][np.ix_(ia, ib)] if b != a else 0.0)
            np.testing.assert_allclose(s[:len(ia), :len(ib)], S[np.ix_(ia, ib)], rtol=0, atol=0)
            np.testing.assert_allclose(h[:len(ia), :len(ib)], hl, rtol=0, atol=1e-14)
            assert not s[len(ia):].any() and not s[:, len(ib):].any() and not h[len(ia):].any()
            flags = np.zeros(2 * len(lg.ELEMENTS))
            flags[lg.ELEMENTS.index(nuc.symbols[a])] = 1.0
            flags[len(lg.ELEMENTS) + lg.ELEMENTS.index(nuc.symbols[b])] = 1.0
            np.testing.assert_array_equal(row[o:o + 2 * len(lg.ELEMENTS)], flags)
            d = np.linalg.norm(R[a] - R[b])
            np.testing.assert_allclose(row[o + 2 * len(lg.ELEMENTS):], [float(a == b), d, np.exp(-d), 1 / (1 + d)],
                                       rtol=1e-15, atol=0)


def test_targets_and_assembly_are_inverse_block_maps():
    nuc, blocks, n, S, T, Vn, O = _case(1)
    rng = np.random.default_rng(2)
    D = rng.normal(size=(n, n))
    D = D + D.T
    Y = lg.pair_targets(blocks, D)
    na = len(blocks)
    for a in range(na):
        for b in range(na):
            blk = Y[a * na + b].reshape(lg.NMAX, lg.NMAX)
            np.testing.assert_array_equal(blk[:len(blocks[a]), :len(blocks[b])], D[np.ix_(blocks[a], blocks[b])])
    np.testing.assert_allclose(lg.assemble_density(blocks, Y, n), D, rtol=0, atol=1e-15)


def test_packaged_cho_network_file():
    """The packaged C/H/O network is plain arrays (loads with allow_pickle=False)
    whose shapes match the CHO_SP layout, with its training record."""
    import json
    z = np.load(lg.CHO_MODEL, allow_pickle=False)
    L = lg.CHO_SP
    assert tuple(str(e) for e in z["elements"]) == L.elements and int(z["nmax"]) == L.nmax
    assert z["W0"].shape == (L.n_in, 256) and z["W1"].shape == (256, 256) and z["W2"].shape == (256, L.n_out)
    assert z["mu"].shape == (L.n_feat,) and z["sd"].shape == (L.n_feat,)
    assert all(np.all(np.isfinite(z[k])) for k in ("W0", "W1", "W2", "b0", "b1", "b2", "mu", "sd"))
    meta = json.loads(str(z["meta"]))
    assert meta["systems"] == 4000 and meta["atoms"] == [8, 16] and meta["pairs"] > 0

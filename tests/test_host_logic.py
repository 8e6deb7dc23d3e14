

def test_parse_autotune_report():
    """the tm_autotune report format: trial lines, then the chosen shapes"""
    from timothy_amd import parse_autotune
    text = ("trial neighbors nbr_block=64 walk_window=0 ms=1.2500\n"
            "trial neighbors walk_f32=1 ms=1.0000\n"
            "trial tiles_sph tile_width=128 unroll=2 tiles_ms=0.1000 sph_ms=0.0800 ms=0.1800\n"
            "chosen precision=mixed nbr_block=256 walk_window=1024 walk_f32=1 sph_tiles=1 tile_width=128 "
            "sph_block=256 unroll=2 lanes=1\n")
    r = parse_autotune(text)
    assert [t["stage"] for t in r["trials"]] == ["neighbors", "neighbors", "tiles_sph"]
    assert r["trials"][2] == dict(stage="tiles_sph", tile_width=128, unroll=2, tiles_ms=0.1, sph_ms=0.08, ms=0.18)
    assert r["chosen"]["precision"] == "mixed" and r["chosen"]["tile_width"] == 128
    import pytest
    with pytest.raises(ValueError):
        parse_autotune("nonsense line\n")


def _radix_tree(keys, f, l, nodes):
    """binary radix tree over the distinct sorted keys[f..l] by recursive
    split at the highest differing bit; appends (f, l) of every internal
    node and returns the subtree as nested (f, l, left, right) tuples"""
    if f == l:
        return (f, l, None, None)
    b = int(keys[f] ^ keys[l]).bit_length() - 1
    g = f
    while (int(keys[g + 1]) >> b) & 1 == 0:
        g += 1
    nodes.append((f, l))
    return (f, l, _radix_tree(keys, f, g, nodes), _radix_tree(keys, g + 1, l, nodes))


def test_walk_tree_slot_formula():
    """the walk's tree places a colony's nodes in descending pre-order (the
    child with the higher keys first).  The build computes each node's slot
    without walking the tree as 2s + 2 (e - l) + (f - s) - (P[f] - P[s]),
    with P the exclusive prefix sum over cells of the number of the colony's
    internal nodes whose range ends there; this checks it against an
    explicit depth-first traversal on random colonies, including colonies
    of one and two cells and lopsided key sets."""
    import numpy as np
    rng = np.random.default_rng(7)
    for trial in range(60):
        sizes = rng.integers(1, 40, size=rng.integers(1, 6))
        if trial % 10 == 0:
            sizes = np.array([1, 2, 1, 3])
        n = int(sizes.sum())
        off = np.concatenate([[0], np.cumsum(sizes)])
        expect = {}
        chain = np.zeros(n, dtype=np.int64)
        for c in range(len(sizes)):
            s, e = int(off[c]), int(off[c + 1]) - 1
            bits = 6 if trial % 3 == 0 else 20  # dense codes give deep, lopsided trees
            keys = np.sort(rng.choice(1 << bits, size=e - s + 1, replace=False)).astype(np.int64)
            keys = np.concatenate([np.zeros(s, dtype=np.int64), keys])  # index by global cell
            nodes = []
            root = _radix_tree(keys, s, e, nodes)
            for (_, l) in nodes:
                chain[l] += 1
            order = []
            stack = [root]
            while stack:
                nd = stack.pop()
                order.append((nd[0], nd[1], nd[2] is None))
                if nd[2] is not None:
                    stack.append(nd[2])  # the lower-key child is visited second
                    stack.append(nd[3])
            assert len(order) == 2 * (e - s) + 1
            for k, key in enumerate(order):
                expect[key] = 2 * s + k
        P = np.concatenate([[0], np.cumsum(chain)])
        for (f, l, leaf), slot in expect.items():
            c = int(np.searchsorted(off, f, side="right")) - 1
            s, e = int(off[c]), int(off[c + 1]) - 1
            assert 2 * s + 2 * (e - l) + (f - s) - (P[f] - P[s]) == slot, (trial, f, l, leaf)

"""Pins the oracle's Barnes-Hut path for D = 3..7 (src/tsne/morton/d3.rs .. d7.rs,
src/tsne/arena.rs:137-392, src/test.rs:1177-1185, 1914-2016).

The Morton codes are checked against the reference's known answers and against spreaders built
differently from the oracle's bit loop (shift-and-mask chains for D = 3, 4 and nibble tables for
D = 5..7, the constructions the reference uses).  The arena is checked structurally from the
quantised coordinates (every internal node's members share its cell, and its children split that
cell by the next digit), and the repulsion against the FP64 all-pairs sum as theta shrinks."""
import numpy as np
import pytest

import oracle
from tests._util import lcg_cloud, lcg_samples

BITS = {2: 32, 3: 21, 4: 16, 5: 25, 6: 21, 7: 18}
DIMS = [3, 4, 5, 6, 7]


# --- Morton codecs ---------------------------------------------------------------------------

def test_morton_bits_per_axis():
    assert {d: oracle.morton_bits(d) for d in BITS} == BITS
    assert oracle.morton_bits(1) == 0 and oracle.morton_bits(8) == 0


@pytest.mark.parametrize("D", DIMS)
def test_morton_known_z_order(D):
    # d3.rs:63-69, d4.rs, d5.rs:140-147, d6.rs:151-159, d7.rs:149-158
    assert oracle.morton_encode(np.zeros(D, np.uint32)) == 0
    for a in range(D):
        e = np.zeros(D, np.uint32)
        e[a] = 1
        assert oracle.morton_encode(e) == 1 << a
    assert oracle.morton_encode(np.ones(D, np.uint32)) == (1 << D) - 1


def _part_1by2(x):  # D = 3: shift-and-mask chain over 21 bits
    x &= 0x1FFFFF
    x = (x | (x << 32)) & 0x001F00000000FFFF
    x = (x | (x << 16)) & 0x001F0000FF0000FF
    x = (x | (x << 8)) & 0x100F00F00F00F00F
    x = (x | (x << 4)) & 0x10C30C30C30C30C3
    return (x | (x << 2)) & 0x1249249249249249


def _part_1by3(x):  # D = 4: shift-and-mask chain over 16 bits
    x &= 0xFFFF
    x = (x | (x << 24)) & 0x000000FF000000FF
    x = (x | (x << 12)) & 0x000F000F000F000F
    x = (x | (x << 6)) & 0x0303030303030303
    return (x | (x << 3)) & 0x1111111111111111


def _nibble_spread(x, D, bits):  # D = 5..7: per-nibble table, nibbles placed 4*D bits apart
    table = [sum(((v >> i) & 1) << (D * i) for i in range(4)) for v in range(16)]
    out = 0
    for k in range((bits + 3) // 4):
        out |= table[(x >> (4 * k)) & 0xF] << (4 * D * k)
    return out


def _encode_alt(c, D):
    bits = BITS[D]
    if D == 3:
        parts = [_part_1by2(int(v)) for v in c]
    elif D == 4:
        parts = [_part_1by3(int(v)) for v in c]
    else:
        parts = [_nibble_spread(int(v), D, bits) for v in c]
    code = 0
    for a, p in enumerate(parts):
        code |= p << a
    return code


def _decode(code, D):
    bits = BITS[D]
    return [sum(((code >> (b * D + a)) & 1) << b for b in range(bits)) for a in range(D)]


@pytest.mark.parametrize("D", DIMS)
def test_morton_matches_independent_spreaders_and_round_trips(D):
    rng = np.random.default_rng(0x9ABC + D)
    bits = BITS[D]
    for _ in range(300):
        c = rng.integers(0, 1 << bits, D).astype(np.uint32)
        code = oracle.morton_encode(c)
        assert code == _encode_alt(c, D)
        assert _decode(code, D) == [int(v) for v in c]
    top = np.full(D, (1 << bits) - 1, np.uint32)
    assert oracle.morton_encode(top) == (1 << (D * bits)) - 1


# --- arena -----------------------------------------------------------------------------------

def _quantise(y, bits):
    """morton.rs:103-120 over the per-axis bounding box (arena.rs:150-165), in f32."""
    y = np.asarray(y, np.float32)
    mn, mx = y.min(0), y.max(0)
    ext = (mx - mn).astype(np.float32)
    scale = np.float32(2.0 ** bits)
    inv = np.where(ext > 0, scale / np.where(ext > 0, ext, 1), 0).astype(np.float32)
    s = np.floor(((y - mn).astype(np.float32) * inv).astype(np.float32))
    q = np.minimum(np.where(s > 0, s, 0).astype(np.uint64), np.uint64((1 << bits) - 1))
    return q, float(ext.max())


@pytest.mark.parametrize("D", DIMS)
def test_arena_cells_split_by_digit(D):
    n = 600
    y = lcg_cloud(n, D, 11 + D) + np.float32(100.0)
    a = oracle.bh_build(y)
    bits = BITS[D]
    q, max_ext = _quantise(y, bits)
    order = a["order"]
    nn = len(a["count"])
    leaves = a["first_child"] == 0xFFFFFFFF
    assert int(a["count"][0]) == n and int(a["count"][leaves].sum()) == n
    # the sorted order is Z-order of the quantised coordinates, ties by index
    codes = [oracle.morton_encode(q[i].astype(np.uint32)) for i in order]
    assert all((codes[t], order[t]) < (codes[t + 1], order[t + 1]) for t in range(n - 1))
    for i in range(nn):
        s, e = (int(v) for v in a["ranges"][i])
        assert int(a["count"][i]) == e - s
        mem = order[s:e]
        np.testing.assert_allclose(a["com"][i], y[mem].astype(np.float64).mean(0), rtol=1e-5)
        if leaves[i]:
            assert e - s == 1 or len({codes[t] for t in range(s, e)}) == 1
            continue
        lvl = int(a["level"][i])
        assert 0 <= lvl < bits
        top = q[mem] >> np.uint64(bits - lvl)
        assert np.all(top == top[0])  # one cell of level lvl
        digit = (q[mem] >> np.uint64(bits - 1 - lvl)) & np.uint64(1)
        dig = (digit * (np.uint64(1) << np.arange(D, dtype=np.uint64))).sum(1)
        fc, cc = int(a["first_child"][i]), int(a["child_count"][i])
        assert 2 <= cc <= 2 ** D and fc > i
        assert cc == len(np.unique(dig))  # the cell splits at this level
        for c in range(fc, fc + cc):
            cs, ce = (int(v) for v in a["ranges"][c])
            assert s <= cs < ce <= e
            assert len(np.unique(dig[cs - s:ce - s])) == 1
        assert int(a["ranges"][fc][0]) == s and int(a["ranges"][fc + cc - 1][1]) == e
    # squared half-width per level (arena.rs:167-174)
    for lvl in range(bits + 1):
        hw = np.float32(max_ext) / np.float32(2.0 ** (lvl + 1))
        assert a["hw2"][lvl] == np.float32(hw * hw)


@pytest.mark.parametrize("D,seed", [(5, 42), (6, 46), (7, 47)])
def test_arena_builds_reference_cases(D, seed):
    # test.rs:1177-1185 (5-D), 1982-1990 (6-D), 2008-2016 (7-D): the root holds every point
    a = oracle.bh_build(lcg_samples(500, D, seed))
    assert int(a["count"][0]) == 500


@pytest.mark.parametrize("D", DIMS)
def test_arena_duplicates_and_single(D):
    a = oracle.bh_build(np.full((300, D), -1.25, np.float32))
    assert len(a["count"]) == 1 and a["count"][0] == 300
    one = np.arange(D, dtype=np.float32)[None]
    a = oracle.bh_build(one)
    assert a["com"][0].tolist() == one[0].tolist()


@pytest.mark.parametrize("D", DIMS)
def test_repulsion_converges_to_exact_as_theta_shrinks(D):
    rng = np.random.default_rng(7 + D)
    y = (rng.standard_normal((400, D)) * 3).astype(np.float32)
    ref, z = oracle.exact_repulsion_f64(y)
    errs = []
    for theta in (1.0, 0.5, 0.1, 0.001):
        neg, q, s = oracle.bh_repulsion(y, theta)
        errs.append(float(np.sqrt(((neg - ref) ** 2).sum() / (ref ** 2).sum())))
        if theta == 0.001:
            assert abs(s - z) / z < 1e-5
    assert errs[-1] < 1e-5 and errs[1] < 0.1
    assert all(a >= b for a, b in zip(errs, errs[1:]))


def test_rejects_unsupported_dimension():
    with pytest.raises(ValueError):
        oracle.bh_repulsion(np.zeros((5, 8), np.float32), 0.5)
    with pytest.raises(ValueError):
        oracle.bh_build(np.zeros((5, 1), np.float32))


@pytest.mark.parametrize("D,din,seed", [(3, 5, 43), (4, 5, 42), (5, 6, 44), (6, 7, 45),
                                         (7, 8, 46)])
def test_run_loop_in_d_dimensions(D, din, seed):
    # test.rs:1914-2004: N = 200, perplexity 10, 50 epochs of Barnes-Hut theta 0.5; the embedding
    # stays finite and holds N * D values (seeded start instead of the unseeded random init)
    x = lcg_samples(200, din, seed)
    rows, cols, vals = oracle.build_affinities(x, 10.0)
    vals = vals.copy()
    oracle.normalize_p(vals, 12.0)
    y = (np.random.default_rng(seed).standard_normal((200, D)) * 1e-2).astype(np.float32)
    y0 = y.copy()
    uy, g = np.zeros_like(y), np.ones_like(y)
    oracle.run_epochs(y, uy, g, rows, cols, vals, method="bh", theta=0.5,
                      learning_rate=oracle.auto_learning_rate(200, 12.0), e0=0, e1=50)
    assert y.shape == (200, D) and np.all(np.isfinite(y))
    assert np.abs(y.mean(0)).max() < 1e-4  # zero-mean after every epoch
    assert np.abs(y - (y0 - y0.mean(0))).max() > 1e-3  # moved from the start

"""Pins the oracle's Philox4x32-10 and its draw maps (CPU)."""
import json
import os

import numpy as np

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))

# Random123 known-answer vectors for philox4x32_10 (Salmon et al., SC'11,
# distributed with Random123 as kat_vectors): counter, key -> output.
R123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_random123_kat():
    for ctr, key, out in R123_KAT:
        got = oracle.philox_block(ctr, *key)
        assert tuple(int(x) for x in got) == out


def test_rocrand_golden():
    with open(os.path.join(HERE, "golden", "philox_rocrand.json")) as f:
        vecs = json.load(f)["vectors"]
    assert len(vecs) == 64
    for v in vecs:
        got = oracle.philox_block(v["ctr"], *v["key"])
        assert [int(x) for x in got] == v["out"]


def test_move_stream_statistics():
    d, n = 37, 200000
    mv = oracle.move_stream(d, n, seed=11, pair_prob=0.3)
    types, i, j = mv[:, 0], mv[:, 1], mv[:, 2]
    assert set(np.unique(types)) <= {0, 1}
    frac = (types == 1).mean()
    assert abs(frac - 0.3) < 4 * np.sqrt(0.3 * 0.7 / n)
    assert i.min() == 0 and i.max() == d - 1
    counts = np.bincount(i, minlength=d)
    # chi-square for uniform coordinates
    exp = n / d
    chi2 = ((counts - exp) ** 2 / exp).sum()
    assert chi2 < d + 6 * np.sqrt(2 * d)
    pair = types == 1
    assert np.all(j[pair] != i[pair]) and np.all((j[pair] >= 0) & (j[pair] < d))
    assert np.all(j[~pair] == -1)


def test_move_stream_ksparse_and_groups():
    d = 10
    a = oracle.move_stream(d, 5000, seed=3, ksparse_prob=0.2, ksparse_k=4)
    frac = (a[:, 0] == 2).mean()
    assert abs(frac - 0.2) < 0.03
    b = oracle.move_stream(d, 5000, seed=3, group=1, ksparse_prob=0.2, ksparse_k=4)
    assert (a != b).any()
    c = oracle.move_stream(d, 5000, seed=3, variant=5, ksparse_prob=0.2, ksparse_k=4)
    assert (a != c).any()
    again = oracle.move_stream(d, 5000, seed=3, ksparse_prob=0.2, ksparse_k=4)
    assert (a == again).all()


def test_move_stream_pair_mode2():
    d = 6
    pi = np.array([0, 2, 4], dtype=np.int32)
    pj = np.array([1, 3, 5], dtype=np.int32)
    th = np.array([0.1, 0.7, 1.3])
    pc, ps = np.cos(th).astype(np.float32), np.sin(th).astype(np.float32)
    mv = oracle.move_stream(d, 30000, seed=5, pair_prob=1.0, pair_mode=2, pair=(pi, pj, pc, ps))
    assert (mv[:, 0] == 1).all()
    pairs = set(map(tuple, mv[:, 1:].tolist()))
    assert pairs == {(0, 1), (2, 3), (4, 5)}


def test_pair_mode2_without_schedule_falls_back():
    mv = oracle.move_stream(6, 1000, seed=5, pair_prob=1.0, pair_mode=2)
    assert (mv[:, 0] == 0).all()

"""Pins the CPU oracle against values the reference repository holds and
against what the mathematics fixes (CPU only)."""
import itertools
import math
import struct

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import PINS, frozen_oracle_config, quickstart_oracle


def test_frozen_trajectory_bit_exact():
    """crates/cpm-core/tests/dimension_consistency.rs:113-131"""
    f = PINS["frozen_trajectory"]
    m = O.Model(frozen_oracle_config(O))
    m.run_mcs(f["mcs"])
    assert m.mcs == 500
    assert m.volume().tolist() == f["volume"]
    assert m.interface().tolist() == f["interface"]
    assert struct.unpack("<Q", struct.pack("<d", m.energy))[0] == f["conservative_energy_bits"]
    assert m.model_hash() == f["model_hash"]


def test_frozen_trajectory_with_mirror_halo_is_identical():
    """Updating only a target's halo mirrors equals the reference's whole-halo
    rescan (lattice.rs:240-254) when the halo was consistent before."""
    a = O.Model(frozen_oracle_config(O))
    b = O.Model(frozen_oracle_config(O), halo_mirror=True)
    a.run_mcs(300)
    b.run_mcs(300)
    assert np.array_equal(a.padded_lattice(), b.padded_lattice())
    assert a.energy == b.energy


def test_derived_seed_known_answers():
    """notebooks/batch_execution_demo.ipynb:179-183"""
    d = PINS["derived_seeds"]
    assert [O.derive_seed_u64(d["master_seed"], i) for i in range(5)] == d["values"]


def test_blake3_published_vectors():
    v = PINS["blake3_vectors"]
    assert O.blake3_hex(b"") == v["empty"]
    assert O.blake3_hex(b"abc") == v["abc"]


def test_blake3_tree_shape_is_consistent_across_chunk_boundaries():
    """A one-byte change anywhere (first chunk, a middle chunk, the last
    partial chunk) changes the digest; lengths at chunk boundaries differ."""
    base = bytes(i % 251 for i in range(5000))
    h = O.blake3_hex(base)
    for pos in (0, 1023, 1024, 3000, 4999):
        mod = bytearray(base)
        mod[pos] ^= 1
        assert O.blake3_hex(bytes(mod)) != h
    digests = {O.blake3_hex(base[:n]) for n in (1023, 1024, 1025, 2048, 2049, 4096)}
    assert len(digests) == 6


def test_rng_seeding_and_range_sampling_statistics():
    """SplitMix64 seeding is fixed by the frozen trajectory; here: gen_range
    is uniform, its rejection rate for a power of two is 1/2, and
    gen_range_incl(hi) covers 0..=hi."""
    r = O.Rng(123)
    n = 60000
    xs = [r.gen_range(6) for _ in range(n)]
    counts = np.bincount(xs, minlength=6)
    assert counts.min() > n / 6 * 0.95 and counts.max() < n / 6 * 1.05
    r1, r2 = O.Rng(9), O.Rng(9)
    draws = 0
    for _ in range(20000):
        r1.gen_range(64)
    while True:  # count raw draws consumed by 20000 gen_range(64) calls
        if np.array_equal(r2.s, r1.s):
            break
        r2.next_u64()
        draws += 1
    assert abs(draws / 20000 - 2.0) < 0.05
    r3 = O.Rng(5)
    assert set(r3.gen_range_incl(3) for _ in range(400)) == {0, 1, 2, 3}
    f = [O.Rng(77).gen_f64() for _ in range(1)]
    assert 0.0 <= f[0] < 1.0


def test_simple_point_slow_and_fast_agree_2d_exhaustively():
    t = O.sp_table2d()
    assert all(O.lib().oc_sp_slow(2, p) == t[p] for p in range(256))
    assert "".join(str(int(b)) for b in t) == PINS["sp_table_2d_bits"]


def test_simple_point_slow_and_fast_agree_3d_random():
    """simple_point.rs:536-553"""
    rng = np.random.default_rng(0xC93D)
    w = O.sp_table3d_words()
    for p in rng.integers(0, 1 << 26, size=3000):
        p = int(p)
        slow = O.lib().oc_sp_slow(3, p)
        assert slow == O.lib().oc_sp_fast(3, p)
        assert slow == int((int(w[p >> 6]) >> (p & 63)) & 1)


def _idx(offs, o):
    return [tuple(x) for x in offs.tolist()].index(tuple(o))


def test_simple_point_hand_built_cases():
    """simple_point.rs:384-438, 585-628"""
    o2 = O.full_offsets(2)
    sp = lambda dim, members: O.lib().oc_sp_slow(  # noqa: E731
        dim, sum(1 << _idx(O.full_offsets(dim), m) for m in members))
    assert sp(2, [(0, -1)])
    assert not sp(2, [(0, -1), (1, 0)])
    assert not sp(2, [(0, -1), (0, 1)])
    assert not sp(2, [])
    assert not O.lib().oc_sp_slow(2, 0xFF)
    assert sp(2, [(0, -1), (1, -1), (1, 0)])
    assert not sp(3, [(1, 0, 0), (-1, 0, 0)])
    assert sp(3, [(1, 0, 0)])
    assert sp(3, [(1, 0, 0), (0, 1, 0), (1, 1, 0)])
    assert not sp(3, [(1, 0, 0), (0, 1, 0)])
    assert not sp(3, [(1, 1, 1)])
    assert o2.shape == (8, 2)


def test_simple_point_2d_ring_walk_cross_check():
    """simple_point.rs:449-485: the ring-walk count disagrees with the
    definition only for patterns with a diagonal-only object component."""
    offs = O.full_offsets(2)
    order = sorted(range(8), key=lambda i: math.atan2(offs[i][1], offs[i][0]))
    t = O.sp_table2d()
    for p in range(255):
        obj = [(p >> i) & 1 for i in range(8)]
        cyc = [obj[i] for i in order]
        if all(cyc):
            runs = 1
        elif not any(cyc):
            runs = 0
        else:
            runs = sum(1 for i in range(8) if cyc[i] and not cyc[i - 1])
        if (runs == 1) == bool(t[p]):
            continue
        # disagreement: some object component has no face neighbour of origin
        face = [i for i in range(8) if abs(offs[i][0]) + abs(offs[i][1]) == 1]
        members = [i for i in range(8) if obj[i]]
        comp_has_face = []
        seen = set()
        for s in members:
            if s in seen:
                continue
            stack, comp = [s], {s}
            while stack:
                a = stack.pop()
                for b in members:
                    if b not in comp and abs(offs[a][0] - offs[b][0]) + abs(offs[a][1] - offs[b][1]) == 1:
                        comp.add(b)
                        stack.append(b)
            seen |= comp
            comp_has_face.append(any(c in face for c in comp))
        assert not all(comp_has_face), f"pattern {p:08b}"


def test_simple_point_3d_table_is_invariant_under_cube_symmetries():
    """Simplicity is a geometric property: permuting and reflecting axes
    must map simple patterns to simple patterns (catches a transposed or
    mis-ordered offset basis)."""
    offs = O.full_offsets(3)
    w = O.sp_table3d_words()
    index = {tuple(o): i for i, o in enumerate(offs.tolist())}
    rng = np.random.default_rng(3)
    pats = [int(p) for p in rng.integers(0, 1 << 26, size=400)]
    # bias towards simple patterns (sparse random patterns are mostly not)
    pats += [int(p) & int(q) for p, q in zip(rng.integers(0, 1 << 26, 400),
                                             rng.integers(0, 1 << 26, 400))]
    bit = lambda p: (int(w[p >> 6]) >> (p & 63)) & 1  # noqa: E731
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            mp = [index[tuple(signs[d] * o[perm[d]] for d in range(3))] for o in offs.tolist()]
            for p in pats:
                q = 0
                for i in range(26):
                    if (p >> i) & 1:
                        q |= 1 << mp[i]
                assert bit(p) == bit(q)


def test_3d_table_digest():
    import hashlib
    w = O.sp_table3d_words()
    assert hashlib.sha256(w.astype("<u8").tobytes()).hexdigest() == PINS["sp_table_3d_sha256"]


def test_readme_quickstart_derived_phi():
    """notebooks/performance_demo.ipynb:24-41: density 0.278, status Ok."""
    cfg = quickstart_oracle(O, windows=(500, 500, 100))
    cfg.warn_on_default_init = False
    r = O.run_with(O.Model(cfg))
    assert f"{r.derived_phi:.3f}" == PINS["readme_quickstart"]["derived_phi_3dp"]
    assert r.status[0] == PINS["readme_quickstart"]["status"]


@pytest.mark.parametrize("boundary", ["periodic", "fixed"])
def test_incremental_bookkeeping_matches_recompute(boundary):
    """checker.rs:36-136: after many accepted moves, interface, contact,
    position sums (up to periodic unwrapping) and energy equal a
    from-scratch recomputation."""
    cfg = quickstart_oracle(O, seed=3)
    cfg.boundary = boundary
    m = O.Model(cfg)
    m.run_mcs(300)
    itf, sp, so, ct = m.recompute()
    assert np.array_equal(itf, m.interface())
    assert np.array_equal(ct, m.contact())
    assert abs(m.global_energy() - m.energy) <= 1e-6 * abs(m.global_energy())
    if boundary == "fixed":
        assert np.array_equal(sp, m.sum_position())
        assert np.array_equal(so, m.sum_outer())
    else:
        L = np.array(cfg.grid)
        assert np.all((sp - m.sum_position()) % L == 0)


def test_volume_and_interface_closed_forms():
    """energy.rs:92-103, 190-202 for a hand-set two-cell lattice: the
    global energy equals the closed-form sum over cells and contacts."""
    lat = np.zeros((6, 6), dtype=np.uint32)
    lat[1:3, 1:3] = 1
    lat[3:5, 2:5] = 2
    cfg = O.Config(grid=(6, 6), boundary="fixed", seed=0,
                   cell_types=[O.CellTypeSpec("a", 5, 6, 2.0, 0.5),
                               O.CellTypeSpec("b", 7, 9, 1.5, 0.25)],
                   cell_counts=[1, 1], adhesion=[[0.0, 3.0, 4.0], [3.0, 1.0, 2.0],
                                                 [4.0, 2.0, 1.5]],
                   init_lattice=lat.ravel().tolist(), init_cell_type_of=[0, 1])
    m = O.Model(cfg)
    v, i = m.volume(), m.interface()
    assert v.tolist() == [4, 6]
    # Moore interface by brute force over the padded lattice
    pad = np.zeros((8, 8), dtype=np.uint32)
    pad[1:7, 1:7] = lat
    bf = [0, 0]
    J = np.array(cfg.adhesion)
    adh = 0.0
    for r in range(1, 7):
        for c in range(1, 7):
            x = pad[r, c]
            for dr, dc in itertools.product((-1, 0, 1), repeat=2):
                if dr == dc == 0:
                    continue
                y = pad[r + dr, c + dc]
                if x and y != x:
                    bf[x - 1] += 1
                if y != x:
                    inner = 1 <= r + dr <= 6 and 1 <= c + dc <= 6
                    if not inner or (r + dr, c + dc) > (r, c):
                        adh += J[x][y]
    assert i.tolist() == bf
    e = 2.0 * (4 - 5) ** 2 + 1.5 * (6 - 7) ** 2 + 0.5 * (bf[0] - 6) ** 2 + 0.25 * (bf[1] - 9) ** 2
    assert m.energy == pytest.approx(e + adh, rel=0, abs=1e-12)


def test_theta_layout():
    """theta.rs:43-78 closed form and round trip."""
    assert [O.theta_len(n) for n in (1, 2, 4)] == [7, 14, 31]
    cfg = quickstart_oracle(O)
    th = O.theta_extract(cfg)
    assert th == [10.0, 2.0, 0.0, 5.0, 2.0, 0.0, 0.0]
    th2 = [1.0, 3.0, 0.0, 6.0, 1.0, 0.0, 2.6]
    assert O.theta_extract(O.theta_inject(cfg, th2)) == [1.0, 3.0, 0.0, 6.0, 1.0, 0.0, 3.0]


def test_shape_descriptors_closed_forms():
    """output.rs:565-611: a 2x1 domino has eigenvalues (1/4, 0); an
    axis-aligned 3D box has eigenvalues (n^2-1)/12 per axis."""
    outer = np.array([[1, 0], [0, 0]])  # sites (0,0),(1,0): sum x^2 = 1
    ev, s, k = O.shape_descriptors(outer, 2, [0.5, 0.0])
    assert ev == pytest.approx([0.25, 0.0]) and s == pytest.approx(0.25) and k == pytest.approx(1.0)
    sites = [(x, y, z) for x in range(4) for y in range(3) for z in range(2)]
    o3 = np.zeros((3, 3))
    for p in sites:
        o3 += np.outer(p, p)
    cen = np.mean(sites, axis=0)
    ev, s, k = O.shape_descriptors(o3, len(sites), cen.tolist())
    assert ev == pytest.approx([(16 - 1) / 12, (9 - 1) / 12, (4 - 1) / 12])

"""Pins of the oracle's Act term and Metropolis-Hastings acceptance (CPU only).

* Act (energy.rs:311-471): the hand-built two-cell scenarios of
  energy.rs:1048-1345 (fixed and periodic 6x6 lattices, lambda_Act = [2, 3],
  Max_Act = [10, 20], MCS 20), plus a fixed-boundary corner where the halo
  is retained as medium.  Each expected value is written out from the sites
  the geometric means retain, listed in the test, with the roots taken in
  60-digit decimal arithmetic.
* Act never reaches the conservative energy (monte_carlo.rs:265-270): after
  an Act run the tracked energy equals the from-scratch global energy.
* Metropolis-Hastings (monte_carlo.rs:168-222): on a 3x3 fixed lattice with
  one cell and the connectivity check off (exact_boltzmann.rs runs its
  checks the same way), the chain's visit frequencies over all 511
  non-empty patterns match exp(-H) normalised by exact enumeration
  (chi-squared; the energy parameters are chosen so that the chain relaxes
  within ~6 MCS, the slowest mode of its transition matrix, and samples are
  taken 30 MCS apart, so that they are close to independent).  The same chain under plain Metropolis, whose proposal
  asymmetry (n_A != n_B) is left uncorrected, fails the same test.
"""
from decimal import Decimal, getcontext

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O

getcontext().prec = 60

LAMBDA_ACT = [2.0, 3.0]
MAX_ACT = [10, 20]
MCS = 20
# von Neumann copy offsets in the oracle's order: (+1,0), (-1,0), (0,+1), (0,-1)
K = {(1, 0): 0, (-1, 0): 1, (0, 1): 2, (0, -1): 3}


def root(x: int, n: int) -> float:
    return float(Decimal(x) ** (Decimal(1) / Decimal(n))) if x else 0.0


def act_model(boundary, cells, births):
    """A 6x6 lattice with the given cells ({id: [(r, c), ...]}, cell i of
    type i - 1) and birth MCS ({(r, c): mcs})."""
    lat = np.zeros((6, 6), dtype=np.uint32)
    for cid, sites in cells.items():
        for r, c in sites:
            lat[r, c] = cid
    n = len(cells)
    cfg = O.Config(grid=(6, 6), boundary=boundary, seed=1,
                   cell_types=[O.CellTypeSpec(f"t{i}", 3, 8, 1.0, 1.0, LAMBDA_ACT[i], MAX_ACT[i])
                               for i in range(2)],
                   cell_counts=[1, 1], adhesion=[[0.0] * 3 for _ in range(3)],
                   init_lattice=lat.ravel().tolist(), init_cell_type_of=list(range(n)),
                   act=True)
    m = O.Model(cfg)
    b = np.zeros((6, 6), dtype=np.uint32)
    for (r, c), v in births.items():
        b[r, c] = v
    m.set_birth_mcs(b)
    return m


TWO_CELLS = {1: [(1, 1), (1, 2), (2, 1)], 2: [(2, 2), (2, 3), (3, 2)]}


def expected(lam, max_act, src_activities, tgt_activities):
    gs = root(int(np.prod(src_activities)), len(src_activities))
    gt = root(int(np.prod(tgt_activities)), len(tgt_activities))
    return -(lam / max_act) * (gs - gt)


def test_act_hand_built_fixed():
    """energy.rs:1084-1122: target (2,1) of cell 1, source (2,2) of cell 2
    (type 1: lambda 3, Max_Act 20).  GM_src keeps (2,2), (3,2), (2,3)
    (activities 18, 10, 19); GM_tgt keeps (2,1), (1,1) (15, 12)."""
    m = act_model("fixed", TWO_CELLS, {(2, 1): 15, (1, 1): 12, (2, 2): 18, (3, 2): 10,
                                       (2, 3): 19})
    got = m.act_delta((2, 1), K[(0, 1)], MCS)
    want = expected(3.0, 20, [18, 10, 19], [15, 12])
    assert abs(got - want) <= 1e-13 * abs(want)
    assert abs(got) > 1e-6


def test_act_zero_activity_zeroes_the_geometric_mean():
    """energy.rs:1124-1142: (3,2) born at MCS 0 has activity 20 - 20 = 0, so
    GM_src is 0 whatever the other sites hold."""
    m = act_model("fixed", TWO_CELLS, {(2, 1): 15, (1, 1): 12, (2, 2): 18, (3, 2): 0,
                                       (2, 3): 19})
    got = m.act_delta((2, 1), K[(0, 1)], MCS)
    assert got == -(3.0 / 20) * (0.0 - root(15 * 12, 2))


def test_act_commit_resets_birth_on_a_cell_to_cell_copy():
    """energy.rs:1144-1159 and 459-464: an accepted copy writes the current
    MCS into the target's birth, also when a cell takes a site from another
    cell.  Checked on a run: every site whose owner changed during MCS k
    carries a birth of at least k."""
    m = act_model("periodic", TWO_CELLS, {})
    before = m.lattice().copy()
    m.run_mcs(1)
    changed = m.lattice() != before
    assert changed.any()
    assert (m.birth_mcs()[changed] == 0).all()  # the first MCS is MCS 0
    before = m.lattice().copy()
    m.run_mcs(1)
    changed = m.lattice() != before
    b = m.birth_mcs()
    assert (b[changed] == 1).all()
    assert (b <= 1).all()


def test_act_gaining_medium_is_zero():
    """energy.rs:1161-1181: medium copying into cell 1 prices no Act."""
    m = act_model("fixed", TWO_CELLS, {(1, 1): 5, (1, 2): 7})
    assert m.act_delta((1, 1), K[(-1, 0)], MCS) == 0.0


def test_act_losing_medium_uses_the_mediums_own_births():
    """energy.rs:1183-1262: cell 1 (type 0: lambda 2, Max_Act 10) at (2,3)
    grows into medium (2,2).  GM_src keeps (2,3) alone (born at 0: activity
    0); GM_tgt keeps the medium sites (2,2), (3,2), (1,2), (2,1) (births
    18, 16, 17, 15: activities 8, 6, 7, 5)."""
    m = act_model("fixed", {1: [(2, 3)]}, {(2, 2): 18, (1, 2): 17, (3, 2): 16, (2, 1): 15})
    got = m.act_delta((2, 2), K[(0, 1)], MCS)
    want = expected(2.0, 10, [0], [8, 6, 7, 5])
    assert abs(got - want) <= 1e-13 * abs(want)
    assert want > 0


def test_act_fixed_halo_counts_as_medium_born_at_zero():
    """A fixed boundary's halo is medium with birth 0 (energy.rs:364-376):
    cell 1 at (0,1) grows into the corner (0,0) at MCS 4 (Max_Act 10):
    GM_tgt keeps (0,0) (birth 3: activity 9), (1,0) (birth 2: 8) and the two
    halo sites (activity 10 - 4 = 6 each)."""
    m = act_model("fixed", {1: [(0, 1), (1, 1)]}, {(0, 0): 3, (1, 0): 2, (0, 1): 1})
    got = m.act_delta((0, 0), K[(0, 1)], 4)
    # GM_src: (0,1) (birth 1: 7) and (1,1) (birth 0: 6); its halo neighbour is not cell 1
    want = expected(2.0, 10, [7, 6], [9, 8, 6, 6])
    assert abs(got - want) <= 1e-13 * abs(want)


def test_act_periodic_reads_through_the_halo():
    """energy.rs:1264-1344: the target (0,1)'s upward neighbour wraps to
    (5,1) of cell 1.  GM_src keeps (0,2), (1,2), (0,3) (18, 10, 19); GM_tgt
    keeps (0,1), (1,1), (5,1) (15, 12, 8).  With (5,1) born at 0 (activity
    0) GM_tgt is 0 and the delta moves."""
    cells = {1: [(0, 1), (1, 1), (5, 1)], 2: [(0, 2), (1, 2), (0, 3)]}
    births = {(0, 1): 15, (1, 1): 12, (5, 1): 8, (0, 2): 18, (1, 2): 10, (0, 3): 19}
    m = act_model("periodic", cells, births)
    got = m.act_delta((0, 1), K[(0, 1)], MCS)
    want = expected(3.0, 20, [18, 10, 19], [15, 12, 8])
    assert abs(got - want) <= 1e-13 * abs(want)
    births[(5, 1)] = 0
    m2 = act_model("periodic", cells, births)
    got2 = m2.act_delta((0, 1), K[(0, 1)], MCS)
    want2 = expected(3.0, 20, [18, 10, 19], [15, 12, 0])
    assert abs(got2 - want2) <= 1e-13 * abs(want2)
    assert abs(got2 - got) > 1e-6


def _sorting_cfg(seed, act, acceptance="metropolis", boundary="periodic"):
    J = [[0.0, 16.0, 16.0], [16.0, 4.0, 11.0], [16.0, 11.0, 6.0]]
    return O.Config(grid=(40, 40), boundary=boundary, seed=seed,
                    cell_types=[O.CellTypeSpec("a", 30, 24, 5.0, 1.0, 40.0, 20),
                                O.CellTypeSpec("b", 30, 24, 5.0, 1.0, 25.0, 30)],
                    cell_counts=[18, 18], adhesion=J, warn_on_default_init=False,
                    act=act, acceptance=acceptance)


@pytest.mark.parametrize("boundary", ["periodic", "fixed"])
def test_act_never_reaches_the_conservative_energy(boundary):
    m = O.Model(_sorting_cfg(5, True, boundary=boundary))
    plain = O.Model(_sorting_cfg(5, False, boundary=boundary))
    m.run_mcs(60)
    plain.run_mcs(60)
    assert abs(m.energy - m.global_energy()) <= 1e-9 * abs(m.global_energy())
    # the term changes the trajectory, and births stay within the run
    assert not np.array_equal(m.lattice(), plain.lattice())
    assert 40 <= m.birth_mcs().max() <= 59


def test_act_with_zero_lambda_leaves_the_trajectory_unchanged():
    """lambda_Act = 0 (Max_Act > 0) prices every copy at -0 * x = +-0, so the
    chain equals the Act-free one draw for draw."""
    cfg = _sorting_cfg(6, True)
    for t in cfg.cell_types:
        t.lambda_act = 0.0
    m, plain = O.Model(cfg), O.Model(_sorting_cfg(6, False))
    m.run_mcs(40)
    plain.run_mcs(40)
    assert np.array_equal(m.lattice(), plain.lattice())
    assert np.array_equal(m.rng_state(), plain.rng_state())
    assert m.energy == plain.energy


# ---------------------------------------------------------------- MH


def _tiny_cfg(acceptance, seed=3):
    return O.Config(grid=(3, 3), boundary="fixed", seed=seed,
                    cell_types=[O.CellTypeSpec("c", 5, 14, 0.05, 0.01)], cell_counts=[1],
                    adhesion=[[0.0, 0.1], [0.1, 0.0]],
                    init_lattice=[0, 0, 0, 0, 1, 0, 0, 0, 0], init_cell_type_of=[0],
                    acceptance=acceptance)


def _boltzmann():
    """exp(-H) over the 511 non-empty patterns, H from the global energy of
    each pattern (energy.rs:106-114, 205-213, 269-308)."""
    m = O.Model(_tiny_cfg("metropolis"))
    w = np.zeros(512)
    for p in range(1, 512):
        lat = np.array([(p >> i) & 1 for i in range(9)], dtype=np.uint32)
        m.load_state(lat, [0])
        w[p] = np.exp(-m.global_energy())
    return w / w.sum()


def _visits(acceptance, n_samples, thin_mcs=30, seed=3):
    m = O.Model(_tiny_cfg(acceptance, seed))
    m.set_connectivity_check(False)
    m.run_mcs(50)
    bits = 1 << np.arange(9, dtype=np.int64)
    counts = np.zeros(512, dtype=np.int64)
    for _ in range(n_samples):
        m.run_mcs(thin_mcs)
        counts[int((m.lattice().ravel().astype(np.int64) * bits).sum())] += 1
    return counts


def _chi2_pvalue(counts, probs):
    n = counts.sum()
    exp_counts = probs * n
    keep = exp_counts >= 10
    obs = np.append(counts[keep], counts[~keep].sum())
    ex = np.append(exp_counts[keep], exp_counts[~keep].sum())
    return stats.chisquare(obs, ex).pvalue, int(keep.sum())


def test_metropolis_hastings_samples_the_exact_boltzmann_distribution():
    probs = _boltzmann()
    counts = _visits("metropolis_hastings", 100_000)
    p, bins = _chi2_pvalue(counts, probs)
    print(f"[mh] chi-squared over {bins} bins: p = {p:.3g}")
    assert counts[0] == 0  # the cell never vanishes (min_cell_volume)
    assert bins > 300
    assert p > 1e-4


def test_uncorrected_metropolis_fails_the_same_check():
    probs = _boltzmann()
    counts = _visits("metropolis", 100_000)
    p, _ = _chi2_pvalue(counts, probs)
    print(f"[mh] plain Metropolis: p = {p:.3g}")
    assert p < 1e-8


def test_metropolis_hastings_rejects_without_a_draw_when_n_a_is_zero():
    """monte_carlo.rs:172-186: a target with no copy neighbour in the losing
    cell is rejected before the uniform draw.  On the 3x3 lattice with the
    cell's single site at the centre, copies of medium onto the centre have
    n_A = 0; the run's draw count is checked through a replay in which each
    attempt's draws are counted from the stream."""
    m = O.Model(_tiny_cfg("metropolis_hastings"))
    m.set_connectivity_check(False)
    m.run_mcs(1)
    # replay the nine attempts by hand from the same seed
    rng = O.Rng(3)
    lat = np.zeros((5, 5), dtype=np.int64)
    lat[2, 2] = 1
    vol = 1
    offs = [(1, 0), (-1, 0), (0, 1), (0, -1)]
    for _ in range(9):
        r, c = rng.gen_range(3) + 1, rng.gen_range(3) + 1
        dr, dc = offs[rng.gen_range(4)]
        L, G = lat[r, c], lat[r + dr, c + dc]
        if L == G or (L and vol <= 1):
            continue
        na = sum(lat[r + a, c + b] == L for a, b in offs)
        if na == 0:
            continue  # no draw
        u = rng.gen_f64()
        m2 = O.Model(_tiny_cfg("metropolis"))
        m2.load_state((lat[1:4, 1:4] != 0).astype(np.uint32).ravel(), [0])
        h0 = m2.global_energy()
        new = lat.copy()
        new[r, c] = G
        m2.load_state((new[1:4, 1:4] != 0).astype(np.uint32).ravel(), [0])
        dh = m2.global_energy() - h0
        nb = sum(lat[r + a, c + b] == G for a, b in offs)
        if u < min(na / nb * np.exp(-dh), 1.0):
            lat = new
            vol += 1 if G else -1
    assert np.array_equal(m.rng_state(), rng.s)
    assert np.array_equal(m.lattice(), (lat[1:4, 1:4] != 0).astype(np.uint32))

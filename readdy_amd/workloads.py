"""Synthetic inputs of the benchmark and test configurations.

c1_system / c1_chemistry: the single large A+B<->C system (periodic 200^3 box, 1,000,000 particles,
40 / 40 / 20 % A / B / C uniform random, seed 1; harmonic repulsion k=10 d=1 on all pairs, A-A weak
interaction k=10 d0=1 depth 0.5 cutoff 2, fusion A+B->C rate 0.1 radius 1, fission C->A+B rate 0.05
distance 1).

system / potentials / reactions: the bead-chain + crowder system with enzymatic binding, at any size.

Chains are 3-D random walks of unit steps (a step closer than 0.8 to the bead two or three back is
drawn again, otherwise Lennard-Jones would blow up); crowders sit on random, jittered sites of a
lattice of spacing >= 1, at least 0.9 from beads and from each other (uniform crowders would overlap
each other and blow Lennard-Jones up as well).  Species: 0 bead (topology flavour), 1 free crowder X_free, 2 bound crowder
X_bound.  Potentials: Lennard-Jones 12-6 (epsilon 1, sigma 1, cutoff 2.5, shifted) between crowders
and between crowders and beads, none between beads (the reference has no bonded exclusions, so
bead-bead pairs are left to the bonded terms); harmonic bonds k=100 l0=1, harmonic angles k=10
theta0=pi, cosine dihedrals k=1 n=3 phi0=0 along every chain.  Reaction: the enzymatic binding
X_free +(1.2) bead -> X_bound + bead (the bead is the catalyst, topologies stay untouched).
"""
import numpy as np

LJ = [12, 6, 2.5, 1.0, 1.0, 1.0]  # m, n, cutoff, shift, epsilon, sigma


def chains(n_chains, beads, box, rng):
    """(n_chains * beads, 3) bead positions wrapped into [-box/2, box/2)."""
    steps = rng.normal(size=(n_chains, beads, 3))
    steps /= np.linalg.norm(steps, axis=2, keepdims=True)
    for _ in range(100):  # redraw steps that fold back onto the bead two or three back
        x = np.cumsum(steps, axis=1)
        bad = np.zeros((n_chains, beads), dtype=bool)
        for back in (2, 3):
            d = np.linalg.norm(x[:, back:] - x[:, :-back], axis=2)
            bad[:, back:] |= d < 0.8
        if not bad.any():
            break
        new = rng.normal(size=(int(bad.sum()), 3))
        steps[bad] = new / np.linalg.norm(new, axis=1, keepdims=True)
    x = np.cumsum(steps, axis=1)
    x = x - x.mean(axis=1, keepdims=True) + (rng.random((n_chains, 1, 3)) - 0.5) * box
    return ((x + box / 2) % box - box / 2).reshape(-1, 3)


def crowders(n, beads_pos, box, rng, excl=0.9):
    """n positions at least excl from every bead and from each other (periodic): random sites of a
    cubic lattice of spacing >= 1, jittered by up to (spacing - excl) / 2 per axis, kept when no bead
    is within excl (checked through a cell grid)"""
    ns = int(box // 1.0)
    a = box / ns
    sites = rng.permutation(ns ** 3)[: min(ns ** 3, 2 * n + 64)]
    ijk = np.stack([sites // (ns * ns), (sites // ns) % ns, sites % ns], axis=1)
    cand = (ijk + 0.5) * a - box / 2 + (rng.random((len(sites), 3)) - 0.5) * (a - excl)
    nc = max(3, int(box // excl))
    cs = box / nc
    bc = np.floor((beads_pos + box / 2) / cs).astype(np.int64) % nc
    key = (bc[:, 0] * nc + bc[:, 1]) * nc + bc[:, 2]
    order = np.argsort(key, kind="stable")
    skey, spos = key[order], beads_pos[order]
    cc = np.floor((cand + box / 2) / cs).astype(np.int64) % nc
    ok = np.ones(len(cand), dtype=bool)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                nk = (((cc[:, 0] + dx) % nc) * nc + (cc[:, 1] + dy) % nc) * nc + (cc[:, 2] + dz) % nc
                lo, hi = np.searchsorted(skey, nk, "left"), np.searchsorted(skey, nk, "right")
                for k in range(int((hi - lo).max(initial=0))):
                    has = lo + k < hi
                    j = np.where(has, lo + k, 0)
                    d = cand - spos[j]
                    d -= box * np.round(d / box)
                    ok &= ~(has & ((d * d).sum(1) < excl * excl))
    out = cand[ok][:n]
    if len(out) < n:
        raise ValueError("box too small for the crowders")
    return out


def system(n_chains, beads, n_crowders, box, seed):
    """positions, types (beads first), bonds, angles, dihedrals as index tuples"""
    rng = np.random.default_rng(seed)
    b = chains(n_chains, beads, box, rng)
    c = crowders(n_crowders, b, box, rng)
    pos = np.concatenate([b, c])
    types = np.concatenate([np.zeros(len(b), np.uint32), np.ones(len(c), np.uint32)])
    idx = np.arange(n_chains * beads).reshape(n_chains, beads)
    bonds = np.stack([idx[:, :-1], idx[:, 1:]], axis=2).reshape(-1, 2)
    angles = np.stack([idx[:, :-2], idx[:, 1:-1], idx[:, 2:]], axis=2).reshape(-1, 3)
    dihedrals = np.stack([idx[:, :-3], idx[:, 1:-2], idx[:, 2:-1], idx[:, 3:]], axis=2).reshape(-1, 4)
    return pos, types, bonds, angles, dihedrals


def c1_system(n=1_000_000, L=200.0, seed=1):
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) - 0.5) * L
    types = rng.choice(np.array([0, 1, 2], dtype=np.uint32), n, p=[0.4, 0.4, 0.2])
    return pos, types


def c1_chemistry(L):
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(3) for b in range(a, 3)]
    pots.append((L.POT_WEAK_INTERACTION, 0, 0, [10.0, 1.0, 0.5, 2.0]))
    rxs = [dict(kind=L.RX_FUSION, educts=[0, 1], products=[2], rate=0.1, educt_distance=1.0),
           dict(kind=L.RX_FISSION, educts=[2], products=[0, 1], rate=0.05, product_distance=1.0)]
    return pots, rxs


def potentials(L):
    """crowder-crowder and crowder-bead Lennard-Jones (species 0 bead, 1 free, 2 bound)"""
    return [(L.POT_LENNARD_JONES, a, b, LJ) for a, b in [(1, 1), (1, 2), (2, 2), (0, 1), (0, 2)]]


def reactions(L, rate):
    return [dict(kind=L.RX_ENZYMATIC, educts=[1, 0], products=[2, 0], rate=rate, educt_distance=1.2)]

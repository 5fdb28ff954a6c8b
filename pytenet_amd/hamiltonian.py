"""
Input Hamiltonians as MPOs: the XXZ Heisenberg chain and the transverse-field
Ising chain (reference pytenet/hamiltonian/heisenberg.py:14-54 and
pytenet/hamiltonian/ising.py:14-69).

The reference builds these through its symbolic operator-graph machinery
(pytenet/opgraph.py:197-375, bipartite vertex cover). That machinery is
one-time host setup and not rebuilt here; instead the tensors it produces are
written down in closed form, with the same virtual-bond order, quantum numbers
and gauge, so environment blocks agree element by element with the reference.
The layouts below are pinned against tensors produced by the reference
(tests/golden/mpo_*.npz) for open chains of 1..10 sites, with and without a
field term.

Virtual-bond states of the XXZ MPO (quantum numbers are 2*Sz):
    up   (+2): an S+ was placed, waiting for its S-
    id   ( 0): nothing placed yet (identity string from the left)
    dn   (-2): an S- was placed, waiting for its S+
    z    ( 0): an Sz was placed, waiting for its partner (and carrying the field
               term of the site it was placed on)
    done ( 0): the term is complete (identity string to the right)
    last ( 0): right-edge state used on the final bond.
"""

import numpy as np
from .mpo import MPO

__all__ = ["heisenberg_xxz_1d_mpo", "ising_1d_mpo"]

_SUP = np.array([[0., 1.], [0., 0.]])
_SDN = np.array([[0., 0.], [1., 0.]])
_SZ = np.array([[0.5, 0.], [0., -0.5]])
_ID = np.identity(2)


def _assemble(qsite, bond_labels, qnum_of, site_terms):
    """Build MPO tensors from per-site transition lists
    [(left_label, right_label, 2x2 operator), ...]."""
    a = []
    for i, terms in enumerate(site_terms):
        left, right = bond_labels[i], bond_labels[i + 1]
        t = np.zeros((len(left), 2, 2, len(right)))
        for (x, y, op) in terms:
            t[left.index(x), :, :, right.index(y)] += op
        a.append(t)
    qbonds = [[qnum_of[lbl] for lbl in labels] for labels in bond_labels]
    return MPO.from_tensors(qsite, qbonds, a)


def heisenberg_xxz_1d_mpo(nsites: int, J: float, D: float, h: float) -> MPO:
    """XXZ chain sum_i J (X X + Y Y) + D Z Z - h Z with spin-1/2 operators,
    i.e. sum_i J/2 (S+ S- + S- S+) + D Sz Sz - h Sz, as an MPO with
    physical quantum numbers [1, -1] (heisenberg.py:14-54).
    Bond dimensions are [1, 4, 5, ..., 5, 4, 1] (3 on the single bond of a
    two-site chain without field).

    A zero J or D keeps the same layout (with zero blocks); the reference
    instead drops such terms and shrinks the bonds. Both represent the same
    operator."""
    assert nsites >= 2, "the reference builder requires at least two sites"
    qsite = [1, -1]
    qn = {"start": 0, "end": 0, "up": 2, "id": 0, "dn": -2, "z": 0, "done": 0, "last": 0}
    jh = 0.5 * J
    field = h != 0
    # operator completing a pending Sz: the Sz Sz coupling plus, through the
    # identity, the field term of the site where the pending Sz sits
    dz = D * _SZ + (-h) * _ID if field else D * _SZ
    L = nsites
    if L == 2:
        if field:
            b1 = ["up", "dn", "z", "id"]
            s0 = [("start", "up", _SUP), ("start", "dn", _SDN), ("start", "z", _SZ),
                  ("start", "id", _ID)]
            s1 = [("up", "end", jh * _SDN), ("dn", "end", jh * _SUP), ("z", "end", dz),
                  ("id", "end", (-h) * _SZ)]
        else:
            b1 = ["up", "dn", "z"]
            s0 = [("start", "up", _SUP), ("start", "dn", _SDN), ("start", "z", _SZ)]
            s1 = [("up", "end", jh * _SDN), ("dn", "end", jh * _SUP), ("z", "end", D * _SZ)]
        return _assemble(qsite, [["start"], b1, ["end"]], qn, [s0, s1])

    bulk = ["up", "id", "dn", "z", "done"]
    labels = [["start"], ["up", "id", "dn", "z"]] + [bulk] * (L - 3)
    labels.append(["up", "dn", "done", "last"] if field else ["up", "dn", "last", "done"])
    labels.append(["end"])

    sites = [[("start", "up", _SUP), ("start", "id", _ID), ("start", "dn", _SDN),
              ("start", "z", _SZ)]]
    for i in range(1, L - 2):
        terms = [("up", "done", jh * _SDN), ("id", "up", _SUP), ("id", "id", _ID),
                 ("id", "dn", _SDN), ("id", "z", _SZ), ("dn", "done", jh * _SUP),
                 ("z", "done", dz)]
        if "done" in labels[i]:
            terms.append(("done", "done", _ID))
        sites.append(terms)
    # site L-2: opens the right-edge state
    terms = [("up", "done", jh * _SDN), ("id", "up", _SUP), ("id", "dn", _SDN),
             ("dn", "done", jh * _SUP), ("z", "done", dz)]
    if field:
        terms += [("id", "done", (-h) * _SZ), ("id", "last", dz)]
    else:
        terms += [("id", "last", _SZ)]
    if "done" in labels[L - 2]:
        terms.append(("done", "done", _ID))
    sites.append(terms)
    # site L-1
    last_op = _SZ if field else D * _SZ
    sites.append([("up", "end", jh * _SDN), ("dn", "end", jh * _SUP), ("done", "end", _ID),
                  ("last", "end", last_op)])
    return _assemble(qsite, labels, qn, sites)


def ising_1d_mpo(nsites: int, J: float, h: float, g: float) -> MPO:
    """Ising chain sum_i J Z Z + h Z + g X with Pauli matrices, all quantum
    numbers zero (ising.py:14-69). Bond states: id (identity string from the
    left), done (term complete), z (a Z waiting for its partner).
    Bond dimensions [1, 3, ..., 3, 1]."""
    assert nsites >= 1
    sx = np.array([[0., 1.], [1., 0.]])
    sz = np.array([[1., 0.], [0., -1.]])
    qn = {"start": 0, "end": 0, "id": 0, "done": 0, "z": 0}
    local = h * sz + g * sx
    if nsites == 1:
        return _assemble([0, 0], [["start"], ["end"]], qn, [[("start", "end", local)]])
    labels = [["start"]] + [["id", "done", "z"]] * (nsites - 1) + [["end"]]
    sites = [[("start", "id", _ID), ("start", "done", local), ("start", "z", J * sz)]]
    for _ in range(1, nsites - 1):
        sites.append([("id", "id", _ID), ("id", "done", local), ("id", "z", J * sz),
                      ("done", "done", _ID), ("z", "done", sz)])
    sites.append([("id", "end", local), ("done", "end", _ID), ("z", "end", sz)])
    return _assemble([0, 0], labels, qn, sites)

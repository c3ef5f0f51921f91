"""The frame-to-frame matching restatement pinned on hand-made contact sets: identity on an
unchanged frame, -1 for a pair without previous contacts, -2 beyond the position or normal
threshold, one previous contact per new contact with the closest claimant winning (ties to the
lower index), and matching only inside a world."""

import numpy as np


def _frame(rows, worlds):
    """rows: (world, s0, s1, mid xyz, normal xyz); worlds: world count."""
    rows = sorted(rows, key=lambda r: r[0])
    cnt = np.bincount([r[0] for r in rows], minlength=worlds)
    return {"off": np.concatenate([[0], np.cumsum(cnt)[:-1]]), "cnt": cnt,
            "s0": np.array([r[1] for r in rows]), "s1": np.array([r[2] for r in rows]),
            "mid": np.array([r[3] for r in rows], dtype=np.float64).reshape(-1, 3),
            "n": np.array([r[4] for r in rows], dtype=np.float64).reshape(-1, 3)}


Z = (0.0, 0.0, 1.0)


def test_identity_new_pair_and_thresholds(O):
    prev = _frame([(0, 1, 2, (0, 0, 0), Z), (0, 1, 3, (1, 0, 0), Z), (1, 5, 6, (0, 0, 0), Z)], 2)
    assert list(O.match_contacts(prev, prev, 5e-4, 0.995)) == [0, 1, 2]
    cur = _frame([(0, 1, 2, (4e-4, 0, 0), Z),            # within 0.5 mm: matched
                  (0, 1, 3, (1.0 + 6e-4, 0, 0), Z),      # 0.6 mm: broken
                  (0, 1, 4, (2, 0, 0), Z),               # pair absent before: not found
                  (1, 5, 6, (0, 0, 0), (0.0, 0.2, 0.98))], 2)  # normal turned: broken
    assert list(O.match_contacts(prev, cur, 5e-4, 0.995)) == [0, -2, -1, -2]


def test_one_previous_contact_per_new_contact(O):
    prev = _frame([(0, 1, 2, (0, 0, 0), Z)], 1)
    cur = _frame([(0, 1, 2, (3e-4, 0, 0), Z), (0, 1, 2, (1e-4, 0, 0), Z), (0, 1, 2, (-1e-4, 0, 0), Z)], 1)
    # the two closest tie at 1e-4; the lower index of them wins, the others are broken
    assert list(O.match_contacts(prev, cur, 5e-4, 0.995)) == [-2, 0, -2]


def test_matching_stays_inside_a_world(O):
    prev = _frame([(0, 1, 2, (0, 0, 0), Z)], 2)
    cur = _frame([(1, 1, 2, (0, 0, 0), Z)], 2)
    assert list(O.match_contacts(prev, cur, 5e-4, 0.995)) == [-1]

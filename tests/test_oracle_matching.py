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


def test_sort_key_layout(O):
    """make_contact_sort_key's documented fields (contact_data.py:62-68): shape_a in bits 62..43,
    shape_b in 42..23, sub key in 22..0, bit 63 zero, each field masked to its width."""
    k = O.contact_sort_key
    assert int(k(1, 0, 0)) == 1 << 43 and int(k(0, 1, 0)) == 1 << 23 and int(k(0, 0, 1)) == 1
    assert int(k(0xFFFFF, 0xFFFFF, 0x7FFFFF)) == O.SORT_KEY_SENTINEL == (1 << 63) - 1
    assert int(k(0x100000 + 3, 0x100000 + 4, 0x800000 + 5)) == int(k(3, 4, 5))  # masked, not carried
    rng = np.random.default_rng(0)
    a, b, c = rng.integers(0, 1 << 20, 500), rng.integers(0, 1 << 20, 500), rng.integers(0, 1 << 23, 500)
    a[:50] = a[50:100]  # ties in the leading fields
    b[:25] = b[50:75]
    # the integer order of the keys is the lexicographic order of (shape_a, shape_b, sub)
    assert np.array_equal(np.argsort(k(a, b, c), kind="stable"), np.lexsort((c, b, a)))


def test_contact_order_ranks_within_a_pair(O):
    s0 = np.array([5, 5, 2, 5, 2, 9])
    s1 = np.array([9, 9, 7, 3, 7, 1])
    keys, perm = O.contact_order(s0, s1)
    assert list(keys & 0x7FFFFF) == [0, 1, 0, 0, 1, 0]
    # (2,7,0) (2,7,1) (5,3,0) (5,9,0) (5,9,1) (9,1,0)
    assert list(perm) == [2, 4, 3, 0, 1, 5]
    assert np.all(np.diff(keys[perm]) > 0)

"""The plain matcher decides acceptance on full channel sums; the reference stops
each channel's running row sum at the first row end above a stop value
(compute_ssd_flat, bm3d.rs:367-391, called from block_matching_joint,
bm3d.rs:318-350).  Both must accept the same candidates with the same
distances.  This checks the two decision rules against each other in float32
on row sums built to land on, and one or two ulps either side of, every stop
value (thr, fl(thr - d0), fl(thr - fl(d0 + d1))), where a rounding argument
would break first.  A rule that accepts at total == thr, or adds the channels
in another order, fails here."""
import numpy as np

F = np.float32


def _early(rows, stop):
    acc = F(0)
    for r in rows:
        acc = F(acc + r)
        if acc > stop:
            break
    return acc


def _full(rows):
    acc = F(0)
    for r in rows:
        acc = F(acc + r)
    return acc


def _reference(ch, thr):
    d0 = _early(ch[0], thr)
    if d0 > thr:
        return None
    d1 = _early(ch[1], F(thr - d0))
    if F(d0 + d1) > thr:
        return None
    d2 = _early(ch[2], F(thr - F(d0 + d1)))
    total = F(F(d0 + d1) + d2)
    return total if total < thr else None


def _full_rule(ch, thr, strict=True, order=0):
    d0, d1, d2 = (_full(c) for c in ch)
    if d0 > thr:
        return None
    if F(d0 + d1) > thr:
        return None
    total = F(F(d0 + d1) + d2) if order == 0 else F(d0 + F(d1 + d2))
    ok = total < thr if strict else total <= thr
    return total if ok else None


def _cases(rng, n):
    for _ in range(n):
        ch = rng.exponential(1.0, (3, 8)).astype(F) * F(rng.choice([1e-3, 1.0, 300.0]))
        ch[rng.random((3, 8)) < 0.2] = 0
        parts = [np.cumsum(c, dtype=F) for c in ch]
        kind = rng.integers(0, 4)
        if kind == 0:  # thr at a channel-0 row sum
            base = parts[0][rng.integers(0, 8)]
        elif kind == 1:  # thr at d0 + a channel-1 row sum
            base = F(parts[0][-1] + parts[1][rng.integers(0, 8)])
        elif kind == 2:  # thr at d0 + d1 + a channel-2 row sum
            base = F(F(parts[0][-1] + parts[1][-1]) + parts[2][rng.integers(0, 8)])
        else:
            base = F(rng.uniform(0, 3) * (parts[0][-1] + parts[1][-1] + parts[2][-1]))
        thr = F(base)
        for _ in range(int(rng.integers(-2, 3))):
            thr = np.nextafter(thr, F(np.inf) if rng.random() < 0.5 else F(0), dtype=F)
        yield ch, thr


def _agree(rule, n=40000, seed=0):
    rng = np.random.default_rng(seed)
    bad = 0
    for ch, thr in _cases(rng, n):
        a, b = _reference(ch, thr), rule(ch, thr)
        if (a is None) != (b is None) or (a is not None and a != b):
            bad += 1
    return bad


def test_full_sum_rule_matches_early_exit_rule():
    assert _agree(_full_rule) == 0


def test_planted_rule_changes_are_caught():
    assert _agree(lambda ch, thr: _full_rule(ch, thr, strict=False), n=20000) > 0
    assert _agree(lambda ch, thr: _full_rule(ch, thr, order=1), n=20000) > 0

"""The speculative split step reuses a parent's bound() samples for its
children: a child's edge samples must equal the parent's bit for bit.  The
split ranges are dyadic (k / 2^d, d <= 21 for every evaluated range), so the
FP32 sample positions u = umin + (umax - umin) * t, t in {0, 1/2, 1}, are
exact; this checks the identities the kernel relies on over random ranges of
every depth, in the float32 arithmetic the kernel uses."""
import numpy as np

F = np.float32


def sample_u(lo, hi, t):
    return F(lo + F(F(hi - lo) * F(t)))


def test_child_edge_samples_repeat_the_parent():
    rng = np.random.default_rng(5)
    for d in range(0, 22):
        n = 1 << d
        k = rng.integers(0, n, size=4000, dtype=np.int64)
        lo = (k / n).astype(F)
        hi = ((k + 1) / n).astype(F)
        mid = F(0.5) * (lo + hi)                  # split_children's cut
        for a, b in zip(lo, hi):
            c = F(F(a + b) * F(0.5))
            # parent column u at t = 0, 1/2, 1
            p0, p1, p2 = sample_u(a, b, 0.0), sample_u(a, b, 0.5), sample_u(a, b, 1.0)
            assert p0 == a and p2 == b and p1 == c
            # child 0 = [a, c], child 1 = [c, b]: edges repeat the parent's samples
            assert sample_u(a, c, 0.0) == p0 and sample_u(a, c, 1.0) == p1
            assert sample_u(c, b, 0.0) == p1 and sample_u(c, b, 1.0) == p2
        assert np.all(mid == F(0.5) * (lo + hi))


def test_child_middle_sample_is_exact_up_to_the_depth_limit():
    # evaluated children sit at depth <= 21 per axis: their middle sample
    # needs 23 significant bits at most
    for d in (19, 20, 21):
        n = 1 << d
        for k in (0, 1, n // 3, n - 2, n - 1):
            a, b = F(k / n), F((k + 1) / n)
            m = sample_u(a, b, 0.5)
            assert float(m) == (k + 0.5) / n

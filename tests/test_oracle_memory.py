"""The oracle's fields with memory (oracle.z2_memory_phi / z2_memory_log_ratio), pinned by what
the definition fixes rather than by restating it: without coupling (b = 0) they are the tilted
sampler's fields and ratios; the implied proposal probabilities of every field history sum to
one (a ratio that used the memory after its update, a sign slip or a clamp that the draw and
the ratio disagree on breaks that); and the fixed-point memory follows the closed-form
geometric sum."""
import itertools

import numpy as np


def test_without_coupling_it_is_the_tilted_sampler(orc):
    seed, c, T, V, a, r = 17, 5, 4, 64, 1 << 23, 29000
    phis, lr = orc.z2_memory_phi(seed, c, T, V, a, 0, r, stream=2)
    thr = (1 << 31) + a
    for t in range(T):
        assert np.array_equal(phis[t], orc.z2_tilt_phi(seed, c, t, V, thr))
    want = orc.z2_tilt_log_ratio(phis, thr / 2.0 ** 32)
    np.testing.assert_allclose(lr, want, rtol=1e-12, atol=1e-12)
    assert np.allclose(orc.z2_memory_log_ratio(phis, a, 0, r), lr, rtol=1e-13, atol=1e-13)


def test_proposal_probabilities_sum_to_one():
    """One site, T = 10 slices, every one of the 2^10 histories: sum of 2^-T exp(-log(p/q))
    = sum of q(history) = 1.  Parameters strong enough that the clamp at 2^25 binds."""
    import oracle

    T = 10
    for a, b, r in ((1 << 22, 3000, 30000), (-(1 << 24), 40000, 31457), (0, -20000, 16384)):
        tot = 0.0
        for h in itertools.product((-1, 1), repeat=T):
            phis = np.array(h + (1,), dtype=np.int8)[:, None]  # one extra slice, not counted
            lr = oracle.z2_memory_log_ratio(phis, a, b, r)
            tot += 2.0 ** -T * np.exp(-lr[T])
        assert abs(tot - 1.0) < 1e-12, (a, b, r, tot)


def test_memory_follows_the_geometric_sum():
    """A field that is +1 at every slice: M(t) = 1024 (1 - rho^t) / (1 - rho) up to the
    fixed-point rounding (at most 1/2 per slice, so 0.5 / (1 - rho) in all); read back
    through the ratio of a sampler with lambda = 0: -log1p(b M / 2^31) per slice."""
    import oracle

    T, b, r = 40, 1000, 28000
    rho = r / 2.0 ** 15
    phis = np.ones((T + 1, 1), dtype=np.int8)
    lr = oracle.z2_memory_log_ratio(phis, 0, b, r)
    inc = np.diff(lr)  # slice t's term
    M = np.expm1(-inc) * 2.0 ** 31 / b  # invert -log1p(b M / 2^31)
    t = np.arange(T)
    closed = 1024 * (1 - rho ** t) / (1 - rho)
    assert np.abs(M - closed).max() <= 0.5 / (1 - rho) + 1e-6


def test_draw_frequencies_follow_the_memory(orc):
    """With a strong coupling the draws copy the site's past: the fraction of sites equal to
    their previous value is (1 + <|eps|>) / 2 in expectation, above 1/2 by far more than the
    noise, and exactly 1/2 + O(noise) without coupling."""
    seed, V, T = 3, 4096, 6
    a, b, r = 0, 1 << 14, 16384  # eps = b M / 2^31 reaches the 2^-6 clamp after one slice
    phis, _ = orc.z2_memory_phi(seed, 1, T, V, a, b, r)
    same = (phis[1:] == phis[:-1]).mean()
    assert abs(same - (1 + 2.0 ** -6) / 2) < 4 * 0.5 / np.sqrt(V * (T - 1))
    phis0, _ = orc.z2_memory_phi(seed, 1, T, V, 0, 0, r)
    assert abs((phis0[1:] == phis0[:-1]).mean() - 0.5) < 4 * 0.5 / np.sqrt(V * (T - 1))

"""Tilted Z2 fields (nrc_ensemble_set_tilt / nrc_ensemble_log_ratio): the importance-sampled
field of production/one_body-reweight/main.C:73-76 in Z2 form.  The fields the GPU draws,
replayed through the oracle, give the same determinants; the log ratios are the oracle's for
those fields; the reweighted correlator agrees with the plain sampler's, and with the tilt
fitted to a plain sample the per-sample variance falls."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import nrcps_amd as na

    c = na.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("prec,conj", [("f64", False), ("f32", True)])
def test_tilted_fields_replayed_by_oracle(ctx, orc, prec, conj):
    import nrcps_amd as na

    L, N, T, C, seed, c0, p = 8, 7, 6, 4, 91, 300, 0.53
    V = L ** 3
    print(f"[gpu] tilted fields {prec} conj={conj} L={L} N={N} T={T} q={p}", flush=True)
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    e = na.Ensemble(ctx, L, moms, kin, 0.4, C,
                    prec=na.PREC_F64 if prec == "f64" else na.PREC_F32)
    e.set_symmetry(conj)
    q = e.set_tilt(p)
    thr = int(round(q * 2.0 ** 32))
    assert q == thr / 2.0 ** 32 and abs(q - p) < 2.0 ** -32
    la, ph, _ = e.run(seed, c0, T)
    lr = e.log_ratio(T)
    e.close()
    tol = 1e-10 if prec == "f64" else 5e-6
    for c in range(C):
        phi = np.stack([orc.z2_tilt_phi(seed, c0 + c, t, V, thr) for t in range(T)])
        ref = orc.run_config(L, N, T, moms, kin, A=0.4, field="array", phi=phi)
        assert np.abs(la[:, c] - ref["logabs"]).max() < tol
        assert np.abs(ph[:, c] - ref["phase"]).max() < tol
        want = orc.z2_tilt_log_ratio(phi, q)
        assert np.allclose(lr[:, c], want, rtol=1e-12, atol=1e-9)
    assert np.all(lr[0] == 0)


def test_tilt_bookkeeping(ctx, orc):
    import nrcps_amd as na

    L, N, C, T = 8, 7, 2, 3
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    e = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F32)
    for bad in (0.0, 1.0, -0.1, float("nan")):
        with pytest.raises(RuntimeError):
            e.set_tilt(bad)
    assert e.set_tilt(0.5) == 0.5
    e.run(1, 0, T)
    assert np.all(e.log_ratio(T) == 0)  # plain run
    e.set_tilt(0.6)
    e.run(1, 0, T)
    assert np.any(e.log_ratio(T) != 0)
    with pytest.raises(RuntimeError):
        e.log_ratio(T + 1)
    # caller fields are never tilted
    words = np.stack([[orc.z2_bits(1, c, t, L ** 3) for c in range(C)] for t in range(T)])
    e.run_fields(words, T)
    assert np.all(e.log_ratio(T) == 0)
    with pytest.raises(RuntimeError):  # a batch refuses a tilted ensemble
        na.Batch(ctx, [e])
    e.set_tilt(0.5)
    b = na.Batch(ctx, [e])
    with pytest.raises(RuntimeError):  # and the tilt cannot change inside a batch
        e.set_tilt(0.6)
    b.close()
    e.close()


def test_fitted_tilt_is_unbiased_and_cuts_variance(ctx, orc):
    """L = 8, N = 7, T = 16: the constant tilt fitted to 16384 plain configurations (their
    field sums and log-determinants), then 32768 tilted and 32768 plain configurations on
    fresh counters.  C(t) agrees within 3.5 sigma at every t measured and the per-sample
    variance of the reweighted estimate falls at least 1.5-fold (about 2.5 in practice) at
    the cost of the plain sampler."""
    import nrcps_amd as na
    from nrcps_amd import proposal as pr

    L, N, T, C, seed = 8, 7, 16, 16384, 7012
    V = L ** 3
    ts = [4, 8, 12, 15]
    print(f"[gpu] fitted tilt L={L} N={N} T={T} C={C}", flush=True)
    moms = orc.lowest_momenta(L, N)
    kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
    e = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F32)
    e.set_symmetry(True)
    la_fit, _, _ = e.run(seed, 0, T)
    lam = pr.best_constant_tilt(la_fit, pr.z2_sums(ctx, seed, 0, C, T, V), ts, V)
    la_p = np.concatenate([e.run(seed, (1 << 22) + k, T)[0] for k in (0, C)], axis=1)
    q = e.set_tilt(0.5 * (1.0 + np.tanh(lam)))
    parts = []
    for k in (0, C):
        la_k, _, _ = e.run(seed + 1, (1 << 24) + k, T)
        parts.append((la_k, e.log_ratio(T)))
    e.close()
    la_t = np.concatenate([a for a, _ in parts], axis=1)
    lr = np.concatenate([b for _, b in parts], axis=1)
    sp = pr.estimator_stats(la_p, None, ts)
    st = pr.estimator_stats(la_t, lr, ts)
    for t in ts:
        sig = np.hypot(sp[t]["rel_se"], st[t]["rel_se"])
        d = abs(st[t]["log_C"] - sp[t]["log_C"])
        gain = (sp[t]["rel_se"] / st[t]["rel_se"]) ** 2
        print(f"[gpu]   q={q:.6f} t={t}: |d log C|/sigma {d / sig:.2f}, variance gain "
              f"{gain:.2f}", flush=True)
        assert d < 3.5 * sig
        assert gain >= 1.5

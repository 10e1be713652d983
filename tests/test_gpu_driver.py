"""The class layer end to end on the GPU: the slater_det driver (the production loop written
against the reference's class API) on a small production-style argument set, checked
against the oracle fed with the same ranlxs stream, kinetic table, smeared interaction and
momenta.  FP64 build: agreement to 1e-10 relative per (n, t)."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "nrcps_amd", "drivers", "slater_det")
PI = 3.1415926535897932

ARGS = {
    "verbose.arg": "0\n0\n0\n0\n0\n",
    "do.arg": "6\n8\n8\n8\nBOUNDARY_TYPE_PRD\nBOUNDARY_TYPE_PRD\nBOUNDARY_TYPE_PRD\n1.0\n",
    "evo.arg": "0\n1\n2\n",
    "lattice.arg": "FIELD_TYPE_ZTWO\n",
    "random.arg": "RANDOM_TYPE_RANLXS2\nSEED_TYPE_INPUT\n7012\nrngs/ranlxs2\n",
    "interaction.arg": "5.0\nINTERACTION_TYPE_ONEMINUSXIINVSQ\n9.82275\n2.06919\n0.45723\n"
                       "0.00223528\n0.0451682\n0.00709158\n0.000244729\n",
    "kinetic.arg": "DISPERSION_TYPE_PERFECT\n1.18343\n",
    "propagator.arg": "1\nresults/prop\n",
    "one_body.arg": "SOURCE_TYPE_MOM\n1.0\n1.0\n1.0\n",
    "momenta.arg": "DISPERSION_TYPE_PERFECT\n0.3\n1.18343\n",
}


def test_slater_det_driver_matches_oracle(orc):
    assert os.path.exists(DRIVER), "the slater_det driver is not built"
    L, T, ncfg = 8, 6, 2
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("args", "results", "rngs"):
            os.makedirs(os.path.join(tmp, d))
        for name, text in ARGS.items():
            with open(os.path.join(tmp, "args", name), "w") as f:
                f.write(text)
        print("[gpu] running the slater_det driver (L=8, T=6, 2 configurations)", flush=True)
        r = subprocess.run([DRIVER], cwd=tmp, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        cut2 = (1.0 * PI) * (1.0 * PI)
        n, moms, par, om = orc.momenta(L, orc.DISP_PERFECT, 1.18343, 0.3, cutoff_sq=cut2)
        assert n == 7
        got = []
        for k in range(1, n + 1):
            rows = [list(map(float, line.split()))
                    for line in open(os.path.join(tmp, "results", f"slater.{k}")) if line.strip()]
            got.append(np.array(rows).reshape(ncfg, T, 2))
        got = np.stack(got)  # [n][block][t][re, im]
        state = open(os.path.join(tmp, "rngs", "ranlxs2.0"), "rb").read()
        sources = open(os.path.join(tmp, "results", "sources")).read().splitlines()
    # results/sources (production/slater_det/main.C:122-140): "n (partner): k1 k2 k3 omega"
    # with the stream's default formatting, omega = log(1 + eps(k)) of the kinetic dispersion
    assert sources == [f"{i} ({par[i]}): {moms[i][0]} {moms[i][1]} {moms[i][2]} {om[i]:g}"
                       for i in range(n)]
    kin = orc.kinetic_xi(L, orc.DISP_PERFECT, 1.18343, 0.5, cutoff_sq=cut2)
    cpl = [9.82275, 2.06919, 0.45723, 0.00223528, 0.0451682, 0.00709158, 0.000244729]
    f, rc = orc.interaction_table(L, 5.0, orc.INTER_ONEMINUSXIINVSQ, cpl, 1.0, cutoff_sq=cut2)
    assert rc == 0
    rng = orc.ranlxs(2, 7012)
    for j in range(ncfg):
        ref = orc.run_config(L, n, T, moms, kin, smear=f, minors=True, field="ranlxs", rng=rng)
        d2 = np.exp(2 * ref["logabs"])  # |det|^2 [t][n]
        for k in range(n):
            np.testing.assert_allclose(got[k, j, :, 0], d2[:, k], rtol=1e-10)
            assert np.all(got[k, j, :, 1] == 0.0)
    # the RNG state written before the last block: RandomType then long[105]
    assert len(state) == 4 + 8 * 105
    st = np.frombuffer(state[4:], dtype=np.int64)
    assert st[0] == 105 and st[101] == 397


def test_slater_det_driver_resumes_from_rng_state_file(orc):
    """SEED_TYPE_FILE (src/random/random.C:78-87,157-181): a first run writes the ranlxs state
    at the start of every block (RandomType, then long[105]); the state left after two blocks
    is the oracle stream's after (T + 1) V draws (the first field is drawn by the Lattice
    constructor).  A second run that reads it continues that stream: its block matches the
    oracle's run from the same state."""
    assert os.path.exists(DRIVER), "the slater_det driver is not built"
    L, T = 8, 6
    V = L ** 3
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("args", "results", "rngs"):
            os.makedirs(os.path.join(tmp, d))
        for name, text in ARGS.items():
            with open(os.path.join(tmp, "args", name), "w") as f:
                f.write(text)
        print("[gpu] slater_det driver: two blocks, then a resumed block", flush=True)
        r = subprocess.run([DRIVER], cwd=tmp, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        state = np.frombuffer(open(os.path.join(tmp, "rngs", "ranlxs2.0"), "rb").read()[4:],
                              dtype=np.int64).copy()
        with open(os.path.join(tmp, "args", "random.arg"), "w") as f:
            f.write("RANDOM_TYPE_RANLXS2\nSEED_TYPE_FILE\n7012\nrngs/ranlxs2\n")
        with open(os.path.join(tmp, "args", "evo.arg"), "w") as f:
            f.write("0\n1\n1\n")
        r = subprocess.run([DRIVER], cwd=tmp, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        cut2 = (1.0 * PI) * (1.0 * PI)
        n, moms, par, om = orc.momenta(L, orc.DISP_PERFECT, 1.18343, 0.3, cutoff_sq=cut2)
        got = []
        for k in range(1, n + 1):
            rows = [list(map(float, line.split()))
                    for line in open(os.path.join(tmp, "results", f"slater.{k}")) if line.strip()]
            assert len(rows) == 3  # two blocks, then the resumed one appended
            got.append(np.array(rows[2]).reshape(T, 2))
        got = np.stack(got)  # [n][t][re, im]
    # the oracle stream: one configuration (T V draws), then the next block's first field
    rng = orc.ranlxs(2, 7012)
    kin = orc.kinetic_xi(L, orc.DISP_PERFECT, 1.18343, 0.5, cutoff_sq=cut2)
    cpl = [9.82275, 2.06919, 0.45723, 0.00223528, 0.0451682, 0.00709158, 0.000244729]
    f, rc = orc.interaction_table(L, 5.0, orc.INTER_ONEMINUSXIINVSQ, cpl, 1.0, cutoff_sq=cut2)
    orc.run_config(L, n, T, moms, kin, smear=f, minors=True, field="ranlxs", rng=rng)
    orc.lib().orc_ranlxs_z2_field(C.byref(rng), V, np.zeros(V, np.int8))
    st = np.zeros(105, np.int64)
    orc.lib().orc_ranlxs_get_state(C.byref(rng), st)
    assert np.array_equal(st, state)
    ref = orc.run_config(L, n, T, moms, kin, smear=f, minors=True, field="ranlxs", rng=rng)
    d2 = np.exp(2 * ref["logabs"])
    for k in range(n):
        np.testing.assert_allclose(got[k, :, 0], d2[:, k], rtol=1e-10)


DRIVER3 = os.path.join(ROOT, "nrcps_amd", "drivers", "slater_det3")


@pytest.mark.parametrize("trapped", [False, True])
def test_slater_det3_driver_matches_oracle(orc, trapped):
    """The pair-sink driver (SlaterDet3 + TwoBody on the class layer, main-cum.C's loop):
    every determinant Run(n+1, -1, -1) and Run(n+1, -1, n) per slice and the six cumulant
    powers against the oracle's pair run with the same SHO sources (OneBody::Set), ranlxs
    stream, smeared interaction and TGND wave function; trapped: with the harmonic Potential
    (Potential::Evolve on the class layer, the K P I P K order) against the oracle's."""
    assert os.path.exists(DRIVER3), "the slater_det3 driver is not built"
    L, T, ncfg, N = 8, 6, 2, 5
    V = L ** 3
    rng = np.random.default_rng(8)
    lam = (1.0, 1.1, 0.9)
    src = np.zeros((N, V), np.complex128)
    lines = []
    for n in range(N):  # SHO sources: random coefficients of the lowest levels
        c = np.zeros((10, 10, 10), np.complex128)
        out = []
        for n1 in range(2):
            for n2 in range(2):
                for n3 in range(3):
                    v = complex(rng.standard_normal(), rng.standard_normal())
                    c[n1, n2, n3] = v
                    out.append(f"{n1} {n2} {n3} {v.real!r} {v.imag!r}\n")
        src[n] = orc.sho_source(c, L, *lam)
        lines.append(out)
    args = dict(ARGS)
    args["one_body.arg"] = f"SOURCE_TYPE_SHO\n{lam[0]}\n{lam[1]}\n{lam[2]}\n"
    args["propagator.arg"] = f"{N}\nresults/prop\n"
    args["two_body.arg"] = ("WAVEFUNC_TYPE_TGND\nDISPERSION_TYPE_STANDARD\n"
                            "DISPERSION_TYPE_STANDARD\n1.0\n1.0\n0.176776695\nresults/two\n")
    spring = (0.04, 0.03, 0.05)
    if trapped:
        args["potential.arg"] = ("POTENTIAL_FORM_HARMONIC\nPOTENTIAL_TYPE_EXP\n"
                                 f"{spring[0]}\n{spring[1]}\n{spring[2]}\n")
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("args", "results", "rngs", "sources"):
            os.makedirs(os.path.join(tmp, d))
        for name, text in args.items():
            with open(os.path.join(tmp, "args", name), "w") as f:
                f.write(text)
        for n in range(N):
            with open(os.path.join(tmp, "sources", f"source.{n}"), "w") as f:
                f.writelines(lines[n])
        print("[gpu] running the slater_det3 driver (L=8, T=6, N=5, 2 configurations)",
              flush=True)
        r = subprocess.run([DRIVER3], cwd=tmp, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]

        def rows(name):
            return np.array([list(map(float, line.split()))
                             for line in open(os.path.join(tmp, "results", name))
                             if line.strip()])
        det = np.stack([rows(f"slater.{m + 1}").reshape(ncfg, T, 2) for m in range(2 * N)])
        z = np.stack([[rows(f"z{k + 1}.{m + 1}").reshape(ncfg, T) for m in range(2 * N)]
                      for k in range(6)])
    cut2 = PI * PI
    kin = orc.kinetic_xi(L, orc.DISP_PERFECT, 1.18343, 0.5, cutoff_sq=cut2)
    cpl = [9.82275, 2.06919, 0.45723, 0.00223528, 0.0451682, 0.00709158, 0.000244729]
    f, rc = orc.interaction_table(L, 5.0, orc.INTER_ONEMINUSXIINVSQ, cpl, 1.0, cutoff_sq=cut2)
    assert rc == 0
    w = orc.two_body_wavefunc(L, orc.WAVEFUNC_TGND, 0.176776695, cutoff_sq=cut2)
    opp = orc.opposite_parity_index(L)
    ranlxs = orc.ranlxs(2, 7012)
    moms = np.zeros((N, 3), np.int32)  # unused: the sources are given
    for j in range(ncfg):
        pot = (orc.potential_table(L, orc.POT_FORM_HARMONIC, orc.POT_TYPE_EXP, spring, 0.5)
               if trapped else None)
        ref = orc.run_config(L, N, T, moms, kin, smear=f, field="ranlxs", rng=ranlxs,
                             pair_w=w, pair_opp=opp, sources=src, pot=pot)
        for m in range(2 * N):
            want = np.exp(ref["logabs"][:, m]) * ref["phase"][:, m]
            got = det[m, j, :, 0] + 1j * det[m, j, :, 1]
            np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-300)
            for k in range(6):
                np.testing.assert_allclose(z[k, m, j], ref["logabs"][:, m] ** (k + 1),
                                           rtol=1e-9, atol=1e-12)


DRIVER2 = os.path.join(ROOT, "nrcps_amd", "drivers", "slater_det2")


def test_slater_det2_driver_matches_oracle(orc):
    """The two-species pair-sink driver (SlaterDet2 + TwoBody on the class layer, the loop of
    production/slater_det2_untrapped/main-cum.C): momentum sources, species-2 orbitals
    reindexed by the opposite-parity map, Run(n+1, -1, -1) and Run(n+1, -1, first index of
    shell s) per slice, against determinants of the oracle's pair matrix and one-body
    overlaps on the same ranlxs fields (src/slater_det/slater_det2.C:11-150)."""
    assert os.path.exists(DRIVER2), "the slater_det2 driver is not built"
    L, T, ncfg = 8, 6, 2
    V = L ** 3
    args = dict(ARGS)
    args["two_body.arg"] = ("WAVEFUNC_TYPE_TGND\nDISPERSION_TYPE_STANDARD\n"
                            "DISPERSION_TYPE_STANDARD\n1.0\n1.0\n0.176776695\nresults/two\n")
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("args", "results", "rngs"):
            os.makedirs(os.path.join(tmp, d))
        for name, text in args.items():
            with open(os.path.join(tmp, "args", name), "w") as f:
                f.write(text)
        with open(os.path.join(tmp, "args", "two_body_def.arg"), "w") as f:
            f.write("1 0 0 0.25\n0 2 7 -0.125\n")  # Deform: additions to the wave function
        print("[gpu] running the slater_det2 driver (L=8, T=6, 2 configurations)", flush=True)
        r = subprocess.run([DRIVER2], cwd=tmp, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]

        def rows(name):
            return np.array([list(map(float, line.split()))
                             for line in open(os.path.join(tmp, "results", name))
                             if line.strip()])
        cut2 = PI * PI
        n, moms, par, om = orc.momenta(L, orc.DISP_PERFECT, 1.18343, 0.3, cutoff_sq=cut2)
        starts = [0] + [i for i in range(1, n) if om[i] != om[i - 1]]  # AccShellCount(s - 1)
        even = np.stack([rows(f"slater2_even.{k + 1}").reshape(ncfg, T, 10) for k in range(n)])
        odd = {(s, k): rows(f"slater2_odd_{s}.{k + 1}").reshape(ncfg, T, 2)
               for s, t0 in enumerate(starts) for k in range(t0, n)}
        srcs = open(os.path.join(tmp, "results", "sources")).read().splitlines()
    assert len(srcs) == n and srcs[1].startswith(f"1 ({par[1]}): ")
    kin = orc.kinetic_xi(L, orc.DISP_PERFECT, 1.18343, 0.5, cutoff_sq=cut2)
    cpl = [9.82275, 2.06919, 0.45723, 0.00223528, 0.0451682, 0.00709158, 0.000244729]
    f, rc = orc.interaction_table(L, 5.0, orc.INTER_ONEMINUSXIINVSQ, cpl, 1.0, cutoff_sq=cut2)
    assert rc == 0
    w = orc.two_body_wavefunc(L, orc.WAVEFUNC_TGND, 0.176776695, cutoff_sq=cut2)
    for k1, k2, k3, re in ((1, 0, 0, 0.25), (0, 2, 7, -0.125)):
        w[(k1 * L + k2) * L + k3] += re
    opp = orc.opposite_parity_index(L)
    rng = orc.ranlxs(2, 7012)
    for j in range(ncfg):
        phi = np.zeros((T, V), np.int8)
        for t in range(T):  # the Lattice's draws: site order, one field per slice
            orc.lib().orc_ranlxs_z2_field(C.byref(rng), V, phi[t])
        mp = orc.run_config(L, n, T, moms, kin, smear=f, field="array", phi=phi, pair_w=w,
                            pair_opp=opp, want_M=True)["M"]
        m1 = orc.run_config(L, n, T, moms, kin, smear=f, field="array", phi=phi,
                            want_M=True)["M"]  # [t][sink][orbital]
        for t in range(T):
            P = mp[t][:, par]  # phi[s][t] = TwoBody::Run(up_s, down_t), down_t = up_par(t)
            for k in range(n):
                want = np.linalg.det(P[:k + 1, :k + 1])
                got = even[k, j, t, 0] + 1j * even[k, j, t, 1]
                assert abs(got - want) <= 1e-9 * abs(want) + 1e-300, (j, t, k)
                z1 = np.log(want.real)
                assert even[k, j, t, 2] == pytest.approx(z1, rel=1e-9, abs=1e-12)
                for s, t0 in enumerate(starts):
                    if k < t0:
                        continue
                    Q = P[:k + 1, :k + 1].copy()
                    Q[:, t0] = m1[t][t0, :k + 1]  # psi1[s + t0 N] = <sink_t0 | up_s>
                    want = np.linalg.det(Q)
                    got = odd[s, k][j, t, 0] + 1j * odd[s, k][j, t, 1]
                    assert abs(got - want) <= 1e-9 * abs(want) + 1e-300, (j, t, k, s)


DRIVER_RW = os.path.join(ROOT, "nrcps_amd", "drivers", "one_body_reweight")


def test_one_body_reweight_driver_matches_oracle(orc):
    """Importance sampling with shifted Gaussian fields (production/one_body-reweight/
    main.C): Lattice::Refresh(phi0) draws polar Box-Muller Gaussians from ranlxs plus the
    shift (src/alg/lattice.C:111-165), the trapped slice K P I P K evolves the one propagator,
    and Propagator::Normalize divides it by the density ratio exp(V phi0 (<phi> - phi0/2))
    (src/alg/propagator.C:116-133, NotImplemented in the reference's GPU build).  The
    overlaps with a separate sink against the oracle fed the same fields and ratios."""
    assert os.path.exists(DRIVER_RW), "the one_body_reweight driver is not built"
    L, T, ncfg, phi0 = 8, 6, 2, 0.0022
    V = L ** 3
    cpl = [0.48725928786895517, 0.29804274431821154, -0.21167497672068536,
           0.040531132841112746, 0.0]
    spring = (0.01, 0.02, 0.015)
    args = {
        "verbose.arg": "0\n0\n0\n0\n0\n",
        "do.arg": f"{T}\n{L}\n{L}\n{L}\nBOUNDARY_TYPE_PRD\nBOUNDARY_TYPE_PRD\n"
                  "BOUNDARY_TYPE_PRD\n1.0\n",
        "evo.arg": f"0\n1\n{ncfg}\n",
        "lattice.arg": "FIELD_TYPE_GAUSS\n",
        "random.arg": "RANDOM_TYPE_RANLXS2\nSEED_TYPE_INPUT\n7012\nrngs/rng\n",
        "interaction.arg": "5.0\nINTERACTION_TYPE_ONEMINUSXIINVSQ\n"
                           + "".join(f"{c!r}\n" for c in cpl),
        "kinetic.arg": "DISPERSION_TYPE_PERFECT\n5.0\n",
        "potential.arg": "POTENTIAL_FORM_HARMONIC\nPOTENTIAL_TYPE_EXP\n"
                         f"{spring[0]}\n{spring[1]}\n{spring[2]}\n",
        "prop.arg": "1\nresults/prop\n",
        "one_body.arg": "SOURCE_TYPE_MOM\n1.0\n1.0\n1.0\n",
        "source.arg": "0 0 0 1.0 0.0\n",
        "sink.arg": "0 0 0 1.0 0.0\n1 0 0 0.5 0.25\n",
    }
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("args", "results", "rngs"):
            os.makedirs(os.path.join(tmp, d))
        for name, text in args.items():
            with open(os.path.join(tmp, "args", name), "w") as f:
                f.write(text)
        print("[gpu] running the one_body_reweight driver (L=8, T=6, 2 configurations)",
              flush=True)
        r = subprocess.run([DRIVER_RW], cwd=tmp, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        got = np.array([list(map(float, line.split()))
                        for line in open(os.path.join(tmp, "results", "prop")) if line.strip()])
        got = got.reshape(ncfg, T, 2)
    cut2 = PI * PI
    kin = orc.kinetic_xi(L, orc.DISP_PERFECT, 5.0, 0.5, cutoff_sq=cut2)
    f, rc = orc.interaction_table(L, 5.0, orc.INTER_ONEMINUSXIINVSQ, cpl, 1.0, cutoff_sq=cut2)
    assert rc == 0
    pot = orc.potential_table(L, orc.POT_FORM_HARMONIC, orc.POT_TYPE_EXP, spring, 0.5)
    src = np.zeros((1, V), np.complex128)
    src[0, 0] = 1.0
    snk = src.copy()
    snk[0, (1 * L + 0) * L + 0] = 0.5 + 0.25j
    rng = orc.ranlxs(2, 7012)
    orc.gauss_field(rng, V, 0.0)  # the Lattice constructor's draw
    moms = np.zeros((1, 3), np.int32)  # unused: the source is given
    for j in range(ncfg):
        fields = np.stack([orc.gauss_field(rng, V, phi0) for _ in range(T)])
        norms = np.exp(V * phi0 * (fields.mean(axis=1) - 0.5 * phi0))
        ref = orc.run_config(L, 1, T, moms, kin, smear=f, field="real", phi_real=fields,
                             sources=src, sinks=snk, norms=norms, pot=pot)
        want = np.exp(ref["logabs"][:, 0]) * ref["phase"][:, 0]
        np.testing.assert_allclose(got[j, :, 0] + 1j * got[j, :, 1], want, rtol=1e-10)
    assert np.all(np.abs(norms - 1) > 0)  # the ratios are not trivial


def _run_slater_det(args: dict, what: str):
    """Runs the slater_det driver on `args` (ARGS with overrides) in a scratch directory and
    returns {k: [block][t][re, im]} of results/slater.<k>."""
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("args", "results", "rngs"):
            os.makedirs(os.path.join(tmp, d))
        for name, text in {**ARGS, **args}.items():
            with open(os.path.join(tmp, "args", name), "w") as f:
                f.write(text)
        print(f"[gpu] running the slater_det driver ({what})", flush=True)
        r = subprocess.run([DRIVER], cwd=tmp, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        out = {}
        for name in os.listdir(os.path.join(tmp, "results")):
            if name.startswith("slater."):
                rows = [list(map(float, line.split())) for line in
                        open(os.path.join(tmp, "results", name)) if line.strip()]
                out[int(name.split(".")[1])] = np.array(rows)
        return out


def test_slater_det_driver_l64_matches_oracle(orc):
    """L = 64 (the side tests/hamiltonian_performance uses): the class layer's Fourier runs
    one launch per axis with the two-step 64-point transform; 7 momenta (Fermi energy 0.006),
    the smeared interaction, ranlxs fields, T = 2, one configuration, against the oracle."""
    L, T = 64, 2
    got = _run_slater_det({"do.arg": f"{T}\n{L}\n{L}\n{L}\nBOUNDARY_TYPE_PRD\nBOUNDARY_TYPE_PRD\n"
                                     "BOUNDARY_TYPE_PRD\n1.0\n",
                           "evo.arg": "0\n1\n1\n",
                           "momenta.arg": "DISPERSION_TYPE_PERFECT\n0.006\n1.18343\n"},
                          "L=64, T=2, 1 configuration")
    cut2 = np.pi ** 2
    n, moms, _, _ = orc.momenta(L, orc.DISP_PERFECT, 1.18343, 0.006, cutoff_sq=cut2)
    assert n == 7 and sorted(got) == list(range(1, n + 1))
    kin = orc.kinetic_xi(L, orc.DISP_PERFECT, 1.18343, 0.5, cutoff_sq=cut2)
    cpl = [9.82275, 2.06919, 0.45723, 0.00223528, 0.0451682, 0.00709158, 0.000244729]
    f, rc = orc.interaction_table(L, 5.0, orc.INTER_ONEMINUSXIINVSQ, cpl, 1.0, cutoff_sq=cut2)
    assert rc == 0
    ref = orc.run_config(L, n, T, moms, kin, smear=f, minors=True, field="ranlxs",
                         rng=orc.ranlxs(2, 7012))
    d2 = np.exp(2 * ref["logabs"])
    for k in range(1, n + 1):
        row = got[k].reshape(T, 2)
        np.testing.assert_allclose(row[:, 0], d2[:, k - 1], rtol=1e-10)


@pytest.mark.parametrize("shape", [(8, 12, 16), (16, 6, 48)])
def test_slater_det_driver_non_cubic_free_gas(orc, shape):
    """An X x Y x Z lattice (src/fourier/noarch/fourier.C:29-39 plans any shape) with the
    interaction's couplings at zero: plane waves are eigenvectors of every factor, so
    |det_k(t)|^2 = prod_{n < k} xi_half(p_n)^(-4 t) for the k lowest momenta of the X x Y x Z
    dispersion (the form the oracle's cubic run gives to 1e-14).  A transposed axis, stride or
    table index moves a wave onto another momentum and fails this."""
    X, Y, Z = shape
    T = 4
    got = _run_slater_det({"do.arg": f"{T}\n{X}\n{Y}\n{Z}\nBOUNDARY_TYPE_PRD\nBOUNDARY_TYPE_PRD\n"
                                     "BOUNDARY_TYPE_PRD\n1.0\n",
                           "evo.arg": "0\n1\n1\n",
                           "interaction.arg": "5.0\nINTERACTION_TYPE_ONEMINUSXIINVSQ\n0.0\n"},
                          f"{X}x{Y}x{Z} free gas")
    cut2 = np.pi ** 2
    n, moms, _, _ = orc.momenta(shape, orc.DISP_PERFECT, 1.18343, 0.3, cutoff_sq=cut2)
    assert n > 3 and sorted(got) == list(range(1, n + 1))
    kin = orc.kinetic_xi(shape, orc.DISP_PERFECT, 1.18343, 0.5, cutoff_sq=cut2)
    idx = (moms[:, 0] * Y + moms[:, 1]) * Z + moms[:, 2]
    for k in range(1, n + 1):
        row = got[k].reshape(T, 2)
        want = np.array([np.prod(kin[idx[:k]] ** (-4.0 * t)) for t in range(T)])
        np.testing.assert_allclose(row[:, 0], want, rtol=1e-11)
        assert np.all(row[:, 1] == 0.0)

"""The check build (libnrcps_hip_check.so, -DNRC_CHECK; selected with NRC_CHECK=1): the same
results as the default build on a normal run, including a captured batch; a NaN that reaches
an overlap matrix, and a pair_opp site outside the lattice, become NRC_ERR_CHECK with the
failing check named (the site array is checked before any kernel reads it).  Run in a child
process, since a process loads one of the two libraries."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import sys
sys.path.insert(0, ROOT)
import numpy as np
import nrcps_amd as na
import oracle as orc

ctx = na.Context(0)
assert ctx.lib.nrc_check_build() == 1, "not the check build"
L, N, C, T = 8, 7, 2, 4
moms = orc.lowest_momenta(L, N)
kin = orc.kinetic_xi(L, orc.DISP_STANDARD, 1 / 0.05, 0.5)
e = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F64)
la, ph, _ = e.run(7, 0, T)
for c in range(C):
    ref = orc.run_config(L, N, T, moms, kin, A=0.4, seed=7, config=c)
    assert np.abs(la[:, c] - ref["logabs"]).max() < 1e-10
g = na.Ensemble(ctx, L, moms, kin, 0.4, C, prec=na.PREC_F32)
b = na.Batch(ctx, [g])
b.run(7, [0], T, graph=True)
lg, _, _ = g.results(T)
assert np.abs(lg - la).max() < 1e-5
b.close(); g.close(); e.close()
print("[child] normal runs pass the checks", flush=True)

ns, nt, V = 3, 4, 512
rng = np.random.default_rng(1)
phi = (rng.standard_normal((ns, V)) + 1j * rng.standard_normal((ns, V))).astype(np.complex64)
psi = (rng.standard_normal((1, nt, V)) + 1j * rng.standard_normal((1, nt, V))).astype(np.complex64)
psi[0, 2, 17] = np.nan
M = ctx.array((1, ns, nt), np.complex128)
try:
    ctx.overlap(ctx.to_device(phi), ctx.to_device(psi), M, ns, nt, V, 1, na.OVERLAP_BF16X6)
    raise SystemExit("NaN overlap not reported")
except na.NrcError as err:
    assert "(-8)" in str(err) and "overlap matrix" in str(err), str(err)
    print("[child] NaN reported:", err, flush=True)

w = np.ones(V)
opp = orc.opposite_parity_index(L).astype(np.int32)
opp[5] = V  # one site outside the lattice
q = ctx.array((1, nt, V), np.complex128)
Mp = ctx.array((1, nt, nt), np.complex128)
rc = ctx.lib.nrc_pair_overlap(ctx.h, q.ptr, q.ptr, ctx.to_device(w).ptr,
                              ctx.to_device(opp).ptr, Mp.ptr, nt, nt, V, 1, na.OVERLAP_F64)
assert rc == -8, rc
assert "pair_opp sites: 1 bad values" in ctx.lib.nrc_last_error_detail().decode()
print("[child] bad site refused:", ctx.lib.nrc_last_error_detail().decode(), flush=True)
ctx.close()
print("CHECK OK", flush=True)
'''


def test_check_build_passes_normal_runs_and_reports_bad_values():
    lib = os.path.join(ROOT, "nrcps_amd", "libnrcps_hip_check.so")
    assert os.path.exists(lib), "the check build is not built"
    env = dict(os.environ, NRC_CHECK="1")
    print("[gpu] check build child process", flush=True)
    r = subprocess.run([sys.executable, "-u", "-c", f"ROOT = {ROOT!r}\n" + CHILD], env=env,
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-3000:], flush=True)
    assert r.returncode == 0 and "CHECK OK" in r.stdout


def test_default_build_is_not_the_check_build():
    import nrcps_amd as na

    if os.environ.get("NRC_CHECK") == "1":
        pytest.skip("running under the check build")
    assert na.load_library().nrc_check_build() == 0

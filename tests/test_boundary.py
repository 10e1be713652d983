"""The drop-in boundary on the CPU: the reference's own drivers and tests compile and link,
unchanged, against the class layer in include/nrcps (libnrcps.so over the C ABI).  Only
compiles here (no GPU); the reference tree is read where it lies and nothing is kept."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
PKG = os.path.join(ROOT, "nrcps_amd")

# drivers and tests whose every call is on the path this build implements
COMPATIBLE = [
    "production/slater_det/main.C",
    "production/one_body/main.C",
    "tests/free_gas/main.C",
    "tests/fourier/main.C",
    "tests/ludcmp_test/main.C",
    "tests/momenta/main.C",
    "tests/dispersion/main.C",
    "tests/lattice/main.C",
    # pair sinks: TwoBody, SlaterDet2, SlaterDet3
    "production/two_body/main.C",
    "production/slater_det2_trapped/main-cum.C",
    "production/slater_det2_trapped/main-no_cum.C",
    "production/slater_det2_untrapped/main-cum.C",
    "production/slater_det2_untrapped/main-no_cum.C",
    "production/slater_det3_trapped/main-cum.C",
    # reweighted fields: Propagator::Normalize(Float), Lattice::Mean
    "production/one_body-reweight/main.C",
    "production/few_body_untrapped-reweight/main.C",
    # the other drivers and tests whose calls the class layer covers
    "production/few_body_untrapped/main.C",
    "production/few_body_untrapped_fast/main.C",
    "tests/dawson/main.C",
    "tests/hamiltonian_class/main.C",
    "tests/hamiltonian_performance/main.C",
    "tests/hermite/main.C",
    "tests/kappa_vs_energy/main.C",
    "tests/mpi/main.C",
    "tests/mpi_glb_sum/main.C",
    "tests/numerical_precision/main.C",
    "tests/one_body_performance/main.C",
    "tests/one_body_performance2/main.C",
    "tests/random/main.C",
    "tests/rngio/main.C",
    "tests/sho/main.C",
    "tests/sho+virial/main.C",
    "tests/slater_det_performance/main.C",
    "tests/toy_model/main.C",
    "tests/two_body_performance/main.C",
    "tests/two_body_performance2/main.C",
    "tests/wave_function/main.C",
]
# sources written for C++98 (a stream returned as bool): compiled as the reference builds them
GNU98 = {"production/few_body_untrapped/main.C", "production/few_body_untrapped_fast/main.C"}


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "production")),
                    reason="reference tree not present")
@pytest.mark.parametrize("src", COMPATIBLE)
def test_reference_driver_compiles_and_links_unchanged(src):
    if not os.path.exists(os.path.join(PKG, "libnrcps.so")):
        pytest.skip("libnrcps.so not built")
    with tempfile.TemporaryDirectory() as tmp:
        obj = os.path.join(tmp, "main.o")
        exe = os.path.join(tmp, "main")
        std = "-std=gnu++98" if src in GNU98 else "-std=c++17"
        r = subprocess.run(["g++", std, "-w", "-I", os.path.join(ROOT, "include", "nrcps"),
                            "-c", os.path.join(REF, src), "-o", obj], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        r = subprocess.run(["g++", "-o", exe, obj, f"-L{PKG}", "-lnrcps", "-lnrcps_hip"],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]


def test_class_layer_and_driver_are_built():
    lib = os.path.join(PKG, "libnrcps.so")
    drv = os.path.join(PKG, "drivers", "slater_det")
    if not os.path.exists(lib):
        pytest.skip("not built")
    out = subprocess.run(["nm", "-DC", lib], capture_output=True, text=True).stdout
    for sym in ["Fourier::Forward()", "Propagator::Run(", "SlaterDet::Run(int)",
                "OneBody::Project(double*)", "ludcmp(std::complex<double>*, int, int*, double*)",
                "Interaction::Initialize()", "Kinetic::Evolve(Fourier*)", "Lattice::Refresh()",
                "Momenta::Momenta(MomentaArg const&)", "Random::WriteState(char const*)",
                "TwoBody::Run(double*, double*)", "SlaterDet3::Initialize(TwoBody*)",
                "SlaterDet2::Run(int, int, int)", "DawsonF(double)"]:
        assert sym in out, sym
    assert os.path.exists(drv)

"""AddressSanitizer and UndefinedBehaviorSanitizer builds on the CPU: the FP64 oracle with a
driver that runs every one of its entries
(tests/sanitize/oracle_drive.c), and the host-only half of the class layer -- argument files,
job parameters, Dispersion, Momenta, ranlxs with its state files, Gaussian draws, special
functions -- with tests/sanitize/host_drive.cpp.  Any out-of-bounds access, use after free,
leak or undefined operation ends the run with a report and a non-zero status."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, (" ".join(cmd) + "\n" + r.stdout[-3000:] + r.stderr[-6000:])
    return r


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no gcc")
def test_oracle_under_address_and_ub_sanitizers():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "oracle_drive")
        _run(["gcc", "-std=gnu11", *SAN, "-I", os.path.join(ROOT, "oracle"),
              os.path.join(ROOT, "tests", "sanitize", "oracle_drive.c"),
              os.path.join(ROOT, "oracle", "oracle.c"), "-o", exe, "-lm", "-lpthread"])
        r = _run([exe], env=ENV, timeout=300)
        assert "oracle clean" in r.stdout


ARGS = {
    "verbose.arg": "0\n0\n0\n0\n0\n",
    "do.arg": "6\n8\n12\n16\nBOUNDARY_TYPE_PRD\nBOUNDARY_TYPE_PRD\nBOUNDARY_TYPE_PRD\n1.0\n",
    "evo.arg": "0\n1\n2\n",
    "kinetic.arg": "DISPERSION_TYPE_PERFECT\n1.18343\n",
    "interaction.arg": "5.0\nINTERACTION_TYPE_ONEMINUSXIINVSQ\n9.82275\n2.06919\n",
    "momenta.arg": "DISPERSION_TYPE_PERFECT\n0.3\n1.18343\n",
    "one_body.arg": "SOURCE_TYPE_MOM\n1.0\n1.0\n1.0\n",
}


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_host_class_layer_under_address_and_ub_sanitizers():
    lib = os.path.join(ROOT, "nrcps_amd", "libnrcps_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libnrcps_hip.so not built (the class layer links it for its device calls)")
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("args", "rngs"):
            os.makedirs(os.path.join(tmp, d))
        for name, text in ARGS.items():
            with open(os.path.join(tmp, "args", name), "w") as f:
                f.write(text)
        exe = os.path.join(tmp, "host_drive")
        pkg = os.path.join(ROOT, "nrcps_amd")
        _run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"),
              "-I", os.path.join(ROOT, "include", "nrcps"),
              os.path.join(ROOT, "tests", "sanitize", "host_drive.cpp"),
              os.path.join(pkg, "host", "util.cpp"), os.path.join(pkg, "host", "physics.cpp"),
              "-o", exe, f"-L{pkg}", "-lnrcps_hip", f"-Wl,-rpath,{pkg}"])
        # the HIP runtime the executable links (never called here) keeps allocations of its
        # own until exit: those are not this code's leaks
        sup = os.path.join(tmp, "lsan.supp")
        with open(sup, "w") as f:
            f.write("leak:libamdhip64\nleak:libhsa-runtime64\nleak:libamd_comgr\n")
        env = dict(ENV, LSAN_OPTIONS=f"suppressions={sup}")
        r = _run([exe], env=env, cwd=tmp, timeout=300)
        assert "host clean" in r.stdout
        assert os.path.exists(os.path.join(tmp, "rngs", "ranlxs.0"))

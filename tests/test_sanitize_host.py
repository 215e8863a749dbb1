"""The host C++ (proto decoders, text-format parser, qubit resolution, gate
building, fuser, pass scheduler, rebind recipes and the device re-evaluation
tables, whose indices are checked against the plan, parameter-shift
rewriting) under
AddressSanitizer and UndefinedBehaviorSanitizer, driven by a mutation fuzz of
valid programs and Pauli sums (tests/sanitize/host_fuzz.cc).  The analogue of
the reference's sanitizer run over core/src (scripts/msan_test.sh:16-35).
CPU only: g++ builds the host sources with the sanitizers."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "quantum_amd", "csrc")
SOURCES = ["wire.cc", "textproto.cc", "circuit.cc", "plan.cc", "large.cc", "ps_ops.cc"]


@pytest.fixture(scope="module")
def fuzz_binary(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the ROCm headers")
    out = str(tmp_path_factory.mktemp("sanitize") / "host_fuzz")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{CSRC}",
           f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "sanitize", "host_fuzz.cc")]
    cmd += [os.path.join(CSRC, s) for s in SOURCES] + ["-o", out, "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


@pytest.mark.parametrize("seed", [11, 12])
def test_host_code_is_clean_under_sanitizers(fuzz_binary, seed):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([fuzz_binary, "150000", str(seed)], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "circuits built" in r.stdout

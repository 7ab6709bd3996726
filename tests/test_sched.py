"""The request loop's scheduler (casphy_amd/csrc/sched.cpp: queue, admission
by device-memory budget, head-of-line reservation, batching, completion
bookkeeping) built with ThreadSanitizer and driven by tests/sched_check.cpp
from many submitting and waiting threads, with a stub runner in place of the
device work.  No GPU.

The build uses the LLVM clang++ of the ROCm tree: the GCC 11 ThreadSanitizer
runtime of this image reports a double lock on a plain std::mutex guarded by
std::lock_guard (a false report), so its verdicts cannot be trusted here."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module")
def sched_check(tmp_path_factory):
    cxx = CLANG if os.path.exists(CLANG) else shutil.which("clang++")
    if cxx is None:
        pytest.skip("clang++ not available")
    out = tmp_path_factory.mktemp("schedbin") / "sched_check"
    src = [os.path.join(ROOT, "tests", "sched_check.cpp"), os.path.join(ROOT, "casphy_amd", "csrc", "sched.cpp")]
    subprocess.check_call([cxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=thread",
                           "-o", str(out)] + src + ["-lpthread"])
    return str(out)


def test_scheduler_under_tsan(sched_check):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1")
    r = subprocess.run([sched_check], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "sched_check ok" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr

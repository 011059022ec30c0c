"""CPU-side race and memory checking (no GPU).

* The oracle's OpenMP loops under ThreadSanitizer, with LLVM's OpenMP
  runtime and its Archer tool (`make -C oracle sanitize`): the reference's own
  OpenMP build races on shared scratch and on achtemp (gppOpenMP3.cpp:306,
  :310); the oracle's restatement must not, and its threaded results must
  equal the sequential loop bit for bit.
* The library's host code -- input generators, the problem-file writer,
  header parser and reader (damaged files included), argument checks -- and
  the driver's command-line and file-header paths under AddressSanitizer and
  UndefinedBehaviorSanitizer (`make -C doe-csfg-hackathon_amd sanitize`;
  host side only, nothing reaches a GPU)."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(directory, timeout):
    jobs = str(min(8, os.cpu_count() or 1))
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, directory), f"-j{jobs}", "sanitize"],
                       capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, f"make sanitize in {directory} failed:\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}"
    return r.stdout


def test_oracle_openmp_under_tsan():
    if not os.path.exists("/opt/rocm/llvm/lib/libarcher.so"):
        pytest.skip("LLVM OpenMP with Archer not installed")
    out = _make("oracle", 300)
    assert "equal one thread bit for bit" in out


def test_host_code_under_asan_ubsan():
    out = _make("doe-csfg-hackathon_amd", 900)
    assert "gpp_host_selftest: all checks passed" in out
    assert "clean under ASan + UBSan" in out

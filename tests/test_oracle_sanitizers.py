"""The CPU oracle under sanitizers, on the CPU: every runner (2D f32/f64/u8,
CLOUDSIM, 3D f32/f64 with COPY and HALO faces, 16-bit storage with both
roundings, residuals, conversions, the MLP checker) on small ragged grids,
built with AddressSanitizer + UndefinedBehaviorSanitizer (gcc) and with
ThreadSanitizer over LLVM OpenMP with the Archer tool (clang), so the OpenMP
loops are checked for races.  oracle/Makefile target `sanitize`."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_asan_ubsan_tsan_clean():
    p = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True,
                       timeout=600)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-4000:]
    assert out.count("sanitize: ok") == 2, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
    assert "WARNING: ThreadSanitizer" not in out

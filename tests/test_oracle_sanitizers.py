"""The C oracle under sanitizers (CPU): every entry point of
oracle/fmm2d_oracle.c on small seeded fields (oracle/sanitize_driver.c),
built with AddressSanitizer + UndefinedBehaviorSanitizer (gcc) and, for its
OpenMP direct sum, with ThreadSanitizer (clang, LLVM OpenMP runtime with its
archer tool).  A control shows the ThreadSanitizer set-up reports a race: an
OpenMP loop that sums into one shared variable without a reduction."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
LLVM = "/opt/rocm/llvm"


def _make(target):
    return subprocess.run(["make", "-s", "-C", ORACLE, target], capture_output=True, text=True,
                          timeout=600)


def test_oracle_under_asan_and_ubsan():
    r = _make("asan")
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "all entry points ran" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


@pytest.mark.skipif(not os.path.exists(os.path.join(LLVM, "lib", "libarcher.so")),
                    reason="no LLVM OpenMP runtime with the archer tool")
def test_oracle_openmp_under_tsan():
    r = _make("tsan")
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "all entry points ran" in r.stdout
    assert "ThreadSanitizer" not in r.stderr


@pytest.mark.skipif(not os.path.exists(os.path.join(LLVM, "lib", "libarcher.so")),
                    reason="no LLVM OpenMP runtime with the archer tool")
def test_tsan_setup_reports_a_planted_race(tmp_path):
    src = tmp_path / "race.c"
    src.write_text(
        "#include <stdio.h>\n"
        "int main(void) { double s = 0.0;\n"
        "#pragma omp parallel for\n"
        "  for (int i = 0; i < 100000; ++i) s += 1.0;\n"
        "  printf(\"%f\\n\", s); return 0; }\n")
    exe = tmp_path / "race"
    subprocess.run([os.path.join(LLVM, "bin", "clang"), "-O1", "-g", "-fsanitize=thread",
                    "-fopenmp", str(src), "-o", str(exe), f"-L{LLVM}/lib",
                    f"-Wl,-rpath,{LLVM}/lib"], check=True, timeout=120)
    env = dict(os.environ, OMP_NUM_THREADS="4",
               OMP_TOOL_LIBRARIES=os.path.join(LLVM, "lib", "libarcher.so"),
               TSAN_OPTIONS="halt_on_error=1:exitcode=66:ignore_noninstrumented_modules=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 66 and "data race" in r.stderr

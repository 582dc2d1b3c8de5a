"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU
only): tests/sanitize/setup_sanitize runs the product's host setup (basis,
box mesh, geometric and cubature factors, masked ids, ogs T maps) and the
oracle's setup, Ax, gather, scatter and CG on sixteen configurations (BP1-BP6,
p = 1 ... 15, affine and perturbed hexahedra; quadrilateral, triangle and
tetrahedron boxes), comparing the two setups' arrays. It is
built by __graft_entry__.build() (g++ -fsanitize=address,undefined)."""
import os
import subprocess


HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "sanitize", "_build", "setup_sanitize")


def test_host_setup_and_oracle_clean_under_asan_ubsan():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-j8"], cwd=os.path.join(HERE, "sanitize"), check=True)
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([EXE], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitized setup ok (0 mismatches)" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr

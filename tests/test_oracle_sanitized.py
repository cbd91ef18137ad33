"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

The oracle's C source is built a second time with -fsanitize=address,undefined
(serial, oracle/Makefile target liboracle_san.so) and the whole derivative
chain (tree, density, cons2prim, force) runs on small cases that reach
every physics branch: periodic hydro, MHD with one-fluid dust, self-gravity
with block timesteps, and the quintic shock tube.  Any out-of-bounds access,
use after free or undefined operation ends the run with a report."""
import os
import shutil
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

CHILD = textwrap.dedent("""
    from oracle import oracle
    from phantom_amd import setups
    from tests import parity
    print("library", oracle.lib()._name, flush=True)
    cases = {
        "box8_periodic": lambda: setups.uniform_box(nx=8, jitter=0.1, seed=3, isothermal=False, vamp=0.5),
        "mhd_dust_8": lambda: setups.mhd_dust_box(nx=8, jitter=0.05),
        "polytrope_1500_bins": lambda: setups.with_bins(setups.polytrope(n=1500, theta=0.5), nbinmax=3,
                                                        istepfrac=2, dtmax=0.05),
        "shock_tube_ny4": lambda: setups.shock_tube(ny=4),
    }
    for name, mk in cases.items():
        o = parity.run_oracle(mk())
        print(name, "ok", int(o["dens"]["nneigh"].sum()), flush=True)
""")


def test_oracle_chain_under_sanitizers():
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc and make are needed to build the sanitized oracle")
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("this gcc has no AddressSanitizer runtime")
    subprocess.run(["make", "-s", "-B", "liboracle_san.so"], cwd=ORACLE, check=True)
    env = dict(os.environ, PHSPH_ORACLE_LIB=os.path.join(ORACLE, "liboracle_san.so"), LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=900)
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert "liboracle_san.so" in r.stdout
    assert r.stdout.count(" ok ") == 4

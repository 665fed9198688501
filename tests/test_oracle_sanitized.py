"""The oracle's C sources built under AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/Makefile, target san) and run over every kind of scene the tests use: boxes, spheres,
capsules and cylinders on a floor, batched worlds with point-constraint chains, a triangle-mesh
terrain with hull variants, throw worlds with forces, torques and iteration overrides, sleeping
and waking.  Any report aborts the child process (-fno-sanitize-recover=all), failing the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_SCRIPT = r"""
import numpy as np
from crashcat_amd import scene as SC
from crashcat_amd import shapes as SH
from oracle import refcpu as O

dt = 1.0 / 60.0
runs = [
    (SC.cube_heap(seed=1001, nx=4, ny=4, nz=4), None, 90),
    (SC.mixed_pile(seed=1002, n=800, extent=(10.0, 12.0, 10.0), cylinders=True), None, 120),
    (SC.batched_worlds(seed=1003, num_worlds=3, min_bodies=8, max_bodies=40), 1, 60),
    (SC.terrain_drop(seed=1004, vertices=24, spacing=0.5, n=60, drop=(1.0, 4.0), hull_variants=4), None, 60),
    (SC.throw_worlds(num_worlds=2), 1, 70),
]
for sc, world, steps in runs:
    o = O.OracleWorld(sc) if world is None else O.OracleWorld(sc, world)
    n = len(o.bodies())
    rng = np.random.default_rng(steps)
    for s in range(steps):
        if s % 10 == 3:
            o.add_forces(rng.normal(0, 50, (n, 3)), rng.normal(0, 5, (n, 3)), wake=True)
        o.step(dt if s % 7 else dt / 2)
    o.contacts()
    o.pairs()
    o.island_iterations(0)
print("sanitized oracle ok")
"""


def test_oracle_under_asan_and_ubsan():
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True, capture_output=True,
                       timeout=300)
    except (subprocess.CalledProcessError, FileNotFoundError) as e:
        pytest.skip(f"no sanitizer build here: {e}")
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("libasan runtime not found")
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               CC_ORACLE_LIB=os.path.join(ROOT, "oracle", "librefcpu_san.so"), PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-c", _SCRIPT], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and "sanitized oracle ok" in r.stdout, (r.stdout[-3000:], r.stderr[-6000:])

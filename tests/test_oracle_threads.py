"""The oracle's OpenMP loops (auxevol, rhs_eval, residual; every iteration writes only its own
grid points) give bit-identical results with 1 and 8 threads: the race check that stands in
for ThreadSanitizer, which needs an instrumented OpenMP runtime this image does not have."""
import os
import subprocess
import sys

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))

_PROBE = r"""
import hashlib, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import oracle
from nrpyelliptic_amd import config
O = oracle.OracleProblem(config.unequal_mass_q4((48, 32, 8)))
h = hashlib.sha256(O.aux.tobytes())
st = np.random.default_rng(3).uniform(-0.1, 0.1, 2 * O.ntot)
h.update(O.rhs(st).tobytes())
h.update(O.residual(st).tobytes())
s = O.new_state()
for _ in range(5):
    O.mol_step(s)
h.update(s["y_n"].tobytes())
for bc in ("sommerfeld6",):
    O2 = oracle.OracleProblem(config.unequal_mass_q4((48, 32, 8)), outer_bc=bc)
    s2 = O2.new_state()
    for _ in range(5):
        O2.mol_step(s2)
    h.update(s2["y_n"].tobytes())
print(h.hexdigest())
"""


def _digest(threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, "-c", _PROBE, ROOT], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    return r.stdout.strip().splitlines()[-1]


def test_oracle_is_thread_count_independent():
    one, many = _digest(1), _digest(8)
    assert len(one) == 64 and one == many

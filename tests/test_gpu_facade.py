"""The C++ facade (include/dgops.hpp) as a drop-in: a caller-side templated
LSRK integrator written against the reference's perform_stage concept
(tests/cpp/facade_lsrk.cpp, built by __graft_entry__.build()) runs three
LSRK(5,4) steps on the GPU; the result matches the oracle's
LowStorageRungeKuttaIntegrator to 1e-12."""
import os
import subprocess

import numpy as np
import pytest

from helpers import Cfg, rel_diff

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "facade_lsrk")


def test_facade_lsrk_driver_matches_oracle(tmp_path):
    assert os.path.exists(DRIVER), \
        "tests/cpp/facade_lsrk is missing: run __graft_entry__.build()"
    out = tmp_path / "u.bin"
    steps = 3
    r = subprocess.run([DRIVER, str(out), str(steps)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    u_gpu = np.fromfile(out, dtype=np.float64)
    c = Cfg(dim=3, degree=3, refine=2, perturbation=0.1, velocity_kind=1,
            solution_kind=3, solution_param=[0.5, 0.5, 0.5, 0.15])
    o = c.oracle_handle()
    u = o.project(0.0)
    ri, ki = o.zeros(), o.zeros()
    dt = 2e-3
    for s in range(steps):
        o.lsrk_time_step(1, s * dt, dt, u, ri, ki)
    assert u_gpu.size == u.size
    assert rel_diff(u_gpu, u) < 1e-12

"""Per-degree tile autotuner: every built tile shape (cells per workgroup)
computes the same right-hand side and LSRK stage as the default shape, bit
for bit (the arithmetic per cell does not depend on the shape); the tuner
times its candidates on meshes of 4096 cells and more and keeps the
fastest."""
import numpy as np
import pytest
import torch

from advection_miniapp_amd import AdvectionOperation
from helpers import Cfg

pytestmark = pytest.mark.gpu


def tile_cells(op):
    """Cells per workgroup of the built shapes, from a first launch."""
    dst = op.initialize_dof_vector()
    op.apply(op.initialize_dof_vector(), dst, 0.0)
    return sorted(op.tile_shape()["candidates_ms"])


CASES = {
    "advection_k4_pert": dict(dim=3, degree=4, refine=2, perturbation=0.1,
                              velocity_kind=1),
    "advection_k2_cart_dirichlet": dict(dim=3, degree=2, n_coarse=[3, 2, 1],
                                        refine=2, periodic=[0, 1, 0],
                                        velocity=[1.1, 0.15, -0.05],
                                        solution_kind=1),
    "advection_k6_fp32": dict(dim=3, degree=6, refine=1, perturbation=0.1,
                              velocity_kind=1, precision=1),
    "sipg_k3": dict(dim=3, degree=3, refine=2, velocity=[1.0, 0.5, 0.25],
                    operator_kind=1, viscosity=1e-2, penalty_factor=16,
                    solution_kind=2, solution_param=[2 * np.pi, 1e-2]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_every_tile_shape_gives_the_same_result(gpu_lib, name):
    op = AdvectionOperation(Cfg(**CASES[name]).dg())
    g = torch.Generator(device="cuda").manual_seed(5)
    src = torch.randn(op.n_dofs, device="cuda", dtype=op.dtype, generator=g)
    ref = op.initialize_dof_vector()
    op.apply(src, ref, 0.3)
    sol0 = torch.randn(op.n_dofs, device="cuda", dtype=op.dtype, generator=g)
    base = None
    cells = tile_cells(op)
    assert len(cells) == 3
    for c in cells:
        op.set_tile_shape(c)
        dst = op.initialize_dof_vector()
        op.apply(src, dst, 0.3)
        assert op.tile_shape()["cells_per_workgroup"] == c
        assert torch.equal(dst, ref), c
        if op.dtype != torch.float64:
            continue      # the fused LSRK stage is built in FP64
        sol = sol0.clone()
        k, nri = op.initialize_dof_vector(), op.initialize_dof_vector()
        op.perform_stage(0.3, 0.01, 0.02, src, k, sol, nri)
        if base is None:
            base = (sol.clone(), nri.clone())
        assert torch.equal(sol, base[0]) and torch.equal(nri, base[1])
    op.set_tile_shape(0)


def test_autotuner_keeps_the_fastest_candidate(gpu_lib):
    op = AdvectionOperation(Cfg(dim=3, degree=3, refine=4, perturbation=0.1,
                                velocity_kind=1).dg())
    src = op.initialize_dof_vector()
    op.project_initial(src, 0.0)
    dst = op.initialize_dof_vector()
    op.apply(src, dst, 0.0)
    t = op.tile_shape()
    assert t["tuned_ms"] > 0
    cand = t["candidates_ms"]
    timed = sorted((ms, c) for c, ms in cand.items() if ms > 0)
    assert len(timed) == 3
    assert t["cells_per_workgroup"] == timed[0][1]
    assert t["tuned_ms"] == timed[0][0]
    # the tuner wrote into scratch vectors: dst is the plain apply
    op.set_tile_shape(sorted(cand)[1])
    ref = op.initialize_dof_vector()
    op.apply(src, ref, 0.0)
    assert torch.equal(dst, ref)


def test_forcing_an_unbuilt_shape_is_an_error(gpu_lib):
    from advection_miniapp_amd import _lib as L
    op = AdvectionOperation(Cfg(dim=3, degree=4, refine=1).dg())
    with pytest.raises(L.DgError):
        op.set_tile_shape(7)

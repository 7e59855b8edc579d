"""The mixed-precision single solve replays one captured PCG iteration
(hipGraph) between its host checks: the same iterations and bitwise the same
head as launching the iteration's kernels one by one, for two solves in a row
on one workspace (each captures again and replays its graph many times, with
host synchronisations between the replays) and over the boundary mixes."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import bc_pair, grid_pair, random_K
from macroflow3d_amd import physics as ph

pytestmark = pytest.mark.gpu


def dev(a, ctx):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to(ctx.device)


@pytest.mark.parametrize("mix", ["darcy_x", "periodic_yz", "all_neumann"])
def test_graph_replay_equals_plain_launches(ctx, mix):
    g, _ = grid_pair(64, 48, 32)
    bc, _ = bc_pair(mix)
    n = g.num_cells()
    cfg = ph.HeadSolveConfig.mixed(1e-10, 4)
    heads = {}
    for graph in (0, 1, -1):  # -1: the default, replay up to 2^21 cells
        ws = ph.FlowWorkspace()
        ws.set_tuning(ph.FlowWorkspace.TUNE_GRAPH, graph)
        out = []
        for seed in (91, 92):  # two solves in a row on one workspace
            K = dev(random_K(n, seed, 1.5), ctx)
            h = ctx.zeros(n)
            res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
            assert res.converged and res.true_residual <= 1e-10 * res.bnorm, (mix, graph, seed)
            torch.cuda.synchronize()
            out.append((h.cpu().numpy(), res.num_iterations, res.restarts))
        heads[graph] = out
        ws.close()
    for other in (1, -1):
        for (h0, it0, rs0), (h1, it1, rs1) in zip(heads[0], heads[other]):
            assert it0 == it1 and rs0 == rs1
            assert np.array_equal(h0, h1)


@pytest.mark.parametrize("graph", [1, 0])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_one_operator_application_per_iteration(ctx, graph, precision):
    """The PCG solve enqueues one operator application per preconditioner
    application; besides them only the initial residual, the true-residual
    checks and the final true residual apply the operator (a doubled A p in
    the iteration would add one per iteration). Every iteration the solve
    reports was enqueued."""
    g, _ = grid_pair(64, 64, 64)
    bc, _ = bc_pair("darcy_x")
    n = g.num_cells()
    K = dev(random_K(n, 93, 1.5), ctx)
    cfg = (ph.HeadSolveConfig.mixed(1e-10, 4) if precision == "mixed"
           else ph.HeadSolveConfig.fp64_direct(1e-10, 4))
    ws = ph.FlowWorkspace()
    ws.set_tuning(ph.FlowWorkspace.TUNE_GRAPH, graph)
    for _ in range(2):  # the counts restart with each solve
        h = ctx.zeros(n)
        res = ph.solve_head(h, K, g, bc, cfg, ctx, ws)
        assert res.converged
        n_op, n_pc = ws.solve_counts()
        passes = res.restarts + 1
        print(f"{precision} graph={graph}: {res.num_iterations} iterations, "
              f"{n_op} operator, {n_pc} preconditioner applications", flush=True)
        assert n_pc - passes >= res.num_iterations
        assert 0 <= n_op - n_pc <= 2
    ws.close()

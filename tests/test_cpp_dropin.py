"""The C++ host surface (include/macroflow3d/*.hpp) used the way MacroFlow3D's
pipeline calls the head solve, checked against the CPU oracle.

CPU: every header compiles on its own. GPU: tests/cpp/dropin_solve.cpp (built
by macroflow3d_amd.build) runs solve_head with the reference and extension
semantics, both velocity layouts, and the generic pcg_solve over
VarCoeffLaplacian + MultigridPreconditioner, the velocity diagnostics, and the
extensions (mixed
precision, solve_head_batched, WarmStartNet); results are compared with the
oracle on the same K."""
from __future__ import annotations

import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from macroflow3d_amd import build as B
from macroflow3d_amd import warmstart as WS
from oracle import cnn_oracle as CO

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["core.hpp", "multigrid.hpp", "numerics.hpp", "physics.hpp"]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("hdr", HEADERS)
def test_header_compiles_alone(tmp_path, hdr):
    src = tmp_path / "t.cpp"
    src.write_text(f'#include "macroflow3d/{hdr}"\nint main() {{ return 0; }}\n')
    r = subprocess.run([B.hipcc(), "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-Wno-unused-command-line-argument",
                        "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_cpp_dropin_matches_oracle(tmp_path):
    assert os.path.exists(B.DROPIN_BIN), "build the package first (macroflow3d_amd.build)"
    n = 32
    prefix = str(tmp_path / "run")
    print(f"dropin_solve n={n}", flush=True)
    r = subprocess.run([B.DROPIN_BIN, str(n), prefix, WS.DEFAULT_WEIGHTS], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    rd = lambda s: np.fromfile(f"{prefix}.{s}.bin", dtype=np.float64)  # noqa: E731
    K = rd("K")
    g, bc = O.grid(n, n, n), O.darcy_bc()
    # reference semantics
    ho, ro = O.solve_head(g, bc, O.mode_r(mg_levels=3, cg_rtol=1e-10), K)
    assert info["ref_converged"] and info["ref_iters"] == ro.num_iterations
    assert _rel(rd("h_ref"), ho) < 1e-10
    assert abs(info["ref_r0"] - ro.initial_residual) <= 1e-12 * ro.initial_residual
    # extension semantics
    hdeep, _ = O.solve_head(g, bc, O.mode_n(mg_levels=3, cg_rtol=1e-13), K)
    h_ns = rd("h_ns")
    assert info["ns_converged"] and info["ns_true_residual"] <= 1.5e-10 * info["ns_bnorm"]
    assert _rel(h_ns, hdeep) < 1e-8
    # generic pcg_solve: same algorithm as the reference-semantics solve
    assert info["pcg_converged"] and info["pcg_iters"] == info["ref_iters"]
    assert _rel(rd("h_pcg"), ho) < 1e-10
    # velocities of h_ns
    U, V, W = O.velocity_compact(g, bc, h_ns, K)
    for name, ref in (("U", U), ("V", V), ("W", W)):
        got = rd(name)
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), name
    Up = rd("Up").reshape(n + 1, n + 1, n + 1)
    assert np.array_equal(Up[:n, :n, :], rd("U").reshape(n, n, n + 1))
    assert np.abs(info["U_sum"] - U.sum()) <= 1e-10 * abs(U.sum()) and info["U_finite"]
    # velocity diagnostics of the padded field vs the oracle's on its own padded
    # velocity of the same head; the divergence is the solve's residual (tiny)
    Uo, Vo, Wo = O.velocity_padded(g, bc, h_ns, K)
    dref, oref, href = O.velocity_diagnostics(g, Uo, Vo, Wo)
    scale = np.abs(U).max()
    assert np.abs(rd("div") - dref).max() <= 1e-10 * scale
    assert np.abs(rd("omag") - oref).max() <= 1e-10 * scale
    assert np.abs(rd("hel") - href).max() <= 1e-10 * scale * scale
    d = rd("div")
    assert info["div_min"] == d.min() and info["div_max"] == d.max()
    assert abs(info["div_mean"] - d.mean()) <= 1e-12 * scale
    assert info["levels"] == 3
    # extensions: mixed precision reaches the FP64 residual bar, the batched
    # solve reproduces the single solves bit for bit, the C++ network loader
    # and prediction match the oracle network, and PCG converges from it
    assert info["mixed_converged"] and info["batched_equal"]
    p = np.load(WS.DEFAULT_WEIGHTS).astype(np.float64)
    lnK = np.log(K).reshape(n, n, n)
    y_o, _ = CO.forward(p, lnK, rb=True)
    ref = CO.output(y_o, 1.0, 0.0, (n, n, n))
    assert np.abs(rd("h0").reshape(n, n, n) - ref).max() < 2e-4 / n
    assert info["warm_converged"] and info["warm_iters"] > 0

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


@pytest.mark.gpu
def test_pipeline_sequence_matches_oracle(tmp_path):
    """tests/cpp/ensemble_dropin.cpp: the realisation loop's calls
    (EnsembleRunner.cu:146-309) through the C++ surface, from_yaml("mg_cg")
    included: each realisation (seed base + r) reaches the oracle's iteration
    count and head in reference semantics, and its K statistics and velocity
    divergence are reported."""
    assert os.path.exists(B.ENSEMBLE_BIN), "build the package first (macroflow3d_amd.build)"
    n, nr, base = 32, 3, 700
    prefix = str(tmp_path / "ens")
    print(f"ensemble_dropin n={n} realisations={nr}", flush=True)
    r = subprocess.run([B.ENSEMBLE_BIN, str(n), str(nr), str(base), prefix], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(s) for s in r.stdout.splitlines() if s.startswith("{")]
    assert [x["r"] for x in lines] == list(range(nr))
    g, bc = O.grid(n, n, n), O.darcy_bc()
    seen = set()
    for x in lines:
        assert x["seed"] == base + x["r"]
        assert x["solver_type"] == 2 and x["mg_levels"] == 3 and x["cg_check_every"] == 10
        rd = lambda s: np.fromfile(f"{prefix}.r{x['r']}.{s}.bin", dtype=np.float64)  # noqa: E731
        K, h = rd("K"), rd("h")
        seen.add(float(K.sum()))
        ho, ro = O.solve_head(g, bc, O.mode_r(mg_levels=3, cg_rtol=1e-10), K)
        assert x["converged"] and x["iters"] == ro.num_iterations
        assert _rel(h, ho) < 1e-10
        assert abs(x["r0"] - ro.initial_residual) <= 1e-12 * ro.initial_residual
        assert x["kmin"] == K.min() and x["kmax"] == K.max()
        assert abs(x["kmean"] - K.mean()) <= 1e-12 * K.mean()
        # the padded velocity's divergence is the solve's residual: tiny
        assert max(abs(x["div_min"]), abs(x["div_max"])) < 1e-6
    assert len(seen) == nr  # distinct fields per realisation


def test_from_yaml_solver_strings():
    """HeadSolveConfig::from_yaml's solver mapping (solve_head.cuh:98-130) and
    the FlowYamlConfig defaults (Config.hpp:86-114), compiled and run on the
    host (no GPU call)."""
    import tempfile
    src = r'''
#include "macroflow3d/physics.hpp"
#include <cstdio>
using namespace macroflow3d;
int main() {
    io::FlowYamlConfig f;
    int bad = 0;
    bad += physics::HeadSolveConfig::from_yaml(f).solver_type != physics::HeadSolverType::MG;
    const char* s[5] = {"cg", "mg", "mg_cg", "pcg_mg", "other"};
    int want[5] = {0, 1, 2, 2, 2};
    for (int i = 0; i < 5; ++i) {
        f.solver = s[i];
        bad += (int)physics::HeadSolveConfig::from_yaml(f).solver_type != want[i];
    }
    f.mg_levels = 6; f.mg_pre_smooth = 4; f.mg_post_smooth = 3; f.mg_max_cycles = 7;
    f.mg_coarse_iters = 9; f.cg_max_iter = 11; f.cg_rtol = 1e-9; f.cg_check_every = 5;
    f.rtol = 1e-7; f.pin.mode = PinMode::On;
    const physics::HeadSolveConfig c = physics::HeadSolveConfig::from_yaml(f);
    bad += c.mg_levels != 6 || c.mg_pre_smooth != 4 || c.mg_post_smooth != 3 ||
           c.mg_max_cycles != 7 || c.mg_coarse_iters != 9 || c.cg_max_iter != 11 ||
           c.cg_rtol != 1e-9 || c.cg_check_every != 5 || c.rtol != 1e-7 ||
           c.pin.mode != PinMode::On;
    // extensions keep the reference behaviour
    bad += c.init_guess != physics::InitGuess::Reference ||
           c.coarse_solver != physics::CoarseSolver::GSRB ||
           c.precond_precision != physics::PrecondPrecision::FP64;
    io::FlowYamlConfig d;
    bad += d.mg_levels != 4 || d.mg_pre_smooth != 2 || d.mg_post_smooth != 2 ||
           d.mg_coarse_iters != 50 || d.mg_max_cycles != 20 || d.cg_max_iter != 1000 ||
           d.cg_rtol != 1e-8 || d.cg_check_every != 10 || d.rtol != 1e-6 || d.verify_velocity;
    std::printf("%d\n", bad);
    return bad;
}
'''
    with tempfile.TemporaryDirectory() as d:
        cpp = os.path.join(d, "t.cpp")
        exe = os.path.join(d, "t")
        with open(cpp, "w") as f:
            f.write(src)
        lib = os.path.dirname(B.LIB)
        r = subprocess.run([B.hipcc(), "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(ROOT, "include"),
                            cpp, "-L" + lib, "-lmf3d_hip", "-Wl,-rpath," + lib, "-o", exe],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout + r.stderr

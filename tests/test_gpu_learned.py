"""The trained warm-start network (gpufem_amd/weights/) on held-out members
of the Brick8 cantilever family.  Measured: the network's prediction is a
better initial guess than the method without it (its untrained output, the
beam-theory estimate, and zero start) by the L2 error, but its residual
||b - K x0|| is larger than ||b|| (the small-scale roughness of a
regression network costs more residual than its large-scale accuracy
saves), so with a stopping test relative to ||b|| it cannot save Jacobi-PCG
iterations.  The solver therefore takes the guess guarded (use_x0 = 2): it
is kept only when its residual is below ||b||, else the solve is the
zero-start one, bitwise.  The test asserts what the network buys (the
closer guess), that a warm-started solve never takes more iterations than
zero start and meets the same true residual, and that the GPU inference
equals the CPU network oracle on the trained weights."""
import os

import numpy as np
import pytest

import gpufem_amd as gf
from gpufem_amd import forms
from gpufem_amd import learned as L
from oracle import nn_ref as R

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "gpufem_amd", "weights", "brick8_cantilever_64x32x32.npz")


@pytest.fixture(scope="module")
def trained(gpu_ctx):
    assert os.path.exists(WEIGHTS), "trained weights missing: " + WEIGHTS
    net, c_u, meta = L.load_weights(gpu_ctx, WEIGHTS, batch_max=1)
    return net, c_u, meta, L.Family()


def test_warm_start_meets_residual_and_never_costs_iterations(gpu_ctx, trained):
    net, c_u, meta, fam = trained
    ks = [L.HOLDOUT_BASE + k for k in range(6)]
    recs = L.evaluate(gpu_ctx, fam, net, c_u, ks)
    for r in recs:
        assert r["zero_true_relres"] <= 1e-8 and r["warm_true_relres"] <= 1e-8, r
        assert r["x0_rel_err"] < 0.25, r
        # a guarded guess never makes the solve longer; a rejected one gives
        # the zero-start solve exactly
        assert r["warm_it"] <= r["zero_it"], r
        if r["warm_x0_rejected"]:
            assert r["warm_it"] == r["zero_it"] and r["solution_rel_diff"] == 0.0, r
        else:
            assert r["solution_rel_diff"] <= 1e-8, r
    s = L.summarize(recs)
    # what the trained network buys on members it never saw: a closer
    # initial guess than its own untrained output (the beam estimate)
    assert s["x0_rel_err_mean"] < 0.8 * s["base_rel_err_mean"] < 0.8, s
    assert s["iteration_ratio"] <= 1.0, s


def test_inference_matches_network_oracle(gpu_ctx, trained):
    net, c_u, meta, fam = trained
    p = fam.problem(L.HOLDOUT_BASE + 7)
    feats, s_f = fam.features(p)
    d_in = gf.DevBuf.from_array(gpu_ctx, feats)
    d_out = gf.DevBuf(gpu_ctx, fam.nodes * 16 * 4)
    net.forward(d_in.ptr, 1, d_out.ptr)
    got = d_out.download(np.float32).reshape(fam.nz + 1, fam.ny + 1, fam.nx + 1, 16)
    x = R.from_bf16_bits(feats).reshape(1, fam.nz + 1, fam.ny + 1, fam.nx + 1, 16)
    ref = R.Net(fam.nx + 1, fam.ny + 1, fam.nz + 1, net.params(), bf16_mode=True).forward(x)[0]
    err = np.linalg.norm(got[..., :3] - ref[..., :3]) / np.linalg.norm(ref[..., :3])
    assert err <= 2e-2, err


def test_combined_forms_quoted_at_the_accuracy_bar(gpu_ctx):
    """The combined forms (warm start + grouped batch, warm start + mixed
    precision) are quoted at the loosest tolerance where every member's
    displacements are within 1e-9 of its FP64 solution converged to 1e-12."""
    r = forms.form_combined(gpu_ctx, WEIGHTS, members=4, group=2)
    for key in ("grouped_learned", "mixed_learned"):
        op = r[key]
        assert op["meets_bar"] and op["displacement_rel_err"] <= forms.ERR_BAR, (key, op)
        assert op["max_true_relres"] <= 1.05 * op["rtol"], (key, op)
        assert op["rtol"] in forms.OP_RTOLS and op["runs"][-1]["rtol"] == op["rtol"]

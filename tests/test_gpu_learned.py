"""The trained warm-start network (gpufem_amd/weights/) on held-out members
of the Brick8 cantilever family: every warm-started solve meets the same true
residual as zero-start PCG; what the network gains is a better initial
guess than the method without it (its untrained output, the beam-theory
estimate, and zero start); the iterations it saves are bounded by the
residual test (a few percent at this accuracy), so the test bounds the
iteration ratio instead of asking for a large saving; and the GPU inference
equals the CPU network oracle on the trained weights."""
import os

import numpy as np
import pytest

import gpufem_amd as gf
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


def test_warm_start_meets_residual_and_saves_iterations(gpu_ctx, trained):
    net, c_u, meta, fam = trained
    ks = [L.HOLDOUT_BASE + k for k in range(6)]
    recs = L.evaluate(gpu_ctx, fam, net, c_u, ks)
    for r in recs:
        assert r["zero_true_relres"] <= 1e-8 and r["warm_true_relres"] <= 1e-8, r
        assert r["x0_rel_err"] < 0.25, r
        assert r["solution_rel_diff"] <= 1e-8, r
    s = L.summarize(recs)
    # what the trained network gains on members it never saw: a closer
    # initial guess than its own untrained output (the beam estimate)
    assert s["x0_rel_err_mean"] < 0.8 * s["base_rel_err_mean"] < 0.8, s
    # and the warm start costs no more than a few percent of iterations
    assert s["iteration_ratio"] <= 1.05, s


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

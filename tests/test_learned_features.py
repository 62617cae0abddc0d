"""Host-side features of the learned warm start (gpufem_amd/learned.py):
the Timoshenko beam estimate against the closed-form cantilever deflection
under a uniform top traction, the residual target / prediction round trip,
and the bf16 rounding used for the network inputs."""
import numpy as np
import pytest

from gpufem_amd import learned as L
from gpufem_amd import problems as P


@pytest.fixture(scope="module")
def fam():
    return L.Family(32, 8, 8)


def _uniform_traction(fam, tz):
    nx, ny, nz = fam.nx, fam.ny, fam.nz
    area = np.ones((ny + 1, nx + 1))
    area[0, :] *= 0.5
    area[-1, :] *= 0.5
    area[:, 0] *= 0.5
    area[:, -1] *= 0.5
    F = np.zeros((nz + 1, ny + 1, nx + 1, 3))
    F[nz, :, :, 2] = tz * area
    return F.reshape(-1)


def test_beam_estimate_matches_timoshenko_closed_form(fam):
    E = P.E_DEFAULT
    f = _uniform_traction(fam, -1.0)
    w = fam.beam_estimate(np.full(fam.nx * fam.ny * fam.nz, E), f)
    Lb, b, h = fam.nx, fam.ny, fam.nz
    q = -1.0 * b                                  # force per unit length
    I, A = b * h ** 3 / 12.0, b * h
    G = E / (2.0 * (1.0 + P.NU_DEFAULT))
    x = np.arange(Lb + 1, dtype=float)
    bend = q * x ** 2 * (6 * Lb ** 2 - 4 * Lb * x + x ** 2) / (24 * E * I)
    shear = q * (Lb * x - x ** 2 / 2) / (5.0 / 6.0 * G * A)
    wz = w[0, 0, :, 2]
    assert np.allclose(wz, bend + shear, rtol=0, atol=2e-3 * abs(bend[-1] + shear[-1]))
    # the whole cross-section moves with the centre line; u_y stays 0
    assert np.allclose(w[..., 2], wz[None, None, :])
    assert np.all(w[..., 1] == 0.0)
    # rotated plane sections: u_x = -(z - h/2) theta, theta(L) = q L^3 / (6 E I)
    theta_tip = q * Lb ** 3 / (6 * E * I)
    ux_top = w[h, 0, -1, 0]
    assert ux_top == pytest.approx(-(h / 2.0) * theta_tip, rel=2e-2)
    assert w[0, 0, -1, 0] == pytest.approx(-ux_top, rel=1e-12)


def test_beam_estimate_scales_with_modulus_and_load(fam):
    f = _uniform_traction(fam, -1.0)
    E1 = np.full(fam.nx * fam.ny * fam.nz, P.E_DEFAULT)
    w1 = fam.beam_estimate(E1, f)
    assert np.allclose(fam.beam_estimate(2 * E1, f), w1 / 2, rtol=1e-12, atol=0)
    assert np.allclose(fam.beam_estimate(E1, 3 * f), 3 * w1, rtol=1e-12, atol=0)
    # a stiffer half near the clamp lowers the tip deflection
    E2 = E1.reshape(fam.nz, fam.ny, fam.nx).copy()
    E2[:, :, : fam.nx // 2] *= 4
    assert abs(fam.beam_estimate(E2.reshape(-1), f)[0, 0, -1, 2]) < abs(w1[0, 0, -1, 2])


def test_target_and_prediction_round_trip(fam):
    p = fam.problem(7)
    feats, s_f = fam.features(p)
    rng = np.random.default_rng(0)
    u = rng.standard_normal(p.ndof) * 1e-3
    u.reshape(-1, 3)[fam.base_flat[:, L.MASK_CHANNEL] > 0.5] = 0.0
    c_u = 4e-4
    t = fam.target(u, feats, s_f, c_u).reshape(-1, 4)
    assert np.all(t[:, 3] == 0)
    x = fam.x0(t, feats, s_f, c_u)
    assert np.linalg.norm(x - u) <= 1e-6 * np.linalg.norm(u)
    # an untrained (zero) prediction is the beam estimate as the features carry it
    x_beam = fam.x0(np.zeros_like(t), feats, s_f, c_u)
    ref = fam.beam_estimate(p.elem_E, p.f).reshape(-1, 3)
    ref[fam.base_flat[:, L.MASK_CHANNEL] > 0.5] = 0.0
    assert np.abs(x_beam - ref.reshape(-1)).max() <= 2 ** -8 * np.abs(ref).max()


def test_bf16_rounding_nearest_even():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0], dtype=np.float32)
    bits = L.bf16_bits(a)
    back = (bits.astype(np.uint32) << 16).view(np.float32)
    assert list(back) == [1.0, 1.0, 1.0 + 2 ** -7, -2.5, 0.0]

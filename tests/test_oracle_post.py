"""Pins of the wall post-process oracle (oracle/orc_post.py) on the CPU.

* Closed forms: on a uniform Cartesian grid (spacings dx, dy) rotated by an
  angle theta in the x-y plane, a wall-parallel velocity that grows linearly
  from the wall, U(j) = a_i j dy, and a linear temperature T(j) = T_w + b_i j dy
  give cf = mu_i a_i / q_dyn and qw = lambda_i b_i for every theta, dx, dy
  (the metrics are xi = (cos, sin)/dx, eta = (-sin, cos)/dy).
* A velocity normal to the wall gives cf = 0 (the projection is on xi).
* The file text follows std::ostream's default notation (%g)."""
import numpy as np
import pytest

import orc_post as P


def block(theta, dx, dy, mx=7, my=5, mz=4, g=2, normal=False):
    nx, ny, nz = mx + 2 * g, my + 2 * g, mz + 2 * g
    c, s = np.cos(theta), np.sin(theta)
    metric = np.zeros((9, nz, ny, nx))
    metric[0], metric[1], metric[2] = c / dx, s / dx, 0.0
    metric[3], metric[4], metric[5] = -s / dy, c / dy, 0.0
    metric[8] = 1.0
    i = np.arange(nx) - g
    j = (np.arange(ny) - g)[:, None]
    a = 3.0 + 0.5 * i            # velocity gradient per i
    b = -40.0 + 7.0 * i          # temperature gradient per i
    mu = 1.8e-5 * (1 + 0.1 * i)
    lam = 2.5e-2 * (1 + 0.05 * i)
    U = a[None, :] * j * dy
    direction = (-s, c) if normal else (c, s)
    bv = np.zeros((6, nz, ny, nx))
    bv[0] = 1.2
    bv[1] = U * direction[0]
    bv[2] = U * direction[1]
    bv[4] = 101325.0
    bv[5] = 300.0 + b[None, :] * j * dy
    mul = np.broadcast_to(mu, (nz, ny, nx)).copy()
    lmb = np.broadcast_to(lam, (nz, ny, nx)).copy()
    inner = slice(g, nx - g)
    return metric, bv, mul, lmb, g, a[inner], b[inner], mu[inner], lam[inner]


@pytest.mark.parametrize("theta,dx,dy", [(0.0, 1e-3, 1e-5), (0.7, 2e-3, 3e-6), (-2.4, 5e-4, 1e-4),
                                         (np.pi / 2, 1e-3, 1e-5)])
def test_closed_form_on_rotated_uniform_grids(theta, dx, dy):
    metric, bv, mul, lam, g, a, b, mu, lm = block(theta, dx, dy)
    q = 0.5 * 0.3 * 600.0 ** 2
    cf, qw = P.wall_friction_heatflux(metric, bv, mul, lam, g, q)
    assert np.allclose(cf, mu * a / q, rtol=1e-12, atol=0)
    # T(1) - T(0) cancels the wall temperature: its rounding scales with T/dy
    eps = np.finfo(np.float64).eps
    assert np.all(np.abs(qw - lm * b) <= 8 * eps * (300.0 + np.abs(b) * dy) * lm / dy + 1e-12 * np.abs(lm * b))


def test_wall_normal_velocity_has_no_friction():
    metric, bv, mul, lam, g, *_ = block(0.9, 1e-3, 1e-5, normal=True)
    cf, _ = P.wall_friction_heatflux(metric, bv, mul, lam, g, 1000.0)
    assert np.all(np.abs(cf) < 1e-12 * np.abs(bv[1]).max())


def test_file_text_uses_the_default_stream_notation():
    text = P.friction_heatflux_file([0.1, 1234567.0], [0.0025, -1e-7], [12.5, 0.0])
    assert text == 'variables = "x", "cf", "qw"\n0.1\t0.0025\t12.5\n1.23457e+06\t-1e-07\t0\n'

"""numpy restatement of the reference's 2-D wall post-process.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

wall_friction_heatflux follows wall_friction_heatFlux_2d
(src/PostProcess.cu:49-70) and the host loop around it
(src/PostProcess.cu:6-47): for every i of the k = 0 line of a block whose
j = 0 plane is the wall,

    u_par    = (xi_x u + xi_y v)(i,1,0) / sqrt(xi_x^2 + xi_y^2)   metrics at (i,0,0)
    grad_eta = sqrt(eta_x^2 + eta_y^2)                             at (i,0,0)
    cf       = mu_l(i,0,0) u_par grad_eta / (0.5 rho_inf v_inf^2)
    qw       = lambda(i,0,0) (T(i,1,0) - T(i,0,0)) grad_eta

Fields are padded by g ghost layers on every side, shaped (..., nz, ny, nx),
metric planes in the order (r, c) = (1,1), (1,2), ..., (3,3).
Parity pins (tests/test_oracle_post.py): closed forms on a uniform Cartesian
grid (linear velocity and temperature profiles) and invariance under a
rotation of the grid and the flow in the x-y plane.
"""
from __future__ import annotations

import numpy as np


def wall_friction_heatflux(metric, bv, mul, lam, g, dyn_pressure):
    """cf, qw (length mx) from the padded fields of one block."""
    w = (g, g, slice(g, metric.shape[-1] - g))          # (k=0, j=0, i)
    a = (g, g + 1, slice(g, metric.shape[-1] - g))      # (k=0, j=1, i)
    xi_x, xi_y = metric[0][w], metric[1][w]
    eta_x, eta_y = metric[3][w], metric[4][w]
    u_par = (xi_x * bv[1][a] + xi_y * bv[2][a]) / np.sqrt(xi_x * xi_x + xi_y * xi_y)
    grad_eta = np.sqrt(eta_x * eta_x + eta_y * eta_y)
    cf = mul[w] * (u_par * grad_eta) / dyn_pressure
    qw = lam[w] * (bv[5][a] - bv[5][w]) * grad_eta
    return cf, qw


def friction_heatflux_file(x_line, cf, qw) -> str:
    """Text of output/wall/friction_heatflux.dat for one block
    (src/PostProcess.cu:38-44): a header line and "x<TAB>cf<TAB>qw" per i in
    std::ostream's default notation (6 significant digits, %g)."""
    lines = ['variables = "x", "cf", "qw"']
    lines += ["%g\t%g\t%g" % (float(a), float(b), float(c)) for a, b, c in zip(x_line, cf, qw)]
    return "\n".join(lines) + "\n"

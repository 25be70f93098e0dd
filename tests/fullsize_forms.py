"""Closed forms used by the full-size GPU checks (tests/test_gpu_full_size.py),
kept apart so the CPU suite can check them against the oracle on small grids."""
import numpy as np


def dirichlet_sine_modes(n, h, modes):
    """Separable sines psi_m(i, j, k) = prod_a sin(pi m_a (i_a + 1) / (n_a + 1)) on an
    n[0] x n[1] x n[2] grid (x fastest), and their eigenvalues under the stencil
    -1/2 Lap_h with zero ghost values: sum_a (1 - cos(pi m_a / (n_a + 1))) / h_a^2."""
    psi = np.empty((len(modes), n[0] * n[1] * n[2]))
    lam = np.empty(len(modes))
    for o, m in enumerate(modes):
        s = [np.sin(np.pi * m[a] * (np.arange(n[a]) + 1) / (n[a] + 1)) for a in range(3)]
        psi[o] = np.multiply.outer(np.multiply.outer(s[2], s[1]), s[0]).ravel()
        lam[o] = sum((1.0 - np.cos(np.pi * m[a] / (n[a] + 1))) / h[a] ** 2 for a in range(3))
    return psi, lam


def boundary_sine_mode(n, h):
    """rho = prod_a sin(pi i_a / (n_a - 1)) (zero on the boundary planes) and the
    Poisson solution -4 pi rho / mu of the interior Jacobi fixed point,
    mu = -sum_a 4 sin^2(pi / (2 (n_a - 1))) / h_a^2."""
    s = [np.sin(np.pi * np.arange(n[a]) / (n[a] - 1)) for a in range(3)]
    for a in range(3):
        s[a][0] = s[a][-1] = 0.0
    rho = np.multiply.outer(np.multiply.outer(s[2], s[1]), s[0]).ravel()
    mu = -sum(4.0 * np.sin(np.pi / (2.0 * (n[a] - 1))) ** 2 / h[a] ** 2 for a in range(3))
    return rho, -4.0 * np.pi * rho / mu

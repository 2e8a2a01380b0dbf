"""Independent numpy statement of the lattice action, used to pin the drift.

Written from the physics, not from the oracle's code: the kinetic operator via
numpy.fft (pocketfft), the interaction as tr[(Psibar^T Psi)^2] with explicit
2x2 matrices, the spin term with explicit Pauli matrices.  The complex-Langevin
drift of psi is -dS/dpsibar, that of psibar is -dS/dpsi (reference drift
Propagator_without_cutoff.cu:60-128 with a_x = factor = 1; its psi(t+1) variant is
compared separately).
"""
from __future__ import annotations

import numpy as np

PAULI = [np.array([[0, 1], [1, 0]], complex),
         np.array([[0, -1j], [1j, 0]], complex),
         np.array([[1, 0], [0, -1]], complex)]


def momenta(n: int) -> np.ndarray:
    """2 pi j / n on numpy's frequency grid (equals the reference's p, :169-194)."""
    return 2 * np.pi * np.fft.fftfreq(n)


def kinetic(f: np.ndarray, a_x: float = 1.0, kin_dims: int = 1) -> np.ndarray:
    """K = (1/a_x) F^-1[-p^2 F[f]] over the spatial axes of f (4, Nt, X, Y, Z)."""
    _, _, X, Y, Z = f.shape
    px, py, pz = momenta(X), momenta(Y), momenta(Z)
    p2 = px[:, None, None] ** 2 + 0 * py[None, :, None] + 0 * pz[None, None, :]
    if kin_dims == 3:
        p2 = px[:, None, None] ** 2 + py[None, :, None] ** 2 + pz[None, None, :] ** 2
    fk = np.fft.fftn(f, axes=(2, 3, 4))
    return np.fft.ifftn(-p2 * fk, axes=(2, 3, 4)) / a_x


def as_matrix(f: np.ndarray) -> np.ndarray:
    """(4, ...) components -> (2, 2, ...) matrix [[psi0, psi1], [psi2, psi3]]."""
    return f.reshape((2, 2) + f.shape[1:])


def h_density(b: np.ndarray, q: np.ndarray, p, include_mu: bool = True) -> np.ndarray:
    """Hamiltonian density on the time-split pair (psibar = b, psi = q).

    h = sum_i [-b_i K_i(q) - s_i b_i q_i] + G tr[(B^T Q)^2] + c1/2 sum_a F_a^2,
    F_a = tr(B^T sigma_a Q), s = (mu+P+Q, mu, mu, mu-P+Q) (mu dropped when
    include_mu is False: the energy estimator).
    """
    mu = p.mu if include_mu else 0.0
    s = np.array([mu + p.P + p.Q, mu, mu, mu - p.P + p.Q])
    K = kinetic(q, p.a_x, p.kin_dims)
    h = np.sum(-b * K - s[:, None, None, None, None] * b * q, axis=0)
    B, Q = as_matrix(b), as_matrix(q)
    A = np.einsum("rc...,rd...->cd...", B, Q)          # B^T Q
    AA = np.einsum("cd...,de...->ce...", A, A)
    h = h + p.G * (AA[0, 0] + AA[1, 1])
    for sig in PAULI:
        Fa = np.einsum("rc...,rs,sc...->...", B, sig, Q)
        h = h + 0.5 * p.c1 * Fa * Fa
    return h


def action(psi: np.ndarray, psib: np.ndarray, p) -> complex:
    """S = sum_t sum_x [psibar(t) (psi(t) - psi(t-1)) + eps h(psibar(t), psi(t-1))]."""
    pm = np.roll(psi, 1, axis=1)
    S = np.sum(psib * (psi - pm))
    S = S + p.eps * np.sum(h_density(psib, pm, p))
    return S


def d_action(psi, psib, p, which: str, idx, h: float = 0.05) -> complex:
    """dS/d(field[idx]) by the five-point stencil -- exact for the quartic S."""
    def S_at(delta):
        a, b = psi.copy(), psib.copy()
        if which == "psi":
            a[idx] += delta
        else:
            b[idx] += delta
        return action(a, b, p)
    return (-S_at(2 * h) + 8 * S_at(h) - 8 * S_at(-h) + S_at(-2 * h)) / (12 * h)

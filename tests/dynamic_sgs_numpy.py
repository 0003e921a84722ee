"""An independent numpy formulation of the dynamic mixed-scale subgrid model,
used only to check the oracle's restatement of compute_dynamic_sgs
(src/uniform/solver.rs:562-828).

Written from the model's definition, vectorised, with different mechanics
from the oracle's node loops: the clamped box filter is a zero-padded sliding
window divided by the padded count of in-domain nodes, derivatives are array
slices against edge-replicated copies, pi_neq is a moment contraction over the
velocity set, and the equilibrium uses the 3 / 4.5 / 1.5 coefficients.
Agreement with the oracle is therefore to round-off, not bitwise.

The model (Germano identity with Lilly's least squares, mixed-scale forms):
  grid level  : S~ = -pi_neq / (2 cs2 tau rho),  |S~| = sqrt(S~:S~)
  test level  : S- = sym grad of the test-filtered velocity, |S-| likewise
  similarity  : q^1/2 = |u - u-| / (2 sqrt 2),  Phi^1/2 = |T - T-| |u - u-| / (2 sqrt 2)
  tensors     : |T| = sqrt(2 |grad T . S|^2) at both levels
  mixed forms : nu_mix = (D^2 |S|)^a (D q^1/2)^(1-a),
                kap_mix = (D^3/dT |T|)^a (D/dT Phi^1/2)^(1-a), D = 1 or eta
  Germano     : M_ij = nu_mix- S-_ij - nu_mix~ S~_ij,  M_j = kap_mix- dT-/dx_j - kap_mix~ dT~/dx_j
  Leonard     : L_ij = (u_i u_j)- - u-_i u-_j (deviatoric, 3-D trace), L_j = (u_j T)- - u-_j T-
  Lilly       : C = clamp(box(L:M) / (2 box(M:M)), 0, 0.04), nu_t = C nu_mix~, floored.
"""
import numpy as np
from numpy.lib.stride_tricks import sliding_window_view

from tests.textbook_d2q9 import CX, CY, equilibrium

CS2 = 1.0 / 3.0
C_MAX = 0.04
TAU_FLOOR = 0.505
TRACE_DIM = 3.0  # the deviatoric part uses the 3-D trace L_kk / 3 (2-D fields)


def box(a, width):
    """Mean over the window [i - h, i - h + width) of each axis, h = (width-1)//2,
    restricted to the domain."""
    h = (width - 1) // 2
    pad = ((h, width - 1 - h), (h, width - 1 - h))
    s = sliding_window_view(np.pad(a, pad), (width, width)).sum(axis=(2, 3))
    n = sliding_window_view(np.pad(np.ones_like(a), pad), (width, width)).sum(axis=(2, 3))
    return s / n


def d_dx(a):
    e = np.concatenate([a[:, :1], a, a[:, -1:]], axis=1)
    return (e[:, 2:] - e[:, :-2]) / 2.0


def d_dy(a):
    e = np.concatenate([a[:1, :], a, a[-1:, :]], axis=0)
    return (e[2:, :] - e[:-2, :]) / 2.0


def lilly(num, den):
    with np.errstate(divide="ignore", invalid="ignore"):
        c = num / (2.0 * den)
    ok = np.isfinite(num) & np.isfinite(den) & (den > 0.0) & np.isfinite(c)
    return np.where(ok, np.clip(c, 0.0, C_MAX), 0.0)


def dynamic_sgs(f, rho, ux, uy, temp, *, tau, nu, alpha_mol, smag_alpha=0.5, nu_sm_min=-0.1,
                eta_test=2.0):
    """f: [9, ny, nx] post-stream populations; fields [ny, nx].  Returns
    (nu_t, alpha_t, C_nu, C_kappa) per node."""
    eta = max(eta_test, 1.0)
    w = max(int(np.floor(eta + 0.5)), 1)
    a = smag_alpha
    fneq = f - equilibrium(rho, ux, uy)
    scale = -1.0 / (2.0 * CS2 * tau * np.maximum(rho, 1e-12))
    sh = np.stack([(CX * CX * fneq).sum(0), (CX * CY * fneq).sum(0), (CY * CY * fneq).sum(0)]) * scale
    smag_h = np.sqrt(sh[0] ** 2 + 2.0 * sh[1] ** 2 + sh[2] ** 2)

    ub, vb, tb = box(ux, w), box(uy, w), box(temp, w)
    sb = np.stack([d_dx(ub), 0.5 * (d_dx(vb) + d_dy(ub)), d_dy(vb)])
    smag_b = np.sqrt(sb[0] ** 2 + 2.0 * sb[1] ** 2 + sb[2] ** 2)
    gt = np.stack([d_dx(temp), d_dy(temp)])
    gb = np.stack([d_dx(tb), d_dy(tb)])

    du = np.hypot(ux - ub, uy - vb)
    q_h = du / (2.0 * np.sqrt(2.0))
    phi_h = np.abs(temp - tb) * du / (2.0 * np.sqrt(2.0))
    q_b, phi_b = box(q_h, w), box(phi_h, w)

    def tmag(g, s):
        tx = g[0] * s[0] + g[1] * s[1]
        ty = g[0] * s[1] + g[1] * s[2]
        return np.sqrt(2.0 * (tx ** 2 + ty ** 2))

    dT = 1.0
    nu_g = smag_h ** a * q_h ** (1.0 - a)
    nu_b = (eta * eta * smag_b) ** a * (eta * q_b) ** (1.0 - a)
    kap_g = (tmag(gt, sh) / dT) ** a * (phi_h / dT) ** (1.0 - a)
    kap_b = (eta ** 3 / dT * tmag(gb, sb)) ** a * (eta / dT * phi_b) ** (1.0 - a)

    m = nu_b * sb - nu_g * sh                 # (xx, xy, yy)
    mj = kap_b * gb - kap_g * gt              # (x, y)
    lt = np.stack([box(ux * ux, w) - ub * ub, box(ux * uy, w) - ub * vb, box(uy * uy, w) - vb * vb])
    tr = (lt[0] + lt[2]) / TRACE_DIM
    ld = np.stack([lt[0] - tr, lt[1], lt[2] - tr])
    lj = np.stack([box(ux * temp, w) - ub * tb, box(uy * temp, w) - vb * tb])

    def contract(x, y):
        return x[0] * y[0] + 2.0 * x[1] * y[1] + x[2] * y[2]

    c_nu = lilly(box(contract(ld, m), w), box(contract(m, m), w))
    c_k = lilly(box((lj * mj).sum(0), w), box((mj * mj).sum(0), w))
    margin = (TAU_FLOOR - 0.5) / 3.0
    nu_t = np.maximum(c_nu * nu_g, max(nu_sm_min * nu, margin - nu))
    alpha_t = np.maximum(c_k * kap_g, max(nu_sm_min * alpha_mol, margin - alpha_mol))
    return nu_t, alpha_t, c_nu, c_k

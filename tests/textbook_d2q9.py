"""An independent textbook D2Q9 double-distribution step in numpy, used only
to check the oracle.

It is written from the physics, vectorised over the lattice and in a different
algebraic form from the oracle: the equilibria use the 3 / 4.5 / 1.5
coefficients, the TRT collision relaxes the even and odd parts of f - feq
separately, the Guo source is split into its even and odd parts by reflection
(S_i -/+ S_opp(i))/2, streaming is an array shift with a bounce-back mask, and
the lattice tables are derived from the velocity set rather than listed.
Agreement with the oracle is therefore to round-off, not bitwise.
"""
import numpy as np

CS2 = 1.0 / 3.0
# velocity set in the C-ABI's documented order
C = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [-1, 1], [-1, 0], [-1, -1], [0, -1], [1, -1]])
SPEED2 = (C ** 2).sum(1)
W = np.where(SPEED2 == 0, 4.0 / 9.0, np.where(SPEED2 == 1, 1.0 / 9.0, 1.0 / 36.0))
OPP = np.array([int(np.nonzero((C == -c).all(1))[0][0]) for c in C])
CX = C[:, 0].astype(float)[:, None, None]
CY = C[:, 1].astype(float)[:, None, None]
WW = W[:, None, None]


def equilibrium(scalar, ux, uy, quadratic=True):
    cu = CX * ux + CY * uy
    if not quadratic:
        return WW * scalar * (1.0 + 3.0 * cu)
    usq = ux * ux + uy * uy
    return WW * scalar * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)


class TextbookCavity:
    """State arrays are [9, ny, nx] populations and [ny, nx] fields."""

    def __init__(self, f, g, rho, ux, uy, temp, fy, nu_t, *, nu, alpha, beta_eff, spin_up,
                 cs, pr_t, t_ref=0.0, theta_hot=None, quadratic_g=True, dynamic=None):
        self.f, self.g = f.copy(), g.copy()
        self.rho, self.ux, self.uy, self.temp = rho.copy(), ux.copy(), uy.copy(), temp.copy()
        self.fy, self.nu_t = fy.copy(), nu_t.copy()
        self.nu, self.alpha, self.beta_eff, self.spin_up = nu, alpha, beta_eff, spin_up
        self.cs, self.pr_t, self.t_ref = cs, pr_t, t_ref
        self.theta_hot = theta_hot
        self.quad = quadratic_g
        # dynamic mixed-scale model parameters (smag_alpha, nu_sm_min, eta_test)
        # as keywords of tests.dynamic_sgs_numpy.dynamic_sgs, or None
        self.dynamic = dynamic
        self.time_step = 0

    def collide(self, step):
        ramp = min(step / self.spin_up, 1.0)
        beta = self.beta_eff * ramp
        tau_mol = 0.5 + 3.0 * max(self.nu, 0.002 / 3.0)
        rho, ux, uy, temp = self.rho, self.ux, self.uy, self.temp
        self.fy = rho * beta * (temp - self.t_ref)
        fneq = self.f - equilibrium(rho, ux, uy)
        pxx = (CX * CX * fneq).sum(0)
        pxy = (CX * CY * fneq).sum(0)
        pyy = (CY * CY * fneq).sum(0)
        smag = np.sqrt(pxx ** 2 + 2.0 * pxy ** 2 + pyy ** 2) / (2.0 * CS2 * tau_mol * np.maximum(rho, 1e-12))
        if self.dynamic is not None:
            from tests.dynamic_sgs_numpy import dynamic_sgs
            self.nu_t, alpha_t, _, _ = dynamic_sgs(self.f, rho, ux, uy, temp, tau=tau_mol,
                                                   nu=self.nu, alpha_mol=self.alpha,
                                                   **self.dynamic)
        else:
            self.nu_t = self.cs ** 2 * smag
            alpha_t = self.nu_t / self.pr_t
        tp = 0.5 + 3.0 * (self.nu + self.nu_t)
        tm = 0.5 + (3.0 / 16.0) / (tp - 0.5)
        even = 0.5 * (fneq + fneq[OPP])
        odd = 0.5 * (fneq - fneq[OPP])
        # Guo source with F = (0, fy): w_i [ (c_i - u).F / cs2 + (c_i.u)(c_i.F) / cs2^2 ]
        cu = CX * ux + CY * uy
        src = WW * ((CY - uy) * self.fy / CS2 + cu * CY * self.fy / CS2 ** 2)
        s_odd = 0.5 * (src - src[OPP])
        s_even = 0.5 * (src + src[OPP])
        self.f = (self.f - even / tp - odd / tm + (1.0 - 0.5 / tm) * s_odd
                  + (1.0 - 0.5 / tp) * s_even)
        tau_t = 0.5 + 3.0 * (self.alpha + alpha_t)
        self.g = self.g - (self.g - equilibrium(temp, ux, uy, self.quad)) / tau_t

    def stream(self):
        ny, nx = self.rho.shape
        iy, ix = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        fn = np.empty_like(self.f)
        gn = np.empty_like(self.g)
        for i, (cx, cy) in enumerate(C):
            sy, sx = iy - cy, ix - cx
            inside = (sx >= 0) & (sx < nx) & (sy >= 0) & (sy < ny)
            syc, sxc = np.clip(sy, 0, ny - 1), np.clip(sx, 0, nx - 1)
            fn[i] = np.where(inside, self.f[i][syc, sxc], self.f[OPP[i]])
            gn[i] = np.where(inside, self.g[i][syc, sxc], self.g[i])
        self.f, self.g = fn, gn

    def thermal_bc(self):
        ny = self.rho.shape[0]
        th = np.full(ny, 0.5) if self.theta_hot is None else np.asarray(self.theta_hot)
        # isothermal walls: g = geq(theta, u = 0) = w_i * theta
        self.g[:, :, 0] = WW[:, :, 0] * th[None, :]
        self.g[:, :, -1] = WW[:, :, 0] * -0.5
        self.g[:, 0, :] = self.g[:, 1, :]
        self.g[:, -1, :] = self.g[:, -2, :]

    def macroscopic(self):
        rho = self.f.sum(0)
        self.rho = rho
        self.ux = (CX * self.f).sum(0) / rho
        self.uy = (CY * self.f).sum(0) / rho + 0.5 * self.fy / rho
        self.temp = self.g.sum(0)
        for a in (self.ux, self.uy):
            a[0, :] = a[-1, :] = 0.0
            a[:, 0] = a[:, -1] = 0.0

    def step(self):
        self.collide(self.time_step)
        self.stream()
        self.thermal_bc()
        self.macroscopic()
        self.time_step += 1

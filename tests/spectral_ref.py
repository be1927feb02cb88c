"""Second restatement of the periodic right-hand side, for pinning the oracle.

TEST INFRASTRUCTURE ONLY.  Written from the reference text, independently
of oracle/st_oracle.c: every periodic compact operator is applied through
its exact discrete symbol in Fourier space instead of a band solve.  A
constant-coefficient cyclic band matrix is a circulant, so the DFT
diagonalises it exactly, and this module reproduces the compact scheme to
rounding.  It restates, from the reference:

  RHS        src/rhsv_periodic_sf.f:9-25 (RHS_x :36-461, RHS_yz :472-1385)
  Compute_coeff1, Compute_Contraction, Compute_Laplacian, Compute_hyper
             src/hyper_dynamic.f:12-318
  Hyper_RHS  src/hyper_rhs.f:12-499
  LDDRK4 stage body and Fix_Pressure
             src/lddrk4_periodic_sf.f:344-356, :436-473

Symbols (theta = k h):
  D1       i k~,   k~ h = [(14/9) sin th + (1/18) sin 2th] / [1 + (2/3) cos th]
           (src/header:193-197: alpha = 1/3, a = 14/18 on f+-1, b = 1/36 on f+-2)
  D2       -k2~^2, k2~^2 h^2 = [(24/11)(1 - cos th) + (3/22)(1 - cos 2th)]
                               / [1 + (4/11) cos th]
           (src/header:198-203 with the consistent centre weight 2(a2+b2))
  filter   [a0 + 2 sum a_m cos m th] / [1 + 2 alpha cos th + 2 beta cos 2th]
           (src/LES_header:107-113)
  Gaussian  a0 + 2 sum a_m cos m th  (src/LES_header:153-157)

Arrays are F(My, Mz, Mx) (axis 0 = y, 1 = z, 2 = x); Q(..., 5) holds
rho u, rho v, rho w, rho, rho E.  Grid spacing is period / N per direction
(src/define_periodic_sf.f:42-44, src/grid.f:15-112).
"""
import numpy as np

from fields import GAUSS_A, sym_d1, sym_d2, sym_filter, sym_gauss

AXIS = {0: 2, 1: 0, 2: 1}


class Spectral:
    def __init__(self, nx, ny, nz, period=(2 * np.pi,) * 3, Re=100.0, Pr=0.7, gamma=1.4, vis_exp=0.76,
                 bulk_ratio=0.0):
        self.n = (nx, ny, nz)
        self.h = tuple(period[d] / self.n[d] for d in range(3))
        self.gamma = gamma
        self.re_inv = 1.0 / Re if Re > 0 else 0.0
        self.pr_inv = 1.0 / Pr if Pr > 0 else 0.0
        self.vis_exp = vis_exp
        self.bulk = bulk_ratio
        self._sym = {}
        for d in range(3):
            n = self.n[d]
            th = 2 * np.pi * np.fft.fftfreq(n)
            h = self.h[d]
            d1 = 1j * sym_d1(th) / h
            if n % 2 == 0:
                d1[n // 2] = 0.0       # the symbol vanishes there; keep it exactly 0
            self._sym[d] = {
                "d1": d1,
                "d2": -sym_d2(th) / (h * h),
                "filt": sym_filter(th).astype(complex),
                "gauss": sym_gauss(th).astype(complex),
            }

    def _apply(self, d, op, f):
        ax = AXIS[d]
        m = self._sym[d][op]
        sh = [1, 1, 1]
        sh[ax] = len(m)
        return np.real(np.fft.ifft(np.fft.fft(f, axis=ax) * m.reshape(sh), axis=ax))

    def d1(self, d, f):
        return self._apply(d, "d1", f)

    def d2(self, d, f):
        return self._apply(d, "d2", f)

    def filt(self, d, f):
        return self._apply(d, "filt", f)

    def gauss(self, d, f):
        return self._apply(d, "gauss", f)

    def lap(self, f):
        """Compute_Laplacian: D2x + D2y + D2z (src/hyper_dynamic.f:41-58)."""
        t1 = self.d2(0, f)
        t2 = self.d2(1, f)
        return self.d2(2, f) + t1 + t2

    # ---- thermodynamics (src/rhsv_periodic_sf.f:89-104) ----
    def prim(self, Q):
        g = self.gamma
        fac1 = g - 1.0
        fac2 = g / fac1
        Vs = 1.0 / Q[..., 3]
        P = fac1 * (Q[..., 4] - 0.5 * Vs * (Q[..., 0] ** 2 + Q[..., 1] ** 2 + Q[..., 2] ** 2))
        T = fac2 * Vs * P
        return Vs, P, T

    # ---- RHS (src/rhsv_periodic_sf.f) ----
    def rhs(self, Q):
        """Returns (Qrhs, grad12) with grad12 in /GRAD/ order: dux_dx duy_dx
        duz_dx dux_dy duy_dy duz_dy dux_dz duy_dz duz_dz dTdx dTdy dTdz."""
        g = self.gamma
        fac1 = g - 1.0
        Vs, P, T = self.prim(Q)
        rho = Q[..., 3]
        E = Q[..., 4]
        m = [Q[..., i] for i in range(3)]
        u = [m[i] * Vs for i in range(3)]
        D1, D2 = self.d1, self.d2
        mu = ((fac1 * T) ** self.vis_exp) * self.re_inv
        mup = ((fac1 * T) ** (self.vis_exp - 1.0)) * self.re_inv * self.vis_exp * fac1
        conv = [np.zeros_like(rho) for _ in range(5)]
        visc = [np.zeros_like(rho) for _ in range(5)]
        # convective terms in skew-symmetric form, direction d: the
        # divergence of the mass flux, -1/2 [D(m_i u_d) + u_i D(m_d) +
        # m_d D(u_i)], -D p, -D((rho E + p) u_d)  (RHS_x :86-234, RHS_yz :530-1000)
        du = [[None] * 3 for _ in range(3)]     # du[i][d] = D_d u_i
        for d in range(3):
            ud = u[d]
            if d == 0:
                conv[3] = conv[3] - D1(0, rho * ud)
            else:
                conv[3] = conv[3] - D1(d, m[d])
            conv[4] = conv[4] - D1(d, (E + P) * ud)
            for i in range(3):
                conv[i] = conv[i] - 0.5 * D1(d, m[i] * ud)
            conv[d] = conv[d] - D1(d, P)
            dm = D1(d, m[d])
            for i in range(3):
                du[i][d] = D1(d, u[i])
                conv[i] = conv[i] - 0.5 * dm * u[i] - 0.5 * rho * ud * du[i][d]
        dT = [D1(d, T) for d in range(3)]
        d2T = [D2(d, T) for d in range(3)]
        # conduction (1/Pr)(mu Lap T + mu' |grad T|^2)  (RHS_x :296-306, RHS_yz :853-864)
        visc[4] = visc[4] + self.pr_inv * (mu * (d2T[0] + d2T[1] + d2T[2])
                                           + mup * (dT[0] ** 2 + dT[1] ** 2 + dT[2] ** 2))
        # expanded viscous stress: mu [Lap u_i + (1/3 + bulk) d_i dil]
        # + mu' grad T . (grad u + grad u^T - (2/3 - bulk) dil I), and the
        # energy flux u . (those) + mu (2 S:S - (2/3 - bulk) dil^2) form
        fA = 4.0 / 3.0 + self.bulk
        f13 = 1.0 / 3.0 + self.bulk
        fm = self.bulk - 2.0 / 3.0
        d2u = [[D2(d, u[i]) for d in range(3)] for i in range(3)]
        dil = du[0][0] + du[1][1] + du[2][2]
        for i in range(3):
            lap_i = sum((fA if d == i else 1.0) * d2u[i][d] for d in range(3))
            # mixed derivatives d_i d_j u_j (j != i).  The reference takes each
            # as a D1 of a stored D1 (RHS_yz :1062-1373: D_y(D_x u), D_z(D_x u),
            # D_y(D_x v), D_z(D_x w), D_y(D_z v), D_y(D_z w)); periodic D1s
            # commute exactly in Fourier space, so the order moves rounding only.
            mix = sum(D1(i, du[j][j]) for j in range(3) if j != i)
            # mu' dT_j (d_j u_i + d_i u_j), the diagonal term 2 d_i u_i + (bulk - 2/3) dil
            s = sum(mup * dT[j] * ((2.0 * du[i][i] + fm * dil) if j == i else (du[i][j] + du[j][i]))
                    for j in range(3))
            f = mu * lap_i + f13 * mu * mix + s
            visc[i] = visc[i] + f
            visc[4] = visc[4] + u[i] * f
        shear = (du[0][1] + du[1][0]) ** 2 + (du[0][2] + du[2][0]) ** 2 + (du[1][2] + du[2][1]) ** 2
        visc[4] = visc[4] + mu * (2.0 * (du[0][0] ** 2 + du[1][1] ** 2 + du[2][2] ** 2) + fm * dil * dil + shear)
        Qrhs = np.zeros(Q.shape, order="F")
        for L in range(5):
            Qrhs[..., L] = conv[L] + visc[L]
        G = np.zeros(Q.shape[:3] + (12,), order="F")
        for d in range(3):
            for i in range(3):
                G[..., 3 * d + i] = du[i][d]
            G[..., 9 + d] = dT[d]
        return Qrhs, G

    # ---- artificial transport (src/hyper_dynamic.f) ----
    def coeff1(self, Q):
        """C_beta of Compute_coeff1 (src/hyper_dynamic.f:72-165)."""
        tmp = 1.0 / Q[..., 3]
        U = [Q[..., i] * tmp for i in range(3)]
        wx = self.d1(1, U[2]) - self.d1(2, U[1])
        wy = -self.d1(0, U[2]) + self.d1(2, U[0])
        wz = self.d1(0, U[1]) - self.d1(1, U[0])
        dil = self.d1(0, U[0]) + self.d1(1, U[1]) + self.d1(2, U[2])
        fac = 0.5 * (1.0 - np.tanh(2.5 + 10.0 * dil))
        om2 = wx * wx + wy * wy + wz * wz
        return fac * dil * dil / (dil * dil + om2 + 1.0e-7) + 1.0e-20

    def strain(self, G):
        """Compute_Contraction of the symmetrised gradient (:12-33, :227-245)."""
        s12 = 0.5 * (G[..., 3] + G[..., 1])
        s13 = 0.5 * (G[..., 6] + G[..., 2])
        s23 = 0.5 * (G[..., 7] + G[..., 5])
        return np.sqrt(0.5 * (G[..., 0] ** 2 + G[..., 4] ** 2 + G[..., 8] ** 2
                              + 2 * (s12 * s12 + s13 * s13 + s23 * s23)))

    def hyper(self, Q, G):
        """beta*, mu*, kappa* of Compute_hyper (src/hyper_dynamic.f:173-318)."""
        g = self.gamma
        Cb = self.coeff1(Q)
        S = self.strain(G)
        Vs = 1.0 / Q[..., 3]
        ke = Q[..., 0] ** 2 + Q[..., 1] ** 2 + Q[..., 2] ** 2
        P = (g - 1.0) * (Q[..., 4] - 0.5 * Vs * ke)
        T = g / (g - 1.0) * Vs * P
        e = Vs * (Q[..., 4] - 0.5 * Vs * ke)
        d4S = np.abs(self.lap(self.lap(S)))
        d4e = np.abs(self.lap(self.lap(e)))
        fac3 = (Q[..., 3] / T) * np.sqrt(g * P * Vs)
        eta = [Q[..., 3] * d4S, fac3 * d4e, Q[..., 3] * d4S]
        etag = [self.gauss(2, self.gauss(1, self.gauss(0, f))) for f in eta]
        DX = self.h[0]
        H = np.zeros(Q.shape[:3] + (3,), order="F")
        H[..., 0] = Cb * etag[2] * DX ** 6
        H[..., 1] = 2.0e-3 * etag[0] * DX ** 6
        H[..., 2] = 1.0e-2 * etag[1] * DX ** 5
        return H

    def hyper_rhs(self, Q, G, H):
        """The artificial-transport terms Qh of Hyper_RHS (src/hyper_rhs.f:12-499)."""
        g = self.gamma
        beta, vis, kap = H[..., 0], H[..., 1], H[..., 2]
        U = [Q[..., i] / Q[..., 3] for i in range(3)]
        du = [[G[..., 3 * d + i] for d in range(3)] for i in range(3)]   # du[i][d] = d u_i / d x_d
        dil = du[0][0] + du[1][1] + du[2][2]
        Sx = {(0, 1): du[0][1] + du[1][0], (0, 2): du[0][2] + du[2][0], (1, 2): du[1][2] + du[2][1]}
        Qh = np.zeros(Q.shape, order="F")
        for i in range(3):
            for d in range(3):
                if d == i:
                    flux = 2.0 * vis * du[i][i] + (beta - (2.0 / 3.0) * vis) * dil
                else:
                    flux = vis * Sx[tuple(sorted((i, d)))]
                Qh[..., i] += self.d1(d, flux)
                Qh[..., 4] += self.d1(d, U[i] * flux)
        for d in range(3):
            Qh[..., 4] += self.d1(d, kap * G[..., 9 + d])
        return Qh

    def full_rhs(self, Q):
        r, G = self.rhs(Q)
        H = self.hyper(Q, G)
        return r + self.hyper_rhs(Q, G, H), G, H

    # ---- stage (src/lddrk4_periodic_sf.f:344-356, :436-473) ----
    def fix_pressure(self, Q):
        g = self.gamma
        Vs = 1.0 / Q[..., 3]
        ke = Q[..., 0] ** 2 + Q[..., 1] ** 2 + Q[..., 2] ** 2
        P = (g - 1.0) * (Q[..., 4] - 0.5 * Vs * ke)
        P = np.where(P < 0.0, 0.71463, P)
        out = Q.copy(order="F")
        out[..., 4] = (1.0 / (g - 1.0)) * P + 0.5 * Vs * ke
        return out

    def stage(self, Q, QW, alpha, beta, dt, hyper=True):
        if hyper:
            dQ, _, _ = self.full_rhs(Q)
        else:
            dQ, _ = self.rhs(Q)
        QW = alpha * QW + dt * dQ
        Qs = Q + beta * QW
        Qf = np.zeros_like(Qs, order="F")
        for L in range(5):
            Qf[..., L] = self.filt(2, self.filt(1, self.filt(0, Qs[..., L])))
        return self.fix_pressure(Qf), QW


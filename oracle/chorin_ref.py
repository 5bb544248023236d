"""numpy restatement of HopeFOAM's incompressible projection solver dgChorinFoam on the pearsonVortex case.

TEST INFRASTRUCTURE ONLY: the checker of the HIP operators in hopefoam_amd/csrc/ins.hip, of the Neumann data of
the SIPG right-hand side, and of the device-resident driver hopefoam_amd/chorin.py.  Only tests/ load it.

It follows (paths relative to HopeFOAM-0.1/):
  tutorials/DG/2D/pearsonVortex/dgChorinFoam/dgChorinFoam.C:63-109   the time loop
  tutorials/DG/2D/pearsonVortex/dgChorinFoam/pEqnCorrect.H            curl curl U on the inlet, nbp
  tutorials/DG/2D/pearsonVortex/dgChorinFoam/pCorrectEquation.H:60-130 the pressure's Neumann data b -= Vf^T W Vf
  tutorials/DG/2D/pearsonVortex/dgChorinFoam/setNonUniformInlet.H, setBoundaryValues.H, setparaT.H, chorinerror.H
  src/DG/DG/convectionSchemes/defaultConvectionScheme/defaultConvectionScheme.C:215-303 (dof-field convection)
  src/DG/DG/simpleFlux/schemes/LFFlux/LFFlux.C:103-211               (LF flux at the face nodes, maxV per face)
  src/DG/DG/gradSchemes/defaultGrad/defaultGrad.C:45-84, divSchemes/defaultDivScheme/defaultDivScheme.C:45-83
  src/DG/DG/laplacianSchemes/defaultLaplacianScheme/defaultLaplacianScheme.C:62-288 (through the C oracle's
      assembled SIPG matrix; the fixedGradient term of :210-220 restated here)
  src/DG/Equation/Equation.C:42-79, EquationEqual.H, EquationAdd.H, dgLduMatrix.C:487-595 (how the terms combine)
Every linear system is solved directly (the tutorial's PETSc LU / preonly).

The convection is restated face by face as the reference evaluates it: one flux per face from the owner's side,
subtracted in the owner and added in the neighbour.  Pins (tests/test_oracle_chorin.py): exact divergence of
quadratic fluxes, conservation, exact nodal derivatives of polynomials, exact reproduction of polynomial Poisson
solutions with Neumann data, and the convergence of the whole scheme to the exact Taylor-Green (Pearson) vortex.
"""
import numpy as np

PI = np.pi


# ---------------------------------------------------------------- operators
def nodal_grad(ops, geo, u):
    """dgc::grad of a dof field (defaultGrad.C:45-84): (du/dx, du/dy) at the nodes; u [K][Np] -> [2][K][Np]"""
    ur, us = u @ ops["Dr"].T, u @ ops["Ds"].T
    rx, ry, sx, sy = (geo[:, i][:, None] for i in range(4))
    return np.stack([rx * ur + sx * us, ry * ur + sy * us])


def nodal_div(ops, geo, U):
    """dgc::div of a dof vector field (defaultDivScheme.C:45-83): du/dx + dv/dy at the nodes"""
    gx = nodal_grad(ops, geo, U[0])[0]
    gy = nodal_grad(ops, geo, U[1])[1]
    return gx + gy


def convection(ops, geo, EToE, EToF, U, bcU=None, fixed=None):
    """convecU of dg::solveEquation(dgm::Sp(convecU) == dgc::div(U, U)) with the LF scheme: U [2][K][Np];
    bcU [2][K][3][Nfp] patch values of fixedValue faces (fixed [K][3] bool), other boundary faces take the own
    trace.  Returns [2][K][Np]."""
    N = ops["fmask"].shape[1] - 1
    fm = ops["fmask"]
    K, Np = U.shape[1], U.shape[2]
    Vq, Drq, Dsq, w, Vf, gw, M = ops["Vq"], ops["Drq"], ops["Dsq"], ops["cw"], ops["Vf"], ops["gw"], ops["M"]
    rx, ry, sx, sy, J = (geo[:, i] for i in range(5))
    b = np.zeros((2, K, Np))
    # cell integration: b_j += sum_k sum_i (d phi_j / d x_k)(x_i) W_i J_i (Vq (U_k u_c))_i
    for e in range(K):
        Dx = Drq * rx[e] + Dsq * sx[e]          # d phi_j / dx at the cubature points [Nq][Np]
        Dy = Drq * ry[e] + Dsq * sy[e]
        for c in range(2):
            gx = Vq @ (U[0, e] * U[c, e]) * w * J[e]
            gy = Vq @ (U[1, e] * U[c, e]) * w * J[e]
            b[c, e] += Dx.T @ gx + Dy.T @ gy
    # face integration, one flux per face from the owner's side
    for e in range(K):
        for f in range(3):
            en = EToE[e, f]
            if en >= 0 and en < e:
                continue                        # the face belongs to its lower-numbered element
            nx, ny, sJ = geo[e, 5 + 3 * f], geo[e, 6 + 3 * f], geo[e, 7 + 3 * f]
            Uo = U[:, e, fm[f]]
            if en >= 0:
                Un = U[:, en, fm[EToF[e, f]][::-1]]
            elif fixed is not None and fixed[e, f] and bcU is not None:
                Un = bcU[:, e, f, :]
            else:
                Un = Uo
            vo = nx * Uo[0] + ny * Uo[1]
            vn = nx * Un[0] + ny * Un[1]
            maxv = max(np.abs(vo).max(), np.abs(vn).max())
            Fn = (vo * Uo + vn * Un) * 0.5 + maxv * (Uo - Un) * 0.5      # [2][Nfp] at the face nodes
            Fg = Fn @ Vf.T                                                 # face Gauss points
            contrib = (Fg * (sJ * gw)) @ Vf                                # [2][Nfp]
            b[:, e, fm[f]] -= contrib
            if en >= 0:
                b[:, en, fm[EToF[e, f]][::-1]] += contrib
    # Sp(convecU) == div: M convecU = -b
    Minv = np.linalg.inv(M)
    return -np.einsum("jk,cek->cej", Minv, b) / J[None, :, None]


def face_mass_term(ops, geo, EToE, g, faces):
    """sum over the selected boundary faces of Vf^T W Vf g (W = face Gauss weights times the edge Jacobian),
    g [K][3][Nfp] nodal face data; returns [K][Np] (defaultLaplacianScheme.C:210-220, pCorrectEquation.H:103-126)"""
    fm = ops["fmask"]
    K, Np = geo.shape[0], ops["M"].shape[0]
    Vf, gw = ops["Vf"], ops["gw"]
    out = np.zeros((K, Np))
    for e in range(K):
        for f in range(3):
            if EToE[e, f] >= 0 or not faces[e, f]:
                continue
            out[e, fm[f]] += Vf.T @ (geo[e, 7 + 3 * f] * gw * (Vf @ g[e, f]))
    return out


# ---------------------------------------------------------------- the pearsonVortex case
def exact_u(x, y, t, nu):
    a = np.exp(-4.0 * PI * PI * nu * t)
    return np.stack([-np.sin(2 * PI * y) * a, np.sin(2 * PI * x) * a])


def exact_p(x, y, t, nu):
    return -np.cos(2 * PI * x) * np.cos(2 * PI * y) * np.exp(-8.0 * PI * PI * nu * t)


class ChorinRef:
    """dgChorinFoam on one mesh.  om_p, om_U: oracle meshes of the same elements whose fixedValue (tag 1) faces
    are the pressure's Dirichlet patch (outlet) and the velocity's (inlet).  Every solve is direct."""

    def __init__(self, om_p, om_U, ops, nu, dt):
        import scipy.sparse.linalg as spla
        self.spla = spla
        self.ops, self.nu, self.dt = ops, nu, dt
        self.om_p, self.om_U = om_p, om_U
        self.geo = om_U.geo()
        self.EToE, self.EToF = om_U.conn()
        self.x, self.y = om_U.nodes()
        fm = ops["fmask"]
        self.K, self.Np = self.x.shape
        self.xb = np.stack([self.x[:, fm[f]] for f in range(3)], axis=1)       # [K][3][Nfp]
        self.yb = np.stack([self.y[:, fm[f]] for f in range(3)], axis=1)
        bnd = self.EToE < 0
        self.inlet = bnd & (om_U.bctag().reshape(self.K, 3) == 1)
        self.outlet = bnd & (om_p.bctag().reshape(self.K, 3) == 1)
        self.g0, self.a0, self.a1, self.b0, self.b1 = 1.0, 1.0, 0.0, 1.0, 0.0
        self.t = 0.0
        # setNonUniformInlet.H
        self.U = exact_u(self.x, self.y, 0.0, nu)
        self.p = exact_p(self.x, self.y, 0.0, nu)
        self.Uold = self.U.copy()
        self.convecU = np.zeros_like(self.U)
        self.nbp = np.zeros((self.K, 3, fm.shape[1]))
        self._bc(0.0)
        self.paraT3 = np.exp(-nu * 8 * PI * PI * dt)
        self.paraT4 = np.exp(-nu * 4 * PI * PI * dt)
        self._Ap = None
        self._AU = {}

    def _bc(self, t):
        """setBoundaryValues.H at time t: U inlet value, U outlet gradient, p outlet value; paraT1, paraT2"""
        nu = self.nu
        self.paraT1 = np.exp(-nu * 8 * PI * PI * t)
        self.paraT2 = np.exp(-nu * 4 * PI * PI * t)
        self.U_in = exact_u(self.xb, self.yb, 0.0, nu) * self.paraT2                     # [2][K][3][Nfp]
        nx = self.geo[:, 5::3][:, :, None]
        ny = self.geo[:, 6::3][:, :, None]
        self.gradU_out = np.stack([2 * PI * ny * (-np.cos(2 * PI * self.yb)) * self.paraT2,
                                   2 * PI * nx * np.cos(2 * PI * self.xb) * self.paraT2])
        self.p_out = -np.cos(2 * PI * self.xb) * np.cos(2 * PI * self.yb) * self.paraT1

    def pressure_neumann(self, U, convecU, paraT2):
        """pEqnCorrect.H + pCorrectEquation.H:91-101: nbp = n . (convecU + nu curl(curl U)) at the inlet face nodes,
        plus n . u_exact (-4 pi^2 nu paraT2) (the exact du/dt); [K][3][Nfp], zero elsewhere.  For the exact
        solution it is -dp/dn."""
        ops, geo, nu = self.ops, self.geo, self.nu
        fm = ops["fmask"]
        Dr, Ds = ops["Dr"], ops["Ds"]
        paraT = -nu * 4 * PI * PI * paraT2
        nbp = np.zeros((self.K, 3, fm.shape[1]))
        for e in range(self.K):
            if not self.inlet[e].any():
                continue
            rx, ry, sx, sy = geo[e, :4]
            Dx, Dy = Dr * rx + Ds * sx, Dr * ry + Ds * sy
            w = -(Dy @ U[0, e]) + Dx @ U[1, e]                              # vorticity at the nodes
            cc = np.stack([Dy @ w, -(Dx @ w)])                             # curl of the vorticity
            curlU = convecU[:, e] + nu * cc
            for f in range(3):
                if not self.inlet[e, f]:
                    continue
                nx, ny = geo[e, 5 + 3 * f], geo[e, 6 + 3 * f]
                xs, ys = self.x[e, fm[f]], self.y[e, fm[f]]
                nbp[e, f] = nx * curlU[0, fm[f]] + ny * curlU[1, fm[f]]
                nbp[e, f] += (-nx * np.sin(2 * PI * ys) + ny * np.sin(2 * PI * xs)) * paraT
        return nbp

    def step(self):
        ops, geo, dt, nu = self.ops, self.geo, self.dt, self.nu
        K, Np = self.K, self.Np
        self.t += dt
        # step 1: convecU = div(U, U)
        convecUold = self.convecU
        self.convecU = convection(ops, geo, self.EToE, self.EToF, self.U, self.U_in * self.inlet[None, :, :, None],
                                  self.inlet)
        UT = (self.a0 * self.U + self.a1 * self.Uold - (self.b0 * self.convecU + self.b1 * convecUold) * dt) / self.g0
        # step 2: pressure.  pEqnCorrect.H: the Neumann data b0 nbp + b1 nbp_old on the velocity's fixedValue faces
        nbp_old = self.nbp
        self.nbp = self.pressure_neumann(self.U, self.convecU, self.paraT2)
        nbp_comb = self.b0 * self.nbp + self.b1 * nbp_old
        # bCorrect(laplacian(p), paraT3) == pEqnCorrect + div(UT) g0 / dt
        fdiv = (self.g0 / dt) * nodal_div(ops, geo, UT)
        if self._Ap is None:
            A, _ = self.om_p.sipg_system(0, np.zeros((K, Np)))
            self._Ap = self.spla.splu(A.tocsc())
        _, bD = self.om_p.sipg_system(0, np.zeros((K, Np)), self.paraT3 * self.p_out * self.outlet[:, :, None])
        _, bf = self.om_p.sipg_system(0, fdiv, None)
        b = bD + bf + face_mass_term(ops, geo, self.EToE, nbp_comb, self.inlet).reshape(-1)
        self.p = self._Ap.solve(b).reshape(K, Np)
        # step 3: UT -= dt / g0 grad p (nodal)
        UT = UT - (dt / self.g0) * nodal_grad(ops, geo, self.p)
        # step 4: Sp(U) g0 / (dt nu) == bCorrect(laplacian(U), paraT4) + UT g0 / (dt nu)
        tmp = self.g0 / (dt * nu)
        if tmp not in self._AU:
            A, _ = self.om_U.sipg_system(0, np.zeros((K, Np)), None, mass_shift=tmp)
            self._AU[tmp] = self.spla.splu(A.tocsc())
        Unew = np.zeros_like(self.U)
        for c in range(2):
            _, bD = self.om_U.sipg_system(0, np.zeros((K, Np)), self.paraT4 * self.U_in[c] * self.inlet[:, :, None])
            _, bf = self.om_U.sipg_system(0, -tmp * UT[c], None)
            bN = -self.paraT4 * face_mass_term(ops, geo, self.EToE, self.gradU_out[c], self.outlet).reshape(-1)
            Unew[c] = self._AU[tmp].solve(bD + bf + bN).reshape(K, Np)
        self.Uold, self.U = self.U, Unew
        self._bc(self.t)
        if self.a0 < 2.0:
            self.g0, self.a0, self.a1, self.b0, self.b1 = 1.5, 2.0, -0.5, 2.0, -1.0

    def errors(self):
        """chorinerror.H: mean and max |p - p_exact| over the nodes; and the same for |U - U_exact|"""
        pe = exact_p(self.x, self.y, self.t, self.nu)
        ue = exact_u(self.x, self.y, self.t, self.nu)
        dp = np.abs(self.p - pe)
        du = np.hypot(*(self.U - ue))
        return dp.mean(), dp.max(), du.mean(), du.max()

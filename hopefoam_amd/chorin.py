"""The incompressible projection solver dgChorinFoam on the device.

HopeFOAM's tutorial solver (HopeFOAM-0.1/tutorials/DG/2D/pearsonVortex/dgChorinFoam/dgChorinFoam.C:63-109) advances
the Navier-Stokes equations by a stiffly stable velocity-correction scheme: explicit convection (BDF/extrapolation
of order 1 in the first step, 2 afterwards), a pressure Poisson solve with a high-order Neumann condition on the
velocity-Dirichlet patches, a nodal pressure-gradient correction and an implicit (Helmholtz) viscous step.  Each
step here is the same sequence of whole-mesh operations on device-resident fields:

  convecU = div(U, U)                                  dgh_ins_convection  (dgChorinFoam.C:74, "div(U,U) LF")
  UT = (a0 U^n + a1 U^(n-1) - dt (b0 convecU + b1 convecUold)) / g0          dgh_field_lincomb (:76)
  laplacian(p) == div(UT) g0 / dt + pressure Neumann data               dgh_nodal_div, dgh_sipg_rhs_neumann,
                                                                          dgh_sipg_solve (:80-85, pEqnCorrect.H)
  UT -= dt / g0 grad(p)                                dgh_nodal_grad, dgh_field_lincomb (:89)
  (g0 / (dt nu) - laplacian) U = g0 / (dt nu) UT       mass-shifted SIPG per component (:92-97)

The tutorial-specific parts are host code, as in the reference application: the exact Taylor-Green (Pearson)
vortex boundary data of setNonUniformInlet.H / setBoundaryValues.H, the time factors of setparaT.H, and the
pressure's Neumann data of pEqnCorrect.H (n . (convecU + nu curl curl U^n) on the inlet face nodes, plus the exact
-n . dU/dt term), evaluated on the boundary elements only.  The laplacian's Dirichlet and gradient data are the
boundary values of t^n scaled by the exact decay over one step (bCorrectEquation with paraT3, paraT4), as the
tutorial does.
"""
import numpy as np

from . import caseio
from .dgh import (DGH_PC_BLOCK_JACOBI, DGH_SIPG_PENALTY_REFERENCE, DGH_SOLVE_FP64, refelem_get)

PI = np.pi


def exact_u(x, y, t, nu):
    """Taylor-Green (Pearson) vortex velocity (setNonUniformInlet.H:22-24, setBoundaryValues.H:46-48)"""
    a = np.exp(-4.0 * PI * PI * nu * t)
    return np.stack([-np.sin(2 * PI * y) * a, np.sin(2 * PI * x) * a])


def exact_p(x, y, t, nu):
    """its pressure (chorinerror.H:10-12)"""
    return -np.cos(2 * PI * x) * np.cos(2 * PI * y) * np.exp(-8.0 * PI * PI * nu * t)


class ChorinSolver:
    """dgChorinFoam on one case directory (the pearsonVortex layout: U fixedValue on the inlet and fixedGradient
    on the outlet, p zeroGradient on the inlet and fixedValue on the outlet).  rtol: the relative tolerance of the
    PCG solves that replace the tutorial's direct (LU) solves, abstol the absolute floor of their preconditioned
    residual norm."""

    def __init__(self, ctx, case_dir, mesh=None, rtol=1e-12, abstol=1e-14, maxit=5000):
        self.ctx = ctx
        self.case = caseio.Case(case_dir, mesh=mesh)
        c = self.case
        self.N = N = c.N
        tp = caseio.read_dict(c.path + "/constant/transportProperties")
        nu = tp["nu"]
        self.nu = float(nu[-1] if isinstance(nu, list) else nu)
        self.dt = float(c.control["deltaT"])
        self.end_time = float(c.control["endTime"])
        m = c.mesh2d()
        self.K = m.K
        tag_p = c.bc_tags(["p"])
        tag_U = c.bc_tags(["U"])
        types_U = self._patch_types("U")
        types_p = self._patch_types("p")
        # the velocity's fixedValue patches (where pEqnCorrect adds the pressure's Neumann data), its fixedGradient
        # patches and the pressure's fixedValue patches
        self.u_fixed = self._faces(m, types_U, "fixedValue")
        self.u_grad = self._faces(m, types_U, "fixedGradient")
        self.p_fixed = self._faces(m, types_p, "fixedValue")
        vxy = m.vertices()
        self.mU = ctx.mesh(N, vxy, m.tris, tag_U)
        self.mp = ctx.mesh(N, vxy, m.tris, tag_p)
        self.Np, self.Nfp = self.mU.Np, self.mU.Nfp
        self.x, self.y = self.mU.nodes()
        self.geo = self.mU.geo()
        self.fm = refelem_get(N, "fmask").astype(int).reshape(3, N + 1)
        self.Dr = refelem_get(N, "Dr").reshape(self.Np, self.Np)
        self.Ds = refelem_get(N, "Ds").reshape(self.Np, self.Np)
        self.xb = np.stack([self.x[:, self.fm[f]] for f in range(3)], axis=1)     # [K][3][Nfp]
        self.yb = np.stack([self.y[:, self.fm[f]] for f in range(3)], axis=1)
        self.inlet_elems = np.nonzero(self.u_fixed.any(axis=1))[0]
        # fields
        self.t = 0.0
        U0 = exact_u(self.x, self.y, 0.0, self.nu)
        self.U = self.mU.field(2, U0)
        self.Uold = self.mU.field(2, U0)
        self.convecU = self.mU.field(2, np.zeros_like(U0))
        self.convecUold = self.mU.field(2, np.zeros_like(U0))
        self.UT = self.mU.field(2)
        self.gradp = self.mU.field(2)
        self.p = self.mp.field(1, exact_p(self.x, self.y, 0.0, self.nu))
        self.divUT = self.mp.field(1)
        self.b = self.mp.field(1)
        self.f = self.mp.field(1)
        self.uc = self.mU.field(1)
        self.bu = self.mU.field(1)
        self.fu = self.mU.field(1)
        self.nbp = np.zeros((self.K, 3, self.Nfp))
        self.g0, self.a0, self.a1, self.b0, self.b1 = 1.0, 1.0, 0.0, 1.0, 0.0
        self.paraT3 = np.exp(-self.nu * 8 * PI * PI * self.dt)
        self.paraT4 = np.exp(-self.nu * 4 * PI * PI * self.dt)
        self.sipg_p = self.mp.sipg(penalty=DGH_SIPG_PENALTY_REFERENCE, pc=DGH_PC_BLOCK_JACOBI, mode=DGH_SOLVE_FP64,
                                   maxit=maxit, rtol=rtol, abstol=abstol)
        self._rtol, self._abstol, self._maxit = rtol, abstol, maxit
        self.sipg_U = {}
        self.stats = []
        self._set_boundary_values(0.0)

    # ---------------------------------------------------------------- case set-up
    def _patch_types(self, field):
        import os
        from .caseio import _parse_dict, _split_entry, _tokens
        with open(os.path.join(self.case.path, "0", field)) as fh:
            text = fh.read()
        bd, _ = _parse_dict(_tokens(text[_split_entry(text, "boundaryField"):]), 1, "}")
        return {p: (s.get("type") if isinstance(s, dict) else None) for p, s in bd.items()}

    @staticmethod
    def _faces(m, types, kind):
        sel = np.zeros((m.K, 3), bool)
        for p, name in enumerate(m.patches):
            if types.get(name) == kind:
                sel |= m.face_patch == p
        return sel

    def _set_boundary_values(self, t):
        """setBoundaryValues.H at time t: the velocity's inlet values, its outlet gradient, the pressure's outlet
        values, and the factors paraT1, paraT2 of t"""
        nu = self.nu
        self.paraT1 = np.exp(-nu * 8 * PI * PI * t)
        self.paraT2 = np.exp(-nu * 4 * PI * PI * t)
        self.U_in = exact_u(self.xb, self.yb, 0.0, nu) * self.paraT2 * self.u_fixed[None, :, :, None]
        nx = self.geo[:, 5::3][:, :, None]
        ny = self.geo[:, 6::3][:, :, None]
        self.gradU = np.stack([2 * PI * ny * (-np.cos(2 * PI * self.yb)), 2 * PI * nx * np.cos(2 * PI * self.xb)]) \
            * self.paraT2 * self.u_grad[None, :, :, None]
        self.p_out = -np.cos(2 * PI * self.xb) * np.cos(2 * PI * self.yb) * self.paraT1 * self.p_fixed[:, :, None]
        self.mU.set_bc_values(np.ascontiguousarray(self.U_in))

    def _sipg_U(self, tmp):
        if tmp not in self.sipg_U:
            self.sipg_U[tmp] = self.mU.sipg(penalty=DGH_SIPG_PENALTY_REFERENCE, pc=DGH_PC_BLOCK_JACOBI,
                                            mode=DGH_SOLVE_FP64, maxit=self._maxit, rtol=self._rtol,
                                            abstol=self._abstol,
                                            mass_shift=tmp)
        return self.sipg_U[tmp]

    def _pressure_neumann(self):
        """pEqnCorrect.H: on the inlet face nodes nbp = n . (convecU + nu curl(curl U^n)) + n . u_exact(t^n) (-4 pi^2
        nu); the pressure's Neumann data are b0 nbp + b1 nbp_old (pCorrectEquation.H:103-117)"""
        U = self.U.download()
        cu = self.convecU.download()
        nbp_old = self.nbp.copy()
        paraT = -self.nu * 4 * PI * PI * self.paraT2
        geo, fm = self.geo, self.fm
        for e in self.inlet_elems:
            rx, ry, sx, sy = geo[e, :4]
            Dx, Dy = self.Dr * rx + self.Ds * sx, self.Dr * ry + self.Ds * sy
            w = Dx @ U[1, e] - Dy @ U[0, e]
            curl = cu[:, e] + self.nu * np.stack([Dy @ w, -(Dx @ w)])
            for f in range(3):
                if not self.u_fixed[e, f]:
                    continue
                nx, ny = geo[e, 5 + 3 * f], geo[e, 6 + 3 * f]
                xs, ys = self.x[e, fm[f]], self.y[e, fm[f]]
                self.nbp[e, f] = nx * curl[0, fm[f]] + ny * curl[1, fm[f]] + \
                    (-nx * np.sin(2 * PI * ys) + ny * np.sin(2 * PI * xs)) * paraT
        return self.b0 * self.nbp + self.b1 * nbp_old

    # ---------------------------------------------------------------- one time step
    def step(self):
        dt, g0 = self.dt, self.g0
        self.t += dt
        # step 1: velocity prediction
        self.convecUold.copy_from(self.convecU)
        self.mU.ins_convection(self.U, self.convecU)
        self.UT.lincomb([(self.a0 / g0, self.U), (self.a1 / g0, self.Uold), (-dt * self.b0 / g0, self.convecU),
                         (-dt * self.b1 / g0, self.convecUold)])
        # step 2: pressure, laplacian(p) paraT3 == pEqnCorrect + div(UT) g0 / dt
        gN = -self._pressure_neumann()
        self.mU.nodal_div(self.UT, self.divUT)
        self.f.lincomb([(g0 / dt, self.divUT)])
        self.sipg_p.rhs(self.f, self.b, gbc=self.paraT3 * self.p_out, gnb=gN)
        st = [self.sipg_p.solve(self.b, self.p, check=False)]
        # step 3: UT -= dt / g0 grad(p)
        self.mp.nodal_grad(self.p, self.gradp)
        self.UT.lincomb([(1.0, self.UT), (-dt / g0, self.gradp)])
        # step 4: viscous step, Sp(U) g0 / (dt nu) == laplacian(U) paraT4 + UT g0 / (dt nu), per component
        tmp = g0 / (dt * self.nu)
        S = self._sipg_U(tmp)
        self.Uold.copy_from(self.U)
        for cpt in range(2):
            self.fu.copy_var(0, self.UT, cpt)
            self.fu.lincomb([(-tmp, self.fu)])
            self.uc.copy_var(0, self.U, cpt)
            S.rhs(self.fu, self.bu, gbc=self.paraT4 * self.U_in[cpt], gnb=self.paraT4 * self.gradU[cpt])
            st.append(S.solve(self.bu, self.uc, check=False))
            self.U.copy_var(cpt, self.uc, 0)
        self.stats.append(st)
        self._set_boundary_values(self.t)
        if self.a0 < 2.0:
            self.g0, self.a0, self.a1, self.b0, self.b1 = 1.5, 2.0, -0.5, 2.0, -1.0

    def run(self, nsteps=None):
        n = int(round(self.end_time / self.dt)) if nsteps is None else nsteps
        for _ in range(n):
            self.step()

    def fields(self):
        return self.U.download(), self.p.download()

    def errors(self):
        """chorinerror.H: mean and max |p - p_exact| over the nodes, and the same for |U - U_exact|"""
        U, p = self.fields()
        dp = np.abs(p - exact_p(self.x, self.y, self.t, self.nu))
        du = np.hypot(*(U - exact_u(self.x, self.y, self.t, self.nu)))
        return dp.mean(), dp.max(), du.mean(), du.max()

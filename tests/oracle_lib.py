"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests call it to produce expected values and
compare the HIP library against them.  It is never part of the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
ORACLE_XP_SO = os.path.join(ORACLE_DIR, "liboracle_xp.so")

_libs = {}

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib(extended=False):
    """extended=False: FP64 arithmetic as in the reference; True: sums accumulated in long double."""
    if extended in _libs:
        return _libs[extended]
    path = ORACLE_XP_SO if extended else ORACLE_SO
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    L.or_jacobi_p.restype = C.c_double
    L.or_jacobi_p.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
    L.or_grad_jacobi_p.restype = C.c_double
    L.or_grad_jacobi_p.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
    L.or_jacobi_gq.argtypes = [C.c_double, C.c_double, C.c_int, dp, dp]
    L.or_jacobi_gl.argtypes = [C.c_double, C.c_double, C.c_int, dp]
    L.or_np.argtypes = [C.c_int]
    L.or_nq.argtypes = [C.c_int]
    L.or_nodes2d.argtypes = [C.c_int, dp, dp]
    L.or_vandermonde2d.argtypes = [C.c_int, C.c_int, dp, dp, dp]
    L.or_grad_vandermonde2d.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp]
    L.or_invert.argtypes = [C.c_int, dp, dp]
    L.or_cubature.argtypes = [C.c_int, dp, dp, dp]
    L.or_fmask.argtypes = [C.c_int, ip]
    L.or_refelem_ops.argtypes = [C.c_int] + [dp] * 13
    L.or_mesh_create.restype = C.c_void_p
    L.or_mesh_create.argtypes = [C.c_int, C.c_int, dp, ip, C.c_void_p]
    L.or_mesh_destroy.argtypes = [C.c_void_p]
    L.or_mesh_info.argtypes = [C.c_void_p, ip]
    L.or_mesh_conn.argtypes = [C.c_void_p, ip, ip]
    L.or_mesh_nodes.argtypes = [C.c_void_p, dp, dp]
    L.or_mesh_geo.argtypes = [C.c_void_p, dp]
    L.or_rhs_advect.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_void_p, dp]
    L.or_rhs_advect_flux.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_void_p, C.c_int, dp]
    L.or_rhs_euler.argtypes = [C.c_void_p, dp, C.c_double, C.c_int, C.c_void_p, dp]
    L.or_lserk4.argtypes = [C.c_void_p, C.c_int, dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                            C.c_void_p]
    L.or_ssprk2.argtypes = L.or_lserk4.argtypes
    L.or_lserk4_coeffs.argtypes = [dp, dp, dp]
    L.or_vortex_ref_run.argtypes = [C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
    L.or_vortex_exact.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_double, dp, dp, dp, dp]
    opt = {
        "or_tvb_limit": [C.c_void_p, dp, C.c_int, C.c_double, C.c_double, C.c_void_p, ip],
        "or_tvb_ops": [C.c_int, dp, dp, dp],
        "or_positivity_points": [C.c_void_p, dp],
        "or_positivity": [C.c_void_p, dp, C.c_double, C.c_double, ip],
        "or_tvb_geometry": [C.c_void_p, dp, ip, dp],
        "or_triangle_limit": [C.c_void_p, dp, C.c_double, C.c_double, C.c_void_p, C.c_void_p, ip],
        "or_triangle_limit_pf": [C.c_void_p, dp, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, ip],
        "or_ldg_lift": [C.c_void_p, dp],
        "or_ldg_rhs": [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, dp,
                       C.c_void_p],
        "or_ldg_lserk4": [C.c_void_p, dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                          C.c_double, C.c_double],
        "or_sipg_nnz": [C.c_void_p],
        "or_sipg_assemble": [C.c_void_p, C.c_int, dp, C.c_void_p, ip, ip, dp, dp],
        "or_cg_jacobi": [C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_double, C.c_int,
                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p],
        "or_csr_matvec": [C.c_int, ip, ip, dp, dp, dp],
    }
    for name, args in opt.items():
        if hasattr(L, name):
            getattr(L, name).argtypes = args
    _libs[extended] = L
    return L


def _ptr(a):
    """raw pointer of a C-contiguous array (the C side reads the buffer in row-major order)"""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("oracle arrays must be C-contiguous (use np.ascontiguousarray)")
    return a.ctypes.data_as(C.c_void_p)


class OracleMesh:
    """One oracle mesh + reference element of order N."""

    def __init__(self, N, vxy, vid, bctag=None, extended=False):
        self.L = L = lib(extended)
        vxy = np.ascontiguousarray(vxy, dtype=np.float64).reshape(-1)
        vid = np.ascontiguousarray(vid, dtype=np.int32).reshape(-1)
        K = vid.size // 3
        self._bctag = None if bctag is None else np.ascontiguousarray(bctag, dtype=np.int32).reshape(-1)
        self.h = L.or_mesh_create(N, K, vxy, vid, _ptr(self._bctag))
        if not self.h:
            raise ValueError("oracle mesh creation failed")
        info = np.zeros(7, np.int32)
        L.or_mesh_info(self.h, info)
        self.N, self.Np, self.Nfp, self.Nq, self.Nfq, self.K, self.nbfaces = [int(v) for v in info]

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_mesh_destroy(self.h)
            self.h = None

    def bctag(self):
        """the boundary tags the mesh was built with, [K*3] (zeros when none were given)"""
        return np.zeros(3 * self.K, np.int32) if self._bctag is None else self._bctag.copy()

    def conn(self):
        E = np.zeros(3 * self.K, np.int32)
        F = np.zeros(3 * self.K, np.int32)
        self.L.or_mesh_conn(self.h, E, F)
        return E.reshape(self.K, 3), F.reshape(self.K, 3)

    def nodes(self):
        x = np.zeros(self.K * self.Np)
        y = np.zeros(self.K * self.Np)
        self.L.or_mesh_nodes(self.h, x, y)
        return x.reshape(self.K, self.Np), y.reshape(self.K, self.Np)

    def geo(self):
        g = np.zeros(self.K * 14)
        self.L.or_mesh_geo(self.h, g)
        return g.reshape(self.K, 14)

    def rhs_advect(self, q, ax, ay, bcq=None, average=False):
        """advection right-hand side with the upwind flux, or the central "average" flux"""
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros_like(q)
        if average:
            self.L.or_rhs_advect_flux(self.h, q.reshape(-1), ax, ay, _ptr(bcq), 1, out.reshape(-1))
        else:
            self.L.or_rhs_advect(self.h, q.reshape(-1), ax, ay, _ptr(bcq), out.reshape(-1))
        return out

    def rhs_euler(self, q, gamma=1.4, flux=1, bcq=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros_like(q)
        self.L.or_rhs_euler(self.h, q.reshape(-1), gamma, flux, _ptr(bcq), out.reshape(-1))
        return out

    def lserk4(self, eq, q, nsteps, dt, a0, a1=0.0, flux=1, bcq=None):
        q = np.array(q, dtype=np.float64, copy=True)
        self.L.or_lserk4(self.h, eq, q.reshape(-1), nsteps, dt, a0, a1, flux, _ptr(bcq))
        return q

    def ssprk2(self, eq, q, nsteps, dt, a0, a1=0.0, flux=1, bcq=None):
        q = np.array(q, dtype=np.float64, copy=True)
        self.L.or_ssprk2(self.h, eq, q.reshape(-1), nsteps, dt, a0, a1, flux, _ptr(bcq))
        return q

    def sipg_system(self, penalty, f, gbc=None, mass_shift=0.0):
        """Assembled reference SIPG matrix A (scipy CSR) and right-hand side b of dgm::laplacian(u) == f;
        gbc: [K,3,Nfp] Dirichlet nodal values (None: zero).  mass_shift sigma: A - sigma blockdiag(J_e M),
        the reference-sign form of the Helmholtz operator -laplacian + sigma."""
        import scipy.sparse as sp
        n = self.K * self.Np
        nnz = self.L.or_sipg_nnz(self.h)
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz)
        b = np.zeros(n)
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        g = None if gbc is None else np.ascontiguousarray(gbc, dtype=np.float64).reshape(-1)
        got = self.L.or_sipg_assemble(self.h, penalty, f, _ptr(g), rp, col, val, b)
        assert got == nnz
        A = sp.csr_matrix((val, col, rp), shape=(n, n))
        if mass_shift != 0.0:
            J = self.geo()[:, 4]
            Mref = refelem_ops(self.N)["M"]
            A = (A - mass_shift * sp.block_diag([J[e] * Mref for e in range(self.K)], format="csr")).tocsr()
        return A, b

    def tvb_limit(self, q, M=50.0, nu=1.5, bcq=None):
        """TVB-limited copy of q ([nvar][K][Np] or [K][Np]) and the troubled-cell flags."""
        q = np.array(q, dtype=np.float64, copy=True)
        nvar = q.size // (self.K * self.Np)
        tr = np.zeros(self.K, np.int32)
        g = None if bcq is None else np.ascontiguousarray(bcq, dtype=np.float64)
        self.L.or_tvb_limit(self.h, q.reshape(-1), nvar, M, nu, _ptr(g), tr)
        return q, tr.astype(bool)

    def triangle_limit(self, q, gamma=1.4, tol=1e-2, bcq=None, sel=None, patch_first=None):
        """Copy of the Euler state q [4][K][Np] rebuilt by the reference's Triangle limiter, and the per-cell
        flags (bit 0 rebuilt, bit 1 density floor, bit 2 pressure below tol).  patch_first [K][3]: the first face
        (e' 3 + f') of each fixedValue face's patch, whose first value is the ghost mean (None: the own face's)."""
        q = np.array(q, dtype=np.float64, copy=True)
        fl = np.zeros(self.K, np.int32)
        g = None if bcq is None else np.ascontiguousarray(bcq, dtype=np.float64)
        s = None if sel is None else np.ascontiguousarray(sel, dtype=np.int32)
        pf = None if patch_first is None else np.ascontiguousarray(patch_first, dtype=np.int32)
        self.L.or_triangle_limit_pf(self.h, q.reshape(-1), gamma, tol, _ptr(g), _ptr(s), _ptr(pf), fl)
        return q, fl

    def ldg_lift(self):
        L = np.zeros(self.Np * 3 * self.Nfp)
        self.L.or_ldg_lift(self.h, L)
        return L.reshape(self.Np, 3 * self.Nfp)

    def ldg_rhs(self, u, ax, ay, nu, tau=1.0, beta=(1.0, 1.0 / np.pi), with_q=False):
        """LDG advection-diffusion right-hand side of u [K][Np] (periodic meshes), and q [2][K][Np]"""
        u = np.ascontiguousarray(u, dtype=np.float64)
        rhs = np.zeros_like(u)
        q = np.zeros((2,) + u.shape) if with_q else None
        rc = self.L.or_ldg_rhs(self.h, u.reshape(-1), ax, ay, nu, tau, beta[0], beta[1], rhs.reshape(-1), _ptr(q))
        if rc != 0:
            raise ValueError("or_ldg_rhs: the mesh has boundary faces")
        return (rhs, q) if with_q else rhs

    def ldg_lserk4(self, u, nsteps, dt, ax, ay, nu, tau=1.0, beta=(1.0, 1.0 / np.pi)):
        u = np.array(u, dtype=np.float64, copy=True)
        rc = self.L.or_ldg_lserk4(self.h, u.reshape(-1), nsteps, dt, ax, ay, nu, tau, beta[0], beta[1])
        if rc != 0:
            raise ValueError("or_ldg_lserk4 failed")
        return u

    def positivity(self, q, gamma=1.4, eps=1e-10):
        """Zhang-Shu scaled copy of the Euler state q [4][K][Np] and the flags of the scaled cells."""
        q = np.array(q, dtype=np.float64, copy=True)
        sc = np.zeros(self.K, np.int32)
        self.L.or_positivity(self.h, q.reshape(-1), gamma, eps, sc)
        return q, sc.astype(bool)

    def positivity_points(self):
        n = self.Np + self.Nq + 3 * self.Nfq
        V = np.zeros(n * self.Np)
        self.L.or_positivity_points(self.h, V)
        return V.reshape(n, self.Np)

    def tvb_geometry(self):
        coef = np.zeros(self.K * 6)
        jsel = np.zeros(self.K * 3, np.int32)
        h = np.zeros(self.K)
        self.L.or_tvb_geometry(self.h, coef, jsel, h)
        return coef.reshape(self.K, 3, 2), jsel.reshape(self.K, 3), h

    def vortex_ref_run(self, dt, nsteps):
        er, eu = C.c_double(), C.c_double()
        self.L.or_vortex_ref_run(self.h, dt, nsteps, C.byref(er), C.byref(eu))
        return er.value, eu.value


def cg_jacobi(A, b, x0, rtol=1e-10, abstol=1e-50, maxit=10000, history=False):
    """PETSc-order point-Jacobi CG of the oracle on a scipy CSR matrix.
    Returns (x, iterations, rnorm0, rnorm), plus the preconditioned residual norms of the iterates
    0..iterations when history is set."""
    L = lib()
    A = A.tocsr()
    x = np.array(x0, dtype=np.float64, copy=True).reshape(-1)
    r0, rn = C.c_double(), C.c_double()
    hist = np.zeros(maxit + 1) if history else None
    its = L.or_cg_jacobi(A.shape[0], np.ascontiguousarray(A.indptr, np.int32),
                         np.ascontiguousarray(A.indices, np.int32), np.ascontiguousarray(A.data),
                         np.ascontiguousarray(b, dtype=np.float64), x, rtol, abstol, maxit, C.byref(r0), C.byref(rn),
                         _ptr(hist))
    if history:
        return x, its, r0.value, rn.value, hist[:its + 1].copy()
    return x, its, r0.value, rn.value


def tvb_ops(N):
    L = lib()
    Np = L.or_np(N)
    pmean, pmid, phi = np.zeros(Np), np.zeros(3 * Np), np.zeros(3 * Np)
    L.or_tvb_ops(N, pmean, pmid, phi)
    return pmean, pmid.reshape(3, Np), phi.reshape(Np, 3)


def refelem_ops(N):
    L = lib()
    Np, Nq, Nfp, Nfq = L.or_np(N), L.or_nq(N), N + 1, N + 2
    a = {k: np.zeros(n) for k, n in dict(r=Np, s=Np, V=Np * Np, invV=Np * Np, Dr=Np * Np, Ds=Np * Np,
                                        Vq=Nq * Np, Drq=Nq * Np, Dsq=Nq * Np, cw=Nq, Vf=Nfq * Nfp,
                                        gw=Nfq, M=Np * Np).items()}
    rc = L.or_refelem_ops(N, a["r"], a["s"], a["V"], a["invV"], a["Dr"], a["Ds"], a["Vq"], a["Drq"],
                          a["Dsq"], a["cw"], a["Vf"], a["gw"], a["M"])
    assert rc == 0
    shp = dict(V=(Np, Np), invV=(Np, Np), Dr=(Np, Np), Ds=(Np, Np), Vq=(Nq, Np), Drq=(Nq, Np),
               Dsq=(Nq, Np), Vf=(Nfq, Nfp), M=(Np, Np))
    for k, s in shp.items():
        a[k] = a[k].reshape(s)
    fm = np.zeros(3 * Nfp, np.int32)
    L.or_fmask(N, fm)
    a["fmask"] = fm.reshape(3, Nfp)
    return a

"""ctypes binding of libdghip.so (C ABI declared in include/dgh.h)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libdghip.so")

DGH_EQ_ADVECT, DGH_EQ_EULER = 0, 1
DGH_FLUX_UPWIND, DGH_FLUX_LF, DGH_FLUX_ROE, DGH_FLUX_AVERAGE = 0, 1, 2, 3
DGH_PREC_FP64, DGH_PREC_FP32, DGH_PREC_BF16X3 = 0, 1, 2
DGH_BC_DEFAULT, DGH_BC_FIXED_VALUE, DGH_BC_ZERO_GRADIENT, DGH_BC_WALL = 0, 1, 2, 3
DGH_SIPG_PENALTY_REFERENCE, DGH_SIPG_PENALTY_H = 0, 1
DGH_PC_JACOBI, DGH_PC_BLOCK_JACOBI = 0, 1
DGH_SOLVE_FP64, DGH_SOLVE_MIXED = 0, 1
DGH_ERR_NOT_CONVERGED = -5
DGH_LIM_TVB, DGH_LIM_MLP, DGH_LIM_ALL = 0, 1, 2
DGH_LIM_TRIANGLE, DGH_LIM_TRIANGLE_TVB, DGH_LIM_TRIANGLE_MLP = 3, 4, 5
DGH_MLP_FP32, DGH_MLP_BF16 = 0, 1
DGH_BCFN_NONE, DGH_BCFN_ISENTROPIC_VORTEX, DGH_BCFN_DOUBLE_MACH = 0, 1, 2

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class DghError(RuntimeError):
    pass


class RhsDesc(C.Structure):
    _fields_ = [("eq", C.c_int), ("flux", C.c_int), ("prec", C.c_int), ("reserved", C.c_int),
                ("ax", C.c_double), ("ay", C.c_double), ("gamma", C.c_double)]


class SipgDesc(C.Structure):
    _fields_ = [("penalty", C.c_int), ("pc", C.c_int), ("mode", C.c_int), ("maxit", C.c_int),
                ("rtol", C.c_double), ("abstol", C.c_double), ("inner_rtol", C.c_double),
                ("max_outer", C.c_int), ("reserved", C.c_int), ("mass_shift", C.c_double)]


class LimDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("detect_only", C.c_int), ("M", C.c_double), ("nu", C.c_double),
                ("mlp", C.c_int), ("positivity", C.c_int), ("threshold", C.c_double), ("gamma", C.c_double),
                ("eps", C.c_double)]


class MlpTrainDesc(C.Structure):
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_float), ("pos_weight", C.c_float),
                ("seed", C.c_uint), ("reserved", C.c_int)]


class EnsCase(C.Structure):
    _fields_ = [("N", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("x0", C.c_double), ("y0", C.c_double),
                ("Lx", C.c_double), ("Ly", C.c_double), ("ax", C.c_double), ("ay", C.c_double), ("nu", C.c_double),
                ("dt", C.c_double)]


class EnsDesc(C.Structure):
    _fields_ = [("tau", C.c_double), ("bx", C.c_double), ("by", C.c_double), ("autotune", C.c_int),
                ("reserved", C.c_int)]


class BcFnDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("reserved", C.c_int), ("gamma", C.c_double), ("x0", C.c_double),
                ("y0", C.c_double), ("beta", C.c_double), ("u_inf", C.c_double), ("v_inf", C.c_double)]


class SipgStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("outer", C.c_int), ("reason", C.c_int), ("reserved", C.c_int),
                ("rnorm0", C.c_double), ("rnorm", C.c_double), ("relres", C.c_double), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


def lib_path():
    return LIB


def build():
    """Compile libdghip.so for gfx950 in-tree (always recompiles what changed)."""
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise DghError("libdghip.so is not built: run __graft_entry__.build() (or make -C hopefoam_amd)")
    L = C.CDLL(LIB)
    vp = C.c_void_p
    sig = {
        "dgh_create": [C.POINTER(vp), C.c_int],
        "dgh_destroy": [vp],
        "dgh_sync": [vp],
        "dgh_mesh_create": [vp, C.c_int, C.c_int, _dp, _ip, vp, C.POINTER(C.c_int)],
        "dgh_mesh_structured": [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)],
        "dgh_mesh_info": [vp, C.c_int, _ip],
        "dgh_mesh_nodes": [vp, C.c_int, _dp, _dp],
        "dgh_mesh_conn": [vp, C.c_int, _ip, _ip],
        "dgh_mesh_geo": [vp, C.c_int, _dp],
        "dgh_mesh_set_bc_values": [vp, C.c_int, C.c_int, vp],
        "dgh_mesh_set_bc_function": [vp, C.c_int, C.POINTER(BcFnDesc), C.c_double],
        "dgh_mesh_get_time": [vp, C.c_int, C.POINTER(C.c_double)],
        "dgh_mesh_set_patch_first": [vp, C.c_int, vp],
        "dgh_set_launch_cap": [vp, C.c_int],
        "dgh_refelem_get": [C.c_int, C.c_char_p, _dp, C.c_int],
        "dgh_field_create": [vp, C.c_int, C.c_int, C.POINTER(C.c_int)],
        "dgh_field_destroy": [vp, C.c_int],
        "dgh_field_upload": [vp, C.c_int, _dp],
        "dgh_field_download": [vp, C.c_int, _dp],
        "dgh_field_copy": [vp, C.c_int, C.c_int],
        "dgh_rhs": [vp, C.c_int, C.POINTER(RhsDesc), C.c_int, C.c_int],
        "dgh_lserk4_stage": [vp, C.c_int, C.POINTER(RhsDesc), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double],
        "dgh_lserk4_run": [vp, C.c_int, C.POINTER(RhsDesc), C.c_int, C.c_int, C.c_double, C.POINTER(C.c_float)],
        "dgh_lserk4_run_limited": [vp, C.c_int, C.POINTER(RhsDesc), C.POINTER(LimDesc), C.c_int, C.c_int, C.c_double,
                                   C.POINTER(C.c_float)],
        "dgh_ssprk2_run": [vp, C.c_int, C.POINTER(RhsDesc), C.POINTER(LimDesc), C.c_int, C.c_int, C.c_double,
                           C.POINTER(C.c_float)],
        "dgh_ssprk2_stage": [vp, C.c_int, C.POINTER(RhsDesc), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double],
        "dgh_limit": [vp, C.c_int, C.POINTER(LimDesc), C.c_int, vp, C.POINTER(C.c_int)],
        "dgh_limit_noalloc": [vp, C.c_int, C.POINTER(LimDesc), C.c_int, vp, C.POINTER(C.c_int)],
        "dgh_mlp_create": [vp, C.c_int, C.c_uint, C.POINTER(C.c_int)],
        "dgh_mlp_destroy": [vp, C.c_int],
        "dgh_mlp_info": [vp, C.c_int, _ip],
        "dgh_mlp_get_params": [vp, C.c_int, _fp],
        "dgh_mlp_set_params": [vp, C.c_int, _fp],
        "dgh_mlp_features": [vp, C.c_int, C.c_int, C.c_double, _fp],
        "dgh_mlp_forward": [vp, C.c_int, _fp, C.c_int, C.c_int, _fp],
        "dgh_mlp_grad": [vp, C.c_int, _fp, _fp, C.c_int, C.c_float, _fp, C.POINTER(C.c_float)],
        "dgh_mlp_train": [vp, C.c_int, _fp, _fp, C.c_int, C.POINTER(MlpTrainDesc), vp],
        "dgh_sipg_create": [vp, C.c_int, C.POINTER(SipgDesc), C.POINTER(C.c_int)],
        "dgh_sipg_destroy": [vp, C.c_int],
        "dgh_sipg_info": [vp, C.c_int, _ip],
        "dgh_sipg_rhs": [vp, C.c_int, C.c_int, vp, C.c_int],
        "dgh_sipg_rhs_neumann": [vp, C.c_int, C.c_int, vp, vp, C.c_int],
        "dgh_sipg_apply": [vp, C.c_int, C.c_int, C.c_int, C.c_int],
        "dgh_ins_convection": [vp, C.c_int, C.c_int, C.c_int],
        "dgh_nodal_grad": [vp, C.c_int, C.c_int, C.c_int],
        "dgh_nodal_div": [vp, C.c_int, C.c_int, C.c_int],
        "dgh_field_lincomb": [vp, C.c_int, C.c_int, _ip, _dp],
        "dgh_field_copy_var": [vp, C.c_int, C.c_int, C.c_int, C.c_int],
        "dgh_sipg_solve": [vp, C.c_int, C.c_int, C.c_int, C.POINTER(SipgStats)],
        "dgh_ensemble_create": [vp, C.POINTER(EnsDesc), C.c_int, C.POINTER(EnsCase), C.POINTER(C.c_int)],
        "dgh_ensemble_destroy": [vp, C.c_int],
        "dgh_ensemble_set": [vp, C.c_int, C.c_int, _dp],
        "dgh_ensemble_get": [vp, C.c_int, C.c_int, C.c_int, _dp],
        "dgh_ensemble_rhs": [vp, C.c_int],
        "dgh_ensemble_run": [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)],
        "dgh_ensemble_info": [vp, C.c_int, vp, vp, vp, vp],
        "dgh_selftest_mfma": [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)],
        "dgh_selftest_mfma_layout": [vp, _dp],
        "dgh_selftest_mfma_bf16x3": [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.dgh_last_error.argtypes = [vp]
    L.dgh_last_error.restype = C.c_char_p
    L.dgh_stream.argtypes = [vp]
    L.dgh_stream.restype = vp
    _lib = L
    return L


def refelem_get(N, name, cap=1 << 16):
    out = np.zeros(cap)
    n = lib().dgh_refelem_get(N, name.encode(), out, cap)
    if n < 0:
        raise DghError("dgh_refelem_get(%d, %s) failed: %d" % (N, name, n))
    return out[:n].copy()


class Context:
    """One device context (owns a HIP stream, meshes and fields)."""

    def __init__(self, device=0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.dgh_create(C.byref(h), device)
        if rc != 0:
            raise DghError("dgh_create(%d) failed with %d (no usable GPU?)" % (device, rc))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.dgh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc, what):
        if rc != 0:
            msg = self.L.dgh_last_error(self.h)
            raise DghError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    # ---- meshes
    def mesh(self, N, vxy, vid, bctag=None):
        vxy = np.ascontiguousarray(vxy, dtype=np.float64).reshape(-1)
        vid = np.ascontiguousarray(vid, dtype=np.int32).reshape(-1)
        bct = None if bctag is None else np.ascontiguousarray(bctag, dtype=np.int32).reshape(-1)
        out = C.c_int()
        self._ck(self.L.dgh_mesh_create(self.h, N, vid.size // 3, vxy, vid,
                                        None if bct is None else bct.ctypes.data_as(C.c_void_p),
                                        C.byref(out)), "dgh_mesh_create")
        m = Mesh(self, out.value)
        m._keep = bct
        return m

    def mesh_structured(self, N, nx, ny, x0, x1, y0, y1, periodic=(True, True), bc_tag=DGH_BC_ZERO_GRADIENT):
        out = C.c_int()
        self._ck(self.L.dgh_mesh_structured(self.h, N, nx, ny, x0, x1, y0, y1, int(periodic[0]),
                                            int(periodic[1]), bc_tag, C.byref(out)), "dgh_mesh_structured")
        return Mesh(self, out.value)

    def sync(self):
        self._ck(self.L.dgh_sync(self.h), "dgh_sync")

    def set_launch_cap(self, blocks):
        """at most `blocks` persistent blocks per stage launch (0: the device's resident maximum)"""
        self._ck(self.L.dgh_set_launch_cap(self.h, int(blocks)), "dgh_set_launch_cap")

    def stream(self):
        return self.L.dgh_stream(self.h)

    def mlp(self, N, seed=0):
        out = C.c_int()
        self._ck(self.L.dgh_mlp_create(self.h, N, seed, C.byref(out)), "dgh_mlp_create")
        return Mlp(self, out.value)

    def ensemble(self, cases, tau=1.0, beta=(1.0, 1.0 / np.pi), autotune=True):
        """LDG advection-diffusion ensemble; cases: dicts with N, nx, ny, x0, y0, Lx, Ly, ax, ay, nu, dt"""
        arr = (EnsCase * len(cases))()
        for i, c in enumerate(cases):
            arr[i] = EnsCase(c["N"], c["nx"], c["ny"], c.get("x0", 0.0), c.get("y0", 0.0), c.get("Lx", 1.0),
                             c.get("Ly", 1.0), c["ax"], c["ay"], c["nu"], c["dt"])
        d = EnsDesc(tau, beta[0], beta[1], int(autotune), 0)
        out = C.c_int()
        self._ck(self.L.dgh_ensemble_create(self.h, C.byref(d), len(cases), arr, C.byref(out)), "dgh_ensemble_create")
        return Ensemble(self, out.value, cases)

    def mfma_layout(self):
        raw = np.zeros(1024)
        self._ck(self.L.dgh_selftest_mfma_layout(self.h, raw), "dgh_selftest_mfma_layout")
        return raw.reshape(4, 64, 4)

    def selftest_mfma(self):
        a, b = C.c_double(), C.c_double()
        self._ck(self.L.dgh_selftest_mfma(self.h, C.byref(a), C.byref(b)), "dgh_selftest_mfma")
        return a.value, b.value

    def selftest_bf16x3(self):
        """(bf16x3 error, hi-only bf16 error) of one 16x16x16 product, relative to its largest entry."""
        a, b = C.c_double(), C.c_double()
        self._ck(self.L.dgh_selftest_mfma_bf16x3(self.h, C.byref(a), C.byref(b)), "dgh_selftest_mfma_bf16x3")
        return a.value, b.value


class Mesh:
    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        info = np.zeros(7, np.int32)
        ctx._ck(ctx.L.dgh_mesh_info(ctx.h, handle, info), "dgh_mesh_info")
        self.N, self.Np, self.Nfp, self.Nq, self.Nfq, self.K, self.nbfaces = [int(v) for v in info]

    def nodes(self):
        x = np.zeros(self.K * self.Np)
        y = np.zeros(self.K * self.Np)
        self.ctx._ck(self.ctx.L.dgh_mesh_nodes(self.ctx.h, self.h, x, y), "dgh_mesh_nodes")
        return x.reshape(self.K, self.Np), y.reshape(self.K, self.Np)

    def conn(self):
        E = np.zeros(3 * self.K, np.int32)
        F = np.zeros(3 * self.K, np.int32)
        self.ctx._ck(self.ctx.L.dgh_mesh_conn(self.ctx.h, self.h, E, F), "dgh_mesh_conn")
        return E.reshape(self.K, 3), F.reshape(self.K, 3)

    def geo(self):
        g = np.zeros(self.K * 14)
        self.ctx._ck(self.ctx.L.dgh_mesh_geo(self.ctx.h, self.h, g), "dgh_mesh_geo")
        return g.reshape(self.K, 14)

    def set_bc_values(self, bcq):
        if bcq is None:
            self.ctx._ck(self.ctx.L.dgh_mesh_set_bc_values(self.ctx.h, self.h, 1, None), "set_bc_values")
            return
        bcq = np.ascontiguousarray(bcq, dtype=np.float64)
        nvar = bcq.size // (self.K * 3 * self.Nfp)
        self._bcq = bcq
        self.ctx._ck(self.ctx.L.dgh_mesh_set_bc_values(self.ctx.h, self.h, nvar, bcq.ctypes.data_as(C.c_void_p)),
                     "dgh_mesh_set_bc_values")

    def set_bc_function(self, kind, t=0.0, gamma=1.4, x0=5.0, y0=0.0, beta=5.0, u_inf=1.0, v_inf=0.0):
        """fixedValue ghost states from a function of time evaluated on the device (DGH_BCFN_*); the defaults
        are the isentropicVortex tutorial's"""
        d = BcFnDesc(kind, 0, gamma, x0, y0, beta, u_inf, v_inf)
        self.ctx._ck(self.ctx.L.dgh_mesh_set_bc_function(self.ctx.h, self.h, C.byref(d), t),
                     "dgh_mesh_set_bc_function")

    def set_patch_first(self, first):
        """[K][3] first face e' 3 + f' of each fixedValue face's patch (the Triangle ghost means take that face's
        first value, as the reference), -1 elsewhere; None clears"""
        if first is None:
            self.ctx._ck(self.ctx.L.dgh_mesh_set_patch_first(self.ctx.h, self.h, None), "dgh_mesh_set_patch_first")
            return
        a = np.ascontiguousarray(first, dtype=np.int32).reshape(-1)
        assert a.size == 3 * self.K
        self.ctx._ck(self.ctx.L.dgh_mesh_set_patch_first(self.ctx.h, self.h, a.ctypes.data_as(C.c_void_p)),
                     "dgh_mesh_set_patch_first")

    def time(self):
        t = C.c_double()
        self.ctx._ck(self.ctx.L.dgh_mesh_get_time(self.ctx.h, self.h, C.byref(t)), "dgh_mesh_get_time")
        return t.value

    def field(self, nvar, data=None):
        out = C.c_int()
        self.ctx._ck(self.ctx.L.dgh_field_create(self.ctx.h, self.h, nvar, C.byref(out)), "dgh_field_create")
        f = Field(self, out.value, nvar)
        if data is not None:
            f.upload(data)
        return f

    def rhs(self, desc, q, out):
        self.ctx._ck(self.ctx.L.dgh_rhs(self.ctx.h, self.h, C.byref(desc), q.h, out.h), "dgh_rhs")

    def ins_convection(self, U, out):
        """out = div(U U) of the 2-variable field U (dgc::div(U, U), LF, then the mass solve)"""
        self.ctx._ck(self.ctx.L.dgh_ins_convection(self.ctx.h, self.h, U.h, out.h), "dgh_ins_convection")

    def nodal_grad(self, u, out):
        self.ctx._ck(self.ctx.L.dgh_nodal_grad(self.ctx.h, self.h, u.h, out.h), "dgh_nodal_grad")

    def nodal_div(self, u, out):
        self.ctx._ck(self.ctx.L.dgh_nodal_div(self.ctx.h, self.h, u.h, out.h), "dgh_nodal_div")

    def lserk4_stage(self, desc, q_in, q_out, res, stage, dt):
        self.ctx._ck(self.ctx.L.dgh_lserk4_stage(self.ctx.h, self.h, C.byref(desc), q_in.h, q_out.h, res.h,
                                                 stage, dt), "dgh_lserk4_stage")

    @staticmethod
    def limiter_desc(M=50.0, nu=1.5, kind=DGH_LIM_TVB, mlp=None, threshold=0.5, positivity=False, gamma=1.4,
                     eps=1e-10):
        """the limiter descriptor of limit() for lserk4_run(limiter=...); mlp makes the learned indicator the
        detection step (of the TVB limiter, or of the Triangle scheme with kind=DGH_LIM_TRIANGLE)"""
        if mlp is not None:
            if kind in (DGH_LIM_TVB, DGH_LIM_MLP):
                kind = DGH_LIM_MLP
            elif kind in (DGH_LIM_TRIANGLE, DGH_LIM_TRIANGLE_MLP):
                kind = DGH_LIM_TRIANGLE_MLP
            else:
                raise ValueError("mlp= needs the TVB or Triangle limiter kind")
        return LimDesc(kind, 0, M, nu, -1 if mlp is None else mlp.h, int(positivity), threshold, gamma, eps)

    def limit(self, q, M=50.0, nu=1.5, kind=DGH_LIM_TVB, flags=True, detect_only=False, mlp=None, threshold=0.5,
              positivity=False, gamma=1.4, eps=1e-10, noalloc=False):
        """Limits field q in place (or only flags, detect_only); the learned indicator decides when mlp is
        given; positivity adds the Zhang-Shu scaling (Euler).  noalloc: through dgh_limit_noalloc (scratch in the
        library's static device array; the mesh's limiter tables must exist).  Returns (troubled flags [K] bool
        or None, number of troubled cells)."""
        if mlp is not None:
            kind = DGH_LIM_MLP
        d = LimDesc(kind, int(detect_only), M, nu, -1 if mlp is None else mlp.h, int(positivity), threshold, gamma,
                    eps)
        tr = np.zeros(self.K, np.int32) if flags else None
        n = C.c_int()
        name = "dgh_limit_noalloc" if noalloc else "dgh_limit"
        self.ctx._ck(getattr(self.ctx.L, name)(self.ctx.h, self.h, C.byref(d), q.h,
                                               None if tr is None else tr.ctypes.data_as(C.c_void_p), C.byref(n)),
                     name)
        return (None if tr is None else tr.astype(bool)), n.value

    def triangle_limit(self, q, gamma=1.4, tol=1e-2, detect=None, M=50.0, nu=1.5, mlp=None, threshold=0.5,
                       positivity=False, flags=True):
        """The reference's Triangle limiter on the 4-variable field q in place (Trianglelimite.C:61-852): every
        cell, or with detect="tvb" / an mlp only the cells the detection step flags.  Returns (flags [K] int:
        bit 0 rebuilt, bit 1 density floor used, bit 2 pressure below tol; number of rebuilt cells), or
        (None, None) with flags=False (the call is then asynchronous)."""
        kind = DGH_LIM_TRIANGLE
        if mlp is not None:
            kind = DGH_LIM_TRIANGLE_MLP
        elif detect == "tvb":
            kind = DGH_LIM_TRIANGLE_TVB
        elif detect is not None:
            raise ValueError("detect: None or 'tvb' (pass mlp= for the learned indicator)")
        d = LimDesc(kind, 0, M, nu, -1 if mlp is None else mlp.h, int(positivity), threshold, gamma, tol)
        if not flags:
            self.ctx._ck(self.ctx.L.dgh_limit(self.ctx.h, self.h, C.byref(d), q.h, None, None), "dgh_limit")
            return None, None
        fl = np.zeros(self.K, np.int32)
        n = C.c_int()
        self.ctx._ck(self.ctx.L.dgh_limit(self.ctx.h, self.h, C.byref(d), q.h, fl.ctypes.data_as(C.c_void_p),
                                          C.byref(n)), "dgh_limit")
        return fl, n.value

    def mlp_features(self, q, M=50.0):
        """MLP input features [K, 2 Np + 8] of the 4-variable field q (TVB constant M)."""
        X = np.zeros((self.K, 2 * self.Np + 8), np.float32)
        self.ctx._ck(self.ctx.L.dgh_mlp_features(self.ctx.h, self.h, q.h, M, X.reshape(-1)), "dgh_mlp_features")
        return X

    def sipg(self, penalty=DGH_SIPG_PENALTY_REFERENCE, pc=DGH_PC_JACOBI, mode=DGH_SOLVE_FP64, maxit=10000,
             rtol=1e-10, abstol=1e-50, inner_rtol=0.0, max_outer=0, mass_shift=0.0):
        """SIPG operator and PCG solver; mass_shift sigma > 0 solves (A + sigma M) u = b (Helmholtz)."""
        d = SipgDesc(penalty, pc, mode, maxit, rtol, abstol, inner_rtol, max_outer, 0, mass_shift)
        out = C.c_int()
        self.ctx._ck(self.ctx.L.dgh_sipg_create(self.ctx.h, self.h, C.byref(d), C.byref(out)), "dgh_sipg_create")
        return Sipg(self, out.value)

    def ssprk2_stage(self, desc, q_in, q_out, res, stage, dt):
        self.ctx._ck(self.ctx.L.dgh_ssprk2_stage(self.ctx.h, self.h, C.byref(desc), q_in.h, q_out.h, res.h,
                                                 stage, dt), "dgh_ssprk2_stage")

    def ssprk2_run(self, desc, q, nsteps, dt, limiter=None):
        """nsteps SSP-RK2 steps on q in place (the reference's Euler driver), limiter (a LimDesc) after each
        stage; returns the device time in ms"""
        ms = C.c_float()
        self.ctx._ck(self.ctx.L.dgh_ssprk2_run(self.ctx.h, self.h, C.byref(desc),
                                               None if limiter is None else C.byref(limiter), q.h, nsteps, dt,
                                               C.byref(ms)), "dgh_ssprk2_run")
        return ms.value

    def lserk4_run(self, desc, q, nsteps, dt, limiter=None):
        """nsteps LSERK4 steps on q; limiter: a LimDesc (see limiter_desc) applied to every stage's new state.
        Returns the device time of the loop in ms."""
        ms = C.c_float()
        if limiter is None:
            self.ctx._ck(self.ctx.L.dgh_lserk4_run(self.ctx.h, self.h, C.byref(desc), q.h, nsteps, dt, C.byref(ms)),
                         "dgh_lserk4_run")
        else:
            self.ctx._ck(self.ctx.L.dgh_lserk4_run_limited(self.ctx.h, self.h, C.byref(desc), C.byref(limiter), q.h,
                                                           nsteps, dt, C.byref(ms)), "dgh_lserk4_run_limited")
        return ms.value


class Field:
    def __init__(self, mesh, handle, nvar):
        self.mesh = mesh
        self.h = handle
        self.nvar = nvar
        self.shape = (nvar, mesh.K, mesh.Np) if nvar > 1 else (mesh.K, mesh.Np)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.nvar * self.mesh.K * self.mesh.Np, (a.shape, self.shape)
        c = self.mesh.ctx
        c._ck(c.L.dgh_field_upload(c.h, self.h, a.reshape(-1)), "dgh_field_upload")

    def download(self):
        out = np.zeros(self.nvar * self.mesh.K * self.mesh.Np)
        c = self.mesh.ctx
        c._ck(c.L.dgh_field_download(c.h, self.h, out), "dgh_field_download")
        return out.reshape(self.shape)

    def copy_from(self, other):
        c = self.mesh.ctx
        c._ck(c.L.dgh_field_copy(c.h, self.h, other.h), "dgh_field_copy")

    def lincomb(self, terms):
        """self = sum of coef * field over terms [(coef, field), ...]"""
        c = self.mesh.ctx
        h = np.array([f.h for _, f in terms], np.int32)
        k = np.array([float(a) for a, _ in terms])
        c._ck(c.L.dgh_field_lincomb(c.h, self.h, len(terms), h, k), "dgh_field_lincomb")

    def copy_var(self, dvar, src, svar):
        c = self.mesh.ctx
        c._ck(c.L.dgh_field_copy_var(c.h, self.h, dvar, src.h, svar), "dgh_field_copy_var")


class Sipg:
    """Matrix-free SIPG operator + PCG solver of dgm::laplacian(u) == f on one mesh."""

    def __init__(self, mesh, handle):
        self.mesh = mesh
        self.h = handle
        info = np.zeros(4, np.int32)
        c = mesh.ctx
        c._ck(c.L.dgh_sipg_info(c.h, handle, info), "dgh_sipg_info")
        self.nblocks = int(info[3])

    def rhs(self, f, b, gbc=None, gnb=None):
        """b = M f + Dirichlet terms of gbc - (face mass) gnb on the other boundary faces (fixedGradient data);
        gbc, gnb: [K][3][Nfp] nodal values or None"""
        c = self.mesh.ctx
        arrs = []
        for a in (gbc, gnb):
            g = None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            if g is not None:
                assert g.size == self.mesh.K * 3 * self.mesh.Nfp
            arrs.append(g)
        ptr = [None if g is None else g.ctypes.data_as(C.c_void_p) for g in arrs]
        if gnb is None:
            c._ck(c.L.dgh_sipg_rhs(c.h, self.h, f.h, ptr[0], b.h), "dgh_sipg_rhs")
        else:
            c._ck(c.L.dgh_sipg_rhs_neumann(c.h, self.h, f.h, ptr[0], ptr[1], b.h), "dgh_sipg_rhs_neumann")

    def apply(self, u, out, prec=DGH_PREC_FP64):
        c = self.mesh.ctx
        c._ck(c.L.dgh_sipg_apply(c.h, self.h, prec, u.h, out.h), "dgh_sipg_apply")

    def solve(self, b, u, check=True):
        """Solves A u = b from the initial guess in u; returns the solver statistics as a dict."""
        c = self.mesh.ctx
        st = SipgStats()
        rc = c.L.dgh_sipg_solve(c.h, self.h, b.h, u.h, C.byref(st))
        if rc != 0 and (check or rc != DGH_ERR_NOT_CONVERGED):
            c._ck(rc, "dgh_sipg_solve")
        return st.as_dict()

    def destroy(self):
        c = self.mesh.ctx
        if self.h is not None and c.h:
            c._ck(c.L.dgh_sipg_destroy(c.h, self.h), "dgh_sipg_destroy")
            self.h = None


class Mlp:
    """Troubled-cell network (inputs -> 64 -> 64 -> 1) on the device."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        info = np.zeros(3, np.int32)
        ctx._ck(ctx.L.dgh_mlp_info(ctx.h, handle, info), "dgh_mlp_info")
        self.N, self.nin, self.nparams = [int(v) for v in info]

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        self.ctx._ck(self.ctx.L.dgh_mlp_get_params(self.ctx.h, self.h, p), "dgh_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        self.ctx._ck(self.ctx.L.dgh_mlp_set_params(self.ctx.h, self.h, p), "dgh_mlp_set_params")

    def forward(self, X, prec=DGH_MLP_FP32):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n = X.shape[0]
        z = np.zeros(n, np.float32)
        self.ctx._ck(self.ctx.L.dgh_mlp_forward(self.ctx.h, self.h, X.reshape(-1), n, prec, z), "dgh_mlp_forward")
        return z

    def grad(self, X, y, pos_weight=1.0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_float()
        self.ctx._ck(self.ctx.L.dgh_mlp_grad(self.ctx.h, self.h, X.reshape(-1), y, X.shape[0], pos_weight, g,
                                             C.byref(loss)), "dgh_mlp_grad")
        return loss.value, g

    def train(self, X, y, epochs=10, batch=1024, lr=1e-3, pos_weight=1.0, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        d = MlpTrainDesc(epochs, batch, lr, pos_weight, seed, 0)
        losses = np.zeros(max(epochs, 1), np.float32)
        self.ctx._ck(self.ctx.L.dgh_mlp_train(self.ctx.h, self.h, X.reshape(-1), y, X.shape[0], C.byref(d),
                                              losses.ctypes.data_as(C.c_void_p)), "dgh_mlp_train")
        return losses[:epochs]


class Ensemble:
    """Independent LDG advection-diffusion cases advanced together (dgh_ensemble_*)."""

    def __init__(self, ctx, handle, cases):
        self.ctx = ctx
        self.h = handle
        self.cases = [dict(c) for c in cases]
        for c in self.cases:
            c["K"] = 2 * c["nx"] * c["ny"]
            c["Np"] = (c["N"] + 1) * (c["N"] + 2) // 2

    def set(self, i, u):
        c = self.cases[i]
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        assert u.size == c["K"] * c["Np"]
        self.ctx._ck(self.ctx.L.dgh_ensemble_set(self.ctx.h, self.h, i, u), "dgh_ensemble_set")

    def get(self, i, what=0):
        c = self.cases[i]
        n = c["K"] * c["Np"] * (2 if what == 2 else 1)
        out = np.zeros(n)
        self.ctx._ck(self.ctx.L.dgh_ensemble_get(self.ctx.h, self.h, i, what, out), "dgh_ensemble_get")
        return out.reshape((2, c["K"], c["Np"]) if what == 2 else (c["K"], c["Np"]))

    def rhs(self):
        self.ctx._ck(self.ctx.L.dgh_ensemble_rhs(self.ctx.h, self.h), "dgh_ensemble_rhs")

    def run(self, nsteps, graph=True):
        ms = C.c_float()
        self.ctx._ck(self.ctx.L.dgh_ensemble_run(self.ctx.h, self.h, nsteps, int(graph), C.byref(ms)),
                     "dgh_ensemble_run")
        return ms.value

    def info(self):
        split = np.zeros(9, np.int32)
        tune = np.zeros(45)
        dofs = C.c_longlong()
        nitems = C.c_int()
        self.ctx._ck(self.ctx.L.dgh_ensemble_info(self.ctx.h, self.h, split.ctypes.data_as(C.c_void_p),
                                                  tune.ctypes.data_as(C.c_void_p), C.byref(dofs), C.byref(nitems)),
                     "dgh_ensemble_info")
        return {"split": split, "tune_us": tune.reshape(9, 5), "dofs": dofs.value, "nitems": nitems.value}

    def close(self):
        if getattr(self, "h", None) is not None and self.ctx.h:
            self.ctx.L.dgh_ensemble_destroy(self.ctx.h, self.h)
            self.h = None

"""Parity of the HIP path (through the C ABI) with the oracle on the same inputs.

The reference is the oracle built with long-double accumulation: the FP64 build reproduces the
reference's own summation round-off, which alone is ~1e-12 relative on the N=6 Euler RHS.

Bars: relative L2 difference <= 1e-12 per RHS evaluation and
<= 1e-10 after 100 LSERK4 steps for the FP64 path.  The FP32 path is checked against
a loose bound that its round-off allows; the bf16x3 path (split products on the bf16 matrix
cores, about 2^-16 relative per product) against bounds of its own, measured on MI355X with
a margin of about 4x: the hi-only bf16 product is ~400x worse than bf16x3, so a lost lo term
fails every one of them.
"""
import numpy as np
import pytest

from hopefoam_amd import (DGH_BC_FIXED_VALUE, DGH_BC_WALL, DGH_BC_ZERO_GRADIENT, DGH_EQ_ADVECT,
                          DGH_EQ_EULER, DGH_FLUX_AVERAGE, DGH_FLUX_LF, DGH_FLUX_ROE, DGH_FLUX_UPWIND, DGH_PREC_BF16X3,
                          DGH_PREC_FP32, DGH_PREC_FP64, Context, RhsDesc)
from hopefoam_amd.cases import gaussian_pulse, rel_l2, stable_dt, vortex, vortex_periodic
from meshgen import load_fixture_mesh, split_quads
from oracle_lib import OracleMesh, refelem_ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def test_mfma_layout(ctx):
    e64, e32 = ctx.selftest_mfma()
    print("mfma probe err", e64, e32, flush=True)
    assert e64 < 1e-12 and e32 < 1e-3


def test_mfma_bf16x3_probe(ctx):
    """One 16x16x16 product through the bf16x3 fragments (host-split A, device-split B): the error is at the
    2^-16 level of the split, and far below the hi-only bf16 product's (which the lo terms correct)."""
    e3, e1 = ctx.selftest_bf16x3()
    print("bf16x3 probe rel err %.2e, hi-only %.2e" % (e3, e1), flush=True)
    assert e3 < 6e-5
    assert e1 > 1e-3 and e1 > 50 * e3


def rounding_sensitivity(om, q, flux, u, shifted=False, seed=7):
    """Relative change of the oracle's Euler right-hand side when every nodal value of the state takes one random
    rounding-sized perturbation: u |q| (one rounding of each operand of the FP32 arithmetic) or, shifted, u |q - q0|
    with q0 the element's node-0 state (one hi/lo split of each operand of the bf16x3 contractions, which see the
    state relative to q0).  The error models of the reduced-precision kernels are multiples of this."""
    ref = om.rhs_euler(q, 1.4, flux)
    xi = np.random.default_rng(seed).standard_normal(q.shape)
    scale = np.abs(q - q[..., :1]) if shifted else np.abs(q)
    return rel_l2(om.rhs_euler(q + u * scale * xi, 1.4, flux), ref)


def fp32_rhs_model(om, q, flux, Np):
    """FP32 kernel: the state's conversion, about 10 roundings in the flux evaluation and the FP32 accumulation of
    the interpolation and projection products (at most Np terms, each rounding at most one partial sum) -- about
    12 + Np rounding-sized perturbations of the operands, independent: sqrt(12 + Np) times one, times 2 for the
    alignment of the largest ones."""
    return 2.0 * np.sqrt(12.0 + Np) * rounding_sensitivity(om, q, flux, 2.0 ** -24)


def bf16x3_rhs_model(om, q, flux):
    """bf16x3 kernel: the three split products drop lo x lo and round each operand's low part to bf16, a relative
    error of at most 2^-17 of the operand; the split operands along the chain are the state (less q0), the fluxes
    at the cubature points (less F(q0)) and the operator matrices: three independent perturbations of that size,
    sqrt(3) times one, times 2 for the alignment of the largest ones (the FP32 accumulation is 2^7 finer)."""
    return 2.0 * np.sqrt(3.0) * rounding_sensitivity(om, q, flux, 2.0 ** -17, shifted=True)


def _pair(ctx, N, vxy, vid, bct=None):
    # the parity reference: the oracle with long-double accumulation (the reference algorithm's
    # result to well below FP64 round-off)
    om = OracleMesh(N, vxy, vid, bct, extended=True)
    gm = ctx.mesh(N, vxy, vid, bct)
    return om, gm


def test_mesh_matches_oracle(ctx):
    vxy, vid = split_quads(7, 5, -1.0, 1.0, 0.0, 2.0, periodic=(True, False))
    bct = np.full(vid.shape, DGH_BC_ZERO_GRADIENT, np.int32)
    om, gm = _pair(ctx, 5, vxy, vid, bct)
    E1, F1 = om.conn()
    E2, F2 = gm.conn()
    assert (E1 == E2).all() and (F1 == F2).all()
    assert np.allclose(om.geo(), gm.geo(), rtol=1e-14, atol=1e-14)
    x1, y1 = om.nodes()
    x2, y2 = gm.nodes()
    assert np.abs(x1 - x2).max() < 1e-14 and np.abs(y1 - y2).max() < 1e-14


@pytest.mark.parametrize("N", [1, 2, 4, 6, 8])
def test_advection_rhs_parity(ctx, N):
    """C1-style: periodic split-quad mesh, constant velocity (1, 0.5), upwind (reference LF)."""
    n = 32 if N == 4 else 9
    vxy, vid = split_quads(n, n, -1.0, 1.0, -1.0, 1.0)
    om, gm = _pair(ctx, N, vxy, vid)
    x, y = om.nodes()
    q = gaussian_pulse(x, y) + 0.1 * np.sin(np.pi * x) * np.cos(np.pi * y)
    ref = om.rhs_advect(q, 1.0, 0.5)
    d = RhsDesc(DGH_EQ_ADVECT, DGH_FLUX_UPWIND, DGH_PREC_FP64, 0, 1.0, 0.5, 0.0)
    fq, fr = gm.field(1, q), gm.field(1)
    gm.rhs(d, fq, fr)
    got = fr.download()
    err = rel_l2(got, ref)
    print("advection N=%d K=%d rel L2 %.3e" % (N, gm.K, err), flush=True)
    assert err <= 1e-12
    # FP32 contractions: bounded by single-precision round-off
    d32 = RhsDesc(DGH_EQ_ADVECT, DGH_FLUX_UPWIND, DGH_PREC_FP32, 0, 1.0, 0.5, 0.0)
    gm.rhs(d32, fq, fr)
    assert rel_l2(fr.download(), ref) < 2e-5
    # bf16x3 split products (measured 2e-5 .. 1.3e-4 for N = 1..8)
    d3 = RhsDesc(DGH_EQ_ADVECT, DGH_FLUX_UPWIND, DGH_PREC_BF16X3, 0, 1.0, 0.5, 0.0)
    gm.rhs(d3, fq, fr)
    e3 = rel_l2(fr.download(), ref)
    print("advection N=%d bf16x3 rel L2 %.3e" % (N, e3), flush=True)
    assert e3 < 5e-4


@pytest.mark.parametrize("N", [1, 4, 6])
def test_average_flux_gradient_parity(ctx, N):
    """The reference's "average" flux (averageFlux.C:45-90) with a = (1, 0) and (0, 1) on a mesh with
    zero-gradient sides: the right-hand sides are minus the DG x and y derivatives with central fluxes (the
    projection solver's grad(p) / div(U)), against the oracle; plus the exact derivative of a degree-N
    polynomial with exact (fixedValue) boundary traces."""
    vxy, vid = split_quads(7, 6, -1.0, 1.0, -1.0, 0.5, periodic=(False, False))
    bct = np.full(vid.shape, DGH_BC_ZERO_GRADIENT, np.int32)
    om, gm = _pair(ctx, N, vxy, vid, bct)
    x, y = om.nodes()
    q = gaussian_pulse(x, y) + 0.1 * np.sin(np.pi * x) * np.cos(np.pi * y)
    fq, fr = gm.field(1, q), gm.field(1)
    for ax, ay in ((1.0, 0.0), (0.0, 1.0)):
        ref = om.rhs_advect(q, ax, ay, average=True)
        gm.rhs(RhsDesc(DGH_EQ_ADVECT, DGH_FLUX_AVERAGE, DGH_PREC_FP64, 0, ax, ay, 0.0), fq, fr)
        err = rel_l2(fr.download(), ref)
        print("average flux N=%d a=(%g,%g) rel L2 %.3e" % (N, ax, ay, err), flush=True)
        assert err <= 1e-12
    # exact derivatives of a polynomial of degree N with its exact boundary traces
    bct1 = np.ones(vid.shape, np.int32)
    om1, gm1 = _pair(ctx, N, vxy, vid, bct1)
    p = x ** N - 3 * x * y ** (N - 1) + 0.5 * y
    gx = N * x ** (N - 1) - 3 * y ** (N - 1)
    fm = refelem_ops(N)["fmask"]
    gm1.set_bc_values(np.ascontiguousarray(np.stack([p[:, fm[f]] for f in range(3)], axis=1))[None])
    fp, fo = gm1.field(1, p), gm1.field(1)
    gm1.rhs(RhsDesc(DGH_EQ_ADVECT, DGH_FLUX_AVERAGE, DGH_PREC_FP64, 0, 1.0, 0.0, 0.0), fp, fo)
    assert np.abs(fo.download() + gx).max() < 1e-10 * max(1.0, np.abs(gx).max())


def test_advection_lserk4_100_steps(ctx):
    """The advection benchmark: 32x32x2 periodic, N=4, a=(1,0.5), upwind, Gaussian, CFL 0.5, 100 steps."""
    vxy, vid = split_quads(32, 32, -1.0, 1.0, -1.0, 1.0)
    om, gm = _pair(ctx, 4, vxy, vid)
    x, y = om.nodes()
    q0 = gaussian_pulse(x, y)
    dt = stable_dt(4, 2.0 / 32, 2.0 / 32, np.hypot(1.0, 0.5), 0.5)
    ref = om.lserk4(0, q0, 100, dt, 1.0, 0.5)
    d = RhsDesc(DGH_EQ_ADVECT, DGH_FLUX_UPWIND, DGH_PREC_FP64, 0, 1.0, 0.5, 0.0)
    fq = gm.field(1, q0)
    gm.lserk4_run(d, fq, 100, dt)
    got = fq.download()
    err = rel_l2(got, ref)
    print("C1 100-step rel L2 %.3e (dt %.6g)" % (err, dt), flush=True)
    assert err <= 1e-10
    # the stage-by-stage entry point gives the same answer
    fa, fb, fr = gm.field(1, q0), gm.field(1), gm.field(1)
    cur, nxt = fa, fb
    for step in range(3):
        for s in range(5):
            gm.lserk4_stage(d, cur, nxt, fr, s, dt)
            cur, nxt = nxt, cur
    ref3 = om.lserk4(0, q0, 3, dt, 1.0, 0.5)
    assert rel_l2(cur.download(), ref3) <= 1e-12


@pytest.mark.parametrize("flux", [DGH_FLUX_LF, DGH_FLUX_ROE])
@pytest.mark.parametrize("N", [3, 6])
def test_euler_rhs_parity_periodic_vortex(ctx, N, flux):
    """The Euler vortex benchmark's physics on a reduced mesh: periodic [0,10]^2 vortex, strength 5."""
    n = 10
    vxy, vid = split_quads(n, n, 0.0, 10.0, 0.0, 10.0)
    om, gm = _pair(ctx, N, vxy, vid)
    x, y = om.nodes()
    q = vortex_periodic(x, y, 0.0)
    ref = om.rhs_euler(q, 1.4, 0 if flux == DGH_FLUX_ROE else 1)
    d = RhsDesc(DGH_EQ_EULER, flux, DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    fq, fr = gm.field(4, q), gm.field(4)
    gm.rhs(d, fq, fr)
    err = rel_l2(fr.download(), ref)
    print("euler N=%d flux=%d rel L2 %.3e" % (N, flux, err), flush=True)
    assert err <= 1e-12
    d32 = RhsDesc(DGH_EQ_EULER, flux, DGH_PREC_FP32, 0, 0.0, 0.0, 1.4)
    gm.rhs(d32, fq, fr)
    e32 = rel_l2(fr.download(), ref)
    print("euler N=%d flux=%d FP32 rel L2 %.3e" % (N, flux, e32), flush=True)
    m32 = fp32_rhs_model(om, q, 0 if flux == DGH_FLUX_ROE else 1, (N + 1) * (N + 2) // 2)
    print("  FP32 error model %.3e" % m32, flush=True)
    assert e32 <= m32
    d3 = RhsDesc(DGH_EQ_EULER, flux, DGH_PREC_BF16X3, 0, 0.0, 0.0, 1.4)
    gm.rhs(d3, fq, fr)
    e3 = rel_l2(fr.download(), ref)
    m3 = bf16x3_rhs_model(om, q, 0 if flux == DGH_FLUX_ROE else 1)
    print("euler N=%d flux=%d bf16x3 rel L2 %.3e, error model %.3e" % (N, flux, e3, m3), flush=True)
    assert e3 <= m3


@pytest.mark.parametrize("N", [3, 6, 8])
def test_bf16x3_preserves_uniform_flow(ctx, N):
    """A uniform state has a zero right-hand side.  The bf16x3 kernel contracts the state and the fluxes
    relative to each element's constant state, so the split products see exact zeros here and only the
    FP32 flux arithmetic remains; unshifted, the split error (2^-16 of the flux times the projection's
    amplification) would be visible."""
    vxy, vid = split_quads(6, 6, 0.0, 10.0, 0.0, 10.0)
    gm = ctx.mesh(N, vxy, vid)
    x, y = gm.nodes()
    rho, u, v, p = 1.3, 0.9, -0.4, 2.0
    q = np.stack([np.full_like(x, rho), np.full_like(x, rho * u), np.full_like(x, rho * v),
                  np.full_like(x, p / 0.4 + 0.5 * rho * (u * u + v * v))])
    fq, fr = gm.field(4, q), gm.field(4)
    scale = np.abs(q).max() * (np.hypot(u, v) + np.sqrt(1.4 * p / rho)) / (10.0 / 6)
    # measured: FP64 2e-13 .. 5e-12, bf16x3 3e-6 .. 6.5e-6 (N = 3 .. 8)
    for prec, bar in ((DGH_PREC_FP64, 2e-11), (DGH_PREC_BF16X3, 2e-5)):
        gm.rhs(RhsDesc(DGH_EQ_EULER, DGH_FLUX_LF, prec, 0, 0.0, 0.0, 1.4), fq, fr)
        r = np.abs(fr.download()).max() / scale
        print("uniform flow N=%d prec=%d max |rhs| / (|q| c / h) = %.2e" % (N, prec, r), flush=True)
        assert r < bar


def test_euler_roe_fixed_value_vortex_mesh(ctx):
    """The reference's own vortex mesh and boundary treatment: vortex1024.msh, N=4, Roe,
    fixedValue exact nodal boundary values (setBoundaryValues.H)."""
    vxy, tri = load_fixture_mesh("vortex1024")
    bct = np.full(tri.shape, DGH_BC_FIXED_VALUE, np.int32)
    om, gm = _pair(ctx, 4, vxy, tri, bct)
    x, y = om.nodes()
    q = vortex(x, y, 0.3, x0=5.0, y0=0.0)
    fm = refelem_ops(4)["fmask"]
    bcq = np.zeros((4, om.K, 3, 5))
    for f in range(3):
        bcq[:, :, f, :] = vortex(x[:, fm[f]], y[:, fm[f]], 0.3, x0=5.0, y0=0.0)
    ref = om.rhs_euler(q, 1.4, 0, bcq)
    gm.set_bc_values(bcq)
    d = RhsDesc(DGH_EQ_EULER, DGH_FLUX_ROE, DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    fq, fr = gm.field(4, q), gm.field(4)
    gm.rhs(d, fq, fr)
    err = rel_l2(fr.download(), ref)
    print("vortex1024 Roe fixedValue rel L2 %.3e" % err, flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("tag", [DGH_BC_ZERO_GRADIENT, DGH_BC_WALL])
def test_euler_rhs_parity_open_boundaries(ctx, tag):
    """The Riemann-problem benchmark's shape on a reduced mesh: [0,2]x[0,0.5], Roe, N=3, piecewise-constant
    Riemann data smoothed slightly, non-periodic boundary faces."""
    vxy, vid = split_quads(16, 4, 0.0, 2.0, 0.0, 0.5, periodic=(False, False))
    bct = np.full(vid.shape, tag, np.int32)
    om, gm = _pair(ctx, 3, vxy, vid, bct)
    x, y = om.nodes()
    s = 0.5 * (1 + np.tanh((x - 1.0) / 0.1)) * 0.5 * (1 + np.tanh((y - 0.25) / 0.05))
    rho = 1.0 + 0.8 * s
    u = 0.3 - 0.5 * s
    v = 0.1 * s
    p = 1.0 + s
    q = np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])
    ref = om.rhs_euler(q, 1.4, 0)
    d = RhsDesc(DGH_EQ_EULER, DGH_FLUX_ROE, DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    fq, fr = gm.field(4, q), gm.field(4)
    gm.rhs(d, fq, fr)
    err = rel_l2(fr.download(), ref)
    print("open-boundary tag %d rel L2 %.3e" % (tag, err), flush=True)
    if tag == DGH_BC_ZERO_GRADIENT:
        assert err <= 1e-12
    else:
        # the oracle has no wall treatment for bctag 3 on the Euler path other than mirroring,
        # which it applies at the Gauss points exactly like the kernel
        assert err <= 1e-12


def test_euler_lserk4_steps_parity(ctx):
    vxy, vid = split_quads(8, 8, 0.0, 10.0, 0.0, 10.0)
    om, gm = _pair(ctx, 6, vxy, vid)
    x, y = om.nodes()
    q0 = vortex_periodic(x, y, 0.0)
    dt = stable_dt(6, 10 / 8, 10 / 8, 2.5, 0.4)
    ref = om.lserk4(1, q0, 20, dt, 1.4, 0.0, 1)
    d = RhsDesc(DGH_EQ_EULER, DGH_FLUX_LF, DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    fq = gm.field(4, q0)
    gm.lserk4_run(d, fq, 20, dt)
    err = rel_l2(fq.download(), ref)
    print("euler 20 LSERK4 steps rel L2 %.3e" % err, flush=True)
    assert err <= 1e-10
    # bf16x3 contractions with FP64 RK registers, 20 steps.  Error model: the 100 stage evaluations' right-hand
    # side errors add up at most coherently (the state moves little in 20 steps), each weighted by dt through
    # the stable scheme: |dq| <= 20 dt |d rhs|, with |d rhs| the kernel's own right-hand-side error at the
    # initial state (itself within bf16x3_rhs_model) and 1.5 for its change along the run.
    d3 = RhsDesc(DGH_EQ_EULER, DGH_FLUX_LF, DGH_PREC_BF16X3, 0, 0.0, 0.0, 1.4)
    fr, f3 = gm.field(4), gm.field(4, q0)
    gm.rhs(d3, f3, fr)
    r0 = om.rhs_euler(q0, 1.4, 1)
    drhs = np.linalg.norm(fr.download() - r0)
    assert drhs <= bf16x3_rhs_model(om, q0, 1) * np.linalg.norm(r0)
    gm.lserk4_run(d3, f3, 20, dt)
    e3 = rel_l2(f3.download(), ref)
    bound = 1.5 * 20 * dt * drhs / np.linalg.norm(q0)
    print("euler 20 LSERK4 steps bf16x3 rel L2 %.3e, error model %.3e" % (e3, bound), flush=True)
    assert e3 <= bound


def test_ragged_element_count(ctx):
    """K not a multiple of the 16-element wave tile, including K < 16."""
    for nx, ny in ((3, 1), (5, 3), (11, 7)):
        vxy, vid = split_quads(nx, ny, 0.0, 1.0, 0.0, 1.0, periodic=(nx >= 3, ny >= 3))
        bct = np.full(vid.shape, DGH_BC_ZERO_GRADIENT, np.int32)
        om, gm = _pair(ctx, 2, vxy, vid, bct)
        x, y = om.nodes()
        q = np.sin(3 * x) + y
        ref = om.rhs_advect(q, -0.7, 0.4)
        d = RhsDesc(DGH_EQ_ADVECT, DGH_FLUX_UPWIND, DGH_PREC_FP64, 0, -0.7, 0.4, 0.0)
        fq, fr = gm.field(1, q), gm.field(1)
        gm.rhs(d, fq, fr)
        assert rel_l2(fr.download(), ref) <= 1e-12


def test_mfma_accumulator_layout_decoded(ctx):
    """Decode, per (lane, register), the row and column of the 16x16x4 accumulators and compare
    with the row map the kernels and fragment builders use (mfma_row in dg_common.hpp)."""
    raw = ctx.mfma_layout()
    for blk, (ra, rb), rowmap in ((0, (0, 1), lambda g, r: g + 4 * r), (1, (2, 3), lambda g, r: 4 * g + r)):
        rows = (raw[ra] - 96) / 4
        cols = (raw[rb] - 96) / 4
        for lane in range(64):
            for r in range(4):
                assert cols[lane, r] == lane % 16, ("col", blk, lane, r, cols[lane, r])
                assert rows[lane, r] == rowmap(lane // 16, r), ("row", blk, lane, r, rows[lane].tolist())


@pytest.mark.parametrize("N", [3, 6])
def test_persistent_loop_several_groups_per_wave(ctx, N):
    """Every wavefront walks several element groups: the stage kernel's grid capped at 8 blocks on K = 2046
    (128 groups, the last one partial; 4 groups per wavefront with 4-wave blocks, 2 with 8-wave blocks).  This
    runs the loop-carried parts of the kernel -- the LDS DMA of the next group's q_in, the next group's
    connectivity and geometric factors, the waits across iterations -- against the oracle: FP64 to the RHS bar
    and one LSERK4 step (five fused stages) to 1e-12, and every precision bit-identical to the uncapped grid
    (a group's arithmetic does not depend on which wavefront runs it)."""
    vxy, vid = split_quads(31, 33, 0.0, 10.0, 0.0, 10.0)
    om, gm = _pair(ctx, N, vxy, vid)
    x, y = om.nodes()
    q = vortex_periodic(x, y, 0.0)
    dt = stable_dt(N, 10 / 31, 10 / 33, 2.5, 0.4)
    ref = om.rhs_euler(q, 1.4, 1)
    ref_step = om.lserk4(1, q, 1, dt, 1.4, 0.0, 1)
    fq, fr = gm.field(4, q), gm.field(4)
    try:
        for prec, bar in ((DGH_PREC_FP64, 1e-12), (DGH_PREC_FP32, 5e-3), (DGH_PREC_BF16X3, 2e-3)):
            d = RhsDesc(DGH_EQ_EULER, DGH_FLUX_LF, prec, 0, 0.0, 0.0, 1.4)
            out = {}
            for cap in (0, 8):
                ctx.set_launch_cap(cap)
                gm.rhs(d, fq, fr)
                fs = gm.field(4, q)
                gm.lserk4_run(d, fs, 1, dt)
                out[cap] = (fr.download(), fs.download())
            ctx.set_launch_cap(0)
            e_rhs = rel_l2(out[8][0], ref)
            e_step = rel_l2(out[8][1], ref_step)
            print("N=%d prec=%d capped grid: rhs rel L2 %.3e, one step %.3e" % (N, prec, e_rhs, e_step), flush=True)
            assert np.array_equal(out[0][0], out[8][0]) and np.array_equal(out[0][1], out[8][1])
            assert e_rhs <= bar
            if prec == DGH_PREC_FP64:
                assert e_step <= 1e-12
    finally:
        ctx.set_launch_cap(0)


def test_persistent_loop_full_grid_fp64(ctx):
    """The uncapped FP64 N = 6 grid (one 4-wave block per CU) on K = 32768: 2048 element groups, two per
    wavefront, against the oracle's right-hand side.

    On this mesh the reference algorithm evaluated in FP64 (the oracle without long-double accumulation) is itself
    2.1e-12 away from the long-double result in relative L2: the vortex's right-hand side vanishes outside its
    core, where the volume and face terms (~|F|/h each) cancel to round-off, so the ratio grows as 1/h (5.3e-13 at
    K = 2048, 1.1e-12 at 8192).  The bar is therefore 1e-12 or twice the reference's own FP64 round-off,
    whichever is larger; and the same launch with the grid capped to 8 blocks (64 groups per wavefront) must give
    the same bits, since no group's arithmetic depends on which wavefront runs it."""
    vxy, vid = split_quads(128, 128, 0.0, 10.0, 0.0, 10.0)
    om, gm = _pair(ctx, 6, vxy, vid)
    x, y = om.nodes()
    q = vortex_periodic(x, y, 0.0)
    ref = om.rhs_euler(q, 1.4, 1)
    ref64 = OracleMesh(6, vxy, vid).rhs_euler(q, 1.4, 1)
    env = rel_l2(ref64, ref)
    d = RhsDesc(DGH_EQ_EULER, DGH_FLUX_LF, DGH_PREC_FP64, 0, 0.0, 0.0, 1.4)
    fq, fr, fc = gm.field(4, q), gm.field(4), gm.field(4)
    gm.rhs(d, fq, fr)
    try:
        ctx.set_launch_cap(8)
        gm.rhs(d, fq, fc)
    finally:
        ctx.set_launch_cap(0)
    got = fr.download()
    err = rel_l2(got, ref)
    print("K=32768 N=6 FP64 rhs rel L2 %.3e (reference algorithm in FP64: %.3e)" % (err, env), flush=True)
    assert np.array_equal(got, fc.download())
    assert err <= max(1e-12, 2.0 * env)

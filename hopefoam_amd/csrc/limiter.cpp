// Host side of the TVB minmod limiter behind dgh_limit (include/dgh.h): per-mesh limiter geometry and
// the launch of the two limiter kernels (limiter.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "dg_common.hpp"
#include "dgh.h"
#include "dgh_internal.hpp"
#include "limiter.hpp"
#include "triangle.hpp"

using namespace dg;

namespace dg {

struct LimDev {
    int device = 0;
    double *lgeo = nullptr, *p4 = nullptr, *phi = nullptr, *wbc = nullptr, *mom = nullptr, *invV = nullptr;
    double* vpos = nullptr;
    double *p4f = nullptr, *cff = nullptr;   // MFMA fragments of the moment and feature operators
    int npos = 0;
    int *jsel = nullptr, *troubled = nullptr;
    unsigned* count = nullptr;
    int mom_nvar = 0;
    // Triangle limiter: geometry, node offsets and the per-cell work arrays (built on first use)
    double *tgeo = nullptr, *rsn = nullptr, *tcell = nullptr, *tcv = nullptr, *tvg = nullptr;
    ~LimDev() {
        (void)hipSetDevice(device);
        for (void* p : {(void*)lgeo, (void*)p4, (void*)phi, (void*)wbc, (void*)mom, (void*)invV, (void*)vpos,
                        (void*)jsel, (void*)troubled, (void*)count, (void*)tgeo, (void*)rsn, (void*)tcell,
                        (void*)tcv, (void*)tvg, (void*)p4f, (void*)cff})
            if (p) (void)hipFree(p);
    }
};

}  // namespace dg

bool dg_mlp_infer(dgh_ctx c, int h, int N, double threshold, MlpInfer& w);   // mlp.cpp

// Barycentres of the neighbours through the shared edge (b_i = m_i + b_n - m_n, periodic-safe), ghosts
// mirrored through boundary edges, and for each edge the neighbour pair with non-negative weights.
static void limiter_geometry(const HostMesh& h, std::vector<double>& coef, std::vector<int>& jsel,
                             std::vector<double>& hlen) {
    const int K = h.K;
    coef.assign(size_t(K) * 6, 0.0);
    jsel.assign(size_t(K) * 3, 0);
    hlen.assign(K, 0.0);
    for (int e = 0; e < K; ++e) {
        const double* v = &h.vxy[size_t(e) * 6];
        const double bx = (v[0] + v[2] + v[4]) / 3.0, by = (v[1] + v[3] + v[5]) / 3.0;
        double dx[3], dy[3], tx[3], ty[3], hmax = 0.0;
        for (int f = 0; f < 3; ++f) {
            const int a = f, b = (f + 1) % 3;
            const double mx = 0.5 * (v[2 * a] + v[2 * b]), my = 0.5 * (v[2 * a + 1] + v[2 * b + 1]);
            const double ex = v[2 * b] - v[2 * a], ey = v[2 * b + 1] - v[2 * a + 1];
            hmax = std::fmax(hmax, std::sqrt(ex * ex + ey * ey));
            tx[f] = mx - bx;
            ty[f] = my - by;
            const int en = h.EToE[e * 3 + f];
            if (en >= 0) {
                const int fn = h.EToF[e * 3 + f];
                const double* w = &h.vxy[size_t(en) * 6];
                const double cx = (w[0] + w[2] + w[4]) / 3.0, cy = (w[1] + w[3] + w[5]) / 3.0;
                const int p = fn, q = (fn + 1) % 3;
                const double nmx = 0.5 * (w[2 * p] + w[2 * q]), nmy = 0.5 * (w[2 * p + 1] + w[2 * q + 1]);
                dx[f] = tx[f] + (cx - nmx);
                dy[f] = ty[f] + (cy - nmy);
            } else {
                dx[f] = 2.0 * tx[f];
                dy[f] = 2.0 * ty[f];
            }
        }
        hlen[e] = hmax;
        for (int f = 0; f < 3; ++f) {
            double best = -1e300, ba1 = 0.0, ba2 = 0.0;
            int bj = (f + 1) % 3;
            for (int k = 1; k <= 2; ++k) {
                const int j = (f + k) % 3;
                const double det = dx[f] * dy[j] - dy[f] * dx[j];
                const double a1 = (tx[f] * dy[j] - ty[f] * dx[j]) / det;
                const double a2 = (dx[f] * ty[f] - dy[f] * tx[f]) / det;
                const double score = std::fmin(a1, a2);
                if (score > best + 1e-12) {
                    best = score;
                    ba1 = a1;
                    ba2 = a2;
                    bj = j;
                }
            }
            coef[(size_t(e) * 3 + f) * 2 + 0] = ba1;
            coef[(size_t(e) * 3 + f) * 2 + 1] = ba2;
            jsel[size_t(e) * 3 + f] = bj;
        }
    }
}

// The moment operator (rows 0..3: mean, edge-midpoint values) and, for orders whose inputs fit the network,
// the feature operator as f64 16x16x4 MFMA A fragments.  Moments: lane l of K step s holds row l % 16 at node
// lim_mnode(s, l / 16) (zero beyond Np).  Features: lane l of K step s holds row l % 16 at node 4 s + l / 16, tiles
// in the row order and slots of lim_cff_slot.
static void limiter_fragments(const RefElem& re, std::vector<double>& p4f, std::vector<double>& cff) {
    const int Np = re.Np, NS = lim_ksteps(Np), NM = lim_msteps(Np);
    p4f.assign(size_t(NM) * 64, 0.0);
    for (int s = 0; s < NM; ++s)
        for (int l = 0; l < 64; ++l) {
            const int i = l & 15, n = lim_mnode(s, l >> 4);
            if (i < 4 && n < Np) p4f[size_t(s) * 64 + l] = re.tvb_p4[size_t(i) * Np + n];
        }
    cff.clear();
    const int nin = mlp_inputs(Np);
    if (nin > kMlpInMax) return;
    const int KS = mlp_pad(nin) / 16;
    cff.assign(size_t(lim_cff_pairs(Np, KS)) * NS * 64, 0.0);
    for (int t = 0; t < KS; ++t)
        for (int w = 0; w < 2; ++w) {
            if (!lim_cff_active(Np, t, w)) continue;
            const int slot = lim_cff_slot(Np, t, w);
            for (int s = 0; s < NS; ++s)
                for (int l = 0; l < 64; ++l) {
                    const int i = l & 15, n = 4 * s + (l >> 4);
                    const int mode = 16 * t + 4 * (i & 3) + (i >> 2) - w * Np;
                    if (n < Np && mode >= 0 && mode < Np)
                        cff[(size_t(slot) * NS + s) * 64 + l] = re.invV[size_t(mode) * Np + n];
                }
        }
}

static LimDev* limiter_for(dgh_ctx c, MeshDev* m) {
    if (m->lim) return m->lim.get();
    const HostMesh& h = m->host;
    const RefElem& re = m->ref->re;
    const int K = h.K, groups = (K + 15) / 16;
    std::vector<double> coef, hlen;
    std::vector<int> js;
    limiter_geometry(h, coef, js, hlen);
    std::vector<double> lgeo(size_t(groups) * kLimGeo * 16, 0.0);
    std::vector<int> jsel(size_t(3) * K);
    for (int e = 0; e < groups * 16; ++e) {
        const int ec = std::min(e, K - 1);
        auto at = [&](int s) -> double& { return lgeo[(size_t(e >> 4) * kLimGeo + s) * 16 + (e & 15)]; };
        for (int k = 0; k < 6; ++k) at(k) = coef[size_t(ec) * 6 + k];
        at(6) = hlen[ec];
        for (int f = 0; f < 3; ++f) {
            at(7 + 2 * f) = h.geo[size_t(ec) * 14 + 5 + 3 * f];
            at(8 + 2 * f) = h.geo[size_t(ec) * 14 + 6 + 3 * f];
        }
    }
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 3; ++f) jsel[size_t(f) * K + e] = js[size_t(e) * 3 + f];
    // point evaluation for the positivity scaling: nodes, cubature points, face Gauss points
    const int Np = re.Np, Nq = re.Nq, Nfq = re.Nfq, Nfp = re.Nfp, npos = Np + Nq + 3 * Nfq;
    std::vector<double> vpos(size_t(npos) * Np, 0.0);
    for (int i = 0; i < Np; ++i) vpos[size_t(i) * Np + i] = 1.0;
    for (int qq = 0; qq < Nq; ++qq)
        for (int j = 0; j < Np; ++j) vpos[size_t(Np + qq) * Np + j] = re.Vq[size_t(qq) * Np + j];
    for (int f = 0; f < 3; ++f)
        for (int gq = 0; gq < Nfq; ++gq)
            for (int k = 0; k < Nfp; ++k)
                vpos[size_t(Np + Nq + f * Nfq + gq) * Np + re.fmask[f * Nfp + k]] += re.Vf[size_t(gq) * Nfp + k];
    std::vector<double> p4f, cff;
    limiter_fragments(re, p4f, cff);
    auto L = std::make_shared<LimDev>();
    L->device = c->device;
    L->npos = npos;
    if (!c->upload(&L->p4f, p4f) || (!cff.empty() && !c->upload(&L->cff, cff))) return nullptr;
    const bool ok = c->upload(&L->lgeo, lgeo) && c->upload(&L->jsel, jsel) && c->upload(&L->p4, re.tvb_p4) &&
                    c->upload(&L->phi, re.tvb_phi) && c->upload(&L->wbc, re.tvb_wbc) && c->upload(&L->invV, re.invV) &&
                    c->upload(&L->vpos, vpos) &&
                    c->hip(hipMalloc(reinterpret_cast<void**>(&L->troubled), size_t(groups) * 16 * sizeof(int)),
                           "hipMalloc") &&
                    c->hip(hipMalloc(reinterpret_cast<void**>(&L->count), 2 * sizeof(unsigned)), "hipMalloc");
    if (!ok) return nullptr;
    m->lim = L;
    return L.get();
}

// the launch arguments shared by dgh_limit, dgh_limit_noalloc and dgh_mlp_features (allocates nothing)
static void limiter_args_fill(MeshDev* m, LimDev* L, int nvar, LimArgs& a) {
    const int K = m->host.K;
    a.K = K;
    a.nvar = nvar;
    a.lgeo = L->lgeo;
    a.jsel = L->jsel;
    a.conn = m->conn;
    a.p4 = L->p4;
    a.phi = L->phi;
    a.wbc = L->wbc;
    a.invV = L->invV;
    a.bcq = (m->bcq && m->bc_nvar == nvar) ? m->bcq : nullptr;
    a.mom = L->mom;
    a.troubled = L->troubled;
    a.count = L->count;
    a.pcount = L->count + 1;
    a.vpos = L->vpos;
    a.npos = L->npos;
    a.p4f = L->p4f;
    a.cff = L->cff;
}

// moments buffer for nvar variables (allocated on first use), and the launch arguments
static bool limiter_args(dgh_ctx c, MeshDev* m, LimDev* L, int nvar, LimArgs& a) {
    const int K = m->host.K, groups = (K + 15) / 16;
    if (L->mom_nvar < nvar) {
        if (L->mom) (void)hipFree(L->mom);
        L->mom = nullptr;
        L->mom_nvar = 0;
        if (!c->hip(hipMalloc(reinterpret_cast<void**>(&L->mom), size_t(groups) * nvar * 4 * 16 * sizeof(double)),
                    "hipMalloc moments"))
            return false;
        L->mom_nvar = nvar;
    }
    limiter_args_fill(m, L, nvar, a);
    return true;
}

// Geometry of the Triangle limiter (Trianglelimite.C:92-133, 150-262, 395-420) in the group layout
// [K/16][kTriGeo][16]: the affine map about the centroid, and per face the centroid offset of the neighbour
// (moved through the shared edge, so periodic images need no special case) or of the ghost cell (the own
// centroid mirrored through the face line), the endpoint difference, the signed diamond area and A_2.
static void triangle_geometry(const HostMesh& h, std::vector<double>& tgeo) {
    const int K = h.K, groups = (K + 15) / 16;
    tgeo.assign(size_t(groups) * kTriGeo * 16, 0.0);
    std::vector<double> cx(K), cy(K), A0(K);
    for (int e = 0; e < K; ++e) {
        const double* v = &h.vxy[size_t(e) * 6];
        cx[e] = (v[0] + v[2] + v[4]) / 3.0;
        cy[e] = (v[1] + v[3] + v[5]) / 3.0;
        const double J = 0.25 * ((v[2] - v[0]) * (v[5] - v[1]) - (v[4] - v[0]) * (v[3] - v[1]));
        A0[e] = J * 2.0 / 3.0;
    }
    for (int e = 0; e < groups * 16; ++e) {
        const int ec = std::min(e, K - 1);
        auto at = [&](int s) -> double& { return tgeo[(size_t(e >> 4) * kTriGeo + s) * 16 + (e & 15)]; };
        const double* v = &h.vxy[size_t(ec) * 6];
        at(0) = 0.5 * (v[2] - v[0]);
        at(1) = 0.5 * (v[4] - v[0]);
        at(2) = 0.5 * (v[3] - v[1]);
        at(3) = 0.5 * (v[5] - v[1]);
        for (int f = 0; f < 3; ++f) {
            const int a0 = f, a1 = (f + 1) % 3;
            const double ex = v[2 * a1] - v[2 * a0], ey = v[2 * a1 + 1] - v[2 * a0 + 1];
            const double nx = h.geo[size_t(ec) * 14 + 5 + 3 * f], ny = h.geo[size_t(ec) * 14 + 6 + 3 * f];
            const int en = h.EToE[ec * 3 + f];
            double ncx, ncy, A_2;
            if (en >= 0) {
                const int fn = h.EToF[ec * 3 + f], p = fn, r = (fn + 1) % 3;
                const double* w = &h.vxy[size_t(en) * 6];
                ncx = cx[en] + (0.5 * (v[2 * a0] + v[2 * a1]) - 0.5 * (w[2 * p] + w[2 * r]));
                ncy = cy[en] + (0.5 * (v[2 * a0 + 1] + v[2 * a1 + 1]) - 0.5 * (w[2 * p + 1] + w[2 * r + 1]));
                A_2 = A0[ec] + A0[en];
            } else {
                const double A = nx, B = ny, C = -v[2 * a0] * A - v[2 * a0 + 1] * B;
                ncx = (B * B - A * A) * cx[ec] - 2 * A * B * cy[ec] - 2 * A * C;
                ncy = (-B * B + A * A) * cy[ec] - 2 * A * B * cx[ec] - 2 * B * C;
                A_2 = A0[ec] + A0[ec];
            }
            const double dxc = ncx - cx[ec], dyc = ncy - cy[ec];
            at(4 + 8 * f) = dxc;
            at(5 + 8 * f) = dyc;
            at(6 + 8 * f) = ex;
            at(7 + 8 * f) = ey;
            at(8 + 8 * f) = (dxc * ey - ex * dyc) * 0.5;
            at(9 + 8 * f) = A_2;
            at(10 + 8 * f) = nx;
            at(11 + 8 * f) = ny;
        }
    }
}

static bool triangle_setup(dgh_ctx c, MeshDev* m, LimDev* L) {
    if (L->tgeo) return true;
    const RefElem& re = m->ref->re;
    const int groups = (m->host.K + 15) / 16;
    std::vector<double> tgeo, rsn(size_t(2) * re.Np);
    triangle_geometry(m->host, tgeo);
    for (int n = 0; n < re.Np; ++n) {
        rsn[2 * n] = re.r[n] + 1.0 / 3.0;
        rsn[2 * n + 1] = re.s[n] + 1.0 / 3.0;
    }
    auto alloc = [&](double** p, size_t n) {
        return c->hip(hipMalloc(reinterpret_cast<void**>(p), n * sizeof(double)), "hipMalloc");
    };
    if (!c->upload(&L->tgeo, tgeo) || !c->upload(&L->rsn, rsn) || !alloc(&L->tcell, size_t(groups) * kTriCell * 16) ||
        !alloc(&L->tcv, size_t(groups) * 8 * 16) || !alloc(&L->tvg, size_t(groups) * 3 * 8 * 16))
        return false;
    return true;
}

// Checks d against a state of nvar variables on mesh m and enqueues the limiter on the context stream for
// the state at q (no synchronisation); with reset_count the troubled-cell counter is cleared first.  Shared
// by dgh_limit and the limited LSERK4 loop (dgh_lserk4_run_limited).
// scratch: null, or the moments, flags and counters of dgh_limit_noalloc (then nothing is allocated: the mesh's
// limiter tables must exist, and the Triangle kinds, whose work arrays are not in it, are refused).
// moments_ready: the stage kernel that produced q has written its moments into the mesh's moments buffer
// (dg_limit_fused_moments), so the moments pass is skipped.
int dg_limit_enqueue(dgh_ctx c, MeshDev* m, const dgh_lim_desc* d, double* q, int nvar, bool reset_count,
                     const LimScratch* scratch, bool moments_ready) {
    if (!d) return c->fail(DGH_ERR_ARG, "dgh_limit: null descriptor"), DGH_ERR_ARG;
    if (d->kind < DGH_LIM_TVB || d->kind > DGH_LIM_TRIANGLE_MLP)
        return c->fail(DGH_ERR_UNSUPPORTED, "dgh_limit: kind %d", d->kind), DGH_ERR_UNSUPPORTED;
    const bool tri = d->kind >= DGH_LIM_TRIANGLE;
    if (tri && (nvar != 4 || d->detect_only || !(d->gamma > 1.0) || !(d->eps > 0.0)))
        return c->fail(DGH_ERR_ARG, "dgh_limit: the Triangle limiter needs 4 variables, gamma > 1, a density "
                                    "floor eps > 0, and rebuilds (no detect_only)"),
               DGH_ERR_ARG;
    if (nvar != 1 && nvar != 4)
        return c->fail(DGH_ERR_UNSUPPORTED, "dgh_limit: %d variables (1 or 4 supported)", nvar),
               DGH_ERR_UNSUPPORTED;
    if ((d->kind == DGH_LIM_MLP || d->kind == DGH_LIM_TRIANGLE_MLP) && nvar != 4)
        return c->fail(DGH_ERR_UNSUPPORTED, "dgh_limit: the learned indicator needs 4 variables"), DGH_ERR_UNSUPPORTED;
    if (!(d->M >= 0.0) || !(d->nu > 0.0)) return c->fail(DGH_ERR_ARG, "dgh_limit: M >= 0 and nu > 0"), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    const int K = m->host.K;
    LimDev* L = nullptr;
    LimArgs a;
    if (scratch) {
        if (tri)
            return c->fail(DGH_ERR_UNSUPPORTED, "dgh_limit_noalloc: the Triangle kinds need work arrays beyond the "
                                                "static scratch (use dgh_limit)"),
                   DGH_ERR_UNSUPPORTED;
        if (K > kNoallocMaxK)
            return c->fail(DGH_ERR_ARG, "dgh_limit_noalloc: K = %d exceeds the static scratch's %d cells", K,
                           kNoallocMaxK),
                   DGH_ERR_ARG;
        L = m->lim.get();
        if (!L)
            return c->fail(DGH_ERR_ARG, "dgh_limit_noalloc: the mesh's limiter tables are not built yet (one "
                                        "dgh_limit call on the mesh builds them)"),
                   DGH_ERR_ARG;
        limiter_args_fill(m, L, nvar, a);
        a.mom = scratch->mom;
        a.troubled = scratch->troubled;
        a.count = scratch->count;
        a.pcount = scratch->count + 1;
    } else {
        L = limiter_for(c, m);
        if (!L) return c->last;
        if (!limiter_args(c, m, L, nvar, a)) return c->last;
    }
    a.M = d->M;
    a.nu = d->nu;
    a.detect_only = d->detect_only != 0;
    if (d->positivity) {
        if (nvar != 4 || !(d->gamma > 1.0) || !(d->eps > 0.0))
            return c->fail(DGH_ERR_ARG, "dgh_limit: positivity needs 4 variables, gamma > 1, eps > 0"), DGH_ERR_ARG;
        a.positivity = 1;
        a.gamma = d->gamma;
        a.eps = d->eps;
    }
    if (d->kind == DGH_LIM_MLP || d->kind == DGH_LIM_TRIANGLE_MLP) {
        if (!dg_mlp_infer(c, d->mlp, m->ref->re.N, d->threshold, a.mlp)) return c->last;
        a.use_mlp = 1;
    }
    a.all = d->kind == DGH_LIM_ALL;
    a.moments_ready = moments_ready && !scratch && !tri;
    if (scratch) {
        if (reset_count && !c->hip(launch_zero_counts(a.count, c->stream), "zero counts")) return c->last;
    } else if (reset_count && !c->hip(hipMemsetAsync(L->count, 0, 2 * sizeof(unsigned), c->stream), "memset")) {
        return c->last;
    }
    if (tri) {
        if (!triangle_setup(c, m, L)) return c->last;
        TriArgs t;
        t.K = K;
        t.gamma = d->gamma;
        t.tol = d->eps;
        t.tgeo = L->tgeo;
        t.pmean = L->p4;   // row 0 of the moment operator: the mean weights
        t.rsn = L->rsn;
        t.conn = m->conn;
        t.bcq = a.bcq;
        t.pfirst = m->patch_first;
        t.cell = L->tcell;
        t.cv = L->tcv;
        t.vg = L->tvg;
        t.flags = L->troubled;
        t.select = d->kind != DGH_LIM_TRIANGLE;
        if (t.select) {
            // detection only: the TVB test or the network flags the cells the reconstruction rebuilds
            LimArgs det = a;
            det.detect_only = 1;
            det.positivity = 0;
            if (!c->hip(launch_tvb(m->ref->re.N, det, q, c->stream), "limiter detection")) return c->last;
        }
        if (!c->hip(launch_triangle(m->ref->re.N, t, q, c->stream), "triangle limiter launch")) return c->last;
        if (a.positivity && !c->hip(launch_positivity(m->ref->re.N, a, q, c->stream), "positivity launch"))
            return c->last;
    } else if (!c->hip(launch_tvb(m->ref->re.N, a, q, c->stream), "limiter launch")) {
        return c->last;
    }
    return DGH_OK;
}

// The limited run loops: when the limiter ld is of a TVB kind (TVB test, network, every cell) and the stages are
// bf16x3, the stage kernel forms the moments of its new state (stage_kernels.hip) into the mesh's moments buffer,
// and the limiter skips its moments pass.  Returns the buffer and the moment fragments in mom and p4f, or
// mom = null where this does not apply (any invalid ld is then reported by dg_limit_enqueue, as before).
int dg_limit_fused_moments(dgh_ctx c, MeshDev* m, const dgh_rhs_desc* d, const dgh_lim_desc* ld, int nvar,
                           double** mom, const double** p4f) {
    *mom = nullptr;
    *p4f = nullptr;
    if (!ld || !d || d->prec != DGH_PREC_BF16X3 || ld->detect_only) return DGH_OK;
    if (ld->kind != DGH_LIM_TVB && ld->kind != DGH_LIM_MLP && ld->kind != DGH_LIM_ALL) return DGH_OK;
    if ((nvar != 1 && nvar != 4) || (ld->kind == DGH_LIM_MLP && nvar != 4)) return DGH_OK;
    (void)hipSetDevice(c->device);
    LimDev* L = limiter_for(c, m);
    if (!L) return c->last;
    LimArgs a;
    if (!limiter_args(c, m, L, nvar, a)) return c->last;
    *mom = L->mom;
    *p4f = L->p4f;
    return DGH_OK;
}

// the flags and the count of a limiter call (flags, count: its device arrays) to the host
static int limit_results(dgh_ctx c, const dgh_lim_desc* d, int K, const int* flags, const unsigned* count,
                         int* troubled, int* ntroubled) {
    if (troubled &&
        !c->hip(hipMemcpyAsync(troubled, flags, size_t(K) * sizeof(int), hipMemcpyDeviceToHost, c->stream),
                "troubled download"))
        return c->last;
    const bool every_cell = d->kind == DGH_LIM_ALL || d->kind == DGH_LIM_TRIANGLE;   // count = K, not counted
    if (ntroubled && every_cell) {
        if (!c->hip(hipStreamSynchronize(c->stream), "sync")) return c->last;
        *ntroubled = K;
    } else if (ntroubled) {
        unsigned n = 0;
        if (!c->hip(hipMemcpyAsync(&n, count, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream), "count") ||
            !c->hip(hipStreamSynchronize(c->stream), "sync"))
            return c->last;
        *ntroubled = int(n);
    } else if (troubled && !c->hip(hipStreamSynchronize(c->stream), "sync")) {
        return c->last;
    }
    return DGH_OK;
}

extern "C" int dgh_limit(dgh_ctx c, int mh, const dgh_lim_desc* d, int q, int* troubled, int* ntroubled) {
    if (!c) return DGH_ERR_ARG;
    if (!d) return c->fail(DGH_ERR_ARG, "dgh_limit: null descriptor"), DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    FieldDev* f = c->field(q);
    if (!m || !f) return c->last;
    if (f->mesh != mh) return c->fail(DGH_ERR_ARG, "dgh_limit: field of another mesh"), DGH_ERR_ARG;
    if (int st = dg_limit_enqueue(c, m, d, f->d, f->nvar, true, nullptr, false)) return st;
    LimDev* L = m->lim.get();
    return limit_results(c, d, m->host.K, L->troubled, L->count, troubled, ntroubled);
}

extern "C" int dgh_limit_noalloc(dgh_ctx c, int mh, const dgh_lim_desc* d, int q, int* troubled, int* ntroubled) {
    if (!c) return DGH_ERR_ARG;
    if (!d) return c->fail(DGH_ERR_ARG, "dgh_limit_noalloc: null descriptor"), DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    FieldDev* f = c->field(q);
    if (!m || !f) return c->last;
    if (f->mesh != mh) return c->fail(DGH_ERR_ARG, "dgh_limit_noalloc: field of another mesh"), DGH_ERR_ARG;
    if (!c->noalloc) return c->fail(DGH_ERR_HIP, "dgh_limit_noalloc: no static scratch on this device"), DGH_ERR_HIP;
    // the device address looked up at dgh_create, never the array's host-side name
    char* base = static_cast<char*>(c->noalloc);
    const LimScratch s{reinterpret_cast<double*>(base), reinterpret_cast<int*>(base + kNoallocFlagOff),
                       reinterpret_cast<unsigned*>(base + kNoallocCountOff)};
    if (int st = dg_limit_enqueue(c, m, d, f->d, f->nvar, true, &s, false)) return st;
    return limit_results(c, d, m->host.K, s.troubled, s.count, troubled, ntroubled);
}

extern "C" int dgh_mlp_features(dgh_ctx c, int mh, int q, double M, float* X) {
    if (!c || !X) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    FieldDev* f = c->field(q);
    if (!m || !f) return c->last;
    const int N = m->ref->re.N, Np = m->ref->re.Np, nin = mlp_inputs(Np);
    if (f->mesh != mh || f->nvar != 4 || nin > kMlpInMax)
        return c->fail(DGH_ERR_UNSUPPORTED, "dgh_mlp_features: 4-variable fields of orders 1..6"), DGH_ERR_UNSUPPORTED;
    (void)hipSetDevice(c->device);
    LimDev* L = limiter_for(c, m);
    if (!L) return c->last;
    LimArgs a;
    if (!limiter_args(c, m, L, 4, a)) return c->last;
    if (!(M > 0.0)) return c->fail(DGH_ERR_ARG, "dgh_mlp_features: M > 0"), DGH_ERR_ARG;
    a.M = M;
    const int K = m->host.K;
    float* dX = nullptr;
    if (!c->hip(hipMalloc(reinterpret_cast<void**>(&dX), size_t(K) * nin * sizeof(float)), "hipMalloc")) return c->last;
    const bool ok = c->hip(launch_features(N, a, f->d, dX, c->stream), "features") &&
                    c->hip(hipMemcpyAsync(X, dX, size_t(K) * nin * sizeof(float), hipMemcpyDeviceToHost, c->stream),
                           "d2h") &&
                    c->hip(hipStreamSynchronize(c->stream), "sync");
    (void)hipFree(dX);
    return ok ? DGH_OK : c->last;
}

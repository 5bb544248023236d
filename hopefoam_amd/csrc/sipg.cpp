// Host side of the implicit SIPG solve: operator set-up, preconditioners and the PCG driver behind
// dgh_sipg_* (include/dgh.h).  The kernels are in sipg.hip.
//
// Replaces, for dgm::laplacian(u) == f (HopeFOAM-0.1/src/DG/DG/dgm/dgmLaplacian.H:61-65):
//   defaultLaplacianScheme::laplacianFunc (defaultLaplacianScheme.C:62-288) -- no matrix is formed;
//   dgMatrix::{petscPreAllocate,assembleMatrix,finalAssemble} (dgMatrix.C:296-446);
//   dgMatrix::kspSolver / dgScalarMatrix::solve (dgMatrix.C:163-283, dgScalarMatrix.C:67-188):
//   PETSc KSPCG with PCJACOBI (or element block Jacobi), preconditioned-norm convergence test
//   relative to the initial residual, nonzero initial guess.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "dg_common.hpp"
#include "dgh.h"
#include "dgh_internal.hpp"
#include "sipg.hpp"

using namespace dg;

namespace dg {

template <typename T>
struct SipgOps {
    T *geo = nullptr, *Dr = nullptr, *Ds = nullptr, *Vf = nullptr, *gw = nullptr;
    T *Srr = nullptr, *Srs = nullptr, *Sss = nullptr, *M = nullptr, *TrF = nullptr, *ApF = nullptr, *MF = nullptr;
    T *dinv = nullptr, *binv = nullptr;   // binv: the inverse blocks as MFMA A fragments (block_fragments)
    T *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *w = nullptr, *ap = nullptr, *rhs = nullptr,
      *trace = nullptr;
};

struct SipgDev {
    int device = 0;
    int mesh = -1, N = 0, K = 0, Np = 0, groups = 0;
    dgh_sipg_desc desc{};
    VecShape vs{};
    size_t ntrace = 0;
    int* conn = nullptr;   // owned by the mesh
    int *fmask = nullptr, *nface = nullptr, *bidx = nullptr;
    int nblocks = 0;       // distinct inverse blocks
    int ncu = 256;
    SipgOps<double> d;
    SipgOps<float> s;
    bool have32 = false;
    double* part = nullptr;
    CgState* st = nullptr;       // [4]: the iteration's two states, the published state, auxiliary norms
    double* part_b[2] = {nullptr, nullptr};   // the update kernel's partials by iteration parity (part holds the
                                              // apply's and the norms')
    CgState* host = nullptr;     // pinned [4]: two pipeline slots + aux
    hipEvent_t slot_ev[2] = {nullptr, nullptr};
    hipGraphExec_t chunk[2] = {nullptr, nullptr};   // kChunk iterations + the publish, FP64 / FP32 vectors
    std::vector<void*> allocs;

    ~SipgDev() {
        (void)hipSetDevice(device);
        for (auto& g : chunk)
            if (g) (void)hipGraphExecDestroy(g);
        for (void* a : allocs) (void)hipFree(a);
        if (host) (void)hipHostFree(host);
        for (auto& e : slot_ev)
            if (e) (void)hipEventDestroy(e);
    }
};

}  // namespace dg

namespace {

constexpr int kChunk = 32;   // CG iterations queued between two looks at the device state
static_assert(kChunk % 2 == 0, "the published state is read after an even number of iterations");

template <typename T>
bool dev_upload(dgh_ctx c, SipgDev& S, T** p, const std::vector<T>& v) {
    if (!c->upload(p, v)) return false;
    S.allocs.push_back(*p);
    return true;
}

template <typename T>
bool dev_alloc(dgh_ctx c, SipgDev& S, T** p, size_t n) {
    if (!c->hip(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc sipg"))
        return false;
    S.allocs.push_back(*p);
    return c->hip(vec_zero(*p, std::max<size_t>(n, 1) * sizeof(T), c->stream), "zero");
}

template <typename T>
std::vector<T> as(const std::vector<double>& v) {
    return std::vector<T>(v.begin(), v.end());
}

// key of an element's diagonal block: its SIPG geometry rounded to 2^-40 relative, plus face kinds
using BlockKey = std::array<long long, 2 * kSipgGeo + 3>;
BlockKey block_key(const double* g, const int* kind) {
    BlockKey k{};
    for (int i = 0; i < kSipgGeo; ++i) {
        int ex = 0;
        const double m = std::frexp(g[i], &ex);
        k[2 * i] = std::llround(std::ldexp(m, 40));
        k[2 * i + 1] = g[i] == 0.0 ? 0 : ex;
    }
    for (int f = 0; f < 3; ++f) k[2 * kSipgGeo + f] = kind[f];
    return k;
}

}  // namespace

// A-operand fragments of the two GEMMs (lane l of a 16x16x4 step holds A[16 tile + l % 16][4 s + l / 16]);
// the tiling is described with SipgTiles (sipg.hpp).
static void sipg_fragments(const RefElem& re, const std::vector<double>& Srr, const std::vector<double>& Srs,
                           const std::vector<double>& Sss, std::vector<double>& TrF, std::vector<double>& ApF,
                           std::vector<double>& MF) {
    const SipgTiles S(re.N);
    const int Np = re.Np, Nfp = re.Nfp, Nfq = re.Nfq;
    // Vface, DrF, DsF: [NFQ3][Np]
    std::vector<double> Vface(size_t(S.NFQ3) * Np, 0.0), DrF(size_t(S.NFQ3) * Np, 0.0), DsF(size_t(S.NFQ3) * Np, 0.0);
    for (int f = 0; f < 3; ++f)
        for (int q = 0; q < Nfq; ++q) {
            const int it = f * Nfq + q;
            for (int kk = 0; kk < Nfp; ++kk) {
                const int node = re.fmask[f * Nfp + kk];
                const double w = re.Vf[size_t(q) * Nfp + kk];
                Vface[size_t(it) * Np + node] += w;
                for (int k = 0; k < Np; ++k) {
                    DrF[size_t(it) * Np + k] += w * re.Dr[size_t(node) * Np + k];
                    DsF[size_t(it) * Np + k] += w * re.Ds[size_t(node) * Np + k];
                }
            }
        }
    const std::vector<double>* blocks[3] = {&Vface, &DrF, &DsF};
    TrF.assign(size_t(3 * S.TI) * S.KS * 64, 0.0);
    for (int b = 0; b < 3; ++b)
        for (int t = 0; t < S.TI; ++t)
            for (int st = 0; st < S.KS; ++st)
                for (int l = 0; l < 64; ++l) {
                    const int item = 16 * t + (l & 15), k = 4 * st + (l >> 4);
                    if (item < S.NFQ3 && k < Np)
                        TrF[((size_t(b * S.TI + t)) * S.KS + st) * 64 + l] = (*blocks[b])[size_t(item) * Np + k];
                }
    // apply: columns [Srr | Srs | Sss] (NPP each) then [Vface^T | DrF^T | DsF^T] (NIP each)
    auto apply_a = [&](int n, int col) -> double {
        if (n >= Np) return 0.0;
        if (col < 3 * S.NPP) {
            const int b = col / S.NPP, k = col % S.NPP;
            if (k >= Np) return 0.0;
            const std::vector<double>& M = b == 0 ? Srr : (b == 1 ? Srs : Sss);
            return M[size_t(n) * Np + k];
        }
        const int c2 = col - 3 * S.NPP, b = c2 / S.NIP, it = c2 % S.NIP;
        if (it >= S.NFQ3) return 0.0;
        return (*blocks[b])[size_t(it) * Np + n];
    };
    ApF.assign(size_t(S.MT) * S.KA * 64, 0.0);
    for (int m = 0; m < S.MT; ++m)
        for (int st = 0; st < S.KA; ++st)
            for (int l = 0; l < 64; ++l) ApF[(size_t(m) * S.KA + st) * 64 + l] = apply_a(16 * m + (l & 15), 4 * st + (l >> 4));
    // the reference mass matrix in the apply's row tiles and its node K steps (the mass shift)
    const int KU = S.NPP / 4;
    MF.assign(size_t(S.MT) * KU * 64, 0.0);
    for (int m = 0; m < S.MT; ++m)
        for (int st = 0; st < KU; ++st)
            for (int l = 0; l < 64; ++l) {
                const int n = 16 * m + (l & 15), k = 4 * st + (l >> 4);
                if (n < Np && k < Np) MF[(size_t(m) * KU + st) * 64 + l] = re.M[size_t(n) * Np + k];
            }
}

// The inverse blocks as A fragments of the 16x16x4 MFMA, [block][MT][NPP / 4][64 lanes]: lane l of node step s
// in row tile m holds row 16 m + rho(l % 16), column 4 s + l / 16 (zero outside Np x Np).  rho is the identity
// for FP64; for FP32 it is the 4 x 4 transpose 4 a + b -> a + 4 b.  Either way accumulator register r of row
// group kq then holds node 16 m + kq + 4 r, which is the node that row group's lanes supply at node step
// 4 m + r of the next product (the B-fragment layout the vectors are read in).
static std::vector<double> block_fragments(int N, int nblocks, const std::vector<double>& binv, bool f32) {
    const SipgTiles S(N);
    const int Np = S.Np, KS = S.NPP / 4;
    std::vector<double> F(size_t(nblocks) * S.MT * KS * 64, 0.0);
    for (int b = 0; b < nblocks; ++b)
        for (int m = 0; m < S.MT; ++m)
            for (int st = 0; st < KS; ++st)
                for (int l = 0; l < 64; ++l) {
                    const int i = l & 15, row = 16 * m + (f32 ? (i >> 2) + 4 * (i & 3) : i), col = 4 * st + (l >> 4);
                    if (row < Np && col < Np)
                        F[((size_t(b) * S.MT + m) * KS + st) * 64 + l] = binv[(size_t(b) * Np + row) * Np + col];
                }
    return F;
}

// Diagonal block A_ee of the reference's SIPG matrix (defaultLaplacianScheme.C:96-102 volume,
// :222-286 interior faces, :153-209 Dirichlet faces), from the element's SIPG geometry.
// kind: 0 interior, 1 Dirichlet, 2 no face term.
static void element_block(const RefElem& re, const std::vector<double>& Srr, const std::vector<double>& Srs,
                          const std::vector<double>& Sss, const double* g, const int* kind, std::vector<double>& blk) {
    const int Np = re.Np, Nfp = re.Nfp, Nfq = re.Nfq;
    blk.assign(size_t(Np) * Np, 0.0);
    for (int i = 0; i < Np * Np; ++i) blk[i] = -g[0] * (g[1] * Srr[i] + g[2] * Srs[i] + g[3] * Sss[i]);
    std::vector<double> V1(size_t(Nfq) * Np), Dn(size_t(Nfq) * Np);
    for (int f = 0; f < 3; ++f) {
        if (kind[f] == 2) continue;
        const double fac = kind[f] == 0 ? 0.5 : 1.0;
        const double ar = g[4 + 2 * f], as = g[5 + 2 * f], js = g[10 + f], tau = g[13 + f];
        std::fill(V1.begin(), V1.end(), 0.0);
        std::fill(Dn.begin(), Dn.end(), 0.0);
        for (int q = 0; q < Nfq; ++q)
            for (int k = 0; k < Nfp; ++k) {
                const int node = re.fmask[f * Nfp + k];
                const double w = re.Vf[size_t(q) * Nfp + k];
                V1[size_t(q) * Np + node] += w;
                for (int j = 0; j < Np; ++j)
                    Dn[size_t(q) * Np + j] += w * (ar * re.Dr[size_t(node) * Np + j] + as * re.Ds[size_t(node) * Np + j]);
            }
        for (int q = 0; q < Nfq; ++q) {
            const double w = js * re.gw[q];
            const double* v = &V1[size_t(q) * Np];
            const double* d = &Dn[size_t(q) * Np];
            for (int i = 0; i < Np; ++i)
                for (int j = 0; j < Np; ++j)
                    blk[size_t(i) * Np + j] += w * fac * (-tau * v[i] * v[j] + v[i] * d[j] + d[i] * v[j]);
        }
    }
}

static int sipg_build(dgh_ctx c, int mh, const dgh_sipg_desc* desc, std::shared_ptr<SipgDev>& out) {
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    const RefElem& re = m->ref->re;
    const HostMesh& h = m->host;
    auto S = std::make_shared<SipgDev>();
    S->device = c->device;
    S->mesh = mh;
    S->desc = *desc;
    S->N = re.N;
    S->K = h.K;
    S->Np = re.Np;
    S->groups = (h.K + 15) / 16;
    S->vs = vec_shape(h.K, re.Np);
    S->ntrace = size_t(S->groups) * 3 * re.Nfq * 2 * 16;
    S->conn = m->conn;
    S->ncu = c->ncu;
    const int Np = re.Np, Nq = re.Nq, K = h.K, N = re.N;

    // stiffness pieces from the cubature rule (long double sums, rounded once)
    std::vector<double> Srr(size_t(Np) * Np), Srs(size_t(Np) * Np), Sss(size_t(Np) * Np);
    for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
            long double a = 0, b = 0, cc = 0;
            for (int q = 0; q < Nq; ++q) {
                const long double w = re.cw[q];
                const long double ri = re.Drq[size_t(q) * Np + i], rj = re.Drq[size_t(q) * Np + j];
                const long double si = re.Dsq[size_t(q) * Np + i], sj = re.Dsq[size_t(q) * Np + j];
                a += w * ri * rj;
                b += w * (ri * sj + si * rj);
                cc += w * si * sj;
            }
            Srr[size_t(i) * Np + j] = double(a);
            Srs[size_t(i) * Np + j] = double(b);
            Sss[size_t(i) * Np + j] = double(cc);
        }
    std::vector<int> nface(size_t(Np) * 2, -1);
    for (int f = 0; f < 3; ++f)
        for (int k = 0; k < re.Nfp; ++k) {
            const int node = re.fmask[f * re.Nfp + k];
            nface[2 * node + (nface[2 * node] >= 0 ? 1 : 0)] = f * re.Nfp + k;
        }

    // per-element SIPG geometry
    const double n1 = N + 1.0;
    const double pen = desc->penalty == DGH_SIPG_PENALTY_REFERENCE ? 20.0 * n1 * n1 : n1 * n1;
    std::vector<double> geo(size_t(K) * kSipgGeo, 0.0);
    std::vector<int> kind(size_t(K) * 3);
    for (int e = 0; e < K; ++e) {
        const double* g = &h.geo[size_t(e) * 14];
        const double rx = g[0], ry = g[1], sx = g[2], sy = g[3], J = g[4];
        double* o = &geo[size_t(e) * kSipgGeo];
        o[0] = J;
        o[1] = rx * rx + ry * ry;
        o[2] = rx * sx + ry * sy;
        o[3] = sx * sx + sy * sy;
        for (int f = 0; f < 3; ++f) {
            const double nx = g[5 + 3 * f], ny = g[6 + 3 * f], js = g[7 + 3 * f];
            o[4 + 2 * f] = nx * rx + ny * ry;
            o[5 + 2 * f] = nx * sx + ny * sy;
            const int en = h.EToE[e * 3 + f];
            double fs = js / J, jown = js;
            if (en >= 0) {
                const double* gn = &h.geo[size_t(en) * 14];
                const double jsn = gn[7 + 3 * h.EToF[e * 3 + f]];
                fs = std::max(fs, jsn / gn[4]);
                if (en < e) jown = jsn;
                kind[e * 3 + f] = 0;
            } else {
                kind[e * 3 + f] = h.bctag[e * 3 + f] == BC_FIXED ? 1 : 2;
            }
            o[10 + f] = jown;
            o[13 + f] = fs * pen;
        }
    }

    // diagonal blocks: exact diagonal per element (point Jacobi); inverse blocks shared by elements
    // whose geometry agrees to 2^-40 relative (block Jacobi)
    const size_t nvec = S->vs.n;
    std::vector<double> dinv(nvec, 0.0);
    std::vector<int> bidx(size_t(S->groups) * 16, 0);
    std::vector<double> binv;
    std::map<BlockKey, int> seen;
    std::vector<double> blk, Sblk(size_t(Np) * Np), inv;
    const bool block = desc->pc == DGH_PC_BLOCK_JACOBI;
    const double sigma = desc->mass_shift;
    for (int e = 0; e < K; ++e) {
        const double* g = &geo[size_t(e) * kSipgGeo];
        element_block(re, Srr, Srs, Sss, g, &kind[size_t(e) * 3], blk);
        if (sigma != 0.0)   // A + sigma M in the reference sign: the block gains -sigma J M
            for (int i = 0; i < Np * Np; ++i) blk[i] -= sigma * g[0] * re.M[i];
        for (int i = 0; i < Np; ++i) dinv[dev_index(e, 0, i, 1, Np)] = -1.0 / blk[size_t(i) * Np + i];
        if (!block) continue;
        const BlockKey key = block_key(g, &kind[size_t(e) * 3]);
        auto it = seen.find(key);
        if (it != seen.end()) {
            bidx[e] = it->second;
            continue;
        }
        for (int i = 0; i < Np * Np; ++i) Sblk[i] = -blk[i];
        if (!invert(Np, Sblk, inv))
            return c->fail(DGH_ERR_ARG, "sipg: singular diagonal block at element %d", e), DGH_ERR_ARG;
        const int id = int(seen.size());
        seen.emplace(key, id);
        bidx[e] = id;
        for (int i = 0; i < Np; ++i)
            for (int j = 0; j < Np; ++j)
                binv.push_back(0.5 * (inv[size_t(i) * Np + j] + inv[size_t(j) * Np + i]));
    }
    for (int e = K; e < S->groups * 16; ++e) bidx[e] = bidx[K - 1];
    S->nblocks = int(seen.size());

    // geometry in the group layout [group][slot][16]; padding repeats element K-1
    std::vector<double> geog(size_t(S->groups) * kSipgGeo * 16);
    for (int e = 0; e < S->groups * 16; ++e) {
        const int ec = std::min(e, K - 1);
        for (int k = 0; k < kSipgGeo; ++k)
            geog[(size_t(e >> 4) * kSipgGeo + k) * 16 + (e & 15)] = geo[size_t(ec) * kSipgGeo + k];
    }

    SipgDev& D = *S;
    std::vector<double> TrF, ApF, MF;
    sipg_fragments(re, Srr, Srs, Sss, TrF, ApF, MF);
    bool ok = dev_upload(c, D, &D.d.TrF, TrF) && dev_upload(c, D, &D.d.ApF, ApF) &&
              dev_upload(c, D, &D.s.TrF, as<float>(TrF)) && dev_upload(c, D, &D.s.ApF, as<float>(ApF)) &&
              dev_upload(c, D, &D.d.MF, MF) && dev_upload(c, D, &D.s.MF, as<float>(MF));
    ok = ok && dev_upload(c, D, &D.d.geo, geog) && dev_upload(c, D, &D.d.Dr, re.Dr) && dev_upload(c, D, &D.d.Ds, re.Ds) &&
              dev_upload(c, D, &D.d.Vf, re.Vf) && dev_upload(c, D, &D.d.gw, re.gw) && dev_upload(c, D, &D.d.Srr, Srr) &&
              dev_upload(c, D, &D.d.Srs, Srs) && dev_upload(c, D, &D.d.Sss, Sss) && dev_upload(c, D, &D.d.M, re.M) &&
              dev_upload(c, D, &D.d.dinv, dinv) && dev_upload(c, D, &D.fmask, re.fmask) &&
              dev_upload(c, D, &D.nface, nface) && dev_upload(c, D, &D.bidx, bidx);
    ok = ok && dev_upload(c, D, &D.s.geo, as<float>(geog)) && dev_upload(c, D, &D.s.Dr, as<float>(re.Dr)) &&
         dev_upload(c, D, &D.s.Ds, as<float>(re.Ds)) && dev_upload(c, D, &D.s.Vf, as<float>(re.Vf)) &&
         dev_upload(c, D, &D.s.gw, as<float>(re.gw)) && dev_upload(c, D, &D.s.Srr, as<float>(Srr)) &&
         dev_upload(c, D, &D.s.Srs, as<float>(Srs)) && dev_upload(c, D, &D.s.Sss, as<float>(Sss)) &&
         dev_upload(c, D, &D.s.dinv, as<float>(dinv));
    if (ok && block)
        ok = dev_upload(c, D, &D.d.binv, block_fragments(N, S->nblocks, binv, false)) &&
             dev_upload(c, D, &D.s.binv, as<float>(block_fragments(N, S->nblocks, binv, true)));
    ok = ok && dev_alloc(c, D, &D.d.r, nvec) && dev_alloc(c, D, &D.d.z, nvec) && dev_alloc(c, D, &D.d.p, nvec) &&
         dev_alloc(c, D, &D.d.w, nvec) && dev_alloc(c, D, &D.d.ap, nvec) && dev_alloc(c, D, &D.d.x, nvec) &&
         dev_alloc(c, D, &D.d.rhs, nvec) && dev_alloc(c, D, &D.d.trace, D.ntrace);
    const size_t npart = 2 * size_t(std::max(S->groups, 2048));
    ok = ok && dev_alloc(c, D, &D.part, npart) && dev_alloc(c, D, &D.part_b[0], npart) &&
         dev_alloc(c, D, &D.part_b[1], npart) && dev_alloc(c, D, &D.st, 4);
    if (!ok) return c->last;
    if (!c->hip(hipHostMalloc(reinterpret_cast<void**>(&D.host), 4 * sizeof(CgState)), "hipHostMalloc")) return c->last;
    std::memset(D.host, 0, 4 * sizeof(CgState));
    for (auto& e : D.slot_ev)
        if (!c->hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event")) return c->last;
    out = S;
    return DGH_OK;
}

template <typename T>
static SipgArgs<T> sipg_args(const SipgDev& S, const SipgOps<T>& o) {
    SipgArgs<T> a;
    a.K = S.K;
    a.geo = o.geo;
    a.conn = S.conn;
    a.Dr = o.Dr;
    a.Ds = o.Ds;
    a.Vf = o.Vf;
    a.gw = o.gw;
    a.Srr = o.Srr;
    a.Srs = o.Srs;
    a.Sss = o.Sss;
    a.M = o.M;
    a.fmask = S.fmask;
    a.nface = S.nface;
    a.TrF = o.TrF;
    a.ApF = o.ApF;
    a.MF = o.MF;
    a.sigma = T(S.desc.mass_shift);
    a.ncu = S.ncu;
    return a;
}

static bool ensure32(dgh_ctx c, SipgDev& S) {
    if (S.have32) return true;
    const size_t n = S.vs.n;
    S.have32 = dev_alloc(c, S, &S.s.x, n) && dev_alloc(c, S, &S.s.r, n) && dev_alloc(c, S, &S.s.z, n) &&
               dev_alloc(c, S, &S.s.p, n) && dev_alloc(c, S, &S.s.w, n) && dev_alloc(c, S, &S.s.ap, n) && dev_alloc(c, S, &S.s.rhs, n) &&
               dev_alloc(c, S, &S.s.trace, S.ntrace);
    return S.have32;
}

// Runs PCG on S x = rhs (S = -A, SPD) from the state prepared by the caller; x, r, z, p, w are the
// precision's work vectors.  Queues kChunk iterations at a time, two chunks in flight.
template <typename T>
static bool run_pcg(dgh_ctx c, SipgDev& S, SipgOps<T>& o, double rtol, double abstol, int maxit, CgState& result) {
    T* x = o.x;
    const SipgArgs<T> a = sipg_args(S, o);
    const bool block = S.desc.pc == DGH_PC_BLOCK_JACOBI;
    const T* binv = block ? o.binv : nullptr;
    CgState init{};
    init.rtol = rtol;
    init.abstol = abstol;
    init.maxit = maxit;
    init.its = -1;   // the first finish is the initial one
    S.host[3] = init;
    if (!c->hip(hipMemcpyAsync(S.st + 1, &S.host[3], sizeof(CgState), hipMemcpyHostToDevice, c->stream), "state") ||
        !c->hip(hipStreamSynchronize(c->stream), "sync"))
        return false;
    const CgLoop<T> loop{x,      o.r,    o.z,    o.w,  o.p,  o.ap,   o.trace, o.rhs, o.dinv, binv, S.bidx,
                         S.st,   S.part, {S.part_b[0], S.part_b[1]}, S.ncu};
    // w = S x, then r = rhs - w, z = B r with its traces, w = S z, norms
    if (!c->hip(sipg_apply<T>(S.N, a, x, o.trace, o.w, 1, c->stream), "sipg apply") ||
        !c->hip(cg_start<T>(S.N, a, loop, c->stream), "cg init"))
        return false;
    // the chunk of kChunk iterations and the publish, captured once per solver and precision (its kernels' arguments
    // are the solver's own buffers) and replayed
    hipGraphExec_t& exec = S.chunk[sizeof(T) == 4];
    if (!exec) {
        hipGraph_t g = nullptr;
        if (!c->hip(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture")) return false;
        bool ok = true;
        for (int i = 0; i < kChunk && ok; ++i) ok = c->hip(cg_iteration<T>(S.N, a, loop, i, c->stream), "cg iteration");
        ok = ok && c->hip(cg_publish<T>(a, loop, c->stream), "cg state");
        const bool ended = c->hip(hipStreamEndCapture(c->stream, &g), "end capture");
        if (!ok || !ended) {
            if (g) (void)hipGraphDestroy(g);
            return false;
        }
        const bool inst = c->hip(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), "graph instantiate");
        (void)hipGraphDestroy(g);
        if (!inst) return false;
    }
    int chunk = 0;
    for (;;) {
        if (!c->hip(hipGraphLaunch(exec, c->stream), "graph launch")) return false;
        const int slot = chunk & 1;
        if (!c->hip(hipMemcpyAsync(&S.host[slot], S.st + 2, sizeof(CgState), hipMemcpyDeviceToHost, c->stream), "state") ||
            !c->hip(hipEventRecord(S.slot_ev[slot], c->stream), "event"))
            return false;
        if (chunk > 0) {
            const int prev = slot ^ 1;
            if (!c->hip(hipEventSynchronize(S.slot_ev[prev]), "event sync")) return false;
            if (S.host[prev].done) break;
        }
        ++chunk;
    }
    if (!c->hip(hipStreamSynchronize(c->stream), "sync")) return false;
    result = S.host[chunk & 1];
    return true;
}

// ||v|| over live entries (FP64 accumulation), via the auxiliary state
template <typename T>
static bool device_norm(dgh_ctx c, SipgDev& S, const T* v, double& out) {
    if (!c->hip(vec_norm<T>(S.vs, v, S.st + 3, S.part, c->stream), "norm") ||
        !c->hip(hipMemcpyAsync(&S.host[2], S.st + 3, sizeof(CgState), hipMemcpyDeviceToHost, c->stream), "state") ||
        !c->hip(hipStreamSynchronize(c->stream), "sync"))
        return false;
    out = S.host[2].dp;
    return true;
}

// r = rhs - S u with rhs = -b already in S.d.rhs; rn = |r| (= |b - A u|), FP64
static bool true_residual(dgh_ctx c, SipgDev& S, const double* u, double& rn) {
    return c->hip(sipg_apply<double>(S.N, sipg_args(S, S.d), u, S.d.trace, S.d.w, 1, c->stream), "apply") &&
           c->hip(vec_sub_live(S.vs, S.d.r, S.d.rhs, S.d.w, c->stream), "residual") &&
           device_norm<double>(c, S, S.d.r, rn);
}

static SipgDev* sipg_get(dgh_ctx c, int h) {
    if (h < 0 || h >= int(c->sipg.size()) || !c->sipg[h]) {
        c->fail(DGH_ERR_ARG, "invalid sipg handle %d", h);
        return nullptr;
    }
    return c->sipg[h].get();
}

static FieldDev* scalar_field(dgh_ctx c, const SipgDev& S, int fh) {
    FieldDev* f = c->field(fh);
    if (!f) return nullptr;
    if (f->mesh != S.mesh || f->nvar != 1) {
        c->fail(DGH_ERR_ARG, "sipg: field %d is not a scalar field of the solver's mesh", fh);
        return nullptr;
    }
    return f;
}

extern "C" {

int dgh_sipg_create(dgh_ctx c, int mesh, const dgh_sipg_desc* d, int* out) {
    if (!c) return DGH_ERR_ARG;
    if (!d || !out) return c->fail(DGH_ERR_ARG, "dgh_sipg_create: null argument"), DGH_ERR_ARG;
    if ((d->penalty != DGH_SIPG_PENALTY_REFERENCE && d->penalty != DGH_SIPG_PENALTY_H) ||
        (d->pc != DGH_PC_JACOBI && d->pc != DGH_PC_BLOCK_JACOBI) ||
        (d->mode != DGH_SOLVE_FP64 && d->mode != DGH_SOLVE_MIXED) || d->maxit < 0 || !(d->rtol >= 0.0) ||
        !(d->abstol >= 0.0) || !(d->mass_shift >= 0.0) || !std::isfinite(d->mass_shift))
        return c->fail(DGH_ERR_ARG, "dgh_sipg_create: bad descriptor"), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    std::shared_ptr<SipgDev> S;
    if (int st = sipg_build(c, mesh, d, S)) return st;
    c->sipg.push_back(S);
    *out = int(c->sipg.size()) - 1;
    return DGH_OK;
}

int dgh_sipg_destroy(dgh_ctx c, int h) {
    if (!c) return DGH_ERR_ARG;
    if (!sipg_get(c, h)) return c->last;
    (void)hipStreamSynchronize(c->stream);
    c->sipg[h].reset();
    return DGH_OK;
}

int dgh_sipg_info(dgh_ctx c, int h, int* info) {
    if (!c || !info) return DGH_ERR_ARG;
    SipgDev* S = sipg_get(c, h);
    if (!S) return c->last;
    info[0] = S->N;
    info[1] = S->K;
    info[2] = S->Np;
    info[3] = S->nblocks;
    return DGH_OK;
}

int dgh_sipg_rhs_neumann(dgh_ctx c, int h, int f, const double* gbc, const double* gnb, int b) {
    if (!c) return DGH_ERR_ARG;
    SipgDev* S = sipg_get(c, h);
    if (!S) return c->last;
    FieldDev* ff = scalar_field(c, *S, f);
    FieldDev* fb = scalar_field(c, *S, b);
    if (!ff || !fb) return c->last;
    if (f == b) return c->fail(DGH_ERR_ARG, "dgh_sipg_rhs: f and b must differ"), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    const size_t n = size_t(S->K) * 3 * (S->N + 1);
    double* dg[2] = {nullptr, nullptr};
    const double* host[2] = {gbc, gnb};
    bool ok = true;
    for (int i = 0; i < 2 && ok; ++i) {
        if (!host[i]) continue;
        ok = c->hip(hipMalloc(reinterpret_cast<void**>(&dg[i]), n * sizeof(double)), "hipMalloc boundary data") &&
             c->hip(hipMemcpyAsync(dg[i], host[i], n * sizeof(double), hipMemcpyHostToDevice, c->stream),
                    "boundary data upload");
    }
    ok = ok && c->hip(sipg_rhs(S->N, sipg_args(*S, S->d), ff->d, dg[0], dg[1], fb->d, c->stream), "sipg rhs") &&
         c->hip(hipStreamSynchronize(c->stream), "sync");
    for (double* p : dg)
        if (p) (void)hipFree(p);
    return ok ? DGH_OK : c->last;
}

int dgh_sipg_rhs(dgh_ctx c, int h, int f, const double* gbc, int b) { return dgh_sipg_rhs_neumann(c, h, f, gbc, nullptr, b); }

int dgh_sipg_apply(dgh_ctx c, int h, int prec, int u, int out) {
    if (!c) return DGH_ERR_ARG;
    SipgDev* S = sipg_get(c, h);
    if (!S) return c->last;
    FieldDev* fu = scalar_field(c, *S, u);
    FieldDev* fo = scalar_field(c, *S, out);
    if (!fu || !fo) return c->last;
    if (u == out) return c->fail(DGH_ERR_ARG, "dgh_sipg_apply: u and out must differ"), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (prec == DGH_PREC_FP64)
        return c->hip(sipg_apply<double>(S->N, sipg_args(*S, S->d), fu->d, S->d.trace, fo->d, 0, c->stream), "apply")
                   ? DGH_OK
                   : c->last;
    if (prec != DGH_PREC_FP32) return c->fail(DGH_ERR_UNSUPPORTED, "dgh_sipg_apply: precision"), DGH_ERR_UNSUPPORTED;
    if (!ensure32(c, *S)) return c->last;
    const bool ok = c->hip(vec_d2f(S->vs, S->s.x, fu->d, c->stream), "d2f") &&
                    c->hip(sipg_apply<float>(S->N, sipg_args(*S, S->s), S->s.x, S->s.trace, S->s.w, 0, c->stream),
                           "apply32") &&
                    c->hip(vec_zero(fo->d, fo->n * sizeof(double), c->stream), "zero") &&
                    c->hip(vec_add_f32(S->vs, fo->d, S->s.w, c->stream), "f2d");
    return ok ? DGH_OK : c->last;
}

int dgh_sipg_solve(dgh_ctx c, int h, int b, int u, dgh_sipg_stats* stats) {
    if (!c) return DGH_ERR_ARG;
    SipgDev* S = sipg_get(c, h);
    if (!S) return c->last;
    FieldDev* fb = scalar_field(c, *S, b);
    FieldDev* fu = scalar_field(c, *S, u);
    if (!fb || !fu) return c->last;
    if (b == u) return c->fail(DGH_ERR_ARG, "dgh_sipg_solve: b and u must differ"), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    const dgh_sipg_desc& d = S->desc;
    dgh_sipg_stats out{};
    if (!c->hip(hipEventRecord(c->ev0, c->stream), "event")) return c->last;
    double bnorm = 0.0;
    if (!device_norm<double>(c, *S, fb->d, bnorm)) return c->last;
    // the SPD system S u = -b, S = -A
    if (!c->hip(vec_neg_live(S->vs, S->d.rhs, fb->d, c->stream), "neg")) return c->last;
    if (d.mode == DGH_SOLVE_FP64) {
        CgState r{};
        // the solve runs on the solver's own x (the captured iteration's argument), from and back to u
        const size_t bytes = S->vs.n * sizeof(double);
        if (!c->hip(hipMemcpyAsync(S->d.x, fu->d, bytes, hipMemcpyDeviceToDevice, c->stream), "copy u") ||
            !run_pcg<double>(c, *S, S->d, d.rtol, d.abstol, d.maxit, r) ||
            !c->hip(hipMemcpyAsync(fu->d, S->d.x, bytes, hipMemcpyDeviceToDevice, c->stream), "copy u"))
            return c->last;
        out.iters = r.its;
        out.reason = r.reason;
        out.rnorm0 = r.rnorm0;
        out.rnorm = r.dp;
    } else {
        if (!ensure32(c, *S)) return c->last;
        const int max_outer = d.max_outer > 0 ? d.max_outer : 20;
        // The FP32 operator's rounding bounds what one sweep can take off the true residual (about eps32 times
        // the condition number), so a sweep asks only a modest reduction before the FP64 residual is formed
        // again; each sweep restarts PCG from e = 0.
        const double inner = d.inner_rtol > 0.0 ? d.inner_rtol : 1e-3;
        int total = 0;
        out.reason = -3;
        for (int k = 0; k <= max_outer; ++k) {
            // FP64 residual of the current iterate in the S system: r = -b - S u (|r| = |b - A u|)
            double rn = 0.0;
            if (!true_residual(c, *S, fu->d, rn)) return c->last;
            const double rel = bnorm > 0.0 ? rn / bnorm : rn;
            out.relres = rel;
            if (k == 0) out.rnorm0 = rn;
            out.rnorm = rn;
            if (rel <= d.rtol || rn <= d.abstol) {
                out.reason = 1;
                break;
            }
            if (k == max_outer || total >= d.maxit) break;
            // correction: S e = r in FP32, from e = 0
            CgState ir{};
            if (!c->hip(vec_d2f(S->vs, S->s.rhs, S->d.r, c->stream), "d2f") ||
                !c->hip(vec_zero(S->s.x, S->vs.n * sizeof(float), c->stream), "zero") ||
                !run_pcg<float>(c, *S, S->s, inner, 0.0, d.maxit - total, ir))
                return c->last;
            total += ir.its;
            out.outer = k + 1;
            if (ir.reason == -9 || ir.reason == -8) {
                out.reason = ir.reason;
                break;
            }
            if (!c->hip(vec_add_f32(S->vs, fu->d, S->s.x, c->stream), "update")) return c->last;
        }
        out.iters = total;
    }
    if (!c->hip(hipEventRecord(c->ev1, c->stream), "event") || !c->hip(hipStreamSynchronize(c->stream), "sync"))
        return c->last;
    float ms = 0.0f;
    if (!c->hip(hipEventElapsedTime(&ms, c->ev0, c->ev1), "elapsed")) return c->last;
    out.ms = ms;
    if (d.mode == DGH_SOLVE_FP64) {
        double rn = 0.0;
        if (!true_residual(c, *S, fu->d, rn)) return c->last;
        out.relres = bnorm > 0.0 ? rn / bnorm : rn;
    }
    if (stats) *stats = out;
    if (out.reason < 0)
        return c->fail(DGH_ERR_NOT_CONVERGED, "dgh_sipg_solve: stopped with reason %d after %d iterations", out.reason,
                       out.iters),
               DGH_ERR_NOT_CONVERGED;
    return DGH_OK;
}

}  // extern "C"

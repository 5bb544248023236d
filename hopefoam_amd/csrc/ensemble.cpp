// Host side of the LDG advection-diffusion ensembles (dgh_ensemble_*, include/dgh.h): per-case periodic
// meshes and geometry, the work-item list with its XCD ranges, the autotuner of the row split per order,
// and the time loop, eager or captured once into a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "dg_common.hpp"
#include "dgh.h"
#include "dgh_internal.hpp"
#include "ensemble.hpp"
#include "mesh.hpp"

using namespace dg;

namespace dg {

static constexpr const double* kA = kLserkA;
static constexpr const double* kB = kLserkB;

struct EnsHostCase {
    dgh_ens_case d;
    int N = 0, Np = 0, K = 0, Kp = 0;
};

struct EnsDev {
    int device = 0;
    std::vector<EnsHostCase> hc;
    std::vector<EnsCase> cases;
    EnsCase* dcases = nullptr;
    double* data = nullptr;
    double* geo = nullptr;
    int* conn = nullptr;
    double* ops[9] = {nullptr};
    double* opsM[9] = {nullptr};
    EnsItem* items = nullptr;
    int nitems = 0;
    int chunk[9][9] = {};       // per order: the 8 XCD ranges of its items (absolute item indices)
    int grid[9] = {0};          // per order: workgroups of its kernels (0: order absent)
    int split[9] = {0};         // rows split P per order (kEnsMfma: the MFMA variant)
    double tune_us[9][5] = {};  // autotuner: microseconds per stage (both passes) of each order's cases,
                                // P = 1, 2, 4, 8, MFMA
    int cur = ENS_U0;
    size_t ndata = 0;
    long long dofs = 0;
    std::map<long long, hipGraphExec_t> graphs;   // key: nsteps * 2 + starting buffer
    ~EnsDev() {
        (void)hipSetDevice(device);
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
        for (void* p : {(void*)dcases, (void*)data, (void*)geo, (void*)conn, (void*)items})
            if (p) (void)hipFree(p);
        for (double* p : ops)
            if (p) (void)hipFree(p);
        for (double* p : opsM)
            if (p) (void)hipFree(p);
    }
};

}  // namespace dg

static constexpr int ens_np_host(int N) { return (N + 1) * (N + 2) / 2; }

// device copy of a host table of plain structs
template <typename T>
static bool upload_table(dgh_ctx c, T** dptr, const std::vector<T>& v) {
    if (!c->hip(hipMalloc(reinterpret_cast<void**>(dptr), std::max<size_t>(1, v.size()) * sizeof(T)), "hipMalloc"))
        return false;
    return v.empty() ||
           c->hip(hipMemcpy(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

static constexpr int kSplits[5] = {1, 2, 4, 8, kEnsMfma};

// default tile shape: MFMA tiles from order 4 up, else the fewest slices with at most 12 rows per lane
static int default_split(int N) {
    if (N >= 4) return kEnsMfma;
    for (int P : {1, 2, 4, 8})
        if (ens_split_supported(N, P) && (ens_np_host(N) + P - 1) / P <= 12) return P;
    for (int P : {1, 2, 4, 8})
        if (ens_split_supported(N, P)) return P;
    return 0;
}

// items of the cases selected by `use` (all when empty), grouped by order and in case order within an order;
// per order, the 8 XCD ranges of equal estimated cost (rows per lane times (Np + 3 Nfp) operator columns)
static void build_items(const EnsDev& E, const std::vector<int>& use, const int* split, std::vector<EnsItem>& items,
                        int (*chunk)[9]) {
    items.clear();
    for (int N = 0; N < 9; ++N)
        for (int x = 0; x < 9; ++x) chunk[N][x] = 0;
    for (int N = 1; N <= 8; ++N) {
        std::vector<double> cost;
        const int first = int(items.size());
        for (int c = 0; c < int(E.hc.size()); ++c) {
            const EnsHostCase& h = E.hc[c];
            if (h.N != N || (!use.empty() && !std::binary_search(use.begin(), use.end(), c))) continue;
            const int P = split[N];
            if (P == kEnsMfma) {   // groups of 16 elements, one per wave
                const int nb = kEnsWaves, nblk = h.Kp / 16;
                for (int b = 0; b < nblk; b += nb) {
                    EnsItem it;
                    it.c = c;
                    it.b0 = b;
                    it.nb = nb;
                    it.P = P;
                    items.push_back(it);
                    cost.push_back(double(std::min(nb, nblk - b)));
                }
                continue;
            }
            const int R = (h.Np + P - 1) / P;
            const int nb = std::max(1, kEnsWaves / P), nblk = h.Kp / kEnsLanes;   // nb blocks of P slices
            for (int b = 0; b < nblk; b += nb) {
                EnsItem it;
                it.c = c;
                it.b0 = b;
                it.nb = nb;
                it.P = P;
                items.push_back(it);
                cost.push_back(double(std::min(nb, nblk - b)) * P * R * (h.Np + 3 * (h.N + 1)));
            }
        }
        const int n = int(items.size()) - first;
        double total = 0.0;
        for (double x : cost) total += x;
        int* ch = chunk[N];
        ch[0] = first;
        double run = 0.0;
        int x = 1;
        for (int i = 0; i < n && x < 8; ++i) {
            run += cost[i];
            while (x < 8 && run >= total * x / 8.0) ch[x++] = first + i + 1;
        }
        while (x < 8) ch[x++] = first + n;
        ch[8] = first + n;
    }
}

static bool upload_items(dgh_ctx c, EnsDev& E, const std::vector<EnsItem>& items) {
    if (!c->hip(hipStreamSynchronize(c->stream), "sync")) return false;   // no launch may still read the old list
    if (E.items) (void)hipFree(E.items);
    E.items = nullptr;
    E.nitems = int(items.size());
    return upload_table(c, &E.items, items);
}

static EnsArgs ens_args(const EnsDev& E) {
    EnsArgs a;
    a.cases = E.dcases;
    a.items = E.items;
    a.data = E.data;
    a.geo = E.geo;
    a.conn = E.conn;
    for (int i = 0; i < 9; ++i) a.ops[i] = E.ops[i];
    for (int i = 0; i < 9; ++i) a.opsM[i] = E.opsM[i];
    return a;
}

// workgroups of each order's kernels: as many as stay resident, but no more than its items
static void ens_grids(dgh_ctx c, EnsDev& E) {
    for (int N = 1; N <= 8; ++N) {
        const int n = E.chunk[N][8] - E.chunk[N][0];
        if (n == 0) {
            E.grid[N] = 0;
            continue;
        }
        const int resident = std::max(8, (ens_blocks_per_cu(N, E.split[N]) * c->ncu) / 8 * 8);
        E.grid[N] = std::min(resident, std::max(8, (n + 7) / 8 * 8));
    }
}

// one LSERK4 stage (mode 1) or the right-hand side into res (mode 0), reading u buffer `cur`: per order,
// the q pass then the right-hand-side pass over that order's cases (cases of different orders are independent)
static bool ens_stage(dgh_ctx c, EnsDev& E, int cur, int mode, double ra, double rb) {
    EnsArgs a = ens_args(E);
    a.cur = cur;
    a.mode = mode;
    a.rk_a = ra;
    a.rk_b = rb;
    for (int N = 1; N <= 8; ++N) {
        if (!E.grid[N]) continue;
        for (int x = 0; x < 9; ++x) a.chunk[x] = E.chunk[N][x];
        if (!c->hip(launch_ens_q(N, E.split[N], a, E.grid[N], c->stream), "ensemble q pass") ||
            !c->hip(launch_ens_rhs(N, E.split[N], a, E.grid[N], c->stream), "ensemble rhs pass"))
            return false;
    }
    return true;
}

// autotuner: per order, the row split with the fastest stage over that order's cases
static bool ens_tune(dgh_ctx c, EnsDev& E) {
    std::map<int, std::vector<int>> byN;
    for (int i = 0; i < int(E.hc.size()); ++i) byN[E.hc[i].N].push_back(i);
    std::vector<EnsItem> items;
    for (auto& kv : byN) {
        const int N = kv.first;
        double best = 1e300;
        int bestP = E.split[N];
        for (int pi = 0; pi < 5; ++pi) {
            const int P = kSplits[pi];
            E.tune_us[N][pi] = 0.0;
            if (!ens_split_supported(N, P)) continue;
            E.split[N] = P;   // the split the items and the launched kernels use
            build_items(E, kv.second, E.split, items, E.chunk);
            if (!upload_items(c, E, items)) return false;
            ens_grids(c, E);
            if (!ens_stage(c, E, E.cur, 0, 0.0, 0.0)) return false;   // warm-up
            if (!c->hip(hipEventRecord(c->ev0, c->stream), "event")) return false;
            constexpr int reps = 5;
            for (int r = 0; r < reps; ++r)
                if (!ens_stage(c, E, E.cur, 0, 0.0, 0.0)) return false;
            float ms = 0.f;
            if (!c->hip(hipEventRecord(c->ev1, c->stream), "event") || !c->hip(hipEventSynchronize(c->ev1), "sync") ||
                !c->hip(hipEventElapsedTime(&ms, c->ev0, c->ev1), "elapsed"))
                return false;
            const double us = 1000.0 * ms / reps;
            E.tune_us[N][pi] = us;
            if (us < best) {
                best = us;
                bestP = P;
            }
        }
        E.split[N] = bestP;
    }
    return true;
}

static bool ens_finalize_items(dgh_ctx c, EnsDev& E) {
    std::vector<EnsItem> items;
    build_items(E, {}, E.split, items, E.chunk);
    if (!upload_items(c, E, items)) return false;
    ens_grids(c, E);
    for (auto& g : E.graphs) (void)hipGraphExecDestroy(g.second);
    E.graphs.clear();
    return true;
}

static EnsDev* ens_get(dgh_ctx c, int h) {
    if (h < 0 || h >= int(c->ens.size()) || !c->ens[h]) {
        c->fail(DGH_ERR_ARG, "bad ensemble handle %d", h);
        return nullptr;
    }
    return c->ens[h].get();
}

extern "C" int dgh_ensemble_create(dgh_ctx c, const dgh_ens_desc* d, int ncases, const dgh_ens_case* cs, int* out) {
    if (!c) return DGH_ERR_ARG;
    if (!d || !cs || !out || ncases < 1) return c->fail(DGH_ERR_ARG, "dgh_ensemble_create: bad arguments"), DGH_ERR_ARG;
    if (!(d->tau >= 0.0) || !std::isfinite(d->bx) || !std::isfinite(d->by))
        return c->fail(DGH_ERR_ARG, "dgh_ensemble_create: tau >= 0 and a finite switch vector"), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    auto E = std::make_shared<EnsDev>();
    E->device = c->device;
    E->hc.resize(ncases);
    size_t nd = 0, ng = 0, nc = 0;
    for (int i = 0; i < ncases; ++i) {
        const dgh_ens_case& x = cs[i];
        if (x.N < 1 || x.N > 8 || x.nx < 3 || x.ny < 3 || !(x.Lx > 0) || !(x.Ly > 0) || !(x.nu >= 0) || !(x.dt > 0) ||
            !std::isfinite(x.ax) || !std::isfinite(x.ay))
            return c->fail(DGH_ERR_ARG, "dgh_ensemble_create: case %d invalid (N 1..8, nx, ny >= 3, L > 0, nu >= 0, "
                                        "dt > 0)", i),
                   DGH_ERR_ARG;
        EnsHostCase& h = E->hc[i];
        h.d = x;
        h.N = x.N;
        h.Np = ens_np_host(x.N);
        h.K = 2 * x.nx * x.ny;
        h.Kp = (h.K + kEnsLanes - 1) / kEnsLanes * kEnsLanes;
        EnsCase ec;
        ec.N = h.N;
        ec.K = h.K;
        ec.Kp = h.Kp;
        ec.off_u = nd;
        ec.off_geo = ng;
        ec.off_conn = nc;
        ec.ax = x.ax;
        ec.ay = x.ay;
        ec.nu = x.nu;
        ec.dt = x.dt;
        E->cases.push_back(ec);
        nd += size_t(kEnsArrays) * h.Kp * h.Np;
        ng += size_t(kEnsGeo) * h.Kp;
        nc += size_t(3) * h.Kp;
        E->dofs += (long long)h.K * h.Np;
    }
    E->ndata = nd;
    // geometry and connectivity of every case (periodic split-quad meshes)
    std::vector<double> geo(ng, 0.0);
    std::vector<int> conn(nc, 0);
    for (int i = 0; i < ncases; ++i) {
        const EnsHostCase& h = E->hc[i];
        std::vector<double> vxy;
        std::vector<int> vid, bct;
        structured_split_quads(h.d.nx, h.d.ny, h.d.x0, h.d.x0 + h.d.Lx, h.d.y0, h.d.y0 + h.d.Ly, true, true, 0, vxy,
                               vid, bct);
        HostMesh m;
        if (!build_mesh(h.N, h.K, vxy.data(), vid.data(), bct.data(), m) || m.nbfaces != 0)
            return c->fail(DGH_ERR_ARG, "dgh_ensemble_create: case %d mesh", i), DGH_ERR_ARG;
        double* g = &geo[E->cases[i].off_geo];
        int* cn = &conn[E->cases[i].off_conn];
        for (int e = 0; e < h.Kp; ++e) {
            const int ec = std::min(e, h.K - 1);
            const double* mg = &m.geo[size_t(ec) * 14];
            for (int s = 0; s < 4; ++s) g[size_t(s) * h.Kp + e] = mg[s];
            for (int f = 0; f < 3; ++f) {
                const int en = m.EToE[ec * 3 + f], fn = m.EToF[ec * 3 + f];
                const double nx = mg[5 + 3 * f], ny = mg[6 + 3 * f];
                const double fs = mg[7 + 3 * f] / mg[4];
                const double fsn = m.geo[size_t(en) * 14 + 7 + 3 * fn] / m.geo[size_t(en) * 14 + 4];
                g[size_t(4 + 4 * f) * h.Kp + e] = nx;
                g[size_t(5 + 4 * f) * h.Kp + e] = ny;
                g[size_t(6 + 4 * f) * h.Kp + e] = fs;
                g[size_t(7 + 4 * f) * h.Kp + e] = d->tau * std::max(fs, fsn);
                const double sgn = d->bx * nx + d->by * ny;
                const bool own = sgn > 0.0 || (sgn == 0.0 && ec < en);
                cn[size_t(f) * h.Kp + e] = (en << 3) | (fn << 1) | (own ? 1 : 0);
            }
        }
    }
    // operators per order: Dr^T, Ds^T, LIFT^T with rows padded to kEnsRows
    for (const EnsHostCase& h : E->hc) {
        if (E->ops[h.N]) continue;
        RefDev* rd = c->refdev(h.N);
        if (!rd) return c->last;
        const RefElem& re = rd->re;
        const int Np = re.Np, Nfp = re.Nfp;
        std::vector<double> op(size_t(2 * Np + 3 * Nfp) * kEnsRows, 0.0);
        for (int j = 0; j < Np; ++j)
            for (int i = 0; i < Np; ++i) {
                op[size_t(j) * kEnsRows + i] = re.Dr[size_t(i) * Np + j];
                op[size_t(Np + j) * kEnsRows + i] = re.Ds[size_t(i) * Np + j];
            }
        for (int col = 0; col < 3 * Nfp; ++col)
            for (int i = 0; i < Np; ++i) op[size_t(2 * Np + col) * kEnsRows + i] = re.LiftN[size_t(i) * 3 * Nfp + col];
        // MFMA A fragments: tile m, K step s, lane i + 16 k holds the entry (16 m + i, 4 s + k)
        const int MT = ens_mfma_mt(h.N), KS = ens_mfma_ks(h.N), KSL = ens_mfma_ksl(h.N);
        std::vector<double> fr(size_t(2 * MT * KS + MT * KSL) * 64, 0.0);
        for (int m = 0; m < MT; ++m)
            for (int l = 0; l < 64; ++l) {
                const int i = 16 * m + (l & 15), k = l >> 4;
                for (int st = 0; st < KS; ++st) {
                    const int j = 4 * st + k;
                    if (i < Np && j < Np) {
                        fr[size_t(m * KS + st) * 64 + l] = re.Dr[size_t(i) * Np + j];
                        fr[size_t(MT * KS + m * KS + st) * 64 + l] = re.Ds[size_t(i) * Np + j];
                    }
                }
                for (int st = 0; st < KSL; ++st) {
                    const int j = 4 * st + k;
                    if (i < Np && j < 3 * Nfp)
                        fr[size_t(2 * MT * KS + m * KSL + st) * 64 + l] = re.LiftN[size_t(i) * 3 * Nfp + j];
                }
            }
        if (!c->upload(&E->ops[h.N], op) || !c->upload(&E->opsM[h.N], fr)) return c->last;
        E->split[h.N] = default_split(h.N);
    }
    if (!upload_table(c, &E->dcases, E->cases) || !c->upload(&E->geo, geo) || !c->upload(&E->conn, conn) ||
        !c->hip(hipMalloc(reinterpret_cast<void**>(&E->data), nd * sizeof(double)), "hipMalloc ensemble") ||
        !c->hip(hipMemsetAsync(E->data, 0, nd * sizeof(double), c->stream), "memset"))
        return c->last;
    if (d->autotune && !ens_tune(c, *E)) return c->last;
    if (!ens_finalize_items(c, *E) || !c->hip(hipStreamSynchronize(c->stream), "sync")) return c->last;
    c->ens.push_back(E);
    *out = int(c->ens.size()) - 1;
    return DGH_OK;
}

extern "C" int dgh_ensemble_destroy(dgh_ctx c, int h) {
    if (!c) return DGH_ERR_ARG;
    if (!ens_get(c, h)) return c->last;
    (void)hipStreamSynchronize(c->stream);
    c->ens[h].reset();
    return DGH_OK;
}

// host [K][Np] <-> device group layout of one nodal array of case i
static int ens_copy(dgh_ctx c, EnsDev& E, int i, int array, double* host, bool to_device) {
    const EnsHostCase& h = E.hc[i];
    const size_t n = size_t(h.Kp) * h.Np;
    std::vector<double> buf(n, 0.0);
    double* dptr = E.data + E.cases[i].off_u + size_t(array) * n;
    if (to_device) {
        for (int e = 0; e < h.Kp; ++e)
            for (int j = 0; j < h.Np; ++j)
                buf[dev_index(e, 0, j, 1, h.Np)] = host[size_t(std::min(e, h.K - 1)) * h.Np + j];
        if (!c->hip(hipMemcpyAsync(dptr, buf.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream), "upload") ||
            !c->hip(hipStreamSynchronize(c->stream), "sync"))
            return c->last;
    } else {
        if (!c->hip(hipMemcpyAsync(buf.data(), dptr, n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "download") ||
            !c->hip(hipStreamSynchronize(c->stream), "sync"))
            return c->last;
        for (int e = 0; e < h.K; ++e)
            for (int j = 0; j < h.Np; ++j) host[size_t(e) * h.Np + j] = buf[dev_index(e, 0, j, 1, h.Np)];
    }
    return DGH_OK;
}

extern "C" int dgh_ensemble_set(dgh_ctx c, int h, int i, const double* u) {
    if (!c) return DGH_ERR_ARG;
    EnsDev* E = ens_get(c, h);
    if (!E) return c->last;
    if (i < 0 || i >= int(E->hc.size()) || !u) return c->fail(DGH_ERR_ARG, "dgh_ensemble_set: case %d", i), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    return ens_copy(c, *E, i, E->cur, const_cast<double*>(u), true);
}

extern "C" int dgh_ensemble_get(dgh_ctx c, int h, int i, int what, double* out) {
    if (!c) return DGH_ERR_ARG;
    EnsDev* E = ens_get(c, h);
    if (!E) return c->last;
    if (i < 0 || i >= int(E->hc.size()) || !out || what < 0 || what > 2)
        return c->fail(DGH_ERR_ARG, "dgh_ensemble_get: case %d, what %d", i, what), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (what == 0) return ens_copy(c, *E, i, E->cur, out, false);
    if (what == 1) return ens_copy(c, *E, i, ENS_RES, out, false);
    const size_t n = size_t(E->hc[i].K) * E->hc[i].Np;
    if (int s = ens_copy(c, *E, i, ENS_QX, out, false)) return s;
    return ens_copy(c, *E, i, ENS_QY, out + n, false);
}

extern "C" int dgh_ensemble_rhs(dgh_ctx c, int h) {
    if (!c) return DGH_ERR_ARG;
    EnsDev* E = ens_get(c, h);
    if (!E) return c->last;
    (void)hipSetDevice(c->device);
    if (!ens_stage(c, *E, E->cur, 0, 0.0, 0.0) || !c->hip(hipStreamSynchronize(c->stream), "sync")) return c->last;
    return DGH_OK;
}

extern "C" int dgh_ensemble_run(dgh_ctx c, int h, int nsteps, int use_graph, float* ms) {
    if (!c) return DGH_ERR_ARG;
    EnsDev* E = ens_get(c, h);
    if (!E) return c->last;
    if (nsteps < 0) return c->fail(DGH_ERR_ARG, "dgh_ensemble_run: nsteps < 0"), DGH_ERR_ARG;
    if (ms) *ms = 0.f;
    if (nsteps == 0) return DGH_OK;
    (void)hipSetDevice(c->device);
    const int start = E->cur;
    auto record = [&](int cur0) {
        int cur = cur0;
        for (int step = 0; step < nsteps; ++step)
            for (int s = 0; s < 5; ++s) {
                if (!ens_stage(c, *E, cur, 1, kA[s], kB[s])) return false;
                cur ^= 1;
            }
        return true;
    };
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
        const long long key = (long long)nsteps * 2 + start;
        auto it = E->graphs.find(key);
        if (it != E->graphs.end()) {
            exec = it->second;
        } else {
            hipGraph_t g = nullptr;
            if (!c->hip(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture"))
                return c->last;
            const bool ok = record(start);
            const bool ended = c->hip(hipStreamEndCapture(c->stream, &g), "end capture");
            if (!ok || !ended) {
                if (g) (void)hipGraphDestroy(g);
                return c->last;
            }
            const bool inst = c->hip(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), "graph instantiate");
            (void)hipGraphDestroy(g);
            if (!inst) return c->last;
            E->graphs[key] = exec;
        }
    }
    if (!c->hip(hipEventRecord(c->ev0, c->stream), "event")) return c->last;
    if (use_graph) {
        if (!c->hip(hipGraphLaunch(exec, c->stream), "graph launch")) return c->last;
    } else if (!record(start)) {
        return c->last;
    }
    if (!c->hip(hipEventRecord(c->ev1, c->stream), "event") || !c->hip(hipEventSynchronize(c->ev1), "sync"))
        return c->last;
    if (ms && !c->hip(hipEventElapsedTime(ms, c->ev0, c->ev1), "elapsed")) return c->last;
    E->cur = start ^ ((5 * nsteps) & 1);
    return DGH_OK;
}

extern "C" int dgh_ensemble_info(dgh_ctx c, int h, int* split, double* tune_us, long long* dofs, int* nitems) {
    if (!c) return DGH_ERR_ARG;
    EnsDev* E = ens_get(c, h);
    if (!E) return c->last;
    if (split)
        for (int i = 0; i < 9; ++i) split[i] = E->split[i];
    if (tune_us)
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 5; ++j) tune_us[i * 5 + j] = E->tune_us[i][j];
    if (dofs) *dofs = E->dofs;
    if (nitems) *nitems = E->nitems;
    return DGH_OK;
}

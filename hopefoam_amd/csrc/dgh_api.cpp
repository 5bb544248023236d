// C ABI of libdghip.so (declared in include/dgh.h): contexts, meshes, fields and the
// explicit DG stage.  Host code is C++17; device work goes to hand-written HIP kernels.
#include "dgh.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "dg_common.hpp"
#include "dgh_internal.hpp"
#include "limiter.hpp"
#include "mesh.hpp"
#include "refelem.hpp"
#include "stage_kernels.hpp"

using namespace dg;

namespace {

template <typename T>
std::vector<T> cast_vec(const std::vector<double>& v) {
    return std::vector<T>(v.begin(), v.end());
}

// Operator fragments in the per-lane order of the 16x16x4 MFMA operands (see stage_kernels.hip);
// PF depends on the accumulator row map of the precision it is built for.
template <typename T>
void build_fragments(const RefElem& re, std::vector<double>& VqF, std::vector<double>& PF, std::vector<double>& LF) {
    const int Np = re.Np, Nq = re.Nq, Nfq = re.Nfq;
    const int MT = round_up(Np, 16) / 16, QT = round_up(Nq, 16) / 16;
    const int KS1 = round_up(Np, 4) / 4, SF = round_up(Nfq, 4) / 4, KSF = 3 * SF;
    const int FT = round_up(KSF, 4) / 4;
    VqF.assign(size_t(QT + FT) * KS1 * 64, 0.0);
    for (int t = 0; t < QT; ++t)
        for (int s = 0; s < KS1; ++s)
            for (int l = 0; l < 64; ++l) {
                const int i = 16 * t + (l & 15), k = gemm1_node<T>(s, l >> 4);
                if (i < Nq && k < Np) VqF[(size_t(t) * KS1 + s) * 64 + l] = re.Vq[size_t(i) * Np + k];
            }
    // face-trace tiles: accumulator row i of tile tf is held by (row group rg, register r); it is the own
    // trace of lift K step 4 tf + r, Gauss point 4 ((4 tf + r) % SF) + rg of face (4 tf + r) / SF
    for (int tf = 0; tf < FT; ++tf)
        for (int s1 = 0; s1 < KS1; ++s1)
            for (int l = 0; l < 64; ++l) {
                const int i = l & 15, node = gemm1_node<T>(s1, l >> 4);
                const int rg = mfma_rg<T>(i), r = mfma_reg<T>(i);
                const int st = 4 * tf + r, g = 4 * (st % SF) + rg, f = st / SF;
                if (st >= KSF || g >= Nfq || node >= Np) continue;
                double w = 0.0;
                for (int k = 0; k < re.Nfp; ++k)
                    if (re.fmask[f * re.Nfp + k] == node) w = re.Vf[size_t(g) * re.Nfp + k];
                VqF[(size_t(QT + tf) * KS1 + s1) * 64 + l] = w;
            }
    PF.assign(size_t(QT) * 4 * MT * 2 * 64, 0.0);
    for (int t = 0; t < QT; ++t)
        for (int r = 0; r < 4; ++r)
            for (int m = 0; m < MT; ++m)
                for (int w = 0; w < 2; ++w)
                    for (int l = 0; l < 64; ++l) {
                        const int j = 16 * m + (l & 15), p = 16 * t + mfma_row<T>(l >> 4, r);
                        const std::vector<double>& P = w == 0 ? re.Pr : re.Ps;
                        if (j < Np && p < Nq)
                            PF[((((size_t(t) * 4 + r) * MT + m) * 2 + w) * 64) + l] = P[size_t(j) * Nq + p];
                    }
    // face items: K step s = f * SF + j covers Gauss points g = 4 j + k of face f (zero when g >= Nfq)
    LF.assign(size_t(KSF) * MT * 64, 0.0);
    for (int s = 0; s < KSF; ++s)
        for (int m = 0; m < MT; ++m)
            for (int l = 0; l < 64; ++l) {
                const int j = 16 * m + (l & 15), f = s / SF, g = 4 * (s % SF) + (l >> 4);
                if (j < Np && g < Nfq) LF[(size_t(s) * MT + m) * 64 + l] = re.Lift[size_t(j) * 3 * Nfq + f * Nfq + g];
            }
}

// bf16 with round-to-nearest-even of a finite float
uint16_t bf16_rne(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    b += 0x7fffu + ((b >> 16) & 1u);
    return uint16_t(b >> 16);
}

float bf16_value(uint16_t h) {
    const uint32_t b = uint32_t(h) << 16;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// One lane's bf16x3 fragment of four values: x = hi + lo, hi = bf16(x), lo = bf16(x - hi); words
// {hi0 | hi1 << 16, hi2 | hi3 << 16, lo0 | lo1 << 16, lo2 | lo3 << 16} (element 0 in the low half)
void put_split(std::vector<uint32_t>& v, size_t slot, const double (&x)[4]) {
    uint16_t h[4], l[4];
    for (int e = 0; e < 4; ++e) {
        h[e] = bf16_rne(float(x[e]));
        l[e] = bf16_rne(float(x[e] - double(bf16_value(h[e]))));
    }
    v[slot * 4 + 0] = h[0] | uint32_t(h[1]) << 16;
    v[slot * 4 + 1] = h[2] | uint32_t(h[3]) << 16;
    v[slot * 4 + 2] = l[0] | uint32_t(l[1]) << 16;
    v[slot * 4 + 3] = l[2] | uint32_t(l[3]) << 16;
}

// The operators of build_fragments<float> as bf16x3 fragments of v_mfma_f32_16x16x16_bf16: element e of K16
// step S is the FP32 kernel's K4 step 4 S + e (interpolation, lift) or register e of cubature tile t
// (projection), so the kernels feed the B operands from the same registers as the FP32 kernel.
void build_split_fragments(const RefElem& re, std::vector<uint32_t>& VqS, std::vector<uint32_t>& PS,
                           std::vector<uint32_t>& LS) {
    using T = float;
    const int Np = re.Np, Nq = re.Nq, Nfq = re.Nfq;
    const int MT = round_up(Np, 16) / 16, QT = round_up(Nq, 16) / 16;
    const int KS1 = round_up(Np, 4) / 4, SF = round_up(Nfq, 4) / 4, KSF = 3 * SF;
    const int FT = round_up(KSF, 4) / 4, KS16 = (KS1 + 3) / 4;
    double x[4];
    VqS.assign(size_t(QT + FT) * KS16 * 64 * 4, 0u);
    for (int t = 0; t < QT; ++t)
        for (int S = 0; S < KS16; ++S)
            for (int l = 0; l < 64; ++l) {
                const int i = 16 * t + (l & 15);
                for (int e = 0; e < 4; ++e) {
                    const int s = 4 * S + e, k = gemm1_node<T>(s, l >> 4);
                    x[e] = (s < KS1 && i < Nq && k < Np) ? re.Vq[size_t(i) * Np + k] : 0.0;
                }
                put_split(VqS, (size_t(t) * KS16 + S) * 64 + l, x);
            }
    for (int tf = 0; tf < FT; ++tf)
        for (int S = 0; S < KS16; ++S)
            for (int l = 0; l < 64; ++l) {
                const int i = l & 15;
                const int rg = mfma_rg<T>(i), r = mfma_reg<T>(i);
                const int st = 4 * tf + r, g = 4 * (st % SF) + rg, f = st / SF;
                for (int e = 0; e < 4; ++e) {
                    const int s1 = 4 * S + e, node = gemm1_node<T>(s1, l >> 4);
                    x[e] = 0.0;
                    if (s1 >= KS1 || st >= KSF || g >= Nfq || node >= Np) continue;
                    for (int k = 0; k < re.Nfp; ++k)
                        if (re.fmask[f * re.Nfp + k] == node) x[e] = re.Vf[size_t(g) * re.Nfp + k];
                }
                put_split(VqS, (size_t(QT + tf) * KS16 + S) * 64 + l, x);
            }
    PS.assign(size_t(QT) * MT * 2 * 64 * 4, 0u);
    for (int t = 0; t < QT; ++t)
        for (int m = 0; m < MT; ++m)
            for (int w = 0; w < 2; ++w)
                for (int l = 0; l < 64; ++l) {
                    const std::vector<double>& P = w == 0 ? re.Pr : re.Ps;
                    const int j = 16 * m + (l & 15);
                    for (int e = 0; e < 4; ++e) {
                        const int p = 16 * t + mfma_row<T>(l >> 4, e);
                        x[e] = (j < Np && p < Nq) ? P[size_t(j) * Nq + p] : 0.0;
                    }
                    put_split(PS, ((size_t(t) * MT + m) * 2 + w) * 64 + l, x);
                }
    LS.assign(size_t(FT) * MT * 64 * 4, 0u);
    for (int S = 0; S < FT; ++S)
        for (int m = 0; m < MT; ++m)
            for (int l = 0; l < 64; ++l) {
                const int j = 16 * m + (l & 15);
                for (int e = 0; e < 4; ++e) {
                    const int st = 4 * S + e, f = st / SF, g = 4 * (st % SF) + (l >> 4);
                    x[e] = (st < KSF && j < Np && g < Nfq) ? re.Lift[size_t(j) * 3 * Nfq + f * Nfq + g] : 0.0;
                }
                put_split(LS, (size_t(S) * MT + m) * 64 + l, x);
            }
}

}  // namespace

// ------------------------------------------------------------------------------------------
bool dgh_ctx_s::fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    last = code;
    return false;
}

bool dgh_ctx_s::hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    return fail(DGH_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

template <typename T>
bool dgh_ctx_s::upload(T** dptr, const std::vector<T>& v) {
    if (!hip(hipMalloc(reinterpret_cast<void**>(dptr), std::max<size_t>(1, v.size()) * sizeof(T)), "hipMalloc"))
        return false;
    if (v.empty()) return true;
    return hip(hipMemcpy(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

template bool dgh_ctx_s::upload<double>(double**, const std::vector<double>&);
template bool dgh_ctx_s::upload<float>(float**, const std::vector<float>&);
template bool dgh_ctx_s::upload<int>(int**, const std::vector<int>&);
template bool dgh_ctx_s::upload<uint32_t>(uint32_t**, const std::vector<uint32_t>&);

RefDev* dgh_ctx_s::refdev(int N) {
    auto it = refs.find(N);
    if (it != refs.end()) return it->second.get();
    auto rd = std::make_unique<RefDev>();
    if (!build_refelem(N, rd->re)) {
        fail(DGH_ERR_UNSUPPORTED, "order N=%d not supported (1..8)", N);
        return nullptr;
    }
    std::vector<double> VqF, PF, LF, VqF32, PF32, LF32;
    std::vector<uint32_t> VqS, PS, LS;
    build_fragments<double>(rd->re, VqF, PF, LF);
    build_fragments<float>(rd->re, VqF32, PF32, LF32);
    build_split_fragments(rd->re, VqS, PS, LS);
    bool ok = upload(&rd->VqF64, VqF) && upload(&rd->PF64, PF) && upload(&rd->LF64, LF) &&
              upload(&rd->Vf64, rd->re.Vf) && upload(&rd->VqF32, cast_vec<float>(VqF32)) &&
              upload(&rd->PF32, cast_vec<float>(PF32)) && upload(&rd->LF32, cast_vec<float>(LF32)) &&
              upload(&rd->Vf32, cast_vec<float>(rd->re.Vf)) && upload(&rd->fmask, rd->re.fmask) &&
              upload(&rd->VqS, VqS) && upload(&rd->PS, PS) && upload(&rd->LS, LS);
    if (!ok) return nullptr;
    RefDev* p = rd.get();
    refs[N] = std::move(rd);
    return p;
}

MeshDev* dgh_ctx_s::mesh(int h) {
    if (h < 0 || h >= int(meshes.size()) || !meshes[h]) {
        fail(DGH_ERR_ARG, "invalid mesh handle %d", h);
        return nullptr;
    }
    return meshes[h].get();
}

FieldDev* dgh_ctx_s::field(int h) {
    if (h < 0 || h >= int(fields.size()) || !fields[h]) {
        fail(DGH_ERR_ARG, "invalid field handle %d", h);
        return nullptr;
    }
    return fields[h].get();
}

static void dfree(void* p) {
    if (p) (void)hipFree(p);
}
MeshDev::~MeshDev() {
    dfree(geo64); dfree(geo32); dfree(conn); dfree(bcq); dfree(bc_xy); dfree(bc_slot); dfree(patch_first);
    for (auto& w : ws) { dfree(w.second.res); dfree(w.second.tmp); }
}
RefDev::~RefDev() {
    dfree(VqF64); dfree(PF64); dfree(LF64); dfree(Vf64);
    dfree(VqF32); dfree(PF32); dfree(LF32); dfree(Vf32); dfree(fmask);
    dfree(VqS); dfree(PS); dfree(LS);
    dfree(insSr); dfree(insSs); dfree(insLift); dfree(insDr); dfree(insDs);
}
FieldDev::~FieldDev() { dfree(d); }

// ------------------------------------------------------------------------------------------
extern "C" {

int dgh_create(dgh_ctx* out, int device) {
    if (!out) return DGH_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return DGH_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return DGH_ERR_HIP;
    auto* c = new dgh_ctx_s();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return DGH_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->ncu = prop.multiProcessorCount;
    // this device's copy of the static scratch of dgh_limit_noalloc (device current: set above)
    if (dg::noalloc_scratch(device, &c->noalloc) != hipSuccess) {
        (void)hipEventDestroy(c->ev0);
        (void)hipEventDestroy(c->ev1);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return DGH_ERR_HIP;
    }
    *out = c;
    return DGH_OK;
}

int dgh_destroy(dgh_ctx c) {
    if (!c) return DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->fields.clear();
    c->meshes.clear();
    c->refs.clear();
    c->sipg.clear();
    c->mlps.clear();
    c->ens.clear();
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return DGH_OK;
}

const char* dgh_last_error(dgh_ctx c) { return c ? c->err.c_str() : "null context"; }

int dgh_sync(dgh_ctx c) {
    if (!c) return DGH_ERR_ARG;
    return c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize") ? DGH_OK : c->last;
}

void* dgh_stream(dgh_ctx c) { return c ? static_cast<void*>(c->stream) : nullptr; }

int dgh_set_launch_cap(dgh_ctx c, int blocks) {
    if (!c) return DGH_ERR_ARG;
    if (blocks < 0) return c->fail(DGH_ERR_ARG, "dgh_set_launch_cap: blocks must be >= 0"), DGH_ERR_ARG;
    c->stage_block_cap = blocks;
    return DGH_OK;
}

int dgh_mesh_create(dgh_ctx c, int N, int K, const double* vxy, const int* vid, const int* bctag, int* out) {
    if (!c || !vxy || !vid || !out || K < 1) return c ? (c->fail(DGH_ERR_ARG, "dgh_mesh_create: bad args"), DGH_ERR_ARG) : DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    RefDev* rd = c->refdev(N);
    if (!rd) return c->last;
    auto m = std::make_unique<MeshDev>();
    m->ref = rd;
    if (!build_mesh(N, K, vxy, vid, bctag, m->host))
        return c->fail(DGH_ERR_ARG, "dgh_mesh_create: non-manifold edge or degenerate/clockwise element"),
               DGH_ERR_ARG;
    const HostMesh& h = m->host;
    std::vector<double> geo(size_t(kGeoSlots) * K, 0.0);
    std::vector<int> conn(size_t(3) * K);
    for (int e = 0; e < K; ++e) {
        const double* g = &h.geo[size_t(e) * 14];
        geo[G_RX * size_t(K) + e] = g[0];
        geo[G_RY * size_t(K) + e] = g[1];
        geo[G_SX * size_t(K) + e] = g[2];
        geo[G_SY * size_t(K) + e] = g[3];
        geo[G_J * size_t(K) + e] = g[4];
        for (int f = 0; f < 3; ++f) {
            geo[(G_NX0 + 3 * f) * size_t(K) + e] = g[5 + 3 * f];
            geo[(G_NY0 + 3 * f) * size_t(K) + e] = g[6 + 3 * f];
            geo[(G_FS0 + 3 * f) * size_t(K) + e] = g[7 + 3 * f] / g[4];
            const int en = h.EToE[e * 3 + f];
            conn[size_t(f) * K + e] = en >= 0 ? en * 4 + h.EToF[e * 3 + f] : -1 - h.bctag[e * 3 + f];
        }
    }
    if (!c->upload(&m->geo64, geo) || !c->upload(&m->geo32, cast_vec<float>(geo)) || !c->upload(&m->conn, conn))
        return c->last;
    c->meshes.push_back(std::move(m));
    *out = int(c->meshes.size()) - 1;
    return DGH_OK;
}

int dgh_mesh_structured(dgh_ctx c, int N, int nx, int ny, double x0, double x1, double y0, double y1, int px,
                        int py, int bc_tag, int* out) {
    if (!c || nx < 1 || ny < 1 || !(x1 > x0) || !(y1 > y0)) return DGH_ERR_ARG;
    if ((px && nx < 3) || (py && ny < 3))
        return c->fail(DGH_ERR_ARG, "periodic directions need at least 3 quads"), DGH_ERR_ARG;
    std::vector<double> vxy;
    std::vector<int> vid, bct;
    structured_split_quads(nx, ny, x0, x1, y0, y1, px != 0, py != 0, bc_tag, vxy, vid, bct);
    return dgh_mesh_create(c, N, 2 * nx * ny, vxy.data(), vid.data(), bct.data(), out);
}

int dgh_mesh_info(dgh_ctx c, int mh, int* info) {
    if (!c || !info) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    const RefElem& re = m->ref->re;
    int v[7] = {re.N, re.Np, re.Nfp, re.Nq, re.Nfq, m->host.K, m->host.nbfaces};
    std::memcpy(info, v, sizeof v);
    return DGH_OK;
}

int dgh_mesh_nodes(dgh_ctx c, int mh, double* x, double* y) {
    if (!c || !x || !y) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    const RefElem& re = m->ref->re;
    for (int e = 0; e < m->host.K; ++e) {
        const double* v = &m->host.vxy[size_t(e) * 6];
        for (int i = 0; i < re.Np; ++i) {
            const double r = re.r[i], s = re.s[i];
            x[size_t(e) * re.Np + i] = -(r + s) * 0.5 * v[0] + (r + 1) * 0.5 * v[2] + (s + 1) * 0.5 * v[4];
            y[size_t(e) * re.Np + i] = -(r + s) * 0.5 * v[1] + (r + 1) * 0.5 * v[3] + (s + 1) * 0.5 * v[5];
        }
    }
    return DGH_OK;
}

int dgh_mesh_conn(dgh_ctx c, int mh, int* EToE, int* EToF) {
    if (!c || !EToE || !EToF) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    std::memcpy(EToE, m->host.EToE.data(), sizeof(int) * m->host.EToE.size());
    std::memcpy(EToF, m->host.EToF.data(), sizeof(int) * m->host.EToF.size());
    return DGH_OK;
}

int dgh_mesh_geo(dgh_ctx c, int mh, double* geo) {
    if (!c || !geo) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    std::memcpy(geo, m->host.geo.data(), sizeof(double) * m->host.geo.size());
    return DGH_OK;
}

int dgh_mesh_set_bc_values(dgh_ctx c, int mh, int nvar, const double* bcq) {
    if (!c || nvar < 1) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    (void)hipSetDevice(c->device);
    dfree(m->bcq);
    m->bcq = nullptr;
    m->bc_nvar = 0;
    m->bcfn = BcFunction{};   // explicit data replace a boundary function
    if (!bcq) return DGH_OK;
    const size_t n = size_t(nvar) * m->host.K * 3 * m->ref->re.Nfp;
    if (!c->upload(&m->bcq, std::vector<double>(bcq, bcq + n))) return c->last;
    m->bc_nvar = nvar;
    return DGH_OK;
}

// ghost states of the mesh's boundary function at time t, on the context stream
static bool eval_bc(dgh_ctx c, MeshDev* m, double t) {
    if (m->bcfn.kind == BCFN_NONE) return true;
    const size_t stride = size_t(m->host.K) * 3 * m->ref->re.Nfp;
    return c->hip(launch_bc_eval(m->bcfn, t, m->bc_n, m->bc_xy, m->bc_slot, stride, m->bcq, c->stream),
                  "boundary function");
}

int dgh_mesh_set_bc_function(dgh_ctx c, int mh, const dgh_bcfn_desc* fn, double t) {
    if (!c) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    (void)hipSetDevice(c->device);
    if (!fn || fn->kind == DGH_BCFN_NONE) {
        m->bcfn = BcFunction{};
        m->bc_time = t;
        return DGH_OK;
    }
    if (fn->kind != DGH_BCFN_ISENTROPIC_VORTEX && fn->kind != DGH_BCFN_DOUBLE_MACH)
        return c->fail(DGH_ERR_ARG, "dgh_mesh_set_bc_function: unknown kind %d", fn->kind), DGH_ERR_ARG;
    if (!(fn->gamma > 1.0))
        return c->fail(DGH_ERR_ARG, "dgh_mesh_set_bc_function: gamma must exceed 1"), DGH_ERR_ARG;
    const HostMesh& h = m->host;
    const RefElem& re = m->ref->re;
    const int Nfp = re.Nfp;
    if (!m->bc_xy) {   // the fixedValue face nodes (the owner's nodes, in its face-node order)
        std::vector<double> xy;
        std::vector<int> slot;
        for (int e = 0; e < h.K; ++e)
            for (int f = 0; f < 3; ++f) {
                if (h.EToE[e * 3 + f] >= 0 || h.bctag[e * 3 + f] != DGH_BC_FIXED_VALUE) continue;
                const double* v = &h.vxy[size_t(e) * 6];
                for (int k = 0; k < Nfp; ++k) {
                    const int i = re.fmask[f * Nfp + k];
                    const double r = re.r[i], s_ = re.s[i];
                    xy.push_back(-(r + s_) * 0.5 * v[0] + (r + 1) * 0.5 * v[2] + (s_ + 1) * 0.5 * v[4]);
                    xy.push_back(-(r + s_) * 0.5 * v[1] + (r + 1) * 0.5 * v[3] + (s_ + 1) * 0.5 * v[5]);
                    slot.push_back((e * 3 + f) * Nfp + k);
                }
            }
        if (!c->upload(&m->bc_xy, xy) || !c->upload(&m->bc_slot, slot)) return c->last;
        m->bc_n = int(slot.size());
    }
    if (!m->bcq || m->bc_nvar != 4) {
        dfree(m->bcq);
        m->bcq = nullptr;
        m->bc_nvar = 0;
        const size_t n = size_t(4) * h.K * 3 * Nfp;
        if (!c->hip(hipMalloc(reinterpret_cast<void**>(&m->bcq), n * sizeof(double)), "hipMalloc bcq") ||
            !c->hip(hipMemsetAsync(m->bcq, 0, n * sizeof(double), c->stream), "memset bcq"))
            return c->last;
        m->bc_nvar = 4;
    }
    BcFunction f;
    f.kind = fn->kind;
    f.gamma = fn->gamma;
    f.x0 = fn->x0;
    f.y0 = fn->y0;
    f.beta = fn->beta;
    f.u_inf = fn->u_inf;
    f.v_inf = fn->v_inf;
    m->bcfn = f;
    m->bc_time = t;
    return eval_bc(c, m, t) ? DGH_OK : c->last;
}

int dgh_mesh_set_patch_first(dgh_ctx c, int mh, const int* first) {
    if (!c) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    (void)hipSetDevice(c->device);
    dfree(m->patch_first);
    m->patch_first = nullptr;
    if (!first) return DGH_OK;
    const int K = m->host.K;
    std::vector<int> v(first, first + size_t(K) * 3);
    for (int i = 0; i < 3 * K; ++i) {
        const int s = v[i];
        if (s < -1 || s >= 3 * K || (s >= 0 && (m->host.EToE[i] >= 0 || m->host.EToE[s] >= 0)))
            return c->fail(DGH_ERR_ARG, "dgh_mesh_set_patch_first: entry %d (%d) is not a boundary face", i, s),
                   DGH_ERR_ARG;
    }
    return c->upload(&m->patch_first, v) ? DGH_OK : c->last;
}

int dgh_mesh_get_time(dgh_ctx c, int mh, double* t) {
    if (!c || !t) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    *t = m->bc_time;
    return DGH_OK;
}

int dgh_refelem_get(int N, const char* name, double* out, int cap) {
    if (!name || !out) return DGH_ERR_ARG;
    RefElem re;
    if (!build_refelem(N, re)) return DGH_ERR_UNSUPPORTED;
    const std::map<std::string, const std::vector<double>*> tab = {
        {"r", &re.r},   {"s", &re.s},     {"Dr", &re.Dr},   {"Ds", &re.Ds}, {"V", &re.V},
        {"M", &re.M},   {"Vq", &re.Vq},   {"Drq", &re.Drq}, {"Dsq", &re.Dsq}, {"cw", &re.cw},
        {"Vf", &re.Vf}, {"gw", &re.gw},   {"Pr", &re.Pr},   {"Ps", &re.Ps}, {"Lift", &re.Lift}};
    auto it = tab.find(name);
    if (it != tab.end()) {
        const int n = std::min<int>(cap, int(it->second->size()));
        std::memcpy(out, it->second->data(), sizeof(double) * n);
        return n;
    }
    if (std::string(name) == "fmask") {
        const int n = std::min<int>(cap, int(re.fmask.size()));
        for (int i = 0; i < n; ++i) out[i] = re.fmask[i];
        return n;
    }
    return DGH_ERR_ARG;
}

int dgh_field_create(dgh_ctx c, int mh, int nvar, int* out) {
    if (!c || !out || nvar < 1) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    (void)hipSetDevice(c->device);
    auto f = std::make_unique<FieldDev>();
    f->mesh = mh;
    f->nvar = nvar;
    f->K = m->host.K;
    f->Np = m->ref->re.Np;
    f->n = size_t(nvar) * round_up(m->host.K, 16) * m->ref->re.Np;   // device layout, K padded to 16
    if (!c->hip(hipMalloc(reinterpret_cast<void**>(&f->d), f->n * sizeof(double)), "hipMalloc field")) return c->last;
    if (!c->hip(hipMemsetAsync(f->d, 0, f->n * sizeof(double), c->stream), "hipMemsetAsync")) return c->last;
    c->fields.push_back(std::move(f));
    *out = int(c->fields.size()) - 1;
    return DGH_OK;
}

int dgh_field_destroy(dgh_ctx c, int fh) {
    if (!c) return DGH_ERR_ARG;
    if (!c->field(fh)) return c->last;
    (void)hipStreamSynchronize(c->stream);
    c->fields[fh].reset();
    return DGH_OK;
}

// host [var][K][Np] <-> device layout (dev_index in dg_common.hpp); padding repeats element K-1
int dgh_field_upload(dgh_ctx c, int fh, const double* host) {
    if (!c || !host) return DGH_ERR_ARG;
    FieldDev* f = c->field(fh);
    if (!f) return c->last;
    const int Kp = round_up(f->K, 16);
    std::vector<double> dev(f->n);
    for (int v = 0; v < f->nvar; ++v)
        for (int e = 0; e < Kp; ++e) {
            const double* src = host + (size_t(v) * f->K + std::min(e, f->K - 1)) * f->Np;
            for (int n = 0; n < f->Np; ++n) dev[dev_index(e, v, n, f->nvar, f->Np)] = src[n];
        }
    if (!c->hip(hipMemcpyAsync(f->d, dev.data(), f->n * sizeof(double), hipMemcpyHostToDevice, c->stream), "upload") ||
        !c->hip(hipStreamSynchronize(c->stream), "upload sync"))
        return c->last;
    return DGH_OK;
}

int dgh_field_download(dgh_ctx c, int fh, double* host) {
    if (!c || !host) return DGH_ERR_ARG;
    FieldDev* f = c->field(fh);
    if (!f) return c->last;
    std::vector<double> dev(f->n);
    if (!c->hip(hipMemcpyAsync(dev.data(), f->d, f->n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "download") ||
        !c->hip(hipStreamSynchronize(c->stream), "download sync"))
        return c->last;
    for (int v = 0; v < f->nvar; ++v)
        for (int e = 0; e < f->K; ++e) {
            double* dst = host + (size_t(v) * f->K + e) * f->Np;
            for (int n = 0; n < f->Np; ++n) dst[n] = dev[dev_index(e, v, n, f->nvar, f->Np)];
        }
    return DGH_OK;
}

int dgh_field_copy(dgh_ctx c, int dst, int src) {
    if (!c) return DGH_ERR_ARG;
    FieldDev* a = c->field(dst);
    FieldDev* b = c->field(src);
    if (!a || !b) return c->last;
    if (a->n != b->n) return c->fail(DGH_ERR_ARG, "dgh_field_copy: size mismatch"), DGH_ERR_ARG;
    return c->hip(hipMemcpyAsync(a->d, b->d, a->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream), "copy")
               ? DGH_OK
               : c->last;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// explicit stage
static constexpr const double* kRK4a = kLserkA;
static constexpr const double* kRK4b = kLserkB;

static int check_desc(dgh_ctx c, const dgh_rhs_desc* d, int* nvar) {
    if (!d) return c->fail(DGH_ERR_ARG, "null rhs descriptor"), DGH_ERR_ARG;
    if (d->eq == DGH_EQ_ADVECT) {
        *nvar = 1;
        if (d->flux != DGH_FLUX_UPWIND && d->flux != DGH_FLUX_LF && d->flux != DGH_FLUX_AVERAGE)
            return c->fail(DGH_ERR_UNSUPPORTED, "advection supports the upwind/LF and average fluxes"),
                   DGH_ERR_UNSUPPORTED;
    } else if (d->eq == DGH_EQ_EULER) {
        *nvar = 4;
        if (d->flux != DGH_FLUX_LF && d->flux != DGH_FLUX_ROE)
            return c->fail(DGH_ERR_UNSUPPORTED, "Euler supports the LF and Roe fluxes"), DGH_ERR_UNSUPPORTED;
    } else {
        return c->fail(DGH_ERR_UNSUPPORTED, "unknown equation %d", d->eq), DGH_ERR_UNSUPPORTED;
    }
    if (d->prec != DGH_PREC_FP64 && d->prec != DGH_PREC_FP32 && d->prec != DGH_PREC_BF16X3)
        return c->fail(DGH_ERR_UNSUPPORTED, "precision %d not available", d->prec), DGH_ERR_UNSUPPORTED;
    return DGH_OK;
}

// mom_out, p4f: the limiter's moments of the new state written by the stage kernel (bf16x3 stage updates only;
// dg_limit_fused_moments says when), else null
static int stage_launch(dgh_ctx c, MeshDev* m, const dgh_rhs_desc* d, const double* qin, double* qout, double* res,
                        int mode, double ra, double rb, double dt, int nvar, double* mom_out = nullptr,
                        const double* p4f = nullptr) {
    const RefDev* rd = m->ref;
    const bool f64 = d->prec == DGH_PREC_FP64;
    StageArgs a{};
    a.K = m->host.K;
    a.flux = d->flux;
    a.mode = mode;
    a.q_in = qin;
    a.q_out = qout;
    a.res = res;
    a.geo = f64 ? static_cast<const void*>(m->geo64) : static_cast<const void*>(m->geo32);
    a.conn = m->conn;
    const bool split = d->prec == DGH_PREC_BF16X3;
    a.VqF = f64 ? static_cast<const void*>(rd->VqF64) : split ? static_cast<const void*>(rd->VqS) : rd->VqF32;
    a.PF = f64 ? static_cast<const void*>(rd->PF64) : split ? static_cast<const void*>(rd->PS) : rd->PF32;
    a.LF = f64 ? static_cast<const void*>(rd->LF64) : split ? static_cast<const void*>(rd->LS) : rd->LF32;
    a.Vf = f64 ? static_cast<const void*>(rd->Vf64) : static_cast<const void*>(rd->Vf32);
    a.fmask = rd->fmask;
    a.bcq = (m->bcq && m->bc_nvar == nvar) ? m->bcq : nullptr;
    a.ax = d->ax;
    a.ay = d->ay;
    a.gamma = d->gamma;
    a.rk_a = ra;
    a.rk_b = rb;
    a.dt = dt;
    a.block_cap = c->stage_block_cap;
    if (mom_out) {
        if (!split || mode != 1 || !p4f)
            return c->fail(DGH_ERR_ARG, "stage kernel: moments output needs a bf16x3 stage update"), DGH_ERR_ARG;
        a.mom_out = mom_out;
        a.p4f = p4f;
    }
    const int phys = d->eq == DGH_EQ_EULER ? PHYS_EULER : PHYS_ADVECT;
    if (!c->hip(launch_stage(rd->re.N, phys, d->prec, a, c->ncu, c->stream), "stage kernel launch")) return c->last;
    return DGH_OK;
}

int dg_limit_enqueue(dgh_ctx c, MeshDev* m, const dgh_lim_desc* d, double* q, int nvar, bool reset_count,
                     const LimScratch* scratch = nullptr, bool moments_ready = false);   // limiter.cpp
// When the stage kernel can form the moments the limiter ld needs (its TVB, network or every-cell kinds after a
// bf16x3 stage): builds the mesh's limiter tables and moments buffer and returns them in mom, p4f; else mom = null.
// Returns a dgh_status.
int dg_limit_fused_moments(dgh_ctx c, MeshDev* m, const dgh_rhs_desc* d, const dgh_lim_desc* ld, int nvar,
                           double** mom, const double** p4f);   // limiter.cpp

extern "C" {

int dgh_rhs(dgh_ctx c, int mh, const dgh_rhs_desc* d, int q, int rhs) {
    if (!c) return DGH_ERR_ARG;
    int nvar = 0;
    if (int s = check_desc(c, d, &nvar)) return s;
    MeshDev* m = c->mesh(mh);
    FieldDev* fq = c->field(q);
    FieldDev* fr = c->field(rhs);
    if (!m || !fq || !fr) return c->last;
    if (fq->mesh != mh || fr->mesh != mh || fq->nvar != nvar || fr->nvar != nvar || q == rhs)
        return c->fail(DGH_ERR_ARG, "dgh_rhs: field/mesh mismatch"), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    return stage_launch(c, m, d, fq->d, fr->d, nullptr, 0, 0.0, 0.0, 0.0, nvar);
}

int dgh_lserk4_stage(dgh_ctx c, int mh, const dgh_rhs_desc* d, int q_in, int q_out, int res, int stage, double dt) {
    if (!c) return DGH_ERR_ARG;
    int nvar = 0;
    if (int s = check_desc(c, d, &nvar)) return s;
    MeshDev* m = c->mesh(mh);
    FieldDev* a = c->field(q_in);
    FieldDev* b = c->field(q_out);
    FieldDev* r = c->field(res);
    if (!m || !a || !b || !r) return c->last;
    if (stage < 0 || stage > 4 || q_in == q_out || res == q_in || res == q_out || a->nvar != nvar ||
        b->nvar != nvar || r->nvar != nvar || a->mesh != mh || b->mesh != mh || r->mesh != mh)
        return c->fail(DGH_ERR_ARG, "dgh_lserk4_stage: bad fields or stage (q_in, q_out and res must differ)"),
               DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    return stage_launch(c, m, d, a->d, b->d, r->d, 1, kRK4a[stage], kRK4b[stage], dt, nvar);
}

// SSP-RK2 of the reference's Euler driver (dgEulerFoam.C:64-123): q1 = q + dt L(q);
// q_new = (q + q1 + dt L(q1)) / 2 = q1 + (dt L(q1) - dt L(q)) / 2, i.e. the low-storage stage with
// (a, b) = (0, 1) then (-1, 1/2)
// stage 0: res = dt L(q^n), q^(1) = q^n + res.  Stage 1 (dgEulerFoam.C:99-118, rho = rho/2 + rho1/2 after
// rho1 is advanced): res = q^n - q^(1) from the fields themselves, then res += dt L(q^(1)) and
// q^(n+1) = q^(1) + res / 2 = (q^n + q^(1) + dt L(q^(1))) / 2 -- right also when a limiter has changed
// q^(1) between the stages.
static const double kSSPa[2] = {0.0, 1.0};
static const double kSSPb[2] = {1.0, 0.5};

int dgh_ssprk2_stage(dgh_ctx c, int mh, const dgh_rhs_desc* d, int q_in, int q_out, int res, int stage, double dt) {
    if (!c) return DGH_ERR_ARG;
    int nvar = 0;
    if (int s = check_desc(c, d, &nvar)) return s;
    MeshDev* m = c->mesh(mh);
    FieldDev* a = c->field(q_in);
    FieldDev* b = c->field(q_out);
    FieldDev* r = c->field(res);
    if (!m || !a || !b || !r) return c->last;
    if (stage < 0 || stage > 1 || q_in == q_out || res == q_in || res == q_out || a->nvar != nvar ||
        b->nvar != nvar || r->nvar != nvar || a->mesh != mh || b->mesh != mh || r->mesh != mh)
        return c->fail(DGH_ERR_ARG, "dgh_ssprk2_stage: bad fields or stage (q_in, q_out and res must differ)"),
               DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (stage == 1 && !c->hip(launch_field_diff(b->d, a->d, r->d, a->n, c->ncu, c->stream), "field difference"))
        return c->last;
    return stage_launch(c, m, d, a->d, b->d, r->d, 1, kSSPa[stage], kSSPb[stage], dt, nvar);
}

// nsteps LSERK4 steps on q, with the limiter ld (or none) applied to every stage's new state
static int lserk4_loop(dgh_ctx c, int mh, const dgh_rhs_desc* d, const dgh_lim_desc* ld, int q, int nsteps,
                       double dt, float* ms, const char* what) {
    if (!c) return DGH_ERR_ARG;
    int nvar = 0;
    if (int s = check_desc(c, d, &nvar)) return s;
    MeshDev* m = c->mesh(mh);
    FieldDev* fq = c->field(q);
    if (!m || !fq) return c->last;
    if (fq->mesh != mh || fq->nvar != nvar || nsteps < 0)
        return c->fail(DGH_ERR_ARG, "%s: field/mesh mismatch", what), DGH_ERR_ARG;
    if (ld && ld->detect_only) return c->fail(DGH_ERR_ARG, "%s: a detect-only limiter changes nothing", what), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    Workspace& w = m->ws[nvar];
    if (!w.res) {
        if (!c->hip(hipMalloc(reinterpret_cast<void**>(&w.res), fq->n * sizeof(double)), "hipMalloc ws") ||
            !c->hip(hipMalloc(reinterpret_cast<void**>(&w.tmp), fq->n * sizeof(double)), "hipMalloc ws"))
            return c->last;
    }
    if (!c->hip(hipMemsetAsync(w.res, 0, fq->n * sizeof(double), c->stream), "memset res")) return c->last;
    // bf16x3 stages with a TVB-family limiter: the stage kernel forms the limiter's moments of its new state, and
    // the limiter skips its moments pass (the same bits as the separate pass)
    double* mom = nullptr;
    const double* p4f = nullptr;
    if (ld) {
        if (int st = dg_limit_fused_moments(c, m, d, ld, nvar, &mom, &p4f)) return st;
    }
    double* cur = fq->d;
    double* nxt = w.tmp;
    const double t0 = m->bc_time;
    if (!c->hip(hipEventRecord(c->ev0, c->stream), "event")) return c->last;
    for (int step = 0; step < nsteps; ++step)
        for (int s = 0; s < 5; ++s) {
            if (!eval_bc(c, m, t0 + (step + kLserkC[s]) * dt)) return c->last;
            if (int st = stage_launch(c, m, d, cur, nxt, w.res, 1, kRK4a[s], kRK4b[s], dt, nvar, mom, p4f)) return st;
            if (ld) {
                if (int st = dg_limit_enqueue(c, m, ld, nxt, nvar, false, nullptr, mom != nullptr)) return st;
            }
            std::swap(cur, nxt);
        }
    if (!c->hip(hipEventRecord(c->ev1, c->stream), "event")) return c->last;
    m->bc_time = t0 + nsteps * dt;
    if (!eval_bc(c, m, m->bc_time)) return c->last;
    if (cur != fq->d &&
        !c->hip(hipMemcpyAsync(fq->d, cur, fq->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream), "copy back"))
        return c->last;
    if (!c->hip(hipStreamSynchronize(c->stream), "sync")) return c->last;
    if (ms && !c->hip(hipEventElapsedTime(ms, c->ev0, c->ev1), "elapsed")) return c->last;
    return DGH_OK;
}

int dgh_ssprk2_run(dgh_ctx c, int mh, const dgh_rhs_desc* d, const dgh_lim_desc* ld, int q, int nsteps, double dt,
                   float* ms) {
    if (!c) return DGH_ERR_ARG;
    int nvar = 0;
    if (int s = check_desc(c, d, &nvar)) return s;
    MeshDev* m = c->mesh(mh);
    FieldDev* fq = c->field(q);
    if (!m || !fq) return c->last;
    if (fq->mesh != mh || fq->nvar != nvar || nsteps < 0)
        return c->fail(DGH_ERR_ARG, "dgh_ssprk2_run: field/mesh mismatch"), DGH_ERR_ARG;
    if (ld && ld->detect_only)
        return c->fail(DGH_ERR_ARG, "dgh_ssprk2_run: a detect-only limiter changes nothing"), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    Workspace& w = m->ws[nvar];
    if (!w.res) {
        if (!c->hip(hipMalloc(reinterpret_cast<void**>(&w.res), fq->n * sizeof(double)), "hipMalloc ws") ||
            !c->hip(hipMalloc(reinterpret_cast<void**>(&w.tmp), fq->n * sizeof(double)), "hipMalloc ws"))
            return c->last;
    }
    // q holds q^n; w.tmp the first stage q^(1); the second stage writes q^(n+1) back into q (dgh_ssprk2_stage).
    // The first stage scales res by 0: it must hold numbers, not whatever the allocation held.
    if (!c->hip(hipMemsetAsync(w.res, 0, fq->n * sizeof(double), c->stream), "memset res")) return c->last;
    double* mom = nullptr;   // moments formed by bf16x3 stage kernels for a TVB-family limiter (lserk4_loop)
    const double* p4f = nullptr;
    if (ld) {
        if (int st = dg_limit_fused_moments(c, m, d, ld, nvar, &mom, &p4f)) return st;
    }
    const double t0 = m->bc_time;
    if (!c->hip(hipEventRecord(c->ev0, c->stream), "event")) return c->last;
    for (int step = 0; step < nsteps; ++step) {
        // both stages see the boundary data of t^n (dgEulerFoam.C:73)
        if (!eval_bc(c, m, t0 + step * dt)) return c->last;
        if (int st = stage_launch(c, m, d, fq->d, w.tmp, w.res, 1, kSSPa[0], kSSPb[0], dt, nvar, mom, p4f)) return st;
        if (ld) {
            if (int st = dg_limit_enqueue(c, m, ld, w.tmp, nvar, false, nullptr, mom != nullptr)) return st;
        }
        if (!c->hip(launch_field_diff(fq->d, w.tmp, w.res, fq->n, c->ncu, c->stream), "field difference"))
            return c->last;
        if (int st = stage_launch(c, m, d, w.tmp, fq->d, w.res, 1, kSSPa[1], kSSPb[1], dt, nvar, mom, p4f)) return st;
        if (ld) {
            if (int st = dg_limit_enqueue(c, m, ld, fq->d, nvar, false, nullptr, mom != nullptr)) return st;
        }
    }
    if (!c->hip(hipEventRecord(c->ev1, c->stream), "event")) return c->last;
    m->bc_time = t0 + nsteps * dt;
    if (!eval_bc(c, m, m->bc_time)) return c->last;
    if (!c->hip(hipStreamSynchronize(c->stream), "sync")) return c->last;
    if (ms && !c->hip(hipEventElapsedTime(ms, c->ev0, c->ev1), "elapsed")) return c->last;
    return DGH_OK;
}

int dgh_lserk4_run(dgh_ctx c, int mh, const dgh_rhs_desc* d, int q, int nsteps, double dt, float* ms) {
    return lserk4_loop(c, mh, d, nullptr, q, nsteps, dt, ms, "dgh_lserk4_run");
}

int dgh_lserk4_run_limited(dgh_ctx c, int mh, const dgh_rhs_desc* d, const dgh_lim_desc* ld, int q, int nsteps,
                           double dt, float* ms) {
    if (c && !ld) return c->fail(DGH_ERR_ARG, "dgh_lserk4_run_limited: null limiter descriptor"), DGH_ERR_ARG;
    return lserk4_loop(c, mh, d, ld, q, nsteps, dt, ms, "dgh_lserk4_run_limited");
}

int dgh_selftest_mfma(dgh_ctx c, double* err64, double* err32) {
    if (!c || !err64 || !err32) return DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    std::vector<double> A(64), B(64), D(256, 0.0);
    std::vector<float> Af(64), Bf(64);
    for (int i = 0; i < 64; ++i) {
        A[i] = 1.0 + i * 0.25;
        B[i] = 2.0 - i * 0.125;
        Af[i] = float(A[i]);
        Bf[i] = float(B[i]);
    }
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j)
            for (int k = 0; k < 4; ++k) D[i * 16 + j] += A[i * 4 + k] * B[k * 16 + j];
    double *dA, *dB, *dD;
    float *fA, *fB, *fD;
    bool ok = c->upload(&dA, A) && c->upload(&dB, B) && c->upload(&fA, Af) && c->upload(&fB, Bf) &&
              c->hip(hipMalloc(reinterpret_cast<void**>(&dD), 256 * sizeof(double)), "malloc") &&
              c->hip(hipMalloc(reinterpret_cast<void**>(&fD), 256 * sizeof(float)), "malloc");
    if (!ok) return c->last;
    std::vector<double> o64(256);
    std::vector<float> o32(256);
    ok = c->hip(launch_mfma_probe(dA, dB, dD, fA, fB, fD, c->stream), "probe") &&
         c->hip(hipMemcpyAsync(o64.data(), dD, 256 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "d2h") &&
         c->hip(hipMemcpyAsync(o32.data(), fD, 256 * sizeof(float), hipMemcpyDeviceToHost, c->stream), "d2h") &&
         c->hip(hipStreamSynchronize(c->stream), "sync");
    dfree(dA); dfree(dB); dfree(dD); dfree(fA); dfree(fB); dfree(fD);
    if (!ok) return c->last;
    *err64 = 0.0;
    *err32 = 0.0;
    for (int i = 0; i < 256; ++i) {
        *err64 = std::max(*err64, std::fabs(o64[i] - D[i]));
        *err32 = std::max(*err32, std::fabs(double(o32[i]) - D[i]));
    }
    return DGH_OK;
}

int dgh_selftest_mfma_bf16x3(dgh_ctx c, double* err3, double* err1) {
    if (!c || !err3 || !err1) return DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    // A, B: 16 x 16 with entries of mixed sign and magnitude that bf16 does not hold exactly
    std::vector<double> A(256), Bd(256), D(256, 0.0);
    std::vector<float> B(256);
    for (int i = 0; i < 256; ++i) {
        A[i] = std::sin(0.37 * i + 0.1) * (1.0 + (i % 7) / 3.0);
        B[i] = float(std::cos(0.23 * i - 0.4) * (0.5 + (i % 5) / 4.0));
        Bd[i] = double(B[i]);
    }
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j)
            for (int k = 0; k < 16; ++k) D[i * 16 + j] += A[i * 16 + k] * Bd[k * 16 + j];
    std::vector<uint32_t> Af(64 * 4);
    for (int l = 0; l < 64; ++l) {
        double x[4];
        for (int e = 0; e < 4; ++e) x[e] = A[(l & 15) * 16 + 4 * (l >> 4) + e];
        put_split(Af, size_t(l), x);
    }
    uint32_t* dA = nullptr;
    float *dB = nullptr, *d3 = nullptr, *d1 = nullptr;
    bool ok = c->upload(&dA, Af) && c->upload(&dB, B) &&
              c->hip(hipMalloc(reinterpret_cast<void**>(&d3), 256 * sizeof(float)), "malloc") &&
              c->hip(hipMalloc(reinterpret_cast<void**>(&d1), 256 * sizeof(float)), "malloc");
    std::vector<float> o3(256), o1(256);
    ok = ok && c->hip(launch_mfma_bf16x3_probe(dA, dB, d3, d1, c->stream), "bf16x3 probe") &&
         c->hip(hipMemcpyAsync(o3.data(), d3, 256 * sizeof(float), hipMemcpyDeviceToHost, c->stream), "d2h") &&
         c->hip(hipMemcpyAsync(o1.data(), d1, 256 * sizeof(float), hipMemcpyDeviceToHost, c->stream), "d2h") &&
         c->hip(hipStreamSynchronize(c->stream), "sync");
    dfree(dA); dfree(dB); dfree(d3); dfree(d1);
    if (!ok) return c->last;
    double scale = 0.0, e3 = 0.0, e1 = 0.0;
    for (int i = 0; i < 256; ++i) {
        scale = std::max(scale, std::fabs(D[i]));
        e3 = std::max(e3, std::fabs(double(o3[i]) - D[i]));
        e1 = std::max(e1, std::fabs(double(o1[i]) - D[i]));
    }
    *err3 = e3 / scale;
    *err1 = e1 / scale;
    return DGH_OK;
}

int dgh_selftest_mfma_layout(dgh_ctx c, double* raw) {
    if (!c || !raw) return DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    double* d = nullptr;
    if (!c->hip(hipMalloc(reinterpret_cast<void**>(&d), 1024 * sizeof(double)), "malloc")) return c->last;
    const bool ok = c->hip(launch_mfma_layout_probe(d, c->stream), "layout probe") &&
                    c->hip(hipMemcpyAsync(raw, d, 1024 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "d2h") &&
                    c->hip(hipStreamSynchronize(c->stream), "sync");
    dfree(d);
    return ok ? DGH_OK : c->last;
}

}  // extern "C"

// Internal state behind the dgh_ctx handle.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "boundary.hpp"
#include "dg_common.hpp"
#include "mesh.hpp"
#include "refelem.hpp"

namespace dg {

// reference-element operators resident on the device (one per order N)
struct RefDev {
    RefElem re;
    double *VqF64 = nullptr, *PF64 = nullptr, *LF64 = nullptr, *Vf64 = nullptr;
    float *VqF32 = nullptr, *PF32 = nullptr, *LF32 = nullptr, *Vf32 = nullptr;
    uint32_t *VqS = nullptr, *PS = nullptr, *LS = nullptr;   // bf16x3 split fragments
    int* fmask = nullptr;
    // projection-solver operators (built on first use): Pr Vq, Ps Vq, Lift, Dr, Ds in FP64
    double *insSr = nullptr, *insSs = nullptr, *insLift = nullptr, *insDr = nullptr, *insDs = nullptr;
    ~RefDev();
};

struct LimDev;   // TVB limiter state of a mesh (limiter.cpp)
struct EnsDev;   // LDG advection-diffusion ensemble (ensemble.cpp)

struct Workspace {
    double* res = nullptr;
    double* tmp = nullptr;
};

struct MeshDev {
    HostMesh host;
    RefDev* ref = nullptr;
    double* geo64 = nullptr;   // [kGeoSlots][K]
    float* geo32 = nullptr;
    int* conn = nullptr;       // [3][K]
    double* bcq = nullptr;     // fixedValue ghost nodal values
    int bc_nvar = 0;
    // time-dependent fixedValue data: the function, the mesh's time, and the fixedValue face nodes (coordinates
    // and slots (e 3 + f) Nfp + k in bcq)
    BcFunction bcfn;
    double bc_time = 0.0;
    double* bc_xy = nullptr;
    int* bc_slot = nullptr;
    int bc_n = 0;
    int* patch_first = nullptr;   // [K][3] first face of each fixedValue face's patch (Triangle ghost means), or null
    std::map<int, Workspace> ws;   // LSERK4 scratch per variable count
    std::shared_ptr<LimDev> lim;
    ~MeshDev();
};

struct FieldDev {
    int mesh = -1, nvar = 0, K = 0, Np = 0;
    size_t n = 0;   // doubles in the device layout (K padded to 16)
    double* d = nullptr;
    ~FieldDev();
};

struct SipgDev;   // implicit solver state (sipg.cpp)
struct MlpDev;    // troubled-cell network (mlp.cpp)

}  // namespace dg

struct dgh_ctx_s {
    int device = 0;
    int ncu = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int stage_block_cap = 0;   // > 0: at most this many persistent blocks per stage launch (dgh_set_launch_cap)
    void* noalloc = nullptr;   // device address of this device's static scratch of dgh_limit_noalloc
    std::string err;
    int last = 0;
    std::map<int, std::unique_ptr<dg::RefDev>> refs;
    std::vector<std::unique_ptr<dg::MeshDev>> meshes;
    std::vector<std::unique_ptr<dg::FieldDev>> fields;
    std::vector<std::shared_ptr<dg::SipgDev>> sipg;
    std::vector<std::shared_ptr<dg::MlpDev>> mlps;
    std::vector<std::shared_ptr<dg::EnsDev>> ens;

    bool fail(int code, const char* fmt, ...);
    bool hip(hipError_t e, const char* what);
    template <typename T>
    bool upload(T** dptr, const std::vector<T>& v);
    dg::RefDev* refdev(int N);
    dg::MeshDev* mesh(int h);
    dg::FieldDev* field(int h);
};

// Host mesh construction (see mesh.hpp).
//
// Geometry follows physicalCellElement.C:72-110 and triangleBaseFunction.C:303-419 for affine
// triangles: x = -(r+s)/2 v0 + (r+1)/2 v1 + (s+1)/2 v2, so xr = (v1-v0)/2, xs = (v2-v0)/2,
// J = xr ys - xs yr, rx = ys/J, sx = -yr/J, ry = -xs/J, sy = xr/J; the outward normals of faces
// 0, 1, 2 are (yr,-xr), (ys-yr, xr-xs), (-ys, xs) normalised, Js = their length (half the edge).
#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <unordered_map>

namespace dg {

bool build_mesh(int N, int K, const double* vxy, const int* vid, const int* bctag, HostMesh& m) {
    m.N = N;
    m.K = K;
    m.vxy.assign(vxy, vxy + size_t(K) * 6);
    m.vid.assign(vid, vid + size_t(K) * 3);
    if (bctag) m.bctag.assign(bctag, bctag + size_t(K) * 3);
    else m.bctag.assign(size_t(K) * 3, 0);
    m.EToE.assign(size_t(K) * 3, -1);
    m.EToF.assign(size_t(K) * 3, -1);
    std::unordered_map<unsigned long long, int> open;
    open.reserve(size_t(K) * 2);
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 3; ++f) {
            const unsigned a = unsigned(vid[e * 3 + f]), b = unsigned(vid[e * 3 + (f + 1) % 3]);
            const unsigned long long key = (static_cast<unsigned long long>(std::min(a, b)) << 32) | std::max(a, b);
            auto it = open.find(key);
            if (it == open.end()) {
                open.emplace(key, e * 3 + f);
            } else {
                const int o = it->second;
                if (o < 0) return false;   // third face on one edge
                m.EToE[e * 3 + f] = o / 3;
                m.EToF[e * 3 + f] = o % 3;
                m.EToE[o] = e;
                m.EToF[o] = f;
                it->second = -1;
            }
        }
    m.nbfaces = 0;
    for (int i = 0; i < 3 * K; ++i) m.nbfaces += m.EToE[i] < 0;
    m.geo.assign(size_t(K) * 14, 0.0);
    for (int e = 0; e < K; ++e) {
        const double* v = vxy + size_t(e) * 6;
        const double xr = 0.5 * (v[2] - v[0]), yr = 0.5 * (v[3] - v[1]);
        const double xs = 0.5 * (v[4] - v[0]), ys = 0.5 * (v[5] - v[1]);
        const double J = xr * ys - xs * yr;
        if (!(J > 0)) return false;
        double* g = &m.geo[size_t(e) * 14];
        g[0] = ys / J;
        g[1] = -xs / J;
        g[2] = -yr / J;
        g[3] = xr / J;
        g[4] = J;
        const double fn[3][2] = {{yr, -xr}, {ys - yr, xr - xs}, {-ys, xs}};
        for (int f = 0; f < 3; ++f) {
            const double Js = std::sqrt(fn[f][0] * fn[f][0] + fn[f][1] * fn[f][1]);
            g[5 + 3 * f] = fn[f][0] / Js;
            g[6 + 3 * f] = fn[f][1] / Js;
            g[7 + 3 * f] = Js;
        }
    }
    return true;
}

void structured_split_quads(int nx, int ny, double x0, double x1, double y0, double y1, bool px, bool py,
                            int bc_tag, std::vector<double>& vxy, std::vector<int>& vid, std::vector<int>& bctag) {
    const double hx = (x1 - x0) / nx, hy = (y1 - y0) / ny;
    const int nvx = px ? nx : nx + 1;
    const int nvy = py ? ny : ny + 1;
    auto id = [&](int i, int j) { return (py ? j % nvy : j) * nvx + (px ? i % nvx : i); };
    const int K = 2 * nx * ny;
    vxy.assign(size_t(K) * 6, 0.0);
    vid.assign(size_t(K) * 3, 0);
    bctag.assign(size_t(K) * 3, bc_tag);
    static const int tri[2][3] = {{0, 1, 2}, {0, 2, 3}};
    int e = 0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int ci[4] = {i, i + 1, i + 1, i}, cj[4] = {j, j, j + 1, j + 1};
            for (int t = 0; t < 2; ++t, ++e)
                for (int k = 0; k < 3; ++k) {
                    const int c = tri[t][k];
                    vxy[size_t(e) * 6 + 2 * k] = x0 + ci[c] * hx;
                    vxy[size_t(e) * 6 + 2 * k + 1] = y0 + cj[c] * hy;
                    vid[size_t(e) * 3 + k] = id(ci[c], cj[c]);
                }
        }
}

}  // namespace dg

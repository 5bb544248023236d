// Host mesh: connectivity and affine geometric factors of 2D triangle meshes.
#pragma once
#include <vector>

namespace dg {

struct HostMesh {
    int N = 0, K = 0, nbfaces = 0;
    std::vector<double> vxy;       // K x 3 x 2
    std::vector<int> vid, bctag;   // K x 3
    std::vector<int> EToE, EToF;   // K x 3 (-1 on boundary)
    std::vector<double> geo;       // K x 14: rx ry sx sy J (nx ny Js) x 3
};

// Face k of element e joins its vertices k and k+1 (dgPolyMesh.C:412-530); two faces match
// when their vertex-id pairs match.  Returns false on a non-manifold edge or degenerate element.
bool build_mesh(int N, int K, const double* vxy, const int* vid, const int* bctag, HostMesh& m);

// nx x ny quads on [x0,x1]x[y0,y1], each split along its (i,j)-(i+1,j+1) diagonal into the
// anticlockwise triangles (v00,v10,v11) and (v00,v11,v01); element 2(j nx + i) is the first.
void structured_split_quads(int nx, int ny, double x0, double x1, double y0, double y1, bool px, bool py,
                            int bc_tag, std::vector<double>& vxy, std::vector<int>& vid, std::vector<int>& bctag);

}  // namespace dg

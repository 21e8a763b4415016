// ne_oracle_bvh.cpp -- TEST INFRASTRUCTURE ONLY (part of the oracle).
//
// The bounding-volume tree of a static triangle mesh, as `MeshData::build`
// makes it (src/mesh.hpp:38-43 and buildNode 62-92): nodes are numbered in
// creation order (a node before its children, the left subtree before the
// right); a node bounds its triangles' boxes; up to 4 triangles make a leaf;
// otherwise the triangles are split at the middle index along the widest axis
// of the node box (x if it is strictly the widest, else y if wider than z,
// else z), by `std::nth_element` on the sum of each triangle box's low and
// high coordinate on that axis.  nth_element leaves the order within each
// half to the standard library (libstdc++'s introselect here, the library
// the reference is built against), and the leaves' contents and the order
// the query reports triangles in depend on it, so this restatement calls the
// same library routine with the same comparison rather than re-deriving it.
#include <algorithm>
#include <cstdint>
#include <vector>

namespace {

struct Tree {
    const double* verts;
    const int32_t* tris;
    std::vector<int32_t> order;
    std::vector<double> bounds;   // 6 per node: lo xyz, hi xyz
    std::vector<int32_t> links;   // 4 per node: left, right, start, count

    // triangle box: componentwise min / max of its three vertices (mesh.hpp:30-35)
    void tri_box(int32_t t, double lo[3], double hi[3]) const {
        const double* a = verts + 3 * (size_t)tris[3 * (size_t)t];
        const double* b = verts + 3 * (size_t)tris[3 * (size_t)t + 1];
        const double* c = verts + 3 * (size_t)tris[3 * (size_t)t + 2];
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min({a[k], b[k], c[k]});
            hi[k] = std::max({a[k], b[k], c[k]});
        }
    }

    int32_t node(int32_t start, int32_t count) {
        const int32_t idx = (int32_t)(links.size() / 4);
        bounds.resize(bounds.size() + 6);
        links.resize(links.size() + 4);
        double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
        for (int32_t i = 0; i < count; ++i) {
            double tl[3], th[3];
            tri_box(order[(size_t)start + i], tl, th);
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], tl[k]);
                hi[k] = std::max(hi[k], th[k]);
            }
        }
        int32_t left = -1, right = -1, s = 0, n = 0;
        if (count <= 4) {
            s = start;
            n = count;
        } else {
            const double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
            const int axis = (ex > ey && ex > ez) ? 0 : (ey > ez ? 1 : 2);
            const int32_t mid = start + count / 2;
            std::nth_element(order.begin() + start, order.begin() + mid, order.begin() + start + count,
                             [&](int32_t A, int32_t B) {
                                 double al[3], ah[3], bl[3], bh[3];
                                 tri_box(A, al, ah);
                                 tri_box(B, bl, bh);
                                 return al[axis] + ah[axis] < bl[axis] + bh[axis];
                             });
            left = node(start, mid - start);
            right = node(mid, start + count - mid);
        }
        for (int k = 0; k < 3; ++k) {
            bounds[6 * (size_t)idx + k] = lo[k];
            bounds[6 * (size_t)idx + 3 + k] = hi[k];
        }
        links[4 * (size_t)idx] = left;
        links[4 * (size_t)idx + 1] = right;
        links[4 * (size_t)idx + 2] = s;
        links[4 * (size_t)idx + 3] = n;
        return idx;
    }
};

}  // namespace

extern "C" {

// Builds the tree of `ntris` triangles.  Writes `order` (ntris entries) and,
// when node_cap is large enough, the nodes; returns the node count (also when
// node_cap is too small: then nothing but `order` is written), or -1.
int32_t ne_oracle_bvh_build(const double* verts, const int32_t* tris, int32_t ntris, int32_t* order,
                            int32_t node_cap, double* bounds, int32_t* links) {
    if (ntris < 0 || (ntris > 0 && (!verts || !tris || !order))) return -1;
    Tree t;
    t.verts = verts;
    t.tris = tris;
    t.order.resize((size_t)ntris);
    for (int32_t i = 0; i < ntris; ++i) t.order[(size_t)i] = i;
    if (ntris > 0) t.node(0, ntris);
    for (int32_t i = 0; i < ntris; ++i) order[i] = t.order[(size_t)i];
    const int32_t nn = (int32_t)(t.links.size() / 4);
    if (nn <= node_cap && bounds && links) {
        std::copy(t.bounds.begin(), t.bounds.end(), bounds);
        std::copy(t.links.begin(), t.links.end(), links);
    }
    return nn;
}

}  // extern "C"

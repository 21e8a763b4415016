// The drop-in header's factories (include/ne_world.hpp) on sizes read from
// stdin, one body per line: "shape density d0 d1 d2 nverts [x y z]..." with
// hex floats; prints "invMass ii0 ii1 ii2 he0 he1 he2 radius halfHeight" in
// hex, for the factory bit checks in tests/test_factories.py.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ne_world.hpp"

int main() {
    int shape = 0;
    char tok[64];
    while (std::scanf("%d", &shape) == 1) {
        double v[4];
        for (double& d : v) {
            if (std::scanf("%63s", tok) != 1) return 2;
            d = std::strtod(tok, nullptr);
        }
        unsigned nv = 0;
        if (std::scanf("%u", &nv) != 1) return 2;
        std::vector<ne::V3> verts(nv);
        for (ne::V3& p : verts) {
            double c[3];
            for (double& d : c) {
                if (std::scanf("%63s", tok) != 1) return 2;
                d = std::strtod(tok, nullptr);
            }
            p = ne::V3{c[0], c[1], c[2]};
        }
        const double density = v[0];
        const ne::V3 x{0.0, 0.0, 0.0};
        const ne::Q q{1.0, 0.0, 0.0, 0.0};
        ne::Body b;
        switch (shape) {
            case NE_SHAPE_SPHERE: b = ne::makeSphere(0, x, v[1], density); break;
            case NE_SHAPE_CYLINDER: b = ne::makeCylinder(0, x, q, v[1], v[2], density); break;
            case NE_SHAPE_BOX: b = ne::makeBox(0, x, q, ne::V3{v[1], v[2], v[3]}, density); break;
            case NE_SHAPE_CAPSULE: b = ne::makeCapsule(0, x, q, v[1], v[2], density); break;
            case NE_SHAPE_CONVEX: b = ne::makeConvex(0, x, q, verts, density); break;
            case NE_SHAPE_PLANE: b = ne::makePlane(0, x, ne::V3{v[1], v[2], v[3]}); break;
            default: return 3;
        }
        std::printf("%a %a %a %a %a %a %a %a %a\n", b.invMass, b.invInertiaBody.x, b.invInertiaBody.y,
                    b.invInertiaBody.z, b.halfExtents.x, b.halfExtents.y, b.halfExtents.z, b.radius, b.halfHeight);
    }
    return 0;
}

// ne_dmath.hpp -- FP64 vector / quaternion math for the gfx950 kernels.
//
// Every operation keeps the reference's evaluation order (src/vmath.hpp:18-132,
// src/box.hpp:24-41, src/body.hpp:162-202) so that, compiled with
// -ffp-contract=off, a kernel produces the same IEEE double bit patterns as the
// single-threaded reference: no fused multiply-add, no reassociation.  Only
// + - * / sqrt floor fabs and comparisons appear.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NE_HD __host__ __device__ __forceinline__

namespace nea {

struct V3 {
    double x, y, z;
};
struct Qt {
    double w, x, y, z;
};

NE_HD V3 v3(double x, double y, double z) { return V3{x, y, z}; }
NE_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
NE_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
NE_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
NE_HD V3 operator*(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
NE_HD V3 operator/(V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; }
NE_HD double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
NE_HD V3 cross(V3 a, V3 o) {
    return V3{a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x};
}
NE_HD double norm2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
NE_HD double norm(V3 a) { return sqrt(norm2(a)); }
NE_HD V3 normalized(V3 a) {
    double n = norm(a);
    return (n > 1e-300) ? V3{a.x / n, a.y / n, a.z / n} : V3{0.0, 0.0, 0.0};
}
NE_HD double comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

NE_HD Qt qmul(Qt a, Qt o) {
    return Qt{a.w * o.w - a.x * o.x - a.y * o.y - a.z * o.z,
              a.w * o.x + a.x * o.w + a.y * o.z - a.z * o.y,
              a.w * o.y - a.x * o.z + a.y * o.w + a.z * o.x,
              a.w * o.z + a.x * o.y - a.y * o.x + a.z * o.w};
}
NE_HD Qt qnormalized(Qt a) {
    double n = sqrt(a.w * a.w + a.x * a.x + a.y * a.y + a.z * a.z);
    return (n > 1e-300) ? Qt{a.w / n, a.x / n, a.y / n, a.z / n} : Qt{1.0, 0.0, 0.0, 0.0};
}
NE_HD Qt qconj(Qt a) { return Qt{a.w, -a.x, -a.y, -a.z}; }
// v' = v + w t + q_vec x t with t = 2 (q_vec x v)
NE_HD V3 qrot(Qt q, V3 v) {
    V3 qv{q.x, q.y, q.z};
    V3 t = cross(qv, v) * 2.0;
    return (v + t * q.w) + cross(qv, t);
}
NE_HD V3 qinvrot(Qt q, V3 v) { return qrot(qconj(q), v); }
// q + (omega_quat * q) * (dt/2), renormalised
NE_HD Qt integrate_orientation(Qt q, V3 om, double dt) {
    Qt oq{0.0, om.x, om.y, om.z};
    Qt m = qmul(oq, q);
    double h = 0.5 * dt;
    Qt dq{m.w * h, m.x * h, m.y * h, m.z * h};
    return qnormalized(Qt{q.w + dq.w, q.x + dq.x, q.y + dq.y, q.z + dq.z});
}

// Domain box with per-axis half extents (a cube when all three are equal).
struct Box3 {
    double half[3];
    double edge[3];       // 2*half
    double inv_edge[3];   // 1/edge: conservative candidate tests only, never the exact path
    int periodic;
};

NE_HD double mi1(double d, double L) { return d - L * floor(d / L + 0.5); }
NE_HD V3 min_image(const Box3& b, V3 d) {
    if (!b.periodic) return d;
    return V3{mi1(d.x, b.edge[0]), mi1(d.y, b.edge[1]), mi1(d.z, b.edge[2])};
}
NE_HD double wr1(double p, double h, double L) { return p - L * floor((p + h) / L); }
NE_HD V3 wrap_pos(const Box3& b, V3 p) {
    if (!b.periodic) return p;
    return V3{wr1(p.x, b.half[0], b.edge[0]), wr1(p.y, b.half[1], b.edge[1]),
              wr1(p.z, b.half[2], b.edge[2])};
}

// R diag(invI) R^T u
NE_HD V3 apply_inv_inertia(Qt q, V3 invI, V3 u) {
    V3 ub = qinvrot(q, u);
    V3 sb{ub.x * invI.x, ub.y * invI.y, ub.z * invI.z};
    return qrot(q, sb);
}
NE_HD V3 inertia_of(V3 invI) {
    return V3{invI.x > 0 ? 1.0 / invI.x : 0.0, invI.y > 0 ? 1.0 / invI.y : 0.0,
              invI.z > 0 ? 1.0 / invI.z : 0.0};
}
NE_HD V3 angular_momentum(Qt q, V3 invI, V3 w) {
    V3 Ib = inertia_of(invI);
    V3 wb = qinvrot(q, w);
    V3 Lb{wb.x * Ib.x, wb.y * Ib.y, wb.z * Ib.z};
    return qrot(q, Lb);
}

}  // namespace nea

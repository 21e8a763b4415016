// ne_joint.hpp -- general joints on the device: World::solveJoints
// (src/world.cpp:206-315), ball / hinge (motor, angle limit) / fixed /
// slider by sequential impulse, 8 iterations over the joints in order.
//
// Joints that share no movable body never interact, so each connected
// component (union-find over the movable bodies, built on the host) is swept
// by one thread in the joints' own order -- the reference's order restricted
// to the component -- and components run in parallel.  The arithmetic is the
// reference's, including its 3x3 matrices (vmath.hpp:81-123): products
// accumulate from zero, inverses by cofactors.  The hinge limit's angle takes
// atan2 rounded to nearest (ne_atan2.hpp) -- the host C library's result
// wherever that is correctly rounded (world.cpp:300).
//
// Included by ne_gpu.hip after its body layout and sleep helpers (BodiesDev,
// BodyView, load_view, store_view, sleep_union).
#pragma once

#include "ne_atan2.hpp"
#include "ne_dmath.hpp"

namespace nea {

struct M3 {
    double m[9];   // m[row * 3 + col]
};

__device__ __forceinline__ M3 m3_zero() {
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = 0.0;
    return r;
}
__device__ __forceinline__ M3 m3_make(double a, double b, double c, double d, double e, double f, double g, double h,
                                      double i) {
    M3 r;
    r.m[0] = a; r.m[1] = b; r.m[2] = c; r.m[3] = d; r.m[4] = e; r.m[5] = f; r.m[6] = g; r.m[7] = h; r.m[8] = i;
    return r;
}
__device__ __forceinline__ M3 m3_skew(V3 v) { return m3_make(0, -v.z, v.y, v.z, 0, -v.x, -v.y, v.x, 0); }
__device__ __forceinline__ V3 m3_mulv(const M3& a, V3 v) {
    return V3{a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
              a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
__device__ __forceinline__ M3 m3_mul(const M3& a, const M3& o) {
    M3 r = m3_zero();
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) r.m[i * 3 + j] += a.m[i * 3 + k] * o.m[k * 3 + j];
    return r;
}
__device__ __forceinline__ M3 m3_add(const M3& a, const M3& o) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] + o.m[i];
    return r;
}
__device__ __forceinline__ M3 m3_sub(const M3& a, const M3& o) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] - o.m[i];
    return r;
}
__device__ __forceinline__ V3 m3_solve(const M3& M, V3 rhs) {
    const double a = M.m[0], b = M.m[1], c = M.m[2], d = M.m[3], e = M.m[4], f = M.m[5], g = M.m[6], h = M.m[7],
                 i = M.m[8];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    if (fabs(det) < 1e-18) return m3_mulv(m3_zero(), rhs);
    const double inv = 1.0 / det;
    const M3 r = m3_make(A * inv, (c * h - b * i) * inv, (b * f - c * e) * inv, B * inv, (a * i - c * g) * inv,
                         (c * d - a * f) * inv, C * inv, (b * g - a * h) * inv, (a * e - b * d) * inv);
    return m3_mulv(r, rhs);
}

// R diag(invI) R^T with R's columns the rotated basis (body.hpp:184-192).
__device__ __forceinline__ M3 inv_inertia_world(Qt q, V3 ii) {
    const V3 cx = qrot(q, V3{1, 0, 0}), cy = qrot(q, V3{0, 1, 0}), cz = qrot(q, V3{0, 0, 1});
    const M3 R = m3_make(cx.x, cy.x, cz.x, cx.y, cy.y, cz.y, cx.z, cy.z, cz.z);
    M3 D = m3_zero();
    D.m[0] = ii.x; D.m[4] = ii.y; D.m[8] = ii.z;
    const M3 Rt = m3_make(R.m[0], R.m[3], R.m[6], R.m[1], R.m[4], R.m[7], R.m[2], R.m[5], R.m[8]);
    return m3_mul(m3_mul(R, D), Rt);
}

__device__ __forceinline__ V3 perpendicular(V3 v) {   // world.cpp:200-203
    const V3 a = fabs(v.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
    return normalized(cross(v, a));
}

__device__ __forceinline__ Qt qconj_(Qt a) { return Qt{a.w, -a.x, -a.y, -a.z}; }
__device__ __forceinline__ Qt qmul_(Qt a, Qt o) {
    return Qt{a.w * o.w - a.x * o.x - a.y * o.y - a.z * o.z, a.w * o.x + a.x * o.w + a.y * o.z - a.z * o.y,
              a.w * o.y - a.x * o.z + a.y * o.w + a.z * o.x, a.w * o.z + a.x * o.y - a.y * o.x + a.z * o.w};
}

// K = (mA + mB) I - [rA]x IA [rA]x - [rB]x IB [rB]x (world.cpp:232-234, 248-250)
__device__ __forceinline__ M3 point_mass_matrix(double ima, double imb, const M3& IA, const M3& IB, V3 rA, V3 rB) {
    M3 s = m3_zero();
    s.m[0] = s.m[4] = s.m[8] = ima + imb;
    const M3 sa = m3_skew(rA), sb = m3_skew(rB);
    return m3_sub(m3_sub(s, m3_mul(m3_mul(sa, IA), sa)), m3_mul(m3_mul(sb, IB), sb));
}

// The joints' device table: the records as given (broken updated in place),
// and the step's accumulators.
struct GJointDev {
    ne_joint* j;
    double *applied, *motor;
    const uint32_t* comp_start;   // ncomp + 1
    const uint32_t* comp_list;    // joint indices, ascending within a component
    uint32_t ncomp;
};

// One component per thread: its joints' accumulators reset, 8 iterations in
// order, then the break test (world.cpp:206-315).
__global__ void __launch_bounds__(64) k_joint_general(BodiesDev B, SimParams P, GJointDev G) {
    const double beta = 0.2, dt = P.dt;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < G.ncomp; c += gridDim.x * blockDim.x) {
        const uint32_t t0 = G.comp_start[c], t1 = G.comp_start[c + 1];
        for (uint32_t t = t0; t < t1; ++t) {
            G.applied[G.comp_list[t]] = 0.0;
            G.motor[G.comp_list[t]] = 0.0;
        }
        for (int it = 0; it < 8; ++it)
            for (uint32_t t = t0; t < t1; ++t) {
                const uint32_t k = G.comp_list[t];
                const ne_joint& j = G.j[k];
                if (j.broken) continue;
                const uint32_t a = j.a, b = j.b;
                BodyView A = load_view(B, a), Bv = load_view(B, b);
                const M3 IA = inv_inertia_world(A.q, A.ii), IB = inv_inertia_world(Bv.q, Bv.ii);
                const V3 rA = qrot(A.q, V3{j.local_a[0], j.local_a[1], j.local_a[2]});
                const V3 rB = qrot(Bv.q, V3{j.local_b[0], j.local_b[1], j.local_b[2]});
                const V3 xa = ld3(B.x0, B.x1, B.x2, a), xb = ld3(B.x0, B.x1, B.x2, b);
                const V3 posErr = min_image(P.box, (xa + rA) - (xb + rB));
                V3 vRel = (A.v + cross(A.w, rA)) - (Bv.v + cross(Bv.w, rB));
                double applied = G.applied[k];
                if (j.type == NE_JOINT_SLIDER) {
                    const V3 ax = normalized(qrot(A.q, V3{j.axis_a[0], j.axis_a[1], j.axis_a[2]}));
                    const V3 t1v = perpendicular(ax), t2v = normalized(cross(ax, t1v));
                    const M3 K = point_mass_matrix(A.im, Bv.im, IA, IB, rA, rB);
                    for (int qq = 0; qq < 2; ++qq) {
                        const V3 tv = qq == 0 ? t1v : t2v;
                        const double jv = dot(vRel, tv) + dot(posErr, tv) * (beta / dt);
                        const double kk = dot(tv, m3_mulv(K, tv));
                        if (kk > 1e-18) {
                            const double l = -jv / kk;
                            const V3 Pi = tv * l;
                            if (A.mv) A.v = A.v + Pi * A.im;
                            if (A.mw) A.w = A.w + m3_mulv(IA, cross(rA, Pi));
                            if (Bv.mv) Bv.v = Bv.v - Pi * Bv.im;
                            if (Bv.mw) Bv.w = Bv.w - m3_mulv(IB, cross(rB, Pi));
                            applied += fabs(l);
                            vRel = (A.v + cross(A.w, rA)) - (Bv.v + cross(Bv.w, rB));
                        }
                    }
                } else {
                    const M3 K = point_mass_matrix(A.im, Bv.im, IA, IB, rA, rB);
                    const V3 Pi = m3_solve(K, -(vRel + posErr * (beta / dt)));
                    if (A.mv) A.v = A.v + Pi * A.im;
                    if (A.mw) A.w = A.w + m3_mulv(IA, cross(rA, Pi));
                    if (Bv.mv) Bv.v = Bv.v - Pi * Bv.im;
                    if (Bv.mw) Bv.w = Bv.w - m3_mulv(IB, cross(rB, Pi));
                    applied += norm(Pi);
                }
                if (j.type == NE_JOINT_FIXED || j.type == NE_JOINT_SLIDER) {
                    const V3 wRel = A.w - Bv.w;
                    const Qt rr = Qt{j.ref_rel[0], j.ref_rel[1], j.ref_rel[2], j.ref_rel[3]};
                    const Qt dq = qmul_(qmul_(A.q, qconj_(Bv.q)), qconj_(rr));
                    V3 angErr{dq.x, dq.y, dq.z};
                    if (dq.w < 0.0) angErr = -angErr;
                    angErr = angErr * 2.0;
                    const M3 Kang = m3_add(IA, IB);
                    const V3 L = m3_solve(Kang, -(wRel + angErr * (beta / dt)));
                    if (A.mw) A.w = A.w + m3_mulv(IA, L);
                    if (Bv.mw) Bv.w = Bv.w - m3_mulv(IB, L);
                } else if (j.type == NE_JOINT_HINGE) {
                    const V3 aA = normalized(qrot(A.q, V3{j.axis_a[0], j.axis_a[1], j.axis_a[2]}));
                    const V3 aB = normalized(qrot(Bv.q, V3{j.axis_b[0], j.axis_b[1], j.axis_b[2]}));
                    const V3 t1v = perpendicular(aA), t2v = normalized(cross(aA, t1v));
                    const V3 alignErr = cross(aA, aB);
                    const M3 Kang = m3_add(IA, IB);
                    for (int qq = 0; qq < 2; ++qq) {
                        const V3 tv = qq == 0 ? t1v : t2v;
                        const V3 wRel = A.w - Bv.w;
                        const double jv = dot(wRel, tv) - dot(alignErr, tv) * (beta / dt);
                        const double kk = dot(tv, m3_mulv(Kang, tv));
                        if (kk > 1e-18) {
                            const V3 L = tv * (-jv / kk);
                            if (A.mw) A.w = A.w + m3_mulv(IA, L);
                            if (Bv.mw) Bv.w = Bv.w - m3_mulv(IB, L);
                        }
                    }
                    if (j.use_motor) {
                        const double kk = dot(aA, m3_mulv(Kang, aA));
                        if (kk > 1e-18) {
                            const double jv = dot(A.w - Bv.w, aA) - j.motor_speed;
                            double l = -jv / kk;
                            const double old = G.motor[k];
                            const double sum = old + l, lim = j.max_motor_impulse;
                            const double nn = sum < -lim ? -lim : (sum > lim ? lim : sum);
                            l = nn - old;
                            G.motor[k] = nn;
                            const V3 L = aA * l;
                            if (A.mw) A.w = A.w + m3_mulv(IA, L);
                            if (Bv.mw) Bv.w = Bv.w - m3_mulv(IB, L);
                        }
                    }
                    if (j.use_limit) {
                        const Qt rr = Qt{j.ref_rel[0], j.ref_rel[1], j.ref_rel[2], j.ref_rel[3]};
                        const Qt dq = qmul_(qmul_(A.q, qconj_(Bv.q)), qconj_(rr));
                        const V3 dv{dq.x, dq.y, dq.z};
                        const double angle = 2.0 * atan2_rn(dot(dv, aA), dq.w);
                        double cc = 0.0;
                        bool active = false;
                        if (angle <= j.lower) { cc = angle - j.lower; active = true; }
                        else if (angle >= j.upper) { cc = angle - j.upper; active = true; }
                        if (active) {
                            const double kk = dot(aA, m3_mulv(Kang, aA));
                            if (kk > 1e-18) {
                                const double jv = dot(A.w - Bv.w, aA) + cc * (beta / dt);
                                const V3 L = aA * (-jv / kk);
                                if (A.mw) A.w = A.w + m3_mulv(IA, L);
                                if (Bv.mw) Bv.w = Bv.w - m3_mulv(IB, L);
                            }
                        }
                    }
                }
                G.applied[k] = applied;
                store_view(B, a, A);
                store_view(B, b, Bv);
            }
        for (uint32_t t = t0; t < t1; ++t) {
            const uint32_t k = G.comp_list[t];
            ne_joint& j = G.j[k];
            if (j.breakable && !j.broken && G.applied[k] > j.break_impulse) j.broken = 1;
        }
    }
}

// Island sleeping: bodies joined by an unbroken general joint share an
// island (world.cpp:343).
__global__ void __launch_bounds__(TPB) k_sleep_join_gjoints(BodiesDev B, const ne_joint* __restrict__ gj, uint32_t n,
                                                            uint32_t* __restrict__ parent) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && !gj[k].broken) sleep_union(B, parent, gj[k].a, gj[k].b);
}

}  // namespace nea

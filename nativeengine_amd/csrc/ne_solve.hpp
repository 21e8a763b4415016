// ne_solve.hpp -- the sequential-impulse primitives of one contact
// (host and device): material combine, effective mass, impulse application,
// warm start and one Gauss-Seidel visit, in the reference's operation order
// (src/world.cpp:42-51, 559-611; src/contact.hpp:28-33).  The kernels use
// them on device; the adjoint checks evaluate them on the host.
#pragma once

#include "../../include/ne_gpu.h"
#include "ne_dmath.hpp"

namespace nea {

NE_HD double combine_mat(int m, double a, double b) {   // world.cpp:42-51
    switch (m) {
        case NE_COMBINE_AVERAGE: return 0.5 * (a + b);
        case NE_COMBINE_MIN: return a < b ? a : b;
        case NE_COMBINE_MAX: return a > b ? a : b;
        case NE_COMBINE_MULTIPLY: return a * b;
        case NE_COMBINE_GEOMETRIC_MEAN: return sqrt((a > 0 ? a : 0) * (b > 0 ? b : 0));
    }
    return 0.5 * (a + b);
}

NE_HD double eff_mass(double ima, Qt qa, V3 iia, double imb, Qt qb, V3 iib, V3 ra, V3 rb, V3 u) {
    V3 angA = cross(apply_inv_inertia(qa, iia, cross(ra, u)), ra);
    V3 angB = cross(apply_inv_inertia(qb, iib, cross(rb, u)), rb);
    return ima + imb + dot(angA + angB, u);
}

// One body as the solver sees it.  Writes of v (w) are skipped for bodies
// whose inverse mass (inverse inertia) is zero: the reference adds +-0 there.
struct BodyView {
    V3 v, w;
    Qt q;
    V3 ii;
    double im;
    bool mv, mw;
};

// a.v += J m_a^-1, a.w += I_a^-1 (rA x J); b gets the opposite (world.cpp:565-566)
NE_HD void apply_impulse(BodyView& a, BodyView& b, V3 rA, V3 rB, V3 J) {
    if (a.mv) a.v = a.v + J * a.im;
    if (a.mw) a.w = a.w + apply_inv_inertia(a.q, a.ii, cross(rA, J));
    if (b.mv) b.v = b.v - J * b.im;
    if (b.mw) b.w = b.w - apply_inv_inertia(b.q, b.ii, cross(rB, J));
}

struct ConsGeom {
    V3 n, rA, rB;
};

NE_HD void warm_apply(BodyView& a, BodyView& b, const ConsGeom& g, double nimp) {
    if (nimp == 0.0) return;
    apply_impulse(a, b, g.rA, g.rB, g.n * nimp);
}

// One Gauss-Seidel visit of constraint k (world.cpp:574-611).
NE_HD void gs_visit(BodyView& a, BodyView& b, const ConsGeom& g, double rbias, double kn, double mu, double& nimp,
                    double& timp) {
    const V3 vc = (a.v + cross(a.w, g.rA)) - (b.v + cross(b.w, g.rB));
    const double vn = dot(vc, g.n);
    if (kn > 1e-18) {
        double dj = (rbias - vn) / kn;
        const double old = nimp;
        nimp = (old + dj > 0.0) ? old + dj : 0.0;
        dj = nimp - old;
        apply_impulse(a, b, g.rA, g.rB, g.n * dj);
    }
    if (mu > 0.0 && nimp > 0.0) {
        const V3 vc2 = (a.v + cross(a.w, g.rA)) - (b.v + cross(b.w, g.rB));
        const V3 vt = vc2 - g.n * dot(vc2, g.n);
        const double vtl = norm(vt);
        if (vtl > 1e-12) {
            const V3 t = vt / vtl;
            const double Kt = eff_mass(a.im, a.q, a.ii, b.im, b.q, b.ii, g.rA, g.rB, t);
            if (Kt > 1e-18) {
                double djt = -vtl / Kt;
                const double maxF = mu * nimp;
                const double old = timp;
                const double sum = old + djt;
                timp = (sum < -maxF) ? -maxF : (sum > maxF ? maxF : sum);
                djt = timp - old;
                apply_impulse(a, b, g.rA, g.rB, t * djt);
            }
        }
    }
}

}  // namespace nea

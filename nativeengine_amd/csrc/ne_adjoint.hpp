// ne_adjoint.hpp -- reverse-mode (vector-Jacobian) rules for the FP64 step
// math of ne_dmath.hpp.
//
// Each *_rev function takes the adjoint of an operation's output and ADDS
// the adjoints of its inputs into the `...b` arguments.  The rules are the
// exact derivatives of the formulas as written (a quaternion rotation is
// differentiated as the expression v + w t + q_vec x t, not as a rotation
// matrix), so they hold for the slightly non-unit quaternions the step
// produces.  Adjoints need no bit-level agreement with anything; they are
// checked against finite differences of the forward step.
#pragma once

#include "ne_dmath.hpp"
#include "ne_solve.hpp"

namespace nea {

NE_HD void cross_rev(V3 a, V3 b, V3 cb, V3& ab, V3& bb) {   // c = a x b
    ab = ab + cross(b, cb);
    bb = bb + cross(cb, a);
}

// out = v + w t + qv x t, t = 2 (qv x v)
NE_HD void qrot_rev(Qt q, V3 v, V3 ob, Qt& qb, V3& vb) {
    const V3 qv{q.x, q.y, q.z};
    const V3 t = cross(qv, v) * 2.0;
    vb = vb + ob;
    V3 tb = ob * q.w;
    qb.w += dot(ob, t);
    V3 qvb{0.0, 0.0, 0.0};
    cross_rev(qv, t, ob, qvb, tb);
    const V3 cb = tb * 2.0;
    cross_rev(qv, v, cb, qvb, vb);
    qb.x += qvb.x;
    qb.y += qvb.y;
    qb.z += qvb.z;
}

NE_HD void qinvrot_rev(Qt q, V3 v, V3 ob, Qt& qb, V3& vb) {
    Qt cb{0.0, 0.0, 0.0, 0.0};
    qrot_rev(qconj(q), v, ob, cb, vb);
    qb.w += cb.w;
    qb.x -= cb.x;
    qb.y -= cb.y;
    qb.z -= cb.z;
}

// out = qrot(q, d (.) qinvrot(q, u)): apply_inv_inertia (d = invI) and
// angular_momentum (d = inertia_of(invI)).
NE_HD void sandwich_rev(Qt q, V3 d, V3 u, V3 ob, Qt& qb, V3& ub) {
    const V3 body = qinvrot(q, u);
    const V3 sb{body.x * d.x, body.y * d.y, body.z * d.z};
    V3 sbb{0.0, 0.0, 0.0};
    qrot_rev(q, sb, ob, qb, sbb);
    const V3 bodyb{sbb.x * d.x, sbb.y * d.y, sbb.z * d.z};
    qinvrot_rev(q, u, bodyb, qb, ub);
}

NE_HD void qmul_rev(Qt a, Qt o, Qt mb, Qt& ab, Qt& obb) {   // m = a (x) o
    ab.w += mb.w * o.w + mb.x * o.x + mb.y * o.y + mb.z * o.z;
    ab.x += -mb.w * o.x + mb.x * o.w - mb.y * o.z + mb.z * o.y;
    ab.y += -mb.w * o.y + mb.x * o.z + mb.y * o.w - mb.z * o.x;
    ab.z += -mb.w * o.z - mb.x * o.y + mb.y * o.x + mb.z * o.w;
    obb.w += mb.w * a.w + mb.x * a.x + mb.y * a.y + mb.z * a.z;
    obb.x += -mb.w * a.x + mb.x * a.w + mb.y * a.z - mb.z * a.y;
    obb.y += -mb.w * a.y - mb.x * a.z + mb.y * a.w + mb.z * a.x;
    obb.z += -mb.w * a.z + mb.x * a.y - mb.y * a.x + mb.z * a.w;
}

// out = normalize(q + (omega (x) q) dt/2)
NE_HD void integrate_orientation_rev(Qt q, V3 om, double dt, Qt ob, Qt& qb, V3& omb) {
    const Qt oq{0.0, om.x, om.y, om.z};
    const Qt m = qmul(oq, q);
    const double h = 0.5 * dt;
    const Qt s{q.w + m.w * h, q.x + m.x * h, q.y + m.y * h, q.z + m.z * h};
    const double n2 = s.w * s.w + s.x * s.x + s.y * s.y + s.z * s.z;
    const double n = sqrt(n2);
    if (!(n > 1e-300)) return;   // the forward returns the identity: no dependence
    const double sd = (s.w * ob.w + s.x * ob.x + s.y * ob.y + s.z * ob.z) / (n2 * n);
    const Qt sb{ob.w / n - s.w * sd, ob.x / n - s.x * sd, ob.y / n - s.y * sd, ob.z / n - s.z * sd};
    qb.w += sb.w;
    qb.x += sb.x;
    qb.y += sb.y;
    qb.z += sb.z;
    const Qt mbb{sb.w * h, sb.x * h, sb.y * h, sb.z * h};
    Qt oqb{0.0, 0.0, 0.0, 0.0};
    qmul_rev(oq, q, mbb, oqb, qb);
    omb.x += oqb.x;
    omb.y += oqb.y;
    omb.z += oqb.z;
}

// ---------------------------------------------------------------------------
// One contact's solver steps
// ---------------------------------------------------------------------------

// apply_impulse(a, b, rA, rB, J) adds terms in J to the velocities, so the
// velocity adjoints pass through unchanged; this adds J's, the lever arms'
// and the orientations' adjoints.
NE_HD void apply_impulse_rev(const BodyView& a, const BodyView& b, V3 rA, V3 rB, V3 J, V3 avb, V3 awb, V3 bvb,
                             V3 bwb, V3& Jb, Qt& qab, Qt& qbb, V3& rAb, V3& rBb) {
    if (a.mv) Jb = Jb + avb * a.im;
    if (a.mw) {
        V3 ub{0.0, 0.0, 0.0};
        sandwich_rev(a.q, a.ii, cross(rA, J), awb, qab, ub);
        cross_rev(rA, J, ub, rAb, Jb);
    }
    if (b.mv) Jb = Jb - bvb * b.im;
    if (b.mw) {
        V3 ub{0.0, 0.0, 0.0};
        sandwich_rev(b.q, b.ii, cross(rB, J), -bwb, qbb, ub);
        cross_rev(rB, J, ub, rBb, Jb);
    }
}

NE_HD void eff_mass_rev(Qt qa, V3 iia, Qt qb, V3 iib, V3 ra, V3 rb, V3 u, double Kb, Qt& qab, Qt& qbb, V3& rab,
                        V3& rbb, V3& ub) {
    const V3 ia = cross(ra, u), ib = cross(rb, u);
    const V3 Xa = apply_inv_inertia(qa, iia, ia), Xb = apply_inv_inertia(qb, iib, ib);
    const V3 angA = cross(Xa, ra), angB = cross(Xb, rb);
    ub = ub + (angA + angB) * Kb;
    const V3 angb = u * Kb;
    V3 Xab{0.0, 0.0, 0.0}, Xbb{0.0, 0.0, 0.0}, iab{0.0, 0.0, 0.0}, ibb{0.0, 0.0, 0.0};
    cross_rev(Xa, ra, angb, Xab, rab);
    cross_rev(Xb, rb, angb, Xbb, rbb);
    sandwich_rev(qa, iia, ia, Xab, qab, iab);
    sandwich_rev(qb, iib, ib, Xbb, qbb, ibb);
    cross_rev(ra, u, iab, rab, ub);
    cross_rev(rb, u, ibb, rbb, ub);
}

// Adjoints of one contact's solver inputs, accumulated by gs_visit_rev.
struct ConsAdj {
    V3 rA, rB;
    double rbias, kn, mu;
};

// Reverse of gs_visit.  a0/b0 hold the pre-visit velocities (and q, ii, im),
// nimp0/timp0 the pre-visit accumulated impulses; the forward is recomputed
// from them.  On entry av..bw, nimpb, timpb are the adjoints after the visit;
// on exit, before it.  Orientation adjoints go to qab/qbb.
NE_HD void gs_visit_rev(const BodyView& a0, const BodyView& b0, const ConsGeom& g, double rbias, double kn,
                        double mu, double nimp0, double timp0, V3& avb, V3& awb, V3& bvb, V3& bwb, double& nimpb,
                        double& timpb, ConsAdj& ca, Qt& qab, Qt& qbb) {
    BodyView a = a0, b = b0;
    const V3 vc = (a.v + cross(a.w, g.rA)) - (b.v + cross(b.w, g.rB));
    const double vn = dot(vc, g.n);
    const bool nrm = kn > 1e-18;
    double dj = 0.0, nimp1 = nimp0, dj1 = 0.0;
    bool pos = false;
    if (nrm) {
        dj = (rbias - vn) / kn;
        pos = nimp0 + dj > 0.0;
        nimp1 = pos ? nimp0 + dj : 0.0;
        dj1 = nimp1 - nimp0;
        apply_impulse(a, b, g.rA, g.rB, g.n * dj1);
    }
    const BodyView a1 = a, b1 = b;
    bool fr = false;
    V3 vt{0.0, 0.0, 0.0}, t{0.0, 0.0, 0.0};
    double vtl = 0.0, Kt = 0.0, djt1 = 0.0;
    int clamp = 0;
    if (mu > 0.0 && nimp1 > 0.0) {
        const V3 vc2 = (a1.v + cross(a1.w, g.rA)) - (b1.v + cross(b1.w, g.rB));
        vt = vc2 - g.n * dot(vc2, g.n);
        vtl = norm(vt);
        if (vtl > 1e-12) {
            t = vt / vtl;
            Kt = eff_mass(a1.im, a1.q, a1.ii, b1.im, b1.q, b1.ii, g.rA, g.rB, t);
            if (Kt > 1e-18) {
                fr = true;
                const double djt = -vtl / Kt;
                const double maxF = mu * nimp1;
                const double sum = timp0 + djt;
                clamp = (sum < -maxF) ? -1 : (sum > maxF ? 1 : 0);
                const double timp1 = clamp < 0 ? -maxF : (clamp > 0 ? maxF : sum);
                djt1 = timp1 - timp0;
            }
        }
    }
    double nimp1b = nimpb, timp0b = timpb;
    if (fr) {
        V3 Jb{0.0, 0.0, 0.0};
        apply_impulse_rev(a1, b1, g.rA, g.rB, t * djt1, avb, awb, bvb, bwb, Jb, qab, qbb, ca.rA, ca.rB);
        V3 tb = Jb * djt1;
        const double djt1b = dot(Jb, t);
        const double timp1b = timpb + djt1b;
        double oldb = -djt1b, maxFb = 0.0, sumb = 0.0;
        if (clamp < 0) maxFb -= timp1b;
        else if (clamp > 0) maxFb += timp1b;
        else sumb += timp1b;
        oldb += sumb;
        const double djtb = sumb;
        ca.mu += maxFb * nimp1;
        nimp1b += maxFb * mu;
        double vtlb = -djtb / Kt;
        const double Ktb = djtb * vtl / (Kt * Kt);
        eff_mass_rev(a1.q, a1.ii, b1.q, b1.ii, g.rA, g.rB, t, Ktb, qab, qbb, ca.rA, ca.rB, tb);
        V3 vtb = tb / vtl;
        vtlb -= dot(tb, vt) / (vtl * vtl);
        vtb = vtb + vt * (vtlb / vtl);
        const V3 vc2b = vtb - g.n * dot(g.n, vtb);
        avb = avb + vc2b;
        bvb = bvb - vc2b;
        cross_rev(a1.w, g.rA, vc2b, awb, ca.rA);
        cross_rev(b1.w, g.rB, -vc2b, bwb, ca.rB);
        timp0b = oldb;
    }
    double nimp0b = nimp1b;
    if (nrm) {
        V3 Jb{0.0, 0.0, 0.0};
        apply_impulse_rev(a0, b0, g.rA, g.rB, g.n * dj1, avb, awb, bvb, bwb, Jb, qab, qbb, ca.rA, ca.rB);
        const double dj1b = dot(Jb, g.n);
        nimp1b += dj1b;
        double oldb = -dj1b, djb = 0.0;
        if (pos) {
            oldb += nimp1b;
            djb += nimp1b;
        }
        ca.rbias += djb / kn;
        const double vnb = -djb / kn;
        ca.kn += -djb * dj / kn;
        const V3 vcb = g.n * vnb;
        avb = avb + vcb;
        bvb = bvb - vcb;
        cross_rev(a0.w, g.rA, vcb, awb, ca.rA);
        cross_rev(b0.w, g.rB, -vcb, bwb, ca.rB);
        nimp0b = oldb;
    }
    nimpb = nimp0b;
    timpb = timp0b;
}

// d combine_mat(m, a, b) / d w, where a (b) is the world default w when a_w
// (b_w) -- a body with no material of its own -- and a constant otherwise.
NE_HD double combine_dw(int m, double a, double b, bool a_w, bool b_w) {
    const double da = a_w ? 1.0 : 0.0, db = b_w ? 1.0 : 0.0;
    switch (m) {
        case NE_COMBINE_AVERAGE: return 0.5 * (da + db);
        case NE_COMBINE_MIN: return a < b ? da : db;
        case NE_COMBINE_MAX: return a > b ? da : db;
        case NE_COMBINE_MULTIPLY: return da * b + db * a;
        case NE_COMBINE_GEOMETRIC_MEAN: {
            if (!(a > 0.0) || !(b > 0.0)) return 0.0;
            const double s = sqrt(a * b);
            return (da * b + db * a) / (2.0 * s);
        }
    }
    return 0.5 * (da + db);
}

}  // namespace nea

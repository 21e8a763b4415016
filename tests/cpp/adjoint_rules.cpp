// Finite-difference check of the reverse-mode rules in
// nativeengine_amd/csrc/ne_adjoint.hpp, evaluated on the host.
//
// For each rule y = f(x): draw x, an output adjoint yb and an input direction
// dx; the reverse rule gives xb = J^T yb, so <xb, dx> must equal
// <yb, (f(x + h dx) - f(x - h dx)) / 2h>.  Prints the worst relative error
// per rule and exits 1 if any exceeds the tolerance.
#include <cmath>
#include <cstdio>
#include <functional>
#include <random>
#include <vector>

#include "../../nativeengine_amd/csrc/ne_adjoint.hpp"

using namespace nea;

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }
static V3 RV(double s = 1.0) { return V3{U(-s, s), U(-s, s), U(-s, s)}; }
static Qt RQ() {
    Qt q{U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)};
    const double n = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    return Qt{q.w / n, q.x / n, q.y / n, q.z / n};
}

// packing helpers: a flat parameter vector <-> structured inputs
struct Pack {
    std::vector<double> v;
    void put(V3 a) { v.push_back(a.x); v.push_back(a.y); v.push_back(a.z); }
    void put(Qt a) { v.push_back(a.w); v.push_back(a.x); v.push_back(a.y); v.push_back(a.z); }
    void put(double a) { v.push_back(a); }
};
struct Unpack {
    const double* p;
    V3 v3() { V3 a{p[0], p[1], p[2]}; p += 3; return a; }
    Qt q() { Qt a{p[0], p[1], p[2], p[3]}; p += 4; return a; }
    double d() { return *p++; }
};

static double worst = 0.0;
static int failures = 0;

// f maps inputs to outputs (flat); rev maps (inputs, output adjoint) to the
// input adjoint (flat, same layout as the inputs).
static void check(const char* name, const std::vector<double>& x0, int nout,
                  const std::function<std::vector<double>(const std::vector<double>&)>& f,
                  const std::function<std::vector<double>(const std::vector<double>&, const std::vector<double>&)>& rev,
                  double tol = 1e-6, int trials = 1, int frozen = -1) {
    double w = 0.0;
    for (int t = 0; t < trials; ++t) {
        std::vector<double> yb(nout), dx(x0.size());
        for (auto& e : yb) e = U(-1, 1);
        for (auto& e : dx) e = U(-1, 1);
        if (frozen >= 0) dx[(size_t)frozen] = 0.0;   // an input held at a kink (e.g. mu == 0)
        const std::vector<double> xb = rev(x0, yb);
        double lhs = 0.0;
        for (size_t k = 0; k < dx.size(); ++k) lhs += xb[k] * dx[k];
        const double h = 1e-6;
        std::vector<double> xp = x0, xm = x0;
        for (size_t k = 0; k < dx.size(); ++k) { xp[k] += h * dx[k]; xm[k] -= h * dx[k]; }
        const std::vector<double> fp = f(xp), fm = f(xm);
        double rhs = 0.0;
        for (int k = 0; k < nout; ++k) rhs += yb[k] * (fp[k] - fm[k]) / (2.0 * h);
        const double err = std::fabs(lhs - rhs) / std::fmax(std::fmax(std::fabs(lhs), std::fabs(rhs)), 1e-6);
        w = std::fmax(w, err);
    }
    std::printf("%-34s rel err %.3e\n", name, w);
    worst = std::fmax(worst, w);
    if (!(w <= tol)) ++failures;
}

static std::vector<double> out3(V3 a) { return {a.x, a.y, a.z}; }

int main() {
    for (int rep = 0; rep < 20; ++rep) {
        // qrot / qinvrot / sandwich / integrate_orientation ------------------
        {
            Pack P; P.put(RQ()); P.put(RV());
            check("qrot", P.v, 3,
                  [](const std::vector<double>& x) { Unpack u{x.data()}; Qt q = u.q(); V3 v = u.v3(); return out3(qrot(q, v)); },
                  [](const std::vector<double>& x, const std::vector<double>& yb) {
                      Unpack u{x.data()}; Qt q = u.q(); V3 v = u.v3();
                      Qt qb{0, 0, 0, 0}; V3 vb{0, 0, 0};
                      qrot_rev(q, v, V3{yb[0], yb[1], yb[2]}, qb, vb);
                      Pack r; r.put(qb); r.put(vb); return r.v;
                  });
            check("qinvrot", P.v, 3,
                  [](const std::vector<double>& x) { Unpack u{x.data()}; Qt q = u.q(); V3 v = u.v3(); return out3(qinvrot(q, v)); },
                  [](const std::vector<double>& x, const std::vector<double>& yb) {
                      Unpack u{x.data()}; Qt q = u.q(); V3 v = u.v3();
                      Qt qb{0, 0, 0, 0}; V3 vb{0, 0, 0};
                      qinvrot_rev(q, v, V3{yb[0], yb[1], yb[2]}, qb, vb);
                      Pack r; r.put(qb); r.put(vb); return r.v;
                  });
            const V3 d{U(0.5, 3), U(0.5, 3), U(0.5, 3)};
            check("apply_inv_inertia (q, u)", P.v, 3,
                  [d](const std::vector<double>& x) { Unpack u{x.data()}; Qt q = u.q(); V3 v = u.v3(); return out3(apply_inv_inertia(q, d, v)); },
                  [d](const std::vector<double>& x, const std::vector<double>& yb) {
                      Unpack u{x.data()}; Qt q = u.q(); V3 v = u.v3();
                      Qt qb{0, 0, 0, 0}; V3 vb{0, 0, 0};
                      sandwich_rev(q, d, v, V3{yb[0], yb[1], yb[2]}, qb, vb);
                      Pack r; r.put(qb); r.put(vb); return r.v;
                  });
            const double dt = 1.0 / 60.0;
            Pack Q; Q.put(RQ()); Q.put(RV(5.0));
            check("integrate_orientation (q, w)", Q.v, 4,
                  [dt](const std::vector<double>& x) {
                      Unpack u{x.data()}; Qt q = u.q(); V3 w = u.v3();
                      Qt o = integrate_orientation(q, w, dt); return std::vector<double>{o.w, o.x, o.y, o.z};
                  },
                  [dt](const std::vector<double>& x, const std::vector<double>& yb) {
                      Unpack u{x.data()}; Qt q = u.q(); V3 w = u.v3();
                      Qt qb{0, 0, 0, 0}; V3 wb{0, 0, 0};
                      integrate_orientation_rev(q, w, dt, Qt{yb[0], yb[1], yb[2], yb[3]}, qb, wb);
                      Pack r; r.put(qb); r.put(wb); return r.v;
                  });
        }
        // eff_mass ------------------------------------------------------------
        {
            const V3 ia{U(1, 8), U(1, 8), U(1, 8)}, ib{U(1, 8), U(1, 8), U(1, 8)};
            Pack P; P.put(RQ()); P.put(RQ()); P.put(RV(0.5)); P.put(RV(0.5)); P.put(RV());
            check("eff_mass (qa, qb, ra, rb, u)", P.v, 1,
                  [ia, ib](const std::vector<double>& x) {
                      Unpack u{x.data()}; Qt qa = u.q(), qb = u.q(); V3 ra = u.v3(), rb = u.v3(), n = u.v3();
                      return std::vector<double>{eff_mass(0.7, qa, ia, 1.3, qb, ib, ra, rb, n)};
                  },
                  [ia, ib](const std::vector<double>& x, const std::vector<double>& yb) {
                      Unpack u{x.data()}; Qt qa = u.q(), qb = u.q(); V3 ra = u.v3(), rb = u.v3(), n = u.v3();
                      Qt qab{0, 0, 0, 0}, qbb{0, 0, 0, 0}; V3 rab{0, 0, 0}, rbb{0, 0, 0}, nb{0, 0, 0};
                      eff_mass_rev(qa, ia, qb, ib, ra, rb, n, yb[0], qab, qbb, rab, rbb, nb);
                      Pack r; r.put(qab); r.put(qbb); r.put(rab); r.put(rbb); r.put(nb); return r.v;
                  });
        }
        // gs_visit in its regimes ---------------------------------------------
        // inputs: va wa vb wb (12) nimp timp (2) rA rB (6) rbias kn mu (3) qa qb (8); n fixed
        for (int regime = 0; regime < 5; ++regime) {
            const V3 n = normalized(RV());
            const V3 iia{U(1, 8), U(1, 8), U(1, 8)}, iib{U(1, 8), U(1, 8), U(1, 8)};
            const double ima = 0.8, imb = (regime == 4) ? 0.0 : 1.1;   // regime 4: b static (ground)
            const V3 iibx = (regime == 4) ? V3{0, 0, 0} : iib;
            Pack P;
            V3 va = RV(), vb = RV();
            // approach along the normal so the normal impulse is active
            const double approach = (regime == 3) ? 0.5 : -1.0;   // regime 3: separating -> clamp to 0
            va = va + n * (approach - dot(va - vb, n));
            P.put(va); P.put(RV()); P.put(vb); P.put(RV());
            P.put(regime == 1 ? 0.3 : 0.0); P.put(0.0);
            P.put(RV(0.4)); P.put(RV(0.4));
            const double kn = eff_mass(ima, RQ(), iia, imb, RQ(), iibx, RV(0.4), RV(0.4), n);
            P.put(regime == 2 ? 0.0 : 0.4); P.put(kn);
            P.put(regime == 0 ? 0.0 : (regime == 2 ? 5.0 : 0.3));   // 0: frictionless, 2: sticking
            P.put(RQ()); P.put(RQ());
            auto unpack = [=](const std::vector<double>& x, BodyView& a, BodyView& b, ConsGeom& g, double& rbias,
                              double& knv, double& mu, double& nimp, double& timp) {
                Unpack u{x.data()};
                a.v = u.v3(); a.w = u.v3(); b.v = u.v3(); b.w = u.v3();
                nimp = u.d(); timp = u.d();
                g.n = n; g.rA = u.v3(); g.rB = u.v3();
                rbias = u.d(); knv = u.d(); mu = u.d();
                a.q = u.q(); b.q = u.q();
                a.ii = iia; b.ii = iibx; a.im = ima; b.im = imb;
                a.mv = ima != 0.0; a.mw = true; b.mv = imb != 0.0; b.mw = b.ii.x != 0.0;
            };
            char name[64];
            std::snprintf(name, sizeof name, "gs_visit regime %d", regime);
            check(name, P.v, 14,
                  [=](const std::vector<double>& x) {
                      BodyView a, b; ConsGeom g; double rbias, knv, mu, nimp, timp;
                      unpack(x, a, b, g, rbias, knv, mu, nimp, timp);
                      gs_visit(a, b, g, rbias, knv, mu, nimp, timp);
                      return std::vector<double>{a.v.x, a.v.y, a.v.z, a.w.x, a.w.y, a.w.z, b.v.x, b.v.y, b.v.z,
                                                 b.w.x, b.w.y, b.w.z, nimp, timp};
                  },
                  [=](const std::vector<double>& x, const std::vector<double>& yb) {
                      BodyView a, b; ConsGeom g; double rbias, knv, mu, nimp, timp;
                      unpack(x, a, b, g, rbias, knv, mu, nimp, timp);
                      V3 avb{yb[0], yb[1], yb[2]}, awb{yb[3], yb[4], yb[5]}, bvb{yb[6], yb[7], yb[8]}, bwb{yb[9], yb[10], yb[11]};
                      double nb = yb[12], tb = yb[13];
                      ConsAdj ca{{0, 0, 0}, {0, 0, 0}, 0, 0, 0};
                      Qt qab{0, 0, 0, 0}, qbb{0, 0, 0, 0};
                      gs_visit_rev(a, b, g, rbias, knv, mu, nimp, timp, avb, awb, bvb, bwb, nb, tb, ca, qab, qbb);
                      Pack r; r.put(avb); r.put(awb); r.put(bvb); r.put(bwb); r.put(nb); r.put(tb);
                      r.put(ca.rA); r.put(ca.rB); r.put(ca.rbias); r.put(ca.kn); r.put(ca.mu); r.put(qab); r.put(qbb);
                      return r.v;
                  },
                  1e-5, 1, regime == 0 ? 22 : -1);
        }
    }
    std::printf("worst %.3e, %d failing\n", worst, failures);
    return failures ? 1 : 0;
}

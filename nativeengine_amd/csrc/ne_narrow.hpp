// ne_narrow.hpp -- narrow phase for the gfx950 kernels (device code).
//
// Restates, in the reference's operation order, the contact routines of
//   src/detect.hpp:19-77        planeVsConvex, detectContact dispatch
//   src/narrowphase.hpp:34-155  sphere-cylinder, segment helpers, capsules
//   src/gjk.hpp:17-46, 120-123  support maps, minimum-image shift
//   src/epa.hpp:15-158          GJK tetrahedron + EPA
//   src/manifold.hpp:16-126     box-box clipping manifold
// Conventions: normal points from body b toward body a; the point is on a's
// side; depth (overlap) is positive when penetrating.
//
// GJK/EPA keep their polytope in a caller-supplied workspace.  Vertices are
// bounded by the iteration caps (<= 4 + 64); faces of a well-conditioned
// polytope by 2V-4, but near-degenerate overlaps (deeply interpenetrating
// boxes) can leave many more, as the reference's growable vectors allow.  A
// polytope that would outgrow the workspace stops with `overflow` set, never
// writing past it, and the caller reruns the pair with a larger workspace.
#pragma once

#include "ne_dmath.hpp"

namespace nea {

enum : int {
    SH_SPHERE = 0, SH_CYLINDER = 1, SH_BOX = 2, SH_CAPSULE = 3,
    SH_CONVEX = 4, SH_PLANE = 5, SH_MESH = 6, SH_COMPOUND = 7
};

enum : uint32_t {
    NE_FAULT_PAIRS = 1u,      // candidate pair buffer overflow
    NE_FAULT_CONTACTS = 2u,   // constraint buffer overflow
    NE_FAULT_EPA = 4u,        // EPA polytope exceeded its workspace
    NE_FAULT_CELLS = 8u,      // grid cell capacity exceeded
    NE_FAULT_EPA_QUEUE = 16u, // more large-workspace EPA reruns than queue slots
    NE_FAULT_EPA_BIG = 32u,   // an EPA rerun outgrew the large workspace
    NE_FAULT_WAKE = 64u,      // a chain of sleeping bodies woken one by another outran the wake rounds
    NE_FAULT_MESH = 128u,     // mesh triangle candidate buffer overflow
    NE_FAULT_WF_PLAN = 256u,  // a step had large islands where none were expected (rerun, reading them back)
};

// Geometry of one body as the narrow phase sees it.
struct Geom {
    V3 x;
    Qt q;
    V3 he;            // box half extents / plane normal
    double radius, hh;
    const double* hull;   // 3 doubles per vertex, body frame
    uint32_t nhull;
    int shape;
};

struct Contact {
    int hit;
    V3 normal, point;
    double overlap;
};

__device__ __forceinline__ Contact no_contact() {
    Contact c;
    c.hit = 0;
    c.normal = V3{0.0, 0.0, 0.0};
    c.point = V3{0.0, 0.0, 0.0};
    c.overlap = 0.0;
    return c;
}

// src/gjk.hpp:17-46
__device__ inline V3 support_world(const Geom& b, V3 d) {
    if (b.shape == SH_SPHERE) return b.x + normalized(d) * b.radius;
    V3 db = qinvrot(b.q, d);
    V3 sb;
    if (b.shape == SH_BOX) {
        sb = V3{db.x >= 0 ? b.he.x : -b.he.x, db.y >= 0 ? b.he.y : -b.he.y,
                db.z >= 0 ? b.he.z : -b.he.z};
    } else if (b.shape == SH_CAPSULE) {
        double sy = (db.y >= 0.0) ? b.hh : -b.hh;
        sb = V3{0.0, sy, 0.0} + normalized(db) * b.radius;
    } else if (b.shape == SH_CONVEX) {
        double best = -1e300;
        sb = V3{0.0, 0.0, 0.0};
        for (uint32_t k = 0; k < b.nhull; ++k) {
            V3 v{b.hull[3 * k], b.hull[3 * k + 1], b.hull[3 * k + 2]};
            double dp = dot(v, db);
            if (dp > best) { best = dp; sb = v; }
        }
    } else {
        double sy = (db.y >= 0.0) ? b.hh : -b.hh;
        double rho = sqrt(db.x * db.x + db.z * db.z);
        sb = (rho > 1e-12) ? V3{db.x / rho * b.radius, sy, db.z / rho * b.radius}
                           : V3{0.0, sy, 0.0};
    }
    return b.x + qrot(b.q, sb);
}

// src/gjk.hpp:120-123
__device__ __forceinline__ V3 min_image_shift(const Box3& box, V3 ax, V3 bx) {
    V3 raw = bx - ax;
    return min_image(box, raw) - raw;
}

struct SupV {
    V3 v, a, b;
};

__device__ __forceinline__ SupV cso_support(const Geom& A, const Geom& B, V3 shift, V3 d) {
    V3 pa = support_world(A, d);
    V3 pb = support_world(B, -d) + shift;
    return SupV{pa - pb, pa, pb};
}

// A Minkowski point with the direction it was found along instead of its two
// witness points: the witnesses are the same support calls again, so a
// polytope of these needs them only for its final face.
struct SupD {
    V3 v, d;
};

__device__ __forceinline__ void make_sup(const Geom& A, const Geom& B, V3 shift, V3 d, SupV& out) {
    out = cso_support(A, B, shift, d);
}
__device__ __forceinline__ void make_sup(const Geom& A, const Geom& B, V3 shift, V3 d, SupD& out) {
    out.v = support_world(A, d) - (support_world(B, -d) + shift);
    out.d = d;
}
// The witness points of a SupD: exactly cso_support's for its direction.
__device__ __forceinline__ SupV witness(const Geom& A, const Geom& B, V3 shift, V3 v, V3 d) {
    SupV s;
    s.a = support_world(A, d);
    s.b = support_world(B, -d) + shift;
    s.v = v;
    return s;
}

// src/epa.hpp:28-64 -- GJK evolving simplex; on overlap returns an
// origin-enclosing tetrahedron (S = SupV with witness points, or SupD).
template <class S>
__device__ inline bool gjk_tetra(const Geom& A, const Geom& B, V3 shift, S tet[4]) {
    V3 d{1.0, 0.0, 0.0};
    S s;
    make_sup(A, B, shift, d, s);
    if (norm2(s.v) < 1e-24) { d = V3{0.0, 1.0, 0.0}; make_sup(A, B, shift, d, s); }
    S W[4];
    int n = 1;
    W[0] = s;
    d = -s.v;
    for (int iter = 0; iter < 64; ++iter) {
        if (norm2(d) < 1e-24) d = V3{1.0, 0.0, 0.0};
        S p;
        make_sup(A, B, shift, d, p);
        if (dot(p.v, d) <= 1e-12) return false;
        // shift the simplex (n <= 3 here) with fixed indices so it stays in
        // registers; slots past n hold stale points that are never read
        W[3] = W[2];
        W[2] = W[1];
        W[1] = W[0];
        W[0] = p;
        ++n;
        V3 a = W[0].v, ao = -a;
        if (n == 2) {
            V3 ab = W[1].v - a;
            d = cross(cross(ab, ao), ab);
            if (norm2(d) < 1e-24) d = V3{ab.y, -ab.x, 0.0};
        } else if (n == 3) {
            V3 ab = W[1].v - a, ac = W[2].v - a;
            V3 abc = cross(ab, ac);
            if (dot(abc, ao) > 0) d = abc;
            else { S t = W[1]; W[1] = W[2]; W[2] = t; d = -abc; }
            if (norm2(d) < 1e-24) return false;
        } else {
            V3 b = W[1].v, c = W[2].v, dd = W[3].v;
            V3 ab = b - a, ac = c - a, ad = dd - a;
            V3 abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
            if (dot(abc, ao) > 0) { n = 3; d = abc; }
            else if (dot(acd, ao) > 0) { S w2 = W[2], w3 = W[3]; W[1] = w2; W[2] = w3; n = 3; d = acd; }
            else if (dot(adb, ao) > 0) { S w1 = W[1], w3 = W[3]; W[1] = w3; W[2] = w1; n = 3; d = adb; }
            else { tet[0] = W[0]; tet[1] = W[1]; tet[2] = W[2]; tet[3] = W[3]; return true; }
        }
    }
    return false;
}

// EPA workspace bounds: <= 4 + 64 vertices; a convex polytope on V vertices
// has <= 2V-4 faces.  Faces/horizon carry headroom for near-degenerate hulls.
constexpr int EPA_MAXV = 68;    // 4 + 64 iterations: every vertex an EPA run can make
constexpr int EPA_MAXF = 160;
constexpr int EPA_MAXH = 96;
struct Face {
    V3 n;
    double dist;
    int a, b, c;
};

// Horizon of one EPA iteration: the edges of the faces that see the new
// point, where an edge whose reverse is already listed removes the first such
// entry instead of being added (the reference's find + erase on a vector).
// The surviving edges, in insertion order, become the new faces.
//
// HorizonList keeps the list compact and searches it linearly (small
// polytopes: per-thread arrays, or one wave with the polytope in LDS).
struct HorizonList {
    int (*hor)[2];
    int cap;
    int n;
    __device__ void begin() { n = 0; }
    __device__ bool add(int u, int w) {
        for (int e = 0; e < n; ++e) {
            if (hor[e][0] == w && hor[e][1] == u) {
                for (int m = e; m + 1 < n; ++m) { hor[m][0] = hor[m + 1][0]; hor[m][1] = hor[m + 1][1]; }
                --n;
                return true;
            }
        }
        if (n >= cap) return false;
        hor[n][0] = u; hor[n][1] = w; ++n;
        return true;
    }
    __device__ int alive() const { return n; }
    __device__ int entries() const { return n; }
    __device__ bool get(int e, int& u, int& w) const { u = hor[e][0]; w = hor[e][1]; return true; }
};

__device__ __forceinline__ Face make_face(const SupV* verts, int i, int j, int k) {
    Face f;
    V3 n = cross(verts[j].v - verts[i].v, verts[k].v - verts[i].v);
    double L = norm(n);
    n = (L > 1e-18) ? n / L : V3{0.0, 0.0, 1.0};
    double dist = dot(n, verts[i].v);
    if (dist < 0) { n = -n; dist = -dist; int t = j; j = k; k = t; }
    f.n = n; f.dist = dist; f.a = i; f.b = j; f.c = k;
    return f;
}

// Barycentric witness points of the closest face (src/epa.hpp:123-149).
__device__ inline Contact epa_result(const SupV* verts, const Face& best) {
    const SupV A0 = verts[best.a], B0 = verts[best.b], C0 = verts[best.c];
    V3 proj = best.n * best.dist;
    V3 v0 = B0.v - A0.v, v1 = C0.v - A0.v, v2 = proj - A0.v;
    double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    double d20 = dot(v2, v0), d21 = dot(v2, v1);
    double den = d00 * d11 - d01 * d01;
    double bv = 0, bw = 0;
    if (fabs(den) > 1e-18) {
        bv = (d11 * d20 - d01 * d21) / den;
        bw = (d00 * d21 - d01 * d20) / den;
    }
    double bu = 1.0 - bv - bw;
    V3 onA = (A0.a * bu + B0.a * bv) + C0.a * bw;
    V3 onB = (A0.b * bu + B0.b * bv) + C0.b * bw;
    Contact c;
    c.hit = 1;
    c.normal = -best.n;
    c.overlap = best.dist;
    c.point = (onA + onB) * 0.5;
    return c;
}


// src/epa.hpp:71-150.  Faces are compacted in place (kept faces in their
// original order, then the new horizon faces), exactly the order the
// reference's `keep` + push_back produces.  `verts` holds EPA_MAXV entries,
// `faces` maxf; exceeding maxf or the horizon's capacity sets `overflow` and
// stops (the result is then not used).  (The overlap queries' per-thread
// tier; the step's EPA tiers are in ne_epa_lds.hpp and ne_gpu.hip.)
template <class Hz>
__device__ inline Contact epa_contact(const Geom& A, const Geom& B, V3 shift, const SupV tet[4], SupV* verts,
                                      Face* faces, int maxf, int maxv, int max_iter, Hz& hz, bool& overflow,
                                      bool& too_long) {
    int nv = 4;
    for (int k = 0; k < 4; ++k) verts[k] = tet[k];
    int nf = 4;
    faces[0] = make_face(verts, 0, 1, 2);
    faces[1] = make_face(verts, 0, 2, 3);
    faces[2] = make_face(verts, 0, 3, 1);
    faces[3] = make_face(verts, 1, 3, 2);
    Face best = faces[0];
    for (int iter = 0; iter < 64; ++iter) {
        if (iter == max_iter) {   // not converged within this launch's budget: rerun later from the start
            too_long = true;
            break;
        }
        int ci = 0;
        for (int i = 1; i < nf; ++i)
            if (faces[i].dist < faces[ci].dist) ci = i;
        best = faces[ci];
        SupV p = cso_support(A, B, shift, best.n);
        double d = dot(p.v, best.n);
        if (d - best.dist < 1e-7) break;
        if (nv == maxv) {   // the workspace holds no more vertices
            overflow = true;
            break;
        }
        int pidx = nv;
        verts[nv++] = p;   // nv <= 4 + 64 == EPA_MAXV
        hz.begin();
        int nk = 0;
        bool over = false;
        for (int fi = 0; fi < nf && !over; ++fi) {
            Face f = faces[fi];
            if (dot(f.n, p.v - verts[f.a].v) > 1e-12) {
                over = !hz.add(f.a, f.b) || !hz.add(f.b, f.c) || !hz.add(f.c, f.a);
            } else {
                if (nk != fi) faces[nk] = f;   // nk <= fi: in-place compaction keeps order
                ++nk;
            }
        }
        if (over || nk + hz.alive() > maxf) {
            overflow = true;
            break;
        }
        nf = nk;
        const int ne = hz.entries();
        for (int e = 0; e < ne; ++e) {
            int u, w;
            if (hz.get(e, u, w)) faces[nf++] = make_face(verts, u, w, pidx);
        }
        if (nf == 0) break;
    }
    return epa_result(verts, best);
}

// GJK tetrahedra of the overlapping convex pairs, by pair index: 24 doubles
// each (v and d of four vertices), contiguous, so the EPA tiers -- which
// take the pairs in queue order, scattered over the pair list -- read one
// and a half cache lines a pair.
struct EpaTets {
    double* d;
    __device__ void put(uint32_t pair, const SupD tet[4]) const {
        double2* f = reinterpret_cast<double2*>(d + (size_t)pair * 24);
        for (int k = 0; k < 4; ++k) {
            f[3 * k] = double2{tet[k].v.x, tet[k].v.y};
            f[3 * k + 1] = double2{tet[k].v.z, tet[k].d.x};
            f[3 * k + 2] = double2{tet[k].d.y, tet[k].d.z};
        }
    }
    __device__ void get(uint32_t pair, SupD tet[4]) const {
        const double2* f = reinterpret_cast<const double2*>(d + (size_t)pair * 24);
        for (int k = 0; k < 4; ++k) {
            const double2 a = f[3 * k], b = f[3 * k + 1], c = f[3 * k + 2];
            tet[k].v = V3{a.x, a.y, b.x};
            tet[k].d = V3{b.y, c.x, c.y};
        }
    }
};

// What the detect path needs for convex pairs: it stops after GJK (EPA runs
// later, in the EPA tiers), and an overlapping pair reports hit == 2 and
// leaves its tetrahedron in `tets` at index `pair` (when tets.d is set).
struct EpaWS {
    EpaTets tets;
    uint32_t pair;
};

// src/epa.hpp:153-158 up to EPA: the minimum-image shift and GJK.
__device__ inline Contact convex_contact(const Box3& box, const Geom& A, const Geom& B, EpaWS& ws) {
    const V3 shift = min_image_shift(box, A.x, B.x);
    SupD tet[4];
    if (!gjk_tetra(A, B, shift, tet)) return no_contact();
    if (ws.tets.d) ws.tets.put(ws.pair, tet);
    Contact c = no_contact();
    c.hit = 2;
    return c;
}

// src/narrowphase.hpp:34-81
__device__ inline Contact sphere_vs_cylinder(const Box3& box, const Geom& sph, const Geom& cyl) {
    Contact c = no_contact();
    const double r = cyl.radius, h = cyl.hh, sr = sph.radius;
    V3 rel = min_image(box, sph.x - cyl.x);
    V3 L = qinvrot(cyl.q, rel);
    double rho = sqrt(L.x * L.x + L.z * L.z);
    bool insideAxial = fabs(L.y) < h;
    bool insideRadial = rho < r;
    V3 Pl, nl;
    if (insideAxial && insideRadial) {
        double dCap = h - fabs(L.y);
        double dSide = r - rho;
        if (dSide < dCap) {
            double s = (rho > 1e-12) ? r / rho : 0.0;
            Pl = V3{L.x * s, L.y, L.z * s};
            nl = (rho > 1e-12) ? V3{L.x / rho, 0.0, L.z / rho} : V3{1.0, 0.0, 0.0};
            c.overlap = sr + dSide;
        } else {
            double sy = (L.y >= 0) ? h : -h;
            Pl = V3{L.x, sy, L.z};
            nl = V3{0.0, (L.y >= 0) ? 1.0 : -1.0, 0.0};
            c.overlap = sr + dCap;
        }
        c.hit = 1;
    } else {
        double yc = (L.y > h) ? h : (L.y < -h ? -h : L.y);
        double rx = L.x, rz = L.z;
        if (rho > r) { double s = r / rho; rx *= s; rz *= s; }
        Pl = V3{rx, yc, rz};
        V3 diff = L - Pl;
        double dist = norm(diff);
        if (dist >= sr) return c;
        c.hit = 1;
        c.overlap = sr - dist;
        nl = (dist > 1e-12) ? diff / dist : V3{0.0, 1.0, 0.0};
    }
    c.normal = normalized(qrot(cyl.q, nl));
    c.point = cyl.x + qrot(cyl.q, Pl);
    return c;
}

__device__ __forceinline__ double np_clamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

// src/narrowphase.hpp:88-93
__device__ __forceinline__ V3 closest_on_segment(V3 p, V3 a, V3 b) {
    V3 ab = b - a;
    double L2 = norm2(ab);
    if (L2 < 1e-18) return a;
    return a + ab * np_clamp(dot(p - a, ab) / L2, 0.0, 1.0);
}

// src/narrowphase.hpp:96-116 (Ericson)
__device__ inline void closest_seg_seg(V3 p1, V3 q1, V3 p2, V3 q2, V3& c1, V3& c2) {
    V3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    const double EPS = 1e-12;
    double s, t;
    if (a <= EPS && e <= EPS) { s = t = 0.0; }
    else if (a <= EPS) { s = 0.0; t = np_clamp(f / e, 0.0, 1.0); }
    else {
        double c = dot(d1, r);
        if (e <= EPS) { t = 0.0; s = np_clamp(-c / a, 0.0, 1.0); }
        else {
            double b = dot(d1, d2), denom = a * e - b * b;
            s = denom > EPS ? np_clamp((b * f - c * e) / denom, 0.0, 1.0) : 0.0;
            t = (b * s + f) / e;
            if (t < 0.0) { t = 0.0; s = np_clamp(-c / a, 0.0, 1.0); }
            else if (t > 1.0) { t = 1.0; s = np_clamp((b - c) / a, 0.0, 1.0); }
        }
    }
    c1 = p1 + d1 * s;
    c2 = p2 + d2 * t;
}

// src/narrowphase.hpp:120-135
__device__ inline Contact capsule_vs_sphere(const Box3& box, const Geom& cap, const Geom& sph) {
    Contact c = no_contact();
    V3 sc = cap.x + min_image(box, sph.x - cap.x);
    V3 axis = qrot(cap.q, V3{0.0, 1.0, 0.0});
    V3 cp = closest_on_segment(sc, cap.x - axis * cap.hh, cap.x + axis * cap.hh);
    V3 e = sc - cp;
    double dist = norm(e);
    double sumR = cap.radius + sph.radius;
    if (dist >= sumR) return c;
    V3 nrm = (dist > 1e-12) ? e / dist : V3{0.0, 1.0, 0.0};
    c.hit = 1;
    c.overlap = sumR - dist;
    c.normal = -nrm;
    c.point = cp + nrm * cap.radius;
    return c;
}

// src/narrowphase.hpp:138-155
__device__ inline Contact capsule_vs_capsule(const Box3& box, const Geom& A, const Geom& B) {
    Contact c = no_contact();
    V3 axA = qrot(A.q, V3{0.0, 1.0, 0.0}), axB = qrot(B.q, V3{0.0, 1.0, 0.0});
    V3 bc = A.x + min_image(box, B.x - A.x);
    V3 pA, pB;
    closest_seg_seg(A.x - axA * A.hh, A.x + axA * A.hh, bc - axB * B.hh, bc + axB * B.hh, pA, pB);
    V3 e = pB - pA;
    double dist = norm(e);
    double sumR = A.radius + B.radius;
    if (dist >= sumR) return c;
    V3 nrm = (dist > 1e-12) ? e / dist : V3{0.0, 1.0, 0.0};
    c.hit = 1;
    c.overlap = sumR - dist;
    c.normal = -nrm;
    c.point = pA + nrm * A.radius;
    return c;
}

// src/detect.hpp:19-31
__device__ inline Contact plane_vs_convex(const Box3& box, const Geom& other, const Geom& plane) {
    Contact c = no_contact();
    V3 n = plane.he;
    V3 sp = support_world(other, -n);
    V3 rel = min_image(box, sp - plane.x);
    double pen = -dot(n, rel);
    if (pen <= 0.0) return c;
    c.hit = 1;
    c.overlap = pen;
    c.normal = n;
    c.point = sp + n * (0.5 * pen);
    return c;
}

// src/detect.hpp:44-54
__device__ __forceinline__ Contact sphere_sphere(const Box3& box, V3 ax, double ra, V3 bx, double rb) {
    Contact c = no_contact();
    V3 d = min_image(box, ax - bx);
    double sumR = ra + rb, dist2 = norm2(d);
    if (dist2 >= sumR * sumR || dist2 < 1e-18) return c;
    double dist = sqrt(dist2);
    c.hit = 1;
    c.normal = d / dist;
    c.overlap = sumR - dist;
    c.point = ax - c.normal * ra;
    return c;
}

// src/detect.hpp:33-77
__device__ inline Contact detect_contact(const Box3& box, const Geom& a, const Geom& b, EpaWS& ws) {
    if (a.shape == SH_PLANE || b.shape == SH_PLANE) {
        if (a.shape == SH_PLANE && b.shape == SH_PLANE) return no_contact();
        const Geom& plane = (a.shape == SH_PLANE) ? a : b;
        const Geom& other = (a.shape == SH_PLANE) ? b : a;
        Contact t = plane_vs_convex(box, other, plane);
        if (t.hit && a.shape == SH_PLANE) t.normal = -t.normal;
        return t;
    }
    if (a.shape == SH_SPHERE && b.shape == SH_SPHERE) return sphere_sphere(box, a.x, a.radius, b.x, b.radius);
    const bool aCyl = a.shape == SH_CYLINDER, bCyl = b.shape == SH_CYLINDER;
    if (a.shape == SH_SPHERE && bCyl) return sphere_vs_cylinder(box, a, b);
    if (b.shape == SH_SPHERE && aCyl) {
        Contact t = sphere_vs_cylinder(box, b, a);
        if (t.hit) t.normal = -t.normal;
        return t;
    }
    const bool aCap = a.shape == SH_CAPSULE, bCap = b.shape == SH_CAPSULE;
    if (aCap && bCap) return capsule_vs_capsule(box, a, b);
    if (aCap && b.shape == SH_SPHERE) return capsule_vs_sphere(box, a, b);
    if (a.shape == SH_SPHERE && bCap) {
        Contact t = capsule_vs_sphere(box, b, a);
        if (t.hit) t.normal = -t.normal;
        return t;
    }
    return convex_contact(box, a, b, ws);
}

// ---- box-box manifold: src/manifold.hpp:16-126 ----------------------------

struct BFace {
    V3 n, c, u, v;
    double eu, ev;
    V3 vert[4];
};

__device__ __forceinline__ V3 axis_vec(int a) {
    return a == 0 ? V3{1.0, 0.0, 0.0} : (a == 1 ? V3{0.0, 1.0, 0.0} : V3{0.0, 0.0, 1.0});
}

__device__ inline BFace box_face(const Geom& b, V3 centre, int axis, int sign) {
    const int t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
    const double h[3] = {b.he.x, b.he.y, b.he.z};
    BFace f;
    f.n = qrot(b.q, axis_vec(axis) * (double)sign);
    f.c = centre + f.n * h[axis];
    f.u = qrot(b.q, axis_vec(t1)); f.eu = h[t1];
    f.v = qrot(b.q, axis_vec(t2)); f.ev = h[t2];
    f.vert[0] = (f.c + f.u * f.eu) + f.v * f.ev;
    f.vert[1] = (f.c - f.u * f.eu) + f.v * f.ev;
    f.vert[2] = (f.c - f.u * f.eu) - f.v * f.ev;
    f.vert[3] = (f.c + f.u * f.eu) - f.v * f.ev;
    return f;
}

__device__ inline BFace best_face(const Geom& b, V3 centre, V3 dir) {
    int bestAxis = 0, bestSign = 1;
    double best = -1e30;
    for (int axis = 0; axis < 3; ++axis)
        for (int sign = -1; sign <= 1; sign += 2) {
            double a = dot(qrot(b.q, axis_vec(axis) * (double)sign), dir);
            if (a > best) { best = a; bestAxis = axis; bestSign = sign; }
        }
    return box_face(b, centre, bestAxis, bestSign);
}

// Sutherland-Hodgman against one plane.  A convex input grows by at most one
// vertex per plane; `cap` guards rounding-made sign flips (fault, no overrun).
constexpr int CLIP_CAP = 16;
__device__ __forceinline__ int clip_plane(const V3* poly, int m, V3 pn, double off, V3* out,
                                          uint32_t* fault) {
    int k = 0;
    for (int i = 0; i < m; ++i) {
        V3 cur = poly[i];
        V3 nxt = poly[(i + 1) % m];
        double dc = dot(cur, pn) - off, dn = dot(nxt, pn) - off;
        bool inC = dc <= 0.0, inN = dn <= 0.0;
        if (inC) {
            if (k >= CLIP_CAP) { atomicOr(fault, NE_FAULT_EPA); return k; }
            out[k++] = cur;
        }
        if (inC != inN) {
            double t = dc / (dc - dn);
            if (k >= CLIP_CAP) { atomicOr(fault, NE_FAULT_EPA); return k; }
            out[k++] = cur + (nxt - cur) * t;
        }
    }
    return k;
}

__device__ inline int box_box_manifold(const Box3& box, const Geom& a, const Geom& b, V3 n,
                                       V3 pts[4], double depth[4], uint32_t* fault) {
    int count = 0;
    const V3 bShift = min_image_shift(box, a.x, b.x);
    const V3 bC = b.x + bShift;
    BFace fa = best_face(a, a.x, -n);
    BFace fb = best_face(b, bC, n);
    double alignA = dot(fa.n, -n), alignB = dot(fb.n, n);
    const bool refA = alignA >= alignB;
    const BFace& ref = refA ? fa : fb;
    BFace inc = refA ? best_face(b, bC, -ref.n) : best_face(a, a.x, -ref.n);
    V3 p0[CLIP_CAP], p1[CLIP_CAP];
    int m = 4;
    for (int k = 0; k < 4; ++k) p0[k] = inc.vert[k];
    m = clip_plane(p0, m, ref.u, dot(ref.c, ref.u) + ref.eu, p1, fault);
    m = clip_plane(p1, m, ref.u * -1.0, -(dot(ref.c, ref.u) - ref.eu), p0, fault);
    m = clip_plane(p0, m, ref.v, dot(ref.c, ref.v) + ref.ev, p1, fault);
    m = clip_plane(p1, m, ref.v * -1.0, -(dot(ref.c, ref.v) - ref.ev), p0, fault);
    const double slop = 0.01;
    for (int k = 0; k < m; ++k) {
        double sd = dot(p0[k] - ref.c, ref.n);
        if (sd > slop) continue;
        if (count >= 4) break;
        pts[count] = p0[k] - ref.n * sd;
        depth[count] = -sd;
        ++count;
    }
    return count;
}

}  // namespace nea

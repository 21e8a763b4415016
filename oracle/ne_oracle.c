/*
 * ne_oracle.c -- CPU restatement of NativeEngine's per-step hot path.
 *
 * ORACLE / TEST INFRASTRUCTURE ONLY (see ne_oracle.h).  Single-threaded C99.
 *
 * Every function cites the reference file:line whose semantics and operation
 * order it restates.  Build with -ffp-contract=off (oracle/Makefile): the
 * point of this file is to produce the same IEEE double bit patterns as the
 * reference.
 *
 * Pinning (tests/test_oracle.py):
 *   - against oracle/_ref (the unmodified reference): bitwise state, contact
 *     list and hit-pair list on gas, pile and narrow-phase scenes;
 *   - against the committed golden fixtures in tests/golden/ (made from
 *     oracle/_ref) when oracle/_ref is absent;
 *   - against closed forms: min-image/wrap KATs, sphere-cylinder KATs, the
 *     2-body elastic exchange, free-spin invariants, gravity, the Coulomb cone.
 */
#include "ne_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Math: src/vmath.hpp:18-132                                                */
/* ------------------------------------------------------------------------ */

typedef struct { double x, y, z; } v3;
typedef struct { double w, x, y, z; } qt;

static v3 mk(double x, double y, double z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }          /* vmath.hpp:24 */
static v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }          /* vmath.hpp:25 */
static v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }                               /* vmath.hpp:26 */
static v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }            /* vmath.hpp:27 */
static v3 dvs(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }            /* vmath.hpp:28 */
static double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }         /* vmath.hpp:32 */
static v3 cross(v3 a, v3 o) {                                                      /* vmath.hpp:33-35 */
    return mk(a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x);
}
static double norm2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }            /* vmath.hpp:36 */
static double norm(v3 a) { return sqrt(norm2(a)); }                                /* vmath.hpp:37 */
static v3 normalized(v3 a) {                                                       /* vmath.hpp:38-41 */
    double n = norm(a);
    return (n > 1e-300) ? mk(a.x / n, a.y / n, a.z / n) : mk(0, 0, 0);
}
static double comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

static qt mq(double w, double x, double y, double z) { qt r; r.w = w; r.x = x; r.y = y; r.z = z; return r; }
static qt qmul(qt a, qt o) {                                                       /* vmath.hpp:53-59 */
    return mq(a.w * o.w - a.x * o.x - a.y * o.y - a.z * o.z,
              a.w * o.x + a.x * o.w + a.y * o.z - a.z * o.y,
              a.w * o.y - a.x * o.z + a.y * o.w + a.z * o.x,
              a.w * o.z + a.x * o.y - a.y * o.x + a.z * o.w);
}
static qt qadd(qt a, qt o) { return mq(a.w + o.w, a.x + o.x, a.y + o.y, a.z + o.z); } /* vmath.hpp:60 */
static qt qscl(qt a, double s) { return mq(a.w * s, a.x * s, a.y * s, a.z * s); }     /* vmath.hpp:61 */
static qt qnormalized(qt a) {                                                         /* vmath.hpp:63-67 */
    double n = sqrt(a.w * a.w + a.x * a.x + a.y * a.y + a.z * a.z);
    return (n > 1e-300) ? mq(a.w / n, a.x / n, a.y / n, a.z / n) : mq(1, 0, 0, 0);
}
static qt qconj(qt a) { return mq(a.w, -a.x, -a.y, -a.z); }                           /* vmath.hpp:68 */
static v3 qrot(qt q, v3 v) {                                                          /* vmath.hpp:71-76 */
    v3 qv = mk(q.x, q.y, q.z);
    v3 t = scl(cross(qv, v), 2.0);
    return add(add(v, scl(t, q.w)), cross(qv, t));
}
static v3 qinvrot(qt q, v3 v) { return qrot(qconj(q), v); }                           /* vmath.hpp:77 */
static qt integrate_orientation(qt q, v3 om, double dt) {                            /* vmath.hpp:128-132 */
    qt oq = mq(0.0, om.x, om.y, om.z);
    qt dq = qscl(qmul(oq, q), 0.5 * dt);
    return qnormalized(qadd(q, dq));
}

/* ------------------------------------------------------------------------ */
/* Box: src/box.hpp:16-42 (per-axis extension; cubic == reference)          */
/* ------------------------------------------------------------------------ */

typedef struct { double half[3]; int periodic; } obox;

static double mi1(double d, double L) { return d - L * floor(d / L + 0.5); }       /* box.hpp:27-29 */
static v3 min_image(const obox* b, v3 d) {                                          /* box.hpp:24-31 */
    if (!b->periodic) return d;
    d.x = mi1(d.x, 2.0 * b->half[0]);
    d.y = mi1(d.y, 2.0 * b->half[1]);
    d.z = mi1(d.z, 2.0 * b->half[2]);
    return d;
}
static double wr1(double p, double h, double L) { return p - L * floor((p + h) / L); } /* box.hpp:37-39 */
static v3 wrap_pos(const obox* b, v3 p) {                                           /* box.hpp:34-41 */
    if (!b->periodic) return p;
    p.x = wr1(p.x, b->half[0], 2.0 * b->half[0]);
    p.y = wr1(p.y, b->half[1], 2.0 * b->half[1]);
    p.z = wr1(p.z, b->half[2], 2.0 * b->half[2]);
    return p;
}

/* ------------------------------------------------------------------------ */
/* Body: src/body.hpp:41-203                                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    v3 x, v, w, force, torque, invI, he, invIStore;
    qt q;
    double invMass, radius, hh, friction, restitution, linDamp, angDamp, gravityScale;
    double sleepTimer, invMassStore;
    int shape;
    uint32_t layer, mask, flags;
    uint32_t hull0, hull1;  /* [hull0, hull1) into world hull vertex array */
    uint32_t nc0, nc1;      /* [nc0, nc1) into world no-collide array      */
    int32_t mesh;           /* mesh body: index into the world's meshes, or -1 */
    const v3* hvp;          /* when set, the hull vertices (a temporary prism) */
    v3 prevX;               /* the pose at the step's start (world.cpp:135), for CCD */
} obody;

/* MeshData (src/mesh.hpp:14-93): vertices, triangles, the tree's triangle
 * order and its nodes (lo, hi; children or a leaf's [start, start+count)). */
typedef struct { v3 lo, hi; int32_t left, right, start, count; } onode;
typedef struct {
    uint32_t nv, nt, nn;
    v3* verts;
    int32_t* tris;
    int32_t* order;
    onode* nodes;
} omesh;

/* the tree build (ne_oracle_bvh.cpp, through std::nth_element) */
int32_t ne_oracle_bvh_build(const double* verts, const int32_t* tris, int32_t ntris, int32_t* order,
                            int32_t node_cap, double* bounds, int32_t* links);

typedef struct { int hit; v3 normal, point; double overlap; } ocontact;   /* narrowphase.hpp:16-21 */

struct ne_oracle {
    ne_world_params P;
    obox box;
    uint32_t n;
    obody* B;
    v3* hull;
    uint32_t nhull;
    uint32_t* nocol;
    uint32_t nnocol;
    /* last step's constraint list */
    uint32_t ncons, capcons;
    struct ocons* cons;
    /* warm cache = previous step's (key, slot, nImp), sorted by key then slot */
    uint32_t ncache;
    uint64_t* cache_key;
    double* cache_imp;
    uint32_t ncand;
    /* contact-set tape: mode 1 records each step's filtered contact lists,
     * mode 2 replays them in place of broadphase + narrow phase */
    int tape_mode;
    uint32_t tape_steps, tape_cap, tape_cursor;
    struct otape_step* tape;
    /* distance joints, in order (world.hpp:129) */
    uint32_t njoints;
    ne_distance_joint* joints;
    /* this step's dynamic-dynamic contact pairs, for island sleeping (world.hpp:158) */
    uint32_t nisp, capisp;
    uint32_t* isp;
    /* static triangle meshes (Body::mesh) */
    uint32_t nmeshes;
    omesh* meshes;
    /* one pair's contacts (gatherContacts' output vector) */
    uint32_t capcs;
    ocontact* cs;
    /* facade contact events: this step's and the last step's overlapping
     * pairs (physics_world.cpp:253-302), each in (i, j) order */
    int events_on;
    uint32_t nccd;   /* bodies with the CCD flag */
    /* general joints (world.hpp:130), with their runtime accumulators */
    uint32_t ngj;
    ne_joint* gj;
    double* gj_applied;
    double* gj_motor;
    uint32_t nev[2], capev[2];
    struct oevent* ev[2];
    int ev_cur;
};

struct oevent {
    uint32_t i, j;
    int trigger;   /* a sensor is in the pair: curTriggers, else curContacts */
    ocontact c;
};

struct otape_pair {
    uint32_t i, j;
    int nc;
    ocontact cs[4];
};

/* Branch bits of one Gauss-Seidel visit (world.cpp:572-612), recorded in
 * mode 1 and replayed in mode 3. */
enum {
    TB_KN = 1,        /* Kn > 1e-18: the normal impulse is applied            */
    TB_UNCLAMPED = 2, /* old + dj > 0: the accumulated normal impulse is kept  */
    TB_FRIC = 4,      /* mu > 0 and nImp > 0: friction is visited             */
    TB_VT = 8,        /* |vt| > 1e-12                                          */
    TB_KT = 16,       /* Kt > 1e-18                                            */
    TB_LO = 32,       /* the friction impulse clamps at -mu nImp               */
    TB_HI = 64        /* ... at +mu nImp                                       */
};

struct otape_step {
    size_t n, cap;
    struct otape_pair* p;
    /* per constraint: the restitution branch (vn < -0.5, world.cpp:544);
     * per constraint and iteration: TB_* bits */
    uint32_t ncons, iters;
    uint8_t* rest;
    uint8_t* visit;
};

static v3 ld3(const double* p, uint32_t i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

static int is_sphere(const obody* b) { return b->shape == NE_SHAPE_SPHERE; }
static int grid_excluded(const obody* b) {                                          /* body.hpp:134 */
    return b->shape == NE_SHAPE_PLANE || b->shape == NE_SHAPE_MESH;
}

static double bounding_radius(const ne_oracle* o, const obody* b) {                 /* body.hpp:137-158 */
    if (b->shape == NE_SHAPE_SPHERE) return b->radius;
    if (b->shape == NE_SHAPE_BOX) return norm(b->he);
    if (b->shape == NE_SHAPE_CAPSULE) return b->hh + b->radius;
    if (b->shape == NE_SHAPE_PLANE || b->shape == NE_SHAPE_MESH) return 1e30;
    if (b->shape == NE_SHAPE_CONVEX) {
        double m = 0.0;
        for (uint32_t k = b->hull0; k < b->hull1; ++k) {
            double r = norm(o->hull[k]);
            m = (m < r) ? r : m;   /* std::max(m, r) */
        }
        return m;
    }
    return sqrt(b->hh * b->hh + b->radius * b->radius);
}

static v3 apply_inv_inertia(const obody* b, v3 u) {                                 /* body.hpp:162-168 */
    v3 ub = qinvrot(b->q, u);
    v3 sb = mk(ub.x * b->invI.x, ub.y * b->invI.y, ub.z * b->invI.z);
    return qrot(b->q, sb);
}

static v3 ibody_of(const obody* b) {
    return mk(b->invI.x > 0 ? 1.0 / b->invI.x : 0.0,
              b->invI.y > 0 ? 1.0 / b->invI.y : 0.0,
              b->invI.z > 0 ? 1.0 / b->invI.z : 0.0);
}

static v3 angular_momentum(const obody* b) {                                        /* body.hpp:195-202 */
    v3 Ib = ibody_of(b);
    v3 wb = qinvrot(b->q, b->w);
    v3 Lb = mk(wb.x * Ib.x, wb.y * Ib.y, wb.z * Ib.z);
    return qrot(b->q, Lb);
}

static double kinetic(const obody* b) {                                             /* body.hpp:170-180 */
    double lin = 0.5 * (b->invMass > 0 ? 1.0 / b->invMass : 0.0) * norm2(b->v);
    v3 Ib = ibody_of(b);
    v3 wb = qinvrot(b->q, b->w);
    v3 Lb = mk(wb.x * Ib.x, wb.y * Ib.y, wb.z * Ib.z);
    double rot = 0.5 * dot(wb, Lb);
    return lin + rot;
}

/* ------------------------------------------------------------------------ */
/* Support maps: src/gjk.hpp:17-46, 120-123; src/epa.hpp:15-25               */
/* ------------------------------------------------------------------------ */

static v3 support_world(const ne_oracle* o, const obody* b, v3 d) {                 /* gjk.hpp:17-46 */
    if (is_sphere(b)) return add(b->x, scl(normalized(d), b->radius));
    v3 db = qinvrot(b->q, d);
    v3 sb;
    if (b->shape == NE_SHAPE_BOX) {
        sb = mk(db.x >= 0 ? b->he.x : -b->he.x, db.y >= 0 ? b->he.y : -b->he.y,
                db.z >= 0 ? b->he.z : -b->he.z);
    } else if (b->shape == NE_SHAPE_CAPSULE) {
        double sy = (db.y >= 0.0) ? b->hh : -b->hh;
        sb = add(mk(0.0, sy, 0.0), scl(normalized(db), b->radius));
    } else if (b->shape == NE_SHAPE_CONVEX) {
        double best = -1e300;
        sb = mk(0, 0, 0);
        const v3* H = b->hvp ? b->hvp : o->hull;
        for (uint32_t k = b->hull0; k < b->hull1; ++k) {
            double dp = dot(H[k], db);
            if (dp > best) { best = dp; sb = H[k]; }
        }
    } else {
        double sy = (db.y >= 0.0) ? b->hh : -b->hh;
        double rho = sqrt(db.x * db.x + db.z * db.z);
        sb = (rho > 1e-12) ? mk(db.x / rho * b->radius, sy, db.z / rho * b->radius) : mk(0.0, sy, 0.0);
    }
    return add(b->x, qrot(b->q, sb));
}

static v3 min_image_shift(const obox* box, const obody* A, const obody* B) {       /* gjk.hpp:120-123 */
    v3 raw = sub(B->x, A->x);
    return sub(min_image(box, raw), raw);
}

typedef struct { v3 v, a, b; } supv;                                                /* epa.hpp:15-19 */

static supv cso_support(const ne_oracle* o, const obody* A, const obody* B, v3 shift, v3 d) { /* epa.hpp:21-25 */
    supv s;
    v3 pa = support_world(o, A, d);
    v3 pb = add(support_world(o, B, neg(d)), shift);
    s.v = sub(pa, pb); s.a = pa; s.b = pb;
    return s;
}

/* GJK returning an origin-enclosing tetrahedron: src/epa.hpp:28-64 */
static int gjk_tetra(const ne_oracle* o, const obody* A, const obody* B, v3 shift, supv tet[4]) {
    v3 d = mk(1, 0, 0);
    supv s = cso_support(o, A, B, shift, d);
    if (norm2(s.v) < 1e-24) { d = mk(0, 1, 0); s = cso_support(o, A, B, shift, d); }
    supv W[4];
    int n = 1;
    W[0] = s;
    d = neg(s.v);
    for (int iter = 0; iter < 64; ++iter) {
        if (norm2(d) < 1e-24) d = mk(1, 0, 0);
        supv p = cso_support(o, A, B, shift, d);
        if (dot(p.v, d) <= 1e-12) return 0;
        for (int k = n; k > 0; --k) W[k] = W[k - 1];     /* W.insert(W.begin(), p) */
        W[0] = p;
        ++n;
        v3 a = W[0].v, ao = neg(a);
        if (n == 2) {
            v3 ab = sub(W[1].v, a);
            d = cross(cross(ab, ao), ab);
            if (norm2(d) < 1e-24) d = mk(ab.y, -ab.x, 0);
        } else if (n == 3) {
            v3 ab = sub(W[1].v, a), ac = sub(W[2].v, a);
            v3 abc = cross(ab, ac);
            if (dot(abc, ao) > 0) d = abc;
            else { supv t = W[1]; W[1] = W[2]; W[2] = t; d = neg(abc); }
            if (norm2(d) < 1e-24) return 0;
        } else {
            v3 b = W[1].v, c = W[2].v, dd = W[3].v;
            v3 ab = sub(b, a), ac = sub(c, a), ad = sub(dd, a);
            v3 abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
            if (dot(abc, ao) > 0) { n = 3; d = abc; }
            else if (dot(acd, ao) > 0) { supv w2 = W[2], w3 = W[3]; W[1] = w2; W[2] = w3; n = 3; d = acd; }
            else if (dot(adb, ao) > 0) { supv w1 = W[1], w3 = W[3]; W[1] = w3; W[2] = w1; n = 3; d = adb; }
            else { tet[0] = W[0]; tet[1] = W[1]; tet[2] = W[2]; tet[3] = W[3]; return 1; }
        }
    }
    return 0;
}

/* EPA: src/epa.hpp:71-150 */
typedef struct { int a, b, c; v3 n; double dist; } oface;

static oface make_face(const supv* verts, int i, int j, int k) {                    /* epa.hpp:76-83 */
    oface f;
    v3 n = cross(sub(verts[j].v, verts[i].v), sub(verts[k].v, verts[i].v));
    double L = norm(n);
    n = (L > 1e-18) ? dvs(n, L) : mk(0, 0, 1);
    double dist = dot(n, verts[i].v);
    if (dist < 0) { n = neg(n); dist = -dist; int t = j; j = k; k = t; }
    f.a = i; f.b = j; f.c = k; f.n = n; f.dist = dist;
    return f;
}

/* Largest EPA working set seen by this process (faces after an iteration,
 * horizon edges): sizes the fixed workspace of a device implementation. */
static int g_epa_max_faces = 0, g_epa_max_horizon = 0;
static long g_epa_calls = 0, g_epa_big = 0, g_epa_big_faces = 160, g_epa_big_horizon = 96;

void ne_oracle_epa_stats(int32_t* max_faces, int32_t* max_horizon) {
    if (max_faces) *max_faces = g_epa_max_faces;
    if (max_horizon) *max_horizon = g_epa_max_horizon;
}

/* Resets the statistics; later EPA calls whose face list exceeds `faces` or
 * whose horizon exceeds `horizon` are counted as big. */
void ne_oracle_epa_stats_reset(int32_t faces, int32_t horizon) {
    g_epa_max_faces = g_epa_max_horizon = 0;
    g_epa_calls = g_epa_big = 0;
    g_epa_big_faces = faces;
    g_epa_big_horizon = horizon;
}

void ne_oracle_epa_counts(int64_t* calls, int64_t* big) {
    if (calls) *calls = g_epa_calls;
    if (big) *big = g_epa_big;
}

static ocontact epa_contact(const ne_oracle* o, const obody* A, const obody* B, v3 shift, const supv tet[4]) {
    int nv = 4, capv = 80;
    supv* verts = (supv*)malloc(sizeof(supv) * capv);
    for (int k = 0; k < 4; ++k) verts[k] = tet[k];
    int nf = 4, capf = 64;
    oface* faces = (oface*)malloc(sizeof(oface) * capf);
    oface* keep = (oface*)malloc(sizeof(oface) * capf);
    int caph = 64;
    int (*hor)[2] = malloc(sizeof(int[2]) * caph);
    faces[0] = make_face(verts, 0, 1, 2);
    faces[1] = make_face(verts, 0, 2, 3);
    faces[2] = make_face(verts, 0, 3, 1);
    faces[3] = make_face(verts, 1, 3, 2);
    oface best;
    memset(&best, 0, sizeof(best));
    int was_big = 0;
    ++g_epa_calls;
    for (int iter = 0; iter < 64; ++iter) {
        int ci = 0;
        for (int i = 1; i < nf; ++i)
            if (faces[i].dist < faces[ci].dist) ci = i;
        best = faces[ci];
        supv p = cso_support(o, A, B, shift, best.n);
        double d = dot(p.v, best.n);
        if (d - best.dist < 1e-7) break;
        int pidx = nv;
        if (nv == capv) { capv *= 2; verts = (supv*)realloc(verts, sizeof(supv) * capv); }
        verts[nv++] = p;
        int nh = 0, nk = 0;
        for (int fi = 0; fi < nf; ++fi) {
            const oface* f = &faces[fi];
            if (dot(f->n, sub(p.v, verts[f->a].v)) > 1e-12) {
                int e3[3][2] = {{f->a, f->b}, {f->b, f->c}, {f->c, f->a}};
                for (int q = 0; q < 3; ++q) {
                    int u = e3[q][0], w = e3[q][1], found = 0;
                    for (int e = 0; e < nh; ++e) {
                        if (hor[e][0] == w && hor[e][1] == u) {        /* vector::erase */
                            for (int m = e; m + 1 < nh; ++m) { hor[m][0] = hor[m + 1][0]; hor[m][1] = hor[m + 1][1]; }
                            --nh; found = 1; break;
                        }
                    }
                    if (!found) {
                        if (nh == caph) { caph *= 2; hor = realloc(hor, sizeof(int[2]) * caph); }
                        hor[nh][0] = u; hor[nh][1] = w; ++nh;
                    }
                }
            } else {
                keep[nk++] = *f;
            }
        }
        int need = nk + nh;
        if (need > g_epa_max_faces) g_epa_max_faces = need;
        if (nh > g_epa_max_horizon) g_epa_max_horizon = nh;
        if (!was_big && (need > g_epa_big_faces || nh > g_epa_big_horizon)) { was_big = 1; ++g_epa_big; }
        if (need > capf) {
            while (capf < need) capf *= 2;
            faces = (oface*)realloc(faces, sizeof(oface) * capf);
            keep = (oface*)realloc(keep, sizeof(oface) * capf);
        }
        memcpy(faces, keep, sizeof(oface) * nk);
        nf = nk;
        for (int e = 0; e < nh; ++e) faces[nf++] = make_face(verts, hor[e][0], hor[e][1], pidx);
        if (nf == 0) break;
    }
    /* barycentric witness points: epa.hpp:123-149 */
    supv A0 = verts[best.a], B0 = verts[best.b], C0 = verts[best.c];
    v3 proj = scl(best.n, best.dist);
    v3 v0 = sub(B0.v, A0.v), v1 = sub(C0.v, A0.v), v2 = sub(proj, A0.v);
    double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    double d20 = dot(v2, v0), d21 = dot(v2, v1);
    double den = d00 * d11 - d01 * d01;
    double bv = 0, bw = 0;
    if (fabs(den) > 1e-18) {
        bv = (d11 * d20 - d01 * d21) / den;
        bw = (d00 * d21 - d01 * d20) / den;
    }
    double bu = 1.0 - bv - bw;
    v3 onA = add(add(scl(A0.a, bu), scl(B0.a, bv)), scl(C0.a, bw));
    v3 onB = add(add(scl(A0.b, bu), scl(B0.b, bv)), scl(C0.b, bw));
    ocontact c;
    c.hit = 1;
    c.normal = neg(best.n);
    c.overlap = best.dist;
    c.point = scl(add(onA, onB), 0.5);
    free(verts); free(faces); free(keep); free(hor);
    return c;
}

static ocontact convex_contact(const ne_oracle* o, const obody* A, const obody* B) { /* epa.hpp:153-158 */
    ocontact none; memset(&none, 0, sizeof(none));
    v3 shift = min_image_shift(&o->box, A, B);
    supv tet[4];
    if (!gjk_tetra(o, A, B, shift, tet)) return none;
    return epa_contact(o, A, B, shift, tet);
}

/* ------------------------------------------------------------------------ */
/* Analytic narrow phase: src/narrowphase.hpp:34-155                         */
/* ------------------------------------------------------------------------ */

static ocontact sphere_vs_cylinder(const obox* box, const obody* sph, const obody* cyl) { /* narrowphase.hpp:34-81 */
    ocontact c; memset(&c, 0, sizeof(c));
    const double r = cyl->radius, h = cyl->hh, sr = sph->radius;
    v3 rel = min_image(box, sub(sph->x, cyl->x));
    v3 L = qinvrot(cyl->q, rel);
    double rho = sqrt(L.x * L.x + L.z * L.z);
    int insideAxial = fabs(L.y) < h;
    int insideRadial = rho < r;
    v3 Pl, nl;
    if (insideAxial && insideRadial) {
        double dCap = h - fabs(L.y);
        double dSide = r - rho;
        if (dSide < dCap) {
            double s = (rho > 1e-12) ? r / rho : 0.0;
            Pl = mk(L.x * s, L.y, L.z * s);
            nl = (rho > 1e-12) ? mk(L.x / rho, 0, L.z / rho) : mk(1, 0, 0);
            c.overlap = sr + dSide;
        } else {
            double sy = (L.y >= 0) ? h : -h;
            Pl = mk(L.x, sy, L.z);
            nl = mk(0, (L.y >= 0) ? 1.0 : -1.0, 0);
            c.overlap = sr + dCap;
        }
        c.hit = 1;
    } else {
        double yc = (L.y > h) ? h : (L.y < -h ? -h : L.y);
        double rx = L.x, rz = L.z;
        if (rho > r) { double s = r / rho; rx *= s; rz *= s; }
        Pl = mk(rx, yc, rz);
        v3 diff = sub(L, Pl);
        double dist = norm(diff);
        if (dist >= sr) return c;
        c.hit = 1;
        c.overlap = sr - dist;
        nl = (dist > 1e-12) ? dvs(diff, dist) : mk(0, 1, 0);
    }
    c.normal = normalized(qrot(cyl->q, nl));
    c.point = add(cyl->x, qrot(cyl->q, Pl));
    return c;
}

static double np_clamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); } /* narrowphase.hpp:86 */

static v3 closest_on_segment(v3 p, v3 a, v3 b) {                                    /* narrowphase.hpp:88-93 */
    v3 ab = sub(b, a);
    double L2 = norm2(ab);
    if (L2 < 1e-18) return a;
    return add(a, scl(ab, np_clamp(dot(sub(p, a), ab) / L2, 0.0, 1.0)));
}

static void closest_seg_seg(v3 p1, v3 q1, v3 p2, v3 q2, v3* c1, v3* c2) {           /* narrowphase.hpp:96-116 */
    v3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
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
    *c1 = add(p1, scl(d1, s));
    *c2 = add(p2, scl(d2, t));
}

static ocontact capsule_vs_sphere(const obox* box, const obody* cap, const obody* sph) { /* narrowphase.hpp:120-135 */
    ocontact c; memset(&c, 0, sizeof(c));
    v3 sc = add(cap->x, min_image(box, sub(sph->x, cap->x)));
    v3 axis = qrot(cap->q, mk(0, 1, 0));
    v3 cp = closest_on_segment(sc, sub(cap->x, scl(axis, cap->hh)), add(cap->x, scl(axis, cap->hh)));
    v3 e = sub(sc, cp);
    double dist = norm(e);
    double sumR = cap->radius + sph->radius;
    if (dist >= sumR) return c;
    v3 nrm = (dist > 1e-12) ? dvs(e, dist) : mk(0, 1, 0);
    c.hit = 1;
    c.overlap = sumR - dist;
    c.normal = neg(nrm);
    c.point = add(cp, scl(nrm, cap->radius));
    return c;
}

static ocontact capsule_vs_capsule(const obox* box, const obody* A, const obody* B) { /* narrowphase.hpp:138-155 */
    ocontact c; memset(&c, 0, sizeof(c));
    v3 axA = qrot(A->q, mk(0, 1, 0)), axB = qrot(B->q, mk(0, 1, 0));
    v3 bc = add(A->x, min_image(box, sub(B->x, A->x)));
    v3 pA, pB;
    closest_seg_seg(sub(A->x, scl(axA, A->hh)), add(A->x, scl(axA, A->hh)),
                    sub(bc, scl(axB, B->hh)), add(bc, scl(axB, B->hh)), &pA, &pB);
    v3 e = sub(pB, pA);
    double dist = norm(e);
    double sumR = A->radius + B->radius;
    if (dist >= sumR) return c;
    v3 nrm = (dist > 1e-12) ? dvs(e, dist) : mk(0, 1, 0);
    c.hit = 1;
    c.overlap = sumR - dist;
    c.normal = neg(nrm);
    c.point = add(pA, scl(nrm, A->radius));
    return c;
}

static ocontact plane_vs_convex(const ne_oracle* o, const obody* other, const obody* plane) { /* detect.hpp:19-31 */
    ocontact c; memset(&c, 0, sizeof(c));
    v3 n = plane->he;
    v3 sp = support_world(o, other, neg(n));
    v3 rel = min_image(&o->box, sub(sp, plane->x));
    double pen = -dot(n, rel);
    if (pen <= 0.0) return c;
    c.hit = 1;
    c.overlap = pen;
    c.normal = n;
    c.point = add(sp, scl(n, 0.5 * pen));
    return c;
}

static ocontact detect_contact(const ne_oracle* o, const obody* a, const obody* b) {  /* detect.hpp:33-77 */
    ocontact c; memset(&c, 0, sizeof(c));
    const obox* box = &o->box;
    if (a->shape == NE_SHAPE_PLANE || b->shape == NE_SHAPE_PLANE) {
        if (a->shape == NE_SHAPE_PLANE && b->shape == NE_SHAPE_PLANE) return c;
        const obody* plane = (a->shape == NE_SHAPE_PLANE) ? a : b;
        const obody* other = (a->shape == NE_SHAPE_PLANE) ? b : a;
        ocontact t = plane_vs_convex(o, other, plane);
        if (t.hit && a->shape == NE_SHAPE_PLANE) t.normal = neg(t.normal);
        return t;
    }
    if (is_sphere(a) && is_sphere(b)) {
        v3 d = min_image(box, sub(a->x, b->x));
        double sumR = a->radius + b->radius, dist2 = norm2(d);
        if (dist2 >= sumR * sumR || dist2 < 1e-18) return c;
        double dist = sqrt(dist2);
        c.hit = 1;
        c.normal = dvs(d, dist);
        c.overlap = sumR - dist;
        c.point = sub(a->x, scl(c.normal, a->radius));
        return c;
    }
    int aCyl = a->shape == NE_SHAPE_CYLINDER, bCyl = b->shape == NE_SHAPE_CYLINDER;
    if (is_sphere(a) && bCyl) return sphere_vs_cylinder(box, a, b);
    if (is_sphere(b) && aCyl) {
        ocontact t = sphere_vs_cylinder(box, b, a);
        if (t.hit) t.normal = neg(t.normal);
        return t;
    }
    int aCap = a->shape == NE_SHAPE_CAPSULE, bCap = b->shape == NE_SHAPE_CAPSULE;
    if (aCap && bCap) return capsule_vs_capsule(box, a, b);
    if (aCap && is_sphere(b)) return capsule_vs_sphere(box, a, b);
    if (is_sphere(a) && bCap) {
        ocontact t = capsule_vs_sphere(box, b, a);
        if (t.hit) t.normal = neg(t.normal);
        return t;
    }
    return convex_contact(o, a, b);
}

/* ------------------------------------------------------------------------ */
/* Box-box manifold: src/manifold.hpp:16-126                                 */
/* ------------------------------------------------------------------------ */

typedef struct { v3 n, c, u, v; double eu, ev; v3 vert[4]; } bface;

static v3 axis_vec(int a) { return a == 0 ? mk(1, 0, 0) : (a == 1 ? mk(0, 1, 0) : mk(0, 0, 1)); } /* manifold.hpp:24-26 */

static bface box_face(const obody* b, v3 centre, int axis, int sign) {              /* manifold.hpp:39-52 */
    const int t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
    const double h[3] = {b->he.x, b->he.y, b->he.z};
    bface f;
    f.n = qrot(b->q, scl(axis_vec(axis), (double)sign));
    f.c = add(centre, scl(f.n, h[axis]));
    f.u = qrot(b->q, axis_vec(t1)); f.eu = h[t1];
    f.v = qrot(b->q, axis_vec(t2)); f.ev = h[t2];
    f.vert[0] = add(add(f.c, scl(f.u, f.eu)), scl(f.v, f.ev));
    f.vert[1] = add(sub(f.c, scl(f.u, f.eu)), scl(f.v, f.ev));
    f.vert[2] = sub(sub(f.c, scl(f.u, f.eu)), scl(f.v, f.ev));
    f.vert[3] = sub(add(f.c, scl(f.u, f.eu)), scl(f.v, f.ev));
    return f;
}

static bface best_face(const obody* b, v3 centre, v3 dir) {                         /* manifold.hpp:55-64 */
    int bestAxis = 0, bestSign = 1;
    double best = -1e30;
    for (int axis = 0; axis < 3; ++axis)
        for (int sign = -1; sign <= 1; sign += 2) {
            double a = dot(qrot(b->q, scl(axis_vec(axis), (double)sign)), dir);
            if (a > best) { best = a; bestAxis = axis; bestSign = sign; }
        }
    return box_face(b, centre, bestAxis, bestSign);
}

static int clip_plane(v3* poly, int m, v3 pn, double off, v3* out) {               /* manifold.hpp:67-82 */
    int k = 0;
    for (int i = 0; i < m; ++i) {
        v3 cur = poly[i];
        v3 nxt = poly[(i + 1) % m];
        double dc = dot(cur, pn) - off, dn = dot(nxt, pn) - off;
        int inC = dc <= 0.0, inN = dn <= 0.0;
        if (inC) out[k++] = cur;
        if (inC != inN) {
            double t = dc / (dc - dn);
            out[k++] = add(cur, scl(sub(nxt, cur), t));
        }
    }
    return k;
}

static int box_box_manifold(const ne_oracle* o, const obody* a, const obody* b, v3 n,
                            v3 pts[4], double depth[4]) {                           /* manifold.hpp:90-126 */
    int count = 0;
    const v3 bShift = min_image_shift(&o->box, a, b);
    const v3 bC = add(b->x, bShift);
    bface fa = best_face(a, a->x, neg(n));
    bface fb = best_face(b, bC, n);
    double alignA = dot(fa.n, neg(n)), alignB = dot(fb.n, n);
    const bface ref = (alignA >= alignB) ? fa : fb;
    const obody* incBody = (alignA >= alignB) ? b : a;
    const v3 incCentre = (alignA >= alignB) ? bC : a->x;
    bface inc = best_face(incBody, incCentre, neg(ref.n));
    v3 p0[64], p1[64];   /* each input vertex emits <= 2 outputs: 4 -> 64 over 4 planes */
    int m = 4;
    for (int k = 0; k < 4; ++k) p0[k] = inc.vert[k];
    m = clip_plane(p0, m, ref.u, dot(ref.c, ref.u) + ref.eu, p1);
    m = clip_plane(p1, m, scl(ref.u, -1.0), -(dot(ref.c, ref.u) - ref.eu), p0);
    m = clip_plane(p0, m, ref.v, dot(ref.c, ref.v) + ref.ev, p1);
    m = clip_plane(p1, m, scl(ref.v, -1.0), -(dot(ref.c, ref.v) - ref.ev), p0);
    const double slop = 0.01;
    for (int k = 0; k < m; ++k) {
        double sd = dot(sub(p0[k], ref.c), ref.n);
        if (sd > slop) continue;
        if (count >= 4) break;
        pts[count] = sub(p0[k], scl(ref.n, sd));
        depth[count] = -sd;
        ++count;
    }
    return count;
}

/* gatherContacts for non-mesh, non-compound shapes: src/world.cpp:122-129 */
static void cs_reserve(ne_oracle* o, uint32_t need) {
    if (need <= o->capcs) return;
    uint32_t cap = o->capcs ? o->capcs : 16;
    while (cap < need) cap *= 2;
    o->cs = (ocontact*)realloc(o->cs, sizeof(ocontact) * cap);
    o->capcs = cap;
}

/* MeshData::query (mesh.hpp:46-59): depth-first with an explicit stack of 64,
 * right child popped first; children are dropped once 62 entries are held. */
static uint32_t mesh_query(const omesh* m, v3 lo, v3 hi, int32_t** out, uint32_t* cap) {
    uint32_t n = 0;
    if (m->nn == 0) return 0;
    int32_t stack[64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const onode* nd = &m->nodes[stack[--sp]];
        if (lo.x > nd->hi.x || hi.x < nd->lo.x || lo.y > nd->hi.y || hi.y < nd->lo.y || lo.z > nd->hi.z ||
            hi.z < nd->lo.z)
            continue;
        if (nd->left < 0) {
            for (int32_t i = 0; i < nd->count; ++i) {
                if (n == *cap) {
                    *cap = *cap ? 2 * *cap : 64;
                    *out = (int32_t*)realloc(*out, sizeof(int32_t) * *cap);
                }
                (*out)[n++] = m->order[nd->start + i];
            }
        } else if (sp < 62) {
            stack[sp++] = nd->left;
            stack[sp++] = nd->right;
        }
    }
    return n;
}

/* meshContacts (world.cpp:56-83): every triangle whose box meets the body's
 * cube of half-side boundingRadius + 0.05 (in the mesh frame) becomes a prism
 * of thickness 2e-3 about its plane, a temporary convex at the origin, and
 * goes through detectContact(other, prism).  Normals point mesh -> other. */
static uint32_t mesh_contacts(ne_oracle* o, const obody* other, const obody* mb, uint32_t at) {
    if (mb->mesh < 0) return 0;
    const omesh* m = &o->meshes[mb->mesh];
    v3 rel = min_image(&o->box, sub(other->x, mb->x));
    v3 cl = qinvrot(mb->q, rel);
    double rr = bounding_radius(o, other) + 0.05;
    int32_t* tris = NULL;
    uint32_t cap = 0;
    uint32_t nt = mesh_query(m, sub(cl, mk(rr, rr, rr)), add(cl, mk(rr, rr, rr)), &tris, &cap);
    const double thick = 1e-3;
    obody tri;
    memset(&tri, 0, sizeof(tri));
    tri.q = mq(1.0, 0.0, 0.0, 0.0);
    tri.shape = NE_SHAPE_CONVEX;
    tri.mesh = -1;
    tri.hull0 = 0;
    tri.hull1 = 6;
    v3 pv[6];
    tri.hvp = pv;
    uint32_t nc = 0;
    for (uint32_t k = 0; k < nt; ++k) {
        const int32_t* t = m->tris + 3 * (size_t)tris[k];
        v3 w0 = add(mb->x, qrot(mb->q, m->verts[t[0]]));
        v3 w1 = add(mb->x, qrot(mb->q, m->verts[t[1]]));
        v3 w2 = add(mb->x, qrot(mb->q, m->verts[t[2]]));
        v3 n = cross(sub(w1, w0), sub(w2, w0));
        double nl = norm(n);
        if (nl < 1e-12) continue;
        n = dvs(n, nl);
        pv[0] = add(w0, scl(n, thick)); pv[1] = add(w1, scl(n, thick)); pv[2] = add(w2, scl(n, thick));
        pv[3] = sub(w0, scl(n, thick)); pv[4] = sub(w1, scl(n, thick)); pv[5] = sub(w2, scl(n, thick));
        ocontact c = detect_contact(o, other, &tri);
        if (c.hit) {
            cs_reserve(o, at + nc + 1);
            o->cs[at + nc++] = c;
        }
    }
    free(tris);
    return nc;
}

/* gatherContacts (world.cpp:88-127) into o->cs; returns the count. */
static int gather_contacts(ne_oracle* o, const obody* a, const obody* b) {
    if (b->shape == NE_SHAPE_MESH) return (int)mesh_contacts(o, a, b, 0);
    if (a->shape == NE_SHAPE_MESH) {
        int m = (int)mesh_contacts(o, b, a, 0);
        for (int k = 0; k < m; ++k) o->cs[k].normal = neg(o->cs[k].normal);
        return m;
    }
    cs_reserve(o, 4);
    ocontact* out = o->cs;
    ocontact c = detect_contact(o, a, b);
    if (!c.hit) return 0;
    if (a->shape == NE_SHAPE_BOX && b->shape == NE_SHAPE_BOX) {
        v3 pts[4]; double dep[4];
        int m = box_box_manifold(o, a, b, c.normal, pts, dep);
        for (int p = 0; p < m; ++p) { out[p].hit = 1; out[p].normal = c.normal; out[p].point = pts[p]; out[p].overlap = dep[p]; }
        if (m > 0) return m;
    }
    out[0] = c;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Broadphase: the candidate set only prunes pairs whose bounding spheres    */
/* (src/body.hpp:137-158) cannot touch, so the sorted pair list that reaches */
/* the narrow phase yields the same constraint list as the reference's       */
/* grid and its brute-force scan (src/world.cpp:418-501, selftest 744-774).   */
/* Here: planes/meshes and bodies much larger than the typical body are      */
/* paired against everything; the rest go through a sorted uniform grid.   */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t i, j; } opair;

static int cmp_pair(const void* A, const void* B) {
    const opair* a = (const opair*)A; const opair* b = (const opair*)B;
    if (a->i != b->i) return a->i < b->i ? -1 : 1;
    if (a->j != b->j) return a->j < b->j ? -1 : 1;
    return 0;
}

typedef struct { uint64_t key; uint32_t idx; } okey;
static int cmp_key(const void* A, const void* B) {
    const okey* a = (const okey*)A; const okey* b = (const okey*)B;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

typedef struct { opair* p; size_t n, cap; } pairvec;
static void pv_push(pairvec* v, uint32_t i, uint32_t j) {
    if (v->n == v->cap) { v->cap = v->cap ? v->cap * 2 : 1024; v->p = (opair*)realloc(v->p, sizeof(opair) * v->cap); }
    v->p[v->n].i = i; v->p[v->n].j = j; v->n++;
}

static int bounds_may_touch(const ne_oracle* o, const double* R, uint32_t i, uint32_t j) {
    if (R[i] >= 1e29 || R[j] >= 1e29) return 1;
    v3 d = min_image(&o->box, sub(o->B[i].x, o->B[j].x));
    double s = (R[i] + R[j]) * (1.0 + 1e-9) + 1e-9;
    return norm2(d) <= s * s;
}

static int64_t floor_div_i64(double c, double cell) { return (int64_t)floor(c / cell); }

static void broadphase(const ne_oracle* o, pairvec* out) {
    const uint32_t n = o->n;
    out->n = 0;
    double* R = (double*)malloc(sizeof(double) * (n ? n : 1));
    double maxDyn = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        R[i] = bounding_radius(o, &o->B[i]);
        if (!grid_excluded(&o->B[i]) && o->B[i].invMass > 0.0 && R[i] > maxDyn) maxDyn = R[i];
    }
    /* large statics (e.g. a ground box) are paired against everything */
    unsigned char* ex = (unsigned char*)calloc(n ? n : 1, 1);
    double maxR = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        const obody* b = &o->B[i];
        if (grid_excluded(b) || R[i] >= 1e29) ex[i] = 1;
        else if (maxDyn > 0.0 && R[i] > 2.0 * maxDyn) ex[i] = 1;
        if (!ex[i] && R[i] > maxR) maxR = R[i];
    }
    double cell = 2.0 * maxR * (1.0 + 1e-6) + 1e-9;
    if (n < 64 || maxR <= 0.0) {
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t j = i + 1; j < n; ++j)
                if (bounds_may_touch(o, R, i, j)) pv_push(out, i, j);
        free(R); free(ex);
        return;
    }
    int64_t nc[3] = {0, 0, 0};
    double cw[3] = {cell, cell, cell};
    if (o->box.periodic) {
        for (int a = 0; a < 3; ++a) {
            double L = 2.0 * o->box.half[a];
            int64_t k = (int64_t)floor(L / cell);
            if (k < 1) k = 1;
            if (k > 2097151) k = 2097151;
            nc[a] = k;
            cw[a] = L / (double)k;
        }
    }
    okey* keys = (okey*)malloc(sizeof(okey) * n);
    int64_t (*cc)[3] = malloc(sizeof(int64_t[3]) * n);
    uint32_t nk = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (ex[i]) continue;
        v3 x = o->B[i].x;
        int64_t c3[3];
        for (int a = 0; a < 3; ++a) {
            double c = comp(x, a);
            if (o->box.periodic) {
                int64_t k = (int64_t)floor((c + o->box.half[a]) / cw[a]);
                k %= nc[a]; if (k < 0) k += nc[a];
                c3[a] = k;
            } else {
                c3[a] = floor_div_i64(c, cell);
            }
        }
        cc[i][0] = c3[0]; cc[i][1] = c3[1]; cc[i][2] = c3[2];
        keys[nk].key = (uint64_t)(c3[0] & 0x1FFFFF) | ((uint64_t)(c3[1] & 0x1FFFFF) << 21) |
                       ((uint64_t)(c3[2] & 0x1FFFFF) << 42);
        keys[nk].idx = i;
        ++nk;
    }
    qsort(keys, nk, sizeof(okey), cmp_key);
    for (uint32_t s = 0; s < nk; ++s) {
        uint32_t i = keys[s].idx;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    int64_t c3[3] = {cc[i][0] + dx, cc[i][1] + dy, cc[i][2] + dz};
                    if (o->box.periodic)
                        for (int a = 0; a < 3; ++a) { c3[a] %= nc[a]; if (c3[a] < 0) c3[a] += nc[a]; }
                    uint64_t key = (uint64_t)(c3[0] & 0x1FFFFF) | ((uint64_t)(c3[1] & 0x1FFFFF) << 21) |
                                   ((uint64_t)(c3[2] & 0x1FFFFF) << 42);
                    /* lower bound */
                    uint32_t lo = 0, hi = nk;
                    while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (keys[mid].key < key) lo = mid + 1; else hi = mid; }
                    for (uint32_t t = lo; t < nk && keys[t].key == key; ++t) {
                        uint32_t j = keys[t].idx;
                        if (j > i && bounds_may_touch(o, R, i, j)) pv_push(out, i, j);
                    }
                }
    }
    for (uint32_t i = 0; i < n; ++i) {
        if (!ex[i]) continue;
        for (uint32_t j = 0; j < n; ++j) {
            if (j == i) continue;
            if (ex[j] && j < i) continue;   /* excluded-excluded pair emitted once */
            uint32_t a = i < j ? i : j, b = i < j ? j : i;
            if (bounds_may_touch(o, R, a, b)) pv_push(out, a, b);
        }
    }
    qsort(out->p, out->n, sizeof(opair), cmp_pair);
    size_t u = 0;
    for (size_t k = 0; k < out->n; ++k)
        if (u == 0 || out->p[k].i != out->p[u - 1].i || out->p[k].j != out->p[u - 1].j) out->p[u++] = out->p[k];
    out->n = u;
    free(keys); free(cc); free(R); free(ex);
}

/* ------------------------------------------------------------------------ */
/* Solver: src/world.cpp:22-51, 503-627; src/contact.hpp:28-33, 78-85         */
/* ------------------------------------------------------------------------ */

struct ocons {                                                                       /* world.cpp:22-32 */
    uint32_t i, j;
    v3 n, rA, rB, point;
    double overlap, restBias, mu, nImp, tImp;
    uint64_t pairKey;
    int slot;
    int bounce;   /* vn < -0.5 at the build: the restitution bias is on (world.cpp:544) */
};

static uint64_t pair_key_of(uint32_t i, uint32_t j) { return ((uint64_t)i << 32) | (uint32_t)j; } /* world.cpp:34-36 */

static double combine_mat(int m, double a, double b) {                               /* world.cpp:42-51 */
    switch (m) {
        case NE_COMBINE_AVERAGE: return 0.5 * (a + b);
        case NE_COMBINE_MIN: return a < b ? a : b;
        case NE_COMBINE_MAX: return a > b ? a : b;
        case NE_COMBINE_MULTIPLY: return a * b;
        case NE_COMBINE_GEOMETRIC_MEAN: return sqrt((a > 0 ? a : 0) * (b > 0 ? b : 0));
    }
    return 0.5 * (a + b);
}

static double eff_mass(const obody* a, const obody* b, v3 ra, v3 rb, v3 u) {         /* contact.hpp:28-33 */
    v3 angA = cross(apply_inv_inertia(a, cross(ra, u)), ra);
    v3 angB = cross(apply_inv_inertia(b, cross(rb, u)), rb);
    return a->invMass + b->invMass + dot(add(angA, angB), u);
}

static void apply_impulse(obody* a, obody* b, v3 ra, v3 rb, v3 J) {                  /* world.cpp:565-566 */
    a->v = add(a->v, scl(J, a->invMass));
    a->w = add(a->w, apply_inv_inertia(a, cross(ra, J)));
    b->v = sub(b->v, scl(J, b->invMass));
    b->w = sub(b->w, apply_inv_inertia(b, cross(rb, J)));
}

static v3 contact_vel(const obody* a, const obody* b, v3 ra, v3 rb) {
    return sub(add(a->v, cross(a->w, ra)), add(b->v, cross(b->w, rb)));
}

static int layers_collide(const obody* a, const obody* b) {                          /* world.hpp:94-96 */
    return (a->mask & (1u << (b->layer & 31u))) && (b->mask & (1u << (a->layer & 31u)));
}

static int pair_excluded(const ne_oracle* o, const obody* a, uint32_t j) {           /* world.hpp:100-103 */
    for (uint32_t k = a->nc0; k < a->nc1; ++k) if (o->nocol[k] == j) return 1;
    return 0;
}

static double eff_friction(const ne_oracle* o, const obody* b) { return b->friction < 0 ? o->P.friction : b->friction; }
static double eff_restitution(const ne_oracle* o, const obody* b) { return b->restitution < 0 ? o->P.restitution : b->restitution; }

static void push_cons(ne_oracle* o, const struct ocons* k) {
    if (o->ncons == o->capcons) {
        o->capcons = o->capcons ? o->capcons * 2 : 1024;
        o->cons = (struct ocons*)realloc(o->cons, sizeof(struct ocons) * o->capcons);
    }
    o->cons[o->ncons++] = *k;
}

/* warm cache lookup: previous step's list is sorted by (key, slot) */
static int warm_lookup(const ne_oracle* o, uint64_t key, int slot, double* out) {
    uint32_t lo = 0, hi = o->ncache;
    while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (o->cache_key[mid] < key) lo = mid + 1; else hi = mid; }
    if (lo >= o->ncache || o->cache_key[lo] != key) return 0;
    uint32_t cnt = 0;
    while (lo + cnt < o->ncache && o->cache_key[lo + cnt] == key) ++cnt;
    if ((uint32_t)slot >= cnt) return 0;
    *out = o->cache_imp[lo + slot];
    return 1;
}

static void wake(obody* b) {                                                         /* world.cpp:317-323 */
    if (!(b->flags & NE_FLAG_SLEEPING)) return;
    b->invMass = b->invMassStore;
    b->invI = b->invIStore;
    b->flags &= ~NE_FLAG_SLEEPING;
    b->sleepTimer = 0.0;
}

static void tape_push(struct otape_step* st, uint32_t i, uint32_t j, int nc, const ocontact* cs) {
    if (st->n == st->cap) {
        st->cap = st->cap ? 2 * st->cap : 256;
        st->p = (struct otape_pair*)realloc(st->p, sizeof(struct otape_pair) * st->cap);
    }
    struct otape_pair* r = &st->p[st->n++];
    r->i = i; r->j = j; r->nc = nc;
    memcpy(r->cs, cs, sizeof(ocontact) * (size_t)nc);
}

static void collide(ne_oracle* o) {                                                  /* world.cpp:503-627 */
    pairvec pv = {0, 0, 0};
    const struct otape_step* replay = NULL;
    struct otape_step* record = NULL;
    if (o->tape_mode >= 2) {
        replay = (o->tape_cursor < o->tape_steps) ? &o->tape[o->tape_cursor] : NULL;
        ++o->tape_cursor;
    } else {
        broadphase(o, &pv);
        if (o->tape_mode == 1) {
            if (o->tape_steps == o->tape_cap) {
                o->tape_cap = o->tape_cap ? 2 * o->tape_cap : 64;
                o->tape = (struct otape_step*)realloc(o->tape, sizeof(struct otape_step) * o->tape_cap);
            }
            record = &o->tape[o->tape_steps++];
            memset(record, 0, sizeof(*record));
        }
    }
    const size_t npairs = (o->tape_mode >= 2) ? (replay ? replay->n : 0) : pv.n;
    const struct otape_step* frozen = (o->tape_mode == 3) ? replay : NULL;
    o->ncand = (uint32_t)npairs;
    o->ncons = 0;
    o->nisp = 0;
    const ne_world_params* P = &o->P;
    for (size_t pi = 0; pi < npairs; ++pi) {
        const uint32_t i = replay ? replay->p[pi].i : pv.p[pi].i, j = replay ? replay->p[pi].j : pv.p[pi].j;
        obody* a = &o->B[i];
        obody* b = &o->B[j];
        int nc;
        if (replay) {                 /* the recorded contacts of this pair: fixed contact set */
            nc = replay->p[pi].nc;
            cs_reserve(o, 4);
            memcpy(o->cs, replay->p[pi].cs, sizeof(ocontact) * (size_t)nc);
        } else {
            if (a->invMass + b->invMass <= 0.0) continue;
            if (!layers_collide(a, b)) continue;
            if (pair_excluded(o, a, j)) continue;
            if ((a->flags & NE_FLAG_SENSOR) || (b->flags & NE_FLAG_SENSOR)) continue;
            nc = gather_contacts(o, a, b);
            if (nc == 0) continue;
            if (record) tape_push(record, i, j, nc > 4 ? 4 : nc, o->cs);   /* tapes hold no mesh pairs */
        }
        const ocontact* cs = o->cs;
        const double e = combine_mat(P->restitution_combine, eff_restitution(o, a), eff_restitution(o, b));
        const double mu = combine_mat(P->friction_combine, eff_friction(o, a), eff_friction(o, b));
        if (P->sleep_enabled) {                                                      /* world.cpp:527-531 */
            int aS = (a->flags & NE_FLAG_SLEEPING) != 0, bS = (b->flags & NE_FLAG_SLEEPING) != 0;
            if (aS && !bS && (b->flags & (NE_FLAG_DYNAMIC | NE_FLAG_KINEMATIC)) && norm(b->v) > P->wake_vel) wake(a);
            aS = (a->flags & NE_FLAG_SLEEPING) != 0;
            if (bS && !aS && (a->flags & (NE_FLAG_DYNAMIC | NE_FLAG_KINEMATIC)) && norm(a->v) > P->wake_vel) wake(b);
            bS = (b->flags & NE_FLAG_SLEEPING) != 0;
            if (aS && bS) continue;
        }
        if (a->invMass > 0.0 && b->invMass > 0.0) {                                 /* world.cpp:533 */
            if (o->nisp == o->capisp) {
                o->capisp = o->capisp ? 2 * o->capisp : 256;
                o->isp = (uint32_t*)realloc(o->isp, sizeof(uint32_t) * 2 * o->capisp);
            }
            o->isp[2 * o->nisp] = i;
            o->isp[2 * o->nisp + 1] = j;
            ++o->nisp;
        }
        const uint64_t key = pair_key_of(i, j);
        int slot = 0;
        for (int c = 0; c < nc; ++c) {
            struct ocons k;
            memset(&k, 0, sizeof(k));
            k.i = i; k.j = j; k.n = cs[c].normal; k.overlap = cs[c].overlap; k.mu = mu;
            k.point = cs[c].point;
            k.rA = min_image(&o->box, sub(cs[c].point, a->x));
            k.rB = min_image(&o->box, sub(cs[c].point, b->x));
            v3 vc = contact_vel(a, b, k.rA, k.rB);
            double vn = dot(vc, k.n);
            int bounce = vn < -0.5;
            if (frozen) bounce = (o->ncons < frozen->ncons) ? frozen->rest[o->ncons] : bounce;
            k.restBias = bounce ? -e * vn : 0.0;                                     /* world.cpp:40, 544 */
            k.bounce = bounce;
            k.pairKey = key; k.slot = slot++;
            k.nImp = 0.0; k.tImp = 0.0;
            if (P->warm_start) {
                double w;
                if (warm_lookup(o, key, k.slot, &w)) k.nImp = w;
            }
            push_cons(o, &k);
        }
    }
    free(pv.p);

    struct ocons* C = o->cons;
    const uint32_t nC = o->ncons;
    const int iters = P->solver_iterations;                                          /* world.cpp:572 */
    if (frozen && (frozen->ncons != nC || frozen->iters != (uint32_t)iters)) frozen = NULL;   /* not the recorded set */
    if (record) {
        record->ncons = nC;
        record->iters = (uint32_t)iters;
        record->rest = (uint8_t*)malloc(nC ? nC : 1);
        record->visit = (uint8_t*)malloc((size_t)nC * (size_t)iters + 1);
        for (uint32_t q = 0; q < nC; ++q) record->rest[q] = (uint8_t)C[q].bounce;
    }
    if (P->warm_start) {                                                             /* world.cpp:559-568 */
        for (uint32_t q = 0; q < nC; ++q) {
            if (C[q].nImp == 0.0) continue;
            v3 J = scl(C[q].n, C[q].nImp);
            apply_impulse(&o->B[C[q].i], &o->B[C[q].j], C[q].rA, C[q].rB, J);
        }
    }
    /* In mode 3 every branch below takes the recorded run's side (frozen
     * active set), so the step is one smooth function of its inputs near the
     * recorded ones; otherwise the branches are evaluated, as in the
     * reference, and mode 1 records them. */
    for (int it = 0; it < iters; ++it) {
        for (uint32_t q = 0; q < nC; ++q) {
            struct ocons* k = &C[q];
            obody* a = &o->B[k->i];
            obody* b = &o->B[k->j];
            const uint8_t fz = frozen ? frozen->visit[(size_t)it * nC + q] : 0;
            uint8_t bits = 0;
            v3 vc = contact_vel(a, b, k->rA, k->rB);
            double vn = dot(vc, k->n);
            double Kn = eff_mass(a, b, k->rA, k->rB, k->n);
            if (frozen ? (fz & TB_KN) : Kn > 1e-18) {
                bits |= TB_KN;
                double dj = (k->restBias - vn) / Kn;
                double old = k->nImp;
                const int keep = frozen ? (fz & TB_UNCLAMPED) != 0 : old + dj > 0.0;
                if (keep) bits |= TB_UNCLAMPED;
                k->nImp = keep ? old + dj : 0.0;
                dj = k->nImp - old;
                apply_impulse(a, b, k->rA, k->rB, scl(k->n, dj));
            }
            if (frozen ? (fz & TB_FRIC) : (k->mu > 0.0 && k->nImp > 0.0)) {
                bits |= TB_FRIC;
                v3 vc2 = contact_vel(a, b, k->rA, k->rB);
                v3 vt = sub(vc2, scl(k->n, dot(vc2, k->n)));
                double vtl = norm(vt);
                if (frozen ? (fz & TB_VT) : vtl > 1e-12) {
                    bits |= TB_VT;
                    v3 t = dvs(vt, vtl);
                    double Kt = eff_mass(a, b, k->rA, k->rB, t);
                    if (frozen ? (fz & TB_KT) : Kt > 1e-18) {
                        bits |= TB_KT;
                        double djt = -vtl / Kt;
                        double maxF = k->mu * k->nImp;
                        double old = k->tImp;
                        double sum = old + djt;
                        const int lo = frozen ? (fz & TB_LO) != 0 : sum < -maxF;
                        const int hi = frozen ? (fz & TB_HI) != 0 : (!lo && sum > maxF);
                        if (lo) bits |= TB_LO;
                        if (hi) bits |= TB_HI;
                        k->tImp = lo ? -maxF : (hi ? maxF : sum);
                        djt = k->tImp - old;
                        apply_impulse(a, b, k->rA, k->rB, scl(t, djt));
                    }
                }
            }
            if (record) record->visit[(size_t)it * nC + q] = bits;
        }
    }
    for (uint32_t q = 0; q < nC; ++q) {                                              /* world.cpp:615-616, contact.hpp:78-85 */
        obody* a = &o->B[C[q].i];
        obody* b = &o->B[C[q].j];
        double invSum = a->invMass + b->invMass;
        if (invSum <= 0.0 || C[q].overlap <= 0.0) continue;
        v3 corr = scl(C[q].n, P->contact_beta * C[q].overlap / invSum);
        a->x = add(a->x, scl(corr, a->invMass));
        b->x = sub(b->x, scl(corr, b->invMass));
    }
    if (P->warm_start) {                                                             /* world.cpp:619-626 */
        o->cache_key = (uint64_t*)realloc(o->cache_key, sizeof(uint64_t) * (nC ? nC : 1));
        o->cache_imp = (double*)realloc(o->cache_imp, sizeof(double) * (nC ? nC : 1));
        for (uint32_t q = 0; q < nC; ++q) { o->cache_key[q] = C[q].pairKey; o->cache_imp[q] = C[q].nImp; }
        o->ncache = nC;
    }
}

/* ------------------------------------------------------------------------ */
/* Integrate, boundaries, step: src/world.cpp:133-146, 373-411, 635-702      */
/* ------------------------------------------------------------------------ */

static void integrate(ne_oracle* o) {                                                /* world.cpp:373-411 */
    const double dt = o->P.dt;
    const v3 g = mk(o->P.gravity[0], o->P.gravity[1], o->P.gravity[2]);
    for (uint32_t i = 0; i < o->n; ++i) {
        obody* b = &o->B[i];
        if (b->flags & NE_FLAG_KINEMATIC) {
            b->x = add(b->x, scl(b->v, dt));
            b->q = integrate_orientation(b->q, b->w, dt);
            b->force = mk(0, 0, 0); b->torque = mk(0, 0, 0);
            continue;
        }
        if (b->invMass <= 0.0) { b->force = mk(0, 0, 0); b->torque = mk(0, 0, 0); continue; }
        b->v = add(b->v, scl(add(scl(g, b->gravityScale), scl(b->force, b->invMass)), dt));
        b->x = add(b->x, scl(b->v, dt));
        v3 L = angular_momentum(b);
        L = add(L, scl(b->torque, dt));
        b->q = integrate_orientation(b->q, b->w, dt);
        b->w = apply_inv_inertia(b, L);
        if (b->linDamp > 0.0) b->v = scl(b->v, 1.0 / (1.0 + b->linDamp * dt));
        if (b->angDamp > 0.0) b->w = scl(b->w, 1.0 / (1.0 + b->angDamp * dt));
        b->force = mk(0, 0, 0); b->torque = mk(0, 0, 0);
    }
}

static void apply_walls(ne_oracle* o) {                                              /* world.cpp:635-657 */
    for (uint32_t i = 0; i < o->n; ++i) {
        obody* b = &o->B[i];
        if (b->invMass <= 0.0) continue;
        for (int axis = 0; axis < 3; ++axis) {
            const double h = o->box.half[axis];
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                v3 outward = mk(axis == 0 ? (double)sgn : 0.0, axis == 1 ? (double)sgn : 0.0,
                                axis == 2 ? (double)sgn : 0.0);
                v3 sp = support_world(o, b, outward);
                double coord = comp(sp, axis) * sgn;
                double pen = coord - h;
                if (pen > 0.0) {
                    double* px = axis == 0 ? &b->x.x : (axis == 1 ? &b->x.y : &b->x.z);
                    double* pvv = axis == 0 ? &b->v.x : (axis == 1 ? &b->v.y : &b->v.z);
                    *px -= (double)sgn * pen;
                    if (*pvv * sgn > 0.0) *pvv = -o->P.restitution * *pvv;
                }
            }
        }
    }
}

static void wrap_positions(ne_oracle* o) {                                           /* world.cpp:698-702 */
    for (uint32_t i = 0; i < o->n; ++i)
        if (o->B[i].invMass > 0.0) o->B[i].x = wrap_pos(&o->box, o->B[i].x);
}

/* Spring-damper joints add force and torque before integration
 * (world.cpp:150-170). */
static void apply_joint_forces(ne_oracle* o) {
    for (uint32_t k = 0; k < o->njoints; ++k) {
        const ne_distance_joint* j = &o->joints[k];
        if (j->stiffness <= 0.0) continue;
        obody* a = &o->B[j->a];
        obody* b = &o->B[j->b];
        const v3 wa = add(a->x, qrot(a->q, ld3(j->local_a, 0)));
        const v3 wb = add(b->x, qrot(b->q, ld3(j->local_b, 0)));
        const v3 d = min_image(&o->box, sub(wa, wb));
        const double dist = norm(d);
        if (dist < 1e-9) continue;
        const v3 nrm = dvs(d, dist);
        const v3 rA = sub(wa, a->x), rB = min_image(&o->box, sub(wb, b->x));
        const v3 vrel = contact_vel(a, b, rA, rB);
        const double vn = dot(vrel, nrm);
        const double f = -j->stiffness * (dist - j->rest) - j->damping * vn;
        const v3 F = scl(nrm, f);
        a->force = add(a->force, F);
        a->torque = add(a->torque, cross(rA, F));
        b->force = sub(b->force, F);
        b->torque = sub(b->torque, cross(rB, F));
    }
}

/* Rigid joints: bilateral velocity constraints with a Baumgarte length term,
 * 8 sequential sweeps after the contact solve (world.cpp:173-197). */
static void solve_rigid_joints(ne_oracle* o, int iterations) {
    for (int it = 0; it < iterations; ++it) {
        for (uint32_t k = 0; k < o->njoints; ++k) {
            const ne_distance_joint* j = &o->joints[k];
            if (j->stiffness > 0.0) continue;
            obody* a = &o->B[j->a];
            obody* b = &o->B[j->b];
            const v3 wa = add(a->x, qrot(a->q, ld3(j->local_a, 0)));
            const v3 wb = add(b->x, qrot(b->q, ld3(j->local_b, 0)));
            const v3 d = min_image(&o->box, sub(wa, wb));
            const double dist = norm(d);
            if (dist < 1e-9) continue;
            const v3 nrm = dvs(d, dist);
            const v3 rA = sub(wa, a->x), rB = min_image(&o->box, sub(wb, b->x));
            const double K = eff_mass(a, b, rA, rB, nrm);
            if (K <= 1e-18) continue;
            const v3 vrel = contact_vel(a, b, rA, rB);
            const double vn = dot(vrel, nrm);
            const double bias = 0.2 * (dist - j->rest) / o->P.dt;
            const double dj = -(vn + bias) / K;
            apply_impulse(a, b, rA, rB, scl(nrm, dj));
        }
    }
}

/* Island sleeping (world.cpp:328-371): awake dynamic bodies joined by this
 * step's contact pairs and by distance joints; an island sleeps when every
 * member has been slow for sleep_time. */
static int awake_dyn(const obody* b) {
    return (b->flags & NE_FLAG_DYNAMIC) && !(b->flags & NE_FLAG_SLEEPING) && b->invMass > 0.0;
}

static uint32_t uf_find(uint32_t* parent, uint32_t x) {
    while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; }
    return x;
}

static void update_sleep(ne_oracle* o) {
    const uint32_t n = o->n;
    uint32_t* parent = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) parent[i] = i;
#define JOIN(a_, b_)                                                                  \
    do {                                                                              \
        const uint32_t ja_ = (a_), jb_ = (b_);                                        \
        if (awake_dyn(&o->B[ja_]) && awake_dyn(&o->B[jb_])) {                         \
            const uint32_t ra_ = uf_find(parent, ja_);                                \
            parent[ra_] = uf_find(parent, jb_);                                       \
        }                                                                             \
    } while (0)
    for (uint32_t k = 0; k < o->nisp; ++k) JOIN(o->isp[2 * k], o->isp[2 * k + 1]);
    for (uint32_t k = 0; k < o->njoints; ++k) JOIN(o->joints[k].a, o->joints[k].b);
    for (uint32_t k = 0; k < o->ngj; ++k)
        if (!o->gj[k].broken) JOIN(o->gj[k].a, o->gj[k].b);
#undef JOIN
    for (uint32_t i = 0; i < n; ++i) {
        obody* b = &o->B[i];
        if (!awake_dyn(b)) continue;
        if (norm(b->v) < o->P.sleep_lin_vel && norm(b->w) < o->P.sleep_ang_vel) b->sleepTimer += o->P.dt;
        else b->sleepTimer = 0.0;
    }
    double* mint = (double*)malloc(sizeof(double) * (n ? n : 1));
    unsigned char* seen = (unsigned char*)calloc(n ? n : 1, 1);
    for (uint32_t i = 0; i < n; ++i) {
        if (!awake_dyn(&o->B[i])) continue;
        const uint32_t r = uf_find(parent, i);
        if (!seen[r] || o->B[i].sleepTimer < mint[r]) { mint[r] = o->B[i].sleepTimer; seen[r] = 1; }
    }
    for (uint32_t i = 0; i < n; ++i) {
        obody* b = &o->B[i];
        if (!awake_dyn(b)) continue;
        if (mint[uf_find(parent, i)] < o->P.sleep_time) continue;
        b->invMassStore = b->invMass;
        b->invIStore = b->invI;
        b->invMass = 0.0;
        b->invI = mk(0.0, 0.0, 0.0);
        b->v = mk(0.0, 0.0, 0.0);
        b->w = mk(0.0, 0.0, 0.0);
        b->flags |= NE_FLAG_SLEEPING;
    }
    free(parent);
    free(mint);
    free(seen);
}

/* ---- general joints: World::solveJoints (world.cpp:206-315) over Mat3
 * (vmath.hpp:81-123), m[row*3 + col] ---------------------------------------- */
typedef struct { double m[9]; } m3;

static m3 m3_zero(void) { m3 r; for (int i = 0; i < 9; ++i) r.m[i] = 0.0; return r; }   /* vmath.hpp:84 */
static m3 m3_scaled(double s) { m3 r = m3_zero(); r.m[0] = r.m[4] = r.m[8] = s; return r; } /* vmath.hpp:86 */
static m3 m3_make(double a, double b, double c, double d, double e, double f, double g, double h, double i) {
    m3 r;
    r.m[0] = a; r.m[1] = b; r.m[2] = c; r.m[3] = d; r.m[4] = e; r.m[5] = f; r.m[6] = g; r.m[7] = h; r.m[8] = i;
    return r;
}
static m3 m3_skew(v3 v) { return m3_make(0, -v.z, v.y, v.z, 0, -v.x, -v.y, v.x, 0); }     /* vmath.hpp:89-91 */
static v3 m3_mulv(const m3* a, v3 v) {                                                  /* vmath.hpp:96-100 */
    return mk(a->m[0] * v.x + a->m[1] * v.y + a->m[2] * v.z, a->m[3] * v.x + a->m[4] * v.y + a->m[5] * v.z,
              a->m[6] * v.x + a->m[7] * v.y + a->m[8] * v.z);
}
static m3 m3_mul(const m3* a, const m3* o) {                                            /* vmath.hpp:101-107 */
    m3 r = m3_zero();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) r.m[i * 3 + j] += a->m[i * 3 + k] * o->m[k * 3 + j];
    return r;
}
static m3 m3_add(const m3* a, const m3* o) { m3 r; for (int i = 0; i < 9; ++i) r.m[i] = a->m[i] + o->m[i]; return r; }
static m3 m3_sub(const m3* a, const m3* o) { m3 r; for (int i = 0; i < 9; ++i) r.m[i] = a->m[i] - o->m[i]; return r; }
static m3 m3_transposed(const m3* a) {                                                  /* vmath.hpp:110 */
    return m3_make(a->m[0], a->m[3], a->m[6], a->m[1], a->m[4], a->m[7], a->m[2], a->m[5], a->m[8]);
}
static m3 m3_inverse(const m3* M) {                                                     /* vmath.hpp:112-121 */
    const double a = M->m[0], b = M->m[1], c = M->m[2], d = M->m[3], e = M->m[4], f = M->m[5], g = M->m[6],
                 h = M->m[7], i = M->m[8];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    if (fabs(det) < 1e-18) return m3_zero();
    double inv = 1.0 / det;
    return m3_make(A * inv, (c * h - b * i) * inv, (b * f - c * e) * inv, B * inv, (a * i - c * g) * inv,
                   (c * d - a * f) * inv, C * inv, (b * g - a * h) * inv, (a * e - b * d) * inv);
}
static v3 m3_solve(const m3* M, v3 rhs) { m3 inv = m3_inverse(M); return m3_mulv(&inv, rhs); } /* vmath.hpp:123 */

static m3 inv_inertia_world(const obody* b) {                                           /* body.hpp:184-192 */
    v3 cx = qrot(b->q, mk(1, 0, 0)), cy = qrot(b->q, mk(0, 1, 0)), cz = qrot(b->q, mk(0, 0, 1));
    m3 R = m3_make(cx.x, cy.x, cz.x, cx.y, cy.y, cz.y, cx.z, cy.z, cz.z);
    m3 D = m3_zero();
    D.m[0] = b->invI.x; D.m[4] = b->invI.y; D.m[8] = b->invI.z;
    m3 RD = m3_mul(&R, &D), Rt = m3_transposed(&R);
    return m3_mul(&RD, &Rt);
}

static v3 perpendicular(v3 v) {                                                        /* world.cpp:200-203 */
    v3 a = fabs(v.x) < 0.9 ? mk(1, 0, 0) : mk(0, 1, 0);
    return normalized(cross(v, a));
}
static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); } /* world.cpp:204 */

/* K = (mA + mB) I - [rA]x IA [rA]x - [rB]x IB [rB]x (world.cpp:232-234, 248-250) */
static m3 point_mass_matrix(const obody* A, const obody* B, const m3* IA, const m3* IB, v3 rA, v3 rB) {
    m3 s = m3_scaled(A->invMass + B->invMass);
    m3 sa = m3_skew(rA), sb = m3_skew(rB);
    m3 t = m3_mul(&sa, IA);
    m3 ta = m3_mul(&t, &sa);
    t = m3_mul(&sb, IB);
    m3 tb = m3_mul(&t, &sb);
    m3 k1 = m3_sub(&s, &ta);
    return m3_sub(&k1, &tb);
}

static void solve_joints(ne_oracle* o, int iterations) {                                /* world.cpp:206-315 */
    const double beta = 0.2, dt = o->P.dt;
    for (uint32_t k = 0; k < o->ngj; ++k) { o->gj_applied[k] = 0.0; o->gj_motor[k] = 0.0; }
    for (int it = 0; it < iterations; ++it) {
        for (uint32_t k = 0; k < o->ngj; ++k) {
            ne_joint* j = &o->gj[k];
            if (j->broken) continue;
            obody* A = &o->B[j->a];
            obody* B = &o->B[j->b];
            const m3 IA = inv_inertia_world(A), IB = inv_inertia_world(B);
            v3 rA = qrot(A->q, ld3(j->local_a, 0)), rB = qrot(B->q, ld3(j->local_b, 0));
            v3 posErr = min_image(&o->box, sub(add(A->x, rA), add(B->x, rB)));
            v3 vRel = sub(add(A->v, cross(A->w, rA)), add(B->v, cross(B->w, rB)));
            if (j->type == NE_JOINT_SLIDER) {
                v3 ax = normalized(qrot(A->q, ld3(j->axis_a, 0)));
                v3 t1 = perpendicular(ax), t2 = normalized(cross(ax, t1));
                m3 K = point_mass_matrix(A, B, &IA, &IB, rA, rB);
                const v3 ts[2] = {t1, t2};
                for (int q = 0; q < 2; ++q) {
                    const v3 t = ts[q];
                    double jv = dot(vRel, t) + dot(posErr, t) * (beta / dt);
                    double kk = dot(t, m3_mulv(&K, t));
                    if (kk > 1e-18) {
                        double l = -jv / kk;
                        v3 P = scl(t, l);
                        A->v = add(A->v, scl(P, A->invMass)); A->w = add(A->w, m3_mulv(&IA, cross(rA, P)));
                        B->v = sub(B->v, scl(P, B->invMass)); B->w = sub(B->w, m3_mulv(&IB, cross(rB, P)));
                        o->gj_applied[k] += fabs(l);
                        vRel = sub(add(A->v, cross(A->w, rA)), add(B->v, cross(B->w, rB)));
                    }
                }
            } else {
                m3 K = point_mass_matrix(A, B, &IA, &IB, rA, rB);
                v3 P = m3_solve(&K, neg(add(vRel, scl(posErr, beta / dt))));
                A->v = add(A->v, scl(P, A->invMass)); A->w = add(A->w, m3_mulv(&IA, cross(rA, P)));
                B->v = sub(B->v, scl(P, B->invMass)); B->w = sub(B->w, m3_mulv(&IB, cross(rB, P)));
                o->gj_applied[k] += norm(P);
            }
            if (j->type == NE_JOINT_FIXED || j->type == NE_JOINT_SLIDER) {
                v3 wRel = sub(A->w, B->w);
                qt dq = qmul(qmul(A->q, qconj(B->q)), qconj(mq(j->ref_rel[0], j->ref_rel[1], j->ref_rel[2], j->ref_rel[3])));
                v3 angErr = mk(dq.x, dq.y, dq.z);
                if (dq.w < 0.0) angErr = neg(angErr);
                angErr = scl(angErr, 2.0);
                m3 Kang = m3_add(&IA, &IB);
                v3 L = m3_solve(&Kang, neg(add(wRel, scl(angErr, beta / dt))));
                A->w = add(A->w, m3_mulv(&IA, L));
                B->w = sub(B->w, m3_mulv(&IB, L));
            } else if (j->type == NE_JOINT_HINGE) {
                v3 aA = normalized(qrot(A->q, ld3(j->axis_a, 0)));
                v3 aB = normalized(qrot(B->q, ld3(j->axis_b, 0)));
                v3 t1 = perpendicular(aA), t2 = normalized(cross(aA, t1));
                v3 alignErr = cross(aA, aB);
                m3 Kang = m3_add(&IA, &IB);
                const v3 ts[2] = {t1, t2};
                for (int q = 0; q < 2; ++q) {
                    const v3 t = ts[q];
                    v3 wRel = sub(A->w, B->w);
                    double jv = dot(wRel, t) - dot(alignErr, t) * (beta / dt);
                    double kk = dot(t, m3_mulv(&Kang, t));
                    if (kk > 1e-18) {
                        v3 L = scl(t, -jv / kk);
                        A->w = add(A->w, m3_mulv(&IA, L));
                        B->w = sub(B->w, m3_mulv(&IB, L));
                    }
                }
                if (j->use_motor) {
                    double kk = dot(aA, m3_mulv(&Kang, aA));
                    if (kk > 1e-18) {
                        double jv = dot(sub(A->w, B->w), aA) - j->motor_speed;
                        double l = -jv / kk;
                        double old = o->gj_motor[k];
                        double nn = clampd(old + l, -j->max_motor_impulse, j->max_motor_impulse);
                        l = nn - old;
                        o->gj_motor[k] = nn;
                        v3 L = scl(aA, l);
                        A->w = add(A->w, m3_mulv(&IA, L));
                        B->w = sub(B->w, m3_mulv(&IB, L));
                    }
                }
                if (j->use_limit) {
                    qt dq = qmul(qmul(A->q, qconj(B->q)), qconj(mq(j->ref_rel[0], j->ref_rel[1], j->ref_rel[2], j->ref_rel[3])));
                    v3 dv = mk(dq.x, dq.y, dq.z);
                    double angle = 2.0 * atan2(dot(dv, aA), dq.w);
                    double c = 0.0;
                    int active = 0;
                    if (angle <= j->lower) { c = angle - j->lower; active = 1; }
                    else if (angle >= j->upper) { c = angle - j->upper; active = 1; }
                    if (active) {
                        double kk = dot(aA, m3_mulv(&Kang, aA));
                        if (kk > 1e-18) {
                            double jv = dot(sub(A->w, B->w), aA) + c * (beta / dt);
                            v3 L = scl(aA, -jv / kk);
                            A->w = add(A->w, m3_mulv(&IA, L));
                            B->w = sub(B->w, m3_mulv(&IB, L));
                        }
                    }
                }
            }
        }
    }
    for (uint32_t k = 0; k < o->ngj; ++k) {
        ne_joint* j = &o->gj[k];
        if (j->breakable && !j->broken && o->gj_applied[k] > j->break_impulse) j->broken = 1;
    }
}

static int ray_vs_body(int shape, qt q, v3 he, double r, double h, v3 c, v3 o, v3 d, double maxT,
                       double* tOut, v3* pOut, v3* nOut);

/* World::ccdPass (world.cpp:658-696): each fast CCD body in index order
 * sweeps its bounding sphere from its step-start position along this step's
 * displacement against every other non-sensor body it collides with
 * (inflated by the sweep radius, at the minimum image), stops at the first
 * impact and loses its velocity into the surface.  Later bodies see the
 * earlier ones' stopped positions. */
static void ccd_pass(ne_oracle* o) {
    for (uint32_t i = 0; i < o->n; ++i) {
        obody* bi = &o->B[i];
        if (!(bi->flags & NE_FLAG_CCD) || bi->invMass <= 0.0 || (bi->flags & NE_FLAG_SLEEPING)) continue;
        v3 disp = sub(bi->x, bi->prevX);
        double dist = norm(disp);
        double R = bounding_radius(o, bi);
        if (dist < R * 0.5 || dist < 1e-9) continue;
        v3 dir = dvs(disp, dist);
        double bestT = dist;
        v3 bestN = mk(0.0, 0.0, 0.0);
        int hit = 0;
        for (uint32_t j = 0; j < o->n; ++j) {
            if (j == i) continue;
            const obody* bj = &o->B[j];
            if (bj->flags & NE_FLAG_SENSOR) continue;
            if (!layers_collide(bi, bj)) continue;
            double r = bj->radius + R, h = bj->hh;
            v3 he = bj->he;
            if (bj->shape == NE_SHAPE_BOX) he = add(bj->he, mk(R, R, R));
            else if (bj->shape == NE_SHAPE_CYLINDER) h = bj->hh + R;
            v3 cImg = add(bi->prevX, min_image(&o->box, sub(bj->x, bi->prevX)));
            double t;
            v3 p, nn;
            if (ray_vs_body(bj->shape, bj->q, he, r, h, cImg, bi->prevX, dir, bestT, &t, &p, &nn) && t < bestT) {
                bestT = t;
                bestN = nn;
                hit = 1;
            }
        }
        if (hit) {
            bi->x = add(bi->prevX, scl(dir, bestT));
            double vn = dot(bi->v, bestN);
            if (vn < 0.0) bi->v = sub(bi->v, scl(bestN, vn));
        }
    }
}

static void world_step(ne_oracle* o) {                                               /* world.cpp:133-146 */
    apply_joint_forces(o);
    for (uint32_t i = 0; i < o->n; ++i) o->B[i].prevX = o->B[i].x;
    integrate(o);
    if (o->nccd) ccd_pass(o);
    collide(o);
    solve_rigid_joints(o, 8);
    if (o->ngj) solve_joints(o, 8);
    if (o->box.periodic) wrap_positions(o);
    else if (!o->P.open_boundary) apply_walls(o);
    if (o->P.sleep_enabled) update_sleep(o);
}

/* PhysicsWorld::processContacts (physics_world.cpp:253-302), step 1: every
 * pair i < j of the post-step state that passes the facade's filters and
 * overlaps under detectContact (one contact, no box-box manifold), split by
 * whether a sensor is in it.  The diff against the last step is done when
 * the lists are read. */
static void process_contacts(ne_oracle* o) {
    const int cur = o->ev_cur ^ 1;
    o->nev[cur] = 0;
    for (uint32_t i = 0; i < o->n; ++i) {
        for (uint32_t j = i + 1; j < o->n; ++j) {
            const obody* a = &o->B[i];
            const obody* b = &o->B[j];
            const int any_sensor = ((a->flags | b->flags) & NE_FLAG_SENSOR) != 0;
            if (!any_sensor && a->invMass + b->invMass <= 0.0) continue;
            if (!layers_collide(a, b)) continue;
            if (pair_excluded(o, a, j)) continue;
            const ocontact c = detect_contact(o, a, b);
            if (!c.hit) continue;
            if (o->nev[cur] == o->capev[cur]) {
                o->capev[cur] = o->capev[cur] ? 2 * o->capev[cur] : 256;
                o->ev[cur] = (struct oevent*)realloc(o->ev[cur], sizeof(struct oevent) * o->capev[cur]);
            }
            struct oevent* e = &o->ev[cur][o->nev[cur]++];
            e->i = i; e->j = j; e->trigger = any_sensor; e->c = c;
        }
    }
    o->ev_cur = cur;
}

/* Whether list l holds pair (i, j) with the same trigger flag (the lists are
 * in (i, j) order: binary search). */
static int ev_has(const ne_oracle* o, int l, uint32_t i, uint32_t j, int trigger) {
    uint32_t lo = 0, hi = o->nev[l];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        const struct oevent* e = &o->ev[l][mid];
        if (e->i < i || (e->i == i && e->j < j)) lo = mid + 1;
        else hi = mid;
    }
    return lo < o->nev[l] && o->ev[l][lo].i == i && o->ev[l][lo].j == j && o->ev[l][lo].trigger == trigger;
}

/* ------------------------------------------------------------------------ */
/* C API                                                                     */
/* ------------------------------------------------------------------------ */

int ne_oracle_create(ne_oracle** out) {
    if (!out) return NE_ERR_INVALID;
    ne_oracle* o = (ne_oracle*)calloc(1, sizeof(ne_oracle));
    if (!o) return NE_ERR_INVALID;
    /* World defaults: src/world.hpp:75-126, src/box.hpp:17-18 */
    o->P.half[0] = o->P.half[1] = o->P.half[2] = 12.0;
    o->P.dt = 1.0 / 120.0;
    o->P.restitution = 1.0;
    o->P.contact_beta = 0.2;
    o->P.restitution_combine = NE_COMBINE_MAX;
    o->P.friction_combine = NE_COMBINE_GEOMETRIC_MEAN;
    o->P.warm_start = 1;
    o->P.solver_iterations = 8;
    o->P.sleep_lin_vel = 0.05; o->P.sleep_ang_vel = 0.05; o->P.sleep_time = 0.5; o->P.wake_vel = 0.5;
    o->box.half[0] = o->box.half[1] = o->box.half[2] = 12.0;
    *out = o;
    return NE_OK;
}

static void free_meshes(ne_oracle* o) {
    for (uint32_t m = 0; m < o->nmeshes; ++m) {
        free(o->meshes[m].verts); free(o->meshes[m].tris); free(o->meshes[m].order); free(o->meshes[m].nodes);
    }
    free(o->meshes);
    o->meshes = NULL;
    o->nmeshes = 0;
    for (uint32_t i = 0; i < o->n; ++i) o->B[i].mesh = -1;
}

int ne_oracle_destroy(ne_oracle* o) {
    if (!o) return NE_OK;
    free_meshes(o);
    free(o->cs);
    free(o->B); free(o->hull); free(o->nocol); free(o->cons); free(o->cache_key); free(o->cache_imp);
    for (uint32_t t = 0; t < o->tape_steps; ++t) { free(o->tape[t].p); free(o->tape[t].rest); free(o->tape[t].visit); }
    free(o->tape);
    free(o->joints);
    free(o->isp);
    free(o->ev[0]); free(o->ev[1]);
    free(o->gj); free(o->gj_applied); free(o->gj_motor);
    free(o);
    return NE_OK;
}

int ne_oracle_set_params(ne_oracle* o, const ne_world_params* p) {
    if (!o || !p) return NE_ERR_INVALID;
    if (p->solver_iterations < 0) return NE_ERR_INVALID;
    o->P = *p;
    for (int a = 0; a < 3; ++a) o->box.half[a] = p->half[a];
    o->box.periodic = p->periodic != 0;
    return NE_OK;
}

int ne_oracle_set_bodies(ne_oracle* o, const ne_bodies* b) {
    if (!o || !b) return NE_ERR_INVALID;
    o->nev[0] = o->nev[1] = 0;   /* a new world: no last-step pairs */
    o->nccd = 0;
    o->ngj = 0;   /* and no general joints */
    free_meshes(o);
    free(o->B); free(o->hull); free(o->nocol);
    o->B = NULL; o->hull = NULL; o->nocol = NULL; o->nhull = 0; o->nnocol = 0;
    o->n = b->n;
    o->B = (obody*)calloc(b->n ? b->n : 1, sizeof(obody));
    if (b->hull_offsets) {
        o->nhull = b->hull_offsets[b->n];
        o->hull = (v3*)malloc(sizeof(v3) * (o->nhull ? o->nhull : 1));
        for (uint32_t k = 0; k < o->nhull; ++k) o->hull[k] = ld3(b->hull_verts, k);
    }
    if (b->nocollide_offsets) {
        o->nnocol = b->nocollide_offsets[b->n];
        o->nocol = (uint32_t*)malloc(sizeof(uint32_t) * (o->nnocol ? o->nnocol : 1));
        memcpy(o->nocol, b->nocollide, sizeof(uint32_t) * o->nnocol);
    }
    for (uint32_t i = 0; i < b->n; ++i) {
        obody* B = &o->B[i];
        B->x = ld3(b->x, i); B->v = ld3(b->v, i); B->w = ld3(b->w, i);
        B->q = mq(b->q[4 * i], b->q[4 * i + 1], b->q[4 * i + 2], b->q[4 * i + 3]);
        B->force = b->force ? ld3(b->force, i) : mk(0, 0, 0);
        B->torque = b->torque ? ld3(b->torque, i) : mk(0, 0, 0);
        B->invMass = b->inv_mass[i];
        B->invI = ld3(b->inv_inertia, i);
        B->shape = b->shape[i];
        if (B->shape < 0 || B->shape > NE_SHAPE_MESH) return NE_ERR_UNSUPPORTED;
        B->mesh = -1;
        B->radius = b->radius[i];
        B->hh = b->half_height[i];
        B->he = ld3(b->half_extents, i);
        B->friction = b->friction ? b->friction[i] : -1.0;
        B->restitution = b->restitution ? b->restitution[i] : -1.0;
        B->linDamp = b->linear_damping ? b->linear_damping[i] : 0.0;
        B->angDamp = b->angular_damping ? b->angular_damping[i] : 0.0;
        B->gravityScale = b->gravity_scale ? b->gravity_scale[i] : 1.0;
        B->layer = b->layer ? b->layer[i] : 0u;
        B->mask = b->mask ? b->mask[i] : 0xFFFFFFFFu;
        B->flags = b->flags ? b->flags[i] : NE_FLAG_DYNAMIC;
        if (B->flags & NE_FLAG_CCD) ++o->nccd;
        B->sleepTimer = b->sleep_timer ? b->sleep_timer[i] : 0.0;
        B->invMassStore = b->inv_mass_store ? b->inv_mass_store[i] : 0.0;
        B->invIStore = b->inv_inertia_store ? ld3(b->inv_inertia_store, i) : mk(0, 0, 0);
        B->hull0 = b->hull_offsets ? b->hull_offsets[i] : 0;
        B->hull1 = b->hull_offsets ? b->hull_offsets[i + 1] : 0;
        B->nc0 = b->nocollide_offsets ? b->nocollide_offsets[i] : 0;
        B->nc1 = b->nocollide_offsets ? b->nocollide_offsets[i + 1] : 0;
    }
    return NE_OK;
}

int ne_oracle_set_state(ne_oracle* o, const double* x, const double* v, const double* q, const double* w) {
    if (!o) return NE_ERR_INVALID;
    for (uint32_t i = 0; i < o->n; ++i) {
        if (x) o->B[i].x = ld3(x, i);
        if (v) o->B[i].v = ld3(v, i);
        if (q) o->B[i].q = mq(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]);
        if (w) o->B[i].w = ld3(w, i);
    }
    return NE_OK;
}

int ne_oracle_step(ne_oracle* o, int32_t nsteps) {
    if (!o) return NE_ERR_INVALID;
    for (int32_t s = 0; s < nsteps; ++s) {
        world_step(o);
        if (o->events_on) process_contacts(o);
    }
    return NE_OK;
}

int ne_oracle_set_events(ne_oracle* o, int32_t enable) {
    if (!o) return NE_ERR_INVALID;
    o->events_on = enable != 0;
    o->nev[0] = o->nev[1] = 0;   /* the facade's maps start empty */
    return NE_OK;
}

/* The step's event lists (physics_world.cpp:277-299): contact = every solid
 * pair; begin / stay = solid pairs absent from / present in the last step's
 * solid list; end = the last step's solid pairs absent now, with their last
 * contact; the trigger kinds likewise over the sensor pairs.  In (i, j)
 * order, which is the reference's map order when handles follow indices. */
int ne_oracle_get_events(ne_oracle* o, int32_t kind, uint32_t cap, uint32_t* count, uint32_t* ci, uint32_t* cj,
                         double* point, double* normal, double* depth) {
    if (!o || !count || kind < NE_EVENT_CONTACT || kind > NE_EVENT_TRIGGER_EXIT) return NE_ERR_INVALID;
    const int cur = o->ev_cur, prev = cur ^ 1;
    const int trig = kind >= NE_EVENT_TRIGGER_ENTER;
    const int from_prev = (kind == NE_EVENT_END || kind == NE_EVENT_TRIGGER_EXIT);
    const int l = from_prev ? prev : cur;
    uint32_t m = 0;
    for (uint32_t k = 0; k < o->nev[l]; ++k) {
        const struct oevent* e = &o->ev[l][k];
        if (e->trigger != trig) continue;
        int take;
        if (from_prev) take = !ev_has(o, cur, e->i, e->j, trig);
        else if (kind == NE_EVENT_CONTACT) take = 1;
        else {
            const int was = ev_has(o, prev, e->i, e->j, trig);
            take = (kind == NE_EVENT_BEGIN || kind == NE_EVENT_TRIGGER_ENTER) ? !was : was;
        }
        if (!take) continue;
        if (m < cap) {
            if (ci) ci[m] = e->i;
            if (cj) cj[m] = e->j;
            if (point) { point[3 * m] = e->c.point.x; point[3 * m + 1] = e->c.point.y; point[3 * m + 2] = e->c.point.z; }
            if (normal) { normal[3 * m] = e->c.normal.x; normal[3 * m + 1] = e->c.normal.y; normal[3 * m + 2] = e->c.normal.z; }
            if (depth) depth[m] = e->c.overlap;
        }
        ++m;
    }
    *count = m;
    return NE_OK;
}

/* Contact-set tape (test infrastructure for the adjoint): mode 1 clears the
 * tape and records every following step's contact lists; mode 2 replays the
 * tape from its first step, using the recorded contacts (pairs, normals,
 * points, depths) instead of broadphase and narrow phase, with the rest of
 * the step unchanged; mode 0 steps normally.  Replaying at the recorded
 * inputs reproduces the recorded run exactly; at perturbed inputs it is the
 * same step with the contact set held fixed.  Mode 3 replays as mode 2 and
 * also holds the solver's branches fixed at the recorded run's sides: the
 * restitution bias (vn < -0.5), each visit's normal clamp, whether friction
 * is visited, |vt| > 1e-12, and the friction cone's clamp side
 * (world.cpp:544, 572-612).  With the active set frozen, the rollout is one
 * smooth function near the recorded inputs, whose derivative is the one an
 * adjoint through the recorded branches computes. */
int ne_oracle_tape(ne_oracle* o, int32_t mode) {
    if (!o || mode < 0 || mode > 3) return NE_ERR_INVALID;
    if (mode == 1) {
        for (uint32_t t = 0; t < o->tape_steps; ++t) { free(o->tape[t].p); free(o->tape[t].rest); free(o->tape[t].visit); }
        o->tape_steps = 0;
    }
    o->tape_cursor = 0;
    o->tape_mode = mode;
    return NE_OK;
}

int ne_oracle_set_joints(ne_oracle* o, uint32_t n, const ne_distance_joint* j) {
    if (!o || (n && !j)) return NE_ERR_INVALID;
    for (uint32_t k = 0; k < n; ++k)
        if (j[k].a >= o->n || j[k].b >= o->n) return NE_ERR_INVALID;
    free(o->joints);
    o->joints = (ne_distance_joint*)malloc(sizeof(ne_distance_joint) * (n ? n : 1));
    if (n) memcpy(o->joints, j, sizeof(ne_distance_joint) * n);
    o->njoints = n;
    return NE_OK;
}

/* Builds each mesh's tree (MeshData::build, mesh.hpp:38-43) and attaches it
 * to its body (Body::mesh). */
int ne_oracle_set_meshes(ne_oracle* o, uint32_t n, const ne_mesh* ms) {
    if (!o || (n && !ms)) return NE_ERR_INVALID;
    for (uint32_t m = 0; m < n; ++m) {
        if (ms[m].body >= o->n || o->B[ms[m].body].shape != NE_SHAPE_MESH) return NE_ERR_INVALID;
        if (ms[m].ntris && (!ms[m].tris || !ms[m].verts)) return NE_ERR_INVALID;
        for (uint64_t k = 0; k < 3ull * ms[m].ntris; ++k)
            if (ms[m].tris[k] < 0 || (uint32_t)ms[m].tris[k] >= ms[m].nverts) return NE_ERR_INVALID;
    }
    free_meshes(o);
    o->meshes = (omesh*)calloc(n ? n : 1, sizeof(omesh));
    o->nmeshes = n;
    for (uint32_t m = 0; m < n; ++m) {
        omesh* M = &o->meshes[m];
        const ne_mesh* src = &ms[m];
        M->nv = src->nverts;
        M->nt = src->ntris;
        M->verts = (v3*)malloc(sizeof(v3) * (M->nv ? M->nv : 1));
        for (uint32_t k = 0; k < M->nv; ++k) M->verts[k] = ld3(src->verts, k);
        M->tris = (int32_t*)malloc(sizeof(int32_t) * 3 * (M->nt ? M->nt : 1));
        if (M->nt) memcpy(M->tris, src->tris, sizeof(int32_t) * 3 * M->nt);
        M->order = (int32_t*)malloc(sizeof(int32_t) * (M->nt ? M->nt : 1));
        const int32_t cap = 2 * (int32_t)M->nt + 1;
        double* bounds = (double*)malloc(sizeof(double) * 6 * (size_t)cap);
        int32_t* links = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)cap);
        int32_t nn = ne_oracle_bvh_build(src->verts, src->tris, (int32_t)M->nt, M->order, cap, bounds, links);
        if (nn < 0 || nn > cap) { free(bounds); free(links); return NE_ERR_INVALID; }
        M->nn = (uint32_t)nn;
        M->nodes = (onode*)malloc(sizeof(onode) * (nn ? (size_t)nn : 1));
        for (int32_t k = 0; k < nn; ++k) {
            onode* d = &M->nodes[k];
            d->lo = mk(bounds[6 * k], bounds[6 * k + 1], bounds[6 * k + 2]);
            d->hi = mk(bounds[6 * k + 3], bounds[6 * k + 4], bounds[6 * k + 5]);
            d->left = links[4 * k]; d->right = links[4 * k + 1];
            d->start = links[4 * k + 2]; d->count = links[4 * k + 3];
        }
        free(bounds);
        free(links);
        o->B[src->body].mesh = (int32_t)m;
    }
    return NE_OK;
}

int ne_oracle_mesh_bvh(ne_oracle* o, uint32_t m, int32_t* order, uint32_t node_cap, uint32_t* nnodes, double* bounds,
                       int32_t* links) {
    if (!o || m >= o->nmeshes) return NE_ERR_INVALID;
    const omesh* M = &o->meshes[m];
    if (nnodes) *nnodes = M->nn;
    if (order) memcpy(order, M->order, sizeof(int32_t) * M->nt);
    if (bounds && links && node_cap >= M->nn)
        for (uint32_t k = 0; k < M->nn; ++k) {
            const onode* d = &M->nodes[k];
            bounds[6 * k] = d->lo.x; bounds[6 * k + 1] = d->lo.y; bounds[6 * k + 2] = d->lo.z;
            bounds[6 * k + 3] = d->hi.x; bounds[6 * k + 4] = d->hi.y; bounds[6 * k + 5] = d->hi.z;
            links[4 * k] = d->left; links[4 * k + 1] = d->right; links[4 * k + 2] = d->start; links[4 * k + 3] = d->count;
        }
    return NE_OK;
}

int ne_oracle_reset_warm_cache(ne_oracle* o) {
    if (!o) return NE_ERR_INVALID;
    o->ncache = 0;
    return NE_OK;
}

int ne_oracle_get_state(ne_oracle* o, double* x, double* v, double* q, double* w) {
    if (!o) return NE_ERR_INVALID;
    for (uint32_t i = 0; i < o->n; ++i) {
        const obody* B = &o->B[i];
        if (x) { x[3 * i] = B->x.x; x[3 * i + 1] = B->x.y; x[3 * i + 2] = B->x.z; }
        if (v) { v[3 * i] = B->v.x; v[3 * i + 1] = B->v.y; v[3 * i + 2] = B->v.z; }
        if (q) { q[4 * i] = B->q.w; q[4 * i + 1] = B->q.x; q[4 * i + 2] = B->q.y; q[4 * i + 3] = B->q.z; }
        if (w) { w[3 * i] = B->w.x; w[3 * i + 1] = B->w.y; w[3 * i + 2] = B->w.z; }
    }
    return NE_OK;
}

int ne_oracle_get_body_aux(ne_oracle* o, uint32_t* flags, double* sleep_timer, double* inv_mass, double* inv_inertia) {
    if (!o) return NE_ERR_INVALID;
    for (uint32_t i = 0; i < o->n; ++i) {
        const obody* B = &o->B[i];
        if (flags) flags[i] = B->flags;
        if (sleep_timer) sleep_timer[i] = B->sleepTimer;
        if (inv_mass) inv_mass[i] = B->invMass;
        if (inv_inertia) { inv_inertia[3 * i] = B->invI.x; inv_inertia[3 * i + 1] = B->invI.y; inv_inertia[3 * i + 2] = B->invI.z; }
    }
    return NE_OK;
}

int ne_oracle_get_contacts(ne_oracle* o, uint32_t cap, uint32_t* count, uint32_t* ci, uint32_t* cj,
                           double* normal, double* point, double* depth, double* nimp) {
    if (!o || !count) return NE_ERR_INVALID;
    *count = o->ncons;
    uint32_t m = o->ncons < cap ? o->ncons : cap;
    for (uint32_t k = 0; k < m; ++k) {
        const struct ocons* c = &o->cons[k];
        if (ci) ci[k] = c->i;
        if (cj) cj[k] = c->j;
        if (normal) { normal[3 * k] = c->n.x; normal[3 * k + 1] = c->n.y; normal[3 * k + 2] = c->n.z; }
        if (point) { point[3 * k] = c->point.x; point[3 * k + 1] = c->point.y; point[3 * k + 2] = c->point.z; }
        if (depth) depth[k] = c->overlap;
        if (nimp) nimp[k] = c->nImp;
    }
    return NE_OK;
}

int ne_oracle_get_pairs(ne_oracle* o, uint32_t cap, uint32_t* count, uint32_t* pi, uint32_t* pj) {
    if (!o || !count) return NE_ERR_INVALID;
    uint32_t m = 0;
    for (uint32_t k = 0; k < o->ncons; ++k) {
        if (k > 0 && o->cons[k].i == o->cons[k - 1].i && o->cons[k].j == o->cons[k - 1].j) continue;
        if (m < cap) { if (pi) pi[m] = o->cons[k].i; if (pj) pj[m] = o->cons[k].j; }
        ++m;
    }
    *count = m;
    return NE_OK;
}

int ne_oracle_pair_contacts(ne_oracle* o, uint32_t a, uint32_t b, uint32_t cap, uint32_t* count,
                            double* normal, double* point, double* depth) {
    if (!o || !count || a >= o->n || b >= o->n) return NE_ERR_INVALID;
    int nc = gather_contacts(o, &o->B[a], &o->B[b]);
    const ocontact* cs = o->cs;
    *count = (uint32_t)nc;
    uint32_t m = (uint32_t)nc < cap ? (uint32_t)nc : cap;
    for (uint32_t k = 0; k < m; ++k) {
        if (normal) { normal[3 * k] = cs[k].normal.x; normal[3 * k + 1] = cs[k].normal.y; normal[3 * k + 2] = cs[k].normal.z; }
        if (point) { point[3 * k] = cs[k].point.x; point[3 * k + 1] = cs[k].point.y; point[3 * k + 2] = cs[k].point.z; }
        if (depth) depth[k] = cs[k].overlap;
    }
    return NE_OK;
}

int ne_oracle_candidate_count(ne_oracle* o, uint32_t* count) {
    if (!o || !count) return NE_ERR_INVALID;
    *count = o->ncand;
    return NE_OK;
}

int ne_oracle_diagnostics(ne_oracle* o, double* p, double* L, double* ke) {          /* world.cpp:704-729 */
    if (!o) return NE_ERR_INVALID;
    v3 P = mk(0, 0, 0), LL = mk(0, 0, 0);
    double e = 0.0;
    for (uint32_t i = 0; i < o->n; ++i) {
        const obody* b = &o->B[i];
        if (b->invMass > 0.0) P = add(P, scl(b->v, 1.0 / b->invMass));
    }
    for (uint32_t i = 0; i < o->n; ++i) {
        const obody* b = &o->B[i];
        if (b->invMass <= 0.0) continue;
        double m = 1.0 / b->invMass;
        LL = add(LL, cross(b->x, scl(b->v, m)));
        LL = add(LL, angular_momentum(b));
    }
    for (uint32_t i = 0; i < o->n; ++i)
        if (o->B[i].invMass > 0.0) e += kinetic(&o->B[i]);
    if (p) { p[0] = P.x; p[1] = P.y; p[2] = P.z; }
    if (L) { L[0] = LL.x; L[1] = LL.y; L[2] = LL.z; }
    if (ke) *ke = e;
    return NE_OK;
}

int ne_oracle_min_image(ne_oracle* o, const double* d, double* out) {
    if (!o) return NE_ERR_INVALID;
    v3 m = min_image(&o->box, mk(d[0], d[1], d[2]));
    out[0] = m.x; out[1] = m.y; out[2] = m.z;
    return NE_OK;
}

int ne_oracle_wrap(ne_oracle* o, const double* p, double* out) {
    if (!o) return NE_ERR_INVALID;
    v3 m = wrap_pos(&o->box, mk(p[0], p[1], p[2]));
    out[0] = m.x; out[1] = m.y; out[2] = m.z;
    return NE_OK;
}

/* ------------------------------------------------------------------------ */
/* Scene queries: PhysicsWorld::raycast / raycastAll / overlapSphere /      */
/* overlapBox / overlapContacts / sphereCast (physics_world.cpp:146-251)    */
/* over rayVsBody (queries.hpp:18-122).                                     */
/* ------------------------------------------------------------------------ */

static int layer_allowed(uint32_t mask, const obody* b) {                         /* physics_world.cpp:67 */
    return (mask & (1u << (b->layer & 31u))) != 0;
}

/* queries.hpp:18-31 */
static int ray_vs_sphere_at(v3 o, v3 d, v3 c, double r, double maxT, double* tOut, v3* nOut) {
    v3 m = sub(o, c);
    double b = dot(m, d), cc = dot(m, m) - r * r;
    if (cc > 0.0 && b > 0.0) return 0;
    double disc = b * b - cc;
    if (disc < 0.0) return 0;
    double t = -b - sqrt(disc);
    if (t < 0.0) t = 0.0;
    if (t > maxT) return 0;
    *tOut = t;
    *nOut = normalized(sub(add(o, scl(d, t)), c));
    return 1;
}

/* queries.hpp:35-59 */
static int ray_vs_box_local(v3 ol, v3 dl, v3 he, double maxT, double* tOut, v3* nLocal) {
    double tmin = 0.0, tmax = maxT;
    int axis = -1;
    double sign = 1.0;
    const double o[3] = {ol.x, ol.y, ol.z}, dd[3] = {dl.x, dl.y, dl.z}, h[3] = {he.x, he.y, he.z};
    for (int a = 0; a < 3; ++a) {
        if (fabs(dd[a]) < 1e-12) {
            if (o[a] < -h[a] || o[a] > h[a]) return 0;
        } else {
            double inv = 1.0 / dd[a];
            double t1 = (-h[a] - o[a]) * inv, t2 = (h[a] - o[a]) * inv;
            double s = -1.0;
            if (t1 > t2) { double tmp = t1; t1 = t2; t2 = tmp; s = 1.0; }
            if (t1 > tmin) { tmin = t1; axis = a; sign = s; }
            if (t2 < tmax) tmax = t2;
            if (tmin > tmax) return 0;
        }
    }
    if (axis < 0) return 0;
    *tOut = tmin;
    *nLocal = mk(axis == 0 ? sign : 0.0, axis == 1 ? sign : 0.0, axis == 2 ? sign : 0.0);
    return 1;
}

/* queries.hpp:61-122: a body of shape, orientation q, half extents he,
 * radius r and half height h, centred at c. */
static int ray_vs_body(int shape, qt q, v3 he, double r, double h, v3 c, v3 o, v3 d, double maxT,
                       double* tOut, v3* pOut, v3* nOut) {
    if (shape == NE_SHAPE_CONVEX || shape == NE_SHAPE_PLANE || shape == NE_SHAPE_MESH) return 0;
    if (shape == NE_SHAPE_SPHERE) {
        if (!ray_vs_sphere_at(o, d, c, r, maxT, tOut, nOut)) return 0;
        *pOut = add(o, scl(d, *tOut));
        return 1;
    }
    v3 ol = qinvrot(q, sub(o, c));
    v3 dl = qinvrot(q, d);
    if (shape == NE_SHAPE_BOX) {
        v3 nl;
        if (!ray_vs_box_local(ol, dl, he, maxT, tOut, &nl)) return 0;
        *nOut = qrot(q, nl);
        *pOut = add(o, scl(d, *tOut));
        return 1;
    }
    double best = maxT + 1.0;
    v3 nl = mk(0.0, 0.0, 0.0);
    double a2 = dl.x * dl.x + dl.z * dl.z;
    double b2 = 2.0 * (ol.x * dl.x + ol.z * dl.z);
    double c2 = ol.x * ol.x + ol.z * ol.z - r * r;
    if (a2 > 1e-18) {
        double disc = b2 * b2 - 4.0 * a2 * c2;
        if (disc >= 0.0) {
            double t = (-b2 - sqrt(disc)) / (2.0 * a2);
            if (t < 0.0) t = (-b2 + sqrt(disc)) / (2.0 * a2);
            if (t >= 0.0 && t <= maxT) {
                double y = ol.y + t * dl.y;
                if (y >= -h && y <= h && t < best) {
                    best = t;
                    nl = mk((ol.x + t * dl.x) / r, 0.0, (ol.z + t * dl.z) / r);
                }
            }
        }
    }
    if (shape == NE_SHAPE_CYLINDER) {
        for (int s = -1; s <= 1; s += 2) {
            double yc = s * h;
            if (fabs(dl.y) < 1e-12) continue;
            double t = (yc - ol.y) / dl.y;
            if (t < 0.0 || t > maxT || t >= best) continue;
            double px = ol.x + t * dl.x, pz = ol.z + t * dl.z;
            if (px * px + pz * pz <= r * r) { best = t; nl = mk(0.0, (double)s, 0.0); }
        }
    } else if (best > maxT) {   /* capsule: the cap spheres only when the side missed */
        for (int s = -1; s <= 1; s += 2) {
            double tS;
            v3 nS;
            if (ray_vs_sphere_at(ol, dl, mk(0.0, (double)s * h, 0.0), r, maxT, &tS, &nS) && tS < best) {
                best = tS;
                nl = nS;
            }
        }
    }
    if (best > maxT) return 0;
    *tOut = best;
    *nOut = normalized(qrot(q, nl));
    *pOut = add(o, scl(d, *tOut));
    return 1;
}

/* Body b as rayVsBody sees it, inflated by R for sphereCast
 * (physics_world.cpp:237-240). */
static int ray_vs_world_body(const ne_oracle* o, const obody* b, double R, int cast, v3 origin, v3 d, double maxT,
                             double* t, v3* p, v3* n) {
    double r = b->radius, h = b->hh;
    v3 he = b->he;
    if (cast) {
        r = b->radius + R;
        if (b->shape == NE_SHAPE_BOX) he = add(b->he, mk(R, R, R));
        else if (b->shape == NE_SHAPE_CYLINDER) h = b->hh + R;
    }
    v3 cImg = add(origin, min_image(&o->box, sub(b->x, origin)));
    return ray_vs_body(b->shape, b->q, he, r, h, cImg, origin, d, maxT, t, p, n);
}

int ne_oracle_raycast(ne_oracle* o, uint32_t n, const double* origin, const double* dir, const double* max_dist,
                      const uint32_t* mask, const double* cast_radius, int32_t* body, double* distance,
                      double* point, double* normal) {
    if (!o || (n && (!origin || !dir || !max_dist || !body || !distance || !point || !normal))) return NE_ERR_INVALID;
    for (uint32_t k = 0; k < n; ++k) {
        const v3 org = ld3(origin, k);
        const v3 d = normalized(ld3(dir, k));
        const uint32_t m = mask ? mask[k] : 0xFFFFFFFFu;
        const int cast = cast_radius != NULL;
        const double R = cast ? cast_radius[k] : 0.0;
        double best = max_dist[k];
        int32_t hb = -1;
        v3 hp = mk(0, 0, 0), hn = mk(0, 0, 0);
        double hd = 0.0;
        for (uint32_t i = 0; i < o->n; ++i) {
            const obody* b = &o->B[i];
            if (!layer_allowed(m, b)) continue;
            double t;
            v3 p, nn;
            if (ray_vs_world_body(o, b, R, cast, org, d, best, &t, &p, &nn) && t < best) {
                best = t;
                hb = (int32_t)i;
                hd = t;
                hn = nn;
                hp = cast ? sub(p, scl(nn, R)) : p;   /* physics_world.cpp:245 */
            }
        }
        body[k] = hb;
        distance[k] = hd;
        point[3 * k] = hp.x; point[3 * k + 1] = hp.y; point[3 * k + 2] = hp.z;
        normal[3 * k] = hn.x; normal[3 * k + 1] = hn.y; normal[3 * k + 2] = hn.z;
    }
    return NE_OK;
}

int ne_oracle_raycast_all(ne_oracle* o, uint32_t n, const double* origin, const double* dir, const double* max_dist,
                          const uint32_t* mask, uint32_t cap, uint32_t* offsets, int32_t* body, double* distance,
                          double* point, double* normal) {
    if (!o || !offsets || (n && (!origin || !dir || !max_dist))) return NE_ERR_INVALID;
    uint32_t total = 0;
    offsets[0] = 0;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && (total > cap || (total && (!body || !distance || !point || !normal))))
            return total > cap ? NE_ERR_CAPACITY : NE_ERR_INVALID;
        for (uint32_t k = 0; k < n; ++k) {
            const v3 org = ld3(origin, k);
            const v3 d = normalized(ld3(dir, k));
            const uint32_t m = mask ? mask[k] : 0xFFFFFFFFu;
            uint32_t at = pass ? offsets[k] : 0, cnt = 0;
            for (uint32_t i = 0; i < o->n; ++i) {
                const obody* b = &o->B[i];
                if (!layer_allowed(m, b)) continue;
                double t;
                v3 p, nn;
                if (!ray_vs_world_body(o, b, 0.0, 0, org, d, max_dist[k], &t, &p, &nn)) continue;
                if (pass) {
                    /* stable insertion by distance (equal distances keep body order) */
                    uint32_t s = at + cnt;
                    while (s > at && distance[s - 1] > t) {
                        body[s] = body[s - 1];
                        distance[s] = distance[s - 1];
                        for (int c = 0; c < 3; ++c) {
                            point[3 * s + c] = point[3 * (s - 1) + c];
                            normal[3 * s + c] = normal[3 * (s - 1) + c];
                        }
                        --s;
                    }
                    body[s] = (int32_t)i;
                    distance[s] = t;
                    point[3 * s] = p.x; point[3 * s + 1] = p.y; point[3 * s + 2] = p.z;
                    normal[3 * s] = nn.x; normal[3 * s + 1] = nn.y; normal[3 * s + 2] = nn.z;
                }
                ++cnt;
            }
            if (!pass) { total += cnt; offsets[k + 1] = total; }
        }
    }
    return NE_OK;
}

/* The query Body (physics_world.cpp:190-227): makeSphere / makeBox at the
 * query pose for overlapSphere / overlapBox, or the caller's Body. */
static int query_body(const ne_query_shape* s, obody* q, v3* hv) {
    memset(q, 0, sizeof(*q));
    if (s->shape < NE_SHAPE_SPHERE || s->shape > NE_SHAPE_PLANE) return 0;
    q->shape = s->shape;
    q->x = mk(s->x[0], s->x[1], s->x[2]);
    q->q = mq(s->q[0], s->q[1], s->q[2], s->q[3]);
    q->he = mk(s->half_extents[0], s->half_extents[1], s->half_extents[2]);
    q->radius = s->radius;
    q->hh = s->half_height;
    q->mesh = -1;
    if (s->shape == NE_SHAPE_CONVEX) {
        if (!s->verts || s->nverts == 0) return 0;
        for (uint32_t k = 0; k < s->nverts; ++k) hv[k] = ld3(s->verts, k);
        q->hvp = hv;
        q->hull0 = 0;
        q->hull1 = s->nverts;
    }
    return 1;
}

/* detectContact(a, b).hit (detect.hpp:33-77): the same dispatch, with the
 * GJK/EPA fallback reduced to GJK -- epaContact reports a hit whenever GJK
 * enclosed the origin (epa.hpp:134-158), so the flag needs no EPA run. */
static int overlap_hit(const ne_oracle* o, const obody* a, const obody* b) {
    const int plane = a->shape == NE_SHAPE_PLANE || b->shape == NE_SHAPE_PLANE;
    const int sph = is_sphere(a) || is_sphere(b);
    const int cyl = a->shape == NE_SHAPE_CYLINDER || b->shape == NE_SHAPE_CYLINDER;
    const int cap = a->shape == NE_SHAPE_CAPSULE || b->shape == NE_SHAPE_CAPSULE;
    const int analytic = plane || (is_sphere(a) && is_sphere(b)) || (sph && cyl) ||
                         (a->shape == NE_SHAPE_CAPSULE && b->shape == NE_SHAPE_CAPSULE) || (sph && cap);
    if (analytic) return detect_contact(o, a, b).hit;
    supv tet[4];
    return gjk_tetra(o, a, b, min_image_shift(&o->box, a, b), tet);   /* convexContact, epa.hpp:153-158 */
}

int ne_oracle_overlap(ne_oracle* o, uint32_t n, const ne_query_shape* shapes, const uint32_t* mask, int32_t contacts,
                      uint32_t cap, uint32_t* offsets, int32_t* body, double* normal, double* depth) {
    if (!o || !offsets || (n && !shapes)) return NE_ERR_INVALID;
    for (uint32_t i = 0; i < o->n; ++i)
        if (o->B[i].shape == NE_SHAPE_MESH) return NE_ERR_UNSUPPORTED;
    uint32_t total = 0, maxv = 0;
    for (uint32_t k = 0; k < n; ++k)
        if (shapes[k].shape == NE_SHAPE_CONVEX && shapes[k].nverts > maxv) maxv = shapes[k].nverts;
    v3* hv = (v3*)malloc(sizeof(v3) * (maxv ? maxv : 1));
    offsets[0] = 0;
    int rc = NE_OK;
    for (int pass = 0; pass < 2 && rc == NE_OK; ++pass) {
        if (pass == 1 && total > cap) { rc = NE_ERR_CAPACITY; break; }
        if (pass == 1 && total && !body) { rc = NE_ERR_INVALID; break; }
        for (uint32_t k = 0; k < n; ++k) {
            obody q;
            if (!query_body(&shapes[k], &q, hv)) { rc = NE_ERR_INVALID; break; }
            const uint32_t m = mask ? mask[k] : 0xFFFFFFFFu;
            uint32_t at = pass ? offsets[k] : 0, cnt = 0;
            for (uint32_t i = 0; i < o->n; ++i) {
                const obody* b = &o->B[i];
                if (contacts && (b->flags & NE_FLAG_SENSOR)) continue;   /* triggers don't block */
                if (!layer_allowed(m, b)) continue;
                ocontact c;
                if (contacts) {
                    c = detect_contact(o, &q, b);
                    if (!c.hit) continue;
                } else {
                    if (!overlap_hit(o, &q, b)) continue;
                    memset(&c, 0, sizeof(c));
                }
                if (pass) {
                    body[at + cnt] = (int32_t)i;
                    if (normal) {
                        normal[3 * (at + cnt)] = c.normal.x;
                        normal[3 * (at + cnt) + 1] = c.normal.y;
                        normal[3 * (at + cnt) + 2] = c.normal.z;
                    }
                    if (depth) depth[at + cnt] = c.overlap;
                }
                ++cnt;
            }
            if (!pass) { total += cnt; offsets[k + 1] = total; }
        }
    }
    free(hv);
    return rc;
}

/* World::joints (world.hpp:130): the general joints, in order. */
int ne_oracle_set_general_joints(ne_oracle* o, uint32_t n, const ne_joint* joints) {
    if (!o || (n && !joints)) return NE_ERR_INVALID;
    for (uint32_t k = 0; k < n; ++k)
        if (joints[k].a >= o->n || joints[k].b >= o->n || joints[k].type < NE_JOINT_BALL || joints[k].type > NE_JOINT_SLIDER)
            return NE_ERR_INVALID;
    free(o->gj); free(o->gj_applied); free(o->gj_motor);
    o->ngj = n;
    o->gj = (ne_joint*)malloc(sizeof(ne_joint) * (n ? n : 1));
    o->gj_applied = (double*)calloc(n ? n : 1, sizeof(double));
    o->gj_motor = (double*)calloc(n ? n : 1, sizeof(double));
    if (n) memcpy(o->gj, joints, sizeof(ne_joint) * n);
    return NE_OK;
}

int ne_oracle_get_general_joints(ne_oracle* o, uint32_t n, int32_t* broken, double* applied, double* motor) {
    if (!o || n > o->ngj) return NE_ERR_INVALID;
    for (uint32_t k = 0; k < n; ++k) {
        if (broken) broken[k] = o->gj[k].broken;
        if (applied) applied[k] = o->gj_applied[k];
        if (motor) motor[k] = o->gj_motor[k];
    }
    return NE_OK;
}

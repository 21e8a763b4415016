// ne_query.hpp -- scene queries on the device: batched rays and swept spheres
// (PhysicsWorld::raycast / raycastAll / sphereCast, src/physics_world.cpp:
// 146-188, 229-251) and overlap tests of query shapes (overlapSphere /
// overlapBox / overlapContacts, :190-227).
//
// Rays.  The reference loops over every body in index order, shrinking the
// search distance to the best hit so far; whether a capsule is hit on its side
// or on a cap depends on that distance (the caps are tried only when the side
// misses within it), so the fold's order is part of the result.  Here one
// thread owns one ray and runs that fold exactly, over body tiles staged in
// LDS: every lane of a wave tests the same body at the same time, so the
// shape branch never diverges.  A tile whose bounding box (centres +- bounding
// radius, padded) the ray's remaining segment misses is skipped -- any body
// in it would fail `t < best` anyway -- and the block loads a tile only when
// one of its rays needs it.  Periodic worlds test every tile (each body is
// taken at its image nearest the ray's origin).
//
// Overlaps.  Thread per (query, body); a padded bounding-sphere test skips
// pairs detectContact cannot report, GJK decides the overlap (EPA always
// reports a hit once GJK encloses the origin, src/epa.hpp:134-153), and for
// overlapContacts the convex pairs queue for EPA: one thread per pair with a
// small workspace in global memory, and the runs that outgrow it once more,
// one block per pair, with the step's large-workspace EPA.
//
// Included by ne_gpu.hip after its body layout and EPA tiers (BodiesDev,
// load_geom, ld3, epa_block).
#pragma once

#include "ne_narrow.hpp"
#include "ne_prims.hpp"

namespace nea {

__device__ __forceinline__ double v3c(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// src/queries.hpp:18-31
__device__ inline bool ray_vs_sphere_at(V3 o, V3 d, V3 c, double r, double maxT, double& tOut, V3& nOut) {
    V3 m = o - c;
    double b = dot(m, d), cc = dot(m, m) - r * r;
    if (cc > 0.0 && b > 0.0) return false;
    double disc = b * b - cc;
    if (disc < 0.0) return false;
    double t = -b - sqrt(disc);
    if (t < 0.0) t = 0.0;
    if (t > maxT) return false;
    tOut = t;
    nOut = normalized((o + d * t) - c);
    return true;
}

// src/queries.hpp:35-59 (the box's own frame)
__device__ inline bool ray_vs_box_local(V3 ol, V3 dl, V3 he, double maxT, double& tOut, V3& nLocal) {
    double tmin = 0.0, tmax = maxT;
    int axis = -1;
    double sign = 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double o = v3c(ol, a), dd = v3c(dl, a), h = v3c(he, a);
        if (fabs(dd) < 1e-12) {
            if (o < -h || o > h) return false;
        } else {
            double inv = 1.0 / dd;
            double t1 = (-h - o) * inv, t2 = (h - o) * inv;
            double s = -1.0;
            if (t1 > t2) { double tmp = t1; t1 = t2; t2 = tmp; s = 1.0; }
            if (t1 > tmin) { tmin = t1; axis = a; sign = s; }
            if (t2 < tmax) tmax = t2;
            if (tmin > tmax) return false;
        }
    }
    if (axis < 0) return false;
    tOut = tmin;
    nLocal = V3{axis == 0 ? sign : 0.0, axis == 1 ? sign : 0.0, axis == 2 ? sign : 0.0};
    return true;
}

// src/queries.hpp:61-122 for a sphere, box, cylinder or capsule (the caller
// skips the shapes rayVsBody never hits) of orientation q, half extents he,
// radius r and half height h, centred at c.
__device__ inline bool ray_vs_body(int shape, Qt q, V3 he, double r, double h, V3 c, V3 o, V3 d, double maxT,
                                   double& tOut, V3& pOut, V3& nOut) {
    if (shape == SH_SPHERE) {
        if (!ray_vs_sphere_at(o, d, c, r, maxT, tOut, nOut)) return false;
        pOut = o + d * tOut;
        return true;
    }
    const V3 ol = qinvrot(q, o - c);
    const V3 dl = qinvrot(q, d);
    if (shape == SH_BOX) {
        V3 nl;
        if (!ray_vs_box_local(ol, dl, he, maxT, tOut, nl)) return false;
        nOut = qrot(q, nl);
        pOut = o + d * tOut;
        return true;
    }
    double best = maxT + 1.0;
    V3 nl{0.0, 0.0, 0.0};
    const double a2 = dl.x * dl.x + dl.z * dl.z;
    const double b2 = 2.0 * (ol.x * dl.x + ol.z * dl.z);
    const double c2 = ol.x * ol.x + ol.z * ol.z - r * r;
    if (a2 > 1e-18) {
        const double disc = b2 * b2 - 4.0 * a2 * c2;
        if (disc >= 0.0) {
            double t = (-b2 - sqrt(disc)) / (2.0 * a2);
            if (t < 0.0) t = (-b2 + sqrt(disc)) / (2.0 * a2);
            if (t >= 0.0 && t <= maxT) {
                const double y = ol.y + t * dl.y;
                if (y >= -h && y <= h && t < best) {
                    best = t;
                    nl = V3{(ol.x + t * dl.x) / r, 0.0, (ol.z + t * dl.z) / r};
                }
            }
        }
    }
    if (shape == SH_CYLINDER) {
        for (int s = -1; s <= 1; s += 2) {
            const double yc = s * h;
            if (fabs(dl.y) < 1e-12) continue;
            const double t = (yc - ol.y) / dl.y;
            if (t < 0.0 || t > maxT || t >= best) continue;
            const double px = ol.x + t * dl.x, pz = ol.z + t * dl.z;
            if (px * px + pz * pz <= r * r) { best = t; nl = V3{0.0, (double)s, 0.0}; }
        }
    } else if (best > maxT) {   // capsule: the cap spheres only when the side missed
        for (int s = -1; s <= 1; s += 2) {
            double tS;
            V3 nS;
            if (ray_vs_sphere_at(ol, dl, V3{0.0, (double)s * h, 0.0}, r, maxT, tS, nS) && tS < best) {
                best = tS;
                nl = nS;
            }
        }
    }
    if (best > maxT) return false;
    tOut = best;
    nOut = normalized(qrot(q, nl));
    pOut = o + d * tOut;
    return true;
}

__device__ __forceinline__ bool ray_shape(int s) {
    return s == SH_SPHERE || s == SH_BOX || s == SH_CYLINDER || s == SH_CAPSULE;
}

// Rays as structure of arrays (n each): origin, direction, search distance,
// cast radius (sphere casts), layer mask (nullable: every layer).
struct RayIn {
    const double *ox, *oy, *oz, *dx, *dy, *dz, *maxd, *rad;
    const uint32_t* mask;
};

// Results: nearest hit (body, distance, point xyz, normal xyz) per ray, or
// every hit at [off[r], off[r+1]) for raycastAll.
struct RayOut {
    int32_t* body;
    double *t, *px, *py, *pz, *nx, *ny, *nz;
    uint32_t* count;       // raycastAll count pass: hits per ray
    const uint32_t* off;   // raycastAll emit pass: exclusive offsets
};

constexpr int RAY_THREADS = 256;
constexpr int RAY_F = 13;   // staged fields per body: x(3) q(4) he(3) radius, half height, bounding radius

// The bounding radius of body i for the skip tests, padded 1e-5 relative;
// infinite (never skipped) for a quaternion far from unit length, whose
// rotation may stretch the body, and for planes and meshes.
__device__ __forceinline__ double skip_radius(const BodiesDev& B, uint32_t i) {
    const double q2 = B.q0[i] * B.q0[i] + B.q1[i] * B.q1[i] + B.q2[i] * B.q2[i] + B.q3[i] * B.q3[i];
    const double r = B.bR[i];
    return (fabs(q2 - 1.0) > 1e-6 || r >= 1e29) ? INFINITY : r * (1.0 + 1e-5);
}

// Per tile of RAY_THREADS bodies in index order: the bounding box of the
// centres of its bodies (RAYS: only the ray-hittable ones) grown by each
// one's skip radius.  A tile with a body of infinite skip radius is never
// skipped (lo = -inf, hi = +inf); a tile with no body to test is always
// skipped (lo = +inf, hi = -inf).
template <bool RAYS>
__global__ void __launch_bounds__(RAY_THREADS) k_query_tiles(BodiesDev B, double* __restrict__ tiles) {
    __shared__ double red[6][RAY_THREADS];
    const uint32_t i = blockIdx.x * RAY_THREADS + threadIdx.x;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (i < B.n && (!RAYS || ray_shape(B.shape[i]))) {
        const double R = skip_radius(B, i);
        const double x[3] = {B.x0[i], B.x1[i], B.x2[i]};
        for (int a = 0; a < 3; ++a) {
            lo[a] = R == INFINITY ? -INFINITY : x[a] - R;
            hi[a] = R == INFINITY ? INFINITY : x[a] + R;
        }
    }
    for (int a = 0; a < 3; ++a) { red[a][threadIdx.x] = lo[a]; red[3 + a][threadIdx.x] = hi[a]; }
    __syncthreads();
    for (int s = RAY_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 3; ++a) {
                red[a][threadIdx.x] = fmin(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = fmax(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) tiles[6 * (size_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Whether the segment o + d t, 0 <= t <= best, can reach a body of the tile
// (its box grown by pad).  Conservative: the pad covers the hit test's
// rounding, relative to the coordinates and the segment's length.
__device__ __forceinline__ bool ray_meets_tile(V3 o, V3 d, double best, double grow, const double* __restrict__ tb) {
    const double lo[3] = {tb[0], tb[1], tb[2]}, hi[3] = {tb[3], tb[4], tb[5]};
    if (!(lo[0] <= hi[0])) return false;   // no ray-hittable body
    double span = fabs(o.x) + fabs(o.y) + fabs(o.z) + best;
#pragma unroll
    for (int a = 0; a < 3; ++a) span += fmax(fabs(lo[a]), fabs(hi[a]));
    const double pad = grow + 1e-6 * (1.0 + span);
    double tmin = -pad, tmax = best + pad;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double oa = v3c(o, a), da = v3c(d, a);
        const double l = lo[a] - pad, h = hi[a] + pad;
        if (da == 0.0) {
            if (oa < l || oa > h) return false;
        } else {
            double t1 = (l - oa) / da, t2 = (h - oa) / da;
            if (t1 > t2) { const double t = t1; t1 = t2; t2 = t; }
            tmin = fmax(tmin, t1);
            tmax = fmin(tmax, t2);
            if (!(tmin <= tmax)) return false;
        }
    }
    return true;
}

// MODE 0: nearest hit (raycast / sphereCast when CAST); MODE 1: count every
// hit within the search distance; MODE 2: write them, sorted by distance
// (raycastAll).  One thread per ray; bodies in index order.
template <int MODE, bool CAST, bool CULL>
__global__ void __launch_bounds__(RAY_THREADS) k_ray(BodiesDev B, Box3 box, uint32_t nr, RayIn R,
                                                     const double* __restrict__ tiles, RayOut O) {
    __shared__ double sf[RAY_F][RAY_THREADS];
    __shared__ int sshape[RAY_THREADS];
    __shared__ uint32_t slayer[RAY_THREADS];
    const uint32_t r = blockIdx.x * RAY_THREADS + threadIdx.x;
    const bool live = r < nr;
    V3 org{0.0, 0.0, 0.0}, d{0.0, 0.0, 0.0};
    double best = 0.0, rc = 0.0;
    uint32_t m = 0;
    if (live) {
        org = V3{R.ox[r], R.oy[r], R.oz[r]};
        d = normalized(V3{R.dx[r], R.dy[r], R.dz[r]});
        best = R.maxd[r];
        rc = CAST ? R.rad[r] : 0.0;
        m = R.mask ? R.mask[r] : 0xFFFFFFFFu;
    }
    const double maxd = best;
    int32_t hb = -1;
    double ht = 0.0;
    V3 hp{0.0, 0.0, 0.0}, hn{0.0, 0.0, 0.0};
    uint32_t nhit = 0, at = 0, lim = 0;
    if (MODE == 2 && live) {   // the count pass's hits; never more are written
        at = O.off[r];
        lim = O.off[r + 1] - at;
    }
    const uint32_t ntiles = (B.n + RAY_THREADS - 1) / RAY_THREADS;
    for (uint32_t tile = 0; tile < ntiles; ++tile) {
        bool want = live;
        if (CULL && want) want = ray_meets_tile(org, d, best, 2.0 * rc, tiles + 6 * (size_t)tile);
        if (!__syncthreads_or(want ? 1 : 0)) continue;
        const uint32_t base = tile * RAY_THREADS;
        const uint32_t i = base + threadIdx.x;
        if (i < B.n) {
            const int t = threadIdx.x;
            sf[0][t] = B.x0[i]; sf[1][t] = B.x1[i]; sf[2][t] = B.x2[i];
            sf[3][t] = B.q0[i]; sf[4][t] = B.q1[i]; sf[5][t] = B.q2[i]; sf[6][t] = B.q3[i];
            sf[7][t] = B.he0[i]; sf[8][t] = B.he1[i]; sf[9][t] = B.he2[i];
            sf[10][t] = B.rad[i]; sf[11][t] = B.hh[i];
            sf[12][t] = skip_radius(B, i);
            sshape[t] = B.shape[i];
            slayer[t] = B.layer[i];
        }
        __syncthreads();
        if (want) {
            const uint32_t cnt = min((uint32_t)RAY_THREADS, B.n - base);
            for (uint32_t j = 0; j < cnt; ++j) {
                const int sh = sshape[j];
                if (!ray_shape(sh)) continue;
                if (!(m & (1u << (slayer[j] & 31u)))) continue;
                const V3 x{sf[0][j], sf[1][j], sf[2][j]};
                const Qt q{sf[3][j], sf[4][j], sf[5][j], sf[6][j]};
                V3 he{sf[7][j], sf[8][j], sf[9][j]};
                double rad = sf[10][j], hh = sf[11][j];
                if (CAST) {   // physics_world.cpp:237-240
                    rad = rad + rc;
                    if (sh == SH_BOX) he = he + V3{rc, rc, rc};
                    else if (sh == SH_CYLINDER) hh = hh + rc;
                }
                const V3 cimg = org + min_image(box, x - org);
                {   // skip a body whose padded bounding sphere (grown by a cast's
                    // inflation) the segment cannot reach: any hit lies within it
                    const double reach = sf[12][j] + 2.0 * rc;
                    const double tl = MODE == 0 ? best : maxd;
                    const V3 mo = cimg - org;
                    const double tc = fmin(fmax(dot(mo, d), 0.0), tl);
                    const V3 cp = mo - d * tc;
                    const double pad = 1e-6 * (1.0 + fabs(org.x) + fabs(org.y) + fabs(org.z) + tl + reach);
                    if (norm2(cp) > (reach + pad) * (reach + pad)) continue;
                }
                double t;
                V3 p, n;
                if (MODE == 0) {
                    if (ray_vs_body(sh, q, he, rad, hh, cimg, org, d, best, t, p, n) && t < best) {
                        best = t;
                        hb = (int32_t)(base + j);
                        ht = t;
                        hn = n;
                        hp = CAST ? p - n * rc : p;
                    }
                } else if (ray_vs_body(sh, q, he, rad, hh, cimg, org, d, maxd, t, p, n)) {
                    if (MODE == 2 && nhit < lim) {   // stable insertion by distance: equal distances keep body order
                        uint32_t s = at + nhit;
                        while (s > at && O.t[s - 1] > t) {
                            O.body[s] = O.body[s - 1];
                            O.t[s] = O.t[s - 1];
                            O.px[s] = O.px[s - 1]; O.py[s] = O.py[s - 1]; O.pz[s] = O.pz[s - 1];
                            O.nx[s] = O.nx[s - 1]; O.ny[s] = O.ny[s - 1]; O.nz[s] = O.nz[s - 1];
                            --s;
                        }
                        O.body[s] = (int32_t)(base + j);
                        O.t[s] = t;
                        O.px[s] = p.x; O.py[s] = p.y; O.pz[s] = p.z;
                        O.nx[s] = n.x; O.ny[s] = n.y; O.nz[s] = n.z;
                    }
                    ++nhit;
                }
            }
        }
        __syncthreads();
    }
    if (!live) return;
    if (MODE == 0) {
        O.body[r] = hb;
        O.t[r] = ht;
        O.px[r] = hp.x; O.py[r] = hp.y; O.pz[r] = hp.z;
        O.nx[r] = hn.x; O.ny[r] = hn.y; O.nz[r] = hn.z;
    } else if (MODE == 1) {
        O.count[r] = nhit;
    }
}

// A query shape on the device: its geometry (hull in device memory), the
// padded bounding radius of the pair test (1e30: never skipped) and its mask.
struct QueryDev {
    Geom g;
    double bR;
    uint32_t mask;
};

constexpr int OVL_THREADS = 256;
static_assert(OVL_THREADS == RAY_THREADS, "an overlap block covers one body tile of k_query_tiles");

// Overlap of query q with body i: a padded bounding-sphere
// test at the minimum image first (bodies with a quaternion far from unit
// length, planes and queries with bR 1e30 are always tested), then
// detectContact with EPA deferred: hit 1 (analytic contact) or 2 (GJK
// overlap of a convex pair).
__device__ __forceinline__ Contact ovl_test(const BodiesDev& B, const Box3& box, const QueryDev& Q, uint32_t i,
                                            int contacts) {
    Contact c = no_contact();
    const uint32_t fl = B.flags[i];
    if (contacts && (fl & NE_FLAG_SENSOR)) return c;   // triggers don't block
    if (!(Q.mask & (1u << (B.layer[i] & 31u)))) return c;
    const V3 x = ld3(B.x0, B.x1, B.x2, i);
    const Qt q = Qt{B.q0[i], B.q1[i], B.q2[i], B.q3[i]};
    const double q2 = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    const double rb = B.bR[i];
    if (fabs(q2 - 1.0) <= 1e-6 && rb < 1e29 && Q.bR < 1e29) {
        const double reach = (Q.bR + rb) * (1.0 + 1e-5) + 1e-9;
        if (norm2(min_image(box, x - Q.g.x)) > reach * reach) return c;
    }
    EpaWS ws{};   // GJK only: overlapping convex pairs are queued for EPA
    return detect_contact(box, Q.g, load_geom(B, i), ws);
}

// Query q0 + blockIdx.y against the bodies of tile blockIdx.x (skipped
// whole when tiles is set and the query cannot reach the tile's box).
// EMIT false: hits per (query, tile) into cnt; EMIT true: the hits' body
// indices at off[...] + rank (body order), contacts of analytic pairs, and
// convex pairs queued (slot, body, query) for EPA.
// Whether query Q can reach a body of a tile (k_query_tiles without RAYS):
// the distance from its centre to the tile's box against its padded
// bounding radius.
__device__ __forceinline__ bool query_meets_tile(const QueryDev& Q, const double* __restrict__ tb) {
    if (!(tb[0] <= tb[3])) return false;
    if (Q.bR >= 1e29) return true;
    const double xq[3] = {Q.g.x.x, Q.g.x.y, Q.g.x.z};
    double d2 = 0.0, span = fabs(xq[0]) + fabs(xq[1]) + fabs(xq[2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double e = fmax(fmax(tb[a] - xq[a], 0.0), xq[a] - tb[3 + a]);
        d2 += e * e;
        span += fmax(fabs(tb[a]), fabs(tb[3 + a]));
    }
    const double reach = Q.bR * (1.0 + 1e-5) + 1e-6 * (1.0 + span + Q.bR);
    return d2 <= reach * reach;
}

template <bool EMIT>
__global__ void __launch_bounds__(OVL_THREADS) k_ovl(BodiesDev B, Box3 box, const QueryDev* __restrict__ Q, uint32_t q0,
                                                     const double* __restrict__ tiles, int contacts, uint32_t* __restrict__ cnt,
                                                     const uint32_t* __restrict__ off, int32_t* __restrict__ obody,
                                                     double* __restrict__ onrm, double* __restrict__ odep,
                                                     uint32_t* __restrict__ epa_ent, uint32_t* __restrict__ epa_n) {
    __shared__ uint32_t lds[PRIM_WAVES + 1];
    const uint32_t q = q0 + blockIdx.y;
    const uint32_t i = blockIdx.x * OVL_THREADS + threadIdx.x;
    Contact c = no_contact();
    const bool reach = !tiles || query_meets_tile(Q[q], tiles + 6 * (size_t)blockIdx.x);   // the same for the block
    if (i < B.n && reach) c = ovl_test(B, box, Q[q], i, contacts);
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(c.hit ? 1u : 0u, &tot, lds);
    const size_t cell = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
    if (!EMIT) {
        if (threadIdx.x == 0) cnt[cell] = tot;
        return;
    }
    if (!c.hit) return;
    const uint32_t k = off[cell] + ex;
    if (k >= off[cell + 1]) return;   // the count pass's hits; never more are written
    obody[k] = (int32_t)i;
    if (!contacts) return;
    if (c.hit == 2) {
        const uint32_t e = atomicAdd(epa_n, 1u);
        epa_ent[3 * (size_t)e] = k;
        epa_ent[3 * (size_t)e + 1] = i;
        epa_ent[3 * (size_t)e + 2] = q;
    } else {
        onrm[3 * (size_t)k] = c.normal.x;
        onrm[3 * (size_t)k + 1] = c.normal.y;
        onrm[3 * (size_t)k + 2] = c.normal.z;
        odep[k] = c.overlap;
    }
}

// out[t] = off[t * stride] for t <= nq: the first offset of each query.
__global__ void k_ovl_qoff(const uint32_t* __restrict__ off, uint32_t stride, uint32_t nq, uint32_t* __restrict__ out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= nq) out[t] = off[(size_t)t * stride];
}

// EPA of queued overlapContacts pairs (src/epa.hpp:71-158), one thread per
// pair, the polytope in this launch's global workspace (maxf faces, maxh
// horizon edges per pair).  A run that outgrows it is listed in retry for
// the block-wide tier (k_ovl_big).
__global__ void __launch_bounds__(64) k_ovl_epa(BodiesDev B, Box3 box, const QueryDev* __restrict__ Q,
                                                const uint32_t* __restrict__ ent, uint32_t nent, SupV* __restrict__ wv,
                                                Face* __restrict__ wf, int (*wh)[2], int maxf, int maxh,
                                                double* __restrict__ onrm, double* __restrict__ odep,
                                                uint32_t* __restrict__ retry, uint32_t* __restrict__ nretry) {
    const uint32_t e = blockIdx.x * 64 + threadIdx.x;
    if (e >= nent) return;
    const uint32_t k = ent[3 * (size_t)e], i = ent[3 * (size_t)e + 1], q = ent[3 * (size_t)e + 2];
    const Geom A = Q[q].g;
    const Geom G = load_geom(B, i);
    const V3 shift = min_image_shift(box, A.x, G.x);
    SupV tet[4];
    if (!gjk_tetra(A, G, shift, tet)) return;   // cannot happen: the pair was queued after this same GJK
    HorizonList hz;
    hz.hor = wh + (size_t)e * maxh;
    hz.cap = maxh;
    hz.n = 0;
    bool overflow = false, too_long = false;
    const Contact c = epa_contact(A, G, shift, tet, wv + (size_t)e * EPA_MAXV, wf + (size_t)e * maxf, maxf, EPA_MAXV, 64,
                                  hz, overflow, too_long);
    if (overflow) {
        const uint32_t s = atomicAdd(nretry, 1u);
        retry[3 * (size_t)s] = k;
        retry[3 * (size_t)s + 1] = i;
        retry[3 * (size_t)s + 2] = q;
        return;
    }
    onrm[3 * (size_t)k] = c.normal.x;
    onrm[3 * (size_t)k + 1] = c.normal.y;
    onrm[3 * (size_t)k + 2] = c.normal.z;
    odep[k] = c.overlap;
}

// The runs the per-thread tier could not hold: one block of BIG_NW waves per
// pair, with the step's large-workspace EPA (epa_block) and faces_per faces
// per block in global memory; a run that needs more stops at its closest
// face and counts in *capped.
__global__ void __launch_bounds__(64 * BIG_NW) k_ovl_big(BodiesDev B, Box3 box, const QueryDev* __restrict__ Q,
                                                         const uint32_t* __restrict__ ent, uint32_t nent,
                                                         Face* __restrict__ faces, uint16_t* __restrict__ occ,
                                                         int faces_per, double* __restrict__ onrm,
                                                         double* __restrict__ odep, uint32_t* __restrict__ capped) {
    __shared__ EpaBlockLds L;
    Face* f = faces + (size_t)blockIdx.x * faces_per;
    uint16_t* oc = occ + 3 * (size_t)blockIdx.x * faces_per;
    for (uint32_t e = blockIdx.x; e < nent; e += gridDim.x) {
        const uint32_t k = ent[3 * (size_t)e], i = ent[3 * (size_t)e + 1], q = ent[3 * (size_t)e + 2];
        const Geom A = Q[q].g;
        const Geom G = load_geom(B, i);
        const V3 shift = min_image_shift(box, A.x, G.x);
        SupV tet[4];
        bool overflow = false;
        Contact c = no_contact();
        if (gjk_tetra(A, G, shift, tet)) c = epa_block(A, G, shift, tet, L, f, oc, faces_per, overflow);
        if (threadIdx.x == 0) {
            if (overflow) atomicAdd(capped, 1u);
            onrm[3 * (size_t)k] = c.normal.x;
            onrm[3 * (size_t)k + 1] = c.normal.y;
            onrm[3 * (size_t)k + 2] = c.normal.z;
            odep[k] = c.overlap;
        }
        __syncthreads();
    }
}

}  // namespace nea

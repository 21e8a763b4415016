// ne_gpu.hip -- the MI355X (gfx950) rigid-body step behind include/ne_gpu.h.
//
// One context = one world resident in HBM as structure-of-arrays FP64, stepped
// on one HIP stream.  A step is the reference's `World::step()`
// (src/world.cpp:133-146) re-designed as a pipeline of data-parallel kernels:
//
//   k_integrate        world.cpp:373-411   per body
//   grid build         world.cpp:424-501   counting sort of bodies into cells
//   candidate rows     world.cpp:472-499   27-cell scan, rows (i, j>i) in order
//   k_narrow           detect.hpp:33-77, world.cpp:88-130   per candidate pair
//   k_cons_build       world.cpp:508-555   per contact, ordered by (i, j, slot)
//   islands            --                  union-find over dynamic bodies
//   k_solve_islands    world.cpp:557-616   warm start, I x GS sweep, correction
//   k_wrap / k_walls   world.cpp:635-702   per body
//
// Exactness: the reference solves constraints strictly in sorted (i, j, slot)
// order (Gauss-Seidel).  Two constraints that share no body whose velocity can
// change commute exactly, so partitioning constraints into connected islands
// and sweeping each island in that same order reproduces the sequential
// result bit for bit, with no float atomics.  All kernels are compiled with
// -ffp-contract=off and follow the reference's operation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ne_gpu.h"
#include "ne_dmath.hpp"
#include "ne_narrow.hpp"
#include "ne_solve.hpp"
#include "ne_adjoint.hpp"
#include "ne_prims.hpp"
#include "ne_atan2.hpp"

using namespace nea;

namespace {

// ---------------------------------------------------------------------------
// Device views
// ---------------------------------------------------------------------------

struct BodiesDev {
    double *x0, *x1, *x2, *v0, *v1, *v2, *q0, *q1, *q2, *q3, *w0, *w1, *w2;
    double *f0, *f1, *f2, *t0, *t1, *t2;
    double *im, *ii0, *ii1, *ii2;
    double *rad, *hh, *he0, *he1, *he2, *bR;
    double *fric, *rest, *ldamp, *adamp, *gs;
    int32_t* shape;
    uint32_t *layer, *mask, *flags;
    uint32_t *hull_off, *nc_off, *nc_idx;
    double* hull;
    double *st, *ims, *iis0, *iis1, *iis2;   // sleep timer, stored inverse mass / inertia (world.cpp:317-371)
    int sleep_on;                            // island sleeping enabled: candidates keep sleeping bodies
    int ev_filter;                           // the facade's event filter instead of the step's (events pass)
    uint32_t n;
    // Batched worlds: world id per body and the world table (both null for
    // one world).  Bodies of a world are contiguous.
    const uint32_t* wid;
    const struct WorldDev* W;
};

struct SimParams {
    Box3 box;
    double dt;
    double g[3];
    double restitution, friction, beta;
    int rcomb, fcomb;
    int warm;
    int iters;
    int open_boundary;
    int sleep;
    double sleep_lin, sleep_ang, sleep_time, wake_vel;
    int no_manifold;   // detectContact's one contact per pair, no box-box manifold (events pass)
};

// Grid description, computed on the device every step.
struct GridDev {
    double origin[3];
    double cw[3];
    double inv_cw[3];   // cell coordinates use x * (1/cw): binning and lookup share one rule
    int32_t nc[3];
    uint32_t ncells;
    int32_t periodic;
};

// One of several independent periodic worlds stepped together: its box, its
// default material (src/world.hpp:91-92), its slice of the shared cell table
// and its range of bodies and of grid-excluded bodies.
struct WorldDev {
    Box3 box;
    GridDev grid;
    double restitution, friction;
    uint32_t cell_base;
    uint32_t body_begin, body_end;
    uint32_t excl_begin, excl_end;   // excl[excl_begin..excl_end) lie in this world
    uint32_t flag_base;              // first flag of this world's excluded-body rows
};

__device__ __forceinline__ const Box3& box_of(const BodiesDev& B, const SimParams& P, uint32_t i) {
    return B.W ? B.W[B.wid[i]].box : P.box;
}

enum CounterSlot : int {
    CNT_P = 0,        // candidate pairs (clamped)
    CNT_HIT = 1,      // hit pairs
    CNT_C = 2,        // constraints (clamped)
    CNT_ISL = 3,      // islands
    CNT_MAXISL = 4,   // largest island (constraints)
    CNT_NGRID = 5,    // gridded bodies
    CNT_CELLS = 6,    // cells in use
    CNT_FAULT = 7,    // NE_FAULT_* bits
    CNT_PRAW = 8,     // candidate pairs before clamping
    CNT_CRAW = 9,     // constraints before clamping
    CNT_NBIG = 10,    // islands handed to the workgroup solver
    CNT_EPABIG = 11,  // pairs queued for the large-workspace EPA rerun (this step)
    CNT_EPATRUNC = 12,   // EPA runs stopped at the largest workspace (this call)
    CNT_EPARERUN = 13,   // large-workspace EPA reruns (this call)
    CNT_EPAQ = 14,       // convex pairs queued for EPA (this step)
    CNT_WFC = 15,        // constraints of the large islands (this step)
    CNT_MQRAW = 16,      // mesh triangle candidates before clamping (this step)
    CNT_MQ = 17,         // mesh triangle candidates (clamped)
    CNT_WFT = 18,        // tasks (contact pairs) of the large islands (this step)
    CNT_EPALONG = 19,    // convex pairs whose EPA ran past the first launch's budget (this step)
    CNT_MQLONG = 20,     // mesh triangle candidates likewise
    CNT_EPABIG_PAIRS = 21,   // pair entries of the large-workspace queue (mesh entries follow them)
    CNT_EPALONG2 = 22,   // convex pairs whose EPA ran past the second launch's budget (this step)
    CNT_HOSTCHK = 23,    // a host call's own check (snapshot loads); no step kernel uses it
    CNT_EV = 24,         // facade events pass: overlapping pairs of the post-step state
    CNT_EPALONG3 = 25,   // convex pairs the second LDS EPA tier left for the group tier
    CNT_MQLONG2 = 26,    // mesh triangle candidates the group tier left for the one-wave tier
    CNT_MQEPA = 27,      // mesh triangle candidates GJK found overlapping (queued for EPA)
    CNT_N = 28
};

constexpr int TPB = 256;
constexpr int WF_REC = 16;   // doubles per constraint record and per body record (one 128-byte line)

// Wave-aggregated atomics.  When one island spans the scene, every lane of a
// wave targets the same counter; lanes with equal keys combine first and one
// lane per key issues the atomic.  All 64 lanes must call these together
// (`active` selects the lanes that take part).
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

// atomicAdd(&ctr[key], 1) for each active lane; returns the lane's old value
// (distinct per lane).  The keys of the first two leaders are combined; the
// remaining lanes (keys that are not shared much, as in a gas of small
// islands) add on their own, so the cost stays two rounds.
__device__ __forceinline__ uint32_t wave_add1(uint32_t* ctr, uint32_t key, bool active) {
    uint64_t pending = __ballot(active);
    uint32_t out = 0;
    const uint64_t below = lane_id() ? (~0ull >> (64 - lane_id())) : 0ull;
    for (int round = 0; round < 2 && pending; ++round) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lkey = __shfl(key, leader);
        const uint64_t same = __ballot(active && key == lkey) & pending;
        uint32_t base = 0;
        if (lane_id() == leader) base = atomicAdd(&ctr[lkey], (uint32_t)__popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane_id()) & 1ull) out = base + (uint32_t)__popcll(same & below);
        pending &= ~same;
    }
    if ((pending >> lane_id()) & 1ull) out = atomicAdd(&ctr[key], 1u);
    return out;
}

// atomicMin(&ctr[key], val) for each active lane (the same two rounds).
__device__ __forceinline__ void wave_min_u64(unsigned long long* ctr, uint32_t key, unsigned long long val, bool active) {
    uint64_t pending = __ballot(active);
    for (int round = 0; round < 2 && pending; ++round) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lkey = __shfl(key, leader);
        const uint64_t same = __ballot(active && key == lkey) & pending;
        unsigned long long m = ((same >> lane_id()) & 1ull) ? val : ~0ull;
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(m, off);
            m = o < m ? o : m;
        }
        if (lane_id() == leader) atomicMin(&ctr[lkey], m);
        pending &= ~same;
    }
    if ((pending >> lane_id()) & 1ull) atomicMin(&ctr[key], val);
}

__device__ __forceinline__ V3 ld3(const double* a, const double* b, const double* c, uint32_t i) {
    return V3{a[i], b[i], c[i]};
}
__device__ __forceinline__ void st3(double* a, double* b, double* c, uint32_t i, V3 v) {
    a[i] = v.x; b[i] = v.y; c[i] = v.z;
}
__device__ __forceinline__ Qt ldq(const BodiesDev& B, uint32_t i) { return Qt{B.q0[i], B.q1[i], B.q2[i], B.q3[i]}; }

__device__ __forceinline__ Geom load_geom(const BodiesDev& B, uint32_t i) {
    Geom g;
    g.x = ld3(B.x0, B.x1, B.x2, i);
    g.q = ldq(B, i);
    g.he = ld3(B.he0, B.he1, B.he2, i);
    g.radius = B.rad[i];
    g.hh = B.hh[i];
    g.shape = B.shape[i];
    uint32_t h0 = B.hull_off[i], h1 = B.hull_off[i + 1];
    g.hull = B.hull + 3u * h0;
    g.nhull = h1 - h0;
    return g;
}

// ---------------------------------------------------------------------------
// Integrate: src/world.cpp:373-411
// ---------------------------------------------------------------------------

constexpr uint32_t FLAG_GRID_EXCLUDED = 1u << 30;   // internal: body kept out of the grid

__device__ __forceinline__ bool same_bits(double a, double b) {
    return __double_as_longlong(a) == __double_as_longlong(b);
}
__device__ __forceinline__ bool same_bits(V3 a, V3 b) { return same_bits(a.x, b.x) && same_bits(a.y, b.y) && same_bits(a.z, b.z); }
__device__ __forceinline__ bool same_bits(Qt a, Qt b) {
    return same_bits(a.w, b.w) && same_bits(a.x, b.x) && same_bits(a.y, b.y) && same_bits(a.z, b.z);
}

struct StaticGrid {   // periodic worlds: the grid never changes between steps
    double origin[3], cw[3], inv_cw[3];
    int32_t nc[3];
};

__device__ __forceinline__ uint32_t static_cell(const StaticGrid& g, V3 x);
__device__ __forceinline__ uint32_t grid_cell(const GridDev& g, V3 x);

// Integrate (src/world.cpp:373-411), optionally preceded by the previous
// step's periodic wrap (world.cpp:698-702, same per-body arithmetic, moved to
// the start of the next step) and followed by binning the body into the
// static periodic grid.
struct SortedBody {
    double x, y, z, r;
};

template <bool FORCES, bool WRAP, bool BIN>
__global__ void __launch_bounds__(TPB) k_integrate(BodiesDev B, SimParams P, StaticGrid G, SortedBody* __restrict__ xp,
                                                   uint32_t* __restrict__ cell_count, uint32_t* __restrict__ parent,
                                                   uint32_t* __restrict__ claimed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    parent[i] = i;      // this step's union-find forest starts as singletons
    claimed[i] = 0u;    // (also: no island claim, and with sleeping no wake yet -- see k_sleep_wake_mark)
    const double dt = P.dt;
    const uint32_t fl = B.flags[i];
    const double im = B.im[i];
    V3 x = ld3(B.x0, B.x1, B.x2, i);
    if (WRAP && im > 0.0) x = wrap_pos(box_of(B, P, i), x);
    V3 f = FORCES ? ld3(B.f0, B.f1, B.f2, i) : V3{0.0, 0.0, 0.0};
    V3 tq = FORCES ? ld3(B.t0, B.t1, B.t2, i) : V3{0.0, 0.0, 0.0};
    if (fl & NE_FLAG_KINEMATIC) {
        V3 v = ld3(B.v0, B.v1, B.v2, i), w = ld3(B.w0, B.w1, B.w2, i);
        x = x + v * dt;
        Qt q = integrate_orientation(ldq(B, i), w, dt);
        B.q0[i] = q.w; B.q1[i] = q.x; B.q2[i] = q.y; B.q3[i] = q.z;
        st3(B.x0, B.x1, B.x2, i, x);
    } else if (!(im <= 0.0)) {
        V3 v = ld3(B.v0, B.v1, B.v2, i), w = ld3(B.w0, B.w1, B.w2, i);
        Qt q = ldq(B, i);
        const V3 w0 = w;
        const Qt q0 = q;
        V3 invI = ld3(B.ii0, B.ii1, B.ii2, i);
        const V3 g{P.g[0], P.g[1], P.g[2]};
        v = v + (g * B.gs[i] + f * im) * dt;
        x = x + v * dt;
        V3 L = angular_momentum(q, invI, w);
        L = L + tq * dt;
        q = integrate_orientation(q, w, dt);
        w = apply_inv_inertia(q, invI, L);
        const double ld = B.ldamp[i], ad = B.adamp[i];
        if (ld > 0.0) v = v * (1.0 / (1.0 + ld * dt));
        if (ad > 0.0) w = w * (1.0 / (1.0 + ad * dt));
        st3(B.x0, B.x1, B.x2, i, x);
        st3(B.v0, B.v1, B.v2, i, v);
        // a body that does not turn keeps q and w bit for bit: skip those stores
        if (!same_bits(w, w0)) st3(B.w0, B.w1, B.w2, i, w);
        if (!same_bits(q, q0)) { B.q0[i] = q.w; B.q1[i] = q.x; B.q2[i] = q.y; B.q3[i] = q.z; }
    } else if (WRAP) {
        // static body: unchanged (wrap only applies to invMass > 0)
    }
    if (FORCES) {
        st3(B.f0, B.f1, B.f2, i, V3{0.0, 0.0, 0.0});
        st3(B.t0, B.t1, B.t2, i, V3{0.0, 0.0, 0.0});
    }
    xp[i] = SortedBody{x.x, x.y, x.z, B.bR[i]};   // packed pose for the grid fill
    if (BIN && !(fl & FLAG_GRID_EXCLUDED)) {
        uint32_t cell;
        if (B.W) {
            const WorldDev& w = B.W[B.wid[i]];
            cell = w.cell_base + grid_cell(w.grid, x);
        } else {
            cell = static_cell(G, x);
        }
        atomicAdd(&cell_count[cell], 1u);
    }
}

// ---------------------------------------------------------------------------
// Uniform grid (counting sort).  Cells are >= 2*maxR of the gridded bodies
// wide, so every pair whose bounding spheres overlap lies in the same or an
// adjacent cell (periodic: modulo the cell count).  Bodies far larger than
// the rest (planes, a ground box) are kept out of the grid and paired with
// everything, so they cannot inflate the cell size the way they do in the
// reference (src/world.cpp:429-433).
// ---------------------------------------------------------------------------

constexpr int AABB_BLOCKS = 256;

__global__ void __launch_bounds__(TPB) k_aabb_partial(BodiesDev B, const uint32_t* __restrict__ gridded,
                                                      uint32_t ngrid, double* __restrict__ part) {
    __shared__ double lo[3][TPB], hi[3][TPB];
    double l0 = 1e300, l1 = 1e300, l2 = 1e300, h0 = -1e300, h1 = -1e300, h2 = -1e300;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ngrid; t += gridDim.x * blockDim.x) {
        uint32_t i = gridded[t];
        double a = B.x0[i], b = B.x1[i], c = B.x2[i];
        l0 = fmin(l0, a); l1 = fmin(l1, b); l2 = fmin(l2, c);
        h0 = fmax(h0, a); h1 = fmax(h1, b); h2 = fmax(h2, c);
    }
    lo[0][threadIdx.x] = l0; lo[1][threadIdx.x] = l1; lo[2][threadIdx.x] = l2;
    hi[0][threadIdx.x] = h0; hi[1][threadIdx.x] = h1; hi[2][threadIdx.x] = h2;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 3; ++a) {
                lo[a][threadIdx.x] = fmin(lo[a][threadIdx.x], lo[a][threadIdx.x + s]);
                hi[a][threadIdx.x] = fmax(hi[a][threadIdx.x], hi[a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; ++a) { part[6 * blockIdx.x + a] = lo[a][0]; part[6 * blockIdx.x + 3 + a] = hi[a][0]; }
}

__global__ void k_grid_setup(SimParams P, double cell, uint32_t cell_cap, const double* __restrict__ part,
                             int nparts, GridDev* G, uint32_t* counters) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    GridDev g;
    g.periodic = P.box.periodic;
    if (P.box.periodic) {
        for (int a = 0; a < 3; ++a) {
            const double L = P.box.edge[a];
            double k = floor(L / cell);
            if (k < 1.0) k = 1.0;
            if (k > 1048576.0) k = 1048576.0;
            g.nc[a] = (int32_t)k;
            g.cw[a] = L / k;
            g.origin[a] = -P.box.half[a];
        }
        // keep the cell table within its capacity by coarsening (still >= 2 maxR)
        while ((double)g.nc[0] * g.nc[1] * g.nc[2] > (double)cell_cap) {
            for (int a = 0; a < 3; ++a) {
                int32_t k = g.nc[a] * 4 / 5;
                if (k < 1) k = 1;
                g.nc[a] = k;
                g.cw[a] = P.box.edge[a] / (double)k;
            }
        }
    } else {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int p = 0; p < nparts; ++p)
            for (int a = 0; a < 3; ++a) {
                lo[a] = fmin(lo[a], part[6 * p + a]);
                hi[a] = fmax(hi[a], part[6 * p + 3 + a]);
            }
        double c = cell;
        for (int tries = 0; tries < 64; ++tries) {
            double tot = 1.0;
            for (int a = 0; a < 3; ++a) {
                double span = hi[a] - lo[a];
                if (!(span >= 0.0)) span = 0.0;
                double k = floor(span / c) + 1.0;
                tot *= k;
            }
            if (tot <= (double)cell_cap) break;
            c *= 1.25;
        }
        for (int a = 0; a < 3; ++a) {
            double span = hi[a] - lo[a];
            if (!(span >= 0.0)) span = 0.0;
            double k = floor(span / c) + 1.0;
            if (k > (double)cell_cap) k = (double)cell_cap;
            g.nc[a] = (int32_t)k;
            g.cw[a] = c;
            g.origin[a] = lo[a] <= hi[a] ? lo[a] : 0.0;
        }
    }
    if ((double)g.nc[0] * (double)g.nc[1] * (double)g.nc[2] > (double)cell_cap) {
        // degenerate extents (e.g. non-finite positions): one cell holds all
        for (int a = 0; a < 3; ++a) { g.nc[a] = 1; g.cw[a] = 1.0; g.origin[a] = 0.0; }
    }
    g.ncells = (uint32_t)g.nc[0] * (uint32_t)g.nc[1] * (uint32_t)g.nc[2];
    for (int a = 0; a < 3; ++a) g.inv_cw[a] = 1.0 / g.cw[a];
    *G = g;
    counters[CNT_CELLS] = g.ncells;
}

// Cell coordinate along axis a, always in [0, nc-1] (NaN maps to 0).  Every
// binning and every lookup go through this one function, so a body's cell is
// consistent; cells are wider than 2*maxR*(1+1e-6), which absorbs the
// rounding of the reciprocal (a touching pair is always <= 1 cell apart).
__device__ __forceinline__ int32_t cell_coord(const GridDev& g, int a, double c) {
    double k = floor((c - g.origin[a]) * g.inv_cw[a]);
    const double m = (double)g.nc[a];
    if (g.periodic && (k < 0.0 || k >= m)) {   // bodies within one box length: one add/subtract
        if (k < 0.0 && k >= -m) k += m;
        else if (k >= m && k < 2.0 * m) k -= m;
        else k = k - m * floor(k / m);
    }
    if (!(k >= 0.0)) k = 0.0;
    if (k > m - 1.0) k = m - 1.0;
    return (int32_t)k;
}

__device__ __forceinline__ uint32_t grid_cell(const GridDev& g, V3 x) {
    const int32_t cx = cell_coord(g, 0, x.x), cy = cell_coord(g, 1, x.y), cz = cell_coord(g, 2, x.z);
    return ((uint32_t)cz * (uint32_t)g.nc[1] + (uint32_t)cy) * (uint32_t)g.nc[0] + (uint32_t)cx;
}

__device__ __forceinline__ uint32_t static_cell(const StaticGrid& sg, V3 x) {
    GridDev g;
    for (int a = 0; a < 3; ++a) {
        g.origin[a] = sg.origin[a]; g.cw[a] = sg.cw[a]; g.inv_cw[a] = sg.inv_cw[a]; g.nc[a] = sg.nc[a];
    }
    g.periodic = 1;
    const int32_t cx = cell_coord(g, 0, x.x), cy = cell_coord(g, 1, x.y), cz = cell_coord(g, 2, x.z);
    return ((uint32_t)cz * (uint32_t)g.nc[1] + (uint32_t)cy) * (uint32_t)g.nc[0] + (uint32_t)cx;
}

__global__ void __launch_bounds__(TPB) k_cell_count(const uint32_t* __restrict__ gridded, uint32_t ngrid,
                                                    const GridDev* __restrict__ Gp, const SortedBody* __restrict__ xp,
                                                    uint32_t* __restrict__ cell_count) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ngrid) return;
    const GridDev g = *Gp;
    const SortedBody p = xp[gridded[t]];
    const uint32_t c = ((uint32_t)cell_coord(g, 2, p.z) * (uint32_t)g.nc[1] + (uint32_t)cell_coord(g, 1, p.y)) *
                           (uint32_t)g.nc[0] + (uint32_t)cell_coord(g, 0, p.x);
    atomicAdd(&cell_count[c], 1u);
}

// cell_start holds exclusive offsets; the atomic bump leaves cell_start[c] at
// the END of cell c, so begin(c) = cell_start[c-1] afterwards.
// Bodies are visited in the previous step's cell order (bodies move little
// per step), so the slot bumps and the sorted-array writes of a wave land
// close together instead of scattering across the whole table.
__global__ void __launch_bounds__(TPB) k_cell_fill(const uint32_t* __restrict__ order, uint32_t ngrid,
                                                   const GridDev* __restrict__ Gp, const uint32_t* __restrict__ wid,
                                                   const WorldDev* __restrict__ W, const SortedBody* __restrict__ xp,
                                                   uint32_t* __restrict__ cell_start, uint32_t* __restrict__ sorted_idx,
                                                   SortedBody* __restrict__ sorted) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ngrid) return;
    const uint32_t i = order[t];
    const SortedBody p = xp[i];
    uint32_t c;
    if (W) {
        const WorldDev& w = W[wid[i]];
        c = w.cell_base + grid_cell(w.grid, V3{p.x, p.y, p.z});
    } else {
        c = grid_cell(*Gp, V3{p.x, p.y, p.z});
    }
    const uint32_t slot = atomicAdd(&cell_start[c], 1u);
    sorted_idx[slot] = i;
    sorted[slot] = p;
}

// Pair filters of src/world.cpp:513-516 (world.hpp:94-103).  With island
// sleeping, a sleeping body counts with its stored mass here: it may be woken
// earlier in the same step's pair order, after which the reference keeps its
// pairs (the exact decision is taken after the narrow phase, k_sleep_fate).
__device__ __forceinline__ bool pair_passes(const BodiesDev& B, uint32_t i, uint32_t j) {
    if (B.ev_filter) {   // PhysicsWorld::processContacts (src/physics_world.cpp:261-266)
        const bool any_sensor = ((B.flags[i] | B.flags[j]) & NE_FLAG_SENSOR) != 0;
        if (!any_sensor && B.im[i] + B.im[j] <= 0.0) return false;
        const uint32_t li = B.layer[i], lj = B.layer[j];
        if (!((B.mask[i] & (1u << (lj & 31u))) && (B.mask[j] & (1u << (li & 31u))))) return false;
        for (uint32_t k = B.nc_off[i]; k < B.nc_off[i + 1]; ++k)
            if (B.nc_idx[k] == j) return false;
        return true;
    }
    double mi = B.im[i], mj = B.im[j];
    if (B.sleep_on) {
        if (B.flags[i] & NE_FLAG_SLEEPING) mi = fmax(mi, B.ims[i]);
        if (B.flags[j] & NE_FLAG_SLEEPING) mj = fmax(mj, B.ims[j]);
    }
    if (mi + mj <= 0.0) return false;
    const uint32_t li = B.layer[i], lj = B.layer[j];
    if (!((B.mask[i] & (1u << (lj & 31u))) && (B.mask[j] & (1u << (li & 31u))))) return false;
    for (uint32_t k = B.nc_off[i]; k < B.nc_off[i + 1]; ++k)
        if (B.nc_idx[k] == j) return false;
    if ((B.flags[i] | B.flags[j]) & NE_FLAG_SENSOR) return false;
    return true;
}

// Conservative bounding-sphere overlap under the minimum image.  Only decides
// which pairs reach the narrow phase (which recomputes the exact minimum
// image), so it may use 1/L: a rounding-chosen image differs from the exact
// one only for |d| ~ L/2, far from touching.
__device__ __forceinline__ bool bounds_touch(const Box3& box, V3 a, double ra, V3 b, double rb) {
    V3 d = a - b;
    if (box.periodic) {
        d.x -= box.edge[0] * floor(d.x * box.inv_edge[0] + 0.5);
        d.y -= box.edge[1] * floor(d.y * box.inv_edge[1] + 0.5);
        d.z -= box.edge[2] * floor(d.z * box.inv_edge[2] + 0.5);
    }
    double s = (ra + rb) * (1.0 + 1e-9) + 1e-9;
    return norm2(d) <= s * s;
}

// Distinct neighbour offsets along one axis (fewer than 3 cells => dedupe).
__device__ __forceinline__ int axis_neighbours(const GridDev& g, int a, int32_t c, int32_t out[3]) {
    const int32_t n = g.nc[a];
    int k = 0;
    for (int d = -1; d <= 1; ++d) {
        int32_t v = c + d;
        if (g.periodic) {
            v = ((v % n) + n) % n;
        } else if (v < 0 || v >= n) {
            continue;
        }
        bool dup = false;
        for (int m = 0; m < k; ++m) dup |= (out[m] == v);
        if (!dup) out[k++] = v;
    }
    return k;
}

// Neighbour coordinates along one axis: v[d] for offset d-1, ok[d] false
// when outside a non-periodic grid or a repeat of an earlier offset (fewer
// than 3 cells along a periodic axis).
__device__ __forceinline__ void axis_offsets(const GridDev& g, int a, int32_t c, int32_t v[3], bool ok[3]) {
    const int32_t n = g.nc[a];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int32_t x = c + d - 1;   // in [-1, n]
        bool good = true;
        if (g.periodic) x = x < 0 ? x + n : (x >= n ? x - n : x);
        else good = x >= 0 && x < n;
        v[d] = x;
        ok[d] = good;
    }
    if (g.periodic) {
        ok[1] = ok[1] && v[1] != v[0];
        ok[2] = ok[2] && v[2] != v[0] && v[2] != v[1];
    }
}

// Visits every candidate partner j > i of gridded body i: bodies in the
// 27-cell neighbourhood and the grid-excluded bodies, keeping pairs whose
// bounding spheres overlap and that pass the reference's filters.  All
// neighbour-range bounds are loaded before any range is walked (the walk is
// latency-bound); when the three x-neighbours are consecutive cells a row is
// one range.
// The grid, cell offset, box and grid-excluded range that body i sees.
struct CellView {
    GridDev g;
    uint32_t base;
    Box3 box;
    uint32_t e0, e1;
};

__device__ __forceinline__ CellView cell_view(const BodiesDev& B, const SimParams& P, const GridDev* Gp, uint32_t i,
                                              uint32_t nexcl) {
    CellView v;
    if (B.W) {
        const WorldDev& w = B.W[B.wid[i]];
        v.g = w.grid; v.base = w.cell_base; v.box = w.box; v.e0 = w.excl_begin; v.e1 = w.excl_end;
    } else {
        v.g = *Gp; v.base = 0; v.box = P.box; v.e0 = 0; v.e1 = nexcl;
    }
    return v;
}

template <class F>
__device__ __forceinline__ void for_each_candidate(const BodiesDev& B, const CellView& cv,
                                                   uint32_t i, const SortedBody& me,
                                                   const uint32_t* __restrict__ sorted_idx,
                                                   const SortedBody* __restrict__ sorted,
                                                   const uint32_t* __restrict__ cell_end,
                                                   const uint32_t* __restrict__ excl, F&& emit) {
    const GridDev& g = cv.g;
    const uint32_t base = cv.base;
    const V3 xi{me.x, me.y, me.z};
    int32_t vx[3], vy[3], vz[3];
    bool ox[3], oy[3], oz[3];
    const int32_t cx = cell_coord(g, 0, me.x);
    axis_offsets(g, 0, cx, vx, ox);
    axis_offsets(g, 1, cell_coord(g, 1, me.y), vy, oy);
    axis_offsets(g, 2, cell_coord(g, 2, me.z), vz, oz);
    const uint32_t nc0 = (uint32_t)g.nc[0], nc1 = (uint32_t)g.nc[1];
    auto walk = [&](uint32_t beg, uint32_t end) {
        for (uint32_t t = beg; t < end; ++t) {
            const uint32_t j = sorted_idx[t];
            const SortedBody o = sorted[t];
            if (j <= i) continue;
            if (!bounds_touch(cv.box, xi, me.r, V3{o.x, o.y, o.z}, o.r)) continue;
            if (!pair_passes(B, i, j)) continue;
            emit(j);
        }
    };
    if (cx >= 1 && cx + 1 < g.nc[0]) {
        // x-interior cell: each (y, z) row of three cells is one contiguous
        // range; the nine ranges are walked as one flattened sequence so a
        // wave waits for one load per candidate, not one per row.
        uint32_t rb[9], rl[9];
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const int b = r % 3, c = r / 3;
            const bool ok = oy[b] && oz[c];
            const uint32_t c0 = base + ((uint32_t)vz[c] * nc1 + (uint32_t)vy[b]) * nc0 + (uint32_t)(cx - 1);
            const uint32_t e = ok ? cell_end[c0 + 2] : 0u;
            const uint32_t bgn = (ok && c0 > 0) ? cell_end[c0 - 1] : 0u;
            rb[r] = bgn;
            rl[r] = ok ? e - bgn : 0u;
        }
        uint32_t total = 0;
#pragma unroll
        for (int r = 0; r < 9; ++r) total += rl[r];
        for (uint32_t u = 0; u < total; ++u) {
            uint32_t off = u, t = 0;
#pragma unroll
            for (int r = 0; r < 9; ++r) {        // select the row holding element u
                const bool here = off < rl[r] && t == 0;
                t = here ? rb[r] + off + 1 : t;  // +1 keeps 0 as "not found"
                off = (t == 0) ? off - rl[r] : off;
            }
            walk(t - 1, t);
        }
    } else {
        uint32_t cb[27], ce[27];
#pragma unroll
        for (int k = 0; k < 27; ++k) {
            const int a = k % 3, b = (k / 3) % 3, c = k / 9;
            const bool ok = ox[a] && oy[b] && oz[c];
            const uint32_t cell = base + ((uint32_t)vz[c] * nc1 + (uint32_t)vy[b]) * nc0 + (uint32_t)vx[a];
            cb[k] = (ok && cell) ? cell_end[cell - 1] : 0u;
            ce[k] = ok ? cell_end[cell] : 0u;
        }
#pragma unroll
        for (int k = 0; k < 27; ++k) walk(cb[k], ce[k]);
    }
    for (uint32_t e = cv.e0; e < cv.e1; ++e) {
        const uint32_t j = excl[e];
        if (j <= i) continue;
        if (!bounds_touch(cv.box, xi, me.r, ld3(B.x0, B.x1, B.x2, j), B.bR[j])) continue;
        if (!pair_passes(B, i, j)) continue;
        emit(j);
    }
}

constexpr uint32_t ROW_SLOTS = 16;   // candidates kept per body by the single walk

// One walk per gridded body (threads in cell-sorted order for locality):
// counts its row (i, j>i) and keeps the first ROW_SLOTS partners in scratch.
__global__ void __launch_bounds__(TPB) k_cand_walk(BodiesDev B, SimParams P, const GridDev* __restrict__ Gp,
                                                   uint32_t ngrid, const uint32_t* __restrict__ sorted_idx,
                                                   const SortedBody* __restrict__ sorted,
                                                   const uint32_t* __restrict__ cell_end,
                                                   const uint32_t* __restrict__ excl, uint32_t nexcl,
                                                   uint32_t* __restrict__ row_count, uint32_t* __restrict__ row_scratch,
                                                   uint32_t* __restrict__ slot_of) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ngrid) return;
    const uint32_t i = sorted_idx[s];
    const SortedBody me = sorted[s];
    const CellView cv = cell_view(B, P, Gp, i, nexcl);
    uint32_t cnt = 0;
    uint32_t* row = row_scratch + (size_t)i * ROW_SLOTS;
    for_each_candidate(B, cv, i, me, sorted_idx, sorted, cell_end, excl, [&](uint32_t j) {
        if (cnt < ROW_SLOTS) row[cnt] = j;
        ++cnt;
    });
    row_count[i] = cnt;
    if (cnt > ROW_SLOTS) slot_of[i] = s;   // the emit kernel walks this row again
}

// LDS-staged neighbourhood walk.  One workgroup per brick of BX x BY x BZ
// cells stages the bodies of the brick and its one-cell halo into LDS (one
// coalesced pass over the cell-sorted array) and walks every interior body's
// 27-cell neighbourhood from LDS.  Emits exactly what k_cand_walk emits.  Used
// when every axis has >= brick + 2 cells (so halo cells never alias);
// a brick whose halo holds more bodies than the LDS pool walks from global
// memory instead.
constexpr int BRICK_X = 8, BRICK_Y = 8, BRICK_Z = 4;
constexpr int HALO_X = BRICK_X + 2, HALO_Y = BRICK_Y + 2, HALO_Z = BRICK_Z + 2;
constexpr int HALO_CELLS = HALO_X * HALO_Y * HALO_Z;   // 600
constexpr int BRICK_POOL = 1024;                       // bodies staged per brick
constexpr int BRICK_THREADS = 256;

__global__ void __launch_bounds__(BRICK_THREADS) k_cand_brick(BodiesDev B, SimParams P, const GridDev* __restrict__ Gp,
                                                              const uint32_t* __restrict__ sorted_idx,
                                                              const SortedBody* __restrict__ sorted,
                                                              const uint32_t* __restrict__ cell_end,
                                                              const uint32_t* __restrict__ excl, uint32_t nexcl,
                                                              uint32_t* __restrict__ row_count,
                                                              uint32_t* __restrict__ row_scratch,
                                                              uint32_t* __restrict__ slot_of) {
    __shared__ uint32_t cbeg[HALO_CELLS];
    __shared__ uint32_t coff[HALO_CELLS + 1];
    __shared__ SortedBody pool[BRICK_POOL];
    __shared__ uint32_t pidx[BRICK_POOL];
    __shared__ uint16_t inner[BRICK_POOL];
    __shared__ uint8_t cint[HALO_CELLS];
    __shared__ uint32_t scan_tmp[PRIM_WAVES + 1];
    __shared__ uint32_t n_inner;
    const GridDev g = *Gp;
    const int nbx = (g.nc[0] + BRICK_X - 1) / BRICK_X, nby = (g.nc[1] + BRICK_Y - 1) / BRICK_Y,
              nbz = (g.nc[2] + BRICK_Z - 1) / BRICK_Z;
    const uint32_t nbricks = (uint32_t)nbx * (uint32_t)nby * (uint32_t)nbz;
    // XCD-aware order: blocks are dealt round-robin to the 8 XCDs; give each
    // XCD a contiguous run of bricks so neighbouring halos share its L2.
    const uint32_t per_xcd = (nbricks + 7u) / 8u;
    for (uint32_t blk = blockIdx.x; blk < per_xcd * 8u; blk += gridDim.x) {
        const uint32_t brick = (blk & 7u) * per_xcd + (blk >> 3);
        if (brick >= nbricks) continue;   // uniform per block
        const int bx = (int)(brick % (uint32_t)nbx), by = (int)((brick / (uint32_t)nbx) % (uint32_t)nby),
                  bz = (int)(brick / ((uint32_t)nbx * (uint32_t)nby));
        // 1. halo cell ranges.  Periodic: every halo position wraps to a
        // distinct cell (nc >= brick + 2); positions of a partial brick past
        // the grid edge are the wrapped neighbours, not interior cells.
        for (int hc = threadIdx.x; hc < HALO_CELLS; hc += BRICK_THREADS) {
            const int hx = hc % HALO_X, hy = (hc / HALO_X) % HALO_Y, hz = hc / (HALO_X * HALO_Y);
            const int ux = bx * BRICK_X + hx - 1, uy = by * BRICK_Y + hy - 1, uz = bz * BRICK_Z + hz - 1;
            int cx = ux, cy = uy, cz = uz;
            bool ok = true;
            if (g.periodic) {
                cx = (cx + g.nc[0]) % g.nc[0];
                cy = (cy + g.nc[1]) % g.nc[1];
                cz = (cz + g.nc[2]) % g.nc[2];
            } else {
                ok = cx >= 0 && cx < g.nc[0] && cy >= 0 && cy < g.nc[1] && cz >= 0 && cz < g.nc[2];
            }
            cint[hc] = (hx >= 1 && hx <= BRICK_X && hy >= 1 && hy <= BRICK_Y && hz >= 1 && hz <= BRICK_Z &&
                        ux < g.nc[0] && uy < g.nc[1] && uz < g.nc[2]) ? 1 : 0;
            uint32_t b0 = 0, b1 = 0;
            if (ok) {
                const uint32_t cell = ((uint32_t)cz * (uint32_t)g.nc[1] + (uint32_t)cy) * (uint32_t)g.nc[0] + (uint32_t)cx;
                b0 = cell ? cell_end[cell - 1] : 0u;
                b1 = cell_end[cell];
            }
            cbeg[hc] = b0;
            coff[hc] = b1 - b0;   // count, scanned below
        }
        if (threadIdx.x == 0) n_inner = 0;
        __syncthreads();
        // 2. exclusive scan of the 600 counts (3 per thread, in order)
        {
            constexpr int PER = (HALO_CELLS + BRICK_THREADS - 1) / BRICK_THREADS;   // 3
            uint32_t loc[PER];
            uint32_t sum = 0;
            for (int k = 0; k < PER; ++k) {
                const int hc = threadIdx.x * PER + k;
                loc[k] = hc < HALO_CELLS ? coff[hc] : 0u;
                sum += loc[k];
            }
            uint32_t tot;
            uint32_t ex = block_exclusive_scan(sum, &tot, scan_tmp);
            for (int k = 0; k < PER; ++k) {
                const int hc = threadIdx.x * PER + k;
                if (hc < HALO_CELLS) coff[hc] = ex;
                ex += loc[k];
            }
            if (threadIdx.x == 0) coff[HALO_CELLS] = tot;
            __syncthreads();
        }
        const uint32_t H = coff[HALO_CELLS];
        const bool staged = H <= (uint32_t)BRICK_POOL;
        // 3. stage halo bodies (consecutive slots of a cell are consecutive in memory)
        if (staged) {
            for (uint32_t h = threadIdx.x; h < H; h += BRICK_THREADS) {
                int lo = 0, hi = HALO_CELLS - 1;   // last cell with coff <= h
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (coff[mid] <= h) lo = mid; else hi = mid - 1;
                }
                const uint32_t src = cbeg[lo] + (h - coff[lo]);
                pool[h] = sorted[src];
                pidx[h] = sorted_idx[src];
                if (cint[lo]) inner[atomicAdd(&n_inner, 1u)] = (uint16_t)h;
            }
        }
        __syncthreads();
        if (staged) {
            // 4. walk every interior body's neighbourhood inside LDS
            const uint32_t ni = n_inner;
            for (uint32_t t = threadIdx.x; t < ni; t += BRICK_THREADS) {
                const uint32_t h = inner[t];
                int lo = 0, hi = HALO_CELLS - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (coff[mid] <= h) lo = mid; else hi = mid - 1;
                }
                const int hx = lo % HALO_X, hy = (lo / HALO_X) % HALO_Y, hz = lo / (HALO_X * HALO_Y);
                const uint32_t i = pidx[h];
                const SortedBody me = pool[h];
                const V3 xi{me.x, me.y, me.z};
                uint32_t cnt = 0;
                uint32_t* row = row_scratch + (size_t)i * ROW_SLOTS;
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy) {
                        const int c0 = ((hz + dz) * HALO_Y + (hy + dy)) * HALO_X + (hx - 1);
                        const uint32_t u0 = coff[c0], u1 = coff[c0 + 3];   // 3 consecutive x-cells
                        for (uint32_t u = u0; u < u1; ++u) {
                            const uint32_t j = pidx[u];
                            if (j <= i) continue;
                            const SortedBody o = pool[u];
                            if (!bounds_touch(P.box, xi, me.r, V3{o.x, o.y, o.z}, o.r)) continue;
                            if (!pair_passes(B, i, j)) continue;
                            if (cnt < ROW_SLOTS) row[cnt] = j;
                            ++cnt;
                        }
                    }
                for (uint32_t e = 0; e < nexcl; ++e) {
                    const uint32_t j = excl[e];
                    if (j <= i) continue;
                    if (!bounds_touch(P.box, xi, me.r, ld3(B.x0, B.x1, B.x2, j), B.bR[j])) continue;
                    if (!pair_passes(B, i, j)) continue;
                    if (cnt < ROW_SLOTS) row[cnt] = j;
                    ++cnt;
                }
                row_count[i] = cnt;
                if (cnt > ROW_SLOTS) {   // the emit kernel walks this row again from its sorted slot
                    int lo2 = 0, hi2 = HALO_CELLS - 1;
                    while (lo2 < hi2) {
                        const int mid = (lo2 + hi2 + 1) >> 1;
                        if (coff[mid] <= h) lo2 = mid; else hi2 = mid - 1;
                    }
                    slot_of[i] = cbeg[lo2] + (h - coff[lo2]);
                }
            }
        } else {
            // crowded brick: walk its interior bodies from global memory
            for (int hc = 0; hc < HALO_CELLS; ++hc) {
                if (!cint[hc]) continue;
                const uint32_t b0 = cbeg[hc], nb = coff[hc + 1] - coff[hc];
                for (uint32_t t = threadIdx.x; t < nb; t += BRICK_THREADS) {
                    const uint32_t sidx = b0 + t;
                    const uint32_t i = sorted_idx[sidx];
                    uint32_t cnt = 0;
                    uint32_t* row = row_scratch + (size_t)i * ROW_SLOTS;
                    const CellView cv = cell_view(B, P, Gp, i, nexcl);
                    for_each_candidate(B, cv, i, sorted[sidx], sorted_idx, sorted, cell_end, excl,
                                       [&](uint32_t j) {
                                           if (cnt < ROW_SLOTS) row[cnt] = j;
                                           ++cnt;
                                       });
                    row_count[i] = cnt;
                    if (cnt > ROW_SLOTS) slot_of[i] = sidx;
                }
            }
        }
        __syncthreads();
    }
}

// Cell-major walk: one thread per grid cell (consecutive lanes = consecutive
// cells along x), so the neighbour-row bounds a wave loads are contiguous and
// the loads coalesce.  Each thread walks the neighbourhood for every body in
// its cell; the rows it produces are exactly k_cand_walk's.
__global__ void __launch_bounds__(TPB) k_cand_cells(BodiesDev B, SimParams P, const GridDev* __restrict__ Gp,
                                                    const uint32_t* __restrict__ sorted_idx,
                                                    const SortedBody* __restrict__ sorted,
                                                    const uint32_t* __restrict__ cell_end,
                                                    const uint32_t* __restrict__ excl, uint32_t nexcl,
                                                    uint32_t* __restrict__ row_count, uint32_t* __restrict__ row_scratch,
                                                    uint32_t* __restrict__ slot_of) {
    const GridDev g = *Gp;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < g.ncells; c += gridDim.x * blockDim.x) {
        const uint32_t cb = c ? cell_end[c - 1] : 0u, ce = cell_end[c];
        if (cb == ce) continue;
        for (uint32_t s = cb; s < ce; ++s) {
            const uint32_t i = sorted_idx[s];
            uint32_t cnt = 0;
            uint32_t* row = row_scratch + (size_t)i * ROW_SLOTS;
            const CellView cv = cell_view(B, P, Gp, i, nexcl);
            for_each_candidate(B, cv, i, sorted[s], sorted_idx, sorted, cell_end, excl, [&](uint32_t j) {
                if (cnt < ROW_SLOTS) row[cnt] = j;
                ++cnt;
            });
            row_count[i] = cnt;
            if (cnt > ROW_SLOTS) slot_of[i] = s;
        }
    }
}

// Writes each gridded body's row at its offset in (i, j) order: from the
// scratch when it fit, else by walking the neighbourhood again; then sorts
// the row by j (rows are short).
__global__ void __launch_bounds__(TPB) k_cand_emit(BodiesDev B, SimParams P, const GridDev* __restrict__ Gp,
                                                   const uint32_t* __restrict__ slot_of,
                                                   const uint32_t* __restrict__ sorted_idx,
                                                   const SortedBody* __restrict__ sorted,
                                                   const uint32_t* __restrict__ cell_end,
                                                   const uint32_t* __restrict__ excl, uint32_t nexcl,
                                                   const uint32_t* __restrict__ row_count,
                                                   const uint32_t* __restrict__ row_scratch,
                                                   const uint32_t* __restrict__ row_start, uint32_t pcap,
                                                   uint32_t* __restrict__ cand_i, uint32_t* __restrict__ cand_j) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    const uint32_t cnt = row_count[i];
    if (cnt == 0 || (B.flags[i] & FLAG_GRID_EXCLUDED)) return;   // empty, or a grid-excluded body's row
    const uint32_t base = row_start[i];
    uint32_t m = 0;
    if (cnt <= ROW_SLOTS) {
        const uint32_t* row = row_scratch + (size_t)i * ROW_SLOTS;
        for (; m < cnt && base + m < pcap; ++m) { cand_i[base + m] = i; cand_j[base + m] = row[m]; }
    } else {
        const CellView cv = cell_view(B, P, Gp, i, nexcl);
        const uint32_t s = slot_of[i];
        for_each_candidate(B, cv, i, sorted[s], sorted_idx, sorted, cell_end, excl, [&](uint32_t j) {
            if (base + m < pcap) { cand_i[base + m] = i; cand_j[base + m] = j; }
            ++m;
        });
        if (base + m > pcap) m = base < pcap ? pcap - base : 0u;
    }
    for (uint32_t a = 1; a < m; ++a) {
        uint32_t key = cand_j[base + a];
        uint32_t b = a;
        while (b > 0 && cand_j[base + b - 1] > key) { cand_j[base + b] = cand_j[base + b - 1]; --b; }
        cand_j[base + b] = key;
    }
}

// Rows of grid-excluded bodies X: (X, j) for every j > X of X's world, in j
// order.  Excluded body e owns one flag per body of its world (every body
// with one world), set when the pair survives the bound test and the filters.
struct ExclRow {
    uint32_t begin, end;   // bodies of X's world
    size_t base;           // first flag of the row
};

__device__ __forceinline__ ExclRow excl_row(const BodiesDev& B, uint32_t e, uint32_t X) {
    if (B.W) {
        const WorldDev& w = B.W[B.wid[X]];
        const uint32_t m = w.body_end - w.body_begin;
        return ExclRow{w.body_begin, w.body_end, (size_t)w.flag_base + (size_t)(e - w.excl_begin) * m};
    }
    return ExclRow{0u, B.n, (size_t)e * B.n};
}

__global__ void __launch_bounds__(TPB) k_excl_flags(BodiesDev B, SimParams P, const uint32_t* __restrict__ excl,
                                                    uint32_t* __restrict__ flags) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t e = blockIdx.y;
    const uint32_t X = excl[e];
    const ExclRow r = excl_row(B, e, X);
    const uint32_t j = r.begin + u;
    if (j >= r.end) return;
    uint32_t f = 0;
    if (j > X && bounds_touch(box_of(B, P, X), ld3(B.x0, B.x1, B.x2, X), B.bR[X], ld3(B.x0, B.x1, B.x2, j), B.bR[j]) &&
        pair_passes(B, X, j))
        f = 1;
    flags[r.base + u] = f;
}

__global__ void k_excl_rowcount(BodiesDev B, const uint32_t* __restrict__ excl, uint32_t nexcl,
                                const uint32_t* __restrict__ ex_off, uint32_t* __restrict__ row_count) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nexcl) return;
    const uint32_t X = excl[e];
    const ExclRow r = excl_row(B, e, X);
    row_count[X] = ex_off[r.base + (r.end - r.begin)] - ex_off[r.base];
}

__global__ void __launch_bounds__(TPB) k_excl_write(BodiesDev B, const uint32_t* __restrict__ excl,
                                                    const uint32_t* __restrict__ flags, const uint32_t* __restrict__ ex_off,
                                                    const uint32_t* __restrict__ row_start, uint32_t pcap,
                                                    uint32_t* __restrict__ cand_i, uint32_t* __restrict__ cand_j) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t e = blockIdx.y;
    const uint32_t X = excl[e];
    const ExclRow r = excl_row(B, e, X);
    const uint32_t j = r.begin + u;
    if (j >= r.end) return;
    const size_t f = r.base + u;
    if (!flags[f]) return;
    const uint32_t pos = row_start[X] + (ex_off[f] - ex_off[r.base]);
    if (pos < pcap) { cand_i[pos] = X; cand_j[pos] = j; }
}

__global__ void k_clamp_count(const uint32_t* __restrict__ raw, uint32_t cap, uint32_t* counters, int slot_eff,
                              int slot_raw, uint32_t fault_bit) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t r = *raw;
    counters[slot_raw] = r;
    if (r > cap) { counters[CNT_FAULT] |= fault_bit; r = cap; }
    counters[slot_eff] = r;
}

// ---------------------------------------------------------------------------
// Narrow phase: one thread per candidate pair, gatherContacts of
// src/world.cpp:122-129 (detectContact + box-box manifold).
// ---------------------------------------------------------------------------

// Contacts of one candidate pair into out[0..27] (<= 4 x (n3, pt3, depth));
// returns how many.  ws.overflow is set when EPA outgrew its workspace; with
// ws.defer, a convex pair that GJK finds overlapping returns NARROW_DEFERRED.
constexpr uint32_t NARROW_DEFERRED = 0xFFFFFFFFu;
template <bool SPHERES_ONLY>
__device__ __forceinline__ uint32_t narrow_pair(const BodiesDev& B, const SimParams& P, uint32_t i, uint32_t j,
                                                double* out, EpaWS& ws, uint32_t* fault) {
    const Box3& box = box_of(B, P, i);   // i and j share a world
    uint32_t nc = 0;
    const int si = SPHERES_ONLY ? (int)SH_SPHERE : B.shape[i];
    const int sj = SPHERES_ONLY ? (int)SH_SPHERE : B.shape[j];
    if (!SPHERES_ONLY && (si == SH_MESH || sj == SH_MESH)) return 0;   // the mesh kernels take these pairs
    if (SPHERES_ONLY || (si == SH_SPHERE && sj == SH_SPHERE)) {
        Contact c = sphere_sphere(box, ld3(B.x0, B.x1, B.x2, i), B.rad[i], ld3(B.x0, B.x1, B.x2, j), B.rad[j]);
        if (c.hit) {
            out[0] = c.normal.x; out[1] = c.normal.y; out[2] = c.normal.z;
            out[3] = c.point.x; out[4] = c.point.y; out[5] = c.point.z; out[6] = c.overlap;
            nc = 1;
        }
    } else if (!SPHERES_ONLY) {
        const Geom a = load_geom(B, i), b = load_geom(B, j);
        Contact c = detect_contact(box, a, b, ws);
        if (c.hit == 2) return NARROW_DEFERRED;   // GJK overlap: EPA (and a box pair's manifold) in the EPA tiers
        if (c.hit) {   // an analytic contact: two boxes only ever reach GJK, so no manifold here
            out[0] = c.normal.x; out[1] = c.normal.y; out[2] = c.normal.z;
            out[3] = c.point.x; out[4] = c.point.y; out[5] = c.point.z; out[6] = c.overlap;
            nc = 1;
        }
    }
    return nc;
}

// Pairs whose EPA polytope outgrows the per-thread and per-wave workspaces
// (near-degenerate contacts, where the reference's vectors grow to tens of
// thousands of faces) are queued for k_narrow_big, which reruns each with one
// block, a workspace in global memory and the horizon built in parallel.
// Both capacities grow when exceeded (the step is then run again).
constexpr uint32_t EPA_BIG_PAIRS = 64;    // initial queue capacity
constexpr int EPA_BIG_FACES_MAX = 1 << 17;   // default cap: beyond it a run stops at its closest face
// Initial faces per queued pair: the default cap (about 6 MiB a slot, 64
// slots).  Starting at 4,096 and growing x4 cost three reruns of the call at
// the first run that grows past each size: the 1024-world batch's first 30
// steps, whose overlapping start queues tens of such runs a step.
constexpr int EPA_BIG_FACES = EPA_BIG_FACES_MAX;
constexpr int EPA_KEYS = EPA_MAXV * EPA_MAXV;   // directed edges u * EPA_MAXV + w of one polytope
constexpr int BIG_NW = 8;                       // waves per large-workspace run

constexpr uint32_t MESH_ENTRY = 0x80000000u;   // queue entry is a mesh triangle candidate, not a pair

struct EpaBig {
    uint32_t* list;   // queued pair indices (or MESH_ENTRY | mesh candidate index)
    uint32_t cap;     // queue slots
    int faces_per;    // faces per slot
    int faces_max;    // the cap faces_per grows to; a run that needs more stops there
    Face* faces;
    uint16_t* occ;    // 3 per face: directed edges of the visible faces, in order
};

// Narrow phase, sorted by pair type: k_narrow computes the analytic pairs
// and runs GJK on the convex ones; convex pairs that overlap are queued for
// the EPA tiers (ne_epa_lds.hpp: per-lane polytopes in LDS, then groups of
// lanes per run, then a wave, then a block per run), so waves doing EPA hold
// only EPA runs of like length.  Their contacts are the same as computed
// inline.
template <bool SPHERES_ONLY>
__global__ void __launch_bounds__(TPB) k_narrow(BodiesDev B, SimParams P, uint32_t* __restrict__ counters,
                                                const uint32_t* __restrict__ cand_i, const uint32_t* __restrict__ cand_j,
                                                uint32_t* __restrict__ ncont, double* __restrict__ ct,
                                                uint32_t* __restrict__ epa_q, uint32_t epa_cap, EpaTets tets) {
    const uint32_t np = counters[CNT_P];
    uint32_t* fault = counters + CNT_FAULT;
    if (!SPHERES_ONLY && blockIdx.x == 0 && threadIdx.x == 0) counters[CNT_EPALONG] = 0u;   // the EPA tiers run after
    EpaWS ws{};
    ws.tets = tets;
    for (uint32_t p0 = blockIdx.x * blockDim.x; p0 < np; p0 += gridDim.x * blockDim.x) {
        const uint32_t p = p0 + threadIdx.x;
        uint32_t nc = 0;
        ws.pair = p;
        if (p < np) nc = narrow_pair<SPHERES_ONLY>(B, P, cand_i[p], cand_j[p], ct + (size_t)p * 28, ws, fault);
        if (!SPHERES_ONLY) {
            const bool defer = p < np && nc == NARROW_DEFERRED;
            const uint32_t slot = wave_add1(counters, CNT_EPAQ, defer);
            if (defer) {
                if (slot < epa_cap) epa_q[slot] = p;
                else atomicOr(fault, NE_FAULT_PAIRS);
                nc = 0;
            }
        }
        if (p < np) ncont[p] = nc;
    }
}

// EPA of one pair by one wave (src/epa.hpp:71-150, the same steps as
// epa_contact), polytope and horizon list in LDS.  Each iteration the 64
// lanes find the first closest face (a min/first-index reduction), test the
// faces for visibility and compact the kept ones in rounds of 64 (order is
// kept by ballot ranks), stage each round's visible faces for lane 0, which
// runs the horizon list, and finally build the new faces from the live
// horizon entries in parallel.  All lanes hold identical copies of the
// uniform values (best face, support point).
template <class Hz>
__device__ Contact epa_wave(const Geom& A, const Geom& B, V3 shift, const SupV tet[4], SupV* verts, Face* faces,
                            int maxf, Hz& hz, int* stage, int* shared, bool truncate, bool& overflow) {
    const int lane = threadIdx.x;
    if (lane == 0)
        for (int k = 0; k < 4; ++k) verts[k] = tet[k];
    __syncthreads();
    int nv = 4, nf = 4;
    if (lane < 4) {
        const int fi[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
        faces[lane] = make_face(verts, fi[lane][0], fi[lane][1], fi[lane][2]);
    }
    __syncthreads();
    Face best = faces[0];
    for (int iter = 0; iter < 64; ++iter) {
        // first face of minimal distance
        double bd = 0.0;
        int bi = -1;
#pragma unroll 8
        for (int i = lane; i < nf; i += 64) {
            const double d = faces[i].dist;
            if (bi < 0 || d < bd) { bd = d; bi = i; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double od = __shfl_xor(bd, off);
            const int oi = __shfl_xor(bi, off);
            const bool take = oi >= 0 && (bi < 0 || od < bd || (od == bd && oi < bi));
            bd = take ? od : bd;
            bi = take ? oi : bi;
        }
        best = faces[bi];
        const SupV p = cso_support(A, B, shift, best.n);
        const double d = dot(p.v, best.n);
        if (d - best.dist < 1e-7) break;
        const int pidx = nv;
        if (lane == 0) verts[nv] = p;
        ++nv;
        if (lane == 0) { hz.begin(); shared[0] = 0; }
        __syncthreads();
        // visibility rounds: compact kept faces, hand visible ones to lane 0
        int nk = 0;
        // the next chunk is loaded while this one is compacted: compaction
        // writes only below this chunk's end, so it never overwrites it
        bool in_next = lane < nf;
        Face f_next{};
        if (in_next) f_next = faces[lane];
        for (int base = 0; base < nf; base += 64) {
            const bool in = in_next;
            const Face f = f_next;
            in_next = base + 64 + lane < nf;
            if (in_next) f_next = faces[base + 64 + lane];
            bool vis = false;
            if (in) vis = dot(f.n, p.v - verts[f.a].v) > 1e-12;
            const uint64_t km = __ballot(in && !vis), vm = __ballot(vis);
            const uint64_t below = lanemask_lt();
            if (in && !vis) faces[nk + (int)__popcll(km & below)] = f;
            if (vis) {
                const int r = (int)__popcll(vm & below);
                stage[3 * r] = f.a; stage[3 * r + 1] = f.b; stage[3 * r + 2] = f.c;
            }
            nk += (int)__popcll(km);
            __syncthreads();
            if (lane == 0 && !shared[0]) {
                const int nvis = (int)__popcll(vm);
                for (int k = 0; k < nvis; ++k) {
                    const int a = stage[3 * k], b = stage[3 * k + 1], c = stage[3 * k + 2];
                    if (!hz.add(a, b) || !hz.add(b, c) || !hz.add(c, a)) { shared[0] = 1; break; }
                }
            }
            __syncthreads();
            if (shared[0]) break;
        }
        if (lane == 0) { shared[1] = hz.entries(); shared[2] = hz.alive(); }
        __syncthreads();
        const int ne = shared[1], nalive = shared[2];
        if (shared[0] || nk + nalive > maxf) {
            overflow = true;
            break;
        }
        __threadfence_block();
        // new faces from the live horizon entries, in insertion order
        int w = nk;
        for (int base = 0; base < ne; base += 64) {
            const int e = base + lane;
            int u = 0, v = 0;
            const bool live = e < ne && hz.get(e, u, v);
            const uint64_t lm = __ballot(live);
            if (live) faces[w + (int)__popcll(lm & lanemask_lt())] = make_face(verts, u, v, pidx);
            w += (int)__popcll(lm);
        }
        nf = w;
        __syncthreads();
        if (nf == 0) break;
    }
    (void)truncate;
    return epa_result(verts, best);
}

// The queue's pair entries, counted into CNT_EPABIG_PAIRS before the mesh
// kernels append theirs: this kernel may run beside them on a second stream.
__global__ void k_snap_counter(uint32_t* counters, int from, int to) {
    if (blockIdx.x == 0 && threadIdx.x == 0) counters[to] = counters[from];
}

// ---- Large-workspace EPA: one block of BIG_NW waves per run --------------
//
// The queued runs are the near-degenerate ones whose polytope grows to tens
// of thousands of faces and whose horizon holds thousands of edges; one wave
// scanning them, with one lane building the horizon, is latency bound.  Here
// a block shares the scans, and the horizon list is built in parallel with
// the same result as the reference's find + erase on a vector
// (src/epa.hpp:95-112):
//
// Edges of different undirected pairs {u, w} never interact, and the list's
// entries of one pair always point the same way (an arriving reverse edge
// removes one).  An arriving reverse edge removes the oldest entry, so after
// all visible faces are processed, with P occurrences of (u, w) and N of
// (w, u), the survivors are the last P - N occurrences of (u, w) when
// P > N (none when P <= N), and the list holds the survivors of all pairs in
// the order they were added.  Each iteration therefore records the three
// edges of each visible face in order (occurrence s = 3 * visible rank + edge)
// and counts each directed edge; an occurrence is kept when its edge has
// P > N and, if N > 0, its rank among the edge's occurrences is >= N.  The
// ranks come from per-edge counters advanced in order (rounds of the block in
// turn, waves of a round in turn, lanes of a wave by their position).  A
// block-wide ordered compaction of the kept occurrences gives the new faces
// in list order.
struct EpaBlockLds {
    SupV verts[EPA_MAXV];
    int cnt[EPA_KEYS];   // occurrences of each directed edge this iteration
    int cur[EPA_KEYS];   // occurrences of each edge ranked so far (edges seen both ways)
    int wsum[2 * BIG_NW];
    double wd[BIG_NW];
    int wi[BIG_NW];
};

// Exclusive block-wide ranks of two predicates in thread order, and totals.
__device__ __forceinline__ void block_rank2(bool p0, bool p1, int* wsum, int& r0, int& r1, int& t0, int& t1) {
    const uint64_t m0 = __ballot(p0), m1 = __ballot(p1);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { wsum[w] = (int)__popcll(m0); wsum[BIG_NW + w] = (int)__popcll(m1); }
    __syncthreads();
    int b0 = 0, b1 = 0;
    t0 = 0; t1 = 0;
    for (int k = 0; k < BIG_NW; ++k) {
        const int c0 = wsum[k], c1 = wsum[BIG_NW + k];
        if (k < w) { b0 += c0; b1 += c1; }
        t0 += c0; t1 += c1;
    }
    __syncthreads();
    const uint64_t below = lanemask_lt();
    r0 = b0 + (int)__popcll(m0 & below);
    r1 = b1 + (int)__popcll(m1 & below);
}

__device__ Contact epa_block(const Geom& A, const Geom& B, V3 shift, const SupV tet[4], EpaBlockLds& L, Face* faces,
                             uint16_t* occ, int maxf, bool& overflow) {
    constexpr int T = 64 * BIG_NW;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    SupV* verts = L.verts;
    if (tid == 0)
        for (int k = 0; k < 4; ++k) verts[k] = tet[k];
    __syncthreads();
    int nv = 4, nf = 4;
    if (tid < 4) {
        const int fi[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
        faces[tid] = make_face(verts, fi[tid][0], fi[tid][1], fi[tid][2]);
    }
    __syncthreads();
    Face best = faces[0];
    for (int iter = 0; iter < 64; ++iter) {
        // first face of minimal distance
        double bd = 0.0;
        int bi = -1;
#pragma unroll 4
        for (int i = tid; i < nf; i += T) {
            const double d = faces[i].dist;
            if (bi < 0 || d < bd) { bd = d; bi = i; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double od = __shfl_xor(bd, off);
            const int oi = __shfl_xor(bi, off);
            const bool take = oi >= 0 && (bi < 0 || od < bd || (od == bd && oi < bi));
            bd = take ? od : bd;
            bi = take ? oi : bi;
        }
        if ((tid & 63) == 0) { L.wd[w] = bd; L.wi[w] = bi; }
        __syncthreads();
        bd = L.wd[0];
        bi = L.wi[0];
        for (int k = 1; k < BIG_NW; ++k) {
            const double od = L.wd[k];
            const int oi = L.wi[k];
            const bool take = oi >= 0 && (bi < 0 || od < bd || (od == bd && oi < bi));
            bd = take ? od : bd;
            bi = take ? oi : bi;
        }
        best = faces[bi];
        const SupV p = cso_support(A, B, shift, best.n);
        const double d = dot(p.v, best.n);
        if (d - best.dist < 1e-7) break;
        const int pidx = nv;
        if (tid == 0) verts[nv] = p;
        ++nv;
        for (int k = tid; k < EPA_KEYS; k += T) { L.cnt[k] = 0; L.cur[k] = 0; }
        __syncthreads();
        // visibility rounds: compact the kept faces in order, record the
        // visible faces' edges in order.  A kept face moves to an index at or
        // below its own, never into the chunk loaded ahead.
        int nk = 0, nvis = 0;
        bool in_next = tid < nf;
        Face f_next{};
        if (in_next) f_next = faces[tid];
        for (int base = 0; base < nf; base += T) {
            const bool in = in_next;
            const Face f = f_next;
            in_next = base + T + tid < nf;
            if (in_next) f_next = faces[base + T + tid];
            const bool vis = in && dot(f.n, p.v - verts[f.a].v) > 1e-12;
            int rk, rv, tk, tv;
            block_rank2(in && !vis, vis, L.wsum, rk, rv, tk, tv);
            if (in && !vis) faces[nk + rk] = f;
            if (vis) {
                const int e0 = f.a * EPA_MAXV + f.b, e1 = f.b * EPA_MAXV + f.c, e2 = f.c * EPA_MAXV + f.a;
                uint16_t* o = occ + 3 * (size_t)(nvis + rv);
                o[0] = (uint16_t)e0; o[1] = (uint16_t)e1; o[2] = (uint16_t)e2;
                atomicAdd(&L.cnt[e0], 1);
                atomicAdd(&L.cnt[e1], 1);
                atomicAdd(&L.cnt[e2], 1);
            }
            nk += tk;
            nvis += tv;
        }
        __syncthreads();
        // the surviving occurrences, in order, become the new faces
        const int nocc = 3 * nvis;
        int wpos = nk;
        for (int base = 0; base < nocc; base += T) {
            const int q = base + tid;
            int e = -1, u = 0, v = 0, np = 0, nn = 0;
            if (q < nocc) {
                e = occ[q];
                u = e / EPA_MAXV;
                v = e - u * EPA_MAXV;
                np = L.cnt[e];
                nn = L.cnt[v * EPA_MAXV + u];
            }
            bool keep = np > nn;
            const bool mixed = keep && nn > 0;
            if (__syncthreads_or(mixed)) {
                // rank among the edge's occurrences: earlier rounds (the
                // edge's counter), earlier waves of this round (the waves
                // take turns), earlier lanes of this wave
                const int key = mixed ? e : -1;
                int before = 0;
                bool last = true;
                for (int l = 0; l < 64; ++l) {
                    const int o = __shfl(key, l);
                    if (o == key) {
                        before += l < lane;
                        last = last && l <= lane;
                    }
                }
                for (int ph = 0; ph < BIG_NW; ++ph) {
                    if (w == ph && mixed) {
                        const int rank = L.cur[e] + before;
                        keep = rank >= nn;
                        if (last) L.cur[e] = rank + 1;
                    }
                    __syncthreads();
                }
            }
            int rs, r1, ts, t1;
            block_rank2(keep, false, L.wsum, rs, r1, ts, t1);
            if (keep && wpos + rs < maxf) faces[wpos + rs] = make_face(verts, u, v, pidx);
            wpos += ts;
        }
        if (wpos > maxf) {
            overflow = true;
            break;
        }
        nf = wpos;
        __syncthreads();
        if (nf == 0) break;
    }
    return epa_result(verts, best);
}

// Large-workspace EPA reruns of the queued pairs, one block per pair.
__global__ void __launch_bounds__(64 * BIG_NW) k_narrow_big(BodiesDev B, SimParams P, uint32_t* __restrict__ counters,
                                                            const uint32_t* __restrict__ cand_i,
                                                            const uint32_t* __restrict__ cand_j,
                                                            uint32_t* __restrict__ ncont, double* __restrict__ ct,
                                                            EpaBig big) {
    __shared__ EpaBlockLds L;
    const uint32_t nbig = min(counters[CNT_EPABIG_PAIRS], big.cap);
    uint32_t* fault = counters + CNT_FAULT;
    if (blockIdx.x == 0 && threadIdx.x == 0 && nbig) atomicAdd(&counters[CNT_EPARERUN], nbig);
    for (uint32_t t = blockIdx.x; t < nbig; t += gridDim.x) {
        const uint32_t p = big.list[t];
        if (p & MESH_ENTRY) continue;   // a mesh triangle: k_mesh_big
        const size_t fo = (size_t)t * (size_t)big.faces_per;
        const bool truncate = big.faces_per >= big.faces_max;
        const uint32_t i = cand_i[p], j = cand_j[p];
        const Box3& box = box_of(B, P, i);
        const Geom a = load_geom(B, i), b = load_geom(B, j);
        // only pairs that reached EPA are queued: convexContact (src/epa.hpp:153-158)
        const V3 shift = min_image_shift(box, a.x, b.x);
        SupV tet[4];
        bool overflow = false;
        Contact c = no_contact();
        if (gjk_tetra(a, b, shift, tet))
            c = epa_block(a, b, shift, tet, L, big.faces + fo, big.occ + 3 * fo, big.faces_per,
                          overflow);
        uint32_t nc = 0;
        double* out = ct + (size_t)p * 28;
        if (overflow && !truncate) {
            if (threadIdx.x == 0) atomicOr(fault, NE_FAULT_EPA_BIG);
        } else if (c.hit) {
            if (overflow && threadIdx.x == 0) atomicAdd(&counters[CNT_EPATRUNC], 1u);
            if (a.shape == SH_BOX && b.shape == SH_BOX && !P.no_manifold) {
                V3 pts[4];
                double dep[4];
                const int m = box_box_manifold(box, a, b, c.normal, pts, dep, fault);
                if (threadIdx.x == 0)
                    for (int k = 0; k < m; ++k) {
                        double* o = out + 7 * k;
                        o[0] = c.normal.x; o[1] = c.normal.y; o[2] = c.normal.z;
                        o[3] = pts[k].x; o[4] = pts[k].y; o[5] = pts[k].z; o[6] = dep[k];
                    }
                nc = (uint32_t)m;
            }
            if (nc == 0) {
                if (threadIdx.x == 0) {
                    out[0] = c.normal.x; out[1] = c.normal.y; out[2] = c.normal.z;
                    out[3] = c.point.x; out[4] = c.point.y; out[5] = c.point.z; out[6] = c.overlap;
                }
                nc = 1;
            }
        }
        if (threadIdx.x == 0) ncont[p] = nc;
        __syncthreads();
    }
}

// The EPA runs that went past the second per-thread budget, one wave per
// pair with the whole polytope (vertices, faces, horizon list) in LDS: a long
// run no longer holds 63 other lanes, and its faces stay out of scratch.
// A run that outgrows the LDS workspace goes to the global-memory queue.
__global__ void __launch_bounds__(64) k_epa_wave(BodiesDev B, SimParams P, uint32_t* __restrict__ counters,
                                                 const uint32_t* __restrict__ cand_i, const uint32_t* __restrict__ cand_j,
                                                 uint32_t* __restrict__ ncont, double* __restrict__ ct,
                                                 const uint32_t* __restrict__ queue, int count_slot, uint32_t cap,
                                                 EpaBig big) {
    __shared__ SupV verts[EPA_MAXV];
    __shared__ Face faces[EPA_MAXF];
    __shared__ int hor[EPA_MAXH][2];
    __shared__ int stage[3 * 64];
    __shared__ int shared[4];
    const uint32_t nq = min(counters[count_slot], cap);
    uint32_t* fault = counters + CNT_FAULT;
    for (uint32_t t = blockIdx.x; t < nq; t += gridDim.x) {
        const uint32_t p = queue[t];
        HorizonList hz;
        hz.hor = hor; hz.cap = EPA_MAXH; hz.n = 0;
        const uint32_t i = cand_i[p], j = cand_j[p];
        const Box3& box = box_of(B, P, i);
        const Geom a = load_geom(B, i), b = load_geom(B, j);
        const V3 shift = min_image_shift(box, a.x, b.x);   // convexContact (src/epa.hpp:153-158)
        SupV tet[4];
        bool overflow = false;
        Contact c = no_contact();
        if (gjk_tetra(a, b, shift, tet))
            c = epa_wave(a, b, shift, tet, verts, faces, EPA_MAXF, hz, stage, shared, false, overflow);
        uint32_t nc = 0;
        double* out = ct + (size_t)p * 28;
        if (overflow) {   // rerun from the start with the large workspace
            if (threadIdx.x == 0) {
                const uint32_t slot = atomicAdd(&counters[CNT_EPABIG], 1u);
                if (slot < big.cap) big.list[slot] = p;
                else atomicOr(fault, NE_FAULT_EPA_QUEUE);
            }
        } else if (c.hit) {
            if (a.shape == SH_BOX && b.shape == SH_BOX && !P.no_manifold) {
                V3 pts[4];
                double dep[4];
                const int m = box_box_manifold(box, a, b, c.normal, pts, dep, fault);
                if (threadIdx.x == 0)
                    for (int k = 0; k < m; ++k) {
                        double* o = out + 7 * k;
                        o[0] = c.normal.x; o[1] = c.normal.y; o[2] = c.normal.z;
                        o[3] = pts[k].x; o[4] = pts[k].y; o[5] = pts[k].z; o[6] = dep[k];
                    }
                nc = (uint32_t)m;
            }
            if (nc == 0) {
                if (threadIdx.x == 0) {
                    out[0] = c.normal.x; out[1] = c.normal.y; out[2] = c.normal.z;
                    out[3] = c.point.x; out[4] = c.point.y; out[5] = c.point.z; out[6] = c.overlap;
                }
                nc = 1;
            }
        }
        if (threadIdx.x == 0) ncont[p] = nc;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Static triangle meshes: src/mesh.hpp:14-93 (tree, query) and
// meshContacts, src/world.cpp:56-83.
//
// All meshes share one set of arrays; indices inside are absolute.  Per step
// every candidate pair with a mesh body queries the tree twice (count, then
// write the triangles in the query's order), each reported triangle becomes
// a candidate of its own for the narrow phase, and a pass per pair keeps the
// hits in that order: the reference's per-pair contact vector.
// ---------------------------------------------------------------------------

struct MeshNode {
    double lo[3], hi[3];
    int32_t left, right, start, count;   // children (absolute) or a leaf's [start, start+count) into order
};

struct MeshDev {
    const double* verts;      // 3 per vertex, body frame
    const int32_t* tris;      // 3 per triangle, absolute vertex indices
    const int32_t* order;     // absolute triangle indices in tree order
    const MeshNode* nodes;
    const int32_t* root;      // per mesh: root node, or -1 for an empty mesh
    const int32_t* of_body;   // per body: mesh index, or -1
};

// The pair's mesh body and the other body, as gatherContacts picks them
// (world.cpp:89-94): b's mesh when b is a mesh, else a's with normals flipped.
__device__ __forceinline__ bool mesh_pair(const BodiesDev& B, uint32_t i, uint32_t j, uint32_t& o, uint32_t& m,
                                          bool& flip) {
    if (B.shape[j] == SH_MESH) { m = j; o = i; flip = false; return true; }
    if (B.shape[i] == SH_MESH) { m = i; o = j; flip = true; return true; }
    return false;
}

// MeshData::query (mesh.hpp:46-59) over the cube of half-side
// boundingRadius + 0.05 about the other body's centre in the mesh frame:
// depth first, right child popped first, children dropped once 62 entries
// are held.
template <class Emit>
__device__ __forceinline__ uint32_t mesh_query(const BodiesDev& B, const SimParams& P, const MeshDev& M,
                                               uint32_t o, uint32_t m, Emit emit) {
    const int mi = M.of_body[m];
    if (mi < 0) return 0;
    const int root = M.root[mi];
    if (root < 0) return 0;
    const Box3& box = box_of(B, P, o);
    const V3 rel = min_image(box, ld3(B.x0, B.x1, B.x2, o) - ld3(B.x0, B.x1, B.x2, m));
    const V3 cl = qinvrot(ldq(B, m), rel);
    const double rr = B.bR[o] + 0.05;
    const V3 lo = cl - V3{rr, rr, rr}, hi = cl + V3{rr, rr, rr};
    int stack[64];
    int sp = 0;
    stack[sp++] = root;
    uint32_t n = 0;
    while (sp > 0) {
        const MeshNode& nd = M.nodes[stack[--sp]];
        if (lo.x > nd.hi[0] || hi.x < nd.lo[0] || lo.y > nd.hi[1] || hi.y < nd.lo[1] || lo.z > nd.hi[2] ||
            hi.z < nd.lo[2])
            continue;
        if (nd.left < 0) {
            for (int k = 0; k < nd.count; ++k) emit(n++, M.order[nd.start + k]);
        } else if (sp < 62) {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    return n;
}

__global__ void __launch_bounds__(TPB) k_mesh_count(BodiesDev B, SimParams P, MeshDev M,
                                                    const uint32_t* __restrict__ counters,
                                                    const uint32_t* __restrict__ cand_i,
                                                    const uint32_t* __restrict__ cand_j, uint32_t* __restrict__ mq_cnt) {
    const uint32_t np = counters[CNT_P];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        uint32_t o, m;
        bool flip;
        uint32_t n = 0;
        if (mesh_pair(B, cand_i[p], cand_j[p], o, m, flip)) n = mesh_query(B, P, M, o, m, [](uint32_t, int32_t) {});
        mq_cnt[p] = n;
    }
}

__global__ void __launch_bounds__(TPB) k_mesh_emit(BodiesDev B, SimParams P, MeshDev M,
                                                   const uint32_t* __restrict__ counters,
                                                   const uint32_t* __restrict__ cand_i, const uint32_t* __restrict__ cand_j,
                                                   const uint32_t* __restrict__ mq_off, uint32_t mqcap,
                                                   uint32_t* __restrict__ mq_pair, int32_t* __restrict__ mq_tri) {
    const uint32_t np = counters[CNT_P];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        uint32_t o, m;
        bool flip;
        const uint32_t base = mq_off[p];
        if (base == mq_off[p + 1] || !mesh_pair(B, cand_i[p], cand_j[p], o, m, flip)) continue;
        mesh_query(B, P, M, o, m, [&](uint32_t k, int32_t t) {
            if (base + k < mqcap) { mq_pair[base + k] = p; mq_tri[base + k] = t; }
        });
    }
}

// The triangle as a thin convex prism (world.cpp:66-79): world vertices,
// unit normal, the vertices pushed 1e-3 along and against it; a temporary
// body at the origin with identity orientation.  False for a degenerate
// triangle (|n| < 1e-12), which the reference skips.
__device__ __forceinline__ bool mesh_prism(const BodiesDev& B, const MeshDev& M, uint32_t m, int32_t t, double* pv,
                                           Geom& g) {
    const V3 xm = ld3(B.x0, B.x1, B.x2, m);
    const Qt qm = ldq(B, m);
    const int32_t* tv = M.tris + 3 * (size_t)t;
    const double* a = M.verts + 3 * (size_t)tv[0];
    const double* b = M.verts + 3 * (size_t)tv[1];
    const double* c = M.verts + 3 * (size_t)tv[2];
    const V3 w0 = xm + qrot(qm, V3{a[0], a[1], a[2]});
    const V3 w1 = xm + qrot(qm, V3{b[0], b[1], b[2]});
    const V3 w2 = xm + qrot(qm, V3{c[0], c[1], c[2]});
    V3 n = cross(w1 - w0, w2 - w0);
    const double nl = norm(n);
    if (nl < 1e-12) return false;
    n = n / nl;
    const double thick = 1e-3;
    const V3 up = n * thick;
    const V3 v[6] = {w0 + up, w1 + up, w2 + up, w0 - up, w1 - up, w2 - up};
    for (int k = 0; k < 6; ++k) { pv[3 * k] = v[k].x; pv[3 * k + 1] = v[k].y; pv[3 * k + 2] = v[k].z; }
    g.x = V3{0.0, 0.0, 0.0};
    g.q = Qt{1.0, 0.0, 0.0, 0.0};
    g.he = V3{0.0, 0.0, 0.0};
    g.radius = 1.0;
    g.hh = 0.0;
    g.hull = pv;
    g.nhull = 6;
    g.shape = SH_CONVEX;
    return true;
}

__device__ __forceinline__ void store_contact(double* out, const Contact& c, bool flip) {
    out[0] = flip ? -c.normal.x : c.normal.x;
    out[1] = flip ? -c.normal.y : c.normal.y;
    out[2] = flip ? -c.normal.z : c.normal.z;
    out[3] = c.point.x; out[4] = c.point.y; out[5] = c.point.z; out[6] = c.overlap;
}

constexpr int PV_STRIDE = 19;   // doubles per thread for a prism's vertices (18, padded against bank conflicts)

#include "ne_epa_lds.hpp"

// detectContact(other, prism) per triangle candidate up to EPA: the
// plane's analytic path, else GJK; candidates GJK finds overlapping are
// queued for the EPA tiers (`epa_q`, counted in CNT_MQEPA).  Without the EPA
// workspace this kernel runs several waves a SIMD over the many candidates
// that do not touch (the prism's six vertices in LDS, 9.5 KiB a block).
__global__ void __launch_bounds__(64) k_mesh_gjk(BodiesDev B, SimParams P, MeshDev M, uint32_t* __restrict__ counters,
                                                 const uint32_t* __restrict__ cand_i, const uint32_t* __restrict__ cand_j,
                                                 const uint32_t* __restrict__ mq_pair, const int32_t* __restrict__ mq_tri,
                                                 uint8_t* __restrict__ mhit, double* __restrict__ mct,
                                                 uint32_t* __restrict__ epa_q) {
    __shared__ double lds_pv[64 * PV_STRIDE];
    double* pv = lds_pv + threadIdx.x * PV_STRIDE;
    const uint32_t nq = counters[CNT_MQ];
    for (uint32_t u0 = blockIdx.x * 64u; u0 < nq; u0 += gridDim.x * 64u) {
        const uint32_t q = u0 + threadIdx.x;
        bool overlap = false;
        if (q < nq) {
            const uint32_t p = mq_pair[q];
            uint32_t o, m;
            bool flip;
            mesh_pair(B, cand_i[p], cand_j[p], o, m, flip);
            Geom tri;
            uint8_t hit = 0;
            if (mesh_prism(B, M, m, mq_tri[q], pv, tri)) {
                const Geom other = load_geom(B, o);
                EpaWS ws{};   // GJK only
                const Contact c = detect_contact(box_of(B, P, o), other, tri, ws);
                if (c.hit == 2) {
                    overlap = true;
                } else if (c.hit) {   // the plane's analytic contact
                    store_contact(mct + 7 * (size_t)q, c, flip);
                    hit = 1;
                }
            }
            mhit[q] = hit;   // an overlapping candidate's EPA tier writes it
        }
        const uint32_t slot = wave_add1(counters, CNT_MQEPA, overlap);
        if (overlap) epa_q[slot] = q;   // at most one entry per candidate: fits the candidate capacity
    }
}

// EPA tier A for the queued mesh candidates (k_epa_lds's workspace beside
// the prism's six vertices: three waves a CU), from GJK on the same inputs
// again.  Runs that outgrow the workspace go to `long_q` for the group tier,
// which reruns them from the start.
__global__ void __launch_bounds__(64) k_mesh_lds(BodiesDev B, SimParams P, MeshDev M, uint32_t* __restrict__ counters,
                                                 const uint32_t* __restrict__ cand_i, const uint32_t* __restrict__ cand_j,
                                                 const uint32_t* __restrict__ mq_pair, const int32_t* __restrict__ mq_tri,
                                                 uint8_t* __restrict__ mhit, double* __restrict__ mct,
                                                 const uint32_t* __restrict__ epa_q, uint32_t* __restrict__ long_q) {
    __shared__ EpaLaneLds<8, 12, 8> lds;
    __shared__ double lds_pv[64 * PV_STRIDE];
    EpaLane<8, 12, 8> E{lds, (int)threadIdx.x};
    double* pv = lds_pv + threadIdx.x * PV_STRIDE;
    const uint32_t nq = counters[CNT_MQEPA];
    V3 dirs[8];
    for (uint32_t u0 = blockIdx.x * 64u; u0 < nq; u0 += gridDim.x * 64u) {
        const uint32_t u = u0 + threadIdx.x;
        bool defer = false;
        const uint32_t q = u < nq ? epa_q[u] : 0u;
        if (u < nq) {
            const uint32_t p = mq_pair[q];
            uint32_t o, m;
            bool flip;
            mesh_pair(B, cand_i[p], cand_j[p], o, m, flip);
            Geom tri;
            mesh_prism(B, M, m, mq_tri[q], pv, tri);   // non-degenerate: it was queued
            const Geom other = load_geom(B, o);
            const V3 shift = min_image_shift(box_of(B, P, o), other.x, tri.x);   // convex_contact (epa.hpp:153-158)
            SupD tet[4];
            Contact c = no_contact();
            if (gjk_tetra(other, tri, shift, tet) && !epa_lane(E, other, tri, shift, tet, dirs, c)) defer = true;
            if (!defer && c.hit) store_contact(mct + 7 * (size_t)q, c, flip);
            if (!defer) mhit[q] = c.hit ? 1 : 0;
        }
        const uint32_t slot = wave_add1(counters, CNT_MQLONG, defer);
        if (defer) long_q[slot] = q;   // at most one entry per candidate: fits the candidate capacity
    }
}

// The queued triangle candidates, one block per candidate (as k_narrow_big
// does for pairs).
__global__ void __launch_bounds__(64 * BIG_NW) k_mesh_big(BodiesDev B, SimParams P, MeshDev M,
                                                          uint32_t* __restrict__ counters,
                                                          const uint32_t* __restrict__ cand_i,
                                                          const uint32_t* __restrict__ cand_j,
                                                          const uint32_t* __restrict__ mq_pair,
                                                          const int32_t* __restrict__ mq_tri, uint8_t* __restrict__ mhit,
                                                          double* __restrict__ mct, EpaBig big) {
    __shared__ EpaBlockLds L;
    const uint32_t nbig = min(counters[CNT_EPABIG], big.cap);
    uint32_t* fault = counters + CNT_FAULT;
    double pv[18];
    for (uint32_t t = blockIdx.x; t < nbig; t += gridDim.x) {
        const uint32_t e = big.list[t];
        if (!(e & MESH_ENTRY)) continue;
        const uint32_t q = e & ~MESH_ENTRY;
        if (threadIdx.x == 0) atomicAdd(&counters[CNT_EPARERUN], 1u);
        const size_t fo = (size_t)t * (size_t)big.faces_per;
        const bool truncate = big.faces_per >= big.faces_max;
        const uint32_t p = mq_pair[q];
        uint32_t o, m;
        bool flip;
        mesh_pair(B, cand_i[p], cand_j[p], o, m, flip);
        Geom tri;
        mesh_prism(B, M, m, mq_tri[q], pv, tri);   // non-degenerate: it was queued
        const Geom a = load_geom(B, o);
        const Box3& box = box_of(B, P, o);
        const V3 shift = min_image_shift(box, a.x, tri.x);
        SupV tet[4];
        bool overflow = false;
        Contact c = no_contact();
        if (gjk_tetra(a, tri, shift, tet))
            c = epa_block(a, tri, shift, tet, L, big.faces + fo, big.occ + 3 * fo, big.faces_per,
                          overflow);
        uint8_t hit = 0;
        if (overflow && !truncate) {
            if (threadIdx.x == 0) atomicOr(fault, NE_FAULT_EPA_BIG);
        } else if (c.hit) {
            if (overflow && threadIdx.x == 0) atomicAdd(&counters[CNT_EPATRUNC], 1u);
            if (threadIdx.x == 0) store_contact(mct + 7 * (size_t)q, c, flip);
            hit = 1;
        }
        if (threadIdx.x == 0) mhit[q] = hit;
        __syncthreads();
    }
}

// The triangle candidates whose EPA went past the first launch's budget:
// one wave per candidate, polytope and horizon list in LDS (as k_epa_wave).
__global__ void __launch_bounds__(64) k_mesh_wave(BodiesDev B, SimParams P, MeshDev M, uint32_t* __restrict__ counters,
                                                  const uint32_t* __restrict__ cand_i, const uint32_t* __restrict__ cand_j,
                                                  const uint32_t* __restrict__ mq_pair, const int32_t* __restrict__ mq_tri,
                                                  uint8_t* __restrict__ mhit, double* __restrict__ mct,
                                                  const uint32_t* __restrict__ list, int count_slot, EpaBig big) {
    __shared__ SupV verts[EPA_MAXV];
    __shared__ Face faces[EPA_MAXF];
    __shared__ int hor[EPA_MAXH][2];
    __shared__ int stage[3 * 64];
    __shared__ int shared[4];
    const uint32_t nq = counters[count_slot];
    uint32_t* fault = counters + CNT_FAULT;
    double pv[18];
    for (uint32_t t = blockIdx.x; t < nq; t += gridDim.x) {
        const uint32_t q = list[t];
        const uint32_t p = mq_pair[q];
        uint32_t o, m;
        bool flip;
        mesh_pair(B, cand_i[p], cand_j[p], o, m, flip);
        Geom tri;
        mesh_prism(B, M, m, mq_tri[q], pv, tri);   // non-degenerate: its EPA ran
        const Geom a = load_geom(B, o);
        const Box3& box = box_of(B, P, o);
        const V3 shift = min_image_shift(box, a.x, tri.x);
        HorizonList hz;
        hz.hor = hor; hz.cap = EPA_MAXH; hz.n = 0;
        SupV tet[4];
        bool overflow = false;
        Contact c = no_contact();
        if (gjk_tetra(a, tri, shift, tet))
            c = epa_wave(a, tri, shift, tet, verts, faces, EPA_MAXF, hz, stage, shared, false, overflow);
        if (threadIdx.x == 0) {
            uint8_t hit = 0;
            if (overflow) {   // rerun from the start with the large workspace
                const uint32_t slot = atomicAdd(&counters[CNT_EPABIG], 1u);
                if (slot < big.cap) big.list[slot] = MESH_ENTRY | q;
                else atomicOr(fault, NE_FAULT_EPA_QUEUE);
            } else if (c.hit) {
                store_contact(mct + 7 * (size_t)q, c, flip);
                hit = 1;
            }
            mhit[q] = hit;
        }
        __syncthreads();
    }
}

// Per mesh pair: its hits, in query order, become the pair's contacts.
__global__ void __launch_bounds__(TPB) k_mesh_gather(BodiesDev B, const uint32_t* __restrict__ counters,
                                                     const uint32_t* __restrict__ cand_i,
                                                     const uint32_t* __restrict__ cand_j,
                                                     const uint32_t* __restrict__ mq_off, uint32_t mqcap,
                                                     const uint8_t* __restrict__ mhit, double* __restrict__ mct,
                                                     uint32_t* __restrict__ ncont) {
    const uint32_t np = counters[CNT_P];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        const uint32_t b0 = mq_off[p], b1 = min(mq_off[p + 1], mqcap);
        if (b0 >= b1) continue;
        uint32_t h = 0;
        for (uint32_t q = b0; q < b1; ++q) {
            if (!mhit[q]) continue;
            const uint32_t d = b0 + h;
            if (d != q)
                for (int k = 0; k < 7; ++k) mct[7 * (size_t)d + k] = mct[7 * (size_t)q + k];
            ++h;
        }
        ncont[p] = h;
    }
}

// ---------------------------------------------------------------------------
// Constraint build: src/world.cpp:522-553 (+ effMass once, contact.hpp:28-33)
// ---------------------------------------------------------------------------

struct ConsDev {
    uint32_t *ci, *cj, *slot;
    double *n0, *n1, *n2, *p0, *p1, *p2, *ra0, *ra1, *ra2, *rb0, *rb1, *rb2;
    double *ov, *rbias, *mu, *nimp, *timp, *kn;
    uint64_t* key;
    uint32_t* src;   // index of the previous step's constraint that warm-started this one, or ~0u
};


__device__ __forceinline__ void uf_union(const BodiesDev& B, uint32_t* parent, uint32_t i, uint32_t j);


__global__ void __launch_bounds__(TPB) k_cons_build(BodiesDev B, SimParams P, const uint32_t* __restrict__ counters_in,
                                                    const uint32_t* __restrict__ cand_i, const uint32_t* __restrict__ cand_j,
                                                    const uint32_t* __restrict__ ncont, const uint32_t* __restrict__ cons_off,
                                                    const double* __restrict__ ct, ConsDev C, uint32_t ccap,
                                                    const uint64_t* __restrict__ prev_key, const double* __restrict__ prev_imp,
                                                    const uint32_t* __restrict__ prev_n, uint32_t* __restrict__ parent,
                                                    uint32_t* __restrict__ claimed, const uint32_t* __restrict__ mq_off,
                                                    const double* __restrict__ mct, bool link,
                                                    double* __restrict__ rec) {
    const uint32_t np = counters_in[CNT_P];
    const uint32_t pn = *prev_n;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        const uint32_t nc = ncont[p];
        if (!nc) continue;
        const uint32_t i = cand_i[p], j = cand_j[p];
        const double ima = B.im[i], imb = B.im[j];
        const WorldDev* wd = B.W ? &B.W[B.wid[i]] : nullptr;
        const Box3& box = wd ? wd->box : P.box;
        const double wrest = wd ? wd->restitution : P.restitution;
        const double wfric = wd ? wd->friction : P.friction;
        const double ra = B.rest[i] < 0 ? wrest : B.rest[i];
        const double rb = B.rest[j] < 0 ? wrest : B.rest[j];
        const double fa = B.fric[i] < 0 ? wfric : B.fric[i];
        const double fb = B.fric[j] < 0 ? wfric : B.fric[j];
        const double e = combine_mat(P.rcomb, ra, rb);
        const double mu = combine_mat(P.fcomb, fa, fb);
        const V3 xa = ld3(B.x0, B.x1, B.x2, i), xb = ld3(B.x0, B.x1, B.x2, j);
        const V3 va = ld3(B.v0, B.v1, B.v2, i), vb = ld3(B.v0, B.v1, B.v2, j);
        const V3 wa = ld3(B.w0, B.w1, B.w2, i), wb = ld3(B.w0, B.w1, B.w2, j);
        const Qt qa = ldq(B, i), qb = ldq(B, j);
        const V3 iia = ld3(B.ii0, B.ii1, B.ii2, i), iib = ld3(B.ii0, B.ii1, B.ii2, j);
        if (link) uf_union(B, parent, i, j);   // islands: join the two bodies when both can move
        const uint64_t key = ((uint64_t)i << 32) | (uint64_t)j;   // world.cpp:34-36
        // warm cache: previous step's constraints are sorted by (key, slot)
        uint32_t lo = 0;
        if (P.warm) {
            uint32_t hi = pn;
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (prev_key[mid] < key) lo = mid + 1; else hi = mid;
            }
        }
        const uint32_t k0 = cons_off[p];
        const bool meshp = mct && (B.shape[i] == SH_MESH || B.shape[j] == SH_MESH);
        for (uint32_t s = 0; s < nc; ++s) {
            const uint32_t k = k0 + s;
            if (k >= ccap) break;
            const double* c = meshp ? mct + 7 * ((size_t)mq_off[p] + s) : ct + (size_t)p * 28 + 7 * s;
            const V3 n{c[0], c[1], c[2]}, pt{c[3], c[4], c[5]};
            const V3 rA = min_image(box, pt - xa);
            const V3 rB = min_image(box, pt - xb);
            const V3 vc = (va + cross(wa, rA)) - (vb + cross(wb, rB));
            const double vn = dot(vc, n);
            double nimp = 0.0;
            if (P.warm && lo + s < pn && prev_key[lo + s] == key) nimp = prev_imp[lo + s];
            C.ci[k] = i; C.cj[k] = j; C.slot[k] = s;
            C.n0[k] = n.x; C.n1[k] = n.y; C.n2[k] = n.z;
            C.p0[k] = pt.x; C.p1[k] = pt.y; C.p2[k] = pt.z;
            C.ov[k] = c[6];
            C.nimp[k] = nimp;
            C.timp[k] = 0.0;
            const double rbias = (vn < -0.5) ? -e * vn : 0.0;                // world.cpp:40, 544
            const double kn = eff_mass(ima, qa, iia, imb, qb, iib, rA, rB, n);
            if (rec) {   // only the dataflow solve reads these: straight into its record
                double2* r = reinterpret_cast<double2*>(rec + (size_t)k * WF_REC);
                r[0] = double2{n.x, n.y};
                r[1] = double2{n.z, rA.x};
                r[2] = double2{rA.y, rA.z};
                r[3] = double2{rB.x, rB.y};
                r[4] = double2{rB.z, rbias};
                r[5] = double2{kn, mu};
                r[6] = double2{nimp, 0.0};
                r[7] = double2{c[6], 0.0};
            } else {
                C.ra0[k] = rA.x; C.ra1[k] = rA.y; C.ra2[k] = rA.z;
                C.rb0[k] = rB.x; C.rb1[k] = rB.y; C.rb2[k] = rB.z;
                C.rbias[k] = rbias;
                C.mu[k] = mu;
                C.kn[k] = kn;
            }
            C.key[k] = key;
            C.src[k] = (P.warm && lo + s < pn && prev_key[lo + s] == key) ? lo + s : 0xFFFFFFFFu;
        }
    }
}

// ---------------------------------------------------------------------------
// Islands: union-find over bodies whose velocity can change.  Static and
// kinematic bodies (invMass == 0 and zero inverse inertia) only receive +-0
// updates in the reference, so they never join islands and are never written.
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool mut_v(const BodiesDev& B, uint32_t b) { return B.im[b] != 0.0; }
__device__ __forceinline__ bool mut_w(const BodiesDev& B, uint32_t b) {
    return B.ii0[b] != 0.0 || B.ii1[b] != 0.0 || B.ii2[b] != 0.0;
}
__device__ __forceinline__ bool mutable_body(const BodiesDev& B, uint32_t b) { return mut_v(B, b) || mut_w(B, b); }

__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
    uint32_t p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
    while (p != x) {
        uint32_t gp = __atomic_load_n(&parent[p], __ATOMIC_RELAXED);
        if (gp != p) __atomic_store_n(&parent[x], gp, __ATOMIC_RELAXED);   // path halving (benign)
        x = p;
        p = gp;
    }
    return x;
}

// Lock-free union of the sets of i and j (both must be able to move).
__device__ __forceinline__ void uf_union(const BodiesDev& B, uint32_t* parent, uint32_t i, uint32_t j) {
    if (!mutable_body(B, i) || !mutable_body(B, j)) return;
    uint32_t a = i, b = j;
    for (int guard = 0; guard < 1 << 20; ++guard) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) break;
        if (a < b) { uint32_t t = a; a = b; b = t; }   // hang the larger root under the smaller
        if (atomicCAS(&parent[a], a, b) == a) break;
    }
}

// Island ids: the first constraint to reach a root claims it and takes the
// next id (ids are a scheduling detail; the results do not depend on them).
__global__ void __launch_bounds__(TPB) k_isl_claim(BodiesDev B, uint32_t* counters, const ConsDev C,
                                                   uint32_t* __restrict__ parent, uint32_t* __restrict__ cons_root,
                                                   uint32_t* __restrict__ claimed, uint32_t* __restrict__ root_id,
                                                   uint32_t* __restrict__ isl_count, uint32_t* __restrict__ isl_fill) {
    const uint32_t nc = counters[CNT_C];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
        const uint32_t i = C.ci[k];
        const uint32_t body = mutable_body(B, i) ? i : C.cj[k];
        const uint32_t r = uf_find(parent, body);
        cons_root[k] = r;
        if (__atomic_load_n(&claimed[r], __ATOMIC_RELAXED) != 0u) continue;   // most lanes of a large island
        if (atomicExch(&claimed[r], 1u) == 0u) {
            const uint32_t id = atomicAdd(&counters[CNT_ISL], 1u);
            root_id[r] = id;
            isl_count[id] = 0u;
            isl_fill[id] = 0u;
        }
    }
}

__global__ void __launch_bounds__(TPB) k_isl_count(const uint32_t* __restrict__ counters, const uint32_t* __restrict__ cons_root,
                                                   const uint32_t* __restrict__ root_id, uint32_t* __restrict__ cons_isl,
                                                   uint32_t* __restrict__ isl_count) {
    const uint32_t nc = counters[CNT_C];
    for (uint32_t k0 = blockIdx.x * blockDim.x; k0 < nc; k0 += gridDim.x * blockDim.x) {
        const uint32_t k = k0 + threadIdx.x;
        const bool on = k < nc;
        const uint32_t id = on ? root_id[cons_root[k]] : 0u;
        if (on) cons_isl[k] = id;
        wave_add1(isl_count, id, on);
    }
}

// Unordered fill of each island's segment; the solvers sort their segment.
__global__ void __launch_bounds__(TPB) k_isl_fill(const uint32_t* __restrict__ counters, const uint32_t* __restrict__ cons_isl,
                                                  const uint32_t* __restrict__ isl_start, uint32_t* __restrict__ isl_fill,
                                                  uint32_t* __restrict__ isl_list) {
    const uint32_t nc = counters[CNT_C];
    for (uint32_t k0 = blockIdx.x * blockDim.x; k0 < nc; k0 += gridDim.x * blockDim.x) {
        const uint32_t k = k0 + threadIdx.x;
        const bool on = k < nc;
        const uint32_t id = on ? cons_isl[k] : 0u;
        const uint32_t pos = wave_add1(isl_fill, id, on);
        if (on) isl_list[isl_start[id] + pos] = k;
    }
}

// ---------------------------------------------------------------------------
// Solve: each island sweeps its constraints in global (i, j, slot) order:
// warm start (world.cpp:559-568), I iterations of normal + Coulomb friction
// (world.cpp:572-612), positional correction (615-616, contact.hpp:78-85).
// ---------------------------------------------------------------------------


__device__ __forceinline__ BodyView load_view(const BodiesDev& B, uint32_t b) {
    BodyView o;
    o.v = ld3(B.v0, B.v1, B.v2, b);
    o.w = ld3(B.w0, B.w1, B.w2, b);
    o.q = ldq(B, b);
    o.ii = ld3(B.ii0, B.ii1, B.ii2, b);
    o.im = B.im[b];
    o.mv = o.im != 0.0;
    o.mw = o.ii.x != 0.0 || o.ii.y != 0.0 || o.ii.z != 0.0;
    return o;
}

__device__ __forceinline__ void store_view(BodiesDev& B, uint32_t b, const BodyView& o) {
    if (o.mv) st3(B.v0, B.v1, B.v2, b, o.v);
    if (o.mw) st3(B.w0, B.w1, B.w2, b, o.w);
}



__device__ __forceinline__ ConsGeom load_geom(const ConsDev& C, uint32_t k) {
    return ConsGeom{V3{C.n0[k], C.n1[k], C.n2[k]}, V3{C.ra0[k], C.ra1[k], C.ra2[k]}, V3{C.rb0[k], C.rb1[k], C.rb2[k]}};
}



// x_a += n (beta overlap / (m_a^-1 + m_b^-1)) m_a^-1, b opposite (contact.hpp:78-85)
__device__ __forceinline__ void correct(BodiesDev& B, const ConsDev& C, uint32_t k, double beta, V3& xa, V3& xb,
                                        double ima, double imb) {
    const double invSum = ima + imb;
    const double ov = C.ov[k];
    if (invSum <= 0.0 || ov <= 0.0) return;
    const V3 n{C.n0[k], C.n1[k], C.n2[k]};
    const V3 corr = n * (beta * ov / invSum);
    if (ima != 0.0) xa = xa + corr * ima;
    if (imb != 0.0) xb = xb - corr * imb;
}

__device__ __forceinline__ void sort_segment(uint32_t* a, uint32_t m) {
    for (uint32_t t = 1; t < m; ++t) {
        uint32_t key = a[t];
        uint32_t u = t;
        while (u > 0 && a[u - 1] > key) { a[u] = a[u - 1]; --u; }
        a[u] = key;
    }
}

constexpr uint32_t NO_BODY = 0xFFFFFFFFu;
constexpr uint32_t WAVE_SORT = 4096;   // wave_sort_list sorts lists up to this size in LDS

__device__ __forceinline__ uint32_t movable_or_none(const BodiesDev& B, uint32_t b) {
    return (B.im[b] != 0.0 || B.ii0[b] != 0.0 || B.ii1[b] != 0.0 || B.ii2[b] != 0.0) ? b : NO_BODY;
}

// Runs visits 0..count-1 of a sequence on one wave (64 threads per block)
// with the sequential result: bodies(v, bi, bj) names the movable bodies
// visit v reads and writes (NO_BODY for none), run(v) performs it.  Visits
// go in windows of 64; a lane runs once the latest earlier lane of its
// window touching each of its bodies has run, so every body sees its visits
// in sequence order and nothing else is shared between visits.
template <class Bodies, class Run>
__device__ __forceinline__ void wave_in_order(uint32_t count, Bodies bodies, Run run) {
    const int lane = threadIdx.x & 63;
    for (uint32_t base = 0; base < count; base += 64) {
        const uint32_t v = base + (uint32_t)lane;
        const bool act = v < count;
        uint32_t bi = NO_BODY, bj = NO_BODY;
        if (act) bodies(v, bi, bj);
        int di = -1, dj = -1;
        for (int l = 0; l < 63; ++l) {
            const uint32_t oi = __shfl(bi, l), oj = __shfl(bj, l);
            if (l < lane) {
                if (bi != NO_BODY && (oi == bi || oj == bi)) di = l;
                if (bj != NO_BODY && (oi == bj || oj == bj)) dj = l;
            }
        }
        bool done = !act;
        uint64_t dm = __ballot(done);
        while (dm != ~0ull) {   // each pass runs at least the lowest lane not done
            const bool ready = !done && (di < 0 || ((dm >> di) & 1ull)) && (dj < 0 || ((dm >> dj) & 1ull));
            if (ready) {
                run(v);
                done = true;
            }
            __threadfence_block();   // this pass's stores are seen by the next
            dm = __ballot(done);
        }
    }
}

// Sorts an island's constraint list ascending (the sequential sweep order),
// one wave: a bitonic sort in LDS, or lane 0 for a very large island.
__device__ void wave_sort_list(uint32_t* list, uint32_t m, uint32_t* sk) {
    const int lane = threadIdx.x & 63;
    if (m <= WAVE_SORT) {
        uint32_t p2 = 1;
        while (p2 < m) p2 <<= 1;
        for (uint32_t t = lane; t < p2; t += 64) sk[t] = t < m ? list[t] : 0xFFFFFFFFu;
        __syncthreads();
        for (uint32_t k = 2; k <= p2; k <<= 1)
            for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
                for (uint32_t t = lane; t < p2; t += 64) {
                    const uint32_t l = t ^ jj;
                    if (l > t) {
                        const uint32_t a = sk[t], b = sk[l];
                        if ((a > b) == ((t & k) == 0)) { sk[t] = b; sk[l] = a; }
                    }
                }
                __syncthreads();
            }
        for (uint32_t t = lane; t < m; t += 64) list[t] = sk[t];
    } else if (lane == 0) {
        sort_segment(list, m);
    }
    __threadfence_block();
    __syncthreads();
}

constexpr int TAPE_DOUBLES = 14;   // pre-visit va, wa, vb, wb, nimp, timp
constexpr size_t ADJ_TAPE_STEP_MAX = size_t(2) << 30;   // bytes of one recomputed step's tape

// The adjoint's tape of a step's solve, written by the solve itself when the
// backward recomputes the step: per (iteration, constraint) the two bodies'
// velocities and the impulses before the visit (d == nullptr: not recorded).
struct Tape {
    double* d = nullptr;
    uint32_t stride = 0;   // constraints per iteration
    __device__ __forceinline__ void put(uint32_t it, uint32_t k, const BodyView& a, const BodyView& b, double nimp,
                                        double timp) const {
        if (!d) return;
        double* tp = d + ((size_t)it * stride + k) * TAPE_DOUBLES;
        tp[0] = a.v.x; tp[1] = a.v.y; tp[2] = a.v.z; tp[3] = a.w.x; tp[4] = a.w.y; tp[5] = a.w.z;
        tp[6] = b.v.x; tp[7] = b.v.y; tp[8] = b.v.z; tp[9] = b.w.x; tp[10] = b.w.y; tp[11] = b.w.z;
        tp[12] = nimp; tp[13] = timp;
    }
};

// Body-state store for the sequential island sweep in global memory.
struct GlobalStore {
    BodiesDev& B;
    __device__ BodyView load(uint32_t b) const { return load_view(B, b); }
    __device__ void store(uint32_t b, const BodyView& o) const { store_view(B, b, o); }
    __device__ V3 loadx(uint32_t b) const { return ld3(B.x0, B.x1, B.x2, b); }
    __device__ void storex(uint32_t b, V3 x) const { st3(B.x0, B.x1, B.x2, b, x); }
};

template <class Store>
__device__ __forceinline__ double C_im(const Store& S, uint32_t b);

// Sequential sweep of one island in constraint order (world.cpp:559-616).
template <class Store>
__device__ void solve_island_seq(const Store& S, const SimParams& P, const ConsDev& C, const uint32_t* list, uint32_t m,
                                 const Tape& tape = Tape{}) {
    if (P.warm) {
        for (uint32_t t = 0; t < m; ++t) {
            const uint32_t k = list[t];
            const double imp = C.nimp[k];
            if (imp == 0.0) continue;
            const uint32_t i = C.ci[k], j = C.cj[k];
            BodyView a = S.load(i), b = S.load(j);
            warm_apply(a, b, load_geom(C, k), imp);
            S.store(i, a);
            S.store(j, b);
        }
    }
    for (int it = 0; it < P.iters; ++it) {
        for (uint32_t t = 0; t < m; ++t) {
            const uint32_t k = list[t];
            const uint32_t i = C.ci[k], j = C.cj[k];
            BodyView a = S.load(i), b = S.load(j);
            double nimp = C.nimp[k], timp = C.timp[k];
            tape.put((uint32_t)it, k, a, b, nimp, timp);
            gs_visit(a, b, load_geom(C, k), C.rbias[k], C.kn[k], C.mu[k], nimp, timp);
            C.nimp[k] = nimp;
            C.timp[k] = timp;
            S.store(i, a);
            S.store(j, b);
        }
    }
    for (uint32_t t = 0; t < m; ++t) {
        const uint32_t k = list[t];
        const uint32_t i = C.ci[k], j = C.cj[k];
        V3 xa = S.loadx(i), xb = S.loadx(j);
        const double ima = C_im(S, i), imb = C_im(S, j);
        correct(S.B, C, k, P.beta, xa, xb, ima, imb);
        if (ima != 0.0) S.storex(i, xa);
        if (imb != 0.0) S.storex(j, xb);
    }
}

template <class Store>
__device__ __forceinline__ double C_im(const Store& S, uint32_t b) { return S.B.im[b]; }

// The same sweep on one wave (64 threads per block), with the sequential
// result bit for bit: each pass goes through wave_in_order, so every movable
// body sees its visits in constraint order, and a constraint between two
// bodies that cannot move still orders its own visits (its impulses).
__device__ void solve_island_wave(BodiesDev& B, const SimParams& P, const ConsDev& C, const uint32_t* list, uint32_t m) {
    auto bodies_of = [&](uint32_t t, uint32_t& bi, uint32_t& bj) {
        const uint32_t k = list[t];
        bi = movable_or_none(B, C.ci[k]);
        bj = movable_or_none(B, C.cj[k]);
        if (bi == NO_BODY && bj == NO_BODY) bi = 0x80000000u | k;
    };
    if (P.warm)
        wave_in_order(
            m,
            [&](uint32_t t, uint32_t& bi, uint32_t& bj) {
                if (C.nimp[list[t]] != 0.0) bodies_of(t, bi, bj);
            },
            [&](uint32_t t) {
                const uint32_t k = list[t];
                const double imp = C.nimp[k];
                if (imp == 0.0) return;
                const uint32_t i = C.ci[k], j = C.cj[k];
                BodyView a = load_view(B, i), b = load_view(B, j);
                warm_apply(a, b, load_geom(C, k), imp);
                store_view(B, i, a);
                store_view(B, j, b);
            });
    wave_in_order(
        (uint32_t)P.iters * m, [&](uint32_t v, uint32_t& bi, uint32_t& bj) { bodies_of(v % m, bi, bj); },
        [&](uint32_t v) {
            const uint32_t k = list[v % m];
            const uint32_t i = C.ci[k], j = C.cj[k];
            BodyView a = load_view(B, i), b = load_view(B, j);
            double nimp = C.nimp[k], timp = C.timp[k];
            gs_visit(a, b, load_geom(C, k), C.rbias[k], C.kn[k], C.mu[k], nimp, timp);
            C.nimp[k] = nimp;
            C.timp[k] = timp;
            store_view(B, i, a);
            store_view(B, j, b);
        });
    wave_in_order(
        m, bodies_of,
        [&](uint32_t t) {
            const uint32_t k = list[t];
            const uint32_t i = C.ci[k], j = C.cj[k];
            V3 xa = ld3(B.x0, B.x1, B.x2, i), xb = ld3(B.x0, B.x1, B.x2, j);
            const double ima = B.im[i], imb = B.im[j];
            correct(B, C, k, P.beta, xa, xb, ima, imb);
            if (ima != 0.0) st3(B.x0, B.x1, B.x2, i, xa);
            if (imb != 0.0) st3(B.x0, B.x1, B.x2, j, xb);
        });
}

// An island whose constraints all join the same two bodies (a sphere pair,
// a box-box manifold: <= 4 contacts): both bodies and the impulses stay in
// registers for the whole solve.
__device__ void solve_pair_island(BodiesDev& B, const SimParams& P, const ConsDev& C, const uint32_t* list, uint32_t m,
                                  const Tape& tape) {
    const uint32_t k0 = list[0];
    const uint32_t i = C.ci[k0], j = C.cj[k0];
    BodyView a = load_view(B, i), b = load_view(B, j);
    double nimp[4], timp[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        nimp[t] = t < (int)m ? C.nimp[list[t]] : 0.0;
        timp[t] = 0.0;
    }
    if (P.warm) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (t < (int)m) warm_apply(a, b, load_geom(C, list[t]), nimp[t]);
    }
    for (int it = 0; it < P.iters; ++it) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (t < (int)m) {
                const uint32_t k = list[t];
                tape.put((uint32_t)it, k, a, b, nimp[t], timp[t]);
                gs_visit(a, b, load_geom(C, k), C.rbias[k], C.kn[k], C.mu[k], nimp[t], timp[t]);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
        if (t < (int)m) { C.nimp[list[t]] = nimp[t]; C.timp[list[t]] = timp[t]; }
    store_view(B, i, a);
    store_view(B, j, b);
    V3 xa = ld3(B.x0, B.x1, B.x2, i), xb = ld3(B.x0, B.x1, B.x2, j);
#pragma unroll
    for (int t = 0; t < 4; ++t)
        if (t < (int)m) correct(B, C, list[t], P.beta, xa, xb, a.im, b.im);
    if (a.im != 0.0) st3(B.x0, B.x1, B.x2, i, xa);
    if (b.im != 0.0) st3(B.x0, B.x1, B.x2, j, xb);
}

// One thread per small island.
// This step's (key, slot)-ordered normal impulses become the next step's warm
// cache (src/world.cpp:619-626): each constraint keeps its global position.
__device__ __forceinline__ void publish_cache(const ConsDev& C, const uint32_t* list, uint32_t m,
                                              uint64_t* __restrict__ prev_key, double* __restrict__ prev_imp) {
    for (uint32_t t = 0; t < m; ++t) {
        const uint32_t k = list[t];
        prev_key[k] = C.key[k];
        prev_imp[k] = C.nimp[k];
    }
}

// Islands of exactly one constraint (isolated contacts: most of a gas): the
// two bodies, the constraint and its impulses stay in registers.
__global__ void __launch_bounds__(TPB) k_solve_single(BodiesDev B, SimParams P, const uint32_t* __restrict__ counters,
                                                      const uint32_t* __restrict__ isl_start,
                                                      const uint32_t* __restrict__ isl_list, ConsDev C,
                                                      uint64_t* __restrict__ prev_key, double* __restrict__ prev_imp,
                                                      Tape tape) {
    const uint32_t nisl = counters[CNT_ISL];
    for (uint32_t isl = blockIdx.x * blockDim.x + threadIdx.x; isl < nisl; isl += gridDim.x * blockDim.x) {
        const uint32_t s = isl_start[isl];
        if (isl_start[isl + 1] - s != 1) continue;
        const uint32_t k = isl_list[s];
        const uint32_t i = C.ci[k], j = C.cj[k];
        BodyView a = load_view(B, i), b = load_view(B, j);
        const ConsGeom g = load_geom(C, k);
        const double rbias = C.rbias[k], kn = C.kn[k], mu = C.mu[k];
        double nimp = C.nimp[k], timp = 0.0;
        if (P.warm) warm_apply(a, b, g, nimp);
        for (int it = 0; it < P.iters; ++it) {
            tape.put((uint32_t)it, k, a, b, nimp, timp);
            gs_visit(a, b, g, rbias, kn, mu, nimp, timp);
        }
        C.nimp[k] = nimp;
        C.timp[k] = timp;
        store_view(B, i, a);
        store_view(B, j, b);
        V3 xa = ld3(B.x0, B.x1, B.x2, i), xb = ld3(B.x0, B.x1, B.x2, j);
        correct(B, C, k, P.beta, xa, xb, a.im, b.im);
        if (a.im != 0.0) st3(B.x0, B.x1, B.x2, i, xa);
        if (b.im != 0.0) st3(B.x0, B.x1, B.x2, j, xb);
        if (P.warm) { prev_key[k] = C.key[k]; prev_imp[k] = nimp; }
    }
}

// Small islands (2..STAGE_CONS constraints over <= STAGE_BODIES bodies): one
// thread stages the whole island -- body state, constraint geometry and
// impulses -- in its slice of LDS and sweeps it there, so the sequential
// Gauss-Seidel chain never waits on global memory.  Larger small islands take
// the global-memory sweep.
constexpr int SOLVE_THREADS = 32;
// Measured on the 1024-world batch (captured steps): staging up to 4
// constraints (40 KiB of LDS a block, four blocks a CU) and handing larger
// islands to k_solve_large's waves took the step from 9.65 to 8.09 ms; 16
// constraints (146 KiB, one block a CU) left most CUs' SIMDs idle.
constexpr int STAGE_BODIES = 5;   // 4 constraints join at most 5 bodies
constexpr int STAGE_CONS = 4;

struct StagedBody {
    V3 v, w, x, ii;
    Qt q;
    double im;
    int mv, mw;
};
struct StagedCons {
    V3 n, rA, rB;
    double rbias, kn, mu, ov, nimp, timp;
    int a, b;   // staged body slots
};

__device__ __forceinline__ BodyView view_of(const StagedBody& s) {
    BodyView o;
    o.v = s.v; o.w = s.w; o.q = s.q; o.ii = s.ii; o.im = s.im; o.mv = s.mv != 0; o.mw = s.mw != 0;
    return o;
}

__device__ void solve_island_staged(BodiesDev& B, const SimParams& P, ConsDev& C, const uint32_t* list, uint32_t m,
                                    const uint32_t* ids, int nb, StagedBody* sb, StagedCons* sc, const Tape& tape) {
    for (int u = 0; u < nb; ++u) {
        const uint32_t b = ids[u];
        BodyView o = load_view(B, b);
        sb[u].v = o.v; sb[u].w = o.w; sb[u].q = o.q; sb[u].ii = o.ii; sb[u].im = o.im;
        sb[u].mv = o.mv; sb[u].mw = o.mw;
        sb[u].x = ld3(B.x0, B.x1, B.x2, b);
    }
    for (uint32_t t = 0; t < m; ++t) {
        const uint32_t k = list[t];
        const ConsGeom g = load_geom(C, k);
        StagedCons& c = sc[t];
        c.n = g.n; c.rA = g.rA; c.rB = g.rB;
        c.rbias = C.rbias[k]; c.kn = C.kn[k]; c.mu = C.mu[k]; c.ov = C.ov[k];
        c.nimp = C.nimp[k]; c.timp = 0.0;
        const uint32_t i = C.ci[k], j = C.cj[k];
        int sa = 0, sbb = 0;
        for (int u = 0; u < nb; ++u) { sa = ids[u] == i ? u : sa; sbb = ids[u] == j ? u : sbb; }
        c.a = sa; c.b = sbb;
    }
    if (P.warm) {
        for (uint32_t t = 0; t < m; ++t) {
            StagedCons& c = sc[t];
            if (c.nimp == 0.0) continue;
            BodyView a = view_of(sb[c.a]), b = view_of(sb[c.b]);
            warm_apply(a, b, ConsGeom{c.n, c.rA, c.rB}, c.nimp);
            sb[c.a].v = a.v; sb[c.a].w = a.w; sb[c.b].v = b.v; sb[c.b].w = b.w;
        }
    }
    for (int it = 0; it < P.iters; ++it) {
        for (uint32_t t = 0; t < m; ++t) {
            StagedCons& c = sc[t];
            BodyView a = view_of(sb[c.a]), b = view_of(sb[c.b]);
            double nimp = c.nimp, timp = c.timp;
            tape.put((uint32_t)it, list[t], a, b, nimp, timp);
            gs_visit(a, b, ConsGeom{c.n, c.rA, c.rB}, c.rbias, c.kn, c.mu, nimp, timp);
            c.nimp = nimp; c.timp = timp;
            sb[c.a].v = a.v; sb[c.a].w = a.w; sb[c.b].v = b.v; sb[c.b].w = b.w;
        }
    }
    for (uint32_t t = 0; t < m; ++t) {           // contact.hpp:78-85
        StagedCons& c = sc[t];
        const double ima = sb[c.a].im, imb = sb[c.b].im;
        const double invSum = ima + imb;
        if (invSum <= 0.0 || c.ov <= 0.0) continue;
        const V3 corr = c.n * (P.beta * c.ov / invSum);
        if (ima != 0.0) sb[c.a].x = sb[c.a].x + corr * ima;
        if (imb != 0.0) sb[c.b].x = sb[c.b].x - corr * imb;
    }
    for (uint32_t t = 0; t < m; ++t) { C.nimp[list[t]] = sc[t].nimp; C.timp[list[t]] = sc[t].timp; }
    for (int u = 0; u < nb; ++u) {               // write back what may change (store_view's rule)
        const uint32_t b = ids[u];
        if (sb[u].mv) { st3(B.v0, B.v1, B.v2, b, sb[u].v); st3(B.x0, B.x1, B.x2, b, sb[u].x); }
        if (sb[u].mw) st3(B.w0, B.w1, B.w2, b, sb[u].w);
    }
}


__global__ void __launch_bounds__(SOLVE_THREADS) k_solve_small(BodiesDev B, SimParams P, uint32_t* counters,
                                                     const uint32_t* __restrict__ isl_start, uint32_t* __restrict__ isl_list,
                                                     ConsDev C, uint32_t* __restrict__ big_list,
                                                     uint64_t* __restrict__ prev_key, double* __restrict__ prev_imp,
                                                     uint32_t* __restrict__ prev_n, uint32_t big_above, Tape tape) {
    __shared__ uint32_t lds_ids[SOLVE_THREADS * STAGE_BODIES];
    __shared__ StagedBody lds_b[SOLVE_THREADS * STAGE_BODIES];
    __shared__ StagedCons lds_c[SOLVE_THREADS * STAGE_CONS];
    const uint32_t nisl = counters[CNT_ISL];
    for (uint32_t isl = blockIdx.x * blockDim.x + threadIdx.x; isl < nisl; isl += gridDim.x * blockDim.x) {
        const uint32_t s = isl_start[isl], e = isl_start[isl + 1];
        const uint32_t m = e - s;
        if (m == 1) continue;   // k_solve_single
        if (m > big_above) {
            big_list[atomicAdd(&counters[CNT_NBIG], 1u)] = isl;
            atomicMax(&counters[CNT_MAXISL], m);
            continue;
        }
        uint32_t* list = isl_list + s;
        sort_segment(list, m);
        bool pair = m <= 4;
        if (pair) {
            const uint32_t i0 = C.ci[list[0]], j0 = C.cj[list[0]];
            for (uint32_t t = 1; t < m; ++t) pair = pair && C.ci[list[t]] == i0 && C.cj[list[t]] == j0;
        }
        if (pair) {
            solve_pair_island(B, P, C, list, m, tape);
        } else {
            uint32_t* ids = lds_ids + threadIdx.x * STAGE_BODIES;
            int nb = 0;
            bool fits = m <= (uint32_t)STAGE_CONS;
            for (uint32_t t = 0; t < m && fits; ++t) {
                const uint32_t k = list[t];
                const uint32_t e2[2] = {C.ci[k], C.cj[k]};
                for (int q = 0; q < 2; ++q) {
                    bool seen = false;
                    for (int u = 0; u < nb; ++u) seen |= ids[u] == e2[q];
                    if (!seen) {
                        if (nb == STAGE_BODIES) { fits = false; break; }
                        ids[nb++] = e2[q];
                    }
                }
            }
            if (fits)
                solve_island_staged(B, P, C, list, m, ids, nb, lds_b + threadIdx.x * STAGE_BODIES,
                                    lds_c + threadIdx.x * STAGE_CONS, tape);
            else
                solve_island_seq(GlobalStore{B}, P, C, list, m, tape);
        }
        if (P.warm) publish_cache(C, list, m, prev_key, prev_imp);
    }
    if (P.warm && blockIdx.x == 0 && threadIdx.x == 0) *prev_n = counters[CNT_C];
}

// One wave per large island (captured steps, which cannot read back the
// island sizes for the dataflow solve): sort the segment, then sweep it.
__global__ void __launch_bounds__(64) k_solve_large(BodiesDev B, SimParams P, uint32_t* counters,
                                                    const uint32_t* __restrict__ isl_start, uint32_t* __restrict__ isl_list,
                                                    ConsDev C, const uint32_t* __restrict__ big_list,
                                                    uint64_t* __restrict__ prev_key, double* __restrict__ prev_imp) {
    __shared__ uint32_t sk[WAVE_SORT];
    const uint32_t nbig = counters[CNT_NBIG];
    for (uint32_t b = blockIdx.x; b < nbig; b += gridDim.x) {
        const uint32_t isl = big_list[b];
        const uint32_t s = isl_start[isl], e = isl_start[isl + 1];
        const uint32_t m = e - s;
        uint32_t* list = isl_list + s;
        wave_sort_list(list, m, sk);
        solve_island_wave(B, P, C, list, m);
        if (P.warm)
            for (uint32_t t = threadIdx.x; t < m; t += 64) {
                const uint32_t k = list[t];
                prev_key[k] = C.key[k];
                prev_imp[k] = C.nimp[k];
            }
        __syncthreads();   // the next island's sort reuses the LDS
    }
}

// ---------------------------------------------------------------------------
// Islands of more than WF_ISLAND constraints in eager steps (or every
// constraint): the sweep's iterations as a dataflow of (iteration, pair)
// visits, one round of independent visits per launch.
//
// A task is one pair's run of constraints (consecutive in (i, j, slot)
// order); a visit is one iteration (world.cpp:572-612) over one task.  On
// each movable body the tasks of an iteration are chained in constraint
// order, and the body's first task of iteration p+1 follows its last task of
// iteration p.  A visit runs in the round after both of its predecessors
// ran, so the visits of one round touch disjoint movable bodies and every
// body receives exactly the sequential sweep's updates in the sequential
// order: the result is the reference's bit for bit.  Static bodies (nothing
// writes them) chain nothing.  The warm start (559-568) and the positional
// correction (615-616) need no such order across bodies: each body applies
// its own share of them, in constraint order, in k_wf_link and k_wf_correct.
// ---------------------------------------------------------------------------

constexpr uint32_t WF_WRAP = 0x80000000u;   // successor is the body's first task of the next pass
constexpr uint32_t WF_NONE = 0xFFFFFFFFu;   // no successor on this side (a static body)
constexpr uint32_t WF_MAX_ROUNDS = 1u << 16;
constexpr int WF_THREADS = 64;
constexpr uint32_t WF_PASS_MAX = 64;        // iterations
constexpr uint32_t WF_ISLAND = 16;          // eager steps: islands above this many constraints take the dataflow solve
constexpr uint32_t CAP_ISLAND = STAGE_CONS;  // captured steps: islands above this many constraints take k_solve_large
// When the large islands hold at least WF_ALL_NUM / WF_ALL_DEN of the
// constraints, the next eager steps put every constraint in the dataflow and
// skip building islands (union-find, island ids and lists, the per-island
// solvers); every WF_ALL_RECHECK-th step builds them again to check.  The
// small islands' visits widen the dataflow's early rounds but add no depth
// beyond the large islands'.
constexpr uint32_t WF_ALL_NUM = 3, WF_ALL_DEN = 4;
constexpr uint32_t WF_ALL_RECHECK = 32;

struct WfDev {
    uint32_t* flag;        // per constraint: 1 if it starts a task of a large island
    uint32_t* tid;         // per constraint: exclusive scan of flag (its task index)
    uint32_t* task_k;      // per task: first constraint
    uint32_t* pend0;       // per task: predecessors in pass 0
    uint32_t* pend;        // per visit (pass * T + task): predecessors not yet run
    uint4* front;          // the visits queued for a launch: two halves of T entries (launch r reads half r & 1 and
                           // fills the other; at most one visit per task is ready at a time), 32 bytes each: the
                           // task's first line, then (successor side 0, successor side 1, task, pass)
    uint32_t* rcnt;        // per round: visits queued for it
    uint32_t* done;        // visits run so far this step (one counter)
    uint32_t* inc_cnt;     // per body: tasks touching it (zero between steps)
    uint32_t* inc_start;   // per body: start of its slice of inc_list
    uint32_t* inc_list;    // per movable body: its tasks
    uint32_t* task_n;      // per task: constraints in it
    uint4* task4;          // per task, 32 bytes: (first constraint, constraints, body i, body j), then the
                           // successor task per side (0 = body i, 1 = body j: a task, WF_WRAP | first, or
                           // WF_NONE) -- one cache line a visit
    double* rec;           // per constraint of a task: n, rA, rB, restBias, Kn, mu, nImp, tImp, overlap (16 doubles)
    double* bp;            // per body: v, w, q, invI, invMass (16 doubles), the sweep's working copy
    Tape tape;             // the adjoint's tape of this solve, when the backward recomputes the step
};


__device__ __forceinline__ BodyView load_packed(const double* __restrict__ bp, uint32_t b) {
    const double2* r = reinterpret_cast<const double2*>(bp + (size_t)b * WF_REC);
    const double2 a0 = r[0], a1 = r[1], a2 = r[2], a3 = r[3], a4 = r[4], a5 = r[5], a6 = r[6];
    BodyView o;
    o.v = V3{a0.x, a0.y, a1.x};
    o.w = V3{a1.y, a2.x, a2.y};
    o.q = Qt{a3.x, a3.y, a4.x, a4.y};
    o.ii = V3{a5.x, a5.y, a6.x};
    o.im = a6.y;
    o.mv = o.im != 0.0;
    o.mw = o.ii.x != 0.0 || o.ii.y != 0.0 || o.ii.z != 0.0;
    return o;
}

__device__ __forceinline__ void store_packed(double* __restrict__ bp, uint32_t b, const BodyView& o) {
    double* r = bp + (size_t)b * WF_REC;
    if (o.mv) { r[0] = o.v.x; r[1] = o.v.y; r[2] = o.v.z; }
    if (o.mw) { r[3] = o.w.x; r[4] = o.w.y; r[5] = o.w.z; }
}

__global__ void __launch_bounds__(TPB) k_wf_pack_bodies(BodiesDev B, double* __restrict__ bp) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B.n) return;
    double2* r = reinterpret_cast<double2*>(bp + (size_t)b * WF_REC);
    r[0] = double2{B.v0[b], B.v1[b]};
    r[1] = double2{B.v2[b], B.w0[b]};
    r[2] = double2{B.w1[b], B.w2[b]};
    r[3] = double2{B.q0[b], B.q1[b]};
    r[4] = double2{B.q2[b], B.q3[b]};
    r[5] = double2{B.ii0[b], B.ii1[b]};
    r[6] = double2{B.ii2[b], B.im[b]};
}

// Back to the body fields: what the sweep may have changed (store_view's rule).
__global__ void __launch_bounds__(TPB) k_wf_unpack_bodies(BodiesDev B, const double* __restrict__ bp) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B.n) return;
    const double* r = bp + (size_t)b * WF_REC;
    if (B.im[b] != 0.0) { B.v0[b] = r[0]; B.v1[b] = r[1]; B.v2[b] = r[2]; }
    if (B.ii0[b] != 0.0 || B.ii1[b] != 0.0 || B.ii2[b] != 0.0) { B.w0[b] = r[3]; B.w1[b] = r[4]; B.w2[b] = r[5]; }
}

// Flags the first constraint of each contact pair in an island of more than
// `big_above` constraints and counts those islands' constraints (one atomic
// per block), or, when `cons_isl` is null (the whole solve as one dataflow,
// no islands built), flags every pair and publishes the warm cache's length,
// which k_solve_small does otherwise.
__global__ void __launch_bounds__(TPB) k_wf_mark(BodiesDev B, uint32_t* __restrict__ counters, ConsDev C,
                                                 const uint32_t* __restrict__ cons_isl,
                                                 const uint32_t* __restrict__ isl_start, uint32_t* __restrict__ flag,
                                                 uint32_t big_above, uint32_t* __restrict__ prev_n) {
    __shared__ uint32_t part[TPB / 64];
    const uint32_t nc = counters[CNT_C];
    if (prev_n && blockIdx.x == 0 && threadIdx.x == 0) *prev_n = nc;
    uint32_t nbig = 0;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
        const bool start = k == 0 || C.ci[k] != C.ci[k - 1] || C.cj[k] != C.cj[k - 1];
        bool big = true;
        if (cons_isl) {
            const uint32_t id = cons_isl[k];
            big = isl_start[id + 1] - isl_start[id] > big_above;
            nbig += big ? 1u : 0u;
        }
        flag[k] = (start && big) ? 1u : 0u;
    }
    if (!cons_isl) return;
    for (int off = 32; off > 0; off >>= 1) nbig += __shfl_xor(nbig, off);
    if (lane_id() == 0) part[threadIdx.x >> 6] = nbig;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
        for (int w = 0; w < TPB / 64; ++w) sum += part[w];
        if (sum) atomicAdd(&counters[CNT_WFC], sum);
    }
}

// `packed`: k_cons_build wrote the constraints' records already (the whole
// solve as one dataflow).
__global__ void __launch_bounds__(TPB) k_wf_tasks(BodiesDev B, const uint32_t* __restrict__ counters, ConsDev C,
                                                  WfDev W, bool packed) {
    const uint32_t nc = counters[CNT_C];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
        if (!W.flag[k]) continue;
        const uint32_t t = W.tid[k];
        W.task_k[t] = k;
        const uint32_t i = C.ci[k], j = C.cj[k];
        const bool mi = mutable_body(B, i), mj = mutable_body(B, j);
        if (mi) atomicAdd(&W.inc_cnt[i], 1u);
        if (mj) atomicAdd(&W.inc_cnt[j], 1u);
        uint32_t k1 = k + 1;
        while (k1 < nc && C.ci[k1] == i && C.cj[k1] == j) ++k1;
        W.task_n[t] = k1 - k;
        W.task4[2 * (size_t)t] = make_uint4(k, k1 - k, i, j);
        // a pair of two bodies that cannot move (the pair filter drops them,
        // world.cpp:513) would still order its own passes: it succeeds itself
        W.task4[2 * (size_t)t + 1] = make_uint4((mi || mj) ? WF_NONE : (WF_WRAP | t), WF_NONE, 0u, 0u);
        if (packed) continue;
        for (uint32_t q = k; q < k1; ++q) {   // the sweep's record of constraint q
            double2* r = reinterpret_cast<double2*>(W.rec + (size_t)q * WF_REC);
            r[0] = double2{C.n0[q], C.n1[q]};
            r[1] = double2{C.n2[q], C.ra0[q]};
            r[2] = double2{C.ra1[q], C.ra2[q]};
            r[3] = double2{C.rb0[q], C.rb1[q]};
            r[4] = double2{C.rb2[q], C.rbias[q]};
            r[5] = double2{C.kn[q], C.mu[q]};
            r[6] = double2{C.nimp[q], C.timp[q]};
            r[7] = double2{C.ov[q], 0.0};
        }
    }
}

// inc_cnt was zeroed by the scan that made inc_start: it counts the fill here.
__global__ void __launch_bounds__(TPB) k_wf_fill(BodiesDev B, uint32_t T, ConsDev C, WfDev W) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
        const uint32_t k = W.task_k[t];
        const uint32_t i = C.ci[k], j = C.cj[k];
        if (mutable_body(B, i)) W.inc_list[W.inc_start[i] + atomicAdd(&W.inc_cnt[i], 1u)] = t;
        if (mutable_body(B, j)) W.inc_list[W.inc_start[j] + atomicAdd(&W.inc_cnt[j], 1u)] = t;
    }
}

// One body's share of apply_impulse (solve: a.v += J m_a^-1, a.w += I_a^-1
// (rA x J) on side a, the opposite on side b), in the same operations.
__device__ __forceinline__ void apply_side(BodyView& x, bool side_a, V3 r, V3 J) {
    if (side_a) {
        if (x.mv) x.v = x.v + J * x.im;
        if (x.mw) x.w = x.w + apply_inv_inertia(x.q, x.ii, cross(r, J));
    } else {
        if (x.mv) x.v = x.v - J * x.im;
        if (x.mw) x.w = x.w - apply_inv_inertia(x.q, x.ii, cross(r, J));
    }
}

// Per movable body: its tasks in order, chained; the last links to the first
// of the next iteration.  With warm starting, the body's share of each warm
// impulse (world.cpp:559-568) is applied here, in constraint order: what
// constraint k adds to a body depends only on k and on that body's
// orientation, inverse mass and inertia, which the solve does not change, so
// the warm start orders nothing across bodies.  Leaves inc_cnt zero for the
// next step.
__global__ void __launch_bounds__(TPB) k_wf_link(uint32_t n, SimParams P, WfDev W) {
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += gridDim.x * blockDim.x) {
        const uint32_t m = W.inc_cnt[x];
        if (m == 0) continue;
        W.inc_cnt[x] = 0u;
        uint32_t* a = W.inc_list + W.inc_start[x];
        sort_segment(a, m);
        BodyView bx{};
        if (P.warm) bx = load_packed(W.bp, x);
        for (uint32_t u = 0; u < m; ++u) {
            const uint32_t t = a[u];
            const uint4 tk = W.task4[2 * (size_t)t];
            const uint32_t side = tk.z == x ? 0u : 1u;
            reinterpret_cast<uint32_t*>(W.task4 + 2 * (size_t)t + 1)[side] = (u + 1 < m) ? a[u + 1] : (WF_WRAP | a[0]);
            if (u > 0) atomicAdd(&W.pend0[t], 1u);
            if (!P.warm) continue;
            for (uint32_t k = tk.x; k < tk.x + tk.y; ++k) {
                const double2* r = reinterpret_cast<const double2*>(W.rec + (size_t)k * WF_REC);
                const double nimp = r[6].x;
                if (nimp == 0.0) continue;   // warm_apply
                const double2 r0 = r[0], r1 = r[1];
                const V3 nrm{r0.x, r0.y, r1.x};
                V3 arm;
                if (side == 0) {
                    const double2 r2 = r[2];
                    arm = V3{r1.y, r2.x, r2.y};
                } else {
                    const double2 r3 = r[3], r4 = r[4];
                    arm = V3{r3.x, r3.y, r4.x};
                }
                apply_side(bx, side == 0, arm, nrm * nimp);
            }
        }
        if (P.warm) store_packed(W.bp, x, bx);
    }
}

// After the iterations, per movable body: its positional correction
// (world.cpp:615-616, contact.hpp:78-85) from its constraints in order.
// What constraint k moves a body by depends only on k and the two inverse
// masses, so, like the warm start, it orders nothing across bodies.
__global__ void __launch_bounds__(TPB) k_wf_correct(BodiesDev B, SimParams P, WfDev W) {
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < B.n; x += gridDim.x * blockDim.x) {
        const uint32_t s0 = W.inc_start[x], m = W.inc_start[x + 1] - s0;
        if (m == 0) continue;
        const double imx = B.im[x];
        if (imx == 0.0) continue;   // its position never changes
        V3 px = ld3(B.x0, B.x1, B.x2, x);
        for (uint32_t u = 0; u < m; ++u) {
            const uint4 tk = W.task4[2 * (size_t)W.inc_list[s0 + u]];
            const bool side_a = tk.z == x;
            const double invSum = B.im[tk.z] + B.im[tk.w];
            if (invSum <= 0.0) continue;
            for (uint32_t k = tk.x; k < tk.x + tk.y; ++k) {
                const double2* r = reinterpret_cast<const double2*>(W.rec + (size_t)k * WF_REC);
                const double ov = r[7].x;
                if (ov <= 0.0) continue;
                const double2 r0 = r[0], r1 = r[1];
                const V3 corr = V3{r0.x, r0.y, r1.x} * (P.beta * ov / invSum);
                px = side_a ? px + corr * imx : px - corr * imx;
            }
        }
        st3(B.x0, B.x1, B.x2, x, px);
    }
}

// The converged impulses back to the constraint list and, with warm
// starting, into the next step's cache (world.cpp:619-626).
__global__ void __launch_bounds__(TPB) k_wf_out(uint32_t T, SimParams P, ConsDev C, WfDev W,
                                                uint64_t* __restrict__ prev_key, double* __restrict__ prev_imp) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
        const uint4 tk = W.task4[2 * (size_t)t];
        for (uint32_t k = tk.x; k < tk.x + tk.y; ++k) {
            const double2 r6 = reinterpret_cast<const double2*>(W.rec + (size_t)k * WF_REC)[6];
            C.nimp[k] = r6.x;
            C.timp[k] = r6.y;
            if (P.warm) { prev_key[k] = C.key[k]; prev_imp[k] = r6.x; }
        }
    }
}

// Pending counts of every visit (pass p = iteration p + 1); the first
// iteration's visits with no predecessor are round 0.
__global__ void __launch_bounds__(TPB) k_wf_init(BodiesDev B, uint32_t T, uint32_t npass, ConsDev C, WfDev W) {
    for (uint32_t t0 = blockIdx.x * blockDim.x; t0 < T; t0 += gridDim.x * blockDim.x) {
        const uint32_t t = t0 + threadIdx.x;
        bool root = false;
        if (t < T) {
            const uint32_t k = W.task_k[t];
            const uint32_t sides = max((mutable_body(B, C.ci[k]) ? 1u : 0u) + (mutable_body(B, C.cj[k]) ? 1u : 0u), 1u);
            const uint32_t p0 = W.pend0[t];
            W.pend[t] = p0;
            for (uint32_t p = 1; p < npass; ++p) W.pend[(size_t)p * T + t] = sides;
            root = p0 == 0u;
        }
        const uint32_t pos = wave_add1(W.rcnt, 0u, root);
        if (root) {
            W.front[2 * (size_t)pos] = W.task4[2 * (size_t)t];
            const uint4 nx = W.task4[2 * (size_t)t + 1];
            W.front[2 * (size_t)pos + 1] = make_uint4(nx.x, nx.y, t, 0u);
        }
    }
}

// One iteration over one task's constraints [k0, k1) on the packed records
// (world.cpp:572-612): the two bodies stay in registers across the run, as
// in the pair solver.
// A visit's inputs: its two bodies and its first constraint's record.
struct WfLoaded {
    BodyView a, b;
    double2 r0, r1, r2, r3, r4, r5, r6;
};

__device__ __forceinline__ void wf_visit_load(const WfDev& W, uint32_t i, uint32_t j, uint32_t k0, WfLoaded& L) {
    L.a = load_packed(W.bp, i);
    L.b = load_packed(W.bp, j);
    const double2* r = reinterpret_cast<const double2*>(W.rec + (size_t)k0 * WF_REC);
    L.r0 = r[0]; L.r1 = r[1]; L.r2 = r[2]; L.r3 = r[3]; L.r4 = r[4]; L.r5 = r[5]; L.r6 = r[6];
}

__device__ __forceinline__ void wf_visit(const WfDev& W, const WfLoaded& L, uint32_t pass, uint32_t i, uint32_t j,
                                         uint32_t k0, uint32_t k1) {
    BodyView a = L.a, b = L.b;
    // the next record is loaded while this one is solved (the records do not
    // depend on the sweep's results; only nImp/tImp are written)
    double2 r0 = L.r0, r1 = L.r1, r2 = L.r2, r3 = L.r3, r4 = L.r4, r5 = L.r5, r6 = L.r6;
    for (uint32_t k = k0; k < k1; ++k) {
        const ConsGeom g{V3{r0.x, r0.y, r1.x}, V3{r1.y, r2.x, r2.y}, V3{r3.x, r3.y, r4.x}};
        const double rbias = r4.y, kn = r5.x, mu = r5.y;
        double nimp = r6.x, timp = r6.y;
        if (k + 1 < k1) {
            const double2* rn = reinterpret_cast<const double2*>(W.rec + (size_t)(k + 1) * WF_REC);
            r0 = rn[0]; r1 = rn[1]; r2 = rn[2]; r3 = rn[3]; r4 = rn[4]; r5 = rn[5]; r6 = rn[6];
        }
        W.tape.put(pass, k, a, b, nimp, timp);
        gs_visit(a, b, g, rbias, kn, mu, nimp, timp);
        reinterpret_cast<double2*>(W.rec + (size_t)k * WF_REC)[6] = double2{nimp, timp};
    }
    store_packed(W.bp, i, a);
    store_packed(W.bp, j, b);
}

// k_solve_small counted the islands above its limit (CNT_NBIG).
__global__ void k_wf_expect_none(uint32_t* counters) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && counters[CNT_NBIG] != 0u) atomicOr(&counters[CNT_FAULT], NE_FAULT_WF_PLAN);
}

// Launch r: runs the visits queued for it, then, in the same block, up to
// WF_LEVELS - 1 further levels of the successors that became ready there.
//
// A visit may run in the block that made it ready only if every predecessor
// that ran in this launch ran in this block: that block's earlier writes are
// visible to it after a barrier (one CU), while another block's are not
// until the launch ends.  So each pending count also records who released
// the visit: a predecessor adds (tag << 2) - 1, tag = (launch + 1) << 12 |
// block, to the count (at most 2, in the low two bits); the predecessor that
// takes it to zero sees the other one's tag in the old value (0 when it ran
// in no launch yet, or never existed).  A visit released by another block in
// this launch, or past the last local level or the block's list capacity,
// is queued for launch r + 1 instead.  Every visit still runs after all of
// its predecessors, on bodies no other ready visit touches, so the result is
// the sequential sweep's bit for bit.
constexpr int WF_LEVELS = 4;     // levels a block runs per launch when rounds are wide
constexpr int WF_LCAP = 512;     // a block's list of locally ready visits (per level)
// Measured on one MI355X: local levels help where rounds are wide (the 1M
// terrain scene, about 78k visits a round: 55.5 -> 52.7 ms a step) and cost
// where they are narrow (the 65,536-body pile, about 1,900 visits a round:
// 23.2 -> 25.9 ms), where a visit's chain of dependent memory accesses, not
// the launch, bounds a round.  The host picks one level below this width.
constexpr uint64_t WF_WIDE_ROUND = 16384;

// A released successor: its entry (the task's first line, then (successor
// side 0, successor side 1, task, pass)), whether it became ready and may
// run in this block.
struct WfNext {
    uint4 a = make_uint4(0u, 0u, 0u, 0u), b = make_uint4(0u, 0u, 0u, 0u);
    bool ready = false, local = false;
};

// Release the successor `nx` of a visit in pass `pass`: the pending count's
// old value, and the successor's entry loaded.
__device__ __forceinline__ uint32_t wf_release_one(const WfDev& W, uint32_t T, uint32_t npass, uint32_t pass,
                                                   uint32_t nx, uint32_t tag, WfNext& s) {
    if (nx == WF_NONE) return 0u;
    uint32_t st = nx, sp = pass;
    if (nx & WF_WRAP) {
        if (pass + 1 >= npass) return 0u;
        st = nx & ~WF_WRAP;
        sp = pass + 1;
    }
    const uint32_t old = atomicAdd(&W.pend[(size_t)sp * T + st], (tag << 2) - 1u);
    s.a = W.task4[2 * (size_t)st];
    const uint4 nxt = W.task4[2 * (size_t)st + 1];
    s.b = make_uint4(nxt.x, nxt.y, st, sp);
    s.ready = true;
    return old;
}

// Whether the successor released with pending count `old` is ready, and
// whether it may run in this block.
__device__ __forceinline__ void wf_settle(uint32_t old, uint32_t tag, WfNext& s) {
    if (!s.ready) return;
    if ((old & 3u) != 1u) {   // another predecessor is still to run
        s.ready = false;
        return;
    }
    const uint32_t other = old >> 2;
    s.local = (other >> 12) != (tag >> 12) || other == tag;
}

// Release the two successors of the visit whose entry is (ea, eb) and load
// the visit's inputs: every round trip of the visit in flight at once.
// (Releasing before the visit runs is safe: a successor made ready here runs
// only at this block's next level, after a barrier, so after this visit's
// writes, or in a later launch, never beside this visit.)
__device__ __forceinline__ void wf_start_visit(const WfDev& W, uint32_t T, uint32_t npass, uint4 ea, uint4 eb,
                                               uint32_t tag, WfNext& s0, WfNext& s1, uint32_t& o0, uint32_t& o1,
                                               WfLoaded& L) {
    // the whole entry first (an empty asm that takes it): otherwise the
    // compiler, which loads its fields separately, waits for every load in
    // flight, the first successor's included, before the second release
    asm volatile("" : : "v"(ea.x), "v"(ea.y), "v"(ea.z), "v"(ea.w), "v"(eb.x), "v"(eb.y), "v"(eb.z), "v"(eb.w));
    o0 = wf_release_one(W, T, npass, eb.w, eb.x, tag, s0);
    o1 = wf_release_one(W, T, npass, eb.w, eb.y, tag, s1);
    wf_visit_load(W, ea.z, ea.w, ea.x, L);
}

// This wave's queue slots for its ready successors: local ones (by visit id)
// go into the block's list for the next level (when `may_local` and it has
// room), the rest are counted for launch r + 1 with one atomic per wave,
// whose result wf_publish reads after the visit, so its round trip overlaps
// the visit's arithmetic.
struct WfSlots {
    bool g0 = false, g1 = false;
    uint64_t b0 = 0, b1 = 0;
    uint32_t slot0;   // lane 0's only (no other lane assigns it: no copy joins the branches)
};

__device__ __forceinline__ WfSlots wf_reserve(const WfDev& W, uint32_t T, uint32_t r, const WfNext& s0,
                                              const WfNext& s1, bool may_local, uint32_t* lq, uint32_t* lqn) {
    auto keep_local = [&](const WfNext& sx) {
        if (!(sx.ready && sx.local && may_local)) return false;
        const uint32_t at = atomicAdd(lqn, 1u);
        if (at >= (uint32_t)WF_LCAP) return false;
        lq[at] = sx.b.w * T + sx.b.z;
        return true;
    };
    WfSlots q;
    q.g0 = s0.ready && !keep_local(s0);
    q.g1 = s1.ready && !keep_local(s1);
    q.b0 = __ballot(q.g0);
    q.b1 = __ballot(q.g1);
    const uint32_t tot = (uint32_t)(__popcll(q.b0) + __popcll(q.b1));
    // The count's index through a register the compiler cannot see into:
    // for a uniform address it would rewrite the atomic to broadcast its
    // result at once, which waits for the round trip here.
    uint32_t zero;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
    if (tot != 0 && lane_id() == 0) q.slot0 = atomicAdd(&W.rcnt[r + 1 + zero], tot);
    return q;
}

// Write the entries of the reserved successors into `outq`.  An index past
// T would break the one-ready-visit-per-task bound; such a visit is dropped
// and the host's completion check reports the stall.
__device__ __forceinline__ void wf_publish(uint4* __restrict__ outq, uint32_t T, const WfSlots& q, const WfNext& s0,
                                           const WfNext& s1) {
    if ((q.b0 | q.b1) == 0) return;
    const uint32_t slot0 = __shfl(q.slot0, 0);
    const uint64_t below = lane_id() ? (~0ull >> (64 - lane_id())) : 0ull;
    const uint32_t at0 = slot0 + (uint32_t)__popcll(q.b0 & below);
    const uint32_t at1 = slot0 + (uint32_t)__popcll(q.b0) + (uint32_t)__popcll(q.b1 & below);
    if (q.g0 && at0 < T) {
        outq[2 * (size_t)at0] = s0.a;
        outq[2 * (size_t)at0 + 1] = s0.b;
    }
    if (q.g1 && at1 < T) {
        outq[2 * (size_t)at1] = s1.a;
        outq[2 * (size_t)at1 + 1] = s1.b;
    }
}

// One visit per lane (active lanes only) and its successors queued: all 64
// lanes call this together.
__device__ __forceinline__ void wf_step(const WfDev& W, uint4* __restrict__ outq, uint32_t T, uint32_t npass,
                                        uint32_t r, uint32_t tag, bool act, uint4 ea, uint4 eb, bool may_local,
                                        uint32_t* lq, uint32_t* lqn, uint32_t& ran) {
    WfNext s0, s1;
    WfLoaded L;
    uint32_t o0 = 0, o1 = 0;
    if (act) wf_start_visit(W, T, npass, ea, eb, tag, s0, s1, o0, o1, L);
    wf_settle(o0, tag, s0);
    wf_settle(o1, tag, s1);
    // Every load of the visit has landed before the queue's atomic is issued
    // (the visit needs them all next anyway).  Memory results return in
    // issue order, so a wait for any load issued before the atomic would
    // otherwise be compiled as a wait for the atomic too, and the visit would
    // no longer overlap it.  (s_waitcnt vmcnt(0), other counters untouched.)
    __builtin_amdgcn_s_waitcnt(0x0F70);
    const WfSlots q = wf_reserve(W, T, r, s0, s1, may_local, lq, lqn);
    if (act) {
        wf_visit(W, L, eb.w, ea.z, ea.w, ea.x, ea.x + ea.y);
        ++ran;
    }
    wf_publish(outq, T, q, s0, s1);
}

// Launch r: runs the visits queued for it, then (levels > 1) the successors
// that became ready in this block.  The first entry is loaded before the
// count it is checked against, so the two loads are in flight together (its
// index is clamped into the half).
__global__ void __launch_bounds__(WF_THREADS) k_wf_round(WfDev W, uint32_t T, uint32_t npass, uint32_t r, int levels) {
    __shared__ uint32_t lq[2][WF_LCAP];
    __shared__ uint32_t lqn[2];
    const uint4* __restrict__ in = W.front + (size_t)(r & 1u) * 2u * T;
    uint4* __restrict__ outq = W.front + (size_t)((r + 1u) & 1u) * 2u * T;
    const uint32_t ub = blockIdx.x * blockDim.x;
    const uint32_t us = min(ub + threadIdx.x, T - 1u);
    uint4 ea = in[2 * (size_t)us], eb = in[2 * (size_t)us + 1];
    const uint32_t cnt = W.rcnt[r];
    // (an empty asm that takes the entry: the compiler would otherwise sink
    // its loads below the count's test)
    asm volatile("" : : "v"(ea.x), "v"(ea.y), "v"(ea.z), "v"(ea.w), "v"(eb.x), "v"(eb.y), "v"(eb.z), "v"(eb.w));
    if (ub >= cnt) return;
    const uint32_t tag = ((r + 1u) << 12) | blockIdx.x;
    if (threadIdx.x == 0) { lqn[0] = 0u; lqn[1] = 0u; }
    __syncthreads();
    uint32_t ran = 0;
    // level 0: this block's share of the visits queued for launch r
    for (uint32_t u0 = ub; u0 < cnt; u0 += gridDim.x * blockDim.x) {
        const uint32_t u = u0 + threadIdx.x;
        if (u0 != ub && u < cnt) {
            ea = in[2 * (size_t)u];
            eb = in[2 * (size_t)u + 1];
        }
        wf_step(W, outq, T, npass, r, tag, u < cnt, ea, eb, levels > 1, lq[0], &lqn[0], ran);
    }
    // levels 1 ..: the visits this block made ready, from its own list
    for (int lev = 1; lev < levels; ++lev) {
        const int cur = (lev - 1) & 1, nxl = lev & 1;
        __syncthreads();   // level lev - 1's writes and list are complete
        const uint32_t n = min(lqn[cur], (uint32_t)WF_LCAP);
        if (n == 0) break;
        if (threadIdx.x == 0) lqn[nxl] = 0u;   // its last reader was level lev - 1
        __syncthreads();
        for (uint32_t u0 = 0; u0 < n; u0 += blockDim.x) {
            const uint32_t u = u0 + threadIdx.x;
            uint4 la = make_uint4(0u, 0u, 0u, 0u), lb = make_uint4(0u, 0u, 0u, 0u);
            if (u < n) {
                const uint32_t v = lq[cur][u];
                const uint32_t pass = v / T, t = v - pass * T;
                la = W.task4[2 * (size_t)t];
                const uint4 nx = W.task4[2 * (size_t)t + 1];
                lb = make_uint4(nx.x, nx.y, t, pass);
            }
            wf_step(W, outq, T, npass, r, tag, u < n, la, lb, lev + 1 < levels, lq[nxl], &lqn[nxl], ran);
        }
    }
    // visits run: one atomic per wave
    uint32_t wsum = ran;
    for (int off = 32; off > 0; off >>= 1) wsum += __shfl_xor(wsum, off);
    if (lane_id() == 0 && wsum) atomicAdd(W.done, wsum);
}

// ---------------------------------------------------------------------------
// Boundaries: src/world.cpp:698-702 (wrap) and 635-657 (reflective walls)
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(TPB) k_wrap(BodiesDev B, SimParams P) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    if (!(B.im[i] > 0.0)) return;
    st3(B.x0, B.x1, B.x2, i, wrap_pos(box_of(B, P, i), ld3(B.x0, B.x1, B.x2, i)));
}

__global__ void __launch_bounds__(TPB) k_walls(BodiesDev B, SimParams P) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    if (B.im[i] <= 0.0) return;
    Geom g = load_geom(B, i);
    V3 x = g.x;
    V3 v = ld3(B.v0, B.v1, B.v2, i);
    for (int axis = 0; axis < 3; ++axis) {
        const double h = P.box.half[axis];
        for (int sgn = -1; sgn <= 1; sgn += 2) {
            V3 outward{axis == 0 ? (double)sgn : 0.0, axis == 1 ? (double)sgn : 0.0, axis == 2 ? (double)sgn : 0.0};
            g.x = x;
            V3 sp = support_world(g, outward);
            double coord = comp(sp, axis) * sgn;
            double pen = coord - h;
            if (pen > 0.0) {
                double* px = axis == 0 ? &x.x : (axis == 1 ? &x.y : &x.z);
                double* pv = axis == 0 ? &v.x : (axis == 1 ? &v.y : &v.z);
                *px -= (double)sgn * pen;
                if (*pv * sgn > 0.0) *pv = -P.restitution * *pv;
            }
        }
    }
    st3(B.x0, B.x1, B.x2, i, x);
    st3(B.v0, B.v1, B.v2, i, v);
}

// Constraint count = cons_off[P] (P read on the device), clamped to capacity.
__global__ void k_cons_count(const uint32_t* __restrict__ cons_off, uint32_t ccap, uint32_t* counters) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t r = cons_off[counters[CNT_P]];
    counters[CNT_CRAW] = r;
    if (r > ccap) { counters[CNT_FAULT] |= NE_FAULT_CONTACTS; r = ccap; }
    counters[CNT_C] = r;
    counters[CNT_ISL] = 0u;
    counters[CNT_NBIG] = 0u;
}

// This step's (key, slot)-ordered normal impulses become the next step's warm
// cache (src/world.cpp:619-626).
__global__ void __launch_bounds__(TPB) k_copy_cache(const uint32_t* __restrict__ counters_in, const uint64_t* __restrict__ key,
                                                    const double* __restrict__ nimp, uint64_t* __restrict__ prev_key,
                                                    double* __restrict__ prev_imp, uint32_t* __restrict__ prev_n) {
    const uint32_t nc = counters_in[CNT_C];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
        prev_key[k] = key[k];
        prev_imp[k] = nimp[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *prev_n = nc;
}

// FP probe: the device's sqrt / divide / floor on host-given operands, for the
// bit-equality checks the whole parity story rests on.
__global__ void k_fp_probe(int op, uint32_t n, const double* a, const double* b, double* out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double r;
    if (op == 0) r = sqrt(a[t]);
    else if (op == 1) r = a[t] / b[t];
    else if (op == 2) r = floor(a[t]);
    else if (op == 4) r = atan2_rn(a[t], b[t]);   // the hinge limit's, rounded to nearest
    else if (op == 5) r = atan2(a[t], b[t]);      // the device library's
    else r = a[t] * b[t] + a[t];
    out[t] = r;
}

}  // namespace

// ---------------------------------------------------------------------------
// Island sleeping (src/world.cpp:317-371, wake in collide 527-531).
//
// The reference walks the pairs in (i, j) order; a sleeping body is woken by
// the first pair with contacts whose other body is awake at that point,
// dynamic or kinematic, and faster than wakeVel (after the pair's mass filter).
// Each sleeper's waking pair is the smallest such index, found by rounds of
// atomic minima: every round only lowers an estimate that stays at or above
// the true index, and a round that finds nothing new is the fixed point.  A
// woken body keeps the velocity it slept with -- zero when the step put it to
// sleep -- so a second round is needed only when a body set asleep with
// velocity above wakeVel is woken, and it may then wake others further on.
// Each pair then sees its bodies as the reference does at its position --
// asleep before the body's waking pair, awake from it on -- which decides
// whether the pair is skipped and whether it joins the sleep islands.
// ---------------------------------------------------------------------------

constexpr int WAKE_ROUNDS = 8;   // rounds after the first; a longer wake chain is reported

__device__ __forceinline__ bool sleeping_now(const BodiesDev& B, uint32_t i) { return B.flags[i] & NE_FLAG_SLEEPING; }

// Inverse mass of a body at a point of the pair walk: its current value while
// asleep, the stored one once woken during this step.
__device__ __forceinline__ double walk_mass(const BodiesDev& B, uint32_t k, bool slept, bool asleep) {
    return (slept && !asleep) ? B.ims[k] : B.im[k];
}

__device__ __forceinline__ bool wakes_others(const BodiesDev& B, const SimParams& P, uint32_t k) {
    return (B.flags[k] & (NE_FLAG_DYNAMIC | NE_FLAG_KINEMATIC)) && norm(ld3(B.v0, B.v1, B.v2, k)) > P.wake_vel;
}

__global__ void __launch_bounds__(TPB) k_sleep_reset(BodiesDev B, uint32_t* __restrict__ wake_at, uint32_t* __restrict__ again) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < B.n) wake_at[i] = 0xFFFFFFFFu;
    if (i <= (uint32_t)WAKE_ROUNDS) again[i] = 0u;
}

__global__ void __launch_bounds__(TPB) k_sleep_wake_mark(BodiesDev B, SimParams P, const uint32_t* __restrict__ counters,
                                                         const uint32_t* __restrict__ cand_i,
                                                         const uint32_t* __restrict__ cand_j,
                                                         const uint32_t* __restrict__ ncont, uint32_t* wake_at,
                                                         uint32_t* again, int round) {
    if (round > 0 && !again[round - 1]) return;
    const uint32_t np = counters[CNT_P];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        if (!ncont[p]) continue;
        const uint32_t i = cand_i[p], j = cand_j[p];
        for (int side = 0; side < 2; ++side) {
            const uint32_t sl = side ? j : i, o = side ? i : j;   // sleeper, other
            if (!sleeping_now(B, sl) || wake_at[sl] <= p) continue;
            const bool o_slept = sleeping_now(B, o);
            const bool o_asleep = o_slept && !(wake_at[o] < p);
            if (o_asleep || !wakes_others(B, P, o)) continue;
            if (B.im[sl] + walk_mass(B, o, o_slept, false) <= 0.0) continue;   // the pair's mass filter
            const uint32_t old = atomicMin(&wake_at[sl], p);
            if (old > p && wakes_others(B, P, sl)) again[round] = 1u;
        }
    }
}

// Decides each pair as the reference's walk does at its position; drops the
// skipped ones (ncont = 0) and flags dynamic-dynamic pairs for the islands.
__global__ void __launch_bounds__(TPB) k_sleep_fate(BodiesDev B, uint32_t* __restrict__ counters,
                                                    const uint32_t* __restrict__ cand_i, const uint32_t* __restrict__ cand_j,
                                                    uint32_t* __restrict__ ncont, const uint32_t* __restrict__ wake_at,
                                                    const uint32_t* __restrict__ again, uint8_t* __restrict__ isp) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && again[WAKE_ROUNDS]) atomicOr(&counters[CNT_FAULT], NE_FAULT_WAKE);
    const uint32_t np = counters[CNT_P];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        isp[p] = 0;
        if (!ncont[p]) continue;
        const uint32_t i = cand_i[p], j = cand_j[p];
        const bool si = sleeping_now(B, i), sj = sleeping_now(B, j);
        const bool si_pre = si && !(wake_at[i] < p), sj_pre = sj && !(wake_at[j] < p);
        if (walk_mass(B, i, si, si_pre) + walk_mass(B, j, sj, sj_pre) <= 0.0) { ncont[p] = 0; continue; }
        const bool si_post = si_pre && wake_at[i] != p, sj_post = sj_pre && wake_at[j] != p;
        if (si_post && sj_post) { ncont[p] = 0; continue; }
        isp[p] = (walk_mass(B, i, si, si_post) > 0.0 && walk_mass(B, j, sj, sj_post) > 0.0) ? 1 : 0;
    }
}

__global__ void __launch_bounds__(TPB) k_sleep_apply_wake(BodiesDev B, const uint32_t* __restrict__ wake_at) {   // world.cpp:317-323
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n || wake_at[i] == 0xFFFFFFFFu) return;
    B.im[i] = B.ims[i];
    B.ii0[i] = B.iis0[i]; B.ii1[i] = B.iis1[i]; B.ii2[i] = B.iis2[i];
    B.flags[i] &= ~NE_FLAG_SLEEPING;
    B.st[i] = 0.0;
}

__device__ __forceinline__ bool awake_dyn(const BodiesDev& B, uint32_t i) {
    return (B.flags[i] & NE_FLAG_DYNAMIC) && !(B.flags[i] & NE_FLAG_SLEEPING) && B.im[i] > 0.0;
}

__device__ __forceinline__ void sleep_union(const BodiesDev& B, uint32_t* parent, uint32_t i, uint32_t j) {
    if (!awake_dyn(B, i) || !awake_dyn(B, j)) return;
    uint32_t a = i, b = j;
    for (int guard = 0; guard < 1 << 20; ++guard) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) break;
        if (a < b) { uint32_t t = a; a = b; b = t; }
        if (atomicCAS(&parent[a], a, b) == a) break;
    }
}

// Doubles as unsigned keys that order as the values (negative ones reversed),
// so an island's smallest timer is an atomic minimum.
__device__ __forceinline__ unsigned long long timer_key(double t) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(t);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}
__device__ __forceinline__ double timer_of_key(unsigned long long k) {
    return __longlong_as_double((long long)((k >> 63) ? (k & ~(1ull << 63)) : ~k));
}

__global__ void __launch_bounds__(TPB) k_sleep_init(BodiesDev B, uint32_t* __restrict__ parent,
                                                    unsigned long long* __restrict__ mint) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    parent[i] = i;
    mint[i] = ~0ull;
}

__global__ void __launch_bounds__(TPB) k_sleep_join_pairs(BodiesDev B, const uint32_t* __restrict__ counters,
                                                          const uint32_t* __restrict__ cand_i,
                                                          const uint32_t* __restrict__ cand_j,
                                                          const uint8_t* __restrict__ isp, uint32_t* __restrict__ parent) {
    const uint32_t np = counters[CNT_P];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x)
        if (isp[p]) sleep_union(B, parent, cand_i[p], cand_j[p]);
}

__global__ void __launch_bounds__(TPB) k_sleep_join_joints(BodiesDev B, const uint32_t* __restrict__ ja,
                                                           const uint32_t* __restrict__ jb, uint32_t nj,
                                                           uint32_t* __restrict__ parent) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nj) sleep_union(B, parent, ja[k], jb[k]);
}

// Slow-timers (world.cpp:346-351) and each island's smallest timer.
__global__ void __launch_bounds__(TPB) k_sleep_timers(BodiesDev B, SimParams P, uint32_t* __restrict__ parent,
                                                      unsigned long long* __restrict__ mint) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = i < B.n && awake_dyn(B, i);
    uint32_t root = 0;
    unsigned long long key = 0;
    if (on) {
        double t = B.st[i];
        if (norm(ld3(B.v0, B.v1, B.v2, i)) < P.sleep_lin && norm(ld3(B.w0, B.w1, B.w2, i)) < P.sleep_ang) t = t + P.dt;
        else t = 0.0;
        B.st[i] = t;
        root = uf_find(parent, i);
        key = timer_key(t);
    }
    wave_min_u64(mint, root, key, on);
}

// Islands whose slowest member has been slow for sleepTime go to sleep
// (world.cpp:360-370).
__global__ void __launch_bounds__(TPB) k_sleep_apply(BodiesDev B, SimParams P, uint32_t* __restrict__ parent,
                                                     const unsigned long long* __restrict__ mint) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n || !awake_dyn(B, i)) return;
    const double m = timer_of_key(mint[uf_find(parent, i)]);
    if (m < P.sleep_time) return;
    B.ims[i] = B.im[i];
    B.iis0[i] = B.ii0[i]; B.iis1[i] = B.ii1[i]; B.iis2[i] = B.ii2[i];
    B.im[i] = 0.0;
    B.ii0[i] = 0.0; B.ii1[i] = 0.0; B.ii2[i] = 0.0;
    st3(B.v0, B.v1, B.v2, i, V3{0.0, 0.0, 0.0});
    st3(B.w0, B.w1, B.w2, i, V3{0.0, 0.0, 0.0});
    B.flags[i] |= NE_FLAG_SLEEPING;
}

// ---------------------------------------------------------------------------
// Distance joints (src/world.cpp:150-197).  Joints are grouped on the host
// into components: joints that share a body that can move are in one
// component, swept by one thread in the joints' own order, so the
// reference's sequential order is kept where it matters and components run
// in parallel.
// ---------------------------------------------------------------------------

struct JointDev {
    const uint32_t *a, *b;
    const double *la0, *la1, *la2, *lb0, *lb1, *lb2, *rest, *stiff, *damp;
    const uint32_t *comp_start, *comp_list;
    uint32_t ncomp;
};

struct JointGeom {
    V3 nrm, rA, rB;
    double dist;
};

// Anchor geometry of joint k; false when the anchors coincide (dist < 1e-9).
__device__ __forceinline__ bool joint_geom(const BodiesDev& B, const SimParams& P, const JointDev& J, uint32_t k,
                                           uint32_t a, uint32_t b, JointGeom& g) {
    const Box3& box = box_of(B, P, a);
    const V3 xa = ld3(B.x0, B.x1, B.x2, a), xb = ld3(B.x0, B.x1, B.x2, b);
    const V3 wa = xa + qrot(ldq(B, a), V3{J.la0[k], J.la1[k], J.la2[k]});
    const V3 wb = xb + qrot(ldq(B, b), V3{J.lb0[k], J.lb1[k], J.lb2[k]});
    const V3 d = min_image(box, wa - wb);
    g.dist = norm(d);
    if (g.dist < 1e-9) return false;
    g.nrm = d / g.dist;
    g.rA = wa - xa;
    g.rB = min_image(box, wb - xb);
    return true;
}

// Spring-damper joints: force and torque added before integrate
// (world.cpp:150-170), in joint order within a component.
__global__ void __launch_bounds__(64) k_joint_springs(BodiesDev B, SimParams P, JointDev J) {
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < J.ncomp; c += gridDim.x * blockDim.x) {
        for (uint32_t t = J.comp_start[c]; t < J.comp_start[c + 1]; ++t) {
            const uint32_t k = J.comp_list[t];
            if (J.stiff[k] <= 0.0) continue;
            const uint32_t a = J.a[k], b = J.b[k];
            JointGeom g;
            if (!joint_geom(B, P, J, k, a, b, g)) continue;
            const V3 va = ld3(B.v0, B.v1, B.v2, a), wa = ld3(B.w0, B.w1, B.w2, a);
            const V3 vb = ld3(B.v0, B.v1, B.v2, b), wb = ld3(B.w0, B.w1, B.w2, b);
            const V3 vrel = (va + cross(wa, g.rA)) - (vb + cross(wb, g.rB));
            const double vn = dot(vrel, g.nrm);
            const double f = -J.stiff[k] * (g.dist - J.rest[k]) - J.damp[k] * vn;
            const V3 F = g.nrm * f;
            // a body that cannot move ignores force and torque (integrate skips it)
            if (mutable_body(B, a)) {
                st3(B.f0, B.f1, B.f2, a, ld3(B.f0, B.f1, B.f2, a) + F);
                st3(B.t0, B.t1, B.t2, a, ld3(B.t0, B.t1, B.t2, a) + cross(g.rA, F));
            }
            if (mutable_body(B, b)) {
                st3(B.f0, B.f1, B.f2, b, ld3(B.f0, B.f1, B.f2, b) - F);
                st3(B.t0, B.t1, B.t2, b, ld3(B.t0, B.t1, B.t2, b) - cross(g.rB, F));
            }
        }
    }
}

// Rigid joints: 8 sweeps of bilateral impulses with a Baumgarte length term
// after the contact solve (world.cpp:173-197).
__global__ void __launch_bounds__(64) k_joint_rigid(BodiesDev B, SimParams P, JointDev J) {
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < J.ncomp; c += gridDim.x * blockDim.x) {
        const uint32_t t0 = J.comp_start[c], t1 = J.comp_start[c + 1];
        for (int it = 0; it < 8; ++it)
            for (uint32_t t = t0; t < t1; ++t) {
                const uint32_t k = J.comp_list[t];
                if (J.stiff[k] > 0.0) continue;
                const uint32_t a = J.a[k], b = J.b[k];
                JointGeom g;
                if (!joint_geom(B, P, J, k, a, b, g)) continue;
                BodyView A = load_view(B, a), Bv = load_view(B, b);
                const double K = eff_mass(A.im, A.q, A.ii, Bv.im, Bv.q, Bv.ii, g.rA, g.rB, g.nrm);
                if (K <= 1e-18) continue;
                const V3 vrel = (A.v + cross(A.w, g.rA)) - (Bv.v + cross(Bv.w, g.rB));
                const double vn = dot(vrel, g.nrm);
                const double bias = 0.2 * (g.dist - J.rest[k]) / P.dt;
                const double dj = -(vn + bias) / K;
                apply_impulse(A, Bv, g.rA, g.rB, g.nrm * dj);
                store_view(B, a, A);
                store_view(B, b, Bv);
            }
    }
}

// ===========================================================================
// Adjoint of the step with the contact set of each step held fixed.
//
// The function differentiated is one step as the kernels above run it, except
// that the contacts gatherContacts returns (pairs, normals, points, depths)
// are constants of the step.  Everything downstream is differentiated:
// integrate (src/world.cpp:373-411), the constraint build (lever arms
// rA = minImage(point - x_a), restitution bias, effective mass, combined
// friction and restitution; world.cpp:535-553), the warm start carried from
// the previous step (559-568), the Gauss-Seidel iterations (572-612).  The
// positional correction (contact.hpp:78-85) and the periodic wrap add terms
// that do not depend on the inputs.
//
// Body adjoints use the field order of the body table (x0..2, v0..2, q0..3,
// w0..2), SoA over ncap.  Per-constraint accumulators (SoA over ccap):
constexpr int CA_NINIT = 0, CA_NIMP = 1, CA_TIMP = 2, CA_RA = 3, CA_RB = 6, CA_RBIAS = 9, CA_KN = 10, CA_MU = 11,
              CA_COUNT = 12;

struct AdjBodies {   // adjoints of one body table
    double* f;       // 13 fields x ncap
    uint32_t ncap;
    __device__ V3 get3(int f0, uint32_t b) const {
        return V3{f[(size_t)f0 * ncap + b], f[(size_t)(f0 + 1) * ncap + b], f[(size_t)(f0 + 2) * ncap + b]};
    }
    __device__ void put3(int f0, uint32_t b, V3 v) const {
        f[(size_t)f0 * ncap + b] = v.x; f[(size_t)(f0 + 1) * ncap + b] = v.y; f[(size_t)(f0 + 2) * ncap + b] = v.z;
    }
    __device__ Qt getq(uint32_t b) const {
        return Qt{f[(size_t)6 * ncap + b], f[(size_t)7 * ncap + b], f[(size_t)8 * ncap + b], f[(size_t)9 * ncap + b]};
    }
    __device__ void putq(uint32_t b, Qt q) const {
        f[(size_t)6 * ncap + b] = q.w; f[(size_t)7 * ncap + b] = q.x; f[(size_t)8 * ncap + b] = q.y;
        f[(size_t)9 * ncap + b] = q.z;
    }
};
constexpr int AF_X = 0, AF_V = 3, AF_W = 10;

__device__ __forceinline__ Qt qadd(Qt a, Qt b) { return Qt{a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z}; }

// One wave per island: replays the island's solve (warm start, iterations)
// recording each visit's inputs, then runs it backwards, then reverses the
// constraint build, each phase in sequence order through wave_in_order.
// Only bodies that can move carry adjoints (a body that cannot move belongs
// to no island and its velocity never changes).
__global__ void __launch_bounds__(64) k_adj_island(BodiesDev B, SimParams P, const uint32_t* __restrict__ counters,
                                                   const uint32_t* __restrict__ isl_start, uint32_t* __restrict__ isl_list,
                                                   ConsDev C, uint32_t ccap, double* __restrict__ tape,
                                                   uint32_t tstride, bool replay,
                                                   const double* __restrict__ pre, uint32_t ncap, AdjBodies A,
                                                   double* __restrict__ ca, const double* __restrict__ nfin,
                                                   double* __restrict__ nprev, double* __restrict__ cw) {
    __shared__ uint32_t sk[WAVE_SORT];
    const uint32_t nisl = counters[CNT_ISL];
    const int lane = threadIdx.x & 63;
    auto CA = [&](int f, uint32_t k) -> double& { return ca[(size_t)f * ccap + k]; };
    for (uint32_t isl = blockIdx.x; isl < nisl; isl += gridDim.x) {
        const uint32_t s = isl_start[isl], e = isl_start[isl + 1];
        const uint32_t m = e - s;
        uint32_t* list = isl_list + s;
        wave_sort_list(list, m, sk);
        const uint32_t iters = (uint32_t)P.iters;
        // the movable bodies of the visit to list[t]; a constraint between two
        // bodies that cannot move still orders its own visits (its impulses)
        auto cons_bodies = [&](uint32_t t, uint32_t& bi, uint32_t& bj) {
            const uint32_t k = list[t];
            bi = movable_or_none(B, C.ci[k]);
            bj = movable_or_none(B, C.cj[k]);
            if (bi == NO_BODY && bj == NO_BODY) bi = 0x80000000u | k;
        };
        // --- forward replay with tape (solve_island_seq on global memory)
        for (uint32_t t = lane; t < m; t += 64) {
            const uint32_t k = list[t];
            CA(CA_NINIT, k) = C.nimp[k];
            for (int f = 1; f < CA_COUNT; ++f) CA(f, k) = 0.0;
            CA(CA_NIMP, k) = nfin[k];
        }
        __threadfence_block();
        // (replay false: the recomputed step's solve wrote the tape, and the
        // reverse passes below read only the tape and the bodies' masses and
        // orientations)
        if (replay && P.warm)
            wave_in_order(
                m,
                [&](uint32_t t, uint32_t& bi, uint32_t& bj) {
                    if (C.nimp[list[t]] != 0.0) cons_bodies(t, bi, bj);
                },
                [&](uint32_t t) {
                    const uint32_t k = list[t];
                    if (C.nimp[k] == 0.0) return;
                    BodyView a = load_view(B, C.ci[k]), b = load_view(B, C.cj[k]);
                    warm_apply(a, b, load_geom(C, k), C.nimp[k]);
                    store_view(B, C.ci[k], a);
                    store_view(B, C.cj[k], b);
                });
        if (replay)
        wave_in_order(
            iters * m, [&](uint32_t v, uint32_t& bi, uint32_t& bj) { cons_bodies(v % m, bi, bj); },
            [&](uint32_t v) {
                const uint32_t it = v / m, k = list[v % m];
                const uint32_t i = C.ci[k], j = C.cj[k];
                BodyView a = load_view(B, i), b = load_view(B, j);
                double* tp = tape + ((size_t)it * tstride + k) * TAPE_DOUBLES;
                tp[0] = a.v.x; tp[1] = a.v.y; tp[2] = a.v.z; tp[3] = a.w.x; tp[4] = a.w.y; tp[5] = a.w.z;
                tp[6] = b.v.x; tp[7] = b.v.y; tp[8] = b.v.z; tp[9] = b.w.x; tp[10] = b.w.y; tp[11] = b.w.z;
                tp[12] = C.nimp[k]; tp[13] = C.timp[k];
                double nimp = C.nimp[k], timp = C.timp[k];
                gs_visit(a, b, load_geom(C, k), C.rbias[k], C.kn[k], C.mu[k], nimp, timp);
                C.nimp[k] = nimp;
                C.timp[k] = timp;
                store_view(B, i, a);
                store_view(B, j, b);
            });
        // --- reverse of the iterations
        const uint32_t nv = iters * m;
        wave_in_order(
            nv, [&](uint32_t v, uint32_t& bi, uint32_t& bj) { cons_bodies((nv - 1 - v) % m, bi, bj); },
            [&](uint32_t v) {
                const uint32_t r = nv - 1 - v;
                const uint32_t it = r / m, k = list[r % m];
                const uint32_t i = C.ci[k], j = C.cj[k];
                BodyView a = load_view(B, i), b = load_view(B, j);
                const double* tp = tape + ((size_t)it * tstride + k) * TAPE_DOUBLES;
                a.v = V3{tp[0], tp[1], tp[2]}; a.w = V3{tp[3], tp[4], tp[5]};
                b.v = V3{tp[6], tp[7], tp[8]}; b.w = V3{tp[9], tp[10], tp[11]};
                const bool ma = a.mv || a.mw, mb = b.mv || b.mw;
                V3 avb = ma ? A.get3(AF_V, i) : V3{0, 0, 0}, awb = ma ? A.get3(AF_W, i) : V3{0, 0, 0};
                V3 bvb = mb ? A.get3(AF_V, j) : V3{0, 0, 0}, bwb = mb ? A.get3(AF_W, j) : V3{0, 0, 0};
                Qt qab{0, 0, 0, 0}, qbb{0, 0, 0, 0};
                ConsAdj cj{V3{CA(CA_RA, k), CA(CA_RA + 1, k), CA(CA_RA + 2, k)},
                           V3{CA(CA_RB, k), CA(CA_RB + 1, k), CA(CA_RB + 2, k)}, CA(CA_RBIAS, k), CA(CA_KN, k),
                           CA(CA_MU, k)};
                double nb = CA(CA_NIMP, k), tb = CA(CA_TIMP, k);
                gs_visit_rev(a, b, load_geom(C, k), C.rbias[k], C.kn[k], C.mu[k], tp[12], tp[13], avb, awb, bvb, bwb,
                             nb, tb, cj, qab, qbb);
                CA(CA_NIMP, k) = nb; CA(CA_TIMP, k) = tb;
                CA(CA_RA, k) = cj.rA.x; CA(CA_RA + 1, k) = cj.rA.y; CA(CA_RA + 2, k) = cj.rA.z;
                CA(CA_RB, k) = cj.rB.x; CA(CA_RB + 1, k) = cj.rB.y; CA(CA_RB + 2, k) = cj.rB.z;
                CA(CA_RBIAS, k) = cj.rbias; CA(CA_KN, k) = cj.kn; CA(CA_MU, k) = cj.mu;
                if (ma) { A.put3(AF_V, i, avb); A.put3(AF_W, i, awb); A.putq(i, qadd(A.getq(i), qab)); }
                if (mb) { A.put3(AF_V, j, bvb); A.put3(AF_W, j, bwb); A.putq(j, qadd(A.getq(j), qbb)); }
            });
        // --- reverse of the warm start: every warm impulse sees the adjoint of
        // the velocities after the whole warm start, so this pass ends before
        // the constraint build adds its terms to the body adjoints
        if (P.warm)
            wave_in_order(
                m, [&](uint32_t v, uint32_t& bi, uint32_t& bj) { cons_bodies(m - 1 - v, bi, bj); },
                [&](uint32_t v) {
                    const uint32_t k = list[m - 1 - v];
                    const uint32_t i = C.ci[k], j = C.cj[k];
                    const BodyView a = load_view(B, i), b = load_view(B, j);
                    const bool ma = a.mv || a.mw, mb = b.mv || b.mw;
                    const ConsGeom g = load_geom(C, k);
                    V3 rAb{CA(CA_RA, k), CA(CA_RA + 1, k), CA(CA_RA + 2, k)};
                    V3 rBb{CA(CA_RB, k), CA(CA_RB + 1, k), CA(CA_RB + 2, k)};
                    Qt qab{0, 0, 0, 0}, qbb{0, 0, 0, 0};
                    const V3 avb = ma ? A.get3(AF_V, i) : V3{0, 0, 0}, awb = ma ? A.get3(AF_W, i) : V3{0, 0, 0};
                    const V3 bvb = mb ? A.get3(AF_V, j) : V3{0, 0, 0}, bwb = mb ? A.get3(AF_W, j) : V3{0, 0, 0};
                    V3 Jb{0, 0, 0};
                    apply_impulse_rev(a, b, g.rA, g.rB, g.n * CA(CA_NINIT, k), avb, awb, bvb, bwb, Jb, qab, qbb, rAb,
                                      rBb);
                    const double ninitb = CA(CA_NIMP, k) + dot(Jb, g.n);
                    const uint32_t src = C.src[k];
                    if (src != 0xFFFFFFFFu) nprev[src] = ninitb;
                    CA(CA_RA, k) = rAb.x; CA(CA_RA + 1, k) = rAb.y; CA(CA_RA + 2, k) = rAb.z;
                    CA(CA_RB, k) = rBb.x; CA(CA_RB + 1, k) = rBb.y; CA(CA_RB + 2, k) = rBb.z;
                    if (ma) A.putq(i, qadd(A.getq(i), qab));
                    if (mb) A.putq(j, qadd(A.getq(j), qbb));
                });
        // --- reverse of the constraint build
        wave_in_order(
            m, [&](uint32_t v, uint32_t& bi, uint32_t& bj) { cons_bodies(m - 1 - v, bi, bj); },
            [&](uint32_t v) {
                const uint32_t k = list[m - 1 - v];
                const uint32_t i = C.ci[k], j = C.cj[k];
                const BodyView a = load_view(B, i), b = load_view(B, j);
                const bool ma = a.mv || a.mw, mb = b.mv || b.mw;
                const ConsGeom g = load_geom(C, k);
                V3 rAb{CA(CA_RA, k), CA(CA_RA + 1, k), CA(CA_RA + 2, k)};
                V3 rBb{CA(CA_RB, k), CA(CA_RB + 1, k), CA(CA_RB + 2, k)};
                Qt qab{0, 0, 0, 0}, qbb{0, 0, 0, 0};
                V3 avb = ma ? A.get3(AF_V, i) : V3{0, 0, 0}, awb = ma ? A.get3(AF_W, i) : V3{0, 0, 0};
                V3 bvb = mb ? A.get3(AF_V, j) : V3{0, 0, 0}, bwb = mb ? A.get3(AF_W, j) : V3{0, 0, 0};
                // restitution bias from the pre-solve velocities (world.cpp:40, 544)
                const V3 va{pre[(size_t)0 * ncap + i], pre[(size_t)1 * ncap + i], pre[(size_t)2 * ncap + i]};
                const V3 wa{pre[(size_t)3 * ncap + i], pre[(size_t)4 * ncap + i], pre[(size_t)5 * ncap + i]};
                const V3 vb{pre[(size_t)0 * ncap + j], pre[(size_t)1 * ncap + j], pre[(size_t)2 * ncap + j]};
                const V3 wb{pre[(size_t)3 * ncap + j], pre[(size_t)4 * ncap + j], pre[(size_t)5 * ncap + j]};
                const V3 vc = (va + cross(wa, g.rA)) - (vb + cross(wb, g.rB));
                const double vn = dot(vc, g.n);
                const WorldDev* wd = B.W ? &B.W[B.wid[i]] : nullptr;
                const double wrest = wd ? wd->restitution : P.restitution;
                const double wfric = wd ? wd->friction : P.friction;
                const double ra = B.rest[i] < 0 ? wrest : B.rest[i], rb = B.rest[j] < 0 ? wrest : B.rest[j];
                const double fa = B.fric[i] < 0 ? wfric : B.fric[i], fb = B.fric[j] < 0 ? wfric : B.fric[j];
                const double ecomb = combine_mat(P.rcomb, ra, rb);
                double eb = 0.0;
                if (vn < -0.5) {
                    const double rbb = CA(CA_RBIAS, k);
                    eb = -vn * rbb;
                    const V3 vcb = g.n * (-ecomb * rbb);
                    avb = avb + vcb;
                    bvb = bvb - vcb;
                    cross_rev(wa, g.rA, vcb, awb, rAb);
                    cross_rev(wb, g.rB, -vcb, bwb, rBb);
                }
                V3 nb{0, 0, 0};
                eff_mass_rev(a.q, a.ii, b.q, b.ii, g.rA, g.rB, g.n, CA(CA_KN, k), qab, qbb, rAb, rBb, nb);
                cw[k] = CA(CA_MU, k) * combine_dw(P.fcomb, fa, fb, B.fric[i] < 0, B.fric[j] < 0);
                cw[(size_t)ccap + k] = eb * combine_dw(P.rcomb, ra, rb, B.rest[i] < 0, B.rest[j] < 0);
                // rA = minImage(point - x_a): the minimum image is a constant shift
                if (ma) {
                    A.put3(AF_V, i, avb); A.put3(AF_W, i, awb); A.putq(i, qadd(A.getq(i), qab));
                    A.put3(AF_X, i, A.get3(AF_X, i) - rAb);
                }
                if (mb) {
                    A.put3(AF_V, j, bvb); A.put3(AF_W, j, bwb); A.putq(j, qadd(A.getq(j), qbb));
                    A.put3(AF_X, j, A.get3(AF_X, j) - rBb);
                }
            });
        __syncthreads();   // the next island's sort reuses the LDS
    }
}

// Reverse of integrate (world.cpp:373-411) per body: adjoints at the end of
// integrate -> adjoints at the start of the step.  `st` is the state at the
// start of the step (same field order).
__global__ void __launch_bounds__(TPB) k_adj_integrate(BodiesDev B, SimParams P, const double* __restrict__ st,
                                                       uint32_t ncap, AdjBodies A) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    auto S3 = [&](int f0) { return V3{st[(size_t)f0 * ncap + i], st[(size_t)(f0 + 1) * ncap + i], st[(size_t)(f0 + 2) * ncap + i]}; };
    const Qt q0{st[(size_t)6 * ncap + i], st[(size_t)7 * ncap + i], st[(size_t)8 * ncap + i], st[(size_t)9 * ncap + i]};
    const V3 w0 = S3(10);
    const double dt = P.dt;
    V3 xb = A.get3(AF_X, i), vb = A.get3(AF_V, i), wb = A.get3(AF_W, i);
    Qt qb = A.getq(i);
    const uint32_t fl = B.flags[i];
    const double im = B.im[i];
    if (fl & NE_FLAG_KINEMATIC) {
        // x1 = x0 + v dt, q1 = integrate_orientation(q0, w, dt); v and w unchanged
        vb = vb + xb * dt;
        Qt q0b{0, 0, 0, 0};
        integrate_orientation_rev(q0, w0, dt, qb, q0b, wb);
        qb = q0b;
    } else if (!(im <= 0.0)) {
        const V3 invI = ld3(B.ii0, B.ii1, B.ii2, i);
        const double ld = B.ldamp[i], ad = B.adamp[i];
        if (ld > 0.0) vb = vb * (1.0 / (1.0 + ld * dt));
        if (ad > 0.0) wb = wb * (1.0 / (1.0 + ad * dt));
        const Qt q1 = integrate_orientation(q0, w0, dt);
        V3 L1 = angular_momentum(q0, invI, w0);   // + torque * dt: a constant
        const V3 tq = ld3(B.t0, B.t1, B.t2, i);
        L1 = L1 + tq * dt;
        // w1 = apply_inv_inertia(q1, invI, L1)
        V3 L1b{0, 0, 0};
        Qt q1b = qb;
        sandwich_rev(q1, invI, L1, wb, q1b, L1b);
        // q1 = integrate_orientation(q0, w0, dt);  L = angular_momentum(q0, invI, w0)
        Qt q0b{0, 0, 0, 0};
        V3 w0b{0, 0, 0};
        integrate_orientation_rev(q0, w0, dt, q1b, q0b, w0b);
        sandwich_rev(q0, inertia_of(invI), w0, L1b, q0b, w0b);
        // x1 = x0 + v1 dt, v1 = v0 + (g gs + f im) dt
        vb = vb + xb * dt;
        wb = w0b;
        qb = q0b;
    }
    A.put3(AF_V, i, vb);
    A.put3(AF_W, i, wb);
    A.putq(i, qb);
    (void)xb;
}

// Per-world sums of the constraints' d loss / d mu_world and d loss /
// d e_world (one block per world, a fixed-order reduction), added to gw.
__global__ void __launch_bounds__(256) k_adj_world_sum(BodiesDev B, const uint32_t* __restrict__ counters, ConsDev C,
                                                       const double* __restrict__ cw, uint32_t ccap,
                                                       double* __restrict__ gw, uint32_t nw) {
    __shared__ double sm[2][256];
    const uint32_t w = blockIdx.x;
    if (w >= nw) return;
    const uint32_t nc = counters[CNT_C];
    uint32_t k0 = 0, k1 = nc;
    if (B.W) {   // constraints are in (i, j) order and a world's bodies are contiguous
        const uint32_t b0 = B.W[w].body_begin, b1 = B.W[w].body_end;
        uint32_t lo = 0, hi = nc;
        while (lo < hi) { const uint32_t mid = (lo + hi) >> 1; if (C.ci[mid] < b0) lo = mid + 1; else hi = mid; }
        k0 = lo;
        hi = nc;
        while (lo < hi) { const uint32_t mid = (lo + hi) >> 1; if (C.ci[mid] < b1) lo = mid + 1; else hi = mid; }
        k1 = lo;
    }
    double a = 0.0, b = 0.0;
    for (uint32_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        a += cw[k];
        b += cw[(size_t)ccap + k];
    }
    sm[0][threadIdx.x] = a;
    sm[1][threadIdx.x] = b;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + st];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        gw[w] += sm[0][0];
        gw[nw + w] += sm[1][0];
    }
}

// n x 3 interleaved <-> three SoA fields
__global__ void __launch_bounds__(TPB) k_scatter3(const double* __restrict__ src, uint32_t n, double* __restrict__ f0,
                                                  double* __restrict__ f1, double* __restrict__ f2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f0[i] = src[3 * (size_t)i]; f1[i] = src[3 * (size_t)i + 1]; f2[i] = src[3 * (size_t)i + 2];
}
__global__ void __launch_bounds__(TPB) k_gather3(const double* __restrict__ f0, const double* __restrict__ f1,
                                                 const double* __restrict__ f2, uint32_t n, double* __restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[3 * (size_t)i] = f0[i]; dst[3 * (size_t)i + 1] = f1[i]; dst[3 * (size_t)i + 2] = f2[i];
}
__global__ void __launch_bounds__(TPB) k_fill_f64(double* __restrict__ p, size_t n, double v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ===========================================================================
// Host side
// ===========================================================================

// A mesh's bounding-volume tree as MeshData::build makes it (src/mesh.hpp:
// 38-43, buildNode 62-92): nodes in creation order (parent, then the left
// subtree, then the right); a node bounds its triangles' boxes; at most 4
// triangles make a leaf; otherwise the triangles are split at the middle
// index along the widest axis of the node box (x when strictly widest, else
// y when wider than z, else z) by std::nth_element on the sum of each
// triangle box's low and high coordinate on that axis -- the same library
// routine and comparison as the reference, so the halves, the leaves and the
// query's triangle order come out the same.
struct MeshTree {
    const double* v = nullptr;
    const int32_t* t = nullptr;
    std::vector<int32_t> order;
    std::vector<double> bounds;   // 6 per node: lo xyz, hi xyz
    std::vector<int32_t> links;   // 4 per node: left, right, start, count (relative)

    void tri_box(int32_t tri, double lo[3], double hi[3]) const {
        const double* a = v + 3 * (size_t)t[3 * (size_t)tri];
        const double* b = v + 3 * (size_t)t[3 * (size_t)tri + 1];
        const double* c = v + 3 * (size_t)t[3 * (size_t)tri + 2];
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min({a[k], b[k], c[k]});
            hi[k] = std::max({a[k], b[k], c[k]});
        }
    }

    int32_t build(int32_t start, int32_t count) {
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
        int32_t left = -1, right = -1, leaf_start = 0, leaf_count = 0;
        if (count <= 4) {
            leaf_start = start;
            leaf_count = count;
        } else {
            const double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
            const int axis = (ex > ey && ex > ez) ? 0 : (ey > ez ? 1 : 2);
            const int32_t mid = start + count / 2;
            std::nth_element(order.begin() + start, order.begin() + mid, order.begin() + start + count,
                             [&](int32_t A, int32_t Bt) {
                                 double al[3], ah[3], bl[3], bh[3];
                                 tri_box(A, al, ah);
                                 tri_box(Bt, bl, bh);
                                 return al[axis] + ah[axis] < bl[axis] + bh[axis];
                             });
            left = build(start, mid - start);
            right = build(mid, start + count - mid);
        }
        for (int k = 0; k < 3; ++k) {
            bounds[6 * (size_t)idx + k] = lo[k];
            bounds[6 * (size_t)idx + 3 + k] = hi[k];
        }
        int32_t* l = &links[4 * (size_t)idx];
        l[0] = left; l[1] = right; l[2] = leaf_start; l[3] = leaf_count;
        return idx;
    }
};

struct ne_gpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    ne_world_params params{};
    bool have_params = false;
    // capacities
    uint32_t n = 0, ncap = 0;
    uint32_t pcap = 0, ccap = 0, cellcap = 0;
    uint32_t user_pcap = 0, user_ccap = 0;
    double cell_scale = 1.0;
    // body storage
    double* dfield = nullptr;     // 38 double fields x ncap
    uint32_t* ufield = nullptr;   // shape, layer, mask, flags x ncap
    uint32_t* hull_off = nullptr;
    double* hull = nullptr;
    uint32_t* nc_off = nullptr;
    uint32_t* nc_idx = nullptr;
    BodiesDev B{};
    bool has_forces = false;
    bool spheres_only = false;
    // broadphase
    uint32_t* gridded = nullptr;
    uint32_t ngrid = 0;
    uint32_t* excl = nullptr;
    uint32_t nexcl = 0;
    double cell = 1.0;
    double grid_maxR = 0.0;
    double* aabb_part = nullptr;
    GridDev* grid = nullptr;
    uint32_t *cell_of = nullptr, *cell_count = nullptr, *cell_start = nullptr;
    uint32_t* sorted_idx = nullptr;
    uint32_t* order_prev = nullptr;   // previous step's cell order (a permutation of `gridded`)
    SortedBody* sorted = nullptr;
    SortedBody* xp = nullptr;         // packed post-integrate positions + bounding radius
    uint32_t *row_count = nullptr, *row_start = nullptr;
    uint32_t *ex_flags = nullptr, *ex_off = nullptr;
    uint32_t *cand_i = nullptr, *cand_j = nullptr;
    uint32_t* epa_q = nullptr;        // convex pairs queued for EPA (pcap entries)
    double* epa_tet = nullptr;        // GJK tetrahedra of overlapping convex pairs, by pair (24 doubles each)
    uint32_t* epa_long = nullptr;     // of those, the ones the first EPA launch left for the second (pcap entries)
    uint32_t* epa_long2 = nullptr;    // the ones the per-thread scratch tier left for the wave tier
    uint32_t* epa_long3 = nullptr;    // the ones the second LDS tier left for the scratch tier
    uint32_t *ncont = nullptr, *cons_off = nullptr;
    double* ct = nullptr;
    // constraints (current) + warm cache (previous)
    ConsDev C{};
    void* cons_mem = nullptr;
    uint64_t* prev_key = nullptr;
    double* prev_imp = nullptr;
    uint32_t* prev_n = nullptr;
    uint32_t prev_cap = 0;
    // islands
    uint32_t *parent = nullptr, *cons_root = nullptr, *used = nullptr, *root_id = nullptr;
    uint32_t *skey0 = nullptr, *sval0 = nullptr, *skey1 = nullptr, *sval1 = nullptr;
    uint32_t* isl_start = nullptr;
    EpaBig epa_big{};                    // pairs rerun with the large EPA workspace
    uint32_t* wake_at = nullptr;         // island sleeping: first waking pair per body
    uint32_t* wake_again = nullptr;      // island sleeping: per-round "a fast body woke" flags
    uint8_t* isp = nullptr;              // island sleeping: dynamic-dynamic pair flags (pcap)
    unsigned long long* mint = nullptr;  // island sleeping: smallest timer per island root
    // start-of-call snapshot (rerun after a capacity overflow)
    double* snap_body = nullptr;
    uint32_t* snap_flags = nullptr;
    uint8_t* nes1 = nullptr;   // snapshot records staged on the device
    uint64_t nes1_cap = 0;
    uint32_t snap_fcap = 0;
    size_t snap_nd = 0;
    uint64_t* snap_key = nullptr;
    double* snap_imp = nullptr;
    uint32_t* snap_n = nullptr;
    uint32_t snap_pcap = 0, snap_prev = 0;
    uint32_t *rs_hist = nullptr, *scan_part = nullptr;
    uint32_t *slot_of = nullptr, *row_scratch = nullptr, *claimed = nullptr, *cons_isl = nullptr;
    uint32_t *isl_count = nullptr, *isl_fill = nullptr, *isl_list = nullptr, *big_list = nullptr;
    StaticGrid sgrid{};
    int broadphase = NE_BROADPHASE_AUTO;
    uint32_t hint_p = 4096, hint_c = 4096;   // last call's pair/constraint counts (launch sizing)
    uint32_t cells_per_body = 2;
    int epa_faces_max = EPA_BIG_FACES_MAX;   // the large EPA workspace's face cap per run
    uint8_t* scratch_dev = nullptr;          // this device's address of `scratch`
    // facade events (PhysicsWorld::processContacts): two lists, this step's
    // and the last step's (ev_cur is the last completed step's), each of
    // pair keys (i << 32 | j), trigger flags, contacts (n, point, depth) and
    // the diff's marks; counts on the device (ev_n) and the host (h_ev_n)
    bool events_on = false;
    int ev_cur = 0;
    uint32_t h_ev_n[2] = {0, 0};
    uint32_t ev_cap = 0;
    uint64_t* ev_key[2] = {nullptr, nullptr};
    uint32_t* ev_trig[2] = {nullptr, nullptr};
    double* ev_dat[2] = {nullptr, nullptr};
    uint32_t* ev_state[2] = {nullptr, nullptr};
    uint32_t *ev_off = nullptr, *ev_n = nullptr, *ev_saved = nullptr;
    // general joints (World::joints): the records on the device, the step's
    // accumulators, and the components (ncomp + 1 starts, then joint lists)
    uint32_t ngj = 0, ngj_comp = 0;
    ne_joint* gj = nullptr;
    double *gj_applied = nullptr, *gj_motor = nullptr;
    uint32_t* gj_comp = nullptr;
    // continuous collision (World::ccdPass): the CCD bodies in index order,
    // their step-start positions, and the candidate pass's counts, offsets,
    // bodies and tile boxes
    uint32_t nccd = 0, ccd_cap = 0;
    uint32_t *ccd_idx = nullptr, *ccd_cnt = nullptr, *ccd_off = nullptr, *ccd_cand = nullptr;
    double *ccd_prev = nullptr, *ccd_tiles = nullptr;
    // scene queries: device arenas, grown on demand and carved per call (a: the
    // call's inputs and counts, b: its outputs
    // and w: EPA workspaces of overlapContacts)
    uint8_t *q_mem_a = nullptr, *q_mem_b = nullptr, *q_mem_w = nullptr;
    size_t q_cap_a = 0, q_cap_b = 0, q_cap_w = 0;
    GridDev grid_host{};
    uint32_t sgrid_cells = 0;
    // host copies from the last upload (grid plans are rebuilt from them)
    std::vector<double> h_R;          // bounding radii
    std::vector<uint32_t> h_excl;     // grid-excluded bodies, ascending
    std::vector<uint8_t> h_isexcl;
    bool grid_dirty = true;           // the static grid / world table must be rebuilt
    size_t ex_cap = 0;                // capacity of ex_flags / ex_off
    // batched worlds (nworlds == 0: one world)
    uint32_t nworlds = 0;
    std::vector<uint32_t> w_off;
    std::vector<double> w_half, w_rest, w_fric;
    uint32_t w_maxsize = 0;
    size_t ex_total = 0;
    uint32_t* d_wid = nullptr;
    WorldDev* d_worlds = nullptr;
    // HIP graphs of one step: [wrap_first][cell-order buffer parity]
    bool use_graph = false;
    hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    uint32_t graph_hint_p = 0, graph_hint_c = 0;
    int order_parity = 0;
    // distance joints
    uint32_t njoints = 0, nspring = 0, nrigid = 0, ncomp = 0;
    uint32_t *j_ab = nullptr, *j_comp = nullptr;   // a[n], b[n]; comp_start[ncomp+1], comp_list[n]
    double* j_f = nullptr;                         // la0..2, lb0..2, rest, stiff, damp (9 x n)
    std::vector<uint8_t> h_mut;                    // bodies that can move (im != 0 or ii != 0)
    // static triangle meshes (all meshes in shared arrays, absolute indices)
    uint32_t nmeshes = 0;
    double* m_verts = nullptr;
    int32_t *m_tris = nullptr, *m_order = nullptr, *m_root = nullptr, *m_of_body = nullptr;
    MeshNode* m_nodes = nullptr;
    std::vector<MeshTree> h_trees;                 // per mesh, relative indices (for ne_gpu_mesh_bvh)
    std::vector<int32_t> h_shape;                  // body shapes as uploaded
    std::vector<uint32_t> h_tris_n;
    // per step: triangle candidates of the mesh pairs
    uint32_t *mq_cnt = nullptr, *mq_off = nullptr, *mq_pair = nullptr;
    int32_t* mq_tri = nullptr;
    uint32_t* mq_long = nullptr;                   // candidates the first mesh launch left for the group tier
    uint32_t* mq_long2 = nullptr;                  // and the ones the group tier left for the one-wave tier
    uint32_t* mq_epa = nullptr;                    // candidates GJK found overlapping, for EPA tier A
    uint8_t* mhit = nullptr;
    double* mct = nullptr;                         // 7 doubles per candidate: normal, point, depth
    uint32_t mqcap = 0, mq_pcap = 0, hint_mq = 4096;
    // rollout checkpoints and adjoint work space
    struct Saved {
        double* state = nullptr;   // x, v, q, w fields (13 x ncap)
        uint64_t* key = nullptr;   // warm cache
        double* imp = nullptr;
        uint32_t* n = nullptr;
        uint32_t cap = 0;
    };
    std::vector<Saved> ckpt, seg;
    // The backward's copy of each recomputed step's front half (integrated
    // body fields, constraints, islands, counters), restored instead of
    // recomputed when the reverse sweep reaches that step.
    struct FrontSave {
        double* body = nullptr;     // 13 x ncap: the fields save_state keeps, after integrate
        double* cd = nullptr;       // the constraint list's 18 double fields, ccap each
        uint32_t* cu = nullptr;     // ci, cj, slot, src, ccap each
        uint64_t* key = nullptr;    // ccap
        uint32_t* isl = nullptr;    // isl_start (ccap + 1) then isl_list (ccap)
        uint32_t* cnt = nullptr;    // the step's counters
        uint32_t ccap = 0, ncap = 0;
        double* tape = nullptr;     // the solve's tape (iterations x tape_stride x TAPE_DOUBLES), or none
        uint32_t tape_stride = 0;
        size_t tape_n = 0;
        bool taped = false;         // this step's solve wrote the tape
    };
    std::vector<FrontSave> fronts;
    FrontSave* front_save = nullptr;                  // set: enqueue_step copies its front half here
    int32_t roll_steps = 0, roll_every = 0;
    double* adj = nullptr;        // 13 x ncap body adjoints
    double* adj_pre = nullptr;    // 6 x ncap pre-solve v, w
    double* adj_tape = nullptr;   // iterations x ccap x TAPE_DOUBLES
    double* adj_ca = nullptr;     // CA_COUNT x ccap
    double* adj_nfin = nullptr;   // ccap
    double* adj_nprev = nullptr;  // ccap
    double* adj_cw = nullptr;     // 2 x ccap
    double* adj_gw = nullptr;     // 2 x worlds
    double* stage = nullptr;      // n x 3 staging for caller buffers (host or device)
    uint32_t stage_n = 0;
    size_t adj_tape_n = 0;
    uint32_t adj_ccap = 0, adj_ncap = 0, adj_nw = 0;
    // large islands: the (pass, task) dataflow solve (eager steps; graphs keep k_solve_large)
    WfDev wf{};
    uint32_t wf_tcap = 0, wf_ccap = 0, wf_ncap = 0;   // capacities: tasks, constraints, bodies
    uint32_t wf_vcap_pass = 0;                        // passes the visit arrays hold per task
    uint32_t wf_rounds_hint = 0;                      // rounds the last solve took
    uint64_t wf_width_hint = 0;                       // visits per level in the last solve
    bool wf_expect = true;                            // expect large islands: read their task count back
    uint32_t wf_quiet = 0;                            // consecutive steps without large islands (checked ones)
    bool wf_seen = false;                             // this call had a step with large islands
    // The whole solve as one dataflow (no islands built, no per-island
    // solvers) while the large islands hold most constraints; eager steps only.
    bool solve_all = false;                           // this step
    bool all_ok = false;                              // the last island step's decision
    uint32_t all_steps = 0;                           // steps since that island step
    uint32_t* h_wf = nullptr;                         // pinned: values read back mid-step
    bool capturing = false;                           // a step is being captured into a graph
    hipStream_t side = nullptr;                       // second stream (default priority) for work beside the main one
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // counters
    uint32_t* counters = nullptr;
    uint32_t* h_counters = nullptr;
    int uf_bits = 1;
    // live kernel timing (HIP events on this stream)
    int timed_kernel = 0;
    std::vector<hipEvent_t> ev_start, ev_stop;
    uint32_t ev_used = 0;
    double timed_ms = 0.0;
    uint32_t timed_launches = 0;
};

// Scene queries (rays, swept spheres, overlaps), continuous collision and
// general joints: kernels over the body layout above.
#include "ne_query.hpp"
#include "ne_ccd.hpp"
#include "ne_joint.hpp"

namespace {

int set_err(ne_gpu_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_err(ctx, NE_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// Event bracket around one kernel launch when that kernel is being timed.
struct KTimer {
    ne_gpu_ctx* c;
    bool on;
    KTimer(ne_gpu_ctx* ctx, int kid) : c(ctx), on(ctx->timed_kernel == kid) {
        if (!on) return;
        if (c->ev_used == c->ev_start.size()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { on = false; return; }
            c->ev_start.push_back(a);
            c->ev_stop.push_back(b);
        }
        (void)hipEventRecord(c->ev_start[c->ev_used], c->stream);
    }
    ~KTimer() {
        if (!on) return;
        (void)hipEventRecord(c->ev_stop[c->ev_used], c->stream);
        ++c->ev_used;
    }
};

// Reads back the events recorded since the last harvest (after a sync).
void harvest_timing(ne_gpu_ctx* c) {
    for (uint32_t k = 0; k < c->ev_used; ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c->ev_start[k], c->ev_stop[k]) == hipSuccess) {
            c->timed_ms += ms;
            c->timed_launches += 1;
        }
    }
    c->ev_used = 0;
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

template <class T>
hipError_t dalloc(T*& p, size_t count) {
    dfree(p);
    if (count == 0) count = 1;
    return hipMalloc((void**)&p, sizeof(T) * count);
}

void drop_graphs(ne_gpu_ctx* c) {
    for (auto& row : c->gexec)
        for (hipGraphExec_t& g : row) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

enum DF : int {
    F_X0, F_X1, F_X2, F_V0, F_V1, F_V2, F_Q0, F_Q1, F_Q2, F_Q3, F_W0, F_W1, F_W2,
    F_F0, F_F1, F_F2, F_T0, F_T1, F_T2, F_IM, F_II0, F_II1, F_II2,
    F_RAD, F_HH, F_HE0, F_HE1, F_HE2, F_BR, F_FRIC, F_REST, F_LD, F_AD, F_GS,
    F_ST, F_IMS, F_IIS0, F_IIS1, F_IIS2, F_COUNT
};
enum UF : int { U_SHAPE, U_LAYER, U_MASK, U_FLAGS, U_COUNT };

// ---- Snapshot in the reference's "NES1" bytes (src/physics_world.cpp:22-64)
//
// After the magic and the body count, each body is one packed record in the
// order saveState writes it: x, q (w, x, y, z), v, w, invMass,
// invInertiaBody, invMassStore, invInertiaStore, sleepTimer (doubles, native
// little-endian), then sleeping as one byte.  The facade's joint count ends
// the blob.
constexpr uint32_t NES1_MAGIC = 0x4E455331u;
constexpr uint64_t NES1_HEAD = 8;       // magic, body count
constexpr uint64_t NES1_RECORD = 177;   // 22 doubles + 1 byte
constexpr int NES1_FIELDS[22] = {F_X0, F_X1,  F_X2,  F_Q0,  F_Q1,  F_Q2,  F_Q3, F_V0, F_V1, F_V2, F_W0,
                                 F_W1, F_W2,  F_IM,  F_II0, F_II1, F_II2, F_IMS, F_IIS0, F_IIS1, F_IIS2, F_ST};

// One thread per body: its record, at 177-byte stride (unaligned, so the
// doubles go through byte copies).
__global__ void k_nes1_pack(const double* __restrict__ F, uint32_t ncap, const uint32_t* __restrict__ flags,
                            uint32_t n, uint8_t* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint8_t* o = out + NES1_RECORD * i;
#pragma unroll
        for (int k = 0; k < 22; ++k) {
            const double v = F[(size_t)NES1_FIELDS[k] * ncap + i];
            __builtin_memcpy(o + 8 * k, &v, 8);
        }
        o[176] = (flags[i] & NE_FLAG_SLEEPING) ? 1 : 0;
    }
}

// Counts the bodies kept out of the grid as static that a snapshot would make
// movable (the grid's layout assumes they never move).
__global__ void k_nes1_check(const uint32_t* __restrict__ flags, uint32_t n, const uint8_t* __restrict__ in,
                             uint32_t* __restrict__ bad) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!(flags[i] & FLAG_GRID_EXCLUDED)) continue;
        const uint8_t* r = in + NES1_RECORD * i;
        double im, ii[3];
        __builtin_memcpy(&im, r + 8 * 13, 8);
        __builtin_memcpy(ii, r + 8 * 14, 24);
        if (im != 0.0 || ii[0] != 0.0 || ii[1] != 0.0 || ii[2] != 0.0 || r[176] != 0) atomicAdd(bad, 1u);
    }
}

// loadState's body loop (src/physics_world.cpp:46-53): every field of the
// record, and the sleeping flag.
__global__ void k_nes1_unpack(double* __restrict__ F, uint32_t ncap, uint32_t* __restrict__ flags, uint32_t n,
                              const uint8_t* __restrict__ in) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint8_t* r = in + NES1_RECORD * i;
#pragma unroll
        for (int k = 0; k < 22; ++k) {
            double v;
            __builtin_memcpy(&v, r + 8 * k, 8);
            F[(size_t)NES1_FIELDS[k] * ncap + i] = v;
        }
        flags[i] = r[176] ? (flags[i] | NE_FLAG_SLEEPING) : (flags[i] & ~NE_FLAG_SLEEPING);
    }
}

// ---------------------------------------------------------------------------
// Facade contact events (PhysicsWorld::processContacts,
// src/physics_world.cpp:253-302).  After a step, the overlapping pairs of the
// post-step state are found again with the step's own grid and narrow-phase
// kernels, run with the facade's filter (a sensor pair always, else not two
// immovable bodies; layers; no-collide lists) and detectContact's single
// contact per pair (no box-box manifold).  They are gathered in (i, j) order
// with their contact and a trigger flag (a sensor in the pair), and each
// list entry is marked by a binary search in the other step's list: a
// current pair also present last step (same kind) stays, else begins or
// enters; a last-step pair absent now ends or exits.
// ---------------------------------------------------------------------------

// The packed pose of every body for the grid fill, binned into the static
// periodic grid when BIN (the step bins inside integrate).
template <bool BIN>
__global__ void __launch_bounds__(TPB) k_ev_bin(BodiesDev B, StaticGrid G, SortedBody* __restrict__ xp,
                                                uint32_t* __restrict__ cell_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    const V3 x = ld3(B.x0, B.x1, B.x2, i);
    xp[i] = SortedBody{x.x, x.y, x.z, B.bR[i]};
    if (BIN && !(B.flags[i] & FLAG_GRID_EXCLUDED)) {
        uint32_t cell;
        if (B.W) {
            const WorldDev& w = B.W[B.wid[i]];
            cell = w.cell_base + grid_cell(w.grid, x);
        } else {
            cell = static_cell(G, x);
        }
        atomicAdd(&cell_count[cell], 1u);
    }
}

// The overlapping pairs, in candidate order (i, j), into list slots ev_off[p].
__global__ void __launch_bounds__(TPB) k_ev_gather(BodiesDev B, const uint32_t* __restrict__ counters,
                                                   const uint32_t* __restrict__ cand_i, const uint32_t* __restrict__ cand_j,
                                                   const uint32_t* __restrict__ ncont, const uint32_t* __restrict__ ev_off,
                                                   const double* __restrict__ ct, uint64_t* __restrict__ key,
                                                   uint32_t* __restrict__ trig, double* __restrict__ dat,
                                                   uint32_t* __restrict__ ev_n) {
    const uint32_t np = counters[CNT_P];
    if (blockIdx.x == 0 && threadIdx.x == 0) *ev_n = counters[CNT_EV];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        if (!ncont[p]) continue;
        const uint32_t k = ev_off[p], i = cand_i[p], j = cand_j[p];
        key[k] = ((uint64_t)i << 32) | (uint64_t)j;
        trig[k] = ((B.flags[i] | B.flags[j]) & NE_FLAG_SENSOR) ? 1u : 0u;
        const double* c = ct + (size_t)p * 28;
#pragma unroll
        for (int f = 0; f < 7; ++f) dat[7 * (size_t)k + f] = c[f];
    }
}

// Whether the (ascending, unique) key list holds `key` with trigger flag t.
__device__ __forceinline__ uint32_t ev_find(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ trig,
                                            uint32_t n, uint64_t key, uint32_t t) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && keys[lo] == key && trig[lo] == t) ? 1u : 0u;
}

// state_cur[k] = 1: the current pair was in the last step's list (stay);
// state_prev[k] = 1: the last step's pair is still there (no end).
__global__ void __launch_bounds__(TPB) k_ev_diff(const uint32_t* __restrict__ ev_n, int cur, int prev,
                                                 const uint64_t* __restrict__ key_c, const uint32_t* __restrict__ trig_c,
                                                 const uint64_t* __restrict__ key_p, const uint32_t* __restrict__ trig_p,
                                                 uint32_t* __restrict__ state_c, uint32_t* __restrict__ state_p) {
    const uint32_t nc = ev_n[cur], np = ev_n[prev];
    const uint32_t m = nc > np ? nc : np;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        if (k < nc) state_c[k] = ev_find(key_p, trig_p, np, key_c[k], trig_c[k]);
        if (k < np) state_p[k] = ev_find(key_c, trig_c, nc, key_p[k], trig_p[k]);
    }
}

// The step's counters back after the events pass, but for what the pass
// adds to: faults, EPA truncations and reruns, and its own pair count.
__global__ void k_ev_restore(uint32_t* __restrict__ counters, const uint32_t* __restrict__ saved) {
    const int s = (int)threadIdx.x;
    if (s < CNT_N && s != CNT_FAULT && s != CNT_EPATRUNC && s != CNT_EPARERUN && s != CNT_EV) counters[s] = saved[s];
}

// Staging of ne_gpu_save_state_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends.  The calls on one GPU share it and
// find the last call's bytes there; every byte a call reads back it has
// written first, so nothing needs clearing, but calls that overlap on one GPU
// (two threads, two contexts) race on it.  Host code never passes the array's
// name to a kernel or a copy (on the host it is a stand-in, not a device
// address): it uses the address looked up once per device by scratch_device().
constexpr uint64_t NOALLOC_SCRATCH_BYTES = 8ull << 20;   // 47,393 NES1 records
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_dev[SCRATCH_MAX_DEVICES] = {};

// The device address of `scratch` on `device`, which must be current; looked
// up on the first call for that device.  Returns nullptr if the lookup fails.
uint8_t* scratch_device(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_dev[device] = static_cast<uint8_t*>(ptr);
    }
    return g_scratch_dev[device];
}

double host_bounding_radius(int shape, double r, double hh, const double* he, const double* hull, uint32_t h0,
                            uint32_t h1) {
    auto nrm = [](double a, double b, double c) { return std::sqrt(a * a + b * b + c * c); };
    switch (shape) {
        case NE_SHAPE_SPHERE: return r;
        case NE_SHAPE_BOX: return nrm(he[0], he[1], he[2]);
        case NE_SHAPE_CAPSULE: return hh + r;
        case NE_SHAPE_PLANE: return 1e30;
        case NE_SHAPE_MESH: return 1e30;
        case NE_SHAPE_CONVEX: {
            double m = 0.0;
            for (uint32_t k = h0; k < h1; ++k) m = std::max(m, nrm(hull[3 * k], hull[3 * k + 1], hull[3 * k + 2]));
            return m;
        }
        default: return std::sqrt(hh * hh + r * r);   // cylinder
    }
}

SimParams make_sim_params(const ne_world_params& p) {
    SimParams s{};
    for (int a = 0; a < 3; ++a) {
        s.box.half[a] = p.half[a];
        s.box.edge[a] = 2.0 * p.half[a];
        s.box.inv_edge[a] = 1.0 / (2.0 * p.half[a]);
    }
    s.box.periodic = p.periodic ? 1 : 0;
    s.dt = p.dt;
    for (int a = 0; a < 3; ++a) s.g[a] = p.gravity[a];
    s.restitution = p.restitution;
    s.friction = p.friction;
    s.beta = p.contact_beta;
    s.rcomb = p.restitution_combine;
    s.fcomb = p.friction_combine;
    s.warm = p.warm_start ? 1 : 0;
    s.iters = p.solver_iterations;
    s.open_boundary = p.open_boundary ? 1 : 0;
    s.sleep = p.sleep_enabled ? 1 : 0;
    s.sleep_lin = p.sleep_lin_vel;
    s.sleep_ang = p.sleep_ang_vel;
    s.sleep_time = p.sleep_time;
    s.wake_vel = p.wake_vel;
    return s;
}

uint32_t blocks_for(uint64_t n, uint32_t tpb = TPB) {
    uint64_t b = (n + tpb - 1) / tpb;
    if (b < 1) b = 1;
    return (uint32_t)b;
}
// Grid for a grid-stride kernel over a device-side count: sized from the
// count the previous call saw (any count is correct; this only sets the
// parallelism), between 64 and 4096 blocks.
uint32_t hint_blocks(uint32_t hint) {
    uint64_t b = ((uint64_t)hint * 2 + TPB - 1) / TPB;
    if (b < 64) b = 64;
    if (b > 4096) b = 4096;
    return (uint32_t)b;
}

// One-wave blocks for the per-thread EPA kernels: they run one wave per SIMD
// and their waves end at very different times, so a SIMD whose wave ended
// takes new work at once instead of waiting for the rest of a larger block.
uint32_t wave_blocks(uint32_t hint) { return std::min<uint32_t>(4u * hint_blocks(hint), 16384u); }


uint32_t gs_blocks(uint64_t n) {   // grid-stride launches over device-side counts
    uint64_t b = (n + TPB - 1) / TPB;
    if (b < 1) b = 1;
    if (b > 4096) b = 4096;
    return (uint32_t)b;
}

int alloc_mesh_pairs(ne_gpu_ctx* c);

// Workgroups of the large-workspace EPA reruns: one wave per queued pair,
// up to the queue's capacity.
uint32_t big_blocks(const ne_gpu_ctx* c);

int alloc_pairs(ne_gpu_ctx* c, uint32_t pcap) {
    drop_graphs(c);
    c->pcap = pcap;
    HIP_TRY(c, dalloc(c->cand_i, pcap));
    HIP_TRY(c, dalloc(c->cand_j, pcap));
    HIP_TRY(c, dalloc(c->epa_q, pcap));
    HIP_TRY(c, dalloc(c->epa_tet, 24 * (size_t)pcap));
    HIP_TRY(c, dalloc(c->epa_long, pcap));
    HIP_TRY(c, dalloc(c->epa_long2, pcap));
    HIP_TRY(c, dalloc(c->epa_long3, pcap));
    HIP_TRY(c, dalloc(c->isp, pcap));
    HIP_TRY(c, dalloc(c->ncont, (size_t)pcap + 1));
    HIP_TRY(c, dalloc(c->cons_off, (size_t)pcap + 1));
    HIP_TRY(c, dalloc(c->ct, (size_t)pcap * 28));
    ScanPlan sp = scan_plan(std::max(pcap + 1, std::max(c->ncap * 4u, c->cellcap + 1)));
    (void)sp;
    if (c->nmeshes) return alloc_mesh_pairs(c);
    return NE_OK;
}

int alloc_cons(ne_gpu_ctx* c, uint32_t ccap) {
    drop_graphs(c);
    c->ccap = ccap;
    ConsDev& C = c->C;
    HIP_TRY(c, dalloc(C.ci, ccap));
    HIP_TRY(c, dalloc(C.cj, ccap));
    HIP_TRY(c, dalloc(C.slot, ccap));
    double** dp[] = {&C.n0, &C.n1, &C.n2, &C.p0, &C.p1, &C.p2, &C.ra0, &C.ra1, &C.ra2, &C.rb0, &C.rb1, &C.rb2,
                     &C.ov, &C.rbias, &C.mu, &C.nimp, &C.timp, &C.kn};
    for (double** p : dp) HIP_TRY(c, dalloc(*p, ccap));
    HIP_TRY(c, dalloc(C.key, ccap));
    HIP_TRY(c, dalloc(C.src, ccap));
    HIP_TRY(c, dalloc(c->cons_root, ccap));
    HIP_TRY(c, dalloc(c->cons_isl, ccap));
    HIP_TRY(c, dalloc(c->isl_list, ccap));
    HIP_TRY(c, dalloc(c->big_list, ccap));
    HIP_TRY(c, dalloc(c->isl_count, (size_t)ccap + 1));
    HIP_TRY(c, dalloc(c->isl_fill, (size_t)ccap + 1));
    HIP_TRY(c, dalloc(c->isl_start, (size_t)ccap + 1));
    return NE_OK;
}

int alloc_mesh_pairs(ne_gpu_ctx* c) {
    if (c->mq_pcap >= c->pcap && c->mq_cnt) return NE_OK;
    drop_graphs(c);
    HIP_TRY(c, dalloc(c->mq_cnt, (size_t)c->pcap + 1));
    HIP_TRY(c, dalloc(c->mq_off, (size_t)c->pcap + 1));
    c->mq_pcap = c->pcap;
    return NE_OK;
}

int alloc_mesh_cands(ne_gpu_ctx* c, uint32_t cap) {
    drop_graphs(c);
    c->mqcap = cap;
    HIP_TRY(c, dalloc(c->mq_pair, cap));
    HIP_TRY(c, dalloc(c->mq_tri, cap));
    HIP_TRY(c, dalloc(c->mq_long, cap));
    HIP_TRY(c, dalloc(c->mq_long2, cap));
    HIP_TRY(c, dalloc(c->mq_epa, cap));
    HIP_TRY(c, dalloc(c->mhit, cap));
    HIP_TRY(c, dalloc(c->mct, (size_t)cap * 7));
    return NE_OK;
}

void free_meshes(ne_gpu_ctx* c) {
    dfree(c->m_verts); dfree(c->m_tris); dfree(c->m_order); dfree(c->m_root); dfree(c->m_of_body); dfree(c->m_nodes);
    c->nmeshes = 0;
    c->h_trees.clear();
    c->h_tris_n.clear();
}

void free_mesh_work(ne_gpu_ctx* c) {
    dfree(c->mq_cnt); dfree(c->mq_off); dfree(c->mq_pair); dfree(c->mq_tri); dfree(c->mq_long); dfree(c->mq_long2); dfree(c->mq_epa); dfree(c->mhit);
    dfree(c->mct);
    c->mqcap = c->mq_pcap = 0;
}

MeshDev mesh_dev(const ne_gpu_ctx* c) {
    MeshDev M;
    M.verts = c->m_verts;
    M.tris = c->m_tris;
    M.order = c->m_order;
    M.nodes = c->m_nodes;
    M.root = c->m_root;
    M.of_body = c->m_of_body;
    return M;
}

int alloc_epa_big(ne_gpu_ctx* c, uint32_t cap, int faces_per) {
    drop_graphs(c);
    EpaBig& g = c->epa_big;
    g.cap = cap;
    g.faces_per = std::min(faces_per, c->epa_faces_max);
    g.faces_max = c->epa_faces_max;
    const size_t nf = (size_t)cap * (size_t)g.faces_per;
    HIP_TRY(c, dalloc(g.list, cap));
    HIP_TRY(c, dalloc(g.faces, nf));
    HIP_TRY(c, dalloc(g.occ, 3 * nf));
    return NE_OK;
}

int ensure_prev(ne_gpu_ctx* c, uint32_t cap) {
    if (c->prev_cap >= cap && c->prev_key) return NE_OK;
    drop_graphs(c);
    uint64_t* nk = nullptr;
    double* ni = nullptr;
    HIP_TRY(c, hipMalloc((void**)&nk, sizeof(uint64_t) * std::max(cap, 1u)));
    HIP_TRY(c, hipMalloc((void**)&ni, sizeof(double) * std::max(cap, 1u)));
    if (c->prev_key && c->prev_cap) {
        HIP_TRY(c, hipMemcpyAsync(nk, c->prev_key, sizeof(uint64_t) * c->prev_cap, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(ni, c->prev_imp, sizeof(double) * c->prev_cap, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    dfree(c->prev_key);
    dfree(c->prev_imp);
    c->prev_key = nk;
    c->prev_imp = ni;
    c->prev_cap = cap;
    return NE_OK;
}

// Everything but the warm cache and the counters (kept across re-uploads).
void free_work(ne_gpu_ctx* c) {
    dfree(c->dfield); dfree(c->ufield); dfree(c->hull_off); dfree(c->hull); dfree(c->nc_off); dfree(c->nc_idx);
    dfree(c->gridded); dfree(c->excl); dfree(c->aabb_part); dfree(c->grid);
    dfree(c->cell_of); dfree(c->cell_count); dfree(c->cell_start); dfree(c->sorted_idx); dfree(c->sorted);
    dfree(c->order_prev); dfree(c->xp);
    dfree(c->row_count); dfree(c->row_start); dfree(c->ex_flags); dfree(c->ex_off);
    dfree(c->cand_i); dfree(c->cand_j); dfree(c->epa_q); dfree(c->epa_tet); dfree(c->epa_long); dfree(c->epa_long2); dfree(c->epa_long3); dfree(c->isp);
    dfree(c->wake_at); dfree(c->mint); dfree(c->wake_again); dfree(c->ncont); dfree(c->cons_off); dfree(c->ct);
    ConsDev& C = c->C;
    dfree(C.ci); dfree(C.cj); dfree(C.slot);
    double** dp[] = {&C.n0, &C.n1, &C.n2, &C.p0, &C.p1, &C.p2, &C.ra0, &C.ra1, &C.ra2, &C.rb0, &C.rb1, &C.rb2,
                     &C.ov, &C.rbias, &C.mu, &C.nimp, &C.timp, &C.kn};
    for (double** p : dp) dfree(*p);
    dfree(C.key);
    dfree(C.src);
    dfree(c->parent); dfree(c->cons_root); dfree(c->used); dfree(c->root_id);
    dfree(c->skey0); dfree(c->sval0); dfree(c->skey1); dfree(c->sval1); dfree(c->isl_start);
    dfree(c->slot_of); dfree(c->row_scratch); dfree(c->claimed); dfree(c->cons_isl);
    dfree(c->isl_count); dfree(c->isl_fill); dfree(c->isl_list); dfree(c->big_list);
    dfree(c->rs_hist); dfree(c->scan_part);
    dfree(c->d_wid); dfree(c->d_worlds);
    dfree(c->j_ab); dfree(c->j_comp); dfree(c->j_f);
    c->njoints = c->nspring = c->nrigid = c->ncomp = 0;
    dfree(c->epa_big.list); dfree(c->epa_big.faces); dfree(c->epa_big.occ);
    c->epa_big.cap = 0;
    c->ex_cap = 0;
    drop_graphs(c);
}

void free_fronts(std::vector<ne_gpu_ctx::FrontSave>& v) {
    for (auto& F : v) {
        dfree(F.body); dfree(F.cd); dfree(F.cu); dfree(F.key); dfree(F.isl); dfree(F.cnt); dfree(F.tape);
    }
    v.clear();
}

void free_saved(std::vector<ne_gpu_ctx::Saved>& v) {
    for (auto& x : v) { dfree(x.state); dfree(x.key); dfree(x.imp); dfree(x.n); }
    v.clear();
}

void free_adjoint(ne_gpu_ctx* c) {
    free_saved(c->ckpt);
    free_saved(c->seg);
    free_fronts(c->fronts);
    dfree(c->adj); dfree(c->adj_pre); dfree(c->adj_tape); dfree(c->adj_ca); dfree(c->adj_nfin);
    dfree(c->adj_nprev); dfree(c->adj_cw); dfree(c->adj_gw); dfree(c->stage);
    c->stage_n = 0;
    c->adj_tape_n = 0;
    c->adj_ccap = c->adj_ncap = c->adj_nw = 0;
    c->roll_steps = 0;
}

void free_wf(ne_gpu_ctx* c) {
    WfDev& W = c->wf;
    dfree(W.flag); dfree(W.tid); dfree(W.task_k); dfree(W.pend0); dfree(W.pend); dfree(W.front);
    dfree(W.rcnt); dfree(W.done); dfree(W.inc_cnt); dfree(W.inc_start); dfree(W.inc_list);
    dfree(W.task_n); dfree(W.task4); dfree(W.rec); dfree(W.bp);
    c->wf_tcap = c->wf_ccap = c->wf_ncap = c->wf_vcap_pass = 0;
}

void free_all(ne_gpu_ctx* c) {
    free_wf(c);
    if (c->h_wf) { (void)hipHostFree(c->h_wf); c->h_wf = nullptr; }
    free_adjoint(c);
    free_work(c);
    free_meshes(c);
    free_mesh_work(c);
    dfree(c->prev_key); dfree(c->prev_imp); dfree(c->prev_n); dfree(c->counters);
    dfree(c->snap_body); dfree(c->snap_key); dfree(c->snap_imp); dfree(c->snap_n); dfree(c->snap_flags);
    dfree(c->nes1); c->nes1_cap = 0;
    for (int b = 0; b < 2; ++b) { dfree(c->ev_key[b]); dfree(c->ev_trig[b]); dfree(c->ev_dat[b]); dfree(c->ev_state[b]); }
    dfree(c->ev_off); dfree(c->ev_n); dfree(c->ev_saved);
    dfree(c->q_mem_a); dfree(c->q_mem_b); dfree(c->q_mem_w);
    dfree(c->ccd_idx); dfree(c->ccd_cnt); dfree(c->ccd_off); dfree(c->ccd_cand); dfree(c->ccd_prev); dfree(c->ccd_tiles);
    dfree(c->gj); dfree(c->gj_applied); dfree(c->gj_motor); dfree(c->gj_comp);
    c->ngj = c->ngj_comp = 0;
    c->nccd = c->ccd_cap = 0;
    c->q_cap_a = c->q_cap_b = c->q_cap_w = 0;
    c->ev_cap = 0;
    c->snap_nd = 0;
    c->snap_fcap = 0;
    c->snap_pcap = 0;
    c->prev_cap = 0;
    if (c->h_counters) { (void)hipHostFree(c->h_counters); c->h_counters = nullptr; }
}

// One reference step (src/world.cpp:133-146) as a kernel pipeline.
template <bool F, bool W>
void launch_integrate(ne_gpu_ctx* c, const SimParams& P, bool bin) {
    if (bin)
        k_integrate<F, W, true><<<blocks_for(c->n), TPB, 0, c->stream>>>(c->B, P, c->sgrid, c->xp, c->cell_count,
                                                                          c->parent, c->claimed);
    else
        k_integrate<F, W, false><<<blocks_for(c->n), TPB, 0, c->stream>>>(c->B, P, c->sgrid, c->xp, c->cell_count,
                                                                           c->parent, c->claimed);
}

// One reference step (src/world.cpp:133-146) as a kernel pipeline.  When
// `wrap_first`, the previous step's periodic wrap runs at the top of this
// step's integrate kernel; the caller wraps after the last step.
JointDev joint_dev(const ne_gpu_ctx* c) {
    const uint32_t n = c->njoints;
    const double* f = c->j_f;
    JointDev J;
    J.a = c->j_ab; J.b = c->j_ab + n;
    J.la0 = f; J.la1 = f + n; J.la2 = f + 2 * (size_t)n;
    J.lb0 = f + 3 * (size_t)n; J.lb1 = f + 4 * (size_t)n; J.lb2 = f + 5 * (size_t)n;
    J.rest = f + 6 * (size_t)n; J.stiff = f + 7 * (size_t)n; J.damp = f + 8 * (size_t)n;
    J.comp_start = c->j_comp; J.comp_list = c->j_comp + c->ncomp + 1;
    J.ncomp = c->ncomp;
    return J;
}

int ensure_side(ne_gpu_ctx* c) {
    if (c->side) return NE_OK;
    HIP_TRY(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIP_TRY(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIP_TRY(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    return NE_OK;
}

// This step solves every constraint as one dataflow: islands are neither
// linked nor built.  Captured steps keep the islands (k_solve_large).
bool whole_dataflow(const ne_gpu_ctx* c) { return c->solve_all && !c->capturing; }

// The grid, the candidate rows and the narrow phase of one step (or of the
// facade events pass, with its own filter and detectContact's one contact
// per pair), from the packed poses in c->xp and, with a static grid, the
// cell counts binned beside them.
void enqueue_detect(ne_gpu_ctx* c, const BodiesDev& B, const SimParams& P, bool static_grid) {
    hipStream_t s = c->stream;
    const uint32_t n = c->n;
    // --- grid
    if (c->ngrid > 0) {
        if (!static_grid) {
            k_aabb_partial<<<AABB_BLOCKS, TPB, 0, s>>>(B, c->gridded, c->ngrid, c->aabb_part);
            k_grid_setup<<<1, 1, 0, s>>>(P, c->cell, c->cellcap, c->aabb_part, AABB_BLOCKS, c->grid, c->counters);
            KTimer kt(c, NE_KERNEL_CELL_COUNT);
            k_cell_count<<<blocks_for(c->ngrid), TPB, 0, s>>>(c->gridded, c->ngrid, c->grid, c->xp, c->cell_count);
        }
        scan_u32(c->cell_count, c->cell_start, static_grid ? nullptr : c->counters + CNT_CELLS,
                 static_grid ? c->sgrid_cells : c->cellcap, c->scan_part, s, true);
        {
            KTimer kt(c, NE_KERNEL_CELL_FILL);
            std::swap(c->sorted_idx, c->order_prev);   // last step's cell order drives this fill
            c->order_parity ^= 1;
            k_cell_fill<<<blocks_for(c->ngrid), TPB, 0, s>>>(c->order_prev, c->ngrid, c->grid, B.wid, B.W, c->xp,
                                                             c->cell_start, c->sorted_idx, c->sorted);
        }
        // --- candidate rows: one neighbourhood walk per body
        KTimer kt(c, NE_KERNEL_CAND_COUNT);
        if (static_grid && c->nworlds == 0 && c->broadphase == NE_BROADPHASE_BRICK && c->sgrid.nc[0] >= BRICK_X + 2 && c->sgrid.nc[1] >= BRICK_Y + 2 &&
            c->sgrid.nc[2] >= BRICK_Z + 2) {
            const uint32_t nb = (uint32_t)((c->sgrid.nc[0] + BRICK_X - 1) / BRICK_X) *
                                (uint32_t)((c->sgrid.nc[1] + BRICK_Y - 1) / BRICK_Y) *
                                (uint32_t)((c->sgrid.nc[2] + BRICK_Z - 1) / BRICK_Z);
            const uint32_t blocks = ((nb + 7u) / 8u) * 8u;
            k_cand_brick<<<blocks, BRICK_THREADS, 0, s>>>(B, P, c->grid, c->sorted_idx, c->sorted, c->cell_start,
                                                          c->excl, c->nexcl, c->row_count, c->row_scratch, c->slot_of);
        } else if (c->broadphase == NE_BROADPHASE_WALK || c->broadphase == NE_BROADPHASE_AUTO) {
            k_cand_walk<<<blocks_for(c->ngrid), TPB, 0, s>>>(B, P, c->grid, c->ngrid, c->sorted_idx, c->sorted,
                                                             c->cell_start, c->excl, c->nexcl, c->row_count,
                                                             c->row_scratch, c->slot_of);
        } else {
            const uint32_t cells = static_grid ? c->sgrid_cells : c->cellcap;
            k_cand_cells<<<gs_blocks(cells), TPB, 0, s>>>(B, P, c->grid, c->sorted_idx, c->sorted, c->cell_start,
                                                          c->excl, c->nexcl, c->row_count, c->row_scratch, c->slot_of);
        }
    }
    const uint32_t ex_width = c->nworlds ? c->w_maxsize : n;   // flags per excluded row (at most)
    if (c->nexcl > 0) {
        dim3 g(blocks_for(ex_width), c->nexcl);
        k_excl_flags<<<g, TPB, 0, s>>>(B, P, c->excl, c->ex_flags);
        scan_u32(c->ex_flags, c->ex_off, nullptr, (uint32_t)c->ex_total, c->scan_part, s);
        k_excl_rowcount<<<blocks_for(c->nexcl), TPB, 0, s>>>(B, c->excl, c->nexcl, c->ex_off, c->row_count);
    }
    scan_u32(c->row_count, c->row_start, nullptr, n, c->scan_part, s, false,
             ScanTotal{c->counters, CNT_PRAW, CNT_P, CNT_FAULT, c->pcap, NE_FAULT_PAIRS, CNT_EPABIG, CNT_EPAQ});
    if (c->ngrid > 0) {
        KTimer kt(c, NE_KERNEL_CAND_WRITE);
        k_cand_emit<<<blocks_for(n), TPB, 0, s>>>(B, P, c->grid, c->slot_of, c->sorted_idx, c->sorted, c->cell_start,
                                                  c->excl, c->nexcl, c->row_count, c->row_scratch, c->row_start, c->pcap,
                                                  c->cand_i, c->cand_j);
    }
    if (c->nexcl > 0) {
        dim3 g(blocks_for(ex_width), c->nexcl);
        k_excl_write<<<g, TPB, 0, s>>>(B, c->excl, c->ex_flags, c->ex_off, c->row_start, c->pcap, c->cand_i, c->cand_j);
    }
    // --- narrow phase + constraints
    bool fork = false;
    {
        KTimer kt(c, NE_KERNEL_NARROW);
        if (c->spheres_only) {
            k_narrow<true><<<hint_blocks(c->hint_p), TPB, 0, s>>>(B, P, c->counters, c->cand_i, c->cand_j, c->ncont, c->ct,
                                                              c->epa_q, c->pcap, EpaTets{c->epa_tet});
        } else {
            const EpaTets tets{c->epa_tet};
            k_narrow<false><<<hint_blocks(c->hint_p), TPB, 0, s>>>(B, P, c->counters, c->cand_i, c->cand_j, c->ncont,
                                                               c->ct, c->epa_q, c->pcap, tets);
            // the EPA tiers start from k_narrow's GJK tetrahedra (tets, by pair)
            k_epa_lds<8, 12, 8><<<wave_blocks(c->hint_c), 64, 0, s>>>(B, P, c->counters, c->cand_i, c->cand_j,
                                                                      c->ncont, c->ct, c->epa_q, CNT_EPAQ, c->pcap,
                                                                      c->epa_long, CNT_EPALONG, tets);
            k_fill_u32<<<1, 64, 0, s>>>(c->counters + CNT_EPALONG3, 0u, nullptr, 1u);
            k_epa_group<4, 24, 40, false><<<wave_blocks(c->hint_c / 4), 64, 0, s>>>(
                B, P, c->counters, c->cand_i, c->cand_j, c->ncont, c->ct, c->epa_long, CNT_EPALONG, c->pcap,
                c->epa_long3, CNT_EPALONG3, MeshSrc{}, tets);
            k_fill_u32<<<1, 64, 0, s>>>(c->counters + CNT_EPALONG2, 0u, nullptr, 1u);
            k_epa_group<8, 36, 64, false><<<wave_blocks(c->hint_c / 8), 64, 0, s>>>(
                B, P, c->counters, c->cand_i, c->cand_j, c->ncont, c->ct, c->epa_long3, CNT_EPALONG3, c->pcap,
                c->epa_long2, CNT_EPALONG2, MeshSrc{}, tets);
            k_epa_wave<<<2048, 64, 0, s>>>(B, P, c->counters, c->cand_i, c->cand_j, c->ncont, c->ct, c->epa_long2,
                                           CNT_EPALONG2, c->pcap, c->epa_big);
            k_snap_counter<<<1, 64, 0, s>>>(c->counters, CNT_EPABIG, CNT_EPABIG_PAIRS);
            // The large-workspace reruns are few pairs on few waves: in eager
            // steps they run on the second stream beside the mesh kernels.
            // (If a stream call fails, the host waits for the second stream instead.)
            fork = !c->capturing && c->nmeshes && ensure_side(c) == NE_OK &&
                   hipEventRecord(c->ev_fork, s) == hipSuccess && hipStreamWaitEvent(c->side, c->ev_fork, 0) == hipSuccess;
            k_narrow_big<<<big_blocks(c), 64 * BIG_NW, 0, fork ? c->side : s>>>(B, P, c->counters, c->cand_i, c->cand_j, c->ncont,
                                                                       c->ct, c->epa_big);
            if (fork && hipEventRecord(c->ev_join, c->side) != hipSuccess) {
                (void)hipStreamSynchronize(c->side);
                fork = false;
            }
        }
        if (c->nmeshes) {   // meshContacts (world.cpp:56-83)
            const MeshDev M = mesh_dev(c);
            k_mesh_count<<<hint_blocks(c->hint_p), TPB, 0, s>>>(B, P, M, c->counters, c->cand_i, c->cand_j, c->mq_cnt);
            scan_u32(c->mq_cnt, c->mq_off, c->counters + CNT_P, c->pcap, c->scan_part, s, false,
                     ScanTotal{c->counters, CNT_MQRAW, CNT_MQ, CNT_FAULT, c->mqcap, NE_FAULT_MESH, -1, -1});
            k_mesh_emit<<<hint_blocks(c->hint_p), TPB, 0, s>>>(B, P, M, c->counters, c->cand_i, c->cand_j, c->mq_off,
                                                               c->mqcap, c->mq_pair, c->mq_tri);
            k_fill_u32<<<1, 64, 0, s>>>(c->counters + CNT_MQLONG, 0u, nullptr, 1u);
            k_fill_u32<<<1, 64, 0, s>>>(c->counters + CNT_MQEPA, 0u, nullptr, 1u);
            k_mesh_gjk<<<wave_blocks(c->hint_mq), 64, 0, s>>>(B, P, M, c->counters, c->cand_i, c->cand_j, c->mq_pair,
                                                               c->mq_tri, c->mhit, c->mct, c->mq_epa);
            k_mesh_lds<<<wave_blocks(c->hint_mq / 4), 64, 0, s>>>(B, P, M, c->counters, c->cand_i, c->cand_j, c->mq_pair,
                                                                   c->mq_tri, c->mhit, c->mct, c->mq_epa, c->mq_long);
            k_fill_u32<<<1, 64, 0, s>>>(c->counters + CNT_MQLONG2, 0u, nullptr, 1u);
            k_epa_group<8, 36, 64, true><<<wave_blocks(c->hint_mq / 16), 64, 0, s>>>(
                B, P, c->counters, c->cand_i, c->cand_j, c->ncont, c->ct, c->mq_long, CNT_MQLONG, c->mqcap,
                c->mq_long2, CNT_MQLONG2, MeshSrc{M, c->mq_pair, c->mq_tri, c->mhit, c->mct}, EpaTets{});
            k_mesh_wave<<<2048, 64, 0, s>>>(B, P, M, c->counters, c->cand_i, c->cand_j, c->mq_pair, c->mq_tri,
                                            c->mhit, c->mct, c->mq_long2, CNT_MQLONG2, c->epa_big);
            k_mesh_big<<<big_blocks(c), 64 * BIG_NW, 0, s>>>(B, P, M, c->counters, c->cand_i, c->cand_j, c->mq_pair, c->mq_tri, c->mhit,
                                         c->mct, c->epa_big);
            k_mesh_gather<<<hint_blocks(c->hint_p), TPB, 0, s>>>(B, c->counters, c->cand_i, c->cand_j, c->mq_off,
                                                                 c->mqcap, c->mhit, c->mct, c->ncont);
        }
        if (fork && hipStreamWaitEvent(s, c->ev_join, 0) != hipSuccess) (void)hipStreamSynchronize(c->side);
    }
}

// World::ccdPass (world.cpp:658-696) after integrate: the candidate pass
// over (CCD body, tile), its scan, and the one-thread fold that stops the
// bodies and moves their packed poses and static-grid cells.
void enqueue_ccd(ne_gpu_ctx* c, const BodiesDev& B, const SimParams& P, bool static_grid) {
    hipStream_t s = c->stream;
    const uint32_t ntiles = (c->n + RAY_THREADS - 1) / RAY_THREADS;
    const dim3 grid(ntiles, std::min<uint32_t>(c->nccd, 65535u));
    const bool cull = !P.box.periodic;
    if (cull) {
        k_query_tiles<true><<<ntiles, RAY_THREADS, 0, s>>>(B, c->ccd_tiles);
        k_ccd_cand<false, true><<<grid, RAY_THREADS, 0, s>>>(B, P.box, c->ccd_idx, c->nccd, c->ccd_prev, c->ccd_tiles,
                                                            c->ccd_cnt, nullptr, nullptr, 0);
    } else {
        k_ccd_cand<false, false><<<grid, RAY_THREADS, 0, s>>>(B, P.box, c->ccd_idx, c->nccd, c->ccd_prev, nullptr,
                                                             c->ccd_cnt, nullptr, nullptr, 0);
    }
    scan_u32(c->ccd_cnt, c->ccd_off, nullptr, c->nccd * ntiles, c->scan_part, s);
    if (cull)
        k_ccd_cand<true, true><<<grid, RAY_THREADS, 0, s>>>(B, P.box, c->ccd_idx, c->nccd, c->ccd_prev, c->ccd_tiles,
                                                           nullptr, c->ccd_off, c->ccd_cand, c->ccd_cap);
    else
        k_ccd_cand<true, false><<<grid, RAY_THREADS, 0, s>>>(B, P.box, c->ccd_idx, c->nccd, c->ccd_prev, nullptr,
                                                            nullptr, c->ccd_off, c->ccd_cand, c->ccd_cap);
    k_ccd_fold<<<1, 64, 0, s>>>(B, P.box, c->ccd_idx, c->nccd, c->ccd_prev, ntiles, c->ccd_off, c->ccd_cand,
                                c->ccd_cap, c->xp, static_grid ? 1 : 0, c->sgrid, c->cell_count);
}

// The step up to the solve: integrate, grid, candidate rows, narrow phase,
// constraint build and islands.
void enqueue_front(ne_gpu_ctx* c, bool wrap_first) {
    hipStream_t s = c->stream;
    const SimParams P = make_sim_params(c->params);
    BodiesDev B = c->B;
    const uint32_t n = c->n;
    const bool static_grid = P.box.periodic && c->ngrid > 0;
    if (c->nspring)   // spring joints add forces before integration (world.cpp:134)
        k_joint_springs<<<blocks_for(c->ncomp, 64), 64, 0, s>>>(B, P, joint_dev(c));
    if (c->nccd)   // step-start positions of the CCD bodies (world.cpp:135)
        k_ccd_save<<<blocks_for(c->nccd), TPB, 0, s>>>(B, c->ccd_idx, c->nccd, c->ccd_prev);
    {
        KTimer kt(c, NE_KERNEL_INTEGRATE);
        const bool w = wrap_first && P.box.periodic && c->njoints == 0 && c->ngj == 0 && !P.sleep && !c->events_on &&
                       !c->nccd;
        if (c->has_forces || c->nspring) {
            if (w) launch_integrate<true, true>(c, P, static_grid);
            else launch_integrate<true, false>(c, P, static_grid);
            c->has_forces = false;
        } else {
            if (w) launch_integrate<false, true>(c, P, static_grid);
            else launch_integrate<false, false>(c, P, static_grid);
        }
    }
    if (c->nccd) enqueue_ccd(c, B, P, static_grid);
    enqueue_detect(c, B, P, static_grid);
    if (P.sleep) {   // wakes in pair order (world.cpp:527-531), then the woken bodies' masses
        k_sleep_reset<<<blocks_for(n), TPB, 0, s>>>(B, c->wake_at, c->wake_again);
        for (int r = 0; r <= WAKE_ROUNDS; ++r)
            k_sleep_wake_mark<<<hint_blocks(c->hint_p), TPB, 0, s>>>(B, P, c->counters, c->cand_i, c->cand_j, c->ncont,
                                                                     c->wake_at, c->wake_again, r);
        k_sleep_fate<<<hint_blocks(c->hint_p), TPB, 0, s>>>(B, c->counters, c->cand_i, c->cand_j, c->ncont,
                                                            c->wake_at, c->wake_again, c->isp);
        k_sleep_apply_wake<<<blocks_for(n), TPB, 0, s>>>(B, c->wake_at);
    }
    scan_u32(c->ncont, c->cons_off, c->counters + CNT_P, c->pcap, c->scan_part, s, false,
             ScanTotal{c->counters, CNT_CRAW, CNT_C, CNT_FAULT, c->ccap, NE_FAULT_CONTACTS, CNT_ISL, CNT_NBIG});
    {
        KTimer kt(c, NE_KERNEL_CONS_BUILD);
        k_cons_build<<<hint_blocks(c->hint_p), TPB, 0, s>>>(B, P, c->counters, c->cand_i, c->cand_j, c->ncont, c->cons_off,
                                                        c->ct, c->C, c->ccap, c->prev_key, c->prev_imp, c->prev_n,
                                                        c->parent, c->claimed, c->nmeshes ? c->mq_off : nullptr,
                                                        c->nmeshes ? c->mct : nullptr, !whole_dataflow(c),
                                                        whole_dataflow(c) ? c->wf.rec : nullptr);
    }
    if (whole_dataflow(c)) return;   // no islands: every constraint goes to the dataflow solve
    // --- islands (linked inside k_cons_build)
    k_isl_claim<<<hint_blocks(c->hint_c), TPB, 0, s>>>(B, c->counters, c->C, c->parent, c->cons_root, c->claimed,
                                                   c->root_id, c->isl_count, c->isl_fill);
    k_isl_count<<<hint_blocks(c->hint_c), TPB, 0, s>>>(c->counters, c->cons_root, c->root_id, c->cons_isl, c->isl_count);
    scan_u32(c->isl_count, c->isl_start, c->counters + CNT_ISL, c->ccap, c->scan_part, s);
    k_isl_fill<<<hint_blocks(c->hint_c), TPB, 0, s>>>(c->counters, c->cons_isl, c->isl_start, c->isl_fill, c->isl_list);
}

// Work space of the large-island solve, grown to the constraint capacity, the
// body capacity and T tasks.
int ensure_wf(ne_gpu_ctx* c, uint32_t T, uint32_t npass) {
    WfDev& W = c->wf;
    if (!c->h_wf) HIP_TRY(c, hipHostMalloc((void**)&c->h_wf, sizeof(uint32_t) * 4, hipHostMallocDefault));
    if (!W.rcnt) {
        HIP_TRY(c, dalloc(W.rcnt, (size_t)WF_MAX_ROUNDS + 2));
        HIP_TRY(c, dalloc(W.done, 4));
    }
    if (c->wf_ccap < c->ccap) {
        HIP_TRY(c, dalloc(W.flag, c->ccap));
        HIP_TRY(c, dalloc(W.tid, (size_t)c->ccap + 1));
        HIP_TRY(c, dalloc(W.rec, (size_t)c->ccap * WF_REC));
        c->wf_ccap = c->ccap;
    }
    if (c->wf_ncap < c->ncap) {
        HIP_TRY(c, dalloc(W.inc_cnt, c->ncap));
        HIP_TRY(c, dalloc(W.inc_start, (size_t)c->ncap + 1));
        HIP_TRY(c, dalloc(W.bp, (size_t)c->ncap * WF_REC));
        c->wf_ncap = c->ncap;
    }
    if (T == 0) return NE_OK;
    if ((uint64_t)T * npass >= (1ull << 31)) return set_err(c, NE_ERR_UNSUPPORTED, "too many contact pairs in large islands");
    if (c->wf_tcap < T) {
        const uint32_t cap = T + T / 2 + 1024;
        HIP_TRY(c, dalloc(W.task_k, cap));
        HIP_TRY(c, dalloc(W.task_n, cap));
        HIP_TRY(c, dalloc(W.task4, 2 * (size_t)cap));
        HIP_TRY(c, dalloc(W.pend0, cap));
        HIP_TRY(c, dalloc(W.inc_list, 2 * (size_t)cap));
        HIP_TRY(c, dalloc(W.front, 4 * (size_t)cap));
        c->wf_tcap = cap;
        c->wf_vcap_pass = 0;
    }
    if (c->wf_vcap_pass < npass) {
        HIP_TRY(c, dalloc(W.pend, (size_t)c->wf_tcap * npass));
        c->wf_vcap_pass = npass;
    }
    return NE_OK;
}

int solve_rounds(ne_gpu_ctx* c, uint32_t T, uint32_t npass, uint32_t cleared);

// Islands of more than WF_ISLAND constraints (or every constraint): the warm
// start and the positional correction per body, the iterations as a
// dataflow of (iteration, task) visits.  Two synchronisations per step.
int solve_large_islands(ne_gpu_ctx* c, const SimParams& P) {
    hipStream_t s = c->stream;
    BodiesDev B = c->B;
    WfDev& W = c->wf;
    const bool all = whole_dataflow(c);
    if (!c->wf_expect && !all) {
        // recent steps had no large islands: do not stop the stream to ask;
        // if this one has some, the call reruns from its start reading them back
        k_wf_expect_none<<<1, 64, 0, s>>>(c->counters);
        return NE_OK;
    }
    int r0 = ensure_wf(c, 0, 1);
    if (r0 != NE_OK) return r0;
    k_fill_u32<<<1, 64, 0, s>>>(c->counters + CNT_WFC, 0u, nullptr, 1u);
    k_wf_mark<<<hint_blocks(c->hint_c), TPB, 0, s>>>(B, c->counters, c->C, all ? nullptr : c->cons_isl, c->isl_start,
                                                     W.flag, WF_ISLAND, (all && P.warm) ? c->prev_n : nullptr);
    scan_u32(W.flag, W.tid, c->counters + CNT_C, c->ccap, c->scan_part, s, false,
             ScanTotal{c->counters, CNT_WFT, CNT_WFT, CNT_FAULT, c->ccap, 0u, -1, -1});
    HIP_TRY(c, hipMemcpyAsync(c->h_wf, c->counters + CNT_WFT, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(c->h_wf + 2, c->counters + CNT_WFC, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(c->h_wf + 3, c->counters + CNT_C, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    const uint32_t T = c->h_wf[0];
    if (!all) {   // the next steps skip the islands while the large ones hold most constraints
        c->all_ok = (uint64_t)c->h_wf[2] * WF_ALL_DEN >= (uint64_t)c->h_wf[3] * WF_ALL_NUM && T > 0;
        c->all_steps = 0;
    }
    if (T == 0) {
        ++c->wf_quiet;   // step_call stops asking after 8 such steps, between calls
        return NE_OK;
    }
    c->wf_quiet = 0;
    c->wf_seen = true;
    const uint32_t npass = (uint32_t)std::max(P.iters, 0);   // the dataflow's passes: the iterations
    if (npass > WF_PASS_MAX) return set_err(c, NE_ERR_UNSUPPORTED, "too many solver iterations for the large-island solve");
    r0 = ensure_wf(c, T, npass);
    if (r0 != NE_OK) return r0;
    k_fill_u32<<<gs_blocks(c->n), TPB, 0, s>>>(W.inc_cnt, 0u, nullptr, c->n);
    k_fill_u32<<<gs_blocks(T), TPB, 0, s>>>(W.pend0, 0u, nullptr, T);
    uint32_t cleared = std::min<uint32_t>(WF_MAX_ROUNDS + 2, 2 * c->wf_rounds_hint + 256);
    k_fill_u32<<<gs_blocks(cleared), TPB, 0, s>>>(W.rcnt, 0u, nullptr, cleared);
    k_wf_tasks<<<hint_blocks(c->hint_c), TPB, 0, s>>>(B, c->counters, c->C, W, all);
    scan_u32(W.inc_cnt, W.inc_start, nullptr, c->n, c->scan_part, s, true);
    k_wf_fill<<<gs_blocks(T), TPB, 0, s>>>(B, T, c->C, W);
    k_wf_pack_bodies<<<blocks_for(c->n), TPB, 0, s>>>(B, W.bp);
    k_wf_link<<<gs_blocks(c->n), TPB, 0, s>>>(c->n, P, W);   // and the warm start
    const uint64_t total = (uint64_t)T * npass;
    if (total > 0) {
        r0 = solve_rounds(c, T, npass, cleared);
        if (r0 != NE_OK) return r0;
    }
    k_wf_correct<<<gs_blocks(c->n), TPB, 0, s>>>(B, P, W);
    k_wf_out<<<gs_blocks(T), TPB, 0, s>>>(T, P, c->C, W, c->prev_key, c->prev_imp);
    k_wf_unpack_bodies<<<blocks_for(c->n), TPB, 0, s>>>(B, W.bp);
    return NE_OK;
}

// The iterations: one launch per round, the host reading back the number of
// visits run after a batch of rounds (the schedule's depth is not known in
// advance).
int solve_rounds(ne_gpu_ctx* c, uint32_t T, uint32_t npass, uint32_t cleared) {
    hipStream_t s = c->stream;
    BodiesDev B = c->B;
    WfDev& W = c->wf;
    k_wf_init<<<gs_blocks(T), TPB, 0, s>>>(B, T, npass, c->C, W);
    k_fill_u32<<<1, 64, 0, s>>>(W.done, 0u, nullptr, 1u);
    const uint64_t total = (uint64_t)T * npass;
    const uint32_t hint = c->wf_rounds_hint ? c->wf_rounds_hint : 64u;
    uint64_t per_round = total / hint + 1;
    uint32_t grid = (uint32_t)std::min<uint64_t>(2048, std::max<uint64_t>(8, (per_round * 3 / 2 + WF_THREADS - 1) / WF_THREADS));
    // round width from the last solve's depth in single-level rounds
    const int levels = (c->wf_width_hint > WF_WIDE_ROUND) ? WF_LEVELS : 1;
    uint32_t r = 0, batch = hint + 16;
    for (;;) {
        if (r + batch + 2 >= WF_MAX_ROUNDS) return set_err(c, NE_ERR_UNSUPPORTED, "large-island solve needs too many rounds");
        if (r + batch + 2 > cleared) {   // round r fills the count of round r + 1
            k_fill_u32<<<gs_blocks(r + batch + 2 - cleared), TPB, 0, s>>>(W.rcnt + cleared, 0u, nullptr,
                                                                          r + batch + 2 - cleared);
            cleared = r + batch + 2;
        }
        for (uint32_t q = 0; q < batch; ++q, ++r) k_wf_round<<<grid, WF_THREADS, 0, s>>>(W, T, npass, r, levels);
        HIP_TRY(c, hipMemcpyAsync(c->h_wf, W.done, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipMemcpyAsync(c->h_wf + 1, W.rcnt + r, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipStreamSynchronize(s));
        if (c->h_wf[0] >= total) break;
        if (c->h_wf[1] == 0)   // nothing queued for the next launch, and nothing else is running
            return set_err(c, NE_ERR_INVALID, "large-island solve stalled (a visit never became ready)");
        batch = 64;
    }
    // the depth: rounds up to the last one that ran a visit
    std::vector<uint32_t> cnt(r);
    HIP_TRY(c, hipMemcpy(cnt.data(), W.rcnt, sizeof(uint32_t) * r, hipMemcpyDeviceToHost));
    uint32_t depth = r;
    while (depth > 0 && cnt[depth - 1] == 0) --depth;
    c->wf_rounds_hint = depth;
    // visits per level: the launches' depth times the levels each ran
    c->wf_width_hint = total / ((uint64_t)std::max<uint32_t>(depth, 1u) * (uint64_t)levels);
    return NE_OK;
}

uint32_t big_blocks(const ne_gpu_ctx* c) { return std::min<uint32_t>(std::max<uint32_t>(c->epa_big.cap, 16u), 1024u); }

int enqueue_solve(ne_gpu_ctx* c) {
    hipStream_t s = c->stream;
    const SimParams P = make_sim_params(c->params);
    BodiesDev B = c->B;
    KTimer kt(c, NE_KERNEL_SOLVE);
    // the adjoint's recompute of a step records the solve's tape
    const ne_gpu_ctx::FrontSave* F = c->front_save;
    const Tape tape = (F && F->taped) ? Tape{F->tape, F->tape_stride} : Tape{};
    c->wf.tape = tape;
    if (whole_dataflow(c)) return solve_large_islands(c, P);
    k_solve_single<<<hint_blocks(c->hint_c), TPB, 0, s>>>(B, P, c->counters, c->isl_start, c->isl_list, c->C,
                                                          c->prev_key, c->prev_imp, tape);
    k_solve_small<<<std::min<uint32_t>(4096u, std::max<uint32_t>(64u, (c->hint_c * 2 + SOLVE_THREADS - 1) / SOLVE_THREADS)),
                    SOLVE_THREADS, 0, s>>>(B, P, c->counters, c->isl_start, c->isl_list, c->C, c->big_list, c->prev_key,
                                           c->prev_imp, c->prev_n, c->capturing ? CAP_ISLAND : WF_ISLAND, tape);
    if (c->capturing) {   // a graph cannot read back mid-step: one wave sweeps each large island
        k_solve_large<<<std::min<uint32_t>(std::max<uint32_t>(c->hint_c / CAP_ISLAND, 64u), 8192u), 64, 0, s>>>(
            B, P, c->counters, c->isl_start, c->isl_list, c->C, c->big_list, c->prev_key, c->prev_imp);
        return NE_OK;
    }
    return solve_large_islands(c, P);
}

// Event lists for `cap` pairs.  Growing keeps the last step's list (the
// current one is rewritten by the pass).
int ensure_events(ne_gpu_ctx* c, uint32_t cap) {
    if (!c->ev_n) {
        HIP_TRY(c, dalloc(c->ev_n, 2));
        HIP_TRY(c, dalloc(c->ev_saved, CNT_N));
        HIP_TRY(c, hipMemsetAsync(c->ev_n, 0, sizeof(uint32_t) * 2, c->stream));
    }
    if (cap <= c->ev_cap) return NE_OK;
    const int prev = c->ev_cur;
    const uint32_t keep = c->h_ev_n[prev];
    for (int b = 0; b < 2; ++b) {
        uint64_t* k = nullptr;
        uint32_t *t = nullptr, *st = nullptr;
        double* d = nullptr;
        HIP_TRY(c, dalloc(k, cap));
        HIP_TRY(c, dalloc(t, cap));
        HIP_TRY(c, dalloc(st, cap));
        HIP_TRY(c, dalloc(d, (size_t)7 * cap));
        if (b == prev && keep) {
            HIP_TRY(c, hipMemcpyAsync(k, c->ev_key[b], sizeof(uint64_t) * keep, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(t, c->ev_trig[b], sizeof(uint32_t) * keep, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(st, c->ev_state[b], sizeof(uint32_t) * keep, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(d, c->ev_dat[b], sizeof(double) * 7 * keep, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
        }
        dfree(c->ev_key[b]); dfree(c->ev_trig[b]); dfree(c->ev_state[b]); dfree(c->ev_dat[b]);
        c->ev_key[b] = k; c->ev_trig[b] = t; c->ev_state[b] = st; c->ev_dat[b] = d;
    }
    HIP_TRY(c, dalloc(c->ev_off, (size_t)cap + 1));
    c->ev_cap = cap;
    return NE_OK;
}

// The facade events pass after a step (PhysicsWorld::processContacts,
// src/physics_world.cpp:253-302): the step's grid and narrow-phase stages on
// the post-step state with the facade's filter and one contact per pair,
// then the gather into the other list and the diff against the last one.
// The step's counters are kept (the pass adds only faults, EPA counts and
// its pair count), so a fault here reruns the call like one in the step.
int enqueue_events(ne_gpu_ctx* c) {
    if (c->nmeshes)
        return set_err(c, NE_ERR_UNSUPPORTED, "facade events with mesh bodies: the reference's processContacts runs "
                                              "detectContact on them as cylinders of their radius and half height");
    int r = ensure_events(c, c->pcap);
    if (r != NE_OK) return r;
    hipStream_t s = c->stream;
    SimParams P = make_sim_params(c->params);
    P.no_manifold = 1;
    BodiesDev B = c->B;
    B.ev_filter = 1;
    const int timed = c->timed_kernel;   // the step's stage timers stay off in this pass
    c->timed_kernel = NE_KERNEL_NONE;
    HIP_TRY(c, hipMemcpyAsync(c->ev_saved, c->counters, sizeof(uint32_t) * CNT_N, hipMemcpyDeviceToDevice, s));
    const bool static_grid = P.box.periodic && c->ngrid > 0;
    if (static_grid) k_ev_bin<true><<<blocks_for(c->n), TPB, 0, s>>>(B, c->sgrid, c->xp, c->cell_count);
    else k_ev_bin<false><<<blocks_for(c->n), TPB, 0, s>>>(B, c->sgrid, c->xp, c->cell_count);
    enqueue_detect(c, B, P, static_grid);
    scan_u32(c->ncont, c->ev_off, c->counters + CNT_P, c->pcap, c->scan_part, s, false,
             ScanTotal{c->counters, CNT_EV, CNT_EV, CNT_FAULT, c->pcap, NE_FAULT_PAIRS, -1, -1});
    const int cur = c->ev_cur ^ 1, prev = c->ev_cur;
    k_ev_gather<<<hint_blocks(c->hint_p), TPB, 0, s>>>(B, c->counters, c->cand_i, c->cand_j, c->ncont, c->ev_off,
                                                        c->ct, c->ev_key[cur], c->ev_trig[cur], c->ev_dat[cur],
                                                        c->ev_n + cur);
    k_ev_diff<<<hint_blocks(c->hint_p), TPB, 0, s>>>(c->ev_n, cur, prev, c->ev_key[cur], c->ev_trig[cur],
                                                      c->ev_key[prev], c->ev_trig[prev], c->ev_state[cur],
                                                      c->ev_state[prev]);
    k_ev_restore<<<1, 64, 0, s>>>(c->counters, c->ev_saved);
    c->timed_kernel = timed;
    HIP_TRY(c, hipGetLastError());
    return NE_OK;
}

int front_io(ne_gpu_ctx* c, ne_gpu_ctx::FrontSave& F, bool save);

int enqueue_step(ne_gpu_ctx* c, bool wrap_first) {
    if (c->n == 0) return NE_OK;
    KTimer whole(c, NE_KERNEL_STEP);
    enqueue_front(c, wrap_first);
    if (c->front_save) {
        const int rf = front_io(c, *c->front_save, true);
        if (rf != NE_OK) return rf;
    }
    const int rs = enqueue_solve(c);
    if (rs != NE_OK) return rs;
    const SimParams P = make_sim_params(c->params);
    if (c->nrigid)   // rigid joints after the contacts (world.cpp:139)
        k_joint_rigid<<<blocks_for(c->ncomp, 64), 64, 0, c->stream>>>(c->B, P, joint_dev(c));
    if (c->ngj) {    // then the general joints (world.cpp:140)
        const GJointDev G{c->gj, c->gj_applied, c->gj_motor, c->gj_comp, c->gj_comp + c->ngj_comp + 1, c->ngj_comp};
        k_joint_general<<<blocks_for(c->ngj_comp, 64), 64, 0, c->stream>>>(c->B, P, G);
    }
    // --- boundary.  The periodic wrap is deferred to the next integrate,
    // except with joints: spring forces at the next step's start read the
    // wrapped positions.
    if (!P.box.periodic && !P.open_boundary) {
        KTimer kt(c, NE_KERNEL_BOUNDARY);
        k_walls<<<blocks_for(c->n), TPB, 0, c->stream>>>(c->B, P);
    } else if (P.box.periodic && (c->njoints || c->ngj || P.sleep || c->events_on || c->nccd)) {
        KTimer kt(c, NE_KERNEL_BOUNDARY);
        k_wrap<<<blocks_for(c->n), TPB, 0, c->stream>>>(c->B, P);
    }
    if (P.sleep) {   // world.cpp:328-371
        hipStream_t s = c->stream;
        k_sleep_init<<<blocks_for(c->n), TPB, 0, s>>>(c->B, c->parent, c->mint);
        k_sleep_join_pairs<<<hint_blocks(c->hint_p), TPB, 0, s>>>(c->B, c->counters, c->cand_i, c->cand_j, c->isp,
                                                                  c->parent);
        if (c->njoints)
            k_sleep_join_joints<<<blocks_for(c->njoints), TPB, 0, s>>>(c->B, c->j_ab, c->j_ab + c->njoints, c->njoints,
                                                                        c->parent);
        if (c->ngj)
            k_sleep_join_gjoints<<<blocks_for(c->ngj), TPB, 0, s>>>(c->B, c->gj, c->ngj, c->parent);
        k_sleep_timers<<<blocks_for(c->n), TPB, 0, s>>>(c->B, P, c->parent, c->mint);
        k_sleep_apply<<<blocks_for(c->n), TPB, 0, s>>>(c->B, P, c->parent, c->mint);
    }
    return NE_OK;
}

int upload_bodies(ne_gpu_ctx* c, const ne_bodies* b);

// Rows of grid-excluded bodies: one flag per (excluded body, body of its
// world); the flag and offset arrays are sized for that.
int ensure_excl_flags(ne_gpu_ctx* c) {
    size_t total = 0;
    if (c->nworlds) {
        total = c->ex_total;
    } else {
        total = (size_t)c->nexcl * c->n;
    }
    if (total >= (1ull << 31))
        return set_err(c, NE_ERR_UNSUPPORTED, "too many grid-excluded bodies for this body count");
    c->ex_total = total;
    if (total == 0 || total <= c->ex_cap) return NE_OK;
    drop_graphs(c);
    HIP_TRY(c, dalloc(c->ex_flags, total));
    HIP_TRY(c, dalloc(c->ex_off, total + 1));
    c->ex_cap = total;
    return NE_OK;
}

// Cell count along one periodic axis of length L for cells >= `cell` wide.
int32_t axis_cells(double L, double cell) {
    double k = std::floor(L / cell);
    if (!(k >= 1.0)) k = 1.0;
    if (k > 1048576.0) k = 1048576.0;
    return (int32_t)k;
}

// Batched worlds: each world's periodic grid (cells >= 2*maxR of its gridded
// bodies, times cell_scale) gets a slice of the shared cell table; all worlds
// are coarsened together until the table fits its capacity.
int prepare_worlds(ne_gpu_ctx* c) {
    const uint32_t nw = c->nworlds;
    std::vector<double> maxR(nw, 0.0);
    std::vector<uint32_t> e0(nw, 0), e1(nw, 0);
    for (uint32_t w = 0; w < nw; ++w) {
        for (uint32_t i = c->w_off[w]; i < c->w_off[w + 1]; ++i)
            if (!c->h_isexcl[i]) maxR[w] = std::max(maxR[w], c->h_R[i]);
        e0[w] = (uint32_t)(std::lower_bound(c->h_excl.begin(), c->h_excl.end(), c->w_off[w]) - c->h_excl.begin());
        e1[w] = (uint32_t)(std::lower_bound(c->h_excl.begin(), c->h_excl.end(), c->w_off[w + 1]) - c->h_excl.begin());
    }
    std::vector<WorldDev> W(nw);
    double grow = 1.0;
    for (int attempt = 0;; ++attempt) {
        uint64_t total = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            const double cell = (2.0 * maxR[w] * (1.0 + 1e-6) + 1e-9) * c->cell_scale * grow;
            GridDev g{};
            for (int a = 0; a < 3; ++a) {
                const double L = 2.0 * c->w_half[3 * w + a];
                g.nc[a] = maxR[w] > 0.0 ? axis_cells(L, cell) : 1;
                g.cw[a] = L / (double)g.nc[a];
                g.inv_cw[a] = 1.0 / g.cw[a];
                g.origin[a] = -c->w_half[3 * w + a];
            }
            g.periodic = 1;
            g.ncells = (uint32_t)g.nc[0] * (uint32_t)g.nc[1] * (uint32_t)g.nc[2];
            W[w].grid = g;
            W[w].cell_base = (uint32_t)std::min<uint64_t>(total, 0xFFFFFFFFull);
            total += g.ncells;
        }
        if (total <= c->cellcap) { c->sgrid_cells = (uint32_t)total; break; }
        if (attempt > 200) return set_err(c, NE_ERR_CAPACITY, "cell table cannot hold the worlds' grids");
        grow *= 1.25;
    }
    size_t flag_base = 0;
    for (uint32_t w = 0; w < nw; ++w) {
        WorldDev& d = W[w];
        for (int a = 0; a < 3; ++a) {
            d.box.half[a] = c->w_half[3 * w + a];
            d.box.edge[a] = 2.0 * c->w_half[3 * w + a];
            d.box.inv_edge[a] = 1.0 / d.box.edge[a];
        }
        d.box.periodic = 1;
        d.restitution = c->w_rest[w];
        d.friction = c->w_fric[w];
        d.body_begin = c->w_off[w];
        d.body_end = c->w_off[w + 1];
        d.excl_begin = e0[w];
        d.excl_end = e1[w];
        d.flag_base = (uint32_t)std::min<size_t>(flag_base, 0xFFFFFFFFu);
        flag_base += (size_t)(e1[w] - e0[w]) * (d.body_end - d.body_begin);
    }
    c->ex_total = flag_base;
    HIP_TRY(c, dalloc(c->d_worlds, std::max<uint32_t>(nw, 1u)));
    HIP_TRY(c, hipMemcpy(c->d_worlds, W.data(), sizeof(WorldDev) * nw, hipMemcpyHostToDevice));
    GridDev gd{};
    gd.periodic = 1;
    gd.ncells = c->sgrid_cells;   // the cell-major walk visits every world's cells
    c->grid_host = gd;
    HIP_TRY(c, hipMemcpy(c->grid, &c->grid_host, sizeof(GridDev), hipMemcpyHostToDevice));
    c->B.W = c->d_worlds;
    c->B.wid = c->d_wid;
    return NE_OK;
}

// Periodic worlds: cells of width L/nc >= the gridded bodies' 2*maxR, with the
// cell table kept within its capacity; the same rule k_grid_setup applies.
int prepare_static_grid(ne_gpu_ctx* c) {
    if (!c->grid_dirty) return NE_OK;
    drop_graphs(c);
    c->sgrid_cells = 0;
    c->B.W = nullptr;
    c->B.wid = nullptr;
    if (c->nworlds) {
        int r = prepare_worlds(c);
        if (r != NE_OK) return r;
    } else if (c->params.periodic && c->ngrid > 0) {
        StaticGrid g{};
        GridDev gd{};
        for (int a = 0; a < 3; ++a) g.nc[a] = axis_cells(2.0 * c->params.half[a], c->cell);
        while ((double)g.nc[0] * g.nc[1] * g.nc[2] > (double)c->cellcap)
            for (int a = 0; a < 3; ++a) g.nc[a] = std::max(1, g.nc[a] * 4 / 5);
        for (int a = 0; a < 3; ++a) {
            g.cw[a] = 2.0 * c->params.half[a] / (double)g.nc[a];
            g.inv_cw[a] = 1.0 / g.cw[a];
            g.origin[a] = -c->params.half[a];
            gd.nc[a] = g.nc[a]; gd.cw[a] = g.cw[a]; gd.origin[a] = g.origin[a];
            gd.inv_cw[a] = 1.0 / g.cw[a];
        }
        gd.periodic = 1;
        gd.ncells = (uint32_t)g.nc[0] * (uint32_t)g.nc[1] * (uint32_t)g.nc[2];
        c->sgrid = g;
        c->sgrid_cells = gd.ncells;
        c->grid_host = gd;
        HIP_TRY(c, hipMemcpy(c->grid, &c->grid_host, sizeof(GridDev), hipMemcpyHostToDevice));
    }
    int r = ensure_excl_flags(c);
    if (r != NE_OK) return r;
    c->grid_dirty = false;
    return NE_OK;
}

// One step, eagerly or as a replay of its captured graph.  A graph bakes the
// cell-order buffers in (the fill reads the previous order and writes the new
// one), so one graph is kept per buffer parity and per wrap-first flag.
int run_step(ne_gpu_ctx* c, bool wrap_first) {
    const bool graphable = c->use_graph && !c->has_forces && c->timed_kernel == NE_KERNEL_NONE && !c->events_on &&
                           !c->front_save;
    if (!graphable) return enqueue_step(c, wrap_first);
    // launch sizes follow the counts seen when captured: recapture when they drift
    auto drift = [](uint32_t now, uint32_t then) { return now > 2u * then + 1024u || 2u * now + 1024u < then; };
    if (drift(c->hint_p, c->graph_hint_p) || drift(c->hint_c, c->graph_hint_c)) drop_graphs(c);
    hipGraphExec_t& ex = c->gexec[wrap_first ? 1 : 0][c->order_parity];
    if (!ex) {
        if (!c->gexec[0][0] && !c->gexec[0][1] && !c->gexec[1][0] && !c->gexec[1][1]) {
            c->graph_hint_p = c->hint_p;
            c->graph_hint_c = c->hint_c;
        }
        uint32_t* si = c->sorted_idx;
        uint32_t* op = c->order_prev;
        const int par = c->order_parity;
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        const int r = enqueue_step(c, wrap_first);
        c->capturing = false;
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        c->sorted_idx = si;   // nothing ran yet: undo the host-side bookkeeping
        c->order_prev = op;
        c->order_parity = par;
        if (r != NE_OK) { if (graph) (void)hipGraphDestroy(graph); return r; }
        if (e != hipSuccess) return set_err(c, NE_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        const hipError_t e2 = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e2 != hipSuccess) {
            ex = nullptr;
            return set_err(c, NE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
        }
    }
    HIP_TRY(c, hipGraphLaunch(ex, c->stream));
    std::swap(c->sorted_idx, c->order_prev);   // what the replayed fill did
    c->order_parity ^= 1;
    return NE_OK;
}


// ---------------------------------------------------------------------------
// Rollouts and their adjoint
// ---------------------------------------------------------------------------

int save_state(ne_gpu_ctx* c, ne_gpu_ctx::Saved& d) {
    const size_t nd = (size_t)13 * c->ncap;
    if (!d.state) HIP_TRY(c, dalloc(d.state, nd));
    if (!d.n) HIP_TRY(c, dalloc(d.n, 1));
    if (d.cap < c->prev_cap) {
        HIP_TRY(c, dalloc(d.key, c->prev_cap));
        HIP_TRY(c, dalloc(d.imp, c->prev_cap));
        d.cap = c->prev_cap;
    }
    HIP_TRY(c, hipMemcpyAsync(d.state, c->dfield, sizeof(double) * nd, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(d.key, c->prev_key, sizeof(uint64_t) * c->prev_cap, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(d.imp, c->prev_imp, sizeof(double) * c->prev_cap, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(d.n, c->prev_n, sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    return NE_OK;
}

int load_state(ne_gpu_ctx* c, const ne_gpu_ctx::Saved& d) {
    const size_t nd = (size_t)13 * c->ncap;
    const uint32_t m = std::min(d.cap, c->prev_cap);
    HIP_TRY(c, hipMemcpyAsync(c->dfield, d.state, sizeof(double) * nd, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->prev_key, d.key, sizeof(uint64_t) * m, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->prev_imp, d.imp, sizeof(double) * m, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->prev_n, d.n, sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    return NE_OK;
}

// Copy a step's front half into F (save) or back (restore): the body
// fields k_integrate wrote, every constraint field, the island lists and the
// counters.  F grows to the current capacities on a save.
int front_io(ne_gpu_ctx* c, ne_gpu_ctx::FrontSave& F, bool save) {
    hipStream_t s = c->stream;
    if (save && (F.ccap < c->ccap || F.ncap < c->ncap)) {
        HIP_TRY(c, dalloc(F.body, (size_t)13 * c->ncap));
        HIP_TRY(c, dalloc(F.cd, (size_t)18 * c->ccap));
        HIP_TRY(c, dalloc(F.cu, (size_t)4 * c->ccap));
        HIP_TRY(c, dalloc(F.key, c->ccap));
        HIP_TRY(c, dalloc(F.isl, (size_t)2 * c->ccap + 1));
        if (!F.cnt) HIP_TRY(c, dalloc(F.cnt, CNT_N));
        F.ccap = c->ccap;
        F.ncap = c->ncap;
    }
    if (save) {   // the tape, within ADJ_TAPE_STEP_MAX bytes a step (else the reverse replays the solve)
        const size_t want = (size_t)std::max(c->params.solver_iterations, 0) * c->ccap * TAPE_DOUBLES;
        F.taped = false;
        if (want * sizeof(double) <= ADJ_TAPE_STEP_MAX) {
            if (F.tape_n < want) {
                HIP_TRY(c, dalloc(F.tape, want));
                F.tape_n = want;
            }
            F.tape_stride = c->ccap;
            F.taped = true;
        }
    }
    if (!save && (F.ccap > c->ccap || F.ncap != c->ncap)) return set_err(c, NE_ERR_INVALID, "adjoint: capacities changed");
    const size_t cc = F.ccap;
    auto cp = [&](void* live, void* kept, size_t bytes) -> hipError_t {
        return save ? hipMemcpyAsync(kept, live, bytes, hipMemcpyDeviceToDevice, s)
                    : hipMemcpyAsync(live, kept, bytes, hipMemcpyDeviceToDevice, s);
    };
    ConsDev& C = c->C;
    double* dp[] = {C.n0, C.n1, C.n2, C.p0, C.p1, C.p2, C.ra0, C.ra1, C.ra2, C.rb0, C.rb1, C.rb2,
                    C.ov, C.rbias, C.mu, C.nimp, C.timp, C.kn};
    uint32_t* up[] = {C.ci, C.cj, C.slot, C.src};
    HIP_TRY(c, cp(c->dfield, F.body, sizeof(double) * 13 * (size_t)F.ncap));
    for (int f = 0; f < 18; ++f) HIP_TRY(c, cp(dp[f], F.cd + (size_t)f * cc, sizeof(double) * cc));
    for (int f = 0; f < 4; ++f) HIP_TRY(c, cp(up[f], F.cu + (size_t)f * cc, sizeof(uint32_t) * cc));
    HIP_TRY(c, cp(C.key, F.key, sizeof(uint64_t) * cc));
    HIP_TRY(c, cp(c->isl_start, F.isl, sizeof(uint32_t) * (cc + 1)));
    HIP_TRY(c, cp(c->isl_list, F.isl + cc + 1, sizeof(uint32_t) * cc));
    HIP_TRY(c, cp(c->counters, F.cnt, sizeof(uint32_t) * CNT_N));
    return NE_OK;
}

int ensure_saved(ne_gpu_ctx* c, std::vector<ne_gpu_ctx::Saved>& v, size_t count) {
    if (v.size() > count) {
        for (size_t k = count; k < v.size(); ++k) { dfree(v[k].state); dfree(v[k].key); dfree(v[k].imp); dfree(v[k].n); }
        v.resize(count);
    }
    v.resize(count);
    return NE_OK;
}

// Caller n x 3 arrays may be host or device memory: they go through a device
// staging buffer with hipMemcpyDefault.
int stage_in(ne_gpu_ctx* c, const double* src, double* f0, double* f1, double* f2) {
    if (c->stage_n < c->n) {
        HIP_TRY(c, dalloc(c->stage, (size_t)3 * c->ncap));
        c->stage_n = c->ncap;
    }
    HIP_TRY(c, hipMemcpyAsync(c->stage, src, sizeof(double) * 3 * c->n, hipMemcpyDefault, c->stream));
    k_scatter3<<<blocks_for(c->n), TPB, 0, c->stream>>>(c->stage, c->n, f0, f1, f2);
    return NE_OK;
}

int stage_out(ne_gpu_ctx* c, const double* f0, const double* f1, const double* f2, double* dst) {
    if (c->stage_n < c->n) {
        HIP_TRY(c, dalloc(c->stage, (size_t)3 * c->ncap));
        c->stage_n = c->ncap;
    }
    k_gather3<<<blocks_for(c->n), TPB, 0, c->stream>>>(f0, f1, f2, c->n, c->stage);
    HIP_TRY(c, hipMemcpyAsync(dst, c->stage, sizeof(double) * 3 * c->n, hipMemcpyDefault, c->stream));
    return NE_OK;
}

void fill(ne_gpu_ctx* c, double* p, size_t n, double v) {
    if (n) k_fill_f64<<<gs_blocks(n), TPB, 0, c->stream>>>(p, n, v);
}

}  // namespace

extern "C" {

int ne_gpu_rollout(ne_gpu_ctx* c, const double* v0, const double* mu, const double* e, int32_t nsteps, int32_t every,
                   double* x_out) {
    if (!c || nsteps <= 0 || every <= 0) return NE_ERR_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    if (c->n == 0) return set_err(c, NE_ERR_INVALID, "no bodies");
    if (!c->params.periodic) return set_err(c, NE_ERR_UNSUPPORTED, "rollouts are differentiated for periodic worlds");
    if (c->has_forces || c->njoints || c->ngj || c->params.sleep_enabled || c->nmeshes || c->nccd)
        return set_err(c, NE_ERR_UNSUPPORTED, "rollouts do not take applied forces, joints, sleeping, meshes or CCD");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const uint32_t nw = c->nworlds ? c->nworlds : 1u;
    if (mu || e) {
        if (c->nworlds) {
            for (uint32_t w = 0; w < nw; ++w) {
                if (mu) c->w_fric[w] = mu[w];
                if (e) c->w_rest[w] = e[w];
            }
        } else {
            if (mu) c->params.friction = mu[0];
            if (e) c->params.restitution = e[0];
        }
        c->grid_dirty = true;
    }
    int r = prepare_static_grid(c);
    if (r != NE_OK) return r;
    if (v0) {
        r = stage_in(c, v0, c->B.v0, c->B.v1, c->B.v2);
        if (r != NE_OK) return r;
    }
    HIP_TRY(c, hipMemsetAsync(c->prev_n, 0, sizeof(uint32_t), c->stream));   // a rollout starts with no warm cache
    const int32_t nck = (nsteps + every - 1) / every;
    ensure_saved(c, c->ckpt, (size_t)nck);
    bool capped = false;
    for (int32_t k = 0; k < nck; ++k) {
        r = save_state(c, c->ckpt[k]);
        if (r != NE_OK) return r;
        const int32_t m = std::min(every, nsteps - k * every);
        r = ne_gpu_step(c, m);
        if (r == NE_WARN_EPA_CAPPED) { capped = true; r = NE_OK; }   // reported once the rollout is complete
        if (r != NE_OK) return r;
    }
    if (x_out) {
        r = stage_out(c, c->B.x0, c->B.x1, c->B.x2, x_out);
        if (r != NE_OK) return r;
    }
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->roll_steps = nsteps;
    c->roll_every = every;
    c->adj_nw = nw;
    if (capped)
        return set_err(c, NE_WARN_EPA_CAPPED, "the rollout completed, but an EPA run stopped at the workspace's face "
                                              "cap in at least one step");
    return NE_OK;
}

int ne_gpu_rollout_backward(ne_gpu_ctx* c, const double* gx, double* gv0, double* gmu, double* ge) {
    if (!c || !gx) return NE_ERR_INVALID;
    if (c->roll_steps <= 0) return set_err(c, NE_ERR_INVALID, "no rollout to differentiate");
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    const uint32_t n = c->n, ncap = c->ncap, ccap = c->ccap, nw = c->adj_nw;
    const int iters = std::max(c->params.solver_iterations, 1);
    // work space
    if (c->adj_ncap != ncap) {
        HIP_TRY(c, dalloc(c->adj, (size_t)13 * ncap));
        HIP_TRY(c, dalloc(c->adj_pre, (size_t)6 * ncap));
        c->adj_ncap = ncap;
    }
    if (c->adj_ccap != ccap) {
        HIP_TRY(c, dalloc(c->adj_ca, (size_t)CA_COUNT * ccap));
        HIP_TRY(c, dalloc(c->adj_nfin, ccap));
        HIP_TRY(c, dalloc(c->adj_nprev, ccap));
        HIP_TRY(c, dalloc(c->adj_cw, (size_t)2 * ccap));
        c->adj_ccap = ccap;
        c->adj_tape_n = 0;
    }
    const size_t tape_n = (size_t)iters * ccap * TAPE_DOUBLES;
    if (c->adj_tape_n < tape_n) {
        HIP_TRY(c, dalloc(c->adj_tape, tape_n));
        c->adj_tape_n = tape_n;
    }
    HIP_TRY(c, dalloc(c->adj_gw, (size_t)2 * nw));
    hipStream_t s = c->stream;
    fill(c, c->adj, (size_t)13 * ncap, 0.0);
    fill(c, c->adj_nfin, ccap, 0.0);
    fill(c, c->adj_gw, (size_t)2 * nw, 0.0);
    {
        int r = stage_in(c, gx, c->adj, c->adj + ncap, c->adj + 2 * (size_t)ncap);
        if (r != NE_OK) return r;
    }
    const int32_t every = c->roll_every, nsteps = c->roll_steps;
    const int32_t nck = (nsteps + every - 1) / every;
    ensure_saved(c, c->seg, (size_t)every);
    if (c->fronts.size() != (size_t)every) {
        free_fronts(c->fronts);
        c->fronts.resize((size_t)every);
    }
    AdjBodies A{c->adj, ncap};
    for (int32_t k = nck - 1; k >= 0; --k) {
        const int32_t t0 = k * every, t1 = std::min(nsteps, t0 + every);
        // states at the start of every step of the segment
        int r = load_state(c, c->ckpt[k]);
        if (r != NE_OK) return r;
        for (int32_t t = t0; t < t1; ++t) {
            r = save_state(c, c->seg[t - t0]);
            if (r != NE_OK) return r;
            if (t + 1 < t1) {   // the step, keeping its front half for the reverse sweep
                c->front_save = &c->fronts[t - t0];
                r = ne_gpu_step(c, 1);
                c->front_save = nullptr;
                if (r == NE_WARN_EPA_CAPPED) r = NE_OK;   // the forward rollout reported it
                if (r != NE_OK) return r;
            }
        }
        for (int32_t t = t1 - 1; t >= t0; --t) {
            if (t + 1 < t1) {   // the front half the recompute kept (the same inputs, the same bits)
                r = front_io(c, c->fronts[t - t0], false);
                if (r != NE_OK) return r;
            } else {            // the segment's last step: its front half now
                r = load_state(c, c->seg[t - t0]);
                if (r != NE_OK) return r;
                r = prepare_static_grid(c);
                if (r != NE_OK) return r;
                HIP_TRY(c, hipMemsetAsync(c->counters, 0, sizeof(uint32_t) * CNT_N, s));
                enqueue_front(c, false);
            }
            const SimParams P = make_sim_params(c->params);
            BodiesDev B = c->B;
            HIP_TRY(c, hipMemcpyAsync(c->adj_pre, B.v0, sizeof(double) * 3 * ncap, hipMemcpyDeviceToDevice, s));
            HIP_TRY(c, hipMemcpyAsync(c->adj_pre + 3 * (size_t)ncap, B.w0, sizeof(double) * 3 * ncap,
                                      hipMemcpyDeviceToDevice, s));
            fill(c, c->adj_nprev, ccap, 0.0);
            // the recompute's tape when its solve wrote one, else the island pass replays the solve
            const ne_gpu_ctx::FrontSave* F = (t + 1 < t1 && c->fronts[t - t0].taped) ? &c->fronts[t - t0] : nullptr;
            k_adj_island<<<std::min<uint32_t>(std::max<uint32_t>(c->hint_c, 64u), 16384u), 64, 0, s>>>(
                B, P, c->counters, c->isl_start, c->isl_list, c->C, ccap, F ? F->tape : c->adj_tape,
                F ? F->tape_stride : ccap, F == nullptr, c->adj_pre, ncap, A, c->adj_ca, c->adj_nfin, c->adj_nprev,
                c->adj_cw);
            k_adj_integrate<<<blocks_for(n), TPB, 0, s>>>(B, P, c->seg[t - t0].state, ncap, A);
            k_adj_world_sum<<<nw, 256, 0, s>>>(B, c->counters, c->C, c->adj_cw, ccap, c->adj_gw, nw);
            std::swap(c->adj_nfin, c->adj_nprev);
            HIP_TRY(c, hipGetLastError());
            HIP_TRY(c, hipMemcpyAsync(c->h_counters, c->counters, sizeof(uint32_t) * CNT_N, hipMemcpyDeviceToHost, s));
            HIP_TRY(c, hipStreamSynchronize(s));
            if (c->h_counters[CNT_FAULT])
                return set_err(c, NE_ERR_CAPACITY, "a step overflowed while recomputing the rollout");
        }
    }
    if (gv0) {
        int r = stage_out(c, c->adj + 3 * (size_t)ncap, c->adj + 4 * (size_t)ncap, c->adj + 5 * (size_t)ncap, gv0);
        if (r != NE_OK) return r;
    }
    std::vector<double> gw(2 * (size_t)nw);
    HIP_TRY(c, hipMemcpyAsync(gw.data(), c->adj_gw, sizeof(double) * 2 * nw, hipMemcpyDeviceToHost, s));
    // leave the world at the rollout's start
    int r = load_state(c, c->ckpt[0]);
    if (r != NE_OK) return r;
    HIP_TRY(c, hipStreamSynchronize(s));
    for (uint32_t w = 0; w < nw; ++w) {
        if (gmu) gmu[w] = gw[w];
        if (ge) ge[w] = gw[nw + w];
    }
    return NE_OK;
}

}  // extern "C"

namespace {

int step_call(ne_gpu_ctx* c, int32_t nsteps);

// Start-of-call copy of what a step changes (pose, velocity, force and torque
// fields, which lead the field table, and the warm cache), taken on the
// stream, so a call that overflows a buffer can be run again from its start.
// With island sleeping, masses, timers and flags change too: then the whole
// body table is kept.
size_t snapshot_fields(const ne_gpu_ctx* c) { return c->params.sleep_enabled ? (size_t)F_COUNT : (size_t)(F_T2 + 1); }

int snapshot_take(ne_gpu_ctx* c) {
    const size_t nd = snapshot_fields(c) * c->ncap;
    if (c->snap_nd < nd) {
        HIP_TRY(c, dalloc(c->snap_body, nd));
        c->snap_nd = nd;
    }
    if (c->snap_pcap < c->prev_cap) {
        HIP_TRY(c, dalloc(c->snap_key, c->prev_cap));
        HIP_TRY(c, dalloc(c->snap_imp, c->prev_cap));
        c->snap_pcap = c->prev_cap;
    }
    if (!c->snap_n) HIP_TRY(c, dalloc(c->snap_n, 1));
    c->snap_prev = c->prev_cap;
    HIP_TRY(c, hipMemcpyAsync(c->snap_body, c->dfield, sizeof(double) * nd, hipMemcpyDeviceToDevice, c->stream));
    if (c->params.sleep_enabled) {
        if (c->snap_fcap < c->ncap) {
            HIP_TRY(c, dalloc(c->snap_flags, c->ncap));
            c->snap_fcap = c->ncap;
        }
        HIP_TRY(c, hipMemcpyAsync(c->snap_flags, c->ufield + (size_t)U_FLAGS * c->ncap, sizeof(uint32_t) * c->ncap,
                                  hipMemcpyDeviceToDevice, c->stream));
    }
    HIP_TRY(c, hipMemcpyAsync(c->snap_key, c->prev_key, sizeof(uint64_t) * c->prev_cap, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->snap_imp, c->prev_imp, sizeof(double) * c->prev_cap, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->snap_n, c->prev_n, sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    return NE_OK;
}

int snapshot_restore(ne_gpu_ctx* c) {
    const size_t nd = snapshot_fields(c) * c->ncap;
    const uint32_t m = std::min(c->snap_prev, c->prev_cap);
    HIP_TRY(c, hipMemcpyAsync(c->dfield, c->snap_body, sizeof(double) * nd, hipMemcpyDeviceToDevice, c->stream));
    if (c->params.sleep_enabled && c->snap_flags)
        HIP_TRY(c, hipMemcpyAsync(c->ufield + (size_t)U_FLAGS * c->ncap, c->snap_flags, sizeof(uint32_t) * c->ncap,
                                  hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->prev_key, c->snap_key, sizeof(uint64_t) * m, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->prev_imp, c->snap_imp, sizeof(double) * m, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->prev_n, c->snap_n, sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    return NE_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

int ne_gpu_create(const ne_gpu_config* cfg, ne_gpu_ctx** out) {
    if (!out) return NE_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return NE_ERR_NO_DEVICE;
    ne_gpu_ctx* c = new ne_gpu_ctx();
    c->device = cfg ? cfg->device : 0;
    if (c->device < 0 || c->device >= ndev) { delete c; return NE_ERR_INVALID; }
    c->user_pcap = cfg ? cfg->pair_capacity : 0;
    c->user_ccap = cfg ? cfg->contact_capacity : 0;
    c->cell_scale = (cfg && cfg->cell_scale >= 1.0) ? cfg->cell_scale : 1.0;
    c->broadphase = cfg ? cfg->broadphase : NE_BROADPHASE_AUTO;
    c->cells_per_body = (cfg && cfg->cells_per_body > 0) ? cfg->cells_per_body : 8u;
    c->epa_faces_max = (cfg && cfg->epa_face_cap > 0) ? (int)std::min<uint32_t>(std::max<uint32_t>(cfg->epa_face_cap, 64u), 1u << 26)
                                                       : EPA_BIG_FACES_MAX;
    if (c->broadphase < NE_BROADPHASE_AUTO || c->broadphase > NE_BROADPHASE_CELLS) { delete c; return NE_ERR_INVALID; }
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return NE_ERR_HIP;
    }
    // World defaults (src/world.hpp:75-126, src/box.hpp:17-18)
    ne_world_params& p = c->params;
    p.half[0] = p.half[1] = p.half[2] = 12.0;
    p.periodic = 0;
    p.dt = 1.0 / 120.0;
    p.restitution = 1.0;
    p.friction = 0.0;
    p.contact_beta = 0.2;
    p.restitution_combine = NE_COMBINE_MAX;
    p.friction_combine = NE_COMBINE_GEOMETRIC_MEAN;
    p.warm_start = 1;
    p.solver_iterations = 8;
    p.sleep_lin_vel = 0.05; p.sleep_ang_vel = 0.05; p.sleep_time = 0.5; p.wake_vel = 0.5;
    if (hipMalloc((void**)&c->counters, sizeof(uint32_t) * CNT_N) != hipSuccess ||
        hipHostMalloc((void**)&c->h_counters, sizeof(uint32_t) * CNT_N, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&c->prev_n, sizeof(uint32_t)) != hipSuccess ||
        hipMemsetAsync(c->prev_n, 0, sizeof(uint32_t), c->stream) != hipSuccess ||
        hipMemsetAsync(c->counters, 0, sizeof(uint32_t) * CNT_N, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        free_all(c);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return NE_ERR_HIP;
    }
    memset(c->h_counters, 0, sizeof(uint32_t) * CNT_N);
    c->scratch_dev = scratch_device(c->device);   // the device is current (hipSetDevice above)
    if (!c->scratch_dev) {
        free_all(c);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return NE_ERR_HIP;
    }
    *out = c;
    return NE_OK;
}

int ne_gpu_destroy(ne_gpu_ctx* c) {
    if (!c) return NE_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->ev_start) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->ev_stop) (void)hipEventDestroy(e);
    if (c->side) {
        (void)hipStreamSynchronize(c->side);
        (void)hipEventDestroy(c->ev_fork);
        (void)hipEventDestroy(c->ev_join);
        (void)hipStreamDestroy(c->side);
    }
    free_all(c);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return NE_OK;
}

int ne_gpu_set_params(ne_gpu_ctx* c, const ne_world_params* p) {
    if (!c || !p) return NE_ERR_INVALID;
    if (!(p->dt > 0.0) || p->solver_iterations < 0)
        return set_err(c, NE_ERR_INVALID, "dt must be > 0 and solver_iterations >= 0");
    for (int a = 0; a < 3; ++a)
        if (!(p->half[a] > 0.0)) return set_err(c, NE_ERR_INVALID, "box half extents must be > 0");
    if (p->restitution_combine < 0 || p->restitution_combine > 4 || p->friction_combine < 0 || p->friction_combine > 4)
        return set_err(c, NE_ERR_INVALID, "unknown combine mode");
    if (p->sleep_enabled && !(p->wake_vel >= 0.0)) return set_err(c, NE_ERR_INVALID, "wake_vel must be >= 0");
    if (c->nworlds && !p->periodic) return set_err(c, NE_ERR_UNSUPPORTED, "batched worlds must be periodic");
    c->params = *p;
    c->have_params = true;
    c->grid_dirty = true;
    c->B.sleep_on = p->sleep_enabled ? 1 : 0;
    drop_graphs(c);
    return NE_OK;
}

int reset_events(ne_gpu_ctx* c) {
    c->h_ev_n[0] = c->h_ev_n[1] = 0;
    if (c->ev_n) {
        HIP_TRY(c, hipMemsetAsync(c->ev_n, 0, sizeof(uint32_t) * 2, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return NE_OK;
}

int ne_gpu_set_bodies(ne_gpu_ctx* c, const ne_bodies* b) {
    if (!c || !b) return NE_ERR_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    const int r = upload_bodies(c, b);
    if (r != NE_OK) return r;
    return reset_events(c);   // a new world: no last-step pairs
}

int ne_gpu_set_events(ne_gpu_ctx* c, int32_t enable) {
    if (!c) return NE_ERR_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    if (enable && c->nmeshes)
        return set_err(c, NE_ERR_UNSUPPORTED, "facade events with mesh bodies: the reference's processContacts runs "
                                              "detectContact on them as cylinders of their radius and half height");
    if (enable && c->nworlds)
        return set_err(c, NE_ERR_UNSUPPORTED, "facade events are for one world (PhysicsWorld), not batched worlds");
    c->events_on = enable != 0;
    drop_graphs(c);
    return reset_events(c);   // the facade's maps start empty
}

int ne_gpu_get_events(ne_gpu_ctx* c, int32_t kind, uint32_t cap, uint32_t* count, uint32_t* ei, uint32_t* ej,
                      double* point, double* normal, double* depth) {
    if (!c || !count || kind < NE_EVENT_CONTACT || kind > NE_EVENT_TRIGGER_EXIT) return NE_ERR_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    const bool from_prev = kind == NE_EVENT_END || kind == NE_EVENT_TRIGGER_EXIT;
    const uint32_t trig = kind >= NE_EVENT_TRIGGER_ENTER ? 1u : 0u;
    const int b = from_prev ? (c->ev_cur ^ 1) : c->ev_cur;
    const uint32_t n = c->h_ev_n[b];
    *count = 0;
    if (!c->events_on || n == 0 || !c->ev_key[b]) return NE_OK;
    std::vector<uint64_t> key(n);
    std::vector<uint32_t> tr(n), st(n);
    std::vector<double> dat(7 * (size_t)n);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(key.data(), c->ev_key[b], sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(tr.data(), c->ev_trig[b], sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(st.data(), c->ev_state[b], sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(dat.data(), c->ev_dat[b], sizeof(double) * 7 * n, hipMemcpyDeviceToHost));
    uint32_t m = 0;
    for (uint32_t k = 0; k < n; ++k) {
        if (tr[k] != trig) continue;
        bool take;
        if (from_prev) take = st[k] == 0;                                    // gone this step
        else if (kind == NE_EVENT_CONTACT) take = true;                      // every solid pair
        else if (kind == NE_EVENT_BEGIN || kind == NE_EVENT_TRIGGER_ENTER) take = st[k] == 0;
        else take = st[k] != 0;                                              // stay
        if (!take) continue;
        if (m < cap) {
            if (ei) ei[m] = (uint32_t)(key[k] >> 32);
            if (ej) ej[m] = (uint32_t)(key[k] & 0xFFFFFFFFu);
            const double* d = dat.data() + 7 * (size_t)k;
            if (normal) { normal[3 * m] = d[0]; normal[3 * m + 1] = d[1]; normal[3 * m + 2] = d[2]; }
            if (point) { point[3 * m] = d[3]; point[3 * m + 1] = d[4]; point[3 * m + 2] = d[5]; }
            if (depth) depth[m] = d[6];
        }
        ++m;
    }
    *count = m;
    return NE_OK;
}

}  // extern "C"

// ---- scene queries (PhysicsWorld, src/physics_world.cpp:146-251) --------

namespace {


// Carves 256-byte aligned pieces from an arena; with base == nullptr it only
// measures.
struct Carve {
    uint8_t* base = nullptr;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += sizeof(T) * (count ? count : 1);
        return p;
    }
};

hipError_t ensure_arena(uint8_t*& mem, size_t& cap, size_t need) {
    if (mem && cap >= need) return hipSuccess;
    dfree(mem);
    cap = 0;
    const hipError_t e = hipMalloc((void**)&mem, need);
    if (e == hipSuccess) cap = need;
    return e;
}

int query_world_check(ne_gpu_ctx* c) {
    if (c->nworlds) return set_err(c, NE_ERR_UNSUPPORTED, "scene queries are for one world (PhysicsWorld), not batched worlds");
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    return NE_OK;
}

// The rays' inputs on the device, as structure of arrays.
struct RayUpload {
    RayIn in{};
    double* tiles = nullptr;
    bool cull = false;
};

int upload_rays(ne_gpu_ctx* c, Carve& cv, uint32_t n, const double* origin, const double* dir, const double* max_dist,
                const uint32_t* mask, const double* rad, RayUpload& u) {
    const uint32_t ntiles = (c->n + RAY_THREADS - 1) / RAY_THREADS;
    double* f = cv.take<double>((size_t)8 * n);
    uint32_t* m = mask ? cv.take<uint32_t>(n) : nullptr;
    u.tiles = cv.take<double>((size_t)6 * (ntiles ? ntiles : 1));
    if (!cv.base) return NE_OK;
    std::vector<double> h((size_t)8 * n);
    for (uint32_t k = 0; k < n; ++k) {
        for (int a = 0; a < 3; ++a) {
            h[(size_t)a * n + k] = origin[3 * (size_t)k + a];
            h[(size_t)(3 + a) * n + k] = dir[3 * (size_t)k + a];
        }
        h[(size_t)6 * n + k] = max_dist[k];
        h[(size_t)7 * n + k] = rad ? rad[k] : 0.0;
    }
    HIP_TRY(c, hipMemcpyAsync(f, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
    if (m) HIP_TRY(c, hipMemcpyAsync(m, mask, sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
    u.in = RayIn{f, f + n, f + 2 * (size_t)n, f + 3 * (size_t)n, f + 4 * (size_t)n, f + 5 * (size_t)n,
                 f + 6 * (size_t)n, f + 7 * (size_t)n, m};
    // Tiles are skipped by their bounding boxes in a closed box; a periodic
    // world tests every body at its image nearest each ray's origin.
    u.cull = !make_sim_params(c->params).box.periodic;
    if (u.cull && ntiles) k_query_tiles<true><<<ntiles, RAY_THREADS, 0, c->stream>>>(c->B, u.tiles);
    return NE_OK;
}

RayOut carve_ray_out(Carve& cv, size_t m) {
    RayOut o{};
    o.body = cv.take<int32_t>(m);
    double* f = cv.take<double>(7 * m);
    o.t = f;
    o.px = f + m; o.py = f + 2 * m; o.pz = f + 3 * m;
    o.nx = f + 4 * m; o.ny = f + 5 * m; o.nz = f + 6 * m;
    return o;
}

// Device results (m hits) back into the caller's packed arrays at `at`.
int download_hits(ne_gpu_ctx* c, const RayOut& o, size_t m, size_t at, int32_t* body, double* distance, double* point,
                  double* normal) {
    if (m == 0) return NE_OK;
    std::vector<double> h(7 * m);
    HIP_TRY(c, hipMemcpyAsync(body + at, o.body, sizeof(int32_t) * m, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(h.data(), o.t, sizeof(double) * 7 * m, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < m; ++k) {
        distance[at + k] = h[k];
        for (int a = 0; a < 3; ++a) {
            point[3 * (at + k) + a] = h[(1 + a) * m + k];
            normal[3 * (at + k) + a] = h[(4 + a) * m + k];
        }
    }
    return NE_OK;
}

template <int MODE, bool CAST>
void launch_ray(ne_gpu_ctx* c, uint32_t n, const RayUpload& u, const RayOut& o) {
    const Box3 box = make_sim_params(c->params).box;
    const uint32_t blocks = (n + RAY_THREADS - 1) / RAY_THREADS;
    if (u.cull)
        k_ray<MODE, CAST, true><<<blocks, RAY_THREADS, 0, c->stream>>>(c->B, box, n, u.in, u.tiles, o);
    else
        k_ray<MODE, CAST, false><<<blocks, RAY_THREADS, 0, c->stream>>>(c->B, box, n, u.in, u.tiles, o);
}

// EPA of the queued overlapContacts pairs: one thread per pair with 256
// faces and 128 horizon edges each (chunks of OVL_CHUNK pairs sharing the
// workspace), then the runs that did not fit, one block per pair with the
// context's face cap (OVL_BIG_BLOCKS blocks at a time).  *capped_out: runs
// stopped at the cap.
constexpr int OVL_FACES = 256;
constexpr uint32_t OVL_CHUNK = 8192, OVL_BIG_BLOCKS = 32;

size_t ovl_epa_bytes(const ne_gpu_ctx* c) {
    const size_t small = OVL_CHUNK * (sizeof(SupV) * EPA_MAXV + sizeof(Face) * OVL_FACES + sizeof(int) * OVL_FACES);
    const size_t big = OVL_BIG_BLOCKS * (size_t)c->epa_faces_max * (sizeof(Face) + 3 * sizeof(uint16_t));
    return std::max(small, big) + 3 * 256;
}

int ovl_epa(ne_gpu_ctx* c, const QueryDev* Q, uint32_t* ent, uint32_t nent, uint32_t* ent2, uint32_t* ctr,
            double* nrm, double* dep, uint8_t* ws, uint32_t* capped_out) {
    const Box3 box = make_sim_params(c->params).box;
    *capped_out = 0;
    HIP_TRY(c, hipMemsetAsync(ctr, 0, 2 * sizeof(uint32_t), c->stream));
    {
        const uint32_t ch = std::min(OVL_CHUNK, nent);
        Carve w{ws, 0};
        SupV* wv = w.take<SupV>((size_t)ch * EPA_MAXV);
        Face* wf = w.take<Face>((size_t)ch * OVL_FACES);
        int(*wh)[2] = reinterpret_cast<int(*)[2]>(w.take<int>((size_t)ch * OVL_FACES));
        for (uint32_t b = 0; b < nent; b += ch) {
            const uint32_t m = std::min(ch, nent - b);
            k_ovl_epa<<<(m + 63) / 64, 64, 0, c->stream>>>(c->B, box, Q, ent + 3 * (size_t)b, m, wv, wf, wh, OVL_FACES,
                                                          OVL_FACES / 2, nrm, dep, ent2, ctr);
        }
    }
    HIP_TRY(c, hipGetLastError());
    uint32_t nbig = 0;
    HIP_TRY(c, hipMemcpyAsync(&nbig, ctr, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (nbig) {
        Carve w{ws, 0};
        const int fp = c->epa_faces_max;
        Face* faces = w.take<Face>((size_t)OVL_BIG_BLOCKS * fp);
        uint16_t* occ = w.take<uint16_t>((size_t)OVL_BIG_BLOCKS * 3 * fp);
        k_ovl_big<<<std::min(nbig, OVL_BIG_BLOCKS), 64 * BIG_NW, 0, c->stream>>>(c->B, box, Q, ent2, nbig, faces, occ,
                                                                                fp, nrm, dep, ctr + 1);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(capped_out, ctr + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return NE_OK;
}

}  // namespace

extern "C" {

int ne_gpu_raycast(ne_gpu_ctx* c, uint32_t n, const double* origin, const double* dir, const double* max_dist,
                   const uint32_t* mask, const double* cast_radius, int32_t* body, double* distance, double* point,
                   double* normal) {
    if (!c || (n && (!origin || !dir || !max_dist || !body || !distance || !point || !normal))) return NE_ERR_INVALID;
    int r = query_world_check(c);
    if (r != NE_OK || n == 0) return r;
    RayUpload u;
    Carve cv;
    upload_rays(c, cv, n, origin, dir, max_dist, mask, cast_radius, u);
    carve_ray_out(cv, n);
    HIP_TRY(c, ensure_arena(c->q_mem_a, c->q_cap_a, cv.off));
    cv = Carve{c->q_mem_a, 0};
    if ((r = upload_rays(c, cv, n, origin, dir, max_dist, mask, cast_radius, u)) != NE_OK) return r;
    const RayOut o = carve_ray_out(cv, n);
    {
        KTimer kt(c, NE_KERNEL_RAY);
        if (cast_radius) launch_ray<0, true>(c, n, u, o);
        else launch_ray<0, false>(c, n, u, o);
    }
    HIP_TRY(c, hipGetLastError());
    if ((r = download_hits(c, o, n, 0, body, distance, point, normal)) != NE_OK) return r;
    harvest_timing(c);
    return NE_OK;
}

int ne_gpu_raycast_all(ne_gpu_ctx* c, uint32_t n, const double* origin, const double* dir, const double* max_dist,
                       const uint32_t* mask, uint32_t cap, uint32_t* offsets, int32_t* body, double* distance,
                       double* point, double* normal) {
    if (!c || !offsets || (n && (!origin || !dir || !max_dist))) return NE_ERR_INVALID;
    offsets[0] = 0;
    int r = query_world_check(c);
    if (r != NE_OK || n == 0) return r;
    RayUpload u;
    Carve cv;
    upload_rays(c, cv, n, origin, dir, max_dist, mask, nullptr, u);
    cv.take<uint32_t>(n);
    cv.take<uint32_t>((size_t)n + 1);
    cv.take<uint32_t>(1025);
    HIP_TRY(c, ensure_arena(c->q_mem_a, c->q_cap_a, cv.off));
    cv = Carve{c->q_mem_a, 0};
    if ((r = upload_rays(c, cv, n, origin, dir, max_dist, mask, nullptr, u)) != NE_OK) return r;
    uint32_t* cnt = cv.take<uint32_t>(n);
    uint32_t* off = cv.take<uint32_t>((size_t)n + 1);
    uint32_t* part = cv.take<uint32_t>(1025);
    RayOut o{};
    o.count = cnt;
    launch_ray<1, false>(c, n, u, o);
    scan_u32(cnt, off, nullptr, n, part, c->stream);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(offsets, off, sizeof(uint32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const size_t total = offsets[n];
    if (total > cap) return set_err(c, NE_ERR_CAPACITY, "raycast_all: more hits than cap (offsets[n] holds the count)");
    if (total == 0) return NE_OK;
    if (!body || !distance || !point || !normal) return NE_ERR_INVALID;
    Carve cb;
    carve_ray_out(cb, total);
    HIP_TRY(c, ensure_arena(c->q_mem_b, c->q_cap_b, cb.off));
    cb = Carve{c->q_mem_b, 0};
    RayOut e = carve_ray_out(cb, total);
    e.off = off;
    launch_ray<2, false>(c, n, u, e);
    HIP_TRY(c, hipGetLastError());
    return download_hits(c, e, total, 0, body, distance, point, normal);
}

int ne_gpu_overlap(ne_gpu_ctx* c, uint32_t n, const ne_query_shape* shapes, const uint32_t* mask, int32_t contacts,
                   uint32_t cap, uint32_t* offsets, int32_t* body, double* normal, double* depth) {
    if (!c || !offsets || (n && !shapes)) return NE_ERR_INVALID;
    offsets[0] = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const ne_query_shape& s = shapes[k];
        if (s.shape < NE_SHAPE_SPHERE || s.shape > NE_SHAPE_PLANE) return set_err(c, NE_ERR_INVALID, "query shape kind");
        if (s.shape == NE_SHAPE_CONVEX && (!s.verts || s.nverts == 0))
            return set_err(c, NE_ERR_INVALID, "convex query shape without vertices");
    }
    int r = query_world_check(c);
    if (r != NE_OK) return r;
    if (c->nmeshes)
        return set_err(c, NE_ERR_UNSUPPORTED, "overlap queries with mesh bodies: detectContact would take a mesh body "
                                              "for a cylinder of its radius and half height");
    if (n == 0) return NE_OK;
    if (c->n == 0) {
        for (uint32_t k = 0; k < n; ++k) offsets[k + 1] = 0;
        return NE_OK;
    }
    const uint32_t nchunks = (c->n + OVL_THREADS - 1) / OVL_THREADS;
    const uint32_t qb = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({(uint64_t)n, 65535u, (16u << 20) / nchunks}));
    size_t nverts = 0;
    for (uint32_t k = 0; k < n; ++k)
        if (shapes[k].shape == NE_SHAPE_CONVEX) nverts += shapes[k].nverts;
    auto carve_a = [&](Carve& cv, QueryDev*& Q, double*& hull, uint32_t*& cnt, uint32_t*& off, uint32_t*& part,
                       uint32_t*& qoff, uint32_t*& ctr, double*& tiles) {
        Q = cv.take<QueryDev>(n);
        hull = cv.take<double>(3 * nverts);
        cnt = cv.take<uint32_t>((size_t)qb * nchunks);
        off = cv.take<uint32_t>((size_t)qb * nchunks + 1);
        part = cv.take<uint32_t>(1025);
        qoff = cv.take<uint32_t>((size_t)qb + 1);
        ctr = cv.take<uint32_t>(4);
        tiles = cv.take<double>(6 * (size_t)nchunks);
    };
    QueryDev* Q;
    double *hull, *tiles;
    uint32_t *cnt, *off, *part, *qoff, *ctr;
    Carve cv;
    carve_a(cv, Q, hull, cnt, off, part, qoff, ctr, tiles);
    HIP_TRY(c, ensure_arena(c->q_mem_a, c->q_cap_a, cv.off));
    cv = Carve{c->q_mem_a, 0};
    carve_a(cv, Q, hull, cnt, off, part, qoff, ctr, tiles);
    // query geometry: the Body the reference builds (fields as given), its
    // hull in device memory, and a bounding radius padded for the pair test
    std::vector<QueryDev> hq(n);
    std::vector<double> hv(3 * nverts);
    size_t vat = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const ne_query_shape& s = shapes[k];
        QueryDev& d = hq[k];
        d.g.x = V3{s.x[0], s.x[1], s.x[2]};
        d.g.q = Qt{s.q[0], s.q[1], s.q[2], s.q[3]};
        d.g.he = V3{s.half_extents[0], s.half_extents[1], s.half_extents[2]};
        d.g.radius = s.radius;
        d.g.hh = s.half_height;
        d.g.shape = s.shape;
        d.g.hull = nullptr;
        d.g.nhull = 0;
        if (s.shape == NE_SHAPE_CONVEX) {
            std::memcpy(hv.data() + 3 * vat, s.verts, sizeof(double) * 3 * s.nverts);
            d.g.hull = hull + 3 * vat;
            d.g.nhull = s.nverts;
            vat += s.nverts;
        }
        const double q2 = s.q[0] * s.q[0] + s.q[1] * s.q[1] + s.q[2] * s.q[2] + s.q[3] * s.q[3];
        d.bR = std::fabs(q2 - 1.0) > 1e-6 ? 1e30
                                          : host_bounding_radius(s.shape, s.radius, s.half_height, s.half_extents,
                                                                 s.verts, 0, s.shape == NE_SHAPE_CONVEX ? s.nverts : 0);
        d.mask = mask ? mask[k] : 0xFFFFFFFFu;
    }
    HIP_TRY(c, hipMemcpyAsync(Q, hq.data(), sizeof(QueryDev) * n, hipMemcpyHostToDevice, c->stream));
    if (nverts) HIP_TRY(c, hipMemcpyAsync(hull, hv.data(), sizeof(double) * hv.size(), hipMemcpyHostToDevice, c->stream));
    const Box3 box = make_sim_params(c->params).box;
    // tiles of bodies skipped whole by their bounds in a closed box (a
    // periodic world tests every pair at its minimum image)
    const double* tl = nullptr;
    if (!box.periodic) {
        k_query_tiles<false><<<nchunks, RAY_THREADS, 0, c->stream>>>(c->B, tiles);
        tl = tiles;
    }
    // pass 1: hits per query (every batch), into offsets
    std::vector<uint32_t> h_qoff((size_t)qb + 1);
    auto count_batch = [&](uint32_t q0, uint32_t m) -> int {
        k_ovl<false><<<dim3(nchunks, m), OVL_THREADS, 0, c->stream>>>(c->B, box, Q, q0, tl, contacts, cnt, nullptr,
                                                                       nullptr, nullptr, nullptr, nullptr, nullptr);
        scan_u32(cnt, off, nullptr, m * nchunks, part, c->stream);
        k_ovl_qoff<<<(m + 1 + 255) / 256, 256, 0, c->stream>>>(off, nchunks, m, qoff);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(h_qoff.data(), qoff, sizeof(uint32_t) * ((size_t)m + 1), hipMemcpyDeviceToHost,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return NE_OK;
    };
    uint64_t total = 0, max_batch = 0;
    for (uint32_t q0 = 0; q0 < n; q0 += qb) {
        const uint32_t m = std::min(qb, n - q0);
        if ((r = count_batch(q0, m)) != NE_OK) return r;
        for (uint32_t t = 1; t <= m; ++t) {
            const uint64_t v = total + h_qoff[t];
            offsets[q0 + t] = (uint32_t)std::min<uint64_t>(v, 0xFFFFFFFFu);
        }
        total += h_qoff[m];
        max_batch = std::max<uint64_t>(max_batch, h_qoff[m]);
    }
    if (total > cap) return set_err(c, NE_ERR_CAPACITY, "overlap: more hits than cap (offsets[n] holds the count)");
    if (total == 0) return NE_OK;
    if (!body) return NE_ERR_INVALID;
    // pass 2: bodies in order, contacts, EPA of the convex pairs
    auto carve_b = [&](Carve& w, int32_t*& ob, double*& on, double*& od, uint32_t*& e1, uint32_t*& e2) {
        ob = w.take<int32_t>(max_batch);
        on = w.take<double>(3 * max_batch);
        od = w.take<double>(max_batch);
        e1 = w.take<uint32_t>(3 * max_batch);
        e2 = w.take<uint32_t>(3 * max_batch);
    };
    int32_t* ob;
    double *on, *od;
    uint32_t *e1, *e2;
    Carve w;
    carve_b(w, ob, on, od, e1, e2);
    HIP_TRY(c, ensure_arena(c->q_mem_b, c->q_cap_b, w.off));
    w = Carve{c->q_mem_b, 0};
    carve_b(w, ob, on, od, e1, e2);
    if (contacts) HIP_TRY(c, ensure_arena(c->q_mem_w, c->q_cap_w, ovl_epa_bytes(c)));
    uint32_t capped_all = 0;
    uint64_t done = 0;
    for (uint32_t q0 = 0; q0 < n; q0 += qb) {
        const uint32_t m = std::min(qb, n - q0);
        if (n > qb && (r = count_batch(q0, m)) != NE_OK) return r;   // one batch: pass 1's offsets still hold
        const uint32_t bt = h_qoff[m];
        if (bt) {
            if (contacts) HIP_TRY(c, hipMemsetAsync(ctr + 2, 0, sizeof(uint32_t), c->stream));
            k_ovl<true><<<dim3(nchunks, m), OVL_THREADS, 0, c->stream>>>(c->B, box, Q, q0, tl, contacts, nullptr, off,
                                                                          ob, on, od, e1, ctr + 2);
            HIP_TRY(c, hipGetLastError());
            if (contacts) {
                uint32_t ne = 0;
                HIP_TRY(c, hipMemcpyAsync(&ne, ctr + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
                HIP_TRY(c, hipStreamSynchronize(c->stream));
                if (ne) {
                    uint32_t capped = 0;
                    if ((r = ovl_epa(c, Q, e1, ne, e2, ctr, on, od, c->q_mem_w, &capped)) != NE_OK) return r;
                    capped_all += capped;
                }
            }
            HIP_TRY(c, hipMemcpyAsync(body + done, ob, sizeof(int32_t) * bt, hipMemcpyDeviceToHost, c->stream));
            if (contacts && normal)
                HIP_TRY(c, hipMemcpyAsync(normal + 3 * done, on, sizeof(double) * 3 * bt, hipMemcpyDeviceToHost, c->stream));
            if (contacts && depth)
                HIP_TRY(c, hipMemcpyAsync(depth + done, od, sizeof(double) * bt, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
        }
        done += bt;
    }
    return capped_all ? NE_WARN_EPA_CAPPED : NE_OK;
}

int ne_gpu_set_general_joints(ne_gpu_ctx* c, uint32_t n, const ne_joint* j) {
    if (!c || (n && !j)) return NE_ERR_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    if (n && c->nworlds) return set_err(c, NE_ERR_UNSUPPORTED, "general joints in batched worlds");
    for (uint32_t k = 0; k < n; ++k) {
        if (j[k].a >= c->n || j[k].b >= c->n) return set_err(c, NE_ERR_INVALID, "joint body index out of range");
        if (j[k].a == j[k].b) return set_err(c, NE_ERR_INVALID, "a joint needs two different bodies");
        if (j[k].type < NE_JOINT_BALL || j[k].type > NE_JOINT_SLIDER) return set_err(c, NE_ERR_INVALID, "joint type");
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    c->ngj = c->ngj_comp = 0;
    if (n == 0) return NE_OK;
    // components: union-find over the bodies that can move; joints between
    // immovable bodies alone form components of their own
    std::vector<uint32_t> parent(c->n);
    for (uint32_t i = 0; i < c->n; ++i) parent[i] = i;
    auto find = [&](uint32_t x) { while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; } return x; };
    for (uint32_t k = 0; k < n; ++k)
        if (c->h_mut[j[k].a] && c->h_mut[j[k].b]) {
            const uint32_t ra = find(j[k].a), rb = find(j[k].b);
            if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
        }
    std::vector<uint32_t> comp(n), count, comp_of_root(c->n, 0xFFFFFFFFu);
    uint32_t ncomp = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t a = j[k].a, b = j[k].b;
        if (!c->h_mut[a] && !c->h_mut[b]) { comp[k] = ncomp++; count.push_back(1); continue; }
        const uint32_t root = c->h_mut[a] ? find(a) : find(b);
        uint32_t& id = comp_of_root[root];
        if (id == 0xFFFFFFFFu) { id = ncomp++; count.push_back(0); }
        comp[k] = id;
        ++count[id];
    }
    std::vector<uint32_t> cl(ncomp + 1 + n, 0), fill(ncomp, 0);
    for (uint32_t q = 0; q < ncomp; ++q) cl[q + 1] = cl[q] + count[q];
    for (uint32_t k = 0; k < n; ++k) cl[ncomp + 1 + cl[comp[k]] + fill[comp[k]]++] = k;   // ascending joint order
    HIP_TRY(c, dalloc(c->gj, n));
    HIP_TRY(c, dalloc(c->gj_applied, n));
    HIP_TRY(c, dalloc(c->gj_motor, n));
    HIP_TRY(c, dalloc(c->gj_comp, cl.size()));
    HIP_TRY(c, hipMemcpy(c->gj, j, sizeof(ne_joint) * n, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->gj_comp, cl.data(), sizeof(uint32_t) * cl.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemset(c->gj_applied, 0, sizeof(double) * n));
    HIP_TRY(c, hipMemset(c->gj_motor, 0, sizeof(double) * n));
    c->ngj = n;
    c->ngj_comp = ncomp;
    return NE_OK;
}

int ne_gpu_get_general_joints(ne_gpu_ctx* c, uint32_t n, int32_t* broken, double* applied, double* motor) {
    if (!c || n > c->ngj) return NE_ERR_INVALID;
    if (n == 0) return NE_OK;
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    std::vector<ne_joint> h(n);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(h.data(), c->gj, sizeof(ne_joint) * n, hipMemcpyDeviceToHost));
    if (applied) HIP_TRY(c, hipMemcpy(applied, c->gj_applied, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (motor) HIP_TRY(c, hipMemcpy(motor, c->gj_motor, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (broken)
        for (uint32_t k = 0; k < n; ++k) broken[k] = h[k].broken;
    return NE_OK;
}

int ne_gpu_set_joints(ne_gpu_ctx* c, uint32_t n, const ne_distance_joint* j) {
    if (!c || (n && !j)) return NE_ERR_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    for (uint32_t k = 0; k < n; ++k)
        if (j[k].a >= c->n || j[k].b >= c->n) return set_err(c, NE_ERR_INVALID, "joint body index out of range");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    c->njoints = c->nspring = c->nrigid = c->ncomp = 0;
    if (n == 0) return NE_OK;
    // components: union-find over the bodies that can move
    std::vector<uint32_t> parent(c->n);
    for (uint32_t i = 0; i < c->n; ++i) parent[i] = i;
    auto find = [&](uint32_t x) { while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; } return x; };
    for (uint32_t k = 0; k < n; ++k)
        if (c->h_mut[j[k].a] && c->h_mut[j[k].b]) {
            const uint32_t ra = find(j[k].a), rb = find(j[k].b);
            if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
        }
    std::vector<uint32_t> root(n), comp_of_root(c->n, 0xFFFFFFFFu), count;
    uint32_t nspring = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t a = j[k].a, b = j[k].b;
        root[k] = c->h_mut[a] ? find(a) : (c->h_mut[b] ? find(b) : c->n + k);
        nspring += j[k].stiffness > 0.0;
    }
    std::vector<uint32_t> comp(n);
    uint32_t ncomp = 0;   // a joint with no movable body gets a component of its own
    for (uint32_t k = 0; k < n; ++k) {
        if (root[k] >= c->n) { comp[k] = ncomp++; count.push_back(1); continue; }
        uint32_t& id = comp_of_root[root[k]];
        if (id == 0xFFFFFFFFu) { id = ncomp++; count.push_back(0); }
        comp[k] = id;
        ++count[id];
    }
    std::vector<uint32_t> start(ncomp + 1, 0), list(n), fill(ncomp, 0);
    for (uint32_t q = 0; q < ncomp; ++q) start[q + 1] = start[q] + count[q];
    for (uint32_t k = 0; k < n; ++k) list[start[comp[k]] + fill[comp[k]]++] = k;   // ascending joint order
    std::vector<uint32_t> ab(2 * (size_t)n);
    std::vector<double> f(9 * (size_t)n);
    for (uint32_t k = 0; k < n; ++k) {
        ab[k] = j[k].a;
        ab[n + k] = j[k].b;
        for (int a = 0; a < 3; ++a) {
            f[(size_t)a * n + k] = j[k].local_a[a];
            f[(size_t)(3 + a) * n + k] = j[k].local_b[a];
        }
        f[6 * (size_t)n + k] = j[k].rest;
        f[7 * (size_t)n + k] = j[k].stiffness;
        f[8 * (size_t)n + k] = j[k].damping;
    }
    std::vector<uint32_t> cl(start.begin(), start.end());
    cl.insert(cl.end(), list.begin(), list.end());
    HIP_TRY(c, dalloc(c->j_ab, ab.size()));
    HIP_TRY(c, dalloc(c->j_f, f.size()));
    HIP_TRY(c, dalloc(c->j_comp, cl.size()));
    HIP_TRY(c, hipMemcpy(c->j_ab, ab.data(), sizeof(uint32_t) * ab.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->j_f, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->j_comp, cl.data(), sizeof(uint32_t) * cl.size(), hipMemcpyHostToDevice));
    c->njoints = n;
    c->nspring = nspring;
    c->nrigid = n - nspring;
    c->ncomp = ncomp;
    return NE_OK;
}

int ne_gpu_set_state(ne_gpu_ctx* c, const double* x, const double* v, const double* q, const double* w) {
    if (!c) return NE_ERR_INVALID;
    const uint32_t n = c->n;
    if (n == 0) return NE_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<double> tmp((size_t)n);
    auto put = [&](const double* src, int comps, int stride, const int* fields) -> int {
        for (int k = 0; k < comps; ++k) {
            for (uint32_t i = 0; i < n; ++i) tmp[i] = src[(size_t)stride * i + k];
            HIP_TRY(c, hipMemcpy(c->dfield + (size_t)fields[k] * c->ncap, tmp.data(), sizeof(double) * n,
                                 hipMemcpyHostToDevice));
        }
        return NE_OK;
    };
    int r;
    const int fx[3] = {F_X0, F_X1, F_X2}, fv[3] = {F_V0, F_V1, F_V2}, fq[4] = {F_Q0, F_Q1, F_Q2, F_Q3},
              fw[3] = {F_W0, F_W1, F_W2};
    if (x && (r = put(x, 3, 3, fx)) != NE_OK) return r;
    if (v && (r = put(v, 3, 3, fv)) != NE_OK) return r;
    if (q && (r = put(q, 4, 4, fq)) != NE_OK) return r;
    if (w && (r = put(w, 3, 3, fw)) != NE_OK) return r;
    return NE_OK;
}

int ne_gpu_step(ne_gpu_ctx* c, int32_t nsteps) {
    if (!c) return NE_ERR_INVALID;
    if (nsteps < 0) return set_err(c, NE_ERR_INVALID, "nsteps < 0");
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    if (c->n == 0 || nsteps == 0) return NE_OK;
    if (c->events_on && nsteps != 1)
        return set_err(c, NE_ERR_INVALID, "with facade events on, ne_gpu_step takes one step per call");
    int r = snapshot_take(c);
    if (r != NE_OK) return r;
    const bool forces = c->has_forces;
    for (int attempt = 0;; ++attempt) {
        r = prepare_static_grid(c);
        if (r != NE_OK) return r;
        r = step_call(c, nsteps);
        if ((r == NE_OK || r == NE_WARN_EPA_CAPPED) && c->events_on) {   // the step's list becomes the last one
            c->ev_cur ^= 1;
            c->h_ev_n[c->ev_cur] = c->h_counters[CNT_EV];
        }
        if (r == NE_OK && c->h_counters[CNT_EPATRUNC]) {
            char buf[160];
            snprintf(buf, sizeof buf, "%u EPA run(s) stopped at the %d-face workspace cap; their contacts can differ "
                     "from the reference's", c->h_counters[CNT_EPATRUNC], c->epa_faces_max);
            (void)set_err(c, NE_WARN_EPA_CAPPED, buf);
            return NE_WARN_EPA_CAPPED;
        }
        if (r != NE_ERR_CAPACITY || attempt == 4) return r;
        // a buffer overflowed and has been enlarged: rerun the call from its start
        r = snapshot_restore(c);
        if (r != NE_OK) return r;
        c->has_forces = forces;
    }
}

}  // extern "C"

namespace {

int step_call(ne_gpu_ctx* c, int32_t nsteps) {
    HIP_TRY(c, hipMemsetAsync(c->counters, 0, sizeof(uint32_t) * CNT_N, c->stream));
    c->wf_seen = false;
    for (int32_t t = 0; t < nsteps; ++t) {
        // (a step whose front half is kept for the adjoint keeps its islands)
        c->solve_all = c->all_ok && !c->front_save && ++c->all_steps % WF_ALL_RECHECK != 0;
        if (c->solve_all) {   // the constraint build writes the solver's records
            const int re = ensure_wf(c, 0, 1);
            if (re != NE_OK) { c->solve_all = false; return re; }
        }
        int r = run_step(c, t > 0);
        c->solve_all = false;
        if (r != NE_OK) return r;
        if (c->events_on) {
            r = enqueue_events(c);
            if (r != NE_OK) return r;
        }
    }
    if (c->params.periodic && nsteps > 0 && c->njoints == 0 && !c->params.sleep_enabled && !c->events_on) {   // else each step wrapped itself
        KTimer kt(c, NE_KERNEL_BOUNDARY);
        k_wrap<<<blocks_for(c->n), TPB, 0, c->stream>>>(c->B, make_sim_params(c->params));
    }
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(c->h_counters, c->counters, sizeof(uint32_t) * CNT_N, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    harvest_timing(c);
    c->hint_p = c->h_counters[CNT_PRAW];
    c->hint_c = c->h_counters[CNT_CRAW];
    if (c->nmeshes) c->hint_mq = c->h_counters[CNT_MQRAW];
    const uint32_t fault = c->h_counters[CNT_FAULT];
    if (!fault && c->wf_expect && !c->wf_seen && c->wf_quiet >= 8) c->wf_expect = false;   // for the next call
    if (fault == NE_FAULT_WF_PLAN) {   // large islands after all: rerun the call, reading their tasks back
        c->wf_expect = true;
        c->wf_quiet = 0;
        return set_err(c, NE_ERR_CAPACITY, "large islands appeared; the call reruns");
    }
    if (fault & NE_FAULT_WF_PLAN) {
        c->wf_expect = true;
        c->wf_quiet = 0;
    }
    if (fault) {
        // grow what overflowed so a re-upload + retry succeeds; state is undefined
        if (fault & NE_FAULT_PAIRS) {
            uint32_t want = std::max(c->pcap * 2u, c->h_counters[CNT_PRAW] + c->h_counters[CNT_PRAW] / 2u + 1024u);
            int r = alloc_pairs(c, want);
            if (r != NE_OK) return r;
        }
        if (fault & (NE_FAULT_EPA_QUEUE | NE_FAULT_EPA_BIG)) {
            uint32_t slots = c->epa_big.cap;
            int per = c->epa_big.faces_per;
            if (fault & NE_FAULT_EPA_QUEUE)
                slots = std::max(slots * 2u, c->h_counters[CNT_EPABIG] + c->h_counters[CNT_EPABIG] / 2u + 16u);
            if (fault & NE_FAULT_EPA_BIG) per = std::min(per * 4, c->epa_faces_max);
            int r = alloc_epa_big(c, slots, per);
            if (r != NE_OK) return r;
        }
        if (fault & NE_FAULT_MESH) {
            uint32_t want = std::max(c->mqcap * 2u, c->h_counters[CNT_MQRAW] + c->h_counters[CNT_MQRAW] / 2u + 1024u);
            int r = alloc_mesh_cands(c, want);
            if (r != NE_OK) return r;
        }
        if (fault & NE_FAULT_CONTACTS) {
            uint32_t want = std::max(c->ccap * 2u, c->h_counters[CNT_CRAW] + c->h_counters[CNT_CRAW] / 2u + 1024u);
            int r = alloc_cons(c, want);
            if (r != NE_OK) return r;
            r = ensure_prev(c, want);
            if (r != NE_OK) return r;
        }
        char buf[160];
        snprintf(buf, sizeof buf, "step overflowed (fault bits 0x%x: pairs %u/%u, contacts %u/%u); capacity raised, re-upload",
                 fault, c->h_counters[CNT_PRAW], c->pcap, c->h_counters[CNT_CRAW], c->ccap);
        if (fault & NE_FAULT_WAKE)
            return set_err(c, NE_ERR_UNSUPPORTED, "a chain of woken bodies longer than the wake rounds; state is undefined");
        return set_err(c, (fault & NE_FAULT_EPA) ? NE_ERR_INVALID : NE_ERR_CAPACITY, buf);
    }
    return NE_OK;
}

}  // namespace

extern "C" {

int ne_gpu_get_state(ne_gpu_ctx* c, double* x, double* v, double* q, double* w) {
    if (!c) return NE_ERR_INVALID;
    const uint32_t n = c->n;
    if (n == 0) return NE_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<double> tmp((size_t)n);
    auto get = [&](double* dst, int comps, int stride, const int* fields) -> int {
        for (int k = 0; k < comps; ++k) {
            HIP_TRY(c, hipMemcpy(tmp.data(), c->dfield + (size_t)fields[k] * c->ncap, sizeof(double) * n,
                                 hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < n; ++i) dst[(size_t)stride * i + k] = tmp[i];
        }
        return NE_OK;
    };
    int r;
    const int fx[3] = {F_X0, F_X1, F_X2}, fv[3] = {F_V0, F_V1, F_V2}, fq[4] = {F_Q0, F_Q1, F_Q2, F_Q3},
              fw[3] = {F_W0, F_W1, F_W2};
    if (x && (r = get(x, 3, 3, fx)) != NE_OK) return r;
    if (v && (r = get(v, 3, 3, fv)) != NE_OK) return r;
    if (q && (r = get(q, 4, 4, fq)) != NE_OK) return r;
    if (w && (r = get(w, 3, 3, fw)) != NE_OK) return r;
    return NE_OK;
}

int ne_gpu_get_body_aux(ne_gpu_ctx* c, uint32_t* flags, double* sleep_timer, double* inv_mass, double* inv_inertia) {
    if (!c) return NE_ERR_INVALID;
    const uint32_t n = c->n;
    if (n == 0) return NE_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (flags) {
        HIP_TRY(c, hipMemcpy(flags, c->ufield + (size_t)U_FLAGS * c->ncap, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) flags[i] &= ~FLAG_GRID_EXCLUDED;
    }
    if (sleep_timer) HIP_TRY(c, hipMemcpy(sleep_timer, c->dfield + (size_t)F_ST * c->ncap, sizeof(double) * n,
                                          hipMemcpyDeviceToHost));
    if (inv_mass) HIP_TRY(c, hipMemcpy(inv_mass, c->dfield + (size_t)F_IM * c->ncap, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (inv_inertia) {
        std::vector<double> tmp(n);
        const int f[3] = {F_II0, F_II1, F_II2};
        for (int k = 0; k < 3; ++k) {
            HIP_TRY(c, hipMemcpy(tmp.data(), c->dfield + (size_t)f[k] * c->ncap, sizeof(double) * n, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < n; ++i) inv_inertia[3 * i + k] = tmp[i];
        }
    }
    return NE_OK;
}

int ne_gpu_get_sleep_store(ne_gpu_ctx* c, double* inv_mass_store, double* inv_inertia_store) {
    if (!c) return NE_ERR_INVALID;
    const uint32_t n = c->n;
    if (n == 0) return NE_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (inv_mass_store)
        HIP_TRY(c, hipMemcpy(inv_mass_store, c->dfield + (size_t)F_IMS * c->ncap, sizeof(double) * n,
                             hipMemcpyDeviceToHost));
    if (inv_inertia_store) {
        std::vector<double> tmp(n);
        for (int k = 0; k < 3; ++k) {
            HIP_TRY(c, hipMemcpy(tmp.data(), c->dfield + (size_t)(F_IIS0 + k) * c->ncap, sizeof(double) * n,
                                 hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < n; ++i) inv_inertia_store[3 * i + k] = tmp[i];
        }
    }
    return NE_OK;
}

// The general joints' broken flags as the snapshot's bytes, and back
// (physics_world.cpp:36-37, 56-61).
__global__ void k_nes1_joints_pack(const ne_joint* __restrict__ gj, uint32_t n, uint8_t* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = gj[k].broken ? 1u : 0u;
}
__global__ void k_nes1_joints_unpack(ne_joint* __restrict__ gj, uint32_t n, const uint8_t* __restrict__ in) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) gj[k].broken = in[k] != 0 ? 1 : 0;
}

int ensure_nes1(ne_gpu_ctx* c, uint64_t bytes) {
    if (c->nes1_cap >= bytes && c->nes1) return NE_OK;
    HIP_TRY(c, dalloc(c->nes1, (size_t)bytes));
    c->nes1_cap = bytes;
    return NE_OK;
}

int ne_gpu_save_state(ne_gpu_ctx* c, void* out, uint64_t cap, uint64_t* size) {
    if (!c || !size) return NE_ERR_INVALID;
    const uint32_t n = c->n;
    const uint64_t recs = NES1_RECORD * (uint64_t)n;
    const uint32_t ngj = c->ngj;
    const uint64_t need = NES1_HEAD + recs + 4 + ngj;
    *size = need;
    if (!out) return NE_OK;
    if (cap < need) return set_err(c, NE_ERR_INVALID, "ne_gpu_save_state: the buffer is smaller than the snapshot");
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    uint8_t* o = static_cast<uint8_t*>(out);
    const uint32_t head[2] = {NES1_MAGIC, n};
    std::memcpy(o, head, sizeof(head));
    if (n || ngj) {
        const int r = ensure_nes1(c, recs + ngj);
        if (r != NE_OK) return r;
    }
    if (n) {
        k_nes1_pack<<<blocks_for(n), TPB, 0, c->stream>>>(c->dfield, c->ncap, c->ufield + (size_t)U_FLAGS * c->ncap, n,
                                                          c->nes1);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(o + NES1_HEAD, c->nes1, (size_t)recs, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    std::memcpy(o + NES1_HEAD + recs, &ngj, 4);   // the general joints' count and broken flags
    if (ngj) {
        k_nes1_joints_pack<<<blocks_for(ngj), TPB, 0, c->stream>>>(c->gj, ngj, c->nes1 + recs);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(o + NES1_HEAD + recs + 4, c->nes1 + recs, ngj, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return NE_OK;
}

int ne_gpu_save_state_noalloc(ne_gpu_ctx* c, void* out, uint64_t cap, uint64_t* size) {
    if (!c || !size) return NE_ERR_INVALID;
    const uint32_t n = c->n;
    const uint64_t recs = NES1_RECORD * (uint64_t)n;
    const uint32_t ngj = c->ngj;
    const uint64_t need = NES1_HEAD + recs + 4 + ngj;
    *size = need;
    if (recs + ngj > NOALLOC_SCRATCH_BYTES) {
        char buf[200];
        snprintf(buf, sizeof buf, "ne_gpu_save_state_noalloc: %u bodies need %llu bytes of staging; its static buffer "
                 "holds %llu (%llu bodies): use ne_gpu_save_state", n, (unsigned long long)(recs + ngj),
                 (unsigned long long)NOALLOC_SCRATCH_BYTES, (unsigned long long)(NOALLOC_SCRATCH_BYTES / NES1_RECORD));
        return set_err(c, NE_ERR_INVALID, buf);
    }
    if (!out) return NE_OK;
    if (cap < need) return set_err(c, NE_ERR_INVALID, "ne_gpu_save_state_noalloc: the buffer is smaller than the snapshot");
    if (!c->scratch_dev) return set_err(c, NE_ERR_HIP, "ne_gpu_save_state_noalloc: no device address for the staging");
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    uint8_t* o = static_cast<uint8_t*>(out);
    const uint32_t head[2] = {NES1_MAGIC, n};
    std::memcpy(o, head, sizeof(head));
    if (n) {
        k_nes1_pack<<<blocks_for(n), TPB, 0, c->stream>>>(c->dfield, c->ncap, c->ufield + (size_t)U_FLAGS * c->ncap, n,
                                                          c->scratch_dev);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(o + NES1_HEAD, c->scratch_dev, (size_t)recs, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    std::memcpy(o + NES1_HEAD + recs, &ngj, 4);   // the general joints' count and broken flags
    if (ngj) {
        k_nes1_joints_pack<<<blocks_for(ngj), TPB, 0, c->stream>>>(c->gj, ngj, c->scratch_dev + recs);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(o + NES1_HEAD + recs + 4, c->scratch_dev + recs, ngj, hipMemcpyDeviceToHost,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return NE_OK;
}

int ne_gpu_load_state(ne_gpu_ctx* c, const void* data, uint64_t size) {
    if (!c || (size && !data)) return NE_ERR_INVALID;
    const uint8_t* d = static_cast<const uint8_t*>(data);
    const uint32_t n = c->n;
    const uint64_t recs = NES1_RECORD * (uint64_t)n;
    uint32_t head[2] = {0, 0};
    if (size < NES1_HEAD) return set_err(c, NE_ERR_INVALID, "ne_gpu_load_state: shorter than the snapshot header");
    std::memcpy(head, d, sizeof(head));
    if (head[0] != NES1_MAGIC) return set_err(c, NE_ERR_INVALID, "ne_gpu_load_state: not an NES1 snapshot");
    if (head[1] != n) return set_err(c, NE_ERR_INVALID, "ne_gpu_load_state: body count differs from the world's");
    if (size < NES1_HEAD + recs + 4) return set_err(c, NE_ERR_INVALID, "ne_gpu_load_state: truncated snapshot");
    uint32_t njoints = 0;
    std::memcpy(&njoints, d + NES1_HEAD + recs, 4);
    if (njoints != c->ngj)
        return set_err(c, NE_ERR_INVALID, "ne_gpu_load_state: the snapshot's general-joint count differs from the world's");
    if (size < NES1_HEAD + recs + 4 + njoints) return set_err(c, NE_ERR_INVALID, "ne_gpu_load_state: truncated snapshot");
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    int r = ensure_nes1(c, recs + njoints);
    if (r != NE_OK) return r;
    auto unpack_joints = [&]() -> int {   // the general joints' broken flags
        if (!njoints) return NE_OK;
        HIP_TRY(c, hipMemcpyAsync(c->nes1 + recs, d + NES1_HEAD + recs + 4, njoints, hipMemcpyHostToDevice, c->stream));
        k_nes1_joints_unpack<<<blocks_for(njoints), TPB, 0, c->stream>>>(c->gj, njoints, c->nes1 + recs);
        HIP_TRY(c, hipGetLastError());
        return NE_OK;
    };
    if (n == 0) {
        if ((r = unpack_joints()) != NE_OK) return r;
        HIP_TRY(c, hipMemsetAsync(c->prev_n, 0, sizeof(uint32_t), c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return NE_OK;
    }
    uint32_t* flags = c->ufield + (size_t)U_FLAGS * c->ncap;
    HIP_TRY(c, hipMemcpyAsync(c->nes1, d + NES1_HEAD, (size_t)recs, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->counters + CNT_HOSTCHK, 0, sizeof(uint32_t), c->stream));
    k_nes1_check<<<blocks_for(n), TPB, 0, c->stream>>>(flags, n, c->nes1, c->counters + CNT_HOSTCHK);
    uint32_t bad = 0;
    HIP_TRY(c, hipMemcpyAsync(&bad, c->counters + CNT_HOSTCHK, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (bad)
        return set_err(c, NE_ERR_UNSUPPORTED,
                       "ne_gpu_load_state: a static body kept out of the grid would move; upload with ne_gpu_set_bodies");
    k_nes1_unpack<<<blocks_for(n), TPB, 0, c->stream>>>(c->dfield, c->ncap, flags, n, c->nes1);
    HIP_TRY(c, hipGetLastError());
    if ((r = unpack_joints()) != NE_OK) return r;
    HIP_TRY(c, hipMemsetAsync(c->prev_n, 0, sizeof(uint32_t), c->stream));   // resetWarmCache (world.hpp:140)
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NE_OK;
}

int ne_gpu_set_meshes(ne_gpu_ctx* c, uint32_t n, const ne_mesh* ms) {
    if (!c || (n && !ms)) return NE_ERR_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    uint64_t tv = 0, tt = 0;
    for (uint32_t m = 0; m < n; ++m) {
        const ne_mesh& e = ms[m];
        if (e.body >= c->n || c->h_shape[e.body] != NE_SHAPE_MESH)
            return set_err(c, NE_ERR_INVALID, "mesh attached to a body that is not a mesh body");
        if ((e.nverts && !e.verts) || (e.ntris && !e.tris)) return set_err(c, NE_ERR_INVALID, "mesh array missing");
        for (uint64_t k = 0; k < 3ull * e.ntris; ++k)
            if (e.tris[k] < 0 || (uint32_t)e.tris[k] >= e.nverts)
                return set_err(c, NE_ERR_INVALID, "mesh triangle vertex index out of range");
        tv += e.nverts;
        tt += e.ntris;
    }
    if (tv >= (1ull << 30) || tt >= (1ull << 29)) return set_err(c, NE_ERR_INVALID, "meshes too large");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    free_meshes(c);
    if (n == 0) return NE_OK;
    std::vector<double> hv;
    std::vector<int32_t> ht, ho, hroot(n, -1), hof(c->n, -1);
    std::vector<MeshNode> hn;
    hv.reserve(3 * tv);
    ht.reserve(3 * tt);
    ho.reserve(tt);
    c->h_trees.resize(n);
    c->h_tris_n.resize(n);
    for (uint32_t m = 0; m < n; ++m) {
        const ne_mesh& e = ms[m];
        MeshTree& T = c->h_trees[m];
        T.v = e.verts;
        T.t = e.tris;
        T.order.resize(e.ntris);
        for (uint32_t k = 0; k < e.ntris; ++k) T.order[k] = (int32_t)k;
        if (e.ntris) T.build(0, (int32_t)e.ntris);
        T.v = nullptr;
        T.t = nullptr;
        c->h_tris_n[m] = e.ntris;
        const int32_t vbase = (int32_t)(hv.size() / 3), tbase = (int32_t)(ht.size() / 3), nbase = (int32_t)hn.size();
        hv.insert(hv.end(), e.verts, e.verts + 3 * (size_t)e.nverts);
        for (uint64_t k = 0; k < 3ull * e.ntris; ++k) ht.push_back(e.tris[k] + vbase);
        for (int32_t t : T.order) ho.push_back(t + tbase);
        const size_t nn = T.links.size() / 4;
        for (size_t k = 0; k < nn; ++k) {
            MeshNode d;
            for (int a = 0; a < 3; ++a) { d.lo[a] = T.bounds[6 * k + a]; d.hi[a] = T.bounds[6 * k + 3 + a]; }
            const int32_t* l = &T.links[4 * k];
            d.left = l[0] < 0 ? -1 : l[0] + nbase;
            d.right = l[1] < 0 ? -1 : l[1] + nbase;
            d.start = l[0] < 0 ? l[2] + tbase : 0;
            d.count = l[3];
            hn.push_back(d);
        }
        hroot[m] = nn ? nbase : -1;
        hof[e.body] = (int32_t)m;
    }
    HIP_TRY(c, dalloc(c->m_verts, std::max<size_t>(hv.size(), 1)));
    HIP_TRY(c, dalloc(c->m_tris, std::max<size_t>(ht.size(), 1)));
    HIP_TRY(c, dalloc(c->m_order, std::max<size_t>(ho.size(), 1)));
    HIP_TRY(c, dalloc(c->m_nodes, std::max<size_t>(hn.size(), 1)));
    HIP_TRY(c, dalloc(c->m_root, n));
    HIP_TRY(c, dalloc(c->m_of_body, std::max<uint32_t>(c->n, 1u)));
    if (!hv.empty()) HIP_TRY(c, hipMemcpy(c->m_verts, hv.data(), sizeof(double) * hv.size(), hipMemcpyHostToDevice));
    if (!ht.empty()) HIP_TRY(c, hipMemcpy(c->m_tris, ht.data(), sizeof(int32_t) * ht.size(), hipMemcpyHostToDevice));
    if (!ho.empty()) HIP_TRY(c, hipMemcpy(c->m_order, ho.data(), sizeof(int32_t) * ho.size(), hipMemcpyHostToDevice));
    if (!hn.empty()) HIP_TRY(c, hipMemcpy(c->m_nodes, hn.data(), sizeof(MeshNode) * hn.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->m_root, hroot.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    if (c->n) HIP_TRY(c, hipMemcpy(c->m_of_body, hof.data(), sizeof(int32_t) * c->n, hipMemcpyHostToDevice));
    c->nmeshes = n;
    int r = alloc_mesh_pairs(c);
    if (r != NE_OK) return r;
    if (c->mqcap == 0) {
        r = alloc_mesh_cands(c, 8u * c->n + 4096u);
        if (r != NE_OK) return r;
    }
    return NE_OK;
}

int ne_gpu_mesh_bvh(ne_gpu_ctx* c, uint32_t m, int32_t* order, uint32_t node_cap, uint32_t* nnodes, double* bounds,
                    int32_t* links) {
    if (!c) return NE_ERR_INVALID;
    if (m >= c->nmeshes) return set_err(c, NE_ERR_INVALID, "no such mesh");
    const MeshTree& T = c->h_trees[m];
    const uint32_t nn = (uint32_t)(T.links.size() / 4);
    if (nnodes) *nnodes = nn;
    if (order) std::copy(T.order.begin(), T.order.end(), order);
    if (bounds && links && node_cap >= nn) {
        std::copy(T.bounds.begin(), T.bounds.end(), bounds);
        std::copy(T.links.begin(), T.links.end(), links);
    }
    return NE_OK;
}

int ne_gpu_reset_warm_cache(ne_gpu_ctx* c) {
    if (!c) return NE_ERR_INVALID;
    HIP_TRY(c, hipMemsetAsync(c->prev_n, 0, sizeof(uint32_t), c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NE_OK;
}

int ne_gpu_export_warm_cache(ne_gpu_ctx* c, uint32_t cap, uint32_t* count, uint64_t* keys, double* imp) {
    if (!c || !count) return NE_ERR_INVALID;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    uint32_t pn = 0;
    HIP_TRY(c, hipMemcpy(&pn, c->prev_n, sizeof(uint32_t), hipMemcpyDeviceToHost));
    *count = pn;
    const uint32_t m = std::min(pn, cap);
    if (m && keys) HIP_TRY(c, hipMemcpy(keys, c->prev_key, sizeof(uint64_t) * m, hipMemcpyDeviceToHost));
    if (m && imp) HIP_TRY(c, hipMemcpy(imp, c->prev_imp, sizeof(double) * m, hipMemcpyDeviceToHost));
    return NE_OK;
}

int ne_gpu_import_warm_cache(ne_gpu_ctx* c, uint32_t count, const uint64_t* keys, const double* imp) {
    if (!c || (count && (!keys || !imp))) return NE_ERR_INVALID;
    for (uint32_t k = 1; k < count; ++k)
        if (keys[k] < keys[k - 1]) return set_err(c, NE_ERR_INVALID, "warm cache keys must be sorted");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    int r = ensure_prev(c, std::max(count, c->ccap));
    if (r != NE_OK) return r;
    if (count) {
        HIP_TRY(c, hipMemcpy(c->prev_key, keys, sizeof(uint64_t) * count, hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(c->prev_imp, imp, sizeof(double) * count, hipMemcpyHostToDevice));
    }
    HIP_TRY(c, hipMemcpy(c->prev_n, &count, sizeof(uint32_t), hipMemcpyHostToDevice));
    return NE_OK;
}

int ne_gpu_get_contacts(ne_gpu_ctx* c, uint32_t cap, uint32_t* count, uint32_t* ci, uint32_t* cj, double* normal,
                        double* point, double* depth, double* normal_impulse) {
    if (!c || !count) return NE_ERR_INVALID;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    uint32_t nc = 0;
    if (c->n) HIP_TRY(c, hipMemcpy(&nc, c->counters + CNT_C, sizeof(uint32_t), hipMemcpyDeviceToHost));
    *count = nc;
    const uint32_t m = std::min(nc, cap);
    if (m == 0) return NE_OK;
    if (ci) HIP_TRY(c, hipMemcpy(ci, c->C.ci, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
    if (cj) HIP_TRY(c, hipMemcpy(cj, c->C.cj, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
    std::vector<double> tmp(m);
    auto get3 = [&](double* dst, double* a, double* b, double* d) -> int {
        double* srcs[3] = {a, b, d};
        for (int k = 0; k < 3; ++k) {
            HIP_TRY(c, hipMemcpy(tmp.data(), srcs[k], sizeof(double) * m, hipMemcpyDeviceToHost));
            for (uint32_t t = 0; t < m; ++t) dst[3 * t + k] = tmp[t];
        }
        return NE_OK;
    };
    int r;
    if (normal && (r = get3(normal, c->C.n0, c->C.n1, c->C.n2)) != NE_OK) return r;
    if (point && (r = get3(point, c->C.p0, c->C.p1, c->C.p2)) != NE_OK) return r;
    if (depth) HIP_TRY(c, hipMemcpy(depth, c->C.ov, sizeof(double) * m, hipMemcpyDeviceToHost));
    if (normal_impulse) HIP_TRY(c, hipMemcpy(normal_impulse, c->C.nimp, sizeof(double) * m, hipMemcpyDeviceToHost));
    return NE_OK;
}

int ne_gpu_get_pairs(ne_gpu_ctx* c, uint32_t cap, uint32_t* count, uint32_t* pi, uint32_t* pj) {
    if (!c || !count) return NE_ERR_INVALID;
    uint32_t nc = 0;
    int r = ne_gpu_get_contacts(c, 0, &nc, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (r != NE_OK) return r;
    std::vector<uint32_t> a(nc), b(nc);
    if (nc) {
        r = ne_gpu_get_contacts(c, nc, &nc, a.data(), b.data(), nullptr, nullptr, nullptr, nullptr);
        if (r != NE_OK) return r;
    }
    uint32_t m = 0;
    for (uint32_t k = 0; k < nc; ++k) {
        if (k > 0 && a[k] == a[k - 1] && b[k] == b[k - 1]) continue;
        if (m < cap) { if (pi) pi[m] = a[k]; if (pj) pj[m] = b[k]; }
        ++m;
    }
    *count = m;
    return NE_OK;
}

// Index-order sums of src/world.cpp:704-729 over the downloaded state.
int ne_gpu_diagnostics(ne_gpu_ctx* c, double* p, double* L, double* ke) {
    if (!c) return NE_ERR_INVALID;
    const uint32_t n = c->n;
    std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n), q(4 * (size_t)n), w(3 * (size_t)n), im(n), ii(3 * (size_t)n);
    int r = ne_gpu_get_state(c, x.data(), v.data(), q.data(), w.data());
    if (r != NE_OK) return r;
    r = ne_gpu_get_body_aux(c, nullptr, nullptr, im.data(), ii.data());
    if (r != NE_OK) return r;
    V3 P{0.0, 0.0, 0.0}, LL{0.0, 0.0, 0.0};
    double e = 0.0;
    for (uint32_t i = 0; i < n; ++i)
        if (im[i] > 0.0) P = P + V3{v[3 * i], v[3 * i + 1], v[3 * i + 2]} * (1.0 / im[i]);
    for (uint32_t i = 0; i < n; ++i) {
        if (im[i] <= 0.0) continue;
        const double m = 1.0 / im[i];
        const V3 xi{x[3 * i], x[3 * i + 1], x[3 * i + 2]}, vi{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
        const V3 wi{w[3 * i], w[3 * i + 1], w[3 * i + 2]}, iii{ii[3 * i], ii[3 * i + 1], ii[3 * i + 2]};
        const Qt qi{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
        LL = LL + cross(xi, vi * m);
        LL = LL + angular_momentum(qi, iii, wi);
    }
    for (uint32_t i = 0; i < n; ++i) {
        if (im[i] <= 0.0) continue;
        const V3 vi{v[3 * i], v[3 * i + 1], v[3 * i + 2]}, wi{w[3 * i], w[3 * i + 1], w[3 * i + 2]};
        const V3 iii{ii[3 * i], ii[3 * i + 1], ii[3 * i + 2]};
        const Qt qi{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
        double lin = 0.5 * (im[i] > 0 ? 1.0 / im[i] : 0.0) * norm2(vi);
        V3 Ib = inertia_of(iii);
        V3 wb = qinvrot(qi, wi);
        V3 Lb{wb.x * Ib.x, wb.y * Ib.y, wb.z * Ib.z};
        double rot = 0.5 * dot(wb, Lb);
        e += lin + rot;
    }
    if (p) { p[0] = P.x; p[1] = P.y; p[2] = P.z; }
    if (L) { L[0] = LL.x; L[1] = LL.y; L[2] = LL.z; }
    if (ke) *ke = e;
    return NE_OK;
}

int ne_gpu_get_counters(ne_gpu_ctx* c, uint32_t out[8]) {
    if (!c || !out) return NE_ERR_INVALID;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    uint32_t h[CNT_N];
    HIP_TRY(c, hipMemcpy(h, c->counters, sizeof(h), hipMemcpyDeviceToHost));
    out[0] = h[CNT_PRAW];
    out[1] = c->nmeshes ? h[CNT_MQRAW] : 0;
    out[2] = h[CNT_CRAW];
    out[3] = h[CNT_ISL];
    out[4] = h[CNT_MAXISL];
    out[5] = c->ngrid;
    out[6] = c->sgrid_cells ? c->sgrid_cells : h[CNT_CELLS];
    out[7] = h[CNT_FAULT];
    return NE_OK;
}

int ne_gpu_get_stream(ne_gpu_ctx* c, void** stream) {
    if (!c || !stream) return NE_ERR_INVALID;
    *stream = (void*)c->stream;
    return NE_OK;
}

const char* ne_gpu_last_error(ne_gpu_ctx* c) {
    if (!c) return "null context";
    return c->err.c_str();
}

int ne_gpu_time_kernel(ne_gpu_ctx* c, int32_t kernel_id) {
    if (!c || kernel_id < 0 || kernel_id > NE_KERNEL_RAY) return NE_ERR_INVALID;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->ev_used = 0;
    c->timed_kernel = kernel_id;
    c->timed_ms = 0.0;
    c->timed_launches = 0;
    return NE_OK;
}

int ne_gpu_kernel_time(ne_gpu_ctx* c, double* total_ms, uint32_t* launches) {
    if (!c) return NE_ERR_INVALID;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    harvest_timing(c);
    if (total_ms) *total_ms = c->timed_ms;
    if (launches) *launches = c->timed_launches;
    return NE_OK;
}


int ne_gpu_set_worlds(ne_gpu_ctx* c, uint32_t nworlds, const uint32_t* offsets, const double* half,
                      const double* restitution, const double* friction) {
    if (!c) return NE_ERR_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    if (nworlds == 0) {
        c->nworlds = 0;
        c->grid_dirty = true;
        return NE_OK;
    }
    if (!offsets || !half) return set_err(c, NE_ERR_INVALID, "offsets and half are required");
    if (!c->params.periodic) return set_err(c, NE_ERR_UNSUPPORTED, "batched worlds must be periodic");
    if (c->nccd) return set_err(c, NE_ERR_UNSUPPORTED, "CCD bodies in batched worlds");
    if (c->ngj) return set_err(c, NE_ERR_UNSUPPORTED, "general joints in batched worlds");
    if (offsets[0] != 0 || offsets[nworlds] != c->n)
        return set_err(c, NE_ERR_INVALID, "world offsets must run from 0 to the body count");
    uint32_t maxsize = 0;
    for (uint32_t w = 0; w < nworlds; ++w) {
        if (offsets[w + 1] < offsets[w]) return set_err(c, NE_ERR_INVALID, "world offsets must not decrease");
        maxsize = std::max(maxsize, offsets[w + 1] - offsets[w]);
        for (int a = 0; a < 3; ++a)
            if (!(half[3 * w + a] > 0.0)) return set_err(c, NE_ERR_INVALID, "world half extents must be > 0");
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->w_off.assign(offsets, offsets + nworlds + 1);
    c->w_half.assign(half, half + 3 * (size_t)nworlds);
    c->w_rest.assign(nworlds, c->params.restitution);
    c->w_fric.assign(nworlds, c->params.friction);
    for (uint32_t w = 0; w < nworlds; ++w) {
        if (restitution) c->w_rest[w] = restitution[w];
        if (friction) c->w_fric[w] = friction[w];
    }
    c->w_maxsize = maxsize;
    std::vector<uint32_t> wid(c->n);
    for (uint32_t w = 0; w < nworlds; ++w)
        for (uint32_t i = offsets[w]; i < offsets[w + 1]; ++i) wid[i] = w;
    HIP_TRY(c, dalloc(c->d_wid, std::max<uint32_t>(c->n, 1u)));
    if (c->n) HIP_TRY(c, hipMemcpy(c->d_wid, wid.data(), sizeof(uint32_t) * c->n, hipMemcpyHostToDevice));
    c->nworlds = nworlds;
    c->grid_dirty = true;
    drop_graphs(c);
    return prepare_static_grid(c);
}

int ne_gpu_epa_stats(ne_gpu_ctx* c, uint32_t* reruns, uint32_t* truncated) {
    if (!c) return NE_ERR_INVALID;
    if (reruns) *reruns = c->h_counters[CNT_EPARERUN];
    if (truncated) *truncated = c->h_counters[CNT_EPATRUNC];
    return NE_OK;
}

int ne_gpu_epa_queues(ne_gpu_ctx* c, uint32_t out[8]) {
    if (!c || !out) return NE_ERR_INVALID;
    const uint32_t* h = c->h_counters;
    const int slot[8] = {CNT_EPAQ, CNT_EPALONG, CNT_EPALONG3, CNT_EPALONG2, CNT_EPABIG, CNT_MQEPA, CNT_MQLONG, CNT_MQLONG2};
    for (int k = 0; k < 8; ++k) out[k] = h[slot[k]];
    return NE_OK;
}

int ne_gpu_use_graph(ne_gpu_ctx* c, int32_t enable) {
    if (!c) return NE_ERR_INVALID;
    c->use_graph = enable != 0;
    if (!c->use_graph) drop_graphs(c);
    return NE_OK;
}

int ne_gpu_autotune(ne_gpu_ctx* c, int32_t steps, ne_tune_result* out) {
    if (!c || steps <= 0) return NE_ERR_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, NE_ERR_HIP, "hipSetDevice failed");
    if (c->n == 0) return set_err(c, NE_ERR_INVALID, "no bodies to tune on");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // snapshot of everything a step changes: body fields, warm cache, flags
    const size_t nd = (size_t)F_COUNT * c->ncap;
    double* snap = nullptr;
    uint64_t* skey = nullptr;
    double* simp = nullptr;
    uint32_t scap = c->prev_cap, spn = 0;
    HIP_TRY(c, hipMalloc((void**)&snap, sizeof(double) * nd));
    if (hipMalloc((void**)&skey, sizeof(uint64_t) * std::max(scap, 1u)) != hipSuccess ||
        hipMalloc((void**)&simp, sizeof(double) * std::max(scap, 1u)) != hipSuccess) {
        (void)hipFree(snap); if (skey) (void)hipFree(skey);
        return set_err(c, NE_ERR_HIP, "autotune snapshot allocation failed");
    }
    auto cleanup = [&]() { (void)hipFree(snap); (void)hipFree(skey); (void)hipFree(simp); };
    HIP_TRY(c, hipMemcpy(snap, c->dfield, sizeof(double) * nd, hipMemcpyDeviceToDevice));
    HIP_TRY(c, hipMemcpy(skey, c->prev_key, sizeof(uint64_t) * scap, hipMemcpyDeviceToDevice));
    HIP_TRY(c, hipMemcpy(simp, c->prev_imp, sizeof(double) * scap, hipMemcpyDeviceToDevice));
    HIP_TRY(c, hipMemcpy(&spn, c->prev_n, sizeof(uint32_t), hipMemcpyDeviceToHost));
    const bool forces = c->has_forces;
    auto restore = [&]() -> int {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipMemcpy(c->dfield, snap, sizeof(double) * nd, hipMemcpyDeviceToDevice));
        const uint32_t m = std::min(scap, c->prev_cap);
        HIP_TRY(c, hipMemcpy(c->prev_key, skey, sizeof(uint64_t) * m, hipMemcpyDeviceToDevice));
        HIP_TRY(c, hipMemcpy(c->prev_imp, simp, sizeof(double) * m, hipMemcpyDeviceToDevice));
        HIP_TRY(c, hipMemcpy(c->prev_n, &spn, sizeof(uint32_t), hipMemcpyHostToDevice));
        c->has_forces = forces;
        return NE_OK;
    };
    const double scales[3] = {1.0, 1.25, 1.5};
    const int32_t walks[2] = {NE_BROADPHASE_WALK, NE_BROADPHASE_CELLS};
    double best = 1e300;
    ne_tune_result res{};
    uint32_t trials = 0;
    for (double sc : scales)
        for (int32_t bp : walks)
            for (int g = 0; g < 2; ++g) {
                c->cell_scale = sc;
                c->cell = (2.0 * c->grid_maxR * (1.0 + 1e-6) + 1e-9) * sc;
                c->broadphase = bp;
                c->use_graph = g != 0;
                c->grid_dirty = true;
                drop_graphs(c);
                double ms = 1e300;
                for (int attempt = 0; attempt < 3; ++attempt) {
                    int r = restore();
                    if (r != NE_OK) { cleanup(); return r; }
                    r = ne_gpu_step(c, 1);   // warm-up: builds graphs, sizes launches
                    if (r == NE_WARN_EPA_CAPPED) r = NE_OK;   // a trial step: only its time is used
                    if (r == NE_ERR_CAPACITY) continue;
                    if (r != NE_OK) { cleanup(); return r; }
                    r = restore();
                    if (r != NE_OK) { cleanup(); return r; }
                    const auto t0 = std::chrono::steady_clock::now();
                    r = ne_gpu_step(c, steps);
                    const auto t1 = std::chrono::steady_clock::now();
                    if (r == NE_WARN_EPA_CAPPED) r = NE_OK;
                    if (r == NE_ERR_CAPACITY) continue;
                    if (r != NE_OK) { cleanup(); return r; }
                    ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
                    break;
                }
                ++trials;
                if (ms < best) {
                    best = ms;
                    res.cell_scale = sc;
                    res.broadphase = bp;
                    res.graph = g;
                    res.ms_per_step = ms;
                }
            }
    res.trials = trials;
    c->cell_scale = res.cell_scale;
    c->cell = (2.0 * c->grid_maxR * (1.0 + 1e-6) + 1e-9) * res.cell_scale;
    c->broadphase = res.broadphase;
    c->use_graph = res.graph != 0;
    c->grid_dirty = true;
    drop_graphs(c);
    int r = restore();
    cleanup();
    if (r != NE_OK) return r;
    if (out) *out = res;
    return best < 1e300 ? NE_OK : set_err(c, NE_ERR_CAPACITY, "every autotune trial overflowed");
}

int ne_gpu_fp_probe(ne_gpu_ctx* c, int32_t op, uint32_t n, const double* a, const double* b, double* out) {
    if (!c || !a || !out || (op == 1 && !b)) return NE_ERR_INVALID;
    if (n == 0) return NE_OK;
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    HIP_TRY(c, hipMalloc((void**)&da, sizeof(double) * n));
    HIP_TRY(c, hipMalloc((void**)&db, sizeof(double) * n));
    HIP_TRY(c, hipMalloc((void**)&dout, sizeof(double) * n));
    HIP_TRY(c, hipMemcpy(da, a, sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(db, b ? b : a, sizeof(double) * n, hipMemcpyHostToDevice));
    k_fp_probe<<<blocks_for(n), TPB, 0, c->stream>>>(op, n, da, db, dout);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost));
    (void)hipFree(da); (void)hipFree(db); (void)hipFree(dout);
    return NE_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Upload: AoS-ish host arrays -> device SoA, plus the static broadphase plan
// (bounding radii, grid-excluded bodies, cell size).
// ---------------------------------------------------------------------------
namespace {

int upload_bodies(ne_gpu_ctx* c, const ne_bodies* b) {
    const uint32_t n = b->n;
    if (n > 0 && (!b->x || !b->v || !b->q || !b->w || !b->inv_mass || !b->inv_inertia || !b->shape || !b->radius ||
                  !b->half_height || !b->half_extents))
        return set_err(c, NE_ERR_INVALID, "required body array missing");
    if (n >= (1u << 30)) return set_err(c, NE_ERR_INVALID, "too many bodies");
    for (uint32_t i = 0; i < n; ++i) {
        const int s = b->shape[i];
        if (s < 0 || s > NE_SHAPE_MESH)
            return set_err(c, NE_ERR_UNSUPPORTED, "compound shapes are not on the GPU path");
        if (s == NE_SHAPE_MESH && (b->inv_mass[i] != 0.0 || b->inv_inertia[3 * i] != 0.0 ||
                                   b->inv_inertia[3 * i + 1] != 0.0 || b->inv_inertia[3 * i + 2] != 0.0))
            return set_err(c, NE_ERR_UNSUPPORTED, "mesh bodies must be static (zero inverse mass and inertia)");
        const uint32_t fl = b->flags ? b->flags[i] : NE_FLAG_DYNAMIC;
        (void)fl;
        if (s == NE_SHAPE_CONVEX && (!b->hull_offsets || b->hull_offsets[i + 1] < b->hull_offsets[i]))
            return set_err(c, NE_ERR_INVALID, "convex body without hull vertices");
    }
    // (re)allocate per-body storage
    if (n > c->ncap || !c->dfield) {
        free_work(c);
        uint32_t ncap = std::max<uint32_t>(std::max<uint32_t>(n, 64u), c->ncap + c->ncap / 2u);
        c->ncap = ncap;
        HIP_TRY(c, dalloc(c->dfield, (size_t)F_COUNT * ncap));
        HIP_TRY(c, dalloc(c->ufield, (size_t)U_COUNT * ncap));
        HIP_TRY(c, dalloc(c->hull_off, (size_t)ncap + 1));
        HIP_TRY(c, dalloc(c->nc_off, (size_t)ncap + 1));
        HIP_TRY(c, dalloc(c->gridded, ncap));
        HIP_TRY(c, dalloc(c->aabb_part, 6 * AABB_BLOCKS));
        HIP_TRY(c, dalloc(c->grid, 1));
        HIP_TRY(c, dalloc(c->cell_of, ncap));
        HIP_TRY(c, dalloc(c->order_prev, ncap));
        HIP_TRY(c, dalloc(c->xp, ncap));
        HIP_TRY(c, dalloc(c->sorted_idx, ncap));
        HIP_TRY(c, dalloc(c->sorted, ncap));
        HIP_TRY(c, dalloc(c->row_count, (size_t)ncap + 1));
        HIP_TRY(c, dalloc(c->row_start, (size_t)ncap + 1));
        HIP_TRY(c, dalloc(c->parent, ncap));
        HIP_TRY(c, dalloc(c->used, (size_t)ncap + 1));
        HIP_TRY(c, dalloc(c->root_id, (size_t)ncap + 1));
        HIP_TRY(c, dalloc(c->slot_of, ncap));
        HIP_TRY(c, dalloc(c->claimed, ncap));
        HIP_TRY(c, dalloc(c->wake_at, ncap));
        if (!c->wake_again) HIP_TRY(c, dalloc(c->wake_again, WAKE_ROUNDS + 1));
        HIP_TRY(c, dalloc(c->mint, ncap));
        HIP_TRY(c, dalloc(c->row_scratch, (size_t)ncap * ROW_SLOTS));
        c->cellcap = (uint32_t)std::min<uint64_t>((uint64_t)c->cells_per_body * ncap + 4096u, 0x7FFFFFFFull);
        HIP_TRY(c, dalloc(c->cell_count, (size_t)c->cellcap + 1));
        HIP_TRY(c, dalloc(c->cell_start, (size_t)c->cellcap + 1));
        HIP_TRY(c, hipMemset(c->cell_count, 0, sizeof(uint32_t) * ((size_t)c->cellcap + 1)));
        uint32_t pcap = c->user_pcap ? c->user_pcap : 8u * ncap + 4096u;
        uint32_t ccap = c->user_ccap ? c->user_ccap : 4u * ncap + 4096u;
        int r = alloc_pairs(c, pcap);
        if (r != NE_OK) return r;
        r = alloc_cons(c, ccap);
        if (r != NE_OK) return r;
        r = ensure_prev(c, ccap);
        if (r != NE_OK) return r;
        r = alloc_epa_big(c, EPA_BIG_PAIRS, std::min(EPA_BIG_FACES, c->epa_faces_max));
        if (r != NE_OK) return r;
        // scan scratch: partial sums (<=1024) and radix histograms
        HIP_TRY(c, dalloc(c->scan_part, 1024));
        HIP_TRY(c, dalloc(c->rs_hist, (size_t)RS_BINS * 1024 + 1));
    }
    c->n = n;
    const uint32_t ncap = c->ncap;
    // --- doubles
    std::vector<double> hd((size_t)F_COUNT * ncap, 0.0);
    auto D = [&](int f, uint32_t i) -> double& { return hd[(size_t)f * ncap + i]; };
    bool forces = false;
    std::vector<uint32_t> hoff(n + 1, 0), noff(n + 1, 0);
    uint32_t nh = 0, nn = 0;
    for (uint32_t i = 0; i < n; ++i) {
        hoff[i] = nh;
        if (b->hull_offsets) nh += b->hull_offsets[i + 1] - b->hull_offsets[i];
        noff[i] = nn;
        if (b->nocollide_offsets) nn += b->nocollide_offsets[i + 1] - b->nocollide_offsets[i];
    }
    hoff[n] = nh;
    noff[n] = nn;
    std::vector<double> hhull(3 * (size_t)std::max(nh, 1u), 0.0);
    std::vector<uint32_t> hnc(std::max(nn, 1u), 0u);
    for (uint32_t i = 0; i < n; ++i) {
        if (b->hull_offsets)
            for (uint32_t k = b->hull_offsets[i], t = hoff[i]; k < b->hull_offsets[i + 1]; ++k, ++t)
                for (int a = 0; a < 3; ++a) hhull[3 * (size_t)t + a] = b->hull_verts[3 * (size_t)k + a];
        if (b->nocollide_offsets)
            for (uint32_t k = b->nocollide_offsets[i], t = noff[i]; k < b->nocollide_offsets[i + 1]; ++k, ++t)
                hnc[t] = b->nocollide[k];
    }
    double maxDyn = 0.0;
    std::vector<double> R(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            D(F_X0 + a, i) = b->x[3 * i + a];
            D(F_V0 + a, i) = b->v[3 * i + a];
            D(F_W0 + a, i) = b->w[3 * i + a];
            D(F_F0 + a, i) = b->force ? b->force[3 * i + a] : 0.0;
            D(F_T0 + a, i) = b->torque ? b->torque[3 * i + a] : 0.0;
            D(F_II0 + a, i) = b->inv_inertia[3 * i + a];
            D(F_HE0 + a, i) = b->half_extents[3 * i + a];
            forces = forces || D(F_F0 + a, i) != 0.0 || D(F_T0 + a, i) != 0.0 || std::signbit(D(F_F0 + a, i)) ||
                     std::signbit(D(F_T0 + a, i));
        }
        for (int a = 0; a < 4; ++a) D(F_Q0 + a, i) = b->q[4 * i + a];
        D(F_IM, i) = b->inv_mass[i];
        D(F_RAD, i) = b->radius[i];
        D(F_HH, i) = b->half_height[i];
        D(F_FRIC, i) = b->friction ? b->friction[i] : -1.0;
        D(F_REST, i) = b->restitution ? b->restitution[i] : -1.0;
        D(F_LD, i) = b->linear_damping ? b->linear_damping[i] : 0.0;
        D(F_AD, i) = b->angular_damping ? b->angular_damping[i] : 0.0;
        D(F_GS, i) = b->gravity_scale ? b->gravity_scale[i] : 1.0;
        D(F_ST, i) = b->sleep_timer ? b->sleep_timer[i] : 0.0;
        D(F_IMS, i) = b->inv_mass_store ? b->inv_mass_store[i] : 0.0;
        for (int a = 0; a < 3; ++a) D(F_IIS0 + a, i) = b->inv_inertia_store ? b->inv_inertia_store[3 * i + a] : 0.0;
        R[i] = host_bounding_radius(b->shape[i], b->radius[i], b->half_height[i], b->half_extents + 3 * i,
                                    hhull.data(), hoff[i], hoff[i + 1]);
        D(F_BR, i) = R[i];
        if (b->shape[i] != NE_SHAPE_PLANE && b->inv_mass[i] > 0.0) maxDyn = std::max(maxDyn, R[i]);
    }
    // grid-excluded: planes, and static bodies much larger than any dynamic one
    std::vector<uint32_t> gridded, excl;
    double maxR = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        bool ex = b->shape[i] == NE_SHAPE_PLANE || R[i] >= 1e29 || (maxDyn > 0.0 && R[i] > 2.0 * maxDyn);
        if (ex) excl.push_back(i);
        else { gridded.push_back(i); maxR = std::max(maxR, R[i]); }
    }
    c->ngrid = (uint32_t)gridded.size();
    c->nexcl = (uint32_t)excl.size();
    c->grid_maxR = maxR;
    c->cell = (2.0 * maxR * (1.0 + 1e-6) + 1e-9) * c->cell_scale;
    c->h_R = R;
    c->h_excl = excl;
    c->h_isexcl.assign(n, 0);
    for (uint32_t e : excl) c->h_isexcl[e] = 1;
    c->nworlds = 0;   // a new body set starts as one world
    free_meshes(c);   // and without meshes
    c->h_shape.assign(b->shape, b->shape + n);
    {   // CCD bodies (World::ccdPass), in index order
        std::vector<uint32_t> ccd;
        for (uint32_t i = 0; i < n; ++i)
            if ((b->flags ? b->flags[i] : NE_FLAG_DYNAMIC) & NE_FLAG_CCD) ccd.push_back(i);
        const uint64_t ntiles = (n + RAY_THREADS - 1) / RAY_THREADS;
        if ((uint64_t)ccd.size() * ntiles >= (1ull << 31))
            return set_err(c, NE_ERR_UNSUPPORTED, "too many CCD bodies for this body count");
        c->nccd = (uint32_t)ccd.size();
        if (c->nccd) {
            c->ccd_cap = std::max<uint32_t>(4096u, 256u * c->nccd);
            HIP_TRY(c, dalloc(c->ccd_idx, c->nccd));
            HIP_TRY(c, dalloc(c->ccd_prev, 3 * (size_t)c->nccd));
            HIP_TRY(c, dalloc(c->ccd_cnt, (size_t)c->nccd * ntiles));
            HIP_TRY(c, dalloc(c->ccd_off, (size_t)c->nccd * ntiles + 1));
            HIP_TRY(c, dalloc(c->ccd_cand, c->ccd_cap));
            HIP_TRY(c, dalloc(c->ccd_tiles, 6 * ntiles));
            HIP_TRY(c, hipMemcpy(c->ccd_idx, ccd.data(), sizeof(uint32_t) * ccd.size(), hipMemcpyHostToDevice));
        }
    }
    c->njoints = c->nspring = c->nrigid = c->ncomp = 0;   // and without joints
    c->ngj = c->ngj_comp = 0;
    c->h_mut.assign(n, 0);
    for (uint32_t i = 0; i < n; ++i)
        c->h_mut[i] = b->inv_mass[i] != 0.0 || b->inv_inertia[3 * i] != 0.0 || b->inv_inertia[3 * i + 1] != 0.0 ||
                      b->inv_inertia[3 * i + 2] != 0.0 ||
                      ((b->flags ? b->flags[i] : NE_FLAG_DYNAMIC) & NE_FLAG_SLEEPING);   // may wake
    c->grid_dirty = true;
    drop_graphs(c);
    c->has_forces = forces;
    c->spheres_only = true;
    for (uint32_t i = 0; i < n; ++i) c->spheres_only = c->spheres_only && b->shape[i] == NE_SHAPE_SPHERE;
    // island-key bits
    int bits = 1;
    while (bits < 32 && (1ull << bits) <= (uint64_t)n) ++bits;
    c->uf_bits = bits;
    std::vector<uint32_t> hu((size_t)U_COUNT * ncap, 0u);
    for (uint32_t i = 0; i < n; ++i) {
        hu[(size_t)U_SHAPE * ncap + i] = (uint32_t)b->shape[i];
        hu[(size_t)U_LAYER * ncap + i] = b->layer ? b->layer[i] : 0u;
        hu[(size_t)U_MASK * ncap + i] = b->mask ? b->mask[i] : 0xFFFFFFFFu;
        hu[(size_t)U_FLAGS * ncap + i] = (b->flags ? b->flags[i] : NE_FLAG_DYNAMIC) & ~FLAG_GRID_EXCLUDED;
    }
    for (uint32_t e : excl) hu[(size_t)U_FLAGS * ncap + e] |= FLAG_GRID_EXCLUDED;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(c->dfield, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->ufield, hu.data(), sizeof(uint32_t) * hu.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, dalloc(c->hull, 3 * (size_t)std::max(nh, 1u)));
    HIP_TRY(c, dalloc(c->nc_idx, std::max(nn, 1u)));
    HIP_TRY(c, hipMemcpy(c->hull, hhull.data(), sizeof(double) * hhull.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->nc_idx, hnc.data(), sizeof(uint32_t) * hnc.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->hull_off, hoff.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->nc_off, noff.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice));
    if (!gridded.empty()) {
        HIP_TRY(c, hipMemcpy(c->gridded, gridded.data(), sizeof(uint32_t) * gridded.size(), hipMemcpyHostToDevice));
        // until a step has sorted them, the cell order is the index order
        HIP_TRY(c, hipMemcpy(c->sorted_idx, gridded.data(), sizeof(uint32_t) * gridded.size(), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(c->order_prev, gridded.data(), sizeof(uint32_t) * gridded.size(), hipMemcpyHostToDevice));
    }
    HIP_TRY(c, hipMemset(c->slot_of, 0xFF, sizeof(uint32_t) * ncap));
    HIP_TRY(c, dalloc(c->excl, std::max<size_t>(excl.size(), 1)));
    if (!excl.empty())
        HIP_TRY(c, hipMemcpy(c->excl, excl.data(), sizeof(uint32_t) * excl.size(), hipMemcpyHostToDevice));
    // device view
    BodiesDev& B = c->B;
    double* d = c->dfield;
    auto F = [&](int f) { return d + (size_t)f * ncap; };
    B.x0 = F(F_X0); B.x1 = F(F_X1); B.x2 = F(F_X2);
    B.v0 = F(F_V0); B.v1 = F(F_V1); B.v2 = F(F_V2);
    B.q0 = F(F_Q0); B.q1 = F(F_Q1); B.q2 = F(F_Q2); B.q3 = F(F_Q3);
    B.w0 = F(F_W0); B.w1 = F(F_W1); B.w2 = F(F_W2);
    B.f0 = F(F_F0); B.f1 = F(F_F1); B.f2 = F(F_F2);
    B.t0 = F(F_T0); B.t1 = F(F_T1); B.t2 = F(F_T2);
    B.im = F(F_IM); B.ii0 = F(F_II0); B.ii1 = F(F_II1); B.ii2 = F(F_II2);
    B.rad = F(F_RAD); B.hh = F(F_HH); B.he0 = F(F_HE0); B.he1 = F(F_HE1); B.he2 = F(F_HE2); B.bR = F(F_BR);
    B.fric = F(F_FRIC); B.rest = F(F_REST); B.ldamp = F(F_LD); B.adamp = F(F_AD); B.gs = F(F_GS);
    B.st = F(F_ST); B.ims = F(F_IMS); B.iis0 = F(F_IIS0); B.iis1 = F(F_IIS1); B.iis2 = F(F_IIS2);
    B.sleep_on = c->params.sleep_enabled ? 1 : 0;
    B.shape = (int32_t*)(c->ufield + (size_t)U_SHAPE * ncap);
    B.layer = c->ufield + (size_t)U_LAYER * ncap;
    B.mask = c->ufield + (size_t)U_MASK * ncap;
    B.flags = c->ufield + (size_t)U_FLAGS * ncap;
    B.hull_off = c->hull_off; B.hull = c->hull; B.nc_off = c->nc_off; B.nc_idx = c->nc_idx;
    B.n = n;
    B.wid = nullptr;
    B.W = nullptr;
    return NE_OK;
}

}  // namespace

// ne_ccd.hpp -- continuous collision on the device: World::ccdPass
// (src/world.cpp:658-696), between integrate and collide.
//
// The reference takes the CCD bodies in index order; each one sweeps its
// bounding sphere from its step-start position along this step's
// displacement, folding rayVsBody over every other body in index order with
// a shrinking search distance (so, as for rays, the fold's order is part of
// the result), stops at the first impact and loses its velocity into the
// surface; later CCD bodies see the earlier ones' stopped positions.
//
// Here: (A) in parallel, for every (CCD body, tile of bodies) the bodies
// that are not CCD bodies and that the sweep hits within its full length --
// the only ones the fold can accept, since a hit within a shorter distance
// is a hit within the full one -- counted, scanned and written in body
// order; (B) one thread runs the reference's fold per CCD body over the
// merge, in body order, of its candidates and the other CCD bodies (whose
// positions (B) itself may have moved), applies the stop, and moves the
// stopped body's packed pose and grid cell.  A candidate list that outgrew
// its buffer makes (B) fold over every body instead.
//
// Included by ne_gpu.hip after the query kernels (ray_vs_body, tiles).
#pragma once

#include "ne_query.hpp"

namespace nea {

// The CCD bodies' step-start positions (world.cpp:135).
__global__ void __launch_bounds__(TPB) k_ccd_save(BodiesDev B, const uint32_t* __restrict__ idx, uint32_t nccd,
                                                  double* __restrict__ prev) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nccd) return;
    const uint32_t i = idx[k];
    prev[3 * (size_t)k] = B.x0[i];
    prev[3 * (size_t)k + 1] = B.x1[i];
    prev[3 * (size_t)k + 2] = B.x2[i];
}

// CCD body k's sweep: origin, unit direction, length and bounding radius, or
// false when it does not sweep this step (world.cpp:666-672).
struct Sweep {
    V3 o, d;
    double len, R;
    uint32_t i;
};

__device__ __forceinline__ bool ccd_sweep(const BodiesDev& B, const uint32_t* __restrict__ idx,
                                          const double* __restrict__ prev, uint32_t k, Sweep& s) {
    const uint32_t i = idx[k];
    s.i = i;
    if (B.im[i] <= 0.0 || (B.flags[i] & NE_FLAG_SLEEPING)) return false;
    s.o = V3{prev[3 * (size_t)k], prev[3 * (size_t)k + 1], prev[3 * (size_t)k + 2]};
    const V3 disp = ld3(B.x0, B.x1, B.x2, i) - s.o;
    s.len = norm(disp);
    s.R = B.bR[i];
    if (s.len < s.R * 0.5 || s.len < 1e-9) return false;
    s.d = disp / s.len;
    return true;
}

// Whether body j can stop sweep s within distance maxT, and where: the
// reference's filters (not the sweeping body, no sensor, layers) and
// rayVsBody against j inflated by the sweep radius at its minimum image
// (world.cpp:675-687).
__device__ __forceinline__ bool ccd_test(const BodiesDev& B, const Box3& box, const Sweep& s, uint32_t j, double maxT,
                                         double& t, V3& n) {
    if (j == s.i) return false;
    const uint32_t fj = B.flags[j];
    if (fj & NE_FLAG_SENSOR) return false;
    const uint32_t li = B.layer[s.i], lj = B.layer[j];
    if (!((B.mask[s.i] & (1u << (lj & 31u))) && (B.mask[j] & (1u << (li & 31u))))) return false;
    const int sh = B.shape[j];
    if (!ray_shape(sh)) return false;   // rayVsBody never hits hulls, planes, meshes
    const double R = s.R;
    const V3 x = ld3(B.x0, B.x1, B.x2, j);
    const V3 cimg = s.o + min_image(box, x - s.o);
    {   // padded bounding-sphere skip, grown by the inflation
        const double reach = skip_radius(B, j) + 2.0 * R;
        const V3 mo = cimg - s.o;
        const double tc = fmin(fmax(dot(mo, s.d), 0.0), maxT);
        const V3 cp = mo - s.d * tc;
        const double pad = 1e-6 * (1.0 + fabs(s.o.x) + fabs(s.o.y) + fabs(s.o.z) + maxT + reach);
        if (norm2(cp) > (reach + pad) * (reach + pad)) return false;
    }
    const double r = B.rad[j] + R;
    double h = B.hh[j];
    V3 he = ld3(B.he0, B.he1, B.he2, j);
    if (sh == SH_BOX) he = he + V3{R, R, R};
    else if (sh == SH_CYLINDER) h = h + R;
    V3 p;
    return ray_vs_body(sh, ldq(B, j), he, r, h, cimg, s.o, s.d, maxT, t, p, n);
}

// (A) EMIT false: per (CCD body k = blockIdx.y + gridDim.y m, tile
// blockIdx.x) the count of non-CCD bodies the full sweep hits; EMIT true:
// those bodies in body order at off[cell] + rank (while below cap).
template <bool EMIT, bool CULL>
__global__ void __launch_bounds__(RAY_THREADS) k_ccd_cand(BodiesDev B, Box3 box, const uint32_t* __restrict__ idx,
                                                          uint32_t nccd, const double* __restrict__ prev,
                                                          const double* __restrict__ tiles, uint32_t* __restrict__ cnt,
                                                          const uint32_t* __restrict__ off, uint32_t* __restrict__ cand,
                                                          uint32_t cap) {
    __shared__ uint32_t lds[PRIM_WAVES + 1];
    const uint32_t j = blockIdx.x * RAY_THREADS + threadIdx.x;
    for (uint32_t k = blockIdx.y; k < nccd; k += gridDim.y) {
        Sweep s;
        bool hit = false;
        if (ccd_sweep(B, idx, prev, k, s) &&
            (!CULL || ray_meets_tile(s.o, s.d, s.len, 2.0 * s.R, tiles + 6 * (size_t)blockIdx.x)) && j < B.n &&
            !(B.flags[j] & NE_FLAG_CCD)) {
            double t;
            V3 n;
            hit = ccd_test(B, box, s, j, s.len, t, n);
        }
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(hit ? 1u : 0u, &tot, lds);
        const size_t cell = (size_t)k * gridDim.x + blockIdx.x;
        if (!EMIT) {
            if (threadIdx.x == 0) cnt[cell] = tot;
        } else if (hit) {
            const uint32_t at = off[cell] + ex;
            if (at < cap) cand[at] = j;
        }
    }
}

// (B) The fold, one thread, CCD bodies in index order (world.cpp:664-696).
__global__ void k_ccd_fold(BodiesDev B, Box3 box, const uint32_t* __restrict__ idx, uint32_t nccd,
                           const double* __restrict__ prev, uint32_t ntiles, const uint32_t* __restrict__ off,
                           const uint32_t* __restrict__ cand, uint32_t cap, SortedBody* __restrict__ xp, int bin,
                           StaticGrid G, uint32_t* __restrict__ cell_count) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (uint32_t k = 0; k < nccd; ++k) {
        Sweep s;
        if (!ccd_sweep(B, idx, prev, k, s)) continue;
        double bestT = s.len;
        V3 bestN{0.0, 0.0, 0.0};
        bool hit = false;
        const uint32_t c0 = off[(size_t)k * ntiles], c1 = off[(size_t)(k + 1) * ntiles];
        auto visit = [&](uint32_t j) {
            double t;
            V3 n;
            if (ccd_test(B, box, s, j, bestT, t, n) && t < bestT) {
                bestT = t;
                bestN = n;
                hit = true;
            }
        };
        if (c1 > cap) {   // the candidate buffer overflowed: every body, in order
            for (uint32_t j = 0; j < B.n; ++j) visit(j);
        } else {          // merge the candidates and the CCD bodies, both ascending
            uint32_t a = c0, b = 0;
            while (a < c1 || b < nccd) {
                const uint32_t ja = a < c1 ? cand[a] : 0xFFFFFFFFu, jb = b < nccd ? idx[b] : 0xFFFFFFFFu;
                if (ja < jb) { visit(ja); ++a; }
                else { visit(jb); ++b; }
            }
        }
        if (!hit) continue;
        const uint32_t i = s.i;
        const V3 x_old = ld3(B.x0, B.x1, B.x2, i);
        const V3 x = s.o + s.d * bestT;
        st3(B.x0, B.x1, B.x2, i, x);
        V3 v = ld3(B.v0, B.v1, B.v2, i);
        const double vn = dot(v, bestN);
        if (vn < 0.0) st3(B.v0, B.v1, B.v2, i, v - bestN * vn);
        xp[i] = SortedBody{x.x, x.y, x.z, B.bR[i]};
        if (bin && !(B.flags[i] & FLAG_GRID_EXCLUDED)) {   // integrate binned the old position
            cell_count[static_cell(G, x_old)] -= 1u;
            cell_count[static_cell(G, x)] += 1u;
        }
    }
}

}  // namespace nea

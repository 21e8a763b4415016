// ne_epa_lds.hpp -- the first EPA tiers: one thread per queued convex pair,
// its polytope in LDS (src/epa.hpp:28-150, the same steps as epa_contact,
// then narrow_pair's box-box manifold, src/manifold.hpp:90-126).
//
// Most EPA runs are short: on the terrain scene about 2.5 iterations, 9 in
// 10 within 4 iterations and 12 faces.  A per-thread polytope in scratch
// memory is laid out with each lane's values 256 bytes apart, so when the
// lanes of a wave read different faces or vertices -- the closest face, the
// vertex a visibility test needs, a compacted or new face -- every lane
// touches a cache line of its own.  Here each wave keeps its 64 polytopes in
// LDS as [slot][lane] arrays: a lane only ever touches its own column, so
// any mix of slots across the lanes is free of bank conflicts.  A wave's
// workspace is 40 KiB, so four one-wave blocks (one per SIMD, which the
// kernel's register use allows anyway) fill a CU's 160 KiB.
//
// The support directions of the vertices stay in (scratch) registers: all
// active lanes add their n-th vertex in the same iteration, so those writes
// are uniform, and only the final face's three are read back, to recompute
// its witness points (the same support calls, so the same bits).
//
// This is tier A (8 vertices, i.e. 4 iterations, 12 faces, 8 horizon
// entries; four waves a CU), over every queued pair.  A run that outgrows it
// goes on to the group tiers (k_epa_group, below: B with 4 lanes a run and
// 20 iterations, C with 8 lanes and 32), then to the one-wave tier
// (k_epa_wave) and the one-block tier (k_narrow_big); each restarts it from
// its GJK tetrahedron, so every tier gives the same bits.  (Tier B as a
// per-lane LDS tier of 16 vertices and 24 faces, two waves a CU, took 50.1
// ms per 1M terrain step against 48.1 with groups of 4; tier A in groups of
// 4 took 50.7.)
//
// Included by ne_gpu.hip inside its anonymous namespace, after k_narrow and
// the mesh helpers (load_geom, box_of, wave_add1, CNT_*, MeshDev, mesh_pair,
// mesh_prism, store_contact).
#pragma once

// [slot][lane] workspace of one wave: MV vertices, MF faces, MH horizon entries.
template <int MV, int MF, int MH>
struct EpaLaneLds {
    double vx[MV][64], vy[MV][64], vz[MV][64];   // Minkowski points
    double fnx[MF][64], fny[MF][64], fnz[MF][64], fd[MF][64];
    uint32_t fabc[MF][64];                       // a | b << 8 | c << 16
    uint16_t hor[MH][64];                        // u | w << 8
};
// Tier A: 4 iterations, 40 KiB a wave (four waves fill a CU's 160 KiB).
static_assert(sizeof(EpaLaneLds<8, 12, 8>) == 40960, "tier A: four waves' workspaces fill a CU's LDS");

template <int MV, int MF, int MH>
struct EpaLane {
    EpaLaneLds<MV, MF, MH>& W;
    int l;
    __device__ V3 vert(int k) const { return V3{W.vx[k][l], W.vy[k][l], W.vz[k][l]}; }
    __device__ void set_vert(int k, V3 v) { W.vx[k][l] = v.x; W.vy[k][l] = v.y; W.vz[k][l] = v.z; }
    // make_face (src/epa.hpp:76-83) into face slot `slot`
    __device__ void make_face(int i, int j, int k, int slot) {
        const V3 vi = vert(i);
        V3 n = cross(vert(j) - vi, vert(k) - vi);
        const double len = norm(n);
        n = (len > 1e-18) ? n / len : V3{0.0, 0.0, 1.0};
        double dist = dot(n, vi);
        if (dist < 0) { n = -n; dist = -dist; const int t = j; j = k; k = t; }
        W.fnx[slot][l] = n.x; W.fny[slot][l] = n.y; W.fnz[slot][l] = n.z; W.fd[slot][l] = dist;
        W.fabc[slot][l] = (uint32_t)i | ((uint32_t)j << 8) | ((uint32_t)k << 16);
    }
    __device__ void copy_face(int from, int to) {
        W.fnx[to][l] = W.fnx[from][l]; W.fny[to][l] = W.fny[from][l]; W.fnz[to][l] = W.fnz[from][l];
        W.fd[to][l] = W.fd[from][l]; W.fabc[to][l] = W.fabc[from][l];
    }
    // the reference's horizon list: an edge whose reverse is listed erases the
    // first such entry, else it is appended (false: the list is full)
    __device__ bool hadd(uint32_t u, uint32_t w, int& nh) {
        const uint16_t rev = (uint16_t)(w | (u << 8));
        for (int e = 0; e < nh; ++e) {
            if (W.hor[e][l] == rev) {
                for (int m = e; m + 1 < nh; ++m) W.hor[m][l] = W.hor[m + 1][l];
                --nh;
                return true;
            }
        }
        if (nh >= MH) return false;
        W.hor[nh][l] = (uint16_t)(u | (w << 8));
        ++nh;
        return true;
    }
};

// EPA (src/epa.hpp:71-150) of one run in the lane's LDS column, from the
// tetrahedron GJK returned.  Returns false when the run outgrows the
// workspace (the caller queues it for a larger tier); else `c` is the
// contact, with the witness points of the final face recomputed from its
// vertices' directions.
template <int MV, int MF, int MH>
__device__ inline bool epa_lane(EpaLane<MV, MF, MH>& E, const Geom& a, const Geom& b, V3 shift, const SupD tet[4],
                                V3* dirs, Contact& c) {
    for (int k = 0; k < 4; ++k) { E.set_vert(k, tet[k].v); dirs[k] = tet[k].d; }
    E.make_face(0, 1, 2, 0);
    E.make_face(0, 2, 3, 1);
    E.make_face(0, 3, 1, 2);
    E.make_face(1, 3, 2, 3);
    int nv = 4, nf = 4;
    V3 bn{0.0, 0.0, 0.0};
    double bdist = 0.0;
    uint32_t babc = 0;
    for (int iter = 0; iter < 64; ++iter) {
        int ci = 0;
        double cd = E.W.fd[0][E.l];
        for (int f = 1; f < nf; ++f) {
            const double d = E.W.fd[f][E.l];
            if (d < cd) { cd = d; ci = f; }
        }
        bn = V3{E.W.fnx[ci][E.l], E.W.fny[ci][E.l], E.W.fnz[ci][E.l]};
        bdist = cd;
        babc = E.W.fabc[ci][E.l];
        const V3 pv = support_world(a, bn) - (support_world(b, -bn) + shift);
        if (dot(pv, bn) - bdist < 1e-7) break;
        if (nv == MV) return false;   // the next tier runs it past this budget
        const int pidx = nv;
        E.set_vert(pidx, pv);
        dirs[pidx] = bn;
        ++nv;
        int nh = 0, nk = 0;
        bool over = false;
        for (int fi = 0; fi < nf && !over; ++fi) {
            const uint32_t abc = E.W.fabc[fi][E.l];
            const V3 fn{E.W.fnx[fi][E.l], E.W.fny[fi][E.l], E.W.fnz[fi][E.l]};
            const uint32_t fa = abc & 0xffu, fb = (abc >> 8) & 0xffu, fc = (abc >> 16) & 0xffu;
            if (dot(fn, pv - E.vert((int)fa)) > 1e-12) {
                over = !E.hadd(fa, fb, nh) || !E.hadd(fb, fc, nh) || !E.hadd(fc, fa, nh);
            } else {
                if (nk != fi) E.copy_face(fi, nk);   // nk <= fi: in-place compaction keeps order
                ++nk;
            }
        }
        if (over || nk + nh > MF) return false;
        nf = nk;
        for (int e = 0; e < nh; ++e) {
            const uint32_t h = E.W.hor[e][E.l];
            E.make_face((int)(h & 0xffu), (int)(h >> 8), pidx, nf++);
        }
        if (nf == 0) break;
    }
    // barycentric witness points of the closest face (src/epa.hpp:123-149)
    const int ia = (int)(babc & 0xffu), ib = (int)((babc >> 8) & 0xffu), ic = (int)((babc >> 16) & 0xffu);
    SupV tri[3];
    tri[0] = witness(a, b, shift, E.vert(ia), dirs[ia]);
    tri[1] = witness(a, b, shift, E.vert(ib), dirs[ib]);
    tri[2] = witness(a, b, shift, E.vert(ic), dirs[ic]);
    Face bf;
    bf.n = bn; bf.dist = bdist; bf.a = 0; bf.b = 1; bf.c = 2;
    c = epa_result(tri, bf);
    return true;
}

// The contacts of a convex pair from its EPA contact (narrow_pair's convex
// branch): the box-box manifold when both are boxes, else the one contact.
__device__ __forceinline__ uint32_t convex_pair_contacts(const Box3& box, const Geom& a, const Geom& b,
                                                         const SimParams& P, const Contact& c, double* out,
                                                         uint32_t* fault) {
    uint32_t nc = 0;
    if (a.shape == SH_BOX && b.shape == SH_BOX && !P.no_manifold) {
        V3 pts[4];
        double dep[4];
        const int m = box_box_manifold(box, a, b, c.normal, pts, dep, fault);
        for (int k = 0; k < m; ++k) {
            double* o = out + 7 * k;
            o[0] = c.normal.x; o[1] = c.normal.y; o[2] = c.normal.z;
            o[3] = pts[k].x; o[4] = pts[k].y; o[5] = pts[k].z; o[6] = dep[k];
        }
        nc = (uint32_t)m;
    }
    if (nc == 0) {
        out[0] = c.normal.x; out[1] = c.normal.y; out[2] = c.normal.z;
        out[3] = c.point.x; out[4] = c.point.y; out[5] = c.point.z; out[6] = c.overlap;
        nc = 1;
    }
    return nc;
}

// One-wave blocks, one queued pair per lane, grid-stride over the queue
// `epa_q` of overlapping convex pairs (its length in counters[count_slot]);
// each starts from its GJK tetrahedron in `tets` (k_narrow's GJK).  Runs
// past the workspace go to `long_q` (counted in counters[long_slot]).
template <int MV, int MF, int MH>
__global__ void __launch_bounds__(64) k_epa_lds(BodiesDev B, SimParams P, uint32_t* __restrict__ counters,
                                                const uint32_t* __restrict__ cand_i,
                                                const uint32_t* __restrict__ cand_j, uint32_t* __restrict__ ncont,
                                                double* __restrict__ ct, const uint32_t* __restrict__ epa_q,
                                                int count_slot, uint32_t epa_cap, uint32_t* __restrict__ long_q,
                                                int long_slot, EpaTets tets) {
    __shared__ EpaLaneLds<MV, MF, MH> lds;
    EpaLane<MV, MF, MH> E{lds, (int)threadIdx.x};
    const uint32_t nq = min(counters[count_slot], epa_cap);
    uint32_t* fault = counters + CNT_FAULT;
    V3 dirs[MV];
    for (uint32_t t0 = blockIdx.x * 64u; t0 < nq; t0 += gridDim.x * 64u) {
        const uint32_t t = t0 + threadIdx.x;
        bool defer = false;
        const uint32_t p = t < nq ? epa_q[t] : 0u;
        if (t < nq) {
            const uint32_t i = cand_i[p], j = cand_j[p];
            const Box3& box = box_of(B, P, i);
            const Geom a = load_geom(B, i), b = load_geom(B, j);
            const V3 shift = min_image_shift(box, a.x, b.x);
            uint32_t nc = 0;
            SupD tet[4];
            tets.get(p, tet);
            Contact c;
            if (epa_lane(E, a, b, shift, tet, dirs, c))
                nc = convex_pair_contacts(box, a, b, P, c, ct + (size_t)p * 28, fault);
            else
                defer = true;
            ncont[p] = nc;   // 0 while deferred: the next tier writes it
        }
        const uint32_t at = wave_add1(counters, long_slot, defer);
        if (defer) long_q[at] = p;   // at most one entry per queued pair: fits the queue's capacity
    }
}

// ---- Tiers B and C: G lanes per run, for the long runs ---------------------
//
// Runs that outgrow tier A are few but long (curved against flat supports --
// sphere or capsule on a box -- refine their closest face for tens of
// iterations over tens of faces).  A group of G lanes shares one run:
//   - the first closest face is a min/first-index reduction over the group;
//   - visibility is tested G faces at a time, and the kept faces are
//     compacted in order by ballot ranks;
//   - the horizon is built in parallel: an occurrence of the directed edge
//     (u, w) among the visible faces' edges survives when (u, w) occurs P
//     times and (w, u) N times with P > N, and at least N occurrences of
//     (u, w) come before it -- exactly the list the reference's find + erase
//     leaves, in its order (each arriving reverse edge erases the oldest
//     entry, so the last P - N occurrences survive);
//   - vertices, faces and the staged edges sit in the group's LDS; the
//     support directions stay with the group's first lane.
// Runs past MV - 4 iterations or MF faces go on to the next tier.

template <int MV, int MF>
struct EpaGroupLds {
    double vx[MV], vy[MV], vz[MV];
    double fnx[MF], fny[MF], fnz[MF], fd[MF];
    double prism[18];            // a mesh candidate's triangle prism (six vertices)
    uint32_t fabc[MF];           // a | b << 8 | c << 16
    uint16_t edge[3 * MF];       // u | w << 8: the visible faces' edges of one iteration
};

// Where the group tier's runs come from when they are mesh triangle
// candidates (k_mesh_lds's leftovers) rather than pairs: the mesh, the
// candidates' pairs and triangles, and their hit flags and contacts.
struct MeshSrc {
    MeshDev M;
    const uint32_t* mq_pair;
    const int32_t* mq_tri;
    uint8_t* mhit;
    double* mct;
};

// LDS written by some lanes of a wave and read by others after this point
// (one-wave blocks: a wave's LDS operations complete in order, so ordering
// the compiler's memory operations is all that is needed).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int MV, int MF>
__device__ __forceinline__ V3 eg_vert(const EpaGroupLds<MV, MF>& L, int k) {
    return V3{L.vx[k], L.vy[k], L.vz[k]};
}

// make_face (src/epa.hpp:76-83) into face slot `slot`.
template <int MV, int MF>
__device__ __forceinline__ void eg_make_face(EpaGroupLds<MV, MF>& L, int i, int j, int k, int slot) {
    const V3 vi = eg_vert(L, i);
    V3 n = cross(eg_vert(L, j) - vi, eg_vert(L, k) - vi);
    const double len = norm(n);
    n = (len > 1e-18) ? n / len : V3{0.0, 0.0, 1.0};
    double dist = dot(n, vi);
    if (dist < 0) { n = -n; dist = -dist; const int t = j; j = k; k = t; }
    L.fnx[slot] = n.x; L.fny[slot] = n.y; L.fnz[slot] = n.z; L.fd[slot] = dist;
    L.fabc[slot] = (uint32_t)i | ((uint32_t)j << 8) | ((uint32_t)k << 16);
}

// One-wave blocks; 64 / G runs per block at a time, grid-stride over the
// queue `epa_q` (length in counters[count_slot]); runs past the workspace go
// to `long_q` (counted in counters[long_slot]).  MESH: the queue holds mesh
// triangle candidates (detectContact(other, prism), contacts to src.mct and
// src.mhit), else pairs (contacts to ct and ncont).
template <int G, int MV, int MF, bool MESH>
__global__ void __launch_bounds__(64) k_epa_group(BodiesDev B, SimParams P, uint32_t* __restrict__ counters,
                                                  const uint32_t* __restrict__ cand_i,
                                                  const uint32_t* __restrict__ cand_j, uint32_t* __restrict__ ncont,
                                                  double* __restrict__ ct, const uint32_t* __restrict__ queue,
                                                  int count_slot, uint32_t epa_cap, uint32_t* __restrict__ long_q,
                                                  int long_slot, MeshSrc src, EpaTets tets) {
    static_assert(G >= 4 && 64 % G == 0, "a group holds the initial tetrahedron's four faces");
    constexpr int NG = 64 / G;
    __shared__ EpaGroupLds<MV, MF> lds[NG];
    const int lane = (int)threadIdx.x, g = lane / G, gl = lane % G;
    const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << (g * G));
    const uint64_t below = lanemask_lt() & gmask;
    EpaGroupLds<MV, MF>& L = lds[g];
    const uint32_t nq = min(counters[count_slot], epa_cap);
    uint32_t* fault = counters + CNT_FAULT;
    V3 dirs[MV];   // the group's first lane keeps the support directions
    for (uint32_t t0 = blockIdx.x * NG; t0 < nq; t0 += gridDim.x * NG) {
        const uint32_t t = t0 + g;
        bool active = t < nq, defer = false, hit = false, flip = false;
        const uint32_t p = active ? queue[t] : 0u;     // a pair (MESH: a triangle candidate)
        Geom a{}, b{};
        V3 shift{0.0, 0.0, 0.0};
        const Box3* box = &P.box;
        if (active) {
            if constexpr (MESH) {
                uint32_t o, m;
                const uint32_t pr = src.mq_pair[p];
                mesh_pair(B, cand_i[pr], cand_j[pr], o, m, flip);
                box = &box_of(B, P, o);
                a = load_geom(B, o);
                mesh_prism(B, src.M, m, src.mq_tri[p], L.prism, b);   // non-degenerate: its EPA ran; every lane writes the same bits
                wave_lds_sync();
            } else {
                const uint32_t i = cand_i[p], j = cand_j[p];
                box = &box_of(B, P, i);
                a = load_geom(B, i);
                b = load_geom(B, j);
            }
            shift = min_image_shift(*box, a.x, b.x);
            SupD tet[4];
            if constexpr (MESH) {
                hit = gjk_tetra(a, b, shift, tet);   // k_mesh_lds's GJK on the same inputs: always true here
            } else {
                tets.get(p, tet);                    // k_narrow's GJK
                hit = true;
            }
            active = hit;
            if (active && gl == 0) {
                for (int k = 0; k < 4; ++k) {
                    L.vx[k] = tet[k].v.x; L.vy[k] = tet[k].v.y; L.vz[k] = tet[k].v.z;
                    dirs[k] = tet[k].d;
                }
            }
        }
        wave_lds_sync();
        if (active && gl < 4) {
            const int fa = gl == 3 ? 1 : 0;
            const int fb = gl == 0 ? 1 : (gl == 1 ? 2 : 3);
            const int fc = gl == 0 ? 2 : (gl == 1 ? 3 : (gl == 2 ? 1 : 2));
            eg_make_face(L, fa, fb, fc, gl);   // {0,1,2}, {0,2,3}, {0,3,1}, {1,3,2}
        }
        wave_lds_sync();
        int nv = 4, nf = 4;
        V3 bn{0.0, 0.0, 0.0};
        double bdist = 0.0;
        uint32_t babc = 0;
        for (int iter = 0; iter < 64; ++iter) {
            if (__ballot(active) == 0ull) break;
            V3 pv{0.0, 0.0, 0.0};
            if (active) {
                // first face of minimal distance
                double bd = 0.0;
                int bi = -1;
                for (int f = gl; f < nf; f += G) {
                    const double d = L.fd[f];
                    if (bi < 0 || d < bd) { bd = d; bi = f; }
                }
#pragma unroll
                for (int off = G / 2; off > 0; off >>= 1) {
                    const double od = __shfl_xor(bd, off);
                    const int oi = __shfl_xor(bi, off);
                    const bool take = oi >= 0 && (bi < 0 || od < bd || (od == bd && oi < bi));
                    bd = take ? od : bd;
                    bi = take ? oi : bi;
                }
                bn = V3{L.fnx[bi], L.fny[bi], L.fnz[bi]};
                bdist = L.fd[bi];
                babc = L.fabc[bi];
                pv = support_world(a, bn) - (support_world(b, -bn) + shift);
                if (dot(pv, bn) - bdist < 1e-7) active = false;           // converged
                else if (nv == MV) { defer = true; active = false; }      // past this tier's budget
            }
            if (active) {
                const int pidx = nv;
                if (gl == 0) {
                    L.vx[pidx] = pv.x; L.vy[pidx] = pv.y; L.vz[pidx] = pv.z;
                    dirs[pidx] = bn;
                }
                ++nv;
                wave_lds_sync();
                // visibility; kept faces compacted in order, visible faces' edges staged in order
                int nk = 0, ne = 0;
                for (int base = 0; base < nf; base += G) {
                    const int f = base + gl;
                    const bool in = f < nf;
                    V3 fn{0.0, 0.0, 0.0};
                    double fdist = 0.0;
                    uint32_t abc = 0;
                    bool vis = false;
                    if (in) {
                        fn = V3{L.fnx[f], L.fny[f], L.fnz[f]};
                        fdist = L.fd[f];
                        abc = L.fabc[f];
                        vis = dot(fn, pv - eg_vert(L, (int)(abc & 0xffu))) > 1e-12;
                    }
                    const uint64_t km = __ballot(in && !vis) & gmask, vm = __ballot(vis) & gmask;
                    wave_lds_sync();   // the chunk is read before any slot at or below it is written
                    if (in && !vis) {
                        const int s = nk + (int)__popcll(km & below);
                        L.fnx[s] = fn.x; L.fny[s] = fn.y; L.fnz[s] = fn.z; L.fd[s] = fdist; L.fabc[s] = abc;
                    }
                    if (vis) {
                        const int e = ne + 3 * (int)__popcll(vm & below);
                        const uint32_t ia = abc & 0xffu, ib = (abc >> 8) & 0xffu, ic = (abc >> 16) & 0xffu;
                        L.edge[e] = (uint16_t)(ia | (ib << 8));
                        L.edge[e + 1] = (uint16_t)(ib | (ic << 8));
                        L.edge[e + 2] = (uint16_t)(ic | (ia << 8));
                    }
                    nk += (int)__popcll(km);
                    ne += 3 * (int)__popcll(vm);
                }
                wave_lds_sync();
                // horizon survivors in order -> new faces after the kept ones
                int ns = 0;
                for (int base = 0; base < ne; base += G) {
                    const int k = base + gl;
                    bool surv = false;
                    uint32_t e = 0;
                    if (k < ne) {
                        e = L.edge[k];
                        const uint32_t rev = (e >> 8) | ((e & 0xffu) << 8);
                        int np = 0, nn = 0, before = 0;
                        for (int m = 0; m < ne; ++m) {
                            const uint32_t o = L.edge[m];
                            np += o == e;
                            nn += o == rev;
                            before += (o == e) & (m < k);
                        }
                        surv = np > nn && before >= nn;
                    }
                    const uint64_t sm = __ballot(surv) & gmask;
                    const int slot = nk + ns + (int)__popcll(sm & below);
                    if (surv && slot < MF) eg_make_face(L, (int)(e & 0xffu), (int)(e >> 8), pidx, slot);
                    ns += (int)__popcll(sm);
                }
                if (nk + ns > MF) { defer = true; active = false; }   // the faces outgrow this tier
                else {
                    nf = nk + ns;
                    if (nf == 0) active = false;
                }
                wave_lds_sync();
            }
        }
        // contacts: the group's first lane (narrow_pair's convex branch after
        // EPA; a mesh candidate's one contact, meshContacts in world.cpp:56-83)
        if (t < nq && gl == 0 && !defer) {
            uint32_t nc = 0;
            Contact c = no_contact();
            if (hit) {
                const int ia = (int)(babc & 0xffu), ib = (int)((babc >> 8) & 0xffu), ic = (int)((babc >> 16) & 0xffu);
                SupV tri[3];
                tri[0] = witness(a, b, shift, eg_vert(L, ia), dirs[ia]);
                tri[1] = witness(a, b, shift, eg_vert(L, ib), dirs[ib]);
                tri[2] = witness(a, b, shift, eg_vert(L, ic), dirs[ic]);
                Face bf;
                bf.n = bn; bf.dist = bdist; bf.a = 0; bf.b = 1; bf.c = 2;
                c = epa_result(tri, bf);
                if constexpr (!MESH) nc = convex_pair_contacts(*box, a, b, P, c, ct + (size_t)p * 28, fault);
            }
            if constexpr (MESH) {
                if (c.hit) store_contact(src.mct + 7 * (size_t)p, c, flip);
                src.mhit[p] = c.hit ? 1 : 0;
            } else {
                ncont[p] = nc;
            }
        }
        const bool push = t < nq && gl == 0 && defer;
        if (push) {   // the next tier writes the result
            if constexpr (MESH) src.mhit[p] = 0;
            else ncont[p] = 0u;
        }
        const uint32_t at = wave_add1(counters, long_slot, push);
        if (push) long_q[at] = p;   // at most one entry per queued run: fits the queue's capacity
        wave_lds_sync();   // this round's LDS reads are done before the next runs' writes
    }
}

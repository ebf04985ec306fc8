// avbd_kernels.hip -- HIP kernels of the device-resident AVBD step (gfx950).
//
// Step pipeline (Solver::Step, solver.cpp:199-627), all scenes at once:
//   k_tile_bounds, k_broad, k_pair_filter -> sort -> k_pair_append
//     -> k_pair_finish                                                 (phase 1)
//   k_manifold_init, k_force_init -> compaction                       (phase 2)
//   incidence lists, greedy colouring, colour lists
//   k_predict                                                          (phase 4)
//   iterations x { per colour: k_primal ; k_dual ; k_fracture_reset }  (phase 5)
//   k_velocity                                                         (phase 6)
//   per colour: k_primal(alpha = 0) ; k_stick                          (phase 7)
#include <hip/hip_runtime.h>

#include "avbd_collide.h"
#include "avbd_fem.h"
#include "avbd_math.h"
#include "avbd_world.h"

#define GRID_STRIDE(i, n) for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += blockDim.x * gridDim.x)

AVBD_DEV v3 f3(float4 a) { return v3{a.x, a.y, a.z}; }
AVBD_DEV qt fq(float4 a) { return qt{a.x, a.y, a.z, a.w}; }
AVBD_DEV float4 v4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
AVBD_DEV float4 q4(qt q) { return make_float4(q.w, q.x, q.y, q.z); }

AVBD_DEV bool scene_runs(const DevWorld& w, int s) {
    const int* st = w.st + s * ST_WORDS;
    return st[ST_ACTIVE] != 0 && st[ST_EMERGENCY] == 0;
}
AVBD_DEV bool is_static(const DevWorld& w, int b) { return (w.flags[b] & AVBD_BODY_STATIC) != 0; }
// The projection a scene's energies use in this sweep: ADAPTIVE resolves by
// the scene's trust-region ratio, CLAMP when |rho - 1| <= 0.01, else
// ABSOLUTE (ProjectEigenvalues, math.cpp:152-160).
AVBD_DEV int proj_mode(const DevWorld& w, int s, int mode) {
    if (mode != AVBD_PROJ_ADAPTIVE) return mode;
    return fabsf(w.rho[s] - 1.f) > 0.01f ? AVBD_PROJ_ABSOLUTE : AVBD_PROJ_CLAMP;
}
// The scene's phases 4-7 run in one workgroup of k_scene_step instead of the
// per-colour bucket launches.
AVBD_DEV bool scene_fused(const DevWorld& w, int s) { return w.st[s * ST_WORDS + ST_FUSED] != 0; }

__global__ void k_fill_int(int* p, int n, int v) {
    GRID_STRIDE(i, n) p[i] = v;
}
__global__ void k_set_word(int* p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Scene run flags of a step: every scene (all = 1), or the n listed scenes
// (the others cleared first by a launch with n = 0, all = 0).
__global__ void k_set_active(DevWorld w, const int* ids, int n, int all) {
    if (n == 0) {
        GRID_STRIDE(s, w.S) w.st[s * ST_WORDS + ST_ACTIVE] = all;
    } else {
        GRID_STRIDE(i, n) w.st[ids[i] * ST_WORDS + ST_ACTIVE] = 1;
    }
}

// ------------------------------------------------------------------ phase 1
// Is there already a force linking bodies i and j?  (isConstrainedTo,
// force.hpp:81-89): scan the shorter incidence list.
AVBD_DEV bool constrained(const DevWorld& w, int i, int j) {
    int li = w.f_off[i + 1] - w.f_off[i];
    int lj = w.f_off[j + 1] - w.f_off[j];
    int self = li <= lj ? i : j;
    int other = li <= lj ? j : i;
    for (int k = w.f_off[self]; k < w.f_off[self + 1]; ++k) {
        uint32_t e = w.f_ent[k];
        uint32_t idx = ENT_INDEX(e);
        int o;
        if (ENT_KIND(e) == ENT_CONTACT) {
            int m = idx >> 3;
            o = ENT_SIDE_B(e) ? w.m.a[m] : w.m.b[m];
        } else if (ENT_KIND(e) == ENT_JOINT) {
            o = ENT_SIDE_B(e) ? w.j_a[idx] : w.j_b[idx];
        } else {
            o = ENT_SIDE_B(e) ? w.s_a[idx] : w.s_b[idx];
        }
        if (o == other) return true;
    }
    return false;
}

// Broad phase: the bounding-sphere test of every body pair of a scene
// (solver.cpp:236-255), tiled.  A scene's bodies form tiles of 256 in index
// order, each of 8 chunks of 32; one work item = (scene, i-tile, j-tile >=
// i-tile), the j-tile staged in LDS.  k_tile_bounds gives every chunk and
// tile the box around its bodies' spheres, widened by 1e-5 (|x| + r) per
// side (far more than the rounding of the bounds).  An item whose tile
// boxes are disjoint is skipped, and inside an item a wave skips the j
// chunks disjoint from the box of its 64 i bodies: no pair skipped can pass
// the exact test, so the pair set equals that of the full loop.
// Overlapping pairs are gathered per block in LDS and go to a candidate list
// (pairs_tmp) with one global atomic per block; k_pair_filter then drops
// those already linked by a force, one thread per candidate, and appends the
// new pairs unordered (creation order is restored by rank later); cnt_i
// counts per lower index.
#define BROAD_HIT_CAP 1024

AVBD_DEV bool boxes_disjoint(float4 alo, float4 ahi, float4 blo, float4 bhi) {
    return ahi.x < blo.x || bhi.x < alo.x || ahi.y < blo.y || bhi.y < alo.y || ahi.z < blo.z || bhi.z < alo.z;
}

// tiles == nullptr: block b bounds tile b; else tile tiles[b] (the tiles of
// the scenes that run).
__global__ __launch_bounds__(256) void k_tile_bounds(DevWorld w, const int* tiles) {
    const int t = tiles ? tiles[blockIdx.x] : (int)blockIdx.x;
    const int s = w.tile_scene[t];
    const int i = w.sc_off[s] + (t - w.tile_off[s]) * 256 + threadIdx.x;
    const float inf = __builtin_huge_valf();
    float lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    if (scene_runs(w, s) && i < w.sc_off[s + 1]) {
        const float4 p = w.pos[i];
        const float r = w.scale[i].w;
        const float c[3] = {p.x, p.y, p.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float m = 1e-5f * (fabsf(c[a]) + r);
            lo[a] = (c[a] - r) - m;
            hi[a] = (c[a] + r) + m;
        }
    }
    // chunks of 32 lanes
#pragma unroll
    for (int off = 16; off > 0; off >>= 1)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], __shfl_xor(lo[a], off));
            hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off));
        }
    __shared__ float red[8][6];
    const int chunk = threadIdx.x / 32;
    if ((threadIdx.x & 31) == 0) {
        w.chunk_lo[t * 8 + chunk] = make_float4(lo[0], lo[1], lo[2], 0.f);
        w.chunk_hi[t * 8 + chunk] = make_float4(hi[0], hi[1], hi[2], 0.f);
        for (int a = 0; a < 3; ++a) {
            red[chunk][a] = lo[a];
            red[chunk][3 + a] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 8; ++k)
            for (int a = 0; a < 3; ++a) {
                lo[a] = fminf(lo[a], red[k][a]);
                hi[a] = fmaxf(hi[a], red[k][3 + a]);
            }
        w.tile_lo[t] = make_float4(lo[0], lo[1], lo[2], 0.f);
        w.tile_hi[t] = make_float4(hi[0], hi[1], hi[2], 0.f);
    }
}

AVBD_DEV void broad_append_global(const DevWorld& w, int2 pr) {
    const int slot = atomicAdd(&w.g[G_CANDIDATES], 1);
    if (slot < w.pair_cap) w.pairs_tmp[slot] = pr;
    else atomicExch(&w.g[G_ERR_PAIR_CAP], 1);
}

__global__ __launch_bounds__(256) void k_broad(DevWorld w, const int4* items) {
    const int4 it = items[blockIdx.x];
    const int s = it.x;
    if (!scene_runs(w, s)) return;
    const int ti = w.tile_off[s] + it.y, tj = w.tile_off[s] + it.z;
    // the same decision for the whole block, before its barriers
    if (it.y != it.z && boxes_disjoint(w.tile_lo[ti], w.tile_hi[ti], w.tile_lo[tj], w.tile_hi[tj])) return;
    const int off = w.sc_off[s];
    const int end = w.sc_off[s + 1];
    __shared__ float4 tile[256];
    __shared__ float4 jlo[8], jhi[8];
    __shared__ int2 hits[BROAD_HIT_CAP];
    __shared__ int n_hits, base;
    const int j0 = off + it.z * 256;
    const int jl = j0 + threadIdx.x;
    if (jl < end) {
        float4 p = w.pos[jl];
        tile[threadIdx.x] = make_float4(p.x, p.y, p.z, w.scale[jl].w);
    }
    if (threadIdx.x < 8) {
        jlo[threadIdx.x] = w.chunk_lo[tj * 8 + threadIdx.x];
        jhi[threadIdx.x] = w.chunk_hi[tj * 8 + threadIdx.x];
    }
    if (threadIdx.x == 0) n_hits = 0;
    __syncthreads();
    const int wave = threadIdx.x / 64;
    const int i = off + it.y * 256 + threadIdx.x;
    const int jn = min(256, end - j0);
    if (off + it.y * 256 + wave * 64 < end) {  // the wave has bodies
        const bool valid = i < end;
        const float4 pi4 = valid ? w.pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const v3 pi = f3(pi4);
        const float ri = valid ? w.scale[i].w : 0.f;
        const float4 c0l = w.chunk_lo[ti * 8 + 2 * wave], c1l = w.chunk_lo[ti * 8 + 2 * wave + 1];
        const float4 c0h = w.chunk_hi[ti * 8 + 2 * wave], c1h = w.chunk_hi[ti * 8 + 2 * wave + 1];
        const float4 ilo = make_float4(fminf(c0l.x, c1l.x), fminf(c0l.y, c1l.y), fminf(c0l.z, c1l.z), 0.f);
        const float4 ihi = make_float4(fmaxf(c0h.x, c1h.x), fmaxf(c0h.y, c1h.y), fmaxf(c0h.z, c1h.z), 0.f);
        for (int c = 0; c < 8 && c * 32 < jn; ++c) {
            if (boxes_disjoint(ilo, ihi, jlo[c], jhi[c])) continue;
            const int t1 = min(c * 32 + 32, jn);
            for (int t = c * 32; t < t1; ++t) {
                const int j = j0 + t;
                const float4 pj = tile[t];
                v3 d = sub(pi, f3(pj));
                float d2 = dot(d, d);
                float rs = ri + pj.w;
                if (!valid || j <= i || d2 > (rs * rs) * 1.00001f) continue;  // conservative pre-test
                if (sqrtf(d2) < rs) {
                    const int k = atomicAdd(&n_hits, 1);
                    if (k < BROAD_HIT_CAP) hits[k] = make_int2(i, j);
                    else broad_append_global(w, make_int2(i, j));
                }
            }
        }
    }
    __syncthreads();
    const int nh = min(n_hits, BROAD_HIT_CAP);
    if (nh == 0) return;
    if (threadIdx.x == 0) base = atomicAdd(&w.g[G_CANDIDATES], nh);
    __syncthreads();
    for (int k = threadIdx.x; k < nh; k += 256) {
        const int slot = base + k;
        if (slot < w.pair_cap) w.pairs_tmp[slot] = hits[k];
        else atomicExch(&w.g[G_ERR_PAIR_CAP], 1);
    }
}

// Keep the candidates not yet linked by a force (isConstrainedTo) as new
// pairs, written as sort keys (i << 32 | j) into the pair buffer.
__global__ void k_pair_filter(DevWorld w) {
    const int C = min(w.g[G_CANDIDATES], w.pair_cap);
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(w.pairs);
    GRID_STRIDE(p, C) {
        const int2 pr = w.pairs_tmp[p];
        if (constrained(w, pr.x, pr.y)) continue;
        int slot = atomicAdd(&w.g[G_NEW_PAIRS], 1);
        if (slot < w.pair_cap) keys[slot] = ((unsigned long long)(unsigned)pr.x << 32) | (unsigned)pr.y;
        else atomicExch(&w.g[G_ERR_PAIR_CAP], 1);
    }
}

// The new-pair keys in ascending order, in place, in one workgroup: each
// key's position is the number of smaller keys (the keys are distinct
// pairs).  For worlds whose possible candidate pairs fit SMALL_PAIR_SORT, so
// the count need not be read back to size a device-wide sort.
__global__ __launch_bounds__(1024) void k_pair_sort_small(DevWorld w) {
    __shared__ unsigned long long sk[SMALL_PAIR_SORT];
    const int K = min(w.g[G_NEW_PAIRS], w.pair_cap);
    if (K > SMALL_PAIR_SORT) {  // cannot happen for the worlds this is used on
        if (threadIdx.x == 0) w.g[G_ERR_PAIR_CAP] = 1;
        return;
    }
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(w.pairs);
    for (int i = threadIdx.x; i < K; i += blockDim.x) sk[i] = keys[i];
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += blockDim.x) {
        const unsigned long long k = sk[i];
        int r = 0;
        for (int j = 0; j < K; ++j) r += sk[j] < k ? 1 : 0;
        keys[r] = k;
    }
}

// Append the new pairs as manifolds in lexicographic (i, j) order -- the
// creation order of the reference's nested loop -- from the sorted keys.
__global__ void k_pair_append(DevWorld w, const unsigned long long* keys) {
    const int K = min(w.g[G_NEW_PAIRS], w.pair_cap);
    const int M0 = w.g[G_MANIFOLDS];
    const long long base = *w.next_serial;
    GRID_STRIDE(p, K) {
        const unsigned long long key = keys[p];
        const int a = (int)(key >> 32), b = (int)(key & 0xFFFFFFFFull);
        const int m = M0 + p;
        if (m < w.Mcap) {
            w.m.a[m] = a;
            w.m.b[m] = b;
            w.m.serial[m] = base + p;
            w.m.n[m] = 0;
            w.m.fric[m] = 0.f;
        }
    }
}

__global__ void k_pair_finish(DevWorld w) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int K = min(w.g[G_NEW_PAIRS], w.pair_cap);
    int M0 = w.g[G_MANIFOLDS];
    int M1 = M0 + K;
    if (M1 > w.Mcap) {
        w.g[G_ERR_MANIFOLD_CAP] = 1;
        M1 = w.Mcap;
    }
    w.g[G_MANIFOLDS_OLD] = M0;
    w.g[G_MANIFOLDS] = M1;
    *w.next_serial += K;
}

// ------------------------------------------------------------------ phase 2
// Manifold::Initialize (manifold.cpp:32-194) + the force warm start of
// solver.cpp:269-292, one thread per manifold.
// An edge (a, b) of the scene's constraint graph appeared or disappeared:
// the higher-index body's set of lower neighbours changed, so its colour may
// change (and through it those of bodies above it); the colouring is updated
// from there.  An edge to the world anchor (-1) is no edge.
AVBD_DEV void colour_seed(const DevWorld& w, int s, int a, int b) {
    if (a < 0 || b < 0) return;
    w.cseed[max(a, b)] = 1;
    atomicOr(&w.st[s * ST_WORDS + ST_DIRTY], DIRTY_SEEDS);
}

// The scene's colouring is updated only when its constraint graph changes:
// a manifold created this step (index >= the count before the broad phase)
// that keeps contacts, or an older one that loses them.  A new pair whose
// boxes do not touch is created and dropped in the same step and leaves the
// graph as it was.
__global__ void k_manifold_init(DevWorld w) {
    const int M = w.g[G_MANIFOLDS];
    const int M_old = w.g[G_MANIFOLDS_OLD];
    GRID_STRIDE(m, w.Mcap) {
        if (m >= M) {
            w.alive[m] = 0;
            continue;
        }
        const int a = w.m.a[m], b = w.m.b[m];
        const int s = w.scene[a];
        if (!scene_runs(w, s)) {
            w.alive[m] = 1;
            continue;
        }
        const avbd_scene& sc = w.sc[s];
        const float fa = w.angvel[a].w, fb = w.angvel[b].w;
        const float fric = sqrtf(fa * fb);
        const v3 pA = f3(w.pos[a]), pB = f3(w.pos[b]);
        const qt qA = fq(w.rot[a]), qB = fq(w.rot[b]);
        DevCollision col;
        const bool spheres = (sc.options & AVBD_SCENE_SPHERE_COLLISION) != 0;
        collide_bodies(pA, qA, f3(w.scale[a]), spheres && (w.flags[a] & AVBD_BODY_SPHERE), pB, qB, f3(w.scale[b]),
                       spheres && (w.flags[b] & AVBD_BODY_SPHERE), col);
        float4* hot = w.m.hot + (size_t)m * 8 * CONTACT_HOT_F4;
        float4* cold = w.m.cold + (size_t)m * 8 * CONTACT_COLD_F4;
        const int oldN = w.m.n[m];
        // match new contacts to old ones by feature id (manifold.cpp:74-99)
        float mk[8][3], ml[8][3];
        int mstick[8];
        bool isNew[8];
        for (int i = 0; i < col.n; ++i) {
            mk[i][0] = mk[i][1] = mk[i][2] = 0.f;
            ml[i][0] = ml[i][1] = ml[i][2] = 0.f;
            mstick[i] = 0;
            isNew[i] = true;
            for (int j = 0; j < oldN; ++j) {
                float4 c0 = cold[j * CONTACT_COLD_F4 + 0];
                if (__float_as_uint(c0.x) == col.c[i].id) {
                    float4 h3 = hot[j * CONTACT_HOT_F4 + 3];
                    float4 h4 = hot[j * CONTACT_HOT_F4 + 4];
                    float4 h5 = hot[j * CONTACT_HOT_F4 + 5];
                    ml[i][0] = h3.w; ml[i][1] = h4.w; ml[i][2] = h5.x;
                    mk[i][0] = h5.y; mk[i][1] = h5.z; mk[i][2] = h5.w;
                    mstick[i] = (int)c0.y;
                    if (mstick[i]) {
                        float4 c1 = cold[j * CONTACT_COLD_F4 + 1];
                        col.c[i].rA = v3{c0.z, c0.w, c1.x};
                        col.c[i].rB = v3{c1.y, c1.z, c1.w};
                    }
                    isNew[i] = false;
                    break;
                }
            }
        }
        const int n = col.n;
        w.m.n[m] = n;
        w.m.fric[m] = fric;
        if (n == 0) {
            w.alive[m] = 0;
            if (m < M_old) colour_seed(w, s, a, b);
            continue;
        }
        w.alive[m] = 1;
        if (m >= M_old) colour_seed(w, s, a, b);
        const m3 RA = rotmat(qA), RB = rotmat(qB);
        const bool decay = sc.post_stabilization == 0;
        for (int i = 0; i < n; ++i) {
            const DevContact& c = col.c[i];
            v3 nr = neg(c.normal);
            v3 t1 = fabsf(nr.x) > 0.9f ? cross(nr, v3{0.f, 1.f, 0.f}) : cross(nr, v3{1.f, 0.f, 0.f});
            t1 = vdiv(t1, norm(t1));
            v3 t2 = cross(nr, t1);
            v3 rAw = mv(RA, c.rA);
            v3 rBw = mv(RB, c.rB);
            v3 diff = sub(add(pA, rAw), add(pB, rBw));
            float c0n = dot(nr, diff) + 0.0005f;
            float c0t1 = dot(t1, diff);
            float c0t2 = dot(t2, diff);
            float k[3] = {mk[i][0], mk[i][1], mk[i][2]};
            float l[3] = {ml[i][0], ml[i][1], ml[i][2]};
            if (isNew[i]) {
                k[0] = clampf(c.penetration * sc.on_penetration_penalty, 1.f, 100000.0f);
                k[1] = 1.f;
                k[2] = 1.f;
            }
            // warm start (solver.cpp:280-291); manifold rows have infinite stiffness
            for (int r = 0; r < 3; ++r) {
                if (decay) l[r] = l[r] * (sc.alpha * sc.gamma);
                float kk = k[r] * sc.gamma;
                if (kk < 1.f) kk = 1.f;
                if (kk > 1000000000.0f) kk = 1000000000.0f;
                k[r] = kk;
            }
            hot[i * CONTACT_HOT_F4 + 0] = v4(nr, c0n);
            hot[i * CONTACT_HOT_F4 + 1] = v4(t1, c0t1);
            hot[i * CONTACT_HOT_F4 + 2] = v4(t2, c0t2);
            hot[i * CONTACT_HOT_F4 + 3] = v4(rAw, l[0]);
            hot[i * CONTACT_HOT_F4 + 4] = v4(rBw, l[1]);
            hot[i * CONTACT_HOT_F4 + 5] = make_float4(l[2], k[0], k[1], k[2]);
            cold[i * CONTACT_COLD_F4 + 0] = make_float4(__uint_as_float(c.id), (float)mstick[i], c.rA.x, c.rA.y);
            cold[i * CONTACT_COLD_F4 + 1] = make_float4(c.rA.z, c.rB.x, c.rB.y, c.rB.z);
            cold[i * CONTACT_COLD_F4 + 2] = v4(c.normal, c.penetration);
            if (w.m.cpos) w.m.cpos[(size_t)m * 8 + i] = v4(c.position, 0.f);
        }
    }
}

AVBD_DEV v3 transform_point(const DevWorld& w, int b, v3 p) {
    return add(qrot(fq(w.rot[b]), cmul(f3(w.scale[b]), p)), f3(w.pos[b]));
}
// AngularError (joint.cpp:13-21)
AVBD_DEV v3 joint_angular_error(qt qa, qt qb, qt rest, float arm) {
    return mul(quaternion_to_rotvec(qmul(qmul(qconj(qa), qb), qconj(rest))), arm);
}

// Joint::Initialize / Spring::Initialize and the force warm start.
__global__ void k_force_init(DevWorld w) {
    GRID_STRIDE(t, w.J + w.SP) {
        if (t < w.J) {
            const int j = t;
            if (w.j_flags[j] & FORCE_DEAD) continue;
            const int b = w.j_b[j];
            const int s = w.scene[b];
            if (!scene_runs(w, s)) continue;
            if (w.j_flags[j] & AVBD_FORCE_DISABLED) {
                w.j_flags[j] |= FORCE_DEAD;
                colour_seed(w, s, w.j_a[j], b);
                continue;
            }
            const int a = w.j_a[j];
            const float4 ra = w.j_ra[j];
            v3 wA = a >= 0 ? transform_point(w, a, f3(ra)) : f3(ra);
            v3 wB = transform_point(w, b, f3(w.j_rb[j]));
            v3 cp = sub(wA, wB);
            qt qa = a >= 0 ? fq(w.rot[a]) : qt{1.f, 0.f, 0.f, 0.f};
            v3 ca = joint_angular_error(qa, fq(w.rot[b]), fq(w.j_rest[j]), ra.w);
            float* c0 = w.j_c0 + j * 6;
            c0[0] = cp.x; c0[1] = cp.y; c0[2] = cp.z;
            c0[3] = ca.x; c0[4] = ca.y; c0[5] = ca.z;
            const avbd_scene& sc = w.sc[s];
            for (int r = 0; r < 6; ++r) {
                float* row = w.j_row + (j * 6 + r) * 6;
                if (sc.post_stabilization == 0) row[4] = row[4] * (sc.alpha * sc.gamma);
                float kk = row[5] * sc.gamma;
                if (kk < 1.f) kk = 1.f;
                if (kk > 1000000000.0f) kk = 1000000000.0f;
                row[5] = fminf(kk, row[0]);
            }
        } else {
            const int sp = t - w.J;
            if (w.s_flags[sp] & FORCE_DEAD) continue;
            const int b = w.s_b[sp];
            const int s = w.scene[b];
            if (!scene_runs(w, s)) continue;
            if (w.s_flags[sp] & AVBD_FORCE_DISABLED) {
                w.s_flags[sp] |= FORCE_DEAD;
                colour_seed(w, s, w.s_a[sp], b);
                continue;
            }
            const avbd_scene& sc = w.sc[s];
            float* row = w.s_row + sp * 6;
            if (sc.post_stabilization == 0) row[4] = row[4] * (sc.alpha * sc.gamma);
            float kk = row[5] * sc.gamma;
            if (kk < 1.f) kk = 1.f;
            if (kk > 1000000000.0f) kk = 1000000000.0f;
            row[5] = fminf(kk, row[0]);
        }
    }
}

// Stable compaction of the live manifolds into the other buffer.
__global__ void k_compact(DevWorld w, ManifoldBuf dst) {
    const int M = w.g[G_MANIFOLDS];
    GRID_STRIDE(m, M) {
        if (!w.alive[m]) continue;
        const int d = w.alive_off[m];
        dst.a[d] = w.m.a[m];
        dst.b[d] = w.m.b[m];
        dst.serial[d] = w.m.serial[m];
        const int n = w.m.n[m];
        dst.n[d] = n;
        dst.fric[d] = w.m.fric[m];
        const float4* hs = w.m.hot + (size_t)m * 8 * CONTACT_HOT_F4;
        float4* hd = dst.hot + (size_t)d * 8 * CONTACT_HOT_F4;
        for (int k = 0; k < n * CONTACT_HOT_F4; ++k) hd[k] = hs[k];
        const float4* cs = w.m.cold + (size_t)m * 8 * CONTACT_COLD_F4;
        float4* cd = dst.cold + (size_t)d * 8 * CONTACT_COLD_F4;
        for (int k = 0; k < n * CONTACT_COLD_F4; ++k) cd[k] = cs[k];
        if (dst.cpos && w.m.cpos)
            for (int k = 0; k < n; ++k) dst.cpos[(size_t)d * 8 + k] = w.m.cpos[(size_t)m * 8 + k];
    }
}

__global__ void k_compact_finish(DevWorld w) {
    if (threadIdx.x == 0 && blockIdx.x == 0) w.g[G_MANIFOLDS] = w.alive_off[w.Mcap];
}

// ------------------------------------------------------- incidence lists
__global__ void k_inc_count(DevWorld w) {
    const int M = w.g[G_MANIFOLDS];
    GRID_STRIDE(t, M + w.J + w.SP) {
        if (t < M) {
            const int n = w.m.n[t];
            atomicAdd(&w.cnt_i[w.m.a[t]], n);
            atomicAdd(&w.cnt_i[w.m.b[t]], n);
        } else if (t < M + w.J) {
            const int j = t - M;
            if (w.j_flags[j] & FORCE_DEAD) continue;
            if (w.j_a[j] >= 0) atomicAdd(&w.cnt_i[w.j_a[j]], 1);
            atomicAdd(&w.cnt_i[w.j_b[j]], 1);
        } else {
            const int sp = t - M - w.J;
            if (w.s_flags[sp] & FORCE_DEAD) continue;
            if (w.s_a[sp] >= 0) atomicAdd(&w.cnt_i[w.s_a[sp]], 1);
            atomicAdd(&w.cnt_i[w.s_b[sp]], 1);
        }
    }
}

__global__ void k_inc_fill(DevWorld w) {
    const int M = w.g[G_MANIFOLDS];
    GRID_STRIDE(t, M + w.J + w.SP) {
        if (t < M) {
            const int n = w.m.n[t];
            const int a = w.m.a[t], b = w.m.b[t];
            int pa = w.f_off[a] + atomicAdd(&w.cur_i[a], n);
            int pb = w.f_off[b] + atomicAdd(&w.cur_i[b], n);
            for (int c = 0; c < n; ++c) {
                w.f_ent[pa + c] = MAKE_ENT(ENT_CONTACT, 0, t * 8 + c);
                w.f_ent[pb + c] = MAKE_ENT(ENT_CONTACT, 1, t * 8 + c);
            }
        } else if (t < M + w.J) {
            const int j = t - M;
            if (w.j_flags[j] & FORCE_DEAD) continue;
            const int a = w.j_a[j], b = w.j_b[j];
            if (a >= 0) w.f_ent[w.f_off[a] + atomicAdd(&w.cur_i[a], 1)] = MAKE_ENT(ENT_JOINT, 0, j);
            w.f_ent[w.f_off[b] + atomicAdd(&w.cur_i[b], 1)] = MAKE_ENT(ENT_JOINT, 1, j);
        } else {
            const int sp = t - M - w.J;
            if (w.s_flags[sp] & FORCE_DEAD) continue;
            const int a = w.s_a[sp], b = w.s_b[sp];
            if (a >= 0) w.f_ent[w.f_off[a] + atomicAdd(&w.cur_i[a], 1)] = MAKE_ENT(ENT_SPRING, 0, sp);
            w.f_ent[w.f_off[b] + atomicAdd(&w.cur_i[b], 1)] = MAKE_ENT(ENT_SPRING, 1, sp);
        }
    }
}

// Creation order of an incidence entry: (force serial, contact index).
AVBD_DEV unsigned long long ent_key(const DevWorld& w, uint32_t e) {
    uint32_t idx = ENT_INDEX(e);
    unsigned long long s;
    unsigned c = 0;
    if (ENT_KIND(e) == ENT_CONTACT) {
        s = (unsigned long long)w.m.serial[idx >> 3];
        c = idx & 7u;
    } else if (ENT_KIND(e) == ENT_JOINT) {
        s = (unsigned long long)w.j_serial[idx];
    } else {
        s = (unsigned long long)w.s_serial[idx];
    }
    return (s << 3) | c;
}

// Creation key of entry e with one load through a selected array (no
// branch around it, so the keys of a list load together).
AVBD_DEV unsigned long long ent_key_flat(const DevWorld& w, uint32_t e) {
    const uint32_t idx = ENT_INDEX(e), kind = ENT_KIND(e);
    const bool contact = kind == ENT_CONTACT;
    const long long* arr = contact ? w.m.serial : (kind == ENT_JOINT ? w.j_serial : w.s_serial);
    const unsigned long long sr = (unsigned long long)arr[contact ? (idx >> 3) : idx];
    return (sr << 3) | (contact ? (idx & 7u) : 0u);
}

// Order each dynamic body's list by creation (Mesh::forces is append-ordered
// with stable erase, force.cpp:5-19): a rank sort in registers for short
// lists, insertion sort for longer ones.
__global__ void k_inc_sort(DevWorld w) {
    constexpr int KMAX = 16;
    GRID_STRIDE(b, w.N) {
        if (is_static(w, b)) continue;
        const int lo = w.f_off[b], hi = w.f_off[b + 1];
        const int n = hi - lo;
        if (n <= 1) continue;
        if (n <= KMAX) {
            // short lists: all entries and keys load at once (indices clamped
            // into the list), each entry is stored at its rank; equal keys
            // (serials are the records' own) rank by position, the order the
            // insertion sort keeps
            uint32_t e[KMAX];
            unsigned long long key[KMAX];
#pragma unroll
            for (int i = 0; i < KMAX; ++i) e[i] = w.f_ent[lo + min(i, n - 1)];
#pragma unroll
            for (int i = 0; i < KMAX; ++i) key[i] = i < n ? ent_key_flat(w, e[i]) : ~0ull;
#pragma unroll
            for (int i = 0; i < KMAX; ++i) {
                int r = 0;
#pragma unroll
                for (int j = 0; j < KMAX; ++j) r += (key[j] < key[i] || (key[j] == key[i] && j < i)) ? 1 : 0;
                if (i < n) w.f_ent[lo + r] = e[i];
            }
            continue;
        }
        for (int i = lo + 1; i < hi; ++i) {
            uint32_t e = w.f_ent[i];
            unsigned long long k = ent_key(w, e);
            int j = i - 1;
            while (j >= lo && ent_key(w, w.f_ent[j]) > k) {
                w.f_ent[j + 1] = w.f_ent[j];
                --j;
            }
            w.f_ent[j + 1] = e;
        }
    }
}

// ---------------------------------------------------------------- colouring
// Other body of incidence entry e seen from body self.
AVBD_DEV int ent_other(const DevWorld& w, uint32_t e) {
    uint32_t idx = ENT_INDEX(e);
    if (ENT_KIND(e) == ENT_CONTACT) {
        int m = idx >> 3;
        return ENT_SIDE_B(e) ? w.m.a[m] : w.m.b[m];
    } else if (ENT_KIND(e) == ENT_JOINT) {
        return ENT_SIDE_B(e) ? w.j_a[idx] : w.j_b[idx];
    }
    return ENT_SIDE_B(e) ? w.s_a[idx] : w.s_b[idx];
}

template <class F>
AVBD_DEV void for_each_neighbour(const DevWorld& w, int v, F f) {
    for (int k = w.f_off[v]; k < w.f_off[v + 1]; ++k) {
        int o = ent_other(w, w.f_ent[k]);
        if (o >= 0) f(o);
    }
    for (int k = w.e_off[v]; k < w.e_off[v + 1]; ++k) {
        uint32_t e = w.e_ent[k];
        int local = EENT_LOCAL(e);
        const int4 t = w.ev_v[k];
        if (EENT_TRI(e)) {
            if (local != 0) f(t.x);
            if (local != 1) f(t.y);
            if (local != 2) f(t.z);
        } else {
            if (local != 0) f(t.x);
            if (local != 1) f(t.y);
            if (local != 2) f(t.z);
            if (local != 3) f(t.w);
        }
    }
}

// Neighbours of v reached through its incidence entries part, part + parts,
// ... (the share of one lane of a group of `parts` lanes).
template <class F>
AVBD_DEV void for_each_neighbour_part(const DevWorld& w, int v, int part, int parts, F f) {
    for (int k = w.f_off[v] + part; k < w.f_off[v + 1]; k += parts) {
        const int o = ent_other(w, w.f_ent[k]);
        if (o >= 0) f(o);
    }
    for (int k = w.e_off[v] + part; k < w.e_off[v + 1]; k += parts) {
        const uint32_t e = w.e_ent[k];
        const int local = EENT_LOCAL(e);
        const int4 t = w.ev_v[k];
        if (local != 0) f(t.x);
        if (local != 1) f(t.y);
        if (local != 2) f(t.z);
        if (!EENT_TRI(e) && local != 3) f(t.w);
    }
}

// The neighbours one lane of a group reaches (for_each_neighbour_part), held
// in registers so that a body's two scans in a colouring round load them
// once, with all loads issued before any is used.  Lanes whose share exceeds
// NB_CACHE entries of a kind reach the rest through the plain loops.
#define NB_CACHE 4
struct NbCache {
    int nf, ne;
    int fo[NB_CACHE];
    uint32_t ee[NB_CACHE];
    int4 ev[NB_CACHE];
};

AVBD_DEV void nb_load(const DevWorld& w, int v, int part, int parts, NbCache& c) {
    const int f0 = w.f_off[v] + part, f1 = w.f_off[v + 1];
    const int e0 = w.e_off[v] + part, e1 = w.e_off[v + 1];
    c.nf = f1 > f0 ? min((f1 - f0 + parts - 1) / parts, NB_CACHE) : 0;
    c.ne = e1 > e0 ? min((e1 - e0 + parts - 1) / parts, NB_CACHE) : 0;
    uint32_t fe[NB_CACHE];
#pragma unroll
    for (int j = 0; j < NB_CACHE; ++j) {
        if (j < c.nf) fe[j] = w.f_ent[f0 + j * parts];
        if (j < c.ne) {
            c.ee[j] = w.e_ent[e0 + j * parts];
            c.ev[j] = w.ev_v[e0 + j * parts];
        }
    }
#pragma unroll
    for (int j = 0; j < NB_CACHE; ++j)
        if (j < c.nf) c.fo[j] = ent_other(w, fe[j]);
}

template <class F>
AVBD_DEV void nb_apply(const DevWorld& w, int v, int part, int parts, const NbCache& c, F f) {
#pragma unroll
    for (int j = 0; j < NB_CACHE; ++j)
        if (j < c.nf && c.fo[j] >= 0) f(c.fo[j]);
#pragma unroll
    for (int j = 0; j < NB_CACHE; ++j) {
        if (j < c.ne) {
            const uint32_t e = c.ee[j];
            const int local = EENT_LOCAL(e);
            const int4 t = c.ev[j];
            if (local != 0) f(t.x);
            if (local != 1) f(t.y);
            if (local != 2) f(t.z);
            if (!EENT_TRI(e) && local != 3) f(t.w);
        }
    }
    // entries beyond the cache
    for (int k = w.f_off[v] + part + NB_CACHE * parts; k < w.f_off[v + 1]; k += parts) {
        const int o = ent_other(w, w.f_ent[k]);
        if (o >= 0) f(o);
    }
    for (int k = w.e_off[v] + part + NB_CACHE * parts; k < w.e_off[v + 1]; k += parts) {
        const uint32_t e = w.e_ent[k];
        const int local = EENT_LOCAL(e);
        const int4 t = w.ev_v[k];
        if (local != 0) f(t.x);
        if (local != 1) f(t.y);
        if (local != 2) f(t.z);
        if (!EENT_TRI(e) && local != 3) f(t.w);
    }
}

// Greedy first-fit colouring of one scene per workgroup, bit-identical to the
// sequential rule "dynamic bodies in ascending index take the smallest colour
// not used by lower-index dynamic neighbours".  A body is coloured as soon as
// all its lower-index neighbours are (a counter per body), so the result does
// not depend on the order threads run in.  Rounds follow the longest chain of
// ascending neighbours (~180 for a 32^3 tet grid), and each round reads every
// neighbour's colour, so those reads decide the cost: scenes of up to
// COLOUR_LDS_MAX bodies keep one word per body in LDS, (colour << 16) |
// pending lower neighbours (colour 0xFFFF = not yet, 0xFFFE = static); larger
// scenes keep colours and counters in global memory.  Each instantiation
// takes only the scenes of its kind.  A scene whose graph only gained or lost
// some edges since its last colouring (DIRTY_SEEDS, the higher endpoints in
// cseed) updates that colouring instead (full = 0 allows it).
template <bool LDS>
__global__ __launch_bounds__(1024) void k_colour(DevWorld w, int full) {
    extern __shared__ unsigned int cw[];
    constexpr int CG = 8;  // lanes sharing one body's incidence list
    const int s = blockIdx.x;
    if (!scene_runs(w, s)) return;
    int* st = w.st + s * ST_WORDS;
    const int dirty = st[ST_DIRTY];
    if (!dirty) return;
    const int off = w.sc_off[s], end = w.sc_off[s + 1];
    const int n = end - off;
    if (LDS != (n <= COLOUR_LDS_MAX)) return;
    int* qa = w.queue + 2 * off;
    int* qb = qa + n;
    __shared__ int qn[2];
    __shared__ int maxc;
    __shared__ int overflow;
    if (threadIdx.x == 0) { qn[0] = 0; qn[1] = 0; maxc = -1; overflow = 0; }
    const unsigned UNCOL = 0xFFFFu, STAT = 0xFFFEu;
    auto nb_static = [&](int o) -> bool { return LDS ? (cw[o - off] >> 16) == STAT : is_static(w, o); };
    auto nb_colour = [&](int o) -> int {
        if (LDS) {
            const unsigned c = cw[o - off] >> 16;
            return c >= STAT ? -1 : (int)c;
        }
        return __atomic_load_n(&w.colour[o], __ATOMIC_RELAXED);
    };
    const int part = threadIdx.x % CG;
    const int groups = blockDim.x / CG;
    // smallest colour not taken by v's lower-index dynamic neighbours (the
    // group's CG lanes each see a share of them)
    auto first_free = [&](int v, bool valid, const NbCache& nc) -> int {
        unsigned long long used = 0ull;
        if (valid)
            nb_apply(w, v, part, CG, nc, [&](int o) {
                if (o < v && !nb_static(o)) {
                    const int c = nb_colour(o);
                    if (c >= 0 && c < 64) used |= (1ull << c);
                }
            });
#pragma unroll
        for (int m = 1; m < CG; m <<= 1) used |= __shfl_xor(used, m, CG);
        int c = __ffsll((long long)~used) - 1;
        if (used == ~0ull) {
            // more than 64 colours: linear search above 64
            c = 64;
            while (true) {
                int taken = 0;
                if (valid)
                    for_each_neighbour_part(w, v, part, CG, [&](int o) {
                        if (o < v && !nb_static(o) && nb_colour(o) == c) taken = 1;
                    });
#pragma unroll
                for (int m = 1; m < CG; m <<= 1) taken |= __shfl_xor(taken, m, CG);
                if (!taken) break;
                ++c;
            }
        }
        return c;
    };
    if ((dirty & DIRTY_FULL) || full) {
        if (LDS) {
            // static flags first: neighbour tests then read LDS only
            for (int v = off + threadIdx.x; v < end; v += blockDim.x)
                cw[v - off] = (is_static(w, v) ? STAT : UNCOL) << 16;
        } else {
            for (int v = off + threadIdx.x; v < end; v += blockDim.x) w.colour[v] = -1;
        }
        __syncthreads();
        // lower-neighbour counts; bodies with none start the first round
        for (int base = off; base < end; base += groups) {
            const int v = base + threadIdx.x / CG;
            const bool dyn = v < end && !nb_static(v);
            int c = 0;
            if (dyn)
                for_each_neighbour_part(w, v, part, CG, [&](int o) {
                    if (o < v && !nb_static(o)) ++c;
                });
#pragma unroll
            for (int m = 1; m < CG; m <<= 1) c += __shfl_xor(c, m, CG);
            if (dyn && part == 0) {
                if (LDS) {
                    if (c > 0xFFFF) overflow = 1;
                    cw[v - off] = (UNCOL << 16) | (unsigned)min(c, 0xFFFF);
                } else {
                    w.ccount[v] = c;
                }
                if (c == 0) qa[atomicAdd(&qn[0], 1)] = v;
            }
        }
        __syncthreads();
        // Rounds: every body of the current queue takes its colour, then at
        // once releases its higher neighbours (their counters); those that
        // reach zero form the next queue.  No barrier is needed between the
        // two steps of one body: a body of the queue never reads the colour
        // of another body of the same queue (they are not neighbours, or one
        // would still be waiting for the other), and a released body is only
        // coloured in a later round.
        int cur = 0;
        while (!overflow) {
            const int len = qn[cur];
            if (len == 0) break;
            int* q = cur ? qb : qa;
            int* nq = cur ? qa : qb;
            for (int kb = 0; kb < len; kb += groups) {
                const int k = kb + threadIdx.x / CG;
                const bool valid = k < len;
                const int v = valid ? q[k] : 0;
                NbCache nc;
                nc.nf = nc.ne = 0;
                if (valid) nb_load(w, v, part, CG, nc);
                const int c = first_free(v, valid, nc);
                if (valid && part == 0) {
                    if (LDS) cw[v - off] = (unsigned)min(c, 0xFFFD) << 16;  // its counter is 0
                    else __atomic_store_n(&w.colour[v], c, __ATOMIC_RELEASE);
                    atomicMax(&maxc, c);
                }
                // the group's colour write precedes every lane's releases
                if (LDS) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __builtin_amdgcn_wave_barrier();
                if (valid)
                    nb_apply(w, v, part, CG, nc, [&](int o) {
                        if (o > v && !nb_static(o)) {
                            const bool ready = LDS ? (atomicSub(&cw[o - off], 1u) & 0xFFFFu) == 1u
                                                   : atomicSub(&w.ccount[o], 1) == 1;
                            if (ready) nq[atomicAdd(&qn[cur ^ 1], 1)] = o;
                        }
                    });
            }
            __syncthreads();
            if (threadIdx.x == 0) qn[cur] = 0;
            cur ^= 1;
            __syncthreads();
        }
        if (LDS) {
            for (int v = off + threadIdx.x; v < end; v += blockDim.x) {
                const unsigned c = cw[v - off] >> 16;
                w.colour[v] = c >= STAT ? -1 : (int)c;
            }
        }
        for (int v = off + threadIdx.x; v < end; v += blockDim.x) w.cseed[v] = 0;
    } else {
        // Update of the last colouring: only bodies whose lower neighbours
        // changed can change colour, and a body whose colour changes can
        // change those of its higher neighbours.  Work lists of such bodies
        // are re-evaluated round by round until no colour changes.  The
        // greedy colouring is the unique fixed point of "colour(v) = first
        // colour free among v's lower neighbours" (by induction over the
        // index), and when a round changes nothing every body last evaluated
        // saw its lower neighbours' final colours (a later change would have
        // queued it again), so the result equals a colouring from scratch.
        // A body is queued at most once per round (ccount = round tag).
        if (LDS) {
            for (int v = off + threadIdx.x; v < end; v += blockDim.x) {
                const int c = w.colour[v];
                cw[v - off] = (is_static(w, v) ? STAT : (unsigned)min(c, 0xFFFD)) << 16;
            }
        }
        for (int v = off + threadIdx.x; v < end; v += blockDim.x) w.ccount[v] = -1;
        __syncthreads();
        for (int v = off + threadIdx.x; v < end; v += blockDim.x) {
            if (w.cseed[v]) {
                w.cseed[v] = 0;
                if (!nb_static(v)) {
                    w.ccount[v] = 0;
                    qa[atomicAdd(&qn[0], 1)] = v;
                }
            }
        }
        __syncthreads();
        int cur = 0;
        for (int round = 0;; ++round) {
            const int len = qn[cur];
            if (len == 0) break;
            int* q = cur ? qb : qa;
            int* nq = cur ? qa : qb;
            for (int kb = 0; kb < len; kb += groups) {
                const int k = kb + threadIdx.x / CG;
                const bool valid = k < len;
                const int v = valid ? q[k] : 0;
                NbCache nc;
                nc.nf = nc.ne = 0;
                if (valid) nb_load(w, v, part, CG, nc);
                const int c = first_free(v, valid, nc);
                const bool changed = valid && c != nb_colour(v);
                // every lane of the group has read the old colour
                __builtin_amdgcn_wave_barrier();
                if (changed && part == 0) {
                    if (LDS) cw[v - off] = (unsigned)min(c, 0xFFFD) << 16;
                    else __atomic_store_n(&w.colour[v], c, __ATOMIC_RELAXED);
                }
                if (changed)
                    nb_apply(w, v, part, CG, nc, [&](int o) {
                        if (o > v && !nb_static(o) && atomicExch(&w.ccount[o], round + 1) != round + 1)
                            nq[atomicAdd(&qn[cur ^ 1], 1)] = o;
                    });
            }
            __syncthreads();
            if (threadIdx.x == 0) qn[cur] = 0;
            cur ^= 1;
            __syncthreads();
        }
        for (int v = off + threadIdx.x; v < end; v += blockDim.x) {
            const int c = nb_colour(v);
            if (!nb_static(v)) atomicMax(&maxc, c);
            if (LDS) w.colour[v] = nb_static(v) ? -1 : c;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        st[ST_NCOL] = maxc + 1;
        st[ST_DIRTY] = 0;
        if (maxc + 1 > AVBD_MAX_COLOURS || overflow) w.g[G_ERR_COLOURS] = 1;
    }
}

template __global__ void k_colour<true>(DevWorld, int);
template __global__ void k_colour<false>(DevWorld, int);

// Colour of body b if it takes part in this step's sweeps, else -1.
// Sweep bucket of body b (2 * colour + 1 for bodies whose only terms are FEM
// energies, 2 * colour otherwise), or -1 if it takes no part in the sweeps.
AVBD_DEV int sweep_bucket(const DevWorld& w, int b) {
    if (b >= w.N || is_static(w, b) || !scene_runs(w, w.scene[b]) || scene_fused(w, w.scene[b])) return -1;
    int c = w.colour[b];
    if (c < 0 || c >= AVBD_MAX_COLOURS) return -1;
    const float4 I = w.inertia[b];
    const bool fem_only = w.f_off[b + 1] == w.f_off[b] && w.e_off[b + 1] > w.e_off[b] && I.x == 0.f && I.y == 0.f &&
                          I.z == 0.f;
    return 2 * c + (fem_only ? 1 : 0);
}

// Bucket histogram: LDS counters per block, one global atomic per block and
// bucket (a single global counter per colour would serialise millions of
// atomics).
__global__ __launch_bounds__(256) void k_colour_hist(DevWorld w) {
    __shared__ int h[AVBD_BUCKETS];
    for (int k = threadIdx.x; k < AVBD_BUCKETS; k += blockDim.x) h[k] = 0;
    __syncthreads();
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < w.N; b += blockDim.x * gridDim.x) {
        int k = sweep_bucket(w, b);
        if (k >= 0) atomicAdd(&h[k], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < AVBD_BUCKETS; k += blockDim.x)
        if (h[k]) atomicAdd(&w.col_cnt[k], h[k]);
}

__global__ __launch_bounds__(256) void k_colour_scan(DevWorld w) {
    __shared__ int buf[AVBD_BUCKETS];
    for (int k = threadIdx.x; k < AVBD_BUCKETS; k += blockDim.x) buf[k] = w.col_cnt[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0, last = -1;
        for (int k = 0; k < AVBD_BUCKETS; ++k) {
            w.col_off[k] = acc;
            if (buf[k] > 0) last = k / 2;
            acc += buf[k];
        }
        w.col_off[AVBD_BUCKETS] = acc;
        w.g[G_NCOLOURS] = last + 1;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < AVBD_BUCKETS; k += blockDim.x) w.col_cnt[k] = 0;
}

// Scatter bodies into the bucket lists: each block reserves one range per
// bucket with a single global atomic, then places its bodies with LDS
// atomics.  The order inside a colour is irrelevant (its bodies are
// independent).
__global__ __launch_bounds__(256) void k_colour_fill(DevWorld w) {
    __shared__ int h[AVBD_BUCKETS];
    __shared__ int base[AVBD_BUCKETS];
    for (int start = blockIdx.x * blockDim.x; start < w.N; start += blockDim.x * gridDim.x) {
        for (int k = threadIdx.x; k < AVBD_BUCKETS; k += blockDim.x) h[k] = 0;
        __syncthreads();
        const int b = start + threadIdx.x;
        const int kb = sweep_bucket(w, b);
        int r = kb >= 0 ? atomicAdd(&h[kb], 1) : 0;
        __syncthreads();
        for (int k = threadIdx.x; k < AVBD_BUCKETS; k += blockDim.x)
            if (h[k]) base[k] = w.col_off[k] + atomicAdd(&w.col_cnt[k], h[k]);
        __syncthreads();
        if (kb >= 0) {
            w.col_list[base[kb] + r] = b;
            if (kb & 1) {
                const int e0 = w.e_off[b];
                w.col_fem[base[kb] + r] = make_int4(b, e0, w.scene[b], ((w.e_off[b + 1] - e0) << 8) | (w.nb_n[b] + 1));
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ phase 4
// Body warm start (solver.cpp:324-384), all bodies of running scenes.
AVBD_DEV void predict_body(const DevWorld& w, int b, int s) {
    {
        const avbd_scene& sc = w.sc[s];
        const float dt = sc.dt;
        const v3 g{sc.gravity[0], sc.gravity[1], sc.gravity[2]};
        const uint32_t fl = w.flags[b];
        const bool st = (fl & AVBD_BODY_STATIC) != 0;
        const bool part = (fl & AVBD_BODY_PARTICLE) != 0;
        float4 av4 = w.angvel[b];
        v3 av = f3(av4);
        av.x = fmaxf(fminf(av.x, 50.f), -50.f);
        av.y = fmaxf(fminf(av.y, 50.f), -50.f);
        av.z = fmaxf(fminf(av.z, 50.f), -50.f);
        w.angvel[b] = v4(av, av4.w);
        const float4 vel4 = w.vel[b];
        const v3 vel = f3(vel4);
        const v3 p = f3(w.pos[b]);
        const qt rq = fq(w.rot[b]);
        v3 ip = add(p, mul(vel, dt));
        if (!st) ip = add(ip, mul(mul(g, dt), dt));
        qt ir = rq;
        qt nr = rq;
        if (!st && !part) {
            qt wq;
            const v3 wv0 = av;
            const v3 qv{rq.x, rq.y, rq.z};
            wq.w = 0.f * rq.w - dot(wv0, qv);
            v3 wv = add(add(smul(0.f, qv), smul(rq.w, wv0)), cross(wv0, qv));
            qt q = rq;
            q.w = q.w + (0.5f * wq.w) * dt;
            q.x = q.x + (0.5f * wv.x) * dt;
            q.y = q.y + (0.5f * wv.y) * dt;
            q.z = q.z + (0.5f * wv.z) * dt;
            ir = qnormalized(q);
            nr = ir;
        }
        v3 acc = vdiv(sub(vel, f3(w.pvel[b])), dt);
        float gm = norm(g);
        float wgt = 0.f;
        if (gm > 1e-6f) {
            v3 gd = vdiv(g, gm);
            wgt = dot(acc, gd) / gm;
            wgt = clampf(wgt, 0.f, 1.f);
        }
        v3 np = add(add(p, mul(vel, dt)), mul(mul(mul(g, dt), dt), wgt));
        w.last_pos[b] = v4(p, 0.f);
        w.last_rot[b] = q4(rq);
        w.inert_pos[b] = v4(ip, 0.f);
        w.inert_rot[b] = q4(ir);
        w.pos[b] = v4(np, 0.f);
        w.rot[b] = q4(nr);
        v3 dl = sub(np, p);
        v3 da = rotation_difference(nr, rq);
        w.diff0[b] = make_float4(dl.x, dl.y, dl.z, da.x);
        w.diff1[b] = make_float4(da.y, da.z, 0.f, 0.f);
    }
}

__global__ void k_predict(DevWorld w) {
    GRID_STRIDE(b, w.N) {
        const int s = w.scene[b];
        if (!scene_runs(w, s) || scene_fused(w, s)) continue;
        predict_body(w, b, s);
    }
}

// ------------------------------------------------------------------ phase 5
struct Acc {
    float L[21];  // packed lower 6x6
    float r[6];
};

AVBD_DEV void acc_zero(Acc& A) {
#pragma unroll
    for (int i = 0; i < 21; ++i) A.L[i] = 0.f;
#pragma unroll
    for (int i = 0; i < 6; ++i) A.r[i] = 0.f;
}

// rhs += J f ; lhs += (J J^T) k  (solver.cpp:417-419)
AVBD_DEV void acc_row(Acc& A, const float J[6], float f, float k) {
#pragma unroll
    for (int i = 0; i < 6; ++i) A.r[i] = A.r[i] + J[i] * f;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) A.L[tri_idx(i, j)] = A.L[tri_idx(i, j)] + (J[i] * J[j]) * k;
}

AVBD_DEV float dot6(const float a[6], const float b[6]) {
    float s = a[0] * b[0];
#pragma unroll
    for (int i = 1; i < 6; ++i) s = s + a[i] * b[i];
    return s;
}

AVBD_DEV void load_diff(const DevWorld& w, int b, float d[6]) {
    float4 x = w.diff0[b], y = w.diff1[b];
    d[0] = x.x; d[1] = x.y; d[2] = x.z; d[3] = x.w; d[4] = y.x; d[5] = y.y;
}

AVBD_DEV void set_J(float J[6], v3 lin, v3 ang) {
    J[0] = lin.x; J[1] = lin.y; J[2] = lin.z; J[3] = ang.x; J[4] = ang.y; J[5] = ang.z;
}

// Constraint values of one contact (Manifold::ComputeConstraints,
// manifold.cpp:197-236): C for the three rows and the friction bound.
struct ContactRows {
    v3 n, t1, t2, rAw, rBw;
    float C[3];
    float lam[3];
    float k[3];
    float bound;
};

AVBD_DEV void contact_rows(const DevWorld& w, int m, int c, float alpha, const float dA[6], const float dB[6],
                           ContactRows& R) {
    const float4* h = w.m.hot + ((size_t)m * 8 + c) * CONTACT_HOT_F4;
    const float4 h0 = h[0], h1 = h[1], h2 = h[2], h3 = h[3], h4 = h[4], h5 = h[5];
    R.n = f3(h0); R.t1 = f3(h1); R.t2 = f3(h2); R.rAw = f3(h3); R.rBw = f3(h4);
    R.lam[0] = h3.w; R.lam[1] = h4.w; R.lam[2] = h5.x;
    R.k[0] = h5.y; R.k[1] = h5.z; R.k[2] = h5.w;
    const float oma = 1.f - alpha;
    const float c0[3] = {h0.w, h1.w, h2.w};
    const v3 dir[3] = {R.n, R.t1, R.t2};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float JA[6], JB[6];
        set_J(JA, dir[r], cross(R.rAw, dir[r]));
        set_J(JB, neg(dir[r]), neg(cross(R.rBw, dir[r])));
        R.C[r] = (oma * c0[r] + dot6(JA, dA)) + dot6(JB, dB);
    }
    R.bound = fabsf(R.lam[0]) * w.m.fric[m];
}

// Joint rows (joint.cpp:97-173): C, J and the (5,5) Hessian entry per row
// for the given side.
struct JointRows {
    float C[6];
    float J[6][6];
    float h55[3];
};

AVBD_DEV void joint_rows(const DevWorld& w, int j, bool sideB, float alpha, JointRows& R) {
    const int a = w.j_a[j], b = w.j_b[j];
    const float4 ra = w.j_ra[j], rb = w.j_rb[j];
    v3 wA = a >= 0 ? transform_point(w, a, f3(ra)) : f3(ra);
    v3 wB = transform_point(w, b, f3(rb));
    v3 cp = sub(wA, wB);
    qt qa = a >= 0 ? fq(w.rot[a]) : qt{1.f, 0.f, 0.f, 0.f};
    qt qb = fq(w.rot[b]);
    v3 ca = joint_angular_error(qa, qb, fq(w.j_rest[j]), ra.w);
    const float* c0 = w.j_c0 + j * 6;
    const float* rows = w.j_row + j * 36;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        R.C[i] = isinf(rows[i * 6 + 0]) ? comp(cp, i) - c0[i] * alpha : comp(cp, i);
        R.C[3 + i] = isinf(rows[(3 + i) * 6 + 0]) ? comp(ca, i) - c0[3 + i] * alpha : comp(ca, i);
    }
    const int self = sideB ? b : a;
    v3 r = qrot(fq(w.rot[self]), cmul(f3(w.scale[self]), sideB ? f3(rb) : f3(ra)));
    const float arm = ra.w;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        v3 e = unit_axis(i);
        if (!sideB) {
            set_J(R.J[i], e, cross(r, e));
            R.h55[i] = -comp(r, i);
            set_J(R.J[3 + i], v3{0.f, 0.f, 0.f}, smul(-arm, e));
        } else {
            set_J(R.J[i], neg(e), neg(cross(r, e)));
            R.h55[i] = comp(r, i);
            set_J(R.J[3 + i], v3{0.f, 0.f, 0.f}, smul(arm, e));
        }
    }
}

AVBD_DEV float spring_C(const DevWorld& w, int sp) {
    const int a = w.s_a[sp], b = w.s_b[sp];
    const float4 ra = w.s_ra[sp];
    v3 pA = a >= 0 ? transform_point(w, a, f3(ra)) : f3(ra);
    v3 pB = transform_point(w, b, f3(w.s_rb[sp]));
    return norm(sub(pB, pA)) - ra.w;
}

// Spring derivatives (spring.cpp:51-109) for one side: J and full 6x6 H.
AVBD_DEV void spring_derivs(const DevWorld& w, int sp, bool sideB, float J[6], float H[6][6]) {
    const int a = w.s_a[sp], b = w.s_b[sp];
    const float4 ra = w.s_ra[sp], rb = w.s_rb[sp];
    v3 pA = a >= 0 ? transform_point(w, a, f3(ra)) : f3(ra);
    v3 pB = transform_point(w, b, f3(rb));
    v3 d = sub(pB, pA);
    float L = norm(d);
    float L2 = dot(d, d);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        J[i] = 0.f;
#pragma unroll
        for (int j = 0; j < 6; ++j) H[i][j] = 0.f;
    }
    if (L2 < 1e-6f) return;
    v3 n = vdiv(d, L);
    m3 dxx;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) dxx.m[i][j] = ((i == j ? 1.f : 0.f) - comp(n, i) * comp(n, j)) / L;
    m3 dxr, drr;
    if (!sideB && a >= 0) {
        v3 r = qrot(fq(w.rot[a]), f3(ra));
        m3 Sr = skew(r);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) Sr.m[i][j] = -Sr.m[i][j];
        dxr = mm(dxx, transpose(Sr));
        m3 t1 = mm(mm(Sr, dxx), transpose(Sr));
        m3 t2 = mm(skew(n), skew(r));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) drr.m[i][j] = t1.m[i][j] - t2.m[i][j];
        set_J(J, neg(n), cross(n, r));
    } else {
        v3 r = qrot(fq(w.rot[b]), f3(rb));
        m3 mSr = skew(r);
        dxr = mm(dxx, transpose(mSr));
        m3 t1 = mm(mm(mSr, dxx), transpose(mSr));
        m3 t2 = mm(skew(n), skew(r));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) drr.m[i][j] = t1.m[i][j] + t2.m[i][j];
        set_J(J, n, cross(r, n));
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            H[i][j] = dxx.m[i][j];
            H[i][3 + j] = dxr.m[i][j];
            H[3 + i][j] = dxr.m[j][i];
            H[3 + i][3 + j] = drr.m[i][j];
        }
}

// Contributions of one incidence entry to the block system of body `self`.
AVBD_DEV void accumulate_force(const DevWorld& w, uint32_t e, float alpha, const float dSelf[6], Acc& A) {
    const uint32_t idx = ENT_INDEX(e);
    const bool sideB = ENT_SIDE_B(e) != 0;
    const uint32_t kind = ENT_KIND(e);
    if (kind == ENT_CONTACT) {
        const int m = idx >> 3, c = idx & 7;
        float dO[6];
        load_diff(w, sideB ? w.m.a[m] : w.m.b[m], dO);
        ContactRows R;
        contact_rows(w, m, c, alpha, sideB ? dO : dSelf, sideB ? dSelf : dO, R);
        const v3 dir[3] = {R.n, R.t1, R.t2};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            float lo = r == 0 ? -INFINITY : -R.bound;
            float hi = r == 0 ? 0.f : R.bound;
            float f = clampf(R.k[r] * R.C[r] + R.lam[r], lo, hi);
            float J[6];
            if (!sideB) set_J(J, dir[r], cross(R.rAw, dir[r]));
            else set_J(J, neg(dir[r]), neg(cross(R.rBw, dir[r])));
            acc_row(A, J, f, R.k[r]);
        }
    } else if (kind == ENT_JOINT) {
        JointRows R;
        joint_rows(w, idx, sideB, alpha, R);
        const float* rows = w.j_row + idx * 36;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const float stiff = rows[r * 6 + 0], fmn = rows[r * 6 + 2], fmx = rows[r * 6 + 3];
            const float lam = rows[r * 6 + 4], k = rows[r * 6 + 5];
            float f = clampf(k * R.C[r] + (isinf(stiff) ? lam : 0.f), fmn, fmx);
            acc_row(A, R.J[r], f, k);
            if (r < 3) A.L[tri_idx(5, 5)] = A.L[tri_idx(5, 5)] + sqrtf(R.h55[r] * R.h55[r]) * fabsf(f);
        }
    } else {
        const float* row = w.s_row + idx * 6;
        const float stiff = row[0], fmn = row[2], fmx = row[3], lam = row[4], k = row[5];
        float C = spring_C(w, idx);
        float J[6], H[6][6];
        spring_derivs(w, idx, sideB, J, H);
        float f = clampf(k * C + (isinf(stiff) ? lam : 0.f), fmn, fmx);
        acc_row(A, J, f, k);
        const float af = fabsf(f);
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            float cn = H[0][j] * H[0][j];
#pragma unroll
            for (int i = 1; i < 6; ++i) cn = cn + H[i][j] * H[i][j];
            A.L[tri_idx(j, j)] = A.L[tri_idx(j, j)] + sqrtf(cn) * af;
        }
    }
}

AVBD_DEV void energy_terms(const DevWorld& w, uint32_t e, int k, int projMode, v3& g, float H[6]);

AVBD_DEV void accumulate_energy(const DevWorld& w, int k, int projMode, Acc& A) {
    v3 g;
    float H[6];
    energy_terms(w, w.e_ent[k], k, projMode, g, H);
    A.r[0] = A.r[0] + g.x;
    A.r[1] = A.r[1] + g.y;
    A.r[2] = A.r[2] + g.z;
#pragma unroll
    for (int i = 0; i < 6; ++i) A.L[i] = A.L[i] + H[i];
}

// Gradient and packed lower 3x3 Hessian of energy incidence entry k (value
// e = e_ent[k]) for its vertex; the element's data comes from the
// incidence-ordered stream.
// The three stages of one energy entry, split so a loop can overlap the
// loads of later entries with the arithmetic of the current one.
struct ElemIn {  // entry value and the element's stream record
    uint32_t e;
    float4 a, b, c;
    int4 v;
};
struct ElemPos {  // the element's vertex positions
    v3 p0, p1, p2, p3;
};

AVBD_DEV void elem_load(const DevWorld& w, int k, ElemIn& x) {
    x.e = w.e_ent[k];
    x.a = w.ev_a[k];
    x.b = w.ev_b[k];
    x.c = w.ev_c[k];
    x.v = w.ev_v[k];
}

AVBD_DEV void elem_pos(const DevWorld& w, const ElemIn& x, ElemPos& p) {
    p.p0 = f3(w.pos[x.v.x]);
    p.p1 = f3(w.pos[x.v.y]);
    p.p2 = f3(w.pos[x.v.z]);
    p.p3 = EENT_TRI(x.e) ? p.p2 : f3(w.pos[x.v.w]);
}

// All four vertex positions of an element, unconditionally (a triangle's
// slot 3 repeats its slot 0 id, so every gather stays inside the element's
// scene, and is not used).
AVBD_DEV void elem_gather(const DevWorld& w, int4 v, ElemPos& p) {
    p.p0 = f3(w.pos[v.x]);
    p.p1 = f3(w.pos[v.y]);
    p.p2 = f3(w.pos[v.z]);
    p.p3 = f3(w.pos[v.w]);
}

AVBD_DEV void elem_terms(const ElemIn& x, const ElemPos& p, int projMode, v3& g, float H[6]) {
    const int local = EENT_LOCAL(x.e);
    if (!EENT_TRI(x.e)) {
        const float4 d0 = x.a, d1 = x.b, d2 = x.c;
        m3 Dm;
        Dm.m[0][0] = d0.x; Dm.m[0][1] = d0.y; Dm.m[0][2] = d0.z;
        Dm.m[1][0] = d0.w; Dm.m[1][1] = d1.x; Dm.m[1][2] = d1.y;
        Dm.m[2][0] = d1.z; Dm.m[2][1] = d1.w; Dm.m[2][2] = d2.x;
        const v3 g1 = row(Dm, 0), g2 = row(Dm, 1), g3 = row(Dm, 2);
        v3 gN = local == 1 ? g1 : (local == 2 ? g2 : (local == 3 ? g3 : sub(sub(neg(g1), g2), g3)));
        neohookean_vertex_terms(p.p0, p.p1, p.p2, p.p3, Dm, d2.y, d2.z, d2.w, gN, projMode, g, H);
    } else {
        const float4 d0 = x.a, d1 = x.b;
        const float Dm[4] = {d0.x, d0.y, d0.z, d0.w};
        float gn0, gn1;
        if (local == 1) { gn0 = Dm[0]; gn1 = Dm[1]; }
        else if (local == 2) { gn0 = Dm[2]; gn1 = Dm[3]; }
        else { gn0 = -Dm[0] - Dm[2]; gn1 = -Dm[1] - Dm[3]; }
        stvk_vertex_terms(p.p0, p.p1, p.p2, Dm, d1.x, d1.y, d1.z, gn0, gn1, projMode, g, H);
    }
}

AVBD_DEV void energy_terms(const DevWorld& w, uint32_t e, int k, int projMode, v3& g, float H[6]) {
    (void)e;
    ElemIn x;
    ElemPos p;
    elem_load(w, k, x);
    elem_pos(w, x, p);
    elem_terms(x, p, projMode, g, H);
}

// E(x - dx) - E(x) of an element for the move of the entry's vertex.
AVBD_DEV float elem_energy_change(const ElemIn& x, const ElemPos& p, v3 dx) {
    const int local = EENT_LOCAL(x.e);
    if (!EENT_TRI(x.e)) {
        const float4 d0 = x.a, d1 = x.b, d2 = x.c;
        m3 Dm;
        Dm.m[0][0] = d0.x; Dm.m[0][1] = d0.y; Dm.m[0][2] = d0.z;
        Dm.m[1][0] = d0.w; Dm.m[1][1] = d1.x; Dm.m[1][2] = d1.y;
        Dm.m[2][0] = d1.z; Dm.m[2][1] = d1.w; Dm.m[2][2] = d2.x;
        const v3 g1 = row(Dm, 0), g2 = row(Dm, 1), g3 = row(Dm, 2);
        const v3 gN = local == 1 ? g1 : (local == 2 ? g2 : (local == 3 ? g3 : sub(sub(neg(g1), g2), g3)));
        return neohookean_energy_change(p.p0, p.p1, p.p2, p.p3, Dm, d2.y, d2.z, d2.w, gN, dx);
    }
    const float4 d0 = x.a, d1 = x.b;
    const float Dm[4] = {d0.x, d0.y, d0.z, d0.w};
    float gn0, gn1;
    if (local == 1) { gn0 = Dm[0]; gn1 = Dm[1]; }
    else if (local == 2) { gn0 = Dm[2]; gn1 = Dm[3]; }
    else { gn0 = -Dm[0] - Dm[2]; gn1 = -Dm[1] - Dm[3]; }
    return stvk_energy_change(p.p0, p.p1, p.p2, Dm, d1.x, d1.y, d1.z, local, gn0, gn1, dx);
}

// One primal block update per body of colour `c` (solver.cpp:392-448 and
// 540-595).  G lanes cooperate on one body: each lane accumulates a strided
// share of the body's incidence entries, the partial systems are summed with
// a butterfly over the lane group, and every lane of the group solves the
// 6x6 system (the result is written by the group's first lane).  G = 1 is the
// sequential order of the reference loop.
//
// The list [lo, hi) is one sweep bucket of the colour.  FEM = true is the
// instantiation for bodies whose only terms are FEM energies: it drops the
// contact/joint/spring code and so needs far fewer registers.
// Block update of FEM-only particles (no forces this step, zero inertia):
// the system is blockdiag(A, 0), so only the translational 3x3 is
// accumulated (9 values through the butterfly instead of 27) and solved with
// the 3x3 pivoted LDL^T, which performs exactly the operations the 6x6 one
// performs on that block; the rotation increment is exactly zero.
// One FEM-only particle: its G lanes (a group within one wave) share the
// particle's energy list.
// Positions of an element from the vertex's neighbour slots in LDS (slot
// NB_MAX holds the vertex itself, which is also a triangle's unused slot 3):
// four unconditional LDS reads, no selects.
AVBD_DEV void elem_pos_lds(const float4* nbp, uint32_t l, ElemPos& p) {
    p.p0 = f3(nbp[l & 0xFFu]);
    p.p1 = f3(nbp[(l >> 8) & 0xFFu]);
    p.p2 = f3(nbp[(l >> 16) & 0xFFu]);
    p.p3 = f3(nbp[l >> 24]);
}

// Order LDS writes of a wave's lanes before other lanes of the same wave
// read them (a lane group never straddles a wave).
AVBD_DEV void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The block solve of a FEM-only particle and its writes (the system is
// blockdiag(A, 0): the rotation increment is exactly zero).
// rot, last_rot and last_pos are the particle's, loaded by the caller.
AVBD_DEV bool identity_bits(float4 q) {
    return __float_as_uint(q.x) == 0x3f800000u && (__float_as_uint(q.y) | __float_as_uint(q.z) | __float_as_uint(q.w)) == 0u;
}
AVBD_DEV void fem_vertex_solve(const DevWorld& w, int b, v3 p, float L[6], float r[3], float4 rot, float4 last_rot,
                               float4 last_pos) {
    float dx[3];
    ldlt3_solve(L, r, dx);
    v3 np = sub(p, v3{dx[0], dx[1], dx[2]});
    if (identity_bits(rot) && identity_bits(last_rot)) {
        // A particle no contact has turned: the update below gives
        // normalize(QuaternionFromDifference(0, -1) * (1, +0, +0, +0)) =
        // (1, +0, +0, +0) again and a rotation difference of +0 from the
        // step start (also the identity), so the same bits are written
        // without that arithmetic (and the unchanged rotation is not stored).
        w.pos[b] = v4(np, 0.f);
        const v3 dl = sub(np, f3(last_pos));
        w.diff0[b] = make_float4(dl.x, dl.y, dl.z, 0.f);
        w.diff1[b] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const qt q = fq(rot);
    const qt lr = fq(last_rot);
    qt dq = quaternion_from_difference(v3{0.f, 0.f, 0.f}, -1.f);
    qt nq = qnormalized(qmul(dq, q));
    w.pos[b] = v4(np, 0.f);
    w.rot[b] = q4(nq);
    v3 dl = sub(np, f3(last_pos));
    v3 da = rotation_difference(nq, lr);
    w.diff0[b] = make_float4(dl.x, dl.y, dl.z, da.x);
    w.diff1[b] = make_float4(da.y, da.z, 0.f, 0.f);
}
AVBD_DEV void fem_vertex_solve(const DevWorld& w, int b, v3 p, float L[6], float r[3]) {
    fem_vertex_solve(w, b, p, L, r, w.rot[b], w.last_rot[b], w.last_pos[b]);
}

// The trust-region terms of one FEM-only block update (scene option
// AVBD_SCENE_TRUST_REGION under ADAPTIVE projection): every lane of the
// group solves the block (L, r: the group's sums) as lane 0 does, sums the
// energy change of its share of the elements for that move, and lane 0 adds
// the inertia term's change and the decrease r . dx / 2 the block model
// predicts to the scene's accumulators (k_rho_update turns them into rho).
template <int G>
AVBD_DEV void fem_trust_terms(const DevWorld& w, int s, int lane, v3 p, v3 ipos, float mass, float dt2, const float L[6],
                              const float r[3], int e0, int e1, const float4* nbp) {
    float dxa[3];
    ldlt3_solve(L, r, dxa);
    const v3 d{dxa[0], dxa[1], dxa[2]};
    float dE = 0.f, dS = 0.f;  // the elements' change and its terms' summed magnitude
    for (int k = e0 + lane; k < e1; k += G) {
        ElemIn x;
        elem_load(w, k, x);
        ElemPos q;
        if (nbp) elem_pos_lds(nbp, w.ev_l[k], q);
        else elem_gather(w, x.v, q);
        const float de = elem_energy_change(x, q, d);
        dE = dE + de;
        dS = dS + fabsf(de);
    }
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
        dE = dE + __shfl_xor(dE, off, G);
        dS = dS + __shfl_xor(dS, off, G);
    }
    if (lane == 0) {
        const float ml = mass / dt2;
        const v3 y = sub(p, ipos);
        const float dEin = (0.5f * ml) * (dot(d, d) - 2.f * dot(d, y));
        const double pred = 0.5 * (((double)r[0] * (double)dxa[0] + (double)r[1] * (double)dxa[1]) + (double)r[2] * (double)dxa[2]);
        atomicAdd(&w.rho_acc[3 * s], -(double)(dEin + dE));
        atomicAdd(&w.rho_acc[3 * s + 1], pred);
        atomicAdd(&w.rho_acc[3 * s + 2], (double)(fabsf(dEin) + dS));
    }
}

template <int G>
AVBD_DEV void primal_fem_vertex_at(const DevWorld& w, int b, int s, int e0, int e1, int nn, int lane, int iter,
                                   int poststab, float4* nbp);

// One FEM-only particle b (its scene, element range and neighbour count
// read here).
template <int G>
AVBD_DEV void primal_fem_vertex(const DevWorld& w, int b, int lane, int iter, int poststab, float4* nbp = nullptr) {
    primal_fem_vertex_at<G>(w, b, w.scene[b], w.e_off[b], w.e_off[b + 1], nbp ? w.nb_n[b] : -1, lane, iter, poststab,
                            nbp);
}

// One FEM-only particle b of scene s with energy entries [e0, e1) and nn
// neighbours in its list (-1: no list, or more than fit).
template <int G>
AVBD_DEV void primal_fem_vertex_at(const DevWorld& w, int b, int s, int e0, int e1, int nn, int lane, int iter,
                                   int poststab, float4* nbp) {
    // What the sweep needs of b is loaded at once, before the scene's run
    // test, so a particle waits on few dependent round trips before its
    // first element: its own records, scene words, neighbour ids and first
    // stream record, then the neighbours' positions.
    if (!nbp) nn = -1;
    const float4 p4 = w.pos[b], vel = w.vel[b], ipos = w.inert_pos[b];
    // the neighbour ids of this lane's slots (clamped into the list; a slot
    // past the list's end later gathers the particle itself, unused)
    constexpr int PER = G >= 4 ? NB_MAX / G : 1;  // narrower groups have no slot
    int nid[PER];
    if constexpr (G >= 4) {
        if (nn >= 0) {
            const int* list = w.nb_id + 2 * (size_t)e0;
            const int jmax = max(nn - 1, 0);
#pragma unroll
            for (int t = 0; t < PER; ++t) nid[t] = list[min(lane + t * G, jmax)];
        }
    }
    const avbd_scene& sc = w.sc[s];
    // the scene's run test from words loaded together (no load waits on
    // the outcome of another)
    const int st_em = w.st[s * ST_WORDS + ST_EMERGENCY], st_oob = w.st[s * ST_WORDS + ST_OOB];
    const int post = sc.post_stabilization, iters = sc.iterations, pmode = sc.projection_mode;
    const uint32_t opts = sc.options;
    const float dt = sc.dt;
    const bool run = (st_em == 0) & (poststab ? (post != 0) & (st_oob == 0) : iter < iters);
    if (!run) return;
    const int pm = proj_mode(w, s, pmode);
    const bool trust = !poststab && pmode == AVBD_PROJ_ADAPTIVE && (opts & AVBD_SCENE_TRUST_REGION);
    const float dt2 = dt * dt;
    float L[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float r[3] = {0.f, 0.f, 0.f};
    const v3 p = f3(p4);
    if (lane == 0) {
        const float ml = vel.w / dt2;
        const v3 dl = sub(p, f3(ipos));
        L[0] = ml; L[2] = ml; L[5] = ml;
        r[0] = ml * dl.x; r[1] = ml * dl.y; r[2] = ml * dl.z;
    }
    // Entries k = e0 + lane, + G, ... in order, software-pipelined: while
    // element k is evaluated, the next element's stream record and vertex
    // positions and the one after's vertex ids are in flight.  Loads are
    // unconditional (indices clamped to the list; a triangle's slot 3 id is
    // its slot 0 id) so that none sits under a branch, which would make every wait
    // drain all loads in flight.
    // With an LDS slot (bucket sweeps, G >= 4): the group loads the vertex's
    // distinct neighbours once and each element reads its positions from
    // there (about one gather per element instead of four).  The positions
    // are the same values the gathers would read: neighbours share an
    // element with the vertex, so none has its colour and none moves in
    // this launch.  All of a lane's gathers are issued before the first
    // LDS write.
    if (G >= 4 && nn >= 0) {
        // the lane's first stream record (clamped into the padded stream)
        // travels with the gathers
        int k = e0 + lane;
        const int last = max(e1 - 1, e0);
        const int kc = min(k, last);
        ElemIn in0;
        in0.e = w.e_ent[kc];
        in0.a = w.ev_a[kc];
        in0.b = w.ev_b[kc];
        in0.c = w.ev_c[kc];
        uint32_t l0 = w.ev_l[kc];
        // every slot of the group's NB_MAX is written (those past the list
        // hold the particle's own position and are never read): with no
        // branch around them the gathers all issue before the first wait
        float4 nv[PER];
#pragma unroll
        for (int t = 0; t < PER; ++t) nv[t] = w.pos[lane + t * G < nn ? nid[t] : b];
#pragma unroll
        for (int t = 0; t < PER; ++t) nbp[lane + t * G] = nv[t];
        if (lane == 0) nbp[NB_MAX] = p4;  // the vertex itself
        wave_lds_sync();
        if (k < e1) {
            bool h0 = true;
            while (h0) {
                const bool h1 = k + G < e1;
                const int k1 = min(k + G, last);
                ElemIn in1;
                in1.e = w.e_ent[k1];
                in1.a = w.ev_a[k1];
                in1.b = w.ev_b[k1];
                in1.c = w.ev_c[k1];
                const uint32_t l1 = w.ev_l[k1];
                ElemPos p0;
                elem_pos_lds(nbp, l0, p0);
                v3 g;
                float H[6];
                elem_terms(in0, p0, pm, g, H);
                r[0] = r[0] + g.x;
                r[1] = r[1] + g.y;
                r[2] = r[2] + g.z;
#pragma unroll
                for (int i = 0; i < 6; ++i) L[i] = L[i] + H[i];
                in0 = in1;
                l0 = l1;
                h0 = h1;
                k += G;
            }
        }
        wave_lds_sync();  // the slot is reused by the group's next vertex
    } else {
    int k = e0 + lane;
    if (k < e1) {
        const int last = e1 - 1;
        ElemIn in0;
        elem_load(w, k, in0);
        int4 ids1 = w.ev_v[min(k + G, last)];
        ElemPos p0;
        elem_gather(w, in0.v, p0);
        bool h0 = true;
        while (h0) {
            const bool h1 = k + G < e1;
            const int k1 = min(k + G, last), k2 = min(k + 2 * G, last);
            ElemIn in1;
            in1.e = w.e_ent[k1];
            in1.a = w.ev_a[k1];
            in1.b = w.ev_b[k1];
            in1.c = w.ev_c[k1];
            in1.v = ids1;
            const int4 ids2 = w.ev_v[k2];
            ElemPos p1;
            elem_gather(w, ids1, p1);
            v3 g;
            float H[6];
            elem_terms(in0, p0, pm, g, H);
            r[0] = r[0] + g.x;
            r[1] = r[1] + g.y;
            r[2] = r[2] + g.z;
#pragma unroll
            for (int i = 0; i < 6; ++i) L[i] = L[i] + H[i];
            in0 = in1;
            p0 = p1;
            ids1 = ids2;
            h0 = h1;
            k += G;
        }
    }
    }
    // the records the solve's writes need, loaded before the lane sums so
    // their round trip overlaps them
    const float4 rot4 = w.rot[b], lrot4 = w.last_rot[b], lpos4 = w.last_pos[b];
    if (G > 1) {
#pragma unroll
        for (int off = 1; off < G; off <<= 1) {
#pragma unroll
            for (int i = 0; i < 6; ++i) L[i] = L[i] + __shfl_xor(L[i], off, G);
#pragma unroll
            for (int i = 0; i < 3; ++i) r[i] = r[i] + __shfl_xor(r[i], off, G);
        }
    }
    if (trust) fem_trust_terms<G>(w, s, lane, p, f3(ipos), vel.w, dt2, L, r, e0, e1, (G >= 4 && nn >= 0) ? nbp : nullptr);
    if (lane == 0) fem_vertex_solve(w, b, p, L, r, rot4, lrot4, lpos4);
}

// Canonical lane sums: the FEM-only update with a result that does not
// depend on G.  The entries are shared out over 32 virtual lanes (virtual
// lane v takes entries e0 + v, + 32, ... in order, lane 0 starting from the
// inertia term) and the 32 partial systems are summed by one fixed tree: a
// butterfly with offsets 16, 8, 4, 2, 1.  A real lane of a G-lane group
// holds the virtual lanes lane + G t (t < T = 32 / G); the tree's first
// log2 T levels pair virtual lanes of the same real lane, so the lane sums
// them itself (leaves in bit-reversed order of t, partial sums kept as in
// pairwise summation), and the last log2 G levels are the shuffles.  Every
// G in {4, 8, 16, 32} performs the same additions in the same tree, so the
// launch width can follow the launch size without changing any result.
template <int T>
AVBD_DEV int bitrev_leaf(int i) {
    if constexpr (T == 8) return ((i & 1) << 2) | (i & 2) | ((i >> 2) & 1);
    else if constexpr (T == 4) return ((i & 1) << 1) | ((i >> 1) & 1);
    else return i;
}

struct FemSum {
    float L[6];
    float r[3];
};
AVBD_DEV void fem_sum_zero(FemSum& s) {
#pragma unroll
    for (int i = 0; i < 6; ++i) s.L[i] = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) s.r[i] = 0.f;
}
AVBD_DEV void fem_sum_add(FemSum& acc, const FemSum& x) {
#pragma unroll
    for (int i = 0; i < 6; ++i) acc.L[i] = x.L[i] + acc.L[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) acc.r[i] = x.r[i] + acc.r[i];
}

template <int G>
AVBD_DEV void primal_fem_vertex_canon(const DevWorld& w, int b, int lane, int iter, int poststab) {
    static_assert(G >= 4 && G <= 32, "canonical sums need 4 to 32 lanes");
    constexpr int T = 32 / G;
    const int s = w.scene[b];
    const avbd_scene& sc = w.sc[s];
    if (w.st[s * ST_WORDS + ST_EMERGENCY]) return;
    if (poststab ? (!sc.post_stabilization || w.st[s * ST_WORDS + ST_OOB]) : iter >= sc.iterations) return;
    const float dt2 = sc.dt * sc.dt;
    const int pm = proj_mode(w, s, sc.projection_mode);
    const v3 p = f3(w.pos[b]);
    const int e0 = w.e_off[b], e1 = w.e_off[b + 1];
    const int last = max(e1 - 1, 0);
    // leaf ti's first entry is e0 + lane + G bitrev(ti); it is loaded one
    // leaf ahead (clamped, unconditionally, as in primal_fem_vertex)
    int k = e0 + lane;
    ElemIn in0;
    elem_load(w, min(k, last), in0);
    ElemPos p0;
    elem_gather(w, in0.v, p0);
    FemSum cur, lv0, lv1, lv2;
    fem_sum_zero(lv0);
    fem_sum_zero(lv1);
    fem_sum_zero(lv2);
#pragma unroll 1
    for (int ti = 0; ti < T; ++ti) {
        const int kn = e0 + lane + G * bitrev_leaf<T>(min(ti + 1, T - 1));
        ElemIn in1;
        elem_load(w, min(kn, last), in1);
        ElemPos p1;
        elem_gather(w, in1.v, p1);
        fem_sum_zero(cur);
        if (lane + G * bitrev_leaf<T>(ti) == 0) {
            const float ml = w.vel[b].w / dt2;
            const v3 dl = sub(p, f3(w.inert_pos[b]));
            cur.L[0] = ml; cur.L[2] = ml; cur.L[5] = ml;
            cur.r[0] = ml * dl.x; cur.r[1] = ml * dl.y; cur.r[2] = ml * dl.z;
        }
        if (k < e1) {
            v3 g;
            float H[6];
            elem_terms(in0, p0, pm, g, H);
            cur.r[0] = cur.r[0] + g.x;
            cur.r[1] = cur.r[1] + g.y;
            cur.r[2] = cur.r[2] + g.z;
#pragma unroll
            for (int i = 0; i < 6; ++i) cur.L[i] = cur.L[i] + H[i];
            // more than 32 entries: the leaf's later ones, in order
            for (int k2 = k + 32; k2 < e1; k2 += 32) {
                ElemIn x;
                elem_load(w, k2, x);
                ElemPos q;
                elem_gather(w, x.v, q);
                elem_terms(x, q, pm, g, H);
                cur.r[0] = cur.r[0] + g.x;
                cur.r[1] = cur.r[1] + g.y;
                cur.r[2] = cur.r[2] + g.z;
#pragma unroll
                for (int i = 0; i < 6; ++i) cur.L[i] = cur.L[i] + H[i];
            }
        }
        // pairwise summation of the leaves: level j holds the sum of the
        // last 2^j leaves until its partner subtree completes
        bool done = true;
        if constexpr (T >= 2) {
            if (!(ti & 1)) { lv0 = cur; done = false; }
            else fem_sum_add(cur, lv0);
        }
        if constexpr (T >= 4) {
            if (done && !(ti & 2)) { lv1 = cur; done = false; }
            else if (done) fem_sum_add(cur, lv1);
        }
        if constexpr (T >= 8) {
            if (done && !(ti & 4)) { lv2 = cur; done = false; }
            else if (done) fem_sum_add(cur, lv2);
        }
        in0 = in1;
        p0 = p1;
        k = kn;
    }
    // cur is the lane's subtree (the loop ends on ti = T - 1, all bits set)
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) {
#pragma unroll
        for (int i = 0; i < 6; ++i) cur.L[i] = cur.L[i] + __shfl_xor(cur.L[i], off, G);
#pragma unroll
        for (int i = 0; i < 3; ++i) cur.r[i] = cur.r[i] + __shfl_xor(cur.r[i], off, G);
    }
    if (!poststab && sc.projection_mode == AVBD_PROJ_ADAPTIVE && (sc.options & AVBD_SCENE_TRUST_REGION))
        fem_trust_terms<G>(w, s, lane, p, f3(w.inert_pos[b]), w.vel[b].w, dt2, cur.L, cur.r, e0, e1, nullptr);
    if (lane == 0) fem_vertex_solve(w, b, p, cur.L, cur.r);
}

// Blocks bid of nblocks stride over the bucket [lo, hi) (canonical lane
// sums: primal_fem_vertex_canon).
template <int G>
AVBD_DEV void primal_fem_body_canon(const DevWorld& w, int lo, int hi, int iter, int poststab, int bid, int nblocks) {
    const int lane = threadIdx.x % G;
    const int gpb = blockDim.x / G;
    for (int base = bid * gpb; base < hi - lo; base += nblocks * gpb) {
        const int item = base + threadIdx.x / G;
        if (item >= hi - lo) continue;
        primal_fem_vertex_canon<G>(w, w.col_list[lo + item], lane, iter, poststab);
    }
}

// Blocks bid of nblocks stride over the bucket [lo, hi).
template <int G>
AVBD_DEV void primal_fem_body(const DevWorld& w, int lo, int hi, int iter, int poststab, int bid, int nblocks) {
    const int lane = threadIdx.x % G;
    const int gpb = blockDim.x / G;
    // neighbour position slots: NB_MAX + 1 per lane group (the vertex's
    // neighbours, then the vertex itself; 33 KiB at G = 4 and 256 threads);
    // narrower groups would need too much LDS and gather
    constexpr int kGroups = G >= 4 ? 256 / G : 1;
    __shared__ float4 nb_lds[kGroups * (NB_MAX + 1)];
    float4* slot = (G >= 4 && w.use_nb) ? nb_lds + (threadIdx.x / G) * (NB_MAX + 1) : nullptr;
    for (int base = bid * gpb; base < hi - lo; base += nblocks * gpb) {
        const int item = base + threadIdx.x / G;
        if (item >= hi - lo) continue;  // lane groups never straddle a wave: the group leaves together
        const int4 ent = w.col_fem[lo + item];
        primal_fem_vertex_at<G>(w, ent.x, ent.z, ent.y, ent.y + (ent.w >> 8), (ent.w & 0xFF) - 1, lane, iter, poststab,
                                slot);
    }
}

// One body with forces (and possibly energies): its G lanes (a group within
// one wave) share the body's incidence list.  valid is false for the padding
// groups of a pass; they take part in nothing.
template <int G>
AVBD_DEV void primal_rigid(const DevWorld& w, int b, bool valid, int iter, int poststab) {
    const int lane = threadIdx.x % G;
    bool run = valid;
    int s = 0;
    if (run) {
        s = w.scene[b];
        const avbd_scene& sc0 = w.sc[s];
        if (w.st[s * ST_WORDS + ST_EMERGENCY]) run = false;
        // a scene with bodies out of bounds post-stabilises after their removal
        if (poststab) { if (!sc0.post_stabilization || w.st[s * ST_WORDS + ST_OOB]) run = false; }
        else if (iter >= sc0.iterations) run = false;
    }
    // all lanes of a group agree on `run`; groups never straddle a wave
    if (run) {
        const avbd_scene& sc = w.sc[s];
        const float alpha = poststab ? 0.f : sc.alpha;
        const float dt = sc.dt;
        const float dt2 = dt * dt;
        float dSelf[6];
        load_diff(w, b, dSelf);
        Acc A;
        acc_zero(A);
        const qt lr = fq(w.last_rot[b]);
        const float4 vm = w.vel[b];
        const v3 I = f3(w.inertia[b]);
        const v3 p = f3(w.pos[b]);
        const qt q = fq(w.rot[b]);
        if (lane == 0) {
            // lhs = M / dt^2 ; rhs = lhs * (x - y, rotdiff(q, q_y))
            const m3 R = rotmat(lr);
            float Iw[3][3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    Iw[i][j] = ((R.m[i][0] * I.x) * R.m[j][0] + (R.m[i][1] * I.y) * R.m[j][1]) +
                               (R.m[i][2] * I.z) * R.m[j][2];
            const float ml = vm.w / dt2;
            v3 dl = sub(p, f3(w.inert_pos[b]));
            v3 da = rotation_difference(q, fq(w.inert_rot[b]));
            A.L[tri_idx(0, 0)] = ml; A.L[tri_idx(1, 1)] = ml; A.L[tri_idx(2, 2)] = ml;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j <= i; ++j) A.L[tri_idx(3 + i, 3 + j)] = Iw[i][j] / dt2;
            A.r[0] = ml * dl.x; A.r[1] = ml * dl.y; A.r[2] = ml * dl.z;
#pragma unroll
            for (int i = 0; i < 3; ++i)
                A.r[3 + i] = ((Iw[i][0] / dt2) * da.x + (Iw[i][1] / dt2) * da.y) + (Iw[i][2] / dt2) * da.z;
        }
        {
            const int f0 = w.f_off[b], f1 = w.f_off[b + 1];
            for (int k = f0 + lane; k < f1; k += G) accumulate_force(w, w.f_ent[k], alpha, dSelf, A);
        }
        const int e0 = w.e_off[b], e1 = w.e_off[b + 1];
        const int pm = proj_mode(w, s, sc.projection_mode);
        for (int k = e0 + lane; k < e1; k += G) accumulate_energy(w, k, pm, A);
        if (G > 1) {
#pragma unroll
            for (int off = 1; off < G; off <<= 1) {
#pragma unroll
                for (int i = 0; i < 21; ++i) A.L[i] = A.L[i] + __shfl_xor(A.L[i], off, G);
#pragma unroll
                for (int i = 0; i < 6; ++i) A.r[i] = A.r[i] + __shfl_xor(A.r[i], off, G);
            }
        }
        float dx[6];
        ldlt6_solve(A.L, A.r, dx);
        if (lane == 0) {
            v3 np = sub(p, v3{dx[0], dx[1], dx[2]});
            qt dq = quaternion_from_difference(v3{dx[3], dx[4], dx[5]}, -1.f);
            qt nq = qnormalized(qmul(dq, q));
            w.pos[b] = v4(np, 0.f);
            w.rot[b] = q4(nq);
            v3 dl = sub(np, f3(w.last_pos[b]));
            v3 da = rotation_difference(nq, lr);
            w.diff0[b] = make_float4(dl.x, dl.y, dl.z, da.x);
            w.diff1[b] = make_float4(da.y, da.z, 0.f, 0.f);
        }
    }
}

template <int G>
AVBD_DEV void primal_rigid_body(const DevWorld& w, int lo, int hi, int iter, int poststab, int bid, int nblocks) {
    const int groups_per_block = blockDim.x / G;
    for (int base = bid * groups_per_block; base < hi - lo; base += nblocks * groups_per_block) {
        const int item = base + threadIdx.x / G;
        const bool valid = item < hi - lo;
        primal_rigid<G>(w, valid ? w.col_list[lo + item] : 0, valid, iter, poststab);
    }
}

template <int G, bool FEM>
__global__ __launch_bounds__(256) void k_primal(DevWorld w, int bucket, int iter, int poststab) {
    const int lo = w.col_off[bucket], hi = w.col_off[bucket + 1];
    if constexpr (FEM)
        primal_fem_body<G>(w, lo, hi, iter, poststab, blockIdx.x, gridDim.x);
    else
        primal_rigid_body<G>(w, lo, hi, iter, poststab, blockIdx.x, gridDim.x);
}

// The FEM-only bucket sweep with its own occupancy target: at least
// AVBD_FEM_MIN_WAVES waves per SIMD, i.e. a budget of 512 / waves VGPRs
// that the compiler keeps (4 waves: 128 VGPRs, met without spills; left
// free the kernel takes 140 and runs 3 waves).  Measured on the 4096-scene
// mix: 3 waves 392.7 us per launch, 4 waves 378.9 us, 5 waves (96 VGPRs,
// 39 spilled) 469.5 us.  Its arithmetic is k_primal<G, true>'s.
#ifndef AVBD_FEM_MIN_WAVES
#define AVBD_FEM_MIN_WAVES 4
#endif
template <int G>
__global__ __launch_bounds__(256, AVBD_FEM_MIN_WAVES) void k_primal_fem(DevWorld w, int bucket, int iter, int poststab) {
    primal_fem_body<G>(w, w.col_off[bucket], w.col_off[bucket + 1], iter, poststab, blockIdx.x, gridDim.x);
}

// Both buckets of a colour in one launch: blocks [0, rigid_blocks) take the
// bodies with forces (GR lanes each), the rest the FEM-only particles (GF
// lanes).  The two sets of bodies are disjoint and neither reads what the
// other writes in this colour, as with the two concurrent launches it
// replaces; for small buckets it saves the fork and join between streams.
template <int GR, int GF, bool CANON>
__global__ __launch_bounds__(256) void k_primal_pair(DevWorld w, int colour, int iter, int poststab, int rigid_blocks) {
    const int bid = blockIdx.x;
    const int lo = w.col_off[2 * colour + 1], hi = w.col_off[2 * colour + 2], nb = (int)gridDim.x - rigid_blocks;
    if (bid < rigid_blocks)
        primal_rigid_body<GR>(w, w.col_off[2 * colour], w.col_off[2 * colour + 1], iter, poststab, bid, rigid_blocks);
    else if constexpr (CANON)
        primal_fem_body_canon<GF>(w, lo, hi, iter, poststab, bid - rigid_blocks, nb);
    else
        primal_fem_body<GF>(w, lo, hi, iter, poststab, bid - rigid_blocks, nb);
}

// FEM-only bucket with canonical lane sums (result independent of G).
template <int G>
__global__ __launch_bounds__(256) void k_primal_canon(DevWorld w, int bucket, int iter, int poststab) {
    primal_fem_body_canon<G>(w, w.col_off[bucket], w.col_off[bucket + 1], iter, poststab, blockIdx.x, gridDim.x);
}

#define AVBD_PAIR_INST(GF)                                                              \
    template __global__ void k_primal_pair<4, GF, false>(DevWorld, int, int, int, int); \
    template __global__ void k_primal_pair<4, GF, true>(DevWorld, int, int, int, int);  \
    template __global__ void k_primal_canon<GF>(DevWorld, int, int, int);
AVBD_PAIR_INST(4)
AVBD_PAIR_INST(8)
AVBD_PAIR_INST(16)
AVBD_PAIR_INST(32)

#define AVBD_PRIMAL_INST(G)                                                        \
    template __global__ void k_primal<G, false>(DevWorld, int, int, int); \
    template __global__ void k_primal<G, true>(DevWorld, int, int, int);
AVBD_PRIMAL_INST(1)
AVBD_PRIMAL_INST(2)
AVBD_PRIMAL_INST(4)
AVBD_PRIMAL_INST(8)
AVBD_PRIMAL_INST(16)
AVBD_PRIMAL_INST(32)
template __global__ void k_primal_fem<1>(DevWorld, int, int, int);
template __global__ void k_primal_fem<2>(DevWorld, int, int, int);
template __global__ void k_primal_fem<4>(DevWorld, int, int, int);
template __global__ void k_primal_fem<8>(DevWorld, int, int, int);
template __global__ void k_primal_fem<16>(DevWorld, int, int, int);
template __global__ void k_primal_fem<32>(DevWorld, int, int, int);

// Dual / penalty update (solver.cpp:457-489), one thread per contact, joint
// or spring.  A fracture disables the force and queues its bodies so that
// k_fracture_reset clears their manifolds (order-independent: a cleared
// manifold stays at lambda = k = 0 for the rest of the sweep).
AVBD_DEV void dual_contact(const DevWorld& w, int m, int c, int iter) {
    if (c >= w.m.n[m]) return;
    const int a = w.m.a[m], b = w.m.b[m];
    const int s = w.scene[a];
    if (!scene_runs(w, s)) return;
    const avbd_scene& sc = w.sc[s];
    if (iter >= sc.iterations) return;
    float dA[6], dB[6];
    load_diff(w, a, dA);
    load_diff(w, b, dB);
    ContactRows R;
    contact_rows(w, m, c, sc.alpha, dA, dB, R);
    float tl = sqrtf(R.lam[1] * R.lam[1] + R.lam[2] * R.lam[2]);
    const float4 h1 = w.m.hot[((size_t)m * 8 + c) * CONTACT_HOT_F4 + 1];
    const float4 h2 = w.m.hot[((size_t)m * 8 + c) * CONTACT_HOT_F4 + 2];
    float tc = sqrtf(h1.w * h1.w + h2.w * h2.w);
    int stick = (tl < R.bound) && (tc < 0.01f);
    float lam[3], k[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float lo = r == 0 ? -INFINITY : -R.bound;
        float hi = r == 0 ? 0.f : R.bound;
        float l = R.lam[r] + R.k[r] * R.C[r];
        if (l < lo) l = lo;
        else if (l > hi) l = hi;
        float kk = R.k[r];
        if (l > lo && l < hi) kk = fminf(kk + sc.beta * fabsf(R.C[r]), 1000000000.0f);
        lam[r] = l;
        k[r] = kk;
    }
    float4* h = w.m.hot + ((size_t)m * 8 + c) * CONTACT_HOT_F4;
    h[3].w = lam[0];
    h[4].w = lam[1];
    h[5] = make_float4(lam[2], k[0], k[1], k[2]);
    w.m.cold[((size_t)m * 8 + c) * CONTACT_COLD_F4].y = (float)stick;
}

AVBD_DEV void dual_joint(const DevWorld& w, int j, int iter) {
    if (w.j_flags[j] & FORCE_DEAD) return;
    const int s = w.scene[w.j_b[j]];
    if (!scene_runs(w, s)) return;
    const avbd_scene& sc = w.sc[s];
    if (iter >= sc.iterations) return;
    JointRows R;
    joint_rows(w, j, true, sc.alpha, R);
    float* rows = w.j_row + j * 36;
    for (int r = 0; r < 6; ++r) {
        float* row = rows + r * 6;
        const float stiff = row[0], frac = row[1], fmn = row[2], fmx = row[3];
        float l = (isinf(stiff) ? row[4] : 0.f) + row[5] * R.C[r];
        if (l < fmn) l = fmn;
        else if (l > fmx) l = fmx;
        row[4] = l;
        if (fabsf(l) >= frac) {
            for (int q = 0; q < 6; ++q) { rows[q * 6 + 4] = 0.f; rows[q * 6 + 5] = 0.f; }
            w.j_flags[j] |= AVBD_FORCE_DISABLED;
            int slot = atomicAdd(&w.g[G_FRACTURES], 2);
            if (slot + 1 < 2 * (w.J + w.SP) + 2) {
                w.fracture_bodies[slot] = w.j_a[j];
                w.fracture_bodies[slot + 1] = w.j_b[j];
            }
            break;
        }
        if (l > fmn && l < fmx) row[5] = fminf(row[5] + sc.beta * fabsf(R.C[r]), fminf(stiff, 1000000000.0f));
    }
}

AVBD_DEV void dual_spring(const DevWorld& w, int sp, int iter) {
    if (w.s_flags[sp] & FORCE_DEAD) return;
    const int s = w.scene[w.s_b[sp]];
    if (!scene_runs(w, s)) return;
    const avbd_scene& sc = w.sc[s];
    if (iter >= sc.iterations) return;
    float* row = w.s_row + sp * 6;
    const float stiff = row[0], frac = row[1], fmn = row[2], fmx = row[3];
    float C = spring_C(w, sp);
    float l = (isinf(stiff) ? row[4] : 0.f) + row[5] * C;
    if (l < fmn) l = fmn;
    else if (l > fmx) l = fmx;
    row[4] = l;
    if (fabsf(l) >= frac) {
        row[4] = 0.f;
        row[5] = 0.f;
        w.s_flags[sp] |= AVBD_FORCE_DISABLED;
        int slot = atomicAdd(&w.g[G_FRACTURES], 2);
        if (slot + 1 < 2 * (w.J + w.SP) + 2) {
            w.fracture_bodies[slot] = w.s_a[sp];
            w.fracture_bodies[slot + 1] = w.s_b[sp];
        }
        return;
    }
    if (l > fmn && l < fmx) row[5] = fminf(row[5] + sc.beta * fabsf(C), fminf(stiff, 1000000000.0f));
}

__global__ void k_dual(DevWorld w, int iter) {
    const int M = w.g[G_MANIFOLDS];
    const int total = M * 8 + w.J + w.SP;
    GRID_STRIDE(t, total) {
        if (t < M * 8) {
            const int m = t >> 3;
            if (scene_fused(w, w.scene[w.m.a[m]])) continue;
            dual_contact(w, m, t & 7, iter);
        } else if (t < M * 8 + w.J) {
            const int j = t - M * 8;
            if (scene_fused(w, w.scene[w.j_b[j]])) continue;
            dual_joint(w, j, iter);
        } else {
            const int sp = t - M * 8 - w.J;
            if (scene_fused(w, w.scene[w.s_b[sp]])) continue;
            dual_spring(w, sp, iter);
        }
    }
}

// Clear lambda and penalty of every manifold touching a fractured force's
// bodies (Force::ResetValues via solver.cpp:473-478), then clear the queue.
// Step 5.2 (solver.cpp:451-455, empty in the reference) for the scenes
// with AVBD_SCENE_TRUST_REGION: rho = actual / predicted decrease of the
// sweep that just ended, then the sums restart.  Both sums cancel terms of
// the summed magnitude acc[2]; a sweep whose predicted decrease is within
// 3e-5 of it (a nearly converged solve: its ratio is FP32 rounding noise)
// and a scene with no FEM-only update keep their rho.
__global__ void k_rho_update(DevWorld w) {
    GRID_STRIDE(s, w.S) {
        double* acc = w.rho_acc + 3 * (size_t)s;
        const double pred = acc[1];
        if (pred > 3e-5 * acc[2]) w.rho[s] = (float)(acc[0] / pred);
        acc[0] = 0.0;
        acc[1] = 0.0;
        acc[2] = 0.0;
    }
}

__global__ void k_fracture_reset(DevWorld w) {
    const int nf = min(w.g[G_FRACTURES], 2 * (w.J + w.SP) + 2);
    if (nf == 0) return;
    GRID_STRIDE(t, nf) {
        const int b = w.fracture_bodies[t];
        if (b < 0) continue;
        for (int k = w.f_off[b]; k < w.f_off[b + 1]; ++k) {
            uint32_t e = w.f_ent[k];
            if (ENT_KIND(e) != ENT_CONTACT) continue;
            const int m = ENT_INDEX(e) >> 3;
            const int n = w.m.n[m];
            for (int c = 0; c < n; ++c) {
                float4* h = w.m.hot + ((size_t)m * 8 + c) * CONTACT_HOT_F4;
                h[3].w = 0.f;
                h[4].w = 0.f;
                h[5] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
    }
}

// ------------------------------------------------------------------ phase 6
// Velocities from the displacement of the step, bounds and explosion checks
// (solver.cpp:493-530, CheckExplosion 182-196).
AVBD_DEV void velocity_body(const DevWorld& w, int b, int s) {
    {
        const avbd_scene& sc = w.sc[s];
        const uint32_t fl = w.flags[b];
        float4 v4v = w.vel[b];
        if (!(fl & AVBD_BODY_STATIC)) {
            float4 av4 = w.angvel[b];
            w.pvel[b] = make_float4(v4v.x, v4v.y, v4v.z, 0.f);
            w.pangvel[b] = make_float4(av4.x, av4.y, av4.z, 0.f);
            const float4 d0 = w.diff0[b], d1 = w.diff1[b];
            const float dt = sc.dt;
            v4v = make_float4(d0.x / dt, d0.y / dt, d0.z / dt, v4v.w);
            const float4 dr = w.drag[b];
            if (dr.w != 0.f) {  // a dragged body: its added velocity, once (solver.cpp:506-510)
                v4v = make_float4(v4v.x + dr.x, v4v.y + dr.y, v4v.z + dr.z, v4v.w);
                w.drag[b] = make_float4(0.f, 0.f, 0.f, dr.w);
            }
            w.vel[b] = v4v;
            if (!(fl & AVBD_BODY_PARTICLE)) w.angvel[b] = make_float4(d0.w / dt, d1.x / dt, d1.y / dt, av4.w);
            const float4 p = w.pos[b];
            if (p.x < sc.bounds_min[0] || p.x > sc.bounds_max[0] || p.y < sc.bounds_min[1] || p.y > sc.bounds_max[1] ||
                p.z < sc.bounds_min[2] || p.z > sc.bounds_max[2]) {
                atomicAdd(&w.st[s * ST_WORDS + ST_OOB], 1);
                w.g[G_ANY_OOB] = 1;
            }
        }
        const float mx = 500.f;
        if (v4v.x > mx || v4v.y > mx || v4v.z > mx || v4v.x < -mx || v4v.y < -mx || v4v.z < -mx) {
            w.st[s * ST_WORDS + ST_EMERGENCY] = 1;
            w.g[G_ANY_EMERGENCY] = 1;
        }
    }
}

__global__ void k_velocity(DevWorld w) {
    GRID_STRIDE(b, w.N) {
        const int s = w.scene[b];
        if (!scene_runs(w, s) || scene_fused(w, s)) continue;
        velocity_body(w, b, s);
    }
}

// ------------------------------------------------------------------ phase 7
// Stick flags as left by the last ComputeConstraints of the step: with
// post-stabilisation that is the post-stab sweep, using the final lambda.
AVBD_DEV void stick_contact(const DevWorld& w, int m, int c) {
    const int a = w.m.a[m], b = w.m.b[m];
    if (is_static(w, a) && is_static(w, b)) return;
    const float4* h = w.m.hot + ((size_t)m * 8 + c) * CONTACT_HOT_F4;
    const float4 h1 = h[1], h2 = h[2], h3 = h[3], h4 = h[4], h5 = h[5];
    float bound = fabsf(h3.w) * w.m.fric[m];
    float tl = sqrtf(h4.w * h4.w + h5.x * h5.x);
    float tc = sqrtf(h1.w * h1.w + h2.w * h2.w);
    w.m.cold[((size_t)m * 8 + c) * CONTACT_COLD_F4].y = (float)((tl < bound) && (tc < 0.01f));
}

__global__ void k_stick(DevWorld w) {
    const int M = w.g[G_MANIFOLDS];
    GRID_STRIDE(t, M * 8) {
        const int m = t >> 3, c = t & 7;
        if (c >= w.m.n[m]) continue;
        const int s = w.scene[w.m.a[m]];
        if (!scene_runs(w, s) || scene_fused(w, s)) continue;
        if (!w.sc[s].post_stabilization || w.st[s * ST_WORDS + ST_OOB]) continue;
        stick_contact(w, m, c);
    }
}

// ------------------------------------------------ per-scene step (small scenes)
// Sweep kind of body b in colour col for the per-scene kernel: 0 = block
// update with forces, 1 = FEM-only particle, -1 = not swept (the bucket
// rule of sweep_bucket without the scene tests).
AVBD_DEV int sweep_kind(const DevWorld& w, int b, int col) {
    if (is_static(w, b) || w.colour[b] != col) return -1;
    const float4 I = w.inertia[b];
    const bool fem_only = w.f_off[b + 1] == w.f_off[b] && w.e_off[b + 1] > w.e_off[b] && I.x == 0.f && I.y == 0.f &&
                          I.z == 0.f;
    return fem_only ? 1 : 0;
}

// Phases 4-7 of one small scene in one workgroup: prediction, the
// iterations (per colour the block updates of the scene's bodies, then the
// dual of its forces), velocities with the bounds and explosion checks, and
// post-stabilisation with the stick flags.  The per-body and per-force
// functions are those of the bucket launches and run in the same order
// (colour-major; bodies of one colour are independent), with workgroup
// barriers in place of launch boundaries, so the result is bit-identical.
// Each force is visited once: a contact point from its body A's incidence
// entry, a joint or spring from its body B's.  Eligible scenes have no breakable joint or spring (a
// fracture resets other bodies' contacts through the global queue).
// mode 1: post-stabilisation only (after the bounds removal).
// Phases 4-7 of one scene in one workgroup (see k_scene_step).
template <int G>
AVBD_DEV void scene_step_body(const DevWorld& w, int s, int mode) {
    const avbd_scene& sc = w.sc[s];
    const int a = w.sc_off[s], e = w.sc_off[s + 1];
    const int ncol = w.st[s * ST_WORDS + ST_NCOL];
    const int gpb = blockDim.x / G;
    auto sweep = [&](int iter, int poststab) {
        for (int col = 0; col < ncol; ++col) {
            for (int base = a; base < e; base += gpb) {
                const int b = base + (int)threadIdx.x / G;  // one lane group per body
                const int k = b < e ? sweep_kind(w, b, col) : -1;
                if (k == 1) primal_fem_vertex<G>(w, b, (int)threadIdx.x % G, iter, poststab);
                else if (k == 0) primal_rigid<G>(w, b, true, iter, poststab);
            }
            __syncthreads();
        }
    };
    // the scene's incidence entries are one contiguous range (bodies of a
    // scene are contiguous): one thread per entry
    const int q0 = w.f_off[a], q1 = w.f_off[e];
    auto each_force = [&](auto f) {
        for (int q = q0 + (int)threadIdx.x; q < q1; q += blockDim.x) {
            const uint32_t ent = w.f_ent[q];
            const bool contact = ENT_KIND(ent) == ENT_CONTACT;
            if (contact == (ENT_SIDE_B(ent) != 0)) continue;
            f(ENT_KIND(ent), ENT_INDEX(ent));
        }
    };
    if (mode == 0) {
        for (int b = a + (int)threadIdx.x; b < e; b += blockDim.x) predict_body(w, b, s);
        __syncthreads();
        for (int it = 0; it < sc.iterations; ++it) {
            sweep(it, 0);
            each_force([&](uint32_t kind, uint32_t idx) {
                if (kind == ENT_CONTACT) dual_contact(w, (int)(idx >> 3), (int)(idx & 7u), it);
                else if (kind == ENT_JOINT) dual_joint(w, (int)idx, it);
                else dual_spring(w, (int)idx, it);
            });
            __syncthreads();
        }
        for (int b = a + (int)threadIdx.x; b < e; b += blockDim.x) velocity_body(w, b, s);
        __syncthreads();
    }
    // post-stabilisation unless the bounds or explosion check fired (flags
    // set by other threads with atomics: read at device scope)
    const int* st = w.st + s * ST_WORDS;
    if (!sc.post_stabilization || __atomic_load_n(&st[ST_OOB], __ATOMIC_RELAXED) != 0 ||
        __atomic_load_n(&st[ST_EMERGENCY], __ATOMIC_RELAXED) != 0)
        return;
    sweep(-1, 1);
    each_force([&](uint32_t kind, uint32_t idx) {
        if (kind == ENT_CONTACT) stick_contact(w, (int)(idx >> 3), (int)(idx & 7u));
    });
}

// `first` rebased so that p[first + i] is lds[i]: the address arithmetic is
// done on integers, and loads through the result are flat accesses that the
// hardware routes to LDS.
template <class T>
AVBD_DEV T* lds_alias(T* lds, int first) {
    return reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(lds) - (uintptr_t)first * sizeof(T));
}

// The per-scene kernel: one workgroup runs a scene's phases 4-7 with
// workgroup barriers between colours instead of launch boundaries.  A
// scene whose bodies and incidence lists fit in lds_bytes works on an LDS
// copy of them: the per-body and per-force functions run unchanged on a
// world whose body arrays and incidence range point into LDS (contact,
// joint and spring records stay in global memory), so each colour's
// dependent loads are LDS latencies; the bodies' state is written back at
// the end.  Every force of a scene joins bodies of that scene, so no access
// leaves the copied ranges.
template <int G>
__global__ __launch_bounds__(256) void k_scene_step(DevWorld w, const int* scenes, int mode, int lds_bytes) {
    extern __shared__ float4 scene_lds[];
    const int s = scenes[blockIdx.x];
    if (!scene_runs(w, s) || !scene_fused(w, s)) return;
    const int a = w.sc_off[s], e = w.sc_off[s + 1];
    const int nb = e - a;
    const int q0 = w.f_off[a], nf = w.f_off[e] - q0;
    constexpr int kF4 = 15;  // float4 body arrays copied
    const size_t need = (size_t)nb * (kF4 * sizeof(float4) + 3 * sizeof(int)) + (size_t)(nb + 1) * sizeof(int) +
                        (size_t)nf * sizeof(uint32_t);
    if (need > (size_t)lds_bytes) {
        scene_step_body<G>(w, s, mode);
        return;
    }
    DevWorld l = w;
    float4* const* src4[kF4] = {&w.pos, &w.rot, &w.diff0, &w.diff1, &w.vel, &w.angvel, &w.pvel, &w.pangvel,
                                &w.last_pos, &w.last_rot, &w.inert_pos, &w.inert_rot, &w.drag, &w.scale, &w.inertia};
    float4** dst4[kF4] = {&l.pos, &l.rot, &l.diff0, &l.diff1, &l.vel, &l.angvel, &l.pvel, &l.pangvel,
                          &l.last_pos, &l.last_rot, &l.inert_pos, &l.inert_rot, &l.drag, &l.scale, &l.inertia};
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
#pragma unroll
    for (int f = 0; f < kF4; ++f) {
        float4* dst = scene_lds + (size_t)f * nb;
        const float4* src = *src4[f];
        for (int i = tid; i < nb; i += nt) dst[i] = src[a + i];
        *dst4[f] = lds_alias(dst, a);
    }
    uint32_t* lflags = reinterpret_cast<uint32_t*>(scene_lds + (size_t)kF4 * nb);
    int* lscene = reinterpret_cast<int*>(lflags + nb);
    int* lcolour = lscene + nb;
    int* lfoff = lcolour + nb;
    uint32_t* lfent = reinterpret_cast<uint32_t*>(lfoff + nb + 1);
    for (int i = tid; i < nb; i += nt) {
        lflags[i] = w.flags[a + i];
        lscene[i] = w.scene[a + i];
        lcolour[i] = w.colour[a + i];
    }
    for (int i = tid; i <= nb; i += nt) lfoff[i] = w.f_off[a + i];
    for (int i = tid; i < nf; i += nt) lfent[i] = w.f_ent[q0 + i];
    l.flags = lds_alias(lflags, a);
    l.scene = lds_alias(lscene, a);
    l.colour = lds_alias(lcolour, a);
    l.f_off = lds_alias(lfoff, a);
    l.f_ent = lds_alias(lfent, q0);
    __syncthreads();
    scene_step_body<G>(l, s, mode);
    __syncthreads();
    // the state the phases write (scale and inertia are read only)
#pragma unroll
    for (int f = 0; f < kF4 - 2; ++f) {
        const float4* src = scene_lds + (size_t)f * nb;
        float4* dst = *src4[f];
        for (int i = tid; i < nb; i += nt) dst[a + i] = src[i];
    }
}

template __global__ void k_scene_step<1>(DevWorld, const int*, int, int);
template __global__ void k_scene_step<2>(DevWorld, const int*, int, int);
template __global__ void k_scene_step<4>(DevWorld, const int*, int, int);
template __global__ void k_scene_step<8>(DevWorld, const int*, int, int);
template __global__ void k_scene_step<16>(DevWorld, const int*, int, int);
template __global__ void k_scene_step<32>(DevWorld, const int*, int, int);

// ------------------------------------------------ out-of-bounds removal
// Solver::RemoveBody for the bodies that left the world bounds
// (solver.cpp:520-534, 104-116): the host computes the surviving order inside
// the affected scenes, the moves (new index <- old index, gathered into a
// buffer, then scattered) and the renumbering (BodyRemap), and these kernels
// apply them to the device state.
__global__ void k_gather_f4(float4* dst, const float4* src, const int* perm, int n) {
    GRID_STRIDE(i, n) dst[i] = src[perm[i]];
}
__global__ void k_gather_u32(uint32_t* dst, const uint32_t* src, const int* perm, int n) {
    GRID_STRIDE(i, n) dst[i] = src[perm[i]];
}
__global__ void k_scatter_f4(float4* dst, const float4* src, const int* to, int n) {
    GRID_STRIDE(i, n) dst[to[i]] = src[i];
}
__global__ void k_scatter_u32(uint32_t* dst, const uint32_t* src, const int* to, int n) {
    GRID_STRIDE(i, n) dst[to[i]] = src[i];
}

// New index of body i (any index; -1 for the world anchor stays -1).  The
// scene of i is read after the move, which keeps every body in its scene's
// range, so scene[i] is the scene of the old index as well.
AVBD_DEV int remap_body(const DevWorld& w, BodyRemap rm, int i) {
    if (i < 0) return -1;
    const int s = w.scene[i];
    const int o = rm.aoff[s];
    return o < 0 ? i : rm.lmap[o + (i - w.sc_off[s])];
}

// Joints and springs: endpoints renumbered; a force that loses a body is
// removed with it (~Mesh deletes the body's forces, geometry.cpp:11-24).
__global__ void k_remap_forces(DevWorld w, BodyRemap rm) {
    GRID_STRIDE(t, w.J + w.SP) {
        int *pa, *pb;
        uint32_t* fl;
        if (t < w.J) { pa = &w.j_a[t]; pb = &w.j_b[t]; fl = &w.j_flags[t]; }
        else { pa = &w.s_a[t - w.J]; pb = &w.s_b[t - w.J]; fl = &w.s_flags[t - w.J]; }
        if (*fl & FORCE_DEAD) continue;
        const int a = *pa, b = *pb;
        const int na = remap_body(w, rm, a), nb = remap_body(w, rm, b);
        if ((a >= 0 && na < 0) || nb < 0) {
            *fl |= FORCE_DEAD;
            continue;
        }
        *pa = na;
        *pb = nb;
    }
}

// Removed bodies left as dead slots: static (no sweep, no colour, no
// bounds or explosion check), a negative radius so no broad-phase test
// passes, at rest.  Mass and friction (the w components) are kept.
__global__ void k_kill_bodies(DevWorld w, const int* list, int n) {
    GRID_STRIDE(k, n) {
        const int b = list[k];
        w.flags[b] |= AVBD_BODY_STATIC;
        const float4 sc = w.scale[b];
        w.scale[b] = make_float4(sc.x, sc.y, sc.z, -1e30f);
        const float4 v = w.vel[b], av = w.angvel[b];
        w.vel[b] = make_float4(0.f, 0.f, 0.f, v.w);
        w.angvel[b] = make_float4(0.f, 0.f, 0.f, av.w);
        w.pvel[b] = make_float4(0.f, 0.f, 0.f, 0.f);
        w.pangvel[b] = make_float4(0.f, 0.f, 0.f, 0.f);
        w.drag[b] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// The element stream in incidence order from the element records: entry k
// of e_ent (kind, local vertex, element index) takes its element's rest-shape
// records and vertex ids.  Run once per upload over the n stream entries.
__global__ void k_element_stream(DevWorld w, int n) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t e = w.e_ent[k];
        const uint32_t idx = e & 0x1FFFFFFFu;
        if (EENT_TRI(e)) {
            w.ev_a[k] = w.r_d[(size_t)idx * 2 + 0];
            w.ev_b[k] = w.r_d[(size_t)idx * 2 + 1];
            w.ev_c[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            int4 v = w.r_v[idx];
            v.w = v.x;  // the unused slot 3 repeats a vertex of the element
            w.ev_v[k] = v;
        } else {
            w.ev_a[k] = w.t_d[(size_t)idx * 3 + 0];
            w.ev_b[k] = w.t_d[(size_t)idx * 3 + 1];
            w.ev_c[k] = w.t_d[(size_t)idx * 3 + 2];
            w.ev_v[k] = w.t_v[idx];
        }
    }
}

// Neighbour lists of the vertices [b0, b1) from their element streams (run
// after the stream is built or rebuilt): one thread per vertex walks its
// entries in order, keeps the distinct other vertices in first-appearance
// order at nb_id[2 e_off[b] ...] (room for two per entry, at most NB_MAX),
// and writes each entry's slot indices into ev_l (NB_MAX for the vertex
// itself, whose position the sweep keeps in the slot after the list).  A vertex whose
// neighbours do not fit gets nb_n = -1 and is swept by element ids.
__global__ void k_nb_build(DevWorld w, int b0, int b1) {
    GRID_STRIDE(t, b1 - b0) {
        const int b = b0 + t;
        const int e0 = w.e_off[b], e1 = w.e_off[b + 1];
        const int cap = min(2 * (e1 - e0), NB_MAX);
        int* list = w.nb_id + 2 * (size_t)e0;
        int n = 0;
        bool fits = true;
        for (int k = e0; k < e1 && fits; ++k) {
            const uint32_t e = w.e_ent[k];
            const int4 v = w.ev_v[k];
            const int local = (int)EENT_LOCAL(e);
            const int slots = EENT_TRI(e) ? 3 : 4;
            const int ids[4] = {v.x, v.y, v.z, v.w};
            // slot NB_MAX: the vertex itself (and a triangle's unused slot 3)
            uint32_t packed = (uint32_t)NB_MAX * 0x01010101u;
            for (int q = 0; q < slots; ++q) {
                if (q == local) continue;
                int j = 0;
                while (j < n && list[j] != ids[q]) ++j;
                if (j == n) {
                    if (n == cap) {
                        fits = false;
                        break;
                    }
                    list[n++] = ids[q];
                }
                packed = (packed & ~(0xFFu << (8 * q))) | ((uint32_t)j << (8 * q));
            }
            w.ev_l[k] = packed;
        }
        w.nb_n[b] = fits ? n : -1;
    }
}

// Element vertex ids renumbered (-1 for a removed body, kept by later
// removals; such an element is removed with its body and no longer read).
__global__ void k_remap_elements(DevWorld w, BodyRemap rm) {
    GRID_STRIDE(t, w.T + w.TR) {
        int4* v = t < w.T ? &w.t_v[t] : &w.r_v[t - w.T];
        const int4 o = *v;
        auto r = [&](int x) { return remap_body(w, rm, x); };
        *v = make_int4(r(o.x), r(o.y), r(o.z), t < w.T ? r(o.w) : 0);
    }
}

// Manifolds: endpoints renumbered (the bodies keep their roles A and B);
// those that lose a body are dropped by the compaction that follows.
__global__ void k_remap_manifolds(DevWorld w, BodyRemap rm) {
    const int M = w.g[G_MANIFOLDS];
    GRID_STRIDE(m, w.Mcap) {
        if (m >= M) {
            w.alive[m] = 0;
            continue;
        }
        const int na = remap_body(w, rm, w.m.a[m]), nb = remap_body(w, rm, w.m.b[m]);
        const bool keep = na >= 0 && nb >= 0;
        w.alive[m] = keep ? 1 : 0;
        if (keep) {
            w.m.a[m] = na;
            w.m.b[m] = nb;
        }
    }
}

// ------------------------------------------------ debug / render output
// The line and point data the reference's Step collects at its warm start
// (solver.cpp:266-313): per force after Initialize, Manifold::AddLineData
// (one line per contact with positive penetration, from the contact point
// along the normal, red; manifold.cpp:259-277), Manifold::AddDebugPointData
// (both contact points, A yellow, B cyan; manifold.cpp:280-302),
// Joint::AddLineData (joint.cpp:176-190, dark red) and Spring::AddLineData
// (spring.cpp:112-125, green) between the world anchor points; per energy,
// the element's edges coloured with the mean of their bodies' colours
// (neohookeanfem.cpp:113-133: six edges, stvkfem.cpp:89-105: three).  Run
// after the step, before the bounds removal: the poses are the step-start
// ones (last_pos / last_rot), the manifolds those the warm start kept, a
// joint or spring that broke during this step still draws.  Scenes that did
// not run are skipped.  Slots are joints, springs, manifolds (Mcap), tets,
// tris; lines come out in that order (the reference draws the same set in
// its force-vector order).
AVBD_DEV v3 start_point(const DevWorld& w, int b, v3 r) {  // Transform::TransformPoint at step start
    const v3 s = f3(w.scale[b]);
    return add(qrot(fq(w.last_rot[b]), v3{r.x * s.x, r.y * s.y, r.z * s.z}), f3(w.last_pos[b]));
}

AVBD_DEV bool dbg_force_alive(const DevWorld& w, uint32_t flags, int b) {
    return !(flags & FORCE_DEAD) && scene_runs(w, w.scene[b]);
}

__global__ void k_dbg_count(DevWorld w, int* lc, int* pc) {
    const int M = w.g[G_MANIFOLDS];
    const int J = w.J, SP = w.SP, MC = w.Mcap, T = w.T, TR = w.TR;
    GRID_STRIDE(t, J + SP + MC + T + TR) {
        int n = 0;
        if (t < J) {
            n = dbg_force_alive(w, w.j_flags[t], w.j_b[t]) ? 1 : 0;
        } else if (t < J + SP) {
            const int k = t - J;
            n = dbg_force_alive(w, w.s_flags[k], w.s_b[k]) ? 1 : 0;
        } else if (t < J + SP + MC) {
            const int m = t - J - SP;
            int np = 0;
            if (m < M && scene_runs(w, w.scene[w.m.a[m]])) {
                const int nc = w.m.n[m];
                np = 2 * nc;
                for (int c = 0; c < nc; ++c)
                    if (w.m.cold[((size_t)m * 8 + c) * CONTACT_COLD_F4 + 2].w > 0.f) ++n;
            }
            pc[m] = np;
        } else if (t < J + SP + MC + T) {
            const int4 v = w.t_v[t - J - SP - MC];
            n = (v.x >= 0 && v.y >= 0 && v.z >= 0 && v.w >= 0 && scene_runs(w, w.scene[v.x])) ? 6 : 0;
        } else {
            const int4 v = w.r_v[t - J - SP - MC - T];
            n = (v.x >= 0 && v.y >= 0 && v.z >= 0 && scene_runs(w, w.scene[v.x])) ? 3 : 0;
        }
        lc[t] = n;
    }
}

AVBD_DEV void put_line(float* o, v3 a, v3 b, float r, float g, float bl, float al) {
    o[0] = a.x; o[1] = a.y; o[2] = a.z;
    o[3] = b.x; o[4] = b.y; o[5] = b.z;
    o[6] = r; o[7] = g; o[8] = bl; o[9] = al;
}

AVBD_DEV float4 body_rgb(const DebugOut& d, int b) {
    return d.rgb ? d.rgb[b] : make_float4(1.f, 1.f, 1.f, 0.f);
}

AVBD_DEV void edge_line(const DevWorld& w, const DebugOut& d, float* o, int a, int b) {
    const float4 ca = body_rgb(d, a), cb = body_rgb(d, b);
    // (color_a + color_b) * 0.5; the reference leaves alpha unset, here 1
    put_line(o, f3(w.last_pos[a]), f3(w.last_pos[b]), (ca.x + cb.x) * 0.5f, (ca.y + cb.y) * 0.5f,
             (ca.z + cb.z) * 0.5f, 1.f);
}

// One workgroup per DBG_SLOTS consecutive slots: their lines are contiguous
// in the output (the offsets are a scan over the slots), so they are built
// in LDS and then stored with consecutive 8-byte stores (a line is 40 bytes;
// one thread writing its slot's lines directly would store 40-240 bytes per
// lane, scattered across a wave).
#define DBG_SLOTS 128
#define DBG_MAX_LINES (8 * DBG_SLOTS)  // a manifold slot has at most 8 lines
__global__ __launch_bounds__(DBG_SLOTS) void k_dbg_write(DevWorld w, DebugOut d, const int* lc, const int* pc) {
    __shared__ float sl[DBG_MAX_LINES * 10];
    const int J = w.J, SP = w.SP, MC = w.Mcap, T = w.T, TR = w.TR;
    const int slots = J + SP + MC + T + TR;
    for (int b0 = blockIdx.x * DBG_SLOTS; b0 < slots; b0 += gridDim.x * DBG_SLOTS) {
        const int b1 = min(b0 + DBG_SLOTS, slots);
        const int base = d.loff[b0];
        const int total = d.loff[b1 - 1] + lc[b1 - 1] - base;
        const int t = b0 + threadIdx.x;
        if (t < b1 && (lc[t] > 0 || (t >= J + SP && t < J + SP + MC && pc[t - J - SP] > 0))) {
            float* o = sl + (size_t)(d.loff[t] - base) * 10;
            if (t < J + SP) {
                const bool jt = t < J;
                const int k = jt ? t : t - J;
                const int a = jt ? w.j_a[k] : w.s_a[k], b = jt ? w.j_b[k] : w.s_b[k];
                const v3 ra = f3(jt ? w.j_ra[k] : w.s_ra[k]), rb = f3(jt ? w.j_rb[k] : w.s_rb[k]);
                const v3 pa = a >= 0 ? start_point(w, a, ra) : ra;
                const v3 pb = start_point(w, b, rb);
                if (jt) put_line(o, pa, pb, 0.75f, 0.f, 0.f, 1.f);
                else put_line(o, pa, pb, 0.f, 1.f, 0.f, 1.f);
            } else if (t < J + SP + MC) {
                const int m = t - J - SP;
                const int a = w.m.a[m], b = w.m.b[m];
                const int nc = w.m.n[m];
                const qt qa = fq(w.last_rot[a]), qb = fq(w.last_rot[b]);
                const v3 xa = f3(w.last_pos[a]), xb = f3(w.last_pos[b]);
                float4* po = reinterpret_cast<float4*>(d.points) + (size_t)d.poff[m] * 2;
                for (int c = 0; c < nc; ++c) {
                    const float4* cold = w.m.cold + ((size_t)m * 8 + c) * CONTACT_COLD_F4;
                    const float4 c0 = cold[0], c1 = cold[1], c2 = cold[2];
                    if (c2.w > 0.f) {
                        const v3 s = w.m.cpos ? f3(w.m.cpos[(size_t)m * 8 + c]) : v3{0.f, 0.f, 0.f};
                        put_line(o, s, add(s, mul(f3(c2), c2.w)), 1.f, 0.f, 0.f, 1.f);
                        o += 10;
                    }
                    // points: posA + R_A rA (yellow), posB + R_B rB (cyan); 32-byte records
                    const v3 wa = add(xa, mv(rotmat(qa), v3{c0.z, c0.w, c1.x}));
                    const v3 wb = add(xb, mv(rotmat(qb), v3{c1.y, c1.z, c1.w}));
                    po[0] = make_float4(wa.x, wa.y, wa.z, 0.f);
                    po[1] = make_float4(1.f, 1.f, 0.f, 1.f);
                    po[2] = make_float4(wb.x, wb.y, wb.z, 0.f);
                    po[3] = make_float4(0.f, 1.f, 1.f, 1.f);
                    po += 4;
                }
            } else if (t < J + SP + MC + T) {
                const int4 v = w.t_v[t - J - SP - MC];
                const int e[6][2] = {{v.x, v.y}, {v.x, v.z}, {v.x, v.w}, {v.y, v.z}, {v.y, v.w}, {v.z, v.w}};
                for (int k = 0; k < 6; ++k) edge_line(w, d, o + 10 * k, e[k][0], e[k][1]);
            } else {
                const int4 v = w.r_v[t - J - SP - MC - T];
                const int e[3][2] = {{v.x, v.y}, {v.y, v.z}, {v.z, v.x}};
                for (int k = 0; k < 3; ++k) edge_line(w, d, o + 10 * k, e[k][0], e[k][1]);
            }
        }
        __syncthreads();
        // a line is 40 bytes, so every line (and the block's first) starts
        // 8-byte aligned
        const float2* src = reinterpret_cast<const float2*>(sl);
        float2* dst = reinterpret_cast<float2*>(d.lines + (size_t)base * 10);
        for (int i = threadIdx.x; i < total * 5; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
}

// Instance data of the renderer (RenderEngine::UpdateInstanceBuffer,
// RenderEngine.cpp:951-1029; GPUInstanceData, RenderEngine.hpp:43-49): per
// live body the column-major model matrix [R diag(scale) | position]
// (Transform::GetModelMatrix, geometry.hpp:112-122), the normal matrix
// (inverse transpose of its 3x3 block = R diag(1 / scale) for a rotation R;
// GetNormalMatrix, geometry.hpp:125-139, columns padded to 4) and the colour
// with alpha 1; 32 floats.  live_off: index among the live bodies (dead
// slots: scale.w = -1e30, marked -1 by the caller's scan of k_live_flag).
__global__ void k_live_flag(DevWorld w, int* flag) {
    GRID_STRIDE(b, w.N) flag[b] = w.scale[b].w > -1e29f ? 1 : 0;
}

// Exclusive scan of n flags by one workgroup of 1024 threads (no temporary
// storage beyond LDS): 1024 flags per pass, a wave-level scan with
// __shfl_up, the 16 wave totals scanned in LDS, a running carry between
// passes.  *total = the sum.  Used where no device allocation may happen.
__global__ __launch_bounds__(1024) void k_scan_flags_1wg(const int* flag, int* off, int n, int* total) {
    __shared__ int wave_sum[16];
    __shared__ int carry_s;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + (int)threadIdx.x;
        const int v = i < n ? flag[i] : 0;
        int x = v;
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        if (lane == 63) wave_sum[wave] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int k = 0; k < 16; ++k) {
                const int t = wave_sum[k];
                wave_sum[k] = acc;
                acc += t;
            }
        }
        __syncthreads();
        const int carry = carry_s;
        if (i < n) off[i] = carry + wave_sum[wave] + x - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry_s = carry + wave_sum[15] + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry_s;
}

// Exclusive scan of n ints in tiles of SCAN_TILE (reduce, then scan):
// k_scan_reduce writes each tile's sum, k_scan_flags_1wg scans the sums in
// place, k_scan_tiles scans each tile from its offset.  Integer sums, so the
// result does not depend on the tiling.
AVBD_DEV int block_sum256(int v) {
    __shared__ int ws[4];
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    return ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ __launch_bounds__(256) void k_scan_reduce(const int* in, int n, int* sums) {
    const int base = blockIdx.x * SCAN_TILE;
    int v = 0;
#pragma unroll
    for (int k = 0; k < SCAN_TILE / 256; ++k) {
        const int i = base + k * 256 + (int)threadIdx.x;
        if (i < n) v += in[i];
    }
    v = block_sum256(v);
    if (threadIdx.x == 0) sums[blockIdx.x] = v;
}

// Thread t scans the 16 consecutive entries base + 16 t ...; the thread
// totals are scanned across the block (waves with __shfl_up, wave totals in
// LDS).  sums == nullptr: a single tile starting at 0.
__global__ __launch_bounds__(256) void k_scan_tiles(const int* in, int* out, int n, const int* sums) {
    static_assert(SCAN_TILE == 256 * 16, "16 entries per thread");
    const int base = blockIdx.x * SCAN_TILE + 16 * (int)threadIdx.x;
    int v[16];
    if (base + 16 <= n) {
        const int4* q = reinterpret_cast<const int4*>(in + base);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int4 x = q[k];
            v[4 * k] = x.x; v[4 * k + 1] = x.y; v[4 * k + 2] = x.z; v[4 * k + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = base + k < n ? in[base + k] : 0;
    }
    int t = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) t += v[k];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = t;
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    __shared__ int ws[4];
    if (lane == 63) ws[wave] = x;
    __syncthreads();
    int acc = (sums ? sums[blockIdx.x] : 0) + x - t;
    for (int k = 0; k < wave; ++k) acc += ws[k];
    int o[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        o[k] = acc;
        acc += v[k];
    }
    if (base + 16 <= n) {
        int4* q = reinterpret_cast<int4*>(out + base);
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = make_int4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (base + k < n) out[base + k] = o[k];
    }
}

__global__ void k_instances(DevWorld w, const int* flag, const int* live_off, const float4* rgb, float4* out) {
    GRID_STRIDE(b, w.N) {
        if (!flag[b]) continue;
        const m3 R = rotmat(fq(w.rot[b]));
        const float4 s = w.scale[b];
        const float4 p = w.pos[b];
        const float4 c = rgb ? rgb[b] : make_float4(1.f, 1.f, 1.f, 0.f);
        const float sc[3] = {s.x, s.y, s.z};
        float4* o = out + (size_t)live_off[b] * 8;
        for (int j = 0; j < 3; ++j)  // model matrix columns 0-2: R e_j s_j
            o[j] = make_float4(R.m[0][j] * sc[j], R.m[1][j] * sc[j], R.m[2][j] * sc[j], 0.f);
        o[3] = make_float4(p.x, p.y, p.z, 1.f);
        for (int j = 0; j < 3; ++j)  // normal matrix columns: R e_j / s_j
            o[4 + j] = make_float4(R.m[0][j] / sc[j], R.m[1][j] / sc[j], R.m[2][j] / sc[j], 0.f);
        o[7] = make_float4(c.x, c.y, c.z, 1.f);
    }
}

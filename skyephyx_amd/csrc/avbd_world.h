// avbd_world.h -- device-resident world layout shared by the kernels and the
// host orchestration.  Everything is structure-of-arrays in HBM; per-body
// records are float4-aligned so one body field is one 16-byte load.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/avbd_capi.h"

// incidence entry encoding (force lists): [31:30] kind, [29] side, [28:0] index
#define ENT_CONTACT 0u
#define ENT_JOINT 1u
#define ENT_SPRING 2u
#define ENT_KIND(e) ((e) >> 30)
#define ENT_SIDE_B(e) (((e) >> 29) & 1u)
#define ENT_INDEX(e) ((e) & 0x1FFFFFFFu)
#define MAKE_ENT(kind, sideB, idx) (((uint32_t)(kind) << 30) | ((uint32_t)(sideB) << 29) | (uint32_t)(idx))
// energy entries: [31] 0 tet / 1 tri, [30:29] local vertex, [28:0] index
#define EENT_TRI(e) ((e) >> 31)
#define EENT_LOCAL(e) (((e) >> 29) & 3u)
#define MAKE_EENT(tri, local, idx) (((uint32_t)(tri) << 31) | ((uint32_t)(local) << 29) | (uint32_t)(idx))
// most neighbours a FEM vertex keeps in its list (and in LDS during a sweep)
#define NB_MAX 32
// entries per workgroup of the tiled exclusive scan (256 threads x 16)
#define SCAN_TILE 4096

// force state flags (device); bit 0 = AVBD_FORCE_DISABLED
#define FORCE_DEAD 2u

// per-scene status words
#define ST_EMERGENCY 0
#define ST_REMOVED 1
#define ST_OOB 2
#define ST_ACTIVE 3
#define ST_DIRTY 4   // constraint graph changed -> recolour (DIRTY_* bits)
#define ST_NCOL 5    // colours of the scene (last colouring)
#define ST_FUSED 6   // phases 4-7 run in the per-scene kernel (k_scene_step)
#define ST_WORDS 8

// ST_DIRTY bits: colour the scene from scratch / update the last colouring
// from the bodies marked in cseed
#define DIRTY_FULL 1
#define DIRTY_SEEDS 2

// scenes of up to this many bodies are coloured with per-body words in LDS
// (4 bytes each); larger ones in global memory
#define COLOUR_LDS_MAX 32768

// worlds with at most this many possible candidate pairs sort their new pairs
// in one workgroup (k_pair_sort_small)
#define SMALL_PAIR_SORT 2048

// global error/status words
#define G_ERR_MANIFOLD_CAP 0
#define G_ERR_PAIR_CAP 1
#define G_ERR_COLOURS 2
#define G_NEW_PAIRS 3
#define G_MANIFOLDS 4     // live manifold count (current buffer)
#define G_MANIFOLDS_OLD 5 // count before this step's appends
#define G_FRACTURES 6
#define G_NCOLOURS 7
#define G_ANY_OOB 8
#define G_ANY_EMERGENCY 9
#define G_CANDIDATES 10   // overlapping pairs found by the broad phase
#define G_WORDS 16

#define AVBD_MAX_COLOURS 256
// sweep lists: bucket 2c holds the colour-c bodies with forces (rigid bodies,
// particles in contact), bucket 2c+1 those with only FEM energies
#define AVBD_BUCKETS (2 * AVBD_MAX_COLOURS)
#define CONTACT_HOT_F4 6   // float4 per contact (hot)
#define CONTACT_COLD_F4 3  // float4 per contact (cold)

// Hot contact record (6 x float4 = 96 B), read by the primal and dual sweeps:
//   h0 = (n, C0_n)  h1 = (t1, C0_t1)  h2 = (t2, C0_t2)
//   h3 = (rA_world, lambda_n)  h4 = (rB_world, lambda_t1)
//   h5 = (lambda_t2, k_n, k_t1, k_t2)
// n is the row direction -(contact normal) (manifold.cpp:113).
// Cold contact record (3 x float4), read at warm start and download:
//   c0 = (id bits, stick, rA.x, rA.y)  c1 = (rA.z, rB.x, rB.y, rB.z)
//   c2 = (normal A->B, penetration)

struct ManifoldBuf {
    int* a;
    int* b;
    long long* serial;
    int* n;
    float* fric;
    float4* hot;   // [cap * 8 * CONTACT_HOT_F4]
    float4* cold;  // [cap * 8 * CONTACT_COLD_F4]
    float4* cpos;  // [cap * 8] world contact points, only with debug output on (else null)
};

struct DevWorld {
    // scenes
    int S;
    const avbd_scene* sc;  // params (device copy)
    int* st;               // [S * ST_WORDS]
    float* rho;            // [S] trustRegionRho (solver.hpp:80)
    double* rho_acc;       // [3S] the current sweep's actual and predicted decrease and their terms' magnitude (AVBD_SCENE_TRUST_REGION)
    const int* sc_off;     // [S+1] body offsets
    // bodies
    int N;
    float4* pos;        // xyz
    float4* rot;        // (w,x,y,z)
    float4* scale;      // xyz, w = detection radius
    float4* vel;        // xyz, w = mass
    float4* angvel;     // xyz, w = friction
    float4* pvel;       // prev velocity
    float4* pangvel;    // prev angular velocity
    float4* inertia;    // body-frame diagonal
    float4* drag;       // xyz: velocity added by the next velocity update, w: dragged (Mesh::isDragged)
    uint32_t* flags;
    int* scene;
    float4* last_pos;
    float4* last_rot;
    float4* inert_pos;
    float4* inert_rot;
    float4* diff0;      // (x - x_last, rotdiff.x)
    float4* diff1;      // (rotdiff.y, rotdiff.z, 0, 0)
    // manifolds
    int Mcap;
    ManifoldBuf m;      // current buffer
    // joints
    int J;
    int* j_a;
    int* j_b;
    long long* j_serial;
    uint32_t* j_flags;
    float4* j_ra;       // xyz, w = torque arm
    float4* j_rb;
    float4* j_rest;     // rest rotation (w,x,y,z)
    float* j_c0;        // [J*6]
    float* j_row;       // [J*6*6]: per row {stiffness, fracture, fmin, fmax, lambda, penalty}
    // springs
    int SP;
    int* s_a;
    int* s_b;
    long long* s_serial;
    uint32_t* s_flags;
    float4* s_ra;       // xyz, w = rest length
    float4* s_rb;
    float* s_row;       // [SP*6]
    // tets
    int T;
    int4* t_v;
    float4* t_d;        // [T*3]: dm_inv row-major (9), V0, mu, lambda
    // tris
    int TR;
    int4* r_v;
    float4* r_d;        // [TR*2]: dm_inv (4), area, mu, lambda, 0
    // force incidence (rebuilt every step)
    int* f_off;         // [N+1]
    uint32_t* f_ent;
    int f_cap;
    // energy incidence (static)
    const int* e_off;   // [N+1]
    const uint32_t* e_ent;
    // element data in incidence order (static): entry k of e_ent finds its
    // element's rest-shape data and vertex ids at index k here, so the
    // elements of one vertex are one contiguous stream.  Tets: a, b, c = the
    // three t_d records, v = t_v; triangles: a, b = the two r_d records,
    // c unused, v = r_v.
    float4* ev_a;
    float4* ev_b;
    float4* ev_c;
    int4* ev_v;
    // Neighbour lists of the FEM vertices: vertex b's distinct element
    // neighbours (first appearance in its element list) at
    // nb_id[2 e_off[b] ...], nb_n[b] of them (-1: more than fit, the
    // vertex gathers by element ids instead); ev_l[k] packs, per slot of
    // entry k's element, the slot vertex's index in that list (byte s;
    // NB_MAX for the vertex itself and a triangle's unused slot 3).
    int* nb_id;         // [2 (4T + 3TR) + 1]
    int* nb_n;          // [N]
    uint32_t* ev_l;     // [4T + 3TR + 1]
    int use_nb;         // FEM-only sweeps read neighbour positions from LDS (option fem_neighbour_lists)
    // colouring
    int* colour;        // [N]
    int* col_cnt;       // [AVBD_BUCKETS + 1]
    int* col_off;       // [AVBD_BUCKETS + 1]
    int* col_list;      // [N]
    // per entry of a FEM-only bucket (at its col_list index): the particle,
    // its first energy entry, its scene and (its element count << 8) |
    // (nb_n + 1), so the sweep starts from one load instead of two levels
    int4* col_fem;      // [N]
    // scratch for the colouring and the broad phase
    int* ccount;        // [N]
    int* cseed;         // [N] bodies whose set of lower-index neighbours changed
    int* queue;         // [2N]
    int* cnt_i;         // [N]
    int* cur_i;         // [N+1]
    int2* pairs;        // [pair_cap]
    int2* pairs_tmp;    // [pair_cap]
    // broad-phase tiles (256 bodies of a scene in index order): first tile
    // of each scene, scene of each tile, widened box of each tile's spheres
    const int* tile_off;    // [S+1]
    const int* tile_scene;  // [tiles]
    float4* tile_lo;        // [tiles]
    float4* tile_hi;        // [tiles]
    float4* chunk_lo;       // [tiles * 8]: the same for each chunk of 32
    float4* chunk_hi;       // [tiles * 8]
    int pair_cap;
    int* alive;         // [Mcap]
    int* alive_off;     // [Mcap+1]
    int* fracture_bodies;  // [2 * (J + SP) + 2]
    int* g;             // global status words [G_WORDS]
    long long* next_serial;
};

// Debug/render output of a step (option debug_output): per line 10 floats
// (start xyz, end xyz, colour rgba), per point 8 (position xyz, 0, colour
// rgba), the layouts of the reference's GPULineData / GPUDebugPointData.
struct DebugOut {
    float* lines;
    float* points;
    const int* loff;    // [J + SP + Mcap + T + TR]: first line of each force / element slot
    const int* poff;    // [Mcap]: first point of each manifold slot
    const float4* rgb;  // per body colour (Mesh::color), or null for white
};

// Body renumbering of a bounds removal, which permutes bodies only inside the
// affected scenes: aoff[s] = start of scene s's part of lmap (-1: scene not
// affected, its indices stay), lmap[aoff[s] + i - sc_off[s]] = new index of
// the scene's old body i (-1: removed).
struct BodyRemap {
    const int* aoff;  // [S]
    const int* lmap;  // [bodies of the affected scenes]
};

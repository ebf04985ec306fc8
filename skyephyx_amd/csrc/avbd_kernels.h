// avbd_kernels.h -- kernel declarations (definitions in avbd_kernels.hip).
#pragma once

#include "avbd_world.h"

__global__ void k_fill_int(int* p, int n, int v);
__global__ void k_set_word(int* p, int v);
__global__ void k_set_active(DevWorld w, const int* ids, int n, int all);
__global__ void k_tile_bounds(DevWorld w, const int* tiles);
__global__ void k_broad(DevWorld w, const int4* items);
__global__ void k_pair_filter(DevWorld w);
__global__ void k_pair_sort_small(DevWorld w);
__global__ void k_pair_append(DevWorld w, const unsigned long long* keys);
__global__ void k_pair_finish(DevWorld w);
__global__ void k_manifold_init(DevWorld w);
__global__ void k_force_init(DevWorld w);
__global__ void k_compact(DevWorld w, ManifoldBuf dst);
__global__ void k_compact_finish(DevWorld w);
__global__ void k_inc_count(DevWorld w);
__global__ void k_inc_fill(DevWorld w);
__global__ void k_inc_sort(DevWorld w);
template <bool LDS>
__global__ void k_colour(DevWorld w, int full);
__global__ void k_colour_hist(DevWorld w);
__global__ void k_colour_scan(DevWorld w);
__global__ void k_colour_fill(DevWorld w);
__global__ void k_predict(DevWorld w);
template <int G, bool FEM>
__global__ void k_primal(DevWorld w, int bucket, int iter, int poststab);
template <int GR, int GF, bool CANON>
__global__ void k_primal_pair(DevWorld w, int colour, int iter, int poststab, int rigid_blocks);
template <int G>
__global__ void k_primal_canon(DevWorld w, int bucket, int iter, int poststab);
template <int G>
__global__ void k_primal_fem(DevWorld w, int bucket, int iter, int poststab);
template <int G>
__global__ void k_scene_step(DevWorld w, const int* scenes, int mode, int lds_bytes);
__global__ void k_dual(DevWorld w, int iter);
__global__ void k_rho_update(DevWorld w);
__global__ void k_fracture_reset(DevWorld w);
__global__ void k_velocity(DevWorld w);
__global__ void k_stick(DevWorld w);
__global__ void k_gather_f4(float4* dst, const float4* src, const int* perm, int n);
__global__ void k_gather_u32(uint32_t* dst, const uint32_t* src, const int* perm, int n);
__global__ void k_scatter_f4(float4* dst, const float4* src, const int* to, int n);
__global__ void k_scatter_u32(uint32_t* dst, const uint32_t* src, const int* to, int n);
__global__ void k_remap_forces(DevWorld w, BodyRemap rm);
__global__ void k_remap_manifolds(DevWorld w, BodyRemap rm);
__global__ void k_kill_bodies(DevWorld w, const int* list, int n);
__global__ void k_element_stream(DevWorld w, int n);
__global__ void k_nb_build(DevWorld w, int b0, int b1);
__global__ void k_remap_elements(DevWorld w, BodyRemap rm);
__global__ void k_dbg_count(DevWorld w, int* lc, int* pc);
__global__ void k_live_flag(DevWorld w, int* flag);
__global__ void k_scan_reduce(const int* in, int n, int* sums);
__global__ void k_scan_tiles(const int* in, int* out, int n, const int* sums);
__global__ void k_scan_flags_1wg(const int* flag, int* off, int n, int* total);
__global__ void k_instances(DevWorld w, const int* flag, const int* live_off, const float4* rgb, float4* out);
__global__ void k_dbg_write(DevWorld w, DebugOut d, const int* lc, const int* pc);

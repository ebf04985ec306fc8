// avbd_internal.h -- library-internal helpers shared by the solver context
// (avbd_host.hip) and the request server (avbd_server.cpp).  Not part of the
// C ABI.
#pragma once

#include <string>

#include "../../include/avbd_capi.h"

// The element counts a device world's allocations depend on.
struct WorldCounts {
    long long S = 0, N = 0, nonpart = 0, J = 0, SP = 0, M = 0, T = 0, TR = 0;
    long long tiles = 0;  // broad-phase tiles (256 bodies), summed over scenes
    long long items = 0;  // broad-phase tile pairs, summed over scenes
};

// Broad-phase tiles and tile pairs (i-tile, j-tile >= i-tile) of one scene
// of n bodies.
inline long long broad_tiles(long long n) { return (n + 255) / 256; }
inline long long broad_items(long long n) {
    const long long t = broad_tiles(n);
    return t * (t + 1) / 2;
}

WorldCounts world_counts(const avbd_world* w);

// Bytes of device memory an upload of a world with these counts allocates
// (manifold_capacity as for avbd_upload; 0 = default).  Negative AVBD_E_*
// code when the counts exceed a device capacity.
long long world_device_bytes(const WorldCounts& k, long long manifold_capacity);

// Record msg as the thread's last error and return code.
int avbd_set_error(int code, const std::string& msg);

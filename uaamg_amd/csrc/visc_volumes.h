// The viscosity level-0 liquid volumes on the device (L_with_level::
// calculate_subcell_sdf_volume, simd_viscosity3d.cpp:258-498): the liquid SDF
// sampled at the half-voxel subcell lattice (BoxSampler, clamped to 1.5 dx),
// each subcell's liquid fraction from its eight corner samples (volume_fraction,
// volume_fractions.cpp:82-126), and the voxel-, face- and edge-centred volumes as
// 1/8 of the eight subcells around each centre. Bit-identical to the host stage.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "leaf_topology.h"

namespace uaamg {

// liquid: the liquid SDF (tree order); sub_origin: the leaves whose 16^3 subcells are
// sampled (the liquid leaves dilated once over 26 neighbours), vol_origin: the
// leaves receiving volumes (dilated twice); both tree ordered. Outputs [nvol][512]
// each, ordered like vol_origin. Returns 0, or a negative uaamg status with err set.
int visc_volumes_device(int device, const HostGrid& liquid, const std::vector<Coord>& sub_origin,
                        const std::vector<Coord>& vol_origin, float dx, std::vector<float>& voxel_vol,
                        std::vector<float> face_vol[3], std::vector<float> edge_vol[3], std::string& err);

}  // namespace uaamg

// RenderSceneIR + RenderAcceleration -> C-ABI descriptors (include/yrh.h).
// The descriptors point into the IR (the reference's own arrays, byte
// compatible) plus a few converted tables owned by Marshal.
#pragma once

#include <vector>

#include "../../include/yrh.h"
#if defined(YAORAY_HIP_DROPIN)
#include <yaoray/accel/acceleration.hpp>
#include <yaoray/scene/render_scene.hpp>
#else
#include "yr_ir.hpp"
#endif

namespace yrhip {

struct Marshal {
    std::vector<yrh_bvh_desc> blas;
    std::vector<yrh_material> materials;
    std::vector<yrh_emissive> emissives;
    std::vector<yrh_texture> textures;
};

yrh_bvh_desc DescribeBvh(const yr::RenderBvh& b);

// Fills `g`; returns false if the acceleration structure is missing.
void DescribeGeometry(const yr::RenderSceneIR& ir, const yr::RenderAcceleration& accel, Marshal& m,
                      yrh_geometry_desc& g);

// Fills `d`; returns an error message (empty on success).
std::string DescribeShading(const yr::RenderSceneIR& ir, Marshal& m, yrh_shading_desc& d);

}  // namespace yrhip

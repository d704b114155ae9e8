// scene_build.h — host-side flattening of a tpt_scene_desc into the device
// arena sections of dev_types.h (cuda::Scene(const cpu::Scene&) replacement,
// li-plus/tinypt src/tinypt/cuda/scene.h:95-174).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tinypt_hip.h"
#include "dev_types.h"

namespace tpt {

struct tpt_error : std::runtime_error {
    int code;
    tpt_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

struct HostScene {
    std::vector<float4> nodes;
    int32_t root = -1;
    std::vector<float4> tris_bvh, tris;
    std::vector<int4> tri_info;
    std::vector<float4> shade_rec; // kShadeRec float4 per triangle (dev_types.h)
    std::vector<float> normals, texcoords;
    std::vector<int32_t> gid_local;
    std::vector<DAnalytic> analytic;
    std::vector<DMaterial> materials;
    std::vector<DImage> images;
    std::vector<uint8_t> texels;
    std::vector<DDelta> delta;
    std::vector<DLight> lights;
    std::vector<float> env_marg, env_cond;
    std::vector<int2> light_shapes;                 // shapes of mesh ObjectLights (dev_types.h)
    std::vector<int32_t> env_gmarg, env_gcond; // guide tables of the CDFs (dev_types.h)
    int32_t env_w = 1, env_h = 1, env_gw = 1, env_gh = 1;
    int32_t bg_map = 0;
    DTex bg{};
    float bg_rot[9];
    DCamera cam{};
    int max_depth_seen = 0; // binary BVH depth
    int max_stack = 0;      // worst-case traversal stack entries of the flattened tree
    int layout = kNodeBvh4Q; // node layout (dev_types.h)
    int bfs_prefix = 0;      // 4-wide layouts: nodes [0, bfs_prefix) are the top of the tree in BFS order
    int loc_bits = 0;        // leaf records carry tri_id | material << loc_bits (DScene::loc_bits)
};

// Number of 4-wide nodes laid out breadth-first at the start of the node array.
static constexpr int kBfsPrefix = 4096;

// Throws tpt_error on invalid input.
void build_host_scene(const tpt_scene_desc &desc, HostScene &out);

} // namespace tpt

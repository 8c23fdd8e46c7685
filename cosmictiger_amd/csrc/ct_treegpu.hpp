// Device tree build (tree_create on the GPU); see ct_treegpu.hip.
#pragma once
#include <cstdint>
#include <vector>

struct ct_box;

namespace ct {

// What the host keeps of a tree built on the device.
struct GpuTreeResult {
	int nnodes = 0;
	int max_depth = 0;
	std::vector<int32_t> level_off;  // breadth-first: level d is [level_off[d], level_off[d+1])
	std::vector<int32_t> leaf_id, leaf_pbegin, leaf_pend;
};

// Builds the tree of the box's device particles (permuting them in Morton
// order) into the box's node, multipole (float and double) and expansion
// buffers.  Returns CT_E_TREE when a node deeper than 63 levels would have to
// split (the host build handles those).
int gpu_tree_build(ct_box* b, double theta, int bucket, double hsoft, GpuTreeResult* res);

// Frees the tree-build scratch a box keeps between builds.
void tree_scratch_free(ct_box* b);

}  // namespace ct

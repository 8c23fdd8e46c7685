// Host-side tree build for the MI355X kick (drop-in for tree_create,
// cosmictiger/tree.hpp:207 and src/tree.cpp:273-601).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace ct {

// One node as built on the host, in breadth-first order after build.
struct HostNode {
	int32_t box_begin[3];  // fixed<int,29> raw (cosmictiger/range.hpp:358-360)
	int32_t box_end[3];
	int32_t child[2];      // -1 for leaves
	int32_t parent;        // -1 for the root
	int32_t pbegin, pend;  // particle range [pbegin, pend)
	uint32_t pos[3];       // fixed32 expansion centre
	float radius;
	int32_t leaf;
	int32_t depth;
	float multi[37];       // trless order-6 multipole in the scaled frame
	double multid[37];     // the same in double (FP64 kicks)
};

struct HostParticles {
	int64_t n = 0;
	std::vector<uint32_t> x, y, z;
	std::vector<float> vx, vy, vz;
	std::vector<int8_t> rung;
	std::vector<int32_t> id;  // original index of each particle
};

struct TreeBuildParams {
	double theta = 0.7;
	int bucket = 128;
	double hsoft = 0.0;
	int nthreads = 0;  // 0: hardware concurrency
	int64_t first = 0; // the tree covers particles [first, n) (the active range)
};

struct HostTree {
	std::vector<HostNode> nodes;     // breadth-first: nodes of depth d are contiguous
	std::vector<int32_t> level_off;  // level_off[d] .. level_off[d+1]
	int max_depth = 0;
	int64_t nleaves = 0;
};

// Builds the tree over particles [prm.first, n) (particles_current_range,
// src/particles.cpp:1020-1025), permuting them in place (the partition of
// src/particles.cpp:979-1018).  Returns 0 or a negative error code.
int build_tree(HostParticles& p, const TreeBuildParams& prm, HostTree& out, std::string* err);

}  // namespace ct

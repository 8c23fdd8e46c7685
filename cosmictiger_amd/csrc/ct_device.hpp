// Device-side data layout of a resident box (HBM), shared by the kernels and
// the host launcher.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ct {

// 64-byte node record read by the walk and the interaction kernels.
// Mirrors the fields of tree_node used on the device (cosmictiger/tree.hpp:98-169).
struct __align__(16) DNode {
	uint32_t pos[3];  // fixed32 expansion centre
	float radius;
	int32_t child[2];
	int32_t pbegin, pend;
	int32_t bb[3];    // box begin, fixed<int,29> raw
	int32_t be[3];    // box end
	int32_t flags;    // bit 0: leaf; bits 1..: depth
	int32_t parent;
};
static_assert(sizeof(DNode) == 64, "DNode is 64 bytes");

// Multipole record: 37 floats of multipole<float> then the fixed32 position
// (the reference's multi_pos, cosmictiger/tree.hpp:86-94), 160 bytes.
constexpr int kMultiStride = 40;
// Local expansion accumulator stride (50 floats padded to 52).
constexpr int kLStride = 52;

// Walk work entry: one sink node with its direct and Ewald check lists.
struct __align__(16) Act {
	int32_t self;
	int32_t db, dn;  // direct checks in dp[db .. db+dn)
	int32_t eb, en;  // Ewald checks in ep[eb .. eb+en)
	int32_t pad[3];
};

// Output list segment: entries out[begin .. begin+len), next segment of the
// same (kind, sink) or -1.
struct __align__(16) Seg {
	int32_t begin, len, next, pad;
};

enum ListKind { kEwald = 0, kCC = 1, kCP = 2, kPC = 3, kPP = 4, kNumLists = 5 };

constexpr int kMaxIter = 300;

// Counter block in device memory (int32 unless noted).
struct Counters {
	int32_t act[kMaxIter + 2];
	int32_t dp[kMaxIter + 2];
	int32_t ep[kMaxIter + 2];
	int32_t out[kNumLists];
	int32_t seg;
	int32_t overflow;   // bit k set: buffer k overflowed
	int32_t walk_error;
	unsigned long long stat[8];  // event counters of the kick
};

struct WalkArgs {
	const DNode* nodes;
	int32_t nnodes;
	Act* act[2];
	int32_t* dp[2];
	int32_t* ep[2];
	int64_t cap_act, cap_dp, cap_ep;
	int32_t* out[kNumLists];
	int64_t cap_out[kNumLists];
	Seg* seg;
	int64_t cap_seg;
	int32_t* head;  // [kNumLists * nnodes]
	Counters* cnt;
	float theta_inv;
	float h;
};

}  // namespace ct

// Device-side data layout of a resident box (HBM), shared by the kernels and
// the host launcher.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/cosmictiger_amd.h"

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

// Walk work entry: one sink node of box `box` with its direct and Ewald check
// lists (node indices local to the box).
struct __align__(16) Act {
	int32_t self;
	int32_t db, dn;  // direct checks in dp[db .. db+dn)
	int32_t eb, en;  // Ewald checks in ep[eb .. eb+en)
	int32_t box;     // index into the batch's BoxDesc table
	int32_t pad[2];
};

// One box of a batched kick: its resident arrays in HBM and its kick
// parameters.  Kernels of a batch index boxes through a table of these; node
// and particle indices stay local to the box.
struct __align__(16) BoxDesc {
	const DNode* nodes;
	const void* multi;  // float (FP32 modes) or double (FP64) records of kMultiStride
	void* L;            // float or double, kLStride per node
	uint32_t* x;
	uint32_t* y;
	uint32_t* z;
	float* vx;
	float* vy;
	float* vz;
	int8_t* rung;
	float* gx;
	float* gy;
	float* gz;
	float* phi;
	int32_t* head;  // [kNumLists * nnodes] list heads in the batch's arena
	void* pp;       // P2P result per particle (float4, or double4 in FP64)
	int32_t nnodes;
	int32_t nparts;
	float theta_inv;
	float h;
	ct_kick_params prm;
	double drift_c1;  // dt / a of the batch drift
	double drag_c;    // 1 / (1 + drag dt)
	int32_t drift_rung;
	int32_t do_drag;
	int32_t drift_first;  // start of the box's active range
};

// Output list segment: entries out[begin .. begin+len), next segment of the
// same (kind, sink) or -1.
struct __align__(16) Seg {
	int32_t begin, len, next, pad;
};

enum ListKind { kEwald = 0, kCC = 1, kCP = 2, kPC = 3, kPP = 4, kNumLists = 5 };

constexpr int kMaxIter = 300;

// A counter alone in its 256-byte block: reservation counters that share a
// cache line serialize each other's atomics.
struct __align__(256) HotCounter {
	int32_t v;
	int32_t pad[63];
};

// A 64-bit event counter alone in its 256-byte block.
constexpr int kStatSlots = 32;
struct __align__(256) StatSlot {
	unsigned long long v;
	unsigned long long pad[31];
};

// Counter block in device memory (int32 unless noted).
struct Counters {
	int32_t act[kMaxIter + 2];
	int32_t dp[kMaxIter + 2];
	int32_t ep[kMaxIter + 2];
	HotCounter out[kNumLists];
	HotCounter seg;
	int32_t overflow;   // bit k set: buffer k overflowed
	int32_t walk_error;
	unsigned long long stat[8];  // event counters of the kick
	// event counters added by every wave of the walk, spread over slots on
	// their own cache lines (one address for all waves serializes their
	// atomics): checks classified, and PC (sink, multipole) pairs
	StatSlot walk_checks[kStatSlots];
	StatSlot walk_pc[kStatSlots];
	// the same for the counts of the kernels with a workgroup per leaf or
	// node: P2P pairs and P2L (CP) particles
	StatSlot pp_pairs[kStatSlots];
	StatSlot cp_parts[kStatSlots];
};

struct WalkArgs {
	const BoxDesc* boxes;
	Act* act[2];
	int32_t* dp[2];
	int32_t* ep[2];
	uint8_t* dcat;  // per direct check of the current level: its class (pass 1 -> pass 2)
	uint8_t* ecat;  // per Ewald check
	int64_t cap_act, cap_dp, cap_ep;
	int32_t* out[kNumLists];
	int64_t cap_out[kNumLists];
	Seg* seg;
	int64_t cap_seg;
	Counters* cnt;
};

}  // namespace ct

// Host-side state of a resident box and the error plumbing shared by the
// HIP sources of the library.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cosmictiger_amd.h"
#include "ct_device.hpp"
#include "ct_tree.hpp"

namespace ct {

// Message of the last failed ABI call on this thread (ct_last_error).
extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
	g_err = msg;
	return code;
}

#define CT_HIP(call)                                                                                                        \
	do {                                                                                                                    \
		hipError_t e_ = (call);                                                                                             \
		if (e_ != hipSuccess) return ::ct::fail(CT_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));            \
	} while (0)

// Device buffer that only grows.
template<class T>
struct DBuf {
	T* p = nullptr;
	int64_t n = 0;
	int grow(int64_t want) {
		if (want <= n) return CT_OK;
		if (p) (void) hipFree(p);
		p = nullptr;
		n = 0;
		if (hipMalloc(&p, std::max<int64_t>(want, 1) * sizeof(T)) != hipSuccess) {
			p = nullptr;
			return fail(CT_E_OOM, "hipMalloc failed");
		}
		n = want;
		return CT_OK;
	}
	void release() {
		if (p) (void) hipFree(p);
		p = nullptr;
		n = 0;
	}
};

// ct_init has been called (device selected, Ewald tables uploaded).
int check_init();

// Source of tree versions: every tree build takes a new one, so that kick
// contexts can tell that a box's tree changed.
extern std::atomic<uint64_t> g_tree_version;

}  // namespace ct

// A resident box: host copy of the particles and tree, device state in HBM.
struct ct_box {
	ct::HostParticles hp;
	ct::HostTree tree;
	bool have_tree = false;
	bool host_stale = false;  // device particles are newer than the host copy
	double theta = 0.7, hsoft = 0.0;
	int bucket = 128;
	uint64_t tree_version = 0;
	// device particle state (tree order)
	ct::DBuf<uint32_t> x, y, z;
	ct::DBuf<float> vx, vy, vz, gx, gy, gz, phi;
	ct::DBuf<int8_t> rung;
	ct::DBuf<int32_t> id;  // index of each particle at ct_box_create
	// tree
	ct::DBuf<ct::DNode> nodes;
	ct::DBuf<float> multi, L;
	ct::DBuf<double4> pp;           // P2P result per particle (float4 or double4 records)
	ct::DBuf<double> multi64, L64;  // FP64 multipoles and expansions (FP64 kicks only)
	bool multi64_ready = false;
	int64_t nnodes = 0;
	bool host_nodes = false;              // tree.nodes holds the node table (host build)
	std::vector<int32_t> leaf_class[16];  // non-empty leaves by size class
	// kick context of single-box kicks (walk arenas, lists of the last kick)
	void* self_ctx = nullptr;
	// active range [rung_begin, n): trees, kicks and drifts cover it
	// (rung_begin of src/particles.cpp:1020-1076); the stack of
	// particles_push_rungs / particles_pop_rungs
	int64_t rung_begin = 0;
	std::vector<int64_t> rung_stack;
	// device tree-build scratch kept between builds (ct_treegpu.hip)
	void* tree_scratch = nullptr;
	// lightcone entries of the last ct_drift_lightcone (grow-only) and their
	// device counter
	ct::DBuf<ct_lc_entry> lc;
	ct::DBuf<unsigned long long> lc_cnt;
	int64_t lc_count = 0;
};

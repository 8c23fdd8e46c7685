// MI355X (gfx950) kernels of the gravity kick and the drift, and the C-ABI.
//
// Kick = five stages, all on one stream, no host synchronisation inside the
// walk:
//   1. k_walk (one launch per tree level): the dual-list walk of
//      src/kick.cu:239-484, done level-synchronously.  A wave takes a batch
//      of (sink node, check lists) entries, classifies their direct and Ewald
//      check lists (two passes: count, then write in order), reserves the
//      batch's output space with one atomic per list, and emits the
//      interaction lists plus the check lists of the next level.  Leaves that
//      still hold unresolved checks stay in the work set, which is the
//      reference's do { } while (checks && leaf) loop (src/kick.cu:344-416).
//   2. k_cells_ewald / k_cells_direct: one wave per node, a source per lane,
//      M2L over its Ewald and direct CC lists and P2L over its CP list
//      (src/gravity.cu:44-155,213-244), then a wave reduction of L.
//   3. k_l2l: per level, L(node) += L2L(L(parent)) (src/kick.cu:246-258).
//   4. k_pp (per leaf-size class) and k_kick: one wave64 per leaf: P2P with
//      sources staged in LDS; then M2P, L2P and the kick update
//      (src/gravity.cu:157-211,246-404; src/kick.cu:53-177).
//   5. k_reduce_*: deterministic FP64 reduction of the per-leaf kick_return.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <array>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cosmictiger_amd.h"
#include "ct_device.hpp"
#include "ct_ewald_tables.hpp"
#include "ct_expansions.hpp"
#include "ct_host.hpp"
#include "ct_tree.hpp"
#include "ct_treegpu.hpp"

namespace ct {

// ---------------------------------------------------------------------------
// constants
// ---------------------------------------------------------------------------

constexpr int kMinPartsPCCP = 38;  // src/kick.cu:32
constexpr float kSelfPhi = 693.0f / 256.0f;  // cosmictiger/gravity.hpp:98-102
constexpr int kNRealMax = 256;
constexpr int kNFourMax = 64;

__constant__ float c_rimg[3 * kNRealMax];
__constant__ float c_fh[3 * kNFourMax];
__constant__ float c_fc[kNFourMax * 50];
__constant__ double c_fc64[kNFourMax * 50];
__constant__ int c_nreal;
__constant__ int c_nfour;
__constant__ float c_rung_dt[16];

enum { K_NONE = 0, K_CC = 1, K_CP = 2, K_PC = 3, K_LEAF = 4, K_NEXT = 5, K_FAIL = 6 };

// ---------------------------------------------------------------------------
// fixed32 and acceptance tests (cosmictiger/fixed.hpp:303-351,
// cosmictiger/range.hpp:444-466, src/kick.cu:286-386)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float fdist(uint32_t a, uint32_t b) {
	return float(int32_t(a - b)) * 2.3283064365386963e-10f;
}

__device__ __forceinline__ float rf_distance(int32_t a, uint32_t b) {
	float f = float(double(a) * (1.0 / 536870912.0) - double(b) * (1.0 / 4294967296.0));
	f = float(double(f) - double(f > 0.5f));
	f = float(double(f) + double(f < -0.5f));
	return f;
}

__device__ __forceinline__ bool box_intersects_sphere(const int32_t* bb, const int32_t* be, const uint32_t* x, float r) {
	float d2 = 0.0f;
	for (int d = 0; d < 3; d++) {
		const float db = -rf_distance(bb[d], x[d]);
		const float de = -rf_distance(be[d], x[d]);
		const int32_t rfx = int32_t(x[d] >> 3);
		const int outside = (rfx < bb[d]) + (rfx > be[d]);
		if (outside) {
			d2 += float(outside) * fminf(db * db, de * de);
		}
	}
	return d2 < r * r;
}

__device__ __forceinline__ float mac_r2(const DNode& a, const DNode& b) {
	const float dx = fdist(a.pos[0], b.pos[0]);
	const float dy = fdist(a.pos[1], b.pos[1]);
	const float dz = fdist(a.pos[2], b.pos[2]);
	return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}

__device__ __forceinline__ int classify_ewald(const DNode& s, const DNode& o, float thinv) {
	if (o.pend - o.pbegin == 0) return K_NONE;
	float R2 = mac_r2(s, o);
	const float m = fmaxf(0.5f - (s.radius + o.radius), 0.f);
	R2 = fmaxf(R2, m * m);
	const float dcc = (s.radius + o.radius) * thinv;
	if (R2 > dcc * dcc) return K_CC;
	return (o.flags & 1) ? K_FAIL : K_NEXT;
}

__device__ __forceinline__ int classify_direct(const DNode& s, const DNode& o, float thinv, float h) {
	if (o.pend - o.pbegin == 0) return K_NONE;
	const float R2 = mac_r2(s, o);
	const float rs = s.radius, ro = o.radius;
	const float mind = (rs + ro) + h;
	const float dcc = fmaxf((rs + ro) * thinv, mind);
	if (R2 > dcc * dcc) return K_CC;
	const bool oleaf = o.flags & 1;
	if (oleaf && (s.flags & 1)) {
		const float dcp = fmaxf(__builtin_fmaf(rs, thinv, ro), mind);
		const float dpc = fmaxf(__builtin_fmaf(ro, thinv, rs), mind);
		const int sp = s.pend - s.pbegin, op = o.pend - o.pbegin;
		bool pc = (R2 > dpc * dpc || !box_intersects_sphere(s.bb, s.be, o.pos, fmaxf(thinv * ro, h))) && sp >= kMinPartsPCCP;
		bool cp = (R2 > dcp * dcp || !box_intersects_sphere(o.bb, o.be, s.pos, fmaxf(thinv * rs, h))) && op >= kMinPartsPCCP;
		if (pc && cp) {
			if (sp < op) {
				cp = false;
			} else if (sp > op) {
				pc = false;
			} else {
				pc = cp = false;
			}
		}
		if (pc) return K_PC;
		if (cp) return K_CP;
	}
	return oleaf ? K_LEAF : K_NEXT;
}

__device__ __forceinline__ int lane_id() {
	return __lane_id();
}

__device__ __forceinline__ int prefix_below(unsigned long long mask) {
	return __builtin_amdgcn_mbcnt_hi((unsigned) (mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned) mask, 0));
}

__device__ __forceinline__ int uniform(int x) {
	return __builtin_amdgcn_readfirstlane(x);
}

// A copy of *p held in scalar registers; p must be the same in every lane.
// (The descriptor tables live in writable global memory, which the compiler
// will not read with scalar loads by itself.)
template<class T>
__device__ __forceinline__ T load_uniform(const T* p) {
	static_assert(sizeof(T) % 16 == 0, "load_uniform reads 16-byte words");
	T out;
	const int4* src = reinterpret_cast<const int4*>(p);
	int* dst = reinterpret_cast<int*>(&out);
#pragma unroll
	for (int i = 0; i < int(sizeof(T) / 16); i++) {
		const int4 v = src[i];
		dst[4 * i + 0] = __builtin_amdgcn_readfirstlane(v.x);
		dst[4 * i + 1] = __builtin_amdgcn_readfirstlane(v.y);
		dst[4 * i + 2] = __builtin_amdgcn_readfirstlane(v.z);
		dst[4 * i + 3] = __builtin_amdgcn_readfirstlane(v.w);
	}
	return out;
}

// ---------------------------------------------------------------------------
// walk
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool reserve_ok(Counters* C, int bit, int64_t off, int64_t cnt, int64_t cap) {
	if (off + cnt > cap) {
		atomicOr(&C->overflow, 1 << bit);
		return false;
	}
	return true;
}

// Classification counts of one (sink, check lists) entry.
struct WalkCounts {
	int ecc, enext, efail;
	int cc, cp, pc, lf, nx;
};

// Pass 1 of an entry: classifies every check, keeps each class for pass 2
// (dcat / ecat at the check's place) and counts per list.
__device__ __forceinline__ WalkCounts walk_count(const DNode* __restrict__ nodes, float thinv, float h, const DNode& s, const int32_t* dp,
		const int32_t* ep, uint8_t* __restrict__ dcat, uint8_t* __restrict__ ecat, int db, int dn, int eb, int en, int lane) {
	WalkCounts w = {0, 0, 0, 0, 0, 0, 0, 0};
	for (int base = 0; base < en; base += 64) {
		const int i = base + lane;
		int cat = K_NONE;
		if (i < en) {
			const DNode o = nodes[ep[eb + i]];
			cat = classify_ewald(s, o, thinv);
			ecat[eb + i] = uint8_t(cat);
		}
		w.ecc += __popcll(__ballot(cat == K_CC));
		w.enext += __popcll(__ballot(cat == K_NEXT));
		w.efail += __popcll(__ballot(cat == K_FAIL));
	}
	for (int base = 0; base < dn; base += 64) {
		const int i = base + lane;
		int cat = K_NONE;
		if (i < dn) {
			const DNode o = nodes[dp[db + i]];
			cat = classify_direct(s, o, thinv, h);
			dcat[db + i] = uint8_t(cat);
		}
		w.cc += __popcll(__ballot(cat == K_CC));
		w.cp += __popcll(__ballot(cat == K_CP));
		w.pc += __popcll(__ballot(cat == K_PC));
		w.lf += __popcll(__ballot(cat == K_LEAF));
		w.nx += __popcll(__ballot(cat == K_NEXT));
	}
	return w;
}

// The fields of a box descriptor the walk reads, in scalar registers.
struct WalkBox {
	const DNode* nodes;
	int32_t* head;
	int32_t nnodes;
	float theta_inv, h;
};

__device__ __forceinline__ uint64_t uniform64(int lo, int hi) {
	return uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(lo))) | (uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(hi))) << 32);
}

__device__ __forceinline__ WalkBox load_walk_box(const BoxDesc* p) {
	static_assert(offsetof(BoxDesc, nodes) == 0 && offsetof(BoxDesc, head) % 16 == 0 && offsetof(BoxDesc, nnodes) % 16 == 0 &&
					offsetof(BoxDesc, theta_inv) == offsetof(BoxDesc, nnodes) + 8 && offsetof(BoxDesc, h) == offsetof(BoxDesc, nnodes) + 12,
			"BoxDesc layout read by load_walk_box");
	const int4* q = reinterpret_cast<const int4*>(p);
	const int4 a = q[0];
	const int4 b = q[offsetof(BoxDesc, head) / 16];
	const int4 c = q[offsetof(BoxDesc, nnodes) / 16];
	WalkBox w;
	w.nodes = reinterpret_cast<const DNode*>(uniform64(a.x, a.y));
	w.head = reinterpret_cast<int32_t*>(uniform64(b.x, b.y));
	w.nnodes = __builtin_amdgcn_readfirstlane(c.x);
	w.theta_inv = __int_as_float(__builtin_amdgcn_readfirstlane(c.z));
	w.h = __int_as_float(__builtin_amdgcn_readfirstlane(c.w));
	return w;
}

// reservation slots of one entry
enum { R_E = 0, R_CC, R_CP, R_PC, R_PP, R_ACT, R_DP, R_EP, R_SEG, R_NUM };
constexpr int kWalkBatch = 64;
// waves per SIMD the walk is compiled for (<= 96 VGPRs); the walk is bound by
// the latency of its dependent loads, so resident waves are what hides it
constexpr int kWalkWaves = 5;

// One level of the dual-list walk (src/kick.cu:239-484).  A wave takes
// kWalkBatch (sink node, check lists) entries at a time: pass 1 classifies
// every check of every entry, keeps each class and counts per list; a wave
// prefix scan turns the counts into offsets and lanes 0..8 reserve the batch's
// space, one atomic per list, in one instruction (each counter in its own
// 256-byte block: counters sharing a line serialize the atomics of every
// wave); pass 2 writes the interaction lists, the check lists of the next
// level and the entries of the next level from the kept classes.  Writes keep
// the order of the check list, so lists match a serial walk entry for entry.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWalkWaves))) void k_walk(WalkArgs A, int k) {
	Counters* C = A.cnt;
	if (__builtin_amdgcn_readfirstlane(C->overflow | C->walk_error)) return;
	const int nact = uniform(min(C->act[k], int(A.cap_act)));
	const int lane = lane_id();
	const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	const int nwaves = (gridDim.x * blockDim.x) >> 6;
	__shared__ DNode s_nodes[4][kWalkBatch];
	DNode* s_node = s_nodes[threadIdx.x >> 6];
	const Act* act = A.act[k & 1];
	const int32_t* dp = A.dp[k & 1];
	const int32_t* ep = A.ep[k & 1];
	Act* nact_out = A.act[(k + 1) & 1];
	int32_t* ndp = A.dp[(k + 1) & 1];
	int32_t* nep = A.ep[(k + 1) & 1];
	unsigned long long nchecks = 0, npcs = 0;
	// entries per batch: enough to give every wave work at this level, at most
	// kWalkBatch (one reservation per list per batch)
	const int bsz = max(1, min(kWalkBatch, (nact + nwaves - 1) / nwaves));
	for (int b0 = wave * bsz; b0 < nact; b0 += nwaves * bsz) {
		const int nb = min(bsz, nact - b0);
		// the batch's entries, entry e in lane e, and the descriptor of their
		// box when they share one (the usual case): the loop over entries then
		// waits on no entry or descriptor load
		Act my_act = {};
		if (lane < nb) my_act = act[b0 + lane];
		const int box0 = uniform(my_act.box);
		const bool one_box = __ballot(lane < nb && my_act.box != box0) == 0ull;
		const WalkBox W0 = load_walk_box(A.boxes + box0);
		// and the entries' sink nodes, staged in LDS (entry e's at s_node[e];
		// the wave's own slice: a wave barrier orders the lanes' writes and
		// the uniform reads below)
		{
			if (lane < nb) {
				const DNode* ln = one_box ? W0.nodes : A.boxes[my_act.box].nodes;
				s_node[lane] = ln[my_act.self];
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		}
		// ---- pass 1: per-entry reservations, entry e's in lane e
		int my[R_NUM];
#pragma unroll
		for (int q = 0; q < R_NUM; q++) my[q] = 0;
		int my_nx = 0, my_lf = 0, my_skip = 1;
		for (int e = 0; e < nb; e++) {
			const int self = __builtin_amdgcn_readlane(my_act.self, e);
			const int edb = __builtin_amdgcn_readlane(my_act.db, e), edn = __builtin_amdgcn_readlane(my_act.dn, e);
			const int eeb = __builtin_amdgcn_readlane(my_act.eb, e), een = __builtin_amdgcn_readlane(my_act.en, e);
			const WalkBox W = one_box ? W0 : load_walk_box(A.boxes + __builtin_amdgcn_readlane(my_act.box, e));
			const DNode* nodes = W.nodes;
			const DNode s = s_node[e];
			if (s.pend - s.pbegin == 0) continue;  // no sinks: nothing to classify (src/kick.cu:268)
			const bool sleaf = s.flags & 1;
			const WalkCounts w = walk_count(nodes, W.theta_inv, W.h, s, dp, ep, A.dcat, A.ecat, edb, edn, eeb, een, lane);
			nchecks += edn + een;
			// PC pairs (leaf sinks only) as the M2P kernel will run them:
			// every sink of the leaf with every multipole
			if (sleaf) npcs += (unsigned long long) w.pc * (unsigned long long) (s.pend - s.pbegin);
			if (w.efail || (sleaf && w.enext)) {
				if (lane == 0) atomicOr(&C->walk_error, 1);
				continue;
			}
			const int cdn = sleaf ? 2 * w.nx : 2 * w.nx + w.lf;  // direct checks per next entry
			const int cen = sleaf ? 0 : 2 * w.enext;             // Ewald checks per next entry
			const int nnext = sleaf ? (w.nx ? 1 : 0) : 2;
			const int npp = sleaf ? w.lf : 0;
			const int nseg = (w.ecc > 0) + (w.cc > 0) + (w.cp > 0) + (w.pc > 0) + (npp > 0);
			if (lane == e) {
				my[R_E] = w.ecc;
				my[R_CC] = w.cc;
				my[R_CP] = w.cp;
				my[R_PC] = w.pc;
				my[R_PP] = npp;
				my[R_ACT] = nnext;
				my[R_DP] = cdn * nnext;
				my[R_EP] = cen * nnext;
				my[R_SEG] = nseg;
				my_nx = w.nx;
				my_lf = w.lf;
				my_skip = 0;
			}
		}
		// ---- wave scans, then one reservation per list for the batch: lane q
		// reserves list q, so the R_NUM atomics are one instruction and one
		// round trip
		int excl[R_NUM];
		int base[R_NUM];
		bool ok[R_NUM];
		int want = 0;
#pragma unroll
		for (int q = 0; q < R_NUM; q++) {
			int v = my[q];
			for (int o = 1; o < 64; o <<= 1) {
				const int t = __shfl_up(v, o);
				if (lane >= o) v += t;
			}
			excl[q] = v - my[q];
			const int tot = __builtin_amdgcn_readlane(v, 63);
			want = lane == q ? tot : want;
		}
		{
			int r = 0;
			int okq = 1;
			if (lane < R_NUM && want > 0) {
				int* ctr;
				int64_t cap;
				switch (lane) {
				case R_E: ctr = &C->out[kEwald].v; cap = A.cap_out[kEwald]; break;
				case R_CC: ctr = &C->out[kCC].v; cap = A.cap_out[kCC]; break;
				case R_CP: ctr = &C->out[kCP].v; cap = A.cap_out[kCP]; break;
				case R_PC: ctr = &C->out[kPC].v; cap = A.cap_out[kPC]; break;
				case R_PP: ctr = &C->out[kPP].v; cap = A.cap_out[kPP]; break;
				case R_ACT: ctr = &C->act[k + 1]; cap = A.cap_act; break;
				case R_DP: ctr = &C->dp[k + 1]; cap = A.cap_dp; break;
				case R_EP: ctr = &C->ep[k + 1]; cap = A.cap_ep; break;
				default: ctr = &C->seg.v; cap = A.cap_seg; break;
				}
				r = atomicAdd(ctr, want);
				okq = reserve_ok(C, lane, r, want, cap);
			}
#pragma unroll
			for (int q = 0; q < R_NUM; q++) {
				base[q] = __builtin_amdgcn_readlane(r, q);
				ok[q] = __builtin_amdgcn_readlane(okq, q);
			}
		}
		// ---- pass 2: writes
		for (int e = 0; e < nb; e++) {
			if (__builtin_amdgcn_readlane(my_skip, e)) continue;
			const int self = __builtin_amdgcn_readlane(my_act.self, e);
			const int box = __builtin_amdgcn_readlane(my_act.box, e);
			const WalkBox W = one_box ? W0 : load_walk_box(A.boxes + box);
			const DNode* nodes = W.nodes;
			const int db = __builtin_amdgcn_readlane(my_act.db, e), dn = __builtin_amdgcn_readlane(my_act.dn, e);
			const int eb = __builtin_amdgcn_readlane(my_act.eb, e), en = __builtin_amdgcn_readlane(my_act.en, e);
			const DNode s = s_node[e];
			const bool sleaf = s.flags & 1;
			int off[R_NUM], cnt[R_NUM];
#pragma unroll
			for (int q = 0; q < R_NUM; q++) {
				off[q] = base[q] + __builtin_amdgcn_readlane(excl[q], e);
				cnt[q] = __builtin_amdgcn_readlane(my[q], e);
			}
			const int n_nx = __builtin_amdgcn_readlane(my_nx, e);
			const int n_lf = __builtin_amdgcn_readlane(my_lf, e);
			const int nnext = cnt[R_ACT];
			const int cdn = nnext ? cnt[R_DP] / nnext : 0;
			const int cen = nnext ? cnt[R_EP] / nnext : 0;
			const int r_e = off[R_E], r_cc = off[R_CC], r_cp = off[R_CP], r_pc = off[R_PC], r_pp = off[R_PP];
			const int r_act = off[R_ACT], r_dp = off[R_DP], r_ep = off[R_EP];
			// Ewald writes
			{
				int run_cc = 0, run_nx = 0;
				for (int bs = 0; bs < en; bs += 64) {
					const int i = bs + lane;
					int cat = K_NONE;
					int oi = 0;
					if (i < en) {
						oi = ep[eb + i];
						cat = A.ecat[eb + i];
					}
					const unsigned long long mcc = __ballot(cat == K_CC);
					const unsigned long long mnx = __ballot(cat == K_NEXT);
					if (cat == K_CC && ok[R_E]) {
						A.out[kEwald][r_e + run_cc + prefix_below(mcc)] = oi;
					}
					if (cat == K_NEXT && ok[R_EP]) {
						const int2 ch = *reinterpret_cast<const int2*>(nodes[oi].child);
						const int pos = 2 * (run_nx + prefix_below(mnx));
						nep[r_ep + pos] = ch.x;
						nep[r_ep + pos + 1] = ch.y;
						nep[r_ep + cen + pos] = ch.x;
						nep[r_ep + cen + pos + 1] = ch.y;
					}
					run_cc += __popcll(mcc);
					run_nx += __popcll(mnx);
				}
			}
			// direct writes
			{
				int run_cc = 0, run_cp = 0, run_pc = 0, run_lf = 0, run_nx = 0;
				for (int bs = 0; bs < dn; bs += 64) {
					const int i = bs + lane;
					int cat = K_NONE;
					int oi = 0;
					if (i < dn) {
						oi = dp[db + i];
						cat = A.dcat[db + i];
					}
					const unsigned long long mcc = __ballot(cat == K_CC);
					const unsigned long long mcp = __ballot(cat == K_CP);
					const unsigned long long mpc = __ballot(cat == K_PC);
					const unsigned long long mlf = __ballot(cat == K_LEAF);
					const unsigned long long mnx = __ballot(cat == K_NEXT);
					if (cat == K_CC && ok[R_CC]) {
						A.out[kCC][r_cc + run_cc + prefix_below(mcc)] = oi;
					} else if (cat == K_CP && ok[R_CP]) {
						A.out[kCP][r_cp + run_cp + prefix_below(mcp)] = oi;
					} else if (cat == K_PC && ok[R_PC]) {
						A.out[kPC][r_pc + run_pc + prefix_below(mpc)] = oi;
					} else if (cat == K_LEAF) {
						const int rank = run_lf + prefix_below(mlf);
						if (sleaf) {
							if (ok[R_PP]) A.out[kPP][r_pp + rank] = oi;
						} else if (ok[R_DP]) {
							const int pos = 2 * n_nx + rank;
							ndp[r_dp + pos] = oi;
							ndp[r_dp + cdn + pos] = oi;
						}
					} else if (cat == K_NEXT && ok[R_DP]) {
						const int2 ch = *reinterpret_cast<const int2*>(nodes[oi].child);
						const int pos = 2 * (run_nx + prefix_below(mnx));
						ndp[r_dp + pos] = ch.x;
						ndp[r_dp + pos + 1] = ch.y;
						if (!sleaf) {
							ndp[r_dp + cdn + pos] = ch.x;
							ndp[r_dp + cdn + pos + 1] = ch.y;
						}
					}
					run_cc += __popcll(mcc);
					run_cp += __popcll(mcp);
					run_pc += __popcll(mpc);
					run_lf += __popcll(mlf);
					run_nx += __popcll(mnx);
				}
			}
			(void) n_lf;
			// next entries and list segments
			if (lane == 0) {
				if (nnext && ok[R_ACT] && ok[R_DP] && ok[R_EP]) {
					Act en_;
					en_.box = box;
					en_.pad[0] = en_.pad[1] = 0;
					if (sleaf) {
						en_.self = self;
						en_.db = r_dp;
						en_.dn = cdn;
						en_.eb = 0;
						en_.en = 0;
						nact_out[r_act] = en_;
					} else {
						en_.self = s.child[0];
						en_.db = r_dp;
						en_.dn = cdn;
						en_.eb = r_ep;
						en_.en = cen;
						nact_out[r_act] = en_;
						en_.self = s.child[1];
						en_.db = r_dp + cdn;
						en_.eb = r_ep + cen;
						nact_out[r_act + 1] = en_;
					}
				}
				if (ok[R_SEG]) {
					int sid = off[R_SEG];
					const int kinds[5] = {kEwald, kCC, kCP, kPC, kPP};
					const int offs[5] = {r_e, r_cc, r_cp, r_pc, r_pp};
					const int cnts[5] = {cnt[R_E], cnt[R_CC], cnt[R_CP], cnt[R_PC], cnt[R_PP]};
					for (int q = 0; q < 5; q++) {
						if (cnts[q] == 0) continue;
						Seg sg;
						sg.begin = offs[q];
						sg.len = cnts[q];
						int32_t* hp = W.head + int64_t(kinds[q]) * W.nnodes + self;
						sg.next = *hp;
						sg.pad = 0;
						A.seg[sid] = sg;
						*hp = sid;
						sid++;
					}
				}
			}
		}
	}
	if (lane == 0) {
		if (nchecks) atomicAdd(&C->walk_checks[wave % kStatSlots].v, nchecks);
		if (npcs) atomicAdd(&C->walk_pc[wave % kStatSlots].v, npcs);
	}
}

struct InitArgs {
	int32_t* head;  // the batch's head arena
	int64_t nhead;
	Counters* cnt;
	Act* act0;
	int32_t* dp0;
	int32_t* ep0;
	int32_t nbox;
};

__global__ void k_walk_init(InitArgs A) {
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	const int64_t stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t j = i; j < A.nhead; j += stride) {
		A.head[j] = -1;
	}
	int32_t* c = reinterpret_cast<int32_t*>(A.cnt);
	const int64_t nc = sizeof(Counters) / sizeof(int32_t);
	for (int64_t j = i; j < nc; j += stride) {
		c[j] = 0;
	}
	// one entry per box: its root with the check lists {root}, {root}
	for (int64_t b = i; b < A.nbox; b += stride) {
		Act e;
		e.self = 0;
		e.db = int32_t(b);
		e.dn = 1;
		e.eb = int32_t(b);
		e.en = 1;
		e.box = int32_t(b);
		e.pad[0] = e.pad[1] = 0;
		A.act0[b] = e;
		A.dp0[b] = 0;
		A.ep0[b] = 0;
	}
}

__global__ void k_walk_seed_counts(Counters* C, int nbox) {
	if (threadIdx.x == 0 && blockIdx.x == 0) {
		C->act[0] = nbox;
		C->dp[0] = nbox;
		C->ep[0] = nbox;
	}
}

// ---------------------------------------------------------------------------
// cell interactions: thread per node
// ---------------------------------------------------------------------------

// Arithmetic modes of the interaction kernels (the walk is the same in all):
//   kF32C  FP32 arithmetic, FP64 sink sums per tile (compensated, the default);
//   kF32U  FP32 arithmetic and FP32 running sums (uncompensated);
//   kF64   FP64 arithmetic throughout (multipoles, expansions, P2P, Ewald).
enum { kF32C = 0, kF32U = 1, kF64 = 2 };
template<int PR> struct Prec {
	using T = float;   // arithmetic
	using A = double;  // sink accumulators
};
template<> struct Prec<kF32U> {
	using T = float;
	using A = float;
};
template<> struct Prec<kF64> {
	using T = double;
	using A = double;
};

// minimum-image difference a - b in box units (cosmictiger/fixed.hpp:303-351)
template<class T> __device__ __forceinline__ T tdist(uint32_t a, uint32_t b);
template<> __device__ __forceinline__ float tdist<float>(uint32_t a, uint32_t b) {
	return fdist(a, b);
}
template<> __device__ __forceinline__ double tdist<double>(uint32_t a, uint32_t b) {
	return double(int32_t(a - b)) * (1.0 / 4294967296.0);
}

template<class T> __device__ __forceinline__ const T* ewald_fc();
template<> __device__ __forceinline__ const float* ewald_fc<float>() {
	return c_fc;
}
template<> __device__ __forceinline__ const double* ewald_fc<double>() {
	return c_fc64;
}

struct CellArgs {
	const BoxDesc* boxes;
	const int32_t* node_off;  // [nbox + 1]: the boxes' nodes in one index space
	int32_t nbox;
	int32_t total;  // node_off[nbox]
	const Seg* seg;
	const int32_t* out_e;
	const int32_t* out_cc;
	const int32_t* out_cp;
	Counters* cnt;
};

// Box b with off[b] <= g < off[b + 1] (off non-decreasing, off[0] = 0).
__device__ __forceinline__ int find_box(const int32_t* __restrict__ off, int nbox, int g) {
	int lo = 0, hi = nbox;
	while (hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if (off[mid] <= g) {
			lo = mid;
		} else {
			hi = mid;
		}
	}
	return lo;
}

__device__ __forceinline__ void load_multi(const float* __restrict__ mp, float* M, uint32_t* pos) {
	const float4* m4 = reinterpret_cast<const float4*>(mp);
	float tmp[40];
#pragma unroll
	for (int q = 0; q < 10; q++) {
		const float4 v = m4[q];
		tmp[4 * q + 0] = v.x;
		tmp[4 * q + 1] = v.y;
		tmp[4 * q + 2] = v.z;
		tmp[4 * q + 3] = v.w;
	}
#pragma unroll
	for (int q = 0; q < 37; q++) {
		M[q] = tmp[q];
	}
	pos[0] = __float_as_uint(tmp[37]);
	pos[1] = __float_as_uint(tmp[38]);
	pos[2] = __float_as_uint(tmp[39]);
}

// FP64 record: 37 moments, then the fixed32 centre as three exact doubles.
__device__ __forceinline__ void load_multi(const double* __restrict__ mp, double* M, uint32_t* pos) {
	const double2* m2 = reinterpret_cast<const double2*>(mp);
	double tmp[40];
#pragma unroll
	for (int q = 0; q < 20; q++) {
		const double2 v = m2[q];
		tmp[2 * q + 0] = v.x;
		tmp[2 * q + 1] = v.y;
	}
#pragma unroll
	for (int q = 0; q < 37; q++) {
		M[q] = tmp[q];
	}
	pos[0] = uint32_t(tmp[37]);
	pos[1] = uint32_t(tmp[38]);
	pos[2] = uint32_t(tmp[39]);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
	for (int o = 32; o > 0; o >>= 1) {
		v += __shfl_xor(v, o);
	}
	return v;
}

// Sum of v over the 64 lanes of the wave, in every lane.
template<class T>
__device__ __forceinline__ T wave_sum_f(T v) {
	for (int o = 32; o > 0; o >>= 1) {
		v += __shfl_xor(v, o);
	}
	return v;
}

// Reduces the per-lane partial expansions over the wave and adds them to the
// node's L (lane q < 50 writes component q).
template<class T>
__device__ __forceinline__ void wave_add_expansion(T* __restrict__ Lo, T (&L)[50], int lane, bool overwrite) {
	T mine = T(0);
#pragma unroll
	for (int q = 0; q < 50; q++) {
		const T t = wave_sum_f(L[q]);
		mine = lane == q ? t : mine;
	}
	if (lane < 50) {
		Lo[lane] = overwrite ? mine : Lo[lane] + mine;
	}
}

constexpr int kCellTile = 512;

// Ewald CC (src/gravity.cu:213-244): one wave per sink node, a source cell per
// lane, then a wave reduction of the 50-float expansion.  Writes L of every
// node (zero when the node has no Ewald list), so it runs first.
template<class T>
__global__ __launch_bounds__(64) void k_cells_ewald(CellArgs A) {
	const int lane = threadIdx.x;
	unsigned long long nev = 0;
	// a node at a time per workgroup; the box descriptor is kept in scalar
	// registers
	for (int g = blockIdx.x; g < A.total; g += gridDim.x) {
		const int bx = uniform(find_box(A.node_off, A.nbox, g));
		const BoxDesc Bv = load_uniform(A.boxes + bx);
		const BoxDesc* B = &Bv;
		const int i = g - uniform(A.node_off[bx]);
		const T* multi = static_cast<const T*>(B->multi);
		T* Lo = static_cast<T*>(B->L) + int64_t(i) * kLStride;
		const bool do_phi = B->prm.do_phi;
		const int h0 = uniform(B->head[int64_t(kEwald) * B->nnodes + i]);
		if (h0 < 0) {
			if (lane < 50) Lo[lane] = T(0);
			continue;
		}
		const uint32_t sp0 = uniform(int(B->nodes[i].pos[0])), sp1 = uniform(int(B->nodes[i].pos[1])), sp2 = uniform(int(B->nodes[i].pos[2]));
		T L[50];
#pragma unroll
		for (int q = 0; q < 50; q++) L[q] = T(0);
		for (int sg = h0; sg >= 0; sg = uniform(A.seg[sg].next)) {
			const int sb = uniform(A.seg[sg].begin), sl = uniform(A.seg[sg].len);
			for (int j = lane; j < sl; j += 64) {
				const int o = A.out_e[sb + j];
				T M[37];
				uint32_t op[3];
				load_multi(multi + int64_t(o) * kMultiStride, M, op);
				T D[50];
				ctfmm::ewald_greens<T, T>(D, tdist<T>(sp0, op[0]), tdist<T>(sp1, op[1]), tdist<T>(sp2, op[2]), c_rimg, c_nreal, c_fh, ewald_fc<T>(),
						c_nfour);
				ctfmm::M2L<T>(L, M, D, do_phi);
				nev++;
			}
		}
		wave_add_expansion(Lo, L, lane, true);
	}
	nev = wave_sum_u64(nev);
	if (lane == 0 && nev) atomicAdd(&A.cnt->stat[2], nev);
}

// Direct CC = M2L (src/gravity.cu:44-73), a source cell per lane (CP =
// false), and CP = P2L (src/gravity.cu:75-155), the particles of the CP
// leaves staged through LDS and a source particle per lane (CP = true); one
// wave per sink node.  Two launches, so that the P2L loop is not held to the
// register budget of M2L.
template<class T, bool CP>
__global__ __launch_bounds__(64) void k_cells_direct(CellArgs A) {
	__shared__ uint32_t tx[kCellTile], ty[kCellTile], tz[kCellTile];
	const int lane = threadIdx.x;
	unsigned long long ncc = 0, ncp = 0;
	// a node at a time per workgroup; the box descriptor is kept in scalar
	// registers
	for (int g = blockIdx.x; g < A.total; g += gridDim.x) {
		const int bx = uniform(find_box(A.node_off, A.nbox, g));
		const BoxDesc Bv = load_uniform(A.boxes + bx);
		const BoxDesc* B = &Bv;
		const int i = g - uniform(A.node_off[bx]);
		const T* multi = static_cast<const T*>(B->multi);
		const bool do_phi = B->prm.do_phi;
		const int hcc = CP ? -1 : uniform(B->head[int64_t(kCC) * B->nnodes + i]);
		const int hcp = CP ? uniform(B->head[int64_t(kCP) * B->nnodes + i]) : -1;
		if (hcc < 0 && hcp < 0) continue;
		const uint32_t sp0 = uniform(int(B->nodes[i].pos[0])), sp1 = uniform(int(B->nodes[i].pos[1])), sp2 = uniform(int(B->nodes[i].pos[2]));
		T L[50];
#pragma unroll
		for (int q = 0; q < 50; q++) L[q] = T(0);
		for (int sg = hcc; sg >= 0; sg = uniform(A.seg[sg].next)) {
			const int sb = uniform(A.seg[sg].begin), sl = uniform(A.seg[sg].len);
			for (int j = lane; j < sl; j += 64) {
				const int o = A.out_cc[sb + j];
				T M[37];
				uint32_t op[3];
				load_multi(multi + int64_t(o) * kMultiStride, M, op);
				T D[50];
				ctfmm::greens<T>(D, tdist<T>(sp0, op[0]), tdist<T>(sp1, op[1]), tdist<T>(sp2, op[2]));
				ctfmm::M2L<T>(L, M, D, do_phi);
				ncc++;
			}
		}
		int tn = 0;
		auto flush = [&]() {
			__syncthreads();
			for (int t = lane; t < tn; t += 64) {
				T D[50];
				ctfmm::greens<T>(D, tdist<T>(sp0, tx[t]), tdist<T>(sp1, ty[t]), tdist<T>(sp2, tz[t]));
#pragma unroll
				for (int q = 0; q < 50; q++) L[q] += D[q];
			}
			ncp += (unsigned long long) (lane < tn ? (tn - lane + 63) / 64 : 0);
			__syncthreads();
			tn = 0;
		};
		for (int sg = hcp; sg >= 0; sg = uniform(A.seg[sg].next)) {
			const int sb = uniform(A.seg[sg].begin), sl = uniform(A.seg[sg].len);
			// the particle ranges of up to 64 CP leaves at once, one per lane,
			// so that staging waits on no list or node load
			int my_pb = 0, my_pe = 0;
			for (int j = 0; j < sl; j++) {
				if ((j & 63) == 0) {
					my_pb = my_pe = 0;
					if (j + lane < sl) {
						const DNode& on = B->nodes[A.out_cp[sb + j + lane]];
						my_pb = on.pbegin;
						my_pe = on.pend;
					}
				}
				int pb = __builtin_amdgcn_readlane(my_pb, j & 63);
				const int pe = __builtin_amdgcn_readlane(my_pe, j & 63);
				while (pb < pe) {
					const int m = min(kCellTile - tn, pe - pb);
					for (int t = lane; t < m; t += 64) {
						tx[tn + t] = B->x[pb + t];
						ty[tn + t] = B->y[pb + t];
						tz[tn + t] = B->z[pb + t];
					}
					tn += m;
					pb += m;
					if (tn == kCellTile) flush();
				}
			}
		}
		if (tn) flush();
		wave_add_expansion(static_cast<T*>(B->L) + int64_t(i) * kLStride, L, lane, false);
	}
	ncc = wave_sum_u64(ncc);
	ncp = wave_sum_u64(ncp);
	if (lane == 0) {
		if (ncc) atomicAdd(&A.cnt->stat[1], ncc);
		if (ncp) atomicAdd(&A.cnt->cp_parts[blockIdx.x % kStatSlots].v, ncp);
	}
}

// ---------------------------------------------------------------------------
// L2L sweep, one launch per level
// ---------------------------------------------------------------------------

// Work list of one tree level of every box: entry t is node lo[b] + t - pref[b]
// of box b, pref[b] <= t < pref[b + 1].
struct L2LArgs {
	const BoxDesc* boxes;
	const int32_t* pref;  // [nbox + 1]
	const int32_t* lo;    // [nbox]
	int32_t nbox;
	int32_t total;
};

template<class T>
__global__ __launch_bounds__(64) void k_l2l(L2LArgs A) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= A.total) return;
	const int bx = find_box(A.pref, A.nbox, t);
	const BoxDesc* B = A.boxes + bx;
	const int i = A.lo[bx] + t - A.pref[bx];
	const DNode s = B->nodes[i];
	if (s.parent < 0) return;
	const DNode p = B->nodes[s.parent];
	T* L = static_cast<T*>(B->L);
	const T* La = L + int64_t(s.parent) * kLStride;
	T Lp[50];
#pragma unroll
	for (int q = 0; q < 50; q++) Lp[q] = La[q];
	T Lb[50];
	ctfmm::L2L<T>(Lb, Lp, tdist<T>(s.pos[0], p.pos[0]), tdist<T>(s.pos[1], p.pos[1]), tdist<T>(s.pos[2], p.pos[2]), B->prm.do_phi);
	T* Lo = L + int64_t(i) * kLStride;
#pragma unroll
	for (int q = 0; q < 50; q++) Lo[q] += Lb[q];
}

// ---------------------------------------------------------------------------
// leaf kernel: wave per leaf
// ---------------------------------------------------------------------------

// Leaf size classes: a leaf of n sinks runs P source phases over S = 64/P
// sink slots, G = ceil(n/S) sinks per lane.  Leaves of 1..96 sinks use P=4
// (G=1..6); 97..256 use P=2 (G=4..8), which keeps the register count (and so
// the occupancy) of the large classes down.
constexpr int kNumLeafClasses = 11;
constexpr int kMaxLeafSinks = 256;

__host__ __device__ inline int leaf_class(int n) {
	return n <= 96 ? (n + 15) / 16 - 1 : 6 + (n + 31) / 32 - 4;
}

struct PPArgs {
	const BoxDesc* boxes;
	const int2* leaves;  // (box, leaf) of this class
	const Seg* seg;
	const int32_t* out_pp;
	Counters* cnt;
};

struct LeafArgs {
	const BoxDesc* boxes;
	const int2* leaves;  // (box, leaf), grouped by box
	const Seg* seg;
	const int32_t* out_pc;
	double* part;  // 8 per entry of leaves
	Counters* cnt;
};

constexpr int kPPTile = 640;  // sources per LDS tile (multiple of 16; ~8 KB per block)

// Positions in the P2P kernel are offsets from the sink leaf's centre in
// units of 2^-32, converted once per particle: int32(p - c) is exact, and its
// float is exact below 2^24 (a 2^-8 box) and off by at most a unit farther
// out, below the rounding of Delta itself.
template<class T>
__device__ __forceinline__ void load4(const T* p, T& a, T& b, T& c, T& d);
template<>
__device__ __forceinline__ void load4<float>(const float* p, float& a, float& b, float& c, float& d) {
	const float4 v = *reinterpret_cast<const float4*>(p);
	a = v.x;
	b = v.y;
	c = v.z;
	d = v.w;
}
template<>
__device__ __forceinline__ void load4<int32_t>(const int32_t* p, int32_t& a, int32_t& b, int32_t& c, int32_t& d) {
	const int4 v = *reinterpret_cast<const int4*>(p);
	a = v.x;
	b = v.y;
	c = v.z;
	d = v.w;
}

// The P2P source tile's element: the offset as a float in the FP32 modes,
// as the exact int32 in the FP64 mode (converted per pair; a double tile
// would double the LDS traffic).
template<int PR>
struct PPTile {
	using type = float;
};
template<>
struct PPTile<kF64> {
	using type = int32_t;
};

// One softened pair in integer units (Delta = sink - source in units of 2^-32),
// g -= Delta/r^3, phi -= 1/r; Ferrers n=4 inside h (cosmictiger/gravity.hpp:76-96).
template<class T>
__device__ __forceinline__ void pp_pair(T sx, T sy, T sz, T X, T Y, T Z, T h2, T hinv, T h2inv, T h3inv, T& fx, T& fy, T& fz, T& pt) {
	const T dx = sx - X;
	const T dy = sy - Y;
	const T dz = sz - Z;
	const T r2 = ctfmm::ct_fma(dx, dx, ctfmm::ct_fma(dy, dy, dz * dz));
	T rinv = ctfmm::ct_rsqrt(r2);
	T r3 = rinv * rinv * rinv;
	if (r2 <= h2) {
		const T q2 = r2 * h2inv;
		T f = T(315.0 / 128.0);
		f = ctfmm::ct_fma(f, q2, T(-385.0 / 32.0));
		f = ctfmm::ct_fma(f, q2, T(1485.0 / 64.0));
		f = ctfmm::ct_fma(f, q2, T(-693.0 / 32.0));
		f = ctfmm::ct_fma(f, q2, T(1155.0 / 128.0));
		T p = T(-63.0 / 256.0);
		p = ctfmm::ct_fma(p, q2, T(385.0 / 256.0));
		p = ctfmm::ct_fma(p, q2, T(-495.0 / 128.0));
		p = ctfmm::ct_fma(p, q2, T(693.0 / 128.0));
		p = ctfmm::ct_fma(p, q2, T(-1155.0 / 256.0));
		p = ctfmm::ct_fma(p, q2, T(693.0 / 256.0));
		r3 = f * h3inv;
		rinv = p * hinv;
	}
	fx = ctfmm::ct_fma(dx, r3, fx);
	fy = ctfmm::ct_fma(dy, r3, fy);
	fz = ctfmm::ct_fma(dz, r3, fz);
	pt -= rinv;
}

// The same pair without the softening test, for source leaves that cannot
// come within h of any sink (no pair of the two leaves has r <= h).
template<class T>
__device__ __forceinline__ void pp_pair_far(T sx, T sy, T sz, T X, T Y, T Z, T& fx, T& fy, T& fz, T& pt) {
	const T dx = sx - X;
	const T dy = sy - Y;
	const T dz = sz - Z;
	const T r2 = ctfmm::ct_fma(dx, dx, ctfmm::ct_fma(dy, dy, dz * dz));
	const T rinv = ctfmm::ct_rsqrt(r2);
	const T r3 = rinv * rinv * rinv;
	fx = ctfmm::ct_fma(dx, r3, fx);
	fy = ctfmm::ct_fma(dy, r3, fy);
	fz = ctfmm::ct_fma(dz, r3, fz);
	pt -= rinv;
}

// Adds FP32/FP64 tile sums in integer units to the sink accumulators in box
// units (force 2^64, potential 2^32) and clears the tile sums.
template<class Acc, class T>
__device__ __forceinline__ void fold_pp(Acc& gx, Acc& gy, Acc& gz, Acc& ph, T& fx, T& fy, T& fz, T& pt) {
	gx -= Acc(fx) * Acc(18446744073709551616.0);
	gy -= Acc(fy) * Acc(18446744073709551616.0);
	gz -= Acc(fz) * Acc(18446744073709551616.0);
	ph += Acc(pt) * Acc(4294967296.0);
	fx = fy = fz = pt = T(0);
}

// P2P (src/gravity.cu:246-404): one wave per leaf.  Lane (slot, phase) holds
// sinks slot, slot+S, ... and runs over the sources of its phase; the sources
// of the leaf's PP list are staged through LDS in tiles of kPPTile, source i of a
// tile at row i%P, column i/P (rows padded by 16 words so the P rows sit in
// different banks).  Each tile is summed in FP32, then added to FP64 sink
// accumulators (compensated at tile granularity); the phases are summed at the
// end and the result is stored per particle for the kick kernel.
template<int G, int P, int PR>
__global__ __launch_bounds__(64) void k_pp(PPArgs A) {
	using T = typename Prec<PR>::T;
	using Acc = typename Prec<PR>::A;
	constexpr int S = 64 / P;
	constexpr int STR = kPPTile / P + 16;
	using TS = typename PPTile<PR>::type;
	__shared__ __align__(16) TS tx[P * STR];
	__shared__ __align__(16) TS ty[P * STR];
	__shared__ __align__(16) TS tz[P * STR];
	const int2 le = A.leaves[blockIdx.x];
	const BoxDesc Bv = load_uniform(A.boxes + uniform(le.x));
	const BoxDesc* B = &Bv;
	const int self = uniform(le.y);
	const DNode* nodes = B->nodes;
	const uint32_t* px = B->x;
	const uint32_t* py = B->y;
	const uint32_t* pz = B->z;
	const int pbeg = uniform(nodes[self].pbegin);
	const int nsink = uniform(nodes[self].pend) - pbeg;
	// the sink leaf's centre: origin of the offsets
	const uint32_t cs0 = uniform(int(nodes[self].pos[0])), cs1 = uniform(int(nodes[self].pos[1])), cs2 = uniform(int(nodes[self].pos[2]));
	const int lane = threadIdx.x;
	const int slot = lane % S, phase = lane / S;
	T sx[G], sy[G], sz[G];
	T fx[G], fy[G], fz[G], pt[G];
	Acc gx[G], gy[G], gz[G], ph[G];
	const float hf = B->h;
#pragma unroll
	for (int g = 0; g < G; g++) {
		int k = g * S + slot;
		if (k >= nsink) k = 0;
		sx[g] = T(int32_t(px[pbeg + k] - cs0));
		sy[g] = T(int32_t(py[pbeg + k] - cs1));
		sz[g] = T(int32_t(pz[pbeg + k] - cs2));
		fx[g] = fy[g] = fz[g] = pt[g] = T(0);
		gx[g] = gy[g] = gz[g] = Acc(0);
		ph[g] = phase == 0 ? Acc(PR == kF64 ? double(kSelfPhi) / double(hf) : double(kSelfPhi / hf)) : Acc(0);
	}
	// integer units: h in units of 2^-32
	const T hi_ = T(hf) * T(4294967296.0);
	const T h2 = hi_ * hi_;
	const T ihinv = T(1) / hi_;
	const T ih2inv = ihinv * ihinv;
	const T ih3inv = ih2inv * ihinv;
	unsigned long long nsrc = 0;
	int tn = 0;
	const TS* bx = tx + phase * STR;
	const TS* by = ty + phase * STR;
	const TS* bz = tz + phase * STR;
	// near = true: a tile of sources that may lie within h of a sink (checked
	// pairs with the softening branch); near = false: the branch-free pair.
	auto flush = [&](auto near_tag) __attribute__((always_inline)) {
		constexpr bool near = decltype(near_tag)::value;
		__syncthreads();
		const int q = (tn & ~(4 * P - 1)) / P;  // full groups of 4 columns in every row
		// sink-major: the tile is read once per sink so that the temporaries do
		// not grow with G (LDS reads are ~4% of the VALU cycles)
#pragma unroll
		for (int g = 0; g < G; g++) {
			T ax = fx[g], ay = fy[g], az = fz[g], ap = pt[g];
			const T X0 = sx[g], Y0 = sy[g], Z0 = sz[g];
			for (int t = 0; t < q; t += 4) {
				struct {
					TS x, y, z, w;
				} X4, Y4, Z4;
				load4(bx + t, X4.x, X4.y, X4.z, X4.w);
				load4(by + t, Y4.x, Y4.y, Y4.z, Y4.w);
				load4(bz + t, Z4.x, Z4.y, Z4.z, Z4.w);
				if constexpr (near) {
					pp_pair(X0, Y0, Z0, T(X4.x), T(Y4.x), T(Z4.x), h2, ihinv, ih2inv, ih3inv, ax, ay, az, ap);
					pp_pair(X0, Y0, Z0, T(X4.y), T(Y4.y), T(Z4.y), h2, ihinv, ih2inv, ih3inv, ax, ay, az, ap);
					pp_pair(X0, Y0, Z0, T(X4.z), T(Y4.z), T(Z4.z), h2, ihinv, ih2inv, ih3inv, ax, ay, az, ap);
					pp_pair(X0, Y0, Z0, T(X4.w), T(Y4.w), T(Z4.w), h2, ihinv, ih2inv, ih3inv, ax, ay, az, ap);
				} else {
					pp_pair_far(X0, Y0, Z0, T(X4.x), T(Y4.x), T(Z4.x), ax, ay, az, ap);
					pp_pair_far(X0, Y0, Z0, T(X4.y), T(Y4.y), T(Z4.y), ax, ay, az, ap);
					pp_pair_far(X0, Y0, Z0, T(X4.z), T(Y4.z), T(Z4.z), ax, ay, az, ap);
					pp_pair_far(X0, Y0, Z0, T(X4.w), T(Y4.w), T(Z4.w), ax, ay, az, ap);
				}
			}
			for (int t = q; t < q + 4; t++) {
				if (t * P + phase < tn) {
					if constexpr (near) {
						pp_pair(X0, Y0, Z0, T(bx[t]), T(by[t]), T(bz[t]), h2, ihinv, ih2inv, ih3inv, ax, ay, az, ap);
					} else {
						pp_pair_far(X0, Y0, Z0, T(bx[t]), T(by[t]), T(bz[t]), ax, ay, az, ap);
					}
				}
			}
			fx[g] = ax;
			fy[g] = ay;
			fz[g] = az;
			pt[g] = ap;
		}
		// tile sums into the sink accumulators (all but the uncompensated mode)
		if constexpr (PR != kF32U) {
#pragma unroll
			for (int g = 0; g < G; g++) {
				fold_pp(gx[g], gy[g], gz[g], ph[g], fx[g], fy[g], fz[g], pt[g]);
			}
		}
		__syncthreads();
		tn = 0;
	};
	// A source leaf is near when its bounding sphere comes within h of the
	// sink leaf's: dist(centres) - r_s - r_o <= h (plus 1e-6 for the float
	// rounding of the centres, radii and distance).  Pass 0 stages the far
	// leaves, pass 1 the near ones.
	const float rs = nodes[self].radius;
	const int hpp = uniform(B->head[int64_t(kPP) * B->nnodes + self]);
	// stage particles [pb, pe) into the tile, flushing whenever it fills
	auto stage = [&](int pb, const int pe, bool nearpass) __attribute__((always_inline)) {
		while (pb < pe) {
			const int m = min(kPPTile - tn, pe - pb);
#pragma unroll 4
			for (int t = lane; t < m; t += 64) {
				const int i = tn + t;
				const int at = (i % P) * STR + i / P;
				tx[at] = TS(int32_t(px[pb + t] - cs0));
				ty[at] = TS(int32_t(py[pb + t] - cs1));
				tz[at] = TS(int32_t(pz[pb + t] - cs2));
			}
			tn += m;
			pb += m;
			if (tn == kPPTile) {
				if (nearpass) {
					flush(std::true_type());
				} else {
					flush(std::false_type());
				}
			}
		}
	};
	for (int pass = 0; pass < 2; pass++) {
		// source leaves adjacent in memory are merged into one range
		int rb = 0, re = 0;
		for (int sg = hpp; sg >= 0; sg = uniform(A.seg[sg].next)) {
			const int sb = uniform(A.seg[sg].begin), sl = uniform(A.seg[sg].len);
			for (int j0 = 0; j0 < sl; j0 += 64) {
				// up to 64 source leaves at once, one per lane: their ranges and
				// near flags, so that the staging loop below waits on no list
				// or node load
				int my_pb = 0, my_pe = 0;
				bool my_near = false;
				if (j0 + lane < sl) {
					const DNode& on = nodes[A.out_pp[sb + j0 + lane]];
					const float ddx = fdist(cs0, on.pos[0]);
					const float ddy = fdist(cs1, on.pos[1]);
					const float ddz = fdist(cs2, on.pos[2]);
					const float dd = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
					my_near = dd - (rs + on.radius) <= hf + 1e-6f;
					my_pb = on.pbegin;
					my_pe = on.pend;
				}
				const unsigned long long mine = __ballot(my_near == (pass == 1));
				const int cnt = min(64, sl - j0);
				for (int jj = 0; jj < cnt; jj++) {
					if (!((mine >> jj) & 1ull)) continue;
					const int pb = __builtin_amdgcn_readlane(my_pb, jj);
					const int pe = __builtin_amdgcn_readlane(my_pe, jj);
					nsrc += pe - pb;
					if (pb == re) {
						re = pe;
					} else {
						stage(rb, re, pass == 1);
						rb = pb;
						re = pe;
					}
				}
			}
		}
		stage(rb, re, pass == 1);
		if (tn) {
			if (pass == 0) {
				flush(std::false_type());
			} else {
				flush(std::true_type());
			}
		}
	}
	if constexpr (PR == kF32U) {
#pragma unroll
		for (int g = 0; g < G; g++) {
			fold_pp(gx[g], gy[g], gz[g], ph[g], fx[g], fy[g], fz[g], pt[g]);
		}
	}
#pragma unroll
	for (int g = 0; g < G; g++) {
		for (int o = S; o < 64; o <<= 1) {
			gx[g] += __shfl_xor(gx[g], o);
			gy[g] += __shfl_xor(gy[g], o);
			gz[g] += __shfl_xor(gz[g], o);
			ph[g] += __shfl_xor(ph[g], o);
		}
		const int k = g * S + slot;
		if (phase == 0 && k < nsink) {
			if constexpr (PR == kF64) {
				static_cast<double4*>(B->pp)[pbeg + k] = make_double4(gx[g], gy[g], gz[g], ph[g]);
			} else {
				static_cast<float4*>(B->pp)[pbeg + k] = make_float4(float(gx[g]), float(gy[g]), float(gz[g]), float(ph[g]));
			}
		}
	}
	if (lane == 0) atomicAdd(&A.cnt->pp_pairs[blockIdx.x % kStatSlots].v, (unsigned long long) nsrc * (unsigned long long) nsink);
}

// M2P (src/gravity.cu:157-211), then L2P and the kick (src/kick.cu:53-177):
// one wave per leaf.  M2P runs over the (sink, multipole) pairs of the leaf,
// pair t on lane t mod 64 with sink t mod nsink, so every lane is busy
// whatever the leaf size; the 4 results of a pair are added to the sink's FP64
// sums in LDS.  L2P and the kick then run a sink per lane.
template<int PR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_kick(LeafArgs A) {
	using T = typename Prec<PR>::T;
	using Acc = typename Prec<PR>::A;
	__shared__ Acc acc[4 * kMaxLeafSinks];
	// a walk that overflowed its arenas left incomplete lists: the kick is
	// redone eagerly, so this pass must not touch the particles
	if (__builtin_amdgcn_readfirstlane(A.cnt->overflow | A.cnt->walk_error)) return;
	const int li = blockIdx.x;
	const int2 le = A.leaves[li];
	const BoxDesc Bv = load_uniform(A.boxes + uniform(le.x));
	const BoxDesc* B = &Bv;
	const int self = uniform(le.y);
	const T* multi = static_cast<const T*>(B->multi);
	const DNode s = B->nodes[self];
	const int pbeg = s.pbegin;
	const int nsink = s.pend - s.pbegin;
	const int lane = threadIdx.x;
	const ct_kick_params& P_ = Bv.prm;
	const uint32_t* px = B->x;
	const uint32_t* py = B->y;
	const uint32_t* pz = B->z;
	const int hpc = uniform(B->head[int64_t(kPC) * B->nnodes + self]);
	if (hpc >= 0) {
		for (int k = lane; k < 4 * nsink; k += 64) acc[k] = Acc(0);
		__syncthreads();
		for (int sg = hpc; sg >= 0; sg = uniform(A.seg[sg].next)) {
			const int sb = uniform(A.seg[sg].begin), sl = uniform(A.seg[sg].len);
			const int npair = nsink * sl;
			for (int t = lane; t < npair; t += 64) {
				const int j = t / nsink;
				const int k = t - j * nsink;
				const int pi = pbeg + k;
				const int o = A.out_pc[sb + j];
				T M[37];
				uint32_t op[3];
				load_multi(multi + int64_t(o) * kMultiStride, M, op);
				T D[50];
				ctfmm::greens<T>(D, tdist<T>(px[pi], op[0]), tdist<T>(py[pi], op[1]), tdist<T>(pz[pi], op[2]));
				T L4[4] = {T(0), T(0), T(0), T(0)};
				ctfmm::M2L2<T>(L4, M, D, P_.do_phi);
				atomicAdd(&acc[4 * k + 0], -Acc(T(1e6) * L4[1]));
				atomicAdd(&acc[4 * k + 1], -Acc(T(1e6) * L4[2]));
				atomicAdd(&acc[4 * k + 2], -Acc(T(1e6) * L4[3]));
				atomicAdd(&acc[4 * k + 3], Acc(T(1e3) * L4[0]));
			}
		}
		__syncthreads();
	}
	// ---- L2P + kick, a sink per lane
	T Ls[50];
	{
		const T* Lp = static_cast<const T*>(B->L) + int64_t(self) * kLStride;
#pragma unroll
		for (int q = 0; q < 50; q++) Ls[q] = Lp[q];
	}
	double kin = 0, dkin = 0, xm = 0, ym = 0, zm = 0, nm = 0, pot = 0;
	int maxr = 0;
	const float hf = P_.h;
	const float sgn = P_.sign * (P_.top ? 1.f : -1.f);
	for (int k = lane; k < nsink; k += 64) {
		const int pi = pbeg + k;
		Acc gxd, gyd, gzd, phd;
		if constexpr (PR == kF64) {
			const double4 pp = static_cast<const double4*>(B->pp)[pi];
			gxd = pp.x;
			gyd = pp.y;
			gzd = pp.z;
			phd = pp.w;
		} else {
			const float4 pp = static_cast<const float4*>(B->pp)[pi];
			gxd = Acc(pp.x);
			gyd = Acc(pp.y);
			gzd = Acc(pp.z);
			phd = Acc(pp.w);
		}
		if (hpc >= 0) {
			gxd += acc[4 * k + 0];
			gyd += acc[4 * k + 1];
			gzd += acc[4 * k + 2];
			phd += acc[4 * k + 3];
		}
		const uint32_t X = px[pi], Y = py[pi], Z = pz[pi];
		T L4[4];
		ctfmm::L2P<T>(L4, Ls, tdist<T>(X, s.pos[0]), tdist<T>(Y, s.pos[1]), tdist<T>(Z, s.pos[2]), P_.do_phi);
		const float Fphi = float((T(phd) + T(1e3) * L4[0]) * T(P_.GM));
		const float Fx = float((T(gxd) - T(1e6) * L4[1]) * T(P_.GM));
		const float Fy = float((T(gyd) - T(1e6) * L4[2]) * T(P_.GM));
		const float Fz = float((T(gzd) - T(1e6) * L4[3]) * T(P_.GM));
		float vx = B->vx[pi], vy = B->vy[pi], vz = B->vz[pi];
		const float kin0 = 0.5f * (vx * vx + vy * vy + vz * vz);
		float dt;
		if (P_.ascending) {
			dt = 0.5f * c_rung_dt[P_.min_rung] * P_.t0;
			if (!P_.first_call) {
				vx = __builtin_fmaf(sgn * Fx, dt, vx);
				vy = __builtin_fmaf(sgn * Fy, dt, vy);
				vz = __builtin_fmaf(sgn * Fz, dt, vz);
			}
		}
		kin += 0.5 * double(vx * vx + vy * vy + vz * vz);
		xm += vx;
		ym += vy;
		zm += vz;
		nm += sqrtf(vx * vx + vy * vy + vz * vz);
		if (P_.save_force) {
			B->gx[pi] = Fx;
			B->gy[pi] = Fy;
			B->gz[pi] = Fz;
			B->phi[pi] = Fphi;
		}
		if (P_.descending) {
			const float g2 = Fx * Fx + Fy * Fy + Fz * Fz;
			dt = fminf(P_.eta * sqrtf(P_.a * hf * rsqrtf(g2)), P_.t0);
			dt = fminf(P_.max_dt, dt);
			int r = B->rung[pi];
			r = max(P_.min_rung + int(int(ceilf(log2f(P_.t0 / dt))) > P_.min_rung), r - 1);
			B->rung[pi] = int8_t(r);
			maxr = max(maxr, r);
			dt = 0.5f * c_rung_dt[P_.min_rung] * P_.t0;
			vx = __builtin_fmaf(sgn * Fx, dt, vx);
			vy = __builtin_fmaf(sgn * Fy, dt, vy);
			vz = __builtin_fmaf(sgn * Fz, dt, vz);
		}
		const float kin1 = 0.5f * (vx * vx + vy * vy + vz * vz);
		B->vx[pi] = vx;
		B->vy[pi] = vy;
		B->vz[pi] = vz;
		pot += 0.5 * double(Fphi);
		dkin += double(kin1) - double(kin0);
	}
	// wave reductions (FP64)
	for (int o = 32; o > 0; o >>= 1) {
		kin += __shfl_xor(kin, o);
		dkin += __shfl_xor(dkin, o);
		xm += __shfl_xor(xm, o);
		ym += __shfl_xor(ym, o);
		zm += __shfl_xor(zm, o);
		nm += __shfl_xor(nm, o);
		pot += __shfl_xor(pot, o);
		maxr = max(maxr, __shfl_xor(maxr, o));
	}
	if (lane == 0) {
		double* part = A.part + int64_t(li) * 8;
		part[0] = double(maxr);
		part[1] = pot;
		part[2] = kin;
		part[3] = dkin;
		part[4] = xm;
		part[5] = ym;
		part[6] = zm;
		part[7] = nm;
	}
}

#define CT_LEAF_CLASSES(X) X(0, 1, 4) X(1, 2, 4) X(2, 3, 4) X(3, 4, 4) X(4, 5, 4) X(5, 6, 4) \
	X(6, 4, 2) X(7, 5, 2) X(8, 6, 2) X(9, 7, 2) X(10, 8, 2)

// Deterministic FP64 reduction of the per-leaf kick_return partials (8
// doubles per leaf: max_rung by max, the rest by sum).  Stage 1: block c
// reduces the records [r.x, r.y) of chunk c into out[8 c ..]; chunks never
// cross a box boundary.  Stage 2: block b reduces the chunk results
// [off[b], off[b + 1]) of box b into out[8 b ..].
constexpr int kReduceChunk = 4096;

__device__ __forceinline__ void reduce8(const double* __restrict__ part, int b, int e, double* __restrict__ out) {
	__shared__ double sh[8][256];
	const int t = threadIdx.x;
	double acc[8];
	for (int q = 0; q < 8; q++) acc[q] = 0.0;
	for (int i = b + t; i < e; i += 256) {
		acc[0] = fmax(acc[0], part[int64_t(i) * 8]);
		for (int q = 1; q < 8; q++) acc[q] += part[int64_t(i) * 8 + q];
	}
	for (int q = 0; q < 8; q++) sh[q][t] = acc[q];
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (t < w) {
			sh[0][t] = fmax(sh[0][t], sh[0][t + w]);
			for (int q = 1; q < 8; q++) sh[q][t] += sh[q][t + w];
		}
		__syncthreads();
	}
	if (t < 8) out[t] = sh[t][0];
}

__global__ __launch_bounds__(256) void k_reduce_chunks(const double* __restrict__ part, const int2* __restrict__ chunks,
		double* __restrict__ out) {
	const int2 r = chunks[blockIdx.x];
	reduce8(part, r.x, r.y, out + int64_t(blockIdx.x) * 8);
}

__global__ __launch_bounds__(256) void k_reduce_boxes(const double* __restrict__ chunk_sums, const int32_t* __restrict__ off,
		double* __restrict__ out) {
	reduce8(chunk_sums, off[blockIdx.x], off[blockIdx.x + 1], out + int64_t(blockIdx.x) * 8);
}

// ---------------------------------------------------------------------------
// drift (src/drift.cu:36-154 without the lightcone branch)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_drift(uint32_t* __restrict__ x, uint32_t* __restrict__ y, uint32_t* __restrict__ z, float* __restrict__ vx,
		float* __restrict__ vy, float* __restrict__ vz, const int8_t* __restrict__ rung, int64_t n, int r, double c1, double dragc, bool do_drag) {
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	if (rung[i] != r) return;
	uint32_t* X[3] = {x, y, z};
	float* V[3] = {vx, vy, vz};
#pragma unroll
	for (int d = 0; d < 3; d++) {
		const double x0 = double(X[d][i]) * (1.0 / 4294967296.0);
		double x1 = __builtin_fma(double(V[d][i]), c1, x0);
		x1 -= double(x1 >= 1.0);
		x1 += double(x1 < 0.0);
		X[d][i] = uint32_t(uint64_t(x1 * 4294967296.0));
	}
	if (do_drag) {
#pragma unroll
		for (int d = 0; d < 3; d++) V[d][i] = float(double(V[d][i]) * dragc);
	}
}

// HEALPix nested pixel of a direction (src/healpix.cu:75-172): the face and
// the in-face edge-line indices at nside 8192, interleaved into the nested
// index and degraded to nside by integer division.
__device__ __forceinline__ uint32_t spread_bits13(uint32_t v) {
	v = (v | (v << 8)) & 0x00FF00FFu;
	v = (v | (v << 4)) & 0x0F0F0F0Fu;
	v = (v | (v << 2)) & 0x33333333u;
	v = (v | (v << 1)) & 0x55555555u;
	return v;
}

__device__ int32_t vec2pix_nest(int nside, double vx, double vy, double vz) {
	constexpr int ns_max = 8192;
	const double z = vz / __dsqrt_rn(vx * vx + vy * vy + vz * vz);
	double phi = 0.0;
	if (vx != 0.0 || vy != 0.0) {
		phi = atan2(vy, vx);
		if (phi < 0.0) phi += 2.0 * M_PI;
	}
	const double za = fabs(z);
	const double tt = phi / (0.5 * M_PI);
	int face, ix, iy;
	if (za <= 2.0 / 3.0) {
		const int jp = int(floor(ns_max * (0.5 + tt - z * 0.75)));
		const int jm = int(floor(ns_max * (0.5 + tt + z * 0.75)));
		const int ifp = jp / ns_max, ifm = jm / ns_max;
		face = ifp == ifm ? ifp % 4 + 4 : (ifp < ifm ? ifp % 4 : ifm % 4 + 8);
		ix = jm % ns_max;
		iy = ns_max - jp % ns_max - 1;
	} else {
		int ntt = int(floor(tt));
		if (ntt >= 4) ntt = 3;
		const double tp = tt - ntt;
		const double tmp = __dsqrt_rn(3.0 * (1.0 - za));
		const int jp = min(int(floor(ns_max * tp * tmp)), ns_max - 1);
		const int jm = min(int(floor(ns_max * (1.0 - tp) * tmp)), ns_max - 1);
		if (z >= 0) {
			face = ntt;
			ix = ns_max - jm - 1;
			iy = ns_max - jp - 1;
		} else {
			face = ntt + 8;
			ix = jp;
			iy = jm;
		}
	}
	const uint32_t ipf = spread_bits13(uint32_t(ix)) + 2u * spread_bits13(uint32_t(iy));
	const uint32_t r = uint32_t(ns_max / nside);
	return int32_t(ipf / (r * r) + uint32_t(face) * uint32_t(nside) * uint32_t(nside));
}

struct LcArgs {
	const uint32_t* x;
	const uint32_t* y;
	const uint32_t* z;
	const float* vx;
	const float* vy;
	const float* vz;
	const int8_t* rung;
	int64_t n;  // particles from the start of the active range
	int r;
	int nside;
	double c1, ainv, c0, tau0, tau1, tau_max;
	ct_lc_entry* out;
	int64_t cap;
	unsigned long long* count;
};

// Crossing of the light shell by periodic image img (offset -1 or 0 per axis,
// bits 2, 1, 0 for x, y, z) of a particle moving from p0 to p1 at velocity v
// (src/drift.cu:75-131).  Writes the entry when e is not null.
__device__ __forceinline__ bool lc_crossing(const LcArgs& A, const double (&p0)[3], const double (&p1)[3], const float (&v)[3], int img,
		ct_lc_entry* e) {
	const double o[3] = {-double((img >> 2) & 1), -double((img >> 1) & 1), -double(img & 1)};
	const double X0[3] = {p0[0] + o[0], p0[1] + o[1], p0[2] + o[2]};
	const double X1[3] = {p1[0] + o[0], p1[1] + o[1], p1[2] + o[2]};
	const double dist0 = __dsqrt_rn(X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2]);
	const double dist1 = __dsqrt_rn(X1[0] * X1[0] + X1[1] * X1[1] + X1[2] * X1[2]);
	const int i0 = int((A.tau0 + dist0) * A.c0);
	const int i1 = int((A.tau1 + dist1) * A.c0);
	if (i0 == i1 || !(dist1 <= 1.0 || dist0 <= 1.0)) return false;
	const float ux = float(double(v[0]) * A.ainv);
	const float uy = float(double(v[1]) * A.ainv);
	const float uz = float(double(v[2]) * A.ainv);
	const double ti = double(i0 + 1) * A.tau_max;
	const double qa = 1.0 - double(ux * ux + uy * uy + uz * uz);
	const double qb = 2.0 * ((A.tau0 - ti) - (double(ux) * X0[0] + double(uy) * X0[1] + double(uz) * X0[2]));
	const double qc = (ti - A.tau0) * (ti - A.tau0) - (X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2]);
	const double t = -(qb + __dsqrt_rn(qb * qb - 4.0 * qa * qc)) / (2.0 * qa);
	const double xc = X0[0] + double(ux) * t, yc = X0[1] + double(uy) * t, zc = X0[2] + double(uz) * t;
	if (!(xc * xc + yc * yc + zc * zc <= 1.0)) return false;
	if (e) {
		e->x = int32_t(2147483648.0 * xc);
		e->y = int32_t(2147483648.0 * yc);
		e->z = int32_t(2147483648.0 * zc);
		e->vx = v[0];
		e->vy = v[1];
		e->vz = v[2];
		e->pix = vec2pix_nest(A.nside, xc, yc, zc);
	}
	return true;
}

// Lightcone crossings of the drift, a particle per thread, read-only (so that
// a run whose entries overflowed `cap` can be repeated before the drift moves
// the particles): the drifted position is the drift's own x0 + v c1 before
// the wrap.  Pass 1 counts each thread's crossings, a block scan turns the
// counts into offsets reserved with one atomic per block, and pass 2
// recomputes the crossings and writes them.
__global__ __launch_bounds__(256) void k_lightcone(LcArgs A) {
	__shared__ unsigned long long wbase[4];
	__shared__ int wtotal[4];
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	const int lane = lane_id(), wave = threadIdx.x >> 6;
	const bool act = i < A.n && A.rung[i] == A.r;
	double p0[3] = {0.0, 0.0, 0.0}, p1[3] = {0.0, 0.0, 0.0};
	float v[3] = {0.f, 0.f, 0.f};
	if (act) {
		const uint32_t X[3] = {A.x[i], A.y[i], A.z[i]};
		v[0] = A.vx[i];
		v[1] = A.vy[i];
		v[2] = A.vz[i];
#pragma unroll
		for (int d = 0; d < 3; d++) {
			p0[d] = double(X[d]) * (1.0 / 4294967296.0);
			p1[d] = __builtin_fma(double(v[d]), A.c1, p0[d]);
		}
	}
	int nh = 0;
	if (act) {
		for (int img = 0; img < 8; img++) nh += lc_crossing(A, p0, p1, v, img, nullptr) ? 1 : 0;
	}
	int incl = nh;
	for (int o = 1; o < 64; o <<= 1) {
		const int t = __shfl_up(incl, o);
		if (lane >= o) incl += t;
	}
	if (lane == 63) wtotal[wave] = incl;
	__syncthreads();
	if (threadIdx.x == 0) {
		unsigned long long tot = 0;
		for (int w = 0; w < 4; w++) tot += (unsigned long long) wtotal[w];
		unsigned long long base = tot ? atomicAdd(A.count, tot) : 0ull;
		for (int w = 0; w < 4; w++) {
			wbase[w] = base;
			base += (unsigned long long) wtotal[w];
		}
	}
	__syncthreads();
	if (nh) {
		unsigned long long slot = wbase[wave] + (unsigned long long) (incl - nh);
		for (int img = 0; img < 8; img++) {
			ct_lc_entry e;
			if (lc_crossing(A, p0, p1, v, img, &e)) {
				if (slot < (unsigned long long) A.cap) A.out[slot] = e;
				slot++;
			}
		}
	}
}

// The drift of every box of a batch in one launch: thread t drifts particle
// t - off[b] of box b (same arithmetic as k_drift, per-box c1 and drag).
// cnt (may be null): the kick's counters; a step whose walk overflowed its
// arenas (and is redone eagerly) does not drift.
__global__ __launch_bounds__(256) void k_drift_batch(BoxDesc* __restrict__ boxes, const int32_t* __restrict__ off, int nbox, int total,
		const Counters* __restrict__ cnt) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= total) return;
	if (cnt && (cnt->overflow | cnt->walk_error)) return;
	const int bx = find_box(off, nbox, t);
	const BoxDesc& B = boxes[bx];
	const int i = t - off[bx];
	if (i < B.drift_first || B.rung[i] != B.drift_rung) return;
	uint32_t* X[3] = {B.x, B.y, B.z};
	float* V[3] = {B.vx, B.vy, B.vz};
#pragma unroll
	for (int d = 0; d < 3; d++) {
		const double x0 = double(X[d][i]) * (1.0 / 4294967296.0);
		double x1 = __builtin_fma(double(V[d][i]), B.drift_c1, x0);
		x1 -= double(x1 >= 1.0);
		x1 += double(x1 < 0.0);
		X[d][i] = uint32_t(uint64_t(x1 * 4294967296.0));
	}
	if (B.do_drag) {
#pragma unroll
		for (int d = 0; d < 3; d++) V[d][i] = float(double(V[d][i]) * B.drag_c);
	}
}

// ---------------------------------------------------------------------------
// expansion kernels on the device, for checks
// ---------------------------------------------------------------------------

__global__ void k_fmm(int op, int64_t n, const double* __restrict__ in, double* __restrict__ out, int in_stride, int out_stride) {
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const double* a = in + i * in_stride;
	double* b = out + i * out_stride;
	float X[3], A_[50], B_[50], C_[50];
	switch (op) {
	case 0:
	case 1:
	case 5: {
		for (int q = 0; q < 3; q++) X[q] = float(a[q]);
		if (op == 0) ctfmm::greens<float>(C_, X[0], X[1], X[2]);
		if (op == 1) ctfmm::ewald_greens<float>(C_, X[0], X[1], X[2], c_rimg, c_nreal, c_fh, c_fc, c_nfour);
		if (op == 5) ctfmm::P2M<float>(C_, X[0], X[1], X[2]);
		const int no = op == 5 ? 37 : 50;
		for (int q = 0; q < no; q++) b[q] = C_[q];
		break;
	}
	case 2:
	case 7: {
		for (int q = 0; q < 37; q++) A_[q] = float(a[q]);
		for (int q = 0; q < 50; q++) B_[q] = float(a[37 + q]);
		for (int q = 0; q < 50; q++) C_[q] = 0.0f;
		if (op == 2) {
			ctfmm::M2L<float>(C_, A_, B_, true);
			for (int q = 0; q < 50; q++) b[q] = C_[q];
		} else {
			ctfmm::M2L2<float>(C_, A_, B_, true);
			for (int q = 0; q < 4; q++) b[q] = C_[q];
		}
		break;
	}
	case 3:
	case 4: {
		for (int q = 0; q < 50; q++) A_[q] = float(a[q]);
		for (int q = 0; q < 3; q++) X[q] = float(a[50 + q]);
		if (op == 3) {
			ctfmm::L2L<float>(C_, A_, X[0], X[1], X[2], true);
			for (int q = 0; q < 50; q++) b[q] = C_[q];
		} else {
			ctfmm::L2P<float>(C_, A_, X[0], X[1], X[2], true);
			for (int q = 0; q < 4; q++) b[q] = C_[q];
		}
		break;
	}
	case 6: {
		for (int q = 0; q < 37; q++) A_[q] = float(a[q]);
		for (int q = 0; q < 3; q++) X[q] = float(a[37 + q]);
		ctfmm::M2M<float>(C_, A_, X[0], X[1], X[2]);
		for (int q = 0; q < 37; q++) b[q] = C_[q];
		break;
	}
	default:
		break;
	}
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

thread_local std::string g_err;
std::atomic<uint64_t> g_tree_version{0};

struct DevState {
	int device = -1;
	bool ewald_ready = false;
	std::mutex mtx;
};
DevState g_dev;

int upload_ewald() {
	const EwaldHost E = make_ewald_tables();
	if (E.nreal > kNRealMax || E.nfour > kNFourMax) return fail(CT_E_ARG, "ewald tables too large");
	CT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_rimg), E.rimg.data(), E.rimg.size() * sizeof(float)));
	CT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_fh), E.fh.data(), E.fh.size() * sizeof(float)));
	CT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_fc), E.fc.data(), E.fc.size() * sizeof(float)));
	CT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_fc64), E.fc64.data(), E.fc64.size() * sizeof(double)));
	CT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_nreal), &E.nreal, sizeof(int)));
	CT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_nfour), &E.nfour, sizeof(int)));
	float rdt[16];
	for (int i = 0; i < 16; i++) rdt[i] = 1.0f / float(1 << i);
	CT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_rung_dt), rdt, sizeof(rdt)));
	g_dev.ewald_ready = true;
	return CT_OK;
}

}  // namespace ct

using namespace ct;

namespace {

int upload_box(ct_box* b) {
	const int64_t n = b->hp.n;
	int rc;
	if ((rc = b->x.grow(n)) || (rc = b->y.grow(n)) || (rc = b->z.grow(n)) || (rc = b->vx.grow(n)) || (rc = b->vy.grow(n)) ||
			(rc = b->vz.grow(n)) || (rc = b->rung.grow(n)) || (rc = b->gx.grow(n)) || (rc = b->gy.grow(n)) || (rc = b->gz.grow(n)) ||
			(rc = b->phi.grow(n)) || (rc = b->id.grow(n)))
		return rc;
	if (n) {
		CT_HIP(hipMemcpy(b->id.p, b->hp.id.data(), n * 4, hipMemcpyHostToDevice));
		CT_HIP(hipMemcpy(b->x.p, b->hp.x.data(), n * 4, hipMemcpyHostToDevice));
		CT_HIP(hipMemcpy(b->y.p, b->hp.y.data(), n * 4, hipMemcpyHostToDevice));
		CT_HIP(hipMemcpy(b->z.p, b->hp.z.data(), n * 4, hipMemcpyHostToDevice));
		CT_HIP(hipMemcpy(b->vx.p, b->hp.vx.data(), n * 4, hipMemcpyHostToDevice));
		CT_HIP(hipMemcpy(b->vy.p, b->hp.vy.data(), n * 4, hipMemcpyHostToDevice));
		CT_HIP(hipMemcpy(b->vz.p, b->hp.vz.data(), n * 4, hipMemcpyHostToDevice));
		CT_HIP(hipMemcpy(b->rung.p, b->hp.rung.data(), n, hipMemcpyHostToDevice));
		CT_HIP(hipMemset(b->gx.p, 0, n * 4));
		CT_HIP(hipMemset(b->gy.p, 0, n * 4));
		CT_HIP(hipMemset(b->gz.p, 0, n * 4));
		CT_HIP(hipMemset(b->phi.p, 0, n * 4));
	}
	return CT_OK;
}

int download_particles(ct_box* b) {
	const int64_t n = b->hp.n;
	if (!b->host_stale || n == 0) return CT_OK;
	CT_HIP(hipMemcpy(b->hp.x.data(), b->x.p, n * 4, hipMemcpyDeviceToHost));
	CT_HIP(hipMemcpy(b->hp.y.data(), b->y.p, n * 4, hipMemcpyDeviceToHost));
	CT_HIP(hipMemcpy(b->hp.z.data(), b->z.p, n * 4, hipMemcpyDeviceToHost));
	CT_HIP(hipMemcpy(b->hp.vx.data(), b->vx.p, n * 4, hipMemcpyDeviceToHost));
	CT_HIP(hipMemcpy(b->hp.vy.data(), b->vy.p, n * 4, hipMemcpyDeviceToHost));
	CT_HIP(hipMemcpy(b->hp.vz.data(), b->vz.p, n * 4, hipMemcpyDeviceToHost));
	CT_HIP(hipMemcpy(b->hp.rung.data(), b->rung.p, n, hipMemcpyDeviceToHost));
	CT_HIP(hipMemcpy(b->hp.id.data(), b->id.p, n * 4, hipMemcpyDeviceToHost));
	b->host_stale = false;
	return CT_OK;
}

// FP64 copies of the multipoles and the FP64 local expansions, made on the
// first FP64 kick after a tree build.
int ensure_fp64(ct_box* b) {
	const int64_t nn = b->nnodes;
	int rc;
	if ((rc = b->L64.grow(std::max<int64_t>(1, nn) * kLStride))) return rc;
	if (b->multi64_ready) return CT_OK;
	if (!b->host_nodes) return fail(CT_E_STATE, "FP64 multipoles missing");
	if ((rc = b->multi64.grow(std::max<int64_t>(1, nn) * kMultiStride))) return rc;
	std::vector<double> mp(nn * kMultiStride, 0.0);
	for (int64_t i = 0; i < nn; i++) {
		const HostNode& h = b->tree.nodes[i];
		for (int q = 0; q < 37; q++) mp[i * kMultiStride + q] = h.multid[q];
		for (int q = 0; q < 3; q++) mp[i * kMultiStride + 37 + q] = double(h.pos[q]);
	}
	if (nn) CT_HIP(hipMemcpy(b->multi64.p, mp.data(), mp.size() * sizeof(double), hipMemcpyHostToDevice));
	b->multi64_ready = true;
	return CT_OK;
}

// The walk arenas, the box table and the work lists of one kick context: a
// box kicked alone, or a batch of boxes kicked together.
struct KickCtx {
	std::vector<ct_box*> boxes;
	std::vector<uint64_t> versions;  // tree version of each box when prepared
	bool prepared = false;
	// box table and index spaces
	std::vector<BoxDesc> hdesc;
	DBuf<BoxDesc> desc;
	DBuf<int32_t> node_off, part_off;
	int32_t total_nodes = 0, total_parts = 0;
	int64_t total_sinks = 0;
	int max_depth = 0;
	DBuf<int32_t> lvl_pref, lvl_lo;  // per level d: [(nbox + 1)] prefix and [nbox] first node
	std::vector<int32_t> lvl_total;
	DBuf<int2> pp_leaves;  // (box, leaf) grouped by size class
	int64_t pp_class_off[kNumLeafClasses + 1] = {0};
	DBuf<int2> kick_leaves;  // (box, leaf) grouped by box
	int64_t nkick = 0;
	DBuf<int2> chunks;  // reduction chunks of kick_leaves, within one box each
	DBuf<int32_t> chunk_off;
	int nchunks = 0;
	DBuf<double> part, chunk_sums, box_sums;
	// walk arenas
	DBuf<Act> act[2];
	DBuf<int32_t> dp[2], ep[2];
	DBuf<uint8_t> dcat, ecat;  // check classes of the level being walked
	DBuf<int32_t> out[kNumLists];
	DBuf<Seg> seg;
	DBuf<int32_t> head;
	DBuf<Counters> cnt;
	int64_t cap_act = 0, cap_dp = 0, cap_ep = 0, cap_out[kNumLists] = {0, 0, 0, 0, 0}, cap_seg = 0;
	// captured kick (batches); captured on `own` when the caller passes the
	// default stream, which cannot be captured
	hipGraphExec_t gexec = nullptr;
	hipStream_t own = nullptr;
	int graph_prec = -1;
	bool graph_drift = false;  // the captured graph ends with the batch drift
	// diagnostics of the last kick
	hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
	double stats[15] = {0};
	Counters last_counters;
	bool have_lists = false;
	std::vector<double> host_sums;

	int nbox() const { return int(boxes.size()); }
	void drop_graph() {
		if (gexec) (void) hipGraphExecDestroy(gexec);
		gexec = nullptr;
		graph_prec = -1;
		graph_drift = false;
	}
	void release() {
		drop_graph();
		desc.release();
		node_off.release();
		part_off.release();
		lvl_pref.release();
		lvl_lo.release();
		pp_leaves.release();
		kick_leaves.release();
		chunks.release();
		chunk_off.release();
		part.release();
		chunk_sums.release();
		box_sums.release();
		for (int q = 0; q < 2; q++) {
			act[q].release();
			dp[q].release();
			ep[q].release();
		}
		dcat.release();
		ecat.release();
		for (int q = 0; q < kNumLists; q++) out[q].release();
		seg.release();
		head.release();
		cnt.release();
		for (int q = 0; q < 7; q++)
			if (ev[q]) (void) hipEventDestroy(ev[q]);
		if (own) (void) hipStreamDestroy(own);
		own = nullptr;
	}
};

int ctx_init(KickCtx* c, const std::vector<ct_box*>& boxes) {
	c->boxes = boxes;
	c->versions.assign(boxes.size(), ~uint64_t(0));
	for (int q = 0; q < 7; q++) CT_HIP(hipEventCreate(&c->ev[q]));
	CT_HIP(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
	return CT_OK;
}

// (Re)builds the index spaces and work lists after a box's tree changed.
int ctx_prepare(KickCtx* c) {
	const int nb = c->nbox();
	bool stale = !c->prepared;
	for (int b = 0; b < nb; b++) {
		if (!c->boxes[b]->have_tree) return fail(CT_E_STATE, "kick: a box has no tree (call ct_tree_build)");
		stale = stale || c->versions[b] != c->boxes[b]->tree_version;
	}
	if (!stale) return CT_OK;
	c->drop_graph();
	int rc;
	std::vector<int32_t> noff(nb + 1, 0), poff(nb + 1, 0);
	int maxd = 0;
	for (int b = 0; b < nb; b++) {
		noff[b + 1] = noff[b] + int32_t(c->boxes[b]->nnodes);
		poff[b + 1] = poff[b] + int32_t(c->boxes[b]->hp.n);
		maxd = std::max(maxd, c->boxes[b]->tree.max_depth);
	}
	c->total_nodes = noff[nb];
	c->total_parts = poff[nb];
	// the sinks of a kick: the particles of every box's active range, which
	// its tree's leaves hold
	c->total_sinks = 0;
	for (int b = 0; b < nb; b++) c->total_sinks += c->boxes[b]->hp.n - c->boxes[b]->rung_begin;
	c->max_depth = maxd;
	// levels 1..maxd of every box
	std::vector<int32_t> pref((maxd + 1) * (nb + 1), 0), lo((maxd + 1) * std::max(nb, 1), 0);
	c->lvl_total.assign(maxd + 1, 0);
	for (int d = 0; d <= maxd; d++) {
		for (int b = 0; b < nb; b++) {
			const HostTree& T = c->boxes[b]->tree;
			int32_t l0 = 0, l1 = 0;
			if (d <= T.max_depth) {
				l0 = T.level_off[d];
				l1 = T.level_off[d + 1];
			}
			lo[d * nb + b] = l0;
			pref[d * (nb + 1) + b + 1] = pref[d * (nb + 1) + b] + (l1 - l0);
		}
		c->lvl_total[d] = pref[d * (nb + 1) + nb];
	}
	// leaf lists
	std::vector<int2> ppl, kl;
	for (int cl = 0; cl < kNumLeafClasses; cl++) {
		c->pp_class_off[cl] = int64_t(ppl.size());
		for (int b = 0; b < nb; b++)
			for (int32_t l : c->boxes[b]->leaf_class[cl]) ppl.push_back(make_int2(b, l));
	}
	c->pp_class_off[kNumLeafClasses] = int64_t(ppl.size());
	std::vector<int2> ch;
	std::vector<int32_t> choff(nb + 1, 0);
	for (int b = 0; b < nb; b++) {
		const int64_t k0 = int64_t(kl.size());
		for (int cl = 0; cl < kNumLeafClasses; cl++)
			for (int32_t l : c->boxes[b]->leaf_class[cl]) kl.push_back(make_int2(b, l));
		const int64_t k1 = int64_t(kl.size());
		for (int64_t q = k0; q < k1; q += kReduceChunk) ch.push_back(make_int2(int(q), int(std::min<int64_t>(k1, q + kReduceChunk))));
		if (k1 == k0) ch.push_back(make_int2(int(k0), int(k0)));
		choff[b + 1] = int32_t(ch.size());
	}
	c->nkick = int64_t(kl.size());
	c->nchunks = int(ch.size());
	if ((rc = c->desc.grow(std::max(nb, 1))) || (rc = c->node_off.grow(nb + 1)) || (rc = c->part_off.grow(nb + 1)) ||
			(rc = c->lvl_pref.grow(int64_t(pref.size()))) || (rc = c->lvl_lo.grow(int64_t(lo.size()))) ||
			(rc = c->pp_leaves.grow(std::max<int64_t>(1, int64_t(ppl.size())))) || (rc = c->kick_leaves.grow(std::max<int64_t>(1, c->nkick))) ||
			(rc = c->chunks.grow(std::max(1, c->nchunks))) || (rc = c->chunk_off.grow(nb + 1)) ||
			(rc = c->part.grow(std::max<int64_t>(1, c->nkick) * 8)) || (rc = c->chunk_sums.grow(int64_t(std::max(1, c->nchunks)) * 8)) ||
			(rc = c->box_sums.grow(int64_t(std::max(1, nb)) * 8)) || (rc = c->head.grow(std::max<int64_t>(1, int64_t(kNumLists) * c->total_nodes))) ||
			(rc = c->cnt.grow(1)))
		return rc;
	CT_HIP(hipMemcpy(c->node_off.p, noff.data(), noff.size() * 4, hipMemcpyHostToDevice));
	CT_HIP(hipMemcpy(c->part_off.p, poff.data(), poff.size() * 4, hipMemcpyHostToDevice));
	CT_HIP(hipMemcpy(c->lvl_pref.p, pref.data(), pref.size() * 4, hipMemcpyHostToDevice));
	CT_HIP(hipMemcpy(c->lvl_lo.p, lo.data(), lo.size() * 4, hipMemcpyHostToDevice));
	if (!ppl.empty()) CT_HIP(hipMemcpy(c->pp_leaves.p, ppl.data(), ppl.size() * sizeof(int2), hipMemcpyHostToDevice));
	if (!kl.empty()) CT_HIP(hipMemcpy(c->kick_leaves.p, kl.data(), kl.size() * sizeof(int2), hipMemcpyHostToDevice));
	if (!ch.empty()) CT_HIP(hipMemcpy(c->chunks.p, ch.data(), ch.size() * sizeof(int2), hipMemcpyHostToDevice));
	CT_HIP(hipMemcpy(c->chunk_off.p, choff.data(), choff.size() * 4, hipMemcpyHostToDevice));
	// box table: everything but the per-kick fields
	c->hdesc.assign(nb, BoxDesc());
	for (int b = 0; b < nb; b++) {
		ct_box* x = c->boxes[b];
		BoxDesc& d = c->hdesc[b];
		std::memset(&d, 0, sizeof(d));
		d.nodes = x->nodes.p;
		d.x = x->x.p;
		d.y = x->y.p;
		d.z = x->z.p;
		d.vx = x->vx.p;
		d.vy = x->vy.p;
		d.vz = x->vz.p;
		d.rung = x->rung.p;
		d.gx = x->gx.p;
		d.gy = x->gy.p;
		d.gz = x->gz.p;
		d.phi = x->phi.p;
		d.head = c->head.p + int64_t(kNumLists) * noff[b];
		d.pp = x->pp.p;
		d.nnodes = int32_t(x->nnodes);
		d.nparts = int32_t(x->hp.n);
		c->versions[b] = x->tree_version;
	}
	// initial list capacities from the node count; grown on overflow
	const int64_t nn = std::max<int64_t>(c->total_nodes, 1);
	c->cap_act = std::max<int64_t>({c->cap_act, 4 * nn, 1024});
	c->cap_dp = std::max<int64_t>({c->cap_dp, 64 * nn, 1 << 16});
	c->cap_ep = std::max<int64_t>({c->cap_ep, 16 * nn, 1 << 14});
	c->cap_out[kEwald] = std::max<int64_t>({c->cap_out[kEwald], 8 * nn, 1 << 14});
	c->cap_out[kCC] = std::max<int64_t>({c->cap_out[kCC], 64 * nn, 1 << 16});
	c->cap_out[kCP] = std::max<int64_t>({c->cap_out[kCP], 4 * nn, 1 << 12});
	c->cap_out[kPC] = std::max<int64_t>({c->cap_out[kPC], 4 * nn, 1 << 12});
	c->cap_out[kPP] = std::max<int64_t>({c->cap_out[kPP], 32 * nn, 1 << 16});
	c->cap_seg = std::max<int64_t>({c->cap_seg, 8 * nn, 1 << 14});
	c->prepared = true;
	c->have_lists = false;
	return CT_OK;
}

int ctx_grow_arenas(KickCtx* c) {
	int rc;
	for (int q = 0; q < 2; q++) {
		if ((rc = c->act[q].grow(c->cap_act))) return rc;
		if ((rc = c->dp[q].grow(c->cap_dp))) return rc;
		if ((rc = c->ep[q].grow(c->cap_ep))) return rc;
	}
	if ((rc = c->dcat.grow(c->cap_dp)) || (rc = c->ecat.grow(c->cap_ep))) return rc;
	for (int q = 0; q < kNumLists; q++) {
		if ((rc = c->out[q].grow(c->cap_out[q]))) return rc;
	}
	if ((rc = c->seg.grow(c->cap_seg))) return rc;
	return CT_OK;
}

// Per-kick fields of the box table: parameters and the arithmetic's buffers.
int ctx_set_params(KickCtx* c, const ct_kick_params* p, int precision) {
	int rc;
	for (int b = 0; b < c->nbox(); b++) {
		ct_box* x = c->boxes[b];
		if (p[b].min_rung < 0 || p[b].min_rung >= 16 || !(p[b].theta > 0.0f) || !(p[b].h > 0.0f)) return fail(CT_E_ARG, "kick: bad params");
		if (precision == CT_PREC_F64 && (rc = ensure_fp64(x))) return rc;
		BoxDesc& d = c->hdesc[b];
		d.multi = precision == CT_PREC_F64 ? static_cast<const void*>(x->multi64.p) : static_cast<const void*>(x->multi.p);
		d.L = precision == CT_PREC_F64 ? static_cast<void*>(x->L64.p) : static_cast<void*>(x->L.p);
		d.theta_inv = 1.0f / p[b].theta;
		d.h = p[b].h;
		d.prm = p[b];
	}
	return CT_OK;
}

WalkArgs walk_args(KickCtx* c) {
	WalkArgs A;
	A.boxes = c->desc.p;
	for (int q = 0; q < 2; q++) {
		A.act[q] = c->act[q].p;
		A.dp[q] = c->dp[q].p;
		A.ep[q] = c->ep[q].p;
	}
	A.dcat = c->dcat.p;
	A.ecat = c->ecat.p;
	A.cap_act = c->cap_act;
	A.cap_dp = c->cap_dp;
	A.cap_ep = c->cap_ep;
	for (int q = 0; q < kNumLists; q++) {
		A.out[q] = c->out[q].p;
		A.cap_out[q] = c->cap_out[q];
	}
	A.seg = c->seg.p;
	A.cap_seg = c->cap_seg;
	A.cnt = c->cnt.p;
	return A;
}

// Workgroups of `kernel` (`block` threads, static LDS only) resident at once
// on the library's device (ct_init): every CU at the occupancy that the
// kernel's resources allow, or `fallback` if the runtime cannot say.
template<class K>
int resident_blocks(K kernel, int block, int fallback) {
	int dev = 0, ncu = 0, per_cu = 0;
	if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
			hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) == hipSuccess && ncu > 0 && per_cu > 0)
		return ncu * per_cu;
	return fallback;
}

// The walk's grid: a larger one would leave a second round of waves behind
// the first in its grid-stride loop.
int walk_grid() {
	static const int g = resident_blocks(k_walk, 256, 256 * kWalkWaves);
	return g;
}

// Grids of the Ewald and M2L cell kernels: 8 workgroups per resident slot,
// each looping over nodes.  A workgroup per node spends time being dispatched
// on the many nodes with short lists; one per slot leaves the slowest slot's
// nodes as a tail.
constexpr int kCellGridPerSlot = 8;
template<class T>
int cell_grid_ewald() {
	static const int g = kCellGridPerSlot * resident_blocks(k_cells_ewald<T>, 64, 4096);
	return g;
}
template<class T, bool CP>
int cell_grid_direct() {
	static const int g = kCellGridPerSlot * resident_blocks(k_cells_direct<T, CP>, 64, 4096);
	return g;
}

int walk_iters(KickCtx* c) {
	return c->max_depth + 3;
}

// The level launches of one walk (no host synchronisation).
void launch_walk(KickCtx* c, hipStream_t st) {
	InitArgs ia;
	ia.head = c->head.p;
	ia.nhead = int64_t(kNumLists) * c->total_nodes;
	ia.cnt = c->cnt.p;
	ia.act0 = c->act[0].p;
	ia.dp0 = c->dp[0].p;
	ia.ep0 = c->ep[0].p;
	ia.nbox = c->nbox();
	hipLaunchKernelGGL(k_walk_init, dim3(256), dim3(256), 0, st, ia);
	hipLaunchKernelGGL(k_walk_seed_counts, dim3(1), dim3(64), 0, st, c->cnt.p, c->nbox());
	const WalkArgs A = walk_args(c);
	const int iters = walk_iters(c);
	const int grid = walk_grid();
	for (int k = 0; k < iters; k++) {
		hipLaunchKernelGGL(k_walk, dim3(grid), dim3(256), 0, st, A, k);
	}
}

// Walk with overflow recovery: arenas that overflowed are grown from the
// counters and the walk is run again.
int run_walk(KickCtx* c, hipStream_t st) {
	if (walk_iters(c) > kMaxIter) return fail(CT_E_TREE, "tree too deep for the walk");
	for (int attempt = 0; attempt < 12; attempt++) {
		int rc = ctx_grow_arenas(c);
		if (rc) return rc;
		launch_walk(c, st);
		Counters C;
		CT_HIP(hipMemcpyAsync(&C, c->cnt.p, sizeof(Counters), hipMemcpyDeviceToHost, st));
		CT_HIP(hipStreamSynchronize(st));
		const int iters = walk_iters(c);
		if (C.walk_error) return fail(CT_E_WALK, "walk: a leaf failed the Ewald acceptance test");
		if (C.act[iters] != 0 && !C.overflow) return fail(CT_E_WALK, "walk did not terminate within the tree depth");
		if (!C.overflow) return CT_OK;
		c->drop_graph();
		auto bump = [](int64_t& cap, int64_t need) { cap = std::max<int64_t>(2 * cap, need + need / 4 + 1024); };
		int maxa = 0, maxd = 0, maxe = 0;
		for (int k = 0; k <= iters; k++) {
			maxa = std::max(maxa, C.act[k]);
			maxd = std::max(maxd, C.dp[k]);
			maxe = std::max(maxe, C.ep[k]);
		}
		if (C.overflow & (1 << 5)) bump(c->cap_act, maxa);
		if (C.overflow & (1 << 6)) bump(c->cap_dp, maxd);
		if (C.overflow & (1 << 7)) bump(c->cap_ep, maxe);
		for (int q = 0; q < kNumLists; q++)
			if (C.overflow & (1 << q)) bump(c->cap_out[q], C.out[q].v);
		if (C.overflow & (1 << 8)) bump(c->cap_seg, C.seg.v);
	}
	return fail(CT_E_ARENA, "walk: list buffers kept overflowing");
}

// The interaction phases of one kick in arithmetic mode PR: cells (Ewald,
// M2L, P2L), the L2L sweep, P2P per leaf class, M2P + L2P + kick, and the
// per-box reduction of the kick_return sums.
template<int PR>
void launch_interactions(KickCtx* c, hipStream_t st, bool events) {
	using T = typename Prec<PR>::T;
	const int nb = c->nbox();
	CellArgs ca;
	ca.boxes = c->desc.p;
	ca.node_off = c->node_off.p;
	ca.nbox = nb;
	ca.total = c->total_nodes;
	ca.seg = c->seg.p;
	ca.out_e = c->out[kEwald].p;
	ca.out_cc = c->out[kCC].p;
	ca.out_cp = c->out[kCP].p;
	ca.cnt = c->cnt.p;
	const int nn = std::max(1, c->total_nodes);
	hipLaunchKernelGGL((k_cells_ewald<T>), dim3(std::min(nn, cell_grid_ewald<T>())), dim3(64), 0, st, ca);
	hipLaunchKernelGGL((k_cells_direct<T, false>), dim3(std::min(nn, cell_grid_direct<T, false>())), dim3(64), 0, st, ca);
	// P2L: a workgroup per node (its nodes' lists are long enough that the
	// dispatcher's balancing beats the saved dispatches: 8.0 vs 8.4 ms on the
	// 256^3 clustered box)
	hipLaunchKernelGGL((k_cells_direct<T, true>), dim3(std::min(nn, 1 << 24)), dim3(64), 0, st, ca);
	if (events) (void) hipEventRecord(c->ev[2], st);
	for (int d = 1; d <= c->max_depth; d++) {
		const int tot = c->lvl_total[d];
		if (tot == 0) continue;
		L2LArgs la;
		la.boxes = c->desc.p;
		la.pref = c->lvl_pref.p + d * (nb + 1);
		la.lo = c->lvl_lo.p + d * nb;
		la.nbox = nb;
		la.total = tot;
		hipLaunchKernelGGL((k_l2l<T>), dim3((tot + 63) / 64), dim3(64), 0, st, la);
	}
	if (events) (void) hipEventRecord(c->ev[3], st);
	PPArgs pa;
	pa.boxes = c->desc.p;
	pa.seg = c->seg.p;
	pa.out_pp = c->out[kPP].p;
	pa.cnt = c->cnt.p;
	for (int cl = 0; cl < kNumLeafClasses; cl++) {
		const int64_t lo = c->pp_class_off[cl], cnt = c->pp_class_off[cl + 1] - lo;
		if (cnt == 0) continue;
		pa.leaves = c->pp_leaves.p + lo;
		switch (cl) {
#define CT_LAUNCH_PP(C, G, P) \
	case C: hipLaunchKernelGGL((k_pp<G, P, PR>), dim3(cnt), dim3(64), 0, st, pa); break;
			CT_LEAF_CLASSES(CT_LAUNCH_PP)
#undef CT_LAUNCH_PP
		}
	}
	if (events) (void) hipEventRecord(c->ev[6], st);
	LeafArgs lk;
	lk.boxes = c->desc.p;
	lk.leaves = c->kick_leaves.p;
	lk.seg = c->seg.p;
	lk.out_pc = c->out[kPC].p;
	lk.part = c->part.p;
	lk.cnt = c->cnt.p;
	if (c->nkick) hipLaunchKernelGGL((k_kick<PR>), dim3(c->nkick), dim3(64), 0, st, lk);
	if (events) (void) hipEventRecord(c->ev[4], st);
	hipLaunchKernelGGL(k_reduce_chunks, dim3(c->nchunks), dim3(256), 0, st, c->part.p, c->chunks.p, c->chunk_sums.p);
	hipLaunchKernelGGL(k_reduce_boxes, dim3(nb), dim3(256), 0, st, c->chunk_sums.p, c->chunk_off.p, c->box_sums.p);
}

void launch_interactions_prec(KickCtx* c, hipStream_t st, int precision, bool events) {
	switch (precision) {
	case CT_PREC_F32U:
		launch_interactions<kF32U>(c, st, events);
		break;
	case CT_PREC_F64:
		launch_interactions<kF64>(c, st, events);
		break;
	default:
		launch_interactions<kF32C>(c, st, events);
		break;
	}
}

// Results of the last kick: returns, counters and phase times.
int ctx_finish(KickCtx* c, hipStream_t st, ct_kick_return* rets, bool events) {
	const int nb = c->nbox();
	c->host_sums.assign(int64_t(std::max(nb, 1)) * 8, 0.0);
	Counters C;
	CT_HIP(hipMemcpyAsync(c->host_sums.data(), c->box_sums.p, int64_t(nb) * 8 * sizeof(double), hipMemcpyDeviceToHost, st));
	CT_HIP(hipMemcpyAsync(&C, c->cnt.p, sizeof(Counters), hipMemcpyDeviceToHost, st));
	CT_HIP(hipEventRecord(c->ev[5], st));
	CT_HIP(hipStreamSynchronize(st));
	CT_HIP(hipGetLastError());
	if (C.walk_error) return fail(CT_E_WALK, "walk: a leaf failed the Ewald acceptance test");
	for (ct_box* x : c->boxes) x->host_stale = true;
	c->have_lists = true;
	c->last_counters = C;
	for (int b = 0; b < nb && rets; b++) {
		const double* r = c->host_sums.data() + int64_t(b) * 8;
		rets[b].max_rung = int(r[0]);
		rets[b].pot = r[1];
		rets[b].kin = r[2];
		rets[b].dkin = r[3];
		rets[b].xmom = r[4];
		rets[b].ymom = r[5];
		rets[b].zmom = r[6];
		rets[b].nmom = r[7];
	}
	float ms = 0.0f;
	if (events) {
		for (int q = 0; q < 4; q++) {
			CT_HIP(hipEventElapsedTime(&ms, c->ev[q], c->ev[q + 1]));
			c->stats[q] = ms;
		}
		CT_HIP(hipEventElapsedTime(&ms, c->ev[3], c->ev[6]));
		c->stats[12] = ms;
		CT_HIP(hipEventElapsedTime(&ms, c->ev[6], c->ev[4]));
		c->stats[13] = ms;
	} else {
		for (int q = 0; q < 4; q++) c->stats[q] = 0.0;
		c->stats[12] = c->stats[13] = 0.0;
	}
	CT_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[5]));
	c->stats[4] = ms;
	unsigned long long checks = 0, pcs = 0, pps = 0, cps = 0;
	for (int q = 0; q < kStatSlots; q++) {
		checks += C.walk_checks[q].v;
		pcs += C.walk_pc[q].v;
		pps += C.pp_pairs[q].v;
		cps += C.cp_parts[q].v;
	}
	c->stats[5] = double(pps);
	c->stats[6] = double(C.stat[1]);
	c->stats[7] = double(C.stat[2]);
	c->stats[8] = double(cps);
	c->stats[9] = double(pcs);
	c->stats[10] = double(c->total_nodes);
	c->stats[11] = double(c->total_sinks);
	c->stats[14] = double(checks);
	return CT_OK;
}

// One kick of every box of the context.  Eager: the walk, a check of its
// counters (arenas grown and the walk repeated on overflow), then the
// interactions.  Graph: the whole kick replayed from a captured graph; if the
// walk overflowed in the replay, the kick is repeated eagerly and the graph
// captured again on the next call.
void launch_drift_batch(KickCtx* c, hipStream_t st, bool guarded) {
	if (c->total_parts > 0)
		hipLaunchKernelGGL(k_drift_batch, dim3((c->total_parts + 255) / 256), dim3(256), 0, st, c->desc.p, c->part_off.p, c->nbox(),
				c->total_parts, guarded ? c->cnt.p : nullptr);
}

// One kick of every box of the context; with drift, the batch drift of the
// box table's drift fields follows the kick in the same stream (and in the
// captured graph).
int ctx_kick(KickCtx* c, const ct_kick_params* p, int precision, ct_kick_return* rets, bool graph, hipStream_t st, bool drift = false) {
	int rc;
	if (precision != CT_PREC_F32C && precision != CT_PREC_F32U && precision != CT_PREC_F64) return fail(CT_E_ARG, "kick: bad precision");
	if ((rc = ctx_prepare(c))) return rc;
	if ((rc = ctx_set_params(c, p, precision))) return rc;
	if (walk_iters(c) > kMaxIter) return fail(CT_E_TREE, "tree too deep for the walk");
	if (graph && st == nullptr) st = c->own;
	CT_HIP(hipMemcpyAsync(c->desc.p, c->hdesc.data(), c->hdesc.size() * sizeof(BoxDesc), hipMemcpyHostToDevice, st));
	if (graph && c->gexec && (c->graph_prec != precision || c->graph_drift != drift)) c->drop_graph();
	if (graph && c->gexec) {
		CT_HIP(hipEventRecord(c->ev[0], st));
		CT_HIP(hipGraphLaunch(c->gexec, st));
		if ((rc = ctx_finish(c, st, rets, false))) return rc;
		if (!c->last_counters.overflow) return CT_OK;
		c->drop_graph();  // overflowed (the drift in the graph did not run): an eager step
	}
	CT_HIP(hipEventRecord(c->ev[0], st));
	if ((rc = run_walk(c, st))) return rc;
	CT_HIP(hipEventRecord(c->ev[1], st));
	launch_interactions_prec(c, st, precision, true);
	if (drift) launch_drift_batch(c, st, false);
	if ((rc = ctx_finish(c, st, rets, true))) return rc;
	if (graph && !c->gexec) {
		// capture walk + interactions (+ drift) for the next calls (arenas now fit)
		hipGraph_t g = nullptr;
		CT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
		launch_walk(c, st);
		launch_interactions_prec(c, st, precision, false);
		if (drift) launch_drift_batch(c, st, true);
		hipError_t e = hipStreamEndCapture(st, &g);
		if (e != hipSuccess) return fail(CT_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
		e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
		(void) hipGraphDestroy(g);
		if (e != hipSuccess) {
			c->gexec = nullptr;
			return fail(CT_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
		}
		c->graph_prec = precision;
		c->graph_drift = drift;
	}
	return CT_OK;
}

KickCtx* box_ctx(ct_box* b) {
	if (!b->self_ctx) {
		KickCtx* c = new (std::nothrow) KickCtx();
		if (!c) return nullptr;
		if (ctx_init(c, {b}) != CT_OK) {
			c->release();
			delete c;
			return nullptr;
		}
		b->self_ctx = c;
	}
	return static_cast<KickCtx*>(b->self_ctx);
}

}  // namespace

struct ct_batch {
	KickCtx ctx;
};

int ct::check_init() {
	if (g_dev.device < 0) return fail(CT_E_STATE, "ct_init has not been called");
	return CT_OK;
}

extern "C" {

int ct_init(int device) {
	std::lock_guard<std::mutex> lk(g_dev.mtx);
	int count = 0;
	CT_HIP(hipGetDeviceCount(&count));
	if (device < 0 || device >= count) return fail(CT_E_ARG, "ct_init: no such device");
	CT_HIP(hipSetDevice(device));
	g_dev.device = device;
	return upload_ewald();
}

int ct_ewald_init(void) {
	int rc = check_init();
	if (rc) return rc;
	return upload_ewald();
}

int ct_box_create(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, const float* vel, const int8_t* rung, ct_box_t* out) {
	int rc = check_init();
	if (rc) return rc;
	if (!out || n < 0 || n > INT32_MAX || (n > 0 && (!x || !y || !z))) return fail(CT_E_ARG, "ct_box_create: bad arguments");
	ct_box* b = new (std::nothrow) ct_box();
	if (!b) return fail(CT_E_OOM, "ct_box_create: host allocation failed");
	HostParticles& P = b->hp;
	P.n = n;
	P.x.assign(x, x + n);
	P.y.assign(y, y + n);
	P.z.assign(z, z + n);
	P.vx.resize(n);
	P.vy.resize(n);
	P.vz.resize(n);
	P.rung.resize(n);
	P.id.resize(n);
	for (int64_t i = 0; i < n; i++) {
		P.vx[i] = vel ? vel[3 * i + 0] : 0.0f;
		P.vy[i] = vel ? vel[3 * i + 1] : 0.0f;
		P.vz[i] = vel ? vel[3 * i + 2] : 0.0f;
		P.rung[i] = rung ? rung[i] : 0;
		P.id[i] = int32_t(i);
	}
	if ((rc = upload_box(b))) {
		ct_box_destroy(b);
		return rc;
	}
	*out = b;
	return CT_OK;
}

int ct_box_destroy(ct_box_t b) {
	if (!b) return CT_OK;
	(void) hipDeviceSynchronize();
	if (b->self_ctx) {
		KickCtx* c = static_cast<KickCtx*>(b->self_ctx);
		c->release();
		delete c;
		b->self_ctx = nullptr;
	}
	b->x.release();
	b->y.release();
	b->z.release();
	b->vx.release();
	b->vy.release();
	b->vz.release();
	b->gx.release();
	b->gy.release();
	b->gz.release();
	b->phi.release();
	b->rung.release();
	b->id.release();
	b->lc.release();
	b->lc_cnt.release();
	b->nodes.release();
	b->multi.release();
	b->L.release();
	b->multi64.release();
	b->L64.release();
	b->pp.release();
	tree_scratch_free(b);
	delete b;
	return CT_OK;
}

int ct_tree_build(ct_box_t b, double theta, int bucket, double hsoft) {
	return ct_tree_build_ex(b, theta, bucket, hsoft, CT_TREE_AUTO);
}

int ct_tree_build_ex(ct_box_t b, double theta, int bucket, double hsoft, int where) {
	int rc = check_init();
	if (rc) return rc;
	if (!b || !(theta > 0.0) || bucket < 1 || bucket > kMaxLeafSinks || hsoft < 0.0 ||
			(where != CT_TREE_AUTO && where != CT_TREE_HOST && where != CT_TREE_DEVICE))
		return fail(CT_E_ARG, "ct_tree_build: bad arguments");
	if (where != CT_TREE_HOST) {
		GpuTreeResult R;
		rc = gpu_tree_build(b, theta, bucket, hsoft, &R);
		if (rc == CT_OK) {
			std::vector<int32_t> cls[kNumLeafClasses];
			for (size_t q = 0; q < R.leaf_id.size(); q++) {
				const int np = R.leaf_pend[q] - R.leaf_pbegin[q];
				if (np == 0) continue;
				if (np > kMaxLeafSinks) return fail(CT_E_TREE, "ct_tree_build: leaf larger than 256 particles");
				cls[leaf_class(np)].push_back(R.leaf_id[q]);
			}
			b->tree.nodes.clear();
			b->tree.level_off = R.level_off;
			b->tree.max_depth = R.max_depth;
			b->tree.nleaves = int64_t(R.leaf_id.size());
			b->host_nodes = false;
			b->theta = theta;
			b->bucket = bucket;
			b->hsoft = hsoft;
			b->nnodes = R.nnodes;
			if ((rc = b->pp.grow(std::max<int64_t>(1, b->hp.n)))) return rc;
			for (int c = 0; c < kNumLeafClasses; c++) b->leaf_class[c] = std::move(cls[c]);
			b->have_tree = true;
			b->multi64_ready = true;  // the device build writes the FP64 multipoles too
			b->host_stale = true;     // particles were permuted on the device
			b->tree_version = ++g_tree_version;
			return CT_OK;
		}
		if (where == CT_TREE_DEVICE || rc != CT_E_TREE) return rc;
		// deeper than the 63-bit key: the host build takes it
	}
	if ((rc = download_particles(b))) return rc;
	TreeBuildParams prm;
	prm.theta = theta;
	prm.bucket = bucket;
	prm.hsoft = hsoft;
	prm.first = b->rung_begin;
	std::string err;
	HostTree T;
	if ((rc = build_tree(b->hp, prm, T, &err))) return fail(rc, err);
	const int64_t nn = int64_t(T.nodes.size());
	std::vector<DNode> dn(nn);
	std::vector<float> mp(nn * kMultiStride, 0.0f);
	std::vector<int32_t> cls[kNumLeafClasses];
	for (int64_t i = 0; i < nn; i++) {
		const HostNode& h = T.nodes[i];
		DNode& d = dn[i];
		for (int q = 0; q < 3; q++) {
			d.pos[q] = h.pos[q];
			d.bb[q] = h.box_begin[q];
			d.be[q] = h.box_end[q];
		}
		d.radius = h.radius;
		d.child[0] = h.child[0];
		d.child[1] = h.child[1];
		d.pbegin = h.pbegin;
		d.pend = h.pend;
		d.flags = (h.leaf ? 1 : 0) | (h.depth << 1);
		d.parent = h.parent;
		std::memcpy(&mp[i * kMultiStride], h.multi, 37 * sizeof(float));
		std::memcpy(&mp[i * kMultiStride + 37], h.pos, 3 * sizeof(uint32_t));
		if (h.leaf && h.pend > h.pbegin) {
			if (h.pend - h.pbegin > kMaxLeafSinks) return fail(CT_E_TREE, "ct_tree_build: leaf larger than 256 particles");
			cls[leaf_class(h.pend - h.pbegin)].push_back(int32_t(i));
		}
	}
	if ((rc = b->nodes.grow(nn)) || (rc = b->multi.grow(nn * kMultiStride)) || (rc = b->L.grow(nn * kLStride)) ||
			(rc = b->pp.grow(std::max<int64_t>(1, b->hp.n))))
		return rc;
	CT_HIP(hipMemcpy(b->nodes.p, dn.data(), nn * sizeof(DNode), hipMemcpyHostToDevice));
	CT_HIP(hipMemcpy(b->multi.p, mp.data(), mp.size() * sizeof(float), hipMemcpyHostToDevice));
	if ((rc = upload_box(b))) return rc;
	b->tree = std::move(T);
	b->host_nodes = true;
	b->theta = theta;
	b->bucket = bucket;
	b->hsoft = hsoft;
	b->nnodes = nn;
	for (int c = 0; c < kNumLeafClasses; c++) b->leaf_class[c] = std::move(cls[c]);
	b->have_tree = true;
	b->multi64_ready = false;
	b->tree_version = ++g_tree_version;
	return CT_OK;
}

int ct_kick(ct_box_t b, const ct_kick_params* p, ct_kick_return* ret, void* stream) {
	return ct_kick_prec(b, p, CT_PREC_F32C, ret, stream);
}

int ct_kick_prec(ct_box_t b, const ct_kick_params* p, int precision, ct_kick_return* ret, void* stream) {
	int rc = check_init();
	if (rc) return rc;
	if (!b || !p) return fail(CT_E_ARG, "ct_kick: bad arguments");
	if (!b->have_tree) return fail(CT_E_STATE, "ct_kick: no tree (call ct_tree_build)");
	KickCtx* c = box_ctx(b);
	if (!c) return fail(CT_E_OOM, "ct_kick: host allocation failed");
	return ctx_kick(c, p, precision, ret, false, reinterpret_cast<hipStream_t>(stream));
}

int ct_batch_create(const ct_box_t* boxes, int nbox, ct_batch_t* out) {
	int rc = check_init();
	if (rc) return rc;
	if (!out || nbox < 1 || !boxes) return fail(CT_E_ARG, "ct_batch_create: bad arguments");
	for (int i = 0; i < nbox; i++) {
		if (!boxes[i]) return fail(CT_E_ARG, "ct_batch_create: null box");
		for (int j = 0; j < i; j++)
			if (boxes[j] == boxes[i]) return fail(CT_E_ARG, "ct_batch_create: a box appears twice");
	}
	ct_batch* bt = new (std::nothrow) ct_batch();
	if (!bt) return fail(CT_E_OOM, "ct_batch_create: host allocation failed");
	if ((rc = ctx_init(&bt->ctx, std::vector<ct_box*>(boxes, boxes + nbox)))) {
		bt->ctx.release();
		delete bt;
		return rc;
	}
	*out = bt;
	return CT_OK;
}

int ct_batch_reset(ct_batch_t bt, const ct_box_t* boxes, int nbox) {
	int rc = check_init();
	if (rc) return rc;
	if (!bt || nbox < 1 || !boxes) return fail(CT_E_ARG, "ct_batch_reset: bad arguments");
	for (int i = 0; i < nbox; i++) {
		if (!boxes[i]) return fail(CT_E_ARG, "ct_batch_reset: null box");
		for (int j = 0; j < i; j++)
			if (boxes[j] == boxes[i]) return fail(CT_E_ARG, "ct_batch_reset: a box appears twice");
	}
	CT_HIP(hipStreamSynchronize(bt->ctx.own));
	KickCtx* c = &bt->ctx;
	c->boxes.assign(boxes, boxes + nbox);
	c->versions.assign(size_t(nbox), ~uint64_t(0));
	c->prepared = false;
	c->drop_graph();
	return CT_OK;
}

int ct_batch_destroy(ct_batch_t bt) {
	if (!bt) return CT_OK;
	(void) hipDeviceSynchronize();
	bt->ctx.release();
	delete bt;
	return CT_OK;
}

int ct_batch_kick(ct_batch_t bt, const ct_kick_params* params, int precision, ct_kick_return* rets, int use_graph, void* stream) {
	int rc = check_init();
	if (rc) return rc;
	if (!bt || !params) return fail(CT_E_ARG, "ct_batch_kick: bad arguments");
	hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : bt->ctx.own;
	return ctx_kick(&bt->ctx, params, precision, rets, use_graph != 0, st);
}

// The drift fields of the box table (cuda_drift's c1 = dt / a and drag
// factor, src/drift.cu:54-73, per box).
static int set_drift_fields(KickCtx* c, int rung, const float* a, const float* dt, const double* drag) {
	for (int b = 0; b < c->nbox(); b++) {
		if (!(a[b] > 0.0f)) return fail(CT_E_ARG, "batch drift: bad scale factor");
		BoxDesc& d = c->hdesc[b];
		const double dg = drag ? drag[b] : 0.0;
		d.drift_c1 = double(dt[b]) / double(a[b]);
		d.drag_c = 1.0 / (1.0 + dg * double(dt[b]));
		d.do_drag = dg != 0.0;
		d.drift_rung = rung;
		d.drift_first = int32_t(c->boxes[b]->rung_begin);
	}
	return CT_OK;
}

int ct_batch_step(ct_batch_t bt, const ct_kick_params* params, int precision, ct_kick_return* rets, int rung, const float* a, const float* dt,
		const double* drag, int use_graph, void* stream) {
	int rc = check_init();
	if (rc) return rc;
	if (!bt || !params || !a || !dt) return fail(CT_E_ARG, "ct_batch_step: bad arguments");
	KickCtx* c = &bt->ctx;
	if ((rc = ctx_prepare(c))) return rc;
	if ((rc = set_drift_fields(c, rung, a, dt, drag))) return rc;
	hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : c->own;
	return ctx_kick(c, params, precision, rets, use_graph != 0, st, true);
}

int ct_batch_drift(ct_batch_t bt, int rung, const float* a, const float* dt, const double* drag, void* stream) {
	int rc = check_init();
	if (rc) return rc;
	if (!bt || !a || !dt) return fail(CT_E_ARG, "ct_batch_drift: bad arguments");
	KickCtx* c = &bt->ctx;
	if ((rc = ctx_prepare(c))) return rc;
	if ((rc = set_drift_fields(c, rung, a, dt, drag))) return rc;
	hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : c->own;
	CT_HIP(hipMemcpyAsync(c->desc.p, c->hdesc.data(), c->hdesc.size() * sizeof(BoxDesc), hipMemcpyHostToDevice, st));
	launch_drift_batch(c, st, false);
	CT_HIP(hipGetLastError());
	CT_HIP(hipStreamSynchronize(st));
	for (ct_box* x : c->boxes) x->host_stale = true;
	return CT_OK;
}

int ct_batch_stats(ct_batch_t bt, double* out, int n) {
	if (!bt || !out) return fail(CT_E_ARG, "ct_batch_stats: bad arguments");
	for (int q = 0; q < n && q < 15; q++) out[q] = bt->ctx.stats[q];
	return CT_OK;
}

int ct_drift_lightcone(ct_box_t b, int rung, float a, float dt, float tau0, float tau1, float tau_max, int nside, double drag,
		int64_t* nentries, void* stream) {
	int rc = check_init();
	if (rc) return rc;
	if (nentries) *nentries = 0;
	if (!b || !(a > 0.0f) || !(tau_max > 0.0f) || nside < 1 || nside > 8192 || (nside & (nside - 1)))
		return fail(CT_E_ARG, "ct_drift_lightcone: bad arguments");
	hipStream_t st = reinterpret_cast<hipStream_t>(stream);
	const int64_t f = b->rung_begin;
	const int64_t n = b->hp.n - f;
	b->lc_count = 0;
	// the lightcone is on only in the last unit of conformal time
	// (src/drift.cu:39), with the float arguments widened to double
	if (n > 0 && double(tau_max) - double(tau1) < 1.0) {
		if ((rc = b->lc_cnt.grow(1)) || (rc = b->lc.grow(std::max<int64_t>(b->lc.n, 4096)))) return rc;
		LcArgs A;
		A.x = b->x.p + f;
		A.y = b->y.p + f;
		A.z = b->z.p + f;
		A.vx = b->vx.p + f;
		A.vy = b->vy.p + f;
		A.vz = b->vz.p + f;
		A.rung = b->rung.p + f;
		A.n = n;
		A.r = rung;
		A.nside = nside;
		A.c1 = double(dt) / double(a);
		A.ainv = 1.0 / double(a);
		A.c0 = 1.0 / double(tau_max);
		A.tau0 = tau0;
		A.tau1 = tau1;
		A.tau_max = tau_max;
		A.count = b->lc_cnt.p;
		for (int attempt = 0; attempt < 2; attempt++) {
			A.out = b->lc.p;
			A.cap = b->lc.n;
			CT_HIP(hipMemsetAsync(b->lc_cnt.p, 0, sizeof(unsigned long long), st));
			hipLaunchKernelGGL(k_lightcone, dim3((n + 255) / 256), dim3(256), 0, st, A);
			CT_HIP(hipGetLastError());
			unsigned long long cnt = 0;
			CT_HIP(hipMemcpyAsync(&cnt, b->lc_cnt.p, sizeof(cnt), hipMemcpyDeviceToHost, st));
			CT_HIP(hipStreamSynchronize(st));
			b->lc_count = int64_t(cnt);
			if (b->lc_count <= b->lc.n) break;
			// more crossings than room: grow and find them again (the pass
			// reads the particles only)
			if ((rc = b->lc.grow(b->lc_count))) return rc;
		}
	}
	if (nentries) *nentries = b->lc_count;
	return ct_drift(b, rung, a, dt, drag, stream);
}

int ct_lc_read(ct_box_t b, ct_lc_entry* out, int64_t cap, int64_t* count) {
	int rc = check_init();
	if (rc) return rc;
	if (!b || cap < 0 || (cap > 0 && !out)) return fail(CT_E_ARG, "ct_lc_read: bad arguments");
	if (count) *count = b->lc_count;
	const int64_t m = std::min(cap, b->lc_count);
	if (m > 0) CT_HIP(hipMemcpy(out, b->lc.p, m * sizeof(ct_lc_entry), hipMemcpyDeviceToHost));
	return CT_OK;
}

int ct_drift(ct_box_t b, int rung, float a, float dt, double drag, void* stream) {
	int rc = check_init();
	if (rc) return rc;
	if (!b || !(a > 0.0f)) return fail(CT_E_ARG, "ct_drift: bad arguments");
	hipStream_t st = reinterpret_cast<hipStream_t>(stream);
	// the active range [rung_begin, n) (cuda_drift, src/drift.cu:170-177)
	const int64_t f = b->rung_begin;
	const int64_t n = b->hp.n - f;
	if (n <= 0) return CT_OK;
	const double c1 = double(dt) / double(a);
	const double dragc = 1.0 / (1.0 + drag * double(dt));
	hipLaunchKernelGGL(k_drift, dim3((n + 255) / 256), dim3(256), 0, st, b->x.p + f, b->y.p + f, b->z.p + f, b->vx.p + f, b->vy.p + f,
			b->vz.p + f, b->rung.p + f, n, rung, c1, dragc, drag != 0.0);
	CT_HIP(hipGetLastError());
	CT_HIP(hipStreamSynchronize(st));
	b->host_stale = true;
	return CT_OK;
}

int ct_box_size(ct_box_t b, int64_t* np, int64_t* nn) {
	if (!b) return fail(CT_E_ARG, "ct_box_size: null box");
	if (np) *np = b->hp.n;
	if (nn) *nn = b->nnodes;
	return CT_OK;
}

int ct_box_read(ct_box_t b, uint32_t* x, uint32_t* y, uint32_t* z, float* vel, int8_t* rung, float* gx, float* gy, float* gz, float* phi, int32_t* id) {
	if (!b) return fail(CT_E_ARG, "ct_box_read: null box");
	const int64_t n = b->hp.n;
	if (n == 0) return CT_OK;
	CT_HIP(hipDeviceSynchronize());
	if (x) CT_HIP(hipMemcpy(x, b->x.p, n * 4, hipMemcpyDeviceToHost));
	if (y) CT_HIP(hipMemcpy(y, b->y.p, n * 4, hipMemcpyDeviceToHost));
	if (z) CT_HIP(hipMemcpy(z, b->z.p, n * 4, hipMemcpyDeviceToHost));
	if (vel) {
		std::vector<float> t(n);
		float* comp[3] = {b->vx.p, b->vy.p, b->vz.p};
		for (int d = 0; d < 3; d++) {
			CT_HIP(hipMemcpy(t.data(), comp[d], n * 4, hipMemcpyDeviceToHost));
			for (int64_t i = 0; i < n; i++) vel[3 * i + d] = t[i];
		}
	}
	if (rung) CT_HIP(hipMemcpy(rung, b->rung.p, n, hipMemcpyDeviceToHost));
	if (gx) CT_HIP(hipMemcpy(gx, b->gx.p, n * 4, hipMemcpyDeviceToHost));
	if (gy) CT_HIP(hipMemcpy(gy, b->gy.p, n * 4, hipMemcpyDeviceToHost));
	if (gz) CT_HIP(hipMemcpy(gz, b->gz.p, n * 4, hipMemcpyDeviceToHost));
	if (phi) CT_HIP(hipMemcpy(phi, b->phi.p, n * 4, hipMemcpyDeviceToHost));
	if (id) CT_HIP(hipMemcpy(id, b->id.p, n * 4, hipMemcpyDeviceToHost));
	return CT_OK;
}

static void host_node_to_abi(const HostNode& h, ct_node& o) {
	for (int q = 0; q < 3; q++) {
		o.box_begin[q] = h.box_begin[q];
		o.box_end[q] = h.box_end[q];
		o.pos[q] = h.pos[q];
	}
	o.child[0] = h.child[0];
	o.child[1] = h.child[1];
	o.parent = h.parent;
	o.pbegin = h.pbegin;
	o.pend = h.pend;
	o.radius = h.radius;
	o.leaf = h.leaf;
	o.depth = h.depth;
	std::memcpy(o.multi, h.multi, sizeof(o.multi));
}

int ct_tree_build_host(int64_t n, uint32_t* x, uint32_t* y, uint32_t* z, int32_t* perm, double theta, int bucket, double hsoft,
		ct_node* nodes, int64_t node_cap, int64_t* nnodes) {
	if (n < 0 || n > INT32_MAX || (n > 0 && (!x || !y || !z)) || !(theta > 0.0) || bucket < 1 || bucket > kMaxLeafSinks || !nnodes)
		return fail(CT_E_ARG, "ct_tree_build_host: bad arguments");
	HostParticles P;
	P.n = n;
	P.x.assign(x, x + n);
	P.y.assign(y, y + n);
	P.z.assign(z, z + n);
	P.vx.assign(n, 0.0f);
	P.vy.assign(n, 0.0f);
	P.vz.assign(n, 0.0f);
	P.rung.assign(n, 0);
	P.id.resize(n);
	for (int64_t i = 0; i < n; i++) P.id[i] = int32_t(i);
	TreeBuildParams prm;
	prm.theta = theta;
	prm.bucket = bucket;
	prm.hsoft = hsoft;
	HostTree T;
	std::string err;
	int rc = build_tree(P, prm, T, &err);
	if (rc) return fail(rc, err);
	*nnodes = int64_t(T.nodes.size());
	if (!nodes || node_cap < *nnodes) return fail(CT_E_ARG, "ct_tree_build_host: node capacity too small");
	std::memcpy(x, P.x.data(), n * 4);
	std::memcpy(y, P.y.data(), n * 4);
	std::memcpy(z, P.z.data(), n * 4);
	if (perm) std::memcpy(perm, P.id.data(), n * 4);
	for (size_t i = 0; i < T.nodes.size(); i++) host_node_to_abi(T.nodes[i], nodes[i]);
	return CT_OK;
}

int ct_tree_read(ct_box_t b, ct_node* nodes, int64_t cap) {
	if (!b || !nodes) return fail(CT_E_ARG, "ct_tree_read: bad arguments");
	if (!b->have_tree) return fail(CT_E_STATE, "ct_tree_read: no tree");
	if (cap < b->nnodes) return fail(CT_E_ARG, "ct_tree_read: capacity too small");
	if (b->host_nodes) {
		for (int64_t i = 0; i < b->nnodes; i++) host_node_to_abi(b->tree.nodes[i], nodes[i]);
		return CT_OK;
	}
	// built on the device: read the node table and multipoles back
	const int64_t nn = b->nnodes;
	std::vector<DNode> dn(nn);
	std::vector<float> mp(nn * kMultiStride);
	CT_HIP(hipDeviceSynchronize());
	if (nn) {
		CT_HIP(hipMemcpy(dn.data(), b->nodes.p, nn * sizeof(DNode), hipMemcpyDeviceToHost));
		CT_HIP(hipMemcpy(mp.data(), b->multi.p, mp.size() * sizeof(float), hipMemcpyDeviceToHost));
	}
	for (int64_t i = 0; i < nn; i++) {
		const DNode& d = dn[i];
		ct_node& o = nodes[i];
		for (int q = 0; q < 3; q++) {
			o.box_begin[q] = d.bb[q];
			o.box_end[q] = d.be[q];
			o.pos[q] = d.pos[q];
		}
		o.child[0] = d.child[0];
		o.child[1] = d.child[1];
		o.parent = d.parent;
		o.pbegin = d.pbegin;
		o.pend = d.pend;
		o.radius = d.radius;
		o.leaf = d.flags & 1;
		o.depth = d.flags >> 1;
		std::memcpy(o.multi, &mp[i * kMultiStride], sizeof(o.multi));
	}
	return CT_OK;
}

int ct_lists_read(ct_box_t b, int kind, int64_t* off, int32_t* idx, int64_t idx_cap) {
	if (!b || kind < 0 || kind >= kNumLists || !off) return fail(CT_E_ARG, "ct_lists_read: bad arguments");
	KickCtx* c = static_cast<KickCtx*>(b->self_ctx);
	if (!c || !c->have_lists) return fail(CT_E_STATE, "ct_lists_read: no kick has run");
	const int64_t nn = b->nnodes;
	std::vector<int32_t> head(std::max<int64_t>(nn, 1));
	const Counters& C = c->last_counters;
	std::vector<Seg> seg(std::max(1, C.seg.v));
	std::vector<int32_t> out(std::max(1, C.out[kind].v));
	CT_HIP(hipDeviceSynchronize());
	if (nn) CT_HIP(hipMemcpy(head.data(), c->head.p + int64_t(kind) * nn, nn * 4, hipMemcpyDeviceToHost));
	if (C.seg.v) CT_HIP(hipMemcpy(seg.data(), c->seg.p, C.seg.v * sizeof(Seg), hipMemcpyDeviceToHost));
	if (C.out[kind].v) CT_HIP(hipMemcpy(out.data(), c->out[kind].p, int64_t(C.out[kind].v) * 4, hipMemcpyDeviceToHost));
	int64_t tot = 0;
	for (int64_t i = 0; i < nn; i++) {
		off[i] = tot;
		// segments are chained newest first; emit oldest first
		std::vector<int32_t> chain;
		for (int s = head[i]; s >= 0; s = seg[s].next) chain.push_back(s);
		for (auto it = chain.rbegin(); it != chain.rend(); ++it) {
			const Seg& sg = seg[*it];
			for (int j = 0; j < sg.len; j++) {
				if (idx) {
					if (tot >= idx_cap) return fail(CT_E_ARG, "ct_lists_read: capacity too small");
					idx[tot] = out[sg.begin + j];
				}
				tot++;
			}
		}
	}
	off[nn] = tot;
	return CT_OK;
}

int ct_kick_stats(ct_box_t b, double* out, int n) {
	if (!b || !out) return fail(CT_E_ARG, "ct_kick_stats: bad arguments");
	KickCtx* c = static_cast<KickCtx*>(b->self_ctx);
	for (int q = 0; q < n && q < 15; q++) out[q] = c ? c->stats[q] : 0.0;
	return CT_OK;
}

int ct_fmm_host(int op, const double* in, double* out) {
	if (!in || !out) return fail(CT_E_ARG, "ct_fmm_host: bad arguments");
	switch (op) {
	case 0:
		ctfmm::greens<double>(out, in[0], in[1], in[2]);
		return CT_OK;
	case 1: {
		const EwaldHost E = make_ewald_tables();
		ctfmm::ewald_greens<double>(out, in[0], in[1], in[2], E.rimg.data(), E.nreal, E.fh.data(), E.fc.data(), E.nfour);
		return CT_OK;
	}
	case 2:
		for (int q = 0; q < 50; q++) out[q] = 0.0;
		ctfmm::M2L<double>(out, in, in + 37, true);
		return CT_OK;
	case 3:
		ctfmm::L2L<double>(out, in, in[50], in[51], in[52], true);
		return CT_OK;
	case 4:
		ctfmm::L2P<double>(out, in, in[50], in[51], in[52], true);
		return CT_OK;
	case 5:
		ctfmm::P2M<double>(out, in[0], in[1], in[2]);
		return CT_OK;
	case 6:
		ctfmm::M2M<double>(out, in, in[37], in[38], in[39]);
		return CT_OK;
	case 7:
		for (int q = 0; q < 4; q++) out[q] = 0.0;
		ctfmm::M2L2<double>(out, in, in + 37, true);
		return CT_OK;
	default:
		return fail(CT_E_ARG, "ct_fmm_host: unknown op");
	}
}

int ct_fmm_device(int op, int64_t n, const double* in, double* out) {
	int rc = check_init();
	if (rc) return rc;
	static const int in_stride[8] = {3, 3, 87, 53, 53, 3, 40, 87};
	static const int out_stride[8] = {50, 50, 50, 50, 4, 37, 37, 4};
	if (op < 0 || op > 7 || n < 0 || !in || !out) return fail(CT_E_ARG, "ct_fmm_device: bad arguments");
	if (n == 0) return CT_OK;
	double *din = nullptr, *dout = nullptr;
	CT_HIP(hipMalloc(&din, n * in_stride[op] * sizeof(double)));
	if (hipMalloc(&dout, n * out_stride[op] * sizeof(double)) != hipSuccess) {
		(void) hipFree(din);
		return fail(CT_E_OOM, "ct_fmm_device: hipMalloc failed");
	}
	(void) hipMemcpy(din, in, n * in_stride[op] * sizeof(double), hipMemcpyHostToDevice);
	hipLaunchKernelGGL(k_fmm, dim3((n + 63) / 64), dim3(64), 0, 0, op, n, din, dout, in_stride[op], out_stride[op]);
	hipError_t e = hipMemcpy(out, dout, n * out_stride[op] * sizeof(double), hipMemcpyDeviceToHost);
	(void) hipFree(din);
	(void) hipFree(dout);
	if (e != hipSuccess) return fail(CT_E_HIP, std::string("ct_fmm_device: ") + hipGetErrorString(e));
	return CT_OK;
}

const char* ct_last_error(void) {
	return g_err.c_str();
}

}  // extern "C"

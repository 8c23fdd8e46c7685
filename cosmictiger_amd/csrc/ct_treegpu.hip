// Tree build on the device (drop-in for tree_create, src/tree.cpp:273-601,
// with the partition of src/particles.cpp:979-1018).
//
// The reference splits a node at the midpoint of the longest side of its
// geometric box, ties going to the first axis (src/tree.cpp:326-347).  Boxes
// start as the unit cube and are halved, so the cuts run x, y, z, x, ... and
// the cut of a node at depth d tests bit 31 - floor(d/3) of coordinate d mod 3:
// the tree is a bisection of the 63-bit Morton key (x31 y31 z31 x30 ...).  The
// build therefore
//   1. sorts the particles by Morton key (rocPRIM radix sort) and gathers the
//      particle arrays in that order (a node's particles are then a contiguous
//      range, as after the reference's in-place partition; the order inside a
//      range differs, the membership does not);
//   2. generates the nodes level by level in breadth-first order, splitting a
//      node when nparts > bucket or (!ewald_satisfied && nparts > 0) (the
//      per-depth test of src/tree.cpp:315-324 is tabulated on the host), its
//      children's ranges found by binary search of the split bit;
//   3. computes geometry bottom-up: leaves (tight box, centre, radius, P2M in
//      double), then internal nodes level by level (the two candidate centres,
//      the clip, M2M in double), with the host build's arithmetic operation for
//      operation so that integer and fixed32 outputs are bit-identical.
// Nodes deeper than 63 cannot be split on the key; such a tree is refused
// (CT_E_TREE) and the caller falls back to the host build.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cmath>
#include <cstring>
#include <vector>

#include "ct_expansions.hpp"
#include "ct_host.hpp"
#include "ct_treegpu.hpp"

namespace ct {
namespace {

constexpr int kKeyBits = 63;

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
	// top 21 bits of v at every third bit position (bit 20 -> 60, ..., bit 0 -> 0)
	uint64_t x = uint64_t(v >> 11) & 0x1fffffull;
	x = (x | (x << 32)) & 0x1f00000000ffffull;
	x = (x | (x << 16)) & 0x1f0000ff0000ffull;
	x = (x | (x << 8)) & 0x100f00f00f00f00full;
	x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
	x = (x | (x << 2)) & 0x1249249249249249ull;
	return x;
}

// key bit 62 - 3j - k = bit 31 - j of coordinate k
__device__ __forceinline__ uint64_t morton63(uint32_t x, uint32_t y, uint32_t z) {
	return (spread21(x) << 2) | (spread21(y) << 1) | spread21(z);
}

__global__ void k_keys(const uint32_t* __restrict__ x, const uint32_t* __restrict__ y, const uint32_t* __restrict__ z, int n,
		uint64_t* __restrict__ key, int32_t* __restrict__ idx) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	key[i] = morton63(x[i], y[i], z[i]);
	idx[i] = i;
}

struct GatherArgs {
	const int32_t* idx;
	const uint32_t *sx, *sy, *sz;
	const float *svx, *svy, *svz;
	const int8_t* srung;
	const int32_t* sid;
	uint32_t *dx, *dy, *dz;
	float *dvx, *dvy, *dvz;
	int8_t* drung;
	int32_t* did;
	int n;
	int first;  // particles below first keep their place
};

__global__ void k_gather(GatherArgs A) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= A.n) return;
	const int j = i < A.first ? i : A.first + A.idx[i - A.first];
	A.dx[i] = A.sx[j];
	A.dy[i] = A.sy[j];
	A.dz[i] = A.sz[j];
	A.dvx[i] = A.svx[j];
	A.dvy[i] = A.svy[j];
	A.dvz[i] = A.svz[j];
	A.drung[i] = A.srung[j];
	A.did[i] = A.sid[j];
}

// topology of one node during the build
struct TNode {
	uint64_t prefix;  // key bits above the node's depth
	int32_t pbegin, pend;
	int32_t depth;
	int32_t parent;
	int32_t child0;  // -1 for leaves
	int32_t pad;
};

// Split decision of every node of one level; flag[q] = 1 if node q splits.
__global__ void k_split_flags(const TNode* __restrict__ nodes, int lb, int le, const uint8_t* __restrict__ ewald_ok, int bucket,
		int32_t* __restrict__ flag, int32_t* __restrict__ err) {
	const int q = lb + blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= le) return;
	const TNode t = nodes[q];
	const int np = t.pend - t.pbegin;
	int s = np > bucket || (!ewald_ok[t.depth] && np > 0);
	if (s && t.depth >= kKeyBits) {
		atomicOr(err, 1);
		s = 0;
	}
	flag[q - lb] = s;
}

// Children of the splitting nodes of one level, written at the next level in
// breadth-first order (the scan of the split flags).
// The number of nodes of a level that split (the scan's last entry plus the
// last flag) and the depth error flag, side by side for one host read.
__global__ void k_level_count(const int32_t* __restrict__ scan, const int32_t* __restrict__ flag, int cnt, int32_t* __restrict__ err) {
	if (threadIdx.x == 0) {
		err[1] = scan[cnt - 1] + flag[cnt - 1];
		err[2] = err[0];
	}
}

__global__ void k_children(TNode* __restrict__ nodes, int lb, int le, const int32_t* __restrict__ flag, const int32_t* __restrict__ scan,
		const uint64_t* __restrict__ key, int koff, int next) {
	const int q = lb + blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= le) return;
	if (!flag[q - lb]) {
		nodes[q].child0 = -1;
		return;
	}
	const TNode t = nodes[q];
	const int bit = kKeyBits - 1 - t.depth;
	const uint64_t hi = t.prefix | (1ull << bit);
	// first particle of the range whose key has the split bit set
	int a = t.pbegin, b = t.pend;
	while (a < b) {
		const int m = a + ((b - a) >> 1);
		if (key[m - koff] < hi) {
			a = m + 1;
		} else {
			b = m;
		}
	}
	const int c = next + 2 * scan[q - lb];
	TNode l, r;
	l.prefix = t.prefix;
	r.prefix = hi;
	l.pbegin = t.pbegin;
	l.pend = a;
	r.pbegin = a;
	r.pend = t.pend;
	l.depth = r.depth = t.depth + 1;
	l.parent = r.parent = q;
	l.child0 = r.child0 = -1;
	l.pad = r.pad = 0;
	nodes[c] = l;
	nodes[c + 1] = r;
	nodes[q].child0 = c;
}

// geometry of one node during the build (double, as the host build keeps it)
struct GNode {
	double bb[3], be[3];  // tight box
	double multi[37];
	uint32_t pos[3];
	float radius;
};

__device__ __forceinline__ double dmax(double a, double b) {
	return (a < b) ? b : a;  // std::max
}
__device__ __forceinline__ double dmin(double a, double b) {
	return (b < a) ? b : a;  // std::min
}
__device__ __forceinline__ float fmax_std(float a, float b) {
	return (a < b) ? b : a;
}
__device__ __forceinline__ double fx2d(uint32_t a) {
	return double(a) * (1.0 / 4294967296.0);
}
__device__ __forceinline__ uint32_t d2fx(double x) {
	return uint32_t(uint64_t(int64_t(x * 4294967296.0)));
}
__device__ __forceinline__ int32_t d2rf(double x) {
	return int32_t(x * 536870912.0);
}
__device__ __forceinline__ float fdist_t(uint32_t a, uint32_t b) {
	return float(int32_t(a - b)) * (1.0f / 4294967296.0f);
}

// geometric box of a node from its key prefix: coordinate k has
// m_k = ceil((depth - k) / 3) bits fixed
__device__ __forceinline__ void node_box(uint64_t prefix, int depth, double* gb, double* ge) {
	for (int k = 0; k < 3; k++) {
		const int m = depth > k ? (depth - k + 2) / 3 : 0;
		uint32_t v = 0;
		for (int j = 0; j < m; j++) {
			v = (v << 1) | uint32_t((prefix >> (kKeyBits - 1 - 3 * j - k)) & 1);
		}
		const double w = ldexp(1.0, -m);
		gb[k] = double(v) * w;
		ge[k] = gb[k] + w;
	}
}

__device__ __forceinline__ double wave_min_d(double v) {
	for (int o = 32; o > 0; o >>= 1) v = dmin(v, __shfl_xor(v, o));
	return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
	for (int o = 32; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o));
	return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
	return v;
}

// Leaves (src/tree.cpp:499-571): tight box, its centre, radius = max distance
// in double, P2M about fixed32(centre) from float offsets, in double.  One
// wave per node; internal nodes are skipped.
__global__ __launch_bounds__(64) void k_leaf_geom(const TNode* __restrict__ nodes, int nn, const uint32_t* __restrict__ x,
		const uint32_t* __restrict__ y, const uint32_t* __restrict__ z, GNode* __restrict__ geo) {
	const int lane = threadIdx.x;
	for (int q = blockIdx.x; q < nn; q += gridDim.x) {
		const TNode t = nodes[q];
		if (t.child0 >= 0) continue;
		double gb[3], ge[3];
		node_box(t.prefix, t.depth, gb, ge);
		double bb[3], be[3];
		for (int d = 0; d < 3; d++) {
			bb[d] = ge[d];
			be[d] = gb[d];
		}
		const uint32_t* P[3] = {x, y, z};
		for (int d = 0; d < 3; d++) {
			double lo = bb[d], hi = be[d];
			for (int i = t.pbegin + lane; i < t.pend; i += 64) {
				const double X = fx2d(P[d][i]);
				hi = dmax(hi, X);
				lo = dmin(lo, X);
			}
			bb[d] = wave_min_d(lo);
			be[d] = wave_max_d(hi);
		}
		double Xc[3];
		for (int d = 0; d < 3; d++) Xc[d] = (be[d] + bb[d]) * 0.5;
		double r = 0.0;
		for (int i = t.pbegin + lane; i < t.pend; i += 64) {
			double qq = 0.0;
			for (int d = 0; d < 3; d++) {
				const double X = fx2d(P[d][i]);
				qq += (X - Xc[d]) * (X - Xc[d]);
			}
			r = dmax(r, qq);
		}
		r = wave_max_d(r);
		GNode& g = geo[q];
		uint32_t pos[3];
		for (int d = 0; d < 3; d++) pos[d] = d2fx(Xc[d]);
		double m[37];
		for (int k = 0; k < 37; k++) m[k] = 0.0;
		for (int i = t.pbegin + lane; i < t.pend; i += 64) {
			double mi[37];
			ctfmm::P2M<double>(mi, double(fdist_t(x[i], pos[0])), double(fdist_t(y[i], pos[1])), double(fdist_t(z[i], pos[2])));
			for (int k = 0; k < 37; k++) m[k] += mi[k];
		}
		for (int k = 0; k < 37; k++) {
			const double s = wave_sum_d(m[k]);
			if (lane == k) g.multi[k] = s;
		}
		if (lane == 0) {
			for (int d = 0; d < 3; d++) {
				g.bb[d] = bb[d];
				g.be[d] = be[d];
				g.pos[d] = pos[d];
			}
			g.radius = float(__dsqrt_rn(r));
		}
	}
}

// Internal nodes of one level (src/tree.cpp:361-495): tight box from the
// children, centre candidates (tight-box centre with the float max particle
// distance; sphere enclosing the children's spheres), clip by the tight box
// half-diagonal, fixed32 centre by truncation, M2M in double.  One wave per
// node.
__global__ __launch_bounds__(64) void k_internal_geom(const TNode* __restrict__ nodes, int lb, int le, int bucket,
		const uint32_t* __restrict__ x, const uint32_t* __restrict__ y, const uint32_t* __restrict__ z, GNode* __restrict__ geo) {
	const int lane = threadIdx.x;
	for (int q = lb + blockIdx.x; q < le; q += gridDim.x) {
		const TNode t = nodes[q];
		if (t.child0 < 0) continue;
		const GNode& L = geo[t.child0];
		const GNode& R = geo[t.child0 + 1];
		const int64_t nparts = t.pend - t.pbegin;
		const int xdim = t.depth % 3;
		double bb[3], be[3];
		const bool lnz = L.multi[0] != 0.0, rnz = R.multi[0] != 0.0;
		for (int d = 0; d < 3; d++) {
			if (lnz && rnz) {
				bb[d] = dmin(L.bb[d], R.bb[d]);
				be[d] = dmax(L.be[d], R.be[d]);
			} else if (lnz) {
				bb[d] = L.bb[d];
				be[d] = L.be[d];
			} else {
				bb[d] = R.bb[d];
				be[d] = R.be[d];
			}
		}
		double Xl[3], Xr[3], Xc1[3] = {0, 0, 0}, Xc2[3], Xc[3];
		for (int d = 0; d < 3; d++) {
			Xl[d] = fx2d(L.pos[d]);
			Xr[d] = fx2d(R.pos[d]);
		}
		double radius1 = 0.0;
		bool tried1 = false;
		if (nparts <= 64 * int64_t(bucket) && (be[xdim] - bb[xdim] <= 0.5)) {
			tried1 = true;
			uint32_t xc[3];
			for (int d = 0; d < 3; d++) {
				Xc1[d] = 0.5 * (bb[d] + be[d]);
				xc[d] = d2fx(Xc1[d]);
			}
			float r2max = 0.0f;
			for (int i = t.pbegin + lane; i < t.pend; i += 64) {
				const float dx = fdist_t(x[i], xc[0]);
				const float dy = fdist_t(y[i], xc[1]);
				const float dz = fdist_t(z[i], xc[2]);
				const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
				r2max = fmax_std(r2max, r2);
			}
			for (int o = 32; o > 0; o >>= 1) r2max = fmax_std(r2max, __shfl_xor(r2max, o));
			// correctly rounded float square root, as the host's sqrtf: the
			// double root rounded to float (double rounding cannot change a
			// square root when 53 >= 2 * 24 + 2); the device's float root
			// can be an ulp low
			radius1 = double(float(__dsqrt_rn(double(r2max))));
		}
		const double Rl = L.radius, Rr = R.radius;
		double N[3];
		double norminv = 0.0;
		for (int d = 0; d < 3; d++) {
			N[d] = Xl[d] - Xr[d];
			norminv += N[d] * N[d];
		}
		norminv = 1.0 / __dsqrt_rn(norminv);
		for (int d = 0; d < 3; d++) N[d] = fabs(N[d] * norminv);
		double r = 0.0;
		if (rnz && lnz) {
			for (int d = 0; d < 3; d++) {
				const double xmax = dmax(Xl[d] + N[d] * Rl, Xr[d] + N[d] * Rr);
				const double xmin = dmin(Xl[d] - N[d] * Rl, Xr[d] - N[d] * Rr);
				Xc2[d] = (xmax + xmin) * 0.5;
				r += ((xmax - xmin) * 0.5) * ((xmax - xmin) * 0.5);
			}
		} else if (rnz) {
			for (int d = 0; d < 3; d++) Xc2[d] = Xr[d];
			r = Rr * Rr;
		} else {
			for (int d = 0; d < 3; d++) Xc2[d] = Xl[d];
			r = Rl * Rl;
		}
		const double radius2 = __dsqrt_rn(r);
		float radius;
		if (tried1 && radius1 < radius2) {
			for (int d = 0; d < 3; d++) Xc[d] = Xc1[d];
			radius = float(radius1);
		} else {
			for (int d = 0; d < 3; d++) Xc[d] = Xc2[d];
			radius = float(radius2);
		}
		r = 0.0;
		for (int d = 0; d < 3; d++) {
			const double half = 0.5 * (be[d] - bb[d]);
			r += half * half;
		}
		r = __dsqrt_rn(r);
		if (r < double(radius)) {
			radius = float(r);
			for (int d = 0; d < 3; d++) Xc[d] = (bb[d] + be[d]) * 0.5;
		}
		if (lane == 0) {
			GNode& g = geo[q];
			double ml[37], mr[37];
			ctfmm::M2M<double>(ml, L.multi, Xl[0] - Xc[0], Xl[1] - Xc[1], Xl[2] - Xc[2]);
			ctfmm::M2M<double>(mr, R.multi, Xr[0] - Xc[0], Xr[1] - Xc[1], Xr[2] - Xc[2]);
			for (int k = 0; k < 37; k++) g.multi[k] = ml[k] + mr[k];
			for (int d = 0; d < 3; d++) {
				g.bb[d] = bb[d];
				g.be[d] = be[d];
				g.pos[d] = d2fx(Xc[d]);
			}
			g.radius = radius;
		}
	}
}

// Device node table and multipole records from the build data.
__global__ void k_emit(const TNode* __restrict__ nodes, const GNode* __restrict__ geo, int nn, DNode* __restrict__ dn, float* __restrict__ multi,
		double* __restrict__ multi64, float* __restrict__ L) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= nn) return;
	const TNode t = nodes[q];
	const GNode& g = geo[q];
	DNode d;
	for (int k = 0; k < 3; k++) {
		d.pos[k] = g.pos[k];
		d.bb[k] = d2rf(g.bb[k]);
		d.be[k] = d2rf(g.be[k]);
	}
	d.radius = g.radius;
	d.child[0] = t.child0;
	d.child[1] = t.child0 < 0 ? -1 : t.child0 + 1;
	d.pbegin = t.pbegin;
	d.pend = t.pend;
	d.flags = (t.child0 < 0 ? 1 : 0) | (t.depth << 1);
	d.parent = t.parent;
	dn[q] = d;
	float* mp = multi + int64_t(q) * kMultiStride;
	double* mp64 = multi64 + int64_t(q) * kMultiStride;
	for (int k = 0; k < 37; k++) {
		mp[k] = float(g.multi[k]);
		mp64[k] = g.multi[k];
	}
	for (int k = 0; k < 3; k++) {
		mp[37 + k] = __uint_as_float(g.pos[k]);
		mp64[37 + k] = double(g.pos[k]);
	}
	float* Lq = L + int64_t(q) * kLStride;
	for (int k = 0; k < kLStride; k++) Lq[k] = 0.0f;
}

}  // namespace

// Per-depth split test of the host build (src/tree.cpp:315-324): the
// geometric box of a depth-d node has the same size everywhere.
static std::vector<uint8_t> ewald_table(double theta, double hsoft) {
	std::vector<uint8_t> t(kKeyBits + 1);
	for (int d = 0; d <= kKeyBits; d++) {
		float box_r = 0.0f;
		for (int k = 0; k < 3; k++) {
			const int m = d > k ? (d - k + 2) / 3 : 0;
			const double s = 0.5 * std::ldexp(1.0, -m);
			box_r = float(double(box_r) + s * s);
		}
		box_r = std::sqrt(box_r);
		t[d] = double(box_r) < 0.25 * (theta / (1.0 + theta)) && double(box_r) < 0.125 - 0.25 * hsoft;
	}
	return t;
}

// Scratch of a box's device tree builds, kept with the box (grow-only) so
// that rebuilding a resident box's tree allocates nothing after the first
// build.  The particle arrays are swapped with the box's on each build.
struct TreeScratch {
	DBuf<uint64_t> kin, kout;
	DBuf<int32_t> iin, iout;
	DBuf<uint32_t> tx, ty, tz;
	DBuf<float> tvx, tvy, tvz;
	DBuf<int8_t> trung;
	DBuf<int32_t> tid;
	DBuf<uint8_t> ewd, tmp;
	DBuf<int32_t> err, flag, scan;
	DBuf<TNode> tn;
	DBuf<GNode> geo;
	// The builds' own stream, non-blocking: a box's rebuild does not wait for
	// (or hold up) work that other host threads put on the null stream.
	hipStream_t st = nullptr;
	void release() {
		if (st) {
			(void) hipStreamDestroy(st);
			st = nullptr;
		}
		kin.release();
		kout.release();
		iin.release();
		iout.release();
		tx.release();
		ty.release();
		tz.release();
		tvx.release();
		tvy.release();
		tvz.release();
		trung.release();
		tid.release();
		ewd.release();
		tmp.release();
		err.release();
		flag.release();
		scan.release();
		tn.release();
		geo.release();
	}
};

void tree_scratch_free(ct_box* b) {
	if (b->tree_scratch) {
		TreeScratch* S = static_cast<TreeScratch*>(b->tree_scratch);
		S->release();
		delete S;
		b->tree_scratch = nullptr;
	}
}

int gpu_tree_build(ct_box* b, double theta, int bucket, double hsoft, GpuTreeResult* res) {
	if (!b->tree_scratch) {
		b->tree_scratch = new (std::nothrow) TreeScratch();
		if (!b->tree_scratch) return fail(CT_E_OOM, "gpu tree: host allocation failed");
	}
	TreeScratch& S = *static_cast<TreeScratch*>(b->tree_scratch);
	const int n = int(b->hp.n);
	// the tree covers the active range [rb, n) (particles_current_range)
	const int rb = int(b->rung_begin);
	const int m = n - rb;
	if (!S.st) CT_HIP(hipStreamCreateWithFlags(&S.st, hipStreamNonBlocking));
	hipStream_t st = S.st;
	int rc;
	auto& kin = S.kin;
	auto& kout = S.kout;
	auto& iin = S.iin;
	auto& iout = S.iout;
	auto& tx = S.tx;
	auto& ty = S.ty;
	auto& tz = S.tz;
	auto& tvx = S.tvx;
	auto& tvy = S.tvy;
	auto& tvz = S.tvz;
	auto& trung = S.trung;
	auto& tid = S.tid;
	auto& tmp = S.tmp;
	auto cleanup = []() {};  // the scratch stays with the box
	const int64_t nn1 = std::max(n, 1);
	const int64_t nm1 = std::max(m, 1);
	if ((rc = kin.grow(nm1)) || (rc = kout.grow(nm1)) || (rc = iin.grow(nm1)) || (rc = iout.grow(nm1)) || (rc = tx.grow(nn1)) ||
			(rc = ty.grow(nn1)) || (rc = tz.grow(nn1)) || (rc = tvx.grow(nn1)) || (rc = tvy.grow(nn1)) || (rc = tvz.grow(nn1)) ||
			(rc = trung.grow(nn1)) || (rc = tid.grow(nn1))) {
		cleanup();
		return rc;
	}
	// 1. sort by Morton key, gather the particle arrays
	if (m > 0) {
		hipLaunchKernelGGL(k_keys, dim3((m + 255) / 256), dim3(256), 0, st, b->x.p + rb, b->y.p + rb, b->z.p + rb, m, kin.p, iin.p);
		size_t tb = 0;
		if (rocprim::radix_sort_pairs(nullptr, tb, kin.p, kout.p, iin.p, iout.p, size_t(m), 0, kKeyBits, st) != hipSuccess) {
			cleanup();
			return fail(CT_E_HIP, "gpu tree: radix sort size query failed");
		}
		if ((rc = tmp.grow(int64_t(tb) + 1))) {
			cleanup();
			return rc;
		}
		hipError_t e = rocprim::radix_sort_pairs(tmp.p, tb, kin.p, kout.p, iin.p, iout.p, size_t(m), 0, kKeyBits, st);
		(void) hipStreamSynchronize(st);
		if (e != hipSuccess) {
			cleanup();
			return fail(CT_E_HIP, std::string("gpu tree: radix sort: ") + hipGetErrorString(e));
		}
		GatherArgs ga;
		ga.idx = iout.p;
		ga.sx = b->x.p;
		ga.sy = b->y.p;
		ga.sz = b->z.p;
		ga.svx = b->vx.p;
		ga.svy = b->vy.p;
		ga.svz = b->vz.p;
		ga.srung = b->rung.p;
		ga.sid = b->id.p;
		ga.dx = tx.p;
		ga.dy = ty.p;
		ga.dz = tz.p;
		ga.dvx = tvx.p;
		ga.dvy = tvy.p;
		ga.dvz = tvz.p;
		ga.drung = trung.p;
		ga.did = tid.p;
		ga.n = n;
		ga.first = rb;
		hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, st, ga);
		CT_HIP(hipGetLastError());
		std::swap(b->x, tx);
		std::swap(b->y, ty);
		std::swap(b->z, tz);
		std::swap(b->vx, tvx);
		std::swap(b->vy, tvy);
		std::swap(b->vz, tvz);
		std::swap(b->rung, trung);
		std::swap(b->id, tid);
	}
	// 2. topology, level by level
	const std::vector<uint8_t> ew = ewald_table(theta, hsoft);
	auto& ewd = S.ewd;
	auto& err = S.err;
	auto& flag = S.flag;
	auto& scan = S.scan;
	std::vector<int32_t> level_off = {0, 1};
	int64_t cap = std::max<int64_t>(4 * int64_t(n) / std::max(1, bucket / 2) + 1024, 1024);
	auto& tn = S.tn;
	cap = std::max<int64_t>(cap, tn.n);
	if ((rc = ewd.grow(int64_t(ew.size()))) || (rc = err.grow(3)) || (rc = tn.grow(cap))) {
		cleanup();
		return rc;
	}
	// (host sources of the asynchronous copies live until the build's final
	// synchronisation)
	TNode root;
	root.prefix = 0;
	root.pbegin = rb;
	root.pend = n;
	root.depth = 0;
	root.parent = -1;
	root.child0 = -1;
	root.pad = 0;
	// declared after ew and root, so that every return waits for the stream
	// before they go
	struct StreamWait {
		hipStream_t s;
		~StreamWait() { (void) hipStreamSynchronize(s); }
	} wait_on_return{st};
	CT_HIP(hipMemcpyAsync(ewd.p, ew.data(), ew.size(), hipMemcpyHostToDevice, st));
	CT_HIP(hipMemsetAsync(err.p, 0, 4, st));
	CT_HIP(hipMemcpyAsync(tn.p, &root, sizeof(TNode), hipMemcpyHostToDevice, st));
	for (int d = 0;; d++) {
		const int lb = level_off[d], le = level_off[d + 1];
		const int cnt = le - lb;
		if ((rc = flag.grow(cnt + 1)) || (rc = scan.grow(cnt + 1))) {
			cleanup();
			return rc;
		}
		hipLaunchKernelGGL(k_split_flags, dim3((cnt + 255) / 256), dim3(256), 0, st, tn.p, lb, le, ewd.p, bucket, flag.p, err.p);
		size_t tb = 0;
		(void) rocprim::exclusive_scan(nullptr, tb, flag.p, scan.p, 0, size_t(cnt) + 0, rocprim::plus<int32_t>(), st);
		if ((rc = tmp.grow(int64_t(tb) + 1))) {
			cleanup();
			return rc;
		}
		if (rocprim::exclusive_scan(tmp.p, tb, flag.p, scan.p, 0, size_t(cnt), rocprim::plus<int32_t>(), st) != hipSuccess) {
			cleanup();
			return fail(CT_E_HIP, "gpu tree: scan failed");
		}
		hipLaunchKernelGGL(k_level_count, dim3(1), dim3(64), 0, st, scan.p, flag.p, cnt, err.p);
		int32_t lc[2] = {0, 0};
		CT_HIP(hipMemcpyAsync(lc, err.p + 1, sizeof(lc), hipMemcpyDeviceToHost, st));
		CT_HIP(hipStreamSynchronize(st));
		if (lc[1]) {
			cleanup();
			return fail(CT_E_TREE, "gpu tree: a node deeper than 63 levels must split");
		}
		const int nsplit = lc[0];
		const int next = le + 2 * nsplit;
		if (next > cap) {
			// grow the node arena, keeping what is built
			DBuf<TNode> bigger;
			const int64_t ncap = std::max<int64_t>(2 * cap, next + 1024);
			if ((rc = bigger.grow(ncap))) {
				cleanup();
				return rc;
			}
			CT_HIP(hipMemcpyAsync(bigger.p, tn.p, int64_t(le) * sizeof(TNode), hipMemcpyDeviceToDevice, st));
			CT_HIP(hipStreamSynchronize(st));
			tn.release();
			tn.p = bigger.p;
			tn.n = bigger.n;
			bigger.p = nullptr;
			cap = ncap;
		}
		hipLaunchKernelGGL(k_children, dim3((cnt + 255) / 256), dim3(256), 0, st, tn.p, lb, le, flag.p, scan.p, kout.p, rb, le);
		CT_HIP(hipGetLastError());
		if (nsplit == 0) break;
		level_off.push_back(next);
	}
	const int nn = level_off.back();
	const int max_depth = int(level_off.size()) - 2;
	// 3. geometry, bottom-up
	auto& geo = S.geo;
	if ((rc = geo.grow(nn)) || (rc = b->nodes.grow(nn)) || (rc = b->multi.grow(int64_t(nn) * kMultiStride)) ||
			(rc = b->multi64.grow(int64_t(nn) * kMultiStride)) || (rc = b->L.grow(int64_t(nn) * kLStride))) {
		cleanup();
		return rc;
	}
	hipLaunchKernelGGL(k_leaf_geom, dim3(std::min(nn, 1 << 15)), dim3(64), 0, st, tn.p, nn, b->x.p, b->y.p, b->z.p, geo.p);
	for (int d = max_depth - 1; d >= 0; d--) {
		const int lb = level_off[d], le = level_off[d + 1];
		hipLaunchKernelGGL(k_internal_geom, dim3(std::min(le - lb, 1 << 15)), dim3(64), 0, st, tn.p, lb, le, bucket, b->x.p, b->y.p, b->z.p,
				geo.p);
	}
	hipLaunchKernelGGL(k_emit, dim3((nn + 255) / 256), dim3(256), 0, st, tn.p, geo.p, nn, b->nodes.p, b->multi.p, b->multi64.p, b->L.p);
	CT_HIP(hipGetLastError());
	// host bookkeeping: per-node (pbegin, pend, leaf) for the leaf lists
	std::vector<TNode> ht(nn);
	CT_HIP(hipMemcpyAsync(ht.data(), tn.p, int64_t(nn) * sizeof(TNode), hipMemcpyDeviceToHost, st));
	CT_HIP(hipStreamSynchronize(st));
	res->nnodes = nn;
	res->max_depth = max_depth;
	res->level_off = level_off;
	res->leaf_pbegin.clear();
	res->leaf_pend.clear();
	res->leaf_id.clear();
	for (int q = 0; q < nn; q++) {
		if (ht[q].child0 < 0) {
			res->leaf_id.push_back(q);
			res->leaf_pbegin.push_back(ht[q].pbegin);
			res->leaf_pend.push_back(ht[q].pend);
		}
	}
	return CT_OK;
}

}  // namespace ct

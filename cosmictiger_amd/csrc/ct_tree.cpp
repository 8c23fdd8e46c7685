// Host tree build (drop-in for tree_create, src/tree.cpp:273-601).
//
// Same decisions as the reference build, so that the integer and fixed32
// outputs (boxes, particle ranges, centres, leaf flags, topology) are
// bit-identical to the oracle's restatement:
//  * split when nparts > bucket or (!ewald_satisfied && nparts > 0), with
//    ewald_satisfied on the geometric box half-diagonal (src/tree.cpp:315-324);
//  * bisect the longest geometric dimension at its midpoint, partitioning on
//    x >= fixed32(xmid) (src/tree.cpp:326-347, src/particles.cpp:979-1018);
//  * internal centre: the tighter of the tight-box centre with the float max
//    particle distance (tried for nparts <= 64 bucket and extent <= 0.5) and
//    the sphere enclosing the two child spheres, then clipped by the tight-box
//    half-diagonal (src/tree.cpp:361-476); centres stored by truncation;
//  * leaf: tight box, its centre, radius = max distance in double, P2M about
//    fixed32(centre) from float offsets (src/tree.cpp:499-571);
//  * M2M in double about the unrounded centre (src/tree.cpp:481-495).
// Subtrees above a size threshold are built on their own threads, as the
// reference forks HPX tasks (src/tree.cpp:349-352).
#pragma clang fp contract(off)

#include "ct_tree.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <string>
#include <thread>

#include "ct_expansions.hpp"

namespace ct {
namespace {

constexpr int kMaxDepth = 128;  // src/tree.cpp:285-287

inline double fx2d(uint32_t a) {
	return double(a) * (1.0 / 4294967296.0);
}

inline uint32_t d2fx(double x) {
	return uint32_t(uint64_t(int64_t(x * 4294967296.0)));
}

inline int32_t d2rf(double x) {
	return int32_t(x * 536870912.0);
}

inline float fdist(uint32_t a, uint32_t b) {
	return float(int32_t(a - b)) * (1.0f / 4294967296.0f);
}

struct Built {
	double multi[37];
	uint32_t pos[3];
	float radius;
	double bb[3], be[3];
	int32_t id;
};

struct Ctx {
	HostParticles* p;
	TreeBuildParams prm;
	std::vector<HostNode>* nodes;
	int64_t cap;
	std::atomic<int64_t> next{0};
	std::atomic<int> err{0};
	std::atomic<int> spare_threads{0};
	int64_t fork_min;
};

void swap_particles(HostParticles& p, int64_t i, int64_t j) {
	std::swap(p.x[i], p.x[j]);
	std::swap(p.y[i], p.y[j]);
	std::swap(p.z[i], p.z[j]);
	std::swap(p.vx[i], p.vx[j]);
	std::swap(p.vy[i], p.vy[j]);
	std::swap(p.vz[i], p.vz[j]);
	std::swap(p.rung[i], p.rung[j]);
	std::swap(p.id[i], p.id[j]);
}

// In-place partition of [b, e): keys < xm first.  Returns the split index.
int64_t split_particles(HostParticles& p, int64_t b, int64_t e, uint32_t xm, int dim) {
	const uint32_t* key = dim == 0 ? p.x.data() : (dim == 1 ? p.y.data() : p.z.data());
	int64_t i = b, j = e;
	for (;;) {
		while (i < j && key[i] < xm) {
			++i;
		}
		while (i < j && key[j - 1] >= xm) {
			--j;
		}
		if (i >= j) {
			break;
		}
		swap_particles(p, i, j - 1);
		++i;
		--j;
	}
	return i;
}

Built build(Ctx& c, int64_t b, int64_t e, const double gb[3], const double ge[3], int depth) {
	Built rc;
	rc.id = -1;
	if (depth >= kMaxDepth) {
		c.err = 1;
		return rc;
	}
	const int64_t index = c.next.fetch_add(1);
	if (index >= c.cap) {
		c.err = 2;
		return rc;
	}
	HostParticles& P = *c.p;
	const int64_t nparts = e - b;
	float box_r = 0.0f;
	for (int d = 0; d < 3; d++) {
		const double s = 0.5 * (ge[d] - gb[d]);
		box_r = float(double(box_r) + s * s);
	}
	box_r = std::sqrt(box_r);
	const double theta = c.prm.theta;
	const bool ewald_ok = double(box_r) < 0.25 * (theta / (1.0 + theta)) && double(box_r) < 0.125 - 0.25 * c.prm.hsoft;
	int32_t child[2] = {-1, -1};
	double Xc[3];
	float radius = 0.0f;
	bool leaf = true;
	if (nparts > c.prm.bucket || (!ewald_ok && nparts > 0)) {
		leaf = false;
		int xdim = 0;
		double span_max = -1.0;
		for (int d = 0; d < 3; d++) {
			const double span = ge[d] - gb[d];
			if (span > span_max) {
				span_max = span;
				xdim = d;
			}
		}
		const double xmid = 0.5 * (ge[xdim] + gb[xdim]);
		const int64_t mid = split_particles(P, b, e, d2fx(xmid), xdim);
		double lb[3], le[3], rb[3], re[3];
		for (int d = 0; d < 3; d++) {
			lb[d] = rb[d] = gb[d];
			le[d] = re[d] = ge[d];
		}
		le[xdim] = rb[xdim] = xmid;
		Built L, R;
		bool forked = false;
		if (nparts >= c.fork_min && c.spare_threads.fetch_sub(1) > 0) {
			forked = true;
			std::thread th([&]() { R = build(c, mid, e, rb, re, depth + 1); });
			L = build(c, b, mid, lb, le, depth + 1);
			th.join();
			c.spare_threads.fetch_add(1);
		} else {
			if (nparts >= c.fork_min) {
				c.spare_threads.fetch_add(1);
			}
			L = build(c, b, mid, lb, le, depth + 1);
			R = build(c, mid, e, rb, re, depth + 1);
		}
		(void) forked;
		if (c.err) {
			return rc;
		}
		const bool lnz = L.multi[0] != 0.0, rnz = R.multi[0] != 0.0;
		for (int d = 0; d < 3; d++) {
			if (lnz && rnz) {
				rc.bb[d] = std::min(L.bb[d], R.bb[d]);
				rc.be[d] = std::max(L.be[d], R.be[d]);
			} else if (lnz) {
				rc.bb[d] = L.bb[d];
				rc.be[d] = L.be[d];
			} else {
				rc.bb[d] = R.bb[d];
				rc.be[d] = R.be[d];
			}
		}
		double Xl[3], Xr[3], Xc1[3] = {0, 0, 0}, Xc2[3];
		for (int d = 0; d < 3; d++) {
			Xl[d] = fx2d(L.pos[d]);
			Xr[d] = fx2d(R.pos[d]);
		}
		double radius1 = 0.0;
		bool tried1 = false;
		if (nparts <= 64 * int64_t(c.prm.bucket) && (rc.be[xdim] - rc.bb[xdim] <= 0.5)) {
			tried1 = true;
			uint32_t xc[3];
			for (int d = 0; d < 3; d++) {
				Xc1[d] = 0.5 * (rc.bb[d] + rc.be[d]);
				xc[d] = d2fx(Xc1[d]);
			}
			float r2max = 0.0f;
			for (int64_t i = b; i < e; i++) {
				const float dx = fdist(P.x[i], xc[0]);
				const float dy = fdist(P.y[i], xc[1]);
				const float dz = fdist(P.z[i], xc[2]);
				const float r2 = std::fma(dx, dx, std::fma(dy, dy, dz * dz));
				r2max = std::max(r2max, r2);
			}
			radius1 = double(std::sqrt(r2max));
		}
		const double Rl = L.radius, Rr = R.radius;
		double N[3];
		double norminv = 0.0;
		for (int d = 0; d < 3; d++) {
			N[d] = Xl[d] - Xr[d];
			norminv += N[d] * N[d];
		}
		norminv = 1.0 / std::sqrt(norminv);
		for (int d = 0; d < 3; d++) {
			N[d] = std::fabs(N[d] * norminv);
		}
		double r = 0.0;
		if (rnz && lnz) {
			for (int d = 0; d < 3; d++) {
				const double xmax = std::max(Xl[d] + N[d] * Rl, Xr[d] + N[d] * Rr);
				const double xmin = std::min(Xl[d] - N[d] * Rl, Xr[d] - N[d] * Rr);
				Xc2[d] = (xmax + xmin) * 0.5;
				r += ((xmax - xmin) * 0.5) * ((xmax - xmin) * 0.5);
			}
		} else if (rnz) {
			for (int d = 0; d < 3; d++) {
				Xc2[d] = Xr[d];
			}
			r = Rr * Rr;
		} else {
			for (int d = 0; d < 3; d++) {
				Xc2[d] = Xl[d];
			}
			r = Rl * Rl;
		}
		const double radius2 = std::sqrt(r);
		if (tried1 && radius1 < radius2) {
			std::memcpy(Xc, Xc1, sizeof(Xc));
			radius = float(radius1);
		} else {
			std::memcpy(Xc, Xc2, sizeof(Xc));
			radius = float(radius2);
		}
		r = 0.0;
		for (int d = 0; d < 3; d++) {
			const double half = 0.5 * (rc.be[d] - rc.bb[d]);
			r += half * half;
		}
		r = std::sqrt(r);
		if (r < double(radius)) {
			radius = float(r);
			for (int d = 0; d < 3; d++) {
				Xc[d] = (rc.bb[d] + rc.be[d]) * 0.5;
			}
		}
		for (int d = 0; d < 3; d++) {
			rc.pos[d] = d2fx(Xc[d]);
		}
		double ml[37], mr[37];
		ctfmm::M2M<double>(ml, L.multi, Xl[0] - Xc[0], Xl[1] - Xc[1], Xl[2] - Xc[2]);
		ctfmm::M2M<double>(mr, R.multi, Xr[0] - Xc[0], Xr[1] - Xc[1], Xr[2] - Xc[2]);
		for (int i = 0; i < 37; i++) {
			rc.multi[i] = ml[i] + mr[i];
		}
		child[0] = L.id;
		child[1] = R.id;
	} else {
		for (int d = 0; d < 3; d++) {
			rc.bb[d] = ge[d];
			rc.be[d] = gb[d];
		}
		for (int64_t i = b; i < e; i++) {
			const double X[3] = {fx2d(P.x[i]), fx2d(P.y[i]), fx2d(P.z[i])};
			for (int d = 0; d < 3; d++) {
				rc.be[d] = std::max(rc.be[d], X[d]);
				rc.bb[d] = std::min(rc.bb[d], X[d]);
			}
		}
		for (int d = 0; d < 3; d++) {
			Xc[d] = (rc.be[d] + rc.bb[d]) * 0.5;
		}
		double r = 0.0;
		for (int64_t i = b; i < e; i++) {
			const double X[3] = {fx2d(P.x[i]), fx2d(P.y[i]), fx2d(P.z[i])};
			double q = 0.0;
			for (int d = 0; d < 3; d++) {
				q += (X[d] - Xc[d]) * (X[d] - Xc[d]);
			}
			r = std::max(r, q);
		}
		radius = float(std::sqrt(r));
		for (int d = 0; d < 3; d++) {
			rc.pos[d] = d2fx(Xc[d]);
		}
		for (int i = 0; i < 37; i++) {
			rc.multi[i] = 0.0;
		}
		for (int64_t i = b; i < e; i++) {
			double m[37];
			ctfmm::P2M<double>(m, double(fdist(P.x[i], rc.pos[0])), double(fdist(P.y[i], rc.pos[1])), double(fdist(P.z[i], rc.pos[2])));
			for (int k = 0; k < 37; k++) {
				rc.multi[k] += m[k];
			}
		}
	}
	rc.radius = radius;
	rc.id = int32_t(index);
	HostNode& n = (*c.nodes)[index];
	for (int d = 0; d < 3; d++) {
		n.box_begin[d] = d2rf(rc.bb[d]);
		n.box_end[d] = d2rf(rc.be[d]);
		n.pos[d] = rc.pos[d];
	}
	n.child[0] = child[0];
	n.child[1] = child[1];
	n.parent = -1;
	n.pbegin = int32_t(b);
	n.pend = int32_t(e);
	n.radius = radius;
	n.leaf = leaf ? 1 : 0;
	n.depth = depth;
	for (int i = 0; i < 37; i++) {
		n.multi[i] = float(rc.multi[i]);
		n.multid[i] = rc.multi[i];
	}
	return rc;
}

}  // namespace

int build_tree(HostParticles& p, const TreeBuildParams& prm, HostTree& out, std::string* err) {
	if (p.n < 0 || p.n > INT32_MAX || prm.bucket < 1 || !(prm.theta > 0.0) || prm.first < 0 || prm.first > p.n) {
		if (err) *err = "build_tree: bad arguments";
		return -1;
	}
	int64_t cap = 8 * (p.n - prm.first) / prm.bucket + 65536;
	for (int attempt = 0; attempt < 6; attempt++) {
		std::vector<HostNode> nodes(cap);
		Ctx c;
		c.p = &p;
		c.prm = prm;
		c.nodes = &nodes;
		c.cap = cap;
		int nth = prm.nthreads > 0 ? prm.nthreads : int(std::thread::hardware_concurrency());
		c.spare_threads = std::max(0, nth - 1);
		c.fork_min = 1 << 16;
		const double gb[3] = {0, 0, 0}, ge[3] = {1, 1, 1};
		// restore the original order is not needed: a retry re-partitions
		Built root = build(c, prm.first, p.n, gb, ge, 0);
		if (c.err == 2) {
			cap *= 2;
			continue;
		}
		if (c.err) {
			if (err) *err = "build_tree: maximum depth exceeded";
			return -6;
		}
		const int64_t nn = c.next.load();
		// breadth-first renumbering
		std::vector<int32_t> order;
		order.reserve(nn);
		order.push_back(root.id);
		std::vector<int32_t> newid(nn, -1);
		for (size_t q = 0; q < order.size(); q++) {
			const HostNode& n = nodes[order[q]];
			newid[order[q]] = int32_t(q);
			if (!n.leaf) {
				order.push_back(n.child[0]);
				order.push_back(n.child[1]);
			}
		}
		out.nodes.assign(order.size(), HostNode());
		out.max_depth = 0;
		out.nleaves = 0;
		for (size_t q = 0; q < order.size(); q++) {
			HostNode n = nodes[order[q]];
			if (!n.leaf) {
				n.child[0] = newid[n.child[0]];
				n.child[1] = newid[n.child[1]];
			} else {
				out.nleaves++;
			}
			out.max_depth = std::max(out.max_depth, int(n.depth));
			out.nodes[q] = n;
		}
		for (size_t q = 0; q < out.nodes.size(); q++) {
			if (!out.nodes[q].leaf) {
				out.nodes[out.nodes[q].child[0]].parent = int32_t(q);
				out.nodes[out.nodes[q].child[1]].parent = int32_t(q);
			}
		}
		out.level_off.assign(out.max_depth + 2, 0);
		for (const HostNode& n : out.nodes) {
			out.level_off[n.depth + 1]++;
		}
		for (int d = 0; d <= out.max_depth; d++) {
			out.level_off[d + 1] += out.level_off[d];
		}
		return 0;
	}
	if (err) *err = "build_tree: node arena exhausted";
	return -4;
}

}  // namespace ct

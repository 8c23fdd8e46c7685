// Active range, rung bookkeeping and the hierarchical step.
//
// The reference kicks and drifts only the particles of the active range
// [rung_begin, N) (particles_current_range, src/particles.cpp:1020-1025): a
// descending kick at rung r first moves the particles with rung >= r to the
// end of the range (particles_sort_by_rung, :1029-1068), builds the tree over
// them alone and kicks that tree.  kick_step_hierarchical
// (src/driver.cpp:382-593) strings these together into the kick ladder of one
// sub-step.  Here the partition is a stable device partition (flags, scan,
// scatter, gather), the range stack lives with the box, and the ladder is a
// host loop over the library's own tree build, kick and drift entry points.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ct_host.hpp"

namespace ct {
namespace {

// flag[j] = 1 when particle first + j stays below the active range
__global__ void k_rung_flags(const int8_t* __restrict__ rung, int m, int min_rung, int32_t* __restrict__ flag) {
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= m) return;
	flag[j] = int(rung[j]) < min_rung ? 1 : 0;
}

// src[k] = the particle (relative to the range start) that goes to place k:
// the particles that stay below first, in order, then the active ones, in order
__global__ void k_rung_scatter(const int32_t* __restrict__ flag, const int32_t* __restrict__ scan, int m, int nlow,
		int32_t* __restrict__ src) {
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= m) return;
	const int s = scan[j];
	const int k = flag[j] ? s : nlow + (j - s);
	src[k] = j;
}

template<class T>
__global__ void k_take(const T* __restrict__ in, const int32_t* __restrict__ src, int m, T* __restrict__ out) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= m) return;
	out[k] = in[src[k]];
}

__global__ void k_set_minrung(int8_t* __restrict__ rung, int64_t n, int min_rung) {
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	rung[i] = int8_t(max(int(rung[i]), min_rung));
}

// per-rung counts: a block histogram in LDS, then one atomic per rung
__global__ __launch_bounds__(256) void k_rung_hist(const int8_t* __restrict__ rung, int64_t n, unsigned long long* __restrict__ counts) {
	__shared__ unsigned int h[128];
	for (int q = threadIdx.x; q < 128; q += blockDim.x) h[q] = 0;
	__syncthreads();
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
		atomicAdd(&h[uint32_t(uint8_t(rung[i])) & 127u], 1u);
	}
	__syncthreads();
	for (int q = threadIdx.x; q < 128; q += blockDim.x)
		if (h[q]) atomicAdd(&counts[q], (unsigned long long) h[q]);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0ull;
}

// Moves [first, n) of one array through the permutation src.
template<class T>
int permute(DBuf<T>& a, int first, int m, const int32_t* src, DBuf<T>& tmp, hipStream_t st) {
	int rc;
	if ((rc = tmp.grow(std::max(m, 1)))) return rc;
	hipLaunchKernelGGL(k_take<T>, dim3((m + 255) / 256), dim3(256), 0, st, a.p + first, src, m, tmp.p);
	CT_HIP(hipGetLastError());
	CT_HIP(hipMemcpyAsync(a.p + first, tmp.p, int64_t(m) * sizeof(T), hipMemcpyDeviceToDevice, st));
	return CT_OK;
}

}  // namespace
}  // namespace ct

using namespace ct;

int ct_box_sort_by_rung(ct_box_t b, int min_rung, int64_t* rung_begin) {
	int rc = check_init();
	if (rc) return rc;
	if (!b || min_rung < 0 || min_rung > 127) return fail(CT_E_ARG, "ct_box_sort_by_rung: bad arguments");
	if (min_rung == 0) {
		b->rung_begin = 0;
		if (rung_begin) *rung_begin = 0;
		return CT_OK;
	}
	const int first = int(b->rung_begin);
	const int m = int(b->hp.n) - first;
	if (m > 0) {
		hipStream_t st = nullptr;
		DBuf<int32_t> flag, scan, src;
		auto cleanup = [&]() {
			flag.release();
			scan.release();
			src.release();
		};
		if ((rc = flag.grow(m)) || (rc = scan.grow(m)) || (rc = src.grow(m))) {
			cleanup();
			return rc;
		}
		hipLaunchKernelGGL(k_rung_flags, dim3((m + 255) / 256), dim3(256), 0, st, b->rung.p + first, m, min_rung, flag.p);
		size_t tb = 0;
		(void) rocprim::exclusive_scan(nullptr, tb, flag.p, scan.p, 0, size_t(m), rocprim::plus<int32_t>(), st);
		DBuf<uint8_t> tmp;
		if ((rc = tmp.grow(int64_t(tb) + 1))) {
			cleanup();
			return rc;
		}
		const hipError_t e = rocprim::exclusive_scan(tmp.p, tb, flag.p, scan.p, 0, size_t(m), rocprim::plus<int32_t>(), st);
		int32_t last_scan = 0, last_flag = 0;
		if (e == hipSuccess) {
			(void) hipMemcpy(&last_scan, scan.p + m - 1, 4, hipMemcpyDeviceToHost);
			(void) hipMemcpy(&last_flag, flag.p + m - 1, 4, hipMemcpyDeviceToHost);
		}
		tmp.release();
		if (e != hipSuccess) {
			cleanup();
			return fail(CT_E_HIP, "ct_box_sort_by_rung: scan failed");
		}
		const int nlow = last_scan + last_flag;
		if (nlow > 0 && nlow < m) {
			// some particles move: permute every particle array of the range
			hipLaunchKernelGGL(k_rung_scatter, dim3((m + 255) / 256), dim3(256), 0, st, flag.p, scan.p, m, nlow, src.p);
			DBuf<uint32_t> t32;
			DBuf<float> tf;
			DBuf<int8_t> t8;
			DBuf<int32_t> ti;
			rc = permute(b->x, first, m, src.p, t32, st);
			if (!rc) rc = permute(b->y, first, m, src.p, t32, st);
			if (!rc) rc = permute(b->z, first, m, src.p, t32, st);
			if (!rc) rc = permute(b->vx, first, m, src.p, tf, st);
			if (!rc) rc = permute(b->vy, first, m, src.p, tf, st);
			if (!rc) rc = permute(b->vz, first, m, src.p, tf, st);
			if (!rc) rc = permute(b->rung, first, m, src.p, t8, st);
			if (!rc) rc = permute(b->id, first, m, src.p, ti, st);
			const hipError_t se = hipStreamSynchronize(st);
			t32.release();
			tf.release();
			t8.release();
			ti.release();
			if (rc) {
				cleanup();
				return rc;
			}
			if (se != hipSuccess) {
				cleanup();
				return fail(CT_E_HIP, std::string("ct_box_sort_by_rung: ") + hipGetErrorString(se));
			}
			b->host_stale = true;
			b->have_tree = false;  // the tree's particle ranges no longer hold
		}
		cleanup();
		if (first + nlow != b->rung_begin) b->have_tree = false;  // the tree covers another range
		b->rung_begin = first + nlow;
	}
	if (rung_begin) *rung_begin = b->rung_begin;
	return CT_OK;
}

int ct_box_current_range(ct_box_t b, int64_t* first, int64_t* second) {
	if (!b) return fail(CT_E_ARG, "ct_box_current_range: null box");
	if (first) *first = b->rung_begin;
	if (second) *second = b->hp.n;
	return CT_OK;
}

int ct_box_push_rungs(ct_box_t b) {
	if (!b) return fail(CT_E_ARG, "ct_box_push_rungs: null box");
	b->rung_stack.push_back(b->rung_begin);
	return CT_OK;
}

int ct_box_pop_rungs(ct_box_t b) {
	if (!b) return fail(CT_E_ARG, "ct_box_pop_rungs: null box");
	if (b->rung_stack.empty()) return fail(CT_E_STATE, "ct_box_pop_rungs: the range stack is empty");
	b->rung_begin = b->rung_stack.back();
	b->rung_stack.pop_back();
	return CT_OK;
}

int ct_box_set_minrung(ct_box_t b, int min_rung) {
	int rc = check_init();
	if (rc) return rc;
	if (!b || min_rung < 0 || min_rung > 127) return fail(CT_E_ARG, "ct_box_set_minrung: bad arguments");
	const int64_t n = b->hp.n;
	if (n == 0) return CT_OK;
	hipLaunchKernelGGL(k_set_minrung, dim3((n + 255) / 256), dim3(256), 0, 0, b->rung.p, n, min_rung);
	CT_HIP(hipGetLastError());
	CT_HIP(hipStreamSynchronize(0));
	b->host_stale = true;
	return CT_OK;
}

int ct_box_rung_counts(ct_box_t b, int64_t* counts, int cap, int* nrungs) {
	int rc = check_init();
	if (rc) return rc;
	if (!b || cap < 0 || (cap > 0 && !counts)) return fail(CT_E_ARG, "ct_box_rung_counts: bad arguments");
	const int64_t n = b->hp.n;
	unsigned long long h[128] = {0};
	if (n > 0) {
		DBuf<unsigned long long> d;
		if ((rc = d.grow(128))) return rc;
		hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(128), 0, 0, d.p, 128);
		const int nb = int(std::min<int64_t>((n + 255) / 256, 4096));
		hipLaunchKernelGGL(k_rung_hist, dim3(nb), dim3(256), 0, 0, b->rung.p, n, d.p);
		const hipError_t e = hipMemcpy(h, d.p, sizeof(h), hipMemcpyDeviceToHost);
		d.release();
		if (e != hipSuccess) return fail(CT_E_HIP, std::string("ct_box_rung_counts: ") + hipGetErrorString(e));
	}
	int top = 0;
	for (int r = 0; r < 128; r++)
		if (h[r]) top = r + 1;
	for (int r = 0; r < cap; r++) counts[r] = r < 128 ? int64_t(h[r]) : 0;
	if (nrungs) *nrungs = top;
	return CT_OK;
}

uint64_t ct_time_inc(uint64_t t, int max_rung) {
	if (max_rung > 0) t += uint64_t(1) << uint64_t(64 - max_rung);
	return t;
}

int ct_time_min_rung(uint64_t t) {
	int r = 64;
	while ((t & 1) == 0 && r != 0) {
		r--;
		t >>= 1;
	}
	return r;
}

// The kick ladder of one sub-step (src/driver.cpp:382-593), step for step:
// the list of levels, the ascending / top / descending flags, the rung sorts
// and range pushes and pops, the top-rung clipping test on the rung counts,
// the kick_params of each kick (driver.cpp:481-508), the levels appended
// after each descending kick and the drifts.
int ct_step_hierarchical(ct_box_t b, const ct_step_params* sp, ct_step_return* out, void* stream) {
	int rc = check_init();
	if (rc) return rc;
	if (!b || !sp || !out || sp->minrung < 0 || sp->minrung > 63 || sp->max_rung < 0 || sp->max_rung > 63 || sp->minrung0 < 0 ||
			!(sp->t0 > 0.0) || !(sp->scale > 0.0))
		return fail(CT_E_ARG, "ct_step_hierarchical: bad arguments");
	int minrung = sp->minrung;
	const int minrung0 = sp->minrung0;
	int max_rung = sp->max_rung;
	std::vector<int> levels;
	for (int i = max_rung; i > minrung; i--) levels.push_back(i);
	levels.push_back(minrung);
	bool ascending = true;
	bool top = false;
	bool clip_top = false;
	ct_step_return R;
	std::memset(&R, 0, sizeof(R));
	ct_kick_return kr;
	std::memset(&kr, 0, sizeof(kr));
	for (size_t li = 0; li < levels.size(); li++) {
		const int level = levels[li];
		if (level == minrung) {
			ascending = false;
			top = true;
		} else {
			top = false;
		}
		if (!ascending && !top && (rc = ct_box_push_rungs(b))) return rc;
		if (!ascending && (rc = ct_box_sort_by_rung(b, level, nullptr))) return rc;
		R.parts_processed += double(b->hp.n - b->rung_begin);
		if (top && minrung == minrung0) {
			int64_t counts[128];
			int nr = 0;
			if ((rc = ct_box_rung_counts(b, counts, 128, &nr))) return rc;
			if (nr > minrung0 + 1) {
				int64_t fast = 0;
				const int64_t slow = counts[minrung0];
				for (int i = minrung0 + 1; i < nr; i++) fast += counts[i];
				if (3 * fast > slow) clip_top = !sp->nocrop;
			}
		}
		if ((rc = ct_tree_build_ex(b, sp->theta, sp->bucket, sp->hsoft, sp->tree_where))) return rc;
		ct_kick_params kp;
		std::memset(&kp, 0, sizeof(kp));
		kp.top = (clip_top && level == minrung0 + 1) ? true : top;
		kp.sign = float(sp->sign);
		kp.ascending = ascending || top;
		kp.descending = (clip_top && top) ? false : (!ascending || top);
		kp.vel_init = false;
		kp.node_load = 0.0f;
		kp.gpu = true;
		kp.min_level = 9;
		kp.save_force = sp->save_force != 0;
		kp.GM = float(sp->GM);
		kp.eta = float(sp->eta);
		kp.h = float(sp->hsoft);
		kp.a = float(sp->scale);
		kp.first_call = sp->tau == 0.0;
		kp.min_rung = level;
		kp.do_phi = sp->do_phi && top;
		kp.t0 = float(sp->t0);
		kp.max_dt = float(sp->max_dt);
		kp.theta = float(sp->theta);
		ct_kick_return tkr;
		if ((rc = ct_kick_prec(b, &kp, sp->precision, &tkr, stream))) return rc;
		R.nkicks++;
		R.tckin += tkr.dkin;
		if (clip_top && top && (rc = ct_box_set_minrung(b, minrung0 + 1))) return rc;
		if (top) kr = tkr;
		if (clip_top && top) {
			levels.push_back(minrung0 + 1);
		} else if (!ascending || top) {
			max_rung = tkr.max_rung;
			kr.max_rung = max_rung;
			if (max_rung > level) levels.push_back(level + 1);
		}
		if ((!ascending || top) && !(clip_top && top)) {
			const double dt = std::ldexp(1.0, -level) * sp->t0;  // rung_dt[level] * t0
			if ((rc = ct_drift(b, level, float(sp->scale), float(dt), sp->drag, stream))) return rc;
		}
		if (ascending && !top && (rc = ct_box_pop_rungs(b))) return rc;
	}
	if (clip_top) minrung++;
	R.kr = kr;
	R.minrung = minrung;
	R.clip_top = clip_top ? 1 : 0;
	*out = R;
	return CT_OK;
}

// FP64 direct-summation periodic Ewald forces on the device: the force check
// of the kick.  Replaces the reference's analytic kernel
// (gravity_analytic_call_kernel, src/analytic.cu:30-116, kernel :119-234),
// which sums each pair's Ewald series in FP32 and passes gsoft's arguments in
// the wrong order; this one is FP64 throughout.
//
// The sum is split with alpha = 9 (box units):
//   real space:  erfc(alpha r)/r over the minimum image only, r < 1/2
//                (erfc(4.5)/0.5 ~ 4e-10 is the largest dropped term);
//   Fourier:     2 exp(-pi^2 h^2/alpha^2)/(pi h^2) cos(2 pi h.(x_s - x_j)) over
//                the half space of integer h with 0 < h^2 <= 220 (the largest
//                dropped coefficient is ~2e-12), through structure factors
//                S(h) = sum_j exp(2 pi i h.x_j);
//   constants:   the r -> 0 limit -2 alpha/sqrt(pi) of the sink's own image and
//                the neutralising background -n pi/alpha^2.
// Pairs closer than h are softened with the Ferrers n=4 kernel
// (cosmictiger/gravity.hpp:76-96).  Phases h.x are formed exactly in
// fixed32 integer arithmetic (mod 2^32) before the sine and cosine.
// phi = -psi, g = -grad(phi) of the unit-mass sum; the caller multiplies by
// GM as analytic_compare does (src/analytic.cpp:75-88).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ct_host.hpp"

namespace ct {
namespace {

constexpr double kDeAlpha = 9.0;
constexpr double kDeRc2 = 0.25;
constexpr int kDeH2Max = 220;
constexpr int kDeSinks = 8;       // sinks per block of the real-space kernel
constexpr int kDeTile = 1024;     // sources staged per LDS tile

struct Wave {
	int32_t h[3];
	int32_t pad;
	double coef;  // 2 exp(-pi^2 h^2 / alpha^2) / (pi h^2)
};

std::vector<Wave> make_waves() {
	std::vector<Wave> w;
	const int hm = int(std::sqrt(double(kDeH2Max))) + 1;
	for (int i = -hm; i <= hm; i++)
		for (int j = -hm; j <= hm; j++)
			for (int k = -hm; k <= hm; k++) {
				const int h2 = i * i + j * j + k * k;
				if (h2 == 0 || h2 > kDeH2Max) continue;
				if (i < 0 || (i == 0 && j < 0) || (i == 0 && j == 0 && k < 0)) continue;
				Wave v;
				v.h[0] = i;
				v.h[1] = j;
				v.h[2] = k;
				v.pad = 0;
				v.coef = 2.0 * std::exp(-M_PI * M_PI * h2 / (kDeAlpha * kDeAlpha)) / (M_PI * h2);
				w.push_back(v);
			}
	return w;
}

// exact phase h.x mod 1 as a signed fraction of 2^32, returned as a multiple of pi
__device__ __forceinline__ double phase_pi(const Wave& w, uint32_t x, uint32_t y, uint32_t z) {
	const uint32_t p = uint32_t(w.h[0]) * x + uint32_t(w.h[1]) * y + uint32_t(w.h[2]) * z;
	return double(int32_t(p)) * (1.0 / 2147483648.0);
}

// S(h) over a slice of the sources; one thread per wave, sources staged in LDS
__global__ __launch_bounds__(256) void k_de_structure(const uint32_t* __restrict__ x, const uint32_t* __restrict__ y,
		const uint32_t* __restrict__ z, int64_t n, const Wave* __restrict__ waves, int nw, double* __restrict__ S) {
	__shared__ uint32_t tx[kDeTile], ty[kDeTile], tz[kDeTile];
	const int w = blockIdx.x * blockDim.x + threadIdx.x;
	Wave wv;
	wv.h[0] = wv.h[1] = wv.h[2] = 0;
	if (w < nw) wv = waves[w];
	const int64_t chunk = (n + gridDim.y - 1) / gridDim.y;
	const int64_t b = int64_t(blockIdx.y) * chunk;
	const int64_t e = b + chunk < n ? b + chunk : n;
	double sc = 0.0, ss = 0.0;
	for (int64_t t0 = b; t0 < e; t0 += kDeTile) {
		const int m = int(e - t0 < kDeTile ? e - t0 : kDeTile);
		__syncthreads();
		for (int q = threadIdx.x; q < m; q += blockDim.x) {
			tx[q] = x[t0 + q];
			ty[q] = y[t0 + q];
			tz[q] = z[t0 + q];
		}
		__syncthreads();
		if (w < nw) {
			for (int q = 0; q < m; q++) {
				double s, c;
				sincospi(phase_pi(wv, tx[q], ty[q], tz[q]), &s, &c);
				sc += c;
				ss += s;
			}
		}
	}
	if (w < nw && e > b) {
		atomicAdd(&S[2 * w], sc);
		atomicAdd(&S[2 * w + 1], ss);
	}
}

// real-space part and softening: kDeSinks sinks per block, sources split over
// blockIdx.y; out[4 s + {0,1,2,3}] += {gx, gy, gz, phi}
__global__ __launch_bounds__(256) void k_de_real(const uint32_t* __restrict__ x, const uint32_t* __restrict__ y, const uint32_t* __restrict__ z,
		int64_t n, const uint32_t* __restrict__ sx, const uint32_t* __restrict__ sy, const uint32_t* __restrict__ sz, int nsink, double h,
		double* __restrict__ out) {
	__shared__ double red[4][kDeSinks * 4];
	const int s0 = blockIdx.x * kDeSinks;
	uint32_t X[kDeSinks], Y[kDeSinks], Z[kDeSinks];
	double acc[kDeSinks][4];
#pragma unroll
	for (int s = 0; s < kDeSinks; s++) {
		const int si = s0 + s < nsink ? s0 + s : nsink - 1;
		X[s] = sx[si];
		Y[s] = sy[si];
		Z[s] = sz[si];
		acc[s][0] = acc[s][1] = acc[s][2] = acc[s][3] = 0.0;
	}
	const double a = kDeAlpha;
	const double twoa_sqrtpi = 2.0 * kDeAlpha / 1.7724538509055160273;
	const double h2 = h * h, hinv = 1.0 / h, h2inv = hinv * hinv, h3inv = h2inv * hinv;
	const int64_t chunk = (n + gridDim.y - 1) / gridDim.y;
	const int64_t b = int64_t(blockIdx.y) * chunk;
	const int64_t e = b + chunk < n ? b + chunk : n;
	for (int64_t j = b + threadIdx.x; j < e; j += blockDim.x) {
		const uint32_t xj = x[j], yj = y[j], zj = z[j];
#pragma unroll
		for (int s = 0; s < kDeSinks; s++) {
			const double dx = double(int32_t(X[s] - xj)) * (1.0 / 4294967296.0);
			const double dy = double(int32_t(Y[s] - yj)) * (1.0 / 4294967296.0);
			const double dz = double(int32_t(Z[s] - zj)) * (1.0 / 4294967296.0);
			const double r2 = dx * dx + dy * dy + dz * dz;
			if (r2 >= kDeRc2 || r2 == 0.0) continue;
			const double r = sqrt(r2);
			const double er = erfc(a * r) / r;
			double dr = -(er + twoa_sqrtpi * exp(-a * a * r2)) / r2;
			double ph = -er;
			if (r2 < h2) {
				const double q2 = r2 * h2inv;
				double f = 315.0 / 128.0;
				f = f * q2 - 385.0 / 32.0;
				f = f * q2 + 1485.0 / 64.0;
				f = f * q2 - 693.0 / 32.0;
				f = f * q2 + 1155.0 / 128.0;
				double p = -63.0 / 256.0;
				p = p * q2 + 385.0 / 256.0;
				p = p * q2 - 495.0 / 128.0;
				p = p * q2 + 693.0 / 128.0;
				p = p * q2 - 1155.0 / 256.0;
				p = p * q2 + 693.0 / 256.0;
				ph -= p * hinv - 1.0 / r;
				dr -= f * h3inv - 1.0 / (r2 * r);
			}
			acc[s][0] += dr * dx;
			acc[s][1] += dr * dy;
			acc[s][2] += dr * dz;
			acc[s][3] += ph;
		}
	}
	// block reduction: wave butterflies, then the 4 waves through LDS
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
	for (int s = 0; s < kDeSinks; s++) {
#pragma unroll
		for (int c = 0; c < 4; c++) {
			double v = acc[s][c];
			for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
			if (lane == 0) red[wid][s * 4 + c] = v;
		}
	}
	__syncthreads();
	if (threadIdx.x < kDeSinks * 4) {
		const int s = threadIdx.x >> 2;
		const double v = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
		if (s0 + s < nsink) atomicAdd(&out[4 * (s0 + s) + (threadIdx.x & 3)], v);
	}
}

// Fourier part and constants, one block per sink, threads over waves
__global__ __launch_bounds__(256) void k_de_fourier(const uint32_t* __restrict__ sx, const uint32_t* __restrict__ sy,
		const uint32_t* __restrict__ sz, int nsink, const Wave* __restrict__ waves, int nw, const double* __restrict__ S, int64_t n,
		double* __restrict__ out) {
	__shared__ double red[4][4];
	const int s = blockIdx.x;
	if (s >= nsink) return;
	const uint32_t X = sx[s], Y = sy[s], Z = sz[s];
	double psi = 0.0, g[3] = {0.0, 0.0, 0.0};
	for (int w = threadIdx.x; w < nw; w += blockDim.x) {
		const Wave wv = waves[w];
		double st, ct;
		sincospi(phase_pi(wv, X, Y, Z), &st, &ct);
		const double Sc = S[2 * w], Ss = S[2 * w + 1];
		psi += wv.coef * (ct * Sc + st * Ss);
		const double gd = wv.coef * (ct * Ss - st * Sc) * 6.283185307179586477;
		g[0] += gd * wv.h[0];
		g[1] += gd * wv.h[1];
		g[2] += gd * wv.h[2];
	}
	double v[4] = {g[0], g[1], g[2], -psi};
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
	for (int c = 0; c < 4; c++) {
		for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
		if (lane == 0) red[wid][c] = v[c];
	}
	__syncthreads();
	if (threadIdx.x < 4) {
		const int c = threadIdx.x;
		double t = red[0][c] + red[1][c] + red[2][c] + red[3][c];
		if (c == 3) t += 2.0 * kDeAlpha / 1.7724538509055160273 + double(n) * M_PI / (kDeAlpha * kDeAlpha);
		out[4 * s + c] += t;
	}
}

}  // namespace
}  // namespace ct

using namespace ct;

extern "C" int ct_direct_ewald(ct_box_t b, int64_t nsink, const uint32_t* sinkx, const uint32_t* sinky, const uint32_t* sinkz, double h,
		double* phi, double* gx, double* gy, double* gz, void* stream) {
	int rc = check_init();
	if (rc) return rc;
	if (!b || nsink < 0 || nsink > (1 << 24) || !(h > 0.0) || (nsink > 0 && (!sinkx || !sinky || !sinkz || !phi || !gx || !gy || !gz)))
		return fail(CT_E_ARG, "ct_direct_ewald: bad arguments");
	if (nsink == 0) return CT_OK;
	const int64_t n = b->hp.n;
	hipStream_t st = reinterpret_cast<hipStream_t>(stream);
	const std::vector<Wave> waves = make_waves();
	const int nw = int(waves.size());
	DBuf<Wave> dw;
	DBuf<double> S, out;
	DBuf<uint32_t> s3;
	if ((rc = dw.grow(nw)) || (rc = S.grow(2 * int64_t(nw))) || (rc = out.grow(4 * nsink)) || (rc = s3.grow(3 * nsink))) {
		dw.release();
		S.release();
		out.release();
		s3.release();
		return rc;
	}
	auto cleanup = [&]() {
		dw.release();
		S.release();
		out.release();
		s3.release();
	};
	hipError_t e = hipSuccess;
	auto chk = [&](hipError_t r) {
		if (e == hipSuccess) e = r;
	};
	chk(hipMemcpyAsync(dw.p, waves.data(), nw * sizeof(Wave), hipMemcpyHostToDevice, st));
	chk(hipMemcpyAsync(s3.p, sinkx, nsink * 4, hipMemcpyHostToDevice, st));
	chk(hipMemcpyAsync(s3.p + nsink, sinky, nsink * 4, hipMemcpyHostToDevice, st));
	chk(hipMemcpyAsync(s3.p + 2 * nsink, sinkz, nsink * 4, hipMemcpyHostToDevice, st));
	chk(hipMemsetAsync(S.p, 0, 2 * nw * sizeof(double), st));
	chk(hipMemsetAsync(out.p, 0, 4 * nsink * sizeof(double), st));
	if (e == hipSuccess && n > 0) {
		const int wblocks = (nw + 255) / 256;
		const int pslices = int(std::min<int64_t>(64, (n + kDeTile - 1) / kDeTile));
		hipLaunchKernelGGL(k_de_structure, dim3(wblocks, pslices), dim3(256), 0, st, b->x.p, b->y.p, b->z.p, n, dw.p, nw, S.p);
		const int sblocks = int((nsink + kDeSinks - 1) / kDeSinks);
		const int rslices = int(std::max<int64_t>(1, std::min<int64_t>((2048 + sblocks - 1) / sblocks, (n + 4095) / 4096)));
		hipLaunchKernelGGL(k_de_real, dim3(sblocks, rslices), dim3(256), 0, st, b->x.p, b->y.p, b->z.p, n, s3.p, s3.p + nsink, s3.p + 2 * nsink,
				int(nsink), h, out.p);
		hipLaunchKernelGGL(k_de_fourier, dim3(int(nsink)), dim3(256), 0, st, s3.p, s3.p + nsink, s3.p + 2 * nsink, int(nsink), dw.p, nw, S.p, n,
				out.p);
		chk(hipGetLastError());
	}
	std::vector<double> host(4 * nsink);
	chk(hipMemcpyAsync(host.data(), out.p, 4 * nsink * sizeof(double), hipMemcpyDeviceToHost, st));
	chk(hipStreamSynchronize(st));
	cleanup();
	if (e != hipSuccess) return fail(CT_E_HIP, std::string("ct_direct_ewald: ") + hipGetErrorString(e));
	for (int64_t s = 0; s < nsink; s++) {
		gx[s] = host[4 * s + 0];
		gy[s] = host[4 * s + 1];
		gz[s] = host[4 * s + 2];
		phi[s] = host[4 * s + 3];
	}
	return CT_OK;
}

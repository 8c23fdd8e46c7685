// Host construction of the Ewald tables shared by the device kick (uploaded
// to constant memory at ct_init) and host code that evaluates the same
// Ewald correction (ctfmm::ewald_greens).
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace ct {

struct EwaldHost {
	std::vector<float> rimg, fh, fc;
	std::vector<double> fc64;  // the Fourier coefficients in double (FP64 kicks)
	int nreal = 0, nfour = 0;
};

// Ewald tables from the reference's rules (src/ewald_indices.cu:30-74):
// real images n != 0 with "box distance" |max(|n|-1/2, 0)| < 2.6, Fourier
// vectors 0 < |h|^2 <= 8; the Fourier part is folded to one of each +-h pair.
inline EwaldHost make_ewald_tables() {
	EwaldHost E;
	const double a = 2.0;
	std::vector<std::array<int, 3>> real;
	for (int i = -4; i <= 4; i++)
		for (int j = -4; j <= 4; j++)
			for (int k = -4; k <= 4; k++) {
				const double x = std::max(std::abs(i) - 0.5, 0.0);
				const double y = std::max(std::abs(j) - 0.5, 0.0);
				const double z = std::max(std::abs(k) - 0.5, 0.0);
				if (std::sqrt(x * x + y * y + z * z) < 2.6 && (i || j || k)) real.push_back({i, j, k});
			}
	// far images first (smallest terms summed first)
	std::stable_sort(real.begin(), real.end(), [](const std::array<int, 3>& p, const std::array<int, 3>& q) {
		return p[0] * p[0] + p[1] * p[1] + p[2] * p[2] > q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
	});
	for (auto& r : real) {
		for (int d = 0; d < 3; d++) E.rimg.push_back(float(r[d]));
	}
	E.nreal = int(real.size());
	std::vector<std::array<int, 3>> four;
	for (int i = -3; i <= 3; i++)
		for (int j = -3; j <= 3; j++)
			for (int k = -3; k <= 3; k++) {
				const int h2 = i * i + j * j + k * k;
				if (h2 == 0 || h2 > 8) continue;
				if (i < 0 || (i == 0 && j < 0) || (i == 0 && j == 0 && k < 0)) continue;
				four.push_back({i, j, k});
			}
	std::stable_sort(four.begin(), four.end(), [](const std::array<int, 3>& p, const std::array<int, 3>& q) {
		return p[0] * p[0] + p[1] * p[1] + p[2] * p[2] > q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
	});
	E.nfour = int(four.size());
	const double sgn4[4] = {1.0, -1.0, -1.0, 1.0};
	for (auto& h : four) {
		for (int d = 0; d < 3; d++) E.fh.push_back(float(h[d]));
		const double h2 = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
		const double c = std::exp(-M_PI * M_PI * h2 / (a * a)) / (M_PI * h2);
		std::vector<double> row(50, 0.0);
		for (int l = 0; l < 7; l++)
			for (int p = l; p >= 0; p--)
				for (int q = l - p; q >= 0; q--) {
					const int r = l - p - q;
					if (!(r <= 1 || (p == 0 && q == 0 && r == 2))) continue;
					const int idx = (p + q) * (p + q + 1) / 2 + q + 28 * (r == 1) + 49 * (r == 2);
					const double hn = std::pow(double(h[0]), p) * std::pow(double(h[1]), q) * std::pow(double(h[2]), r);
					row[idx] = -2.0 * c * std::pow(2.0 * M_PI, l) * hn * sgn4[l & 3];
				}
		for (double v : row) E.fc.push_back(float(v));
		E.fc64.insert(E.fc64.end(), row.begin(), row.end());
	}
	return E;
}

}  // namespace ct

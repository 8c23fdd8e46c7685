/*
 * CPU oracle for the CosmicTiger gravity kick -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, the smoke check and bench.py's CPU-baseline leg may load this
 * library; the product path never calls it.
 *
 * Plain C (C11), FP64 for all expansion and interaction arithmetic, explicit
 * float arithmetic where the reference decides integer outcomes in FP32 (the
 * multipole acceptance tests of src/kick.cu:286-386 and the box/sphere test of
 * cosmictiger/range.hpp:455-466).  Compiled with -ffp-contract=off so that
 * every fused multiply-add below is one that is written out (fma/fmaf).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   fixed32 / distance / drift ......... closed forms and special cases (wrap
 *                                          at 0 and 1, differences across 2^31,
 *                                          truncation of 2^32 x).
 *   HEALPix nested pixels ................ pix2vec/vec2pix round trip of every
 *                                          pixel, the 12 base-pixel centres,
 *                                          the nested hierarchy (pixel at nside
 *                                          = pixel at 2 nside / 4), equal
 *                                          areas (counts of uniform directions).
 *   lightcone crossing ................... each entry lies on the light shell
 *                                          (|X(t)| = t_i - tau0 - t), a radial
 *                                          mover against its closed form, and
 *                                          crossings counted by a bisection of
 *                                          the shell condition.
 *   gsoft / self_phi ..................... closed form: f(1)=phi(1)=1, f=-2P',
 *                                          Ferrers n=4 density integrates to 1,
 *                                          self_phi = 693/256.
 *   greens ............................... sympy derivatives of -1/r.
 *   trless storage / M2L / L2L / L2P ..... exact translation identities and
 *                                          convergence to direct 1/r sums.
 *   P2M / M2M ............................ P2M at c1 then M2M to c2 equals P2M
 *                                          at c2; tracelessness; far field.
 *   Ewald tables ......................... counts 178 and 92 from the rules.
 *   ewald_greens ......................... sympy/finite differences of the
 *                                          FP64 Ewald sum at two alphas; the
 *                                          cubic-lattice constant -2.8372974795.
 *   tree build ........................... invariants (partition, bucket,
 *                                          Ewald size rule, radius covers
 *                                          members, boxes contain members).
 *   walk ................................. every (sink, source) pair covered
 *                                          exactly once (mass count = N).
 *   kick forces .......................... FP64 direct Ewald below.
 *   direct Ewald ......................... alpha independence, lattice
 *                                          constant, lattice zero force,
 *                                          translation invariance, two-body
 *                                          closed form at small separation.
 * The reference ships no golden vectors or fixtures for any of these.
 */
#include "ct_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORDER 7
#define MORDER 6
#define NEXP 50
#define NMULT 37
#define SCALE 1000.0
#define EWALD_ALPHA 2.0
#define EWALD_REAL_CUTOFF 2.6
#define MIN_PARTS_PCCP 38 /* src/kick.cu:32 */
#define MAX_DEPTH 128     /* cosmictiger/defs.hpp:95 (tree depth limit, src/tree.cpp:285) */

/* ------------------------------------------------------------------ */
/* multi-index helpers                                                  */
/* ------------------------------------------------------------------ */

static double g_fact[32];
static int g_init = 0;

static double dfact(int n) {
	double r = 1;
	while (n > 1) {
		r *= n;
		n -= 2;
	}
	return r;
}

static double mfact(const int n[3]) {
	return g_fact[n[0]] * g_fact[n[1]] * g_fact[n[2]];
}

int oc_trless_index(int a, int b, int c, int P) {
	return (a + b) * (a + b + 1) / 2 + b + (P * (P + 1) / 2) * (c == 1) + P * P * (c == 2);
}

/* full component of a stored trless tensor (cosmictiger/tensor.hpp:98-118) */
static double full_rec(const double* T, int P, int a, int b, int c) {
	if (c >= 2 && !(a == 0 && b == 0 && c == 2)) {
		return -(full_rec(T, P, a + 2, b, c - 2) + full_rec(T, P, a, b + 2, c - 2));
	}
	return T[oc_trless_index(a, b, c, P)];
}

double oc_full(const double* T, int P, int a, int b, int c) {
	return full_rec(T, P, a, b, c);
}

/* expand[P][a][b][c]: the full component as a short list of (stored index, coefficient) */
typedef struct {
	int n;
	int idx[16];
	double coef[16];
} full_expr_t;
static full_expr_t g_fx7[8][8][8], g_fx6[8][8][8];

static void build_full_exprs(int P, full_expr_t tab[8][8][8]) {
	double unit[64];
	for (int a = 0; a < P; a++)
		for (int b = 0; a + b < P; b++)
			for (int c = 0; a + b + c < P; c++) {
				full_expr_t* e = &tab[a][b][c];
				e->n = 0;
				for (int k = 0; k < P * P + 1; k++) {
					for (int q = 0; q < P * P + 1; q++)
						unit[q] = (q == k) ? 1.0 : 0.0;
					const double v = full_rec(unit, P, a, b, c);
					if (v != 0.0) {
						e->idx[e->n] = k;
						e->coef[e->n] = v;
						e->n++;
					}
				}
			}
}

static inline double full_fast(const full_expr_t tab[8][8][8], const double* T, int a, int b, int c) {
	const full_expr_t* e = &tab[a][b][c];
	double s = 0.0;
	for (int i = 0; i < e->n; i++)
		s += e->coef[i] * T[e->idx[i]];
	return s;
}

/* monomials x^n for all n with |n| <= 6 (index [a][b][c]) */
typedef struct {
	double v[8][8][8];
} mono_t;

static void monomials(mono_t* m, const double x[3], int maxdeg) {
	double px[8], py[8], pz[8];
	px[0] = py[0] = pz[0] = 1.0;
	for (int i = 1; i <= maxdeg; i++) {
		px[i] = px[i - 1] * x[0];
		py[i] = py[i - 1] * x[1];
		pz[i] = pz[i - 1] * x[2];
	}
	for (int a = 0; a <= maxdeg; a++)
		for (int b = 0; a + b <= maxdeg; b++)
			for (int c = 0; a + b + c <= maxdeg; c++)
				m->v[a][b][c] = px[a] * py[b] * pz[c];
}

/* n!/(2^|m| m! (n-2m)!) */
static double radial_coef(const int n[3], const int m[3]) {
	int d[3] = { n[0] - 2 * m[0], n[1] - 2 * m[1], n[2] - 2 * m[2] };
	return mfact(n) / (ldexp(1.0, m[0] + m[1] + m[2]) * mfact(m) * mfact(d));
}

/* traceless-projection coefficient (-1)^|m| (2l-2|m|-1)!!/(2l-1)!! * radial_coef */
static double tf_coef(const int n[3], const int m[3]) {
	const int l = n[0] + n[1] + n[2];
	const int j = m[0] + m[1] + m[2];
	const double s = (j & 1) ? -1.0 : 1.0;
	return s * dfact(2 * l - 2 * j - 1) / dfact(2 * l - 1) * radial_coef(n, m);
}

/* d^n g(|x|) from the radial derivative functions gk[k] = (1/r d/dr)^k g */
typedef struct {
	int nt;
	int mono[12][3];
	int k[12];
	double coef[12];
} radial_terms_t;
static radial_terms_t g_rt7[64]; /* by position in the trless-7 list */

static void build_radial_terms(const int n[3], radial_terms_t* t) {
	const int l = n[0] + n[1] + n[2];
	int m[3];
	t->nt = 0;
	for (m[0] = 0; 2 * m[0] <= n[0]; m[0]++)
		for (m[1] = 0; 2 * m[1] <= n[1]; m[1]++)
			for (m[2] = 0; 2 * m[2] <= n[2]; m[2]++) {
				const int j = m[0] + m[1] + m[2];
				t->mono[t->nt][0] = n[0] - 2 * m[0];
				t->mono[t->nt][1] = n[1] - 2 * m[1];
				t->mono[t->nt][2] = n[2] - 2 * m[2];
				t->k[t->nt] = l - j;
				t->coef[t->nt] = radial_coef(n, m);
				t->nt++;
			}
}

static inline double radial_fast(const radial_terms_t* t, const mono_t* x, const double* gk) {
	double sum = 0.0;
	for (int i = 0; i < t->nt; i++)
		sum += t->coef[i] * x->v[t->mono[i][0]][t->mono[i][1]][t->mono[i][2]] * gk[t->k[i]];
	return sum;
}

static double radial_derivative(const int n[3], const mono_t* x, const double* gk) {
	radial_terms_t t;
	build_radial_terms(n, &t);
	return radial_fast(&t, x, gk);
}

/* iterate over the trless set of order P: callback style via arrays */
static int trless_list(int P, int out[][3]) {
	int cnt = 0;
	for (int a = 0; a < P; a++)
		for (int b = 0; a + b < P; b++)
			for (int c = 0; a + b + c < P; c++) {
				if (c <= 1 || (a == 0 && b == 0 && c == 2)) {
					out[cnt][0] = a;
					out[cnt][1] = b;
					out[cnt][2] = c;
					cnt++;
				}
			}
	return cnt;
}

static int g_tl7[64][3], g_n7;
static int g_tl6[64][3], g_n6;
static int g_tl2[8][3], g_n2;

/* ------------------------------------------------------------------ */
/* Ewald tables (src/ewald_indices.cu:30-74)                            */
/* ------------------------------------------------------------------ */

static double g_real[256][3];
static int g_nreal = 0;
static double g_four[128][3];
static int g_nfour = 0;

static int cmp_desc_norm(const void* pa, const void* pb) {
	const double* a = (const double*) pa;
	const double* b = (const double*) pb;
	const double a2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
	const double b2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
	if (a2 > b2) return -1;
	if (a2 < b2) return 1;
	for (int d = 0; d < 3; d++) {
		if (a[d] < b[d]) return -1;
		if (a[d] > b[d]) return 1;
	}
	return 0;
}

static void ewald_tables(void) {
	const int nmax = (int) sqrt(12.0) + 1;
	g_nreal = 0;
	for (int i = -nmax; i <= nmax; i++)
		for (int j = -nmax; j <= nmax; j++)
			for (int k = -nmax; k <= nmax; k++) {
				const int i2 = i * i + j * j + k * k;
				const double x = fmax(abs(i) - 0.5, 0.0);
				const double y = fmax(abs(j) - 0.5, 0.0);
				const double z = fmax(abs(k) - 0.5, 0.0);
				const double d = sqrt(x * x + y * y + z * z);
				if (d < EWALD_REAL_CUTOFF && i2 > 0) {
					g_real[g_nreal][0] = i;
					g_real[g_nreal][1] = j;
					g_real[g_nreal][2] = k;
					g_nreal++;
				}
			}
	qsort(g_real, g_nreal, sizeof(g_real[0]), cmp_desc_norm);
	const int n2max = 8;
	const int fmax_ = (int) sqrt((double) n2max) + 1;
	g_nfour = 0;
	for (int i = -fmax_; i <= fmax_; i++)
		for (int j = -fmax_; j <= fmax_; j++)
			for (int k = -fmax_; k <= fmax_; k++) {
				const int h2 = i * i + j * j + k * k;
				if (h2 <= n2max && h2 > 0) {
					g_four[g_nfour][0] = i;
					g_four[g_nfour][1] = j;
					g_four[g_nfour][2] = k;
					g_nfour++;
				}
			}
	qsort(g_four, g_nfour, sizeof(g_four[0]), cmp_desc_norm);
}

int oc_ewald_nreal(void) {
	oc_init();
	return g_nreal;
}
int oc_ewald_nfour(void) {
	oc_init();
	return g_nfour;
}
void oc_ewald_real_index(int i, double out[3]) {
	oc_init();
	for (int d = 0; d < 3; d++)
		out[d] = g_real[i][d];
}
void oc_ewald_four_index(int i, double out[3]) {
	oc_init();
	for (int d = 0; d < 3; d++)
		out[d] = g_four[i][d];
}

void oc_init(void) {
	if (g_init) return;
	g_fact[0] = 1.0;
	for (int i = 1; i < 32; i++)
		g_fact[i] = g_fact[i - 1] * i;
	g_n7 = trless_list(7, g_tl7);
	g_n6 = trless_list(6, g_tl6);
	g_n2 = trless_list(2, g_tl2);
	build_full_exprs(7, g_fx7);
	for (int i = 0; i < g_n7; i++)
		build_radial_terms(g_tl7[i], &g_rt7[i]);
	build_full_exprs(6, g_fx6);
	ewald_tables();
	g_init = 1;
}

/* ------------------------------------------------------------------ */
/* expansions                                                           */
/* ------------------------------------------------------------------ */

/* D_n = d^n(-1/r), X scaled by 1000 (cosmictiger/fmm_kernels.hpp:52-308) */
void oc_greens(double D[50], const double X[3]) {
	oc_init();
	double x[3] = { X[0] * SCALE, X[1] * SCALE, X[2] * SCALE };
	const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
	if (r2 == 0.0) {
		for (int i = 0; i < NEXP; i++)
			D[i] = 0.0;
		return;
	}
	const double r = sqrt(r2);
	double gk[ORDER];
	/* g = -1/r: g_k = -(-1)^k (2k-1)!! r^-(2k+1) */
	for (int k = 0; k < ORDER; k++) {
		gk[k] = -((k & 1) ? -1.0 : 1.0) * dfact(2 * k - 1) / pow(r, 2 * k + 1);
	}
	mono_t m;
	monomials(&m, x, ORDER - 1);
	for (int i = 0; i < g_n7; i++) {
		D[oc_trless_index(g_tl7[i][0], g_tl7[i][1], g_tl7[i][2], ORDER)] = radial_fast(&g_rt7[i], &m, gk);
	}
}

/* radial functions of erf(a r)/r: series for small r, closed form above */
static void erf_over_r_derivs(double r, double a, double* hk) {
	if (r < 0.3) {
		const double s = r * r;
		for (int k = 0; k < ORDER; k++) {
			/* h_k = sum_{j>=k} a_j 2^k j!/(j-k)! s^(j-k),
			 * a_j = 2a/sqrt(pi) (-1)^j a^(2j) / (j! (2j+1)) */
			/* evaluate with recurrence on terms to avoid huge factorials */
			double term = 2.0 * a / sqrt(M_PI); /* a_0 */
			/* a_j / a_{j-1} = -a^2 (2j-1) / (j (2j+1)) */
			double aj = term;
			double acc = 0.0;
			for (int j = 0; j < 80; j++) {
				if (j > 0) aj *= -a * a * (2.0 * j - 1.0) / (j * (2.0 * j + 1.0));
				if (j >= k) {
					/* 2^k j!/(j-k)! */
					double c = ldexp(1.0, k);
					for (int q = 0; q < k; q++)
						c *= (double) (j - q);
					acc += aj * c * pow(s, j - k);
				}
			}
			hk[k] = acc;
		}
	} else {
		/* erf(ar)/r = 1/r - erfc(ar)/r ; 1/r: (-1)^k (2k-1)!! r^-(2k+1);
		 * erfc(ar)/r: g_k = (-1)^k B_k, B_k = ((2k-1)B_{k-1} + (2a^2)^k/(a sqrt pi) e^{-a^2 r^2})/r^2 */
		const double r2 = r * r;
		double b = erfc(a * r) / r;
		double e = exp(-a * a * r2) / (a * sqrt(M_PI));
		for (int k = 0; k < ORDER; k++) {
			if (k > 0) {
				e *= 2.0 * a * a;
				b = ((2 * k - 1) * b + e) / r2;
			}
			const double s = (k & 1) ? -1.0 : 1.0;
			hk[k] = s * (dfact(2 * k - 1) / pow(r, 2 * k + 1) - b);
		}
	}
}

/* D_n of the Ewald correction: periodic zero-mean f minus nearest image, with
 * f = -1/r; alpha = 2, 178-image / r < 2.6 real sum, 92-vector Fourier sum,
 * constant pi/alpha^2 (cosmictiger/fmm_kernels.hpp:310-1490,
 * src/ewald_indices.cu:30-114).  X in box units; D in the scaled frame. */
void oc_ewald_greens(double D[50], const double X[3]) {
	oc_init();
	const double a = EWALD_ALPHA;
	double acc[NEXP];
	for (int i = 0; i < NEXP; i++)
		acc[i] = 0.0;
	/* primary image: + erf(a r)/r */
	{
		const double r = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
		double hk[ORDER];
		erf_over_r_derivs(r, a, hk);
		mono_t m;
		monomials(&m, X, ORDER - 1);
		for (int i = 0; i < g_n7; i++) {
			const int* n = g_tl7[i];
			acc[oc_trless_index(n[0], n[1], n[2], ORDER)] += radial_fast(&g_rt7[i], &m, hk);
		}
	}
	/* real-space images: - erfc(a r)/r */
	for (int i = 0; i < g_nreal; i++) {
		double d[3];
		for (int k = 0; k < 3; k++)
			d[k] = X[k] - g_real[i][k];
		const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
		if (r2 < EWALD_REAL_CUTOFF * EWALD_REAL_CUTOFF) {
			const double r = sqrt(r2);
			double gk[ORDER];
			double b = erfc(a * r) / r;
			double e = exp(-a * a * r2) / (a * sqrt(M_PI));
			for (int k = 0; k < ORDER; k++) {
				if (k > 0) {
					e *= 2.0 * a * a;
					b = ((2 * k - 1) * b + e) / r2;
				}
				gk[k] = ((k & 1) ? -1.0 : 1.0) * b;
			}
			mono_t m;
			monomials(&m, d, ORDER - 1);
			for (int j = 0; j < g_n7; j++) {
				const int* n = g_tl7[j];
				acc[oc_trless_index(n[0], n[1], n[2], ORDER)] -= radial_fast(&g_rt7[j], &m, gk);
			}
		}
	}
	/* Fourier: - sum_h c(h) d^n cos(2 pi h.x), c(h) = exp(-pi^2 h^2/a^2)/(pi h^2);
	 * d^l/dth^l cos(th) = cos(th + l pi/2) */
	static double four_coef[128][64];
	static int four_ready = 0;
	if (!four_ready) {
#pragma omp critical(oc_four_init)
		{
			if (!four_ready) {
				for (int i = 0; i < g_nfour; i++) {
					const double* h = g_four[i];
					const double h2 = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
					const double c = exp(-M_PI * M_PI * h2 / (a * a)) / (M_PI * h2);
					for (int j = 0; j < g_n7; j++) {
						const int* n = g_tl7[j];
						const int l = n[0] + n[1] + n[2];
						const double hn = pow(h[0], n[0]) * pow(h[1], n[1]) * pow(h[2], n[2]);
						four_coef[i][j] = -c * pow(2.0 * M_PI, l) * hn;
					}
				}
				__atomic_store_n(&four_ready, 1, __ATOMIC_RELEASE);
			}
		}
	}
	for (int i = 0; i < g_nfour; i++) {
		const double* h = g_four[i];
		const double th = 2.0 * M_PI * (h[0] * X[0] + h[1] * X[1] + h[2] * X[2]);
		const double ct = cos(th), st = sin(th);
		const double trig[4] = { ct, -st, -ct, st };
		for (int j = 0; j < g_n7; j++) {
			const int* n = g_tl7[j];
			const int l = n[0] + n[1] + n[2];
			acc[oc_trless_index(n[0], n[1], n[2], ORDER)] += four_coef[i][j] * trig[l & 3];
		}
	}
	acc[0] += M_PI / (a * a);
	for (int j = 0; j < g_n7; j++) {
		const int* n = g_tl7[j];
		const int l = n[0] + n[1] + n[2];
		const int idx = oc_trless_index(n[0], n[1], n[2], ORDER);
		D[idx] = acc[idx] * pow(SCALE, -(l + 1));
	}
}

/* L_n += sum_m M_m D_(n+m)/m!, |m| <= min(6-|n|, 5) (src/fmm_kernel_codegen.cpp:1310-1345) */
void oc_m2l(double* L, int Pout, const double M[37], const double D[50], int do_phi) {
	oc_init();
	const int (*outs)[3] = Pout == 2 ? (const int (*)[3]) g_tl2 : (const int (*)[3]) g_tl7;
	const int nout = Pout == 2 ? g_n2 : g_n7;
	for (int i = 0; i < nout; i++) {
		const int* n = outs[i];
		const int l = n[0] + n[1] + n[2];
		if (l == 0 && !do_phi) continue;
		const int mmax = (ORDER - 1 - l) < (MORDER - 1) ? (ORDER - 1 - l) : (MORDER - 1);
		double sum = 0.0;
		for (int a = 0; a <= mmax; a++)
			for (int b = 0; a + b <= mmax; b++)
				for (int c = 0; a + b + c <= mmax; c++) {
					const int m[3] = { a, b, c };
					sum += full_fast(g_fx6, M, a, b, c) * full_fast(g_fx7, D, n[0] + a, n[1] + b, n[2] + c) / mfact(m);
				}
		L[oc_trless_index(n[0], n[1], n[2], Pout)] += sum;
	}
}

static void shift_local(double* Lb, int Pout, const double La[50], const double X[3], int do_phi) {
	oc_init();
	const double x[3] = { X[0] * SCALE, X[1] * SCALE, X[2] * SCALE };
	mono_t m;
	monomials(&m, x, ORDER - 1);
	const int (*outs)[3] = Pout == 2 ? (const int (*)[3]) g_tl2 : (const int (*)[3]) g_tl7;
	const int nout = Pout == 2 ? g_n2 : g_n7;
	for (int i = 0; i < nout; i++) {
		const int* n = outs[i];
		const int l = n[0] + n[1] + n[2];
		double sum = 0.0;
		for (int a = 0; a + l < ORDER; a++)
			for (int b = 0; a + b + l < ORDER; b++)
				for (int c = 0; a + b + c + l < ORDER; c++) {
					if (l == 0 && !do_phi && a + b + c > 0) continue;
					const int k[3] = { a, b, c };
					sum += full_fast(g_fx7, La, n[0] + a, n[1] + b, n[2] + c) * m.v[a][b][c] / mfact(k);
				}
		Lb[oc_trless_index(n[0], n[1], n[2], Pout)] = sum;
	}
}

void oc_l2l(double Lb[50], const double La[50], const double X[3], int do_phi) {
	shift_local(Lb, ORDER, La, X, do_phi);
}

void oc_l2p(double L4[4], const double La[50], const double X[3], int do_phi) {
	shift_local(L4, 2, La, X, do_phi);
}

/* trless multipole of a unit mass at X = particle - centre (moments of -1000 X):
 * degree <= 2 raw, degree >= 3 traceless projection (fmm_kernel_codegen.cpp:1467-1483) */
void oc_p2m(double M[37], const double X[3]) {
	oc_init();
	const double y[3] = { -X[0] * SCALE, -X[1] * SCALE, -X[2] * SCALE };
	const double r2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
	mono_t m;
	monomials(&m, y, MORDER - 1);
	for (int i = 0; i < g_n6; i++) {
		const int* n = g_tl6[i];
		const int l = n[0] + n[1] + n[2];
		const int idx = oc_trless_index(n[0], n[1], n[2], MORDER);
		if (l <= 2) {
			M[idx] = m.v[n[0]][n[1]][n[2]];
			continue;
		}
		double sum = 0.0;
		int k[3];
		for (k[0] = 0; 2 * k[0] <= n[0]; k[0]++)
			for (k[1] = 0; 2 * k[1] <= n[1]; k[1]++)
				for (k[2] = 0; 2 * k[2] <= n[2]; k[2]++) {
					const int j = k[0] + k[1] + k[2];
					sum += tf_coef(n, k) * pow(r2, j) * m.v[n[0] - 2 * k[0]][n[1] - 2 * k[1]][n[2] - 2 * k[2]];
				}
		M[idx] = sum;
	}
}

/* shift a multipole by X = child centre - parent centre (fmm_kernel_codegen.cpp:1485-1608) */
void oc_m2m(double Mb[37], const double Ma[37], const double X[3]) {
	oc_init();
	const double y[3] = { -X[0] * SCALE, -X[1] * SCALE, -X[2] * SCALE };
	mono_t m;
	monomials(&m, y, MORDER - 1);
	static __thread double raw[8][8][8];
	for (int a = 0; a < MORDER; a++)
		for (int b = 0; a + b < MORDER; b++)
			for (int c = 0; a + b + c < MORDER; c++) {
				double sum = 0.0;
				for (int i = 0; i <= a; i++)
					for (int j = 0; j <= b; j++)
						for (int k = 0; k <= c; k++) {
							const double binom = g_fact[a] / (g_fact[i] * g_fact[a - i]) * g_fact[b] / (g_fact[j] * g_fact[b - j])
									* g_fact[c] / (g_fact[k] * g_fact[c - k]);
							sum += binom * m.v[a - i][b - j][c - k] * full_fast(g_fx6, Ma, i, j, k);
						}
				raw[a][b][c] = sum;
			}
	for (int i = 0; i < g_n6; i++) {
		const int* n = g_tl6[i];
		const int l = n[0] + n[1] + n[2];
		const int idx = oc_trless_index(n[0], n[1], n[2], MORDER);
		if (l <= 2) {
			Mb[idx] = raw[n[0]][n[1]][n[2]];
			continue;
		}
		double sum = 0.0;
		int k[3];
		for (k[0] = 0; 2 * k[0] <= n[0]; k[0]++)
			for (k[1] = 0; 2 * k[1] <= n[1]; k[1]++)
				for (k[2] = 0; 2 * k[2] <= n[2]; k[2]++) {
					const int j = k[0] + k[1] + k[2];
					const int base[3] = { n[0] - 2 * k[0], n[1] - 2 * k[1], n[2] - 2 * k[2] };
					/* Tr^j(raw)_base = sum_{|q|=j} j!/q! raw(base + 2q) */
					double tr = 0.0;
					for (int qa = 0; qa <= j; qa++)
						for (int qb = 0; qa + qb <= j; qb++) {
							const int qc = j - qa - qb;
							tr += g_fact[j] / (g_fact[qa] * g_fact[qb] * g_fact[qc]) * raw[base[0] + 2 * qa][base[1] + 2 * qb][base[2] + 2 * qc];
						}
					sum += tf_coef(n, k) * tr;
				}
		Mb[idx] = sum;
	}
}

/* Ferrers n=4 softened force and potential (cosmictiger/gravity.hpp:76-96) */
void oc_gsoft(double* f, double* phi, double q2, double hinv, double h2inv, double h3inv) {
	static const double Af[5] = { 1155. / 128., -693. / 32., 1485. / 64., -385. / 32., 315. / 128. };
	static const double Aphi[6] = { 693. / 256., -1155. / 256., 693. / 128., -495. / 128., 385. / 256., -63. / 256. };
	q2 *= h2inv;
	double ff = Af[4];
	for (int n = 3; n >= 0; n--)
		ff = ff * q2 + Af[n];
	*f = ff * h3inv;
	double pp = Aphi[5];
	for (int n = 4; n >= 0; n--)
		pp = pp * q2 + Aphi[n];
	*phi = pp * hinv;
}

double oc_self_phi(void) {
	double f, phi;
	oc_gsoft(&f, &phi, 0.0, 1.0, 1.0, 1.0);
	return phi;
}

/* ------------------------------------------------------------------ */
/* fixed32 (cosmictiger/fixed.hpp)                                      */
/* ------------------------------------------------------------------ */

float oc_distance(uint32_t a, uint32_t b) {
	return (float) (int32_t) (a - b) * (1.0f / 4294967296.0f);
}

uint32_t oc_fixed32_from_double(double x) {
	const double v = x * 4294967296.0;
	return (uint32_t) (uint64_t) (int64_t) v;
}

static double fixed_to_double(uint32_t a) {
	return (double) a * (1.0 / 4294967296.0);
}

/* fixed<int,29> from double: truncation of 2^29 x */
static int32_t rf_from_double(double x) {
	return (int32_t) (x * 536870912.0);
}

static double rf_to_double(int32_t a) {
	return (double) a * (1.0 / 536870912.0);
}

/* ------------------------------------------------------------------ */
/* tree build (src/tree.cpp:273-601)                                    */
/* ------------------------------------------------------------------ */

typedef struct {
	oc_tree* t;
	uint32_t* x[3];
	int32_t* id;
	int bucket;
	double theta;
	double h;
} build_ctx;

typedef struct {
	double multi[37];
	uint32_t pos[3];
	float radius;
	double box_begin[3], box_end[3];
	int id;
} build_ret;

static int node_alloc(oc_tree* t) {
	const int64_t i = __atomic_fetch_add(&t->nnodes, 1, __ATOMIC_RELAXED);
	if (i >= t->cap) {
		t->error = 2;
		return -1;
	}
	return (int) i;
}

/* two-pointer partition on x[dim] >= fixed32(xm) (src/particles.cpp:979-1018) */
static int64_t partition(build_ctx* c, int64_t b, int64_t e, double xm, int dim) {
	const uint32_t xmid = oc_fixed32_from_double(xm);
	uint32_t* xd = c->x[dim];
	int64_t lo = b, hi = e;
	while (lo < hi) {
		if (xd[lo] >= xmid) {
			while (lo != hi) {
				hi--;
				if (xd[hi] < xmid) {
					for (int d = 0; d < 3; d++) {
						const uint32_t tmp = c->x[d][hi];
						c->x[d][hi] = c->x[d][lo];
						c->x[d][lo] = tmp;
					}
					const int32_t ti = c->id[hi];
					c->id[hi] = c->id[lo];
					c->id[lo] = ti;
					break;
				}
			}
		}
		lo++;
	}
	return hi;
}

static build_ret build_node(build_ctx* c, int64_t b, int64_t e, const double box_b[3], const double box_e[3], int depth) {
	build_ret rc;
	oc_tree* t = c->t;
	if (depth >= MAX_DEPTH) {
		t->error = 1;
		memset(&rc, 0, sizeof(rc));
		rc.id = -1;
		return rc;
	}
	const int index = node_alloc(t);
	if (index < 0) {
		memset(&rc, 0, sizeof(rc));
		rc.id = -1;
		return rc;
	}
	const int64_t nparts = e - b;
	float box_r = 0.0f;
	for (int d = 0; d < 3; d++) {
		const double s = 0.5 * (box_e[d] - box_b[d]);
		box_r = (float) ((double) box_r + s * s);
	}
	box_r = sqrtf(box_r);
	const int ewald_satisfied = ((double) box_r < 0.25 * (c->theta / (1.0 + c->theta)) && (double) box_r < 0.125 - 0.25 * c->h);
	int32_t children[2] = { -1, -1 };
	double Xc[3];
	float radius;
	int isleaf = 1;
	if (nparts > c->bucket || (!ewald_satisfied && nparts > 0)) {
		isleaf = 0;
		/* longest dimension, first max (cosmictiger/range.hpp:212-223) */
		int xdim = 0;
		double max_span = -1.0;
		for (int d = 0; d < 3; d++) {
			const double span = box_e[d] - box_b[d];
			if (span > max_span) {
				max_span = span;
				xdim = d;
			}
		}
		const double xmid = 0.5 * (box_e[xdim] + box_b[xdim]);
		const int64_t mid = partition(c, b, e, xmid, xdim);
		double lb[3], le[3], rb[3], re[3];
		for (int d = 0; d < 3; d++) {
			lb[d] = rb[d] = box_b[d];
			le[d] = re[d] = box_e[d];
		}
		le[xdim] = rb[xdim] = xmid;
		build_ret rcl, rcr;
		if (nparts > 100000) {
#pragma omp task shared(rcl) firstprivate(b, mid, depth)
			rcl = build_node(c, b, mid, lb, le, depth + 1);
#pragma omp task shared(rcr) firstprivate(mid, e, depth)
			rcr = build_node(c, mid, e, rb, re, depth + 1);
#pragma omp taskwait
		} else {
			rcl = build_node(c, b, mid, lb, le, depth + 1);
			rcr = build_node(c, mid, e, rb, re, depth + 1);
		}
		if (t->error) {
			rc.id = -1;
			return rc;
		}
		const double ml0 = rcl.multi[0], mr0 = rcr.multi[0];
		/* union of the non-empty children's boxes */
		if (ml0 != 0.0 && mr0 != 0.0) {
			for (int d = 0; d < 3; d++) {
				rc.box_begin[d] = fmin(rcl.box_begin[d], rcr.box_begin[d]);
				rc.box_end[d] = fmax(rcl.box_end[d], rcr.box_end[d]);
			}
		} else if (ml0 != 0.0) {
			for (int d = 0; d < 3; d++) {
				rc.box_begin[d] = rcl.box_begin[d];
				rc.box_end[d] = rcl.box_end[d];
			}
		} else {
			for (int d = 0; d < 3; d++) {
				rc.box_begin[d] = rcr.box_begin[d];
				rc.box_end[d] = rcr.box_end[d];
			}
		}
		double Xl[3], Xr[3], Xc1[3], Xc2[3];
		for (int d = 0; d < 3; d++) {
			Xl[d] = fixed_to_double(rcl.pos[d]);
			Xr[d] = fixed_to_double(rcr.pos[d]);
		}
		double radius1 = 0.0, radius2;
		int tried1 = 0;
		/* (a) tight-box centre with the max particle distance in float (:380-415) */
		if (nparts <= 64 * (int64_t) c->bucket && (rc.box_end[xdim] - rc.box_begin[xdim] <= 0.5)) {
			tried1 = 1;
			int32_t xc[3];
			for (int d = 0; d < 3; d++) {
				Xc1[d] = 0.5 * (rc.box_begin[d] + rc.box_end[d]);
				xc[d] = (int32_t) oc_fixed32_from_double(Xc1[d]);
			}
			float r2max = 0.0f;
			for (int64_t i = b; i < e; i++) {
				float dx[3];
				for (int d = 0; d < 3; d++)
					dx[d] = (float) (int32_t) ((uint32_t) c->x[d][i] - (uint32_t) xc[d]) * (1.0f / 4294967296.0f);
				const float r2 = fmaf(dx[0], dx[0], fmaf(dx[1], dx[1], dx[2] * dx[2]));
				r2max = fmaxf(r2max, r2);
			}
			radius1 = (double) sqrtf(r2max);
		}
		/* (b) enclosing sphere of the child spheres (:416-456) */
		const double Rl = rcl.radius, Rr = rcr.radius;
		double N[3];
		double norminv = 0.0;
		for (int d = 0; d < 3; d++) {
			N[d] = Xl[d] - Xr[d];
			norminv += N[d] * N[d];
		}
		norminv = 1.0 / sqrt(norminv);
		for (int d = 0; d < 3; d++) {
			N[d] *= norminv;
			N[d] = fabs(N[d]);
		}
		double r = 0.0;
		if (mr0 != 0.0) {
			if (ml0 != 0.0) {
				for (int d = 0; d < 3; d++) {
					const double xmax = fmax(Xl[d] + N[d] * Rl, Xr[d] + N[d] * Rr);
					const double xmin = fmin(Xl[d] - N[d] * Rl, Xr[d] - N[d] * Rr);
					Xc2[d] = (xmax + xmin) * 0.5;
					r += ((xmax - xmin) * 0.5) * ((xmax - xmin) * 0.5);
				}
			} else {
				for (int d = 0; d < 3; d++)
					Xc2[d] = Xr[d];
				r = Rr * Rr;
			}
		} else {
			for (int d = 0; d < 3; d++)
				Xc2[d] = Xl[d];
			r = Rl * Rl;
		}
		radius2 = sqrt(r);
		if (tried1 && radius1 < radius2) {
			for (int d = 0; d < 3; d++)
				Xc[d] = Xc1[d];
			radius = (float) radius1;
		} else {
			for (int d = 0; d < 3; d++)
				Xc[d] = Xc2[d];
			radius = (float) radius2;
		}
		/* clip by the half-diagonal of the tight box (:464-476) */
		r = 0.0;
		for (int d = 0; d < 3; d++) {
			const double span = rc.box_end[d] - rc.box_begin[d];
			r += (span * 0.5) * (span * 0.5);
		}
		r = sqrt(r);
		if (r < (double) radius) {
			radius = (float) r;
			for (int d = 0; d < 3; d++)
				Xc[d] = (rc.box_begin[d] + rc.box_end[d]) * 0.5;
		}
		for (int d = 0; d < 3; d++)
			rc.pos[d] = oc_fixed32_from_double(Xc[d]);
		/* M2M in double about the unrounded centre (:481-495) */
		double ml[37], mr[37];
		double dl[3], dr[3];
		for (int d = 0; d < 3; d++) {
			dl[d] = Xl[d] - Xc[d];
			dr[d] = Xr[d] - Xc[d];
		}
		oc_m2m(ml, rcl.multi, dl);
		oc_m2m(mr, rcr.multi, dr);
		for (int i = 0; i < NMULT; i++)
			rc.multi[i] = ml[i] + mr[i];
		children[0] = rcl.id;
		children[1] = rcr.id;
	} else {
		/* leaf: tight box, centre, radius in double, P2M about fixed32(centre) (:499-571) */
		for (int d = 0; d < 3; d++) {
			rc.box_begin[d] = box_e[d];
			rc.box_end[d] = box_b[d];
		}
		for (int64_t i = b; i < e; i++) {
			for (int d = 0; d < 3; d++) {
				const double x = fixed_to_double(c->x[d][i]);
				rc.box_end[d] = fmax(rc.box_end[d], x);
				rc.box_begin[d] = fmin(rc.box_begin[d], x);
			}
		}
		for (int d = 0; d < 3; d++)
			Xc[d] = (rc.box_end[d] + rc.box_begin[d]) * 0.5;
		double r = 0.0;
		for (int64_t i = b; i < e; i++) {
			double this_radius = 0.0;
			for (int d = 0; d < 3; d++) {
				const double x = fixed_to_double(c->x[d][i]);
				this_radius += (x - Xc[d]) * (x - Xc[d]);
			}
			r = fmax(r, this_radius);
		}
		radius = (float) sqrt(r);
		uint32_t Y[3];
		for (int d = 0; d < 3; d++) {
			Y[d] = oc_fixed32_from_double(Xc[d]);
			rc.pos[d] = Y[d];
		}
		for (int i = 0; i < NMULT; i++)
			rc.multi[i] = 0.0;
		for (int64_t i = b; i < e; i++) {
			double dx[3];
			for (int d = 0; d < 3; d++)
				dx[d] = (double) oc_distance(c->x[d][i], Y[d]);
			double m[37];
			oc_p2m(m, dx);
			for (int k = 0; k < NMULT; k++)
				rc.multi[k] += m[k];
		}
	}
	rc.radius = radius;
	rc.id = index;
	oc_node* node = &t->nodes[index];
	node->radius = radius;
	node->child[0] = children[0];
	node->child[1] = children[1];
	node->pbegin = (int32_t) b;
	node->pend = (int32_t) e;
	for (int d = 0; d < 3; d++) {
		node->pos[d] = rc.pos[d];
		node->box_begin[d] = rf_from_double(rc.box_begin[d]);
		node->box_end[d] = rf_from_double(rc.box_end[d]);
	}
	node->leaf = isleaf;
	node->depth = depth;
	for (int i = 0; i < NMULT; i++)
		node->multi[i] = rc.multi[i];
	return rc;
}

int oc_tree_build(oc_tree* t, int64_t n, uint32_t* x, uint32_t* y, uint32_t* z, int32_t* id, int bucket, double theta,
		double h, int nthreads, int64_t cap_hint) {
	oc_init();
	if (nthreads <= 0) nthreads = 1;
	/* node budget; the build reports error 2 when it is exceeded and the caller
	 * retries from the original arrays with a larger budget */
	t->cap = cap_hint > 0 ? cap_hint : 8 * n / (bucket > 0 ? bucket : 1) + 65536;
	t->nodes = (oc_node*) malloc(t->cap * sizeof(oc_node));
	t->nnodes = 0;
	t->error = 0;
	build_ctx c;
	c.t = t;
	c.x[0] = x;
	c.x[1] = y;
	c.x[2] = z;
	c.id = id;
	c.bucket = bucket;
	c.theta = theta;
	c.h = h;
	const double bb[3] = { 0, 0, 0 }, be[3] = { 1, 1, 1 };
	build_ret rc;
#pragma omp parallel num_threads(nthreads)
#pragma omp single
	rc = build_node(&c, 0, n, bb, be, 0);
	t->root = rc.id;
	return t->error;
}

void oc_tree_free(oc_tree* t) {
	free(t->nodes);
	t->nodes = NULL;
	t->nnodes = t->cap = 0;
}

/* ------------------------------------------------------------------ */
/* kick walk (src/kick.cu:179-506) and interactions (src/gravity.cu)    */
/* ------------------------------------------------------------------ */

static const float rung_dt[16] = { 1.0f / (1 << 0), 1.0f / (1 << 1), 1.0f / (1 << 2), 1.0f / (1 << 3), 1.0f / (1 << 4), 1.0f / (1 << 5),
		1.0f / (1 << 6), 1.0f / (1 << 7), 1.0f / (1 << 8), 1.0f / (1 << 9), 1.0f / (1 << 10), 1.0f / (1 << 11), 1.0f / (1 << 12),
		1.0f / (1 << 13), 1.0f / (1 << 14), 1.0f / (1 << 15) };

typedef struct {
	int32_t* v;
	int64_t n, cap;
} ivec;

static void ivec_push(ivec* a, int32_t x) {
	if (a->n >= a->cap) {
		a->cap = a->cap ? 2 * a->cap : 16;
		a->v = (int32_t*) realloc(a->v, a->cap * sizeof(int32_t));
	}
	a->v[a->n++] = x;
}

typedef struct {
	const oc_tree* t;
	const oc_kick_params* p;
	const uint32_t* x[3];
	float* vel;
	int8_t* rung;
	double *gx, *gy, *gz, *phi;
	ivec* lists[5]; /* per node, may be NULL */
	int64_t sample_b, sample_e; /* only sinks in [sample_b, sample_e) are kicked */
	int error;
	oc_kick_return ret;
} kick_ctx;

/* distance(range_fixed a, fixed32 b) (cosmictiger/range.hpp:444-450) */
static float rf_distance(int32_t a, uint32_t b) {
	float f = (float) (rf_to_double(a) - fixed_to_double(b));
	f = (float) ((double) f - (double) (f > 0.5f));
	f = (float) ((double) f + (double) (f < -0.5f));
	return f;
}

/* box_intersects_sphere (cosmictiger/range.hpp:455-466) */
static int box_intersects_sphere(const oc_node* box, const uint32_t x[3], float r) {
	float d2 = 0.0f;
	for (int d = 0; d < 3; d++) {
		const float db = -rf_distance(box->box_begin[d], x[d]);
		const float de = -rf_distance(box->box_end[d], x[d]);
		const float d2b = db * db;
		const float d2e = de * de;
		const int32_t rfx = (int32_t) (x[d] >> 3);
		const int outside = (rfx < box->box_begin[d]) + (rfx > box->box_end[d]);
		if (outside) d2 += (float) outside * fminf(d2b, d2e);
	}
	return d2 < r * r;
}

static float mac_r2(const oc_node* a, const oc_node* b) {
	const float dx = oc_distance(a->pos[0], b->pos[0]);
	const float dy = oc_distance(a->pos[1], b->pos[1]);
	const float dz = oc_distance(a->pos[2], b->pos[2]);
	return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

enum { K_NONE = 0, K_CC = 1, K_CP = 2, K_PC = 3, K_LEAF = 4, K_NEXT = 5 };

/* Ewald acceptance (src/kick.cu:286-299) */
static int classify_ewald(const oc_node* self, const oc_node* other, float thetainv) {
	if (other->pend - other->pbegin == 0) return K_NONE;
	float R2 = mac_r2(self, other);
	const float m = fmaxf(0.5f - (self->radius + other->radius), 0.f);
	R2 = fmaxf(R2, m * m);
	const float dcc = (self->radius + other->radius) * thetainv;
	if (R2 > dcc * dcc) return K_CC;
	return K_NEXT;
}

/* direct-list classification (src/kick.cu:353-386) */
static int classify_direct(const oc_node* self, const oc_node* other, float thetainv, float h) {
	if (other->pend - other->pbegin == 0) return K_NONE;
	const float R2 = mac_r2(self, other);
	const float rs = self->radius, ro = other->radius;
	const float mind = (rs + ro) + h;
	const float dcc = fmaxf((rs + ro) * thetainv, mind);
	if (R2 > dcc * dcc) return K_CC;
	int pc = 0, cp = 0;
	if (other->leaf && self->leaf) {
		const float dcp = fmaxf(fmaf(rs, thetainv, ro), mind);
		const float dpc = fmaxf(fmaf(ro, thetainv, rs), mind);
		const int self_parts = self->pend - self->pbegin;
		const int other_parts = other->pend - other->pbegin;
		pc = (R2 > dpc * dpc || !box_intersects_sphere(self, other->pos, fmaxf(thetainv * ro, h))) && self_parts >= MIN_PARTS_PCCP;
		cp = (R2 > dcp * dcp || !box_intersects_sphere(other, self->pos, fmaxf(thetainv * rs, h))) && other_parts >= MIN_PARTS_PCCP;
		if (pc && cp) {
			if (self_parts < other_parts) {
				cp = 0;
			} else if (self_parts > other_parts) {
				pc = 0;
			} else {
				cp = pc = 0;
			}
		}
	}
	if (pc) return K_PC;
	if (cp) return K_CP;
	return other->leaf ? K_LEAF : K_NEXT;
}

static void do_kick(kick_ctx* c, const oc_node* self, const double L[50], double* fx, double* fy, double* fz, double* fp,
		oc_kick_return* kr) {
	const oc_kick_params* p = c->p;
	const int n = self->pend - self->pbegin;
	for (int k = 0; k < n; k++) {
		const int i = self->pbegin + k;
		double dx[3];
		for (int d = 0; d < 3; d++)
			dx[d] = (double) oc_distance(c->x[d][i], self->pos[d]);
		double L4[4];
		oc_l2p(L4, L, dx, p->do_phi);
		double phi = fp[k] + SCALE * L4[0];
		double g[3] = { fx[k] - SCALE * SCALE * L4[1], fy[k] - SCALE * SCALE * L4[2], fz[k] - SCALE * SCALE * L4[3] };
		for (int d = 0; d < 3; d++)
			g[d] *= p->GM;
		phi *= p->GM;
		float* v = c->vel + 3 * (int64_t) i;
		double vx = v[0], vy = v[1], vz = v[2];
		const double kin0 = 0.5 * (vx * vx + vy * vy + vz * vz);
		const double sgn = p->sign * (p->top ? 1.0 : -1.0);
		if (p->ascending) {
			const double dt = 0.5 * rung_dt[p->min_rung] * p->t0;
			if (!p->first_call) {
				vx += sgn * g[0] * dt;
				vy += sgn * g[1] * dt;
				vz += sgn * g[2] * dt;
			}
		}
		kr->kin += 0.5 * (vx * vx + vy * vy + vz * vz);
		kr->xmom += vx;
		kr->ymom += vy;
		kr->zmom += vz;
		kr->nmom += sqrt(vx * vx + vy * vy + vz * vz);
		if (c->gx) {
			c->gx[i] = g[0];
			c->gy[i] = g[1];
			c->gz[i] = g[2];
			c->phi[i] = phi;
		}
		if (p->descending) {
			const double g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
			double dt = fmin(p->eta * sqrt(p->a * p->h / sqrt(g2)), p->t0);
			dt = fmin(p->max_dt, dt);
			int r = c->rung[i];
			const int up = (int) ceil(log2(p->t0 / dt)) > p->min_rung;
			const int cand = p->min_rung + up;
			r = cand > r - 1 ? cand : r - 1;
			c->rung[i] = (int8_t) r;
			if (r > kr->max_rung) kr->max_rung = r;
			dt = 0.5 * rung_dt[p->min_rung] * p->t0;
			vx += sgn * g[0] * dt;
			vy += sgn * g[1] * dt;
			vz += sgn * g[2] * dt;
		}
		const double kin1 = 0.5 * (vx * vx + vy * vy + vz * vz);
		v[0] = (float) vx;
		v[1] = (float) vy;
		v[2] = (float) vz;
		kr->pot += 0.5 * phi;
		kr->dkin += kin1 - kin0;
	}
}

typedef struct {
	int32_t* v;
	int64_t n;
} span;

static void walk(kick_ctx* c, int self_id, const double* Lin, const uint32_t Lpos[3], const int32_t* dchk, int64_t nd, const int32_t* echk,
		int64_t ne, int fork_depth) {
	const oc_tree* t = c->t;
	const oc_kick_params* p = c->p;
	const oc_node* self = &t->nodes[self_id];
	if (self->pend <= c->sample_b || self->pbegin >= c->sample_e) return;
	const float thetainv = 1.f / p->theta;
	const float h = p->h;
	double L[50];
	{
		double dx[3];
		for (int d = 0; d < 3; d++)
			dx[d] = (double) oc_distance(self->pos[d], Lpos[d]);
		oc_l2l(L, Lin, dx, p->do_phi);
	}
	const int nsinks = self->pend - self->pbegin;
	ivec enext = { 0 }, dnext = { 0 };
	ivec cc = { 0 }, cp = { 0 }, pc = { 0 }, leaf = { 0 }, ecc = { 0 };
	double *fx = NULL, *fy = NULL, *fz = NULL, *fp = NULL;
	if (self->leaf && nsinks > 0) {
		fx = (double*) calloc(nsinks, sizeof(double));
		fy = (double*) calloc(nsinks, sizeof(double));
		fz = (double*) calloc(nsinks, sizeof(double));
		fp = (double*) malloc(nsinks * sizeof(double));
		for (int k = 0; k < nsinks; k++)
			fp[k] = oc_self_phi() / p->h;
	}
	if (nsinks) {
		/* Ewald list: one pass */
		for (int64_t i = 0; i < ne; i++) {
			const oc_node* o = &t->nodes[echk[i]];
			const int k = classify_ewald(self, o, thetainv);
			if (k == K_CC) {
				ivec_push(&ecc, echk[i]);
			} else if (k == K_NEXT) {
				if (o->leaf) {
					c->error = 3;
					continue;
				}
				ivec_push(&enext, o->child[0]);
				ivec_push(&enext, o->child[1]);
			}
		}
		if (self->leaf && enext.n) {
			c->error = 3;
		}
		for (int64_t i = 0; i < ecc.n; i++) {
			const oc_node* o = &t->nodes[ecc.v[i]];
			double dx[3], D[50];
			for (int d = 0; d < 3; d++)
				dx[d] = (double) oc_distance(self->pos[d], o->pos[d]);
			oc_ewald_greens(D, dx);
			oc_m2l(L, ORDER, o->multi, D, p->do_phi);
		}
		/* direct list: until empty for a leaf */
		ivec cur = { 0 };
		for (int64_t i = 0; i < nd; i++)
			ivec_push(&cur, dchk[i]);
		do {
			ivec nxt = { 0 };
			for (int64_t i = 0; i < cur.n; i++) {
				const oc_node* o = &t->nodes[cur.v[i]];
				const int k = classify_direct(self, o, thetainv, h);
				switch (k) {
				case K_CC:
					ivec_push(&cc, cur.v[i]);
					break;
				case K_CP:
					ivec_push(&cp, cur.v[i]);
					break;
				case K_PC:
					ivec_push(&pc, cur.v[i]);
					break;
				case K_LEAF:
					ivec_push(&leaf, cur.v[i]);
					break;
				case K_NEXT:
					ivec_push(&nxt, o->child[0]);
					ivec_push(&nxt, o->child[1]);
					break;
				default:
					break;
				}
			}
			free(cur.v);
			cur = nxt;
		} while (cur.n && self->leaf);
		dnext = cur;
		for (int64_t i = 0; i < cc.n; i++) {
			const oc_node* o = &t->nodes[cc.v[i]];
			double dx[3], D[50];
			for (int d = 0; d < 3; d++)
				dx[d] = (double) oc_distance(self->pos[d], o->pos[d]);
			oc_greens(D, dx);
			oc_m2l(L, ORDER, o->multi, D, p->do_phi);
		}
		if (self->leaf) {
			/* CP: source particles into L (src/gravity.cu:75-155) */
			for (int64_t i = 0; i < cp.n; i++) {
				const oc_node* o = &t->nodes[cp.v[i]];
				for (int j = o->pbegin; j < o->pend; j++) {
					double dx[3], D[50];
					for (int d = 0; d < 3; d++)
						dx[d] = (double) oc_distance(self->pos[d], c->x[d][j]);
					oc_greens(D, dx);
					for (int q = 0; q < NEXP; q++)
						L[q] += D[q];
				}
			}
			/* PC: multipoles onto sinks (src/gravity.cu:157-211) */
			for (int64_t i = 0; i < pc.n; i++) {
				const oc_node* o = &t->nodes[pc.v[i]];
				for (int k = 0; k < nsinks; k++) {
					const int si = self->pbegin + k;
					double dx[3], D[50], L4[4] = { 0, 0, 0, 0 };
					for (int d = 0; d < 3; d++)
						dx[d] = (double) oc_distance(c->x[d][si], o->pos[d]);
					oc_greens(D, dx);
					oc_m2l(L4, 2, o->multi, D, p->do_phi);
					fx[k] -= SCALE * SCALE * L4[1];
					fy[k] -= SCALE * SCALE * L4[2];
					fz[k] -= SCALE * SCALE * L4[3];
					fp[k] += SCALE * L4[0];
				}
			}
			/* PP (src/gravity.cu:246-404) */
			const double h2 = (double) (h * h);
			const double hinv = 1.0 / h, h2inv = hinv * hinv, h3inv = h2inv * hinv;
			for (int k = 0; k < nsinks; k++) {
				const int si = self->pbegin + k;
				double ax = 0, ay = 0, az = 0, pot = 0;
				for (int64_t i = 0; i < leaf.n; i++) {
					const oc_node* o = &t->nodes[leaf.v[i]];
					for (int j = o->pbegin; j < o->pend; j++) {
						const double dx = oc_distance(c->x[0][si], c->x[0][j]);
						const double dy = oc_distance(c->x[1][si], c->x[1][j]);
						const double dz = oc_distance(c->x[2][si], c->x[2][j]);
						const double r2 = dx * dx + dy * dy + dz * dz;
						double r1inv, r3inv;
						if (r2 > h2) {
							r1inv = 1.0 / sqrt(r2);
							r3inv = r1inv * r1inv * r1inv;
						} else {
							oc_gsoft(&r3inv, &r1inv, r2, hinv, h2inv, h3inv);
						}
						ax += dx * r3inv;
						ay += dy * r3inv;
						az += dz * r3inv;
						pot -= r1inv;
					}
				}
				fx[k] -= ax;
				fy[k] -= ay;
				fz[k] -= az;
				fp[k] += pot;
			}
		} else {
			for (int64_t i = 0; i < leaf.n; i++)
				ivec_push(&dnext, leaf.v[i]);
		}
	}
	if (c->lists[0]) {
		ivec* srcs[5] = { &ecc, &cc, &cp, &pc, &leaf };
		for (int q = 0; q < (self->leaf ? 5 : 2); q++) {
			ivec* dst = &c->lists[q][self_id];
			for (int64_t i = 0; i < srcs[q]->n; i++)
				ivec_push(dst, srcs[q]->v[i]);
		}
	}
	if (self->leaf) {
		if (nsinks) {
			oc_kick_return kr = { 0 };
			do_kick(c, self, L, fx, fy, fz, fp, &kr);
#pragma omp critical(oc_kick_ret)
			{
				if (kr.max_rung > c->ret.max_rung) c->ret.max_rung = kr.max_rung;
				c->ret.pot += kr.pot;
				c->ret.kin += kr.kin;
				c->ret.dkin += kr.dkin;
				c->ret.xmom += kr.xmom;
				c->ret.ymom += kr.ymom;
				c->ret.zmom += kr.zmom;
				c->ret.nmom += kr.nmom;
			}
		}
	} else if (nsinks || 1) {
		const int c0 = self->child[0], c1 = self->child[1];
		if (fork_depth > 0) {
#pragma omp task firstprivate(c0) shared(c, L, self, dnext, enext)
			walk(c, c0, L, self->pos, dnext.v, dnext.n, enext.v, enext.n, fork_depth - 1);
#pragma omp task firstprivate(c1) shared(c, L, self, dnext, enext)
			walk(c, c1, L, self->pos, dnext.v, dnext.n, enext.v, enext.n, fork_depth - 1);
#pragma omp taskwait
		} else {
			walk(c, c0, L, self->pos, dnext.v, dnext.n, enext.v, enext.n, 0);
			walk(c, c1, L, self->pos, dnext.v, dnext.n, enext.v, enext.n, 0);
		}
	}
	free(fx);
	free(fy);
	free(fz);
	free(fp);
	free(enext.v);
	free(dnext.v);
	free(cc.v);
	free(cp.v);
	free(pc.v);
	free(leaf.v);
	free(ecc.v);
}

int oc_kick_sample(const oc_tree* t, const oc_kick_params* p, const uint32_t* x, const uint32_t* y, const uint32_t* z, float* vel,
		int8_t* rung, double* gx, double* gy, double* gz, double* phi, oc_kick_return* ret, oc_lists* lists, int nthreads,
		int64_t sample_b, int64_t sample_e);

int oc_kick(const oc_tree* t, const oc_kick_params* p, const uint32_t* x, const uint32_t* y, const uint32_t* z, float* vel,
		int8_t* rung, double* gx, double* gy, double* gz, double* phi, oc_kick_return* ret, oc_lists* lists, int nthreads) {
	return oc_kick_sample(t, p, x, y, z, vel, rung, gx, gy, gz, phi, ret, lists, nthreads, 0, INT64_MAX);
}

int oc_kick_sample(const oc_tree* t, const oc_kick_params* p, const uint32_t* x, const uint32_t* y, const uint32_t* z, float* vel,
		int8_t* rung, double* gx, double* gy, double* gz, double* phi, oc_kick_return* ret, oc_lists* lists, int nthreads,
		int64_t sample_b, int64_t sample_e) {
	oc_init();
	kick_ctx c;
	memset(&c, 0, sizeof(c));
	c.t = t;
	c.p = p;
	c.x[0] = x;
	c.x[1] = y;
	c.x[2] = z;
	c.vel = vel;
	c.rung = rung;
	c.gx = gx;
	c.gy = gy;
	c.gz = gz;
	c.phi = phi;
	c.sample_b = sample_b;
	c.sample_e = sample_e;
	if (lists) {
		for (int q = 0; q < 5; q++)
			c.lists[q] = (ivec*) calloc(t->nnodes, sizeof(ivec));
	}
	double L[50];
	for (int i = 0; i < 50; i++)
		L[i] = 0.0;
	const uint32_t pos0[3] = { 0, 0, 0 };
	const int32_t root = t->root;
	if (nthreads <= 0) nthreads = 1;
#ifdef _OPENMP
	const int fork_depth = nthreads > 1 ? 12 : 0;
#pragma omp parallel num_threads(nthreads)
#pragma omp single
	walk(&c, root, L, pos0, &root, 1, &root, 1, fork_depth);
#else
	walk(&c, root, L, pos0, &root, 1, &root, 1, 0);
#endif
	if (ret) *ret = c.ret;
	if (lists) {
		lists->nnodes = t->nnodes;
		for (int q = 0; q < 5; q++) {
			lists->off[q] = (int64_t*) malloc((t->nnodes + 1) * sizeof(int64_t));
			int64_t tot = 0;
			for (int64_t i = 0; i < t->nnodes; i++) {
				lists->off[q][i] = tot;
				tot += c.lists[q][i].n;
			}
			lists->off[q][t->nnodes] = tot;
			lists->idx[q] = (int32_t*) malloc((tot ? tot : 1) * sizeof(int32_t));
			for (int64_t i = 0; i < t->nnodes; i++) {
				if (c.lists[q][i].n) memcpy(lists->idx[q] + lists->off[q][i], c.lists[q][i].v, c.lists[q][i].n * sizeof(int32_t));
				free(c.lists[q][i].v);
			}
			free(c.lists[q]);
		}
	}
	return c.error;
}

void oc_lists_free(oc_lists* l) {
	for (int q = 0; q < 5; q++) {
		free(l->off[q]);
		free(l->idx[q]);
		l->off[q] = NULL;
		l->idx[q] = NULL;
	}
}

/* ------------------------------------------------------------------ */
/* drift (src/drift.cu:36-154)                                          */
/* ------------------------------------------------------------------ */

void oc_drift(int64_t n, uint32_t* x, uint32_t* y, uint32_t* z, const float* vel, const int8_t* rungs, int rung, float a,
		float dt) {
	const double c1 = (double) dt / (double) a;
	uint32_t* X[3] = { x, y, z };
	for (int64_t i = 0; i < n; i++) {
		if (rungs[i] != rung) continue;
		for (int d = 0; d < 3; d++) {
			const double x0 = fixed_to_double(X[d][i]);
			double x1 = fma((double) vel[3 * i + d], c1, x0);
			x1 -= (double) (x1 >= 1.0);
			x1 += (double) (x1 < 0.0);
			X[d][i] = (uint32_t) (uint64_t) (x1 * 4294967296.0);
		}
	}
}

/* ------------------------------------------------------------------ */
/* FP64 direct Ewald                                                    */
/* ------------------------------------------------------------------ */

#define DE_ALPHA 9.0
#define DE_RC 0.5
#define DE_H2MAX 220

typedef struct {
	int h[3];
	double coef; /* 2 c(h) for the half-space vector h */
} de_wave;

static de_wave* de_waves = NULL;
static int de_nwave = 0;

static void de_init(void) {
	if (de_waves) return;
	const int hm = (int) sqrt((double) DE_H2MAX) + 1;
	de_wave* w = (de_wave*) malloc(sizeof(de_wave) * (2 * hm + 1) * (2 * hm + 1) * (2 * hm + 1));
	int n = 0;
	for (int i = -hm; i <= hm; i++)
		for (int j = -hm; j <= hm; j++)
			for (int k = -hm; k <= hm; k++) {
				const int h2 = i * i + j * j + k * k;
				if (h2 == 0 || h2 > DE_H2MAX) continue;
				/* canonical half space: first nonzero component positive */
				if (i < 0 || (i == 0 && j < 0) || (i == 0 && j == 0 && k < 0)) continue;
				w[n].h[0] = i;
				w[n].h[1] = j;
				w[n].h[2] = k;
				w[n].coef = 2.0 * exp(-M_PI * M_PI * h2 / (DE_ALPHA * DE_ALPHA)) / (M_PI * h2);
				n++;
			}
	de_nwave = n;
	de_waves = w;
}

double oc_ewald_psi0(void) {
	de_init();
	double s = 0.0;
	for (int i = 0; i < de_nwave; i++)
		s += de_waves[i].coef;
	/* real images n != 0 */
	for (int i = -3; i <= 3; i++)
		for (int j = -3; j <= 3; j++)
			for (int k = -3; k <= 3; k++) {
				if (!i && !j && !k) continue;
				const double r = sqrt((double) (i * i + j * j + k * k));
				s += erfc(DE_ALPHA * r) / r;
			}
	return s - 2.0 * DE_ALPHA / sqrt(M_PI) - M_PI / (DE_ALPHA * DE_ALPHA);
}

void oc_ewald_psi(const double X[3], double* psi, double grad[3]) {
	de_init();
	double s = 0.0, g[3] = { 0, 0, 0 };
	for (int i = -3; i <= 3; i++)
		for (int j = -3; j <= 3; j++)
			for (int k = -3; k <= 3; k++) {
				const double d[3] = { X[0] - i, X[1] - j, X[2] - k };
				const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
				if (r > 3.0) continue;
				const double er = erfc(DE_ALPHA * r) / r;
				s += er;
				const double dr = -(er + 2.0 * DE_ALPHA / sqrt(M_PI) * exp(-DE_ALPHA * DE_ALPHA * r * r)) / (r * r);
				for (int q = 0; q < 3; q++)
					g[q] += dr * d[q];
			}
	for (int i = 0; i < de_nwave; i++) {
		const int* h = de_waves[i].h;
		const double th = 2.0 * M_PI * (h[0] * X[0] + h[1] * X[1] + h[2] * X[2]);
		s += de_waves[i].coef * cos(th);
		for (int q = 0; q < 3; q++)
			g[q] -= de_waves[i].coef * 2.0 * M_PI * h[q] * sin(th);
	}
	*psi = s - M_PI / (DE_ALPHA * DE_ALPHA);
	for (int q = 0; q < 3; q++)
		grad[q] = g[q];
}

void oc_direct_ewald(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int64_t nsink, const int64_t* sinks,
		double h, double GM, double* gx, double* gy, double* gz, double* phi, int nthreads) {
	de_init();
	if (nthreads <= 0) nthreads = 1;
	const int nw = de_nwave;
	const int hm = (int) sqrt((double) DE_H2MAX) + 1;
	/* structure factors S(h) = sum_j exp(2 pi i h.x_j) */
	double* Sc = (double*) calloc(nw, sizeof(double));
	double* Ss = (double*) calloc(nw, sizeof(double));
#pragma omp parallel num_threads(nthreads)
	{
		double* lc = (double*) calloc(nw, sizeof(double));
		double* ls = (double*) calloc(nw, sizeof(double));
		double ec[3][64], es[3][64]; /* index k + hm */
#pragma omp for schedule(static)
		for (int64_t j = 0; j < n; j++) {
			const double X[3] = { fixed_to_double(x[j]), fixed_to_double(y[j]), fixed_to_double(z[j]) };
			for (int d = 0; d < 3; d++) {
				for (int k = -hm; k <= hm; k++) {
					const double th = 2.0 * M_PI * k * X[d];
					ec[d][k + hm] = cos(th);
					es[d][k + hm] = sin(th);
				}
			}
			for (int w = 0; w < nw; w++) {
				const int* hh = de_waves[w].h;
				/* exp(i(a+b+c)) */
				const double c0 = ec[0][hh[0] + hm], s0 = es[0][hh[0] + hm];
				const double c1 = ec[1][hh[1] + hm], s1 = es[1][hh[1] + hm];
				const double c2 = ec[2][hh[2] + hm], s2 = es[2][hh[2] + hm];
				const double c01 = c0 * c1 - s0 * s1, s01 = s0 * c1 + c0 * s1;
				lc[w] += c01 * c2 - s01 * s2;
				ls[w] += s01 * c2 + c01 * s2;
			}
		}
#pragma omp critical(oc_sf)
		{
			for (int w = 0; w < nw; w++) {
				Sc[w] += lc[w];
				Ss[w] += ls[w];
			}
		}
		free(lc);
		free(ls);
	}
	const double psi0 = oc_ewald_psi0();
	const double hinv = 1.0 / h, h2inv = hinv * hinv, h3inv = h2inv * hinv;
	const double a = DE_ALPHA;
	const double twoa_sqrtpi = 2.0 * a / sqrt(M_PI);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
	for (int64_t s = 0; s < nsink; s++) {
		const int64_t i = sinks[s];
		const double X[3] = { fixed_to_double(x[i]), fixed_to_double(y[i]), fixed_to_double(z[i]) };
		double psi = 0.0, g[3] = { 0, 0, 0 };
		double soft_phi = 0.0, soft_g[3] = { 0, 0, 0 };
		/* Fourier part over all sources (includes j = i) */
		for (int w = 0; w < nw; w++) {
			const int* hh = de_waves[w].h;
			const double th = 2.0 * M_PI * (hh[0] * X[0] + hh[1] * X[1] + hh[2] * X[2]);
			const double ct = cos(th), st = sin(th);
			const double cf = de_waves[w].coef;
			psi += cf * (ct * Sc[w] + st * Ss[w]);
			const double gd = cf * (-st * Sc[w] + ct * Ss[w]);
			for (int q = 0; q < 3; q++)
				g[q] += gd * 2.0 * M_PI * hh[q];
		}
		/* real part: minimum image within rc, j != i */
		for (int64_t j = 0; j < n; j++) {
			if (j == i) continue;
			const double d[3] = { (double) (int32_t) (x[i] - x[j]) * (1.0 / 4294967296.0), (double) (int32_t) (y[i] - y[j]) * (1.0
					/ 4294967296.0), (double) (int32_t) (z[i] - z[j]) * (1.0 / 4294967296.0) };
			const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
			if (r2 >= DE_RC * DE_RC) continue;
			const double r = sqrt(r2);
			if (r2 == 0.0) continue;
			const double er = erfc(a * r) / r;
			psi += er;
			const double dr = -(er + twoa_sqrtpi * exp(-a * a * r2)) / r2;
			for (int q = 0; q < 3; q++)
				g[q] += dr * d[q];
			if (r2 < h * h) {
				double fs, ps;
				oc_gsoft(&fs, &ps, r2, hinv, h2inv, h3inv);
				soft_phi -= (ps - 1.0 / r);
				const double r3 = 1.0 / (r2 * r);
				for (int q = 0; q < 3; q++)
					soft_g[q] -= d[q] * (fs - r3);
			}
		}
		/* self: real-space limit of erfc(ar)/r - 1/r and the constant of every source */
		psi += -twoa_sqrtpi;
		psi -= (double) n * M_PI / (a * a);
		(void) psi0;
		phi[s] = GM * (-psi + soft_phi);
		gx[s] = GM * (g[0] + soft_g[0]);
		gy[s] = GM * (g[1] + soft_g[1]);
		gz[s] = GM * (g[2] + soft_g[2]);
	}
	free(Sc);
	free(Ss);
}

/* ------------------------------------------------------------------ */
/* HEALPix nested pixels (src/healpix.cu:36-172, the HEALPix C        */
/* library's vec2pix_nest; pix2vec_nest is its published inverse)      */
/* ------------------------------------------------------------------ */

#define HP_NS_MAX 8192

/* bits of i spread to the even bit positions (the x2pix table of
 * src/healpix.cu:36-72; y2pix = 2 x2pix) */
static int64_t hp_spread(int64_t i) {
	int64_t k = 0;
	for (int b = 0; b < 32; b++) k |= ((i >> b) & 1) << (2 * b);
	return k;
}

static int64_t hp_compress(int64_t k) {
	int64_t i = 0;
	for (int b = 0; b < 32; b++) i |= ((k >> (2 * b)) & 1) << b;
	return i;
}

int64_t oc_vec2pix_nest(int64_t nside, const double vec[3]) {
	if (nside < 1 || nside > HP_NS_MAX) return -1;
	const double piover2 = 0.5 * M_PI, twopi = 2.0 * M_PI;
	const double z = vec[2] / sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
	double phi = 0.0;
	if (vec[0] != 0.0 || vec[1] != 0.0) {
		phi = atan2(vec[1], vec[0]);
		if (phi < 0.0) phi += twopi;
	}
	const double za = fabs(z);
	const double tt = phi / piover2;
	int face, ix, iy;
	if (za <= 2.0 / 3.0) {
		/* equatorial region: edge lines of the faces */
		const int jp = (int) floor(HP_NS_MAX * (0.5 + tt - z * 0.75));
		const int jm = (int) floor(HP_NS_MAX * (0.5 + tt + z * 0.75));
		const int ifp = jp / HP_NS_MAX, ifm = jm / HP_NS_MAX;
		if (ifp == ifm) {
			face = ifp % 4 + 4;
		} else if (ifp < ifm) {
			face = ifp % 4;
		} else {
			face = ifm % 4 + 8;
		}
		ix = jm % HP_NS_MAX;
		iy = HP_NS_MAX - jp % HP_NS_MAX - 1;
	} else {
		/* polar caps */
		int ntt = (int) floor(tt);
		if (ntt >= 4) ntt = 3;
		const double tp = tt - ntt;
		const double tmp = sqrt(3.0 * (1.0 - za));
		int jp = (int) floor(HP_NS_MAX * tp * tmp);
		int jm = (int) floor(HP_NS_MAX * (1.0 - tp) * tmp);
		if (jp > HP_NS_MAX - 1) jp = HP_NS_MAX - 1;
		if (jm > HP_NS_MAX - 1) jm = HP_NS_MAX - 1;
		if (z >= 0) {
			face = ntt;
			ix = HP_NS_MAX - jm - 1;
			iy = HP_NS_MAX - jp - 1;
		} else {
			face = ntt + 8;
			ix = jp;
			iy = jm;
		}
	}
	/* nested index at nside 8192, degraded by integer division */
	const int64_t ipf = hp_spread(ix) + 2 * hp_spread(iy);
	const int64_t r = HP_NS_MAX / nside;
	return ipf / (r * r) + (int64_t) face * nside * nside;
}

/* centre of nested pixel ipix: ring index from the face's corner ring and the
 * in-face coordinates, then z and phi of the ring (HEALPix pix2ang_nest) */
void oc_pix2vec_nest(int64_t nside, int64_t ipix, double vec[3]) {
	static const int jrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
	static const int jpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
	const int64_t npface = nside * nside;
	const int face = (int) (ipix / npface);
	const int64_t ipf = ipix % npface;
	const int64_t ix = hp_compress(ipf), iy = hp_compress(ipf >> 1);
	const int64_t jr = (int64_t) jrll[face] * nside - ix - iy - 1;
	int64_t nr, kshift;
	double z;
	if (jr < nside) {
		nr = jr;
		z = 1.0 - (double) (nr * nr) / (3.0 * (double) npface);
		kshift = 0;
	} else if (jr > 3 * nside) {
		nr = 4 * nside - jr;
		z = -(1.0 - (double) (nr * nr) / (3.0 * (double) npface));
		kshift = 0;
	} else {
		nr = nside;
		z = (double) (2 * nside - jr) * 2.0 / (3.0 * (double) nside);
		kshift = (jr - nside) & 1;
	}
	int64_t jp = ((int64_t) jpll[face] * nr + ix - iy + 1 + kshift) / 2;
	if (jp > 4 * nr) jp -= 4 * nr;
	if (jp < 1) jp += 4 * nr;
	const double phi = ((double) jp - (double) (kshift + 1) * 0.5) * (0.5 * M_PI / (double) nr);
	const double st = sqrt((1.0 - z) * (1.0 + z));
	vec[0] = st * cos(phi);
	vec[1] = st * sin(phi);
	vec[2] = z;
}

/* ------------------------------------------------------------------ */
/* lightcone crossings of the drift (src/drift.cu:36-131)              */
/* ------------------------------------------------------------------ */

/* fixed<int,31> of a double: truncation of 2^31 x (cosmictiger/fixed.hpp:97-100) */
static int32_t lc_fixed31(double x) {
	return (int32_t) (2147483648.0 * x);
}

int64_t oc_lightcone(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, const float* vel, const int8_t* rungs, int rung,
		float a_, float dt_, float tau0_, float tau1_, float taumax_, int nside, oc_lc_entry* out, int64_t cap) {
	/* the float arguments of cuda_drift (cosmictiger/drift.hpp:54), in double
	 * inside the kernel */
	const double a = a_, dt = dt_, tau0 = tau0_, tau1 = tau1_, tau_max = taumax_;
	if (!(tau_max - tau1 < 1.0)) return 0;
	const double c0 = 1.0 / tau_max;
	const double ainv = 1.0 / a;
	const double c1 = dt / a;
	int64_t cnt = 0;
	for (int64_t i = 0; i < n; i++) {
		if (rungs[i] != rung) continue;
		const uint32_t* P[3] = {x + i, y + i, z + i};
		double p0[3], p1[3];
		for (int d = 0; d < 3; d++) {
			p0[d] = fixed_to_double(*P[d]);
			p1[d] = fma((double) vel[3 * i + d], c1, p0[d]); /* the drift's own sum, before the wrap */
		}
		for (int xi = -1; xi <= 0; xi++)
			for (int yi = -1; yi <= 0; yi++)
				for (int zi = -1; zi <= 0; zi++) {
					const double X0[3] = {p0[0] + xi, p0[1] + yi, p0[2] + zi};
					const double X1[3] = {p1[0] + xi, p1[1] + yi, p1[2] + zi};
					const double dist0 = sqrt(X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2]);
					const double dist1 = sqrt(X1[0] * X1[0] + X1[1] * X1[1] + X1[2] * X1[2]);
					const int i0 = (int) ((tau0 + dist0) * c0);
					const int i1 = (int) ((tau1 + dist1) * c0);
					if (i0 == i1 || !(dist1 <= 1.0 || dist0 <= 1.0)) continue;
					const float vx = (float) (vel[3 * i + 0] * ainv);
					const float vy = (float) (vel[3 * i + 1] * ainv);
					const float vz = (float) (vel[3 * i + 2] * ainv);
					const double ti = (double) (i0 + 1) * tau_max;
					const double A = 1.0 - (double) (vx * vx + vy * vy + vz * vz);
					const double B = 2.0 * ((tau0 - ti) - ((double) vx * X0[0] + (double) vy * X0[1] + (double) vz * X0[2]));
					const double C = (ti - tau0) * (ti - tau0) - (X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2]);
					const double t = -(B + sqrt(B * B - 4.0 * A * C)) / (2.0 * A);
					const double Xc[3] = {X0[0] + (double) vx * t, X0[1] + (double) vy * t, X0[2] + (double) vz * t};
					const double R2 = Xc[0] * Xc[0] + Xc[1] * Xc[1] + Xc[2] * Xc[2];
					if (!(R2 <= 1.0)) continue;
					if (cnt < cap) {
						oc_lc_entry e;
						e.x = lc_fixed31(Xc[0]);
						e.y = lc_fixed31(Xc[1]);
						e.z = lc_fixed31(Xc[2]);
						e.vx = vel[3 * i + 0];
						e.vy = vel[3 * i + 1];
						e.vz = vel[3 * i + 2];
						e.pix = (int32_t) oc_vec2pix_nest(nside, Xc);
						e.t = t;
						out[cnt] = e;
					}
					cnt++;
				}
	}
	return cnt;
}

void oc_vec2pix_nest_many(int64_t nside, int64_t n, const double* vec, int64_t* out) {
	for (int64_t i = 0; i < n; i++) out[i] = oc_vec2pix_nest(nside, vec + 3 * i);
}

void oc_pix2vec_nest_many(int64_t nside, int64_t n, const int64_t* pix, double* vec) {
	for (int64_t i = 0; i < n; i++) oc_pix2vec_nest(nside, pix[i], vec + 3 * i);
}

/*
 * CPU oracle for the CosmicTiger gravity kick -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker for the MI355X kick.  Only tests/, the smoke
 * check and bench.py's CPU-baseline leg may load it.  The product never calls
 * it, and it never calls the product.
 *
 * It restates, in plain C with FP64 arithmetic, the reference's algorithm on
 * the kick path:
 *   - fixed32 positions and minimum-image distance (cosmictiger/fixed.hpp:303-351)
 *   - tensor storage and the order-7 expansion operators, as generic loops
 *     over multi-indices (cosmictiger/tensor.hpp:63-319, cosmictiger/
 *     fmm_kernels.hpp, src/fmm_kernel_codegen.cpp)
 *   - Ewald tables and the Ewald Green's function (src/ewald_indices.cu:30-114,
 *     cosmictiger/fmm_kernels.hpp:310-1490)
 *   - the tree build (src/tree.cpp:273-601, src/particles.cpp:979-1018)
 *   - the kick walk and interactions (src/kick.cu:53-506, src/gravity.cu:44-404,
 *     cosmictiger/gravity.hpp:76-102)
 *   - the drift (src/drift.cu:36-154, narrowing at cosmictiger/drift.hpp:54)
 * and adds an independent FP64 direct-summation Ewald force/potential
 * (the physics oracle), written from the Ewald formulas.
 *
 * How each part is pinned is recorded in the header of ct_oracle.c.
 */
#ifndef CT_ORACLE_H
#define CT_ORACLE_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- tensors and expansions (double) ---------------- */
void oc_init(void);
int oc_trless_index(int a, int b, int c, int P);
double oc_full(const double* T, int P, int a, int b, int c);
void oc_greens(double D[50], const double X[3]);
void oc_ewald_greens(double D[50], const double X[3]);
void oc_m2l(double* L, int Pout, const double M[37], const double D[50], int do_phi);
void oc_l2l(double Lb[50], const double La[50], const double X[3], int do_phi);
void oc_l2p(double L4[4], const double La[50], const double X[3], int do_phi);
void oc_p2m(double M[37], const double X[3]);
void oc_m2m(double Mb[37], const double Ma[37], const double X[3]);
void oc_gsoft(double* f, double* phi, double q2, double hinv, double h2inv, double h3inv);
double oc_self_phi(void);
int oc_ewald_nreal(void);
int oc_ewald_nfour(void);
void oc_ewald_real_index(int i, double out[3]);
void oc_ewald_four_index(int i, double out[3]);

/* ---------------- fixed32 ---------------- */
float oc_distance(uint32_t a, uint32_t b);
uint32_t oc_fixed32_from_double(double x);

/* ---------------- tree ---------------- */
typedef struct {
	int32_t box_begin[3]; /* fixed<int,29> raw */
	int32_t box_end[3];
	int32_t child[2];     /* -1 for leaves */
	int32_t pbegin, pend;
	uint32_t pos[3];      /* fixed32 expansion centre */
	float radius;
	int32_t leaf;
	int32_t depth;
	double multi[37];
} oc_node;

typedef struct {
	oc_node* nodes;
	int64_t nnodes;
	int64_t cap;
	int root;
	int error; /* 0 ok, 1 depth exceeded, 2 arena full */
} oc_tree;

/* Builds the tree over particles [0, n).  Permutes x,y,z and id in place
 * (the partition of src/particles.cpp:979-1018).  Returns 0 on success. */
int oc_tree_build(oc_tree* t, int64_t n, uint32_t* x, uint32_t* y, uint32_t* z, int32_t* id, int bucket, double theta,
		double h, int nthreads, int64_t cap_hint);
void oc_tree_free(oc_tree* t);

/* ---------------- kick ---------------- */
/* Field order of kick_params (cosmictiger/kick.hpp:138-157). */
typedef struct {
	float sign;
	int min_level;
	int min_rung;
	float a;
	float t0;
	float theta;
	float h;
	float eta;
	float GM;
	bool vel_init;
	bool save_force;
	bool first_call;
	bool gpu;
	float node_load;
	float max_dt;
	bool ascending;
	bool descending;
	bool top;
	bool do_phi;
} oc_kick_params;

/* Field order of kick_return (cosmictiger/kick.hpp:92-100). */
typedef struct {
	int max_rung;
	double pot;
	double kin;
	double dkin;
	double xmom;
	double ymom;
	double zmom;
	double nmom;
} oc_kick_return;

/* Interaction lists of one kick, CSR by node: for node i the entries
 * [off[i], off[i+1]) of each list. */
typedef struct {
	int64_t nnodes;
	int64_t* off[5];
	int32_t* idx[5];
} oc_lists;
enum { OC_LIST_EWALD = 0, OC_LIST_CC = 1, OC_LIST_CP = 2, OC_LIST_PC = 3, OC_LIST_PP = 4 };

/* One kick of every particle below the root (the reference's top call,
 * src/driver.cpp:481-528).  vel is AoS float[3] per particle, rung int8.
 * gx..phi (optional, may be NULL) receive the forces in double.
 * lists (optional) receives the interaction lists.  Returns 0 or an error. */
int oc_kick(const oc_tree* t, const oc_kick_params* p, const uint32_t* x, const uint32_t* y, const uint32_t* z,
		float* vel, int8_t* rung, double* gx, double* gy, double* gz, double* phi, oc_kick_return* ret,
		oc_lists* lists, int nthreads);
/* As oc_kick, but only the sinks in the particle range [sample_b, sample_e)
 * are kicked (a bounded sample of the same walk, for CPU timing). */
int oc_kick_sample(const oc_tree* t, const oc_kick_params* p, const uint32_t* x, const uint32_t* y, const uint32_t* z,
		float* vel, int8_t* rung, double* gx, double* gy, double* gz, double* phi, oc_kick_return* ret, oc_lists* lists,
		int nthreads, int64_t sample_b, int64_t sample_e);
void oc_lists_free(oc_lists* l);

/* ---------------- lightcone ---------------- */
/* HEALPix nested pixel of a direction (src/healpix.cu:75-172), and the
 * centre of a pixel (HEALPix pix2ang_nest, for the round-trip pin). */
int64_t oc_vec2pix_nest(int64_t nside, const double vec[3]);
void oc_pix2vec_nest(int64_t nside, int64_t ipix, double vec[3]);
void oc_vec2pix_nest_many(int64_t nside, int64_t n, const double* vec, int64_t* out);
void oc_pix2vec_nest_many(int64_t nside, int64_t n, const int64_t* pix, double* vec);

/* lc_entry (cosmictiger/lightcone.hpp:54-58): position on the lightcone as
 * fixed<int,31> raw words, velocity, pixel; t is the crossing time
 * (diagnostic). */
typedef struct {
	int32_t x, y, z;
	float vx, vy, vz;
	int32_t pix;
	double t;
} oc_lc_entry;

/* Lightcone crossings of the drift of the particles of rung `rung`
 * (src/drift.cu:75-131, float arguments as cuda_drift takes them); the
 * drifted position is the drift's x0 + v c1 (one fma, before the wrap).
 * Writes up to cap entries, returns the number found. */
int64_t oc_lightcone(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, const float* vel, const int8_t* rungs, int rung,
		float a, float dt, float tau0, float tau1, float tau_max, int nside, oc_lc_entry* out, int64_t cap);

/* ---------------- drift ---------------- */
/* Drift particles [0,n) with rung == rung by dt (float narrowing as the
 * reference's cuda_drift signature).  Exact integer semantics. */
void oc_drift(int64_t n, uint32_t* x, uint32_t* y, uint32_t* z, const float* vel, const int8_t* rungs, int rung, float a,
		float dt);

/* ---------------- FP64 direct Ewald (physics oracle) ---------------- */
/* Periodic force and potential on the sinks from all n sources (unit masses,
 * times GM), with Ferrers n=4 softening h for near pairs and the self term
 * convention of the kick.  Output in double. */
void oc_direct_ewald(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, int64_t nsink,
		const int64_t* sinks, double h, double GM, double* gx, double* gy, double* gz, double* phi, int nthreads);
/* Zero-mean periodic 1/r, psi(X) (X != 0), and its gradient; alpha-independent. */
void oc_ewald_psi(const double X[3], double* psi, double grad[3]);
/* lim_{r->0} psi(r) - 1/r */
double oc_ewald_psi0(void);

#ifdef __cplusplus
}
#endif
#endif

/*
 * C-ABI of the MI355X-native CosmicTiger gravity kick (libcosmictiger_amd.so).
 *
 * Plain C types only: pointers, sizes, POD structs.  Every entry point
 * returns CT_OK (0) or a negative CT_E_* code; ct_last_error() gives the
 * message of the last failure on the calling thread.  No exception crosses
 * this boundary.  A box handle may be used by one host thread at a time;
 * different boxes are independent.
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef COSMICTIGER_AMD_H
#define COSMICTIGER_AMD_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CT_OK 0
#define CT_E_ARG (-1)   /* invalid argument */
#define CT_E_OOM (-2)   /* device or host allocation failed */
#define CT_E_HIP (-3)   /* HIP runtime error */
#define CT_E_ARENA (-4) /* interaction-list arena overflow that retries could not absorb */
#define CT_E_STATE (-5) /* call out of order (e.g. kick before tree build) */
#define CT_E_TREE (-6)  /* tree build failed (depth limit) */
#define CT_E_WALK (-7)  /* walk invariant violated (a leaf failed the Ewald test) */

/* Mirror of kick_params, same fields in the same order
 * (cosmictiger/kick.hpp:138-157). */
typedef struct ct_kick_params {
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
} ct_kick_params;

/* Mirror of kick_return (cosmictiger/kick.hpp:92-100). */
typedef struct ct_kick_return {
	int max_rung;
	double pot;
	double kin;
	double dkin;
	double xmom;
	double ymom;
	double zmom;
	double nmom;
} ct_kick_return;

/* Node as read back by ct_tree_read (the fields of tree_node,
 * cosmictiger/tree.hpp:98-169, with the multipole of multi_pos inline). */
typedef struct ct_node {
	int32_t box_begin[3]; /* fixed<int,29> raw: units of 2^-29 */
	int32_t box_end[3];
	int32_t child[2]; /* -1 for leaves */
	int32_t parent;   /* -1 for the root */
	int32_t pbegin;   /* part_range [pbegin, pend) */
	int32_t pend;
	uint32_t pos[3]; /* fixed32 expansion centre */
	float radius;
	int32_t leaf;
	int32_t depth;
	float multi[37]; /* multipole<float> (trless order 6, scaled frame) */
} ct_node;

typedef struct ct_box* ct_box_t;

/* Select the device and build the Ewald tables.  Replaces cuda_init /
 * cuda_set_device (cosmictiger/cuda.hpp:124-130) and ewald_const::init()
 * (cosmictiger/ewald_indices.hpp:32-41, src/ewald_indices.cu:30-114). */
int ct_init(int device);

/* Rebuild the Ewald tables (ewald_const::init_gpu, src/ewald_indices.cu:30). */
int ct_ewald_init(void);

/* Create a resident box from host SoA positions (fixed32), AoS velocities
 * float[3] and rungs.  The library owns the device copies.  Replaces the
 * particle store setup (particles_resize / the globals particles_x/v/r,
 * cosmictiger/particles.hpp:159-163, src/particles.cpp:774-839). */
int ct_box_create(int64_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, const float* vel, const int8_t* rung,
		ct_box_t* out);

/* Free a box (particles_destroy / tree_destroy). */
int ct_box_destroy(ct_box_t box);

/* Build the tree over the box's active range (all particles unless
 * ct_box_sort_by_rung narrowed it) with opening angle theta, leaf bucket size
 * and softening hsoft; permutes the particles of the range as the reference's
 * partition does.  Replaces tree_create (cosmictiger/tree.hpp:207,
 * src/tree.cpp:273, whose root covers particles_current_range(), :292). */
int ct_tree_build(ct_box_t box, double theta, int bucket, double hsoft);

/* Where ct_tree_build_ex builds the tree: AUTO = on the device (particles
 * sorted by Morton key, which is the order of the reference's bisections),
 * falling back to the host for trees deeper than 63 levels; HOST = the host
 * build (in-place partition as src/particles.cpp:979-1018); DEVICE = device
 * only (CT_E_TREE when too deep).  Both give the same nodes (boxes, centres,
 * radii, particle ranges, topology); the order of particles inside a leaf
 * differs. */
#define CT_TREE_AUTO 0
#define CT_TREE_HOST 1
#define CT_TREE_DEVICE 2
int ct_tree_build_ex(ct_box_t box, double theta, int bucket, double hsoft, int where);

/* Host-only tree build over caller arrays (no device needed): the same build
 * as ct_tree_build.  Permutes x, y, z in place; perm[i] receives the original
 * index of the particle now at i.  Writes up to node_cap nodes (breadth-first,
 * root = 0) and the node count to *nnodes; returns CT_E_ARG with *nnodes set
 * when node_cap is too small.  Replaces tree_create (src/tree.cpp:273). */
int ct_tree_build_host(int64_t n, uint32_t* x, uint32_t* y, uint32_t* z, int32_t* perm, double theta, int bucket, double hsoft,
		ct_node* nodes, int64_t node_cap, int64_t* nnodes);

/* One kick of every particle of the box from the root, i.e. the reference's
 * kick(params, L=0, pos=0, root, {root}, {root}, nullptr)
 * (cosmictiger/kick.hpp:199, src/kick.cpp:150; device part src/kick.cu:508
 * cuda_execute_kicks).  stream: a hipStream_t or NULL for the default stream.
 * The call is synchronous with respect to the returned kick_return. */
int ct_kick(ct_box_t box, const ct_kick_params* params, ct_kick_return* ret, void* stream);

/* Arithmetic of the interaction kernels (the walk and its acceptance tests
 * are the same in all modes). */
#define CT_PREC_F32C 0 /* FP32, FP64 sink accumulation per source tile (ct_kick) */
#define CT_PREC_F32U 1 /* FP32, FP32 running sums */
#define CT_PREC_F64 2  /* FP64 multipoles, expansions, Ewald and P2P */

/* ct_kick in a chosen arithmetic mode, for the precision sweep.  No
 * counterpart in the reference, whose kick is FP32 only (src/gravity.cu). */
int ct_kick_prec(ct_box_t box, const ct_kick_params* params, int precision, ct_kick_return* ret, void* stream);

/* A batch of resident boxes kicked together: every phase of the kick runs as
 * one launch over all the boxes (one walk per tree level over every box's
 * sinks, one cell launch, one P2P launch per leaf-size class, ...).  The
 * reference hands the GPU a list of subtrees of one box in one persistent
 * launch (kick_workitem, cosmictiger/kick.hpp:188-194; cuda_execute_kicks,
 * src/kick.cu:508-609); a batch extends that list to subtrees of many boxes.
 * A box may belong to several batches but must not be kicked concurrently. */
typedef struct ct_batch* ct_batch_t;
int ct_batch_create(const ct_box_t* boxes, int nbox, ct_batch_t* out);
int ct_batch_destroy(ct_batch_t batch);

/* Point an existing batch at another set of boxes, keeping its walk arenas,
 * list buffers and stream (a serving loop steps a different set of resident
 * boxes each time without reallocating).  The next kick rebuilds the batch's
 * index spaces.  No counterpart in the reference. */
int ct_batch_reset(ct_batch_t batch, const ct_box_t* boxes, int nbox);

/* One kick of every box of the batch with params[i] for box i; rets[i]
 * receives box i's kick_return.  use_graph != 0 replays a captured HIP graph
 * of the whole kick (captured on the first call after the batch's trees or
 * arenas change; a replay whose walk overflowed its arenas is redone eagerly). */
int ct_batch_kick(ct_batch_t batch, const ct_kick_params* params, int precision, ct_kick_return* rets, int use_graph, void* stream);

/* A kick of every box of the batch followed by the batch drift (rung, a[i],
 * dt[i], drag[i] as ct_batch_drift), in one stream; with use_graph != 0 the
 * kick and the drift are captured in one HIP graph and replayed.  A replay
 * whose walk overflowed its arenas moves no particle and is redone eagerly.
 * The serving form of one step of every box (kick then drift,
 * src/driver.cpp:528,562). */
int ct_batch_step(ct_batch_t batch, const ct_kick_params* params, int precision, ct_kick_return* rets, int rung, const float* a,
		const float* dt, const double* drag, int use_graph, void* stream);

/* Drift of every box of the batch in one launch (a[i], dt[i], drag[i] for
 * box i; drag may be NULL), with cuda_drift's arithmetic (src/drift.cu:36-154). */
int ct_batch_drift(ct_batch_t batch, int rung, const float* a, const float* dt, const double* drag, void* stream);

/* Phase times and event counts of the batch's last kick, as ct_kick_stats. */
int ct_batch_stats(ct_batch_t batch, double* out, int n);

/* Drift the particles of rung `rung` in the active range by dt at scale
 * factor a (float arguments as the reference narrows them), with optional
 * drag.  Replaces
 * cuda_drift(char, float a, float dt, ...) (cosmictiger/drift.hpp:54,
 * src/drift.cu:156) as called by drift() (src/drift.cpp:33). */
int ct_drift(ct_box_t box, int rung, float a, float dt, double drag, void* stream);

/* ---- lightcone ---- */

/* One particle crossing the lightcone: lc_entry (cosmictiger/lightcone.hpp:
 * 54-58).  x, y, z are the crossing position relative to the observer at the
 * origin as fixed<int,31> raw words (position * 2^31, truncated); vx, vy, vz
 * the particle's velocity; pix its HEALPix nested pixel. */
typedef struct {
	int32_t x, y, z;
	float vx, vy, vz;
	int32_t pix;
} ct_lc_entry;

/* The drift of ct_drift with the lightcone of cuda_drift_kernel
 * (src/drift.cu:36-154, do_lc = true): for the particles of rung `rung` in the
 * active range, every periodic image (offsets -1 and 0 per axis) whose
 * conformal-time shell index floor((tau + |X|) / tau_max) changes over the step
 * and that lies within the unit sphere yields an entry at its crossing of the
 * light shell.  Only in the last unit of conformal time (tau_max - tau1 < 1),
 * as in the reference.  nside: HEALPix resolution, a power of 2 in
 * [1, 8192].  The entries stay with the box (ct_lc_read); *nentries (may be
 * NULL) receives their number.  a, dt, tau0, tau1, tau_max are float as
 * cuda_drift takes them (cosmictiger/drift.hpp:54). */
int ct_drift_lightcone(ct_box_t box, int rung, float a, float dt, float tau0, float tau1, float tau_max, int nside, double drag,
		int64_t* nentries, void* stream);

/* Copies min(cap, count) entries of the box's last ct_drift_lightcone to out
 * (host memory, in no particular order); *count (may be NULL) receives their
 * number. */
int ct_lc_read(ct_box_t box, ct_lc_entry* out, int64_t cap, int64_t* count);

/* ---- active range and rungs ----
 * A box keeps an active range [rung_begin, N): tree builds, kicks and drifts
 * cover it (particles_current_range, src/particles.cpp:1020-1025).  It starts
 * as the whole box. */

/* Partition the active range so that particles with rung >= min_rung come
 * last and make them the active range; min_rung = 0 makes the whole box
 * active without moving anything.  *rung_begin (may be NULL) receives the new
 * start.  The order inside each part is not the reference's (compare by id);
 * the saved forces are not moved, as in the reference.  The box's tree is
 * invalidated when particles move.  Replaces particles_sort_by_rung
 * (src/particles.cpp:1029-1068). */
int ct_box_sort_by_rung(ct_box_t box, int min_rung, int64_t* rung_begin);

/* The active range [*first, *second).  Replaces particles_current_range
 * (src/particles.cpp:1020-1025). */
int ct_box_current_range(ct_box_t box, int64_t* first, int64_t* second);

/* Save / restore the start of the active range.  Replace particles_push_rungs
 * and particles_pop_rungs (src/particles.cpp:218-235); popping an empty stack
 * is CT_E_STATE. */
int ct_box_push_rungs(ct_box_t box);
int ct_box_pop_rungs(ct_box_t box);

/* rung = max(rung, min_rung) for every particle.  Replaces
 * particles_set_minrung (src/particles.cpp:628-649). */
int ct_box_set_minrung(ct_box_t box, int min_rung);

/* counts[r] = number of particles on rung r for r < cap; *nrungs (may be
 * NULL) receives 1 + the highest rung present (0 for an empty box).  Replaces
 * particles_rung_counts (src/particles.cpp:583-624). */
int ct_box_rung_counts(ct_box_t box, int64_t* counts, int cap, int* nrungs);

/* One call of the hierarchical step (the kick ladder of
 * kick_step_hierarchical, src/driver.cpp:382-593): ascending kicks from
 * max_rung down to minrung, the top kick at minrung, then descending kicks
 * and drifts of every rung the top kick and those after it raise, with the
 * reference's tree builds over the active range, its rung sorts and its
 * push/pop of the range, and its clipping of the top rung.  The expansion
 * factor stays `scale` over the step (the static-box case; the reference's
 * cosmological a(t) update, driver.cpp:561, belongs to its cosmology). */
typedef struct ct_step_params {
	int minrung;   /* min_rung(itime) raised to minrung0 (driver.cpp:990) */
	int max_rung;  /* max_rung of the previous call */
	int minrung0;  /* the run's lowest rung */
	double scale;  /* a */
	double tau;    /* time at the start of the call; first_call = (tau == 0) */
	double t0;     /* top-level step */
	double theta;
	int bucket;
	double hsoft;
	double GM;
	double eta;
	double max_dt; /* SCALE_DT / adot (driver.cpp:501) */
	double sign;   /* 1, or -1 for glass creation */
	int do_phi;    /* full_eval */
	int nocrop;
	double drag;
	int save_force;
	int precision; /* CT_PREC_* */
	int tree_where;/* CT_TREE_* */
} ct_step_params;

typedef struct ct_step_return {
	ct_kick_return kr; /* the top kick's kick_return, max_rung = the new max rung */
	int minrung;       /* minrung after the call (raised by one when the top was clipped) */
	int nkicks;        /* kicks (tree builds) made */
	int clip_top;
	double tckin;      /* sum of dkin over the kicks (energies->tckin) */
	double parts_processed;
} ct_step_return;

int ct_step_hierarchical(ct_box_t box, const ct_step_params* params, ct_step_return* out, void* stream);

/* The integer time of the driver: inc() and min_rung() of
 * cosmictiger/time.hpp:28-43. */
uint64_t ct_time_inc(uint64_t t, int max_rung);
int ct_time_min_rung(uint64_t t);

/* ---- persistence ----
 * A box in the reference's particle file layout (particles_save /
 * particles_load, src/particles.cpp:1182-1227): int64 count; x, y, z
 * (uint32 each); velocities (float[3] each); rungs (int8); int32 size of the
 * active-range stack, int32 rung_begin, the stack (int32 each).  Particles are
 * written in the box's current order; no tree is stored (a kick needs a tree
 * build after a load, as after the reference's read_checkpoint). */

/* Write the box to path (append != 0 appends, e.g. after a checkpoint header
 * such as write_checkpoint's driver_params, src/driver.cpp:1161). */
int ct_box_save(ct_box_t box, const char* path, int append);

/* Create a box from the record at byte offset of path; *end_offset (may be
 * NULL) receives the offset just past it.  Particle ids restart at 0 in file
 * order. */
int ct_box_load(const char* path, int64_t offset, ct_box_t* out, int64_t* end_offset);

/* Number of particles and nodes of a box. */
int ct_box_size(ct_box_t box, int64_t* nparts, int64_t* nnodes);

/* Copy the box back to the host in the current (tree) order; any pointer may
 * be NULL.  id[i] is the index the particle had in ct_box_create.  Replaces
 * the particle accessors (cosmictiger/particles.hpp:221-306). */
int ct_box_read(ct_box_t box, uint32_t* x, uint32_t* y, uint32_t* z, float* vel, int8_t* rung, float* gx, float* gy, float* gz,
		float* phi, int32_t* id);

/* Copy the node table (breadth-first order, root = 0) to the host.
 * Replaces tree_get_node / tree_data (cosmictiger/tree.hpp:213,224). */
int ct_tree_read(ct_box_t box, ct_node* nodes, int64_t cap);

/* Interaction lists of the last kick, CSR by node, for the list kinds
 * 0 = Ewald CC, 1 = direct CC, 2 = CP, 3 = PC, 4 = PP (the lists of
 * src/kick.cu:269-423).  off has nnodes+1 entries; idx capacity idx_cap. */
int ct_lists_read(ct_box_t box, int kind, int64_t* off, int32_t* idx, int64_t idx_cap);

/* Per-phase device times (ms) of the last kick, measured with HIP events on
 * the kick's stream: [0] walk, [1] cell interactions (M2L, Ewald, P2L),
 * [2] L2L sweep, [3] leaf kernel (P2P, M2P, L2P, kick), [4] whole kick.
 * Counters: [5] PP pairs, [6] direct CC, [7] Ewald CC, [8] CP particles,
 * [9] PC sink-multipoles, [10] node visits, [11] sinks; [12] P2P kernels,
 * [13] M2P + L2P + kick kernels (ms; [3] = [12] + [13]); [14] checks
 * classified by the walk. */
int ct_kick_stats(ct_box_t box, double* out, int n);

/* Host evaluation of the expansion kernels in double, for checks of the
 * layout: op 0 greens (in X[3] -> out D[50]), 1 ewald_greens (X -> D),
 * 2 M2L (in M[37] D[50] -> L[50]), 3 L2L (La[50] X[3] -> Lb[50]),
 * 4 L2P (La[50] X[3] -> L[4]), 5 P2M (X -> M[37]), 6 M2M (Ma[37] X[3] -> Mb[37]),
 * 7 M2L order 2 (M[37] D[50] -> L[4]).  These are the templates of
 * cosmictiger/fmm_kernels.hpp instantiated for double. */
int ct_fmm_host(int op, const double* in, double* out);

/* The same kernels on the device in float, for n independent inputs
 * (in: n * in_stride doubles, out: n * out_stride doubles, host memory). */
int ct_fmm_device(int op, int64_t n, const double* in, double* out);

/* FP64 direct-summation periodic Ewald force and potential at nsink sink
 * positions (fixed32) from every particle of the box, unit mass per particle
 * (multiply by GM), pairs at zero separation excluded, Ferrers softening h.
 * Host arrays in and out.  Replaces the analytic force check
 * gravity_analytic_call_kernel (src/analytic.cu:30-116) used by
 * analytic_compare (src/analytic.cpp:59-113), in FP64 throughout. */
int ct_direct_ewald(ct_box_t box, int64_t nsink, const uint32_t* sinkx, const uint32_t* sinky, const uint32_t* sinkz, double h,
		double* phi, double* gx, double* gy, double* gz, void* stream);

const char* ct_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

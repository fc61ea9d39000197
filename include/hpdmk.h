/* hpdmk.h -- C ABI of the MI355X DMK Monte Carlo library (libhpdmk_amd.so).
 *
 * Drop-in for the reference's hpdmk surface (ArrogantGao/PDMK4MC):
 *   include/hpdmk.h:6-25      HPDMKParams, hpdmk_init, hpdmk_tree, hpdmk_tree_create
 *   include/tree.hpp:98-132   HPDMKPtTree<Real> members used on the Monte Carlo
 *                             path (form_outgoing_pw, form_incoming_pw,
 *                             eval_energy, eval_shift_energy, update_shift)
 * The reference's tree is a C++ class; here it is an opaque handle, every call
 * returns an int status (0 = HPDMK_OK) and never throws, and the message of the
 * last failure on the calling thread is hpdmk_last_error().
 *
 * Coordinates are N x 3 interleaved doubles in [0, L); charges are N doubles;
 * particle indices are the caller's original indices (the reference maps them
 * through indices_invmap, src/energy_shift.cpp:21).  The caller keeps
 * ownership of all input arrays, which are copied to HBM at create.
 *
 * One handle = one thread at a time.  Calls are synchronous with respect to
 * the host unless stated otherwise.
 */
#ifndef HPDMK_H
#define HPDMK_H

#ifdef __cplusplus
extern "C" {
#endif

/* include/hpdmk.h:6-9 */
typedef enum { DIRECT = 1, PROXY = 2 } hpdmk_init;

/* include/hpdmk.h:11-17.  C has no default member initialisers: start from
 * hpdmk_default_params(), which fills the reference's defaults
 * (n_per_leaf 200, digits 3, prolate_order 16, init PROXY). */
typedef struct HPDMKParams {
    int n_per_leaf;
    int digits;
    double L;
    double prolate_order;
    hpdmk_init init;
} HPDMKParams;

/* include/hpdmk.h:19 */
typedef void* hpdmk_tree;

/* arithmetic of the device state (no reference counterpart: the reference is
 * templated on float/double, src/tree.cpp:599-600) */
typedef enum {
    HPDMK_FP64 = 0,  /* all-FP64 expansions and sums */
    HPDMK_MIXED = 1, /* FP32 bulk, FP64 phase arguments and accumulation */
} hpdmk_precision;

/* status codes */
#define HPDMK_OK 0
#define HPDMK_EINVAL (-1)    /* bad argument (digits, coordinates, sizes, handle) */
#define HPDMK_ERUNTIME (-2)  /* method error the reference throws (tree too shallow, ...) */
#define HPDMK_EDEVICE (-3)   /* HIP runtime error */
#define HPDMK_ECAPACITY (-4) /* an accepted move would overflow a leaf's slot capacity */
#define HPDMK_ENOTIMPL (-5)  /* PROXY above 3 digits (src/upward_pass.cpp:9-14) */

HPDMKParams hpdmk_default_params(void);
const char* hpdmk_last_error(void);

/* Builds the octree and kernel tables and uploads everything to HBM.
 * Replaces include/hpdmk.h:25 (which returns void and discards the tree) and
 * the HPDMKPtTree constructor (src/tree.cpp:268-444).  The vestigial MPI_Comm
 * argument is dropped (the reference runs single-rank, src/tree.cpp:266).
 * Precision is HPDMK_FP64. */
int hpdmk_tree_create(HPDMKParams params, int n_src, const double* r_src, const double* charge,
                      hpdmk_tree* tree);
/* Same with an explicit precision (hpdmk_precision). */
int hpdmk_tree_create_ex(HPDMKParams params, int precision, int n_src, const double* r_src,
                         const double* charge, hpdmk_tree* tree);
int hpdmk_tree_destroy(hpdmk_tree tree);

/* info[10]: n_nodes, n_levels, n_window, n_diff, window elements, level
 * elements, boxes carrying expansions, max near-list length, leaves, slots */
int hpdmk_tree_info(hpdmk_tree tree, int* info);

/* src/planewave.cpp:86-97 (root expansion exact, not NUFFT) */
int hpdmk_form_outgoing_pw(hpdmk_tree tree);
/* src/planewave.cpp:99-155 */
int hpdmk_form_incoming_pw(hpdmk_tree tree);
/* The expansions the reference keeps in HPDMKPtTree::outgoing_pw and
 * incoming_pw (include/tree.hpp:86-87) and its tests read directly
 * (test/hpdmk_test.cpp:43, 84): which 0 the root outgoing expansion, 1 the
 * outgoing and 2 the incoming expansion of node `node` (Morton pre-order, a
 * box of levels 2..n_levels-2).  Copies the half grid (index i + j d + k d^2,
 * k_z <= 0) into out as interleaved re/im doubles; *count = complex elements. */
int hpdmk_get_expansion(hpdmk_tree tree, int which, int node, double* out, int* count);
/* Rebuilds the octree of a system from its current particle positions into a
 * new handle: fresh topology, leaf slots and box lists, expansions formed when
 * the old system had them formed, and the Monte Carlo settings carried over
 * (model, proposal with its MLP weights, launch shape, profiling level; move
 * lists are not: load them again).  The old handle stays valid; destroy it
 * with hpdmk_tree_destroy.  No counterpart in the reference, which never
 * re-refines (src/tree.cpp:447-597 keeps the topology of the first build):
 * this is the remedy for HPDMK_ECAPACITY after long runs and gives the energy
 * of a fresh build to check the incremental one against. */
int hpdmk_tree_rebuild(hpdmk_tree tree, hpdmk_tree* rebuilt);
/* include/tree.hpp:105 eval_energy; parts (optional, 3 doubles): window,
 * difference, residual (src/energy.cpp:21-173) */
int hpdmk_eval_energy(hpdmk_tree tree, double* energy, double* parts);

/* src/energy_shift.cpp:18-70: energy change of moving particle i by (dx,dy,dz)
 * (target wrapped into the box).  parts (optional, 5 doubles): window,
 * difference, residual at target, residual at origin, self correction. */
int hpdmk_eval_shift_energy(hpdmk_tree tree, long i, double dx, double dy, double dz, double* dE,
                            double* parts);
/* src/tree.cpp:447-597: applies the move.  Unlike the reference it does not
 * depend on a preceding eval_shift_energy call with the same arguments. */
int hpdmk_update_shift(hpdmk_tree tree, long i, double dx, double dy, double dz);

/* current coordinates in the caller's original order (N x 3) */
int hpdmk_get_positions(hpdmk_tree tree, double* r_src);

/* ---- extensions without a reference counterpart ----------------------- */

/* n independent trial moves evaluated against the same state (no update). */
int hpdmk_eval_shift_energy_batch(hpdmk_tree tree, int n, const long* idx, const double* d, double* dE);
/* The same without any device allocation: no counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The move list and the
 * energies are staged in a fixed device array of the library (144 KiB per
 * GPU, held until the process ends), so n is at most 4096 (HPDMK_EINVAL
 * beyond).  Every call on one GPU shares that array: calls that overlap on
 * one GPU (several threads, or several handles on different streams) race
 * and must be serialised by the caller.  The handle's own state was
 * allocated at create. */
int hpdmk_eval_shift_energy_batch_noalloc(hpdmk_tree tree, int n, const long* idx, const double* d, double* dE);

/* Metropolis chain on the device.  Move k proposes particle idx[k] displaced
 * by d[3k..3k+2] and is accepted when u[k] < exp(-beta dE).
 *   mode 0 (plain):   one trial kernel and one update kernel per move
 *   mode 1 (batched): `batch` trials per step evaluated together, their mutual
 *                     effects added exactly, decisions taken in order; the
 *                     chain is identical to the plain one up to roundoff.
 * hpdmk_mc_load copies the move list to HBM; hpdmk_mc_run runs it without
 * host transfers and returns the GPU time (ms) between events around the
 * run; hpdmk_mc_results copies dE and acceptance back. */
int hpdmk_mc_load(hpdmk_tree tree, int n_moves, const long* idx, const double* d, const double* u);
/* Model of later hpdmk_mc_run calls (no reference counterpart).  Energies
 * enter the decision multiplied by the Bjerrum length `bjerrum` (the
 * reference's Ewald uses eps = 1 / l_B for the same scaling, src/ewald.cpp:187),
 * and a trial whose target lies closer than `sigma` to another particle
 * (minimum image) is rejected as a hard-sphere overlap; its reported dE is
 * +infinity.  Reported chain dE values are bjerrum x dE.  Defaults 1 and 0
 * (no hard core); sigma may not exceed the finest box size. */
int hpdmk_mc_set_model(hpdmk_tree tree, double bjerrum, double sigma);
int hpdmk_mc_run(hpdmk_tree tree, double beta, int mode, int batch, double* gpu_ms);  /* batch 0: tuned */
int hpdmk_mc_results(hpdmk_tree tree, double* dE, int* accept);

/* Per-move counts of the last profiled run (8 ints per move): difference levels read
 * at the target and origin, near-field particles scanned at the target and
 * origin, expansion arrays an acceptance rewrites for the target and origin
 * paths, the difference-level terms of the batch's pair corrections that this
 * move received as the later trial, and (on the first move of each batch or
 * single-move step that accepts anything, else 0) the distinct expansion
 * arrays that step rewrites.  They size the algorithmic bytes and flops of
 * each move. */
int hpdmk_mc_stats(hpdmk_tree tree, int* stats);

/* Profiling level of later hpdmk_mc_run calls: 0 off; 1 kernel times
 * (start/end timestamps of every dispatch) and per-move counts; 2 also the
 * device-side phase timers, which slow the kernels that carry them.
 * hpdmk_kernel_times: per-kernel GPU time of the last profiled run, kinds
 * 0 trial, 1 update (apply), 2 pair corrections, 3 decision scan, 4 trial
 * set-up, 5 proposal (features, learned displacements).  ms[6], count[6]. */
int hpdmk_set_profiling(hpdmk_tree tree, int level);
int hpdmk_kernel_times(hpdmk_tree tree, double* ms, int* count);

/* Average duration (us) of device-side phases of the last profiled run:
 * 0 batch scan assembly, 1 batch decisions, 2 update marking, 3 update
 * compaction, 4 particle relocation, 5 single-move finalisation, 6 update
 * contribution lists, 7 update accumulation, 8 update write-back (per work
 * item), 9 batch decisions' set-up (masks, known outcomes), 10 the walk over
 * the batch, 11 the union of the accepted trials' update arrays (part of 2).
 * us[12], count[12]. */
int hpdmk_phase_times(hpdmk_tree tree, double* us, int* count);

/* Host-only inspection (no device): the octree the library would build.
 * Returns the node count through *n_nodes; when nodes is non-NULL it must
 * hold 9 ints per node, written in Morton pre-order as: level, is_leaf,
 * parent, cell x, cell y, cell z, particles, same-level neighbours, near-list
 * length.  particle_leaf (optional, n_src ints) receives each particle's leaf. */
/* src/utils.cpp:14-28 read_particle_info: a text file whose first line is a
 * header and whose other lines hold "x y z q" (lines that do not parse as four
 * numbers are skipped).  Writes the first min(capacity, count) particles as
 * r_src (N x 3) and charge, and the file's particle count to *count (call with
 * capacity 0 to size the arrays).  Host only; no GPU needed. */
int hpdmk_read_particle_info(const char* path, int capacity, double* r_src, double* charge, int* count);

int hpdmk_plan_tree(HPDMKParams params, int n_src, const double* r_src, int* n_nodes, int* nodes,
                    int* particle_leaf);
/* The same octree built on the GPU (Morton keys, radix sort, refinement,
 * periodic 2:1 balance, neighbour and near lists; the build hpdmk_tree_create
 * uses), node for node equal to hpdmk_plan_tree's.  *ms (optional): host wall
 * time of the build including the copies.  HPDMK_ENOTIMPL for trees of more
 * than 20 levels (hpdmk_tree_create then builds them on the host). */
int hpdmk_plan_tree_device(HPDMKParams params, int n_src, const double* r_src, int* n_nodes, int* nodes,
                           int* particle_leaf, double* ms);
/* PSWF constants of a digits setting: c, lambda, C0, diff0, psi0(0),
 * n_window, n_diff (out[7]). */
int hpdmk_pswf_constants(int digits, double* out);

/* ---- proposals (no reference counterpart) ------------------------------
 * learned = 0: hpdmk_mc_load's d holds the displacements (uniform or any
 * symmetric proposal).  learned = 1: d holds standard normal triples z and the
 * displacement is delta = mu + sigma z, with mu = max_mean tanh(o[0:3]),
 * sigma = sigma_min + (sigma_max - sigma_min) sigmoid(o[3:6]) and o the output
 * of an MLP 8 -> 64 -> 64 -> 64 -> 6 (tanh hidden layers; float weights
 * W1[8][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][6] b4[6], 9286
 * values) applied to features of the particle's neighbourhood within radius
 * rf (the particle itself removed): its charge, the softened local field (3),
 * the neighbour count and the crowding direction (3).  Inference runs on the
 * matrix cores with BF16 operands and FP32 accumulation.  The decision
 * adds the exact proposal ratio: accept iff bjerrum dE - H / beta < -ln(u) /
 * beta, H = log q(-delta | features at the target) - log q(delta | features
 * at the origin).  rf > 0 with learned = 0 records the origin features of
 * every move (training data).  In the batched chain a step decides its
 * trials up to the first whose proposal an earlier accepted move of the step
 * made stale (same particle, or an endpoint within rf of its endpoints); the
 * rest are proposed again from the new state, so the chain is the plain one. */
int hpdmk_mc_set_proposal(hpdmk_tree tree, int learned, const float* weights, double rf, double max_mean,
                          double sigma_min, double sigma_max);
/* Of the last hpdmk_mc_run (each optional): delta[3n] displacements used,
 * feat[8n] features at the origins (rf > 0), logratio[n] H (learned). */
int hpdmk_mc_proposal_results(hpdmk_tree tree, double* delta, double* feat, double* logratio);
/* Of the last hpdmk_mc_run with the learned proposal: feat_t[8n] the features
 * at each move's target (the particle removed, the rest as the move saw
 * them), i.e. the input of its reverse-move density (training data of the
 * policy gradient through H).  No reference counterpart. */
int hpdmk_mc_target_features(hpdmk_tree tree, double* feat_t);

/* Proposal MLP on the GPU (no reference counterpart).  Samples: features
 * feat[n][8] (as recorded by hpdmk_mc_proposal_results), displacements
 * delta[n][3] and weights weight[n].
 * hpdmk_mlp_eval: mu_sigma[n][6] (optional) from the deployed BF16
 * matrix-core inference, grad[9286] (optional) = the FP32 sum over samples of
 * weight x d NLL / d w with NLL = sum_a log sigma_a + (delta_a - mu_a)^2 /
 * (2 sigma_a^2), and *loss (optional) = the weighted NLL sum.  cfg3:
 * max_mean, sigma_min, sigma_max.
 * hpdmk_mlp_train: Adam on minibatches of the weighted NLL; cfg[8]:
 * max_mean, sigma_min, sigma_max, learning rate, epochs, minibatch, seed,
 * normalisation (0: by the minibatch's weight sum, for 0/1 likelihood
 * weights; 1: by its sample count, for signed weights such as a reward minus
 * a baseline, whose descent is the policy gradient of the mean reward);
 * weights in/out; epoch_loss[epochs] (optional) = mean NLL per epoch. */
int hpdmk_mlp_eval(int n, const float* weights, const double* feat, const double* delta, const double* weight,
                   const double* cfg3, double* mu_sigma, float* grad, double* loss);
int hpdmk_mlp_train(int n, const double* feat, const double* delta, const double* weight, const double* cfg,
                    float* weights, double* epoch_loss);

/* ---- launch shape and tile autotuner (no reference counterpart) --------
 * tiles[4]: blocks per trial for the window sum, per (endpoint, difference
 * level), per endpoint's near field (each 1..32), and blocks of the update
 * kernel; batch: trials per batched step used when hpdmk_mc_run gets batch 0.
 * The chain's results do not depend on them (up to FP summation order). */
int hpdmk_set_tiles(hpdmk_tree tree, const int* tiles, int batch);
/* out[5]: the four tiles and the batch size */
int hpdmk_get_tiles(hpdmk_tree tree, int* out);
/* Times the batched chain on a twin of the current state (this system is
 * not changed) over the given sample moves, searching batch size then each
 * tile in turn, keeps the fastest shape (hpdmk_get_tiles) and returns it in
 * out[5] and its GPU time in *best_ms.  Each timing run applies the sample's
 * displacements and uniforms to other particles (indices shifted), so that no
 * run finds the previous run's data in cache. */
int hpdmk_autotune(hpdmk_tree tree, int n_moves, const long* idx, const double* d, const double* u, double beta,
                   int* out, double* best_ms);

/* ---- replica sets (no reference counterpart) --------------------------
 * Many independent systems of different sizes (ragged), each a full
 * hpdmk_tree, advanced together: one Metropolis move of every replica per
 * step, with one launch of each kernel over all replicas per step, and runs
 * of steps captured into a HIP graph and replayed.  The chains are exactly
 * the plain chains (hpdmk_mc_run mode 0) of the replicas. */
typedef void* hpdmk_replicas;
/* params[r], n_src[r]: replica r; r_src and charge concatenate the replicas
 * (3 n_src[r] and n_src[r] doubles each).  One precision for the set. */
int hpdmk_replicas_create(int n_rep, const HPDMKParams* params, int precision, const int* n_src,
                          const double* r_src, const double* charge, hpdmk_replicas* set);
int hpdmk_replicas_destroy(hpdmk_replicas set);
int hpdmk_replicas_count(hpdmk_replicas set, int* n_rep);
/* Replica r as an hpdmk_tree (energies, positions, single moves,
 * hpdmk_mc_set_model); it belongs to the set and is destroyed with it. */
int hpdmk_replicas_tree(hpdmk_replicas set, int r, hpdmk_tree* tree);
/* outgoing and incoming expansions of every replica */
int hpdmk_replicas_form_pw(hpdmk_replicas set);
/* n_moves per replica, replica-major: idx[r * n_moves + k] (replica-local
 * index), d[3 (r * n_moves + k) + a], u[r * n_moves + k]. */
int hpdmk_replicas_mc_load(hpdmk_replicas set, int n_moves, const long* idx, const double* d, const double* u);
/* Runs every replica's chain; graph_steps > 1 captures that many steps into
 * one HIP graph and replays it; graph_steps >= n_moves captures the whole
 * run (every step of every replica and the periodic slot compactions) into
 * one graph that runs as a single launch (kept instantiated for later runs
 * of as many moves). */
int hpdmk_replicas_mc_run(hpdmk_replicas set, double beta, int graph_steps, double* gpu_ms);
/* dE and decisions, replica-major like the move list */
int hpdmk_replicas_mc_results(hpdmk_replicas set, double* dE, int* accept);

/* ---- FP64 Ewald reporter -------------------------------------------------
 * The reference's Ewald oracle (include/ewald.hpp:14-57, src/ewald.cpp:17-308)
 * on the GPU, all-FP64: the ground truth for DMK energies at sizes where the
 * CPU sum is too slow.  r_c = s/alpha must be below L/2 (minimum image) and
 * 2 s alpha L / (2 pi) at most 127. */
typedef void* hpdmk_ewald;
/* Ewald(L, s, alpha, eps, q, r, n) and its structure factor */
int hpdmk_ewald_create(double L, double s, double alpha, double eps, int n, const double* q, const double* r,
                       hpdmk_ewald* ewald);
int hpdmk_ewald_destroy(hpdmk_ewald ewald);
/* compute_energy (src/ewald.cpp:178-191); parts (optional, 3): short, long, self */
int hpdmk_ewald_compute_energy(hpdmk_ewald ewald, double* energy, double* parts);
/* compute_potential (src/ewald.cpp:281-308) at n points (3 doubles each) */
int hpdmk_ewald_compute_potential(hpdmk_ewald ewald, int n, const double* points, double* phi);
/* Energy change of each move of a chain (particle idx[k] displaced by
 * d[3k..3k+2], wrapped); moves with apply[k] != 0 are applied before the
 * next move (none when apply is NULL). */
int hpdmk_ewald_shift_chain(hpdmk_ewald ewald, int n_moves, const long* idx, const double* d, const int* apply,
                            double* dE);

/* HBM streaming-copy probe (no counterpart in the reference; a measurement
 * aid): copies `bytes` between two fresh device buffers `reps` times and
 * returns read+write GB/s.  Each launch reads and writes `bytes` once, which
 * calibrates hardware byte counters (FETCH_SIZE, WRITE_SIZE). */
int hpdmk_copy_probe(long bytes, int reps, double* gbs);

/* Clears a sticky device error (HPDMK_ECAPACITY) after the caller rebuilt or
 * gave up on the chain. */
int hpdmk_clear_error(hpdmk_tree tree);

#ifdef __cplusplus
}
#endif

#endif /* HPDMK_H */

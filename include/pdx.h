/* pdx - MI355X Projective-Dynamics substep engine, C ABI.
 *
 * Drop-in for the reference's substep and global solve (12and94/pd_ipc):
 *   int pd::stepOnce(SimContext&)                       core/sim/Integrator.h:14
 *   IGlobalSolver::analyze/factorize/solve/solveComponent/stats
 *                                                       core/solver/IGlobalSolver.h:45-127
 *   EigenDirectSolver (AMD + LDL^T, compact shared factor)
 *                                                       core/solver/EigenDirectSolver.h:33-54
 *
 * Conventions: every entry point returns an int status (PDX_OK or a negative
 * code) where the reference aborts; host buffers are caller-owned and copied;
 * device memory is owned by the context; there is no global mutable state, so
 * distinct contexts may be driven from distinct host threads.  Calls on one
 * context must not overlap.
 *
 * A context holds one or more cloth instances that step together in one
 * persistent launch (one workgroup per instance).  A context with one instance
 * in PDX_PREC_FP64 is the plain drop-in for one SimContext.
 */
#ifndef PDX_H
#define PDX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PDX_OK = 0,
  PDX_EINVAL = -1,      /* bad argument (sizes, indices, NULL) */
  PDX_ENOTREADY = -2,   /* solve before factor, or no instance */
  PDX_ENOTSPD = -3,     /* LDL^T hit a zero pivot (EigenDirectSolver.cpp:284-291) */
  PDX_EHIP = -4,        /* HIP runtime error */
  PDX_EUNSUPPORTED = -5,/* component factors differ / size beyond the build's limits */
  PDX_ENOMEM = -6
};

enum {
  PDX_PREC_FP64 = 0,    /* reference arithmetic in FP64 (the drop-in) */
  PDX_PREC_MIXED32 = 1  /* FP32 factor/solve on the update, FP64 state, gather, residual */
};

enum {
  PDX_SOLVE_EXACT = 0,  /* level-scheduled gather sweeps in the reference's sum order */
  PDX_SOLVE_FAST = 1,   /* chain-blocked sweeps (forward order kept, backward reordered) */
  PDX_SOLVE_DENSE = 2   /* one instance: the inverse of the component matrix is formed once
                           from the factor (host, m x m, resident in HBM) and each PD
                           iteration's solve is one dense product on the whole GPU
                           (cooperative launch); m <= 32768.  Same arithmetic as the fast
                           mode otherwise (FP64 state and residual; mixed32 solves the
                           update in FP32).  No differentiable rollout, no Chebyshev blend. */
};

enum {
  PDX_ORDER_AMD = 0, /* the reference's fill-reducing order (Eigen AMD on the 3n pattern,
                        EigenDirectSolver.cpp:230-262, restated); the default */
  PDX_ORDER_ND = 1   /* nested dissection by coordinate bisection (fast solve mode only):
                        ~25 % more factor entries, less than half the dependent sweep
                        steps on cloth grids; results differ from AMD's by rounding */
};

enum {
  PDX_LAUNCH_BATCH = 0, /* one workgroup per instance, persistent over the substeps */
  PDX_LAUNCH_GRID = 1   /* the grid's workgroups on one instance: cooperative launch; one sweep
                           step per level of whole elimination-tree chains, paced by per-step
                           completion counters, grid barriers only around the vertex phases
                           (fast solve mode, exactly one instance, fewer than 2^24 vertices) */
};

typedef struct pdx_ctx pdx_ctx;

/* Mirrors SceneConfig (core/sim/Scene.h:14-137): the fields the substep reads. */
typedef struct {
  double dt;               /* substep h */
  double gravity[3];
  double damping;          /* velocityDamping k_d */
  int max_iterations;      /* PD iterations per substep */
  double rel_tol;          /* relTolerance (<=0 disables) */
  double abs_tol;          /* absTolerance (<=0 disables) */
  double residual_tol;     /* residualTolerance (<=0 disables; then the displacement test decides) */
  /* Chebyshev blend (Integrator.cpp:474-617; the reference reads PD_CHEB, PD_CHEB_RHO,
   * PD_CHEB_START from the environment): chebyshev_rho > 0 turns it on; from
   * iteration chebyshev_start+1 the solve's output is blended with the previous
   * iterate, x = prev + (x_sol - prev) w, w = 2/(2 - rho^2) then 4/(4 - rho^2 w).
   * The residual stays that of the unblended solve, as in the reference. */
  double chebyshev_rho;
  int chebyshev_start;
} pdx_config;

/* One cloth: mirrors Mesh (core/mesh/Mesh.h:96-217).  velocities/pin_positions may
 * be NULL (zero / current positions); bends may be absent (n_bends = 0). */
typedef struct {
  int n_vertices;
  const double* positions;      /* 3n, xyz interleaved (dof 3v+d) */
  const double* velocities;     /* 3n or NULL */
  const double* masses;         /* n */
  const uint8_t* pinned;        /* n, 1 = pinned */
  const double* pin_positions;  /* 3n or NULL */
  int n_edges;                  /* distance + shear constraints (Edge, Mesh.h:21-26) */
  const int32_t* edge_a;
  const int32_t* edge_b;
  const double* edge_rest;
  const double* edge_kappa;
  int n_bends;                  /* linear midpoint bend stencils (BendStencil, Mesh.h:53-58) */
  const int32_t* bend_a;
  const int32_t* bend_b;
  const int32_t* bend_c;
  const double* bend_k;
} pdx_scene;

/* Per-instance statistics (Scene.h:161-169 + IGlobalSolver::Stats). */
typedef struct {
  int iterations_used;     /* last substep */
  int converged;           /* last substep reached the criterion */
  int early_exit_count;    /* substeps that converged before max_iterations */
  int factorize_count;     /* numeric factorisations so far */
  int analyze_count;       /* symbolic analyses so far */
  int color_count;         /* edge colours */
  long long total_iterations;
  long long substeps;
  double last_residual;    /* m/s^2; negative when the residual test is off */
  long long factor_nnz;    /* strictly-lower nnz of one component (reference factorNnz/3) */
  int fwd_levels;          /* level count of the forward sweep */
  int bwd_levels;
  int fwd_steps;           /* barrier-separated steps actually launched per sweep */
  int bwd_steps;
  double analyze_seconds;
  double factorize_seconds;
  /* device time per phase, summed over substeps: gather (rhs + residual), forward
   * sweep, backward sweep, predict + velocity, criterion reductions (always kept,
   * the reference's stage buckets).  Batch form: the instance's workgroup; grid
   * form and dense mode: workgroup 0 between the grid's barriers, with the whole
   * solve in the forward bucket (its sweeps overlap across workgroups);
   * then, only in a library built with PDX_PHASE_TIMERS=1 (zero otherwise), fast
   * mode, both sweeps: run streaming of thread 0, its wait at the step barrier,
   * finalisation of wavefront 0, its wait at the second barrier */
  double phase_seconds[9];
  /* PDX_PHASE_TIMERS=1 builds: start and end of the instance's workgroup in the
   * last batch launch, and of its partner when split over two (device clock,
   * seconds); zero otherwise */
  double workgroup_seconds[4];
} pdx_stats;

/* --- context ---------------------------------------------------------------- */
int pdx_create(pdx_ctx** out, int device, int precision, int solve_mode);
int pdx_destroy(pdx_ctx* ctx);
const char* pdx_last_error(const pdx_ctx* ctx);

/* Adds an instance (topology upload: Scene.cpp makeScene's products).  Returns
 * its index through *instance.  An instance added after the context has stepped
 * is analysed, factored and uploaded with the next step (the other instances keep
 * their state, factors and counters). */
int pdx_add_instance(pdx_ctx* ctx, const pdx_scene* scene, const pdx_config* cfg, int* instance);
int pdx_instance_count(const pdx_ctx* ctx);

/* Analyse + factor every instance whose stamp changed (host FP64 AMD + LDL^T,
 * Assembler.cpp:29-92 stamps), build the sweep schedules, upload.  stepping
 * calls this implicitly, as stepOnce does on its first call. */
int pdx_factor(pdx_ctx* ctx);

/* Runs nsub substeps on every instance (Integrator.cpp:312-797 per substep) in one
 * persistent launch.  Blocks until done. */
int pdx_step(pdx_ctx* ctx, int nsub);
/* Same, asynchronous on the context's stream (no host synchronisation). */
int pdx_step_async(pdx_ctx* ctx, int nsub);
int pdx_sync(pdx_ctx* ctx);
/* Captures `nsub` substeps as a HIP graph and replays it `replays` times. */
int pdx_step_graph(pdx_ctx* ctx, int nsub, int replays);

/* --- state ------------------------------------------------------------------ */
int pdx_get_state(pdx_ctx* ctx, int instance, double* positions, double* velocities);
int pdx_set_state(pdx_ctx* ctx, int instance, const double* positions, const double* velocities);
/* Moves pins without refactoring (pin positions are not in the stamp). */
int pdx_set_pin_positions(pdx_ctx* ctx, int instance, const double* pin_positions);
/* Changes edge stiffness; the next step refactors once (stamp change). */
int pdx_set_edge_kappa(pdx_ctx* ctx, int instance, const double* kappa);
/* Changes bend stencil stiffness (BendStencil::stiffness); refactors once likewise. */
int pdx_set_bend_k(pdx_ctx* ctx, int instance, const double* k);
/* Changes the lumped masses (Mesh::masses, n values >= 0).  Masses enter
 * L = M/h^2 + ..., so the next step factors once more, without a new analysis
 * (the numeric stamp changes, the structure stamp does not: Assembler.cpp:29-92). */
int pdx_set_masses(pdx_ctx* ctx, int instance, const double* masses);
/* Changes the pin mask (Mesh::pinned, n bytes, nonzero = pinned).  The mask
 * decides which rows of L are overwritten and which coupling blocks exist, so
 * both stamps change and the next step analyses and factors again
 * (Assembler.cpp:29-92; tests/chain/spring_vertical.cpp:366-371, :528-563).
 * A newly pinned vertex is held at its pin position (pdx_set_pin_positions to
 * move it); a pin trajectory set for the instance is cleared when the pin count
 * changes.  An unchanged mask is a no-op. */
int pdx_set_pinned(pdx_ctx* ctx, int instance, const uint8_t* pinned);
/* Changes the substep configuration.  Only a dt change refactors (stamp); an
 * unchanged configuration is a no-op. */
int pdx_set_config(pdx_ctx* ctx, int instance, const pdx_config* cfg);
int pdx_get_stats(pdx_ctx* ctx, int instance, pdx_stats* out);

/* --- global solve (EigenDirectSolver::solve, EigenDirectSolver.cpp:299-357) ---- */
/* x = L^{-1} b for one instance, all three components, on the GPU. */
int pdx_solve3(pdx_ctx* ctx, int instance, const double* b, double* x);
/* pdx_solve3 without allocating: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  b and x are staged through a static
 * device array of the library (393,216 + 393,216 FP64 values, so at most 131,072
 * vertices; PDX_EINVAL beyond that), held on every device the library runs on
 * until the process ends, instead of pdx_solve3's per-context buffers.  The
 * context must already be factored and uploaded (pdx_factor or a step, and no
 * change since): PDX_ENOTREADY otherwise, as an upload would allocate.  Calls
 * that overlap on one device share the array and race; same results as
 * pdx_solve3 bit for bit. */
int pdx_solve3_noalloc(pdx_ctx* ctx, int instance, const double* b, double* x);
/* Host copy of the factor (rank space): ptr[m+1], rows[nnz], vals[nnz], dinv[m],
 * rank_vertex[m], perm3[3m]; any pointer may be NULL.  *m / *nnz receive sizes. */
int pdx_get_factor(pdx_ctx* ctx, int instance, int* m, long long* nnz, int32_t* ptr,
                   int32_t* rows, double* vals, double* dinv, int32_t* rank_vertex,
                   int32_t* perm3);

/* Topology of one instance (ConstraintColoring.h:34-81, Mesh::constraintColoring and
 * Mesh::bendAdjacency): colour of each edge color_of[E], edges by colour order[E],
 * colour ranges start[C+1], vertex -> edge incidence in edge order vertex_start[n+1] /
 * vertex_edges[2E], vertex -> stencil incidence bend_vertex_start[n+1] /
 * bend_stencils[3B].  Any pointer may be NULL. */
int pdx_get_topology(pdx_ctx* ctx, int instance, int* color_count, uint32_t* color_of,
                     uint32_t* order, uint32_t* start, uint32_t* vertex_start,
                     uint32_t* vertex_edges, uint32_t* bend_vertex_start,
                     uint32_t* bend_stencils);

/* Launch shape: threads per workgroup (64..512, multiple of 64) and the longest
 * elimination-tree chain one wavefront takes as a block (1..32, fast mode).
 * Changing the chain cap re-packs the schedules (no new factorisation is counted). */
int pdx_set_schedule(pdx_ctx* ctx, int block_threads, int chain_cap);

/* Fill-reducing order of every instance's factor (PDX_ORDER_*).  No counterpart in
 * the reference, whose order is always AMD (EigenDirectSolver.cpp:240-241).
 * PDX_ORDER_ND needs the fast solve mode.  A change re-analyses every instance at
 * the next pdx_factor / pdx_step (counted in analyze_count and factorize_count). */
int pdx_set_ordering(pdx_ctx* ctx, int ordering);

/* Splits every instance whose factor has at least min_factor_entries strictly-lower
 * entries (pdx_stats.factor_nnz) over two workgroups (0, the default, keeps one
 * workgroup per instance; 1 splits every instance that can be).  No counterpart in the
 * reference, whose solve is one CPU call per cloth (EigenDirectSolver.cpp:411-447).
 * Applies to the batch launch of PDX_PREC_MIXED32 / PDX_SOLVE_FAST contexts in
 * PDX_ORDER_ND whose factor has a top-level split and W in LDS: the two
 * workgroups sweep one half of the split each and meet at its separator, and
 * each gathers the right-hand sides and residuals of its own vertices, so the
 * instance's critical path is about halved.  Both workgroups run in one launch;
 * they wait for each other with a bound, and a launch where a wait gave up
 * returns PDX_EHIP.  Results are bit for bit those of one workgroup (the
 * schedules cut the same blocks either way).  The differentiable rollout runs
 * every instance on one workgroup.  Re-packs the schedules (no new factorisation). */
int pdx_set_split(pdx_ctx* ctx, long long min_factor_entries);
/* Number of instances currently split over two workgroups (after pdx_factor or a step). */
int pdx_split_count(const pdx_ctx* ctx);
/* The current split threshold (factor entries; 0 = no split; -1 for a NULL context). */
long long pdx_get_split(const pdx_ctx* ctx);
/* The context's launch form (PDX_LAUNCH_*), grid workgroups (0 = all that fit)
 * and fill order (PDX_ORDER_*), as set or as the autotuner chose them.  Any
 * pointer may be NULL.  No counterpart in the reference. */
int pdx_get_launch(const pdx_ctx* ctx, int* mode, int* grid_blocks, int* ordering);
/* Host check of an instance's fast-mode sweep schedules (no device use): replays
 * them in FP64 in the kernel's phase order - for a split instance, each half seeing
 * only what its workgroup holds - and fails with PDX_EUNSUPPORTED (message in
 * pdx_last_error) when a step reads a value before it is final, reads a rank its
 * workgroup does not hold, leaves a rank unsolved or indexes outside an array.
 * *rel_err receives max |x - x_direct| / max |x_direct| against a plain LDL^T solve
 * of the same factor.  No counterpart in the reference. */
int pdx_verify_schedule(pdx_ctx* ctx, int instance, double* rel_err);

/* --- differentiable rollout and adjoint ---------------------------------------
 * No counterpart in the reference (it has no adjoint); these entry points are the
 * "seed dLoss/dx_T, return dLoss/dv_0, dLoss/dkappa, dLoss/dq(t)" boundary of the
 * differentiable batched form.  The rollout runs every instance for a fixed
 * iteration count (max_iterations every substep: no stop test), so the map it
 * differentiates is the unrolled substep. */
/* Number of pinned vertices of an instance (the pin index order is ascending vertex). */
int pdx_pin_count(const pdx_ctx* ctx, int instance);
/* Pin targets for each substep of the next rollout: traj[nsub][pins][3].  NULL
 * clears it (the pins then stay at their current targets). */
int pdx_set_pin_trajectory(pdx_ctx* ctx, int instance, int nsub, const double* traj);
/* Runs nsub substeps from the current state on every instance, checkpointing the
 * state every `checkpoint_every` substeps; the state advances as pdx_step would. */
int pdx_rollout(pdx_ctx* ctx, int nsub, int checkpoint_every);
/* dLoss/dx and dLoss/dv of the state after the rollout (3n each, NULL = zero). */
int pdx_set_adjoint_seed(pdx_ctx* ctx, int instance, const double* dx, const double* dv);
/* Reverse pass of the last rollout: per checkpoint segment, recompute the forward
 * recording every iterate, then run the adjoint kernel (transposed solves on the
 * same factor + projection Jacobians).  Leaves the state as after the rollout. */
int pdx_adjoint(pdx_ctx* ctx);
/* Gradients of the last pdx_adjoint: dLoss/dx_0 and dLoss/dv_0 (3n), dLoss/dkappa
 * per edge (E), dLoss/dk per bend stencil (B), dLoss/dq [nsub][pins][3].  Any
 * pointer may be NULL. */
int pdx_get_gradients(pdx_ctx* ctx, int instance, double* g_x0, double* g_v0, double* g_kappa,
                      double* g_bend, double* g_pin);
/* Device time (ms) of the last pdx_rollout and of the last pdx_adjoint. */
int pdx_rollout_timing(const pdx_ctx* ctx, double* forward_ms, double* adjoint_ms);

/* Fast-mode schedule of one instance, 8 values per sweep (forward, then backward):
 * steps, tasks (in the grid form's level sweeps: with the dense external tasks),
 * chain tasks, rows in chains, runs, dense values (chain inverses, and in the
 * level sweeps also inverse and external panels), most tasks in one step, most
 * runs in one step. */
int pdx_get_schedule_stats(pdx_ctx* ctx, int instance, int64_t* out16);

/* Launch form (PDX_LAUNCH_*); grid_blocks = workgroups of the grid form, 0 = all
 * that can be co-resident.  Re-packs the schedules for the form.  In the dense
 * solve mode (always on the grid) only grid_blocks applies. */
int pdx_set_launch(pdx_ctx* ctx, int mode, int grid_blocks);

/* Autotuner: times `probe_substeps` substeps of the whole context under each
 * choice and keeps the fastest.  A context of one cloth in the fast mode first
 * probes the fill order (PDX_ORDER_AMD / PDX_ORDER_ND; a change re-analyses and
 * refactors, which factorize_count and analyze_count count) and the launch form
 * (PDX_LAUNCH_BATCH at the current shape, or PDX_LAUNCH_GRID with 16, 32, ...,
 * 256 workgroups of 256 or 512 threads; a probe within 1.5 x of the best so far
 * runs twice and keeps the faster); pdx_get_launch reads the choice.  Then every context probes the launch shape
 * (threads per workgroup 256 / 512 x chain cap 8 / 16 / 24 / 32 in fast mode; in
 * the dense solve mode, 16, 32, 64, ... workgroups, reported in best_block with
 * best_cap 0); then, for a FP32-solve fast batch, the split threshold
 * (pdx_set_split: none, or the factor size of the 1st, 2nd, 4th, 8th, ..., 64th
 * largest instance that can split; pdx_get_split reads the choice).  State and
 * the step statistics are restored, so the probes leave no trace in the
 * simulation.  No counterpart in the reference (its CPU path has no launch
 * schedule). */
int pdx_autotune(pdx_ctx* ctx, int probe_substeps, int* best_block, int* best_cap, double* best_ms);

/* --- timing ---------------------------------------------------------------- */
/* Average duration (ms) of the substep kernel over the last `pdx_step*` call,
 * measured with HIP events on the stream the kernel runs on. */
double pdx_last_kernel_ms(const pdx_ctx* ctx);
/* The device's own figures (no counterpart in the reference; for reports):
 * compute units, peak shader clock (kHz), peak memory clock (kHz) and memory bus
 * width (bits), as the HIP runtime states them.  Any pointer may be NULL. */
int pdx_device_info(int device, int* compute_units, int* clock_khz, int* mem_clock_khz, int* bus_width_bits);
/* Algorithmic bytes of one PD iteration (factor both sweeps, maps, gather of
 * edges/incidence for the right-hand side and the residual, vector passes),
 * summed over instances; each array counted once per phase at storage width. */
double pdx_bytes_per_iteration(const pdx_ctx* ctx);
/* Same for one substep: iterations x the above + predict/velocity passes. */
double pdx_bytes_per_substep(const pdx_ctx* ctx);

/* --- scene builders (grid cloth, Mesh.cpp:12-54, :183-207, :219-285) ---------- */
/* Fills caller buffers for an nx*ny grid in the xz-plane, spacing sx, sz.
 * Edges: row/column neighbours in the reference order, then (shear>0) one
 * diagonal per quad; bends (bend>0): rows then columns.  Sizes:
 *   n = nx*ny, E = 2nxny-nx-ny (+(nx-1)(ny-1)), B = (nx-2)ny+(ny-2)nx.
 * Pass NULL arrays to query sizes only. */
int pdx_grid_counts(int nx, int ny, int shear, int bend, int* n, int* E, int* B);
int pdx_grid_fill(int nx, int ny, double sx, double sz, double kappa, double shear_kappa,
                  double bend_k, double* positions, int32_t* ea, int32_t* eb, double* rest,
                  double* kap, int32_t* sa, int32_t* sb, int32_t* sc, double* bk);

/* --- unstructured input (Mesh::loadObj, Mesh.cpp:56-125) -------------------- */
/* Reads a Wavefront OBJ triangle mesh the way Mesh::loadObj does: "v" lines are
 * vertices, "f" lines give 1-based indices (text before the first '/'; the first
 * three corners of a polygon), unique undirected edges in face order with rest
 * length |x_b - x_a|.  PDX_EINVAL when the file cannot be opened, a face index
 * does not parse, there are no vertices or no triangles, or an index is outside
 * [1, n] (the reference reads such an index out of bounds).
 * pdx_obj_counts returns the sizes; pdx_obj_fill fills 3n positions, 3T
 * triangle corners (0-based), E edges with kappa on every edge. */
int pdx_obj_counts(const char* path, int* n, int* n_edges, int* n_tris);
int pdx_obj_fill(const char* path, double kappa, double* positions, int32_t* tris,
                 int32_t* ea, int32_t* eb, double* rest, double* kap);

/* Lumped masses.  Surfaces (Mesh::computeMassesFromSurfaces, Mesh.cpp:127-153):
 * density * triangle area / 3 to each corner.  Edges (computeMassesFromEdges,
 * :155-181): density * rest^2 / 2 to each end.  Both then give every mass not
 * above 1e-12 the smallest positive mass (1 if there is none). */
int pdx_masses_from_surfaces(int n, const double* rest_positions, int n_tris,
                             const int32_t* tris, double density, double* masses);
int pdx_masses_from_edges(int n, int n_edges, const int32_t* ea, const int32_t* eb,
                          const double* rest, double density, double* masses);

/* --- physics diagnostics ----------------------------------------------------- */
/* Computed on the host in FP64 from the instance's current state, in the
 * reference's summation order (edge order, then stencil order, then vertex order). */
typedef struct {
  double max_strain;       /* Mesh::maxRelativeStrain (Mesh.cpp:351-360) */
  double mean_strain;      /* Mesh::meanRelativeStrain (Mesh.cpp:362-372) */
  double elastic_energy;   /* surrogateEnergy (Integrator.cpp:804-826): edges + bend stencils */
  double total_energy;     /* totalEnergy (Integrator.cpp:828-840): kinetic + elastic + gravity */
  double max_speed;        /* maxSpeed (Integrator.cpp:842-846) */
  double total_mass;       /* Mesh::totalMass (Mesh.cpp:345-349) */
} pdx_diag;
int pdx_diagnostics(pdx_ctx* ctx, int instance, pdx_diag* out);
/* Structural nonzeros of the 3n x 3n matrix L (IGlobalSolver::Stats::nnz); needs
 * pdx_factor (or a step) first. */
int pdx_matrix_nnz(const pdx_ctx* ctx, int instance, long long* nnz);

/* --- headless report (pd_bench, app/bench/main.cpp:53-264) ------------------ */
/* Runs one cloth the way pd_bench does and writes its report.  Options as pd_bench:
 *   --grid NX NY, --mesh PATH (OBJ; falls back to the grid when it does not load),
 *   --spacing S, --steps N, --substeps N, --dt H, --stiffness K (default 1e4),
 *   --density RHO, --damping KD, --iters N, --preset preview|realtime|accurate,
 *   --residual-tol T, --tol T, --threads N (accepted, no effect), --pin-top,
 *   --pin-all, --report N, --help
 * plus --shear K and --bend K (the reference reads PD_SHEAR / PD_BEND from the
 * environment; both are honoured too), --cheb-rho R and --cheb-start S (the
 * Chebyshev blend; PD_CHEB=1, PD_CHEB_RHO, PD_CHEB_START as the reference's
 * substep reads them are honoured too), --precision fp64|mixed32,
 * --solve exact|fast|dense and --device D.  A preset sets the iteration cap and the
 * residual tolerance; explicit --iters / --residual-tol override it whatever the
 * order.  Unknown options are ignored, as in pd_bench.
 * The report text (NUL-terminated, truncated to text_cap - 1 bytes) goes to
 * `text` when it is not NULL; the figures go to `summary` when it is not NULL.
 * Returns PDX_OK, PDX_EINVAL for an unknown preset, or the failing call's status. */
typedef struct {
  int n_vertices, n_edges, n_bends, n_pinned;
  int substeps, converged_substeps, max_iterations_used;
  long long total_iterations;
  double last_residual;
  double max_strain, mean_strain, run_max_strain;
  double elastic_energy, total_energy, max_speed, total_mass;
  double wall_seconds;             /* whole stepping loop, diagnostics included */
  double device_ms_per_substep;    /* substep kernel, HIP events */
  int analyze_count, factorize_count, color_count;
  long long matrix_nnz, factor_nnz;   /* 3n x 3n structure; 3 x one component's strictly-lower factor */
  double analyze_seconds, factorize_seconds;
  int loaded_mesh;                 /* 1 when --mesh loaded, 0 for the grid */
} pdx_bench_summary;
int pdx_bench_run(int argc, const char* const* argv, char* text, long long text_cap,
                  pdx_bench_summary* summary);

#ifdef __cplusplus
}
#endif
#endif

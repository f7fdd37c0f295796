/* FP64 CPU oracle of the reference Projective-Dynamics substep path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference algorithm
 * (12and94/pd_ipc: core/sim/Integrator.cpp stepOnce, core/energy/DistanceTerm.cpp,
 * core/assemble/Assembler.cpp, core/solver/EigenDirectSolver.cpp,
 * core/mesh/ConstraintColoring.cpp) in plain serial/OpenMP C++ with the reference's
 * arithmetic order, so that the GPU library can be checked against it.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load it.
 * The product library never links or calls it.
 *
 * The fill-reducing ordering (AMD on the 3n pattern) is an input: the caller hands
 * in the pivot order (origOfPerm, length 3n).  Everything else - colouring,
 * incidence CSRs, assembly, the up-looking LDL^T of the permuted 3n matrix, the
 * compact shared factor, the substep - is computed here.
 */
#ifndef PD_ORACLE_H
#define PD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

typedef struct {
  double dt;
  double gravity[3];
  double damping;
  int max_iterations;
  double rel_tol;
  double abs_tol;
  double residual_tol;
  /* Chebyshev blend (Integrator.cpp:474-617, the reference's PD_CHEB / PD_CHEB_RHO /
   * PD_CHEB_START): 0 = off, else from iteration cheb_start+1 on the solve's
   * output is blended with the previous iterate, x = prev + (x_sol - prev) w. */
  int cheb;
  double cheb_rho;
  int cheb_start;
} orc_config;

/* Scene arrays are copied.  pin_positions may be NULL (then = positions). */
orc_ctx* orc_create(int n, const double* positions, const double* velocities,
                    const double* masses, const uint8_t* pinned, const double* pin_positions,
                    int n_edges, const int32_t* ea, const int32_t* eb, const double* rest,
                    const double* kappa, int n_bends, const int32_t* sa, const int32_t* sb,
                    const int32_t* sc, const double* bk, const orc_config* cfg);
void orc_destroy(orc_ctx* c);

/* Pivot order of the 3n system (pivot k eliminates dof perm3[k]).  Must be set
 * before the first step.  Returns 0, or -1 if it is not a permutation. */
int orc_set_ordering(orc_ctx* c, const int32_t* perm3);
void orc_set_threads(orc_ctx* c, int threads);

/* Runs nsub substeps (reference stepOnce each).  Returns the iterations of the
 * last substep, or a negative error code. */
int orc_step(orc_ctx* c, int nsub);

void orc_get_positions(const orc_ctx* c, double* out3n);
void orc_get_velocities(const orc_ctx* c, double* out3n);
void orc_set_positions(orc_ctx* c, const double* x3n);
void orc_set_velocities(orc_ctx* c, const double* v3n);
void orc_set_pin_positions(orc_ctx* c, const double* q3n);
void orc_set_edge_kappa(orc_ctx* c, const double* kappa);
void orc_set_config(orc_ctx* c, const orc_config* cfg);
void orc_set_bend_k(orc_ctx* c, const double* k);
void orc_set_masses(orc_ctx* c, const double* masses);
void orc_set_pinned(orc_ctx* c, const uint8_t* pinned);

/* FP64 reverse-mode derivative of nsub substeps run from the current state with
 * a fixed iteration count (max_iterations every substep; returns -5 if the stop
 * test fired).  pin_traj (nsub x pins x 3, pins in ascending vertex order, or
 * NULL = current pin positions) sets the pin targets of each substep.  Seeds:
 * dLoss/dx_T (3n) and dLoss/dv_T (3n or NULL).  Outputs (any may be NULL):
 * dLoss/dx_0 (3n), dLoss/dv_0 (3n), dLoss/dkappa_e (E), dLoss/dk_s (B),
 * dLoss/dq (nsub x pins x 3).  The state is advanced by the nsub substeps. */
int orc_rollout_adjoint(orc_ctx* c, int nsub, const double* pin_traj, const double* seed_x,
                        const double* seed_v, double* g_x0, double* g_v0, double* g_kappa,
                        double* g_bend, double* g_pin);

/* stats[0]=iterationsUsed, [1]=converged, [2]=earlyExitCount, [3]=factorizeCount,
 * [4]=analyzeCount, [5]=colorCount, [6]=compact factor nnz (one component),
 * [7]=components (3 if split), [8]=compact ready */
void orc_stats(const orc_ctx* c, int64_t* stats);
double orc_last_residual(const orc_ctx* c);

/* Reference whole-run diagnostics (Mesh.cpp maxRelativeStrain/meanRelativeStrain,
 * Integrator.cpp surrogateEnergy/totalEnergy). */
double orc_max_strain(const orc_ctx* c);
double orc_mean_strain(const orc_ctx* c);
double orc_surrogate_energy(const orc_ctx* c);
double orc_total_energy(const orc_ctx* c);

/* Colouring arrays (ConstraintColoring.cpp:121-259): colorOf[E], order[E],
 * start[C+1].  Returns the colour count. */
int orc_coloring(const orc_ctx* c, uint32_t* color_of, uint32_t* order, uint32_t* start);

/* Assembles and factors if needed (stamp logic) without stepping. */
int orc_factor(orc_ctx* c);
/* Compact shared factor of component 0 (EigenDirectSolver.cpp:157-225).
 * Any pointer may be NULL.  Returns m (columns), writes nnz into *nnz. */
int orc_compact_factor(const orc_ctx* c, int32_t* ptr, int32_t* row_rel, double* val,
                       double* dinv, int32_t* comp_dof /* 3*m: dof of rank k in comp c */,
                       int64_t* nnz);
/* Scalar assembled matrix values (one component) as CSC over vertices:
 * col ptr n+1, rows, values.  Returns nnz. */
int64_t orc_matrix(const orc_ctx* c, int32_t* colptr, int32_t* rows, double* vals);
/* Component solve of L x = b (EigenDirectSolver::solveComponent compact path). */
int orc_solve3(orc_ctx* c, const double* b, double* x);
/* Serial nonlinear residual at the current positions (Integrator.cpp:113-205),
 * using `predicted` from the last substep. */
double orc_nonlinear_residual(orc_ctx* c);
/* Bending constant right-hand side (Assembler.cpp:94-137) for the current pins. */
void orc_bending_rhs(const orc_ctx* c, double* out3n);

#ifdef __cplusplus
}
#endif
#endif

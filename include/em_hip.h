/* C ABI of the MI355X (gfx950) compressible finite-volume solver.
 *
 * Drop-in for the reference's GPU solver object
 *   GpuSolver<f64> (fluid/crates/compressible-cuda/src/lib.rs:155-477),
 * which a host binds for the Euler system the way the reference's CPU
 * `Solver` (fluid/crates/compressible/src/solver.rs:139-795) is driven:
 * plain pointers and sizes, caller-owned host buffers, library-owned device
 * buffers, not thread-safe per handle.  A handle owns its HIP streams, all at
 * default priority: one main stream for stepping (and everything a call
 * without a stream note does), plus a second stream on which a solver's
 * backward sweep recomputes checkpoint segments, and four sweep streams in an
 * ensemble's backward.  Every call returns with its streams' work ordered
 * before the handle's next call.
 *
 * Host state arrays use the reference's CPU layout: ghost-padded AoS float64,
 * cell index (i*m1 + j)*m2 + k with m_d = n_d + 2*ng on active axes (n_d > 1)
 * and 1 otherwise, five conserved values (rho, rho u, rho v, rho w, E) per cell
 * (solver.rs:40-77).
 *
 * Error convention: every call returns 0 on success and a negative code on
 * failure (the reference returns Result<_, Box<dyn Error>>); em_last_error()
 * gives the message of the calling thread's last failure.
 */
#ifndef EM_HIP_H
#define EM_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct em_solver em_solver;

enum {
    EM_OK = 0,
    EM_ERR_ARG = -1,       /* invalid argument or unsupported option */
    EM_ERR_HIP = -2,       /* HIP runtime / launch failure */
    EM_ERR_COLLAPSE = -3,  /* "time step collapsed" (lib.rs:447-449) */
    EM_ERR_STATE = -4      /* call order / state error */
};

/* Boundary kinds (solver.rs:8-20): 0 periodic, 1 outflow, 2 reflect,
 * 3 Dirichlet (fixed conserved state), 4 characteristic (far-field state). */
/* Flux modes (scheme.rs:39-47): 0 Hybrid, 1 Weno (WCNS everywhere),
 * 2 Central (KEEP-PE everywhere). */

/* Thread-local message of the last failure ("" if none). */
const char* em_last_error(void);

/* GpuSolver::from_cpu (lib.rs:186-282): snapshot of mesh, boundary
 * conditions and scheme options plus the conserved state.
 *   n[3], lo[3], hi[3]  mesh (Mesh::new, solver.rs:31-34), ng ghost layers
 *                       (>= 5 for Hybrid/Weno, >= 3 for Central)
 *   eos[4]              gamma, p_inf, b, q (eos.rs:12-24)
 *   bc_kind[6]          [axis][side]
 *   bc_state[30]        [axis][side][5] conserved state for Dirichlet and
 *                       characteristic sides (ignored otherwise)
 *   mode, positivity, cfl, tol, t0   Solver fields (solver.rs:144-152)
 *   u                   ghost-padded AoS conserved state (ghosts ignored) */
int em_solver_create(em_solver** out, const int64_t n[3], const double lo[3], const double hi[3], int32_t ng,
                     const double eos[4], const int32_t bc_kind[6], const double bc_state[30], int32_t mode,
                     int32_t positivity, double cfl, double tol, double t0, const double* u);

/* GpuSolver::<f32>::from_cpu (lib.rs:26-48 Real for f32, 186-282): the same
 * snapshot with state, residual and arithmetic in FP32 on the device (EOS,
 * mesh spacing, boundary states and tolerance rounded to float, the
 * limiter's lambda formed in double from dt and rounded, as the reference's
 * R::from_f64).  The host side still passes and receives doubles: upload
 * rounds to float, download/rhs widen.  The returned handle takes every
 * em_solver_* call below except the adjoint ones (FP64 only, EM_ERR_ARG).
 * Multiply-adds are contracted by default, as the reference's kernels are
 * compiled (em_solver_set_variant(s, 1) turns contraction off). */
int em_solver_create_f32(em_solver** out, const int64_t n[3], const double lo[3], const double hi[3], int32_t ng,
                         const double eos[4], const int32_t bc_kind[6], const double bc_state[30], int32_t mode,
                         int32_t positivity, double cfl, double tol, double t0, const double* u);

int em_solver_destroy(em_solver* s);

/* Arithmetic variant: precise != 0 evaluates every expression without FMA
 * contraction (reproduces the reference CPU arithmetic; the default); 0 =
 * contracted (fast).  No counterpart in the reference (its GPU path has one
 * build). */
int em_solver_set_variant(em_solver* s, int32_t precise);

/* Tile shape per axis (index into the library's tile table), and the
 * project autotuner that times every shape on the current state and keeps
 * the fastest per axis.  cfg_out may be NULL. */
int em_solver_set_tiles(em_solver* s, const int32_t cfg[3]);
int em_solver_autotune(em_solver* s, int32_t repeats, int32_t cfg_out[3]);

/* Replace the conserved state (ghosts ignored). */
int em_solver_upload(em_solver* s, const double* u);

/* GpuSolver::download (lib.rs:457-461): conserved state in the CPU layout;
 * ghost cells are filled from the current interior with the boundary
 * conditions, axes in order, as Solver::fill_ghosts (solver.rs:259-348). */
int em_solver_download(em_solver* s, double* u);

/* Per-stage residual -div F of the current state (Solver::rhs,
 * solver.rs:351-475) with the positivity limiter's lambda from stage_dt;
 * written in the CPU layout with zero ghosts.  stats (may be NULL) receives
 * [faces, shock-path faces, first-order fallbacks, limited faces]
 * (scheme.rs:101-117) of this evaluation. */
int em_solver_rhs(em_solver* s, double stage_dt, double* rhs, uint64_t stats[4]);

/* GpuSolver::step (lib.rs:401-412): one SSP-RK3 step of size dt. */
int em_solver_step(em_solver* s, double dt);

/* GpuSolver::dt_cfl (lib.rs:415-441): cfl / max over interior cells of
 * sum_d (|u_d| + c) / dx_d. */
int em_solver_dt_cfl(em_solver* s, double* dt);

/* GpuSolver::run_to (lib.rs:443-454): CFL steps until t_end (stop rule of
 * Solver::run_to, solver.rs:758-768).  dt lives on the device; steps run as
 * replayed HIP graphs with one host synchronisation per chunk of steps. */
int em_solver_run_to(em_solver* s, double t_end, int64_t* steps);

/* nsteps CFL-limited steps (dt = cfl / rate of the state at each step),
 * device-resident dt, graph-replayed.  With pass_ms != NULL the steps are
 * launched eagerly with HIP events around every axis pass on the solver's
 * stream and pass_ms[3] receives the mean duration (ms) of one pass along
 * each axis.  Returns EM_ERR_COLLAPSE if dt became non-finite or <= 0. */
int em_solver_run_steps(em_solver* s, int64_t nsteps, double pass_ms[3]);

/* GpuSolver::counters (lib.rs:464-467): [shock-path faces, first-order
 * fallbacks, positivity-limited faces] accumulated since creation. */
int em_solver_counters(em_solver* s, uint64_t out[3]);

/* Solver time t (lib.rs:164). */
int em_solver_time(em_solver* s, double* t);

/* GpuSolver::synchronize (lib.rs:469-472). */
int em_solver_synchronize(em_solver* s);

/* ------------------------------------------------------------------------
 * Discrete adjoint (no counterpart in the reference, which has none: a
 * checkpointed forward/backward pair for sensitivities of the flow to its
 * initial state).  FluxMode::Weno
 * only, periodic / outflow / reflect (n >= 5) / Dirichlet boundaries, axes of
 * any length (pencils are cut into tiles of at most 128 cells with their
 * halo; an axis periodic on one side only keeps whole pencils in LDS, up to
 * 544 cells: longer ones are refused with EM_ERR_ARG).  Conventions: branches
 * follow the forward run, min/max differentiate the returned operand (the
 * first on a tie), d|x|/dx = sign(x) with sign(0) = 0; the positivity
 * limiter's blend weight is differentiated implicitly at the admissibility
 * constraint that bounds it (its bisection resolves it to 2^-30) and its
 * Lax-Friedrichs speed through the max; WENO-Z weights take the linear
 * weights' derivative on stencils flat to rounding and are held on partly
 * flat ones; the dt sequence is frozen (taken from the forward run). */

/* Interior cells per adjoint tile at most (default 128, at least 10): the
 * VJP's work unit along each axis.  Results do not depend on it; it exists
 * for tuning and for the test that shows so.  No counterpart in the
 * reference. */
int em_solver_set_adjoint_tile(em_solver* s, int32_t cells);

/* (dR/du)^T lam of one residual evaluation at the current state, with the
 * limiter's lambda from stage_dt; lam and out in the CPU layout (ghosts
 * ignored / written as zero). */
int em_solver_adjoint_rhs(em_solver* s, double stage_dt, const double* lam, double* out);

/* Forward run of nsteps SSP-RK3 steps from the current state, keeping a
 * checkpoint every checkpoint_every steps.  checkpoint_every = 0 keeps every
 * step's u_n and both RK stage states in device memory instead (3 states per
 * step, 120 bytes per cell and step; EM_ERR_ARG if they do not fit), so that
 * the backward sweep recomputes nothing.  dt per step from dts (frozen
 * sequence) or, with dts == NULL, the CFL dt of each step; dts_out (may be
 * NULL) receives the sequence used. */
int em_solver_forward(em_solver* s, int64_t nsteps, int32_t checkpoint_every, const double* dts, double* dts_out);

/* Backward sweep of the last forward run: dJ/du_0 from dJ/du_N (CPU
 * layout).  Each checkpoint segment is recomputed forward keeping u_n and
 * both RK stage states of every step, then the steps are reversed (after a
 * forward run with checkpoint_every = 0, over the kept states, with the same
 * result bit for bit).  The state and time are left as after the forward
 * run. */
int em_solver_backward(em_solver* s, const double* dj_final, double* dj_initial);
/* em_solver_backward on device arrays (no counterpart in the reference):
 * dj_final and dj_initial hold the interior cells variable-major, value q of
 * cell (i, j, k) at q * ncell + (i * n1 + j) * n2 + k, so nothing is
 * converted on the host.  Returns when dj_initial is written. */
int em_solver_backward_device(em_solver* s, const double* dj_final, double* dj_initial);

/* ------------------------------------------------------------------------
 * Ensemble: independent uniform-grid solvers (own mesh extents, domain,
 * state, dt and CFL rate each; shared EOS, boundary kinds and options)
 * advanced together.  Every axis pass of every instance runs in one
 * persistent launch and a step is replayed as one HIP graph; the adjoint
 * batches its VJPs the same way.  No counterpart in the reference (its
 * GpuSolver holds one mesh).  Host arrays are per instance, in that
 * instance's CPU layout; u / dj arrays are arrays of per-instance pointers. */
typedef struct em_ensemble em_ensemble;

int em_ensemble_create(em_ensemble** out, int32_t ninst, const int64_t* n /* [ninst][3] */,
                       const double* lo /* [ninst][3] */, const double* hi /* [ninst][3] */, int32_t ng,
                       const double eos[4], const int32_t bc_kind[6], const double bc_state[30], int32_t mode,
                       int32_t positivity, double cfl, double tol, const double* const* u);
int em_ensemble_destroy(em_ensemble* h);

/* 1 (default): one persistent launch per pass over all instances; 0: one
 * launch per instance (same results). */
int em_ensemble_set_batched(em_ensemble* h, int32_t batched);

/* Tile shape of every pass (index into the library's tile table). */
int em_ensemble_set_tiles(em_ensemble* h, int32_t cfg);
/* Project autotuner for the ensemble: times every tile shape on the stage
 * passes of all instances and keeps the fastest (states and counters
 * untouched); *cfg_out (may be NULL) receives it.  No counterpart in the
 * reference (one GpuSolver per grid, compressible-cuda/src/lib.rs:155-477). */
int em_ensemble_autotune(em_ensemble* h, int32_t repeats, int32_t* cfg_out);

/* nsteps CFL steps of every instance (each its own dt). */
int em_ensemble_run_steps(em_ensemble* h, int64_t nsteps);

/* Replace / read instance i's conserved state (its CPU layout). */
int em_ensemble_upload(em_ensemble* h, int32_t i, const double* u);
int em_ensemble_download(em_ensemble* h, int32_t i, double* u);
int em_ensemble_time(em_ensemble* h, int32_t i, double* t);

/* Forward/backward pair of the ensemble (FluxMode::Weno, conventions as
 * em_solver_forward/backward): CFL dt per instance recorded and frozen,
 * checkpoints every checkpoint_every steps. */
int em_ensemble_forward(em_ensemble* h, int64_t nsteps, int32_t checkpoint_every);
int em_ensemble_dts(em_ensemble* h, int32_t i, double* dts);
int em_ensemble_backward(em_ensemble* h, const double* const* dj_final, double* const* dj_initial);

/* Workload initialisers (host, ghost-padded AoS output of the given mesh),
 * mirroring the reference's example set-ups:
 *   case 0  2-D four-quadrant Riemann problem, params {x_split, y_split}
 *           (compressible-cuda/examples/gpu_validate.rs:100-116)
 *   case 1  3-D periodic Taylor-Green velocity (U0 = params[0]) with
 *           p = 1/gamma and rho = 1 + 0.1 * sum of 8 random Fourier modes
 *           (|k| <= 4, SplitMix64 seed params[1] as an integer) on [0,2pi]^3
 *   case 2  cpu_validate smooth wave (examples/cpu_validate.rs:30-39)
 * lo/hi give the domain (Mesh::x cell centres). */
int em_case_init(int32_t case_id, const int64_t n[3], const double lo[3], const double hi[3], int32_t ng,
                 const double* params, double* u);

/* Allocation-free per-stage residual.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * residual and stats as em_solver_rhs for the same mesh, boundaries and
 * options (ng >= 5 assumed), but on caller-owned device buffers and without a
 * handle:
 *   u       device, interior-only SoA conserved state: 5 planes of
 *           n0*n1*n2 doubles, cell index (i*n1 + j)*n2 + k (read only)
 *   rhs     device, same layout, receives the residual (must not alias u)
 *   stats   device uint64[4] or NULL: [faces, shock-path faces, first-order
 *           fallbacks, limited faces] of this evaluation
 *   stream  hipStream_t the work is queued on (NULL: the null stream); the
 *           call returns without waiting for it
 *   precise 1: reference arithmetic (no FMA contraction), 0: fast variant
 * It calls no hipMalloc/hipFree.  Its only scratch is a fixed
 * EM_NOALLOC_SCRATCH_BYTES static device array per GPU (the counter slots),
 * independent of the mesh; the counters are cleared by a kernel at the start
 * of every call.  Calls that overlap on one GPU share that array and race.
 * Meshes whose axis-pass launch would exceed 2^31 - 1 workgroups, or with
 * more than 2^40 cells, are rejected with EM_ERR_ARG. */
#define EM_NOALLOC_SCRATCH_BYTES 65536
int em_rhs_noalloc(const int64_t n[3], const double lo[3], const double hi[3], const double eos[4],
                   const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity, double tol,
                   int32_t precise, double stage_dt, const double* u, double* rhs, uint64_t* stats, void* stream);

/* ------------------------------------------------------------------------
 * Block-structured AMR (global SSP-RK3 step), drop-in for
 *   GpuAmr<f64, 5, Euler> (fluid/crates/compressible-cuda/src/amr.rs:95-858)
 * with the arithmetic of the CPU `Amr` (fluid/crates/compressible/src/amr.rs),
 * which is its parity target.  Blocks are b^D cells (b = 8, 16 or 32; 1 on
 * axes with base[d] == 0), ids follow the reference's order (children
 * appended, coarsened children swap-removed).  Host block data is AoS float64
 * per block: index ((id*b0 + i)*b1 + j)*b2 + k, five conserved values.
 * Every block (leaves and ancestors) lives on the device; a regrid moves only
 * the per-leaf details to the host. */
typedef struct em_amr em_amr;

/* Message of the calling thread's last failed em_amr_* call. */
const char* em_amr_last_error(void);

/* GpuAmr::new(Amr, ..) (amr.rs:132-184 of compressible-cuda): a hierarchy
 * handed over block by block, as the reference's Amr holds it (level, block
 * coordinates idx[3*i..], leaf flag, data).  Configuration as AmrConfig
 * (compressible/src/amr.rs:49-63): base blocks per axis (0 = inactive axis),
 * block size b, levels, ghost depth (>= 5 used), refinement threshold eps,
 * regrid interval; domain lo/hi; eos {gamma, p_inf, b, q}; boundaries and
 * flux mode as em_solver_create; positivity limiter; cfl; hybrid tolerance. */
int em_amr_create(em_amr** out, const int64_t base[3], int32_t b, int32_t levels, int32_t ng, double eps,
                  int32_t regrid_every, const double lo[3], const double hi[3], const double eos[4],
                  const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity, double cfl,
                  double tol, int64_t nblocks, const int32_t* level, const int64_t* idx, const uint8_t* leaf,
                  const double* u);

/* Amr::new (compressible/src/amr.rs:289-343) on the device: level-0 blocks
 * from a built-in initial condition, refined where the wavelet details exceed
 * eps (children re-evaluated from the initial condition).  Cases (primitive
 * states, params[0..14]):
 *   0 four-quadrant 2-D Riemann problem split at (params[0], params[1])
 *   1 smooth wave rho = 1 + 0.1 sin(2 pi x) cos(2 pi y), u = (0.3, 0.2, 0.1)
 *   2 shock-bubble: Mach params[0] shock at x = params[1] moving into
 *     (rho 1, p 1), bubble centre params[2..4], radius params[5], density
 *     params[6] (gamma 1.4)
 *   3 isentropic vortex (compressible/examples/amr_validate.rs:44-51)
 *   4 state at rest rho = 1 + sum c_d x_d^5 / 5, p = 1 + sum c_d x_d^4 / 4,
 *     c = params[0..2] */
int em_amr_create_case(em_amr** out, const int64_t base[3], int32_t b, int32_t levels, int32_t ng, double eps,
                       int32_t regrid_every, const double lo[3], const double hi[3], const double eos[4],
                       const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity,
                       double cfl, double tol, int32_t case_id, const double* params);
/* GpuAmr::<f32>::new (compressible-cuda/src/amr.rs:95-184 with R = f32,
 * lib.rs:26-48): the same constructors with the pool (state, residual,
 * ghost slabs, face fluxes) and its arithmetic in FP32 on the device; EOS,
 * boundary states, mesh spacing and the prediction/restriction weights
 * rounded to float, the float WENO-Z epsilon 1e-30 (kernels.cu:263).  Host
 * arrays, times, dt and wavelet details stay double (details are computed
 * in float and widened).  The handle takes every em_amr_* call; an AMR group
 * takes FP64 handles only (EM_ERR_ARG). */
int em_amr_create_f32(em_amr** out, const int64_t base[3], int32_t b, int32_t levels, int32_t ng, double eps,
                      int32_t regrid_every, const double lo[3], const double hi[3], const double eos[4],
                      const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity,
                      double cfl, double tol, int64_t nblocks, const int32_t* level, const int64_t* idx,
                      const uint8_t* leaf, const double* u);
int em_amr_create_case_f32(em_amr** out, const int64_t base[3], int32_t b, int32_t levels, int32_t ng, double eps,
                           int32_t regrid_every, const double lo[3], const double hi[3], const double eos[4],
                           const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity,
                           double cfl, double tol, int32_t case_id, const double* params);

int em_amr_destroy(em_amr* h);

/* 1 (default): no FMA contraction, the CPU Amr's arithmetic bit for bit;
 * 0: contracted flux pass. */
int em_amr_set_variant(em_amr* h, int32_t precise);

/* 1 (default): one persistent launch per stage phase over every leaf of
 * every level; 0: one launch per leaf (the unbatched comparison). */
int em_amr_set_batched(em_amr* h, int32_t batched);

/* 1 (default): batched em_amr_run_steps replays each step as one captured
 * HIP graph (captured again after every hierarchy edit, and the executable
 * graph updated in place when the launches keep their topology); 0: eager
 * launches. */
int em_amr_set_graph(em_amr* h, int32_t on);

/* Pass shape per axis (pencils per work item and threads of the batched
 * flux pass; 4 shapes for 8^3 blocks, 2 for 16^3 and 32^3, 0 the default).
 * Results do not depend on it.  No counterpart in the reference, whose
 * kernel shapes are fixed (compressible-cuda/src/amr.rs:572-818). */
int em_amr_set_shapes(em_amr* h, const int32_t shape[3]);
/* Project autotuner: times every pass shape of every active axis on the
 * residual of the current hierarchy and keeps the fastest (state and
 * counters untouched); shape_out (may be NULL) receives the choice. */
int em_amr_autotune(em_amr* h, int32_t repeats, int32_t shape_out[3]);

/* GpuAmr::profile / GpuAmr::prof (amr.rs:110-115): when on, record device
 * time per phase; em_amr_profile returns the accumulated milliseconds of
 * [restriction, ghost slabs, flux passes, refluxing + RK update, regrid
 * (host wall time)].  Profiled steps are launched eagerly. */
int em_amr_set_profile(em_amr* h, int32_t on);
int em_amr_profile(em_amr* h, double out[5]);

/* Step count of the next regrid check (Amr::next_regrid). */
int em_amr_set_next_regrid(em_amr* h, int64_t step);

/* Block and leaf counts (GpuAmr::leaf_count, amr.rs:855-857). */
int em_amr_nblocks(em_amr* h, int64_t* nblocks, int64_t* leaves);

/* GpuAmr::download (amr.rs:821-843): hierarchy and data of every block in id
 * order; any pointer may be NULL. */
int em_amr_download(em_amr* h, int32_t* level, int64_t* idx, uint8_t* leaf, double* u);

/* Replace the data of every block (same layout as download). */
int em_amr_upload(em_amr* h, const double* u);

/* Residual of every leaf for one stage (restriction of the edge ancestors,
 * ghost slabs, fluxes, refluxing; Amr::compute_rhs, amr.rs:1302-1323) with
 * the limiter's lambda from stage_dt; ancestors are written as zero.
 * stats = {faces, shock-path faces, fallbacks, limited faces} of this call. */
int em_amr_rhs(em_amr* h, double stage_dt, double* rhs, uint64_t stats[4]);

/* GpuAmr::step (amr.rs:683-700): SSP-RK3 step of size dt, regrid when due. */
int em_amr_step(em_amr* h, double dt);

/* nsteps CFL steps with dt = cfl / rate computed on the device each step
 * (no host round trip between steps except regrids). */
int em_amr_run_steps(em_amr* h, int64_t nsteps);

/* GpuAmr::run_to (amr.rs:807-819). */
int em_amr_run_to(em_amr* h, double t_end, int64_t* steps);

/* GpuAmr::dt_cfl (amr.rs:769-805): cfl / max over leaf cells of
 * sum_d (|u_d| + c) / dx_d(level). */
int em_amr_dt_cfl(em_amr* h, double* dt);

/* Wavelet details of every block in id order (0 for ancestors), after
 * restricting the edge ancestors (GpuAmr::details, amr.rs:702-723). */
int em_amr_details(em_amr* h, double* det);

/* Amr::regrid (compressible/src/amr.rs:1039-1046) on the device data. */
int em_amr_regrid(em_amr* h);

/* GpuAmr::counters (amr.rs:850-853). */
int em_amr_counters(em_amr* h, uint64_t out[3]);

/* GpuAmr::t (amr.rs:845-847). */
int em_amr_time(em_amr* h, double* t);

int em_amr_synchronize(em_amr* h);

/* Hierarchies of uneven block size in one batched step (no counterpart in
 * the reference, whose hierarchy has one block size, amr.rs:49-63): a group
 * of AMR handles, each with its own block size (8, 16 or 32 cells a side),
 * dt, tables and regrid schedule, stepped together.  Per RK stage the flux
 * pass is one persistent launch per axis over the leaves of every member
 * (all block sizes); restriction, ghosts, refluxing and the RK update are
 * the members' own launches.  Members must share the flux mode, arithmetic
 * variant and active axes and run batched; the group borrows them (destroy
 * the group first) and a member's results are those of em_amr_run_steps on
 * it alone, bit for bit.  Error messages come from em_amr_last_error(). */
typedef struct em_amr_group em_amr_group;
int em_amr_group_create(em_amr_group** out, em_amr* const* members, int32_t n);
int em_amr_group_destroy(em_amr_group* g);
/* nsteps CFL steps of every member (each member's dt from its own state). */
int em_amr_group_run_steps(em_amr_group* g, int64_t nsteps);

/* ------------------------------------------------------------------------
 * Two-phase Kapila five-equation system on a uniform grid: drop-in for
 *   GpuSolver<f64> built with GpuSystem::FiveEq { eos, kapila }
 *   (fluid/crates/compressible-cuda/src/lib.rs:50-82, 155-477),
 * with the arithmetic of the CPU Solver<7, FiveEq>
 * (fluid/crates/compressible/src/multiphase.rs), its parity target.
 * Conserved variables (a1, a2, rho u, rho v, rho w, E, alpha1) with
 * a_k = alpha_k rho_k; host arrays are ghost-padded AoS with 7 values per
 * cell, same cell order as the Euler solver.  Same error convention; the
 * messages come from em_fiveeq_last_error(). */
typedef struct em_fiveeq em_fiveeq;

const char* em_fiveeq_last_error(void);

/* GpuSolver::from_cpu for FiveEq (lib.rs:186-282).
 *   phases[8]     gamma, p_inf, b, q of phase 1 then phase 2 (Nasg, eos.rs)
 *   kapila        1: Kapila's K term in the volume-fraction equation
 *                 (multiphase.rs:92-104), 0: Allaire's model
 *   bc_state[42]  [axis][side][7] conserved Dirichlet / far-field states
 * other arguments as em_solver_create; u has 7 values per cell. */
int em_fiveeq_create(em_fiveeq** out, const int64_t n[3], const double lo[3], const double hi[3], int32_t ng,
                     const double phases[8], int32_t kapila, const int32_t bc_kind[6], const double bc_state[42],
                     int32_t mode, int32_t positivity, double cfl, double tol, double t0, const double* u);
/* GpuSolver::<f32>::from_cpu with GpuSystem::FiveEq (lib.rs:26-48 Real for
 * f32; the reference compiles its FiveEq kernels for float, kernels.cu:1157):
 * state, residual and arithmetic in FP32 on the device, phase parameters,
 * boundary states, 1/dx and the tolerance rounded to float, the limiter's
 * lambda formed in double from dt and rounded, the float WENO-Z epsilon
 * 1e-30 (kernels.cu:263).  Host arrays stay double (upload rounds, download
 * and rhs widen); the handle takes every em_fiveeq_* call.  Built without
 * multiply-add contraction, like the FP64 kernels of this system. */
int em_fiveeq_create_f32(em_fiveeq** out, const int64_t n[3], const double lo[3], const double hi[3], int32_t ng,
                         const double phases[8], int32_t kapila, const int32_t bc_kind[6], const double bc_state[42],
                         int32_t mode, int32_t positivity, double cfl, double tol, double t0, const double* u);
int em_fiveeq_destroy(em_fiveeq* h);
/* upload / download the conserved state (download leaves the ghosts zero) */
int em_fiveeq_upload(em_fiveeq* h, const double* u);
int em_fiveeq_download(em_fiveeq* h, double* u);
/* Per-stage residual of the current state, including the non-conservative
 * (alpha1 + K) div u term (solver.rs:350-475); stats as em_solver_rhs. */
int em_fiveeq_rhs(em_fiveeq* h, double stage_dt, double* rhs, uint64_t stats[4]);
/* GpuSolver::step / dt_cfl / run_to: dt_cfl includes the dilatation-rate
 * limit of the stiff non-conservative term (solver.rs:478-523). */
int em_fiveeq_step(em_fiveeq* h, double dt);
int em_fiveeq_dt_cfl(em_fiveeq* h, double* dt);
int em_fiveeq_run_to(em_fiveeq* h, double t_end, int64_t* steps);
/* Tile shape per axis (index into the FiveEq tile table) and the project
 * autotuner over it, as em_solver_set_tiles / em_solver_autotune (no
 * reference counterpart).  cfg_out may be NULL. */
int em_fiveeq_set_tiles(em_fiveeq* h, const int32_t cfg[3]);
int em_fiveeq_autotune(em_fiveeq* h, int32_t repeats, int32_t cfg_out[3]);
int em_fiveeq_counters(em_fiveeq* h, uint64_t out[3]);
int em_fiveeq_time(em_fiveeq* h, double* t);
int em_fiveeq_synchronize(em_fiveeq* h);

#ifdef __cplusplus
}
#endif
#endif /* EM_HIP_H */

/*
 * glide_hip.h - C ABI of the MI355X-native GLIDE SSA / FASCD solver.
 *
 * All arrays are FP64, row-major MAC-grid fields in device (HBM) memory:
 *   u  : ny x (nx+1)  x-velocity on the left face of each cell
 *   v  : (ny+1) x nx  y-velocity on the top face of each cell
 *   H  : ny x nx      thickness (and every other cell field)
 * A "state vector" is [u | v | H] contiguous (n_total = nu + nv + nc
 * values), the layout of glide/grid.py:79-165.
 *
 * Conventions for every entry point:
 *   - plain device pointers and sizes, no framework types;
 *   - no memory is allocated except by the *_create calls: scratch buffers
 *     are caller-provided (a batch plan and a V-cycle workspace own theirs);
 *   - work is enqueued on the caller's stream (hipStream_t passed as void*,
 *     NULL = the default stream); only the entry points that read results
 *     back to the host synchronise with it: glide_residual_norm,
 *     glide_count_nonfinite, glide_batch_residual_norms, and the FASCD
 *     V-cycles when their schedule asks for the idle-pass check
 *     (GlideSchedule.local_skip);
 *   - return 0 on success, a negative GLIDE_E_* code for invalid arguments,
 *     or a positive hipError_t from the runtime; glide_last_error() copies
 *     the message of the last failure on the calling thread.
 */
#ifndef GLIDE_HIP_H
#define GLIDE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GLIDE_OK 0
#define GLIDE_E_NULL (-1)      /* required pointer is NULL */
#define GLIDE_E_SIZE (-2)      /* bad grid size / level count */
#define GLIDE_E_ARG (-3)       /* other invalid argument */
#define GLIDE_E_NONFINITE (-4) /* non-finite input detected */

/* Physical parameters: replaces the PhysicsParams struct
 * (glide/cuda/common.cu:240-247) packed by make_physics_params
 * (glide/kernels.py:10-49); FP64 here. */
typedef struct GlideParams {
    double n;            /* Glen exponent */
    double eps_reg;      /* strain-rate regularisation */
    double water_drag;   /* drag of floating ice */
    double calving_rate; /* calving rate of floating ice */
    double gl_sigmoid_c; /* grounding-line sigmoid sharpness */
    int gl_derivatives;  /* include d(sigmoid)/dH in Jacobians */
} GlideParams;

/* One grid level's geometry: the (ny, nx, dx, dt) kernel arguments of
 * glide/grid.py:195-203. */
typedef struct GlideGeom {
    int ny, nx;
    double dx, dt;
} GlideGeom;

/* All buffers of one multigrid level (glide/grid.py:79-158). Adjoint
 * buffers may be NULL when no adjoint is run. */
typedef struct GlideLevel {
    GlideGeom g;
    GlideParams p;
    double *U, *f, *r, *w, *z;                   /* n_total each */
    double *Lambda, *Lambda_out, *f_adj, *r_adj; /* n_total each (adjoint) */
    double *H_prev, *bed, *B, *beta, *smb, *gamma, *chi, *mask; /* nc each */
    double *scratch;      /* 5 * nc doubles (Vanka per-cell updates; the small-level
                             sweeps' alternate state) */
    unsigned char *active;/* (ny/14+1)*(nx/14+1) bytes: local-smoother blocks */
} GlideLevel;

/* One FASCD / adjoint V-cycle's settings.  The sweep counts are hard-coded
 * in the reference (glide/solver.py:96,109,144,147-155,177,183,219;
 * glide/grid.py:305); the values in the comments are the reference's and
 * what a NULL schedule means.
 *
 * boundary_weight: factor on the boundary-face rows (u faces j = 0 and
 *   j = nx, v faces i = 0 and i = ny) of the fine residual restricted to the
 *   child (glide/solver.py:34-44,119-121 restricts them like interior rows:
 *   1.0).  A boundary face borders one cell, so its momentum row is a
 *   one-sided stress over dx (scaling as 1/h) where interior rows are stress
 *   differences over dx (1/h^2); with 1.0 the coarse correction of those
 *   rows is twice too strong and the V-cycle stalls wherever the solution
 *   carries normal stress at the domain edge.  0.5 gives the interior rows'
 *   scaling.  Values other than 1.0 depart from the reference's iterates.
 * local_skip: nonzero stops a level's local passes at the first pass whose
 *   activation marks no block (glide/solver.py:147-155 always runs all of
 *   them; an idle pass changes nothing, and neither can any pass after it).
 *   The host reads each pass's 4-byte any-block flag: it waits for pass 0's
 *   before enqueuing its sweeps and for pass k-1's (already computed) when
 *   enqueuing pass k, so at most one idle pass per level is still run, and
 *   the V-cycle call returns only after the GPU has reached its last local
 *   pass.  The only difference from running every pass is the sign of zero
 *   entries of U that idle passes would turn from -0.0 into +0.0.  The
 *   single-domain V-cycle needs a GlideVcycleWork for it; batch plans own
 *   their flags. */
typedef struct GlideSchedule {
    int pre, post, coarse;          /* 10, 10, 200 */
    int local_passes, local_sweeps; /* 10, 10 */
    int n_newton;                   /* 10 */
    int adj_sweeps, adj_coarse;     /* 10, 100 */
    double boundary_weight;         /* 1.0 */
    int local_skip;                 /* 0 */
} GlideSchedule;

/* ---- level operators ------------------------------------------------- */

/* r = F(U) - f on one level (f == NULL gives F(U), i.e. compute_F);
 * when sub != NULL, r = F(U) - f - sub (the FAS right-hand side
 * f_c = F_c(U_c) - r_c in one pass).
 * Replaces kernel compute_residual (glide/cuda/residuals.cu:6-292) as
 * launched by Grid.compute_residual / compute_F (glide/grid.py:190-223). */
int glide_residual(const GlideGeom *g, const GlideParams *p, double *r, const double *U,
                   const double *f, const double *sub, const double *bed, const double *B,
                   const double *beta, const double *mask, const double *gamma, void *stream);

/* j = J(U) dU (exact forward-mode linearisation of the residual).
 * Replaces kernel compute_jvp (glide/cuda/residuals.cu:299-579),
 * Grid.compute_jvp (glide/grid.py:225-238). */
int glide_jvp(const GlideGeom *g, const GlideParams *p, double *j, const double *U,
              const double *dU, const double *bed, const double *B, const double *beta,
              const double *mask, const double *gamma, void *stream);

/* l = J(U)^T lambda, gather form (deterministic, no atomics).
 * Replaces kernel compute_vjp (glide/cuda/residuals.cu:586-1023),
 * Grid.compute_vjp (glide/grid.py:240-254, including its l.fill(0)). */
int glide_vjp(const GlideGeom *g, const GlideParams *p, double *l, const double *U,
              const double *lam, const double *bed, const double *B, const double *beta,
              const double *mask, const double *gamma, void *stream);

/* n_sweeps forward Vanka sweeps, each = per-cell damped Newton on the 5
 * local unknowns with eta frozen, then U += 0.5 * delta.  active == NULL
 * smooths every cell; otherwise only cells whose 14x14 reference block has
 * active[block] != 0 (see glide_local_activation).  scratch: 5*nc doubles.
 * Replaces kernels vanka_smooth / vanka_smooth_local
 * (glide/cuda/vanka.cu:3-439, 844-1295) and the host loops
 * Grid.vanka_sweep / vanka_sweep_local (glide/grid.py:256-317). */
int glide_vanka_sweep(const GlideGeom *g, const GlideParams *p, double *U, const double *f,
                      const double *bed, const double *B, const double *beta,
                      const double *gamma, const unsigned char *active, int n_sweeps,
                      int n_newton, double *scratch, void *stream);

/* glide_vanka_sweep without the scratch argument, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): no counterpart in the reference.  The per-cell
 * updates (and the small levels' alternate state, see
 * glide_set_small_levels) go to a static device array of 5 x 2^20 doubles (40 MiB) that
 * each GPU the library runs on holds until the process ends, so levels up
 * to 2^20 cells (1024 x 1024) are accepted and larger ones fail with
 * GLIDE_E_SIZE.  Nothing is allocated; the array is shared by every call on
 * a GPU, so calls that overlap on one GPU (other streams, other threads)
 * race and must be serialised by the caller.  Same results as
 * glide_vanka_sweep, bit for bit. */
int glide_vanka_sweep_noalloc(const GlideGeom *g, const GlideParams *p, double *U, const double *f,
                              const double *bed, const double *B, const double *beta,
                              const double *gamma, const unsigned char *active, int n_sweeps,
                              int n_newton, void *stream);

/* Local 5x5 block (J, r) of every cell after k_updates damped Newton
 * updates of its 5 unknowns (out: 30 doubles per cell, J row-major then r).
 * Exposes the per-cell assembly of glide/cuda/vanka.cu:56-411 for testing. */
int glide_vanka_block(const GlideGeom *g, const GlideParams *p, const double *U, const double *f,
                      const double *bed, const double *B, const double *beta, const double *gamma,
                      int k_updates, double *out, void *stream);

/* Block activation of the local smoother: error = |r_u right| + |r_u left|
 * + |r_v bottom| + |r_v top| + |FB(r_H, H - gamma)| > 0.01 anywhere in a
 * 14x14 block's 16x16 launch footprint.  Replaces the error-mask driver
 * (glide/solver.py:147-154) plus the block test of vanka_smooth_local
 * (glide/cuda/vanka.cu:870-888). */
int glide_local_activation(const GlideGeom *g, const double *r, const double *U,
                           const double *gamma, unsigned char *active, void *stream);

/* One colour of the adjoint red-black Vanka smoother: lam_out = lam +
 * omega * (1/2 faces, 1 H) * J_loc^-T r_adj on cells with (i+j)%2 == color.
 * lam_out must hold a copy of lam on entry.  Replaces kernel
 * vanka_smooth_adjoint (glide/cuda/vanka.cu:441-840) / Grid.vanka_smooth_adjoint
 * (glide/grid.py:286-303). */
int glide_vanka_adjoint_pass(const GlideGeom *g, const GlideParams *p, double *lam_out,
                             const double *lam, const double *mask, const double *r_adj,
                             const double *U, const double *bed, const double *B,
                             const double *beta, const double *gamma, int color,
                             double omega, void *stream);

/* grad = lambda . dR/dbeta.  reference_edges != 0 reproduces the
 * reference's skipped edge terms (glide/cuda/grad.cu:52-53,70-71);
 * 0 gives the exact transpose.  Replaces kernel compute_grad_beta
 * (glide/cuda/grad.cu:1-74) / Grid.compute_grad_beta (glide/grid.py:336-355). */
int glide_grad_beta(const GlideGeom *g, const GlideParams *p, double *grad, const double *U,
                    const double *lam, const double *bed, const double *beta,
                    int reference_edges, void *stream);

/* grad = lambda . dR/dB (B: the rate factor of the viscosity).  No reference
 * counterpart: glide/physics.py:211-242 returns dL/dbeta only.  B enters the
 * residual only through eta (glide/cuda/viscosity.cu:61-112), d eta / d B =
 * eta / B, and eta through the eta*H of sigma_xx, sigma_yy and the vertex
 * sigma_xy (glide/cuda/stress.cu:44-183). */
int glide_grad_B(const GlideGeom *g, const GlideParams *p, double *grad, const double *U,
                 const double *lam, const double *B, void *stream);

/* Transfer operators (glide/cuda/utility.cu; launchers glide/kernels.py:90-236).
 * Sizes are the coarse (restrict) or fine (prolong) grid. */
int glide_restrict_u(const double *uf, double *uc, int nyc, int nxc, void *stream);     /* restrict_u :1-19 */
int glide_restrict_v(const double *vf, double *vc, int nyc, int nxc, void *stream);     /* restrict_v :125-143 */
int glide_restrict_cell(const double *hf, double *hc, int nyc, int nxc, void *stream);  /* restrict_cell_centered :56-75 */
int glide_restrict_max(const double *hf, double *hc, int nyc, int nxc, void *stream);   /* restrict_max_pool :77-101 */
int glide_prolong_inject(const double *hc, double *hf, int nyf, int nxf, double sign,
                         void *stream);                                                 /* prolongate_cell_centered :104-123 */
int glide_prolong_u(const double *uc, double *uf, int nyf, int nxf, void *stream);      /* prolongate_u_smooth :180-232 */
int glide_prolong_v(const double *vc, double *vf, int nyf, int nxf, void *stream);      /* prolongate_v_smooth :235-287 */
int glide_prolong_H(const double *hc, double *hf, int nyf, int nxf, void *stream);      /* prolongate_H_smooth :289-337 */

/* ---- hierarchy drivers ------------------------------------------------ */

/* Restrict bed, B, beta, H_prev, smb down the hierarchy
 * (glide/solver.py:53-68, restrict_parameters_to_hierarchy). */
int glide_restrict_parameters(const GlideLevel *levels, int n_levels, void *stream);

/* Restrict bed, H (into U), H_prev and reset gamma = thklim on the children
 * (glide/physics.py:131-139). */
int glide_propagate_geometry(const GlideLevel *levels, int n_levels, double thklim,
                             void *stream);

/* Host-side state of the single-domain V-cycle's idle-pass check: two
 * pinned 4-byte flag slots, two events and the pass counters.  Created for
 * the device current at creation (the only allocating V-cycle call; no
 * reference counterpart).  A workspace serves one V-cycle at a time: calls
 * that share one must not overlap (one host thread, or the caller's lock);
 * moving it to another stream is safe (the first call there waits for flag
 * copies still in flight on the previous one). */
typedef struct GlideVcycleWork GlideVcycleWork;
int glide_vcycle_work_create(GlideVcycleWork **out);
int glide_vcycle_work_destroy(GlideVcycleWork *work);

/* Local passes run and skipped by V-cycles that used this workspace since
 * the last reset (reset != 0 zeroes them after reading). */
int glide_vcycle_work_counts(GlideVcycleWork *work, long *run, long *skipped, int reset);

/* One FASCD V-cycle from the finest level (glide/solver.py:71-157,
 * fascd_vcycle(grid, thklim, finest=True)).  sched == NULL uses the
 * reference's settings.  work may be NULL unless sched->local_skip is set;
 * without the idle-pass check the call never synchronises or allocates. */
int glide_fascd_vcycle(const GlideLevel *levels, int n_levels, double thklim,
                       const GlideSchedule *sched, GlideVcycleWork *work, void *stream);

/* f_H = H_prev / dt + smb on the finest level and dt on every level
 * (glide/physics.py:166-173). */
int glide_forward_setup(GlideLevel *levels, int n_levels, double dt, void *stream);

/* ||(r_u, r_v, FB(r_H, H - gamma))||_2 of the finest level into *out_host
 * (the verbose norm of glide/physics.py:176-203); writes the residual into
 * level->r and partial sums into level->scratch, synchronises the stream. */
int glide_residual_norm(const GlideLevel *level, double *out_host, void *stream);

/* Number of NaN / +-Inf entries among x[0..n) into *count_host, using the
 * caller's 8-byte device counter; synchronises the stream.  No reference
 * counterpart (the reference never checks its inputs): IcePhysics uses it to
 * reject non-finite geometry and parameters. */
int glide_count_nonfinite(const double *x, long n, unsigned long long *counter, unsigned long long *count_host,
                          void *stream);

/* One adjoint V-cycle (glide/solver.py:160-220). */
int glide_adjoint_vcycle(const GlideLevel *levels, int n_levels, const GlideSchedule *sched,
                         void *stream);

/* mask = (H <= gamma + 0.1) on one level: the adjoint's active set
 * (glide/grid.py:332-334). */
int glide_compute_mask(const GlideLevel *level, void *stream);

/* r_adj = f_adj - J(U)^T Lambda on one level, with the level's current mask
 * (glide/solver.py:187-189, glide/grid.py:322-324). */
int glide_adjoint_residual(const GlideLevel *level, void *stream);

/* ---- batched ensembles ------------------------------------------------
 * An ensemble of independent members, each a level hierarchy of its own
 * size (GlideLevel arrays as above, finest first), solved together: every
 * operation of the V-cycle is one persistent launch over all members' tiles.
 * Member m's level k runs in phase k + (L_max - L_m): hierarchies are aligned
 * at their coarsest level.  The reference has no batched path: a member's
 * results equal those of glide_fascd_vcycle / glide_forward_setup /
 * glide_residual_norm applied to that member alone (bitwise).  Not
 * thread-safe per plan; distinct plans may run on distinct streams. */
typedef struct GlideBatch GlideBatch;

/* Kernel ids for the tile autotuner (glide_batch_set_tiles / _time_kernel). */
#define GLIDE_BK_VANKA 0          /* full Vanka sweep, tile height 1/2/4/8 */
#define GLIDE_BK_VANKA_LOCAL 1    /* local Vanka sweep, tile height 1/2/4/8 */
#define GLIDE_BK_RESIDUAL 2       /* residual, tile height 2/4/8/16 */
#define GLIDE_BK_RESIDUAL_CHILD 3 /* FAS right-hand side on the child level */

/* Creates the plan: copies the member descriptors, builds the tile tables and
 * allocates the plan's own small device tables and pinned staging (the only
 * allocating batched call).  members[m] points to n_levels[m] levels. */
int glide_batch_create(const GlideLevel *const *members, const int *n_levels, int n_members,
                       GlideBatch **out);
int glide_batch_destroy(GlideBatch *batch);

/* Phase count, entries (members present) and cells per phase. */
int glide_batch_info(const GlideBatch *batch, int *n_phases, int *entries_per_phase,
                     int *cells_per_phase);

/* dt on every level of every member and f_H = H_prev / dt + smb on each
 * finest level: glide_forward_setup for all members (glide/physics.py:166-173). */
int glide_batch_forward_setup(GlideBatch *batch, double dt, void *stream);

/* One FASCD V-cycle of every enabled member (glide/solver.py:71-157).
 * With sched->local_skip a pass is idle when no member's activation marks a
 * block; the flags are the plan's. */
int glide_batch_vcycle(GlideBatch *batch, double thklim, const GlideSchedule *sched,
                       void *stream);

/* Per-member ||(r_u, r_v, FB(r_H, H - gamma))||_2 of the finest level into
 * out_host[n_members] (glide/physics.py:176-203); synchronises the stream. */
int glide_batch_residual_norms(GlideBatch *batch, double *out_host, void *stream);

/* r = F(U) - f on every entry of one phase (phase n_phases = every member's
 * finest level): the ensemble-wide residual stencil sweep. */
int glide_batch_residual(GlideBatch *batch, int phase, void *stream);

/* on[n_members]: members with on[m] == 0 are skipped by every batched call
 * (converged members stop iterating); synchronises the stream. */
int glide_batch_set_enabled(GlideBatch *batch, const unsigned char *on, void *stream);

/* Autotuner hooks: tile height ty and workgroups per CU (0 = register-
 * budget default) of one kernel in one phase; _time_kernel runs it reps
 * times with those settings (state-preserving: the Vanka timing writes only
 * scratch, the residual timings only r and the child's f) and returns the
 * mean time per launch in ms, leaving the plan's settings unchanged. */
int glide_batch_set_tiles(GlideBatch *batch, int kernel, int phase, int ty, int wg_per_cu);
int glide_batch_time_kernel(GlideBatch *batch, int kernel, int phase, int ty, int wg_per_cu,
                            int reps, double *ms, void *stream);

/* Times the batched restriction of the residual (r_c = R r, every member of
 * `phase` that has a child level, one launch) over reps launches on the
 * stream: *ms per launch and *bytes = 8 (N_total + N_total,child) summed
 * over those members, the compulsory traffic.  Writes the child levels' r
 * only (a V-cycle recomputes it before use).  The restriction is the
 * reference's residual transfer (glide/solver.py:119-121 with
 * glide/cuda/utility.cu:1-19,125-143), here as a benchmark hook: no
 * reference counterpart. */
int glide_batch_time_restriction(GlideBatch *batch, int phase, int reps, double *ms, double *bytes,
                                 void *stream);

/* Workgroups per CU of the persistent single-level launches: kernel 0 =
 * the Vanka sweep (default 64, i.e. one tile per workgroup), 1 = the
 * residual (default 8).  Each launch has
 * min(tiles, CUs x wg_per_cu) workgroups looping over its tiles.  A launch
 * shape only: results do not depend on it.  Process-wide (atomic); no
 * reference counterpart. */
int glide_set_persistence(int kernel, int wg_per_cu);

/* Tile heights of the single-level stencils (tiles of 64 x ty points):
 * kernel 0 = the residual (4, 8 or 16; default 0 = by level size: 16 from
 * 2^16 to 2^18 cells, 8 above, 4 below), 1 = the JVP and VJP
 * (4, 6 or 8, default 4; 0 = the per-point form without LDS staging).  A launch
 * shape only: results are bitwise the same.  Process-wide (atomic); no
 * reference counterpart. */
int glide_set_tile_height(int kernel, int ty);

/* Vanka kernel of the small levels, whose sweeps are bound by the latency
 * of one wave's Newton chain: levels with at most max_cells cells (default
 * 128 x 128) sweep with `variant` 0 = one lane per cell with the whole
 * register file, 1 = row split (five waves per 64 cells, each assembling one
 * row of the local block), or 2 (default) = variant 0 with the apply pass
 * folded into the sweep (each workgroup also runs the ring of cells around
 * the 62 x 2 cells it owns; the state alternates between U and the scratch);
 * larger levels sweep one lane per cell at two waves per SIMD.  A launch
 * shape only: results are bitwise the same.
 * Process-wide (atomic); no reference counterpart. */
int glide_set_small_levels(long max_cells, int variant);

/* Copies the message of the last failure on this thread ("" if none) into
 * buf (at most cap - 1 characters and a terminating NUL; buf may be NULL
 * when cap is 0) and returns the message's full length. */
size_t glide_last_error(char *buf, size_t cap);

/* Library version string. */
const char *glide_version(void);

#ifdef __cplusplus
}
#endif

#endif

/* ubg_c.h -- C ABI of the MI355X-native ubootgl fluid step (libubg.so).
 *
 * The boundary replaces the reference's in-process C++ classes
 *   Simulation        simulation.hpp:18-141
 *   MG (pressure solver, there is no class named PressureSolver)
 *                     pressure_solver.hpp:13-76
 *   rbgs / calculateResidualField   pressure_solver.hpp:3-11
 * with opaque handles that own device-resident fields.  Every entry point
 * returns an int status (UBG_OK = 0); no C++ exception crosses the ABI and the
 * message of the last failure on the calling thread is ubg_last_error().
 *
 * Fields are float32 row-major arrays, index y*width + x, exactly the layout of
 * Single2DGrid (db2dgrid.hpp:12-50).  Staggered sizes: vx is (W-1) x H, vy is
 * W x (H-1), everything else W x H.  Flags are 0/1 (any other value is
 * rejected with UBG_E_FLAG_VALUE); on the device they are stored as one byte
 * per cell.
 *
 * Threading: one handle is owned by one host thread at a time; calls on a
 * handle are ordered on the handle's own HIP stream.  Host-pointer entry
 * points copy synchronously; compute entry points are asynchronous until
 * ubg_*_sync() or a host copy.
 */
#ifndef UBG_C_H
#define UBG_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UBG_API __attribute__((visibility("default")))

/* ---- status codes ------------------------------------------------------ */
#define UBG_OK 0
#define UBG_E_INVALID 1          /* bad argument or handle                  */
#define UBG_E_DOMAIN_TOO_SMALL 2 /* W or H < 8: the reference's MG indexes   */
                                 /* past its level arrays there             */
#define UBG_E_FLAG_VALUE 3       /* a flag value other than 0 or 1          */
#define UBG_E_HIP 4              /* HIP runtime failure                     */
#define UBG_E_OOM 5              /* device allocation failed                */
#define UBG_E_NONFINITE 6        /* NaN/Inf detected in a norm              */
#define UBG_E_BUDGET 7           /* request exceeds the memory budget       */
#define UBG_E_NOT_CONVERGED 8    /* iteration limit reached                 */
#define UBG_E_STATE 9            /* call not valid in the current state     */
#define UBG_E_NO_DEVICE 10       /* no HIP device visible                   */

/* ---- field, stage and operation ids ------------------------------------ */
enum {
  UBG_FIELD_VX = 0,      /* DoubleBuffered2DGrid vx, front  (simulation.hpp:128) */
  UBG_FIELD_VX_BACK,     /* vx back buffer: live state (db2dgrid.hpp:52-109)     */
  UBG_FIELD_VY,
  UBG_FIELD_VY_BACK,
  UBG_FIELD_VX_ACCUM,    /* simulation.hpp:129 */
  UBG_FIELD_VY_ACCUM,
  UBG_FIELD_VX_CURRENT,  /* simulation.hpp:130 */
  UBG_FIELD_VY_CURRENT,
  UBG_FIELD_P,           /* simulation.hpp:132 */
  UBG_FIELD_F,
  UBG_FIELD_FLAG,
  UBG_FIELD_R,
  UBG_FIELD_COUNT
};

enum {
  UBG_STAGE_ACCUM = 0,        /* applyAccumulatedVelocity  simulation.cpp:376-396 */
  UBG_STAGE_DIFFUSE,          /* diffuse                   simulation.cpp:104-162 */
  UBG_STAGE_ADVECT,           /* advect                    simulation.cpp:241-354 */
  UBG_STAGE_SET_VBCS,         /* setVBCs                   simulation.cpp:80-102  */
  UBG_STAGE_PROJECT,          /* project                   simulation.cpp:164-208 */
  UBG_STAGE_SET_PBC,          /* setPBC                    simulation.cpp:36-45   */
  UBG_STAGE_SAVE_CURRENT,     /* saveCurrentVelocityFields simulation.cpp:16-19   */
  UBG_STAGE_DIVERGENCE,       /* project's divergence + sinks simulation.cpp:166-187 */
  UBG_STAGE_CORRECT_VELOCITY, /* project's correction      simulation.cpp:196-207 */
  UBG_STAGE_MG_SOLVE,         /* one mg.solve(p,f,flag,h,true) simulation.cpp:189 */
  UBG_STAGE_COUNT
};

enum {
  UBG_MGOP_RBGS = 0,          /* rbgs, one red+black sweep   pressure_solver.cpp:49-89   */
  UBG_MGOP_ZERO_GRADIENT,     /* setZeroGradientBC           pressure_solver.cpp:183-192 */
  UBG_MGOP_RESIDUAL,          /* calculateResidualField -> r pressure_solver.cpp:91-116  */
  UBG_MGOP_RESTRICT_RESIDUAL, /* residual + restrict into level+1's f, level+1's p = 0
                                 pressure_solver.cpp:217-227 */
  UBG_MGOP_PROLONG_CORRECT,   /* prolongate level+1's p, add to level's p
                                 pressure_solver.cpp:231-235 */
  UBG_MGOP_COUNT
};

enum { UBG_LEVEL_P = 0, UBG_LEVEL_F = 1, UBG_LEVEL_R = 2 };

/* Solver parameters; ubg_params_default() gives the reference's values. */
typedef struct ubg_params {
  int nu_pre;           /* pre-smoothing sweeps, 3   (pressure_solver.cpp:211) */
  int nu_post;          /* post-smoothing sweeps, 3  (pressure_solver.cpp:243) */
  int nu_coarse;        /* sweeps on level L-2, 5    (pressure_solver.cpp:204) */
  int vcycles_per_step; /* mg.solve calls per step, 2 (simulation.cpp:189-190) */
  float alpha;          /* smoother relaxation, 1.0  (pressure_solver.cpp:204) */
  int use_graph;        /* replay the step as a captured hipGraph, 1 */
  int save_current;     /* copy vx/vy to *_current at step end, 1 */
} ubg_params;

typedef struct ubg_mg ubg_mg;
typedef struct ubg_sim ubg_sim;

/* ---- library ------------------------------------------------------------ */
UBG_API int ubg_version(void);
UBG_API const char *ubg_last_error(void);
UBG_API void ubg_params_default(ubg_params *out);
UBG_API int ubg_device_count(int *count);
/* Level sizes of the MG hierarchy for a W x H grid: halve (integer) while
 * both sizes exceed 3 (pressure_solver.hpp:16-31).  Returns the count. */
UBG_API int ubg_mg_level_sizes(int width, int height, int *widths, int *heights, int cap);

/* ---- MG: pressure_solver.hpp:13-76 -------------------------------------- */
/* MG(w,h): hierarchy with all-fluid coarse flags (pressure_solver.hpp:16-31) */
UBG_API int ubg_mg_create(int width, int height, ubg_mg **out);
UBG_API int ubg_mg_destroy(ubg_mg *mg);
UBG_API int ubg_mg_set_params(ubg_mg *mg, const ubg_params *prm);
UBG_API int ubg_mg_num_levels(const ubg_mg *mg, int *levels);
/* updateFields(flag) (pressure_solver.hpp:34-57): level-0 flag and the coarse
 * classification of every level, on the device. */
UBG_API int ubg_mg_update_fields(ubg_mg *mg, const float *flag);
/* Level fields (p = ecs/p, f = rcs/f, r = residual scratch) as host copies. */
UBG_API int ubg_mg_set_level_field(ubg_mg *mg, int level, int which, const float *src);
UBG_API int ubg_mg_get_level_field(ubg_mg *mg, int level, int which, float *dst);
UBG_API int ubg_mg_get_level_flag(ubg_mg *mg, int level, float *dst);
/* One stage of the V-cycle on the resident level fields (stage-wise parity). */
UBG_API int ubg_mg_level_op(ubg_mg *mg, int op, int level, float h, float alpha);
/* ncycles x solve(p, f, flag, h, zeroGradientBC) on the resident level-0
 * fields (pressure_solver.hpp:59-62, pressure_solver.cpp:201-248). */
UBG_API int ubg_mg_solve(ubg_mg *mg, float h, int zero_gradient_bc, int ncycles);
/* MG::solve(p, f, flag, h, zeroGradientBC) with host grids: uploads p, f and
 * (if non-NULL) the level-0 flag, runs one V-cycle, downloads p.  Coarse
 * flags come from the last update_fields, as in the reference. */
UBG_API int ubg_mg_solve_host(ubg_mg *mg, float *p, const float *f, const float *flag,
                              float h, int zero_gradient_bc);
/* rbgs(p, f, flag, h, alpha) (pressure_solver.hpp:5-6), n sweeps, level 0. */
UBG_API int ubg_mg_rbgs(ubg_mg *mg, float h, float alpha, int nsweeps);
/* gs(p, f, flag, h, alpha) (pressure_solver.hpp:3-4, pressure_solver.cpp:
 * 26-33): n lexicographic Gauss-Seidel sweeps on level 0 (x fastest, then
 * y), the values of the sequential sweep. */
UBG_API int ubg_mg_gs(ubg_mg *mg, float h, float alpha, int nsweeps);
/* calculateResidualField (pressure_solver.hpp:9-10) on level 0: writes r
 * (if dst != NULL) and the L2 norm accumulated in double. */
UBG_API int ubg_mg_residual(ubg_mg *mg, float h, float *r_dst, double *norm);
UBG_API int ubg_mg_sync(ubg_mg *mg);

/* ---- allocation-free solve ------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  ncycles x MG::solve(p, f, flag, h, zeroGradientBC)
 * (pressure_solver.hpp:59-62) on caller-owned DEVICE arrays: p (W x H fp32,
 * updated in place), f (W x H fp32, read only) and flag (W x H bytes, 0 solid
 * / nonzero fluid), with the hierarchy built from the flag on every call (MG
 * (flag): updateFields, pressure_solver.hpp:32-57).  The results equal
 * ubg_mg_solve_host's on a new MG of that flag.  prm: NULL for the reference's
 * smoothing counts (3/3/5, alpha 1).  stream: a hipStream_t, or NULL for the
 * default stream; the call only enqueues work there (no allocation, memset
 * or synchronisation), so it may be captured into a graph.
 * Every intermediate buffer lives in one static device array of
 * UBG_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a solve
 * needs about 10.7 bytes per cell (W*H up to about 3.1 M cells, e.g.
 * 2048 x 1024); larger sizes fail with UBG_E_BUDGET.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams or
 * host threads) race and must be ordered by the caller. */
#define UBG_NOALLOC_SCRATCH_BYTES (32u << 20)
UBG_API int ubg_mg_solve_device_noalloc(float *p, const float *f, const uint8_t *flag, int width,
                                        int height, float h, int zero_gradient_bc, int ncycles,
                                        const ubg_params *prm, void *stream);

/* ---- FP64 iterative refinement around FP32 V-cycles ---------------------
 * No counterpart in the reference (an extension of MG::solve,
 * pressure_solver.hpp:59-62).  Solves the flag-masked Poisson problem of
 * calculateResidualField (pressure_solver.cpp:91-116),
 *     F(c) * (f(c) + sum_n F(n) (x(n) - x(c)) / h^2) = 0   on interior cells,
 * with x on the outer ring held as Dirichlet data (mgtest/mgtest.cpp:18-29),
 * in float64: x and the fine residual r are float64, each refinement step runs
 * FP32 V-cycles on A e = r (level-0 rhs = (float)r, e starts at 0) and adds e
 * to x in float64; the convergence norm ||r||_2 is summed in float64 in a
 * fixed order.  The rhs f is float32 (the problem's data).  The x of the
 * previous solve is the next solve's initial guess unless
 * ubg_mg_ir_set_guess is called; the level-0 p/f of the MG are scratch. */
#define UBG_IR_HISTORY 64
typedef struct ubg_ir_params {
  double tol;           /* stop when ||r_k||_2 <= tol * ||r_0||_2, 1e-10 */
  int vcycles_per_iter; /* FP32 V-cycles per refinement step, 4 */
  int max_iter;         /* refinement steps at most, 200 */
  double ref_norm;      /* > 0: residuals relative to this norm instead of
                           ||r_0|| (a warm start measured against ||F f||
                           of its rhs, as served solves are); 0 */
} ubg_ir_params;
typedef struct ubg_ir_result {
  int iterations;       /* refinement steps taken */
  int vcycles;          /* FP32 V-cycles run */
  int converged;        /* 1 if the tolerance was met */
  double r0_norm;       /* ||r_0||_2 of the initial guess (= ||F f|| for x = 0) */
  double rel_residual;  /* ||r_k||_2 / ||r_0||_2 at exit */
  double history[UBG_IR_HISTORY]; /* rel. residual after steps 1..64 */
} ubg_ir_result;
UBG_API void ubg_ir_params_default(ubg_ir_params *out);
/* the float32 rhs f, W x H host array (copied to the device) */
UBG_API int ubg_mg_ir_set_rhs(ubg_mg *mg, const float *f);
/* the float64 initial guess incl. the Dirichlet ring, W x H host array;
 * NULL sets x = 0 */
UBG_API int ubg_mg_ir_set_guess(ubg_mg *mg, const double *x);
/* refine until tol or max_iter; UBG_E_NOT_CONVERGED (result filled) if the
 * limit is reached or the residual stops decreasing */
UBG_API int ubg_mg_ir_solve(ubg_mg *mg, float h, const ubg_ir_params *prm, ubg_ir_result *res);
UBG_API int ubg_mg_ir_get_solution(ubg_mg *mg, double *x);
/* device addresses of the resident float64 solution and float32 rhs */
UBG_API int ubg_mg_ir_device_ptrs(ubg_mg *mg, double **x, float **f);

/* ---- Batched solves -------------------------------------------------------
 * No counterpart in the reference: n independent refinement solves (the
 * problem of ubg_mg_ir_solve) on domains of uneven sizes, advanced together.
 * Every multigrid stage is one launch over a tile list spanning all domains
 * (sweeps, residual+restriction, prolongation+correction; the coarse levels
 * of each domain in one LDS workgroup), and a refinement step of all domains
 * (V-cycles, FP64 update, per-domain norm and convergence test) is replayed as
 * a captured hipGraph; a converged domain drops out on the device.  Each
 * domain's result is bit-identical to ubg_mg_ir_solve on that domain alone.
 * Domains need W, H >= 16. */
typedef struct ubg_batch ubg_batch;
typedef struct ubg_batch_tuning {
  int tile_w, tile_h; /* smoother tile (cells), one of the compiled shapes;
                         used when a smoothing count exceeds 3            */
  int blocks;         /* workgroups per launch (0: one per tile)            */
  int pass_rows;      /* rows of the fused level-pass tiles: 8, 16 or 32,
                         or 0: chosen per level from the level's tile count
                         (the default); each domain's columns are split
                         evenly within the width the smoothing count allows */
  double sweep_ms;    /* measured time of the tuning pass (output)          */
} ubg_batch_tuning;
/* device bytes a batch of these domains allocates */
UBG_API int ubg_batch_bytes(int n, const int *widths, const int *heights, size_t *bytes);
UBG_API int ubg_batch_create(int n, const int *widths, const int *heights, const ubg_params *prm,
                             ubg_batch **out);
UBG_API int ubg_batch_destroy(ubg_batch *b);
/* domain d: flag (W x H, 0/1; the outer ring must be fluid), float32 rhs, h */
UBG_API int ubg_batch_set_domain(ubg_batch *b, int d, const float *flag, const float *rhs, float h);
UBG_API int ubg_batch_set_rhs(ubg_batch *b, int d, const float *rhs);
/* float64 initial guess incl. the Dirichlet ring; NULL = 0 */
UBG_API int ubg_batch_set_guess(ubg_batch *b, int d, const double *x);
/* solve every domain to prm->tol (relative to its own initial residual);
 * UBG_E_NOT_CONVERGED if any domain stopped short (see ubg_batch_result) */
UBG_API int ubg_batch_solve(ubg_batch *b, const ubg_ir_params *prm, int *graph_replays);
UBG_API int ubg_batch_result(ubg_batch *b, int d, ubg_ir_result *res);
UBG_API int ubg_batch_get_solution(ubg_batch *b, int d, double *x);
/* time the candidate shapes and workgroup counts on this batch (the fused
 * level passes' tile rows, or the smoother tiles when a smoothing count
 * exceeds 3) and keep the fastest; with db_path, reuse / append the choice
 * keyed by the batch's size signature (one "key tile_w tile_h blocks
 * pass_rows ms" line each) */
UBG_API int ubg_batch_autotune(ubg_batch *b, const char *db_path, ubg_batch_tuning *out);
UBG_API int ubg_batch_set_tuning(ubg_batch *b, const ubg_batch_tuning *t);
UBG_API int ubg_batch_get_tuning(ubg_batch *b, ubg_batch_tuning *t);

/* ---- Fluid components ----------------------------------------------------
 * With solid walls each fluid component of the flag is a pure-Neumann
 * problem; its rhs is compatible only if it sums to zero over the component.
 * No counterpart in the reference (its solver never makes the rhs
 * compatible).  Components under 4-neighbour connectivity (the coupling of
 * calculateResidualField, pressure_solver.cpp:100-108), labelled 0..n-1 in
 * raster order of each component's first cell, -1 on solid cells; integer
 * and exact (computed on the GPU by a lock-free union-find).  W x H host
 * arrays. */
UBG_API int ubg_flag_components(const float *flag, int width, int height, int32_t *labels,
                                int *ncomp);
/* f minus, on every fluid cell, the mean of f over its component: sums in
 * 64-bit fixed point (f 2^s rounded, s from max|f| and W*H so no sum
 * overflows; the result does not depend on summation order), means in
 * float64, f = (float)(f - mean); solid cells unchanged.  *ncomp may be
 * NULL. */
UBG_API int ubg_remove_component_means(const float *flag, int width, int height, float *f,
                                       int *ncomp);

/* ---- Batched projections --------------------------------------------------
 * The simulation's own pressure solve for many domains at once: MG::solve
 * (pressure_solver.cpp:250-257 -> solveLevel :201-248) repeated ncycles times
 * on every domain's float32 pressure, with setZeroGradientBC on level 0 as
 * Simulation::project calls it (simulation.cpp:189-190), one launch per
 * multigrid stage over all domains.  The reference steps one Simulation at a
 * time; each domain's result is bit-identical to ubg_mg_solve /
 * Simulation::step on that domain alone.  Needs 1..3 pre- and post-smoothing
 * sweeps (the reference's 3/3 included).
 *
 * a projection domain: any flag (the simulation's walls included), float32
 * rhs (the divergence), h; ubg_batch_solve refuses such a domain unless its
 * outer ring is fluid */
UBG_API int ubg_batch_set_projection_domain(ubg_batch *b, int d, const float *flag,
                                            const float *rhs, float h);
/* the float32 pressure the V-cycles start from (W x H host array; NULL = 0),
 * and the one they leave */
UBG_API int ubg_batch_set_pressure(ubg_batch *b, int d, const float *p);
UBG_API int ubg_batch_get_pressure(ubg_batch *b, int d, float *p);
/* ncycles V-cycles of every domain on its pressure and rhs; zero_gradient_bc
 * as MG::solve's argument */
UBG_API int ubg_batch_vcycles(ubg_batch *b, int ncycles, int zero_gradient_bc);
/* A batch whose domain i is simulation i's pressure solve (size, flag, h;
 * the simulations must share their multigrid parameters; each appears once)
 * ... */
UBG_API int ubg_batch_create_for_sims(ubg_sim *const *sims, int n, ubg_batch **out);
/* ... and one Simulation::step (simulation.cpp:356-374) of each of the n
 * simulations with timestep dts[i], every stage batched: each per-cell stage
 * (two diffuse iterations, advect, setVBCs + divergence, the correction tail)
 * is one launch over all n on the batch's stream, and so is each stage of
 * their V-cycles.  A simulation's own pending work (accumulated velocity,
 * boundary refresh) runs first on its own stream, and its next own work
 * waits for the batched step.  Equal, bit for bit, to ubg_sim_step on each
 * simulation.  A simulation whose flags changed since the last call (or
 * another simulation at slot i) gets its domain rebuilt; a simulation listed
 * twice is UBG_E_INVALID. */
UBG_API int ubg_sims_step(ubg_batch *b, ubg_sim *const *sims, int n, const float *dts);

/* ---- Served solves ----------------------------------------------------------
 * No counterpart in the reference: a request loop on its own host thread.
 * A request is `timesteps` consecutive refinement solves on one registered
 * geometry (flag + h), each to tol relative to ||F f|| of its rhs, the next
 * warm-started from the previous solution; its fields stay resident on the
 * device until the last one is done.  Requests are admitted in arrival order
 * while the device arena (hbm_budget bytes) and max_active allow, and every
 * refinement step advances all admitted requests together (continuous
 * batching: a request arriving while others run joins at the step after
 * next; the worker admits and retires while the current step runs on the
 * device).  A request's results do not depend on what else is in flight
 * (bit-identical to running it alone). */
typedef struct ubg_server ubg_server;
typedef struct ubg_server_params {
  size_t hbm_budget;    /* device bytes for geometries + resident requests  */
  int max_active;       /* requests on the device at once                  */
  double tol;           /* per timestep, relative to ||F f||, 1e-6          */
  int vcycles_per_iter; /* FP32 V-cycles per refinement step, 4            */
  int max_iter;         /* refinement steps per timestep at most, 200      */
  int steps_per_turn;   /* refinement steps the worker enqueues per turn
                           before it syncs and retires finished timesteps
                           (finished domains skip the rest on the device), 1 */
} ubg_server_params;
typedef struct ubg_request_info {
  uint64_t id;
  int state;            /* 0 queued, 1 running, 2 done, 3 failed           */
  int timesteps_done;
  int iterations;       /* refinement steps over all timesteps             */
  int converged;        /* every timestep met tol                          */
  double rel_residual;  /* of the last timestep                            */
  double t_submit, t_admit, t_done; /* seconds on the server's clock       */
} ubg_request_info;
typedef struct ubg_server_stats {
  long submitted, completed, failed, solves, steps;
  int active, queued, peak_active;
  size_t bytes_in_use, peak_bytes;
  double busy_s;        /* worker time in turns (sum of the four below)    */
  double setup_s;       /* member tables, schedule, launches of a step     */
  double wait_s;        /* waiting for the step in flight                  */
  double retire_s;      /* reading states, next rhs, results (while the
                           next step runs)                                 */
  double admit_s;       /* admitting requests (while the next step runs)   */
} ubg_server_stats;
UBG_API void ubg_server_params_default(ubg_server_params *out);
UBG_API int ubg_server_create(const ubg_server_params *prm, const ubg_params *mg,
                              ubg_server **out);
/* stops the worker after the requests in flight finish; queued ones fail */
UBG_API int ubg_server_destroy(ubg_server *srv);
/* register a geometry (flag W x H, 0/1, fluid outer ring; h) shared by
 * requests; resident until the server is destroyed */
UBG_API int ubg_server_add_geometry(ubg_server *srv, int width, int height, const float *flag,
                                    float h, int *geom_id);
/* rhs[t], t < timesteps: W x H float32 host arrays the caller keeps alive
 * until the request is done; fnorm: ||F f|| per timestep or NULL (computed
 * here); x_out: W x H float64 host array for the final solution or NULL */
UBG_API int ubg_server_submit(ubg_server *srv, int geom_id, int timesteps, const float *const *rhs,
                              const double *fnorm, double *x_out, uint64_t *id);
/* wait up to timeout_s (< 0: forever) for request id to finish; returns
 * UBG_E_STATE on timeout with info filled so far */
UBG_API int ubg_server_wait(ubg_server *srv, uint64_t id, double timeout_s, ubg_request_info *info);
UBG_API int ubg_server_drain(ubg_server *srv);
UBG_API int ubg_server_stats_get(ubg_server *srv, ubg_server_stats *st);
UBG_API double ubg_server_clock(ubg_server *srv);

/* ---- Simulation: simulation.hpp:18-141 ---------------------------------- */
/* Simulation(flag, pwidth, mu) (simulation.hpp:32-67): BCs S/N no-slip,
 * W inflow, E zero-pressure outflow; vx(0,y)=1, vy(x,0)=0 in both buffers;
 * MG(flag); h = pwidth/(W-1). */
UBG_API int ubg_sim_create(const float *flag, int width, int height, float pwidth,
                           float mu, ubg_sim **out);
/* Simulation(pwidth, mu, width, height) (simulation.hpp:20-30): every field
 * zero, the flag too, no inflow column; MG(width, height) keeps its coarse
 * flags at 1 until the flag is set (which runs updateFields).  The
 * reference leaves the boundary conditions uninitialised; here they start as
 * the other constructor sets them. */
UBG_API int ubg_sim_create_blank(int width, int height, float pwidth, float mu, ubg_sim **out);
UBG_API int ubg_sim_destroy(ubg_sim *sim);
UBG_API int ubg_sim_set_params(ubg_sim *sim, const ubg_params *prm);
UBG_API int ubg_sim_get_params(const ubg_sim *sim, ubg_params *out);
UBG_API int ubg_sim_dims(const ubg_sim *sim, int *width, int *height, float *h,
                         float *pwidth, float *mu);
UBG_API int ubg_sim_field_shape(const ubg_sim *sim, int field, int *width, int *height);
/* Host copies of any field.  Setting UBG_FIELD_FLAG re-runs the coarse
 * classification, as the game does with memcpy + mg.updateFields
 * (ubootgl_app.cpp:111-112). */
UBG_API int ubg_sim_set_field(ubg_sim *sim, int field, const float *src);
UBG_API int ubg_sim_get_field(ubg_sim *sim, int field, float *dst);
/* Device pointer of a field for zero-copy interop (valid until the next
 * step; vx/vy front and back swap roles during a step). */
UBG_API int ubg_sim_field_device_ptr(ubg_sim *sim, int field, void **dptr);
/* step(timestep) (simulation.cpp:356-374) */
UBG_API int ubg_sim_step(ubg_sim *sim, float dt);
/* nsteps x step(dt) (+ tracer advection when tracers are attached) */
UBG_API int ubg_sim_run(ubg_sim *sim, float dt, int nsteps);
UBG_API int ubg_sim_stage(ubg_sim *sim, int stage, float dt);
/* setGrids(c, val) (simulation.hpp:82-98) */
UBG_API int ubg_sim_set_grids(ubg_sim *sim, int x, int y, float val);
/* boundary conditions bcWest/bcEast/bcNorth/bcSouth (simulation.hpp:124),
 * values of Simulation::BC: 0 INFLOW, 1 OUTFLOW, 2 OUTFLOW_ZERO_PRESSURE,
 * 3 NOSLIP (simulation.hpp:69) */
UBG_API int ubg_sim_set_bcs(ubg_sim *sim, int west, int east, int north, int south);
UBG_API int ubg_sim_get_bcs(const ubg_sim *sim, int *west, int *east, int *north, int *south);
/* sinks (simulation.hpp:140, consumed by project simulation.cpp:173-187) */
UBG_API int ubg_sim_add_sink(ubg_sim *sim, float x, float y, float z);
UBG_API int ubg_sim_set_sinks(ubg_sim *sim, const float *xyz, int n);
UBG_API int ubg_sim_get_sinks(const ubg_sim *sim, float *xyz, int cap, int *count);
/* the simulation's MG (Simulation::mg, simulation.hpp:133); borrowed */
UBG_API int ubg_sim_get_mg(ubg_sim *sim, ubg_mg **mg);
/* bilinearVel(c) (advect_floating_items.cpp:11-14) for n grid-space points
 * c = (x0,y0,x1,y1,...); host arrays. */
UBG_API int ubg_sim_bilinear_vel(ubg_sim *sim, const float *c, float *out, int n);
/* Flag samplers at n physical positions pc = (x0,y0,x1,y1,...), host arrays:
 *   UBG_FLAG_NEAREST psampleFlagNearest (simulation.cpp:422-425): the flag of
 *                    the cell at the truncated position pc / (pwidth / W);
 *                    positions whose cell is outside the grid give 0 (the
 *                    reference indexes out of bounds there);
 *   UBG_FLAG_LINEAR  psampleFlagLinear (simulation.cpp:398-411);
 *   UBG_FLAG_NORMAL  psampleFlagNormal (simulation.cpp:413-420), two values
 *                    (x, y) per position in out.
 * The game calls psampleFlagLinear at ubootgl_app.cpp:144,166. */
enum { UBG_FLAG_NEAREST = 0, UBG_FLAG_LINEAR = 1, UBG_FLAG_NORMAL = 2 };
UBG_API int ubg_sim_sample_flag(ubg_sim *sim, int mode, const float *pc, float *out, int n);
/* getDT() (simulation.cpp:226-238): pwidth / (W-1) / max_vel * 2.5 with
 * max_vel the largest of 1e-7 and vx(x,y), vy(x,y) over x in [1, W-2],
 * y in [1, H-2] (signed values, as the reference takes them). */
UBG_API int ubg_sim_get_dt(ubg_sim *sim, float *dt);
/* centerP() (simulation.cpp:210-224): p -= (sum of interior p) / W / H on
 * interior cells.  The sum is taken in float64 in a fixed order (the
 * reference's float OpenMP reduction has no fixed order) and rounded to
 * float before the two divisions. */
UBG_API int ubg_sim_center_p(ubg_sim *sim);
/* Floating items (components.hpp:6-42: CoItem + CoKinematics /
 * CoKinematicsSimple), physical coordinates. */
typedef struct ubg_item {
  float size[2];
  float pos[2];
  float rotation;
  float mass;
  float vel[2];
  float force[2];
  float ang_vel;
  float ang_force;
  int bump_count;
} ubg_item;
/* Simulation::advectFloatingItems (simple = 0) or advectFloatingItemsSimple
 * (simple = 1) (advect_floating_items.cpp:16-274) for n items in the order
 * the reference's entity view visits them: the items are integrated on the
 * device (one thread each, items updated in place) and their reaction is
 * scattered into vx_accum / vy_accum (float atomics: the sums' order differs
 * from the reference's sequential loop), to be applied by the next step. */
UBG_API int ubg_sim_advect_items(ubg_sim *sim, ubg_item *items, int n, float game_dt, int simple);
/* The simulation part of UbootGlApp::shiftMap (ubootgl_app.cpp:248-296),
 * called after Terrain::shiftMap: vx (both buffers, from the front one), vy
 * and p shift one cell left, setGrids(c, new_flag(c)) for every cell
 * (simulation.hpp:82-98), the inlet velocity 0.07 H / (1 + sum flag(0, y))
 * on vx(0, y), saveCurrentVelocityFields, mg.updateFields.  W <= 16384. */
UBG_API int ubg_sim_shift_map(ubg_sim *sim, const float *new_flag);
/* Terrain flags from a decoded RGBA8 image (lodepng's layout, row 0 at the
 * top): flag_full (W x H, may be NULL) = red / 255 with rows flipped
 * (terrain.cpp:34-39); flag_sim ((W/scale) x (H/scale)) = Terrain::subSample,
 * block mean > 0.99 (terrain.cpp:3-11).  Host arrays. */
UBG_API int ubg_terrain_flags(const uint8_t *rgba, int width, int height, int scale,
                              float *flag_full, float *flag_sim);
/* Collocated velocity for rendering: VelocityTextures::updateFromStaggered
 * (velocity_textures.cpp:63-93) + interp_shader.cs:14-33 on vx_current /
 * vy_current (ubootgl_app.cpp:114).  vxy: (2W-1) x (2H-1) interleaved
 * (vx, vy) pairs (the RG32F image), mag: their lengths (the R32F image);
 * linear texel filtering with GL_REPEAT as the default sampler does.
 * on_device != 0: vxy / mag are device addresses, written asynchronously on
 * the simulation's stream; else host arrays (synchronous). */
UBG_API int ubg_sim_collocate(ubg_sim *sim, float *vxy, float *mag, int on_device);
/* Tracers: n points in physical coordinates (x,y) and ages, resident on the
 * device.  Each ubg_sim_run step advects them with RK2 midpoint integration
 * of bilinearVel, the freeze/age rule and Wang-hash respawn of
 * advect_tracer_points.cs:42-82. */
UBG_API int ubg_sim_set_tracers(ubg_sim *sim, const float *xy, const float *age, int n);
UBG_API int ubg_sim_get_tracers(ubg_sim *sim, float *xy, float *age, int n);
UBG_API int ubg_sim_advect_tracers(ubg_sim *sim, float dt, uint32_t seed);
UBG_API int ubg_sim_sync(ubg_sim *sim);

/* ---- measurement ---------------------------------------------------------- */
/* Kernel timing with HIP events attached to every dispatch of one kernel
 * class while enabled: level-0 stand-alone smoother passes, the level-0
 * fused pre-smoothing + residual + restriction pass, the level-0 fused
 * prolongation + correction + post-smoothing pass, and advect. */
enum { UBG_TIMER_SMOOTH_L0 = 0, UBG_TIMER_L0_PRE, UBG_TIMER_L0_POST, UBG_TIMER_ADVECT,
       UBG_TIMER_COUNT };
UBG_API int ubg_sim_timer_enable(ubg_sim *sim, int timer, int on);
UBG_API int ubg_sim_timer_read(ubg_sim *sim, int timer, double *total_ms, long *launches);

#ifdef __cplusplus
}
#endif
#endif /* UBG_C_H */

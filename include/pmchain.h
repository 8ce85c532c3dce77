/*
 * pmchain.h — C ABI of the MI355X particle-mesh force chain (libpmchain.so).
 *
 * Every entry point takes caller-owned device pointers (plain float / int / double
 * pointers, e.g. from a torch tensor's data_ptr()), sizes, and the HIP stream the work is
 * enqueued on. Only the calls that return a host scalar synchronize that stream. A
 * non-zero pm_status is an error; pm_last_error() returns the calling thread's message.
 * Handles are not thread-safe.
 *
 * Grid convention (reference deposit.py:16-26): node (i,j,k) sits at (i,j,k)*dx, fields
 * are n*n*n float32 arrays in C order with k (z) fastest.
 *
 * Citations name the reference file:line each entry point replaces.
 */
#ifndef PMCHAIN_H
#define PMCHAIN_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  PM_OK = 0,
  PM_ERR_INVALID = 1,  /* bad argument (null pointer, size out of range, ...) */
  PM_ERR_HIP = 2,      /* a HIP runtime call or kernel launch failed */
  PM_ERR_NOMEM = 3,    /* device allocation failed */
  PM_ERR_CAPACITY = 4  /* input larger than the plan was created for */
} pm_status;

/* Message of the last error raised on the calling thread ("" if none). */
const char* pm_last_error(void);
/* ABI version: bumped on any signature change. */
int32_t pm_abi_version(void);

/* ------------------------------------------------------------------------------------
 * CIC deposit (reference deposit.py:47-82 `_deposit_cic`, wrapper deposit.py:177-182
 * `deposit_density(parts, rho, dx)`).
 *
 * rho[node] = sum over particles of mass*(1/dx^3)*wx*wy*wz, out-of-box corners dropped,
 * computed as: integer cell key per particle -> stable LSD radix sort -> segmented
 * reduction of the 8 corner sums per occupied cell -> node assembly. The result is
 * deterministic (no float atomics): the summation order is fixed by the sorted order.
 * ---------------------------------------------------------------------------------- */
typedef struct pm_deposit_plan pm_deposit_plan;

/* Scratch for grids of `grid_size` nodes per axis and up to `max_particles` particles. */
pm_status pm_deposit_plan_create(int32_t grid_size, int64_t max_particles,
                                 pm_deposit_plan** out);
pm_status pm_deposit_plan_destroy(pm_deposit_plan* plan);

/* Deposit modes. PM_DEPOSIT_SORTED (default): cell-key sort + ordered segmented
 * reduction; also produces the keys and the stable permutation. PM_DEPOSIT_FIXED: every
 * corner contribution (same float32 arithmetic) is rounded to a 64-bit fixed-point integer
 * (scale 2^s, s = 61 - a - b for the largest m' = mass/dx^3 in [2^a, 2^(a+1)) and N < 2^b
 * particles) and summed per node with integer atomics after a per-wave aggregation of
 * particles that share a cell: deterministic in any order, rho = float(double(sum)*2^-s);
 * keys_out is available, perm_out is not. Non-finite contributions are dropped. The first
 * switch to PM_DEPOSIT_FIXED allocates n^3 64-bit accumulators and synchronizes. Measured at
 * 256^3 with 10M cell-ordered particles: 5.2 ms (the 64-bit device-scope atomics, ~0.4 G/s
 * across the XCDs) vs about 0.35 ms for the PM_DEPOSIT_SORTED deposit.
 * PM_DEPOSIT_BRICKED: the PM_DEPOSIT_FIXED result, bit for bit (same terms, same scale,
 * exact integer sums), computed without device-scope atomics: the particles are binned by
 * 8^3-cell bricks (a stable radix sort of brick keys), each unit of at most 8192 binned
 * particles sums its brick's region (the brick's nodes plus one cell of slack) in LDS, and a
 * combine pass adds the units' region sums per node. pm_deposit_cic bins on every call;
 * pm_deposit_cic_step reuses the last binning, which stays exact however far the particles
 * have moved since (a particle outside its unit's region takes a global-atomic path). The
 * first switch to it allocates the binning tables, about 10.6 KB per possible unit
 * (max_particles/8192 + the number of cell bricks) and n^3 64-bit sums, and synchronizes.
 * keys_out and perm_out are not available in this mode. */
enum { PM_DEPOSIT_SORTED = 0, PM_DEPOSIT_FIXED = 1, PM_DEPOSIT_BRICKED = 2 };
pm_status pm_deposit_plan_set_mode(pm_deposit_plan* plan, int32_t mode);

/* keys_out (optional, n_particles uint32): extended cell key per particle in input order,
 *   key = ((ix+1)*(n+1) + (iy+1))*(n+1) + (iz+1) with i* = floor(pos/dx), or (n+1)^3 when
 *   no corner of the particle's cell lies in the box.
 * perm_out (optional, n_particles int32): particle indices in stable key order. */
pm_status pm_deposit_cic(pm_deposit_plan* plan, const float* pos_x, const float* pos_y,
                         const float* pos_z, const float* mass, int64_t n_particles, float dx,
                         float* rho, uint32_t* keys_out, int32_t* perm_out, hipStream_t stream);

/* pm_deposit_cic for particles whose cell keys the caller already holds (`keys`, as
 * pm_deposit_cic's keys_out gives them for these positions, e.g. written by
 * pm_gather_kick_drift4 when it moved them): the key pass is skipped and `keys` is only
 * read. Same result, bit for bit. Replaces the same deposit call as pm_deposit_cic
 * (reference deposit.py:47-82, sim.py:318-325); no counterpart of its own. */
pm_status pm_deposit_cic_keyed(pm_deposit_plan* plan, const float* pos_x, const float* pos_y,
                               const float* pos_z, const float* mass, int64_t n_particles,
                               float dx, float* rho, const uint32_t* keys, int32_t* perm_out,
                               hipStream_t stream);

/* pm_deposit_cic without a plan and without allocating: no counterpart in the reference. It
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The sort and reduction scratch is a
 * static device array of PM_DEPOSIT_NOALLOC_SCRATCH_BYTES (96 MiB) that each device the
 * library runs on holds until the process ends; one call needs about 55.3 bytes per
 * particle plus 4 bytes per cell of the (n+1)^3 extended grid (at most 1,799,896 particles
 * on 64^3, 1,664,192 on 128^3, 589,312 on 256^3) and returns PM_ERR_CAPACITY, before any device
 * work, when that exceeds the array. The result (rho, keys_out, perm_out) is bit for bit
 * pm_deposit_cic's in sorted mode. Calls no hipMalloc/hipFree and records no timers. Every
 * call on one device shares that device's array and finds the last call's contents there
 * (the call resets what it needs itself): calls that overlap on one device (other streams
 * or threads) race and must be serialized by the caller. */
#define PM_DEPOSIT_NOALLOC_SCRATCH_BYTES (96ll << 20)
pm_status pm_deposit_cic_noalloc(int32_t grid_size, const float* pos_x, const float* pos_y,
                                 const float* pos_z, const float* mass, int64_t n_particles,
                                 float dx, float* rho, uint32_t* keys_out, int32_t* perm_out,
                                 hipStream_t stream);

/* The deposit of one step of a particle run (PM_DEPOSIT_BRICKED plans only): rho as
 * pm_deposit_cic gives it, bit for bit, reusing the plan's last binning unless
 * flags & PM_STEP_BIN (or the plan has no binning for n_particles particles at this dx),
 * which bins again first; then pm_deposit_plan_order gives the new binning's particle order.
 * The masses must be those of the last binning (their maximum sets the fixed-point scale).
 * flags & PM_STEP_RHO_KEPT: rho is the array of this plan's last deposit, unchanged since;
 * the node bricks far from every particle then keep their zeros instead of being stored
 * again (rho is still written whole after a new binning, and around particles that were
 * outside their unit's region in this or the last deposit).
 * Replaces the per-step deposit call (reference deposit.py:47-82 via sim.py:318-325). */
enum { PM_STEP_BIN = 1, PM_STEP_RHO_KEPT = 2 };
pm_status pm_deposit_cic_step(pm_deposit_plan* plan, const float* pos_x, const float* pos_y,
                              const float* pos_z, const float* mass, int64_t n_particles,
                              float dx, float* rho, int32_t flags, hipStream_t stream);

/* PM_DEPOSIT_BRICKED: the caller has moved the particles into the order of the plan's last
 * binning (pm_deposit_plan_order, e.g. by pm_gather_kick_drift4_reorder), so slot i now
 * holds binned position i: later pm_deposit_cic_step calls read the particles in place.
 * Part of the same per-step deposit (reference deposit.py:47-82); no counterpart of its
 * own. */
pm_status pm_deposit_plan_mark_ordered(pm_deposit_plan* plan);

/* Timing hook: per-phase device times (ms) of the last pm_deposit_cic on this plan when
 * pm_deposit_plan_set_timing(plan, 1) was set before it; phases = key, sort, reduce, assemble. */
pm_status pm_deposit_plan_set_timing(pm_deposit_plan* plan, int32_t enable);

/* The node bricks of pm_deposit_plan_node_bricks that are marked, compacted: list[0..*count)
 * (device memory; *count is a device value, at most *max_count = the number of bricks), in
 * no particular order. Owned by the plan, valid until its next deposit; NULL after a
 * fixed-mode deposit. In PM_DEPOSIT_BRICKED mode: the node bricks within one brick of a
 * binned brick (every node brick when a particle was outside its unit's region), which
 * hold every CIC corner node of the deposited particles. */
pm_status pm_deposit_plan_brick_list(pm_deposit_plan* plan, const int32_t** list,
                                     const uint32_t** count, int64_t* max_count);

/* Particle order of the last pm_deposit_cic / pm_deposit_cic_keyed with this plan: order[i]
 * is the particle at sorted position i (the perm_out array), or NULL after a fixed-mode or
 * empty deposit. Device memory owned by the plan, valid until its next deposit. In
 * PM_DEPOSIT_BRICKED mode: the binning's order (particles stably sorted by cell brick) when
 * the last deposit binned, else NULL. */
pm_status pm_deposit_plan_order(pm_deposit_plan* plan, const int32_t** order);

/* Node bricks of the last deposit with this plan: bricks[b] != 0 when some deposited
 * particle has a CIC corner node in the brick_edge^3 node brick b = (bi*nb + bj)*nb + bk,
 * nb = ceil(grid_size / brick_edge). Device memory owned by the plan, valid until its next
 * deposit; *bricks is NULL after a PM_DEPOSIT_FIXED deposit. For pm_grad4_bricks. No
 * reference counterpart (the reference's gradient covers every node, deposit.py:204-207). */
pm_status pm_deposit_plan_node_bricks(pm_deposit_plan* plan, const uint8_t** bricks,
                                      int32_t* brick_edge);
pm_status pm_deposit_last_timing(pm_deposit_plan* plan, float out_ms[4]);

/* ------------------------------------------------------------------------------------
 * Gradient, gather, integrator (reference deposit.py:132-174, 85-129; integrator.py:27-60,
 * 63-97).
 * ---------------------------------------------------------------------------------- */

/* g = -grad(phi): central differences inside, one-sided on the six faces
 * (reference deposit.py:132-174 `_potential_to_accel`, wrapper deposit.py:204-207). */
pm_status pm_grad(const float* phi, float* gx, float* gy, float* gz, int32_t grid_size,
                  float dx, hipStream_t stream);

/* acc = sum over in-box corners of w * g[node] (reference deposit.py:85-129 `_gather_accel`,
 * wrapper deposit.py:185-201 `gather_acceleration`). */
pm_status pm_gather(const float* pos_x, const float* pos_y, const float* pos_z,
                    const float* gx, const float* gy, const float* gz, float* acc_x,
                    float* acc_y, float* acc_z, int64_t n_particles, int32_t grid_size,
                    float dx, hipStream_t stream);

/* v += a*dt (reference integrator.py:27-42 `kick`). */
pm_status pm_kick(float* vel_x, float* vel_y, float* vel_z, const float* acc_x,
                  const float* acc_y, const float* acc_z, int64_t n_particles, float dt,
                  hipStream_t stream);

/* x += v*dt (reference integrator.py:45-60 `drift`). */
pm_status pm_drift(float* pos_x, float* pos_y, float* pos_z, const float* vel_x,
                   const float* vel_y, const float* vel_z, int64_t n_particles, float dt,
                   hipStream_t stream);

/* Fused kick(h) then drift(dt): bitwise the same as pm_kick followed by pm_drift
 * (the opening half of reference integrator.py:100-131 `kdk_step`). */
pm_status pm_kick_drift(float* pos_x, float* pos_y, float* pos_z, float* vel_x, float* vel_y,
                        float* vel_z, const float* acc_x, const float* acc_y,
                        const float* acc_z, int64_t n_particles, float h, float dt,
                        hipStream_t stream);

/* The fused step's interleaved variants (no reference counterpart as separate calls):
 * pm_grad4 writes g4[node] = (gx, gy, gz, 0) (float32, 16-byte aligned n^3*4 floats), and
 * pm_gather_kick4 reads it; the results are bitwise those of pm_grad + pm_gather_kick. */
pm_status pm_grad4(const float* phi, float* g4, int32_t grid_size, float dx, hipStream_t stream);
/* pm_grad4 on the nodes of the bricks marked in `bricks` (pm_deposit_plan_node_bricks):
 * every node the gather of those deposited particles reads; g4 elsewhere is left as it was.
 * Bitwise pm_grad4 on those nodes. */
pm_status pm_grad4_bricks(const float* phi, float* g4, int32_t grid_size, float dx,
                          const uint8_t* bricks, hipStream_t stream);

/* pm_grad4_bricks on the listed bricks only (pm_deposit_plan_brick_list: list, its device
 * count and max_count, the largest count it can hold): the same values at the same nodes,
 * without a workgroup per unmarked brick. No counterpart in the reference. */
pm_status pm_grad4_listed(const float* phi, float* g4, int32_t grid_size, float dx,
                          const int32_t* list, const uint32_t* count, int64_t max_count,
                          hipStream_t stream);
pm_status pm_gather_kick4(const float* pos_x, const float* pos_y, const float* pos_z,
                          const float* g4, float* acc_x, float* acc_y, float* acc_z,
                          float* vel_x, float* vel_y, float* vel_z, int64_t n_particles,
                          int32_t grid_size, float dx, float h, hipStream_t stream);

/* Between two consecutive steps: gather from g4, closing kick(h_close) of step t, opening
 * kick(h_open) and drift(dt) of step t+1, in one pass. Bitwise the same as pm_gather_kick4
 * followed by pm_kick_drift (reference integrator.py:100-131 `kdk_step`, its closing kick
 * followed by the next call's opening kick and drift); acc is not written, so the caller
 * ends a run of steps with pm_gather_kick4. keys_out (optional, n_particles uint32): the
 * drifted particles' cell keys on this grid, as pm_deposit_cic's keys_out would give them,
 * for the next step's pm_deposit_cic_keyed. */
pm_status pm_gather_kick_drift4(float* pos_x, float* pos_y, float* pos_z, const float* g4,
                                float* vel_x, float* vel_y, float* vel_z, int64_t n_particles,
                                int32_t grid_size, float dx, float h_close, float h_open,
                                float dt, uint32_t* keys_out, hipStream_t stream);

/* The per-particle arrays of one particle set (structure of arrays, n_particles each). pid
 * (optional) is carried along untouched, e.g. a particle's index in the initial conditions. */
typedef struct {
  float* pos_x;
  float* pos_y;
  float* pos_z;
  float* vel_x;
  float* vel_y;
  float* vel_z;
  float* mass;
  int32_t* gid;
  int32_t* pid;
} pm_particle_arrays;

/* pm_gather_kick_drift4 (keys_out required) that also moves particle order[i] of `src` to
 * slot i of `dst`, with every array (out of place: no destination may overlap a source,
 * keys_out or order; pid must be given in both or neither). With order = the last deposit's
 * particle order (pm_deposit_plan_order), particles of one cell become neighbours in memory
 * again, which keeps the next deposit's and gather's loads local over long runs. Each
 * particle's new position, velocity and key are bitwise those of pm_gather_kick_drift4;
 * entries of order outside [0, n_particles) keep the slot's own particle. Replaces the same
 * reference calls as pm_gather_kick_drift4 (integrator.py:100-131); the reordering itself
 * has no counterpart (the reference keeps its particles in load order). */
pm_status pm_gather_kick_drift4_reorder(const int32_t* order, const pm_particle_arrays* src,
                                        const pm_particle_arrays* dst, const float* g4,
                                        int64_t n_particles, int32_t grid_size, float dx,
                                        float h_close, float h_open, float dt,
                                        uint32_t* keys_out, hipStream_t stream);

/* Fused gather then kick(h): bitwise the same as pm_gather followed by pm_kick
 * (the closing half of reference integrator.py:100-131 `kdk_step`). */
pm_status pm_gather_kick(const float* pos_x, const float* pos_y, const float* pos_z,
                         const float* gx, const float* gy, const float* gz, float* acc_x,
                         float* acc_y, float* acc_z, float* vel_x, float* vel_y, float* vel_z,
                         int64_t n_particles, int32_t grid_size, float dx, float h,
                         hipStream_t stream);

/* Direct O(N^2) Plummer-softened acceleration (reference integrator.py:63-97
 * `direct_accel`); a small-N validation path, not part of the PM step. */
pm_status pm_direct_accel(const float* pos_x, const float* pos_y, const float* pos_z,
                          const float* mass, float* acc_x, float* acc_y, float* acc_z,
                          int64_t n_particles, float grav, float eps2, hipStream_t stream);

/* ------------------------------------------------------------------------------------
 * Open-boundary multigrid (reference solver/multigrid.py:46-494).
 * ---------------------------------------------------------------------------------- */
typedef struct pm_mg pm_mg;

enum { PM_MG_PLAIN = 0, PM_MG_MIXED = 1 };

typedef struct {
  int32_t grid_size;     /* finest nodes per axis */
  float dx;              /* finest node spacing */
  double grav;           /* G */
  int32_t n_cycles;      /* max V-cycles per solve (reference default 20) */
  int32_t pre_sweeps;    /* nu1 (2) */
  int32_t post_sweeps;   /* nu2 (2) */
  int32_t coarse_sweeps; /* RB sweeps on the coarsest level in PLAIN mode (40) */
  int32_t min_cycles;    /* (1) */
  double tol;            /* stop when ||r|| < tol*||rhs|| (1e-3) */
  int32_t mode;          /* PM_MG_PLAIN or PM_MG_MIXED */
} pm_mg_desc;

typedef struct {
  int32_t cycles;        /* V-cycles run by the last solve (reference `last_cycles`) */
  int32_t levels;
  double resid_norm;     /* PLAIN: final residual L2 norm; MIXED: norm of the residual the
                            last correction was computed from (float64) */
  double rhs_norm;       /* ||4*pi*G*rho|| (float64) */
  double err_est;        /* MIXED: estimated relative L2 error of phi vs the exact discrete
                            solution when the solve stopped; PLAIN: -1 */
} pm_mg_stats;

/* Builds the level hierarchy. PLAIN (reference multigrid.py:301-360): sizes halve while
 * even and the coarser size is >= 3, coarse node J at fine node 2J. MIXED: level 1 has
 * n/2+1 nodes (n even) or (n-1)/2+1 (n odd) and every further level (m-1)/2+1, down to 3
 * nodes, so every level below the finest spans the whole box; coarse J is still fine 2J.
 * MIXED stops when the estimated relative phi error (from successive correction norms,
 * rho = ||e_k||/||e_{k-1}||, err = ||e_k|| rho/(1-rho)) drops below `tol`. */
pm_status pm_mg_create(const pm_mg_desc* desc, pm_mg** out);
pm_status pm_mg_destroy(pm_mg* mg);
pm_status pm_mg_num_levels(pm_mg* mg, int32_t* out);

/* Solve lap(phi) = 4*pi*G*rho with multipole Dirichlet faces (reference
 * multigrid.py:455-487 `solve`). phi is read as the initial guess when warm_start != 0 and
 * overwritten with the result. Synchronizes `stream` once per V-cycle (convergence test). */
pm_status pm_mg_solve(pm_mg* mg, const float* rho, float* phi, int32_t warm_start,
                      pm_mg_stats* stats, hipStream_t stream);

/* ------------------------------------------------------------------------------------
 * Warm starts across time steps (no reference counterpart: the reference warm-starts from
 * the previous phi, multigrid.py:465-466, which is PM_WARM_PREVIOUS here).
 *
 * A sequence of pm_mg_solve_warm calls on one handle, one per time step with phi holding
 * the previous step's solution, keeps the previous phi in the handle. PM_WARM_EXTRAPOLATE
 * then starts from 2*phi(t-1) - phi(t-2) on the interior; PM_WARM_LEARNED adds the
 * warm-start network's prediction of that guess's error (network input: the residual of
 * the extrapolated guess, phi(t-1) - phi(t-2) and 4*pi*G*rho, each pooled by pool^3 and
 * divided by its RMS; output: float32 at grid_size/pool, scaled by RMS(pooled residual) *
 * (pool*dx)^2 and interpolated trilinearly onto the fine interior). The multigrid judges
 * every inference: the correction is kept only if the norm of the pooled residual (the mean
 * residual over each network voxel, float64; the smooth part the correction addresses) after
 * it is below fallback_ratio times the norm before it, otherwise the solve proceeds from the
 * extrapolated guess alone. Either way the solve runs to the handle's tolerance, so the
 * result meets the same accuracy. The first call of a sequence (or after pm_mg_solve or
 * pm_mg_warm_reset) has no history and starts from phi like PM_WARM_PREVIOUS.
 * pair_input / pair_target, when set, receive the network input and the training target
 * (the pooled error of the extrapolated guess against the converged phi, divided by the
 * output scale) of this solve; they are written only when `extrapolated` comes back 1.
 * Buffers for the history (2 grids) and the network grid are allocated on first use.
 * ---------------------------------------------------------------------------------- */
typedef struct pm_wsnet pm_wsnet;

enum { PM_WARM_PREVIOUS = 0, PM_WARM_EXTRAPOLATE = 1, PM_WARM_LEARNED = 2 };

typedef struct {
  int32_t mode;            /* PM_WARM_* */
  int32_t pool;            /* fine nodes per network voxel per axis; grid_size/pool % 8 == 0 */
  pm_wsnet* net;           /* PM_WARM_LEARNED: network of resolution grid_size/pool */
  double fallback_ratio;   /* keep the correction iff |pooled r| after < ratio * before (1.0) */
  uint16_t* pair_input;    /* optional out: bf16 [R^3][8] network input (R = grid_size/pool) */
  float* pair_target;      /* optional out: float32 [R^3] training target */
  int32_t extrapolated;    /* out: 1 if the guess was extrapolated in time */
  int32_t used_net;        /* out: 1 if the learned correction was kept */
  double resid_before;     /* out (LEARNED): pooled residual norm of the extrapolated guess */
  double resid_after;      /* out (LEARNED): the same after the learned correction */
} pm_mg_warm;

/* Warm-started solve (see above). Synchronizes like pm_mg_solve, plus once before the
 * V-cycles when the network grid is used. */
pm_status pm_mg_solve_warm(pm_mg* mg, const float* rho, float* phi, pm_mg_warm* warm,
                           pm_mg_stats* stats, hipStream_t stream);
/* Forget the time history (the next warm solve starts from phi alone). */
pm_status pm_mg_warm_reset(pm_mg* mg);

/* Execution options of a solver handle (no reference counterpart: the reference's solver
 * is a chain of Taichi kernel calls, multigrid.py:429-440). Results are bitwise the same
 * with every option on or off.
 *   PM_MG_OPT_GRAPHS (default 1): solves replay each V-cycle (with the MIXED correction's
 *     fill, update and norms) from a HIP graph captured on first use for the buffers
 *     involved, launched on the solve's stream; eager launches while per-kernel timing is
 *     enabled.
 *   PM_MG_OPT_SPLIT_FINE (default 1): MIXED solves keep the finest level's correction and
 *     defect in colour-split layout, with the fill, the colour-0 prolongation and the
 *     update folded into the smoothing passes (needs pre/post sweeps >= 1 and a level below
 *     the finest; otherwise ignored).
 *   PM_MG_OPT_BOTTOM (default 1): the levels of size <= 17 at the bottom of the hierarchy
 *     run their part of each V-cycle in one workgroup with every level in LDS; 0 runs them
 *     with the per-level kernels.
 *   PM_MG_OPT_SMALL_SWEEP (default 1): levels of size <= 65 below the finest smooth with one
 *     launch per red-black sweep (out of place, each colour-1 node recomputing its
 *     colour-0 neighbours) when the cycle's pre + post sweeps are even; 0 uses two colour
 *     passes per sweep.
 * Setting an option drops the handle's captured cycles; they are captured again on use. */
#define PM_MG_OPT_GRAPHS 0
#define PM_MG_OPT_SPLIT_FINE 1
#define PM_MG_OPT_BOTTOM 2
#define PM_MG_OPT_SMALL_SWEEP 3
pm_status pm_mg_set_option(pm_mg* mg, int32_t option, int32_t value);

/* One more V-cycle on the state of the last solve (reference multigrid.py:429-440
 * `_vcycle(0)` as called by tests). */
pm_status pm_mg_vcycle(pm_mg* mg, hipStream_t stream);

/* L2 norm (float64) of the finest residual of the last solve's phi (reference
 * multigrid.py:489-494 `residual_norm`). Synchronizes. */
pm_status pm_mg_residual_norm(pm_mg* mg, double* out, hipStream_t stream);

/* Ten float64 raw moments [M, Smx, Smy, Smz, Smxx, Smyy, Smzz, Smxy, Smxz, Smyz] of rho
 * (reference multigrid.py:133-178 `_accumulate_moments`). Synchronizes. */
pm_status pm_mg_moments(pm_mg* mg, const float* rho, double out[10], hipStream_t stream);

/* ---- single-kernel hooks (reference multigrid.py kernels), on caller buffers ---- */
/* one red-black colour pass (multigrid.py:46-59) */
pm_status pm_rb_pass(float* phi, const float* rhs, int32_t n, float h2, int32_t color,
                     hipStream_t stream);
/* one full sweep, colour 0 then colour 1, in a single streaming pass from phi_in into
 * phi_out (every node written; distinct buffers); phi_out equals phi_in after two
 * pm_rb_pass calls, bit for bit (multigrid.py:422-427 `_smooth` with sweeps=1) */
pm_status pm_rb_sweep(const float* phi_in, float* phi_out, const float* rhs, int32_t n,
                      float h2, hipStream_t stream);
/* res = rhs - lap(phi) inside, 0 on the boundary (multigrid.py:62-79) */
pm_status pm_residual(const float* phi, const float* rhs, float* res, int32_t n,
                      float inv_h2, hipStream_t stream);
/* full-weighting restriction onto the coarse rhs, coarse J at fine 2J, 2*nc-2 <= nf <= 2*nc
 * (multigrid.py:82-97) */
pm_status pm_restrict(const float* res_f, float* rhs_c, int32_t nf, int32_t nc,
                      hipStream_t stream);
/* trilinear prolongation added into the fine interior, 2*nc-2 <= nf <= 2*nc
 * (multigrid.py:100-130) */
pm_status pm_prolong_add(const float* phi_c, float* phi_f, int32_t nf, int32_t nc,
                         hipStream_t stream);
/* multipole face potential; params = [G, M, cx, cy, cz, sxx, syy, szz, sxy, sxz, syz]
 * (multigrid.py:261-295) */
pm_status pm_set_faces(float* phi, int32_t n, float dx, const float params[11],
                       hipStream_t stream);

/* ------------------------------------------------------------------------------------
 * Float64 conservation reductions (reference diagnostics_device.py:41-87, 152-184).
 * ---------------------------------------------------------------------------------- */
/* out[11] = [M, KE, Px, Py, Pz, Lx, Ly, Lz, Smx, Smy, Smz] (diagnostics_device.py:41-74).
 * Synchronizes. */
pm_status pm_reduce_particles_f64(const float* pos_x, const float* pos_y, const float* pos_z,
                                  const float* vel_x, const float* vel_y, const float* vel_z,
                                  const float* mass, int64_t n_particles, double out[11],
                                  hipStream_t stream);
/* out = sum rho*phi (diagnostics_device.py:77-87). Synchronizes. */
pm_status pm_reduce_grid_pe_f64(const float* rho, const float* phi, int64_t n_nodes,
                                double* out, hipStream_t stream);
/* float64 mass histogram of radius about (cx,cy,cz); bin = floor(r*inv_binw) clamped to
 * [0, nbins-1] (diagnostics_device.py:152-184). hist_dev: nbins doubles on the device. */
pm_status pm_radial_histogram_f64(const float* pos_x, const float* pos_y, const float* pos_z,
                                  const float* mass, int64_t n_particles, float cx, float cy,
                                  float cz, float inv_binw, int32_t nbins, double* hist_dev,
                                  hipStream_t stream);

/* ------------------------------------------------------------------------------------
 * Device 2D projection (reference viz/project.py:26-83 `_project` / `scatter_density`).
 * img[ia*bins + ib] = mass in bin (ia, ib) (float32 [bins][bins], axis a major), ia =
 * floor((pos_a - lo_a)/binw_a), likewise ib; particles outside [0, bins) on either axis are
 * dropped. Deterministic: masses are summed as 64-bit fixed-point integers (absolute error
 * per particle <= 2^-35 of the largest mass). scratch: pm_project_scratch_bytes(bins) bytes of
 * device memory, 8-byte aligned. n <= 2^28.
 * ---------------------------------------------------------------------------------- */
pm_status pm_project_scratch_bytes(int32_t bins, int64_t* out);
pm_status pm_project_mass(const float* pos_a, const float* pos_b, const float* mass,
                          int64_t n_particles, float lo_a, float lo_b, float binw_a,
                          float binw_b, int32_t bins, float* img, void* scratch,
                          hipStream_t stream);

/* dst = src (count floats): a plain streaming copy of known size, used to calibrate
 * profiler byte counters (no reference counterpart). */
pm_status pm_copy_f32(float* dst, const float* src, int64_t count, hipStream_t stream);

/* ------------------------------------------------------------------------------------
 * BF16 3x3x3 convolutions on the matrix cores (no reference counterpart: they serve the
 * learned warm start).
 *
 * Tensors are channels-last bf16 bit patterns (uint16_t): element (voxel v, channel c) of
 * an r^3 tensor with C channels sits at v*C + c, v = (z*r + y)*r + x. Stored channel counts
 * are multiples of 8 and pointers 16-byte aligned. A convolution is a "same" (padding 1)
 * cross-correlation with torch.nn.functional.conv3d's weight layout w[cout][cin][3][3][3]
 * (float32 master copy), accumulated in float32 (v_mfma_f32_16x16x32_bf16). Its input is
 * xa (ca channels) followed by xb (cb channels, may be absent); with up_a != 0, xa has
 * resolution r/2 and is read nearest-neighbour upsampled.
 * ---------------------------------------------------------------------------------- */
/* Elements of the packed bf16 operand of w: transpose 0 for the forward pass, 1 for the
 * backward-data pass (weights flipped in space, channels transposed). */
pm_status pm_conv3d_packed_size(int32_t cout, int32_t cin, int32_t transpose, int64_t* elems);
pm_status pm_conv3d_pack_weights(const float* w, int32_t cout, int32_t cin, int32_t transpose,
                                 uint16_t* packed, hipStream_t stream);
/* Float32 scratch that lets the two calls below split their reduction over workgroups on
 * small grids (split-K, partial sums folded in a fixed order); without it (scratch null or
 * smaller) they run unsplit. Same result either way up to float32 summation order. */
pm_status pm_conv3d_scratch_size(int32_t cin, int32_t cout, int32_t r, int64_t* floats);
/* y = act(conv(x) + bias), act 1 = leaky ReLU with negative slope `slope`, 0 = identity.
 * Outputs (either may be null): y bf16 [r^3][roundup8(cout)] (pad channels 0), y32 float32
 * [r^3][cout]. bias and scratch may be null. */
pm_status pm_conv3d_fwd_bf16(const uint16_t* xa, int32_t ca, const uint16_t* xb, int32_t cb,
                             int32_t r, int32_t up_a, const uint16_t* packed, const float* bias,
                             int32_t cout, int32_t act, float slope, uint16_t* y, float* y32,
                             float* scratch, int64_t scratch_floats, hipStream_t stream);
/* dx float32 [r^3][cin] = gradient of the (un-upsampled, concatenated) convolution input
 * for the output gradient dz (bf16 [r^3][roundup8(cout)]); packed_t from transpose = 1. */
pm_status pm_conv3d_bwd_data_bf16(const uint16_t* dz, int32_t cout, int32_t r,
                                  const uint16_t* packed_t, int32_t cin, float* dx,
                                  float* scratch, int64_t scratch_floats, hipStream_t stream);
/* Floats of scratch pm_conv3d_bwd_weight_bf16 needs for these shapes. */
pm_status pm_conv3d_wgrad_partial_size(int32_t cin, int32_t cout, int32_t r, int64_t* floats);
/* dw float32 [cout][cin][27] and db float32 [cout] (may be null) of the convolution for the
 * input (xa, xb, up_a) and output gradient dz; accumulate != 0 adds to dw and db. The sum
 * over voxels is split in a fixed order (deterministic). roundup8(cout) must be a power of
 * two. */
pm_status pm_conv3d_bwd_weight_bf16(const uint16_t* xa, int32_t ca, const uint16_t* xb,
                                    int32_t cb, int32_t r, int32_t up_a, const uint16_t* dz,
                                    int32_t cout, float* dw, float* db, int32_t accumulate,
                                    float* partial, int64_t partial_floats, hipStream_t stream);
/* 2x2x2 average pooling, bf16 [r^3][c] -> [(r/2)^3][c] (float32 sum, fixed order). */
pm_status pm_pool2_bf16(const uint16_t* x, int32_t r, int32_t c, uint16_t* y, hipStream_t stream);
/* dz[v][c] = bf16((g1[v*g1_stride + c] + g2[parent(v)*gc + c]/8) * act'(y[v][c])) for
 * c < gc and 0 for gc <= c < dz_c; g1, g2 (at r/2: the gradient through a 2x2x2 average
 * pool) and y (null: no activation) may be null. act' = 1 where y > 0, else slope. */
pm_status pm_act_bwd_bf16(const float* g1, int32_t g1_stride, const float* g2, const uint16_t* y,
                          float slope, int32_t r, int32_t gc, int32_t dz_c, uint16_t* dz,
                          hipStream_t stream);
/* ga[vc][c] = sum over the 8 children of dx[child*cstride + c], c < ca (gradient of a
 * nearest-upsampled tensor), at resolution r/2. */
pm_status pm_upsum_f32(const float* dx, int32_t r, int32_t cstride, int32_t ca, float* ga,
                       hipStream_t stream);
/* y = bf16(x), round to nearest even. */
pm_status pm_f32_to_bf16(const float* x, int64_t count, uint16_t* y, hipStream_t stream);

/* ------------------------------------------------------------------------------------
 * The warm-start network (no reference counterpart): a 4-level 3D convolutional
 * encoder-decoder, 15 conv3x3x3 layers e0a e0b e1a e1b e2a e2b e3a e3b (encoder, 2x2x2
 * average pooling between levels) d2a d2b d1a d1b d0a d0b (decoder, nearest upsampling
 * concatenated with the encoder tensor of the same level) and a linear 1-channel output
 * layer; leaky ReLU elsewhere; bf16 activations, float32 accumulation and output.
 * Parameters: one float32 vector, per layer w [cout][cin][27] then b [cout].
 * ---------------------------------------------------------------------------------- */
typedef struct {
  int32_t res;           /* finest resolution (multiple of 8, <= 256) */
  int32_t in_channels;   /* channels the caller fills (<= 8; the stored input has 8) */
  int32_t width[4];      /* channels per level (16, 32, 64 or 128 each) */
  float slope;           /* leaky ReLU negative slope */
} pm_wsnet_desc;

typedef struct {
  char name[8];
  int32_t cin_a, cin_b, cout, res, upsample_a, act;
  int64_t w_offset, b_offset;  /* into the parameter vector */
} pm_wsnet_layer;

pm_status pm_wsnet_create(const pm_wsnet_desc* desc, pm_wsnet** out);
pm_status pm_wsnet_destroy(pm_wsnet* net);
pm_status pm_wsnet_num_params(const pm_wsnet* net, int64_t* out);
/* index 0..14 */
pm_status pm_wsnet_layer_info(const pm_wsnet* net, int32_t index, pm_wsnet_layer* out);
/* Copies the float32 parameter vector (device) and packs the bf16 operands. */
pm_status pm_wsnet_set_params(pm_wsnet* net, const float* params, hipStream_t stream);
/* out float32 [res^3] from input bf16 [res^3][8]; keeps the activations for backward. */
pm_status pm_wsnet_forward(pm_wsnet* net, const uint16_t* input, float* out, hipStream_t stream);
/* grads (float32, parameter layout) of sum(dout * out) for the last forward pass;
 * accumulate != 0 adds to grads. */
pm_status pm_wsnet_backward(pm_wsnet* net, const float* dout, float* grads, int32_t accumulate,
                            hipStream_t stream);

/* ------------------------------------------------------------------------------------
 * Kernel timing (no reference counterpart; the reference times stages with wall clocks
 * around ti.sync(), bench.py:66-72). When enabled, each library kernel launch records a
 * HIP event pair on its own stream, tagged with the kernel's algorithmic bytes.
 * ---------------------------------------------------------------------------------- */
typedef struct {
  char name[48];
  int64_t launches;
  double total_ms;   /* summed event time */
  double bytes;      /* summed algorithmic (compulsory) bytes of those launches; for the
                        net.*conv kernels the summed floating-point operations */
} pm_kernel_time;

/* on: 0 off; 1 every instrumented launch (multigrid V-cycles then run eagerly, not from
 * their graphs); 2 only the launches whose own dispatch records the pair (the deposit's radix
 * scatter, unit pass, brick reduce and combine), which adds nothing to the streams, so the
 * timed steps run exactly as untimed ones. */
pm_status pm_timing_enable(int32_t on);
/* Drops accumulated totals (waits for pending events). */
pm_status pm_timing_reset(void);
/* Writes up to `capacity` entries (sorted by name) and the total count. Synchronizes. */
pm_status pm_timing_report(int32_t capacity, pm_kernel_time* out, int32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* PMCHAIN_H */

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

/* keys_out (optional, n_particles uint32): extended cell key per particle in input order,
 *   key = ((ix+1)*(n+1) + (iy+1))*(n+1) + (iz+1) with i* = floor(pos/dx), or (n+1)^3 when
 *   no corner of the particle's cell lies in the box.
 * perm_out (optional, n_particles int32): particle indices in stable key order. */
pm_status pm_deposit_cic(pm_deposit_plan* plan, const float* pos_x, const float* pos_y,
                         const float* pos_z, const float* mass, int64_t n_particles, float dx,
                         float* rho, uint32_t* keys_out, int32_t* perm_out, hipStream_t stream);

/* Timing hook: per-phase device times (ms) of the last pm_deposit_cic on this plan when
 * pm_deposit_plan_set_timing(plan, 1) was set before it; phases = key, sort, reduce, assemble. */
pm_status pm_deposit_plan_set_timing(pm_deposit_plan* plan, int32_t enable);
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

/* dst = src (count floats): a plain streaming copy of known size, used to calibrate
 * profiler byte counters (no reference counterpart). */
pm_status pm_copy_f32(float* dst, const float* src, int64_t count, hipStream_t stream);

/* ------------------------------------------------------------------------------------
 * Kernel timing (no reference counterpart; the reference times stages with wall clocks
 * around ti.sync(), bench.py:66-72). When enabled, each library kernel launch records a
 * HIP event pair on its own stream, tagged with the kernel's algorithmic bytes.
 * ---------------------------------------------------------------------------------- */
typedef struct {
  char name[48];
  int64_t launches;
  double total_ms;   /* summed event time */
  double bytes;      /* summed algorithmic (compulsory) bytes of those launches */
} pm_kernel_time;

pm_status pm_timing_enable(int32_t on);
/* Drops accumulated totals (waits for pending events). */
pm_status pm_timing_reset(void);
/* Writes up to `capacity` entries (sorted by name) and the total count. Synchronizes. */
pm_status pm_timing_report(int32_t capacity, pm_kernel_time* out, int32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* PMCHAIN_H */

/*
 * pm_oracle.c — CPU restatement of the reference PM force chain. TEST INFRASTRUCTURE ONLY:
 * used by tests/, the smoke check and the CPU-baseline leg of bench.py as the checker;
 * never linked into or called by the product library.
 *
 * Two parts:
 *
 * 1. oracle_deposit_cic: the deterministic deposit order the GPU deposit implements,
 *    written from its definition (stable counting sort, sequential pieces inside aligned
 *    16-position chunks, recursive pairwise tree over a cell's pieces, fixed corner order
 *    per node). Per-particle arithmetic follows reference deposit.py:58-82.
 *    oracle_deposit_fixed: the fixed-point mode (integer sums of rounded corner terms).
 *
 * 2. cpu_*: an OpenMP port of the reference kernel sequence used as the CPU baseline:
 *    float-atomic CIC scatter (deposit.py:47-82), central-difference gradient
 *    (deposit.py:132-174), CIC gather (deposit.py:85-129), kick/drift
 *    (integrator.py:27-60), and the open-BC multigrid solve (solver/multigrid.py:46-487)
 *    with red-black Gauss-Seidel V(2,2) cycles, 40 coarsest sweeps, tol 1e-3, 20-cycle cap.
 *
 * Build (oracle/Makefile): gcc -O2 -fopenmp -ffp-contract=off — one rounding per
 * operation, IEEE float division and sqrt, as the reference's float32 kernels.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PIECE 16

static int cic_axis(float x, float dx, int n, int* i, float* f) {
  const float g = x / dx;
  if (!(g >= -1.0f && g < (float)n)) return 0;
  const float fl = floorf(g);
  *i = (int)fl;
  *f = g - (float)(*i);
  return 1;
}

int oracle_abi_version(void) { return 1; }

/* ---------------------------------------------------------------------------------------
 * 1. deterministic deposit
 * ------------------------------------------------------------------------------------- */

static void corner_terms(const float r[4], float t[8]) {
  const float wx0 = 1.0f - r[0], wy0 = 1.0f - r[1], wz0 = 1.0f - r[2];
  for (int c = 0; c < 8; ++c) {
    const float wx = (c & 4) ? r[0] : wx0;
    const float wy = (c & 2) ? r[1] : wy0;
    const float wz = (c & 1) ? r[2] : wz0;
    t[c] = r[3] * wx * wy * wz;
  }
}

/* Pairwise tree over pieces[0..k): split at the largest power of two below k. */
static void tree_sum(const float* pieces, int64_t k, float out[8]) {
  if (k == 1) {
    for (int c = 0; c < 8; ++c) out[c] = pieces[c];
    return;
  }
  int64_t m = 1;
  while (m * 2 < k) m *= 2;
  float a[8], b[8];
  tree_sum(pieces, m, a);
  tree_sum(pieces + 8 * m, k - m, b);
  for (int c = 0; c < 8; ++c) out[c] = a[c] + b[c];
}

/* rho: n^3 floats; keys (optional): N uint32; perm (optional): N int32. Returns 0 on
 * success, -1 on allocation failure. */
int oracle_deposit_cic(const float* px, const float* py, const float* pz, const float* mass,
                       int64_t np, int n, float dx, float* rho, uint32_t* keys_out,
                       int32_t* perm_out) {
  const int64_t m = (int64_t)n + 1;
  const uint32_t sentinel = (uint32_t)(m * m * m);
  const float inv_vol = 1.0f / ((dx * dx) * dx);
  uint32_t* key = (uint32_t*)malloc(sizeof(uint32_t) * (np > 0 ? np : 1));
  float* rec = (float*)malloc(sizeof(float) * 4 * (np > 0 ? np : 1));
  int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (np > 0 ? np : 1));
  int64_t* count = (int64_t*)calloc((size_t)sentinel + 2, sizeof(int64_t));
  int32_t* map = (int32_t*)malloc(sizeof(int32_t) * (size_t)sentinel);
  float* sums = (float*)malloc(sizeof(float) * 8 * (np > 0 ? np : 1));
  float* pieces = (float*)malloc(sizeof(float) * 8 * (np / PIECE + 2));
  if (!key || !rec || !perm || !count || !map || !sums || !pieces) {
    free(key); free(rec); free(perm); free(count); free(map); free(sums); free(pieces);
    return -1;
  }
  for (int64_t p = 0; p < np; ++p) {
    int ix = 0, iy = 0, iz = 0;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const int in = cic_axis(px[p], dx, n, &ix, &fx) && cic_axis(py[p], dx, n, &iy, &fy) &&
                   cic_axis(pz[p], dx, n, &iz, &fz);
    key[p] = in ? (uint32_t)(((int64_t)(ix + 1) * m + (iy + 1)) * m + (iz + 1)) : sentinel;
    rec[4 * p + 0] = fx;
    rec[4 * p + 1] = fy;
    rec[4 * p + 2] = fz;
    rec[4 * p + 3] = mass[p] * inv_vol;
    if (keys_out) keys_out[p] = key[p];
  }
  /* stable counting sort */
  for (int64_t p = 0; p < np; ++p) count[key[p] + 1]++;
  for (uint32_t k = 0; k <= sentinel; ++k) count[k + 1] += count[k];
  for (int64_t p = 0; p < np; ++p) perm[count[key[p]]++] = (int32_t)p;
  if (perm_out) memcpy(perm_out, perm, sizeof(int32_t) * np);
  for (uint32_t k = 0; k < sentinel; ++k) map[k] = -1;
  /* per-cell corner sums */
  int64_t a = 0;
  while (a < np) {
    const uint32_t k = key[perm[a]];
    if (k >= sentinel) break;
    int64_t b = a + 1;
    while (b < np && key[perm[b]] == k) ++b;
    int64_t npieces = 0;
    int64_t s = a;
    while (s < b) {
      int64_t e = (s / PIECE + 1) * PIECE;
      if (e > b) e = b;
      float acc[8], t[8];
      corner_terms(&rec[4 * (int64_t)perm[s]], acc);
      for (int64_t q = s + 1; q < e; ++q) {
        corner_terms(&rec[4 * (int64_t)perm[q]], t);
        for (int c = 0; c < 8; ++c) acc[c] = acc[c] + t[c];
      }
      for (int c = 0; c < 8; ++c) pieces[8 * npieces + c] = acc[c];
      ++npieces;
      s = e;
    }
    float out[8];
    tree_sum(pieces, npieces, out);
    for (int c = 0; c < 8; ++c) sums[8 * a + c] = out[c];
    map[k] = (int32_t)a;
    a = b;
  }
  /* node assembly */
  const int64_t n2 = (int64_t)n * n;
  for (int64_t idx = 0; idx < n2 * n; ++idx) {
    const int kk = (int)(idx % n), j = (int)((idx / n) % n), i = (int)(idx / n2);
    float acc = 0.0f;
    for (int c = 0; c < 8; ++c) {
      const int64_t ex = i + 1 - (c >> 2), ey = j + 1 - ((c >> 1) & 1), ez = kk + 1 - (c & 1);
      const int32_t slot = map[(ex * m + ey) * m + ez];
      if (slot >= 0) acc = acc + sums[8 * (int64_t)slot + c];
    }
    rho[idx] = acc;
  }
  free(key); free(rec); free(perm); free(count); free(map); free(sums); free(pieces);
  return 0;
}

/* Fixed-point deposit: the same float32 corner terms, each rounded to a 64-bit integer at
 * scale 2^s (s = 61 - a - b, largest finite positive m' in [2^a, 2^(a+1)), np < 2^b), summed
 * per node as integers (any order gives the same sum), rho = (float)((double)sum * 2^-s).
 * Non-finite or non-positive terms are dropped. keys (optional) as oracle_deposit_cic. */
int oracle_deposit_fixed(const float* px, const float* py, const float* pz, const float* mass,
                         int64_t np, int n, float dx, float* rho, uint32_t* keys_out) {
  const int64_t m = (int64_t)n + 1;
  const uint32_t sentinel = (uint32_t)(m * m * m);
  const float inv_vol = 1.0f / ((dx * dx) * dx);
  const int64_t nodes = (int64_t)n * n * n;
  uint64_t* acc = (uint64_t*)calloc((size_t)(nodes > 0 ? nodes : 1), sizeof(uint64_t));
  if (!acc) return -1;
  float mmax = 0.0f;
  for (int64_t p = 0; p < np; ++p) {
    const float v = mass[p] * inv_vol;
    if (v > 0.0f && isfinite(v) && v > mmax) mmax = v;
  }
  int b = 0;
  for (int64_t v = np; v > 0; v >>= 1) ++b;
  int s = 0;
  if (mmax > 0.0f) {
    int e;
    (void)frexp((double)mmax, &e);
    s = 61 - (e - 1) - b;
  }
  const int64_t n2 = (int64_t)n * n;
  for (int64_t p = 0; p < np; ++p) {
    int ix = 0, iy = 0, iz = 0;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const int in = cic_axis(px[p], dx, n, &ix, &fx) && cic_axis(py[p], dx, n, &iy, &fy) &&
                   cic_axis(pz[p], dx, n, &iz, &fz);
    if (keys_out)
      keys_out[p] = in ? (uint32_t)(((int64_t)(ix + 1) * m + (iy + 1)) * m + (iz + 1)) : sentinel;
    if (!in) continue;
    const float r[4] = {fx, fy, fz, mass[p] * inv_vol};
    float t[8];
    corner_terms(r, t);
    for (int c = 0; c < 8; ++c) {
      const int jx = ix + (c >> 2), jy = iy + ((c >> 1) & 1), jz = iz + (c & 1);
      if (jx < 0 || jx >= n || jy < 0 || jy >= n || jz < 0 || jz >= n) continue;
      if (!(t[c] > 0.0f && isfinite(t[c]))) continue;
      acc[jx * n2 + (int64_t)jy * n + jz] += (uint64_t)llrint(ldexp((double)t[c], s));
    }
  }
  for (int64_t i = 0; i < nodes; ++i) rho[i] = (float)ldexp((double)acc[i], -s);
  free(acc);
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * 2. OpenMP port of the reference step (CPU baseline)
 * ------------------------------------------------------------------------------------- */

int cpu_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void cpu_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

/* deposit.py:47-82: zero rho, then 8 float atomic adds per particle */
void cpu_deposit_atomic(const float* px, const float* py, const float* pz, const float* mass,
                        int64_t np, int n, float dx, float* rho) {
  const int64_t n2 = (int64_t)n * n;
  const float inv_vol = 1.0f / ((dx * dx) * dx);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n2 * n; ++i) rho[i] = 0.0f;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < np; ++p) {
    const float gx = px[p] / dx, gy = py[p] / dx, gz = pz[p] / dx;
    if (!(gx >= -1.0f && gx < (float)n && gy >= -1.0f && gy < (float)n && gz >= -1.0f &&
          gz < (float)n))
      continue;
    const int ix = (int)floorf(gx), iy = (int)floorf(gy), iz = (int)floorf(gz);
    const float fx = gx - (float)ix, fy = gy - (float)iy, fz = gz - (float)iz;
    const float mm = mass[p] * inv_vol;
    for (int c = 0; c < 8; ++c) {
      const int cx = c >> 2, cy = (c >> 1) & 1, cz = c & 1;
      const int jx = ix + cx, jy = iy + cy, jz = iz + cz;
      if (jx < 0 || jx >= n || jy < 0 || jy >= n || jz < 0 || jz >= n) continue;
      const float wx = cx ? fx : 1.0f - fx, wy = cy ? fy : 1.0f - fy, wz = cz ? fz : 1.0f - fz;
      const float v = mm * wx * wy * wz;
#pragma omp atomic
      rho[jx * n2 + (int64_t)jy * n + jz] += v;
    }
  }
}

/* deposit.py:132-174 */
void cpu_grad(const float* phi, float* gx, float* gy, float* gz, int n, float dx) {
  const int64_t n2 = (int64_t)n * n;
  const float inv_dx = 1.0f / dx, inv_2dx = 1.0f / (2.0f * dx);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      for (int k = 0; k < n; ++k) {
        const int64_t c = i * n2 + (int64_t)j * n + k;
        float a, b, s;
        if (i == 0) { a = phi[c + n2]; b = phi[c]; s = inv_dx; }
        else if (i == n - 1) { a = phi[c]; b = phi[c - n2]; s = inv_dx; }
        else { a = phi[c + n2]; b = phi[c - n2]; s = inv_2dx; }
        gx[c] = -(a - b) * s;
        if (j == 0) { a = phi[c + n]; b = phi[c]; s = inv_dx; }
        else if (j == n - 1) { a = phi[c]; b = phi[c - n]; s = inv_dx; }
        else { a = phi[c + n]; b = phi[c - n]; s = inv_2dx; }
        gy[c] = -(a - b) * s;
        if (k == 0) { a = phi[c + 1]; b = phi[c]; s = inv_dx; }
        else if (k == n - 1) { a = phi[c]; b = phi[c - 1]; s = inv_dx; }
        else { a = phi[c + 1]; b = phi[c - 1]; s = inv_2dx; }
        gz[c] = -(a - b) * s;
      }
    }
  }
}

/* deposit.py:85-129 */
void cpu_gather(const float* px, const float* py, const float* pz, const float* gx,
                const float* gy, const float* gz, float* ax, float* ay, float* az, int64_t np,
                int n, float dx) {
  const int64_t n2 = (int64_t)n * n;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < np; ++p) {
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    int ix, iy, iz;
    float fx, fy, fz;
    if (cic_axis(px[p], dx, n, &ix, &fx) && cic_axis(py[p], dx, n, &iy, &fy) &&
        cic_axis(pz[p], dx, n, &iz, &fz)) {
      for (int c = 0; c < 8; ++c) {
        const int cx = c >> 2, cy = (c >> 1) & 1, cz = c & 1;
        const int jx = ix + cx, jy = iy + cy, jz = iz + cz;
        if (jx < 0 || jx >= n || jy < 0 || jy >= n || jz < 0 || jz >= n) continue;
        const float wx = cx ? fx : 1.0f - fx, wy = cy ? fy : 1.0f - fy, wz = cz ? fz : 1.0f - fz;
        const float w = wx * wy * wz;
        const int64_t node = jx * n2 + (int64_t)jy * n + jz;
        a0 = a0 + w * gx[node];
        a1 = a1 + w * gy[node];
        a2 = a2 + w * gz[node];
      }
    }
    ax[p] = a0;
    ay[p] = a1;
    az[p] = a2;
  }
}

/* integrator.py:27-42 / 45-60 */
void cpu_kick(float* vx, float* vy, float* vz, const float* ax, const float* ay, const float* az,
              int64_t np, float dt) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < np; ++p) {
    vx[p] = vx[p] + ax[p] * dt;
    vy[p] = vy[p] + ay[p] * dt;
    vz[p] = vz[p] + az[p] * dt;
  }
}

void cpu_drift(float* px, float* py, float* pz, const float* vx, const float* vy,
               const float* vz, int64_t np, float dt) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < np; ++p) {
    px[p] = px[p] + vx[p] * dt;
    py[p] = py[p] + vy[p] * dt;
    pz[p] = pz[p] + vz[p] * dt;
  }
}

/* ---- multigrid (solver/multigrid.py) ---- */

typedef struct {
  int n;
  float h2, inv_h2;
  float *phi, *rhs, *res;
} cpu_level;

typedef struct {
  int levels;
  cpu_level lv[16];
  float dx;
  double G;
  int n_cycles, nu1, nu2, coarse_sweeps, min_cycles;
  double tol;
  int last_cycles;
} cpu_mg;

static void rb_pass(float* phi, const float* rhs, int n, float h2, int color) {
  const int64_t n2 = (int64_t)n * n;
#pragma omp parallel for schedule(static)
  for (int i = 1; i < n - 1; ++i) {
    for (int j = 1; j < n - 1; ++j) {
      int k0 = 1 + ((i + j + 1 + color) & 1);
      for (int k = k0; k < n - 1; k += 2) {
        const int64_t c = i * n2 + (int64_t)j * n + k;
        const float s = phi[c + n2] + phi[c - n2] + phi[c + n] + phi[c - n] + phi[c + 1] + phi[c - 1];
        phi[c] = (s - h2 * rhs[c]) / 6.0f;
      }
    }
  }
}

static void smooth(cpu_level* L, int sweeps) {
  for (int q = 0; q < sweeps; ++q) {
    rb_pass(L->phi, L->rhs, L->n, L->h2, 0);
    rb_pass(L->phi, L->rhs, L->n, L->h2, 1);
  }
}

static void residual(const float* phi, const float* rhs, float* res, int n, float inv_h2) {
  const int64_t n2 = (int64_t)n * n;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      for (int k = 0; k < n; ++k) {
        const int64_t c = i * n2 + (int64_t)j * n + k;
        float r = 0.0f;
        if (i >= 1 && i <= n - 2 && j >= 1 && j <= n - 2 && k >= 1 && k <= n - 2) {
          const float lap = (phi[c + n2] + phi[c - n2] + phi[c + n] + phi[c - n] + phi[c + 1] +
                             phi[c - 1] - 6.0f * phi[c]) * inv_h2;
          r = rhs[c] - lap;
        }
        res[c] = r;
      }
    }
  }
}

static void restrict_fw(const float* res_f, float* rhs_c, int nc) {
  const int nf = 2 * nc;
  const int64_t nf2 = (int64_t)nf * nf, nc2 = (int64_t)nc * nc;
#pragma omp parallel for schedule(static)
  for (int ic = 0; ic < nc; ++ic) {
    for (int jc = 0; jc < nc; ++jc) {
      for (int kc = 0; kc < nc; ++kc) {
        float out = 0.0f;
        if (ic >= 1 && ic <= nc - 2 && jc >= 1 && jc <= nc - 2 && kc >= 1 && kc <= nc - 2) {
          const int64_t f0 = (2 * ic) * nf2 + (int64_t)(2 * jc) * nf + 2 * kc;
          float acc = 0.0f;
          for (int a = -1; a <= 1; ++a)
            for (int b = -1; b <= 1; ++b)
              for (int d = -1; d <= 1; ++d) {
                const float w = (a == 0 ? 2.0f : 1.0f) * (b == 0 ? 2.0f : 1.0f) * (d == 0 ? 2.0f : 1.0f);
                acc = acc + w * res_f[f0 + a * nf2 + b * nf + d];
              }
          out = acc / 64.0f;
        }
        rhs_c[ic * nc2 + (int64_t)jc * nc + kc] = out;
      }
    }
  }
}

static void prolong_add(const float* phi_c, float* phi_f, int nf) {
  const int nc = nf / 2;
  const int64_t nf2 = (int64_t)nf * nf, nc2 = (int64_t)nc * nc;
#pragma omp parallel for schedule(static)
  for (int i = 1; i < nf - 1; ++i) {
    for (int j = 1; j < nf - 1; ++j) {
      for (int k = 1; k < nf - 1; ++k) {
        const int ci = i >> 1, cj = j >> 1, ck = k >> 1;
        const int oi = i & 1, oj = j & 1, ok = k & 1;
        float val = 0.0f;
        for (int di = 0; di < 2; ++di) {
          const float wi = oi ? 0.5f : (di == 0 ? 1.0f : 0.0f);
          for (int dj = 0; dj < 2; ++dj) {
            const float wj = oj ? 0.5f : (dj == 0 ? 1.0f : 0.0f);
            for (int dk = 0; dk < 2; ++dk) {
              const float wk = ok ? 0.5f : (dk == 0 ? 1.0f : 0.0f);
              const float w = wi * wj * wk;
              if (w > 0.0f) val = val + w * phi_c[(ci + di) * nc2 + (int64_t)(cj + dj) * nc + (ck + dk)];
            }
          }
        }
        const int64_t c = i * nf2 + (int64_t)j * nf + k;
        phi_f[c] = phi_f[c] + val;
      }
    }
  }
}

static void vcycle(cpu_mg* mg, int l) {
  cpu_level* L = &mg->lv[l];
  if (l == mg->levels - 1) {
    smooth(L, mg->coarse_sweeps);
    return;
  }
  smooth(L, mg->nu1);
  residual(L->phi, L->rhs, L->res, L->n, L->inv_h2);
  cpu_level* C = &mg->lv[l + 1];
  restrict_fw(L->res, C->rhs, C->n);
  memset(C->phi, 0, sizeof(float) * (size_t)C->n * C->n * C->n);
  vcycle(mg, l + 1);
  prolong_add(C->phi, L->phi, L->n);
  smooth(L, mg->nu2);
}

static double sumsq(const float* x, int64_t count) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < count; ++i) s += (double)x[i] * (double)x[i];
  return s;
}

void* cpu_mg_create(int n, float dx, double G, int n_cycles, int nu1, int nu2, int coarse_sweeps,
                    int min_cycles, double tol) {
  cpu_mg* mg = (cpu_mg*)calloc(1, sizeof(cpu_mg));
  if (!mg) return NULL;
  mg->dx = dx;
  mg->G = G;
  mg->n_cycles = n_cycles;
  mg->nu1 = nu1;
  mg->nu2 = nu2;
  mg->coarse_sweeps = coarse_sweeps;
  mg->min_cycles = min_cycles < 1 ? 1 : min_cycles;
  mg->tol = tol;
  int s = n, l = 0;
  for (;;) {
    cpu_level* L = &mg->lv[l];
    L->n = s;
    const double h = (double)dx * (double)(1 << l);
    L->h2 = (float)(h * h);
    L->inv_h2 = (float)(1.0 / (h * h));
    const size_t cnt = (size_t)s * s * s;
    L->phi = (float*)calloc(cnt, sizeof(float));
    L->rhs = (float*)calloc(cnt, sizeof(float));
    L->res = (float*)calloc(cnt, sizeof(float));
    if (!L->phi || !L->rhs || !L->res) return NULL;
    ++l;
    if (!(s % 2 == 0 && s / 2 >= 3) || l >= 16) break;
    s /= 2;
  }
  mg->levels = l;
  return mg;
}

void cpu_mg_destroy(void* handle) {
  cpu_mg* mg = (cpu_mg*)handle;
  if (!mg) return;
  for (int l = 0; l < mg->levels; ++l) {
    free(mg->lv[l].phi);
    free(mg->lv[l].rhs);
    free(mg->lv[l].res);
  }
  free(mg);
}

/* solver/multigrid.py:455-487; returns the number of V-cycles run */
int cpu_mg_solve(void* handle, const float* rho, float* phi, int warm) {
  cpu_mg* mg = (cpu_mg*)handle;
  cpu_level* L0 = &mg->lv[0];
  const int n = L0->n;
  const int64_t n2 = (int64_t)n * n, cnt = n2 * n;
  const double dxd = (double)mg->dx, vol = dxd * dxd * dxd;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0, s8 = 0, s9 = 0;
#pragma omp parallel for reduction(+ : s0, s1, s2, s3, s4, s5, s6, s7, s8, s9) schedule(static)
  for (int64_t c = 0; c < cnt; ++c) {
    const int k = (int)(c % n), j = (int)((c / n) % n), i = (int)(c / n2);
    const double m = (double)rho[c] * vol, x = i * dxd, y = j * dxd, z = k * dxd;
    s0 += m; s1 += m * x; s2 += m * y; s3 += m * z;
    s4 += m * x * x; s5 += m * y * y; s6 += m * z * z;
    s7 += m * x * y; s8 += m * x * z; s9 += m * y * z;
  }
  double cx = 0, cy = 0, cz = 0, sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0, M = 0;
  if (s0 > 0.0) {
    M = s0; cx = s1 / s0; cy = s2 / s0; cz = s3 / s0;
    sxx = s4 - s0 * cx * cx; syy = s5 - s0 * cy * cy; szz = s6 - s0 * cz * cz;
    sxy = s7 - s0 * cx * cy; sxz = s8 - s0 * cx * cz; syz = s9 - s0 * cy * cz;
  }
  const float factor = (float)(4.0 * 3.14159265358979323846 * mg->G);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < cnt; ++c) L0->rhs[c] = factor * rho[c];
  if (warm) memcpy(L0->phi, phi, sizeof(float) * cnt);
  else memset(L0->phi, 0, sizeof(float) * cnt);
  {
    const float g = (float)mg->G, mf = (float)M, fcx = (float)cx, fcy = (float)cy, fcz = (float)cz;
    const float fxx = (float)sxx, fyy = (float)syy, fzz = (float)szz;
    const float fxy = (float)sxy, fxz = (float)sxz, fyz = (float)syz;
    const float tr = fxx + fyy + fzz;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
          if (!(i == 0 || i == n - 1 || j == 0 || j == n - 1 || k == 0 || k == n - 1)) continue;
          const float rx = (float)i * mg->dx - fcx, ry = (float)j * mg->dx - fcy, rz = (float)k * mg->dx - fcz;
          const float d2 = rx * rx + ry * ry + rz * rz + 1e-12f;
          const float d = sqrtf(d2);
          const float nx = rx / d, ny = ry / d, nz = rz / d;
          const float nsn = nx * nx * fxx + ny * ny * fyy + nz * nz * fzz +
                            2.0f * (nx * ny * fxy + nx * nz * fxz + ny * nz * fyz);
          const float quad = 3.0f * nsn - tr;
          L0->phi[i * n2 + (int64_t)j * n + k] = -g * (mf / d + quad / (2.0f * d2 * d));
        }
  }
  const double rhs_norm = sqrt(sumsq(L0->rhs, cnt));
  mg->last_cycles = mg->n_cycles;
  for (int c = 0; c < mg->n_cycles; ++c) {
    vcycle(mg, 0);
    if (c + 1 >= mg->min_cycles && rhs_norm > 0.0) {
      residual(L0->phi, L0->rhs, L0->res, n, L0->inv_h2);
      if (sqrt(sumsq(L0->res, cnt)) < mg->tol * rhs_norm) {
        mg->last_cycles = c + 1;
        break;
      }
    }
  }
  memcpy(phi, L0->phi, sizeof(float) * cnt);
  return mg->last_cycles;
}

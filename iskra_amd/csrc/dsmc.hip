/*
 * dsmc.hip -- direct simulation Monte Carlo pairing per cell on gfx950.
 *
 * Reference: PIC.perform!(::DirectSimulationMonteCarlo) (Chemistry/src/
 * dsmc.jl:81-142) with perform!(::DSMC.ElasticCollision) (dsmc.jl:32-79).
 * Per cell (i, j) with Na source and Nb target particles (both at least 2):
 *     na = Na Wa / (dx dy),  Nc = na Nb dt sgmax / (Pab + (Wb/Wa) Pba)
 *     (doubled when source and target differ), plus the cell's remainder
 *     carried from the previous call; floor(Nc) trials, each drawing a source
 *     and a target particle of the cell uniformly (a target index equal to
 *     the source index is drawn again when the species differ), accepted
 *     when sigma(g) g / sgmax >= U, and then scattered isotropically in the
 *     centre-of-mass frame (the vss_inv = 1 branch):
 *         B = 2U - 1, A = sqrt(1 - B^2), C = 2 pi U,
 *         v_r' = |g| (B, A cos C, A sin C),
 *         v_s' = v_cm + mr2 v_r',  v_t' = v_cm - mr1 v_r'.
 * sgmax is the reference's max(sigma nodes) * (node of that maximum).
 *
 * Layout: the source and target species are sorted by exact cell first (the
 * library's stable counting sort), so the particles of a cell are one
 * contiguous range in particle order, the order in which the reference's
 * cache! lists them.  One work-item owns one cell and runs that cell's trials
 * in order: cells share no particle, so the result does not depend on the
 * schedule, and a CPU twin that takes the same draws reproduces it bit for
 * bit.  Draws: Philox4x32-10, key = seed, counter = (cell, call, stream,
 * block), two 52-bit uniforms per block in the order the reference draws.
 * Kinematics in FP64 from the FP32 velocities, rounded back to FP32.
 */
#include <math.h>

#include <algorithm>
#include <vector>

#include "isk_fp64math.h"
#include "isk_internal.h"

#define DSMC_BLOCK 256

struct DsDraws {
  uint32_t cell, step, stream, k0, k1;
  int block;
  int half;
  IskU4 cur;
  __device__ double next() {
    if (half == 0) {
      IskU4 c = {cell, step, stream, (uint32_t)block};
      cur = isk_philox(c, k0, k1);
      ++block;
      half = 1;
      return isk_u01(cur.x, cur.y);
    }
    half = 0;
    return isk_u01(cur.z, cur.w);
  }
};

/* rand(list) of a list of n entries: entry floor(n U) */
__device__ __forceinline__ int ds_pick(double u, int n) {
  int k = (int)floor((double)n * u);
  return k < n - 1 ? k : n - 1;
}

/* sigma(g): piecewise linear over the table, flat outside (cross_section.jl:3-17) */
__device__ __forceinline__ double ds_sigma(const double *tg, const double *ts, int n, double g) {
  if (g <= tg[0]) return ts[0];
  if (g >= tg[n - 1]) return ts[n - 1];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (tg[mid] <= g) lo = mid;
    else hi = mid;
  }
  double t = (g - tg[lo]) / (tg[hi] - tg[lo]);
  return ts[lo] * (1.0 - t) + ts[hi] * t;
}

/* cell ranges of a sorted species: first and one-past-last particle of every
 * cell (cells without particles keep 0, 0 from the fill) */
__global__ void k_ds_ranges(const int32_t *cell, const long long *hdr, int ncx, int *start,
                            int *end) {
  const long long np = hdr[HDR_NP];
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < np;
       p += (long long)gridDim.x * blockDim.x) {
    const int c = cell[p];
    if (c < 0) continue;
    const int key = (c & 0xFFFF) + ncx * (c >> 16);
    if (p == 0 || cell[p - 1] != c) start[key] = (int)p;
    if (p == np - 1 || cell[p + 1] != c) end[key] = (int)(p + 1);
  }
}

__global__ void k_ds_zero(int *a, int n) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) a[k] = 0;
}

struct DsArgs {
  float *avx, *avy, *avz, *bvx, *bvy, *bvz;
  const int *as, *ae, *bs, *be;
  int same;
  int ncx, ncy, nx;
  double Wa, Wb, Pab, Pba, dx, dy, dt, sgmax, mr1, mr2;
  const double *tg, *ts;
  int ntab;
  double *rem;
  unsigned long long *nu, *events;
  uint32_t step, stream, k0, k1;
};

__global__ __launch_bounds__(DSMC_BLOCK) void k_ds_cells(DsArgs a) {
  const int t = blockIdx.x * DSMC_BLOCK + threadIdx.x;
  if (t >= a.ncx * a.ncy) return;
  const int a0 = a.as[t], Na = a.ae[t] - a0;
  const int b0 = a.bs[t], Nb = a.be[t] - b0;
  if (Na < 2 || Nb < 2) return;
  /* dsmc.jl:112-119, in the reference's operation order */
  const double na = (double)Na * a.Wa / (a.dx * a.dy);
  double Nc = na * (double)Nb * a.dt * a.sgmax;
  Nc /= a.Pab + (a.Wb / a.Wa) * a.Pba;
  if (!a.same) Nc *= 2.0;
  Nc += a.rem[t];
  const double fl = floor(Nc);
  a.rem[t] = Nc - fl;
  const long long ntr = (long long)fl;
  DsDraws d;
  d.cell = (uint32_t)t;
  d.step = a.step;
  d.stream = a.stream;
  d.k0 = a.k0;
  d.k1 = a.k1;
  d.block = 0;
  d.half = 0;
  unsigned long long acc = 0;
  for (long long k = 0; k < ntr; ++k) {
    const int sR = a0 + ds_pick(d.next(), Na);
    int tR = b0 + ds_pick(d.next(), Nb);
    while (!a.same && sR == tR) tR = b0 + ds_pick(d.next(), Nb);
    const double sx = a.avx[sR], sy = a.avy[sR], sz = a.avz[sR];
    const double tx = a.bvx[tR], ty = a.bvy[tR], tz = a.bvz[tR];
    const double gx = sx - tx, gy = sy - ty, gz = sz - tz;
    const double gn = sqrt(gx * gx + gy * gy + gz * gz);
    const double sg = ds_sigma(a.tg, a.ts, a.ntab, gn) * gn;
    const double P = sg / a.sgmax;
    const double R = d.next();
    if (P < R) continue;
    /* perform!(::DSMC.ElasticCollision) (dsmc.jl:32-79), equal weights */
    const double B = 2.0 * d.next() - 1.0;
    const double A = sqrt(1.0 - B * B);
    double sC, cC;
    isk_sincos2pi(d.next(), &sC, &cC);
    const double rx = gn * B, ry = gn * (A * cC), rz = gn * (A * sC);
    const double cx = a.mr1 * sx + a.mr2 * tx, cy = a.mr1 * sy + a.mr2 * ty,
                 cz = a.mr1 * sz + a.mr2 * tz;
    a.avx[sR] = (float)(cx + a.mr2 * rx);
    a.avy[sR] = (float)(cy + a.mr2 * ry);
    a.avz[sR] = (float)(cz + a.mr2 * rz);
    a.bvx[tR] = (float)(cx - a.mr1 * rx);
    a.bvy[tR] = (float)(cy - a.mr1 * ry);
    a.bvz[tR] = (float)(cz - a.mr1 * rz);
    ++acc;
  }
  if (acc) {
    const int i = t % a.ncx, j = t / a.ncx;
    a.nu[j * a.nx + i] += acc; /* the only writer of this node */
    atomicAdd(a.events, acc);
  }
}

int isk_dsmc_add(iskra_ctx *c, int src, int tgt, int ntab, const double *tab_g,
                 const double *tab_sig, int *id) {
  if (c->g.dim != 2) {
    isk_set_error("dsmc: the reference's perform! needs a 2-D grid (dsmc.jl:82-83)");
    return ISKRA_E_UNSUPPORTED;
  }
  if (c->ndsmc >= ISK_MAX_DSMC) {
    isk_set_error("dsmc: at most %d interactions", ISK_MAX_DSMC);
    return ISKRA_E_ARG;
  }
  if (ntab < 1 || !tab_g || !tab_sig) {
    isk_set_error("dsmc: empty cross-section table");
    return ISKRA_E_ARG;
  }
  for (int k = 1; k < ntab; ++k)
    if (!(tab_g[k] > tab_g[k - 1])) {
      isk_set_error("dsmc: table speeds must increase");
      return ISKRA_E_ARG;
    }
  const IskSpecies &A = c->sp[src], &B = c->sp[tgt];
  if (A.w0 != B.w0) {
    /* dsmc.jl:71-77 updates target.v[s,:] with mr2, an index of the source
     * list, which can land on a particle of another cell */
    isk_set_error("dsmc: source and target weights differ (%g, %g): the reference's "
                  "unequal-weight update writes into the wrong particle (dsmc.jl:71-77)",
                  A.w0, B.w0);
    return ISKRA_E_UNSUPPORTED;
  }
  const int idx = c->ndsmc;
  IskDsmc &m = c->dsmc[idx];
  m = IskDsmc();
  m.src = src;
  m.tgt = tgt;
  m.ntab = ntab;
  /* sgmax = maximum(rate) * argmax(rate): the largest sigma node times the
   * speed node where it first occurs (cross_section.jl:16-17, dsmc.jl:104) */
  int km = 0;
  for (int k = 1; k < ntab; ++k)
    if (tab_sig[k] > tab_sig[km]) km = k;
  m.sgmax = tab_sig[km] * tab_g[km];
  if (!(m.sgmax > 0.0)) {
    isk_set_error("dsmc: max(sigma) * argmax(sigma) must be positive");
    return ISKRA_E_ARG;
  }
  const int ncell = c->g.ncx * c->g.ncy, nn = isk_nodes(c);
  if (isk_alloc((void **)&m.d_g, ntab * 8) || isk_alloc((void **)&m.d_sig, ntab * 8) ||
      isk_alloc((void **)&m.rem, (size_t)ncell * 8) || isk_alloc((void **)&m.nu, (size_t)nn * 8) ||
      isk_alloc((void **)&m.nu_acc, (size_t)nn * 8) ||
      isk_alloc((void **)&m.events, 8) || isk_alloc((void **)&m.rng, (size_t)ncell * 4 * 4))
    return ISKRA_E_HIP;
  ISK_HIP(hipMemcpy(m.d_g, tab_g, ntab * 8, hipMemcpyHostToDevice));
  ISK_HIP(hipMemcpy(m.d_sig, tab_sig, ntab * 8, hipMemcpyHostToDevice));
  ISK_HIP(hipMemset(m.rem, 0, (size_t)ncell * 8));
  ISK_HIP(hipMemset(m.nu, 0, (size_t)nn * 8));
  ISK_HIP(hipMemset(m.nu_acc, 0, (size_t)nn * 8));
  ISK_HIP(hipMemset(m.events, 0, 8));
  m.step = 0;
  m.stream = 0x44530000u + (uint32_t)idx;
  c->ndsmc = idx + 1;
  *id = idx;
  return ISKRA_OK;
}

void isk_dsmc_release(IskDsmc *m) {
  void *p[] = {m->d_g, m->d_sig, m->rem, m->nu, m->nu_acc, m->events, m->rng};
  for (void *q : p) isk_free(q);
  *m = IskDsmc();
}

int isk_dsmc_run(iskra_ctx *c, int idx, double dt, int64_t *events) {
  IskDsmc &m = c->dsmc[idx];
  const int ncell = c->g.ncx * c->g.ncy;
  const bool same = m.src == m.tgt;
  int rc;
  /* cache! (dsmc.jl:25-30): particles of a cell as one range, in order */
  if ((rc = isk_sort_species(c, m.src, 1))) return rc;
  if (!same && (rc = isk_sort_species(c, m.tgt, 1))) return rc;
  int *as = m.rng, *ae = m.rng + ncell, *bs = m.rng + 2 * ncell, *be = m.rng + 3 * ncell;
  const int zb = std::min((4 * ncell + 255) / 256, 4096);
  hipLaunchKernelGGL(k_ds_zero, dim3(zb), dim3(256), 0, c->stream, m.rng, 4 * ncell);
  ISK_CHECK_LAUNCH();
  const IskSpecies &A = c->sp[m.src], &B = c->sp[m.tgt];
  auto ranges = [&](const IskSpecies &S, int *st, int *en) -> int {
    const int g = (int)std::min<int64_t>((S.cap + 255) / 256, 8192);
    hipLaunchKernelGGL(k_ds_ranges, dim3(std::max(g, 1)), dim3(256), 0, c->stream,
                       S.cell[S.cur], (const long long *)S.hdr, c->g.ncx, st, en);
    ISK_CHECK_LAUNCH();
    return ISKRA_OK;
  };
  if ((rc = ranges(A, as, ae))) return rc;
  if (same) {
    bs = as;
    be = ae;
  } else if ((rc = ranges(B, bs, be))) {
    return rc;
  }
  DsArgs a;
  a.avx = A.vx[A.cur];
  a.avy = A.vy[A.cur];
  a.avz = A.vz[A.cur];
  a.bvx = B.vx[B.cur];
  a.bvy = B.vy[B.cur];
  a.bvz = B.vz[B.cur];
  a.as = as;
  a.ae = ae;
  a.bs = bs;
  a.be = be;
  a.same = same ? 1 : 0;
  a.ncx = c->g.ncx;
  a.ncy = c->g.ncy;
  a.nx = c->g.nx;
  a.Wa = A.w0;
  a.Wb = B.w0;
  if (a.Wa > a.Wb) {
    a.Pab = a.Wb / a.Wa;
    a.Pba = 1.0;
  } else {
    a.Pab = 1.0;
    a.Pba = a.Wa / a.Wb;
  }
  a.dx = c->g.dx;
  a.dy = c->g.dy;
  a.dt = dt;
  a.sgmax = m.sgmax;
  a.mr1 = A.m / (A.m + B.m);
  a.mr2 = B.m / (A.m + B.m);
  a.tg = m.d_g;
  a.ts = m.d_sig;
  a.ntab = m.ntab;
  a.rem = m.rem;
  a.nu = m.nu;
  a.events = m.events;
  a.step = m.step;
  a.stream = m.stream;
  a.k0 = (uint32_t)c->seed;
  a.k1 = (uint32_t)(c->seed >> 32);
  if (events) {
    hipLaunchKernelGGL(k_ds_zero, dim3(1), dim3(64), 0, c->stream, (int *)m.events, 2);
    ISK_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_ds_cells, dim3((ncell + DSMC_BLOCK - 1) / DSMC_BLOCK), dim3(DSMC_BLOCK),
                     0, c->stream, a);
  ISK_CHECK_LAUNCH();
  ++m.step;
  if (events) {
    unsigned long long e = 0;
    ISK_HIP(hipMemcpyAsync(&e, m.events, 8, hipMemcpyDeviceToHost, c->stream));
    ISK_HIP(hipStreamSynchronize(c->stream));
    *events = (int64_t)e;
  }
  return ISKRA_OK;
}

/* reset 0: running sum + live counts; 1: the same, then both cleared; 2: the
 * live counts alone (since the last fold) */
int isk_dsmc_counts(iskra_ctx *c, int idx, int reset, int64_t *out) {
  IskDsmc &m = c->dsmc[idx];
  const int nn = isk_nodes(c);
  ISK_HIP(hipStreamSynchronize(c->stream));
  std::vector<unsigned long long> nu(nn), acc(nn);
  ISK_HIP(hipMemcpy(nu.data(), m.nu, (size_t)nn * 8, hipMemcpyDeviceToHost));
  if (reset != 2) ISK_HIP(hipMemcpy(acc.data(), m.nu_acc, (size_t)nn * 8, hipMemcpyDeviceToHost));
  if (out)
    for (int n = 0; n < nn; ++n) out[n] = (int64_t)(nu[n] + acc[n]);
  if (reset == 1) {
    ISK_HIP(hipMemset(m.nu, 0, (size_t)nn * 8));
    ISK_HIP(hipMemset(m.nu_acc, 0, (size_t)nn * 8));
  }
  return ISKRA_OK;
}

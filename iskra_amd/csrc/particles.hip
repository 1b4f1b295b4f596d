/*
 * particles.hip -- particle kernels of the electrostatic PIC step on gfx950.
 *
 *  k_push<DIM>      fused gather + push + boundary + CIC deposit, one
 *                   work-group per chunk of a tile; E window and deposit
 *                   window in LDS; deposit accumulated as 64-bit fixed point
 *                   (integer adds are associative, so the result does not
 *                   depend on particle order or scheduling).
 *  k_merge<DIM>     per node, sums the chunk windows that cover it in a fixed
 *                   order plus the global fallback accumulator.
 *  k_sort_*         stable counting sort by tile (or by cell in tile):
 *                   block histograms -> column scans -> stable block scatter.
 *  k_rho            rho = sum_s q_s n_s + background,  n_s = deposit / volume.
 *
 * Reference semantics followed (paths relative to the reference tree):
 *  particle_cell            ParticleInCell/src/ParticleInCell.jl:28-35
 *  grid_to_particle (CIC)   ParticleInCell/src/pic/cloud_in_cell.jl:20-36
 *  push_in_cartesian! (B=0) ParticleInCell/src/pic/pushers.jl:37-50
 *  wrap! / discard!         ParticleInCell/src/pic/surfaces/wrap.jl:1-33
 *  particle_to_grid (CIC)   ParticleInCell/src/pic/cloud_in_cell.jl:1-18
 *  density, cell_volume     pic/kinetic.jl:53, RegularGrids/src/RegularGrids.jl:26-38
 *  rho accumulation         ParticleInCell/src/ParticleInCell.jl:118-124
 */
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <vector>

#include "isk_fp64math.h"
#include "isk_internal.h"

/* push work-group shape: threads, and particles per thread per batch (one
 * batch covers a 2048-particle chunk). 2-D: 512 x 4 keeps the kernel at 59
 * VGPRs (8 waves per SIMD), 2 % faster than 256 x 8 at 96 VGPRs on the
 * 256^2-cell slab; 1-D keeps 256 x 8 */
template <int DIM>
struct PushShape {
  static constexpr int block = DIM == 2 ? 512 : 256;
  static constexpr int unroll = DIM == 2 ? 4 : 8;
};
/* the largest float below 1 (1 - 2^-24): an offset on the upper edge of its cell */
#define ISK_BELOW_ONE 0x1.fffffep-1f
#define SORT_BLOCK 256
#define SORT_PT 32
#define SORT_CHUNK (SORT_BLOCK * SORT_PT)
#define SORT_SEG 32
#define SORT_DEPTH 8 /* scatter rounds whose loads are in flight together */
static_assert(SORT_PT % 2 == 0, "the scatter handles rounds in pairs");

/* ------------------------------------------------------------------------ */
/* small utility kernels                                                     */
/* ------------------------------------------------------------------------ */
__global__ void k_fill_i64(long long *p, long long n, long long v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

int isk_fill_i64(iskra_ctx *c, void *p, int64_t n, long long v) {
  if (n <= 0) return ISKRA_OK;
  int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_i64, dim3(blocks), dim3(256), 0, c->stream,
                     (long long *)p, (long long)n, v);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

/* ------------------------------------------------------------------------ */
/* upload (host conversion) and Maxwellian loader                           */
/* ------------------------------------------------------------------------ */
/* particle_cell with the grid origin removed: s = (x - x0)/dh, i = floor(s),
 * offset = s - i rounded to FP32 (an offset that rounds up to 1 moves to the
 * next cell with offset 0). */
static int cell_of(double x, double x0, double dh, int nc, int *i, float *o) {
  double s = (x - x0) / dh;
  double fl = floor(s);
  if (!(fl >= 0.0) || fl >= (double)nc) return -1;
  int ii = (int)fl;
  float of = (float)(s - fl);
  if (of >= 1.0f) {
    of = 0.0f;
    ii += 1;
    if (ii >= nc) return -1;
  }
  *i = ii;
  *o = of;
  return 0;
}

int isk_convert_upload(iskra_ctx *c, int sid, int64_t np, const double *x,
                       const double *v, const uint32_t *id) {
  IskSpecies *s = &c->sp[sid];
  const IskGrid &g = c->g;
  if (np < 0 || np > s->cap) {
    isk_set_error("upload: np=%lld outside [0, capacity=%lld]", (long long)np,
                  (long long)s->cap);
    return ISKRA_E_CAPACITY;
  }
  std::vector<int32_t> cell(np);
  std::vector<float> ox(np), oy(g.dim == 2 ? np : 0), vx(np), vy(np), vz(np);
  std::vector<uint32_t> ids(s->track_ids ? np : 0);
  uint32_t maxid = 0;
  for (int64_t p = 0; p < np; ++p) {
    int i = 0, j = 0;
    float fx = 0.f, fy = 0.f;
    if (cell_of(x[p], g.x0, g.dx, g.ncx, &i, &fx) != 0) {
      isk_set_error("upload: particle %lld has x=%g outside the grid",
                    (long long)p, x[p]);
      return ISKRA_E_ARG;
    }
    if (g.dim == 2) {
      if (cell_of(x[np + p], g.y0, g.dy, g.ncy, &j, &fy) != 0) {
        isk_set_error("upload: particle %lld has y=%g outside the grid",
                      (long long)p, x[np + p]);
        return ISKRA_E_ARG;
      }
      oy[p] = fy;
    }
    cell[p] = i | (j << 16);
    ox[p] = fx;
    vx[p] = (float)v[p];
    vy[p] = (float)v[np + p];
    vz[p] = (float)v[2 * np + p];
    if (s->track_ids) {
      ids[p] = id ? id[p] : (uint32_t)p;
      if (ids[p] > maxid) maxid = ids[p];
    }
  }
  int b = s->cur;
  if (np > 0) {
    ISK_HIP(hipMemcpyAsync(s->cell[b], cell.data(), np * 4, hipMemcpyHostToDevice, c->stream));
    ISK_HIP(hipMemcpyAsync(s->ox[b], ox.data(), np * 4, hipMemcpyHostToDevice, c->stream));
    if (g.dim == 2)
      ISK_HIP(hipMemcpyAsync(s->oy[b], oy.data(), np * 4, hipMemcpyHostToDevice, c->stream));
    ISK_HIP(hipMemcpyAsync(s->vx[b], vx.data(), np * 4, hipMemcpyHostToDevice, c->stream));
    ISK_HIP(hipMemcpyAsync(s->vy[b], vy.data(), np * 4, hipMemcpyHostToDevice, c->stream));
    ISK_HIP(hipMemcpyAsync(s->vz[b], vz.data(), np * 4, hipMemcpyHostToDevice, c->stream));
    if (s->track_ids)
      ISK_HIP(hipMemcpyAsync(s->id[b], ids.data(), np * 4, hipMemcpyHostToDevice, c->stream));
  }
  long long hdr[HDR_WORDS];
  ISK_HIP(hipMemcpyAsync(hdr, s->hdr, sizeof(hdr), hipMemcpyDeviceToHost, c->stream));
  ISK_HIP(hipStreamSynchronize(c->stream));
  hdr[HDR_NP] = np;
  hdr[HDR_NP_SORTED] = 0;
  hdr[HDR_NCHUNKS] = 0;
  hdr[HDR_NEXT_ID] = s->track_ids ? (long long)maxid + 1 : np;
  hdr[HDR_OVERFLOW] = 0;
  ISK_HIP(hipMemcpyAsync(s->hdr, hdr, sizeof(hdr), hipMemcpyHostToDevice, c->stream));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return isk_sort_species(c, sid, 0);
}

/* sample!(::MaxwellianSource): x = U*wx + dx, v = N(0,1)*wv + dv.
 * Draws: Philox4x32-10, key = seed, counter = (k_lo, k_hi, stream, block). */
struct SampleArgs {
  int32_t *cell;
  float *ox, *oy, *vx, *vy, *vz;
  uint32_t *id;
  long long *hdr;
  long long n;
  double wx[2], dxo[2], wv[3], dv[3];
  double x0, y0, dx, dy;
  int ncx, ncy, dim;
  uint32_t k0, k1, stream;
};

__global__ __launch_bounds__(256) void k_sample(SampleArgs a) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  long long np0 = a.hdr[HDR_NP];
  long long p = np0 + k;
  IskU4 c0 = {(uint32_t)k, (uint32_t)(k >> 32), a.stream, 0u};
  IskU4 r0 = isk_philox(c0, a.k0, a.k1);
  c0.w = 1u;
  IskU4 r1 = isk_philox(c0, a.k0, a.k1);
  c0.w = 2u;
  IskU4 r2 = isk_philox(c0, a.k0, a.k1);
  /* positions */
  double ux = isk_u01(r0.x, r0.y);
  double px = ux * a.wx[0] + a.dxo[0];
  double s = (px - a.x0) / a.dx;
  double fl = floor(s);
  int i = (int)fl;
  float of = (float)(s - fl);
  if (of >= 1.0f) { of = 0.0f; i += 1; }
  /* the host keeps the box inside the grid; these only catch a position
   * that rounding put just past an edge of it */
  if (i < 0) { i = 0; of = 0.0f; }
  if (i > a.ncx - 1) { i = a.ncx - 1; of = ISK_BELOW_ONE; }
  int j = 0;
  if (a.dim == 2) {
    double uy = isk_u01(r0.z, r0.w);
    double py = uy * a.wx[1] + a.dxo[1];
    double sy = (py - a.y0) / a.dy;
    double fly = floor(sy);
    j = (int)fly;
    float ofy = (float)(sy - fly);
    if (ofy >= 1.0f) { ofy = 0.0f; j += 1; }
    if (j < 0) { j = 0; ofy = 0.0f; }
    if (j > a.ncy - 1) { j = a.ncy - 1; ofy = ISK_BELOW_ONE; }
    a.oy[p] = ofy;
  }
  /* velocities: Box-Muller on (r1.xy, r1.zw) and (r2.xy, r2.zw) */
  double u1 = isk_u01(r1.x, r1.y), u2 = isk_u01(r1.z, r1.w);
  double u3 = isk_u01(r2.x, r2.y), u4 = isk_u01(r2.z, r2.w);
  double rad1 = sqrt(-2.0 * isk_log(u1)), rad2 = sqrt(-2.0 * isk_log(u3));
  double s1, c1, s2, c2;
  isk_sincos2pi(u2, &s1, &c1);
  isk_sincos2pi(u4, &s2, &c2);
  a.cell[p] = i | (j << 16);
  a.ox[p] = of;
  a.vx[p] = (float)(rad1 * c1 * a.wv[0] + a.dv[0]);
  a.vy[p] = (float)(rad1 * s1 * a.wv[1] + a.dv[1]);
  a.vz[p] = (float)(rad2 * c2 * a.wv[2] + a.dv[2]);
  if (a.id) a.id[p] = (uint32_t)(a.hdr[HDR_NEXT_ID] + k);
}

__global__ void k_sample_commit(long long *hdr, long long n) {
  hdr[HDR_NP] += n;
  hdr[HDR_NEXT_ID] += n;
}

int isk_sample_maxwellian(iskra_ctx *c, int sid, int64_t n, const double *wx,
                          const double *dx, const double *wv, const double *dv,
                          uint64_t seed, uint32_t stream) {
  IskSpecies *s = &c->sp[sid];
  long long hdr[HDR_WORDS];
  ISK_HIP(hipMemcpyAsync(hdr, s->hdr, sizeof(hdr), hipMemcpyDeviceToHost, c->stream));
  ISK_HIP(hipStreamSynchronize(c->stream));
  int64_t room = s->cap - hdr[HDR_NP];
  if (n > room) n = room; /* sample! truncates to the free capacity (sources.jl:30) */
  if (n <= 0) return ISKRA_OK;
  SampleArgs a;
  int b = s->cur;
  a.cell = s->cell[b];
  a.ox = s->ox[b];
  a.oy = s->oy[b];
  a.vx = s->vx[b];
  a.vy = s->vy[b];
  a.vz = s->vz[b];
  a.id = s->track_ids ? s->id[b] : nullptr;
  a.hdr = s->hdr;
  a.n = n;
  for (int d = 0; d < 2; ++d) {
    a.wx[d] = d < c->g.dim ? wx[d] : 0.0;
    a.dxo[d] = d < c->g.dim ? (dx ? dx[d] : 0.0) : 0.0;
  }
  for (int d = 0; d < 3; ++d) {
    a.wv[d] = wv[d];
    a.dv[d] = dv ? dv[d] : 0.0;
  }
  a.x0 = c->g.x0;
  a.y0 = c->g.y0;
  a.dx = c->g.dx;
  a.dy = c->g.dy;
  a.ncx = c->g.ncx;
  a.ncy = c->g.ncy;
  a.dim = c->g.dim;
  a.k0 = (uint32_t)seed;
  a.k1 = (uint32_t)(seed >> 32);
  a.stream = stream;
  int blocks = (int)((n + 255) / 256);
  hipLaunchKernelGGL(k_sample, dim3(blocks), dim3(256), 0, c->stream, a);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sample_commit, dim3(1), dim3(1), 0, c->stream, s->hdr,
                     (long long)n);
  ISK_CHECK_LAUNCH();
  return isk_sort_species(c, sid, 0);
}

/* ------------------------------------------------------------------------ */
/* counting sort                                                             */
/* ------------------------------------------------------------------------ */
struct KeyDesc {
  int mode; /* 0: tile, 1: cell inside tile */
  int dim, T, log2T, ntx, nbins;
};

__device__ __forceinline__ int sort_key(int c, const KeyDesc &k) {
  if (c < 0) return k.nbins; /* discarded -> last bin */
  int i = c & 0xFFFF;
  int j = c >> 16;
  if (k.mode == 0) {
    return (i >> k.log2T) + k.ntx * (j >> k.log2T);
  }
  return (i & (k.T - 1)) + (k.dim == 2 ? ((j & (k.T - 1)) << k.log2T) : 0);
}

__device__ void sort_hist_body(const int32_t *cell, const long long *hdr, int *H,
                               const KeyDesc &kd, const int blk) {
  extern __shared__ int sh[];
  const int nb1 = kd.nbins + 1;
  for (int b = threadIdx.x; b < nb1; b += SORT_BLOCK) sh[b] = 0;
  __syncthreads();
  const long long np = hdr[HDR_NP];
  const long long base = (long long)blk * SORT_CHUNK;
  const int lane = threadIdx.x & 63;
  /* keys are loaded 16 rounds at a time (index clamped to the last particle,
   * validity applied at use), so 16 loads per lane are in flight at once */
  if (np > 0 && base < np) {
    for (int r0 = 0; r0 < SORT_PT; r0 += 16) {
      int cv[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const long long p = base + (long long)(r0 + u) * SORT_BLOCK + threadIdx.x;
        cv[u] = cell[p < np ? p : np - 1];
      }
      /* valid lanes are a prefix of the wave, so lane 0 is valid when any
       * lane is; lane 0's key (in cell-ordered data mostly the key of the
       * whole wave) is counted with one ballot, the other lanes add 1 each:
       * a fixed instruction count per round instead of one pass per
       * distinct key */
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const long long p = base + (long long)(r0 + u) * SORT_BLOCK + threadIdx.x;
        const bool valid = p < np;
        const int key = valid ? sort_key(cv[u], kd) : -1;
        if (__ballot(valid) == 0ull) continue;
        const int k0 = __builtin_amdgcn_readlane(key, 0);
        const unsigned long long m0 = __ballot(key == k0);
        if (lane == 0) atomicAdd(&sh[k0], (int)__popcll(m0));
        if (valid && key != k0) atomicAdd(&sh[key], 1);
      }
    }
  }
  __syncthreads();
  int *Hb = H + (size_t)blk * nb1;
  for (int b = threadIdx.x; b < nb1; b += SORT_BLOCK) Hb[b] = sh[b];
}

__global__ __launch_bounds__(SORT_BLOCK) void k_sort_hist(const int32_t *cell,
                                                         const long long *hdr,
                                                         int *H, KeyDesc kd) {
  sort_hist_body(cell, hdr, H, kd, blockIdx.x);
}

/* per (bin, segment of SORT_SEG blocks): exclusive prefix over the blocks of
 * the segment, segment totals into S */
__device__ void sort_colscan_body(int *H, int *S, int nb1, int nblk, int bx, int seg) {
  int b = bx * blockDim.x + threadIdx.x;
  if (b >= nb1) return;
  int run = 0;
  const int k0 = seg * SORT_SEG, n = min(SORT_SEG, nblk - k0);
  /* all loads of the column before its stores (independent, in flight together) */
  int v[SORT_SEG];
#pragma unroll
  for (int k = 0; k < SORT_SEG; ++k) v[k] = k < n ? H[(size_t)(k0 + k) * nb1 + b] : 0;
#pragma unroll
  for (int k = 0; k < SORT_SEG; ++k) {
    if (k < n) H[(size_t)(k0 + k) * nb1 + b] = run;
    run += v[k];
  }
  S[(size_t)seg * nb1 + b] = run;
}

__global__ void k_sort_colscan(int *H, int *S, int nb1, int nblk) {
  sort_colscan_body(H, S, nb1, nblk, blockIdx.x, blockIdx.y);
}

/* per bin: exclusive prefix over segments, bin totals */
__device__ void sort_segscan_body(int *S, int *tot, int nb1, int nseg, int bx) {
  int b = bx * blockDim.x + threadIdx.x;
  if (b >= nb1) return;
  int run = 0;
  /* 32 segments at a time: loads first, then the stores */
  for (int s0 = 0; s0 < nseg; s0 += 32) {
    const int n = min(32, nseg - s0);
    int v[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) v[k] = k < n ? S[(size_t)(s0 + k) * nb1 + b] : 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      if (k < n) S[(size_t)(s0 + k) * nb1 + b] = run;
      run += v[k];
    }
  }
  tot[b] = run;
}

__global__ void k_sort_segscan(int *S, int *tot, int nb1, int nseg) {
  sort_segscan_body(S, tot, nb1, nseg, blockIdx.x);
}

/* exclusive scan of the bin totals (single work-group of 1024 threads) */
__device__ void sort_binscan_body(const int *tot, int *start, int nb1) {
  __shared__ int sh[1024];
  int per = (nb1 + 1023) / 1024;
  int lo = threadIdx.x * per;
  int hi = min(lo + per, nb1);
  int sum = 0;
  for (int b = lo; b < hi; ++b) sum += tot[b];
  sh[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - sum;
  for (int b = lo; b < hi; ++b) {
    start[b] = run;
    run += tot[b];
  }
  if (threadIdx.x == 1023) start[nb1] = sh[1023];
}

__global__ __launch_bounds__(1024) void k_sort_binscan(const int *tot, int *start,
                                                      int nb1) {
  sort_binscan_body(tot, start, nb1);
}

struct ScatterArgs {
  const int32_t *cell;
  const float *ox, *oy, *vx, *vy, *vz;
  const uint32_t *id;
  int32_t *cell_o;
  float *ox_o, *oy_o, *vx_o, *vy_o, *vz_o;
  uint32_t *id_o;
  const long long *hdr;
  const int *H, *S, *start;
  int nblk;
  KeyDesc kd;
};

/* stable scatter: particles of a block are ranked in index order; within a
 * round of 256, lanes of a wave are matched by key with ballots and waves are
 * ordered through LDS group lists. */
struct SortPay {
  int c;
  float ox, oy, vx, vy, vz;
  uint32_t id;
};

/* one particle's record; the index is clamped to the last particle so the
 * loads carry no condition on their results and stay in flight */
__device__ __forceinline__ void scatter_load(const ScatterArgs &a, long long p, long long np,
                                             SortPay &q) {
  const long long pc = p < np ? p : np - 1;
  q.c = a.cell[pc];
  q.ox = a.ox[pc];
  q.oy = a.kd.dim == 2 ? a.oy[pc] : 0.f;
  q.vx = a.vx[pc];
  q.vy = a.vy[pc];
  q.vz = a.vz[pc];
  q.id = a.id ? a.id[pc] : 0u;
}

/* The group lists are double-buffered by round parity, so a round needs two
 * barriers: after the lists are written (then base is read), and after base
 * has been read (then it is advanced); the next round's first barrier orders
 * the advance before the next reads. */
__device__ __forceinline__ void scatter_round(const ScatterArgs &a, const SortPay &q, bool valid,
                                              int *base, int (*gkey)[64], int (*gcnt)[64],
                                              int *ng, int lane, int w) {
  const int key = valid ? sort_key(q.c, a.kd) : -1;
  unsigned long long todo = __ballot(valid);
  int myrank = 0;
  int ngr = 0;
  while (todo) {
    int leader = __ffsll((long long)todo) - 1;
    int lk = __shfl(key, leader);
    unsigned long long m = __ballot(valid && key == lk);
    if (valid && key == lk) myrank = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) {
      gkey[w][ngr] = lk;
      gcnt[w][ngr] = __popcll(m);
    }
    ++ngr;
    todo &= ~m;
  }
  if (lane == 0) ng[w] = ngr;
  isk_lds_barrier();
  if (valid && key < a.kd.nbins) {
    int pre = 0;
    for (int w2 = 0; w2 < w; ++w2) {
      int n2 = ng[w2];
      for (int g = 0; g < n2; ++g)
        if (gkey[w2][g] == key) pre += gcnt[w2][g];
    }
    long long d = (long long)base[key] + pre + myrank;
    a.cell_o[d] = q.c;
    a.ox_o[d] = q.ox;
    if (a.kd.dim == 2) a.oy_o[d] = q.oy;
    a.vx_o[d] = q.vx;
    a.vy_o[d] = q.vy;
    a.vz_o[d] = q.vz;
    if (a.id) a.id_o[d] = q.id;
  }
  isk_lds_barrier();
  if (lane < ng[w]) atomicAdd(&base[gkey[w][lane]], gcnt[w][lane]);
}

static inline size_t sort_scatter_shm(int nb1) { return (size_t)nb1 * sizeof(int); }

__device__ void sort_scatter_body(const ScatterArgs &a, const int blk) {
  extern __shared__ int base[];
  __shared__ int gkey[2][4][64];
  __shared__ int gcnt[2][4][64];
  __shared__ int ng[2][4];
  const int nb1 = a.kd.nbins + 1;
  const int seg = blk / SORT_SEG;
  for (int b = threadIdx.x; b < nb1; b += SORT_BLOCK)
    base[b] = a.start[b] + a.S[(size_t)seg * nb1 + b] + a.H[(size_t)blk * nb1 + b];
  __syncthreads();
  const long long np = a.hdr[HDR_NP];
  const long long b0 = (long long)blk * SORT_CHUNK;
  if (np <= 0 || b0 >= np) return;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  /* SORT_DEPTH records in registers: the loads of round r + SORT_DEPTH are
   * issued as soon as round r is written, so SORT_DEPTH rounds of loads are
   * in flight per work-group; round parity (SORT_DEPTH and SORT_PT are even)
   * picks the group-list buffer */
  SortPay R[SORT_DEPTH];
#pragma unroll
  for (int u = 0; u < SORT_DEPTH; ++u)
    scatter_load(a, b0 + (long long)u * SORT_BLOCK + threadIdx.x, np, R[u]);
  for (int r = 0; r < SORT_PT; r += SORT_DEPTH) {
#pragma unroll
    for (int u = 0; u < SORT_DEPTH; ++u) {
      const long long pr = b0 + (long long)(r + u) * SORT_BLOCK + threadIdx.x;
      scatter_round(a, R[u], pr < np, base, gkey[u & 1], gcnt[u & 1], ng[u & 1], lane, w);
      if (r + u + SORT_DEPTH < SORT_PT)
        scatter_load(a, pr + (long long)SORT_DEPTH * SORT_BLOCK, np, R[u]);
    }
  }
}

__global__ __launch_bounds__(SORT_BLOCK) void k_sort_scatter(ScatterArgs a) {
  sort_scatter_body(a, blockIdx.x);
}

/* new particle count; with build != 0 also the chunk table of the tile sort */
__device__ void sort_finish_body(long long *hdr, const int *tot, const int *start, int nbins,
                                 int build, int CH, IskChunk *chunks, int *tcb) {
  __shared__ int sh[1024];
  long long live = start[nbins];
  if (!build) {
    if (threadIdx.x == 0) {
      hdr[HDR_NP] = live;
      hdr[HDR_NP_SORTED] = 0;
      hdr[HDR_NCHUNKS] = 0;
    }
    return;
  }
  int per = (nbins + 1023) / 1024;
  int lo = threadIdx.x * per;
  int hi = min(lo + per, nbins);
  int sum = 0;
  for (int t = lo; t < hi; ++t) sum += (tot[t] + CH - 1) / CH;
  sh[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - sum;
  for (int t = lo; t < hi; ++t) {
    tcb[t] = run;
    int n = tot[t];
    int st = start[t];
    for (int k = 0; k * CH < n; ++k) {
      IskChunk ch;
      ch.tile = t;
      ch.start = st + k * CH;
      ch.count = min(CH, n - k * CH);
      ch.pad = 0;
      chunks[run++] = ch;
    }
  }
  if (threadIdx.x == 1023) {
    tcb[nbins] = sh[1023];
    hdr[HDR_NP] = live;
    hdr[HDR_NP_SORTED] = live;
    hdr[HDR_NCHUNKS] = sh[1023];
  }
}

__global__ __launch_bounds__(1024) void k_sort_finish(
    long long *hdr, const int *tot, const int *start, int nbins, int build,
    int CH, IskChunk *chunks, int *tcb) {
  sort_finish_body(hdr, tot, start, nbins, build, CH, chunks, tcb);
}

/* everything one sort pass of one species needs (single or batched launch) */
struct SortArgs {
  KeyDesc kd;
  const int32_t *cell;
  const long long *hdr;
  int *H, *S, *tot, *start;
  int nb1, nblk, nseg;
  ScatterArgs sc;
  long long *hdr_w;
  int build, CH;
  IskChunk *chunks;
  int *tcb;
};

/* scratch sizing and arguments of a pass reading buffer `b` */
static int sort_prepare(iskra_ctx *c, int sid, int mode, int build, int b, SortArgs *A) {
  IskSpecies *s = &c->sp[sid];
  KeyDesc kd;
  kd.mode = mode;
  kd.dim = c->g.dim;
  kd.T = c->t.T;
  kd.log2T = c->t.log2T;
  kd.ntx = c->ntx;
  kd.nbins = mode == 0 ? c->ntiles : (c->g.dim == 2 ? c->t.T * c->t.T : c->t.T);
  int nb1 = kd.nbins + 1;
  if (nb1 > ISK_MAX_BINS + 1) {
    isk_set_error("sort: %d bins exceed %d", kd.nbins, ISK_MAX_BINS);
    return ISKRA_E_UNSUPPORTED;
  }
  int nblk = (int)((s->cap + SORT_CHUNK - 1) / SORT_CHUNK);
  if (nblk < 1) nblk = 1;
  int nseg = (nblk + SORT_SEG - 1) / SORT_SEG;
  int64_t need_H = (int64_t)nblk * nb1, need_S = (int64_t)nseg * nb1;
  if (need_H > c->sort_H_words) {
    isk_free(c->sort_H);
    c->sort_H = nullptr;
    if (isk_alloc((void **)&c->sort_H, need_H * 4)) return ISKRA_E_HIP;
    c->sort_H_words = need_H;
  }
  if (need_S > c->sort_S_words) {
    isk_free(c->sort_S);
    isk_free(c->sort_tot);
    isk_free(c->sort_start);
    c->sort_S = c->sort_tot = c->sort_start = nullptr;
    if (isk_alloc((void **)&c->sort_S, need_S * 4)) return ISKRA_E_HIP;
    /* bin arrays sized for the largest bin count any pass can use */
    if (isk_alloc((void **)&c->sort_tot, (int64_t)(ISK_MAX_BINS + 2) * 4)) return ISKRA_E_HIP;
    if (isk_alloc((void **)&c->sort_start, (int64_t)(ISK_MAX_BINS + 2) * 4)) return ISKRA_E_HIP;
    c->sort_S_words = need_S;
  }
  int o = 1 - b;
  A->kd = kd;
  A->cell = s->cell[b];
  A->hdr = s->hdr;
  A->H = c->sort_H;
  A->S = c->sort_S;
  A->tot = c->sort_tot;
  A->start = c->sort_start;
  A->nb1 = nb1;
  A->nblk = nblk;
  A->nseg = nseg;
  ScatterArgs &a = A->sc;
  a.cell = s->cell[b];
  a.ox = s->ox[b];
  a.oy = s->oy[b];
  a.vx = s->vx[b];
  a.vy = s->vy[b];
  a.vz = s->vz[b];
  a.id = s->track_ids ? s->id[b] : nullptr;
  a.cell_o = s->cell[o];
  a.ox_o = s->ox[o];
  a.oy_o = s->oy[o];
  a.vx_o = s->vx[o];
  a.vy_o = s->vy[o];
  a.vz_o = s->vz[o];
  a.id_o = s->track_ids ? s->id[o] : nullptr;
  a.hdr = s->hdr;
  a.H = c->sort_H;
  a.S = c->sort_S;
  a.start = c->sort_start;
  a.nblk = nblk;
  a.kd = kd;
  A->hdr_w = s->hdr;
  A->build = build;
  A->CH = c->t.CH;
  A->chunks = s->chunks;
  A->tcb = s->tile_chunk_begin;
  return ISKRA_OK;
}

static int sort_pass(iskra_ctx *c, int sid, int mode, int build) {
  IskSpecies *s = &c->sp[sid];
  SortArgs A;
  int rc = sort_prepare(c, sid, mode, build, s->cur, &A);
  if (rc) return rc;
  const int nb1 = A.nb1, nblk = A.nblk, nseg = A.nseg;
  size_t shm = (size_t)nb1 * sizeof(int);
  hipLaunchKernelGGL(k_sort_hist, dim3(nblk), dim3(SORT_BLOCK), shm, c->stream,
                     A.cell, A.hdr, A.H, A.kd);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sort_colscan, dim3((nb1 + 255) / 256, nseg), dim3(256), 0,
                     c->stream, A.H, A.S, nb1, nblk);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sort_segscan, dim3((nb1 + 255) / 256), dim3(256), 0,
                     c->stream, A.S, A.tot, nb1, nseg);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sort_binscan, dim3(1), dim3(1024), 0, c->stream,
                     (const int *)A.tot, A.start, nb1);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sort_scatter, dim3(nblk), dim3(SORT_BLOCK), sort_scatter_shm(nb1),
                     c->stream, A.sc);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sort_finish, dim3(1), dim3(1024), 0, c->stream, A.hdr_w,
                     (const int *)A.tot, (const int *)A.start, A.kd.nbins, build, A.CH,
                     A.chunks, A.tcb);
  ISK_CHECK_LAUNCH();
  s->cur = 1 - s->cur;
  return ISKRA_OK;
}

int isk_sort_species(iskra_ctx *c, int sid, int exact_cell) {
  int rc;
  hipEvent_t ev = nullptr;
  if ((rc = isk_prof_begin(c, &ev))) return rc;
  if (exact_cell) {
    if ((rc = sort_pass(c, sid, 1, 0))) return rc;
  }
  if ((rc = sort_pass(c, sid, 0, 1))) return rc;
  return isk_prof_end(c, ev, ISKRA_K_SORT, 0);
}

/* ------------------------------------------------------------------------ */
/* fused gather + push + boundary + deposit                                  */
/* ------------------------------------------------------------------------ */
struct PushArgs {
  int32_t *cell;
  float *ox, *oy, *vx, *vy;
  float *vz;     /* RZ pusher: v_theta, rotated with v_r */
  double dtdr;   /* RZ pusher: dt / dr */
  /* surface tracker (PM_TRACK): edge codes per cell and direction, tracked
   * cells, per-electrode hit counters of this species, dt and the spacing
   * the reference's check uses */
  const uint8_t *surf, *tracked;
  unsigned long long *ehit;
  double dt_d, dh_d;
  long long *hdr;
  const IskChunk *chunks;
  long long *partial;
  unsigned long long *gfall;
  const float *E;
  int nx, ny, ncx, ncy;
  int T, log2T, M, W, CH, ntx;
  float qmdt, dtdx, dtdy;
  int bndx, bndy;
  int do_push;
  int *live_cnt; /* per work-group live particles pushed */
};

/* rn(a 2^32) and rn(a b 2^32) for offsets a, b in [0, 1): the FP64
 * products are exact (24 x 24 bit mantissas), rint rounds half to even, and
 * the results are below 2^32, so they convert to uint32 exactly */
__device__ __forceinline__ long long fix32_dev(float a) {
  return (long long)(unsigned int)__builtin_rint((double)a * 4294967296.0);
}
__device__ __forceinline__ long long fixprod32_dev(float a, float b) {
  return (long long)(unsigned int)__builtin_rint(((double)a * (double)b) * 4294967296.0);
}

/* CIC weights in 2^-32 units: Wx = rn(hx 2^32), Wy = rn(hy 2^32), Wxy =
 * rn(hx hy 2^32); the four corners sum to exactly 2^32 */
__device__ __forceinline__ void fix_weights2(float hx, float hy, long long *w00,
                                             long long *w10, long long *w01,
                                             long long *w11) {
  long long Wx = fix32_dev(hx);
  long long Wy = fix32_dev(hy);
  long long Wxy = fixprod32_dev(hx, hy);
  *w11 = Wxy;
  *w10 = Wx - Wxy;
  *w01 = Wy - Wxy;
  *w00 = (1LL << 32) - Wx - Wy + Wxy;
}

/* one step of a coordinate: o += v*dtd, split off the integer part */
__device__ __forceinline__ int advance_coord(float v, float dtd, float *o) {
  float h = fmaf(v, dtd, *o);
  float f = floorf(h);
  h = h - f;
  if (h >= 1.0f) {
    h = 0.0f;
    f += 1.0f;
  }
  *o = h;
  return (int)f;
}

/* transform_from_cartesian_to_cylindrical! (pushers.jl:52-66) in cell units,
 * after the Cartesian step of the radial coordinate: the particle sits at
 * X = i + hx + v_r dt/dr in the plane through its old position and the axis
 * and has moved Y = v_theta dt/dr out of it; it is rotated back to radius
 * R = sqrt(X^2 + Y^2) with sin = Y/R (0 when R = 0, the reference's
 * r .~ 0 test) and cos = +sqrt(1 - sin^2) (also for X < 0, as the reference
 * takes it), and (v_r, v_theta) rotated by the same angle.  FP64 throughout
 * (correctly rounded sqrt and division); R beyond 2^30 cells maps to a cell
 * index no grid has. */
__device__ __forceinline__ void rz_transform(int *i, float *hx, float *vr, float *vt,
                                             double dtdr) {
  const double X = (double)*i + ((double)*hx + (double)*vr * dtdr);
  const double Y = (double)*vt * dtdr;
  const double R = sqrt(X * X + Y * Y);
  const double sn = R == 0.0 ? 0.0 : Y / R;
  const double cs = sqrt(1.0 - sn * sn);
  const double nvr = cs * (double)*vr + sn * (double)*vt;
  const double nvt = cs * (double)*vt - sn * (double)*vr;
  int ii;
  float h;
  if (R < 1073741824.0) {
    const double fl = floor(R);
    ii = (int)fl;
    h = (float)(R - fl);
    if (h >= 1.0f) {
      h = 0.0f;
      ++ii;
    }
  } else {
    ii = 1 << 30;
    h = 0.0f;
  }
  *i = ii;
  *hx = h;
  *vr = (float)nvr;
  *vt = (float)nvt;
}

/* pusher modes of the push kernels */
enum { PM_XY = 0, PM_RZ = 1, PM_TRACK = 2 };

/* track! + check! + hit! of the reference's surface tracker
 * (pic/surfaces/track.jl:152-161, check.jl:17-68, hit.jl:1-56,
 * circuit_coupling.jl:47-62) for one particle that starts in a tracked cell:
 * from its offsets (hx, hy) before the push it follows the straight path
 * with the pushed velocity (vx, vy) cell by cell for dt, in FP64 and in the
 * reference's operation order (one spacing dh for both axes, as check uses
 * st.dh).  At each edge crossed: no surface -> keep tracing from the next
 * cell; absorbing surface or fixed electrode -> absorbed; floating electrode
 * k -> absorbed and counted on electrode k; reflective surface -> the
 * velocity components along the edge normal flip, the position is corrected
 * by the reflected rest of the path (the reference's px -= v dt' ... px +=
 * v' dt'), and tracing resumes in the cell the particle came from
 * (scattered!); any other surface ends the tracing.  Returns 1 if absorbed;
 * *cx, *cy receive the position corrections in cells. */
#define ISK_TRACK_MAX 64
__device__ __forceinline__ int track_path(const PushArgs &a, int i, int j, float hx0, float hy0,
                                          float *vx, float *vy, double *cx, double *cy) {
  const double h = a.dh_d;
  double T = a.dt_d, hx = hx0, hy = hy0;
  *cx = 0.0;
  *cy = 0.0;
  for (int it = 0; it < ISK_TRACK_MAX; ++it) {
    const double vxd = *vx, vyd = *vy;
    const double dxs = vxd > 0 ? h * (1.0 - hx) : h * hx;
    const double dys = vyd > 0 ? h * (1.0 - hy) : h * hy;
    const double tx = dxs / fabs(vxd), ty = dys / fabs(vyd);
    if (T < tx && T < ty) return 0; /* stays in this cell */
    int ni = i, nj = j, dir;
    double nhx, nhy, Tn;
    if (tx < ty) {
      Tn = T - tx;
      nhy = hy + vyd * tx / h;
      if (vxd > 0) { ni = i + 1; nhx = 0.0; dir = 1; }
      else { ni = i - 1; nhx = 1.0; dir = 0; }
    } else {
      Tn = T - ty;
      nhx = hx + vxd * ty / h;
      if (vyd > 0) { nj = j + 1; nhy = 0.0; dir = 3; }
      else { nj = j - 1; nhy = 1.0; dir = 2; }
    }
    if (i < 0 || i >= a.ncx || j < 0 || j >= a.ncy) return 0; /* left the grid untracked */
    const int code = a.surf[(i + a.ncx * j) * 4 + dir];
    if (code == ISK_SURF_NONE) {
      i = ni;
      j = nj;
      hx = nhx;
      hy = nhy;
      T = Tn;
      continue;
    }
    if (code == ISK_SURF_ABSORB) return 1;
    if (code >= ISK_SURF_ELECTRODE) {
      atomicAdd(&a.ehit[code - ISK_SURF_ELECTRODE], 1ull);
      return 1;
    }
    if (code != ISK_SURF_REFLECT) return 0; /* hit!(::Surface) does nothing */
    if (ni != i) {
      *vx = -*vx;
      *cx += 2.0 * (double)*vx * Tn / h;
    }
    if (nj != j) {
      *vy = -*vy;
      *cy += 2.0 * (double)*vy * Tn / h;
    }
    /* scattered! (hit.jl:13-21) */
    int si = ni, sj = nj;
    double shx = nhx, shy = nhy;
    if (nhx == 0.0) { si = ni - 1; shx = 1.0; }
    if (nhx == 1.0) { si = ni + 1; shx = 0.0; }
    if (nhy == 0.0) { sj = nj - 1; shy = 1.0; }
    if (nhy == 1.0) { sj = nj + 1; shy = 0.0; }
    i = si;
    j = sj;
    hx = shx;
    hy = shy;
    T = Tn;
  }
  return 0;
}

/* a position correction of c cells added to (cell, offset), FP64 */
__device__ __forceinline__ int shift_coord(double c, float *o) {
  const double t = (double)*o + c;
  const double f = floor(t);
  float h = (float)(t - f);
  int k = (int)f;
  if (h >= 1.0f) {
    h = 0.0f;
    ++k;
  }
  *o = h;
  return k;
}

/* boundary on one axis; returns false if the particle is discarded */
__device__ __forceinline__ bool apply_bnd(int *i, int n, int mode) {
  if (*i >= 0 && *i < n) return true;
  if (mode == ISKRA_BND_WRAP) {
    int r = *i % n;
    if (r < 0) r += n;
    *i = r;
    return true;
  }
  return false; /* discard (NONE also removes particles that leave the grid) */
}

template <int DIM>
struct PushRegs {
  static constexpr int U = PushShape<DIM>::unroll;
  int32_t c[U];
  float hx[U], hy[U], vx[U], vy[U];
};

/* pointers are the work-group's first particle (uniform), indices 32-bit */
template <int DIM>
__device__ __forceinline__ void push_load(PushRegs<DIM> &r, const int32_t *__restrict__ cb,
                                          const float *__restrict__ oxb,
                                          const float *__restrict__ oyb,
                                          const float *__restrict__ vxb,
                                          const float *__restrict__ vyb, unsigned count,
                                          unsigned b0) {
  /* branch-free: slots past the end load the first particle (the caller
   * skips them); nothing here consumes a loaded value, so all loads of the
   * batch are in flight together */
#pragma unroll
  for (int u = 0; u < PushShape<DIM>::unroll; ++u) {
    const unsigned k = b0 + u * PushShape<DIM>::block + threadIdx.x;
    __builtin_assume(k < (1u << 24));
    const unsigned q = k < count ? k : 0u;
    r.c[u] = cb[q];
    r.hx[u] = oxb[q];
    r.vx[u] = vxb[q];
    r.hy[u] = DIM == 2 ? oyb[q] : 0.f;
    r.vy[u] = DIM == 2 ? vyb[q] : 0.f;
  }
}

template <int DIM, int PM = PM_XY>
__device__ __forceinline__ void push_body(const PushArgs &a, const int blk) {
  constexpr bool RZ = PM == PM_RZ, TRACK = PM == PM_TRACK;
  constexpr int BLK = PushShape<DIM>::block, UNR = PushShape<DIM>::unroll;
  extern __shared__ unsigned long long sD[]; /* WW deposit counters, then E */
  const int W = a.W;
  const int WW = DIM == 2 ? W * W : W;
  float *sE = reinterpret_cast<float *>(sD + WW);
  int32_t *__restrict__ cellp = a.cell;
  float *__restrict__ oxp = a.ox;
  float *__restrict__ oyp = a.oy;
  float *__restrict__ vxp = a.vx;
  float *__restrict__ vyp = a.vy;
  const long long nch = a.hdr[HDR_NCHUNKS];
  const long long np = a.hdr[HDR_NP];
  const long long nps = a.hdr[HDR_NP_SORTED];
  long long start, count;
  bool windowed;
  int wx0 = 0, wy0 = 0;
  if ((long long)blk < nch) {
    IskChunk ch = a.chunks[blk];
    start = ch.start;
    count = ch.count;
    windowed = true;
    if (DIM == 2) {
      int ti = ch.tile % a.ntx, tj = ch.tile / a.ntx;
      wx0 = ti * a.T - a.M;
      wy0 = tj * a.T - a.M;
    } else {
      wx0 = ch.tile * a.T - a.M;
    }
  } else {
    start = nps + ((long long)blk - nch) * a.CH;
    if (start >= np) {
      if (threadIdx.x == 0) a.live_cnt[blk] = 0;
      return;
    }
    count = min((long long)a.CH, np - start);
    windowed = false;
  }
  /* the first batch of particles is requested before the E window, so the
   * two memory latencies overlap instead of following each other */
  int32_t *__restrict__ cb = cellp + start;
  float *__restrict__ oxb = oxp + start;
  float *__restrict__ oyb = DIM == 2 ? oyp + start : oyp;
  float *__restrict__ vxb = vxp + start;
  float *__restrict__ vyb = DIM == 2 ? vyp + start : vyp;
  const unsigned ucount = (unsigned)count;
  PushRegs<DIM> cur;
  push_load<DIM>(cur, cb, oxb, oyb, vxb, vyb, ucount, 0u);
  if (windowed) {
    for (int k = threadIdx.x; k < WW; k += BLK) {
      sD[k] = 0ull;
      if (a.do_push) {
        int li = DIM == 2 ? k % W : k;
        int lj = DIM == 2 ? k / W : 0;
        int gi = wx0 + li, gj = wy0 + lj;
        bool in = gi >= 0 && gi < a.nx && (DIM == 1 || (gj >= 0 && gj < a.ny));
        if (DIM == 2) {
          float2 e = in ? reinterpret_cast<const float2 *>(a.E)[gj * a.nx + gi]
                        : make_float2(0.f, 0.f);
          sE[2 * k] = e.x;
          sE[2 * k + 1] = e.y;
        } else {
          sE[k] = in ? a.E[gi] : 0.f;
        }
      }
    }
    __syncthreads();
  }
  int live = 0;
  const long long per = (long long)BLK * UNR;
  for (long long b0 = 0; b0 < count; b0 += per) {
    /* a batch of UNR particles per thread: all loads, then compute + stores */
    if (b0 > 0) push_load<DIM>(cur, cb, oxb, oyb, vxb, vyb, ucount, (unsigned)b0);
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const unsigned kslot = (unsigned)b0 + u * BLK + threadIdx.x;
      if (kslot >= ucount) continue;
      int32_t c = cur.c[u];
      if (c < 0) continue;
      int i = c & 0xFFFF;
      int j = c >> 16;
      float hx = cur.hx[u], hy = cur.hy[u], vx = cur.vx[u], vy = cur.vy[u];
      float vt = 0.f;
      bool alive = true;
      if (a.do_push) {
        /* gather (cloud_in_cell.jl:20-36) */
        float ax = 1.0f - hx;
        float Ex, Ey = 0.f;
        if (DIM == 2) {
          float ay = 1.0f - hy;
          float w00 = ax * ay, w10 = hx * ay, w01 = ax * hy, w11 = hx * hy;
          float2 e00, e10, e01, e11;
          int li = i - wx0, lj = j - wy0;
          if (windowed && (unsigned)li < (unsigned)(W - 1) && (unsigned)lj < (unsigned)(W - 1)) {
            int q = lj * W + li;
            e00 = make_float2(sE[2 * q], sE[2 * q + 1]);
            e10 = make_float2(sE[2 * q + 2], sE[2 * q + 3]);
            e01 = make_float2(sE[2 * (q + W)], sE[2 * (q + W) + 1]);
            e11 = make_float2(sE[2 * (q + W) + 2], sE[2 * (q + W) + 3]);
          } else {
            const float2 *E2 = reinterpret_cast<const float2 *>(a.E);
            int q = j * a.nx + i;
            e00 = E2[q];
            e10 = E2[q + 1];
            e01 = E2[q + a.nx];
            e11 = E2[q + a.nx + 1];
          }
          Ex = fmaf(w11, e11.x, fmaf(w01, e01.x, fmaf(w10, e10.x, w00 * e00.x)));
          Ey = fmaf(w11, e11.y, fmaf(w01, e01.y, fmaf(w10, e10.y, w00 * e00.y)));
        } else {
          float e0, e1;
          int li = i - wx0;
          if (windowed && (unsigned)li < (unsigned)(W - 1)) {
            e0 = sE[li];
            e1 = sE[li + 1];
          } else {
            e0 = a.E[i];
            e1 = a.E[i + 1];
          }
          Ex = fmaf(hx, e1, ax * e0);
        }
        /* push with B = 0 (pushers.jl:37-50): v += (q/m) E dt, x += v dt */
        vx = fmaf(a.qmdt, Ex, vx);
        if (TRACK && DIM == 2) {
          /* track! before the push, check! after it (ParticleInCell.jl:55-60) */
          vy = fmaf(a.qmdt, Ey, vy);
          const float vx0 = vx, vy0 = vy;
          double cx = 0.0, cy = 0.0;
          int absorbed = 0;
          if (a.tracked[i + a.ncx * j])
            absorbed = track_path(a, i, j, hx, hy, &vx, &vy, &cx, &cy);
          i += advance_coord(vx0, a.dtdx, &hx);
          j += advance_coord(vy0, a.dtdy, &hy);
          if (cx != 0.0) i += shift_coord(cx, &hx);
          if (cy != 0.0) j += shift_coord(cy, &hy);
          alive = apply_bnd(&i, a.ncx, a.bndx);
          alive = apply_bnd(&j, a.ncy, a.bndy) && alive;
          if (absorbed) alive = false;
        } else {
          if (RZ) {
            vt = a.vz[start + kslot];
            rz_transform(&i, &hx, &vx, &vt, a.dtdr);
          } else {
            i += advance_coord(vx, a.dtdx, &hx);
          }
          alive = apply_bnd(&i, a.ncx, a.bndx);
          if (DIM == 2) {
            vy = fmaf(a.qmdt, Ey, vy);
            j += advance_coord(vy, a.dtdy, &hy);
            alive = apply_bnd(&j, a.ncy, a.bndy) && alive;
          }
        }
      }
      if (alive) {
        ++live;
        /* deposit (cloud_in_cell.jl:1-18) */
        if (DIM == 2) {
          long long w00, w10, w01, w11;
          fix_weights2(hx, hy, &w00, &w10, &w01, &w11);
          int li = i - wx0, lj = j - wy0;
          if (windowed && (unsigned)li < (unsigned)(W - 1) && (unsigned)lj < (unsigned)(W - 1)) {
            int q = lj * W + li;
            atomicAdd(&sD[q], (unsigned long long)w00);
            atomicAdd(&sD[q + 1], (unsigned long long)w10);
            atomicAdd(&sD[q + W], (unsigned long long)w01);
            atomicAdd(&sD[q + W + 1], (unsigned long long)w11);
          } else {
            int q = j * a.nx + i;
            atomicAdd(&a.gfall[q], (unsigned long long)w00);
            atomicAdd(&a.gfall[q + 1], (unsigned long long)w10);
            atomicAdd(&a.gfall[q + a.nx], (unsigned long long)w01);
            atomicAdd(&a.gfall[q + a.nx + 1], (unsigned long long)w11);
          }
        } else {
          long long w1 = fix32_dev(hx);
          long long w0 = (1LL << 32) - w1;
          int li = i - wx0;
          if (windowed && (unsigned)li < (unsigned)(W - 1)) {
            atomicAdd(&sD[li], (unsigned long long)w0);
            atomicAdd(&sD[li + 1], (unsigned long long)w1);
          } else {
            atomicAdd(&a.gfall[i], (unsigned long long)w0);
            atomicAdd(&a.gfall[i + 1], (unsigned long long)w1);
          }
        }
      }
      if (a.do_push) {
        const unsigned k = (unsigned)b0 + u * BLK + threadIdx.x;
        /* the cell is stored only when it changed: slow particles (the ions)
         * then leave whole sectors of the cell array unwritten */
        const int32_t nc = alive ? (i | (j << 16)) : -1;
        if (nc != c) cb[k] = nc;
        oxb[k] = hx;
        vxb[k] = vx;
        if (DIM == 2) {
          oyb[k] = hy;
          vyb[k] = vy;
        }
        if (RZ) a.vz[start + k] = vt;
      }
    }
  }
  /* per-work-group live count (summed by the merge kernel, no atomics) */
  __shared__ int wl[BLK / 64];
  for (int off = 32; off > 0; off >>= 1) live += __shfl_xor(live, off);
  if ((threadIdx.x & 63) == 0) wl[threadIdx.x >> 6] = live;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < BLK / 64; ++k) t += wl[k];
    a.live_cnt[blk] = t;
  }
  if (windowed) {
    long long *dst = a.partial + (size_t)blk * WW;
    for (int k = threadIdx.x; k < WW; k += BLK) dst[k] = (long long)sD[k];
  }
}

template <int DIM, int PM = PM_XY>
__global__ __launch_bounds__(PushShape<DIM>::block) void k_push(PushArgs a) {
  push_body<DIM, PM>(a, blockIdx.x);
}

/* arguments of one push launch reading buffer `b`; grid and LDS size */
static void push_prepare(iskra_ctx *c, int sid, double dt, int do_push, int b, PushArgs *pa,
                         int *grid, size_t *shm) {
  IskSpecies *s = &c->sp[sid];
  const IskGrid &g = c->g;
  PushArgs &a = *pa;
  a.cell = s->cell[b];
  a.ox = s->ox[b];
  a.oy = s->oy[b];
  a.vx = s->vx[b];
  a.vy = s->vy[b];
  a.vz = s->vz[b];
  a.dtdr = dt / g.dx;
  a.surf = c->trk.surf;
  a.tracked = c->trk.tracked;
  a.ehit = s->ehit;
  a.dt_d = dt;
  a.dh_d = g.dx;
  a.hdr = s->hdr;
  a.chunks = s->chunks;
  a.partial = s->partial;
  a.gfall = s->gfall;
  a.E = c->E;
  a.nx = g.nx;
  a.ny = g.ny;
  a.ncx = g.ncx;
  a.ncy = g.ncy;
  a.T = c->t.T;
  a.log2T = c->t.log2T;
  a.M = isk_sp_margin(c, sid);
  a.W = isk_sp_window(c, sid);
  a.CH = c->t.CH;
  a.ntx = c->ntx;
  a.qmdt = (float)(s->q / s->m * dt);
  a.dtdx = (float)(dt / g.dx);
  a.dtdy = (float)(dt / g.dy);
  a.bndx = s->bnd[0];
  a.bndy = s->bnd[1];
  a.do_push = do_push;
  a.live_cnt = s->live_cnt;
  *grid = (int)s->max_chunks;
  int WW = g.dim == 2 ? a.W * a.W : a.W;
  *shm = (size_t)WW * (8 + 4 * g.dim);
}

/* the push (or deposit-only) launch of one species; fill != 0 clears the
 * fallback grid first (every merge clears it after reading, so a push that
 * follows a merge finds it zero) */
static int push_launch(iskra_ctx *c, int sid, double dt, int do_push, int fill) {
  IskSpecies *s = &c->sp[sid];
  int rc;
  if (fill && (rc = isk_fill_i64(c, s->gfall, isk_nodes(c), 0))) return rc;
  PushArgs a;
  int grid;
  size_t shm;
  push_prepare(c, sid, dt, do_push, s->cur, &a, &grid, &shm);
  hipEvent_t ev = nullptr;
  if (do_push) {
    int rc2 = isk_prof_begin(c, &ev);
    if (rc2) return rc2;
  }
  if (c->g.dim == 2 && c->pusher == ISKRA_PUSHER_RZ && do_push)
    hipLaunchKernelGGL((k_push<2, PM_RZ>), dim3(grid), dim3(PushShape<2>::block), shm, c->stream,
                       a);
  else if (c->g.dim == 2 && c->trk.on && do_push)
    hipLaunchKernelGGL((k_push<2, PM_TRACK>), dim3(grid), dim3(PushShape<2>::block), shm,
                       c->stream, a);
  else if (c->g.dim == 2)
    hipLaunchKernelGGL(k_push<2>, dim3(grid), dim3(PushShape<2>::block), shm, c->stream, a);
  else
    hipLaunchKernelGGL(k_push<1>, dim3(grid), dim3(PushShape<1>::block), shm, c->stream, a);
  ISK_CHECK_LAUNCH();
  if (do_push) {
    int rc2 = isk_prof_end(c, ev, ISKRA_K_PUSH, 0);
    if (rc2) return rc2;
  }
  return ISKRA_OK;
}

/* the pushes of all species in one launch: work-group g belongs to species
 * s with bstart[s] <= g < bstart[s+1] and acts as work-group g - bstart[s]
 * of that species' own launch (one launch boundary less per step, and the
 * tail of one species overlaps the next) */
struct PushMulti {
  PushArgs a[ISK_MAX_SPECIES];
  int bstart[ISK_MAX_SPECIES + 1];
  int n;
};

template <int DIM, int PM = PM_XY>
__global__ __launch_bounds__(PushShape<DIM>::block) void k_push_multi(PushMulti m) {
  int s = 0;
  while (s + 1 < m.n && (int)blockIdx.x >= m.bstart[s + 1]) ++s;
  /* a copy in registers: read through a reference into the argument block,
   * the fields would be re-read from memory after every particle's stores */
  const PushArgs a = m.a[s];
  push_body<DIM, PM>(a, blockIdx.x - m.bstart[s]);
}

static int push_launch_all(iskra_ctx *c, double dt) {
  if (c->nspecies == 0) return ISKRA_OK;
  PushMulti m;
  size_t shm = 0;
  m.n = c->nspecies;
  m.bstart[0] = 0;
  for (int s = 0; s < c->nspecies; ++s) {
    int grid;
    size_t sh;
    push_prepare(c, s, dt, 1, c->sp[s].cur, &m.a[s], &grid, &sh);
    m.bstart[s + 1] = m.bstart[s] + grid;
    shm = std::max(shm, sh);
  }
  if (m.bstart[c->nspecies] == 0) return ISKRA_OK;
  hipEvent_t ev = nullptr;
  int rc = isk_prof_begin(c, &ev);
  if (rc) return rc;
  if (c->g.dim == 2 && c->pusher == ISKRA_PUSHER_RZ)
    hipLaunchKernelGGL((k_push_multi<2, PM_RZ>), dim3(m.bstart[m.n]), dim3(PushShape<2>::block),
                       shm, c->stream, m);
  else if (c->g.dim == 2 && c->trk.on)
    hipLaunchKernelGGL((k_push_multi<2, PM_TRACK>), dim3(m.bstart[m.n]),
                       dim3(PushShape<2>::block), shm, c->stream, m);
  else if (c->g.dim == 2)
    hipLaunchKernelGGL(k_push_multi<2>, dim3(m.bstart[m.n]), dim3(PushShape<2>::block), shm,
                       c->stream, m);
  else
    hipLaunchKernelGGL(k_push_multi<1>, dim3(m.bstart[m.n]), dim3(PushShape<1>::block), shm,
                       c->stream, m);
  ISK_CHECK_LAUNCH();
  return isk_prof_end(c, ev, ISKRA_K_PUSH, 0);
}

int isk_push_species(iskra_ctx *c, int sid, double dt, int do_push) {
  int rc = push_launch(c, sid, dt, do_push, 1);
  if (rc) return rc;
  c->sp[sid].deposited = 1;
  hipEvent_t em = nullptr;
  if ((rc = isk_prof_begin(c, &em))) return rc;
  if ((rc = isk_merge_species(c, sid, do_push))) return rc;
  return isk_prof_end(c, em, ISKRA_K_MERGE, 0);
}

/* ------------------------------------------------------------------------ */
/* deterministic merge of the chunk windows                                  */
/* ------------------------------------------------------------------------ */
struct MergeArgs {
  const long long *tilewin; /* 2-D: per-tile sums of the chunk windows (null: per chunk) */
  const int *live_cnt;
  int nlive; /* entries of live_cnt to add to HDR_PUSHED (0: deposit only) */
  long long *hdr_rw;
  const long long *partial;
  unsigned long long *gfall; /* read, then cleared for the next push */
  const int *tcb;
  const long long *hdr;
  long long *cnt;
  int nx, ny, T, M, W, ntx, nty;
};

/* pushed-particle count of the step: sum of the per-chunk live counts */
__device__ __forceinline__ void merge_live(const MergeArgs &a, const int blk, const int nblk) {
  __shared__ unsigned long long ls[4];
  unsigned long long sum = 0;
  for (int k = blk * blockDim.x + threadIdx.x; k < a.nlive; k += nblk * blockDim.x)
    sum += (unsigned long long)a.live_cnt[k];
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
  if ((threadIdx.x & 63) == 0) ls[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = ls[0] + ls[1] + ls[2] + ls[3];
    if (t) atomicAdd((unsigned long long *)&a.hdr_rw[HDR_PUSHED], t);
  }
  __syncthreads();
}

/* deposit count of one node: fallback grid + the windows that cover it.
 * Window of tile t spans nodes [tT - M, tT - M + W - 1] per axis, so the
 * covering tiles are t in [ceil((i + M - W + 1) / T), floor((i + M) / T)]:
 * at most 4 per axis because the margin M is at most T.  All terms are
 * int64, so the sum does not depend on the order of the additions. */
__device__ __forceinline__ int merge_tlo(int i, const MergeArgs &a) {
  const int u = i + a.M - a.W + 1;
  return u <= 0 ? 0 : (u + a.T - 1) / a.T;
}

template <int DIM>
__device__ __forceinline__ long long merge_node(const MergeArgs &a, const int node) {
  const int i = node % a.nx, j = node / a.nx;
  const unsigned long long g = a.gfall[node];
  if (g) a.gfall[node] = 0ull;
  long long s = (long long)g;
  if (a.hdr[HDR_NCHUNKS] <= 0) return s;
  const int WW = DIM == 2 ? a.W * a.W : a.W;
  const int tlo = merge_tlo(i, a), thi = min((i + a.M) / a.T, a.ntx - 1);
  const int ulo = DIM == 2 ? merge_tlo(j, a) : 0;
  const int uhi = DIM == 2 ? min((j + a.M) / a.T, a.nty - 1) : 0;
  if (DIM == 1 && a.tilewin) {
    for (int ti = tlo; ti <= thi; ++ti) s += a.tilewin[(size_t)ti * WW + (i - (ti * a.T - a.M))];
    return s;
  }
  if (DIM == 2 && a.tilewin) {
    long long v[16];
#pragma unroll
    for (int dy = 0; dy < 4; ++dy)
#pragma unroll
      for (int dx = 0; dx < 4; ++dx) {
        const int ti = tlo + dx, tj = ulo + dy;
        const int li = i - (ti * a.T - a.M), lj = j - (tj * a.T - a.M);
        v[4 * dy + dx] = (ti <= thi && tj <= uhi)
                             ? a.tilewin[(size_t)(ti + a.ntx * tj) * WW + lj * a.W + li]
                             : 0ll;
      }
#pragma unroll
    for (int k = 0; k < 16; ++k) s += v[k];
    return s;
  }
  for (int tj = ulo; tj <= uhi; ++tj) {
    const int lj = j - (tj * a.T - a.M);
    for (int ti = tlo; ti <= thi; ++ti) {
      const int li = i - (ti * a.T - a.M);
      const int t = ti + a.ntx * tj;
      const int q = DIM == 2 ? lj * a.W + li : li;
      for (int k = a.tcb[t]; k < a.tcb[t + 1]; ++k) s += a.partial[(size_t)k * WW + q];
    }
  }
  return s;
}

template <int DIM>
__device__ __forceinline__ void merge_body(const MergeArgs &a, const int blk, const int nblk) {
  if (a.nlive > 0) merge_live(a, blk, nblk);
  const int node = blk * blockDim.x + threadIdx.x;
  if (node >= a.nx * a.ny) return;
  a.cnt[node] = merge_node<DIM>(a, node);
}

template <int DIM>
__global__ __launch_bounds__(256) void k_merge(MergeArgs a) {
  merge_body<DIM>(a, blockIdx.x, gridDim.x);
}

/* stage 1 of the 2-D merge: one work-group per tile sums its chunks' windows
 * (contiguous in memory, chunks of a tile are consecutive) */
__device__ __forceinline__ void merge_tile_body(const long long *__restrict__ partial,
                                                const int *__restrict__ tcb,
                                                const long long *__restrict__ hdr,
                                                long long *__restrict__ tilewin, int WW, int t) {
  const int k0 = hdr[HDR_NCHUNKS] > 0 ? tcb[t] : 0, k1 = hdr[HDR_NCHUNKS] > 0 ? tcb[t + 1] : 0;
  if (2 * WW <= (int)blockDim.x) {
    /* narrow (1-D) windows: G = blockDim / WW thread groups take every G-th
     * chunk of the tile, then the G partial sums are added in LDS (integer
     * sums: any grouping gives the same result) */
    __shared__ long long red[256];
    const int G = (int)blockDim.x / WW, q = (int)threadIdx.x % WW, g = (int)threadIdx.x / WW;
    long long s0 = 0, s1 = 0;
    if (g < G) {
      int k = k0 + g;
      for (; k + G < k1; k += 2 * G) {
        s0 += partial[(size_t)k * WW + q];
        s1 += partial[(size_t)(k + G) * WW + q];
      }
      if (k < k1) s0 += partial[(size_t)k * WW + q];
    }
    red[threadIdx.x] = s0 + s1;
    __syncthreads();
    if (g == 0) {
      long long v = red[threadIdx.x];
      for (int h = 1; h < G; ++h) v += red[threadIdx.x + h * WW];
      tilewin[(size_t)t * WW + q] = v;
    }
    return;
  }
  for (int q = threadIdx.x; q < WW; q += blockDim.x) {
    /* four independent partial sums keep four loads in flight (integer
     * sums: the grouping does not change the result) */
    long long s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int k = k0;
    for (; k + 4 <= k1; k += 4) {
      s0 += partial[(size_t)k * WW + q];
      s1 += partial[(size_t)(k + 1) * WW + q];
      s2 += partial[(size_t)(k + 2) * WW + q];
      s3 += partial[(size_t)(k + 3) * WW + q];
    }
    for (; k < k1; ++k) s0 += partial[(size_t)k * WW + q];
    tilewin[(size_t)t * WW + q] = (s0 + s1) + (s2 + s3);
  }
}

__global__ __launch_bounds__(256) void k_merge_tiles(const long long *__restrict__ partial,
                                                     const int *__restrict__ tcb,
                                                     const long long *__restrict__ hdr,
                                                     long long *__restrict__ tilewin, int WW) {
  merge_tile_body(partial, tcb, hdr, tilewin, WW, blockIdx.x);
}

static void merge_prepare(iskra_ctx *c, int sid, int count_live, MergeArgs *pa, int *grid) {
  IskSpecies *s = &c->sp[sid];
  MergeArgs &a = *pa;
  a.live_cnt = s->live_cnt;
  a.nlive = count_live ? (int)s->max_chunks : 0;
  a.hdr_rw = s->hdr;
  a.partial = s->partial;
  a.gfall = s->gfall;
  a.tcb = s->tile_chunk_begin;
  a.hdr = s->hdr;
  a.cnt = s->cnt;
  a.nx = c->g.nx;
  a.ny = c->g.ny;
  a.T = c->t.T;
  a.M = isk_sp_margin(c, sid);
  a.W = isk_sp_window(c, sid);
  a.ntx = c->ntx;
  a.nty = c->nty;
  a.tilewin = nullptr;
  *grid = (isk_nodes(c) + 255) / 256;
}

int isk_merge_species(iskra_ctx *c, int sid, int count_live) {
  MergeArgs a;
  int grid;
  merge_prepare(c, sid, count_live, &a, &grid);
  IskSpecies *s = &c->sp[sid];
  const int W = isk_sp_window(c, sid);
  hipLaunchKernelGGL(k_merge_tiles, dim3(c->ntiles), dim3(256), 0, c->stream,
                     (const long long *)s->partial, (const int *)s->tile_chunk_begin,
                     (const long long *)s->hdr, s->tilewin, c->g.dim == 2 ? W * W : W);
  ISK_CHECK_LAUNCH();
  a.tilewin = s->tilewin;
  if (c->g.dim == 2)
    hipLaunchKernelGGL(k_merge<2>, dim3(grid), dim3(256), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_merge<1>, dim3(grid), dim3(256), 0, c->stream, a);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

/* ------------------------------------------------------------------------ */
/* rho = sum_s q_s n_s + background                                          */
/* ------------------------------------------------------------------------ */
struct RhoArgs {
  const long long *cnt[ISK_MAX_SPECIES];
  double qw[ISK_MAX_SPECIES]; /* w0 */
  double q[ISK_MAX_SPECIES];
  int ns;
  const double *vol, *bg;
  double *rho;
  int nn;
};

/* r + q_s n_s with n_s = count * w0 / V (count in 2^-32 units) */
__device__ __forceinline__ double rho_term(double r, long long cnt, double qw, double q, double V) {
  const double n = ((double)cnt * (1.0 / ISK_FIX_ONE)) * qw / V;
  return r + q * n;
}

__device__ __forceinline__ void rho_body(const RhoArgs &a, const int blk) {
  int node = blk * blockDim.x + threadIdx.x;
  if (node >= a.nn) return;
  double V = a.vol[node];
  double r = 0.0;
  for (int s = 0; s < a.ns; ++s) r = rho_term(r, a.cnt[s][node], a.qw[s], a.q[s], V);
  a.rho[node] = r + a.bg[node];
}

__global__ void k_rho(RhoArgs a) { rho_body(a, blockIdx.x); }

/* the time loop's deposit tail for all species in two launches: the tile
 * windows of every species (grid tiles x species), then per node the counts
 * of every species and rho (same arithmetic as k_merge + k_rho) */
struct MergeTilesN {
  const long long *partial[ISK_MAX_SPECIES];
  const int *tcb[ISK_MAX_SPECIES];
  const long long *hdr[ISK_MAX_SPECIES];
  long long *tilewin[ISK_MAX_SPECIES];
  int WW[ISK_MAX_SPECIES]; /* window nodes of each species */
};

__global__ __launch_bounds__(256) void k_merge_tiles_n(MergeTilesN a) {
  const int s = blockIdx.y;
  merge_tile_body(a.partial[s], a.tcb[s], a.hdr[s], a.tilewin[s], a.WW[s], blockIdx.x);
}

struct MergeRhoArgs {
  MergeArgs m[ISK_MAX_SPECIES];
  RhoArgs r;
  IskSpPack pk; /* F != null: also the spectral solve's packed right-hand side */
};

template <int DIM>
__global__ __launch_bounds__(256) void k_merge_rho(MergeRhoArgs a) {
  for (int s = 0; s < a.r.ns; ++s)
    if (a.m[s].nlive > 0) merge_live(a.m[s], blockIdx.x, gridDim.x);
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= a.r.nn) return;
  const double V = a.r.vol[node];
  double r = 0.0;
  for (int s = 0; s < a.r.ns; ++s) {
    const long long cnt = merge_node<DIM>(a.m[s], node);
    a.m[s].cnt[node] = cnt;
    r = rho_term(r, cnt, a.r.qw[s], a.r.q[s], V);
  }
  const double rho = r + a.r.bg[node];
  a.r.rho[node] = rho;
  if (DIM == 2 && a.pk.F) {
    const int nx = a.pk.s.nx, i = node % nx, j = node / nx;
    if (i > 0 && j > 0 && i < nx - 1 && j < a.pk.s.ny - 1)
      a.pk.F[(j - 1) * (nx - 2) + (i - 1)] = isk_sp_pack_value(a.pk, rho, i, j);
  }
}

int isk_rho(iskra_ctx *c) {
  RhoArgs a;
  int ns = 0;
  for (int s = 0; s < c->nspecies; ++s) {
    if (!c->sp[s].deposited) {
      int rc = isk_push_species(c, s, 0.0, 0);
      if (rc) return rc;
    }
    a.cnt[ns] = c->sp[s].cnt;
    a.qw[ns] = c->sp[s].w0;
    a.q[ns] = c->sp[s].q;
    ++ns;
  }
  a.ns = ns;
  a.vol = c->vol;
  a.bg = c->bg;
  a.rho = c->rho;
  a.nn = isk_nodes(c);
  hipLaunchKernelGGL(k_rho, dim3((a.nn + 255) / 256), dim3(256), 0, c->stream, a);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

/* time loop: push every species, then the tile windows of all species in
 * one launch and counts + rho in another */
int isk_push_deposit_all(iskra_ctx *c, double dt) {
  int rc;
  if ((rc = push_launch_all(c, dt))) return rc;
  hipEvent_t em = nullptr;
  if ((rc = isk_prof_begin(c, &em))) return rc;
  MergeRhoArgs a;
  int grid = 0;
  for (int s = 0; s < c->nspecies; ++s) {
    merge_prepare(c, s, 1, &a.m[s], &grid);
    a.m[s].tilewin = c->sp[s].tilewin;
    a.r.cnt[s] = c->sp[s].cnt;
    a.r.qw[s] = c->sp[s].w0;
    a.r.q[s] = c->sp[s].q;
  }
  a.r.ns = c->nspecies;
  a.r.vol = c->vol;
  a.r.bg = c->bg;
  a.r.rho = c->rho;
  a.r.nn = isk_nodes(c);
  isk_poisson_pack_desc(c, &a.pk);
  if (c->nspecies > 0) {
    MergeTilesN t;
    for (int s = 0; s < c->nspecies; ++s) {
      t.partial[s] = c->sp[s].partial;
      t.tcb[s] = c->sp[s].tile_chunk_begin;
      t.hdr[s] = c->sp[s].hdr;
      t.tilewin[s] = c->sp[s].tilewin;
      t.WW[s] = c->g.dim == 2 ? isk_sp_window(c, s) * isk_sp_window(c, s) : isk_sp_window(c, s);
    }
    hipLaunchKernelGGL(k_merge_tiles_n, dim3(c->ntiles, c->nspecies), dim3(256), 0, c->stream,
                       t);
    ISK_CHECK_LAUNCH();
  }
  grid = (a.r.nn + 255) / 256;
  if (c->g.dim == 2)
    hipLaunchKernelGGL(k_merge_rho<2>, dim3(grid), dim3(256), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_merge_rho<1>, dim3(grid), dim3(256), 0, c->stream, a);
  ISK_CHECK_LAUNCH();
  for (int s = 0; s < c->nspecies; ++s) c->sp[s].deposited = 1;
  return isk_prof_end(c, em, ISKRA_K_MERGE, 0);
}

/* ------------------------------------------------------------------------ */
/* gather only (grid_to_particle), for inspection                           */
/* ------------------------------------------------------------------------ */
template <int DIM>
__global__ void k_gather(const int32_t *cell, const float *ox, const float *oy,
                         const long long *hdr, const float *E, int nx,
                         float *out, long long cap) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= hdr[HDR_NP]) return;
  int cc = cell[p];
  float Ex = 0.f, Ey = 0.f;
  if (cc >= 0) {
    int i = cc & 0xFFFF, j = cc >> 16;
    float hx = ox[p];
    float ax = 1.0f - hx;
    if (DIM == 2) {
      float hy = oy[p];
      float ay = 1.0f - hy;
      float w00 = ax * ay, w10 = hx * ay, w01 = ax * hy, w11 = hx * hy;
      const float2 *E2 = reinterpret_cast<const float2 *>(E);
      int q = j * nx + i;
      float2 e00 = E2[q], e10 = E2[q + 1], e01 = E2[q + nx], e11 = E2[q + nx + 1];
      Ex = fmaf(w11, e11.x, fmaf(w01, e01.x, fmaf(w10, e10.x, w00 * e00.x)));
      Ey = fmaf(w11, e11.y, fmaf(w01, e01.y, fmaf(w10, e10.y, w00 * e00.y)));
    } else {
      Ex = fmaf(hx, E[i + 1], ax * E[i]);
    }
  }
  out[p] = Ex;
  out[cap + p] = Ey;
}

int isk_gather_species(iskra_ctx *c, int sid) {
  IskSpecies *s = &c->sp[sid];
  if (c->gather_cap < s->cap) {
    isk_free(c->gather_out);
    c->gather_out = nullptr;
    if (isk_alloc((void **)&c->gather_out, (size_t)s->cap * 2 * sizeof(float))) return ISKRA_E_HIP;
    c->gather_cap = s->cap;
  }
  int b = s->cur;
  int blocks = (int)((s->cap + 255) / 256);
  if (c->g.dim == 2)
    hipLaunchKernelGGL(k_gather<2>, dim3(blocks), dim3(256), 0, c->stream, s->cell[b],
                       s->ox[b], s->oy[b], s->hdr, c->E, c->g.nx, c->gather_out,
                       (long long)s->cap);
  else
    hipLaunchKernelGGL(k_gather<1>, dim3(blocks), dim3(256), 0, c->stream, s->cell[b],
                       s->ox[b], s->oy[b], s->hdr, c->E, c->g.nx, c->gather_out,
                       (long long)s->cap);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

/* ------------------------------------------------------------------------ */
/* batched launches over instances (one work-group = one work-group of one  */
/* instance's own launch, same device function)                             */
/* ------------------------------------------------------------------------ */
struct FillArgs {
  long long *p;
  long long n;
};

__global__ void kb_fill(const FillArgs *A, const int *bstart, int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  const long long k = (long long)(blockIdx.x - bstart[i]) * blockDim.x + threadIdx.x;
  if (k < A[i].n) A[i].p[k] = 0;
}

__global__ __launch_bounds__(PushShape<1>::block) void kb_push1(const PushArgs *A, const int *bstart,
                                                        int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  const PushArgs a = A[i]; /* in registers, not re-read after each store */
  push_body<1>(a, blockIdx.x - bstart[i]);
}

__global__ __launch_bounds__(256) void kb_merge1(const MergeArgs *A, const int *bstart, int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  const MergeArgs a = A[i];
  merge_body<1>(a, blockIdx.x - bstart[i], bstart[i + 1] - bstart[i]);
}

__global__ void kb_rho(const RhoArgs *A, const int *bstart, int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  rho_body(A[i], blockIdx.x - bstart[i]);
}

__global__ __launch_bounds__(SORT_BLOCK) void kb_sort_hist(const SortArgs *A, const int *bstart,
                                                            int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  sort_hist_body(A[i].cell, A[i].hdr, A[i].H, A[i].kd, blockIdx.x - bstart[i]);
}

__global__ void kb_sort_colscan(const SortArgs *A, const int *bstart, int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  const int gx = (A[i].nb1 + 255) / 256;
  const int l = blockIdx.x - bstart[i];
  sort_colscan_body(A[i].H, A[i].S, A[i].nb1, A[i].nblk, l % gx, l / gx);
}

__global__ void kb_sort_segscan(const SortArgs *A, const int *bstart, int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  sort_segscan_body(A[i].S, A[i].tot, A[i].nb1, A[i].nseg, blockIdx.x - bstart[i]);
}

__global__ __launch_bounds__(1024) void kb_sort_binscan(const SortArgs *A, const int *bstart,
                                                         int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  sort_binscan_body(A[i].tot, A[i].start, A[i].nb1);
}

__global__ __launch_bounds__(SORT_BLOCK) void kb_sort_scatter(const SortArgs *A,
                                                               const int *bstart, int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  sort_scatter_body(A[i].sc, blockIdx.x - bstart[i]);
}

__global__ __launch_bounds__(1024) void kb_sort_finish(const SortArgs *A, const int *bstart,
                                                        int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  sort_finish_body(A[i].hdr_w, A[i].tot, A[i].start, A[i].kd.nbins, A[i].build, A[i].CH,
                   A[i].chunks, A[i].tcb);
}

int isk_bbuild_fill(iskra_ctx **cs, int n, int sid, IskBLaunch *L) {
  std::vector<FillArgs> A(n);
  std::vector<int> blocks(n);
  for (int i = 0; i < n; ++i) {
    A[i].p = (long long *)cs[i]->sp[sid].gfall;
    A[i].n = isk_nodes(cs[i]);
    blocks[i] = (int)((A[i].n + 255) / 256);
  }
  return isk_blaunch_upload(L, ISK_B_FILL, A.data(), sizeof(FillArgs), blocks, 256, 0);
}

int isk_bbuild_push(iskra_ctx **cs, int n, int sid, double dt, const int *cur, IskBLaunch *L) {
  std::vector<PushArgs> A(n);
  std::vector<int> blocks(n);
  size_t shm = 0;
  for (int i = 0; i < n; ++i) {
    int grid;
    size_t sh;
    push_prepare(cs[i], sid, dt, 1, cur[i], &A[i], &grid, &sh);
    blocks[i] = grid;
    shm = std::max(shm, sh);
  }
  return isk_blaunch_upload(L, ISK_B_PUSH1, A.data(), sizeof(PushArgs), blocks, PushShape<1>::block,
                            shm);
}

int isk_bbuild_merge(iskra_ctx **cs, int n, int sid, IskBLaunch *L) {
  std::vector<MergeArgs> A(n);
  std::vector<int> blocks(n);
  for (int i = 0; i < n; ++i) merge_prepare(cs[i], sid, 1, &A[i], &blocks[i]);
  return isk_blaunch_upload(L, ISK_B_MERGE1, A.data(), sizeof(MergeArgs), blocks, 256, 0);
}

int isk_bbuild_rho(iskra_ctx **cs, int n, IskBLaunch *L) {
  std::vector<RhoArgs> A(n);
  std::vector<int> blocks(n);
  for (int i = 0; i < n; ++i) {
    iskra_ctx *c = cs[i];
    RhoArgs &a = A[i];
    a.ns = c->nspecies;
    for (int s = 0; s < c->nspecies; ++s) {
      a.cnt[s] = c->sp[s].cnt;
      a.qw[s] = c->sp[s].w0;
      a.q[s] = c->sp[s].q;
    }
    a.vol = c->vol;
    a.bg = c->bg;
    a.rho = c->rho;
    a.nn = isk_nodes(c);
    blocks[i] = (a.nn + 255) / 256;
  }
  return isk_blaunch_upload(L, ISK_B_RHO, A.data(), sizeof(RhoArgs), blocks, 256, 0);
}

/* the six launches of one tile-sort pass (with chunk table) */
int isk_bbuild_sort(iskra_ctx **cs, int n, int sid, const int *cur, IskBLaunch L[6]) {
  std::vector<SortArgs> A(n);
  std::vector<int> bh(n), bc(n), bs(n), one(n, 1);
  size_t shm = 0;
  for (int i = 0; i < n; ++i) {
    int rc = sort_prepare(cs[i], sid, 0, 1, cur[i], &A[i]);
    if (rc) return rc;
    bh[i] = A[i].nblk;
    bc[i] = ((A[i].nb1 + 255) / 256) * A[i].nseg;
    bs[i] = (A[i].nb1 + 255) / 256;
    shm = std::max(shm, (size_t)A[i].nb1 * sizeof(int));
  }
  int rc;
  if ((rc = isk_blaunch_upload(&L[0], ISK_B_SORT_HIST, A.data(), sizeof(SortArgs), bh,
                               SORT_BLOCK, shm)))
    return rc;
  if ((rc = isk_blaunch_upload(&L[1], ISK_B_SORT_COLSCAN, A.data(), sizeof(SortArgs), bc, 256,
                               0)))
    return rc;
  if ((rc = isk_blaunch_upload(&L[2], ISK_B_SORT_SEGSCAN, A.data(), sizeof(SortArgs), bs, 256,
                               0)))
    return rc;
  if ((rc = isk_blaunch_upload(&L[3], ISK_B_SORT_BINSCAN, A.data(), sizeof(SortArgs), one, 1024,
                               0)))
    return rc;
  if ((rc = isk_blaunch_upload(&L[4], ISK_B_SORT_SCATTER, A.data(), sizeof(SortArgs), bh,
                               SORT_BLOCK, sort_scatter_shm((int)(shm / sizeof(int))))))
    return rc;
  return isk_blaunch_upload(&L[5], ISK_B_SORT_FINISH, A.data(), sizeof(SortArgs), one, 1024, 0);
}

int isk_blaunch_particles(const IskBLaunch &L, hipStream_t st) {
  if (L.total <= 0) return ISKRA_OK;
  dim3 g((unsigned)L.total), b((unsigned)L.block);
  switch (L.kind) {
    case ISK_B_FILL:
      hipLaunchKernelGGL(kb_fill, g, b, 0, st, (const FillArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_PUSH1:
      hipLaunchKernelGGL(kb_push1, g, b, L.shm, st, (const PushArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_MERGE1:
      hipLaunchKernelGGL(kb_merge1, g, b, 0, st, (const MergeArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_RHO:
      hipLaunchKernelGGL(kb_rho, g, b, 0, st, (const RhoArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_SORT_HIST:
      hipLaunchKernelGGL(kb_sort_hist, g, b, L.shm, st, (const SortArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_SORT_COLSCAN:
      hipLaunchKernelGGL(kb_sort_colscan, g, b, 0, st, (const SortArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_SORT_SEGSCAN:
      hipLaunchKernelGGL(kb_sort_segscan, g, b, 0, st, (const SortArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_SORT_BINSCAN:
      hipLaunchKernelGGL(kb_sort_binscan, g, b, 0, st, (const SortArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_SORT_SCATTER:
      hipLaunchKernelGGL(kb_sort_scatter, g, b, L.shm, st, (const SortArgs *)L.args, L.bstart,
                         L.n);
      break;
    case ISK_B_SORT_FINISH:
      hipLaunchKernelGGL(kb_sort_finish, g, b, 0, st, (const SortArgs *)L.args, L.bstart, L.n);
      break;
    default:
      isk_set_error("batched launch: kind %d is not a particle kernel", L.kind);
      return ISKRA_E_ARG;
  }
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

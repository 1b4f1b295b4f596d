/*
 * mcc.hip -- null-collision Monte Carlo collisions on gfx950.
 *
 * Reference: PIC.perform!(::MonteCarloCollisions) (Chemistry/src/mcc.jl:231-289)
 * and the collision kernels (mcc.jl:53-229).  The reference draws
 * Nc = N*max_P*np candidates by array index with replacement; here every
 * particle is a candidate with probability N*max_P (the same expected number
 * of candidates per particle, and no order dependence), after which process
 * selection and acceptance follow the reference exactly:
 *     k = floor(N U) + 1,  P_k = (1 - exp(-n sigma_k(eps) g dt)) / (N max_P),
 *     accept if U > k/N - P_k.
 * All per-candidate arithmetic is FP64 with the exp/log/sincos of
 * isk_fp64math.h, so a CPU run that consumes the same draws reproduces every
 * decision and every post-collision velocity bit for bit.
 *
 * Draws: Philox4x32-10, key = seed, counter = (id, step, stream, block);
 * block 0 gives the candidate and process uniforms, blocks 1..5 the collision
 * draws in the order the reference consumes them.  Uniforms are 52-bit, in (0,1)
 * (isk_u01); normals are Box-Muller pairs.
 */
#include <math.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "isk_fp64math.h"
#include "isk_internal.h"

/* the pointers in the argument block point to global memory: said so, the
 * batched kernel (arguments read from a table in memory) loads and stores
 * with global instructions instead of flat ones */
#define ISK_GL(ptr) ((__attribute__((address_space(1))) __typeof__(*(ptr)) *)(ptr))
#define MCC_BLOCK 256
/* particles per MCC work-group: MCC_RPT rounds of MCC_BLOCK (the loads of
 * all rounds are issued together) */
#define MCC_RPT 4
#define MCC_PPB (MCC_BLOCK * MCC_RPT)
/* cross-section tables up to this many entries are staged in LDS */
#define MCC_TAB_LDS 1024
#define KB_MCC 1.3806503e-23      /* mcc.jl:75 */
#define QE_MCC 1.60217646e-19     /* mcc.jl:26 */
#define TWO_PI 6.283185307179586

struct MccArgs {
  int32_t *cell;
  float *ox, *oy, *vx, *vy, *vz;
  const uint32_t *id;
  const long long *hdr;
  const float *E;
  const double *tn;
  int dim, nx;
  int ncoll;
  int types[ISK_MAX_COLL];
  double energy[ISK_MAX_COLL];
  const int *toff;
  const double *teps, *tsig;
  double candP;  /* N * max_P */
  double dt;
  double m_eV;   /* source mass / qe */
  double ms, mt; /* source and target masses (kg) */
  double vth_t;  /* target thermal speed sqrt(2 kB T / m) */
  double tqm;    /* target q/m */
  uint32_t k0, k1, step, stream;
  const long long *step_dev; /* batched runs: draws use step + *step_dev */
  unsigned long long *nu, *events;
  int *errflag;
  int *blk_cnt;
  float *stage;
  int32_t *stage_cell;
  int *cand_i;   /* accepted candidates of each work-group (list records) */
  int *cand_cnt; /* per work-group accepted count */
  uint32_t *stage_pid; /* parent id of each staged product */
};

struct Draws {
  uint32_t id, step, stream, k0, k1;
  int block;
  int half;
  IskU4 cur;
  __device__ double next() {
    if (half == 0) {
      IskU4 c = {id, step, stream, (uint32_t)block};
      cur = isk_philox(c, k0, k1);
      ++block;
      half = 1;
      return isk_u01(cur.x, cur.y);
    }
    half = 0;
    return isk_u01(cur.z, cur.w);
  }
};

/* piecewise-linear sigma(e) of process k, flat outside the table
 * (cross_section.jl:3-17); the tables are read from LDS when the work-group
 * staged them, else from global memory */
__device__ __forceinline__ double interp_sigma(const int *toff, const double *teps,
                                              const double *tsig, int k, double e) {
  int lo = toff[k], hi = toff[k + 1] - 1;
  if (e <= teps[lo]) return tsig[lo];
  if (e >= teps[hi]) return tsig[hi];
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (teps[mid] <= e) lo = mid;
    else hi = mid;
  }
  double t = (e - teps[lo]) / (teps[hi] - teps[lo]);
  return tsig[lo] * (1.0 - t) + tsig[hi] * t;
}

struct Ang {
  double st, ct, sp, cp;
};

__device__ Ang euler_angles(double vx, double vy, double vz) {
  Ang g;
  double nv = sqrt(vx * vx + vy * vy + vz * vz);
  g.ct = vz / nv;
  g.st = sqrt(1.0 - g.ct * g.ct);
  if (g.st == 0.0) {
    g.cp = 1.0;
    g.sp = 0.0;
  } else {
    g.cp = vx / nv / g.st;
    g.sp = vy / nv / g.st;
  }
  return g;
}

/* [sinX cosE, sinX sinE, cosX] * unrotated(theta, phi)  (mcc.jl:83-87) */
__device__ void rotate(const Ang &g, double sX, double cX, double sE, double cE,
                       double *r) {
  double a0 = sX * cE, a1 = sX * sE, a2 = cX;
  r[0] = a0 * (g.cp * g.ct) + a1 * (-g.sp) + a2 * (g.cp * g.st);
  r[1] = a0 * (-g.sp * g.ct) + a1 * g.cp + a2 * (g.sp * g.st);
  r[2] = a0 * (-g.st) + a1 * 0.0 + a2 * g.ct;
}

__device__ void isotropic_scattering(double vx, double vy, double vz, Draws &d,
                                     double *r) {
  Ang g = euler_angles(vx, vy, vz);
  double sX, cX, sE, cE;
  isk_sincos2pi(d.next(), &sX, &cX);
  isk_sincos2pi(d.next(), &sE, &cE);
  rotate(g, sX, cX, sE, cE, r);
}

__device__ void diffuse_reflection(double vx, double vy, double vz, Draws &d,
                                   double *r) {
  Ang g = euler_angles(vx, vy, vz);
  double sX = sqrt(d.next());
  double cX = -sqrt(1.0 - sX * sX);
  double sE, cE;
  isk_sincos2pi(d.next(), &sE, &cE);
  rotate(g, sX, cX, sE, cE, r);
}

__device__ void randn3(Draws &d, double *n) {
  double u1 = d.next(), u2 = d.next(), u3 = d.next(), u4 = d.next();
  double r1 = sqrt(-2.0 * isk_log(u1)), r2 = sqrt(-2.0 * isk_log(u3));
  double s1, c1, s2, c2;
  isk_sincos2pi(u2, &s1, &c1);
  isk_sincos2pi(u4, &s2, &c2);
  n[0] = r1 * c1;
  n[1] = r1 * s1;
  n[2] = r2 * c2;
}

/* Two kernels per MCC run, on work-groups of MCC_PPB particles.
 * k_mcc_cand: every live particle draws (Uc, U) from Philox block 0 of its
 * stream (the cell and id loads of the work-group's MCC_RPT rounds are issued
 * together); the candidates (Uc < candP) are compacted in LDS in particle
 * order and then tested 256 at a time with all lanes busy: process k,
 * acceptance, event counts, and for an ionisation the product slot, ranked
 * in candidate order so the products keep particle order.  The accepted
 * candidates go to the block's list (local index | process << 11 |
 * (slot + 1) << 16).  k_mcc_coll: one wave per list groups the accepted
 * candidates by collision type in LDS and runs each type's FP64 collision
 * code 64 at a time, so a wave runs one type's code instead of every type's
 * branch in turn; each candidate's draws continue at Philox block 1, as they
 * would in one pass. */
#define MCC_NTYPE 5
__device__ __forceinline__ void mcc_cand_body(const MccArgs &a, const int blk) {
  __shared__ int wcount[MCC_BLOCK / 64];
  __shared__ int wprod[MCC_BLOCK / 64];
  __shared__ int s_ci[MCC_PPB];    /* candidates: local particle index */
  __shared__ double s_cu[MCC_PPB]; /* and their second uniform */
  __shared__ unsigned int evl[ISK_MAX_COLL];
  __shared__ int s_off[ISK_MAX_COLL + 1];
  /* the cross-section tables in LDS when they fit (dynamic LDS of
   * mcc_tab_shm bytes): the binary search then waits on LDS, not on a chain
   * of global loads */
  extern __shared__ double s_tab[];
  const long long np = a.hdr[HDR_NP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t step = a.step_dev ? (uint32_t)(a.step + *a.step_dev) : a.step;
  const long long b0 = (long long)blk * MCC_PPB;
  /* blocks past the live particles (the grid covers the capacity) */
  if (b0 >= np) {
    if (threadIdx.x == 0) {
      ISK_GL(a.cand_cnt)[blk] = 0;
      a.blk_cnt[blk] = 0;
    }
    return;
  }
  if ((int)threadIdx.x < ISK_MAX_COLL) evl[threadIdx.x] = 0u;
  const int ntab = a.toff[a.ncoll];
  const bool tab_lds = ntab <= MCC_TAB_LDS;
  double *s_eps = s_tab, *s_sig = s_tab + ntab;
  if ((int)threadIdx.x <= a.ncoll) s_off[threadIdx.x] = a.toff[threadIdx.x];
  if (tab_lds)
    for (int q = threadIdx.x; q < ntab; q += MCC_BLOCK) {
      s_eps[q] = ISK_GL(a.teps)[q];
      s_sig[q] = ISK_GL(a.tsig)[q];
    }
  /* cell and id of every round first (loads in flight together) */
  int32_t cr[MCC_RPT];
  uint32_t idr[MCC_RPT];
#pragma unroll
  for (int r = 0; r < MCC_RPT; ++r) {
    const long long p = b0 + r * MCC_BLOCK + threadIdx.x;
    const long long pc = p < np ? p : np - 1; /* np > b0 >= 0 here */
    cr[r] = ISK_GL(a.cell)[pc];
    idr[r] = ISK_GL(a.id)[pc];
  }
  int ncand = 0;
  for (int r = 0; r < MCC_RPT; ++r) {
    const int lt = r * MCC_BLOCK + threadIdx.x;
    bool cand = false;
    double U1 = 0.0;
    if (b0 + lt < np && cr[r] >= 0) {
      Draws d;
      d.id = idr[r];
      d.step = step;
      d.stream = a.stream;
      d.k0 = a.k0;
      d.k1 = a.k1;
      d.block = 0;
      d.half = 0;
      const double Uc = d.next();
      U1 = d.next();
      cand = Uc < a.candP;
    }
    /* candidates compacted in particle order */
    const unsigned long long m = __ballot(cand);
    if (lane == 0) wcount[w] = __popcll(m);
    __syncthreads();
    int off = ncand;
    for (int k = 0; k < w; ++k) off += wcount[k];
    if (cand) {
      const int q = off + __popcll(m & below);
      s_ci[q] = lt;
      s_cu[q] = U1;
    }
    for (int k = 0; k < MCC_BLOCK / 64; ++k) ncand += wcount[k];
    __syncthreads();
  }
  /* acceptance, MCC_BLOCK candidates at a time */
  const int N = a.ncoll;
  int nacc = 0, nprod = 0;
  for (int t0 = 0; t0 < ncand; t0 += MCC_BLOCK) {
    const int tcand = t0 + (int)threadIdx.x;
    int evkey = -1; /* node * N + process of this thread's collision */
    int k = 0, lt = 0;
    bool acc = false, prod = false;
    if (tcand < ncand) {
      lt = s_ci[tcand];
      const double U = s_cu[tcand];
      const long long p = b0 + lt;
      const int32_t c = ISK_GL(a.cell)[p];
      const int i = c & 0xFFFF, j = c >> 16;
      const int node = a.dim == 2 ? j * a.nx + i : i;
      const double n = ISK_GL(a.tn)[node];
      if (n >= 0.0) {
        k = (int)floor((double)N * U);
        if (k > N - 1) k = N - 1;
        const double vx = ISK_GL(a.vx)[p], vy = ISK_GL(a.vy)[p], vz = ISK_GL(a.vz)[p];
        double tvx = 0.0, tvy = 0.0;
        if (a.tqm != 0.0) {
          if (a.dim == 2) {
            tvx = a.tqm * (double)ISK_GL(a.E)[2 * node] * a.dt;
            tvy = a.tqm * (double)ISK_GL(a.E)[2 * node + 1] * a.dt;
          } else {
            tvx = a.tqm * (double)ISK_GL(a.E)[node] * a.dt;
          }
        }
        const double gx = tvx - vx, gy = tvy - vy, gz = 0.0 - vz;
        const double g = sqrt(gx * gx + gy * gy + gz * gz);
        const double eps = (0.5 * a.m_eV) * (g * g);
        const double sg = (tab_lds ? interp_sigma(s_off, s_eps, s_sig, k, eps)
                                   : interp_sigma(s_off, a.teps, a.tsig, k, eps)) * g;
        double P = 1.0 - isk_exp(((-n) * sg) * a.dt);
        P /= a.candP;
        if (P > 1.0) atomicOr(a.errflag, 1);
        if (U > (double)(k + 1) / (double)N - P) {
          acc = true;
          evkey = node * N + k;
          if (a.types[k] == ISKRA_COLL_IONIZATION) {
            const double sE = (0.5 * a.m_eV) * (vx * vx + vy * vy + vz * vz) - a.energy[k];
            prod = sE >= 0.0;
          }
        }
      }
    }
    /* accepted candidates and product slots, both in candidate order */
    const unsigned long long ma = __ballot(acc), mp = __ballot(prod);
    if (lane == 0) {
      wcount[w] = __popcll(ma);
      wprod[w] = __popcll(mp);
    }
    __syncthreads();
    int aoff = nacc, poff = nprod;
    for (int q = 0; q < w; ++q) {
      aoff += wcount[q];
      poff += wprod[q];
    }
    if (acc) {
      const int slot = prod ? poff + __popcll(mp & below) : -1;
      ISK_GL(a.cand_i)[b0 + aoff + __popcll(ma & below)] = lt | (k << 11) | ((slot + 1) << 16);
    }
    for (int q = 0; q < MCC_BLOCK / 64; ++q) {
      nacc += wcount[q];
      nprod += wprod[q];
    }
    /* event counts: equal (node, process) keys of a wave are counted with one
     * atomic, the per-process totals of the work-group go through LDS
     * (integer sums, so the grouping does not change them) */
    unsigned long long todo = __ballot(evkey >= 0);
    while (todo) {
      const int leader = __ffsll((long long)todo) - 1;
      const int lk = __shfl(evkey, leader);
      const unsigned long long mk = __ballot(evkey == lk);
      if (lane == leader) {
        atomicAdd(&a.nu[lk], (unsigned long long)__popcll(mk));
        atomicAdd(&evl[lk % N], (unsigned int)__popcll(mk));
      }
      todo &= ~mk;
    }
    __syncthreads(); /* wcount and wprod are rewritten next round */
  }
  if ((int)threadIdx.x < N && evl[threadIdx.x])
    atomicAdd(&a.events[threadIdx.x], (unsigned long long)evl[threadIdx.x]);
  if (threadIdx.x == 0) {
    ISK_GL(a.cand_cnt)[blk] = nacc;
    a.blk_cnt[blk] = nprod;
  }
}

/* k_mcc_coll: one wave per work-group's list of accepted candidates */
#define MCC_COLL 64
#define MCC_SEG 256 /* accepted candidates grouped by type per segment */
__device__ __forceinline__ void mcc_coll_body(const MccArgs &a, const int blk) {
  /* accepted candidates of a segment by type: list position */
  __shared__ unsigned short s_list[MCC_NTYPE][MCC_SEG];
  __shared__ int s_cnt[MCC_NTYPE];
  const int lane = threadIdx.x;
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t step = a.step_dev ? (uint32_t)(a.step + *a.step_dev) : a.step;
  const long long b0 = (long long)blk * MCC_PPB;
  const int nacc = ISK_GL(a.cand_cnt)[blk];
  for (int s0 = 0; s0 < nacc; s0 += MCC_SEG) {
    const int s1 = min(nacc, s0 + MCC_SEG);
    int lcnt[MCC_NTYPE] = {0, 0, 0, 0, 0};
    for (int t0 = s0; t0 < s1; t0 += MCC_COLL) {
      const int q = t0 + lane;
      const int type = q < s1 ? a.types[(ISK_GL(a.cand_i)[b0 + q] >> 11) & 0x1F] : -1;
#pragma unroll
      for (int T = 0; T < MCC_NTYPE; ++T) {
        const unsigned long long mt = __ballot(type == T);
        if (type == T) s_list[T][lcnt[T] + __popcll(mt & below)] = (unsigned short)(q - s0);
        lcnt[T] += __popcll(mt);
      }
    }
    if (lane < MCC_NTYPE)
      s_cnt[lane] = lane == 0   ? lcnt[0]
                    : lane == 1 ? lcnt[1]
                    : lane == 2 ? lcnt[2]
                    : lane == 3 ? lcnt[3]
                                : lcnt[4];
    __syncthreads();
    /* T is uniform: the type branches below are scalar branches */
#pragma unroll 1
    for (int T = 0; T < MCC_NTYPE; ++T) {
      const int nT = s_cnt[T];
      for (int q0 = 0; q0 < nT; q0 += MCC_COLL) {
        if (q0 + lane >= nT) continue;
        const int rec = ISK_GL(a.cand_i)[b0 + s0 + s_list[T][q0 + lane]];
        const int lt = rec & 0x7FF, k = (rec >> 11) & 0x1F, slot = (rec >> 16) - 1;
        const long long p = b0 + lt;
        Draws d;
        d.id = ISK_GL(a.id)[p];
        d.step = step;
        d.stream = a.stream;
        d.k0 = a.k0;
        d.k1 = a.k1;
        d.block = 1;
        d.half = 0;
        const double vx = ISK_GL(a.vx)[p], vy = ISK_GL(a.vy)[p], vz = ISK_GL(a.vz)[p];
        if (T == ISKRA_COLL_ELASTIC_ISOTROPIC || T == ISKRA_COLL_ELASTIC_BACKWARD ||
            T == ISKRA_COLL_INELASTIC_BACKWARD) {
          const double sv[3] = {vx, vy, vz};
          double mr1 = a.ms / (a.ms + a.mt), mr2 = a.mt / (a.ms + a.mt);
          double tv[3];
          randn3(d, tv);
          for (int r = 0; r < 3; ++r) tv[r] *= a.vth_t;
          double vr[3], wv[3];
          for (int r = 0; r < 3; ++r) {
            vr[r] = sv[r] - tv[r];
            wv[r] = mr1 * sv[r] + mr2 * tv[r];
          }
          double nvr = sqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
          double dir[3];
          if (T == ISKRA_COLL_ELASTIC_ISOTROPIC) {
            isotropic_scattering(vr[0], vr[1], vr[2], d, dir);
          } else {
            /* InelasticBackward (mcc.jl:144-157): |vr| scaled by a uniform
             * drawn before the direction */
            if (T == ISKRA_COLL_INELASTIC_BACKWARD) nvr = d.next() * nvr;
            diffuse_reflection(vr[0], vr[1], vr[2], d, dir);
          }
          ISK_GL(a.vx)[p] = (float)(wv[0] + mr2 * (nvr * dir[0]));
          ISK_GL(a.vy)[p] = (float)(wv[1] + mr2 * (nvr * dir[1]));
          ISK_GL(a.vz)[p] = (float)(wv[2] + mr2 * (nvr * dir[2]));
        } else if (T == ISKRA_COLL_EXCITATION) {
          double sE = (0.5 * a.m_eV) * (vx * vx + vy * vy + vz * vz) - a.energy[k];
          if (sE >= 0.0) {
            double al = sqrt(2.0 / a.m_eV);
            double ev = al * sqrt(sE);
            double dir[3];
            isotropic_scattering(vx, vy, vz, d, dir);
            ISK_GL(a.vx)[p] = (float)(ev * dir[0]);
            ISK_GL(a.vy)[p] = (float)(ev * dir[1]);
            ISK_GL(a.vz)[p] = (float)(ev * dir[2]);
          }
        } else if (slot >= 0) { /* ionisation with sE >= 0, decided by k_mcc_cand */
          double sE = (0.5 * a.m_eV) * (vx * vx + vy * vy + vz * vz) - a.energy[k];
          double e1E = sE * d.next();
          double e2E = sE - e1E;
          double al = sqrt(2.0 / a.m_eV);
          double e1v = al * sqrt(e1E), e2v = al * sqrt(e2E);
          double dir[3];
          diffuse_reflection(vx, vy, vz, d, dir);
          double nv[3] = {e1v * dir[0], e1v * dir[1], e1v * dir[2]};
          ISK_GL(a.vx)[p] = (float)nv[0];
          ISK_GL(a.vy)[p] = (float)nv[1];
          ISK_GL(a.vz)[p] = (float)nv[2];
          double dir2[3];
          diffuse_reflection(nv[0], nv[1], nv[2], d, dir2);
          double tv[3];
          randn3(d, tv);
          const size_t sidx = (size_t)b0 + slot;
          ISK_GL(a.stage_cell)[sidx] = ISK_GL(a.cell)[p];
          ISK_GL(a.stage_pid)[sidx] = d.id;
          auto *st = ISK_GL(a.stage) + sidx * 8;
          st[0] = ISK_GL(a.ox)[p];
          st[1] = a.dim == 2 ? ISK_GL(a.oy)[p] : 0.f;
          for (int r = 0; r < 3; ++r) {
            st[2 + r] = (float)(e2v * dir2[r]);
            st[5 + r] = (float)(tv[r] * a.vth_t);
          }
        }
      }
    }
    __syncthreads(); /* the next segment rewrites the lists */
  }
}

__global__ __launch_bounds__(MCC_BLOCK) void k_mcc_cand(MccArgs a) { mcc_cand_body(a, blockIdx.x); }
__global__ __launch_bounds__(MCC_COLL) void k_mcc_coll(MccArgs a) { mcc_coll_body(a, blockIdx.x); }

/* exclusive scan of the block product counts (one work-group); cnt[nblk] = total */
__device__ __forceinline__ void mcc_scan_body(int *cnt, int nblk) {
  __shared__ int sh[1024];
  int per = (nblk + 1023) / 1024;
  int lo = threadIdx.x * per, hi = min(lo + per, nblk);
  int sum = 0;
  for (int b = lo; b < hi; ++b) sum += cnt[b];
  sh[threadIdx.x] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int v = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - sum;
  for (int b = lo; b < hi; ++b) {
    int v = cnt[b];
    cnt[b] = run;
    run += v;
  }
  if (threadIdx.x == 1023) cnt[nblk] = sh[1023];
}

__global__ __launch_bounds__(1024) void k_mcc_scan(int *cnt, int nblk) { mcc_scan_body(cnt, nblk); }

struct AppendArgs {
  const int *blk_off; /* scanned, [nblk] = total */
  const float *stage;
  const int32_t *stage_cell;
  int nblk;
  /* electron (source) species */
  int32_t *ec;
  float *eox, *eoy, *evx, *evy, *evz;
  uint32_t *eid;
  long long *ehdr;
  long long ecap;
  /* ion species */
  int32_t *ic;
  float *iox, *ioy, *ivx, *ivy, *ivz;
  uint32_t *iid;
  long long *ihdr;
  long long icap;
  int R;
  int dim;
  const uint32_t *stage_pid; /* parent id per staged product */
  uint32_t *pid_c;           /* out: parent id of product g */
  long long *snap;           /* out (work-group 0): counts and next ids */
};

/* one work-group appends the products of MCC_APPEND_GROUP consecutive MCC
 * work-groups (their products are contiguous in the global order); product
 * g comes from the last of them whose offset is <= g.  Positions and
 * velocities go to slot count + g of each species; the ids are handed out
 * afterwards by the commit kernel in the order of the parents' ids, so a
 * product's id (and with it every later draw of it) does not depend on the
 * order the particles happen to be stored in. */
#define MCC_APPEND_GROUP 16
static inline int mcc_append_blocks(int nblk) {
  return (nblk + MCC_APPEND_GROUP - 1) / MCC_APPEND_GROUP;
}

__device__ __forceinline__ void mcc_append_one(const AppendArgs &a, int g, int b, int k);

__device__ __forceinline__ void mcc_append_body(const AppendArgs &a, const int b) {
  const int m0 = b * MCC_APPEND_GROUP, m1 = min(m0 + MCC_APPEND_GROUP, a.nblk);
  const int g1 = a.blk_off[m1];
  /* the counts and next ids the commit kernel hands ids out from (it must
   * not read the headers it rewrites) */
  if (b == 0 && threadIdx.x == 0) {
    a.snap[0] = a.ehdr[HDR_NP];
    a.snap[1] = a.ehdr[HDR_NEXT_ID];
    a.snap[2] = a.ihdr ? a.ihdr[HDR_NP] : 0;
    a.snap[3] = a.ihdr ? a.ihdr[HDR_NEXT_ID] : 0;
  }
  for (int g = a.blk_off[m0] + (int)threadIdx.x; g < g1; g += blockDim.x) {
    int lo = m0, hi = m1 - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.blk_off[mid] <= g) lo = mid;
      else hi = mid - 1;
    }
    mcc_append_one(a, g, lo, g - a.blk_off[lo]);
  }
}

/* product g = product k of MCC work-group b */
__device__ __forceinline__ void mcc_append_one(const AppendArgs &a, int g, int b, int k) {
  const size_t s = (size_t)b * MCC_PPB + k;
  const float *st = a.stage + s * 8;
  const int32_t c = a.stage_cell[s];
  long long enp = a.ehdr[HDR_NP];
  long long pe = enp + g;
  if (pe < a.ecap) {
    a.ec[pe] = c;
    a.eox[pe] = st[0];
    if (a.dim == 2) a.eoy[pe] = st[1];
    a.evx[pe] = st[2];
    a.evy[pe] = st[3];
    a.evz[pe] = st[4];
  }
  a.pid_c[g] = a.stage_pid[s];
  if (a.ic && a.R > 0) {
    long long inp = a.ihdr[HDR_NP];
    for (int r = 0; r < a.R; ++r) {
      long long pi = inp + (long long)g * a.R + r;
      if (pi >= a.icap) break;
      a.ic[pi] = c;
      a.iox[pi] = st[0];
      if (a.dim == 2) a.ioy[pi] = st[1];
      a.ivx[pi] = st[5];
      a.ivy[pi] = st[6];
      a.ivz[pi] = st[7];
    }
  }
}

__global__ void k_mcc_append(AppendArgs a) { mcc_append_body(a, blockIdx.x); }

struct CommitArgs {
  const int *blk_off;
  int nblk;
  long long *ehdr;
  long long ecap;
  long long *ihdr;
  long long icap;
  int R;
  uint32_t *eid, *iid;     /* id arrays of the buffers the products went to */
  const uint32_t *pid_c;   /* parent id of product g */
  const long long *snap;   /* counts and next ids before the append */
};

/* work-groups of the commit kernel when the interaction ionises (above
 * MCC_SORT_MAX products the ranking is O(products^2), spread over them) */
#define MCC_RANK_GROUPS 64
#define MCC_RANK_TILE 1024
/* threads per commit work-group; products ranked by one work-group's LDS
 * bitonic sort when there are at most MCC_SORT_MAX of them */
#define MCC_COMMIT_BLOCK 1024
#define MCC_SORT_MAX 8192

/* ids of product g at rank r */
__device__ __forceinline__ void mcc_commit_ids(const CommitArgs &a, int g, long long rank) {
  const long long enp = a.snap[0], enext = a.snap[1], inp = a.snap[2], inext = a.snap[3];
  if (a.eid && enp + g < a.ecap) a.eid[enp + g] = (uint32_t)(enext + rank);
  if (a.iid)
    for (int r = 0; r < a.R; ++r) {
      const long long pi = inp + (long long)g * a.R + r;
      if (pi >= a.icap) break;
      a.iid[pi] = (uint32_t)(inext + rank * a.R + r);
    }
}

/* Ids of the products and the species headers.  Product g receives
 * next_id + rank(g), rank = the number of products whose parent id is below
 * its parent's (parent ids are unique, so ranks are 0..P-1), and its R ions
 * next_ion_id + rank R + r.  Up to MCC_SORT_MAX products, work-group 0 sorts
 * the keys (parent id, g) in LDS (bitonic) and the rank is the sorted
 * position; above, every work-group counts smaller parent ids directly.
 * Work-group 0 then updates the headers; no work-group reads them (they use
 * the append kernel's snapshot). */
__device__ __forceinline__ void mcc_commit_body(const CommitArgs &a, const int blk,
                                                const int nblk_grid) {
  __shared__ unsigned long long keys[MCC_SORT_MAX]; /* also the rank tile */
  uint32_t *tile = (uint32_t *)keys;
  const long long tot = a.blk_off[a.nblk];
  if (tot == 0) return;
  const int P = (int)tot;
  const int bd = (int)blockDim.x;
  if ((a.eid || (a.iid && a.R > 0)) && P <= MCC_SORT_MAX) {
    if (blk == 0) {
      int n2 = 1;
      while (n2 < P) n2 <<= 1;
      for (int i = threadIdx.x; i < n2; i += bd)
        keys[i] = i < P ? ((unsigned long long)a.pid_c[i] << 32) | (unsigned)i : ~0ull;
      __syncthreads();
      for (int k = 2; k <= n2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int i = threadIdx.x; i < n2; i += bd) {
            const int l = i ^ j;
            if (l > i) {
              const unsigned long long x = keys[i], y = keys[l];
              if ((x > y) == ((i & k) == 0)) {
                keys[i] = y;
                keys[l] = x;
              }
            }
          }
          __syncthreads();
        }
      for (int r = threadIdx.x; r < P; r += bd)
        mcc_commit_ids(a, (int)(keys[r] & 0xffffffffu), r);
    }
  } else if (a.eid || (a.iid && a.R > 0)) {
    for (int g0 = blk * bd; g0 < P; g0 += nblk_grid * bd) {
      const int g = g0 + (int)threadIdx.x;
      const uint32_t mine = g < P ? a.pid_c[g] : 0u;
      int rank = 0;
      for (int t0 = 0; t0 < P; t0 += MCC_RANK_TILE) {
        const int tn = min(MCC_RANK_TILE, P - t0);
        __syncthreads(); /* the previous tile has been read */
        for (int q = threadIdx.x; q < tn; q += blockDim.x) tile[q] = a.pid_c[t0 + q];
        __syncthreads();
        for (int q = 0; q < tn; ++q) rank += tile[q] < mine ? 1 : 0;
      }
      if (g < P) mcc_commit_ids(a, g, rank);
    }
  }
  if (blk != 0 || threadIdx.x != 0) return;
  long long ne = a.snap[0] + tot;
  if (ne > a.ecap) {
    a.ehdr[HDR_OVERFLOW] = 1;
    ne = a.ecap;
  }
  a.ehdr[HDR_NP] = ne;
  a.ehdr[HDR_NEXT_ID] = a.snap[1] + tot;
  if (a.ihdr && a.R > 0) {
    long long ni = a.snap[2] + tot * a.R;
    if (ni > a.icap) {
      a.ihdr[HDR_OVERFLOW] = 1;
      ni = a.icap;
    }
    a.ihdr[HDR_NP] = ni;
    a.ihdr[HDR_NEXT_ID] = a.snap[3] + tot * a.R;
  }
}

__global__ __launch_bounds__(MCC_COMMIT_BLOCK) void k_mcc_commit(CommitArgs a) {
  mcc_commit_body(a, blockIdx.x, gridDim.x);
}

/* ------------------------------------------------------------------------ */
int isk_mcc_setup(iskra_ctx *c, int idx) {
  IskMcc &m = c->mcc[idx];
  const IskSpecies &src = c->sp[m.src];
  int nn = isk_nodes(c);
  m.m_eV = src.m / QE_MCC;
  /* max sigma_t g over the union of table nodes and 0 (mcc.jl:27-51) */
  std::vector<double> eps = {0.0};
  for (int k = 0; k < m.ncoll; ++k)
    for (int q = m.tab_off[k]; q < m.tab_off[k + 1]; ++q) eps.push_back(m.tab_eps[q]);
  std::sort(eps.begin(), eps.end());
  eps.erase(std::unique(eps.begin(), eps.end()), eps.end());
  double alpha = sqrt(2.0 / m.m_eV);
  double best = 0.0;
  for (double e : eps) {
    double v = alpha * sqrt(e);
    double s = 0.0;
    for (int k = 0; k < m.ncoll; ++k) {
      int lo = m.tab_off[k], hi = m.tab_off[k + 1] - 1;
      double sig;
      if (e <= m.tab_eps[lo]) sig = m.tab_sig[lo];
      else if (e >= m.tab_eps[hi]) sig = m.tab_sig[hi];
      else {
        int a = lo, b = hi;
        while (b - a > 1) {
          int mid = (a + b) / 2;
          if (m.tab_eps[mid] <= e) a = mid;
          else b = mid;
        }
        double t = (e - m.tab_eps[a]) / (m.tab_eps[b] - m.tab_eps[a]);
        sig = m.tab_sig[a] * (1.0 - t) + m.tab_sig[b] * t;
      }
      s += sig * v;
    }
    if (s > best) best = s;
  }
  m.max_sigma_g = best;
  size_t ntab = m.tab_eps.size();
  if (isk_alloc((void **)&m.d_eps, ntab * 8) || isk_alloc((void **)&m.d_sig, ntab * 8) ||
      isk_alloc((void **)&m.d_off, (m.ncoll + 1) * 4) ||
      isk_alloc((void **)&m.target_n, (size_t)nn * 8) ||
      isk_alloc((void **)&m.nu, (size_t)nn * m.ncoll * 8) ||
      isk_alloc((void **)&m.nu_acc, (size_t)nn * m.ncoll * 8) ||
      isk_alloc((void **)&m.events, ISK_MAX_COLL * 8) || isk_alloc((void **)&m.errflag, 64))
    return ISKRA_E_HIP;
  ISK_HIP(hipMemcpy(m.d_eps, m.tab_eps.data(), ntab * 8, hipMemcpyHostToDevice));
  ISK_HIP(hipMemcpy(m.d_sig, m.tab_sig.data(), ntab * 8, hipMemcpyHostToDevice));
  ISK_HIP(hipMemcpy(m.d_off, m.tab_off.data(), (m.ncoll + 1) * 4, hipMemcpyHostToDevice));
  ISK_HIP(hipMemset(m.target_n, 0, (size_t)nn * 8));
  ISK_HIP(hipMemset(m.nu, 0, (size_t)nn * m.ncoll * 8));
  ISK_HIP(hipMemset(m.nu_acc, 0, (size_t)nn * m.ncoll * 8));
  ISK_HIP(hipMemset(m.events, 0, ISK_MAX_COLL * 8));
  ISK_HIP(hipMemset(m.errflag, 0, 64));
  m.max_n = 0.0;
  m.step = 0;
  m.stream = 0x4D430000u + (uint32_t)idx;
  m.stage_cap = 0;
  m.prod_flag = nullptr;
  m.stage = nullptr;
  m.blk_cnt = nullptr;
  m.cand_i = nullptr;
  m.cand_cnt = nullptr;
  return ISKRA_OK;
}

/* arguments of the five launches of one MCC run, the source species reading
 * buffer b_src and the ion species buffer b_ion; allocates the staging area */
static int mcc_prepare(iskra_ctx *c, int idx, double dt, int b_src, int b_ion, MccArgs *pa,
                       AppendArgs *pp, CommitArgs *pc, int *nblk_out) {
  IskMcc &m = c->mcc[idx];
  IskSpecies &s = c->sp[m.src];
  const int N = m.ncoll;
  /* max_P = 1 - exp(-max_n max_sigma_g dt); assert max_P <= 1/N (mcc.jl:242-246) */
  double maxP = 1.0 - isk_exp(-m.max_n * m.max_sigma_g * dt);
  if (maxP > 1.0 / N) {
    isk_set_error("mcc: maximum probability %.6g is greater than 1/%d (mcc.jl:244)", maxP, N);
    return ISKRA_E_PROB;
  }
  int64_t cap = s.cap;
  int nblk = (int)((cap + MCC_PPB - 1) / MCC_PPB);
  if (m.stage_cap < cap) {
    isk_free(m.stage);
    isk_free(m.blk_cnt);
    isk_free(m.prod_flag);
    isk_free(m.cand_i);
    isk_free(m.cand_cnt);
    isk_free(m.stage_pid);
    isk_free(m.pid_c);
    isk_free(m.snap);
    m.cand_i = nullptr;
    m.cand_cnt = nullptr;
    m.stage = nullptr;
    m.blk_cnt = nullptr;
    m.prod_flag = nullptr;
    m.stage_pid = nullptr;
    m.pid_c = nullptr;
    m.snap = nullptr;
    if (isk_alloc((void **)&m.stage, (size_t)nblk * MCC_PPB * 8 * 4) ||
        isk_alloc((void **)&m.blk_cnt, (size_t)(nblk + 1) * 4) ||
        isk_alloc((void **)&m.prod_flag, (size_t)nblk * MCC_PPB * 4) ||
        isk_alloc((void **)&m.cand_i, (size_t)nblk * MCC_PPB * 4) ||
        isk_alloc((void **)&m.cand_cnt, (size_t)nblk * 4) ||
        isk_alloc((void **)&m.stage_pid, (size_t)nblk * MCC_PPB * 4) ||
        isk_alloc((void **)&m.pid_c, (size_t)nblk * MCC_PPB * 4) ||
        isk_alloc((void **)&m.snap, 64))
      return ISKRA_E_HIP;
    m.stage_cap = cap;
  }
  MccArgs &a = *pa;
  int b = b_src;
  a.cell = s.cell[b];
  a.ox = s.ox[b];
  a.oy = s.oy[b];
  a.vx = s.vx[b];
  a.vy = s.vy[b];
  a.vz = s.vz[b];
  a.id = s.id[b];
  a.hdr = s.hdr;
  a.E = c->E;
  a.tn = m.target_n;
  a.dim = c->g.dim;
  a.nx = c->g.nx;
  a.ncoll = N;
  for (int k = 0; k < ISK_MAX_COLL; ++k) {
    a.types[k] = k < N ? m.types[k] : 0;
    a.energy[k] = k < N ? m.energy[k] : 0.0;
  }
  a.toff = m.d_off;
  a.teps = m.d_eps;
  a.tsig = m.d_sig;
  a.candP = (double)N * maxP;
  a.dt = dt;
  a.m_eV = m.m_eV;
  a.ms = s.m;
  a.mt = m.target_m;
  a.vth_t = sqrt(2.0 * KB_MCC * m.target_T / m.target_m);
  a.tqm = m.target_q / m.target_m;
  a.k0 = (uint32_t)c->seed;
  a.k1 = (uint32_t)(c->seed >> 32);
  a.step = m.step;
  a.step_dev = nullptr;
  a.stream = m.stream;
  a.nu = m.nu;
  a.events = m.events;
  a.errflag = m.errflag;
  a.blk_cnt = m.blk_cnt;
  a.stage = m.stage;
  a.stage_cell = m.prod_flag;
  a.cand_i = m.cand_i;
  a.cand_cnt = m.cand_cnt;
  a.stage_pid = m.stage_pid;
  AppendArgs &p = *pp;
  p.stage_pid = m.stage_pid;
  p.pid_c = m.pid_c;
  p.snap = m.snap;
  p.blk_off = m.blk_cnt;
  p.stage = m.stage;
  p.stage_cell = m.prod_flag;
  p.nblk = nblk;
  p.ec = s.cell[b];
  p.eox = s.ox[b];
  p.eoy = s.oy[b];
  p.evx = s.vx[b];
  p.evy = s.vy[b];
  p.evz = s.vz[b];
  p.eid = s.track_ids ? s.id[b] : nullptr;
  p.ehdr = s.hdr;
  p.ecap = s.cap;
  int R = 0;
  p.ic = nullptr;
  p.ihdr = nullptr;
  p.icap = 0;
  bool has_ion = false;
  for (int k = 0; k < N; ++k) has_ion |= m.types[k] == ISKRA_COLL_IONIZATION;
  if (has_ion && m.ion >= 0 && m.ion != m.src) {
    IskSpecies &is = c->sp[m.ion];
    R = (int)rint(s.w0 / is.w0);
    int ib = b_ion;
    p.ic = is.cell[ib];
    p.iox = is.ox[ib];
    p.ioy = is.oy[ib];
    p.ivx = is.vx[ib];
    p.ivy = is.vy[ib];
    p.ivz = is.vz[ib];
    p.iid = is.track_ids ? is.id[ib] : nullptr;
    p.ihdr = is.hdr;
    p.icap = is.cap;
  }
  p.R = R;
  p.dim = c->g.dim;
  pc->blk_off = m.blk_cnt;
  pc->nblk = nblk;
  pc->ehdr = s.hdr;
  pc->ecap = (long long)s.cap;
  pc->ihdr = p.ihdr;
  pc->icap = (long long)p.icap;
  pc->R = R;
  pc->eid = p.eid;
  pc->iid = p.ic ? p.iid : nullptr;
  pc->pid_c = m.pid_c;
  pc->snap = m.snap;
  *nblk_out = nblk;
  return ISKRA_OK;
}

/* work-groups of the commit kernel: one when nothing is appended, else enough
 * to spread the ranking of the products */
static int mcc_commit_groups(const IskMcc &m, int nblk) {
  bool ionises = false;
  for (int k = 0; k < m.ncoll; ++k) ionises |= m.types[k] == ISKRA_COLL_IONIZATION;
  return ionises ? std::max(1, std::min(MCC_RANK_GROUPS, mcc_append_blocks(nblk))) : 1;
}

/* dynamic LDS of the candidate kernel: the cross-section tables when they fit */
static size_t mcc_tab_shm(const IskMcc &m) {
  const size_t nt = m.tab_eps.size();
  return nt <= MCC_TAB_LDS ? nt * 2 * sizeof(double) : 0;
}

int isk_mcc_run(iskra_ctx *c, int idx, double dt, int64_t *events) {
  IskMcc &m = c->mcc[idx];
  IskSpecies &s = c->sp[m.src];
  const int N = m.ncoll;
  MccArgs a;
  AppendArgs p;
  CommitArgs cm;
  int nblk;
  int b_ion = (m.ion >= 0 && m.ion < c->nspecies) ? c->sp[m.ion].cur : 0;
  int rc = mcc_prepare(c, idx, dt, s.cur, b_ion, &a, &p, &cm, &nblk);
  if (rc) return rc;
  if (events) ISK_HIP(hipMemsetAsync(m.events, 0, ISK_MAX_COLL * 8, c->stream));
  hipLaunchKernelGGL(k_mcc_cand, dim3(nblk), dim3(MCC_BLOCK), mcc_tab_shm(m), c->stream, a);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mcc_coll, dim3(nblk), dim3(MCC_COLL), 0, c->stream, a);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mcc_scan, dim3(1), dim3(1024), 0, c->stream, m.blk_cnt, nblk);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mcc_append, dim3(mcc_append_blocks(nblk)), dim3(MCC_BLOCK), 0,
                     c->stream, p);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mcc_commit, dim3(mcc_commit_groups(m, nblk)), dim3(MCC_COMMIT_BLOCK), 0,
                     c->stream, cm);
  ISK_CHECK_LAUNCH();
  if (p.ihdr) c->sp[m.ion].deposited = 0;
  s.deposited = 0;
  m.step += 1;
  if (events) {
    unsigned long long ev[ISK_MAX_COLL];
    int err = 0;
    ISK_HIP(hipMemcpyAsync(ev, m.events, sizeof(ev), hipMemcpyDeviceToHost, c->stream));
    ISK_HIP(hipMemcpyAsync(&err, m.errflag, 4, hipMemcpyDeviceToHost, c->stream));
    ISK_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < N; ++k) events[k] = (int64_t)ev[k];
    if (err) {
      isk_set_error("mcc: collision probability above 1 (mcc.jl:273-279)");
      return ISKRA_E_PROB;
    }
  }
  return ISKRA_OK;
}

/* ---- batched launches ---- */
__global__ __launch_bounds__(MCC_BLOCK) void kb_mcc_cand(const MccArgs *A, const int *bstart,
                                                         int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  mcc_cand_body(A[i], blockIdx.x - bstart[i]);
}

__global__ __launch_bounds__(MCC_COLL) void kb_mcc_coll(const MccArgs *A, const int *bstart,
                                                        int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  mcc_coll_body(A[i], blockIdx.x - bstart[i]);
}

__global__ __launch_bounds__(1024) void kb_mcc_scan(const CommitArgs *A, const int *bstart,
                                                     int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  mcc_scan_body(const_cast<int *>(A[i].blk_off), A[i].nblk);
}

__global__ __launch_bounds__(MCC_BLOCK) void kb_mcc_append(const AppendArgs *A, const int *bstart,
                                                            int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  const AppendArgs a = A[i];
  mcc_append_body(a, blockIdx.x - bstart[i]);
}

__global__ __launch_bounds__(MCC_COMMIT_BLOCK) void kb_mcc_commit(const CommitArgs *A,
                                                            const int *bstart, int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  const CommitArgs a = A[i];
  mcc_commit_body(a, blockIdx.x - bstart[i], bstart[i + 1] - bstart[i]);
}

int isk_bbuild_mcc(iskra_ctx **cs, int n, int m, double dt, const int *cur_src,
                   const int *cur_ion, long long *step_dev, IskBLaunch L[5]) {
  std::vector<MccArgs> A(n);
  std::vector<AppendArgs> P(n);
  std::vector<CommitArgs> CM(n);
  std::vector<int> blocks(n), ablocks(n), cblocks(n), one(n, 1);
  size_t tab_shm = 0;
  for (int i = 0; i < n; ++i) {
    int rc = mcc_prepare(cs[i], m, dt, cur_src[i], cur_ion[i], &A[i], &P[i], &CM[i], &blocks[i]);
    if (rc) return rc;
    tab_shm = std::max(tab_shm, mcc_tab_shm(cs[i]->mcc[m]));
    A[i].step_dev = step_dev;
    ablocks[i] = mcc_append_blocks(blocks[i]);
    cblocks[i] = mcc_commit_groups(cs[i]->mcc[m], blocks[i]);
  }
  int rc;
  if ((rc = isk_blaunch_upload(&L[0], ISK_B_MCC, A.data(), sizeof(MccArgs), blocks, MCC_BLOCK,
                               tab_shm)))
    return rc;
  if ((rc = isk_blaunch_upload(&L[1], ISK_B_MCC_COLL, A.data(), sizeof(MccArgs), blocks,
                               MCC_COLL, 0)))
    return rc;
  if ((rc = isk_blaunch_upload(&L[2], ISK_B_MCC_SCAN, CM.data(), sizeof(CommitArgs), one, 1024,
                               0)))
    return rc;
  if ((rc = isk_blaunch_upload(&L[3], ISK_B_MCC_APPEND, P.data(), sizeof(AppendArgs), ablocks,
                               MCC_BLOCK, 0)))
    return rc;
  return isk_blaunch_upload(&L[4], ISK_B_MCC_COMMIT, CM.data(), sizeof(CommitArgs), cblocks,
                            MCC_COMMIT_BLOCK, 0);
}

int isk_blaunch_mcc(const IskBLaunch &L, hipStream_t st) {
  if (L.total <= 0) return ISKRA_OK;
  dim3 g((unsigned)L.total), b((unsigned)L.block);
  switch (L.kind) {
    case ISK_B_MCC:
      hipLaunchKernelGGL(kb_mcc_cand, g, b, L.shm, st, (const MccArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_MCC_COLL:
      hipLaunchKernelGGL(kb_mcc_coll, g, b, 0, st, (const MccArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_MCC_SCAN:
      hipLaunchKernelGGL(kb_mcc_scan, g, b, 0, st, (const CommitArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_MCC_APPEND:
      hipLaunchKernelGGL(kb_mcc_append, g, b, 0, st, (const AppendArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_MCC_COMMIT:
      hipLaunchKernelGGL(kb_mcc_commit, g, b, 0, st, (const CommitArgs *)L.args, L.bstart, L.n);
      break;
    default:
      isk_set_error("batched launch: kind %d is not an MCC kernel", L.kind);
      return ISKRA_E_ARG;
  }
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

/* ------------------------------------------------------------------------ */
__global__ void k_rng_export(const uint32_t *ids, long long n, int nblocks,
                             uint32_t step, uint32_t stream, uint32_t k0,
                             uint32_t k1, uint32_t *out) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * nblocks) return;
  long long p = q / nblocks;
  int b = (int)(q % nblocks);
  IskU4 c = {ids[p], step, stream, (uint32_t)b};
  IskU4 r = isk_philox(c, k0, k1);
  uint32_t *o = out + q * 4;
  o[0] = r.x;
  o[1] = r.y;
  o[2] = r.z;
  o[3] = r.w;
}

int isk_rng_export(iskra_ctx *c, uint32_t stream, uint32_t step, const uint32_t *ids,
                   int64_t n, int nblocks, uint32_t *out) {
  if (n <= 0 || nblocks <= 0) return ISKRA_OK;
  uint32_t *dids = nullptr, *dout = nullptr;
  if (isk_alloc((void **)&dids, n * 4) || isk_alloc((void **)&dout, n * nblocks * 16))
    return ISKRA_E_HIP;
  ISK_HIP(hipMemcpy(dids, ids, n * 4, hipMemcpyHostToDevice));
  long long tot = n * nblocks;
  hipLaunchKernelGGL(k_rng_export, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     c->stream, dids, (long long)n, nblocks, step, stream,
                     (uint32_t)c->seed, (uint32_t)(c->seed >> 32), dout);
  ISK_CHECK_LAUNCH();
  ISK_HIP(hipStreamSynchronize(c->stream));
  ISK_HIP(hipMemcpy(out, dout, n * nblocks * 16, hipMemcpyDeviceToHost));
  isk_free(dids);
  isk_free(dout);
  return ISKRA_OK;
}

/* ---- the draw export without device allocation ----
 * The staging of iskra_rng_export_noalloc (the ids at offset 0, the draws
 * after them at the next 256-byte boundary) is this file-scope array: every
 * GPU holds its own copy until the process ends.  In host code the array's
 * name is only a stand-in; the device address of each GPU's copy comes from
 * hipGetSymbolAddress, taken once per device at context creation with that
 * device current, and only that address is handed to copies and kernels. */
#define ISK_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[ISKRA_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[ISK_NOALLOC_MAX_DEVICES];

int isk_noalloc_init(int device) {
  if (device < 0 || device >= ISK_NOALLOC_MAX_DEVICES) {
    isk_set_error("device %d: the no-allocation scratch supports devices 0..%d", device,
                  ISK_NOALLOC_MAX_DEVICES - 1);
    return ISKRA_E_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return ISKRA_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return isk_hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
  if (!p) {
    isk_set_error("hipGetSymbolAddress(scratch) returned a null address on device %d", device);
    return ISKRA_E_HIP;
  }
  g_scratch_ptr[device] = (unsigned char *)p;
  return ISKRA_OK;
}

int isk_rng_export_noalloc(iskra_ctx *c, uint32_t stream, uint32_t step, const uint32_t *ids,
                           int64_t n, int nblocks, uint32_t *out) {
  if (n <= 0 || nblocks <= 0) return ISKRA_OK;
  const int64_t cap = ISKRA_NOALLOC_SCRATCH_BYTES;
  /* the size test cannot overflow: n <= cap/4 and nblocks <= cap/16 bound the
   * product by 2^38 */
  const int64_t ids_bytes = (n * 4 + 255) / 256 * 256;
  if (n > cap / 4 || nblocks > cap / 16 || ids_bytes + n * (int64_t)nblocks * 16 > cap) {
    isk_set_error("rng export without allocation: %lld ids x %d blocks exceed the %lld-byte "
                  "scratch (it needs round_up(4 n, 256) + 16 n nblocks bytes)",
                  (long long)n, nblocks, (long long)cap);
    return ISKRA_E_ARG;
  }
  unsigned char *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (c->device >= 0 && c->device < ISK_NOALLOC_MAX_DEVICES) base = g_scratch_ptr[c->device];
  }
  if (!base) {
    isk_set_error("no-allocation scratch of device %d was not set up", c->device);
    return ISKRA_E_STATE;
  }
  uint32_t *dids = (uint32_t *)base, *dout = (uint32_t *)(base + ids_bytes);
  ISK_HIP(hipMemcpyAsync(dids, ids, n * 4, hipMemcpyHostToDevice, c->stream));
  const long long tot = n * nblocks;
  hipLaunchKernelGGL(k_rng_export, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     c->stream, dids, (long long)n, nblocks, step, stream,
                     (uint32_t)c->seed, (uint32_t)(c->seed >> 32), dout);
  ISK_CHECK_LAUNCH();
  ISK_HIP(hipMemcpyAsync(out, dout, tot * 16, hipMemcpyDeviceToHost, c->stream));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

/*
 * poisson.hip -- finite-difference Poisson solve and E = -grad(phi) on gfx950.
 *
 * System (FiniteDifferenceMethod/src/generalized_poisson.jl:34-68,205-215,
 * 271-324,372-378): A phi = b with b = -rho/eps0 at free nodes, A the
 * natural-edge 3/5-point Laplacian scaled by 1/dx^2 (also along y,
 * generalized_poisson.jl:65), periodic ring links from apply_periodic and
 * Dirichlet rows (phi = value) from apply_dirichlet.  Multiplying by -dx^2
 * gives the symmetric positive form used here,
 *     L phi = f,   (L phi)_i = sum_links c (phi_i - phi_nb),   f = dx^2 rho/eps0,
 * with Dirichlet values moved to the right-hand side (same solution as the
 * reference's non-symmetric A).
 *
 * Solvers:
 *  DIRECT  1-D tridiagonal elimination (the reference's A\b is a direct solve),
 *          FP64, factor built once on the host, substitution on the device.
 *  PCG     Jacobi-preconditioned conjugate gradients: x, r, Ap and all inner
 *          products in FP64, search direction p and z in FP32; a dot product
 *          is per-work-group partials, summed in block order by the kernel
 *          that needs it (every work-group of it gets the same bits).
 *  MG      geometric multigrid V(2,2) cycles with red-black Gauss-Seidel on FP32
 *          levels as a defect correction of the FP64 iterate (residual, norm
 *          and update in FP64); levels <= 65^2 run in one work-group from LDS.
 *  MGPCG   PCG with one symmetric V-cycle as preconditioner.
 * Stopping rule: ||f_elim - L x||_2 / ||f_elim||_2 <= rtol in FP64 (f_elim is
 * f with the Dirichlet contributions moved over).  A system without any
 * Dirichlet node is singular (periodic); its right-hand side is projected to
 * zero mean and the mean-free solution is returned (the reference's A\b fixes
 * the constant by round-off).
 */
#include <math.h>

#include <algorithm>
#include <vector>

#include "isk_internal.h"

#define RED_BLOCK 256
#define PCG_BATCH 32
#define MG_BATCH 4
#define MG_COARSE_MAX 65

/* scalars: S_RZ0/S_RZ1 hold (r,z) of even/odd CG iterations, S_RTOL2 = rtol^2,
 * S_TOL2 = rtol^2 (f,f) */
enum { S_PAP = 0, S_ALPHA, S_BETA, S_RR, S_FF, S_TOL2, S_MEAN, S_RELRES, S_RZ0, S_RZ1, S_RTOL2,
       S_WORDS = 16 };
/* iflag words: [0] done, [1] iterations, [2] singular (PCG/MG); spectral solve:
 * [4] this solve met rtol, [5] sticky: an in-loop solve missed rtol,
 * [6] transform passes of this solve */
enum { IF_SP_DONE = 4, IF_SP_FAIL = 5, IF_SP_PASSES = 6 };


template <typename T>
__device__ __forceinline__ double apply_L(const Sten &s, const T *v, int i, int j,
                                          double *diag_out) {
  int n = j * s.nx + i;
  double acc = 0.0, diag = 0.0;
  if (s.axial) {
    const double cz = isk_rz_w(i, s.nx) * s.az;
    if (i > 0) { acc += ((double)i - 0.5) * (double)v[n - 1]; diag += (double)i - 0.5; }
    if (i < s.nx - 1) { acc += ((double)i + 0.5) * (double)v[n + 1]; diag += (double)i + 0.5; }
    if (j > 0) { acc += cz * (double)v[n - s.nx]; diag += cz; }
    if (j < s.ny - 1) { acc += cz * (double)v[n + s.nx]; diag += cz; }
    if (diag_out) *diag_out = diag;
    return diag * (double)v[n] - acc;
  }
  if (i > 0) { acc += (double)v[n - 1]; diag += 1.0; }
  if (i < s.nx - 1) { acc += (double)v[n + 1]; diag += 1.0; }
  if (s.per_x) {
    if (i == 0) { acc += s.cxp * (double)v[n + s.nx - 1]; diag += s.cxp; }
    if (i == s.nx - 1) { acc += s.cxp * (double)v[n - (s.nx - 1)]; diag += s.cxp; }
  }
  if (s.dim == 2) {
    if (j > 0) { acc += (double)v[n - s.nx]; diag += 1.0; }
    if (j < s.ny - 1) { acc += (double)v[n + s.nx]; diag += 1.0; }
    if (s.per_y) {
      if (j == 0) { acc += s.cyp * (double)v[n + (s.ny - 1) * s.nx]; diag += s.cyp; }
      if (j == s.ny - 1) { acc += s.cyp * (double)v[n - (s.ny - 1) * s.nx]; diag += s.cyp; }
    }
  }
  if (diag_out) *diag_out = diag;
  return diag * (double)v[n] - acc;
}

/* E = -grad phi at node n: central differences inside, one-sided at the ends */
__device__ __forceinline__ void efield_node(const double *phi, float *E, int dim, int nx, int ny,
                                            double dx, double dy, int n) {
  const int i = n % nx, j = n / nx;
  double ex;
  if (nx < 2) ex = 0.0;
  else if (i == 0) ex = (phi[n] - phi[n + 1]) / dx;
  else if (i == nx - 1) ex = (phi[n - 1] - phi[n]) / dx;
  else ex = (phi[n - 1] - phi[n + 1]) / (2.0 * dx);
  if (dim == 1) {
    E[n] = (float)ex;
    return;
  }
  double ey;
  if (ny < 2) ey = 0.0;
  else if (j == 0) ey = (phi[n] - phi[n + nx]) / dy;
  else if (j == ny - 1) ey = (phi[n - nx] - phi[n]) / dy;
  else ey = (phi[n - nx] - phi[n + nx]) / (2.0 * dy);
  E[2 * n] = (float)ex;
  E[2 * n + 1] = (float)ey;
}



/* ------------------------------------------------------------------------ */
/* 1-D direct solve                                                          */
/* ------------------------------------------------------------------------ */
struct TriArgs {
  const double *rho;
  const int8_t *tags;
  const double *tagval;
  const double *a, *cp, *inv;
  double *y, *phi;
  int n;
  double scale; /* dx^2 / eps0 */
  int singular;
};

/* one thread: rhs, forward elimination, back substitution (FP64, explicit fma) */
__device__ __forceinline__ void tri_solve_body(const TriArgs &t) {
  const int n = t.n;
  if (!t.singular) {
    double yprev = 0.0;
    for (int i = 0; i < n; ++i) {
      int tg = t.tags[i];
      double r = tg ? t.tagval[tg] : -(t.rho[i] * t.scale);
      double y = __builtin_fma(-t.a[i], yprev, r) * t.inv[i];
      t.y[i] = y;
      yprev = y;
    }
    double xn = 0.0;
    for (int i = n - 1; i >= 0; --i) {
      double x = __builtin_fma(-t.cp[i], xn, t.y[i]);
      t.phi[i] = x;
      xn = x;
    }
    return;
  }
  /* periodic ring without Dirichlet nodes: project, pin node 0, solve nodes 1..n-1 */
  double sum = 0.0;
  for (int i = 0; i < n; ++i) sum += -(t.rho[i] * t.scale);
  double mean = sum / (double)n;
  double yprev = 0.0;
  for (int i = 1; i < n; ++i) {
    double r = -(t.rho[i] * t.scale) - mean;
    double y = __builtin_fma(-t.a[i], yprev, r) * t.inv[i];
    t.y[i] = y;
    yprev = y;
  }
  double xn = 0.0;
  double s2 = 0.0;
  for (int i = n - 1; i >= 1; --i) {
    double x = __builtin_fma(-t.cp[i], xn, t.y[i]);
    t.phi[i] = x;
    xn = x;
  }
  t.phi[0] = 0.0;
  for (int i = 0; i < n; ++i) s2 += t.phi[i];
  double m2 = s2 / (double)n;
  for (int i = 0; i < n; ++i) t.phi[i] = t.phi[i] - m2;
}

/* The same system solved by a whole work-group of TRI_BLOCK threads when it
 * has Dirichlet nodes (not the singular ring) and n <= TRI_BLOCK * TRI_K.
 * Both recurrences of the factored system are first-order affine maps,
 * forward y_i = A_i y_(i-1) + B_i with A_i = -a_i inv_i, B_i = r_i inv_i, and
 * backward x_i = C_i x_(i+1) + D_i with C_i = -cp_i, D_i = y_i, so each is a
 * prefix composition of maps, (A2, B2) after (A1, B1) = (A2 A1, fma(A2, B1,
 * B2)): every thread composes its TRI_K consecutive maps, a Hillis-Steele
 * scan over the TRI_BLOCK segment maps gives each segment its incoming value,
 * and every element is then one fma away.  log2(TRI_BLOCK) + 2 TRI_K
 * dependent steps instead of 3n.  The operations are fixed, so the oracle
 * restates them bit for bit.  Otherwise thread 0 runs tri_solve_body. */
#define TRI_BLOCK 256
#define TRI_K 8

/* one prefix composition of n affine maps (ma[i], mb[i]) applied to 0;
 * rev != 0 scans from the last element down.  out[i] = value at element i. */
__device__ __forceinline__ void tri_scan(const double *ma, const double *mb, int n, int rev,
                                         double *out, double *ga, double *gb) {
  const int tid = threadIdx.x, j0 = tid * TRI_K;
  double la[TRI_K], lb[TRI_K];
  double ca = 1.0, cb = 0.0;
#pragma unroll
  for (int k = 0; k < TRI_K; ++k) {
    const int j = j0 + k;
    if (j < n) {
      const int i = rev ? n - 1 - j : j;
      const double A = ma[i], B = mb[i];
      cb = __builtin_fma(A, cb, B);
      ca = A * ca;
    }
    la[k] = ca;
    lb[k] = cb;
  }
  ga[tid] = ca;
  gb[tid] = cb;
  __syncthreads();
  for (int d = 1; d < TRI_BLOCK; d <<= 1) {
    double pa = 1.0, pb = 0.0;
    if (tid >= d) {
      pa = ga[tid - d];
      pb = gb[tid - d];
    }
    __syncthreads();
    if (tid >= d) {
      const double a2 = ga[tid], b2 = gb[tid];
      gb[tid] = __builtin_fma(a2, pb, b2);
      ga[tid] = a2 * pa;
    }
    __syncthreads();
  }
  const double vin = tid > 0 ? gb[tid - 1] : 0.0;
#pragma unroll
  for (int k = 0; k < TRI_K; ++k) {
    const int j = j0 + k;
    if (j < n) out[rev ? n - 1 - j : j] = __builtin_fma(la[k], vin, lb[k]);
  }
  __syncthreads(); /* out, ga and gb are read or rewritten next */
}

/* sum of v[0..n) in a fixed order: each thread adds its TRI_K consecutive
 * elements, then a pairwise tree over the TRI_BLOCK partials (every thread
 * returns the same bits) */
__device__ __forceinline__ double tri_sum(const double *v, int n, double *g) {
  const int tid = threadIdx.x, j0 = tid * TRI_K;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < TRI_K; ++k)
    if (j0 + k < n) s += v[j0 + k];
  g[tid] = s;
  __syncthreads();
  for (int d = TRI_BLOCK / 2; d > 0; d >>= 1) {
    if (tid < d) g[tid] += g[tid + d];
    __syncthreads();
  }
  const double r = g[0];
  __syncthreads();
  return r;
}

/* the periodic ring without Dirichlet nodes (A1/A2): right-hand side
 * projected to zero mean, node 0 pinned, rows 1..n-1 solved by the two
 * scans, then the mean of phi removed; sums in tri_sum's order */
__device__ __forceinline__ void tri_solve_ring_group(const TriArgs &t, double *sA, double *sB,
                                                     double *sy, double *ga, double *gb) {
  const int n = t.n, m = n - 1;
  for (int i = threadIdx.x; i < n; i += blockDim.x) sy[i] = -(t.rho[i] * t.scale);
  __syncthreads();
  const double mean = tri_sum(sy, n, ga) / (double)n;
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    const int i = j + 1;
    const double inv = t.inv[i];
    sA[j] = -t.a[i] * inv;
    sB[j] = (sy[i] - mean) * inv;
  }
  __syncthreads();
  tri_scan(sA, sB, m, 0, sy, ga, gb);
  for (int j = threadIdx.x; j < m; j += blockDim.x) sA[j] = -t.cp[j + 1];
  __syncthreads();
  tri_scan(sA, sy, m, 1, sB, ga, gb);
  /* phi over all n nodes in sy: node 0 pinned at 0 */
  for (int i = threadIdx.x; i < n; i += blockDim.x) sy[i] = i == 0 ? 0.0 : sB[i - 1];
  __syncthreads();
  const double m2 = tri_sum(sy, n, ga) / (double)n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) t.phi[i] = sy[i] - m2;
}

__device__ __forceinline__ void tri_solve_group(const TriArgs &t) {
  const int n = t.n;
  __shared__ double sA[TRI_BLOCK * TRI_K], sB[TRI_BLOCK * TRI_K], sy[TRI_BLOCK * TRI_K];
  __shared__ double ga[TRI_BLOCK], gb[TRI_BLOCK];
  if (t.singular && n <= TRI_BLOCK * TRI_K) {
    tri_solve_ring_group(t, sA, sB, sy, ga, gb);
    return;
  }
  if (t.singular || n > TRI_BLOCK * TRI_K) {
    if (threadIdx.x == 0) tri_solve_body(t);
    return;
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int tg = t.tags[i];
    const double r = tg ? t.tagval[tg] : -(t.rho[i] * t.scale);
    const double inv = t.inv[i];
    sA[i] = -t.a[i] * inv;
    sB[i] = r * inv;
  }
  __syncthreads();
  tri_scan(sA, sB, n, 0, sy, ga, gb);
  for (int i = threadIdx.x; i < n; i += blockDim.x) sA[i] = -t.cp[i];
  __syncthreads();
  tri_scan(sA, sy, n, 1, sB, ga, gb);
  for (int i = threadIdx.x; i < n; i += blockDim.x) t.phi[i] = sB[i];
}

__global__ __launch_bounds__(TRI_BLOCK) void k_tri_solve(TriArgs t) {
  if (blockIdx.x != 0) return;
  tri_solve_group(t);
}

/* ------------------------------------------------------------------------ */
/* PCG                                                                       */
/* ------------------------------------------------------------------------ */
struct PcgArgs {
  Sten s;
  const double *rho;
  double *x, *r, *Ap;
  float *z, *p0, *p1; /* search directions of even / odd iterations */
  float *f0;  /* MG preconditioner input (MGPCG) */
  float *e0;  /* MG finest-level correction, zeroed before each V-cycle */
  double *scal, *partials;
  int *iflag;
  double *cgp; /* CG partials, 7 per work-group */
  int nblk;    /* work-groups of the CG kernels */
  double scale;
  int warm;
  const double *phi_in;
  double *phi_out; /* spectral solve: the residual kernels write phi here */
  int mgpre;
  int max_iter;
};

/* x = Dirichlet values / warm start / 0 */
__global__ void k_pcg_setx(PcgArgs a) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.s.nx * a.s.ny) return;
  int t = a.s.tags[n];
  a.x[n] = t ? a.s.tagval[t] : (a.warm ? a.phi_in[n] : 0.0);
}

/* scalars and flags of a new solve */
__global__ void k_pcg_init(double *scal, int *iflag, double tol2, int singular) {
  if (threadIdx.x < S_WORDS)
    scal[threadIdx.x] = (threadIdx.x == S_TOL2 || threadIdx.x == S_RTOL2) ? tol2 : 0.0;
  if (threadIdx.x < 4) iflag[threadIdx.x] = threadIdx.x == 2 ? singular : 0;
}

/* CG without device-wide reductions: a kernel that needs a dot product sums
 * the per-work-group partials that the previous kernel wrote, in block order,
 * in every work-group (the same data summed the same way gives the same bits
 * everywhere), so no kernel waits on the others through fences or counters.
 * Partials, RED_BLOCK-node work-groups: two slots by iteration parity of
 * {(r,r), (r,z), (f,f)} per work-group, then one (p,Ap) per work-group.
 * Iteration it: A(it) takes (r,r) and (r,z) of x_it, decides convergence,
 * forms p_it = z + beta p_{it-1} in p[it & 1] and (p, Ap); B(it) updates
 * x, r, z and writes the partials of x_{it+1} into slot (it + 1) & 1. */

/* sums of 3 doubles over the work-group: an xor butterfly within each wave
 * (every lane ends with the same bits), then the waves in index order */
template <int NT = RED_BLOCK>
__device__ __forceinline__ void wg_sum3(double &a, double &b, double &c) {
  __shared__ double w3[3][NT / 64];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
    c += __shfl_xor(c, off);
  }
  __syncthreads(); /* a previous call's readers are done with w3 */
  if ((threadIdx.x & 63) == 0) {
    const int w = threadIdx.x >> 6;
    w3[0][w] = a;
    w3[1][w] = b;
    w3[2][w] = c;
  }
  __syncthreads();
  double sa = 0.0, sb = 0.0, sc = 0.0;
  for (int k = 0; k < NT / 64; ++k) {
    sa += w3[0][k];
    sb += w3[1][k];
    sc += w3[2][k];
  }
  a = sa;
  b = sb;
  c = sc;
}

__device__ __forceinline__ double *cg_slot(const PcgArgs &a, int q) {
  return a.cgp + (size_t)q * 3 * a.nblk;
}
__device__ __forceinline__ double *cg_pap(const PcgArgs &a) { return a.cgp + (size_t)6 * a.nblk; }

/* this work-group's partials (thread 0 holds the work-group sums) */
__device__ __forceinline__ void cg_put3(double *slot, double x, double y, double z) {
  wg_sum3(x, y, z);
  if (threadIdx.x == 0) {
    slot[3 * blockIdx.x] = x;
    slot[3 * blockIdx.x + 1] = y;
    slot[3 * blockIdx.x + 2] = z;
  }
}

/* singular systems: per-work-group partial sums of the right-hand side (the
 * residual kernel that follows sums them for the zero-mean projection) */
__global__ __launch_bounds__(RED_BLOCK) void k_pcg_mean(PcgArgs a) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0, d1 = 0.0, d2 = 0.0;
  if (n < a.s.nx * a.s.ny) v = a.rho[n] * a.scale;
  wg_sum3(v, d1, d2);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = v;
}

/* r = f_elim - L x at free nodes; partials of (r,r), (r,z) (Jacobi z = r/diag;
 * the MG preconditioner's z comes later) and (f,f) into slot 0 */
__global__ __launch_bounds__(RED_BLOCK) void k_pcg_resid(PcgArgs a) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int nn = a.s.nx * a.s.ny;
  double rr = 0.0, rz = 0.0, ff = 0.0;
  /* singular system (flag set by an earlier kernel, the same for every
   * thread): the mean of the right-hand side from k_pcg_mean's partials */
  double mean = 0.0;
  if (a.iflag[2]) {
    double d1 = 0.0, d2 = 0.0;
    for (int k = threadIdx.x; k < a.nblk; k += RED_BLOCK) mean += a.partials[k];
    wg_sum3(mean, d1, d2);
    mean /= (double)nn;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.scal[S_MEAN] = mean;
  }
  if (n < nn) {
    int i = n % a.s.nx, j = n / a.s.nx;
    if (a.s.tags[n] == 0) {
      double f = a.rho[n] * a.scale;
      if (a.s.axial) f *= isk_rz_w(i, a.s.nx);
      if (a.iflag[2]) f -= mean;
      double diag;
      /* x holds the Dirichlet values, so L x already carries them */
      double Lx = apply_L(a.s, a.x, i, j, &diag);
      double fe = f + dirichlet_rhs(a.s, i, j);
      double r = f - Lx;
      a.r[n] = r;
      if (!a.mgpre) {
        float z = (float)(r / diag);
        a.z[n] = z;
        rz = r * (double)z;
      } else {
        a.f0[n] = (float)r;
        a.e0[n] = 0.0f;
      }
      rr = r * r;
      ff = fe * fe;
    } else {
      a.r[n] = 0.0;
      a.z[n] = 0.0f;
      if (a.mgpre) {
        a.f0[n] = 0.0f;
        a.e0[n] = 0.0f;
      }
    }
    a.p0[n] = 0.0f;
    a.p1[n] = 0.0f;
  }
  cg_put3(cg_slot(a, 0), rr, rz, ff);
}

/* MG-preconditioned CG: z = the V-cycle's finest correction, (r,z) partials
 * into slot q */
__global__ __launch_bounds__(RED_BLOCK) void k_cg_toz(PcgArgs a, int q) {
  if (a.iflag[0]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, d1 = 0.0, d2 = 0.0;
  if (n < a.s.nx * a.s.ny) {
    float z = a.e0[n];
    a.z[n] = z;
    rz = a.r[n] * (double)z;
  }
  wg_sum3(rz, d1, d2);
  if (threadIdx.x == 0) cg_slot(a, q)[3 * blockIdx.x + 1] = rz;
}

/* A(it): scalars of x_it, convergence, p_it = z + beta p_{it-1} (recomputed
 * at the neighbours), Ap = L p_it, (p, Ap) partials */
__global__ __launch_bounds__(RED_BLOCK) void k_cg_a(PcgArgs a, int it) {
  __shared__ int s_stop;
  if (threadIdx.x == 0) s_stop = a.iflag[0];
  double rr = 0.0, rz = 0.0, ff = 0.0;
  {
    const double *sl = cg_slot(a, it & 1);
    for (int k = threadIdx.x; k < a.nblk; k += RED_BLOCK) {
      rr += sl[3 * k];
      rz += sl[3 * k + 1];
      ff += sl[3 * k + 2];
    }
  }
  wg_sum3(rr, rz, ff); /* its barriers also publish s_stop */
  if (s_stop) return;
  const double sff = it == 0 ? ff : a.scal[S_FF];
  const double tol2 = it == 0 ? a.scal[S_RTOL2] * ff : a.scal[S_TOL2];
  const bool done = rr <= tol2 || it >= a.max_iter;
  double beta = 0.0;
  if (it > 0) {
    const double old = a.scal[S_RZ0 + ((it - 1) & 1)];
    beta = old != 0.0 ? rz / old : 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (it == 0) {
      a.scal[S_FF] = ff;
      a.scal[S_TOL2] = tol2;
    }
    a.scal[S_RR] = rr;
    a.scal[S_RELRES] = sff > 0.0 ? sqrt(rr / sff) : 0.0;
    a.scal[S_RZ0 + (it & 1)] = rz;
    a.scal[S_BETA] = beta;
    a.iflag[1] = it;
    if (done) a.iflag[0] = 1;
  }
  if (done) return;
  const float *pold = (it & 1) ? a.p0 : a.p1;
  float *pnew = (it & 1) ? a.p1 : a.p0;
  const float fb = (float)beta;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int nn = a.s.nx * a.s.ny;
  double pap = 0.0;
  if (n < nn) {
    int i = n % a.s.nx, j = n / a.s.nx;
    if (a.s.tags[n] == 0) {
#define PN(m) (a.s.tags[m] ? 0.0 : (double)fmaf(fb, pold[m], a.z[m]))
      double pc = PN(n);
      double acc = 0.0, diag = 0.0;
      if (a.s.axial) {
        const double cz = isk_rz_w(i, a.s.nx) * a.s.az;
        if (i > 0) { acc += ((double)i - 0.5) * PN(n - 1); diag += (double)i - 0.5; }
        if (i < a.s.nx - 1) { acc += ((double)i + 0.5) * PN(n + 1); diag += (double)i + 0.5; }
        if (j > 0) { acc += cz * PN(n - a.s.nx); diag += cz; }
        if (j < a.s.ny - 1) { acc += cz * PN(n + a.s.nx); diag += cz; }
      } else {
      if (i > 0) { acc += PN(n - 1); diag += 1.0; }
      if (i < a.s.nx - 1) { acc += PN(n + 1); diag += 1.0; }
      if (a.s.per_x) {
        if (i == 0) { acc += a.s.cxp * PN(n + a.s.nx - 1); diag += a.s.cxp; }
        if (i == a.s.nx - 1) { acc += a.s.cxp * PN(n - (a.s.nx - 1)); diag += a.s.cxp; }
      }
      if (a.s.dim == 2) {
        if (j > 0) { acc += PN(n - a.s.nx); diag += 1.0; }
        if (j < a.s.ny - 1) { acc += PN(n + a.s.nx); diag += 1.0; }
        if (a.s.per_y) {
          if (j == 0) { acc += a.s.cyp * PN(n + (a.s.ny - 1) * a.s.nx); diag += a.s.cyp; }
          if (j == a.s.ny - 1) { acc += a.s.cyp * PN(n - (a.s.ny - 1) * a.s.nx); diag += a.s.cyp; }
        }
      }
      }
#undef PN
      double ap = diag * pc - acc;
      pnew[n] = (float)pc;
      a.Ap[n] = ap;
      pap = pc * ap;
    } else {
      pnew[n] = 0.0f;
      a.Ap[n] = 0.0;
    }
  }
  double d1 = 0.0, d2 = 0.0;
  wg_sum3(pap, d1, d2);
  if (threadIdx.x == 0) cg_pap(a)[blockIdx.x] = pap;
}

/* B(it): alpha = (r,z) / (p,Ap); x += alpha p, r -= alpha Ap, z = M^-1 r;
 * partials of x_{it+1} into slot (it + 1) & 1 */
__global__ __launch_bounds__(RED_BLOCK) void k_cg_b(PcgArgs a, int it) {
  if (a.iflag[0]) return; /* set by an earlier kernel: the same for every thread */
  double pap = 0.0, d1 = 0.0, d2 = 0.0;
  {
    const double *pp = cg_pap(a);
    for (int k = threadIdx.x; k < a.nblk; k += RED_BLOCK) pap += pp[k];
  }
  wg_sum3(pap, d1, d2);
  const double rzc = a.scal[S_RZ0 + (it & 1)];
  const double alpha = pap != 0.0 ? rzc / pap : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.scal[S_PAP] = pap;
    a.scal[S_ALPHA] = alpha;
  }
  const float *pc = (it & 1) ? a.p1 : a.p0;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int nn = a.s.nx * a.s.ny;
  double rr = 0.0, rz = 0.0;
  if (n < nn && a.s.tags[n] == 0) {
    int i = n % a.s.nx, j = n / a.s.nx;
    double p = (double)pc[n];
    a.x[n] = __builtin_fma(alpha, p, a.x[n]);
    double r = __builtin_fma(-alpha, a.Ap[n], a.r[n]);
    a.r[n] = r;
    rr = r * r;
    if (!a.mgpre) {
      double diag = 0.0;
      if (a.s.axial) {
        const double cz = isk_rz_w(i, a.s.nx) * a.s.az;
        if (i > 0) diag += (double)i - 0.5;
        if (i < a.s.nx - 1) diag += (double)i + 0.5;
        if (j > 0) diag += cz;
        if (j < a.s.ny - 1) diag += cz;
      } else {
        if (i > 0) diag += 1.0;
        if (i < a.s.nx - 1) diag += 1.0;
        if (a.s.per_x && (i == 0 || i == a.s.nx - 1)) diag += a.s.cxp;
        if (a.s.dim == 2) {
          if (j > 0) diag += 1.0;
          if (j < a.s.ny - 1) diag += 1.0;
          if (a.s.per_y && (j == 0 || j == a.s.ny - 1)) diag += a.s.cyp;
        }
      }
      float z = (float)(r / diag);
      a.z[n] = z;
      rz = r * (double)z;
    } else {
      a.f0[n] = (float)r;
    }
  }
  if (a.mgpre && n < nn) a.e0[n] = 0.0f;
  cg_put3(cg_slot(a, (it + 1) & 1), rr, rz, 0.0);
}

/* final: mean removal for singular systems, copy to phi */
__global__ void k_pcg_out(PcgArgs a, double *phi, double mean_sub) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < a.s.nx * a.s.ny) phi[n] = a.x[n] - mean_sub;
}

/* ------------------------------------------------------------------------ */
/* spectral direct solve: phi = Sy ((Sy F Sx) / lambda) Sx                   */
/* ------------------------------------------------------------------------ */
/* F (interior, row-major [j-1][i-1]) = f_elim = rho scale + Dirichlet
 * neighbours (the time loop's deposit kernel writes the same values) */
__global__ void k_sp_pack(IskSpPack pk, const double *rho) {
  const int mx = pk.s.nx - 2, my = pk.s.ny - 2;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= mx * my) return;
  const int i = e % mx + 1, j = e / mx + 1;
  pk.F[e] = isk_sp_pack_value(pk, rho[j * pk.s.nx + i], i, j);
}

struct SpGemm {
  const double *A, *B, *L; /* L: elementwise factor (mode 1) */
  double *C;               /* modes 0, 1 */
  double *x;               /* modes 2, 3: node array, interior (i, j) = (col+1, row+1) */
  int M, N, K, nx, mode;
  Sten s;                  /* mode 2: the boundary ring's Dirichlet values */
};

typedef double f64x4 __attribute__((ext_vector_type(4)));

/* B operand: the matrix B (RES = 0), or the residual node array read at the
 * interior nodes, row k and column c = node (c+1, k+1) (RES = 1) */
template <int RES>
__device__ __forceinline__ double sp_bload(const SpGemm &g, int k, int col) {
  if (RES) return g.B[(size_t)(k + 1) * g.nx + (col + 1)];
  return g.B[(size_t)k * g.N + col];
}

/* C = A B (row-major FP64) on the matrix cores (v_mfma_f64_16x16x4f64) for
 * one 16 x 16 block of C: the K range is split over 4 waves (w = 0..3) and
 * each wave's partial block goes to red[w]; sp_tile_store sums them in wave
 * order (deterministic).  Operand lanes: A[row0 + l%16][k + l/16],
 * B[k + l/16][col0 + l%16]; result lane l holds C[row0 + l/16 + 4i][col0 +
 * l%16], i = 0..3 (rows interleaved).  All loads of a 64-wide k block are
 * issued before its MFMAs; the accumulation order is k ascending. */
template <int RES>
__device__ __forceinline__ void sp_tile_mfma(const SpGemm &g, int row0, int col0, int w, int lane,
                                             double (*red)[256]) {
  const int lr = lane & 15, kq = lane >> 4;
  const int arow = row0 + lr, bcol = col0 + lr;
  const int kper = ((g.K + 15) / 16) * 4;
  const int kbeg = w * kper;
  const int kend = min(g.K, kbeg + kper);
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = kbeg; k0 < kend; k0 += 64) {
    double av[16], bv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int kk = k0 + 4 * u + kq;
      av[u] = (arow < g.M && kk < kend) ? g.A[(size_t)arow * g.K + kk] : 0.0;
      bv[u] = (kk < kend && bcol < g.N) ? sp_bload<RES>(g, kk, bcol) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      if (k0 + 4 * u >= kend) break;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) red[w][(kq + 4 * i) * 16 + lr] = acc[i];
}

/* wave 0 of the group, after a barrier: sum the 4 partials, apply the mode */
__device__ __forceinline__ void sp_tile_store(const SpGemm &g, int row0, int col0, int lane,
                                              double (*red)[256]) {
  const int lr = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int e = (kq + 4 * i) * 16 + lr;
    const double v = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
    const int row = row0 + kq + 4 * i, col = col0 + lr;
    if (row >= g.M || col >= g.N) continue;
    const size_t o = (size_t)row * g.N + col;
    if (g.mode == 0) g.C[o] = v;
    else if (g.mode == 1) g.C[o] = v * g.L[o];
    else {
      const size_t n = (size_t)(row + 1) * g.nx + (col + 1);
      if (g.mode == 2) g.x[n] = v;
      else g.x[n] += v;
    }
  }
}

/* the boundary ring of x = its Dirichlet values */
__device__ __forceinline__ void sp_ring_setx(const SpGemm &g, int tid, int nthr) {
  const int nx = g.s.nx, ny = g.s.ny;
  const int ring = 2 * nx + 2 * (ny - 2);
  for (int q = tid; q < ring; q += nthr) {
    int n;
    if (q < nx) n = q;                                     /* row 0 */
    else if (q < 2 * nx) n = (ny - 1) * nx + (q - nx);     /* row ny-1 */
    else {
      const int r = q - 2 * nx;                            /* columns 0, nx-1 */
      n = (1 + (r >> 1)) * nx + ((r & 1) ? nx - 1 : 0);
    }
    const int t = g.s.tags[n];
    g.x[n] = t ? g.s.tagval[t] : 0.0;
  }
}

/* one block of C per work-group; mode 2 (first pass, x = result) also sets
 * the boundary ring of x (work-group (0, 0)) */
__global__ __launch_bounds__(256) void k_sp_gemm(SpGemm g) {
  __shared__ double red[4][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row0 = blockIdx.y * 16, col0 = blockIdx.x * 16;
  sp_tile_mfma<0>(g, row0, col0, w, lane, red);
  __syncthreads();
  if (w == 0) sp_tile_store(g, row0, col0, lane, red);
  if (g.mode == 2 && blockIdx.x == 0 && blockIdx.y == 0) sp_ring_setx(g, threadIdx.x, 256);
}

static SpGemm sp_desc(iskra_ctx *c, const PcgArgs &a, const double *A, const double *B,
                      const double *L, double *C, int M, int N, int K, int mode) {
  SpGemm g;
  g.A = A;
  g.B = B;
  g.L = L;
  g.C = C;
  g.x = a.x;
  g.M = M;
  g.N = N;
  g.K = K;
  g.nx = c->g.nx;
  g.mode = mode;
  g.s = a.s;
  return g;
}

/* the four GEMMs of one transform pass: x = L^-1 F (first, with the
 * boundary ring) or x += L^-1 r (correction, r read at the interior nodes) */
static void sp_pass_desc(iskra_ctx *c, const PcgArgs &a, int correction, SpGemm g[4]) {
  IskPoisson &p = c->ps;
  const int mx = p.sp_nx, my = p.sp_ny;
  g[0] = sp_desc(c, a, p.sp_Sy, correction ? a.r : p.sp_F, nullptr, p.sp_T1, my, mx, my, 0);
  g[1] = sp_desc(c, a, p.sp_T1, p.sp_Sx, p.sp_Li, p.sp_T2, my, mx, mx, 1);
  g[2] = sp_desc(c, a, p.sp_Sy, p.sp_T2, nullptr, p.sp_T1, my, mx, my, 0);
  g[3] = sp_desc(c, a, p.sp_T1, p.sp_Sx, nullptr, nullptr, my, mx, mx, correction ? 3 : 2);
}

/* residual of node n (free nodes: r = f - L x, contributions to ||r||^2 and
 * ||f_elim||^2; Dirichlet nodes: r = 0); also phi = x and, with E, the field
 * of node n from x (the arithmetic of efield_body) */
__device__ __forceinline__ void sp_resid_node(const PcgArgs &a, int n, float *E, double dx,
                                              double dy, double *rr, double *ff) {
  const int i = n % a.s.nx, j = n / a.s.nx;
  if (a.s.tags[n] == 0) {
    const double f = a.rho[n] * a.scale;
    const double Lx = apply_L(a.s, a.x, i, j, nullptr);
    const double fe = f + dirichlet_rhs(a.s, i, j);
    const double r = f - Lx;
    a.r[n] = r;
    *rr += r * r;
    *ff += fe * fe;
  } else {
    a.r[n] = 0.0;
  }
  a.phi_out[n] = a.x[n];
  if (E) efield_node(a.x, E, 2, a.s.nx, a.s.ny, dx, dy, n);
}

/* r = f_elim - L x; phi = x; with E != null also E = -grad phi.  Each
 * work-group writes its sums of r^2 and f_elim^2 to partials[2b], [2b+1]
 * (summed in a fixed order by k_sp_correct, which runs next: no device-wide
 * fence and counter here) */
__global__ __launch_bounds__(RED_BLOCK) void k_sp_resid(PcgArgs a, float *E, double dx,
                                                        double dy) {
  __shared__ double ra[RED_BLOCK], rb[RED_BLOCK];
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, ff = 0.0;
  if (n < a.s.nx * a.s.ny) sp_resid_node(a, n, E, dx, dy, &rr, &ff);
  ra[threadIdx.x] = rr;
  rb[threadIdx.x] = ff;
  __syncthreads();
  for (int off = RED_BLOCK / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      ra[threadIdx.x] += ra[threadIdx.x + off];
      rb[threadIdx.x] += rb[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.partials[2 * blockIdx.x] = ra[0];
    a.partials[2 * blockIdx.x + 1] = rb[0];
  }
}

/* A correction pass in one work-group of 16 waves (4 groups of 4 waves, each
 * group computing 16 x 16 blocks with the same tile code as k_sp_gemm, so the
 * results are those of the multi-work-group GEMMs), then the residual, phi and
 * E of the corrected x.  It returns at once when iflag[IF_SP_DONE] is set,
 * which is the usual case: one transform pass is exact to FP64 round-off, and
 * a correction is only needed when rtol is below that (then it takes about a
 * millisecond at 257^2).  second != 0: a miss after the correction sets
 * iflag[IF_SP_FAIL]. */
#define SPC_THREADS 1024
struct SpCorrect {
  SpGemm g[4];
  PcgArgs a;
  double rtol, dx, dy;
  float *E;
  int second;
  int nparts; /* > 0: first sum k_sp_resid's partials (the first pass's residual) */
};

/* fixed-order sum over the work-group of two values per thread; every
 * thread gets the totals */
__device__ __forceinline__ void spc_sum2(double *x, double *y, double (*sred)[SPC_THREADS / 64]) {
  double rr = *x, ff = *y;
  for (int off = 32; off > 0; off >>= 1) {
    rr += __shfl_xor(rr, off);
    ff += __shfl_xor(ff, off);
  }
  if ((threadIdx.x & 63) == 0) {
    sred[0][threadIdx.x >> 6] = rr;
    sred[1][threadIdx.x >> 6] = ff;
  }
  __syncthreads();
  double srr = 0.0, sff = 0.0;
  for (int k = 0; k < SPC_THREADS / 64; ++k) {
    srr += sred[0][k];
    sff += sred[1][k];
  }
  __syncthreads();
  *x = srr;
  *y = sff;
}

__global__ __launch_bounds__(SPC_THREADS) void k_sp_correct(SpCorrect c) {
  __shared__ double red[4][4][256];
  __shared__ double sred[2][SPC_THREADS / 64];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6, grp = wv >> 2, w = wv & 3;
  bool done;
  if (c.nparts > 0) {
    double rr = 0.0, ff = 0.0;
    for (int k = tid; k < c.nparts; k += SPC_THREADS) {
      rr += c.a.partials[2 * k];
      ff += c.a.partials[2 * k + 1];
    }
    spc_sum2(&rr, &ff, sred);
    const double rel = ff > 0.0 ? sqrt(rr / ff) : 0.0;
    done = rel <= c.rtol;
    if (tid == 0) {
      c.a.scal[S_RELRES] = rel;
      c.a.iflag[IF_SP_DONE] = done;
      c.a.iflag[IF_SP_PASSES] = 1;
    }
  } else {
    done = c.a.iflag[IF_SP_DONE] != 0;
  }
  if (done) return;
  for (int st = 0; st < 4; ++st) {
    const SpGemm &g = c.g[st];
    const int tx = (g.N + 15) / 16, ntile = tx * ((g.M + 15) / 16);
    for (int t0 = 0; t0 < ntile; t0 += 4) {
      const int t = t0 + grp;
      const bool on = t < ntile;
      const int row0 = (t / tx) * 16, col0 = (t % tx) * 16;
      if (on) {
        if (st == 0) sp_tile_mfma<1>(g, row0, col0, w, lane, red[grp]);
        else sp_tile_mfma<0>(g, row0, col0, w, lane, red[grp]);
      }
      __syncthreads();
      if (on && w == 0) sp_tile_store(g, row0, col0, lane, red[grp]);
      __syncthreads();
    }
    /* the next stage reads what this one wrote (global memory) */
    __threadfence();
    __syncthreads();
  }
  const int nn = c.a.s.nx * c.a.s.ny;
  double rr = 0.0, ff = 0.0;
  for (int n = tid; n < nn; n += SPC_THREADS) sp_resid_node(c.a, n, c.E, c.dx, c.dy, &rr, &ff);
  spc_sum2(&rr, &ff, sred);
  if (tid == 0) {
    const double rel = ff > 0.0 ? sqrt(rr / ff) : 0.0;
    c.a.scal[S_RELRES] = rel;
    const int ok = rel <= c.rtol;
    c.a.iflag[IF_SP_DONE] = ok;
    c.a.iflag[IF_SP_PASSES] += 1;
    if (c.second && !ok) c.a.iflag[IF_SP_FAIL] = 1;
  }
}

/* ------------------------------------------------------------------------ */
/* multigrid                                                                 */
/* ------------------------------------------------------------------------ */
struct MgLevel {
  float *e, *f;
  int nx, ny;
};

__global__ void k_mg_rbgs(MgLevel L, int color, const int *flag) {
  if (flag[0]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int half = (L.nx * L.ny + 1) / 2;
  if (n >= half) return;
  /* map n to the n-th node of this colour in row-major order */
  int k = 2 * n;
  int j = k / L.nx;
  int i = k % L.nx;
  if (((i + j) & 1) != color) {
    ++i;
    if (i >= L.nx) { i -= L.nx; ++j; }
  }
  if (((i + j) & 1) != color) return;
  if (i <= 0 || j <= 0 || i >= L.nx - 1 || j >= L.ny - 1) return;
  int m = j * L.nx + i;
  float s = (L.e[m - 1] + L.e[m + 1]) + (L.e[m - L.nx] + L.e[m + L.nx]);
  L.e[m] = (L.f[m] + s) * 0.25f;
}

/* coarse rhs = 4 * full-weighting(fine residual); coarse e = 0 */
__global__ void k_mg_restrict(MgLevel F, MgLevel C, const int *flag) {
  if (flag[0]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= C.nx * C.ny) return;
  int I = n % C.nx, J = n / C.nx;
  C.e[n] = 0.0f;
  if (I <= 0 || J <= 0 || I >= C.nx - 1 || J >= C.ny - 1) {
    C.f[n] = 0.0f;
    return;
  }
  int i = 2 * I, j = 2 * J;
  float acc[3][3];
#pragma unroll
  for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
    for (int di = -1; di <= 1; ++di) {
      int m = (j + dj) * F.nx + (i + di);
      float s = (F.e[m - 1] + F.e[m + 1]) + (F.e[m - F.nx] + F.e[m + F.nx]);
      acc[dj + 1][di + 1] = F.f[m] - (4.0f * F.e[m] - s);
    }
  float edge = (acc[0][1] + acc[2][1]) + (acc[1][0] + acc[1][2]);
  float corner = (acc[0][0] + acc[0][2]) + (acc[2][0] + acc[2][2]);
  C.f[n] = (4.0f * acc[1][1] + 2.0f * edge + corner) * 0.25f;
}

__global__ void k_mg_prolong(MgLevel C, MgLevel F, const int *flag) {
  if (flag[0]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= F.nx * F.ny) return;
  int i = n % F.nx, j = n / F.nx;
  if (i <= 0 || j <= 0 || i >= F.nx - 1 || j >= F.ny - 1) return;
  int I = i >> 1, J = j >> 1;
  const float *e = C.e;
  int m = J * C.nx + I;
  float v;
  if (!(i & 1) && !(j & 1)) v = e[m];
  else if ((i & 1) && !(j & 1)) v = 0.5f * (e[m] + e[m + 1]);
  else if (!(i & 1) && (j & 1)) v = 0.5f * (e[m] + e[m + C.nx]);
  else v = 0.25f * ((e[m] + e[m + 1]) + (e[m + C.nx] + e[m + C.nx + 1]));
  F.e[n] += v;
}

/* levels [cl, nlev) as one V-cycle inside one work-group (LDS resident) */
struct MgCoarseArgs {
  float *e_out, *f_in;
  int nlev_c;
  int nx[12], ny[12], off[12];
  int total;
  const int *flag;
  /* decide != 0: this kernel takes the cycle's convergence verdict */
  int decide, nparts, dfirst, max_iter;
  const double *partials;
  double *scal;
  int *iflag;
};

/* defect-correction multigrid: the convergence verdict on the FP64 residual
 * of x, taken by the kernel after the one that wrote the per-work-group
 * (r,r), (f,f) partials, in every work-group alike (the same sums in the
 * same order give the same bits); work-group 0 records it.  Returns true
 * when the cycle stops here.  The iteration limit is applied by work-group 0
 * alone through the flag, which the later kernels of the cycle read. */
template <int NT>
__device__ bool mg_verdict(const double *partials, int nparts, int first, double *scal,
                           int *iflag, int max_iter) {
  double rr = 0.0, ff = 0.0, d = 0.0;
  for (int k = threadIdx.x; k < nparts; k += NT) {
    rr += partials[2 * k];
    ff += partials[2 * k + 1];
  }
  wg_sum3<NT>(rr, ff, d);
  const double sff = first ? ff : scal[S_FF];
  const double tol2 = first ? scal[S_RTOL2] * ff : scal[S_TOL2];
  const bool done = rr <= tol2;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    if (first) {
      scal[S_FF] = ff;
      scal[S_TOL2] = tol2;
    }
    scal[S_RR] = rr;
    scal[S_RELRES] = sff > 0.0 ? sqrt(rr / sff) : 0.0;
    const int it = iflag[1];
    if (done || it >= max_iter) iflag[0] = 1;
    else iflag[1] = it + 1;
  }
  return done;
}

/* threads as a 32 x 32 grid; interior nodes of one colour, no divisions */
__device__ __forceinline__ void lds_rbgs(float *e, const float *f, int nx, int ny, int color) {
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = 1 + ty; j < ny - 1; j += 32) {
    int i0 = 2 - ((color + j) & 1);
    for (int i = i0 + 2 * tx; i < nx - 1; i += 64) {
      int m = j * nx + i;
      float s = (e[m - 1] + e[m + 1]) + (e[m - nx] + e[m + nx]);
      e[m] = (f[m] + s) * 0.25f;
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(1024) void k_mg_coarse(MgCoarseArgs a) {
  if (a.decide) {
    if (mg_verdict<1024>(a.partials, a.nparts, a.dfirst, a.scal, a.iflag, a.max_iter)) return;
  } else if (a.flag[0]) {
    return;
  }
  extern __shared__ float sm[];
  float *E = sm;
  float *Fv = sm + a.total;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  int n0 = a.nx[0] * a.ny[0];
  for (int k = threadIdx.x; k < a.total; k += blockDim.x) {
    E[k] = 0.0f;
    Fv[k] = k < n0 ? a.f_in[k] : 0.0f;
  }
  __syncthreads();
  /* down */
  for (int l = 0; l < a.nlev_c - 1; ++l) {
    float *e = E + a.off[l];
    float *f = Fv + a.off[l];
    int nx = a.nx[l], ny = a.ny[l];
    lds_rbgs(e, f, nx, ny, 0);
    lds_rbgs(e, f, nx, ny, 1);
    lds_rbgs(e, f, nx, ny, 0);
    lds_rbgs(e, f, nx, ny, 1);
    float *fc = Fv + a.off[l + 1];
    int cx = a.nx[l + 1], cy = a.ny[l + 1];
    for (int J = 1 + ty; J < cy - 1; J += 32)
      for (int I = 1 + tx; I < cx - 1; I += 32) {
        int i = 2 * I, j = 2 * J;
        float acc[3][3];
        for (int dj = -1; dj <= 1; ++dj)
          for (int di = -1; di <= 1; ++di) {
            int m = (j + dj) * nx + (i + di);
            float s = (e[m - 1] + e[m + 1]) + (e[m - nx] + e[m + nx]);
            acc[dj + 1][di + 1] = f[m] - (4.0f * e[m] - s);
          }
        float edge = (acc[0][1] + acc[2][1]) + (acc[1][0] + acc[1][2]);
        float corner = (acc[0][0] + acc[0][2]) + (acc[2][0] + acc[2][2]);
        fc[J * cx + I] = (4.0f * acc[1][1] + 2.0f * edge + corner) * 0.25f;
      }
    __syncthreads();
  }
  /* coarsest: exact for one interior node, otherwise Gauss-Seidel sweeps */
  {
    int l = a.nlev_c - 1;
    float *e = E + a.off[l];
    float *f = Fv + a.off[l];
    int nx = a.nx[l], ny = a.ny[l];
    if (nx == 3 && ny == 3) {
      if (threadIdx.x == 0) e[4] = 0.25f * f[4];
      __syncthreads();
    } else {
      for (int it = 0; it < 16; ++it) {
        lds_rbgs(e, f, nx, ny, 0);
        lds_rbgs(e, f, nx, ny, 1);
      }
    }
  }
  /* up */
  for (int l = a.nlev_c - 2; l >= 0; --l) {
    float *e = E + a.off[l];
    float *f = Fv + a.off[l];
    float *ec = E + a.off[l + 1];
    int nx = a.nx[l], ny = a.ny[l], cx = a.nx[l + 1];
    for (int j = 1 + ty; j < ny - 1; j += 32)
      for (int i = 1 + tx; i < nx - 1; i += 32) {
        int I = i >> 1, J = j >> 1;
        int m = J * cx + I;
        float v;
        if (!(i & 1) && !(j & 1)) v = ec[m];
        else if ((i & 1) && !(j & 1)) v = 0.5f * (ec[m] + ec[m + 1]);
        else if (!(i & 1) && (j & 1)) v = 0.5f * (ec[m] + ec[m + cx]);
        else v = 0.25f * ((ec[m] + ec[m + 1]) + (ec[m + cx] + ec[m + cx + 1]));
        e[j * nx + i] += v;
      }
    __syncthreads();
    lds_rbgs(e, f, nx, ny, 1);
    lds_rbgs(e, f, nx, ny, 0);
    lds_rbgs(e, f, nx, ny, 1);
    lds_rbgs(e, f, nx, ny, 0);
  }
  for (int k = threadIdx.x; k < n0; k += blockDim.x) a.e_out[k] = E[k];
}

/* ------------------------------------------------------------------------ */
/* fused fine-level multigrid kernels (temporal blocking in LDS)            */
/* ------------------------------------------------------------------------ */
/* A work-group owns CT x CT coarse nodes and the 2CT x 2CT fine nodes starting
 * at twice their index; it loads a fine tile with a halo of H nodes, runs the
 * four red-black half-sweeps of one smoothing stage there (each half-sweep
 * leaves one fewer halo ring exact), and writes only owned nodes, so the
 * result equals the global red-black Gauss-Seidel sweep. */
#define MGT_CT 8
#define MGT_H 6
#define MGT_S (2 * MGT_CT + 2 * MGT_H)

struct MgFused {
  MgLevel F, C;
  const int *flag;
  /* finest level of the FP64 defect correction */
  int mode_in;  /* down: 0 = f from F.f, 1 = f = dx^2 rho/eps0 - L x (FP64) */
  int mode_out; /* up: 0 = write F.e, 1 = x += e (FP64) */
  Sten s;
  double *x;
  const double *rho;
  double scale;
  double *scal, *partials;
  int *iflag;
  int first, max_iter;
  /* decide != 0: this kernel takes the cycle's convergence verdict from the
   * nparts partials of the finest residual (dfirst: first cycle of the solve) */
  int decide, nparts, dfirst;
};

__device__ __forceinline__ void tile_sweep(float *se, const float *sf, int ox0, int oy0, int nx,
                                           int ny, int color) {
  for (int k = threadIdx.x; k < MGT_S * MGT_S; k += blockDim.x) {
    int lj = k / MGT_S, li = k - lj * MGT_S;
    int gi = ox0 + li, gj = oy0 + lj;
    if (li < 1 || lj < 1 || li > MGT_S - 2 || lj > MGT_S - 2) continue;
    if (gi < 1 || gj < 1 || gi > nx - 2 || gj > ny - 2) continue;
    if (((gi + gj) & 1) != color) continue;
    float sum = (se[k - 1] + se[k + 1]) + (se[k - MGT_S] + se[k + MGT_S]);
    se[k] = (sf[k] + sum) * 0.25f;
  }
  __syncthreads();
}

/* pre-smooth from e = 0, restrict the residual, write owned e (and f) */
__global__ __launch_bounds__(256) void k_mgf_down(MgFused a) {
  if (a.decide) {
    if (mg_verdict<256>(a.partials, a.nparts, a.dfirst, a.scal, a.iflag, a.max_iter)) return;
  } else if (!a.first && a.flag[0]) {
    return;
  }
  __shared__ float se[MGT_S * MGT_S], sf[MGT_S * MGT_S];
  const int nx = a.F.nx, ny = a.F.ny;
  const int I0 = blockIdx.x * MGT_CT, J0 = blockIdx.y * MGT_CT;
  const int ox0 = 2 * I0 - MGT_H, oy0 = 2 * J0 - MGT_H;
  double rr = 0.0, ff = 0.0;
  for (int k = threadIdx.x; k < MGT_S * MGT_S; k += blockDim.x) {
    int lj = k / MGT_S, li = k - lj * MGT_S;
    int gi = ox0 + li, gj = oy0 + lj;
    bool interior = gi >= 1 && gj >= 1 && gi <= nx - 2 && gj <= ny - 2;
    float f = 0.0f;
    if (interior) {
      int m = gj * nx + gi;
      if (a.mode_in == 1) {
        double fs = a.rho[m] * a.scale;
        double r = fs - apply_L(a.s, a.x, gi, gj, nullptr);
        f = (float)r;
        bool owned = li >= MGT_H && lj >= MGT_H && li < MGT_H + 2 * MGT_CT &&
                     lj < MGT_H + 2 * MGT_CT;
        if (owned) {
          double fe = fs + dirichlet_rhs(a.s, gi, gj);
          rr += r * r;
          ff += fe * fe;
        }
      } else {
        f = a.F.f[m];
      }
    }
    sf[k] = f;
    se[k] = 0.0f;
  }
  __syncthreads();
  tile_sweep(se, sf, ox0, oy0, nx, ny, 0);
  tile_sweep(se, sf, ox0, oy0, nx, ny, 1);
  tile_sweep(se, sf, ox0, oy0, nx, ny, 0);
  tile_sweep(se, sf, ox0, oy0, nx, ny, 1);
  /* residual at fine nodes -> full weighting -> coarse rhs (x4) */
  const int cx = a.C.nx, cy = a.C.ny;
  for (int q = threadIdx.x; q < MGT_CT * MGT_CT; q += blockDim.x) {
    int I = I0 + (q % MGT_CT), J = J0 + (q / MGT_CT);
    if (I >= cx || J >= cy) continue;
    int n = J * cx + I;
    a.C.e[n] = 0.0f;
    if (I <= 0 || J <= 0 || I >= cx - 1 || J >= cy - 1) {
      a.C.f[n] = 0.0f;
      continue;
    }
    int lc = (2 * J - oy0) * MGT_S + (2 * I - ox0);
    float acc[3][3];
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int di = -1; di <= 1; ++di) {
        int k = lc + dj * MGT_S + di;
        float sum = (se[k - 1] + se[k + 1]) + (se[k - MGT_S] + se[k + MGT_S]);
        acc[dj + 1][di + 1] = sf[k] - (4.0f * se[k] - sum);
      }
    float edge = (acc[0][1] + acc[2][1]) + (acc[1][0] + acc[1][2]);
    float corner = (acc[0][0] + acc[0][2]) + (acc[2][0] + acc[2][2]);
    a.C.f[n] = (4.0f * acc[1][1] + 2.0f * edge + corner) * 0.25f;
  }
  /* owned fine nodes: smoothed e (and, at the finest level, f) */
  for (int q = threadIdx.x; q < 4 * MGT_CT * MGT_CT; q += blockDim.x) {
    int li = MGT_H + (q % (2 * MGT_CT)), lj = MGT_H + (q / (2 * MGT_CT));
    int gi = ox0 + li, gj = oy0 + lj;
    if (gi >= nx || gj >= ny) continue;
    int m = gj * nx + gi;
    int k = lj * MGT_S + li;
    a.F.e[m] = se[k];
    if (a.mode_in == 1) a.F.f[m] = sf[k];
  }
  if (a.mode_in == 1) {
    /* partials of the residual of x; the next kernel takes the verdict */
    double d = 0.0;
    wg_sum3<256>(rr, ff, d);
    if (threadIdx.x == 0) {
      const int bid = blockIdx.x + gridDim.x * blockIdx.y;
      a.partials[2 * bid] = rr;
      a.partials[2 * bid + 1] = ff;
    }
  }
}

/* e += P e_coarse, post-smooth, write owned e (or x += e at the finest level) */
__global__ __launch_bounds__(256) void k_mgf_up(MgFused a) {
  if (a.flag[0]) return;
  __shared__ float se[MGT_S * MGT_S], sf[MGT_S * MGT_S];
  const int nx = a.F.nx, ny = a.F.ny, cx = a.C.nx;
  const int I0 = blockIdx.x * MGT_CT, J0 = blockIdx.y * MGT_CT;
  const int ox0 = 2 * I0 - MGT_H, oy0 = 2 * J0 - MGT_H;
  for (int k = threadIdx.x; k < MGT_S * MGT_S; k += blockDim.x) {
    int lj = k / MGT_S, li = k - lj * MGT_S;
    int i = ox0 + li, j = oy0 + lj;
    bool interior = i >= 1 && j >= 1 && i <= nx - 2 && j <= ny - 2;
    float e = 0.0f, f = 0.0f;
    if (interior) {
      int m = j * nx + i;
      int I = i >> 1, J = j >> 1;
      const float *ec = a.C.e;
      int mc = J * cx + I;
      float v;
      if (!(i & 1) && !(j & 1)) v = ec[mc];
      else if ((i & 1) && !(j & 1)) v = 0.5f * (ec[mc] + ec[mc + 1]);
      else if (!(i & 1) && (j & 1)) v = 0.5f * (ec[mc] + ec[mc + cx]);
      else v = 0.25f * ((ec[mc] + ec[mc + 1]) + (ec[mc + cx] + ec[mc + cx + 1]));
      e = a.F.e[m] + v;
      f = a.F.f[m];
    }
    se[k] = e;
    sf[k] = f;
  }
  __syncthreads();
  tile_sweep(se, sf, ox0, oy0, nx, ny, 1);
  tile_sweep(se, sf, ox0, oy0, nx, ny, 0);
  tile_sweep(se, sf, ox0, oy0, nx, ny, 1);
  tile_sweep(se, sf, ox0, oy0, nx, ny, 0);
  for (int q = threadIdx.x; q < 4 * MGT_CT * MGT_CT; q += blockDim.x) {
    int li = MGT_H + (q % (2 * MGT_CT)), lj = MGT_H + (q / (2 * MGT_CT));
    int gi = ox0 + li, gj = oy0 + lj;
    if (gi >= nx || gj >= ny) continue;
    int m = gj * nx + gi;
    int k = lj * MGT_S + li;
    if (a.mode_out == 1) {
      if (gi >= 1 && gj >= 1 && gi <= nx - 2 && gj <= ny - 2) a.x[m] += (double)se[k];
    } else {
      a.F.e[m] = se[k];
    }
  }
}

/* FP64 defect-correction residual: f0 = r = f_elim - L x, e0 = 0 */
__global__ __launch_bounds__(RED_BLOCK) void k_mg_resid0(PcgArgs a, float *e0, int first) {
  if (!first && a.iflag[0]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int nn = a.s.nx * a.s.ny;
  double rr = 0.0, ff = 0.0;
  if (n < nn) {
    e0[n] = 0.0f;
    if (a.s.tags[n] == 0) {
      int i = n % a.s.nx, j = n / a.s.nx;
      double f = a.rho[n] * a.scale;
      double fe = f + dirichlet_rhs(a.s, i, j);
      double r = f - apply_L(a.s, a.x, i, j, nullptr);
      a.f0[n] = (float)r;
      rr = r * r;
      ff = fe * fe;
    } else {
      a.f0[n] = 0.0f;
    }
  }
  /* partials of the residual of x; the next kernel takes the verdict */
  double d = 0.0;
  wg_sum3<RED_BLOCK>(rr, ff, d);
  if (threadIdx.x == 0) {
    a.partials[2 * blockIdx.x] = rr;
    a.partials[2 * blockIdx.x + 1] = ff;
  }
}

__global__ void k_mg_update0(PcgArgs a, const float *e0) {
  if (a.iflag[0]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < a.s.nx * a.s.ny && a.s.tags[n] == 0) a.x[n] += (double)e0[n];
}

/* z = e0 after a preconditioning V-cycle */
__global__ void k_mg_toz(const float *e0, float *z, int nn, const int *flag) {
  if (flag[0]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < nn) z[n] = e0[n];
}

/* ------------------------------------------------------------------------ */
/* E = -grad(phi)                                                            */
/* ------------------------------------------------------------------------ */
struct EfieldArgs {
  const double *phi;
  float *E;
  int dim, nx, ny;
  double dx, dy;
};

__device__ __forceinline__ void efield_body(const double *phi, float *E, int dim, int nx, int ny,
                                            double dx, double dy, const int blk) {
  int n = blk * blockDim.x + threadIdx.x;
  if (n >= nx * ny) return;
  efield_node(phi, E, dim, nx, ny, dx, dy, n);
}

__global__ void k_efield(const double *phi, float *E, int dim, int nx, int ny, double dx,
                         double dy) {
  efield_body(phi, E, dim, nx, ny, dx, dy, blockIdx.x);
}

int isk_efield(iskra_ctx *c) {
  int nn = isk_nodes(c);
  hipLaunchKernelGGL(k_efield, dim3((nn + 255) / 256), dim3(256), 0, c->stream,
                     c->phi, c->E, c->g.dim, c->g.nx, c->g.ny, c->g.dx, c->g.dy);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

/* ------------------------------------------------------------------------ */
/* host side                                                                 */
/* ------------------------------------------------------------------------ */
/* ------------------------------------------------------------------------ */
/* BiCGStab for the systems with floating-electrode rows                     */
/* ------------------------------------------------------------------------ */
/* calculate_electric_potential (generalized_poisson.jl:372-378) when
 * apply_neumann has replaced rows (floating electrodes): the system is not
 * symmetric, so it is solved by right-preconditioned BiCGStab (van der
 * Vorst 1992) with the Jacobi preconditioner, FP64 throughout.  Unknowns are
 * the non-Dirichlet nodes; the sigma degrees of freedom are identity rows in
 * the reference (b = sigma), so they enter as known right-hand-side terms.
 * Like the CG kernels, every kernel sums the previous kernel's per-work-group
 * partials in block order itself (no device-wide reduction); iteration it:
 *   B1: beta, p = r + beta (p - omega v), v = A K^-1 p, partials (rh, v)
 *   B2: alpha = rho / (rh, v), s = r - alpha v
 *   B3: t = A K^-1 s, partials (t, s), (t, t)
 *   B4: omega = (t, s) / (t, t), x += alpha K^-1 p + omega K^-1 s,
 *       r = s - omega t, partials (rh, r), (r, r)
 * p and v alternate between two buffers by iteration parity. */
enum { BS_RHO0 = 0, BS_RHO1, BS_ALPHA, BS_OMEGA, BS_BB, BS_RR, BS_RELRES, BS_WORDS = 8 };

static Sten make_sten(iskra_ctx *c);

struct BiArgs {
  Sten s;
  const double *rho;
  double scale;
  const int *srow, *snb, *sdof;
  const double *sd, *sc, *ssig, *sigma, *dg;
  double *x, *r, *rh, *pp[2], *v[2], *sv, *tv;
  double *p0, *pa, *pt; /* partials: 3 x nblk (init / B4), nblk (B1), 2 x nblk (B3) */
  double *scal;
  int *iflag;
  int nblk, max_iter, warm;
  const double *phi_in;
  double rtol2;
};

/* row n of the unknowns' operator applied to get(m) (callers return 0 at
 * Dirichlet nodes, whose values are on the right-hand side) */
template <typename G>
__device__ __forceinline__ double bi_row(const BiArgs &a, int n, G get) {
  const int k = a.srow[n];
  if (k >= 0) {
    double acc = a.sd[k] * get(n);
    for (int q = 0; q < 2; ++q) {
      const int m = a.snb[2 * k + q];
      if (m >= 0) acc -= a.sc[2 * k + q] * get(m);
    }
    return acc;
  }
  const Sten &s = a.s;
  const int i = n % s.nx, j = n / s.nx;
  double acc = 0.0, diag = 0.0;
  if (i > 0) { acc += get(n - 1); diag += 1.0; }
  if (i < s.nx - 1) { acc += get(n + 1); diag += 1.0; }
  if (s.per_x) {
    if (i == 0) { acc += s.cxp * get(n + s.nx - 1); diag += s.cxp; }
    if (i == s.nx - 1) { acc += s.cxp * get(n - (s.nx - 1)); diag += s.cxp; }
  }
  if (s.dim == 2) {
    if (j > 0) { acc += get(n - s.nx); diag += 1.0; }
    if (j < s.ny - 1) { acc += get(n + s.nx); diag += 1.0; }
    if (s.per_y) {
      if (j == 0) { acc += s.cyp * get(n + (s.ny - 1) * s.nx); diag += s.cyp; }
      if (j == s.ny - 1) { acc += s.cyp * get(n - (s.ny - 1) * s.nx); diag += s.cyp; }
    }
  }
  return diag * get(n) - acc;
}

/* right-hand side of unknown n */
__device__ __forceinline__ double bi_rhs(const BiArgs &a, int n) {
  const double f = a.rho[n] * a.scale;
  const int k = a.srow[n];
  if (k >= 0) {
    double b = f + a.ssig[k] * a.sigma[a.sdof[k]];
    for (int q = 0; q < 2; ++q) {
      const int m = a.snb[2 * k + q];
      if (m >= 0 && a.s.tags[m] && a.srow[m] < 0) b += a.sc[2 * k + q] * a.s.tagval[a.s.tags[m]];
    }
    return b;
  }
  return f + dirichlet_rhs(a.s, n % a.s.nx, n / a.s.nx);
}

__global__ __launch_bounds__(RED_BLOCK) void k_bi_init(BiArgs a) {
  const int n = blockIdx.x * RED_BLOCK + threadIdx.x, nn = a.s.nx * a.s.ny;
  double rr = 0.0, bb = 0.0, rhor = 0.0;
  if (blockIdx.x == 0 && threadIdx.x < BS_WORDS) a.scal[threadIdx.x] = 0.0;
  if (n < nn) {
    const int t = a.s.tags[n];
    const bool unk = t == 0 || a.srow[n] >= 0;
    if (!unk) {
      a.x[n] = a.s.tagval[t];
      a.r[n] = a.rh[n] = 0.0;
    } else {
      const double x0 = a.warm ? a.phi_in[n] : 0.0;
      a.x[n] = x0;
      auto gx = [&](int m) -> double {
        return (a.s.tags[m] && a.srow[m] < 0) ? 0.0 : (a.warm ? a.phi_in[m] : 0.0);
      };
      const double b = bi_rhs(a, n);
      const double r = b - bi_row(a, n, gx);
      a.r[n] = r;
      a.rh[n] = r;
      rr = r * r;
      bb = b * b;
      rhor = r * r;
    }
    a.pp[1][n] = 0.0;
    a.v[1][n] = 0.0;
  }
  wg_sum3(rhor, bb, rr);
  if (threadIdx.x == 0) {
    a.p0[3 * blockIdx.x] = rhor;
    a.p0[3 * blockIdx.x + 1] = bb;
    a.p0[3 * blockIdx.x + 2] = rr;
  }
}

__global__ __launch_bounds__(RED_BLOCK) void k_bi_1(BiArgs a, int it) {
  __shared__ int s_stop;
  if (threadIdx.x == 0) s_stop = a.iflag[0];
  double rho = 0.0, bb = 0.0, rr = 0.0;
  for (int k = threadIdx.x; k < a.nblk; k += RED_BLOCK) {
    rho += a.p0[3 * k];
    bb += a.p0[3 * k + 1];
    rr += a.p0[3 * k + 2];
  }
  wg_sum3(rho, bb, rr);
  if (s_stop) return;
  if (it > 0) bb = a.scal[BS_BB];
  const bool done = rr <= a.rtol2 * bb || it >= a.max_iter || !(rr == rr);
  double beta = 0.0;
  const double omega = a.scal[BS_OMEGA];
  if (it > 0) {
    const double rho_old = a.scal[BS_RHO0 + ((it - 1) & 1)];
    beta = (rho_old != 0.0 && omega != 0.0) ? (rho / rho_old) * (a.scal[BS_ALPHA] / omega) : 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (it == 0) a.scal[BS_BB] = bb;
    a.scal[BS_RR] = rr;
    a.scal[BS_RELRES] = bb > 0.0 ? sqrt(rr / bb) : 0.0;
    a.scal[BS_RHO0 + (it & 1)] = rho;
    a.iflag[1] = it;
    if (done) a.iflag[0] = 1;
  }
  if (done) return;
  const double *pold = a.pp[(it + 1) & 1], *vold = a.v[(it + 1) & 1];
  double *pnew = a.pp[it & 1], *vnew = a.v[it & 1];
  const int n = blockIdx.x * RED_BLOCK + threadIdx.x, nn = a.s.nx * a.s.ny;
  double rv = 0.0, d1 = 0.0, d2 = 0.0;
  if (n < nn) {
    auto unk = [&](int m) { return a.s.tags[m] == 0 || a.srow[m] >= 0; };
    auto gp = [&](int m) -> double {
      if (!unk(m)) return 0.0;
      const double p = a.r[m] + beta * (pold[m] - omega * vold[m]);
      return p / a.dg[m];
    };
    if (unk(n)) {
      const double p = a.r[n] + beta * (pold[n] - omega * vold[n]);
      pnew[n] = p;
      const double v = bi_row(a, n, gp);
      vnew[n] = v;
      rv = a.rh[n] * v;
    } else {
      pnew[n] = 0.0;
      vnew[n] = 0.0;
    }
  }
  wg_sum3(rv, d1, d2);
  if (threadIdx.x == 0) a.pa[blockIdx.x] = rv;
}

__global__ __launch_bounds__(RED_BLOCK) void k_bi_2(BiArgs a, int it) {
  if (a.iflag[0]) return;
  double rv = 0.0, d1 = 0.0, d2 = 0.0;
  for (int k = threadIdx.x; k < a.nblk; k += RED_BLOCK) rv += a.pa[k];
  wg_sum3(rv, d1, d2);
  const double rho = a.scal[BS_RHO0 + (it & 1)];
  const double alpha = rv != 0.0 ? rho / rv : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.scal[BS_ALPHA] = alpha;
    if (rv == 0.0) a.iflag[3] = 1; /* breakdown */
  }
  const int n = blockIdx.x * RED_BLOCK + threadIdx.x, nn = a.s.nx * a.s.ny;
  if (n < nn) a.sv[n] = a.r[n] - alpha * a.v[it & 1][n];
}

__global__ __launch_bounds__(RED_BLOCK) void k_bi_3(BiArgs a, int it) {
  if (a.iflag[0]) return;
  (void)it;
  const int n = blockIdx.x * RED_BLOCK + threadIdx.x, nn = a.s.nx * a.s.ny;
  double ts = 0.0, tt = 0.0, d2 = 0.0;
  if (n < nn) {
    auto unk = [&](int m) { return a.s.tags[m] == 0 || a.srow[m] >= 0; };
    if (unk(n)) {
      auto gs = [&](int m) -> double { return unk(m) ? a.sv[m] / a.dg[m] : 0.0; };
      const double t = bi_row(a, n, gs);
      a.tv[n] = t;
      ts = t * a.sv[n];
      tt = t * t;
    } else {
      a.tv[n] = 0.0;
    }
  }
  wg_sum3(ts, tt, d2);
  if (threadIdx.x == 0) {
    a.pt[2 * blockIdx.x] = ts;
    a.pt[2 * blockIdx.x + 1] = tt;
  }
}

__global__ __launch_bounds__(RED_BLOCK) void k_bi_4(BiArgs a, int it) {
  if (a.iflag[0]) return;
  double ts = 0.0, tt = 0.0, d2 = 0.0;
  for (int k = threadIdx.x; k < a.nblk; k += RED_BLOCK) {
    ts += a.pt[2 * k];
    tt += a.pt[2 * k + 1];
  }
  wg_sum3(ts, tt, d2);
  const double omega = tt != 0.0 ? ts / tt : 0.0;
  const double alpha = a.scal[BS_ALPHA];
  if (blockIdx.x == 0 && threadIdx.x == 0) a.scal[BS_OMEGA] = omega;
  const int n = blockIdx.x * RED_BLOCK + threadIdx.x, nn = a.s.nx * a.s.ny;
  double rhor = 0.0, rr = 0.0, z = 0.0;
  if (n < nn && (a.s.tags[n] == 0 || a.srow[n] >= 0)) {
    const double dg = a.dg[n];
    a.x[n] += alpha * (a.pp[it & 1][n] / dg) + omega * (a.sv[n] / dg);
    const double r = a.sv[n] - omega * a.tv[n];
    a.r[n] = r;
    rhor = a.rh[n] * r;
    rr = r * r;
  }
  wg_sum3(rhor, z, rr);
  if (threadIdx.x == 0) {
    a.p0[3 * blockIdx.x] = rhor;
    a.p0[3 * blockIdx.x + 1] = 0.0;
    a.p0[3 * blockIdx.x + 2] = rr;
  }
}

__global__ void k_bi_out(const double *x, double *phi, int nn) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < nn) phi[n] = x[n];
}

int isk_bicgstab_solve(iskra_ctx *c, int warm, int *iters, double *relres) {
  IskPoisson &p = c->ps;
  const int nn = isk_nodes(c), nblk = (nn + RED_BLOCK - 1) / RED_BLOCK;
  BiArgs a;
  a.s = make_sten(c);
  a.rho = c->rho;
  a.scale = (c->g.dx * c->g.dx) / p.eps0;
  a.srow = p.d_srow;
  a.snb = p.d_snb;
  a.sdof = p.d_sdof;
  a.sd = p.d_sd;
  a.sc = p.d_sc;
  a.ssig = p.d_ssig;
  a.sigma = c->d_elec + EL_SIGMA;
  a.dg = p.d_dg;
  double *b = p.bi;
  a.x = p.x;
  a.r = b;
  a.rh = b + nn;
  a.pp[0] = b + 2 * (size_t)nn;
  a.pp[1] = b + 3 * (size_t)nn;
  a.v[0] = b + 4 * (size_t)nn;
  a.v[1] = b + 5 * (size_t)nn;
  a.sv = b + 6 * (size_t)nn;
  a.tv = b + 7 * (size_t)nn;
  a.p0 = b + 8 * (size_t)nn;
  a.pa = a.p0 + 3 * nblk;
  a.pt = a.pa + nblk;
  a.scal = a.pt + 2 * nblk;
  a.iflag = p.iflag;
  a.nblk = nblk;
  a.max_iter = p.max_iter;
  a.warm = warm == ISKRA_WARM_PREV ? 1 : 0;
  a.phi_in = c->phi;
  a.rtol2 = p.rtol * p.rtol;
  hipLaunchKernelGGL(k_pcg_init, dim3(1), dim3(64), 0, c->stream, p.scal, p.iflag, 0.0, 0);
  hipLaunchKernelGGL(k_bi_init, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a);
  ISK_CHECK_LAUNCH();
  int it = 0, hflag[4] = {0, 0, 0, 0};
  while (true) {
    for (int k = 0; k < 8; ++k, ++it) {
      hipLaunchKernelGGL(k_bi_1, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, it);
      hipLaunchKernelGGL(k_bi_2, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, it);
      hipLaunchKernelGGL(k_bi_3, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, it);
      hipLaunchKernelGGL(k_bi_4, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, it);
      ISK_CHECK_LAUNCH();
    }
    ISK_HIP(hipMemcpyAsync(hflag, p.iflag, sizeof(hflag), hipMemcpyDeviceToHost, c->stream));
    ISK_HIP(hipStreamSynchronize(c->stream));
    if (hflag[0]) break;
    if (it > p.max_iter + 16) {
      isk_set_error("poisson: BiCGStab stopped after %d host iterations without a verdict", it);
      return ISKRA_E_CONVERGE;
    }
  }
  double sc[BS_WORDS];
  ISK_HIP(hipMemcpy(sc, a.scal, sizeof(sc), hipMemcpyDeviceToHost));
  hipLaunchKernelGGL(k_bi_out, dim3((nn + 255) / 256), dim3(256), 0, c->stream, p.x, c->phi, nn);
  ISK_CHECK_LAUNCH();
  const int n_it = hflag[1];
  const double rel = sc[BS_RELRES];
  p.last_iters = n_it;
  p.last_relres = rel;
  c->stats.poisson_solves++;
  c->stats.poisson_iters += n_it;
  c->stats.last_iters = n_it;
  c->stats.last_relres = rel;
  if (iters) *iters = n_it;
  if (relres) *relres = rel;
  if (!(rel <= p.rtol)) {
    isk_set_error("poisson: BiCGStab reached relative residual %.3g > %.3g after %d iterations%s",
                  rel, p.rtol, n_it, hflag[3] ? " (breakdown)" : "");
    return ISKRA_E_CONVERGE;
  }
  return ISKRA_OK;
}

int isk_poisson_free(iskra_ctx *c) {
  IskPoisson &p = c->ps;
  void *ptrs[] = {p.d_tags, p.d_tag_val, p.d_tri_a, p.d_tri_cp, p.d_tri_inv,
                  p.d_work, p.x, p.r, p.Ap, p.z, p.p[0], p.p[1], p.scal,
                  p.partials, p.cgp, p.iflag, p.sp_Sx, p.sp_Sy, p.sp_Li,
                  p.sp_F, p.sp_T1, p.sp_T2, p.d_srow, p.d_snb, p.d_sdof, p.d_sd,
                  p.d_sc, p.d_ssig, p.bi, p.d_dg};
  for (void *q : ptrs) isk_free(q);
  p.d_srow = p.d_snb = p.d_sdof = nullptr;
  p.d_sd = p.d_sc = p.d_ssig = p.bi = p.d_dg = nullptr;
  p.sp_Sx = p.sp_Sy = p.sp_Li = p.sp_F = p.sp_T1 = p.sp_T2 = nullptr;
  p.sp_nx = p.sp_ny = 0;
  for (int l = 0; l < p.nlev; ++l) {
    isk_free(p.me[l]);
    isk_free(p.mf[l]);
  }
  p.d_tags = nullptr;
  p.d_tag_val = nullptr;
  p.d_tri_a = p.d_tri_cp = p.d_tri_inv = p.d_work = nullptr;
  p.x = p.r = p.Ap = nullptr;
  p.z = p.p[0] = p.p[1] = nullptr;
  p.scal = p.partials = p.cgp = nullptr;
  p.iflag = nullptr;
  p.nlev = 0;
  return ISKRA_OK;
}

static Sten make_sten(iskra_ctx *c) {
  Sten s;
  s.dim = c->g.dim;
  s.nx = c->g.nx;
  s.ny = c->g.ny;
  s.per_x = c->ps.per_x;
  s.per_y = c->ps.per_y;
  /* apply_periodic axis labels: wrapping i uses dy^2, wrapping j uses dx^2
   * (generalized_poisson.jl:291-323); in 1-D the ring uses dx^2 (:279-283);
   * all entries were divided by dx^2 at assembly (:65) */
  s.cxp = c->g.dim == 2 ? (c->g.dx * c->g.dx) / (c->g.dy * c->g.dy) : 1.0;
  s.cyp = 1.0;
  s.tags = c->ps.d_tags;
  s.tagval = c->ps.d_tag_val;
  s.axial = c->g.axial;
  s.az = c->g.axial ? (c->g.dx * c->g.dx) / (c->g.dy * c->g.dy) : 1.0;
  return s;
}

int isk_poisson_setup(iskra_ctx *c) {
  IskPoisson &p = c->ps;
  const int nn = isk_nodes(c);
  int ndir = 0, interior_tag = 0;
  for (int k = 0; k < nn; ++k) {
    if (p.tags[k]) {
      ++ndir;
      int i = k % c->g.nx, j = k / c->g.nx;
      bool edge = i == 0 || i == c->g.nx - 1 ||
                  (c->g.dim == 2 && (j == 0 || j == c->g.ny - 1));
      if (!edge) interior_tag = 1;
    }
  }
  p.singular = ndir == 0;
  if (c->g.axial) {
    /* the axisymmetric operator (generalized_poisson.jl:70-199) is solved by
     * CG on its symmetric weighted form only */
    if (p.kind != ISKRA_POISSON_PCG) {
      isk_set_error("poisson: an axial grid is solved by PCG only");
      return ISKRA_E_UNSUPPORTED;
    }
    if (p.per_x || p.per_y) {
      isk_set_error("poisson: no periodic link on an axial grid (the reference's "
                    "apply_periodic for RZ cannot run)");
      return ISKRA_E_UNSUPPORTED;
    }
    if (p.singular) {
      isk_set_error("poisson: an axial grid needs a Dirichlet node (the reference's "
                    "matrix is singular without one)");
      return ISKRA_E_UNSUPPORTED;
    }
  }
  if (isk_alloc((void **)&p.d_tags, nn)) return ISKRA_E_HIP;
  ISK_HIP(hipMemcpy(p.d_tags, p.tags.data(), nn, hipMemcpyHostToDevice));
  if (isk_alloc((void **)&p.d_tag_val, sizeof(double) * (ISK_MAX_TAGS + 1))) return ISKRA_E_HIP;
  ISK_HIP(hipMemcpy(p.d_tag_val, p.tag_val, sizeof(double) * (ISK_MAX_TAGS + 1),
                    hipMemcpyHostToDevice));
  if (p.kind == ISKRA_POISSON_DIRECT) {
    if (c->g.dim != 1) {
      isk_set_error("poisson: DIRECT is the 1-D tridiagonal solver; use PCG/MG in 2-D");
      return ISKRA_E_UNSUPPORTED;
    }
    if (p.per_x && !p.singular) {
      isk_set_error("poisson: periodic ring with Dirichlet nodes is not supported by DIRECT; use PCG");
      return ISKRA_E_UNSUPPORTED;
    }
    const int n = c->g.nx;
    std::vector<double> a(n, 0.0), d(n, 0.0), cc(n, 0.0), cp(n, 0.0), inv(n, 0.0);
    if (!p.singular) {
      for (int i = 0; i < n; ++i) {
        if (p.tags[i]) {
          d[i] = 1.0;
          continue;
        }
        a[i] = i > 0 ? 1.0 : 0.0;
        cc[i] = i < n - 1 ? 1.0 : 0.0;
        d[i] = -(a[i] + cc[i]);
      }
      /* elimination: inv_i = 1/(d_i - a_i cp_{i-1}), cp_i = c_i inv_i */
      double cprev = 0.0;
      for (int i = 0; i < n; ++i) {
        double den = d[i] - a[i] * cprev;
        if (den == 0.0) {
          isk_set_error("poisson: singular tridiagonal system (row %d)", i);
          return ISKRA_E_ARG;
        }
        inv[i] = 1.0 / den;
        cp[i] = cc[i] * inv[i];
        cprev = cp[i];
      }
    } else {
      /* ring, node 0 pinned: rows 1..n-1, each with two neighbours */
      double cprev = 0.0;
      for (int i = 1; i < n; ++i) {
        a[i] = i > 1 ? 1.0 : 0.0;
        cc[i] = i < n - 1 ? 1.0 : 0.0;
        d[i] = -2.0;
        double den = d[i] - a[i] * cprev;
        inv[i] = 1.0 / den;
        cp[i] = cc[i] * inv[i];
        cprev = cp[i];
      }
    }
    if (isk_alloc((void **)&p.d_tri_a, n * 8) || isk_alloc((void **)&p.d_tri_cp, n * 8) ||
        isk_alloc((void **)&p.d_tri_inv, n * 8) || isk_alloc((void **)&p.d_work, n * 8))
      return ISKRA_E_HIP;
    ISK_HIP(hipMemcpy(p.d_tri_a, a.data(), n * 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(p.d_tri_cp, cp.data(), n * 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(p.d_tri_inv, inv.data(), n * 8, hipMemcpyHostToDevice));
    return ISKRA_OK;
  }
  /* iterative solvers */
  int nblk = (nn + RED_BLOCK - 1) / RED_BLOCK;
  {
    /* the finest fused multigrid kernel reduces over one block per tile */
    int cx = (c->g.nx - 1) / 2 + 1, cy = c->g.dim == 2 ? (c->g.ny - 1) / 2 + 1 : 1;
    int tiles = ((cx + MGT_CT - 1) / MGT_CT) * ((cy + MGT_CT - 1) / MGT_CT);
    p.partials_blocks = std::max(nblk, tiles);
  }
  if (isk_alloc((void **)&p.x, nn * 8) || isk_alloc((void **)&p.r, nn * 8) ||
      isk_alloc((void **)&p.Ap, nn * 8) || isk_alloc((void **)&p.z, nn * 4) ||
      isk_alloc((void **)&p.p[0], nn * 4) || isk_alloc((void **)&p.p[1], nn * 4) ||
      isk_alloc((void **)&p.scal, S_WORDS * 8) ||
      isk_alloc((void **)&p.partials, (size_t)p.partials_blocks * 2 * 8) ||
      isk_alloc((void **)&p.cgp, (size_t)nblk * 7 * 8) ||
      isk_alloc((void **)&p.iflag, 64))
    return ISKRA_E_HIP;
  ISK_HIP(hipMemset(p.iflag, 0, 64));
  p.nlev = 0;
  if (!p.h_sd.empty()) {
    /* floating electrodes: apply_neumann rows, solved by BiCGStab */
    if (c->g.dim != 2 || c->g.axial ||
        (p.kind != ISKRA_POISSON_PCG && p.kind != ISKRA_POISSON_MGPCG)) {
      isk_set_error("poisson: floating electrodes need a 2-D Cartesian grid and the PCG kind "
                    "(their rows make the system non-symmetric: BiCGStab solves it)");
      return ISKRA_E_UNSUPPORTED;
    }
    const int ns = (int)p.h_sd.size();
    std::vector<double> dg(nn);
    for (int k = 0; k < nn; ++k) {
      if (p.h_srow[k] >= 0) {
        dg[k] = p.h_sd[p.h_srow[k]];
        continue;
      }
      const int i = k % c->g.nx, j = k / c->g.nx;
      double d = 0.0;
      if (i > 0) d += 1.0;
      if (i < c->g.nx - 1) d += 1.0;
      if (p.per_x && (i == 0 || i == c->g.nx - 1)) d += (c->g.dx * c->g.dx) / (c->g.dy * c->g.dy);
      if (j > 0) d += 1.0;
      if (j < c->g.ny - 1) d += 1.0;
      if (p.per_y && (j == 0 || j == c->g.ny - 1)) d += 1.0;
      dg[k] = d > 0.0 ? d : 1.0;
    }
    const size_t nbuf = (size_t)8 * nn + 6 * (size_t)nblk + 16;
    if (isk_alloc((void **)&p.d_srow, (size_t)nn * 4) ||
        isk_alloc((void **)&p.d_snb, (size_t)2 * ns * 4) ||
        isk_alloc((void **)&p.d_sdof, (size_t)ns * 4) || isk_alloc((void **)&p.d_sd, (size_t)ns * 8) ||
        isk_alloc((void **)&p.d_sc, (size_t)2 * ns * 8) ||
        isk_alloc((void **)&p.d_ssig, (size_t)ns * 8) ||
        isk_alloc((void **)&p.d_dg, (size_t)nn * 8) || isk_alloc((void **)&p.bi, nbuf * 8))
      return ISKRA_E_HIP;
    ISK_HIP(hipMemcpy(p.d_srow, p.h_srow.data(), (size_t)nn * 4, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(p.d_snb, p.h_snb.data(), (size_t)2 * ns * 4, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(p.d_sdof, p.h_sdof.data(), (size_t)ns * 4, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(p.d_sd, p.h_sd.data(), (size_t)ns * 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(p.d_sc, p.h_sc.data(), (size_t)2 * ns * 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(p.d_ssig, p.h_ssig.data(), (size_t)ns * 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(p.d_dg, dg.data(), (size_t)nn * 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemset(p.bi, 0, nbuf * 8));
    return ISKRA_OK;
  }
  if (p.kind == ISKRA_POISSON_SPECTRAL) {
    if (c->g.dim != 2 || p.per_x || p.per_y || interior_tag) {
      isk_set_error("poisson: the spectral solve needs a 2-D grid with Dirichlet walls only");
      return ISKRA_E_UNSUPPORTED;
    }
    for (int k = 0; k < nn; ++k) {
      int i = k % c->g.nx, j = k / c->g.nx;
      bool edge = i == 0 || i == c->g.nx - 1 || j == 0 || j == c->g.ny - 1;
      if (edge && !p.tags[k]) {
        isk_set_error("poisson: the spectral solve needs Dirichlet values on all four walls");
        return ISKRA_E_UNSUPPORTED;
      }
    }
    const int mx = c->g.nx - 2, my = c->g.ny - 2;
    if (mx < 1 || my < 1) {
      isk_set_error("poisson: the spectral solve needs interior nodes");
      return ISKRA_E_ARG;
    }
    /* S[a][b] = sqrt(2/(m+1)) sin(pi (a+1)(b+1)/(m+1)): symmetric, S S = I;
     * eigenvalues of the unit 5-point operator: (2-2cos pi p/(mx+1)) + (2-2cos pi q/(my+1)) */
    auto basis = [](int m, std::vector<double> &S) {
      S.assign((size_t)m * m, 0.0);
      const double c0 = sqrt(2.0 / (m + 1)), w = M_PI / (m + 1);
      for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b) {
          /* reduce (a+1)(b+1) modulo 2(m+1) so the sine argument stays small */
          long long k = (long long)(a + 1) * (b + 1) % (2ll * (m + 1));
          S[(size_t)a * m + b] = c0 * sin(w * (double)k);
        }
    };
    std::vector<double> Sx, Sy, Li((size_t)mx * my);
    basis(mx, Sx);
    basis(my, Sy);
    for (int q = 0; q < my; ++q)
      for (int pp = 0; pp < mx; ++pp) {
        double lx = 2.0 - 2.0 * cos(M_PI * (pp + 1) / (mx + 1));
        double ly = 2.0 - 2.0 * cos(M_PI * (q + 1) / (my + 1));
        Li[(size_t)q * mx + pp] = 1.0 / (lx + ly);
      }
    size_t m2 = (size_t)mx * my;
    if (isk_alloc((void **)&p.sp_Sx, (size_t)mx * mx * 8) ||
        isk_alloc((void **)&p.sp_Sy, (size_t)my * my * 8) ||
        isk_alloc((void **)&p.sp_Li, m2 * 8) || isk_alloc((void **)&p.sp_F, m2 * 8) ||
        isk_alloc((void **)&p.sp_T1, m2 * 8) || isk_alloc((void **)&p.sp_T2, m2 * 8))
      return ISKRA_E_HIP;
    ISK_HIP(hipMemcpy(p.sp_Sx, Sx.data(), (size_t)mx * mx * 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(p.sp_Sy, Sy.data(), (size_t)my * my * 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(p.sp_Li, Li.data(), m2 * 8, hipMemcpyHostToDevice));
    p.sp_nx = mx;
    p.sp_ny = my;
    return ISKRA_OK;
  }
  if (p.kind == ISKRA_POISSON_MG || p.kind == ISKRA_POISSON_MGPCG) {
    if (c->g.dim != 2 || p.per_x || p.per_y || interior_tag) {
      isk_set_error("poisson: multigrid needs a 2-D grid with Dirichlet walls only");
      return ISKRA_E_UNSUPPORTED;
    }
    /* every edge node must be Dirichlet */
    for (int k = 0; k < nn; ++k) {
      int i = k % c->g.nx, j = k / c->g.nx;
      bool edge = i == 0 || i == c->g.nx - 1 || j == 0 || j == c->g.ny - 1;
      if (edge && !p.tags[k]) {
        isk_set_error("poisson: multigrid needs Dirichlet values on all four walls");
        return ISKRA_E_UNSUPPORTED;
      }
    }
    if (fabs(c->g.dx - c->g.dy) > 1e-12 * c->g.dx) {
      /* the reference scales all links by 1/dx^2 (isotropic stencil): fine */
    }
    int nx = c->g.nx, ny = c->g.ny;
    int l = 0;
    p.ln[0] = nx;
    while (true) {
      if (isk_alloc((void **)&p.me[l], (size_t)nx * ny * 4) ||
          isk_alloc((void **)&p.mf[l], (size_t)nx * ny * 4))
        return ISKRA_E_HIP;
      p.ln[l] = nx | (ny << 16);
      ++l;
      if ((nx - 1) % 2 || (ny - 1) % 2 || nx <= 3 || ny <= 3 || l >= 12) break;
      nx = (nx - 1) / 2 + 1;
      ny = (ny - 1) / 2 + 1;
    }
    p.nlev = l;
    /* first level that fits the single work-group coarse solver */
    int cl = 0;
    while (cl < p.nlev) {
      int cx = p.ln[cl] & 0xFFFF, cy = p.ln[cl] >> 16;
      if (cx <= MG_COARSE_MAX && cy <= MG_COARSE_MAX) break;
      ++cl;
    }
    if (cl >= p.nlev) {
      isk_set_error("poisson: grid does not coarsen to <= %d nodes per axis", MG_COARSE_MAX);
      return ISKRA_E_UNSUPPORTED;
    }
    p.coarse_lev = cl;
  }
  return ISKRA_OK;
}

static PcgArgs make_pcg(iskra_ctx *c, int mgpre) {
  IskPoisson &p = c->ps;
  PcgArgs a;
  a.s = make_sten(c);
  a.rho = c->rho;
  a.x = p.x;
  a.r = p.r;
  a.Ap = p.Ap;
  a.z = p.z;
  a.p0 = p.p[0];
  a.p1 = p.p[1];
  a.cgp = p.cgp;
  a.nblk = (isk_nodes(c) + RED_BLOCK - 1) / RED_BLOCK;
  a.f0 = p.nlev ? p.mf[0] : nullptr;
  a.e0 = p.nlev ? p.me[0] : nullptr;
  a.scal = p.scal;
  a.partials = p.partials;
  a.iflag = p.iflag;
  a.scale = (c->g.dx * c->g.dx) / p.eps0;
  a.warm = 0;
  a.phi_in = c->phi;
  a.phi_out = c->phi;
  a.mgpre = mgpre;
  a.max_iter = p.max_iter;
  return a;
}

/* one V-cycle; mode_in0 = 1 makes the finest down-kernel compute the FP64
 * residual of x (defect correction, first = 1 on the solve's first cycle);
 * mode_out0 = 1 adds the correction to x. */
static int mg_vcycle(iskra_ctx *c, int mode_in0, int mode_out0, int first) {
  IskPoisson &p = c->ps;
  const int *flag = p.iflag;
  auto lev = [&](int l) {
    MgLevel L;
    L.e = p.me[l];
    L.f = p.mf[l];
    L.nx = p.ln[l] & 0xFFFF;
    L.ny = p.ln[l] >> 16;
    return L;
  };
  MgFused a;
  a.flag = flag;
  a.s = make_sten(c);
  a.x = p.x;
  a.rho = c->rho;
  a.scale = (c->g.dx * c->g.dx) / p.eps0;
  a.scal = p.scal;
  a.partials = p.partials;
  a.iflag = p.iflag;
  a.max_iter = p.max_iter;
  const int nn0 = (p.ln[0] & 0xFFFF) * (p.ln[0] >> 16);
  const int nblk0 = (nn0 + RED_BLOCK - 1) / RED_BLOCK;
  PcgArgs pa = make_pcg(c, 0);
  /* with mode_in0 = 1 the finest residual kernel writes per-work-group
   * partials and the kernel after it takes the convergence verdict: the
   * second down kernel, or the single-work-group kernel */
  int nparts = 0;
  a.decide = 0;
  a.nparts = 0;
  a.dfirst = first;
  if (p.coarse_lev == 0 && mode_in0 == 1) {
    /* the whole grid is solved by the single work-group kernel: FP64 residual here */
    hipLaunchKernelGGL(k_mg_resid0, dim3(nblk0), dim3(RED_BLOCK), 0, c->stream, pa, p.me[0], first);
    nparts = nblk0;
  }
  for (int l = 0; l < p.coarse_lev; ++l) {
    a.F = lev(l);
    a.C = lev(l + 1);
    a.mode_in = l == 0 ? mode_in0 : 0;
    a.first = l == 0 ? first : 0;
    a.decide = (l == 1 && nparts > 0) ? 1 : 0;
    a.nparts = nparts;
    dim3 g((a.C.nx + MGT_CT - 1) / MGT_CT, (a.C.ny + MGT_CT - 1) / MGT_CT);
    if (a.mode_in == 1 && (int)(g.x * g.y) > p.partials_blocks) {
      isk_set_error("poisson: reduction scratch holds %d blocks, kernel needs %d",
                    p.partials_blocks, (int)(g.x * g.y));
      return ISKRA_E_STATE;
    }
    hipLaunchKernelGGL(k_mgf_down, g, dim3(256), 0, c->stream, a);
    if (a.mode_in == 1) nparts = (int)(g.x * g.y);
  }
  {
    MgCoarseArgs ca;
    int tot = 0;
    ca.nlev_c = p.nlev - p.coarse_lev;
    for (int k = 0; k < ca.nlev_c; ++k) {
      int l = p.coarse_lev + k;
      ca.nx[k] = p.ln[l] & 0xFFFF;
      ca.ny[k] = p.ln[l] >> 16;
      ca.off[k] = tot;
      tot += ca.nx[k] * ca.ny[k];
    }
    ca.total = tot;
    ca.e_out = p.me[p.coarse_lev];
    ca.f_in = p.mf[p.coarse_lev];
    ca.flag = flag;
    /* the verdict here when no second down kernel took it */
    ca.decide = (nparts > 0 && p.coarse_lev <= 1) ? 1 : 0;
    ca.nparts = nparts;
    ca.dfirst = first;
    ca.max_iter = p.max_iter;
    ca.partials = p.partials;
    ca.scal = p.scal;
    ca.iflag = p.iflag;
    hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(1024), (size_t)tot * 2 * 4, c->stream, ca);
  }
  for (int l = p.coarse_lev - 1; l >= 0; --l) {
    a.F = lev(l);
    a.C = lev(l + 1);
    a.mode_out = l == 0 ? mode_out0 : 0;
    dim3 g((a.C.nx + MGT_CT - 1) / MGT_CT, (a.C.ny + MGT_CT - 1) / MGT_CT);
    hipLaunchKernelGGL(k_mgf_up, g, dim3(256), 0, c->stream, a);
  }
  if (p.coarse_lev == 0 && mode_out0 == 1)
    hipLaunchKernelGGL(k_mg_update0, dim3(nblk0), dim3(RED_BLOCK), 0, c->stream, pa, p.me[0]);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

static void tri_prepare(iskra_ctx *c, TriArgs *pt) {
  IskPoisson &p = c->ps;
  TriArgs &t = *pt;
  t.rho = c->rho;
  t.tags = p.d_tags;
  t.tagval = p.d_tag_val;
  t.a = p.d_tri_a;
  t.cp = p.d_tri_cp;
  t.inv = p.d_tri_inv;
  t.y = p.d_work;
  t.phi = c->phi;
  t.n = c->g.nx;
  t.scale = (c->g.dx * c->g.dx) / p.eps0;
  t.singular = p.singular;
}

int isk_poisson_solve(iskra_ctx *c, int warm, int *iters, double *relres) {
  return isk_poisson_solve_ex(c, warm, iters, relres, 0, nullptr, 0);
}

/* the spectral right-hand side packing, for the deposit kernel of the time
 * loop; returns 0 (and F = null) for the other solvers */
int isk_poisson_pack_desc(iskra_ctx *c, IskSpPack *pk) {
  IskPoisson &p = c->ps;
  pk->F = nullptr;
  if (p.kind != ISKRA_POISSON_SPECTRAL || !p.sp_F) return 0;
  pk->F = p.sp_F;
  pk->s = make_sten(c);
  pk->scale = (c->g.dx * c->g.dx) / p.eps0;
  return 1;
}

/* with_E != 0: a solver that can form E in the same pass does so and sets
 * *E_done (the spectral solve); otherwise *E_done stays 0 */
int isk_poisson_solve_ex(iskra_ctx *c, int warm, int *iters, double *relres, int with_E,
                         int *E_done, int packed) {
  IskPoisson &p = c->ps;
  if (E_done) *E_done = 0;
  const int nn = isk_nodes(c);
  if (p.kind == ISKRA_POISSON_DIRECT) {
    TriArgs t;
    tri_prepare(c, &t);
    hipLaunchKernelGGL(k_tri_solve, dim3(1), dim3(TRI_BLOCK), 0, c->stream, t);
    ISK_CHECK_LAUNCH();
    p.last_iters = 1;
    p.last_relres = 0.0;
    if (iters) *iters = 1;
    if (relres) *relres = 0.0;
    c->stats.poisson_solves++;
    c->stats.poisson_iters += 1;
    c->stats.last_iters = 1;
    c->stats.last_relres = 0.0;
    return ISKRA_OK;
  }
  const int nblk = (nn + RED_BLOCK - 1) / RED_BLOCK;
  if (p.kind == ISKRA_POISSON_SPECTRAL) {
    /* transform pass (F packed by the deposit kernel in the time loop, else
     * here; 4 GEMM launches, the last also sets the boundary ring of x),
     * FP64 residual (also phi, and E when asked), then
     * the correction kernel, which does nothing when the pass met rtol.
     * Without iters/relres (the time loop) one correction is enqueued and
     * a miss after it is reported by iskra_sync; with them the residual is
     * read back after each pass and up to 3 corrections are made. */
    PcgArgs a = make_pcg(c, 0);
    float *E = with_E ? c->E : nullptr;
    if (!packed) {
      IskSpPack pk;
      isk_poisson_pack_desc(c, &pk);
      const int m2 = p.sp_nx * p.sp_ny;
      hipLaunchKernelGGL(k_sp_pack, dim3((m2 + 255) / 256), dim3(256), 0, c->stream, pk, c->rho);
      ISK_CHECK_LAUNCH();
    }
    SpGemm g[4];
    sp_pass_desc(c, a, 0, g);
    for (int k = 0; k < 4; ++k) {
      hipLaunchKernelGGL(k_sp_gemm, dim3((g[k].N + 15) / 16, (g[k].M + 15) / 16), dim3(256), 0,
                         c->stream, g[k]);
      ISK_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_sp_resid, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, E, c->g.dx,
                       c->g.dy);
    ISK_CHECK_LAUNCH();
    SpCorrect sc;
    sp_pass_desc(c, a, 1, sc.g);
    sc.a = a;
    sc.rtol = p.rtol;
    sc.dx = c->g.dx;
    sc.dy = c->g.dy;
    sc.E = E;
    sc.nparts = nblk; /* the first call sums the residual of the first pass */
    sc.second = (!iters && !relres) ? 1 : 0;
    hipLaunchKernelGGL(k_sp_correct, dim3(1), dim3(SPC_THREADS), 0, c->stream, sc);
    ISK_CHECK_LAUNCH();
    int rounds = 1;
    double rel = 0.0;
    if (iters || relres) {
      sc.nparts = 0;
      while (true) {
        double scal[S_WORDS];
        int fl[8];
        ISK_HIP(hipMemcpyAsync(scal, p.scal, sizeof(scal), hipMemcpyDeviceToHost, c->stream));
        ISK_HIP(hipMemcpyAsync(fl, p.iflag, sizeof(fl), hipMemcpyDeviceToHost, c->stream));
        ISK_HIP(hipStreamSynchronize(c->stream));
        rel = scal[S_RELRES];
        rounds = fl[IF_SP_PASSES];
        if (fl[IF_SP_DONE]) break;
        if (rounds >= 4) {
          isk_set_error("poisson: spectral solve reached relative residual %.3g > %.3g", rel,
                        p.rtol);
          return ISKRA_E_CONVERGE;
        }
        hipLaunchKernelGGL(k_sp_correct, dim3(1), dim3(SPC_THREADS), 0, c->stream, sc);
        ISK_CHECK_LAUNCH();
      }
    }
    if (E_done) *E_done = with_E;
    p.last_iters = rounds;
    p.last_relres = rel;
    if (iters) *iters = rounds;
    if (relres) *relres = rel;
    c->stats.poisson_solves++;
    c->stats.poisson_iters += rounds;
    c->stats.last_iters = rounds;
    c->stats.last_relres = rel;
    return ISKRA_OK;
  }
  if (!p.h_sd.empty()) {
    int rc = isk_electrodes_after_push(c, 0.0); /* fold in charge collected since */
    if (rc) return rc;
    return isk_bicgstab_solve(c, warm, iters, relres);
  }
  const int mgpre = p.kind == ISKRA_POISSON_MGPCG;
  PcgArgs a = make_pcg(c, mgpre);
  a.warm = warm;
  hipLaunchKernelGGL(k_pcg_setx, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a);
  hipLaunchKernelGGL(k_pcg_init, dim3(1), dim3(64), 0, c->stream, p.scal, p.iflag,
                     p.rtol * p.rtol, p.singular ? 1 : 0);
  ISK_CHECK_LAUNCH();
  if (p.singular) {
    hipLaunchKernelGGL(k_pcg_mean, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a);
    ISK_CHECK_LAUNCH();
  }
  int hflag[4] = {0, 0, 0, 0};
  if (p.kind == ISKRA_POISSON_MG) {
    /* cycles per host check: one more than the last solve needed, then 2 */
    int batch = p.last_iters > 0 ? p.last_iters + 1 : MG_BATCH;
    int first = 1;
    int launched = batch;
    while (true) {
      for (int b = 0; b < batch; ++b) {
        int rc = mg_vcycle(c, 1, 1, first);
        if (rc) return rc;
        first = 0;
      }
      /* the next cycle's finest down-kernel evaluates the residual of x */
      int rc = mg_vcycle(c, 1, 1, 0);
      if (rc) return rc;
      ISK_HIP(hipMemcpyAsync(hflag, p.iflag, sizeof(hflag), hipMemcpyDeviceToHost, c->stream));
      ISK_HIP(hipStreamSynchronize(c->stream));
      if (hflag[0]) break;
      batch = 1;
      launched += 2;
      if (launched > p.max_iter + 4) {
        isk_set_error("poisson: multigrid stopped after %d host cycles without a device verdict", launched);
        return ISKRA_E_CONVERGE;
      }
    }
  } else {
    hipLaunchKernelGGL(k_pcg_resid, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a);
    ISK_CHECK_LAUNCH();
    if (mgpre) {
      int rc = mg_vcycle(c, 0, 0, 0);
      if (rc) return rc;
      hipLaunchKernelGGL(k_cg_toz, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, 0);
    }
    /* every batch ends with an A kernel, which decides convergence */
    hipLaunchKernelGGL(k_cg_a, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, 0);
    ISK_CHECK_LAUNCH();
    int it = 0;
    while (true) {
      for (int b = 0; b < (mgpre ? MG_BATCH : PCG_BATCH); ++b, ++it) {
        hipLaunchKernelGGL(k_cg_b, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, it);
        if (mgpre) {
          int rc = mg_vcycle(c, 0, 0, 0);
          if (rc) return rc;
          hipLaunchKernelGGL(k_cg_toz, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a,
                             (it + 1) & 1);
        }
        hipLaunchKernelGGL(k_cg_a, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, it + 1);
        ISK_CHECK_LAUNCH();
      }
      ISK_HIP(hipMemcpyAsync(hflag, p.iflag, sizeof(hflag), hipMemcpyDeviceToHost, c->stream));
      ISK_HIP(hipStreamSynchronize(c->stream));
      if (hflag[0]) break;
      if (it > p.max_iter + 2 * PCG_BATCH) {
        isk_set_error("poisson: CG stopped after %d host iterations without a device verdict", it);
        return ISKRA_E_CONVERGE;
      }
    }
  }
  double scal[S_WORDS];
  ISK_HIP(hipMemcpyAsync(scal, p.scal, sizeof(scal), hipMemcpyDeviceToHost, c->stream));
  ISK_HIP(hipStreamSynchronize(c->stream));
  double mean_sub = 0.0;
  if (p.singular) {
    std::vector<double> hx(nn);
    ISK_HIP(hipMemcpy(hx.data(), p.x, nn * 8, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (int k = 0; k < nn; ++k) s += hx[k];
    mean_sub = s / nn;
  }
  hipLaunchKernelGGL(k_pcg_out, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, c->phi, mean_sub);
  ISK_CHECK_LAUNCH();
  p.last_iters = hflag[1];
  p.last_relres = scal[S_RELRES];
  if (iters) *iters = hflag[1];
  if (relres) *relres = scal[S_RELRES];
  c->stats.poisson_solves++;
  c->stats.poisson_iters += hflag[1];
  c->stats.last_iters = hflag[1];
  c->stats.last_relres = scal[S_RELRES];
  if (!(scal[S_RELRES] <= p.rtol)) {
    isk_set_error("poisson: relative residual %.3e after %d iterations (rtol %.1e)",
                  scal[S_RELRES], hflag[1], p.rtol);
    return ISKRA_E_CONVERGE;
  }
  return ISKRA_OK;
}

/* ---- batched launches ---- */
__global__ __launch_bounds__(TRI_BLOCK) void kb_tri(const TriArgs *A, const int *bstart, int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  tri_solve_group(A[i]);
}

__global__ void kb_efield(const EfieldArgs *A, const int *bstart, int n) {
  const int i = isk_bfind(bstart, n, blockIdx.x);
  const EfieldArgs &a = A[i];
  efield_body(a.phi, a.E, a.dim, a.nx, a.ny, a.dx, a.dy, blockIdx.x - bstart[i]);
}

int isk_bbuild_tri(iskra_ctx **cs, int n, IskBLaunch *L) {
  std::vector<TriArgs> A(n);
  std::vector<int> one(n, 1);
  for (int i = 0; i < n; ++i) {
    if (cs[i]->ps.kind != ISKRA_POISSON_DIRECT) {
      isk_set_error("batch: instance %d does not use the direct 1-D Poisson solve", i);
      return ISKRA_E_UNSUPPORTED;
    }
    tri_prepare(cs[i], &A[i]);
  }
  return isk_blaunch_upload(L, ISK_B_TRI, A.data(), sizeof(TriArgs), one, TRI_BLOCK, 0);
}

int isk_bbuild_efield(iskra_ctx **cs, int n, IskBLaunch *L) {
  std::vector<EfieldArgs> A(n);
  std::vector<int> blocks(n);
  for (int i = 0; i < n; ++i) {
    iskra_ctx *c = cs[i];
    A[i].phi = c->phi;
    A[i].E = c->E;
    A[i].dim = c->g.dim;
    A[i].nx = c->g.nx;
    A[i].ny = c->g.ny;
    A[i].dx = c->g.dx;
    A[i].dy = c->g.dy;
    blocks[i] = (isk_nodes(c) + 255) / 256;
  }
  return isk_blaunch_upload(L, ISK_B_EFIELD, A.data(), sizeof(EfieldArgs), blocks, 256, 0);
}

int isk_blaunch_poisson(const IskBLaunch &L, hipStream_t st) {
  if (L.total <= 0) return ISKRA_OK;
  dim3 g((unsigned)L.total), b((unsigned)L.block);
  switch (L.kind) {
    case ISK_B_TRI:
      hipLaunchKernelGGL(kb_tri, g, b, 0, st, (const TriArgs *)L.args, L.bstart, L.n);
      break;
    case ISK_B_EFIELD:
      hipLaunchKernelGGL(kb_efield, g, b, 0, st, (const EfieldArgs *)L.args, L.bstart, L.n);
      break;
    default:
      isk_set_error("batched launch: kind %d is not a Poisson kernel", L.kind);
      return ISKRA_E_ARG;
  }
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

/* in-loop spectral solves that missed rtol since the last check (clears it) */
int isk_poisson_check(iskra_ctx *c, double *relres) {
  IskPoisson &p = c->ps;
  if (p.kind != ISKRA_POISSON_SPECTRAL || !p.iflag) return ISKRA_OK;
  int fl[8];
  double sc[S_WORDS];
  ISK_HIP(hipMemcpy(fl, p.iflag, sizeof(fl), hipMemcpyDeviceToHost));
  ISK_HIP(hipMemcpy(sc, p.scal, sizeof(sc), hipMemcpyDeviceToHost));
  if (relres) *relres = sc[S_RELRES];
  p.last_relres = sc[S_RELRES];
  c->stats.last_relres = sc[S_RELRES];
  if (fl[IF_SP_FAIL]) {
    int zero = 0;
    ISK_HIP(hipMemcpy(p.iflag + IF_SP_FAIL, &zero, 4, hipMemcpyHostToDevice));
    isk_set_error("poisson: a spectral solve in the time loop missed rtol %.3g after "
                  "correction (relative residual %.3g)", p.rtol, sc[S_RELRES]);
    return ISKRA_E_CONVERGE;
  }
  return ISKRA_OK;
}

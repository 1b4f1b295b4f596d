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
 *          products in FP64, search direction p and z in FP32; reductions are
 *          block partial sums combined in a fixed order by the last block.
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

enum { S_RZ = 0, S_PAP, S_ALPHA, S_BETA, S_RR, S_FF, S_TOL2, S_MEAN, S_RELRES, S_RZNEW, S_WORDS = 16 };

struct Sten {
  int dim, nx, ny, per_x, per_y;
  double cxp, cyp;
  const int8_t *tags;
  const double *tagval;
};

template <typename T>
__device__ __forceinline__ double apply_L(const Sten &s, const T *v, int i, int j,
                                          double *diag_out) {
  int n = j * s.nx + i;
  double acc = 0.0, diag = 0.0;
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

/* sum over the Dirichlet neighbours of c * value (moves known values to the rhs) */
__device__ __forceinline__ double dirichlet_rhs(const Sten &s, int i, int j) {
  int n = j * s.nx + i;
  double acc = 0.0;
#define DIRNB(cond, m, coef)                                      \
  if (cond) {                                                    \
    int t = s.tags[m];                                           \
    if (t) acc += (coef) * s.tagval[t];                          \
  }
  DIRNB(i > 0, n - 1, 1.0)
  DIRNB(i < s.nx - 1, n + 1, 1.0)
  if (s.per_x) {
    DIRNB(i == 0, n + s.nx - 1, s.cxp)
    DIRNB(i == s.nx - 1, n - (s.nx - 1), s.cxp)
  }
  if (s.dim == 2) {
    DIRNB(j > 0, n - s.nx, 1.0)
    DIRNB(j < s.ny - 1, n + s.nx, 1.0)
    if (s.per_y) {
      DIRNB(j == 0, n + (s.ny - 1) * s.nx, s.cyp)
      DIRNB(j == s.ny - 1, n - (s.ny - 1) * s.nx, s.cyp)
    }
  }
#undef DIRNB
  return acc;
}

/* block sum of up to 2 doubles; the last block to finish reduces the per-block
 * partials in index order and returns true (deterministic). */
__device__ bool block_reduce2(double a, double b, double *partials,
                              unsigned int *counter, double *out_a, double *out_b) {
  __shared__ double ra[RED_BLOCK], rb[RED_BLOCK];
  __shared__ bool last;
  ra[threadIdx.x] = a;
  rb[threadIdx.x] = b;
  __syncthreads();
  for (int off = RED_BLOCK / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      ra[threadIdx.x] += ra[threadIdx.x + off];
      rb[threadIdx.x] += rb[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = ra[0];
    partials[2 * blockIdx.x + 1] = rb[0];
    __threadfence();
    unsigned int t = atomicAdd(counter, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double sa = 0.0, sb = 0.0;
  for (int k = threadIdx.x; k < (int)gridDim.x; k += RED_BLOCK) {
    sa += __hip_atomic_load(&partials[2 * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sb += __hip_atomic_load(&partials[2 * k + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  ra[threadIdx.x] = sa;
  rb[threadIdx.x] = sb;
  __syncthreads();
  for (int off = RED_BLOCK / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      ra[threadIdx.x] += ra[threadIdx.x + off];
      rb[threadIdx.x] += rb[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *counter = 0u;
  *out_a = ra[0];
  *out_b = rb[0];
  return true;
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
__global__ void k_tri_solve(TriArgs t) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
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

/* ------------------------------------------------------------------------ */
/* PCG                                                                       */
/* ------------------------------------------------------------------------ */
struct PcgArgs {
  Sten s;
  const double *rho;
  double *x, *r, *Ap;
  float *z, *pold, *pnew;
  float *f0;  /* MG preconditioner input (MGPCG) */
  float *e0;  /* MG finest-level correction, zeroed before each V-cycle */
  double *scal, *partials;
  unsigned int *counter;
  int *iflag;
  double scale;
  int warm;
  const double *phi_in;
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
  if (threadIdx.x < S_WORDS) scal[threadIdx.x] = threadIdx.x == S_TOL2 ? tol2 : 0.0;
  if (threadIdx.x < 4) iflag[threadIdx.x] = threadIdx.x == 2 ? singular : 0;
}

/* sum of the eliminated rhs (for the singular projection) */
__global__ void k_pcg_mean(PcgArgs a) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (n < a.s.nx * a.s.ny) v = a.rho[n] * a.scale;
  double s, dummy;
  if (block_reduce2(v, 0.0, a.partials, a.counter, &s, &dummy) && threadIdx.x == 0)
    a.scal[S_MEAN] = s / (double)(a.s.nx * a.s.ny);
}

/* r = f_elim - L x at free nodes; ||f_elim||^2; z = M^-1 r (Jacobi) */
__global__ void k_pcg_resid(PcgArgs a) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int nn = a.s.nx * a.s.ny;
  double rr = 0.0, ff = 0.0;
  if (n < nn) {
    int i = n % a.s.nx, j = n / a.s.nx;
    if (a.s.tags[n] == 0) {
      double f = a.rho[n] * a.scale;
      if (a.iflag[2]) f -= a.scal[S_MEAN];
      double diag;
      /* x holds the Dirichlet values, so L x already carries them */
      double Lx = apply_L(a.s, a.x, i, j, &diag);
      double fe = f + dirichlet_rhs(a.s, i, j);
      double r = f - Lx;
      a.r[n] = r;
      if (!a.mgpre) a.z[n] = (float)(r / diag);
      else a.f0[n] = (float)r;
      if (a.mgpre) a.e0[n] = 0.0f;
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
    a.pold[n] = 0.0f;
    a.pnew[n] = 0.0f;
  }
  double srr, sff;
  if (block_reduce2(rr, ff, a.partials, a.counter, &srr, &sff) && threadIdx.x == 0) {
    a.scal[S_RR] = srr;
    a.scal[S_FF] = sff;
    a.scal[S_TOL2] = a.scal[S_TOL2] * sff; /* rtol^2 preset by the host */
    a.scal[S_BETA] = 0.0;
    a.scal[S_RELRES] = sff > 0.0 ? sqrt(srr / sff) : 0.0;
    if (srr <= a.scal[S_TOL2]) a.iflag[0] = 1;
  }
}

/* rz = (r, z) after the preconditioner */
__global__ void k_pcg_rz(PcgArgs a) {
  if (a.iflag[0]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (n < a.s.nx * a.s.ny) v = a.r[n] * (double)a.z[n];
  double s, dummy;
  if (block_reduce2(v, 0.0, a.partials, a.counter, &s, &dummy) && threadIdx.x == 0) {
    double old = a.scal[S_RZ];
    if (a.iflag[3]) a.scal[S_BETA] = old != 0.0 ? s / old : 0.0;
    else a.scal[S_BETA] = 0.0;
    a.scal[S_RZ] = s;
    a.iflag[3] = 1;
  }
}

/* p_new = z + beta p_old (recomputed at the neighbours), Ap = L p_new, (p, Ap) */
__global__ void k_pcg_a(PcgArgs a) {
  if (a.iflag[0]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int nn = a.s.nx * a.s.ny;
  float beta = (float)a.scal[S_BETA];
  double pap = 0.0;
  if (n < nn) {
    int i = n % a.s.nx, j = n / a.s.nx;
    if (a.s.tags[n] == 0) {
#define PN(m) (a.s.tags[m] ? 0.0 : (double)fmaf(beta, a.pold[m], a.z[m]))
      double pc = PN(n);
      double acc = 0.0, diag = 0.0;
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
#undef PN
      double ap = diag * pc - acc;
      a.pnew[n] = (float)pc;
      a.Ap[n] = ap;
      pap = pc * ap;
    } else {
      a.pnew[n] = 0.0f;
      a.Ap[n] = 0.0;
    }
  }
  double s, dummy;
  if (block_reduce2(pap, 0.0, a.partials, a.counter, &s, &dummy) && threadIdx.x == 0) {
    a.scal[S_PAP] = s;
    a.scal[S_ALPHA] = s != 0.0 ? a.scal[S_RZ] / s : 0.0;
  }
}

/* x += alpha p, r -= alpha Ap, z = M^-1 r, (r,r), (r,z) */
__global__ void k_pcg_b(PcgArgs a) {
  if (a.iflag[0]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int nn = a.s.nx * a.s.ny;
  double alpha = a.scal[S_ALPHA];
  double rr = 0.0, rz = 0.0;
  if (n < nn && a.s.tags[n] == 0) {
    int i = n % a.s.nx, j = n / a.s.nx;
    double p = (double)a.pnew[n];
    a.x[n] = __builtin_fma(alpha, p, a.x[n]);
    double r = __builtin_fma(-alpha, a.Ap[n], a.r[n]);
    a.r[n] = r;
    rr = r * r;
    if (!a.mgpre) {
      double diag = 0.0;
      if (i > 0) diag += 1.0;
      if (i < a.s.nx - 1) diag += 1.0;
      if (a.s.per_x && (i == 0 || i == a.s.nx - 1)) diag += a.s.cxp;
      if (a.s.dim == 2) {
        if (j > 0) diag += 1.0;
        if (j < a.s.ny - 1) diag += 1.0;
        if (a.s.per_y && (j == 0 || j == a.s.ny - 1)) diag += a.s.cyp;
      }
      float z = (float)(r / diag);
      a.z[n] = z;
      rz = r * (double)z;
    } else {
      a.f0[n] = (float)r;
    }
  }
  if (a.mgpre && n < nn) a.e0[n] = 0.0f;
  double srr, srz;
  if (block_reduce2(rr, rz, a.partials, a.counter, &srr, &srz) && threadIdx.x == 0) {
    a.scal[S_RR] = srr;
    a.scal[S_RELRES] = a.scal[S_FF] > 0.0 ? sqrt(srr / a.scal[S_FF]) : 0.0;
    a.iflag[1] += 1;
    if (!a.mgpre) {
      double old = a.scal[S_RZ];
      a.scal[S_BETA] = old != 0.0 ? srz / old : 0.0;
      a.scal[S_RZ] = srz;
    }
    if (srr <= a.scal[S_TOL2] || a.iflag[1] >= a.max_iter) a.iflag[0] = 1;
  }
}

/* pold <- pnew (swap by copy keeps kernel arguments fixed) */
__global__ void k_pcg_swap(PcgArgs a) {
  if (a.iflag[0]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < a.s.nx * a.s.ny) a.pold[n] = a.pnew[n];
}

/* final: mean removal for singular systems, copy to phi */
__global__ void k_pcg_out(PcgArgs a, double *phi, double mean_sub) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < a.s.nx * a.s.ny) phi[n] = a.x[n] - mean_sub;
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
};

__device__ void lds_rbgs(float *e, const float *f, int nx, int ny, int color) {
  int half = (nx * ny + 1) / 2;
  for (int n = threadIdx.x; n < half; n += blockDim.x) {
    int k = 2 * n;
    int j = k / nx, i = k % nx;
    if (((i + j) & 1) != color) {
      ++i;
      if (i >= nx) { i -= nx; ++j; }
    }
    if (((i + j) & 1) != color) continue;
    if (i <= 0 || j <= 0 || i >= nx - 1 || j >= ny - 1) continue;
    int m = j * nx + i;
    float s = (e[m - 1] + e[m + 1]) + (e[m - nx] + e[m + nx]);
    e[m] = (f[m] + s) * 0.25f;
  }
  __syncthreads();
}

__global__ __launch_bounds__(1024) void k_mg_coarse(MgCoarseArgs a) {
  if (a.flag[0]) return;
  extern __shared__ float sm[];
  float *E = sm;
  float *Fv = sm + a.total;
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
    float *ec = E + a.off[l + 1];
    float *fc = Fv + a.off[l + 1];
    int cx = a.nx[l + 1], cy = a.ny[l + 1];
    for (int n = threadIdx.x; n < cx * cy; n += blockDim.x) {
      int I = n % cx, J = n / cx;
      ec[n] = 0.0f;
      if (I <= 0 || J <= 0 || I >= cx - 1 || J >= cy - 1) { fc[n] = 0.0f; continue; }
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
      fc[n] = (4.0f * acc[1][1] + 2.0f * edge + corner) * 0.25f;
    }
    __syncthreads();
  }
  /* coarsest: Gauss-Seidel to convergence (tiny) */
  {
    int l = a.nlev_c - 1;
    float *e = E + a.off[l];
    float *f = Fv + a.off[l];
    for (int it = 0; it < 32; ++it) {
      lds_rbgs(e, f, a.nx[l], a.ny[l], 0);
      lds_rbgs(e, f, a.nx[l], a.ny[l], 1);
    }
  }
  /* up */
  for (int l = a.nlev_c - 2; l >= 0; --l) {
    float *e = E + a.off[l];
    float *f = Fv + a.off[l];
    float *ec = E + a.off[l + 1];
    int nx = a.nx[l], ny = a.ny[l], cx = a.nx[l + 1];
    for (int n = threadIdx.x; n < nx * ny; n += blockDim.x) {
      int i = n % nx, j = n / nx;
      if (i <= 0 || j <= 0 || i >= nx - 1 || j >= ny - 1) continue;
      int I = i >> 1, J = j >> 1;
      int m = J * cx + I;
      float v;
      if (!(i & 1) && !(j & 1)) v = ec[m];
      else if ((i & 1) && !(j & 1)) v = 0.5f * (ec[m] + ec[m + 1]);
      else if (!(i & 1) && (j & 1)) v = 0.5f * (ec[m] + ec[m + cx]);
      else v = 0.25f * ((ec[m] + ec[m + 1]) + (ec[m + cx] + ec[m + cx + 1]));
      e[n] += v;
    }
    __syncthreads();
    lds_rbgs(e, f, nx, ny, 1);
    lds_rbgs(e, f, nx, ny, 0);
    lds_rbgs(e, f, nx, ny, 1);
    lds_rbgs(e, f, nx, ny, 0);
  }
  for (int k = threadIdx.x; k < n0; k += blockDim.x) a.e_out[k] = E[k];
}

/* FP64 defect-correction residual: f0 = r = f_elim - L x, e0 = 0 */
__global__ void k_mg_resid0(PcgArgs a, float *e0, int first) {
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
  double srr, sff;
  if (block_reduce2(rr, ff, a.partials, a.counter, &srr, &sff) && threadIdx.x == 0) {
    if (first) {
      a.scal[S_FF] = sff;
      a.scal[S_TOL2] = a.scal[S_TOL2] * sff;
    }
    a.scal[S_RR] = srr;
    a.scal[S_RELRES] = a.scal[S_FF] > 0.0 ? sqrt(srr / a.scal[S_FF]) : 0.0;
    if (srr <= a.scal[S_TOL2] || a.iflag[1] >= a.max_iter) a.iflag[0] = 1;
    else a.iflag[1] += 1;
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
__global__ void k_efield(const double *phi, float *E, int dim, int nx, int ny,
                         double dx, double dy) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nx * ny) return;
  int i = n % nx, j = n / nx;
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
int isk_poisson_free(iskra_ctx *c) {
  IskPoisson &p = c->ps;
  void *ptrs[] = {p.d_tags, p.d_tag_val, p.d_tri_a, p.d_tri_cp, p.d_tri_inv,
                  p.d_work, p.x, p.r, p.Ap, p.z, p.p[0], p.p[1], p.scal,
                  p.partials, p.counter, p.iflag};
  for (void *q : ptrs) isk_free(q);
  for (int l = 0; l < p.nlev; ++l) {
    isk_free(p.me[l]);
    isk_free(p.mf[l]);
  }
  p.d_tags = nullptr;
  p.d_tag_val = nullptr;
  p.d_tri_a = p.d_tri_cp = p.d_tri_inv = p.d_work = nullptr;
  p.x = p.r = p.Ap = nullptr;
  p.z = p.p[0] = p.p[1] = nullptr;
  p.scal = p.partials = nullptr;
  p.counter = nullptr;
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
  if (isk_alloc((void **)&p.x, nn * 8) || isk_alloc((void **)&p.r, nn * 8) ||
      isk_alloc((void **)&p.Ap, nn * 8) || isk_alloc((void **)&p.z, nn * 4) ||
      isk_alloc((void **)&p.p[0], nn * 4) || isk_alloc((void **)&p.p[1], nn * 4) ||
      isk_alloc((void **)&p.scal, S_WORDS * 8) ||
      isk_alloc((void **)&p.partials, (size_t)nblk * 2 * 8) ||
      isk_alloc((void **)&p.counter, 64) || isk_alloc((void **)&p.iflag, 64))
    return ISKRA_E_HIP;
  ISK_HIP(hipMemset(p.counter, 0, 64));
  ISK_HIP(hipMemset(p.iflag, 0, 64));
  p.nlev = 0;
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
  a.pold = p.p[0];
  a.pnew = p.p[1];
  a.f0 = p.nlev ? p.mf[0] : nullptr;
  a.e0 = p.nlev ? p.me[0] : nullptr;
  a.scal = p.scal;
  a.partials = p.partials;
  a.counter = p.counter;
  a.iflag = p.iflag;
  a.scale = (c->g.dx * c->g.dx) / p.eps0;
  a.warm = 0;
  a.phi_in = c->phi;
  a.mgpre = mgpre;
  a.max_iter = p.max_iter;
  return a;
}

static int mg_vcycle(iskra_ctx *c) {
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
  for (int l = 0; l < p.coarse_lev; ++l) {
    MgLevel L = lev(l);
    int half = (L.nx * L.ny + 1) / 2;
    for (int s = 0; s < 2; ++s) {
      hipLaunchKernelGGL(k_mg_rbgs, dim3((half + 255) / 256), dim3(256), 0, c->stream, L, 0, flag);
      hipLaunchKernelGGL(k_mg_rbgs, dim3((half + 255) / 256), dim3(256), 0, c->stream, L, 1, flag);
    }
    MgLevel C = lev(l + 1);
    hipLaunchKernelGGL(k_mg_restrict, dim3((C.nx * C.ny + 255) / 256), dim3(256), 0,
                       c->stream, L, C, flag);
  }
  {
    MgCoarseArgs a;
    int tot = 0;
    a.nlev_c = p.nlev - p.coarse_lev;
    for (int k = 0; k < a.nlev_c; ++k) {
      int l = p.coarse_lev + k;
      a.nx[k] = p.ln[l] & 0xFFFF;
      a.ny[k] = p.ln[l] >> 16;
      a.off[k] = tot;
      tot += a.nx[k] * a.ny[k];
    }
    a.total = tot;
    a.e_out = p.me[p.coarse_lev];
    a.f_in = p.mf[p.coarse_lev];
    a.flag = flag;
    hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(1024), (size_t)tot * 2 * 4, c->stream, a);
  }
  for (int l = p.coarse_lev - 1; l >= 0; --l) {
    MgLevel L = lev(l), C = lev(l + 1);
    hipLaunchKernelGGL(k_mg_prolong, dim3((L.nx * L.ny + 255) / 256), dim3(256), 0,
                       c->stream, C, L, flag);
    int half = (L.nx * L.ny + 1) / 2;
    for (int s = 0; s < 2; ++s) {
      hipLaunchKernelGGL(k_mg_rbgs, dim3((half + 255) / 256), dim3(256), 0, c->stream, L, 1, flag);
      hipLaunchKernelGGL(k_mg_rbgs, dim3((half + 255) / 256), dim3(256), 0, c->stream, L, 0, flag);
    }
  }
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

int isk_poisson_solve(iskra_ctx *c, int warm, int *iters, double *relres) {
  IskPoisson &p = c->ps;
  const int nn = isk_nodes(c);
  if (p.kind == ISKRA_POISSON_DIRECT) {
    TriArgs t;
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
    hipLaunchKernelGGL(k_tri_solve, dim3(1), dim3(64), 0, c->stream, t);
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
    hipLaunchKernelGGL(k_mg_resid0, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, p.me[0], 1);
    ISK_CHECK_LAUNCH();
    while (true) {
      for (int b = 0; b < MG_BATCH; ++b) {
        int rc = mg_vcycle(c);
        if (rc) return rc;
        hipLaunchKernelGGL(k_mg_update0, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, p.me[0]);
        hipLaunchKernelGGL(k_mg_resid0, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a, p.me[0], 0);
        ISK_CHECK_LAUNCH();
      }
      ISK_HIP(hipMemcpyAsync(hflag, p.iflag, sizeof(hflag), hipMemcpyDeviceToHost, c->stream));
      ISK_HIP(hipStreamSynchronize(c->stream));
      if (hflag[0]) break;
    }
  } else {
    hipLaunchKernelGGL(k_pcg_resid, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a);
    ISK_CHECK_LAUNCH();
    if (mgpre) {
      int rc = mg_vcycle(c);
      if (rc) return rc;
      hipLaunchKernelGGL(k_mg_toz, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, p.me[0], p.z, nn, p.iflag);
    }
    hipLaunchKernelGGL(k_pcg_rz, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a);
    ISK_CHECK_LAUNCH();
    while (true) {
      for (int b = 0; b < (mgpre ? MG_BATCH : PCG_BATCH); ++b) {
        hipLaunchKernelGGL(k_pcg_a, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a);
        hipLaunchKernelGGL(k_pcg_b, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a);
        if (mgpre) {
          int rc = mg_vcycle(c);
          if (rc) return rc;
          hipLaunchKernelGGL(k_mg_toz, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, p.me[0], p.z, nn, p.iflag);
          hipLaunchKernelGGL(k_pcg_rz, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a);
        }
        hipLaunchKernelGGL(k_pcg_swap, dim3(nblk), dim3(RED_BLOCK), 0, c->stream, a);
        ISK_CHECK_LAUNCH();
      }
      ISK_HIP(hipMemcpyAsync(hflag, p.iflag, sizeof(hflag), hipMemcpyDeviceToHost, c->stream));
      ISK_HIP(hipStreamSynchronize(c->stream));
      if (hflag[0]) break;
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

/*
 * fluid.hip -- mobile fluid species on gfx950.
 *
 * Reference: advance!(::FluidSpecies) (ParticleInCell/src/ParticleInCell.jl:
 * 74-82) with calculate_advection_diffusion (FiniteDifferenceMethod/src/
 * advection_diffusion.jl:4-38): with the E the step starts from,
 *     v  = (E dt) (q/m)
 *     1-D, interior nodes: dn = dt mu (n[i-1] - 2 n[i] + n[i+1]) / dx^2
 *         (the advection terms are commented out in the reference)
 *     2-D, interior nodes:
 *         dn  = dt mu (n[i-1,j] + n[i,j-1] - 4 n[i,j] + n[i+1,j] + n[i,j+1]) / (dx dy)
 *         dn -= dt n[i,j] (vx[i+1,j] - vx[i-1,j] + vy[i,j+1] - vy[i,j-1]) / (2 dx)
 *         dn -= dt (vx[i,j] (n[i+1,j] - n[i-1,j]) + vy[i,j] (n[i,j+1] - n[i,j-1])) / dx
 *     boundary nodes: dn = 0;  then n += dn
 * The FP64 operations are those of the reference's expressions in their
 * order (the library is built with FP contraction off), so a host
 * restatement that does the same operations on the same E reproduces the
 * result bit for bit.  E is the library's FP32 node field.
 *
 * The density of a fluid enters rho as q n (no volume division,
 * pic/fluid.jl:11); the fluids' charge is kept in the background field that
 * the deposit merge adds to rho: bg = bg_fixed + q_0 n_0 + q_1 n_1 + ...
 * Layout: one FP64 value per node, x fastest; two launches per step (the
 * stencil reads the old densities of all neighbours, so the update waits
 * for every dn).
 */
#include <math.h>
#include <string.h>

#include "isk_internal.h"

struct FluidArgs {
  double *n[ISK_MAX_FLUIDS];
  double *dn[ISK_MAX_FLUIDS];
  double q[ISK_MAX_FLUIDS], qm[ISK_MAX_FLUIDS], mu[ISK_MAX_FLUIDS];
  int nf;
  const float *E;
  const double *bg_fixed;
  double *bg;
  int nx, ny, dim;
  double dt, dx, dy;
};

#define FLUID_BLOCK 256

/* dn of fluid blockIdx.y at every node */
__global__ __launch_bounds__(FLUID_BLOCK) void k_fluid_dn(FluidArgs a) {
  const int node = blockIdx.x * FLUID_BLOCK + threadIdx.x;
  const int f = blockIdx.y;
  if (node >= a.nx * a.ny || f >= a.nf) return;
  const int i = node % a.nx, j = node / a.nx;
  const double *n = a.n[f];
  double d = 0.0;
  if (a.dim == 1) {
    if (i >= 1 && i <= a.nx - 2)
      d = 0.0 + ((a.dt * a.mu[f]) * ((n[i - 1] - 2.0 * n[i]) + n[i + 1])) / (a.dx * a.dx);
  } else if (i >= 1 && i <= a.nx - 2 && j >= 1 && j <= a.ny - 2) {
    const int L = node - 1, R = node + 1, D = node - a.nx, U = node + a.nx;
    const double qm = a.qm[f];
    /* v = (E dt) (q/m) at the nodes the stencil reads */
    const double vxL = ((double)a.E[2 * L] * a.dt) * qm;
    const double vxR = ((double)a.E[2 * R] * a.dt) * qm;
    const double vyD = ((double)a.E[2 * D + 1] * a.dt) * qm;
    const double vyU = ((double)a.E[2 * U + 1] * a.dt) * qm;
    const double vx = ((double)a.E[2 * node] * a.dt) * qm;
    const double vy = ((double)a.E[2 * node + 1] * a.dt) * qm;
    const double lap = (((n[L] + n[D]) - 4.0 * n[node]) + n[R]) + n[U];
    d = 0.0 + ((a.dt * a.mu[f]) * lap) / (a.dx * a.dy);
    d = d - ((a.dt * n[node]) * (((vxR - vxL) + vyU) - vyD)) / (2.0 * a.dx);
    d = d - (a.dt * ((vx * (n[R] - n[L])) + (vy * (n[U] - n[D])))) / a.dx;
  }
  a.dn[f][node] = d;
}

/* n += dn (when apply) and bg = bg_fixed + sum q n, per node */
__global__ __launch_bounds__(FLUID_BLOCK) void k_fluid_apply(FluidArgs a, int apply) {
  const int node = blockIdx.x * FLUID_BLOCK + threadIdx.x;
  if (node >= a.nx * a.ny) return;
  double b = a.bg_fixed[node];
  for (int f = 0; f < a.nf; ++f) {
    double v = a.n[f][node];
    if (apply) {
      v = v + a.dn[f][node];
      a.n[f][node] = v;
    }
    b = b + a.q[f] * v;
  }
  a.bg[node] = b;
}

static FluidArgs fluid_args(iskra_ctx *c, double dt) {
  FluidArgs a;
  memset(&a, 0, sizeof a);
  for (int f = 0; f < c->nfluid; ++f) {
    a.n[f] = c->fl[f].n;
    a.dn[f] = c->fl[f].dn;
    a.q[f] = c->fl[f].q;
    a.qm[f] = c->fl[f].qm;
    a.mu[f] = c->fl[f].mu;
  }
  a.nf = c->nfluid;
  a.E = c->E;
  a.bg_fixed = c->bg_fixed;
  a.bg = c->bg;
  a.nx = c->g.nx;
  a.ny = c->g.ny;
  a.dim = c->g.dim;
  a.dt = dt;
  a.dx = c->g.dx;
  a.dy = c->g.dy;
  return a;
}

int isk_fluid_step(iskra_ctx *c, double dt) {
  if (!c->nfluid) return ISKRA_OK;
  FluidArgs a = fluid_args(c, dt);
  const int nn = c->g.nx * c->g.ny;
  const int blocks = (nn + FLUID_BLOCK - 1) / FLUID_BLOCK;
  hipLaunchKernelGGL(k_fluid_dn, dim3(blocks, c->nfluid), dim3(FLUID_BLOCK), 0, c->stream, a);
  ISK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_fluid_apply, dim3(blocks), dim3(FLUID_BLOCK), 0, c->stream, a, 1);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

int isk_fluid_bg(iskra_ctx *c) {
  if (!c->nfluid) return ISKRA_OK;
  FluidArgs a = fluid_args(c, 0.0);
  const int nn = c->g.nx * c->g.ny;
  hipLaunchKernelGGL(k_fluid_apply, dim3((nn + FLUID_BLOCK - 1) / FLUID_BLOCK),
                     dim3(FLUID_BLOCK), 0, c->stream, a, 0);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

void isk_fluid_release(iskra_ctx *c) {
  for (int f = 0; f < c->nfluid; ++f) {
    isk_free(c->fl[f].n);
    isk_free(c->fl[f].dn);
    c->fl[f] = IskFluid();
  }
  c->nfluid = 0;
  isk_free(c->bg_fixed);
  c->bg_fixed = nullptr;
}

/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY.  CPU oracle of the electrostatic PIC
 * step used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker; nothing in the product (iskra_amd/) links, loads or calls it.
 *
 * Two families of functions:
 *
 *  ref_*   literal FP64 restatement of the reference Julia code (absolute
 *          positions, column-major arrays, 1-based logic translated to
 *          0-based indices).  Each function cites the reference lines it
 *          follows.  This is the "reference" every GPU result is compared to
 *          within a stated tolerance.
 *
 *  mir_*   the same algorithms in the GPU's number format (int cell index +
 *          FP32 offsets, FP32 velocities, 64-bit fixed-point CIC weights,
 *          FP64 Poisson / E-field, Philox4x32-10 draws), written
 *          independently of the HIP sources but with the same operation
 *          order, so that GPU results can be compared bit for bit.
 *
 * The FP64 transcendental functions used by the MCC mirror come from
 * iskra_amd/csrc/isk_fp64math.h (shared on purpose: bit-identical exp/log/
 * sin/cos on CPU and GPU); the test suite pins that header against libm.
 *
 * Parity status: see the header of oracle/__init__.py for which functions are
 * pinned by which check.
 *
 * Build: oracle/build.py (gcc -O2 -ffp-contract=off).
 */
#include <math.h>
#ifndef M_PI
#define M_PI 3.14159265358979323846 /* pi rounded to double, as Julia converts it */
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../iskra_amd/csrc/isk_fp64math.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXPORT __attribute__((visibility("default")))
#define FIX_ONE 4294967296.0

/* ======================================================================== */
/* Philox (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as  */
/* 1, 2, 3", SC'11)                                                         */
/* ======================================================================== */
EXPORT void orc_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Philox4x64-10, same construction with 64-bit words (cross-check against
 * numpy.random.Philox, which implements this variant) */
EXPORT void orc_philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
  uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint64_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    __uint128_t p0 = (__uint128_t)0xD2E7470EE14C6C93ULL * c0;
    __uint128_t p1 = (__uint128_t)0xCA5A826395121157ULL * c2;
    uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
    uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
    uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B97F4A7C15ULL;
    k1 += 0xBB67AE8584CAA73BULL;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* uniform in (0,1) from two words: 52 bits, centred in its bin */
EXPORT double orc_u01(uint32_t a, uint32_t b) {
  uint64_t k = ((uint64_t)(a >> 6) << 26) | (uint64_t)(b >> 6);
  return ((double)k + 0.5) * (1.0 / 4503599627370496.0);
}

/* thin wrappers so tests can pin the shared FP64 math against libm */
EXPORT /* fixed-point CIC weights of the particle kernels (shared header) */
long long orc_fix32(float h) { return isk_fix32(h); }
long long orc_fixprod32(float a, float b) { return isk_fixprod32(a, b); }

double orc_exp(double x) { return isk_exp(x); }
EXPORT double orc_log(double x) { return isk_log(x); }
EXPORT void orc_sincos2pi(double u, double *s, double *c) { isk_sincos2pi(u, s, c); }

/* ======================================================================== */
/* reference mode (FP64)                                                    */
/* ======================================================================== */

/* cell_volume (RegularGrids/src/RegularGrids.jl:26-38); the 1-D version is
 * inferred by analogy (no 1-D method exists in the reference). */
EXPORT void ref_cell_volume(int dim, int nx, int ny, double dx, double dy,
                            const int32_t volper[4], double *V) {
  double V0 = dim == 2 ? dx * dy : dx;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) V[j * nx + i] = V0;
  for (int j = 0; j < ny; ++j) {
    if (!volper[0]) V[j * nx + 0] *= 0.5;
    if (!volper[1]) V[j * nx + nx - 1] *= 0.5;
  }
  if (dim == 2) {
    for (int i = 0; i < nx; ++i) {
      if (!volper[2]) V[i] *= 0.5;
      if (!volper[3]) V[(ny - 1) * nx + i] *= 0.5;
    }
  }
}

/* cell_volume(::AxialGrid{2}) (RegularGrids/src/RegularGrids.jl:40-53), in
 * the reference's 1-based node numbers and operation order: annuli
 * pi dz ((i dr - dr/2)^2 - (i dr - 3dr/2)^2) for i = 2..nr-1, the axis disc
 * pi dz (dr/2)^2, the outer node pi dz ((nr dr - dr)^2 - (nr dr - 3dr/2)^2);
 * rows on a non-periodic z edge halved (volper_z = {bottom, top}). */
EXPORT void ref_cell_volume_axial(int nr, int nz, double dr, double dz, const int32_t volper_z[2],
                                  double *V) {
  for (int j = 1; j <= nz; ++j) {
    for (int i = 2; i <= nr - 1; ++i) {
      double a = i * dr - 0.5 * dr, b = i * dr - 1.5 * dr;
      V[(j - 1) * nr + (i - 1)] = M_PI * dz * (a * a - b * b);
    }
    double h = 0.5 * dr;
    V[(j - 1) * nr] = M_PI * dz * (h * h);
    double a = nr * dr - 1.0 * dr, b = nr * dr - 1.5 * dr;
    V[(j - 1) * nr + nr - 1] = M_PI * dz * (a * a - b * b);
  }
  for (int i = 0; i < nr; ++i) {
    if (!volper_z[0]) V[i] *= 0.5;
    if (!volper_z[1]) V[(nz - 1) * nr + i] *= 0.5;
  }
}

/* particle_cell (ParticleInCell/src/ParticleInCell.jl:28-35): f = 1 + x/dh,
 * i = floor(f) (1-based), h = f - i.  Returns the 0-based i. */
static inline void ref_particle_cell(double x, double dh, int *i, double *h) {
  double f = 1.0 + x / dh;
  double fl = floor(f);
  *i = (int)fl - 1;
  *h = f - fl;
}

/* particle_to_grid (ParticleInCell/src/pic/cloud_in_cell.jl:1-18) with
 * pu(p) = wg[p]; u is nx x ny column-major (i fastest) and is overwritten.
 * 1-D: weights (1-h), h by analogy (no 1-D method in the reference). */
EXPORT void ref_particle_to_grid(int dim, int64_t np, const double *x, const double *wg,
                                 int nx, int ny, double dx, double dy, double *u) {
  memset(u, 0, sizeof(double) * (size_t)nx * ny);
  for (int64_t p = 0; p < np; ++p) {
    int i, j = 0;
    double hx, hy = 0.0;
    ref_particle_cell(x[p], dx, &i, &hx);
    if (dim == 2) ref_particle_cell(x[np + p], dy, &j, &hy);
    double w = wg[p];
    if (dim == 2) {
      u[j * nx + i] += (1.0 - hx) * (1.0 - hy) * w;
      u[j * nx + i + 1] += (hx) * (1.0 - hy) * w;
      u[(j + 1) * nx + i] += (1.0 - hx) * (hy)*w;
      u[(j + 1) * nx + i + 1] += (hx) * (hy)*w;
    } else {
      u[i] += (1.0 - hx) * w;
      u[i + 1] += hx * w;
    }
  }
}

/* grid_to_particle (cloud_in_cell.jl:20-36): E is nx x ny x 3 column-major
 * (component slowest); pE is np x 3 column-major. */
EXPORT void ref_grid_to_particle(int dim, int64_t np, const double *x, int nx, int ny,
                                 double dx, double dy, const double *E, double *pE) {
  size_t nn = (size_t)nx * ny;
  for (int64_t p = 0; p < np; ++p) {
    int i, j = 0;
    double hx, hy = 0.0;
    ref_particle_cell(x[p], dx, &i, &hx);
    if (dim == 2) ref_particle_cell(x[np + p], dy, &j, &hy);
    for (int c = 0; c < 3; ++c) {
      const double *u = E + c * nn;
      double v;
      if (dim == 2)
        v = (1.0 - hx) * (1.0 - hy) * u[j * nx + i] + (hx) * (1.0 - hy) * u[j * nx + i + 1] +
            (1.0 - hx) * (hy)*u[(j + 1) * nx + i] + (hx) * (hy)*u[(j + 1) * nx + i + 1];
      else
        v = (1.0 - hx) * u[i] + hx * u[i + 1];
      pE[c * np + p] = v;
    }
  }
}

/* push_in_cartesian! (ParticleInCell/src/pic/pushers.jl:37-50), Boris as the
 * reference writes it (v- x B, not v- x t); x: np x D, v, E, B: np x 3. */
EXPORT void ref_push(int64_t np, int D, double *x, double *v, const double *E, const double *B,
                     double qm, double dt) {
  for (int64_t p = 0; p < np; ++p) {
    double vm[3], t[3], vp[3], s[3], vpl[3];
    for (int c = 0; c < 3; ++c) {
      vm[c] = 0.5 * dt * qm * E[c * np + p] + v[c * np + p];
      t[c] = 0.5 * dt * qm * B[c * np + p];
    }
    double t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    /* v' = v- + v- x B */
    vp[0] = vm[0] + (vm[1] * B[2 * np + p] - vm[2] * B[1 * np + p]);
    vp[1] = vm[1] + (vm[2] * B[0 * np + p] - vm[0] * B[2 * np + p]);
    vp[2] = vm[2] + (vm[0] * B[1 * np + p] - vm[1] * B[0 * np + p]);
    for (int c = 0; c < 3; ++c) s[c] = 2.0 / (1.0 + t2) * t[c];
    vpl[0] = vm[0] + (vp[1] * s[2] - vp[2] * s[1]);
    vpl[1] = vm[1] + (vp[2] * s[0] - vp[0] * s[2]);
    vpl[2] = vm[2] + (vp[0] * s[1] - vp[1] * s[0]);
    for (int c = 0; c < 3; ++c) v[c * np + p] = dt * E[c * np + p] * qm * 0.5 + vpl[c];
    for (int c = 0; c < D; ++c) x[c * np + p] = dt * v[c * np + p] + x[c * np + p];
  }
}

/* transform_from_cartesian_to_cylindrical! (ParticleInCell/src/pic/
 * pushers.jl:52-66) after push_in_cartesian! (the RZ pusher, :13-17):
 * y = dt v_3; r = sqrt(x_1^2 + y^2); sin = y/r (0 where r == 0, the
 * reference's r .~ 0.0), cos = sqrt(1 - sin^2); x_1 = r and
 * (v_1, v_3) = (cos v_1 + sin v_3, -sin v_1 + cos v_3).  x: np x 2, v: np x 3. */
EXPORT void ref_rz_transform(int64_t np, double *x, double *v, double dt) {
  for (int64_t p = 0; p < np; ++p) {
    double y = dt * v[2 * np + p];
    double r = sqrt(x[p] * x[p] + y * y);
    double sn = y / r;
    if (r == 0.0) sn = 0.0;
    double cs = sqrt(1.0 - sn * sn);
    double vr = cs * v[p] + sn * v[2 * np + p];
    double vy = -sn * v[p] + cs * v[2 * np + p];
    x[p] = r;
    v[p] = vr;
    v[2 * np + p] = vy;
  }
}

/* Julia fld(a, b) for b > 0 */
static inline double jl_fld(double a, double b) { return floor(a / b); }

/* The per-step particle work of the reference on a 2-D grid -- grid_to_particle
 * of E (cloud_in_cell.jl:20-36), push_in_cartesian! with B = 0
 * (pushers.jl:37-50), discard! on both axes (wrap.jl:1-18) and
 * particle_to_grid of wg (cloud_in_cell.jl:1-18) -- with the same FP64
 * arithmetic per particle as ref_grid_to_particle / ref_push / ref_discard /
 * ref_particle_to_grid, spread over nthreads OpenMP threads: particles in
 * parallel, the deposit into one grid per thread summed afterwards, discarded
 * particles replaced by survivors from the end of the arrays (the set the
 * reference's swap-remove leaves; the order may differ).  CPU baseline only.
 * x: 2 x np
 * and v: 3 x np column-major with leading dimension np; E: nx x ny x 3; u
 * (nx x ny) receives the deposit.  Returns the particles left. */
EXPORT int64_t ref_particles_omp(int64_t np, double *x, double *v, double *wg, int nx, int ny,
                                 double dx, double dy, const double *E, double qm, double dt,
                                 double *u, int nthreads) {
  const size_t nn = (size_t)nx * ny;
  const double Lx = (double)(nx - 1) * dx, Ly = (double)(ny - 1) * dy;
  unsigned char *keep = (unsigned char *)malloc((size_t)(np > 0 ? np : 1));
  double *part = (double *)calloc(nn * (size_t)nthreads, sizeof(double));
  if (!keep || !part) {
    free(keep);
    free(part);
    return -1;
  }
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t p = 0; p < np; ++p) {
    int i, j;
    double hx, hy;
    ref_particle_cell(x[p], dx, &i, &hx);
    ref_particle_cell(x[np + p], dy, &j, &hy);
    double pe[3];
    for (int c = 0; c < 3; ++c) {
      const double *g = E + c * nn;
      pe[c] = (1.0 - hx) * (1.0 - hy) * g[j * nx + i] + (hx) * (1.0 - hy) * g[j * nx + i + 1] +
              (1.0 - hx) * (hy)*g[(j + 1) * nx + i] + (hx) * (hy)*g[(j + 1) * nx + i + 1];
    }
    for (int c = 0; c < 3; ++c) {
      double vm = 0.5 * dt * qm * pe[c] + v[c * np + p];
      v[c * np + p] = dt * pe[c] * qm * 0.5 + vm;
    }
    for (int c = 0; c < 2; ++c) x[c * np + p] = dt * v[c * np + p] + x[c * np + p];
    keep[p] = jl_fld(x[p], Lx) == 0.0 && jl_fld(x[np + p], Ly) == 0.0;
  }
  /* removal as remove! does it (kinetic.jl:20-27): a discarded slot below
   * the new count takes a surviving particle from above it */
  int64_t m = 0;
#pragma omp parallel for num_threads(nthreads) reduction(+ : m) schedule(static)
  for (int64_t p = 0; p < np; ++p) m += keep[p];
  for (int64_t h = 0, t = m; h < m; ++h) {
    if (keep[h]) continue;
    while (!keep[t]) ++t;
    for (int c = 0; c < 2; ++c) x[c * np + h] = x[c * np + t];
    for (int c = 0; c < 3; ++c) v[c * np + h] = v[c * np + t];
    wg[h] = wg[t];
    ++t;
  }
#pragma omp parallel num_threads(nthreads)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    double *ut = part + (size_t)t * nn;
#pragma omp for schedule(static)
    for (int64_t p = 0; p < m; ++p) {
      int i, j;
      double hx, hy;
      ref_particle_cell(x[p], dx, &i, &hx);
      ref_particle_cell(x[np + p], dy, &j, &hy);
      const double w = wg[p];
      ut[j * nx + i] += (1.0 - hx) * (1.0 - hy) * w;
      ut[j * nx + i + 1] += (hx) * (1.0 - hy) * w;
      ut[(j + 1) * nx + i] += (1.0 - hx) * (hy)*w;
      ut[(j + 1) * nx + i + 1] += (hx) * (hy)*w;
    }
  }
  for (size_t k = 0; k < nn; ++k) {
    double s = 0.0;
    for (int t = 0; t < nthreads; ++t) s += part[(size_t)t * nn + k];
    u[k] = s;
  }
  free(keep);
  free(part);
  return m;
}

/* wrap! (ParticleInCell/src/pic/surfaces/wrap.jl:20-33) over the listed dims */
EXPORT void ref_wrap(int64_t np, int D, double *x, const int64_t *n, const double *dh,
                     const double *origin, int ndims, const int32_t *dims) {
  for (int k = 0; k < ndims; ++k) {
    int d = dims[k];
    double Lx = (double)(n[d] - 1) * dh[d];
    for (int64_t p = 0; p < np; ++p) {
      double a = jl_fld(x[d * np + p] - origin[d], Lx);
      if (a != 0.0) x[d * np + p] -= a * Lx;
    }
  }
}

/* remove! (ParticleInCell/src/pic/kinetic.jl:20-27): slot i takes the last
 * particle; ids are swapped; wg[np-1] = w0. Arrays have leading dimension cap. */
static void ref_remove(int64_t *np, int64_t cap, int D, double *x, double *v, double *wg,
                       uint32_t *id, double w0, int64_t i) {
  int64_t last = *np - 1;
  for (int c = 0; c < D; ++c) x[c * cap + i] = x[c * cap + last];
  for (int c = 0; c < 3; ++c) v[c * cap + i] = v[c * cap + last];
  wg[i] = wg[last];
  wg[last] = w0;
  uint32_t t = id[i];
  id[i] = id[last];
  id[last] = t;
  *np = last;
}

/* discard! (wrap.jl:1-18): reverse scan per dim, swap-remove when
 * fld(x - origin, L) != 0.  Returns the number removed. */
EXPORT int64_t ref_discard(int64_t *np, int64_t cap, int D, double *x, double *v, double *wg,
                           uint32_t *id, double w0, const int64_t *n, const double *dh,
                           const double *origin, int ndims, const int32_t *dims) {
  int64_t before = *np;
  for (int k = 0; k < ndims; ++k) {
    int d = dims[k];
    double Lx = (double)(n[d] - 1) * dh[d];
    for (int64_t p = *np - 1; p >= 0; --p) {
      double a = jl_fld(x[d * cap + p] - origin[d], Lx);
      if (a != 0.0) ref_remove(np, cap, D, x, v, wg, id, w0, p);
    }
  }
  return before - *np;
}

/* create_generalized_poisson_solver (FiniteDifferenceMethod/src/
 * generalized_poisson.jl:34-68): dense nn x nn A, column-major, node
 * k = i + nx*j.  For dim 1, ny = 1. */
EXPORT void ref_poisson_assemble(int nx, int ny, double dx, double *A) {
  size_t nn = (size_t)nx * ny;
  memset(A, 0, sizeof(double) * nn * nn);
#define AE(r, c) A[(size_t)(c)*nn + (r)]
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      int k = i + nx * j;
      if (i < nx - 1) { AE(k, k) -= 1.0; AE(k, k + 1) += 1.0; }
      if (i > 0) { AE(k, k) -= 1.0; AE(k, k - 1) += 1.0; }
      if (j < ny - 1) { AE(k, k) -= 1.0; AE(k, k + nx) += 1.0; }
      if (j > 0) { AE(k, k) -= 1.0; AE(k, k - nx) += 1.0; }
    }
  double d2 = dx * dx;
  for (size_t q = 0; q < nn * nn; ++q) A[q] /= d2;
}

/* create_poisson_solver(::AxialGrid{2}, eps0) (generalized_poisson.jl:70-199):
 * dense nn x nn A (column-major, node k = i + nr j), entries 1/dr^2, 1/dz^2
 * and the first-derivative terms +-0.5/(dr r) with r = i dr (0-based i) on
 * the rows off the axis and off the outer wall.  The reference writes nine
 * loops (interior, bottom, top, axis, side, four corners); they are one rule:
 * a missing neighbour's link is mirrored back onto the diagonal
 * (+1/dh^2 on top of the -2/dh^2), and the derivative terms exist exactly on
 * rows 0 < i < nr-1. */
EXPORT void ref_poisson_assemble_rz(int nr, int nz, double dr, double dz, double *A) {
  size_t nn = (size_t)nr * nz;
  memset(A, 0, sizeof(double) * nn * nn);
  double cr = 1.0 / (dr * dr), cz = 1.0 / (dz * dz);
  for (int j = 0; j < nz; ++j)
    for (int i = 0; i < nr; ++i) {
      int k = i + nr * j;
      /* radial second difference */
      AE(k, (i < nr - 1) ? k + 1 : k) += cr;
      AE(k, k) -= 2.0 * cr;
      AE(k, (i > 0) ? k - 1 : k) += cr;
      /* axial second difference */
      AE(k, (j < nz - 1) ? k + nr : k) += cz;
      AE(k, k) -= 2.0 * cz;
      AE(k, (j > 0) ? k - nr : k) += cz;
      if (i > 0 && i < nr - 1) {
        double r = i * dr;
        AE(k, k + 1) += 0.5 / dr / r;
        AE(k, k - 1) -= 0.5 / dr / r;
      }
    }
}

/* apply_periodic 1-D (generalized_poisson.jl:271-284) */
EXPORT void ref_apply_periodic_1d(int nx, double dx, double *A) {
  size_t nn = (size_t)nx;
  double c = (0.5 * 1.0 + 0.5 * 1.0) / (dx * dx);
  AE(0, 0) -= c;
  AE(0, nx - 1) += c;
  AE(nx - 1, nx - 1) -= c;
  AE(nx - 1, 0) += c;
}

/* apply_periodic 2-D (generalized_poisson.jl:286-324); axis 1 links j=1<->ny
 * with dx^2, axis 2 links i=1<->nx with dy^2 (as written in the reference) */
EXPORT void ref_apply_periodic_2d(int nx, int ny, double dx, double dy, int axis, double *A) {
  size_t nn = (size_t)nx * ny;
  if (axis == 1) {
    double c = (0.5 + 0.5) / (dx * dx);
    for (int jj = 0; jj < 2; ++jj) {
      int j = jj == 0 ? 0 : ny - 1;
      for (int i = 0; i < nx; ++i) {
        int k = i + nx * j;
        if (j == ny - 1) { AE(k, k) -= c; AE(k, i) += c; }
        if (j == 0) { AE(k, k) -= c; AE(k, i + nx * (ny - 1)) += c; }
      }
      if (ny == 1) break;
    }
  } else {
    double c = (0.5 + 0.5) / (dy * dy);
    for (int j = 0; j < ny; ++j)
      for (int ii = 0; ii < 2; ++ii) {
        int i = ii == 0 ? 0 : nx - 1;
        int k = i + nx * j;
        if (i == nx - 1) { AE(k, k) -= c; AE(k, nx * j) += c; }
        if (i == 0) { AE(k, k) -= c; AE(k, nx - 1 + nx * j) += c; }
        if (nx == 1) break;
      }
  }
}

/* apply_dirichlet (generalized_poisson.jl:205-215): row cleared, diagonal 1,
 * b = value, node removed from the rho dofs (isdof[k] = 0) */
EXPORT void ref_apply_dirichlet(int nn, const uint8_t *mask, double value, double *A, double *b,
                                uint8_t *isdof) {
  for (int k = 0; k < nn; ++k) {
    if (!mask[k]) continue;
    for (int c = 0; c < nn; ++c) AE(k, c) = 0.0;
    AE(k, k) = 1.0;
    b[k] = value;
    isdof[k] = 0;
  }
}

/* LU with partial pivoting (the algorithm of LAPACK dgetrf/dgetrs that
 * A\b uses, generalized_poisson.jl:201-203), in place on a copy of A.
 * Returns 0, or 1 if a pivot is exactly zero. */
EXPORT int ref_lu_solve(int n, const double *Ain, const double *b, double *x) {
  double *A = (double *)malloc(sizeof(double) * (size_t)n * n);
  int *piv = (int *)malloc(sizeof(int) * n);
  memcpy(A, Ain, sizeof(double) * (size_t)n * n);
  size_t nn = (size_t)n;
  int sing = 0;
  for (int k = 0; k < n; ++k) {
    int pr = k;
    double mx = fabs(AE(k, k));
    for (int r = k + 1; r < n; ++r)
      if (fabs(AE(r, k)) > mx) { mx = fabs(AE(r, k)); pr = r; }
    piv[k] = pr;
    if (pr != k)
      for (int c = 0; c < n; ++c) { double t = AE(k, c); AE(k, c) = AE(pr, c); AE(pr, c) = t; }
    if (AE(k, k) == 0.0) { sing = 1; continue; }
    double inv = 1.0 / AE(k, k);
    for (int r = k + 1; r < n; ++r) AE(r, k) *= inv;
    for (int c = k + 1; c < n; ++c) {
      double akc = AE(k, c);
      if (akc == 0.0) continue;
      for (int r = k + 1; r < n; ++r) AE(r, c) -= AE(r, k) * akc;
    }
  }
  for (int k = 0; k < n; ++k) x[k] = b[k];
  for (int k = 0; k < n; ++k) if (piv[k] != k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
  for (int k = 0; k < n; ++k) for (int r = k + 1; r < n; ++r) x[r] -= AE(r, k) * x[k];
  for (int k = n - 1; k >= 0; --k) {
    x[k] /= AE(k, k);
    for (int r = 0; r < k; ++r) x[r] -= AE(r, k) * x[k];
  }
  free(A);
  free(piv);
  return sing;
}
#undef AE

/* calculate_electric_potential for non-periodic grids of any size: the same
 * rows as ref_poisson_assemble + ref_apply_dirichlet (generalized_poisson.jl:
 * 42-68,205-215), stored as a band of half-width nx and eliminated without
 * pivoting (every row is diagonally dominant: Laplacian rows or identity
 * rows), then back-substituted.  mask[k] = tag (0 free), tagval[tag] = value;
 * b[k] = -rho[k]/eps0 at free nodes.  Returns 0 or 1 on a zero pivot. */
EXPORT int ref_poisson_band_solve(int nx, int ny, double dx, const int8_t *mask,
                                  const double *tagval, const double *rho, double eps0,
                                  double *x) {
  const int n = nx * ny, bw = nx, w = 2 * bw + 1;
  double *B = (double *)calloc((size_t)n * w, sizeof(double));
  double *b = (double *)malloc(sizeof(double) * n);
  if (!B || !b) { free(B); free(b); return 2; }
#define BE(r, c) B[(size_t)(r) * w + ((c) - (r) + bw)]
  double d2 = dx * dx;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      int k = i + nx * j;
      if (mask[k]) {
        BE(k, k) = 1.0;
        b[k] = tagval[mask[k]];
        continue;
      }
      double dg = 0.0;
      if (i < nx - 1) { dg -= 1.0; BE(k, k + 1) += 1.0 / d2; }
      if (i > 0) { dg -= 1.0; BE(k, k - 1) += 1.0 / d2; }
      if (j < ny - 1) { dg -= 1.0; BE(k, k + nx) += 1.0 / d2; }
      if (j > 0) { dg -= 1.0; BE(k, k - nx) += 1.0 / d2; }
      BE(k, k) = dg / d2;
      b[k] = -rho[k] / eps0;
    }
  int sing = 0;
  for (int k = 0; k < n; ++k) {
    double piv = BE(k, k);
    if (piv == 0.0) { sing = 1; break; }
    int rmax = k + bw < n - 1 ? k + bw : n - 1;
    for (int r = k + 1; r <= rmax; ++r) {
      double a = BE(r, k);
      if (a == 0.0) continue;
      double l = a / piv;
      BE(r, k) = 0.0;
      for (int c = k + 1; c <= rmax; ++c) BE(r, c) -= l * BE(k, c);
      b[r] -= l * b[k];
    }
  }
  if (!sing)
    for (int k = n - 1; k >= 0; --k) {
      double s = b[k];
      int cmax = k + bw < n - 1 ? k + bw : n - 1;
      for (int c = k + 1; c <= cmax; ++c) s -= BE(k, c) * x[c];
      x[k] = s / BE(k, k);
    }
#undef BE
  free(B);
  free(b);
  return sing;
}

/* calculate_electric_field! (generalized_poisson.jl:385-410): E is nx x ny x 3
 * column-major; E[...,3] = 0; 1-D fills only component 1. */
EXPORT void ref_efield(int dim, int nx, int ny, double dx, double dy, const double *phi,
                       double *E) {
  size_t nn = (size_t)nx * ny;
  memset(E, 0, sizeof(double) * nn * 3);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      int k = i + nx * j;
      double ex;
      if (i == 0) ex = (phi[k] - phi[k + 1]) / dx;
      else if (i == nx - 1) ex = (phi[k - 1] - phi[k]) / dx;
      else ex = (phi[k - 1] - phi[k + 1]) / (2.0 * dx);
      E[k] = ex;
      if (dim == 2) {
        double ey;
        if (j == 0) ey = (phi[k] - phi[k + nx]) / dy;
        else if (j == ny - 1) ey = (phi[k - nx] - phi[k]) / dy;
        else ey = (phi[k - nx] - phi[k + nx]) / (2.0 * dy);
        E[nn + k] = ey;
      }
    }
}

/* CrossSection call (Chemistry/src/cross_section.jl:3-17): linear
 * interpolation between nodes, flat outside */
EXPORT double ref_sigma(int n, const double *eps, const double *sig, double e) {
  if (e <= eps[0]) return sig[0];
  if (e >= eps[n - 1]) return sig[n - 1];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (eps[mid] <= e) lo = mid; else hi = mid;
  }
  double t = (e - eps[lo]) / (eps[hi] - eps[lo]);
  return sig[lo] * (1.0 - t) + sig[hi] * t;
}

/* MonteCarloCollisions constructor (Chemistry/src/mcc.jl:27-51): max over the
 * union of table energies (and 0) of sum_k sigma_k(eps) sqrt(2 eps / m) */
EXPORT double ref_max_sigma_g(int ncoll, const int32_t *off, const double *eps,
                              const double *sig, double m_eV, int nunion, double *work) {
  (void)nunion;
  int cnt = 0;
  work[cnt++] = 0.0;
  for (int k = 0; k < ncoll; ++k)
    for (int q = off[k]; q < off[k + 1]; ++q) work[cnt++] = eps[q];
  /* sort + unique (insertion sort: tables are small) */
  for (int a = 1; a < cnt; ++a) {
    double v = work[a];
    int b = a - 1;
    while (b >= 0 && work[b] > v) { work[b + 1] = work[b]; --b; }
    work[b + 1] = v;
  }
  int u = 0;
  for (int a = 0; a < cnt; ++a)
    if (u == 0 || work[a] != work[u - 1]) work[u++] = work[a];
  double alpha = sqrt(2.0 / m_eV), best = 0.0;
  for (int a = 0; a < u; ++a) {
    double v = alpha * sqrt(work[a]);
    double s = 0.0;
    for (int k = 0; k < ncoll; ++k)
      s += ref_sigma(off[k + 1] - off[k], eps + off[k], sig + off[k], work[a]) * v;
    if (s > best) best = s;
  }
  return best;
}

/* ======================================================================== */
/* mirror mode: the GPU number format and operation order                   */
/* ======================================================================== */

static inline int mir_advance(float v, float dtd, float *o) {
  float h = fmaf(v, dtd, *o);
  float f = floorf(h);
  h = h - f;
  if (h >= 1.0f) { h = 0.0f; f += 1.0f; }
  *o = h;
  return (int)f;
}

static inline int mir_bnd(int *i, int n, int mode) {
  if (*i >= 0 && *i < n) return 1;
  if (mode == 1) {
    int r = *i % n;
    if (r < 0) r += n;
    *i = r;
    return 1;
  }
  return 0;
}

/* CIC gather of E (FP32) at a particle: the GPU's formula */
static inline void mir_gather1(int dim, const float *E, int nx, int i, int j, float hx, float hy,
                               float *Ex, float *Ey) {
  float ax = 1.0f - hx;
  if (dim == 2) {
    float ay = 1.0f - hy;
    float w00 = ax * ay, w10 = hx * ay, w01 = ax * hy, w11 = hx * hy;
    int q = j * nx + i;
    const float *e00 = E + 2 * q, *e10 = E + 2 * (q + 1), *e01 = E + 2 * (q + nx),
                *e11 = E + 2 * (q + nx + 1);
    *Ex = fmaf(w11, e11[0], fmaf(w01, e01[0], fmaf(w10, e10[0], w00 * e00[0])));
    *Ey = fmaf(w11, e11[1], fmaf(w01, e01[1], fmaf(w10, e10[1], w00 * e00[1])));
  } else {
    *Ex = fmaf(hx, E[i + 1], ax * E[i]);
    *Ey = 0.0f;
  }
}

EXPORT void mir_gather(int dim, int64_t np, const int32_t *cell, const float *ox, const float *oy,
                       const float *E, int nx, float *outx, float *outy) {
  for (int64_t p = 0; p < np; ++p) {
    float ex = 0.f, ey = 0.f;
    if (cell[p] >= 0) {
      int i = cell[p] & 0xFFFF, j = cell[p] >> 16;
      mir_gather1(dim, E, nx, i, j, ox[p], dim == 2 ? oy[p] : 0.f, &ex, &ey);
    }
    outx[p] = ex;
    if (outy) outy[p] = ey;
  }
}

/* fixed-point CIC deposit of one particle into cnt (nodes) */
static inline void mir_deposit1(int dim, int nx, int i, int j, float hx, float hy, int64_t *cnt) {
  if (dim == 2) {
    double dxh = (double)hx, dyh = (double)hy;
    int64_t Wx = llrint(dxh * FIX_ONE), Wy = llrint(dyh * FIX_ONE);
    int64_t Wxy = llrint((dxh * dyh) * FIX_ONE);
    int q = j * nx + i;
    cnt[q] += ((int64_t)1 << 32) - Wx - Wy + Wxy;
    cnt[q + 1] += Wx - Wxy;
    cnt[q + nx] += Wy - Wxy;
    cnt[q + nx + 1] += Wxy;
  } else {
    int64_t W1 = llrint((double)hx * FIX_ONE);
    cnt[i] += ((int64_t)1 << 32) - W1;
    cnt[i + 1] += W1;
  }
}

/* gather + push (B = 0) + boundary + deposit, the GPU's arithmetic.
 * cnt is accumulated (caller zeroes it).  Returns the live count. */
EXPORT int64_t mir_push_deposit(int dim, int64_t np, int32_t *cell, float *ox, float *oy,
                                float *vx, float *vy, const float *E, int nx, int ncx, int ncy,
                                float qmdt, float dtdx, float dtdy, int bndx, int bndy,
                                int do_push, int64_t *cnt) {
  int64_t live = 0;
  for (int64_t p = 0; p < np; ++p) {
    if (cell[p] < 0) continue;
    int i = cell[p] & 0xFFFF, j = cell[p] >> 16;
    float hx = ox[p], hy = dim == 2 ? oy[p] : 0.f;
    int alive = 1;
    if (do_push) {
      float Ex, Ey;
      mir_gather1(dim, E, nx, i, j, hx, hy, &Ex, &Ey);
      vx[p] = fmaf(qmdt, Ex, vx[p]);
      i += mir_advance(vx[p], dtdx, &hx);
      alive = mir_bnd(&i, ncx, bndx);
      if (dim == 2) {
        vy[p] = fmaf(qmdt, Ey, vy[p]);
        j += mir_advance(vy[p], dtdy, &hy);
        alive = mir_bnd(&j, ncy, bndy) && alive;
      }
      cell[p] = alive ? (i | (j << 16)) : -1;
      ox[p] = hx;
      if (dim == 2) oy[p] = hy;
    }
    if (alive) {
      ++live;
      mir_deposit1(dim, nx, i, j, hx, hy, cnt);
    }
  }
  return live;
}

/* the RZ transform in cell units as the device computes it (FP64 from the
 * pushed FP32 velocity; see mir_push_deposit_rz) */
static inline void mir_rz_transform(int *i, float *hx, float *vr, float *vt, double dtdr) {
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
    if (h >= 1.0f) { h = 0.0f; ++ii; }
  } else {
    ii = 1 << 30;
    h = 0.0f;
  }
  *i = ii;
  *hx = h;
  *vr = (float)nvr;
  *vt = (float)nvt;
}

/* mir_push_deposit for the RZ pusher on a 2-D (r, z) grid: the radial
 * coordinate goes through the transform instead of the Cartesian step, and
 * v_theta (vz) turns with v_r. */
EXPORT int64_t mir_push_deposit_rz(int64_t np, int32_t *cell, float *ox, float *oy, float *vx,
                                   float *vy, float *vz, const float *E, int nx, int ncx, int ncy,
                                   float qmdt, double dtdr, float dtdz, int bndr, int bndz,
                                   int64_t *cnt) {
  int64_t live = 0;
  for (int64_t p = 0; p < np; ++p) {
    if (cell[p] < 0) continue;
    int i = cell[p] & 0xFFFF, j = cell[p] >> 16;
    float hx = ox[p], hy = oy[p];
    float Ex, Ey;
    mir_gather1(2, E, nx, i, j, hx, hy, &Ex, &Ey);
    vx[p] = fmaf(qmdt, Ex, vx[p]);
    mir_rz_transform(&i, &hx, &vx[p], &vz[p], dtdr);
    int alive = mir_bnd(&i, ncx, bndr);
    vy[p] = fmaf(qmdt, Ey, vy[p]);
    j += mir_advance(vy[p], dtdz, &hy);
    alive = mir_bnd(&j, ncy, bndz) && alive;
    cell[p] = alive ? (i | (j << 16)) : -1;
    ox[p] = hx;
    oy[p] = hy;
    if (alive) {
      ++live;
      mir_deposit1(2, nx, i, j, hx, hy, cnt);
    }
  }
  return live;
}

/* the device's surface tracer (track_path in particles.hip): the reference's
 * track!/check!/hit! (pic/surfaces/track.jl:152-161, check.jl:17-68,
 * hit.jl:1-56, circuit_coupling.jl:47-62) in FP64 from the FP32 state.
 * surf: 4 codes per cell (-x, +x, -y, +y): 0 none, 1 absorb, 2 reflect,
 * 3 pass, 16 + k floating electrode k.  Returns 1 if absorbed (electrode k
 * in *elec, else -1). */
#define MIR_TRACK_MAX 64
static int mir_track_path(const uint8_t *surf, int ncx, int ncy, double dt, double h, int i, int j,
                          float hx0, float hy0, float *vx, float *vy, double *cx, double *cy,
                          int *elec) {
  double T = dt, hx = hx0, hy = hy0;
  *cx = 0.0;
  *cy = 0.0;
  *elec = -1;
  for (int it = 0; it < MIR_TRACK_MAX; ++it) {
    const double vxd = *vx, vyd = *vy;
    const double dxs = vxd > 0 ? h * (1.0 - hx) : h * hx;
    const double dys = vyd > 0 ? h * (1.0 - hy) : h * hy;
    const double tx = dxs / fabs(vxd), ty = dys / fabs(vyd);
    if (T < tx && T < ty) return 0;
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
    if (i < 0 || i >= ncx || j < 0 || j >= ncy) return 0;
    const int code = surf[(i + ncx * j) * 4 + dir];
    if (code == 0) {
      i = ni; j = nj; hx = nhx; hy = nhy; T = Tn;
      continue;
    }
    if (code == 1) return 1;
    if (code >= 16) { *elec = code - 16; return 1; }
    if (code != 2) return 0;
    if (ni != i) { *vx = -*vx; *cx += 2.0 * (double)*vx * Tn / h; }
    if (nj != j) { *vy = -*vy; *cy += 2.0 * (double)*vy * Tn / h; }
    int si = ni, sj = nj;
    double shx = nhx, shy = nhy;
    if (nhx == 0.0) { si = ni - 1; shx = 1.0; }
    if (nhx == 1.0) { si = ni + 1; shx = 0.0; }
    if (nhy == 0.0) { sj = nj - 1; shy = 1.0; }
    if (nhy == 1.0) { sj = nj + 1; shy = 0.0; }
    i = si; j = sj; hx = shx; hy = shy; T = Tn;
  }
  return 0;
}

static inline int mir_shift(double c, float *o) {
  const double t = (double)*o + c;
  const double f = floor(t);
  float h = (float)(t - f);
  int k = (int)f;
  if (h >= 1.0f) { h = 0.0f; ++k; }
  *o = h;
  return k;
}

/* mir_push_deposit (2-D) with the surface tracker of the push kernels:
 * particles starting in a tracked cell are traced; ehit[k] counts the
 * particles floating electrode k absorbed. */
EXPORT int64_t mir_push_deposit_track(int64_t np, int32_t *cell, float *ox, float *oy, float *vx,
                                      float *vy, const float *E, int nx, int ncx, int ncy,
                                      float qmdt, float dtdx, float dtdy, int bndx, int bndy,
                                      const uint8_t *surf, const uint8_t *tracked, double dt,
                                      double h, uint64_t *ehit, int64_t *cnt) {
  int64_t live = 0;
  for (int64_t p = 0; p < np; ++p) {
    if (cell[p] < 0) continue;
    int i = cell[p] & 0xFFFF, j = cell[p] >> 16;
    float hx = ox[p], hy = oy[p];
    float Ex, Ey;
    mir_gather1(2, E, nx, i, j, hx, hy, &Ex, &Ey);
    vx[p] = fmaf(qmdt, Ex, vx[p]);
    vy[p] = fmaf(qmdt, Ey, vy[p]);
    const float vx0 = vx[p], vy0 = vy[p];
    double cx = 0.0, cy = 0.0;
    int absorbed = 0, el = -1;
    if (tracked[i + ncx * j])
      absorbed = mir_track_path(surf, ncx, ncy, dt, h, i, j, hx, hy, &vx[p], &vy[p], &cx, &cy, &el);
    if (absorbed && el >= 0) ehit[el]++;
    i += mir_advance(vx0, dtdx, &hx);
    j += mir_advance(vy0, dtdy, &hy);
    if (cx != 0.0) i += mir_shift(cx, &hx);
    if (cy != 0.0) j += mir_shift(cy, &hy);
    int alive = mir_bnd(&i, ncx, bndx);
    alive = mir_bnd(&j, ncy, bndy) && alive;
    if (absorbed) alive = 0;
    cell[p] = alive ? (i | (j << 16)) : -1;
    ox[p] = hx;
    oy[p] = hy;
    if (alive) {
      ++live;
      mir_deposit1(2, nx, i, j, hx, hy, cnt);
    }
  }
  return live;
}

/* rho = sum_s q_s n_s + bg, n_s = (cnt_s 2^-32) w0_s / V (GPU order) */
EXPORT void mir_rho(int nn, int ns, const int64_t *cnt /* ns x nn */, const double *w0,
                    const double *q, const double *vol, const double *bg, double *rho) {
  for (int k = 0; k < nn; ++k) {
    double r = 0.0;
    for (int s = 0; s < ns; ++s) {
      double n = ((double)cnt[(size_t)s * nn + k] * (1.0 / FIX_ONE)) * w0[s] / vol[k];
      r += q[s] * n;
    }
    rho[k] = r + bg[k];
  }
}

/* 1-D tridiagonal factor of the reference matrix scaled by dx^2
 * (generalized_poisson.jl:22-25,42-68,205-215,271-284): a = sub-diagonal,
 * inv = 1/pivot, cp = super-diagonal * inv.  singular = periodic ring with
 * no Dirichlet node: rows 1..n-1 with node 0 pinned. */
EXPORT int mir_tri_factor(int n, const int8_t *tags, int singular, double *a, double *cp,
                          double *inv) {
  for (int i = 0; i < n; ++i) { a[i] = 0.0; cp[i] = 0.0; inv[i] = 0.0; }
  double cprev = 0.0;
  int i0 = singular ? 1 : 0;
  for (int i = i0; i < n; ++i) {
    double ai, ci, di;
    if (!singular && tags[i]) {
      ai = 0.0; ci = 0.0; di = 1.0;
    } else if (!singular) {
      ai = i > 0 ? 1.0 : 0.0;
      ci = i < n - 1 ? 1.0 : 0.0;
      di = -(ai + ci);
    } else {
      ai = i > 1 ? 1.0 : 0.0;
      ci = i < n - 1 ? 1.0 : 0.0;
      di = -2.0;
    }
    double den = di - ai * cprev;
    if (den == 0.0) return 1;
    a[i] = ai;
    inv[i] = 1.0 / den;
    cp[i] = ci * inv[i];
    cprev = cp[i];
  }
  return 0;
}

/* The device's work-group solve of the factored system (systems with
 * Dirichlet nodes and n <= MIR_TRI_T * MIR_TRI_K): each recurrence is a
 * prefix composition of affine maps (A, B), (A2, B2) after (A1, B1) =
 * (A2 A1, fma(A2, B1, B2)), computed as segments of MIR_TRI_K maps per
 * thread, an inclusive Hillis-Steele scan over the MIR_TRI_T segment maps
 * (every step reads the previous step's values), and one fma per element. */
#define MIR_TRI_T 256
#define MIR_TRI_K 8
static void mir_tri_scan(const double *ma, const double *mb, int n, int rev, double *out) {
  double la[MIR_TRI_T][MIR_TRI_K], lb[MIR_TRI_T][MIR_TRI_K];
  double ga[MIR_TRI_T], gb[MIR_TRI_T], pa[MIR_TRI_T], pb[MIR_TRI_T];
  for (int t = 0; t < MIR_TRI_T; ++t) {
    double ca = 1.0, cb = 0.0;
    for (int k = 0; k < MIR_TRI_K; ++k) {
      int j = t * MIR_TRI_K + k;
      if (j < n) {
        int i = rev ? n - 1 - j : j;
        cb = fma(ma[i], cb, mb[i]);
        ca = ma[i] * ca;
      }
      la[t][k] = ca;
      lb[t][k] = cb;
    }
    ga[t] = ca;
    gb[t] = cb;
  }
  for (int d = 1; d < MIR_TRI_T; d <<= 1) {
    for (int t = 0; t < MIR_TRI_T; ++t) {
      pa[t] = t >= d ? ga[t - d] : 1.0;
      pb[t] = t >= d ? gb[t - d] : 0.0;
    }
    for (int t = d; t < MIR_TRI_T; ++t) {
      double a2 = ga[t], b2 = gb[t];
      gb[t] = fma(a2, pb[t], b2);
      ga[t] = a2 * pa[t];
    }
  }
  for (int t = 0; t < MIR_TRI_T; ++t) {
    double vin = t > 0 ? gb[t - 1] : 0.0;
    for (int k = 0; k < MIR_TRI_K; ++k) {
      int j = t * MIR_TRI_K + k;
      if (j < n) out[rev ? n - 1 - j : j] = fma(la[t][k], vin, lb[t][k]);
    }
  }
}

/* the device's tri_sum: per-thread sums of MIR_TRI_K consecutive elements,
 * then a pairwise tree over the MIR_TRI_T partials */
static double mir_tri_sum(const double *v, int n) {
  double g[MIR_TRI_T];
  for (int t = 0; t < MIR_TRI_T; ++t) {
    double s = 0.0;
    for (int k = 0; k < MIR_TRI_K; ++k)
      if (t * MIR_TRI_K + k < n) s += v[t * MIR_TRI_K + k];
    g[t] = s;
  }
  for (int d = MIR_TRI_T / 2; d > 0; d >>= 1)
    for (int t = 0; t < d; ++t) g[t] += g[t + d];
  return g[0];
}

EXPORT void mir_tri_solve(int n, const double *rho, const int8_t *tags, const double *tagval,
                          const double *a, const double *cp, const double *inv, double scale,
                          int singular, double *phi) {
  double *y = (double *)malloc(sizeof(double) * n);
  if (!singular && n <= MIR_TRI_T * MIR_TRI_K) {
    double *mA = (double *)malloc(sizeof(double) * n), *mB = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
      int tg = tags[i];
      double r = tg ? tagval[tg] : -(rho[i] * scale);
      mA[i] = -a[i] * inv[i];
      mB[i] = r * inv[i];
    }
    mir_tri_scan(mA, mB, n, 0, y);
    for (int i = 0; i < n; ++i) mA[i] = -cp[i];
    mir_tri_scan(mA, y, n, 1, phi);
    free(mA);
    free(mB);
  } else if (singular && n <= MIR_TRI_T * MIR_TRI_K) {
    /* the device's ring scan: tree sums, node 0 pinned, rows 1..n-1 */
    const int m = n - 1;
    double *f = (double *)malloc(sizeof(double) * n);
    double *mA = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
    double *mB = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
    for (int i = 0; i < n; ++i) f[i] = -(rho[i] * scale);
    const double mean = mir_tri_sum(f, n) / (double)n;
    for (int j = 0; j < m; ++j) {
      mA[j] = -a[j + 1] * inv[j + 1];
      mB[j] = (f[j + 1] - mean) * inv[j + 1];
    }
    mir_tri_scan(mA, mB, m, 0, y);
    for (int j = 0; j < m; ++j) mA[j] = -cp[j + 1];
    mir_tri_scan(mA, y, m, 1, mB);
    f[0] = 0.0;
    for (int i = 1; i < n; ++i) f[i] = mB[i - 1];
    const double m2 = mir_tri_sum(f, n) / (double)n;
    for (int i = 0; i < n; ++i) phi[i] = f[i] - m2;
    free(f);
    free(mA);
    free(mB);
  } else if (!singular) {
    double yprev = 0.0;
    for (int i = 0; i < n; ++i) {
      int tg = tags[i];
      double r = tg ? tagval[tg] : -(rho[i] * scale);
      double v = fma(-a[i], yprev, r) * inv[i];
      y[i] = v;
      yprev = v;
    }
    double xn = 0.0;
    for (int i = n - 1; i >= 0; --i) {
      double v = fma(-cp[i], xn, y[i]);
      phi[i] = v;
      xn = v;
    }
  } else {
    double sum = 0.0;
    for (int i = 0; i < n; ++i) sum += -(rho[i] * scale);
    double mean = sum / (double)n;
    double yprev = 0.0;
    for (int i = 1; i < n; ++i) {
      double r = -(rho[i] * scale) - mean;
      double v = fma(-a[i], yprev, r) * inv[i];
      y[i] = v;
      yprev = v;
    }
    double xn = 0.0;
    for (int i = n - 1; i >= 1; --i) {
      double v = fma(-cp[i], xn, y[i]);
      phi[i] = v;
      xn = v;
    }
    phi[0] = 0.0;
    double s2 = 0.0;
    for (int i = 0; i < n; ++i) s2 += phi[i];
    double m2 = s2 / (double)n;
    for (int i = 0; i < n; ++i) phi[i] = phi[i] - m2;
  }
  free(y);
}

/* E = -grad phi in FP64, stored FP32 (1-D: nx values; 2-D: interleaved) */
EXPORT void mir_efield(int dim, int nx, int ny, double dx, double dy, const double *phi, float *E) {
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      int n = j * nx + i;
      double ex;
      if (nx < 2) ex = 0.0;
      else if (i == 0) ex = (phi[n] - phi[n + 1]) / dx;
      else if (i == nx - 1) ex = (phi[n - 1] - phi[n]) / dx;
      else ex = (phi[n - 1] - phi[n + 1]) / (2.0 * dx);
      if (dim == 1) { E[n] = (float)ex; continue; }
      double ey;
      if (ny < 2) ey = 0.0;
      else if (j == 0) ey = (phi[n] - phi[n + nx]) / dy;
      else if (j == ny - 1) ey = (phi[n - nx] - phi[n]) / dy;
      else ey = (phi[n - nx] - phi[n + nx]) / (2.0 * dy);
      E[2 * n] = (float)ex;
      E[2 * n + 1] = (float)ey;
    }
}

/* ---- MCC mirror ---- */
typedef struct {
  uint32_t id, step, stream, k0, k1;
  int block, half;
  uint32_t cur[4];
} MirDraws;

static double md_next(MirDraws *d) {
  if (d->half == 0) {
    uint32_t c[4] = {d->id, d->step, d->stream, (uint32_t)d->block};
    uint32_t k[2] = {d->k0, d->k1};
    orc_philox4x32(c, k, d->cur);
    d->block++;
    d->half = 1;
    return orc_u01(d->cur[0], d->cur[1]);
  }
  d->half = 0;
  return orc_u01(d->cur[2], d->cur[3]);
}

/* euler_angles (mcc.jl:89-106) */
static void m_angles(double vx, double vy, double vz, double *st, double *ct, double *sp, double *cp) {
  double nv = sqrt(vx * vx + vy * vy + vz * vz);
  *ct = vz / nv;
  *st = sqrt(1.0 - *ct * *ct);
  if (*st == 0.0) { *cp = 1.0; *sp = 0.0; }
  else { *cp = vx / nv / *st; *sp = vy / nv / *st; }
}

/* [sX cE, sX sE, cX] * unrotated (mcc.jl:83-87,108-113) */
static void m_rotate(double st, double ct, double sp, double cp, double sX, double cX, double sE,
                     double cE, double *r) {
  double a0 = sX * cE, a1 = sX * sE, a2 = cX;
  r[0] = a0 * (cp * ct) + a1 * (-sp) + a2 * (cp * st);
  r[1] = a0 * (-sp * ct) + a1 * cp + a2 * (sp * st);
  r[2] = a0 * (-st) + a1 * 0.0 + a2 * ct;
}

/* scattering direction [sin chi cos eta, sin chi sin eta, cos chi] *
 * unrotated(euler_angles(v)) (mcc.jl:53-72,83-127) from two uniforms:
 * kind 0 is the reference's "isotropic_distribution" (chi = 2 pi u1),
 * kind 1 its "cosine_distribution" of diffuse_reflection (sin chi = sqrt(u1),
 * cos chi = -sqrt(1 - sin chi^2)); eta = 2 pi u2 for both.  unrotated is the
 * reference's matrix as written, which is not orthogonal (its first row holds
 * -sin phi cos theta), so r is a unit vector only in special cases. */
EXPORT void mir_scatter(int kind, double vx, double vy, double vz, double u1, double u2,
                        double *r) {
  double st, ct, sp, cp, sX, cX, sE, cE;
  m_angles(vx, vy, vz, &st, &ct, &sp, &cp);
  if (kind == 0) {
    isk_sincos2pi(u1, &sX, &cX);
  } else {
    sX = sqrt(u1);
    cX = -sqrt(1.0 - sX * sX);
  }
  isk_sincos2pi(u2, &sE, &cE);
  m_rotate(st, ct, sp, cp, sX, cX, sE, cE, r);
}

/* maxwellian_velocity / randn(3) (mcc.jl:79-81) from four uniforms by
 * Box-Muller: (r1 cos, r1 sin) of the first pair, r2 cos of the second */
static void m_randn3(const double *u, double *n) {
  double r1 = sqrt(-2.0 * isk_log(u[0])), r2 = sqrt(-2.0 * isk_log(u[2]));
  double s1, c1, s2, c2;
  isk_sincos2pi(u[1], &s1, &c1);
  isk_sincos2pi(u[3], &s2, &c2);
  n[0] = r1 * c1;
  n[1] = r1 * s1;
  n[2] = r2 * c2;
}

/* One collision of the reference's kernels (mcc.jl:129-229), FP64, with the
 * uniforms u[] that follow the candidate and process draws, consumed in the
 * reference's order:
 *  ELASTIC_ISOTROPIC (0) / ELASTIC_BACKWARD (1) (mcc.jl:129-142,159-172):
 *     u[0..3] target velocity tv = vth_t randn(3), u[4..5] the direction of
 *     the relative velocity vr = v - tv; v' = w + mr2 |vr| dir with the
 *     centre-of-mass velocity w = mr1 v + mr2 tv;
 *  EXCITATION (2) (mcc.jl:215-229): v' = sqrt(2 (eps - E)/m) iso(v), u[0..1];
 *  IONIZATION (3) (mcc.jl:174-213): split fraction u[0]; primary
 *     e1v diffuse(v) with u[1..2]; secondary e2v diffuse(v') with u[3..4]
 *     (relative to the updated primary); ions' velocity vth_t randn(3) from
 *     u[5..8];
 *  INELASTIC_BACKWARD (4) (mcc.jl:144-157): as ELASTIC_BACKWARD with
 *     |vr| scaled by the uniform u[4] (drawn before the direction, u[5..6]).
 * out[0..2] velocity after the collision; out[3..5] the target velocity drawn
 * (elastic kinds) or the ions' velocity (ionisation); out[6..8] the secondary
 * electron.  Returns the number of uniforms used, or 0 when the energy is
 * below the threshold (mcc.jl:178-182,219-223: velocity unchanged). */
EXPORT int mir_coll_kin(int type, double ms, double mt, double m_eV, double energy,
                        double vth_t, const double *v, const double *u, double *out) {
  const double v0 = v[0], v1 = v[1], v2 = v[2];
  if (type == 0 || type == 1 || type == 4) {
    double mr1 = ms / (ms + mt), mr2 = mt / (ms + mt);
    double tv[3], vr[3], w[3], dir[3];
    m_randn3(u, tv);
    for (int q = 0; q < 3; ++q) tv[q] *= vth_t;
    for (int q = 0; q < 3; ++q) { vr[q] = v[q] - tv[q]; w[q] = mr1 * v[q] + mr2 * tv[q]; }
    double nvr = sqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
    int used = 6;
    if (type == 4) {
      nvr = u[4] * nvr;
      mir_scatter(1, vr[0], vr[1], vr[2], u[5], u[6], dir);
      used = 7;
    } else {
      mir_scatter(type, vr[0], vr[1], vr[2], u[4], u[5], dir);
    }
    for (int q = 0; q < 3; ++q) {
      out[q] = w[q] + mr2 * (nvr * dir[q]);
      out[3 + q] = tv[q];
    }
    return used;
  }
  double sE = (0.5 * m_eV) * (v0 * v0 + v1 * v1 + v2 * v2) - energy;
  if (sE < 0.0) return 0;
  if (type == 2) {
    double ev = sqrt(2.0 / m_eV) * sqrt(sE);
    double dir[3];
    mir_scatter(0, v0, v1, v2, u[0], u[1], dir);
    for (int q = 0; q < 3; ++q) out[q] = ev * dir[q];
    return 2;
  }
  double e1E = sE * u[0];
  double e2E = sE - e1E;
  double al = sqrt(2.0 / m_eV);
  double e1v = al * sqrt(e1E), e2v = al * sqrt(e2E);
  double dir[3], dir2[3], tv[3];
  mir_scatter(1, v0, v1, v2, u[1], u[2], dir);
  for (int q = 0; q < 3; ++q) out[q] = e1v * dir[q];
  mir_scatter(1, out[0], out[1], out[2], u[3], u[4], dir2);
  m_randn3(u + 5, tv);
  for (int q = 0; q < 3; ++q) {
    out[6 + q] = e2v * dir2[q];
    out[3 + q] = tv[q] * vth_t;
  }
  return 9;
}

typedef struct {
  int ncoll;
  int32_t types[8];
  double energy[8];
  const int32_t *off;
  const double *eps, *sig;
  double candP, dt, m_eV, ms, mt, vth_t, tqm;
  uint32_t k0, k1, step, stream;
  int dim, nx;
} MirMcc;

/* PIC.perform!(::MonteCarloCollisions) with per-particle candidates
 * (mcc.jl:231-289; kinematics mcc.jl:129-229).  Products (ionisation) are
 * returned in particle order: pcell/pox/poy/pe (3 per product)/pi (3 per
 * product)/ppid (the parent's id; nullable).  nu: nodes x ncoll counts
 * (accumulated), events: per process.  Returns the number of products, or -1
 * if some P_k > 1. */
EXPORT int64_t mir_mcc(const MirMcc *a, int64_t np, const int32_t *cell, const float *ox,
                       const float *oy, float *vx, float *vy, float *vz, const uint32_t *id,
                       const float *E, const double *tn, int64_t *nu, int64_t *events,
                       int32_t *pcell, float *pox, float *poy, float *pe, float *pi,
                       uint32_t *ppid) {
  int64_t nprod = 0;
  int err = 0;
  const int N = a->ncoll;
  for (int64_t p = 0; p < np; ++p) {
    int32_t c = cell[p];
    if (c < 0) continue;
    MirDraws d = {id[p], a->step, a->stream, a->k0, a->k1, 0, 0, {0, 0, 0, 0}};
    double Uc = md_next(&d);
    double U = md_next(&d);
    if (!(Uc < a->candP)) continue;
    int i = c & 0xFFFF, j = c >> 16;
    int node = a->dim == 2 ? j * a->nx + i : i;
    double n = tn[node];
    if (n < 0.0) continue;
    int k = (int)floor((double)N * U);
    if (k > N - 1) k = N - 1;
    double v0 = vx[p], v1 = vy[p], v2 = vz[p];
    double tvx = 0.0, tvy = 0.0;
    if (a->tqm != 0.0) {
      if (a->dim == 2) {
        tvx = a->tqm * (double)E[2 * node] * a->dt;
        tvy = a->tqm * (double)E[2 * node + 1] * a->dt;
      } else {
        tvx = a->tqm * (double)E[node] * a->dt;
      }
    }
    double gx = tvx - v0, gy = tvy - v1, gz = 0.0 - v2;
    double g = sqrt(gx * gx + gy * gy + gz * gz);
    double eps = (0.5 * a->m_eV) * (g * g);
    double sg = ref_sigma(a->off[k + 1] - a->off[k], a->eps + a->off[k], a->sig + a->off[k], eps) * g;
    double P = 1.0 - isk_exp(((-n) * sg) * a->dt);
    P /= a->candP;
    if (P > 1.0) err = 1;
    if (!(U > (double)(k + 1) / (double)N - P)) continue;
    nu[(size_t)node * N + k] += 1;
    events[k] += 1;
    /* the collision's uniforms: Philox blocks 1..5 of the particle */
    double u[10], sv[3] = {v0, v1, v2}, o[9];
    for (int q = 0; q < 10; ++q) u[q] = md_next(&d);
    int type = a->types[k];
    if (mir_coll_kin(type, a->ms, a->mt, a->m_eV, a->energy[k], a->vth_t, sv, u, o) == 0)
      continue;
    vx[p] = (float)o[0];
    vy[p] = (float)o[1];
    vz[p] = (float)o[2];
    if (type == 3) {
      pcell[nprod] = c;
      if (ppid) ppid[nprod] = id[p];
      pox[nprod] = ox[p];
      poy[nprod] = a->dim == 2 ? oy[p] : 0.f;
      for (int q = 0; q < 3; ++q) {
        pe[3 * nprod + q] = (float)o[6 + q];
        pi[3 * nprod + q] = (float)o[3 + q];
      }
      ++nprod;
    }
  }
  return err ? -1 : nprod;
}

/* Maxwellian loader: the GPU's draws and arithmetic (sources.jl:24-34) */
EXPORT void mir_sample(int dim, int64_t n, int64_t first_id, double x0, double y0, double dx,
                       double dy, int ncx, int ncy, const double *wx, const double *dxo,
                       const double *wv, const double *dv, uint64_t seed, uint32_t stream,
                       int32_t *cell, float *ox, float *oy, float *vx, float *vy, float *vz,
                       uint32_t *id) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t k = 0; k < n; ++k) {
    uint32_t c0[4] = {(uint32_t)k, (uint32_t)(k >> 32), stream, 0u}, r0[4], r1[4], r2[4];
    orc_philox4x32(c0, key, r0);
    c0[3] = 1u;
    orc_philox4x32(c0, key, r1);
    c0[3] = 2u;
    orc_philox4x32(c0, key, r2);
    double ux = orc_u01(r0[0], r0[1]);
    double px = ux * wx[0] + dxo[0];
    double s = (px - x0) / dx;
    double fl = floor(s);
    int i = (int)fl;
    float of = (float)(s - fl);
    if (of >= 1.0f) { of = 0.0f; i += 1; }
    if (i < 0) { i = 0; of = 0.0f; }
    if (i > ncx - 1) { i = ncx - 1; of = 0x1.fffffep-1f; }
    int j = 0;
    if (dim == 2) {
      double uy = orc_u01(r0[2], r0[3]);
      double py = uy * wx[1] + dxo[1];
      double sy = (py - y0) / dy;
      double fly = floor(sy);
      j = (int)fly;
      float ofy = (float)(sy - fly);
      if (ofy >= 1.0f) { ofy = 0.0f; j += 1; }
      if (j < 0) { j = 0; ofy = 0.0f; }
      if (j > ncy - 1) { j = ncy - 1; ofy = 0x1.fffffep-1f; }
      oy[k] = ofy;
    }
    double u1 = orc_u01(r1[0], r1[1]), u2 = orc_u01(r1[2], r1[3]);
    double u3 = orc_u01(r2[0], r2[1]), u4 = orc_u01(r2[2], r2[3]);
    double rad1 = sqrt(-2.0 * isk_log(u1)), rad2 = sqrt(-2.0 * isk_log(u3));
    double s1, c1, s2, c2;
    isk_sincos2pi(u2, &s1, &c1);
    isk_sincos2pi(u4, &s2, &c2);
    cell[k] = i | (j << 16);
    ox[k] = of;
    vx[k] = (float)(rad1 * c1 * wv[0] + dv[0]);
    vy[k] = (float)(rad1 * s1 * wv[1] + dv[1]);
    vz[k] = (float)(rad2 * c2 * wv[2] + dv[2]);
    if (id) id[k] = (uint32_t)(first_id + k);
  }
}

/* stable counting sort by key = tile of the cell (or the dead bin) with
 * compaction of discarded particles: the GPU's ordering rule.  Arrays are
 * permuted in place; returns the live count. */
/* stable counting sort of the live particles by tile (exact = 0) or by
 * (tile, cell inside the tile) (exact = 1: the device's two passes, cell in
 * tile then tile, both stable); discarded particles are dropped */
static int64_t mir_sort_key(int dim, int64_t np, int T, int ntx, int exact, int32_t *cell,
                            float *ox, float *oy, float *vx, float *vy, float *vz, uint32_t *id) {
  int log2T = 0;
  while ((1 << log2T) < T) ++log2T;
  int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (np > 0 ? np : 1));
  int maxkey = 0;
  int *key = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  for (int64_t p = 0; p < np; ++p) {
    int c = cell[p];
    if (c < 0) { key[p] = -1; continue; }
    int i = c & 0xFFFF, j = c >> 16;
    key[p] = (i >> log2T) + ntx * (j >> log2T);
    if (exact) {
      int local = (i & (T - 1)) + (dim == 2 ? ((j & (T - 1)) << log2T) : 0);
      key[p] = key[p] * (dim == 2 ? T * T : T) + local;
    }
    if (key[p] > maxkey) maxkey = key[p];
  }
  int64_t *cnt = (int64_t *)calloc((size_t)maxkey + 2, sizeof(int64_t));
  for (int64_t p = 0; p < np; ++p) if (key[p] >= 0) cnt[key[p] + 1]++;
  for (int k = 0; k <= maxkey; ++k) cnt[k + 1] += cnt[k];
  int64_t live = cnt[maxkey + 1];
  for (int64_t p = 0; p < np; ++p) if (key[p] >= 0) perm[cnt[key[p]]++] = p;
#define PERMUTE(T_, arr)                                                    \
  if (arr) {                                                                \
    T_ *tmp = (T_ *)malloc(sizeof(T_) * (live > 0 ? live : 1));              \
    for (int64_t q = 0; q < live; ++q) tmp[q] = arr[perm[q]];               \
    memcpy(arr, tmp, sizeof(T_) * live);                                    \
    free(tmp);                                                              \
  }
  PERMUTE(int32_t, cell)
  PERMUTE(float, ox)
  if (dim == 2) { PERMUTE(float, oy) }
  PERMUTE(float, vx)
  PERMUTE(float, vy)
  PERMUTE(float, vz)
  PERMUTE(uint32_t, id)
#undef PERMUTE
  free(perm);
  free(key);
  free(cnt);
  return live;
}

EXPORT int64_t mir_sort_tile(int dim, int64_t np, int T, int ntx, int32_t *cell, float *ox,
                             float *oy, float *vx, float *vy, float *vz, uint32_t *id) {
  return mir_sort_key(dim, np, T, ntx, 0, cell, ox, oy, vx, vy, vz, id);
}

EXPORT int64_t mir_sort_cell(int dim, int64_t np, int T, int ntx, int32_t *cell, float *ox,
                             float *oy, float *vx, float *vy, float *vz, uint32_t *id) {
  return mir_sort_key(dim, np, T, ntx, 1, cell, ox, oy, vx, vy, vz, id);
}

/* PIC.perform!(::DirectSimulationMonteCarlo) (Chemistry/src/dsmc.jl:81-142)
 * with one DSMC.ElasticCollision of equal weights (dsmc.jl:32-79), as the
 * device runs it: both species already sorted by exact cell (mir_sort_cell),
 * cell t = i + ncx j owns the draws (t, step, stream, block); per cell the
 * reference's Nc and remainder, then floor(Nc) trials in order.  rem: cells,
 * nu: nodes (accumulated).  Returns the accepted collisions. */
EXPORT int64_t mir_dsmc(int ncx, int ncy, int nx, int64_t npa, const int32_t *cella, float *avx,
                        float *avy, float *avz, int64_t npb, const int32_t *cellb, float *bvx,
                        float *bvy, float *bvz, int same, double Wa, double Wb, double dx,
                        double dy, double dt, double sgmax, double mr1, double mr2, int ntab,
                        const double *tg, const double *ts, double *rem, int64_t *nu,
                        uint32_t step, uint32_t stream, uint32_t k0, uint32_t k1) {
  const int ncell = ncx * ncy;
  int *rng = (int *)calloc((size_t)4 * ncell, sizeof(int));
  int *as = rng, *ae = rng + ncell, *bs = rng + 2 * ncell, *be = rng + 3 * ncell;
  for (int64_t p = 0; p < npa; ++p) {
    int c = cella[p];
    if (c < 0) continue;
    int key = (c & 0xFFFF) + ncx * (c >> 16);
    if (p == 0 || cella[p - 1] != c) as[key] = (int)p;
    if (p == npa - 1 || cella[p + 1] != c) ae[key] = (int)(p + 1);
  }
  if (same) {
    bs = as;
    be = ae;
  } else {
    for (int64_t p = 0; p < npb; ++p) {
      int c = cellb[p];
      if (c < 0) continue;
      int key = (c & 0xFFFF) + ncx * (c >> 16);
      if (p == 0 || cellb[p - 1] != c) bs[key] = (int)p;
      if (p == npb - 1 || cellb[p + 1] != c) be[key] = (int)(p + 1);
    }
  }
  double Pab, Pba;
  if (Wa > Wb) { Pab = Wb / Wa; Pba = 1.0; } else { Pab = 1.0; Pba = Wa / Wb; }
  int64_t total = 0;
  for (int t = 0; t < ncell; ++t) {
    const int a0 = as[t], Na = ae[t] - a0, b0 = bs[t], Nb = be[t] - b0;
    if (Na < 2 || Nb < 2) continue;
    const double na = (double)Na * Wa / (dx * dy);
    double Nc = na * (double)Nb * dt * sgmax;
    Nc /= Pab + (Wb / Wa) * Pba;
    if (!same) Nc *= 2.0;
    Nc += rem[t];
    const double fl = floor(Nc);
    rem[t] = Nc - fl;
    const long long ntr = (long long)fl;
    MirDraws d = {(uint32_t)t, step, stream, k0, k1, 0, 0, {0, 0, 0, 0}};
    int64_t acc = 0;
    for (long long k = 0; k < ntr; ++k) {
      int ia = (int)floor((double)Na * md_next(&d));
      if (ia > Na - 1) ia = Na - 1;
      const int sR = a0 + ia;
      int ib = (int)floor((double)Nb * md_next(&d));
      if (ib > Nb - 1) ib = Nb - 1;
      int tR = b0 + ib;
      while (!same && sR == tR) {
        ib = (int)floor((double)Nb * md_next(&d));
        if (ib > Nb - 1) ib = Nb - 1;
        tR = b0 + ib;
      }
      const double sx = avx[sR], sy = avy[sR], sz = avz[sR];
      const double tx = bvx[tR], ty = bvy[tR], tz = bvz[tR];
      const double gx = sx - tx, gy = sy - ty, gz = sz - tz;
      const double gn = sqrt(gx * gx + gy * gy + gz * gz);
      const double sg = ref_sigma(ntab, tg, ts, gn) * gn;
      const double P = sg / sgmax;
      const double R = md_next(&d);
      if (P < R) continue;
      const double B = 2.0 * md_next(&d) - 1.0;
      const double A = sqrt(1.0 - B * B);
      double sC, cC;
      orc_sincos2pi(md_next(&d), &sC, &cC);
      const double rx = gn * B, ry = gn * (A * cC), rz = gn * (A * sC);
      const double cx = mr1 * sx + mr2 * tx, cy = mr1 * sy + mr2 * ty, cz = mr1 * sz + mr2 * tz;
      avx[sR] = (float)(cx + mr2 * rx);
      avy[sR] = (float)(cy + mr2 * ry);
      avz[sR] = (float)(cz + mr2 * rz);
      bvx[tR] = (float)(cx - mr1 * rx);
      bvy[tR] = (float)(cy - mr1 * ry);
      bvz[tR] = (float)(cz - mr1 * rz);
      ++acc;
    }
    if (acc) {
      nu[(t / ncx) * nx + (t % ncx)] += acc;
      total += acc;
    }
  }
  free(rng);
  return total;
}

/*
 * api.hip -- the extern "C" entry points of libiskra_hip.so (declared in
 * include/iskra.h) and the host-side bookkeeping of a context.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "isk_fp64math.h"
#include "isk_internal.h"

static thread_local char g_err[1024] = "";

void isk_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void isk_clear_error(void) { g_err[0] = 0; }

int isk_hip_fail(hipError_t e, const char *what, const char *file, int line) {
  isk_set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what,
                file, line);
  return ISKRA_E_HIP;
}

int isk_alloc(void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 64;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    isk_set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    *p = nullptr;
    return ISKRA_E_HIP;
  }
  return ISKRA_OK;
}

int isk_free(void *p) {
  if (p) (void)hipFree(p);
  return ISKRA_OK;
}

int isk_nodes(const iskra_ctx *c) { return c->g.nx * c->g.ny; }

int isk_species_ok(iskra_ctx *c, int sid) {
  if (!c) {
    isk_set_error("null context");
    return 0;
  }
  if (sid < 0 || sid >= c->nspecies) {
    isk_set_error("species id %d out of range (have %d)", sid, c->nspecies);
    return 0;
  }
  return 1;
}

#define CHECK_CTX(c)                     \
  do {                                   \
    if (!(c)) {                          \
      isk_set_error("null context");     \
      return ISKRA_E_ARG;                \
    }                                    \
    g_err[0] = 0;                        \
    ISK_HIP(hipSetDevice((c)->device));  \
  } while (0)
#define CHECK_GRID(c)                                     \
  do {                                                    \
    if (!(c)->have_grid) {                                \
      isk_set_error("no grid: call iskra_grid_set first"); \
      return ISKRA_E_STATE;                               \
    }                                                     \
  } while (0)
#define RC(x)               \
  do {                      \
    int _rc = (x);          \
    if (_rc) return _rc;    \
  } while (0)

/* ---- species ---- */
void isk_species_release(IskSpecies *s) {
  for (int b = 0; b < 2; ++b) {
    isk_free(s->cell[b]);
    isk_free(s->ox[b]);
    isk_free(s->oy[b]);
    isk_free(s->vx[b]);
    isk_free(s->vy[b]);
    isk_free(s->vz[b]);
    isk_free(s->id[b]);
    s->cell[b] = nullptr;
    s->ox[b] = s->oy[b] = s->vx[b] = s->vy[b] = s->vz[b] = nullptr;
    s->id[b] = nullptr;
  }
  isk_free(s->hdr);
  isk_free(s->chunks);
  isk_free(s->tile_chunk_begin);
  isk_free(s->partial);
  isk_free(s->tilewin);
  isk_free(s->live_cnt);
  isk_free(s->gfall);
  isk_free(s->cnt);
  isk_free(s->ehit);
  s->ehit = nullptr;
  s->hdr = nullptr;
  s->chunks = nullptr;
  s->tile_chunk_begin = nullptr;
  s->partial = nullptr;
  s->tilewin = nullptr;
  s->live_cnt = nullptr;
  s->gfall = nullptr;
  s->cnt = nullptr;
}

static int alloc_chunk_tables(iskra_ctx *c, IskSpecies *s) {
  isk_free(s->chunks);
  isk_free(s->tile_chunk_begin);
  isk_free(s->partial);
  isk_free(s->tilewin);
  isk_free(s->live_cnt);
  s->chunks = nullptr;
  s->tile_chunk_begin = nullptr;
  s->partial = nullptr;
  s->tilewin = nullptr;
  s->live_cnt = nullptr;
  int WW = c->g.dim == 2 ? c->t.W * c->t.W : c->t.W;
  s->max_chunks = (s->cap + c->t.CH - 1) / c->t.CH + c->ntiles + 1;
  if (isk_alloc((void **)&s->chunks, (size_t)s->max_chunks * sizeof(IskChunk)) ||
      isk_alloc((void **)&s->tile_chunk_begin, (size_t)(c->ntiles + 1) * 4) ||
      isk_alloc((void **)&s->partial, (size_t)s->max_chunks * WW * 8) ||
      isk_alloc((void **)&s->tilewin, (size_t)c->ntiles * WW * 8) ||
      isk_alloc((void **)&s->live_cnt, (size_t)s->max_chunks * 4))
    return ISKRA_E_HIP;
  return ISKRA_OK;
}

int isk_species_alloc(iskra_ctx *c, IskSpecies *s) {
  size_t cap = (size_t)std::max<int64_t>(s->cap, 1);
  for (int b = 0; b < 2; ++b) {
    if (isk_alloc((void **)&s->cell[b], cap * 4) || isk_alloc((void **)&s->ox[b], cap * 4) ||
        isk_alloc((void **)&s->vx[b], cap * 4) || isk_alloc((void **)&s->vy[b], cap * 4) ||
        isk_alloc((void **)&s->vz[b], cap * 4))
      return ISKRA_E_HIP;
    if (c->g.dim == 2 && isk_alloc((void **)&s->oy[b], cap * 4)) return ISKRA_E_HIP;
    if (s->track_ids && isk_alloc((void **)&s->id[b], cap * 4)) return ISKRA_E_HIP;
  }
  int nn = isk_nodes(c);
  if (isk_alloc((void **)&s->hdr, HDR_WORDS * 8) ||
      isk_alloc((void **)&s->gfall, (size_t)nn * 8) ||
      isk_alloc((void **)&s->cnt, (size_t)nn * 8) ||
      isk_alloc((void **)&s->ehit, ISK_MAX_ELECTRODES * 8))
    return ISKRA_E_HIP;
  ISK_HIP(hipMemset(s->ehit, 0, ISK_MAX_ELECTRODES * 8));
  ISK_HIP(hipMemset(s->hdr, 0, HDR_WORDS * 8));
  ISK_HIP(hipMemset(s->cnt, 0, (size_t)nn * 8));
  ISK_HIP(hipMemset(s->gfall, 0, (size_t)nn * 8));
  return alloc_chunk_tables(c, s);
}

int isk_species_reset_chunks(iskra_ctx *c, int sid) {
  IskSpecies *s = &c->sp[sid];
  RC(alloc_chunk_tables(c, s));
  return isk_sort_species(c, sid, 0);
}

/* ---- profiling ---- */
static int64_t g_prof_ps0 = 0;

int isk_prof_begin(iskra_ctx *c, hipEvent_t *a) {
  *a = nullptr;
  if (!c->prof) return ISKRA_OK;
  if (c->prof_pool.empty()) {
    hipEvent_t e;
    ISK_HIP(hipEventCreate(&e));
    c->prof_pool.push_back(e);
  }
  *a = c->prof_pool.back();
  c->prof_pool.pop_back();
  ISK_HIP(hipEventRecord(*a, c->stream));
  return ISKRA_OK;
}

int isk_prof_end(iskra_ctx *c, hipEvent_t a, int kind, int64_t particles) {
  if (!c->prof || !a) return ISKRA_OK;
  if (c->prof_pool.empty()) {
    hipEvent_t e;
    ISK_HIP(hipEventCreate(&e));
    c->prof_pool.push_back(e);
  }
  hipEvent_t b = c->prof_pool.back();
  c->prof_pool.pop_back();
  ISK_HIP(hipEventRecord(b, c->stream));
  iskra_ctx::Pending pd;
  pd.a = a;
  pd.b = b;
  pd.kind = kind;
  pd.particles = particles;
  c->prof_pending.push_back(pd);
  return ISKRA_OK;
}

static int prof_resolve(iskra_ctx *c) {
  for (auto &pd : c->prof_pending) {
    ISK_HIP(hipEventSynchronize(pd.b));
    float ms = 0.f;
    ISK_HIP(hipEventElapsedTime(&ms, pd.a, pd.b));
    if (pd.kind >= 0 && pd.kind < ISKRA_K_COUNT) {
      c->stats.kernel_ms[pd.kind] += ms;
      c->stats.kernel_calls[pd.kind] += 1;
    }
    c->prof_pool.push_back(pd.a);
    c->prof_pool.push_back(pd.b);
  }
  c->prof_pending.clear();
  return ISKRA_OK;
}

static void prof_release(iskra_ctx *c) {
  for (auto &pd : c->prof_pending) {
    (void)hipEventDestroy(pd.a);
    (void)hipEventDestroy(pd.b);
  }
  c->prof_pending.clear();
  for (auto e : c->prof_pool) (void)hipEventDestroy(e);
  c->prof_pool.clear();
}

typedef float isk_f4 __attribute__((ext_vector_type(4)));
/* streaming copy: each work-group moves 4 x 256 x 16 B (16 KiB), loads first */
template <int NT>
__global__ __launch_bounds__(256) void k_copy(const isk_f4 *__restrict__ src,
                                              isk_f4 *__restrict__ dst, long long n) {
  long long base = (long long)blockIdx.x * 1024 + threadIdx.x;
  isk_f4 v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    long long i = base + u * 256;
    if (i < n) v[u] = NT ? __builtin_nontemporal_load(&src[i]) : src[i];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    long long i = base + u * 256;
    if (i < n) {
      if (NT) __builtin_nontemporal_store(v[u], &dst[i]);
      else dst[i] = v[u];
    }
  }
}

extern "C" {

const char *iskra_last_error(void) { return g_err; }
int iskra_version(void) { return 100; }

int iskra_ctx_create(int device, iskra_ctx **out) {
  g_err[0] = 0;
  if (!out) {
    isk_set_error("out is null");
    return ISKRA_E_ARG;
  }
  int n = 0;
  ISK_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    isk_set_error("device %d not available (%d devices)", device, n);
    return ISKRA_E_ARG;
  }
  ISK_HIP(hipSetDevice(device));
  /* once per device, with the device current: the address of its copy of
   * the no-allocation scratch array */
  RC(isk_noalloc_init(device));
  iskra_ctx *c = new (std::nothrow) iskra_ctx();
  if (!c) {
    isk_set_error("out of host memory");
    return ISKRA_E_ARG;
  }
  c->device = device;
  hipError_t e = hipStreamCreate(&c->stream);
  if (e != hipSuccess) {
    delete c;
    return isk_hip_fail(e, "hipStreamCreate", __FILE__, __LINE__);
  }
  memset(&c->stats, 0, sizeof(c->stats));
  c->t.T = 8;
  c->t.log2T = 3;
  c->t.CH = 2048;
  c->t.M = 2;
  c->t.W = c->t.T + 2 * c->t.M + 1;
  c->t.sort_every = 8;
  for (int k = 0; k <= ISK_MAX_TAGS; ++k) {
    c->ps.tag_val[k] = 0.0;
    c->ps.wf_on[k] = 0;
  }
  *out = c;
  return ISKRA_OK;
}

static void release_fields(iskra_ctx *c) {
  for (int s = 0; s < c->nspecies; ++s) isk_species_release(&c->sp[s]);
  c->nspecies = 0;
  for (int m = 0; m < c->nmcc; ++m) {
    IskMcc &q = c->mcc[m];
    void *ptrs[] = {q.d_eps, q.d_sig, q.d_off, q.target_n, q.nu, q.nu_acc, q.events, q.errflag,
                    q.prod_flag, q.stage, q.blk_cnt, q.cand_i, q.cand_cnt,
                    q.stage_pid, q.pid_c, q.snap};
    for (void *p : ptrs) isk_free(p);
    q.tab_off.clear();
    q.tab_eps.clear();
    q.tab_sig.clear();
  }
  c->nmcc = 0;
  for (int m = 0; m < c->ndsmc; ++m) isk_dsmc_release(&c->dsmc[m]);
  c->ndsmc = 0;
  isk_electrodes_release(c);
  isk_fluid_release(c);
  isk_poisson_free(c);
  c->ps.kind = -1;
  void *ptrs[] = {c->rho, c->phi, c->bg, c->vol, c->E, c->sort_H, c->sort_S, c->sort_tot,
                  c->sort_start, c->gather_out};
  for (void *p : ptrs) isk_free(p);
  c->rho = c->phi = c->bg = c->vol = nullptr;
  c->E = nullptr;
  c->sort_H = c->sort_S = c->sort_tot = c->sort_start = nullptr;
  c->sort_H_words = c->sort_S_words = 0;
  c->gather_out = nullptr;
  c->gather_cap = 0;
}

int iskra_ctx_destroy(iskra_ctx *c) {
  if (!c) return ISKRA_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  release_fields(c);
  prof_release(c);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return ISKRA_OK;
}

int iskra_sync(iskra_ctx *c) {
  CHECK_CTX(c);
  ISK_HIP(hipStreamSynchronize(c->stream));
  RC(isk_poisson_check(c, nullptr));
  for (int s = 0; s < c->nspecies; ++s) {
    long long of = 0;
    ISK_HIP(hipMemcpy(&of, c->sp[s].hdr + HDR_OVERFLOW, 8, hipMemcpyDeviceToHost));
    if (of) {
      isk_set_error("species %d: particle capacity exceeded (ionisation products dropped)", s);
      return ISKRA_E_CAPACITY;
    }
  }
  for (int m = 0; m < c->nmcc; ++m) {
    int err = 0;
    ISK_HIP(hipMemcpy(&err, c->mcc[m].errflag, 4, hipMemcpyDeviceToHost));
    if (err) {
      isk_set_error("mcc %d: collision probability above 1 (mcc.jl:273-279)", m);
      return ISKRA_E_PROB;
    }
  }
  return ISKRA_OK;
}

static int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

static int grid_setup(iskra_ctx *c, int dim, const int64_t nodes[2], const double dh[2],
                      const double origin[2], const int32_t volper[4], int axial) {
  if (dim != 1 && dim != 2) {
    isk_set_error("grid: dim must be 1 or 2");
    return ISKRA_E_ARG;
  }
  int64_t nx = nodes[0], ny = dim == 2 ? nodes[1] : 1;
  if (nx < 2 || (dim == 2 && ny < 2) || nx - 1 > 65535 || ny - 1 > 32767) {
    isk_set_error("grid: nodes (%lld, %lld) out of range", (long long)nx, (long long)ny);
    return ISKRA_E_ARG;
  }
  if (!(dh[0] > 0) || (dim == 2 && !(dh[1] > 0))) {
    isk_set_error("grid: spacing must be positive");
    return ISKRA_E_ARG;
  }
  ISK_HIP(hipStreamSynchronize(c->stream));
  release_fields(c);
  IskGrid &g = c->g;
  g.dim = dim;
  g.nx = (int)nx;
  g.ny = (int)ny;
  g.ncx = g.nx - 1;
  g.ncy = dim == 2 ? g.ny - 1 : 1;
  g.dx = dh[0];
  g.dy = dim == 2 ? dh[1] : 1.0;
  g.x0 = origin ? origin[0] : 0.0;
  g.y0 = (origin && dim == 2) ? origin[1] : 0.0;
  for (int k = 0; k < 4; ++k) g.volper[k] = volper ? volper[k] : 0;
  g.axial = axial;
  c->pusher = axial ? ISKRA_PUSHER_RZ : ISKRA_PUSHER_XY;
  /* default tiling */
  if (dim == 2) {
    c->t.T = 8;
    c->t.M = 2;
  } else {
    c->t.T = 64;
    c->t.M = 4;
  }
  c->t.log2T = ilog2(c->t.T);
  c->t.W = c->t.T + 2 * c->t.M + 1;
  c->ntx = (g.ncx + c->t.T - 1) / c->t.T;
  c->nty = dim == 2 ? (g.ncy + c->t.T - 1) / c->t.T : 1;
  c->ntiles = c->ntx * c->nty;
  if (c->ntiles > ISK_MAX_BINS) {
    c->have_grid = 0;
    isk_set_error("grid: %d tiles of %d cells exceed %d", c->ntiles, c->t.T, ISK_MAX_BINS);
    return ISKRA_E_UNSUPPORTED;
  }
  int nn = g.nx * g.ny;
  if (isk_alloc((void **)&c->rho, nn * 8) || isk_alloc((void **)&c->phi, nn * 8) ||
      isk_alloc((void **)&c->bg, nn * 8) || isk_alloc((void **)&c->vol, nn * 8) ||
      isk_alloc((void **)&c->E, (size_t)nn * 4 * dim))
    return ISKRA_E_HIP;
  ISK_HIP(hipMemset(c->rho, 0, nn * 8));
  ISK_HIP(hipMemset(c->phi, 0, nn * 8));
  ISK_HIP(hipMemset(c->bg, 0, nn * 8));
  ISK_HIP(hipMemset(c->E, 0, (size_t)nn * 4 * dim));
  /* cell_volume (RegularGrids.jl:26-38); 1-D by analogy: dx, halved at
   * non-periodic ends.  Axial grid (RegularGrids.jl:40-53): the annulus
   * pi dz ((i dr - dr/2)^2 - (i dr - 3dr/2)^2) of 1-based node i, a disc of
   * radius dr/2 on the axis, the outer half annulus from (nr-3/2) dr to
   * (nr-1) dr, each halved on a non-periodic z edge */
  std::vector<double> V(nn);
  double V0 = dim == 2 ? g.dx * g.dy : g.dx;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      double v = V0;
      if (axial) {
        const double dr = g.dx, dz = g.dy, pdz = M_PI * dz;
        const int i1 = i + 1, nr = g.nx;
        if (i == 0) {
          v = pdz * ((0.5 * dr) * (0.5 * dr));
        } else if (i == nr - 1) {
          const double a = nr * dr - 1.0 * dr, b = nr * dr - 1.5 * dr;
          v = pdz * (a * a - b * b);
        } else {
          const double a = i1 * dr - 0.5 * dr, b = i1 * dr - 1.5 * dr;
          v = pdz * (a * a - b * b);
        }
      } else {
        if (i == 0 && !g.volper[0]) v *= 0.5;
        if (i == g.nx - 1 && !g.volper[1]) v *= 0.5;
      }
      if (dim == 2) {
        if (j == 0 && !g.volper[2]) v *= 0.5;
        if (j == g.ny - 1 && !g.volper[3]) v *= 0.5;
      }
      V[j * g.nx + i] = v;
    }
  ISK_HIP(hipMemcpy(c->vol, V.data(), nn * 8, hipMemcpyHostToDevice));
  c->have_grid = 1;
  c->step_count = 0;
  c->t_now = 0.0;
  memset(&c->stats, 0, sizeof(c->stats));
  return ISKRA_OK;
}

int iskra_grid_set(iskra_ctx *c, int dim, const int64_t nodes[2], const double dh[2],
                   const double origin[2], const int32_t vol_periodic[4]) {
  CHECK_CTX(c);
  return grid_setup(c, dim, nodes, dh, origin, vol_periodic, 0);
}

int iskra_grid_set_axial(iskra_ctx *c, const int64_t nodes[2], const double dh[2],
                         const double origin[2], const int32_t vol_periodic_z[2]) {
  CHECK_CTX(c);
  if (!nodes || !dh) {
    isk_set_error("grid_set_axial: nodes and dh are required");
    return ISKRA_E_ARG;
  }
  /* the radial bcs are (:other, :other): no radial halving flags */
  int32_t vp[4] = {0, 0, vol_periodic_z ? vol_periodic_z[0] : 0,
                   vol_periodic_z ? vol_periodic_z[1] : 0};
  return grid_setup(c, 2, nodes, dh, origin, vp, 1);
}

int iskra_pusher_set(iskra_ctx *c, int kind) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (kind != ISKRA_PUSHER_XY && kind != ISKRA_PUSHER_RZ) {
    isk_set_error("pusher_set: unknown kind %d", kind);
    return ISKRA_E_ARG;
  }
  if (kind == ISKRA_PUSHER_RZ && c->g.dim != 2) {
    isk_set_error("pusher_set: the RZ pusher needs a 2-D grid");
    return ISKRA_E_ARG;
  }
  c->pusher = kind;
  return ISKRA_OK;
}

int iskra_tuning_get(iskra_ctx *c, iskra_tuning *t) {
  CHECK_CTX(c);
  if (!t) return ISKRA_E_ARG;
  t->tile = c->t.T;
  t->chunk = c->t.CH;
  t->margin = c->t.M;
  t->sort_every = c->t.sort_every;
  return ISKRA_OK;
}

int iskra_tuning_set(iskra_ctx *c, const iskra_tuning *t) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (!t) return ISKRA_E_ARG;
  IskTune n = c->t;
  if (t->tile > 0) n.T = t->tile;
  if (t->chunk > 0) n.CH = t->chunk;
  if (t->margin > 0) n.M = t->margin;
  if (t->sort_every >= 0) n.sort_every = t->sort_every;
  if ((n.T & (n.T - 1)) || n.T < 2 || n.T > 1024 || n.CH < 256 || n.CH > (1 << 20) ||
      n.M < 1 || n.M > n.T) {
    isk_set_error("tuning: tile must be a power of two in [2,1024], chunk in [256,2^20], margin in [1,tile]");
    return ISKRA_E_ARG;
  }
  n.log2T = ilog2(n.T);
  n.W = n.T + 2 * n.M + 1;
  if (c->g.dim == 2 && n.W > 48) {
    isk_set_error("tuning: 2-D window %d exceeds 48 nodes", n.W);
    return ISKRA_E_ARG;
  }
  int ntx = (c->g.ncx + n.T - 1) / n.T;
  int nty = c->g.dim == 2 ? (c->g.ncy + n.T - 1) / n.T : 1;
  if ((int64_t)ntx * nty > ISK_MAX_BINS || (c->g.dim == 2 ? n.T * n.T : n.T) > ISK_MAX_BINS) {
    isk_set_error("tuning: %d tiles (or %d cells per tile) exceed %d", ntx * nty,
                  c->g.dim == 2 ? n.T * n.T : n.T, ISK_MAX_BINS);
    return ISKRA_E_ARG;
  }
  ISK_HIP(hipStreamSynchronize(c->stream));
  bool relayout = n.T != c->t.T || n.CH != c->t.CH || n.M != c->t.M;
  c->t = n;
  if (relayout) {
    c->ntx = ntx;
    c->nty = nty;
    c->ntiles = ntx * nty;
    for (int s = 0; s < c->nspecies; ++s) RC(isk_species_reset_chunks(c, s));
  }
  return ISKRA_OK;
}

int iskra_species_add(iskra_ctx *c, double q, double m, double w0, int64_t capacity,
                      int track_ids, int *sid) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (!sid || capacity < 0 || capacity > 2147483000LL || !(m > 0)) {
    isk_set_error("species_add: bad arguments (capacity %lld, m %g)", (long long)capacity, m);
    return ISKRA_E_ARG;
  }
  if (c->nspecies >= ISK_MAX_SPECIES) {
    isk_set_error("species_add: at most %d species", ISK_MAX_SPECIES);
    return ISKRA_E_ARG;
  }
  IskSpecies *s = &c->sp[c->nspecies];
  memset(s, 0, sizeof(*s));
  s->q = q;
  s->m = m;
  s->w0 = w0;
  s->cap = capacity;
  s->track_ids = track_ids ? 1 : 0;
  s->bnd[0] = s->bnd[1] = ISKRA_BND_WRAP;
  s->margin = -1;
  s->cur = 0;
  int rc = isk_species_alloc(c, s);
  if (rc) {
    isk_species_release(s);
    return rc;
  }
  *sid = c->nspecies++;
  RC(isk_sort_species(c, *sid, 0));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

int iskra_species_set_boundary(iskra_ctx *c, int sid, const int32_t mode[2]) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  for (int d = 0; d < 2; ++d) {
    if (mode[d] < 0 || mode[d] > 2) {
      isk_set_error("boundary mode %d invalid", mode[d]);
      return ISKRA_E_ARG;
    }
    c->sp[sid].bnd[d] = mode[d];
  }
  return ISKRA_OK;
}

int iskra_species_margin(iskra_ctx *c, int sid, int32_t margin) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  if (margin < -1) {
    isk_set_error("species_margin: %d < -1", margin);
    return ISKRA_E_ARG;
  }
  c->sp[sid].margin = margin;
  return ISKRA_OK;
}

int iskra_species_sort_every(iskra_ctx *c, int sid, int32_t every) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  if (every < 0) {
    isk_set_error("species_sort_every: %d < 0", every);
    return ISKRA_E_ARG;
  }
  c->sp[sid].sort_every = every;
  return ISKRA_OK;
}

int iskra_species_upload(iskra_ctx *c, int sid, int64_t np, const double *x, const double *v,
                         const uint32_t *id) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  if (np > 0 && (!x || !v)) {
    isk_set_error("upload: null x or v");
    return ISKRA_E_ARG;
  }
  RC(isk_convert_upload(c, sid, np, x, v, id));
  c->sp[sid].deposited = 0;
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

int iskra_species_download(iskra_ctx *c, int sid, int64_t cap, int64_t *np, double *x,
                           double *v, uint32_t *id) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  IskSpecies *s = &c->sp[sid];
  ISK_HIP(hipStreamSynchronize(c->stream));
  long long hdr[HDR_WORDS];
  ISK_HIP(hipMemcpy(hdr, s->hdr, sizeof(hdr), hipMemcpyDeviceToHost));
  int64_t n = hdr[HDR_NP];
  int b = s->cur;
  std::vector<int32_t> cell(n);
  std::vector<float> ox(n), oy(n, 0.f), vx(n), vy(n), vz(n);
  std::vector<uint32_t> ids(n, 0);
  if (n > 0) {
    ISK_HIP(hipMemcpy(cell.data(), s->cell[b], n * 4, hipMemcpyDeviceToHost));
    ISK_HIP(hipMemcpy(ox.data(), s->ox[b], n * 4, hipMemcpyDeviceToHost));
    if (c->g.dim == 2) ISK_HIP(hipMemcpy(oy.data(), s->oy[b], n * 4, hipMemcpyDeviceToHost));
    ISK_HIP(hipMemcpy(vx.data(), s->vx[b], n * 4, hipMemcpyDeviceToHost));
    ISK_HIP(hipMemcpy(vy.data(), s->vy[b], n * 4, hipMemcpyDeviceToHost));
    ISK_HIP(hipMemcpy(vz.data(), s->vz[b], n * 4, hipMemcpyDeviceToHost));
    if (s->track_ids) ISK_HIP(hipMemcpy(ids.data(), s->id[b], n * 4, hipMemcpyDeviceToHost));
  }
  int64_t live = 0;
  for (int64_t p = 0; p < n; ++p) live += cell[p] >= 0;
  if (np) *np = live;
  if (live > cap) {
    isk_set_error("download: %lld particles do not fit cap %lld", (long long)live, (long long)cap);
    return ISKRA_E_CAPACITY;
  }
  const IskGrid &g = c->g;
  int64_t k = 0;
  for (int64_t p = 0; p < n; ++p) {
    if (cell[p] < 0) continue;
    int i = cell[p] & 0xFFFF, j = cell[p] >> 16;
    if (x) {
      x[k] = g.x0 + ((double)i + (double)ox[p]) * g.dx;
      if (g.dim == 2) x[cap + k] = g.y0 + ((double)j + (double)oy[p]) * g.dy;
    }
    if (v) {
      v[k] = vx[p];
      v[cap + k] = vy[p];
      v[2 * cap + k] = vz[p];
    }
    if (id) id[k] = ids[p];
    ++k;
  }
  return ISKRA_OK;
}

int iskra_species_get_state(iskra_ctx *c, int sid, int64_t cap, int64_t *np, int32_t *cell,
                            float *ox, float *oy, float *vx, float *vy, float *vz,
                            uint32_t *id) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  IskSpecies *s = &c->sp[sid];
  ISK_HIP(hipStreamSynchronize(c->stream));
  long long hdr[HDR_WORDS];
  ISK_HIP(hipMemcpy(hdr, s->hdr, sizeof(hdr), hipMemcpyDeviceToHost));
  int64_t n = hdr[HDR_NP];
  if (np) *np = n;
  if (!cell && !ox && !oy && !vx && !vy && !vz && !id) return ISKRA_OK; /* size query */
  if (n > cap) {
    isk_set_error("get_state: %lld particles do not fit cap %lld", (long long)n, (long long)cap);
    return ISKRA_E_CAPACITY;
  }
  if (n == 0) return ISKRA_OK;
  int b = s->cur;
  if (cell) ISK_HIP(hipMemcpy(cell, s->cell[b], n * 4, hipMemcpyDeviceToHost));
  if (ox) ISK_HIP(hipMemcpy(ox, s->ox[b], n * 4, hipMemcpyDeviceToHost));
  if (oy && c->g.dim == 2) ISK_HIP(hipMemcpy(oy, s->oy[b], n * 4, hipMemcpyDeviceToHost));
  if (vx) ISK_HIP(hipMemcpy(vx, s->vx[b], n * 4, hipMemcpyDeviceToHost));
  if (vy) ISK_HIP(hipMemcpy(vy, s->vy[b], n * 4, hipMemcpyDeviceToHost));
  if (vz) ISK_HIP(hipMemcpy(vz, s->vz[b], n * 4, hipMemcpyDeviceToHost));
  if (id && s->track_ids) ISK_HIP(hipMemcpy(id, s->id[b], n * 4, hipMemcpyDeviceToHost));
  return ISKRA_OK;
}

int iskra_species_set_state(iskra_ctx *c, int sid, int64_t np, const int32_t *cell,
                            const float *ox, const float *oy, const float *vx, const float *vy,
                            const float *vz, const uint32_t *id) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  IskSpecies *s = &c->sp[sid];
  if (np < 0 || np > s->cap) {
    isk_set_error("set_state: np=%lld outside [0, capacity=%lld]", (long long)np,
                  (long long)s->cap);
    return ISKRA_E_CAPACITY;
  }
  if (np > 0 && (!cell || !ox || !vx || !vy || !vz || (c->g.dim == 2 && !oy))) {
    isk_set_error("set_state: null array");
    return ISKRA_E_ARG;
  }
  const IskGrid &g = c->g;
  uint32_t maxid = 0;
  for (int64_t p = 0; p < np; ++p) {
    int32_t cc = cell[p];
    if (cc < 0) continue;
    int i = cc & 0xFFFF, j = cc >> 16;
    bool bad = i >= g.ncx || j >= g.ncy || !(ox[p] >= 0.0f && ox[p] < 1.0f) ||
               (g.dim == 2 && !(oy[p] >= 0.0f && oy[p] < 1.0f));
    if (bad) {
      isk_set_error("set_state: particle %lld has cell (%d,%d) or offset outside the grid",
                    (long long)p, i, j);
      return ISKRA_E_ARG;
    }
    if (id && id[p] > maxid) maxid = id[p];
  }
  ISK_HIP(hipStreamSynchronize(c->stream));
  int b = s->cur;
  if (np > 0) {
    ISK_HIP(hipMemcpy(s->cell[b], cell, np * 4, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(s->ox[b], ox, np * 4, hipMemcpyHostToDevice));
    if (g.dim == 2) ISK_HIP(hipMemcpy(s->oy[b], oy, np * 4, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(s->vx[b], vx, np * 4, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(s->vy[b], vy, np * 4, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(s->vz[b], vz, np * 4, hipMemcpyHostToDevice));
    if (s->track_ids) {
      if (id) {
        ISK_HIP(hipMemcpy(s->id[b], id, np * 4, hipMemcpyHostToDevice));
      } else {
        std::vector<uint32_t> seq(np);
        for (int64_t p = 0; p < np; ++p) seq[p] = (uint32_t)p;
        ISK_HIP(hipMemcpy(s->id[b], seq.data(), np * 4, hipMemcpyHostToDevice));
      }
    }
  }
  long long hdr[HDR_WORDS];
  ISK_HIP(hipMemcpy(hdr, s->hdr, sizeof(hdr), hipMemcpyDeviceToHost));
  hdr[HDR_NP] = np;
  hdr[HDR_NP_SORTED] = 0;
  hdr[HDR_NCHUNKS] = 0;
  hdr[HDR_NEXT_ID] = (s->track_ids && id) ? (long long)maxid + 1 : np;
  hdr[HDR_OVERFLOW] = 0;
  ISK_HIP(hipMemcpy(s->hdr, hdr, sizeof(hdr), hipMemcpyHostToDevice));
  s->deposited = 0;
  RC(isk_sort_species(c, sid, 0));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

__global__ void k_count_live(const int32_t *cell, const long long *hdr, unsigned long long *out) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int live = p < hdr[HDR_NP] && cell[p] >= 0;
  unsigned long long m = __ballot(live);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(out, (unsigned long long)__popcll(m));
}

int iskra_species_count(iskra_ctx *c, int sid, int64_t *np) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  if (!np) return ISKRA_E_ARG;
  IskSpecies *s = &c->sp[sid];
  unsigned long long *d = nullptr;
  RC(isk_alloc((void **)&d, 8));
  ISK_HIP(hipMemsetAsync(d, 0, 8, c->stream));
  int blocks = (int)((std::max<int64_t>(s->cap, 1) + 255) / 256);
  hipLaunchKernelGGL(k_count_live, dim3(blocks), dim3(256), 0, c->stream, s->cell[s->cur],
                     s->hdr, d);
  ISK_CHECK_LAUNCH();
  unsigned long long h = 0;
  ISK_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, c->stream));
  ISK_HIP(hipStreamSynchronize(c->stream));
  isk_free(d);
  *np = (int64_t)h;
  return ISKRA_OK;
}

int iskra_species_sample_maxwellian(iskra_ctx *c, int sid, int64_t n, const double *wx,
                                    const double *dx, const double *wv, const double *dv,
                                    uint64_t seed, uint32_t stream) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  if (!wx || !wv || n < 0) {
    isk_set_error("sample: null wx/wv or negative n");
    return ISKRA_E_ARG;
  }
  CHECK_GRID(c);
  /* positions are a cell and an offset inside it, so the box [dx, dx + wx)
   * must lie in the grid: a particle outside it has no representation (the
   * reference would wrap or discard it only after the next push,
   * pic/surfaces/wrap.jl:1-33) */
  for (int d = 0; d < c->g.dim; ++d) {
    const double lo = dx ? dx[d] : 0.0, hi = lo + wx[d];
    const double g0 = d == 0 ? c->g.x0 : c->g.y0;
    const double g1 = g0 + (d == 0 ? c->g.ncx * c->g.dx : c->g.ncy * c->g.dy);
    /* slack for a box given as the domain length, which can differ from
     * cells * spacing by round-off (the kernel keeps such positions in the
     * edge cells) */
    const double tol = 1e-12 * std::max(fabs(g0), fabs(g1));
    if (!(wx[d] >= 0.0) || !(lo >= g0 - tol) || !(hi <= g1 + tol)) {
      isk_set_error("sample: source box [%g, %g) of dimension %d leaves the grid [%g, %g]", lo,
                    hi, d + 1, g0, g1);
      return ISKRA_E_ARG;
    }
  }
  RC(isk_sample_maxwellian(c, sid, n, wx, dx, wv, dv, seed, stream));
  c->sp[sid].deposited = 0;
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

/* ---- step operators ---- */
int iskra_sort(iskra_ctx *c, int sid, int exact_cell) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  RC(isk_sort_species(c, sid, exact_cell));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

int iskra_gather(iskra_ctx *c, int sid, int64_t cap, double *E_out) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  IskSpecies *s = &c->sp[sid];
  RC(isk_gather_species(c, sid));
  ISK_HIP(hipStreamSynchronize(c->stream));
  long long hdr[HDR_WORDS];
  ISK_HIP(hipMemcpy(hdr, s->hdr, sizeof(hdr), hipMemcpyDeviceToHost));
  int64_t n = hdr[HDR_NP];
  std::vector<float> ex(n), ey(n);
  std::vector<int32_t> cell(n);
  if (n > 0) {
    ISK_HIP(hipMemcpy(ex.data(), c->gather_out, n * 4, hipMemcpyDeviceToHost));
    ISK_HIP(hipMemcpy(ey.data(), c->gather_out + s->cap, n * 4, hipMemcpyDeviceToHost));
    ISK_HIP(hipMemcpy(cell.data(), s->cell[s->cur], n * 4, hipMemcpyDeviceToHost));
  }
  int64_t k = 0;
  for (int64_t p = 0; p < n; ++p) {
    if (cell[p] < 0) continue;
    if (k >= cap) {
      isk_set_error("gather: cap %lld too small", (long long)cap);
      return ISKRA_E_CAPACITY;
    }
    E_out[k] = ex[p];
    E_out[cap + k] = ey[p];
    E_out[2 * cap + k] = 0.0;
    ++k;
  }
  return ISKRA_OK;
}

int iskra_push(iskra_ctx *c, int sid, double dt) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  RC(isk_push_species(c, sid, dt, 1));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

int iskra_deposit_species(iskra_ctx *c, int sid) {
  CHECK_CTX(c);
  if (!isk_species_ok(c, sid)) return ISKRA_E_ARG;
  RC(isk_push_species(c, sid, 0.0, 0));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

int iskra_deposit(iskra_ctx *c) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  RC(isk_rho(c));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

/* ---- fields ---- */
int iskra_field_download(iskra_ctx *c, int which, int index, double *out) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (!out) return ISKRA_E_ARG;
  ISK_HIP(hipStreamSynchronize(c->stream));
  int nn = isk_nodes(c);
  int dim = c->g.dim;
  switch (which) {
    case ISKRA_F_RHO: ISK_HIP(hipMemcpy(out, c->rho, nn * 8, hipMemcpyDeviceToHost)); break;
    case ISKRA_F_PHI: ISK_HIP(hipMemcpy(out, c->phi, nn * 8, hipMemcpyDeviceToHost)); break;
    case ISKRA_F_BG:
      ISK_HIP(hipMemcpy(out, c->nfluid ? c->bg_fixed : c->bg, nn * 8, hipMemcpyDeviceToHost));
      break;
    case ISKRA_F_FLUID_N:
    case ISKRA_F_FLUID_DN:
      if (index < 0 || index >= c->nfluid) {
        isk_set_error("field_download: fluid %d of %d", index, c->nfluid);
        return ISKRA_E_ARG;
      }
      ISK_HIP(hipMemcpy(out, which == ISKRA_F_FLUID_N ? c->fl[index].n : c->fl[index].dn,
                        nn * 8, hipMemcpyDeviceToHost));
      break;
    case ISKRA_F_EX:
    case ISKRA_F_EY: {
      if (which == ISKRA_F_EY && dim == 1) {
        for (int k = 0; k < nn; ++k) out[k] = 0.0;
        break;
      }
      std::vector<float> e((size_t)nn * dim);
      ISK_HIP(hipMemcpy(e.data(), c->E, (size_t)nn * dim * 4, hipMemcpyDeviceToHost));
      int comp = which == ISKRA_F_EX ? 0 : 1;
      for (int k = 0; k < nn; ++k) out[k] = e[(size_t)k * dim + comp];
      break;
    }
    case ISKRA_F_DENSITY: {
      if (!isk_species_ok(c, index)) return ISKRA_E_ARG;
      IskSpecies *s = &c->sp[index];
      if (!s->deposited) {
        RC(isk_push_species(c, index, 0.0, 0));
        ISK_HIP(hipStreamSynchronize(c->stream));
      }
      std::vector<long long> cnt(nn);
      std::vector<double> V(nn);
      ISK_HIP(hipMemcpy(cnt.data(), s->cnt, nn * 8, hipMemcpyDeviceToHost));
      ISK_HIP(hipMemcpy(V.data(), c->vol, nn * 8, hipMemcpyDeviceToHost));
      for (int k = 0; k < nn; ++k)
        out[k] = ((double)cnt[k] * (1.0 / ISK_FIX_ONE)) * s->w0 / V[k];
      break;
    }
    case ISKRA_F_TARGET: {
      if (index < 0 || index >= c->nmcc) return ISKRA_E_ARG;
      ISK_HIP(hipMemcpy(out, c->mcc[index].target_n, nn * 8, hipMemcpyDeviceToHost));
      break;
    }
    default:
      isk_set_error("field_download: unknown field %d", which);
      return ISKRA_E_ARG;
  }
  return ISKRA_OK;
}

int iskra_field_upload(iskra_ctx *c, int which, int index, const double *in) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (!in) return ISKRA_E_ARG;
  ISK_HIP(hipStreamSynchronize(c->stream));
  int nn = isk_nodes(c);
  int dim = c->g.dim;
  switch (which) {
    case ISKRA_F_RHO: ISK_HIP(hipMemcpy(c->rho, in, nn * 8, hipMemcpyHostToDevice)); break;
    case ISKRA_F_PHI: ISK_HIP(hipMemcpy(c->phi, in, nn * 8, hipMemcpyHostToDevice)); break;
    case ISKRA_F_BG:
      ISK_HIP(hipMemcpy(c->nfluid ? c->bg_fixed : c->bg, in, nn * 8, hipMemcpyHostToDevice));
      RC(isk_fluid_bg(c));
      ISK_HIP(hipStreamSynchronize(c->stream));
      break;
    case ISKRA_F_FLUID_N:
      if (index < 0 || index >= c->nfluid) {
        isk_set_error("field_upload: fluid %d of %d", index, c->nfluid);
        return ISKRA_E_ARG;
      }
      ISK_HIP(hipMemcpy(c->fl[index].n, in, nn * 8, hipMemcpyHostToDevice));
      RC(isk_fluid_bg(c));
      ISK_HIP(hipStreamSynchronize(c->stream));
      break;
    case ISKRA_F_EX:
    case ISKRA_F_EY: {
      if (which == ISKRA_F_EY && dim == 1) {
        isk_set_error("field_upload: no E_y in 1-D");
        return ISKRA_E_ARG;
      }
      std::vector<float> e((size_t)nn * dim);
      ISK_HIP(hipMemcpy(e.data(), c->E, (size_t)nn * dim * 4, hipMemcpyDeviceToHost));
      int comp = which == ISKRA_F_EX ? 0 : 1;
      for (int k = 0; k < nn; ++k) e[(size_t)k * dim + comp] = (float)in[k];
      ISK_HIP(hipMemcpy(c->E, e.data(), (size_t)nn * dim * 4, hipMemcpyHostToDevice));
      break;
    }
    case ISKRA_F_TARGET: {
      if (index < 0 || index >= c->nmcc) {
        isk_set_error("field_upload: mcc index %d out of range", index);
        return ISKRA_E_ARG;
      }
      IskMcc &m = c->mcc[index];
      ISK_HIP(hipMemcpy(m.target_n, in, nn * 8, hipMemcpyHostToDevice));
      double mx = -INFINITY;
      for (int k = 0; k < nn; ++k) mx = std::max(mx, in[k]);
      m.max_n = mx;
      break;
    }
    default:
      isk_set_error("field_upload: field %d cannot be uploaded", which);
      return ISKRA_E_ARG;
  }
  return ISKRA_OK;
}

/* ---- Poisson ---- */
int iskra_poisson_set(iskra_ctx *c, int kind, double eps0, const int8_t *tags, int periodic_x,
                      int periodic_y, double rtol, int max_iter) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (kind < ISKRA_POISSON_DIRECT || kind > ISKRA_POISSON_SPECTRAL || !(eps0 > 0)) {
    isk_set_error("poisson_set: bad kind %d or eps0 %g", kind, eps0);
    return ISKRA_E_ARG;
  }
  ISK_HIP(hipStreamSynchronize(c->stream));
  isk_poisson_free(c);
  IskPoisson &p = c->ps;
  int nn = isk_nodes(c);
  /* a new solver drops the electrodes built on the old one (their rows and
   * tags); the tracker keeps its surfaces */
  c->nelec = 0;
  c->nfloat = 0;
  c->circuit_on = 0;
  p.h_srow.clear();
  p.h_sd.clear();
  p.h_sc.clear();
  p.h_ssig.clear();
  p.h_snb.clear();
  p.h_sdof.clear();
  p.kind = kind;
  p.eps0 = eps0;
  p.rtol = rtol > 0 ? rtol : 1e-8;
  p.max_iter = max_iter > 0 ? max_iter : 20000;
  p.per_x = periodic_x ? 1 : 0;
  p.per_y = (c->g.dim == 2 && periodic_y) ? 1 : 0;
  p.tags.assign(nn, 0);
  if (tags) {
    for (int k = 0; k < nn; ++k) {
      if (tags[k] < 0 || tags[k] > ISK_MAX_TAGS) {
        isk_set_error("poisson_set: tag %d at node %d outside [0,%d]", tags[k], k, ISK_MAX_TAGS);
        return ISKRA_E_ARG;
      }
      p.tags[k] = tags[k];
    }
  }
  return isk_poisson_setup(c);
}

int iskra_poisson_set_tag_value(iskra_ctx *c, int tag, double value) {
  CHECK_CTX(c);
  if (tag < 1 || tag > ISK_MAX_TAGS) {
    isk_set_error("tag %d out of range", tag);
    return ISKRA_E_ARG;
  }
  c->ps.tag_val[tag] = value;
  c->ps.wf_on[tag] = 0;
  if (c->ps.d_tag_val) {
    ISK_HIP(hipStreamSynchronize(c->stream));
    ISK_HIP(hipMemcpy(c->ps.d_tag_val + tag, &value, 8, hipMemcpyHostToDevice));
  }
  return ISKRA_OK;
}

int iskra_poisson_set_tag_waveform(iskra_ctx *c, int tag, double offset, double amplitude,
                                   double freq, double phase) {
  CHECK_CTX(c);
  if (tag < 1 || tag > ISK_MAX_TAGS) {
    isk_set_error("tag %d out of range", tag);
    return ISKRA_E_ARG;
  }
  c->ps.wf_on[tag] = 1;
  c->ps.wf_off[tag] = offset;
  c->ps.wf_amp[tag] = amplitude;
  c->ps.wf_f[tag] = freq;
  c->ps.wf_ph[tag] = phase / 6.283185307179586;
  return ISKRA_OK;
}

int iskra_poisson_solve(iskra_ctx *c, int warm, int *iters, double *relres) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (c->ps.kind < 0) {
    isk_set_error("poisson_solve: no solver (iskra_poisson_set)");
    return ISKRA_E_STATE;
  }
  RC(isk_poisson_solve(c, warm, iters, relres));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

int iskra_efield(iskra_ctx *c) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  RC(isk_efield(c));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

/* ---- MCC ---- */
int iskra_mcc_add(iskra_ctx *c, int src_sid, int n_coll, const int32_t *types,
                  const double *energy, const int32_t *tab_off, const double *tab_eps,
                  const double *tab_sigma, int ion_sid, double target_m, double target_q,
                  double target_T, int *mcc_id) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (!isk_species_ok(c, src_sid)) return ISKRA_E_ARG;
  if (!c->sp[src_sid].track_ids) {
    isk_set_error("mcc_add: the source species must track ids (its draws are keyed by id)");
    return ISKRA_E_ARG;
  }
  if (n_coll < 1 || n_coll > ISK_MAX_COLL || !types || !tab_off || !tab_eps || !tab_sigma ||
      !mcc_id || !(target_m > 0)) {
    isk_set_error("mcc_add: bad arguments");
    return ISKRA_E_ARG;
  }
  if (c->nmcc >= ISK_MAX_MCC) {
    isk_set_error("mcc_add: at most %d interactions", ISK_MAX_MCC);
    return ISKRA_E_ARG;
  }
  if (ion_sid >= 0 && !isk_species_ok(c, ion_sid)) return ISKRA_E_ARG;
  IskMcc &m = c->mcc[c->nmcc];
  m = IskMcc();
  m.src = src_sid;
  m.ion = ion_sid;
  m.ncoll = n_coll;
  for (int k = 0; k < n_coll; ++k) {
    if (types[k] < 0 || types[k] > ISKRA_COLL_INELASTIC_BACKWARD) {
      isk_set_error("mcc_add: bad collision type %d", types[k]);
      return ISKRA_E_ARG;
    }
    m.types[k] = types[k];
    m.energy[k] = energy ? energy[k] : 0.0;
    if (tab_off[k + 1] - tab_off[k] < 1) {
      isk_set_error("mcc_add: empty cross-section table %d", k);
      return ISKRA_E_ARG;
    }
  }
  m.tab_off.assign(tab_off, tab_off + n_coll + 1);
  int nt = tab_off[n_coll] - tab_off[0];
  if (tab_off[0] != 0) {
    isk_set_error("mcc_add: tab_off[0] must be 0");
    return ISKRA_E_ARG;
  }
  m.tab_eps.assign(tab_eps, tab_eps + nt);
  m.tab_sig.assign(tab_sigma, tab_sigma + nt);
  for (int k = 0; k < n_coll; ++k)
    for (int q = tab_off[k] + 1; q < tab_off[k + 1]; ++q)
      if (!(m.tab_eps[q] > m.tab_eps[q - 1])) {
        isk_set_error("mcc_add: energies of table %d are not increasing", k);
        return ISKRA_E_ARG;
      }
  m.target_m = target_m;
  m.target_q = target_q;
  m.target_T = target_T;
  int rc = isk_mcc_setup(c, c->nmcc);
  if (rc) return rc;
  *mcc_id = c->nmcc++;
  return ISKRA_OK;
}

int iskra_mcc_run(iskra_ctx *c, int mcc_id, double dt, int64_t *events) {
  CHECK_CTX(c);
  if (mcc_id < 0 || mcc_id >= c->nmcc) {
    isk_set_error("mcc id %d out of range", mcc_id);
    return ISKRA_E_ARG;
  }
  RC(isk_mcc_run(c, mcc_id, dt, events));
  return iskra_sync(c);
}

/* mode 0: running sum + live counts; 1: the same, then both cleared;
 * 2: the live counts alone (since the last fold) */
static int mcc_counts_mode(iskra_ctx *c, int mcc_id, int mode, int64_t *out) {
  CHECK_CTX(c);
  if (mcc_id < 0 || mcc_id >= c->nmcc) {
    isk_set_error("mcc id %d out of range", mcc_id);
    return ISKRA_E_ARG;
  }
  IskMcc &m = c->mcc[mcc_id];
  int nn = isk_nodes(c);
  ISK_HIP(hipStreamSynchronize(c->stream));
  std::vector<unsigned long long> nu((size_t)nn * m.ncoll), acc(nu.size(), 0ull);
  ISK_HIP(hipMemcpy(nu.data(), m.nu, nu.size() * 8, hipMemcpyDeviceToHost));
  if (mode != 2) ISK_HIP(hipMemcpy(acc.data(), m.nu_acc, acc.size() * 8, hipMemcpyDeviceToHost));
  if (out)
    for (int n = 0; n < nn; ++n)
      for (int k = 0; k < m.ncoll; ++k) {
        const size_t q = (size_t)n * m.ncoll + k;
        out[(size_t)k * nn + n] = (int64_t)(nu[q] + acc[q]);
      }
  if (mode == 1) {
    ISK_HIP(hipMemset(m.nu, 0, nu.size() * 8));
    ISK_HIP(hipMemset(m.nu_acc, 0, nu.size() * 8));
  }
  return ISKRA_OK;
}

int iskra_mcc_counts(iskra_ctx *c, int mcc_id, int reset, int64_t *out) {
  return mcc_counts_mode(c, mcc_id, reset ? 1 : 0, out);
}

int iskra_mcc_counts_last(iskra_ctx *c, int mcc_id, int64_t *out) {
  if (!out) return ISKRA_E_ARG;
  return mcc_counts_mode(c, mcc_id, 2, out);
}

/* move every interaction's event counts into its running sum */
__global__ void k_counts_fold(unsigned long long *acc, unsigned long long *nu, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    acc[k] += nu[k];
    nu[k] = 0ull;
  }
}

static int isk_counts_fold(iskra_ctx *c) {
  const int nn = isk_nodes(c);
  for (int m = 0; m < c->nmcc; ++m) {
    const int n = nn * c->mcc[m].ncoll;
    hipLaunchKernelGGL(k_counts_fold, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                       c->mcc[m].nu_acc, c->mcc[m].nu, n);
    ISK_CHECK_LAUNCH();
  }
  for (int m = 0; m < c->ndsmc; ++m) {
    hipLaunchKernelGGL(k_counts_fold, dim3((nn + 255) / 256), dim3(256), 0, c->stream,
                       c->dsmc[m].nu_acc, c->dsmc[m].nu, nn);
    ISK_CHECK_LAUNCH();
  }
  return ISKRA_OK;
}

/* ---- DSMC ---- */
int iskra_dsmc_add(iskra_ctx *c, int src_sid, int tgt_sid, int n_tab, const double *tab_g,
                   const double *tab_sigma, int *dsmc_id) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (!isk_species_ok(c, src_sid) || !isk_species_ok(c, tgt_sid)) return ISKRA_E_ARG;
  if (!dsmc_id) {
    isk_set_error("dsmc_add: dsmc_id is required");
    return ISKRA_E_ARG;
  }
  RC(isk_dsmc_add(c, src_sid, tgt_sid, n_tab, tab_g, tab_sigma, dsmc_id));
  return ISKRA_OK;
}

int iskra_dsmc_run(iskra_ctx *c, int dsmc_id, double dt, int64_t *events) {
  CHECK_CTX(c);
  if (dsmc_id < 0 || dsmc_id >= c->ndsmc) {
    isk_set_error("dsmc id %d out of range", dsmc_id);
    return ISKRA_E_ARG;
  }
  if (!(dt > 0)) {
    isk_set_error("dsmc_run: dt must be positive");
    return ISKRA_E_ARG;
  }
  RC(isk_dsmc_run(c, dsmc_id, dt, events));
  return iskra_sync(c);
}

int iskra_dsmc_counts(iskra_ctx *c, int dsmc_id, int reset, int64_t *out) {
  CHECK_CTX(c);
  if (dsmc_id < 0 || dsmc_id >= c->ndsmc) {
    isk_set_error("dsmc id %d out of range", dsmc_id);
    return ISKRA_E_ARG;
  }
  return isk_dsmc_counts(c, dsmc_id, reset ? 1 : 0, out);
}

int iskra_dsmc_counts_last(iskra_ctx *c, int dsmc_id, int64_t *out) {
  CHECK_CTX(c);
  if (dsmc_id < 0 || dsmc_id >= c->ndsmc || !out) {
    isk_set_error("dsmc id %d out of range or NULL out", dsmc_id);
    return ISKRA_E_ARG;
  }
  return isk_dsmc_counts(c, dsmc_id, 2, out);
}

/* ---- RNG ---- */
int iskra_rng_set(iskra_ctx *c, uint64_t seed) {
  CHECK_CTX(c);
  c->seed = seed;
  return ISKRA_OK;
}

int iskra_rng_export(iskra_ctx *c, uint32_t stream, uint32_t step, const uint32_t *ids,
                     int64_t n, int nblocks, uint32_t *out) {
  CHECK_CTX(c);
  if ((n > 0 && (!ids || !out)) || nblocks < 0) return ISKRA_E_ARG;
  return isk_rng_export(c, stream, step, ids, n, nblocks, out);
}

int iskra_rng_export_noalloc(iskra_ctx *c, uint32_t stream, uint32_t step, const uint32_t *ids,
                             int64_t n, int nblocks, uint32_t *out) {
  CHECK_CTX(c);
  if (n < 0 || nblocks < 0 || (n > 0 && nblocks > 0 && (!ids || !out))) {
    isk_set_error("rng export: negative size or null array");
    return ISKRA_E_ARG;
  }
  return isk_rng_export_noalloc(c, stream, step, ids, n, nblocks, out);
}

/* ---- the time loop ---- */
__global__ void k_tag_waveform(double *tagval, int tag, double off, double amp, double f,
                               double ph_turns, double t) {
  double u = f * t + ph_turns;
  u = u - floor(u);
  double s, co;
  isk_sincos2pi(u, &s, &co);
  tagval[tag] = off + amp * s;
}

/* ---- fluid species ---- */
int iskra_fluid_add(iskra_ctx *c, double q, double m, double mu, const double *n, int *fid) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (!n || !fid || !(m > 0.0) || !isfinite(q) || !isfinite(mu) || !isfinite(m)) {
    isk_set_error("fluid_add: need n, fid, finite q and mu, and m > 0");
    return ISKRA_E_ARG;
  }
  if (c->nfluid >= ISK_MAX_FLUIDS) {
    isk_set_error("fluid_add: at most %d fluid species", ISK_MAX_FLUIDS);
    return ISKRA_E_UNSUPPORTED;
  }
  const int nn = isk_nodes(c);
  ISK_HIP(hipStreamSynchronize(c->stream));
  if (!c->nfluid) {
    /* the background set so far becomes the fixed part */
    if (isk_alloc((void **)&c->bg_fixed, (size_t)nn * 8)) return ISKRA_E_HIP;
    ISK_HIP(hipMemcpy(c->bg_fixed, c->bg, (size_t)nn * 8, hipMemcpyDeviceToDevice));
  }
  IskFluid &F = c->fl[c->nfluid];
  if (isk_alloc((void **)&F.n, (size_t)nn * 8) || isk_alloc((void **)&F.dn, (size_t)nn * 8))
    return ISKRA_E_HIP;
  ISK_HIP(hipMemcpy(F.n, n, (size_t)nn * 8, hipMemcpyHostToDevice));
  ISK_HIP(hipMemset(F.dn, 0, (size_t)nn * 8));
  F.q = q;
  F.qm = q / m;
  F.mu = mu;
  *fid = c->nfluid++;
  RC(isk_fluid_bg(c));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

int iskra_fluid_advance(iskra_ctx *c, double dt) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (!(dt > 0)) {
    isk_set_error("fluid_advance: dt must be positive");
    return ISKRA_E_ARG;
  }
  RC(isk_fluid_step(c, dt));
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

int iskra_step(iskra_ctx *c, double dt, int64_t nsteps) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (c->ps.kind < 0) {
    isk_set_error("step: no Poisson solver configured");
    return ISKRA_E_STATE;
  }
  if (!(dt > 0) || nsteps < 0) {
    isk_set_error("step: dt must be positive");
    return ISKRA_E_ARG;
  }
  for (int64_t it = 0; it < nsteps; ++it) {
    hipEvent_t ev = nullptr;
    /* the counts of the call's last step stay apart (iskra_*_counts_last) */
    if (it == nsteps - 1 && (c->nmcc || c->ndsmc)) RC(isk_counts_fold(c));
    if (c->nmcc) {
      RC(isk_prof_begin(c, &ev));
      for (int m = 0; m < c->nmcc; ++m) RC(isk_mcc_run(c, m, dt, nullptr));
      RC(isk_prof_end(c, ev, ISKRA_K_MCC, 0));
    }
    for (int m = 0; m < c->ndsmc; ++m) RC(isk_dsmc_run(c, m, dt, nullptr));
    /* fluids before the deposit merge adds their charge, with the previous E */
    RC(isk_fluid_step(c, dt));
    RC(isk_push_deposit_all(c, dt));
    if (c->d_elec) RC(isk_electrodes_after_push(c, dt));
    int iters = 0;
    double rr = 0.0;
    int E_done = 0;
    const bool async = c->ps.kind == ISKRA_POISSON_SPECTRAL;
    RC(isk_prof_begin(c, &ev));
    /* the deposit kernel above packed the spectral right-hand side */
    RC(isk_poisson_solve_ex(c, ISKRA_WARM_PREV, async ? nullptr : &iters,
                            async ? nullptr : &rr, 1, &E_done, async ? 1 : 0));
    RC(isk_prof_end(c, ev, ISKRA_K_POISSON, 0));
    if (!E_done) {
      RC(isk_prof_begin(c, &ev));
      RC(isk_efield(c));
      RC(isk_prof_end(c, ev, ISKRA_K_EFIELD, 0));
    }
    /* after_loop: electrode values for the next step at t = (it-1) dt */
    double t = (double)c->step_count * dt;
    for (int k = 1; k <= ISK_MAX_TAGS; ++k)
      if (c->ps.wf_on[k] && c->ps.d_tag_val) {
        hipLaunchKernelGGL(k_tag_waveform, dim3(1), dim3(1), 0, c->stream, c->ps.d_tag_val, k,
                           c->ps.wf_off[k], c->ps.wf_amp[k], c->ps.wf_f[k], c->ps.wf_ph[k], t);
        ISK_CHECK_LAUNCH();
      }
    c->step_count++;
    c->stats.steps++;
    for (int s = 0; s < c->nspecies; ++s) {
      const int every = c->sp[s].sort_every > 0 ? c->sp[s].sort_every : c->t.sort_every;
      if (every > 0 && c->step_count % every == 0) RC(isk_sort_species(c, s, 0));
    }
  }
  return iskra_sync(c);
}

int iskra_stats_get(iskra_ctx *c, iskra_stats *out) {
  CHECK_CTX(c);
  if (!out) return ISKRA_E_ARG;
  ISK_HIP(hipStreamSynchronize(c->stream));
  RC(prof_resolve(c));
  int64_t ps = 0;
  for (int s = 0; s < c->nspecies; ++s) {
    long long v = 0;
    ISK_HIP(hipMemcpy(&v, c->sp[s].hdr + HDR_PUSHED, 8, hipMemcpyDeviceToHost));
    ps += v;
  }
  c->stats.particle_steps = ps;
  c->stats.push_particles = c->prof ? ps - g_prof_ps0 : 0;
  *out = c->stats;
  return ISKRA_OK;
}

int iskra_profile(iskra_ctx *c, int enable) {
  CHECK_CTX(c);
  ISK_HIP(hipStreamSynchronize(c->stream));
  RC(prof_resolve(c));
  c->prof = enable ? 1 : 0;
  for (int k = 0; k < ISKRA_K_COUNT; ++k) {
    c->stats.kernel_ms[k] = 0.0;
    c->stats.kernel_calls[k] = 0;
  }
  int64_t ps = 0;
  for (int s = 0; s < c->nspecies; ++s) {
    long long v = 0;
    ISK_HIP(hipMemcpy(&v, c->sp[s].hdr + HDR_PUSHED, 8, hipMemcpyDeviceToHost));
    ps += v;
  }
  g_prof_ps0 = ps;
  return ISKRA_OK;
}

int iskra_bench_copy(iskra_ctx *c, int64_t bytes, int iters, int nontemporal, double *ms) {
  CHECK_CTX(c);
  if (bytes < 16 || iters < 1 || !ms) return ISKRA_E_ARG;
  long long n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  RC(isk_alloc(&a, (size_t)n * 16));
  RC(isk_alloc(&b, (size_t)n * 16));
  ISK_HIP(hipMemsetAsync(a, 1, (size_t)n * 16, c->stream));
  long long blocks = (n + 1023) / 1024;
  auto kern = nontemporal ? k_copy<1> : k_copy<0>;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), 0, c->stream, (const isk_f4 *)a, (isk_f4 *)b, n);
  ISK_CHECK_LAUNCH();
  hipEvent_t e0, e1;
  ISK_HIP(hipEventCreate(&e0));
  ISK_HIP(hipEventCreate(&e1));
  ISK_HIP(hipEventRecord(e0, c->stream));
  for (int k = 0; k < iters; ++k)
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), 0, c->stream, (const isk_f4 *)a,
                       (isk_f4 *)b, n);
  ISK_CHECK_LAUNCH();
  ISK_HIP(hipEventRecord(e1, c->stream));
  ISK_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  ISK_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = (double)t / iters;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  isk_free(a);
  isk_free(b);
  return ISKRA_OK;
}

} /* extern "C" */

/* ---- checkpoint / resume ---- */
/* The dynamic state of a context as one host blob: the particles (live ones,
 * in the current device order), the counters that draws, product ids and
 * the sort schedule depend on, the fields the next step reads (phi, E), the
 * collision counts and DSMC remainders, fluid densities, electrode and
 * circuit state and the electrode values.  The configuration (grid, species,
 * interactions, solver, tuning) is not in it: load expects a context set up
 * the same way and checks a fingerprint of it.  Results after a load equal
 * those of the uninterrupted run: particles are re-sorted by cell on load,
 * and no result depends on particle order (integer deposit, id-keyed draws,
 * products numbered by parent id). */
namespace {
struct CkHeader {
  char magic[8];
  int32_t version, dim;
  int64_t nx, ny;
  int32_t nspecies, nmcc, ndsmc, nfluid, has_elec, has_tags, track[ISK_MAX_SPECIES], pad;
  int64_t step_count;
  uint64_t seed;
  double t_now;
};
const char CK_MAGIC[8] = {'I', 'S', 'K', 'R', 'A', 'C', 'K', '1'};

void ck_put(std::vector<char> &b, const void *p, size_t n) {
  const char *q = (const char *)p;
  b.insert(b.end(), q, q + n);
}

int ck_put_dev(std::vector<char> &b, const void *dev, size_t n) {
  size_t at = b.size();
  b.resize(at + n);
  if (n) ISK_HIP(hipMemcpy(b.data() + at, dev, n, hipMemcpyDeviceToHost));
  return ISKRA_OK;
}

CkHeader ck_header(const iskra_ctx *c) {
  CkHeader h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, CK_MAGIC, 8);
  h.version = 1;
  h.dim = c->g.dim;
  h.nx = c->g.nx;
  h.ny = c->g.ny;
  h.nspecies = c->nspecies;
  h.nmcc = c->nmcc;
  h.ndsmc = c->ndsmc;
  h.nfluid = c->nfluid;
  h.has_elec = c->d_elec != nullptr;
  h.has_tags = c->ps.d_tag_val != nullptr;
  for (int s = 0; s < c->nspecies; ++s) h.track[s] = c->sp[s].track_ids;
  h.step_count = c->step_count;
  h.seed = c->seed;
  h.t_now = c->t_now;
  return h;
}

int ck_build(iskra_ctx *c, std::vector<char> &b) {
  ISK_HIP(hipStreamSynchronize(c->stream));
  const int nn = isk_nodes(c);
  const size_t ncell = (size_t)c->g.ncx * c->g.ncy;
  CkHeader h = ck_header(c);
  ck_put(b, &h, sizeof h);
  for (int s = 0; s < c->nspecies; ++s) {
    IskSpecies *sp = &c->sp[s];
    long long hdr[HDR_WORDS];
    ISK_HIP(hipMemcpy(hdr, sp->hdr, sizeof(hdr), hipMemcpyDeviceToHost));
    const int64_t n = hdr[HDR_NP];
    std::vector<int32_t> cell(n);
    std::vector<float> ox(n), oy(n, 0.0f), vx(n), vy(n), vz(n);
    std::vector<uint32_t> id(n, 0u);
    RC(iskra_species_get_state(c, s, n, nullptr, cell.data(), ox.data(), oy.data(), vx.data(),
                               vy.data(), vz.data(), id.data()));
    int64_t live = 0;
    for (int64_t p = 0; p < n; ++p)
      if (cell[p] >= 0) {
        cell[live] = cell[p];
        ox[live] = ox[p];
        oy[live] = oy[p];
        vx[live] = vx[p];
        vy[live] = vy[p];
        vz[live] = vz[p];
        id[live] = id[p];
        ++live;
      }
    const int64_t words[3] = {live, (int64_t)hdr[HDR_NEXT_ID], (int64_t)hdr[HDR_PUSHED]};
    ck_put(b, words, sizeof words);
    ck_put(b, cell.data(), live * 4);
    ck_put(b, ox.data(), live * 4);
    ck_put(b, oy.data(), live * 4);
    ck_put(b, vx.data(), live * 4);
    ck_put(b, vy.data(), live * 4);
    ck_put(b, vz.data(), live * 4);
    ck_put(b, id.data(), live * 4);
  }
  RC(ck_put_dev(b, c->phi, (size_t)nn * 8));
  RC(ck_put_dev(b, c->E, (size_t)nn * 4 * c->g.dim));
  for (int m = 0; m < c->nmcc; ++m) {
    const IskMcc &q = c->mcc[m];
    const int32_t w[2] = {q.ncoll, (int32_t)q.step};
    ck_put(b, w, sizeof w);
    RC(ck_put_dev(b, q.nu, (size_t)nn * q.ncoll * 8));
    RC(ck_put_dev(b, q.nu_acc, (size_t)nn * q.ncoll * 8));
  }
  for (int m = 0; m < c->ndsmc; ++m) {
    const IskDsmc &q = c->dsmc[m];
    const uint32_t w[2] = {q.step, 0u};
    ck_put(b, w, sizeof w);
    RC(ck_put_dev(b, q.rem, ncell * 8));
    RC(ck_put_dev(b, q.nu, (size_t)nn * 8));
    RC(ck_put_dev(b, q.nu_acc, (size_t)nn * 8));
  }
  for (int f = 0; f < c->nfluid; ++f) {
    RC(ck_put_dev(b, c->fl[f].n, (size_t)nn * 8));
    RC(ck_put_dev(b, c->fl[f].dn, (size_t)nn * 8));
  }
  if (c->d_elec) RC(ck_put_dev(b, c->d_elec, EL_WORDS * 8));
  if (c->ps.d_tag_val) RC(ck_put_dev(b, c->ps.d_tag_val, (ISK_MAX_TAGS + 1) * 8));
  return ISKRA_OK;
}

struct CkReader {
  const char *p, *end;
  bool take(void *dst, size_t n) {
    if ((size_t)(end - p) < n) return false;
    memcpy(dst, p, n);
    p += n;
    return true;
  }
  const char *skip(size_t n) {
    if ((size_t)(end - p) < n) return nullptr;
    const char *q = p;
    p += n;
    return q;
  }
};
}  // namespace

int iskra_checkpoint_size(iskra_ctx *c, int64_t *bytes) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (!bytes) return ISKRA_E_ARG;
  std::vector<char> b;
  RC(ck_build(c, b));
  *bytes = (int64_t)b.size();
  return ISKRA_OK;
}

int iskra_checkpoint_save(iskra_ctx *c, void *buf, int64_t bytes) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (!buf) return ISKRA_E_ARG;
  std::vector<char> b;
  RC(ck_build(c, b));
  if ((int64_t)b.size() > bytes) {
    isk_set_error("checkpoint_save: %lld bytes needed, %lld given", (long long)b.size(),
                  (long long)bytes);
    return ISKRA_E_ARG;
  }
  memcpy(buf, b.data(), b.size());
  return ISKRA_OK;
}

int iskra_checkpoint_load(iskra_ctx *c, const void *buf, int64_t bytes) {
  CHECK_CTX(c);
  CHECK_GRID(c);
  if (!buf || bytes < (int64_t)sizeof(CkHeader)) return ISKRA_E_ARG;
  CkReader r{(const char *)buf, (const char *)buf + bytes};
  CkHeader h, mine = ck_header(c);
  r.take(&h, sizeof h);
  if (memcmp(h.magic, CK_MAGIC, 8) || h.version != 1) {
    isk_set_error("checkpoint_load: not a checkpoint of this library");
    return ISKRA_E_ARG;
  }
  if (h.dim != mine.dim || h.nx != mine.nx || h.ny != mine.ny || h.nspecies != mine.nspecies ||
      h.nmcc != mine.nmcc || h.ndsmc != mine.ndsmc || h.nfluid != mine.nfluid ||
      h.has_elec != mine.has_elec || h.has_tags != mine.has_tags ||
      memcmp(h.track, mine.track, sizeof h.track)) {
    isk_set_error("checkpoint_load: the context is not set up like the saved one (grid, "
                  "species, interactions, fluids, electrodes or solver differ)");
    return ISKRA_E_STATE;
  }
  const int nn = isk_nodes(c);
  const size_t ncell = (size_t)c->g.ncx * c->g.ncy;
  const char *bad = "checkpoint_load: truncated checkpoint";
  for (int s = 0; s < c->nspecies; ++s) {
    int64_t w[3];
    if (!r.take(w, sizeof w) || w[0] < 0 || w[0] > c->sp[s].cap) {
      isk_set_error("checkpoint_load: species %d: bad particle count", s);
      return ISKRA_E_ARG;
    }
    const int64_t n = w[0];
    const char *cell = r.skip(n * 4), *ox = r.skip(n * 4), *oy = r.skip(n * 4),
               *vx = r.skip(n * 4), *vy = r.skip(n * 4), *vz = r.skip(n * 4),
               *id = r.skip(n * 4);
    if (!id) {
      isk_set_error("%s", bad);
      return ISKRA_E_ARG;
    }
    RC(iskra_species_set_state(c, s, n, (const int32_t *)cell, (const float *)ox,
                               (const float *)oy, (const float *)vx, (const float *)vy,
                               (const float *)vz,
                               c->sp[s].track_ids ? (const uint32_t *)id : nullptr));
    long long hdr[HDR_WORDS];
    ISK_HIP(hipMemcpy(hdr, c->sp[s].hdr, sizeof(hdr), hipMemcpyDeviceToHost));
    hdr[HDR_NEXT_ID] = w[1];
    hdr[HDR_PUSHED] = w[2];
    ISK_HIP(hipMemcpy(c->sp[s].hdr, hdr, sizeof(hdr), hipMemcpyHostToDevice));
  }
  const char *phi = r.skip((size_t)nn * 8), *E = r.skip((size_t)nn * 4 * c->g.dim);
  if (!E) {
    isk_set_error("%s", bad);
    return ISKRA_E_ARG;
  }
  ISK_HIP(hipMemcpy(c->phi, phi, (size_t)nn * 8, hipMemcpyHostToDevice));
  ISK_HIP(hipMemcpy(c->E, E, (size_t)nn * 4 * c->g.dim, hipMemcpyHostToDevice));
  for (int m = 0; m < c->nmcc; ++m) {
    IskMcc &q = c->mcc[m];
    int32_t w[2];
    if (!r.take(w, sizeof w) || w[0] != q.ncoll) {
      isk_set_error("checkpoint_load: interaction %d has another process count", m);
      return ISKRA_E_STATE;
    }
    const size_t sz = (size_t)nn * q.ncoll * 8;
    const char *nu = r.skip(sz), *acc = r.skip(sz);
    if (!acc) {
      isk_set_error("%s", bad);
      return ISKRA_E_ARG;
    }
    q.step = (uint32_t)w[1];
    ISK_HIP(hipMemcpy(q.nu, nu, sz, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(q.nu_acc, acc, sz, hipMemcpyHostToDevice));
  }
  for (int m = 0; m < c->ndsmc; ++m) {
    IskDsmc &q = c->dsmc[m];
    uint32_t w[2];
    const char *rem = r.take(w, sizeof w) ? r.skip(ncell * 8) : nullptr;
    const char *nu = rem ? r.skip((size_t)nn * 8) : nullptr;
    const char *acc = nu ? r.skip((size_t)nn * 8) : nullptr;
    if (!acc) {
      isk_set_error("%s", bad);
      return ISKRA_E_ARG;
    }
    q.step = w[0];
    ISK_HIP(hipMemcpy(q.rem, rem, ncell * 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(q.nu, nu, (size_t)nn * 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(q.nu_acc, acc, (size_t)nn * 8, hipMemcpyHostToDevice));
  }
  for (int f = 0; f < c->nfluid; ++f) {
    const char *n = r.skip((size_t)nn * 8), *dn = r.skip((size_t)nn * 8);
    if (!dn) {
      isk_set_error("%s", bad);
      return ISKRA_E_ARG;
    }
    ISK_HIP(hipMemcpy(c->fl[f].n, n, (size_t)nn * 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(c->fl[f].dn, dn, (size_t)nn * 8, hipMemcpyHostToDevice));
  }
  RC(isk_fluid_bg(c));
  if (c->d_elec) {
    const char *el = r.skip(EL_WORDS * 8);
    if (!el) {
      isk_set_error("%s", bad);
      return ISKRA_E_ARG;
    }
    ISK_HIP(hipMemcpy(c->d_elec, el, EL_WORDS * 8, hipMemcpyHostToDevice));
  }
  if (c->ps.d_tag_val) {
    const char *tv = r.skip((ISK_MAX_TAGS + 1) * 8);
    if (!tv) {
      isk_set_error("%s", bad);
      return ISKRA_E_ARG;
    }
    ISK_HIP(hipMemcpy(c->ps.d_tag_val, tv, (ISK_MAX_TAGS + 1) * 8, hipMemcpyHostToDevice));
    memcpy(c->ps.tag_val, tv, (ISK_MAX_TAGS + 1) * 8);
  }
  if (r.p != r.end) {
    isk_set_error("checkpoint_load: %lld bytes left over", (long long)(r.end - r.p));
    return ISKRA_E_ARG;
  }
  c->step_count = h.step_count;
  c->seed = h.seed;
  c->t_now = h.t_now;
  ISK_HIP(hipStreamSynchronize(c->stream));
  return ISKRA_OK;
}

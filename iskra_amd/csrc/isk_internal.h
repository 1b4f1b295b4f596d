/*
 * isk_internal.h -- device data structures and helpers shared by the HIP
 * translation units of libiskra_hip.so (not part of the public ABI).
 *
 * Particle state of a species lives in HBM as structure-of-arrays:
 *   cell  int32   packed node index of the lower-left node: i | (j << 16);
 *                 -1 marks a discarded particle (compacted by the next sort)
 *   ox,oy float   in-cell offset in [0,1) along x / y  (oy unused in 1-D)
 *   vx,vy,vz float velocity in m/s
 *   id    uint32  particle identifier (optional; keys the MCC random draws)
 * Two copies of every array exist; the counting sort scatters from one into
 * the other and flips `cur`.
 *
 * Cells are grouped into square tiles of T x T cells (T cells in 1-D).  After
 * a sort the particles of a tile are contiguous; the chunk table splits every
 * tile into chunks of at most CH particles, one work-group each.  A chunk's
 * work-group stages E of its tile plus a margin of M cells in LDS (window of
 * W = T + 2M + 1 nodes per axis) and accumulates its CIC deposit there.
 */
#ifndef ISK_INTERNAL_H
#define ISK_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/iskra.h"

#define ISK_MAX_SPECIES 8
#define ISK_MAX_MCC 4
#define ISK_MAX_DSMC 4
#define ISK_MAX_FLUIDS 4
/* floating electrodes (each one sigma degree of freedom) */
#define ISK_MAX_ELECTRODES 8

/* surface codes of the tracker's edge table (per cell, per direction -x, +x,
 * -y, +y): none, absorbing (absorbing surfaces and fixed electrodes),
 * reflective, a surface whose hit does nothing, floating electrode k */
enum { ISK_SURF_NONE = 0, ISK_SURF_ABSORB = 1, ISK_SURF_REFLECT = 2, ISK_SURF_PASS = 3,
       ISK_SURF_ELECTRODE = 16 };
#define ISK_MAX_COLL 8
#define ISK_MAX_TAGS 16
/* largest number of sort bins (tiles, or cells in a tile); LDS-resident */
#define ISK_MAX_BINS 8192

/* fixed-point scale of CIC weights: one particle deposits 2^32 in total */
#define ISK_FIX_ONE 4294967296.0

/* header words kept in device memory per species */
enum {
  HDR_NP = 0,        /* live + not yet compacted particle count */
  HDR_NP_SORTED = 1, /* particles covered by the chunk table */
  HDR_NCHUNKS = 2,   /* chunks in the chunk table */
  HDR_NEXT_ID = 3,   /* next particle id handed out */
  HDR_PUSHED = 4,    /* accumulated live particles pushed (particle-steps) */
  HDR_OVERFLOW = 5,  /* set when an append would exceed the capacity */
  HDR_WORDS = 8
};

struct IskChunk {
  int tile;
  int start;
  int count;
  int pad;
};

struct IskGrid {
  int dim;
  int nx, ny;   /* nodes (ny = 1 in 1-D) */
  int ncx, ncy; /* cells (ncy = 1 in 1-D: the single row j = 0) */
  double dx, dy;
  double x0, y0;
  int volper[4];
  int axial; /* (r, z) grid: axis 0 radial, symmetry axis at i = 0 */
};

struct IskTune {
  int T, log2T, CH, M, W;
  int sort_every;
};

struct IskSpecies {
  double q, m, w0;
  int64_t cap;
  int track_ids;
  int sort_every; /* 0: the context's */
  int margin;     /* window margin in cells, -1: the context's (never above it) */
  int bnd[2];
  int cur;
  int32_t *cell[2];
  float *ox[2], *oy[2], *vx[2], *vy[2], *vz[2];
  uint32_t *id[2];
  long long *hdr;            /* HDR_WORDS int64 */
  IskChunk *chunks;          /* max_chunks */
  int *tile_chunk_begin;     /* ntiles + 1 */
  long long *partial;        /* max_chunks * W^dim */
  long long *tilewin;        /* ntiles * W^dim: chunk windows summed per tile (2-D) */
  unsigned long long *gfall; /* nodes */
  long long *cnt;            /* nodes: merged deposit (fixed point) */
  int *live_cnt;             /* max_chunks: live particles per push work-group */
  unsigned long long *ehit;  /* ISK_MAX_ELECTRODES: particles absorbed by each floating electrode */
  int64_t max_chunks;
  int deposited;             /* cnt holds this step's deposit */
};

struct IskMcc {
  int src, ion;
  int ncoll;
  int types[ISK_MAX_COLL];
  double energy[ISK_MAX_COLL];
  std::vector<int> tab_off;
  std::vector<double> tab_eps, tab_sig;
  double *d_eps, *d_sig;
  int *d_off;
  double target_m, target_q, target_T;
  double max_sigma_g;
  double m_eV;               /* source mass in eV units (mcc.jl:26) */
  double *target_n;          /* nodes */
  double max_n;
  unsigned long long *nu;    /* nodes * ncoll event counts since the last fold */
  unsigned long long *nu_acc;/* the counts folded out of nu (iskra_step) */
  unsigned long long *events;/* ncoll */
  int *errflag;
  /* ionisation product staging */
  int *prod_flag;            /* per source particle: 1 if it ionised */
  float *stage;              /* per source particle: 2 offsets + 3 v of the new electron + 3 v of the ions */
  int *blk_cnt;              /* per block product count, scanned in place */
  int *cand_i;               /* per block: accepted candidates (list records), in order */
  int *cand_cnt;             /* per block accepted count */
  uint32_t *stage_pid;       /* per staged product: the parent's id */
  uint32_t *pid_c;           /* parent ids of the products in append order */
  long long *snap;           /* species counts and next ids before the append */
  int64_t stage_cap;
  uint32_t stream;
  uint32_t step;
};

/* one DSMC interaction (a single DSMC.ElasticCollision) */
struct IskDsmc {
  int src = -1, tgt = -1;
  int ntab = 0;
  double sgmax = 0.0;              /* maximum(rate) * argmax(rate) */
  double *d_g = nullptr, *d_sig = nullptr;
  double *rem = nullptr;           /* cells: collisions_remaining */
  unsigned long long *nu = nullptr;     /* nodes: accepted events since the last fold */
  unsigned long long *nu_acc = nullptr; /* the counts folded out of nu (iskra_step) */
  unsigned long long *events = nullptr; /* accepted events of the last call */
  int *rng = nullptr;              /* 4 x cells: cell ranges of source, target */
  uint32_t stream = 0, step = 0;
};

/* d_elec layout (doubles): sigma, collected charge and area per sigma dof,
 * then the RLC circuit: current, charge, time, last external voltage */
enum { EL_SIGMA = 0, EL_DQ = ISK_MAX_ELECTRODES, EL_AREA = 2 * ISK_MAX_ELECTRODES,
       EL_CIRC = 3 * ISK_MAX_ELECTRODES, EL_WORDS = 3 * ISK_MAX_ELECTRODES + 8 };

/* the surface tracker: edge table and tracked cells on the device (host copy
 * of the codes kept for updates) */
struct IskTracker {
  int on = 0;
  std::vector<uint8_t> h_surf;
  uint8_t *surf = nullptr, *tracked = nullptr;
};

/* electrodes (create_electrode, problem/configuration.jl:17-76) */
struct IskElectrode {
  int fixed = 0;
  int tag = 0;        /* fixed: its Dirichlet tag */
  int ref_node = 0;   /* first node of the mask (column-major order) */
  double area = 0.0;
  double sigma0 = 0.0;
};

struct IskPoisson {
  int kind = -1;
  double eps0 = 1.0;
  double rtol = 1e-8;
  int max_iter = 10000;
  int per_x = 0, per_y = 0;
  std::vector<int8_t> tags;
  int8_t *d_tags = nullptr;
  double tag_val[ISK_MAX_TAGS + 1];
  double *d_tag_val = nullptr;
  /* waveforms */
  int wf_on[ISK_MAX_TAGS + 1];
  double wf_off[ISK_MAX_TAGS + 1], wf_amp[ISK_MAX_TAGS + 1],
      wf_f[ISK_MAX_TAGS + 1], wf_ph[ISK_MAX_TAGS + 1];
  int singular = 0; /* no Dirichlet node at all -> mean-free solution */
  /* 1-D direct: factored tridiagonal (host-built) */
  double *d_tri_a = nullptr, *d_tri_cp = nullptr, *d_tri_inv = nullptr;
  double *d_work = nullptr;
  /* CG state */
  double *x = nullptr, *r = nullptr, *Ap = nullptr;
  float *z = nullptr, *p[2] = {nullptr, nullptr};
  double *scal = nullptr;      /* scalars, see poisson.hip */
  double *partials = nullptr;  /* per-block partial sums (2 per block) */
  double *cgp = nullptr;       /* CG per-work-group partials (7 per work-group) */
  int partials_blocks = 0;     /* blocks the partials array can hold */
  int *iflag = nullptr;        /* [0] done, [1] iterations */
  /* multigrid */
  int nlev = 0;
  int ln[16];                  /* nodes per axis on each level */
  float *me[16], *mf[16];      /* correction and rhs per level (FP32) */
  int coarse_lev = 0;          /* first level solved by the single-WG kernel */
  /* spectral (DST-I) direct solve: interior sizes, sine bases, 1/eigenvalue,
   * right-hand side and two work matrices (row-major, rows = j) */
  int sp_nx = 0, sp_ny = 0;
  double *sp_Sx = nullptr, *sp_Sy = nullptr, *sp_Li = nullptr, *sp_F = nullptr,
         *sp_T1 = nullptr, *sp_T2 = nullptr;
  int last_iters = 0;
  double last_relres = 0.0;
  /* floating electrodes: rows replaced as apply_neumann does
   * (generalized_poisson.jl:222-269), kept in the solver's scaled form
   * sd x_n - sum_q sc_q x_{nb_q} = rho_n dx^2/eps0 + ssig sigma[sdof]; every
   * other row is the 5-point row.  Such a system is not symmetric: it is
   * solved by BiCGStab (poisson.hip) */
  std::vector<int> h_srow;           /* nodes: special row index or -1 */
  std::vector<double> h_sd, h_sc, h_ssig;
  std::vector<int> h_snb, h_sdof;
  int *d_srow = nullptr, *d_snb = nullptr, *d_sdof = nullptr;
  double *d_sd = nullptr, *d_sc = nullptr, *d_ssig = nullptr;
  double *bi = nullptr;              /* BiCGStab vectors and partials */
  double *d_dg = nullptr;            /* row diagonals (Jacobi) */
};

struct IskFluid {
  double *n = nullptr, *dn = nullptr;
  double q = 0.0, qm = 0.0, mu = 0.0;
};

struct iskra_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int have_grid = 0;
  IskGrid g;
  int pusher = ISKRA_PUSHER_XY;
  IskTune t;
  int ntx = 1, nty = 1, ntiles = 1;
  int nspecies = 0;
  IskSpecies sp[ISK_MAX_SPECIES];
  int nmcc = 0;
  IskMcc mcc[ISK_MAX_MCC];
  int ndsmc = 0;
  IskDsmc dsmc[ISK_MAX_DSMC];
  IskTracker trk;
  /* mobile fluid species (advance!(::FluidSpecies)): density and the last
   * step's change per node; bg then holds bg_fixed + sum q n */
  int nfluid = 0;
  IskFluid fl[ISK_MAX_FLUIDS];
  double *bg_fixed = nullptr;
  int nelec = 0;
  IskElectrode elec[ISK_MAX_ELECTRODES];
  int nfloat = 0;              /* floating electrodes = sigma degrees of freedom */
  int float_elec[ISK_MAX_ELECTRODES]; /* electrode id of each sigma dof */
  /* device state of the floating electrodes, per sigma dof: sigma (C/m^2),
   * charge collected (C), area; the RLC circuit (RLC_WORDS doubles) */
  double *d_elec = nullptr;
  int circuit_on = 0;
  double circ_R = 0, circ_L = 0, circ_C = 0, circ_v[4] = {0, 0, 0, 0};
  int circ_pos = -1, circ_neg = -1;
  IskPoisson ps;
  /* fields */
  double *rho = nullptr, *phi = nullptr, *bg = nullptr, *vol = nullptr;
  float *E = nullptr; /* 1-D: Ex[nx]; 2-D: (Ex,Ey) interleaved per node */
  /* sort scratch */
  int *sort_H = nullptr, *sort_S = nullptr, *sort_tot = nullptr,
      *sort_start = nullptr;
  int64_t sort_H_words = 0, sort_S_words = 0;
  /* misc scratch */
  float *gather_out = nullptr;
  int64_t gather_cap = 0;
  uint64_t seed = 0;
  int64_t step_count = 0;
  iskra_stats stats;
  double t_now = 0.0;
  /* profiling: event pairs recorded around launches, resolved at stats time */
  int prof = 0;
  struct Pending {
    hipEvent_t a, b;
    int kind;
    int64_t particles;
  };
  std::vector<Pending> prof_pending;
  std::vector<hipEvent_t> prof_pool;
};

/* bracket a launch with events when profiling (api.hip) */
int isk_prof_begin(iskra_ctx *c, hipEvent_t *a);
int isk_prof_end(iskra_ctx *c, hipEvent_t a, int kind, int64_t particles);

/* ---- error plumbing (api.hip) ---- */
void isk_set_error(const char *fmt, ...);
void isk_clear_error(void);
int isk_hip_fail(hipError_t e, const char *what, const char *file, int line);
#define ISK_HIP(call)                                                   \
  do {                                                                  \
    hipError_t _e = (call);                                             \
    if (_e != hipSuccess) return isk_hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)
#define ISK_CHECK_LAUNCH() ISK_HIP(hipGetLastError())

/* ---- Philox4x32-10 (Salmon et al., SC'11) ---- */
struct IskU4 {
  uint32_t x, y, z, w;
};
__host__ __device__ inline void isk_mulhilo(uint32_t a, uint32_t b, uint32_t *hi,
                                            uint32_t *lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}
__host__ __device__ inline IskU4 isk_philox(IskU4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    isk_mulhilo(0xD2511F53u, c.x, &hi0, &lo0);
    isk_mulhilo(0xCD9E8D57u, c.z, &hi1, &lo1);
    IskU4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}
/* uniform double strictly inside (0,1) from two words: 52 bits, centred in
 * its bin (the largest value, 1 - 2^-53, is representable) */
__host__ __device__ inline double isk_u01(uint32_t a, uint32_t b) {
  uint64_t k = ((uint64_t)(a >> 6) << 26) | (uint64_t)(b >> 6);
  return ((double)k + 0.5) * (1.0 / 4503599627370496.0);
}
/* uniform float in [0,1) from 24 bits */
__host__ __device__ inline float isk_u24(uint32_t a) {
  return (float)(a >> 8) * (1.0f / 16777216.0f);
}

/* ---- small host helpers ---- */
int isk_alloc(void **p, size_t bytes);
int isk_free(void *p);
int isk_nodes(const iskra_ctx *c);
/* deposit-window margin of a species and the window width T + 2M + 1 */
inline int isk_sp_margin(const iskra_ctx *c, int sid) {
  const int m = c->sp[sid].margin;
  return (m >= 0 && m < c->t.M) ? m : c->t.M;
}
inline int isk_sp_window(const iskra_ctx *c, int sid) {
  return c->t.T + 2 * isk_sp_margin(c, sid) + 1;
}
int isk_species_ok(iskra_ctx *c, int sid);

/* ---- kernels exposed across translation units ---- */
int isk_fill_i64(iskra_ctx *c, void *p, int64_t n, long long v);
int isk_sort_species(iskra_ctx *c, int sid, int exact_cell);
int isk_push_species(iskra_ctx *c, int sid, double dt, int do_push);
int isk_merge_species(iskra_ctx *c, int sid, int count_live);
int isk_rho(iskra_ctx *c);
int isk_push_deposit_all(iskra_ctx *c, double dt);
int isk_gather_species(iskra_ctx *c, int sid);
int isk_convert_upload(iskra_ctx *c, int sid, int64_t np, const double *x,
                       const double *v, const uint32_t *id);
int isk_sample_maxwellian(iskra_ctx *c, int sid, int64_t n, const double *wx,
                          const double *dx, const double *wv, const double *dv,
                          uint64_t seed, uint32_t stream);
int isk_poisson_setup(iskra_ctx *c);
int isk_poisson_free(iskra_ctx *c);
int isk_poisson_solve(iskra_ctx *c, int warm, int *iters, double *relres);
int isk_poisson_check(iskra_ctx *c, double *relres);
int isk_bicgstab_solve(iskra_ctx *c, int warm, int *iters, double *relres);
int isk_electrodes_after_push(iskra_ctx *c, double dt);
void isk_electrodes_release(iskra_ctx *c);
int isk_poisson_solve_ex(iskra_ctx *c, int warm, int *iters, double *relres, int with_E,
                         int *E_done, int packed);
struct IskSpPack;
int isk_poisson_pack_desc(iskra_ctx *c, IskSpPack *pk);
int isk_efield(iskra_ctx *c);
int isk_mcc_setup(iskra_ctx *c, int id);
int isk_mcc_run(iskra_ctx *c, int id, double dt, int64_t *events);
int isk_rng_export(iskra_ctx *c, uint32_t stream, uint32_t step,
                   const uint32_t *ids, int64_t n, int nblocks, uint32_t *out);
int isk_dsmc_add(iskra_ctx *c, int src, int tgt, int ntab, const double *tab_g,
                 const double *tab_sig, int *id);
int isk_dsmc_run(iskra_ctx *c, int idx, double dt, int64_t *events);
int isk_dsmc_counts(iskra_ctx *c, int idx, int reset, int64_t *out);
void isk_dsmc_release(IskDsmc *m);
/* fluid species (fluid.hip) */
int isk_fluid_step(iskra_ctx *c, double dt);
int isk_fluid_bg(iskra_ctx *c);
void isk_fluid_release(iskra_ctx *c);
int isk_noalloc_init(int device);
int isk_rng_export_noalloc(iskra_ctx *c, uint32_t stream, uint32_t step,
                           const uint32_t *ids, int64_t n, int nblocks, uint32_t *out);
int isk_species_alloc(iskra_ctx *c, IskSpecies *s);
void isk_species_release(IskSpecies *s);
int isk_species_reset_chunks(iskra_ctx *c, int sid);

/* ---- batched launches over instances (iskra_batch_*) ----
 * A batched launch runs one kernel type for n contexts at once: work-group
 * g belongs to instance i with bstart[i] <= g < bstart[i+1] and acts as
 * work-group g - bstart[i] of that instance's own launch, through the same
 * device function as the single-context kernel, so every instance computes
 * exactly what it computes alone. */
enum IskBKind {
  ISK_B_FILL = 0, ISK_B_PUSH1, ISK_B_MERGE1, ISK_B_RHO,
  ISK_B_SORT_HIST, ISK_B_SORT_COLSCAN, ISK_B_SORT_SEGSCAN, ISK_B_SORT_BINSCAN,
  ISK_B_SORT_SCATTER, ISK_B_SORT_FINISH,
  ISK_B_MCC, ISK_B_MCC_COLL, ISK_B_MCC_SCAN, ISK_B_MCC_APPEND, ISK_B_MCC_COMMIT,
  ISK_B_TRI, ISK_B_EFIELD, ISK_B_WAVEFORM, ISK_B_STEPINC
};
struct IskBLaunch {
  int kind = -1;
  void *args = nullptr;   /* device array of n per-instance argument structs */
  int *bstart = nullptr;  /* device, n + 1 work-group offsets */
  int n = 0, total = 0;   /* instances, work-groups of the launch */
  int block = 256;        /* work-group size */
  size_t shm = 0;         /* dynamic LDS bytes (maximum over instances) */
  long long *step_dev = nullptr; /* device step counter of the batch */
};
/* the finite-difference Poisson operator on the node grid: periodic
 * couplings cxp / cyp, node tags (0 free, k > 0 Dirichlet with tagval[k]) */
struct Sten {
  int dim, nx, ny, per_x, per_y;
  double cxp, cyp;
  const int8_t *tags;
  const double *tagval;
  /* axisymmetric operator (rows multiplied by w_i): radial link i -- i+1 of
   * weight i + 1/2, axial links of weight w_i az, az = dr^2 / dz^2, with
   * w_0 = 1/2, w_i = i, w_{nx-1} = nx - 3/2 */
  int axial;
  double az;
};

/* row weight w_i of the axisymmetric operator */
__host__ __device__ __forceinline__ double isk_rz_w(int i, int nx) {
  return i == 0 ? 0.5 : (i == nx - 1 ? (double)nx - 1.5 : (double)i);
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
  if (s.axial) {
    const double cz = isk_rz_w(i, s.nx) * s.az;
    DIRNB(i > 0, n - 1, (double)i - 0.5)
    DIRNB(i < s.nx - 1, n + 1, (double)i + 0.5)
    DIRNB(j > 0, n - s.nx, cz)
    DIRNB(j < s.ny - 1, n + s.nx, cz)
    return acc;
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

/* the spectral solve's right-hand side, packed over the interior nodes
 * (row-major [j-1][i-1]): F = rho scale + Dirichlet neighbours; written by
 * the time loop's deposit kernel when F is not null */
struct IskSpPack {
  double *F;
  Sten s;
  double scale;
};

/* F at interior node (i, j).  The spectral solve's Dirichlet nodes are the
 * boundary ring, so only ring neighbours contribute: nodes two or more away
 * from it add +0.0, which is what dirichlet_rhs returns for them */
__device__ __forceinline__ double isk_sp_pack_value(const IskSpPack &pk, double rho, int i, int j) {
  const bool near = i == 1 || j == 1 || i == pk.s.nx - 2 || j == pk.s.ny - 2;
  return rho * pk.scale + (near ? dirichlet_rhs(pk.s, i, j) : 0.0);
}

/* work-group barrier that orders LDS accesses only: global loads still in
 * flight stay in flight across it (a __syncthreads after a global store
 * waits for every outstanding global access of the wave, loads included) */
__device__ __forceinline__ void isk_lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ int isk_bfind(const int *bstart, int n, int g) {
  int lo = 0, hi = n; /* largest i with bstart[i] <= g */
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (bstart[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}
/* uploads args (n structs of `size` bytes) and the block prefix of `blocks` */
int isk_blaunch_upload(IskBLaunch *L, int kind, const void *args, size_t size,
                       const std::vector<int> &blocks, int block, size_t shm);
void isk_blaunch_free(IskBLaunch *L);
/* builders (cur: the species buffer each instance reads, per context) */
int isk_bbuild_fill(iskra_ctx **cs, int n, int sid, IskBLaunch *L);
int isk_bbuild_push(iskra_ctx **cs, int n, int sid, double dt, const int *cur, IskBLaunch *L);
int isk_bbuild_merge(iskra_ctx **cs, int n, int sid, IskBLaunch *L);
int isk_bbuild_rho(iskra_ctx **cs, int n, IskBLaunch *L);
int isk_bbuild_sort(iskra_ctx **cs, int n, int sid, const int *cur, IskBLaunch L[6]);
int isk_bbuild_mcc(iskra_ctx **cs, int n, int m, double dt, const int *cur_src,
                   const int *cur_ion, long long *step_dev, IskBLaunch L[5]);
int isk_bbuild_tri(iskra_ctx **cs, int n, IskBLaunch *L);
int isk_bbuild_efield(iskra_ctx **cs, int n, IskBLaunch *L);
int isk_bbuild_waveform(iskra_ctx **cs, int n, double dt, long long *step_dev, IskBLaunch *L);
/* launchers, one per translation unit (by kind) */
int isk_blaunch_particles(const IskBLaunch &L, hipStream_t st);
int isk_blaunch_mcc(const IskBLaunch &L, hipStream_t st);
int isk_blaunch_poisson(const IskBLaunch &L, hipStream_t st);

#endif /* ISK_INTERNAL_H */

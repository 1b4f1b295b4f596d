/*
 * electrode.hip -- the surface tracker's tables, electrodes and the RLC
 * circuit of the reference (ParticleInCell/src/pic/surfaces/{track,check,hit}.jl,
 * pic/circuit_coupling.jl, problem/configuration.jl:17-76,
 * Circuit/src/Circuit.jl:117-136).
 *
 * The tracker itself runs inside the push kernel (particles.hip,
 * track_path): here are the edge table it reads (one code per cell and
 * direction -x, +x, -y, +y, built like build_default_surface! and
 * build_surface_lookup!, later calls overriding earlier ones) and the tracked
 * cells (every cell that appears in an edge entry, as the reference's
 * `(i, j) in st` tests both cells of every key).
 *
 * Electrodes: a fixed electrode is a Dirichlet tag plus an absorbing surface;
 * a floating one is a sigma degree of freedom, the rows of apply_neumann for
 * its nodes (generalized_poisson.jl:222-269) and a surface that absorbs and
 * counts particles.  After each push one work-item folds the counts into the
 * electrodes' charge (dq = q w0 per particle, sigma += dq / area) and, when a
 * circuit is attached, advances it by advance_circuit! and couples it as
 * advance!(::CircuitRLC) does (sigma of the first floating electrode +=
 * -dt i / area of the positive electrode).  No host round trip per step.
 */
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "isk_fp64math.h"
#include "isk_internal.h"

static int tracker_upload(iskra_ctx *c) {
  IskTracker &t = c->trk;
  const int ncx = c->g.ncx, ncy = c->g.ncy, ncell = ncx * ncy;
  std::vector<uint8_t> tracked(ncell, 0);
  for (int cj = 0; cj < ncy; ++cj)
    for (int ci = 0; ci < ncx; ++ci) {
      const int cell = ci + ncx * cj;
      for (int d = 0; d < 4; ++d) {
        if (!t.h_surf[4 * cell + d]) continue;
        tracked[cell] = 1; /* the key's first cell */
        const int ni = ci + (d == 0 ? -1 : d == 1 ? 1 : 0);
        const int nj = cj + (d == 2 ? -1 : d == 3 ? 1 : 0);
        if (ni >= 0 && ni < ncx && nj >= 0 && nj < ncy) tracked[ni + ncx * nj] = 1;
      }
    }
  if (!t.surf) {
    if (isk_alloc((void **)&t.surf, (size_t)ncell * 4) || isk_alloc((void **)&t.tracked, ncell))
      return ISKRA_E_HIP;
  }
  ISK_HIP(hipStreamSynchronize(c->stream));
  ISK_HIP(hipMemcpy(t.surf, t.h_surf.data(), (size_t)ncell * 4, hipMemcpyHostToDevice));
  ISK_HIP(hipMemcpy(t.tracked, tracked.data(), ncell, hipMemcpyHostToDevice));
  return ISKRA_OK;
}

/* create_surface_tracker(grid, ds) (track.jl:101-114) with
 * build_default_surface! (track.jl:27-43): every edge of the grid boundary
 * gets ds */
static int tracker_create(iskra_ctx *c, int code) {
  IskTracker &t = c->trk;
  const int ncx = c->g.ncx, ncy = c->g.ncy;
  t.h_surf.assign((size_t)ncx * ncy * 4, 0);
  for (int ci = 0; ci < ncx; ++ci) {
    t.h_surf[4 * (ci + ncx * 0) + 2] = (uint8_t)code;
    t.h_surf[4 * (ci + ncx * (ncy - 1)) + 3] = (uint8_t)code;
  }
  for (int cj = 0; cj < ncy; ++cj) {
    t.h_surf[4 * (0 + ncx * cj) + 0] = (uint8_t)code;
    t.h_surf[4 * (ncx - 1 + ncx * cj) + 1] = (uint8_t)code;
  }
  t.on = 1;
  return tracker_upload(c);
}

/* build_surface_lookup! (track.jl:45-60): a cell edge whose two end nodes
 * are both in the mask gets the surface, seen from that cell */
static void surface_lookup(iskra_ctx *c, const uint8_t *nodes, int code) {
  IskTracker &t = c->trk;
  const int nx = c->g.nx, ncx = c->g.ncx, ncy = c->g.ncy;
  for (int ci = 0; ci < ncx; ++ci)
    for (int cj = 0; cj < ncy; ++cj) {
      const bool a = nodes[ci + nx * cj], b = nodes[ci + 1 + nx * cj],
                 cc = nodes[ci + 1 + nx * (cj + 1)], d = nodes[ci + nx * (cj + 1)];
      uint8_t *e = &t.h_surf[4 * (ci + ncx * cj)];
      if (a && b) e[2] = (uint8_t)code;
      if (b && cc) e[1] = (uint8_t)code;
      if (cc && d) e[3] = (uint8_t)code;
      if (d && a) e[0] = (uint8_t)code;
    }
}

static int surf_code(int surface) {
  switch (surface) {
    case ISKRA_SURF_ABSORBING: return ISK_SURF_ABSORB;
    case ISKRA_SURF_REFLECTIVE: return ISK_SURF_REFLECT;
    case ISKRA_SURF_PERIODIC: return ISK_SURF_PASS;
  }
  return -1;
}

static int need_2d_cartesian(iskra_ctx *c, const char *what) {
  if (!c->have_grid || c->g.dim != 2 || c->g.axial) {
    isk_set_error("%s: the surface tracker and electrodes need a 2-D Cartesian grid", what);
    return ISKRA_E_UNSUPPORTED;
  }
  if (c->pusher != ISKRA_PUSHER_XY) {
    isk_set_error("%s: the surface tracker works with the Cartesian pusher", what);
    return ISKRA_E_UNSUPPORTED;
  }
  return ISKRA_OK;
}

int iskra_tracker_create(iskra_ctx *c, int default_surface) {
  if (!c) return ISKRA_E_ARG;
  int rc = need_2d_cartesian(c, "tracker_create");
  if (rc) return rc;
  const int code = surf_code(default_surface);
  if (code < 0) {
    isk_set_error("tracker_create: unknown surface %d", default_surface);
    return ISKRA_E_ARG;
  }
  return tracker_create(c, code);
}

int iskra_tracker_surface(iskra_ctx *c, const uint8_t *nodes, int surface) {
  if (!c || !nodes) return ISKRA_E_ARG;
  int rc = need_2d_cartesian(c, "tracker_surface");
  if (rc) return rc;
  if (!c->trk.on) {
    isk_set_error("tracker_surface: no tracker (iskra_tracker_create first)");
    return ISKRA_E_STATE;
  }
  const int code = surf_code(surface);
  if (code < 0) {
    isk_set_error("tracker_surface: unknown surface %d", surface);
    return ISKRA_E_ARG;
  }
  surface_lookup(c, nodes, code);
  return tracker_upload(c);
}

int iskra_tracker_table(iskra_ctx *c, uint8_t *codes, uint8_t *tracked) {
  if (!c) return ISKRA_E_ARG;
  if (!c->trk.on) {
    isk_set_error("tracker_table: no tracker");
    return ISKRA_E_STATE;
  }
  const int ncell = c->g.ncx * c->g.ncy;
  ISK_HIP(hipStreamSynchronize(c->stream));
  if (codes) ISK_HIP(hipMemcpy(codes, c->trk.surf, (size_t)ncell * 4, hipMemcpyDeviceToHost));
  if (tracked) ISK_HIP(hipMemcpy(tracked, c->trk.tracked, ncell, hipMemcpyDeviceToHost));
  return ISKRA_OK;
}

/* apply_neumann (generalized_poisson.jl:222-269) for one floating electrode,
 * each replaced row written as sd x - sum sc x_nb = f + ssig sigma in the
 * solver's scaling (the reference row times -dx^2; eps_r = 1) */
static int neumann_rows(iskra_ctx *c, const uint8_t *nodes, int dof) {
  IskPoisson &p = c->ps;
  const int nx = c->g.nx, ny = c->g.ny, nn = nx * ny;
  const double dx = c->g.dx, dy = c->g.dy, h2 = dx * dx;
  if ((int)p.h_srow.size() != nn) p.h_srow.assign(nn, -1);
  for (int j1 = 1; j1 <= ny; ++j1)
    for (int i1 = 1; i1 <= nx; ++i1) {
      const int n = (i1 - 1) + nx * (j1 - 1);
      if (!nodes[n]) continue;
      auto in = [&](int i, int j) { return nodes[(i - 1) + nx * (j - 1)] != 0; };
      const bool a = i1 > 1 ? in(i1 - 1, j1) : false;
      const bool b = i1 < nx ? in(i1 + 1, j1) : false;
      const bool cc = j1 > 1 ? in(i1, j1 - 1) : true;
      const bool d = j1 < ny ? in(i1, j1 + 1) : true;
      double ra, rn1, rn2 = 0.0, rs;
      int m1, m2 = -1;
      if (cc && d && !a && !b) {
        const int ip = i1 == 1 ? i1 + 1 : i1 - 1;
        ra = -2.0 / dx;
        rn1 = 2.0 / dx;
        m1 = (ip - 1) + nx * (j1 - 1);
        rs = 2.0;
      } else if (cc != d) {
        const int ip = i1 == 1 ? i1 + 1 : i1 - 1;
        const int jp = cc ? j1 + 1 : j1 - 1;
        ra = -4.0 / (3.0 * dx * dx) - 4.0 / (3.0 * dy * dy);
        rn1 = 4.0 / (3.0 * dx * dx);
        rn2 = 4.0 / (3.0 * dy * dy);
        m1 = (ip - 1) + nx * (j1 - 1);
        m2 = (i1 - 1) + nx * (jp - 1);
        rs = (2.0 / 3.0) * (dx + dy) / (dx * dy);
      } else {
        continue; /* the reference keeps the Laplacian row */
      }
      if (p.tags[n] || p.h_srow[n] >= 0) {
        isk_set_error("electrode_create: node %d is already a Dirichlet or electrode node", n);
        return ISKRA_E_UNSUPPORTED;
      }
      const int k = (int)p.h_sd.size();
      p.h_srow[n] = k;
      p.h_sd.push_back(-h2 * ra);
      p.h_snb.push_back(m1);
      p.h_snb.push_back(m2);
      p.h_sc.push_back(h2 * rn1);
      p.h_sc.push_back(h2 * rn2);
      p.h_ssig.push_back(h2 * rs);
      p.h_sdof.push_back(dof);
    }
  return ISKRA_OK;
}

static int poisson_rebuild(iskra_ctx *c) {
  ISK_HIP(hipStreamSynchronize(c->stream));
  isk_poisson_free(c);
  return isk_poisson_setup(c);
}

int iskra_electrode_create(iskra_ctx *c, const uint8_t *nodes, int fixed, double sigma,
                           double phi, int *electrode_id) {
  if (!c || !nodes || !electrode_id) return ISKRA_E_ARG;
  int rc = need_2d_cartesian(c, "electrode_create");
  if (rc) return rc;
  IskPoisson &p = c->ps;
  if (p.kind < 0) {
    isk_set_error("electrode_create: no field solver (iskra_poisson_set first)");
    return ISKRA_E_STATE;
  }
  if (c->nelec >= ISK_MAX_ELECTRODES) {
    isk_set_error("electrode_create: at most %d electrodes", ISK_MAX_ELECTRODES);
    return ISKRA_E_ARG;
  }
  const int nx = c->g.nx, ny = c->g.ny, nn = nx * ny;
  /* calculate_area and find_reference_node (configuration.jl:41-56), dz = 1 */
  double area = 0.0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int n = i + nx * j;
      if (!nodes[n]) continue;
      area += (i < nx - 1 && nodes[n + 1]) ? c->g.dx * 1.0 : 0.0;
      area += (j < ny - 1 && nodes[n + nx]) ? c->g.dy * 1.0 : 0.0;
    }
  /* findfirst runs in column-major order: smallest j, then smallest i */
  int ref = -1;
  for (int j = 0; j < ny && ref < 0; ++j)
    for (int i = 0; i < nx; ++i)
      if (nodes[i + nx * j]) {
        ref = i + nx * j;
        break;
      }
  if (ref < 0) {
    isk_set_error("electrode_create: empty node mask");
    return ISKRA_E_ARG;
  }
  /* create_electrode(nodes, config) creates a default tracker when none exists */
  if (!c->trk.on && (rc = tracker_create(c, ISK_SURF_ABSORB))) return rc;
  IskElectrode e;
  e.fixed = fixed ? 1 : 0;
  e.ref_node = ref;
  e.area = area;
  e.sigma0 = sigma;
  int code;
  if (fixed) {
    /* apply_dirichlet(ps, nodes, phi) under a tag of its own */
    int tag = 0;
    for (int t = 1; t <= ISK_MAX_TAGS && !tag; ++t) {
      bool used = false;
      for (int k = 0; k < nn && !used; ++k) used = p.tags[k] == t;
      for (int q = 0; q < c->nelec && !used; ++q) used = c->elec[q].fixed && c->elec[q].tag == t;
      if (!used) tag = t;
    }
    if (!tag) {
      isk_set_error("electrode_create: no free Dirichlet tag");
      return ISKRA_E_ARG;
    }
    for (int k = 0; k < nn; ++k)
      if (nodes[k]) {
        if ((int)p.h_srow.size() == nn && p.h_srow[k] >= 0) {
          isk_set_error("electrode_create: node %d already belongs to a floating electrode", k);
          return ISKRA_E_UNSUPPORTED;
        }
        p.tags[k] = (int8_t)tag;
      }
    p.tag_val[tag] = phi;
    p.wf_on[tag] = 0;
    e.tag = tag;
    code = ISK_SURF_ABSORB;
  } else {
    if (c->nfloat >= ISK_MAX_ELECTRODES) {
      isk_set_error("electrode_create: at most %d floating electrodes", ISK_MAX_ELECTRODES);
      return ISKRA_E_ARG;
    }
    if (!(area > 0.0)) {
      isk_set_error("electrode_create: a floating electrode needs a mask with an edge (area 0)");
      return ISKRA_E_ARG;
    }
    const int dof = c->nfloat;
    if ((rc = neumann_rows(c, nodes, dof))) return rc;
    if (!c->d_elec) {
      if (isk_alloc((void **)&c->d_elec, EL_WORDS * 8)) return ISKRA_E_HIP;
      ISK_HIP(hipMemset(c->d_elec, 0, EL_WORDS * 8));
    }
    double v[2] = {sigma, area};
    ISK_HIP(hipStreamSynchronize(c->stream));
    ISK_HIP(hipMemcpy(c->d_elec + EL_SIGMA + dof, &v[0], 8, hipMemcpyHostToDevice));
    ISK_HIP(hipMemcpy(c->d_elec + EL_AREA + dof, &v[1], 8, hipMemcpyHostToDevice));
    c->float_elec[dof] = c->nelec;
    c->nfloat = dof + 1;
    e.tag = dof; /* the sigma dof */
    code = ISK_SURF_ELECTRODE + dof;
  }
  c->elec[c->nelec] = e;
  surface_lookup(c, nodes, code);
  if ((rc = tracker_upload(c))) return rc;
  if ((rc = poisson_rebuild(c))) return rc;
  *electrode_id = c->nelec++;
  return ISKRA_OK;
}

int iskra_electrode_state(iskra_ctx *c, int id, double *sigma, double *phi, double *dq,
                          double *area) {
  if (!c || id < 0 || id >= c->nelec) {
    isk_set_error("electrode id %d out of range", id);
    return ISKRA_E_ARG;
  }
  const IskElectrode &e = c->elec[id];
  ISK_HIP(hipStreamSynchronize(c->stream));
  double s = 0.0, q = 0.0;
  if (!e.fixed) {
    ISK_HIP(hipMemcpy(&s, c->d_elec + EL_SIGMA + e.tag, 8, hipMemcpyDeviceToHost));
    ISK_HIP(hipMemcpy(&q, c->d_elec + EL_DQ + e.tag, 8, hipMemcpyDeviceToHost));
  }
  if (sigma) *sigma = s;
  if (dq) *dq = q;
  if (area) *area = e.area;
  if (phi) {
    if (e.fixed) *phi = c->ps.tag_val[e.tag];
    else ISK_HIP(hipMemcpy(phi, c->phi + e.ref_node, 8, hipMemcpyDeviceToHost));
  }
  return ISKRA_OK;
}

int iskra_circuit_set(iskra_ctx *c, double R, double L, double C, double v_offset,
                      double v_amplitude, double v_freq, double v_phase, int positive,
                      int negative) {
  if (!c) return ISKRA_E_ARG;
  if (positive >= c->nelec || negative >= c->nelec ||
      (positive >= 0 && c->elec[positive].fixed) || (negative >= 0 && !c->elec[negative].fixed) ||
      ((positive >= 0) != (negative >= 0))) {
    isk_set_error("circuit_set: the plasma device is a floating (positive) and a fixed "
                  "(negative) electrode, or none (-1, -1) for a shorted connection");
    return ISKRA_E_ARG;
  }
  if (!(L > 0.0 || R > 0.0)) {
    isk_set_error("circuit_set: L or R must be positive (advance_circuit! divides by L/dt + R/2)");
    return ISKRA_E_ARG;
  }
  if (!c->d_elec) {
    if (isk_alloc((void **)&c->d_elec, EL_WORDS * 8)) return ISKRA_E_HIP;
    ISK_HIP(hipMemset(c->d_elec, 0, EL_WORDS * 8));
  }
  c->circ_R = R;
  c->circ_L = L;
  c->circ_C = C;
  c->circ_v[0] = v_offset;
  c->circ_v[1] = v_amplitude;
  c->circ_v[2] = v_freq;
  c->circ_v[3] = v_phase / 6.283185307179586;
  c->circ_pos = positive;
  c->circ_neg = negative;
  c->circuit_on = 1;
  /* rlc(...) starts from i = q = t = 0 (Circuit.jl:58-66) */
  ISK_HIP(hipStreamSynchronize(c->stream));
  ISK_HIP(hipMemset(c->d_elec + EL_CIRC, 0, 4 * 8));
  return ISKRA_OK;
}

int iskra_circuit_state(iskra_ctx *c, double *i, double *q, double *t, double *vext) {
  if (!c || !c->circuit_on) {
    isk_set_error("circuit_state: no circuit");
    return ISKRA_E_STATE;
  }
  double s[4];
  ISK_HIP(hipStreamSynchronize(c->stream));
  ISK_HIP(hipMemcpy(s, c->d_elec + EL_CIRC, sizeof(s), hipMemcpyDeviceToHost));
  if (i) *i = s[0];
  if (q) *q = s[1];
  if (t) *t = s[2];
  if (vext) *vext = s[3];
  return ISKRA_OK;
}

struct ElArgs {
  unsigned long long *ehit[ISK_MAX_SPECIES];
  double qw[ISK_MAX_SPECIES];
  int nspecies, nfloat;
  double *el;
  /* circuit */
  int circuit, pos_ref, pos_dof;
  double neg_phi, R, L, C, v0, v1, vf, vph, dt;
  const double *phi;
};

/* hits -> charge (hit!(::FloatingPotentialElectrode), circuit_coupling.jl:47-56,
 * summed per species: dq = n q w0), then advance!(::CircuitRLC)
 * (circuit_coupling.jl:35-43) with advance_circuit! (Circuit.jl:117-136) */
__global__ void k_el_update(ElArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int k = 0; k < a.nfloat; ++k) {
    double dq = 0.0;
    for (int s = 0; s < a.nspecies; ++s) {
      const unsigned long long n = a.ehit[s][k];
      if (n) {
        dq += (double)n * a.qw[s];
        a.ehit[s][k] = 0ull;
      }
    }
    if (dq != 0.0) {
      a.el[EL_DQ + k] += dq;
      a.el[EL_SIGMA + k] += dq / a.el[EL_AREA + k];
    }
  }
  if (!a.circuit) return;
  double *cs = a.el + EL_CIRC;
  const double i = cs[0], q = cs[1], t = cs[2];
  const double vext = a.pos_ref >= 0 ? a.phi[a.pos_ref] - a.neg_phi : 0.0;
  double u = a.vf * t + a.vph;
  u = u - floor(u);
  double sn, cn;
  isk_sincos2pi(u, &sn, &cn);
  const double vt = a.v0 + a.v1 * sn;
  double ni = (a.L / a.dt - a.R / 2.0) * i + vext - vt;
  double nq = q;
  if (a.C > 0.0) {
    ni -= q / a.C;
    ni /= (a.L / a.dt + a.R / 2.0);
    nq = q + a.dt * i;
  } else {
    ni /= (a.L / a.dt + a.R / 2.0);
  }
  cs[0] = ni;
  cs[1] = nq;
  cs[2] = t + a.dt;
  cs[3] = vext;
  if (a.pos_dof >= 0) {
    /* foo!(::PlasmaDevice): positive.dq = 0, dsigma = -dt i / area; the
     * reference adds it to the first sigma dof (get_rhs(solver, :sigma, 1)) */
    a.el[EL_DQ + a.pos_dof] = 0.0;
    a.el[EL_SIGMA + 0] += -a.dt * ni / a.el[EL_AREA + a.pos_dof];
  }
}

static int el_launch(iskra_ctx *c, double dt, int with_circuit) {
  if (!c->d_elec) return ISKRA_OK;
  ElArgs a;
  memset(&a, 0, sizeof(a));
  a.nspecies = c->nspecies;
  for (int s = 0; s < c->nspecies; ++s) {
    a.ehit[s] = c->sp[s].ehit;
    a.qw[s] = c->sp[s].q * c->sp[s].w0;
  }
  a.nfloat = c->nfloat;
  a.el = c->d_elec;
  a.circuit = with_circuit && c->circuit_on;
  a.pos_ref = -1;
  a.pos_dof = -1;
  if (a.circuit && c->circ_pos >= 0) {
    a.pos_ref = c->elec[c->circ_pos].ref_node;
    a.pos_dof = c->elec[c->circ_pos].tag;
    a.neg_phi = c->ps.tag_val[c->elec[c->circ_neg].tag];
  }
  a.R = c->circ_R;
  a.L = c->circ_L;
  a.C = c->circ_C;
  a.v0 = c->circ_v[0];
  a.v1 = c->circ_v[1];
  a.vf = c->circ_v[2];
  a.vph = c->circ_v[3];
  a.dt = dt;
  a.phi = c->phi;
  hipLaunchKernelGGL(k_el_update, dim3(1), dim3(64), 0, c->stream, a);
  ISK_CHECK_LAUNCH();
  return ISKRA_OK;
}

int isk_electrodes_after_push(iskra_ctx *c, double dt) { return el_launch(c, dt, dt > 0.0); }

void isk_electrodes_release(iskra_ctx *c) {
  isk_free(c->trk.surf);
  isk_free(c->trk.tracked);
  c->trk.surf = c->trk.tracked = nullptr;
  c->trk.h_surf.clear();
  c->trk.on = 0;
  isk_free(c->d_elec);
  c->d_elec = nullptr;
  c->nelec = 0;
  c->nfloat = 0;
  c->circuit_on = 0;
  c->circ_pos = c->circ_neg = -1;
  c->ps.h_srow.clear();
  c->ps.h_sd.clear();
  c->ps.h_sc.clear();
  c->ps.h_ssig.clear();
  c->ps.h_snb.clear();
  c->ps.h_sdof.clear();
}

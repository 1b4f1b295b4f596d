/*
 * iskra.h -- C ABI of libiskra_hip.so, the MI355X (gfx950) electrostatic
 * particle-in-cell step: gather + push + boundary + CIC deposit, counting sort
 * by cell, finite-difference Poisson solve, E = -grad(phi), null-collision MCC.
 *
 * Every entry point returns an int status (ISKRA_OK == 0); the message of the
 * last failure on the calling thread is iskra_last_error().  No C++ exception
 * crosses this boundary.  Host arrays passed in are only read/written during the
 * call.  Arrays use the reference's layouts: particle positions x are np x D
 * column-major FP64 (ParticleInCell/src/pic/kinetic.jl:2), velocities np x 3
 * column-major FP64 (kinetic.jl:3), node fields nx x ny column-major FP64
 * (FiniteDifferenceMethod/src/generalized_poisson.jl:40), node indices 0-based.
 *
 * A context owns one HIP stream and the device state of one simulation; it is
 * not thread-safe (one host thread per context).
 *
 * Reference interfaces each entry point replaces are cited per declaration.
 */
#ifndef ISKRA_H
#define ISKRA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum iskra_status {
  ISKRA_OK = 0,
  ISKRA_E_ARG = 1,         /* invalid argument */
  ISKRA_E_HIP = 2,         /* HIP runtime failure */
  ISKRA_E_CAPACITY = 3,    /* particle capacity exceeded (sample!, ionisation) */
  ISKRA_E_PROB = 4,        /* MCC probability assertion (Chemistry/src/mcc.jl:244-246,273-279) */
  ISKRA_E_STATE = 5,       /* call out of order (no grid / solver / species) */
  ISKRA_E_CONVERGE = 6,    /* Poisson solve did not reach its tolerance */
  ISKRA_E_UNSUPPORTED = 7, /* configuration this build does not implement */
  ISKRA_E_IO = 8           /* file output failed (the openPMD writer, iskra_pmd.h) */
};

typedef struct iskra_ctx iskra_ctx;

/* Message of the last failing call on this thread ("" if none). */
const char *iskra_last_error(void);
/* Library ABI version (major*100 + minor). */
int iskra_version(void);

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart: the reference runs on the host only
 * (ParticleInCell/src/ParticleInCell.jl:84-139 owns all arrays). */
int iskra_ctx_create(int device, iskra_ctx **out);
int iskra_ctx_destroy(iskra_ctx *ctx);
/* Wait for all work queued on the context's stream.  Reports errors the
 * device recorded meanwhile: particle-capacity overflow (ISKRA_E_CAPACITY),
 * a collision probability above 1 (ISKRA_E_PROB), and a spectral Poisson solve inside iskra_step that missed rtol after its
 * correction pass (ISKRA_E_CONVERGE; the flag is cleared once reported). */
int iskra_sync(iskra_ctx *ctx);

/* ---- grid --------------------------------------------------------------- */
/* RegularGrids.create_uniform_grid (RegularGrids/src/RegularGrids.jl:55-82).
 * dim 1 or 2; nodes[d] node counts (cells = nodes-1); dh[d] spacing;
 * origin[d]; vol_periodic[4] = {left,right,bottom,top} edges declared
 * :periodic, which disables the edge-volume halving of cell_volume
 * (RegularGrids.jl:26-38).  Resets species, solver and fields. */
int iskra_grid_set(iskra_ctx *ctx, int dim, const int64_t nodes[2],
                   const double dh[2], const double origin[2],
                   const int32_t vol_periodic[4]);
/* RegularGrids.create_axial_grid (RegularGrids/src/RegularGrids.jl:84-97):
 * a 2-D (r, z) grid, axis 0 radial with the symmetry axis at node i = 0
 * (r = i dr), axis 1 axial.  Node volumes are the annuli of
 * cell_volume(::AxialGrid{2}) (RegularGrids.jl:40-53), halved on a z edge
 * unless vol_periodic_z[0] (bottom) / [1] (top) declares it :periodic.
 * Particles, deposit and gather work in (r, z) exactly as in (x, y)
 * (cloud_in_cell.jl:38-73); iskra_poisson_set assembles the axisymmetric
 * operator of generalized_poisson.jl:70-199.  Resets species, solver and
 * fields, and selects the RZ pusher. */
int iskra_grid_set_axial(iskra_ctx *ctx, const int64_t nodes[2],
                         const double dh[2], const double origin[2],
                         const int32_t vol_periodic_z[2]);
/* Boris pusher geometry (ParticleInCell/src/pic/pushers.jl:4-17): XY is
 * create_boris_pusher (push_in_cartesian!), RZ is create_axial_boris_pusher
 * (push_in_cartesian! followed by transform_from_cartesian_to_cylindrical!,
 * pushers.jl:52-66: the particle moves in the plane through its position
 * and the axis, then is rotated back into the (r, z) half plane together with
 * (v_r, v_theta) = (v_x, v_z)).  RZ needs a 2-D grid. */
enum { ISKRA_PUSHER_XY = 0, ISKRA_PUSHER_RZ = 1 };
int iskra_pusher_set(iskra_ctx *ctx, int kind);

/* Work-decomposition parameters (the autotuner's knobs).  tile: cells per
 * tile edge (power of two); chunk: particles per work-group; margin: window
 * halo in cells; sort_every: steps between counting sorts inside
 * iskra_step.  Zero keeps the current value. */
typedef struct {
  int32_t tile;
  int32_t chunk;
  int32_t margin;
  int32_t sort_every;
} iskra_tuning;
int iskra_tuning_set(iskra_ctx *ctx, const iskra_tuning *t);
int iskra_tuning_get(iskra_ctx *ctx, iskra_tuning *t);
/* Steps between counting sorts of one species inside iskra_step (0: the
 * context's sort_every).  Slow species (heavy, cold ions) stay tile-ordered
 * for many more steps than electrons.  No reference counterpart: the
 * device-side particle order is this library's (the reference keeps its
 * particles unsorted).  Batches need every species on the context value. */
int iskra_species_sort_every(iskra_ctx *ctx, int sid, int32_t every);
/* Deposit-window margin of one species in cells (-1: the context's margin;
 * values above it are clipped to it).  A chunk's window spans its tile plus
 * the margin, so a species that moves well under a cell between sorts
 * (cold ions) needs margin 0 and its push and merge move less window data.
 * No reference counterpart (the windows are this library's). */
int iskra_species_margin(iskra_ctx *ctx, int sid, int32_t margin);

/* ---- species ------------------------------------------------------------ */
/* KineticSpecies{D,3} + create_kinetic_species
 * (ParticleInCell/src/pic/kinetic.jl:1-18, problem/configuration.jl:95-102).
 * q, m in SI; w0 = weight of every macro-particle (the reference's wg, which
 * only ever holds w0); capacity = maximum particle count; track_ids != 0 keeps
 * a uint32 id per particle (needed by MCC, whose random draws are keyed by
 * id). */
int iskra_species_add(iskra_ctx *ctx, double q, double m, double w0,
                      int64_t capacity, int track_ids, int *sid);
/* Boundary applied after the push (the after_push hook:
 * ParticleInCell/src/ParticleInCell.jl:41, wrap!/discard!
 * ParticleInCell/src/pic/surfaces/wrap.jl:1-33), one mode per dimension. */
enum { ISKRA_BND_NONE = 0, ISKRA_BND_WRAP = 1, ISKRA_BND_DISCARD = 2 };
int iskra_species_set_boundary(iskra_ctx *ctx, int sid, const int32_t mode[2]);
/* Replace the species' particles by np host particles (x: np x D col-major,
 * v: np x 3 col-major, id: np or NULL -> ids 0..np-1).  Positions must lie in
 * the grid.  Leaves the species cell-sorted. */
int iskra_species_upload(iskra_ctx *ctx, int sid, int64_t np, const double *x,
                         const double *v, const uint32_t *id);
/* Copy the species out in the current device order (x: cap x D col-major with
 * leading dimension cap, v: cap x 3, id may be NULL). */
int iskra_species_download(iskra_ctx *ctx, int sid, int64_t cap, int64_t *np,
                           double *x, double *v, uint32_t *id);
int iskra_species_count(iskra_ctx *ctx, int sid, int64_t *np);
/* Raw device state in the library's own format (checkpoint / resume; no
 * reference counterpart: the reference has no checkpoint, TODO:3):
 * cell = i | (j << 16) with (i, j) the 0-based lower-left node, -1 for a
 * discarded particle; ox, oy in [0,1) the in-cell offsets (oy ignored in
 * 1-D); vx, vy, vz in m/s; id may be NULL.  get copies the current device
 * order (discarded particles included) and sets *np (all arrays NULL: size
 * query only); set replaces the
 * particles and leaves them cell-sorted (stable). */
int iskra_species_get_state(iskra_ctx *ctx, int sid, int64_t cap, int64_t *np,
                            int32_t *cell, float *ox, float *oy, float *vx,
                            float *vy, float *vz, uint32_t *id);
int iskra_species_set_state(iskra_ctx *ctx, int sid, int64_t np,
                            const int32_t *cell, const float *ox,
                            const float *oy, const float *vx, const float *vy,
                            const float *vz, const uint32_t *id);
/* sample!(::MaxwellianSource) (ParticleInCell/src/pic/sources.jl:24-34):
 * append n = min(free capacity, n) particles with x = U*wx + dx and
 * v = N(0,1)*wv + dv per component, drawn on the device from Philox4x32-10
 * keyed by (seed, stream).  wx, dx: D values; wv, dv: 3 values. */
int iskra_species_sample_maxwellian(iskra_ctx *ctx, int sid, int64_t n,
                                    const double *wx, const double *dx,
                                    const double *wv, const double *dv,
                                    uint64_t seed, uint32_t stream);

/* ---- step operators ----------------------------------------------------- */
/* Counting sort by cell (no reference counterpart; closest is the DSMC cell
 * binning Chemistry/src/dsmc.jl:25-30).  exact_cell = 0 orders by cell tile
 * (stable), 1 by exact cell; discarded particles are compacted away. */
int iskra_sort(iskra_ctx *ctx, int sid, int exact_cell);
/* grid_to_particle (ParticleInCell/src/pic/cloud_in_cell.jl:20-36) of the
 * current E at the current positions; E_out: cap x 3 col-major. */
int iskra_gather(iskra_ctx *ctx, int sid, int64_t cap, double *E_out);
/* One advance! of a kinetic species (ParticleInCell.jl:51-61): gather E,
 * Boris push with B = 0 (pushers.jl:37-50), after_push boundary (wrap.jl),
 * fused with this species' CIC deposit at the new positions
 * (cloud_in_cell.jl:1-18).  dt in seconds. */
int iskra_push(iskra_ctx *ctx, int sid, double dt);
/* CIC deposit of the current positions without pushing (density(),
 * kinetic.jl:53). */
int iskra_deposit_species(iskra_ctx *ctx, int sid);
/* Merge all species' deposits: n_s = deposit_s / cell_volume and
 * rho = sum_s q_s n_s + background (ParticleInCell.jl:118-124; fluid
 * species enter as q*n without volume division, pic/fluid.jl:11). */
int iskra_deposit(iskra_ctx *ctx);

/* ---- fields ------------------------------------------------------------- */
enum {
  ISKRA_F_RHO = 0,      /* charge density rho (C/m^D), FP64 */
  ISKRA_F_PHI = 1,      /* potential phi, FP64 */
  ISKRA_F_EX = 2,       /* E_x at nodes (stored FP32) */
  ISKRA_F_EY = 3,       /* E_y at nodes (stored FP32) */
  ISKRA_F_BG = 4,       /* fixed background charge density (fluid species) */
  ISKRA_F_DENSITY = 5,  /* number density of species `index` (download only) */
  ISKRA_F_TARGET = 6,   /* MCC target gas density of interaction `index` */
  ISKRA_F_FLUID_N = 7,  /* density n of mobile fluid `index` (iskra_fluid_add) */
  ISKRA_F_FLUID_DN = 8  /* its change dn in the last advance (download only) */
};
int iskra_field_download(iskra_ctx *ctx, int which, int index, double *out);
int iskra_field_upload(iskra_ctx *ctx, int which, int index, const double *in);

/* ---- Poisson ------------------------------------------------------------ */
/* create_poisson_solver + apply_periodic + apply_dirichlet
 * (FiniteDifferenceMethod/src/generalized_poisson.jl:22-68,205-215,271-324).
 * tags: one int8 per node, 0 = free, t > 0 = Dirichlet node whose value is
 * tag t's value (each apply_dirichlet mask is one tag); periodic_x/y link the
 * first and last node of that axis as the reference's ring.
 * kind: DIRECT (1-D tridiagonal), PCG (Jacobi-preconditioned CG), MG
 * (geometric multigrid V-cycles, 2-D with Dirichlet walls), MGPCG (CG
 * preconditioned by one V-cycle), SPECTRAL (2-D direct solve by the discrete
 * sine transform, which diagonalises the operator when all four walls are
 * Dirichlet and no interior node is: FP64 dense transforms, then the FP64
 * residual is checked and corrected until it meets rtol).  eps0 in F/m; rtol on ||b-A phi||2/||b||2
 * evaluated in FP64.
 * On an axial grid (iskra_grid_set_axial) the operator is the reference's
 * axisymmetric one (generalized_poisson.jl:70-199), which becomes symmetric
 * once row i is multiplied by w_i = i (w_0 = 1/2 on the axis, w_{nr-1} =
 * nr - 3/2 on the outer wall): radial links of weight (i + 1/2)/dr^2 between
 * i and i+1, axial links of weight w_i/dz^2.  Only PCG solves it (on that
 * weighted system, relres measured there); it needs a Dirichlet node and no
 * periodic link (the reference's apply_periodic for RZ cannot run:
 * generalized_poisson.jl:326-361 reads undefined nr, nz, dr, dz). */
enum {
  ISKRA_POISSON_DIRECT = 0,
  ISKRA_POISSON_PCG = 1,
  ISKRA_POISSON_MG = 2,
  ISKRA_POISSON_MGPCG = 3,
  ISKRA_POISSON_SPECTRAL = 4
};
int iskra_poisson_set(iskra_ctx *ctx, int kind, double eps0, const int8_t *tags,
                      int periodic_x, int periodic_y, double rtol,
                      int max_iter);
/* apply_dirichlet(ps, mask, phi0) value update of an existing tag. */
int iskra_poisson_set_tag_value(iskra_ctx *ctx, int tag, double value);
/* value(t) = offset + amplitude*sin(2 pi freq t + phase), set inside
 * iskra_step after each step's solve with t = (it-1)*dt, exactly like the
 * after_loop hook of problem/11_rf_discharge.jl:95 (one-step lag). */
int iskra_poisson_set_tag_waveform(iskra_ctx *ctx, int tag, double offset,
                                   double amplitude, double freq, double phase);
/* calculate_electric_potential(ps, -rho) (generalized_poisson.jl:372-378).
 * warm: ZERO = start from phi = 0, PREV = start from the current phi field
 * (previous solve or an uploaded warm start). */
enum { ISKRA_WARM_ZERO = 0, ISKRA_WARM_PREV = 1 };
int iskra_poisson_solve(iskra_ctx *ctx, int warm, int *iters, double *relres);
/* calculate_electric_field (generalized_poisson.jl:385-410). */
int iskra_efield(iskra_ctx *ctx);

/* ---- Monte Carlo collisions -------------------------------------------- */
/* mcc(@reactions ...) -> MonteCarloCollisions
 * (Chemistry/src/mcc.jl:18-51,291-320).  n_coll processes in reaction order;
 * types[k] one of ISKRA_COLL_*, energy[k] threshold in eV (excitation /
 * ionisation); cross-section tables (CrossSection, cross_section.jl:3-17):
 * process k owns nodes [tab_off[k], tab_off[k+1]) of tab_eps (eV) and
 * tab_sigma (m^2).  Products of ionisation: electron -> the source species,
 * ion -> species ion_sid (round(w0_src/w0_ion) ions each, mcc.jl:205-211).
 * Target: fluid species with mass, charge, temperature; its density per node
 * is the ISKRA_F_TARGET field of this interaction (default 0). */
enum {
  ISKRA_COLL_ELASTIC_ISOTROPIC = 0,
  ISKRA_COLL_ELASTIC_BACKWARD = 1,
  ISKRA_COLL_EXCITATION = 2,
  ISKRA_COLL_IONIZATION = 3,
  ISKRA_COLL_INELASTIC_BACKWARD = 4 /* MCC.InelasticBackward (mcc.jl:144-157) */
};
int iskra_mcc_add(iskra_ctx *ctx, int src_sid, int n_coll, const int32_t *types,
                  const double *energy, const int32_t *tab_off,
                  const double *tab_eps, const double *tab_sigma, int ion_sid,
                  double target_m, double target_q, double target_T,
                  int *mcc_id);
/* PIC.perform!(::MonteCarloCollisions, E, dt, config) (mcc.jl:231-289) with a
 * per-particle null-collision rule; events[k] (nullable) receives this
 * call's accepted events per process. */
int iskra_mcc_run(iskra_ctx *ctx, int mcc_id, double dt, int64_t *events);
/* Per-node event counts nu[node][k] accumulated since the last reset
 * (mcc.jl:283-288), out: nodes x n_coll col-major. */
int iskra_mcc_counts(iskra_ctx *ctx, int mcc_id, int reset, int64_t *out);
/* The same counts for the last step of the latest iskra_step call only (the
 * reference's nu, which perform! recomputes every step): iskra_step sets
 * the earlier steps' counts apart before it runs its last step.  Events of
 * iskra_mcc_run calls made since then are included; batch steps do not set
 * counts apart. */
int iskra_mcc_counts_last(iskra_ctx *ctx, int mcc_id, int64_t *out);

/* ---- direct simulation Monte Carlo ------------------------------------- */
/* dsmc(@reactions ...) with one DSMC.ElasticCollision
 * (Chemistry/src/dsmc.jl:1-21,144-167) between two kinetic species of equal
 * weight (src_sid may equal tgt_sid): the rate is a CrossSection over the
 * relative speed g (cross_section.jl:3-17): n_tab nodes tab_g (m/s,
 * increasing) and tab_sigma (m^2).  2-D grids only (the reference's
 * perform! destructures a 2-D grid).  ISKRA_E_UNSUPPORTED for unequal
 * weights: the reference's update for them writes the target velocity into
 * the particle with the source's index (dsmc.jl:71-77).  One reaction per
 * interaction: add one interaction per reaction (the reference's lists are
 * shared between the reactions of one interaction, dsmc.jl:94-96). */
int iskra_dsmc_add(iskra_ctx *ctx, int src_sid, int tgt_sid, int n_tab,
                   const double *tab_g, const double *tab_sigma, int *dsmc_id);
/* PIC.perform!(::DirectSimulationMonteCarlo, E, dt, config)
 * (dsmc.jl:81-142): both species are sorted by exact cell (their device
 * order changes), then every cell runs its floor(Nc) pairing trials; the
 * fractional remainder of Nc is carried per cell to the next call.
 * events (nullable) receives this call's accepted collisions.  iskra_step
 * runs every DSMC interaction after the MCC ones, before the push. */
int iskra_dsmc_run(iskra_ctx *ctx, int dsmc_id, double dt, int64_t *events);
/* Accepted collisions per cell since the last reset, at the cell's lower-left
 * node (the reference's nuDSMC field, dsmc.jl:124-141): out has one entry
 * per node (nodes x 1, col-major). */
int iskra_dsmc_counts(iskra_ctx *ctx, int dsmc_id, int reset, int64_t *out);
/* The counts of the last step of the latest iskra_step call (see
 * iskra_mcc_counts_last; dsmc.jl:124-141). */
int iskra_dsmc_counts_last(iskra_ctx *ctx, int dsmc_id, int64_t *out);

/* ---- surfaces, electrodes and the external circuit --------------------- */
/* Surfaces of the cell-edge tracker (ParticleInCell/src/pic/surfaces/
 * track.jl:9-25, hit.jl:30-56): absorbing (the particle is removed),
 * reflective (the velocity component along the edge normal flips and the
 * rest of the step is mirrored), periodic (a hit does nothing: the particle
 * continues and the after_push boundary applies). */
enum { ISKRA_SURF_ABSORBING = 1, ISKRA_SURF_REFLECTIVE = 2, ISKRA_SURF_PERIODIC = 3 };
/* create_surface_tracker(grid, ds) (track.jl:101-114): every edge of the grid
 * boundary gets default_surface.  From then on every push of a 2-D Cartesian
 * grid runs track! + check! (ParticleInCell.jl:55-60; check.jl:17-68): a
 * particle that starts the step in a cell next to a surface edge follows its
 * straight path cell by cell (FP64, the reference's arithmetic with the grid
 * spacing dx along both axes as check uses it) and meets every surface edge
 * it crosses.  Absorbed particles are discarded (removed at the next sort,
 * like discard!); the reference's "too fast" message is not printed. */
int iskra_tracker_create(iskra_ctx *ctx, int default_surface);
/* track_surface!(tracker, nodes, surface) (track.jl:116-118): every cell edge
 * whose two end nodes are in the mask (one uint8 per node, x fastest) gets
 * the surface (build_surface_lookup!, track.jl:45-60; later calls override). */
int iskra_tracker_surface(iskra_ctx *ctx, const uint8_t *nodes, int surface);
/* The tracker as the push kernels read it (no reference counterpart: the
 * reference keeps a Dict of cell pairs): per cell (x fastest) four codes for
 * the edges -x, +x, -y, +y (0 none, 1 absorbing or fixed electrode,
 * 2 reflective, 3 periodic, 16 + k floating electrode with sigma dof k) and
 * one tracked flag (the cell appears in a surface entry on either side). */
int iskra_tracker_table(iskra_ctx *ctx, uint8_t *codes, uint8_t *tracked);
/* create_electrode(nodes, config; fixed, sigma, phi) (problem/configuration.jl
 * :17-76; circuit_coupling.jl:4-15,45-62).  Needs the Poisson solver set
 * (iskra_poisson_set first; setting it again removes the electrodes) and
 * creates the default absorbing tracker if none exists.  fixed != 0:
 * apply_dirichlet(nodes, phi) under a Dirichlet tag of its own, and an
 * absorbing surface on the mask's edges.  fixed == 0 (floating): a surface
 * charge sigma (C/m^2, initial value `sigma`) and the rows apply_neumann
 * writes for the mask (generalized_poisson.jl:222-269; the system is then
 * not symmetric and is solved by BiCGStab, PCG kind only); every particle
 * its edges absorb adds q w0 to the electrode's charge and q w0 / area to
 * sigma (area from calculate_area, dz = 1). */
int iskra_electrode_create(iskra_ctx *ctx, const uint8_t *nodes, int fixed, double sigma,
                           double phi, int *electrode_id);
/* sigma and collected charge (floating; 0 for fixed), phi (fixed: its value;
 * floating: the potential at its reference node, the first mask node in
 * column-major order), area.  Any argument may be NULL. */
int iskra_electrode_state(iskra_ctx *ctx, int electrode_id, double *sigma, double *phi,
                          double *dq, double *area);
/* rlc(@netlist ...) (Circuit/src/Circuit.jl:58-66) with V(t) = v_offset +
 * v_amplitude sin(2 pi v_freq t + v_phase) and, as external device, the
 * PlasmaDevice of a floating (positive) and a fixed (negative) electrode
 * (circuit_coupling.jl:17-33), or none (-1, -1: ShortedConnection, zero
 * voltage).  Each iskra_step then runs advance!(::CircuitRLC)
 * (circuit_coupling.jl:35-43) after the push: advance_circuit! (Circuit.jl
 * :117-136) with the external voltage phi(positive) - phi(negative) of the
 * previous solve, and sigma of the first floating electrode (the reference's
 * get_rhs(solver, :sigma, 1)) += -dt i / area(positive).  i = q = t = 0 at
 * the call. */
int iskra_circuit_set(iskra_ctx *ctx, double R, double L, double C, double v_offset,
                      double v_amplitude, double v_freq, double v_phase, int positive,
                      int negative);
/* the circuit's current, capacitor charge, time and last external voltage */
int iskra_circuit_state(iskra_ctx *ctx, double *i, double *q, double *t, double *vext);

/* ---- fluid species --------------------------------------------------------- */
/* A mobile FluidSpecies (pic/fluid.jl:1-11, problem/configuration.jl:104-108):
 * charge q, mass m, mobility mu (the diffusion coefficient of
 * calculate_advection_diffusion) and node densities n (nodes, x fastest).
 * Its charge q n enters rho with the background (ISKRA_F_BG, which then
 * names the fixed part of the background only).  iskra_step advances every
 * fluid once per step before the deposit merge, with the E the step starts
 * from: advance!(::FluidSpecies) (ParticleInCell.jl:74-82), v = (E dt)(q/m),
 * dn = calculate_advection_diffusion(n, mu, v, dh, dt)
 * (FiniteDifferenceMethod/src/advection_diffusion.jl:4-38), n += dn.  A
 * fluid whose density never changes can stay in ISKRA_F_BG instead.  Not
 * available in batches. */
int iskra_fluid_add(iskra_ctx *ctx, double q, double m, double mu, const double *n, int *fid);
/* The same update of every fluid once, outside the step (advance! of the
 * fluids alone, with the current E). */
int iskra_fluid_advance(iskra_ctx *ctx, double dt);

/* ---- checkpoint / resume -------------------------------------------------- */
/* No reference counterpart (the reference has no checkpoint, TODO:3).  The
 * dynamic state of a context as one host blob: the live particles of every
 * species with their next id, the step counter and seed, phi and E, the
 * draw counters and event counts of the MCC and DSMC interactions and the
 * DSMC remainders, fluid densities, electrode and circuit state and the
 * electrode values.  Load it into a context set up like the saved one (same
 * grid, species, interactions, fluids, electrodes, solver and tuning; a
 * mismatch is ISKRA_E_STATE): the run then continues bit for bit as it would
 * have without the interruption (particles re-sorted by cell; results do
 * not depend on their order).  size gives the bytes save writes. */
int iskra_checkpoint_size(iskra_ctx *ctx, int64_t *bytes);
int iskra_checkpoint_save(iskra_ctx *ctx, void *buf, int64_t bytes);
int iskra_checkpoint_load(iskra_ctx *ctx, const void *buf, int64_t bytes);

/* ---- random numbers ------------------------------------------------------ */
/* The seed of all device draws (Random.seed!, src/iskra.jl:22-26). */
int iskra_rng_set(iskra_ctx *ctx, uint64_t seed);
/* Export the device's Philox4x32-10 draws: for each id, blocks
 * [0, nblocks) of counter (id, step, stream, block) under key = seed, 4 uint32
 * words per block -> out[n][nblocks][4].  The CPU oracle consumes these. */
int iskra_rng_export(iskra_ctx *ctx, uint32_t stream, uint32_t step,
                     const uint32_t *ids, int64_t n, int nblocks, uint32_t *out);
/* iskra_rng_export with no device allocation.  No counterpart in the
 * reference (which has no device memory): it is for callers that cannot
 * allocate device memory -- code under stream or graph capture, tight
 * serving or test loops, harnesses that forbid hipMalloc.  Same arguments
 * and the same draws as iskra_rng_export; the ids and draws are staged in a
 * fixed per-GPU scratch array of ISKRA_NOALLOC_SCRATCH_BYTES, held by each
 * GPU the library runs on until the process ends, so
 * round_up(4 n, 256) + 16 n nblocks must not exceed it (ISKRA_E_ARG
 * otherwise, nothing launched).  Every context on one GPU shares that array,
 * which keeps the last call's contents: calls that overlap on one GPU (other
 * host threads or contexts) race. */
#define ISKRA_NOALLOC_SCRATCH_BYTES (4 << 20)
int iskra_rng_export_noalloc(iskra_ctx *ctx, uint32_t stream, uint32_t step,
                             const uint32_t *ids, int64_t n, int nblocks,
                             uint32_t *out);

/* ---- the time loop ------------------------------------------------------- */
/* nsteps iterations of ParticleInCell.solve's loop body
 * (ParticleInCell.jl:102-135): MCC interactions, advance! of every species,
 * deposit, Poisson, E; tag waveforms are applied after each solve.  With
 * the spectral solver no step reads anything back; the call waits for its
 * stream once, at its end (iskra_sync), which reports device-side failures. */
int iskra_step(iskra_ctx *ctx, double dt, int64_t nsteps);

/* Kernel classes timed when profiling is on (HIP events recorded on the
 * context's stream around each launch). */
enum {
  ISKRA_K_PUSH = 0,    /* fused gather + push + boundary + deposit */
  ISKRA_K_SORT = 1,    /* counting sort (all passes) */
  ISKRA_K_POISSON = 2, /* one Poisson solve (all kernels) */
  ISKRA_K_MCC = 3,     /* collisions + product append */
  ISKRA_K_MERGE = 4,   /* window merge + rho */
  ISKRA_K_EFIELD = 5,
  ISKRA_K_WARMSTART = 6, /* learned warm-start inference */
  ISKRA_K_COUNT = 8
};
typedef struct {
  int64_t steps;            /* steps taken by iskra_step */
  int64_t particle_steps;   /* sum over steps and species of live particles pushed */
  int64_t poisson_solves;
  int64_t poisson_iters;    /* summed over solves */
  int32_t last_iters;
  int32_t pad;
  double last_relres;
  double kernel_ms[ISKRA_K_COUNT];     /* summed durations while profiling */
  int64_t kernel_calls[ISKRA_K_COUNT]; /* launches timed */
  int64_t push_particles;   /* particles handled by the timed push launches */
} iskra_stats;
int iskra_stats_get(iskra_ctx *ctx, iskra_stats *out);
/* Turn per-kernel timing on (1) or off (0); turning it on clears the totals.
 * No reference counterpart (the reference only @time's the whole solve,
 * src/iskra.jl:51). */
int iskra_profile(iskra_ctx *ctx, int enable);

/* ---- batched ensembles of 1-D discharges --------------------------------- */
/* No reference counterpart: the reference runs one simulation per process
 * (ParticleInCell/src/ParticleInCell.jl:84-139).  A batch groups n contexts
 * set up through the calls above (each a 1-D grid with the DIRECT Poisson
 * solve; all with the same number of species and interactions, the same
 * species roles in each interaction, the same sort interval and step count).
 * iskra_batch_step advances all of them by nsteps iterations of iskra_step:
 * each kernel type of the step is one launch covering every instance's
 * work-groups, each instance computing exactly what iskra_step computes for
 * it alone (particles, fields and collision counts are bit for bit those of
 * the same context stepped alone with the same tuning).  Cycles of
 * 2 x sort_every steps are captured in a hipGraph and replayed
 * (iskra_batch_set_graph(b, 0) launches step by step instead).  The contexts
 * stay usable through every other call between batch steps; a context must
 * not be destroyed while its batch exists. */
typedef struct iskra_batch iskra_batch;
int iskra_batch_create(iskra_ctx **ctxs, int n, iskra_batch **batch);
int iskra_batch_destroy(iskra_batch *batch);
int iskra_batch_set_graph(iskra_batch *batch, int enable);
/* ISKRA_E_UNSUPPORTED for a one-member batch when a rocprofiler-sdk tool is
 * loaded (ROCP_TOOL_LIBRARIES set): that case ended once in a segmentation
 * fault inside the HIP runtime; profile the context's iskra_step instead. */
int iskra_batch_step(iskra_batch *batch, double dt, int64_t nsteps);
/* graph replays issued so far (each one 2 x sort_every steps) */
int iskra_batch_graphs_launched(iskra_batch *batch, int64_t *count);

/* ---- learned warm start for the 2-D Poisson solve ----------------------- */
/* No reference counterpart: the reference solves A\b directly every step
 * (FiniteDifferenceMethod/src/generalized_poisson.jl:201-203,372-378), so it
 * has no iterative solve to seed.  A convolutional encoder-decoder maps the
 * right-hand side f = dx^2 rho / eps0 of a square 2-D grid of (side+1)^2
 * nodes with phi = 0 on all four walls to phi at the interior nodes; the
 * iterative solver then starts from that phi (ISKRA_WARM_PREV) and iterates
 * to its usual tolerance.  levels encoder/decoder levels with base*2^min(l,3)
 * channels, 3x3 convolutions, BF16 activations and weights on MFMA with FP32
 * accumulation, FP32 master weights trained with Adam on the device.
 * Parameters are float32, layer by layer in the order enc0a, enc0b, ...,
 * enc(L-1)b, dec(L-2)a, dec(L-2)b, ..., dec0b, out; each layer [cout][3x3][cin]
 * (cin of enc0a is 8: channel 0 carries the input; cout of out is 16:
 * channel 0 is the output).  The network is bound to ctx (its stream and,
 * for capture/predict, its grid and fields). */
typedef struct iskra_ws iskra_ws;
int iskra_ws_create(iskra_ctx *ctx, int levels, int base_channels, int side,
                    int max_batch, iskra_ws **net);
int iskra_ws_destroy(iskra_ws *net);
int iskra_ws_param_count(iskra_ws *net, int64_t *n);
int iskra_ws_set_params(iskra_ws *net, const float *params);
int iskra_ws_get_params(iskra_ws *net, float *params);
/* phi = output * gain * rms(f) (the network sees f / rms(f)) */
int iskra_ws_set_gain(iskra_ws *net, double gain);
/* Training objective of backward/train/eval:
 *   ISKRA_WS_PHI_MSE (default): mean squared error of the output against the
 *     target phi / (rms(f) gain) over the interior (side-1)^2 pixels (the
 *     warm-start objective; omega and sweeps unused);
 *   ISKRA_WS_ENERGY: the network as the coarse solver of a two-level cycle
 *     (iskra_ws_solve's preconditioner): k = sweeps damped Jacobi sweeps
 *     from 0, the network on the residual they leave, k sweeps again.  With
 *     A v = 4 v - the 4 neighbours on the interior pixels (0 outside) and
 *     P = (I - omega/4 A)^k, each sample (x, t) becomes (P x, P t) (what
 *     the pre-sweeps leave and its exact correction), the network sees
 *     bf16(P x), y_0 = gain * output, y_{j+1} = y_j + omega/4 (P x - A y_j)
 *     for j < k, e = y_k - gain P t (= the cycle's error), and the loss is
 *     e.Ae / (gain P t).A(gain P t) summed over the batch: the squared
 *     A-norm of the error relative to the correction's, the quantity that
 *     sets flexible CG's rate.  omega in [0, 2), sweeps in [0, 16].
 *   ISKRA_WS_ENERGY_NET: the same pairs (P x, P t), scored on the network's
 *     own output: e = gain * output - gain P t, no post-sweeps in the loss,
 *     so rough error the post-sweeps would hide is charged in full. */
enum { ISKRA_WS_PHI_MSE = 0, ISKRA_WS_ENERGY = 1, ISKRA_WS_ENERGY_NET = 2 };
int iskra_ws_set_objective(iskra_ws *net, int objective, double omega,
                           int sweeps);
/* Host batches (side x side planes per sample, already normalised input):
 * forward -> out; backward: forward, the objective's loss over the interior
 * (side-1)^2 pixels against target, gradients of all parameters -> grads. */
int iskra_ws_forward(iskra_ws *net, int batch, const float *in, float *out);
int iskra_ws_backward(iskra_ws *net, int batch, const float *in,
                      const float *target, float *grads, double *loss);
/* Device-resident training set: reserve n samples; capture appends the
 * context's current (f / rms(f), phi / rms(f)) (call after a solve). */
int iskra_ws_dataset_reserve(iskra_ws *net, int64_t n);
int iskra_ws_capture(iskra_ws *net, int64_t *index);
int iskra_ws_dataset_size(iskra_ws *net, int64_t *n);
int iskra_ws_dataset_rms(iskra_ws *net, int64_t first, int64_t count,
                         double *rms);
/* steps Adam steps on random batches of samples [first, first+count)
 * (targets divided by gain), batch draws from Philox4x32-10 keyed by seed;
 * losses (nullable) receives each step's batch loss.  eval: mean loss. */
int iskra_ws_train(iskra_ws *net, int64_t first, int64_t count, int batch,
                   int steps, float lr, uint64_t seed, double *losses);
int iskra_ws_eval(iskra_ws *net, int64_t first, int64_t count, double *mse);
/* phi at the interior nodes of the context <- prediction from its rho */
int iskra_ws_predict(iskra_ws *net);
/* calculate_electric_potential (generalized_poisson.jl:372-378) with the
 * network inside the iteration: flexible conjugate gradients on the
 * context's system (each new direction A-orthogonalised against the previous
 * one), every iteration preconditioned by z = the network applied to the
 * residual: when omega > 0, `sweeps` damped Jacobi sweeps
 * z += omega/4 (r - A z) from z = 0, then z += the network applied to
 * r - A z (scaled by its RMS and the gain), then `sweeps` sweeps again (train
 * with ISKRA_WS_ENERGY and the same omega and sweeps; sweeps in [0, 16]).
 * The system must be the network's: phi = 0 on all four
 * walls, no other Dirichlet node and no periodic link (ISKRA_E_UNSUPPORTED
 * otherwise).  Iterates until the FP64 relative residual ||f - A phi|| /
 * ||f|| meets the solver's rtol (ISKRA_E_CONVERGE after max_iter
 * iterations; 0: the solver's max_iter); warm: ISKRA_WARM_ZERO or
 * ISKRA_WARM_PREV (the current phi).  Leaves phi like iskra_poisson_solve
 * (E is not recomputed). */
int iskra_ws_solve(iskra_ws *net, int warm, double omega, int sweeps,
                   int max_iter, int *iters, double *relres);
/* Training pairs from the solve's own iterates (the residuals the
 * preconditioner meets inside the iteration): the context's phi must hold
 * the solution for its rho (call after an exact solve); flexible CG as in
 * iskra_ws_solve runs from phi = 0, and after iters[k] iterations (n_iters
 * increasing counts) the pair (r / rms(r), (phi - x) / rms(r)) of iterate x
 * and residual r = f - A x is appended to the dataset; stops early when r
 * vanishes.  *captured (nullable) = pairs appended.  phi is left as it was. */
int iskra_ws_capture_iterates(iskra_ws *net, double omega, int sweeps,
                              const int *iters, int n_iters, int64_t *captured);
/* sample index of the dataset: in = the stored (bf16) input plane, out =
 * the target plane, side x side floats each */
int iskra_ws_dataset_get(iskra_ws *net, int64_t index, float *in, float *out);
/* ---- diagnostics ---------------------------------------------------------- */
/* Device-to-device copy of `bytes` bytes, `iters` times, with the library's
 * own streaming copy kernel (nontemporal != 0: non-temporal loads/stores,
 * else plain ones, as the particle kernels use); *ms = mean duration of one
 * copy (HIP events).  Measures the achievable HBM bandwidth and calibrates
 * the HBM traffic counters. */
int iskra_bench_copy(iskra_ctx *ctx, int64_t bytes, int iters, int nontemporal,
                     double *ms);

#ifdef __cplusplus
}
#endif
#endif /* ISKRA_H */

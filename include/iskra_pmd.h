/* iskra_pmd.h -- openPMD 1.1.0 / HDF5 writer of the diagnostics records.
 *
 * The reference writes its diagnostics registry (Diagnostics/src/
 * Diagnostics.jl:24-47) with HDF5.jl: one file <prefix>/hdf5/data<i>.h5 per
 * saved iteration, holding /data/<i>/fields/<record> and
 * /data/<i>/particles/<species>/<record> with the openPMD attributes of
 * Diagnostics/src/openpmd/{root,fields,particles}.jl.  This library writes
 * the same files from host arrays with libhdf5's C API; callers download the
 * device state (iskra_field_download, iskra_species_download, ...) and hand
 * the arrays over.  It is a separate library (libiskra_pmd.so) so that the
 * device library does not depend on HDF5.
 *
 * Layout conventions (those of HDF5.jl writing Julia arrays):
 *  - a node field of a 2-D grid is an HDF5 dataset of dims (ny, nx) whose
 *    x index varies fastest (Julia's column-major nx x ny array); a 1-D
 *    field is a dataset of dims (1, nx);
 *  - a particle component is a 1-D dataset of np values;
 *  - scalar strings are fixed-length UTF-8 strings of their byte length (an
 *    empty string has a null dataspace), string tuples are 1-D arrays of
 *    variable-length UTF-8 strings, axis labels are 1-D arrays of one-byte
 *    strings.
 *
 * Every function returns ISKRA_OK (0) or an error code of iskra.h
 * (ISKRA_E_ARG for bad arguments or unknown units, ISKRA_E_STATE for a record
 * written twice or a closed file, ISKRA_E_IO for a failed HDF5 call);
 * iskra_pmd_last_error() gives the message of the calling thread's last
 * failure.  Calls on one file are not thread-safe.
 */
#ifndef ISKRA_PMD_H
#define ISKRA_PMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct iskra_pmd iskra_pmd; /* one open iteration file */

/* Root attributes that vary between runs (RootMetadata,
 * Diagnostics/src/openpmd/root.jl:46-63).  NULL pointers take the defaults:
 * author "", software "iskra", software_version "git+?", date = local time
 * formatted as the reference's Dates format "Y/m/d HH:MM". */
typedef struct {
  const char *author;
  int64_t seed;
  const char *software;
  const char *software_version;
  const char *date;
} iskra_pmd_root;

/* new_iteration(f, prefix, i, t, dt) (Diagnostics/src/hdf5.jl:47-66): create
 * the directory <prefix>/hdf5 (and its parents), create or truncate
 * <prefix>/hdf5/data<iteration>.h5, the groups /data/<iteration>/fields and
 * /data/<iteration>/particles, and the root, iteration (dt, time,
 * timeUnitSI = 1), meshes (FieldMetadata, fields.jl:1-12) and particles
 * (ParticleMetadata, particles.jl:1-8) attributes.  root may be NULL. */
int iskra_pmd_open(const char *prefix, int64_t iteration, double time, double dt,
                   const iskra_pmd_root *root, iskra_pmd **out);
/* Close the file (the end of new_iteration's do-block). */
int iskra_pmd_close(iskra_pmd *f);

/* usi(units) (Diagnostics/src/openpmd/root.jl:17-38): the openPMD
 * unitDimension (powers of length, mass, time, current, temperature, amount,
 * luminous intensity) and unitSI (the unit's value in coherent SI units) of a
 * unit string such as "1", "kg*m/s", "1/m^2", "C/m^2", "V/m", "eV", "cm^-3".
 * Grammar: products and quotients (* /) of factors, a factor being a
 * positive number, a parenthesised expression or an SI unit symbol with an
 * optional SI prefix, each with an optional integer power (^2, ^-1, ^(-3)). */
int iskra_pmd_units(const char *units, double dimension[7], double *unit_si);

/* The mesh geometry of a field record (geo(grid), root.jl:9-13, with the
 * grid's spacing): dim 1 or 2, axial 0 for x(,y) axes, 1 for r(,z) axes;
 * nodes and spacing per axis, x (r) first. */
typedef struct {
  int32_t dim;
  int32_t axial;
  int64_t nodes[2];
  double spacing[2];
} iskra_pmd_grid;

/* save_record(it, key, ::FieldRecord) (Diagnostics/src/hdf5.jl:86-95, records
 * built by init!(FieldRecord), openpmd/fields.jl:27-63).  data holds ncomp
 * node planes one after another, each nodes[0]*nodes[1] values with the
 * first axis fastest.  with_components = 0 (ncomp must be 1) writes the
 * dataset fields/<key>; with_components = 1 writes fields/<key>/x, /y, /z
 * from the first three planes, planes beyond ncomp being zero (the
 * reference's E and B records).  position (dim values, NULL: zeros) is the
 * components' in-cell position and time_offset the record's timeOffset. */
int iskra_pmd_field(iskra_pmd *f, const char *key, const char *units,
                    const iskra_pmd_grid *grid, int ncomp, int with_components,
                    const double *data, const double *position, double time_offset);

enum {
  ISKRA_PMD_F64 = 0, /* double */
  ISKRA_PMD_U32 = 1, /* uint32_t (particle ids, kinetic.jl:10) */
  ISKRA_PMD_I64 = 2  /* int64_t */
};

/* save_record(it, key, ::ParticleRecord) (Diagnostics/src/hdf5.jl:69-84,
 * records built by init!(ParticleRecord), openpmd/particles.jl:27-51), key
 * being "<species>/<record>".  data: ncomp columns of np values of type
 * dtype, column k starting at element k*ld (ld >= np; the reference's
 * N x M arrays have ld = N).  with_components = 0 (ncomp must be 1) writes
 * the dataset particles/<key>; with_components = 1 writes particles/<key>/x,
 * /y, /z, columns beyond ncomp being zero.  weighted sets macroWeighted. */
int iskra_pmd_particle(iskra_pmd *f, const char *key, const char *units, int64_t np,
                       int ncomp, int with_components, int dtype, const void *data,
                       int64_t ld, int weighted, double time_offset);
/* The same for a record whose input is a single value (mass, charge,
 * positionOffset/x, ...): the group particles/<key> with the attributes
 * value and shape = [np] (hdf5.jl:76-79). */
int iskra_pmd_particle_constant(iskra_pmd *f, const char *key, const char *units,
                                double value, int64_t np, int weighted, double time_offset);

/* save_record(it, key, ::CircuitProbeRecord) (Diagnostics/src/hdf5.jl:97-101,
 * Diagnostics/src/circuit.jl:1-19): the 1 x 1 dataset fields/<key> holding
 * value with the field metadata of a 2-D record without axes (axis labels
 * two empty strings, zero spacing and offset). */
int iskra_pmd_probe(iskra_pmd *f, const char *key, const char *units, double value,
                    double time_offset);

/* Message of the calling thread's last failed call ("" if none). */
const char *iskra_pmd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ISKRA_PMD_H */

/* openPMD 1.1.0 files of the diagnostics records, written with libhdf5.
 *
 * Restates the file layout of the reference's Diagnostics package
 * (Diagnostics/src/hdf5.jl:47-101, openpmd/root.jl, fields.jl,
 * particles.jl, circuit.jl) on libhdf5's C API: the reference writes through
 * HDF5.jl, whose conventions for Julia values (reversed dimensions,
 * fixed-length UTF-8 scalar strings, variable-length string vectors,
 * intermediate groups created on demand) are reproduced here so that the
 * files read the same in openPMD readers.  Plain C: the library links only
 * libhdf5 and the C runtime.
 */
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <time.h>

#include <hdf5.h>

#include "../../include/iskra.h"
#include "../../include/iskra_pmd.h"

struct iskra_pmd {
  hid_t file;
  hid_t base; /* /data/<i> */
};

static __thread char g_err[512];

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *iskra_pmd_last_error(void) { return g_err; }

/* HDF5 prints its error stack by default; every entry point silences it and
 * reports through the return code instead. */
typedef struct {
  H5E_auto2_t fn;
  void *data;
} quiet_t;

static quiet_t quiet_on(void) {
  quiet_t q;
  H5Eget_auto2(H5E_DEFAULT, &q.fn, &q.data);
  H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
  g_err[0] = 0;
  return q;
}

static int quiet_off(quiet_t q, int rc) {
  H5Eset_auto2(H5E_DEFAULT, q.fn, q.data);
  return rc;
}

/* ---- units (usi, root.jl:17-38) ----------------------------------------- */

/* dimension order of openPMD's unitDimension (root.jl:1-7) */
enum { D_L, D_M, D_T, D_I, D_TH, D_N, D_J };

/* a unit's value in coherent SI units is m * 10^e: powers of ten are kept
 * apart so that prefixed units come out exact (cm^-3 is 1e6, not
 * 999999.9999999999), as Unitful's rational arithmetic gives them */
typedef struct {
  double d[7];
  double m, e;
} uval;

typedef struct {
  const char *sym;
  double m;
  int e;
  signed char d[7];
  int prefixable;
} unit_t;

/* SI base and derived units and a few others Unitful parses, with their
 * value in coherent SI units (what upreferred gives). */
static const unit_t UNITS[] = {
    {"m", 1.0, 0, {1, 0, 0, 0, 0, 0, 0}, 1},
    {"g", 1.0, -3, {0, 1, 0, 0, 0, 0, 0}, 1},
    {"s", 1.0, 0, {0, 0, 1, 0, 0, 0, 0}, 1},
    {"A", 1.0, 0, {0, 0, 0, 1, 0, 0, 0}, 1},
    {"K", 1.0, 0, {0, 0, 0, 0, 1, 0, 0}, 1},
    {"mol", 1.0, 0, {0, 0, 0, 0, 0, 1, 0}, 1},
    {"cd", 1.0, 0, {0, 0, 0, 0, 0, 0, 1}, 1},
    {"Hz", 1.0, 0, {0, 0, -1, 0, 0, 0, 0}, 1},
    {"N", 1.0, 0, {1, 1, -2, 0, 0, 0, 0}, 1},
    {"Pa", 1.0, 0, {-1, 1, -2, 0, 0, 0, 0}, 1},
    {"J", 1.0, 0, {2, 1, -2, 0, 0, 0, 0}, 1},
    {"W", 1.0, 0, {2, 1, -3, 0, 0, 0, 0}, 1},
    {"C", 1.0, 0, {0, 0, 1, 1, 0, 0, 0}, 1},
    {"V", 1.0, 0, {2, 1, -3, -1, 0, 0, 0}, 1},
    {"F", 1.0, 0, {-2, -1, 4, 2, 0, 0, 0}, 1},
    {"Ohm", 1.0, 0, {2, 1, -3, -2, 0, 0, 0}, 1},
    {"\xce\xa9", 1.0, 0, {2, 1, -3, -2, 0, 0, 0}, 1}, /* U+03A9 OMEGA */
    {"S", 1.0, 0, {-2, -1, 3, 2, 0, 0, 0}, 1},
    {"Wb", 1.0, 0, {2, 1, -2, -1, 0, 0, 0}, 1},
    {"T", 1.0, 0, {0, 1, -2, -1, 0, 0, 0}, 1},
    {"H", 1.0, 0, {2, 1, -2, -2, 0, 0, 0}, 1},
    {"eV", 1.602176634e-19, 0, {2, 1, -2, 0, 0, 0, 0}, 1},
    {"L", 1.0, -3, {3, 0, 0, 0, 0, 0, 0}, 1},
    {"bar", 1.0, 5, {-1, 1, -2, 0, 0, 0, 0}, 1},
    {"min", 60.0, 0, {0, 0, 1, 0, 0, 0, 0}, 0},
    {"h", 3600.0, 0, {0, 0, 1, 0, 0, 0, 0}, 0},
    {"atm", 101325.0, 0, {-1, 1, -2, 0, 0, 0, 0}, 0},
    {"Torr", 101325.0 / 760.0, 0, {-1, 1, -2, 0, 0, 0, 0}, 0},
};

typedef struct {
  const char *sym;
  int e; /* power of ten */
} prefix_t;

static const prefix_t PREFIXES[] = {
    {"da", 1},  {"y", -24}, {"z", -21}, {"a", -18}, {"f", -15}, {"p", -12},
    {"n", -9},  {"\xce\xbc", -6} /* U+03BC */, {"\xc2\xb5", -6} /* U+00B5 */,
    {"m", -3},  {"c", -2},  {"d", -1},  {"h", 2},   {"k", 3},   {"M", 6},
    {"G", 9},   {"T", 12},  {"P", 15},  {"E", 18},  {"Z", 21},  {"Y", 24},
};

static const unit_t *unit_exact(const char *s, size_t n) {
  for (size_t k = 0; k < sizeof UNITS / sizeof UNITS[0]; ++k)
    if (strlen(UNITS[k].sym) == n && memcmp(UNITS[k].sym, s, n) == 0) return &UNITS[k];
  return NULL;
}

static int unit_symbol(const char *s, size_t n, uval *out) {
  const unit_t *u = unit_exact(s, n);
  int pe = 0;
  if (!u) {
    for (size_t k = 0; k < sizeof PREFIXES / sizeof PREFIXES[0] && !u; ++k) {
      size_t pl = strlen(PREFIXES[k].sym);
      if (pl < n && memcmp(PREFIXES[k].sym, s, pl) == 0) {
        const unit_t *c = unit_exact(s + pl, n - pl);
        if (c && c->prefixable) {
          u = c;
          pe = PREFIXES[k].e;
        }
      }
    }
  }
  if (!u) return 0;
  for (int k = 0; k < 7; ++k) out->d[k] = u->d[k];
  out->m = u->m;
  out->e = u->e + pe;
  return 1;
}

typedef struct {
  const char *p;
  const char *err;
} uparse_t;

static void skip_ws(uparse_t *P) {
  while (*P->p == ' ' || *P->p == '\t') P->p++;
}

static int is_sym_byte(unsigned char c) {
  return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c >= 0x80;
}

static int parse_expr(uparse_t *P, uval *out);

static int parse_int(uparse_t *P, long *out) {
  char *e;
  errno = 0;
  long v = strtol(P->p, &e, 10);
  if (e == P->p || errno) return 0;
  P->p = e;
  *out = v;
  return 1;
}

/* power := int | '(' ['-'|'+'] int ['/' int] ')' */
static int parse_power(uparse_t *P, double *out) {
  skip_ws(P);
  int paren = 0;
  if (*P->p == '(') {
    paren = 1;
    P->p++;
    skip_ws(P);
  }
  long num, den = 1;
  if (!parse_int(P, &num)) return 0;
  skip_ws(P);
  if (paren && *P->p == '/') {
    P->p++;
    skip_ws(P);
    if (!parse_int(P, &den) || den == 0) return 0;
    skip_ws(P);
  }
  if (paren) {
    if (*P->p != ')') return 0;
    P->p++;
  }
  *out = (double)num / (double)den;
  return 1;
}

static int parse_factor(uparse_t *P, uval *out) {
  skip_ws(P);
  const char *s = P->p;
  if (*s == '(') {
    P->p++;
    if (!parse_expr(P, out)) return 0;
    skip_ws(P);
    if (*P->p != ')') {
      P->err = "missing ')'";
      return 0;
    }
    P->p++;
  } else if ((*s >= '0' && *s <= '9') || *s == '.') {
    char *e;
    double v = strtod(s, &e);
    if (e == s || !(v > 0.0) || !isfinite(v)) {
      P->err = "bad number";
      return 0;
    }
    P->p = e;
    for (int k = 0; k < 7; ++k) out->d[k] = 0.0;
    out->m = v;
    out->e = 0.0;
  } else if (is_sym_byte((unsigned char)*s)) {
    const char *e = s;
    while (is_sym_byte((unsigned char)*e)) ++e;
    if (!unit_symbol(s, (size_t)(e - s), out)) {
      P->err = "unknown unit symbol";
      return 0;
    }
    P->p = e;
  } else {
    P->err = "expected a unit, a number or '('";
    return 0;
  }
  skip_ws(P);
  if (*P->p == '^') {
    P->p++;
    double pw;
    if (!parse_power(P, &pw)) {
      P->err = "bad power";
      return 0;
    }
    for (int k = 0; k < 7; ++k) out->d[k] *= pw;
    out->m = pow(out->m, pw);
    out->e *= pw;
  }
  return 1;
}

static int parse_expr(uparse_t *P, uval *out) {
  if (!parse_factor(P, out)) return 0;
  for (;;) {
    skip_ws(P);
    char op = *P->p;
    if (op != '*' && op != '/') return 1;
    P->p++;
    uval r;
    if (!parse_factor(P, &r)) return 0;
    const double sg = op == '*' ? 1.0 : -1.0;
    for (int k = 0; k < 7; ++k) out->d[k] += sg * r.d[k];
    out->m = op == '*' ? out->m * r.m : out->m / r.m;
    out->e += sg * r.e;
  }
}

int iskra_pmd_units(const char *units, double dimension[7], double *unit_si) {
  g_err[0] = 0;
  if (!units || !dimension || !unit_si) return fail(ISKRA_E_ARG, "units: NULL argument");
  uparse_t P = {units, NULL};
  uval u;
  if (!parse_expr(&P, &u) || (skip_ws(&P), *P.p != 0))
    return fail(ISKRA_E_ARG, "units \"%s\": %s at offset %d", units,
                P.err ? P.err : "unexpected text", (int)(P.p - units));
  for (int k = 0; k < 7; ++k) dimension[k] = u.d[k] == 0.0 ? 0.0 : u.d[k];
  /* 10^k is exact in double for 0 <= k <= 22, and one product or quotient
   * by it is correctly rounded */
  if (u.e == floor(u.e) && fabs(u.e) <= 22.0)
    *unit_si = u.e >= 0 ? u.m * pow(10.0, u.e) : u.m / pow(10.0, -u.e);
  else
    *unit_si = u.m * pow(10.0, u.e);
  return ISKRA_OK;
}

/* ---- attributes ---------------------------------------------------------- */

static int attr_write(hid_t obj, const char *name, hid_t type, hid_t space, const void *buf) {
  if (H5Aexists(obj, name) > 0 && H5Adelete(obj, name) < 0) return -1;
  hid_t a = H5Acreate2(obj, name, type, space, H5P_DEFAULT, H5P_DEFAULT);
  if (a < 0) return -1;
  herr_t e = buf ? H5Awrite(a, type, buf) : 0;
  H5Aclose(a);
  return e < 0 ? -1 : 0;
}

static int attr_scalar(hid_t obj, const char *name, hid_t type, const void *v) {
  hid_t s = H5Screate(H5S_SCALAR);
  int e = attr_write(obj, name, type, s, v);
  H5Sclose(s);
  return e;
}

static int attr_f64(hid_t obj, const char *name, double v) {
  return attr_scalar(obj, name, H5T_NATIVE_DOUBLE, &v);
}

static int attr_vec(hid_t obj, const char *name, hid_t type, const void *v, hsize_t n) {
  hid_t s = H5Screate_simple(1, &n, NULL);
  int e = attr_write(obj, name, type, s, v);
  H5Sclose(s);
  return e;
}

/* a Julia String: fixed length = byte length, UTF-8, scalar; "" has a null
 * dataspace */
static int attr_str(hid_t obj, const char *name, const char *v) {
  size_t n = strlen(v);
  hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, n ? n : 1);
  H5Tset_cset(t, H5T_CSET_UTF8);
  H5Tset_strpad(t, H5T_STR_NULLTERM);
  hid_t s = H5Screate(n ? H5S_SCALAR : H5S_NULL);
  int e = attr_write(obj, name, t, s, n ? v : NULL);
  H5Sclose(s);
  H5Tclose(t);
  return e;
}

/* a tuple of Strings: 1-D array of variable-length UTF-8 strings */
static int attr_strs(hid_t obj, const char *name, const char *const *v, hsize_t n) {
  hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, H5T_VARIABLE);
  H5Tset_cset(t, H5T_CSET_UTF8);
  int e = attr_vec(obj, name, t, v, n);
  H5Tclose(t);
  return e;
}

/* a tuple of Chars (hdf5.jl:5-19): 1-D array of one-byte UTF-8 strings */
static int attr_chars(hid_t obj, const char *name, const char *v, hsize_t n) {
  hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, 1);
  H5Tset_cset(t, H5T_CSET_UTF8);
  H5Tset_strpad(t, H5T_STR_NULLTERM);
  int e = attr_vec(obj, name, t, v, n);
  H5Tclose(t);
  return e;
}

/* ---- objects ------------------------------------------------------------- */

static hid_t lcpl_intermediate(void) {
  hid_t p = H5Pcreate(H5P_LINK_CREATE);
  if (p >= 0) H5Pset_create_intermediate_group(p, 1);
  return p;
}

/* does the path (relative to loc, '/'-separated) name an existing object? */
static int path_exists(hid_t loc, const char *path) {
  char buf[1024];
  size_t n = strlen(path);
  if (n >= sizeof buf) return 0;
  memcpy(buf, path, n + 1);
  for (size_t k = 0; k <= n; ++k) {
    if (buf[k] == '/' || buf[k] == 0) {
      char c = buf[k];
      buf[k] = 0;
      htri_t e = H5Lexists(loc, buf, H5P_DEFAULT);
      buf[k] = c;
      if (e <= 0) return 0;
    }
  }
  return 1;
}

static int mkdirs(const char *path) {
  char buf[4096];
  size_t n = strlen(path);
  if (n == 0 || n >= sizeof buf) return -1;
  memcpy(buf, path, n + 1);
  for (size_t k = 1; k <= n; ++k) {
    if (buf[k] == '/' || buf[k] == 0) {
      char c = buf[k];
      buf[k] = 0;
      if (mkdir(buf, 0777) != 0 && errno != EEXIST) return -1;
      buf[k] = c;
    }
  }
  return 0;
}

static int check_key(const char *key) {
  if (!key || !key[0] || key[0] == '/' || strlen(key) > 900 || strstr(key, "//")) return 0;
  return key[strlen(key) - 1] != '/';
}

/* ---- files ----------------------------------------------------------------- */

int iskra_pmd_open(const char *prefix, int64_t iteration, double t, double dt,
                   const iskra_pmd_root *root, iskra_pmd **out) {
  quiet_t q = quiet_on();
  if (!prefix || !out || iteration < 0)
    return quiet_off(q, fail(ISKRA_E_ARG, "open: NULL prefix or out, or negative iteration"));
  *out = NULL;
  char dir[4096], path[4200], grp[64];
  if (snprintf(dir, sizeof dir, "%s/hdf5", prefix) >= (int)sizeof dir)
    return quiet_off(q, fail(ISKRA_E_ARG, "open: prefix too long"));
  if (mkdirs(dir) != 0)
    return quiet_off(q, fail(ISKRA_E_IO, "open: cannot create %s: %s", dir, strerror(errno)));
  snprintf(path, sizeof path, "%s/data%lld.h5", dir, (long long)iteration);
  hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0) return quiet_off(q, fail(ISKRA_E_IO, "open: cannot create %s", path));
  snprintf(grp, sizeof grp, "data/%lld", (long long)iteration);
  hid_t lcpl = lcpl_intermediate();
  hid_t base = H5Gcreate2(file, grp, lcpl, H5P_DEFAULT, H5P_DEFAULT);
  hid_t meshes = base >= 0 ? H5Gcreate2(base, "fields", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT) : -1;
  hid_t parts = base >= 0 ? H5Gcreate2(base, "particles", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT) : -1;
  H5Pclose(lcpl);

  /* RootMetadata (root.jl:46-63) */
  char date[64];
  const char *author = root && root->author ? root->author : "";
  const char *software = root && root->software ? root->software : "iskra";
  const char *version = root && root->software_version ? root->software_version : "git+?";
  if (root && root->date) {
    snprintf(date, sizeof date, "%s", root->date);
  } else {
    time_t now = time(NULL);
    struct tm tm;
    localtime_r(&now, &tm);
    snprintf(date, sizeof date, "%d/%d/%d %02d:%02d", tm.tm_year + 1900, tm.tm_mon + 1,
             tm.tm_mday, tm.tm_hour, tm.tm_min);
  }
  const uint32_t ext = 1;
  const int64_t seed = root ? root->seed : 0;
  int e = meshes < 0 || parts < 0;
  e = e || attr_str(file, "openPMD", "1.1.0");
  e = e || attr_scalar(file, "openPMDextension", H5T_NATIVE_UINT32, &ext);
  e = e || attr_str(file, "iterationEncoding", "fileBased");
  e = e || attr_str(file, "iterationFormat", "data%T.h5");
  e = e || attr_str(file, "basePath", "/data/%T");
  e = e || attr_str(file, "meshesPath", "fields/");
  e = e || attr_str(file, "particlesPath", "particles/");
  e = e || attr_str(file, "author", author);
  e = e || attr_scalar(file, "seed", H5T_NATIVE_INT64, &seed);
  e = e || attr_str(file, "software", software);
  e = e || attr_str(file, "softwareVersion", version);
  e = e || attr_str(file, "date", date);
  /* FieldMetadata (fields.jl:1-12) */
  static const char *const open4[4] = {"open", "open", "open", "open"};
  static const char *const periodic4[4] = {"periodic", "periodic", "periodic", "periodic"};
  e = e || attr_str(meshes, "fieldSolver", "other");
  e = e || attr_str(meshes, "fieldSolverParameters", "Nagel");
  e = e || attr_strs(meshes, "fieldBoundary", open4, 4);
  e = e || attr_strs(meshes, "particleBoundary", periodic4, 4);
  e = e || attr_str(meshes, "currentSmoothing", "none");
  e = e || attr_str(meshes, "chargeCorrection", "none");
  /* Iteration(dt, t, 1.0) (hdf5.jl:48, root.jl:40-44) */
  e = e || attr_f64(base, "dt", dt);
  e = e || attr_f64(base, "time", t);
  e = e || attr_f64(base, "timeUnitSI", 1.0);
  /* ParticleMetadata (particles.jl:1-8) */
  e = e || attr_f64(parts, "particleShape", 0.0);
  e = e || attr_str(parts, "currentDeposition", "none");
  e = e || attr_str(parts, "particlePush", "Boris");
  e = e || attr_str(parts, "particleInterpolation", "other");
  e = e || attr_str(parts, "particleSmoothing", "none");
  if (meshes >= 0) H5Gclose(meshes);
  if (parts >= 0) H5Gclose(parts);
  if (e) {
    if (base >= 0) H5Gclose(base);
    H5Fclose(file);
    return quiet_off(q, fail(ISKRA_E_IO, "open: writing the groups and attributes of %s failed",
                             path));
  }
  iskra_pmd *f = (iskra_pmd *)calloc(1, sizeof *f);
  if (!f) {
    H5Gclose(base);
    H5Fclose(file);
    return quiet_off(q, fail(ISKRA_E_IO, "open: out of memory"));
  }
  f->file = file;
  f->base = base;
  *out = f;
  return quiet_off(q, ISKRA_OK);
}

int iskra_pmd_close(iskra_pmd *f) {
  quiet_t q = quiet_on();
  if (!f) return quiet_off(q, fail(ISKRA_E_ARG, "close: NULL file"));
  herr_t a = H5Gclose(f->base);
  herr_t b = H5Fclose(f->file);
  free(f);
  if (a < 0 || b < 0) return quiet_off(q, fail(ISKRA_E_IO, "close: HDF5 close failed"));
  return quiet_off(q, ISKRA_OK);
}

/* ---- records ------------------------------------------------------------- */

/* create a dataset (with intermediate groups) at base/path and write buf */
static int dataset(hid_t base, const char *path, hid_t type, int rank, const hsize_t *dims,
                   const void *buf) {
  hid_t lcpl = lcpl_intermediate();
  hid_t s = H5Screate_simple(rank, dims, NULL);
  hid_t d = H5Dcreate2(base, path, type, s, lcpl, H5P_DEFAULT, H5P_DEFAULT);
  H5Pclose(lcpl);
  H5Sclose(s);
  if (d < 0) return -1;
  hsize_t n = 1;
  for (int k = 0; k < rank; ++k) n *= dims[k];
  herr_t e = n ? H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) : 0;
  H5Dclose(d);
  return e < 0 ? -1 : 0;
}

static int group(hid_t base, const char *path) {
  hid_t lcpl = lcpl_intermediate();
  hid_t g = H5Gcreate2(base, path, lcpl, H5P_DEFAULT, H5P_DEFAULT);
  H5Pclose(lcpl);
  if (g < 0) return -1;
  H5Gclose(g);
  return 0;
}

/* FieldRecordMetadata (fields.jl:15-31) split as save_record does: position
 * and unitSI on each component, the rest on the record. */
typedef struct {
  double dim7[7];
  double unit_si;
  double time_offset;
  int n;               /* axes */
  const char *labels;  /* n one-byte labels ("" each for probe records) */
  double offset[2], spacing[2], position[2];
} field_meta;

static int field_component_attrs(hid_t obj, const field_meta *m) {
  int e = attr_vec(obj, "position", H5T_NATIVE_DOUBLE, m->position, (hsize_t)m->n);
  return e || attr_f64(obj, "unitSI", m->unit_si);
}

static int field_record_attrs(hid_t obj, const field_meta *m) {
  int e = attr_vec(obj, "unitDimension", H5T_NATIVE_DOUBLE, m->dim7, 7);
  e = e || attr_f64(obj, "timeOffset", m->time_offset);
  e = e || attr_chars(obj, "axisLabels", m->labels, (hsize_t)m->n);
  e = e || attr_str(obj, "dataOrder", "C");
  e = e || attr_str(obj, "geometry", "cartesian");
  e = e || attr_str(obj, "geometryParameters", "");
  e = e || attr_vec(obj, "gridGlobalOffset", H5T_NATIVE_DOUBLE, m->offset, (hsize_t)m->n);
  e = e || attr_vec(obj, "gridSpacing", H5T_NATIVE_DOUBLE, m->spacing, (hsize_t)m->n);
  e = e || attr_f64(obj, "gridUnitSI", 1.0);
  e = e || attr_str(obj, "fieldSmoothing", "none");
  return e;
}

static int with_obj(hid_t base, const char *path, int (*fn)(hid_t, const void *),
                    const void *arg) {
  hid_t o = H5Oopen(base, path, H5P_DEFAULT);
  if (o < 0) return -1;
  int e = fn(o, arg);
  H5Oclose(o);
  return e;
}

static int fc_attrs(hid_t o, const void *m) { return field_component_attrs(o, (const field_meta *)m); }
static int fr_attrs(hid_t o, const void *m) { return field_record_attrs(o, (const field_meta *)m); }
static int fall_attrs(hid_t o, const void *m) {
  return field_component_attrs(o, (const field_meta *)m) || field_record_attrs(o, (const field_meta *)m);
}

int iskra_pmd_field(iskra_pmd *f, const char *key, const char *units,
                    const iskra_pmd_grid *g, int ncomp, int with_components,
                    const double *data, const double *position, double time_offset) {
  quiet_t q = quiet_on();
  if (!f) return quiet_off(q, fail(ISKRA_E_STATE, "field: closed or NULL file"));
  if (!check_key(key) || !units || !g || !data)
    return quiet_off(q, fail(ISKRA_E_ARG, "field: bad key or NULL argument"));
  if ((g->dim != 1 && g->dim != 2) || (g->axial != 0 && g->axial != 1) || g->nodes[0] < 1 ||
      (g->dim == 2 && g->nodes[1] < 1))
    return quiet_off(q, fail(ISKRA_E_ARG, "field %s: bad grid", key));
  if (ncomp < 1 || (!with_components && ncomp != 1) || ncomp > 3)
    return quiet_off(q, fail(ISKRA_E_ARG, "field %s: ncomp %d with_components %d", key, ncomp,
                             with_components));
  field_meta m;
  memset(&m, 0, sizeof m);
  int rc = iskra_pmd_units(units, m.dim7, &m.unit_si);
  if (rc) return quiet_off(q, rc);
  m.time_offset = time_offset;
  m.n = g->dim;
  m.labels = g->axial ? "rz" : "xy";
  for (int k = 0; k < g->dim; ++k) {
    m.spacing[k] = g->spacing[k];
    m.position[k] = position ? position[k] : 0.0;
  }
  char rec[1024], comp[1040];
  snprintf(rec, sizeof rec, "fields/%s", key);
  if (path_exists(f->base, rec))
    return quiet_off(q, fail(ISKRA_E_STATE, "field %s: already written in this iteration", key));
  const int64_t nx = g->nodes[0], ny = g->dim == 2 ? g->nodes[1] : 1;
  const hsize_t dims[2] = {(hsize_t)ny, (hsize_t)nx};
  const size_t plane = (size_t)nx * (size_t)ny;
  int e = 0;
  if (!with_components) {
    e = dataset(f->base, rec, H5T_NATIVE_DOUBLE, 2, dims, data);
    e = e || with_obj(f->base, rec, fall_attrs, &m);
  } else {
    double *zero = NULL;
    for (int c = 0; c < 3 && !e; ++c) {
      snprintf(comp, sizeof comp, "%s/%c", rec, "xyz"[c]);
      const double *src = data + (size_t)c * plane;
      if (c >= ncomp) {
        if (!zero && !(zero = (double *)calloc(plane, sizeof(double)))) {
          e = 1;
          break;
        }
        src = zero;
      }
      e = dataset(f->base, comp, H5T_NATIVE_DOUBLE, 2, dims, src);
      e = e || with_obj(f->base, comp, fc_attrs, &m);
    }
    free(zero);
    e = e || with_obj(f->base, rec, fr_attrs, &m);
  }
  if (e) return quiet_off(q, fail(ISKRA_E_IO, "field %s: HDF5 write failed", key));
  return quiet_off(q, ISKRA_OK);
}

/* ParticleRecordMetadata (particles.jl:10-18) */
typedef struct {
  double dim7[7];
  double unit_si;
  double time_offset;
  uint8_t weighted;
} part_meta;

static int pc_attrs(hid_t o, const void *a) {
  return attr_f64(o, "unitSI", ((const part_meta *)a)->unit_si);
}

static int pr_attrs(hid_t o, const void *a) {
  const part_meta *m = (const part_meta *)a;
  int e = attr_vec(o, "unitDimension", H5T_NATIVE_DOUBLE, m->dim7, 7);
  e = e || attr_f64(o, "timeOffset", m->time_offset);
  e = e || attr_scalar(o, "macroWeighted", H5T_NATIVE_UINT8, &m->weighted);
  e = e || attr_f64(o, "weightingPower", 1.0);
  return e;
}

static int part_begin(iskra_pmd *f, const char *key, const char *units, int weighted,
                      double time_offset, part_meta *m, char *rec, size_t recn) {
  if (!f) return fail(ISKRA_E_STATE, "particle record: closed or NULL file");
  if (!check_key(key) || !strchr(key, '/') || !units)
    return fail(ISKRA_E_ARG, "particle record: key must be <species>/<record>");
  memset(m, 0, sizeof *m);
  int rc = iskra_pmd_units(units, m->dim7, &m->unit_si);
  if (rc) return rc;
  m->time_offset = time_offset;
  m->weighted = weighted ? 1 : 0;
  snprintf(rec, recn, "particles/%s", key);
  if (path_exists(f->base, rec))
    return fail(ISKRA_E_STATE, "particle record %s: already written in this iteration", key);
  return ISKRA_OK;
}

int iskra_pmd_particle(iskra_pmd *f, const char *key, const char *units, int64_t np,
                       int ncomp, int with_components, int dtype, const void *data,
                       int64_t ld, int weighted, double time_offset) {
  quiet_t q = quiet_on();
  part_meta m;
  char rec[1024], comp[1040];
  int rc = part_begin(f, key, units, weighted, time_offset, &m, rec, sizeof rec);
  if (rc) return quiet_off(q, rc);
  if (np < 0 || ld < np || (np > 0 && !data) || ncomp < 1 || ncomp > 3 ||
      (!with_components && ncomp != 1))
    return quiet_off(q, fail(ISKRA_E_ARG, "particle record %s: np %lld ld %lld ncomp %d", key,
                             (long long)np, (long long)ld, ncomp));
  hid_t type;
  size_t esz;
  switch (dtype) {
    case ISKRA_PMD_F64: type = H5T_NATIVE_DOUBLE; esz = 8; break;
    case ISKRA_PMD_U32: type = H5T_NATIVE_UINT32; esz = 4; break;
    case ISKRA_PMD_I64: type = H5T_NATIVE_INT64; esz = 8; break;
    default: return quiet_off(q, fail(ISKRA_E_ARG, "particle record %s: dtype %d", key, dtype));
  }
  const hsize_t dims[1] = {(hsize_t)np};
  int e = 0;
  if (!with_components) {
    e = dataset(f->base, rec, type, 1, dims, data);
    e = e || with_obj(f->base, rec, pc_attrs, &m);
  } else {
    void *zero = NULL;
    for (int c = 0; c < 3 && !e; ++c) {
      snprintf(comp, sizeof comp, "%s/%c", rec, "xyz"[c]);
      const void *src = (const char *)data + (size_t)c * (size_t)ld * esz;
      if (c >= ncomp) {
        if (!zero && !(zero = calloc(np ? (size_t)np : 1, esz))) {
          e = 1;
          break;
        }
        src = zero;
      }
      e = dataset(f->base, comp, type, 1, dims, src);
      e = e || with_obj(f->base, comp, pc_attrs, &m);
    }
    free(zero);
  }
  e = e || with_obj(f->base, rec, pr_attrs, &m);
  if (e) return quiet_off(q, fail(ISKRA_E_IO, "particle record %s: HDF5 write failed", key));
  return quiet_off(q, ISKRA_OK);
}

int iskra_pmd_particle_constant(iskra_pmd *f, const char *key, const char *units,
                                double value, int64_t np, int weighted, double time_offset) {
  quiet_t q = quiet_on();
  part_meta m;
  char rec[1024];
  int rc = part_begin(f, key, units, weighted, time_offset, &m, rec, sizeof rec);
  if (rc) return quiet_off(q, rc);
  if (np < 0) return quiet_off(q, fail(ISKRA_E_ARG, "particle record %s: np < 0", key));
  hid_t o;
  int e = group(f->base, rec);
  if (!e && (o = H5Oopen(f->base, rec, H5P_DEFAULT)) >= 0) {
    const int64_t shape = np;
    e = attr_f64(o, "value", value);
    e = e || attr_vec(o, "shape", H5T_NATIVE_INT64, &shape, 1);
    e = e || pc_attrs(o, &m) || pr_attrs(o, &m);
    H5Oclose(o);
  } else {
    e = 1;
  }
  if (e) return quiet_off(q, fail(ISKRA_E_IO, "particle record %s: HDF5 write failed", key));
  return quiet_off(q, ISKRA_OK);
}

int iskra_pmd_probe(iskra_pmd *f, const char *key, const char *units, double value,
                    double time_offset) {
  quiet_t q = quiet_on();
  if (!f) return quiet_off(q, fail(ISKRA_E_STATE, "probe: closed or NULL file"));
  if (!check_key(key) || !units) return quiet_off(q, fail(ISKRA_E_ARG, "probe: bad key"));
  field_meta m;
  memset(&m, 0, sizeof m);
  int rc = iskra_pmd_units(units, m.dim7, &m.unit_si);
  if (rc) return quiet_off(q, rc);
  m.time_offset = time_offset;
  m.n = 2;
  m.labels = "\0\0";
  char rec[1024];
  snprintf(rec, sizeof rec, "fields/%s", key);
  if (path_exists(f->base, rec))
    return quiet_off(q, fail(ISKRA_E_STATE, "probe %s: already written in this iteration", key));
  const hsize_t dims[2] = {1, 1};
  int e = dataset(f->base, rec, H5T_NATIVE_DOUBLE, 2, dims, &value);
  e = e || with_obj(f->base, rec, fall_attrs, &m);
  if (e) return quiet_off(q, fail(ISKRA_E_IO, "probe %s: HDF5 write failed", key));
  return quiet_off(q, ISKRA_OK);
}

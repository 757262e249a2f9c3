/* stokes_nc.c -- libstokes trajectory files in the netCDF classic format
 * with 64-bit offsets (CDF-2), written without the netCDF library.
 *
 * The file layout (dimensions, variables, their long_name attributes, the
 * flag table of stokes_nc_init and the index variables filled at the end
 * of the definitions) is that of the reference's stokes-nc.c:716-1366 and
 * 1445-1598, so the files read the same in tools that read the reference's
 * files.  The encoding is the published netCDF classic format: a header
 * ("CDF" 2, the record count, the dimension list, an absent global
 * attribute list, the variable list with each variable's dimension ids,
 * attributes, type, padded size vsize and 8-byte begin offset), then the
 * fixed-size variables one after the other, then the records, each record
 * holding every record variable's slab in definition order; all values
 * big-endian.  As the netCDF library does in its default fill mode,
 * fixed-size variables and new records are filled with the type's fill
 * value before data are written.  Errors print a message and exit, as
 * stokes_nc_error does (stokes-nc.c:33-48).
 */
#define _POSIX_C_SOURCE 200809L
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "stokes-nc.h"
#include "stokes.h"

enum { NCT_CHAR = 2, NCT_INT = 4, NCT_DOUBLE = 6 };
enum { NC_TAG_DIM = 0x0A, NC_TAG_VAR = 0x0B, NC_TAG_ATT = 0x0C };
#define NC_MAXDIM 8
#define NC_MAXVAR 72
/* default fill values of the classic format */
#define FILL_DOUBLE 9.9692099683868690e+36
#define FILL_INT (-2147483647)

typedef struct {
  char name[16];
  int len;                      /* 0: the record dimension */
} nc_dim;

typedef struct {
  char name[16];
  int type;
  int nd;
  int dim[3];
  const char *longname;         /* NULL: no attribute */
  int rec;                      /* first dimension is the record one */
  size_t count;                 /* values (per record for record vars) */
  size_t vsize;                 /* bytes, padded to 4 */
  long long begin;              /* file offset (record vars: of record 0) */
} nc_var;

/* the file: dimensions, variables, record bookkeeping */
typedef struct {
  FILE *fp;
  char fname[256];
  int ndims, nvars;
  nc_dim dims[NC_MAXDIM];
  nc_var vars[NC_MAXVAR];
  int numrecs;
  long long recsize;
  int *ntime;                   /* the public struct's ntime */
} nc_file;

#define NCF(nc) ((nc_file *) (nc)->priv)

static void
nc_die (const char *fname, const char *what, const char *name)
{
  fprintf (stderr, "libstokes-hip stokes_nc (%s): %s%s%s\n",
           fname ? fname : "?", what, name ? " for " : "", name ? name : "");
  exit (1);
}

/* ---- big-endian encoding ---- */

static void
be32 (unsigned char *p, uint32_t v)
{
  p[0] = (unsigned char) (v >> 24);
  p[1] = (unsigned char) (v >> 16);
  p[2] = (unsigned char) (v >> 8);
  p[3] = (unsigned char) v;
}

static void
be64 (unsigned char *p, uint64_t v)
{
  be32 (p, (uint32_t) (v >> 32));
  be32 (p + 4, (uint32_t) v);
}

/* a growing byte buffer for the header */
typedef struct {
  unsigned char *b;
  size_t n, cap;
} nc_buf;

static void
buf_put (nc_buf *h, const void *p, size_t n)
{
  if (h->n + n > h->cap)
    {
      size_t cap = h->cap ? 2 * h->cap : 1024;
      while (cap < h->n + n) cap *= 2;
      unsigned char *b = realloc (h->b, cap);
      if (!b) nc_die (NULL, "header allocation failed", NULL);
      h->b = b;
      h->cap = cap;
    }
  memcpy (h->b + h->n, p, n);
  h->n += n;
}

static void
buf_u32 (nc_buf *h, uint32_t v)
{
  unsigned char t[4];
  be32 (t, v);
  buf_put (h, t, 4);
}

static void
buf_u64 (nc_buf *h, uint64_t v)
{
  unsigned char t[8];
  be64 (t, v);
  buf_put (h, t, 8);
}

/* name or text: length, bytes, zero padding to 4 */
static void
buf_text (nc_buf *h, const char *s)
{
  size_t n = strlen (s);
  static const unsigned char zero[4] = {0, 0, 0, 0};
  buf_u32 (h, (uint32_t) n);
  buf_put (h, s, n);
  if (n % 4) buf_put (h, zero, 4 - n % 4);
}

/* ---- define mode ---- */

static int
def_dim (nc_file *nc, const char *name, int len)
{
  if (nc->ndims >= NC_MAXDIM) nc_die (nc->fname, "too many dimensions", name);
  nc_dim *d = &nc->dims[nc->ndims];
  snprintf (d->name, sizeof (d->name), "%s", name);
  d->len = len;
  return nc->ndims++;
}

static void
def_var (nc_file *nc, const char *name, int type, int nd,
         const int *dims, const char *longname)
{
  if (nc->nvars >= NC_MAXVAR) nc_die (nc->fname, "too many variables", name);
  nc_var *v = &nc->vars[nc->nvars++];
  memset (v, 0, sizeof (*v));
  snprintf (v->name, sizeof (v->name), "%s", name);
  v->type = type;
  v->nd = nd;
  v->count = 1;
  for (int k = 0; k < nd; k++)
    {
      v->dim[k] = dims[k];
      if (nc->dims[dims[k]].len == 0)
        v->rec = k == 0;
      else
        v->count *= (size_t) nc->dims[dims[k]].len;
    }
  v->longname = longname;
  const size_t bytes = v->count * (type == NCT_INT ? 4 : 8);
  v->vsize = (bytes + 3) / 4 * 4;
}

static nc_var *
var_of (nc_file *nc, const char *name)
{
  for (int k = 0; k < nc->nvars; k++)
    if (strcmp (nc->vars[k].name, name) == 0) return &nc->vars[k];
  nc_die (nc->fname, "variable not defined in this file", name);
  return NULL;
}

/* write count values of v (doubles, or ints when v is NC_INT) at offset */
static void
write_values (nc_file *nc, const nc_var *v, long long off,
              const double *d, const int *iv)
{
  unsigned char *b = malloc (v->vsize ? v->vsize : 4);
  if (!b) nc_die (nc->fname, "allocation failed", v->name);
  memset (b, 0, v->vsize);
  for (size_t k = 0; k < v->count; k++)
    {
      if (v->type == NCT_INT)
        be32 (b + 4 * k, (uint32_t) (iv ? iv[k] : FILL_INT));
      else
        {
          double x = d ? d[k] : FILL_DOUBLE;
          uint64_t u;
          memcpy (&u, &x, 8);
          be64 (b + 8 * k, u);
        }
    }
  if (fseeko (nc->fp, (off_t) off, SEEK_SET) != 0
      || fwrite (b, 1, v->vsize, nc->fp) != v->vsize)
    nc_die (nc->fname, "write failed", v->name);
  free (b);
}

/* leave define mode: header, offsets, fill of the fixed-size variables */
static void
end_def (nc_file *nc)
{
  /* the header size is independent of the begin values (8 bytes each) */
  nc_buf h = {0};
  for (int pass = 0; pass < 2; pass++)
    {
      h.n = 0;
      buf_put (&h, "CDF\002", 4);
      buf_u32 (&h, 0);                         /* numrecs */
      buf_u32 (&h, NC_TAG_DIM);
      buf_u32 (&h, (uint32_t) nc->ndims);
      for (int k = 0; k < nc->ndims; k++)
        {
          buf_text (&h, nc->dims[k].name);
          buf_u32 (&h, (uint32_t) nc->dims[k].len);
        }
      buf_u32 (&h, 0);                         /* no global attributes */
      buf_u32 (&h, 0);
      buf_u32 (&h, NC_TAG_VAR);
      buf_u32 (&h, (uint32_t) nc->nvars);
      for (int k = 0; k < nc->nvars; k++)
        {
          const nc_var *v = &nc->vars[k];
          buf_text (&h, v->name);
          buf_u32 (&h, (uint32_t) v->nd);
          for (int j = 0; j < v->nd; j++) buf_u32 (&h, (uint32_t) v->dim[j]);
          if (v->longname)
            {
              buf_u32 (&h, NC_TAG_ATT);
              buf_u32 (&h, 1);
              buf_text (&h, "long_name");
              buf_u32 (&h, NCT_CHAR);
              buf_text (&h, v->longname);
            }
          else
            {
              buf_u32 (&h, 0);
              buf_u32 (&h, 0);
            }
          buf_u32 (&h, (uint32_t) v->type);
          buf_u32 (&h, (uint32_t) v->vsize);
          buf_u64 (&h, (uint64_t) v->begin);
        }
      if (pass == 1) break;
      /* offsets: the fixed-size variables after the header, then the
       * records */
      long long off = (long long) h.n;
      for (int k = 0; k < nc->nvars; k++)
        if (!nc->vars[k].rec)
          {
            nc->vars[k].begin = off;
            off += (long long) nc->vars[k].vsize;
          }
      nc->recsize = 0;
      for (int k = 0; k < nc->nvars; k++)
        if (nc->vars[k].rec)
          {
            nc->vars[k].begin = off + nc->recsize;
            nc->recsize += (long long) nc->vars[k].vsize;
          }
    }
  if (fwrite (h.b, 1, h.n, nc->fp) != h.n)
    nc_die (nc->fname, "header write failed", NULL);
  free (h.b);
  for (int k = 0; k < nc->nvars; k++)
    if (!nc->vars[k].rec)
      write_values (nc, &nc->vars[k], nc->vars[k].begin, NULL, NULL);
}

/* ---- data ---- */

static void
put_fixed (nc_file *nc, const char *name, const double *d,
           const int *iv)
{
  const nc_var *v = var_of (nc, name);
  if (v->rec) nc_die (nc->fname, "a record variable needs a step", name);
  write_values (nc, v, v->begin, d, iv);
}

static void
put_rec (nc_file *nc, const char *name, int step, const double *d,
         const int *iv)
{
  const nc_var *v = var_of (nc, name);
  if (!v->rec) nc_die (nc->fname, "not a record variable", name);
  if (step < 0) nc_die (nc->fname, "negative step", name);
  if (step >= nc->numrecs)
    {
      /* new records: every record variable filled first */
      for (int r = nc->numrecs; r <= step; r++)
        for (int k = 0; k < nc->nvars; k++)
          if (nc->vars[k].rec)
            write_values (nc, &nc->vars[k],
                          nc->vars[k].begin + (long long) r * nc->recsize,
                          NULL, NULL);
      nc->numrecs = step + 1;
      if (nc->ntime) *nc->ntime = nc->numrecs;
      unsigned char t[4];
      be32 (t, (uint32_t) nc->numrecs);
      if (fseeko (nc->fp, 4, SEEK_SET) != 0 || fwrite (t, 1, 4, nc->fp) != 4)
        nc_die (nc->fname, "record count write failed", name);
    }
  write_values (nc, v, v->begin + (long long) step * nc->recsize, d, iv);
}

/* ---- stokes_nc_init_ (stokes-nc.c:716-1366) ---- */

typedef struct {
  int ui0, oi0, ei0, ui, oi, ei;
  int x0, u0, o0, e0, f0, t0, s0;
  int xf0, uf0, of0, ef0, ff0, tf0, sf0;
  int x, u, o, e, f, t, s, q;
  int xf, uf, of, ef, ff, tf, sf;
  int a, af;
} nc_flags;

static struct stokes_nc *nc_public (nc_file *f);

static struct stokes_nc *
nc_create (const char *filename, int nm, int nf, const nc_flags *fl,
           int shear_mode, int flag_rng)
{
  nc_file *nc = calloc (1, sizeof (*nc));
  if (!nc) nc_die (filename, "allocation failed", NULL);
  snprintf (nc->fname, sizeof (nc->fname), "%s", filename);
  /* NC_NOCLOBBER: an existing file is an error */
  nc->fp = fopen (filename, "w+xb");
  if (!nc->fp) nc_die (filename, "cannot create (exists, or no access)", NULL);
  int dv[3];
  const int p_dim = def_dim (nc, "p", nm);
  dv[0] = p_dim;
  def_var (nc, "p", NCT_INT, 1, dv, NULL);
  int pf_dim = -1;
  if (nf > 0)
    {
      pf_dim = def_dim (nc, "pf", nf);
      dv[0] = pf_dim;
      def_var (nc, "pf", NCT_INT, 1, dv, NULL);
    }
  const int vec_dim = def_dim (nc, "vec", 3);
  dv[0] = vec_dim;
  def_var (nc, "vec", NCT_INT, 1, dv, NULL);
  const int stt_dim = def_dim (nc, "stt", 5);
  dv[0] = stt_dim;
  def_var (nc, "stt", NCT_INT, 1, dv, NULL);
  const int quat_dim = def_dim (nc, "quat", 4);
  dv[0] = quat_dim;
  def_var (nc, "quat", NCT_INT, 1, dv, NULL);
  int rng_dim = -1;
  if (flag_rng)
    {
      rng_dim = def_dim (nc, "rng", 624);      /* MT19937 state words */
      dv[0] = rng_dim;
      def_var (nc, "rng", NCT_INT, 1, dv, NULL);
    }
  const int t = def_dim (nc, "time", 0);
  dv[0] = t;
  def_var (nc, "time", NCT_DOUBLE, 1, dv, NULL);

#define V1(n, a, ln) do { int d_[1] = {a}; def_var (nc, n, NCT_DOUBLE, 1, d_, ln); } while (0)
#define V2(n, a, b, ln) do { int d_[2] = {a, b}; def_var (nc, n, NCT_DOUBLE, 2, d_, ln); } while (0)
#define V3(n, a, b, c, ln) do { int d_[3] = {a, b, c}; def_var (nc, n, NCT_DOUBLE, 3, d_, ln); } while (0)
  V1 ("l", vec_dim, "lattice vector");
  def_var (nc, "shear_mode", NCT_INT, 0, dv, "shear mode");
  if (shear_mode != 0)
    {
      def_var (nc, "shear_rate", NCT_DOUBLE, 0, dv, "shear rate");
      V1 ("shear_shift", t, "cell shift for shear boundary condition");
    }
  if (fl->ui0) V1 ("Ui0", vec_dim, "constant imposed translational velocity");
  if (fl->oi0) V1 ("Oi0", vec_dim, "constant imposed angular velocity");
  if (fl->ei0) V1 ("Ei0", stt_dim, "constant imposed strain");
  if (fl->ui) V2 ("Ui", t, vec_dim, "(changing) imposed translational velocity");
  if (fl->oi) V2 ("Oi", t, vec_dim, "(changing) imposed angular velocity");
  if (fl->ei) V2 ("Ei", t, stt_dim, "(changing) imposed strain");
  if (fl->x0) V2 ("x0", p_dim, vec_dim, "positions of particle center at t0");
  if (fl->u0) V2 ("U0", p_dim, vec_dim, "translational velocities at t0");
  if (fl->o0) V2 ("O0", p_dim, vec_dim, "angular velocities at t0");
  if (fl->e0) V2 ("E0", p_dim, stt_dim, "rate of strain at t0");
  if (fl->f0) V2 ("F0", p_dim, vec_dim, "forces at t0");
  if (fl->t0) V2 ("T0", p_dim, vec_dim, "torques at t0");
  if (fl->s0) V2 ("S0", p_dim, stt_dim, "stresslets at t0");
  if (fl->x) V3 ("x", t, p_dim, vec_dim, "positions of particle center");
  if (fl->q) V3 ("q", t, p_dim, quat_dim, "quaternion of particle direction");
  if (fl->a) V1 ("a", p_dim, "radii of particle");
  if (fl->u) V3 ("U", t, p_dim, vec_dim, "translational velocities");
  if (fl->o) V3 ("O", t, p_dim, vec_dim, "angular velocities");
  if (fl->e) V3 ("E", t, p_dim, stt_dim, "rate of strain");
  if (fl->f) V3 ("F", t, p_dim, vec_dim, "forces");
  if (fl->t) V3 ("T", t, p_dim, vec_dim, "torques");
  if (fl->s) V3 ("S", t, p_dim, stt_dim, "stresslets");
  if (nf > 0)
    {
      const int q = pf_dim;
      if (fl->xf0) V2 ("xf0", q, vec_dim, "positions of fixed particle center at t0");
      if (fl->af) V1 ("af", q, "radii of particle");
      if (fl->uf0) V2 ("Uf0", q, vec_dim, "translational velocities for fixed particles at t0");
      if (fl->of0) V2 ("Of0", q, vec_dim, "angular velocities for fixed particles at t0");
      if (fl->ef0) V2 ("Ef0", q, stt_dim, "rate of strain for fixed particles at t0");
      if (fl->ff0) V2 ("Ff0", q, vec_dim, "forces for fixed particles at t0");
      if (fl->tf0) V2 ("Tf0", q, vec_dim, "torques for fixed particles at t0");
      if (fl->sf0) V2 ("Sf0", q, stt_dim, "stresslets for fixed particles at t0");
      if (fl->xf) V3 ("xf", t, q, vec_dim, "positions of fixed particle center");
      if (fl->uf) V3 ("Uf", t, q, vec_dim, "translational velocities for fixed particles");
      if (fl->of) V3 ("Of", t, q, vec_dim, "angular velocities for fixed particles");
      if (fl->ef) V3 ("Ef", t, q, stt_dim, "rate of strain for fixed particles");
      if (fl->ff) V3 ("Ff", t, q, vec_dim, "forces for fixed particles");
      if (fl->tf) V3 ("Tf", t, q, vec_dim, "torques for fixed particles");
      if (fl->sf) V3 ("Sf", t, q, stt_dim, "stresslets for fixed particles");
    }
  if (flag_rng)
    {
      int d2[2] = {t, rng_dim};
      def_var (nc, "mt", NCT_INT, 2, d2, "random number generator table");
      def_var (nc, "mti", NCT_INT, 1, &t, "random number generator index");
      def_var (nc, "mt_Ghs", NCT_INT, 1, &t, "Gaussian random number generator flag");
      V1 ("mt_Gs", t, "Gaussian random number generator saved value");
    }
#undef V1
#undef V2
#undef V3
  end_def (nc);

  /* (0, 0, 0): a non-periodic system until stokes_nc_set_l */
  const double zero[3] = {0.0, 0.0, 0.0};
  put_fixed (nc, "l", zero, NULL);
  const int nidx = nm > 624 ? nm : 624;
  const int nmax = nidx > nf ? nidx : nf;
  int *idx = malloc (sizeof (int) * (size_t) nmax);
  if (!idx) nc_die (filename, "allocation failed", NULL);
  for (int i = 0; i < nmax; i++) idx[i] = i;
  put_fixed (nc, "p", NULL, idx);
  put_fixed (nc, "vec", NULL, idx);
  put_fixed (nc, "stt", NULL, idx);
  put_fixed (nc, "quat", NULL, idx);
  if (nf > 0) put_fixed (nc, "pf", NULL, idx);
  if (flag_rng) put_fixed (nc, "rng", NULL, idx);
  free (idx);
  put_fixed (nc, "shear_mode", NULL, &shear_mode);
  return nc_public (nc);
}

/* the public struct: sizes, this file's dimension and variable indices in
 * the reference's *_dim / *_id fields, and a flag per data variable */
typedef struct {
  const char *name;
  size_t id;                    /* offset of the *_id field */
  long flag;                    /* offset of the flag_* field, or -1 */
} nc_field;

#define FD(nm, f) {nm, offsetof (struct stokes_nc, f##_id), \
                   (long) offsetof (struct stokes_nc, flag_##f)}
#define FA(nm, f) {nm, offsetof (struct stokes_nc, f##_id), -1}
static const nc_field nc_fields[] = {
  FA ("p", p), FA ("pf", pf), FA ("vec", vec), FA ("stt", stt),
  FA ("quat", quat), FA ("rng", rng), FA ("time", time), FA ("l", l),
  FA ("shear_mode", shear_mode), FA ("shear_rate", shear_rate),
  FA ("shear_shift", shear_shift), FA ("mt", mt), FA ("mti", mti),
  FA ("mt_Ghs", mt_Ghs), FA ("mt_Gs", mt_Gs),
  FD ("Ui0", ui0), FD ("Oi0", oi0), FD ("Ei0", ei0), FD ("Ui", ui),
  FD ("Oi", oi), FD ("Ei", ei), FD ("x0", x0), FD ("U0", u0), FD ("O0", o0),
  FD ("E0", e0), FD ("F0", f0), FD ("T0", t0), FD ("S0", s0),
  FD ("xf0", xf0), FD ("Uf0", uf0), FD ("Of0", of0), FD ("Ef0", ef0),
  FD ("Ff0", ff0), FD ("Tf0", tf0), FD ("Sf0", sf0), FD ("x", x),
  FD ("U", u), FD ("O", o), FD ("E", e), FD ("F", f), FD ("T", t),
  FD ("S", s), FD ("q", q), FD ("xf", xf), FD ("Uf", uf), FD ("Of", of),
  FD ("Ef", ef), FD ("Ff", ff), FD ("Tf", tf), FD ("Sf", sf), FD ("a", a),
  FD ("af", af),
};
#undef FD
#undef FA

static int
dim_of (const nc_file *f, const char *name)
{
  for (int k = 0; k < f->ndims; k++)
    if (strcmp (f->dims[k].name, name) == 0) return k;
  return -1;
}

static int
var_index (const nc_file *f, const char *name)
{
  for (int k = 0; k < f->nvars; k++)
    if (strcmp (f->vars[k].name, name) == 0) return k;
  return -1;
}

static struct stokes_nc *
nc_public (nc_file *f)
{
  struct stokes_nc *nc = calloc (1, sizeof (*nc));
  if (!nc) nc_die (f->fname, "allocation failed", NULL);
  nc->priv = f;
  nc->id = fileno (f->fp);
  const int pd = dim_of (f, "p"), qd = dim_of (f, "pf"), rd = dim_of (f, "rng");
  nc->p_dim = pd;
  nc->pf_dim = qd;
  nc->vec_dim = dim_of (f, "vec");
  nc->stt_dim = dim_of (f, "stt");
  nc->quat_dim = dim_of (f, "quat");
  nc->rng_dim = rd;
  nc->time_dim = dim_of (f, "time");
  nc->np = pd >= 0 ? f->dims[pd].len : 0;
  nc->npf = qd >= 0 ? f->dims[qd].len : 0;
  nc->nvec = 3;
  nc->nstt = 5;
  nc->nquat = 4;
  nc->nrng = rd >= 0 ? f->dims[rd].len : 0;
  nc->flag_rng = rd >= 0;
  nc->ntime = f->numrecs;
  f->ntime = &nc->ntime;
  for (size_t k = 0; k < sizeof (nc_fields) / sizeof (nc_fields[0]); k++)
    {
      const int v = var_index (f, nc_fields[k].name);
      *(int *) ((char *) nc + nc_fields[k].id) = v;
      if (nc_fields[k].flag >= 0)
        *(int *) ((char *) nc + nc_fields[k].flag) = v >= 0;
    }
  return nc;
}

struct stokes_nc *
stokes_nc_x_init (const char *filename, int np)
{
  nc_flags fl;
  memset (&fl, 0, sizeof (fl));
  fl.x = 1;
  return nc_create (filename, np, 0, &fl, 0, 0);
}

/* the flag table of stokes-nc.c:1445-1598 */
struct stokes_nc *
stokes_nc_init (const char *filename, int np, int nf, int version,
                int flag_poly, int flag_Q, int flag_it, int shear_mode,
                int flag_rng)
{
  nc_flags fl;
  memset (&fl, 0, sizeof (fl));
  if (flag_it == 0) fl.ui0 = fl.oi0 = fl.ei0 = 1;
  else fl.ui = fl.oi = fl.ei = 1;
  fl.q = flag_Q;
  fl.a = flag_poly;
  fl.f0 = fl.x = fl.u = 1;
  if (version >= 1) fl.t0 = fl.o = 1;
  if (version == 2) fl.e0 = fl.s = 1;
  if (nf > 0)
    {
      fl.af = flag_poly;
      fl.xf0 = fl.uf0 = fl.ff = 1;
      if (version >= 1) fl.of0 = fl.tf = 1;
      if (version == 2) fl.ef0 = fl.sf = 1;
    }
  return nc_create (filename, np, nf, &fl, shear_mode, flag_rng);
}

void
stokes_nc_free (struct stokes_nc *nc)
{
  if (!nc) return;
  nc_file *f = NCF (nc);
  if (fclose (f->fp) != 0) nc_die (f->fname, "close failed", NULL);
  free (f);
  free (nc);
}

/* ---- setters (stokes-nc.c:1625-2645) ---- */

void
stokes_nc_set_shear_rate (struct stokes_nc *nc, double shear_rate)
{
  put_fixed (NCF (nc), "shear_rate", &shear_rate, NULL);
}

#define FIXED(fn, var)                                                     \
  void fn (struct stokes_nc *nc, const double *v)                         \
  { put_fixed (NCF (nc), var, v, NULL); }
#define RECORD(fn, var)                                                    \
  void fn (struct stokes_nc *nc, int step, const double *v)                \
  { put_rec (NCF (nc), var, step, v, NULL); }

FIXED (stokes_nc_set_l, "l")
FIXED (stokes_nc_set_ui0, "Ui0")
FIXED (stokes_nc_set_oi0, "Oi0")
FIXED (stokes_nc_set_ei0, "Ei0")
FIXED (stokes_nc_set_x0, "x0")
FIXED (stokes_nc_set_u0, "U0")
FIXED (stokes_nc_set_o0, "O0")
FIXED (stokes_nc_set_e0, "E0")
FIXED (stokes_nc_set_f0, "F0")
FIXED (stokes_nc_set_t0, "T0")
FIXED (stokes_nc_set_s0, "S0")
FIXED (stokes_nc_set_xf0, "xf0")
FIXED (stokes_nc_set_uf0, "Uf0")
FIXED (stokes_nc_set_of0, "Of0")
FIXED (stokes_nc_set_ef0, "Ef0")
FIXED (stokes_nc_set_ff0, "Ff0")
FIXED (stokes_nc_set_tf0, "Tf0")
FIXED (stokes_nc_set_sf0, "Sf0")
FIXED (stokes_nc_set_a, "a")
FIXED (stokes_nc_set_af, "af")
RECORD (stokes_nc_set_ui, "Ui")
RECORD (stokes_nc_set_oi, "Oi")
RECORD (stokes_nc_set_ei, "Ei")
RECORD (stokes_nc_set_x, "x")
RECORD (stokes_nc_set_q, "q")
RECORD (stokes_nc_set_u, "U")
RECORD (stokes_nc_set_o, "O")
RECORD (stokes_nc_set_e, "E")
RECORD (stokes_nc_set_f, "F")
RECORD (stokes_nc_set_t, "T")
RECORD (stokes_nc_set_s, "S")
RECORD (stokes_nc_set_xf, "xf")
RECORD (stokes_nc_set_uf, "Uf")
RECORD (stokes_nc_set_of, "Of")
RECORD (stokes_nc_set_ef, "Ef")
RECORD (stokes_nc_set_ff, "Ff")
RECORD (stokes_nc_set_tf, "Tf")
RECORD (stokes_nc_set_sf, "Sf")

void
stokes_nc_set_time (struct stokes_nc *nc, int step, double time)
{
  put_rec (NCF (nc), "time", step, &time, NULL);
}

void
stokes_nc_set_shear_shift (struct stokes_nc *nc, int step, double shear_shift)
{
  put_rec (NCF (nc), "shear_shift", step, &shear_shift, NULL);
}

/* the whole MT19937 state (the reference stores half of it on LP64,
 * stokes-nc.c:2667-2671): the 624 words as their 32-bit patterns, mti, and
 * the polar method's saved-value flag and value */
void
stokes_nc_set_rng_state (struct stokes_nc *pub, int step, const uint32_t *mt,
                         int mti, int gauss_has_saved, double gauss_saved)
{
  nc_file *nc = NCF (pub);
  int *w = malloc (sizeof (int) * 624);
  if (!w) nc_die (nc->fname, "allocation failed", "mt");
  for (int k = 0; k < 624; k++) memcpy (&w[k], &mt[k], 4);
  put_rec (nc, "mt", step, NULL, w);
  free (w);
  put_rec (nc, "mti", step, NULL, &mti);
  put_rec (nc, "mt_Ghs", step, NULL, &gauss_has_saved);
  put_rec (nc, "mt_Gs", step, &gauss_saved, NULL);
}

/* ---- reading (stokes-nc-read.c) ---- */

static uint32_t
rd32 (nc_file *f)
{
  unsigned char t[4];
  if (fread (t, 1, 4, f->fp) != 4) nc_die (f->fname, "truncated header", NULL);
  return ((uint32_t) t[0] << 24) | ((uint32_t) t[1] << 16)
    | ((uint32_t) t[2] << 8) | t[3];
}

static void
rdname (nc_file *f, char *out, size_t cap)
{
  uint32_t n = rd32 (f);
  size_t pad = (4 - n % 4) % 4;
  char *b = malloc (n + pad + 1);
  if (!b || fread (b, 1, n + pad, f->fp) != n + pad)
    nc_die (f->fname, "truncated header", NULL);
  b[n] = 0;
  snprintf (out, cap, "%s", b);
  free (b);
}

static size_t
type_size (int type)
{
  return type == 1 || type == 2 ? 1 : (type == 3 ? 2 : (type == 6 ? 8 : 4));
}

/* an attribute list: skipped (names and values are not needed here) */
static void
skip_atts (nc_file *f)
{
  uint32_t tag = rd32 (f), n = rd32 (f);
  if (tag == 0 && n == 0) return;
  if (tag != NC_TAG_ATT) nc_die (f->fname, "bad attribute list", NULL);
  for (uint32_t k = 0; k < n; k++)
    {
      char nm[256];
      rdname (f, nm, sizeof (nm));
      const int type = (int) rd32 (f);
      const size_t bytes = rd32 (f) * type_size (type);
      if (fseeko (f->fp, (off_t) ((bytes + 3) / 4 * 4), SEEK_CUR) != 0)
        nc_die (f->fname, "truncated header", nm);
    }
}

static struct stokes_nc *
nc_open (const char *filename, int writable)
{
  nc_file *f = calloc (1, sizeof (*f));
  if (!f) nc_die (filename, "allocation failed", NULL);
  snprintf (f->fname, sizeof (f->fname), "%s", filename);
  f->fp = fopen (filename, writable ? "r+b" : "rb");
  if (!f->fp) nc_die (filename, "cannot open", NULL);
  unsigned char magic[4];
  if (fread (magic, 1, 4, f->fp) != 4 || memcmp (magic, "CDF", 3) != 0
      || (magic[3] != 1 && magic[3] != 2))
    nc_die (filename, "not a netCDF classic or 64-bit-offset file", NULL);
  const int off64 = magic[3] == 2;
  f->numrecs = (int) rd32 (f);
  uint32_t tag = rd32 (f), n = rd32 (f);
  if (tag == NC_TAG_DIM)
    for (uint32_t k = 0; k < n; k++)
      {
        if (f->ndims >= NC_MAXDIM) nc_die (filename, "too many dimensions", NULL);
        nc_dim *d = &f->dims[f->ndims++];
        rdname (f, d->name, sizeof (d->name));
        d->len = (int) rd32 (f);
      }
  skip_atts (f);                                /* global attributes */
  tag = rd32 (f);
  n = rd32 (f);
  if (tag == NC_TAG_VAR)
    for (uint32_t k = 0; k < n; k++)
      {
        if (f->nvars >= NC_MAXVAR) nc_die (filename, "too many variables", NULL);
        nc_var *v = &f->vars[f->nvars++];
        memset (v, 0, sizeof (*v));
        rdname (f, v->name, sizeof (v->name));
        v->nd = (int) rd32 (f);
        if (v->nd > 3) nc_die (filename, "variable of more than 3 dimensions", v->name);
        v->count = 1;
        for (int j = 0; j < v->nd; j++)
          {
            v->dim[j] = (int) rd32 (f);
            if (v->dim[j] < 0 || v->dim[j] >= f->ndims)
              nc_die (filename, "bad dimension id", v->name);
            if (f->dims[v->dim[j]].len == 0) v->rec = j == 0;
            else v->count *= (size_t) f->dims[v->dim[j]].len;
          }
        skip_atts (f);
        v->type = (int) rd32 (f);
        v->vsize = rd32 (f);
        if (off64)
          {
            uint64_t hi = rd32 (f), lo = rd32 (f);
            v->begin = (long long) ((hi << 32) | lo);
          }
        else
          v->begin = (long long) rd32 (f);
      }
  for (int k = 0; k < f->nvars; k++)
    if (f->vars[k].rec) f->recsize += (long long) f->vars[k].vsize;
  return nc_public (f);
}

struct stokes_nc *
stokes_nc_open (const char *filename)
{
  return nc_open (filename, 0);
}

struct stokes_nc *
stokes_nc_reopen (const char *filename)
{
  return nc_open (filename, 1);
}

/* count values of v at offset as doubles (NC_INT, NC_FLOAT, NC_DOUBLE) */
static void
read_values (nc_file *f, const nc_var *v, long long off, double *x, int *iv)
{
  const size_t ts = type_size (v->type), bytes = v->count * ts;
  unsigned char *b = malloc (bytes ? bytes : 1);
  if (!b) nc_die (f->fname, "allocation failed", v->name);
  if (fseeko (f->fp, (off_t) off, SEEK_SET) != 0
      || fread (b, 1, bytes, f->fp) != bytes)
    nc_die (f->fname, "read failed", v->name);
  for (size_t k = 0; k < v->count; k++)
    {
      const unsigned char *q = b + ts * k;
      uint32_t w = ((uint32_t) q[0] << 24) | ((uint32_t) q[1] << 16)
        | ((uint32_t) q[2] << 8) | q[3];
      double val;
      if (v->type == NCT_DOUBLE)
        {
          uint64_t u = ((uint64_t) w << 32)
            | ((uint32_t) q[4] << 24) | ((uint32_t) q[5] << 16)
            | ((uint32_t) q[6] << 8) | q[7];
          memcpy (&val, &u, 8);
        }
      else if (v->type == 5)
        {
          float fl;
          memcpy (&fl, &w, 4);
          val = fl;
        }
      else if (v->type == NCT_INT)
        {
          int32_t i;
          memcpy (&i, &w, 4);
          val = i;
          if (iv) iv[k] = i;
        }
      else
        nc_die (f->fname, "unsupported variable type", v->name);
      if (x) x[k] = val;
    }
  free (b);
}

/* a variable of the reference's reading functions by name, with the
 * shape the function takes (rec: a record variable; nd: its dimensions) */
static const nc_var *
want_var (nc_file *f, const char *name, int rec, int nd, const char *who)
{
  const int k = var_index (f, name);
  if (k < 0) nc_die (f->fname, "variable not in this file", name);
  const nc_var *v = &f->vars[k];
  if (v->rec != rec || v->nd != nd)
    {
      fprintf (stderr, "invalid name %s for %s()\n", name, who);
      return NULL;
    }
  return v;
}

void
stokes_nc_get_array1d (const struct stokes_nc *nc, const char *name, double *x)
{
  nc_file *f = NCF (nc);
  static const char *ok[] = {"Ui0", "Oi0", "Ei0", "Ui", "Oi", "Ei", "a", "af", "l"};
  int found = 0;
  for (size_t k = 0; k < sizeof (ok) / sizeof (ok[0]); k++)
    found |= strcmp (ok[k], name) == 0;
  if (!found)
    {
      fprintf (stderr, "invalid name for stokes_nc_get_array1d()\n"
               "which is for Ui0, Oi0, Ei0, Ui, Oi, Ei, a, af, or l.\n");
      return;
    }
  const int k = var_index (f, name);
  if (k < 0) nc_die (f->fname, "variable not in this file", name);
  /* Ui, Oi, Ei: their first record */
  read_values (f, &f->vars[k], f->vars[k].begin, x, NULL);
}

void
stokes_nc_get_data0 (const struct stokes_nc *nc, const char *name, double *x)
{
  nc_file *f = NCF (nc);
  const nc_var *v = want_var (f, name, 0, 2, "stokes_nc_get_data0");
  if (v) read_values (f, v, v->begin, x, NULL);
}

void
stokes_nc_get_data (const struct stokes_nc *nc, const char *name, int step,
                    double *x)
{
  nc_file *f = NCF (nc);
  const nc_var *v = want_var (f, name, 1, 3, "stokes_nc_get_data");
  if (!v) return;
  if (step < 0 || step >= f->numrecs) nc_die (f->fname, "step out of range", name);
  read_values (f, v, v->begin + (long long) step * f->recsize, x, NULL);
}

void
stokes_nc_get_time (const struct stokes_nc *nc, double *time)
{
  for (int s = 0; s < nc->ntime; s++) time[s] = stokes_nc_get_time_step (nc, s);
}

double
stokes_nc_get_time_step (const struct stokes_nc *nc, int step)
{
  nc_file *f = NCF (nc);
  const nc_var *v = want_var (f, "time", 1, 1, "stokes_nc_get_time_step");
  if (step < 0 || step >= f->numrecs) nc_die (f->fname, "step out of range", "time");
  double t = 0.0;
  read_values (f, v, v->begin + (long long) step * f->recsize, &t, NULL);
  return t;
}

void
stokes_nc_get_rng_state (const struct stokes_nc *nc, int step, uint32_t *mt,
                         int *mti, int *gauss_has_saved, double *gauss_saved)
{
  nc_file *f = NCF (nc);
  if (step < 0 || step >= f->numrecs) nc_die (f->fname, "step out of range", "mt");
  const long long r = (long long) step * f->recsize;
  const nc_var *v = want_var (f, "mt", 1, 2, "stokes_nc_get_rng_state");
  int *w = malloc (sizeof (int) * v->count);
  if (!w) nc_die (f->fname, "allocation failed", "mt");
  read_values (f, v, v->begin + r, NULL, w);
  for (size_t k = 0; k < v->count; k++) memcpy (&mt[k], &w[k], 4);
  free (w);
  v = want_var (f, "mti", 1, 1, "stokes_nc_get_rng_state");
  read_values (f, v, v->begin + r, NULL, mti);
  v = want_var (f, "mt_Ghs", 1, 1, "stokes_nc_get_rng_state");
  read_values (f, v, v->begin + r, NULL, gauss_has_saved);
  v = want_var (f, "mt_Gs", 1, 1, "stokes_nc_get_rng_state");
  read_values (f, v, v->begin + r, gauss_saved, NULL);
}

/* ---- BD parameter files (stokes-nc.c:52-115, 2731-3540) ---- */

/* the scalar int / double variable name of the file (0 / 0.0 and a message
 * when the file has no such variable) */
static int
scalar_int (const struct stokes_nc *nc, const char *name, const char *who)
{
  nc_file *f = NCF (nc);
  const int k = var_index (f, name);
  if (k < 0)
    {
      fprintf (stderr, "at reading %s in %s\n", name, who);
      return 0;
    }
  int v = 0;
  read_values (f, &f->vars[k], f->vars[k].begin, NULL, &v);
  return v;
}

static double
scalar_double (const struct stokes_nc *nc, const char *name, const char *who)
{
  nc_file *f = NCF (nc);
  const int k = var_index (f, name);
  if (k < 0)
    {
      fprintf (stderr, "at reading %s in %s\n", name, who);
      return 0.0;
    }
  double v = 0.0;
  read_values (f, &f->vars[k], f->vars[k].begin, &v, NULL);
  return v;
}

void
stokes_nc_print_actives (struct stokes_nc *nc, FILE *out)
{
  fprintf (out, "    | x | u o e | f t s | ui oi ei\n");
  fprintf (out, "----+---+-------+-------+----------\n");
  fprintf (out, "x0  | %d | %d %d %d | %d %d %d | %d  %d  %d\n", nc->flag_x0, nc->flag_u0,
           nc->flag_o0, nc->flag_e0, nc->flag_f0, nc->flag_t0, nc->flag_s0, nc->flag_ui0,
           nc->flag_oi0, nc->flag_ei0);
  fprintf (out, "xf0 | %d | %d %d %d | %d %d %d |\n", nc->flag_xf0, nc->flag_uf0,
           nc->flag_of0, nc->flag_ef0, nc->flag_ff0, nc->flag_tf0, nc->flag_sf0);
  fprintf (out, "----+---+-------+-------+----------\n");
  fprintf (out, "x   | %d | %d %d %d | %d %d %d | %d  %d  %d\n", nc->flag_x, nc->flag_u,
           nc->flag_o, nc->flag_e, nc->flag_f, nc->flag_t, nc->flag_s, nc->flag_ui,
           nc->flag_oi, nc->flag_ei);
  fprintf (out, "xf  | %d | %d %d %d | %d %d %d |\n", nc->flag_xf, nc->flag_uf, nc->flag_of,
           nc->flag_ef, nc->flag_ff, nc->flag_tf, nc->flag_sf);
  fprintf (out, "----+---+-------+-------+----------\n");
  fprintf (out, "q   | %d |\n", nc->flag_q);
  fprintf (out, "a   | %d |\n", nc->flag_a);
  fprintf (out, "af  | %d |\n", nc->flag_af);
  fprintf (out, "----+---+-------+------------------\n");
  fprintf (out, "shear_mode      | %d\n",
           scalar_int (nc, "shear_mode", "stokes_nc_print_actives"));
  fprintf (out, "----------------+------------------\n");
  fprintf (out, "states of RNG   | %d\n", nc->flag_rng);
}

struct stokes_nc *
stokes_nc_set_by_params (const char *out_file, const struct stokes *sys, int flag_Q,
                         const double *Ui, const double *Oi, const double *Ei,
                         const double *F, const double *T, const double *E,
                         const double *uf, const double *of, const double *ef,
                         const double *xf, const double *lat, int shear_mode,
                         double shear_rate, int flag_rng)
{
  const int nm = sys->nm, nf = sys->np - nm;
  const int poly = sys->a != NULL;
  /* constant imposed flow (Ui0, Oi0, Ei0) */
  struct stokes_nc *nc = stokes_nc_init (out_file, nm, nf, sys->version, poly, flag_Q, 0,
                                         shear_mode, flag_rng);
  stokes_nc_set_ui0 (nc, Ui);
  stokes_nc_set_oi0 (nc, Oi);
  stokes_nc_set_ei0 (nc, Ei);
  if (sys->periodic == 1) stokes_nc_set_l (nc, lat);
  if (poly) stokes_nc_set_a (nc, sys->a);
  stokes_nc_set_f0 (nc, F);
  if (sys->version >= 1) stokes_nc_set_t0 (nc, T);
  if (sys->version == 2) stokes_nc_set_e0 (nc, E);
  if (nf > 0)
    {
      if (poly) stokes_nc_set_af (nc, sys->a + nm);
      stokes_nc_set_uf0 (nc, uf);
      if (sys->version >= 1) stokes_nc_set_of0 (nc, of);
      if (sys->version == 2) stokes_nc_set_ef0 (nc, ef);
      stokes_nc_set_xf0 (nc, xf);
    }
  if (shear_mode != 0) stokes_nc_set_shear_rate (nc, shear_rate);
  return nc;
}

/* 1 if some |x[i] - y[i]| >= tiny */
static int
differ (const double *x, const double *y, int n, double tiny)
{
  for (int i = 0; i < n; i++)
    if (fabs (x[i] - y[i]) >= tiny) return 1;
  return 0;
}

int
stokes_nc_check_params (const struct stokes_nc *nc, const struct stokes *sys, int flag_Q,
                        const double *Ui, const double *Oi, const double *Ei,
                        const double *F, const double *T, const double *E,
                        const double *uf, const double *of, const double *ef,
                        const double *xf, const double *lat, double tiny)
{
  (void) flag_Q;
  const int nm = sys->nm, nf = sys->np - nm, v = sys->version;
  int bad = 0;
  if (nm != nc->np) { fprintf (stderr, "mismatch np %d != %d\n", nm, nc->np); bad++; }
  if (nf != nc->npf) { fprintf (stderr, "mismatch npf %d != %d\n", nf, nc->npf); bad++; }
  if (nc->flag_ui0 != 1 || nc->flag_oi0 != 1 || nc->flag_ei0 != 1)
    {
      fprintf (stderr, "imposed flows are not set (%d,%d,%d)\n", nc->flag_ui0, nc->flag_oi0,
               nc->flag_ei0);
      bad++;
    }
  /* the entries the version needs, without (nf == 0) or with fixed
   * particles (then also their radii flag) */
  int want = nc->flag_f0 == 1 && nc->flag_x == 1 && nc->flag_u == 1;
  if (v >= 1) want = want && nc->flag_t0 == 1 && nc->flag_o == 1;
  if (v == 2) want = want && nc->flag_e0 == 1 && nc->flag_s == 1;
  if (nf == 0)
    {
      if ((sys->a == NULL) != (nc->flag_a == 0)) { fprintf (stderr, "mismatch flag_a\n"); bad++; }
      if (!want)
        {
          fprintf (stderr, "%s params are not set properly.\n",
                   v == 0 ? "F" : (v == 1 ? "FT" : "FTS"));
          bad++;
        }
    }
  else
    {
      if ((sys->a == NULL) != (nc->flag_af == 0)) { fprintf (stderr, "mismatch flag_af\n"); bad++; }
      want = want && nc->flag_uf0 == 1 && nc->flag_xf0 == 1 && nc->flag_ff == 1;
      if (v >= 1) want = want && nc->flag_of0 == 1 && nc->flag_tf == 1;
      if (v == 2) want = want && nc->flag_ef0 == 1 && nc->flag_sf == 1;
      if (!want)
        {
          fprintf (stderr, "%s-fix params are not set properly.\n",
                   v == 0 ? "F" : (v == 1 ? "FT" : "FTS"));
          bad++;
        }
    }
  double c[5];
  stokes_nc_get_array1d (nc, "Ui0", c);
  if (differ (Ui, c, 3, tiny)) { fprintf (stderr, "Ui mismatch\n"); bad++; }
  stokes_nc_get_array1d (nc, "Oi0", c);
  if (differ (Oi, c, 3, tiny)) { fprintf (stderr, "Oi mismatch\n"); bad++; }
  stokes_nc_get_array1d (nc, "Ei0", c);
  if (differ (Ei, c, 5, tiny)) { fprintf (stderr, "Ei mismatch\n"); bad++; }
  if (sys->periodic == 1)
    {
      stokes_nc_get_array1d (nc, "l", c);
      if (differ (lat, c, 3, tiny)) { fprintf (stderr, "l mismatch\n"); bad++; }
    }
  if (sys->shear_mode != 0)
    {
      if (sys->shear_mode != scalar_int (nc, "shear_mode", "stokes_nc_check_params"))
        { fprintf (stderr, "shear_mode mismatch\n"); bad++; }
      c[0] = scalar_double (nc, "shear_rate", "stokes_nc_check_params");
      if (differ (&sys->shear_rate, c, 1, tiny)) { fprintf (stderr, "shear_rate mismatch\n"); bad++; }
    }
  double *w = malloc (sizeof (double) * 5 * (size_t) (sys->np > 0 ? sys->np : 1));
  if (!w) nc_die (NCF (nc)->fname, "allocation failed", NULL);
  if (sys->a != NULL)
    {
      stokes_nc_get_array1d (nc, "a", w);
      if (differ (sys->a, w, nm, tiny)) { fprintf (stderr, "a mismatch\n"); bad++; }
    }
  stokes_nc_get_data0 (nc, "F0", w);
  if (differ (F, w, 3 * nm, tiny)) { fprintf (stderr, "F0 mismatch\n"); bad++; }
  if (v >= 1)
    {
      stokes_nc_get_data0 (nc, "T0", w);
      if (differ (T, w, 3 * nm, tiny)) { fprintf (stderr, "T0 mismatch\n"); bad++; }
    }
  if (v == 2)
    {
      stokes_nc_get_data0 (nc, "E0", w);
      if (differ (E, w, 5 * nm, tiny)) { fprintf (stderr, "E0 mismatch\n"); bad++; }
    }
  if (nf > 0)
    {
      if (sys->a != NULL)
        {
          stokes_nc_get_array1d (nc, "af", w);
          if (differ (sys->a + nm, w, nf, tiny)) { fprintf (stderr, "af mismatch\n"); bad++; }
        }
      stokes_nc_get_data0 (nc, "Uf0", w);
      if (differ (uf, w, 3 * nf, tiny)) { fprintf (stderr, "Uf0 mismatch\n"); bad++; }
      if (v >= 1)
        {
          stokes_nc_get_data0 (nc, "Of0", w);
          if (differ (of, w, 3 * nf, tiny)) { fprintf (stderr, "Of0 mismatch\n"); bad++; }
        }
      if (v == 2)
        {
          stokes_nc_get_data0 (nc, "Ef0", w);
          if (differ (ef, w, 5 * nf, tiny)) { fprintf (stderr, "Ef0 mismatch\n"); bad++; }
        }
      stokes_nc_get_data0 (nc, "xf0", w);
      if (differ (xf, w, 3 * nf, tiny)) { fprintf (stderr, "xf0 mismatch\n"); bad++; }
    }
  free (w);
  return bad;
}

/* the version from the flags: exactly the entries stokes_nc_init defines
 * for F, FT or FTS (with or without fixed particles), else 0 */
int
stokes_nc_get_version (const struct stokes_nc *nc)
{
  const int fix = nc->npf != 0;
  for (int v = 2; v >= 0; v--)
    {
      const int t = v >= 1, s = v == 2;
      if (nc->flag_f0 == 1 && nc->flag_t0 == t && nc->flag_e0 == s && nc->flag_x == 1
          && nc->flag_u == 1 && nc->flag_o == t && nc->flag_s == s && nc->flag_uf0 == fix
          && nc->flag_of0 == (fix && t) && nc->flag_ef0 == (fix && s) && nc->flag_xf0 == fix
          && nc->flag_ff == fix && nc->flag_tf == (fix && t) && nc->flag_sf == (fix && s))
        return v;
    }
  return 0;
}

void
stokes_nc_get_params (const struct stokes_nc *nc, struct stokes *sys, int *flag_Q,
                      double *Ui, double *Oi, double *Ei, double *F, double *T, double *E,
                      double *uf, double *of, double *ef, double *xf, double *lat,
                      int *shear_mode, double *shear_rate, int *flag_rng)
{
  *flag_Q = nc->flag_q;
  *flag_rng = nc->flag_rng;
  sys->version = stokes_nc_get_version (nc);
  const int nm = nc->np, nf = nc->npf, np = nm + nf;
  stokes_set_np (sys, np, nm);
  if (nc->flag_a != 0)
    {
      double *a = malloc (sizeof (double) * (size_t) (np > 0 ? np : 1));
      if (!a) nc_die (NCF (nc)->fname, "allocation failed", NULL);
      stokes_nc_get_array1d (nc, "a", a);
      if (nc->flag_af != 0) stokes_nc_get_array1d (nc, "af", a + nm);
      stokes_set_radius (sys, a);
      free (a);
    }
  /* the lattice is (0, 0, 0) until set: an open system */
  stokes_nc_get_array1d (nc, "l", lat);
  if (lat[0] == 0.0 && lat[1] == 0.0 && lat[2] == 0.0)
    sys->periodic = 0;
  else
    {
      sys->periodic = 1;
      stokes_set_l (sys, lat[0], lat[1], lat[2]);
    }
  stokes_nc_get_array1d (nc, "Ui0", Ui);
  stokes_nc_get_array1d (nc, "Oi0", Oi);
  stokes_nc_get_array1d (nc, "Ei0", Ei);
  stokes_nc_get_data0 (nc, "F0", F);
  if (sys->version >= 1) stokes_nc_get_data0 (nc, "T0", T);
  if (sys->version == 2) stokes_nc_get_data0 (nc, "E0", E);
  if (nf > 0)
    {
      stokes_nc_get_data0 (nc, "xf0", xf);
      stokes_nc_get_data0 (nc, "Uf0", uf);
      if (sys->version >= 1) stokes_nc_get_data0 (nc, "Of0", of);
      if (sys->version == 2) stokes_nc_get_data0 (nc, "Ef0", ef);
    }
  *shear_mode = scalar_int (nc, "shear_mode", "stokes_nc_get_params");
  sys->shear_mode = *shear_mode;
  if (sys->shear_mode != 0)
    {
      *shear_rate = scalar_double (nc, "shear_rate", "stokes_nc_get_params");
      sys->shear_rate = *shear_rate;
    }
}

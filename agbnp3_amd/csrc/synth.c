/* Seeded synthetic globular solutes (the globular-solute recipe); see synth.h.
 *
 * Growth model
 *   backbone : self-avoiding chain, bond 1.53 A, bond angle U[104.5,114.5] deg,
 *              dihedral U[0,360); every backbone atom stays inside a sphere of
 *              radius Rs = (3N / (4 pi 0.1))^(1/3).
 *   pendants : placed together with the next backbone atom, once both bonds of
 *              their parent exist (carbonyl O on the bisector, hydroxyl O and
 *              sulfone O pairs on the sp3 directions).
 *   clashes  : non-bonded heavy-heavy pairs (1-2 and 1-3 pairs excluded) must be
 *              >= 2.5 A apart; checked with a cell hash.  A blocked step pops a
 *              few placements and retries; persistent jams inflate Rs by 2%.
 *   hydrogens: 1.0 A from the parent on ideal sp3 / sp2 directions.
 * Composition targets (all atoms): C 32%, N 8%, O 9%, S 1%, H ~50%.
 */
#include "synth.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BOND_HH   1.53
#define BOND_XH   1.00
#define MIN_NB    2.5
#define DENSITY   0.1
#define TETRA     1.9106332362490186   /* 109.47 deg */
#define HALF_TETRA 0.9553166181245093  /* 54.74 deg */

enum { EL_C = 0, EL_N, EL_O, EL_S, EL_H };
static const double el_radius[5] = {1.7, 1.55, 1.52, 1.8, 1.2};

/* hydrogen-bond type codes of the AGBNP3 interface */
enum { HB_INACTIVE = 0, HB_POLARH = 1, HB_TRIGONAL = 10, HB_TRIGONAL_OOP = 13,
       HB_TETRAHEDRAL = 20, HB_TRIGONAL_S = 30 };

/* backbone motifs */
enum { M_CH2 = 0, M_CH_SP2, M_C_CARBONYL, M_C_HYDROXYL, M_NH, M_N_SP2, M_O_ETHER,
       M_S_THIO, M_S_SULFONE, M_CH3_END };

/* ------------------------------------------------------------------ RNG */
typedef struct { uint64_t s[4]; int has_spare; double spare; } Rng;

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static void rng_seed(Rng *g, uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; i++) g->s[i] = splitmix64(&x);
  g->has_spare = 0;
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t rng_next(Rng *g) {
  uint64_t *s = g->s;
  uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
  s[2] ^= t; s[3] = rotl(s[3], 45);
  return r;
}
static double rng_u01(Rng *g) { return (double)(rng_next(g) >> 11) * (1.0 / 9007199254740992.0); }
static double rng_uniform(Rng *g, double a, double b) { return a + (b - a) * rng_u01(g); }
static double rng_normal(Rng *g) {
  if (g->has_spare) { g->has_spare = 0; return g->spare; }
  double u1, u2;
  do { u1 = rng_u01(g); } while (u1 <= 0.0);
  u2 = rng_u01(g);
  double m = sqrt(-2.0 * log(u1));
  g->spare = m * sin(2.0 * M_PI * u2);
  g->has_spare = 1;
  return m * cos(2.0 * M_PI * u2);
}

/* --------------------------------------------------------------- vectors */
typedef struct { double v[3]; } V3;
static V3 v3(double a, double b, double c) { V3 r = {{a, b, c}}; return r; }
static V3 vadd(V3 a, V3 b) { return v3(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
static V3 vsub(V3 a, V3 b) { return v3(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
static V3 vscale(V3 a, double s) { return v3(a.v[0] * s, a.v[1] * s, a.v[2] * s); }
static double vdot(V3 a, V3 b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
static double vnorm(V3 a) { return sqrt(vdot(a, a)); }
static V3 vunit(V3 a) { double n = vnorm(a); return n > 0 ? vscale(a, 1.0 / n) : v3(1, 0, 0); }
static V3 vcross(V3 a, V3 b) {
  return v3(a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2],
            a.v[0] * b.v[1] - a.v[1] * b.v[0]);
}
/* orthonormal pair perpendicular to unit vector b */
static void vperp(V3 b, V3 *p, V3 *q) {
  V3 t = fabs(b.v[0]) < 0.9 ? v3(1, 0, 0) : v3(0, 1, 0);
  *p = vunit(vcross(b, t));
  *q = vcross(b, *p);
}
static V3 random_unit(Rng *g) {
  double z = rng_uniform(g, -1.0, 1.0), ph = rng_uniform(g, 0.0, 2.0 * M_PI);
  double s = sqrt(1.0 - z * z);
  return v3(s * cos(ph), s * sin(ph), z);
}

/* ------------------------------------------------------------ heavy atoms */
typedef struct {
  V3 pos;
  int el;
  int motif;      /* backbone motif, -1 for pendants */
  int parent;     /* pendant: backbone parent; backbone: -1 */
  int bb_index;   /* backbone position (backbone atoms), parent's for pendants */
  int nh;         /* hydrogens to attach */
  int hb;         /* hydrogen-bond type */
  int pend[2];    /* pendant heavy atoms (backbone atoms) */
  int npend;
  int next_in_cell;
} Heavy;

typedef struct {
  Heavy *h;
  int n, cap;
  int *bb;        /* backbone atom heavy indices */
  int nbb;
  /* cell hash */
  double lo, csize, min_nb;
  int dim;
  int *head;
} Chain;

static int cell_of(const Chain *c, V3 p) {
  int ix[3];
  for (int d = 0; d < 3; d++) {
    int k = (int)floor((p.v[d] - c->lo) / c->csize);
    if (k < 0) k = 0;
    if (k >= c->dim) k = c->dim - 1;
    ix[d] = k;
  }
  return (ix[2] * c->dim + ix[1]) * c->dim + ix[0];
}

static void cells_reset(Chain *c, double rs) {
  c->csize = c->min_nb;
  c->lo = -rs - 4.0;
  c->dim = (int)ceil((2.0 * rs + 8.0) / c->csize) + 1;
  free(c->head);
  c->head = (int *)malloc(sizeof(int) * (size_t)c->dim * c->dim * c->dim);
  for (int i = 0; i < c->dim * c->dim * c->dim; i++) c->head[i] = -1;
  for (int i = 0; i < c->n; i++) {
    int k = cell_of(c, c->h[i].pos);
    c->h[i].next_in_cell = c->head[k];
    c->head[k] = i;
  }
}

/* push a heavy atom (always the newest); cells are LIFO so pops stay O(1) */
static int push_heavy(Chain *c, V3 p, int el, int motif, int parent, int bbi) {
  if (c->n == c->cap) {
    c->cap = c->cap * 2 + 64;
    c->h = (Heavy *)realloc(c->h, sizeof(Heavy) * c->cap);
  }
  Heavy *a = &c->h[c->n];
  memset(a, 0, sizeof(*a));
  a->pos = p; a->el = el; a->motif = motif; a->parent = parent; a->bb_index = bbi;
  int k = cell_of(c, p);
  a->next_in_cell = c->head[k];
  c->head[k] = c->n;
  return c->n++;
}
static void pop_heavy(Chain *c) {
  Heavy *a = &c->h[c->n - 1];
  int k = cell_of(c, a->pos);
  c->head[k] = a->next_in_cell;
  c->n--;
}

/* true if p is >= MIN_NB from every heavy atom not in skip[] */
static int clash_free(const Chain *c, V3 p, const int *skip, int nskip) {
  int ix[3];
  for (int d = 0; d < 3; d++) ix[d] = (int)floor((p.v[d] - c->lo) / c->csize);
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        int cx = ix[0] + dx, cy = ix[1] + dy, cz = ix[2] + dz;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= c->dim || cy >= c->dim || cz >= c->dim) continue;
        for (int j = c->head[(cz * c->dim + cy) * c->dim + cx]; j >= 0; j = c->h[j].next_in_cell) {
          int s = 0;
          for (int q = 0; q < nskip; q++) if (skip[q] == j) { s = 1; break; }
          if (s) continue;
          V3 d = vsub(c->h[j].pos, p);
          if (vdot(d, d) < c->min_nb * c->min_nb) return 0;
        }
      }
  return 1;
}

/* heavy atoms within 4 A of p */
static int crowding(const Chain *c, V3 p) {
  int ix[3], n = 0, rc = (int)ceil(4.0 / c->csize);
  for (int d = 0; d < 3; d++) ix[d] = (int)floor((p.v[d] - c->lo) / c->csize);
  for (int dz = -rc; dz <= rc; dz++)
    for (int dy = -rc; dy <= rc; dy++)
      for (int dx = -rc; dx <= rc; dx++) {
        int cx = ix[0] + dx, cy = ix[1] + dy, cz = ix[2] + dz;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= c->dim || cy >= c->dim || cz >= c->dim) continue;
        for (int j = c->head[(cz * c->dim + cy) * c->dim + cx]; j >= 0; j = c->h[j].next_in_cell) {
          V3 d = vsub(c->h[j].pos, p);
          if (vdot(d, d) < 16.0) n++;
        }
      }
  return n;
}

/* distance from p to the nearest heavy atom, capped at 2 cells */
static double nearest_heavy(const Chain *c, V3 p) {
  int ix[3];
  double best = 2.0 * c->csize;
  for (int d = 0; d < 3; d++) ix[d] = (int)floor((p.v[d] - c->lo) / c->csize);
  for (int dz = -2; dz <= 2; dz++)
    for (int dy = -2; dy <= 2; dy++)
      for (int dx = -2; dx <= 2; dx++) {
        int cx = ix[0] + dx, cy = ix[1] + dy, cz = ix[2] + dz;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= c->dim || cy >= c->dim || cz >= c->dim) continue;
        for (int j = c->head[(cz * c->dim + cy) * c->dim + cx]; j >= 0; j = c->h[j].next_in_cell) {
          double dd = vnorm(vsub(c->h[j].pos, p));
          if (dd < best) best = dd;
        }
      }
  return best;
}

static int motif_element(int m) {
  switch (m) {
    case M_NH: case M_N_SP2: return EL_N;
    case M_O_ETHER: return EL_O;
    case M_S_THIO: case M_S_SULFONE: return EL_S;
    default: return EL_C;
  }
}
/* hydrogens carried by the backbone atom itself (pendant H counted separately) */
static int motif_nh(int m) {
  switch (m) {
    case M_CH2: return 2;
    case M_CH_SP2: case M_C_HYDROXYL: case M_NH: return 1;
    case M_CH3_END: return 3;
    default: return 0;
  }
}
static int motif_npend(int m) {
  switch (m) {
    case M_C_CARBONYL: case M_C_HYDROXYL: return 1;
    case M_S_SULFONE: return 2;
    default: return 0;
  }
}
/* atoms (heavy + H) a motif adds, pendants and their H included */
static int motif_atoms(int m) {
  int n = 1 + motif_nh(m) + motif_npend(m);
  if (m == M_C_HYDROXYL) n += 1;
  return n;
}

static int choose_motif(Rng *g) {
  /* backbone element split C 0.751 / N 0.188 / O 0.0376 / S 0.0235 gives
     heavy-atom fractions C:N:O:S = 64:16:18:2 once pendant O are counted */
  double u = rng_u01(g);
  if (u < 0.751) {
    double v = rng_u01(g);
    if (v < 0.48) return M_CH2;
    if (v < 0.80) return M_CH_SP2;
    if (v < 0.92) return M_C_CARBONYL;
    return M_C_HYDROXYL;
  }
  if (u < 0.751 + 0.188) return rng_u01(g) < 0.7 ? M_NH : M_N_SP2;
  if (u < 0.751 + 0.188 + 0.0376) return M_O_ETHER;
  return rng_u01(g) < 0.5 ? M_S_THIO : M_S_SULFONE;
}

/* pendant directions for a backbone atom with bond unit vectors e1 (to prev)
   and e2 (to next) */
static int pendant_dirs(int motif, V3 e1, V3 e2, V3 out[2]) {
  V3 u = vunit(vscale(vadd(e1, e2), -1.0));
  V3 w = vunit(vcross(e1, e2));
  switch (motif) {
    case M_C_CARBONYL: out[0] = u; return 1;
    case M_C_HYDROXYL:
      out[0] = vadd(vscale(u, cos(HALF_TETRA)), vscale(w, sin(HALF_TETRA)));
      return 1;
    case M_S_SULFONE:
      out[0] = vadd(vscale(u, cos(HALF_TETRA)), vscale(w, sin(HALF_TETRA)));
      out[1] = vsub(vscale(u, cos(HALF_TETRA)), vscale(w, sin(HALF_TETRA)));
      return 2;
    default: return 0;
  }
}

/* hydrogen directions around an atom with m existing unit bond vectors */
static int hydrogen_dirs(Rng *g, const V3 *e, int m, int nh, V3 *out) {
  if (nh <= 0) return 0;
  if (m == 0) {                    /* isolated: spread around a random axis */
    V3 a = random_unit(g), p, q;
    vperp(a, &p, &q);
    out[0] = a;
    for (int k = 1; k < nh; k++) {
      double ph = 2.0 * M_PI * (k - 1) / 3.0;
      out[k] = vadd(vscale(a, cos(TETRA)),
                    vscale(vadd(vscale(p, cos(ph)), vscale(q, sin(ph))), sin(TETRA)));
    }
    return nh;
  }
  if (m == 1) {
    V3 p, q;
    vperp(e[0], &p, &q);
    double ph0 = rng_uniform(g, 0.0, 2.0 * M_PI);
    for (int k = 0; k < nh; k++) {
      double ph = ph0 + 2.0 * M_PI * k / 3.0;
      out[k] = vadd(vscale(e[0], cos(TETRA)),
                    vscale(vadd(vscale(p, cos(ph)), vscale(q, sin(ph))), sin(TETRA)));
    }
    return nh;
  }
  V3 s = v3(0, 0, 0);
  for (int k = 0; k < m; k++) s = vadd(s, e[k]);
  V3 u = vunit(vscale(s, -1.0));
  if (m == 2 && nh >= 2) {
    V3 w = vunit(vcross(e[0], e[1]));
    out[0] = vadd(vscale(u, cos(HALF_TETRA)), vscale(w, sin(HALF_TETRA)));
    out[1] = vsub(vscale(u, cos(HALF_TETRA)), vscale(w, sin(HALF_TETRA)));
    return 2;
  }
  out[0] = u;
  return 1;
}

/* ------------------------------------------------------------- generator */
/* non-bonded heavy-heavy exclusion distance (A) */
static double min_nb_env(void) {
  const char *e = getenv("AGBSYNTH_MIN_NB");
  return e ? atof(e) : MIN_NB;
}

int agbsynth_generate(uint64_t seed, int natoms,
                      double *x, double *y, double *z, double *r,
                      double *charge, double *igamma, double *sgamma,
                      double *ialpha, double *salpha,
                      int *hbtype, double *hbcorr, int *ishydrogen,
                      int *nne, int *neighl1, int neighl1_cap,
                      int *nconn_out, double *info) {
  if (natoms < 8 || !x || !y || !z || !r || !charge || !nne || !neighl1) return -1;
  Rng grow, par;
  rng_seed(&grow, seed);
  rng_seed(&par, seed ^ 0x5DEECE66DULL);

  double rs = cbrt(3.0 * natoms / (4.0 * M_PI * DENSITY));
  Chain c;
  memset(&c, 0, sizeof(c));
  c.min_nb = min_nb_env();
  c.bb = (int *)malloc(sizeof(int) * (size_t)natoms);
  cells_reset(&c, rs);

  long backtracks = 0, fails_in_row = 0, nsegments = 0;
  int total = 0;              /* atoms committed so far, hydrogens included */
  int *tot_at = (int *)malloc(sizeof(int) * (size_t)(natoms + 1)); /* total before bb k */
  unsigned char *seg_start = (unsigned char *)calloc((size_t)natoms + 1, 1);
  int seg_begin = 0;          /* backbone index where the current segment starts */
  int seg_best = 0;           /* longest extent the current segment reached */
  long stall = 0;             /* blocked steps since seg_best last grew */
  int need_segment = 1;
  int done = 0;

  while (!done) {
    if (need_segment) {
      /* start a chain segment (CH3 cap) at the emptiest of 256 random points */
      V3 best = v3(0, 0, 0);
      double bestd = -1.0;
      for (int t = 0; t < 256; t++) {
        V3 p = vscale(random_unit(&grow), rs * cbrt(rng_u01(&grow)));
        double dn = nearest_heavy(&c, p);
        if (dn > bestd) { bestd = dn; best = p; }
      }
      if (bestd < c.min_nb) {           /* sphere is full: inflate it */
        rs *= 1.02;
        cells_reset(&c, rs);
        continue;
      }
      int h = push_heavy(&c, best, EL_C, M_CH3_END, -1, c.nbb);
      seg_start[c.nbb] = 1;
      seg_begin = c.nbb;
      tot_at[c.nbb] = total;
      c.bb[c.nbb++] = h;
      total += motif_atoms(M_CH3_END);
      nsegments++;
      need_segment = 0;
      fails_in_row = 0;
      seg_best = c.nbb;
      stall = 0;
      if (total >= natoms) done = 1;
      continue;
    }
    int k = c.nbb;                       /* index of the backbone atom to place */
    int prev = c.bb[k - 1];
    int seglen = k - seg_begin;
    int is_last = (total + 4 >= natoms);
    int motif = is_last ? M_CH3_END : choose_motif(&grow);
    int el = motif_element(motif);
    /* configurational bias: among valid trials keep the least crowded */
    int nvalid = 0, best_crowd = 1 << 30;
    V3 best_pn = v3(0, 0, 0), best_pp[2];
    int best_npd = 0;
    for (int trial = 0; trial < 24; trial++) {
      V3 xp = c.h[prev].pos, d;
      if (seglen == 1) {
        d = random_unit(&grow);
      } else {
        V3 b = vunit(vsub(xp, c.h[c.bb[k - 2]].pos)), p, q;
        vperp(b, &p, &q);
        double th = rng_uniform(&grow, 104.5, 114.5) * M_PI / 180.0;
        double ph = rng_uniform(&grow, 0.0, 2.0 * M_PI);
        d = vadd(vscale(b, cos(M_PI - th)),
                 vscale(vadd(vscale(p, cos(ph)), vscale(q, sin(ph))), sin(M_PI - th)));
      }
      V3 pn = vadd(xp, vscale(d, BOND_HH));
      if (vnorm(pn) > rs) continue;
      int skip[2], ns = 0;
      skip[ns++] = prev;
      if (seglen >= 2) skip[ns++] = c.bb[k - 2];
      /* pendants of prev are placed in this same step */
      int npd = 0;
      V3 pdir[2], ppos[2];
      if (seglen >= 2) {
        V3 e1 = vunit(vsub(c.h[c.bb[k - 2]].pos, xp));
        V3 e2 = vunit(vsub(pn, xp));
        npd = pendant_dirs(c.h[prev].motif, e1, e2, pdir);
      }
      int ok = clash_free(&c, pn, skip, ns);
      for (int q = 0; q < npd && ok; q++) {
        ppos[q] = vadd(xp, vscale(pdir[q], BOND_HH));
        if (vnorm(ppos[q]) > rs + 1.0) ok = 0;
        else ok = clash_free(&c, ppos[q], skip, ns);
        V3 dd = vsub(ppos[q], pn);         /* pendant vs the new atom: 1-3 pair */
        if (ok && vdot(dd, dd) < 1.0) ok = 0;
      }
      if (!ok) continue;
      int cr = crowding(&c, pn);
      nvalid++;
      if (cr < best_crowd) {
        best_crowd = cr; best_pn = pn; best_npd = npd;
        for (int q = 0; q < npd; q++) best_pp[q] = ppos[q];
      }
      if (nvalid >= 6) break;
    }
    if (nvalid > 0) {
      for (int q = 0; q < best_npd; q++) {
        int ph = push_heavy(&c, best_pp[q], EL_O, -1, prev, k - 1);
        c.h[prev].pend[c.h[prev].npend++] = ph;
      }
      int hn = push_heavy(&c, best_pn, el, motif, -1, k);
      tot_at[k] = total;
      c.bb[c.nbb++] = hn;
      total += motif_atoms(motif);
      fails_in_row = 0;
      if (c.nbb > seg_best) { seg_best = c.nbb; stall = 0; }
      if (is_last) done = 1;
      continue;
    }
    /* blocked: pop a few placements of this segment; a frontier that stays
       blocked ends the segment (its last atom becomes a CH3 cap) */
    fails_in_row++;
    backtracks++;
    stall++;
    if (stall > 48 || seglen <= 2) {
      int last = c.bb[c.nbb - 1];
      total = tot_at[c.nbb - 1];
      c.h[last].motif = M_CH3_END;
      c.h[last].el = EL_C;
      total += motif_atoms(M_CH3_END);
      need_segment = 1;
      if (total >= natoms) done = 1;
      continue;
    }
    int npop = 1 + (int)(fails_in_row / 3);
    for (int t = 0; t < npop && c.nbb - seg_begin > 2; t++) {
      int before = c.bb[c.nbb - 2];
      pop_heavy(&c);                          /* the backbone atom (newest) */
      for (int q = c.h[before].npend - 1; q >= 0; q--) pop_heavy(&c);
      c.h[before].npend = 0;
      c.nbb--;
      total = tot_at[c.nbb];
    }
    if (backtracks > 400L * natoms + 4000) {
      free(c.h); free(c.bb); free(c.head); free(tot_at); free(seg_start);
      return -2;
    }
  }

  /* ---------------------------------------------------- assemble atoms */
  int nheavy = c.n;
  /* hydrogen budget per heavy atom; trim the excess from C atoms at the end */
  int *nhh = (int *)calloc((size_t)nheavy, sizeof(int));
  int htotal = 0;
  for (int i = 0; i < nheavy; i++) {
    if (c.h[i].parent < 0) nhh[i] = motif_nh(c.h[i].motif);
    else if (c.h[c.h[i].parent].motif == M_C_HYDROXYL) nhh[i] = 1;   /* O-H */
    htotal += nhh[i];
  }
  int excess = nheavy + htotal - natoms;
  for (int k = c.nbb - 1; k >= 0 && excess > 0; k--) {
    int i = c.bb[k];
    if (c.h[i].el != EL_C) continue;
    while (nhh[i] > 0 && excess > 0) { nhh[i]--; htotal--; excess--; }
  }
  if (nheavy + htotal != natoms) {   /* cannot happen for natoms >= 8 */
    free(c.h); free(c.bb); free(c.head); free(tot_at); free(nhh); free(seg_start);
    return -2;
  }

  /* output order: backbone atom, its pendants (each followed by its H), its H */
  int *out_of = (int *)malloc(sizeof(int) * (size_t)nheavy);  /* heavy -> output index */
  int *hpar = (int *)malloc(sizeof(int) * (size_t)natoms);    /* output index -> heavy parent (H) or -1 */
  int *hofs = (int *)malloc(sizeof(int) * (size_t)nheavy);    /* first H output index of heavy */
  int *heavy_at = (int *)malloc(sizeof(int) * (size_t)natoms); /* output index -> heavy, -1 for H */
  int nout = 0;
  for (int k = 0; k < c.nbb; k++) {
    int i = c.bb[k];
    out_of[i] = nout; hpar[nout] = -1; heavy_at[nout] = i; nout++;
    for (int q = 0; q < c.h[i].npend; q++) {
      int p = c.h[i].pend[q];
      out_of[p] = nout; hpar[nout] = -1; heavy_at[nout] = p; nout++;
      hofs[p] = nout;
      for (int t = 0; t < nhh[p]; t++) { hpar[nout] = p; heavy_at[nout] = -1; nout++; }
    }
    hofs[i] = nout;
    for (int t = 0; t < nhh[i]; t++) { hpar[nout] = i; heavy_at[nout] = -1; nout++; }
  }

  /* positions, elements and bonded heavy neighbours */
  int *el = (int *)malloc(sizeof(int) * (size_t)natoms);
  for (int i = 0; i < nheavy; i++) {
    int o = out_of[i];
    x[o] = c.h[i].pos.v[0]; y[o] = c.h[i].pos.v[1]; z[o] = c.h[i].pos.v[2];
    el[o] = c.h[i].el;
  }
  /* heavy adjacency (output indices), ordered prev, next, pendants */
  int *hadj = (int *)malloc(sizeof(int) * 4 * (size_t)nheavy);
  int *nhadj = (int *)calloc((size_t)nheavy, sizeof(int));
  for (int k = 0; k < c.nbb; k++) {
    int i = c.bb[k];
    if (k > 0 && !seg_start[k]) hadj[4 * i + nhadj[i]++] = c.bb[k - 1];
    if (k + 1 < c.nbb && !seg_start[k + 1]) hadj[4 * i + nhadj[i]++] = c.bb[k + 1];
    for (int q = 0; q < c.h[i].npend; q++) {
      int p = c.h[i].pend[q];
      hadj[4 * i + nhadj[i]++] = p;
      hadj[4 * p + nhadj[p]++] = i;
    }
  }
  for (int i = 0; i < nheavy; i++) {
    V3 e[4];
    for (int q = 0; q < nhadj[i]; q++) e[q] = vunit(vsub(c.h[hadj[4 * i + q]].pos, c.h[i].pos));
    /* directions follow the designed hydrogen count, so a trimmed CH3 keeps
       its tetrahedral geometry */
    int ndesign = c.h[i].parent < 0 ? motif_nh(c.h[i].motif) : nhh[i];
    V3 hd[4];
    int nd = nhh[i] > 0 ? hydrogen_dirs(&grow, e, nhadj[i], ndesign, hd) : 0;
    for (int t = 0; t < nhh[i] && t < nd; t++) {
      int o = hofs[i] + t;
      V3 p = vadd(c.h[i].pos, vscale(hd[t], BOND_XH));
      x[o] = p.v[0]; y[o] = p.v[1]; z[o] = p.v[2];
      el[o] = EL_H;
    }
  }

  /* connection table */
  int nc = 0;
  for (int o = 0; o < natoms; o++) {
    nne[o] = 0;
    if (hpar[o] >= 0) {                         /* hydrogen: its parent */
      if (nc + 1 > neighl1_cap) goto too_small;
      neighl1[nc++] = out_of[hpar[o]];
      nne[o] = 1;
      continue;
    }
    int i = heavy_at[o];
    for (int q = 0; q < nhadj[i]; q++) {
      if (nc + 1 > neighl1_cap) goto too_small;
      neighl1[nc++] = out_of[hadj[4 * i + q]];
      nne[o]++;
    }
    for (int t = 0; t < nhh[i]; t++) {
      if (nc + 1 > neighl1_cap) goto too_small;
      neighl1[nc++] = hofs[i] + t;
      nne[o]++;
    }
  }
  if (nconn_out) *nconn_out = nc;

  /* per-atom parameters, hydrogen-bond typing */
  double qsum = 0.0;
  for (int o = 0; o < natoms; o++) {
    int e2 = el[o];
    r[o] = el_radius[e2];
    charge[o] = 0.35 * rng_normal(&par);
    qsum += charge[o];
    int heavy = (e2 != EL_H);
    if (igamma) igamma[o] = heavy ? 0.117 : 0.0;
    if (sgamma) sgamma[o] = 0.0;
    if (ialpha) ialpha[o] = heavy ? rng_uniform(&par, -0.6, -0.2) : 0.0;
    if (salpha) salpha[o] = 0.0;
    if (ishydrogen) ishydrogen[o] = !heavy;
    int hb = HB_INACTIVE;
    if (!heavy) {
      int pe = el[out_of[hpar[o]]];
      if (pe == EL_N || pe == EL_O) hb = HB_POLARH;
    } else {
      const Heavy *a = &c.h[heavy_at[o]];
      if (a->parent >= 0) {                         /* pendant oxygen */
        int pm = c.h[a->parent].motif;
        if (pm == M_C_CARBONYL) hb = rng_u01(&par) < 0.3 ? HB_TRIGONAL_S : HB_TRIGONAL;
        else hb = HB_TETRAHEDRAL;                   /* hydroxyl O (2 bonds), sulfone O (1 bond) */
      } else if (a->motif == M_NH) {
        hb = rng_u01(&par) < 0.3 ? HB_TRIGONAL_OOP : HB_TETRAHEDRAL;
      } else if (a->motif == M_N_SP2) {
        hb = HB_TRIGONAL;
      } else if (a->motif == M_O_ETHER) {
        hb = HB_TETRAHEDRAL;
      }
    }
    if (hbtype) hbtype[o] = hb;
    if (hbcorr) hbcorr[o] = hb != HB_INACTIVE ? rng_uniform(&par, -1.0, -0.5) : 0.0;
  }
  for (int o = 0; o < natoms; o++) charge[o] -= qsum / natoms;

  if (info) {
    info[AGBSYNTH_INFO_SPHERE_R] = rs;
    info[AGBSYNTH_INFO_MIN_NB] = c.min_nb;
    info[AGBSYNTH_INFO_BACKTRACKS] = (double)backtracks;
    info[AGBSYNTH_INFO_NHEAVY] = nheavy;
    info[AGBSYNTH_INFO_NSEG] = (double)nsegments;
  }
  free(c.h); free(c.bb); free(c.head); free(tot_at); free(nhh); free(seg_start);
  free(out_of); free(hpar); free(hofs); free(heavy_at); free(el); free(hadj); free(nhadj);
  return 0;

too_small:
  free(c.h); free(c.bb); free(c.head); free(tot_at); free(nhh); free(seg_start);
  free(out_of); free(hpar); free(hofs); free(heavy_at); free(el); free(hadj); free(nhadj);
  return -3;
}

void agbsynth_perturb(uint64_t seed, int natoms, double sigma,
                      double *x, double *y, double *z) {
  Rng g;
  rng_seed(&g, seed);
  for (int i = 0; i < natoms; i++) {
    x[i] += sigma * rng_normal(&g);
    y[i] += sigma * rng_normal(&g);
    z[i] += sigma * rng_normal(&g);
  }
}

void agbsynth_sizes(uint64_t seed, int ninst, int nmin, int nmax, int *sizes) {
  Rng g;
  rng_seed(&g, seed);
  double a = log((double)nmin), b = log((double)nmax + 1.0);
  for (int k = 0; k < ninst; k++) {
    int n = (int)floor(exp(rng_uniform(&g, a, b)));
    if (n < nmin) n = nmin;
    if (n > nmax) n = nmax;
    sizes[k] = n;
  }
}

/* FP64 CPU oracle for AGBNP3 energy + gradients.
 *
 * TEST INFRASTRUCTURE ONLY (see agbnp3_oracle.h).  Not part of the product.
 *
 * Restates the reference algorithm of egallicc/AGBNP3 in double precision.
 * Reference citations are file:line into the reference tree.  Two places keep
 * the reference's FP32 arithmetic on purpose, because they define discrete
 * structure rather than values:
 *   - the neighbour-list distance test and row ordering (agbnp3.c:4028-4066),
 *   - the Q4 spline tables, which are model data (agbnp3_utils.c:380-422).
 * Build with -ffp-contract=off so those FP32 steps round like the reference.
 */
#include "agbnp3_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants of the model (agbnp3_private.h:57-137, agbnp3.c:1230-1236) */
#define RPI        3.14159265359      /* the reference's pi */
#define KFC        2.5
#define PFC        2.97354
#define RADINC     0.5
#define NBOFFSET   1.6f
#define MAXLEVEL   6                  /* orders 2..5 are built */
#define VOLMINA    0.01
#define VOLMINB    0.1
#define HB_LENGTH  2.5
#define HB_RADIUS  1.4
#define HB_SWA     0.65
#define HB_SWB     1.00
#define HB_SWA0    0.60
#define TOKCAL     332.0
#define RW_NP      1.4
#define TBL_N      512
#define TBL_AMAX   20.0f

enum { HB_INACTIVE = 0, HB_POLARH = 1, HB_TRIGONAL = 10, HB_TRIGONAL1 = 11,
       HB_TRIGONAL2 = 12, HB_TRIGONAL_OOP = 13, HB_TETRAHEDRAL = 20,
       HB_TETRAHEDRAL1 = 21, HB_TETRAHEDRAL2 = 22, HB_TETRAHEDRAL3 = 23,
       HB_TRIGONAL_S = 30 };

typedef struct {
  int nparents;
  int parent[4];
  int type;
  double pos[3];
  double dpos[4][3][3];   /* d pos_i / d x_{parent p, j} */
  double khb;
  double sp;
} WSite;

typedef struct {          /* one kept overlap of the Gaussian tree */
  int order;
  int atoms[MAXLEVEL];
  double a, p, c[3];      /* overlap Gaussian */
  double v, vp, fp, fpp;  /* raw / switched volume, switch derivatives */
} Ovl;

struct OracleSys {
  int natoms, nheavy, nhyd;
  int *int2ext, *ext2int;
  float *xf, *yf, *zf;           /* FP32 coordinates, internal order */
  float *R;                      /* r + 0.5, FP32 (agbnp3.c:210) */
  double *q, *igamma, *sgamma, *ialpha, *salpha, *hbcorr;
  int *hbtype;
  double eps_in, eps_out;
  int *cnne, *coff, *cnl;        /* connection table, internal indices */
  int *rtype, ntypes;
  float *radii;
  float *ty, *ty2;               /* [ntypes*ntypes][512] */
  float tdx, tdxinv;
  /* per-call state */
  int *nne, *nbr, nbrcap;        /* near list */
  Ovl *ov; long nov, ovcap, neval;
  double *vols, *volumep_raw, *volumep, *sa_raw, *sa_f, *sp, *spe, *psvol,
         *gammap, *beta, *br, *brw, *swf, *A, *q2ab, *abrw, *deru, *derv,
         *derus, *dervs, *derh;
  double (*g_gb)[3], (*g_vw)[3], (*g_cav)[3], (*g_hb)[3];
  WSite *ws; int nws, wscap;
  double e[ORC_NENERGY];
  int spline_fixed;
};

/* ------------------------------------------------------------------ utils */
static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); exit(1); }
  return p;
}

/* -------------------------------------------- FP32 spline tables (model) */
/* agbnp3_i4 in the reference's mixed float/double arithmetic
   (agbnp3.c:3886-3944); used only to build the tables. */
static float i4_f32(float rij, float Ri, float Rj, float *dr) {
  float u1, u2, u3, u4, u5, u6, a, ad, u4sq, u5sq, q;
  float rij2 = rij * rij;
  static const float twopi = 2.0 * RPI;
  static const float twothirds = 2.0 / 3.0;
  if (rij > (Ri + Rj)) {
    u1 = rij + Rj;
    u2 = rij - Rj;
    u3 = u1 * u2;
    u4 = 0.5 * log(u1 / u2);
    q = twopi * (Rj / u3 - u4 / rij);
    *dr = twopi * ((Rj / (rij * u3)) * (1. - 2. * rij2 / u3) + u4 / rij2);
  } else {
    u1 = Rj - Ri;
    if (rij2 > u1 * u1) {
      u1 = rij + Rj;
      u2 = rij - Rj;
      u3 = u1 * u2;
      u4 = 1. / u1;
      u4sq = u4 * u4;
      u5 = 1. / Ri;
      u5sq = u5 * u5;
      u6 = 0.5 * log(u1 / Ri);
      q = twopi * (-(u4 - u5) + (0.25 * u3 * (u4sq - u5sq) - u6) / rij);
      *dr = twopi * (0.5 * (1. - 0.5 * u3 / rij2) * (u4sq - u5sq) + u6 / rij2);
    } else {
      if (Ri > Rj) {
        q = 0.0;
        *dr = 0.0;
      } else {
        u1 = rij + Rj;
        u2 = Rj - rij;
        u3 = -u1 * u2;
        if (rij < .001 * Rj) {
          a = rij / Rj;
          ad = a * a - 1;
          u6 = (1. + twothirds * a * a) / Rj;
          q = twopi * (2. / Ri + Rj / u3 - u6);
          *dr = -(2. * twopi * a / (Rj * Rj)) * (1. / ad + twothirds);
        } else {
          u6 = 0.5 * log(u1 / u2);
          q = twopi * (2. / Ri + Rj / u3 - u6 / (rij));
          *dr = twopi * (-(Rj / u3) * (2. * rij / u3 - 1. / (rij)) + u6 / (rij2));
        }
      }
    }
  }
  return q;
}

/* agbnp3_i4ov (agbnp3.c:3964-3975) with agbnp3_ogauss_2body_smpl
   (agbnp3.c:3946-3956); both Gaussian exponents come from Ri (quirk Q4). */
float oracle_i4ov_f32(float rij, float Ri, float Rj, float *dr) {
  float ai = KFC / (Ri * Ri);
  float pii = PFC;
  float aj = KFC / (Ri * Ri);
  float pjj = PFC;
  float d2 = rij * rij;
  float deltai = 1. / (ai + aj);
  float p = pii * pjj;
  float kappa = exp(-ai * aj * d2 * deltai);
  float gvol = p * kappa * pow(RPI * deltai, 1.5);
  float volj = 4. * RPI * Rj * Rj * Rj / 3.;
  float newRj = pow((volj + 2. * gvol) / volj, 1. / 3.) * Rj;
  return i4_f32(rij, Ri, newRj, dr);
}

/* agbnp3_cspline_setup (agbnp3_utils.c:193-227), incl. its end condition
   (divides by 2*dx: quirk Q5) */
static void cspline_setup_f32(float dx, int n, const float *y, float yp1, float ypn, float *y2) {
  float *u = (float *)xcalloc((size_t)n, sizeof(float));
  float p, qn, sig, un;
  if (yp1 > 0.99e30)
    y2[0] = u[0] = 0.0;
  else {
    y2[0] = -0.5;
    u[0] = (3.0 / (dx)) * ((y[1] - y[0]) / (dx) - yp1);
  }
  for (int i = 1; i < n - 1; i++) {
    sig = 0.5;
    p = sig * y2[i - 1] + 2.0;
    y2[i] = (sig - 1.0) / p;
    u[i] = (y[i + 1] - y[i]) / (dx) - (y[i] - y[i - 1]) / (dx);
    u[i] = (6.0 * u[i] / (2. * dx) - sig * u[i - 1]) / p;
  }
  if (ypn > 0.99e30)
    qn = un = 0.0;
  else {
    qn = 0.5;
    un = (3.0 / (dx)) * (ypn - (y[n - 1] - y[n - 2]) / (2. * dx));
  }
  y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
  for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
  free(u);
}

/* agbnp3_create_ctablef4 (agbnp3_utils.c:380-422): 511 nodes of i4ov on
   [0, 20) at accumulated spacing, y[511] = 0, clamped slopes. */
static void build_table(float b, float *y, float *y2, float *dx_out, float *dxinv_out) {
  const int n = TBL_N;
  float amax = TBL_AMAX;
  float da = amax / (n - 1);
  float a = 0.0, dr, yp1 = 0.0, Rj = 1.0;
  for (int i = 0; i < n - 1; i++) {
    float q1 = oracle_i4ov_f32(a, b, Rj, &dr);
    if (i == 0) yp1 = dr;
    y[i] = q1;
    a += da;
  }
  y[n - 1] = 0.0f;
  cspline_setup_f32(da, n, y, yp1, 0.0f, y2);
  *dx_out = da;
  *dxinv_out = 1. / da;
}

/* --------------------------------------------------------- geometry util */
static void mat_unit_der(const double u[3], double invr, double der[3][3]) {
  /* d u / d r for u = r/|r| (agbnp3.c:3251-3264) */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) der[i][j] = -invr * u[i] * u[j] + (i == j ? invr : 0.0);
}
static void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static void matmul3(double a[3][3], double b[3][3], double c[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += a[i][k] * b[k][j];
      c[i][j] = s;
    }
}
static double unitvec(const double *from, const double *to, double u[3]) {
  double d[3] = {to[0] - from[0], to[1] - from[1], to[2] - from[2]};
  double inv = 1.0 / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int k = 0; k < 3; k++) u[k] = d[k] * inv;
  return inv;
}

/* ------------------------------------------------ Gaussian pair overlap */
/* quintic switch of the overlap volume on [0.01, 0.1]
   (agbnp3_cpu_simd_kernels.c:2448-2471, agbnp3.c:1120-1145) */
static void vol_switch(double v, double *s, double *sp, double *spp) {
  if (v > VOLMINB) { *s = 1.0; *sp = 0.0; *spp = 0.0; return; }
  if (v < VOLMINA) { *s = 0.0; *sp = 0.0; *spp = 0.0; return; }
  double d = 1.0 / (VOLMINB - VOLMINA), u = (v - VOLMINA) * d, u2 = u * u;
  *s = u2 * u * (10.0 - 15.0 * u + 6.0 * u2);
  *sp = d * 30.0 * u2 * (1.0 - 2.0 * u + u2);
  *spp = d * d * 60.0 * u * (1.0 - 3.0 * u + 2.0 * u2);
}

/* agbnp3_ogauss_soa (agbnp3_cpu_simd_kernels.c:2418-2477) */
double oracle_ogauss(double a1, double p1, const double *c1, double a2, double p2,
                     const double *c2, double *out) {
  double dx = c2[0] - c1[0], dy = c2[1] - c1[1], dz = c2[2] - c1[2];
  double d2 = dx * dx + dy * dy + dz * dz;
  double a12 = a1 + a2, deltai = 1.0 / a12;
  double p12 = p1 * p2 * exp(-a1 * a2 * d2 * deltai);
  double v = p12 * pow(RPI * deltai, 1.5);
  double s, sp, spp;
  vol_switch(v, &s, &sp, &spp);
  out[0] = a12;
  out[1] = p12;
  for (int k = 0; k < 3; k++) out[2 + k] = deltai * (a1 * c1[k] + a2 * c2[k]);
  out[5] = s * v;
  out[6] = s + v * sp;
  out[7] = 2.0 * sp + v * spp;
  return v;
}

/* polynomial switch for E_HB (agbnp3.c:1120-1145) */
static double pol_switch(double x, double xa, double xb, double *fp) {
  if (x > xb) { *fp = 0.0; return 1.0; }
  if (x < xa) { *fp = 0.0; return 0.0; }
  double d = 1.0 / (xb - xa), u = (x - xa) * d, u2 = u * u;
  *fp = d * 30.0 * u2 * (1.0 - 2.0 * u + u2);
  return u2 * u * (10.0 - 15.0 * u + 6.0 * u2);
}

/* ------------------------------------------------------------- i4 (double) */
/* agbnp3_i4 (agbnp3.c:3886-3944) in double, for closed-form tests */
double oracle_i4(double rij, double Ri, double Rj, double *dr) {
  const double twopi = 2.0 * RPI, tt = 2.0 / 3.0;
  double rij2 = rij * rij, q;
  if (rij > Ri + Rj) {
    double u1 = rij + Rj, u2 = rij - Rj, u3 = u1 * u2, u4 = 0.5 * log(u1 / u2);
    q = twopi * (Rj / u3 - u4 / rij);
    *dr = twopi * ((Rj / (rij * u3)) * (1. - 2. * rij2 / u3) + u4 / rij2);
  } else if (rij2 > (Rj - Ri) * (Rj - Ri)) {
    double u1 = rij + Rj, u2 = rij - Rj, u3 = u1 * u2, u4 = 1. / u1, u5 = 1. / Ri;
    double u6 = 0.5 * log(u1 / Ri);
    q = twopi * (-(u4 - u5) + (0.25 * u3 * (u4 * u4 - u5 * u5) - u6) / rij);
    *dr = twopi * (0.5 * (1. - 0.5 * u3 / rij2) * (u4 * u4 - u5 * u5) + u6 / rij2);
  } else if (Ri > Rj) {
    q = 0.0; *dr = 0.0;
  } else {
    double u1 = rij + Rj, u2 = Rj - rij, u3 = -u1 * u2;
    if (rij < .001 * Rj) {
      double a = rij / Rj, ad = a * a - 1, u6 = (1. + tt * a * a) / Rj;
      q = twopi * (2. / Ri + Rj / u3 - u6);
      *dr = -(2. * twopi * a / (Rj * Rj)) * (1. / ad + tt);
    } else {
      double u6 = 0.5 * log(u1 / u2);
      q = twopi * (2. / Ri + Rj / u3 - u6 / rij);
      *dr = twopi * (-(Rj / u3) * (2. * rij / u3 - 1. / rij) + u6 / rij2);
    }
  }
  return q;
}

/* --------------------------------------------------------------- Q4 lookup */
/* agbnp3_i4p_soa + agbnp3_interpolate_ctablef42d_soa + cspline interpolation
   (agbnp3_utils.c:297-377, 266-295) in double from the FP32 table */
double oracle_q4(const OracleSys *s, int t, double d, double Rj, double *dq4) {
  double a = d / Rj;
  double dx = s->tdx, dxinv = s->tdxinv;
  double xh = a * dxinv;
  int k = (int)xh;
  if (k > TBL_N - 2) { *dq4 = 0.0; return 0.0; }
  const float *y = s->ty + (size_t)t * TBL_N, *y2 = s->ty2 + (size_t)t * TBL_N;
  double A = (k + 1.0) - xh, B = xh - k, A2 = A * A, B2 = B * B;
  double dp1 = dx / 6.0, dp2 = dx * dp1;
  double f = A * y[k] + B * y[k + 1] + ((A2 * A - A) * y2[k] + (B2 * B - B) * y2[k + 1]) * dp2;
  /* the reference subtracts the y2[k+1] term (agbnp3_utils.c:292-293);
     the derivative of the interpolant adds it */
  double sgn = s->spline_fixed ? -1.0 : 1.0;
  double fp = (y[k + 1] - (double)y[k]) * dxinv -
              ((3.0 * A2 - 1.0) * y2[k] + sgn * (3.0 * B2 - 1.0) * y2[k + 1]) * dp1;
  *dq4 = fp / (Rj * Rj);
  return f / Rj;
}

/* ------------------------------------------------------------ oracle_new */
OracleSys *oracle_new(int natoms, const double *x, const double *y, const double *z,
                      const double *r, const double *charge, double eps_in, double eps_out,
                      const double *igamma, const double *sgamma,
                      const double *ialpha, const double *salpha,
                      const int *hbtype, const double *hbcorr,
                      int nhydrogen, const int *ihydrogen,
                      const int *conn_nne, const int *conn_neighl1) {
  if (natoms <= 0 || !conn_nne || !conn_neighl1) {
    fprintf(stderr, "oracle_new(): bad arguments\n");
    return NULL;
  }
  OracleSys *s = (OracleSys *)xcalloc(1, sizeof(OracleSys));
  int n = natoms;
  s->natoms = n;
  s->nhyd = nhydrogen;
  s->nheavy = n - nhydrogen;
  s->int2ext = (int *)xcalloc(n, sizeof(int));
  s->ext2int = (int *)xcalloc(n, sizeof(int));
  /* heavy atoms first, hydrogens last, both stable (agbnp3.c:739-779) */
  int *ish = (int *)xcalloc(n, sizeof(int));
  for (int i = 0; i < nhydrogen; i++) ish[ihydrogen[i]] = 1;
  int nc = 0, nh = 0;
  for (int i = 0; i < n; i++) {
    int slot = ish[i] ? s->nheavy + nh++ : nc++;
    s->int2ext[slot] = i;
    s->ext2int[i] = slot;
  }
  free(ish);
  s->xf = (float *)xcalloc(n, sizeof(float));
  s->yf = (float *)xcalloc(n, sizeof(float));
  s->zf = (float *)xcalloc(n, sizeof(float));
  s->R = (float *)xcalloc(n, sizeof(float));
  s->q = (double *)xcalloc(n, sizeof(double));
  s->igamma = (double *)xcalloc(n, sizeof(double));
  s->sgamma = (double *)xcalloc(n, sizeof(double));
  s->ialpha = (double *)xcalloc(n, sizeof(double));
  s->salpha = (double *)xcalloc(n, sizeof(double));
  s->hbcorr = (double *)xcalloc(n, sizeof(double));
  s->hbtype = (int *)xcalloc(n, sizeof(int));
  for (int i = 0; i < n; i++) {
    int e = s->int2ext[i];
    s->xf[i] = (float)x[e]; s->yf[i] = (float)y[e]; s->zf[i] = (float)z[e];
    s->R[i] = (float)(r[e] + RADINC);                 /* agbnp3.c:210 */
    s->q[i] = (float)charge[e];                       /* float_a storage */
    s->igamma[i] = (float)igamma[e];
    s->sgamma[i] = (float)sgamma[e];
    s->ialpha[i] = (float)ialpha[e];
    s->salpha[i] = (float)salpha[e];
    s->hbcorr[i] = (float)hbcorr[e];
    s->hbtype[i] = hbtype[e];
  }
  s->eps_in = (float)eps_in;
  s->eps_out = (float)eps_out;
  /* connection table remapped to internal indices (agbnp3.c:301-317) */
  s->cnne = (int *)xcalloc(n, sizeof(int));
  s->coff = (int *)xcalloc(n + 1, sizeof(int));
  int *eoff = (int *)xcalloc(n + 1, sizeof(int));
  for (int i = 0; i < n; i++) eoff[i + 1] = eoff[i] + conn_nne[i];
  for (int i = 0; i < n; i++) {
    s->cnne[i] = conn_nne[s->int2ext[i]];
    s->coff[i + 1] = s->coff[i] + s->cnne[i];
  }
  s->cnl = (int *)xcalloc(s->coff[n], sizeof(int));
  for (int i = 0; i < n; i++) {
    int e = s->int2ext[i];
    for (int k = 0; k < s->cnne[i]; k++) s->cnl[s->coff[i] + k] = s->ext2int[conn_neighl1[eoff[e] + k]];
  }
  free(eoff);
  /* radius types by exact equality (agbnp3_utils.c:808-838) */
  s->rtype = (int *)xcalloc(n, sizeof(int));
  s->radii = (float *)xcalloc(n, sizeof(float));
  for (int i = 0; i < n; i++) {
    int found = -1;
    for (int t = 0; t < s->ntypes; t++)
      if (fabsf(s->R[i] - s->radii[t]) < FLT_MIN) { found = t; break; }
    if (found < 0) { found = s->ntypes; s->radii[s->ntypes++] = s->R[i]; }
    s->rtype[i] = found;
  }
  /* one table per ordered type pair (agbnp3_utils.c:840-883) */
  int nt = s->ntypes;
  s->ty = (float *)xcalloc((size_t)nt * nt * TBL_N, sizeof(float));
  s->ty2 = (float *)xcalloc((size_t)nt * nt * TBL_N, sizeof(float));
  for (int i = 0; i < nt; i++)
    for (int j = 0; j < nt; j++) {
      float c = RADINC;
      float b = (s->radii[i] - c) / s->radii[j];
      size_t t = (size_t)(i * nt + j);
      build_table(b, s->ty + t * TBL_N, s->ty2 + t * TBL_N, &s->tdx, &s->tdxinv);
    }
  /* per-call buffers */
  s->vols = (double *)xcalloc(n, sizeof(double));
  for (int i = 0; i < n; i++) s->vols[i] = 4.0 * RPI / 3.0 * pow((double)s->R[i], 3);
#define ALLOC_D(f) s->f = (double *)xcalloc(n, sizeof(double))
  ALLOC_D(volumep_raw); ALLOC_D(volumep); ALLOC_D(sa_raw); ALLOC_D(sa_f); ALLOC_D(sp);
  ALLOC_D(spe); ALLOC_D(psvol); ALLOC_D(gammap); ALLOC_D(beta); ALLOC_D(br); ALLOC_D(brw);
  ALLOC_D(swf); ALLOC_D(A); ALLOC_D(q2ab); ALLOC_D(abrw); ALLOC_D(deru); ALLOC_D(derv);
  ALLOC_D(derus); ALLOC_D(dervs); ALLOC_D(derh);
#undef ALLOC_D
  s->g_gb = xcalloc(n, sizeof(double[3]));
  s->g_vw = xcalloc(n, sizeof(double[3]));
  s->g_cav = xcalloc(n, sizeof(double[3]));
  s->g_hb = xcalloc(n, sizeof(double[3]));
  return s;
}

void oracle_free(OracleSys *s) {
  if (!s) return;
  free(s->int2ext); free(s->ext2int); free(s->xf); free(s->yf); free(s->zf); free(s->R);
  free(s->q); free(s->igamma); free(s->sgamma); free(s->ialpha); free(s->salpha);
  free(s->hbcorr); free(s->hbtype); free(s->cnne); free(s->coff); free(s->cnl);
  free(s->rtype); free(s->radii); free(s->ty); free(s->ty2); free(s->nne); free(s->nbr);
  free(s->ov); free(s->vols); free(s->volumep_raw); free(s->volumep); free(s->sa_raw);
  free(s->sa_f); free(s->sp); free(s->spe); free(s->psvol); free(s->gammap); free(s->beta);
  free(s->br); free(s->brw); free(s->swf); free(s->A); free(s->q2ab); free(s->abrw);
  free(s->deru); free(s->derv); free(s->derus); free(s->dervs); free(s->derh);
  free(s->g_gb); free(s->g_vw); free(s->g_cav); free(s->g_hb); free(s->ws);
  free(s);
}

/* ------------------------------------------------------------ near list */
/* agbnp3_neighbor_lists (agbnp3.c:3978-4115): heavy j > i with
   d2 < ((R_i + R_j) * 1.6f)^2 in FP32, each row sorted by (d2, j)
   (agbnp3_fcompare / agbnp3_nblist_reorder, agbnp3.c:3766-3859). */
typedef struct { float d2; int j; } NbEntry;
static int nb_cmp(const void *a, const void *b) {
  const NbEntry *p = (const NbEntry *)a, *q = (const NbEntry *)b;
  if (p->d2 < q->d2) return -1;
  if (p->d2 > q->d2) return 1;
  return (p->j > q->j) - (p->j < q->j);
}
static void near_list(OracleSys *s) {
  int nh = s->nheavy;
  free(s->nne); free(s->nbr);
  s->nne = (int *)xcalloc(nh, sizeof(int));
  /* first pass: row lengths */
  int cap = 1;
  for (int i = 0; i < nh; i++) {
    int c = 0;
    for (int j = i + 1; j < nh; j++) {
      float dx = s->xf[j] - s->xf[i], dy = s->yf[j] - s->yf[i], dz = s->zf[j] - s->zf[i];
      float d2 = dx * dx + dy * dy + dz * dz;
      float u = (s->R[i] + s->R[j]) * NBOFFSET;
      if (d2 < u * u) c++;
    }
    s->nne[i] = c;
    if (c > cap) cap = c;
  }
  s->nbrcap = cap;
  s->nbr = (int *)xcalloc((size_t)nh * cap, sizeof(int));
  NbEntry *row = (NbEntry *)xcalloc(cap, sizeof(NbEntry));
  for (int i = 0; i < nh; i++) {
    int c = 0;
    for (int j = i + 1; j < nh; j++) {
      float dx = s->xf[j] - s->xf[i], dy = s->yf[j] - s->yf[i], dz = s->zf[j] - s->zf[i];
      float d2 = dx * dx + dy * dy + dz * dz;
      float u = (s->R[i] + s->R[j]) * NBOFFSET;
      if (d2 < u * u) { row[c].d2 = d2; row[c].j = j; c++; }
    }
    qsort(row, c, sizeof(NbEntry), nb_cmp);
    for (int k = 0; k < c; k++) s->nbr[(size_t)i * cap + k] = row[k].j;
  }
  free(row);
}

/* ------------------------------------------------- overlap tree (pass 1) */
static void push_ovl(OracleSys *s, const Ovl *o) {
  if (s->nov == s->ovcap) {
    s->ovcap = s->ovcap * 2 + 1024;
    s->ov = (Ovl *)realloc(s->ov, sizeof(Ovl) * s->ovcap);
    if (!s->ov) { fprintf(stderr, "oracle: out of memory\n"); exit(1); }
  }
  s->ov[s->nov++] = *o;
}

/* Level-by-level inclusion-exclusion of agbnp3_self_volumes_rooti
   (agbnp3_cpu_simd_kernels.c:1064-1598): children of an overlap are formed
   with the last atom of each later sibling in its root; an overlap is kept if
   its switched volume exceeds FLT_MIN; orders 2..5. */
static void overlap_tree(OracleSys *s, const double *ga, const double (*gc)[3]) {
  int nh = s->nheavy;
  s->nov = 0;
  s->neval = 0;
  /* lists of one root level: records are kept overlaps, grouped in roots */
  Ovl *cur = NULL, *nxt = NULL;
  int *croot = NULL, *nroot = NULL;
  long nnxt = 0, capc = 0, capn = 0;
  int ncroot = 0, nnroot = 0;
  long rcap = 0, rcapn = 0;
  /* order 2 */
  for (int i = 0; i < nh; i++) {
    if (nnroot + 2 > rcapn) { rcapn = 2 * rcapn + 1024; nroot = realloc(nroot, sizeof(int) * rcapn); }
    nroot[nnroot] = (int)nnxt;
    for (int k = 0; k < s->nne[i]; k++) {
      int j = s->nbr[(size_t)i * s->nbrcap + k];
      double out[8];
      double v = oracle_ogauss(ga[i], PFC, gc[i], ga[j], PFC, gc[j], out);
      s->neval++;
      if (out[5] > FLT_MIN) {
        Ovl o;
        memset(&o, 0, sizeof(o));
        o.order = 2; o.atoms[0] = i; o.atoms[1] = j;
        o.a = out[0]; o.p = out[1]; o.c[0] = out[2]; o.c[1] = out[3]; o.c[2] = out[4];
        o.v = v; o.vp = out[5]; o.fp = out[6]; o.fpp = out[7];
        if (nnxt == capn) { capn = 2 * capn + 1024; nxt = realloc(nxt, sizeof(Ovl) * capn); }
        nxt[nnxt++] = o;
        push_ovl(s, &o);
      }
    }
    nnroot++;
  }
  nroot[nnroot] = (int)nnxt;
  /* orders 3..5 */
  for (int order = 3; order < MAXLEVEL && nnxt > 0; order++) {
    /* swap */
    Ovl *t = cur; cur = nxt; nxt = t;
    long tc = capc; capc = capn; capn = tc;
    int *tr = croot; croot = nroot; nroot = tr;
    long trc = rcap; rcap = rcapn; rcapn = trc;
    ncroot = nnroot;
    nnxt = 0; nnroot = 0;
    for (int ir = 0; ir < ncroot; ir++) {
      for (int a = croot[ir]; a < croot[ir + 1]; a++) {
        if (nnroot + 2 > rcapn) { rcapn = 2 * rcapn + 1024; nroot = realloc(nroot, sizeof(int) * rcapn); }
        nroot[nnroot] = (int)nnxt;
        const Ovl *oa = &cur[a];
        for (int b = a + 1; b < croot[ir + 1]; b++) {
          int j = cur[b].atoms[order - 2];
          double out[8];
          double v = oracle_ogauss(oa->a, oa->p, oa->c, ga[j], PFC, gc[j], out);
          s->neval++;
          if (out[5] > FLT_MIN) {
            Ovl o = *oa;
            o.order = order;
            o.atoms[order - 1] = j;
            o.a = out[0]; o.p = out[1]; o.c[0] = out[2]; o.c[1] = out[3]; o.c[2] = out[4];
            o.v = v; o.vp = out[5]; o.fp = out[6]; o.fpp = out[7];
            if (nnxt == capn) { capn = 2 * capn + 1024; nxt = realloc(nxt, sizeof(Ovl) * capn); }
            nxt[nnxt++] = o;
            push_ovl(s, &o);
          }
        }
        nnroot++;
      }
    }
    if (nnroot + 1 > rcapn) { rcapn = 2 * rcapn + 1024; nroot = realloc(nroot, sizeof(int) * rcapn); }
    nroot[nnroot] = (int)nnxt;
  }
  free(cur); free(nxt); free(croot); free(nroot);
}

/* -------------------------------------------------------- water sites */
/* site geometries of agbnp3.c:2014-3761; dpos[p][i][j] = d pos_i / d x_{p,j} */
static int ws_polarh(const double *X, int h, int d, WSite *w) {
  /* agbnp3_place_wat_hydrogen (agbnp3.c:2098-2150) */
  const double *xd = X + 3 * d, *xh = X + 3 * h;
  double u[3];
  double inv = unitvec(xd, xh, u);
  double wl = HB_LENGTH * inv;
  for (int k = 0; k < 3; k++) w->pos[k] = xd[k] + HB_LENGTH * u[k];
  memset(w->dpos, 0, sizeof(w->dpos));
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double m = wl * u[i] * u[j];
      w->dpos[0][i][j] = -m + (i == j ? wl : 0.0);
      w->dpos[1][i][j] = m + (i == j ? 1.0 - wl : 0.0);
    }
  w->nparents = 2; w->parent[0] = h; w->parent[1] = d;
  return 1;
}

static void ws_trigonal1_pair(const double *X, int a, int r, int r1, int r2, WSite *w) {
  /* agbnp3_place_wat_trigonal1 (agbnp3.c:2295-2366) */
  const int rr[2] = {r1, r2};
  for (int s = 0; s < 2; s++) {
    double u[3];
    double inv = unitvec(X + 3 * r, X + 3 * rr[s], u);
    double c = HB_LENGTH * inv;
    for (int k = 0; k < 3; k++) w[s].pos[k] = X[3 * a + k] - HB_LENGTH * u[k];
    memset(w[s].dpos, 0, sizeof(w[s].dpos));
    for (int i = 0; i < 3; i++) {
      w[s].dpos[0][i][i] = 1.0;
      for (int j = 0; j < 3; j++) {
        double m = c * ((i == j) - u[i] * u[j]);
        w[s].dpos[1][i][j] = m;
        w[s].dpos[2 + s][i][j] = -m;
      }
    }
    w[s].nparents = 4;
    w[s].parent[0] = a; w[s].parent[1] = r; w[s].parent[2] = r1; w[s].parent[3] = r2;
  }
}

static void ws_trigonal_s_out(const double *X, int a, int r, int r1, int r2, WSite *w) {
  /* agbnp3_place_wat_trigonal_s (agbnp3.c:2590-2697) */
  const double costh = 0.5, sinth = 0.866025;
  const double *xa = X + 3 * a, *xr = X + 3 * r, *xr1 = X + 3 * r1;
  double dx0[3], dx1[3], uin[3], c[3], uout[3];
  for (int k = 0; k < 3; k++) { dx0[k] = xa[k] - xr[k]; dx1[k] = xr1[k] - xr[k]; }
  double d0i = 1.0 / sqrt(dx0[0] * dx0[0] + dx0[1] * dx0[1] + dx0[2] * dx0[2]);
  for (int k = 0; k < 3; k++) uin[k] = dx0[k] * d0i;
  cross3(dx1, dx0, c);
  double d2i = 1.0 / sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  for (int k = 0; k < 3; k++) uout[k] = c[k] * d2i;
  for (int k = 0; k < 3; k++) {
    w[0].pos[k] = xa[k] + HB_LENGTH * (costh * uin[k] + sinth * uout[k]);
    w[1].pos[k] = xa[k] + HB_LENGTH * (costh * uin[k] - sinth * uout[k]);
  }
  double Din[3][3], Dout[3][3], t1[3][3], t2[3][3];
  mat_unit_der(uin, d0i, Din);
  mat_unit_der(uout, d2i, Dout);
  for (int i = 0; i < 3; i++) { cross3(dx1, Dout[i], t1[i]); cross3(dx0, Dout[i], t2[i]); }
  for (int s = 0; s < 2; s++) {
    double sg = s == 0 ? 1.0 : -1.0;
    memset(w[s].dpos, 0, sizeof(w[s].dpos));
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double win = HB_LENGTH * costh * Din[i][j];
        double w1 = -HB_LENGTH * sinth * t1[i][j], w2 = HB_LENGTH * sinth * t2[i][j];
        w[s].dpos[0][i][j] = win + sg * w1 + (i == j);
        w[s].dpos[1][i][j] = -win - sg * (w1 + w2);
        w[s].dpos[2][i][j] = sg * w2;
      }
    w[s].nparents = 4;
    w[s].parent[0] = a; w[s].parent[1] = r; w[s].parent[2] = r1; w[s].parent[3] = r2;
  }
}

static void ws_trigonal_oop(const double *X, int a, int r1, int r2, int r3, WSite *w) {
  /* agbnp3_place_wat_trigonal_oop (agbnp3.c:2815-2942): sites at 2.7 A
     across A along the normal of the R1 R2 R3 plane */
  const double d = HB_LENGTH + 0.2;
  const double *x1 = X + 3 * r1, *x2 = X + 3 * r2, *x3 = X + 3 * r3;
  double v1[3], v2[3], nu[3];
  for (int k = 0; k < 3; k++) { v1[k] = x2[k] - x1[k]; v2[k] = x3[k] - x1[k]; }
  cross3(v2, v1, nu);
  double nun = sqrt(nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2]);
  for (int k = 0; k < 3; k++) nu[k] /= nun;
  for (int k = 0; k < 3; k++) {
    w[0].pos[k] = X[3 * a + k] + d * nu[k];
    w[1].pos[k] = X[3 * a + k] - d * nu[k];
  }
  /* d(v2 x v1)/dv1 and /dv2 */
  double J1[3][3] = {{0, -v2[2], v2[1]}, {v2[2], 0, -v2[0]}, {-v2[1], v2[0], 0}};
  double J2[3][3] = {{0, v1[2], -v1[1]}, {-v1[2], 0, v1[0]}, {v1[1], -v1[0], 0}};
  double P[3][3], M1[3][3], M2[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) P[i][j] = (i == j) - nu[i] * nu[j];
  matmul3(P, J1, M1);
  matmul3(P, J2, M2);
  double dn = d / nun;
  for (int s = 0; s < 2; s++) {
    double sg = s == 0 ? 1.0 : -1.0;
    memset(w[s].dpos, 0, sizeof(w[s].dpos));
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        w[s].dpos[0][i][j] = (i == j);
        w[s].dpos[2][i][j] = sg * dn * M1[i][j];
        w[s].dpos[3][i][j] = sg * dn * M2[i][j];
        w[s].dpos[1][i][j] = -(w[s].dpos[2][i][j] + w[s].dpos[3][i][j]);
      }
    w[s].nparents = 4;
    w[s].parent[0] = a; w[s].parent[1] = r1; w[s].parent[2] = r2; w[s].parent[3] = r3;
  }
}

static void ws_bisector(const double *X, int a, const int *rr, int nr, WSite *w) {
  /* agbnp3_place_wat_trigonal2 (agbnp3.c:3024-3090) for nr = 2 and
     agbnp3_place_wat_tetrahedral3 (agbnp3.c:3476-3556) for nr = 3:
     site opposite the sum of the bond unit vectors */
  double d[3][3], inv[3], dw[3] = {0, 0, 0};
  for (int k = 0; k < nr; k++) {
    inv[k] = unitvec(X + 3 * a, X + 3 * rr[k], d[k]);
    for (int c = 0; c < 3; c++) dw[c] += d[k][c];
  }
  double iw = 1.0 / sqrt(dw[0] * dw[0] + dw[1] * dw[1] + dw[2] * dw[2]);
  for (int c = 0; c < 3; c++) dw[c] *= iw;
  for (int c = 0; c < 3; c++) w->pos[c] = X[3 * a + c] - HB_LENGTH * dw[c];
  double Dw[3][3];
  mat_unit_der(dw, iw, Dw);
  memset(w->dpos, 0, sizeof(w->dpos));
  for (int i = 0; i < 3; i++) w->dpos[0][i][i] = 1.0;
  for (int k = 0; k < nr; k++) {
    double Dk[3][3], M[3][3];
    mat_unit_der(d[k], inv[k], Dk);
    matmul3(Dw, Dk, M);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        w->dpos[1 + k][i][j] = -HB_LENGTH * M[i][j];
        w->dpos[0][i][j] += HB_LENGTH * M[i][j];
      }
  }
  w->nparents = nr + 1;
  w->parent[0] = a;
  for (int k = 0; k < nr; k++) w->parent[1 + k] = rr[k];
}

static void ws_tetrahedral2(const double *X, int a, int r1, int r2, WSite *w) {
  /* agbnp3_place_wat_tetrahedral2 (agbnp3.c:3266-3390) */
  const double costh = -0.577350269, sinth = 0.816496581;
  const double *xa = X + 3 * a;
  double r1v[3], r2v[3], d1[3], d2[3], pin[3], out[3];
  for (int k = 0; k < 3; k++) { r1v[k] = X[3 * r1 + k] - xa[k]; r2v[k] = X[3 * r2 + k] - xa[k]; }
  double i1 = 1.0 / sqrt(r1v[0] * r1v[0] + r1v[1] * r1v[1] + r1v[2] * r1v[2]);
  double i2 = 1.0 / sqrt(r2v[0] * r2v[0] + r2v[1] * r2v[1] + r2v[2] * r2v[2]);
  for (int k = 0; k < 3; k++) { d1[k] = r1v[k] * i1; d2[k] = r2v[k] * i2; pin[k] = d1[k] + d2[k]; }
  double iin = 1.0 / sqrt(pin[0] * pin[0] + pin[1] * pin[1] + pin[2] * pin[2]);
  for (int k = 0; k < 3; k++) pin[k] *= iin;
  cross3(r2v, r1v, out);
  double iout = 1.0 / sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
  for (int k = 0; k < 3; k++) out[k] *= iout;
  for (int k = 0; k < 3; k++) {
    w[0].pos[k] = xa[k] + HB_LENGTH * (costh * pin[k] + sinth * out[k]);
    w[1].pos[k] = xa[k] + HB_LENGTH * (costh * pin[k] - sinth * out[k]);
  }
  double Dp[3][3], D1[3][3], D2[3][3], P1[3][3], P2[3][3], Do[3][3], t1[3][3], t2[3][3];
  mat_unit_der(pin, iin, Dp);
  mat_unit_der(d1, i1, D1);
  mat_unit_der(d2, i2, D2);
  matmul3(Dp, D1, P1);
  matmul3(Dp, D2, P2);
  mat_unit_der(out, iout, Do);
  for (int i = 0; i < 3; i++) { cross3(Do[i], r2v, t1[i]); cross3(r1v, Do[i], t2[i]); }
  for (int s = 0; s < 2; s++) {
    double sg = s == 0 ? 1.0 : -1.0;
    memset(w[s].dpos, 0, sizeof(w[s].dpos));
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double a1 = HB_LENGTH * costh * P1[i][j], a2 = HB_LENGTH * costh * P2[i][j];
        double b1 = HB_LENGTH * sinth * t1[i][j], b2 = HB_LENGTH * sinth * t2[i][j];
        w[s].dpos[1][i][j] = a1 + sg * b1;
        w[s].dpos[2][i][j] = a2 + sg * b2;
        w[s].dpos[0][i][j] = -(a1 + a2) - sg * (b1 + b2) + (i == j);
      }
    w[s].nparents = 3;
    w[s].parent[0] = a; w[s].parent[1] = r1; w[s].parent[2] = r2;
  }
}

static void ws_tetrahedral1_one(const double *X, int a, int r, int rk, WSite *w) {
  /* agbnp3_place_wat_tetrahedral1_one (agbnp3.c:3689-3761) */
  double u[3];
  double inv = unitvec(X + 3 * r, X + 3 * rk, u);
  for (int k = 0; k < 3; k++) w->pos[k] = X[3 * a + k] - HB_LENGTH * u[k];
  memset(w->dpos, 0, sizeof(w->dpos));
  double c = HB_LENGTH * inv;
  for (int i = 0; i < 3; i++) {
    w->dpos[0][i][i] = 1.0;
    for (int j = 0; j < 3; j++) {
      double m = c * ((i == j) - u[i] * u[j]);
      w->dpos[1][i][j] = m;
      w->dpos[2][i][j] = -m;
    }
  }
  w->nparents = 3;
  w->parent[0] = a; w->parent[1] = r; w->parent[2] = rk;
}

/* agbnp3_create_ws_ofatom and the per-type creators (agbnp3.c:1920-3700).
   Returns the number of sites, or -1 for the TRIGONAL_OOP topology error. */
static int place_sites(const OracleSys *s, int iat, const double *X, WSite *w) {
  const int *nb = s->cnl + s->coff[iat];
  int m = s->cnne[iat];
  int t = s->hbtype[iat];
  int ns = 0;
  switch (t) {
    case HB_POLARH: {
      if (m <= 0) return 0;
      int d = -1;
      for (int k = 0; k < m; k++) if (nb[k] < s->nheavy) { d = nb[k]; break; }
      if (d < 0) return 0;
      ns = ws_polarh(X, iat, d, w);
      w[0].type = HB_POLARH;
      break;
    }
    case HB_TRIGONAL:
    case HB_TRIGONAL_S: {
      if (t == HB_TRIGONAL && m == 2) {
        ws_bisector(X, iat, nb, 2, w);
        w[0].type = HB_TRIGONAL2;
        ns = 1;
        break;
      }
      if (m != 1) return 0;
      int r = nb[0];
      if (s->cnne[r] != 3) return 0;
      const int *nr = s->cnl + s->coff[r];
      int r1 = -1, r2 = -1;
      for (int k = 0; k < 3; k++) {
        if (nr[k] == iat) continue;
        if (r1 < 0) r1 = nr[k]; else if (r2 < 0 && nr[k] != r1) r2 = nr[k];
      }
      if (r1 < 0 || r2 < 0) return 0;
      ws_trigonal1_pair(X, iat, r, r1, r2, w);
      w[0].type = w[1].type = (t == HB_TRIGONAL) ? HB_TRIGONAL1 : HB_TRIGONAL_S;
      ns = 2;
      if (t == HB_TRIGONAL_S) {
        ws_trigonal_s_out(X, iat, r, r1, r2, w + 2);
        w[2].type = w[3].type = HB_TRIGONAL_S;
        ns = 4;
      }
      break;
    }
    case HB_TRIGONAL_OOP:
      if (m != 3) {
        fprintf(stderr, "oracle: TRIGONAL_OOP atom %d needs 3 bonds, has %d\n", iat, m);
        return -1;
      }
      ws_trigonal_oop(X, iat, nb[0], nb[1], nb[2], w);
      w[0].type = w[1].type = HB_TRIGONAL_OOP;
      ns = 2;
      break;
    case HB_TETRAHEDRAL:
      if (m == 2) {
        ws_tetrahedral2(X, iat, nb[0], nb[1], w);
        w[0].type = w[1].type = HB_TETRAHEDRAL2;
        ns = 2;
      } else if (m == 3) {
        ws_bisector(X, iat, nb, 3, w);
        w[0].type = HB_TETRAHEDRAL3;
        ns = 1;
      } else if (m == 1) {
        int r = nb[0];
        if (s->cnne[r] != 4) return 0;
        const int *nr = s->cnl + s->coff[r];
        int rk[3], c = 0;
        for (int k = 0; k < 4; k++) if (nr[k] != iat && c < 3) rk[c++] = nr[k];
        if (c != 3) return 0;
        for (int k = 0; k < 3; k++) {
          ws_tetrahedral1_one(X, iat, r, rk[k], w + k);
          w[k].type = HB_TETRAHEDRAL1;
        }
        ns = 3;
      }
      break;
    default:
      return 0;
  }
  for (int k = 0; k < ns; k++) w[k].khb = s->hbcorr[iat];
  return ns;
}

int oracle_place_ws(const OracleSys *s, int iat, const double *X, double *pos, double *dpos,
                    int *parents, int *nparents) {
  WSite w[4];
  memset(w, 0, sizeof(w));
  int ns = place_sites(s, iat, X, w);
  for (int k = 0; k < ns; k++) {
    memcpy(pos + 3 * k, w[k].pos, sizeof(w[k].pos));
    memcpy(dpos + 36 * k, w[k].dpos, sizeof(w[k].dpos));
    memcpy(parents + 4 * k, w[k].parent, sizeof(w[k].parent));
    nparents[k] = w[k].nparents;
  }
  return ns;
}

/* ------------------------------------------------------------ the energy */
int oracle_ener(OracleSys *s, const double *x, const double *y, const double *z, int flags,
                double *energies, double *sp_out, double *br_out, double *sa_out,
                double *dgbdr, double *dvwdr, double *decav, double *dehb) {
  const int n = s->natoms, nh = s->nheavy;
  const double fourpi1 = 1.0 / (4.0 * RPI);
  const double df = -0.5 * (1.0 / s->eps_in - 1.0 / s->eps_out);   /* agbnp3.c:1237-1238 */

  s->spline_fixed = (flags & ORACLE_SPLINE_FIXED) != 0;
  /* coordinates in, FP32 as the reference stores them (agbnp3.c:520-525) */
  for (int i = 0; i < n; i++) {
    int e = s->int2ext[i];
    s->xf[i] = (float)x[e]; s->yf[i] = (float)y[e]; s->zf[i] = (float)z[e];
  }
  double *X = (double *)xcalloc(3 * (size_t)n, sizeof(double));
  for (int i = 0; i < n; i++) { X[3 * i] = s->xf[i]; X[3 * i + 1] = s->yf[i]; X[3 * i + 2] = s->zf[i]; }
  double *ga = (double *)xcalloc(n, sizeof(double));
  double (*gc)[3] = xcalloc(n, sizeof(double[3]));
  for (int i = 0; i < n; i++) {
    double R = s->R[i];
    ga[i] = KFC / (R * R);                                /* agbnp3.c:1087-1091 */
    for (int k = 0; k < 3; k++) gc[i][k] = X[3 * i + k];
  }
  memset(s->g_gb, 0, sizeof(double[3]) * n);
  memset(s->g_vw, 0, sizeof(double[3]) * n);
  memset(s->g_cav, 0, sizeof(double[3]) * n);
  memset(s->g_hb, 0, sizeof(double[3]) * n);

  /* near list, then self volumes and surface areas (agbnp3.c:1326-1388) */
  near_list(s);
  overlap_tree(s, ga, (const double (*)[3])gc);
  for (int i = 0; i < n; i++) {
    double R = s->R[i];
    s->volumep_raw[i] = i < nh ? s->vols[i] : 0.0;       /* agbnp3_utils.c:496-504 */
    s->sa_raw[i] = i < nh ? 4.0 * RPI * R * R : 0.0;
  }
  static const double vcoef[MAXLEVEL] = {1.0, -0.5, 1.0 / 3.0, -0.25, 0.2, -1.0 / 6.0};
  for (long t = 0; t < s->nov; t++) {
    const Ovl *o = &s->ov[t];
    int nord = o->order;
    double c = vcoef[nord - 1], v = nord * c;             /* (-1)^(n+1)/n and (-1)^(n+1) */
    double h3 = v * o->fp * 3.0 / o->a, h4 = 0.5 * v * o->fp;
    for (int m = 0; m < nord; m++) {
      int k = o->atoms[m];
      s->volumep_raw[k] += c * o->vp;
      double u1 = ga[k] * o->v, d2 = 0.0;
      for (int q = 0; q < 3; q++) { double t2 = -2.0 * u1 * (gc[k][q] - o->c[q]); d2 += t2 * t2; }
      s->sa_raw[k] += (h3 * u1 + h4 * d2 / u1) / s->R[k];  /* kernels.c:1554-1569 */
    }
  }

  /* scaling factors (agbnp3_utils.c:524-613) */
  double ecav = 0.0, ecorr_cav = 0.0, molvol = 0.0;
  for (int i = 0; i < n; i++) {
    s->sa_f[i] = s->sp[i] = s->spe[i] = s->psvol[i] = s->gammap[i] = 0.0;
    s->volumep[i] = 0.0;
  }
  for (int i = 0; i < nh; i++) {
    double a = s->sa_raw[i], f, fp;
    if (a < 0.0) { f = 0.0; fp = 0.0; }                    /* agbnp3_swf_area, agbnp3.c:1104-1117 */
    else { double t = a / (64.0 + a * a); f = a * t; fp = 2.0 * t * (1.0 - f); }
    s->sa_f[i] = a * f;
    s->gammap[i] = (s->igamma[i] + s->sgamma[i]) * (f + a * fp);
    s->spe[i] = s->volumep_raw[i] / s->vols[i];
    double R = s->R[i], rvdw = R - RADINC;
    double pr = R * (1.0 - pow(rvdw / R, 3)) / 3.0;
    s->psvol[i] = (fp * a + f) * pr;
    s->volumep[i] = s->volumep_raw[i] - s->sa_f[i] * pr;
    s->sp[i] = s->volumep[i] / s->vols[i];
    ecav += s->igamma[i] * s->sa_f[i];                    /* agbnp3.c:1395-1401 */
    ecorr_cav += s->sgamma[i] * s->sa_f[i];
    molvol += s->volumep[i];                              /* agbnp3.c:1404-1410 */
  }

  /* inverse Born radii: pairwise descreening by heavy atoms
     (agbnp3_cpu_simd_kernels.c:500-678) */
  const int nt = s->ntypes;
  #pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; i++) {
    double b = 1.0 / (s->R[i] - RADINC);                  /* agbnp3_utils.c:507-511 */
    double acc = 0.0;
    for (int j = 0; j < nh; j++) {
      if (j == i) continue;
      double dx = X[3 * j] - X[3 * i], dy = X[3 * j + 1] - X[3 * i + 1], dz = X[3 * j + 2] - X[3 * i + 2];
      double d = sqrt(dx * dx + dy * dy + dz * dz), dq;
      double q4 = oracle_q4(s, s->rtype[i] * nt + s->rtype[j], d, s->R[j], &dq);
      acc += q4 * s->sp[j];
    }
    s->beta[i] = b - fourpi1 * acc;
  }

  /* Born radii and np quantities (agbnp3_utils.c:615-670, agbnp3.c:1174-1188) */
  double evdw = 0.0, ecorr_vdw = 0.0;
  for (int i = 0; i < n; i++) {
    double b = s->beta[i], bf, fp;
    if (b < 0.0) { bf = 0.02; fp = 0.0; }
    else { bf = sqrt(0.02 * 0.02 + b * b); fp = b / bf; }
    s->swf[i] = fp;
    s->br[i] = 1.0 / bf;
    double B = s->br[i], t = B + RW_NP;
    s->brw[i] = 3.0 * B * B / (t * t * t * t);             /* AGBNP_BRW, agbnp3_private.h:772 */
    double a3 = 1.0 / (t * t * t);
    evdw += s->ialpha[i] * a3;                             /* agbnp3.c:1478-1484 */
    ecorr_vdw += s->salpha[i] * a3;
  }

  /* GB energy, A_i and gradient at constant Born radii
     (agbnp3_cpu_simd_kernels.c:66-151, 329-460) */
  double eself = 0.0, epair = 0.0;
  for (int i = 0; i < n; i++) eself += df * s->q[i] * s->q[i] / s->br[i];
  #pragma omp parallel for schedule(dynamic, 16) reduction(+ : epair)
  for (int i = 0; i < n; i++) {
    double xi = X[3 * i], yi = X[3 * i + 1], zi = X[3 * i + 2], qi = s->q[i], bi = s->br[i];
    double gx = 0, gy = 0, gz = 0, ai = 0, ei = 0;
    for (int j = 0; j < n; j++) {
      if (j == i) continue;
      double dx = X[3 * j] - xi, dy = X[3 * j + 1] - yi, dz = X[3 * j + 2] - zi;
      double d2 = dx * dx + dy * dy + dz * dz;
      if (!(d2 > 0.0)) continue;
      double qq = qi * s->q[j], qqf = 2.0 * df * qq, bb = bi * s->br[j];
      double et = exp(-0.25 * d2 / bb);
      double fgb = 1.0 / sqrt(d2 + bb * et);
      double fgb3 = fgb * fgb * fgb;
      double mw = -qqf * (1.0 - 0.25 * et) * fgb3;
      ei += qqf * fgb;
      gx -= mw * dx; gy -= mw * dy; gz -= mw * dz;
      ai += qq * (bb + 0.25 * d2) * et * fgb3;
    }
    s->g_gb[i][0] = gx; s->g_gb[i][1] = gy; s->g_gb[i][2] = gz;
    s->A[i] = ai;
    epair += 0.5 * ei;
  }
  double egb = TOKCAL * (eself + epair);                      /* agbnp3.c:1513-1520 */
  for (int i = 0; i < n; i++) {
    s->q2ab[i] = (s->q[i] * s->q[i] + s->A[i] * s->br[i]) * s->swf[i];   /* kernels.c:442-448 */
    s->abrw[i] = (s->ialpha[i] + s->salpha[i]) * s->brw[i] * s->swf[i];
  }

  /* chain rule at constant self volumes (kernels.c:681-805) and U_i, V_i
     (kernels.c:809-904), row-owner form over directed descreening pairs */
  #pragma omp parallel for schedule(dynamic, 16)
  for (int k = 0; k < n; k++) {
    double gbx = 0, gby = 0, gbz = 0, vwx = 0, vwy = 0, vwz = 0, U = 0, V = 0;
    int kh = k < nh;
    for (int j = 0; j < n; j++) {
      if (j == k) continue;
      int jh = j < nh;
      if (!kh && !jh) continue;
      double dx = X[3 * j] - X[3 * k], dy = X[3 * j + 1] - X[3 * k + 1], dz = X[3 * j + 2] - X[3 * k + 2];
      double d = sqrt(dx * dx + dy * dy + dz * dz);
      double hgb = 0.0, hvw = 0.0;
      if (jh) {            /* k descreened by j */
        double dq;
        oracle_q4(s, s->rtype[k] * nt + s->rtype[j], d, s->R[j], &dq);
        hgb += s->q2ab[k] * dq * s->sp[j];
        hvw += s->abrw[k] * dq * s->sp[j];
      }
      if (kh) {            /* j descreened by k */
        double dq;
        double q4 = oracle_q4(s, s->rtype[j] * nt + s->rtype[k], d, s->R[k], &dq);
        hgb += s->q2ab[j] * dq * s->sp[k];
        hvw += s->abrw[j] * dq * s->sp[k];
        U += s->q2ab[j] * q4;
        V += s->abrw[j] * q4;
      }
      hgb *= fourpi1 * df / d;
      hvw *= fourpi1 / d;
      gbx += hgb * dx; gby += hgb * dy; gbz += hgb * dz;
      vwx += hvw * dx; vwy += hvw * dy; vwz += hvw * dz;
    }
    s->g_gb[k][0] += gbx; s->g_gb[k][1] += gby; s->g_gb[k][2] += gbz;
    s->g_vw[k][0] = vwx; s->g_vw[k][1] = vwy; s->g_vw[k][2] = vwz;
    if (kh) {
      s->deru[k] = U / s->vols[k];
      s->derv[k] = V / s->vols[k];
      s->derus[k] = df * fourpi1 * s->psvol[k] * s->deru[k];
      s->dervs[k] = fourpi1 * s->psvol[k] * s->derv[k];
    } else {
      s->deru[k] = s->derv[k] = s->derus[k] = s->dervs[k] = 0.0;
    }
  }

  /* water sites (agbnp3.c:1837-1914) in internal atom order */
  s->nws = 0;
  for (int i = 0; i < n; i++) {
    if (s->hbtype[i] == HB_INACTIVE) continue;
    if (s->nws + 4 > s->wscap) {
      s->wscap = 2 * s->wscap + 64;
      s->ws = (WSite *)realloc(s->ws, sizeof(WSite) * s->wscap);
    }
    WSite *w = s->ws + s->nws;
    memset(w, 0, 4 * sizeof(WSite));
    int ns = place_sites(s, i, X, w);
    if (ns < 0) { free(X); free(ga); free(gc); return -1; }
    s->nws += ns;
  }

  /* HB free volumes and energy (kernels.c:2637-2954) */
  double ehb = 0.0;
  memset(s->derh, 0, sizeof(double) * n);
  {
    const double aw = KFC / (HB_RADIUS * HB_RADIUS);
    const double vw = 4.0 / 3.0 * RPI * pow(HB_RADIUS, 3);
    int stale = nh;                   /* index the reference's loop leaves behind */
    int *bi = (int *)xcalloc(nh + 1, sizeof(int));
    double *bv = (double *)xcalloc(nh + 1, sizeof(double));
    double (*bgw)[3] = xcalloc(nh + 1, sizeof(double[3]));
    double (*bgi)[3] = xcalloc(nh + 1, sizeof(double[3]));
    for (int w = 0; w < s->nws; w++) {
      WSite *ws = s->ws + w;
      double freev = vw, cutoff = HB_SWA0 * vw;
      int nc = 0;
      for (int k = 0; k < nh; k++) {
        double dx = X[3 * k] - ws->pos[0], dy = X[3 * k + 1] - ws->pos[1], dz = X[3 * k + 2] - ws->pos[2];
        double d2 = dx * dx + dy * dy + dz * dz;
        double u = (s->R[k] + HB_RADIUS) * 1.6;
        if (!(d2 < u * u)) continue;
        if (!(freev > cutoff)) continue;
        double out[8];
        double v = oracle_ogauss(aw, PFC, ws->pos, ga[k], PFC, gc[k], out);
        if (!(out[5] > 0.0)) continue;
        int idx = (flags & ORACLE_HB_FIXED) ? k : stale;   /* quirk Q3 */
        freev -= (idx < n ? s->spe[idx] : 0.0) * out[5];
        stale = k;
        bi[nc] = k;
        bv[nc] = out[5];
        for (int q = 0; q < 3; q++) {
          bgw[nc][q] = -2.0 * aw * v * out[6] * (ws->pos[q] - out[2 + q]);
          bgi[nc][q] = -2.0 * ga[k] * v * out[6] * (gc[k][q] - out[2 + q]);
        }
        nc++;
      }
      ws->sp = freev / vw;
      double fp, sw = pol_switch(ws->sp, HB_SWA, HB_SWB, &fp);
      double dhw = ws->khb * fp / vw;
      ehb += ws->khb * sw;
      double jw[3] = {0, 0, 0};
      for (int c = 0; c < nc; c++) {
        int k = bi[c];
        for (int q = 0; q < 3; q++) {
          s->g_hb[k][q] += -s->spe[k] * dhw * bgi[c][q];
          jw[q] += -dhw * s->spe[k] * bgw[c][q];
        }
        s->derh[k] += dhw / s->vols[k] * bv[c];
      }
      for (int p = 0; p < ws->nparents; p++) {
        int k = ws->parent[p];
        for (int j = 0; j < 3; j++) {
          double g = 0.0;
          for (int i = 0; i < 3; i++) g += jw[i] * ws->dpos[p][i][j];   /* agbnp3_rtvec */
          s->g_hb[k][j] += g;
        }
        stale = k;
      }
    }
    free(bi); free(bv); free(bgw); free(bgi);
  }

  /* derivatives through the self volumes (kernels.c:1601-1956) and the
     surface areas (kernels.c:1959-2405), walking the stored tree */
  for (long t = 0; t < s->nov; t++) {
    const Ovl *o = &s->ov[t];
    int nord = o->order;
    double c = vcoef[nord - 1], v = nord * c;
    double U = 0, Vv = 0, H = 0;
    for (int m = 0; m < nord; m++) {
      int k = o->atoms[m];
      U += s->deru[k]; Vv += s->derv[k]; H += s->derh[k];
    }
    U *= -df * fourpi1 * c;
    Vv *= -fourpi1 * c;
    H *= -c;
    double gdr[MAXLEVEL][3], gdR[MAXLEVEL];
    double dai = 1.0 / o->a;
    for (int m = 0; m < nord; m++) {
      int k = o->atoms[m];
      double u1 = ga[k] * o->v, d2 = 0.0;
      for (int q = 0; q < 3; q++) {
        gdr[m][q] = -2.0 * u1 * (gc[k][q] - o->c[q]);
        d2 += gdr[m][q] * gdr[m][q];
        double ur = gdr[m][q] * o->fp;
        s->g_gb[k][q] += U * ur;
        s->g_vw[k][q] += Vv * ur;
        s->g_hb[k][q] += H * ur;
      }
      gdR[m] = (3.0 * dai * u1 + 0.5 * d2 / u1) / s->R[k];
    }
    double h3 = o->fpp + o->fp / o->v, h4 = 2.0 * o->fp * dai;
    for (int m = 0; m < nord; m++) {
      int k = o->atoms[m];
      double u1 = -2.0 * o->fp * (1.0 - ga[k] * dai) / s->R[k] + h3 * gdR[m];
      for (int q = 0; q < 3; q++) {
        double ur = u1 * gdr[m][q];
        s->g_cav[k][q] += v * s->gammap[k] * ur;
        s->g_vw[k][q] += v * s->dervs[k] * ur;
        s->g_gb[k][q] += v * s->derus[k] * ur;
      }
      for (int l = 0; l < nord; l++) {
        if (l == m) continue;
        int kl = o->atoms[l];
        double w1 = h4 * ga[k] / s->R[kl], w2 = h3 * gdR[l];
        for (int q = 0; q < 3; q++) {
          double ur = w1 * gdr[l][q] + w2 * gdr[m][q];
          s->g_cav[k][q] += v * s->gammap[kl] * ur;
          s->g_vw[k][q] += v * s->dervs[kl] * ur;
          s->g_gb[k][q] += v * s->derus[kl] * ur;
        }
      }
    }
  }

  /* outputs in caller order (agbnp3.c:537-563, 1762-1786) */
  s->e[ORC_EGB] = egb;
  s->e[ORC_EVDW] = evdw;
  s->e[ORC_ECORR_VDW] = ecorr_vdw;
  s->e[ORC_ECAV] = ecav;
  s->e[ORC_ECORR_CAV] = ecorr_cav;
  s->e[ORC_EHB] = ehb;
  s->e[ORC_MOLVOL] = molvol;
  if (energies) memcpy(energies, s->e, sizeof(s->e));
  for (int i = 0; i < n; i++) {
    int e = s->int2ext[i];
    if (sp_out) sp_out[e] = i < nh ? s->sp[i] : 1.0;
    if (br_out) br_out[e] = s->br[i];
    if (sa_out) sa_out[e] = i < nh ? s->sa_f[i] : 0.0;
    for (int q = 0; q < 3; q++) {
      if (dgbdr) dgbdr[3 * e + q] = TOKCAL * s->g_gb[i][q];
      if (dvwdr) dvwdr[3 * e + q] = s->g_vw[i][q];
      if (decav) decav[3 * e + q] = s->g_cav[i][q];
      if (dehb) dehb[3 * e + q] = s->g_hb[i][q];
    }
  }
  free(X); free(ga); free(gc);
  return 0;
}

/* --------------------------------------------------------- introspection */
int oracle_natoms(const OracleSys *s) { return s->natoms; }
int oracle_nheavy(const OracleSys *s) { return s->nheavy; }
void oracle_int2ext(const OracleSys *s, int *int2ext) {
  memcpy(int2ext, s->int2ext, sizeof(int) * s->natoms);
}
int oracle_near_list(const OracleSys *s, int *nne, int *nbr, int cap) {
  if (!s->nne) return -1;
  int maxn = 0;
  for (int i = 0; i < s->nheavy; i++) {
    if (s->nne[i] > maxn) maxn = s->nne[i];
    if (nne) nne[i] = s->nne[i];
    if (nbr)
      for (int k = 0; k < s->nne[i] && k < cap; k++) nbr[(size_t)i * cap + k] = s->nbr[(size_t)i * s->nbrcap + k];
  }
  return maxn;
}
int oracle_tables(const OracleSys *s, float *y, float *y2, float *radii) {
  size_t m = (size_t)s->ntypes * s->ntypes * TBL_N;
  if (y) memcpy(y, s->ty, m * sizeof(float));
  if (y2) memcpy(y2, s->ty2, m * sizeof(float));
  if (radii) memcpy(radii, s->radii, s->ntypes * sizeof(float));
  return s->ntypes;
}
int oracle_rtype(const OracleSys *s, int *rtype) {
  memcpy(rtype, s->rtype, sizeof(int) * s->natoms);
  return s->ntypes;
}
int oracle_internal(const OracleSys *s, const char *name, double *out) {
  const double *src = NULL;
#define PICK(f) if (!strcmp(name, #f)) src = s->f
  PICK(volumep_raw); PICK(volumep); PICK(sa_raw); PICK(sa_f); PICK(sp); PICK(spe);
  PICK(psvol); PICK(gammap); PICK(beta); PICK(br); PICK(brw); PICK(swf); PICK(A);
  PICK(q2ab); PICK(abrw); PICK(deru); PICK(derv); PICK(derus); PICK(dervs); PICK(derh);
#undef PICK
  if (!src) return -1;
  memcpy(out, src, sizeof(double) * s->natoms);
  return 0;
}
void oracle_counts(const OracleSys *s, long *counts) {
  long nl = 0;
  for (int i = 0; s->nne && i < s->nheavy; i++) nl += s->nne[i];
  counts[0] = s->nov;
  counts[1] = s->neval;
  counts[2] = s->nws;
  counts[3] = nl;
}
int oracle_water_sites(const OracleSys *s, double *pos, int *parent0, int *type, double *sp) {
  for (int w = 0; w < s->nws; w++) {
    if (pos) memcpy(pos + 3 * w, s->ws[w].pos, sizeof(double) * 3);
    if (parent0) parent0[w] = s->ws[w].parent[0];
    if (type) type[w] = s->ws[w].type;
    if (sp) sp[w] = s->ws[w].sp;
  }
  return s->nws;
}

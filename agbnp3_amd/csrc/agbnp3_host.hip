// Host side of libagbnp3: the agbnp3.h C API (instance registry, agbnp3_new /
// agbnp3_ener / agbnp3_delete), the Q4 spline tables, the static water-site
// topology, device arenas and the per-call kernel pipeline on one stream.
#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "agb_common.h"
#include "agbnp3.h"
#include "agbnp3_amd.h"

using namespace agb;

namespace {

void errprint(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      errprint("agbnp3: HIP error %s at %s:%d: %s\n", hipGetErrorString(e_), __FILE__, \
               __LINE__, #expr);                                                       \
      return AGBNP_ERR;                                                                \
    }                                                                                  \
  } while (0)

// ------------------------------------------------------------ Q4 tables
// The tables are model data; they are built with the reference's own mixed
// float/double arithmetic so every node is the reference's FP32 value:
// agbnp3_i4 (agbnp3.c:3886-3944), agbnp3_i4ov (agbnp3.c:3964-3975, both
// Gaussian exponents taken from Ri), agbnp3_create_ctablef4
// (agbnp3_utils.c:380-422: 511 nodes on [0, 20) at an accumulated spacing,
// y[511] = 0) and agbnp3_cspline_setup (agbnp3_utils.c:193-227, whose end
// condition divides by 2 dx).  The reference is C, where log/exp/pow of a
// float promote to double; the explicit casts keep that under C++ overloads.
float q4_integral(float rij, float Ri, float Rj, float *dr) {
  const float twopi = 2.0 * kPi;
  const float twothirds = 2.0 / 3.0;
  const float rij2 = rij * rij;
  float q;
  if (rij > (Ri + Rj)) {                         // no overlap
    float u1 = rij + Rj, u2 = rij - Rj, u3 = u1 * u2;
    float u4 = 0.5 * log((double)(u1 / u2));
    q = twopi * (Rj / u3 - u4 / rij);
    *dr = twopi * ((Rj / (rij * u3)) * (1. - 2. * rij2 / u3) + u4 / rij2);
    return q;
  }
  float u1 = Rj - Ri;
  if (rij2 > u1 * u1) {                          // partial overlap
    u1 = rij + Rj;
    float u2 = rij - Rj, u3 = u1 * u2;
    float u4 = 1. / u1, u4sq = u4 * u4;
    float u5 = 1. / Ri, u5sq = u5 * u5;
    float u6 = 0.5 * log((double)(u1 / Ri));
    q = twopi * (-(u4 - u5) + (0.25 * u3 * (u4sq - u5sq) - u6) / rij);
    *dr = twopi * (0.5 * (1. - 0.5 * u3 / rij2) * (u4sq - u5sq) + u6 / rij2);
    return q;
  }
  if (Ri > Rj) {                                 // i contains j
    *dr = 0.0f;
    return 0.0f;
  }
  u1 = rij + Rj;                                 // j contains i
  float u2 = Rj - rij, u3 = -u1 * u2;
  if (rij < .001 * Rj) {                         // series at the centre
    float a = rij / Rj, ad = a * a - 1;
    float u6 = (1. + twothirds * a * a) / Rj;
    q = twopi * (2. / Ri + Rj / u3 - u6);
    *dr = -(2. * twopi * a / (Rj * Rj)) * (1. / ad + twothirds);
  } else {
    float u6 = 0.5 * log((double)(u1 / u2));
    q = twopi * (2. / Ri + Rj / u3 - u6 / (rij));
    *dr = twopi * (-(Rj / u3) * (2. * rij / u3 - 1. / (rij)) + u6 / (rij2));
  }
  return q;
}

float q4_inflated(float rij, float Ri, float Rj, float *dr) {
  const float ai = kKFC / (Ri * Ri);
  const float pfc = kPFC;
  const float d2 = rij * rij;
  const float deltai = 1. / (ai + ai);
  const float p = pfc * pfc;
  const float kappa = exp((double)(-ai * ai * d2 * deltai));
  const float gvol = p * kappa * pow(kPi * deltai, 1.5);
  const float volj = 4. * kPi * Rj * Rj * Rj / 3.;
  const float newRj = pow((volj + 2. * gvol) / volj, 1. / 3.) * Rj;
  return q4_integral(rij, Ri, newRj, dr);
}

void spline_coeffs(float dx, int n, const float *y, float yp1, float ypn, float *y2) {
  std::vector<float> u(n);
  y2[0] = -0.5;
  u[0] = (3.0 / (dx)) * ((y[1] - y[0]) / (dx)-yp1);
  for (int i = 1; i < n - 1; i++) {
    const float sig = 0.5;
    const float p = sig * y2[i - 1] + 2.0;
    y2[i] = (sig - 1.0) / p;
    u[i] = (y[i + 1] - y[i]) / (dx) - (y[i] - y[i - 1]) / (dx);
    u[i] = (6.0 * u[i] / (2. * dx) - sig * u[i - 1]) / p;
  }
  const float qn = 0.5;
  const float un = (3.0 / (dx)) * (ypn - (y[n - 1] - y[n - 2]) / (2. * dx));
  y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
  for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
}

void build_table(float b, float *y, float *y2, float &dx, float &dxinv) {
  const int n = kTblN;
  const float amax = kTblAmax;
  const float da = amax / (n - 1);
  float a = 0.0f, dr = 0.0f, yp1 = 0.0f;
  for (int i = 0; i < n - 1; i++) {
    const float q = q4_inflated(a, b, 1.0f, &dr);
    if (i == 0) yp1 = dr;
    y[i] = q;
    a += da;
  }
  y[n - 1] = 0.0f;
  spline_coeffs(da, n, y, yp1, 0.0f, y2);
  dx = da;
  dxinv = 1. / da;
}

// ------------------------------------------------------------- instance
struct Instance {
  int natoms = 0, nheavy = 0;
  int verbose = 0;
  int options = 0;
  int ws_error_atom = -1;              // TRIGONAL_OOP topology error (reported by ener)
  std::vector<int> int2ext, ext2int;
  std::vector<float> R, q, igamma, sgamma, ialpha, salpha, hbcorr, galpha;
  std::vector<int> hbtype, rtype;
  std::vector<double> vols;
  std::vector<int> cnne, coff, cnl;    // connection table, internal indices
  int ntypes = 0;
  std::vector<float> radii, ty, ty2;
  float tdx = 0, tdxinv = 0;
  float eps_in = 1, eps_out = 1;
  std::vector<int> row_perm, block_type, block_hrows;
  std::vector<WsDef> wsdef;
  std::vector<int> ws_stale0;
  int nws = 0;
  // device
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  Dev d{};
  double *dx = nullptr, *dy = nullptr, *dz = nullptr;      // coordinate staging (device)
  double *dout = nullptr;                                  // outputs: 3n + 12n + 7
  double *hstage = nullptr;                                // pinned host staging
  size_t hstage_len = 0;
  // timing mode: one event set per recorded call, averaged when read back
  std::vector<std::vector<hipEvent_t>> ev_sets;
  std::vector<std::string> ev_names;
  int ncalls_timed = 0;
};

std::mutex g_mutex;
bool g_initialized = false;
std::vector<Instance *> g_slots;

Instance *get_instance(int tag, const char *who) {
  if (!g_initialized) {
    errprint("%s: agbnp library is not initialized.\n", who);
    return nullptr;
  }
  if (tag < 0 || tag >= (int)g_slots.size() || !g_slots[tag]) {
    errprint("%s: invalid tag %d.\n", who, tag);
    return nullptr;
  }
  return g_slots[tag];
}

// device allocation helpers: 0 on success, non-zero on failure
template <typename T>
int dalloc(Instance *I, T **p, size_t count) {
  void *q = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e != hipSuccess) {
    errprint("agbnp3_new(): device allocation of %zu bytes failed: %s\n", count * sizeof(T), hipGetErrorString(e));
    return 1;
  }
  I->allocs.push_back(q);
  *p = (T *)q;
  return 0;
}

template <typename T>
int upload(Instance *I, const T **dst, const std::vector<T> &src) {
  T *p = nullptr;
  if (dalloc(I, &p, src.size())) return 1;
  if (!src.empty() && hipMemcpy(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
    errprint("agbnp3_new(): upload of %zu bytes failed\n", src.size() * sizeof(T));
    return 1;
  }
  *dst = p;
  return 0;
}

void free_instance(Instance *I) {
  if (!I) return;
  if (I->stream) (void)hipStreamSynchronize(I->stream);
  for (void *p : I->allocs) (void)hipFree(p);
  if (I->hstage) (void)hipHostFree(I->hstage);
  for (auto &set : I->ev_sets)
    for (hipEvent_t e : set) (void)hipEventDestroy(e);
  if (I->stream) (void)hipStreamDestroy(I->stream);
  delete I;
}

// static water-site topology: geometry per HB-active atom, following the
// dispatch of agbnp3_create_ws_ofatom (agbnp3.c:1920-2012); atoms whose
// topology does not fit get no sites, as in the reference
void build_water_sites(Instance *I) {
  const int n = I->natoms, nh = I->nheavy;
  I->wsdef.clear();
  I->ws_stale0.clear();
  I->ws_error_atom = -1;
  int nsite = 0;
  auto nb = [&](int a, int k) { return I->cnl[I->coff[a] + k]; };
  for (int a = 0; a < n; a++) {
    const int t = I->hbtype[a], m = I->cnne[a];
    WsDef w{};
    w.atom = a;
    w.p[0] = a;
    int ns = 0;
    if (t == AGBNP_HB_POLARH) {
      int dd = -1;
      for (int k = 0; k < m; k++)
        if (nb(a, k) < nh) { dd = nb(a, k); break; }
      if (dd >= 0) { w.kind = WS_POLARH; w.p[1] = dd; ns = 1; }
    } else if (t == AGBNP_HB_TRIGONAL && m == 2) {
      w.kind = WS_TRIGONAL2; w.p[1] = nb(a, 0); w.p[2] = nb(a, 1); ns = 1;
    } else if ((t == AGBNP_HB_TRIGONAL || t == AGBNP_HB_TRIGONAL_S) && m == 1) {
      const int r = nb(a, 0);
      if (I->cnne[r] == 3) {
        int r1 = -1, r2 = -1;
        for (int k = 0; k < 3; k++) {
          const int c = nb(r, k);
          if (c == a) continue;
          if (r1 < 0) r1 = c;
          else if (r2 < 0 && c != r1) r2 = c;
        }
        if (r1 >= 0 && r2 >= 0) {
          w.kind = t == AGBNP_HB_TRIGONAL ? WS_TRIGONAL1 : WS_TRIGONAL_S;
          w.p[1] = r; w.p[2] = r1; w.p[3] = r2;
          ns = t == AGBNP_HB_TRIGONAL ? 2 : 4;
        }
      }
    } else if (t == AGBNP_HB_TRIGONAL_OOP) {
      if (m != 3) {
        if (I->ws_error_atom < 0) I->ws_error_atom = a;
      } else {
        w.kind = WS_TRIGONAL_OOP; w.p[1] = nb(a, 0); w.p[2] = nb(a, 1); w.p[3] = nb(a, 2); ns = 2;
      }
    } else if (t == AGBNP_HB_TETRAHEDRAL) {
      if (m == 2) {
        w.kind = WS_TETRAHEDRAL2; w.p[1] = nb(a, 0); w.p[2] = nb(a, 1); ns = 2;
      } else if (m == 3) {
        w.kind = WS_TETRAHEDRAL3; w.p[1] = nb(a, 0); w.p[2] = nb(a, 1); w.p[3] = nb(a, 2); ns = 1;
      } else if (m == 1) {
        const int r = nb(a, 0);
        if (I->cnne[r] == 4) {
          int c = 0;
          w.p[1] = r;
          for (int k = 0; k < 4 && c < 3; k++)
            if (nb(r, k) != a) w.p[2 + c++] = nb(r, k);
          if (c == 3) { w.kind = WS_TETRAHEDRAL1; ns = 3; }
        }
      }
    }
    if (ns == 0) continue;
    w.nsite = ns;
    w.first = nsite;
    I->wsdef.push_back(w);
    // last parent of each site: the stale index the next site starts from
    for (int s = 0; s < ns; s++) {
      int last;
      switch (w.kind) {
        case WS_POLARH: last = w.p[1]; break;
        case WS_TRIGONAL1: case WS_TRIGONAL_S: case WS_TRIGONAL_OOP: case WS_TETRAHEDRAL3: last = w.p[3]; break;
        case WS_TRIGONAL2: case WS_TETRAHEDRAL2: last = w.p[2]; break;
        default: last = w.p[2 + s]; break;      // tetrahedral1: (A, R, R_k)
      }
      I->ws_stale0.push_back(last);
    }
    nsite += ns;
  }
  // site s reads the last parent of site s-1 first; site 0 the index after
  // the heavy atoms (a hydrogen, spe = 0)
  std::vector<int> st(nsite);
  for (int s = 0; s < nsite; s++) st[s] = s == 0 ? nh : I->ws_stale0[s - 1];
  I->ws_stale0 = st;
  I->nws = nsite;
}

int choose_chunks(int rowblocks, int ncols) {
  int c = (1024 + rowblocks - 1) / std::max(rowblocks, 1);
  c = std::max(1, std::min(c, std::max(1, ncols / 512)));
  return c;
}

int setup_device(Instance *I) {
  const int n = I->natoms, nh = I->nheavy;
  Dev &d = I->d;
  d = Dev{};
  d.n = n;
  d.nh = nh;
  d.ntypes = I->ntypes;
  d.cap = 128;
  d.ncell_max = 8192;
  d.options = I->options;
  d.df = -0.5 * (1.0 / (double)I->eps_in - 1.0 / (double)I->eps_out);
  d.eps_df = (float)d.df;
  float rmax = 0.f;
  for (int i = 0; i < nh; i++) rmax = std::max(rmax, I->R[i]);
  d.rmax_heavy = rmax;
  if (upload(I, &d.int2ext, I->int2ext) || upload(I, &d.R, I->R) || upload(I, &d.q, I->q) ||
      upload(I, &d.rtype, I->rtype) || upload(I, &d.galpha, I->galpha) || upload(I, &d.igamma, I->igamma) ||
      upload(I, &d.sgamma, I->sgamma) || upload(I, &d.ialpha, I->ialpha) || upload(I, &d.salpha, I->salpha) ||
      upload(I, &d.hbcorr_f, I->hbcorr) || upload(I, &d.vols, I->vols))
    return AGBNP_ERR;
  {
    std::vector<float2> tb(I->ty.size());
    for (size_t k = 0; k < tb.size(); k++) tb[k] = make_float2(I->ty[k], I->ty2[k]);
    if (upload(I, &d.tbl, tb)) return AGBNP_ERR;
    d.tdx = I->tdx;
    d.tdxinv = I->tdxinv;
  }
  // rows grouped by radius type, 256-row blocks of a single type
  I->row_perm.clear();
  I->block_type.clear();
  I->block_hrows.clear();
  for (int t = 0; t < I->ntypes; t++) {
    std::vector<int> rows;
    for (int i = 0; i < n; i++)
      if (I->rtype[i] == t) rows.push_back(i);
    for (size_t b = 0; b < rows.size(); b += 256) {
      int allh = 1;
      for (size_t k = 0; k < 256; k++) {
        const int r = b + k < rows.size() ? rows[b + k] : -1;
        I->row_perm.push_back(r);
        if (r >= 0 && r < nh) allh = 0;
      }
      I->block_type.push_back(t);
      I->block_hrows.push_back(allh);
    }
  }
  d.nrow_blocks = (int)I->block_type.size();
  if (upload(I, &d.row_perm, I->row_perm) || upload(I, &d.block_type, I->block_type) ||
      upload(I, &d.block_hrows, I->block_hrows))
    return AGBNP_ERR;
  d.nwsdef = (int)I->wsdef.size();
  d.nws = I->nws;
  d.ws_cap = 128;
  if (upload(I, &d.wsdef, I->wsdef) || upload(I, &d.ws_stale0, I->ws_stale0)) return AGBNP_ERR;
  // per-call buffers
  const int nblk = (n + 255) / 256;
  d.nchunk_desc = choose_chunks(d.nrow_blocks, nh);
  {
    std::vector<int2> tiles;
    const int nb = (n + 255) / 256;
    for (int a = 0; a < nb; a++)
      for (int b = a; b < nb; b++) tiles.push_back(make_int2(a, b));
    d.ntiles = (int)tiles.size();
    if (upload(I, &d.tile_ij, tiles)) return AGBNP_ERR;
  }
  d.nchunk_ch = choose_chunks(d.nrow_blocks, n);
  d.epart_stride = nblk;
  const size_t nn = n > 0 ? n : 1, nhh = nh > 0 ? nh : 1;
  if (dalloc(I, &d.pos, nn) || dalloc(I, &d.bbox_part, 6 * (size_t)nblk) || dalloc(I, &d.grid, 1) ||
      dalloc(I, &d.cell_start, d.ncell_max + 2 + nhh) || dalloc(I, &d.cell_atoms, nhh) ||
      dalloc(I, &d.nne, nhh) || dalloc(I, &d.nbr, nhh * 256) || dalloc(I, &d.status, 1) ||
      dalloc(I, &d.acc_vol, nhh) || dalloc(I, &d.acc_sa, nhh) || dalloc(I, &d.acc_grad, nhh * TG_N) ||
      dalloc(I, &d.acc_hb, 3 * nn) || dalloc(I, &d.acc_derh, nhh) || dalloc(I, &d.counters, 4) ||
      dalloc(I, &d.volumep_raw, nn) || dalloc(I, &d.sa_raw, nn) || dalloc(I, &d.sa_f, nn) ||
      dalloc(I, &d.sp, nn) || dalloc(I, &d.spe, nn) || dalloc(I, &d.psvol, nn) || dalloc(I, &d.gammap, nn) ||
      dalloc(I, &d.sp_f, nn) || dalloc(I, &d.spe_f, nn) ||
      dalloc(I, &d.beta_part, (size_t)d.nchunk_desc * nn) || dalloc(I, &d.beta, nn) || dalloc(I, &d.br, nn) ||
      dalloc(I, &d.brw, nn) || dalloc(I, &d.swf, nn) || dalloc(I, &d.br_f, nn) ||
      dalloc(I, &d.acc_gbA, nn) || dalloc(I, &d.acc_gbg, 3 * nn) ||
      dalloc(I, &d.gb_epart, (size_t)std::max(1, d.ntiles)) || dalloc(I, &d.A, nn) || dalloc(I, &d.q2ab, nn) ||
      dalloc(I, &d.abrw, nn) || dalloc(I, &d.q2ab_f, nn) || dalloc(I, &d.abrw_f, nn) ||
      dalloc(I, &d.ch_part, (size_t)d.nchunk_ch * 8 * nn) || dalloc(I, &d.deru, nn) || dalloc(I, &d.derv, nn) ||
      dalloc(I, &d.derus, nn) || dalloc(I, &d.dervs, nn) || dalloc(I, &d.derh, nn) ||
      dalloc(I, &d.deru_f, nn) || dalloc(I, &d.derv_f, nn) || dalloc(I, &d.derus_f, nn) ||
      dalloc(I, &d.dervs_f, nn) || dalloc(I, &d.derh_f, nn) || dalloc(I, &d.gammap_f, nn) ||
      dalloc(I, &d.ws, (size_t)std::max(1, I->nws)) || dalloc(I, &d.ws_part, (size_t)std::max(1, I->nws)) ||
      dalloc(I, &d.epart, (size_t)EP_COUNT * nblk) || dalloc(I, &I->dx, nn) || dalloc(I, &I->dy, nn) ||
      dalloc(I, &I->dz, nn) || dalloc(I, &I->dout, 15 * nn + 7))
    return AGBNP_ERR;
  d.cell_count = nullptr;
  HIPCHK(hipMemset(d.epart, 0, sizeof(double) * EP_COUNT * nblk));
  HIPCHK(hipMemset(d.ws_part, 0, sizeof(double) * std::max(1, I->nws)));
  I->hstage_len = 15 * nn + 7 + 3 * nn;
  HIPCHK(hipHostMalloc((void **)&I->hstage, I->hstage_len * sizeof(double), hipHostMallocDefault));
  HIPCHK(hipStreamCreateWithFlags(&I->stream, hipStreamNonBlocking));
  return AGBNP_OK;
}

// ---------------------------------------------------------- the pipeline
struct Step {
  const char *name;
};

int run_pipeline(Instance *I, const double *x, const double *y, const double *z, double *sp, double *br, double *sa,
                 double *dgbdr, double *dvwdr, double *decav, double *dehb, double *energies) {
  Dev &d = I->d;
  d.options = I->options;
  hipStream_t s = I->stream;
  const bool timed = (I->options & AGBNP3_OPT_TIMING) != 0;
  int k = 0;
  std::vector<hipEvent_t> *set = nullptr;
  if (timed) {
    if ((size_t)I->ncalls_timed >= I->ev_sets.size()) I->ev_sets.emplace_back();
    set = &I->ev_sets[I->ncalls_timed++];
  }
  auto mark = [&](const char *name) {
    if (!timed) return;
    if ((size_t)k >= set->size()) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      set->push_back(e);
    }
    if ((size_t)k >= I->ev_names.size()) I->ev_names.push_back(name);
    I->ev_names[k] = name;
    (void)hipEventRecord((*set)[k], s);
    k++;
  };
  mark("start");
  launch_ingest(d, x, y, z, s);
  mark("ingest");
  launch_cells(d, s);
  mark("cells");
  launch_nblist(d, s);
  mark("nblist");
  launch_tree(d, 0, s);
  mark("tree_volumes");
  launch_scale(d, s);
  mark("scale");
  launch_descreen(d, s);
  mark("descreen");
  launch_born(d, s);
  mark("born");
  launch_gbpair(d, s);
  mark("gbpair");
  launch_q2ab(d, s);
  mark("q2ab");
  launch_chain(d, s);
  mark("chain");
  launch_ws_place(d, s);
  mark("ws_place");
  launch_ws_free(d, s);
  mark("ws_free");
  launch_uv(d, s);
  mark("uv");
  launch_tree(d, 1, s);
  mark("tree_derivs");
  launch_final(d, sp, br, sa, dgbdr, dvwdr, decav, dehb, energies, s);
  mark("final");
  HIPCHK(hipGetLastError());
  return AGBNP_OK;
}

int check_status(Instance *I, int *status_out) {
  int st = 0;
  HIPCHK(hipMemcpyAsync(&st, I->d.status, sizeof(int), hipMemcpyDeviceToHost, I->stream));
  HIPCHK(hipStreamSynchronize(I->stream));
  *status_out = st;
  return AGBNP_OK;
}

}  // namespace

// ======================================================================== API
extern "C" {

int agbnp3_initialize(void) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (g_initialized) return AGBNP_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    errprint("agbnp3_initialize(): no HIP device available.\n");
    return AGBNP_ERR;
  }
  g_initialized = true;
  return AGBNP_OK;
}

void agbnp3_terminate(void) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_initialized) {
    errprint("agbnp3_terminate(): agbnp library is not initialized.\n");
    return;
  }
  for (auto &p : g_slots) {
    free_instance(p);
    p = nullptr;
  }
  g_slots.clear();
  g_initialized = false;
}

int agbnp3_new(int *tag, int natoms, float_i *x, float_i *y, float_i *z, float_i *r, float_i *charge,
               float_i dielectric_in, float_i dielectric_out, float_i *igamma, float_i *sgamma, float_i *ialpha,
               float_i *salpha, int *hbtype, float_i *hbcorr, int nhydrogen, int *ihydrogen, NeighList *conntbl,
               int verbose) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_initialized) {
    errprint("agbnp3_new(): agbnp library is not initialized.\n");
    return AGBNP_ERR;
  }
  if (!conntbl) {
    errprint("agbnp3_new(): error: NULL connection table. AGBNP v.3 requires a connection table.\n");
    return AGBNP_ERR;
  }
  if (natoms <= 0 || nhydrogen < 0 || nhydrogen > natoms || !x || !y || !z || !r || !charge || !igamma ||
      !sgamma || !ialpha || !salpha || !hbtype || !hbcorr || (nhydrogen > 0 && !ihydrogen)) {
    errprint("agbnp3_new(): invalid arguments.\n");
    return AGBNP_ERR;
  }
  Instance *I = new Instance();
  const int n = natoms;
  I->natoms = n;
  I->verbose = verbose;
  // heavy atoms first, hydrogens last, both in caller order (agbnp3.c:739-779)
  std::vector<char> ish(n, 0);
  for (int k = 0; k < nhydrogen; k++) {
    if (ihydrogen[k] < 0 || ihydrogen[k] >= n) {
      errprint("agbnp3_new(): hydrogen index %d out of range.\n", ihydrogen[k]);
      delete I;
      return AGBNP_ERR;
    }
    ish[ihydrogen[k]] = 1;
  }
  int nh_count = 0;
  for (int i = 0; i < n; i++) nh_count += ish[i];
  I->nheavy = n - nh_count;
  I->int2ext.resize(n);
  I->ext2int.resize(n);
  {
    int c = 0, h = I->nheavy;
    for (int i = 0; i < n; i++) {
      const int slot = ish[i] ? h++ : c++;
      I->int2ext[slot] = i;
      I->ext2int[i] = slot;
    }
  }
  auto f = [](double v) { return (float)v; };
  I->R.resize(n); I->q.resize(n); I->igamma.resize(n); I->sgamma.resize(n); I->ialpha.resize(n);
  I->salpha.resize(n); I->hbcorr.resize(n); I->hbtype.resize(n); I->vols.resize(n); I->galpha.resize(n);
  for (int i = 0; i < n; i++) {
    const int e = I->int2ext[i];
    I->R[i] = f((double)r[e] + (double)kRadInc);     // radius + 0.5 in double, stored FP32 (agbnp3.c:210)
    if (!(I->R[i] > kRadInc)) {
      errprint("agbnp3_new(): atom %d has a non-positive radius.\n", e);
      delete I;
      return AGBNP_ERR;
    }
    I->q[i] = f(charge[e]);
    I->igamma[i] = f(igamma[e]);
    I->sgamma[i] = f(sgamma[e]);
    I->ialpha[i] = f(ialpha[e]);
    I->salpha[i] = f(salpha[e]);
    I->hbcorr[i] = f(hbcorr[e]);
    I->hbtype[i] = hbtype[e];
    const double R = I->R[i];
    I->vols[i] = 4.0 * kPi / 3.0 * R * R * R;
    I->galpha[i] = kKFC / (I->R[i] * I->R[i]);
  }
  I->eps_in = f(dielectric_in);
  I->eps_out = f(dielectric_out);
  // connection table, remapped (agbnp3.c:296-317)
  if (conntbl->natoms < n) {
    errprint("agbnp3_new(): connection table has %d rows for %d atoms.\n", conntbl->natoms, n);
    delete I;
    return AGBNP_ERR;
  }
  I->cnne.resize(n);
  I->coff.assign(n + 1, 0);
  for (int i = 0; i < n; i++) {
    I->cnne[i] = conntbl->nne[I->int2ext[i]];
    I->coff[i + 1] = I->coff[i] + I->cnne[i];
  }
  I->cnl.resize(I->coff[n]);
  for (int i = 0; i < n; i++) {
    const int e = I->int2ext[i];
    for (int k = 0; k < I->cnne[i]; k++) {
      const int nbr = conntbl->neighl[e][k];
      if (nbr < 0 || nbr >= n) {
        errprint("agbnp3_new(): connection table entry %d of atom %d out of range.\n", nbr, e);
        delete I;
        return AGBNP_ERR;
      }
      I->cnl[I->coff[i] + k] = I->ext2int[nbr];
    }
  }
  // radius types by exact equality (agbnp3_utils.c:808-838) and one table per
  // ordered type pair (agbnp3_utils.c:840-883)
  I->rtype.resize(n);
  for (int i = 0; i < n; i++) {
    int found = -1;
    for (int t = 0; t < I->ntypes; t++)
      if (fabsf(I->R[i] - I->radii[t]) < FLT_MIN) { found = t; break; }
    if (found < 0) {
      found = I->ntypes++;
      I->radii.push_back(I->R[i]);
    }
    I->rtype[i] = found;
  }
  if (I->ntypes > 16) {
    errprint("agbnp3_new(): %d distinct radii; at most 16 are supported.\n", I->ntypes);
    delete I;
    return AGBNP_ERR;
  }
  const int nt = I->ntypes;
  I->ty.assign((size_t)nt * nt * kTblN, 0.f);
  I->ty2.assign((size_t)nt * nt * kTblN, 0.f);
  for (int a = 0; a < nt; a++)
    for (int b = 0; b < nt; b++) {
      const float c = kRadInc;
      const float ratio = (I->radii[a] - c) / I->radii[b];
      const size_t t = (size_t)(a * nt + b) * kTblN;
      build_table(ratio, &I->ty[t], &I->ty2[t], I->tdx, I->tdxinv);
    }
  build_water_sites(I);
  if (setup_device(I) != AGBNP_OK) {
    free_instance(I);
    return AGBNP_ERR;
  }
  // lowest free slot (agbnp3.c:144-170)
  int slot = 0;
  while (slot < (int)g_slots.size() && g_slots[slot]) slot++;
  if (slot == (int)g_slots.size()) g_slots.push_back(nullptr);
  g_slots[slot] = I;
  *tag = slot;
  return AGBNP_OK;
}

int agbnp3_delete(int tag) {
  std::lock_guard<std::mutex> lk(g_mutex);
  Instance *I = get_instance(tag, "agbnp3_delete()");
  if (!I) return AGBNP_ERR;
  free_instance(I);
  g_slots[tag] = nullptr;
  return AGBNP_OK;
}

int agbnp_initialize(void) { return agbnp3_initialize(); }
void agbnp_terminate(void) { agbnp3_terminate(); }
int agbnp_delete(int tag) { return agbnp3_delete(tag); }

int agbnp3_set_options(int tag, int options) {
  Instance *I = get_instance(tag, "agbnp3_set_options()");
  if (!I) return AGBNP_ERR;
  I->options = options;
  return AGBNP_OK;
}

int agbnp3_ener_device(int tag, const double *x, const double *y, const double *z, double *sp, double *br,
                       double *surf_area, double *dgbdr, double *dvwdr, double *decav, double *dehb,
                       double *energies) {
  Instance *I = get_instance(tag, "agbnp3_ener_device()");
  if (!I) return AGBNP_ERR;
  if (I->ws_error_atom >= 0) {
    errprint("agbnp3_create_ws_atoms(): error in agbnp3_create_ws_atoms_trigonal_oop(): atom %d needs 3 bonds.\n",
             I->int2ext[I->ws_error_atom]);
    errprint("agbnp3_ener(): error in agbnp3_total_energy().\n");
    return AGBNP_ERR;
  }
  return run_pipeline(I, x, y, z, sp, br, surf_area, dgbdr, dvwdr, decav, dehb, energies);
}

int agbnp3_sync(int tag) {
  Instance *I = get_instance(tag, "agbnp3_sync()");
  if (!I) return AGBNP_ERR;
  int st = 0;
  if (check_status(I, &st) != AGBNP_OK) return AGBNP_ERR;
  if (st) {
    errprint("agbnp3: device status 0x%x (capacity exceeded)\n", st);
    return AGBNP_ERR;
  }
  return AGBNP_OK;
}

void *agbnp3_stream(int tag) {
  Instance *I = get_instance(tag, "agbnp3_stream()");
  return I ? (void *)I->stream : nullptr;
}

int agbnp3_ener(int tag, int init, float_i *x, float_i *y, float_i *z, float_i *sp, float_i *br,
                float_i *mol_volume, float_i *surf_area, float_i *egb, float_i (*dgbdr)[3], float_i *evdw,
                float_i *ecorr_vdw, float_i (*dvwdr)[3], float_i *ecav, float_i *ecorr_cav, float_i (*decav)[3],
                float_i *ehb, float_i (*dehb)[3]) {
  (void)init;
  Instance *I = get_instance(tag, "agbnp3_agb_energy()");
  if (!I) return AGBNP_ERR;
  const size_t n = I->natoms;
  double *h = I->hstage;
  double *hin = h + 15 * n + 7;
  memcpy(hin, x, n * sizeof(double));
  memcpy(hin + n, y, n * sizeof(double));
  memcpy(hin + 2 * n, z, n * sizeof(double));
  hipStream_t s = I->stream;
  HIPCHK(hipMemcpyAsync(I->dx, hin, n * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(I->dy, hin + n, n * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(I->dz, hin + 2 * n, n * sizeof(double), hipMemcpyHostToDevice, s));
  double *o = I->dout;
  for (int attempt = 0; attempt < 2; attempt++) {
    if (agbnp3_ener_device(tag, I->dx, I->dy, I->dz, o, o + n, o + 2 * n, o + 3 * n, o + 6 * n, o + 9 * n,
                           o + 12 * n, o + 15 * n) != AGBNP_OK) {
      errprint("agbnp3_ener(): error in agbnp3_total_energy().\n");
      return AGBNP_ERR;
    }
    int st = 0;
    if (check_status(I, &st) != AGBNP_OK) return AGBNP_ERR;
    if (st == 0) break;
    if ((st & ST_NBLIST_OVERFLOW) && I->d.cap < 256 && attempt == 0) {
      I->d.cap = 256;   // wider rows, larger LDS variants; re-run the call
      continue;
    }
    errprint("agbnp3_ener(): device capacity exceeded (status 0x%x).\n", st);
    return AGBNP_ERR;
  }
  HIPCHK(hipMemcpyAsync(h, o, (15 * n + 7) * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  memcpy(sp, h, n * sizeof(double));
  memcpy(br, h + n, n * sizeof(double));
  memcpy(surf_area, h + 2 * n, n * sizeof(double));
  memcpy(dgbdr, h + 3 * n, 3 * n * sizeof(double));
  memcpy(dvwdr, h + 6 * n, 3 * n * sizeof(double));
  memcpy(decav, h + 9 * n, 3 * n * sizeof(double));
  memcpy(dehb, h + 12 * n, 3 * n * sizeof(double));
  const double *e = h + 15 * n;
  *egb = e[0];
  *evdw = e[1];
  *ecorr_vdw = e[2];
  *ecav = e[3];
  *ecorr_cav = e[4];
  *ehb = e[5];
  *mol_volume = e[6];
  return AGBNP_OK;
}

int agbnp3_q4_tables(int ntypes, const float *radii, float *y, float *y2, float *dx) {
  if (ntypes <= 0 || !radii) return AGBNP_ERR;
  float tdx = 0.f, tdxinv = 0.f;
  for (int a = 0; a < ntypes; a++)
    for (int b = 0; b < ntypes; b++) {
      const float c = kRadInc;
      const float ratio = (radii[a] - c) / radii[b];
      const size_t t = (size_t)(a * ntypes + b) * kTblN;
      std::vector<float> ty(kTblN), ty2(kTblN);
      build_table(ratio, ty.data(), ty2.data(), tdx, tdxinv);
      if (y) memcpy(y + t, ty.data(), sizeof(float) * kTblN);
      if (y2) memcpy(y2 + t, ty2.data(), sizeof(float) * kTblN);
    }
  if (dx) { dx[0] = tdx; dx[1] = tdxinv; }
  return AGBNP_OK;
}

// ------------------------------------------------------------ introspection
int agbnp3_natoms(int tag, int *natoms, int *nheavy) {
  Instance *I = get_instance(tag, "agbnp3_natoms()");
  if (!I) return AGBNP_ERR;
  if (natoms) *natoms = I->natoms;
  if (nheavy) *nheavy = I->nheavy;
  return AGBNP_OK;
}

int agbnp3_get_int2ext(int tag, int *int2ext) {
  Instance *I = get_instance(tag, "agbnp3_get_int2ext()");
  if (!I) return AGBNP_ERR;
  memcpy(int2ext, I->int2ext.data(), sizeof(int) * I->natoms);
  return AGBNP_OK;
}

int agbnp3_get_near_list(int tag, int *nne, int *nbr, int cap) {
  Instance *I = get_instance(tag, "agbnp3_get_near_list()");
  if (!I) return AGBNP_ERR;
  const int nh = I->nheavy, dc = I->d.cap;
  std::vector<int> hn(std::max(nh, 1)), hb((size_t)std::max(nh, 1) * dc);
  HIPCHK(hipStreamSynchronize(I->stream));
  HIPCHK(hipMemcpy(hn.data(), I->d.nne, sizeof(int) * nh, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(hb.data(), I->d.nbr, sizeof(int) * (size_t)nh * dc, hipMemcpyDeviceToHost));
  for (int i = 0; i < nh; i++) {
    if (nne) nne[i] = hn[i];
    if (nbr)
      for (int k = 0; k < hn[i] && k < cap; k++) nbr[(size_t)i * cap + k] = hb[(size_t)i * dc + k];
  }
  return AGBNP_OK;
}

int agbnp3_get_tables(int tag, float *y, float *y2, float *radii) {
  Instance *I = get_instance(tag, "agbnp3_get_tables()");
  if (!I) return AGBNP_ERR;
  if (y) memcpy(y, I->ty.data(), sizeof(float) * I->ty.size());
  if (y2) memcpy(y2, I->ty2.data(), sizeof(float) * I->ty2.size());
  if (radii) memcpy(radii, I->radii.data(), sizeof(float) * I->ntypes);
  return I->ntypes;
}

int agbnp3_get_internal(int tag, const char *name, double *out) {
  Instance *I = get_instance(tag, "agbnp3_get_internal()");
  if (!I) return AGBNP_ERR;
  const Dev &d = I->d;
  const double *src = nullptr;
  struct { const char *n; const double *p; } tab[] = {
      {"volumep_raw", d.volumep_raw}, {"sa_raw", d.sa_raw}, {"sa_f", d.sa_f}, {"sp", d.sp}, {"spe", d.spe},
      {"psvol", d.psvol}, {"gammap", d.gammap}, {"beta", d.beta}, {"br", d.br}, {"brw", d.brw}, {"swf", d.swf},
      {"A", d.A}, {"q2ab", d.q2ab}, {"abrw", d.abrw}, {"deru", d.deru}, {"derv", d.derv}, {"derus", d.derus},
      {"dervs", d.dervs}, {"derh", d.derh}};
  for (auto &e : tab)
    if (!strcmp(e.n, name)) src = e.p;
  if (!src) {
    errprint("agbnp3_get_internal(): unknown array '%s'\n", name);
    return AGBNP_ERR;
  }
  HIPCHK(hipStreamSynchronize(I->stream));
  HIPCHK(hipMemcpy(out, src, sizeof(double) * I->natoms, hipMemcpyDeviceToHost));
  return AGBNP_OK;
}

int agbnp3_get_counts(int tag, long long *counts) {
  Instance *I = get_instance(tag, "agbnp3_get_counts()");
  if (!I) return AGBNP_ERR;
  unsigned long long c[4];
  HIPCHK(hipStreamSynchronize(I->stream));
  HIPCHK(hipMemcpy(c, I->d.counters, sizeof(c), hipMemcpyDeviceToHost));
  counts[0] = (long long)c[0];
  counts[1] = (long long)c[1];
  counts[2] = I->nws;
  counts[3] = (long long)c[2];
  counts[4] = (long long)c[3];
  return AGBNP_OK;
}

int agbnp3_kernel_times(int tag, float *ms, int nmax, char *names, int names_len) {
  Instance *I = get_instance(tag, "agbnp3_kernel_times()");
  if (!I) return AGBNP_ERR;
  HIPCHK(hipStreamSynchronize(I->stream));
  const int nk = I->ev_names.empty() ? 0 : (int)I->ev_names.size() - 1;
  const int k = std::min(nmax, nk);
  std::vector<double> acc(std::max(nk, 1), 0.0);
  for (int c = 0; c < I->ncalls_timed; c++) {
    const auto &set = I->ev_sets[c];
    for (int i = 0; i < nk && i + 1 < (int)set.size(); i++) {
      float t = 0.f;
      (void)hipEventElapsedTime(&t, set[i], set[i + 1]);
      acc[i] += t;
    }
  }
  std::string all;
  for (int i = 0; i < k; i++) {
    ms[i] = I->ncalls_timed ? (float)(acc[i] / I->ncalls_timed) : 0.f;
    if (i) all += ";";
    all += I->ev_names[i + 1];
  }
  if (names && names_len > 0) {
    strncpy(names, all.c_str(), names_len - 1);
    names[names_len - 1] = 0;
  }
  I->ncalls_timed = 0;
  return k;
}

}  // extern "C"

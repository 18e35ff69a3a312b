// Shared definitions of the gfx950 AGBNP3 engine: model constants, the
// per-instance device view passed by value to every kernel, and the launcher
// prototypes implemented by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace agb {

// ---- model constants (reference agbnp3_private.h:57-137, agbnp3.c:1230-1236)
constexpr double kPi = 3.14159265359;   // the reference's pi
constexpr float kPiF = 3.14159265359f;
constexpr float kKFC = 2.5f;            // Gaussian exponent factor a = KFC / R^2
constexpr float kPFC = 2.97354f;        // Gaussian prefactor
constexpr float kRadInc = 0.5f;         // R = r + 0.5
constexpr float kNbOffset = 1.6f;       // neighbour test d < 1.6 (Ri + Rj)
constexpr float kVolMinA = 0.01f, kVolMinB = 0.1f;  // overlap-volume switch window
constexpr int kMaxOrder = 5;            // overlaps of order 2..5
constexpr double kTokcal = 332.0;
constexpr double kRwNp = 1.4;           // water radius in the vdW term
constexpr float kHbLength = 2.5f, kHbRadius = 1.4f;
constexpr float kHbSwA = 0.65f, kHbSwB = 1.0f, kHbSwA0 = 0.6f;
constexpr int kTblN = 512;              // spline nodes per table (y[511] = 0)
constexpr float kTblAmax = 20.0f;

// fixed-point scale of the order-independent (deterministic) scatter sums
constexpr double kFix = 4294967296.0;   // 2^32
constexpr double kFixInv = 1.0 / 4294967296.0;

// status word bits (device side capacity / topology checks)
enum : int {
  ST_NBLIST_OVERFLOW = 1,   // a neighbour-list row exceeded the row capacity
  ST_TREE_OVERFLOW = 2,     // an overlap group exceeded its LDS capacity
  ST_WS_OVERFLOW = 4,       // a water site had more candidates than its buffer
  ST_GRID_OVERFLOW = 8,     // cell grid larger than allocated
};

// energy partial-sum slots (per block, reduced in a fixed order)
enum : int {
  EP_CAV = 0, EP_CORR_CAV, EP_MOLVOL, EP_VDW, EP_CORR_VDW, EP_SELF, EP_PAIR, EP_HB, EP_COUNT
};

// per-overlap-tree accumulator layout (int64 fixed point per heavy atom)
enum : int { TG_GB = 0, TG_VW = 3, TG_CAV = 6, TG_HB = 9, TG_N = 12 };

// water-site record written by the placement kernel
struct WSite {
  float pos[3];
  float khb;
  int parent[4];
  int nparents;
  int type;
  int stale0;        // atom whose spe the reference reads first (quirk Q3)
  int pad;
  float dpos[4][3][3];   // d pos_i / d x_{parent p, j}
};

// static description of the water sites of one HB-active atom
struct WsDef {
  int atom;          // internal index of the acceptor / polar hydrogen
  int kind;          // placement geometry (WS_* below)
  int nsite;         // sites this atom creates
  int first;         // index of its first site
  int p[5];          // topology atoms used by the geometry
};
enum : int { WS_POLARH = 0, WS_TRIGONAL1, WS_TRIGONAL_S, WS_TRIGONAL_OOP, WS_TRIGONAL2,
             WS_TETRAHEDRAL2, WS_TETRAHEDRAL3, WS_TETRAHEDRAL1 };

struct GridParams {
  float lo[3];
  float cs, inv_cs;
  int dim[3];
  int ncell;
};

// Device view of one instance. Internal order: heavy atoms first.
struct Dev {
  int n, nh;             // atoms, heavy atoms
  int ntypes;
  int cap;               // neighbour-list row capacity
  int ncell_max;
  int nws, nwsdef;       // water sites, HB-active atoms
  int ws_cap;            // candidate capacity per water site
  int options;
  float eps_df;          // dielectric factor -0.5 (1/ein - 1/eout), FP32
  double df;             // same in FP64
  float rmax_heavy;      // largest heavy-atom R
  // static inputs
  const int *int2ext;
  const float *R;        // r + 0.5
  const float *q;
  const int *rtype;
  const float *galpha;   // KFC / R^2 (heavy)
  const float *igamma, *sgamma, *ialpha, *salpha, *hbcorr_f;
  const double *vols;    // 4/3 pi R^3
  const float2 *tbl;     // [ntypes*ntypes][512] (y, y2)
  float tdx, tdxinv;
  const int *row_perm;   // type-grouped row order, -1 padded
  const int *block_type; // row type of each 256-row block
  const int *block_hrows;  // 1 if every row of the block is a hydrogen
  int nrow_blocks;
  const WsDef *wsdef;
  const int *ws_stale0;  // per site: atom whose spe the first subtraction reads
  // per-call state
  float4 *pos;           // x, y, z, R
  float *bbox_part;      // [6][nblk]
  GridParams *grid;
  int *cell_count, *cell_start, *cell_atoms;
  int *nne, *nbr;
  int *status;
  long long *acc_vol, *acc_sa;         // tree pass A [nh]
  long long *acc_grad;                 // tree pass B [nh][TG_N]
  long long *acc_hb;                   // HB direct gradient [n][3]
  long long *acc_derh;                 // [nh]
  unsigned long long *counters;        // [0] kept overlaps [1] evaluated [2] nbl entries [3] max row
  double *volumep_raw, *sa_raw, *sa_f, *sp, *spe, *psvol, *gammap;   // [n]
  float *sp_f, *spe_f;                 // FP32 copies for pair kernels
  double *beta_part;                   // [nchunk_desc][n]
  int nchunk_desc;
  double *beta, *br, *brw, *swf;       // [n]
  float *br_f;
  long long *acc_gbA, *acc_gbg;        // GB pairs: A_i [n], gradient [n][3] (fixed point)
  double *gb_epart;                    // per tile pair: sum of q_i q_j / f
  const int2 *tile_ij;                 // tile pairs (I <= J) of 256-atom blocks
  int ntiles;
  double *A, *q2ab, *abrw;             // [n]
  float *q2ab_f, *abrw_f;
  double *ch_part;                     // [nchunk_ch][8][n]: gbx,gby,gbz,vwx,vwy,vwz,U,V
  int nchunk_ch;
  double *deru, *derv, *derus, *dervs, *derh;   // [n]
  float *deru_f, *derv_f, *derus_f, *dervs_f, *derh_f, *gammap_f;
  WSite *ws;
  double *ws_part;                     // per-site energy khb*s
  double *epart;                       // [EP_COUNT][epart_stride]
  int epart_stride;
};

// ---- launchers (kernel translation units) ------------------------------
void launch_ingest(const Dev &d, const double *x, const double *y, const double *z, hipStream_t s);
void launch_grid(const Dev &d, hipStream_t s);
void launch_cells(const Dev &d, hipStream_t s);
void launch_nblist(const Dev &d, hipStream_t s);
void launch_tree(const Dev &d, int pass, hipStream_t s);
void launch_scale(const Dev &d, hipStream_t s);
void launch_descreen(const Dev &d, hipStream_t s);
void launch_born(const Dev &d, hipStream_t s);
void launch_gbpair(const Dev &d, hipStream_t s);
void launch_q2ab(const Dev &d, hipStream_t s);
void launch_chain(const Dev &d, hipStream_t s);
void launch_uv(const Dev &d, hipStream_t s);
void launch_ws_place(const Dev &d, hipStream_t s);
void launch_ws_free(const Dev &d, hipStream_t s);
void launch_final(const Dev &d, double *sp, double *br, double *sa, double *dgbdr, double *dvwdr,
                  double *decav, double *dehb, double *energies, hipStream_t s);

// block counts used by the energy partial sums
int scale_blocks(const Dev &d);
int born_blocks(const Dev &d);

}  // namespace agb

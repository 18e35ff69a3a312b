// Per-call setup kernels: coordinate ingest, cell grid, heavy-atom neighbour
// list.
//
// The neighbour list must equal the reference's exactly (it fixes the shape
// of the overlap tree): heavy j > i with
//     d2 = (dx*dx + dy*dy) + dz*dz  <  ((R_i + R_j) * 1.6f)^2
// evaluated in FP32 without contraction, each row sorted by (d2, j)
// (reference agbnp3.c:4028-4066, agbnp3_fcompare agbnp3.c:3766-3784).
#include "agb_common.h"
#include "agb_device.h"

namespace agb {

// ---------------------------------------------------------------- ingest
// caller-order FP64 coordinates -> internal FP32 (x, y, z, R); clears the
// scatter accumulators of the call; per-block bounding box of heavy atoms
__global__ __launch_bounds__(256) void k_ingest(Dev d, const double *__restrict__ x,
                                                const double *__restrict__ y,
                                                const double *__restrict__ z) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (i < d.n) {
    const int e = d.int2ext[i];
    float4 p = make_float4((float)x[e], (float)y[e], (float)z[e], d.R[i]);
    d.pos[i] = p;
    if (i < d.nh) {
      lo[0] = hi[0] = p.x;
      lo[1] = hi[1] = p.y;
      lo[2] = hi[2] = p.z;
      d.acc_vol[i] = 0;
      d.acc_sa[i] = 0;
      d.acc_derh[i] = 0;
#pragma unroll
      for (int k = 0; k < TG_N; k++) d.acc_grad[(size_t)i * TG_N + k] = 0;
    }
    d.acc_gbA[i] = 0;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      d.acc_hb[3 * (size_t)i + q] = 0;
      d.acc_gbg[3 * (size_t)i + q] = 0;
    }
  }
  if (i == 0) {
    *d.status = 0;
    for (int k = 0; k < 4; k++) d.counters[k] = 0;
  }
  __shared__ float red[6][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float a = wave_min(lo[k]), b = wave_max(hi[k]);
    if (lane == 0) { red[k][w] = a; red[3 + k][w] = b; }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    float v = red[k][0];
    for (int q = 1; q < 4; q++) v = k < 3 ? fminf(v, red[k][q]) : fmaxf(v, red[k][q]);
    d.bbox_part[(size_t)k * gridDim.x + blockIdx.x] = v;
  }
}

void launch_ingest(const Dev &d, const double *x, const double *y, const double *z, hipStream_t s) {
  const int nblk = (d.n + 255) / 256;
  hipLaunchKernelGGL(k_ingest, dim3(nblk), dim3(256), 0, s, d, x, y, z);
}

// ------------------------------------------------------------ cell grid
// One workgroup: reduce the box, size the grid (cell edge >= largest
// heavy-heavy cutoff, so a 27-cell stencil is complete), counting sort of
// heavy atoms into cells.  kMaxCells bounds the LDS histogram; the edge grows
// until the grid fits.
constexpr int kMaxCells = 8192;

__global__ __launch_bounds__(1024) void k_cells(Dev d, int nblk) {
  __shared__ int cnt[kMaxCells];
  __shared__ int wsum[16], wpre[16];
  __shared__ GridParams g;
  __shared__ float box[6];
  const int t = threadIdx.x;
  if (t < 6) {
    float v = t < 3 ? INFINITY : -INFINITY;
    for (int b = 0; b < nblk; b++) {
      float u = d.bbox_part[(size_t)t * nblk + b];
      v = t < 3 ? fminf(v, u) : fmaxf(v, u);
    }
    box[t] = v;
  }
  __syncthreads();
  if (t == 0) {
    // every heavy-heavy and water-site-heavy cutoff is below 1.6 * 2 * Rmax
    float cs = fmaxf(kNbOffset * 2.0f * d.rmax_heavy, kNbOffset * (d.rmax_heavy + kHbRadius)) * 1.001f + 1e-3f;
    float ext[3];
    for (int k = 0; k < 3; k++) {
      float lo = box[k], hi = box[k + 3];
      if (!(hi >= lo)) { lo = 0.f; hi = 0.f; }    // no heavy atoms
      g.lo[k] = lo - 1e-3f;
      ext[k] = hi - lo + 2e-3f;
      if (!(ext[k] < 1e7f)) ext[k] = 1e7f;          // non-finite input: keep the loop finite
    }
    for (;;) {
      long long nc = 1;
      for (int k = 0; k < 3; k++) {
        g.dim[k] = (int)floorf(ext[k] / cs) + 1;
        nc *= g.dim[k];
      }
      if (nc <= kMaxCells && nc <= d.ncell_max) { g.ncell = (int)nc; break; }
      cs *= 1.05f;
    }
    g.cs = cs;
    g.inv_cs = 1.0f / cs;
    *d.grid = g;
  }
  __syncthreads();
  const int ncell = g.ncell;
  for (int c = t; c < ncell; c += 1024) cnt[c] = 0;
  __syncthreads();
  for (int i = t; i < d.nh; i += 1024) {
    float4 p = d.pos[i];
    int c = cell_index(g, p.x, p.y, p.z);
    d.cell_atoms[i] = c;            // temporary: cell of atom i
    atomicAdd(&cnt[c], 1);
  }
  __syncthreads();
  // exclusive scan of cnt over ncell cells: 8 cells per thread
  const int per = (ncell + 1023) / 1024;
  int local = 0;
  for (int k = 0; k < per; k++) {
    int c = t * per + k;
    if (c < ncell) local += cnt[c];
  }
  int incl = wave_incl_scan(local);
  const int lane = t & 63, w = t >> 6;
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (t < 16) {
    int s = 0;
    for (int q = 0; q < t; q++) s += wsum[q];
    wpre[t] = s;
  }
  __syncthreads();
  int run = incl - local + wpre[w];
  for (int k = 0; k < per; k++) {
    int c = t * per + k;
    if (c < ncell) {
      int v = cnt[c];
      d.cell_start[c] = run;
      cnt[c] = run;                  // cursor
      run += v;
    }
  }
  if (t == 0) d.cell_start[ncell] = d.nh;
  __syncthreads();
  // scatter; the order inside a cell is irrelevant (rows are re-sorted)
  for (int i = t; i < d.nh; i += 1024) {
    int c = d.cell_atoms[i];
    int slot = atomicAdd(&cnt[c], 1);
    d.cell_start[ncell + 1 + slot] = i;   // staged behind the starts
  }
  __syncthreads();
  for (int k = t; k < d.nh; k += 1024) d.cell_atoms[k] = d.cell_start[ncell + 1 + k];
}

void launch_cells(const Dev &d, hipStream_t s) {
  const int nblk = (d.n + 255) / 256;
  hipLaunchKernelGGL(k_cells, dim3(1), dim3(1024), 0, s, d, nblk);
}
void launch_grid(const Dev &, hipStream_t) {}

// ------------------------------------------------------- neighbour list
// one wavefront per heavy row; candidates from the 27-cell stencil are
// compacted into LDS with ballots, then rank-sorted by (d2, j)
constexpr int kNbWaves = 4;

template <int CAP>
__global__ __launch_bounds__(64 * kNbWaves) void k_nblist(Dev d) {
#pragma clang fp contract(off)
  __shared__ float sd2[kNbWaves][CAP];
  __shared__ int sj[kNbWaves][CAP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * kNbWaves + w;
  const bool active = i < d.nh;             // wave-uniform
  const GridParams g = *d.grid;
  const float4 pi = d.pos[active ? i : 0];
  int cx, cy, cz;
  cell_coords(g, pi.x, pi.y, pi.z, cx, cy, cz);
  int cnt = 0;
  for (int dz = -1; dz <= 1 && active; dz++) {
    const int zz = cz + dz;
    if (zz < 0 || zz >= g.dim[2]) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int yy = cy + dy;
      if (yy < 0 || yy >= g.dim[1]) continue;
      for (int dx = -1; dx <= 1; dx++) {
        const int xx = cx + dx;
        if (xx < 0 || xx >= g.dim[0]) continue;
        const int c = (zz * g.dim[1] + yy) * g.dim[0] + xx;
        const int beg = d.cell_start[c], end = d.cell_start[c + 1];
        for (int base = beg; base < end; base += 64) {
          const int k = base + lane;
          bool pass = false;
          int j = -1;
          float d2 = 0.f;
          if (k < end) {
            j = d.cell_atoms[k];
            if (j > i) {
              const float4 pj = d.pos[j];
              const float ex = __fsub_rn(pj.x, pi.x), ey = __fsub_rn(pj.y, pi.y), ez = __fsub_rn(pj.z, pi.z);
              d2 = __fadd_rn(__fadd_rn(__fmul_rn(ex, ex), __fmul_rn(ey, ey)), __fmul_rn(ez, ez));
              const float u = __fmul_rn(__fadd_rn(pi.w, pj.w), kNbOffset);
              pass = d2 < __fmul_rn(u, u);
            }
          }
          const unsigned long long m = __ballot(pass);
          if (pass) {
            const int slot = cnt + __popcll(m & ((1ull << lane) - 1ull));
            if (slot < CAP) { sd2[w][slot] = d2; sj[w][slot] = j; }
          }
          cnt += __popcll(m);
        }
      }
    }
  }
  int m = cnt;
  if (m > CAP) {
    if (lane == 0) atomicOr(d.status, ST_NBLIST_OVERFLOW);
    m = CAP;
  }
  __syncthreads();
  if (!active) return;
  for (int e = lane; e < m; e += 64) {
    const float de = sd2[w][e];
    const int je = sj[w][e];
    int rank = 0;
    for (int f = 0; f < m; f++) {
      const float df = sd2[w][f];
      rank += (df < de) || (df == de && sj[w][f] < je);
    }
    d.nbr[(size_t)i * d.cap + rank] = je;
  }
  if (lane == 0) {
    d.nne[i] = m;
    atomicAdd(&d.counters[2], (unsigned long long)m);
    atomicMax(&d.counters[3], (unsigned long long)m);
  }
}

void launch_nblist(const Dev &d, hipStream_t s) {
  const int nblk = (d.nh + kNbWaves - 1) / kNbWaves;
  if (nblk == 0) return;
  if (d.cap <= 128)
    hipLaunchKernelGGL(k_nblist<128>, dim3(nblk), dim3(64 * kNbWaves), 0, s, d);
  else
    hipLaunchKernelGGL(k_nblist<256>, dim3(nblk), dim3(64 * kNbWaves), 0, s, d);
}

}  // namespace agb

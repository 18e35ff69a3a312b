// Per-atom epilogues in FP64: volume scaling factors and surface filter,
// Born radii, chain-rule factors, U/V coefficients, final gradient assembly
// (caller order) and the fixed-order energy reduction.
#include "agb_common.h"
#include "agb_device.h"

namespace agb {

constexpr int kEpi = 256;

// block sum of up to 4 doubles into epart[slot][blockIdx.x]
__device__ void block_partials(const Dev &d, const int *slots, const double *v, int nv) {
  __shared__ double red[4][kEpi / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < nv; k++) {
    const double s = wave_sum(v[k]);
    if (lane == 0) red[k][w] = s;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    double s = 0.0;
    for (int q = 0; q < kEpi / 64; q++) s += red[threadIdx.x][q];
    d.epart[(size_t)slots[threadIdx.x] * d.epart_stride + blockIdx.x] = s;
  }
}

// ------------------------------------------------------- scaling factors
// agbnp3_scaling_factors (agbnp3_utils.c:524-613) with agbnp3_swf_area
// (agbnp3.c:1104-1117); E_cav, E_corr_cav, molecular volume
// (agbnp3.c:1390-1411)
__global__ __launch_bounds__(kEpi) void k_scale(Dev d) {
  const int i = blockIdx.x * kEpi + threadIdx.x;
  double ecav = 0.0, ecorr = 0.0, mv = 0.0;
  if (i < d.n) {
    if (i < d.nh) {
      const double R = d.R[i];
      const double vraw = d.vols[i] + fix_get(d.acc_vol[i]);
      const double a = 4.0 * kPi * R * R + fix_get(d.acc_sa[i]);
      double f = 0.0, fp = 0.0;
      if (a >= 0.0) {
        const double t = a / (64.0 + a * a);
        f = a * t;
        fp = 2.0 * t * (1.0 - f);
      }
      const double saf = a * f;
      const double rv = (R - kRadInc) / R;
      const double pr = R * (1.0 - rv * rv * rv) / 3.0;
      const double vp = vraw - saf * pr;
      const double gp = ((double)d.igamma[i] + (double)d.sgamma[i]) * (f + a * fp);
      d.volumep_raw[i] = vraw;
      d.sa_raw[i] = a;
      d.sa_f[i] = saf;
      d.spe[i] = vraw / d.vols[i];
      d.psvol[i] = (fp * a + f) * pr;
      d.sp[i] = vp / d.vols[i];
      d.gammap[i] = gp;
      d.sp_f[i] = (float)d.sp[i];
      d.spe_f[i] = (float)d.spe[i];
      d.gammap_f[i] = (float)gp;
      ecav = (double)d.igamma[i] * saf;
      ecorr = (double)d.sgamma[i] * saf;
      mv = vp;
    } else {
      d.volumep_raw[i] = d.sa_raw[i] = d.sa_f[i] = d.spe[i] = d.psvol[i] = d.sp[i] = d.gammap[i] = 0.0;
      d.sp_f[i] = d.spe_f[i] = d.gammap_f[i] = 0.f;
    }
  }
  const int slots[3] = {EP_CAV, EP_CORR_CAV, EP_MOLVOL};
  const double v[3] = {ecav, ecorr, mv};
  block_partials(d, slots, v, 3);
}

void launch_scale(const Dev &d, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(scale_blocks(d)), dim3(kEpi), 0, s, d);
}
int scale_blocks(const Dev &d) { return (d.n + kEpi - 1) / kEpi; }

// ------------------------------------------------------------ Born radii
// beta filter agbnp3_swf_invbr (agbnp3.c:1174-1188), agbnp3_born_radii
// (agbnp3_utils.c:615-670), E_vdW (agbnp3.c:1466-1487), GB self energy
__global__ __launch_bounds__(kEpi) void k_born(Dev d) {
  const int i = blockIdx.x * kEpi + threadIdx.x;
  double evdw = 0.0, ecorr = 0.0, eself = 0.0;
  if (i < d.n) {
    double s = 0.0;
    for (int c = 0; c < d.nchunk_desc; c++) s += d.beta_part[(size_t)c * d.n + i];
    const double beta = 1.0 / ((double)d.R[i] - kRadInc) - s / (4.0 * kPi);
    double bf, sw;
    if (beta < 0.0) {
      bf = 0.02;
      sw = 0.0;
    } else {
      bf = sqrt(0.02 * 0.02 + beta * beta);
      sw = beta / bf;
    }
    const double B = 1.0 / bf, t = B + kRwNp;
    d.beta[i] = beta;
    d.br[i] = B;
    d.br_f[i] = (float)B;
    d.swf[i] = sw;
    d.brw[i] = 3.0 * B * B / (t * t * t * t);
    const double a3 = 1.0 / (t * t * t);
    evdw = (double)d.ialpha[i] * a3;
    ecorr = (double)d.salpha[i] * a3;
    const double q = d.q[i];
    eself = d.df * q * q / B;
  }
  const int slots[3] = {EP_VDW, EP_CORR_VDW, EP_SELF};
  const double v[3] = {evdw, ecorr, eself};
  block_partials(d, slots, v, 3);
}

void launch_born(const Dev &d, hipStream_t s) {
  hipLaunchKernelGGL(k_born, dim3(born_blocks(d)), dim3(kEpi), 0, s, d);
}
int born_blocks(const Dev &d) { return (d.n + kEpi - 1) / kEpi; }

// -------------------------------------------------- chain-rule factors
// q2ab = (q^2 + A B) swf', abrw = alpha brw swf' (kernels.c:433-450)
__global__ __launch_bounds__(kEpi) void k_q2ab(Dev d) {
  const int i = blockIdx.x * kEpi + threadIdx.x;
  if (i < d.n) {
    const double A = fix_get(d.acc_gbA[i]);
    const double q = d.q[i];
    d.A[i] = A;
    d.q2ab[i] = (q * q + A * d.br[i]) * d.swf[i];
    d.abrw[i] = ((double)d.ialpha[i] + (double)d.salpha[i]) * d.brw[i] * d.swf[i];
    d.q2ab_f[i] = (float)d.q2ab[i];
    d.abrw_f[i] = (float)d.abrw[i];
  }
}

void launch_q2ab(const Dev &d, hipStream_t s) {
  hipLaunchKernelGGL(k_q2ab, dim3((d.n + kEpi - 1) / kEpi), dim3(kEpi), 0, s, d);
}

// ------------------------------------------------------ U, V coefficients
// deru = U/V_i, derus = df/(4pi) psvol deru, dervs = psvol derv/(4pi)
// (kernels.c:861-875); H_i from the water-site pass
__global__ __launch_bounds__(kEpi) void k_uv(Dev d) {
  const int i = blockIdx.x * kEpi + threadIdx.x;
  if (i >= d.n) return;
  double U = 0.0, V = 0.0;
  if (i < d.nh) {
    for (int c = 0; c < d.nchunk_ch; c++) {
      const double *p = d.ch_part + (size_t)c * 8 * d.n;
      U += p[6 * (size_t)d.n + i];
      V += p[7 * (size_t)d.n + i];
    }
  }
  const double vol = i < d.nh ? d.vols[i] : 1.0;
  const double du = i < d.nh ? U / vol : 0.0, dv = i < d.nh ? V / vol : 0.0;
  const double f4 = 1.0 / (4.0 * kPi);
  d.deru[i] = du;
  d.derv[i] = dv;
  d.derus[i] = d.df * f4 * d.psvol[i] * du;
  d.dervs[i] = f4 * d.psvol[i] * dv;
  d.derh[i] = i < d.nh ? fix_get(d.acc_derh[i]) : 0.0;
  d.deru_f[i] = (float)d.deru[i];
  d.derv_f[i] = (float)d.derv[i];
  d.derus_f[i] = (float)d.derus[i];
  d.dervs_f[i] = (float)d.dervs[i];
  d.derh_f[i] = (float)d.derh[i];
}

void launch_uv(const Dev &d, hipStream_t s) {
  hipLaunchKernelGGL(k_uv, dim3((d.n + kEpi - 1) / kEpi), dim3(kEpi), 0, s, d);
}

// ----------------------------------------------------------- final output
// gradients: GB = 332 (pair + chain + tree), vdW = chain + tree,
// cavity = tree, HB = direct + tree (agbnp3.c:1718-1786); caller order
__global__ __launch_bounds__(kEpi) void k_final(Dev d, double *sp, double *br, double *sa, double *dgbdr,
                                                double *dvwdr, double *decav, double *dehb) {
  const int i = blockIdx.x * kEpi + threadIdx.x;
  if (i >= d.n) return;
  const int e = d.int2ext[i];
  const size_t n = d.n;
  double g_gb[3] = {0, 0, 0}, g_vw[3] = {0, 0, 0};
  for (int q = 0; q < 3; q++) g_gb[q] = 2.0 * d.df * fix_get(d.acc_gbg[3 * (size_t)i + q]);
  const double f4 = 1.0 / (4.0 * kPi);
  for (int c = 0; c < d.nchunk_ch; c++) {
    const double *p = d.ch_part + (size_t)c * 8 * n;
    for (int q = 0; q < 3; q++) {
      g_gb[q] += d.df * f4 * p[q * n + i];
      g_vw[q] += f4 * p[(3 + q) * n + i];
    }
  }
  double g_cav[3] = {0, 0, 0}, g_hb[3];
  for (int q = 0; q < 3; q++) g_hb[q] = fix_get(d.acc_hb[3 * (size_t)i + q]);
  if (i < d.nh) {
    const long long *t = d.acc_grad + (size_t)i * TG_N;
    for (int q = 0; q < 3; q++) {
      g_gb[q] += fix_get(t[TG_GB + q]);
      g_vw[q] += fix_get(t[TG_VW + q]);
      g_cav[q] = fix_get(t[TG_CAV + q]);
      g_hb[q] += fix_get(t[TG_HB + q]);
    }
  }
  for (int q = 0; q < 3; q++) {
    dgbdr[3 * (size_t)e + q] = kTokcal * g_gb[q];
    dvwdr[3 * (size_t)e + q] = g_vw[q];
    decav[3 * (size_t)e + q] = g_cav[q];
    dehb[3 * (size_t)e + q] = g_hb[q];
  }
  sp[e] = i < d.nh ? d.sp[i] : 1.0;
  br[e] = d.br[i];
  sa[e] = i < d.nh ? d.sa_f[i] : 0.0;
}

// fixed-order reduction of the per-block energy partials and site energies
__global__ __launch_bounds__(256) void k_energy(Dev d, double *energies, int nscale, int nborn) {
  __shared__ double red[EP_COUNT][4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double acc[EP_COUNT];
  for (int k = 0; k < EP_COUNT; k++) acc[k] = 0.0;
  const int nb[EP_COUNT] = {nscale, nscale, nscale, nborn, nborn, nborn, 0, 0};
  for (int k = 0; k < EP_PAIR; k++) {
    // contiguous ranges per thread keep the order fixed
    const int per = (nb[k] + 255) / 256;
    for (int b = t * per; b < min(nb[k], (t + 1) * per); b++) acc[k] += d.epart[(size_t)k * d.epart_stride + b];
  }
  {
    const int per = (d.ntiles + 255) / 256;
    for (int b = t * per; b < min(d.ntiles, (t + 1) * per); b++) acc[EP_PAIR] += d.df * d.gb_epart[b];
  }
  {
    const int per = (d.nws + 255) / 256;
    for (int s = t * per; s < min(d.nws, (t + 1) * per); s++) acc[EP_HB] += d.ws_part[s];
  }
  for (int k = 0; k < EP_COUNT; k++) {
    const double s = wave_sum(acc[k]);
    if (lane == 0) red[k][w] = s;
  }
  __syncthreads();
  if (t == 0) {
    double e[EP_COUNT];
    for (int k = 0; k < EP_COUNT; k++) e[k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
    energies[0] = kTokcal * (e[EP_SELF] + e[EP_PAIR]);
    energies[1] = e[EP_VDW];
    energies[2] = e[EP_CORR_VDW];
    energies[3] = e[EP_CAV];
    energies[4] = e[EP_CORR_CAV];
    energies[5] = e[EP_HB];
    energies[6] = e[EP_MOLVOL];
  }
}

void launch_final(const Dev &d, double *sp, double *br, double *sa, double *dgbdr, double *dvwdr, double *decav,
                  double *dehb, double *energies, hipStream_t s) {
  hipLaunchKernelGGL(k_final, dim3((d.n + kEpi - 1) / kEpi), dim3(kEpi), 0, s, d, sp, br, sa, dgbdr, dvwdr, decav,
                     dehb);
  hipLaunchKernelGGL(k_energy, dim3(1), dim3(256), 0, s, d, energies, scale_blocks(d), born_blocks(d));
}

}  // namespace agb

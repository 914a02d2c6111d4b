// gp_kernels.hip -- CDNA4 (gfx950) kernels of the batched backward-Euler DAE integrator.
//
// Mapping.  Device, cluster and bus kernels put one scenario on each lane:
// a 64-lane wavefront evaluates one device/cluster/bus for 64 consecutive
// scenarios, every table/parameter read is wave-uniform (scalar loads) and
// every state read/write is one coalesced 512-byte access (arrays are
// [row][ldM]).  The sparse LU kernels put one scenario on each workgroup:
// the factorisation walks the elimination levels with workgroup barriers,
// staging each column in LDS, and the triangular solves keep the whole
// solution vector in LDS.
//
// Device equations restate the reference leaves (paths relative to the
// GradPower.jl repository): src/kernels/genrou.jl:48-141,407-559,
// ieesgo.jl:133-224, tgov1.jl:73-146, sexs.jl:74-156, esdc1a.jl:97-196,
// ieeest.jl:207-359, zipload.jl:33-179, static_gen.jl:107-220; network and
// faults src/dynamics.jl:265-327,823-894; backward Euler src/dynamics.jl:339-500.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "../../include/gradpower_hip.h"
#include "gp_internal.hpp"
#include "gp_splu.hpp"

namespace gp {

#define TWOPI60 (2.0 * 3.14159265358979323846 * 60.0)

__device__ __forceinline__ int lane_scenario() { return blockIdx.y * WAVE + threadIdx.x; }

// ---------------------------------------------------------------- GENROU
struct GenVals {
  // parameters
  double x_d, x_q, x_dp, x_qp, x_ddp, xl, H, D, T_d0p, T_q0p, T_d0dp, T_q0dp, S1, S2;
};

__device__ __forceinline__ void sat_coeff(double s1, double s2, double &sa, double &sb) {
  sa = 0.0; sb = 0.0;
  if (s1 <= 0.0 || s2 <= 0.0) return;
  double a = sqrt(s1 * 1.0 / (s2 * 1.2));
  if (a == 1.0) return;
  sa = 1.2 - (1.0 - 1.2) / (a - 1.0);
  sb = s2 * 1.2 * (a - 1.0) * (a - 1.0) / ((1.0 - 1.2) * (1.0 - 1.2));
}
__device__ __forceinline__ double sat_se(double psi, double sa, double sb) {
  if (sb == 0.0 || psi <= sa || psi == 0.0) return 0.0;
  return sb * (psi - sa) * (psi - sa) / psi;
}

// Residual of one GENROU (unscaled f for diff rows, g for alg rows) and its
// network current injection.
__device__ __forceinline__ void genrou_residual(const double *P, const double *zs, double e_fd, double p_m,
                                                double vr, double vi, double *fr, double &ir, double &ii) {
  double x_d = P[0], x_q = P[1], x_dp = P[2], x_qp = P[3], x_ddp = P[4], xl = P[5], H = P[6], Dm = P[7];
  double T_d0p = P[8], T_q0p = P[9], T_d0dp = P[10], T_q0dp = P[11], S1 = P[12], S2 = P[13], x_qdp = x_ddp;
  double e_qp = zs[0], e_dp = zs[1], phi_1d = zs[2], phi_2q = zs[3], w = zs[4], delta = zs[5];
  double v_q = zs[6], v_d = zs[7], i_q = zs[8], i_d = zs[9];
  double psi_de = (x_ddp - xl) / (x_dp - xl) * e_qp + (x_dp - x_ddp) / (x_dp - xl) * phi_1d;
  double psi_qe = -(x_ddp - xl) / (x_qp - xl) * e_dp + (x_qp - x_ddp) / (x_qp - xl) * phi_2q;
  double sa, sb;
  sat_coeff(S1, S2, sa, sb);
  double Se = sat_se(sqrt(psi_de * psi_de + psi_qe * psi_qe), sa, sb);
  fr[0] = (-e_qp + e_fd - (i_d - (-x_ddp + x_dp) * (-e_qp + i_d * (x_dp - xl) + phi_1d) / ((x_dp - xl) * (x_dp - xl))) * (x_d - x_dp) - Se * psi_de) / T_d0p;
  fr[1] = (-e_dp + (i_q - (-x_qdp + x_qp) * (e_dp + i_q * (x_qp - xl) + phi_2q) / ((x_qp - xl) * (x_qp - xl))) * (x_q - x_qp)) / T_q0p;
  fr[2] = (e_qp - i_d * (x_dp - xl) - phi_1d) / T_d0dp;
  fr[3] = (-e_dp - i_q * (x_qp - xl) - phi_2q) / T_q0dp;
  fr[4] = (p_m - Dm * w - psi_de * i_q + psi_qe * i_d) / (2.0 * H);
  fr[5] = TWOPI60 * w;
  fr[6] = i_d - ((x_ddp - xl) / (x_dp - xl) * e_qp + (x_dp - x_ddp) / (x_dp - xl) * phi_1d - v_q) / x_ddp;
  fr[7] = i_q - (-(x_qdp - xl) / (x_qp - xl) * e_dp + (x_qp - x_qdp) / (x_qp - xl) * phi_2q + v_d) / x_qdp;
  double sd, cd;
  sincos(delta, &sd, &cd);
  fr[8] = v_d - (vr * sd - vi * cd);
  fr[9] = v_q - (vr * cd + vi * sd);
  ir = sd * i_d + cd * i_q;
  ii = -cd * i_d + sd * i_q;
}

// Jacobian of one GENROU, unscaled.  jd[] = 47 slots in the order of the
// host slot list (see gp_host.cpp, genrou_slots).
__device__ __forceinline__ void genrou_jac(const double *P, const double *zs, double vr, double vi, double *j) {
  double x_d = P[0], x_q = P[1], x_dp = P[2], x_qp = P[3], x_ddp = P[4], xl = P[5], H = P[6], Dm = P[7];
  double T_d0p = P[8], T_q0p = P[9], T_d0dp = P[10], T_q0dp = P[11], S1 = P[12], S2 = P[13], x_qdp = x_ddp;
  double e_qp = zs[0], e_dp = zs[1], phi_1d = zs[2], phi_2q = zs[3], delta = zs[5];
  double i_q = zs[8], i_d = zs[9];
  double sd, cd;
  sincos(delta, &sd, &cd);
  double psi_de = (x_ddp - xl) / (x_dp - xl) * e_qp + (x_dp - x_ddp) / (x_dp - xl) * phi_1d;
  double psi_qe = -(x_ddp - xl) / (x_qp - xl) * e_dp + (x_qp - x_ddp) / (x_qp - xl) * phi_2q;
  double sa, sb;
  sat_coeff(S1, S2, sa, sb);
  double psi2 = sqrt(psi_de * psi_de + psi_qe * psi_qe);
  double Se = sat_se(psi2, sa, sb), dSe;
  if (sb == 0.0 || psi2 <= sa || psi2 == 0.0) dSe = 0.0;
  else { double g = psi2 - sa; dSe = sb * (2.0 * g * psi2 - g * g) / (psi2 * psi2); }
  double dpd = psi2 == 0.0 ? 0.0 : psi_de / psi2, dpq = psi2 == 0.0 ? 0.0 : psi_qe / psi2;
  double a_eq = (x_ddp - xl) / (x_dp - xl), a_p1 = (x_dp - x_ddp) / (x_dp - xl);
  double a_ed = -(x_ddp - xl) / (x_qp - xl), a_p2 = (x_qp - x_ddp) / (x_qp - xl);
  double dT_de = -(dSe * dpd * psi_de + Se), dT_qe = -(dSe * dpq * psi_de);
  double kd = (x_dp - xl) * (x_dp - xl), kq = (x_qp - xl) * (x_qp - xl);
  j[0] = (-(x_d - x_dp) * (-x_ddp + x_dp) / kd - 1) / T_d0p + dT_de * a_eq / T_d0p;   // R0 eq
  j[1] = (x_d - x_dp) * (-x_ddp + x_dp) / kd / T_d0p + dT_de * a_p1 / T_d0p;         // R0 p1
  j[2] = -(x_d - x_dp) * (-(-x_ddp + x_dp) / (x_dp - xl) + 1) / T_d0p;               // R0 id
  j[3] = dT_qe * a_ed / T_d0p;                                                        // R0 ed
  j[4] = dT_qe * a_p2 / T_d0p;                                                        // R0 p2
  j[5] = 1.0 / T_d0p;                                                                 // R0 efd
  j[6] = (-(x_q - x_qp) * (-x_qdp + x_qp) / kq - 1) / T_q0p;                          // R1 ed
  j[7] = -(x_q - x_qp) * (-x_qdp + x_qp) / kq / T_q0p;                                // R1 p2
  j[8] = (x_q - x_qp) * (-(-x_qdp + x_qp) / (x_qp - xl) + 1) / T_q0p;                 // R1 iq
  j[9] = 1.0 / T_d0dp;                                                                // R2 eq
  j[10] = -1.0 / T_d0dp;                                                              // R2 p1
  j[11] = (-x_dp + xl) / T_d0dp;                                                      // R2 id
  j[12] = -1.0 / T_q0dp;                                                              // R3 ed
  j[13] = -1.0 / T_q0dp;                                                              // R3 p2
  j[14] = (-x_qp + xl) / T_q0dp;                                                      // R3 iq
  j[15] = -0.5 * i_q * (x_ddp - xl) / (H * (x_dp - xl));                              // R4 eq
  j[16] = 0.5 * i_d * (-x_ddp + xl) / (H * (x_qp - xl));                              // R4 ed
  j[17] = -0.5 * i_q * (-x_ddp + x_dp) / (H * (x_dp - xl));                           // R4 p1
  j[18] = 0.5 * i_d * (-x_ddp + x_qp) / (H * (x_qp - xl));                            // R4 p2
  j[19] = -0.5 * Dm / H;                                                              // R4 w
  j[20] = 0.5 * (-e_qp * (x_ddp - xl) / (x_dp - xl) - phi_1d * (-x_ddp + x_dp) / (x_dp - xl)) / H;  // R4 iq
  j[21] = 0.5 * (e_dp * (-x_ddp + xl) / (x_qp - xl) + phi_2q * (-x_ddp + x_qp) / (x_qp - xl)) / H;  // R4 id
  j[22] = 0.5 / H;                                                                    // R4 pm
  j[23] = TWOPI60;                                                                    // R5 w
  j[24] = 0.0;                                                                        // R5 delta
  j[25] = -(x_ddp - xl) / (x_ddp * (x_dp - xl));                                      // A1 eq
  j[26] = -(-x_ddp + x_dp) / (x_ddp * (x_dp - xl));                                   // A1 p1
  j[27] = 1.0 / x_ddp;                                                                // A1 vq
  j[28] = 1.0;                                                                        // A1 id
  j[29] = -(-x_qdp + xl) / (x_qdp * (x_qp - xl));                                     // A2 ed
  j[30] = -(-x_qdp + x_qp) / (x_qdp * (x_qp - xl));                                   // A2 p2
  j[31] = -1.0 / x_qdp;                                                               // A2 vd
  j[32] = 1.0;                                                                        // A2 iq
  j[33] = -vr * cd - vi * sd;                                                         // A3 delta
  j[34] = 1.0;                                                                        // A3 vd
  j[35] = -sd;                                                                        // A3 vr
  j[36] = cd;                                                                         // A3 vi
  j[37] = vr * sd - vi * cd;                                                          // A4 delta
  j[38] = 1.0;                                                                        // A4 vq
  j[39] = -cd;                                                                        // A4 vr
  j[40] = -sd;                                                                        // A4 vi
  j[41] = i_d * cd - i_q * sd;                                                        // NR delta
  j[42] = cd;                                                                         // NR iq
  j[43] = sd;                                                                         // NR id
  j[44] = i_d * sd + i_q * cd;                                                        // NI delta
  j[45] = sd;                                                                         // NI iq
  j[46] = -cd;                                                                        // NI id
}
// diff-row flags of the 47 GENROU slots (rows 0..5 are differential)
__constant__ unsigned char GEN_SLOT_DIFF[47] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                               0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
__constant__ unsigned char GEN_SLOT_DIAG[47] = {1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1,
                                               0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

// ---------------------------------------------------------------- controllers
__device__ __forceinline__ void ieesgo_residual(const double *P, const double *x, double p_m, double w, double *fr) {
  double T1 = P[0], T2 = P[1], T3 = P[2], T4 = P[3], T5 = P[4], T6 = P[5], K1 = P[6], K2 = P[7], K3 = P[8], pref = P[11];
  double PF0 = x[0], PLL = x[1], TP1 = x[2], TP2 = x[3], TP3 = x[4];
  fr[0] = (1.0 / T1) * (K1 * w - PF0);
  fr[1] = (1.0 / T3) * ((1.0 - (T2 / T3)) * PF0 - PLL);
  double SatP = pref - (T2 / T3) * PF0 - PLL;
  fr[2] = (1.0 / T4) * (SatP - TP1);
  fr[3] = (1.0 / T5) * (K2 * TP1 - TP2);
  fr[4] = (1.0 / T6) * (K3 * TP2 - TP3);
  fr[5] = TP1 * (1.0 - K2) + TP2 * (1.0 - K3) + TP3 - p_m;
}
__device__ __forceinline__ void tgov1_residual(const double *P, const double *x, double p_m, double w, double *fr) {
  double R = P[0], T1 = P[1], T2 = P[4], T3 = P[5], DT = P[6], pref = P[7], t23 = T2 / T3;
  fr[0] = (-x[0] + (1.0 - t23) * x[1]) / T3;
  fr[1] = ((pref - w) / R - x[1]) / T1;
  fr[2] = x[0] + t23 * x[1] - DT * w - p_m;
}
__device__ __forceinline__ void sexs_residual(const double *P, const double *x, double vr, double vi, double vs, double *fr) {
  double TA_TB = P[0], TB = P[1], K = P[2], TE = P[3], vref = P[6];
  double vm = sqrt(vr * vr + vi * vi);
  if (vm == 0.0) vm = 1e-12;
  double err = vref - vm + vs;
  fr[0] = (-x[0] + (1.0 - TA_TB) * err) / TB;
  fr[1] = (-x[1] + K * (x[0] + TA_TB * err)) / TE;
}
__device__ __forceinline__ void esdc1a_residual(const double *P, const double *x, double vr, double vi, double vs, double *fr) {
  double Ka = P[0], Ta = P[1], Kf = P[2], Tf = P[3], Ke = P[4], Te = P[5], sa = P[7], sb = P[8], vref = P[9];
  double vm = sqrt(vr * vr + vi * vi);
  if (vm == 0.0) vm = 1e-12;
  double e_fd = x[2];
  double sat = (sb == 0.0 || e_fd <= sa) ? 0.0 : sb * (e_fd - sa) * (e_fd - sa);
  fr[0] = (Ka * (vref - vm + vs - x[1] - (Kf / Tf) * e_fd) - x[0]) / Ta;
  fr[1] = -((Kf / Tf) * e_fd + x[1]) / Tf;
  fr[2] = (x[0] - Ke * e_fd - sat) / Te;
}
__device__ __forceinline__ void ieeest_residual(const double *P, const double *x, double sig, double *fr) {
  double A1 = P[0], A2 = P[1], A3 = P[2], A4 = P[3], A5 = P[4], A6 = P[5];
  double T1 = P[6], T2 = P[7], T3 = P[8], T4 = P[9], T5 = P[10], T6 = P[11], KS = P[12];
  double s0 = x[0], s1 = x[1], s2 = x[2], s3 = x[3], s4 = x[4], s5 = x[5], s6 = x[6], vs = x[7];
  double qq = s1 - s3 - A3 * s2;
  double y2 = s3 + A5 * s2 + (A6 / A4) * qq;
  double y3 = s4 + (T1 / T2) * (y2 - s4);
  double y4 = s5 + (T3 / T4) * (y3 - s5);
  fr[0] = (sig - s1 - A1 * s0) / A2;
  fr[1] = s0;
  fr[2] = (s1 - s3 - A3 * s2) / A4;
  fr[3] = s2;
  fr[4] = (y2 - s4) / T2;
  fr[5] = (y3 - s5) / T4;
  fr[6] = (KS * y4 - s6) / T6;
  fr[7] = vs - (T5 / T6) * (KS * y4 - s6);
}

// Exciter `row` of type et (GP_SEXS / GP_ESDC1A) taken offline in scenario m.
// An offline exciter's rows of F are zero (the reference's kernels skip a
// device whose online flag is false into a zeroed F, ext:1340,1072-1199),
// so backward Euler holds its states: F = z - zold, J rows = identity.
__device__ __forceinline__ bool exc_offline(const FaultState &fs, int ldM, int m, int et, int row) {
  if (!fs.nexc) return false;
  const int code = kExcCode0 - (2 * row + (et == GP_ESDC1A ? 1 : 0));
  bool off = false;
#pragma unroll
  for (int s = 0; s < MAXFAULT; s++) off = off || fs.bus[s * ldM + m] == code;
  return off;
}

// ---------------------------------------------------------------- residual kernels
// One wave = one cluster x 64 scenarios.  Writes the backward-Euler residual
// of every cluster row and the generator's current injection.
__global__ void __launch_bounds__(64) k_cluster_residual(DevTables T, const double *__restrict__ p,
                                                         const double *__restrict__ z, const double *__restrict__ zold,
                                                         double dt0, const double *__restrict__ dtv, FaultState fs,
                                                         double *__restrict__ f, double *__restrict__ inj,
                                                         const uint8_t *__restrict__ work, double *__restrict__ nrm) {
  const int c = T.cl_nt[blockIdx.x];
  const int m = lane_scenario();
  if (m >= T.M || !work[m]) return;
  const double dt = dtv ? dtv[m] : dt0;   // per-scenario step (0: event re-solve)
  const int ldM = T.ldM;
  // |F| of the rows written here, as the bit patterns of non-negative
  // doubles (whose order is the integers'; a NaN is the largest)
  unsigned long long bits = 0;
  auto put = [&](size_t o, double v) {
    f[o] = v;
    const unsigned long long u = __double_as_longlong(fabs(v));
    bits = u > bits ? u : bits;
  };
  const int *cl = T.cl + c * GP_CLUSTER_COLS;
  const int ws = cl[GP_CL_WSTART], gk = cl[GP_CL_GENROU];
  const int *g = T.genrou + gk * GP_COLS_GENROU;
  const int dp = g[0], ap = g[1], bus = g[3], pm = g[4], efd = g[5];
  const double *Pg = p + g[2];
  const int vrz = T.net_ptr + 2 * bus;
  const double vr = z[(size_t)vrz * ldM + m], vi = z[(size_t)(vrz + 1) * ldM + m];
  double zs[10], fr[10];
#pragma unroll
  for (int i = 0; i < 6; i++) zs[i] = z[(size_t)(dp + i) * ldM + m];
#pragma unroll
  for (int i = 0; i < 4; i++) zs[6 + i] = z[(size_t)(ap + i) * ldM + m];
  const double e_fd = efd >= 0 ? z[(size_t)efd * ldM + m] : T.u0[2 * gk];
  const double p_m = pm >= 0 ? z[(size_t)pm * ldM + m] : T.u0[2 * gk + 1];
  double ir, ii;
  genrou_residual(Pg, zs, e_fd, p_m, vr, vi, fr, ir, ii);
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const size_t o = (size_t)(dp + i) * ldM + m;
    put(o, zs[i] - zold[o] - dt * fr[i]);
  }
#pragma unroll
  for (int i = 0; i < 4; i++) put((size_t)(ap + i) * ldM + m, fr[6 + i]);
  inj[(size_t)(2 * c) * ldM + m] = ir;
  inj[(size_t)(2 * c + 1) * ldM + m] = ii;
  const double w = zs[4];
  const int gt = cl[GP_CL_GOVTYPE];
  if (gt == GP_IEESGO) {
    const int *d = T.ieesgo + cl[GP_CL_GOV] * GP_COLS_IEESGO;
    double x[5], r[6];
#pragma unroll
    for (int i = 0; i < 5; i++) x[i] = z[(size_t)(d[0] + i) * ldM + m];
    const double pmv = z[(size_t)d[1] * ldM + m];
    const double wv = d[3] >= 0 ? z[(size_t)d[3] * ldM + m] : 0.0;
    ieesgo_residual(p + d[2], x, pmv, wv, r);
#pragma unroll
    for (int i = 0; i < 5; i++) {
      const size_t o = (size_t)(d[0] + i) * ldM + m;
      put(o, x[i] - zold[o] - dt * r[i]);
    }
    put((size_t)d[1] * ldM + m, r[5]);
  } else if (gt == GP_TGOV1) {
    const int *d = T.tgov1 + cl[GP_CL_GOV] * GP_COLS_TGOV1;
    double x[2], r[3];
    x[0] = z[(size_t)d[0] * ldM + m];
    x[1] = z[(size_t)(d[0] + 1) * ldM + m];
    const double pmv = z[(size_t)d[1] * ldM + m];
    const double wv = d[3] >= 0 ? z[(size_t)d[3] * ldM + m] : 0.0;
    tgov1_residual(p + d[2], x, pmv, wv, r);
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const size_t o = (size_t)(d[0] + i) * ldM + m;
      put(o, x[i] - zold[o] - dt * r[i]);
    }
    put((size_t)d[1] * ldM + m, r[2]);
  }
  const int et = cl[GP_CL_EXCTYPE];
  if (et == GP_SEXS || et == GP_ESDC1A) {
    const int *d = (et == GP_SEXS ? T.sexs + cl[GP_CL_EXC] * GP_COLS_SEXS : T.esdc1a + cl[GP_CL_EXC] * GP_COLS_ESDC1A);
    const int ne = et == GP_SEXS ? 2 : 3;
    const double evr = z[(size_t)d[2] * ldM + m], evi = z[(size_t)(d[2] + 1) * ldM + m];
    const double vs = d[3] >= 0 ? z[(size_t)d[3] * ldM + m] : 0.0;
    double x[3], r[3];
    for (int i = 0; i < ne; i++) x[i] = z[(size_t)(d[0] + i) * ldM + m];
    if (et == GP_SEXS) sexs_residual(p + d[1], x, evr, evi, vs, r);
    else esdc1a_residual(p + d[1], x, evr, evi, vs, r);
    if (exc_offline(fs, ldM, m, et, cl[GP_CL_EXC])) r[0] = r[1] = r[2] = 0.0;
    for (int i = 0; i < ne; i++) {
      const size_t o = (size_t)(d[0] + i) * ldM + m;
      put(o, x[i] - zold[o] - dt * r[i]);
    }
  }
  if (cl[GP_CL_PSS] >= 0) {
    const int *d = T.ieeest + cl[GP_CL_PSS] * GP_COLS_IEEEST;
    double x[8], r[8];
#pragma unroll
    for (int i = 0; i < 7; i++) x[i] = z[(size_t)(d[0] + i) * ldM + m];
    x[7] = z[(size_t)d[1] * ldM + m];
    const double sig = d[3] >= 0 ? z[(size_t)d[3] * ldM + m] : 0.0;
    ieeest_residual(p + d[2], x, sig, r);
#pragma unroll
    for (int i = 0; i < 7; i++) {
      const size_t o = (size_t)(d[0] + i) * ldM + m;
      put(o, x[i] - zold[o] - dt * r[i]);
    }
    put((size_t)d[1] * ldM + m, r[7]);
  }
  (void)w;
  (void)ws;
  if (nrm) nrm[(size_t)blockIdx.x * ldM + m] = __longlong_as_double(bits);
}

// ZIP load k disconnected in scenario m (DisconnectDeviceEvent,
// src/GradPower.jl:130-143,681-698: its residual and Jacobian leave the bus).
__device__ __forceinline__ bool zip_dropped(const FaultState &fs, int ldM, int m, int k) {
  bool off = false;
#pragma unroll
  for (int s = 0; s < MAXFAULT; s++) off = off || fs.bus[s * ldM + m] == kDropCode0 - k;
  return off;
}

// Line trips active in scenario m (event slots holding -2 - t): a tripped
// line's stamp leaves Y.  The network rows hold -Y v, so the rows of the
// line's buses gain +y v: at the from bus yff v_from + yft v_to, at the to
// bus ytf v_from + ytt v_to (src/GradPower.jl:556-568 takes the stamp out of
// ybus_real instead).
__device__ __forceinline__ void trip_residual(const DevTables &T, const FaultState &fs, const double *__restrict__ z,
                                              int b, int m, double vr, double vi, double &fr, double &fi) {
  const int ldM = T.ldM;
#pragma unroll
  for (int k = 0; k < MAXFAULT; k++) {
    const int code = fs.bus[k * ldM + m];
    if (code > -2 || code <= kDropCode0) continue;
    const int t = -2 - code;
    const int a = T.tr_bus[2 * t], c = T.tr_bus[2 * t + 1];
    if (a != b && c != b) continue;
    const double *y = T.tr_y + 8 * t;
    const double *ys = y + (a == b ? 0 : 6);   // own bus: yff or ytt
    const double *yo = y + (a == b ? 2 : 4);   // other bus: yft or ytf
    const int oz = T.net_ptr + 2 * (a == b ? c : a);
    const double ur = z[(size_t)oz * ldM + m], ui = z[(size_t)(oz + 1) * ldM + m];
    fr += ys[0] * vr - ys[1] * vi + yo[0] * ur - yo[1] * ui;
    fi += ys[1] * vr + ys[0] * vi + yo[1] * ur + yo[0] * ui;
  }
}

// Entry (row bus b, column bus col, component comp = 2 rr + cc of the 2x2
// block) of -Y with the lines tripped in scenario m taken out: -Y + y_line,
// y_line realified as [g -b; b g].
__device__ __forceinline__ double trip_offdiag(const DevTables &T, const FaultState &fs, int b, int m, int col,
                                               int comp, double base) {
  const int ldM = T.ldM;
  double v = base;
#pragma unroll
  for (int k = 0; k < MAXFAULT; k++) {
    const int code = fs.bus[k * ldM + m];
    if (code > -2 || code <= kDropCode0) continue;
    const int t = -2 - code;
    const int a = T.tr_bus[2 * t], c = T.tr_bus[2 * t + 1];
    const double *y = T.tr_y + 8 * t;
    const double *yy = (a == b && c == col) ? y + 2 : ((c == b && a == col) ? y + 4 : nullptr);
    if (!yy) continue;
    v += (comp == 0 || comp == 3) ? yy[0] : (comp == 1 ? -yy[1] : yy[1]);
  }
  return v;
}

// One wave = one bus x 64 scenarios: network rows (-Y v + injections, faults)
// and the static generator's own rows.  Summation order per bus follows the
// reference: -Y v, machines, loads, static generator, faults.
__global__ void __launch_bounds__(64) k_bus_residual(DevTables T, const double *__restrict__ p,
                                                     const double *__restrict__ z, const double *__restrict__ inj,
                                                     FaultState fs, double *__restrict__ f,
                                                     const uint8_t *__restrict__ work, double *__restrict__ pm) {
  const int b = blockIdx.x;
  const int m = lane_scenario();
  if (m >= T.M || !work[m]) return;
  const int ldM = T.ldM;
  const int vrz = T.net_ptr + 2 * b;
  const double vr = z[(size_t)vrz * ldM + m], vi = z[(size_t)(vrz + 1) * ldM + m];
  double fr = 0.0, fi = 0.0;
  for (int q = T.yrp[2 * b]; q < T.yrp[2 * b + 1]; q++) fr += T.yv[q] * z[(size_t)(T.net_ptr + T.yci[q]) * ldM + m];
  for (int q = T.yrp[2 * b + 1]; q < T.yrp[2 * b + 2]; q++) fi += T.yv[q] * z[(size_t)(T.net_ptr + T.yci[q]) * ldM + m];
  fr = -fr;
  fi = -fi;
  for (int q = T.bus_cl_p[b]; q < T.bus_cl_p[b + 1]; q++) {
    const int c = T.bus_cl[q];
    fr += inj[(size_t)(2 * c) * ldM + m];
    fi += inj[(size_t)(2 * c + 1) * ldM + m];
  }
  const double vm2r = vr * vr + vi * vi;
  for (int q = T.bus_zip_p[b]; q < T.bus_zip_p[b + 1]; q++) {
    if (zip_dropped(fs, ldM, m, T.bus_zip[q])) continue;
    const double *P = p + T.zip[T.bus_zip[q] * GP_COLS_ZIPLOAD + 1];
    const double pl = P[0], ql = P[1], al = P[2], yre = al * P[7], yim = al * P[8];
    const double den = vm2r > 0.2 ? vm2r : 0.2;
    fr -= vr * yre - vi * yim;
    fi -= vr * yim + vi * yre;
    fr -= (1.0 - al) * (pl * vr - ql * vi) / den;
    fi -= (1.0 - al) * (ql * vr + pl * vi) / den;
  }
  const int sgk = T.bus_sg[b];
  double sg0 = 0.0, sg1 = 0.0;   // the static generator's rows (zero without)
  if (sgk >= 0) {
    const int *d = T.sg + sgk * GP_COLS_STATICGEN;
    const int ap = d[1], bt = d[3];
    const double *P = p + d[2];
    const double vm2 = vm2r > 0.2 ? vm2r : 0.2;
    double pv, qv;
    if (bt == 2) { pv = P[0]; qv = z[(size_t)ap * ldM + m]; }
    else if (bt == 3) { pv = z[(size_t)ap * ldM + m]; qv = z[(size_t)(ap + 1) * ldM + m]; }
    else { pv = P[0]; qv = P[1]; }
    fr += (pv * vr + qv * vi) / vm2;
    fi += (pv * vi - qv * vr) / vm2;
    if (bt == 2) {
      sg0 = vr * vr + vi * vi - P[2] * P[2];
      f[(size_t)ap * ldM + m] = sg0;
    } else if (bt == 3) {
      sg0 = vr - P[2] * cos(P[3]);
      sg1 = vi - P[2] * sin(P[3]);
      f[(size_t)ap * ldM + m] = sg0;
      f[(size_t)(ap + 1) * ldM + m] = sg1;
    }
  }
#pragma unroll
  for (int k = 0; k < MAXFAULT; k++) {
    if (fs.bus[k * ldM + m] == b) {
      const double gf = fs.g[k * ldM + m];
      fr -= gf * vr;
      fi -= gf * vi;
    }
  }
  trip_residual(T, fs, z, b, m, vr, vi, fr, fi);
  f[(size_t)vrz * ldM + m] = fr;
  f[(size_t)(vrz + 1) * ldM + m] = fi;
  if (pm) {   // |F| of the rows written here (bit patterns, as k_cluster_residual)
    unsigned long long bits = 0;
    const double r4[4] = {fr, fi, sg0, sg1};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const unsigned long long u = __double_as_longlong(fabs(r4[k]));
      bits = u > bits ? u : bits;
    }
    pm[(size_t)(T.n_cl_nt + b) * ldM + m] = __longlong_as_double(bits);
  }
}

// Max-norm of F per scenario.  Grid (ldM/64, chunks); partial maxima are
// combined with an integer atomicMax on the bit pattern (|x| >= 0, so the
// integer order is the float order; NaN patterns sort above +inf).
// |F|inf per scenario: block (x: 64 scenarios, y: a chunk of rows), its 4
// wavefronts stride over the chunk's rows; the per-lane maxima meet in LDS
// and one atomicMax per scenario and block publishes them.  |F| >= 0, so the
// IEEE bit patterns order like the values.
__global__ void __launch_bounds__(256) k_rownorm(int n, int ldM, int M, const double *__restrict__ f,
                                                 const uint8_t *__restrict__ work,
                                                 unsigned long long *__restrict__ nb) {
  __shared__ unsigned long long red[4][WAVE];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
  const int m = blockIdx.x * WAVE + lane;
  const bool on = m < M && work[m];
  const int chunk = (n + gridDim.y - 1) / gridDim.y;
  const int r0 = blockIdx.y * chunk, r1 = min(n, r0 + chunk);
  unsigned long long bits = 0;
  if (on)
    for (int i = r0 + w; i < r1; i += 4) {
      const unsigned long long u = __double_as_longlong(fabs(f[(size_t)i * ldM + m]));
      bits = u > bits ? u : bits;
    }
  red[w][lane] = bits;
  __syncthreads();
  if (w == 0 && on) {
    unsigned long long v = red[0][lane];
    for (int q = 1; q < 4; q++) v = red[q][lane] > v ? red[q][lane] : v;
    if (v) atomicMax(&nb[m], v);
  }
}

// Finishes the max-norm and decides, per working scenario, what this round
// does.  Converged scenarios (|F| < tol) leave `work`.  A scenario that has
// used its itermax corrections, or (stall > 0) whose |F| is not finite,
// exceeds 1e10 or has not halved within `stall` of its corrections, leaves
// `work` with failed[m] = 1 (3 when |F| is not finite).  refac[m] asks for a numeric factorisation: at
// the first iteration refac0[m] (or 1 without refac0); later when every
// iteration refactors (theta <= 0) or |F| fell by less than theta since the
// previous iteration, except while one is pending (pend[m]).  With
// async_below > 0 a later request from a scenario whose |F| is below it is
// asynchronous (areq[m] = 1, refac[m] = 0).  upd[m] = 1 when the scenario
// applies a correction this round: a scenario whose factorisation is pending,
// or that asks for an asynchronous one, waits for it instead, so that its
// iterates are those of the synchronous policy.  sch[m] = 1 where the Schur
// kernels run (corrections and new asynchronous requests).  iters[m] counts
// corrections; first (per scenario with firstv) resets the per-solve history.
// exact[m] != 0 (optional) runs scenario m's solve as exact Newton: a
// factorisation at every iteration, whatever theta says (asynchronous ones
// are waited for, as above).  switch_window > 0: a scenario not in exact
// mode whose |F| has not halved within that many corrections stops as failed
// too (the host then runs its solve again as exact Newton).
// n_unconv: [0] working scenarios, [1] synchronous requests, [2]
// corrections, [3] asynchronous requests.
__global__ void k_norm_finish(int M, double tol, unsigned long long *__restrict__ nb, double *__restrict__ fnorm,
                              uint8_t *__restrict__ work, int *__restrict__ n_unconv, uint8_t *__restrict__ refac,
                              const uint8_t *__restrict__ refac0, int first0, const uint8_t *__restrict__ firstv,
                              double theta, int stall, int itermax,
                              double *__restrict__ fbest, int *__restrict__ nstall, int *__restrict__ iters,
                              uint8_t *__restrict__ failed, uint8_t *__restrict__ pend,
                              uint8_t *__restrict__ upd, double async_below, uint8_t *__restrict__ areq,
                              uint8_t *__restrict__ sch, const uint8_t *__restrict__ exact, int switch_window) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const bool ex = exact && exact[m];
  if (ex) theta = 0.0;
  const bool first = firstv ? firstv[m] != 0 : first0 != 0;   // first round of this scenario's solve
  const bool was_upd = first ? true : upd[m] != 0;
  upd[m] = 0;
  refac[m] = 0;
  areq[m] = 0;
  sch[m] = 0;
  if (!work[m]) return;
  const double v = __longlong_as_double(nb[m]);
  const double prev = fnorm[m];
  fnorm[m] = v;
  nb[m] = 0ull;
  if (v < tol) {                 // converged (NaN compares false)
    work[m] = 0;
    return;
  }
  const int it = first ? 0 : iters[m];
  bool fail = it >= itermax;
  if (!ex && switch_window > 0) stall = stall > 0 ? min(stall, switch_window) : switch_window;
  if (stall > 0) {
    double best = first ? 1e300 : fbest[m];
    int ns = first ? 0 : nstall[m];
    if (v < 0.5 * best) {
      best = v;
      ns = 0;
    } else if (was_upd) {
      ns++;
    }
    fbest[m] = best;
    nstall[m] = ns;
    fail = fail || !(v <= 1e10) || ns >= stall;
  }
  // on the shared factorisation (theta < 0) a correction that raises |F|
  // gives the solve up: it runs again as exact Newton from the step's start
  // (an inexact iteration that wanders may settle on another root)
  if (theta < 0.0 && !first && !(v <= prev)) fail = true;
  if (fail) {
    work[m] = 0;
    failed[m] = isfinite(v) ? 1 : 3;   // bit 1: |F| not finite
    return;
  }
  atomicAdd(n_unconv, 1);
  const bool pending = pend && pend[m];
  // a scenario that waited last round compares against an unchanged |F|:
  // its new factorisation has just arrived, so it asks for none.  theta < 0:
  // one shared factorisation serves the scenario (only exact Newton factorises)
  const int rf = theta < 0.0 ? 0
                 : first ? (refac0 ? refac0[m] : 1)
                         : ((pending || !was_upd) ? 0 : ((theta <= 0.0 || !(v * theta < prev)) ? 1 : 0));
  // a solve that starts with a factorisation of its own (after an event or
  // a change of dt) does not wait for one asked for under the old system:
  // the host lands and drops that request before this round's factorisation
  bool wait = pending && !(first && rf);
  if (rf && !first && async_below > 0.0 && v < async_below) {
    areq[m] = 1;   // requested: pending until its batch lands
    pend[m] = 1;
    atomicAdd(n_unconv + 3, 1);
    wait = true;
  } else if (rf) {
    refac[m] = 1;
    atomicAdd(n_unconv + 1, 1);
  }
  const bool u = !wait;
  sch[m] = (u || areq[m]) ? 1 : 0;   // the Schur kernels also run for a new request's values
  if (u) {
    upd[m] = 1;
    iters[m] = it + 1;
    atomicAdd(n_unconv + 2, 1);
  } else {
    iters[m] = it;
  }
}

// ---------------------------------------------------------------- monolithic Jacobian
__device__ __forceinline__ double be(double v, bool diff, bool diag, double dt) {
  return diff ? (diag ? 1.0 - dt * v : -dt * v) : v;
}

__global__ void __launch_bounds__(64) k_cluster_jac_mono(DevTables T, const double *__restrict__ p,
                                                         const double *__restrict__ z, double dt0,
                                                         const double *__restrict__ dtv, FaultState fs,
                                                         double *__restrict__ jv, const uint8_t *__restrict__ work) {
  const int c = T.cl_nt[blockIdx.x];
  const int m = lane_scenario();
  if (m >= T.M || !work[m]) return;
  const double dt = dtv ? dtv[m] : dt0;
  const int ldM = T.ldM;
  const int *cl = T.cl + c * GP_CLUSTER_COLS;
  const int gk = cl[GP_CL_GENROU];
  const int *g = T.genrou + gk * GP_COLS_GENROU;
  const int dp = g[0], ap = g[1], bus = g[3];
  const int vrz = T.net_ptr + 2 * bus;
  const double vr = z[(size_t)vrz * ldM + m], vi = z[(size_t)(vrz + 1) * ldM + m];
  double zs[10], j[47];
#pragma unroll
  for (int i = 0; i < 6; i++) zs[i] = z[(size_t)(dp + i) * ldM + m];
#pragma unroll
  for (int i = 0; i < 4; i++) zs[6 + i] = z[(size_t)(ap + i) * ldM + m];
  genrou_jac(p + g[2], zs, vr, vi, j);
  const int *jp = T.jp_genrou + gk * NS_GENROU;
#pragma unroll
  for (int s = 0; s < NS_GENROU; s++) {
    const int pos = jp[s];
    if (pos >= 0) jv[(size_t)pos * ldM + m] = be(j[s], GEN_SLOT_DIFF[s], GEN_SLOT_DIAG[s], dt);
  }
  const int gt = cl[GP_CL_GOVTYPE];
  if (gt == GP_IEESGO) {
    const int *d = T.ieesgo + cl[GP_CL_GOV] * GP_COLS_IEESGO;
    const double *P = p + d[2];
    const double T1 = P[0], T2 = P[1], T3 = P[2], T4 = P[3], T5 = P[4], T6 = P[5], K1 = P[6], K2 = P[7], K3 = P[8];
    const double v[15] = {-1.0 / T1, K1 / T1, (1.0 - T2 / T3) / T3, -1.0 / T3, -T2 / (T3 * T4), -1.0 / T4, -1.0 / T4,
                          K2 / T5, -1.0 / T5, K3 / T6, -1.0 / T6, 1.0 - K2, 1.0 - K3, 1.0, -1.0};
    const bool diff[15] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0};
    const bool diag[15] = {1, 0, 0, 1, 0, 0, 1, 0, 1, 0, 1, 0, 0, 0, 0};
    const int *jpd = T.jp_ieesgo + cl[GP_CL_GOV] * NS_IEESGO;
#pragma unroll
    for (int s = 0; s < 15; s++)
      if (jpd[s] >= 0) jv[(size_t)jpd[s] * ldM + m] = be(v[s], diff[s], diag[s], dt);
  } else if (gt == GP_TGOV1) {
    const int *d = T.tgov1 + cl[GP_CL_GOV] * GP_COLS_TGOV1;
    const double *P = p + d[2];
    const double R = P[0], T1 = P[1], T2 = P[4], T3 = P[5], DT = P[6], t23 = T2 / T3;
    const double v[8] = {-1.0 / T3, (1.0 - t23) / T3, -1.0 / T1, -1.0 / (R * T1), 1.0, t23, -DT, -1.0};
    const bool diff[8] = {1, 1, 1, 1, 0, 0, 0, 0};
    const bool diag[8] = {1, 0, 1, 0, 0, 0, 0, 0};
    const int *jpd = T.jp_tgov1 + cl[GP_CL_GOV] * NS_TGOV1;
#pragma unroll
    for (int s = 0; s < 8; s++)
      if (jpd[s] >= 0) jv[(size_t)jpd[s] * ldM + m] = be(v[s], diff[s], diag[s], dt);
  }
  const int et = cl[GP_CL_EXCTYPE];
  if (et == GP_SEXS) {
    const int *d = T.sexs + cl[GP_CL_EXC] * GP_COLS_SEXS;
    const double *P = p + d[1];
    const double TA_TB = P[0], TB = P[1], K = P[2], TE = P[3];
    const double evr = z[(size_t)d[2] * ldM + m], evi = z[(size_t)(d[2] + 1) * ldM + m];
    double vm = sqrt(evr * evr + evi * evi);
    if (vm == 0.0) vm = 1e-12;
    const double dr = evr / vm, di = evi / vm;
    const double v[9] = {-1.0 / TB, -(1.0 - TA_TB) * dr / TB, -(1.0 - TA_TB) * di / TB, (1.0 - TA_TB) / TB,
                         K / TE, -1.0 / TE, -K * TA_TB * dr / TE, -K * TA_TB * di / TE, K * TA_TB / TE};
    const bool diag[9] = {1, 0, 0, 0, 0, 1, 0, 0, 0};
    const int *jpd = T.jp_sexs + cl[GP_CL_EXC] * NS_SEXS;
    const double on = exc_offline(fs, ldM, m, et, cl[GP_CL_EXC]) ? 0.0 : 1.0;
#pragma unroll
    for (int s = 0; s < 9; s++)
      if (jpd[s] >= 0) jv[(size_t)jpd[s] * ldM + m] = be(on * v[s], true, diag[s], dt);
  } else if (et == GP_ESDC1A) {
    const int *d = T.esdc1a + cl[GP_CL_EXC] * GP_COLS_ESDC1A;
    const double *P = p + d[1];
    const double Ka = P[0], Ta = P[1], Kf = P[2], Tf = P[3], Ke = P[4], Te = P[5], sa = P[7], sb = P[8];
    const double e_fd = z[(size_t)(d[0] + 2) * ldM + m];
    const double evr = z[(size_t)d[2] * ldM + m], evi = z[(size_t)(d[2] + 1) * ldM + m];
    double vm = sqrt(evr * evr + evi * evi);
    if (vm == 0.0) vm = 1e-12;
    const double dsat = (sb == 0.0 || e_fd <= sa) ? 0.0 : 2.0 * sb * (e_fd - sa);
    const double v[10] = {-1.0 / Ta, -Ka / Ta, -Ka * Kf / (Ta * Tf), -Ka * (evr / vm) / Ta, -Ka * (evi / vm) / Ta,
                          Ka / Ta, -1.0 / Tf, -Kf / (Tf * Tf), 1.0 / Te, -(Ke + dsat) / Te};
    const bool diag[10] = {1, 0, 0, 0, 0, 0, 1, 0, 0, 1};
    const int *jpd = T.jp_esdc1a + cl[GP_CL_EXC] * NS_ESDC1A;
    const double on = exc_offline(fs, ldM, m, et, cl[GP_CL_EXC]) ? 0.0 : 1.0;
#pragma unroll
    for (int s = 0; s < 10; s++)
      if (jpd[s] >= 0) jv[(size_t)jpd[s] * ldM + m] = be(on * v[s], true, diag[s], dt);
  }
  if (cl[GP_CL_PSS] >= 0) {
    const int *d = T.ieeest + cl[GP_CL_PSS] * GP_COLS_IEEEST;
    const double *P = p + d[2];
    const double A1 = P[0], A2 = P[1], A3 = P[2], A4 = P[3], A5 = P[4], A6 = P[5];
    const double T1 = P[6], T2 = P[7], T3 = P[8], T4 = P[9], T5 = P[10], T6 = P[11], KS = P[12];
    const double rA = A6 / A4, y21 = rA, y22 = A5 - A3 * rA, y23 = 1.0 - rA;
    const double r12 = T1 / T2, y31 = r12 * y21, y32 = r12 * y22, y33 = r12 * y23, y34 = 1.0 - r12;
    const double r34 = T3 / T4, y41 = r34 * y31, y42 = r34 * y32, y43 = r34 * y33, y44 = r34 * y34, y45 = 1.0 - r34;
    const double r56 = T5 / T6;
    const double v[32] = {-A1 / A2, -1.0 / A2, 1.0 / A2, 1.0, 1.0 / A4, -A3 / A4, -1.0 / A4, 1.0,
                          y21 / T2, y22 / T2, y23 / T2, -1.0 / T2,
                          y31 / T4, y32 / T4, y33 / T4, y34 / T4, -1.0 / T4,
                          KS * y41 / T6, KS * y42 / T6, KS * y43 / T6, KS * y44 / T6, KS * y45 / T6, -1.0 / T6,
                          -r56 * KS * y41, -r56 * KS * y42, -r56 * KS * y43, -r56 * KS * y44, -r56 * KS * y45, r56, 1.0,
                          0.0, 0.0};
    const int *jpd = T.jp_ieeest + cl[GP_CL_PSS] * NS_IEEEST;
#pragma unroll
    for (int s = 0; s < 32; s++) {
      if (jpd[s] < 0) continue;
      const bool diff = !(s >= 23 && s <= 29);
      const bool diag = (s == 0 || s == 5 || s == 11 || s == 16 || s == 22 || s == 30 || s == 31);
      jv[(size_t)jpd[s] * ldM + m] = be(v[s], diff, diag, dt);
    }
  }
}

// Static generator Jacobian entries (src/kernels/static_gen.jl:168-220).
// out: drr, dri, dir, dii, vr_ap, vi_ap, vr_ap1, vi_ap1, ap_vr, ap_vi, ap1_vi
__device__ __forceinline__ void sg_jac(const double *P, int bt, double vr, double vi, double za, double za1, double *o) {
  const double vm2r = vr * vr + vi * vi, vm2 = vm2r > 0.2 ? vm2r : 0.2;
  double pv, qv;
  if (bt == 2) { pv = P[0]; qv = za; }
  else if (bt == 3) { pv = za; qv = za1; }
  else { pv = P[0]; qv = P[1]; }
  const double irn = pv * vr + qv * vi, iin = pv * vi - qv * vr;
  if (vm2r > 0.2) {
    const double vm4 = vm2 * vm2;
    o[0] = pv / vm2 - 2.0 * vr * irn / vm4; o[1] = qv / vm2 - 2.0 * vi * irn / vm4;
    o[2] = -qv / vm2 - 2.0 * vr * iin / vm4; o[3] = pv / vm2 - 2.0 * vi * iin / vm4;
  } else {
    o[0] = pv / vm2; o[1] = qv / vm2; o[2] = -qv / vm2; o[3] = pv / vm2;
  }
  for (int i = 4; i < 11; i++) o[i] = 0.0;
  if (bt == 2) {
    o[4] = vi / vm2; o[5] = -vr / vm2; o[8] = 2.0 * vr; o[9] = 2.0 * vi;
  } else if (bt == 3) {
    o[4] = vr / vm2; o[5] = vi / vm2; o[6] = vi / vm2; o[7] = -vr / vm2; o[8] = 1.0; o[10] = 1.0;
  }
}

// ZIP-load Jacobian on the bus diagonal block (src/kernels/zipload.jl:133-179).
__device__ __forceinline__ void zip_jac(const double *P, double vr, double vi, double *o) {
  const double pl = P[0], ql = P[1], al = P[2], yre = P[7], yim = P[8];
  const double vm2 = vr * vr + vi * vi;
  double prr, pri, pir, pii;
  if (vm2 > 0.2) {
    const double iv4 = 1.0 / (vm2 * vm2);
    prr = (1.0 - al) * ((pl * vr - ql * vi) * 2.0 * vr - pl * vm2) * iv4;
    pri = (1.0 - al) * ((pl * vr - ql * vi) * 2.0 * vi + ql * vm2) * iv4;
    pir = (1.0 - al) * ((ql * vr + pl * vi) * 2.0 * vr - ql * vm2) * iv4;
    pii = (1.0 - al) * ((ql * vr + pl * vi) * 2.0 * vi - pl * vm2) * iv4;
  } else {
    const double it = 1.0 / 0.2;
    prr = (1.0 - al) * (-pl) * it; pri = (1.0 - al) * ql * it;
    pir = (1.0 - al) * (-ql) * it; pii = (1.0 - al) * (-pl) * it;
  }
  o[0] += -al * yre + prr;
  o[1] += al * yim + pri;
  o[2] += -al * yim + pir;
  o[3] += -al * yre + pii;
}

// Bus diagonal block of the network rows: -Y_bb + loads + static generator - faults.
__device__ __forceinline__ void bus_diag(const DevTables &T, const double *p, const double *z, const FaultState &fs,
                                         int b, int m, double vr, double vi, double *dg, double *sgo) {
  const int ldM = T.ldM;
  dg[0] = -T.ydiag[4 * b]; dg[1] = -T.ydiag[4 * b + 1]; dg[2] = -T.ydiag[4 * b + 2]; dg[3] = -T.ydiag[4 * b + 3];
  for (int q = T.bus_zip_p[b]; q < T.bus_zip_p[b + 1]; q++)
    if (!zip_dropped(fs, ldM, m, T.bus_zip[q])) zip_jac(p + T.zip[T.bus_zip[q] * GP_COLS_ZIPLOAD + 1], vr, vi, dg);
  const int sgk = T.bus_sg[b];
  for (int i = 0; i < 11; i++) sgo[i] = 0.0;
  if (sgk >= 0) {
    const int *d = T.sg + sgk * GP_COLS_STATICGEN;
    const int ap = d[1], bt = d[3];
    const double za = ap >= 0 ? z[(size_t)ap * ldM + m] : 0.0;
    const double za1 = bt == 3 ? z[(size_t)(ap + 1) * ldM + m] : 0.0;
    sg_jac(p + d[2], bt, vr, vi, za, za1, sgo);
    dg[0] += sgo[0]; dg[1] += sgo[1]; dg[2] += sgo[2]; dg[3] += sgo[3];
  }
#pragma unroll
  for (int k = 0; k < MAXFAULT; k++) {
    const int code = fs.bus[k * ldM + m];
    if (code == b) {
      const double gf = fs.g[k * ldM + m];
      dg[0] -= gf;
      dg[3] -= gf;
    } else if (code <= -2 && code > kDropCode0) {   // a tripped line at this bus: -Y_bb + yff (or ytt)
      const int t = -2 - code;
      const int a = T.tr_bus[2 * t], c = T.tr_bus[2 * t + 1];
      if (a == b || c == b) {
        const double *yy = T.tr_y + 8 * t + (a == b ? 0 : 6);
        dg[0] += yy[0];
        dg[1] -= yy[1];
        dg[2] += yy[1];
        dg[3] += yy[0];
      }
    }
  }
}

__global__ void __launch_bounds__(64) k_bus_jac_mono(DevTables T, const double *__restrict__ p,
                                                     const double *__restrict__ z, FaultState fs,
                                                     double *__restrict__ jv, const uint8_t *__restrict__ work) {
  const int b = blockIdx.x;
  const int m = lane_scenario();
  if (m >= T.M || !work[m]) return;
  const int ldM = T.ldM;
  const int vrz = T.net_ptr + 2 * b;
  const double vr = z[(size_t)vrz * ldM + m], vi = z[(size_t)(vrz + 1) * ldM + m];
  double dg[4], sgo[11];
  bus_diag(T, p, z, fs, b, m, vr, vi, dg, sgo);
  const int *jb = T.jp_bus + 15 * b;
#pragma unroll
  for (int i = 0; i < 4; i++) jv[(size_t)jb[i] * ldM + m] = dg[i];
#pragma unroll
  for (int i = 4; i < 11; i++)
    if (jb[i] >= 0) jv[(size_t)jb[i] * ldM + m] = sgo[i];
  for (int q = T.jp_yoff_p[b]; q < T.jp_yoff_p[b + 1]; q++) jv[(size_t)T.jp_yoff_pos[q] * ldM + m] = -T.jp_yoff_val[q];
  // off-diagonal entries that line trips change (written again, after the constants)
  for (int q = T.tv_p[b]; q < T.tv_p[b + 1]; q++) {
    const int *e = T.tv_i + 4 * q;
    jv[(size_t)e[2] * ldM + m] = trip_offdiag(T, fs, b, m, e[0], e[1], T.tv_base[q]);
  }
}

// ---------------------------------------------------------------- small dense solves
template <int N, int NR, typename TF>
__device__ __forceinline__ void dense_solve_nopiv(TF (&a)[N][N], TF (&r)[N][NR]) {
#pragma unroll
  for (int k = 0; k < N; k++) {
    const TF inv = TF(1) / a[k][k];
#pragma unroll
    for (int i = k + 1; i < N; i++) {
      const TF l = a[i][k] * inv;
#pragma unroll
      for (int j = k + 1; j < N; j++) a[i][j] -= l * a[k][j];
#pragma unroll
      for (int c = 0; c < NR; c++) r[i][c] -= l * r[k][c];
    }
  }
#pragma unroll
  for (int k = N - 1; k >= 0; k--) {
    const TF inv = TF(1) / a[k][k];
#pragma unroll
    for (int c = 0; c < NR; c++) {
      TF s = r[k][c];
#pragma unroll
      for (int j = k + 1; j < N; j++) s -= a[k][j] * r[j][c];
      r[k][c] = s * inv;
    }
  }
}

// Gaussian elimination with partial pivoting; row exchanges are done with
// selects so every index stays a compile-time constant (registers only).
template <int N, int NR, typename TF>
__device__ __forceinline__ void dense_solve_pp(TF (&a)[N][N], TF (&r)[N][NR]) {
#pragma unroll
  for (int k = 0; k < N; k++) {
    TF best = fabs(a[k][k]);
    int piv = k;
#pragma unroll
    for (int i = k + 1; i < N; i++) {
      const TF v = fabs(a[i][k]);
      if (v > best) { best = v; piv = i; }
    }
#pragma unroll
    for (int i = k + 1; i < N; i++) {
      const bool sw = (piv == i);
#pragma unroll
      for (int j = k; j < N; j++) {
        const TF ta = a[k][j], tb = a[i][j];
        a[k][j] = sw ? tb : ta;
        a[i][j] = sw ? ta : tb;
      }
#pragma unroll
      for (int c = 0; c < NR; c++) {
        const TF ta = r[k][c], tb = r[i][c];
        r[k][c] = sw ? tb : ta;
        r[i][c] = sw ? ta : tb;
      }
    }
    const TF inv = TF(1) / a[k][k];
#pragma unroll
    for (int i = k + 1; i < N; i++) {
      const TF l = a[i][k] * inv;
#pragma unroll
      for (int j = k + 1; j < N; j++) a[i][j] -= l * a[k][j];
#pragma unroll
      for (int c = 0; c < NR; c++) r[i][c] -= l * r[k][c];
    }
  }
#pragma unroll
  for (int k = N - 1; k >= 0; k--) {
    const TF inv = TF(1) / a[k][k];
#pragma unroll
    for (int c = 0; c < NR; c++) {
      TF s = r[k][c];
#pragma unroll
      for (int j = k + 1; j < N; j++) s -= a[k][j] * r[j][c];
      r[k][c] = s * inv;
    }
  }
}

// ---------------------------------------------------------------- controller blocks (BE-scaled)
// Each builder fills the block matrix a (states x states), the column of the
// block rows at the generator speed w (aw) and, for exciters, at the PSS
// output v_s (av) and at the bus voltage (bvr, bvi).
template <typename TF>
__device__ __forceinline__ void ieesgo_block(const double *P, double dt, TF (&a)[6][6], TF (&aw)[6]) {
  const double T1 = P[0], T2 = P[1], T3 = P[2], T4 = P[3], T5 = P[4], T6 = P[5], K1 = P[6], K2 = P[7], K3 = P[8];
#pragma unroll
  for (int i = 0; i < 6; i++) {
    aw[i] = 0;
#pragma unroll
    for (int j = 0; j < 6; j++) a[i][j] = 0;
  }
  a[0][0] = TF(1.0 + dt / T1);                   aw[0] = TF(-dt * K1 / T1);
  a[1][0] = TF(-dt * (1.0 - T2 / T3) / T3);      a[1][1] = TF(1.0 + dt / T3);
  a[2][0] = TF(dt * T2 / (T3 * T4));             a[2][1] = TF(dt / T4);  a[2][2] = TF(1.0 + dt / T4);
  a[3][2] = TF(-dt * K2 / T5);                   a[3][3] = TF(1.0 + dt / T5);
  a[4][3] = TF(-dt * K3 / T6);                   a[4][4] = TF(1.0 + dt / T6);
  a[5][2] = TF(1.0 - K2); a[5][3] = TF(1.0 - K3); a[5][4] = TF(1.0); a[5][5] = TF(-1.0);
}
template <typename TF>
__device__ __forceinline__ void tgov1_block(const double *P, double dt, TF (&a)[3][3], TF (&aw)[3]) {
  const double R = P[0], T1 = P[1], T2 = P[4], T3 = P[5], DT = P[6], t23 = T2 / T3;
  a[0][0] = TF(1.0 + dt / T3); a[0][1] = TF(-dt * (1.0 - t23) / T3); a[0][2] = 0;  aw[0] = 0;
  a[1][0] = 0; a[1][1] = TF(1.0 + dt / T1); a[1][2] = 0;                          aw[1] = TF(dt / (R * T1));
  a[2][0] = TF(1.0); a[2][1] = TF(t23); a[2][2] = TF(-1.0);                        aw[2] = TF(-DT);
}
template <typename TF>
__device__ __forceinline__ void ieeest_block(const double *P, double dt, TF (&a)[8][8], TF (&aw)[8]) {
  const double A1 = P[0], A2 = P[1], A3 = P[2], A4 = P[3], A5 = P[4], A6 = P[5];
  const double T1 = P[6], T2 = P[7], T3 = P[8], T4 = P[9], T5 = P[10], T6 = P[11], KS = P[12];
  const double rA = A6 / A4, y2[3] = {rA, A5 - A3 * rA, 1.0 - rA};
  const double r12 = T1 / T2, y3[4] = {r12 * y2[0], r12 * y2[1], r12 * y2[2], 1.0 - r12};
  const double r34 = T3 / T4, y4[5] = {r34 * y3[0], r34 * y3[1], r34 * y3[2], r34 * y3[3], 1.0 - r34};
  const double r56 = T5 / T6;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    aw[i] = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) a[i][j] = 0;
  }
  a[0][0] = TF(1.0 + dt * A1 / A2); a[0][1] = TF(dt / A2); aw[0] = TF(-dt / A2);
  a[1][0] = TF(-dt); a[1][1] = TF(1.0);
  a[2][1] = TF(-dt / A4); a[2][2] = TF(1.0 + dt * A3 / A4); a[2][3] = TF(dt / A4);
  a[3][2] = TF(-dt); a[3][3] = TF(1.0);
#pragma unroll
  for (int j = 0; j < 3; j++) a[4][1 + j] = TF(-dt * y2[j] / T2);
  a[4][4] = TF(1.0 + dt / T2);
#pragma unroll
  for (int j = 0; j < 4; j++) a[5][1 + j] = TF(-dt * y3[j] / T4);
  a[5][5] = TF(1.0 + dt / T4);
#pragma unroll
  for (int j = 0; j < 5; j++) a[6][1 + j] = TF(-dt * KS * y4[j] / T6);
  a[6][6] = TF(1.0 + dt / T6);
#pragma unroll
  for (int j = 0; j < 5; j++) a[7][1 + j] = TF(-r56 * KS * y4[j]);
  a[7][6] = TF(r56); a[7][7] = TF(1.0);
}
// SEXS / ESDC1A (2 or 3 states, padded to 3 with an identity row)
template <typename TF>
__device__ __forceinline__ void exciter_block(int et, const double *P, double dt, double e_fd, double vr, double vi,
                                              bool off, TF (&a)[3][3], TF (&av)[3], TF (&bvr)[3], TF (&bvi)[3]) {
  double vm = sqrt(vr * vr + vi * vi);
  if (vm == 0.0) vm = 1e-12;
  const double dr = vr / vm, di = vi / vm;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    av[i] = 0; bvr[i] = 0; bvi[i] = 0;
#pragma unroll
    for (int j = 0; j < 3; j++) a[i][j] = 0;
  }
  if (off) {   // offline: identity rows, no coupling
    a[0][0] = a[1][1] = a[2][2] = TF(1.0);
  } else if (et == GP_SEXS) {
    const double TA_TB = P[0], TB = P[1], K = P[2], TE = P[3];
    a[0][0] = TF(1.0 + dt / TB);
    bvr[0] = TF(dt * (1.0 - TA_TB) * dr / TB); bvi[0] = TF(dt * (1.0 - TA_TB) * di / TB);
    av[0] = TF(-dt * (1.0 - TA_TB) / TB);
    a[1][0] = TF(-dt * K / TE); a[1][1] = TF(1.0 + dt / TE);
    bvr[1] = TF(dt * K * TA_TB * dr / TE); bvi[1] = TF(dt * K * TA_TB * di / TE);
    av[1] = TF(-dt * K * TA_TB / TE);
    a[2][2] = TF(1.0);
  } else {
    const double Ka = P[0], Ta = P[1], Kf = P[2], Tf = P[3], Ke = P[4], Te = P[5], sa = P[7], sb = P[8];
    const double dsat = (sb == 0.0 || e_fd <= sa) ? 0.0 : 2.0 * sb * (e_fd - sa);
    a[0][0] = TF(1.0 + dt / Ta); a[0][1] = TF(dt * Ka / Ta); a[0][2] = TF(dt * Ka * Kf / (Ta * Tf));
    bvr[0] = TF(dt * Ka * dr / Ta); bvi[0] = TF(dt * Ka * di / Ta); av[0] = TF(-dt * Ka / Ta);
    a[1][1] = TF(1.0 + dt / Tf); a[1][2] = TF(dt * Kf / (Tf * Tf));
    a[2][0] = TF(-dt / Te); a[2][2] = TF(1.0 + dt * (Ke + dsat) / Te);
  }
}

// ---------------------------------------------------------------- Schur cluster kernel
// For one non-trivial cluster and 64 scenarios: builds the BE-scaled block
// A_k, the voltage columns B_k and the network rows C_k, and computes
//   X = A_k^{-1} [B_vr | B_vi | f_k],  D_k = C_k X[:,0:2],  g_k = C_k X[:,2]
// (src/schur.jl:195-247,321-348).  The block is eliminated controller by
// controller: each controller's states are affine in the generator speed w
// (and the exciter's in the PSS output), so the generator block absorbs them
// as a rank-one change of its w column; the generator's four algebraic
// states are substituted exactly and the remaining 6x6 differential block is
// solved with partial pivoting.
template <typename TF>
__global__ void __launch_bounds__(64) k_cluster_schur(DevTables T, const double *__restrict__ p,
                                                      const double *__restrict__ z, const double *__restrict__ f,
                                                      double dt0, const double *__restrict__ dtv, FaultState fs,
                                                      TF *__restrict__ X, double *__restrict__ Dg,
                                                      const uint8_t *__restrict__ work) {
  const int c = T.cl_nt[blockIdx.x];
  const int m = lane_scenario();
  if (m >= T.M || !work[m]) return;
  const double dt = dtv ? dtv[m] : dt0;
  const int ldM = T.ldM;
  const int *cl = T.cl + c * GP_CLUSTER_COLS;
  const int gk = cl[GP_CL_GENROU];
  const int *g = T.genrou + gk * GP_COLS_GENROU;
  const int dp = g[0], ap = g[1], bus = g[3];
  const int vrz = T.net_ptr + 2 * bus;
  const double vr = z[(size_t)vrz * ldM + m], vi = z[(size_t)(vrz + 1) * ldM + m];
  double zs[10], j[47];
#pragma unroll
  for (int i = 0; i < 6; i++) zs[i] = z[(size_t)(dp + i) * ldM + m];
#pragma unroll
  for (int i = 0; i < 4; i++) zs[6 + i] = z[(size_t)(ap + i) * ldM + m];
  genrou_jac(p + g[2], zs, vr, vi, j);
  // BE scaling of the generator's differential rows
#pragma unroll
  for (int s = 0; s < 25; s++) j[s] = be(j[s], true, GEN_SLOT_DIAG[s], dt);
  // RHS of the generator rows: columns (vr, vi, f)
  TF R[10][3];
#pragma unroll
  for (int i = 0; i < 10; i++) {
    R[i][0] = 0; R[i][1] = 0;
    R[i][2] = TF(f[(size_t)((i < 6 ? dp + i : ap + i - 6)) * ldM + m]);
  }
  R[8][0] = TF(j[35]); R[8][1] = TF(j[36]);
  R[9][0] = TF(j[39]); R[9][1] = TF(j[40]);
  TF a04 = 0, a44 = TF(j[19]);

  // ---- governor: p_m = alpha_pm - beta_pm * w
  const int gt = cl[GP_CL_GOVTYPE];
  const int *gd = nullptr;
  TF alpha_pm[3] = {0, 0, 0}, beta_pm = 0;
  if (gt == GP_IEESGO) {
    gd = T.ieesgo + cl[GP_CL_GOV] * GP_COLS_IEESGO;
    TF a[6][6], aw[6], r[6][4];
    ieesgo_block<TF>(p + gd[2], dt, a, aw);
#pragma unroll
    for (int i = 0; i < 6; i++) {
      r[i][0] = 0; r[i][1] = 0;
      r[i][2] = TF(f[(size_t)(i < 5 ? gd[0] + i : gd[1]) * ldM + m]);
      r[i][3] = aw[i];
    }
    dense_solve_nopiv<6, 4, TF>(a, r);
    alpha_pm[0] = r[5][0]; alpha_pm[1] = r[5][1]; alpha_pm[2] = r[5][2]; beta_pm = r[5][3];
  } else if (gt == GP_TGOV1) {
    gd = T.tgov1 + cl[GP_CL_GOV] * GP_COLS_TGOV1;
    TF a[3][3], aw[3], r[3][4];
    tgov1_block<TF>(p + gd[2], dt, a, aw);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      r[i][0] = 0; r[i][1] = 0;
      r[i][2] = TF(f[(size_t)(i < 2 ? gd[0] + i : gd[1]) * ldM + m]);
      r[i][3] = aw[i];
    }
    dense_solve_nopiv<3, 4, TF>(a, r);
    alpha_pm[0] = r[2][0]; alpha_pm[1] = r[2][1]; alpha_pm[2] = r[2][2]; beta_pm = r[2][3];
  }
  if (gt != 0) {
    const TF j4pm = TF(j[22]);   // d f5 / d p_m, already BE-scaled
#pragma unroll
    for (int col = 0; col < 3; col++) R[4][col] -= j4pm * alpha_pm[col];
    a44 -= j4pm * beta_pm;
  }
  // ---- PSS: v_s = alpha_vs - beta_vs * w
  const int *pd = cl[GP_CL_PSS] >= 0 ? T.ieeest + cl[GP_CL_PSS] * GP_COLS_IEEEST : nullptr;
  TF alpha_vs[3] = {0, 0, 0}, beta_vs = 0;
  if (pd) {
    TF a[8][8], aw[8], r[8][4];
    ieeest_block<TF>(p + pd[2], dt, a, aw);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      r[i][0] = 0; r[i][1] = 0;
      r[i][2] = TF(f[(size_t)(i < 7 ? pd[0] + i : pd[1]) * ldM + m]);
      r[i][3] = aw[i];
    }
    dense_solve_nopiv<8, 4, TF>(a, r);
    alpha_vs[0] = r[7][0]; alpha_vs[1] = r[7][1]; alpha_vs[2] = r[7][2]; beta_vs = r[7][3];
  }
  // ---- exciter: e_fd = a_e - gamma * v_s
  const int et = cl[GP_CL_EXCTYPE];
  const int *ed = et == GP_SEXS ? T.sexs + cl[GP_CL_EXC] * GP_COLS_SEXS
                                : (et == GP_ESDC1A ? T.esdc1a + cl[GP_CL_EXC] * GP_COLS_ESDC1A : nullptr);
  const int ne = et == GP_SEXS ? 2 : 3;
  const bool eoff = ed && exc_offline(fs, ldM, m, et, cl[GP_CL_EXC]);
  if (ed) {
    TF a[3][3], av[3], bvr[3], bvi[3], r[3][4];
    const double e_fd0 = et == GP_ESDC1A ? z[(size_t)(ed[0] + 2) * ldM + m] : 0.0;
    exciter_block<TF>(et, p + ed[1], dt, e_fd0, vr, vi, eoff, a, av, bvr, bvi);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      r[i][0] = bvr[i]; r[i][1] = bvi[i];
      r[i][2] = i < ne ? TF(f[(size_t)(ed[0] + i) * ldM + m]) : TF(0);
      r[i][3] = av[i];
    }
    dense_solve_nopiv<3, 4, TF>(a, r);
    const bool is_sexs = et == GP_SEXS;   // e_fd is state 1 of SEXS, state 2 of ESDC1A
    const TF gam = is_sexs ? r[1][3] : r[2][3];
    const TF j0e = TF(j[5]);   // d f1 / d e_fd, already BE-scaled
#pragma unroll
    for (int col = 0; col < 3; col++) R[0][col] -= j0e * ((is_sexs ? r[1][col] : r[2][col]) - gam * alpha_vs[col]);
    a04 += j0e * gam * beta_vs;
  }

  // ---- generator: substitute the algebraic states, solve the 6x6 block
  const TF a60 = TF(j[25]), a62 = TF(j[26]), a66 = TF(j[27]), a69 = TF(j[28]);
  const TF a71 = TF(j[29]), a73 = TF(j[30]), a77 = TF(j[31]), a78 = TF(j[32]);
  const TF a85 = TF(j[33]), a87 = TF(j[34]), a95 = TF(j[37]), a96 = TF(j[38]);
  const TF cd_eq = -a60 / a69, cd_p1 = -a62 / a69, cd_de = a66 * a95 / (a96 * a69);
  const TF cq_ed = -a71 / a78, cq_p2 = -a73 / a78, cq_de = a77 * a85 / (a87 * a78);
  TF rd[3], rq[3];
#pragma unroll
  for (int col = 0; col < 3; col++) {
    rd[col] = (R[6][col] - a66 * R[9][col] / a96) / a69;
    rq[col] = (R[7][col] - a77 * R[8][col] / a87) / a78;
  }
  const TF a0id = TF(j[2]), a1iq = TF(j[8]), a2id = TF(j[11]), a3iq = TF(j[14]), a4iq = TF(j[20]), a4id = TF(j[21]);
  TF A[6][6], B[6][3];
  A[0][0] = TF(j[0]) + a0id * cd_eq; A[0][1] = TF(j[3]); A[0][2] = TF(j[1]) + a0id * cd_p1; A[0][3] = TF(j[4]);
  A[0][4] = a04; A[0][5] = a0id * cd_de;
  A[1][0] = 0; A[1][1] = TF(j[6]) + a1iq * cq_ed; A[1][2] = 0; A[1][3] = TF(j[7]) + a1iq * cq_p2; A[1][4] = 0;
  A[1][5] = a1iq * cq_de;
  A[2][0] = TF(j[9]) + a2id * cd_eq; A[2][1] = 0; A[2][2] = TF(j[10]) + a2id * cd_p1; A[2][3] = 0; A[2][4] = 0;
  A[2][5] = a2id * cd_de;
  A[3][0] = 0; A[3][1] = TF(j[12]) + a3iq * cq_ed; A[3][2] = 0; A[3][3] = TF(j[13]) + a3iq * cq_p2; A[3][4] = 0;
  A[3][5] = a3iq * cq_de;
  A[4][0] = TF(j[15]) + a4id * cd_eq; A[4][1] = TF(j[16]) + a4iq * cq_ed; A[4][2] = TF(j[17]) + a4id * cd_p1;
  A[4][3] = TF(j[18]) + a4iq * cq_p2; A[4][4] = a44; A[4][5] = a4iq * cq_de + a4id * cd_de;
  A[5][0] = 0; A[5][1] = 0; A[5][2] = 0; A[5][3] = 0; A[5][4] = TF(j[23]); A[5][5] = TF(j[24]);
  const TF rowid[6] = {a0id, 0, a2id, 0, a4id, 0};
  const TF rowiq[6] = {0, a1iq, 0, a3iq, a4iq, 0};
#pragma unroll
  for (int r = 0; r < 6; r++)
#pragma unroll
    for (int col = 0; col < 3; col++) B[r][col] = R[r][col] - rowid[r] * rd[col] - rowiq[r] * rq[col];
  dense_solve_pp<6, 3, TF>(A, B);
  TF xg[10][3];
#pragma unroll
  for (int col = 0; col < 3; col++) {
#pragma unroll
    for (int i = 0; i < 6; i++) xg[i][col] = B[i][col];
    const TF de = B[5][col];
    const TF vd = (R[8][col] - a85 * de) / a87;
    const TF vq = (R[9][col] - a95 * de) / a96;
    xg[6][col] = vq;
    xg[7][col] = vd;
    xg[8][col] = rq[col] + cq_ed * B[1][col] + cq_p2 * B[3][col] + cq_de * de;
    xg[9][col] = rd[col] + cd_eq * B[0][col] + cd_p1 * B[2][col] + cd_de * de;
  }
  // ---- outputs: X rows in cluster order, D_k, g_k
  TF *Xc = X + (size_t)T.cl_xoff[c] * 3 * ldM;
#pragma unroll
  for (int i = 0; i < 10; i++)
#pragma unroll
    for (int col = 0; col < 3; col++) Xc[(size_t)(i * 3 + col) * ldM + m] = xg[i][col];
  int row = 10;
  // governor back-substitution: x = A^{-1}(r - aw w)
  if (gt == GP_IEESGO) {
    TF a[6][6], aw[6], r[6][3];
    ieesgo_block<TF>(p + gd[2], dt, a, aw);
#pragma unroll
    for (int i = 0; i < 6; i++) {
      const TF fi = TF(f[(size_t)(i < 5 ? gd[0] + i : gd[1]) * ldM + m]);
      r[i][0] = -aw[i] * xg[4][0]; r[i][1] = -aw[i] * xg[4][1]; r[i][2] = fi - aw[i] * xg[4][2];
    }
    dense_solve_nopiv<6, 3, TF>(a, r);
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
      for (int col = 0; col < 3; col++) Xc[(size_t)((row + i) * 3 + col) * ldM + m] = r[i][col];
    row += 6;
  } else if (gt == GP_TGOV1) {
    TF a[3][3], aw[3], r[3][3];
    tgov1_block<TF>(p + gd[2], dt, a, aw);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const TF fi = TF(f[(size_t)(i < 2 ? gd[0] + i : gd[1]) * ldM + m]);
      r[i][0] = -aw[i] * xg[4][0]; r[i][1] = -aw[i] * xg[4][1]; r[i][2] = fi - aw[i] * xg[4][2];
    }
    dense_solve_nopiv<3, 3, TF>(a, r);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int col = 0; col < 3; col++) Xc[(size_t)((row + i) * 3 + col) * ldM + m] = r[i][col];
    row += 3;
  }
  TF vs[3] = {0, 0, 0};
  TF xp[8][3];
  if (pd) {
    TF a[8][8], aw[8];
    ieeest_block<TF>(p + pd[2], dt, a, aw);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const TF fi = TF(f[(size_t)(i < 7 ? pd[0] + i : pd[1]) * ldM + m]);
      xp[i][0] = -aw[i] * xg[4][0]; xp[i][1] = -aw[i] * xg[4][1]; xp[i][2] = fi - aw[i] * xg[4][2];
    }
    dense_solve_nopiv<8, 3, TF>(a, xp);
    vs[0] = xp[7][0]; vs[1] = xp[7][1]; vs[2] = xp[7][2];
  }
  if (ed) {
    TF a[3][3], av[3], bvr[3], bvi[3], r[3][3];
    const double e_fd0 = et == GP_ESDC1A ? z[(size_t)(ed[0] + 2) * ldM + m] : 0.0;
    exciter_block<TF>(et, p + ed[1], dt, e_fd0, vr, vi, eoff, a, av, bvr, bvi);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const TF fi = i < ne ? TF(f[(size_t)(ed[0] + i) * ldM + m]) : TF(0);
      r[i][0] = bvr[i] - av[i] * vs[0]; r[i][1] = bvi[i] - av[i] * vs[1]; r[i][2] = fi - av[i] * vs[2];
    }
    dense_solve_nopiv<3, 3, TF>(a, r);
#pragma unroll
    for (int i = 0; i < 3; i++)
      if (i < ne)
#pragma unroll
        for (int col = 0; col < 3; col++) Xc[(size_t)((row + i) * 3 + col) * ldM + m] = r[i][col];
    row += ne;
  }
  if (pd) {
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
      for (int col = 0; col < 3; col++) Xc[(size_t)((row + i) * 3 + col) * ldM + m] = xp[i][col];
  }
  // C_k: network rows at (delta, i_q, i_d)
  const double cr[3] = {j[41], j[42], j[43]}, ci[3] = {j[44], j[45], j[46]};
  const int cidx[3] = {5, 8, 9};
  double D[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int t = 0; t < 3; t++) {
    const double x0 = (double)xg[cidx[t]][0], x1 = (double)xg[cidx[t]][1], x2 = (double)xg[cidx[t]][2];
    D[0] += cr[t] * x0; D[1] += cr[t] * x1;
    D[2] += ci[t] * x0; D[3] += ci[t] * x1;
    D[4] += cr[t] * x2; D[5] += ci[t] * x2;
  }
#pragma unroll
  for (int k = 0; k < 6; k++) Dg[(size_t)(6 * c + k) * ldM + m] = D[k];
}

// Bus kernel of the Schur path: S entries of the bus (diagonal block and the
// static generator's couplings) and the reduced right-hand side
//   rhs = -(F_red - sum_k C_k A_k^{-1} F_k).
__global__ void __launch_bounds__(64) k_bus_schur(DevTables T, const double *__restrict__ p,
                                                  const double *__restrict__ z, const double *__restrict__ f,
                                                  const double *__restrict__ Dg, FaultState fs,
                                                  double *__restrict__ svar, double *__restrict__ rhs,
                                                  const uint8_t *__restrict__ work) {
  const int b = blockIdx.x;
  const int m = lane_scenario();
  if (m >= T.M || !work[m]) return;
  const int ldM = T.ldM;
  const int vrz = T.net_ptr + 2 * b;
  const double vr = z[(size_t)vrz * ldM + m], vi = z[(size_t)(vrz + 1) * ldM + m];
  double dg[4], sgo[11];
  bus_diag(T, p, z, fs, b, m, vr, vi, dg, sgo);
  double rr = f[(size_t)vrz * ldM + m], ri = f[(size_t)(vrz + 1) * ldM + m];
  for (int q = T.bus_cl_p[b]; q < T.bus_cl_p[b + 1]; q++) {
    const int c = T.bus_cl[q];
    dg[0] -= Dg[(size_t)(6 * c) * ldM + m];
    dg[1] -= Dg[(size_t)(6 * c + 1) * ldM + m];
    dg[2] -= Dg[(size_t)(6 * c + 2) * ldM + m];
    dg[3] -= Dg[(size_t)(6 * c + 3) * ldM + m];
    rr -= Dg[(size_t)(6 * c + 4) * ldM + m];
    ri -= Dg[(size_t)(6 * c + 5) * ldM + m];
  }
  const int *sv = T.sv_bus + 15 * b;
#pragma unroll
  for (int i = 0; i < 4; i++) svar[(size_t)sv[i] * ldM + m] = dg[i];
#pragma unroll
  for (int i = 4; i < 11; i++)
    if (sv[i] >= 0) svar[(size_t)sv[i] * ldM + m] = sgo[i];
  // off-diagonal entries that line trips change (variable slots instead of constants)
  for (int q = T.tv_p[b]; q < T.tv_p[b + 1]; q++) {
    const int *e = T.tv_i + 4 * q;
    svar[(size_t)e[3] * ldM + m] = trip_offdiag(T, fs, b, m, e[0], e[1], T.tv_base[q]);
  }
  rhs[(size_t)T.red_of_z[vrz] * ldM + m] = -rr;
  rhs[(size_t)T.red_of_z[vrz + 1] * ldM + m] = -ri;
  const int sgk = T.bus_sg[b];
  if (sgk >= 0) {
    const int *d = T.sg + sgk * GP_COLS_STATICGEN;
    const int ap = d[1], bt = d[3];
    if (bt == 2 || bt == 3) rhs[(size_t)T.red_of_z[ap] * ldM + m] = -f[(size_t)ap * ldM + m];
    if (bt == 3) rhs[(size_t)T.red_of_z[ap + 1] * ldM + m] = -f[(size_t)(ap + 1) * ldM + m];
  }
}

// dw_k = -(A_k^{-1} F_k + A_k^{-1} B_k dv_bus);  z_w += dw_k  (src/schur.jl:389-421)
template <typename TF>
__global__ void __launch_bounds__(64) k_cluster_backsub(DevTables T, const TF *__restrict__ X,
                                                        const double *__restrict__ dv, double *__restrict__ z,
                                                        const uint8_t *__restrict__ work) {
  const int c = T.cl_nt[blockIdx.x];
  const int m = lane_scenario();
  if (m >= T.M || !work[m]) return;
  const int ldM = T.ldM;
  const int *cl = T.cl + c * GP_CLUSTER_COLS;
  const int ws = cl[GP_CL_WSTART], wsz = cl[GP_CL_WSIZE];
  const int vrz = T.net_ptr + 2 * cl[GP_CL_BUS];
  const double d0 = dv[(size_t)T.red_of_z[vrz] * ldM + m], d1 = dv[(size_t)T.red_of_z[vrz + 1] * ldM + m];
  const TF *Xc = X + (size_t)T.cl_xoff[c] * 3 * ldM;
  for (int i = 0; i < wsz; i++) {
    const double x0 = Xc[(size_t)(3 * i) * ldM + m], x1 = Xc[(size_t)(3 * i + 1) * ldM + m];
    const double x2 = Xc[(size_t)(3 * i + 2) * ldM + m];
    z[(size_t)(ws + i) * ldM + m] -= x2 + x0 * d0 + x1 * d1;
  }
}

// ---------------------------------------------------------------- batched sparse LU
// One workgroup per scenario.  Workgroups are dispatched to the 8 XCDs round
// robin; the map below gives the workgroups of one XCD consecutive scenarios,
// so the scenario-major vectors they gather from share that XCD's L2 lines.

// Status bits of a pivot: 1 exactly zero (the static pivot sequence does not
// fit this matrix), 2 not finite (a non-finite operand).
// An entry of a read-only plan table at a wave-uniform address, through a
// scalar load (constant address space).  The vector-memory counter does not
// track scalar loads, so waiting for a step header never waits for the next
// step's prefetches still in flight.
__device__ __forceinline__ int ldu(const int *p) { return *(const __attribute__((address_space(4))) int *)p; }

template <typename TF>
__device__ __forceinline__ int pivot_flag(TF piv) {
  return piv == TF(0) ? 1 : (isfinite((double)piv) ? 0 : 2);
}

__device__ __forceinline__ int scenario_of_block(int b, int nb) {
  if (nb & 7) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// Compiler barrier plus wavefront-scope ordering: LDS operations of one
// wavefront execute in order, so this is all one lane needs to see the
// writes other lanes of its wavefront issued before.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// LDS-only workgroup barrier: orders the workgroup's LDS accesses across the
// barrier and leaves its global loads in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A's values into the LU storage, 8 entries per thread in flight (the
// index loads, then the value loads, then the stores).
// Matrix value s of scenario m: var in rows [s][ldM] (ldM > 0, the Schur
// kernels' layout) or in scenario-major rows [m][-ldM] (ldM < 0: the
// factorisation lanes' transposed snapshots, one contiguous range per
// scenario instead of one cache line per value).
__device__ __forceinline__ double var_at(const double *__restrict__ var, int s, int ldM, int m) {
  // one load at a selected index (two loads under a condition would be
  // branches, each waited for)
  const size_t i = ldM > 0 ? (size_t)s * ldM + m : (size_t)m * (size_t)(-ldM) + s;
  return var[i];
}

template <typename TF>
__device__ __forceinline__ void scatter_a(int nnz_a, const int *__restrict__ a_src, const int *__restrict__ a2,
                                          const double *__restrict__ var, const double *__restrict__ cst, int ldM,
                                          int m, TF *lu, int tid, int nt) {
  constexpr int U = 8;
  for (int e0 = 0; e0 < nnz_a; e0 += U * nt) {
    int src[U], dst[U];
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = min(e0 + tid + u * nt, nnz_a - 1);
      src[u] = a_src[e];
      dst[u] = a2[e];
    }
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = src[u] >= 0 ? var_at(var, src[u], ldM, m) : cst[-src[u] - 1];
#pragma unroll
    for (int u = 0; u < U; u++)
      if (e0 + tid + u * nt < nnz_a) lu[dst[u]] = TF(v[u]);
  }
}

// Copy of `bytes` (a multiple of 16, both ends 16-byte aligned) in 16-byte
// pieces, 8 per thread in flight.
__device__ __forceinline__ void copy16(void *dst, const void *__restrict__ src, size_t bytes, int tid, int nt) {
  const int n16 = (int)(bytes / 16);
  const int4 *s4 = reinterpret_cast<const int4 *>(src);
  int4 *d4 = reinterpret_cast<int4 *>(dst);
  for (int i0 = 0; i0 < n16; i0 += 8 * nt) {
    int4 t[8];
#pragma unroll
    for (int u = 0; u < 8; u++) t[u] = s4[min(i0 + tid + u * nt, n16 - 1)];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (i0 + tid + u * nt < n16) d4[i0 + tid + u * nt] = t[u];
  }
}

struct ColRec {
  int k, c0, len, dl, o0, nops, s0, s1;
};
__device__ __forceinline__ ColRec load_rec(const int *__restrict__ crec, int idx) {
  const int4 a = reinterpret_cast<const int4 *>(crec)[2 * idx];
  const int4 b = reinterpret_cast<const int4 *>(crec)[2 * idx + 1];
  return ColRec{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Thread mode: one thread runs the whole update program of one column on x
// (the column staged in LDS, or in place in global memory), 4 ops at a time
// with their operand loads issued together.  The L multipliers it reads
// belong to earlier columns and are final.
// Thread mode: one thread runs the whole update program of one column on x
// (the column staged in LDS, or in place in global memory), 4 ops at a time;
// the next batch's op words and multipliers are loaded while the current
// batch is applied.  The multipliers belong to earlier columns and are final;
// op indices are clamped to the column's last op (the op arrays carry one
// padding entry), so every load is unconditional.
template <typename TF, typename XP>
__device__ __forceinline__ void col_ops_thread(const SpluDev &S, const TF *__restrict__ lu, XP x, int o0, int o1) {
  if (o1 <= o0) return;
  const int ol = o1 - 1;
  int loc[4];
  TF lv[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int o = min(o0 + u, ol);
    loc[u] = S.op_loc[o];
    lv[u] = lu[S.op_sl[o]];
  }
  for (int o = o0; o < o1; o += 4) {
    int nloc[4], nsl[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int on = min(o + 4 + u, ol);
      nloc[u] = S.op_loc[on];
      nsl[u] = S.op_sl[on];
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (o + u < o1) x[loc[u] & 0xffff] -= lv[u] * x[loc[u] >> 16];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      loc[u] = nloc[u];
      lv[u] = lu[nsl[u]];
    }
  }
}

template <typename TF>
__device__ __forceinline__ int thread_column(const SpluDev &S, TF *__restrict__ lu, TF *__restrict__ so,
                                             const ColRec &r, TF *buf) {
  TF piv;
  if (r.len <= S.t_cols) {
    for (int t0 = 0; t0 < r.len; t0 += 4) {
      TF v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = lu[r.c0 + min(t0 + u, r.len - 1)];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (t0 + u < r.len) buf[t0 + u] = v[u];
    }
    col_ops_thread<TF>(S, lu, buf, r.o0, r.o0 + r.nops);
    piv = buf[r.dl];
    const TF inv = TF(1) / piv;
    for (int t = 0; t < r.len; t++) {
      TF v = buf[t];
      if (t > r.dl) {
        v *= inv;
        lu[r.c0 + t] = v;
      }
      const int d = S.lu2s[r.c0 + t];
      if (d >= 0) so[d] = v;
    }
  } else {
    TF *x = lu + r.c0;
    col_ops_thread<TF>(S, lu, x, r.o0, r.o0 + r.nops);
    piv = x[r.dl];
    const TF inv = TF(1) / piv;
    for (int t = 0; t < r.len; t++) {
      TF v = x[t];
      if (t > r.dl) {
        v *= inv;
        x[t] = v;
      }
      const int d = S.lu2s[r.c0 + t];
      if (d >= 0) so[d] = v;
    }
  }
  return pivot_flag(piv);
}

// Wavefront mode: the column, its op locations and every multiplier it needs
// are loaded at once into the wavefront's LDS buffers; then each step applies
// up to 64 independent updates (one per lane) with the next step's operands
// already in registers.  ext: the column belongs to the dense trailing block
// and only receives the updates from columns outside it (r.dl = number of
// its rows above the block, which are final afterwards).
template <typename TF>
__device__ __forceinline__ int wave_column(const SpluDev &S, TF *__restrict__ lu, TF *__restrict__ so, const ColRec &r,
                                           bool ext, TF *col, TF *lvs, int *locs, int lane) {
  const int len = r.len, nops = r.nops, c0 = r.c0, o0 = r.o0, nst = r.s1 - r.s0;
  TF piv = TF(1);
  if (len <= S.w_col && len <= 128 && nops <= S.w_ops && nops <= 512) {
    // every load of the column at once, at clamped (always valid) indices so
    // that none of them waits before the LDS stores below
    TF cv[2];
    int l2s[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int t = min(lane + WAVE * j, len - 1);
      cv[j] = lu[c0 + t];
      l2s[j] = S.lu2s[c0 + t];
    }
    int locr[8], slr[8];
    const int olast = max(nops - 1, 0);   // op arrays carry one padding entry at the end
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int t = min(lane + WAVE * j, olast);
      locr[j] = S.op_loc[o0 + t];
      slr[j] = S.op_sl[o0 + t];
    }
    const int sn0 = min(WAVE, nst);
    int bnd = S.st_o[r.s0 + min(lane, sn0)] - o0;
    TF lvr[8];
#pragma unroll
    for (int j = 0; j < 8; j++) lvr[j] = lu[slr[j]];
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int t = lane + WAVE * j;
      if (t < len) col[t] = cv[j];
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int t = lane + WAVE * j;
      if (t < nops) {
        locs[t] = locr[j];
        lvs[t] = lvr[j];
      }
    }
    wave_sync();
    for (int sb = 0; sb < nst; sb += WAVE) {
      const int sn = min(WAVE, nst - sb);
      if (sb) bnd = S.st_o[r.s0 + sb + min(lane, sn)] - o0;
      const int last = S.st_o[r.s0 + sb + sn] - o0;
      // step s covers op slots [bnd(s), bnd(s+1)); bnd(64) is `last`
      int t = __builtin_amdgcn_readlane(bnd, 0) + lane;
      bool v = t < __builtin_amdgcn_readlane(bnd, 1);
      int loc = v ? locs[t] : 0;
      TF lv = v ? lvs[t] : TF(0);
      for (int s = 0; s < sn; s++) {
        const bool vc = v;
        const int locc = loc;
        const TF lvc = lv;
        if (s + 1 < sn) {
          const int tb = __builtin_amdgcn_readlane(bnd, s + 1);
          const int te = s + 2 < WAVE ? __builtin_amdgcn_readlane(bnd, s + 2) : last;
          t = tb + lane;
          v = t < te;
          loc = v ? locs[t] : 0;
          lv = v ? lvs[t] : TF(0);
        }
        if (vc) col[locc & 0xffff] -= lvc * col[locc >> 16];
        wave_sync();
      }
    }
    if (!ext) {
      piv = col[r.dl];
      const TF inv = TF(1) / piv;
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const int t = lane + WAVE * j;
        if (t < len) {
          TF x = col[t];
          if (t > r.dl) {
            x *= inv;
            lu[c0 + t] = x;
          }
          if (l2s[j] >= 0) so[l2s[j]] = x;
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const int t = lane + WAVE * j;
        if (t < len) {
          const TF x = col[t];
          lu[c0 + t] = x;
          if (t < r.dl && l2s[j] >= 0) so[l2s[j]] = x;
        }
      }
    }
  } else {
    // in place in global memory, operands loaded per step
    TF *x = lu + c0;
    for (int s = r.s0; s < r.s1; s++) {
      const int t = S.st_o[s] + lane;
      if (t < S.st_o[s + 1]) {
        const int loc = S.op_loc[t];
        x[loc & 0xffff] -= lu[S.op_sl[t]] * x[loc >> 16];
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      __builtin_amdgcn_wave_barrier();
    }
    if (!ext) {
      piv = x[r.dl];
      const TF inv = TF(1) / piv;
      for (int t = lane; t < len; t += WAVE) {
        TF v = x[t];
        if (t > r.dl) {
          v *= inv;
          x[t] = v;
        }
        const int d = S.lu2s[c0 + t];
        if (d >= 0) so[d] = v;
      }
    } else {
      for (int t = lane; t < r.dl; t += WAVE) {
        const int d = S.lu2s[c0 + t];
        if (d >= 0) so[d] = x[t];
      }
    }
  }
  wave_sync();
  return pivot_flag(piv);
}

// Right-looking LU without pivoting of the nd x nd block D (column-major,
// LDS, nd <= 128) by a 512-thread workgroup, the block held in registers:
// thread (i = lane, w = wavefront) owns rows i and i + 64 of columns
// w, w + 8, ..., w + 120.  For pivot p the wavefront owning column p scales
// it into multipliers and publishes them (zero for rows <= p), lane p % 64
// of every wavefront publishes its part of row p (zero for columns <= p);
// after one barrier everyone applies a -= l * u to its columns right of p
// (a zero operand leaves an entry exactly as it is).  The multipliers and
// the row go through double buffers (lbuf: 2 x 128 values, ubuf: 2 x 8 x 16
// values), so one barrier per pivot suffices.  Same arithmetic as a scalar
// right-looking LU (l = a * (1/piv), a -= l * u).  L below the diagonal, U
// on and above it, written back to D.  Returns 1 if a pivot is zero or not
// finite (a non-finite operand also spreads through the zero products).
// Ends with a barrier.
// Pivot p in register column K of its owner wavefront: scale the column
// below the diagonal and publish the multipliers (zero for rows <= p).
template <typename TF, int K>
__device__ __forceinline__ void dense_pivot(TF (&a)[2][16], int p, int nd, int i0, TF *lb, int &bad) {
  const TF piv = __shfl(p < 64 ? a[0][K] : a[1][K], p & 63);
  bad |= pivot_flag(piv);
  const TF inv = TF(1) / piv;
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int i = i0 + 64 * r;
    const bool below = i > p && i < nd;
    if (below) a[r][K] *= inv;
    lb[i] = below ? a[r][K] : TF(0);
  }
}

template <typename TF>
__device__ __forceinline__ int dense_lu(TF *D, int nd, int tid, TF *lbuf, TF *ubuf) {
  constexpr int NR = 2, NC = 16;
  const int i0 = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  TF a[NR][NC];
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int k = 0; k < NC; k++) {
      const int i = i0 + 64 * r, j = w + 8 * k;
      a[r][k] = (i < nd && j < nd) ? D[i + j * nd] : TF(0);
    }
  int bad = 0;
  for (int p = 0; p < nd; p++) {
    TF *lb = lbuf + (p & 1) * 128;
    TF *ub = ubuf + (p & 1) * 128;
    if (w == (p & 7)) {
      switch (p >> 3) {
        case 0: dense_pivot<TF, 0>(a, p, nd, i0, lb, bad); break;
        case 1: dense_pivot<TF, 1>(a, p, nd, i0, lb, bad); break;
        case 2: dense_pivot<TF, 2>(a, p, nd, i0, lb, bad); break;
        case 3: dense_pivot<TF, 3>(a, p, nd, i0, lb, bad); break;
        case 4: dense_pivot<TF, 4>(a, p, nd, i0, lb, bad); break;
        case 5: dense_pivot<TF, 5>(a, p, nd, i0, lb, bad); break;
        case 6: dense_pivot<TF, 6>(a, p, nd, i0, lb, bad); break;
        case 7: dense_pivot<TF, 7>(a, p, nd, i0, lb, bad); break;
        case 8: dense_pivot<TF, 8>(a, p, nd, i0, lb, bad); break;
        case 9: dense_pivot<TF, 9>(a, p, nd, i0, lb, bad); break;
        case 10: dense_pivot<TF, 10>(a, p, nd, i0, lb, bad); break;
        case 11: dense_pivot<TF, 11>(a, p, nd, i0, lb, bad); break;
        case 12: dense_pivot<TF, 12>(a, p, nd, i0, lb, bad); break;
        case 13: dense_pivot<TF, 13>(a, p, nd, i0, lb, bad); break;
        case 14: dense_pivot<TF, 14>(a, p, nd, i0, lb, bad); break;
        default: dense_pivot<TF, 15>(a, p, nd, i0, lb, bad); break;
      }
    }
    if (i0 == (p & 63)) {
      // row p of this wavefront's columns right of p
#pragma unroll
      for (int k = 0; k < NC; k++) {
        const int j = w + 8 * k;
        ub[w * 16 + k] = (j > p && j < nd) ? (p < 64 ? a[0][k] : a[1][k]) : TF(0);
      }
    }
    __syncthreads();
    // a -= l * u everywhere: rows <= p have l = 0 and columns <= p have u = 0
    TF l[NR], u[NC];
#pragma unroll
    for (int r = 0; r < NR; r++) l[r] = lb[i0 + 64 * r];
#pragma unroll
    for (int k = 0; k < NC; k++) u[k] = ub[w * 16 + k];
#pragma unroll
    for (int k = 0; k < NC; k++)
#pragma unroll
      for (int r = 0; r < NR; r++) a[r][k] = a[r][k] - l[r] * u[k];
  }
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int k = 0; k < NC; k++) {
      const int i = i0 + 64 * r, j = w + 8 * k;
      if (i < nd && j < nd) D[i + j * nd] = a[r][k];
    }
  __syncthreads();
  return bad;
}

// The same LU for nd <= 64 on the first wavefront alone, with no workgroup
// barrier per pivot: lane i holds row i of the block in registers (NDC >= nd
// columns, the rest zero); the pivot row is read lane by lane (readlane),
// each lane below the pivot scales its entry in the pivot column and updates
// the columns right of it.  Same arithmetic as dense_lu (l = a * (1/piv),
// a -= l * u, rows on and above the pivot with l = 0).  Ends with a
// workgroup barrier.
template <typename TF>
__device__ __forceinline__ TF readlane_t(TF v, int lane) {
  if constexpr (sizeof(TF) == 8) {
    return readlane_d(v, lane);
  } else {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
  }
}
template <typename TF, int NDC>
__device__ int dense_lu_wave(TF *D, int nd, int tid) {
  int bad = 0;
  if (tid < WAVE) {
    const int i = tid;
    TF a[NDC];
#pragma unroll
    for (int j = 0; j < NDC; j++) a[j] = (i < nd && j < nd) ? D[i + j * nd] : TF(0);
#pragma unroll
    for (int p = 0; p < NDC; p++) {
      if (p < nd) {
        const TF piv = readlane_t(a[p], p);
        bad |= pivot_flag(piv);
        const TF inv = TF(1) / piv;
        const TF l = i > p ? a[p] * inv : TF(0);
        if (i > p) a[p] = l;
#pragma unroll
        for (int j = p + 1; j < NDC; j++) a[j] = a[j] - l * readlane_t(a[j], p);
      }
    }
#pragma unroll
    for (int j = 0; j < NDC; j++)
      if (i < nd && j < nd) D[i + j * nd] = a[j];
  }
  __syncthreads();
  return bad;
}
template <typename TF>
__device__ int dense_lu_any(TF *D, int nd, int tid, TF *lbuf) {
  if (nd <= 16) return dense_lu_wave<TF, 16>(D, nd, tid);
  if (nd <= 32) return dense_lu_wave<TF, 32>(D, nd, tid);
  if (nd <= 48) return dense_lu_wave<TF, 48>(D, nd, tid);
  // (64 FP64 values per lane would spill)
  if (sizeof(TF) == 4 && nd <= 64) return dense_lu_wave<TF, 64>(D, nd, tid);
  return dense_lu<TF>(D, nd, tid, lbuf, lbuf + 256);
}

// svv/dnv hold two buffers per scenario ([2][M][...]); tgt[m] selects the
// one written (NULL: buffer 0).
template <typename TF>
__global__ void __launch_bounds__(512) k_splu_factor(SpluDev S, const double *__restrict__ var,
                                                     const double *__restrict__ cst, int ldM, TF *__restrict__ luv,
                                                     TF *__restrict__ svv, TF *__restrict__ dnv,
                                                     const uint8_t *__restrict__ refac, const uint8_t *__restrict__ tgt,
                                                     int *__restrict__ status) {
  const int m = scenario_of_block(blockIdx.x, gridDim.x);
  if (!refac[m]) return;
  extern __shared__ __align__(16) char smem_raw[];
  const int Mb = gridDim.x;
  const int bsel = tgt ? tgt[m] : 0;
  TF *lu = luv + (size_t)m * S.nnz_lu;
  TF *so = svv + ((size_t)bsel * Mb + m) * S.ns;
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
  for (int p = tid; p < S.nnz_lu; p += nt) lu[p] = TF(0);
  __syncthreads();
  scatter_a<TF>(S.nnz_a, S.a_src, S.a2lu, var, cst, ldM, m, lu, tid, nt);
  __syncthreads();
  int bad = 0;
  for (int l = 0; l < S.n_flev; l++) {
    const int rb = S.flp[l], re = S.flp[l + 1], mode = S.fac_mode[l];
    if (mode == 0) {
      TF *buf = reinterpret_cast<TF *>(smem_raw) + (size_t)tid * S.t_cols;
      for (int idx = rb + tid; idx < re; idx += nt) bad |= thread_column<TF>(S, lu, so, load_rec(S.crec, idx), buf);
    } else {
      char *wb = smem_raw + (size_t)wave * S.w_bytes;
      TF *col = reinterpret_cast<TF *>(wb);
      TF *lvs = col + S.w_col;
      int *locs = reinterpret_cast<int *>(lvs + S.w_ops);
      for (int idx = rb + wave; idx < re; idx += nw)
        bad |= wave_column<TF>(S, lu, so, load_rec(S.crec, idx), mode == 2, col, lvs, locs, lane);
    }
    __syncthreads();
  }
  if (S.nd) {
    // dense trailing block: right-looking LU without pivoting, in LDS
    const int nd = S.nd;
    TF *D = reinterpret_cast<TF *>(smem_raw);
    for (int i = tid; i < nd * nd; i += nt) D[i] = TF(0);
    __syncthreads();
    for (int q0 = 0; q0 < S.nnz_top; q0 += 8 * nt) {
      int d[8];
      TF v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int q = min(q0 + tid + u * nt, S.nnz_top - 1);
        d[u] = S.dmap[q];
        v[u] = lu[S.qb + q];
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (q0 + tid + u * nt < S.nnz_top && d[u] >= 0) D[d[u]] = v[u];
    }
    __syncthreads();
    TF *lbuf = D + (((size_t)nd * nd + 3) & ~(size_t)3);
    bad |= dense_lu_any<TF>(D, nd, tid, lbuf);
    TF *dn = dnv + ((size_t)bsel * Mb + m) * nd * nd;
    for (int i = tid; i < nd * nd; i += nt) dn[i] = D[i];
  }
  if (bad) atomicOr(&status[m], bad);   // bit 1 zero pivot, bit 2 not finite
}

// ---- LDS-resident factorisation -----------------------------------------
// The multipliers (L entries of the columns below the dense block) stay in
// LDS (Lv) once their column is final, so the updates read no global
// memory.  The schedule is cut into steps; the static data of step s + 1
// (column records, update program, step bounds, output map) and its initial
// column values are loaded into registers while step s computes, and stored
// into the LDS staging area between two barriers.

// One column, a group of g lanes (j: lane within the group).  The column's
// updates come in rows of g (one step's ops, padded with no-ops on the spare
// slot x[len]); lane j applies entry j of each row, reading the next row's
// op word while the current row updates.  Every group of the wavefront runs
// the same number of rows R (the largest column of the wavefront in this
// round); rows past the column's own are read clamped and not applied.
// Then (sparse columns) the pivot and the scaling of L, whose entries also
// go to their multiplier slots Lv[lbase + t - dl - 1].
template <typename TF>
__device__ __forceinline__ int res_group_column(const int *rec, const int2 *ops, TF *x, TF *Lv, bool ext, int j,
                                                int g) {
  const int len = rec[1], dl = rec[2], nrow = rec[4], lbase = rec[6];
  const int R = __builtin_amdgcn_readfirstlane(rec[5]);
  ops += rec[3] * g + j;
  // straight-line rows: the operand reads first, then the next row's op
  // word (its read completes behind the update), and a write in every row
  // (rows past the column's own write the value back unchanged)
  int2 w = ops[0];
  for (int r = 0; r < R; r++) {
    const int dst = w.x & 0xffff;
    const TF xs = x[w.x >> 16];
    const TF d = x[dst];
    const TF l = Lv[w.y];
    w = ops[min(r + 1, nrow - 1) * g];
    const TF v = d - l * xs;
    x[dst] = r < nrow ? v : d;
  }
  if (ext) return 0;
  const TF piv = x[dl];
  const TF inv = TF(1) / piv;
  for (int t = dl + 1 + j; t < len; t += g) {
    const TF v = x[t] * inv;
    x[t] = v;
    Lv[lbase + t - dl - 1] = v;
  }
  return pivot_flag(piv);
}


// Same interface as k_splu_factor; luv holds the step-ordered values.
// Step s: its column values sit in xb[s & 1]; while it computes (LDS only),
// the blob and initial values of step s + 1 and the output map of step s
// are loaded into registers; after a barrier they go to the staging area and
// xb[(s + 1) & 1], and the finished values of step s are written out (solve
// array, dense copy), so that no wait for those stores ever holds up a step.
__device__ double g_dzero = 0.0;

template <typename TF>
__global__ void __launch_bounds__(512) k_splu_factor_res(SpluDev S, const double *__restrict__ var,
                                                         const double *__restrict__ cst, int ldM,
                                                         TF *__restrict__ luv, TF *__restrict__ svv,
                                                         TF *__restrict__ dnv, const uint8_t *__restrict__ refac,
                                                         const uint8_t *__restrict__ tgt, int *__restrict__ status) {
  const int m = scenario_of_block(blockIdx.x, gridDim.x);
  if (!refac[m]) return;
  extern __shared__ __align__(16) char smem_raw[];
  const int Mb = gridDim.x;
  const int bsel = tgt ? tgt[m] : 0;
  TF *so = svv + ((size_t)bsel * Mb + m) * S.ns;
  TF *dn = dnv + ((size_t)bsel * Mb + m) * (size_t)S.nd * S.nd;
  const int tid = threadIdx.x, nt = blockDim.x;
  // LDS: multipliers (later the dense block), staging blob, two value buffers
  TF *Lv = reinterpret_cast<TF *>(smem_raw);
  const size_t lvb = ((size_t)S.res_lv * sizeof(TF) + 15) & ~(size_t)15;
  int *stg = reinterpret_cast<int *>(smem_raw + lvb);
  constexpr int CX = res_cx(sizeof(TF));
  TF *xbuf = reinterpret_cast<TF *>(smem_raw + lvb + (size_t)kResCB * nt * 4);
  const int xcap = CX * nt;
  TF *lu = luv + (size_t)m * S.res_xtot;
  // the dense copy starts at zero: the external updates write only the
  // block's pattern entries, and the block is read back whole
  for (int i = tid; i < S.nd * S.nd; i += nt) dn[i] = TF(0);
  if (tid == 0) Lv[S.res_nl] = TF(0);   // multiplier of the no-op updates
  // the step-ordered initial values in one pass (res_xsrc: the source of
  // each slot, zero for fill), groups of four slots per thread, four groups
  // in flight.  Plans without constants (J, J^T) pass cst NULL: their loads
  // of a constant read a zero of this file instead
  const double *cbase = cst ? cst : &g_dzero;
  for (int t0 = 4 * tid; t0 < S.res_xtot; t0 += 16 * nt) {
    int4 sv[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      sv[u] = *reinterpret_cast<const int4 *>(S.res_xsrc + min(t0 + 4 * u * nt, (S.res_xtot - 1) & ~3));
    // all 32 loads (matrix value and constant, at clamped indices) are
    // issued before any is used, and both enter every value through 0/1
    // weights: a value chosen by a branch would let the compiler sink each
    // load into its branch and wait for it there (a scenario with a
    // non-finite value is non-finite throughout anyway)
    int sx[16];
    double a[16], c[16];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      sx[4 * u] = sv[u].x;
      sx[4 * u + 1] = sv[u].y;
      sx[4 * u + 2] = sv[u].z;
      sx[4 * u + 3] = sv[u].w;
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int s = sx[q];
      a[q] = var_at(var, max(s, 0), ldM, m);
      c[q] = cbase[(s < 0 && s != kResNoSrc) ? -1 - s : 0];
    }
    TF v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const double wa = sx[q] >= 0 ? 1.0 : 0.0, wc = (sx[q] < 0 && sx[q] != kResNoSrc) ? 1.0 : 0.0;
      v[q] = (TF)fma(a[q], wa, c[q] * wc);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int t = t0 + 4 * u * nt;
      if (t < S.res_xtot)
#pragma unroll
        for (int k = 0; k < 4; k++) lu[t + k] = v[4 * u + k];
    }
  }
  __syncthreads();
  // per thread: blob dwords 4 (tid + j nt) .. + 3 (j < kResCB / 4), value
  // and output-map slots 4 (tid + j nt) .. + 3 (j < CX / 4), at clamped
  // addresses (blobs and value ranges start at multiples of 4 and are
  // followed by at least 4 readable entries); the LDS stores are not masked
  // (slots past a step's length are never read)
  constexpr int PB = kResCB / 4, PX = CX / 4;
  int pb[kResCB];
  TF px[CX];
  int po[CX];
  const int nst = S.res_nsteps;
  auto fetch_blob = [&](int st) {
    const int *h = S.res_hdr + kResHdr * st;
    const int bo = ldu(h + 2), bl = ldu(h + 3), xo = ldu(h + 4), xl = ldu(h + 5);
#pragma unroll
    for (int j = 0; j < PB; j++) {
      const int4 v = *reinterpret_cast<const int4 *>(S.res_blob + bo + min(4 * (tid + j * nt), bl - 4));
      pb[4 * j] = v.x;
      pb[4 * j + 1] = v.y;
      pb[4 * j + 2] = v.z;
      pb[4 * j + 3] = v.w;
    }
#pragma unroll
    for (int j = 0; j < PX; j++) {
      const TF *q = lu + xo + min(4 * (tid + j * nt), xl - 4);
#pragma unroll
      for (int u = 0; u < 4; u++) px[4 * j + u] = q[u];
    }
  };
  auto fetch_omap = [&](int st) {
    const int *h = S.res_hdr + kResHdr * st;
    const int xo = ldu(h + 4), xl = ldu(h + 5);
#pragma unroll
    for (int j = 0; j < PX; j++) {
      const int4 v = *reinterpret_cast<const int4 *>(S.res_omap + xo + min(4 * (tid + j * nt), xl - 4));
      po[4 * j] = v.x;
      po[4 * j + 1] = v.y;
      po[4 * j + 2] = v.z;
      po[4 * j + 3] = v.w;
    }
  };
  auto store = [&](int st) {
    TF *xd = xbuf + (st & 1) * xcap;
#pragma unroll
    for (int j = 0; j < PB; j++)
      *reinterpret_cast<int4 *>(stg + 4 * (tid + j * nt)) =
          make_int4(pb[4 * j], pb[4 * j + 1], pb[4 * j + 2], pb[4 * j + 3]);
#pragma unroll
    for (int j = 0; j < PX; j++)
#pragma unroll
      for (int u = 0; u < 4; u++) xd[4 * (tid + j * nt) + u] = px[4 * j + u];
  };
  auto flush = [&](int st) {
    const int xl = ldu(S.res_hdr + kResHdr * st + 5);
    const TF *xs = xbuf + (st & 1) * xcap;
    // every load of this step has landed by now (the staging stores waited
    // for the younger ones); waiting here, before the first store, keeps
    // the output stores from being waited for until the next step
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt 0, expcnt 7, lgkmcnt 15
#pragma unroll
    for (int j = 0; j < PX; j++)
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int t = 4 * (tid + j * nt) + u, o = po[4 * j + u];
        if (t < xl && o != -1) {
          if (o >= 0)
            so[o] = xs[t];
          else
            dn[-2 - o] = xs[t];
        }
      }
  };
  fetch_blob(0);
  store(0);
  __syncthreads();
  int bad = 0;
  for (int st = 0; st < nst; st++) {
    const int *h = S.res_hdr + kResHdr * st;
    const int mode = ldu(h), nc = ldu(h + 1);
    const int2 *ops = reinterpret_cast<const int2 *>(stg + ldu(h + 6));
    TF *xb = xbuf + (st & 1) * xcap;
    fetch_omap(st);
    if (st + 1 < nst) fetch_blob(st + 1);
    {
      const int g = ldu(h + 9), G = nt / g;
      for (int c = tid / g; c < nc; c += G) {
        const int *rec = stg + 8 * c;
        bad |= res_group_column<TF>(rec, ops, xb + rec[0], Lv, mode == 2, tid & (g - 1), g);
      }
    }
    lds_barrier();
    if (st + 1 < nst) store(st + 1);
    flush(st);
    lds_barrier();
  }
  if (S.nd) {
    // dense trailing block (over the multipliers, no longer needed): the
    // block columns' values as the external updates left them in dn
    __syncthreads();
    const int nd = S.nd;
    TF *D = Lv;
    copy16(D, dn, (size_t)nd * nd * sizeof(TF), tid, nt);
    __syncthreads();
    TF *lbuf = reinterpret_cast<TF *>(stg);   // the staging area is free now
    bad |= dense_lu_any<TF>(D, nd, tid, lbuf);
    for (int i = tid; i < nd * nd; i += nt) dn[i] = D[i];
  }
  if (bad) atomicOr(&status[m], bad);   // bit 1 zero pivot, bit 2 not finite
}

// DPP lane exchange of a double within a row of 16 lanes
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over aligned groups of g lanes (g uniform, a power of two <= 64);
// every lane of a group ends with the group's sum.  Lanes 1, 2 apart swap by
// quad permutes, 4 and 8 apart by the half-row and row mirrors (after the
// previous steps each half holds one partial sum, so a mirror pairs halves).
__device__ __forceinline__ double group_sum(double acc, int g) {
  if (g > 1) acc += dpp_d<0xB1>(acc);    // quad_perm [1,0,3,2]
  if (g > 2) acc += dpp_d<0x4E>(acc);    // quad_perm [2,3,0,1]
  if (g > 4) acc += dpp_d<0x141>(acc);   // row_half_mirror
  if (g > 8) acc += dpp_d<0x140>(acc);   // row_mirror
  if (g > 16) acc += __shfl_xor(acc, 16);
  if (g > 32) acc += __shfl_xor(acc, 32);
  return acc;
}

// ---- staged triangular solve ----------------------------------------------
// The levels of each side are grouped into steps (host plan); the static
// data of step s + 1 (level and chunk tables, value columns, lane rows) and
// its values (and diagonals) are loaded into registers while step s
// computes from LDS, and stored into the staging area between two barriers.
// Per thread: kSolCB blob dwords, sol_cv(sizeof(TF)) values, kSolCD diagonals.
struct SolStage {
  int b[kSolCB];
};

// Loads of step st of a side into registers (clamped, always valid).
template <typename TF, int CV = sol_cv(sizeof(TF))>
__device__ __forceinline__ void sol_fetch(const int *__restrict__ hdr, const int *__restrict__ blob, int st,
                                          const TF *__restrict__ vals, const TF *__restrict__ dg, int tid, int nt,
                                          SolStage &P, TF (&pv)[CV], TF (&pd)[kSolCD]) {
  const int *h = hdr + kSolHdr * st;
  const int bo = ldu(h + 1), bl = ldu(h + 2), vo = ldu(h + 3), vl = ldu(h + 4), dgo = ldu(h + 5), dl = ldu(h + 6);
#pragma unroll
  for (int j = 0; j < kSolCB / 4; j++) {
    const int4 v = *reinterpret_cast<const int4 *>(blob + bo + min(4 * (tid + j * nt), bl - 4));
    P.b[4 * j] = v.x;
    P.b[4 * j + 1] = v.y;
    P.b[4 * j + 2] = v.z;
    P.b[4 * j + 3] = v.w;
  }
  const int vend = ldu(h + 10);   // values actually present (the rest of vl is padding)
#pragma unroll
  for (int j = 0; j < CV; j++) pv[j] = vals[vo + min(tid + j * nt, max(vend - 1, 0))];
  const int dend = ldu(h + 11);
  // loaded either way and used as loaded: a branch around the load, or a
  // select on its value here, would wait for it before the step computes.
  // A step without diagonals (dl = 0: the forward side, whose dg_off counts
  // its lanes, not diagonals) loads its own values again instead (the
  // addresses of the loads above) and never reads them.
  const TF *dgp = dl ? dg + dgo : vals + vo;
  const int dlast = dl ? max(dend - 1, 0) : max(vend - 1, 0);
#pragma unroll
  for (int j = 0; j < kSolCD; j++) pd[j] = dgp[min(tid + j * nt, dlast)];
  (void)vl;
}

template <typename TF, int CV = sol_cv(sizeof(TF))>
__device__ __forceinline__ void sol_store(int *sb, TF *svl, TF *sdg, int tid, int nt, const SolStage &P,
                                          const TF (&pv)[CV], const TF (&pd)[kSolCD]) {
#pragma unroll
  for (int j = 0; j < kSolCB / 4; j++)
    *reinterpret_cast<int4 *>(sb + 4 * (tid + j * nt)) = make_int4(P.b[4 * j], P.b[4 * j + 1], P.b[4 * j + 2],
                                                                   P.b[4 * j + 3]);
#pragma unroll
  for (int j = 0; j < CV; j++) svl[tid + j * nt] = pv[j];
#pragma unroll
  for (int j = 0; j < kSolCD; j++) sdg[tid + j * nt] = pd[j];
}

// The passes of one staged step.  A pass record gives each wavefront one
// chunk:
// lane i (i < nl) sums vals[E + i + 64 j] * v[col] over j < pf, the g lanes
// of a row add up, and the row's first lane writes v[row] - sum (divided by
// the diagonal where the level's flags have bit 1).  The row's value is
// read together with the entries.  Unpartitioned, v is x throughout.
// Partitioned, the residual vector r (x forward, y backward) and the final
// one f (the other): an A-level (flag bit 0) reads r for both and writes
// f[row], a B-level reads f at the columns and updates r[row].  A barrier
// follows the last pass of each level.
template <typename TF, bool BW>
__device__ __forceinline__ void sol_step(const int *h, const int *sb, const TF *svl,
                                         const TF *sdg, double *x, double *y, int tid) {
  const int np = ldu(h);
  const uint16_t *cols = reinterpret_cast<const uint16_t *>(sb + ldu(h + 8));
  const int *rows = sb + ldu(h + 9);
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  // pass records from the staged blob (LDS: in-order, unlike scalar loads,
  // so waiting for them never waits for anything else); the next pass's
  // records are read while the current pass runs
  const int4 *L = reinterpret_cast<const int4 *>(sb);
  int4 hd = L[0], rc = L[1 + w];
  for (int k = 0; k < np; k++) {
    const int4 hdc = hd, rcc = rc;
    if (k + 1 < np) {
      hd = L[9 * (k + 1)];
      rc = L[9 * (k + 1) + 1 + w];
    }
    const int g = __builtin_amdgcn_readfirstlane(hdc.x), E = __builtin_amdgcn_readfirstlane(rcc.x);
    const int fl = __builtin_amdgcn_readfirstlane(hdc.z);
    const int pf = __builtin_amdgcn_readfirstlane(rcc.y), rs = __builtin_amdgcn_readfirstlane(rcc.z);
    const int nl = __builtin_amdgcn_readfirstlane(rcc.w);
    if (nl > 0) {
      double *rv = BW ? y : x, *fv = BW ? x : y;   // the same vector unpartitioned
      const double *cv = (fl & 1) ? rv : fv;
      double *ov = (fl & 1) ? fv : rv;
      const int uc = min(lane, nl - 1);
      const int base = E + uc;
      const int row = rows[rs + uc];
      double acc = 0.0;
      double xr = 0.0;
      for (int j0 = 0; j0 < pf; j0 += 4) {
        int cc[4];
        TF vv[4];
        double xv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) cc[q] = cols[base + WAVE * min(j0 + q, pf - 1)];
#pragma unroll
        for (int q = 0; q < 4; q++) vv[q] = svl[base + WAVE * min(j0 + q, pf - 1)];
#pragma unroll
        for (int q = 0; q < 4; q++) xv[q] = cv[cc[q]];
        if (j0 == 0) xr = rv[row];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const double t = (double)vv[q] * xv[q];
          acc += j0 + q < pf ? t : 0.0;
        }
      }
      acc = group_sum(acc, g);
      if (lane < nl && (lane & (g - 1)) == 0) {
        if (fl & 2)
          ov[row] = (xr - acc) / (double)sdg[rs + uc];
        else
          ov[row] = xr - acc;
      }
    }
    if (__builtin_amdgcn_readfirstlane(hdc.y)) lds_barrier();
  }
}

// Composite values of the partitioned solves after a factorisation (one
// workgroup per scenario whose refac flag is set, the buffer the
// factorisation wrote).  The blocks run in order (a block's entries are of
// one depth in their partitions, so it reads only earlier blocks'
// composites): every term of the block is computed into LDS, four per
// thread with their loads in flight together, then one thread per entry
// adds its terms in their fixed order (the result does not depend on
// timing) and stores the sum.  A term reads raw entries, diagonals and
// composites of the same buffer.
template <typename TF>
__global__ void __launch_bounds__(512) k_splu_compose(SpluDev S, TF *__restrict__ svv,
                                                      const uint8_t *__restrict__ refac,
                                                      const uint8_t *__restrict__ tgt) {
  const int m = scenario_of_block(blockIdx.x, gridDim.x);
  if (!refac[m]) return;
  __shared__ double tv[kCmpBlock];
  const int bsel = tgt ? tgt[m] : 0;
  TF *v = svv + ((size_t)bsel * gridDim.x + m) * S.ns;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int p = 0; p < S.cmp_nph; p++) {
    const int e0 = S.cmp_ph[p], e1 = S.cmp_ph[p + 1];
    const int t0 = S.cmp_tp[e0], t1 = S.cmp_tp[e1];
    for (int t = t0 + tid; t < t1; t += 4 * nt) {
      int4 q[4];
      TF va[4], vb[4], vd[4];
#pragma unroll
      for (int u = 0; u < 4; u++) q[u] = *reinterpret_cast<const int4 *>(S.cmp_ops + 4 * min(t + u * nt, t1 - 1));
#pragma unroll
      for (int u = 0; u < 4; u++) {
        va[u] = v[q[u].y];
        vb[u] = v[max(q[u].z, 0)];
        vd[u] = v[max(q[u].w, 0)];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (t + u * nt < t1) {
          const double a = q[u].w >= 0 ? (double)va[u] / (double)vd[u] : (double)va[u];
          tv[t + u * nt - t0] = q[u].z >= 0 ? -a * (double)vb[u] : a;
        }
    }
    __syncthreads();
    for (int e = e0 + tid; e < e1; e += nt) {
      double acc = 0.0;
      for (int t = S.cmp_tp[e]; t < S.cmp_tp[e + 1]; t++) acc += tv[t - t0];
      v[S.cmp_dst[e]] = (TF)acc;
    }
    __syncthreads();
  }
}

// The solve's outputs from x (LDS): xout[q[k]] = x[k] and zupd[zr] +=
// sign * x[k], zr = scatter[q[k]] (SC) or q[k].
template <bool SC, bool ZU, bool XO>
__device__ __forceinline__ void sol_output(const SpluDev &S, const double *x, const int *__restrict__ scatter,
                                           double *__restrict__ zupd, double *__restrict__ xout, int ldM, int m,
                                           double sign, int tid, int nt) {
  const int n = S.n;
  for (int k0 = 0; k0 < n; k0 += 16 * nt) {
    int j[16], zr[16];
    double v[16], z[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int k = min(k0 + tid + u * nt, n - 1);
      j[u] = S.q[k];
      v[u] = x[k];
    }
#pragma unroll
    for (int u = 0; u < 16; u++) zr[u] = SC ? scatter[j[u]] : j[u];
    if (ZU) {
#pragma unroll
      for (int u = 0; u < 16; u++) z[u] = zupd[(size_t)zr[u] * ldM + m];
    }
#pragma unroll
    for (int u = 0; u < 16; u++)
      if (k0 + tid + u * nt < n) {
        if (XO) xout[(size_t)j[u] * ldM + m] = v[u];
        if (ZU) zupd[(size_t)zr[u] * ldM + m] = z[u] + sign * v[u];
      }
  }
}

// One workgroup per scenario: x = A^{-1} b with b gathered from rhs (row
// space), the solution kept in LDS: forward steps over the sparse part and
// the block rows' external entries, the dense block's forward and backward
// substitution (one wavefront, x of the block in registers), then the
// backward steps; then zupd[scatter[j]] += sign * x[j] and xout[j] = x[j].
// svv/dnv hold two buffers per scenario; cur[m] selects the one read (NULL: 0).
template <typename TF>
__global__ void __launch_bounds__(512) k_splu_solve(SpluDev S, const TF *__restrict__ svv, const TF *__restrict__ dnv,
                                                    const double *__restrict__ rhs, int ldM, double sign,
                                                    const int *__restrict__ scatter, double *__restrict__ zupd,
                                                    double *__restrict__ xout, const uint8_t *__restrict__ work,
                                                    const uint8_t *__restrict__ cur, int shared, WbTab wb) {
  const int m = scenario_of_block(blockIdx.x, gridDim.x);
  if (!work[m]) return;
  extern __shared__ __align__(16) char smem_raw[];
  const int n = S.n, nd = S.nd, B = n - nd;
  const int Mb = gridDim.x;
  const int bsel = cur ? cur[m] : 0;
  const int fm = shared ? 0 : m;   // shared: one factorisation (slot 0) serves every scenario
  const int tid = threadIdx.x, nt = blockDim.x;
  // LDS: x (and, partitioned, y), then the staging area (blob, values,
  // diagonals), which the dense block reuses between the two sides
  const size_t xb = ((size_t)n * 8 + 15) & ~(size_t)15;
  double *x = reinterpret_cast<double *>(smem_raw);
  double *y = S.part ? reinterpret_cast<double *>(smem_raw + xb) : x;
  char *area = smem_raw + (S.part ? 2 : 1) * xb;
  int *sb = reinterpret_cast<int *>(area);
  TF *svl = reinterpret_cast<TF *>(area + (size_t)kSolCB * nt * 4);
  TF *sdg = svl + (size_t)sol_cv(sizeof(TF)) * nt;
  TF *Dl = reinterpret_cast<TF *>(area);
  const TF *fv = svv + ((size_t)bsel * Mb + fm) * S.ns;
  const TF *bv = fv + S.fw_nvals;
  const TF *bd = bv + S.bw_nvals;
  const int nfs = S.sol_nfw, nbs = S.sol_nbw;
  SolStage P;
  TF pv[sol_cv(sizeof(TF))], pd[kSolCD];
  if (nfs) sol_fetch<TF>(S.sol_fw_hdr, S.sol_fw_blob, 0, fv, bd, tid, nt, P, pv, pd);
  {
    // gather b, sixteen rows per thread in flight (one round up to 8k rows)
    for (int r0 = 0; r0 < n; r0 += 16 * nt) {
      double b8[16];
      int p8[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int r = min(r0 + tid + u * nt, n - 1);
        b8[u] = rhs[(size_t)r * ldM + m];
        p8[u] = S.pinv[r];
      }
#pragma unroll
      for (int u = 0; u < 16; u++)
        if (r0 + tid + u * nt < n) x[p8[u]] = b8[u];
    }
  }
  if (nfs) sol_store<TF>(sb, svl, sdg, tid, nt, P, pv, pd);
  __syncthreads();
  for (int st = 0; st < nfs; st++) {
    if (st + 1 < nfs) sol_fetch<TF>(S.sol_fw_hdr, S.sol_fw_blob, st + 1, fv, bd, tid, nt, P, pv, pd);
    sol_step<TF, false>(S.sol_fw_hdr + kSolHdr * st, sb, svl, sdg, x, y, tid);
    if (st + 1 < nfs) {
      sol_store<TF>(sb, svl, sdg, tid, nt, P, pv, pd);
      lds_barrier();
    }
  }
  if (nbs) sol_fetch<TF>(S.sol_bw_hdr, S.sol_bw_blob, 0, bv, bd, tid, nt, P, pv, pd);
  if (nd) {
    const TF *dn = dnv + ((size_t)bsel * Mb + fm) * nd * nd;
    copy16(Dl, dn, (size_t)nd * nd * sizeof(TF), tid, nt);
    __syncthreads();
  }
  if (nd && tid < WAVE) {
    // dense block, one wavefront: lane i holds x[B + i] and x[B + 64 + i];
    // the block's values for 8 pivots are read ahead of their use
    const int i0 = tid, i1 = tid + WAVE;
    double x0 = i0 < nd ? x[B + i0] : 0.0, x1 = i1 < nd ? x[B + i1] : 0.0;
    for (int p0 = 0; p0 < nd; p0 += 8) {
      TF a0[8], a1[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int p = min(p0 + u, nd - 1);
        a0[u] = Dl[min(i0, nd - 1) + p * nd];
        a1[u] = Dl[min(i1, nd - 1) + p * nd];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int p = p0 + u;
        if (p < nd) {
          const double xp = p < WAVE ? readlane_d(x0, p) : readlane_d(x1, p - WAVE);
          if (i0 > p && i0 < nd) x0 -= (double)a0[u] * xp;
          if (i1 > p && i1 < nd) x1 -= (double)a1[u] * xp;
        }
      }
    }
    for (int p0 = nd - 1; p0 >= 0; p0 -= 8) {
      TF a0[8], a1[8], dgv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int p = max(p0 - u, 0);
        a0[u] = Dl[min(i0, nd - 1) + p * nd];
        a1[u] = Dl[min(i1, nd - 1) + p * nd];
        dgv[u] = Dl[p + p * nd];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int p = p0 - u;
        if (p >= 0) {
          const double xp = (p < WAVE ? readlane_d(x0, p) : readlane_d(x1, p - WAVE)) / (double)dgv[u];
          if (i0 == p) x0 = xp;
          if (i1 == p) x1 = xp;
          if (i0 < p) x0 -= (double)a0[u] * xp;
          if (i1 < p && i1 < nd) x1 -= (double)a1[u] * xp;
        }
      }
    }
    if (i0 < nd) x[B + i0] = x0;
    if (i1 < nd) x[B + i1] = x1;
  }
  __syncthreads();
  if (nbs) {
    sol_store<TF>(sb, svl, sdg, tid, nt, P, pv, pd);
    __syncthreads();
  }
  for (int st = 0; st < nbs; st++) {
    if (st + 1 < nbs) sol_fetch<TF>(S.sol_bw_hdr, S.sol_bw_blob, st + 1, bv, bd, tid, nt, P, pv, pd);
    sol_step<TF, true>(S.sol_bw_hdr + kSolHdr * st, sb, svl, sdg, x, y, tid);
    if (st + 1 < nbs) {
      sol_store<TF>(sb, svl, sdg, tid, nt, P, pv, pd);
      lds_barrier();
    }
  }
  // Woodbury correction of the shared factorisation (uniform per workgroup):
  // s = H^{-1} x[pos] by the first nc threads into the free staging area,
  // then x -= W s, the columns of W read contiguously
  const int wnc = wb.nc ? wb.nc[m] : 0;
  if (wnc > 0) {
    double *ws = reinterpret_cast<double *>(area);
    const int *wcol = wb.col + (size_t)m * wb.kw;
    __syncthreads();
    if (tid < wnc) {
      const double *h = wb.h + ((size_t)m * wb.kw + tid) * wb.kw;
      const int *wpos = wb.pos + (size_t)m * wb.kw;
      double acc = 0.0;
      for (int c = 0; c < wnc; c++) acc = fma(h[c], x[wpos[c]], acc);
      ws[tid] = acc;
    }
    __syncthreads();
    for (int k = tid; k < n; k += nt) {
      double acc = 0.0;
      for (int c = 0; c < wnc; c++) acc = fma(wb.W[(size_t)wcol[c] * n + k], ws[c], acc);
      x[k] -= acc;
    }
    __syncthreads();
  }
  // outputs, sixteen rows per thread in flight (the variants without a
  // scatter table or a state update carry no conditional loads, which
  // would be waited for one by one)
  switch ((scatter ? 4 : 0) | (zupd ? 2 : 0) | (xout ? 1 : 0)) {
    case 7: sol_output<true, true, true>(S, x, scatter, zupd, xout, ldM, m, sign, tid, nt); break;
    case 6: sol_output<true, true, false>(S, x, scatter, zupd, xout, ldM, m, sign, tid, nt); break;
    case 5: sol_output<true, false, true>(S, x, scatter, zupd, xout, ldM, m, sign, tid, nt); break;
    case 3: sol_output<false, true, true>(S, x, scatter, zupd, xout, ldM, m, sign, tid, nt); break;
    case 2: sol_output<false, true, false>(S, x, scatter, zupd, xout, ldM, m, sign, tid, nt); break;
    case 1: sol_output<false, false, true>(S, x, scatter, zupd, xout, ldM, m, sign, tid, nt); break;
    default: break;   // nothing to write
  }
}

// Systems whose solution does not fit in LDS (tens of thousands of rows and
// more): the same two triangular solves with x in global memory (S.gx_ws,
// scenario m's x at j*ldM + m, pivot order) and the level tables read where
// they lie (sliced ELL: lane l of a level holds the entries E + u + 64 j of
// its chunk, u its lane in the chunk, the g lanes of a row add up, the row's
// first lane updates x[row]).  The wide levels near the leaves (the first
// forward and the last backward levels, S.gx_fw_wide and S.gx_bw_wide of
// them) run as launches of their own across the whole device; the narrow
// levels near the root and the dense block run in one workgroup per
// scenario, where a workgroup barrier (which orders the workgroup's global
// memory too) closes each level.  The gather of b and the scatter of the
// result are launches of their own, lane per scenario.
__device__ __forceinline__ int gx_lane_row(int l, int T, int c0, const int *__restrict__ chunk, int &E, int &pf,
                                           int &uc) {
  const int lc = min(l, T - 1);   // spare lanes of the last chunk read lane T - 1's entries
  const int c = c0 + (lc >> 6);
  uc = lc & 63;
  E = chunk[2 * c];
  pf = chunk[2 * c + 1];
  return lc;
}

// sum over j < pf of vals[e] x[cols[e]] (e = e0 + 64 j), U entries in
// flight (their column indices, then their x values, are loaded together)
template <typename TF, int U>
__device__ __forceinline__ double gx_row_sum(int e0, int pf, const int *__restrict__ cols, const TF *__restrict__ vals,
                                             const double *x, int ldM) {
  double acc = 0.0;
  int j = 0;
  for (; j + U <= pf; j += U) {
    int cc[U];
    TF vv[U];
    double xv[U];
#pragma unroll
    for (int q = 0; q < U; q++) {
      const int e = e0 + WAVE * (j + q);
      cc[q] = cols[e];
      vv[q] = vals[e];
    }
#pragma unroll
    for (int q = 0; q < U; q++) xv[q] = x[(size_t)cc[q] * ldM];
#pragma unroll
    for (int q = 0; q < U; q++) acc += (double)vv[q] * xv[q];
  }
  for (; j < pf; j++) {
    const int e = e0 + WAVE * j;
    acc += (double)vals[e] * x[(size_t)cols[e] * ldM];
  }
  return acc;
}

template <typename TF, bool BW>
__device__ __forceinline__ void gx_levels(int l0, int l1, const int *__restrict__ lvl, const int *__restrict__ chunk,
                                          const int *__restrict__ cols, const int *__restrict__ rows,
                                          const TF *__restrict__ vals, const TF *__restrict__ diag, double *x,
                                          int ldM, int tid, int nt) {
  for (int lv = l0; lv < l1; lv++) {
    const int T = lvl[4 * lv], g = lvl[4 * lv + 1], R = lvl[4 * lv + 2], c0 = lvl[4 * lv + 3];
    const int Tp = (T + 63) & ~63;
    for (int l = tid; l < Tp; l += nt) {
      int E, pf, uc;
      const int lc = gx_lane_row(l, T, c0, chunk, E, pf, uc);
      const double acc = group_sum(gx_row_sum<TF, 8>(E + uc, pf, cols, vals, x, ldM), g);
      if (l < T && (uc & (g - 1)) == 0) {
        const int row = rows[R + lc];
        const double xr = x[(size_t)row * ldM];
        x[(size_t)row * ldM] = BW ? (xr - acc) / (double)diag[R + lc] : xr - acc;
      }
    }
    __syncthreads();
  }
}

// One wide level lv of one side: block (b, m) takes lanes 256 b .. of the
// level for scenario m (whole wavefronts: the level's lane count is padded
// to 64).  vals / diag at off_vals / off_diag of scenario m's solve array.
template <typename TF, bool BW>
__global__ void __launch_bounds__(256) k_gx_level(const int *__restrict__ lvl, int lv, const int *__restrict__ chunk,
                                                  const int *__restrict__ cols, const int *__restrict__ rows,
                                                  const TF *__restrict__ svv, int ns, int off_vals, int off_diag,
                                                  int Mb, double *__restrict__ xw, int ldM,
                                                  const uint8_t *__restrict__ work, const uint8_t *__restrict__ cur) {
  const int m = blockIdx.y;
  if (!work[m]) return;
  const int T = lvl[4 * lv], g = lvl[4 * lv + 1], R = lvl[4 * lv + 2], c0 = lvl[4 * lv + 3];
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= ((T + 63) & ~63)) return;   // whole wavefronts
  const int bsel = cur ? cur[m] : 0;
  const TF *base = svv + ((size_t)bsel * Mb + m) * ns;
  const TF *vals = base + off_vals;
  double *x = xw + m;
  int E, pf, uc;
  const int lc = gx_lane_row(l, T, c0, chunk, E, pf, uc);
  const double acc = group_sum(gx_row_sum<TF, 4>(E + uc, pf, cols, vals, x, ldM), g);
  if (l < T && (uc & (g - 1)) == 0) {
    const int row = rows[R + lc];
    const double xr = x[(size_t)row * ldM];
    x[(size_t)row * ldM] = BW ? (xr - acc) / (double)base[off_diag + R + lc] : xr - acc;
  }
}

// b into x (pivot order), and the result out of x, lane per scenario
__global__ void k_gx_gather(int n, int ldM, const int *__restrict__ pinv, const double *__restrict__ rhs,
                            double *__restrict__ x, const uint8_t *__restrict__ work) {
  const int m = blockIdx.y * WAVE + threadIdx.x;
  if (m >= ldM || !work[m]) return;
  for (int r = blockIdx.x; r < n; r += gridDim.x) x[(size_t)pinv[r] * ldM + m] = rhs[(size_t)r * ldM + m];
}
__global__ void k_gx_output(int n, int ldM, const int *__restrict__ q, const int *__restrict__ scatter,
                            const double *__restrict__ x, double sign, double *__restrict__ zupd,
                            double *__restrict__ xout, const uint8_t *__restrict__ work) {
  const int m = blockIdx.y * WAVE + threadIdx.x;
  if (m >= ldM || !work[m]) return;
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    const int j = q[k];
    const double v = x[(size_t)k * ldM + m];
    if (xout) xout[(size_t)j * ldM + m] = v;
    if (zupd) {
      const int zr = scatter ? scatter[j] : j;
      zupd[(size_t)zr * ldM + m] += sign * v;
    }
  }
}

// The narrow levels (forward from S.gx_fw_wide on, backward up to
// n_bwlev - S.gx_bw_wide) and the dense block, one workgroup per scenario.
template <typename TF>
__global__ void __launch_bounds__(512) k_splu_solve_gx(SpluDev S, const TF *__restrict__ svv,
                                                       const TF *__restrict__ dnv, const uint8_t *__restrict__ work,
                                                       const uint8_t *__restrict__ cur, int ldM) {
  const int m = scenario_of_block(blockIdx.x, gridDim.x);
  if (!work[m]) return;
  extern __shared__ __align__(16) char smem_raw[];
  const int n = S.n, nd = S.nd, B = n - nd;
  const int Mb = gridDim.x;
  const int bsel = cur ? cur[m] : 0;
  const int tid = threadIdx.x, nt = blockDim.x;
  double *x = S.gx_ws + m;
  const TF *fv = svv + ((size_t)bsel * Mb + m) * S.ns;
  const TF *bv = fv + S.fw_nvals;
  const TF *bd = bv + S.bw_nvals;
  gx_levels<TF, false>(S.gx_fw_wide, S.n_fwlev, S.fw_lvl, S.fw_chunk, S.fw_cols, S.fw_rows, fv, nullptr, x, ldM, tid,
                       nt);
  if (nd) {
    TF *Dl = reinterpret_cast<TF *>(smem_raw);
    const TF *dn = dnv + ((size_t)bsel * Mb + m) * nd * nd;
    for (int i = tid; i < nd * nd; i += nt) Dl[i] = dn[i];
    __syncthreads();
    if (tid < WAVE) {   // lane i holds x[B + i] and x[B + 64 + i]
      const int i0 = tid, i1 = tid + WAVE;
      double x0 = i0 < nd ? x[(size_t)(B + i0) * ldM] : 0.0, x1 = i1 < nd ? x[(size_t)(B + i1) * ldM] : 0.0;
      for (int p = 0; p < nd; p++) {
        const double xp = p < WAVE ? readlane_d(x0, p) : readlane_d(x1, p - WAVE);
        if (i0 > p && i0 < nd) x0 -= (double)Dl[min(i0, nd - 1) + p * nd] * xp;
        if (i1 > p && i1 < nd) x1 -= (double)Dl[min(i1, nd - 1) + p * nd] * xp;
      }
      for (int p = nd - 1; p >= 0; p--) {
        const double xp = (p < WAVE ? readlane_d(x0, p) : readlane_d(x1, p - WAVE)) / (double)Dl[p + p * nd];
        if (i0 == p) x0 = xp;
        if (i1 == p) x1 = xp;
        if (i0 < p) x0 -= (double)Dl[min(i0, nd - 1) + p * nd] * xp;
        if (i1 < p && i1 < nd) x1 -= (double)Dl[min(i1, nd - 1) + p * nd] * xp;
      }
      if (i0 < nd) x[(size_t)(B + i0) * ldM] = x0;
      if (i1 < nd) x[(size_t)(B + i1) * ldM] = x1;
    }
    __syncthreads();
  }
  gx_levels<TF, true>(0, S.n_bwlev - S.gx_bw_wide, S.bw_lvl, S.bw_chunk, S.bw_cols, S.bw_rows, bv, bd, x, ldM, tid,
                      nt);
}

// ---------------------------------------------------------------- discrete adjoint
// Backward step k of the adjoint of the backward-Euler sweep
//   F_k(z_k, z_{k-1}, p) = 0:  z - z_{k-1} - dt f(z, p) on differential rows,
//                              g(z, p) on algebraic rows,
// objective  J = sum_k w sum_r (z_k[r] - off)^2  (w = scale * dt):
//   (dF_k/dz_k)^T lam_k = dJ/dz_k + lam_{k+1} restricted to differential rows
//   dJ/dp -= lam_k^T dF_k/dp.
// Scenario m takes part in backward step k when k <= hz[m] (its horizon; -1
// for a scenario left out).
__global__ void k_adj_mask(int ldM, int k, const int *__restrict__ hz, uint8_t *__restrict__ act) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < ldM) act[m] = k <= hz[m] ? 1 : 0;
}

// rhs = lam_{k+1} on the differential rows, 0 on the algebraic rows
__global__ void k_adj_carry(int n, int ldM, const uint8_t *__restrict__ is_diff, const double *__restrict__ lam,
                            double *__restrict__ rhs, const uint8_t *__restrict__ act) {
  const int m = blockIdx.y * WAVE + threadIdx.x;
  if (!act[m]) return;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const size_t o = (size_t)i * ldM + m;
    rhs[o] = is_diff[i] ? lam[o] : 0.0;
  }
}

// Objective rows at z_k: rhs += 2 w (z - off), val += w sum (z - off)^2,
// rows in list order (one thread per scenario, a fixed summation order).
__global__ void k_adj_objective(int nr, const int *__restrict__ rows, int ldM, double w, double off,
                                const double *__restrict__ z, double *__restrict__ rhs, double *__restrict__ val,
                                const uint8_t *__restrict__ act) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= ldM || !act[m]) return;
  double s = 0.0;
  for (int r = 0; r < nr; r++) {
    const size_t o = (size_t)rows[r] * ldM + m;
    const double e = z[o] - off;
    s += e * e;
    rhs[o] += 2.0 * w * e;
  }
  val[m] += w * s;
}

// Parameter pullbacks.  Entry e = (kind, device row); g[e][m] -= lam . dF/dp
// at z_k, from the device equations above:
//   GENROU H  (P[6]):  F_w = w - w_old - dt N/(2H)            dF_w/dH  =  dt f_w / H
//   GENROU D  (P[7]):                                        dF_w/dD  =  dt w / (2H)
//   SEXS K    (P[2]):  f_1 = (-x1 + K(x0 + TA/TB err))/TE    dF_1/dK  = -dt (x0 + TA/TB err) / TE
//   ESDC1A KA (P[0]):  f_0 = (KA u - x0)/TA                  dF_0/dKA = -dt u / TA
//   TGOV1 R   (P[0]):  f_1 = ((pref - w)/R - x1)/T1          dF_1/dR  =  dt (pref - w) / (R^2 T1)
//   IEESGO K1 (P[6]):  f_0 = (K1 w - PF0)/T1                 dF_0/dK1 = -dt w / T1
//   IEEEST KS (P[12]): f_6 = (KS y4 - s6)/T6, g = vs - T5/T6 (KS y4 - s6)
//                      dF_6/dKS = -dt y4 / T6,  dg/dKS = -(T5/T6) y4
__global__ void __launch_bounds__(64) k_adj_param(DevTables T, const double *__restrict__ p, const int2 *__restrict__ ent,
                                                  const double *__restrict__ z, const double *__restrict__ lam,
                                                  double dt, FaultState fs, double *__restrict__ g,
                                                  const uint8_t *__restrict__ act) {
  const int e = blockIdx.x;
  const int m = lane_scenario();
  if (m >= T.M || !act[m]) return;
  const int ldM = T.ldM;
  const int kind = ent[e].x, row = ent[e].y;
  auto Z = [&](int i) { return z[(size_t)i * ldM + m]; };
  auto LAM = [&](int i) { return lam[(size_t)i * ldM + m]; };
  double acc = 0.0;
  if (kind == GP_ADJ_GENROU_H || kind == GP_ADJ_GENROU_D) {
    const int *gr = T.genrou + row * GP_COLS_GENROU;
    const int dp = gr[0], ap = gr[1], bus = gr[3], pm = gr[4], efd = gr[5];
    const double *P = p + gr[2];
    const double H = P[6], w = Z(dp + 4);
    if (kind == GP_ADJ_GENROU_D) {
      acc = LAM(dp + 4) * dt * w / (2.0 * H);
    } else {
      double zs[10], fr[10], ir, ii;
#pragma unroll
      for (int i = 0; i < 6; i++) zs[i] = Z(dp + i);
#pragma unroll
      for (int i = 0; i < 4; i++) zs[6 + i] = Z(ap + i);
      const int vrz = T.net_ptr + 2 * bus;
      const double e_fd = efd >= 0 ? Z(efd) : T.u0[2 * row];
      const double p_m = pm >= 0 ? Z(pm) : T.u0[2 * row + 1];
      genrou_residual(P, zs, e_fd, p_m, Z(vrz), Z(vrz + 1), fr, ir, ii);
      acc = LAM(dp + 4) * dt * fr[4] / H;
    }
  } else if (kind == GP_ADJ_SEXS_K || kind == GP_ADJ_ESDC1A_KA) {
    const int *d = kind == GP_ADJ_SEXS_K ? T.sexs + row * GP_COLS_SEXS : T.esdc1a + row * GP_COLS_ESDC1A;
    const double *P = p + d[1];
    const double vr = Z(d[2]), vi = Z(d[2] + 1), vs = d[3] >= 0 ? Z(d[3]) : 0.0;
    double vm = sqrt(vr * vr + vi * vi);
    if (vm == 0.0) vm = 1e-12;
    if (exc_offline(fs, ldM, m, kind == GP_ADJ_SEXS_K ? GP_SEXS : GP_ESDC1A, row)) {
      acc = 0.0;   // offline: its rows of F do not depend on its parameters
    } else if (kind == GP_ADJ_SEXS_K) {
      const double err = P[6] - vm + vs;
      acc = -LAM(d[0] + 1) * dt * (Z(d[0]) + P[0] * err) / P[3];
    } else {
      const double u = P[9] - vm + vs - Z(d[0] + 1) - (P[2] / P[3]) * Z(d[0] + 2);
      acc = -LAM(d[0]) * dt * u / P[1];
    }
  } else if (kind == GP_ADJ_TGOV1_R) {
    const int *d = T.tgov1 + row * GP_COLS_TGOV1;
    const double *P = p + d[2];
    const double w = d[3] >= 0 ? Z(d[3]) : 0.0, R = P[0];
    acc = LAM(d[0] + 1) * dt * (P[7] - w) / (R * R * P[1]);
  } else if (kind == GP_ADJ_IEESGO_K1) {
    const int *d = T.ieesgo + row * GP_COLS_IEESGO;
    const double *P = p + d[2];
    const double w = d[3] >= 0 ? Z(d[3]) : 0.0;
    acc = -LAM(d[0]) * dt * w / P[0];
  } else if (kind == GP_ADJ_IEEEST_KS) {
    const int *d = T.ieeest + row * GP_COLS_IEEEST;
    const double *P = p + d[2];
    const double A3 = P[2], A4 = P[3], A5 = P[4], A6 = P[5];
    const double T1 = P[6], T2 = P[7], T3 = P[8], T4 = P[9], T5 = P[10], T6 = P[11];
    const double s1 = Z(d[0] + 1), s2 = Z(d[0] + 2), s3 = Z(d[0] + 3), s4 = Z(d[0] + 4), s5 = Z(d[0] + 5);
    const double qq = s1 - s3 - A3 * s2;
    const double y2 = s3 + A5 * s2 + (A6 / A4) * qq;
    const double y3 = s4 + (T1 / T2) * (y2 - s4);
    const double y4 = s5 + (T3 / T4) * (y3 - s5);
    acc = -LAM(d[0] + 6) * dt * y4 / T6 - LAM(d[1]) * (T5 / T6) * y4;
  }
  g[(size_t)e * ldM + m] -= acc;
}

// ---------------------------------------------------------------- small utilities
// Flags of the asynchronous refactorisation for the scenarios in a:
// mode 0 (a batch landed): cur ^= 1, pend = 0; mode 1 (requested): pend = 1;
// mode 2 (request dropped): pend = 0.
__global__ void k_async_flags(int L, int mode, const uint8_t *__restrict__ a, uint8_t *__restrict__ cur,
                              uint8_t *__restrict__ pend) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= L || !a[m]) return;
  if (mode == 0) {
    cur[m] = (uint8_t)(1 - cur[m]);
    pend[m] = 0;
  } else {
    pend[m] = mode == 1 ? 1 : 0;
  }
}

// A batch's masks from a bit set passed by value (no copy from the host):
// req[m] = bit m, tgt[m] = the buffer the solves of m do not read, and the
// requested scenarios become pending.
__global__ void k_stage_batch(int L, ScenarioBits bits, uint8_t *__restrict__ req, uint8_t *__restrict__ tgt,
                              const uint8_t *__restrict__ cur, uint8_t *__restrict__ pend) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= L) return;
  const uint8_t r = (uint8_t)((bits.w[m >> 5] >> (m & 31)) & 1u);
  req[m] = r;
  tgt[m] = (uint8_t)(1 - cur[m]);
  if (r) pend[m] = 1;
}
void launch_stage_batch(int L, const ScenarioBits &bits, uint8_t *req, uint8_t *tgt, const uint8_t *cur,
                        uint8_t *pend, hipStream_t s) {
  hipLaunchKernelGGL(k_stage_batch, dim3((L + 255) / 256), dim3(256), 0, s, L, bits, req, tgt, cur, pend);
}

// Splits the round's Schur, refactorisation and correction masks by the
// precision of each scenario's factors: out[0..2] = (sch, refac, upd) of the
// scenarios with FP32 factors, out[3..5] those with FP64 factors (x64[m]).
__global__ void k_split_precision(int L, const uint8_t *__restrict__ x64, const uint8_t *__restrict__ sch,
                                  const uint8_t *__restrict__ refac, const uint8_t *__restrict__ upd,
                                  uint8_t *__restrict__ out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= L) return;
  const uint8_t x = x64[m] ? 1 : 0, k = (uint8_t)(1 - x);
  out[m] = sch[m] & k;
  out[L + m] = refac[m] & k;
  out[2 * L + m] = upd[m] & k;
  out[3 * L + m] = sch[m] & x;
  out[4 * L + m] = refac[m] & x;
  out[5 * L + m] = upd[m] & x;
}

// Columns of the scenarios in mask into their own slot of a stack of
// scenario-major arrays: dst[slot[m]*stride + i*ldM + m] = src[i*ldM + m].
__global__ void k_copy_cols(int n, int ldM, const double *__restrict__ src, double *__restrict__ dst, size_t stride,
                            const int *__restrict__ slot, const uint8_t *__restrict__ mask) {
  const int m = blockIdx.y * WAVE + threadIdx.x;
  if (!mask[m]) return;
  double *d = dst + (size_t)slot[m] * stride;
  for (int i = blockIdx.x; i < n; i += gridDim.x) d[(size_t)i * ldM + m] = src[(size_t)i * ldM + m];
}

// Scenario-major copy of the scenarios in act: dst[m*n + r] = src[r*ldM + m],
// through an LDS tile of 64 rows x 64 scenarios (rows read for the active
// lanes only, each active scenario's 64 rows written as one run).  A block
// whose 64 scenarios are all inactive returns at once (the four waves see
// the same mask).
__global__ void __launch_bounds__(256) k_copy_transpose(int n, int ldM, const double *__restrict__ src,
                                                        double *__restrict__ dst, const uint8_t *__restrict__ act) {
  __shared__ double tile[64][65];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
  const int r0 = blockIdx.x * 64, s0 = blockIdx.y * WAVE;
  const bool on = act[s0 + lane] != 0;
  if (!__any(on)) return;
  for (int i = w; i < 64; i += 4) {
    const int r = r0 + i;
    tile[i][lane] = (on && r < n) ? src[(size_t)r * ldM + s0 + lane] : 0.0;
  }
  __syncthreads();
  for (int j = w; j < WAVE; j += 4) {
    if (!act[s0 + j]) continue;
    const int r = r0 + lane;
    if (r < n) dst[(size_t)(s0 + j) * n + r] = tile[lane][j];
  }
}

__global__ void k_copy_active(int n, int ldM, const double *__restrict__ src, double *__restrict__ dst,
                              const uint8_t *__restrict__ act) {
  const int m = blockIdx.y * WAVE + threadIdx.x;
  if (!act[m]) return;
  for (int i = blockIdx.x; i < n; i += gridDim.x) dst[(size_t)i * ldM + m] = src[(size_t)i * ldM + m];
}

// The scenarios in act start a new step: zold = z, and where pred[m] the
// first iterate is the linear extrapolation z = 2 z - zold (the previous
// step's increment once more).
// The accepted steps of a round, one pass over the rows.  Scenarios in acc
// record z_k into trajectory column col[m]: row r goes to the channels
// rec_ch[rec_p[r] .. rec_p[r+1]) (rec_p NULL: channel r).  Scenarios in adv
// then start their next step: z_k becomes zold, and z starts at z_k
// (pred 0), at the linear extrapolation 2 z_k - z_(k-1) (pred 1) or at the
// quadratic one 3 z_k - 3 z_(k-1) + z_(k-2) (pred 2); zold2 (optional) keeps
// z_(k-1) for the quadratic one.
__global__ void k_advance(int n, int ldM, double *__restrict__ z, double *__restrict__ zold,
                          double *__restrict__ zold2, const uint8_t *__restrict__ adv,
                          const uint8_t *__restrict__ pred, const uint8_t *__restrict__ acc,
                          const int *__restrict__ col, const int *__restrict__ rec_p,
                          const int *__restrict__ rec_ch, int nrec, double *__restrict__ hist) {
  const int m = blockIdx.y * WAVE + threadIdx.x;
  const bool a = adv[m] != 0, r = acc && acc[m] != 0;
  if (!a && !r) return;
  const int p = pred[m];
  const size_t base = r ? (size_t)col[m] * nrec : 0;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const size_t o = (size_t)i * ldM + m;
    const double v = z[o];
    if (r) {
      if (rec_p) {
        for (int c = rec_p[i]; c < rec_p[i + 1]; c++) hist[(base + rec_ch[c]) * ldM + m] = v;
      } else {
        hist[(base + i) * ldM + m] = v;
      }
    }
    if (a) {
      const double w = zold[o];
      if (p == 1) z[o] = 2.0 * v - w;
      if (p == 2) z[o] = 3.0 * (v - w) + zold2[o];
      if (zold2) zold2[o] = w;
      zold[o] = v;
    }
  }
}

__global__ void k_record(int nrec, const int *__restrict__ idx, int ldM, int M, const double *__restrict__ z,
                         double *__restrict__ hist, const uint8_t *__restrict__ act) {
  const int m = blockIdx.y * WAVE + threadIdx.x;
  if (m >= M || !act[m]) return;
  for (int j = blockIdx.x; j < nrec; j += gridDim.x) {
    const int r = idx ? idx[j] : j;
    hist[(size_t)j * ldM + m] = z[(size_t)r * ldM + m];
  }
}

// One scenario's recorded history, transposed: out[j*len + k] = hist[(k*nrec + j)*ldM + m].
__global__ void k_gather_traj(int nrec, int len, int ldM, int m, const double *__restrict__ hist,
                              double *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nrec * len) return;
  const int j = (int)(i / len), k = (int)(i % len);
  out[i] = hist[((size_t)k * nrec + j) * ldM + m];
}

// Plain copy (calibration of byte counters).
__global__ void k_copy_plain(const double *__restrict__ src, double *__restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------- launchers
static inline dim3 lane_grid(int items, int ldM) { return dim3(items, ldM / WAVE); }

void launch_cluster_residual(const DevTables &T, const double *p, const double *z, const double *zold, double dt,
                             FaultState fs, double *f, double *inj, const uint8_t *work, hipStream_t s,
                             const double *dtv, double *pm) {
  if (T.n_cl_nt == 0) return;
  hipLaunchKernelGGL(k_cluster_residual, lane_grid(T.n_cl_nt, T.ldM), dim3(WAVE), 0, s, T, p, z, zold, dt, dtv, fs, f,
                     inj, work, pm);
}
void launch_bus_residual(const DevTables &T, const double *p, const double *z, const double *inj, FaultState fs,
                         double *f, const uint8_t *work, hipStream_t s, double *pm) {
  hipLaunchKernelGGL(k_bus_residual, lane_grid(T.nbus, T.ldM), dim3(WAVE), 0, s, T, p, z, inj, fs, f, work, pm);
}
void launch_rownorm_impl(int n, int M, int ldM, const double *f, uint8_t *work, double tol, double *fnorm,
                         unsigned long long *nb, int *n_unconv, uint8_t *refac, const uint8_t *refac0, int first,
                         double theta, int stall, int itermax, double *fbest, int *nstall, int *iters,
                         uint8_t *failed, uint8_t *pend, uint8_t *upd, double async_below, uint8_t *areq,
                         uint8_t *sch, hipStream_t s, const uint8_t *firstv, const uint8_t *exact,
                         int switch_window) {
  const int chunks = std::max(1, std::min((n + 15) / 16, 1024 / (ldM / WAVE)));
  hipLaunchKernelGGL(k_rownorm, dim3(ldM / WAVE, chunks), dim3(256), 0, s, n, ldM, M, f, work, nb);
  hipLaunchKernelGGL(k_norm_finish, dim3((M + 255) / 256), dim3(256), 0, s, M, tol, nb, fnorm, work, n_unconv, refac,
                     refac0, first, firstv, theta, stall, itermax, fbest, nstall, iters, failed, pend, upd,
                     async_below, areq, sch, exact, switch_window);
}
void launch_jac_mono(const DevTables &T, const double *p, const double *z, double dt, FaultState fs, double *jv,
                     const uint8_t *work, hipStream_t s, const double *dtv) {
  if (T.n_cl_nt)
    hipLaunchKernelGGL(k_cluster_jac_mono, lane_grid(T.n_cl_nt, T.ldM), dim3(WAVE), 0, s, T, p, z, dt, dtv, fs, jv,
                       work);
  hipLaunchKernelGGL(k_bus_jac_mono, lane_grid(T.nbus, T.ldM), dim3(WAVE), 0, s, T, p, z, fs, jv, work);
}
template <typename TF>
void launch_cluster_schur(const DevTables &T, const double *p, const double *z, const double *f, double dt,
                          FaultState fs, TF *X, double *Dg, const uint8_t *work, hipStream_t s, const double *dtv) {
  if (T.n_cl_nt == 0) return;
  hipLaunchKernelGGL(k_cluster_schur<TF>, lane_grid(T.n_cl_nt, T.ldM), dim3(WAVE), 0, s, T, p, z, f, dt, dtv, fs, X,
                     Dg, work);
}
void launch_bus_schur(const DevTables &T, const double *p, const double *z, const double *f, const double *Dg,
                      FaultState fs, double *svar, double *rhs, const uint8_t *work, hipStream_t s) {
  hipLaunchKernelGGL(k_bus_schur, lane_grid(T.nbus, T.ldM), dim3(WAVE), 0, s, T, p, z, f, Dg, fs, svar, rhs, work);
}
template <typename TF>
void launch_cluster_backsub(const DevTables &T, const TF *X, const double *f, const double *dv, double *z,
                            const uint8_t *work, hipStream_t s) {
  (void)f;
  if (T.n_cl_nt == 0) return;
  hipLaunchKernelGGL(k_cluster_backsub<TF>, lane_grid(T.n_cl_nt, T.ldM), dim3(WAVE), 0, s, T, X, dv, z, work);
}
// Dynamic LDS above 64 KB needs the kernel's attribute raised once per
// device; raised here to the whole 160 KB of a CU the first time a launch
// needs more than 64 KB (an attribute call on every launch costs host time
// on the Newton rounds' critical path).
static void allow_lds(const void *f, std::atomic<unsigned> &done, int bytes) {
  if (bytes <= 65536) return;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const unsigned bit = 1u << (dev & 31);
  if (done.load(std::memory_order_acquire) & bit) return;
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  done.fetch_or(bit, std::memory_order_acq_rel);
}

template <typename TF>
void launch_splu_factor_impl(const SpluDev &S, const double *var, const double *cst, int ldM, TF *luv, TF *sv, TF *dn,
                             const uint8_t *refac, const uint8_t *tgt, int M, int *status, hipStream_t s,
                             hipEvent_t factored) {
  if (S.res_nsteps > 0) {
    static std::atomic<unsigned> done_res{0};
    allow_lds(reinterpret_cast<const void *>(k_splu_factor_res<TF>), done_res, S.lds_factor);
    hipLaunchKernelGGL(k_splu_factor_res<TF>, dim3(M), dim3(S.threads), S.lds_factor, s, S, var, cst, ldM, luv, sv,
                       dn, refac, tgt, status);
  } else {
    static std::atomic<unsigned> done_fac{0};
    allow_lds(reinterpret_cast<const void *>(k_splu_factor<TF>), done_fac, S.lds_factor);
    hipLaunchKernelGGL(k_splu_factor<TF>, dim3(M), dim3(S.threads), S.lds_factor, s, S, var, cst, ldM, luv, sv, dn,
                       refac, tgt, status);
  }
  if (factored) (void)hipEventRecord(factored, s);
  if (S.cmp_nph > 0) {   // the partitioned solves' composites, from the values just written
    hipLaunchKernelGGL(k_splu_compose<TF>, dim3(M), dim3(S.threads), 0, s, S, sv, refac, tgt);
  }
}
template <typename TF>
void launch_splu_solve(const SpluDev &S, const TF *sv, const TF *dn, const double *rhs, int ldM, double sign,
                       const int *scatter, double *zupd, double *xout, const uint8_t *work, const uint8_t *cur, int M,
                       hipStream_t s, bool shared, const WbTab *wb) {
  if (S.gx) {   // (the host refuses a shared factorisation with x in global memory)
    static std::atomic<unsigned> done_gx{0};
    allow_lds(reinterpret_cast<const void *>(k_splu_solve_gx<TF>), done_gx, S.lds_solve);
    const dim3 lanes(std::min(S.n, 4096), (ldM + WAVE - 1) / WAVE);
    hipLaunchKernelGGL(k_gx_gather, lanes, dim3(WAVE), 0, s, S.n, ldM, S.pinv, rhs, S.gx_ws, work);
    for (int l = 0; l < S.gx_fw_wide; l++)
      hipLaunchKernelGGL((k_gx_level<TF, false>), dim3((S.h_fw_lvl[4 * l] + 255) / 256, M), dim3(256), 0, s,
                         S.fw_lvl, l, S.fw_chunk, S.fw_cols, S.fw_rows, sv, S.ns, 0, 0, M, S.gx_ws, ldM, work, cur);
    hipLaunchKernelGGL(k_splu_solve_gx<TF>, dim3(M), dim3(S.threads), S.lds_solve, s, S, sv, dn, work, cur, ldM);
    for (int l = S.n_bwlev - S.gx_bw_wide; l < S.n_bwlev; l++)
      hipLaunchKernelGGL((k_gx_level<TF, true>), dim3((S.h_bw_lvl[4 * l] + 255) / 256, M), dim3(256), 0, s,
                         S.bw_lvl, l, S.bw_chunk, S.bw_cols, S.bw_rows, sv, S.ns, S.fw_nvals,
                         S.fw_nvals + S.bw_nvals, M, S.gx_ws, ldM, work, cur);
    hipLaunchKernelGGL(k_gx_output, lanes, dim3(WAVE), 0, s, S.n, ldM, S.q, scatter, S.gx_ws, sign, zupd, xout, work);
    return;
  }
  static std::atomic<unsigned> done_sol{0};
  allow_lds(reinterpret_cast<const void *>(k_splu_solve<TF>), done_sol, S.lds_solve);
  hipLaunchKernelGGL(k_splu_solve<TF>, dim3(M), dim3(S.threads), S.lds_solve, s, S, sv, dn, rhs, ldM, sign, scatter,
                     zupd, xout, work, cur, shared ? 1 : 0, wb ? *wb : WbTab{});
}
void launch_adj_mask(int ldM, int k, const int *hz, uint8_t *act, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_mask, dim3((ldM + 255) / 256), dim3(256), 0, s, ldM, k, hz, act);
}
void launch_adj_carry(int n, int ldM, const uint8_t *is_diff, const double *lam, double *rhs, const uint8_t *act,
                      hipStream_t s) {
  const int gx = n < 1024 ? n : 1024;
  hipLaunchKernelGGL(k_adj_carry, dim3(gx, ldM / WAVE), dim3(WAVE), 0, s, n, ldM, is_diff, lam, rhs, act);
}
void launch_adj_objective(int nr, const int *rows, int ldM, double w, double off, const double *z, double *rhs,
                          double *val, const uint8_t *act, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_objective, dim3((ldM + 63) / 64), dim3(64), 0, s, nr, rows, ldM, w, off, z, rhs, val, act);
}
void launch_adj_param(const DevTables &T, const double *p, int ne, const int2 *ent, const double *z,
                      const double *lam, double dt, FaultState fs, double *g, const uint8_t *act, hipStream_t s) {
  if (ne > 0)
    hipLaunchKernelGGL(k_adj_param, lane_grid(ne, T.ldM), dim3(WAVE), 0, s, T, p, ent, z, lam, dt, fs, g, act);
}
void launch_copy_plain(const double *src, double *dst, long long n, hipStream_t s) {
  if (n <= 0) return;
  const long long blocks = std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_copy_plain, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
}
void launch_gather_traj(int nrec, int len, int ldM, int m, const double *hist, double *out, hipStream_t s) {
  const size_t tot = (size_t)nrec * len;
  if (tot == 0) return;
  hipLaunchKernelGGL(k_gather_traj, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, nrec, len, ldM, m, hist, out);
}
void launch_copy_cols(int n, int ldM, const double *src, double *dst, size_t stride, const int *slot,
                      const uint8_t *mask, hipStream_t s) {
  const int gx = n < 1024 ? n : 1024;
  hipLaunchKernelGGL(k_copy_cols, dim3(gx, ldM / WAVE), dim3(WAVE), 0, s, n, ldM, src, dst, stride, slot, mask);
}
void launch_async_flags(int L, int mode, const uint8_t *a, uint8_t *cur, uint8_t *pend, hipStream_t s) {
  hipLaunchKernelGGL(k_async_flags, dim3((L + 255) / 256), dim3(256), 0, s, L, mode, a, cur, pend);
}
void launch_split_precision(int L, const uint8_t *x64, const uint8_t *sch, const uint8_t *refac, const uint8_t *upd,
                            uint8_t *out, hipStream_t s) {
  hipLaunchKernelGGL(k_split_precision, dim3((L + 255) / 256), dim3(256), 0, s, L, x64, sch, refac, upd, out);
}
void launch_copy_transpose(int n, int ldM, const double *src, double *dst, const uint8_t *act, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_transpose, dim3((n + 63) / 64, ldM / WAVE), dim3(256), 0, s, n, ldM, src, dst, act);
}
void launch_copy_active(int n, int ldM, const double *src, double *dst, const uint8_t *act, hipStream_t s) {
  const int gx = n < 1024 ? n : 1024;
  hipLaunchKernelGGL(k_copy_active, dim3(gx, ldM / WAVE), dim3(WAVE), 0, s, n, ldM, src, dst, act);
}
void launch_advance(int n, int ldM, double *z, double *zold, double *zold2, const uint8_t *adv, const uint8_t *pred,
                    const uint8_t *acc, const int *col, const int *rec_p, const int *rec_ch, int nrec, double *hist,
                    hipStream_t s) {
  const int gx = n < 1024 ? n : 1024;
  hipLaunchKernelGGL(k_advance, dim3(gx, ldM / WAVE), dim3(WAVE), 0, s, n, ldM, z, zold, zold2, adv, pred, acc, col,
                     rec_p, rec_ch, nrec, hist);
}
void launch_record_impl(int nrec, const int *idx, int ldM, int M, const double *z, double *hist, const uint8_t *act,
                        hipStream_t s) {
  const int gx = nrec < 1024 ? (nrec > 0 ? nrec : 1) : 1024;
  hipLaunchKernelGGL(k_record, dim3(gx, ldM / WAVE), dim3(WAVE), 0, s, nrec, idx, ldM, M, z, hist, act);
}

// ---------------------------------------------------------------- allocation-free residual
// Device memory of gp_eval_residual_noalloc: the scenario mask, the fault
// tables and the generators' current injections of one evaluation.  One copy
// per GPU, held until the process ends; every call on a GPU shares it (calls
// that overlap on one GPU race).  Host code reaches it only through the
// address scratch_setup() takes once per device.
static __device__ __align__(256) unsigned char scratch[GP_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int scratch_setup(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return e != hipSuccess ? (int)e : -1;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

// mask[m] = (m < M), m < ldM
__global__ void k_prefix_mask(int ldM, int M, uint8_t *__restrict__ mask) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < ldM) mask[m] = m < M ? 1 : 0;
}

// The fault tables of an evaluation: scenario m's events (ev_bus / ev_g,
// [MAXFAULT][ldM]) where fault_on[m] != 0 (fault_on NULL: none), empty slots
// elsewhere.
__global__ void k_select_faults(int ldM, const int *__restrict__ ev_bus, const double *__restrict__ ev_g,
                                const uint8_t *__restrict__ fault_on, int *__restrict__ bus, double *__restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MAXFAULT * ldM) return;
  const int m = i % ldM;
  const bool on = fault_on && fault_on[m];
  bus[i] = on ? ev_bus[i] : -1;
  g[i] = on ? ev_g[i] : 0.0;
}

void launch_prefix_mask(int ldM, int M, uint8_t *mask, hipStream_t s) {
  hipLaunchKernelGGL(k_prefix_mask, dim3((ldM + 255) / 256), dim3(256), 0, s, ldM, M, mask);
}
void launch_select_faults(int ldM, const int *ev_bus, const double *ev_g, const uint8_t *fault_on, int *bus,
                          double *g, hipStream_t s) {
  hipLaunchKernelGGL(k_select_faults, dim3((MAXFAULT * ldM + 255) / 256), dim3(256), 0, s, ldM, ev_bus, ev_g,
                     fault_on, bus, g);
}

template void launch_cluster_schur<double>(const DevTables &, const double *, const double *, const double *, double,
                                             FaultState, double *, double *, const uint8_t *, hipStream_t, const double *);
template void launch_cluster_schur<float>(const DevTables &, const double *, const double *, const double *, double,
                                             FaultState, float *, double *, const uint8_t *, hipStream_t, const double *);
template void launch_cluster_backsub<double>(const DevTables &, const double *, const double *, const double *,
                                             double *, const uint8_t *, hipStream_t);
template void launch_cluster_backsub<float>(const DevTables &, const float *, const double *, const double *,
                                            double *, const uint8_t *, hipStream_t);
template void launch_splu_factor_impl<double>(const SpluDev &, const double *, const double *, int, double *,
                                              double *, double *, const uint8_t *, const uint8_t *, int, int *,
                                              hipStream_t, hipEvent_t);
template void launch_splu_factor_impl<float>(const SpluDev &, const double *, const double *, int, float *, float *,
                                             float *, const uint8_t *, const uint8_t *, int, int *, hipStream_t,
                                             hipEvent_t);
template void launch_splu_solve<double>(const SpluDev &, const double *, const double *, const double *, int, double,
                                        const int *, double *, double *, const uint8_t *, const uint8_t *, int,
                                        hipStream_t, bool, const WbTab *);
template void launch_splu_solve<float>(const SpluDev &, const float *, const float *, const double *, int, double,
                                       const int *, double *, double *, const uint8_t *, const uint8_t *, int,
                                       hipStream_t, bool, const WbTab *);

}  // namespace gp

/*
 * gp_oracle.c -- CPU oracle for the batched backward-Euler DAE integrator.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates GradPower.jl's CPU path so
 * that tests, the smoke check and the benchmark's CPU baseline can check the
 * HIP library against it.  Nothing in the product links or calls it.
 *
 * Restated from (paths relative to the GradPower.jl repository):
 *   Jacobian pattern          src/dynamics.jl:687-794 (+ per-device *_preallocate!)
 *   residual                  src/dynamics.jl:265-327 and the device leaves
 *                             src/kernels/{genrou,ieesgo,tgov1,sexs,esdc1a,ieeest,
 *                             zipload,static_gen}.jl
 *   Jacobian values           src/dynamics.jl:823-894 and the same leaves
 *   backward Euler            src/dynamics.jl:339-500
 *   Newton                    src/numerics.jl:55-176
 *   time / event loop         src/dynamics.jl:502-655
 *   line trips / reclosures   src/GradPower.jl:140-167,509-623 (TripLineEvent,
 *                             _apply_trip_line!, _apply_reclose_line!)
 *   ZIP load disconnects      src/GradPower.jl:130-143,681-698 (DisconnectDeviceEvent)
 *   discrete adjoint          examples/adjoint.jl:34-51,106 and
 *                             test/test_adjoint_event_offset.jl:3-24 (the
 *                             adjoint itself is absent from the reference
 *                             snapshot: pinned by central differences)
 *   Schur Newton step         src/schur.jl:195-493
 * The sparse LU plays the role of KLU (an external dependency of the
 * reference): fill-reducing minimum-degree column order on A+A^T, then a
 * left-looking Gilbert-Peierls factorisation with threshold partial pivoting
 * (threshold 1e-3, diagonal preferred, the KLU settings of src/dynamics.jl:577-580).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/gradpower_hip.h"

#define TWOPI60 (2.0 * 3.14159265358979323846 * 60.0)

typedef struct {
  int n, nbus, net_ptr, npar;
  uint8_t *is_diff;
  int *yrp, *yci;
  double *yv;
  int ndev[GP_NDEVTYPES];
  int *dev[GP_NDEVTYPES];
  double *u0;
  int ncl;
  int *cl;
  /* Jacobian pattern, CSC, rows sorted within each column */
  int nnz;
  int *cp, *ri;
  /* sparse LU */
  int *q;          /* column order: k-th pivot column is q[k] */
  int lu_ready;
  /* line trips of the next sweeps (orc_set_trips) */
  int ntr, ynnz;
  int *tr_fr, *tr_to, *tr_on, *tr_off;
  double *tr_y;
  /* device disconnects of the next sweeps (orc_set_drops) and, during a
   * sweep, the online flag of every device that can go offline (NULL: all
   * online), indexed ZIP loads first, then SEXS, then ESDC1A rows */
  int ndrop;
  /* 1: a sweep stops at its first Newton solve that does not converge (the
   * HIP library's rule; the reference itself keeps integrating) */
  int stop_on_fail;
  int *drop_k, *drop_step;
  uint8_t *zon;
} orc_sys;

/* ---------------------------------------------------------------- utils */
static void *xmalloc(size_t s) { void *p = calloc(1, s ? s : 1); if (!p) abort(); return p; }
static int *idup(const int32_t *a, size_t n) { int *b = xmalloc(n * sizeof(int)); for (size_t i = 0; i < n; i++) b[i] = a[i]; return b; }

static int cmp_int(const void *a, const void *b) { int x = *(const int *)a, y = *(const int *)b; return (x > y) - (x < y); }

/* position of (r, c) in the CSC pattern, -1 if absent */
static int jpos(const orc_sys *s, int r, int c) {
  int lo = s->cp[c], hi = s->cp[c + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    if (s->ri[mid] == r) return mid;
    if (s->ri[mid] < r) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}
static void jset(const orc_sys *s, double *J, int r, int c, double v) { int p = jpos(s, r, c); if (p < 0) abort(); J[p] = v; }
static void jadd(const orc_sys *s, double *J, int r, int c, double v) { int p = jpos(s, r, c); if (p < 0) abort(); J[p] += v; }

/* --------------------------------------------------- pattern (coord list) */
typedef struct { int *r, *c; int n, cap; } coo;
static void cpush(coo *a, int r, int c) {
  if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 1024; a->r = realloc(a->r, a->cap * sizeof(int)); a->c = realloc(a->c, a->cap * sizeof(int)); }
  a->r[a->n] = r; a->c[a->n] = c; a->n++;
}
static void crow(coo *a, int r, const int *cols, int k) { for (int i = 0; i < k; i++) cpush(a, r, cols[i]); }

#define D(t, k, j) (s->dev[t][(k) * cols_##t + (j)])
static const int cols_0 = GP_COLS_GENROU, cols_1 = GP_COLS_IEESGO, cols_2 = GP_COLS_TGOV1,
                 cols_3 = GP_COLS_SEXS, cols_4 = GP_COLS_ESDC1A, cols_5 = GP_COLS_IEEEST,
                 cols_6 = GP_COLS_ZIPLOAD, cols_7 = GP_COLS_STATICGEN;

/* preallocate_jacobian, src/dynamics.jl:687-794 */
static void build_pattern(orc_sys *s) {
  coo a = {0};
  int n = s->n, npt = s->net_ptr;
  for (int i = 0; i < npt; i++) if (s->is_diff[i]) cpush(&a, i, i);
  for (int b = 0; b < s->nbus; b++)
    for (int u = 0; u < 2; u++) for (int v = 0; v < 2; v++) cpush(&a, npt + 2 * b + u, npt + 2 * b + v);
  for (int r = 0; r < 2 * s->nbus; r++)
    for (int p = s->yrp[r]; p < s->yrp[r + 1]; p++) {
      int c = s->yci[p];
      /* bus-level adjacency: whole 2x2 blocks */
      int rb = r / 2, cb = c / 2;
      for (int u = 0; u < 2; u++) for (int v = 0; v < 2; v++) cpush(&a, npt + 2 * rb + u, npt + 2 * cb + v);
    }
  for (int k = 0; k < s->ndev[GP_GENROU]; k++) {   /* src/generators.jl:445-530 + src/kernels/genrou.jl:242-254 */
    int dp = D(0, k, 0), ap = D(0, k, 1), bus = D(0, k, 3), pm = D(0, k, 4), efd = D(0, k, 5);
    int eq = dp, ed = dp + 1, p1 = dp + 2, p2 = dp + 3, w = dp + 4, de = dp + 5;
    int vq = ap, vd = ap + 1, iq = ap + 2, id = ap + 3, vr = npt + 2 * bus, vi = vr + 1;
    int r0[] = {eq, ed, p1, p2, id}; crow(&a, dp, r0, 5);
    if (efd >= 0) cpush(&a, dp, efd);
    int r1[] = {ed, p2, iq}; crow(&a, dp + 1, r1, 3);
    int r2[] = {eq, p1, id}; crow(&a, dp + 2, r2, 3);
    int r3[] = {ed, p2, iq}; crow(&a, dp + 3, r3, 3);
    int r4[] = {eq, ed, p1, p2, iq, id, w}; crow(&a, dp + 4, r4, 7);
    if (pm >= 0) cpush(&a, dp + 4, pm);
    cpush(&a, dp + 5, w);
    int a0[] = {eq, p1, vq, id}; crow(&a, ap, a0, 4);
    int a1[] = {ed, p2, vd, iq}; crow(&a, ap + 1, a1, 4);
    int a2[] = {de, vd, vr, vi}; crow(&a, ap + 2, a2, 4);
    int a3[] = {de, vq, vr, vi}; crow(&a, ap + 3, a3, 4);
    int nr[] = {de, iq, id}; crow(&a, vr, nr, 3); crow(&a, vi, nr, 3);
  }
  for (int k = 0; k < s->ndev[GP_IEESGO]; k++) {   /* src/kernels/ieesgo.jl:52-84 */
    int dp = D(1, k, 0), ap = D(1, k, 1), w = D(1, k, 3);
    cpush(&a, dp, dp); if (w >= 0) cpush(&a, dp, w);
    cpush(&a, dp + 1, dp); cpush(&a, dp + 1, dp + 1);
    cpush(&a, dp + 2, dp); cpush(&a, dp + 2, dp + 1); cpush(&a, dp + 2, dp + 2);
    cpush(&a, dp + 3, dp + 2); cpush(&a, dp + 3, dp + 3);
    cpush(&a, dp + 4, dp + 3); cpush(&a, dp + 4, dp + 4);
    cpush(&a, ap, dp + 2); cpush(&a, ap, dp + 3); cpush(&a, ap, dp + 4); cpush(&a, ap, ap);
  }
  for (int k = 0; k < s->ndev[GP_TGOV1]; k++) {    /* src/kernels/tgov1.jl:37-60 */
    int dp = D(2, k, 0), ap = D(2, k, 1), w = D(2, k, 3);
    cpush(&a, dp, dp); cpush(&a, dp, dp + 1);
    cpush(&a, dp + 1, dp + 1); if (w >= 0) cpush(&a, dp + 1, w);
    cpush(&a, ap, dp); cpush(&a, ap, dp + 1); if (w >= 0) cpush(&a, ap, w); cpush(&a, ap, ap);
  }
  for (int k = 0; k < s->ndev[GP_IEEEST]; k++) {   /* src/kernels/ieeest.jl:75-124 */
    int dp = D(5, k, 0), ap = D(5, k, 1), w = D(5, k, 3);
    cpush(&a, dp, dp); cpush(&a, dp, dp + 1); if (w >= 0) cpush(&a, dp, w);
    cpush(&a, dp + 1, dp);
    cpush(&a, dp + 2, dp + 1); cpush(&a, dp + 2, dp + 2); cpush(&a, dp + 2, dp + 3);
    cpush(&a, dp + 3, dp + 2);
    for (int j = 1; j <= 4; j++) cpush(&a, dp + 4, dp + j);
    for (int j = 1; j <= 5; j++) cpush(&a, dp + 5, dp + j);
    for (int j = 1; j <= 6; j++) cpush(&a, dp + 6, dp + j);
    for (int j = 1; j <= 6; j++) cpush(&a, ap, dp + j);
    cpush(&a, ap, ap);
  }
  for (int t = GP_SEXS; t <= GP_ESDC1A; t++)
    for (int k = 0; k < s->ndev[t]; k++) {           /* src/kernels/sexs.jl:29-52, esdc1a.jl:32-58 */
      const int *row = s->dev[t] + k * 4;
      int dp = row[0], vr = row[2], vi = vr + 1, vs = row[3];
      if (t == GP_SEXS) {
        cpush(&a, dp, dp); cpush(&a, dp, vr); cpush(&a, dp, vi); if (vs >= 0) cpush(&a, dp, vs);
        cpush(&a, dp + 1, dp); cpush(&a, dp + 1, dp + 1); cpush(&a, dp + 1, vr); cpush(&a, dp + 1, vi);
        if (vs >= 0) cpush(&a, dp + 1, vs);
      } else {
        cpush(&a, dp, dp); cpush(&a, dp, dp + 1); cpush(&a, dp, dp + 2); cpush(&a, dp, vr); cpush(&a, dp, vi);
        if (vs >= 0) cpush(&a, dp, vs);
        cpush(&a, dp + 1, dp + 1); cpush(&a, dp + 1, dp + 2);
        cpush(&a, dp + 2, dp); cpush(&a, dp + 2, dp + 2);
      }
    }
  for (int k = 0; k < s->ndev[GP_STATICGEN]; k++) { /* src/kernels/static_gen.jl:14-42 */
    int bus = D(7, k, 0), ap = D(7, k, 1), bt = D(7, k, 3);
    int vr = npt + 2 * bus, vi = vr + 1;
    if (bt == 2) { cpush(&a, ap, vr); cpush(&a, ap, vi); cpush(&a, vr, ap); cpush(&a, vi, ap); }
    else if (bt == 3) { cpush(&a, ap, vr); cpush(&a, ap + 1, vi); cpush(&a, vr, ap); cpush(&a, vr, ap + 1); cpush(&a, vi, ap); cpush(&a, vi, ap + 1); }
  }
  /* coo -> CSC with duplicates merged */
  int *cnt = xmalloc((n + 1) * sizeof(int));
  for (int i = 0; i < a.n; i++) cnt[a.c[i] + 1]++;
  for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  int *tmp = xmalloc((a.n + 1) * sizeof(int)), *fill = xmalloc((n + 1) * sizeof(int));
  memcpy(fill, cnt, (n + 1) * sizeof(int));
  for (int i = 0; i < a.n; i++) tmp[fill[a.c[i]]++] = a.r[i];
  s->cp = xmalloc((n + 1) * sizeof(int));
  s->ri = xmalloc((a.n + 1) * sizeof(int));
  int nz = 0;
  for (int c = 0; c < n; c++) {
    s->cp[c] = nz;
    int b = cnt[c], e = cnt[c + 1];
    qsort(tmp + b, e - b, sizeof(int), cmp_int);
    for (int p = b; p < e; p++) if (p == b || tmp[p] != tmp[p - 1]) s->ri[nz++] = tmp[p];
  }
  s->cp[n] = nz;
  s->nnz = nz;
  free(cnt); free(tmp); free(fill); free(a.r); free(a.c);
}

/* ------------------------------------------------------------ residual */
static void sat_coeff(double s1, double s2, double *sa, double *sb) {
  *sa = 0; *sb = 0;
  if (s1 <= 0.0 || s2 <= 0.0) return;
  double a = sqrt(s1 * 1.0 / (s2 * 1.2));
  if (a == 1.0) return;
  *sa = 1.2 - (1.0 - 1.2) / (a - 1.0);
  *sb = s2 * 1.2 * (a - 1.0) * (a - 1.0) / ((1.0 - 1.2) * (1.0 - 1.2));
}
static double sat_se(double psi, double sa, double sb) {
  if (sb == 0.0 || psi <= sa || psi == 0.0) return 0.0;
  return sb * (psi - sa) * (psi - sa) / psi;
}

/* rhs_fun (src/dynamics.jl:265-327); fault_g[b] = sum of 1/r_f active at bus b */
static void rhs(const orc_sys *s, const double *z, const double *p, const double *fault_g, double *f) {
  int npt = s->net_ptr;
  for (int r = 0; r < 2 * s->nbus; r++) {
    double acc = 0.0;
    for (int q = s->yrp[r]; q < s->yrp[r + 1]; q++) acc += s->yv[q] * z[npt + s->yci[q]];
    f[npt + r] = -acc;
  }
  for (int k = 0; k < s->ndev[GP_GENROU]; k++) {    /* src/kernels/genrou.jl:48-141 */
    int dp = D(0, k, 0), ap = D(0, k, 1), pp = D(0, k, 2), bus = D(0, k, 3), pm = D(0, k, 4), efd = D(0, k, 5);
    const double *P = p + pp;
    double x_d = P[0], x_q = P[1], x_dp = P[2], x_qp = P[3], x_ddp = P[4], xl = P[5], H = P[6], Dm = P[7];
    double T_d0p = P[8], T_q0p = P[9], T_d0dp = P[10], T_q0dp = P[11], S1 = P[12], S2 = P[13], x_qdp = x_ddp;
    double e_qp = z[dp], e_dp = z[dp + 1], phi_1d = z[dp + 2], phi_2q = z[dp + 3], w = z[dp + 4], delta = z[dp + 5];
    double v_q = z[ap], v_d = z[ap + 1], i_q = z[ap + 2], i_d = z[ap + 3];
    int vri = npt + 2 * bus;
    double vr = z[vri], vi = z[vri + 1];
    double e_fd = efd >= 0 ? z[efd] : s->u0[2 * k], p_m = pm >= 0 ? z[pm] : s->u0[2 * k + 1];
    double psi_de = (x_ddp - xl) / (x_dp - xl) * e_qp + (x_dp - x_ddp) / (x_dp - xl) * phi_1d;
    double psi_qe = -(x_ddp - xl) / (x_qp - xl) * e_dp + (x_qp - x_ddp) / (x_qp - xl) * phi_2q;
    double sa, sb; sat_coeff(S1, S2, &sa, &sb);
    double Se = sat_se(sqrt(psi_de * psi_de + psi_qe * psi_qe), sa, sb);
    f[dp] = (-e_qp + e_fd - (i_d - (-x_ddp + x_dp) * (-e_qp + i_d * (x_dp - xl) + phi_1d) / ((x_dp - xl) * (x_dp - xl))) * (x_d - x_dp) - Se * psi_de) / T_d0p;
    f[dp + 1] = (-e_dp + (i_q - (-x_qdp + x_qp) * (e_dp + i_q * (x_qp - xl) + phi_2q) / ((x_qp - xl) * (x_qp - xl))) * (x_q - x_qp)) / T_q0p;
    f[dp + 2] = (e_qp - i_d * (x_dp - xl) - phi_1d) / T_d0dp;
    f[dp + 3] = (-e_dp - i_q * (x_qp - xl) - phi_2q) / T_q0dp;
    f[dp + 4] = (p_m - Dm * w - psi_de * i_q + psi_qe * i_d) / (2.0 * H);
    f[dp + 5] = TWOPI60 * w;
    f[ap] = i_d - ((x_ddp - xl) / (x_dp - xl) * e_qp + (x_dp - x_ddp) / (x_dp - xl) * phi_1d - v_q) / x_ddp;
    f[ap + 1] = i_q - (-(x_qdp - xl) / (x_qp - xl) * e_dp + (x_qp - x_qdp) / (x_qp - xl) * phi_2q + v_d) / x_qdp;
    double sd = sin(delta), cd = cos(delta);
    f[ap + 2] = v_d - (vr * sd - vi * cd);
    f[ap + 3] = v_q - (vr * cd + vi * sd);
    f[vri] += sd * i_d + cd * i_q;
    f[vri + 1] += -cd * i_d + sd * i_q;
  }
  for (int k = 0; k < s->ndev[GP_IEESGO]; k++) {    /* src/kernels/ieesgo.jl:105-131 */
    int dp = D(1, k, 0), ap = D(1, k, 1), w_i = D(1, k, 3);
    const double *P = p + D(1, k, 2);
    double T1 = P[0], T2 = P[1], T3 = P[2], T4 = P[3], T5 = P[4], T6 = P[5], K1 = P[6], K2 = P[7], K3 = P[8], pref = P[11];
    double PF0 = z[dp], PLL = z[dp + 1], TP1 = z[dp + 2], TP2 = z[dp + 3], TP3 = z[dp + 4], p_m = z[ap];
    double w = w_i >= 0 ? z[w_i] : 0.0;
    f[dp] = (1.0 / T1) * (K1 * w - PF0);
    f[dp + 1] = (1.0 / T3) * ((1.0 - (T2 / T3)) * PF0 - PLL);
    double SatP = pref - (T2 / T3) * PF0 - PLL;
    f[dp + 2] = (1.0 / T4) * (SatP - TP1);
    f[dp + 3] = (1.0 / T5) * (K2 * TP1 - TP2);
    f[dp + 4] = (1.0 / T6) * (K3 * TP2 - TP3);
    f[ap] = TP1 * (1.0 - K2) + TP2 * (1.0 - K3) + TP3 - p_m;
  }
  for (int k = 0; k < s->ndev[GP_TGOV1]; k++) {     /* src/kernels/tgov1.jl:73-100 */
    int dp = D(2, k, 0), ap = D(2, k, 1), w_i = D(2, k, 3);
    const double *P = p + D(2, k, 2);
    double R = P[0], T1 = P[1], T2 = P[4], T3 = P[5], DT = P[6], pref = P[7];
    double x1 = z[dp], x2 = z[dp + 1], p_m = z[ap], w = w_i >= 0 ? z[w_i] : 0.0, t23 = T2 / T3;
    f[dp] = (-x1 + (1.0 - t23) * x2) / T3;
    f[dp + 1] = ((pref - w) / R - x2) / T1;
    f[ap] = x1 + t23 * x2 - DT * w - p_m;
  }
  for (int k = 0; k < s->ndev[GP_IEEEST]; k++) {    /* src/kernels/ieeest.jl:207-241 */
    int dp = D(5, k, 0), ap = D(5, k, 1), w_i = D(5, k, 3);
    const double *P = p + D(5, k, 2);
    double A1 = P[0], A2 = P[1], A3 = P[2], A4 = P[3], A5 = P[4], A6 = P[5];
    double T1 = P[6], T2 = P[7], T3 = P[8], T4 = P[9], T5 = P[10], T6 = P[11], KS = P[12];
    double s0 = z[dp], s1 = z[dp + 1], s2 = z[dp + 2], s3 = z[dp + 3], s4 = z[dp + 4], s5 = z[dp + 5], s6 = z[dp + 6];
    double vs = z[ap], sig = w_i >= 0 ? z[w_i] : 0.0;
    double qq = s1 - s3 - A3 * s2;
    double y2 = s3 + A5 * s2 + (A6 / A4) * qq;
    double y3 = s4 + (T1 / T2) * (y2 - s4);
    double y4 = s5 + (T3 / T4) * (y3 - s5);
    f[dp] = (sig - s1 - A1 * s0) / A2;
    f[dp + 1] = s0;
    f[dp + 2] = (s1 - s3 - A3 * s2) / A4;
    f[dp + 3] = s2;
    f[dp + 4] = (y2 - s4) / T2;
    f[dp + 5] = (y3 - s5) / T4;
    f[dp + 6] = (KS * y4 - s6) / T6;
    f[ap] = vs - (T5 / T6) * (KS * y4 - s6);
  }
  for (int k = 0; k < s->ndev[GP_SEXS]; k++) {      /* src/kernels/sexs.jl:74-103 */
    int dp = D(3, k, 0), vri = D(3, k, 2), vsi = D(3, k, 3);
    if (s->zon && !s->zon[s->ndev[GP_ZIPLOAD] + k]) {   /* offline: skipped into a zeroed F (ext:1340) */
      f[dp] = f[dp + 1] = 0.0;
      continue;
    }
    const double *P = p + D(3, k, 1);
    double TA_TB = P[0], TB = P[1], K = P[2], TE = P[3], vref = P[6];
    double x1 = z[dp], e_fd = z[dp + 1], vr = z[vri], vi = z[vri + 1];
    double vm = sqrt(vr * vr + vi * vi); if (vm == 0.0) vm = 1e-12;
    double vs = vsi >= 0 ? z[vsi] : 0.0, err = vref - vm + vs;
    f[dp] = (-x1 + (1.0 - TA_TB) * err) / TB;
    f[dp + 1] = (-e_fd + K * (x1 + TA_TB * err)) / TE;
  }
  for (int k = 0; k < s->ndev[GP_ESDC1A]; k++) {    /* src/kernels/esdc1a.jl:97-131 */
    int dp = D(4, k, 0), vri = D(4, k, 2), vsi = D(4, k, 3);
    if (s->zon && !s->zon[s->ndev[GP_ZIPLOAD] + s->ndev[GP_SEXS] + k]) {
      f[dp] = f[dp + 1] = f[dp + 2] = 0.0;
      continue;
    }
    const double *P = p + D(4, k, 1);
    double Ka = P[0], Ta = P[1], Kf = P[2], Tf = P[3], Ke = P[4], Te = P[5], sa = P[7], sb = P[8], vref = P[9];
    double vr1 = z[dp], vr2 = z[dp + 1], e_fd = z[dp + 2], vr = z[vri], vi = z[vri + 1];
    double vm = sqrt(vr * vr + vi * vi); if (vm == 0.0) vm = 1e-12;
    double vs = vsi >= 0 ? z[vsi] : 0.0;
    double sat = (sb == 0.0 || e_fd <= sa) ? 0.0 : sb * (e_fd - sa) * (e_fd - sa);
    f[dp] = (Ka * (vref - vm + vs - vr2 - (Kf / Tf) * e_fd) - vr1) / Ta;
    f[dp + 1] = -((Kf / Tf) * e_fd + vr2) / Tf;
    f[dp + 2] = (vr1 - Ke * e_fd - sat) / Te;
  }
  for (int k = 0; k < s->ndev[GP_ZIPLOAD]; k++) {   /* src/kernels/zipload.jl:33-85 */
    if (s->zon && !s->zon[k]) continue;
    int bus = D(6, k, 0);
    const double *P = p + D(6, k, 1);
    int vri = npt + 2 * bus;
    double vr = z[vri], vi = z[vri + 1], pl = P[0], ql = P[1], al = P[2], yre = al * P[7], yim = al * P[8];
    double vm2 = vr * vr + vi * vi, den = vm2 > 0.2 ? vm2 : 0.2;
    f[vri] -= vr * yre - vi * yim;
    f[vri + 1] -= vr * yim + vi * yre;
    f[vri] -= (1.0 - al) * (pl * vr - ql * vi) / den;
    f[vri + 1] -= (1.0 - al) * (ql * vr + pl * vi) / den;
  }
  for (int k = 0; k < s->ndev[GP_STATICGEN]; k++) { /* src/kernels/static_gen.jl:107-148 */
    int bus = D(7, k, 0), ap = D(7, k, 1), bt = D(7, k, 3);
    const double *P = p + D(7, k, 2);
    int vri = npt + 2 * bus;
    double vr = z[vri], vi = z[vri + 1];
    double vm2r = vr * vr + vi * vi, vm2 = vm2r > 0.2 ? vm2r : 0.2;
    double pv, qv;
    if (bt == 2) { pv = P[0]; qv = z[ap]; } else if (bt == 3) { pv = z[ap]; qv = z[ap + 1]; } else { pv = P[0]; qv = P[1]; }
    f[vri] += (pv * vr + qv * vi) / vm2;
    f[vri + 1] += (pv * vi - qv * vr) / vm2;
    if (bt == 2) f[ap] = vr * vr + vi * vi - P[2] * P[2];
    else if (bt == 3) { f[ap] = vr - P[2] * cos(P[3]); f[ap + 1] = vi - P[2] * sin(P[3]); }
  }
  for (int b = 0; b < s->nbus; b++)                  /* src/dynamics.jl:314-327 */
    if (fault_g && fault_g[b] != 0.0) {
      f[npt + 2 * b] -= fault_g[b] * z[npt + 2 * b];
      f[npt + 2 * b + 1] -= fault_g[b] * z[npt + 2 * b + 1];
    }
}

/* beuler (src/dynamics.jl:366-391): diff rows become z - zold - dt*f */
static void beuler(const orc_sys *s, const double *z, const double *zold, const double *p,
                   const double *fault_g, double dt, double *f) {
  rhs(s, z, p, fault_g, f);
  for (int i = 0; i < s->net_ptr; i++)
    if (s->is_diff[i]) f[i] = z[i] - zold[i] - dt * f[i];
}

/* ------------------------------------------------------------ Jacobian */
static void rhs_jac(const orc_sys *s, const double *z, const double *p, const double *fault_g, double *J) {
  int npt = s->net_ptr;
  memset(J, 0, s->nnz * sizeof(double));
  for (int r = 0; r < 2 * s->nbus; r++)              /* current_injection_jacobian!, src/dynamics.jl:865-893 */
    for (int q = s->yrp[r]; q < s->yrp[r + 1]; q++) jset(s, J, npt + r, npt + s->yci[q], -s->yv[q]);
  for (int k = 0; k < s->ndev[GP_GENROU]; k++) {    /* src/kernels/genrou.jl:407-559 */
    int dp = D(0, k, 0), ap = D(0, k, 1), pp = D(0, k, 2), bus = D(0, k, 3), pm = D(0, k, 4), efd = D(0, k, 5);
    const double *P = p + pp;
    double x_d = P[0], x_q = P[1], x_dp = P[2], x_qp = P[3], x_ddp = P[4], xl = P[5], H = P[6], Dm = P[7];
    double T_d0p = P[8], T_q0p = P[9], T_d0dp = P[10], T_q0dp = P[11], S1 = P[12], S2 = P[13], x_qdp = x_ddp;
    double e_qp = z[dp], e_dp = z[dp + 1], phi_1d = z[dp + 2], phi_2q = z[dp + 3], delta = z[dp + 5];
    double i_q = z[ap + 2], i_d = z[ap + 3];
    int eq = dp, ed = dp + 1, p1 = dp + 2, p2 = dp + 3, w = dp + 4, de = dp + 5;
    int vq = ap, vd = ap + 1, iq = ap + 2, id = ap + 3, vr = npt + 2 * bus, vi = vr + 1;
    double sd = sin(delta), cd = cos(delta);
    double psi_de = (x_ddp - xl) / (x_dp - xl) * e_qp + (x_dp - x_ddp) / (x_dp - xl) * phi_1d;
    double psi_qe = -(x_ddp - xl) / (x_qp - xl) * e_dp + (x_qp - x_ddp) / (x_qp - xl) * phi_2q;
    double sa, sb; sat_coeff(S1, S2, &sa, &sb);
    double psi2 = sqrt(psi_de * psi_de + psi_qe * psi_qe);
    double Se = sat_se(psi2, sa, sb), dSe;
    if (sb == 0.0 || psi2 <= sa || psi2 == 0.0) dSe = 0.0;
    else { double g = psi2 - sa; dSe = sb * (2.0 * g * psi2 - g * g) / (psi2 * psi2); }
    double dpsi_de = psi2 == 0.0 ? 0.0 : psi_de / psi2, dpsi_qe = psi2 == 0.0 ? 0.0 : psi_qe / psi2;
    double a_eq = (x_ddp - xl) / (x_dp - xl), a_p1 = (x_dp - x_ddp) / (x_dp - xl);
    double a_ed = -(x_ddp - xl) / (x_qp - xl), a_p2 = (x_qp - x_ddp) / (x_qp - xl);
    double dT_de = -(dSe * dpsi_de * psi_de + Se), dT_qe = -(dSe * dpsi_qe * psi_de);
    double kd = (x_dp - xl) * (x_dp - xl), kq = (x_qp - xl) * (x_qp - xl);
    jset(s, J, dp, eq, (-(x_d - x_dp) * (-x_ddp + x_dp) / kd - 1) / T_d0p + dT_de * a_eq / T_d0p);
    jset(s, J, dp, p1, (x_d - x_dp) * (-x_ddp + x_dp) / kd / T_d0p + dT_de * a_p1 / T_d0p);
    jset(s, J, dp, id, -(x_d - x_dp) * (-(-x_ddp + x_dp) / (x_dp - xl) + 1) / T_d0p);
    jset(s, J, dp, ed, dT_qe * a_ed / T_d0p);
    jset(s, J, dp, p2, dT_qe * a_p2 / T_d0p);
    if (efd >= 0) jset(s, J, dp, efd, 1.0 / T_d0p);
    jset(s, J, dp + 1, ed, (-(x_q - x_qp) * (-x_qdp + x_qp) / kq - 1) / T_q0p);
    jset(s, J, dp + 1, p2, -(x_q - x_qp) * (-x_qdp + x_qp) / kq / T_q0p);
    jset(s, J, dp + 1, iq, (x_q - x_qp) * (-(-x_qdp + x_qp) / (x_qp - xl) + 1) / T_q0p);
    jset(s, J, dp + 2, eq, 1.0 / T_d0dp);
    jset(s, J, dp + 2, p1, -1.0 / T_d0dp);
    jset(s, J, dp + 2, id, (-x_dp + xl) / T_d0dp);
    jset(s, J, dp + 3, ed, -1.0 / T_q0dp);
    jset(s, J, dp + 3, p2, -1.0 / T_q0dp);
    jset(s, J, dp + 3, iq, (-x_qp + xl) / T_q0dp);
    jset(s, J, dp + 4, eq, -0.5 * i_q * (x_ddp - xl) / (H * (x_dp - xl)));
    jset(s, J, dp + 4, ed, 0.5 * i_d * (-x_ddp + xl) / (H * (x_qp - xl)));
    jset(s, J, dp + 4, p1, -0.5 * i_q * (-x_ddp + x_dp) / (H * (x_dp - xl)));
    jset(s, J, dp + 4, p2, 0.5 * i_d * (-x_ddp + x_qp) / (H * (x_qp - xl)));
    jset(s, J, dp + 4, w, -0.5 * Dm / H);
    if (pm >= 0) jset(s, J, dp + 4, pm, 0.5 / H);
    jset(s, J, dp + 4, iq, 0.5 * (-e_qp * (x_ddp - xl) / (x_dp - xl) - phi_1d * (-x_ddp + x_dp) / (x_dp - xl)) / H);
    jset(s, J, dp + 4, id, 0.5 * (e_dp * (-x_ddp + xl) / (x_qp - xl) + phi_2q * (-x_ddp + x_qp) / (x_qp - xl)) / H);
    jset(s, J, dp + 5, w, TWOPI60);
    jset(s, J, ap, eq, -(x_ddp - xl) / (x_ddp * (x_dp - xl)));
    jset(s, J, ap, p1, -(-x_ddp + x_dp) / (x_ddp * (x_dp - xl)));
    jset(s, J, ap, vq, 1.0 / x_ddp);
    jset(s, J, ap, id, 1.0);
    jset(s, J, ap + 1, ed, -(-x_qdp + xl) / (x_qdp * (x_qp - xl)));
    jset(s, J, ap + 1, p2, -(-x_qdp + x_qp) / (x_qdp * (x_qp - xl)));
    jset(s, J, ap + 1, vd, -1.0 / x_qdp);
    jset(s, J, ap + 1, iq, 1.0);
    double zvr = z[vr], zvi = z[vi];
    jset(s, J, ap + 2, de, -zvr * cd - zvi * sd);
    jset(s, J, ap + 2, vd, 1.0);
    jset(s, J, ap + 2, vr, -sd);
    jset(s, J, ap + 2, vi, cd);
    jset(s, J, ap + 3, de, zvr * sd - zvi * cd);
    jset(s, J, ap + 3, vq, 1.0);
    jset(s, J, ap + 3, vr, -cd);
    jset(s, J, ap + 3, vi, -sd);
    jadd(s, J, vr, de, i_d * cd - i_q * sd);
    jadd(s, J, vr, iq, cd);
    jadd(s, J, vr, id, sd);
    jadd(s, J, vi, de, i_d * sd + i_q * cd);
    jadd(s, J, vi, iq, sd);
    jadd(s, J, vi, id, -cd);
  }
  for (int k = 0; k < s->ndev[GP_IEESGO]; k++) {    /* src/kernels/ieesgo.jl:184-224 */
    int dp = D(1, k, 0), ap = D(1, k, 1), w = D(1, k, 3);
    const double *P = p + D(1, k, 2);
    double T1 = P[0], T2 = P[1], T3 = P[2], T4 = P[3], T5 = P[4], T6 = P[5], K1 = P[6], K2 = P[7], K3 = P[8];
    jset(s, J, dp, dp, -1.0 / T1);
    if (w >= 0) jset(s, J, dp, w, K1 / T1);
    jset(s, J, dp + 1, dp, (1.0 - T2 / T3) / T3);
    jset(s, J, dp + 1, dp + 1, -1.0 / T3);
    jset(s, J, dp + 2, dp, -T2 / (T3 * T4));
    jset(s, J, dp + 2, dp + 1, -1.0 / T4);
    jset(s, J, dp + 2, dp + 2, -1.0 / T4);
    jset(s, J, dp + 3, dp + 2, K2 / T5);
    jset(s, J, dp + 3, dp + 3, -1.0 / T5);
    jset(s, J, dp + 4, dp + 3, K3 / T6);
    jset(s, J, dp + 4, dp + 4, -1.0 / T6);
    jset(s, J, ap, dp + 2, 1.0 - K2);
    jset(s, J, ap, dp + 3, 1.0 - K3);
    jset(s, J, ap, dp + 4, 1.0);
    jset(s, J, ap, ap, -1.0);
  }
  for (int k = 0; k < s->ndev[GP_TGOV1]; k++) {     /* src/kernels/tgov1.jl:116-146 */
    int dp = D(2, k, 0), ap = D(2, k, 1), w = D(2, k, 3);
    const double *P = p + D(2, k, 2);
    double R = P[0], T1 = P[1], T2 = P[4], T3 = P[5], DT = P[6], t23 = T2 / T3;
    jset(s, J, dp, dp, -1.0 / T3);
    jset(s, J, dp, dp + 1, (1.0 - t23) / T3);
    jset(s, J, dp + 1, dp + 1, -1.0 / T1);
    if (w >= 0) jset(s, J, dp + 1, w, -1.0 / (R * T1));
    jset(s, J, ap, dp, 1.0);
    jset(s, J, ap, dp + 1, t23);
    if (w >= 0) jset(s, J, ap, w, -DT);
    jset(s, J, ap, ap, -1.0);
  }
  for (int k = 0; k < s->ndev[GP_IEEEST]; k++) {    /* src/kernels/ieeest.jl:276-359 */
    int dp = D(5, k, 0), ap = D(5, k, 1), w = D(5, k, 3);
    const double *P = p + D(5, k, 2);
    double A1 = P[0], A2 = P[1], A3 = P[2], A4 = P[3], A5 = P[4], A6 = P[5];
    double T1 = P[6], T2 = P[7], T3 = P[8], T4 = P[9], T5 = P[10], T6 = P[11], KS = P[12];
    double rA = A6 / A4, y2_1 = rA, y2_2 = A5 - A3 * rA, y2_3 = 1.0 - rA;
    double r12 = T1 / T2, y3_1 = r12 * y2_1, y3_2 = r12 * y2_2, y3_3 = r12 * y2_3, y3_4 = 1.0 - r12;
    double r34 = T3 / T4, y4_1 = r34 * y3_1, y4_2 = r34 * y3_2, y4_3 = r34 * y3_3, y4_4 = r34 * y3_4, y4_5 = 1.0 - r34;
    jset(s, J, dp, dp, -A1 / A2);
    jset(s, J, dp, dp + 1, -1.0 / A2);
    if (w >= 0) jset(s, J, dp, w, 1.0 / A2);
    jset(s, J, dp + 1, dp, 1.0);
    jset(s, J, dp + 2, dp + 1, 1.0 / A4);
    jset(s, J, dp + 2, dp + 2, -A3 / A4);
    jset(s, J, dp + 2, dp + 3, -1.0 / A4);
    jset(s, J, dp + 3, dp + 2, 1.0);
    jset(s, J, dp + 4, dp + 1, y2_1 / T2);
    jset(s, J, dp + 4, dp + 2, y2_2 / T2);
    jset(s, J, dp + 4, dp + 3, y2_3 / T2);
    jset(s, J, dp + 4, dp + 4, -1.0 / T2);
    jset(s, J, dp + 5, dp + 1, y3_1 / T4);
    jset(s, J, dp + 5, dp + 2, y3_2 / T4);
    jset(s, J, dp + 5, dp + 3, y3_3 / T4);
    jset(s, J, dp + 5, dp + 4, y3_4 / T4);
    jset(s, J, dp + 5, dp + 5, -1.0 / T4);
    double y4[5] = {y4_1, y4_2, y4_3, y4_4, y4_5};
    for (int j = 0; j < 5; j++) jset(s, J, dp + 6, dp + 1 + j, KS * y4[j] / T6);
    jset(s, J, dp + 6, dp + 6, -1.0 / T6);
    double r56 = T5 / T6;
    for (int j = 0; j < 5; j++) jset(s, J, ap, dp + 1 + j, -r56 * KS * y4[j]);
    jset(s, J, ap, dp + 6, r56);
    jset(s, J, ap, ap, 1.0);
  }
  for (int k = 0; k < s->ndev[GP_SEXS]; k++) {      /* src/kernels/sexs.jl:118-156 */
    if (s->zon && !s->zon[s->ndev[GP_ZIPLOAD] + k]) continue;
    int dp = D(3, k, 0), vri = D(3, k, 2), vsi = D(3, k, 3);
    const double *P = p + D(3, k, 1);
    double TA_TB = P[0], TB = P[1], K = P[2], TE = P[3];
    double vr = z[vri], vi = z[vri + 1], vm = sqrt(vr * vr + vi * vi); if (vm == 0.0) vm = 1e-12;
    double dr = vr / vm, di = vi / vm;
    jset(s, J, dp, dp, -1.0 / TB);
    jset(s, J, dp, vri, -(1.0 - TA_TB) * dr / TB);
    jset(s, J, dp, vri + 1, -(1.0 - TA_TB) * di / TB);
    if (vsi >= 0) jset(s, J, dp, vsi, (1.0 - TA_TB) / TB);
    jset(s, J, dp + 1, dp, K / TE);
    jset(s, J, dp + 1, dp + 1, -1.0 / TE);
    jset(s, J, dp + 1, vri, -K * TA_TB * dr / TE);
    jset(s, J, dp + 1, vri + 1, -K * TA_TB * di / TE);
    if (vsi >= 0) jset(s, J, dp + 1, vsi, K * TA_TB / TE);
  }
  for (int k = 0; k < s->ndev[GP_ESDC1A]; k++) {    /* src/kernels/esdc1a.jl:150-196 */
    if (s->zon && !s->zon[s->ndev[GP_ZIPLOAD] + s->ndev[GP_SEXS] + k]) continue;
    int dp = D(4, k, 0), vri = D(4, k, 2), vsi = D(4, k, 3);
    const double *P = p + D(4, k, 1);
    double Ka = P[0], Ta = P[1], Kf = P[2], Tf = P[3], Ke = P[4], Te = P[5], sa = P[7], sb = P[8];
    double e_fd = z[dp + 2], vr = z[vri], vi = z[vri + 1], vm = sqrt(vr * vr + vi * vi); if (vm == 0.0) vm = 1e-12;
    double dsat = (sb == 0.0 || e_fd <= sa) ? 0.0 : 2.0 * sb * (e_fd - sa);
    jset(s, J, dp, dp, -1.0 / Ta);
    jset(s, J, dp, dp + 1, -Ka / Ta);
    jset(s, J, dp, dp + 2, -Ka * Kf / (Ta * Tf));
    jset(s, J, dp, vri, -Ka * (vr / vm) / Ta);
    jset(s, J, dp, vri + 1, -Ka * (vi / vm) / Ta);
    if (vsi >= 0) jset(s, J, dp, vsi, Ka / Ta);
    jset(s, J, dp + 1, dp + 1, -1.0 / Tf);
    jset(s, J, dp + 1, dp + 2, -Kf / (Tf * Tf));
    jset(s, J, dp + 2, dp, 1.0 / Te);
    jset(s, J, dp + 2, dp + 2, -(Ke + dsat) / Te);
  }
  for (int k = 0; k < s->ndev[GP_ZIPLOAD]; k++) {   /* src/kernels/zipload.jl:133-179 */
    if (s->zon && !s->zon[k]) continue;
    int bus = D(6, k, 0), vri = npt + 2 * bus;
    const double *P = p + D(6, k, 1);
    double vr = z[vri], vi = z[vri + 1], pl = P[0], ql = P[1], al = P[2], yre = P[7], yim = P[8];
    double vm2 = vr * vr + vi * vi;
    double a_rr = -al * yre, a_ri = al * yim, a_ir = -al * yim, a_ii = -al * yre, p_rr, p_ri, p_ir, p_ii;
    if (vm2 > 0.2) {
      double iv4 = 1.0 / (vm2 * vm2);
      p_rr = (1.0 - al) * ((pl * vr - ql * vi) * 2.0 * vr - pl * vm2) * iv4;
      p_ri = (1.0 - al) * ((pl * vr - ql * vi) * 2.0 * vi + ql * vm2) * iv4;
      p_ir = (1.0 - al) * ((ql * vr + pl * vi) * 2.0 * vr - ql * vm2) * iv4;
      p_ii = (1.0 - al) * ((ql * vr + pl * vi) * 2.0 * vi - pl * vm2) * iv4;
    } else {
      double it = 1.0 / 0.2;
      p_rr = (1.0 - al) * (-pl) * it; p_ri = (1.0 - al) * ql * it;
      p_ir = (1.0 - al) * (-ql) * it; p_ii = (1.0 - al) * (-pl) * it;
    }
    jadd(s, J, vri, vri, a_rr + p_rr);
    jadd(s, J, vri, vri + 1, a_ri + p_ri);
    jadd(s, J, vri + 1, vri, a_ir + p_ir);
    jadd(s, J, vri + 1, vri + 1, a_ii + p_ii);
  }
  for (int k = 0; k < s->ndev[GP_STATICGEN]; k++) { /* src/kernels/static_gen.jl:168-220 */
    int bus = D(7, k, 0), ap = D(7, k, 1), bt = D(7, k, 3), vri = npt + 2 * bus;
    const double *P = p + D(7, k, 2);
    double vr = z[vri], vi = z[vri + 1], vm2r = vr * vr + vi * vi, vm2 = vm2r > 0.2 ? vm2r : 0.2, pv, qv;
    if (bt == 2) { pv = P[0]; qv = z[ap]; } else if (bt == 3) { pv = z[ap]; qv = z[ap + 1]; } else { pv = P[0]; qv = P[1]; }
    double irn = pv * vr + qv * vi, iin = pv * vi - qv * vr, drr, dri, dir, dii;
    if (vm2r > 0.2) {
      double vm4 = vm2 * vm2;
      drr = pv / vm2 - 2.0 * vr * irn / vm4; dri = qv / vm2 - 2.0 * vi * irn / vm4;
      dir = -qv / vm2 - 2.0 * vr * iin / vm4; dii = pv / vm2 - 2.0 * vi * iin / vm4;
    } else { drr = pv / vm2; dri = qv / vm2; dir = -qv / vm2; dii = pv / vm2; }
    jadd(s, J, vri, vri, drr); jadd(s, J, vri, vri + 1, dri);
    jadd(s, J, vri + 1, vri, dir); jadd(s, J, vri + 1, vri + 1, dii);
    if (bt == 2) {
      jset(s, J, ap, vri, 2.0 * vr); jset(s, J, ap, vri + 1, 2.0 * vi);
      jadd(s, J, vri, ap, vi / vm2); jadd(s, J, vri + 1, ap, -vr / vm2);
    } else if (bt == 3) {
      jset(s, J, ap, vri, 1.0); jset(s, J, ap + 1, vri + 1, 1.0);
      jadd(s, J, vri, ap, vr / vm2); jadd(s, J, vri + 1, ap, vi / vm2);
      jadd(s, J, vri, ap + 1, vi / vm2); jadd(s, J, vri + 1, ap + 1, -vr / vm2);
    }
  }
  for (int b = 0; b < s->nbus; b++)                  /* src/dynamics.jl:844-857 */
    if (fault_g && fault_g[b] != 0.0) {
      jadd(s, J, npt + 2 * b, npt + 2 * b, -fault_g[b]);
      jadd(s, J, npt + 2 * b + 1, npt + 2 * b + 1, -fault_g[b]);
    }
}

/* beuler_jac (src/dynamics.jl:418-500): diff rows scaled by -dt, +1 on the diagonal */
static void beuler_jac(const orc_sys *s, const double *z, const double *p, const double *fault_g,
                       double dt, double *J) {
  rhs_jac(s, z, p, fault_g, J);
  for (int c = 0; c < s->n; c++)
    for (int q = s->cp[c]; q < s->cp[c + 1]; q++) {
      int r = s->ri[q];
      if (r < s->net_ptr && s->is_diff[r]) {
        J[q] *= -dt;
        if (r == c) J[q] += 1.0;
      }
    }
}

/* ------------------------------------------------------------ sparse LU */
/* Minimum-degree ordering on the symmetric pattern of A + A^T with an
 * explicit elimination graph (fill is small for these matrices). */
typedef struct { int *a; int n, cap; } ivec;
static void iv_push(ivec *v, int x) { if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 8; v->a = realloc(v->a, v->cap * sizeof(int)); } v->a[v->n++] = x; }
static int iv_has(const ivec *v, int x) { for (int i = 0; i < v->n; i++) if (v->a[i] == x) return 1; return 0; }
static void iv_del(ivec *v, int x) { for (int i = 0; i < v->n; i++) if (v->a[i] == x) { v->a[i] = v->a[--v->n]; return; } }

int *orc_min_degree(int n, const int *cp, const int *ri) {
  ivec *g = xmalloc(n * sizeof(ivec));
  for (int c = 0; c < n; c++)
    for (int q = cp[c]; q < cp[c + 1]; q++) {
      int r = ri[q];
      if (r == c) continue;
      if (!iv_has(&g[c], r)) iv_push(&g[c], r);
      if (!iv_has(&g[r], c)) iv_push(&g[r], c);
    }
  int *done = xmalloc(n * sizeof(int)), *order = xmalloc(n * sizeof(int));
  /* bucket by degree with lazy updates */
  for (int k = 0; k < n; k++) {
    int best = -1, bd = 1 << 30;
    for (int v = 0; v < n; v++) if (!done[v] && g[v].n < bd) { bd = g[v].n; best = v; }
    order[k] = best;
    done[best] = 1;
    ivec *nb = &g[best];
    for (int i = 0; i < nb->n; i++) {
      int u = nb->a[i];
      iv_del(&g[u], best);
      for (int j = 0; j < nb->n; j++) {
        int w = nb->a[j];
        if (w != u && !iv_has(&g[u], w)) iv_push(&g[u], w);
      }
    }
    free(nb->a); nb->a = NULL; nb->n = nb->cap = 0;
  }
  for (int v = 0; v < n; v++) free(g[v].a);
  free(g); free(done);
  return order;
}

typedef struct {
  int n;
  int *lp, *li; double *lx;   /* L by columns (unit diagonal implied), row indices are pivot positions */
  int *up, *ui; double *ux;   /* U by columns, rows = pivot positions, diagonal stored last */
  int *pinv;                  /* pinv[row] = pivot position */
  int lcap, ucap;
} orc_lu;

static void lu_free(orc_lu *f) { free(f->lp); free(f->li); free(f->lx); free(f->up); free(f->ui); free(f->ux); free(f->pinv); memset(f, 0, sizeof(*f)); }

/* Left-looking LU with threshold partial pivoting (Gilbert-Peierls). */
static int lu_factor(int n, const int *cp, const int *ri, const double *ax, const int *q, orc_lu *F) {
  lu_free(F);
  F->n = n;
  F->lcap = 4 * (cp[n] + n); F->ucap = 4 * (cp[n] + n);
  F->lp = xmalloc((n + 1) * sizeof(int)); F->li = xmalloc(F->lcap * sizeof(int)); F->lx = xmalloc(F->lcap * sizeof(double));
  F->up = xmalloc((n + 1) * sizeof(int)); F->ui = xmalloc(F->ucap * sizeof(int)); F->ux = xmalloc(F->ucap * sizeof(double));
  F->pinv = xmalloc(n * sizeof(int));
  int *pinv = F->pinv;
  for (int i = 0; i < n; i++) pinv[i] = -1;
  double *x = xmalloc(n * sizeof(double));
  int *xi = xmalloc(2 * n * sizeof(int)), *mark = xmalloc(n * sizeof(int)), *stk = xmalloc(n * sizeof(int)), *pst = xmalloc(n * sizeof(int));
  int *prow = xmalloc(n * sizeof(int)); /* prow[k] = original row pivotal at step k */
  int lnz = 0, unz = 0;
  for (int i = 0; i < n; i++) mark[i] = -1;
  for (int k = 0; k < n; k++) {
    F->lp[k] = lnz; F->up[k] = unz;
    if (lnz + n > F->lcap) { F->lcap = 2 * F->lcap + n; F->li = realloc(F->li, F->lcap * sizeof(int)); F->lx = realloc(F->lx, F->lcap * sizeof(double)); }
    if (unz + n > F->ucap) { F->ucap = 2 * F->ucap + n; F->ui = realloc(F->ui, F->ucap * sizeof(int)); F->ux = realloc(F->ux, F->ucap * sizeof(double)); }
    int col = q[k];
    /* reach: DFS from the rows of A(:,col) through the graph of L.  During
     * the factorisation L's row indices are original rows; the columns of
     * L for pivots jp < k are complete at li[lp[jp] .. lp[jp+1]). */
    int top = n;
    for (int p = cp[col]; p < cp[col + 1]; p++) {
      int r = ri[p];
      if (mark[r] == k) continue;
      int head = 0;
      stk[0] = r;
      mark[r] = k;
      pst[0] = pinv[r] >= 0 ? F->lp[pinv[r]] : 0;
      while (head >= 0) {
        int j = stk[head], jp = pinv[j], pushed = 0;
        if (jp >= 0) {
          int end = F->lp[jp + 1];
          for (int pp = pst[head]; pp < end; pp++) {
            int i = F->li[pp];
            if (mark[i] != k) {
              pst[head] = pp + 1;
              stk[++head] = i;
              mark[i] = k;
              pst[head] = pinv[i] >= 0 ? F->lp[pinv[i]] : 0;
              pushed = 1;
              break;
            }
          }
        }
        if (!pushed) { head--; xi[--top] = j; }
      }
    }
    /* numeric: x = L \ A(:,col) over the reach in topological order */
    for (int p = top; p < n; p++) x[xi[p]] = 0.0;
    for (int p = cp[col]; p < cp[col + 1]; p++) x[ri[p]] = ax[p];
    for (int p = top; p < n; p++) {
      int j = xi[p], jp = pinv[j];
      if (jp < 0) continue;
      double xj = x[j];
      for (int pp = F->lp[jp]; pp < F->lp[jp + 1]; pp++) x[F->li[pp]] -= F->lx[pp] * xj;
    }
    /* pivot: largest among non-pivotal rows; keep diagonal if within threshold */
    int ipiv = -1; double amax = -1.0;
    for (int p = top; p < n; p++) {
      int i = xi[p];
      if (pinv[i] < 0) { double t = fabs(x[i]); if (t > amax) { amax = t; ipiv = i; } }
      else { F->ui[unz] = pinv[i]; F->ux[unz++] = x[i]; }
    }
    if (ipiv < 0 || amax <= 0.0) { free(x); free(xi); free(mark); free(stk); free(pst); free(prow); return -1; }
    if (pinv[col] < 0 && fabs(x[col]) >= 1e-3 * amax) ipiv = col;
    double piv = x[ipiv];
    F->ui[unz] = k; F->ux[unz++] = piv;
    pinv[ipiv] = k; prow[k] = ipiv;
    for (int p = top; p < n; p++) {
      int i = xi[p];
      if (pinv[i] < 0) { F->li[lnz] = i; F->lx[lnz++] = x[i] / piv; }
      x[i] = 0.0;
    }
  }
  F->lp[n] = lnz; F->up[n] = unz;
  /* convert L row indices from original rows to pivot positions */
  for (int p = 0; p < lnz; p++) F->li[p] = pinv[F->li[p]];
  free(x); free(xi); free(mark); free(stk); free(pst); free(prow);
  return 0;
}

/* x = A \ b using the factorisation (b in original row space, x in original column space) */
static void lu_solve(const orc_lu *F, const int *q, const double *b, double *x, double *w) {
  int n = F->n;
  for (int i = 0; i < n; i++) w[F->pinv[i]] = b[i];
  for (int k = 0; k < n; k++) {           /* L y = Pb */
    double yk = w[k];
    for (int p = F->lp[k]; p < F->lp[k + 1]; p++) w[F->li[p]] -= F->lx[p] * yk;
  }
  for (int k = n - 1; k >= 0; k--) {      /* U x = y, diagonal is last in column k */
    int pd = F->up[k + 1] - 1;
    w[k] /= F->ux[pd];
    double xk = w[k];
    for (int p = F->up[k]; p < pd; p++) w[F->ui[p]] -= F->ux[p] * xk;
  }
  for (int k = 0; k < n; k++) x[q[k]] = w[k];
}

/* ------------------------------------------------------------ public API */
void *orc_create(const gp_system_desc *d) {
  orc_sys *s = xmalloc(sizeof(orc_sys));
  s->n = d->n; s->nbus = d->nbus; s->net_ptr = d->net_ptr; s->npar = d->n_par;
  s->is_diff = xmalloc(d->n); memcpy(s->is_diff, d->is_diff, d->n);
  s->yrp = idup(d->y_rowptr, 2 * d->nbus + 1);
  s->yci = idup(d->y_colind, d->y_nnz);
  s->yv = xmalloc(d->y_nnz * sizeof(double)); memcpy(s->yv, d->y_val, d->y_nnz * sizeof(double));
  s->ynnz = d->y_nnz;
  static const int cols[GP_NDEVTYPES] = {GP_COLS_GENROU, GP_COLS_IEESGO, GP_COLS_TGOV1, GP_COLS_SEXS,
                                         GP_COLS_ESDC1A, GP_COLS_IEEEST, GP_COLS_ZIPLOAD, GP_COLS_STATICGEN};
  for (int t = 0; t < GP_NDEVTYPES; t++) { s->ndev[t] = d->n_dev[t]; s->dev[t] = idup(d->dev[t], (size_t)d->n_dev[t] * cols[t]); }
  s->u0 = xmalloc((2 * d->n_dev[GP_GENROU] + 2) * sizeof(double));
  memcpy(s->u0, d->genrou_u0, 2 * d->n_dev[GP_GENROU] * sizeof(double));
  s->ncl = d->n_clusters;
  s->cl = idup(d->clusters, (size_t)d->n_clusters * GP_CLUSTER_COLS);
  build_pattern(s);
  s->q = orc_min_degree(s->n, s->cp, s->ri);
  return s;
}

void orc_destroy(void *h) {
  orc_sys *s = h;
  if (!s) return;
  free(s->is_diff); free(s->yrp); free(s->yci); free(s->yv);
  for (int t = 0; t < GP_NDEVTYPES; t++) free(s->dev[t]);
  free(s->u0); free(s->cl); free(s->cp); free(s->ri); free(s->q);
  free(s->tr_fr); free(s->tr_to); free(s->tr_on); free(s->tr_off); free(s->tr_y);
  free(s->drop_k); free(s->drop_step); free(s);
}

/* Device disconnects of the following sweeps (DisconnectDeviceEvent,
 * src/GradPower.jl:130-143,681-698): device k[i] goes offline for good at
 * the end of step step[i]; k indexes the ZIP table's rows, then the SEXS
 * rows (nzip + row), then the ESDC1A rows (nzip + nsexs + row).  An offline
 * device is skipped by the residual and Jacobian (the `online` checks of
 * src/kernels/); an exciter's rows of F are then zero, as in the GPU
 * extension's zeroed F (ext:1340), so backward Euler holds its states. */
void orc_set_stop_on_fail(void *h, int flag) { ((orc_sys *)h)->stop_on_fail = flag != 0; }

void orc_set_drops(void *h, int n, const int *k, const int *step) {
  orc_sys *s = h;
  free(s->drop_k); free(s->drop_step);
  s->ndrop = n;
  s->drop_k = idup(k, n);
  s->drop_step = idup(step, n);
}

/* Line trips of the following sweeps (TripLineEvent, src/GradPower.jl:140-167):
 * n rows of (from, to, step_trip, step_reclose; 0 = never) with the line's
 * stamp y[8*i..] = yff, yft, ytf, ytt as (re, im) (create_trip_line_event,
 * src/GradPower.jl:509-542).  n = 0 clears them. */
void orc_set_trips(void *h, int n, const int *fr, const int *to, const int *on, const int *off, const double *y) {
  orc_sys *s = h;
  free(s->tr_fr); free(s->tr_to); free(s->tr_on); free(s->tr_off); free(s->tr_y);
  s->ntr = n;
  s->tr_fr = idup(fr, n); s->tr_to = idup(to, n); s->tr_on = idup(on, n); s->tr_off = idup(off, n);
  s->tr_y = xmalloc(8 * (size_t)n * sizeof(double) + 8);
  if (n) memcpy(s->tr_y, y, 8 * (size_t)n * sizeof(double));
}

/* Add sign * y (complex) to the 2x2 block (bus rb, bus cb) of the realified
 * Y, [re -im; im re] (_add_complex_to_realified! / _subtract_complex_from_
 * realified!, src/GradPower.jl:580-623). */
static void y_block_add(orc_sys *s, int rb, int cb, const double *yy, double sign) {
  for (int rr = 0; rr < 2; rr++)
    for (int cc = 0; cc < 2; cc++) {
      int r = 2 * rb + rr, c = 2 * cb + cc, q;
      double v = rr == cc ? yy[0] : (rr == 0 ? -yy[1] : yy[1]);
      for (q = s->yrp[r]; q < s->yrp[r + 1]; q++) if (s->yci[q] == c) break;
      if (q == s->yrp[r + 1]) abort();
      s->yv[q] += sign * v;
    }
}

/* Take trip t out of Y (sign -1, _apply_trip_line!, src/GradPower.jl:556-568)
 * or put it back (+1, _apply_reclose_line!, :570-578), in place. */
static void trip_apply(orc_sys *s, int t, double sign) {
  const double *y = s->tr_y + 8 * t;
  int a = s->tr_fr[t], b = s->tr_to[t];
  y_block_add(s, a, a, y, sign);
  y_block_add(s, a, b, y + 2, sign);
  y_block_add(s, b, a, y + 4, sign);
  y_block_add(s, b, b, y + 6, sign);
}

int orc_nnz(void *h) { return ((orc_sys *)h)->nnz; }
void orc_pattern(void *h, int *cp, int *ri) {
  orc_sys *s = h;
  memcpy(cp, s->cp, (s->n + 1) * sizeof(int));
  memcpy(ri, s->ri, s->nnz * sizeof(int));
}
void orc_residual(void *h, const double *z, const double *zold, const double *p,
                  const double *fault_g, double dt, double *f) {
  beuler((orc_sys *)h, z, zold, p, fault_g, dt, f);
}
void orc_jacobian(void *h, const double *z, const double *p, const double *fault_g, double dt, double *J) {
  beuler_jac((orc_sys *)h, z, p, fault_g, dt, J);
}

/* Solve J x = b with the oracle's sparse LU (test hook). Returns 0 on success. */
int orc_linsolve(void *h, const double *J, const double *b, double *x) {
  orc_sys *s = h;
  orc_lu F = {0};
  if (lu_factor(s->n, s->cp, s->ri, J, s->q, &F)) return -1;
  double *w = xmalloc(s->n * sizeof(double));
  lu_solve(&F, s->q, b, x, w);
  free(w); lu_free(&F);
  return 0;
}

/* newton_step! (src/numerics.jl:55-176): z <- z - J^{-1} F until |F|inf < tol.
 * Returns the number of iterations that computed a correction, or -1 when
 * itermax was exhausted (the reference keeps the last iterate either way),
 * -2 on a singular factorisation.  *fnorm gets the last |F|inf. */
static int newton(const orc_sys *s, double *z, const double *zold, const double *p, const double *fault_g,
                  double dt, double tol, int itermax, double *f, double *J, double *dx, double *w,
                  orc_lu *F, double *fnorm) {
  for (int it = 0; it < itermax; it++) {
    beuler(s, z, zold, p, fault_g, dt, f);
    double nf = 0.0;
    for (int i = 0; i < s->n; i++) { double a = fabs(f[i]); if (!(a <= nf)) nf = a; }
    *fnorm = nf;
    if (nf < tol) return it;
    if (!isfinite(nf)) return -3;
    beuler_jac(s, z, p, fault_g, dt, J);
    if (lu_factor(s->n, s->cp, s->ri, J, s->q, F)) return -2;
    lu_solve(F, s->q, f, dx, w);
    for (int i = 0; i < s->n; i++) z[i] -= dx[i];
  }
  beuler(s, z, zold, p, fault_g, dt, f);
  double nf = 0.0;
  for (int i = 0; i < s->n; i++) { double a = fabs(f[i]); if (!(a <= nf)) nf = a; }
  *fnorm = nf;
  return nf < tol ? itermax : -1;
}

/* integrate! (src/dynamics.jl:502-655) for one scenario.
 * events: n_ev rows of (bus, r_fault, step_on, step_off).
 * traj: (nsteps+1) x n, row k = state after step k (row 0 = z0), or NULL.
 * fgs: (nsteps+1) x nbus, row k = fault conductances in force during step k
 * (row 0 unused), or NULL.
 * stats[0] = Newton iterations, stats[1] = non-converged solves, stats[2] = max accepted |F|,
 * stats[3] = steps completed before the first Newton solve that did not
 * converge (k-1 for the solve of step k, k for its event re-solve; -1: none). */
static int sweep(orc_sys *s, const double *z0, const double *p, double dt, int nsteps,
                 const int *ev_bus, const double *ev_r, const int *ev_on, const int *ev_off, int n_ev,
                 double tol, int itermax, double *traj, double *fgs, double *ys, uint8_t *zons, double *zfinal,
                 double *stats) {
  /* the line trips change Y in place during the sweep, as integrate! does
   * (src/dynamics.jl:616-624); Y is restored when the sweep ends */
  double *y_saved = xmalloc((size_t)s->ynnz * sizeof(double) + 8);
  memcpy(y_saved, s->yv, (size_t)s->ynnz * sizeof(double));
  int nz = s->ndev[GP_ZIPLOAD] + s->ndev[GP_SEXS] + s->ndev[GP_ESDC1A];   /* devices that can go offline */
  uint8_t *zon = xmalloc((size_t)nz + 1);
  memset(zon, 1, (size_t)nz + 1);
  s->zon = zon;
  int n = s->n;
  double *z = xmalloc(n * sizeof(double)), *zold = xmalloc(n * sizeof(double));
  double *f = xmalloc(n * sizeof(double)), *dx = xmalloc(n * sizeof(double)), *w = xmalloc(n * sizeof(double));
  double *J = xmalloc(s->nnz * sizeof(double)), *fg = xmalloc(s->nbus * sizeof(double));
  int *active = xmalloc((n_ev + 1) * sizeof(int));
  orc_lu F = {0};
  memcpy(z, z0, n * sizeof(double));
  if (traj) memcpy(traj, z, n * sizeof(double));
  double iters = 0, fails = 0, maxf = 0.0, fn, first_fail = -1.0;
  int rc = 0;
  for (int k = 1; k <= nsteps; k++) {
    memcpy(zold, z, n * sizeof(double));
    for (int b = 0; b < s->nbus; b++) fg[b] = 0.0;
    for (int e = 0; e < n_ev; e++) if (active[e]) fg[ev_bus[e]] += 1.0 / ev_r[e];
    if (fgs) memcpy(fgs + (size_t)k * s->nbus, fg, s->nbus * sizeof(double));
    if (ys) memcpy(ys + (size_t)k * s->ynnz, s->yv, (size_t)s->ynnz * sizeof(double));
    if (zons) memcpy(zons + (size_t)k * (nz + 1), zon, (size_t)nz + 1);
    int it = newton(s, z, zold, p, fg, dt, tol, itermax, f, J, dx, w, &F, &fn);
    if (it < 0) { fails++; if (first_fail < 0) first_fail = k - 1; if (it <= -2) { rc = it; } } else iters += it;
    if (it < 0 && s->stop_on_fail) break;
    if (fn > maxf) maxf = fn;
    if (traj) memcpy(traj + (size_t)k * n, z, n * sizeof(double));
    int any = 0;
    for (int e = 0; e < n_ev; e++) {
      if (ev_on[e] == k) { active[e] = 1; any = 1; }
      if (ev_off[e] == k) { active[e] = 0; any = 1; }
    }
    for (int d = 0; d < s->ndrop; d++)   /* disconnects (src/dynamics.jl:533-535,612-615) */
      if (s->drop_step[d] == k) { zon[s->drop_k[d]] = 0; any = 1; }
    for (int t = 0; t < s->ntr; t++) {   /* trip, then reclose (src/dynamics.jl:536-540) */
      if (s->tr_on[t] == k) { trip_apply(s, t, -1.0); any = 1; }
      if (s->tr_off[t] == k) { trip_apply(s, t, 1.0); any = 1; }
    }
    if (any) {
      memcpy(zold, z, n * sizeof(double));
      for (int b = 0; b < s->nbus; b++) fg[b] = 0.0;
      for (int e = 0; e < n_ev; e++) if (active[e]) fg[ev_bus[e]] += 1.0 / ev_r[e];
      it = newton(s, z, zold, p, fg, 0.0, tol, itermax, f, J, dx, w, &F, &fn);
      if (it < 0) { fails++; if (first_fail < 0) first_fail = k; if (it <= -2) rc = it; } else iters += it;
      if (it < 0 && s->stop_on_fail) break;
      if (fn > maxf) maxf = fn;
    }
    if (rc) break;
  }
  if (zfinal) memcpy(zfinal, z, n * sizeof(double));
  if (stats) { stats[0] = iters; stats[1] = fails; stats[2] = maxf; stats[3] = first_fail; }
  memcpy(s->yv, y_saved, (size_t)s->ynnz * sizeof(double));
  s->zon = NULL;
  lu_free(&F);
  free(z); free(zold); free(f); free(dx); free(w); free(J); free(fg); free(active); free(y_saved); free(zon);
  return rc;
}

int orc_integrate(void *h, const double *z0, const double *p, double dt, int nsteps,
                  const int *ev_bus, const double *ev_r, const int *ev_on, const int *ev_off, int n_ev,
                  double tol, int itermax, double *traj, double *zfinal, double *stats) {
  return sweep((orc_sys *)h, z0, p, dt, nsteps, ev_bus, ev_r, ev_on, ev_off, n_ev, tol, itermax, traj, NULL, NULL,
               NULL, zfinal, stats);
}

/* y = A^{-T} c with the factorisation of A (P A Q = L U): Q U^T L^T P y = c,
 * so v = Q^T c, U^T u = v (forward), L^T t = u (backward), y = P^T t. */
static void lu_solve_t(const orc_lu *F, const int *q, const double *c, double *y, double *w) {
  int n = F->n;
  for (int k = 0; k < n; k++) w[k] = c[q[k]];
  for (int k = 0; k < n; k++) {           /* U^T: row k of U^T is column k of U */
    int pd = F->up[k + 1] - 1;
    double t = w[k];
    for (int p = F->up[k]; p < pd; p++) t -= F->ux[p] * w[F->ui[p]];
    w[k] = t / F->ux[pd];
  }
  for (int k = n - 1; k >= 0; k--) {      /* L^T, unit diagonal */
    double t = w[k];
    for (int p = F->lp[k]; p < F->lp[k + 1]; p++) t -= F->lx[p] * w[F->li[p]];
    w[k] = t;
  }
  for (int i = 0; i < n; i++) y[i] = w[F->pinv[i]];
}

/* Solve J^T y = c with the oracle's sparse LU (test hook). Returns 0 on success. */
int orc_linsolve_t(void *h, const double *J, const double *c, double *y) {
  orc_sys *s = h;
  orc_lu F = {0};
  if (lu_factor(s->n, s->cp, s->ri, J, s->q, &F)) return -1;
  double *w = xmalloc(s->n * sizeof(double));
  lu_solve_t(&F, s->q, c, y, w);
  free(w); lu_free(&F);
  return 0;
}

/* Discrete adjoint of the backward-Euler sweep of one scenario: the
 * GradPower.adjoint of examples/adjoint.jl:34-51,106 (called there, absent
 * from the reference snapshot), z0 held fixed when p moves.
 * Objective  J = scale sum_{k=1..nsteps} dt sum_r (z_k[rows[r]] - off)^2
 * over the converged states z_k of the forward sweep (traj rows 1..nsteps,
 * test/test_adjoint_event_offset.jl:58-69).  For k = nsteps..1, with the
 * Jacobian, fault state, Y (line trips) and online loads of forward step k:
 *   J_k^T lam_k = dJ/dz_k + lam_{k+1} on the differential rows
 *   grad_j     -= lam_k . dF_k/dp_j
 * where dF_k/dp_j is a central difference of the residual in p_j (step
 * 1e-6 |p_j|, or 1e-6 when p_j = 0; the zold terms cancel).  dz0 = lam_1 on the
 * differential rows.  fault_shift (0 for the adjoint proper) takes the fault
 * state of step k + fault_shift instead: a planted misalignment for tests.
 * Returns the forward sweep's status. */
int orc_adjoint(void *h, const double *z0, const double *p, double dt, int nsteps,
                const int *ev_bus, const double *ev_r, const int *ev_on, const int *ev_off, int n_ev,
                double tol, int itermax, int n_rows, const int *rows, double scale, double off,
                int n_par, const int *par, int fault_shift, double *value, double *grad, double *dz0) {
  orc_sys *s = h;
  int n = s->n, nb = s->nbus;
  double *traj = xmalloc((size_t)(nsteps + 1) * n * sizeof(double));
  double *fgs = xmalloc((size_t)(nsteps + 1) * nb * sizeof(double));
  double *ys = xmalloc((size_t)(nsteps + 1) * s->ynnz * sizeof(double) + 8);
  double *y_saved = xmalloc((size_t)s->ynnz * sizeof(double) + 8);
  memcpy(y_saved, s->yv, (size_t)s->ynnz * sizeof(double));
  int nz = s->ndev[GP_ZIPLOAD] + s->ndev[GP_SEXS] + s->ndev[GP_ESDC1A];   /* devices that can go offline */
  uint8_t *zons = xmalloc((size_t)(nsteps + 1) * (nz + 1));
  double st[4];
  int rc = sweep(s, z0, p, dt, nsteps, ev_bus, ev_r, ev_on, ev_off, n_ev, tol, itermax, traj, fgs, ys, zons, NULL, st);
  if (rc) { free(traj); free(fgs); free(ys); free(y_saved); free(zons); return rc; }
  double *lam = xmalloc(n * sizeof(double)), *a = xmalloc(n * sizeof(double)), *w = xmalloc(n * sizeof(double));
  double *J = xmalloc(s->nnz * sizeof(double)), *fp = xmalloc(n * sizeof(double)), *fm = xmalloc(n * sizeof(double));
  double *pp = xmalloc((s->npar + 1) * sizeof(double));
  memcpy(pp, p, s->npar * sizeof(double));
  orc_lu F = {0};
  double val = 0.0;
  for (int k = 1; k <= nsteps; k++)
    for (int r = 0; r < n_rows; r++) { double e = traj[(size_t)k * n + rows[r]] - off; val += scale * dt * e * e; }
  for (int j = 0; j < n_par; j++) grad[j] = 0.0;
  for (int i = 0; i < n; i++) lam[i] = 0.0;
  for (int k = nsteps; k >= 1; k--) {
    int kf = k + fault_shift;   /* fault_shift != 0 only plants a misalignment in tests */
    kf = kf < 1 ? 1 : (kf > nsteps ? nsteps : kf);
    const double *zk = traj + (size_t)k * n, *zkm = traj + (size_t)(k - 1) * n, *fg = fgs + (size_t)kf * nb;
    memcpy(s->yv, ys + (size_t)kf * s->ynnz, (size_t)s->ynnz * sizeof(double));   /* Y of step kf */
    s->zon = zons + (size_t)kf * (nz + 1);                                         /* its devices */
    for (int i = 0; i < n; i++) a[i] = s->is_diff[i] ? lam[i] : 0.0;
    for (int r = 0; r < n_rows; r++) a[rows[r]] += 2.0 * scale * dt * (zk[rows[r]] - off);
    beuler_jac(s, zk, p, fg, dt, J);
    if (lu_factor(n, s->cp, s->ri, J, s->q, &F)) { rc = -2; break; }
    lu_solve_t(&F, s->q, a, lam, w);
    for (int j = 0; j < n_par; j++) {
      int q = par[j];
      double hq = p[q] != 0.0 ? 1e-6 * fabs(p[q]) : 1e-6;
      pp[q] = p[q] + hq;
      beuler(s, zk, zkm, pp, fg, dt, fp);
      pp[q] = p[q] - hq;
      beuler(s, zk, zkm, pp, fg, dt, fm);
      pp[q] = p[q];
      double g = 0.0;
      for (int i = 0; i < n; i++) g += lam[i] * (fp[i] - fm[i]) / (2.0 * hq);
      grad[j] -= g;
    }
  }
  if (value) *value = val;
  if (dz0) for (int i = 0; i < n; i++) dz0[i] = s->is_diff[i] ? lam[i] : 0.0;
  memcpy(s->yv, y_saved, (size_t)s->ynnz * sizeof(double));
  s->zon = NULL;
  lu_free(&F);
  free(traj); free(fgs); free(ys); free(y_saved); free(zons); free(lam); free(a); free(w); free(J); free(fp); free(fm); free(pp);
  return rc;
}

/* ------------------------------------------------------- Schur step */
/* Dense LU with partial pivoting, column-major n x n (src/schur.jl:436-493). */
static void dlu_factor(double *A, int *ipiv, int n) {
  for (int k = 0; k < n; k++) {
    double mv = fabs(A[k + k * n]); int mr = k;
    for (int i = k + 1; i < n; i++) { double v = fabs(A[i + k * n]); if (v > mv) { mv = v; mr = i; } }
    ipiv[k] = mr;
    if (mr != k) for (int j = 0; j < n; j++) { double t = A[k + j * n]; A[k + j * n] = A[mr + j * n]; A[mr + j * n] = t; }
    double akk = A[k + k * n];
    if (akk != 0.0) {
      double inv = 1.0 / akk;
      for (int i = k + 1; i < n; i++) {
        A[i + k * n] *= inv;
        for (int j = k + 1; j < n; j++) A[i + j * n] -= A[i + k * n] * A[k + j * n];
      }
    }
  }
}
static void dlu_solve(const double *A, const int *ipiv, double *b, int n) {
  for (int k = 0; k < n; k++) { int p = ipiv[k]; if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; } }
  for (int k = 0; k < n; k++) for (int i = k + 1; i < n; i++) b[i] -= A[i + k * n] * b[k];
  for (int k = n - 1; k >= 0; k--) { b[k] /= A[k + k * n]; for (int i = 0; i < k; i++) b[i] -= A[i + k * n] * b[k]; }
}

/* One Schur-reduced Newton correction (src/schur.jl:253-426): given F and J
 * at z, computes dz with J dz = -F by eliminating every non-trivial cluster
 * block.  Reduced system = trivial-cluster states + all bus voltages. */
int orc_schur_step(void *h, const double *J, const double *f, double *dz) {
  orc_sys *s = h;
  int n = s->n, npt = s->net_ptr;
  int *g2r = xmalloc(n * sizeof(int)), *red = xmalloc(n * sizeof(int));
  int nr = 0;
  for (int i = 0; i < n; i++) g2r[i] = -1;
  for (int c = 0; c < s->ncl; c++) {
    const int *cl = s->cl + c * GP_CLUSTER_COLS;
    if (cl[GP_CL_TRIVIAL] || cl[GP_CL_WSIZE] == 0)
      for (int j = 0; j < cl[GP_CL_WSIZE]; j++) g2r[cl[GP_CL_WSTART] + j] = 1;
  }
  for (int i = npt; i < n; i++) g2r[i] = 1;
  for (int i = 0; i < n; i++) if (g2r[i] > 0) { g2r[i] = nr; red[nr++] = i; } else g2r[i] = -1;
  /* dense reduced matrix (test sizes only) */
  double *S = xmalloc((size_t)nr * nr * sizeof(double)), *rhs = xmalloc(nr * sizeof(double));
  for (int c = 0; c < n; c++) if (g2r[c] >= 0)
    for (int q = s->cp[c]; q < s->cp[c + 1]; q++) if (g2r[s->ri[q]] >= 0) S[g2r[s->ri[q]] + (size_t)g2r[c] * nr] = J[q];
  for (int i = 0; i < nr; i++) rhs[i] = f[red[i]];
  int wmax = 1;
  for (int c = 0; c < s->ncl; c++) if (s->cl[c * GP_CLUSTER_COLS + GP_CL_WSIZE] > wmax) wmax = s->cl[c * GP_CLUSTER_COLS + GP_CL_WSIZE];
  double *A = xmalloc(wmax * wmax * sizeof(double)), *B = xmalloc(2 * wmax * sizeof(double)), *Cm = xmalloc(2 * wmax * sizeof(double));
  double *t = xmalloc(wmax * sizeof(double));
  int *ipiv = xmalloc(wmax * sizeof(int));
  double **keepA = xmalloc(s->ncl * sizeof(double *)), **keepB = xmalloc(s->ncl * sizeof(double *));
  int **keepP = xmalloc(s->ncl * sizeof(int *));
  for (int c = 0; c < s->ncl; c++) {
    const int *cl = s->cl + c * GP_CLUSTER_COLS;
    if (cl[GP_CL_TRIVIAL] || cl[GP_CL_WSIZE] == 0) continue;
    int ws = cl[GP_CL_WSTART], wk = cl[GP_CL_WSIZE], vr = npt + 2 * cl[GP_CL_BUS], vi = vr + 1;
    memset(A, 0, wk * wk * sizeof(double)); memset(B, 0, 2 * wk * sizeof(double)); memset(Cm, 0, 2 * wk * sizeof(double));
    for (int j = 0; j < wk; j++)
      for (int q = s->cp[ws + j]; q < s->cp[ws + j + 1]; q++) {
        int r = s->ri[q];
        if (r >= ws && r < ws + wk) A[(r - ws) + j * wk] = J[q];
        else if (r == vr) Cm[0 + 2 * j] = J[q];
        else if (r == vi) Cm[1 + 2 * j] = J[q];
      }
    for (int cc = 0; cc < 2; cc++)
      for (int q = s->cp[vr + cc]; q < s->cp[vr + cc + 1]; q++) {
        int r = s->ri[q];
        if (r >= ws && r < ws + wk) B[(r - ws) + cc * wk] = J[q];
      }
    dlu_factor(A, ipiv, wk);
    double Dk[4];
    for (int cc = 0; cc < 2; cc++) {
      for (int i = 0; i < wk; i++) t[i] = B[i + cc * wk];
      dlu_solve(A, ipiv, t, wk);
      double d0 = 0, d1 = 0;
      for (int i = 0; i < wk; i++) { d0 += Cm[2 * i] * t[i]; d1 += Cm[2 * i + 1] * t[i]; }
      Dk[0 + 2 * cc] = d0; Dk[1 + 2 * cc] = d1;
    }
    int lr = g2r[vr], li = g2r[vi];
    S[lr + (size_t)lr * nr] -= Dk[0]; S[li + (size_t)lr * nr] -= Dk[1];
    S[lr + (size_t)li * nr] -= Dk[2]; S[li + (size_t)li * nr] -= Dk[3];
    for (int i = 0; i < wk; i++) t[i] = f[ws + i];
    dlu_solve(A, ipiv, t, wk);
    for (int i = 0; i < wk; i++) { rhs[lr] -= Cm[2 * i] * t[i]; rhs[li] -= Cm[2 * i + 1] * t[i]; }
    keepA[c] = xmalloc(wk * wk * sizeof(double)); memcpy(keepA[c], A, wk * wk * sizeof(double));
    keepB[c] = xmalloc(2 * wk * sizeof(double)); memcpy(keepB[c], B, 2 * wk * sizeof(double));
    keepP[c] = xmalloc(wk * sizeof(int)); memcpy(keepP[c], ipiv, wk * sizeof(int));
  }
  for (int i = 0; i < nr; i++) rhs[i] = -rhs[i];
  int *sp = xmalloc(nr * sizeof(int));
  dlu_factor(S, sp, nr);
  dlu_solve(S, sp, rhs, nr);
  for (int i = 0; i < n; i++) dz[i] = 0.0;
  for (int i = 0; i < nr; i++) dz[red[i]] = rhs[i];
  for (int c = 0; c < s->ncl; c++) {
    const int *cl = s->cl + c * GP_CLUSTER_COLS;
    if (cl[GP_CL_TRIVIAL] || cl[GP_CL_WSIZE] == 0) continue;
    int ws = cl[GP_CL_WSTART], wk = cl[GP_CL_WSIZE], vr = npt + 2 * cl[GP_CL_BUS];
    for (int i = 0; i < wk; i++) t[i] = f[ws + i] + keepB[c][i] * dz[vr] + keepB[c][i + wk] * dz[vr + 1];
    dlu_solve(keepA[c], keepP[c], t, wk);
    for (int i = 0; i < wk; i++) dz[ws + i] = -t[i];
    free(keepA[c]); free(keepB[c]); free(keepP[c]);
  }
  free(g2r); free(red); free(S); free(rhs); free(A); free(B); free(Cm); free(t); free(ipiv);
  free(keepA); free(keepB); free(keepP); free(sp);
  return 0;
}

// gp_host.cpp -- C-ABI runtime of the batched backward-Euler DAE integrator.
//
// Owns the context: device copies of the system tables, the Jacobian and
// Schur-complement patterns with their slot/position tables, the symbolic
// LU analyses, and the host time loop that drives the kernels
// (integrate!/integrate_gpu_schur_cudss! semantics: src/dynamics.jl:502-655,
// ext:3008-3080, with per-scenario faults and horizons).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gradpower_hip.h"
#include "gp_internal.hpp"
#include "gp_splu.hpp"

using namespace gp;

namespace {

thread_local std::string g_err;

struct GpError {
  int code;
  std::string msg;
};

#define HIPCHK(x)                                                                                 \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess)                                                                         \
      throw GpError{GP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)};                  \
  } while (0)

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  bool owned = true;
  // a window into memory owned elsewhere (never freed here)
  void view(void *base, size_t count) {
    release();
    p = static_cast<T *>(base);
    n = count;
    owned = false;
  }
  void alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw GpError{GP_ERR_OOM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed"};
    }
  }
  // allocate unless a buffer of exactly this size is already held
  void ensure(size_t count) {
    if (!p || n != count) alloc(count);
  }
  void upload(const std::vector<T> &v) {
    alloc(v.size() ? v.size() : 1);
    if (v.size()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    n = 0;
    owned = true;
  }
  ~DBuf() { release(); }
};

// Pinned host memory (page-locked: copies to and from it are plain DMA in
// stream order).
struct HBuf {
  uint8_t *p = nullptr;
  void alloc(size_t bytes) {
    release();
    if (hipHostMalloc(reinterpret_cast<void **>(&p), bytes, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      throw GpError{GP_ERR_OOM, "hipHostMalloc of " + std::to_string(bytes) + " bytes failed"};
    }
    std::memset(p, 0, bytes);
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
  }
  ~HBuf() { release(); }
};

// Per-round exchange of the scenario-asynchronous time loop: one block in
// device memory and its pinned host mirror (L = padded scenario count, a
// multiple of 64).  The accepted steps of round r (what the kernels that
// record and advance them read) come first, then what round r + 1's
// evaluation reads, then what it decides:
//   [acc L][adv L][col 4L][sv L][svslot 4L][ck L][ckslot 4L][pred L]      (P = 17L)
//   [first L][refac0 L][exact L][dt 8L][work L][counts 16][failed L][areq L]
// so that one copy uploads [0, P + 13L + 16) (the failed flags, which only
// the evaluation sets, go up cleared) and one copy brings back
// [P + 11L, P + 14L + 16).
struct RoundLayout {
  size_t L;
  explicit RoundLayout(size_t l) : L(l) {}
  size_t acc() const { return 0; }
  size_t adv() const { return L; }
  size_t col() const { return 2 * L; }
  size_t sv() const { return 6 * L; }
  size_t svslot() const { return 7 * L; }
  size_t ck() const { return 11 * L; }
  size_t ckslot() const { return 12 * L; }
  size_t pred() const { return 16 * L; }
  size_t P() const { return 17 * L; }
  size_t first() const { return P(); }
  size_t refac0() const { return P() + L; }
  size_t exact() const { return P() + 2 * L; }
  size_t dt() const { return P() + 3 * L; }
  size_t work() const { return P() + 11 * L; }
  size_t cnt() const { return P() + 12 * L; }
  size_t failed() const { return P() + 12 * L + 16; }
  size_t areq() const { return P() + 13 * L + 16; }
  size_t bytes() const { return P() + 14 * L + 16; }
  size_t up_bytes() const { return P() + 13 * L + 16; }
  size_t down_off() const { return P() + 11 * L; }
  size_t down_bytes() const { return 3 * L + 16; }
};

struct LUPlan {
  SpluAnalysis A;
  SpluDev dev{};
  DBuf<int> a2lu, a_src, op_loc, op_sl, st_o, flp, fac_mode, crec, lu2s, dmap, fw_lvl, fw_chunk, fw_cols, fw_rows,
      bw_lvl, bw_chunk, bw_cols, bw_rows, q, pinv, cmp_ph, cmp_tp, cmp_dst, cmp_ops;
  // staged-solve plans, one per factor precision (their steps hold
  // sol_cv(elem) values per thread)
  struct Sol {
    DBuf<int> fh, fb, bh, bb;
    int nf = 0, nb = 0;
  } sol[2];   // [0] FP32, [1] FP64
  // LDS-resident factorisation plans, one per factor precision (their
  // steps are cut at the precision's value staging capacity)
  struct Res {
    ResPlan rp;
    bool ok = false;
    DBuf<int> hdr, blob, a2x, omap, xsrc;
  } res[2];   // [0] FP32, [1] FP64
  // LU values per scenario the factorisation kernels use
  size_t lu_elems() const {
    size_t e = (size_t)A.nnz_lu;
    for (const Res &r : res)
      if (r.ok) e = std::max<size_t>(e, r.rp.xtot);
    return e;
  }
  bool ready = false;
  int trip_gen = -1;                              // trip set its pivots were chosen for
  double dt = -1.0;                               // J^T plan: dt of its pivot matrix
  static constexpr int kThreads = 512;
  static constexpr int kLdsBudget = 120 * 1024;   // staging / dense block
  static constexpr int kLdsMax = 160 * 1024;      // LDS of one CU

  static std::vector<int> nonempty(const std::vector<int> &v) { return v.empty() ? std::vector<int>{0} : v; }
  // LDS bytes of the solve's x (16-byte aligned)
  int x_bytes() const { return (int)(((size_t)A.n * 8 + 15) / 16 * 16); }
  // Partitioned solves: up to part_k levels per partition (0 or 1: one
  // level per step, unpartitioned); on when x, y and the staging area fit LDS
  int part_k = 4;
  bool part_tune = true;   // the autotuner also picks part_k (gp_options.solve_levels 0)
  bool part_on = false;
  int x_vecs() const { return part_on ? 2 : 1; }
  // LDS bytes of a solve step's staging area (blob, values, diagonals)
  static long long stage_bytes(int elem) {
    return (long long)kSolCB * kThreads * 4 + (long long)(sol_cv(elem) + kSolCD) * kThreads * elem;
  }
  static long long stage_max() { return std::max(stage_bytes(4), stage_bytes(8)); }
  bool part_fits() const { return !gx_needed() && 2LL * x_bytes() + stage_max() + 16384 <= kLdsMax; }
  // The solve keeps x in global memory when x and the staging area of FP64
  // values do not fit in LDS together (above about 12k rows)
  bool force_gx = false;   // gp_options.global_x
  bool gx_needed() const {
    return force_gx || (long long)x_bytes() + stage_max() > kLdsMax;
  }

  // Largest dense trailing block the solve's LDS allows (x, 8 bytes per row,
  // stays resident beside it unless it lives in global memory).
  int nd_bound() const {
    const long long room =
        std::min<long long>(kLdsBudget, (long long)kLdsMax - (gx_needed() ? 0LL : 8LL * x_vecs() * A.n) - 16384);
    int nd_max = 0;
    while ((long long)(nd_max + 16) * (nd_max + 16) * 8 <= room) nd_max += 16;
    return nd_max;
  }
  // Schedules and uploads the analysis, the dense block at most nd_cap rows
  // (and what the solve's LDS allows).
  void upload(const std::vector<int> &asrc, int nd_cap = 1 << 30) {
    part_on = part_k > 1 && part_fits();
    splu_schedule(A, kThreads, std::min(nd_bound(), nd_cap), part_on ? part_k : 0);
    if (!gx_needed() && (long long)x_vecs() * x_bytes() + (long long)A.nd * A.nd * 8 > kLdsMax)
      throw GpError{GP_ERR_UNSUPPORTED, "system too large for the in-LDS triangular solve (" +
                                            std::to_string(A.n) + " rows)"};
    a2lu.upload(A.a2lu); a_src.upload(asrc);
    std::vector<int> pl = A.op_loc, ps = A.op_sl;   // one padding entry: clamped reads stay in bounds
    pl.push_back(0);
    ps.push_back(0);
    op_loc.upload(pl); op_sl.upload(ps); st_o.upload(A.st_o);
    flp.upload(A.flp); fac_mode.upload(nonempty(A.fac_mode)); crec.upload(nonempty(A.crec));
    lu2s.upload(A.lu2s); dmap.upload(nonempty(A.dmap));
    fw_lvl.upload(nonempty(A.sf.lvl)); fw_chunk.upload(nonempty(A.sf.chunk));
    fw_cols.upload(nonempty(A.sf.cols)); fw_rows.upload(nonempty(A.sf.rows));
    bw_lvl.upload(nonempty(A.sb.lvl)); bw_chunk.upload(nonempty(A.sb.chunk));
    bw_cols.upload(nonempty(A.sb.cols)); bw_rows.upload(nonempty(A.sb.rows));
    q.upload(A.q); pinv.upload(A.pinv);
    cmp_ph.upload(nonempty(A.cmp_ph)); cmp_tp.upload(nonempty(A.cmp_tp));
    cmp_dst.upload(nonempty(A.cmp_dst)); cmp_ops.upload(nonempty(A.cmp_ops));
    // the global-memory solve reads the level tables where they lie; the
    // LDS one runs a staged plan per factor precision (shape() picks it)
    for (int pr = 0; pr < 2; pr++) {
      Sol &P = sol[pr];
      P.nf = P.nb = 0;
      if (gx_needed()) continue;
      const int elem = pr ? 8 : 4;
      SolvePlan pf, pb;
      if (!solve_plan(A.sf, false, 0, kSolCB * kThreads, sol_cv(elem) * kThreads, kSolCD * kThreads, pf) ||
          !solve_plan(A.sb, true, 0, kSolCB * kThreads, sol_cv(elem) * kThreads, kSolCD * kThreads, pb))
        throw GpError{GP_ERR_UNSUPPORTED, "triangular-solve staging plan failed"};
      P.fh.upload(nonempty(pf.hdr));
      P.fb.upload(nonempty(pf.blob));
      P.bh.upload(nonempty(pb.hdr));
      P.bb.upload(nonempty(pb.blob));
      P.nf = pf.nsteps();
      P.nb = pb.nsteps();
    }
    for (int pr = 0; pr < 2; pr++) {
      Res &R = res[pr];
      R.ok = res_plan(A, kResCB * kThreads, res_cx(pr ? 8 : 4) * kThreads, R.rp);
      if (!R.ok) continue;
      const ResPlan &rp = R.rp;
      R.hdr.upload(rp.hdr);
      R.blob.upload(rp.blob);
      R.a2x.upload(rp.a2x);
      R.omap.upload(rp.omap);
      // initial value source of each step-ordered slot (four spare slots:
      // the staged loads read whole groups of four)
      std::vector<int> xs((size_t)rp.xtot + 4, kResNoSrc);
      for (int e = 0; e < A.nnz_a; e++) xs[rp.a2x[e]] = asrc[e];
      R.xsrc.upload(xs);
    }
    SpluDev &d = dev;
    d.n = A.n; d.nnz_a = A.nnz_a; d.nnz_lu = A.nnz_lu; d.ns = std::max(1, A.ns);
    d.n_flev = (int)A.flp.size() - 1; d.nd = A.nd;
    d.qb = A.lu_cp[A.n - A.nd]; d.nnz_top = A.lu_cp[A.n] - d.qb;
    d.a2lu = a2lu.p; d.a_src = a_src.p; d.op_loc = op_loc.p; d.op_sl = op_sl.p; d.st_o = st_o.p;
    d.flp = flp.p; d.fac_mode = fac_mode.p; d.crec = crec.p; d.lu2s = lu2s.p; d.dmap = dmap.p;
    d.n_fwlev = (int)A.sf.lvl.size() / 4; d.n_bwlev = (int)A.sb.lvl.size() / 4;
    d.n_fwch = (int)A.sf.chunk.size() / 2; d.n_bwch = (int)A.sb.chunk.size() / 2;
    d.fw_nvals = A.sf.nvals; d.bw_nvals = A.sb.nvals;
    d.fw_lvl = fw_lvl.p; d.fw_chunk = fw_chunk.p; d.fw_cols = fw_cols.p; d.fw_rows = fw_rows.p;
    d.bw_lvl = bw_lvl.p; d.bw_chunk = bw_chunk.p; d.bw_cols = bw_cols.p; d.bw_rows = bw_rows.p;
    d.q = q.p; d.pinv = pinv.p;
    d.part = A.part > 1 ? 1 : 0;
    d.cmp_nph = (int)A.cmp_ph.size() - 1;
    d.cmp_ph = cmp_ph.p; d.cmp_tp = cmp_tp.p; d.cmp_dst = cmp_dst.p; d.cmp_ops = cmp_ops.p;
    d.res_xsrc = nullptr;   // set by shape() with the precision's plan
    d.threads = kThreads;
    d.gx = gx_needed() ? 1 : 0;
    d.gx_ws = nullptr;   // the context's [n][ldM] workspace, set after the plan is made
    // levels of at least kWideLanes lanes take a launch of their own: one
    // workgroup would need more than four passes of its 512 threads
    d.gx_fw_wide = d.gx_bw_wide = 0;
    d.h_fw_lvl = A.sf.lvl.data();
    d.h_bw_lvl = A.sb.lvl.data();
    if (d.gx) {
      constexpr int kWideLanes = 2048;
      while (d.gx_fw_wide < d.n_fwlev && A.sf.lvl[4 * d.gx_fw_wide] >= kWideLanes) d.gx_fw_wide++;
      while (d.gx_bw_wide < d.n_bwlev && A.sb.lvl[4 * (d.n_bwlev - 1 - d.gx_bw_wide)] >= kWideLanes) d.gx_bw_wide++;
    }
  }
  // LDS staging sizes for element size `elem`; columns (or op lists) larger
  // than the staging buffers run in place in global memory.
  void shape(size_t elem) {
    const int W = kThreads / 64;
    int tc = std::min<long long>(A.t_cols_max, (long long)kLdsBudget / ((long long)kThreads * (long long)elem));
    int wc = A.w_col_max, wo = A.w_ops_max;
    auto wbytes = [&](int c, int o) { return (int)(((size_t)c * elem + (size_t)o * (elem + 4) + 15) / 16 * 16); };
    while ((wc > 0 || wo > 0) && (long long)W * wbytes(wc, wo) > kLdsBudget) {
      wc = wc * 3 / 4;
      wo = wo * 3 / 4;
    }
    dev.t_cols = tc;
    dev.w_col = wc;
    dev.w_ops = wo;
    dev.w_bytes = wbytes(wc, wo);
    dev.lds_factor = std::max<int>(std::max<int>((int)((size_t)kThreads * tc * elem), W * dev.w_bytes),
                                   (int)((((size_t)A.nd * A.nd + 3) / 4 * 4 + 512) * elem));   // dense block, buffers
    // solve: x, then the staging area, which the dense block reuses; with x
    // in global memory only the dense block
    {
      const Sol &P = sol[elem == 8 ? 1 : 0];
      dev.sol_nfw = P.nf;
      dev.sol_nbw = P.nb;
      dev.sol_fw_hdr = P.nf ? P.fh.p : nullptr;
      dev.sol_fw_blob = P.nf ? P.fb.p : nullptr;
      dev.sol_bw_hdr = P.nb ? P.bh.p : nullptr;
      dev.sol_bw_blob = P.nb ? P.bb.p : nullptr;
    }
    dev.lds_solve = dev.gx ? (int)std::max<size_t>(16, (size_t)A.nd * A.nd * elem)
                           : x_vecs() * x_bytes() + (int)std::max<size_t>((size_t)stage_bytes((int)elem),
                                                                          (size_t)A.nd * A.nd * elem);
    if (dev.lds_solve > kLdsMax)
      throw GpError{GP_ERR_UNSUPPORTED, "system too large for the in-LDS triangular solve (" +
                                            std::to_string(A.n) + " rows)"};
    // LDS-resident factorisation when its multipliers (or the dense block),
    // staging blob and column values fit
    dev.res_nsteps = 0;
    dev.res_xsrc = nullptr;
    const Res &R = res[elem == 8 ? 1 : 0];
    if (R.ok) {
      const ResPlan &rp = R.rp;
      const int lv = std::max(rp.nL + 1, A.nd * A.nd);
      const size_t need = ((size_t)lv * elem + 15) / 16 * 16 + (size_t)kResCB * kThreads * 4 +
                          2 * (size_t)res_cx(elem) * kThreads * elem;
      if (need <= (size_t)kLdsMax) {
        dev.res_nsteps = rp.nsteps();
        dev.res_lv = lv;
        dev.res_nl = rp.nL;
        dev.res_xtot = rp.xtot;
        dev.res_hdr = R.hdr.p;
        dev.res_blob = R.blob.p;
        dev.res_a2x = R.a2x.p;
        dev.res_omap = R.omap.p;
        dev.res_xsrc = R.xsrc.p;
        dev.lds_factor = (int)need;
      }
    }
  }

  // LU-plan autotuner.  The schedule's cost model picks the dense trailing
  // block within a bound; the block's size also decides whether the
  // multipliers fit in LDS (the resident factorisation) and how long the
  // solves' two dense substitutions take, which the model does not see.  So
  // each bound in 16-row steps is planned and timed on the device -- one
  // factorisation and one solve of one scenario's matrix (var0: the value
  // array's column of scenario 0), median of 3 launches each -- and the bound
  // with the least factorisation + w_solve * solve time is kept; then, at
  // that bound, the levels per partition of the partitioned solves (1, 2, 4,
  // 6, 8) the same way.
  double tuned_us = 0.0;
  int tuned_cap = -1;
  template <typename TF>
  double time_one(const double *var, const double *cst, const double *rhs, double *x, const uint8_t *one, int *st,
                  double w_solve, hipStream_t s) {
    DBuf<TF> lu, sv, dn;
    DBuf<double> ws;
    if (dev.gx) {   // one scenario: ldM 1
      ws.alloc(A.n);
      dev.gx_ws = ws.p;
    }
    lu.alloc(lu_elems());
    sv.alloc(dev.ns);
    HIPCHK(hipMemsetAsync(sv.p, 0, sv.n * sizeof(TF), s));   // padding slots stay zero
    dn.alloc((size_t)std::max(1, A.nd * A.nd));
    auto fac = [&]() { launch_splu_factor_impl<TF>(dev, var, cst, 1, lu.p, sv.p, dn.p, one, nullptr, 1, st, s); };
    auto sol = [&]() {
      launch_splu_solve<TF>(dev, sv.p, dn.p, rhs, 1, 1.0, nullptr, nullptr, x, one, nullptr, 1, s);
    };
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto med = [&](const std::function<void()> &f) {
      float t[3];
      for (int r = 0; r < 3; r++) {
        HIPCHK(hipEventRecord(e0, s));
        f();
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&t[r], e0, e1));
      }
      std::sort(t, t + 3);
      return 1e3 * (double)t[1];
    };
    fac();
    sol();
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    const double tf = med(fac), ts = med(sol);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dev.gx_ws = nullptr;
    return tf + w_solve * ts;
  }
  // mode (gp_options.dense_block): 0 tune, < 0 the cost model alone, > 0 a
  // fixed bound of that many rows
  void tune(const std::vector<int> &asrc, const std::vector<double> &var0, const double *cst, size_t elem,
            double w_solve, hipStream_t s, int mode) {
    if (mode != 0) {
      upload(asrc, mode > 0 ? mode : 1 << 30);
      shape(elem);
      return;
    }
    DBuf<double> v1, r1, x1;
    DBuf<uint8_t> one;
    DBuf<int> st;
    v1.upload(var0);
    r1.upload(std::vector<double>(A.n, 1.0));
    x1.alloc(A.n);
    one.upload(std::vector<uint8_t>(1, 1));
    st.upload(std::vector<int>(1, 0));
    double best = 1e300;
    int best_cap = -1;
    std::vector<int> seen;
    std::string why;
    for (int cap = nd_bound(); cap >= 0; cap -= 16) {
      try {
        upload(asrc, cap);
        if (std::find(seen.begin(), seen.end(), A.nd) != seen.end()) continue;
        seen.push_back(A.nd);
        shape(elem);
      } catch (const GpError &e) {
        why = e.msg;
        continue;
      }
      const double t = elem == 4 ? time_one<float>(v1.p, cst, r1.p, x1.p, one.p, st.p, w_solve, s)
                                 : time_one<double>(v1.p, cst, r1.p, x1.p, one.p, st.p, w_solve, s);
      if (t < best) {
        best = t;
        best_cap = cap;
      }
    }
    if (best_cap < 0) throw GpError{GP_ERR_UNSUPPORTED, "no dense-block size fits the LU kernels (" + why + ")"};
    // then the levels per partition of the solves at that block size (the
    // composites' cost is the factorisation's, their benefit the solves')
    if (part_tune && part_fits()) {
      const int k0 = part_k;
      int best_k = k0;
      for (int k : {1, 2, 4, 6, 8}) {
        if (k == k0) continue;
        part_k = k;
        try {
          upload(asrc, best_cap);
          shape(elem);
        } catch (const GpError &) {
          continue;
        }
        const double t = elem == 4 ? time_one<float>(v1.p, cst, r1.p, x1.p, one.p, st.p, w_solve, s)
                                   : time_one<double>(v1.p, cst, r1.p, x1.p, one.p, st.p, w_solve, s);
        if (t < best) {
          best = t;
          best_k = k;
        }
      }
      part_k = best_k;
    }
    upload(asrc, best_cap);
    shape(elem);
    tuned_us = best;
    tuned_cap = best_cap;
  }
};

}  // namespace

struct gp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int n = 0, nbus = 0, net_ptr = 0, npar = 0, M = 0, ldM = 0;
  gp_options opt{};
  // host copies
  std::vector<uint8_t> is_diff;
  std::vector<int> yrp, yci;
  std::vector<double> yv;
  std::vector<int> dev[GP_NDEVTYPES];
  std::vector<double> u0;
  std::vector<int> cl;
  int ncl = 0;
  std::vector<int> cl_nt, cl_xoff;
  int xrows = 0;
  // monolithic Jacobian pattern (CSC) and per-device positions
  std::vector<int> jcp, jri;
  std::vector<int> jp[6], jp_bus, jp_yoff_p, jp_yoff_pos;
  std::vector<double> jp_yoff_val, ydiag;
  // reduced (Schur) system
  int nred = 0;
  std::vector<int> red_of_z, z_of_red, scp, sri, s_src, sv_bus;
  std::vector<double> s_cst;
  int n_svar = 0;
  // per-bus attachments
  std::vector<int> bus_cl_p, bus_cl, bus_zip_p, bus_zip, bus_sg;
  // events / horizons: faults of gp_set_events, then one record per line
  // trip (bus = -2 - t for trip record t, see FaultState)
  std::vector<gp_event> events, faults_user;
  std::vector<gp_line_trip> trips;
  std::vector<gp_device_drop> drops;   // ZIP loads and exciters taken offline
  int nexc = 0;                        // exciter drops among the events (FaultState::nexc)
  // line trips: records (from, to; stamp) and the off-diagonal Y entries
  // that trips change, kept once made variable (see DevTables)
  std::vector<int> tr_bus, tv_p, tv_i;
  std::vector<double> tr_y, tv_base;
  std::map<std::pair<int, int>, int> tv_of;   // (row z, column z) -> entry
  int trip_gen = 0;                             // bumped when the set of tripped lines changes
  std::vector<int> nsteps;
  std::vector<int> rec_idx;
  int nrec = 0;
  // device memory
  DBuf<uint8_t> d_is_diff;
  DBuf<int> d_dev[GP_NDEVTYPES], d_cl, d_yrp, d_yci, d_bus_cl_p, d_bus_cl, d_bus_zip_p, d_bus_zip, d_bus_sg;
  DBuf<int> d_jp[6], d_jp_bus, d_jp_yoff_p, d_jp_yoff_pos, d_cl_nt, d_cl_xoff, d_red_of_z, d_z_of_red, d_sv_bus;
  DBuf<int> d_rec_idx;
  DBuf<int> d_rec_p, d_rec_ch;   // recorded channels of each row (CSR; empty when every row is recorded)
  DBuf<double> d_yv, d_u0, d_jp_yoff_val, d_ydiag, d_s_cst;
  DBuf<int> d_tr_bus, d_tv_p, d_tv_i;
  DBuf<double> d_tr_y, d_tv_base;
  // every scenario's events in slot form ([MAXFAULT][ldM], the first
  // MAXFAULT events of each scenario, as the parity hooks apply them) and
  // the GPU's allocation-free scratch array (gp_eval_residual_noalloc)
  DBuf<int> d_evbus;
  DBuf<double> d_evg;
  void *scratch = nullptr;
  DBuf<double> d_zold2;   // z_(k-2) for the quadratic predictor
  DBuf<double> d_p, d_z, d_zold, d_f, d_inj, d_jv, d_svar, d_rhs, d_dv, d_Dg, d_fnorm, d_hist, d_luJ, d_luS64;
  DBuf<double> d_svJ, d_svS64, d_dnJ, d_dnS64;
  DBuf<float> d_luS32, d_svS32, d_dnS32, d_X32;
  DBuf<double> d_X64;
  DBuf<uint8_t> d_work, d_act, d_refac, d_refac0, d_sel;
  DBuf<uint8_t> d_round;   // per-round exchange block (RoundLayout); views point into it
  HBuf h_round;
  // a batch's requests and target buffers [req L][tgt L] per lane, uploaded
  // from a ring of pinned slots: each round synchronises the stream at least
  // once and launches at most kLanes batches, so a slot is reused only after
  // its copy has completed
  DBuf<uint8_t> d_batch;
  HBuf h_batch;
  int batch_slot = 0;
  // lane 0 factorises in the method's precision (measured at C2: a second
  // concurrent lane of those batches saves rounds but slows them as much);
  // lane 1 takes the FP32 method's FP64 factorisations of the scenarios
  // running exact Newton, so that they do not queue behind lane 0's batches
  static constexpr int kLanes = 2;
  static constexpr int kBatchSlots = 4 * kLanes;
  // asynchronous refactorisation (GP_REFACTOR_ASYNC): up to kLanes batches in
  // flight, each on a stream of its own, factorising into the buffers their
  // scenarios' solves do not read; a scenario is in at most one batch
  struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t ev_snap = nullptr, ev_done = nullptr;
    bool inflight = false;
    std::vector<uint8_t> batch;
    int nb = 0;                               // scenarios in the batch
    uint8_t *req = nullptr, *tgt = nullptr;   // in d_batch
    DBuf<double> snap;                        // the batch's matrix values
  } lanes[kLanes];
  DBuf<int> d_fbus, d_nunconv, d_status, d_nstall, d_iters;
  DBuf<double> d_fg, d_fbest, d_traj_tmp;
  DBuf<uint8_t> d_failed, d_upd, d_cur, d_pend, d_areq, d_sch;
  // asynchronous refactorisation (GP_REFACTOR_ASYNC): one batch at a time on
  // a second stream, into the buffer the solves do not read
  std::vector<uint8_t> cur, pend, queue;
  // kernel timing (gp_options.profile): event pairs around each launch
  struct KProf {
    std::vector<hipEvent_t> ev;
    std::vector<int64_t> scen;   // per pair: the scenarios its launch worked on (0: none, not counted)
    size_t used = 0;
    int64_t calls = 0;           // launches seen; every `every`-th one is timed
    int every = 1;
  } prof_solve, prof_factor;
  DBuf<unsigned long long> d_nb;
  // per working scenario, max |F| of each non-trivial cluster's rows and of
  // each bus's rows, written by the residual kernels ([n_cl_nt + nbus][ldM]);
  // used for the Newton max-norm when those rows are all n rows (pm_rows)
  DBuf<double> d_pm;
  bool pm_rows = false;
  DevTables T{};
  LUPlan luJ, luS;
  // the plans' representative matrices at analysis (scenario 0, dt = 0, no
  // event in force), values in the J (CSC) and S patterns
  std::vector<double> rep0_J, rep0_S;
  SpluDev luS64{};   // luS shaped for FP64 factors (the FP64 methods' and the FP32 method's exact retries)
  DBuf<double> d_gxJ, d_gxS, d_gxJT;   // solves' x in global memory [n][ldM] (large systems)
  // the FP32 method's FP64 exact-Newton kernels run beside the FP32 ones on
  // a stream of their own: forked after the round's precision split and
  // after its Schur assembly, joined before the next use of their outputs
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_side = nullptr, ev_fork2 = nullptr, ev_join = nullptr;
  hipEvent_t ev_round = nullptr;   // end of a round's exchange, polled by the host
  bool has64 = false;   // luS64 fits the LU kernels' LDS
  // shared-factor method (GP_METHOD_SCHUR_SHARED): the one FP64 factorisation
  // P of the reduced network (slot 0 of lu/sv/dn), the Woodbury columns
  // W = P^{-1} U of every faulted bus (U: unit columns at the bus's vr and vi
  // rows; W [column][nred] in the solve's pivot order), and per scenario the
  // columns, rows and inverse capacitance matrix of its faults in force
  struct Shared {
    DBuf<double> lu, sv, dn, W, h;
    DBuf<int> nc, col, pos;
    std::map<int, int> col_of_bus;   // faulted bus -> its first column (vr; vi next)
    std::vector<double> G;           // U^T W: G[i*ncol + j] = W[row of column i][j]
    std::vector<int> upos;           // pivot-order position of each column's row
    int ncol = 0, kw = 2;            // columns; per-scenario table width (2 x events in force)
    std::vector<int> hnc, hcol, hpos;
    std::vector<double> hh;
    WbTab tab() const {
      WbTab t;
      t.nc = nc.p; t.col = col.p; t.pos = pos.p; t.h = h.p; t.W = W.p; t.kw = kw;
      return t;
    }
  } sh;
  DBuf<uint8_t> d_split;   // a round's masks split by factor precision [6][L]
  // discrete adjoint: LU plan of J^T (values read from d_jv), its factors,
  // the initial state of the last gp_set_state, checkpoints, the states of
  // one segment, multipliers, parameter pullbacks and objective per scenario
  LUPlan luJT;
  DBuf<double> d_luJT, d_svJT, d_dnJT, d_z0, d_ckpt, d_seg, d_lam, d_gacc, d_val;
  DBuf<int> d_hz, d_objrows;
  DBuf<int2> d_ent;
  DBuf<uint8_t> d_bact;
  // scenario-asynchronous time loop: per-scenario step, first-round flags,
  // record columns, accepted / advancing masks
  DBuf<double> d_dtv;
  DBuf<uint8_t> d_first, d_acc, d_adv, d_sv, d_ck, d_exact, d_pred;
  DBuf<int> d_col, d_svslot, d_ckslot;
  bool state_set = false;
  int hist_steps = 0;
  std::vector<int> traj_len;
  std::vector<int> scen_status, scen_steps;   // per scenario, from the last gp_integrate
};

namespace {

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------- patterns
// Per-device Jacobian slot lists.  Each slot is (row, col) in z indices; the
// order matches the value order of the kernels (gp_kernels.hip) and covers
// the reference pattern of src/dynamics.jl:687-794 plus the backward-Euler
// diagonal of every differential state.
void genrou_slots(const int *g, int npt, std::vector<std::pair<int, int>> &s) {
  int dp = g[0], ap = g[1], bus = g[3], pm = g[4], efd = g[5];
  int eq = dp, ed = dp + 1, p1 = dp + 2, p2 = dp + 3, w = dp + 4, de = dp + 5;
  int vq = ap, vd = ap + 1, iq = ap + 2, id = ap + 3, vr = npt + 2 * bus, vi = vr + 1;
  s = {{dp, eq}, {dp, p1}, {dp, id}, {dp, ed}, {dp, p2}, {dp, efd},
       {dp + 1, ed}, {dp + 1, p2}, {dp + 1, iq},
       {dp + 2, eq}, {dp + 2, p1}, {dp + 2, id},
       {dp + 3, ed}, {dp + 3, p2}, {dp + 3, iq},
       {dp + 4, eq}, {dp + 4, ed}, {dp + 4, p1}, {dp + 4, p2}, {dp + 4, w}, {dp + 4, iq}, {dp + 4, id}, {dp + 4, pm},
       {dp + 5, w}, {dp + 5, de},
       {ap, eq}, {ap, p1}, {ap, vq}, {ap, id},
       {ap + 1, ed}, {ap + 1, p2}, {ap + 1, vd}, {ap + 1, iq},
       {ap + 2, de}, {ap + 2, vd}, {ap + 2, vr}, {ap + 2, vi},
       {ap + 3, de}, {ap + 3, vq}, {ap + 3, vr}, {ap + 3, vi},
       {vr, de}, {vr, iq}, {vr, id}, {vi, de}, {vi, iq}, {vi, id}};
}
void ieesgo_slots(const int *d, std::vector<std::pair<int, int>> &s) {
  int dp = d[0], ap = d[1], w = d[3];
  s = {{dp, dp}, {dp, w}, {dp + 1, dp}, {dp + 1, dp + 1}, {dp + 2, dp}, {dp + 2, dp + 1}, {dp + 2, dp + 2},
       {dp + 3, dp + 2}, {dp + 3, dp + 3}, {dp + 4, dp + 3}, {dp + 4, dp + 4},
       {ap, dp + 2}, {ap, dp + 3}, {ap, dp + 4}, {ap, ap}};
}
void tgov1_slots(const int *d, std::vector<std::pair<int, int>> &s) {
  int dp = d[0], ap = d[1], w = d[3];
  s = {{dp, dp}, {dp, dp + 1}, {dp + 1, dp + 1}, {dp + 1, w}, {ap, dp}, {ap, dp + 1}, {ap, w}, {ap, ap}};
}
void sexs_slots(const int *d, std::vector<std::pair<int, int>> &s) {
  int dp = d[0], vr = d[2], vi = vr + 1, vs = d[3];
  s = {{dp, dp}, {dp, vr}, {dp, vi}, {dp, vs}, {dp + 1, dp}, {dp + 1, dp + 1}, {dp + 1, vr}, {dp + 1, vi}, {dp + 1, vs}};
}
void esdc1a_slots(const int *d, std::vector<std::pair<int, int>> &s) {
  int dp = d[0], vr = d[2], vi = vr + 1, vs = d[3];
  s = {{dp, dp}, {dp, dp + 1}, {dp, dp + 2}, {dp, vr}, {dp, vi}, {dp, vs},
       {dp + 1, dp + 1}, {dp + 1, dp + 2}, {dp + 2, dp}, {dp + 2, dp + 2}};
}
void ieeest_slots(const int *d, std::vector<std::pair<int, int>> &s) {
  int dp = d[0], ap = d[1], w = d[3];
  s = {{dp, dp}, {dp, dp + 1}, {dp, w}, {dp + 1, dp}, {dp + 2, dp + 1}, {dp + 2, dp + 2}, {dp + 2, dp + 3},
       {dp + 3, dp + 2}};
  for (int j = 1; j <= 4; j++) s.push_back({dp + 4, dp + j});
  for (int j = 1; j <= 5; j++) s.push_back({dp + 5, dp + j});
  for (int j = 1; j <= 6; j++) s.push_back({dp + 6, dp + j});
  for (int j = 1; j <= 6; j++) s.push_back({ap, dp + j});
  s.push_back({ap, ap});
  s.push_back({dp + 1, dp + 1});
  s.push_back({dp + 3, dp + 3});
}
// static generator slots: vr_ap, vi_ap, vr_ap1, vi_ap1, ap_vr, ap_vi, ap1_vi
void sg_slots(const int *d, int npt, std::vector<std::pair<int, int>> &s) {
  int bus = d[0], ap = d[1], bt = d[3], vr = npt + 2 * bus, vi = vr + 1;
  s.assign(7, {-1, -1});
  if (bt == 2) { s[0] = {vr, ap}; s[1] = {vi, ap}; s[4] = {ap, vr}; s[5] = {ap, vi}; }
  if (bt == 3) { s[0] = {vr, ap}; s[1] = {vi, ap}; s[2] = {vr, ap + 1}; s[3] = {vi, ap + 1}; s[4] = {ap, vr}; s[6] = {ap + 1, vi}; }
}

struct Pattern {
  int n;
  std::vector<std::vector<int>> cols;  // rows per column
  explicit Pattern(int n_) : n(n_), cols(n_) {}
  void add(int r, int c) {
    if (r >= 0 && c >= 0) cols[c].push_back(r);
  }
  void finish(std::vector<int> &cp, std::vector<int> &ri) {
    cp.assign(n + 1, 0);
    ri.clear();
    for (int c = 0; c < n; c++) {
      auto &v = cols[c];
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      cp[c + 1] = cp[c] + (int)v.size();
      ri.insert(ri.end(), v.begin(), v.end());
    }
  }
};

int find_pos(const std::vector<int> &cp, const std::vector<int> &ri, int r, int c) {
  if (r < 0 || c < 0) return -1;
  auto b = ri.begin() + cp[c], e = ri.begin() + cp[c + 1];
  auto it = std::lower_bound(b, e, r);
  if (it == e || *it != r) return -2;
  return (int)(it - ri.begin());
}

void build_structures(gp_ctx *C) {
  const int npt = C->net_ptr, n = C->n, nbus = C->nbus;
  // ---- monolithic pattern
  Pattern P(n);
  for (int i = 0; i < npt; i++)
    if (C->is_diff[i]) P.add(i, i);
  for (int b = 0; b < nbus; b++)
    for (int u = 0; u < 2; u++)
      for (int v = 0; v < 2; v++) P.add(npt + 2 * b + u, npt + 2 * b + v);
  for (int r = 0; r < 2 * nbus; r++)
    for (int q = C->yrp[r]; q < C->yrp[r + 1]; q++) P.add(npt + r, npt + C->yci[q]);
  std::vector<std::pair<int, int>> sl;
  static const int ncols[6] = {GP_COLS_GENROU, GP_COLS_IEESGO, GP_COLS_TGOV1, GP_COLS_SEXS, GP_COLS_ESDC1A, GP_COLS_IEEEST};
  auto slots_of = [&](int t, const int *d) {
    switch (t) {
      case GP_GENROU: genrou_slots(d, npt, sl); break;
      case GP_IEESGO: ieesgo_slots(d, sl); break;
      case GP_TGOV1: tgov1_slots(d, sl); break;
      case GP_SEXS: sexs_slots(d, sl); break;
      case GP_ESDC1A: esdc1a_slots(d, sl); break;
      default: ieeest_slots(d, sl); break;
    }
  };
  for (int t = 0; t < 6; t++)
    for (size_t k = 0; k < C->dev[t].size() / ncols[t]; k++) {
      slots_of(t, C->dev[t].data() + k * ncols[t]);
      for (auto &e : sl) P.add(e.first, e.second);
    }
  const auto &SG = C->dev[GP_STATICGEN];
  const int nsg = (int)SG.size() / GP_COLS_STATICGEN;
  for (int k = 0; k < nsg; k++) {
    sg_slots(SG.data() + k * GP_COLS_STATICGEN, npt, sl);
    for (auto &e : sl) P.add(e.first, e.second);
  }
  P.finish(C->jcp, C->jri);
  static const int nslot[6] = {NS_GENROU, NS_IEESGO, NS_TGOV1, NS_SEXS, NS_ESDC1A, NS_IEEEST};
  for (int t = 0; t < 6; t++) {
    int nd = (int)C->dev[t].size() / ncols[t];
    C->jp[t].assign((size_t)std::max(1, nd) * nslot[t], -1);
    for (int k = 0; k < nd; k++) {
      slots_of(t, C->dev[t].data() + k * ncols[t]);
      if ((int)sl.size() != nslot[t]) throw GpError{GP_ERR_BADARG, "internal slot count mismatch"};
      for (int s = 0; s < nslot[t]; s++) {
        int pos = find_pos(C->jcp, C->jri, sl[s].first, sl[s].second);
        if (pos == -2) throw GpError{GP_ERR_BADARG, "internal: Jacobian slot missing"};
        C->jp[t][(size_t)k * nslot[t] + s] = pos;
      }
    }
  }
  // ---- per-bus attachments
  C->bus_sg.assign(nbus, -1);
  for (int k = 0; k < nsg; k++) {
    int b = SG[k * GP_COLS_STATICGEN];
    if (b < 0 || b >= nbus) throw GpError{GP_ERR_BADARG, "static generator bus out of range"};
    if (C->bus_sg[b] >= 0) throw GpError{GP_ERR_UNSUPPORTED, "more than one static generator on a bus"};
    C->bus_sg[b] = k;
  }
  const auto &ZP = C->dev[GP_ZIPLOAD];
  const int nzip = (int)ZP.size() / GP_COLS_ZIPLOAD;
  std::vector<std::vector<int>> bz(nbus), bc(nbus);
  for (int k = 0; k < nzip; k++) {
    int b = ZP[k * GP_COLS_ZIPLOAD];
    if (b < 0 || b >= nbus) throw GpError{GP_ERR_BADARG, "ZIP load bus out of range"};
    bz[b].push_back(k);
  }
  // machines at a bus in GENROU-table order (the reference's summation order)
  std::vector<std::pair<int, int>> gen_cl;
  for (int c = 0; c < C->ncl; c++) {
    const int *r = C->cl.data() + c * GP_CLUSTER_COLS;
    if (!r[GP_CL_TRIVIAL] && r[GP_CL_GENROU] >= 0) gen_cl.push_back({r[GP_CL_GENROU], c});
  }
  std::sort(gen_cl.begin(), gen_cl.end());
  for (auto &gc : gen_cl) bc[C->cl[gc.second * GP_CLUSTER_COLS + GP_CL_BUS]].push_back(gc.second);
  C->bus_cl_p.assign(nbus + 1, 0);
  C->bus_zip_p.assign(nbus + 1, 0);
  C->bus_cl.clear();
  C->bus_zip.clear();
  for (int b = 0; b < nbus; b++) {
    C->bus_cl.insert(C->bus_cl.end(), bc[b].begin(), bc[b].end());
    C->bus_zip.insert(C->bus_zip.end(), bz[b].begin(), bz[b].end());
    C->bus_cl_p[b + 1] = (int)C->bus_cl.size();
    C->bus_zip_p[b + 1] = (int)C->bus_zip.size();
  }
  // ---- bus rows of the monolithic Jacobian
  C->jp_bus.assign((size_t)15 * nbus, -1);
  C->ydiag.assign((size_t)4 * nbus, 0.0);
  C->jp_yoff_p.assign(nbus + 1, 0);
  C->jp_yoff_pos.clear();
  C->jp_yoff_val.clear();
  for (int b = 0; b < nbus; b++) {
    int vr = npt + 2 * b;
    int dpos[4] = {find_pos(C->jcp, C->jri, vr, vr), find_pos(C->jcp, C->jri, vr, vr + 1),
                   find_pos(C->jcp, C->jri, vr + 1, vr), find_pos(C->jcp, C->jri, vr + 1, vr + 1)};
    for (int i = 0; i < 4; i++) C->jp_bus[15 * b + i] = dpos[i];
    if (C->bus_sg[b] >= 0) {
      sg_slots(SG.data() + C->bus_sg[b] * GP_COLS_STATICGEN, npt, sl);
      for (int i = 0; i < 7; i++) C->jp_bus[15 * b + 4 + i] = find_pos(C->jcp, C->jri, sl[i].first, sl[i].second);
    }
    for (int rr = 0; rr < 2; rr++)
      for (int q = C->yrp[2 * b + rr]; q < C->yrp[2 * b + rr + 1]; q++) {
        int c = C->yci[q];
        if (c / 2 == b) C->ydiag[4 * b + 2 * rr + (c & 1)] += C->yv[q];
        else {
          C->jp_yoff_pos.push_back(find_pos(C->jcp, C->jri, vr + rr, npt + c));
          C->jp_yoff_val.push_back(C->yv[q]);
        }
      }
    C->jp_yoff_p[b + 1] = (int)C->jp_yoff_pos.size();
  }
  // ---- clusters: non-trivial list and X offsets
  C->cl_nt.clear();
  C->cl_xoff.assign(std::max(1, C->ncl), 0);
  C->xrows = 0;
  for (int c = 0; c < C->ncl; c++) {
    const int *r = C->cl.data() + c * GP_CLUSTER_COLS;
    if (r[GP_CL_TRIVIAL] || r[GP_CL_WSIZE] == 0) continue;
    if (r[GP_CL_GENROU] < 0) throw GpError{GP_ERR_UNSUPPORTED, "non-trivial cluster without a GENROU"};
    C->cl_nt.push_back(c);
    C->cl_xoff[c] = C->xrows;
    C->xrows += r[GP_CL_WSIZE];
  }
  // ---- reduced system: trivial-cluster states + bus voltages (src/schur.jl:53-168)
  C->red_of_z.assign(n, -1);
  C->z_of_red.clear();
  for (int c = 0; c < C->ncl; c++) {
    const int *r = C->cl.data() + c * GP_CLUSTER_COLS;
    if (r[GP_CL_TRIVIAL] || r[GP_CL_WSIZE] == 0)
      for (int j = 0; j < r[GP_CL_WSIZE]; j++) C->red_of_z[r[GP_CL_WSTART] + j] = 1;
  }
  for (int i = npt; i < n; i++) C->red_of_z[i] = 1;
  for (int i = 0; i < n; i++)
    if (C->red_of_z[i] > 0) {
      C->red_of_z[i] = (int)C->z_of_red.size();
      C->z_of_red.push_back(i);
    }
  C->nred = (int)C->z_of_red.size();
  // S pattern: J restricted to the reduced set + every bus diagonal block
  Pattern SP(C->nred);
  for (int c = 0; c < n; c++) {
    int rc = C->red_of_z[c];
    if (rc < 0) continue;
    for (int q = C->jcp[c]; q < C->jcp[c + 1]; q++) {
      int rr = C->red_of_z[C->jri[q]];
      if (rr >= 0) SP.add(rr, rc);
    }
  }
  SP.finish(C->scp, C->sri);
  const int nnzS = C->scp[C->nred];
  C->s_src.assign(nnzS, 0);
  std::vector<char> assigned(nnzS, 0);
  C->sv_bus.assign((size_t)15 * nbus, -1);
  C->n_svar = 0;
  auto spos = [&](int zr, int zc) {
    int rr = C->red_of_z[zr], rc = C->red_of_z[zc];
    return find_pos(C->scp, C->sri, rr, rc);
  };
  for (int b = 0; b < nbus; b++) {
    int vr = npt + 2 * b;
    int pr[4] = {spos(vr, vr), spos(vr, vr + 1), spos(vr + 1, vr), spos(vr + 1, vr + 1)};
    for (int i = 0; i < 4; i++) {
      C->sv_bus[15 * b + i] = C->n_svar;
      C->s_src[pr[i]] = C->n_svar++;
      assigned[pr[i]] = 1;
    }
    if (C->bus_sg[b] >= 0) {
      sg_slots(SG.data() + C->bus_sg[b] * GP_COLS_STATICGEN, npt, sl);
      for (int i = 0; i < 7; i++) {
        if (sl[i].first < 0) continue;
        int ps = spos(sl[i].first, sl[i].second);
        C->sv_bus[15 * b + 4 + i] = C->n_svar;
        C->s_src[ps] = C->n_svar++;
        assigned[ps] = 1;
      }
    }
  }
  C->s_cst.clear();
  for (int b = 0; b < nbus; b++) {
    int vr = npt + 2 * b;
    for (int rr = 0; rr < 2; rr++)
      for (int q = C->yrp[2 * b + rr]; q < C->yrp[2 * b + rr + 1]; q++) {
        int c = C->yci[q];
        if (c / 2 == b) continue;
        int ps = spos(vr + rr, npt + c);
        C->s_cst.push_back(-C->yv[q]);
        C->s_src[ps] = -(int)C->s_cst.size();
        assigned[ps] = 1;
      }
  }
  for (int e = 0; e < nnzS; e++)
    if (!assigned[e]) {
      // entries of the reduced pattern that no device writes are structural zeros
      C->s_cst.push_back(0.0);
      C->s_src[e] = -(int)C->s_cst.size();
    }
  if (C->s_cst.empty()) C->s_cst.push_back(0.0);
}

template <typename T>
void up(DBuf<T> &d, const std::vector<T> &v) { d.upload(v); }

void upload_trip_tables(gp_ctx *C) {
  up(C->d_tr_bus, C->tr_bus.empty() ? std::vector<int>{0, 0} : C->tr_bus);
  up(C->d_tr_y, C->tr_y.empty() ? std::vector<double>(8, 0.0) : C->tr_y);
  up(C->d_tv_p, C->tv_p);
  up(C->d_tv_i, C->tv_i.empty() ? std::vector<int>(4, 0) : C->tv_i);
  up(C->d_tv_base, C->tv_base.empty() ? std::vector<double>{0.0} : C->tv_base);
  DevTables &T = C->T;
  T.tr_bus = C->d_tr_bus.p; T.tr_y = C->d_tr_y.p;
  T.tv_p = C->d_tv_p.p; T.tv_i = C->d_tv_i.p; T.tv_base = C->d_tv_base.p;
}

void upload_tables(gp_ctx *C) {
  up(C->d_is_diff, C->is_diff);
  for (int t = 0; t < GP_NDEVTYPES; t++) {
    std::vector<int> v = C->dev[t];
    if (v.empty()) v.push_back(0);
    up(C->d_dev[t], v);
  }
  up(C->d_cl, C->cl.empty() ? std::vector<int>(GP_CLUSTER_COLS, 0) : C->cl);
  up(C->d_yrp, C->yrp); up(C->d_yci, C->yci.empty() ? std::vector<int>{0} : C->yci);
  up(C->d_yv, C->yv.empty() ? std::vector<double>{0.0} : C->yv);
  up(C->d_u0, C->u0.empty() ? std::vector<double>{0.0, 0.0} : C->u0);
  up(C->d_bus_cl_p, C->bus_cl_p); up(C->d_bus_cl, C->bus_cl.empty() ? std::vector<int>{0} : C->bus_cl);
  up(C->d_bus_zip_p, C->bus_zip_p); up(C->d_bus_zip, C->bus_zip.empty() ? std::vector<int>{0} : C->bus_zip);
  up(C->d_bus_sg, C->bus_sg);
  for (int t = 0; t < 6; t++) up(C->d_jp[t], C->jp[t]);
  up(C->d_jp_bus, C->jp_bus); up(C->d_jp_yoff_p, C->jp_yoff_p);
  up(C->d_jp_yoff_pos, C->jp_yoff_pos.empty() ? std::vector<int>{0} : C->jp_yoff_pos);
  up(C->d_jp_yoff_val, C->jp_yoff_val.empty() ? std::vector<double>{0.0} : C->jp_yoff_val);
  up(C->d_ydiag, C->ydiag);
  up(C->d_cl_nt, C->cl_nt.empty() ? std::vector<int>{0} : C->cl_nt);
  up(C->d_cl_xoff, C->cl_xoff);
  up(C->d_red_of_z, C->red_of_z); up(C->d_z_of_red, C->z_of_red); up(C->d_sv_bus, C->sv_bus);
  up(C->d_s_cst, C->s_cst);
  DevTables &T = C->T;
  T.n = C->n; T.nbus = C->nbus; T.net_ptr = C->net_ptr; T.ldM = C->ldM; T.M = C->M;
  T.is_diff = C->d_is_diff.p;
  T.genrou = C->d_dev[GP_GENROU].p; T.ieesgo = C->d_dev[GP_IEESGO].p; T.tgov1 = C->d_dev[GP_TGOV1].p;
  T.sexs = C->d_dev[GP_SEXS].p; T.esdc1a = C->d_dev[GP_ESDC1A].p; T.ieeest = C->d_dev[GP_IEEEST].p;
  T.zip = C->d_dev[GP_ZIPLOAD].p; T.sg = C->d_dev[GP_STATICGEN].p;
  T.u0 = C->d_u0.p; T.cl = C->d_cl.p; T.n_cl = C->ncl;
  T.yrp = C->d_yrp.p; T.yci = C->d_yci.p; T.yv = C->d_yv.p;
  T.bus_cl_p = C->d_bus_cl_p.p; T.bus_cl = C->d_bus_cl.p; T.bus_zip_p = C->d_bus_zip_p.p; T.bus_zip = C->d_bus_zip.p;
  T.bus_sg = C->d_bus_sg.p;
  T.jp_genrou = C->d_jp[0].p; T.jp_ieesgo = C->d_jp[1].p; T.jp_tgov1 = C->d_jp[2].p; T.jp_sexs = C->d_jp[3].p;
  T.jp_esdc1a = C->d_jp[4].p; T.jp_ieeest = C->d_jp[5].p; T.jp_bus = C->d_jp_bus.p;
  T.jp_yoff_p = C->d_jp_yoff_p.p; T.jp_yoff_pos = C->d_jp_yoff_pos.p; T.jp_yoff_val = C->d_jp_yoff_val.p;
  T.ydiag = C->d_ydiag.p;
  T.cl_nt = C->d_cl_nt.p; T.n_cl_nt = (int)C->cl_nt.size(); T.cl_xoff = C->d_cl_xoff.p;
  T.red_of_z = C->d_red_of_z.p; T.sv_bus = C->d_sv_bus.p;
  if (C->tv_p.empty()) C->tv_p.assign(C->nbus + 1, 0);
  upload_trip_tables(C);
}

void alloc_state(gp_ctx *C) {
  const size_t L = (size_t)C->ldM;
  C->d_p.alloc(std::max(1, C->npar));
  C->d_z.alloc((size_t)C->n * L);
  C->d_z0.alloc((size_t)C->n * L);
  C->d_zold.alloc((size_t)C->n * L);
  C->d_f.alloc((size_t)C->n * L);
  C->d_inj.alloc((size_t)std::max(1, 2 * C->ncl) * L);
  C->d_Dg.alloc((size_t)std::max(1, 6 * C->ncl) * L);
  C->d_fnorm.alloc(L);
  C->d_act.alloc(L); C->d_refac.alloc(L); C->d_sel.alloc(L);
  C->d_split.alloc(6 * L);
  {
    const RoundLayout R(L);
    C->d_round.alloc(R.bytes());
    HIPCHK(hipMemset(C->d_round.p, 0, R.bytes()));
    C->h_round.alloc(R.bytes());
    uint8_t *b = C->d_round.p, *q = C->d_round.p;
    C->d_first.view(b + R.first(), L);
    C->d_refac0.view(b + R.refac0(), L);
    C->d_exact.view(b + R.exact(), L);
    C->d_dtv.view(b + R.dt(), L);
    C->d_work.view(b + R.work(), L);
    C->d_nunconv.view(b + R.cnt(), 4);
    C->d_failed.view(b + R.failed(), L);
    C->d_areq.view(b + R.areq(), L);
    C->d_acc.view(q + R.acc(), L);
    C->d_adv.view(q + R.adv(), L);
    C->d_col.view(q + R.col(), L);
    C->d_sv.view(q + R.sv(), L);
    C->d_svslot.view(q + R.svslot(), L);
    C->d_ck.view(q + R.ck(), L);
    C->d_ckslot.view(q + R.ckslot(), L);
    C->d_pred.view(q + R.pred(), L);
  }
  C->d_fbus.alloc((size_t)MAXFAULT * L);
  C->d_fg.alloc((size_t)MAXFAULT * L);
  C->d_sch.alloc(L);
  C->d_nstall.alloc(L);
  C->d_iters.alloc(L);
  C->d_upd.alloc(L);
  C->d_cur.alloc(L);
  C->d_pend.alloc(L);
  C->d_batch.alloc((size_t)gp_ctx::kLanes * 2 * L);
  C->h_batch.alloc((size_t)gp_ctx::kBatchSlots * 2 * L);
  for (int l = 0; l < gp_ctx::kLanes; l++) {
    C->lanes[l].req = C->d_batch.p + (size_t)l * 2 * L;
    C->lanes[l].tgt = C->lanes[l].req + L;
    C->lanes[l].batch.assign(L, 0);
  }
  HIPCHK(hipMemset(C->d_cur.p, 0, L));
  HIPCHK(hipMemset(C->d_pend.p, 0, L));
  HIPCHK(hipMemset(C->d_upd.p, 0, L));
  C->cur.assign(L, 0);
  C->pend.assign(L, 0);
  C->queue.assign(L, 0);
  C->d_fbest.alloc(L);
  C->d_status.alloc(L);
  C->d_nb.alloc(L);
  HIPCHK(hipMemset(C->d_nb.p, 0, L * sizeof(unsigned long long)));
  {
    // the residual kernels write every row of F: the non-trivial clusters'
    // windows, the PV and slack static generators' algebraic rows and two
    // rows per bus
    int64_t rows = 2 * (int64_t)C->nbus;
    for (int c : C->cl_nt) rows += C->cl[(size_t)c * GP_CLUSTER_COLS + GP_CL_WSIZE];
    const auto &SG = C->dev[GP_STATICGEN];
    for (size_t k = 0; k < SG.size() / GP_COLS_STATICGEN; k++) {
      const int bt = SG[k * GP_COLS_STATICGEN + 3];
      rows += bt == 2 ? 1 : bt == 3 ? 2 : 0;
    }
    C->pm_rows = rows == C->n;
    C->d_pm.alloc((size_t)(C->cl_nt.size() + C->nbus) * L);
    HIPCHK(hipMemset(C->d_pm.p, 0, (size_t)(C->cl_nt.size() + C->nbus) * L * sizeof(double)));
  }
  HIPCHK(hipMemset(C->d_status.p, 0, L * sizeof(int)));
  HIPCHK(hipMemset(C->d_z.p, 0, (size_t)C->n * L * sizeof(double)));
  std::vector<int> fb((size_t)MAXFAULT * L, -1);
  HIPCHK(hipMemcpy(C->d_fbus.p, fb.data(), fb.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(C->d_fg.p, 0, (size_t)MAXFAULT * L * sizeof(double)));
}

FaultState faults(gp_ctx *C) { return FaultState{C->d_fbus.p, C->d_fg.p, C->nexc}; }

void set_mask(gp_ctx *C, DBuf<uint8_t> &d, const std::vector<uint8_t> &h) {
  HIPCHK(hipMemcpyAsync(d.p, h.data(), h.size(), hipMemcpyHostToDevice, C->stream));
}

// The pivot analysis's extra matrices (see splu_analyse): the
// representative values `a` (J in CSC order, or S when reduced) with a set
// of tripped lines taken out at once -- the diagonal blocks of its buses gain
// +yff / +ytt and the off-diagonal blocks +yft / +ytf, realified [g -b; b g].
// One matrix per distinct set: every line a scenario trips (at once), and
// every tripped line alone.  Empty when no line is tripped.  only >= 0: the
// lines of that scenario alone.
std::vector<std::vector<double>> trip_matrices(const gp_ctx *C, const std::vector<double> &a, bool reduced,
                                               int only = -1) {
  std::vector<std::vector<double>> out;
  std::vector<gp_line_trip> trips;
  for (const gp_line_trip &t : C->trips)
    if (only < 0 || t.scenario == only) trips.push_back(t);
  if (trips.empty()) return out;
  const int npt = C->net_ptr;
  auto pos = [&](int zr, int zc) {
    return reduced ? find_pos(C->scp, C->sri, C->red_of_z[zr], C->red_of_z[zc]) : find_pos(C->jcp, C->jri, zr, zc);
  };
  auto add = [&](std::vector<double> &m, int rb, int cb, const double *y) {
    const double v[4] = {y[0], -y[1], y[1], y[0]};
    for (int rr = 0; rr < 2; rr++)
      for (int cc = 0; cc < 2; cc++) {
        const int p = pos(npt + 2 * rb + rr, npt + 2 * cb + cc);
        if (p < 0) throw GpError{GP_ERR_HIP, "internal: tripped line's entry missing from a pattern"};
        m[p] += v[2 * rr + cc];
      }
  };
  // distinct lines (a line tripped in several scenarios, or tripped again
  // after a reclosure, is one line)
  std::vector<std::vector<double>> lines;
  std::vector<int> line_of(trips.size());
  for (size_t t = 0; t < trips.size(); t++) {
    const gp_line_trip &tr = trips[t];
    std::vector<double> key{(double)tr.from_bus, (double)tr.to_bus};
    key.insert(key.end(), tr.y, tr.y + 8);
    auto it = std::find(lines.begin(), lines.end(), key);
    line_of[t] = (int)(it - lines.begin());
    if (it == lines.end()) lines.push_back(key);
  }
  std::vector<std::vector<int>> sets;
  for (int l = 0; l < (int)lines.size(); l++) sets.push_back({l});
  std::map<int, std::vector<int>> per_scen;
  for (size_t t = 0; t < trips.size(); t++) per_scen[trips[t].scenario].push_back(line_of[t]);
  for (auto &kv : per_scen) {
    std::vector<int> st = kv.second;
    std::sort(st.begin(), st.end());
    st.erase(std::unique(st.begin(), st.end()), st.end());
    if (std::find(sets.begin(), sets.end(), st) == sets.end()) sets.push_back(st);
  }
  for (const auto &st : sets) {
    std::vector<double> m = a;
    for (int l : st) {
      const std::vector<double> &k = lines[l];
      const int fb = (int)k[0], tb = (int)k[1];
      const double *y = k.data() + 2;
      add(m, fb, fb, y);
      add(m, fb, tb, y + 2);
      add(m, tb, fb, y + 4);
      add(m, tb, tb, y + 6);
    }
    out.push_back(std::move(m));
  }
  return out;
}

static std::vector<const double *> ptrs(const std::vector<std::vector<double>> &v) {
  std::vector<const double *> p;
  for (const auto &x : v) p.push_back(x.data());
  return p;
}

// Representative matrix of scenario 0 for the pivot analysis, evaluated by
// the product's own kernels at the current state with dt = 0 (the matrix of
// the event re-solves, whose differential rows are the identity).  On that
// matrix no entry that scales with dt (a differential row's -dt df/dz) can
// be chosen as a pivot, so one set of static pivots serves every dt.  A
// pivot chosen on a dt > 0 matrix can be such an entry -- in the column of a
// bus voltage whose own diagonal is zero, as on a generator bus fed only by
// lossless transformers (G = 0, and the machine's subtransient admittance is
// imaginary) -- and vanishes at dt = 0, which ruins the re-solve's factor.
// The plan does not depend on dt and is made once.
void analyse(gp_ctx *C, int method) {
  const double dt = 0.0;
  const int L = C->ldM;
  std::vector<uint8_t> sel(L, 0);
  sel[0] = 1;
  set_mask(C, C->d_sel, sel);
  if (method == GP_METHOD_MONO_FP64) {
    if (C->luJ.ready && C->luJ.trip_gen == C->trip_gen) return;
    const int nnz = C->jcp[C->n];
    C->d_jv.alloc((size_t)nnz * L);
    HIPCHK(hipMemsetAsync(C->d_jv.p, 0, (size_t)nnz * L * sizeof(double), C->stream));
    launch_jac_mono(C->T, C->d_p.p, C->d_z.p, dt, faults(C), C->d_jv.p, C->d_sel.p, C->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> jv(nnz);
    HIPCHK(hipMemcpy2DAsync(jv.data(), sizeof(double), C->d_jv.p, (size_t)L * sizeof(double), sizeof(double), nnz,
                            hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    const std::vector<std::vector<double>> jx =
        C->opt.pivot_rule == 1 ? std::vector<std::vector<double>>{} : trip_matrices(C, jv, false);
    std::string e = splu_analyse(C->n, C->jcp.data(), C->jri.data(), jv.data(), C->luJ.A, 1e-3, ptrs(jx));
    C->rep0_J = jv;
    if (!e.empty()) throw GpError{GP_ERR_SINGULAR_PIVOT, "Jacobian analysis: " + e};
    std::vector<int> asrc(nnz);
    for (int i = 0; i < nnz; i++) asrc[i] = i;
    C->luJ.tune(asrc, jv, nullptr, sizeof(double), 1.0, C->stream, C->opt.dense_block);   // a solve per factorisation
    if (C->luJ.dev.gx) {
      C->d_gxJ.ensure((size_t)C->n * L);
      C->luJ.dev.gx_ws = C->d_gxJ.p;
    }
    C->luJ.ready = true;
    C->luJ.trip_gen = C->trip_gen;
    C->d_luJ.alloc(C->luJ.lu_elems() * C->M);
    C->d_svJ.alloc((size_t)C->luJ.dev.ns * C->M);
    HIPCHK(hipMemset(C->d_svJ.p, 0, C->d_svJ.n * sizeof(double)));   // padding slots stay zero
    C->d_dnJ.alloc((size_t)std::max(1, C->luJ.A.nd * C->luJ.A.nd) * C->M);
  } else {
    if (C->luS.ready && C->luS.trip_gen == C->trip_gen) return;
    const int nnzS = C->scp[C->nred];
    C->d_svar.alloc((size_t)std::max(1, C->n_svar) * L);
    for (auto &ln : C->lanes) ln.snap.alloc((size_t)std::max(1, C->n_svar) * L);
    C->d_rhs.alloc((size_t)C->nred * L);
    C->d_dv.alloc((size_t)C->nred * L);
    C->d_X64.alloc((size_t)std::max(1, C->xrows) * 3 * L);
    HIPCHK(hipMemcpyAsync(C->d_zold.p, C->d_z.p, (size_t)C->n * L * sizeof(double), hipMemcpyDeviceToDevice, C->stream));
    launch_cluster_residual(C->T, C->d_p.p, C->d_z.p, C->d_zold.p, dt, faults(C),
                            C->d_f.p, C->d_inj.p, C->d_sel.p, C->stream);
    launch_bus_residual(C->T, C->d_p.p, C->d_z.p, C->d_inj.p, faults(C), C->d_f.p, C->d_sel.p, C->stream);
    launch_cluster_schur<double>(C->T, C->d_p.p, C->d_z.p, C->d_f.p, dt, faults(C),
                                 C->d_X64.p, C->d_Dg.p, C->d_sel.p, C->stream);
    launch_bus_schur(C->T, C->d_p.p, C->d_z.p, C->d_f.p, C->d_Dg.p, faults(C), C->d_svar.p, C->d_rhs.p, C->d_sel.p,
                     C->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> sv(std::max(1, C->n_svar));
    HIPCHK(hipMemcpy2DAsync(sv.data(), sizeof(double), C->d_svar.p, (size_t)L * sizeof(double), sizeof(double),
                            std::max(1, C->n_svar), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    std::vector<double> sx(nnzS);
    for (int e = 0; e < nnzS; e++) {
      int s = C->s_src[e];
      sx[e] = s >= 0 ? sv[s] : C->s_cst[-s - 1];
    }
    const std::vector<std::vector<double>> sxx =
        C->opt.pivot_rule == 1 ? std::vector<std::vector<double>>{} : trip_matrices(C, sx, true);
    std::string e = splu_analyse(C->nred, C->scp.data(), C->sri.data(), sx.data(), C->luS.A, 1e-3, ptrs(sxx));
    C->rep0_S = sx;
    if (!e.empty()) throw GpError{GP_ERR_SINGULAR_PIVOT, "reduced-network analysis: " + e};
    // a factorisation serves several solves (refactorisation policies), and
    // the solves are on every round's critical path
    C->luS.tune(C->s_src, sv, C->d_s_cst.p, method == GP_METHOD_SCHUR_FP32 ? sizeof(float) : sizeof(double), 2.0,
                C->stream, C->opt.dense_block);
    if (C->luS.dev.gx) {
      C->d_gxS.ensure((size_t)C->nred * L);
      C->luS.dev.gx_ws = C->d_gxS.p;
    }
    C->luS.ready = true;
    C->luS.trip_gen = C->trip_gen;
  }
}

// Factor buffers of the Schur methods.  The FP64 ones exist for both: the
// FP32 method runs a solve that failed again with FP64 factors.
void ensure_schur_buffers(gp_ctx *C, bool fp32) {
  const size_t L = C->ldM;
  C->d_luS64.ensure(C->luS.lu_elems() * C->M);
  if (C->d_svS64.n != (size_t)2 * C->luS.dev.ns * C->M) {   // two buffers per scenario
    C->d_svS64.alloc((size_t)2 * C->luS.dev.ns * C->M);
    HIPCHK(hipMemset(C->d_svS64.p, 0, C->d_svS64.n * sizeof(double)));   // padding slots stay zero
  }
  C->d_dnS64.ensure((size_t)2 * std::max(1, C->luS.A.nd * C->luS.A.nd) * C->M);
  try {
    C->luS.shape(sizeof(double));
    C->luS64 = C->luS.dev;
    C->has64 = true;
  } catch (const GpError &) {   // the FP64 solve's LDS does not fit: FP32 only, no FP64 retries
    C->has64 = false;
    if (!fp32) throw;
  }
  if (fp32) {
    C->d_X32.ensure((size_t)std::max(1, C->xrows) * 3 * L);
    C->d_luS32.ensure(C->luS.lu_elems() * C->M);
    if (C->d_svS32.n != (size_t)2 * C->luS.dev.ns * C->M) {   // two buffers per scenario
      C->d_svS32.alloc((size_t)2 * C->luS.dev.ns * C->M);
      HIPCHK(hipMemset(C->d_svS32.p, 0, C->d_svS32.n * sizeof(float)));   // padding slots stay zero
    }
    C->d_dnS32.ensure((size_t)2 * std::max(1, C->luS.A.nd * C->luS.A.nd) * C->M);
    C->luS.shape(sizeof(float));
  }
}

// Shared-factor method: scenario m's Woodbury table row from its event slots
// (fbus/fg [MAXFAULT][ldM]: bus and 1/r_fault, bus < 0 free).  The faults
// of a bus add -g_b to the bus's two diagonal entries of S, so that
// S = P + U D U^T with D = diag(-g_b) and, by Woodbury,
//   S^{-1} b = y - W H^{-1} U^T y,  y = P^{-1} b,  H = D^{-1} + U^T W
// (_precompute_woodbury_multi_gpu!, ext:2441-2495: H = -r I + U^T W).
void shared_set(gp_ctx *C, int m, const int *fbus, const double *fg) {
  gp_ctx::Shared &W = C->sh;
  const int L = C->ldM;
  std::vector<std::pair<int, double>> gb;   // faulted bus, summed admittance
  for (int q = 0; q < MAXFAULT; q++) {
    const int b = fbus[(size_t)q * L + m];
    if (b < 0) continue;
    auto it = std::find_if(gb.begin(), gb.end(), [b](const std::pair<int, double> &x) { return x.first == b; });
    if (it == gb.end())
      gb.emplace_back(b, fg[(size_t)q * L + m]);
    else
      it->second += fg[(size_t)q * L + m];
  }
  const int nc = 2 * (int)gb.size();
  if (nc > W.kw) throw GpError{GP_ERR_HIP, "internal: Woodbury table too narrow"};
  int cols[2 * MAXFAULT];
  double dinv[2 * MAXFAULT], H[2 * MAXFAULT][4 * MAXFAULT];
  for (int i = 0; i < (int)gb.size(); i++) {
    auto it = W.col_of_bus.find(gb[i].first);
    if (it == W.col_of_bus.end()) throw GpError{GP_ERR_HIP, "internal: faulted bus without Woodbury columns"};
    cols[2 * i] = it->second;
    cols[2 * i + 1] = it->second + 1;
    dinv[2 * i] = dinv[2 * i + 1] = -1.0 / gb[i].second;
  }
  // H | I, inverted by Gauss-Jordan elimination with partial pivoting
  for (int i = 0; i < nc; i++)
    for (int j = 0; j < 2 * nc; j++)
      H[i][j] = j < nc ? W.G[(size_t)cols[i] * W.ncol + cols[j]] + (i == j ? dinv[i] : 0.0)
                       : (j - nc == i ? 1.0 : 0.0);
  for (int p = 0; p < nc; p++) {
    int r = p;
    for (int i = p + 1; i < nc; i++)
      if (std::fabs(H[i][p]) > std::fabs(H[r][p])) r = i;
    if (!(std::fabs(H[r][p]) > 0.0) || !std::isfinite(H[r][p]))
      throw GpError{GP_ERR_SINGULAR_PIVOT, "shared factorisation: singular Woodbury capacitance matrix (scenario " +
                                               std::to_string(m) + ")"};
    if (r != p)
      for (int j = 0; j < 2 * nc; j++) std::swap(H[p][j], H[r][j]);
    const double inv = 1.0 / H[p][p];
    for (int j = 0; j < 2 * nc; j++) H[p][j] *= inv;
    for (int i = 0; i < nc; i++) {
      if (i == p || H[i][p] == 0.0) continue;
      const double f = H[i][p];
      for (int j = 0; j < 2 * nc; j++) H[i][j] -= f * H[p][j];
    }
  }
  const int kw = W.kw;
  W.hnc[m] = nc;
  for (int c = 0; c < kw; c++) {
    W.hcol[(size_t)m * kw + c] = c < nc ? cols[c] : 0;
    W.hpos[(size_t)m * kw + c] = c < nc ? W.upos[cols[c]] : 0;
    for (int j = 0; j < kw; j++) W.hh[((size_t)m * kw + c) * kw + j] = (c < nc && j < nc) ? H[c][nc + j] : 0.0;
  }
}
// the tables on the device, in stream order (the host rows stay valid until
// the stream has passed the copies: callers change them only after a later
// synchronisation)
void shared_upload(gp_ctx *C) {
  gp_ctx::Shared &W = C->sh;
  HIPCHK(hipMemcpyAsync(W.nc.p, W.hnc.data(), W.hnc.size() * sizeof(int), hipMemcpyHostToDevice, C->stream));
  HIPCHK(hipMemcpyAsync(W.col.p, W.hcol.data(), W.hcol.size() * sizeof(int), hipMemcpyHostToDevice, C->stream));
  HIPCHK(hipMemcpyAsync(W.pos.p, W.hpos.data(), W.hpos.size() * sizeof(int), hipMemcpyHostToDevice, C->stream));
  HIPCHK(hipMemcpyAsync(W.h.p, W.hh.data(), W.hh.size() * sizeof(double), hipMemcpyHostToDevice, C->stream));
}

// Events around one launch of a timed kernel (gp_options.profile).  A
// launch whose scenario count is not known when it goes out gets it later
// (prof_count); launches that worked on no scenario are left out.
int prof_begin(gp_ctx *C, gp_ctx::KProf &P, hipStream_t s) {
  if (!C->opt.profile) return -1;
  if (P.calls++ % P.every) return -1;
  if (P.used + 2 > P.ev.size()) {
    for (int k = 0; k < 64; k++) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      P.ev.push_back(e);
    }
    P.scen.resize(P.ev.size() / 2, 0);
  }
  const int i = (int)P.used;
  P.used += 2;
  P.scen[i / 2] = 0;
  HIPCHK(hipEventRecord(P.ev[i], s));
  return i;
}
void prof_end(gp_ctx::KProf &P, int i, hipStream_t s, int64_t scenarios) {
  if (i < 0) return;
  HIPCHK(hipEventRecord(P.ev[i + 1], s));
  P.scen[i / 2] = scenarios;
}
// The end event of pair i (NULL when not timed), for a factorisation
// launch to record after its factor kernel and before the kernel of the
// partitioned solves' composite entries: the pair times the factor kernel.
hipEvent_t prof_end_ev(gp_ctx::KProf &P, int i, int64_t scenarios) {
  if (i < 0) return nullptr;
  P.scen[i / 2] = scenarios;
  return P.ev[i + 1];
}
void prof_count(gp_ctx::KProf &P, int i, int64_t scenarios) {
  if (i >= 0) P.scen[i / 2] = scenarios;
}
// Total milliseconds of the recorded pairs (all complete), then reset.
double prof_collect(gp_ctx::KProf &P, int64_t &launches, int64_t &scen) {
  double tot = 0.0;
  launches = scen = 0;
  for (size_t i = 0; i + 1 < P.used; i += 2) {
    if (P.scen[i / 2] <= 0) continue;
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, P.ev[i], P.ev[i + 1]));
    tot += ms;
    launches++;
    scen += P.scen[i / 2];
  }
  P.used = 0;
  return tot;
}

// Waits for an event recorded on a stream by polling it:
// a round's exchange is short, and polling wakes the host sooner than a
// blocking synchronisation.
void spin_wait(hipEvent_t ev) {
  for (;;) {
    const hipError_t q = hipEventQuery(ev);
    if (q == hipSuccess) return;
    if (q != hipErrorNotReady) HIPCHK(q);
  }
}

struct NewtonOut {
  int64_t corrections = 0;
  int64_t refac = 0;
  int64_t rounds = 0, factor_launches = 0;
  int64_t exact_retries = 0;     // solves run again as exact Newton
  std::vector<uint8_t> failed;   // scenarios that did not converge (empty: none)
};

// Asynchronous refactorisation (GP_REFACTOR_ASYNC).  A batch copies the
// requested scenarios' reduced-network values (made by this round's Schur
// kernels) aside, then factorises them on the second stream into the buffer
// their solves do not read; when the batch has finished, those scenarios
// switch buffers.  The copy is taken on the main stream and the factorisation
// waits for it, so it also waits for every solve that was still reading the
// buffer it overwrites.
bool any_inflight(const gp_ctx *C) {
  for (const auto &ln : C->lanes)
    if (ln.inflight) return true;
  return false;
}
int free_lane(const gp_ctx *C) {
  for (int l = 0; l < gp_ctx::kLanes; l++)
    if (!C->lanes[l].inflight) return l;
  return -1;
}

// A batch is staged on the main stream (its masks, the pending flags, the
// snapshot of its matrix values) and then started on its lane: the caller
// launches the round's solve between the two, so that the host calls that
// start the factorisation overlap the solve instead of delaying it.  Lane l
// takes the queued requests of its class: with FP32 factors, those of the
// scenarios running exact Newton (exacth[m]) go to lane 1 in FP64, the
// others to lane 0.  Returns l, or -1 (lane busy or nothing queued for it).
int lane_class(const gp_ctx *C, bool fp32, const uint8_t *exacth, int m) {
  return (fp32 && C->has64 && exacth && exacth[m]) ? 1 : 0;
}
int stage_batch(gp_ctx *C, int l, bool fp32, const uint8_t *exacth) {
  const int L = C->ldM;
  gp_ctx::Lane &ln = C->lanes[l];
  if (ln.inflight) return -1;
  int nb = 0;
  for (int m = 0; m < L; m++) nb += (C->queue[m] && lane_class(C, fp32, exacth, m) == l) ? 1 : 0;
  if (nb == 0) return -1;
  for (int m = 0; m < L; m++) {
    ln.batch[m] = (C->queue[m] && lane_class(C, fp32, exacth, m) == l) ? 1 : 0;
    if (ln.batch[m]) {
      C->pend[m] = 1;
      C->queue[m] = 0;
    }
  }
  if (L <= kStageBits) {   // the mask travels as kernel arguments
    ScenarioBits bits{};
    for (int m = 0; m < L; m++)
      if (ln.batch[m]) bits.w[m >> 5] |= 1u << (m & 31);
    launch_stage_batch(L, bits, ln.req, ln.tgt, C->d_cur.p, C->d_pend.p, C->stream);
  } else {
    uint8_t *slot = C->h_batch.p + (size_t)C->batch_slot * 2 * L;
    C->batch_slot = (C->batch_slot + 1) % gp_ctx::kBatchSlots;
    for (int m = 0; m < L; m++) {
      slot[m] = ln.batch[m];
      slot[L + m] = (uint8_t)(1 - C->cur[m]);
    }
    HIPCHK(hipMemcpyAsync(ln.req, slot, 2 * L, hipMemcpyHostToDevice, C->stream));
    launch_async_flags(L, 1, ln.req, C->d_cur.p, C->d_pend.p, C->stream);
  }
  launch_copy_transpose(std::max(1, C->n_svar), L, C->d_svar.p, ln.snap.p, ln.req, C->stream);   // [m][n_svar]
  HIPCHK(hipEventRecord(ln.ev_snap, C->stream));
  ln.nb = nb;
  return l;
}

template <typename TF>
void start_batch(gp_ctx *C, int l, TF *luv, TF *sv, TF *dn) {
  gp_ctx::Lane &ln = C->lanes[l];
  const int L = C->ldM, nb = ln.nb;
  HIPCHK(hipStreamWaitEvent(ln.stream, ln.ev_snap, 0));
  // timed: lane 0, the method's own precision (lane 1 holds the FP32
  // method's FP64 exact-Newton factorisations)
  const int pi = l == 0 ? prof_begin(C, C->prof_factor, ln.stream) : -1;
  launch_splu_factor_impl<TF>(sizeof(TF) == 8 ? C->luS64 : C->luS.dev, ln.snap.p, C->d_s_cst.p,
                              -std::max(1, C->n_svar), luv, sv, dn,
                              ln.req, ln.tgt, C->M, C->d_status.p, ln.stream, prof_end_ev(C->prof_factor, pi, nb));
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ln.ev_done, ln.stream));
  ln.inflight = true;
}

// The batches in flight that have finished (all of them with `wait`): their
// scenarios switch buffers.  With `wait` the context's stream waits for the
// batches (the host does not): what it runs next sees them landed.  Returns
// whether any landed.
bool land_lane(gp_ctx *C, int l, bool wait) {
  const int L = C->ldM;
  gp_ctx::Lane &ln = C->lanes[l];
  if (!ln.inflight) return false;
  if (wait) {
    HIPCHK(hipStreamWaitEvent(C->stream, ln.ev_done, 0));
  } else {
    const hipError_t q = hipEventQuery(ln.ev_done);
    if (q == hipErrorNotReady) return false;
    HIPCHK(q);
  }
  for (int m = 0; m < L; m++)
    if (ln.batch[m]) {
      C->cur[m] = (uint8_t)(1 - C->cur[m]);
      C->pend[m] = 0;
      ln.batch[m] = 0;
    }
  // the same update on the device, from the batch's request mask (no
  // longer read by the finished factorisation)
  launch_async_flags(L, 0, ln.req, C->d_cur.p, C->d_pend.p, C->stream);
  ln.inflight = false;
  return true;
}
bool land_batch(gp_ctx *C, bool wait) {
  bool any = false;
  for (int l = 0; l < gp_ctx::kLanes; l++) any = land_lane(C, l, wait) || any;
  return any;
}

// One Newton round engine (newton_step!, src/numerics.jl:55-176, batched):
// eval() evaluates F, its max-norm and the per-scenario decisions of
// k_norm_finish (converged scenarios leave `work`); correct() applies one
// correction to the scenarios that asked for one and runs the
// factorisations the refactorisation policy asks for.  Used in lock step by
// newton() and by the scenario-asynchronous time loop.
struct Newton {
  gp_ctx *C;
  int method, itermax, stall;
  double dt, tol, theta, async_below;
  bool schur, reuse, async, fp32;
  bool shared;                   // one shared factorisation with Woodbury corrections: nothing factorises
  bool lagged;                   // a factorisation may serve several iterations
  bool all64 = false;            // FP32 method running exact Newton: every factor in FP64
  const double *dtv = nullptr;   // per-scenario step on the device (NULL: dt for all)
  const uint8_t *exactv = nullptr;   // per-scenario exact-Newton flags on the device (NULL: none)
  const uint8_t *exacth = nullptr;   // the same flags on the host
  const uint8_t *workh = nullptr;    // host copy of this round's working flags (NULL: unknown)
  std::vector<uint8_t> req;
  const uint8_t *reqp;           // host copy of this round's asynchronous requests
  NewtonOut out;

  Newton(gp_ctx *c, int meth, double dt_, double tol_, int itmax, bool has_refac0)
      : C(c), method(meth), itermax(itmax), dt(dt_), tol(tol_) {
    const int pol = C->opt.refactor_policy;
    schur = method != GP_METHOD_MONO_FP64;
    shared = method == GP_METHOD_SCHUR_SHARED;
    reuse = has_refac0 && schur && !shared && (pol == GP_REFACTOR_ADAPTIVE || pol == GP_REFACTOR_ASYNC);
    async = schur && !shared && pol == GP_REFACTOR_ASYNC;
    stall = C->opt.stall_window;
    // refactor policy: theta <= 0 refactors every iteration; otherwise after
    // the first iteration only when |F| fell by less than theta (FIRST_ITER:
    // a factor 1e-300, which every iteration reaches, so never)
    theta = (!schur || pol == GP_REFACTOR_EVERY_ITER)
                ? 0.0
                : ((pol == GP_REFACTOR_ADAPTIVE || pol == GP_REFACTOR_ASYNC)
                       ? (C->opt.refactor_theta > 0.0 ? C->opt.refactor_theta : 30.0)
                       : 1e-300);
    if (shared) theta = -1.0;   // one factorisation, made before the loop: no requests
    lagged = theta > 0.0;
    fp32 = method == GP_METHOD_SCHUR_FP32;
    switch_window = C->opt.retry_window < 0 ? 0 : (C->opt.retry_window > 0 ? C->opt.retry_window : 8);
    if (!retryable()) switch_window = 0;
    async_below = C->opt.async_below > 0.0 ? C->opt.async_below : 1e300;
    req.assign(C->ldM, 0);
    reqp = req.data();
  }
  // exact Newton for every scenario of this engine: a synchronous FP64
  // factorisation at every iteration
  void make_exact() {
    theta = 0.0;
    shared = false;   // (the shared method's exact Newton is the FP64 Schur method's)
    reuse = async = false;
    all64 = fp32 && C->has64;
    switch_window = 0;
  }
  // A failed solve runs again as exact Newton: when factorisations are
  // reused, or are FP32 (the reference factors in FP64 at every iteration).
  bool retryable() const { return shared || lagged || (fp32 && C->has64); }
  // corrections without halving |F| after which an inexact solve is given
  // up for exact Newton (0: only at itermax)
  int switch_window = 0;

  // first: every working scenario starts its solve (lock step); firstv: the
  // scenarios that do (device flags).  cnt[] as k_norm_finish's n_unconv,
  // read back here unless cnt is NULL (the caller uploaded zero counts and
  // reads them back with the rest of the round's exchange block).
  void eval(int first, const uint8_t *firstv, const uint8_t *refac0, int *cnt, bool read_req) {
    const int L = C->ldM;
    double *z = C->d_z.p;
    double *pm = C->pm_rows ? C->d_pm.p : nullptr;
    launch_cluster_residual(C->T, C->d_p.p, z, C->d_zold.p, dt, faults(C),
                            C->d_f.p, C->d_inj.p, C->d_work.p, C->stream, dtv, pm);
    launch_bus_residual(C->T, C->d_p.p, z, C->d_inj.p, faults(C), C->d_f.p, C->d_work.p, C->stream, pm);
    if (cnt) HIPCHK(hipMemsetAsync(C->d_nunconv.p, 0, 4 * sizeof(int), C->stream));
    // max-norm over the kernels' per-block maxima (the same maximum as over F)
    launch_rownorm_impl(pm ? C->T.n_cl_nt + C->nbus : C->n, C->M, L, pm ? pm : C->d_f.p, C->d_work.p, tol, C->d_fnorm.p, C->d_nb.p, C->d_nunconv.p,
                        C->d_refac.p, reuse ? refac0 : nullptr, first, theta, stall, itermax, C->d_fbest.p,
                        C->d_nstall.p, C->d_iters.p, C->d_failed.p, async ? C->d_pend.p : nullptr, C->d_upd.p,
                        async ? async_below : 0.0, C->d_areq.p, C->d_sch.p, C->stream, firstv, exactv,
                        switch_window);
    if (!cnt) return;
    HIPCHK(hipMemcpyAsync(cnt, C->d_nunconv.p, 4 * sizeof(int), hipMemcpyDeviceToHost, C->stream));
    if (async && read_req) HIPCHK(hipMemcpyAsync(req.data(), C->d_areq.p, L, hipMemcpyDeviceToHost, C->stream));
  }

  bool queued() const {
    for (int m = 0; m < C->ldM; m++)
      if (C->queue[m]) return true;
    return false;
  }
  // stage the queued requests as batches on the free lanes of their class
  // (bit l of the result: lane l staged), then start them
  int stage_queue() {
    int staged = 0;
    for (int l = 0; l < gp_ctx::kLanes; l++)
      if (stage_batch(C, l, fp32, exacth) >= 0) {
        staged |= 1 << l;
        out.refac += C->lanes[l].nb;
        out.factor_launches++;
      }
    return staged;
  }
  void start_queue(int staged) {
    for (int l = 0; l < gp_ctx::kLanes; l++) {
      if (!(staged >> l & 1)) continue;
      if (fp32 && l == 0)
        start_batch<float>(C, l, C->d_luS32.p, C->d_svS32.p, C->d_dnS32.p);
      else
        start_batch<double>(C, l, C->d_luS64.p, C->d_svS64.p, C->d_dnS64.p);
    }
  }
  void launch_queue() { start_queue(stage_queue()); }

  // One round's corrections after eval(), in two halves: launch_pre() puts
  // the Schur kernels and the synchronous factorisation on the streams,
  // launch_post() the staging of asynchronous requests, the solve and the
  // back-substitution.  The scenario-asynchronous loop calls launch_pre()
  // before the round's decisions come back (cnt NULL: every kernel goes out
  // and works from the device masks alone, a launch with an empty mask
  // returning at once), so the device runs it while the host waits for the
  // decisions, and launch_post() with k_norm_finish's counts (cnt[] as
  // n_unconv), which stages this round's requests at once.  absorb() then
  // takes the counts into the statistics and handles the rounds in which
  // every remaining scenario waits for a factorisation.  land_first: a
  // scenario starts a solve that refactorises (an event or a change of dt):
  // a batch in flight that holds one of the scenarios in `fresh` lands first
  // and their queued requests (made for the old system) are dropped.
  int prof_f = -1;
  bool landed = false, skip = false;
  bool mixed = false, run32 = false, run64 = false;
  bool shmix = false;   // shared method: some scenarios run exact Newton on their own factors
  const uint8_t *upd32 = nullptr, *upd64 = nullptr;
  void launch_pre(const int *cnt, bool land_first, const std::vector<uint8_t> *fresh) {
    const int L = C->ldM;
    const double *p = C->d_p.p;
    double *z = C->d_z.p;
    const bool known = cnt != nullptr;
    prof_f = -1;
    mixed = false;
    skip = known && cnt[2] == 0 && cnt[3] == 0;   // every remaining scenario waits (absorb)
    if (skip) return;
    // with the counts unknown: under ASYNC only a solve that starts afresh
    // (an event or a change of dt: land_first) factorises in the round
    // itself, every other request goes to a lane
    const bool fac = !shared && (known ? cnt[1] > 0 : (!async || land_first || C->opt.async_below > 0.0));
    fac_done = fac;
    if (method == GP_METHOD_MONO_FP64) {
      launch_jac_mono(C->T, p, z, dt, faults(C), C->d_jv.p, C->d_upd.p, C->stream, dtv);
      prof_f = prof_begin(C, C->prof_factor, C->stream);
      launch_splu_factor_impl<double>(C->luJ.dev, C->d_jv.p, nullptr, L, C->d_luJ.p, C->d_svJ.p, C->d_dnJ.p,
                                      C->d_refac.p, nullptr, C->M, C->d_status.p, C->stream,
                                      prof_end_ev(C->prof_factor, prof_f, known ? cnt[1] : 0));
      return;
    }
    if (shared) {
      // the shared factorisation serves every scenario except those running
      // exact Newton (exacth), which factorise in FP64 at every iteration
      // (the synchronous factorisation, launched here whatever the counts:
      // an empty mask returns at once)
      fac_done = true;
      shmix = false;
      if (exacth)
        for (int m = 0; m < L && !shmix; m++) shmix = exacth[m] && (!workh || workh[m]);
      run32 = false;
      run64 = true;
      upd32 = upd64 = C->d_upd.p;
      const uint8_t *ref64 = C->d_refac.p;
      if (shmix) {
        launch_split_precision(L, exactv, C->d_sch.p, C->d_refac.p, C->d_upd.p, C->d_split.p, C->stream);
        upd32 = C->d_split.p + 2 * L;
        ref64 = C->d_split.p + 4 * L;
        upd64 = C->d_split.p + 5 * L;
      }
      launch_cluster_schur<double>(C->T, p, z, C->d_f.p, dt, faults(C),
                                   C->d_X64.p, C->d_Dg.p, C->d_sch.p, C->stream, dtv);
      launch_bus_schur(C->T, p, z, C->d_f.p, C->d_Dg.p, faults(C), C->d_svar.p, C->d_rhs.p, C->d_sch.p, C->stream);
      if (shmix) {
        prof_f = prof_begin(C, C->prof_factor, C->stream);
        launch_splu_factor_impl<double>(C->luS64, C->d_svar.p, C->d_s_cst.p, L, C->d_luS64.p, C->d_svS64.p,
                                        C->d_dnS64.p, ref64, C->d_cur.p, C->M, C->d_status.p, C->stream,
                                        prof_end_ev(C->prof_factor, prof_f, known ? cnt[1] : 0));
      }
      HIPCHK(hipGetLastError());
      return;
    }
    // factor precision per scenario: FP32 for the FP32 method, except the
    // scenarios running exact Newton (exacth), whose factors are FP64
    const bool use32 = fp32 && !all64;
    // (a scenario waiting for a factorisation in flight applies nothing this
    // round: no FP64 launches for it)
    if (use32 && exacth && C->has64)
      for (int m = 0; m < L && !mixed; m++) mixed = exacth[m] && (!workh || workh[m]) && !C->pend[m];
    const uint8_t *sch32 = C->d_sch.p, *ref32 = C->d_refac.p;
    upd32 = C->d_upd.p;
    const uint8_t *sch64 = sch32, *ref64 = ref32;
    upd64 = upd32;
    if (mixed) {
      launch_split_precision(L, exactv, C->d_sch.p, C->d_refac.p, C->d_upd.p, C->d_split.p, C->stream);
      const uint8_t *b = C->d_split.p;
      sch32 = b; ref32 = b + L; upd32 = b + 2 * L;
      sch64 = b + 3 * L; ref64 = b + 4 * L; upd64 = b + 5 * L;
    }
    run32 = use32;
    run64 = !use32 || mixed;
    // the FP64 kernels' stream: the side stream beside the FP32 kernels
    // (they touch other scenarios' columns only), else the main stream
    hipStream_t s64 = mixed ? C->side : C->stream;
    if (mixed) {
      HIPCHK(hipEventRecord(C->ev_fork, C->stream));
      HIPCHK(hipStreamWaitEvent(C->side, C->ev_fork, 0));
    }
    if (run64) launch_cluster_schur<double>(C->T, p, z, C->d_f.p, dt, faults(C),
        C->d_X64.p, C->d_Dg.p, sch64, s64, dtv);
    if (run32)
      launch_cluster_schur<float>(C->T, p, z, C->d_f.p, dt, faults(C), C->d_X32.p, C->d_Dg.p, sch32, C->stream, dtv);
    if (mixed) {
      HIPCHK(hipEventRecord(C->ev_side, C->side));
      HIPCHK(hipStreamWaitEvent(C->stream, C->ev_side, 0));
    }
    launch_bus_schur(C->T, p, z, C->d_f.p, C->d_Dg.p, faults(C), C->d_svar.p, C->d_rhs.p, C->d_sch.p, C->stream);
    // synchronous factorisation into the buffer the solves read.  Later
    // iterations ask only for scenarios with nothing pending; the first one
    // of a solve after an event or a change of dt may include scenarios in
    // the batch in flight, which lands first, and their queued requests
    // (made for the old system) are dropped
    if (fac && async && land_first) {
      // a batch in flight that holds a scenario starting a solve of its
      // own lands first (the factorisation below writes the buffer that
      // landing would switch that scenario to); the others stay in flight
      for (int l = 0; l < gp_ctx::kLanes; l++) {
        const gp_ctx::Lane &ln = C->lanes[l];
        bool hit = ln.inflight && !fresh;
        for (int m = 0; m < L && ln.inflight && !hit; m++) hit = ln.batch[m] && (*fresh)[m];
        if (hit) land_lane(C, l, true);
      }
      if (fresh) {   // fresh is on the device as this round's refac0 mask
        for (int m = 0; m < L; m++)
          if ((*fresh)[m]) C->queue[m] = C->pend[m] = 0;
        launch_async_flags(L, 2, C->d_refac0.p, C->d_cur.p, C->d_pend.p, C->stream);
      }
    }
    if (mixed) {   // the FP64 scenarios' factorisation, solve and back-substitution on the side stream
      HIPCHK(hipEventRecord(C->ev_fork2, C->stream));
      HIPCHK(hipStreamWaitEvent(C->side, C->ev_fork2, 0));
      launch_splu_factor_impl<double>(C->luS64, C->d_svar.p, C->d_s_cst.p, L, C->d_luS64.p, C->d_svS64.p,
                                      C->d_dnS64.p, ref64, C->d_cur.p, C->M, C->d_status.p, s64);
      launch_splu_solve<double>(C->luS64, C->d_svS64.p, C->d_dnS64.p, C->d_rhs.p, L, 1.0, C->d_z_of_red.p, z,
                                C->d_dv.p, upd64, C->d_cur.p, C->M, s64);
      launch_cluster_backsub<double>(C->T, C->d_X64.p, C->d_f.p, C->d_dv.p, z, upd64, s64);
      HIPCHK(hipEventRecord(C->ev_join, C->side));
    }
    ref32_ = ref32;
    ref64_ = ref64;
    if (fac) sync_factor(known ? cnt[1] : 0);
    HIPCHK(hipGetLastError());
  }
  // the synchronous factorisation of the main stream (the FP64 one of a
  // mixed round runs on the side stream)
  const uint8_t *ref32_ = nullptr, *ref64_ = nullptr;
  bool fac_done = false;
  void sync_factor(int64_t nscen) {
    const int L = C->ldM;
    // timed: the launches in the method's own precision
    prof_f = (run32 || !fp32) ? prof_begin(C, C->prof_factor, C->stream) : -1;
    hipEvent_t fe = prof_end_ev(C->prof_factor, prof_f, nscen);
    if (run32)
      launch_splu_factor_impl<float>(C->luS.dev, C->d_svar.p, C->d_s_cst.p, L, C->d_luS32.p, C->d_svS32.p,
                                     C->d_dnS32.p, ref32_, C->d_cur.p, C->M, C->d_status.p, C->stream, fe);
    if (run64 && !mixed)
      launch_splu_factor_impl<double>(C->luS64, C->d_svar.p, C->d_s_cst.p, L, C->d_luS64.p, C->d_svS64.p,
                                      C->d_dnS64.p, ref64_, C->d_cur.p, C->M, C->d_status.p, C->stream, fe);
  }
  void launch_post(const int cnt[4]) {
    const int L = C->ldM;
    double *z = C->d_z.p;
    // the batches that have finished by now switch their scenarios' buffers
    // before the solve (none of those scenarios refactorised this round:
    // pending scenarios ask for nothing, and a fresh one's batch landed in
    // launch_pre)
    landed = async && land_batch(C, false);
    if (skip) return;
    if (method == GP_METHOD_MONO_FP64) {
      if (cnt[2] > 0) {
        const int ps = prof_begin(C, C->prof_solve, C->stream);
        launch_splu_solve<double>(C->luJ.dev, C->d_svJ.p, C->d_dnJ.p, C->d_f.p, L, -1.0, nullptr, z, nullptr,
                                  C->d_upd.p, nullptr, C->M, C->stream);
        prof_end(C->prof_solve, ps, C->stream, cnt[2]);
      }
      HIPCHK(hipGetLastError());
      return;
    }
    // a synchronous request launch_pre did not expect (a residual that is
    // not finite under exact Newton, whose request is not sent to a lane):
    // its factorisation goes out now, still before the solve
    if (cnt[1] > 0 && !fac_done && !shared) {
      sync_factor(cnt[1]);
      fac_done = true;
    }
    // asynchronous requests (pending on the device from k_norm_finish on:
    // the scenario waits when its |F| stops decreasing) join the host queue
    // and are staged when a lane is free
    if (async && cnt[3] > 0) take_requests();
    const int lane = (async && free_lane(C) >= 0 && queued()) ? stage_queue() : 0;
    if (cnt[2] > 0) {
      const int ps = prof_begin(C, C->prof_solve, C->stream);
      if (run32)
        launch_splu_solve<float>(C->luS.dev, C->d_svS32.p, C->d_dnS32.p, C->d_rhs.p, L, 1.0, C->d_z_of_red.p, z,
                                 C->d_dv.p, upd32, C->d_cur.p, C->M, C->stream);
      if (shared) {
        const WbTab wb = C->sh.tab();
        launch_splu_solve<double>(C->luS64, C->sh.sv.p, C->sh.dn.p, C->d_rhs.p, L, 1.0, C->d_z_of_red.p, z,
                                  C->d_dv.p, upd32, nullptr, C->M, C->stream, true, &wb);
        if (shmix)
          launch_splu_solve<double>(C->luS64, C->d_svS64.p, C->d_dnS64.p, C->d_rhs.p, L, 1.0, C->d_z_of_red.p, z,
                                    C->d_dv.p, upd64, C->d_cur.p, C->M, C->stream);
      } else if (run64 && !mixed) {
        launch_splu_solve<double>(C->luS64, C->d_svS64.p, C->d_dnS64.p, C->d_rhs.p, L, 1.0, C->d_z_of_red.p, z,
                                  C->d_dv.p, upd64, C->d_cur.p, C->M, C->stream);
      }
      prof_end(C->prof_solve, ps, C->stream, cnt[2]);
      if (shared)
        launch_cluster_backsub<double>(C->T, C->d_X64.p, C->d_f.p, C->d_dv.p, z, C->d_upd.p, C->stream);
      if (run32)
        launch_cluster_backsub<float>(C->T, C->d_X32.p, C->d_f.p, C->d_dv.p, z, upd32, C->stream);
      if (run64 && !mixed && !shared)
        launch_cluster_backsub<double>(C->T, C->d_X64.p, C->d_f.p, C->d_dv.p, z, upd64, C->stream);
    }
    start_queue(lane);
    if (mixed) HIPCHK(hipStreamWaitEvent(C->stream, C->ev_join, 0));   // the side stream's work joins the round
    HIPCHK(hipGetLastError());
  }
  // the round's asynchronous requests (host copy reqp) into the queue
  void take_requests() {
    for (int m = 0; m < C->ldM; m++)
      if (reqp[m]) C->queue[m] = C->pend[m] = 1;
  }
  // After launch() and the counts: false when nothing can change any more.
  bool absorb(const int cnt[4]) {
    out.corrections += cnt[2];
    out.rounds++;
    if (cnt[1] > 0) {
      out.refac += cnt[1];
      out.factor_launches++;
    }
    prof_count(C->prof_factor, prof_f, cnt[1]);
    if (cnt[2] == 0 && cnt[3] == 0) {
      // every remaining scenario waits for its factorisation
      if (async && free_lane(C) >= 0 && queued()) launch_queue();
      if (any_inflight(C))
        land_batch(C, true);
      else if (!landed)
        return false;   // nothing will change (cannot happen)
    }
    return true;
  }
  bool correct(const int cnt[4], bool land_first, const std::vector<uint8_t> *fresh) {
    launch_pre(cnt, land_first, fresh);
    launch_post(cnt);
    return absorb(cnt);
  }
};

// Drops the asynchronous refactorisations asked for by the scenarios in
// `sel` (a batch in flight lands first) and restarts their Newton solves
// from the step's initial iterate (z = zold).  dsel: `sel` on the device.
void restart_solves(gp_ctx *C, const std::vector<uint8_t> &sel, const uint8_t *dsel) {
  const int L = C->ldM;
  land_batch(C, true);
  for (int m = 0; m < L; m++)
    if (sel[m]) C->queue[m] = C->pend[m] = 0;
  launch_async_flags(L, 2, dsel, C->d_cur.p, C->d_pend.p, C->stream);
  launch_copy_active(C->n, L, C->d_zold.p, C->d_z.p, dsel, C->stream);
}

// newton_step! (src/numerics.jl:55-176) for every scenario in `mask`, in lock
// step: each round evaluates F, drops the scenarios with |F|inf < tol (and
// those that failed), and applies one correction to the rest.  With
// `retry_exact`, a scenario whose solve failed under a policy that reuses
// factorisations solves again from the step's initial iterate as exact
// Newton (a factorisation at every iteration, the reference's method) before
// it counts as failed.
NewtonOut newton(gp_ctx *C, int method, double dt, double tol, int itermax, const std::vector<uint8_t> &mask,
                 const std::vector<uint8_t> *refac0 = nullptr, bool retry_exact = false, bool exact = false) {
  const int L = C->ldM;
  Newton N(C, method, dt, tol, itermax, refac0 != nullptr);
  if (exact) N.make_exact();
  if (!retry_exact) N.switch_window = 0;
  set_mask(C, C->d_work, mask);
  if (N.reuse) set_mask(C, C->d_refac0, *refac0);
  HIPCHK(hipMemsetAsync(C->d_failed.p, 0, L, C->stream));
  // a scenario may sit out rounds while its factorisation is in flight, so
  // rounds are bounded separately from each scenario's itermax corrections
  const int max_rounds = 4 * itermax + 16;
  bool done = false;
  for (int it = 0; it < max_rounds; it++) {
    int cnt[4] = {0, 0, 0, 0};
    N.eval(it == 0, nullptr, C->d_refac0.p, cnt, it > 0);
    HIPCHK(hipStreamSynchronize(C->stream));
    if (cnt[0] == 0) {
      done = true;
      break;
    }
    if (!N.correct(cnt, it == 0, refac0)) break;
  }
  // scenarios stopped by the itermax / stall rules, plus (round cap) any
  // still working
  NewtonOut out = N.out;
  std::vector<uint8_t> w(L, 0), fl(L, 0);
  if (!done) HIPCHK(hipMemcpyAsync(w.data(), C->d_work.p, L, hipMemcpyDeviceToHost, C->stream));
  HIPCHK(hipMemcpyAsync(fl.data(), C->d_failed.p, L, hipMemcpyDeviceToHost, C->stream));
  HIPCHK(hipStreamSynchronize(C->stream));
  bool any = false;
  for (int m = 0; m < L; m++) {
    w[m] = (w[m] || fl[m]) ? 1 : 0;
    any = any || w[m];
  }
  if (any) out.failed = w;
  if (any && retry_exact && N.retryable() && !exact) {
    set_mask(C, C->d_sel, w);
    restart_solves(C, w, C->d_sel.p);
    const NewtonOut r = newton(C, method, dt, tol, itermax, w, nullptr, false, true);
    out.corrections += r.corrections;
    out.refac += r.refac;
    out.rounds += r.rounds;
    out.factor_launches += r.factor_launches;
    out.exact_retries += r.exact_retries;
    for (int m = 0; m < L; m++) out.exact_retries += w[m];
    out.failed = r.failed;
  }
  return out;
}

// Symbolic analysis and buffers of `method`.
void prepare_method(gp_ctx *C, int method) {
  analyse(C, method);
  if (method != GP_METHOD_MONO_FP64) ensure_schur_buffers(C, method == GP_METHOD_SCHUR_FP32);
  else C->luJ.shape(sizeof(double));
}

// Column 0 of the recorded trajectory: the initial state of every scenario.
void record_initial(gp_ctx *C) {
  std::vector<uint8_t> all(C->ldM, 0);
  std::fill(all.begin(), all.begin() + C->M, 1);
  set_mask(C, C->d_act, all);
  launch_record_impl(C->nrec, C->rec_idx.empty() ? nullptr : C->d_rec_idx.p, C->ldM, C->M, C->d_z.p, C->d_hist.p,
                     C->d_act.p, C->stream);
  HIPCHK(hipStreamSynchronize(C->stream));   // the host mask stays valid until then
}

// Device time between construction and ms() on one stream.
struct DevTimer {
  hipStream_t s;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  explicit DevTimer(hipStream_t s_) : s(s_) {
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s));
  }
  double ms() {
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    return t;
  }
  ~DevTimer() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

// Zero pivots reported by the factorisations of the sweep, then the
// scenario outcome as the call's status.
int finish_status(gp_ctx *C, const gp_stats &S) {
  const int L = C->ldM, M = C->M;
  std::vector<int> status(L);
  HIPCHK(hipMemcpy(status.data(), C->d_status.p, L * sizeof(int), hipMemcpyDeviceToHost));
  int nsing = 0;
  for (int m = 0; m < M; m++)
    if (status[m] != 0) {
      nsing++;
      C->scen_status[m] = GP_ERR_SINGULAR_PIVOT;
    }
  HIPCHK(hipMemset(C->d_status.p, 0, L * sizeof(int)));
  if (nsing) return fail(GP_ERR_SINGULAR_PIVOT, std::to_string(nsing) + " scenario(s) hit a zero or non-finite pivot");
  if (S.n_nan) return fail(GP_ERR_NAN, std::to_string(S.n_nan) + " scenario(s) produced a non-finite residual");
  if (S.n_not_converged)
    return fail(GP_ERR_NEWTON_NOT_CONVERGED,
                std::to_string(S.n_not_converged) + " scenario(s) had a Newton solve that did not converge");
  return GP_OK;
}

// The time loop of integrate! (src/dynamics.jl:548-640) in lock step
// (gp_options.lock_step), one step at a time for every scenario still
// running: the Newton solve of step k, the record of z_k, then the fault
// events at step k (activation, then deactivation) and the dt = 0 re-solve
// of the scenarios they touch.
struct Stepper {
  gp_ctx *C;
  int method, itermax;
  double dt, tol;
  std::map<int, std::vector<const gp_event *>> on_at, off_at;
  // per-scenario state
  struct State {
    std::vector<std::vector<const gp_event *>> active_ev;   // active faults, in event order
    std::vector<uint8_t> dead, everfail;
    std::vector<uint8_t> stale;   // factorisation made with another dt or before an event
    std::vector<int> fbus;        // fault tables [MAXFAULT][ldM]
    std::vector<double> fg;
    int ftab = 0;                 // index of the fault tables in tab_bus / tab_g
  } st;
  std::vector<uint8_t> act, res, need0;
  gp_stats S{};
  // every fault table the sweep used (version 0: no fault), kept so that
  // the asynchronous uploads read memory that stays valid
  std::vector<std::vector<int>> tab_bus;
  std::vector<std::vector<double>> tab_g;

  Stepper(gp_ctx *c, int meth, double dt_, double tol_, int itmax)
      : C(c), method(meth), itermax(itmax), dt(dt_), tol(tol_) {
    const int L = C->ldM, M = C->M;
    for (auto &e : C->events) {
      on_at[e.step_on].push_back(&e);
      off_at[e.step_off].push_back(&e);
    }
    st.active_ev.assign(M, {});
    st.dead.assign(L, 0);
    st.everfail.assign(L, 0);
    st.stale.assign(L, 1);
    st.fbus.assign((size_t)MAXFAULT * L, -1);
    st.fg.assign((size_t)MAXFAULT * L, 0.0);
    act.assign(L, 0);
    res.assign(L, 0);
    need0.assign(L, 0);
    tab_bus.push_back(st.fbus);
    tab_g.push_back(st.fg);
    upload_faults();
  }
  void upload_faults() {
    HIPCHK(hipMemcpyAsync(C->d_fbus.p, tab_bus[st.ftab].data(), st.fbus.size() * sizeof(int), hipMemcpyHostToDevice,
                          C->stream));
    HIPCHK(hipMemcpyAsync(C->d_fg.p, tab_g[st.ftab].data(), st.fg.size() * sizeof(double), hipMemcpyHostToDevice,
                          C->stream));
  }
  void add_stats(const NewtonOut &no) {
    S.newton_iters += no.corrections;
    S.refactorizations += no.refac;
    S.newton_rounds += no.rounds;
    S.factor_launches += no.factor_launches;
    S.exact_retries += no.exact_retries;
  }
  // Step k for the scenarios with k <= nsteps[m] that have not stopped.
  // hist_k: recorded channels of z_k (NULL: none); zsave: all of z_k,
  // scenario-major (NULL: none).  Returns false when no scenario ran.
  bool step(int k, double *hist_k, double *zsave) {
    const int L = C->ldM, M = C->M;
    for (int m = 0; m < M; m++) act[m] = (k <= C->nsteps[m] && !st.dead[m]) ? 1 : 0;
    int nact = 0;
    for (int m = 0; m < M; m++) nact += act[m];
    if (nact == 0) return false;
    set_mask(C, C->d_act, act);
    launch_copy_active(C->n, L, C->d_z.p, C->d_zold.p, C->d_act.p, C->stream);
    for (int m = 0; m < L; m++) need0[m] = act[m] && st.stale[m];
    NewtonOut no = newton(C, method, dt, tol, itermax, act, &need0, true);
    for (int m = 0; m < M; m++)
      if (act[m]) st.stale[m] = 0;
    add_stats(no);
    S.scenario_steps += nact;
    if (!no.failed.empty())
      for (int m = 0; m < M; m++)
        if (no.failed[m]) {
          S.scenario_steps--;
          st.everfail[m] = 1;
          st.dead[m] = 1;   // a scenario whose Newton solve failed is stopped and reported
          C->scen_status[m] = GP_ERR_NEWTON_NOT_CONVERGED;
          C->scen_steps[m] = k - 1;
        }
    if (hist_k)
      launch_record_impl(C->nrec, C->rec_idx.empty() ? nullptr : C->d_rec_idx.p, L, M, C->d_z.p, hist_k,
                         C->d_act.p, C->stream);
    if (zsave) launch_copy_active(C->n, L, C->d_z.p, zsave, C->d_act.p, C->stream);
    // events at step k: activation then deactivation (src/dynamics.jl:599-629)
    bool changed = false;
    std::fill(res.begin(), res.end(), 0);
    auto ion = on_at.find(k);
    if (ion != on_at.end())
      for (const gp_event *e : ion->second) {
        if (!act[e->scenario]) continue;
        st.active_ev[e->scenario].push_back(e);
        res[e->scenario] = 1;
        changed = true;
      }
    auto ioff = off_at.find(k);
    if (ioff != off_at.end())
      for (const gp_event *e : ioff->second) {
        if (!act[e->scenario]) continue;
        auto &v = st.active_ev[e->scenario];
        auto it = std::find(v.begin(), v.end(), e);
        if (it != v.end()) v.erase(it);
        res[e->scenario] = 1;
        changed = true;
      }
    if (changed) {
      for (int m = 0; m < M; m++) {
        if (!res[m]) continue;
        if ((int)st.active_ev[m].size() > MAXFAULT)
          throw GpError{GP_ERR_UNSUPPORTED, "more than " + std::to_string(MAXFAULT) + " simultaneous faults, line trips or load drops in one scenario"};
        for (int q = 0; q < MAXFAULT; q++) {
          const bool on = q < (int)st.active_ev[m].size();
          st.fbus[(size_t)q * L + m] = on ? st.active_ev[m][q]->bus : -1;
          st.fg[(size_t)q * L + m] = on ? 1.0 / st.active_ev[m][q]->r_fault : 0.0;
        }
      }
      // a new version of the tables; kept, so that the asynchronous copies
      // read memory that stays valid and a backward sweep can reuse them
      tab_bus.push_back(st.fbus);
      tab_g.push_back(st.fg);
      st.ftab = (int)tab_bus.size() - 1;
      upload_faults();
      int nres = 0;
      for (int m = 0; m < M; m++) nres += res[m];
      S.event_resolves += nres;
      // re-solve at dt = 0: differential states frozen, algebraic states re-solved
      set_mask(C, C->d_sel, res);
      launch_copy_active(C->n, L, C->d_z.p, C->d_zold.p, C->d_sel.p, C->stream);
      NewtonOut nr = newton(C, method, 0.0, tol, itermax, res, &res, true);
      for (int m = 0; m < M; m++)
        if (res[m]) st.stale[m] = 1;   // factorised at dt = 0: the next step refactorises
      add_stats(nr);
      if (!nr.failed.empty())
        for (int m = 0; m < M; m++)
          if (nr.failed[m]) {
            st.everfail[m] = 1;
            st.dead[m] = 1;
            C->scen_status[m] = GP_ERR_NEWTON_NOT_CONVERGED;
            C->scen_steps[m] = k;
          }
    }
    // non-finite states stop a scenario
    if (!no.failed.empty()) {
      std::vector<double> fn(L);
      HIPCHK(hipMemcpy(fn.data(), C->d_fnorm.p, L * sizeof(double), hipMemcpyDeviceToHost));
      for (int m = 0; m < M; m++)
        if (no.failed[m] && !std::isfinite(fn[m])) {
          st.dead[m] = 1;
          S.n_nan++;
          C->scen_status[m] = GP_ERR_NAN;
        }
    }
    return true;
  }
};


// What a scenario-asynchronous sweep runs and keeps.
// A scenario stopped by an exactly zero pivot: it continues from the
// start of solve (k, mode) -- mode 1: step k's dt = 0 re-solve under the
// events `active` -- on a plan of its own (repivot_runs).
struct RepivotReq {
  int m = 0, k = 0, mode = 0;
  std::vector<const gp_event *> active;
};

struct FreeSpec {
  std::vector<int> k_first, k_last;   // per scenario: steps k_first..k_last (none when k_last < k_first)
  std::vector<std::vector<const gp_event *>> active0;   // faults in force at k_first (empty: none)
  std::vector<uint8_t> resolve0;      // per scenario (empty: none): the first solve is step k_first's
                                      // dt = 0 re-solve (its events already in active0)
  std::vector<RepivotReq> *repivot = nullptr;   // scenarios stopped by an exactly zero pivot go here
                                                // instead of failing
  bool record = true;                 // recorded channels of z_k into trajectory column k
  double *seg = nullptr;              // all of z_k, seg_k0 < k <= seg_k1, into slot k - seg_k0 - 1
  int seg_k0 = 0, seg_k1 = -1;
  double *ckpt = nullptr;             // the state after step s*every (its events included) into slot s,
  int every = 0, nck = 0;             // 1 <= s < nck
};

// The time loop with every scenario advancing on its own: each scenario
// follows integrate! (src/dynamics.jl:548-640) step by step, but a round
// evaluates F for every running scenario at its own step k_m, and the
// scenarios that converged accept their step at once (record z_k, fire
// their events of step k with the dt = 0 re-solve, or start step k+1) while
// the others receive a correction.  No scenario waits for a slower one, so
// the rounds of a call are those of its slowest scenario, not the sum over
// steps of the slowest scenario of each step.
gp_stats integrate_free(gp_ctx *C, int method, double dt, double tol, int itermax, const FreeSpec &F) {
  const int L = C->ldM, M = C->M, n = C->n;
  const size_t nL = (size_t)n * L;
  gp_stats S{};
  std::vector<std::map<int, std::vector<const gp_event *>>> on(M), off(M);
  for (auto &e : C->events) {
    on[e.scenario][e.step_on].push_back(&e);
    off[e.scenario][e.step_off].push_back(&e);
  }
  std::vector<std::vector<const gp_event *>> active_ev(M);
  std::vector<int> k(L, 1);
  std::vector<uint8_t> mode(L, 0), run(L, 0), first(L, 0), stale(L, 1), fresh(L, 0), wk(L, 0), fl(L, 0),
      everfail(L, 0), exact(L, 0), restart(L, 0);
  const RoundLayout R(L);
  uint8_t *hr = C->h_round.p, *hp = C->h_round.p;   // pinned staging (RoundLayout)
  uint8_t *acc = hp + R.acc(), *adv = hp + R.adv(), *sv = hp + R.sv(), *ck = hp + R.ck(), *pred = hp + R.pred();
  int *col = reinterpret_cast<int *>(hp + R.col()), *svslot = reinterpret_cast<int *>(hp + R.svslot()),
      *ckslot = reinterpret_cast<int *>(hp + R.ckslot());
  std::vector<double> dtv(L, dt);
  std::vector<int> fbus((size_t)MAXFAULT * L, -1);
  std::vector<double> fg((size_t)MAXFAULT * L, 0.0);
  const bool shared = method == GP_METHOD_SCHUR_SHARED;
  auto set_faults = [&](int m) {
    if ((int)active_ev[m].size() > MAXFAULT)
      throw GpError{GP_ERR_UNSUPPORTED, "more than " + std::to_string(MAXFAULT) + " simultaneous faults, line trips or load drops in one scenario"};
    for (int q = 0; q < MAXFAULT; q++) {
      const bool f_on = q < (int)active_ev[m].size();
      fbus[(size_t)q * L + m] = f_on ? active_ev[m][q]->bus : -1;
      fg[(size_t)q * L + m] = f_on ? 1.0 / active_ev[m][q]->r_fault : 0.0;
    }
    if (shared) shared_set(C, m, fbus.data(), fg.data());
  };
  int kmax = 0;
  for (int m = 0; m < M; m++) {
    k[m] = F.k_first[m];
    run[m] = F.k_first[m] <= F.k_last[m] ? 1 : 0;
    first[m] = 1;
    if (!F.resolve0.empty() && F.resolve0[m]) {
      mode[m] = 1;
      dtv[m] = 0.0;
    }
    kmax = std::max(kmax, F.k_last[m]);
    if (!F.active0.empty()) active_ev[m] = F.active0[m];
    set_faults(m);
  }
  std::memcpy(adv, run.data(), L);
  std::memset(pred, 0, L);
  HIPCHK(hipMemcpyAsync(C->d_round.p, hp, R.P(), hipMemcpyHostToDevice, C->stream));
  launch_copy_active(n, L, C->d_z.p, C->d_zold.p, C->d_adv.p, C->stream);
  HIPCHK(hipMemsetAsync(C->d_failed.p, 0, L, C->stream));
  HIPCHK(hipStreamSynchronize(C->stream));   // the staging block stays valid until then
  std::memset(adv, 0, L);
  Newton N(C, method, dt, tol, itermax, true);
  N.dtv = C->d_dtv.p;
  N.exactv = C->d_exact.p;
  N.exacth = exact.data();
  // first iterate of a step by extrapolation (gp_options.predictor) under
  // the policies that reuse factorisations
  const bool extrapolate = (C->opt.predictor == 1 || C->opt.predictor == 2) && N.reuse;
  // quadratic extrapolation (predictor 2) after two smooth steps: z_(k-2)
  // is kept in zold2; nplain counts the accepted steps since the start or
  // since the last event's re-solve
  const bool quad = extrapolate && C->opt.predictor == 2;
  if (quad) C->d_zold2.ensure(nL);
  std::vector<int> nplain(L, 0);
  bool fault_dirty = true;
  const long long max_rounds = (long long)(kmax + 1) * 2 * (4 * itermax + 16);
  // the previous round's accepted steps (record, segment and checkpoint
  // copies, advance), launched after the copy that uploads them together
  // with this round's flags
  bool any_acc = false, any_adv = false, any_sv = false, any_ck = false;
  auto launch_accepted = [&]() {
    if (any_sv) launch_copy_cols(n, L, C->d_z.p, F.seg, nL, C->d_svslot.p, C->d_sv.p, C->stream);
    if (any_ck) launch_copy_cols(n, L, C->d_z.p, F.ckpt, nL, C->d_ckslot.p, C->d_ck.p, C->stream);
    if (any_acc || any_adv)
      launch_advance(n, L, C->d_z.p, C->d_zold.p, quad ? C->d_zold2.p : nullptr, C->d_adv.p, C->d_pred.p,
                     any_acc ? C->d_acc.p : nullptr, C->d_col.p, C->rec_idx.empty() ? nullptr : C->d_rec_p.p,
                     C->d_rec_ch.p, C->nrec, C->d_hist.p, C->stream);
  };
  for (long long round = 0;; round++) {
    int nrun = 0;
    bool land_first = false;
    double *hdt = reinterpret_cast<double *>(hr + R.dt());
    for (int m = 0; m < L; m++) {
      nrun += run[m];
      fresh[m] = (run[m] && first[m] && stale[m]) ? 1 : 0;
      land_first = land_first || fresh[m];
      hr[R.first() + m] = first[m];
      hr[R.refac0() + m] = fresh[m];
      hr[R.exact() + m] = exact[m];
      hr[R.work() + m] = run[m];
      hdt[m] = dtv[m];
    }
    if (nrun == 0) {
      if (any_acc || any_sv || any_ck) {
        HIPCHK(hipMemcpyAsync(C->d_round.p, hp, R.P(), hipMemcpyHostToDevice, C->stream));
        launch_accepted();
      }
      break;
    }
    if (round >= max_rounds) throw GpError{GP_ERR_HIP, "internal: scenario-asynchronous loop did not finish"};
    S.newton_rounds++;
    std::memset(hr + R.cnt(), 0, 16);
    std::memset(hr + R.failed(), 0, L);
    HIPCHK(hipMemcpyAsync(C->d_round.p, hr, R.up_bytes(), hipMemcpyHostToDevice, C->stream));
    launch_accepted();
    if (fault_dirty) {
      HIPCHK(hipMemcpyAsync(C->d_fbus.p, fbus.data(), fbus.size() * sizeof(int), hipMemcpyHostToDevice, C->stream));
      HIPCHK(hipMemcpyAsync(C->d_fg.p, fg.data(), fg.size() * sizeof(double), hipMemcpyHostToDevice, C->stream));
      if (shared) shared_upload(C);
      fault_dirty = false;
    }
    N.eval(0, C->d_first.p, C->d_refac0.p, nullptr, true);
    HIPCHK(hipMemcpyAsync(hr + R.down_off(), C->d_round.p + R.down_off(), R.down_bytes(), hipMemcpyDeviceToHost,
                          C->stream));
    HIPCHK(hipEventRecord(C->ev_round, C->stream));
    // the round's corrections go out before its decisions come back (the
    // device masks select their scenarios), so the device runs them while
    // the host reads the decisions and accepts the converged scenarios
    N.workh = run.data();
    N.launch_pre(nullptr, land_first, &fresh);
    spin_wait(C->ev_round);
    int cnt[4];
    std::memcpy(cnt, hr + R.cnt(), sizeof(cnt));
    std::memcpy(wk.data(), hr + R.work(), L);
    std::memcpy(fl.data(), hr + R.failed(), L);
    N.reqp = hr + R.areq();
    N.launch_post(cnt);
    if (cnt[0] > 0 && !N.absorb(cnt)) throw GpError{GP_ERR_HIP, "internal: no scenario can make progress"};
    std::memset(acc, 0, L);
    std::memset(adv, 0, L);
    std::memset(sv, 0, L);
    std::memset(ck, 0, L);
    std::memset(pred, 0, L);
    any_acc = any_adv = any_sv = any_ck = false;
    bool any_fail = false, any_restart = false;
    std::fill(restart.begin(), restart.end(), 0);
    // zero or non-finite pivots (the factorisations' status bits 1, 2), read
    // when a solve failed
    std::vector<int> pst;
    if (F.repivot)
      for (int m = 0; m < M && pst.empty(); m++)
        if (run[m] && fl[m]) {
          pst.resize(L);
          HIPCHK(hipMemcpyAsync(pst.data(), C->d_status.p, L * sizeof(int), hipMemcpyDeviceToHost, C->stream));
          HIPCHK(hipStreamSynchronize(C->stream));
        }
    for (int m = 0; m < M; m++) {
      if (!run[m]) continue;
      first[m] = 0;
      if (fl[m] && !pst.empty() && (pst[m] & 1)) {
        // the static pivot sequence met an exactly zero pivot: the scenario
        // leaves this run and continues on a plan of its own
        run[m] = 0;
        RepivotReq q;
        q.m = m;
        q.k = k[m];
        q.mode = mode[m];
        q.active = active_ev[m];
        F.repivot->push_back(q);
        C->queue[m] = 0;   // its queued refactorisation is not staged any more
        continue;
      }
      if (fl[m] && N.retryable() && !exact[m]) {
        // a solve on reused or FP32 factorisations failed: it runs again
        // from the step's initial iterate as exact Newton with FP64 factors
        // (the reference's method)
        exact[m] = 1;
        first[m] = 1;
        stale[m] = 1;
        restart[m] = 1;
        any_restart = true;
        S.exact_retries++;
        continue;
      }
      if (fl[m]) {   // itermax or stall rule: the scenario stops and is reported
        run[m] = 0;
        everfail[m] = 1;
        any_fail = true;
        C->scen_status[m] = GP_ERR_NEWTON_NOT_CONVERGED;
        C->scen_steps[m] = mode[m] ? k[m] : k[m] - 1;
        continue;
      }
      if (wk[m]) continue;   // still iterating
      const bool was_exact = exact[m] != 0;
      exact[m] = 0;          // solve accepted
      bool next = true;
      const int mode_was = mode[m];
      if (mode[m] == 0) {    // step k accepted: record z_k, then its events
        if (F.record) {
          acc[m] = 1;
          col[m] = k[m];
          any_acc = true;
        }
        if (F.seg && k[m] > F.seg_k0 && k[m] <= F.seg_k1) {
          sv[m] = 1;
          svslot[m] = k[m] - F.seg_k0 - 1;
          any_sv = true;
        }
        stale[m] = was_exact ? 1 : 0;   // after an exact solve the next step factorises afresh
        S.scenario_steps++;
        bool changed = false;
        auto ion = on[m].find(k[m]);
        if (ion != on[m].end())
          for (const gp_event *e : ion->second) {
            active_ev[m].push_back(e);
            changed = true;
          }
        auto ioff = off[m].find(k[m]);
        if (ioff != off[m].end())
          for (const gp_event *e : ioff->second) {
            auto &v = active_ev[m];
            auto it = std::find(v.begin(), v.end(), e);
            if (it != v.end()) v.erase(it);
            changed = true;
          }
        nplain[m] = changed ? 0 : nplain[m] + 1;
        if (changed) {       // re-solve at dt = 0 under the new fault state
          set_faults(m);
          fault_dirty = true;
          S.event_resolves++;
          mode[m] = 1;
          dtv[m] = 0.0;
          first[m] = 1;
          stale[m] = 1;
          adv[m] = 1;
          any_adv = true;
          next = false;
        }
      } else {               // re-solve done: factorised at dt = 0, the next step refactorises
        mode[m] = 0;
        stale[m] = 1;
      }
      if (next) {            // step k complete, events included
        if (F.ckpt && F.every > 0 && k[m] % F.every == 0 && k[m] / F.every < F.nck) {
          ck[m] = 1;
          ckslot[m] = k[m] / F.every;
          any_ck = true;
        }
        // z_k - zold is the increment of a step of length dt unless z_k
        // came from an event's re-solve
        const bool plain = mode_was == 0;
        k[m]++;
        if (k[m] > F.k_last[m]) {
          run[m] = 0;
        } else {
          dtv[m] = dt;
          first[m] = 1;
          adv[m] = 1;
          pred[m] = (extrapolate && plain) ? ((quad && nplain[m] >= 2) ? 2 : 1) : 0;
          any_adv = true;
        }
      }
    }
    if (any_restart) {
      set_mask(C, C->d_sel, restart);
      restart_solves(C, restart, C->d_sel.p);
    }
    // scenarios that stopped: non-finite |F| (failed flag bit 1)
    if (any_fail)
      for (int m = 0; m < M; m++)
        if (fl[m] && everfail[m] && C->scen_status[m] == GP_ERR_NEWTON_NOT_CONVERGED && !run[m] && (fl[m] & 2)) {
          S.n_nan++;
          C->scen_status[m] = GP_ERR_NAN;
        }
  }
  HIPCHK(hipStreamSynchronize(C->stream));
  S.newton_iters = N.out.corrections;
  S.refactorizations = N.out.refac;
  S.factor_launches = N.out.factor_launches;
  for (int m = 0; m < M; m++) S.n_not_converged += everfail[m];
  return S;
}

// ---- pivot fallback -------------------------------------------------------
// The reference's Newton refactors its KLU factorisation in place and, when
// that meets a zero pivot (SingularException), factorises afresh with new
// pivots (src/numerics.jl:144-167).  Here one static pivot sequence serves a
// batch; a scenario whose factorisation meets an exactly zero pivot
// leaves the batch (integrate_free) and continues from the start of the
// failed solve, alone, on a plan analysed for its own matrix.

// Values of scenario m's matrix (J in CSC order, or S) at its current state
// with dt = 0, the matrix kind the plans' pivots are chosen on, evaluated by
// the product's kernels under the scenario's event slots.
std::vector<double> scenario_matrix(gp_ctx *C, int method, int m) {
  const int L = C->ldM;
  std::vector<uint8_t> sel(L, 0);
  sel[m] = 1;
  set_mask(C, C->d_sel, sel);
  if (method == GP_METHOD_MONO_FP64) {
    const int nnz = C->jcp[C->n];
    HIPCHK(hipMemsetAsync(C->d_jv.p, 0, (size_t)nnz * L * sizeof(double), C->stream));
    launch_jac_mono(C->T, C->d_p.p, C->d_z.p, 0.0, faults(C), C->d_jv.p, C->d_sel.p, C->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> jv(nnz);
    HIPCHK(hipMemcpy2DAsync(jv.data(), sizeof(double), C->d_jv.p + m, (size_t)L * sizeof(double), sizeof(double),
                            nnz, hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return jv;
  }
  // at dt = 0 the matrix does not depend on zold (z stands in for it)
  launch_cluster_residual(C->T, C->d_p.p, C->d_z.p, C->d_z.p, 0.0, faults(C),
                          C->d_f.p, C->d_inj.p, C->d_sel.p, C->stream);
  launch_bus_residual(C->T, C->d_p.p, C->d_z.p, C->d_inj.p, faults(C), C->d_f.p, C->d_sel.p, C->stream);
  launch_cluster_schur<double>(C->T, C->d_p.p, C->d_z.p, C->d_f.p, 0.0, faults(C),
                               C->d_X64.p, C->d_Dg.p, C->d_sel.p, C->stream);
  launch_bus_schur(C->T, C->d_p.p, C->d_z.p, C->d_f.p, C->d_Dg.p, faults(C), C->d_svar.p, C->d_rhs.p, C->d_sel.p,
                   C->stream);
  HIPCHK(hipGetLastError());
  const int nv = std::max(1, C->n_svar);
  std::vector<double> sv(nv);
  HIPCHK(hipMemcpy2DAsync(sv.data(), sizeof(double), C->d_svar.p + m, (size_t)L * sizeof(double), sizeof(double), nv,
                          hipMemcpyDeviceToHost, C->stream));
  HIPCHK(hipStreamSynchronize(C->stream));
  const int nnzS = C->scp[C->nred];
  std::vector<double> sx(nnzS);
  for (int e = 0; e < nnzS; e++) {
    const int q = C->s_src[e];
    sx[e] = q >= 0 ? sv[q] : C->s_cst[-q - 1];
  }
  return sx;
}

// The method's plan made for scenario m: pivots large in its own matrices
// (own[0] the latest, own[1..] those of its earlier pivot failures), in the
// batch plan's representative matrix and in the matrix of every set of lines
// m trips; the dense block within the batch plan's bound (no timing).  The
// next gp_integrate makes the batch plan again.
void analyse_own(gp_ctx *C, int method, int m, const std::vector<std::vector<double>> &own) {
  const bool mono = method == GP_METHOD_MONO_FP64;
  LUPlan &P = mono ? C->luJ : C->luS;
  int cap = P.tuned_cap >= 0 ? P.tuned_cap : C->opt.dense_block;
  if (cap == 0) cap = -1;
  const std::vector<double> &r0 = mono ? C->rep0_J : C->rep0_S;
  std::vector<std::vector<double>> extra(own.begin() + 1, own.end());
  extra.push_back(r0);
  for (auto &t : trip_matrices(C, r0, !mono, m)) extra.push_back(std::move(t));
  if (mono) {
    std::string e = splu_analyse(C->n, C->jcp.data(), C->jri.data(), own[0].data(), C->luJ.A, 1e-3, ptrs(extra));
    if (!e.empty()) throw GpError{GP_ERR_SINGULAR_PIVOT, "Jacobian analysis of scenario " + std::to_string(m) + ": " + e};
    const int nnz = C->jcp[C->n];
    std::vector<int> asrc(nnz);
    for (int i = 0; i < nnz; i++) asrc[i] = i;
    C->luJ.tune(asrc, own[0], nullptr, sizeof(double), 1.0, C->stream, cap);
    if (C->luJ.dev.gx) {
      C->d_gxJ.ensure((size_t)C->n * C->ldM);
      C->luJ.dev.gx_ws = C->d_gxJ.p;
    }
    C->d_luJ.alloc(C->luJ.lu_elems() * C->M);
    C->d_svJ.alloc((size_t)C->luJ.dev.ns * C->M);
    HIPCHK(hipMemset(C->d_svJ.p, 0, C->d_svJ.n * sizeof(double)));   // padding slots stay zero
    C->d_dnJ.alloc((size_t)std::max(1, C->luJ.A.nd * C->luJ.A.nd) * C->M);
    C->luJ.shape(sizeof(double));
  } else {
    std::string e = splu_analyse(C->nred, C->scp.data(), C->sri.data(), own[0].data(), C->luS.A, 1e-3, ptrs(extra));
    if (!e.empty())
      throw GpError{GP_ERR_SINGULAR_PIVOT, "reduced-network analysis of scenario " + std::to_string(m) + ": " + e};
    const bool fp32 = method == GP_METHOD_SCHUR_FP32;
    C->luS.tune(C->s_src, {}, C->d_s_cst.p, fp32 ? sizeof(float) : sizeof(double), 2.0, C->stream, cap);
    if (C->luS.dev.gx) {
      C->d_gxS.ensure((size_t)C->nred * C->ldM);
      C->luS.dev.gx_ws = C->d_gxS.p;
    }
    ensure_schur_buffers(C, fp32);
    // the new plan's padding slots of the solve arrays must read zero
    HIPCHK(hipMemset(C->d_svS64.p, 0, C->d_svS64.n * sizeof(double)));
    if (fp32) HIPCHK(hipMemset(C->d_svS32.p, 0, C->d_svS32.n * sizeof(float)));
  }
  P.ready = true;
  P.trip_gen = -2;   // not the batch plan: the next call analyses again
}

gp_stats integrate_free(gp_ctx *C, int method, double dt, double tol, int itermax, const FreeSpec &F);

// Runs the scenarios of `reqs` one after another on plans of their own, each
// from the start of its failed solve to its horizon; a scenario that meets a
// zero pivot again is analysed once more with that matrix added (at most
// kTries plans).  Their statistics join S.
void repivot_runs(gp_ctx *C, int method, double dt, double tol, int itermax, std::vector<RepivotReq> reqs,
                  gp_stats &S) {
  constexpr int kTries = 4;
  const int L = C->ldM, M = C->M;
  land_batch(C, true);
  HIPCHK(hipStreamSynchronize(C->stream));
  std::map<int, std::vector<std::vector<double>>> own;
  while (!reqs.empty()) {
    const RepivotReq q = reqs.front();
    reqs.erase(reqs.begin());
    const int m = q.m;
    // back to the start of the failed solve (zold: z_(k-1), or z_k for a
    // re-solve); the scenario's event slots are those of that solve
    std::vector<uint8_t> sel(L, 0);
    sel[m] = 1;
    set_mask(C, C->d_sel, sel);
    launch_copy_active(C->n, L, C->d_zold.p, C->d_z.p, C->d_sel.p, C->stream);
    if ((int)own[m].size() >= kTries) {   // its status bit stays: reported as a singular pivot
      C->scen_steps[m] = q.mode ? q.k : q.k - 1;
      continue;
    }
    HIPCHK(hipMemsetAsync(C->d_status.p + m, 0, sizeof(int), C->stream));   // every batch has landed
    own[m].insert(own[m].begin(), scenario_matrix(C, method, m));
    analyse_own(C, method, m, own[m]);
    FreeSpec F;
    F.k_first.assign(M, 1);
    F.k_last.assign(M, 0);
    F.k_first[m] = q.k;
    F.k_last[m] = C->nsteps[m];
    F.active0.assign(M, {});
    F.active0[m] = q.active;
    F.resolve0.assign(M, 0);
    F.resolve0[m] = (uint8_t)q.mode;
    std::vector<RepivotReq> again;
    F.repivot = &again;
    const gp_stats s2 = integrate_free(C, method, dt, tol, itermax, F);
    land_batch(C, true);
    S.newton_iters += s2.newton_iters;
    S.scenario_steps += s2.scenario_steps;
    S.event_resolves += s2.event_resolves;
    S.refactorizations += s2.refactorizations;
    S.n_not_converged += s2.n_not_converged;
    S.n_nan += s2.n_nan;
    S.newton_rounds += s2.newton_rounds;
    S.factor_launches += s2.factor_launches;
    S.exact_retries += s2.exact_retries;
    S.repivots++;
    for (const RepivotReq &r : again) reqs.insert(reqs.begin(), r);
  }
  C->luS.ready = C->luJ.ready = false;
}

// The fault state in force during each step of every scenario, from the
// event lists (activation then deactivation at the end of a step, as the
// time loop applies them): one table version per change, the version of
// every step, and the active lists at the requested steps.
struct EventTimeline {
  std::vector<int> ver;                           // [k] -> version in force during step k
  std::vector<std::vector<int>> bus;              // per version: [MAXFAULT][ldM]
  std::vector<std::vector<double>> g;
  std::map<int, std::vector<std::vector<const gp_event *>>> active_at;   // k -> per scenario
  EventTimeline(const gp_ctx *C, int maxsteps, const std::vector<int> &want) {
    const int L = C->ldM, M = C->M;
    std::vector<std::map<int, std::vector<const gp_event *>>> on(M), off(M);
    for (auto &e : C->events) {
      on[e.scenario][e.step_on].push_back(&e);
      off[e.scenario][e.step_off].push_back(&e);
    }
    std::vector<std::vector<const gp_event *>> act(M);
    std::vector<int> fb((size_t)MAXFAULT * L, -1);
    std::vector<double> fgv((size_t)MAXFAULT * L, 0.0);
    bus.push_back(fb);
    g.push_back(fgv);
    ver.assign(maxsteps + 2, 0);
    for (int k = 1; k <= maxsteps + 1; k++) {
      ver[k] = (int)bus.size() - 1;
      if (std::find(want.begin(), want.end(), k) != want.end()) active_at[k] = act;
      bool changed = false;
      for (int m = 0; m < M; m++) {
        bool ch = false;
        auto ion = on[m].find(k);
        if (ion != on[m].end())
          for (const gp_event *e : ion->second) {
            act[m].push_back(e);
            ch = true;
          }
        auto ioff = off[m].find(k);
        if (ioff != off[m].end())
          for (const gp_event *e : ioff->second) {
            auto it = std::find(act[m].begin(), act[m].end(), e);
            if (it != act[m].end()) act[m].erase(it);
            ch = true;
          }
        if (!ch) continue;
        changed = true;
        for (int q = 0; q < MAXFAULT; q++) {
          const bool f_on = q < (int)act[m].size();
          fb[(size_t)q * L + m] = f_on ? act[m][q]->bus : -1;
          fgv[(size_t)q * L + m] = f_on ? 1.0 / act[m][q]->r_fault : 0.0;
        }
      }
      if (changed) {
        bus.push_back(fb);
        g.push_back(fgv);
      }
    }
  }
};

// LU plan of J^T for the adjoint's transposed solves: the CSC pattern of J^T
// (the CSR pattern of J) whose values the factorisation reads from the
// Jacobian value array in place (a_src), pivots chosen on scenario 0's matrix
// at the current state.
void analyse_transposed(gp_ctx *C, double dt) {
  analyse(C, GP_METHOD_MONO_FP64);
  C->luJ.shape(sizeof(double));
  if (C->luJT.ready && C->luJT.dt == dt && C->luJT.trip_gen == C->trip_gen) return;
  const int n = C->n, nnz = C->jcp[n], L = C->ldM;
  std::vector<uint8_t> sel(L, 0);
  sel[0] = 1;
  set_mask(C, C->d_sel, sel);
  launch_jac_mono(C->T, C->d_p.p, C->d_z.p, dt, faults(C), C->d_jv.p, C->d_sel.p, C->stream);
  HIPCHK(hipGetLastError());
  std::vector<double> jv(nnz);
  HIPCHK(hipMemcpy2DAsync(jv.data(), sizeof(double), C->d_jv.p, (size_t)L * sizeof(double), sizeof(double), nnz,
                          hipMemcpyDeviceToHost, C->stream));
  HIPCHK(hipStreamSynchronize(C->stream));
  std::vector<int> tcp(n + 1, 0), tri(nnz), tsrc(nnz);
  for (int q = 0; q < nnz; q++) tcp[C->jri[q] + 1]++;
  for (int i = 0; i < n; i++) tcp[i + 1] += tcp[i];
  std::vector<int> next(tcp.begin(), tcp.end() - 1);
  for (int c = 0; c < n; c++)
    for (int q = C->jcp[c]; q < C->jcp[c + 1]; q++) {
      const int t = next[C->jri[q]]++;
      tri[t] = c;   // rows of each J^T column ascend: c ascends
      tsrc[t] = q;
    }
  std::vector<double> tv(nnz);
  for (int t = 0; t < nnz; t++) tv[t] = jv[tsrc[t]];
  const std::vector<std::vector<double>> jx = trip_matrices(C, jv, false);
  std::vector<std::vector<double>> tx(jx.size(), std::vector<double>(nnz));
  for (size_t k = 0; k < jx.size(); k++)
    for (int t = 0; t < nnz; t++) tx[k][t] = jx[k][tsrc[t]];
  std::string e = splu_analyse(n, tcp.data(), tri.data(), tv.data(), C->luJT.A, 1e-3, ptrs(tx));
  if (!e.empty()) throw GpError{GP_ERR_SINGULAR_PIVOT, "transposed Jacobian analysis: " + e};
  C->luJT.tune(tsrc, jv, nullptr, sizeof(double), 1.0, C->stream, C->opt.dense_block);
  if (C->luJT.dev.gx) {
    C->d_gxJT.ensure((size_t)n * L);
    C->luJT.dev.gx_ws = C->d_gxJT.p;
  }
  C->luJT.ready = true;
  C->luJT.dt = dt;
  C->luJT.trip_gen = C->trip_gen;
  C->d_luJT.alloc(C->luJT.lu_elems() * C->M);
  C->d_svJT.alloc((size_t)C->luJT.dev.ns * C->M);
  HIPCHK(hipMemset(C->d_svJT.p, 0, C->d_svJT.n * sizeof(double)));   // padding slots stay zero
  C->d_dnJT.alloc((size_t)std::max(1, C->luJT.A.nd * C->luJT.A.nd) * C->M);
}

// Parameter index -> pullback entry (kind, device row) of k_adj_param.
bool adj_entry(const gp_ctx *C, int q, int2 &e) {
  struct Slot {
    int type, pcol, off, kind;
  };
  static const Slot slots[] = {{GP_GENROU, 2, 6, GP_ADJ_GENROU_H}, {GP_GENROU, 2, 7, GP_ADJ_GENROU_D},
                               {GP_SEXS, 1, 2, GP_ADJ_SEXS_K},     {GP_ESDC1A, 1, 0, GP_ADJ_ESDC1A_KA},
                               {GP_TGOV1, 2, 0, GP_ADJ_TGOV1_R},   {GP_IEESGO, 2, 6, GP_ADJ_IEESGO_K1},
                               {GP_IEEEST, 2, 12, GP_ADJ_IEEEST_KS}};
  static const int ncols[GP_NDEVTYPES] = {GP_COLS_GENROU, GP_COLS_IEESGO, GP_COLS_TGOV1, GP_COLS_SEXS,
                                          GP_COLS_ESDC1A, GP_COLS_IEEEST, GP_COLS_ZIPLOAD, GP_COLS_STATICGEN};
  for (const Slot &s : slots) {
    const auto &d = C->dev[s.type];
    const int nc = ncols[s.type], rows = (int)d.size() / nc;
    for (int r = 0; r < rows; r++)
      if (d[(size_t)r * nc + s.pcol] + s.off == q) {
        e = make_int2(s.kind, r);
        return true;
      }
  }
  return false;
}

// Restores the refactorisation policy on scope exit.
struct PolicyGuard {
  gp_ctx *C;
  int saved;
  PolicyGuard(gp_ctx *c, int pol) : C(c), saved(c->opt.refactor_policy) { C->opt.refactor_policy = pol; }
  ~PolicyGuard() { C->opt.refactor_policy = saved; }
};

}  // namespace

// The parity hooks' fault tables: scenario m's first MAXFAULT events where
// fault_on[m] != 0 (all scenarios' with all_on).
static void eval_fault_tables(const gp_ctx *C, const uint8_t *fault_on, bool all_on, std::vector<int> &fbus,
                              std::vector<double> &fg) {
  const int L = C->ldM;
  fbus.assign((size_t)MAXFAULT * L, -1);
  fg.assign((size_t)MAXFAULT * L, 0.0);
  for (int m = 0; m < C->M; m++) {
    if (!all_on && !(fault_on && fault_on[m])) continue;
    int q = 0;
    for (auto &e : C->events)
      if (e.scenario == m && q < MAXFAULT) {
        fbus[(size_t)q * L + m] = e.bus;
        fg[(size_t)q * L + m] = 1.0 / e.r_fault;
        q++;
      }
  }
}

// Every scenario's events in slot form on the device (after any change of
// the event tables), for gp_eval_residual_noalloc.
static void upload_event_slots(gp_ctx *C) {
  std::vector<int> fbus;
  std::vector<double> fg;
  eval_fault_tables(C, nullptr, true, fbus, fg);
  C->d_evbus.upload(fbus);
  C->d_evg.upload(fg);
}

// ================================================================= C ABI
extern "C" {

const char *gp_last_error(void) { return g_err.c_str(); }
const char *gp_version(void) { return "gradpower_hip 0.1 (gfx950)"; }

int gp_create(const gp_system_desc *d, int32_t M, const gp_options *opt, gp_ctx **out) {
  if (!d || !out || M <= 0) return fail(GP_ERR_BADARG, "gp_create: null argument or M <= 0");
  if (d->n <= 0 || d->nbus <= 0 || d->net_ptr + 2 * d->nbus != d->n)
    return fail(GP_ERR_BADARG, "gp_create: inconsistent sizes (n != net_ptr + 2*nbus)");
  std::unique_ptr<gp_ctx> C(new gp_ctx());
  try {
    HIPCHK(hipGetDevice(&C->device));
    HIPCHK(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&C->side, hipStreamNonBlocking));
    for (hipEvent_t *e : {&C->ev_fork, &C->ev_side, &C->ev_fork2, &C->ev_join, &C->ev_round})
      HIPCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    for (auto &ln : C->lanes) {
      HIPCHK(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&ln.ev_snap, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&ln.ev_done, hipEventDisableTiming));
    }
    C->n = d->n; C->nbus = d->nbus; C->net_ptr = d->net_ptr; C->npar = d->n_par; C->M = M;
    C->ldM = (M + WAVE - 1) / WAVE * WAVE;
    if (opt) C->opt = *opt;
    else { C->opt.record_all = 1; C->opt.refactor_policy = GP_REFACTOR_EVERY_ITER; }
    C->luJ.force_gx = C->luS.force_gx = C->luJT.force_gx = C->opt.global_x != 0;
    C->luJ.part_k = C->luS.part_k = C->luJT.part_k = C->opt.solve_levels > 0 ? C->opt.solve_levels : 4;
    C->luJ.part_tune = C->luS.part_tune = C->luJT.part_tune = C->opt.solve_levels <= 0;
    // profile = k: every k-th solve launch and every k-th factorisation launch timed
    C->prof_solve.every = C->prof_factor.every = std::max(1, C->opt.profile);
    C->is_diff.assign(d->is_diff, d->is_diff + d->n);
    C->yrp.assign(d->y_rowptr, d->y_rowptr + 2 * d->nbus + 1);
    C->yci.assign(d->y_colind, d->y_colind + d->y_nnz);
    C->yv.assign(d->y_val, d->y_val + d->y_nnz);
    static const int ncols[GP_NDEVTYPES] = {GP_COLS_GENROU, GP_COLS_IEESGO, GP_COLS_TGOV1, GP_COLS_SEXS,
                                            GP_COLS_ESDC1A, GP_COLS_IEEEST, GP_COLS_ZIPLOAD, GP_COLS_STATICGEN};
    for (int t = 0; t < GP_NDEVTYPES; t++) {
      if (d->n_dev[t] < 0) return fail(GP_ERR_BADARG, "negative device count");
      if (d->n_dev[t] && !d->dev[t]) return fail(GP_ERR_BADARG, "missing device table");
      C->dev[t].assign(d->dev[t], d->dev[t] + (size_t)d->n_dev[t] * ncols[t]);
    }
    C->u0.assign(d->genrou_u0, d->genrou_u0 + 2 * (size_t)d->n_dev[GP_GENROU]);
    C->ncl = d->n_clusters;
    C->cl.assign(d->clusters, d->clusters + (size_t)d->n_clusters * GP_CLUSTER_COLS);
    // bounds checks of every index the kernels will dereference
    auto zok = [&](int i) { return i >= -1 && i < C->n; };
    for (int t = 0; t < 6; t++)
      for (int v : C->dev[t])
        if (!zok(v) && !(v >= 0 && v < std::max(C->npar, C->n)))
          return fail(GP_ERR_BADARG, "device table index out of range");
    for (int c = 0; c < C->ncl; c++) {
      const int *r = C->cl.data() + c * GP_CLUSTER_COLS;
      if (r[GP_CL_WSTART] < 0 || r[GP_CL_WSTART] + r[GP_CL_WSIZE] > C->net_ptr || r[GP_CL_BUS] < 0 ||
          r[GP_CL_BUS] >= C->nbus)
        return fail(GP_ERR_BADARG, "cluster table out of range");
    }
    for (int q = 0; q < (int)C->yci.size(); q++)
      if (C->yci[q] < 0 || C->yci[q] >= 2 * C->nbus) return fail(GP_ERR_BADARG, "Y column out of range");
    build_structures(C.get());
    upload_tables(C.get());
    alloc_state(C.get());
    if (C->opt.record_all || !C->opt.record_idx) {
      C->nrec = C->n;
      C->rec_idx.clear();
    } else {
      C->nrec = C->opt.n_record;
      C->rec_idx.assign(C->opt.record_idx, C->opt.record_idx + C->nrec);
      for (int r : C->rec_idx)
        if (r < 0 || r >= C->n) return fail(GP_ERR_BADARG, "record index out of range");
      up(C->d_rec_idx, C->rec_idx);
      std::vector<int> rp(C->n + 1, 0), rc(C->nrec);
      for (int r : C->rec_idx) rp[r + 1]++;
      for (int i = 0; i < C->n; i++) rp[i + 1] += rp[i];
      std::vector<int> nx(rp.begin(), rp.end() - 1);
      for (int j = 0; j < C->nrec; j++) rc[nx[C->rec_idx[j]]++] = j;
      up(C->d_rec_p, rp);
      up(C->d_rec_ch, rc);
    }
    C->nsteps.assign(M, 0);
    upload_event_slots(C.get());
    if (scratch_setup(C->device, &C->scratch) != 0) C->scratch = nullptr;   // the noalloc entry point refuses
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  } catch (const std::exception &e) {
    return fail(GP_ERR_OOM, e.what());
  }
  *out = C.release();
  return GP_OK;
}

void gp_destroy(gp_ctx *C) {
  if (!C) return;
  for (auto &ln : C->lanes)
    if (ln.stream) (void)hipStreamSynchronize(ln.stream);
  if (C->stream) (void)hipStreamSynchronize(C->stream);
  hipStream_t s = C->stream;
  std::vector<hipStream_t> ls;
  std::vector<hipEvent_t> le;
  for (auto &ln : C->lanes) {
    ls.push_back(ln.stream);
    le.push_back(ln.ev_snap);
    le.push_back(ln.ev_done);
  }
  if (C->side) (void)hipStreamSynchronize(C->side);
  for (hipEvent_t e : {C->ev_fork, C->ev_side, C->ev_fork2, C->ev_join, C->ev_round})
    if (e) (void)hipEventDestroy(e);
  if (C->side) (void)hipStreamDestroy(C->side);
  for (hipEvent_t e : C->prof_solve.ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : C->prof_factor.ev) (void)hipEventDestroy(e);
  delete C;
  for (hipEvent_t e : le)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t x : ls)
    if (x) (void)hipStreamDestroy(x);
  if (s) (void)hipStreamDestroy(s);
}

int gp_set_state(gp_ctx *C, const double *z0, int32_t per_scenario, const double *p) {
  if (!C || !z0 || !p) return fail(GP_ERR_BADARG, "gp_set_state: null argument");
  try {
    const int L = C->ldM;
    std::vector<double> zt((size_t)C->n * L, 0.0);
    for (int m = 0; m < C->M; m++) {
      const double *src = per_scenario ? z0 + (size_t)m * C->n : z0;
      for (int i = 0; i < C->n; i++) zt[(size_t)i * L + m] = src[i];
    }
    HIPCHK(hipMemcpy(C->d_z.p, zt.data(), zt.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(C->d_z0.p, zt.data(), zt.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(C->d_p.p, p, (size_t)C->npar * sizeof(double), hipMemcpyHostToDevice));
    C->state_set = true;
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

// The event list the time loops read: the faults, then one record per line
// trip (slot code -2 - t) and per ZIP load disconnect (kDropCode0 - load).
static std::vector<gp_event> event_list(const std::vector<gp_event> &faults, const std::vector<gp_line_trip> &trips,
                                        const std::vector<gp_device_drop> &drops) {
  std::vector<gp_event> ev = faults;
  for (size_t i = 0; i < trips.size(); i++) {
    gp_event e{};
    e.scenario = trips[i].scenario;
    e.bus = -2 - (int)i;
    e.r_fault = 1.0;
    e.step_on = trips[i].step_trip;
    e.step_off = trips[i].step_reclose > 0 ? trips[i].step_reclose : std::numeric_limits<int32_t>::max();
    ev.push_back(e);
  }
  for (const gp_device_drop &d : drops) {
    gp_event e{};
    e.scenario = d.scenario;
    e.bus = d.type == GP_ZIPLOAD ? kDropCode0 - d.row : kExcCode0 - (2 * d.row + (d.type == GP_ESDC1A ? 1 : 0));
    e.r_fault = 1.0;
    e.step_on = d.step;
    e.step_off = std::numeric_limits<int32_t>::max();
    ev.push_back(e);
  }
  return ev;
}

// The time loops apply a step's activations, then its deactivations, so an
// event is in force after step k when step_on <= k and it was not switched
// off at or after its activation by step k.  The most events in force at
// once in any scenario (checked when the events are set, so that a sweep
// never meets more than its MAXFAULT slots).
static int max_in_force(const std::vector<gp_event> &ev, int M) {
  std::vector<std::vector<const gp_event *>> by(M);
  for (const gp_event &e : ev) by[e.scenario].push_back(&e);
  int best = 0;
  for (const auto &v : by)
    for (const gp_event *a : v) {
      const int k = a->step_on;   // the count only rises at an activation
      int cnt = 0;
      for (const gp_event *e : v) cnt += (e->step_on <= k && !(e->step_on <= e->step_off && e->step_off <= k)) ? 1 : 0;
      best = std::max(best, cnt);
    }
  return best;
}

static bool events_fit(const std::vector<gp_event> &ev, int M) { return max_in_force(ev, M) <= MAXFAULT; }

// Shared-factor method, set-up of one call (_factor_shared_P_gpu! and
// _precompute_woodbury_multi_gpu!, ext:2149-2173, 2441-2495): P = S of
// scenario 0 at its current state, step dt, no event in force, factorised
// once in FP64; W = P^{-1} U for the two unit columns of every bus that
// one of the events faults, by one solve with the columns as its
// scenarios; U^T W on the host for the per-scenario capacitance matrices.
// Every scenario's table row starts with no fault in force.
static void prepare_shared(gp_ctx *C, double dt, int in_force = -1) {
  if (!C->trips.empty() || !C->drops.empty())
    throw GpError{GP_ERR_UNSUPPORTED, "shared-factor method: line trips and load drops change P itself"};
  if (C->opt.lock_step) throw GpError{GP_ERR_UNSUPPORTED, "shared-factor method: not in lock step"};
  const SpluDev &D = C->luS64;
  if (D.gx) throw GpError{GP_ERR_UNSUPPORTED, "shared-factor method: the reduced network's solution exceeds LDS"};
  gp_ctx::Shared &W = C->sh;
  const int L = C->ldM, nr = C->nred;
  const std::vector<int> fb0((size_t)MAXFAULT * L, -1);
  const std::vector<double> fg0((size_t)MAXFAULT * L, 0.0);
  HIPCHK(hipMemcpyAsync(C->d_fbus.p, fb0.data(), fb0.size() * sizeof(int), hipMemcpyHostToDevice, C->stream));
  HIPCHK(hipMemcpyAsync(C->d_fg.p, fg0.data(), fg0.size() * sizeof(double), hipMemcpyHostToDevice, C->stream));
  std::vector<uint8_t> sel(L, 0);
  sel[0] = 1;
  set_mask(C, C->d_sel, sel);
  launch_copy_active(C->n, L, C->d_z.p, C->d_zold.p, C->d_sel.p, C->stream);
  launch_cluster_residual(C->T, C->d_p.p, C->d_z.p, C->d_zold.p, dt, faults(C),
                          C->d_f.p, C->d_inj.p, C->d_sel.p, C->stream);
  launch_bus_residual(C->T, C->d_p.p, C->d_z.p, C->d_inj.p, faults(C), C->d_f.p, C->d_sel.p, C->stream);
  launch_cluster_schur<double>(C->T, C->d_p.p, C->d_z.p, C->d_f.p, dt, faults(C),
                               C->d_X64.p, C->d_Dg.p, C->d_sel.p, C->stream);
  launch_bus_schur(C->T, C->d_p.p, C->d_z.p, C->d_f.p, C->d_Dg.p, faults(C), C->d_svar.p, C->d_rhs.p, C->d_sel.p,
                   C->stream);
  W.lu.ensure(std::max<size_t>(1, C->luS.lu_elems()));
  if (W.sv.n != (size_t)D.ns) {
    W.sv.alloc(D.ns);
    HIPCHK(hipMemset(W.sv.p, 0, W.sv.n * sizeof(double)));   // padding slots stay zero
  }
  W.dn.ensure((size_t)std::max(1, D.nd * D.nd));
  HIPCHK(hipMemsetAsync(C->d_status.p, 0, sizeof(int), C->stream));
  launch_splu_factor_impl<double>(D, C->d_svar.p, C->d_s_cst.p, L, W.lu.p, W.sv.p, W.dn.p, C->d_sel.p, nullptr, 1,
                                  C->d_status.p, C->stream);
  HIPCHK(hipGetLastError());
  int st0 = 0;
  HIPCHK(hipMemcpyAsync(&st0, C->d_status.p, sizeof(int), hipMemcpyDeviceToHost, C->stream));
  HIPCHK(hipStreamSynchronize(C->stream));
  if (st0) throw GpError{GP_ERR_SINGULAR_PIVOT, "shared factorisation: zero or non-finite pivot in P"};
  // Woodbury columns of the faulted buses
  std::vector<int> buses;
  for (const gp_event &e : C->events)
    if (e.bus >= 0) buses.push_back(e.bus);
  std::sort(buses.begin(), buses.end());
  buses.erase(std::unique(buses.begin(), buses.end()), buses.end());
  const int nb = (int)buses.size();
  W.ncol = 2 * nb;
  W.kw = 2 * std::max(1, std::min(MAXFAULT, in_force >= 0 ? in_force : max_in_force(C->events, C->M)));
  W.col_of_bus.clear();
  W.G.clear();
  W.upos.assign(W.ncol, 0);
  if (nb > 0) {
    const int ldw = (W.ncol + 63) / 64 * 64;
    std::vector<int> urow(W.ncol);
    for (int b = 0; b < nb; b++) {
      W.col_of_bus[buses[b]] = 2 * b;
      for (int t = 0; t < 2; t++) urow[2 * b + t] = C->red_of_z[C->net_ptr + 2 * buses[b] + t];
    }
    std::vector<double> E((size_t)nr * ldw, 0.0);
    for (int c = 0; c < W.ncol; c++) E[(size_t)urow[c] * ldw + c] = 1.0;
    std::vector<uint8_t> on(ldw, 0);
    std::fill(on.begin(), on.begin() + W.ncol, 1);
    DBuf<double> dE, dWn;
    DBuf<uint8_t> dOn;
    dE.upload(E);
    dOn.upload(on);
    dWn.alloc((size_t)nr * ldw);
    launch_splu_solve<double>(D, W.sv.p, W.dn.p, dE.p, ldw, 1.0, nullptr, nullptr, dWn.p, dOn.p, nullptr, W.ncol,
                              C->stream, true);
    HIPCHK(hipGetLastError());
    std::vector<double> Wn((size_t)nr * ldw);
    HIPCHK(hipMemcpyAsync(Wn.data(), dWn.p, Wn.size() * sizeof(double), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    const SpluAnalysis &A = C->luS.A;
    std::vector<double> Wi((size_t)W.ncol * nr);
    for (int c = 0; c < W.ncol; c++)
      for (int k = 0; k < nr; k++) Wi[(size_t)c * nr + k] = Wn[(size_t)A.q[k] * ldw + c];
    W.W.upload(Wi);
    W.G.assign((size_t)W.ncol * W.ncol, 0.0);
    for (int i = 0; i < W.ncol; i++) {
      W.upos[i] = A.qinv[urow[i]];
      for (int j = 0; j < W.ncol; j++) W.G[(size_t)i * W.ncol + j] = Wn[(size_t)urow[i] * ldw + j];
    }
  } else {
    W.W.ensure(1);
  }
  const int kw = W.kw;
  W.hnc.assign(L, 0);
  W.hcol.assign((size_t)L * kw, 0);
  W.hpos.assign((size_t)L * kw, 0);
  W.hh.assign((size_t)L * kw * kw, 0.0);
  W.nc.ensure(L);
  W.col.ensure((size_t)L * kw);
  W.pos.ensure((size_t)L * kw);
  W.h.ensure((size_t)L * kw * kw);
  shared_upload(C);
  HIPCHK(hipStreamSynchronize(C->stream));
}

static void rebuild_events(gp_ctx *C) {
  C->events = event_list(C->faults_user, C->trips, C->drops);
  C->nexc = 0;
  for (const gp_device_drop &d : C->drops) C->nexc += d.type != GP_ZIPLOAD;
}

int gp_set_events(gp_ctx *C, const gp_event *ev, int32_t n_ev, const int32_t *nsteps) {
  if (!C || (n_ev > 0 && !ev) || !nsteps) return fail(GP_ERR_BADARG, "gp_set_events: null argument");
  for (int i = 0; i < n_ev; i++) {
    const gp_event &e = ev[i];
    if (e.scenario < 0 || e.scenario >= C->M || e.bus < 0 || e.bus >= C->nbus)
      return fail(GP_ERR_BADARG, "gp_set_events: scenario or bus out of range");
    if (!(e.r_fault > 0.0)) return fail(GP_ERR_BADARG, "gp_set_events: r_fault must be > 0");
    if (e.step_on < 1 || e.step_off < 1)
      return fail(GP_ERR_BADARG, "gp_set_events: event steps must be >= 1 (an event at t < dt/2 never fires)");
  }
  for (int m = 0; m < C->M; m++)
    if (nsteps[m] < 0) return fail(GP_ERR_BADARG, "gp_set_events: negative nsteps");
  if (!events_fit(std::vector<gp_event>(ev, ev + n_ev), C->M))
    return fail(GP_ERR_UNSUPPORTED, "gp_set_events: more than " + std::to_string(MAXFAULT) +
                                        " faults in force at once in a scenario");
  C->faults_user.assign(ev, ev + n_ev);
  C->events = C->faults_user;
  if (!C->trips.empty()) C->trip_gen++;   // the LU plans' pivots were chosen with the trips
  C->trips.clear();
  C->drops.clear();
  C->tr_bus.clear();
  C->tr_y.clear();
  C->nsteps.assign(nsteps, nsteps + C->M);
  try {
    upload_trip_tables(C);
    upload_event_slots(C);
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  }
  return GP_OK;
}


int gp_set_device_drops(gp_ctx *C, const gp_device_drop *d, int32_t n) {
  if (!C || (n > 0 && !d) || n < 0) return fail(GP_ERR_BADARG, "gp_set_device_drops: bad argument");
  for (int i = 0; i < n; i++) {
    const int t = d[i].type;
    if (t != GP_ZIPLOAD && t != GP_SEXS && t != GP_ESDC1A)
      return fail(GP_ERR_UNSUPPORTED, "gp_set_device_drops: only ZIP loads and SEXS / ESDC1A exciters can go offline "
                                      "(a device with algebraic states would leave them undetermined)");
    const int rows = (int)(C->dev[t].size() / (t == GP_ZIPLOAD ? GP_COLS_ZIPLOAD
                                                               : (t == GP_SEXS ? GP_COLS_SEXS : GP_COLS_ESDC1A)));
    if (d[i].scenario < 0 || d[i].scenario >= C->M || d[i].row < 0 || d[i].row >= rows)
      return fail(GP_ERR_BADARG, "gp_set_device_drops: scenario or device row out of range");
    if (d[i].step < 1) return fail(GP_ERR_BADARG, "gp_set_device_drops: step must be >= 1");
  }
  if (!events_fit(event_list(C->faults_user, C->trips, std::vector<gp_device_drop>(d, d + n)), C->M))
    return fail(GP_ERR_UNSUPPORTED, "gp_set_device_drops: more than " + std::to_string(MAXFAULT) +
                                        " faults, tripped lines and offline devices in force at once in a scenario");
  C->drops.assign(d, d + n);
  rebuild_events(C);
  try {
    upload_event_slots(C);
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

int gp_set_load_drops(gp_ctx *C, const gp_load_drop *d, int32_t n) {
  if (!C || (n > 0 && !d) || n < 0) return fail(GP_ERR_BADARG, "gp_set_load_drops: bad argument");
  const int nzip = (int)(C->dev[GP_ZIPLOAD].size() / GP_COLS_ZIPLOAD);
  std::vector<gp_device_drop> dd(n);
  for (int i = 0; i < n; i++) {
    if (d[i].scenario < 0 || d[i].scenario >= C->M || d[i].load < 0 || d[i].load >= nzip)
      return fail(GP_ERR_BADARG, "gp_set_load_drops: scenario or load out of range");
    if (d[i].step < 1) return fail(GP_ERR_BADARG, "gp_set_load_drops: step must be >= 1");
    dd[i] = gp_device_drop{d[i].scenario, GP_ZIPLOAD, d[i].load, d[i].step};
  }
  if (!events_fit(event_list(C->faults_user, C->trips, dd), C->M))
    return fail(GP_ERR_UNSUPPORTED, "gp_set_load_drops: more than " + std::to_string(MAXFAULT) +
                                        " faults, tripped lines and dropped loads in force at once in a scenario");
  C->drops = dd;
  rebuild_events(C);
  try {
    upload_event_slots(C);
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

int gp_set_line_trips(gp_ctx *C, const gp_line_trip *tr, int32_t n) {
  if (!C || (n > 0 && !tr) || n < 0) return fail(GP_ERR_BADARG, "gp_set_line_trips: bad argument");
  const int npt = C->net_ptr;
  // Y entry (row, column) of the realified Y, or none
  auto yval = [&](int r, int c, double &v) {
    for (int q = C->yrp[r]; q < C->yrp[r + 1]; q++)
      if (C->yci[q] == c) { v = C->yv[q]; return true; }
    return false;
  };
  for (int i = 0; i < n; i++) {
    const gp_line_trip &t = tr[i];
    if (t.scenario < 0 || t.scenario >= C->M || t.from_bus < 0 || t.from_bus >= C->nbus || t.to_bus < 0 ||
        t.to_bus >= C->nbus || t.from_bus == t.to_bus)
      return fail(GP_ERR_BADARG, "gp_set_line_trips: scenario or bus out of range");
    if (t.step_trip < 1 || t.step_reclose < 0)
      return fail(GP_ERR_BADARG, "gp_set_line_trips: step_trip must be >= 1 and step_reclose >= 0");
    for (int k = 0; k < 8; k++)
      if (!std::isfinite(t.y[k])) return fail(GP_ERR_BADARG, "gp_set_line_trips: non-finite admittance");
    for (int d = 0; d < 2; d++)
      for (int rr = 0; rr < 2; rr++)
        for (int cc = 0; cc < 2; cc++) {
          const int rb = d ? t.to_bus : t.from_bus, cb = d ? t.from_bus : t.to_bus;
          double v;
          if (!yval(2 * rb + rr, 2 * cb + cc, v) ||
              find_pos(C->jcp, C->jri, npt + 2 * rb + rr, npt + 2 * cb + cc) < 0)
            return fail(GP_ERR_BADARG, "gp_set_line_trips: no Y entry between buses " + std::to_string(t.from_bus) +
                                           " and " + std::to_string(t.to_bus));
        }
  }
  if (!events_fit(event_list(C->faults_user, std::vector<gp_line_trip>(tr, tr + n), C->drops), C->M))
    return fail(GP_ERR_UNSUPPORTED, "gp_set_line_trips: more than " + std::to_string(MAXFAULT) +
                                        " faults, tripped lines and dropped loads in force at once in a scenario");
  try {
    // the off-diagonal entries these trips change become per-scenario values:
    // J positions, and S slots taken out of the constants (the reduced
    // system's LU plan is then made again)
    bool s_changed = false;
    std::vector<std::vector<int>> add_i(C->nbus);
    std::vector<std::vector<double>> add_b(C->nbus);
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 2; d++) {
        const int rb = d ? tr[i].to_bus : tr[i].from_bus, cb = d ? tr[i].from_bus : tr[i].to_bus;
        for (int rr = 0; rr < 2; rr++)
          for (int cc = 0; cc < 2; cc++) {
            const int zr = npt + 2 * rb + rr, zc = npt + 2 * cb + cc;
            if (C->tv_of.count({zr, zc})) continue;
            double v = 0.0;
            yval(2 * rb + rr, 2 * cb + cc, v);
            const int jp = find_pos(C->jcp, C->jri, zr, zc);
            const int ps = find_pos(C->scp, C->sri, C->red_of_z[zr], C->red_of_z[zc]);
            if (ps < 0) throw GpError{GP_ERR_HIP, "internal: Y entry missing from the reduced pattern"};
            if (C->s_src[ps] < 0) {
              C->s_src[ps] = C->n_svar++;
              s_changed = true;
            }
            C->tv_of[{zr, zc}] = 1;
            add_i[rb].insert(add_i[rb].end(), {cb, 2 * rr + cc, jp, C->s_src[ps]});
            add_b[rb].push_back(-v);
          }
      }
    // per-bus lists: the entries kept so far, then the new ones
    std::vector<int> tv_p(C->nbus + 1, 0), tv_i;
    std::vector<double> tv_base;
    for (int b = 0; b < C->nbus; b++) {
      for (int q = C->tv_p[b]; q < C->tv_p[b + 1]; q++) {
        tv_i.insert(tv_i.end(), C->tv_i.begin() + 4 * q, C->tv_i.begin() + 4 * q + 4);
        tv_base.push_back(C->tv_base[q]);
      }
      tv_i.insert(tv_i.end(), add_i[b].begin(), add_i[b].end());
      tv_base.insert(tv_base.end(), add_b[b].begin(), add_b[b].end());
      tv_p[b + 1] = (int)tv_base.size();
    }
    C->tv_p = tv_p;
    C->tv_i = tv_i;
    C->tv_base = tv_base;
    if (s_changed) C->luS.ready = false;
    if (n > 0 || !C->trips.empty()) C->trip_gen++;   // pivots are chosen with the tripped lines
    C->trips.assign(tr, tr + n);
    C->tr_bus.clear();
    C->tr_y.clear();
    for (int i = 0; i < n; i++) {
      C->tr_bus.push_back(tr[i].from_bus);
      C->tr_bus.push_back(tr[i].to_bus);
      C->tr_y.insert(C->tr_y.end(), tr[i].y, tr[i].y + 8);
    }
    rebuild_events(C);
    upload_trip_tables(C);
    upload_event_slots(C);
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

int gp_integrate(gp_ctx *C, int32_t method, double dt, double tol, int32_t itermax, gp_stats *st) {
  if (!C) return fail(GP_ERR_BADARG, "null context");
  if (method < 0 || method > 3) return fail(GP_ERR_BADARG, "unknown method");
  if (!(dt > 0.0) || !(tol > 0.0) || itermax <= 0) return fail(GP_ERR_BADARG, "dt, tol and itermax must be > 0");
  if (!C->state_set) return fail(GP_ERR_BADARG, "gp_set_state was not called");
  gp_stats S{};
  try {
    const int L = C->ldM, M = C->M;
    const bool shared = method == GP_METHOD_SCHUR_SHARED;
    int maxsteps = 0;
    for (int m = 0; m < M; m++) maxsteps = std::max(maxsteps, C->nsteps[m]);
    if (shared && (C->opt.lock_step || !C->trips.empty() || !C->drops.empty()))
      throw GpError{GP_ERR_UNSUPPORTED, "shared-factor method: not in lock step, nor with line trips or load drops"};
    Stepper sp(C, method, dt, tol, itermax);   // no fault active at t = 0
    prepare_method(C, method);
    if (shared) prepare_shared(C, dt);
    // trajectory storage [step][channel][ldM]
    C->d_hist.alloc((size_t)(maxsteps + 1) * C->nrec * L);
    C->hist_steps = maxsteps + 1;
    C->traj_len.assign(M, 0);
    for (int m = 0; m < M; m++) C->traj_len[m] = C->nsteps[m] + 1;
    C->scen_status.assign(M, GP_OK);
    C->scen_steps.assign(C->nsteps.begin(), C->nsteps.begin() + M);
    DevTimer tm(C->stream);
    record_initial(C);
    if (C->opt.lock_step) {
      for (int k = 1; k <= maxsteps; k++)
        if (!sp.step(k, C->d_hist.p + (size_t)k * C->nrec * L, nullptr)) break;
      S = sp.S;
      for (int m = 0; m < M; m++) S.n_not_converged += sp.st.everfail[m];
    } else {
      FreeSpec F;
      F.k_first.assign(M, 1);
      F.k_last.assign(C->nsteps.begin(), C->nsteps.begin() + M);
      std::vector<RepivotReq> reqs;
      F.repivot = shared ? nullptr : &reqs;   // (the shared method factorises nothing in the loop)
      S = integrate_free(C, method, dt, tol, itermax, F);
      if (!reqs.empty()) repivot_runs(C, method, dt, tol, itermax, reqs, S);
    }
    land_batch(C, true);   // no factorisation outlives the call
    S.elapsed_ms = S.solve_ms = tm.ms();
    if (C->opt.profile) {
      for (auto &ln : C->lanes) HIPCHK(hipStreamSynchronize(ln.stream));
      S.solve_kernel_ms = prof_collect(C->prof_solve, S.solve_kernel_launches, S.solve_kernel_scenarios);
      S.factor_kernel_ms = prof_collect(C->prof_factor, S.factor_kernel_launches, S.factor_kernel_scenarios);
      const bool mono = method == GP_METHOD_MONO_FP64;
      const SpluAnalysis &A = mono ? C->luJ.A : C->luS.A;
      const double bf = method == GP_METHOD_SCHUR_FP32 ? 4.0 : 8.0;
      S.solve_bytes_per_scenario = bf * A.nnz_lu + (mono ? 24.0 : 32.0) * A.n;
      S.factor_bytes_per_scenario = 8.0 * A.nnz_a + bf * A.nnz_lu;
    }
    if (st) *st = S;
    const int rc = finish_status(C, S);
    if (rc != GP_OK) return rc;
  } catch (const GpError &e) {
    try {
      land_batch(C, true);
    } catch (const GpError &) {
      for (auto &ln : C->lanes) ln.inflight = false;
    }
    if (st) *st = S;
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

int gp_adjoint(gp_ctx *C, int32_t method, double dt, double tol, int32_t itermax, int32_t segment,
               const gp_objective *obj, int32_t n_params, const int32_t *params, double *value, double *grad,
               double *grad_scen, double *dz0, gp_adjoint_stats *ast) {
  if (!C) return fail(GP_ERR_BADARG, "null context");
  if (method == GP_METHOD_SCHUR_SHARED)
    return fail(GP_ERR_UNSUPPORTED, "gp_adjoint: the shared-factor method has no adjoint (use a Schur method)");
  if (method < 0 || method > 2) return fail(GP_ERR_BADARG, "unknown method");
  if (!(dt > 0.0) || !(tol > 0.0) || itermax <= 0) return fail(GP_ERR_BADARG, "dt, tol and itermax must be > 0");
  if (!C->state_set) return fail(GP_ERR_BADARG, "gp_set_state was not called");
  if (segment < 0 || n_params < 0 || (n_params > 0 && (!params || !grad)))
    return fail(GP_ERR_BADARG, "gp_adjoint: bad segment, parameter list or gradient output");
  // objective rows (default: the GENROU speed deviations)
  std::vector<int> rows;
  if (obj && obj->rows) {
    if (obj->n_rows < 0) return fail(GP_ERR_BADARG, "gp_adjoint: negative objective row count");
    rows.assign(obj->rows, obj->rows + obj->n_rows);
    for (int r : rows)
      if (r < 0 || r >= C->n) return fail(GP_ERR_BADARG, "gp_adjoint: objective row out of range");
  } else {
    const auto &g = C->dev[GP_GENROU];
    for (size_t k = 0; k < g.size() / GP_COLS_GENROU; k++) rows.push_back(g[k * GP_COLS_GENROU] + 4);
  }
  const double scale = (obj && obj->scale != 0.0) ? obj->scale : 1.0;
  const double off = obj ? obj->offset : 0.0;
  std::vector<int2> ent(n_params);
  for (int j = 0; j < n_params; j++)
    if (params[j] < 0 || params[j] >= C->npar || !adj_entry(C, params[j], ent[j]))
      return fail(GP_ERR_UNSUPPORTED, "gp_adjoint: parameter " + std::to_string(params[j]) +
                                          " has no pullback (GENROU H, D; SEXS K; ESDC1A KA; TGOV1 R; "
                                          "IEESGO K1; IEEEST KS are supported)");
  const int SEG = segment > 0 ? segment : 30;
  gp_adjoint_stats A{};
  // factorisations start afresh in every Newton solve, so that a recomputed
  // segment repeats the forward sweep's iterates exactly
  PolicyGuard pg(C, GP_REFACTOR_EVERY_ITER);
  try {
    const int L = C->ldM, M = C->M, n = C->n, ne = n_params;
    const size_t nL = (size_t)n * L;
    int maxsteps = 0;
    for (int m = 0; m < M; m++) maxsteps = std::max(maxsteps, C->nsteps[m]);
    HIPCHK(hipMemcpyAsync(C->d_z.p, C->d_z0.p, nL * sizeof(double), hipMemcpyDeviceToDevice, C->stream));
    {   // no fault active at t = 0 (the analyses evaluate scenario 0 there)
      std::vector<int> fb((size_t)MAXFAULT * L, -1);
      std::vector<double> fgz((size_t)MAXFAULT * L, 0.0);
      HIPCHK(hipMemcpyAsync(C->d_fbus.p, fb.data(), fb.size() * sizeof(int), hipMemcpyHostToDevice, C->stream));
      HIPCHK(hipMemcpyAsync(C->d_fg.p, fgz.data(), fgz.size() * sizeof(double), hipMemcpyHostToDevice, C->stream));
      HIPCHK(hipStreamSynchronize(C->stream));
    }
    prepare_method(C, method);
    analyse_transposed(C, dt);
    C->d_hist.alloc((size_t)(maxsteps + 1) * C->nrec * L);
    C->hist_steps = maxsteps + 1;
    C->traj_len.assign(M, 0);
    for (int m = 0; m < M; m++) C->traj_len[m] = C->nsteps[m] + 1;
    C->scen_status.assign(M, GP_OK);
    C->scen_steps.assign(C->nsteps.begin(), C->nsteps.begin() + M);
    const int nseg = (maxsteps + SEG - 1) / SEG;
    C->d_ckpt.ensure((size_t)(nseg + 1) * nL);   // checkpoints, then the final state
    C->d_seg.ensure((size_t)std::min(SEG, std::max(1, maxsteps)) * nL);
    C->d_lam.ensure(nL);
    C->d_gacc.ensure((size_t)std::max(1, ne) * L);
    C->d_val.ensure(L);
    C->d_hz.ensure(L);
    C->d_bact.ensure(L);
    C->d_objrows.upload(rows);
    C->d_ent.upload(ent);
    // fault tables of every step, and the faults in force where each
    // segment's recomputation starts
    std::vector<int> starts;
    for (int sg = 0; sg < nseg; sg++) starts.push_back(sg * SEG + 1);
    const EventTimeline TL(C, maxsteps, starts);
    DevTimer total(C->stream);
    // ---- forward sweep (every scenario advancing on its own): the state
    // after step s*SEG (and its events) is checkpoint s; the last segment's
    // states are kept for the backward sweep
    record_initial(C);
    const int klast = (nseg - 1) * SEG;
    launch_copy_plain(C->d_z.p, C->d_ckpt.p, (long long)nL, C->stream);
    FreeSpec F;
    F.k_first.assign(M, 1);
    F.k_last.assign(C->nsteps.begin(), C->nsteps.begin() + M);
    F.seg = C->d_seg.p;
    F.seg_k0 = klast;
    F.seg_k1 = maxsteps;
    F.ckpt = C->d_ckpt.p;
    F.every = SEG;
    F.nck = nseg;
    const gp_stats fwd = integrate_free(C, method, dt, tol, itermax, F);
    land_batch(C, true);
    launch_copy_plain(C->d_z.p, C->d_ckpt.p + (size_t)nseg * nL, (long long)nL, C->stream);
    int64_t newton_total = fwd.newton_iters;
    A.scenario_steps = fwd.scenario_steps;
    A.forward_ms = total.ms();
    // scenarios in the backward sweep: those whose forward sweep finished
    std::vector<int> status(L);
    HIPCHK(hipMemcpy(status.data(), C->d_status.p, L * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemset(C->d_status.p, 0, L * sizeof(int)));
    std::vector<int> hz(L, -1);
    int nsing = 0;
    for (int m = 0; m < M; m++) {
      if (status[m] != 0) {
        nsing++;
        C->scen_status[m] = GP_ERR_SINGULAR_PIVOT;
      }
      if (C->scen_status[m] == GP_OK) hz[m] = C->nsteps[m];
      else A.n_excluded++;
    }
    HIPCHK(hipMemcpy(C->d_hz.p, hz.data(), L * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(C->d_lam.p, 0, nL * sizeof(double), C->stream));
    HIPCHK(hipMemsetAsync(C->d_gacc.p, 0, C->d_gacc.n * sizeof(double), C->stream));
    HIPCHK(hipMemsetAsync(C->d_val.p, 0, L * sizeof(double), C->stream));
    DevTimer back(C->stream);
    const SpluDev &JT = C->luJT.dev;
    int hmax = 0;   // no scenario of the backward sweep reaches past hmax
    for (int m = 0; m < M; m++) hmax = std::max(hmax, hz[m]);
    for (int s = nseg - 1; s >= 0; s--) {
      const int k0 = s * SEG, k1 = std::min(k0 + SEG, hmax);
      if (k1 <= k0) continue;
      if (s != nseg - 1) {   // recompute the segment from its checkpoint
        FreeSpec Rs;
        Rs.k_first.assign(M, k0 + 1);
        Rs.k_last.assign(M, 0);
        for (int m = 0; m < M; m++)
          if (hz[m] > k0) Rs.k_last[m] = std::min(k1, hz[m]);
        Rs.active0 = TL.active_at.at(k0 + 1);
        Rs.record = false;
        Rs.seg = C->d_seg.p;
        Rs.seg_k0 = k0;
        Rs.seg_k1 = k1;
        launch_copy_plain(C->d_ckpt.p + (size_t)s * nL, C->d_z.p, (long long)nL, C->stream);
        newton_total += integrate_free(C, method, dt, tol, itermax, Rs).newton_iters;
        land_batch(C, true);
      }
      int tab = -1;
      for (int k = k1; k > k0; k--) {
        // fault tables of forward step k (the "+2" alignment of
        // test/test_adjoint_event_offset.jl:3-24)
        if (TL.ver[k] != tab) {
          tab = TL.ver[k];
          HIPCHK(hipMemcpyAsync(C->d_fbus.p, TL.bus[tab].data(), TL.bus[tab].size() * sizeof(int),
                                hipMemcpyHostToDevice, C->stream));
          HIPCHK(hipMemcpyAsync(C->d_fg.p, TL.g[tab].data(), TL.g[tab].size() * sizeof(double),
                                hipMemcpyHostToDevice, C->stream));
        }
        const double *zk = C->d_seg.p + (size_t)(k - k0 - 1) * nL;
        uint8_t *act = C->d_bact.p;
        launch_adj_mask(L, k, C->d_hz.p, act, C->stream);
        launch_adj_carry(n, L, C->d_is_diff.p, C->d_lam.p, C->d_f.p, act, C->stream);
        launch_adj_objective((int)rows.size(), C->d_objrows.p, L, scale * dt, off, zk, C->d_f.p, C->d_val.p, act,
                             C->stream);
        launch_jac_mono(C->T, C->d_p.p, zk, dt, faults(C), C->d_jv.p, act, C->stream);
        launch_splu_factor_impl<double>(JT, C->d_jv.p, nullptr, L, C->d_luJT.p, C->d_svJT.p, C->d_dnJT.p, act,
                                        nullptr, M, C->d_status.p, C->stream);
        launch_splu_solve<double>(JT, C->d_svJT.p, C->d_dnJT.p, C->d_f.p, L, 1.0, nullptr, nullptr, C->d_lam.p, act,
                                  nullptr, M, C->stream);
        launch_adj_param(C->T, C->d_p.p, ne, C->d_ent.p, zk, C->d_lam.p, dt, faults(C), C->d_gacc.p, act, C->stream);
        HIPCHK(hipGetLastError());
      }
      for (int k = k0 + 1; k <= k1; k++)
        for (int m = 0; m < M; m++) A.adjoint_steps += (k <= hz[m]) ? 1 : 0;
      // the next segment's forward recomputation uploads its own tables
    }
    launch_copy_plain(C->d_ckpt.p + (size_t)nseg * nL, C->d_z.p, (long long)nL, C->stream);
    A.backward_ms = back.ms();
    A.elapsed_ms = total.ms();
    A.n_segments = nseg;
    A.newton_iters = newton_total;
    // outputs
    std::vector<double> gacc((size_t)std::max(1, ne) * L), val(L), lam;
    HIPCHK(hipMemcpy(gacc.data(), C->d_gacc.p, gacc.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(val.data(), C->d_val.p, L * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(status.data(), C->d_status.p, L * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemset(C->d_status.p, 0, L * sizeof(int)));
    for (int m = 0; m < M; m++)
      if (status[m] != 0 && hz[m] >= 0) {   // a zero pivot of J^T
        nsing++;
        C->scen_status[m] = GP_ERR_SINGULAR_PIVOT;
        hz[m] = -1;
        A.n_excluded++;
      }
    for (int j = 0; j < ne; j++) {
      double acc = 0.0;
      for (int m = 0; m < M; m++) acc += hz[m] >= 0 ? gacc[(size_t)j * L + m] : 0.0;
      grad[j] = acc;
    }
    if (grad_scen)
      for (int m = 0; m < M; m++)
        for (int j = 0; j < ne; j++) grad_scen[(size_t)m * ne + j] = hz[m] >= 0 ? gacc[(size_t)j * L + m] : 0.0;
    if (value)
      for (int m = 0; m < M; m++) value[m] = hz[m] >= 0 ? val[m] : 0.0;
    if (dz0) {
      lam.resize(nL);
      HIPCHK(hipMemcpy(lam.data(), C->d_lam.p, nL * sizeof(double), hipMemcpyDeviceToHost));
      for (int m = 0; m < M; m++)
        for (int i = 0; i < n; i++)
          dz0[(size_t)m * n + i] = (hz[m] >= 0 && C->is_diff[i]) ? lam[(size_t)i * L + m] : 0.0;
    }
    if (ast) *ast = A;
    if (nsing) return fail(GP_ERR_SINGULAR_PIVOT, std::to_string(nsing) + " scenario(s) hit a zero pivot");
    if (fwd.n_nan) return fail(GP_ERR_NAN, std::to_string(fwd.n_nan) + " scenario(s) produced a non-finite residual");
    if (A.n_excluded)
      return fail(GP_ERR_NEWTON_NOT_CONVERGED, std::to_string(A.n_excluded) +
                                                   " scenario(s) had a Newton solve that did not converge and "
                                                   "were left out of the gradient");
  } catch (const GpError &e) {
    try {
      land_batch(C, true);
    } catch (const GpError &) {
      for (auto &ln : C->lanes) ln.inflight = false;
    }
    if (ast) *ast = A;
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

int gp_get_trajectory(gp_ctx *C, int32_t m, double *out, int64_t ld) {
  if (!C || !out || m < 0 || m >= C->M) return fail(GP_ERR_BADARG, "gp_get_trajectory: bad argument");
  if (C->traj_len.empty()) return fail(GP_ERR_BADARG, "gp_get_trajectory: nothing integrated yet");
  const int len = C->traj_len[m];
  if (ld < len) return fail(GP_ERR_BADARG, "gp_get_trajectory: ld < nsteps[m]+1");
  try {
    const int L = C->ldM;
    const size_t tot = (size_t)len * C->nrec;
    C->d_traj_tmp.ensure(std::max<size_t>(1, tot));
    launch_gather_traj(C->nrec, len, L, m, C->d_hist.p, C->d_traj_tmp.p, C->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> tmp(tot);
    HIPCHK(hipMemcpyAsync(tmp.data(), C->d_traj_tmp.p, tot * sizeof(double), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    for (int j = 0; j < C->nrec; j++)
      for (int k = 0; k < len; k++) out[(size_t)j * ld + k] = tmp[(size_t)j * len + k];
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

int gp_analyse_pivots(int32_t n, const int32_t *colptr, const int32_t *rowind, const double *val, int32_t n_extra,
                      const double *const *extra, double threshold, int32_t *col_order, int32_t *row_of_pivot) {
  if (n <= 0 || !colptr || !rowind || !val || !col_order || !row_of_pivot || n_extra < 0 || (n_extra > 0 && !extra))
    return fail(GP_ERR_BADARG, "gp_analyse_pivots: bad argument");
  for (int c = 0; c < n; c++)
    if (colptr[c + 1] < colptr[c]) return fail(GP_ERR_BADARG, "gp_analyse_pivots: colptr not ascending");
  for (int p = 0; p < colptr[n]; p++)
    if (rowind[p] < 0 || rowind[p] >= n) return fail(GP_ERR_BADARG, "gp_analyse_pivots: row index out of range");
  try {
    std::vector<const double *> ex(extra, extra + n_extra);
    SpluAnalysis A;
    const std::string e = splu_analyse(n, colptr, rowind, val, A, threshold > 0.0 ? threshold : 1e-3, ex);
    if (!e.empty()) return fail(GP_ERR_SINGULAR_PIVOT, "gp_analyse_pivots: " + e);
    for (int k = 0; k < n; k++) {
      col_order[k] = A.q[k];
      row_of_pivot[k] = A.prow[k];
    }
  } catch (const std::exception &e) {
    return fail(GP_ERR_OOM, e.what());
  }
  return GP_OK;
}

int gp_copy_kernel(gp_ctx *C, const double *src, double *dst, int64_t n) {
  if (!C || !src || !dst || n < 0) return fail(GP_ERR_BADARG, "gp_copy_kernel: bad argument");
  try {
    launch_copy_plain(src, dst, n, C->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

int gp_get_scenario_status(gp_ctx *C, int32_t *status, int32_t *steps_done) {
  if (!C) return fail(GP_ERR_BADARG, "gp_get_scenario_status: null context");
  if (C->scen_status.empty()) return fail(GP_ERR_BADARG, "gp_get_scenario_status: nothing integrated yet");
  for (int m = 0; m < C->M; m++) {
    if (status) status[m] = C->scen_status[m];
    if (steps_done) steps_done[m] = C->scen_steps[m];
  }
  return GP_OK;
}

int gp_get_state(gp_ctx *C, double *out) {
  if (!C || !out) return fail(GP_ERR_BADARG, "gp_get_state: null argument");
  try {
    const int L = C->ldM;
    std::vector<double> zt((size_t)C->n * L);
    HIPCHK(hipMemcpy(zt.data(), C->d_z.p, zt.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int m = 0; m < C->M; m++)
      for (int i = 0; i < C->n; i++) out[(size_t)m * C->n + i] = zt[(size_t)i * L + m];
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

static void load_eval_state(gp_ctx *C, const double *z, const uint8_t *fault_on) {
  const int L = C->ldM;
  std::vector<double> zt((size_t)C->n * L, 0.0);
  for (int m = 0; m < C->M; m++)
    for (int i = 0; i < C->n; i++) zt[(size_t)i * L + m] = z[(size_t)m * C->n + i];
  HIPCHK(hipMemcpy(C->d_z.p, zt.data(), zt.size() * sizeof(double), hipMemcpyHostToDevice));
  std::vector<int> fbus;
  std::vector<double> fg;
  eval_fault_tables(C, fault_on, false, fbus, fg);
  HIPCHK(hipMemcpy(C->d_fbus.p, fbus.data(), fbus.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(C->d_fg.p, fg.data(), fg.size() * sizeof(double), hipMemcpyHostToDevice));
}

int gp_eval_residual_noalloc(gp_ctx *C, const double *z, const double *zold, double dt, const uint8_t *fault_on,
                             double *f, void *stream) {
  if (!C || !z || !zold || !f) return fail(GP_ERR_BADARG, "gp_eval_residual_noalloc: null argument");
  if (!C->state_set) return fail(GP_ERR_BADARG, "gp_eval_residual_noalloc: gp_set_state was not called");
  if (!C->scratch) return fail(GP_ERR_HIP, "gp_eval_residual_noalloc: no scratch array on this device");
  const size_t L = (size_t)C->ldM;
  auto up256 = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t o_fb = up256(L), o_fg = o_fb + up256((size_t)MAXFAULT * L * sizeof(int)),
               o_inj = o_fg + up256((size_t)MAXFAULT * L * sizeof(double)),
               need = o_inj + up256((size_t)std::max(1, 2 * C->ncl) * L * sizeof(double));
  if (need > (size_t)GP_NOALLOC_SCRATCH_BYTES)
    return fail(GP_ERR_UNSUPPORTED, "gp_eval_residual_noalloc: this context needs " + std::to_string(need) +
                                        " bytes of scratch, the maximum is " +
                                        std::to_string((size_t)GP_NOALLOC_SCRATCH_BYTES));
  uint8_t *base = static_cast<uint8_t *>(C->scratch);
  uint8_t *mask = base;
  int *fbus = reinterpret_cast<int *>(base + o_fb);
  double *fg = reinterpret_cast<double *>(base + o_fg);
  double *inj = reinterpret_cast<double *>(base + o_inj);
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : C->stream;
  launch_prefix_mask((int)L, C->M, mask, s);
  launch_select_faults((int)L, C->d_evbus.p, C->d_evg.p, fault_on, fbus, fg, s);
  launch_cluster_residual(C->T, C->d_p.p, z, zold, dt, FaultState{fbus, fg, C->nexc}, f, inj, mask, s);
  launch_bus_residual(C->T, C->d_p.p, z, inj, FaultState{fbus, fg, C->nexc}, f, mask, s);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(GP_ERR_HIP, std::string("gp_eval_residual_noalloc: ") + hipGetErrorString(e));
  return GP_OK;
}

int gp_eval_residual(gp_ctx *C, const double *z, const double *zold, double dt, const uint8_t *fault_on, double *f) {
  if (!C || !z || !zold || !f) return fail(GP_ERR_BADARG, "gp_eval_residual: null argument");
  try {
    const int L = C->ldM;
    std::vector<double> zt((size_t)C->n * L, 0.0);
    for (int m = 0; m < C->M; m++)
      for (int i = 0; i < C->n; i++) zt[(size_t)i * L + m] = zold[(size_t)m * C->n + i];
    HIPCHK(hipMemcpy(C->d_zold.p, zt.data(), zt.size() * sizeof(double), hipMemcpyHostToDevice));
    load_eval_state(C, z, fault_on);
    std::vector<uint8_t> all(L, 0);
    for (int m = 0; m < C->M; m++) all[m] = 1;
    set_mask(C, C->d_sel, all);
    launch_cluster_residual(C->T, C->d_p.p, C->d_z.p, C->d_zold.p, dt, faults(C),
                            C->d_f.p, C->d_inj.p, C->d_sel.p, C->stream);
    launch_bus_residual(C->T, C->d_p.p, C->d_z.p, C->d_inj.p, faults(C), C->d_f.p, C->d_sel.p, C->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipMemcpy(zt.data(), C->d_f.p, zt.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int m = 0; m < C->M; m++)
      for (int i = 0; i < C->n; i++) f[(size_t)m * C->n + i] = zt[(size_t)i * L + m];
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

int gp_jacobian_pattern(gp_ctx *C, int32_t *nnz, int32_t *rowptr, int32_t *colind) {
  if (!C || !nnz) return fail(GP_ERR_BADARG, "gp_jacobian_pattern: null argument");
  const int nz = C->jcp[C->n];
  *nnz = nz;
  if (rowptr && colind) {
    // the library's pattern is CSC; report it transposed as CSR of J
    std::vector<int> cnt(C->n + 1, 0);
    for (int q = 0; q < nz; q++) cnt[C->jri[q] + 1]++;
    for (int i = 0; i < C->n; i++) cnt[i + 1] += cnt[i];
    std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
    for (int c = 0; c < C->n; c++)
      for (int q = C->jcp[c]; q < C->jcp[c + 1]; q++) colind[fillp[C->jri[q]]++] = c;
    for (int i = 0; i <= C->n; i++) rowptr[i] = cnt[i];
  }
  return GP_OK;
}

int gp_eval_jacobian(gp_ctx *C, const double *z, double dt, const uint8_t *fault_on, double *jval) {
  if (!C || !z || !jval) return fail(GP_ERR_BADARG, "gp_eval_jacobian: null argument");
  try {
    const int L = C->ldM, nz = C->jcp[C->n];
    if (!C->d_jv.p || C->d_jv.n < (size_t)nz * L) C->d_jv.alloc((size_t)nz * L);
    load_eval_state(C, z, fault_on);
    std::vector<uint8_t> all(L, 0);
    for (int m = 0; m < C->M; m++) all[m] = 1;
    set_mask(C, C->d_sel, all);
    HIPCHK(hipMemsetAsync(C->d_jv.p, 0, (size_t)nz * L * sizeof(double), C->stream));
    launch_jac_mono(C->T, C->d_p.p, C->d_z.p, dt, faults(C), C->d_jv.p, C->d_sel.p, C->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    std::vector<double> jt((size_t)nz * L);
    HIPCHK(hipMemcpy(jt.data(), C->d_jv.p, jt.size() * sizeof(double), hipMemcpyDeviceToHost));
    // CSC position -> CSR position
    std::vector<int> cnt(C->n + 1, 0);
    for (int q = 0; q < nz; q++) cnt[C->jri[q] + 1]++;
    for (int i = 0; i < C->n; i++) cnt[i + 1] += cnt[i];
    std::vector<int> fillp(cnt.begin(), cnt.end() - 1), csr_of(nz);
    for (int c = 0; c < C->n; c++)
      for (int q = C->jcp[c]; q < C->jcp[c + 1]; q++) csr_of[q] = fillp[C->jri[q]]++;
    for (int m = 0; m < C->M; m++)
      for (int q = 0; q < nz; q++) jval[(size_t)m * nz + csr_of[q]] = jt[(size_t)q * L + m];
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

int gp_newton_step(gp_ctx *C, int32_t method, const double *z, const double *zold, double dt,
                   const uint8_t *fault_on, double *dz) {
  if (!C || !z || !zold || !dz) return fail(GP_ERR_BADARG, "gp_newton_step: null argument");
  if (method < 0 || method > 3) return fail(GP_ERR_BADARG, "unknown method");
  if (!C->state_set) return fail(GP_ERR_BADARG, "gp_set_state was not called");
  try {
    const int L = C->ldM;
    // the symbolic analysis evaluates its representative matrix with
    // zold = z, so zold is loaded after it
    load_eval_state(C, z, nullptr);
    analyse(C, method);
    if (method != GP_METHOD_MONO_FP64) ensure_schur_buffers(C, method == GP_METHOD_SCHUR_FP32);
    // shared: P of scenario 0 at z with step dt
    const bool shared = method == GP_METHOD_SCHUR_SHARED;
    if (shared) {   // (the parity hook puts each scenario's first events in force at once)
      std::vector<int> per(C->M, 0);
      int most = 0;
      for (const gp_event &e : C->events) most = std::max(most, ++per[e.scenario]);
      prepare_shared(C, dt, most);
    }
    std::vector<double> zt((size_t)C->n * L, 0.0);
    for (int m = 0; m < C->M; m++)
      for (int i = 0; i < C->n; i++) zt[(size_t)i * L + m] = zold[(size_t)m * C->n + i];
    HIPCHK(hipMemcpy(C->d_zold.p, zt.data(), zt.size() * sizeof(double), hipMemcpyHostToDevice));
    load_eval_state(C, z, fault_on);
    if (shared) {   // the Woodbury rows of the faults in force, as load_eval_state set them
      std::vector<int> fbus;
      std::vector<double> fg;
      eval_fault_tables(C, fault_on, false, fbus, fg);
      for (int m = 0; m < C->M; m++) shared_set(C, m, fbus.data(), fg.data());
      shared_upload(C);
    }
    std::vector<uint8_t> all(L, 0);
    for (int m = 0; m < C->M; m++) all[m] = 1;
    (void)newton(C, method, dt, -1.0, 1, all);   // one correction, nothing counts as converged
    HIPCHK(hipMemcpy(zt.data(), C->d_z.p, zt.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int m = 0; m < C->M; m++)
      for (int i = 0; i < C->n; i++) dz[(size_t)m * C->n + i] = zt[(size_t)i * L + m] - z[(size_t)m * C->n + i];
  } catch (const GpError &e) {
    return fail(e.code, e.msg);
  }
  return GP_OK;
}

int gp_get_symbolic_info(gp_ctx *C, gp_symbolic_info *info) {
  if (!C || !info) return fail(GP_ERR_BADARG, "gp_get_symbolic_info: null argument");
  gp_symbolic_info I{};
  I.n = C->n;
  I.nnz_j = C->jcp[C->n];
  I.nnz_lu_j = C->luJ.ready ? C->luJ.A.nnz_lu : 0;
  I.levels_j = C->luJ.ready ? (int)C->luJ.A.fac_lp.size() - 1 : 0;
  I.n_red = C->nred;
  I.nnz_s = C->scp[C->nred];
  I.nnz_lu_s = C->luS.ready ? C->luS.A.nnz_lu : 0;
  I.levels_s = C->luS.ready ? (int)C->luS.A.fac_lp.size() - 1 : 0;
  I.n_clusters_nt = (int)C->cl_nt.size();
  I.w_total_nt = C->xrows;
  I.nd_j = C->luJ.ready ? C->luJ.A.nd : -1;
  I.nd_s = C->luS.ready ? C->luS.A.nd : -1;
  I.nd_jt = C->luJT.ready ? C->luJT.A.nd : -1;
  I.res_j = C->luJ.ready && C->luJ.dev.res_nsteps > 0;
  I.res_s = C->luS.ready && C->luS.dev.res_nsteps > 0;
  I.ld_scen = C->ldM;
  I.solve_levels_j = C->luJ.ready ? std::max(1, C->luJ.A.part) : 0;
  I.solve_levels_s = C->luS.ready ? std::max(1, C->luS.A.part) : 0;
  *info = I;
  return GP_OK;
}

}  // extern "C"

// gp_internal.hpp -- shared definitions of the host runtime and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gp {

// Scenario-major ("entry-major") arrays are [rows][ldM]: scenario m of row i
// lives at i*ldM + m, so the 64 lanes of a wavefront (64 consecutive
// scenarios) touch 512 contiguous bytes.
constexpr int WAVE = 64;
constexpr int MAXFAULT = 8;   // simultaneously active events (faults, tripped lines, dropped loads) per scenario
constexpr int kDropCode0 = -(1 << 24);   // event slot code of ZIP load 0 disconnected
constexpr int kExcCode0 = -(1 << 25);    // event slot code of SEXS 0 taken offline (see FaultState)

// Jacobian slot counts per device (order of the slot lists built on the host).
constexpr int NS_GENROU = 47;   // incl. optional pm, efd slots (-1 when unwired) and the delta diagonal
constexpr int NS_IEESGO = 15;
constexpr int NS_TGOV1 = 8;
constexpr int NS_SEXS = 9;
constexpr int NS_ESDC1A = 10;
constexpr int NS_IEEEST = 32;   // incl. the two BE diagonals (s1, s3) absent from the device rows

// Device-resident tables (read-only after gp_create).
struct DevTables {
  int n, nbus, net_ptr, ldM, M;
  const uint8_t *is_diff;
  const int *genrou, *ieesgo, *tgov1, *sexs, *esdc1a, *ieeest, *zip, *sg;
  int n_genrou, n_ieesgo, n_tgov1, n_sexs, n_esdc1a, n_ieeest, n_zip, n_sg;
  const double *u0;
  const int *cl;          // cluster rows, GP_CLUSTER_COLS ints each
  int n_cl;
  // Y realified CSR
  const int *yrp, *yci;
  const double *yv;
  // per-bus attachments (CSR over buses)
  const int *bus_cl_p, *bus_cl;      // non-trivial clusters at each bus
  const int *bus_zip_p, *bus_zip;    // ZIP loads at each bus
  const int *bus_sg;                 // static generator at each bus (-1)
  // Jacobian position tables for the monolithic path
  const int *jp_genrou, *jp_ieesgo, *jp_tgov1, *jp_sexs, *jp_esdc1a, *jp_ieeest;
  const int *jp_bus;      // per bus: 4 diag positions + 6 static-gen positions
  const int *jp_yoff_p, *jp_yoff_pos;   // per bus row pair: off-diagonal Y entries
  const double *jp_yoff_val;
  const double *ydiag;    // per bus: 4 diagonal Y values (rr, ri, ir, ii)
  // Schur path: reduced indexing and S slots
  const int *cl_nt;       // list of non-trivial cluster ids
  int n_cl_nt;
  const int *cl_xoff;     // per cluster: offset (in 3-column rows) into X
  const int *red_of_z;    // z index -> reduced index (-1 if eliminated)
  const int *sv_bus;      // per bus: 4 var slots of S diag + 6 static-gen var slots
  // line trips: per trip record t, its buses (from, to) and its admittance
  // stamp (yff, yft, ytf, ytt as re, im); per bus, the off-diagonal Y entries
  // of its rows that some trip changes: 4 ints each (column bus, component
  // 2*rr + cc of the 2x2 block, J position, S slot) and the entry of -Y
  const int *tr_bus;
  const double *tr_y;
  const int *tv_p, *tv_i;
  const double *tv_base;
};

// Per-scenario event slots: bus[k*ldM+m] >= 0 is a fault on that bus with
// conductance g[k*ldM+m]; kDropCode0 < bus <= -2 is the trip of line record
// t = -2 - bus; kExcCode0 < bus <= kDropCode0 is ZIP load kDropCode0 - bus
// disconnected; bus <= kExcCode0 is exciter x = kExcCode0 - bus taken
// offline (SEXS row x/2 for even x, ESDC1A row x/2 for odd x); -1 is an
// empty slot.  nexc: whether any exciter codes are in the tables (0: the
// cluster kernels skip the search).
struct FaultState {
  int *bus;
  double *g;
  int nexc = 0;
};

// Woodbury correction of a shared factorisation P (one per scenario, applied
// by the solve before its outputs): x -= W_c * s with s = H^{-1} x[pos_c],
// over the nc[m] columns c of scenario m.  W: [column][n] in the solve's
// internal (pivot) order; col/pos: [m][kw]; h: [m][kw][kw] row-major.
// nc == NULL: no correction.
struct WbTab {
  const int *nc = nullptr, *col = nullptr, *pos = nullptr;
  const double *h = nullptr, *W = nullptr;
  int kw = 0;
};

struct SpluDev {
  int n, nnz_a, nnz_lu, ns, n_flev, nd, qb, nnz_top;
  const int *a2lu, *a_src;      // per A entry: LU position, value source
  const int *op_loc, *op_sl;    // update program (compact form)
  const int *st_o;              // wavefront-mode step bounds
  const int *flp, *fac_mode, *crec;   // factor levels: column records, mode
  const int *lu2s;              // LU position -> solve slot (-1: not used by the solves)
  const int *dmap;              // dense block: entry of the block's columns -> dense index
  // solves, sliced ELL: per level (T, g, R, first chunk), per chunk (E, pf),
  // per value slot its column, per lane its row
  int n_fwlev, n_bwlev, n_fwch, n_bwch, fw_nvals, bw_nvals;
  const int *fw_lvl, *fw_chunk, *fw_cols, *fw_rows;
  const int *bw_lvl, *bw_chunk, *bw_cols, *bw_rows;
  const int *q, *pinv;
  // partitioned solves (part > 1): a second vector y in LDS, levels flagged
  // A (a partition's rows from the values before it, x -> y forward, y -> x
  // backward) or B (the other rows' entries in its columns); the compose
  // program fills the A-level values after every factorisation
  int part, cmp_nph;
  const int *cmp_ph, *cmp_tp, *cmp_dst, *cmp_ops;
  int threads;                  // workgroup size of both kernels
  int t_cols;                   // LDS staging per thread (thread mode), 0 = none
  int w_col, w_ops, w_bytes;    // LDS staging per wavefront (wavefront mode)
  int lds_factor, lds_solve;    // dynamic LDS bytes
  // LDS-resident factorisation (res_nsteps > 0): step headers, blobs, A
  // entry -> step-ordered position, LDS elements of the multiplier / dense
  // block region
  int res_nsteps, res_lv, res_nl, res_xtot;
  // staged triangular solve: steps per side, headers and blobs
  int sol_nfw, sol_nbw;
  // systems whose solution does not fit in LDS: the solve keeps x in global
  // memory, gx_ws [n][ldM] (scenario m's x at j*ldM + m, pivot order)
  int gx;
  double *gx_ws;
  // the leading forward / trailing backward levels wide enough to run as
  // launches of their own across the device, and (host memory, read by the
  // launcher only) the level tables for their lane counts
  int gx_fw_wide, gx_bw_wide;
  const int *h_fw_lvl, *h_bw_lvl;
  const int *sol_fw_hdr, *sol_fw_blob, *sol_bw_hdr, *sol_bw_blob;
  const int *res_hdr, *res_blob, *res_a2x, *res_omap;
  // per step-ordered value slot of the LDS-resident factorisation: where its
  // initial value comes from (a_src code: var row >= 0, constant -1 - i, or
  // kResNoSrc for a fill entry, which starts at zero)
  const int *res_xsrc;
};
constexpr int kResNoSrc = -2147483647 - 1;
// staging of the LDS-resident factorisation, per thread of a 512-thread
// workgroup: blob dwords and initial column values
constexpr int kResCB = 24, kResCX = 8;
// value slots per thread of the resident factorisation's step staging: FP32
// kResCX, FP64 half as many (its multipliers and values take twice the LDS)
constexpr int res_cx(size_t elem) { return elem == 8 ? kResCX / 2 : kResCX; }
// staging of the triangular solve, per thread: blob dwords, values, diagonals
constexpr int kSolCB = 8, kSolCD = 4;
// values per thread of a solve step: 16 for FP32 factors, 8 for FP64 (the
// staging area of either fits beside x and y in LDS)
constexpr int sol_cv(int elem) { return elem == 4 ? 16 : 8; }

// Parameters with a hand-written pullback in the discrete adjoint.
enum AdjKind {
  GP_ADJ_GENROU_H = 0, GP_ADJ_GENROU_D, GP_ADJ_SEXS_K, GP_ADJ_ESDC1A_KA, GP_ADJ_TGOV1_R, GP_ADJ_IEESGO_K1,
  GP_ADJ_IEEEST_KS
};

// ---- kernel launchers (gp_kernels.hip) ----------------------------------
// dtv (optional): per-scenario step, replacing dt (0 for an event re-solve)
// pm (optional): per working scenario, max |F| over each non-trivial
// cluster's rows (row c of [n_cl_nt + nbus][ldM]) and each bus's rows (row
// n_cl_nt + b), for the max-norm
void launch_cluster_residual(const DevTables &T, const double *p, const double *z, const double *zold,
                             double dt, FaultState fs, double *f, double *inj, const uint8_t *work, hipStream_t s,
                             const double *dtv = nullptr, double *pm = nullptr);
void launch_bus_residual(const DevTables &T, const double *p, const double *z, const double *inj,
                         FaultState fs, double *f, const uint8_t *work, hipStream_t s, double *pm = nullptr);
void launch_rownorm_impl(int n, int M, int ldM, const double *f, uint8_t *work, double tol, double *fnorm,
                         unsigned long long *nb, int *n_unconv, uint8_t *refac, const uint8_t *refac0, int first,
                         double theta, int stall, int itermax, double *fbest, int *nstall, int *iters,
                         uint8_t *failed, uint8_t *pend, uint8_t *upd, double async_below, uint8_t *areq,
                         uint8_t *sch, hipStream_t s, const uint8_t *firstv = nullptr,
                         const uint8_t *exact = nullptr, int switch_window = 0);
void launch_jac_mono(const DevTables &T, const double *p, const double *z, double dt, FaultState fs,
                     double *jv, const uint8_t *work, hipStream_t s, const double *dtv = nullptr);
template <typename TF>
void launch_cluster_schur(const DevTables &T, const double *p, const double *z, const double *f,
                          double dt, FaultState fs, TF *X, double *Dg, const uint8_t *work, hipStream_t s,
                          const double *dtv = nullptr);
void launch_bus_schur(const DevTables &T, const double *p, const double *z, const double *f,
                      const double *Dg, FaultState fs, double *svar, double *rhs, const uint8_t *work,
                      hipStream_t s);
template <typename TF>
void launch_cluster_backsub(const DevTables &T, const TF *X, const double *f, const double *dv,
                            double *z, const uint8_t *work, hipStream_t s);
// luv: [M][nnz_lu] factor workspace; sv: [2][M][ns] solve values; dn:
// [2][M][nd*nd] dense trailing block (column-major LU, unit lower part
// implied).  The factorisation writes buffer tgt[m], the solve reads cur[m]
// (NULL: buffer 0).  The matrix values var are rows [s][ldM] (ldM > 0) or
// scenario-major rows [m][-ldM] (ldM < 0).  A plan with partitioned solves
// then computes their composite entries (k_splu_compose); `factored`, when
// given, is recorded between the two kernels.
template <typename TF>
void launch_splu_factor_impl(const SpluDev &S, const double *var, const double *cst, int ldM, TF *luv, TF *sv,
                             TF *dn, const uint8_t *refac, const uint8_t *tgt, int M, int *status, hipStream_t s,
                             hipEvent_t factored = nullptr);
template <typename TF>
void launch_splu_solve(const SpluDev &S, const TF *sv, const TF *dn, const double *rhs, int ldM, double sign,
                       const int *scatter, double *zupd, double *xout, const uint8_t *work, const uint8_t *cur, int M,
                       hipStream_t s, bool shared = false,   // shared: slot 0's factors serve every scenario
                       const WbTab *wb = nullptr);           // with a Woodbury correction per scenario
void launch_copy_active(int n, int ldM, const double *src, double *dst, const uint8_t *act, hipStream_t s);
// dst[m*n + r] = src[r*ldM + m] for the scenarios in act
void launch_copy_transpose(int n, int ldM, const double *src, double *dst, const uint8_t *act, hipStream_t s);
void launch_advance(int n, int ldM, double *z, double *zold, double *zold2, const uint8_t *adv, const uint8_t *pred,
                    const uint8_t *acc, const int *col, const int *rec_p, const int *rec_ch, int nrec, double *hist,
                    hipStream_t s);
// columns of the scenarios in mask into slot[m] of a stack of scenario-major arrays
void launch_copy_cols(int n, int ldM, const double *src, double *dst, size_t stride, const int *slot,
                      const uint8_t *mask, hipStream_t s);
// asynchronous-refactorisation flags: mode 0 landed (cur ^= 1, pend = 0),
// 1 requested (pend = 1), 2 dropped (pend = 0), for the scenarios in a
void launch_async_flags(int L, int mode, const uint8_t *a, uint8_t *cur, uint8_t *pend, hipStream_t s);
// a batch's request mask as kernel arguments: one bit per scenario, up to
// kStageBits scenarios (larger batches upload their mask instead)
constexpr int kStageBits = 2048;
struct ScenarioBits {
  uint32_t w[kStageBits / 32];
};
void launch_stage_batch(int L, const ScenarioBits &bits, uint8_t *req, uint8_t *tgt, const uint8_t *cur,
                        uint8_t *pend, hipStream_t s);
// the round's (sch, refac, upd) masks split by factor precision: out [6][L],
// FP32 scenarios' masks then those of the scenarios with x64[m] != 0
void launch_split_precision(int L, const uint8_t *x64, const uint8_t *sch, const uint8_t *refac, const uint8_t *upd,
                            uint8_t *out, hipStream_t s);
// discrete adjoint: backward-step mask, carried multipliers, objective rows,
// parameter pullbacks
void launch_adj_mask(int ldM, int k, const int *hz, uint8_t *act, hipStream_t s);
void launch_adj_carry(int n, int ldM, const uint8_t *is_diff, const double *lam, double *rhs, const uint8_t *act,
                      hipStream_t s);
void launch_adj_objective(int nr, const int *rows, int ldM, double w, double off, const double *z, double *rhs,
                          double *val, const uint8_t *act, hipStream_t s);
void launch_adj_param(const DevTables &T, const double *p, int ne, const int2 *ent, const double *z,
                      const double *lam, double dt, FaultState fs, double *g, const uint8_t *act, hipStream_t s);
void launch_gather_traj(int nrec, int len, int ldM, int m, const double *hist, double *out, hipStream_t s);
void launch_copy_plain(const double *src, double *dst, long long n, hipStream_t s);
// allocation-free residual: the address of this GPU's scratch array (taken
// once per device, that device current; 0 or an error code), the all-M
// scenario mask and the selected fault tables
int scratch_setup(int device, void **out);
void launch_prefix_mask(int ldM, int M, uint8_t *mask, hipStream_t s);
void launch_select_faults(int ldM, const int *ev_bus, const double *ev_g, const uint8_t *fault_on, int *bus,
                          double *g, hipStream_t s);
void launch_record_impl(int nrec, const int *idx, int ldM, int M, const double *z, double *hist,
                        const uint8_t *act, hipStream_t s);
}  // namespace gp

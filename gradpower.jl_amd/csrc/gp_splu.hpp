// gp_splu.hpp -- host-side symbolic analysis for the batched sparse LU.
//
// One analysis serves every scenario and every Newton iteration: the
// matrices of all scenarios share a pattern, so the fill-reducing column
// order, the pivot rows, the L/U patterns and the level schedules are
// computed once on the host (the role KLU's symbolic analysis and first
// factorisation play for the reference CPU path, src/numerics.jl:138-167,
// and cuDSS's "analysis" phase for its GPU path, ext:776-819).  The device
// then only runs numeric refactorisations with the fixed pivot sequence and
// triangular solves.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gp {

constexpr int kCmpBlock = 4096;   // terms of one block of the compose program (LDS of the compose kernel)

struct SpluAnalysis {
  int n = 0;
  int nnz_a = 0;
  // permutations: pivot position k has column q[k] and row prow[k]
  std::vector<int> q, qinv, prow, pinv;
  // LU storage in pivot space, by columns: column k holds U(0..k-1,k) rows
  // (ascending), the diagonal, then L(k+1..,k) rows (ascending).
  std::vector<int> lu_cp;   // [n+1]
  std::vector<int> lu_ri;   // [nnz_lu] pivot-space row
  std::vector<int> lu_dp;   // [n] position of the diagonal
  int nnz_lu = 0;
  std::vector<int> a2lu;    // [nnz_a] destination of each A entry (CSC order)
  // factorisation program: for column k, ops [op_cp[k], op_cp[k+1]) are
  // triples (dst, srcL, srcU): lu[dst] -= lu[srcL] * lu[srcU], grouped by
  // source column in ascending order so every multiplier is final when used.
  std::vector<int> op_cp;
  std::vector<int> op;      // 3 ints per op
  // factorisation levels: columns of level l are fac_cols[fac_lp[l] .. fac_lp[l+1])
  std::vector<int> fac_lp, fac_cols;
  int max_col = 0;          // largest column (entries of U + diag + L)
  // row-oriented access for the solves (pivot space):
  //   forward:  y[i] -= sum_k L(i,k) y[k]   rows of L:  lrow_p / lrow_k / lrow_pos
  //   backward: x[k] = (y[k] - sum_j U(k,j) x[j]) / U(k,k)
  std::vector<int> lrow_p, lrow_k, lrow_pos;
  std::vector<int> urow_p, urow_j, urow_pos;
  // Solve streams.  The factorisation also writes every final value into a
  // solve array laid out in the order the solves read it: L rows in forward
  // level order, then U rows (off-diagonal) in backward level order, then
  // the n diagonal entries.  lu2s maps an LU position to its solve slot.
  std::vector<int> lu2s;             // [nnz_lu]
  int nnz_l = 0, nnz_u = 0;
  std::vector<int> ls_k;             // [nnz_l] column of each L stream entry
  std::vector<int> us_j;             // [nnz_u] column of each U stream entry
  std::vector<int> fw_info;          // [3n] per forward slot: row, stream begin, end
  std::vector<int> bw_info;          // [3n] per backward slot: row, stream begin, end
  // factorisation ops in compact form: op_loc[o] = (dst - c0) | (srcU - c0) << 16
  // (offsets inside the target column), op_sl[o] = position of L(i,j)
  std::vector<int> op_loc, op_sl;
  std::vector<int> fw_lp, fw_rows;   // forward levels
  std::vector<int> bw_lp, bw_rows;   // backward levels
  long long flops = 0;               // multiply-add pairs of one factorisation
  // Launch schedule (splu_schedule).  The trailing nd pivots form a dense
  // block (B = n - nd): its columns first receive the updates from columns
  // < B (one extra level), then a dense LU without pivoting finishes it.
  // A factorisation level runs one thread per column (mode 0: wide levels of
  // short columns), one wavefront per column (mode 1: narrow levels of long
  // columns) or the block's external updates (mode 2, wavefront per column).
  // In wavefront mode the updates of a column are split into steps: step s
  // applies ops [st_o[s], st_o[s+1]), at most 64 of them, all with the same
  // source column, one op per lane.
  int nd = 0;
  long long est_cycles = 0;          // cost-model estimate of one factorisation
  std::vector<int> st_cp, st_o;      // [n+1] first step of column k; [steps+1] op bounds
  std::vector<int> fac_mode;         // [levels]
  std::vector<int> flp;              // [levels+1] first column record of each level
  std::vector<int> crec;             // 8 ints per column record: k, c0, len, dl (mode 2: rows
                                     // above the block), o0, nops, first step, end step
  std::vector<int> dmap;             // per entry of the block's columns: dense index or -1
  // solve levels: rows < B (forward, then one level of the block rows'
  // entries in columns < B) and rows < B (backward); 3 ints per slot: row,
  // stream begin, stream end; g = lanes cooperating on one row
  std::vector<int> sfw_lp, sfw_info, sbw_lp, sbw_info;
  // Sliced-ELL solve layout (forward side sf over L, backward side sb over U).
  struct SolveSide {
    std::vector<int> lvl;     // 4 ints per level: lanes T, lanes per row g, first lane R, first chunk
    std::vector<int> lflag;   // per level: bit 0 A-level of a partition (partitioned solves), bit 1 divide
    std::vector<int> chunk;   // 2 ints per chunk of 64 lanes: first value E, entries per lane pf
    std::vector<int> cols;    // [nvals] column of each value slot (0 in padding)
    std::vector<int> rows;    // per lane (all levels): its row
    int nvals = 0;
  };
  SolveSide sf, sb;
  // Partitioned solves (part > 1: up to part levels per partition).  The
  // A-level values are composites the compose program computes after each
  // factorisation from the raw entries inside the partitions (nraw slots
  // after the diagonals): block p covers entries [cmp_ph[p], cmp_ph[p+1])
  // (blocks in phase order, each of one depth, at most kCmpBlock terms);
  // entry e writes solve slot cmp_dst[e], the sum of its terms
  // [cmp_tp[e], cmp_tp[e+1]), 4 ints each (0, a, b, d):
  // v[a] / (d >= 0 ? v[d] : 1) for b < 0, -v[a] / (d >= 0 ? v[d] : 1) * v[b]
  // otherwise.
  int part = 0, nraw = 0;
  std::vector<int> cmp_ph, cmp_tp, cmp_dst, cmp_ops;
  int ns = 0;                        // solve values per scenario: sf values, sb values, sb diagonals, raw
  int t_cols_max = 0;                // largest column in a thread-mode level
  int w_col_max = 0, w_ops_max = 0;  // largest column / op count in a wavefront-mode level
};

// Build the launch schedule for workgroups of `threads` threads, with a
// dense trailing block of at most nd_max pivots.
void splu_schedule(SpluAnalysis &A, int threads, int nd_max = 112, int part_k = 0);

// Analyse the CSC matrix (cp, ri, representative values ax).  With extra
// representative matrices of the same pattern (the line trips make one per
// distinct set of lines out at once), each pivot is chosen large in all of
// them: the largest min over the matrices of |x|/max|x| among the
// candidates, the diagonal when that measure reaches the threshold -- so no
// pivot relies on an entry that a trip, or the fill it feeds, makes zero.
// A matrix in which no candidate of a column is nonzero (a trip set that
// islands part of the network) leaves the rule for that column; one whose
// chosen pivot is zero leaves it for good (its elimination cannot go on).
// Where no candidate is nonzero in every matrix still in the rule, the
// candidate nonzero in the most of them is taken.  Returns an empty string
// on success, otherwise an error message.
std::string splu_analyse(int n, const int *cp, const int *ri, const double *ax,
                         SpluAnalysis &out, double pivot_threshold = 1e-3,
                         const std::vector<const double *> &extra = {});

// Plan of the LDS-resident factorisation.  Every L entry of the columns
// below the dense block (B = n - nd) that an update reads is kept in LDS
// (Lv, nL values, slot nL holding zero) once its column is final, so an
// update reads only LDS (a column's L entries take consecutive slots from
// its lbase).  The levels of the schedule are cut into steps whose static
// data (a blob of dwords) and initial column values fit the staging
// capacities; each column has len + 1 value slots (the last one a spare that
// padding updates write), stored contiguously per step in a step-ordered
// copy of the LU values (a2x: A entry -> position, xtot values).  A level
// runs g lanes per column (a power of two): the updates of one step (one
// source entry, distinct destinations) are laid out in rows of g, padded
// with no-op updates (spare slot, multiplier nL), and a group of g lanes
// applies one row at a time.  Blob of a step, in dwords:
//   8 per column: x offset, len, dl, first row, rows, rows of its
//                 wavefront in this round (the loop bound), lbase, 0
//   2 per op:     op_loc, Lv index of the multiplier (rows of g)
// omap, per value slot: its solve slot (lu2s), -2 - d for entry d of the
// dense block, or -1.  Step blob and x range start at multiples of 4.
// Step header (kResHdr ints): mode, ncols, blob offset, blob length, x
// offset, x length, ops offset in the blob, 0, 0, g.
constexpr int kResHdr = 12;
struct ResPlan {
  int nL = 0;
  int maxb = 0, maxx = 0;          // largest step blob (dwords) / x length
  int xtot = 0;                    // step-ordered values per scenario
  std::vector<int> a2x;            // [nnz_a]
  std::vector<int> omap;           // [xtot]
  std::vector<int> hdr;            // kResHdr per step
  std::vector<int> blob;
  int nsteps() const { return (int)hdr.size() / kResHdr; }
};
// Plan of the staged triangular solve (one side).  Consecutive levels are
// grouped into steps whose data fit the staging capacities; a level larger
// than that is cut into sub-levels of whole chunks (its rows are
// independent).  A (sub-)level runs in passes of 8 chunks, one per
// wavefront.  Step blob, in dwords: 36 per pass (lanes per row g, 1 if a
// barrier follows, the level's flags (SolveSide::lflag), 0, then per
// wavefront: first value of its chunk relative to the step, entries per
// lane pf, first lane relative to the step, active lanes (0: idle)), the
// step's value columns (16 bits each, two per dword), then its lanes'
// rows.  The step's values (and, for the backward side, diagonals) are the
// ranges [val_off, + val_len) and [dg_off, + dg_len) of the solve array;
// a step holds at most vcap values, dcap diagonals and bcap blob dwords.
// Step header (kSolHdr ints): passes, blob offset, blob length, val_off,
// val_len, dg_off, dg_len, 0, cols offset, rows offset, values and
// diagonals actually present (the lengths before padding to 4).  All
// offsets and lengths of blob sections are multiples of 4.
constexpr int kSolHdr = 12, kSolLvl = 4;
struct SolvePlan {
  std::vector<int> hdr, blob;
  int maxb = 0, maxv = 0;
  int nsteps() const { return (int)hdr.size() / kSolHdr; }
};
bool solve_plan(const SpluAnalysis::SolveSide &S, bool diag, int dg_base, int bcap, int vcap, int dcap,
                SolvePlan &P);

// Returns false (and leaves R empty) if a column alone exceeds a capacity.
bool res_plan(const SpluAnalysis &A, int bcap, int xcap, ResPlan &R);

// Minimum-degree order of the symmetric pattern of A + A^T.
std::vector<int> min_degree_order(int n, const int *cp, const int *ri);

}  // namespace gp

// gp_splu.cpp -- symbolic analysis for the batched sparse LU (see gp_splu.hpp).
#include "gp_splu.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <queue>
#include <utility>

namespace gp {

std::vector<int> min_degree_order(int n, const int *cp, const int *ri) {
  std::vector<std::vector<int>> adj(n);
  for (int c = 0; c < n; c++)
    for (int p = cp[c]; p < cp[c + 1]; p++) {
      int r = ri[p];
      if (r == c) continue;
      adj[c].push_back(r);
      adj[r].push_back(c);
    }
  for (auto &a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  using Item = std::pair<int, int>;  // (degree, node)
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  std::vector<int> deg(n);
  std::vector<char> done(n, 0);
  for (int v = 0; v < n; v++) {
    deg[v] = (int)adj[v].size();
    pq.push({deg[v], v});
  }
  std::vector<int> order;
  order.reserve(n);
  std::vector<int> merged;
  while (!pq.empty()) {
    auto [d, v] = pq.top();
    pq.pop();
    if (done[v] || d != deg[v]) continue;
    done[v] = 1;
    order.push_back(v);
    const std::vector<int> nb = adj[v];
    for (int u : nb) {
      auto &au = adj[u];
      // au <- (au U nb) \ {u, v}
      merged.clear();
      std::set_union(au.begin(), au.end(), nb.begin(), nb.end(), std::back_inserter(merged));
      au.clear();
      for (int w : merged)
        if (w != u && w != v && !done[w]) au.push_back(w);
      deg[u] = (int)au.size();
      pq.push({deg[u], u});
    }
    adj[v].clear();
    adj[v].shrink_to_fit();
  }
  return order;
}

std::string splu_analyse(int n, const int *cp, const int *ri, const double *ax, SpluAnalysis &A,
                         double thr, const std::vector<const double *> &extra) {
  A = SpluAnalysis();
  A.n = n;
  A.nnz_a = cp[n];
  A.q = min_degree_order(n, cp, ri);
  if ((int)A.q.size() != n) return "ordering failed";
  A.qinv.assign(n, 0);
  for (int k = 0; k < n; k++) A.qinv[A.q[k]] = k;

  // ---- numeric left-looking LU with threshold partial pivoting ----------
  const int ne = (int)extra.size();
  std::vector<int> pinv(n, -1), prow(n, -1);
  std::vector<std::vector<std::pair<int, double>>> Lcol(n), Ucol(n);  // L: original rows; U: pivot rows
  std::vector<double> x(n, 0.0);
  // the extra matrices: column work vectors and L values parallel to Lcol
  std::vector<std::vector<double>> xe(ne, std::vector<double>(n, 0.0));
  std::vector<std::vector<std::vector<double>>> Le(ne, std::vector<std::vector<double>>(n));
  std::vector<uint8_t> alive(ne, 1), inrule(ne, 0);
  std::vector<double> amaxe(ne, 0.0);
  std::vector<int> xi(n), mark(n, -1), stk(n), pst(n);
  for (int k = 0; k < n; k++) {
    int col = A.q[k];
    int top = n;
    for (int p = cp[col]; p < cp[col + 1]; p++) {
      int r = ri[p];
      if (mark[r] == k) continue;
      int head = 0;
      stk[0] = r;
      mark[r] = k;
      pst[0] = 0;
      while (head >= 0) {
        int j = stk[head], jp = pinv[j];
        bool pushed = false;
        if (jp >= 0) {
          auto &Lj = Lcol[jp];
          for (int pp = pst[head]; pp < (int)Lj.size(); pp++) {
            int i = Lj[pp].first;
            if (mark[i] != k) {
              pst[head] = pp + 1;
              stk[++head] = i;
              mark[i] = k;
              pst[head] = 0;
              pushed = true;
              break;
            }
          }
        }
        if (!pushed) {
          head--;
          xi[--top] = j;
        }
      }
    }
    for (int p = top; p < n; p++) x[xi[p]] = 0.0;
    for (int p = cp[col]; p < cp[col + 1]; p++) x[ri[p]] = ax[p];
    for (int p = top; p < n; p++) {
      int j = xi[p], jp = pinv[j];
      if (jp < 0) continue;
      double xj = x[j];
      for (auto &e : Lcol[jp]) x[e.first] -= e.second * xj;
    }
    for (int e = 0; e < ne; e++) {   // the same column of each extra matrix, same pivots so far
      if (!alive[e]) continue;
      std::vector<double> &y = xe[e];
      for (int p = top; p < n; p++) y[xi[p]] = 0.0;
      for (int p = cp[col]; p < cp[col + 1]; p++) y[ri[p]] = extra[e][p];
      for (int p = top; p < n; p++) {
        int j = xi[p], jp = pinv[j];
        if (jp < 0) continue;
        double yj = y[j];
        const auto &lc = Lcol[jp];
        const auto &lv = Le[e][jp];
        for (size_t q = 0; q < lc.size(); q++) y[lc[q].first] -= lv[q] * yj;
      }
    }
    int ipiv = -1;
    double amax = -1.0;
    for (int e = 0; e < ne; e++) amaxe[e] = 0.0;
    for (int p = top; p < n; p++) {
      int i = xi[p];
      if (pinv[i] < 0) {
        double t = std::fabs(x[i]);
        if (t > amax) { amax = t; ipiv = i; }
        for (int e = 0; e < ne; e++)
          if (alive[e]) amaxe[e] = std::max(amaxe[e], std::fabs(xe[e][i]));
      } else {
        Ucol[k].push_back({pinv[i], x[i]});
      }
    }
    if (ipiv < 0 || !(amax > 0.0)) return "structurally or numerically singular matrix at pivot " + std::to_string(k);
    int nin = 0;
    for (int e = 0; e < ne; e++) {
      inrule[e] = alive[e] && amaxe[e] > 0.0;
      nin += inrule[e];
    }
    if (nin == 0) {
      if (pinv[col] < 0 && std::fabs(x[col]) >= thr * amax) ipiv = col;
    } else {
      // the measure of a candidate: min over the matrices in the rule of
      // |x|/max|x|, and (ties, all zero) how many of them it is nonzero in
      auto score = [&](int i, int &cnt) {
        double sc = std::fabs(x[i]) / amax;
        cnt = 0;
        for (int e = 0; e < ne; e++) {
          if (!inrule[e]) continue;
          const double v = std::fabs(xe[e][i]) / amaxe[e];
          cnt += v > 0.0;
          sc = std::min(sc, v);
        }
        return sc;
      };
      double best = -1.0;
      int bcnt = -1, ib = -1;
      for (int p = top; p < n; p++) {
        int i = xi[p];
        if (pinv[i] >= 0 || x[i] == 0.0) continue;
        int cnt;
        const double sc = score(i, cnt);
        if (cnt > bcnt || (cnt == bcnt && sc > best)) {
          best = sc;
          bcnt = cnt;
          ib = i;
        }
      }
      if (ib >= 0) ipiv = ib;
      if (pinv[col] < 0 && x[col] != 0.0) {
        int cnt;
        const double sc = score(col, cnt);
        if (cnt == bcnt && sc >= thr) ipiv = col;
      }
    }
    double piv = x[ipiv];
    pinv[ipiv] = k;
    prow[k] = ipiv;
    for (int e = 0; e < ne; e++)
      if (alive[e] && xe[e][ipiv] == 0.0) alive[e] = 0;   // its elimination cannot go on
    for (int p = top; p < n; p++) {
      int i = xi[p];
      if (pinv[i] < 0) {
        Lcol[k].push_back({i, x[i] / piv});
        for (int e = 0; e < ne; e++)
          if (alive[e]) Le[e][k].push_back(xe[e][i] / xe[e][ipiv]);
      }
      x[i] = 0.0;
    }
    for (int e = 0; e < ne; e++)
      if (alive[e])
        for (int p = top; p < n; p++) xe[e][xi[p]] = 0.0;
    // Le[e][k] of a matrix that left the rule for good stays empty: never read
    for (int e = 0; e < ne; e++)
      if (!alive[e]) Le[e][k].clear();
  }
  A.pinv = pinv;
  A.prow = prow;

  // ---- LU storage in pivot space -------------------------------------------
  A.lu_cp.assign(n + 1, 0);
  A.lu_dp.assign(n, 0);
  std::vector<std::vector<int>> urows(n), lrows(n);
  for (int k = 0; k < n; k++) {
    for (auto &e : Ucol[k]) urows[k].push_back(e.first);
    for (auto &e : Lcol[k]) lrows[k].push_back(pinv[e.first]);
    std::sort(urows[k].begin(), urows[k].end());
    std::sort(lrows[k].begin(), lrows[k].end());
  }
  for (int k = 0; k < n; k++) {
    A.lu_cp[k + 1] = A.lu_cp[k] + (int)urows[k].size() + 1 + (int)lrows[k].size();
    A.max_col = std::max(A.max_col, A.lu_cp[k + 1] - A.lu_cp[k]);
  }
  A.nnz_lu = A.lu_cp[n];
  A.lu_ri.assign(A.nnz_lu, 0);
  for (int k = 0; k < n; k++) {
    int p = A.lu_cp[k];
    for (int r : urows[k]) A.lu_ri[p++] = r;
    A.lu_dp[k] = p;
    A.lu_ri[p++] = k;
    for (int r : lrows[k]) A.lu_ri[p++] = r;
  }
  auto pos = [&](int i, int k) -> int {  // position of pivot row i in column k
    const int *b = A.lu_ri.data() + A.lu_cp[k];
    const int *e = A.lu_ri.data() + A.lu_cp[k + 1];
    // U part ascending, then diag (=k), then L part ascending: the whole
    // column is ascending in pivot rows.
    const int *it = std::lower_bound(b, e, i);
    if (it == e || *it != i) return -1;
    return (int)(it - A.lu_ri.data());
  };
  A.a2lu.assign(A.nnz_a, -1);
  for (int c = 0; c < n; c++)
    for (int p = cp[c]; p < cp[c + 1]; p++) {
      int d = pos(pinv[ri[p]], A.qinv[c]);
      if (d < 0) return "internal: A entry missing from LU pattern";
      A.a2lu[p] = d;
    }

  // ---- factorisation program + levels ------------------------------------
  A.op_cp.assign(n + 1, 0);
  std::vector<int> flev(n, 0);
  int nlev = 0;
  for (int k = 0; k < n; k++) {
    int lev = 0;
    for (int pu = A.lu_cp[k]; pu < A.lu_dp[k]; pu++) {
      int j = A.lu_ri[pu];
      lev = std::max(lev, flev[j] + 1);
      for (int pl = A.lu_dp[j] + 1; pl < A.lu_cp[j + 1]; pl++) {
        int i = A.lu_ri[pl];
        int d = pos(i, k);
        if (d < 0) return "internal: fill entry missing";
        A.op.push_back(d);
        A.op.push_back(pl);
        A.op.push_back(pu);
        A.flops++;
      }
    }
    flev[k] = lev;
    nlev = std::max(nlev, lev + 1);
    A.op_cp[k + 1] = (int)(A.op.size() / 3);
  }
  A.fac_lp.assign(nlev + 1, 0);
  for (int k = 0; k < n; k++) A.fac_lp[flev[k] + 1]++;
  for (int l = 0; l < nlev; l++) A.fac_lp[l + 1] += A.fac_lp[l];
  A.fac_cols.assign(n, 0);
  {
    std::vector<int> fillp(A.fac_lp.begin(), A.fac_lp.end() - 1);
    for (int k = 0; k < n; k++) A.fac_cols[fillp[flev[k]]++] = k;
  }

  // ---- row-oriented solve structures ----------------------------------------
  std::vector<std::vector<std::pair<int, int>>> lr(n), ur(n);
  for (int k = 0; k < n; k++) {
    for (int p = A.lu_dp[k] + 1; p < A.lu_cp[k + 1]; p++) lr[A.lu_ri[p]].push_back({k, p});
    for (int p = A.lu_cp[k]; p < A.lu_dp[k]; p++) ur[A.lu_ri[p]].push_back({k, p});
  }
  A.lrow_p.assign(n + 1, 0);
  A.urow_p.assign(n + 1, 0);
  for (int i = 0; i < n; i++) {
    A.lrow_p[i + 1] = A.lrow_p[i] + (int)lr[i].size();
    A.urow_p[i + 1] = A.urow_p[i] + (int)ur[i].size();
    for (auto &e : lr[i]) { A.lrow_k.push_back(e.first); A.lrow_pos.push_back(e.second); }
    for (auto &e : ur[i]) { A.urow_j.push_back(e.first); A.urow_pos.push_back(e.second); }
  }
  A.nnz_l = A.lrow_p[n];
  A.nnz_u = A.urow_p[n];
  std::vector<int> lv(n, 0);
  int nl = 0;
  for (int i = 0; i < n; i++) {
    int l = 0;
    for (auto &e : lr[i]) l = std::max(l, lv[e.first] + 1);
    lv[i] = l;
    nl = std::max(nl, l + 1);
  }
  A.fw_lp.assign(nl + 1, 0);
  for (int i = 0; i < n; i++) A.fw_lp[lv[i] + 1]++;
  for (int l = 0; l < nl; l++) A.fw_lp[l + 1] += A.fw_lp[l];
  A.fw_rows.assign(n, 0);
  {
    std::vector<int> f(A.fw_lp.begin(), A.fw_lp.end() - 1);
    for (int i = 0; i < n; i++) A.fw_rows[f[lv[i]]++] = i;
  }
  std::fill(lv.begin(), lv.end(), 0);
  nl = 0;
  for (int k = n - 1; k >= 0; k--) {
    int l = 0;
    for (auto &e : ur[k]) l = std::max(l, lv[e.first] + 1);
    lv[k] = l;
    nl = std::max(nl, l + 1);
  }
  A.bw_lp.assign(nl + 1, 0);
  for (int i = 0; i < n; i++) A.bw_lp[lv[i] + 1]++;
  for (int l = 0; l < nl; l++) A.bw_lp[l + 1] += A.bw_lp[l];
  A.bw_rows.assign(n, 0);
  {
    std::vector<int> f(A.bw_lp.begin(), A.bw_lp.end() - 1);
    for (int i = 0; i < n; i++) A.bw_rows[f[lv[i]]++] = i;
  }
  // ---- solve streams in level order -----------------------------------------
  A.lu2s.assign(A.nnz_lu, -1);
  A.ls_k.clear();
  A.us_j.clear();
  A.fw_info.assign(3 * n, 0);
  A.bw_info.assign(3 * n, 0);
  for (int idx = 0; idx < n; idx++) {
    int i = A.fw_rows[idx];
    A.fw_info[3 * idx] = i;
    A.fw_info[3 * idx + 1] = (int)A.ls_k.size();
    for (int e = A.lrow_p[i]; e < A.lrow_p[i + 1]; e++) {
      A.lu2s[A.lrow_pos[e]] = (int)A.ls_k.size();
      A.ls_k.push_back(A.lrow_k[e]);
    }
    A.fw_info[3 * idx + 2] = (int)A.ls_k.size();
  }
  for (int idx = 0; idx < n; idx++) {
    int k = A.bw_rows[idx];
    A.bw_info[3 * idx] = k;
    A.bw_info[3 * idx + 1] = (int)A.us_j.size();
    for (int e = A.urow_p[k]; e < A.urow_p[k + 1]; e++) {
      A.lu2s[A.urow_pos[e]] = A.nnz_l + (int)A.us_j.size();
      A.us_j.push_back(A.urow_j[e]);
    }
    A.bw_info[3 * idx + 2] = (int)A.us_j.size();
  }
  for (int k = 0; k < n; k++) A.lu2s[A.lu_dp[k]] = A.nnz_l + A.nnz_u + k;
  // ---- compact ops -------------------------------------------------------------
  const int nop = (int)A.op.size() / 3;
  A.op_loc.assign(nop, 0);
  A.op_sl.assign(nop, 0);
  for (int k = 0; k < n; k++) {
    const int c0 = A.lu_cp[k];
    for (int o = A.op_cp[k]; o < A.op_cp[k + 1]; o++) {
      const int d = A.op[3 * o] - c0, su = A.op[3 * o + 2] - c0;
      if (d < 0 || d > 65535 || su < 0 || su > 65535) return "column too large for compact ops";
      A.op_loc[o] = d | (su << 16);
      A.op_sl[o] = A.op[3 * o + 1];
    }
  }
  return "";
}

// ---- launch schedule ------------------------------------------------------------
// Cost model in GPU cycles, fitted to level timings of the factor kernel on a
// 4,300-row reduced network: a thread-mode round pays a fixed latency plus
// ~220 cycles per op of its longest column; a wavefront-mode column pays its
// staging plus ~40 cycles per op and ~150 per step; the dense block pays
// ~400 cycles per pivot.
namespace {
long long thread_level_cost(long long rounds, long long maxops) { return rounds * (10000 + 220 * maxops); }
long long wave_col_cost(long long nops, long long steps) { return 4000 + 40 * nops + 150 * steps; }
long long dense_cost(int nd) { return nd ? 6000 + 400LL * nd : 0; }
}  // namespace

void splu_schedule(SpluAnalysis &A, int threads, int nd_max, int part_k) {
  const int n = A.n, W = threads / 64;
  const int nl = (int)A.fac_lp.size() - 1;
  // steps: maximal runs of ops with the same source entry, cut at 64
  A.st_cp.assign(n + 1, 0);
  A.st_o.clear();
  for (int k = 0; k < n; k++) {
    A.st_cp[k] = (int)A.st_o.size();
    int o = A.op_cp[k];
    while (o < A.op_cp[k + 1]) {
      const int su = A.op_loc[o] >> 16;
      A.st_o.push_back(o);
      int cnt = 0;
      while (o < A.op_cp[k + 1] && (A.op_loc[o] >> 16) == su && cnt < 64) { o++; cnt++; }
    }
  }
  A.st_cp[n] = (int)A.st_o.size();
  A.st_o.push_back(A.op_cp[n]);
  // external steps of column k >= B: the steps whose source column is < B
  // (sources are visited in ascending order, so they form a prefix)
  auto ext_steps = [&](int k, int B) {
    const int c0 = A.lu_cp[k];
    int s = A.st_cp[k];
    while (s < A.st_cp[k + 1] && A.lu_ri[c0 + (A.op_loc[A.st_o[s]] >> 16)] < B) s++;
    return s;
  };
  auto level_cost = [&](const std::vector<int> &cols, int B, bool ext, int &mode) {
    long long maxops = 0;
    std::vector<long long> wsum(W, 0);
    for (size_t i = 0; i < cols.size(); i++) {
      const int k = cols[i];
      const int se = ext ? ext_steps(k, B) : A.st_cp[k + 1];
      const int nops = A.st_o[se] - A.op_cp[k];
      maxops = std::max<long long>(maxops, nops);
      wsum[i % W] += wave_col_cost(nops, se - A.st_cp[k]);
    }
    const long long tc = thread_level_cost((long long)(cols.size() + threads - 1) / threads, maxops);
    const long long wc = *std::max_element(wsum.begin(), wsum.end());
    mode = ext ? 2 : (wc < tc ? 1 : 0);
    return ext ? wc : std::min(tc, wc);
  };
  // sparse levels restricted to columns < B
  auto sparse_levels = [&](int B) {
    std::vector<std::vector<int>> lv;
    for (int l = 0; l < nl; l++) {
      std::vector<int> c;
      for (int i = A.fac_lp[l]; i < A.fac_lp[l + 1]; i++)
        if (A.fac_cols[i] < B) c.push_back(A.fac_cols[i]);
      if (!c.empty()) lv.push_back(std::move(c));
    }
    return lv;
  };
  // dense block size: the trailing nd pivots, nd * nd * 8 bytes within LDS
  int best_nd = 0;
  long long best = -1;
  for (int nd = 0; nd <= std::min(std::min(nd_max, 128), n / 2); nd += 16) {
    const int B = n - nd;
    long long cost = dense_cost(nd);
    int mode;
    for (auto &c : sparse_levels(B)) cost += 2000 + level_cost(c, B, false, mode);
    if (nd) {
      std::vector<int> top;
      for (int k = B; k < n; k++) top.push_back(k);
      cost += 2000 + level_cost(top, B, true, mode);
    }
    if (best < 0 || cost < best) { best = cost; best_nd = nd; }
  }
  A.nd = best_nd;
  const int B = n - A.nd;
  A.est_cycles = best;
  // factor levels and column records
  auto lvls = sparse_levels(B);
  if (A.nd) {
    std::vector<int> top;
    for (int k = B; k < n; k++) top.push_back(k);
    lvls.push_back(top);
  }
  A.flp.assign(1, 0);
  A.fac_mode.clear();
  A.crec.clear();
  A.t_cols_max = A.w_col_max = A.w_ops_max = 0;
  for (size_t l = 0; l < lvls.size(); l++) {
    const bool ext = A.nd && l + 1 == lvls.size();
    int mode;
    level_cost(lvls[l], B, ext, mode);
    A.fac_mode.push_back(mode);
    for (int k : lvls[l]) {
      const int c0 = A.lu_cp[k], len = A.lu_cp[k + 1] - c0;
      const int se = ext ? ext_steps(k, B) : A.st_cp[k + 1];
      const int nops = A.st_o[se] - A.op_cp[k];
      int d = A.lu_dp[k] - c0;
      if (ext) {   // rows above the block: final after the external updates
        d = 0;
        while (d < len && A.lu_ri[c0 + d] < B) d++;
      }
      const int r[8] = {k, c0, len, d, A.op_cp[k], nops, A.st_cp[k], se};
      A.crec.insert(A.crec.end(), r, r + 8);
      if (mode) {
        A.w_col_max = std::max(A.w_col_max, len);
        A.w_ops_max = std::max(A.w_ops_max, nops);
      } else {
        A.t_cols_max = std::max(A.t_cols_max, len);
      }
    }
    A.flp.push_back((int)A.crec.size() / 8);
  }
  // dense block map: entry q of the block's columns -> (r - B) + (k - B) * nd, or -1 above the block
  A.dmap.clear();
  for (int k = B; k < n; k++)
    for (int q = A.lu_cp[k]; q < A.lu_cp[k + 1]; q++) {
      const int r = A.lu_ri[q];
      A.dmap.push_back(r >= B ? (r - B) + (k - B) * A.nd : -1);
    }
  // stream ranges per row
  std::vector<int> lb(n), le(n), ub(n), ue(n);
  for (int i = 0; i < n; i++) {
    lb[A.fw_info[3 * i]] = A.fw_info[3 * i + 1];
    le[A.fw_info[3 * i]] = A.fw_info[3 * i + 2];
    ub[A.bw_info[3 * i]] = A.bw_info[3 * i + 1];
    ue[A.bw_info[3 * i]] = A.bw_info[3 * i + 2];
  }
  // forward levels: rows < B keep their levels; the block rows form one more
  // level that applies only their entries in columns < B
  std::vector<int> flp{0}, finfo;
  for (int l = 0; l + 1 < (int)A.fw_lp.size(); l++) {
    bool any = false;
    for (int i = A.fw_lp[l]; i < A.fw_lp[l + 1]; i++) {
      const int r = A.fw_rows[i];
      if (r >= B) continue;
      finfo.insert(finfo.end(), {r, lb[r], le[r]});
      any = true;
    }
    if (any) flp.push_back((int)finfo.size() / 3);
  }
  if (A.nd) {
    for (int r = B; r < n; r++) {
      int e = lb[r];
      while (e < le[r] && A.ls_k[e] < B) e++;
      finfo.insert(finfo.end(), {r, lb[r], e});
    }
    flp.push_back((int)finfo.size() / 3);
  }
  // backward levels of rows < B (the block rows are solved densely first)
  std::vector<int> blev(n, -1);
  int nbl = 0;
  for (int k = B - 1; k >= 0; k--) {
    int l = 0;
    for (int e = ub[k]; e < ue[k]; e++) {
      const int j = A.us_j[e];
      if (j < B) l = std::max(l, blev[j] + 1);
    }
    blev[k] = l;
    nbl = std::max(nbl, l + 1);
  }
  std::vector<int> blp(nbl + 1, 0), binfo(3 * (size_t)B);
  for (int k = 0; k < B; k++) blp[blev[k] + 1]++;
  for (int l = 0; l < nbl; l++) blp[l + 1] += blp[l];
  {
    std::vector<int> f(blp.begin(), blp.end() - 1);
    for (int k = B - 1; k >= 0; k--) {
      const int slot = f[blev[k]]++;
      binfo[3 * slot] = k;
      binfo[3 * slot + 1] = ub[k];
      binfo[3 * slot + 2] = ue[k];
    }
  }
  A.sfw_lp = flp;
  A.sfw_info = finfo;
  A.sbw_lp = blp;
  A.sbw_info = binfo;
  // ---- sliced-ELL solve layout ------------------------------------------------
  // Each level's rows (longest first) get g lanes each; lanes are cut into
  // chunks of 64 (one wavefront), and a chunk stores its entries lane-major,
  // padded to its longest lane: entry j of lane u of chunk c sits at
  // E_c + 64 j + (u mod 64).  Every address a lane reads is then known
  // without reading the matrix, so a level's loads are issued during the
  // previous level.  An entry's source is an LU position (lu2s maps it to
  // its slot) or, partitioned, a composite value (-1 - id; cslot gets its
  // slot).
  A.lu2s.assign(A.nnz_lu, -1);
  struct Lrow {
    int row;
    std::vector<std::pair<int, int>> ent;   // (column, source)
  };
  std::vector<int> cslot;
  auto build = [&](const std::vector<std::vector<Lrow>> &levels, const std::vector<int> &flags,
                   SpluAnalysis::SolveSide &S, int base, std::vector<int> *lane_of_a) {
    S = SpluAnalysis::SolveSide{};
    int R = 0;
    for (size_t l = 0; l < levels.size(); l++) {
      std::vector<const Lrow *> rows;
      for (const Lrow &r : levels[l]) rows.push_back(&r);
      std::stable_sort(rows.begin(), rows.end(),
                       [](const Lrow *a, const Lrow *b) { return a->ent.size() > b->ent.size(); });
      const int nr = (int)rows.size();
      if (nr == 0) continue;
      const int maxlen = (int)rows[0]->ent.size();
      int g = 1;
      long long bestc = -1;
      for (int gg = 1, lg = 0; gg <= 64; gg *= 2, lg++) {
        const long long passes = ((long long)nr * gg + threads - 1) / threads;
        const long long c = passes * ((maxlen + gg - 1) / gg + lg + 2);
        if (bestc < 0 || c < bestc) { bestc = c; g = gg; }
      }
      const int T = nr * g;
      const int nch = (T + 63) / 64;
      S.lvl.insert(S.lvl.end(), {T, g, R, (int)S.chunk.size() / 2});
      S.lflag.push_back(flags[l]);
      for (int c = 0; c < nch; c++) {
        // longest lane of the chunk: its first row is the longest
        const int r0 = (64 * c) / g;
        const int len0 = (int)rows[r0]->ent.size();
        const int pf = std::max(1, (len0 + g - 1) / g);   // >= 1: every lane has a readable slot
        const int E = S.nvals;
        S.chunk.insert(S.chunk.end(), {E, pf});
        S.nvals += 64 * pf;
        S.cols.resize(S.nvals, 0);
        for (int u = 64 * c; u < std::min(T, 64 * c + 64); u++) {
          const int r = u / g, sub = u % g;
          const auto &ent = rows[r]->ent;
          for (int e = sub, jj = 0; e < (int)ent.size(); e += g, jj++) {
            const int at = E + 64 * jj + (u - 64 * c);
            S.cols[at] = ent[e].first;
            const int src = ent[e].second;
            if (src >= 0)
              A.lu2s[src] = base + at;
            else
              cslot[-1 - src] = base + at;
          }
        }
      }
      for (int u = 0; u < T; u++) {
        S.rows.push_back(rows[u / g]->row);
        if (lane_of_a && flags[l] && u % g == 0) (*lane_of_a)[rows[u / g]->row] = R + u;
      }
      R += T;
    }
  };
  // stream index -> LU position (streams were laid out from lrow/urow order)
  std::vector<int> lpos(A.nnz_l), upos(A.nnz_u);
  for (int i = 0; i < n; i++) {
    int sb = lb[i];
    for (int e = A.lrow_p[i]; e < A.lrow_p[i + 1]; e++) lpos[sb++] = A.lrow_pos[e];
    sb = ub[i];
    for (int e = A.urow_p[i]; e < A.urow_p[i + 1]; e++) upos[sb++] = A.urow_pos[e];
  }
  A.part = 0;
  A.cmp_ph.assign(1, 0);
  A.cmp_tp.assign(1, 0);
  A.cmp_dst.clear();
  A.cmp_ops.clear();
  int nraw = 0;
  if (part_k < 2) {
    std::vector<std::vector<Lrow>> lf, lbk;
    for (int l = 0; l + 1 < (int)flp.size(); l++) {
      lf.emplace_back();
      for (int s = flp[l]; s < flp[l + 1]; s++) {
        Lrow x{finfo[3 * s], {}};
        for (int e = finfo[3 * s + 1]; e < finfo[3 * s + 2]; e++) x.ent.push_back({A.ls_k[e], lpos[e]});
        lf.back().push_back(std::move(x));
      }
    }
    for (int l = 0; l + 1 < (int)blp.size(); l++) {
      lbk.emplace_back();
      for (int s = blp[l]; s < blp[l + 1]; s++) {
        Lrow x{binfo[3 * s], {}};
        for (int e = binfo[3 * s + 1]; e < binfo[3 * s + 2]; e++) x.ent.push_back({A.us_j[e], upos[e]});
        lbk.back().push_back(std::move(x));
      }
    }
    build(lf, std::vector<int>(lf.size(), 0), A.sf, 0, nullptr);
    build(lbk, std::vector<int>(lbk.size(), 2), A.sb, A.sf.nvals, nullptr);   // backward rows divide
    // diagonal of row k stored per backward lane (the group's first lane carries it)
    const int dbase = A.sf.nvals + A.sb.nvals;
    std::vector<int> first(n, -1);
    for (int u = (int)A.sb.rows.size() - 1; u >= 0; u--) first[A.sb.rows[u]] = u;
    for (int k = 0; k < B; k++)
      if (first[k] >= 0) A.lu2s[A.lu_dp[k]] = dbase + first[k];
  } else {
    // Partitioned solves: the levels of rows < B are grouped part_k at a
    // time (fewer where a row's chain grows long).  Within partition P a
    // forward row i is finished from the values before P alone:
    //   y[i] = x[i] - sum_{j in reach(i)} V(i,j) x[j],
    //   V(i,j) = [j in dep(i)] L(i,j) - sum_{k in dep(i) & P} L(i,k) V(k,j),
    // (reach(i): the rows of P that i depends on, transitively), written to
    // the second vector y (A-level); the rows outside P then take their own
    // entries in P's columns from y (B-level).  Backward alike with
    //   W(i,j) = [j in dep(i)] U(i,j)/U(j,j) - sum_k U(i,k)/U(k,k) W(k,j)
    // and the division by U(i,i), from y (the forward result, updated in
    // place) into x.  A factorisation computes the composites V, W from the
    // raw entries inside the partitions (the compose program, phases by depth
    // inside a partition).
    A.part = part_k;
    std::vector<int> rawof(A.nnz_lu, -1);   // LU position of an entry inside a partition -> raw index
    struct Comp { int a, b, d, cid_b; };     // one term: raw a, composite cid_b (-1: direct), diag row d (-1)
    std::vector<std::vector<Comp>> cterms;   // per composite id
    std::vector<int> cdepth;                 // per composite id: depth of its row in the partition
    std::vector<int> a_row_of_diag;          // per composite term: row whose diagonal divides (bw), via lane map
    auto side = [&](bool fw, std::vector<std::vector<Lrow>> &levels, std::vector<int> &flags) {
      // rows < B by level
      std::vector<int> lev(n, -1);
      int nlev = 0;
      const std::vector<int> &lp = fw ? flp : blp;
      const std::vector<int> &inf = fw ? finfo : binfo;
      for (int l = 0; l + 1 < (int)lp.size(); l++)
        for (int s = lp[l]; s < lp[l + 1]; s++)
          if (inf[3 * s] < B) { lev[inf[3 * s]] = l; nlev = std::max(nlev, l + 1); }
      std::vector<std::vector<int>> byl(nlev);
      for (int r = 0; r < B; r++)
        if (lev[r] >= 0) byl[lev[r]].push_back(r);
      // dependencies (column, LU position) of row r in the sparse part
      auto deps = [&](int r) {
        std::vector<std::pair<int, int>> d;
        if (fw) {
          for (int e = A.lrow_p[r]; e < A.lrow_p[r + 1]; e++) d.push_back({A.lrow_k[e], A.lrow_pos[e]});
        } else {
          for (int e = A.urow_p[r]; e < A.urow_p[r + 1]; e++)
            if (A.urow_j[e] < B) d.push_back({A.urow_j[e], A.urow_pos[e]});
        }
        return d;
      };
      if (!fw) {   // the rows' entries in the dense block's columns (final before the sparse part)
        std::vector<Lrow> lv0;
        for (int r = 0; r < B; r++) {
          Lrow x{r, {}};
          for (int e = A.urow_p[r]; e < A.urow_p[r + 1]; e++)
            if (A.urow_j[e] >= B) x.ent.push_back({A.urow_j[e], A.urow_pos[e]});
          if (!x.ent.empty()) lv0.push_back(std::move(x));
        }
        if (!lv0.empty()) { levels.push_back(std::move(lv0)); flags.push_back(0); }
      }
      std::vector<int> pid(n, -1);
      std::vector<std::vector<int>> reach(n);   // sorted rows of the partition
      int l0 = 0, np = 0;
      while (l0 < nlev) {
        // grow the partition level by level while the chains stay short
        int l1 = l0 + 1;
        for (int r : byl[l0]) { pid[r] = np; reach[r].clear(); }
        while (l1 < nlev && l1 - l0 < part_k) {
          bool ok = true;
          std::vector<std::pair<int, std::vector<int>>> add;
          for (int r : byl[l1]) {
            std::vector<int> rc;
            for (auto &d : deps(r))
              if (pid[d.first] == np) {
                rc.push_back(d.first);
                rc.insert(rc.end(), reach[d.first].begin(), reach[d.first].end());
              }
            std::sort(rc.begin(), rc.end());
            rc.erase(std::unique(rc.begin(), rc.end()), rc.end());
            if ((int)rc.size() > 48) { ok = false; break; }
            add.push_back({r, std::move(rc)});
          }
          if (!ok) break;
          for (auto &x : add) { pid[x.first] = np; reach[x.first] = std::move(x.second); }
          l1++;
        }
        // A-level: the partition's rows; composites in depth order
        std::vector<Lrow> al;
        for (int l = l0; l < l1; l++)
          for (int r : byl[l]) {
            Lrow x{r, {}};
            for (int j : reach[r]) {
              const int cid = (int)cterms.size();
              cterms.emplace_back();
              cdepth.push_back(l - l0);
              x.ent.push_back({j, -1 - cid});
            }
            al.push_back(std::move(x));
          }
        // composite terms (ids of a row's reach in order)
        {
          std::map<std::pair<int, int>, int> cid_of;
          for (auto &x : al)
            for (auto &e : x.ent) cid_of[{x.row, e.first}] = -1 - e.second;
          for (auto &x : al)
            for (auto &e : x.ent) {
              const int j = e.first, cid = -1 - e.second;
              for (auto &d : deps(x.row)) {
                const int k = d.first;
                if (pid[k] != np) continue;
                int ra = rawof[d.second];
                if (ra < 0) ra = rawof[d.second] = nraw++;
                if (k == j) cterms[cid].push_back({ra, -1, fw ? -1 : k, -1});
                else if (std::binary_search(reach[k].begin(), reach[k].end(), j))
                  cterms[cid].push_back({ra, -1, fw ? -1 : k, cid_of.at({k, j})});
              }
            }
        }
        levels.push_back(std::move(al));
        flags.push_back(fw ? 1 : 3);   // A-level (backward: dividing by the row's diagonal)
        // B-level: later rows with entries in the partition's columns
        std::vector<Lrow> blv;
        auto outside = [&](int r) {
          Lrow x{r, {}};
          if (fw) {
            for (int e = A.lrow_p[r]; e < A.lrow_p[r + 1]; e++)
              if (A.lrow_k[e] < B && pid[A.lrow_k[e]] == np) x.ent.push_back({A.lrow_k[e], A.lrow_pos[e]});
          } else {
            for (int e = A.urow_p[r]; e < A.urow_p[r + 1]; e++)
              if (A.urow_j[e] < B && pid[A.urow_j[e]] == np) x.ent.push_back({A.urow_j[e], A.urow_pos[e]});
          }
          if (!x.ent.empty()) blv.push_back(std::move(x));
        };
        for (int l = l1; l < nlev; l++)
          for (int r : byl[l]) outside(r);
        if (fw)
          for (int r = B; r < n; r++) outside(r);
        if (!blv.empty()) { levels.push_back(std::move(blv)); flags.push_back(0); }
        np++;
        l0 = l1;
      }
    };
    std::vector<std::vector<Lrow>> lf, lbk;
    std::vector<int> ff, fb;
    side(true, lf, ff);
    const int ncf = (int)cterms.size();
    side(false, lbk, fb);
    cslot.assign(cterms.size(), -1);
    std::vector<int> alane(n, -1);
    build(lf, ff, A.sf, 0, nullptr);
    build(lbk, fb, A.sb, A.sf.nvals, &alane);
    const int dbase = A.sf.nvals + A.sb.nvals;
    for (int k = 0; k < B; k++)
      if (alane[k] >= 0) A.lu2s[A.lu_dp[k]] = dbase + alane[k];
    const int rawbase = dbase + (int)A.sb.rows.size();
    for (int p = 0; p < A.nnz_lu; p++)
      if (rawof[p] >= 0) A.lu2s[p] = rawbase + rawof[p];
    // compose program: phases by depth (forward and backward composites
    // together), per entry its terms (4 ints: 0, a, b, d); a phase is cut
    // into blocks of at most kCmpBlock terms at entry boundaries (the blocks
    // of a phase are independent), cmp_ph listing every block's first entry
    int maxd = 0;
    for (int c = 0; c < (int)cterms.size(); c++) maxd = std::max(maxd, cdepth[c]);
    A.cmp_tp.assign(1, 0);
    for (int d = 1; d <= maxd; d++) {
      for (int c = 0; c < (int)cterms.size(); c++) {
        if (cdepth[c] != d) continue;
        if ((int)A.cmp_ops.size() / 4 - A.cmp_tp[A.cmp_ph.back()] + (int)cterms[c].size() > kCmpBlock &&
            (int)A.cmp_dst.size() > A.cmp_ph.back())
          A.cmp_ph.push_back((int)A.cmp_dst.size());   // a new block
        A.cmp_dst.push_back(cslot[c]);
        for (const Comp &t : cterms[c]) {
          const int dg = t.d >= 0 ? dbase + alane[t.d] : -1;
          A.cmp_ops.insert(A.cmp_ops.end(), {0, rawbase + t.a, t.cid_b >= 0 ? cslot[t.cid_b] : -1, dg});
        }
        A.cmp_tp.push_back((int)A.cmp_ops.size() / 4);
      }
      if ((int)A.cmp_dst.size() > A.cmp_ph.back()) A.cmp_ph.push_back((int)A.cmp_dst.size());
    }
    (void)ncf;
  }
  A.nraw = nraw;
  A.ns = A.sf.nvals + A.sb.nvals + (int)A.sb.rows.size() + nraw;
  // the dense block's own entries are read from the dense copy, not the streams
  for (int k = B; k < n; k++)
    for (int q = A.lu_cp[k]; q < A.lu_cp[k + 1]; q++)
      if (A.lu_ri[q] >= B) A.lu2s[q] = -1;
}

bool res_plan(const SpluAnalysis &A, int bcap, int xcap, ResPlan &R) {
  R = ResPlan{};
  const int n = A.n, B = n - A.nd;
  const int qb = A.lu_cp[B];
  const int nlev = (int)A.flp.size() - 1;
  // Lv slots: the L entries of the columns below the block, each column's in
  // one contiguous range, live from the level that writes them to the last
  // level whose updates read them; ranges are reused once dead (first fit),
  // which keeps about a third of the L entries in LDS at once.  Slot nL is
  // zero.  A range is released only for levels after its last reader, so a
  // column written in level l never overwrites what another column of
  // level l reads.
  std::vector<int> lslot(A.nnz_lu, -1);
  {
    std::vector<int> lev_of(n, -1), colof(A.nnz_lu, -1), last(B, -1);
    for (int l = 0; l < nlev; l++)
      for (int i = A.flp[l]; i < A.flp[l + 1]; i++) lev_of[A.crec[8 * i]] = l;
    for (int k = 0; k < n; k++)
      for (int p = A.lu_cp[k]; p < A.lu_cp[k + 1]; p++) colof[p] = k;
    for (int l = 0; l < nlev; l++)
      for (int i = A.flp[l]; i < A.flp[l + 1]; i++) {
        const int *r = &A.crec[8 * i];
        for (int o = r[4]; o < r[4] + r[5]; o++) {
          const int j = colof[A.op_sl[o]];
          if (j >= 0 && j < B) last[j] = std::max(last[j], l);
        }
      }
    std::vector<std::pair<int, int>> freel;   // free ranges (start, length), by start
    std::vector<std::vector<std::pair<int, int>>> release(nlev + 1);
    auto give_back = [&](int st, int ln) {
      auto it = std::lower_bound(freel.begin(), freel.end(), std::make_pair(st, 0));
      it = freel.insert(it, {st, ln});
      if (it + 1 != freel.end() && it->first + it->second == (it + 1)->first) {
        it->second += (it + 1)->second;
        freel.erase(it + 1);
      }
      if (it != freel.begin() && (it - 1)->first + (it - 1)->second == it->first) {
        (it - 1)->second += it->second;
        freel.erase(it);
      }
    };
    for (int l = 0; l < nlev; l++) {
      for (const auto &rg : release[l]) give_back(rg.first, rg.second);
      for (int i = A.flp[l]; i < A.flp[l + 1]; i++) {
        const int k = A.crec[8 * i];
        if (k >= B) continue;
        const int ln = A.lu_cp[k + 1] - A.lu_dp[k] - 1;
        if (ln <= 0) continue;
        int base = -1;
        for (auto &fr : freel)
          if (fr.second >= ln) {
            base = fr.first;
            fr.first += ln;
            fr.second -= ln;
            break;
          }
        freel.erase(std::remove_if(freel.begin(), freel.end(), [](const std::pair<int, int> &q) { return q.second == 0; }),
                    freel.end());
        if (base < 0) {
          base = R.nL;
          R.nL += ln;
        }
        for (int p = A.lu_dp[k] + 1; p < A.lu_cp[k + 1]; p++) lslot[p] = base + (p - A.lu_dp[k] - 1);
        const int e = std::max(last[k], l);   // written in level l even if never read
        release[std::min(e + 1, nlev)].push_back({base, ln});
      }
    }
  }
  std::vector<int> xstart(n, -1);
  auto al4 = [](size_t v) { return (v + 3) & ~(size_t)3; };
  // Rows of a column with g lanes: its ops (in program order) packed into
  // rows of at most g, an op going to the first row after every op that
  // writes its source entry, with room left and no other op writing its
  // destination.  Ops on one destination commute and the sources ascend, so
  // any such packing gives the sequential result.  At least one row.
  auto pack_rows = [&](int i2, int g, std::vector<std::vector<int>> *out) {
    const int *r = &A.crec[8 * i2];
    const int len = r[2], o0 = r[4], nops = r[5];
    std::vector<int> ready(len, 0);           // first row that may read entry t
    std::vector<int> cnt;                     // ops per row
    std::vector<std::vector<int>> dsts;       // destinations per row
    if (out) out->clear();
    for (int o = o0; o < o0 + nops; o++) {
      const int d = A.op_loc[o] & 0xffff, src = A.op_loc[o] >> 16;
      int row = ready[src];
      while (row < (int)cnt.size() &&
             (cnt[row] >= g || std::find(dsts[row].begin(), dsts[row].end(), d) != dsts[row].end()))
        row++;
      if (row == (int)cnt.size()) {
        cnt.push_back(0);
        dsts.emplace_back();
        if (out) out->emplace_back();
      }
      cnt[row]++;
      dsts[row].push_back(d);
      if (out) (*out)[row].push_back(o);
      ready[d] = std::max(ready[d], row + 1);
    }
    return std::max<long long>((long long)cnt.size(), 1);
  };
  auto col_rows = [&](int i2, int g) { return pack_rows(i2, g, nullptr); };
  auto col_cost = [&](int i2, int g) { return col_rows(i2, g) + (A.crec[8 * i2 + 2] + g - 1) / g + 4; };
  auto col_blob = [&](int i2, int g) { return 8 + 2 * (size_t)g * col_rows(i2, g); };
  int xtot = 0;
  for (int l = 0; l < nlev; l++) {
    const int mode = A.fac_mode[l];
    std::vector<int> lev;
    for (int i = A.flp[l]; i < A.flp[l + 1]; i++) lev.push_back(i);
    // lanes per column g: the groups of the workgroup take the columns in
    // rounds, longest first; cost in row latencies, ties to the smaller blob
    std::stable_sort(lev.begin(), lev.end(), [&](int x, int y) { return col_cost(x, 1) > col_cost(y, 1); });
    int g = 1;
    long long cbest = -1;
    size_t bbest = 0;
    for (int gg = 1; gg <= 64; gg *= 2) {
      const int G = 512 / gg;
      long long t = 0;
      size_t bl = 0;
      for (size_t c = 0; c < lev.size(); c += G) {
        long long mx = 0;
        for (size_t c2 = c; c2 < std::min(lev.size(), c + G); c2++) mx = std::max(mx, col_cost(lev[c2], gg));
        t += mx + 8;
      }
      for (int i2 : lev) bl += col_blob(i2, gg);
      if (cbest < 0 || t < cbest || (t == cbest && bl < bbest)) { cbest = t; g = gg; bbest = bl; }
    }
    const int G = 512 / g, gw = 64 / g;
    size_t i = 0;
    while (i < lev.size()) {
      // one step: as many of the level's columns as fit
      std::vector<int> cols;
      size_t blen = 0;
      int xlen = 0;
      for (; i < lev.size(); i++) {
        const int len = A.crec[8 * lev[i] + 2];
        const size_t cb = col_blob(lev[i], g);
        if (al4(blen + cb) > (size_t)bcap || (int)al4(xlen + len + 1) > xcap) {
          if (cols.empty()) { R = ResPlan{}; return false; }
          break;
        }
        cols.push_back(lev[i]);
        blen += cb;
        xlen += len + 1;
      }
      const int nc = (int)cols.size();
      // rows every group of a wavefront runs: the largest among the columns
      // the wavefront holds in the same round
      std::vector<int> rows(nc), wrows(nc);
      for (int c = 0; c < nc; c++) rows[c] = (int)col_rows(cols[c], g);
      for (int c = 0; c < nc; c++) {
        const int round = c / G, wave = (c % G) / gw;
        int mx = 0;
        for (int c2 = round * G + wave * gw; c2 < std::min(nc, round * G + (wave + 1) * gw); c2++)
          mx = std::max(mx, rows[c2]);
        wrows[c] = mx;
      }
      const int boff = (int)R.blob.size();
      std::vector<int> recs, ops;
      int xo = 0;
      for (int c = 0; c < nc; c++) {
        const int *r = &A.crec[8 * cols[c]];
        const int k = r[0], c0 = r[1], len = r[2], dl = r[3];
        xstart[k] = xtot + xo;
        const int lbase = (mode != 2 && dl + 1 < len) ? lslot[c0 + dl + 1] : 0;
        recs.insert(recs.end(), {xo, len, dl, (int)ops.size() / (2 * g), rows[c], wrows[c], lbase, 0});
        std::vector<std::vector<int>> prow;
        pack_rows(cols[c], g, &prow);
        int nrow = (int)prow.size();
        for (const auto &row : prow)
          for (int jj = 0; jj < g; jj++) {
            if (jj < (int)row.size()) {
              const int o = row[jj];
              const int li = lslot[A.op_sl[o]];
              if (li < 0) { R = ResPlan{}; return false; }   // multiplier outside the kept L entries
              ops.insert(ops.end(), {A.op_loc[o], li});
            } else {
              ops.insert(ops.end(), {len, R.nL});   // no-op on the spare slot
            }
          }
        if (nrow == 0) {   // a column without updates still has one (no-op) row
          for (int jj = 0; jj < g; jj++) ops.insert(ops.end(), {len, R.nL});
          nrow = 1;
        }
        if (nrow != rows[c]) { R = ResPlan{}; return false; }
        for (int t = 0; t < len; t++) {
          const int p = c0 + t;
          if (mode != 2 && t > dl && lslot[p] != lbase + (t - dl - 1)) { R = ResPlan{}; return false; }
          int o = A.lu2s[p];
          if (mode == 2 && t >= dl) o = A.dmap[p - qb] >= 0 ? -2 - A.dmap[p - qb] : -1;
          R.omap.push_back(o);
        }
        R.omap.push_back(-1);   // spare slot
        xo += len + 1;
      }
      const int oo = 8 * nc;
      R.blob.insert(R.blob.end(), recs.begin(), recs.end());
      R.blob.insert(R.blob.end(), ops.begin(), ops.end());
      R.blob.resize(al4(R.blob.size()), 0);
      const int blen2 = (int)R.blob.size() - boff;
      const int xl = (int)al4(xo);
      R.omap.resize(xtot + xl, -1);
      R.hdr.insert(R.hdr.end(), {mode, nc, boff, blen2, xtot, xl, oo, 0, 0, g, 0, 0});
      R.maxb = std::max(R.maxb, blen2);
      R.maxx = std::max(R.maxx, xl);
      xtot += xl;
    }
  }
  R.blob.resize(R.blob.size() + 4, 0);   // vector loads of the last step may read past it
  R.xtot = xtot + 4;
  R.omap.resize(R.xtot, -1);
  // A entry -> step-ordered position
  std::vector<int> colof(A.nnz_lu);
  for (int k = 0; k < n; k++)
    for (int p = A.lu_cp[k]; p < A.lu_cp[k + 1]; p++) colof[p] = k;
  R.a2x.resize(A.nnz_a);
  for (int e = 0; e < A.nnz_a; e++) {
    const int p = A.a2lu[e], k = colof[p];
    if (xstart[k] < 0) { R = ResPlan{}; return false; }
    R.a2x[e] = xstart[k] + (p - A.lu_cp[k]);
  }
  return true;
}

bool solve_plan(const SpluAnalysis::SolveSide &S, bool diag, int dg_base, int bcap, int vcap, int dcap,
                SolvePlan &P) {
  P = SolvePlan{};
  const int nlev = (int)S.lvl.size() / 4;
  auto al4 = [](int v) { return (v + 3) & ~3; };
  auto half = [](int v) { return (v + 1) / 2; };   // dwords of v 16-bit columns
  for (int c : S.cols)
    if (c < 0 || c > 65535) return false;
  // (sub-)levels as chunk ranges [c0, c1) of a level
  struct Sub { int l, c0, c1; };
  std::vector<Sub> subs;
  auto chunk_vals = [&](int c) { return 64 * S.chunk[2 * c + 1]; };
  for (int l = 0; l < nlev; l++) {
    const int T = S.lvl[4 * l], cb = S.lvl[4 * l + 3], nch = (T + 63) / 64;
    int c = cb;
    while (c < cb + nch) {
      int c1 = c, v = 0;
      // whole chunks while the values, columns and rows fit one step
      while (c1 < cb + nch && v + chunk_vals(c1) <= vcap - 8 && (!diag || 64 * (c1 - c + 1) <= dcap) &&
             36 * ((c1 - c + 8) / 8 + 1) + half(v + chunk_vals(c1)) + 64 * (c1 - c + 1) <= bcap - 16) {
        v += chunk_vals(c1);
        c1++;
      }
      if (c1 == c) return false;   // one chunk does not fit
      subs.push_back({l, c, c1});
      c = c1;
    }
  }
  // steps: consecutive sub-levels while their data fit
  size_t i = 0;
  while (i < subs.size()) {
    size_t j = i;
    int nch = 0, nv = 0, nl = 0;
    while (j < subs.size()) {
      const int ch = subs[j].c1 - subs[j].c0;
      int v = 0;
      for (int c = subs[j].c0; c < subs[j].c1; c++) v += chunk_vals(c);
      const int lanes = 64 * (nch + ch);   // rows: at most 64 per chunk
      const int blen = 36 * (nl + (nch + ch + 7) / 8 + 1) + al4(half(nv + v)) + al4(lanes);
      if (blen > bcap || al4(nv + v) > vcap || (diag && al4(lanes) > dcap)) break;
      nch += ch;
      nv += v;
      nl++;
      j++;
    }
    if (j == i) return false;
    // the step's ranges
    const int cfirst = subs[i].c0, clast = subs[j - 1].c1;   // chunks [cfirst, clast)
    const int E0 = S.chunk[2 * cfirst];
    const int E1 = E0 + nv;
    auto lane_of = [&](const Sub &sb, int c) {   // first lane of chunk c of a sub-level
      const int l = sb.l;
      return S.lvl[4 * l + 2] + 64 * (c - S.lvl[4 * l + 3]);
    };
    const int R0 = lane_of(subs[i], cfirst);
    const Sub &sl = subs[j - 1];
    const int Tl = S.lvl[4 * sl.l];
    const int R1 = std::min(lane_of(sl, sl.c1 - 1) + 64, S.lvl[4 * sl.l + 2] + Tl);
    const int boff = (int)P.blob.size();
    std::vector<int> lv, cols, rows;
    int nlp = 0;
    for (size_t k = i; k < j; k++) {
      const Sub &sb = subs[k];
      const int l = sb.l, T = S.lvl[4 * l], g = S.lvl[4 * l + 1], cb = S.lvl[4 * l + 3];
      // a sub-level of one chunk runs on wavefront 0 alone: when the next
      // sub-level of the step is one chunk too, wavefront 0 goes straight on
      // (a wavefront's LDS accesses complete in order), and the barrier
      // comes after the last of the run
      const bool single = sb.c1 - sb.c0 == 1;
      const bool chain = single && k + 1 < j && subs[k + 1].c1 - subs[k + 1].c0 == 1;
      // passes of 8 chunks (one per wavefront); a barrier after the last
      for (int c = sb.c0; c < sb.c1; c += 8, nlp++) {
        lv.insert(lv.end(), {g, (c + 8 >= sb.c1 && !chain) ? 1 : 0, S.lflag.empty() ? 0 : S.lflag[l], 0});
        for (int w = 0; w < 8; w++) {
          const int cw = c + w;
          if (cw < sb.c1) {
            const int u0 = 64 * (cw - cb);   // first lane of the chunk in its level
            lv.insert(lv.end(), {S.chunk[2 * cw] - E0, S.chunk[2 * cw + 1], S.lvl[4 * l + 2] + u0 - R0,
                                 std::min(64, T - u0)});
          } else {
            lv.insert(lv.end(), {0, 0, 0, 0});
          }
        }
      }
    }
    if (S.chunk[2 * (clast - 1)] + chunk_vals(clast - 1) != E1) return false;   // values contiguous
    cols.assign(half(E1 - E0), 0);   // 16-bit columns, two per dword (low half first)
    for (int e = E0; e < E1; e++) cols[(e - E0) / 2] |= S.cols[e] << (16 * ((e - E0) & 1));
    rows.assign(S.rows.begin() + R0, S.rows.begin() + R1);
    auto put = [&](std::vector<int> &v) {
      v.resize(al4((int)v.size()), 0);
      P.blob.insert(P.blob.end(), v.begin(), v.end());
    };
    put(lv);
    const int o_ch = 0;
    const int o_cols = (int)P.blob.size() - boff;
    put(cols);
    const int o_rows = (int)P.blob.size() - boff;
    put(rows);
    const int blen = (int)P.blob.size() - boff;
    const int vlen = al4(E1 - E0), dlen = diag ? al4(R1 - R0) : 0;
    P.hdr.insert(P.hdr.end(), {nlp, boff, blen, E0, vlen, dg_base + R0, dlen, o_ch, o_cols, o_rows, E1 - E0,
                               diag ? R1 - R0 : 0});
    P.maxb = std::max(P.maxb, blen);
    P.maxv = std::max(P.maxv, vlen + dlen);
    i = j;
  }
  P.blob.resize(P.blob.size() + 4, 0);
  return true;
}

}  // namespace gp

// gp_splu.cpp -- symbolic analysis for the batched sparse LU (see gp_splu.hpp).
#include "gp_splu.hpp"

#include <algorithm>
#include <cmath>
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
                         double thr) {
  A = SpluAnalysis();
  A.n = n;
  A.nnz_a = cp[n];
  A.q = min_degree_order(n, cp, ri);
  if ((int)A.q.size() != n) return "ordering failed";
  A.qinv.assign(n, 0);
  for (int k = 0; k < n; k++) A.qinv[A.q[k]] = k;

  // ---- numeric left-looking LU with threshold partial pivoting ----------
  std::vector<int> pinv(n, -1), prow(n, -1);
  std::vector<std::vector<std::pair<int, double>>> Lcol(n), Ucol(n);  // L: original rows; U: pivot rows
  std::vector<double> x(n, 0.0);
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
    int ipiv = -1;
    double amax = -1.0;
    for (int p = top; p < n; p++) {
      int i = xi[p];
      if (pinv[i] < 0) {
        double t = std::fabs(x[i]);
        if (t > amax) { amax = t; ipiv = i; }
      } else {
        Ucol[k].push_back({pinv[i], x[i]});
      }
    }
    if (ipiv < 0 || !(amax > 0.0)) return "structurally or numerically singular matrix at pivot " + std::to_string(k);
    if (pinv[col] < 0 && std::fabs(x[col]) >= thr * amax) ipiv = col;
    double piv = x[ipiv];
    pinv[ipiv] = k;
    prow[k] = ipiv;
    for (int p = top; p < n; p++) {
      int i = xi[p];
      if (pinv[i] < 0) Lcol[k].push_back({i, x[i] / piv});
      x[i] = 0.0;
    }
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

// Cost units below are one dependent update (an LDS read-modify-write, or a
// gathered load followed by one).  Thread mode pays every op of its longest
// column; wavefront mode pays one unit per step plus the column's staging.
void splu_schedule(SpluAnalysis &A, int threads) {
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
  A.fac_mode.assign(nl, 0);
  A.t_cols_max = A.w_col_max = A.w_ops_max = 0;
  std::vector<long long> wsum(W);
  for (int l = 0; l < nl; l++) {
    const int c0 = A.fac_lp[l], c1 = A.fac_lp[l + 1];
    int maxops = 0;
    std::fill(wsum.begin(), wsum.end(), 0);
    for (int i = c0; i < c1; i++) {
      const int k = A.fac_cols[i];
      maxops = std::max(maxops, A.op_cp[k + 1] - A.op_cp[k]);
      wsum[(i - c0) % W] += (A.st_cp[k + 1] - A.st_cp[k]) + 6;
    }
    const long long tcost = (long long)((c1 - c0 + threads - 1) / threads) * (maxops + 4);
    const long long wcost = *std::max_element(wsum.begin(), wsum.end());
    A.fac_mode[l] = wcost < tcost ? 1 : 0;
    for (int i = c0; i < c1; i++) {
      const int k = A.fac_cols[i];
      const int len = A.lu_cp[k + 1] - A.lu_cp[k], nops = A.op_cp[k + 1] - A.op_cp[k];
      if (A.fac_mode[l]) {
        A.w_col_max = std::max(A.w_col_max, len);
        A.w_ops_max = std::max(A.w_ops_max, nops);
      } else {
        A.t_cols_max = std::max(A.t_cols_max, len);
      }
    }
  }
  // solve levels: lanes per row g minimising passes * (ceil(len/g) + log2 g)
  auto pick = [&](const std::vector<int> &lp, const std::vector<int> &rows, const std::vector<int> &rp,
                  std::vector<int> &g) {
    const int nlev = (int)lp.size() - 1;
    g.assign(nlev, 1);
    for (int l = 0; l < nlev; l++) {
      int maxlen = 0;
      for (int i = lp[l]; i < lp[l + 1]; i++) maxlen = std::max(maxlen, rp[rows[i] + 1] - rp[rows[i]]);
      const int nr = lp[l + 1] - lp[l];
      long long best = -1;
      for (int gg = 1, lg = 0; gg <= 64; gg *= 2, lg++) {
        const long long passes = ((long long)nr * gg + threads - 1) / threads;
        const long long c = passes * ((maxlen + gg - 1) / gg + lg + 2);
        if (best < 0 || c < best) { best = c; g[l] = gg; }
      }
    }
  };
  pick(A.fw_lp, A.fw_rows, A.lrow_p, A.fw_g);
  pick(A.bw_lp, A.bw_rows, A.urow_p, A.bw_g);
}

}  // namespace gp

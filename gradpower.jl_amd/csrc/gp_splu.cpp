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
  return "";
}

}  // namespace gp

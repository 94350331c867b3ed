// Host symbolic analysis (see symbolic.hpp for the reference functions each
// routine restates).  Every routine is written for this engine; where the
// reference's tie-breaking decides the output (postorder child order,
// supernode relaxation, coarse level partitioning) the same decision rules
// are followed so that the supernode partition and level schedule match
// bit for bit.
#include "symbolic.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <stdexcept>

namespace lblg {

// ---------------------------------------------------------------------------
// permuted transposes (ptranspose, src/common/Transpose.cpp:399-570)
// ---------------------------------------------------------------------------
void upper_of_permuted(const CscView& A, const int* pinv, const double* Ax,
                       std::vector<int>& Up, std::vector<int>& Ui, std::vector<double>* Ux) {
  const int n = A.n;
  Up.assign(n + 1, 0);
  for (int j = 0; j < n; ++j) {
    const int fj = pinv ? pinv[j] : j;
    for (int q = A.p[j]; q < A.p[j + 1]; ++q) {
      const int i = A.i[q];
      if (i < j) continue;  // lower part only (stype = -1)
      const int fi = pinv ? pinv[i] : i;
      Up[std::max(fi, fj) + 1]++;
    }
  }
  for (int j = 0; j < n; ++j) Up[j + 1] += Up[j];
  Ui.resize(Up[n]);
  if (Ux) Ux->resize(Up[n]);
  std::vector<int> fill(Up.begin(), Up.end() - 1);
  for (int j = 0; j < n; ++j) {
    const int fj = pinv ? pinv[j] : j;
    for (int q = A.p[j]; q < A.p[j + 1]; ++q) {
      const int i = A.i[q];
      if (i < j) continue;
      const int fi = pinv ? pinv[i] : i;
      const int c = std::max(fi, fj), r = std::min(fi, fj);
      const int dst = fill[c]++;
      Ui[dst] = r;
      if (Ux) (*Ux)[dst] = Ax[q];
    }
  }
}

void transpose_pattern(int n, const int* Ap, const int* Ai, const double* Ax,
                       std::vector<int>& Tp, std::vector<int>& Ti, std::vector<double>* Tx) {
  Tp.assign(n + 1, 0);
  for (int q = 0; q < Ap[n]; ++q) Tp[Ai[q] + 1]++;
  for (int j = 0; j < n; ++j) Tp[j + 1] += Tp[j];
  Ti.resize(Ap[n]);
  if (Tx) Tx->resize(Ap[n]);
  std::vector<int> fill(Tp.begin(), Tp.end() - 1);
  for (int j = 0; j < n; ++j)
    for (int q = Ap[j]; q < Ap[j + 1]; ++q) {
      const int dst = fill[Ai[q]]++;
      Ti[dst] = j;  // source columns visited in order: rows come out sorted
      if (Tx) (*Tx)[dst] = Ax[q];
    }
}

// ---------------------------------------------------------------------------
// elimination tree of triu(S) with path compression (Liu's algorithm)
// ---------------------------------------------------------------------------
void etree_upper(int n, const int* Up, const int* Ui, std::vector<int>& parent) {
  parent.assign(n, -1);
  std::vector<int> anc(n, -1);
  for (int j = 0; j < n; ++j) {
    for (int q = Up[j]; q < Up[j + 1]; ++q) {
      int k = Ui[q];
      if (k >= j) continue;
      while (true) {
        const int a = anc[k];
        if (a == j) break;
        anc[k] = j;
        if (a == -1) {
          parent[k] = j;
          break;
        }
        k = a;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// postorder with the reference's child ordering rules
// ---------------------------------------------------------------------------
int postorder(int n, const int* parent, const int* weight, std::vector<int>& post) {
  std::vector<int> head(n + 1, -1), next(n, -1), stack(n > 0 ? n : 1);
  if (!weight) {
    for (int j = n - 1; j >= 0; --j) {
      const int p = parent[j];
      if (p >= 0 && p < n) {
        next[j] = head[p];
        head[p] = j;
      }
    }
  } else {
    // bucket by clipped weight, then push buckets from heaviest to lightest so
    // each child list ends up lightest-first, ties by increasing index
    std::vector<int> bucket(n > 0 ? n : 1, -1);
    for (int j = 0; j < n; ++j) {
      const int p = parent[j];
      if (p >= 0 && p < n) {
        int w = std::min(std::max(0, weight[j]), n - 1);
        next[j] = bucket[w];
        bucket[w] = j;
      }
    }
    for (int w = n - 1; w >= 0; --w) {
      for (int j = bucket[w]; j != -1;) {
        const int nj = next[j];
        const int p = parent[j];
        next[j] = head[p];
        head[p] = j;
        j = nj;
      }
    }
  }
  post.assign(n, 0);
  int k = 0;
  for (int r = 0; r < n; ++r) {
    if (parent[r] != -1) continue;
    int top = 0;
    stack[0] = r;
    while (top >= 0) {
      const int p = stack[top];
      const int c = head[p];
      if (c == -1) {
        --top;
        post[k++] = p;
      } else {
        head[p] = next[c];
        stack[++top] = c;
      }
    }
  }
  return k;
}

// ---------------------------------------------------------------------------
// column counts of L (Gilbert-Ng-Peyton row-subtree leaf counting)
// ---------------------------------------------------------------------------
void column_counts(int n, const int* Fp, const int* Fi, const int* parent, const int* post,
                   std::vector<int>& colcount) {
  std::vector<int> first(n, -1), delta(n, 0), prevleaf(n, -1), prevnbr(n, -1), uf(n);
  for (int k = 0; k < n; ++k) {
    const int j = post[k];
    delta[j] = (first[j] == -1) ? 1 : 0;  // leaf of the etree
    for (int r = j; r != -1 && first[r] == -1; r = parent[r]) first[r] = k;
  }
  for (int i = 0; i < n; ++i) uf[i] = i;
  for (int k = 0; k < n; ++k) {
    const int j = post[k];
    if (parent[j] != -1) delta[parent[j]]--;
    prevnbr[j] = k;
    for (int q = Fp[j]; q < Fp[j + 1]; ++q) {
      const int i = Fi[q];
      if (i <= j) continue;
      if (first[j] > prevnbr[i]) {
        delta[j]++;
        const int pl = prevleaf[i];
        if (pl != -1) {
          int root = pl;
          while (root != uf[root]) root = uf[root];
          for (int s = pl; s != root;) {
            const int nx = uf[s];
            uf[s] = root;
            s = nx;
          }
          delta[root]--;
        }
        prevleaf[i] = j;
      }
      prevnbr[i] = k;
    }
    if (parent[j] != -1) uf[j] = parent[j];
  }
  colcount = delta;
  for (int j = 0; j < n; ++j)
    if (parent[j] != -1) colcount[parent[j]] += colcount[j];
}

// ---------------------------------------------------------------------------
// relaxed supernode partition and supernodal row structure
// (super_node_detection, src/symbolic/supernode_detection.cpp:43-746)
// ---------------------------------------------------------------------------
static void supernodes(int n, const int* Sp, const int* Si, const int* parent,
                       const int* colcount, const int* extra, const SymbolicOptions& opt,
                       Symbolic& L) {
  std::vector<int> nchild(n, 0);
  for (int j = 0; j < n; ++j)
    if (parent[j] != -1) nchild[parent[j]]++;

  // fundamental supernodes; constraint ("extra") columns are kept alone and
  // flagged non-mergeable
  std::vector<int> fsuper;
  std::vector<char> nomerge;
  fsuper.reserve(n + 1);
  if (n > 0) {
    fsuper.push_back(0);
    nomerge.push_back(0);
  }
  for (int j = 1; j < n; ++j) {
    const bool ex = extra[j] == 1 || extra[j - 1] == 1;
    if (parent[j - 1] != j || colcount[j - 1] != colcount[j] + 1 || nchild[j] > 1 || ex) {
      fsuper.push_back(j);
      nomerge.push_back(ex ? 1 : 0);
    }
  }
  const int nf = (int)fsuper.size();
  fsuper.push_back(n);

  std::vector<int> fmap(n), fpar(nf);
  for (int s = 0; s < nf; ++s)
    for (int k = fsuper[s]; k < fsuper[s + 1]; ++k) fmap[k] = s;
  for (int s = 0; s < nf; ++s) {
    const int par = parent[fsuper[s + 1] - 1];
    fpar[s] = par == -1 ? -1 : fmap[par];
  }

  // relaxed amalgamation of s into its (current) parent s+1
  std::vector<int> merged(nf, -1), nscol(nf), snz(nf);
  std::vector<int64_t> zeros(nf, 0);
  for (int s = 0; s < nf; ++s) {
    nscol[s] = fsuper[s + 1] - fsuper[s];
    snz[s] = colcount[fsuper[s]];
  }
  const double size_cap = (double)(INT_MAX / sizeof(double));
  for (int s = nf - 2; s >= 0; --s) {
    if (fpar[s] == -1) continue;
    if (nomerge[s] || nomerge[s + 1]) continue;
    int cur = fpar[s];
    while (merged[cur] != -1) cur = merged[cur];
    for (int t = fpar[s]; merged[t] != -1;) {
      const int nx = merged[t];
      merged[t] = cur;
      t = nx;
    }
    if (cur != s + 1) continue;
    const int c0 = nscol[s], c1 = nscol[s + 1], ns = c0 + c1;
    int64_t totzeros = zeros[s + 1];
    const double lnz1 = (double)snz[s + 1];
    bool merge;
    if (ns <= opt.nrelax[0]) {
      merge = true;
    } else {
      const double xnew = (double)c0 * (lnz1 + c0 - (double)snz[s]);
      const int64_t inew = (int64_t)c0 * ((int64_t)snz[s + 1] + c0 - snz[s]);
      if (xnew == 0) {
        merge = true;
      } else {
        const double xtotz = (double)totzeros + xnew;
        const double xns = (double)ns;
        const double xtots = xns * (xns + 1) / 2 + xns * (lnz1 - c1);
        const double z = xtotz / xtots;
        totzeros += inew;
        merge = ((ns <= opt.nrelax[1] && z < opt.zrelax[0]) ||
                 (ns <= opt.nrelax[2] && z < opt.zrelax[1]) || (z < opt.zrelax[2])) &&
                (xtots < size_cap);
      }
    }
    if (merge) {
      zeros[s] = totzeros;
      merged[s + 1] = s;
      snz[s] = c0 + snz[s + 1];
      nscol[s] += nscol[s + 1];
    }
  }

  std::vector<int> rsnz;
  L.super.clear();
  for (int s = 0; s < nf; ++s)
    if (merged[s] == -1) {
      L.super.push_back(fsuper[s]);
      rsnz.push_back(snz[s]);
    }
  const int ns = (int)L.super.size();
  L.super.push_back(n);
  L.nsuper = ns;
  L.col2sup.assign(n, 0);
  for (int s = 0; s < ns; ++s)
    for (int k = L.super[s]; k < L.super[s + 1]; ++k) L.col2sup[k] = s;
  L.sparent.assign(ns, -1);
  for (int s = 0; s < ns; ++s) {
    const int par = parent[L.super[s + 1] - 1];
    L.sparent[s] = par == -1 ? -1 : L.col2sup[par];
  }
  int64_t ssize = 0, xsize = 0;
  for (int s = 0; s < ns; ++s) {
    ssize += rsnz[s];
    xsize += (int64_t)(L.super[s + 1] - L.super[s]) * rsnz[s];
  }
  L.ssize = std::max<int64_t>(1, ssize);
  L.xsize = std::max<int64_t>(1, xsize);
  L.pi.assign(ns + 1, 0);
  for (int s = 0; s < ns; ++s) L.pi[s + 1] = L.pi[s] + rsnz[s];

  // row structure: union of row subtrees walked on the supernodal tree
  L.s.assign(std::max<int64_t>(1, L.pi[ns]), 0);
  std::vector<int64_t> fillp(L.pi.begin(), L.pi.end());
  std::vector<int> stamp(std::max(ns, 1), -1);
  for (int s = 0; s < ns; ++s) {
    const int k1 = L.super[s], k2 = L.super[s + 1];
    for (int k = k1; k < k2; ++k) L.s[fillp[s]++] = k;
    for (int k = k1; k < k2; ++k) {
      stamp[s] = k;
      for (int q = Sp[k]; q < Sp[k + 1]; ++q) {
        const int i = Si[q];
        if (i >= k1) continue;
        for (int t = L.col2sup[i]; stamp[t] != k; t = L.sparent[t]) {
          L.s[fillp[t]++] = k;
          stamp[t] = k;
        }
      }
    }
  }
  // largest update block and off-diagonal height (maxcsize / maxesize)
  int64_t maxc = 1, maxe = 1;
  for (int d = 0; d < ns; ++d) {
    const int w = L.super[d + 1] - L.super[d];
    int64_t q = L.pi[d] + w, qlast = q;
    const int64_t qend = L.pi[d + 1];
    maxe = std::max(maxe, qend - q);
    int slast = (q == qend) ? -1 : L.col2sup[L.s[q]];
    for (; q <= qend; ++q) {
      const int t = (q == qend) ? -1 : L.col2sup[L.s[q]];
      if (t != slast) {
        maxc = std::max(maxc, (qend - qlast) * (q - qlast));
        qlast = q;
        slast = t;
      }
    }
  }
  L.maxcsize = maxc;
  L.maxesize = maxe;
}

// ---------------------------------------------------------------------------
// coarse level set (getCoarseLevelSet_6 and its helpers getLevelSet,
// heightPartitioning (sw=false), postOrderSpliting, worstFitBinPack)
// ---------------------------------------------------------------------------
namespace {

struct Part {
  double cost;
  int idx;
};

// one leaf-first sweep of a cluster forest: builds the initial parts
void split_parts(int n, const std::vector<int>& tree, const double* cost,
                 const std::vector<int>& cptr, const std::vector<int>& cno, std::vector<int>& nchild,
                 std::vector<double>& part_cost, std::vector<std::vector<int>>& parts) {
  std::vector<char> visited(n, 0), onpath(n, 0);
  std::vector<int> work;  // front of the work list is element 0
  parts.clear();
  parts.emplace_back();
  int cur = 0;
  for (int start = 0; start < n; ++start) {
    if (tree[start] == -2 || visited[start]) continue;
    work.assign(1, start);
    while (!work.empty()) {
      const int k = work.front();
      if (nchild[k] == 0) {
        work.erase(work.begin());
        parts[cur].push_back(k);
        part_cost[cur] += cost[k];
        visited[k] = 1;
        const int par = tree[k];
        if (par >= 0) {
          nchild[par]--;
          if (!visited[par]) {
            if (!onpath[par]) work.insert(work.begin(), par);
            onpath[par] = 1;
          }
        } else {
          break;
        }
      } else {
        for (int q = cptr[k]; q < cptr[k + 1]; ++q) {
          const int c = cno[q];
          if (visited[c]) continue;
          if (nchild[c] == 0) {
            if (!visited[k]) {
              int t = k;
              if (!work.empty()) {
                while (t >= 0 && t != work.back()) {
                  onpath[t] = 1;
                  t = tree[t];
                }
              }
              if (t >= 0) onpath[t] = 1;
            }
            parts[cur].push_back(c);
            part_cost[cur] += cost[c];
            visited[c] = 1;
            nchild[tree[c]]--;
          } else {
            work.insert(work.begin(), c);
          }
        }
      }
    }
    ++cur;
    parts.emplace_back();
  }
  parts.pop_back();
}

}  // namespace

void coarse_level_set(int n, const int* tree, const double* node_cost, int inner_parts,
                      int min_level_dist, int div_rate, LevelSchedule& out) {
  out.level_ptr.assign(n + 1, 0);
  out.par_ptr.assign(n + 1, 0);
  out.par_set.assign(n, 0);
  std::vector<int> nchild(n, 0);
  for (int v = 0; v < n; ++v)
    if (tree[v] >= 0) nchild[tree[v]]++;
  // tree height = deepest leaf
  int height = 0;
  for (int v = 0; v < n; ++v) {
    if (nchild[v] != 0) continue;
    int d = 0;
    for (int u = v; tree[u] >= 0; u = tree[u]) ++d;
    height = std::max(height, d);
  }
  if (height == 0 && n == 1) {
    out.n_level = 1;
    out.level_ptr = {0, 1};
    out.par_ptr = {0, 1};
    out.par_set = {0};
    return;
  }
  // leaf-first levels: level(v) = 1 + max level(children); members by index
  std::vector<int> lev(n, 0);
  int nlev = 0;
  for (int v = 0; v < n; ++v) {  // children precede parents (postordered tree)
    nlev = std::max(nlev, lev[v] + 1);
    if (tree[v] >= 0) lev[tree[v]] = std::max(lev[tree[v]], lev[v] + 1);
  }
  std::vector<int> lptr(nlev + 1, 0), lset(n);
  for (int v = 0; v < n; ++v) lptr[lev[v] + 1]++;
  for (int l = 0; l < nlev; ++l) lptr[l + 1] += lptr[l];
  {
    std::vector<int> f(lptr.begin(), lptr.end() - 1);
    for (int v = 0; v < n; ++v) lset[f[lev[v]]++] = v;
  }
  std::vector<double> lcost(nlev, 0.0);
  for (int v = 0; v < n; ++v) lcost[lev[v]] += node_cost[v];

  // height partitioning (sw = false)
  std::vector<int> p2l(nlev + 2, 0), psize;
  int ncl = 0;
  if (nlev <= 2) {
    p2l[0] = 0;
    p2l[1] = nlev;
    psize.push_back(1);
    ncl = 1;
  } else {
    int parts = inner_parts, cut = 0;
    while (parts > 1) {
      while ((lptr[height - cut] - lptr[height - cut - 1]) <= parts && cut < nlev) {
        ++cut;
        if (height - cut - 1 < 0 || height - cut - 1 >= nlev) break;
      }
      psize.push_back(parts);
      const int t = height - cut - min_level_dist;
      if (t > p2l[ncl] && t < nlev) p2l[++ncl] = t;
      parts /= div_rate;
      cut = 0;
    }
    p2l[++ncl] = height + 1;
    psize.push_back(1);
  }

  std::vector<int> ctree(n, -2), cptr(n + 1), cno(n), cnch(n);
  std::vector<double> pcost(n), bins(n, 0.0);  // bins persist across clusters
  std::vector<std::vector<int>> parts, merged;
  int npart = 0;
  out.level_ptr[0] = 0;
  for (int l = 0; l < ncl; ++l) {
    int cl_cost = 0;  // int accumulation (truncating), as the reference
    const int ip = psize[l];
    for (int i = p2l[l]; i < p2l[l + 1]; ++i) {
      cl_cost = (int)(cl_cost + lcost[i]);
      for (int q = lptr[i]; q < lptr[i + 1]; ++q) {
        const int v = lset[q];
        const int par = tree[v];
        const int lp = par >= 0 ? lev[par] : height + 1;
        ctree[v] = (lp < p2l[l + 1] && l + 1 <= ncl && lp >= p2l[l]) ? par : -1;
      }
    }
    std::fill(cnch.begin(), cnch.end(), 0);
    for (int v = 0; v < n; ++v)
      if (ctree[v] >= 0) cnch[ctree[v]]++;
    cptr[0] = 0;
    for (int v = 0; v < n; ++v) cptr[v + 1] = cptr[v] + cnch[v];
    {
      std::vector<int> f(cptr.begin(), cptr.end() - 1);
      for (int v = 0; v < n; ++v)
        if (ctree[v] >= 0) cno[f[ctree[v]]++] = v;
    }
    std::fill(pcost.begin(), pcost.end(), 0.0);
    split_parts(n, ctree, node_cost, cptr, cno, cnch, pcost, parts);
    int nout;
    if ((int)parts.size() > ip) {
      std::vector<Part> order(parts.size());
      for (size_t i = 0; i < parts.size(); ++i) order[i] = {pcost[i], (int)i};
      std::sort(order.begin(), order.end(), [](const Part& a, const Part& b) { return a.cost > b.cost; });
      merged.assign(ip, {});
      for (const Part& pt : order) {
        double mn = (double)INT_MAX;
        int mb = 0;
        for (int b = 0; b < ip; ++b)
          if (bins[b] < mn) {
            mn = bins[b];
            mb = b;
          }
        bins[mb] += pt.cost;
        merged[mb].insert(merged[mb].end(), parts[pt.idx].begin(), parts[pt.idx].end());
      }
      nout = ip;
    } else {
      merged = parts;
      nout = (int)parts.size();
    }
    std::fill(ctree.begin(), ctree.end(), -2);
    out.level_ptr[l + 1] = out.level_ptr[l] + nout;
    for (int i = 0; i < nout; ++i) {
      int e = 0;
      for (int v : merged[i]) out.par_set[out.par_ptr[npart] + e++] = v;
      out.par_ptr[npart + 1] = out.par_ptr[npart] + e;
      ++npart;
    }
  }
  out.n_level = ncl;
  out.level_ptr.resize(ncl + 1);
  out.par_ptr.resize(npart + 1);
}

// ---------------------------------------------------------------------------
// full analysis (symbolic_analysis_lin_solve)
// ---------------------------------------------------------------------------
bool analyze(const CscView& A, const int* order, const int* extra_in, const SymbolicOptions& opt,
             Symbolic& L) {
  const int n = A.n;
  L = Symbolic();
  L.n = n;
  if (n <= 0) return false;
  L.perm.assign(order, order + n);
  std::vector<int> pinv(n, -1);
  for (int k = 0; k < n; ++k) {
    if (L.perm[k] < 0 || L.perm[k] >= n || pinv[L.perm[k]] != -1)
      throw std::invalid_argument("ordering is not a permutation");
    pinv[L.perm[k]] = k;
  }
  // etree, postorder and column counts of P A P^T
  std::vector<int> Up, Ui, Fp, Fi;
  upper_of_permuted(A, pinv.data(), nullptr, Up, Ui, nullptr);
  transpose_pattern(n, Up.data(), Ui.data(), nullptr, Fp, Fi, nullptr);
  std::vector<int> parent, post, cc;
  etree_upper(n, Up.data(), Ui.data(), parent);
  postorder(n, parent.data(), nullptr, post);
  column_counts(n, Fp.data(), Fi.data(), parent.data(), post.data(), cc);

  // weighted postorder, combined into the ordering
  std::vector<int> wpost;
  if (postorder(n, parent.data(), cc.data(), wpost) == n) {
    std::vector<int> w(n), ipost(n);
    for (int k = 0; k < n; ++k) w[k] = L.perm[wpost[k]];
    L.perm = w;
    for (int k = 0; k < n; ++k) w[k] = cc[wpost[k]];
    cc = w;
    for (int k = 0; k < n; ++k) ipost[wpost[k]] = k;
    for (int k = 0; k < n; ++k) {
      const int op = parent[wpost[k]];
      w[k] = op == -1 ? -1 : ipost[op];
    }
    parent = w;
  }
  L.iperm.assign(n, 0);
  for (int k = 0; k < n; ++k) L.iperm[L.perm[k]] = k;
  L.extra_cols.assign(n, 0);
  if (extra_in)
    for (int k = 0; k < n; ++k) L.extra_cols[k] = extra_in[L.perm[k]];
  L.parent = parent;
  L.colcount = cc;
  for (int j = 0; j < n; ++j) {
    L.lnz += (double)cc[j];
    L.fl += (double)cc[j] * (double)cc[j];
  }

  // supernodes on the final ordering
  upper_of_permuted(A, L.iperm.data(), nullptr, Up, Ui, nullptr);
  supernodes(n, Up.data(), Ui.data(), parent.data(), cc.data(), L.extra_cols.data(), opt, L);

  // BCSC pointers (symbolic_phase.cpp:680-711)
  L.max_sup_wid = 0;
  L.max_col = 0;
  for (int s = 0; s < L.nsuper; ++s) L.max_sup_wid = std::max(L.max_sup_wid, L.super[s + 1] - L.super[s]);
  L.p.assign(n + 1, 0);
  L.i_ptr.assign(n + 1, 0);
  int64_t len = 0;
  for (int s = 0; s < L.nsuper; ++s) {
    const int c0 = L.super[s], c1 = L.super[s + 1];
    len = L.pi[s + 1] - L.pi[s];
    L.max_col = std::max<int64_t>(L.max_col, len);
    for (int i = c0 + 1; i <= c1; ++i) {
      L.i_ptr[i - 1] = L.pi[s];
      L.p[i] = L.p[c0] + (int64_t)(i - c0) * len;
    }
  }
  L.i_ptr[n] = L.pi[L.nsuper];
  L.p[n] = L.p[n - 1] + len;

  // node cost: the reference's descendant walk (computeCostperBlock over the
  // lower-triangular pattern) never leaves the supernode, so each cost is
  // width * rows (computeCostColFact4)
  L.node_cost.assign(L.nsuper, 0.0);
  for (int s = 0; s < L.nsuper; ++s) {
    const double w = L.super[s + 1] - L.super[s];
    const double r = (double)(L.pi[s + 1] - L.pi[s]);
    L.node_cost[s] = std::max(0.0, w * r);
  }
  coarse_level_set(L.nsuper, L.sparent.data(), L.node_cost.data(), opt.cost_param, opt.level_param,
                   opt.final_seq_node, L.lbc);
  return true;
}

}  // namespace lblg

// Fill-reducing nested-dissection ordering.
//
// The reference orders the KKT graph with METIS_NodeND
// (src/symbolic/symbolic_phase.cpp:448-527); METIS is not part of this
// engine, so this file provides a multilevel nested dissection of its own:
//   * connected components are ordered independently;
//   * a component is bisected by heavy-edge-matching coarsening, greedy graph
//     growing on the coarsest graph, and boundary refinement during projection;
//   * the edge cut is turned into a vertex separator by a greedy cover that
//     prefers vertices incident to many cut edges (constraint rows of a KKT
//     matrix touch many cut edges and therefore land in separators);
//   * part 0, part 1, then the separator are numbered; small pieces are
//     ordered by exact minimum degree on their elimination graph.
// The result is deterministic for a given seed.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <set>
#include <vector>

#include "symbolic.hpp"

namespace lblg {
namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed ? seed : 0x9E3779B97F4A7C15ull) {}
  uint64_t next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  }
  int below(int m) { return (int)(next() % (uint64_t)m); }
};

// weighted graph in CSR form
struct Graph {
  int n = 0;
  std::vector<int> xadj, adj, ew, vw;
  int64_t total_vw() const { return std::accumulate(vw.begin(), vw.end(), (int64_t)0); }
};

// heavy-edge matching coarsening; returns coarse graph and fine->coarse map
void coarsen(const Graph& g, Rng& rng, Graph& c, std::vector<int>& cmap) {
  const int n = g.n;
  std::vector<int> match(n, -1), order(n);
  std::iota(order.begin(), order.end(), 0);
  for (int i = n - 1; i > 0; --i) std::swap(order[i], order[rng.below(i + 1)]);
  for (int v : order) {
    if (match[v] != -1) continue;
    int best = -1, bw = -1;
    for (int q = g.xadj[v]; q < g.xadj[v + 1]; ++q) {
      const int u = g.adj[q];
      if (match[u] == -1 && u != v && g.ew[q] > bw) {
        bw = g.ew[q];
        best = u;
      }
    }
    if (best == -1) {
      match[v] = v;
    } else {
      match[v] = best;
      match[best] = v;
    }
  }
  cmap.assign(n, -1);
  int nc = 0;
  for (int v = 0; v < n; ++v) {
    if (cmap[v] != -1) continue;
    cmap[v] = nc;
    if (match[v] != v) cmap[match[v]] = nc;
    ++nc;
  }
  c.n = nc;
  c.vw.assign(nc, 0);
  c.xadj.assign(nc + 1, 0);
  c.adj.clear();
  c.ew.clear();
  std::vector<int> mark(nc, -1), slot(nc, 0);
  std::vector<int> members;  // fine vertices of coarse vertex, via scan
  std::vector<int> first(nc, -1), second(nc, -1);
  for (int v = 0; v < n; ++v) {
    const int cv = cmap[v];
    if (first[cv] == -1) first[cv] = v; else second[cv] = v;
  }
  for (int cv = 0; cv < nc; ++cv) {
    const int base = (int)c.adj.size();
    for (int t = 0; t < 2; ++t) {
      const int v = t == 0 ? first[cv] : second[cv];
      if (v < 0) continue;
      c.vw[cv] += g.vw[v];
      for (int q = g.xadj[v]; q < g.xadj[v + 1]; ++q) {
        const int cu = cmap[g.adj[q]];
        if (cu == cv) continue;
        if (mark[cu] != cv) {
          mark[cu] = cv;
          slot[cu] = (int)c.adj.size();
          c.adj.push_back(cu);
          c.ew.push_back(g.ew[q]);
        } else {
          c.ew[slot[cu]] += g.ew[q];
        }
      }
    }
    (void)base;
    c.xadj[cv + 1] = (int)c.adj.size();
  }
}

int64_t cut_of(const Graph& g, const std::vector<int>& part) {
  int64_t cut = 0;
  for (int v = 0; v < g.n; ++v)
    for (int q = g.xadj[v]; q < g.xadj[v + 1]; ++q)
      if (part[v] != part[g.adj[q]]) cut += g.ew[q];
  return cut / 2;
}

// greedy boundary refinement: move boundary vertices with positive gain (or
// zero gain that improves balance) while respecting the balance bound
void refine(const Graph& g, std::vector<int>& part, Rng& rng, int passes) {
  const int64_t tot = g.total_vw();
  const int64_t maxw = (int64_t)(0.55 * (double)tot) + 1;
  int64_t pw[2] = {0, 0};
  for (int v = 0; v < g.n; ++v) pw[part[v]] += g.vw[v];
  std::vector<int> order(g.n);
  std::iota(order.begin(), order.end(), 0);
  for (int pass = 0; pass < passes; ++pass) {
    for (int i = g.n - 1; i > 0; --i) std::swap(order[i], order[rng.below(i + 1)]);
    int moved = 0;
    for (int v : order) {
      int64_t ext = 0, in = 0;
      for (int q = g.xadj[v]; q < g.xadj[v + 1]; ++q) {
        if (part[g.adj[q]] == part[v]) in += g.ew[q]; else ext += g.ew[q];
      }
      if (ext == 0) continue;
      const int from = part[v], to = 1 - from;
      const int64_t gain = ext - in;
      if (pw[to] + g.vw[v] > maxw) continue;
      const bool better_balance = pw[to] + g.vw[v] < pw[from];
      if (gain > 0 || (gain == 0 && better_balance)) {
        part[v] = to;
        pw[from] -= g.vw[v];
        pw[to] += g.vw[v];
        ++moved;
      }
    }
    if (!moved) break;
  }
}

// greedy graph growing from a seed vertex until half of the weight is reached
void grow(const Graph& g, int seed, std::vector<int>& part) {
  const int64_t half = g.total_vw() / 2;
  part.assign(g.n, 1);
  std::vector<int64_t> gain(g.n, 0);
  std::vector<char> inq(g.n, 0);
  std::set<std::pair<int64_t, int>> frontier;  // (-gain, v)
  int64_t w = 0;
  auto add = [&](int v) {
    part[v] = 0;
    w += g.vw[v];
    for (int q = g.xadj[v]; q < g.xadj[v + 1]; ++q) {
      const int u = g.adj[q];
      if (part[u] != 1) continue;
      if (inq[u]) frontier.erase({-gain[u], u});
      gain[u] += 2 * g.ew[q];
      inq[u] = 1;
      frontier.insert({-gain[u], u});
    }
  };
  // gain of moving u into part 0 = (edges to part 0) - (edges to part 1)
  for (int v = 0; v < g.n; ++v) {
    int64_t s = 0;
    for (int q = g.xadj[v]; q < g.xadj[v + 1]; ++q) s += g.ew[q];
    gain[v] = -s;
  }
  add(seed);
  while (w < half) {
    int v;
    if (!frontier.empty()) {
      v = frontier.begin()->second;
      frontier.erase(frontier.begin());
      inq[v] = 0;
      if (part[v] == 0) continue;
    } else {
      v = -1;
      for (int u = 0; u < g.n; ++u)
        if (part[u] == 1) {
          v = u;
          break;
        }
      if (v < 0) break;
    }
    add(v);
  }
}

int bfs_far(const Graph& g, int s) {
  std::vector<int> dist(g.n, -1), q;
  q.push_back(s);
  dist[s] = 0;
  int last = s;
  for (size_t h = 0; h < q.size(); ++h) {
    const int v = q[h];
    last = v;
    for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      const int u = g.adj[k];
      if (dist[u] < 0) {
        dist[u] = dist[v] + 1;
        q.push_back(u);
      }
    }
  }
  return last;
}

// multilevel bisection of a connected graph
void bisect(const Graph& g0, Rng& rng, std::vector<int>& part) {
  std::vector<Graph> levels;
  std::vector<std::vector<int>> maps;
  levels.push_back(g0);
  while (levels.back().n > 160) {
    Graph c;
    std::vector<int> cmap;
    coarsen(levels.back(), rng, c, cmap);
    if (c.n > (int)(0.92 * levels.back().n)) break;
    maps.push_back(std::move(cmap));
    levels.push_back(std::move(c));
  }
  const Graph& gc = levels.back();
  std::vector<int> best;
  int64_t bestcut = -1;
  const int tries = 6;
  for (int t = 0; t < tries; ++t) {
    int seed = t == 0 ? bfs_far(gc, bfs_far(gc, 0)) : rng.below(gc.n);
    std::vector<int> p;
    grow(gc, seed, p);
    refine(gc, p, rng, 8);
    const int64_t c = cut_of(gc, p);
    if (bestcut < 0 || c < bestcut) {
      bestcut = c;
      best = p;
    }
  }
  part = best;
  for (int l = (int)levels.size() - 2; l >= 0; --l) {
    std::vector<int> fine(levels[l].n);
    for (int v = 0; v < levels[l].n; ++v) fine[v] = part[maps[l][v]];
    part.swap(fine);
    refine(levels[l], part, rng, 6);
  }
}

// exact minimum degree on a small graph (elimination graph kept explicitly)
void min_degree(const Graph& g, std::vector<int>& order) {
  const int n = g.n;
  std::vector<std::set<int>> nb(n);
  for (int v = 0; v < n; ++v)
    for (int q = g.xadj[v]; q < g.xadj[v + 1]; ++q)
      if (g.adj[q] != v) nb[v].insert(g.adj[q]);
  std::vector<char> done(n, 0);
  order.clear();
  for (int step = 0; step < n; ++step) {
    int best = -1;
    size_t bd = 0;
    for (int v = 0; v < n; ++v)
      if (!done[v] && (best < 0 || nb[v].size() < bd)) {
        best = v;
        bd = nb[v].size();
      }
    done[best] = 1;
    order.push_back(best);
    std::vector<int> cl(nb[best].begin(), nb[best].end());
    for (int a : cl) {
      nb[a].erase(best);
      for (int b : cl)
        if (a != b) nb[a].insert(b);
    }
    nb[best].clear();
  }
}

// induced subgraph on `verts` of the global graph
void induced(const Graph& G, const std::vector<int>& verts, std::vector<int>& loc, Graph& g) {
  g.n = (int)verts.size();
  for (int k = 0; k < g.n; ++k) loc[verts[k]] = k;
  g.xadj.assign(g.n + 1, 0);
  g.adj.clear();
  g.ew.clear();
  g.vw.assign(g.n, 1);
  for (int k = 0; k < g.n; ++k) {
    const int v = verts[k];
    for (int q = G.xadj[v]; q < G.xadj[v + 1]; ++q) {
      const int l = loc[G.adj[q]];
      if (l >= 0) {
        g.adj.push_back(l);
        g.ew.push_back(1);
      }
    }
    g.xadj[k + 1] = (int)g.adj.size();
  }
  for (int k = 0; k < g.n; ++k) loc[verts[k]] = -1;
}

}  // namespace

void nested_dissection(const CscView& A, std::vector<int>& perm, uint64_t seed) {
  const int n = A.n;
  Graph G;
  G.n = n;
  {
    std::vector<int> deg(n, 0);
    for (int j = 0; j < n; ++j)
      for (int q = A.p[j]; q < A.p[j + 1]; ++q) {
        const int i = A.i[q];
        if (i > j) {
          deg[i]++;
          deg[j]++;
        }
      }
    G.xadj.assign(n + 1, 0);
    for (int v = 0; v < n; ++v) G.xadj[v + 1] = G.xadj[v] + deg[v];
    G.adj.resize(G.xadj[n]);
    std::vector<int> f(G.xadj.begin(), G.xadj.end() - 1);
    for (int j = 0; j < n; ++j)
      for (int q = A.p[j]; q < A.p[j + 1]; ++q) {
        const int i = A.i[q];
        if (i > j) {
          G.adj[f[i]++] = j;
          G.adj[f[j]++] = i;
        }
      }
    // drop duplicate edges
    std::vector<int> mark(n, -1);
    std::vector<int> nx(1, 0), na;
    na.reserve(G.adj.size());
    for (int v = 0; v < n; ++v) {
      for (int q = G.xadj[v]; q < G.xadj[v + 1]; ++q) {
        const int u = G.adj[q];
        if (mark[u] != v) {
          mark[u] = v;
          na.push_back(u);
        }
      }
      nx.push_back((int)na.size());
    }
    G.xadj = nx;
    G.adj = na;
    G.vw.assign(n, 1);
    G.ew.assign(G.adj.size(), 1);
  }
  Rng rng(seed);
  perm.assign(n, -1);
  std::vector<int> loc(n, -1);
  // task = (vertex set, end position in perm)
  struct Task {
    std::vector<int> verts;
    int hi;
  };
  std::vector<Task> stack;
  {
    std::vector<int> all(n);
    std::iota(all.begin(), all.end(), 0);
    stack.push_back({std::move(all), n});
  }
  const int leaf = 96;
  while (!stack.empty()) {
    Task t = std::move(stack.back());
    stack.pop_back();
    const int m = (int)t.verts.size();
    if (m == 0) continue;
    Graph g;
    induced(G, t.verts, loc, g);
    // connected components first
    std::vector<int> comp(m, -1);
    int ncomp = 0;
    {
      std::vector<int> q;
      for (int s = 0; s < m; ++s) {
        if (comp[s] >= 0) continue;
        q.assign(1, s);
        comp[s] = ncomp;
        for (size_t h = 0; h < q.size(); ++h)
          for (int k = g.xadj[q[h]]; k < g.xadj[q[h] + 1]; ++k)
            if (comp[g.adj[k]] < 0) {
              comp[g.adj[k]] = ncomp;
              q.push_back(g.adj[k]);
            }
        ++ncomp;
      }
    }
    if (ncomp > 1) {
      std::vector<std::vector<int>> parts(ncomp);
      for (int k = 0; k < m; ++k) parts[comp[k]].push_back(t.verts[k]);
      int hi = t.hi;
      for (int c = ncomp - 1; c >= 0; --c) {
        const int sz = (int)parts[c].size();
        stack.push_back({std::move(parts[c]), hi});
        hi -= sz;
      }
      continue;
    }
    if (m <= leaf) {
      std::vector<int> ord;
      min_degree(g, ord);
      const int lo = t.hi - m;
      for (int k = 0; k < m; ++k) perm[lo + k] = t.verts[ord[k]];
      continue;
    }
    std::vector<int> part;
    bisect(g, rng, part);
    // vertex separator: cover cut edges, preferring high cut-degree vertices
    std::vector<int> cutdeg(m, 0);
    for (int v = 0; v < m; ++v)
      for (int q = g.xadj[v]; q < g.xadj[v + 1]; ++q)
        if (part[g.adj[q]] != part[v]) cutdeg[v]++;
    std::vector<int> cand;
    for (int v = 0; v < m; ++v)
      if (cutdeg[v] > 0) cand.push_back(v);
    std::sort(cand.begin(), cand.end(), [&](int a, int b) {
      return cutdeg[a] != cutdeg[b] ? cutdeg[a] > cutdeg[b] : a < b;
    });
    std::vector<char> sep(m, 0);
    for (int v : cand) {
      bool needed = false;
      for (int q = g.xadj[v]; q < g.xadj[v + 1] && !needed; ++q) {
        const int u = g.adj[q];
        if (part[u] != part[v] && !sep[u]) needed = true;
      }
      if (needed) sep[v] = 1;
    }
    // release separator vertices with no neighbour on the opposite side
    for (int v : cand) {
      if (!sep[v]) continue;
      bool touch[2] = {false, false};
      for (int q = g.xadj[v]; q < g.xadj[v + 1]; ++q) {
        const int u = g.adj[q];
        if (!sep[u]) touch[part[u]] = true;
      }
      if (!touch[0]) {
        sep[v] = 0;
        part[v] = 1;
      } else if (!touch[1]) {
        sep[v] = 0;
        part[v] = 0;
      }
    }
    std::vector<int> p0, p1, ps;
    for (int v = 0; v < m; ++v) {
      if (sep[v]) ps.push_back(t.verts[v]);
      else if (part[v] == 0) p0.push_back(t.verts[v]);
      else p1.push_back(t.verts[v]);
    }
    if (p0.empty() || p1.empty()) {  // degenerate split: order by minimum degree
      std::vector<int> ord;
      if (m <= 4000) {
        min_degree(g, ord);
      } else {
        ord.resize(m);
        std::iota(ord.begin(), ord.end(), 0);
      }
      const int lo = t.hi - m;
      for (int k = 0; k < m; ++k) perm[lo + k] = t.verts[ord[k]];
      continue;
    }
    const int hs = t.hi - (int)ps.size();
    for (size_t k = 0; k < ps.size(); ++k) perm[hs + k] = ps[k];
    const int h1 = hs, h0 = hs - (int)p1.size();
    stack.push_back({std::move(p1), h1});
    stack.push_back({std::move(p0), h0});
  }
}

}  // namespace lblg

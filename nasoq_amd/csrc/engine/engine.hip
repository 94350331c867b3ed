// Host orchestration of the device LBL engine: uploads the symbolic result,
// builds the level schedules, runs the level-scheduled supernodal
// factorization, the blocked sweeps and the device GMRES refinement.
//
// Reference call flow replaced here (src/QP/linear_solver_wrapper.cpp):
//   numerical_factorization (:602-768) -> ldl_left_sn_parallel_02 + reorder_matrix
//   solve_only (:1305-1404) -> solve_phase_ldl_blocked_parallel_permuted +
//                              pmgmres_ldlt_auto (refinement on the unperturbed K)
#include "engine/engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "engine/device.hpp"
#include "engine/handle.hpp"
#include "engine/kernels.hpp"

namespace lblg {

#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

DeviceFactor::~DeviceFactor() {
  resolve_timing();
  for (GraphCache* g : {&gfactor, &gsolve, &gcycle})
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
  for (void* p : owned) hipFree(p);
  if (blob_d) hipFree(blob_d);
  if (blob_h) hipHostFree(blob_h);
  if (stream) hipStreamDestroy(stream);
}

void DeviceFactor::resolve_timing() {
  if (pending.empty()) return;
  hipStreamSynchronize(stream);
  for (auto& e : pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, e.a, e.b);
    auto& t = ktime[e.name];
    t.first += ms;
    t.second += 1;
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  pending.clear();
}

// run `launch` bracketed by events on the engine stream when timing is on
template <typename F>
static void timed(DeviceFactor& D, const char* name, F launch) {
  if (!D.timing || (!D.timing_only.empty() && !D.timing_only.count(name))) {
    launch();
    return;
  }
  DeviceFactor::Pending p{name, nullptr, nullptr};
  hipEventCreate(&p.a);
  hipEventCreate(&p.b);
  hipEventRecord(p.a, D.stream);
  launch();
  hipEventRecord(p.b, D.stream);
  D.pending.push_back(p);
  if (D.pending.size() > 4096) D.resolve_timing();
}

// Run a launch sequence through the graph cache G: eagerly on the first call
// with this key, captured (stream capture, then instantiate) and launched on
// the second, replayed afterwards.  The sequence must only enqueue kernels on
// D.stream whose arguments are fixed by the key.
template <typename F>
static void run_graph(DeviceFactor& D, DeviceFactor::GraphCache& G, const std::string& key, F body) {
  if (!D.graphs || D.timing) {
    body();
    return;
  }
  if (G.key != key) {
    if (G.exec) HIPCHK(hipGraphExecDestroy(G.exec));
    G.exec = nullptr;
    G.key = key;
    G.seen = 0;
  }
  if (G.exec) {
    HIPCHK(hipGraphLaunch(G.exec, D.stream));
    ++D.graph_launches;
    return;
  }
  if (G.seen++ == 0) {
    body();
    return;
  }
  HIPCHK(hipStreamBeginCapture(D.stream, hipStreamCaptureModeRelaxed));
  try {
    body();
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(D.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamEndCapture(D.stream, &g));
  const hipError_t e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHK(e);
  ++D.graph_captures;
  HIPCHK(hipGraphLaunch(G.exec, D.stream));
  ++D.graph_launches;
}

static std::string graph_key(std::initializer_list<double> v, std::initializer_list<const void*> p) {
  std::string k;
  char buf[40];
  for (double x : v) {
    std::snprintf(buf, sizeof(buf), "%a;", x);
    k += buf;
  }
  for (const void* q : p) {
    std::snprintf(buf, sizeof(buf), "%p;", q);
    k += buf;
  }
  return k;
}

template <typename T>
static T* dalloc(DeviceFactor& D, int64_t count) {
  if (count <= 0) count = 1;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(T) * (size_t)count);
  if (e != hipSuccess) throw std::bad_alloc();
  D.owned.push_back(p);
  return static_cast<T*>(p);
}
template <typename T>
static T* dupload(DeviceFactor& D, const std::vector<T>& v) {
  T* p = dalloc<T>(D, (int64_t)v.size());
  if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, D.stream));
  return p;
}

// an array of the factor's value type
static void* valloc(DeviceFactor& D, int64_t count) {
  return D.prec != 0 ? (void*)dalloc<float>(D, count) : (void*)dalloc<double>(D, count);
}

// the device factor's precision code of a setting: 0 FP64, 1 FP32, 2 BF16
// operands in the 128-tile updates (FP32 storage) with diagonal scaling
static int prec_code(int precision) { return precision == LBLG_F32 ? 1 : precision == LBLG_BF16 ? 2 : 0; }

static SnArgs snargs(const DeviceFactor& D) {
  return SnArgs{D.super, D.pi, D.px, D.rows, D.col2sup, D.n};
}

static int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v ? std::atoi(v) : dflt;
}

// Level schedule of the one-workgroup sweeps over the supernodes [s0, s1)
// (the whole factor, or one system of a block-diagonal batch): order lists
// them level by level (level l in [ptr[l], ptr[l+1])), mode[l] is the level's
// sweep mode (bit 0) and thin count (bits 1..).
static void ss_schedule(const Symbolic& S, const std::vector<int>& lev, int s0, int s1, std::vector<int>& order,
                        std::vector<int>& ptr, std::vector<int>& mode) {
  int nl = 0;
  for (int s = s0; s < s1; ++s) nl = std::max(nl, lev[s] + 1);
  ptr.assign(nl + 1, 0);
  order.assign(s1 - s0, 0);
  mode.assign(nl, 0);
  for (int s = s0; s < s1; ++s) ++ptr[lev[s] + 1];
  for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
  std::vector<int> pos(ptr.begin(), ptr.end() - 1);
  // "thin" supernodes (one column, at most SS_THIN_ROWS rows: the singleton
  // constraint rows of the QP's KKT) are swept one per lane; they come first
  // within their level
  auto thin = [&](int sn) {
    return S.super[sn + 1] - S.super[sn] == 1 && S.pi[sn + 1] - S.pi[sn] <= SS_THIN_ROWS;
  };
  std::vector<int> nthin(nl, 0);
  for (int s = s0; s < s1; ++s)
    if (thin(s)) {
      order[pos[lev[s]]++] = s;
      ++nthin[lev[s]];
    }
  for (int s = s0; s < s1; ++s)
    if (!thin(s)) order[pos[lev[s]]++] = s;
  // mode of each level: per-wave supernodes, or the whole workgroup on one
  // supernode at a time.  Cost model in dependent global-load steps: a
  // per-wave triangle costs a step per column, a 64-row block below it a
  // quarter step per column (independent loads); the workgroup mode stages
  // the triangle (one step), substitutes from LDS (a fifth of a step per
  // column), splits the row blocks over 16 waves and pays three barriers
  // per 64-column chunk.
  // one lane per thin supernode pays where a level holds many of them; a few
  // (chain tails) stay with the prefetched per-wave sweep
  const bool thin_on = env_int("LBLG_SMALL_SOLVE_THIN", 1) != 0;
  std::vector<int> lthin(nl, 0);
  for (int l = 0; l < nl; ++l) lthin[l] = (thin_on && nthin[l] >= SS_THIN_MIN) ? nthin[l] : 0;
  for (int l = 0; l < nl; ++l) {
    double wave_max = 0.0, wave_sum = 0.0, wg = 0.0;
    for (int k = ptr[l] + lthin[l]; k < ptr[l + 1]; ++k) {
      const int sn = order[k];
      const int w = S.super[sn + 1] - S.super[sn];
      const int64_t r = S.pi[sn + 1] - S.pi[sn];
      double cw_cost = 0.0, cg_cost = 0.0;
      for (int cb = 0; cb < w; cb += 64) {
        const int cw = std::min(64, w - cb);
        const int64_t rb = r - cb - cw;
        cw_cost += (cw - 1) + 0.25 * cw * (double)((rb + 63) / 64);
        cg_cost += 1.0 + 0.2 * (cw - 1) + 0.25 * cw * (double)((rb + 1023) / 1024) + 1.5;
      }
      wave_max = std::max(wave_max, cw_cost);
      wave_sum += cw_cost;
      wg += cg_cost;
    }
    const double per_wave = std::max(wave_max, wave_sum / 16.0) + 0.5;
    mode[l] = (ptr[l + 1] - ptr[l] > lthin[l] && wg < per_wave) ? 1 : 0;
  }
  const int force = env_int("LBLG_SMALL_SOLVE_MODE", -1);  // 0 / 1: every level in that mode
  if (force == 0 || force == 1) std::fill(mode.begin(), mode.end(), force);
  for (int l = 0; l < nl; ++l) mode[l] |= lthin[l] << 1;  // bit 0: mode, bits 1..: thin count
}


// ---------------------------------------------------------------------------
// one-time device setup for an analysed handle
// ---------------------------------------------------------------------------
// schedule knobs not set on the handle take the environment's or built-in default
void resolve_tuning(Handle& h) {
  Handle::Tuning& t = h.tune;
  // defaults: the autotuner's choice on the 64^3 KKT (3-D Laplacian + 16,384
  // constraint rows), the largest workload measured
  if (t.group_w < 0) t.group_w = env_int("LBLG_GROUP_W", 1024);
  if (t.chain_w < 0) t.chain_w = env_int("LBLG_CHAIN_W", 1024);
  if (t.small_nnz < 0) t.small_nnz = env_int("LBLG_SMALL_SOLVE_NNZ", 1 << 20);
  if (t.big_m < 0) t.big_m = env_int("LBLG_SCHUR_BIG_M", 256);
  if (t.wide_sub < 0) t.wide_sub = env_int("LBLG_WIDE_SUB", 8);
  t.wide_sub = std::max(1, std::min(t.wide_sub, 16));
}

static void setup(Handle& h) {
  resolve_tuning(h);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw std::runtime_error("no HIP device");
  h.dev.reset(new DeviceFactor);
  DeviceFactor& D = *h.dev;
  const Symbolic& S = h.sym;
  const int n = h.n, ns = S.nsuper;
  D.n = n;
  D.nsuper = ns;
  HIPCHK(hipStreamCreateWithFlags(&D.stream, hipStreamNonBlocking));

  std::vector<int64_t> px(ns + 1);
  for (int s = 0; s < ns; ++s) px[s] = S.p[S.super[s]];
  px[ns] = S.p[n];
  D.super = dupload(D, S.super);
  D.pi = dupload(D, S.pi);
  D.hsuper = S.super;
  D.hpi = S.pi;
  D.px = dupload(D, px);
  std::vector<int> rows0(S.s.begin(), S.s.begin() + S.pi[ns]);
  D.rows0 = dupload(D, rows0);
  D.rows = dalloc<int>(D, (int64_t)rows0.size());
  D.col2sup = dupload(D, S.col2sup);
  D.perm0 = dupload(D, S.perm);
  D.prec = prec_code(h.set.precision);
  D.valL = valloc(D, S.p[n]);
  D.dval = valloc(D, 2 * (int64_t)n);
  D.perm_piv = dalloc<int>(D, n);
  D.newpos = dalloc<int>(D, n);
  D.perm_f = dalloc<int>(D, n);
  D.iperm_f = dalloc<int>(D, n);
  D.status = dalloc<int>(D, 16);

  // entry -> panel slot map (the destination of each entry of P K P^T)
  const int nnz = h.Ap[n];
  D.nnzK = nnz;
  std::vector<int64_t> amap(nnz);
  std::vector<unsigned char> isdiag(nnz, 0);
  D.hasup.assign(nnz, 0);
  for (int j = 0; j < n; ++j) {
    for (int q = h.Ap[j]; q < h.Ap[j + 1]; ++q) {
      const int i = h.Ai[q];
      isdiag[q] = (i == j);
      const int a = S.iperm[i], b = S.iperm[j];
      const int c = std::min(a, b), rw = std::max(a, b);
      const int t = S.col2sup[c];
      const int* rb = S.s.data() + S.pi[t];
      const int* re = S.s.data() + S.pi[t + 1];
      const int64_t pos = std::lower_bound(rb, re, rw) - rb;
      const int64_t rt = S.pi[t + 1] - S.pi[t];
      amap[q] = px[t] + (int64_t)(c - S.super[t]) * rt + pos;
      D.hasup[q] = t;
    }
  }
  D.amap = dupload(D, amap);
  D.asup = dupload(D, D.hasup);
  D.tmask = dalloc<unsigned char>(D, std::max(ns, 1));
  if (D.prec != 0) {  // the scaled matrix S K S of the FP32 / BF16 factors
    std::vector<int> kcol(nnz);
    for (int j = 0; j < n; ++j)
      for (int q = h.Ap[j]; q < h.Ap[j + 1]; ++q) kcol[q] = j;
    D.Krow = dupload(D, h.Ai);
    D.Kcol = dupload(D, kcol);
    D.Kf = dalloc<double>(D, nnz);
    D.scale = dalloc<double>(D, n);
    D.scale_ord = dalloc<double>(D, n);
  }
  D.Kx = dalloc<double>(D, nnz);
  D.pert = dalloc<unsigned char>(D, nnz);
  D.isdiag = dupload(D, isdiag);
  D.red = dalloc<double>(D, 1024);
  D.k0norm = dalloc<double>(D, 1);
  if (h.batch_off.size() > 2) {  // a batch: backward errors are checked per system
    D.nseg = (int)h.batch_off.size() - 1;
    D.seg = dupload(D, h.batch_off);
    D.segnorm = dalloc<double>(D, D.nseg);
    D.segsums = dalloc<double>(D, (int64_t)D.nseg * 3 * MAX_RHS);
  }

  // full CSR of K in the original ordering (structure) and its source entries
  {
    std::vector<int> cnt(n + 1, 0);
    for (int j = 0; j < n; ++j)
      for (int q = h.Ap[j]; q < h.Ap[j + 1]; ++q) {
        const int i = h.Ai[q];
        cnt[i + 1]++;
        if (i != j) cnt[j + 1]++;
      }
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> ci(cnt[n]);
    std::vector<int64_t> src(cnt[n]);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int j = 0; j < n; ++j)
      for (int q = h.Ap[j]; q < h.Ap[j + 1]; ++q) {
        const int i = h.Ai[q];
        int d = fill[i]++;
        ci[d] = j;
        src[d] = q;
        if (i != j) {
          d = fill[j]++;
          ci[d] = i;
          src[d] = q;
        }
      }
    D.Krp = dupload(D, cnt);
    D.Kci = dupload(D, ci);
    D.Ksrc = dupload(D, src);
    D.Kv = dalloc<double>(D, cnt[n]);
    D.nnzf = cnt[n];
    h.csr_src = std::move(src);
    h.csr_rp = std::move(cnt);
  }

  // ---- factor levels: height of each supernode above the leaves
  std::vector<int> lev(ns, 0);
  int nlev = 0;
  for (int s = 0; s < ns; ++s) {
    nlev = std::max(nlev, lev[s] + 1);
    if (S.sparent[s] >= 0) lev[S.sparent[s]] = std::max(lev[S.sparent[s]], lev[s] + 1);
  }
  std::vector<std::vector<int>> bylev(nlev);
  for (int s = 0; s < ns; ++s) bylev[lev[s]].push_back(s);
  const int group_w = h.tune.group_w;
  // LBLG_WIDE_TALL=1: every cooperative panel takes the tall launch shape
  // (tests run it at small sizes)
  const bool wide_tall = env_int("LBLG_WIDE_TALL", 0) != 0;
  int64_t ld_max = 0;
  int64_t wmax_total = 0;
  std::vector<int64_t> woff(ns, 0);
  D.flevels.resize(nlev);
  D.s_kind.assign(ns, 0);
  D.s_off.assign(ns, 0);
  D.s_sch.assign(ns, 0);
  D.s_large_q.assign(ns, -1);
  D.s_ldoff.assign(ns, 0);
  D.s_mark.assign(ns, 0);
  D.s_send.assign(ns, 0);
  for (int l = 0; l < nlev; ++l) {
    FactorLevel& F = D.flevels[l];
    std::vector<int> small, off, sch, schb, offblk, big;
    std::vector<int64_t> tiles, tilesb, ldoffb;
    std::vector<int> colsb;
    int blk = 0, ncolb = 0;
    int64_t tl = 0, tlb = 0, wsum = 0, ldsum = 0;
    for (int s : bylev[l]) {
      const int w = S.super[s + 1] - S.super[s];
      const int64_t r = S.pi[s + 1] - S.pi[s];
      F.max_w = std::max(F.max_w, w);
      if (w <= SMALL_W) {
        small.push_back(s);
      } else {
        woff[s] = wsum;
        if (w <= group_w) {
          F.medium.push_back(s);
          D.s_kind[s] = 1;
          wsum += (int64_t)w * PANEL_NB + 2 * (int64_t)w + 16;
        } else {
          // cooperative panel over G workgroups, <= WIDE_ROWS far rows each
          // (the diagonal block below the panel and the off-diagonal rows);
          // beyond 256 workgroups (49,152 rows), 512-thread workgroups with
          // WIDE_ROWS_TALL far rows each and 256-byte panel rows
          int nt = 256, nb = WIDE_NB, rows = WIDE_ROWS;
          int64_t G = std::max<int64_t>(1, (r - nb + rows - 1) / rows);
          if (G > 256 || wide_tall) {
            nt = WIDE_NT_TALL;
            nb = wide_panel_nb(nt, D.prec != 0 ? sizeof(float) : sizeof(double));
            rows = WIDE_ROWS_TALL;
            G = std::max<int64_t>(1, (r - nb + rows - 1) / rows);
          }
          if (G > 256)
            throw std::runtime_error("supernode too tall for the cooperative panel kernel (more than 256 x " +
                                     std::to_string(WIDE_ROWS_TALL) + " rows)");
          // trailing updates of rank wide_sub x 64 whatever the panel width
          const int wsub = std::min(16, h.tune.wide_sub * WIDE_NB / nb);
          const int64_t wcols = (int64_t)wsub * nb;
          wsum += (int64_t)w * wcols + r * nb + (int64_t)WIDE_MOVED * (w + r + wcols) + 64;
          F.large.push_back(s);
          D.has_coop = true;
          D.s_kind[s] = 2;
          D.s_large_q[s] = (int)F.large_G.size();
          F.large_G.push_back((int)G);
          F.large_nb.push_back(nb);
          F.large_nt.push_back(nt);
          F.large_wsub.push_back(wsub);
        }
      }
      if (r > w) {
        // the cooperative panel computes the off-diagonal rows of wide supernodes itself
        if (w <= group_w) {
          off.push_back(s);
          D.s_off[s] = 1;
          offblk.push_back(blk);
          blk += (int)((r - w + 63) / 64);
        }
        if (r - w >= h.tune.big_m) {
          schb.push_back(s);
          D.s_sch[s] = 2;
          D.s_ldoff[s] = ldsum;
          tilesb.push_back(tlb);
          colsb.push_back(ncolb);
          ldoffb.push_back(ldsum);
          ncolb += w;
          ldsum += (r - w) * (int64_t)w;
          const int64_t T = (r - w + 127) / 128;
          tlb += T * (T + 1) / 2;
        } else {
          sch.push_back(s);
          D.s_sch[s] = 1;
          tiles.push_back(tl);
          const int64_t T = (r - w + 63) / 64;
          tl += T * (T + 1) / 2;
        }
      }
    }
    wmax_total = std::max(wmax_total, wsum);
    F.h_small = small;
    F.h_off = off;
    F.h_sch = sch;
    F.h_schb = schb;
    F.h_schb_ldoff = ldoffb;
    F.small.count = (int)small.size();
    F.small_wmax = 0;
    for (int s : small) F.small_wmax = std::max(F.small_wmax, S.super[s + 1] - S.super[s]);
    F.small.d = dupload(D, small);
    F.offdiag.count = (int)off.size();
    F.offdiag.d = dupload(D, off);
    F.offdiag_wmax = 0;
    for (int s : off) F.offdiag_wmax = std::max(F.offdiag_wmax, S.super[s + 1] - S.super[s]);
    F.offdiag_blocks = blk;
    F.d_offdiag_blk = dupload(D, offblk);
    {  // row block -> position of its supernode in the off-diagonal list
      std::vector<int> pos(blk);
      for (size_t k = 0; k < off.size(); ++k) {
        const int e = k + 1 < off.size() ? offblk[k + 1] : blk;
        for (int b = offblk[k]; b < e; ++b) pos[b] = (int)k;
      }
      F.d_offdiag_pos = dupload(D, pos);
    }
    F.schur.count = (int)sch.size();
    F.schur.d = dupload(D, sch);
    F.schur_tiles = tl;
    F.d_schur_tile = dupload(D, tiles);
    {  // tile -> position of its supernode in the Schur list
      if (tl > 0x7fffffffLL) throw std::runtime_error("schur: too many tiles");
      std::vector<int> pos((size_t)tl);
      for (size_t k = 0; k < sch.size(); ++k) {
        const int64_t e = k + 1 < sch.size() ? tiles[k + 1] : tl;
        for (int64_t b = tiles[k]; b < e; ++b) pos[(size_t)b] = (int)k;
      }
      F.d_schur_pos = dupload(D, pos);
    }
    F.schur_big.count = (int)schb.size();
    F.schur_big.d = dupload(D, schb);
    F.schur_big_tiles = tlb;
    F.d_schur_big_tile = dupload(D, tilesb);
    F.d_schur_big_cols = dupload(D, colsb);
    F.schur_big_ncols = ncolb;
    F.d_schur_big_ldoff = dupload(D, ldoffb);
    ld_max = std::max(ld_max, ldsum);
    F.d_medium = dupload(D, F.medium);
    F.d_large = dupload(D, F.large);
  }
  D.hlevel = lev;
  {  // supernodes with rows in each supernode's columns (distinct, ascending)
    std::vector<int> cnt(ns + 1, 0);
    auto each_target = [&](int d, auto f) {
      const int w = S.super[d + 1] - S.super[d];
      int last = -1;
      for (int64_t q = S.pi[d] + w; q < S.pi[d + 1]; ++q) {
        const int t = S.col2sup[S.s[q]];
        if (t != last) f(t);
        last = t;
      }
    };
    for (int d = 0; d < ns; ++d) each_target(d, [&](int t) { ++cnt[t + 1]; });
    for (int t = 0; t < ns; ++t) cnt[t + 1] += cnt[t];
    D.upd_ptr = cnt;
    D.upd_idx.resize(cnt[ns]);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int d = 0; d < ns; ++d) each_target(d, [&](int t) { D.upd_idx[fill[t]++] = d; });
    D.smark.assign(ns, 0);
    D.smark_epoch = 0;
  }
  D.W = valloc(D, wmax_total);
  D.LD = valloc(D, std::max<int64_t>(ld_max, 1));
  D.W_elems = wmax_total;
  D.woff = dupload(D, woff);
  D.kpos = dalloc<int>(D, ns);
  D.kstart = dalloc<int>(D, ns);
  D.sync = dalloc<int>(D, std::max<int64_t>(ns, 2));
  launch_fill_int(D.sync, (int)std::max<int64_t>(ns, 2), 0, D.stream);  // the wide panel's two counters start at 0
  D.wide_par = 0;
  D.ll = reinterpret_cast<uint64_t*>(dalloc<double>(D, LL_WORDS));
  launch_fill(reinterpret_cast<double*>(D.ll), LL_WORDS, 0.0, D.stream);
  D.ll_epoch = 0;

  // ---- solve pieces and their levels
  std::vector<int> psn, pc0, first_piece(ns), piece_par;
  for (int s = 0; s < ns; ++s) {
    const int w = S.super[s + 1] - S.super[s];
    first_piece[s] = (int)psn.size();
    for (int c = 0; c < w; c += SOLVE_W) {
      psn.push_back(s);
      pc0.push_back(c);
    }
  }
  const int np = (int)psn.size();
  piece_par.assign(np, -1);
  for (int p = 0; p < np; ++p) {
    const int s = psn[p];
    const int w = S.super[s + 1] - S.super[s];
    if (pc0[p] + SOLVE_W < w) piece_par[p] = p + 1;
    else if (S.sparent[s] >= 0) piece_par[p] = first_piece[S.sparent[s]];
  }
  // wide supernodes are swept by one cooperative launch ("chain"): all their
  // pieces take one level
  const int chain_w = h.tune.chain_w;
  std::vector<char> chain(ns, 0);
  int chain_pmax = 1;
  for (int s = 0; s < ns; ++s) {
    const int w = S.super[s + 1] - S.super[s];
    const int64_t r = S.pi[s + 1] - S.pi[s];
    if (w > SOLVE_W && w >= chain_w && chain_groups(w, r, true) > 0 && chain_groups(w, r, false) > 0) {
      chain[s] = 1;
      chain_pmax = std::max(chain_pmax, (w + SOLVE_W - 1) / SOLVE_W);
    }
  }
  std::vector<int> plev(np, 0);
  int nplev = 0;
  for (int p = 0; p < np; ++p) {  // parents come later in this numbering
    const int s = psn[p];
    if (chain[s]) {
      if (p != first_piece[s]) continue;
      const int w = S.super[s + 1] - S.super[s];
      const int lastp = p + (w + SOLVE_W - 1) / SOLVE_W - 1;
      for (int q = p + 1; q <= lastp; ++q) plev[q] = plev[p];
      nplev = std::max(nplev, plev[p] + 1);
      const int par = piece_par[lastp];
      if (par >= 0) plev[par] = std::max(plev[par], plev[p] + 1);
      continue;
    }
    nplev = std::max(nplev, plev[p] + 1);
    if (piece_par[p] >= 0) plev[piece_par[p]] = std::max(plev[piece_par[p]], plev[p] + 1);
  }
  std::vector<std::vector<int>> pbylev(nplev);
  D.slevels.resize(nplev);
  for (int p = 0; p < np; ++p) {
    const int s = psn[p];
    if (!chain[s]) pbylev[plev[p]].push_back(p);
    else if (p == first_piece[s]) {
      D.slevels[plev[p]].chains.push_back(s);
      D.has_chains = true;
    }
  }
  {  // supernode levels (parents come later in the numbering) for the one-workgroup sweeps
    std::vector<int> lev(ns, 0);
    int nl = 0;
    for (int s = 0; s < ns; ++s) {
      nl = std::max(nl, lev[s] + 1);
      const int par = S.sparent[s];
      if (par >= 0) {
        if (par <= s) throw std::runtime_error("supernode tree: parent numbered before child");
        lev[par] = std::max(lev[par], lev[s] + 1);
      }
    }
    std::vector<int> ptr, order, mode;
    ss_schedule(S, lev, 0, ns, order, ptr, mode);
    std::vector<int> info(4 * (size_t)ns);
    std::vector<int64_t> ipx(ns);
    bool fits = nl <= SOLVE_SMALL_LEVELS && S.pi[ns] < (int64_t)INT32_MAX;
    for (int k = 0; k < ns; ++k) {
      const int sn = order[k];
      const int64_t r = S.pi[sn + 1] - S.pi[sn];
      info[4 * k] = S.super[sn];
      info[4 * k + 1] = S.super[sn + 1] - S.super[sn];
      info[4 * k + 2] = (int)r;
      info[4 * k + 3] = (int)S.pi[sn];
      ipx[k] = px[sn];
    }
    D.ss_info = reinterpret_cast<int4*>(dupload(D, info));
    D.ss_ipx = dupload(D, ipx);
    D.ss_levptr = dupload(D, ptr);
    D.ss_mode = dupload(D, mode);
    if (!fits) nl = 0;  // too many levels for the LDS copy: the level sweeps only
    D.ss_nlev = nl;
    // 1: the block in memory even where it fits LDS; 2: in memory only where it does not
    D.ss_global = env_int("LBLG_SMALL_SOLVE_GLOBAL", 0);
    // A block-diagonal batch: one schedule per system, swept by a workgroup
    // of its own.  Each system's supernodes must form a contiguous range
    // whose rows stay within the system's columns (its levels are then the
    // global ones); otherwise the batch keeps the level sweeps.
    const int nsys = h.batch_off.size() > 2 ? (int)h.batch_off.size() - 1 : 0;
    D.ssb_nsys = 0;
    if (nsys > 0 && env_int("LBLG_BATCH_SWEEPS", 1) != 0 && S.pi[ns] < (int64_t)INT32_MAX) {
      std::vector<int> sb(nsys + 1, ns);
      bool ok = true;
      int g = 0;
      for (int s = 0; s < ns && ok; ++s) {
        while (g < nsys && S.super[s] >= h.batch_off[g + 1]) ++g;
        if (g >= nsys || S.super[s] < h.batch_off[g] || S.super[s + 1] > h.batch_off[g + 1] ||
            S.s[S.pi[s + 1] - 1] >= h.batch_off[g + 1])
          ok = false;
        else if (sb[g] == ns) sb[g] = s;
      }
      for (int q = 0; q < nsys && ok; ++q) ok = sb[q] < ns && (q == 0 || sb[q] > sb[q - 1]);
      std::vector<int> binfo, blp, bmd, bsys;
      std::vector<int64_t> bipx;
      std::vector<double> work(nsys, 0.0);  // stored entries of each system's factor
      int maxlev = 0;
      for (int q = 0; q < nsys && ok; ++q) {
        const int s0 = sb[q], s1 = q + 1 < nsys ? sb[q + 1] : ns;
        std::vector<int> o, lptr, md;
        ss_schedule(S, lev, s0, s1, o, lptr, md);
        const int nlq = (int)md.size();
        if (nlq < 1 || nlq > SOLVE_SMALL_LEVELS) {
          ok = false;
          break;
        }
        maxlev = std::max(maxlev, nlq);
        const int base = (int)(binfo.size() / 4);
        bsys.insert(bsys.end(), {(int)blp.size(), (int)bmd.size(), nlq, (int)h.batch_off[q],
                                 (int)(h.batch_off[q + 1] - h.batch_off[q]), 0});
        for (int v : lptr) blp.push_back(base + v);
        bmd.insert(bmd.end(), md.begin(), md.end());
        for (int sn : o) {
          binfo.insert(binfo.end(), {S.super[sn], S.super[sn + 1] - S.super[sn], (int)(S.pi[sn + 1] - S.pi[sn]),
                                     (int)S.pi[sn]});
          bipx.push_back(px[sn]);
          work[q] += (double)(S.super[sn + 1] - S.super[sn]) * (double)(S.pi[sn + 1] - S.pi[sn]);
        }
      }
      if (ok) {
        // workgroups start in system order: the largest systems first, so that
        // the small ones fill in behind them instead of a large one starting last
        std::vector<int> by(nsys);
        for (int q = 0; q < nsys; ++q) by[q] = q;
        std::stable_sort(by.begin(), by.end(), [&](int x, int y) { return work[x] > work[y]; });
        std::vector<int> sorted(bsys.size());
        for (int q = 0; q < nsys; ++q)
          for (int k = 0; k < 6; ++k) sorted[6 * (size_t)q + k] = bsys[6 * (size_t)by[q] + k];
        bsys.swap(sorted);
        D.ssb_sysd = dupload(D, bsys);
        D.ssb_info = reinterpret_cast<int4*>(dupload(D, binfo));
        D.ssb_ipx = dupload(D, bipx);
        D.ssb_levptr = dupload(D, blp);
        D.ssb_mode = dupload(D, bmd);
        D.ssb_maxlev = maxlev;
        D.ssb_lds_kb = env_int("LBLG_BATCH_SWEEP_LDS_KB", (int)(SOLVE_SMALL_LDS / 1024));
        D.ssb_cols.resize(nsys);
        for (int q = 0; q < nsys; ++q) D.ssb_cols[q] = (int)(h.batch_off[q + 1] - h.batch_off[q]);
        std::sort(D.ssb_cols.begin(), D.ssb_cols.end());
        D.ssb_nsys = nsys;
      }
    }
    D.ss_nnzL = S.p[n];
    D.ss_max_nnz = h.tune.small_nnz;
    D.graphs = env_int("LBLG_GRAPH", 1) != 0;
    D.schur_direct = env_int("LBLG_SCHUR_DIRECT", 1) != 0;
    D.panel_lds = env_int("LBLG_PANEL_LDS", 1) != 0;
    D.path_refactor = env_int("LBLG_PATH_REFACTOR", 1) != 0;
    D.trailing_direct = env_int("LBLG_TRAILING_DIRECT", 1) != 0;
    D.strip64 = env_int("LBLG_STRIP_TILE64", 1) != 0;
  }
  {  // one published block of SOLVE_W x MAX_RHS values per piece, two words per value
    const int64_t words = (int64_t)chain_pmax * SOLVE_W * MAX_RHS * 2;
    D.chain_xb = reinterpret_cast<uint64_t*>(dalloc<double>(D, words));
    launch_fill(reinterpret_cast<double*>(D.chain_xb), words, 0.0, D.stream);
  }
  D.chain_epoch = 0;
  for (int l = 0; l < nplev; ++l) {
    SolveLevel& L = D.slevels[l];
    // narrowest pieces first: the triangle kernels launch the narrow ones
    // apart, with LDS sized for them (many workgroups per CU)
    auto pwidth = [&](int p) { return std::min(SOLVE_W, S.super[psn[p] + 1] - S.super[psn[p]] - pc0[p]); };
    std::stable_sort(pbylev[l].begin(), pbylev[l].end(), [&](int x, int y) { return pwidth(x) < pwidth(y); });
    std::vector<int> pre;
    int nb = 0;
    for (int p : pbylev[l]) {
      const int s = psn[p];
      const int w = S.super[s + 1] - S.super[s];
      const int64_t r = S.pi[s + 1] - S.pi[s];
      const int pw = std::min(SOLVE_W, w - pc0[p]);
      pre.push_back(nb);
      nb += (int)((r - pc0[p] - pw + SOLVE_ROWS - 1) / SOLVE_ROWS);
      L.maxpw = std::max(L.maxpw, pw);
      if (pw <= TRI_NARROW) {
        ++L.nnarrow;
        L.maxpw_narrow = std::max(L.maxpw_narrow, pw);
        L.nblocks_narrow = nb;  // row blocks of the narrow pieces (they come first)
      }
    }
    L.npieces = (int)pbylev[l].size();
    L.nblocks = nb;
    L.d_piece = dupload(D, pbylev[l]);
    L.d_blk = dupload(D, pre);
    std::vector<int> bpos(nb);  // row block -> position of its piece in the level's list
    for (int k = 0; k < L.npieces; ++k) {
      const int e = k + 1 < L.npieces ? pre[k + 1] : nb;
      for (int b = pre[k]; b < e; ++b) bpos[b] = k;
    }
    L.d_blkpos = dupload(D, bpos);
  }
  D.piece_sn = dupload(D, psn);
  D.piece_c0 = dupload(D, pc0);
  D.npieces = np;
  HIPCHK(hipStreamSynchronize(D.stream));
}

// ---------------------------------------------------------------------------
// numeric factorization
// ---------------------------------------------------------------------------
// the lists one factor level launches over: all of the level's supernodes,
// or the subsets of a path-restricted refactorization
struct LevelLists {
  const int* small = nullptr;
  int nsmall = 0, small_wmax = 0;
  const int* medium = nullptr;
  int nmedium = 0, medium_wmax = 0;
  std::vector<int> large_q;  // positions in FactorLevel::large
  const int* off = nullptr;
  int noff = 0, off_wmax = 0, off_blocks = 0;
  const int* off_blk = nullptr;
  const int* off_pos = nullptr;
  const int* sch = nullptr;
  int nsch = 0;
  int64_t sch_tiles = 0;
  const int64_t* sch_tile = nullptr;
  const int* sch_pos = nullptr;
  const int* schb = nullptr;
  int nschb = 0, schb_ncols = 0;
  int64_t schb_tiles = 0;
  const int64_t* schb_tile = nullptr;
  const int* schb_cols = nullptr;
  const int64_t* schb_ldoff = nullptr;
};

static LevelLists full_lists(const Symbolic& S, const FactorLevel& F) {
  LevelLists L;
  L.small = F.small.d;
  L.nsmall = F.small.count;
  L.small_wmax = F.small_wmax;
  L.medium = F.d_medium;
  L.nmedium = (int)F.medium.size();
  for (int s : F.medium) L.medium_wmax = std::max(L.medium_wmax, S.super[s + 1] - S.super[s]);
  for (size_t q = 0; q < F.large.size(); ++q) L.large_q.push_back((int)q);
  L.off = F.offdiag.d;
  L.noff = F.offdiag.count;
  L.off_wmax = F.offdiag_wmax;
  L.off_blocks = F.offdiag_blocks;
  L.off_blk = F.d_offdiag_blk;
  L.off_pos = F.d_offdiag_pos;
  L.sch = F.schur.d;
  L.nsch = F.schur.count;
  L.sch_tiles = F.schur_tiles;
  L.sch_tile = F.d_schur_tile;
  L.sch_pos = F.d_schur_pos;
  L.schb = F.schur_big.d;
  L.nschb = F.schur_big.count;
  L.schb_ncols = F.schur_big_ncols;
  L.schb_tiles = F.schur_big_tiles;
  L.schb_tile = F.d_schur_big_tile;
  L.schb_cols = F.d_schur_big_cols;
  L.schb_ldoff = F.d_schur_big_ldoff;
  return L;
}

// the launches of one factor level: Bunch-Kaufman of the level's diagonal
// blocks (narrow in LDS, medium panel by panel, wide by cooperative groups),
// the rows below them, and their Schur updates into the ancestors
template <typename T>
static void factor_level(Handle& h, FactorLevel& F, const LevelLists& L, const SnArgs& a, int gemm_schur,
                         int gemm_trailing) {
  DeviceFactor& D = *h.dev;
  const Symbolic& S = h.sym;
  T* const valL = static_cast<T*>(D.valL);
  T* const dval = static_cast<T*>(D.dval);
  T* const Wb = static_cast<T*>(D.W);
  T* const LDb = static_cast<T*>(D.LD);
  hipStream_t st = D.stream;
  timed(D, "bk_small", [&] { launch_bk_small(L.small, L.nsmall, L.small_wmax, a, valL, dval, D.perm_piv, D.status, st); });
  // medium supernodes: one workgroup each, panel by panel
  if (L.nmedium > 0) {
    const int wmax = L.medium_wmax;
    const int rounds = (wmax + PANEL_NB - 2) / (PANEL_NB - 1) + 1;
    // diagonal blocks of at most panel_lds_rows rows: the panel in LDS
    const int lds_rows = (D.panel_lds && wmax <= panel_lds_rows<T>()) ? wmax : 0;
    for (int p = 0; p < rounds; ++p) {
      timed(D, "bk_panel", [&] {
        launch_bk_panel(L.medium, L.nmedium, 1, a, valL, dval, D.perm_piv, Wb, D.woff, D.kpos, D.kstart, D.sync,
                        D.status, false, lds_rows, st);
      });
      const int lo = std::min(wmax, (p + 1) * (PANEL_NB - 1));
      const int nt = (wmax - lo + 63) / 64;
      if (nt > 0)
        timed(D, "trailing", [&] {
          launch_trailing(L.medium, L.nmedium, nt * (nt + 1) / 2, lo, a, valL, Wb, D.woff, D.kpos, D.kstart, st);
        });
    }
  }
  // large supernodes: a cooperative group of workgroups each, one at a time
  for (int q : L.large_q) {
    const int s = F.large[q];
    const int w = S.super[s + 1] - S.super[s];
    const int64_t r = S.pi[s + 1] - S.pi[s];
    const int G = F.large_G[q], nb = F.large_nb[q], nt = F.large_nt[q];
    const int rows = nt - 64;  // far rows per workgroup
    const int rounds = (w + nb - 2) / (nb - 1) + 1;
    const int wsub = F.large_wsub[q];
    const int wcols = wsub * nb;
    for (int p = 0; p < rounds; ++p) {
      // panels run in blocks of wsub: before panel p of a block, its columns
      // receive the block's earlier panels; the rest of the trailing block is
      // updated once at the end of the block (rank up to wsub * nb)
      const int sub = p % wsub;
      const int lo = std::min(w, p * (nb - 1));  // panel p starts at column >= lo
      if (sub > 0) {
        if (D.strip64 && (gemm_trailing == GEMM_DIRECT || gemm_trailing == GEMM_STAGED)) {
          const int T64 = (int)((r - lo + 63) / 64);
          if (T64 > 0)
            timed(D, "trailing", [&] {
              launch_strip_update(F.d_large + q, T64, nb, wcols, a, (const T*)valL, Wb, D.woff, D.kpos, D.kstart, st);
            });
        } else {
          const int Tr = (int)((r - lo + 127) / 128);
          if (Tr > 0)
            timed(D, "trailing", [&] {
              launch_trailing_big(F.d_large + q, Tr, lo, nb, wcols, a, valL, Wb, D.woff, D.kpos, D.kstart,
                                  gemm_trailing, st);
            });
        }
      }
      // far rows of this panel: r - min(w, k0 + nb) with k0 >= p (nb - 1) (each
      // panel eliminates >= nb - 1 columns); the last panel may be narrower
      const int far = (int)(r - std::min<int64_t>(w, (int64_t)p * (nb - 1) + nb));
      const int Gp = std::min(G, std::max(1, (far + rows - 1) / rows));
      const uint32_t eb = D.ll_epoch;
      D.ll_epoch += 3u * (uint32_t)nb + 8u;
      timed(D, "bk_panel_group", [&] {
        launch_bk_wide(F.d_large + q, Gp, nt, nb, sub, wcols, a, valL, dval, D.perm_piv, Wb, D.woff, D.kpos,
                       D.kstart, D.sync + D.wide_par, D.sync + (D.wide_par ^ 1), D.ll, eb, D.status, st);
      });
      D.wide_par ^= 1;
      if (sub == wsub - 1 || p == rounds - 1) {
        const int lo2 = std::min(w, (p + 1) * (nb - 1));
        const int Tj = (w - lo2 + 127) / 128, Ti = (int)((r - lo2 + 127) / 128);
        const int ntiles = Tj * (Tj + 1) / 2 + (Ti - Tj) * Tj;
        if (Tj > 0)
          timed(D, "trailing", [&] {
            launch_trailing_big(F.d_large + q, ntiles, lo2, 0, wcols, a, valL, Wb, D.woff, D.kpos, D.kstart,
                                gemm_trailing, st);
          });
      }
    }
  }
  timed(D, "offdiag", [&] {
    launch_offdiag(L.off, L.noff, L.off_wmax, L.off_blk, L.off_pos, L.off_blocks, a, valL, dval, D.perm_piv, st);
  });
  timed(D, "schur", [&] { launch_schur(L.sch, L.nsch, L.sch_tile, L.sch_pos, L.sch_tiles, a, valL, dval, st); });
  if (L.nschb > 0) {
    timed(D, "scale_ld", [&] {
      launch_scale_ld(L.schb, L.nschb, L.schb_cols, L.schb_ncols, a, valL, dval, LDb, L.schb_ldoff, st);
    });
    timed(D, "schur_big", [&] {
      launch_schur_big(L.schb, L.nschb, L.schb_tile, L.schb_tiles, a, valL, LDb, L.schb_ldoff, gemm_schur, st);
    });
  }
}

// static-pivoting threshold and GEMM operand modes of the handle's factor
struct FactorModes {
  double reg, reg_factor;
  int gemm_schur, gemm_trailing;
};
static FactorModes factor_modes(const Handle& h) {
  const DeviceFactor& D = *h.dev;
  FactorModes m;
  m.reg = h.set.reg_diag;
  // The FP32 and BF16 factors take S K S (S: max-norm diagonal scaling) and
  // perturb its diagonals below the static-pivoting threshold 2^pexp
  // (sqrt(eps_f32) by default; pivots of reg_diag = 1e-9 would give growth
  // ~1e9, beyond FP32); the refinement still runs on the reference's K0, so
  // only the preconditioner changes.  The entries of S K S have magnitude
  // <= 1 (max-norm scaling): the threshold is 2^pexp itself.
  const int pexp = env_int("LBLG_F32_PIVOT_EXP", -12);
  m.reg_factor = D.prec != 0 ? std::ldexp(1.0, pexp) : m.reg;
  // BF16 operands: split into two BF16 terms (three MFMAs per product) by
  // default, one rounded term with LBLG_BF16_TERMS=1
  const int bf16_mode = env_int("LBLG_BF16_TERMS", 3) == 1 ? GEMM_BF16 : GEMM_BF16X3;
  m.gemm_schur = D.prec == 2 ? bf16_mode : (D.schur_direct ? GEMM_DIRECT : GEMM_STAGED);
  m.gemm_trailing = D.prec == 2 ? bf16_mode : (D.trailing_direct ? GEMM_DIRECT : GEMM_STAGED);
  return m;
}

// status words of a finished factorization: zero pivots and the wide panels'
// counters; a bounded wait that ran out or a launch shape that did not fit
// ends the factorization with an error
static void finish_factor(Handle& h, bool partial) {
  DeviceFactor& D = *h.dev;
  hipStream_t st = D.stream;
  int stat[12] = {0};
  HIPCHK(hipMemcpyAsync(stat, D.status, sizeof(stat), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&h.k0_norm_inf, D.k0norm, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // a restricted refactorization reports the refactored supernodes' zero
  // pivots on top of those of the factorization it updates
  h.zero_pivot = (partial ? h.zero_pivot : 0) | (stat[0] & 1);
  h.second_col_steps = stat[2];
  h.interchanges = stat[3];
  h.wide_panels = stat[4];
  h.wide_panels_pivoted = stat[5];
  if (stat[0] & 4) {
    // status[8..11]: the first exchange wait that ran out (bk_wide.hip, LlDiag);
    // zero when a group barrier ran out instead
    char msg[256];
    std::snprintf(msg, sizeof(msg),
                  "cross-workgroup exchange timed out (wait site %d at word %d, workgroup %d, waited for flag %u, "
                  "saw %u, epoch base now %u)",
                  stat[8] & 15, stat[8] >> 4, stat[9], (unsigned)stat[10], (unsigned)stat[11], D.ll_epoch);
    throw std::runtime_error(msg);
  }
  if (stat[0] & 8) throw std::runtime_error("wide panel launch shape does not fit the supernode");
  h.factored = true;
  D.factor_valid = true;
  D.factor_key[0] = h.set.reg_diag;
  D.factor_key[1] = h.set.regularization;
  D.factor_key[2] = h.set.precision;
}

template <typename T>
static void factor_device_t(Handle& h) {
  DeviceFactor& D = *h.dev;
  T* const valL = static_cast<T*>(D.valL);
  T* const dval = static_cast<T*>(D.dval);
  const Symbolic& S = h.sym;
  const int n = h.n;
  hipStream_t st = D.stream;
  const FactorModes fm = factor_modes(h);
  SnArgs a = snargs(D);
  a.rows = D.rows0;  // pre-pivot labels during the factorization
  D.factor_valid = false;
  // the whole launch sequence depends only on the symbolic structure, the
  // buffers and (reg, reg_factor, regularization): captured once and replayed
  // when no cooperative wide panel (epoch-stamped) takes part
  auto seq = [&]() {
    launch_fill(valL, S.p[n], 0.0, st);
    launch_fill(dval, 2 * (int64_t)n, 0.0, st);
    launch_iota(D.perm_piv, n, st);
    launch_fill_int(D.status, 16, 0, st);
    launch_fill_int(D.kpos, D.nsuper, 0, st);
    const double* Kfac = D.Kx;  // the matrix the factor takes: K, or S K S for FP32 / BF16
    if (D.prec != 0) {
      launch_diag_scaling(n, D.nnzK, D.Krp, D.Ksrc, D.Krow, D.Kcol, D.Kx, D.scale, D.Kf, st);
      Kfac = D.Kf;
    }
    launch_pert_flags(D.nnzK, D.isdiag, D.Kx, Kfac, fm.reg, fm.reg_factor, h.set.regularization, D.pert, st);
    launch_assemble<T>(D.amap, Kfac, D.pert, D.nnzK, fm.reg_factor, valL, nullptr, nullptr, st);
    for (FactorLevel& F : D.flevels) factor_level<T>(h, F, full_lists(S, F), a, fm.gemm_schur, fm.gemm_trailing);
    launch_relabel(a, D.nsuper, D.rows0, D.rows, D.perm_piv, D.newpos, D.perm0, D.perm_f, D.iperm_f, st);
    if (D.prec != 0) launch_gather_scale(n, D.perm_f, D.scale, D.scale_ord, st);
    // unperturbed K0 (perturbed diagonals set to zero, set_diags(0)) for the SpMV
    launch_k0(n, D.nnzf, D.Krp, D.Ksrc, D.Kx, D.pert, D.Kv, D.red, 512, D.k0norm, st);
    launch_seg_rowabs_max(D.nseg, D.seg, D.Krp, D.Kv, D.segnorm, st);
  };
  if (D.has_coop) {
    seq();
  } else {
    run_graph(D, D.gfactor,
              graph_key({(double)D.prec, fm.reg, fm.reg_factor, (double)h.set.regularization}, {}), seq);
  }
  finish_factor(h, false);
  h.refactored = D.nsuper;
}

// staging of one restricted refactorization: the lists of every level and
// the changed values, packed host side and sent in one copy
struct Blob {
  std::vector<char> bytes;
  template <typename V>
  size_t add(const std::vector<V>& v) {
    const size_t off = (bytes.size() + 15) & ~(size_t)15;
    bytes.resize(off + sizeof(V) * v.size());
    if (!v.empty()) std::memcpy(bytes.data() + off, v.data(), sizeof(V) * v.size());
    return off;
  }
};

// The SoMod update (update_factorization with update_ldl_left_sn_02_v2,
// Serial_update_ldl.cpp:17-277; modified supernodes as add_del_matrix_qp
// collects them, linear_solver_wrapper.cpp:1058-1119): the supernodes on the
// tree paths from the changed entries to the root are zeroed, re-assembled
// from K and refactored level by level with the same kernels as a full
// factorization.  Supernodes off those paths keep their panels, D and
// pivots; the ones whose rows reach a refactored supernode send their Schur
// updates again, to the refactored supernodes only (tmask), from their kept
// factor.  Only the changed values travel to the device.
template <typename T>
static void path_factor_t(Handle& h) {
  DeviceFactor& D = *h.dev;
  const Symbolic& S = h.sym;
  const int n = h.n, ns = D.nsuper;
  hipStream_t st = D.stream;
  T* const valL = static_cast<T*>(D.valL);
  T* const dval = static_cast<T*>(D.dval);
  const FactorModes fm = factor_modes(h);
  // modified supernodes: the root paths from the supernode of the first
  // (in the ordering) of each changed entry's row and column
  if (++D.smark_epoch == 0) {
    std::fill(D.smark.begin(), D.smark.end(), 0u);
    D.smark_epoch = 1;
  }
  const unsigned ep = D.smark_epoch;
  std::vector<int> mod;
  std::vector<double> vals(h.changed.size());
  for (size_t k = 0; k < h.changed.size(); ++k) {
    const int64_t q = h.changed[k];
    vals[k] = h.Ax[q];
    for (int t = D.hasup[q]; t >= 0 && D.smark[t] != ep; t = S.sparent[t]) {
      D.smark[t] = ep;
      mod.push_back(t);
    }
  }
  // supernodes outside the paths with rows in a modified one: they keep
  // their factor and send their Schur updates again, to the modified
  // supernodes only.  Those updates depend on nothing this refactorization
  // computes, so the ones in 64-row tiles go out in one launch before the
  // first level; 128-row-tile senders stay in their level (their L D blocks
  // share the level's LD space).
  std::vector<unsigned char>& hmask = D.s_mark;
  std::vector<unsigned char>& sends = D.s_send;
  for (int t : mod) hmask[t] = 1;
  std::vector<int> senders;
  for (int t : mod)
    for (int e = D.upd_ptr[t]; e < D.upd_ptr[t + 1]; ++e) {
      const int d = D.upd_idx[e];
      if (!hmask[d] && !sends[d]) {
        sends[d] = 1;
        senders.push_back(d);
      }
    }
  const int L = (int)D.flevels.size();
  std::vector<std::vector<int>> lsmall(L), lmed(L), lqlarge(L), loff(L), lsch(L), lschb(L);
  std::vector<std::vector<int64_t>> lldoff(L);
  std::vector<int> lsch_send;
  std::vector<int> touched;
  auto touch = [&](int l) {
    if (lsmall[l].empty() && lmed[l].empty() && lqlarge[l].empty() && loff[l].empty() && lsch[l].empty() &&
        lschb[l].empty())
      touched.push_back(l);
  };
  for (int t : mod) {
    const int l = D.hlevel[t];
    touch(l);
    if (D.s_kind[t] == 0) lsmall[l].push_back(t);
    else if (D.s_kind[t] == 1) lmed[l].push_back(t);
    else lqlarge[l].push_back(D.s_large_q[t]);
    if (D.s_off[t]) loff[l].push_back(t);
    if (D.s_sch[t] == 1) lsch[l].push_back(t);
    if (D.s_sch[t] == 2) {
      lschb[l].push_back(t);
      lldoff[l].push_back(D.s_ldoff[t]);
    }
  }
  for (int d : senders) {
    if (D.s_sch[d] == 1) {
      lsch_send.push_back(d);
    } else if (D.s_sch[d] == 2) {
      const int l = D.hlevel[d];
      touch(l);
      lschb[l].push_back(d);
      lldoff[l].push_back(D.s_ldoff[d]);
    }
  }
  std::sort(touched.begin(), touched.end());
  std::vector<unsigned char> mask_bytes(ns, 0);
  for (int t : mod) mask_bytes[t] = 1;
  for (int t : mod) hmask[t] = 0;  // the marks are per refactorization
  for (int d : senders) sends[d] = 0;
  // pack: reset list, mask, changed entries, then per level the lists and
  // their prefixes (the same forms setup builds for full levels)
  Blob B;
  const size_t o_mod = B.add(mod), o_mask = B.add(mask_bytes), o_idx = B.add(h.changed), o_val = B.add(vals);
  struct Offs {
    size_t small, med, off, offblk, offpos, sch, schtile, schpos, schb, schbtile, schbcols, schbld;
    int off_blocks = 0, schb_ncols = 0, small_wmax = 0, med_wmax = 0, off_wmax = 0;
    int64_t sch_tiles = 0, schb_tiles = 0;
  };
  std::vector<Offs> O(L);
  auto width = [&](int s) { return S.super[s + 1] - S.super[s]; };
  auto offrows = [&](int s) { return S.pi[s + 1] - S.pi[s] - width(s); };
  // the senders' 64-row-tile updates: one list, tile prefix and tile map
  int64_t send_tiles = 0;
  size_t o_send, o_sendtile, o_sendpos;
  {
    std::vector<int64_t> tiles;
    std::vector<int> tpos;
    for (size_t k = 0; k < lsch_send.size(); ++k) {
      tiles.push_back(send_tiles);
      const int64_t T64 = (offrows(lsch_send[k]) + 63) / 64, nt = T64 * (T64 + 1) / 2;
      for (int64_t b = 0; b < nt; ++b) tpos.push_back((int)k);
      send_tiles += nt;
    }
    o_send = B.add(lsch_send);
    o_sendtile = B.add(tiles);
    o_sendpos = B.add(tpos);
  }
  for (int l : touched) {
    Offs& o = O[l];
    for (int s : lsmall[l]) o.small_wmax = std::max(o.small_wmax, width(s));
    for (int s : lmed[l]) o.med_wmax = std::max(o.med_wmax, width(s));
    o.small = B.add(lsmall[l]);
    o.med = B.add(lmed[l]);
    std::vector<int> blk, pos;
    for (size_t k = 0; k < loff[l].size(); ++k) {
      const int s = loff[l][k];
      o.off_wmax = std::max(o.off_wmax, width(s));
      blk.push_back(o.off_blocks);
      const int nb = (int)((offrows(s) + 63) / 64);
      for (int b = 0; b < nb; ++b) pos.push_back((int)k);
      o.off_blocks += nb;
    }
    o.off = B.add(loff[l]);
    o.offblk = B.add(blk);
    o.offpos = B.add(pos);
    std::vector<int64_t> tiles;
    std::vector<int> tpos;
    for (size_t k = 0; k < lsch[l].size(); ++k) {
      tiles.push_back(o.sch_tiles);
      const int64_t T64 = (offrows(lsch[l][k]) + 63) / 64, nt = T64 * (T64 + 1) / 2;
      for (int64_t b = 0; b < nt; ++b) tpos.push_back((int)k);
      o.sch_tiles += nt;
    }
    o.sch = B.add(lsch[l]);
    o.schtile = B.add(tiles);
    o.schpos = B.add(tpos);
    std::vector<int64_t> tilesb;
    std::vector<int> colsb;
    for (int s : lschb[l]) {
      tilesb.push_back(o.schb_tiles);
      colsb.push_back(o.schb_ncols);
      o.schb_ncols += width(s);
      const int64_t T128 = (offrows(s) + 127) / 128;
      o.schb_tiles += T128 * (T128 + 1) / 2;
    }
    o.schb = B.add(lschb[l]);
    o.schbtile = B.add(tilesb);
    o.schbcols = B.add(colsb);
    o.schbld = B.add(lldoff[l]);
  }
  const size_t bytes = B.bytes.size();
  if (bytes > D.blob_cap) {
    if (D.blob_d) HIPCHK(hipFree(D.blob_d));
    if (D.blob_h) HIPCHK(hipHostFree(D.blob_h));
    D.blob_d = nullptr;
    D.blob_h = nullptr;
    const size_t cap = std::max<size_t>(bytes + bytes / 2, 1 << 16);
    HIPCHK(hipMalloc((void**)&D.blob_d, cap));
    HIPCHK(hipHostMalloc((void**)&D.blob_h, cap, hipHostMallocDefault));
    D.blob_cap = cap;
  }
  std::memcpy(D.blob_h, B.bytes.data(), bytes);
  char* const bd = D.blob_d;
  auto dptr = [&](size_t off) { return bd + off; };
  HIPCHK(hipMemcpyAsync(bd, D.blob_h, bytes, hipMemcpyHostToDevice, st));
  D.factor_valid = false;
  SnArgs a = snargs(D);
  a.rows = D.rows0;
  launch_fill_int(D.status, 16, 0, st);
  launch_fill_int(D.kpos, ns, 0, st);
  launch_scatter_values((int64_t)h.changed.size(), reinterpret_cast<const int64_t*>(dptr(o_idx)),
                        reinterpret_cast<const double*>(dptr(o_val)), D.Kx, st);
  const int* dmod = reinterpret_cast<const int*>(dptr(o_mod));
  const unsigned char* dmask = reinterpret_cast<const unsigned char*>(dptr(o_mask));
  launch_reset_supernodes<T>(dmod, (int)mod.size(), a, valL, dval, D.perm_piv, st);
  const double* Kfac = D.Kx;
  if (D.prec != 0) {  // S K S with the scaling of the last full factorization
    launch_scale_entries(D.nnzK, D.Krow, D.Kcol, D.Kx, D.scale, D.Kf, st);
    Kfac = D.Kf;
  }
  launch_pert_flags(D.nnzK, D.isdiag, D.Kx, Kfac, fm.reg, fm.reg_factor, h.set.regularization, D.pert, st);
  launch_assemble<T>(D.amap, Kfac, D.pert, D.nnzK, fm.reg_factor, valL, D.asup, dmask, st);
  a.tmask = dmask;
  timed(D, "schur", [&] {
    launch_schur(reinterpret_cast<const int*>(dptr(o_send)), (int)lsch_send.size(),
                 reinterpret_cast<const int64_t*>(dptr(o_sendtile)), reinterpret_cast<const int*>(dptr(o_sendpos)),
                 send_tiles, a, valL, dval, st);
  });
  for (int l : touched) {
    const Offs& o = O[l];
    LevelLists LL;
    LL.small = reinterpret_cast<const int*>(dptr(o.small));
    LL.nsmall = (int)lsmall[l].size();
    LL.small_wmax = o.small_wmax;
    LL.medium = reinterpret_cast<const int*>(dptr(o.med));
    LL.nmedium = (int)lmed[l].size();
    LL.medium_wmax = o.med_wmax;
    LL.large_q = lqlarge[l];
    LL.off = reinterpret_cast<const int*>(dptr(o.off));
    LL.noff = (int)loff[l].size();
    LL.off_wmax = o.off_wmax;
    LL.off_blocks = o.off_blocks;
    LL.off_blk = reinterpret_cast<const int*>(dptr(o.offblk));
    LL.off_pos = reinterpret_cast<const int*>(dptr(o.offpos));
    LL.sch = reinterpret_cast<const int*>(dptr(o.sch));
    LL.nsch = (int)lsch[l].size();
    LL.sch_tiles = o.sch_tiles;
    LL.sch_tile = reinterpret_cast<const int64_t*>(dptr(o.schtile));
    LL.sch_pos = reinterpret_cast<const int*>(dptr(o.schpos));
    LL.schb = reinterpret_cast<const int*>(dptr(o.schb));
    LL.nschb = (int)lschb[l].size();
    LL.schb_ncols = o.schb_ncols;
    LL.schb_tiles = o.schb_tiles;
    LL.schb_tile = reinterpret_cast<const int64_t*>(dptr(o.schbtile));
    LL.schb_cols = reinterpret_cast<const int*>(dptr(o.schbcols));
    LL.schb_ldoff = reinterpret_cast<const int64_t*>(dptr(o.schbld));
    factor_level<T>(h, D.flevels[l], LL, a, fm.gemm_schur, fm.gemm_trailing);
  }
  a.tmask = nullptr;
  launch_relabel(a, ns, D.rows0, D.rows, D.perm_piv, D.newpos, D.perm0, D.perm_f, D.iperm_f, st);
  if (D.prec != 0) launch_gather_scale(n, D.perm_f, D.scale, D.scale_ord, st);
  launch_k0(n, D.nnzf, D.Krp, D.Ksrc, D.Kx, D.pert, D.Kv, D.red, 512, D.k0norm, st);
  launch_seg_rowabs_max(D.nseg, D.seg, D.Krp, D.Kv, D.segnorm, st);
  finish_factor(h, true);  // synchronizes: the staging buffers are free again
  h.refactored = (int64_t)mod.size();
}

void factor(Handle& h) {
  if (h.dev && h.dev->prec != prec_code(h.set.precision)) h.dev.reset();  // value type changed
  if (!h.dev) {
    setup(h);
    h.values_dirty = true;
  }
  DeviceFactor& D = *h.dev;
  double md = 0.0;
  for (double v : h.Ax) md = std::max(md, std::fabs(v));
  h.max_abs_entry = md;
  // a restricted refactorization updates the device factor of the values
  // last uploaded (same precision and perturbation settings) after
  // lblg_update_values changes; the FP32 / BF16 factors keep the diagonal
  // scaling of the last full factorization (any positive scaling gives a
  // valid preconditioner; refinement runs on K itself).  Without changes a
  // factorization recomputes everything, as the reference's
  // numerical_factorization does.
  const bool partial = D.path_refactor && !h.values_dirty && !h.changed.empty() && D.factor_valid &&
                       D.factor_key[0] == h.set.reg_diag && D.factor_key[1] == h.set.regularization &&
                       D.factor_key[2] == h.set.precision;
  if (partial) {
    if (D.prec != 0)
      path_factor_t<float>(h);
    else
      path_factor_t<double>(h);
  } else {
    HIPCHK(hipMemcpyAsync(D.Kx, h.Ax.data(), sizeof(double) * D.nnzK, hipMemcpyHostToDevice, D.stream));
    factor_device(h);
  }
  for (int64_t q : h.changed) h.changed_flag[q] = 0;
  h.changed.clear();
  h.values_dirty = false;
}

void factor_device(Handle& h) {
  if (h.dev->prec != 0)
    factor_device_t<float>(h);
  else
    factor_device_t<double>(h);
}

// ---------------------------------------------------------------------------
// M^{-1}: P L^{-T} D^{-1} L^{-1} P^T applied to an ordered block (in -> out)
// ---------------------------------------------------------------------------
static void ensure_solve_buffers(DeviceFactor& D, int nrhs, int mr) {
  if (D.nrhs_cap >= nrhs && D.mr_cap >= mr && D.X) return;
  const int64_t nn = (int64_t)D.n * std::max(nrhs, D.nrhs_cap);
  const int cap = std::max(nrhs, D.nrhs_cap);
  const int mrc = std::max(mr, std::max(D.mr_cap, 1));
  D.X = valloc(D, nn);
  D.Y = valloc(D, nn);
  D.Z = valloc(D, nn);
  D.acc = valloc(D, nn);
  D.O = valloc(D, nn);
  D.R = dalloc<double>(D, nn);
  D.B = dalloc<double>(D, nn);
  D.T = dalloc<double>(D, nn);
  D.Xs = dalloc<double>(D, nn);
  D.V = dalloc<double>(D, nn * (mrc + 1));
  D.gm = dalloc<double>(D, gmres_state_size(mrc) * cap);
  D.nblk_red = 512;
  D.part = dalloc<double>(D, (int64_t)D.nblk_red * 3 * cap * (mrc + 2));
  D.dots = dalloc<double>(D, 3 * cap * (mrc + 2));
  D.acc2 = dalloc<double>(D, nn);
  D.stage_b = dalloc<double>(D, nn);
  D.stage_x = dalloc<double>(D, nn);
  if (D.prec != 0) D.sbuf = dalloc<double>(D, nn);
  D.nrhs_cap = cap;
  D.mr_cap = mrc;
}

// sweeps in the factor's type T; in / out are the FP64 ordered blocks of the
// refinement (converted at entry and exit when T is float)
// whether a solve of nrhs columns takes the one-workgroup sweeps
// (a block beyond the LDS takes the level sweeps unless LBLG_SMALL_SOLVE_GLOBAL
// asks for the one-workgroup sweeps with the block in memory: measured on a
// 21k-row QP KKT, the level sweeps' many workgroups were faster, 1.5 vs 2.5 ms)
static bool small_sweeps(const DeviceFactor& D, int nrhs) {
  const size_t vs = D.prec != 0 ? sizeof(float) : sizeof(double);
  return D.ss_nlev > 0 && D.ss_nnzL <= D.ss_max_nnz &&
         (D.ss_global != 0 || vs * (size_t)D.n * nrhs <= SOLVE_SMALL_LDS);
}

template <typename T>
static void apply_solve_t(DeviceFactor& D, const double* in, double* out, int nrhs, const double* active,
                          int64_t astride) {
  hipStream_t st = D.stream;
  SnArgs a = snargs(D);
  const int64_t nn = (int64_t)D.n * nrhs;
  const T* valL = static_cast<const T*>(D.valL);
  const T* dval = static_cast<const T*>(D.dval);
  T* X = static_cast<T*>(D.X);
  T* Y = static_cast<T*>(D.Y);
  T* Z = static_cast<T*>(D.Z);
  T* acc = static_cast<T*>(D.acc);
  // the FP32 / BF16 factors are of S K S: M^-1 = S (L D L^T)^-1 S in the pivoted order
  const bool scaled = D.prec != 0;
  if (D.ssb_nsys > 0 && nrhs <= MAX_RHS) {  // a batch: one workgroup per system
    const double* sin = in;
    if (scaled) {
      launch_row_scale_copy(D.sbuf, in, D.scale_ord, D.n, nrhs, st);
      sin = D.sbuf;
    }
    // the largest systems whose blocks fit the LDS set the launch's LDS; larger ones use Z
    const size_t cap = std::min<size_t>(SOLVE_SMALL_LDS, (size_t)std::max(1, D.ssb_lds_kb) * 1024);
    const size_t lim = cap / (sizeof(T) * (size_t)nrhs);
    auto it = std::upper_bound(D.ssb_cols.begin(), D.ssb_cols.end(), (int)std::min<size_t>(lim, INT32_MAX));
    const int fit_cols = it == D.ssb_cols.begin() ? 0 : *(it - 1);
    timed(D, "solve_batch", [&] {
      launch_solve_batch<T>(D.ssb_sysd, D.ssb_nsys, D.ssb_info, D.ssb_ipx, D.ssb_levptr, D.ssb_mode, fit_cols,
                            D.ssb_maxlev, a, valL, dval, sin, out, nrhs, active, astride, Z, st);
    });
    if (scaled) launch_row_scale_copy(out, out, D.scale_ord, D.n, nrhs, st);
    return;
  }
  if (small_sweeps(D, nrhs)) {
    const double* sin = in;
    if (scaled) {
      launch_row_scale_copy(D.sbuf, in, D.scale_ord, D.n, nrhs, st);
      sin = D.sbuf;
    }
    // the ordered block in LDS, or in memory (Z) when asked for or when it
    // does not fit (small_sweeps admits that only when asked for)
    T* xg = (D.ss_global == 1 || sizeof(T) * (size_t)D.n * nrhs > SOLVE_SMALL_LDS) ? Z : nullptr;
    timed(D, "solve_small", [&] {
      launch_solve_small<T>(D.ss_info, D.ss_ipx, D.ss_levptr, D.ss_mode, D.ss_nlev, a, static_cast<const T*>(D.valL),
                            static_cast<const T*>(D.dval), sin, out, nrhs, active, astride, xg, st);
    });
    if (scaled) launch_row_scale_copy(out, out, D.scale_ord, D.n, nrhs, st);
    return;
  }
  T* O;
  if constexpr (std::is_same<T, double>::value) {
    launch_copy(X, in, nn, st);
    O = out;
  } else {
    if (scaled)
      launch_row_scale_copy(X, in, D.scale_ord, D.n, nrhs, st);
    else
      launch_convert(X, in, nn, st);
    O = static_cast<T*>(D.O);
  }
  for (size_t l = 0; l < D.slevels.size(); ++l) {
    const SolveLevel& L = D.slevels[l];
    if (L.npieces > 0) {
      timed(D, "fwd_tri", [&] {
        if (L.nnarrow > 0)
          launch_fwd_tri(L.d_piece, L.nnarrow, L.maxpw_narrow, D.piece_sn, D.piece_c0, a, valL, (const T*)X, Y, nrhs,
                         st);
        if (L.npieces > L.nnarrow)
          launch_fwd_tri(L.d_piece + L.nnarrow, L.npieces - L.nnarrow, L.maxpw, D.piece_sn, D.piece_c0, a, valL,
                         (const T*)X, Y, nrhs, st);
      });
      timed(D, "fwd_gemv", [&] {
        launch_fwd_gemv(L.d_piece, L.d_blk, L.d_blkpos, L.nblocks, D.piece_sn, D.piece_c0, a, valL, (const T*)Y, X,
                        nrhs, st);
      });
    }
    for (int s : L.chains) {
      const int w = D.hsuper[s + 1] - D.hsuper[s];
      const int64_t r = D.hpi[s + 1] - D.hpi[s];
      const uint32_t ep = ++D.chain_epoch;
      timed(D, "fwd_chain", [&] {
        launch_fwd_chain(s, w, r, a, valL, X, Y, nrhs, D.chain_xb, ep, D.status, st);
      });
    }
  }
  launch_dsolve(D.n, dval, (const T*)Y, Z, nrhs, st);
  launch_fill(acc, nn, 0.0, st);
  for (size_t l = D.slevels.size(); l-- > 0;) {
    const SolveLevel& L = D.slevels[l];
    for (int s : L.chains) {
      const int w = D.hsuper[s + 1] - D.hsuper[s];
      const int64_t r = D.hpi[s + 1] - D.hpi[s];
      const uint32_t ep = ++D.chain_epoch;
      timed(D, "bwd_chain", [&] {
        launch_bwd_chain(s, w, r, a, valL, (const T*)Z, acc, O, nrhs, D.chain_xb, ep, D.status, st);
      });
    }
    if (L.npieces > 0) {
      timed(D, "bwd_gemv", [&] {
        // narrow pieces apart, with LDS sized for them (as the triangles)
        if (L.nblocks_narrow > 0)
          launch_bwd_gemv(L.d_piece, L.maxpw_narrow, L.d_blk, L.d_blkpos, L.nblocks_narrow, 0, D.piece_sn,
                          D.piece_c0, a, valL, (const T*)O, acc, nrhs, st);
        if (L.nblocks > L.nblocks_narrow)
          launch_bwd_gemv(L.d_piece, L.maxpw, L.d_blk, L.d_blkpos, L.nblocks - L.nblocks_narrow, L.nblocks_narrow,
                          D.piece_sn, D.piece_c0, a, valL, (const T*)O, acc, nrhs, st);
      });
      timed(D, "bwd_tri", [&] {
        if (L.nnarrow > 0)
          launch_bwd_tri(L.d_piece, L.nnarrow, L.maxpw_narrow, D.piece_sn, D.piece_c0, a, valL, (const T*)Z,
                         (const T*)acc, O, nrhs, st);
        if (L.npieces > L.nnarrow)
          launch_bwd_tri(L.d_piece + L.nnarrow, L.npieces - L.nnarrow, L.maxpw, D.piece_sn, D.piece_c0, a, valL,
                         (const T*)Z, (const T*)acc, O, nrhs, st);
      });
    }
  }
  if constexpr (!std::is_same<T, double>::value) {
    if (scaled)
      launch_row_scale_copy(out, (const T*)O, D.scale_ord, D.n, nrhs, st);
    else
      launch_convert(out, (const T*)O, nn, st);
  }
}

// active: see launch_solve_small (only the one-workgroup sweeps use it; the
// level sweeps always run)
void apply_solve(DeviceFactor& D, const double* in, double* out, int nrhs, const double* active, int64_t astride) {
  if (D.prec != 0)
    apply_solve_t<float>(D, in, out, nrhs, active, astride);
  else
    apply_solve_t<double>(D, in, out, nrhs, active, astride);
}

static void spmv_ordered(DeviceFactor& D, const double* x, double* y, int nrhs) {
  timed(D, "spmv", [&] { launch_spmv(D.n, D.Krp, D.Kci, D.Kv, D.perm_f, D.iperm_f, x, y, nrhs, D.stream); });
}


// one GMRES(mr) cycle per column, pmgmres_ldlt_auto (mgmres.cpp:1015-1211)
static void gmres_cycle(DeviceFactor& D, const double* b, double* x, int nrhs, int mr, int itr,
                        double tol_abs, double tol_rel) {
  hipStream_t st = D.stream;
  const int n = D.n;
  const int64_t nn = (int64_t)n * nrhs;
  const int MR = D.mr_cap;
  const int64_t SS = gmres_state_size(MR);
  double* mask = D.gm + gmres_off_S(MR) + 7;
  double* coef0 = D.gm + gmres_off_D(MR);
  double* reo = D.gm + gmres_off_D(MR) + 1;
  auto Vk = [&](int k) { return D.V + (int64_t)k * nn; };
  spmv_ordered(D, x, D.T, nrhs);
  launch_sub(nn, b, D.T, D.R, st);
  apply_solve(D, D.R, D.Xs, nrhs, nullptr, 0);
  launch_dot_scalar(n, nrhs, D.Xs, D.Xs, D.part, D.nblk_red, D.dots, 0, MR, D.gm, 0, 0,
                    itr, tol_abs, tol_rel, st);
  launch_fill(D.V, nn * (MR + 1), 0.0, st);
  launch_axpby(n, nrhs, coef0, (int)SS, D.Xs, Vk(0), mask, (int)SS, 0, 1, st);
  for (int k = 0; k < mr; ++k) {
    spmv_ordered(D, Vk(k), D.T, nrhs);
    apply_solve(D, D.T, Vk(k + 1), nrhs, mask, SS);  // columns out of the step keep V(k + 1) = 0
    launch_dot_scalar(n, nrhs, Vk(k + 1), Vk(k + 1), D.part, D.nblk_red, D.dots, 1, MR, D.gm, 0, k,
                      itr, tol_abs, tol_rel, st);
    for (int j = 0; j <= k; ++j) {
      launch_dot_scalar(n, nrhs, Vk(k + 1), Vk(j), D.part, D.nblk_red, D.dots, 2, MR, D.gm, j, k,
                        itr, tol_abs, tol_rel, st);
      launch_axpby(n, nrhs, coef0, (int)SS, Vk(j), Vk(k + 1), mask, (int)SS, 0, 0, st);
    }
    launch_dot_scalar(n, nrhs, Vk(k + 1), Vk(k + 1), D.part, D.nblk_red, D.dots, 3, MR, D.gm, 0, k,
                      itr, tol_abs, tol_rel, st);
    for (int j = 0; j <= k; ++j) {  // conditional re-orthogonalisation (flag on device)
      launch_dot_scalar(n, nrhs, Vk(k + 1), Vk(j), D.part, D.nblk_red, D.dots, 4, MR, D.gm, j, k,
                        itr, tol_abs, tol_rel, st);
      launch_axpby(n, nrhs, coef0, (int)SS, Vk(j), Vk(k + 1), reo, (int)SS, 0, 0, st);
    }
    launch_dot_scalar(n, nrhs, Vk(k + 1), Vk(k + 1), D.part, D.nblk_red, D.dots, 5, MR, D.gm, 0, k,
                      itr, tol_abs, tol_rel, st);
    launch_axpby(n, nrhs, coef0, (int)SS, Vk(k + 1), Vk(k + 1), mask, (int)SS, 0, 1, st);
    launch_gmres_scalar(6, nrhs, MR, D.gm, D.dots, 0, k, itr, tol_abs, tol_rel, st);
  }
  launch_gmres_scalar(7, nrhs, MR, D.gm, D.dots, 0, 0, itr, tol_abs, tol_rel, st);
  launch_gmres_update(n, nrhs, MR, D.V, D.gm, x, st);
}

// backward error per column in the ordered space (permutation invariant):
// ||b - K0 x||_1 / (||K0||_inf ||x||_1 + ||b||_1)
static void backward_errors(Handle& h, const double* b, const double* x, int nrhs, double* err) {
  DeviceFactor& D = *h.dev;
  const int64_t nn = (int64_t)D.n * nrhs;
  spmv_ordered(D, x, D.T, nrhs);
  launch_sub(nn, b, D.T, D.R, D.stream);
  if (D.nseg > 0) {
    // a block-diagonal batch: every system's own backward error, the worst per column
    // (the systems occupy the same contiguous row ranges in both orderings)
    launch_seg_abs_sums(D.nseg, D.seg, nrhs, D.R, x, b, D.segsums, D.stream);
    std::vector<double> sums((size_t)D.nseg * 3 * nrhs), nrm(D.nseg);
    HIPCHK(hipMemcpyAsync(sums.data(), D.segsums, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, D.stream));
    HIPCHK(hipMemcpyAsync(nrm.data(), D.segnorm, sizeof(double) * nrm.size(), hipMemcpyDeviceToHost, D.stream));
    HIPCHK(hipStreamSynchronize(D.stream));
    for (int c = 0; c < nrhs; ++c) {
      double worst = 0.0;
      for (int k = 0; k < D.nseg; ++k) {
        const double* sk = sums.data() + (size_t)k * 3 * nrhs;
        const double den = nrm[k] * sk[nrhs + c] + sk[2 * nrhs + c];
        worst = std::max(worst, den > 0 ? sk[c] / den : 0.0);
      }
      err[c] = worst;
    }
    return;
  }
  launch_abs_sums(D.n, nrhs, D.R, x, b, D.part, D.nblk_red, D.dots, D.stream);
  std::vector<double> sums(3 * nrhs);
  HIPCHK(hipMemcpyAsync(sums.data(), D.dots, sizeof(double) * 3 * nrhs, hipMemcpyDeviceToHost, D.stream));
  HIPCHK(hipStreamSynchronize(D.stream));
  const double an = h.k0_norm_inf;
  for (int c = 0; c < nrhs; ++c) {
    const double den = an * sums[nrhs + c] + sums[2 * nrhs + c];
    err[c] = den > 0 ? sums[c] / den : 0.0;
  }
}


// refinement of the ordered solution xo of the ordered block D.B to a target
// backward error: GMRES cycles until every column meets it (host decision
// after each cycle; the cycle itself goes through the graph cache)
static void refine_to_target(Handle& h, double* xo, int nrhs, bool capturable, double* bwd_out) {
  DeviceFactor& D = *h.dev;
  const lblg_settings& s = h.set;
  const int mr = std::max(s.max_inner_iter, 1);
  hipStream_t st = D.stream;
  std::vector<double> err(nrhs);
  backward_errors(h, D.B, xo, nrhs, err.data());
  int rounds = 0;
  auto cycle = [&]() {
    // each round: one reference GMRES cycle (itr_max = 1 with mr inner steps)
    launch_gmres_reset(D.gm, gmres_state_size(D.mr_cap), nrhs, gmres_off_S(D.mr_cap) + 3, st);
    gmres_cycle(D, D.B, xo, nrhs, mr, 0, 0.0, 0.0);
  };
  while (*std::max_element(err.begin(), err.end()) > s.target_bwd_err && rounds < s.max_refine_rounds) {
    if (capturable)
      run_graph(D, D.gcycle, graph_key({(double)nrhs, (double)mr, 1.0}, {xo, D.X, D.V, D.gm, D.part}), cycle);
    else
      cycle();
    ++rounds;
    backward_errors(h, D.B, xo, nrhs, err.data());
  }
  h.num_ref_iter = rounds;
  if (bwd_out) std::copy(err.begin(), err.end(), bwd_out);
}

// a chain sweep whose bounded wait ran out leaves status bit 16
static void check_solve_status(DeviceFactor& D) {
  int st0 = 0;
  HIPCHK(hipMemcpyAsync(&st0, D.status, sizeof(int), hipMemcpyDeviceToHost, D.stream));
  HIPCHK(hipStreamSynchronize(D.stream));
  if (st0 & 16) throw std::runtime_error("chain sweep: cross-workgroup wait timed out");
}

void solve(Handle& h, int nrhs, const double* B, int64_t ldb, double* X, int64_t ldx, bool device_ptrs,
           double* bwd_out) {
  DeviceFactor& D = *h.dev;
  const lblg_settings& s = h.set;
  const int mr = std::max(s.max_inner_iter, 1);
  ensure_solve_buffers(D, nrhs, mr);
  hipStream_t st = D.stream;
  const double* Bd = B;
  double* Xd = X;
  if (!device_ptrs) {  // staged through the handle's blocks (sized with the rhs capacity)
    if (ldb == D.n) {
      HIPCHK(hipMemcpyAsync(D.stage_b, B, sizeof(double) * D.n * nrhs, hipMemcpyHostToDevice, st));
    } else {
      for (int c = 0; c < nrhs; ++c)
        HIPCHK(hipMemcpyAsync(D.stage_b + (int64_t)c * D.n, B + (int64_t)c * ldb, sizeof(double) * D.n,
                              hipMemcpyHostToDevice, st));
    }
    Bd = D.stage_b;
    Xd = D.stage_x;
  }
  const int64_t lb = device_ptrs ? ldb : D.n, lx = device_ptrs ? ldx : D.n;
  double* xo = D.acc2;
  // status word 0 reports this solve's chain-sweep time-outs only (the
  // factor's flags were read back when it ended)
  launch_fill_int(D.status, 1, 0, st);
  const bool to_target = s.target_bwd_err > 0;
  const int fixed_cycles = to_target ? 0 : s.req_ref_iter;
  // every launch below is fixed by (nrhs, refinement budget and tolerances,
  // buffers); the level sweeps of chained supernodes are epoch-stamped and
  // keep the eager path
  const bool capturable = !D.has_chains || small_sweeps(D, nrhs);
  auto head = [&]() {
    launch_gather_rhs(D.n, nrhs, Bd, lb, D.perm_f, D.B, st);
    apply_solve(D, D.B, xo, nrhs, nullptr, 0);
    if (fixed_cycles > 0) {
      // zero state, every column marked running
      launch_gmres_reset(D.gm, gmres_state_size(D.mr_cap), nrhs, gmres_off_S(D.mr_cap) + 3, st);
      for (int itr = 0; itr < fixed_cycles; ++itr) gmres_cycle(D, D.B, xo, nrhs, mr, itr, s.tol_abs, s.tol_rel);
    }
    if (!to_target) launch_scatter_sol(D.n, nrhs, xo, D.perm_f, Xd, lx, st);
  };
  if (capturable)
    run_graph(D, D.gsolve,
              graph_key({(double)nrhs, (double)mr, (double)fixed_cycles, s.tol_abs, s.tol_rel, (double)lb,
                         (double)lx, to_target ? 1.0 : 0.0},
                        {Bd, Xd, xo, D.X, D.V, D.gm, D.part}),
              head);
  else
    head();
  h.num_ref_iter = fixed_cycles;
  if (to_target) {
    refine_to_target(h, xo, nrhs, capturable, bwd_out);
    launch_scatter_sol(D.n, nrhs, xo, D.perm_f, Xd, lx, st);
  } else if (bwd_out) {
    backward_errors(h, D.B, xo, nrhs, bwd_out);
  }
  if (!device_ptrs) {
    if (ldx == D.n) {
      HIPCHK(hipMemcpyAsync(X, D.stage_x, sizeof(double) * D.n * nrhs, hipMemcpyDeviceToHost, st));
    } else {
      for (int c = 0; c < nrhs; ++c)
        HIPCHK(hipMemcpyAsync(X + (int64_t)c * ldx, D.stage_x + (int64_t)c * D.n, sizeof(double) * D.n,
                              hipMemcpyDeviceToHost, st));
    }
  }
  check_solve_status(D);
}

// ---------------------------------------------------------------------------
// autotuner of the device schedule
// ---------------------------------------------------------------------------
// factor and solve times (ms, best of reps after two warm runs, which also
// capture the graphs) of the schedule in h.tune
static void time_schedule(Handle& h, int nrhs, int reps, const std::vector<double>& B, std::vector<double>& X,
                          double& fms, double& sms) {
  h.dev.reset();
  factor(h);  // device setup and first factor, untimed
  for (int w = 0; w < 2; ++w) {
    factor_device(h);
    solve(h, nrhs, B.data(), h.n, X.data(), h.n, false, nullptr);
  }
  fms = sms = 1e300;
  for (int r = 0; r < reps; ++r) {
    auto t0 = std::chrono::steady_clock::now();
    factor_device(h);  // reads back the status word: ends synchronised
    auto t1 = std::chrono::steady_clock::now();
    solve(h, nrhs, B.data(), h.n, X.data(), h.n, false, nullptr);  // synchronised on return
    auto t2 = std::chrono::steady_clock::now();
    fms = std::min(fms, std::chrono::duration<double, std::milli>(t1 - t0).count());
    sms = std::min(sms, std::chrono::duration<double, std::milli>(t2 - t1).count());
  }
}

void autotune(Handle& h, int nrhs, int reps, double& factor_ms, double& solve_ms, int& tried) {
  resolve_tuning(h);
  const Symbolic& S = h.sym;
  int wmax = 0;
  for (int s = 0; s < S.nsuper; ++s) wmax = std::max(wmax, S.super[s + 1] - S.super[s]);
  const int64_t nnzL = S.p[h.n];
  std::vector<double> B((size_t)h.n * nrhs), X((size_t)h.n * nrhs);
  uint64_t seed = 0x9e3779b97f4a7c15ull;
  for (double& v : B) {  // fixed uniform(-1, 1) right-hand sides
    seed = seed * 6364136223846793005ull + 1442695040888963407ull;
    v = (double)(seed >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
  tried = 0;
  auto attempt = [&](const Handle::Tuning& t, double& f, double& sl) {
    h.tune = t;
    try {
      time_schedule(h, nrhs, reps, B, X, f, sl);
      ++tried;
      return true;
    } catch (const std::exception&) {  // a schedule the matrix does not admit
      h.dev.reset();
      return false;
    }
  };
  // factor schedule, one knob at a time (coordinate descent), timed by the
  // factor: panel group width (only values below the widest supernode change
  // the partition), then the Schur tile threshold, then the rank of the wide
  // panels' trailing updates where cooperative panels remain
  Handle::Tuning best = h.tune;
  double bf = 1e300, bs = 1e300;
  auto consider = [&](const Handle::Tuning& t) {
    double f, sl;
    if (attempt(t, f, sl) && f < bf) {
      bf = f;
      bs = sl;
      best = t;
    }
  };
  consider(best);
  for (int g : {128, 256, 512, 1024, 2048, 4096})
    if (g != best.group_w && g < wmax) {
      Handle::Tuning t = best;
      t.group_w = g;
      consider(t);
    }
  for (int m : {128, 256, 512})
    if (m != best.big_m) {
      Handle::Tuning t = best;
      t.big_m = m;
      consider(t);
    }
  if (wmax > best.group_w) {
    const int cur = best.wide_sub;
    for (int ws : {2, 4, 8})
      if (ws != cur) {
        Handle::Tuning t = best;
        t.wide_sub = ws;
        consider(t);
      }
  }
  // solve schedule: chain width and the one-workgroup sweep limit, timed by the solve
  const Handle::Tuning fbest = best;
  std::vector<int> cws = {fbest.chain_w};
  for (int c : {256, 512, 1024, 1 << 30})  // 1 << 30: no chain sweeps (only tried where some exist now)
    if (c != fbest.chain_w && (c < wmax || (c == (1 << 30) && wmax >= fbest.chain_w))) cws.push_back(c);
  std::vector<int> sns = {fbest.small_nnz};
  for (int v : {0, 1 << 20, 1 << 22})
    if (v != fbest.small_nnz && (v == 0 || nnzL <= v)) sns.push_back(v);
  for (int c : cws)
    for (int v : sns) {
      if (c == fbest.chain_w && v == fbest.small_nnz) continue;  // timed above
      Handle::Tuning t = fbest;
      t.chain_w = c;
      t.small_nnz = v;
      double f, sl;
      if (attempt(t, f, sl) && sl < bs) {
        bs = sl;
        best.chain_w = c;
        best.small_nnz = v;
      }
    }
  h.tune = best;
  double f, sl;
  time_schedule(h, nrhs, reps, B, X, f, sl);  // leave the handle factored with the chosen schedule
  factor_ms = f;
  solve_ms = sl;
}

}  // namespace lblg

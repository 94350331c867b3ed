// Device-resident LBL factor and the host-side schedules that drive it.
//
// HBM layout (one handle):
//   valL     supernode panels, column-major r_s x w_s, at px[s] (the
//            reference's BCSC value layout, include/nasoq/common/def.h:131-220)
//   dval     2n: diagonal of D in [0,n), 2x2 sub-diagonals in [n,2n)
//   rows     row labels of every supernode (BCSC s[]), relabelled after
//            pivoting so no panel row is ever moved
//   perm_piv local pivot permutation (pivoted position -> pre-pivot position)
//   amap     for every stored entry of K: its slot in valL
//   K (CSR)  full symmetric K for the SpMV of the refinement
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <vector>

namespace lblg {

// list-driven launches: a launch processes the supernodes in a device list
struct SnList {
  int* d = nullptr;  // device array of supernode ids
  int count = 0;
};

// one GPU level of the factorization
struct FactorLevel {
  SnList small;                    // w <= SMALL_W: LDS Bunch-Kaufman
  std::vector<int> medium;         // SMALL_W < w <= GROUP_W: one workgroup each
  std::vector<int> large;          // w > GROUP_W: a cooperative group each
  SnList offdiag;                  // supernodes with rows below the diagonal block
  int offdiag_blocks = 0;          // total 64-row blocks of those
  int offdiag_wmax = 0;            // widest of those
  int small_wmax = 0;              // widest supernode of the small list
  int* d_offdiag_blk = nullptr;    // prefix of 64-row blocks per offdiag supernode
  int* d_offdiag_pos = nullptr;    // 64-row block -> position of its supernode in the list
  SnList schur;                    // supernodes with an update (m > 0)
  int64_t schur_tiles = 0;
  int64_t* d_schur_tile = nullptr; // prefix of 64x64 lower tiles per schur supernode
  int* d_schur_pos = nullptr;      // tile -> position of its supernode in the list
  SnList schur_big;                // Schur updates with >= 256 off-diagonal rows: 128x128 tiles
  int64_t schur_big_tiles = 0;
  int64_t* d_schur_big_tile = nullptr;
  int* d_schur_big_cols = nullptr;    // prefix of columns (for the L D scaling launch)
  int schur_big_ncols = 0;
  int64_t* d_schur_big_ldoff = nullptr;  // offset of each supernode's L D block in DeviceFactor::LD
  int max_w = 0;
  std::vector<int> medium_dummy;
  int* d_medium = nullptr;
  int* d_large = nullptr;
  std::vector<int> large_G, large_nb;  // workgroups and panel width per large supernode
  std::vector<int> large_nt, large_wsub;  // workgroup size and panels per trailing update
  // host copies of the lists (the path-restricted refactorization filters them)
  std::vector<int> h_small, h_off, h_sch, h_schb;
  std::vector<int64_t> h_schb_ldoff;  // offset of each schur_big supernode's L D block
};

// solve pieces: supernodes split into column blocks of at most SOLVE_W
struct SolveLevel {
  std::vector<int> chains;  // wide supernodes swept by one cooperative launch each
  int npieces = 0;
  int maxpw = 1;         // widest piece of the level
  int nnarrow = 0;       // pieces of at most TRI_NARROW columns: the first nnarrow of d_piece
  int maxpw_narrow = 1;  // widest of those
  int nblocks_narrow = 0;  // their row blocks (the first of the level's)
  int nblocks = 0;       // total row blocks (SOLVE_ROWS rows) of all pieces
  int* d_piece = nullptr;   // piece ids, narrowest first
  int* d_blk = nullptr;     // prefix of row blocks per piece
  int* d_blkpos = nullptr;  // row block -> position of its piece in d_piece
};

struct DeviceFactor {
  int n = 0, nsuper = 0, nrhs_cap = 0;
  hipStream_t stream = nullptr;
  // symbolic (device)
  int* super = nullptr;
  int64_t* pi = nullptr;
  int64_t* px = nullptr;   // value offset of supernode s
  int* rows = nullptr;     // row labels (ssize), relabelled after factor
  int* rows0 = nullptr;    // pre-pivot row labels (kept for re-factorization)
  int* col2sup = nullptr;
  std::vector<int> hsuper;      // host copies of super / pi (launch shapes)
  std::vector<int64_t> hpi;
  // numeric: the factor's value type is fixed at setup (prec 0: double,
  // 1: float, 2: float with BF16 operands in the 128-tile updates; both
  // float forms factor the diagonally scaled S K S); valL, dval, W, LD and
  // the sweep vectors X, Y, Z, acc, O hold that type
  int prec = 0;
  void* valL = nullptr;
  void* dval = nullptr;
  int* perm_piv = nullptr;   // local: new position -> old position (global indices)
  int* newpos = nullptr;     // inverse of perm_piv
  int* perm_f = nullptr;     // final ordering: ordered index -> original index
  int* iperm_f = nullptr;
  int* perm0 = nullptr;      // symbolic ordering before pivoting
  int64_t* amap = nullptr;   // entry -> valL slot
  double* Kx = nullptr;      // values of K (lower CSC, as given)
  int* Kdiag = nullptr;      // position of each diagonal entry in Kx
  unsigned char* pert = nullptr;  // perturbed-diagonal flags
  int nnzK = 0;
  // full CSR of K0 (unperturbed; perturbed diagonals zeroed) in the original ordering
  int* Krp = nullptr;
  int* Kci = nullptr;
  double* Kv = nullptr;
  int64_t* Ksrc = nullptr;   // CSR entry -> lower CSC entry (to refresh values)
  // workspace
  void* W = nullptr;            // BK panel workspace (max over concurrent supernodes)
  void* LD = nullptr;           // L_off D of the large Schur updates of one level
  int64_t W_elems = 0;
  int64_t* woff = nullptr;      // per-supernode offset into W
  int* kpos = nullptr;          // next column to factor (wide supernodes)
  int* kstart = nullptr;        // first column of the current block of panels
  unsigned char* isdiag = nullptr;
  int64_t nnzf = 0;             // entries of the full CSR
  double* k0norm = nullptr;     // ||K0||_inf on the device
  int nseg = 0;                 // systems of a block-diagonal batch (0: one system)
  int64_t* seg = nullptr;       // their first rows (nseg + 1), in both orderings
  double* segnorm = nullptr;    // ||K0||_inf of every system
  double* segsums = nullptr;    // |r|, |x|, |b| sums per system and column
  int* sync = nullptr;          // group-barrier counters
  int wide_par = 0;             // which of sync[0], sync[1] the next wide panel counts on
  uint64_t* ll = nullptr;       // flag-in-word exchange buffer of the wide panel kernel
  uint32_t ll_epoch = 0;        // next epoch base (monotone for the handle's lifetime)
  // path-restricted refactorization (the SoMod update): the supernodes on
  // the tree paths from changed entries to the root are refactored; the
  // others keep their panels, and those with rows in a refactored supernode
  // send their Schur updates again, to the refactored supernodes only
  bool path_refactor = true;    // LBLG_PATH_REFACTOR=0: every factorization is a full one
  bool factor_valid = false;    // the device factor is that of the uploaded values
  double factor_key[3] = {0, 0, 0};  // reg_diag, regularization, precision of that factor
  int* asup = nullptr;          // entry of K -> supernode of its slot in valL
  std::vector<int> hasup;
  std::vector<int> hlevel;      // factor level of every supernode
  // per supernode: panel family (0 LDS, 1 one-workgroup panel, 2 cooperative:
  // its position in the level's large list is s_large_q), whether it has
  // off-diagonal rows for k_offdiag, its Schur form (0 none, 1 64-row tiles,
  // 2 128-row tiles) and the offset of its L D block in LD
  std::vector<unsigned char> s_kind, s_off, s_sch;
  std::vector<int> s_large_q;
  std::vector<int64_t> s_ldoff;
  std::vector<unsigned char> s_mark, s_send;  // per refactorization: modified / sending supernodes
  std::vector<int> upd_ptr, upd_idx;  // per supernode: the supernodes with rows in its columns
  std::vector<unsigned> smark;  // marks of the current refactorization (epoch stamps)
  unsigned smark_epoch = 0;
  unsigned char* tmask = nullptr;  // device: flag of every supernode refactored
  char* blob_h = nullptr;       // pinned staging of the restricted lists and changed values
  char* blob_d = nullptr;
  size_t blob_cap = 0;
  double* red = nullptr;        // reduction scratch
  int* status = nullptr;        // device flags (zero pivots etc.)
  // FP32 / BF16 factors (prec 1, 2): the symmetric diagonal scaling S and S K S
  int* Krow = nullptr;          // row and column of every entry of K (lower CSC)
  int* Kcol = nullptr;
  double* Kf = nullptr;         // S K S, lower CSC values
  double* scale = nullptr;      // s_i, original ordering
  double* scale_ord = nullptr;  // s in the factor's pivoted ordering
  double* sbuf = nullptr;       // scaled block for the one-workgroup sweeps (n x nrhs_cap)
  // solve
  void* X = nullptr;            // n x nrhs sweep vectors in the factor's type (row-interleaved)
  void* Y = nullptr;
  void* Z = nullptr;
  void* acc = nullptr;
  void* O = nullptr;            // sweep result (float factor; double writes the caller's block)
  double* acc2 = nullptr;       // ordered solution block
  double* T = nullptr;
  double* Xs = nullptr;
  double* V = nullptr;          // GMRES basis (mr+1) x n x nrhs
  double* R = nullptr;
  double* B = nullptr;
  double* gm = nullptr;         // GMRES small state
  double* part = nullptr;       // reduction partials
  double* dots = nullptr;       // reduction results
  int nblk_red = 512;
  int mr_cap = 0;
  // schedules
  std::vector<FactorLevel> flevels;
  std::vector<SolveLevel> slevels;
  // supernodes by level for the one-workgroup sweeps of small factors
  int4* ss_info = nullptr;      // (first column, width, rows, first row index) in level order
  int64_t* ss_ipx = nullptr;   // first value of each, in level order
  int* ss_levptr = nullptr;
  int* ss_mode = nullptr;  // per level: 1 = whole-workgroup sweeps
  int ss_nlev = 0;
  int ss_global = 0;            // 1: the block in memory even where it fits LDS (tests); 2: beyond the LDS only
  // one-workgroup sweeps of each system of a block-diagonal batch (one workgroup per system)
  int ssb_nsys = 0;             // systems with a schedule (0: the level sweeps)
  int ssb_maxlev = 0;           // most levels of a system
  int ssb_lds_kb = 104;         // LDS for a system's block (larger blocks stay in memory)
  int* ssb_sysd = nullptr;      // 6 ints per system: level-pointer and mode offsets, levels, first column, columns
  int4* ssb_info = nullptr;     // supernodes system by system, each system level by level
  int64_t* ssb_ipx = nullptr;
  int* ssb_levptr = nullptr;    // per system: levels + 1 pointers into ssb_info
  int* ssb_mode = nullptr;
  std::vector<int> ssb_cols;    // columns of each system, ascending (the LDS size of a launch)
  int64_t ss_nnzL = 0, ss_max_nnz = 0;
  double *stage_b = nullptr, *stage_x = nullptr;  // host-pointer solves: staged blocks (n x nrhs_cap)
  // residual norms of caller-given solutions (residual.hip), allocated at first use
  double* res_part = nullptr;   // per-workgroup partial sums
  double* res_out = nullptr;    // final sums and errors
  double* res_bx = nullptr;     // staged B and X of the host-pointer call (n x MAX_RHS each)
  uint64_t* chain_xb = nullptr;     // flag-in-word exchange of the chain sweeps (epoch-stamped words)
  uint32_t chain_epoch = 0;
  int* piece_sn = nullptr;      // piece -> supernode
  int* piece_c0 = nullptr;      // piece -> first local column within the supernode
  int npieces = 0;
  std::vector<void*> owned;     // device allocations freed at destruction
  // captured launch sequences (hipGraph) of the factor and of the
  // fixed-refinement solve; a sequence is captured on its second call with
  // the same key and replayed from then on
  struct GraphCache {
    std::string key;
    hipGraphExec_t exec = nullptr;
    int seen = 0;
  };
  GraphCache gfactor, gsolve, gcycle;
  bool graphs = true;       // LBLG_GRAPH=0 turns capture off
  bool panel_lds = true;         // one-workgroup panels of diagonal blocks that fit: staged in LDS
  bool schur_direct = true;      // 128-tile Schur updates with operands loaded lane by lane (no LDS)
  bool trailing_direct = true;   // the same for the wide panels' trailing updates
  bool strip64 = true;           // the wide panels' strip pre-updates in 64 x 64 tiles
  bool has_coop = false;    // wide panels (cooperative, epoch-stamped): factor not captured
  bool has_chains = false;  // chain sweeps (epoch-stamped): level sweeps not captured
  int64_t graph_launches = 0, graph_captures = 0;
  // optional per-kernel timing with events on the engine stream
  bool timing = false;
  std::set<std::string> timing_only;  // when not empty, only these kernel families are timed
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::map<std::string, std::pair<double, int64_t>> ktime;  // total ms, launches
  void resolve_timing();
  ~DeviceFactor();
};

}  // namespace lblg

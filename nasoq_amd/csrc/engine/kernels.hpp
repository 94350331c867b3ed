// Launch wrappers of the gfx950 kernels (kernels.hip).  All launches are
// asynchronous on the given stream.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lblg {

constexpr int SMALL_W = 64;    // widest supernode factored entirely in LDS
constexpr int PANEL_NB = 32;   // Bunch-Kaufman panel width for wider supernodes
constexpr int SOLVE_W = 128;   // column width of a solve piece
constexpr int TRI_NARROW = 32; // triangle launches of a level: pieces up to this width, then the wider ones
constexpr int MAX_RHS = 8;     // right-hand sides per solve call

struct SnArgs {  // symbolic arrays on the device
  const int* super;
  const int64_t* pi;
  const int64_t* px;
  const int* rows;
  const int* col2sup;
  int n;
  // path-restricted refactorization: Schur updates reach only supernodes
  // whose flag is set (nullptr: all)
  const unsigned char* tmask = nullptr;
};

template <typename T>
void launch_fill(T* p, int64_t count, double v, hipStream_t st);
void launch_fill_int(int* p, int64_t count, int v, hipStream_t st);
void launch_iota(int* p, int count, hipStream_t st);
void launch_gmres_reset(double* gm, int64_t ss, int nrhs, int64_t flag_off, hipStream_t st);
void launch_copy(double* dst, const double* src, int64_t count, hipStream_t st);
template <typename TO, typename TI>
void launch_convert(TO* dst, const TI* src, int64_t count, hipStream_t st);
// The value kernels are templates over the factor's type T (double: plain
// FP64 factor, float: FP32 factor), instantiated for both in kernels.hip,
// bk_wide.hip and gemm_big.hip.
// asup / tmask (both or neither): only entries whose slot lies in a supernode
// with its tmask flag set are written
template <typename T>
void launch_assemble(const int64_t* amap, const double* Kx, const unsigned char* pert, int nnz, double reg,
                     T* valL, const int* asup, const unsigned char* tmask, hipStream_t st);
// the panels, D entries and local pivots of the listed supernodes back to
// zero / identity before they are refactored
template <typename T>
void launch_reset_supernodes(const int* list, int count, const SnArgs& a, T* valL, T* dval, int* perm_piv,
                             hipStream_t st);
// Kx[idx[k]] = vals[k]
void launch_scatter_values(int64_t count, const int64_t* idx, const double* vals, double* Kx, hipStream_t st);
template <typename T>
void launch_bk_small(const int* list, int count, int wmax, const SnArgs& a, T* valL, T* dval,
                     int* perm_piv, int* status, hipStream_t st);
// lds_rows > 0: one workgroup per supernode with the panel staged in LDS
// (every supernode of the list has at most lds_rows <= panel_lds_rows<T>()
// columns); 0: the panel stays in memory
constexpr size_t PANEL_LDS_BYTES = 150 * 1024;
template <typename T>
int panel_lds_rows();
template <typename T>
void launch_bk_panel(const int* list, int count, int G, const SnArgs& a, T* valL, T* dval,
                     int* perm_piv, T* W, const int64_t* woff, int* kpos, int* kstart, int* sync,
                     int* status, bool cooperative, int lds_rows, hipStream_t st);
// wide panels with flag-in-word exchange (bk_wide.hip)
constexpr int WIDE_NB = 64;         // panel width
constexpr int WIDE_ROWS = 192;      // far rows per workgroup (256-thread workgroups)
constexpr int WIDE_NT_TALL = 512;   // workgroup size for supernodes beyond 256 x WIDE_ROWS far rows
constexpr int WIDE_ROWS_TALL = WIDE_NT_TALL - 64;  // their far rows per workgroup
constexpr int WIDE_MOVED = 128;     // moved indices per panel (scratch rows)
// widest panel of a launch shape: 64 columns, or for the tall shape the
// columns of a 256-byte LDS row (32 FP64, 64 FP32)
inline int wide_panel_nb(int nt, size_t value_size) { return nt == 256 ? 64 : (value_size == 8 ? 32 : 64); }
constexpr int64_t LL_WORDS = 2 * 256 * 4 + 2 * 256 * 8 + 2 * 256 * 2 * 64;
template <typename T>
void launch_bk_wide(const int* list, int G, int nt, int nb, int sub, int wcols, const SnArgs& a, T* valL, T* dval,
                    int* perm_piv, T* W, const int64_t* woff, int* kpos, int* kstart, int* sync,
                    int* sync_next, uint64_t* ll, uint32_t ebase, int* status, hipStream_t st);
template <typename T>
void launch_trailing(const int* list, int count, int tiles_per_sn, int lo, const SnArgs& a, T* valL,
                     const T* W, const int64_t* woff, const int* kpos, const int* kstart,
                     hipStream_t st);
// 128 x 128 MFMA tiles (gemm_big.hip); operand modes of the GEMM
constexpr int GEMM_STAGED = 0;  // chunks of the value type staged in LDS
constexpr int GEMM_DIRECT = 1;  // every lane loads its own operands from memory
constexpr int GEMM_BF16 = 2;    // float operands rounded to BF16 in LDS, FP32 accumulation
constexpr int GEMM_BF16X3 = 3;  // float operands split into BF16 hi + lo, three MFMA terms
template <typename T>
void launch_scale_ld(const int* list, int count, const int* col_prefix, int ncols, const SnArgs& a,
                     const T* valL, const T* dval, T* LD, const int64_t* ldoff, hipStream_t st);
template <typename T>
void launch_schur_big(const int* list, int count, const int64_t* tile_prefix, int64_t ntiles, const SnArgs& a,
                      T* valL, const T* LD, const int64_t* ldoff, int mode, hipStream_t st);
// the strip pre-update of a wide panel (launch_trailing_big's strip form) in
// 64 x 64 tiles: row_tiles >= ceil((r - first panel column) / 64)
template <typename T>
void launch_strip_update(const int* list, int row_tiles, int strip, int wcols, const SnArgs& a, const T* valL, T* W,
                         const int64_t* woff, const int* kpos, const int* kstart, hipStream_t st);
template <typename T>
void launch_trailing_big(const int* list, int ntiles, int lo, int strip, int wcols, const SnArgs& a, T* valL,
                         T* W, const int64_t* woff, const int* kpos, const int* kstart, int mode, hipStream_t st);
template <typename T>
void launch_offdiag(const int* list, int count, int wmax, const int* blk_prefix, const int* blkpos, int nblocks,
                    const SnArgs& a, T* valL, const T* dval, const int* perm_piv, hipStream_t st);
template <typename T>
void launch_schur(const int* list, int count, const int64_t* tile_prefix, const int* tilepos, int64_t ntiles,
                  const SnArgs& a, T* valL, const T* dval, hipStream_t st);
void launch_relabel(const SnArgs& a, int nsuper, const int* rows0, int* rows, const int* perm_piv,
                    int* newpos, const int* perm0, int* perm_f, int* iperm_f, hipStream_t st);
constexpr int SOLVE_ROWS = 64;  // rows per workgroup in the sweeps below a piece
template <typename T>
void launch_fwd_tri(const int* pieces, int np, int maxpw, const int* piece_sn, const int* piece_c0, const SnArgs& a,
                    const T* valL, const T* X, T* Y, int nrhs, hipStream_t st);
template <typename T>
void launch_fwd_gemv(const int* pieces, const int* blk_prefix, const int* blkpos, int nblk, const int* piece_sn,
                     const int* piece_c0, const SnArgs& a, const T* valL, const T* Y, T* X,
                     int nrhs, hipStream_t st);
template <typename T>
void launch_dsolve(int n, const T* dval, const T* Y, T* Z, int nrhs, hipStream_t st);
template <typename T>
void launch_bwd_gemv(const int* pieces, int maxpw, const int* blk_prefix, const int* blkpos, int nblk, int blk_base,
                     const int* piece_sn, const int* piece_c0, const SnArgs& a, const T* valL, const T* Xf, T* acc,
                     int nrhs, hipStream_t st);
template <typename T>
void launch_bwd_tri(const int* pieces, int np, int maxpw, const int* piece_sn, const int* piece_c0, const SnArgs& a,
                    const T* valL, const T* Z, const T* acc, T* Xf, int nrhs,
                    hipStream_t st);
// one wide supernode's sweep in a single cooperative launch (solve_chain.hip)
int chain_groups(int w, int64_t r, bool forward);  // workgroups, or -1 if too large
template <typename T>
void launch_fwd_chain(int s, int w, int64_t r, const SnArgs& a, const T* valL, T* X, T* Y, int nrhs,
                      uint64_t* xb, uint32_t epoch, int* status, hipStream_t st);
template <typename T>
void launch_bwd_chain(int s, int w, int64_t r, const SnArgs& a, const T* valL, const T* Z, T* offacc, T* Xf,
                      int nrhs, uint64_t* xb, uint32_t epoch, int* status, hipStream_t st);
void launch_spmv(int n, const int* rp, const int* ci, const double* v, const int* perm_f, const int* iperm_f,
                 const double* x, double* y, int nrhs, hipStream_t st);
void launch_gather_rhs(int n, int nrhs, const double* B, int64_t ldb, const int* perm_f, double* out,
                       hipStream_t st);
void launch_scatter_sol(int n, int nrhs, const double* Xo, const int* perm_f, double* X, int64_t ldx,
                        hipStream_t st);
// L, D and L^T sweeps of a small factor in one workgroup with the ordered
// block (n x nrhs, rows interleaved) resident in LDS; supernodes level by
// level in order[levptr[l] .. levptr[l + 1]), mode[l] = 1 where the whole
// workgroup takes the level's supernodes one at a time.  The block is held in
// LDS when n * nrhs * sizeof(T) <= SOLVE_SMALL_LDS, otherwise in xg (n x nrhs
// values of T in memory; nullptr: LDS).
constexpr size_t SOLVE_SMALL_LDS = 104 * 1024;
constexpr int SOLVE_SMALL_LEVELS = 2048;  // level pointers and modes kept in LDS
constexpr int SS_THIN_ROWS = 32;          // one-column supernodes up to this many rows: one lane each,
constexpr int SS_THIN_MIN = 64;           // in levels holding at least this many of them
// info[k] = (first column, width, rows, first row index), ipx[k] = first value,
// for the k-th supernode of the level order; mode[l]: bit 0 = whole-workgroup
// level, bits 1.. = number of thin supernodes, which lead the level's list
template <typename T>
// active (stride astride per column, may be null): a GMRES step's per-column
// participation flags; when none is set the launch returns at once
void launch_solve_small(const int4* info, const int64_t* ipx, const int* levptr, const int* mode, int nlev,
                        const SnArgs& a, const T* valL, const T* dval, const double* in, double* out, int nrhs,
                        const double* active, int64_t astride, T* xg, hipStream_t st);
// the same sweeps for every system of a block-diagonal batch, one workgroup
// per system (sysd: 6 ints per system, see solve_small.hip; its level
// pointers index info / ipx); systems of at most fit_cols rows keep their
// block in LDS, the others in xg (n x nrhs values of T in memory)
template <typename T>
void launch_solve_batch(const int* sysd, int nsys, const int4* info, const int64_t* ipx, const int* levptr,
                        const int* mode, int fit_cols, int maxlev, const SnArgs& a, const T* valL, const T* dval,
                        const double* in, double* out, int nrhs, const double* active, int64_t astride, T* xg,
                        hipStream_t st);
// per-column dot products out[c] = sum_i a(i, c) b(i, c) (rows interleaved by
// column), reduced in a fixed order through at most nblk partials
void launch_dot(int n, int nrhs, const double* a, const double* b, double* part, int nblk, double* out,
                hipStream_t st);
void launch_sub(int64_t count, const double* b, const double* y, double* r, hipStream_t st);
void launch_axpby(int n, int nrhs, const double* coefs, int cstride, const double* x, double* y,
                  const double* mask, int mstride, int mask_off, int mode, hipStream_t st);
// out = column dots of a and b, then GMRES scalar step op on them (one launch for short vectors)
void launch_dot_scalar(int n, int nrhs, const double* a, const double* b, double* part, int nblk, double* out,
                       int op, int MR, double* gm, int j, int k, int itr, double tol_abs, double tol_rel,
                       hipStream_t st);
void launch_gmres_scalar(int op, int nrhs, int MR, double* gm, const double* dots, int j, int k, int itr,
                         double tol_abs, double tol_rel, hipStream_t st);
void launch_gmres_update(int n, int nrhs, int MR, const double* V, const double* gm, double* x,
                         hipStream_t st);
void launch_abs_sums(int n, int nrhs, const double* r, const double* x, const double* b, double* part,
                     int nblk, double* out, hipStream_t st);
// per-system reductions of a block-diagonal batch (segments of rows)
void launch_seg_abs_sums(int nseg, const int64_t* seg, int nrhs, const double* r, const double* x, const double* b,
                         double* out, hipStream_t st);
void launch_seg_rowabs_max(int nseg, const int64_t* seg, const int* rp, const double* Kv, double* out,
                           hipStream_t st);
void launch_pert_flags(int nnz, const unsigned char* isdiag, const double* Kx, const double* Kf, double reg,
                       double reg_factor, int regularization, unsigned char* pert, hipStream_t st);
// symmetric diagonal scaling of the BF16 factor (kernels.hip)
void launch_diag_scaling(int n, int nnz, const int* rp, const int64_t* src, const int* row, const int* col,
                         const double* Kx, double* s, double* Kf, hipStream_t st);
// Kf = S K S with the scaling s as it is (the path-restricted refactorization
// of the FP32 / BF16 factors keeps the scaling of the last full one)
void launch_scale_entries(int nnz, const int* row, const int* col, const double* Kx, const double* s, double* Kf,
                          hipStream_t st);
void launch_gather_scale(int n, const int* perm_f, const double* s, double* s_ord, hipStream_t st);
template <typename TO, typename TI>
void launch_row_scale_copy(TO* dst, const TI* src, const double* s, int n, int nrhs, hipStream_t st);
void launch_k0(int n, int64_t nnzf, const int* rp, const int64_t* src, const double* Kx,
               const unsigned char* pert, double* Kv, double* part, int nblk, double* norm_out, hipStream_t st);
void launch_mfma_selftest(const double* A, const double* B, double* C, int f32, hipStream_t st);

// GMRES state per right-hand-side column (doubles)
inline int64_t gmres_state_size(int MR) { return (int64_t)(MR + 1) * MR + 4 * (MR + 1) + (MR + 2) + 8; }
inline int64_t gmres_off_D(int MR) { return (int64_t)(MR + 1) * MR + 4 * (MR + 1); }
inline int64_t gmres_off_S(int MR) { return gmres_off_D(MR) + (MR + 2); }

}  // namespace lblg

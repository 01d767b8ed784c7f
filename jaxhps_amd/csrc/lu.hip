// Batched dense LU with partial pivoting and the matching solve, gfx950.
//
// Factor format: P A = L U stored LAPACK-style in A, plus
//   perm[b][i]        row of A that became row i of P A,
//   dinv[b][blk]      for every NB x NB diagonal block, the packed inverses
//                     (strict lower part: inv(L11); upper part: inv(U11)),
// so every triangular solve is a parallel block product instead of a
// sequential substitution.
//
// getrf: blocked right-looking, panel width NB = 32.
//   1. panel (one workgroup per matrix), column by column with a wave-level
//      argmax (shuffle reduction inside each 64-wide wave, then across waves
//      through LDS):
//        - short panels are staged whole in LDS;
//        - tall panels (merge matrices) keep 4-column sub-panels in
//          registers (rows distributed over 1024 threads) and update each
//          sub-panel left-looking from the columns already factored.
//      The kernel ends by inverting the diagonal block into dinv and writing
//      the panel's net row permutation as a (row, source) table.
//   2. swap + U12: applies that table to every column outside the panel in
//      parallel and forms U12 = inv(L11) A12 as a block product.
//   3. trailing update A22 -= L21 U12 with the batched MFMA GEMM.
// getrs: gather rows by perm; per block X_k = inv(L_kk) X_k, then the MFMA
//   GEMM update below; backward X_k = inv(U_kk) X_k, GEMM update above.
#include <string>
#include <utility>

#include "hps_kernels.h"
#include "hps_prof.h"

namespace hps {

#define HPS_LU_CHECK(expr)              \
  do {                                  \
    hipError_t _e = (expr);             \
    if (_e != hipSuccess) return _e;    \
  } while (0)

#ifndef HPS_LU_TYPES
#define HPS_LU_TYPES 15
#endif

constexpr int LU_NB = 32;
// with row exchanges off, batches this small take the unpivoted panel kernel
// for every panel: the pivoted register panels then run one workgroup per
// matrix on an otherwise idle GPU
constexpr int LU_NOPIV_SMALL_BATCH = 16;
// largest multiplier an unpivoted register panel accepts (the leaf and merge
// matrices of the HPS tree stay below 2)
constexpr double LU_NOPIV_BOUND = 64.0;
constexpr size_t LU_PANEL_LDS_MAX = 150 * 1024;
constexpr int SWTAB = 1 + 4 * LU_NB;  // count, then up to 2*NB (row, src) int pairs
template <typename T> __device__ __forceinline__ T make_conj(T x) { return x; }
template <> __device__ __forceinline__ c128 make_conj<c128>(c128 x) { return c128{x.re, -x.im}; }
template <> __device__ __forceinline__ c64 make_conj<c64>(c64 x) { return c64{x.re, -x.im}; }

constexpr int LU_NBO = 256;           // column block of the blocked factorisation
constexpr int LU_BLOCKED_MIN = 768;   // matrices larger than this are factored in column blocks
constexpr int TS_CHUNK = 256;  // rows per first-round tournament chunk
static size_t tslu_cand_ints(int n, int batch) { return (size_t)batch * ((n + TS_CHUNK - 1) / TS_CHUNK) * LU_NB; }
// tournament pivoting for panels taller than the register panel holds when
// batch x rows is small enough that one workgroup per matrix would leave most
// of the GPU idle
constexpr long LU_TSLU_MAX_WORK = 64L * 1024;

// Column block of the two-level factorisation and solves: 32-wide steps
// update only their block; the rest is updated once per block by a GEMM with
// K = the block width.
static int lu_outer_block(int n) { return n > LU_BLOCKED_MIN ? LU_NBO : (n >= 256 ? 128 : n); }
// the factorisation and the many-column solves of large matrices take column
// blocks of 512: their trailing updates are single large GEMMs whose K = 512
// halves the C-tile traffic per flop of K = 256 (the few-column solves keep
// 256-blocks, which their block kernel holds)
static_assert(LU_WIDE_MIN == 3072 && BLOCKINV_NB == 2 * LU_NBO, "block sizes");
static int lu_outer_block_wide(int n) { return n >= LU_WIDE_MIN ? 2 * LU_NBO : lu_outer_block(n); }

static __global__ void fill_int_kernel(int* p, long n, int v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

static __global__ void iota_kernel(int* perm, int n, long total) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) perm[i] = (int)(i % n);
}

// selects instead of a branch: inside the wave reductions a branch costs a
// dozen exec-mask instructions per merge
template <typename R>
__device__ __forceinline__ void argmax_merge(R& bv, int& bi, R v, int i) {
  const bool take = v > bv || (v == bv && i < bi);
  bv = take ? v : bv;
  bi = take ? i : bi;
}

// ---------------------------------------------------------------------------
// Wave-wide argmax of (|value|, position) pairs with DPP lane moves (no LDS
// round trips): quad exchanges, row shifts by 4 and 8, then the row
// broadcasts; the wave's winner ends in lane 63 and is read back as a
// uniform value.  Lanes whose DPP source is outside the pattern keep the
// identity (-1, NONE), which never wins a merge.
// ---------------------------------------------------------------------------
template <int CTRL, int RM>
__device__ __forceinline__ int dpp_int(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, RM, 0xf, false);
}
template <int CTRL, int RM>
__device__ __forceinline__ float dpp_move(float old, float v) {
  return __int_as_float(dpp_int<CTRL, RM>(__float_as_int(old), __float_as_int(v)));
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_move(double old, double v) {
  const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
  const int lo = dpp_int<CTRL, RM>((int)o, (int)x);
  const int hi = dpp_int<CTRL, RM>((int)(o >> 32), (int)(x >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL, int RM, typename R>
__device__ __forceinline__ void dpp_argmax_step(R& bv, int& bp) {
  const R ov = dpp_move<CTRL, RM>(R(-1), bv);
  const int op = dpp_int<CTRL, RM>(0x7fffffff, bp);
  argmax_merge(bv, bp, ov, op);
}
__device__ __forceinline__ float readlane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }
__device__ __forceinline__ double readlane63(double v) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)x, 63), hi = __builtin_amdgcn_readlane((int)(x >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <typename R>
__device__ __forceinline__ void wave_argmax(R& bv, int& bp) {
  dpp_argmax_step<0xb1, 0xf>(bv, bp);   // quad_perm [1,0,3,2]
  dpp_argmax_step<0x4e, 0xf>(bv, bp);   // quad_perm [2,3,0,1]
  dpp_argmax_step<0x114, 0xf>(bv, bp);  // row_shr:4
  dpp_argmax_step<0x118, 0xf>(bv, bp);  // row_shr:8
  dpp_argmax_step<0x142, 0xa>(bv, bp);  // row_bcast:15 into rows 1, 3
  dpp_argmax_step<0x143, 0xc>(bv, bp);  // row_bcast:31 into rows 2, 3
  bv = readlane63(bv);
  bp = __builtin_amdgcn_readlane(bp, 63);
}

// 1 / a for a value every lane holds alike: the branch of Smith's algorithm is
// taken on the first lane's comparison, so only one side executes.
__device__ __forceinline__ double s_inv_uniform(double a) { return 1.0 / a; }
__device__ __forceinline__ float s_inv_uniform(float a) { return 1.0f / a; }
template <typename C>
__device__ __forceinline__ C s_inv_uniform_c(C a) {
  using R = decltype(a.re);
  if (__builtin_amdgcn_readfirstlane((int)(fabs(a.re) >= fabs(a.im)))) {
    const R r = a.im / a.re, d = a.re + a.im * r;
    const R id = R(1) / d;
    return C{id, -r * id};
  }
  const R r = a.re / a.im, d = a.re * r + a.im;
  const R id = R(1) / d;
  return C{r * id, -id};
}
__device__ __forceinline__ c128 s_inv_uniform(c128 a) { return s_inv_uniform_c(a); }
__device__ __forceinline__ c64 s_inv_uniform(c64 a) { return s_inv_uniform_c(a); }

// Block-wide argmax of (value, index) pairs; every thread gets the winner.
template <typename R, int THREADS>
__device__ __forceinline__ int block_argmax(R bv, int bi, R* red_v, int* red_i, int* s_out, R* s_val) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    R ov = __shfl_xor(bv, off, 64);
    int oi = __shfl_xor(bi, off, 64);
    argmax_merge(bv, bi, ov, oi);
  }
  if (lane == 0) {
    red_v[wave] = bv;
    red_i[wave] = bi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    R fv = red_v[0];
    int fi = red_i[0];
    for (int w = 1; w < THREADS / 64; ++w) argmax_merge(fv, fi, red_v[w], red_i[w]);
    *s_out = fi;
    *s_val = fv;
  }
  __syncthreads();
  return *s_out;
}

// Record pivot j of the panel: swap sequence and breakdown flag.  The row
// permutation is assembled once from the swap sequence after the last panel.
__device__ __forceinline__ void record_pivot(int k, int j, int fi, bool zero, int* ipivb, int* info, int b) {
  ipivb[k + j] = k + fi;
  if (zero && info[b] == 0) info[b] = k + j + 1;
}

// perm = the composition of all row swaps: one workgroup per matrix, the
// permutation built in LDS.
static __global__ __launch_bounds__(256) void perm_from_ipiv_kernel(int n, const int* ipiv, int* perm) {
  extern __shared__ int sp[];
  const int b = blockIdx.x;
  const int* ib = ipiv + (long)b * n;
  for (int i = threadIdx.x; i < n; i += 256) sp[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < n; ++i) {
      int p = ib[i];
      if (p != i) {
        int t = sp[i];
        sp[i] = sp[p];
        sp[p] = t;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 256) perm[(long)b * n + i] = sp[i];
}

// After the panel: (row, source) table of its net row permutation, from the
// swap sequence ipiv[k .. k+nb), into tab (LDS: count, then pairs).  Single
// thread.
__device__ void build_swap_table(int k, int nb, const int* ipivb, int* tab) {
  // working lists in LDS: a per-thread array indexed at run time would live in scratch
  __shared__ int rows[2 * LU_NB], src[2 * LU_NB];
  int cnt = 0;
  for (int j = 0; j < nb; ++j) {
    int a = k + j, bb = ipivb[k + j];
    if (a == bb) continue;
    int ia = -1, ib = -1;
    for (int t = 0; t < cnt; ++t) {
      if (rows[t] == a) ia = t;
      if (rows[t] == bb) ib = t;
    }
    if (ia < 0) {
      rows[cnt] = a;
      src[cnt] = a;
      ia = cnt++;
    }
    if (ib < 0) {
      rows[cnt] = bb;
      src[cnt] = bb;
      ib = cnt++;
    }
    int t = src[ia];
    src[ia] = src[ib];
    src[ib] = t;
  }
  tab[0] = cnt;
  for (int t = 0; t < cnt; ++t) {
    tab[1 + 2 * t] = rows[t];
    tab[2 + 2 * t] = src[t];
  }
}

// Invert the factored nb x nb diagonal block D (row stride ldd) into the
// packed form out (row stride ldo): strict lower = inv(L11), upper = inv(U11).
// Four consecutive lanes (a DPP quad) form one column; lane s of the quad
// keeps the column's entries j = 4q + s in registers.  Each entry is one unrolled
// substitution step whose dot product is split over the quad and summed with
// two quad exchanges.  Entries not yet formed are zero, so the sums need no
// bounds: for column c the rows above c stay zero in inv(L11) and the rows
// below c stay zero in inv(U11).  Rows and columns >= nb hold garbage the
// caller discards.
template <typename T>
__device__ __forceinline__ T quad_sum(T v);
template <>
__device__ __forceinline__ double quad_sum(double v) {
  v += dpp_move<0xb1, 0xf>(0.0, v);
  return v + dpp_move<0x4e, 0xf>(0.0, v);
}
template <>
__device__ __forceinline__ float quad_sum(float v) {
  v += dpp_move<0xb1, 0xf>(0.0f, v);
  return v + dpp_move<0x4e, 0xf>(0.0f, v);
}
template <>
__device__ __forceinline__ c128 quad_sum(c128 v) { return c128{quad_sum(v.re), quad_sum(v.im)}; }
template <>
__device__ __forceinline__ c64 quad_sum(c64 v) { return c64{quad_sum(v.re), quad_sum(v.im)}; }

template <typename T>
__device__ __forceinline__ void invert_diag_block(const T* D, int ldd, T* out, int ldo, int nb) {
  constexpr int P = 4, Q = LU_NB / P;
  __shared__ T invd[LU_NB];
  const int t = threadIdx.x;
  if (t < LU_NB) invd[t] = t < nb ? s_inv(D[t * ldd + t]) : s_zero<T>();
  __syncthreads();
  // 64 column jobs (32 of inv(L11), then 32 of inv(U11)) over the quads; with
  // 64, 128 or 256 threads every wave takes whole rounds of one kind
  const int nquad = (int)blockDim.x / P < 64 ? (int)blockDim.x / P : 64;
  const int sl = t % P;
  for (int job = t / P; job < 64 && t < nquad * P; job += nquad) {
    const bool upper = job >= LU_NB;
    const int c = job % LU_NB;
    T xs[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) xs[q] = s_zero<T>();
    if (!upper) {
#pragma unroll
      for (int i = 0; i < LU_NB; ++i) {
        T acc = s_zero<T>();
#pragma unroll
        for (int q = 0; q * P < i; ++q) acc = s_fnms(acc, D[i * ldd + q * P + sl], xs[q]);
        acc = quad_sum(acc);
        if (sl == i % P) xs[i / P] = i == c ? s_one<T>() : acc;
      }
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (q * P + sl > c) out[(q * P + sl) * ldo + c] = xs[q];
    } else {
#pragma unroll
      for (int i = LU_NB - 1; i >= 0; --i) {
        T acc = s_zero<T>();
#pragma unroll
        for (int q = (i + 1) / P; q < Q; ++q) acc = s_fnms(acc, D[i * ldd + q * P + sl], xs[q]);
        acc = quad_sum(acc);
        if (i == c) acc = s_add(acc, s_one<T>());
        if (sl == i % P) xs[i / P] = s_mul(acc, invd[i]);
      }
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (q * P + sl <= c) out[(q * P + sl) * ldo + c] = xs[q];
    }
  }
}

// Diagonal-block inverse -> dinv, and the panel's swap table tab (LDS:
// count, then (row, source) pairs) -> swtab.  D is the final factored
// diagonal block in LDS (a 32 x 32 array; zeroed outside nb x nb here); Out
// is LDS scratch of nb rows.  64, 128, 256 or more threads.
template <typename T>
__device__ __forceinline__ void finish_panel_tab(T* D, int ldd, T* Out, int ldo, int k, int nb, int n, int b,
                                                 const int* tab, T* dinv, int* swtab) {
  __syncthreads();
  if (nb < LU_NB) {
    // the inversion reads the whole 32 x 32 block: zero it outside nb x nb
    for (int idx = threadIdx.x; idx < LU_NB * LU_NB; idx += blockDim.x) {
      const int r = idx / LU_NB, c = idx - r * LU_NB;
      if (r >= nb || c >= nb) D[r * ldd + c] = s_zero<T>();
    }
    __syncthreads();
  }
  const int nblk = (n + LU_NB - 1) / LU_NB;
  T* db = dinv + ((long)b * nblk + k / LU_NB) * LU_NB * LU_NB;
  // with D zero outside nb x nb the inverse is zero there too, so without
  // LDS scratch (Out == nullptr) it is written straight to dinv
  invert_diag_block<T>(D, ldd, Out ? Out : db, Out ? ldo : LU_NB, nb);
  const int cnt = tab[0] < 2 * LU_NB ? tab[0] : 2 * LU_NB;
  for (int i = threadIdx.x; i < 1 + 2 * cnt; i += blockDim.x) swtab[(long)b * SWTAB + i] = i == 0 ? cnt : tab[i];
  if (!Out) return;
  __syncthreads();
  for (int idx = threadIdx.x; idx < LU_NB * LU_NB; idx += blockDim.x) {
    int r = idx / LU_NB, c = idx - r * LU_NB;
    db[idx] = (r < nb && c < nb) ? Out[r * ldo + c] : s_zero<T>();
  }
}

// finish_panel_tab with the swap table built from ipiv, for panels that swap
// rows explicitly.
template <typename T>
__device__ __forceinline__ void finish_panel(T* D, int ldd, T* Out, int ldo, int k, int nb, int n, int b,
                                             const int* ipivb, T* dinv, int* swtab) {
  __shared__ int tab[SWTAB];
  __syncthreads();
  if (threadIdx.x == 0) build_swap_table(k, nb, ipivb, tab);
  finish_panel_tab<T>(D, ldd, Out, ldo, k, nb, n, b, tab, dinv, swtab);
}

// Short panels: the whole (n-k) x nb panel in LDS.
template <typename T, int THREADS>
__global__ __launch_bounds__(THREADS) void getrf_panel_lds_kernel(int n, T* A, long lda, long sA, int k, int nb,
                                                                  int* ipiv, int* info, T* dinv, int* swtab) {
  using R = typename real_of<T>::type;
  extern __shared__ char smem_raw[];
  __shared__ R red_v[THREADS / 64];
  __shared__ int red_i[THREADS / 64];
  __shared__ int s_piv;
  __shared__ R s_val;
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int m = n - k;
  T* Ag = A + (long)b * sA + (long)k * lda + k;
  T* P = reinterpret_cast<T*>(smem_raw);
  for (int idx = tid; idx < m * nb; idx += THREADS) {
    int r = idx / nb, c = idx - r * nb;
    P[r * LU_NB + c] = Ag[(long)r * lda + c];
  }
  __syncthreads();
  int* ipivb = ipiv + (long)b * n;
  for (int j = 0; j < nb; ++j) {
    R bv = R(-1);
    int bi = 0x7fffffff;
    for (int r = j + tid; r < m; r += THREADS) argmax_merge(bv, bi, s_abs1(P[r * LU_NB + j]), r);
    int pr = block_argmax<R, THREADS>(bv, bi, red_v, red_i, &s_piv, &s_val);
    // NaN entries never win a comparison; keep the diagonal then so the
    // pivot row always lies inside the panel.
    if (pr < j || pr >= m) pr = j;
    if (tid == 0) record_pivot(k, j, pr, s_val == R(0), ipivb, info, b);
    if (pr != j) {
      for (int c = tid; c < nb; c += THREADS) {
        T t = P[j * LU_NB + c];
        P[j * LU_NB + c] = P[pr * LU_NB + c];
        P[pr * LU_NB + c] = t;
      }
    }
    __syncthreads();
    T piv = P[j * LU_NB + j];
    if (!s_is_zero(piv)) {
      T inv = s_inv(piv);
      for (int r = j + 1 + tid; r < m; r += THREADS) {
        T l = s_mul(P[r * LU_NB + j], inv);
        P[r * LU_NB + j] = l;
        for (int c = j + 1; c < nb; ++c) P[r * LU_NB + c] = s_fnms(P[r * LU_NB + c], l, P[j * LU_NB + c]);
      }
    }
    __syncthreads();
  }
  for (int idx = tid; idx < m * nb; idx += THREADS) {
    int r = idx / nb, c = idx - r * nb;
    Ag[(long)r * lda + c] = P[r * LU_NB + c];
  }
  // the dynamic allocation holds m + LU_NB rows: the extra rows take the inverse
  finish_panel<T>(P, LU_NB, P + (long)m * LU_NB, LU_NB, k, nb, n, b, ipivb, dinv, swtab);
}

// Tall panels: rows distributed cyclically over THREADS (thread t owns panel
// rows t + THREADS*i, i < RPT); 4-column sub-panels live in registers and are
// brought up to date left-looking before they are factored.  Rows are
// addressed through an LDS row map (logical row -> memory row), so a pivot
// swap costs a register exchange and two LDS writes; the panel is put into
// physical order once at the end (only the <= 2*nb rows the swaps touched
// move).
// Sub-panel width and left-looking prefetch depth keep the per-thread
// register arrays near 48 (x) + 16 (prefetch) VGPRs: a 1024-thread workgroup
// leaves 128 VGPRs per thread, and larger arrays spill to scratch.
template <typename T, int RPT>
constexpr int reg_panel_sw() {
  constexpr int w = 48 / (RPT * (int)(sizeof(T) / 4));
  return w >= 4 ? 4 : (w >= 2 ? 2 : 1);
}
template <typename T, int RPT>
constexpr int reg_panel_jc() {
  constexpr int w = 16 / (RPT * (int)(sizeof(T) / 4));
  return w >= 4 ? 4 : (w >= 2 ? 2 : 1);
}

template <typename T, int RPT, int THREADS>
__global__ __launch_bounds__(THREADS) void getrf_panel_reg_kernel(int n, T* A, long lda, long sA, int k, int nb,
                                                                  int* ipiv, int* info, T* dinv, int* swtab) {
  using R = typename real_of<T>::type;
  constexpr int SW = reg_panel_sw<T, RPT>();
  __shared__ R red_v[THREADS / 64];
  __shared__ int red_i[THREADS / 64];
  __shared__ int s_piv;
  __shared__ R s_val;
  __shared__ int rowmap[RPT * THREADS];
  __shared__ T Lt[LU_NB][LU_NB + 1];  // top nb logical rows of the panel
  __shared__ T Ut[LU_NB][SW];         // U rows above the current sub-panel
  __shared__ T xa[SW], xb[SW];        // row exchange buffers
  __shared__ T prow[SW];
  __shared__ T Dout[LU_NB][LU_NB + 1];
  __shared__ T fix[2 * LU_NB][LU_NB];
  __shared__ int fix_rows[2 * LU_NB];
  __shared__ int fix_cnt;
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int m = n - k;
  T* P = A + (long)b * sA + (long)k * lda + k;
  int* ipivb = ipiv + (long)b * n;
  for (int r = tid; r < m; r += THREADS) rowmap[r] = r;
  __syncthreads();

  T x[RPT][SW];
  for (int c0 = 0; c0 < nb; c0 += SW) {
    const int sw = nb - c0 < SW ? nb - c0 : SW;
    // U rows above: the sub-panel's top c0 rows are staged in LDS in one
    // parallel load, solved with the unit-lower top block, written back
    for (int idx = tid; idx < c0 * sw; idx += THREADS) {
      int j = idx / sw, q = idx - j * sw;
      Ut[j][q] = P[(long)rowmap[j] * lda + c0 + q];
    }
    __syncthreads();
    if (tid < sw) {
      for (int j = 0; j < c0; ++j) {
        T acc = Ut[j][tid];
        for (int i = 0; i < j; ++i) acc = s_fnms(acc, Lt[j][i], Ut[i][tid]);
        Ut[j][tid] = acc;
      }
    }
    __syncthreads();
    for (int idx = tid; idx < c0 * sw; idx += THREADS) {
      int j = idx / sw, q = idx - j * sw;
      P[(long)rowmap[j] * lda + c0 + q] = Ut[j][q];
    }
    // load my rows of the sub-panel and apply the left-looking update; the
    // j-chunk loop is outside the row loop so RPT*4 loads are in flight
    const T* Prow[RPT];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = tid + THREADS * i;
      const bool live = r >= c0 && r < m;
      Prow[i] = live ? P + (long)rowmap[r] * lda : nullptr;
#pragma unroll
      for (int q = 0; q < SW; ++q) x[i][q] = (live && q < sw) ? Prow[i][c0 + q] : s_zero<T>();
    }
    constexpr int JC = reg_panel_jc<T, RPT>();
    for (int j = 0; j < c0; j += JC) {
      T l[RPT][JC];
#pragma unroll
      for (int i = 0; i < RPT; ++i)
#pragma unroll
        for (int t = 0; t < JC; ++t) l[i][t] = (Prow[i] && j + t < c0) ? Prow[i][j + t] : s_zero<T>();
#pragma unroll
      for (int i = 0; i < RPT; ++i)
#pragma unroll
        for (int t = 0; t < JC; ++t)
          if (j + t < c0)
#pragma unroll
            for (int q = 0; q < SW; ++q) x[i][q] = s_fnms(x[i][q], l[i][t], Ut[j + t][q]);
    }
    // factor the sub-panel columns (unrolled: x[i][q] needs compile-time q)
#pragma unroll
    for (int q = 0; q < SW; ++q) {
      if (q >= sw) break;
      const int j = c0 + q;
      R bv = R(-1);
      int bi = 0x7fffffff;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int r = tid + THREADS * i;
        if (r >= j && r < m) argmax_merge(bv, bi, s_abs1(x[i][q]), r);
      }
      int pr = block_argmax<R, THREADS>(bv, bi, red_v, red_i, &s_piv, &s_val);
      if (pr < j || pr >= m) pr = j;
      if (tid == 0) record_pivot(k, j, pr, s_val == R(0), ipivb, info, b);
      if (pr != j) {
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          const int r = tid + THREADS * i;
          if (r == j)
#pragma unroll
            for (int t = 0; t < SW; ++t) xa[t] = x[i][t];
          if (r == pr)
#pragma unroll
            for (int t = 0; t < SW; ++t) xb[t] = x[i][t];
        }
        if (tid == 0) {
          int t = rowmap[j];
          rowmap[j] = rowmap[pr];
          rowmap[pr] = t;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          const int r = tid + THREADS * i;
          if (r == j)
#pragma unroll
            for (int t = 0; t < SW; ++t) x[i][t] = xb[t];
          if (r == pr)
#pragma unroll
            for (int t = 0; t < SW; ++t) x[i][t] = xa[t];
        }
      }
      // pivot row broadcast
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int r = tid + THREADS * i;
        if (r == j)
#pragma unroll
          for (int t = 0; t < SW; ++t) prow[t] = x[i][t];
      }
      __syncthreads();
      T piv = prow[q];
      if (!s_is_zero(piv)) {
        T inv = s_inv(piv);
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          const int r = tid + THREADS * i;
          if (r > j && r < m) {
            T l = s_mul(x[i][q], inv);
            x[i][q] = l;
#pragma unroll
            for (int t = 0; t < SW; ++t)
              if (t > q && t < sw) x[i][t] = s_fnms(x[i][t], l, prow[t]);
          }
        }
      }
      __syncthreads();
    }
    // write the sub-panel back; refresh the top block
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = tid + THREADS * i;
      if (r >= c0 && r < m) {
        T* Pr = P + (long)rowmap[r] * lda;
#pragma unroll
        for (int q = 0; q < SW; ++q)
          if (q < sw) Pr[c0 + q] = x[i][q];
      }
    }
    __syncthreads();
    for (int idx = tid; idx < nb * (c0 + sw); idx += THREADS) {
      int r = idx / (c0 + sw), c = idx - r * (c0 + sw);
      Lt[r][c] = P[(long)rowmap[r] * lda + c];
    }
    __syncthreads();
  }
  // physical order: move the rows whose memory row differs from their position
  if (tid == 0) fix_cnt = 0;
  __syncthreads();
  for (int r = tid; r < m; r += THREADS) {
    if (rowmap[r] != r) {
      int slot = atomicAdd(&fix_cnt, 1);
      if (slot < 2 * LU_NB) fix_rows[slot] = r;
    }
  }
  __syncthreads();
  const int cnt = fix_cnt < 2 * LU_NB ? fix_cnt : 2 * LU_NB;
  for (int idx = tid; idx < cnt * nb; idx += THREADS) {
    int a = idx / nb, c = idx - a * nb;
    fix[a][c] = P[(long)rowmap[fix_rows[a]] * lda + c];
  }
  __syncthreads();
  for (int idx = tid; idx < cnt * nb; idx += THREADS) {
    int a = idx / nb, c = idx - a * nb;
    P[(long)fix_rows[a] * lda + c] = fix[a][c];
  }
  // Lt holds the whole factored diagonal block
  finish_panel<T>(&Lt[0][0], LU_NB + 1, &Dout[0][0], LU_NB + 1, k, nb, n, b, ipivb, dinv, swtab);
}

// Register panel state in LDS: each wave's argmax candidate (|value|,
// position, and its row from the column on), double-buffered by column parity
// so a column needs a single barrier.
template <typename T, int NW>
struct PanelLDS {
  using R = typename real_of<T>::type;
  R cv[2][NW];
  int cp[2][NW];
  T crow[2][NW][LU_NB];
};

// One column step J of partial pivoting over rows held in registers, RPT rows
// per thread of a workgroup of NW waves.  Rows never move between threads: each
// row's current position in the partially pivoted panel is tracked in pos[]
// (a row swap is two position updates).  Each wave publishes its best
// candidate row; after the one barrier every thread picks the winner among the
// NW candidates and reads the pivot row from its slot.  J is a template parameter so
// every access to x[] is resolved at compile time.  on_pivot(J, pq, |pivot|,
// mask) runs after the step with the pivot's position pq (J when no candidate
// is left: the column is NaN or no live row remains) and the mask of this
// thread's rows that held it.
template <typename T, int RPT, int NW, int J, bool NOPIV, class F>
__device__ __forceinline__ void panel_step(T (&x)[RPT][LU_NB], int (&pos)[RPT], const bool (&live)[RPT],
                                           PanelLDS<T, NW>& sh, int nb, F& on_pivot,
                                           typename real_of<T>::type& lmax) {
  using R = typename real_of<T>::type;
  constexpr int NONE = 0x7fffffff;
  constexpr int par = J & 1;
  if (J >= nb) return;
  if constexpr (NOPIV) {
    // natural order: the row at position J is the pivot row; lmax keeps the
    // largest multiplier magnitude for the caller's stability check
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      if (live[r] && pos[r] == J) {
#pragma unroll
        for (int c = J; c < LU_NB; ++c) sh.crow[par][0][c] = x[r][c];
      }
    __syncthreads();
    const T* prow = sh.crow[par][0];
    const T piv = prow[J];
    on_pivot(J, J, s_abs1(piv), 0u);
    if (!s_is_zero(piv)) {
      const T inv = s_inv_uniform(piv);
#pragma unroll
      for (int r = 0; r < RPT; ++r)
        if (live[r] && pos[r] > J) {
          const T l = s_mul(x[r][J], inv);
          const R al = s_abs1(l);
          lmax = (al > lmax || al != al) ? al : lmax;
          x[r][J] = l;
#pragma unroll
          for (int c = J + 1; c < LU_NB; ++c) x[r][c] = s_fnms(x[r][c], l, prow[c]);
        }
    }
    return;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  R bv = R(-1);
  int bp = NONE;
#pragma unroll
  for (int r = 0; r < RPT; ++r)
    if (live[r] && pos[r] >= J) argmax_merge(bv, bp, s_abs1(x[r][J]), pos[r]);
  // ties go to the lower position, as in a sequential scan; NaN never wins
  wave_argmax(bv, bp);
#pragma unroll
  for (int r = 0; r < RPT; ++r)
    if (pos[r] == bp) {
#pragma unroll
      for (int c = J; c < LU_NB; ++c) sh.crow[par][wave][c] = x[r][c];
    }
  if (lane == 0) {
    sh.cv[par][wave] = bv;
    sh.cp[par][wave] = bp;
  }
  __syncthreads();
  // selects, not branches: every candidate's read is unconditional, so the
  // reads are issued together
  R fv = sh.cv[par][0];
  int fp = sh.cp[par][0], fw = 0;
#pragma unroll
  for (int w = 1; w < NW; ++w) {
    const R v = sh.cv[par][w];
    const int q = sh.cp[par][w];
    const bool take = v > fv || (v == fv && q < fp);
    fv = take ? v : fv;
    fp = take ? q : fp;
    fw = take ? w : fw;
  }
  // without a candidate the row at position J stays; the column's updates
  // are then NaN whatever the pivot slot holds
  const int pq = fp != NONE ? fp : J;
  unsigned mask = 0;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    if (pos[r] == pq) {
      mask |= 1u << r;
      pos[r] = J;
    } else if (pos[r] == J) {
      pos[r] = pq;
    }
  }
  on_pivot(J, pq, fv, mask);
  const T* prow = sh.crow[par][fw];
  const T piv = prow[J];
  if (!s_is_zero(piv)) {
    const T inv = s_inv_uniform(piv);
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      if (live[r] && pos[r] > J) {
        const T l = s_mul(x[r][J], inv);
        x[r][J] = l;
#pragma unroll
        for (int c = J + 1; c < LU_NB; ++c) x[r][c] = s_fnms(x[r][c], l, prow[c]);
      }
  }
}

template <typename T, int RPT, int NW, bool NOPIV = false, class F, int... J>
__device__ __forceinline__ void panel_steps(std::integer_sequence<int, J...>, T (&x)[RPT][LU_NB], int (&pos)[RPT],
                                            const bool (&live)[RPT], PanelLDS<T, NW>& sh, int nb, F& on_pivot,
                                            typename real_of<T>::type& lmax) {
  (panel_step<T, RPT, NW, J, NOPIV>(x, pos, live, sh, nb, on_pivot, lmax), ...);
}

// Register panels (m <= THREADS RPT): thread t owns panel rows t + THREADS r
// in registers for the whole factorisation and writes each to its final
// position at the end.  No panel storage in LDS.  Columns beyond nb are zero
// and stay zero (their updates subtract l * 0).  The workgroup then applies
// the panel's row permutation to the rest of its matrix itself.
template <typename T, int RPT, int THREADS, bool NOPIV = false>
__global__ __launch_bounds__(THREADS) void getrf_panel_rows_kernel(int n, T* A, long lda, long sA, int k, int nb,
                                                                   int* ipiv, int* info, T* dinv, int* swtab) {
  using R = typename real_of<T>::type;
  __shared__ PanelLDS<T, THREADS / 64> sh;
  __shared__ T Dblk[LU_NB][LU_NB + 1];
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  const int m = n - k;
  T* P = A + (long)b * sA + (long)k * lda + k;
  int* ipivb = ipiv + (long)b * n;
  T x[RPT][LU_NB];
  int pos[RPT];
  bool live[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    pos[r] = t + THREADS * r;
    live[r] = pos[r] < m;
#pragma unroll
    for (int c = 0; c < LU_NB; ++c) x[r][c] = (live[r] && c < nb) ? P[(long)pos[r] * lda + c] : s_zero<T>();
  }
  auto on_pivot = [&](int J, int pq, R fv, unsigned) {
    if (t == 0) record_pivot(k, J, pq, fv == R(0), ipivb, info, b);
  };
  R lmax = R(0);
  panel_steps<T, RPT, THREADS / 64, NOPIV>(std::make_integer_sequence<int, LU_NB>{}, x, pos, live, sh, nb, on_pivot,
                                            lmax);
  if constexpr (NOPIV) {
    // stability check of the natural-order elimination: a multiplier above
    // LU_NOPIV_BOUND (or not finite) marks the matrix with info = -1, and the
    // caller refactors with partial pivoting
    __shared__ R wmax[THREADS / 64];
    R v = lmax;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const R o = __shfl_xor(v, off, 64);
      v = (o > v || o != o) ? o : v;
    }
    if ((t & 63) == 0) wmax[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
      R mx = wmax[0];
      for (int w = 1; w < THREADS / 64; ++w) mx = (wmax[w] > mx || wmax[w] != wmax[w]) ? wmax[w] : mx;
      // a zero pivot in the natural order (info > 0) is no verdict on the
      // matrix either: pivoting decides
      if (!(mx <= R(LU_NOPIV_BOUND)) || info[b] > 0) info[b] = -1;
    }
  }
  // net row permutation of the panel: every row whose position changed
  __shared__ int tab[SWTAB];
  if (t == 0) tab[0] = 0;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int orig = t + THREADS * r;
    if (live[r]) {
#pragma unroll
      for (int c = 0; c < LU_NB; ++c)
        if (c < nb) P[(long)pos[r] * lda + c] = x[r][c];
      if (pos[r] != orig) {
        // at most 2 nb rows move (two per swap)
        const int i = atomicAdd(&tab[0], 1);
        if (i < 2 * LU_NB) {
          tab[1 + 2 * i] = k + pos[r];
          tab[2 + 2 * i] = k + orig;
        }
      }
    }
    if (pos[r] < nb)
#pragma unroll
      for (int c = 0; c < LU_NB; ++c) Dblk[pos[r]][c] = c < nb ? x[r][c] : s_zero<T>();
  }
  finish_panel_tab<T>(&Dblk[0][0], LU_NB + 1, nullptr, 0, k, nb, n, b, tab, dinv, swtab);
  if constexpr (NOPIV) return;  // no row moved
  // the panel's row permutation on every other column of the matrix, CW
  // columns at a time: all moved rows are read before any is written
  constexpr int CW = THREADS >= 128 ? 32 : 16, RG = THREADS / CW, NV = 2 * LU_NB / RG;
  const int cnt = tab[0] < 2 * LU_NB ? tab[0] : 2 * LU_NB;
  const int others = n - nb;
  T* Ab = A + (long)b * sA;
  const int cc = t % CW, a0 = t / CW;
  for (int c0 = 0; c0 < others; c0 += CW) {
    const int oc = c0 + cc;
    const int col = oc < k ? oc : oc + nb;
    const bool ok = oc < others;
    T v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int a = a0 + RG * i;
      v[i] = (ok && a < cnt) ? Ab[(long)tab[2 + 2 * a] * lda + col] : s_zero<T>();
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int a = a0 + RG * i;
      if (ok && a < cnt) Ab[(long)tab[1 + 2 * a] * lda + col] = v[i];
    }
  }
}

// Apply the panel's row permutation table to columns outside the panel and,
// for the trailing columns, U12 = inv(L11) A12.  256 threads = 32 rows x 8
// columns per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void getrf_swap_u12_kernel(int n, T* A, long lda, long sA, int k, int nb,
                                                             const T* dinv, const int* swtab, int u12_end) {
  __shared__ T Li[LU_NB][LU_NB + 1];
  __shared__ T vals[2 * LU_NB][8];
  __shared__ T X[LU_NB][8];
  __shared__ int tab[SWTAB];
  const int b = blockIdx.y;
  T* Ab = A + (long)b * sA;
  const int others = n - nb;
  const int cbase = blockIdx.x * 8;
  const int tid = threadIdx.x;
  const int cc = tid & 7, rr = tid >> 3;
  const int t = cbase + cc;
  const int c = t < k ? t : t + nb;  // column index in A
  const bool valid = t < others;
  for (int i = tid; i < SWTAB; i += 256) tab[i] = swtab[(long)b * SWTAB + i];
  __syncthreads();
  const int cnt = tab[0];
  // gather the permuted rows, then scatter
  for (int a = rr; a < cnt; a += 32)
    if (valid) vals[a][cc] = Ab[(long)tab[2 + 2 * a] * lda + c];
  __syncthreads();
  for (int a = rr; a < cnt; a += 32)
    if (valid) Ab[(long)tab[1 + 2 * a] * lda + c] = vals[a][cc];
  __syncthreads();
  // U12 for the trailing columns up to u12_end (the end of the current column
  // block); slabs never straddle k or u12_end - nb (multiples of 8)
  if (cbase >= k && cbase + nb < u12_end) {
    const int nblk = (n + LU_NB - 1) / LU_NB;
    const T* db = dinv + ((long)b * nblk + k / LU_NB) * LU_NB * LU_NB;
    for (int idx = tid; idx < LU_NB * LU_NB; idx += 256) {
      int r = idx / LU_NB, q = idx - r * LU_NB;
      Li[r][q] = q < r ? db[idx] : (q == r ? s_one<T>() : s_zero<T>());
    }
    const bool in_u12 = valid && c < u12_end;
    X[rr][cc] = (in_u12 && rr < nb) ? Ab[(long)(k + rr) * lda + c] : s_zero<T>();
    __syncthreads();
    if (in_u12 && rr < nb) {
      T acc = s_zero<T>();
      for (int q = 0; q <= rr; ++q) acc = s_add(acc, s_mul(Li[rr][q], X[q][cc]));
      Ab[(long)(k + rr) * lda + c] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Tall panels of a few matrices (the top merge levels): tournament pivoting
// (communication-avoiding LU, Grigori, Demmel & Xiang, SIAM J. Matrix Anal.
// Appl. 32 (2011)).  Every 256-row chunk of the panel picks nb candidate
// pivot rows by partial pivoting on its own rows (one workgroup per chunk,
// so the whole GPU works on one panel); groups of 8 candidate sets are
// reduced the same way until nb winners remain.  The winners are swapped to
// the top in their selection order, the top block is factored without
// further pivoting (it reproduces the final round's elimination), and every
// other row gets its multipliers L21 = A21 inv(U11) independently.
// ---------------------------------------------------------------------------

// Partial pivoting over the rows held one per thread of a THREADS-wide
// workgroup (x = row, valid if row_id >= 0): the nb pivot rows' ids, in
// order, to out[0..nb) (-1 once no valid row remains); pos ends as the row's
// position in the eliminated candidate block, its x as that row of the block's
// LU.  With info set, an exactly zero pivot is reported there.
template <typename T, int THREADS>
__device__ __forceinline__ void tslu_select(T (&x)[1][LU_NB], int (&pos)[1], int row_id, int nb, int* out,
                                            int* info, int b, int k) {
  using R = typename real_of<T>::type;
  __shared__ PanelLDS<T, THREADS / 64> sh;
  pos[0] = (int)threadIdx.x;
  const bool live[1] = {row_id >= 0};
  auto on_pivot = [&](int J, int, R fv, unsigned mask) {
    if (mask) out[J] = row_id;
    if (info && threadIdx.x == 0 && fv == R(0) && info[b] == 0) info[b] = k + J + 1;
  };
  typename real_of<T>::type lmax_unused = 0;
  panel_steps<T, 1, THREADS / 64>(std::make_integer_sequence<int, LU_NB>{}, x, pos, live, sh, nb, on_pivot,
                                   lmax_unused);
  __syncthreads();
}

// Round 0: chunk c of the panel rows [k, n) -> cand_out[b][c][0..nb)
template <typename T>
__global__ __launch_bounds__(256) void tslu_leaf_kernel(int n, const T* A, long lda, long sA, int k, int nb,
                                                        int nchunks, int* cand_out) {
  __shared__ int out[LU_NB];
  const int b = blockIdx.y, c = blockIdx.x, t = threadIdx.x;
  const int r = k + c * TS_CHUNK + t;
  const T* Ab = A + (long)b * sA;
  T x[1][LU_NB];
  int pos[1];
  const bool live = r < n;
#pragma unroll
  for (int q = 0; q < LU_NB; ++q) x[0][q] = (live && q < nb) ? Ab[(long)r * lda + k + q] : s_zero<T>();
  tslu_select<T, 256>(x, pos, live ? r : -1, nb, out, nullptr, b, k);
  if (t < nb) cand_out[((long)b * nchunks + c) * LU_NB + t] = out[t];
}

// Later rounds: group g merges up to TS_GROUP candidate sets of cand_in
// (nsets per matrix).  The last round (one group) also finishes the panel:
// its winners' rows already hold the LU of the top block in selection order
// (the elimination the unpivoted factorisation would repeat), so it writes
// them to the top rows, runs the LAPACK-style swap sequence on the at most 64
// positions it touches (one lane each in the first wave: lane j < 32 is
// position k+j, lane 32+j the original position of winner j below the top
// block), moves the displaced top rows, and writes dinv and the swap table.
constexpr int TS_GROUP = 16;

template <typename T, bool LAST, int THREADS = 512>
__global__ __launch_bounds__(THREADS) void tslu_reduce_kernel(int n, T* A, long lda, long sA, int k, int nb, int nsets,
                                                          const int* cand_in, int ngroups, int* cand_out, int* ipiv,
                                                          int* info, T* dinv, int* swtab) {
  __shared__ int out[LU_NB];
  const int b = blockIdx.y, g = blockIdx.x, t = threadIdx.x;
  const int set = g * TS_GROUP + t / LU_NB, slot = t % LU_NB;
  int r = -1;
  if (set < nsets && slot < nb) r = cand_in[((long)b * nsets + set) * LU_NB + slot];
  T* Ab = A + (long)b * sA;
  T x[1][LU_NB];
  int pos[1];
#pragma unroll
  for (int q = 0; q < LU_NB; ++q) x[0][q] = (r >= 0 && q < nb) ? Ab[(long)r * lda + k + q] : s_zero<T>();
  tslu_select<T, THREADS>(x, pos, r, nb, out, LAST ? info : nullptr, b, k);
  if constexpr (!LAST) {
    if (t < nb) cand_out[((long)b * ngroups + g) * LU_NB + t] = out[t];
  } else {
    __shared__ int tab[SWTAB];
    __shared__ T vals[2 * LU_NB][LU_NB];
    __shared__ T Dblk[LU_NB][LU_NB + 1];
    __shared__ T Dout[LU_NB][LU_NB + 1];
    int* ipivb = ipiv + (long)b * n;
    const int lane = t;
    auto winner = [&](int j) {
      const int w = out[j];
      return (w < k || w >= n) ? k + j : w;  // no candidate left (cannot happen for m >= nb): keep the row
    };
    if (t < 64) {
      int spos = -1;
      if (lane < nb) {
        spos = k + lane;
      } else if (lane >= LU_NB && lane - LU_NB < nb) {
        const int w = winner(lane - LU_NB);
        if (w >= k + nb) spos = w;
      }
      int scont = spos;
      for (int j = 0; j < nb; ++j) {
        const int w = winner(j);
        const unsigned long long hit = __ballot(spos >= 0 && scont == w);
        if (hit == 0ull) {
          if (lane == 0) ipivb[k + j] = k + j;  // not reached: every winner holds one lane
          continue;
        }
        const int L = __ffsll((long long)hit) - 1;
        const int p = __shfl(spos, L);
        if (lane == 0) ipivb[k + j] = p;
        const int cj = __shfl(scont, j), cL = __shfl(scont, L);
        if (lane == j)
          scont = cL;
        else if (lane == L)
          scont = cj;
      }
      const bool moved = spos >= 0 && scont != spos;
      const unsigned long long mv = __ballot(moved);
      if (moved) {
        const int i = __popcll(mv & ((1ull << lane) - 1ull));
        tab[1 + 2 * i] = spos;
        tab[2 + 2 * i] = scont;
      }
      if (lane == 0) tab[0] = __popcll(mv);
    }
    __syncthreads();
    const int cnt = tab[0];
    // rows displaced below the top block come from original top rows: read
    // their panel columns before the winners overwrite the top rows
    for (int idx = t; idx < cnt * LU_NB; idx += THREADS) {
      const int a = idx / LU_NB, c = idx - a * LU_NB;
      if (tab[1 + 2 * a] >= k + nb && c < nb) vals[a][c] = Ab[(long)tab[2 + 2 * a] * lda + k + c];
    }
    __syncthreads();
    for (int idx = t; idx < cnt * LU_NB; idx += THREADS) {
      const int a = idx / LU_NB, c = idx - a * LU_NB;
      if (tab[1 + 2 * a] >= k + nb && c < nb) Ab[(long)tab[1 + 2 * a] * lda + k + c] = vals[a][c];
    }
    if (r >= 0 && pos[0] < nb) {
#pragma unroll
      for (int c = 0; c < LU_NB; ++c) {
        if (c < nb) Ab[(long)(k + pos[0]) * lda + k + c] = x[0][c];
        Dblk[pos[0]][c] = c < nb ? x[0][c] : s_zero<T>();
      }
    }
    finish_panel_tab<T>(&Dblk[0][0], LU_NB + 1, &Dout[0][0], LU_NB + 1, k, nb, n, b, tab, dinv, swtab);
  }
}

// L21 = A21 inv(U11) for the panel rows below the top block, one row per thread
template <typename T>
__global__ __launch_bounds__(256) void tslu_l21_kernel(int n, T* A, long lda, long sA, int k, int nb,
                                                       const T* dinv) {
  __shared__ T Ui[LU_NB][LU_NB + 1];
  const int b = blockIdx.y, t = threadIdx.x;
  const int nblk = (n + LU_NB - 1) / LU_NB;
  const T* db = dinv + ((long)b * nblk + k / LU_NB) * LU_NB * LU_NB;
  for (int idx = t; idx < LU_NB * LU_NB; idx += 256) {
    const int r = idx / LU_NB, c = idx - r * LU_NB;
    Ui[r][c] = c >= r ? db[idx] : s_zero<T>();
  }
  __syncthreads();
  const int r = k + nb + blockIdx.x * 256 + t;
  if (r >= n) return;
  T* row = A + (long)b * sA + (long)r * lda + k;
  T a[LU_NB];
#pragma unroll
  for (int q = 0; q < LU_NB; ++q) a[q] = q < nb ? row[q] : s_zero<T>();
#pragma unroll
  for (int c = 0; c < LU_NB; ++c) {
    if (c < nb) {
      T acc = s_zero<T>();
#pragma unroll
      for (int q = 0; q <= c; ++q) acc = s_add(acc, s_mul(a[q], Ui[q][c]));
      row[c] = acc;
    }
  }
}

template <typename T>
static hipError_t launch_tslu_panel(int n, T* A, long lda, long sA, int k, int nb, int* ipiv, int* info, T* dinv,
                                    int* swtab, int* cand, int batch, hipStream_t s) {
  const int m = n - k;
  int nsets = (m + TS_CHUNK - 1) / TS_CHUNK;
  int* cur = cand;
  int* nxt = cand + tslu_cand_ints(n, batch);
  hipLaunchKernelGGL((tslu_leaf_kernel<T>), dim3(nsets, batch), dim3(256), 0, s, n, A, lda, sA, k, nb, nsets, cur);
  for (;;) {
    const int ngroups = (nsets + TS_GROUP - 1) / TS_GROUP;
    if (ngroups == 1) {
      // up to 8 candidate sets fit 256 threads: one wave per SIMD instead of two
      if (nsets * LU_NB <= 256)
        hipLaunchKernelGGL((tslu_reduce_kernel<T, true, 256>), dim3(1, batch), dim3(256), 0, s, n, A, lda, sA, k, nb,
                           nsets, cur, 1, nullptr, ipiv, info, dinv, swtab);
      else
        hipLaunchKernelGGL((tslu_reduce_kernel<T, true>), dim3(1, batch), dim3(512), 0, s, n, A, lda, sA, k, nb, nsets,
                           cur, 1, nullptr, ipiv, info, dinv, swtab);
      break;
    }
    hipLaunchKernelGGL((tslu_reduce_kernel<T, false>), dim3(ngroups, batch), dim3(512), 0, s, n, A, lda, sA, k, nb,
                       nsets, cur, ngroups, nxt, ipiv, info, dinv, swtab);
    int* tmp = cur;
    cur = nxt;
    nxt = tmp;
    nsets = ngroups;
  }
  if (m > nb)
    hipLaunchKernelGGL((tslu_l21_kernel<T>), dim3(ceil_div(m - nb, 256), batch), dim3(256), 0, s, n, A, lda, sA, k,
                       nb, dinv);
  return hipGetLastError();
}

template <typename T, int RPT>
static void launch_reg_panel(int n, T* A, long lda, long sA, int k, int nb, int* ipiv, int* info, T* dinv,
                             int* swtab, int batch, hipStream_t s) {
  hipLaunchKernelGGL((getrf_panel_reg_kernel<T, RPT, 1024>), dim3(batch), dim3(1024), 0, s, n, A, lda, sA, k, nb,
                     ipiv, info, dinv, swtab);
}

// rows the register panel holds: one per thread, up to 1024 threads (512 for
// complex128, whose 32-column rows take 128 registers)
template <typename T>
static int rows_panel_max() {
  return sizeof(T) >= 16 ? 512 : 1024;
}

// ---------------------------------------------------------------------------
// Tall panels without pivoting, for matrices whose elimination is stable in
// the natural order (the merge matrices of the HPS tree: the ItI W = I -
// D12 D21 and the DtN interface blocks show pivot growth ~1 and multipliers
// below 4; the caller checks the factors and refactors with pivoting
// otherwise).  Every workgroup factors the nb x nb diagonal block (one wave,
// rows in registers; redundant work costs less than a second launch) and
// inverts its U11; its NOPIV_ROWS panel rows below the block then take
// L21 = A21 inv(U11), four lanes per row, each entry an independent dot product over
// the row held in the quad's registers.  The last workgroup of the matrix to
// read the diagonal block (a per-matrix arrival counter) writes the factored
// block, its packed inverse (dinv) and the identity pivots.  The rows are
// independent: grid
// (ceil((n - k - nb) / NOPIV_ROWS), batch).
// ---------------------------------------------------------------------------
// lane `lane`'s value of v in every lane of the wave (all lanes active)
__device__ __forceinline__ float lane_read(float v, int lane) { return __shfl(v, lane, 64); }
__device__ __forceinline__ double lane_read(double v, int lane) { return __shfl(v, lane, 64); }
__device__ __forceinline__ c64 lane_read(c64 v, int lane) { return c64{lane_read(v.re, lane), lane_read(v.im, lane)}; }
__device__ __forceinline__ c128 lane_read(c128 v, int lane) {
  return c128{lane_read(v.re, lane), lane_read(v.im, lane)};
}

constexpr int NOPIV_ROWS = 64;  // panel rows per workgroup of the unpivoted panel

// the value of lane CTRL & 3 of each quad (CTRL = 0x55 * lane: quad_perm [l,l,l,l])
template <int CTRL>
__device__ __forceinline__ float quad_bcast(float v) { return dpp_move<CTRL, 0xf>(0.0f, v); }
template <int CTRL>
__device__ __forceinline__ double quad_bcast(double v) { return dpp_move<CTRL, 0xf>(0.0, v); }
template <int CTRL>
__device__ __forceinline__ c64 quad_bcast(c64 v) { return c64{quad_bcast<CTRL>(v.re), quad_bcast<CTRL>(v.im)}; }
template <int CTRL>
__device__ __forceinline__ c128 quad_bcast(c128 v) { return c128{quad_bcast<CTRL>(v.re), quad_bcast<CTRL>(v.im)}; }

// x[j] += sum_t a_S[t] W[8 S + t][8 sub + j], a_S = lane S's entries of the quad's row
template <int S, typename T>
__device__ __forceinline__ void nopiv_rows_accum(const T (&a)[8], T (&x)[8], const T (*W)[LU_NB + 1], int sub) {
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const T ai = quad_bcast<S * 0x55>(a[t]);
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = s_fma(ai, W[8 * S + t][8 * sub + j], x[j]);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void getrf_panel_nopiv_kernel(int n, T* A, long lda, long sA, int k, int nb,
                                                                int* ipiv, T* dinv, int* arrivals, int row_wgs,
                                                                int u12_end) {
  __shared__ T D[LU_NB][LU_NB + 1];
  __shared__ T W[LU_NB][LU_NB + 1];  // inv(U11)
  __shared__ T P[LU_NB][LU_NB + 1];  // packed inverse
  __shared__ int last;
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  T* Ab = A + (long)b * sA;
  for (int idx = tid; idx < LU_NB * LU_NB; idx += 256) {
    const int r = idx / LU_NB, c = idx - r * LU_NB;
    D[r][c] = (r < nb && c < nb) ? Ab[(long)(k + r) * lda + k + c] : (r == c ? s_one<T>() : s_zero<T>());
  }
  __syncthreads();
  // every workgroup of the matrix has read the original block once the
  // counter reaches gridDim.x: only the last to arrive may overwrite it with
  // the factors (workgroups of one matrix need not be resident together)
  if (tid == 0) {
    const int before = atomicAdd(&arrivals[b], 1);
    last = before == (int)gridDim.x - 1;
    if (last) atomicExch(&arrivals[b], 0);  // every workgroup has arrived: ready for the next launch
  }
  __syncthreads();
  // elimination of the diagonal block by wave 0 alone, no barriers: lane r
  // holds row r in registers; step j reads the pivot row from lane j
  // (cross-lane shuffles) and the rows below subtract their multiple of it.
  // The factored block (multipliers below the diagonal) goes back to D, zero
  // outside nb x nb for the inversion.
  if (tid < 64) {
    const int r = tid;
    T x[LU_NB];
#pragma unroll
    for (int c = 0; c < LU_NB; ++c) x[c] = r < LU_NB ? D[r][c] : s_zero<T>();
#pragma unroll
    for (int j = 0; j < LU_NB - 1; ++j) {
      // rows at or above j (and the idle lanes) take a zero multiplier, so the
      // step runs without divergence and lane j's row stays as it is; the
      // shuffles stay outside any branch (lane j must take part for its value
      // to be read)
      const bool below = r > j && r < LU_NB;
      const T inv = s_inv(lane_read(x[j], j));
      const T l = below ? s_mul(x[j], inv) : s_zero<T>();
      x[j] = below ? l : x[j];
#pragma unroll
      for (int c = j + 1; c < LU_NB; ++c) x[c] = s_fnms(x[c], l, lane_read(x[c], j));
    }
    if (r < LU_NB) {
#pragma unroll
      for (int c = 0; c < LU_NB; ++c) D[r][c] = (r < nb && c < nb) ? x[c] : s_zero<T>();
    }
  }
  __syncthreads();
  invert_diag_block<T>(&D[0][0], LU_NB + 1, &P[0][0], LU_NB + 1, nb);
  __syncthreads();
  for (int idx = tid; idx < LU_NB * LU_NB; idx += 256) {
    const int r = idx / LU_NB, c = idx - r * LU_NB;
    W[r][c] = (r <= c && c < nb) ? P[r][c] : s_zero<T>();  // P holds garbage beyond nb
  }
  __syncthreads();
  if (blockIdx.x >= row_wgs) {
    // U12 = inv(L11) A12 for 32 columns of the block row right of the panel
    // (columns up to u12_end): the chunk staged in W (free here), then each
    // entry a dot product with the unit-lower inv(L11) held in P
    const int c0 = k + nb + (int)(blockIdx.x - row_wgs) * LU_NB;
    const int wc = u12_end - c0 < LU_NB ? u12_end - c0 : LU_NB;
    for (int idx = tid; idx < LU_NB * LU_NB; idx += 256) {
      const int r = idx / LU_NB, c = idx - r * LU_NB;
      W[r][c] = (r < nb && c < wc) ? Ab[(long)(k + r) * lda + c0 + c] : s_zero<T>();
    }
    __syncthreads();
    for (int idx = tid; idx < LU_NB * LU_NB; idx += 256) {
      const int r = idx / LU_NB, c = idx - r * LU_NB;
      if (r < nb && c < wc) {
        T acc = W[r][c];
        for (int q = 0; q < r; ++q) acc = s_fma(P[r][q], W[q][c], acc);
        Ab[(long)(k + r) * lda + c0 + c] = acc;
      }
    }
  } else {
    // panel rows below the block: x = a inv(U11), four lanes per row, each
    // holding eight consecutive entries of the row; entry i comes from lane
    // i / 8 of the quad by a DPP broadcast, and each lane forms the row's
    // outputs 8 sub .. 8 sub + 7
    const int sub = tid & 3;
    const int rg = k + nb + blockIdx.x * NOPIV_ROWS + (tid >> 2);
    // a quad's lanes share their row, so whole quads are live or not and the
    // broadcasts read live lanes only
    if (rg < n) {
      T* rowp = Ab + (long)rg * lda + k + 8 * sub;
      T a[8], x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        a[j] = 8 * sub + j < nb ? rowp[j] : s_zero<T>();
        x[j] = s_zero<T>();
      }
      nopiv_rows_accum<0>(a, x, W, sub);
      nopiv_rows_accum<1>(a, x, W, sub);
      nopiv_rows_accum<2>(a, x, W, sub);
      nopiv_rows_accum<3>(a, x, W, sub);
      if (nb == LU_NB) {
#pragma unroll
        for (int j = 0; j < 8; ++j) rowp[j] = x[j];
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (8 * sub + j < nb) rowp[j] = x[j];
      }
    }
  }
  if (!last) return;
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int rr = idx / nb, c = idx - rr * nb;
    Ab[(long)(k + rr) * lda + k + c] = D[rr][c];
  }
  if (tid < nb) ipiv[(long)b * n + k + tid] = k + tid;
  const int nblk = (n + LU_NB - 1) / LU_NB;
  T* db = dinv + ((long)b * nblk + k / LU_NB) * LU_NB * LU_NB;
  for (int idx = tid; idx < LU_NB * LU_NB; idx += 256) {
    const int r = idx / LU_NB, c = idx - r * LU_NB;
    db[idx] = (r < nb && c < nb) ? P[r][c] : s_zero<T>();
  }
}

// *swapped: the panel kernel already applied its row permutation to the other
// columns (the register panels do).
template <typename T>
static hipError_t launch_panel(int n, T* A, long lda, long sA, int k, int nb, int* ipiv, int* info, T* dinv,
                               int* swtab, int* cand, int batch, hipStream_t s, bool* swapped,
                               bool unpivoted = false) {
  const int m = n - k;
  const double pf = flop_mult<T>() * batch * ((double)m * nb * nb - (double)nb * nb * nb / 3.0);
  *swapped = m <= rows_panel_max<T>();
  if (unpivoted && m <= rows_panel_max<T>()) {
    // natural-order register panels (checked: info = -1 past LU_NOPIV_BOUND)
    static const std::string nm_rows = std::string("getrf_panel_rows_nopiv<") + tname<T>() + ">";
    prof::Scope ps(s, nm_rows.c_str(), pf, 2.0 * sizeof(T) * (double)m * nb * batch);
    if (m <= 128)
      hipLaunchKernelGGL((getrf_panel_rows_kernel<T, 1, 128, true>), dim3(batch), dim3(128), 0, s, n, A, lda, sA, k,
                         nb, ipiv, info, dinv, swtab);
    else if (m <= 256)
      hipLaunchKernelGGL((getrf_panel_rows_kernel<T, 1, 256, true>), dim3(batch), dim3(256), 0, s, n, A, lda, sA, k,
                         nb, ipiv, info, dinv, swtab);
    else if (m <= 512)
      hipLaunchKernelGGL((getrf_panel_rows_kernel<T, 1, 512, true>), dim3(batch), dim3(512), 0, s, n, A, lda, sA, k,
                         nb, ipiv, info, dinv, swtab);
    else if constexpr (sizeof(T) < 16)
      hipLaunchKernelGGL((getrf_panel_rows_kernel<T, 1, 1024, true>), dim3(batch), dim3(1024), 0, s, n, A, lda, sA,
                         k, nb, ipiv, info, dinv, swtab);
    else
      return hipErrorInvalidValue;  // not reached: rows_panel_max<c128>() is 512
    return hipGetLastError();
  }
  if (m <= rows_panel_max<T>()) {
    static const std::string nm_rows = std::string("getrf_panel_rows<") + tname<T>() + ">";
    prof::Scope ps(s, nm_rows.c_str(), pf, 2.0 * sizeof(T) * (double)m * nb * batch);
    // short panels in smaller workgroups: more of them share a CU
    if (m <= 128)
      hipLaunchKernelGGL((getrf_panel_rows_kernel<T, 1, 128>), dim3(batch), dim3(128), 0, s, n, A, lda, sA, k, nb,
                         ipiv, info, dinv, swtab);
    else if (m <= 256)
      hipLaunchKernelGGL((getrf_panel_rows_kernel<T, 1, 256>), dim3(batch), dim3(256), 0, s, n, A, lda, sA, k, nb,
                         ipiv, info, dinv, swtab);
    else if (m <= 512)
      hipLaunchKernelGGL((getrf_panel_rows_kernel<T, 1, 512>), dim3(batch), dim3(512), 0, s, n, A, lda, sA, k, nb,
                         ipiv, info, dinv, swtab);
    else if constexpr (sizeof(T) < 16)
      hipLaunchKernelGGL((getrf_panel_rows_kernel<T, 1, 1024>), dim3(batch), dim3(1024), 0, s, n, A, lda, sA, k, nb,
                         ipiv, info, dinv, swtab);
    else
      return hipErrorInvalidValue;  // not reached: rows_panel_max<c128>() is 512
    return hipGetLastError();
  }
  if ((long)batch * m <= LU_TSLU_MAX_WORK) {
    static const std::string nm_ts = std::string("getrf_panel_tslu<") + tname<T>() + ">";
    prof::Scope ps(s, nm_ts.c_str(), pf, 2.0 * sizeof(T) * (double)m * nb * batch);
    return launch_tslu_panel<T>(n, A, lda, sA, k, nb, ipiv, info, dinv, swtab, cand, batch, s);
  }
  const size_t lds = (size_t)(m + LU_NB) * LU_NB * sizeof(T);
  static const std::string nm_lds = std::string("getrf_panel_lds<") + tname<T>() + ">";
  static const std::string nm_reg = std::string("getrf_panel_reg<") + tname<T>() + ">";
  prof::Scope ps(s, (lds <= LU_PANEL_LDS_MAX ? nm_lds : nm_reg).c_str(), pf,
                 2.0 * sizeof(T) * (double)m * nb * batch);
  if (lds <= LU_PANEL_LDS_MAX) {
    static bool attr_set = false;
    if (!attr_set) {
      hipFuncSetAttribute((const void*)getrf_panel_lds_kernel<T, 256>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)LU_PANEL_LDS_MAX);
      attr_set = true;
    }
    hipLaunchKernelGGL((getrf_panel_lds_kernel<T, 256>), dim3(batch), dim3(256), lds, s, n, A, lda, sA, k, nb, ipiv,
                       info, dinv, swtab);
  } else {
    const int rpt = (m + 1023) / 1024;
    if (rpt <= 1)
      launch_reg_panel<T, 1>(n, A, lda, sA, k, nb, ipiv, info, dinv, swtab, batch, s);
    else if (rpt <= 2)
      launch_reg_panel<T, 2>(n, A, lda, sA, k, nb, ipiv, info, dinv, swtab, batch, s);
    else if (rpt <= 4)
      launch_reg_panel<T, 4>(n, A, lda, sA, k, nb, ipiv, info, dinv, swtab, batch, s);
    else if (rpt <= 8)
      launch_reg_panel<T, 8>(n, A, lda, sA, k, nb, ipiv, info, dinv, swtab, batch, s);
    else if (rpt <= 12)
      launch_reg_panel<T, 12>(n, A, lda, sA, k, nb, ipiv, info, dinv, swtab, batch, s);
    else
      return hipErrorInvalidValue;  // n > 12288: beyond this library's matrix sizes
  }
  return hipGetLastError();
}

// tournament-pivoting candidate lists: two buffers of batch x chunks x NB row indices

#if HPS_LU_TYPES & 1
size_t getrf_scratch_bytes(int n, int batch) {
  // ipiv, swap tables, tournament candidates (two sets), then one arrival
  // counter per matrix for the unpivoted panel kernel
  return ((size_t)batch * n + (size_t)batch * SWTAB + 2 * tslu_cand_ints(n, batch) + (size_t)batch) * sizeof(int);
}
#endif

// X[k:k+nb, :] = tri(Dinv_k) X[k:k+nb, :] in place; tri = unit lower (forward)
// or upper (backward).  With CT the conjugate transpose of the other stored
// inverse is used: lower = inv(U_kk)^H (the U^H solve), upper = inv(L_kk)^H
// (the L^H solve).  256 threads = 32 rows x 8 columns per workgroup.
template <typename T, bool UPPER, bool CT = false>
__global__ __launch_bounds__(256) void trsm_inv_kernel(int n, int nrhs, const T* dinv, T* X, long ldx, long sX, int k,
                                                       int nb) {
  __shared__ T Di[LU_NB][LU_NB + 1];
  __shared__ T Xs[LU_NB][8];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int cc = tid & 7, rr = tid >> 3;
  const int c = blockIdx.x * 8 + cc;
  const int nblk = (n + LU_NB - 1) / LU_NB;
  const T* db = dinv + ((long)b * nblk + k / LU_NB) * LU_NB * LU_NB;
  for (int idx = tid; idx < LU_NB * LU_NB; idx += 256) {
    int r = idx / LU_NB, q = idx - r * LU_NB;
    T v;
    if constexpr (CT) {
      const T t = db[q * LU_NB + r];
      const T ct = make_conj(t);
      if (UPPER)
        v = q > r ? ct : (q == r ? s_one<T>() : s_zero<T>());
      else
        v = q <= r ? ct : s_zero<T>();
    } else {
      if (UPPER)
        v = q >= r ? db[idx] : s_zero<T>();
      else
        v = q < r ? db[idx] : (q == r ? s_one<T>() : s_zero<T>());
    }
    Di[r][q] = v;
  }
  T* Xb = X + (long)b * sX;
  Xs[rr][cc] = (c < nrhs && rr < nb) ? Xb[(long)(k + rr) * ldx + c] : s_zero<T>();
  __syncthreads();
  if (c < nrhs && rr < nb) {
    T acc = s_zero<T>();
    for (int q = 0; q < nb; ++q) acc = s_add(acc, s_mul(Di[rr][q], Xs[q][cc]));
    Xb[(long)(k + rr) * ldx + c] = acc;
  }
}

// X[k:k+nb, :ncols] = tri(Dinv_k) X[k:k+nb, :ncols] in place, tri = unit lower
// or upper: an MFMA product (the K <= 32 GEMM reading the packed inverse as a
// triangle) for wide X, the LDS kernel above for ncols <= 16.
template <typename T>
static hipError_t trsm_block(bool upper, int n, int ncols, const T* dinv, T* X, long ldx, long sX, int k, int nb,
                             int batch, hipStream_t s, const std::string& label) {
  if (ncols > 16) {
    GemmArgs<T> g;
    g.M = nb;
    g.N = ncols;
    g.K = nb;
    g.A = dinv + (long)(k / LU_NB) * LU_NB * LU_NB;
    g.lda = LU_NB;
    g.sA = (long)ceil_div(n, LU_NB) * LU_NB * LU_NB;
    g.B = X + (long)k * ldx;
    g.ldb = ldx;
    g.sB = sX;
    g.D = X + (long)k * ldx;
    g.ldd = ldx;
    g.sD = sX;
    g.batch = batch;
    g.triA = upper ? 2 : 1;
    return gemm<T>(g, s);
  }
  prof::Scope ps(s, label.c_str(), flop_mult<T>() * batch * (double)nb * nb * ncols,
                 2.0 * sizeof(T) * (double)nb * ncols * batch);
  dim3 tgrid(ceil_div(ncols, 8), batch);
  if (upper)
    hipLaunchKernelGGL((trsm_inv_kernel<T, true>), tgrid, dim3(256), 0, s, n, ncols, dinv, X, ldx, sX, k, nb);
  else
    hipLaunchKernelGGL((trsm_inv_kernel<T, false>), tgrid, dim3(256), 0, s, n, ncols, dinv, X, ldx, sX, k, nb);
  return hipGetLastError();
}

template <typename T>
hipError_t getrf_ws(int n, T* A, long lda, long sA, int* perm, T* dinv, int* scratch, int* info, int batch,
                    hipStream_t s, bool nopiv, bool nopiv_all) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  int* ipiv = scratch;
  int* swtab = scratch + (long)batch * n;
  int* cand = swtab + (long)batch * SWTAB;
  int* arrivals = cand + 2 * tslu_cand_ints(n, batch);  // zero between unpivoted panel launches
  hipLaunchKernelGGL(fill_int_kernel, dim3(ceil_div(batch, 256)), dim3(256), 0, s, info, (long)batch, 0);
  if (nopiv) hipLaunchKernelGGL(fill_int_kernel, dim3(ceil_div(batch, 256)), dim3(256), 0, s, arrivals, (long)batch, 0);
  auto gemm_update = [&](int M, int N, int K, const T* Ap, const T* Bp, T* Cp) -> hipError_t {
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    GemmArgs<T> g;
    g.M = M;
    g.N = N;
    g.K = K;
    g.alpha = -1.0;
    g.beta = 1.0;
    g.A = Ap;
    g.lda = lda;
    g.sA = sA;
    g.B = Bp;
    g.ldb = lda;
    g.sB = sA;
    g.C = Cp;
    g.ldc = lda;
    g.sC = sA;
    g.D = Cp;
    g.ldd = lda;
    g.sD = sA;
    g.batch = batch;
    return gemm<T>(g, s);
  };
  // Large matrices are factored in column blocks of LU_NBO: inside a block the
  // 32-wide panels update only the block's columns; the block row's U12 and the
  // trailing update then run once per block, the latter as a GEMM with K = LU_NBO.
  const int nbo = lu_outer_block_wide(n);
  for (int k0 = 0; k0 < n; k0 += nbo) {
    const int bend = n - k0 < nbo ? n : k0 + nbo;
    for (int k = k0; k < bend; k += LU_NB) {
      const int nb = bend - k < LU_NB ? bend - k : LU_NB;
      bool swapped = false, u12_done = false;
      hipError_t e = hipSuccess;
      if (nopiv && (n - k > rows_panel_max<T>() || batch <= LU_NOPIV_SMALL_BATCH)) {
        static const std::string nm_np = std::string("getrf_panel_nopiv<") + tname<T>() + ">";
        const int m = n - k;
        prof::Scope ps(s, nm_np.c_str(), flop_mult<T>() * batch * ((double)m * nb * nb - (double)nb * nb * nb / 3.0),
                       2.0 * sizeof(T) * (double)m * nb * batch);
        // row workgroups, then one workgroup per 32 columns of the block's U12
        const int row_wgs = ceil_div(m - nb, NOPIV_ROWS) > 0 ? ceil_div(m - nb, NOPIV_ROWS) : 1;
        const int u12_wgs = bend - k - nb > 0 ? ceil_div(bend - k - nb, LU_NB) : 0;
        for (int b0 = 0; b0 < batch; b0 += 65535) {
          const int bc = batch - b0 < 65535 ? batch - b0 : 65535;
          hipLaunchKernelGGL((getrf_panel_nopiv_kernel<T>), dim3(row_wgs + u12_wgs, bc), dim3(256), 0, s, n,
                             A + (long)b0 * sA, lda, sA, k, nb, ipiv + (long)b0 * n,
                             dinv + (long)b0 * ((n + LU_NB - 1) / LU_NB) * LU_NB * LU_NB, arrivals + b0, row_wgs,
                             bend);
        }
        e = hipGetLastError();
        swapped = true;  // nothing to swap
        u12_done = true;
      } else {
        e = launch_panel<T>(n, A, lda, sA, k, nb, ipiv, info, dinv, swtab, cand, batch, s, &swapped, nopiv_all);
      }
      if (e != hipSuccess) return e;
      const int others = n - nb;
      if (swapped) {
        // U12 = inv(L11) A12 for the block's columns right of the panel (the
        // unpivoted panel kernel forms it itself)
        if (bend - k - nb > 0 && !u12_done) {
          static const std::string nm = std::string("getrf_u12<") + tname<T>() + ">";
          e = trsm_block<T>(false, n, bend - k - nb, dinv, A + k + nb, lda, sA, k, nb, batch, s, nm);
          if (e != hipSuccess) return e;
        }
      } else if (others > 0) {
        static const std::string nm = std::string("getrf_swap_u12<") + tname<T>() + ">";
        prof::Scope ps(s, nm.c_str(), flop_mult<T>() * batch * (double)(bend - k - nb) * nb * nb,
                       3.0 * sizeof(T) * (double)nb * others * batch);
        dim3 grid(ceil_div(others, 8), batch);
        hipLaunchKernelGGL((getrf_swap_u12_kernel<T>), grid, dim3(256), 0, s, n, A, lda, sA, k, nb, dinv, swtab,
                           bend);
      }
      e = gemm_update(n - k - nb, bend - k - nb, nb, A + (long)(k + nb) * lda + k, A + (long)k * lda + k + nb,
                      A + (long)(k + nb) * lda + k + nb);
      if (e != hipSuccess) return e;
    }
    if (bend < n) {
      // U12 = inv(L11) A12 for the block row, block by block with the inverted diagonal blocks
      const int ncols = n - bend;
      for (int k = k0; k < bend; k += LU_NB) {
        const int nb = bend - k < LU_NB ? bend - k : LU_NB;
        static const std::string nm = std::string("getrf_u12<") + tname<T>() + ">";
        hipError_t e = trsm_block<T>(false, n, ncols, dinv, A + bend, lda, sA, k, nb, batch, s, nm);
        if (e != hipSuccess) return e;
        e = gemm_update(bend - k - nb, ncols, nb, A + (long)(k + nb) * lda + k, A + (long)k * lda + bend,
                                   A + (long)(k + nb) * lda + bend);
        if (e != hipSuccess) return e;
      }
      hipError_t e = gemm_update(n - bend, ncols, bend - k0, A + (long)bend * lda + k0, A + (long)k0 * lda + bend,
                                 A + (long)bend * lda + bend);
      if (e != hipSuccess) return e;
    }
  }
  hipLaunchKernelGGL(perm_from_ipiv_kernel, dim3(batch), dim3(256), (size_t)n * sizeof(int), s, n, ipiv, perm);
  return hipGetLastError();
}

// X[b][i][c] = R[b][perm[b][i]][c]
template <typename T>
__global__ void gather_rows_kernel(int n, int nrhs, const int* perm, const T* R, long ldr, long sR, T* X, long ldx,
                                   long sX) {
  const int b = blockIdx.z;
  const int i = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  int src = perm[(long)b * n + i];
  X[(long)b * sX + (long)i * ldx + c] = R[(long)b * sR + (long)src * ldr + c];
}

// Small systems with few right-hand sides (the leaf solves of the up pass,
// refinement and adjoint): one workgroup per matrix keeps X (n x nrhs) in LDS
// and streams the factors once, 32-row block by block: the diagonal block
// through its stored inverse, the rows below (above) by dot products over the
// block's 32 columns, each row segment read once by four lanes for all the
// right-hand sides (offblock_rows).  CT solves with A^H instead (U^H forward,
// L^H backward, then X = P^T Z), whose column reads are coalesced across
// threads already.
constexpr int LU_SMALL_MAX_N = 1024;
constexpr int LU_SMALL_MAX_RHS = 16;


// xs[i][c] -= sum_q Ab[i * lda + q] xs[kx + q][c] for rows r0 <= i < r1 and
// the nb <= 32 block columns q, by a 256-thread workgroup: four lanes per
// row, eight block columns each, so a row's block segment is read
// contiguously by its lanes and their products summed across them (a
// row-per-thread loop strides lda between lanes).  The rows go in chunks of
// 64; each chunk's loads are issued before the previous chunk's products, and
// the first chunk comes in `a`, loaded by offblock_load before the caller's
// own work of the step, so no step waits on a cold global load.  The block
// rows of xs are only read.  Every thread of the workgroup must call it.
template <typename T>
__device__ __forceinline__ void offblock_load(const T* Ab, long lda, int i, int r1, int nb, int sub, T (&a)[8]) {
  const bool live = i < r1;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int q = 8 * sub + j;
    a[j] = live && q < nb ? Ab[(long)i * lda + q] : s_zero<T>();
  }
}

template <typename T>
__device__ __forceinline__ void offblock_rows(const T* Ab, long lda, int r0, int r1, int nb, int nrhs, T* xs, int kx,
                                              int tid, T (&a)[8]) {
  const int sub = tid & 3;
  const int row0 = r0 + (tid >> 2);
  const int iters = (r1 - r0 + 63) / 64;  // uniform across the workgroup
  for (int it = 0; it < iters; ++it) {
    const int i = row0 + 64 * it;
    const bool live = i < r1;
    T an[8];
    offblock_load(Ab, lda, i + 64, it + 1 < iters ? r1 : 0, nb, sub, an);
    for (int c = 0; c < nrhs; ++c) {
      T prod = s_zero<T>();
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int q = 8 * sub + j;
        if (q < nb) prod = s_add(prod, s_mul(a[j], xs[(kx + q) * nrhs + c]));
      }
      prod = quad_sum(prod);  // DPP: every lane of the quad runs this
      if (live && sub == 0) xs[i * nrhs + c] = s_sub(xs[i * nrhs + c], prod);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = an[j];
  }
}

// The 32 x 32 triangular factor the solve applies to a diagonal block, from
// the packed inverse dk: unit-lower inv(L_kk) (LOWER) or inv(U_kk), or for
// CT their conjugate transposes (inv(U_kk)^H lower, inv(L_kk)^H unit upper).
// PER = 1024 / (threads of the workgroup) entries per thread, so the next
// step's block can be loaded into registers while the current one is applied.
template <typename T, bool CT, int PER>
__device__ __forceinline__ void diag_inverse_load(const T* dk, bool lower, int tid, T (&d)[PER]) {
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + (LU_NB * LU_NB / PER) * e;
    const int r = idx / LU_NB, q = idx - r * LU_NB;
    T v;
    if (lower) {
      if (CT)
        v = q <= r ? make_conj(dk[q * LU_NB + r]) : s_zero<T>();
      else
        v = q < r ? dk[idx] : (q == r ? s_one<T>() : s_zero<T>());
    } else {
      if (CT)
        v = q > r ? make_conj(dk[q * LU_NB + r]) : (q == r ? s_one<T>() : s_zero<T>());
      else
        v = q >= r ? dk[idx] : s_zero<T>();
    }
    d[e] = v;
  }
}

// xs[k + r][c] = sum_q Di[r][q] xs[k + q][c] for the nb rows of a diagonal
// block: one quad per output (r, c), lane s summing q = s, s + 4, ... with
// the loads of all eight terms issued together, then a DPP quad sum.  Every
// thread of the workgroup calls it; it ends with xs updated and synchronised.
template <typename T, int THREADS = 256>
__device__ __forceinline__ void diag_block_apply(const T (*Di)[LU_NB + 1], T* xs, int k, int nb, int nrhs, int tid) {
  constexpr int QUADS = THREADS / 4;
  constexpr int MAXP = LU_NB * LU_SMALL_MAX_RHS / QUADS;
  const int sub = tid & 3, quad = tid >> 2;
  const int nout = nb * nrhs;
  const int per = (nout + QUADS - 1) / QUADS;  // <= MAXP for nb <= 32, nrhs <= 16
  T y[MAXP];
#pragma unroll
  for (int e = 0; e < MAXP; ++e) {
    y[e] = s_zero<T>();
    if (e < per) {
      const int o = quad + QUADS * e;
      const bool live = o < nout;
      const int r = live ? o / nrhs : 0, c = live ? o - r * nrhs : 0;
      T acc = s_zero<T>();
#pragma unroll
      for (int t = 0; t < LU_NB / 4; ++t) {
        const int q = 4 * t + sub;
        if (live && q < nb) acc = s_fma(Di[r][q], xs[(k + q) * nrhs + c], acc);
      }
      y[e] = quad_sum(acc);
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < MAXP; ++e) {
    const int o = quad + QUADS * e;
    if (e < per && o < nout && sub == 0) xs[(k + o / nrhs) * nrhs + o % nrhs] = y[e];
  }
  __syncthreads();
}

// ONE: a single right-hand side (nrhs == 1 at compile time: the refinement's
// leaf solves), which needs fewer registers, so more workgroups per CU
// stream their factors at once.
template <typename T, bool CT, bool ONE = false>
__global__ __launch_bounds__(256) void getrs_small_kernel(int n, int nrhs_arg, const T* LU, long lda, long sA,
                                                          const int* perm, const T* dinv, const T* R, long ldr,
                                                          long sR, T* X, long ldx, long sX) {
  const int nrhs = ONE ? 1 : nrhs_arg;
  extern __shared__ char smem_raw[];
  T* xs = (T*)smem_raw;  // (n, nrhs)
  __shared__ T Di[LU_NB][LU_NB + 1];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const T* A = LU + (long)b * sA;
  const int* pb = perm + (long)b * n;
  const int nblk = (n + LU_NB - 1) / LU_NB;
  const T* db = dinv + (long)b * nblk * LU_NB * LU_NB;
  const T* Rb = R + (long)b * sR;
  // steps t < nblk: lower (L, or U^H for CT), forward; then upper (U, or
  // L^H), backward.  Step t + 1's diagonal inverse and first chunk of rows
  // are loaded into registers at the start of step t.
  const int sub = tid & 3;
  auto geom = [&](int t, int& k, int& nb, int& r0, int& r1) {
    const bool lower = t < nblk;
    const int kb = lower ? t : 2 * nblk - 1 - t;
    k = kb * LU_NB;
    nb = n - k < LU_NB ? n - k : LU_NB;
    r0 = lower ? k + nb : 0;  // rows outside the block: below (lower) or above
    r1 = lower ? n : k;
  };
  T d[4], a[8];
  {
    int k, nb, r0, r1;
    geom(0, k, nb, r0, r1);
    diag_inverse_load<T, CT, 4>(db, true, tid, d);
    if (!CT) offblock_load(A + k, lda, r0 + (tid >> 2), r1, nb, sub, a);
  }
  // (the first step's factor loads above go out before these)
  for (int idx = tid; idx < n * nrhs; idx += 256) {
    const int i = idx / nrhs, c = idx - i * nrhs;
    xs[idx] = CT ? Rb[(long)i * ldr + c] : Rb[(long)pb[i] * ldr + c];
  }
  __syncthreads();
  for (int t = 0; t < 2 * nblk; ++t) {
    int k, nb, r0, r1;
    geom(t, k, nb, r0, r1);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      Di[idx / LU_NB][idx % LU_NB] = d[e];
    }
    T a_next[8];
    if (t + 1 < 2 * nblk) {
      int k1, nb1, r01, r11;
      geom(t + 1, k1, nb1, r01, r11);
      diag_inverse_load<T, CT, 4>(db + (long)(k1 / LU_NB) * LU_NB * LU_NB, t + 1 < nblk, tid, d);
      if (!CT) offblock_load(A + k1, lda, r01 + (tid >> 2), r11, nb1, sub, a_next);
    }
    __syncthreads();
    diag_block_apply<T>(Di, xs, k, nb, nrhs, tid);  // diagonal block: x_k = Di x_k
    if constexpr (!CT) {
      offblock_rows(A + k, lda, r0, r1, nb, nrhs, xs, k, tid, a);
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = a_next[j];
    } else {
      for (int i = r0 + tid; i < r1; i += 256) {
        for (int c = 0; c < nrhs; ++c) {
          T acc = xs[i * nrhs + c];
          for (int q = 0; q < nb; ++q) {
            // conj(U[k+q][i]) (lower step) / conj(L[k+q][i]) (upper step)
            const T av = make_conj(A[(long)(k + q) * lda + i]);
            acc = s_fnms(acc, av, xs[(k + q) * nrhs + c]);
          }
          xs[i * nrhs + c] = acc;
        }
      }
    }
    __syncthreads();
  }
  T* Xb = X + (long)b * sX;
  for (int idx = tid; idx < n * nrhs; idx += 256) {
    const int i = idx / nrhs, c = idx - i * nrhs;
    if (CT)
      Xb[(long)pb[i] * ldx + c] = xs[idx];
    else
      Xb[(long)i * ldx + c] = xs[idx];
  }
}

// In-place triangular solve of the diagonal block [k0, k0 + nbk <= 256) of X
// for few right-hand sides (the inner steps of the blocked solve of large
// systems): unit-lower L forward or U backward, 32-row sub-blocks through the
// stored inverses; the X block lives in LDS.  One 1024-thread workgroup per
// matrix: the rows of the block below (above) a sub-block, at most 224, take
// one quad each, so a sub-block's factor entries are all in flight at once,
// and the next sub-block's entries and inverse are loaded while the current
// one is applied (one global latency per sub-block at most, not one per
// 64-row chunk).
constexpr int GETRS_BLOCK_THREADS = 1024;
constexpr int GETRS_BLOCK_MAX = 256;

template <typename T, bool UPPER>
__global__ __launch_bounds__(GETRS_BLOCK_THREADS) void getrs_block_kernel(int n, int nrhs, const T* LU, long lda,
                                                                          long sA, const T* dinv, T* X, long ldx,
                                                                          long sX, int k0, int nbk) {
  extern __shared__ char smem_raw[];
  T* xs = (T*)smem_raw;  // (nbk, nrhs)
  __shared__ T Di[LU_NB][LU_NB + 1];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const T* A = LU + (long)b * sA;
  const int nblk = (n + LU_NB - 1) / LU_NB;
  const T* db = dinv + (long)b * nblk * LU_NB * LU_NB;
  T* Xb = X + (long)b * sX;
  const int nsub = (nbk + LU_NB - 1) / LU_NB;
  const int sub = tid & 3;
  const T* Ak = A + (long)k0 * lda + k0;
  auto geom = [&](int sb, int& kl, int& nb, int& r0, int& r1) {
    kl = (UPPER ? nsub - 1 - sb : sb) * LU_NB;  // sub-block start within the block
    nb = nbk - kl < LU_NB ? nbk - kl : LU_NB;
    r0 = UPPER ? 0 : kl + nb;
    r1 = UPPER ? kl : nbk;
  };
  // the first sub-block's factor loads go out before the X block's, so the
  // two global latencies overlap
  T d[1], a[8];
  {
    int kl, nb, r0, r1;
    geom(0, kl, nb, r0, r1);
    diag_inverse_load<T, false, 1>(db + (long)((k0 + kl) / LU_NB) * LU_NB * LU_NB, !UPPER, tid, d);
    offblock_load(Ak + kl, lda, r0 + (tid >> 2), r1, nb, sub, a);
  }
  for (int idx = tid; idx < nbk * nrhs; idx += GETRS_BLOCK_THREADS) {
    const int i = idx / nrhs, c = idx - i * nrhs;
    xs[idx] = Xb[(long)(k0 + i) * ldx + c];
  }
  for (int sb = 0; sb < nsub; ++sb) {
    int kl, nb, r0, r1;
    geom(sb, kl, nb, r0, r1);
    Di[tid / LU_NB][tid % LU_NB] = d[0];
    T a_next[8];
    if (sb + 1 < nsub) {
      int kl1, nb1, r01, r11;
      geom(sb + 1, kl1, nb1, r01, r11);
      diag_inverse_load<T, false, 1>(db + (long)((k0 + kl1) / LU_NB) * LU_NB * LU_NB, !UPPER, tid, d);
      offblock_load(Ak + kl1, lda, r01 + (tid >> 2), r11, nb1, sub, a_next);
    }
    __syncthreads();
    diag_block_apply<T, GETRS_BLOCK_THREADS>(Di, xs, kl, nb, nrhs, tid);
    // the rows outside the sub-block: one quad per row, r1 - r0 <= 224
    const int i = r0 + (tid >> 2);
    const bool live = i < r1;
    for (int c = 0; c < nrhs; ++c) {
      T prod = s_zero<T>();
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int q = 8 * sub + j;
        if (q < nb) prod = s_add(prod, s_mul(a[j], xs[(kl + q) * nrhs + c]));
      }
      prod = quad_sum(prod);  // DPP: every lane of the quad runs this
      if (live && sub == 0) xs[i * nrhs + c] = s_sub(xs[i * nrhs + c], prod);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = a_next[j];
    __syncthreads();
  }
  for (int idx = tid; idx < nbk * nrhs; idx += GETRS_BLOCK_THREADS) {
    const int i = idx / nrhs, c = idx - i * nrhs;
    Xb[(long)(k0 + i) * ldx + c] = xs[idx];
  }
}

template <typename T, bool UPPER>
static hipError_t getrs_block(int n, int nrhs, const T* LU, long lda, long sA, const T* dinv, T* X, long ldx, long sX,
                              int k0, int nbk, int batch, hipStream_t s) {
  // the kernel's quads cover at most GETRS_BLOCK_MAX rows, for few columns
  if (nrhs > LU_SMALL_MAX_RHS || nbk > GETRS_BLOCK_MAX) return hipErrorInvalidValue;
  static const std::string nm = std::string(UPPER ? "getrs_block_upper<" : "getrs_block_lower<") + tname<T>() + ">";
  prof::Scope ps(s, nm.c_str(), flop_mult<T>() * batch * (double)nbk * nbk * nrhs,
                 sizeof(T) * (double)batch * ((double)nbk * nbk / 2 + 2.0 * nbk * nrhs));
  const int nblk = (n + LU_NB - 1) / LU_NB;
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    const int bc = batch - b0 < 65535 ? batch - b0 : 65535;
    hipLaunchKernelGGL((getrs_block_kernel<T, UPPER>), dim3(bc), dim3(GETRS_BLOCK_THREADS),
                       (size_t)nbk * nrhs * sizeof(T), s, n,
                       nrhs, LU + (long)b0 * sA, lda, sA, dinv + (long)b0 * nblk * LU_NB * LU_NB, X + (long)b0 * sX,
                       ldx, sX, k0, nbk);
  }
  return hipGetLastError();
}

// getrs_small streams a whole matrix through one workgroup: for large
// batches that keeps every CU busy; a few large matrices take the blocked
// solve instead, whose GEMM updates spread each matrix over the GPU.
constexpr int GETRS_SMALL_MIN_BATCH = 128;

template <typename T>
static bool getrs_small_ok(int n, int nrhs, int batch) {
  return n <= LU_SMALL_MAX_N && nrhs <= LU_SMALL_MAX_RHS && (size_t)n * nrhs * sizeof(T) <= 64 * 1024 &&
         (batch >= GETRS_SMALL_MIN_BATCH || n <= 256);
}

template <typename T, bool CT>
static hipError_t getrs_small(int n, int nrhs, const T* LU, long lda, long sA, const int* perm, const T* dinv,
                              const T* R, long ldr, long sR, T* X, long ldx, long sX, int batch, hipStream_t s) {
  static const std::string nm = std::string(CT ? "getrs_small_conjt<" : "getrs_small<") + tname<T>() + ">";
  prof::Scope ps(s, nm.c_str(), flop_mult<T>() * 2.0 * batch * (double)n * n * nrhs,
                 sizeof(T) * (double)batch * ((double)n * n + 2.0 * n * nrhs));
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    const int bc = batch - b0 < 65535 ? batch - b0 : 65535;
    auto kern = nrhs == 1 ? getrs_small_kernel<T, CT, true> : getrs_small_kernel<T, CT, false>;
    hipLaunchKernelGGL(kern, dim3(bc), dim3(256), (size_t)n * nrhs * sizeof(T), s, n, nrhs,
                       LU + (long)b0 * sA, lda, sA, perm + (long)b0 * n,
                       dinv + (long)b0 * ((n + LU_NB - 1) / LU_NB) * LU_NB * LU_NB, R + (long)b0 * sR, ldr, sR,
                       X + (long)b0 * sX, ldx, sX);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t getrs_perm(int n, int nrhs, const T* LU, long lda, long sA, const int* perm, const T* dinv, const T* R,
                      long ldr, long sR, T* X, long ldx, long sX, int batch, hipStream_t s) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return hipSuccess;
  if (getrs_small_ok<T>(n, nrhs, batch))
    return getrs_small<T, false>(n, nrhs, LU, lda, sA, perm, dinv, R, ldr, sR, X, ldx, sX, batch, s);
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    static const std::string nm = std::string("getrs_gather<") + tname<T>() + ">";
    prof::Scope ps(s, nm.c_str(), 0.0, 2.0 * sizeof(T) * (double)n * nrhs * batch);
    int bc = batch - b0 < 65535 ? batch - b0 : 65535;
    dim3 grid(ceil_div(nrhs, 64), n, bc);
    hipLaunchKernelGGL((gather_rows_kernel<T>), grid, dim3(64), 0, s, n, nrhs, perm + (long)b0 * n,
                       R + (long)b0 * sR, ldr, sR, X + (long)b0 * sX, ldx, sX);
  }
  return getrs_inplace<T>(n, nrhs, LU, lda, sA, dinv, X, ldx, sX, batch, s);
}

template <typename T>
hipError_t getrs_inplace(int n, int nrhs, const T* LU, long lda, long sA, const T* dinv, T* X, long ldx, long sX,
                         int batch, hipStream_t s) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return hipSuccess;
  auto gemm_update = [&](int M, int K, const T* Ap, const T* Bp, T* Cp) -> hipError_t {
    if (M <= 0 || K <= 0) return hipSuccess;
    GemmArgs<T> g;
    g.M = M;
    g.N = nrhs;
    g.K = K;
    g.alpha = -1.0;
    g.beta = 1.0;
    g.A = Ap;
    g.lda = lda;
    g.sA = sA;
    g.B = Bp;
    g.ldb = ldx;
    g.sB = sX;
    g.C = Cp;
    g.ldc = ldx;
    g.sC = sX;
    g.D = Cp;
    g.ldd = ldx;
    g.sD = sX;
    g.batch = batch;
    return gemm<T>(g, s);
  };
  // Column blocks of LU_NBO for large matrices (as in getrf): the 32-row
  // steps update only the rows of their block; the rest of X is updated once
  // per block by a GEMM with K = LU_NBO.
  // few right-hand sides: each outer block's inner steps in one kernel
  const int nbo_few = lu_outer_block(n);
  const bool few = nrhs <= LU_SMALL_MAX_RHS && (size_t)nbo_few * nrhs * sizeof(T) <= 64 * 1024;
  const int nbo = few ? nbo_few : lu_outer_block_wide(n);
  for (int k0 = 0; k0 < n; k0 += nbo) {
    const int bend = n - k0 < nbo ? n : k0 + nbo;
    if (few) {
      hipError_t e = getrs_block<T, false>(n, nrhs, LU, lda, sA, dinv, X, ldx, sX, k0, bend - k0, batch, s);
      if (e != hipSuccess) return e;
    }
    for (int k = k0; k < bend && !few; k += LU_NB) {
      const int nb = bend - k < LU_NB ? bend - k : LU_NB;
      static const std::string nm = std::string("trsm_lower<") + tname<T>() + ">";
      hipError_t e = trsm_block<T>(false, n, nrhs, dinv, X, ldx, sX, k, nb, batch, s, nm);
      if (e != hipSuccess) return e;
      e = gemm_update(bend - k - nb, nb, LU + (long)(k + nb) * lda + k, X + (long)k * ldx,
                                 X + (long)(k + nb) * ldx);
      if (e != hipSuccess) return e;
    }
    hipError_t e = gemm_update(n - bend, bend - k0, LU + (long)bend * lda + k0, X + (long)k0 * ldx,
                               X + (long)bend * ldx);
    if (e != hipSuccess) return e;
  }
  const int nblo = (n + nbo - 1) / nbo;
  for (int bo = nblo - 1; bo >= 0; --bo) {
    const int k0 = bo * nbo;
    const int bend = n - k0 < nbo ? n : k0 + nbo;
    const int last = ((bend - 1 - k0) / LU_NB) * LU_NB + k0;
    if (few) {
      hipError_t e = getrs_block<T, true>(n, nrhs, LU, lda, sA, dinv, X, ldx, sX, k0, bend - k0, batch, s);
      if (e != hipSuccess) return e;
    }
    for (int k = last; k >= k0 && !few; k -= LU_NB) {
      const int nb = bend - k < LU_NB ? bend - k : LU_NB;
      static const std::string nm = std::string("trsm_upper<") + tname<T>() + ">";
      hipError_t e = trsm_block<T>(true, n, nrhs, dinv, X, ldx, sX, k, nb, batch, s, nm);
      if (e != hipSuccess) return e;
      e = gemm_update(k - k0, nb, LU + (long)k0 * lda + k, X + (long)k * ldx, X + (long)k0 * ldx);
      if (e != hipSuccess) return e;
    }
    hipError_t e = gemm_update(k0, bend - k0, LU + k0, X + (long)k0 * ldx, X);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Many-column solves with large factors in low precision (the S solves of the
// top merge levels, refined in FP64 afterwards): the inverses of the 512 x 512
// diagonal blocks of L (unit lower) and U are formed once per solve (all
// blocks of a matrix as one batched 512-column triangular solve of the
// identity), after which every block step is a product with K = 512 instead
// of sixteen 32-row steps whose updates stream the block's columns through
// HBM.  The products read the explicit block inverses, so the rounding is
// that of the inverses (cond(L_kk), cond(U_kk): pivot growth ~1 for the merge
// matrices), corrected by the refinement.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void eye_kernel(T* X, int n, long sX, int batch) {
  const long per = (long)n * n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < per * batch;
       idx += (long)gridDim.x * blockDim.x) {
    const long b = idx / per, e = idx - b * per;
    const int r = (int)(e / n), c = (int)(e - (long)r * n);
    X[b * sX + e] = r == c ? s_one<T>() : s_zero<T>();
  }
}

// Linv_j = inv(L_jj) (unit lower) and Uinv_j = inv(U_jj) of `count` NB x NB
// diagonal blocks at once: block j's factor entries at Lb + j sDiag (row
// stride lda), its packed 32 x 32 diagonal inverses at db + j sDinv; the
// inverses go to Linv / Uinv + j NB^2 (row stride NB).  Each 32-row step
// applies the packed inverse (a triangular product) and updates the rows
// below (L) or above (U) with K = 32.
template <typename T>
static hipError_t form_diag_inverses(int NB, int count, const T* Lb, long lda, long sDiag, const T* db, long sDinv,
                                     T* Linv, T* Uinv, hipStream_t s) {
  const long sBlk = (long)NB * NB;
  auto gemm = [&](int M, int N, int K, double alpha, double beta, const T* A, long la, long sa, const T* B, long lb,
                  long sb, const T* C, long lc, long sc, T* D, long ld, long sd, int bt, int tri) -> hipError_t {
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    GemmArgs<T> g;
    g.M = M;
    g.N = N;
    g.K = K;
    g.alpha = alpha;
    g.beta = beta;
    g.A = A;
    g.lda = la;
    g.sA = sa;
    g.B = B;
    g.ldb = lb;
    g.sB = sb;
    g.C = C;
    g.ldc = lc;
    g.sC = sc;
    g.D = D;
    g.ldd = ld;
    g.sD = sd;
    g.batch = bt;
    g.triA = tri;
    return hps::gemm<T>(g, s);
  };
  static const std::string nm_inv = std::string("blockinv_form<") + tname<T>() + ">";
  {
    prof::Scope ps(s, nm_inv.c_str(), 0.0, 2.0 * sizeof(T) * (double)count * sBlk);
    hipLaunchKernelGGL((eye_kernel<T>), dim3(1024), dim3(256), 0, s, Linv, NB, sBlk, count);
    hipLaunchKernelGGL((eye_kernel<T>), dim3(1024), dim3(256), 0, s, Uinv, NB, sBlk, count);
  }
  HPS_LU_CHECK(hipGetLastError());
  for (int k = 0; k < NB; k += LU_NB) {
    // rows k .. k + 32 of Linv: inv(L_kk) times them, then the rows below
    HPS_LU_CHECK(gemm(LU_NB, NB, LU_NB, 1.0, 0.0, db + (long)(k / LU_NB) * LU_NB * LU_NB, LU_NB, sDinv,
                      Linv + (long)k * NB, NB, sBlk, nullptr, 0, 0, Linv + (long)k * NB, NB, sBlk, count, 1));
    HPS_LU_CHECK(gemm(NB - k - LU_NB, NB, LU_NB, -1.0, 1.0, Lb + (long)(k + LU_NB) * lda + k, lda, sDiag,
                      Linv + (long)k * NB, NB, sBlk, Linv + (long)(k + LU_NB) * NB, NB, sBlk,
                      Linv + (long)(k + LU_NB) * NB, NB, sBlk, count, 0));
  }
  for (int k = NB - LU_NB; k >= 0; k -= LU_NB) {
    HPS_LU_CHECK(gemm(LU_NB, NB, LU_NB, 1.0, 0.0, db + (long)(k / LU_NB) * LU_NB * LU_NB, LU_NB, sDinv,
                      Uinv + (long)k * NB, NB, sBlk, nullptr, 0, 0, Uinv + (long)k * NB, NB, sBlk, count, 2));
    HPS_LU_CHECK(gemm(k, NB, LU_NB, -1.0, 1.0, Lb + k, lda, sDiag, Uinv + (long)k * NB, NB, sBlk, Uinv, NB, sBlk,
                      Uinv, NB, sBlk, count, 0));
  }
  return hipGetLastError();
}

template <typename T>
hipError_t getrs_perm_blockinv(int n, int nrhs, const T* LU, long lda, long sA, const int* perm, const T* dinv,
                               const T* R, long ldr, long sR, T* X, long ldx, long sX, int batch, T* wsp,
                               hipStream_t s) {
  if (!blockinv_eligible(n, nrhs) || wsp == nullptr)
    return getrs_perm<T>(n, nrhs, LU, lda, sA, perm, dinv, R, ldr, sR, X, ldx, sX, batch, s);
  const int NB = BLOCKINV_NB;
  const int nblk = n / NB;
  const long sBlk = (long)NB * NB;  // one block inverse
  T* Linv = wsp;
  T* Uinv = wsp + (long)nblk * sBlk;
  T* tmp = Uinv + (long)nblk * sBlk;  // NB x nrhs
  const long dstride = (long)((n + LU_NB - 1) / LU_NB) * LU_NB * LU_NB;  // dinv per matrix
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    static const std::string nm = std::string("getrs_gather<") + tname<T>() + ">";
    prof::Scope ps(s, nm.c_str(), 0.0, 2.0 * sizeof(T) * (double)n * nrhs * batch);
    int bc = batch - b0 < 65535 ? batch - b0 : 65535;
    dim3 grid(ceil_div(nrhs, 64), n, bc);
    hipLaunchKernelGGL((gather_rows_kernel<T>), grid, dim3(64), 0, s, n, nrhs, perm + (long)b0 * n,
                       R + (long)b0 * sR, ldr, sR, X + (long)b0 * sX, ldx, sX);
  }
  auto gemm = [&](int M, int N, int K, double alpha, double beta, const T* A, long la, long sa, const T* B, long lb,
                  long sb, const T* C, long lc, long sc, T* D, long ld, long sd, int bt, int tri) -> hipError_t {
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    GemmArgs<T> g;
    g.M = M;
    g.N = N;
    g.K = K;
    g.alpha = alpha;
    g.beta = beta;
    g.A = A;
    g.lda = la;
    g.sA = sa;
    g.B = B;
    g.ldb = lb;
    g.sB = sb;
    g.C = C;
    g.ldc = lc;
    g.sC = sc;
    g.D = D;
    g.ldd = ld;
    g.sD = sd;
    g.batch = bt;
    g.triA = tri;
    return hps::gemm<T>(g, s);
  };
  for (int b = 0; b < batch; ++b) {
    const T* Lb = LU + (long)b * sA;
    const T* db = dinv + (long)b * dstride;
    T* Xb = X + (long)b * sX;
    // block inverses: Linv_j = inv(L_jj), Uinv_j = inv(U_jj), all j at once
    HPS_LU_CHECK(form_diag_inverses<T>(NB, nblk, Lb, lda, (long)NB * (lda + 1), db,
                                       (long)(NB / LU_NB) * LU_NB * LU_NB, Linv, Uinv, s));
    // forward: X_j = Linv_j X_j, then the rows below
    for (int j = 0; j < nblk; ++j) {
      const int k0 = j * NB;
      HPS_LU_CHECK(gemm(NB, nrhs, NB, 1.0, 0.0, Linv + (long)j * sBlk, NB, 0, Xb + (long)k0 * ldx, ldx, 0, nullptr,
                        0, 0, tmp, nrhs, 0, 1, 0));
      HPS_LU_CHECK(copy2d<T>(NB, nrhs, 1.0, tmp, nrhs, 0, Xb + (long)k0 * ldx, ldx, 0, 1, s));
      HPS_LU_CHECK(gemm(n - k0 - NB, nrhs, NB, -1.0, 1.0, Lb + (long)(k0 + NB) * lda + k0, lda, 0,
                        Xb + (long)k0 * ldx, ldx, 0, Xb + (long)(k0 + NB) * ldx, ldx, 0,
                        Xb + (long)(k0 + NB) * ldx, ldx, 0, 1, 0));
    }
    // backward: X_j = Uinv_j X_j, then the rows above
    for (int j = nblk - 1; j >= 0; --j) {
      const int k0 = j * NB;
      HPS_LU_CHECK(gemm(NB, nrhs, NB, 1.0, 0.0, Uinv + (long)j * sBlk, NB, 0, Xb + (long)k0 * ldx, ldx, 0, nullptr,
                        0, 0, tmp, nrhs, 0, 1, 0));
      HPS_LU_CHECK(copy2d<T>(NB, nrhs, 1.0, tmp, nrhs, 0, Xb + (long)k0 * ldx, ldx, 0, 1, s));
      HPS_LU_CHECK(gemm(k0, nrhs, NB, -1.0, 1.0, Lb + k0, lda, 0, Xb + (long)k0 * ldx, ldx, 0, Xb, ldx, 0, Xb, ldx,
                        0, 1, 0));
    }
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Few-column solves of the top merge levels' low-precision factors (the
// refinement's up passes) through explicit inverses of the FEWINV_NB x
// FEWINV_NB diagonal blocks of L and U, formed once per build: each block
// step is one product with the block inverse instead of eight dependent
// 32-row steps in one workgroup.
// ---------------------------------------------------------------------------

// The diagonal blocks of every matrix, copied contiguously (block bb = b nblk
// + j): G[bb] = the NB x NB factor entries (zero past n), Gd[bb] = its packed
// 32 x 32 inverses (identity past n), so all blocks form their inverses in one
// batched pass and the last, partial block acts as [[A_r, 0], [0, I]].
template <typename T>
__global__ void diag_blocks_gather_kernel(int n, int nblk, const T* LU, long lda, long sA, const T* dinv, T* G,
                                          T* Gd) {
  constexpr int NB = FEWINV_NB;
  const int bb = blockIdx.y;
  const int b = bb / nblk, j = bb - b * nblk;
  const int k0 = j * NB;
  const T* A = LU + (long)b * sA;
  const long dstride = (long)((n + LU_NB - 1) / LU_NB) * LU_NB * LU_NB;
  T* g = G + (long)bb * NB * NB;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < NB * NB; idx += gridDim.x * blockDim.x) {
    const int r = idx / NB, c = idx - r * NB;
    const int gr = k0 + r, gc = k0 + c;
    g[idx] = gr < n && gc < n ? A[(long)gr * lda + gc] : s_zero<T>();
  }
  constexpr int ND = (NB / LU_NB) * LU_NB * LU_NB;
  T* gd = Gd + (long)bb * ND;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < ND; idx += gridDim.x * blockDim.x) {
    const int sb = idx / (LU_NB * LU_NB), e = idx - sb * LU_NB * LU_NB;
    const int r = e / LU_NB, c = e - r * LU_NB;
    gd[idx] = k0 + sb * LU_NB < n ? dinv[(long)b * dstride + (long)(k0 / LU_NB + sb) * LU_NB * LU_NB + e]
                                  : (r == c ? s_one<T>() : s_zero<T>());
  }
}

template <typename T>
hipError_t form_fewinv(int n, const T* LU, long lda, long sA, const T* dinv, T* inv, T* ws, int batch,
                       hipStream_t s) {
  if (!fewinv_eligible(n) || batch <= 0) return hipErrorInvalidValue;
  constexpr int NB = FEWINV_NB;
  const int nblk = fewinv_blocks(n);
  const int count = batch * nblk;
  if (count > 65535) return hipErrorInvalidValue;
  T* G = ws;
  T* Gd = ws + (long)count * NB * NB;
  {
    static const std::string nm = std::string("blockinv_form<") + tname<T>() + ">";
    prof::Scope ps(s, nm.c_str(), 0.0, 2.0 * sizeof(T) * (double)count * NB * NB);
    hipLaunchKernelGGL((diag_blocks_gather_kernel<T>), dim3(64, count), dim3(256), 0, s, n, nblk, LU, lda, sA, dinv,
                       G, Gd);
    HPS_LU_CHECK(hipGetLastError());
  }
  return form_diag_inverses<T>(NB, count, G, NB, (long)NB * NB, Gd, (long)(NB / LU_NB) * LU_NB * LU_NB, inv,
                               inv + (long)count * NB * NB, s);
}

template <typename T>
hipError_t getrs_perm_fewinv(int n, int nrhs, const T* LU, long lda, long sA, const int* perm, const T* inv,
                             const T* R, long ldr, long sR, T* X, long ldx, long sX, T* Y, int batch,
                             hipStream_t s) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return hipSuccess;
  if (!fewinv_eligible(n)) return hipErrorInvalidValue;
  constexpr int NB = FEWINV_NB;
  const int nblk = fewinv_blocks(n);
  const long sI = (long)nblk * NB * NB;  // one matrix's block inverses of L (or U)
  const T* Li = inv;
  const T* Ui = inv + (long)batch * sI;
  const long ldy = nrhs, sY = (long)n * nrhs;
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    static const std::string nm = std::string("getrs_gather<") + tname<T>() + ">";
    prof::Scope ps(s, nm.c_str(), 0.0, 2.0 * sizeof(T) * (double)n * nrhs * batch);
    int bc = batch - b0 < 65535 ? batch - b0 : 65535;
    dim3 grid(ceil_div(nrhs, 64), n, bc);
    hipLaunchKernelGGL((gather_rows_kernel<T>), grid, dim3(64), 0, s, n, nrhs, perm + (long)b0 * n,
                       R + (long)b0 * sR, ldr, sR, X + (long)b0 * sX, ldx, sX);
  }
  auto gemm = [&](int M, int K, double beta, const T* A, long la, long sa, const T* B, long lb, long sb, T* D, long ld,
                  long sd) -> hipError_t {
    if (M <= 0 || K <= 0) return hipSuccess;
    GemmArgs<T> g;
    g.M = M;
    g.N = nrhs;
    g.K = K;
    g.alpha = beta != 0.0 ? -1.0 : 1.0;
    g.beta = beta;
    g.A = A;
    g.lda = la;
    g.sA = sa;
    g.B = B;
    g.ldb = lb;
    g.sB = sb;
    g.C = beta != 0.0 ? D : nullptr;
    g.ldc = ld;
    g.sC = sd;
    g.D = D;
    g.ldd = ld;
    g.sD = sd;
    g.batch = batch;
    return hps::gemm<T>(g, s);
  };
  // forward (L): Y_j = inv(L_jj) X_j, then the rows below: X_>j -= L[>j, j] Y_j
  // (every update spans all remaining rows, so its products fill the GPU)
  for (int j = 0; j < nblk; ++j) {
    const int k0 = j * NB, nb = n - k0 < NB ? n - k0 : NB, kend = k0 + nb;
    HPS_LU_CHECK(gemm(nb, nb, 0.0, Li + (long)j * NB * NB, NB, sI, X + (long)k0 * ldx, ldx, sX, Y + (long)k0 * ldy,
                      ldy, sY));
    HPS_LU_CHECK(gemm(n - kend, nb, 1.0, LU + (long)kend * lda + k0, lda, sA, Y + (long)k0 * ldy, ldy, sY,
                      X + (long)kend * ldx, ldx, sX));
  }
  // backward (U): X_j = inv(U_jj) Y_j, then the rows above: Y_<j -= U[<j, j] X_j
  for (int j = nblk - 1; j >= 0; --j) {
    const int k0 = j * NB, nb = n - k0 < NB ? n - k0 : NB;
    HPS_LU_CHECK(gemm(nb, nb, 0.0, Ui + (long)j * NB * NB, NB, sI, Y + (long)k0 * ldy, ldy, sY, X + (long)k0 * ldx,
                      ldx, sX));
    HPS_LU_CHECK(gemm(k0, nb, 1.0, LU + k0, lda, sA, X + (long)k0 * ldx, ldx, sX, Y, ldy, sY));
  }
  return hipGetLastError();
}

// X[b][perm[b][i]][c] = W[b][i][c]  (the inverse of gather_rows_kernel)
template <typename T>
__global__ void scatter_rows_kernel(int n, int nrhs, const int* perm, const T* W, long ldw, long sW, T* X, long ldx,
                                    long sX) {
  const int b = blockIdx.z;
  const int i = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  int dst = perm[(long)b * n + i];
  X[(long)b * sX + (long)dst * ldx + c] = W[(long)b * sW + (long)i * ldw + c];
}

// A^H X = R with P A = L U:  X = P^T L^{-H} U^{-H} R.  U^H is lower
// triangular (forward, with inv(U_kk)^H), L^H unit upper (backward, with
// inv(L_kk)^H); the off-diagonal updates are conjugate-transposed GEMMs on
// the stored factors, blocked like getrs_perm.
template <typename T>
hipError_t getrs_conjt(int n, int nrhs, const T* LU, long lda, long sA, const int* perm, const T* dinv, const T* R,
                       long ldr, long sR, T* X, long ldx, long sX, T* W, int batch, hipStream_t s) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return hipSuccess;
  // the blocked conjugate-transposed solve has no block kernel: its 32-row
  // steps cost more launches than getrs_small costs latency at any batch
  if (getrs_small_ok<T>(n, nrhs, GETRS_SMALL_MIN_BATCH))
    return getrs_small<T, true>(n, nrhs, LU, lda, sA, perm, dinv, R, ldr, sR, X, ldx, sX, batch, s);
  const long ldw = nrhs, sW = (long)n * nrhs;
  HPS_LU_CHECK(copy2d<T>(n, nrhs, 1.0, R, ldr, sR, W, ldw, sW, batch, s));
  dim3 tgrid(ceil_div(nrhs, 8), batch);
  auto gemm_ct = [&](int M, int K, const T* Ap, const T* Bp, T* Cp) -> hipError_t {
    if (M <= 0 || K <= 0) return hipSuccess;
    GemmArgs<T> g;
    g.M = M;
    g.N = nrhs;
    g.K = K;
    g.alpha = -1.0;
    g.beta = 1.0;
    g.A = Ap;
    g.lda = lda;
    g.sA = sA;
    g.transA = 1;
    g.B = Bp;
    g.ldb = ldw;
    g.sB = sW;
    g.C = Cp;
    g.ldc = ldw;
    g.sC = sW;
    g.D = Cp;
    g.ldd = ldw;
    g.sD = sW;
    g.batch = batch;
    return gemm<T>(g, s);
  };
  const int nbo = lu_outer_block(n);
  // forward: U^H Y = R
  for (int k0 = 0; k0 < n; k0 += nbo) {
    const int bend = n - k0 < nbo ? n : k0 + nbo;
    for (int k = k0; k < bend; k += LU_NB) {
      const int nb = bend - k < LU_NB ? bend - k : LU_NB;
      {
        static const std::string nm = std::string("trsm_lower_conjt<") + tname<T>() + ">";
        prof::Scope ps(s, nm.c_str(), flop_mult<T>() * batch * (double)nb * nb * nrhs,
                       2.0 * sizeof(T) * (double)nb * nrhs * batch);
        hipLaunchKernelGGL((trsm_inv_kernel<T, false, true>), tgrid, dim3(256), 0, s, n, nrhs, dinv, W, ldw, sW, k,
                           nb);
      }
      HPS_LU_CHECK(gemm_ct(bend - k - nb, nb, LU + (long)k * lda + k + nb, W + (long)k * ldw, W + (long)(k + nb) * ldw));
    }
    HPS_LU_CHECK(gemm_ct(n - bend, bend - k0, LU + (long)k0 * lda + bend, W + (long)k0 * ldw, W + (long)bend * ldw));
  }
  // backward: L^H Z = Y
  const int nblo = (n + nbo - 1) / nbo;
  for (int bo = nblo - 1; bo >= 0; --bo) {
    const int k0 = bo * nbo;
    const int bend = n - k0 < nbo ? n : k0 + nbo;
    const int last = ((bend - 1 - k0) / LU_NB) * LU_NB + k0;
    for (int k = last; k >= k0; k -= LU_NB) {
      const int nb = bend - k < LU_NB ? bend - k : LU_NB;
      {
        static const std::string nm = std::string("trsm_upper_conjt<") + tname<T>() + ">";
        prof::Scope ps(s, nm.c_str(), flop_mult<T>() * batch * (double)nb * nb * nrhs,
                       2.0 * sizeof(T) * (double)nb * nrhs * batch);
        hipLaunchKernelGGL((trsm_inv_kernel<T, true, true>), tgrid, dim3(256), 0, s, n, nrhs, dinv, W, ldw, sW, k,
                           nb);
      }
      HPS_LU_CHECK(gemm_ct(k - k0, nb, LU + (long)k * lda + k0, W + (long)k * ldw, W + (long)k0 * ldw));
    }
    HPS_LU_CHECK(gemm_ct(k0, bend - k0, LU + (long)k0 * lda, W + (long)k0 * ldw, W));
  }
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    static const std::string nm = std::string("getrs_scatter<") + tname<T>() + ">";
    prof::Scope ps(s, nm.c_str(), 0.0, 2.0 * sizeof(T) * (double)n * nrhs * batch);
    int bc = batch - b0 < 65535 ? batch - b0 : 65535;
    dim3 grid(ceil_div(nrhs, 64), n, bc);
    hipLaunchKernelGGL((scatter_rows_kernel<T>), grid, dim3(64), 0, s, n, nrhs, perm + (long)b0 * n,
                       W + (long)b0 * sW, ldw, sW, X + (long)b0 * sX, ldx, sX);
  }
  return hipGetLastError();
}

#if HPS_LU_TYPES & 1
size_t dinv_elems(int n) { return (size_t)((n + LU_NB - 1) / LU_NB) * LU_NB * LU_NB; }
#endif

#define HPS_INST_LU(T)                                                                                          \
  template hipError_t getrf_ws<T>(int, T*, long, long, int*, T*, int*, int*, int, hipStream_t, bool, bool);           \
  template hipError_t getrs_perm<T>(int, int, const T*, long, long, const int*, const T*, const T*, long, long, \
                                    T*, long, long, int, hipStream_t);                                          \
  template hipError_t getrs_inplace<T>(int, int, const T*, long, long, const T*, T*, long, long, int, hipStream_t); \
  template hipError_t getrs_perm_blockinv<T>(int, int, const T*, long, long, const int*, const T*, const T*, long,     \
                                             long, T*, long, long, int, T*, hipStream_t);                             \
  template hipError_t getrs_conjt<T>(int, int, const T*, long, long, const int*, const T*, const T*, long, long, \
                                     T*, long, long, T*, int, hipStream_t);                                     \
  template hipError_t form_fewinv<T>(int, const T*, long, long, const T*, T*, T*, int, hipStream_t);            \
  template hipError_t getrs_perm_fewinv<T>(int, int, const T*, long, long, const int*, const T*, const T*, long,  \
                                           long, T*, long, long, T*, int, hipStream_t);
// HPS_LU_TYPES selects the element types this object instantiates (bit 0
// double, 1 float, 2 c128, 3 c64; default all): the build compiles one
// object per type in parallel.  The type-independent functions live in the
// object that holds double.
#if HPS_LU_TYPES & 1
HPS_INST_LU(double)
#endif
#if HPS_LU_TYPES & 2
HPS_INST_LU(float)
#endif
#if HPS_LU_TYPES & 4
HPS_INST_LU(c128)
#endif
#if HPS_LU_TYPES & 8
HPS_INST_LU(c64)
#endif

}  // namespace hps

// Coarsest-level direct solver: sparse Cholesky L L' = P K P' in 64x64 dense
// tiles, replacing cholesky_factorization.m:26-36 (AMD ordering + CHOLMOD
// lchol) and cholesky_solve.m:27-43 (x(perm) = L'\(L\b(perm))).
//
// Ordering: geometric nested dissection of the coarse nodes (recursive
// coordinate bisection, separator = nodes of one half adjacent to the other),
// all free dofs of a node kept together.  Symbolic factorisation runs at tile
// granularity on the host (tile elimination tree), and the numeric phases run
// as level-scheduled batches of tile tasks on the device:
//   update  C_IJ = A_IJ - sum_K L_IK L_JK'   (one workgroup per target tile)
//   potrf   L_JJ = chol(C_JJ), and inv(L_JJ) (one workgroup per column)
//   trsm    L_IJ = C_IJ inv(L_JJ)'           (one workgroup per tile)
// The triangular solves run one level-batch per launch; the diagonal block
// solve is a product with the stored inv(L_JJ).
#include "m3_internal.h"

#include <algorithm>
#include <numeric>

namespace {

constexpr int TS = 64;
constexpr int TT = TS * TS;

// ---- host: nested dissection ----------------------------------------------------------
struct NDState {
  const std::vector<int64_t> *ptr;
  const std::vector<int32_t> *adj;
  const std::vector<double> *xyz;
  std::vector<int32_t> stamp;
  int cur = 0;
  std::vector<int32_t> order;
};

void nd_rec(NDState &st, std::vector<int32_t> nodes) {
  if (nodes.size() <= 40) {
    st.order.insert(st.order.end(), nodes.begin(), nodes.end());
    return;
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int n : nodes)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], (*st.xyz)[3 * n + k]);
      hi[k] = std::max(hi[k], (*st.xyz)[3 * n + k]);
    }
  int ax = 0;
  for (int k = 1; k < 3; ++k)
    if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
  const size_t mid = nodes.size() / 2;
  std::nth_element(nodes.begin(), nodes.begin() + mid, nodes.end(), [&](int a, int b) {
    const double xa = (*st.xyz)[3 * a + ax], xb = (*st.xyz)[3 * b + ax];
    return xa < xb || (xa == xb && a < b);
  });
  std::vector<int32_t> A(nodes.begin(), nodes.begin() + mid), B(nodes.begin() + mid, nodes.end());
  const int sB = ++st.cur;
  for (int n : B) st.stamp[n] = sB;
  std::vector<int32_t> A2, S;
  for (int n : A) {
    bool touch = false;
    for (int64_t k = (*st.ptr)[n]; k < (*st.ptr)[n + 1] && !touch; ++k)
      if (st.stamp[(*st.adj)[k]] == sB) touch = true;
    (touch ? S : A2).push_back(n);
  }
  ++st.cur;  // invalidate stamps of B for nested calls
  nd_rec(st, A2);
  nd_rec(st, B);
  st.order.insert(st.order.end(), S.begin(), S.end());
}

// ---- device kernels -------------------------------------------------------------------
__global__ void k_gather(int64_t n, const int64_t *src, const int64_t *dst, const double *val,
                         int64_t nslot, double *tiles) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) tiles[dst[i]] = val[src[i]];
}

// C_IJ -= sum_pairs L_IK L_JK'   (tiles column-major)
__global__ __launch_bounds__(256) void k_update(const int32_t *task_tile, const int64_t *pair_ptr,
                                                const int32_t *pairs, int task0, double *tiles) {
  const int task = task0 + blockIdx.x;
  const int64_t p0 = pair_ptr[task], p1 = pair_ptr[task + 1];
  if (p0 == p1) return;
  double *C = tiles + (int64_t)task_tile[task] * TT;
  __shared__ double sA[TS][TS + 1];
  __shared__ double sB[TS][TS + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // output rows 4*ty.., cols 4*tx..
  double acc[4][4] = {};
  for (int64_t p = p0; p < p1; ++p) {
    const double *A = tiles + (int64_t)pairs[2 * p] * TT;      // L_IK
    const double *B = tiles + (int64_t)pairs[2 * p + 1] * TT;  // L_JK
    __syncthreads();
    for (int t = threadIdx.x; t < TT; t += 256) {
      const int r = t & 63, c = t >> 6;  // column-major element (r,c)
      sA[r][c] = A[t];
      sB[r][c] = B[t];
    }
    __syncthreads();
    for (int k = 0; k < TS; ++k) {
      double a[4], b[4];
      for (int i = 0; i < 4; ++i) a[i] = sA[4 * ty + i][k];
      for (int j = 0; j < 4; ++j) b[j] = sB[4 * tx + j][k];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const int r = 4 * ty + i, c = 4 * tx + j;
      C[r + TS * c] -= acc[i][j];
    }
}

// Cholesky of the diagonal tile (in LDS) and inverse of the factor.
__global__ __launch_bounds__(256) void k_potrf(const int32_t *cols, int col0, const int32_t *diagtile,
                                               const int32_t *pos2dof, double *tiles, double *dinv,
                                               double *dinvT, int *err) {
  const int J = cols[col0 + blockIdx.x];
  double *A = tiles + (int64_t)diagtile[J] * TT;
  __shared__ double L[TS][TS + 1];
  __shared__ double X[TS][TS + 1];
  for (int t = threadIdx.x; t < TT; t += 256) {
    const int r = t & 63, c = t >> 6;
    double v = A[t];
    const bool pr = pos2dof[J * TS + r] < 0, pc = pos2dof[J * TS + c] < 0;
    if (pr || pc) v = (r == c) ? 1.0 : 0.0;
    L[r][c] = v;
  }
  __syncthreads();
  for (int k = 0; k < TS; ++k) {
    if (threadIdx.x == 0) {
      const double d = L[k][k];
      if (!(d > 0)) atomicOr(err, 16);
      L[k][k] = sqrt(d > 0 ? d : 1.0);
    }
    __syncthreads();
    const double dk = L[k][k];
    for (int r = k + 1 + threadIdx.x; r < TS; r += 256) L[r][k] /= dk;
    __syncthreads();
    for (int t = threadIdx.x; t < TT; t += 256) {
      const int r = t & 63, c = t >> 6;
      if (c > k && r >= c) L[r][c] -= L[r][k] * L[c][k];
    }
    __syncthreads();
  }
  // X = inv(L): column j by forward substitution (thread j < 64)
  if (threadIdx.x < TS) {
    const int j = threadIdx.x;
    for (int i = 0; i < TS; ++i) {
      double s = (i == j) ? 1.0 : 0.0;
      for (int k = j; k < i; ++k) s -= L[i][k] * X[k][j];
      X[i][j] = (i >= j) ? s / L[i][i] : 0.0;
    }
  }
  __syncthreads();
  double *Di = dinv + (int64_t)J * TT;
  double *DiT = dinvT + (int64_t)J * TT;
  for (int t = threadIdx.x; t < TT; t += 256) {
    const int r = t & 63, c = t >> 6;
    A[t] = (r >= c) ? L[r][c] : 0.0;
    Di[t] = X[r][c];
    DiT[t] = X[c][r];
  }
}

// L_IJ = C_IJ inv(L_JJ)'  and its transposed copy
__global__ __launch_bounds__(256) void k_trsm(const int32_t *tile, const int32_t *col, int task0,
                                              double *tiles, double *tilesT, const double *dinv) {
  const int task = task0 + blockIdx.x;
  double *C = tiles + (int64_t)tile[task] * TT;
  const double *D = dinv + (int64_t)col[task] * TT;
  __shared__ double sC[TS][TS + 1];
  __shared__ double sD[TS][TS + 1];
  for (int t = threadIdx.x; t < TT; t += 256) {
    const int r = t & 63, c = t >> 6;
    sC[r][c] = C[t];
    sD[r][c] = D[t];
  }
  __syncthreads();
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4] = {};
  for (int k = 0; k < TS; ++k) {
    double a[4], b[4];
    for (int i = 0; i < 4; ++i) a[i] = sC[4 * ty + i][k];
    for (int j = 0; j < 4; ++j) b[j] = sD[4 * tx + j][k];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
  }
  double *CT = tilesT + (int64_t)tile[task] * TT;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const int r = 4 * ty + i, c = 4 * tx + j;
      C[r + TS * c] = acc[i][j];
      CT[c + TS * r] = acc[i][j];
    }
}

// forward: y_J = inv(L_JJ) (b_J - sum_K L_JK y_K)
__global__ __launch_bounds__(64) void k_fwd(const int32_t *cols, int col0, const int32_t *rptr,
                                            const int32_t *rtile, const int32_t *rcol,
                                            const double *tiles, const double *dinv, double *w) {
  const int J = cols[col0 + blockIdx.x];
  const int i = threadIdx.x;
  __shared__ double sy[TS];
  double acc = w[(int64_t)J * TS + i];
  for (int k = rptr[J]; k < rptr[J + 1]; ++k) {
    const double *Lt = tiles + (int64_t)rtile[k] * TT;
    const double *y = w + (int64_t)rcol[k] * TS;
    __syncthreads();
    sy[i] = y[i];
    __syncthreads();
    double s = 0;
    for (int c = 0; c < TS; ++c) s += Lt[i + TS * c] * sy[c];
    acc -= s;
  }
  __syncthreads();
  sy[i] = acc;
  __syncthreads();
  const double *D = dinv + (int64_t)J * TT;
  double s = 0;
  for (int c = 0; c < TS; ++c) s += D[i + TS * c] * sy[c];
  w[(int64_t)J * TS + i] = s;
}

// backward: x_J = inv(L_JJ)' (y_J - sum_I L_IJ' x_I)
__global__ __launch_bounds__(64) void k_bwd(const int32_t *cols, int col0, const int32_t *cptr,
                                            const int32_t *ctile, const int32_t *crow,
                                            const double *tilesT, const double *dinvT, double *w) {
  const int J = cols[col0 + blockIdx.x];
  const int j = threadIdx.x;
  __shared__ double sx[TS];
  double acc = w[(int64_t)J * TS + j];
  for (int k = cptr[J]; k < cptr[J + 1]; ++k) {
    const double *LT = tilesT + (int64_t)ctile[k] * TT;  // L_IJ' column-major
    const double *x = w + (int64_t)crow[k] * TS;
    __syncthreads();
    sx[j] = x[j];
    __syncthreads();
    double s = 0;
    for (int i = 0; i < TS; ++i) s += LT[j + TS * i] * sx[i];
    acc -= s;
  }
  __syncthreads();
  sx[j] = acc;
  __syncthreads();
  const double *DT = dinvT + (int64_t)J * TT;
  double s = 0;
  for (int i = 0; i < TS; ++i) s += DT[j + TS * i] * sx[i];
  w[(int64_t)J * TS + j] = s;
}

__global__ void k_to_tiles(int64_t npos, const int32_t *pos2dof, const double *x, double *w) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= npos) return;
  const int d = pos2dof[i];
  w[i] = d >= 0 ? x[d] : 0.0;
}

__global__ void k_from_tiles(int64_t npos, const int32_t *pos2dof, const double *w, double *x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= npos) return;
  const int d = pos2dof[i];
  if (d >= 0) x[d] = w[i];
}

inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

namespace m3 {

// Symbolic analysis: ordering, tile structure, schedules.  Needs the coarsest
// level pattern, its mask and the node coordinates.
void chol_analyse(m3_ctx *c) {
  LevelDev &cl = c->L[c->nmg - 1];
  CholTiles &ch = c->chol;
  const int64_t nn = cl.nnod;
  const std::vector<double> mask = cl.mask.to_host();
  // node adjacency (CSR) from the sliced pattern
  std::vector<int64_t> ptr(nn + 1, 0);
  std::vector<int32_t> adj;
  const Sell &K = cl.K;
  for (int64_t r = 0; r < nn; ++r) {
    const int64_t base = K.h_slice_ptr[r >> 6] + (r & 63);
    for (int k = 0; k < K.h_rowlen[r]; ++k) adj.push_back(K.h_col[base + (int64_t)k * 64]);
    ptr[r + 1] = (int64_t)adj.size();
  }
  NDState st;
  st.ptr = &ptr;
  st.adj = &adj;
  st.xyz = &c->h_coord;
  st.stamp.assign(nn, 0);
  std::vector<int32_t> nodes;
  for (int64_t n = 0; n < nn; ++n)
    if (mask[3 * n] + mask[3 * n + 1] + mask[3 * n + 2] > 0) nodes.push_back((int32_t)n);
  nd_rec(st, nodes);
  std::vector<int32_t> perm;  // position -> dof
  for (int n : st.order)
    for (int k = 0; k < 3; ++k)
      if (mask[3 * n + k] != 0.0) perm.push_back(3 * n + k);
  ch.n = (int64_t)perm.size();
  ch.T = (int)((ch.n + TS - 1) / TS);
  const int T = ch.T;
  std::vector<int32_t> inv(3 * nn, -1);
  for (size_t i = 0; i < perm.size(); ++i) inv[perm[i]] = (int32_t)i;
  std::vector<int32_t> pos2dof((size_t)T * TS, -1);
  for (size_t i = 0; i < perm.size(); ++i) pos2dof[i] = perm[i];
  // tile structure of A (lower)
  std::vector<std::vector<int32_t>> Acol(T);
  for (int64_t n = 0; n < nn; ++n)
    for (int64_t k = ptr[n]; k < ptr[n + 1]; ++k) {
      const int m = adj[k];
      for (int a = 0; a < 3; ++a) {
        const int pa = inv[3 * n + a];
        if (pa < 0) continue;
        for (int b = 0; b < 3; ++b) {
          const int pb = inv[3 * (int64_t)m + b];
          if (pb < 0 || pa < pb) continue;
          Acol[pb / TS].push_back(pa / TS);
        }
      }
    }
  for (int J = 0; J < T; ++J) {
    Acol[J].push_back(J);
    std::sort(Acol[J].begin(), Acol[J].end());
    Acol[J].erase(std::unique(Acol[J].begin(), Acol[J].end()), Acol[J].end());
  }
  // tile-level symbolic factorisation
  std::vector<std::vector<int32_t>> S(T);
  std::vector<std::vector<int32_t>> children(T);
  for (int J = 0; J < T; ++J) {
    std::vector<int32_t> s = Acol[J];
    for (int K : children[J])
      for (int I : S[K])
        if (I > K && I >= J) s.push_back(I);
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    S[J] = s;
    if (s.size() > 1) children[s[1]].push_back(J);
  }
  ch.h_colptr.assign(T + 1, 0);
  for (int J = 0; J < T; ++J) ch.h_colptr[J + 1] = ch.h_colptr[J] + (int32_t)S[J].size();
  ch.ntile = ch.h_colptr[T];
  ch.h_rowtile.clear();
  for (int J = 0; J < T; ++J) ch.h_rowtile.insert(ch.h_rowtile.end(), S[J].begin(), S[J].end());
  auto tile_of = [&](int I, int J) -> int32_t {
    const auto b = ch.h_rowtile.begin() + ch.h_colptr[J], e = ch.h_rowtile.begin() + ch.h_colptr[J + 1];
    const auto it = std::lower_bound(b, e, I);
    return (it != e && *it == I) ? (int32_t)(it - ch.h_rowtile.begin()) : -1;
  };
  // row lists: for row I, (K, t_IK) with K < I
  std::vector<std::vector<std::pair<int32_t, int32_t>>> rows(T);
  for (int K = 0; K < T; ++K)
    for (int32_t t = ch.h_colptr[K] + 1; t < ch.h_colptr[K + 1]; ++t)
      rows[ch.h_rowtile[t]].push_back({K, t});
  // levels
  std::vector<int> lev(T, 0), blev(T, 0);
  int nlev = 0;
  for (int J = 0; J < T; ++J) {
    int l = 0;
    for (auto &kt : rows[J]) l = std::max(l, lev[kt.first] + 1);
    lev[J] = l;
    nlev = std::max(nlev, l + 1);
  }
  int nblev = 0;
  for (int J = T - 1; J >= 0; --J) {
    int l = 0;
    for (int32_t t = ch.h_colptr[J] + 1; t < ch.h_colptr[J + 1]; ++t) l = std::max(l, blev[ch.h_rowtile[t]] + 1);
    blev[J] = l;
    nblev = std::max(nblev, l + 1);
  }
  ch.nlev = nlev;
  ch.nflev = nlev;
  ch.nblev = nblev;
  std::vector<int32_t> lev_cols, blev_cols;
  ch.h_lev_ptr.assign(nlev + 1, 0);
  ch.h_blev_ptr.assign(nblev + 1, 0);
  for (int l = 0; l < nlev; ++l) {
    for (int J = 0; J < T; ++J)
      if (lev[J] == l) lev_cols.push_back(J);
    ch.h_lev_ptr[l + 1] = (int32_t)lev_cols.size();
  }
  for (int l = 0; l < nblev; ++l) {
    for (int J = 0; J < T; ++J)
      if (blev[J] == l) blev_cols.push_back(J);
    ch.h_blev_ptr[l + 1] = (int32_t)blev_cols.size();
  }
  ch.h_flev_ptr = ch.h_lev_ptr;
  // update and trsm tasks by level
  std::vector<int32_t> upd_tile, trsm_tile, trsm_col, pairs;
  std::vector<int64_t> pair_ptr{0};
  ch.h_upd_ptr.assign(nlev + 1, 0);
  ch.h_trsm_ptr.assign(nlev + 1, 0);
  for (int l = 0; l < nlev; ++l) {
    for (int32_t q = ch.h_lev_ptr[l]; q < ch.h_lev_ptr[l + 1]; ++q) {
      const int J = lev_cols[q];
      for (int32_t t = ch.h_colptr[J]; t < ch.h_colptr[J + 1]; ++t) {
        const int I = ch.h_rowtile[t];
        upd_tile.push_back(t);
        for (auto &kt : rows[J]) {
          const int32_t tIK = tile_of(I, kt.first);
          if (tIK >= 0) {
            pairs.push_back(tIK);
            pairs.push_back(kt.second);
          }
        }
        pair_ptr.push_back((int64_t)pairs.size() / 2);
        if (I != J) {
          trsm_tile.push_back(t);
          trsm_col.push_back(J);
        }
      }
    }
    ch.h_upd_ptr[l + 1] = (int32_t)upd_tile.size();
    ch.h_trsm_ptr[l + 1] = (int32_t)trsm_tile.size();
  }
  // solve lists
  std::vector<int32_t> frow_ptr{0}, frow_tile, frow_col, bcol_ptr{0}, bcol_tile, bcol_row, diagt(T);
  for (int J = 0; J < T; ++J) {
    for (auto &kt : rows[J]) {
      frow_tile.push_back(kt.second);
      frow_col.push_back(kt.first);
    }
    frow_ptr.push_back((int32_t)frow_tile.size());
    for (int32_t t = ch.h_colptr[J] + 1; t < ch.h_colptr[J + 1]; ++t) {
      bcol_tile.push_back(t);
      bcol_row.push_back(ch.h_rowtile[t]);
    }
    bcol_ptr.push_back((int32_t)bcol_tile.size());
    diagt[J] = ch.h_colptr[J];
  }
  // gather map from the coarse operator values (SoA, 9*nslot)
  std::vector<int64_t> gsrc, gdst;
  for (int64_t n = 0; n < nn; ++n) {
    const int64_t base = K.h_slice_ptr[n >> 6] + (n & 63);
    for (int k = 0; k < K.h_rowlen[n]; ++k) {
      const int64_t slot = base + (int64_t)k * 64;
      const int m = K.h_col[slot];
      for (int a = 0; a < 3; ++a) {
        const int pa = inv[3 * n + a];
        if (pa < 0) continue;
        for (int b = 0; b < 3; ++b) {
          const int pb = inv[3 * (int64_t)m + b];
          if (pb < 0 || pa < pb) continue;
          const int32_t t = tile_of(pa / TS, pb / TS);
          gsrc.push_back((int64_t)(3 * a + b) * K.nslot + slot);
          gdst.push_back((int64_t)t * TT + (pa % TS) + (int64_t)TS * (pb % TS));
        }
      }
    }
  }
  ch.pos2dof.upload(pos2dof);
  ch.lev_cols.upload(lev_cols);
  ch.blev_cols.upload(blev_cols);
  ch.flev_cols.upload(lev_cols);
  ch.upd_tile.upload(upd_tile);
  ch.upd_pair_ptr.upload(pair_ptr);
  ch.upd_pairs.upload(pairs.empty() ? std::vector<int32_t>{0, 0} : pairs);
  ch.trsm_tile.upload(trsm_tile.empty() ? std::vector<int32_t>{0} : trsm_tile);
  ch.trsm_col.upload(trsm_col.empty() ? std::vector<int32_t>{0} : trsm_col);
  ch.frow_ptr.upload(frow_ptr);
  ch.frow_tile.upload(frow_tile.empty() ? std::vector<int32_t>{0} : frow_tile);
  ch.frow_col.upload(frow_col.empty() ? std::vector<int32_t>{0} : frow_col);
  ch.bcol_ptr.upload(bcol_ptr);
  ch.bcol_tile.upload(bcol_tile.empty() ? std::vector<int32_t>{0} : bcol_tile);
  ch.bcol_row.upload(bcol_row.empty() ? std::vector<int32_t>{0} : bcol_row);
  ch.diagtile.upload(diagt);
  ch.gsrc.upload(gsrc);
  ch.gdst.upload(gdst);
  ch.tiles.alloc((size_t)ch.ntile * TT);
  ch.tilesT.alloc((size_t)ch.ntile * TT);
  ch.dinv.alloc((size_t)T * TT);
  ch.dinvT.alloc((size_t)T * TT);
  ch.work.alloc((size_t)T * TS);
  c->factored = false;
}

void chol_factor(m3_ctx *c) {
  CholTiles &ch = c->chol;
  LevelDev &cl = c->L[c->nmg - 1];
  hipStream_t s = c->stream;
  ch.tiles.zero(s);
  // dinvT lives in the upper half of dinv's allocation pair: use tilesT for
  // off-diagonal transposes and a second buffer for inverse transposes
  static_assert(TS == 64, "tile size");
  if (ch.work.n < (size_t)ch.T * TS) ch.work.alloc((size_t)ch.T * TS);
  const int64_t ng = (int64_t)ch.gsrc.n;
  hipLaunchKernelGGL(k_gather, dim3(nblk(ng)), dim3(256), 0, s, ng, ch.gsrc.p, ch.gdst.p, cl.K.val.p,
                     cl.K.nslot, ch.tiles.p);
  for (int l = 0; l < ch.nlev; ++l) {
    const int u0 = ch.h_upd_ptr[l], u1 = ch.h_upd_ptr[l + 1];
    if (u1 > u0)
      hipLaunchKernelGGL(k_update, dim3(u1 - u0), dim3(256), 0, s, ch.upd_tile.p, ch.upd_pair_ptr.p,
                         ch.upd_pairs.p, u0, ch.tiles.p);
    const int c0 = ch.h_lev_ptr[l], c1 = ch.h_lev_ptr[l + 1];
    hipLaunchKernelGGL(k_potrf, dim3(c1 - c0), dim3(256), 0, s, ch.lev_cols.p, c0, ch.diagtile.p,
                       ch.pos2dof.p, ch.tiles.p, ch.dinv.p, ch.dinvT.p, c->errflag.p);
    const int t0 = ch.h_trsm_ptr[l], t1 = ch.h_trsm_ptr[l + 1];
    if (t1 > t0)
      hipLaunchKernelGGL(k_trsm, dim3(t1 - t0), dim3(256), 0, s, ch.trsm_tile.p, ch.trsm_col.p, t0,
                         ch.tiles.p, ch.tilesT.p, ch.dinv.p);
  }
  HIPCHECK(hipGetLastError());
  c->factored = true;
}

void chol_solve_dev(m3_ctx *c, const double *rhs, double *x) {
  CholTiles &ch = c->chol;
  LevelDev &cl = c->L[c->nmg - 1];
  hipStream_t s = c->stream;
  const int64_t npos = (int64_t)ch.T * TS;
  hipLaunchKernelGGL(k_to_tiles, dim3(nblk(npos)), dim3(256), 0, s, npos, ch.pos2dof.p, rhs, ch.work.p);
  for (int l = 0; l < ch.nflev; ++l) {
    const int c0 = ch.h_lev_ptr[l], c1 = ch.h_lev_ptr[l + 1];
    hipLaunchKernelGGL(k_fwd, dim3(c1 - c0), dim3(64), 0, s, ch.flev_cols.p, c0, ch.frow_ptr.p,
                       ch.frow_tile.p, ch.frow_col.p, ch.tiles.p, ch.dinv.p, ch.work.p);
  }
  for (int l = 0; l < ch.nblev; ++l) {
    const int c0 = ch.h_blev_ptr[l], c1 = ch.h_blev_ptr[l + 1];
    hipLaunchKernelGGL(k_bwd, dim3(c1 - c0), dim3(64), 0, s, ch.blev_cols.p, c0, ch.bcol_ptr.p,
                       ch.bcol_tile.p, ch.bcol_row.p, ch.tilesT.p, ch.dinvT.p, ch.work.p);
  }
  HIPCHECK(hipMemsetAsync(x, 0, sizeof(double) * 3 * cl.nnod, s));
  hipLaunchKernelGGL(k_from_tiles, dim3(nblk(npos)), dim3(256), 0, s, npos, ch.pos2dof.p, ch.work.p, x);
  HIPCHECK(hipGetLastError());
}

}  // namespace m3

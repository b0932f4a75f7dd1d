// dcgrid_solve.hpp — gfx950 kernels of the DCGrid projection solve (extension; the reference
// only runs a fixed number of Jacobi sweeps, dcgrid_multigrid_solver.cu:5-41).
//
// The system is the fixed point of those sweeps: on every cell c of level l
//     R_c(x) = (sum of the 6 apron values of x) - 6 x_c - alpha_l div_c = 0,   alpha_l = (2^l dx)^2,
// with apron entries on coarser levels as ghosts (the coarser levels are solved first). Each
// level is solved by iterative refinement: the residual, its norms and the update x += e are
// FP64 (or FP32 / FP64 throughout), the correction e of A e = R comes from one multigrid
// V-cycle in FP32 whose grids are the level itself, the coarser levels restricted to the
// ancestors of the level's cells ("reach"), and below the top level a uniform pyramid of the
// top level (dense by construction). Smoothing is red-black Gauss-Seidel; restriction is the
// mean of the 8 cells of a restriction group, prolongation the trilinear 27/9/3/1 stencil.
// Residuals on coarse grids are kept in density units (R / h^2), so restriction is a plain
// mean and each grid uses its own cell size in the operator.
#pragma once

#include "dcgrid_fluid.hpp"

namespace dcg {

// Face neighbour (dir 0..5: x-, x+, y-, y+, z-, z+) of local cell t of block b, as the apron
// table holds it; `coarser` tells whether it lies on a coarser level.
template <int BW>
__device__ __forceinline__ uint32_t faceNb(const DGBase& g, uint32_t b, int t, int dir, bool& coarser) {
  constexpr int BB = Geo<BW>::BB, BV = Geo<BW>::BV;
  int x, y, z;
  locPos<BB>((uint32_t)t, x, y, z);
  int ai = x + 1, aj = y + 1, ak = z + 1;
  switch (dir) {
    case 0: ai--; break;
    case 1: ai++; break;
    case 2: aj--; break;
    case 3: aj++; break;
    case 4: ak--; break;
    default: ak++; break;
  }
  coarser = false;
  if (ai >= 1 && ai <= BW && aj >= 1 && aj <= BW && ak >= 1 && ak <= BW) return b * BV + loc<BB>(ai - 1, aj - 1, ak - 1);
  const int dx = ai == 0 ? -1 : (ai == BW + 1 ? 1 : 0);
  const int dy = aj == 0 ? -1 : (aj == BW + 1 ? 1 : 0);
  const int dz = ak == 0 ? -1 : (ak == BW + 1 ? 1 : 0);
  const uint32_t sh = ndShift(g.nd[b * 27 + dirIndex(dx, dy, dz)]);
  coarser = sh != 0 && sh != ND_OUT;
  return apronCell<BW>(g, b, ai, aj, ak);
}

// Parent cell (level + 1) of cell t of block b: the restriction-group mapping of accumulate<T>.
template <int BW>
__device__ __forceinline__ uint32_t parentCellOf(const DGBase& g, uint32_t b, int t) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB;
  const uint32_t ps = g.par[b];
  return (ps / 8) * BV + ((ps % 8) << SBB) + (uint32_t)(t >> 3);
}

// reach[c] = finest level with cells inside cell c (its own level for a leaf cell), for the
// cells of `level`; levels are processed finest first.
template <int BW>
__global__ void __launch_bounds__(256) k_sv_reach(DGBase g, int level, uint8_t* __restrict__ reach) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB;
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t c = g.offs[level] * BV + i;
    const uint32_t cb = g.chi[c >> SBB];
    uint8_t r = (uint8_t)level;
    if (cb != NFU) {
      const uint32_t first = cb * BV + 8 * (c & ((1u << SBB) - 1));
#pragma unroll
      for (int j = 0; j < 8; j++) r = min(r, reach[first + j]);
    }
    reach[c] = r;
  }
}

// Fine-level residual R = sum(apron x) - 6 x - alpha div (precision TX), stored as
// rho = (R - shift) / h^2 (precision T); per-workgroup partial sums of R, R^2 and of the
// zero-guess residual R0 and R0^2 go to part[4 * blockIdx.x ...]. A singular level passes the
// mean of R as `shift`, so the zero-mean part is formed before the rounding to T: rounding R
// first would leave an error of T's epsilon times that mean, which does not shrink with R.
template <int BW, class TX, class T, class G>
__global__ void __launch_bounds__(256) k_sv_residual(G g, int level, const TX* __restrict__ x,
                                                     T* __restrict__ rho, double* __restrict__ part, float alpha,
                                                     float h2, double shift) {
  constexpr int BV = Geo<BW>::BV;
  __shared__ double red[4][256];
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  double s[4] = {0, 0, 0, 0};
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t b = g.offs[level] + i / BV;
    const int t = (int)(i % BV);
    const uint32_t c = b * BV + t;
    TX sum = 0, ghosts = 0;
#pragma unroll
    for (int d = 0; d < 6; d++) {
      bool co;
      const uint32_t nb = faceNb<BW>(g, b, t, d, co);
      const TX v = x[nb];
      sum += v;
      if (co) ghosts += v;
    }
    const TX ad = (TX)alpha * (TX)g.div[c];
    const TX R = sum - (TX)6 * x[c] - ad;
    const TX R0 = ghosts - ad;
    rho[c] = (T)((R - (TX)shift) / (TX)h2);
    s[0] += (double)R;
    s[1] += (double)R * (double)R;
    s[2] += (double)R0;
    s[3] += (double)R0 * (double)R0;
  }
#pragma unroll
  for (int k = 0; k < 4; k++) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// rho -= shift on the cells of a level (zero-mean right-hand side of a singular level).
template <int BW, class T>
__global__ void __launch_bounds__(256) k_sv_shift(DGBase g, int level, T* __restrict__ v, double shift) {
  constexpr int BV = Geo<BW>::BV;
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t c = g.offs[level] * BV + i;
    v[c] = (T)((double)v[c] - shift);
  }
}

// Sum of v over the cells of a level (per-workgroup partials).
template <int BW, class T>
__global__ void __launch_bounds__(256) k_sv_sum(DGBase g, int level, const T* __restrict__ v, double* __restrict__ part) {
  constexpr int BV = Geo<BW>::BV;
  __shared__ double red[256];
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  double s = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += (double)v[g.offs[level] * BV + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// e = 0, rho = 0 on the cells of a level.
template <int BW, class T>
__global__ void __launch_bounds__(256) k_sv_zero(DGBase g, int level, T* __restrict__ e, T* __restrict__ rho) {
  constexpr int BV = Geo<BW>::BV;
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t c = g.offs[level] * BV + i;
    e[c] = (T)0;
    rho[c] = (T)0;
  }
}

// (A e)_c / h^2 pieces of an active cell: sum of the coupled neighbours and the diagonal.
// Neighbours on coarser levels and inactive ones are the Dirichlet boundary of the region being
// solved. On the level itself they hold a zero correction; on coarser grids the zero lies
// between the boundary cell and that neighbour (where the finest grid puts it), which a ghost
// value gm * e_c (gm <= 0, from the grid's depth) reproduces.
template <int BW, class T>
__device__ __forceinline__ void svStencil(const DGBase& g, uint32_t b, int t, int lmin, const T* __restrict__ e,
                                          const uint8_t* __restrict__ reach, T gm, T& sum, T& diag) {
  const uint32_t c = b * Geo<BW>::BV + t;
  sum = (T)0;
  diag = (T)6;
#pragma unroll
  for (int d = 0; d < 6; d++) {
    bool co;
    const uint32_t nb = faceNb<BW>(g, b, t, d, co);
    if (nb == c) {
      diag -= (T)1;
    } else if (!co && reach[nb] <= lmin) {
      sum += e[nb];
    } else {
      diag -= gm;
    }
  }
}

// One red-black Gauss-Seidel half-sweep on the active cells of `level` with parity `color`.
template <int BW, class T>
__global__ void __launch_bounds__(256) k_sv_rbgs(DGBase g, int level, int lmin, int color, T* __restrict__ e,
                                                 const T* __restrict__ rho, const uint8_t* __restrict__ reach,
                                                 float h2, float gm) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB;
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t b = g.offs[level] + i / BV;
    const int t = (int)(i % BV);
    const uint32_t c = b * BV + t;
    if (reach[c] > lmin) continue;
    const int3 p0 = g.pos[b];
    int x, y, z;
    locPos<BB>((uint32_t)t, x, y, z);
    if (((p0.x + x + p0.y + y + p0.z + z) & 1) != color) continue;
    T sum, diag;
    svStencil<BW, T>(g, b, t, lmin, e, reach, (T)gm, sum, diag);
    e[c] = (sum + (T)h2 * rho[c]) / diag;
  }
}

// Residual of the correction equation on the active cells of `level`, restricted to the parent
// cells (level + 1): the mean over each restriction group (inactive cells contribute 0). The
// parents' correction starts at 0.
template <int BW, class T>
__global__ void __launch_bounds__(256) k_sv_restrict(DGBase g, int level, int lmin, T* __restrict__ e,
                                                     T* __restrict__ rho, const uint8_t* __restrict__ reach,
                                                     float h2, float gm) {
  constexpr int BV = Geo<BW>::BV, GPB = BV / 8;
  const uint32_t groups = g.ctl[CTL_LOADS + level] * GPB;
  const T inv = (T)1 / (T)h2;
  for (uint32_t gi = blockIdx.x * 256 + threadIdx.x; gi < groups; gi += gridDim.x * 256) {
    const uint32_t b = g.offs[level] + gi / GPB;
    const int t0 = 8 * (int)(gi % GPB);
    if (g.par[b] == NFU) continue;
    T acc = (T)0;
    for (int j = 0; j < 8; j++) {
      const uint32_t c = b * BV + t0 + j;
      if (reach[c] > lmin) continue;
      T sum, diag;
      svStencil<BW, T>(g, b, t0 + j, lmin, e, reach, (T)gm, sum, diag);
      acc += rho[c] - (diag * e[c] - sum) * inv;
    }
    const uint32_t pc = parentCellOf<BW>(g, b, t0);
    rho[pc] = acc * (T).125;
    e[pc] = (T)0;
  }
}

// e += P e(parent level) on the active cells of `level`: the 27/9/3/1 trilinear stencil of
// k_prolongate over the parent's apron, with any stencil cell that is not an active cell of the
// parent level (coarser, or outside the region being solved) replaced by the own parent cell.
template <int BW, class T>
__global__ void __launch_bounds__(256) k_sv_prolong(DGBase g, int level, int lmin, T* __restrict__ e,
                                                    const uint8_t* __restrict__ reach) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB;
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t b = g.offs[level] + i / BV;
    const int t = (int)(i % BV);
    const uint32_t c = b * BV + t;
    const uint32_t ps = g.par[b];
    if (reach[c] > lmin || ps == NFU) continue;
    const int3 p0 = g.pos[b];
    int lx, ly, lz;
    locPos<BB>((uint32_t)t, lx, ly, lz);
    uint32_t q[8];
    parentStencil<BW>(g, ps, p0.x | lx, p0.y | ly, p0.z | lz, q);
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const bool ok = g.lev[q[k] / BV] == level + 1 && reach[q[k]] <= lmin;
      v[k] = e[ok ? q[k] : q[0]];
    }
    e[c] = e[c] + ((T)27 * v[0] + (T)9 * (v[1] + v[2] + v[3]) + (T)3 * (v[4] + v[5] + v[6]) + v[7]) / (T)64;
  }
}

// ---------------------------------------------------------------- uniform pyramid of the top level
struct UDim {
  int w, h, d;
  uint32_t off;
};
__device__ __forceinline__ uint32_t uAt(const UDim& D, int x, int y, int z) {
  return D.off + ((uint32_t)z * D.h + y) * D.w + x;
}

// Top-level residual restricted into the first pyramid grid (dims of the top level / 2): the
// 8 cells of a restriction group form a 2x2x2 cube.
template <int BW, class T>
__global__ void __launch_bounds__(256) k_sv_restrict_top(DGBase g, int lmin, const T* __restrict__ e,
                                                         const T* __restrict__ rho,
                                                         const uint8_t* __restrict__ reach, T* __restrict__ ue,
                                                         T* __restrict__ urho, uint8_t* __restrict__ uact, UDim D1,
                                                         float h2, float gm) {
  constexpr int BV = Geo<BW>::BV, GPB = BV / 8, BB = Geo<BW>::BB;
  const int top = g.L - 1;
  const uint32_t groups = g.ctl[CTL_LOADS + top] * GPB;
  const T inv = (T)1 / (T)h2;
  for (uint32_t gi = blockIdx.x * 256 + threadIdx.x; gi < groups; gi += gridDim.x * 256) {
    const uint32_t b = g.offs[top] + gi / GPB;
    const int t0 = 8 * (int)(gi % GPB);
    T acc = (T)0;
    bool any = false;
    for (int j = 0; j < 8; j++) {
      const uint32_t c = b * BV + t0 + j;
      if (reach[c] > lmin) continue;
      any = true;
      T sum, diag;
      svStencil<BW, T>(g, b, t0 + j, lmin, e, reach, (T)gm, sum, diag);
      acc += rho[c] - (diag * e[c] - sum) * inv;
    }
    const int3 p0 = g.pos[b];
    int x, y, z;
    locPos<BB>((uint32_t)t0, x, y, z);
    const uint32_t u = uAt(D1, (p0.x + x) >> 1, (p0.y + y) >> 1, (p0.z + z) >> 1);
    urho[u] = acc * (T).125;
    ue[u] = (T)0;
    uact[u] = any ? 1 : 0;
  }
}

// Trilinear 27/9/3/1 prolongation from pyramid grid Dc at fine cell (x,y,z) of a grid of size
// (w,h,d) (edge rows collapse onto their own parent, as k_u_prolongate); inactive coarse cells
// are replaced by the own parent.
template <class T>
__device__ __forceinline__ T uProlong(const UDim& Dc, int x, int y, int z, int w, int h, int d,
                                      const T* __restrict__ ue, const uint8_t* __restrict__ uact) {
  const int X = x >> 1, Y = y >> 1, Z = z >> 1;
  const int sx = (x == 0 || x == w - 1) ? 0 : 2 * (x & 1) - 1;
  const int sy = (y == 0 || y == h - 1) ? 0 : 2 * (y & 1) - 1;
  const int sz = (z == 0 || z == d - 1) ? 0 : 2 * (z & 1) - 1;
  const uint32_t u0 = uAt(Dc, X, Y, Z);
  const uint32_t q[8] = {u0,
                         uAt(Dc, X + sx, Y, Z),
                         uAt(Dc, X, Y + sy, Z),
                         uAt(Dc, X, Y, Z + sz),
                         uAt(Dc, X + sx, Y + sy, Z),
                         uAt(Dc, X + sx, Y, Z + sz),
                         uAt(Dc, X, Y + sy, Z + sz),
                         uAt(Dc, X + sx, Y + sy, Z + sz)};
  T v[8];
#pragma unroll
  for (int k = 0; k < 8; k++) v[k] = ue[uact[q[k]] ? q[k] : u0];
  return ((T)27 * v[0] + (T)9 * (v[1] + v[2] + v[3]) + (T)3 * (v[4] + v[5] + v[6]) + v[7]) / (T)64;
}

// e += P e(pyramid) on the active top-level cells.
template <int BW, class T>
__global__ void __launch_bounds__(256) k_sv_prolong_top(DGBase g, int lmin, T* __restrict__ e,
                                                        const uint8_t* __restrict__ reach, const T* __restrict__ ue,
                                                        const uint8_t* __restrict__ uact, UDim D1) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB;
  const int top = g.L - 1;
  const uint32_t n = g.ctl[CTL_LOADS + top] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t b = g.offs[top] + i / BV;
    const int t = (int)(i % BV);
    const uint32_t c = b * BV + t;
    if (reach[c] > lmin) continue;
    const int3 p0 = g.pos[b];
    int x, y, z;
    locPos<BB>((uint32_t)t, x, y, z);
    e[c] = e[c] + uProlong<T>(D1, p0.x + x, p0.y + y, p0.z + z, 2 * D1.w, 2 * D1.h, 2 * D1.d, ue, uact);
  }
}

// Pyramid stencil: clamped (Neumann) at the domain walls; inactive cells are the Dirichlet
// boundary of the solved region (ghost gm * e_c, as in svStencil).
template <class T>
__device__ __forceinline__ void uStencil(const UDim& D, int x, int y, int z, const T* __restrict__ e,
                                         const uint8_t* __restrict__ act, T gm, T& sum, T& diag) {
  sum = (T)0;
  diag = (T)6;
  const int nx[6] = {x - 1, x + 1, x, x, x, x}, ny[6] = {y, y, y - 1, y + 1, y, y}, nz[6] = {z, z, z, z, z - 1, z + 1};
#pragma unroll
  for (int d = 0; d < 6; d++) {
    if (nx[d] < 0 || ny[d] < 0 || nz[d] < 0 || nx[d] >= D.w || ny[d] >= D.h || nz[d] >= D.d) {
      diag -= (T)1;
      continue;
    }
    const uint32_t u = uAt(D, nx[d], ny[d], nz[d]);
    if (act[u])
      sum += e[u];
    else
      diag -= gm;
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_svu_rbgs(UDim D, int color, T* __restrict__ e, const T* __restrict__ rho,
                                                  const uint8_t* __restrict__ act, float h2, float gm) {
  const uint32_t n = (uint32_t)D.w * D.h * D.d;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % D.w), y = (int)((i / D.w) % D.h), z = (int)(i / ((uint32_t)D.w * D.h));
  if (((x + y + z) & 1) != color) return;
  const uint32_t u = D.off + i;
  if (!act[u]) return;
  T sum, diag;
  uStencil<T>(D, x, y, z, e, act, (T)gm, sum, diag);
  e[u] = (sum + (T)h2 * rho[u]) / diag;
}

// Residual of pyramid grid D restricted into the next grid Dc (thread per coarse cell).
template <class T>
__global__ void __launch_bounds__(256) k_svu_restrict(UDim D, UDim Dc, T* __restrict__ e, T* __restrict__ rho,
                                                      uint8_t* __restrict__ act, float h2, float gm) {
  const uint32_t n = (uint32_t)Dc.w * Dc.h * Dc.d;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int X = (int)(i % Dc.w), Y = (int)((i / Dc.w) % Dc.h), Z = (int)(i / ((uint32_t)Dc.w * Dc.h));
  const T inv = (T)1 / (T)h2;
  T acc = (T)0;
  bool any = false;
  for (int j = 0; j < 8; j++) {
    const int x = 2 * X + ((j >> 2) & 1), y = 2 * Y + ((j >> 1) & 1), z = 2 * Z + (j & 1);
    const uint32_t u = uAt(D, x, y, z);
    if (!act[u]) continue;
    any = true;
    T sum, diag;
    uStencil<T>(D, x, y, z, e, act, (T)gm, sum, diag);
    acc += rho[u] - (diag * e[u] - sum) * inv;
  }
  rho[Dc.off + i] = acc * (T).125;
  e[Dc.off + i] = (T)0;
  act[Dc.off + i] = any ? 1 : 0;
}

template <class T>
__global__ void __launch_bounds__(256) k_svu_prolong(UDim D, UDim Dc, T* __restrict__ e, const uint8_t* __restrict__ act) {
  const uint32_t n = (uint32_t)D.w * D.h * D.d;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t u = D.off + i;
  if (!act[u]) return;
  const int x = (int)(i % D.w), y = (int)((i / D.w) % D.h), z = (int)(i / ((uint32_t)D.w * D.h));
  e[u] = e[u] + uProlong<T>(Dc, x, y, z, D.w, D.h, D.d, e, act);
}

// Coarsest pyramid grid: `sweeps` red-black Gauss-Seidel sweeps in one workgroup.
template <class T>
__global__ void __launch_bounds__(1024) k_svu_bottom(UDim D, T* __restrict__ e, const T* __restrict__ rho,
                                                     const uint8_t* __restrict__ act, float h2, float gm, int sweeps) {
  const uint32_t n = (uint32_t)D.w * D.h * D.d;
  for (int s = 0; s < 2 * sweeps; s++) {
    const int color = s & 1;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      const int x = (int)(i % D.w), y = (int)((i / D.w) % D.h), z = (int)(i / ((uint32_t)D.w * D.h));
      if (((x + y + z) & 1) != color) continue;
      const uint32_t u = D.off + i;
      if (!act[u]) continue;
      T sum, diag;
      uStencil<T>(D, x, y, z, e, act, (T)gm, sum, diag);
      e[u] = (sum + (T)h2 * rho[u]) / diag;
    }
    __syncthreads();
  }
}

// x += e on the cells of a level (the refinement update; FP64 in the mixed solve).
template <int BW, class TX, class T>
__global__ void __launch_bounds__(256) k_sv_update(DGBase g, int level, TX* __restrict__ x, const T* __restrict__ e,
                                                   const uint8_t* __restrict__ reach) {
  constexpr int BV = Geo<BW>::BV;
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t c = g.offs[level] * BV + i;
    x[c] = x[c] + (TX)e[c];
  }
}

// Initial guess of a level from its parent level: the 27/9/3/1 stencil of k_prolongate on x.
template <int BW, class TX>
__global__ void __launch_bounds__(256) k_sv_prolong_x(DGBase g, int level, TX* __restrict__ x) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB;
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t b = g.offs[level] + i / BV;
    const int t = (int)(i % BV);
    const uint32_t ps = g.par[b];
    if (ps == NFU) continue;
    const int3 p0 = g.pos[b];
    int lx, ly, lz;
    locPos<BB>((uint32_t)t, lx, ly, lz);
    uint32_t q[8];
    parentStencil<BW>(g, ps, p0.x | lx, p0.y | ly, p0.z | lz, q);
    x[b * BV + t] = ((TX)27 * x[q[0]] + (TX)9 * (x[q[1]] + x[q[2]] + x[q[3]]) + (TX)3 * (x[q[4]] + x[q[5]] + x[q[6]]) +
                     x[q[7]]) /
                    (TX)64;
  }
}

// pressure = (float) x on the cells of a level.
template <int BW, class TX, class G>
__global__ void __launch_bounds__(256) k_sv_store(G g, int level, const TX* __restrict__ x) {
  constexpr int BV = Geo<BW>::BV;
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t c = g.offs[level] * BV + i;
    g.pre[c] = (float)x[c];
  }
}

}  // namespace dcg

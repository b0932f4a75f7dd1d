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
// top level (dense by construction). Restriction is the mean of the 8 cells of a restriction
// group, prolongation the trilinear 27/9/3/1 stencil. Residuals on coarse grids are kept in
// density units (R / h^2), so restriction is a plain mean and each grid uses its own cell size
// in the operator.
//
// Smoother: red-black Gauss-Seidel inside each 8^3 (BW^3) block, blocks in two colours (block
// coordinate parity). Face-adjacent blocks always have opposite colours and the operator couples
// face neighbours only, so the blocks of one colour are independent: a launch stages each of its
// blocks with the apron faces (the other colour's current values) in LDS and runs all its inner
// sweeps there. One smoothing pass is two launches (one per colour) instead of two launches per
// sweep, and the result does not depend on the order in which workgroups run.
//
// The refinement loop runs on the device: the residual kernel writes per-workgroup partial
// norms, and the next launch (k_sv_decide, one workgroup) sums them in workgroup order, decides
// convergence (or the stall of the working precision) and publishes it in the level's SolveCtl;
// every later kernel of the level's loop returns at once when the level is done, so the host
// enqueues cycles ahead and reads the control block once it expects the level to have converged.
#pragma once

#include "dcgrid_fluid.hpp"

namespace dcg {

// A grid of the uniform pyramid below the top level: size and offset in the pyramid arrays.
struct UDimArg {
  int w, h, d;
  uint32_t off;
};

// Refinement state of one level (device memory, one per level).
struct SolveCtl {
  double r0;      // |R(0)|: residual norm of a zero level pressure
  double rr;      // |R| of the last residual
  double prev;    // |R| of the residual before
  double shift;   // mean of R the residual kernel subtracts before rounding (singular levels)
  double delta;   // (mean - shift) / h^2 still to take from rho on the level's cells
  double mean;    // mean of x after the last update (normalisation of a singular level)
  uint32_t cycles, stall, done, singular;
  uint32_t pad[4];
};
static_assert(sizeof(SolveCtl) == 80, "SolveCtl layout");

__device__ __forceinline__ bool solveDone(const SolveCtl* c) { return *(volatile const uint32_t*)&c->done != 0; }

// Parent cell (level + 1) of cell t of block b: the restriction-group mapping of accumulate<T>.
template <int BW>
__device__ __forceinline__ uint32_t parentCellOf(const DGBase& g, uint32_t b, int t) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB;
  const uint32_t ps = g.par[b];
  return (ps / 8) * BV + ((ps % 8) << SBB) + (uint32_t)(t >> 3);
}

// borderApronCell with the level shift of the descriptor (ND_OUT outside the domain).
template <int BW>
__device__ __forceinline__ uint32_t borderApronCellShiftD(uint32_t b, int ai, int aj, int ak, int4 d, uint32_t& sh) {
  constexpr int BB = Geo<BW>::BB, BV = Geo<BW>::BV;
  sh = ndShift(d);
  const int ci = clampi(ai, 1, BW), cj = clampi(aj, 1, BW), ck = clampi(ak, 1, BW);
  const uint32_t spread = loc<BB>(ci & (BW - 1), cj & (BW - 1), ck & (BW - 1));
  const uint32_t clamped = loc<BB>(ci - 1, cj - 1, ck - 1);
  const uint32_t outCell = b * BV + (d.y ? spread : clamped);
  const uint32_t s = sh == ND_OUT ? 0u : sh;
  const uint32_t inCell = ndBlock(d) * BV + loc<BB>((ai - 1 - d.y) >> s, (aj - 1 - d.z) >> s, (ak - 1 - d.w) >> s);
  return sh == ND_OUT ? outCell : inCell;
}
template <int BW>
__device__ __forceinline__ uint32_t borderApronCellShift(const DGBase& g, uint32_t b, int ai, int aj, int ak,
                                                         uint32_t& sh) {
  const int dx = ai == 0 ? -1 : (ai == BW + 1 ? 1 : 0);
  const int dy = aj == 0 ? -1 : (aj == BW + 1 ? 1 : 0);
  const int dz = ak == 0 ? -1 : (ak == BW + 1 ? 1 : 0);
  return borderApronCellShiftD<BW>(b, ai, aj, ak, g.nd[b * 27 + dirIndex(dx, dy, dz)], sh);
}
// Neighbour-descriptor direction of block face `face` (0..5: x-, x+, y-, y+, z-, z+; the face
// numbering of faceCell).
__device__ __forceinline__ int faceDir(int face) {
  const int s = (face & 1) ? 1 : -1, axis = face >> 1;
  return dirIndex(axis == 0 ? s : 0, axis == 1 ? s : 0, axis == 2 ? s : 0);
}
// The face descriptors a thread of a block's staging needs (its face cells t, t + BV, ..): from
// the six listed beside the block's list entry `entry` when `faces` is given (k_sv_lists), else
// from the block's own descriptors. Neither load depends on the block index of a list.
template <int BW>
__device__ __forceinline__ void svFaceDescs(const DGBase& g, const int4* __restrict__ faces, uint32_t entry, uint32_t b,
                                            int t, int4* dd) {
  constexpr int BV = Geo<BW>::BV, FACE = Geo<BW>::FACE, FPT = (FACE + BV - 1) / BV;
#pragma unroll
  for (int j = 0; j < FPT; j++) {
    const int fc = t + j * BV < FACE ? t + j * BV : FACE - 1;
    const int face = fc / (BW * BW);
    dd[j] = faces ? faces[(size_t)entry * 6 + face] : g.nd[b * 27 + faceDir(face)];
  }
}

// svFaceDescs for a listed block: the descriptors beside its list entry (no load depends on the
// block index).
template <int BW>
__device__ __forceinline__ void svFaceDescsListed(const int4* __restrict__ faces, uint32_t entry, int t, int4* dd) {
  constexpr int BV = Geo<BW>::BV, FACE = Geo<BW>::FACE, FPT = (FACE + BV - 1) / BV;
#pragma unroll
  for (int j = 0; j < FPT; j++) {
    const int fc = t + j * BV < FACE ? t + j * BV : FACE - 1;
    dd[j] = faces[(size_t)entry * 6 + fc / (BW * BW)];
  }
}

// Block origin and colour (parity of the block coordinates) of block b of `level`.
template <int BW>
__device__ __forceinline__ int3 svOrigin(const DGBase& g, int level, uint32_t b) {
  return level >= g.S ? denseOrigin<BW>(g, level, b) : g.pos[b];
}
template <int BW>
__device__ __forceinline__ int blockColour(int3 p0) {
  constexpr int BB = Geo<BW>::BB;
  return ((p0.x >> BB) + (p0.y >> BB) + (p0.z >> BB)) & 1;
}

// reach[c] = finest level with cells inside cell c (its own level for a leaf cell), and
// breach[b] = the finest of its block; levels are processed finest first.
template <int BW>
__global__ void __launch_bounds__(Geo<BW>::WG) k_sv_reach(DGBase g, int level, uint8_t* __restrict__ reach,
                                                        uint8_t* __restrict__ breach) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB, BPW = Geo<BW>::BPW;
  __shared__ int bmin[BPW];
  const uint32_t total = slotTotal(g, level);
  const int sub = threadIdx.x / BV, t = threadIdx.x % BV;
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();
    if (t == 0) bmin[sub] = 255;
    __syncthreads();
    const uint32_t slot = base + sub;
    int r = 255;
    uint32_t b = 0;
    if (slot < total) {
      b = g.offs[level] + slot;
      const uint32_t c = b * BV + t;
      const uint32_t cb = g.chi[c >> SBB];
      r = level;
      if (cb != NFU) {
        const uint32_t first = cb * BV + 8 * (c & ((1u << SBB) - 1));
#pragma unroll
        for (int j = 0; j < 8; j++) r = min(r, (int)reach[first + j]);
      }
      reach[c] = (uint8_t)r;
      atomicMin(&bmin[sub], r);
    }
    __syncthreads();
    if (slot < total && t == 0) breach[b] = (uint8_t)bmin[sub];
  }
}

// Active blocks of the V-cycle grids of every level solve: for lmin <= k, the blocks of level k
// with breach <= lmin, per colour c, at list[(lmin * 2 + c) * B + offs[k] + i] for
// i < count[(lmin * L + k) * 2 + c] (2 L B entries in all), and beside each entry the block's six
// face descriptors (faces[entry * 6 + face]), so that a kernel walking a list reads a block's
// index and its faces together. The order inside a list follows the atomic increments and does
// not matter: the blocks of one list are independent.
template <int BW>
__global__ void __launch_bounds__(256) k_sv_lists(DGBase g, const uint8_t* __restrict__ breach,
                                                  uint32_t* __restrict__ list, int4* __restrict__ faces,
                                                  uint32_t* __restrict__ count) {
  for (int k = 0; k < g.L; k++) {
    const uint32_t n = g.ctl[CTL_LOADS + k];
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
      const uint32_t b = g.offs[k] + i;
      const int c = blockColour<BW>(svOrigin<BW>(g, k, b));
      int4 fd[6];
#pragma unroll
      for (int f = 0; f < 6; f++) fd[f] = g.nd[b * 27 + faceDir(f)];
      for (int lmin = breach[b]; lmin <= k; lmin++) {
        const uint32_t slot = atomicAdd(&count[(lmin * g.L + k) * 2 + c], 1u);
        const size_t entry = (size_t)(lmin * 2 + c) * g.B + g.offs[k] + slot;
        list[entry] = b;
#pragma unroll
        for (int f = 0; f < 6; f++) faces[entry * 6 + f] = fd[f];
      }
    }
  }
}
// The list counts into the host-visible mirror (one thread, after k_sv_lists).
__global__ void k_sv_mirror_counts(const uint32_t* __restrict__ count, uint32_t* __restrict__ mirror, int n) {
  for (int i = 0; i < n; i++) mirror[i] = count[i];
}

// ---------------------------------------------------------------- block staging of a correction
// Coupling codes of a staged block tile (apron index a in [0, A^3)): a value >= 0 is the tile
// entry whose value the neighbour contributes (the entry itself for interior and same-level
// halo cells, an own interior cell for the domain-edge clamp of blocks placed by a move or
// refinement); SV_NEU marks the clamp onto the asking cell itself (Neumann: the diagonal loses
// 1); SV_GHOST a coarser or inactive neighbour (the Dirichlet boundary of the solved region: the
// diagonal loses gm). Tile entry 0 (a corner, never a face) holds 0 for uncoupled neighbours.
enum : int { SV_NEU = -1, SV_GHOST = -2 };

// What a thread loads to stage block b around its local cell t: its own cell's correction and
// reach, and for each of its face cells the neighbour cell, its level shift, reach and correction.
template <int BW, class T>
struct SvStage {
  static constexpr int FPT = (Geo<BW>::FACE + Geo<BW>::BV - 1) / Geo<BW>::BV;
  uint32_t nb[FPT], sh[FPT];
  int rc, rn[FPT];
  T vc, vn[FPT];
};

// The loads of svStageBlock, all issued together (no LDS access).
template <int BW, class T>
__device__ __forceinline__ void svLoadBlock(uint32_t b, int t, const T* __restrict__ e, const uint8_t* __restrict__ reach,
                                            bool ownZero, bool haloZero, const int4* dd, SvStage<BW, T>& S) {
  constexpr int BV = Geo<BW>::BV, FACE = Geo<BW>::FACE, FPT = SvStage<BW, T>::FPT;
  const uint32_t c = b * BV + t;
#pragma unroll
  for (int j = 0; j < FPT; j++) {
    const int fc = t + j * BV < FACE ? t + j * BV : FACE - 1;
    int ai, aj, ak;
    faceCell<BW>(fc, ai, aj, ak);
    S.nb[j] = borderApronCellShiftD<BW>(b, ai, aj, ak, dd[j], S.sh[j]);
  }
  S.rc = reach[c];
  S.vc = ownZero ? (T)0 : e[c];
#pragma unroll
  for (int j = 0; j < FPT; j++) {
    S.rn[j] = reach[S.nb[j]];
    S.vn[j] = haloZero ? (T)0 : e[S.nb[j]];
  }
}

// The LDS half of svStageBlock: V (values as T) and K (codes) of the block's tile from what
// svLoadBlock loaded.
template <int BW, class T>
__device__ __forceinline__ void svWriteBlock(uint32_t b, int t, int lmin, bool valid, T* V, int16_t* K,
                                             const SvStage<BW, T>& S) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, FACE = Geo<BW>::FACE, BB = Geo<BW>::BB, FPT = SvStage<BW, T>::FPT;
  if (!valid) return;
  const int a = apronIndexOfLocal<BW>(t);
  V[a] = S.vc;
  K[a] = S.rc <= lmin ? (int16_t)a : (int16_t)SV_GHOST;
  if (t == 0) {
    V[0] = (T)0;
    K[0] = (int16_t)SV_GHOST;
  }
#pragma unroll
  for (int j = 0; j < FPT; j++) {
    if (t + j * BV >= FACE) continue;
    int ai, aj, ak;
    faceCell<BW>(t + j * BV, ai, aj, ak);
    const int fa = (ai * A + aj) * A + ak;
    int code;
    if (S.sh[j] == ND_OUT) {
      // the clamp onto an own cell: the asking cell itself (Neumann) or another own cell
      const uint32_t own = S.nb[j] - b * BV;
      const uint32_t ask = loc<BB>(clampi(ai, 1, BW) - 1, clampi(aj, 1, BW) - 1, clampi(ak, 1, BW) - 1);
      code = own == ask ? SV_NEU : (S.rn[j] <= lmin ? apronIndexOfLocal<BW>((int)own) : SV_GHOST);
    } else if (S.sh[j] != 0) {
      code = SV_GHOST;
    } else {
      code = S.rn[j] <= lmin ? fa : SV_GHOST;
      V[fa] = S.vn[j];
    }
    K[fa] = (int16_t)code;
  }
}

// Stages the correction e of block b around local cell t: V (values as T), K (codes). Loads of
// a thread are issued together.
template <int BW, class T>
__device__ __forceinline__ void svStageBlock(uint32_t b, int t, int lmin, const T* __restrict__ e,
                                             const uint8_t* __restrict__ reach, bool ownZero, bool haloZero,
                                             bool valid, T* V, int16_t* K, const int4* dd) {
  SvStage<BW, T> S;
  svLoadBlock<BW, T>(b, t, e, reach, ownZero, haloZero, dd, S);
  __builtin_amdgcn_sched_barrier(0);
  svWriteBlock<BW, T>(b, t, lmin, valid, V, K, S);
}

// Neighbour tile entries (uncoupled ones point at the zero entry) and diagonal of the cell at
// tile entry a.
template <int A, class T>
__device__ __forceinline__ T svCoupling(const int16_t* K, int a, T gm, int idx[6]) {
  const int n[6] = {a - A * A, a + A * A, a - A, a + A, a - 1, a + 1};
  T diag = (T)6;
#pragma unroll
  for (int d = 0; d < 6; d++) {
    const int k = K[n[d]];
    idx[d] = k >= 0 ? k : 0;
    diag -= k >= 0 ? (T)0 : (k == SV_NEU ? (T)1 : gm);
  }
  return diag;
}

// One smoothing pass on the listed blocks (one colour of grid `level`, active cells: reach <=
// lmin): nu red-black Gauss-Seidel sweeps of the block in LDS against its faces. ownZero: the
// block's correction starts at 0 this pass; haloZero: so do its neighbours'. rho is in density
// units (R / h^2); on the solved level itself the pending mean correction delta is taken from it.
template <int BW, class T>
__global__ void __launch_bounds__(Geo<BW>::WG) k_sv_smooth(DGBase g, int level, int lmin, const uint32_t* __restrict__ list,
                                                         const int4* __restrict__ faces, uint32_t count, int nu,
                                                         int ownZero, int haloZero, T* __restrict__ e,
                                                         const T* __restrict__ rho, const uint8_t* __restrict__ reach,
                                                         float h2, float gm, const SolveCtl* __restrict__ ctl) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW, BB = Geo<BW>::BB;
  constexpr int FPT = (Geo<BW>::FACE + BV - 1) / BV;
  __shared__ T Vs[BPW][A3];
  __shared__ int16_t Ks[BPW][A3];
  const int sub = threadIdx.x / BV, t = threadIdx.x % BV;
  const int a = apronIndexOfLocal<BW>(t);
  int lx, ly, lz;
  locPos<BB>((uint32_t)t, lx, ly, lz);
  const int par = (lx + ly + lz) & 1;
  T* V = Vs[sub];
  int16_t* K = Ks[sub];
  const uint32_t stride = gridDim.x * BPW;
  uint32_t base = blockIdx.x * BPW;
  if (base >= count) return;
  // The loads of a block are issued before the sweeps of the block before it (same colour: the
  // blocks of one launch never read each other's cells, so the early loads see the same values).
  // Nothing is computed from the loaded values before the sweeps, so no wait for them precedes
  // the sweeps; a workgroup's last block runs a copy of the sweeps without loads.
  uint32_t i = base + sub;
  bool work = i < count;
  uint32_t entry = g.offs[level] + (work ? i : count - 1);
  uint32_t b = list[entry];
  int4 dd[FPT];
  svFaceDescsListed<BW>(faces, entry, t, dd);
  if (solveDone(ctl)) return;  // the control block, read while the list loads are in flight
  const T delta = level == lmin ? (T)ctl->delta : (T)0;
  T rv = rho[b * BV + t];
  SvStage<BW, T> S;
  svLoadBlock<BW, T>(b, t, e, reach, ownZero != 0, haloZero != 0, dd, S);
  // nu red-black sweeps of the staged block cb and the store of its active cells
  const auto sweep = [&](uint32_t cb, bool cwork, T hr) {
    int idx[6];
    const T diag = svCoupling<A, T>(K, a, (T)gm, idx);
    const bool act = cwork && K[a] >= 0;
    const T inv = (T)1 / diag;
    for (int s = 0; s < nu; s++) {
#pragma unroll
      for (int col = 0; col < 2; col++) {
        if (act && par == col) V[a] = (V[idx[0]] + V[idx[1]] + V[idx[2]] + V[idx[3]] + V[idx[4]] + V[idx[5]] + hr) * inv;
        __syncthreads();
      }
    }
    if (act) e[cb * BV + t] = V[a];
  };
  for (;;) {
    __syncthreads();  // the previous block's tile is no longer read
    svWriteBlock<BW, T>(b, t, lmin, work, V, K, S);
    __syncthreads();
    const uint32_t cb = b;
    const bool cwork = work;
    const T hr = (T)h2 * (cwork ? rv - delta : (T)0);
    base += stride;
    if (base >= count) {  // the workgroup's last block: nothing to load
      sweep(cb, cwork, hr);
      break;
    }
    i = base + sub;
    work = i < count;
    entry = g.offs[level] + (work ? i : count - 1);
    b = list[entry];
    svFaceDescsListed<BW>(faces, entry, t, dd);
    rv = rho[b * BV + t];
    svLoadBlock<BW, T>(b, t, e, reach, ownZero != 0, haloZero != 0, dd, S);
    sweep(cb, cwork, hr);
  }
}

// Residual of the correction equation on the active cells of grid `level`, restricted to the
// parent cells: the mean over each restriction group (inactive cells contribute 0). TOP: every
// block of the (dense) top level, the parents are the first pyramid grid D1, whose activity is
// recorded too (it changes with lmin); else the listed blocks (both colours, n0 + n1) and the
// cells of level + 1.
template <int BW, class T, bool TOP>
__global__ void __launch_bounds__(Geo<BW>::WG) k_sv_restrict(DGBase g, int level, int lmin, const uint32_t* __restrict__ list0,
                                                           const uint32_t* __restrict__ list1,
                                                           const int4* __restrict__ faces0,
                                                           const int4* __restrict__ faces1, uint32_t n0, uint32_t n1,
                                                           const T* __restrict__ e, T* __restrict__ rho,
                                                           const uint8_t* __restrict__ reach, float h2, float gm,
                                                           T* __restrict__ urho, uint8_t* __restrict__ uact,
                                                           UDimArg D1, const SolveCtl* __restrict__ ctl) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW, BB = Geo<BW>::BB;
  constexpr int FPT = (Geo<BW>::FACE + BV - 1) / BV;
  __shared__ T Vs[BPW][A3];
  __shared__ int16_t Ks[BPW][A3];
  __shared__ T Rs[BPW][BV];
  const uint32_t total = TOP ? g.ctl[CTL_LOADS + level] : n0 + n1;
  const int sub = threadIdx.x / BV, t = threadIdx.x % BV;
  const int a = apronIndexOfLocal<BW>(t);
  T* V = Vs[sub];
  int16_t* K = Ks[sub];
  T delta = (T)0;
  bool first = true;
  for (uint32_t base = blockIdx.x * BPW; base < total; base += gridDim.x * BPW) {
    const uint32_t i = base + sub;
    bool work = i < total;
    const uint32_t j = work ? i : total - 1;
    const uint32_t entry = TOP ? 0u : g.offs[level] + (j < n0 ? j : j - n0);
    const uint32_t b = TOP ? g.offs[level] + j : (j < n0 ? list0[entry] : list1[entry]);
    int4 dd[FPT];
    svFaceDescs<BW>(g, TOP ? nullptr : (j < n0 ? faces0 : faces1), entry, b, t, dd);
    if (first) {  // the control block, read while the list loads are in flight
      if (solveDone(ctl)) return;
      delta = level == lmin ? (T)ctl->delta : (T)0;
      first = false;
    }
    if (!TOP && work) work = g.par[b] != NFU;
    __syncthreads();
    const T r = work ? rho[b * BV + t] - delta : (T)0;
    svStageBlock<BW, T>(b, t, lmin, e, reach, false, false, work, V, K, dd);
    __syncthreads();
    int idx[6];
    const T diag = svCoupling<A, T>(K, a, (T)gm, idx);
    const bool act = work && K[a] >= 0;
    const T sum = V[idx[0]] + V[idx[1]] + V[idx[2]] + V[idx[3]] + V[idx[4]] + V[idx[5]];
    Rs[sub][t] = act ? r - (diag * V[a] - sum) / (T)h2 : (T)0;
    __syncthreads();
    if (work && t < BV / 8) {
      T acc = (T)0;
      bool anyAct = false;
#pragma unroll
      for (int j8 = 0; j8 < 8; j8++) {
        acc += Rs[sub][8 * t + j8];
        anyAct |= K[apronIndexOfLocal<BW>(8 * t + j8)] >= 0;
      }
      if (TOP) {
        const int3 p0 = svOrigin<BW>(g, level, b);
        int x, y, z;
        locPos<BB>((uint32_t)(8 * t), x, y, z);
        const uint32_t u = D1.off + ((uint32_t)((p0.z + z) >> 1) * D1.h + ((p0.y + y) >> 1)) * D1.w + ((p0.x + x) >> 1);
        urho[u] = acc * (T).125;
        uact[u] = anyAct ? 1 : 0;
      } else {
        rho[parentCellOf<BW>(g, b, 8 * t)] = acc * (T).125;
      }
    }
  }
}

// e += P e(parent level) on the active cells of the listed blocks of `level` (both colours):
// the 27/9/3/1 trilinear stencil of k_prolongate over the parent's apron, with any stencil cell
// that is not an active cell of the parent level (coarser, or outside the region being solved)
// replaced by the own parent cell. As in k_prolongate, the (BW/2 + 2)^3 region of the parent's
// apron around the block's octant is staged once per block (one apron lookup, level and reach
// check per entry) and every cell takes its eight stencil values from LDS.
template <int BW, class T>
__global__ void __launch_bounds__(Geo<BW>::WG) k_sv_prolong(DGBase g, int level, int lmin, const uint32_t* __restrict__ list0,
                                                          const uint32_t* __restrict__ list1, uint32_t n0, uint32_t n1,
                                                          T* __restrict__ e, const uint8_t* __restrict__ reach,
                                                          const SolveCtl* __restrict__ ctl) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB, BPW = Geo<BW>::BPW, H = BW / 2 + 2, H3 = H * H * H;
  static_assert(H3 <= BV, "a block's threads stage the parent region");
  __shared__ T vals[BPW][H3];
  __shared__ uint8_t oks[BPW][H3];
  const uint32_t total = n0 + n1;
  const int sub = threadIdx.x / BV, t = threadIdx.x % BV;
  int lx, ly, lz;
  locPos<BB>((uint32_t)t, lx, ly, lz);
  bool first = true;
  for (uint32_t base = blockIdx.x * BPW; base < total; base += gridDim.x * BPW) {
    const uint32_t j = base + sub;
    const bool work = j < total;
    const uint32_t jj = work ? j : total - 1;
    const uint32_t b = jj < n0 ? list0[g.offs[level] + jj] : list1[g.offs[level] + jj - n0];
    if (first) {  // the control block, read while the list load is in flight
      if (solveDone(ctl)) return;
      first = false;
    }
    __syncthreads();  // the previous block's region is no longer read
    const uint32_t ps = work ? g.par[b] : NFU;
    const uint32_t c = b * BV + t;
    const int rc = reach[c];
    int3 p0 = make_int3(0, 0, 0);
    int ox = 0, oy = 0, oz = 0;  // parent apron coordinates of the region's first entry
    if (ps != NFU) {
      p0 = g.pos[b];
      ox = (p0.x / 2) % BW;
      oy = (p0.y / 2) % BW;
      oz = (p0.z / 2) % BW;
      if (t < H3) {
        const int i = t / (H * H), jy = (t / H) % H, k = t % H;
        const uint32_t q = apronCell<BW>(g, ps / 8, ox + i, oy + jy, oz + k);
        oks[sub][t] = g.lev[q / BV] == level + 1 && reach[q] <= lmin;
        vals[sub][t] = e[q];
      }
    }
    __syncthreads();
    if (ps == NFU || rc > lmin) continue;
    const int x = p0.x | lx, y = p0.y | ly, z = p0.z | lz;
    // parentStencil's apron coordinates, relative to the staged region
    const int ti = 1 + (x / 2) % BW - ox, tj = 1 + (y / 2) % BW - oy, tk = 1 + (z / 2) % BW - oz;
    const int si = x % 2 ? 1 : -1, sj = y % 2 ? 1 : -1, sk = z % 2 ? 1 : -1;
    auto at = [&](int i, int jy, int k) { return (i * H + jy) * H + k; };
    const int r[8] = {at(ti, tj, tk),           at(ti, tj, tk + sk),      at(ti, tj + sj, tk),
                      at(ti + si, tj, tk),      at(ti, tj + sj, tk + sk), at(ti + si, tj, tk + sk),
                      at(ti + si, tj + sj, tk), at(ti + si, tj + sj, tk + sk)};
    T v[8];
#pragma unroll
    for (int m = 0; m < 8; m++) v[m] = vals[sub][oks[sub][r[m]] ? r[m] : r[0]];
    e[c] = e[c] + ((T)27 * v[0] + (T)9 * (v[1] + v[2] + v[3]) + (T)3 * (v[4] + v[5] + v[6]) + v[7]) / (T)64;
  }
}

// ---------------------------------------------------------------- uniform pyramid of the top level
__device__ __forceinline__ uint32_t uAt(const UDimArg& D, int x, int y, int z) {
  return D.off + ((uint32_t)z * D.h + y) * D.w + x;
}

// Trilinear 27/9/3/1 prolongation from pyramid grid Dc at fine cell (x,y,z) of a grid of size
// (w,h,d) (edge rows collapse onto their own parent, as k_u_prolongate); inactive coarse cells
// are replaced by the own parent.
template <class T>
__device__ __forceinline__ T uProlong(const UDimArg& Dc, int x, int y, int z, int w, int h, int d,
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
                                                        const uint8_t* __restrict__ uact, UDimArg D1,
                                                        const SolveCtl* __restrict__ ctl) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB;
  if (solveDone(ctl)) return;
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
// boundary of the solved region (ghost gm * e_c).
template <class T>
__device__ __forceinline__ void uStencil(const UDimArg& D, int x, int y, int z, const T* __restrict__ e,
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

// One smoothing pass on the tiles of `colour` of pyramid grid D (8^3 tiles, colour = tile
// coordinate parity): nu red-black sweeps of the tile in LDS against its faces, as k_sv_smooth.
template <class T>
__global__ void __launch_bounds__(512) k_svu_smooth(UDimArg D, int colour, int nu, int ownZero, int haloZero,
                                                    T* __restrict__ e, const T* __restrict__ rho,
                                                    const uint8_t* __restrict__ act, float h2, float gm,
                                                    const SolveCtl* __restrict__ ctl) {
  constexpr int TW = 8, A = TW + 2, A3 = A * A * A, FACE = 6 * TW * TW;
  __shared__ T V[A3];
  __shared__ int16_t K[A3];
  if (solveDone(ctl)) return;
  const int ntx = (D.w + TW - 1) / TW, nty = (D.h + TW - 1) / TW, ntz = (D.d + TW - 1) / TW;
  const int tiles = ntx * nty * ntz;
  const int t = threadIdx.x;
  const int lx = t % TW, ly = (t / TW) % TW, lz = t / (TW * TW);
  const int a = ((lx + 1) * A + (ly + 1)) * A + (lz + 1);
  for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    if (((tx + ty + tz) & 1) != colour) continue;  // uniform over the workgroup
    __syncthreads();
    const int x0 = tx * TW, y0 = ty * TW, z0 = tz * TW;
    // own cell
    const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
    const bool in = x < D.w && y < D.h && z < D.d;
    const uint32_t u = in ? uAt(D, x, y, z) : 0;
    const int ac = in ? act[u] : 0;
    const T vc = (in && !ownZero) ? e[u] : (T)0;
    const T r = in ? rho[u] : (T)0;
    // face entries of the halo (the clamp and inactive codes as in uStencil)
    int fx = 0, fy = 0, fz = 0, fa = 0, fcode = SV_GHOST;
    T fv = (T)0;
    if (t < FACE) {
      int ai, aj, ak;
      faceCell<TW>(t, ai, aj, ak);
      fa = (ai * A + aj) * A + ak;
      fx = x0 + ai - 1;
      fy = y0 + aj - 1;
      fz = z0 + ak - 1;
      if (fx < 0 || fy < 0 || fz < 0 || fx >= D.w || fy >= D.h || fz >= D.d) {
        fcode = SV_NEU;
      } else {
        const uint32_t fu = uAt(D, fx, fy, fz);
        const int fact = act[fu];
        fv = haloZero ? (T)0 : e[fu];
        fcode = fact ? fa : SV_GHOST;
      }
    }
    V[a] = vc;
    K[a] = in ? (ac ? (int16_t)a : (int16_t)SV_GHOST) : (int16_t)SV_NEU;
    if (t < FACE) {
      V[fa] = fv;
      K[fa] = (int16_t)fcode;
    }
    if (t == 0) {
      V[0] = (T)0;
      K[0] = (int16_t)SV_GHOST;
    }
    __syncthreads();
    int idx[6];
    const T diag = svCoupling<A, T>(K, a, (T)gm, idx);
    const bool live = in && ac;
    const T hr = (T)h2 * r;
    const T inv = (T)1 / diag;
    const int par = (x + y + z) & 1;
    for (int s = 0; s < nu; s++) {
#pragma unroll
      for (int col = 0; col < 2; col++) {
        if (live && par == col) V[a] = (V[idx[0]] + V[idx[1]] + V[idx[2]] + V[idx[3]] + V[idx[4]] + V[idx[5]] + hr) * inv;
        __syncthreads();
      }
    }
    if (live) e[u] = V[a];
  }
}

// Residual of pyramid grid D restricted into the next grid Dc: a thread per fine cell (8
// consecutive lanes per coarse cell, child order 000 .. 111), every load of a cell issued at
// once; the group's first lane sums the active children in child order, as a loop over the
// children would.
template <class T>
__global__ void __launch_bounds__(256) k_svu_restrict(UDimArg D, UDimArg Dc, const T* __restrict__ e,
                                                      T* __restrict__ rho, uint8_t* __restrict__ act, float h2, float gm,
                                                      const SolveCtl* __restrict__ ctl) {
  if (solveDone(ctl)) return;
  const uint32_t n = (uint32_t)Dc.w * Dc.h * Dc.d;
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  const uint32_t i = t >> 3;  // the same for the 8 lanes of a group
  const int j = (int)(t & 7u);
  const bool ok = i < n;
  const uint32_t ic = ok ? i : 0u;
  const int X = (int)(ic % Dc.w), Y = (int)((ic / Dc.w) % Dc.h), Z = (int)(ic / ((uint32_t)Dc.w * Dc.h));
  const int x = 2 * X + ((j >> 2) & 1), y = 2 * Y + ((j >> 1) & 1), z = 2 * Z + (j & 1);
  const uint32_t u = uAt(D, x, y, z);
  const int nx[6] = {x - 1, x + 1, x, x, x, x}, ny[6] = {y, y, y - 1, y + 1, y, y}, nz[6] = {z, z, z, z, z - 1, z + 1};
  bool in[6];
  uint8_t na[6];
  T ne[6];
#pragma unroll
  for (int d = 0; d < 6; d++) {
    in[d] = !(nx[d] < 0 || ny[d] < 0 || nz[d] < 0 || nx[d] >= D.w || ny[d] >= D.h || nz[d] >= D.d);
    const uint32_t q = in[d] ? uAt(D, nx[d], ny[d], nz[d]) : u;
    na[d] = act[q];
    ne[d] = e[q];
  }
  const bool a = ok && act[u] != 0;
  const T ec = e[u], rc = rho[u];
  // uStencil's sum and diagonal, in its order
  T sum = (T)0, diag = (T)6;
#pragma unroll
  for (int d = 0; d < 6; d++) {
    if (!in[d])
      diag -= (T)1;
    else if (na[d])
      sum += ne[d];
    else
      diag -= (T)gm;
  }
  const T r = a ? rc - (diag * ec - sum) * ((T)1 / (T)h2) : (T)0;
  const int base = (int)(threadIdx.x & ~7u);
  T acc = (T)0;
  bool any = false;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const T rk = __shfl(r, base + k, 64);
    const int ak = __shfl(a ? 1 : 0, base + k, 64);
    if (ak) acc += rk;
    any |= ak != 0;
  }
  if (ok && j == 0) {
    rho[Dc.off + i] = acc * (T).125;
    act[Dc.off + i] = any ? 1 : 0;
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_svu_prolong(UDimArg D, UDimArg Dc, T* __restrict__ e,
                                                     const uint8_t* __restrict__ act, const SolveCtl* __restrict__ ctl) {
  if (solveDone(ctl)) return;
  const uint32_t n = (uint32_t)D.w * D.h * D.d;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t u = D.off + i;
  if (!act[u]) return;
  const int x = (int)(i % D.w), y = (int)((i / D.w) % D.h), z = (int)(i / ((uint32_t)D.w * D.h));
  e[u] = e[u] + uProlong<T>(Dc, x, y, z, D.w, D.h, D.d, e, act);
}

// The coarse end of the pyramid (grids jc..J-1, at most kCoarseCells cells together) in one
// workgroup with every grid in LDS: sweeps, restriction of the residual, down to the coarsest
// grid (bottom sweeps there), then prolongation and sweeps back up. Reads rho and the active
// cells of grid jc, writes its correction. Each grid's couplings (a neighbour flag per face,
// parity, activity) and inverse diagonal are computed once per call; grids of at most 64 cells
// are swept by one wavefront without workgroup barriers.
static constexpr int kCoarseCells = 4800;
static constexpr int kCoarseThreads = 1024;
static constexpr int kMaxPyr = 12;
struct CoarseArgs {
  UDimArg D[kMaxPyr];   // grids jc.., offsets into the workgroup's LDS arrays
  float h2[kMaxPyr];
  float gm[kMaxPyr];
  uint32_t goff;        // offset of grid jc in the pyramid arrays
  int n, sweeps, bottom;
  int slot;             // cache slot of the level solve (its lmin)
  uint32_t tag;         // the solve's tag: a slot holding it is current
};
__device__ __forceinline__ void cCoord(const UDimArg& D, uint32_t i, int& x, int& y, int& z) {
  x = (int)(i % (uint32_t)D.w);
  const uint32_t q = i / (uint32_t)D.w;
  y = (int)(q % (uint32_t)D.h);
  z = (int)(q / (uint32_t)D.h);
}
// fl: bits 0-5 neighbour d (x-, x+, y-, y+, z-, z+) coupled (inside the grid and active), bit 6
// parity, bit 7 active; inv = 1 / diagonal (walls Neumann, inactive neighbours ghosts gm e_c).
// The active cells of each parity are listed in lst (the grid's range of the LDS arrays: parity
// 0 from its start, parity 1 from its end; cnt[0], cnt[1] their numbers, zero on entry). The
// order within a list does not matter: a half-sweep's cells read only cells of the other parity.
// A thread takes cells threadIdx.x + k * blockDim.x, k < K (K = ceil(n / blockDim.x)): the loads
// of all its cells are issued together, and each wavefront reserves its list entries of all K
// rounds with one LDS atomic per parity.
template <class T, int K>
__device__ __forceinline__ void cPrepareK(const UDimArg& D, const uint8_t* act, uint8_t* fl, T* inv, T gm,
                                          uint16_t* lst, uint32_t* cnt) {
  const uint32_t n = (uint32_t)D.w * D.h * D.d;
  uint32_t f[K];
  T diag[K];
  bool valid[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const uint32_t i = threadIdx.x + (uint32_t)k * blockDim.x;
    valid[k] = i < n;
    int x, y, z;
    cCoord(D, valid[k] ? i : 0u, x, y, z);
    const uint32_t u = D.off + (valid[k] ? i : 0u);
    const int nx[6] = {x - 1, x + 1, x, x, x, x}, ny[6] = {y, y, y - 1, y + 1, y, y}, nz[6] = {z, z, z, z, z - 1, z + 1};
    bool in[6];
    uint8_t na[6];
#pragma unroll
    for (int d = 0; d < 6; d++) {
      in[d] = !(nx[d] < 0 || ny[d] < 0 || nz[d] < 0 || nx[d] >= D.w || ny[d] >= D.h || nz[d] >= D.d);
      na[d] = act[in[d] ? uAt(D, nx[d], ny[d], nz[d]) : u];
    }
    const uint8_t self = act[u];
    uint32_t fk = 0;
    T dg = (T)6;
#pragma unroll
    for (int d = 0; d < 6; d++) {
      if (!in[d]) {
        dg -= (T)1;
      } else if (na[d]) {
        fk |= 1u << d;
      } else {
        dg -= gm;
      }
    }
    fk |= (uint32_t)((x + y + z) & 1) << 6;
    fk |= (valid[k] && self) ? 0x80u : 0u;
    f[k] = fk;
    diag[k] = dg;
  }
  const uint32_t lane = threadIdx.x & 63u;
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int par = 0; par < 2; par++) {
    unsigned long long m[K];
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
      m[k] = __ballot((f[k] & 0x80u) && ((f[k] >> 6) & 1u) == (uint32_t)par);
      tot += (uint32_t)__popcll(m[k]);
    }
    uint32_t base = 0;
    if (tot && lane == (uint32_t)(__ffsll((long long)__ballot(1)) - 1)) base = atomicAdd(&cnt[par], tot);
    base = __builtin_amdgcn_readfirstlane(base);
#pragma unroll
    for (int k = 0; k < K; k++) {
      const uint32_t i = threadIdx.x + (uint32_t)k * blockDim.x;
      if ((m[k] >> lane) & 1ull) {
        const uint32_t pos = base + (uint32_t)__popcll(m[k] & below);
        lst[par ? D.off + n - 1 - pos : D.off + pos] = (uint16_t)(D.off + i);
      }
      base += (uint32_t)__popcll(m[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    if (!valid[k]) continue;
    const uint32_t u = D.off + threadIdx.x + (uint32_t)k * blockDim.x;
    fl[u] = (uint8_t)f[k];
    inv[u] = (T)1 / diag[k];
  }
}
static constexpr int kCPrepRounds = (kCoarseCells + kCoarseThreads - 1) / kCoarseThreads;
static_assert(kCPrepRounds == 5, "one case per round count");
template <class T>
__device__ __forceinline__ void cPrepare(const UDimArg& D, const uint8_t* act, uint8_t* fl, T* inv, T gm,
                                         uint16_t* lst, uint32_t* cnt) {
  const uint32_t n = (uint32_t)D.w * D.h * D.d;
  const uint32_t rounds = (n + blockDim.x - 1) / blockDim.x;
  if (rounds <= 1)
    cPrepareK<T, 1>(D, act, fl, inv, gm, lst, cnt);
  else if (rounds == 2)
    cPrepareK<T, 2>(D, act, fl, inv, gm, lst, cnt);
  else if (rounds == 3)
    cPrepareK<T, 3>(D, act, fl, inv, gm, lst, cnt);
  else if (rounds == 4)
    cPrepareK<T, 4>(D, act, fl, inv, gm, lst, cnt);
  else
    cPrepareK<T, 5>(D, act, fl, inv, gm, lst, cnt);
}
// Sum of the coupled neighbours of LDS cell u: all six loads issued unconditionally (indices
// clamped into the arrays) and selected by the flags, so a cell costs one round of LDS latency.
template <class T>
__device__ __forceinline__ T cNeighbourSum(const UDimArg& D, uint32_t u, uint32_t f, const T* e) {
  const int sy = D.w, sz = D.w * D.h, last = kCoarseCells - 1, iu = (int)u;
  const T a0 = e[max(iu - 1, 0)], a1 = e[min(iu + 1, last)];
  const T a2 = e[max(iu - sy, 0)], a3 = e[min(iu + sy, last)];
  const T a4 = e[max(iu - sz, 0)], a5 = e[min(iu + sz, last)];
  T sum = (f & 1u) ? a0 : (T)0;
  sum += (f & 2u) ? a1 : (T)0;
  sum += (f & 4u) ? a2 : (T)0;
  sum += (f & 8u) ? a3 : (T)0;
  sum += (f & 16u) ? a4 : (T)0;
  sum += (f & 32u) ? a5 : (T)0;
  return sum;
}
// Red-black sweeps of one grid over its parity lists: each thread keeps its (at most S) cells of
// each parity in registers (LDS index, flags, h^2 rho, inverse diagonal), so a half-sweep is one
// round of LDS loads for all of them, then the stores. A grid whose parities have at most
// kCWaveSlots<T> * 64 active cells each is swept by the first wavefront alone, with wavefront
// ordering instead of workgroup barriers (two slots in FP64 keep the kernel within its registers).
static constexpr int kCSlots = 3;
template <class T>
constexpr int kCWaveSlots = sizeof(T) == 4 ? 4 : 2;
static_assert(kCSlots * kCoarseThreads >= (kCoarseCells + 1) / 2, "a parity of a coarse grid exceeds the slots");
template <class T, int S>
struct CParity {
  uint32_t u[S], f[S];
  T hr[S], iv[S];
  int ns;  // slots in use (the same in every thread): loads are issued for these only
  __device__ __forceinline__ void load(const UDimArg& D, uint32_t n, uint32_t nt, int par, uint32_t count,
                                       const uint16_t* lst, const uint8_t* fl, const T* rho, const T* inv, T h2) {
    ns = (int)((count + nt - 1) / nt);
#pragma unroll
    for (int k = 0; k < S; k++) {
      const uint32_t j = threadIdx.x + (uint32_t)k * nt;
      const bool ok = j < count;
      const uint32_t li = ok ? (par ? D.off + n - 1 - j : D.off + j) : D.off;
      const uint32_t uu = ok ? (uint32_t)lst[li] : D.off;
      u[k] = uu;
      f[k] = ok ? (uint32_t)fl[uu] : 0u;
      hr[k] = h2 * rho[uu];
      iv[k] = inv[uu];
    }
  }
  template <int N>
  __device__ __forceinline__ void sweepN(const UDimArg& D, T* e) const {
    T v[N];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = (cNeighbourSum<T>(D, u[k], f[k], e) + hr[k]) * iv[k];
#pragma unroll
    for (int k = 0; k < N; k++)
      if (f[k] & 0x80u) e[u[k]] = v[k];
  }
  __device__ __forceinline__ void sweep(const UDimArg& D, T* e) const {
    static_assert(S >= 2 && S <= 4, "one case per slot count");
    if (ns == 1)
      sweepN<1>(D, e);
    else if (ns == 2 || (S == 2 && ns > 2))
      sweepN<2>(D, e);
    else if (ns == 3 || (S == 3 && ns > 3))
      sweepN<(S < 3 ? S : 3)>(D, e);
    else if (ns >= 4)
      sweepN<S>(D, e);
  }
};
template <class T, int S>
__device__ __forceinline__ void cSweepLoop(const UDimArg& D, T* e, const T* rho, const uint8_t* fl, const T* inv,
                                           const uint16_t* lst, const uint32_t* cnt, T h2, int sweeps, bool wave) {
  const uint32_t n = (uint32_t)D.w * D.h * D.d;
  const uint32_t nt = wave ? 64u : blockDim.x;
  CParity<T, S> p0, p1;
  p0.load(D, n, nt, 0, cnt[0], lst, fl, rho, inv, h2);
  p1.load(D, n, nt, 1, cnt[1], lst, fl, rho, inv, h2);
  auto sync = [&]() {
    if (wave) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      __syncthreads();
    }
  };
  for (int s = 0; s < sweeps; s++) {
    p0.sweep(D, e);
    sync();
    p1.sweep(D, e);
    sync();
  }
}
template <class T>
__device__ __forceinline__ void cSweeps(const UDimArg& D, T* e, const T* rho, const uint8_t* fl, const T* inv,
                                        const uint16_t* lst, const uint32_t* cnt, T h2, int sweeps) {
  const bool wave = max(cnt[0], cnt[1]) <= (uint32_t)(kCWaveSlots<T> * 64);
  if (wave) {
    if (threadIdx.x < 64) cSweepLoop<T, kCWaveSlots<T>>(D, e, rho, fl, inv, lst, cnt, h2, sweeps, true);
    __syncthreads();
  } else {
    cSweepLoop<T, kCSlots>(D, e, rho, fl, inv, lst, cnt, h2, sweeps, false);
  }
}
// What a coarse call computes from the activity alone (couplings, inverse diagonals, parity
// lists, the activity of every grid) is the same in every cycle of a level solve: the first
// call of a solve stores that part of its LDS in a slot of device memory tagged with the solve,
// later calls copy it back (16-byte loads, all issued at once) instead of recomputing it.
template <class T>
struct alignas(16) CoarseLds {
  T e[kCoarseCells];
  T rho[kCoarseCells];
  // the cached image: from inv to the end
  T inv[kCoarseCells];
  uint8_t act[kCoarseCells];
  uint8_t fl[kCoarseCells];
  uint16_t lst[kCoarseCells];
  uint32_t cnt[kMaxPyr][2];
};
template <class T>
constexpr size_t kCoarseImage = sizeof(T) * kCoarseCells + (size_t)kCoarseCells * 4 + sizeof(uint32_t) * 2 * kMaxPyr;
static constexpr size_t kCoarseCacheBytes = (kCoarseImage<double> + 255) & ~(size_t)255;
static_assert(kCoarseImage<float> % 16 == 0 && kCoarseImage<double> % 16 == 0, "cached image in 16-byte units");
template <class T>
__device__ __forceinline__ void cCopyImage(const uint4* __restrict__ src, uint4* __restrict__ dst) {
  constexpr int N = (int)(kCoarseImage<T> / 16), R = (N + kCoarseThreads - 1) / kCoarseThreads;
  uint4 v[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = (int)threadIdx.x + r * kCoarseThreads;
    if (i < N) v[r] = src[i];
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = (int)threadIdx.x + r * kCoarseThreads;
    if (i < N) dst[i] = v[r];
  }
}
template <class T>
__global__ void __launch_bounds__(kCoarseThreads) k_svu_coarse(CoarseArgs ca, T* __restrict__ ue, const T* __restrict__ urho,
                                                               const uint8_t* __restrict__ uact,
                                                               const SolveCtl* __restrict__ ctl,
                                                               uint8_t* __restrict__ cache, uint32_t* __restrict__ tags) {
  __shared__ CoarseLds<T> S;
  static_assert(offsetof(CoarseLds<T>, inv) % 16 == 0 &&
                    offsetof(CoarseLds<T>, cnt) + sizeof(S.cnt) - offsetof(CoarseLds<T>, inv) == kCoarseImage<T>,
                "the cached fields are one contiguous image");
  T* e = S.e;
  T* rho = S.rho;
  T* inv = S.inv;
  uint8_t* act = S.act;
  uint8_t* fl = S.fl;
  uint16_t* lst = S.lst;
  uint32_t(*cnt)[2] = S.cnt;
  const uint32_t tagNow = tags[ca.slot];
  if (solveDone(ctl)) return;
  const bool hit = tagNow == ca.tag;
  uint4* img = (uint4*)(cache + (size_t)ca.slot * kCoarseCacheBytes);
  const UDimArg D0 = ca.D[0];
  const uint32_t n0 = (uint32_t)D0.w * D0.h * D0.d;
  if (hit) cCopyImage<T>(img, (uint4*)inv);
  for (uint32_t i = threadIdx.x; i < n0; i += blockDim.x) {
    e[i] = (T)0;
    rho[i] = urho[ca.goff + i];
    if (!hit) act[i] = uact[ca.goff + i];
  }
  if (!hit && threadIdx.x < 2 * kMaxPyr) cnt[threadIdx.x / 2][threadIdx.x % 2] = 0u;
  __syncthreads();
  for (int j = 0; j + 1 < ca.n; j++) {
    const UDimArg F = ca.D[j], Cg = ca.D[j + 1];
    const T h2j = (T)ca.h2[j];
    if (!hit) {
      cPrepare<T>(F, act, fl, inv, (T)ca.gm[j], lst, cnt[j]);
      __syncthreads();
    }
    cSweeps<T>(F, e, rho, fl, inv, lst, cnt[j], h2j, ca.sweeps);
    // residual restricted into the next grid (mean over the 2x2x2 children, inactive ones 0)
    const uint32_t nc = (uint32_t)Cg.w * Cg.h * Cg.d;
    const T rh2 = (T)1 / h2j;
    for (uint32_t i = threadIdx.x; i < nc; i += blockDim.x) {
      int X, Y, Z;
      cCoord(Cg, i, X, Y, Z);
      T acc = (T)0;
      bool any = false;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const uint32_t u = uAt(F, 2 * X + ((k >> 2) & 1), 2 * Y + ((k >> 1) & 1), 2 * Z + (k & 1));
        const uint32_t f = fl[u];
        const T r = rho[u] - (e[u] / inv[u] - cNeighbourSum<T>(F, u, f, e)) * rh2;
        const bool a = (f & 0x80u) != 0;
        any |= a;
        acc += a ? r : (T)0;
      }
      rho[Cg.off + i] = acc * (T).125;
      e[Cg.off + i] = (T)0;
      if (!hit) act[Cg.off + i] = any ? 1 : 0;
    }
    __syncthreads();
  }
  {
    const int l = ca.n - 1;
    const UDimArg Dl = ca.D[l];
    if (!hit) {
      cPrepare<T>(Dl, act, fl, inv, (T)ca.gm[l], lst, cnt[l]);
      __syncthreads();
    }
    cSweeps<T>(Dl, e, rho, fl, inv, lst, cnt[l], (T)ca.h2[l], ca.bottom);
  }
  for (int j = ca.n - 2; j >= 0; j--) {
    const UDimArg F = ca.D[j], Cg = ca.D[j + 1];
    const uint32_t nf = (uint32_t)F.w * F.h * F.d;
    for (uint32_t i = threadIdx.x; i < nf; i += blockDim.x) {
      const uint32_t u = F.off + i;
      if (!act[u]) continue;
      int x, y, z;
      cCoord(F, i, x, y, z);
      e[u] = e[u] + uProlong<T>(Cg, x, y, z, F.w, F.h, F.d, e, act);
    }
    __syncthreads();
    cSweeps<T>(F, e, rho, fl, inv, lst, cnt[j], (T)ca.h2[j], ca.sweeps);
  }
  for (uint32_t i = threadIdx.x; i < n0; i += blockDim.x)
    if (act[i]) ue[ca.goff + i] = e[i];
  if (!hit) {
    cCopyImage<T>((const uint4*)inv, img);
    if (threadIdx.x == 0) tags[ca.slot] = ca.tag;
  }
}

// ---------------------------------------------------------------- refinement loop of a level
// Residual after the update x' = x + e (first: x' = x, no correction yet) on the cells of
// `level`, block by block with the faces staged in LDS: writes x' to xout,
// rho = (R - shift) / h^2, and per workgroup the partial sums of R, R^2, R0, R0^2 and x' (R0:
// the residual of a zero level pressure). Same-level neighbours come from xin (+ e), coarser
// ones (already solved ghosts) from xc. The partials are combined by k_sv_decide in the next
// launch: a kernel boundary makes them visible without a device-scope fence per workgroup (on
// MI355X such a fence writes the L2 back, which cost this kernel 130 us per launch).
template <int BW, class TX, class T, class G>
__global__ void __launch_bounds__(Geo<BW>::WG) k_sv_residual(G g, int level, const TX* __restrict__ xc,
                                                           const TX* __restrict__ xin, TX* __restrict__ xout,
                                                           const T* __restrict__ e, int first,
                                                           T* __restrict__ rho, double* __restrict__ part, float alpha,
                                                           float h2, const SolveCtl* __restrict__ ctl) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW, WG = Geo<BW>::WG;
  constexpr int FACE = Geo<BW>::FACE, FPT = (FACE + BV - 1) / BV;
  __shared__ TX Xs[BPW][A3];
  __shared__ uint8_t Gs[BPW][A3];  // 1: coarser-level ghost
  __shared__ double red[5][WG / 64];
  if (!first && solveDone(ctl)) return;
  const double shift = first ? 0.0 : ctl->shift;
  const uint32_t total = slotTotal(g, level);
  const int sub = threadIdx.x / BV, t = threadIdx.x % BV;
  const int a = apronIndexOfLocal<BW>(t);
  TX* X = Xs[sub];
  uint8_t* Gh = Gs[sub];
  double s[5] = {0, 0, 0, 0, 0};
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    const uint32_t slot = base + sub;
    const bool valid = slot < total;
    const uint32_t b = g.offs[level] + (valid ? slot : total - 1);
    const uint32_t c = b * BV + t;
    __syncthreads();
    uint32_t nb[FPT], sh[FPT];
    int fa[FPT];
#pragma unroll
    for (int j = 0; j < FPT; j++) {
      const int fc = t + j * BV < FACE ? t + j * BV : FACE - 1;
      int ai, aj, ak;
      faceCell<BW>(fc, ai, aj, ak);
      fa[j] = (ai * A + aj) * A + ak;
      nb[j] = borderApronCellShift<BW>(g, b, ai, aj, ak, sh[j]);
    }
    const TX xv = xin[c] + (first ? (TX)0 : (TX)e[c]);
    const float dv = (float)g.div[c];
    TX fv[FPT];
#pragma unroll
    for (int j = 0; j < FPT; j++) {
      const bool co = sh[j] != 0 && sh[j] != ND_OUT;
      fv[j] = co ? xc[nb[j]] : xin[nb[j]] + (first ? (TX)0 : (TX)e[nb[j]]);
    }
    __builtin_amdgcn_sched_barrier(0);
    X[a] = xv;
    Gh[a] = 0;
#pragma unroll
    for (int j = 0; j < FPT; j++) {
      if (t + j * BV >= FACE) continue;
      X[fa[j]] = fv[j];
      Gh[fa[j]] = sh[j] != 0 && sh[j] != ND_OUT;
    }
    __syncthreads();
    if (!valid) continue;
    if (!first) xout[c] = xv;
    const int n[6] = {a - A * A, a + A * A, a - A, a + A, a - 1, a + 1};
    TX sum = 0, ghosts = 0;
#pragma unroll
    for (int d = 0; d < 6; d++) {
      sum += X[n[d]];
      ghosts += Gh[n[d]] ? X[n[d]] : (TX)0;
    }
    const TX ad = (TX)alpha * (TX)dv;
    const TX R = sum - (TX)6 * xv - ad;
    const TX R0 = ghosts - ad;
    rho[c] = (T)((R - (TX)shift) / (TX)h2);
    s[0] += (double)R;
    s[1] += (double)R * (double)R;
    s[2] += (double)R0;
    s[3] += (double)R0 * (double)R0;
    s[4] += (double)xv;
  }
  // workgroup sums in a fixed order: wavefront, then across wavefronts
#pragma unroll
  for (int k = 0; k < 5; k++) {
    double v = s[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double v = 0;
    for (int w = 0; w < WG / 64; w++) v += red[threadIdx.x][w];
    part[5 * blockIdx.x + threadIdx.x] = v;
  }
}

// One workgroup after k_sv_residual (nparts of its workgroups): sums the partials in workgroup
// order (the result does not depend on scheduling) and decides in ctl whether the level is done
// (|R| <= tol |R(0)|, the cycle limit, or three cycles that did not reduce |R| by 0.7 — the
// working precision is exhausted); if not, the next cycle runs. A level that covers the domain
// (loads == capacity == full) is pure Neumann: its residual's mean is removed. The control block
// is mirrored into host memory (`mirror`), which the host reads after a synchronisation.
__global__ void __launch_bounds__(256) k_sv_decide(DGBase g, int level, int bv, int first, int nparts,
                                                   const double* __restrict__ part, float h2, double tol,
                                                   int maxCycles, SolveCtl* __restrict__ ctl,
                                                   SolveCtl* __restrict__ mirror) {
  __shared__ double red[5][4];
  if (!first && solveDone(ctl)) return;
  double S[5] = {0, 0, 0, 0, 0};
  for (int w = threadIdx.x; w < nparts; w += 256)
#pragma unroll
    for (int k = 0; k < 5; k++) S[k] += part[5 * w + k];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    double v = S[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double T5[5];
  for (int k = 0; k < 5; k++) T5[k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
  const uint32_t loads = g.ctl[CTL_LOADS + level];
  const bool singular = loads == g.maxL[level] && g.maxL[level] == g.fullL[level];
  const double N = (double)loads * bv;
  double rr, r0c;
  const double m = N > 0 ? T5[0] / N : 0.0;
  if (singular) {
    const double m0 = N > 0 ? T5[2] / N : 0.0;
    rr = sqrt(fmax(T5[1] - N * m * m, 0.0));
    r0c = sqrt(fmax(T5[3] - N * m0 * m0, 0.0));
  } else {
    rr = sqrt(T5[1]);
    r0c = sqrt(T5[3]);
  }
  SolveCtl c = *ctl;
  if (first) {
    c.r0 = r0c;
    c.cycles = 0;
    c.stall = 0;
    c.prev = 0;
    c.shift = 0;
    c.singular = singular ? 1u : 0u;
  }
  c.mean = N > 0 ? T5[4] / N : 0.0;
  c.rr = rr;
  bool done = rr <= tol * c.r0 || (int)c.cycles >= maxCycles;
  if (!done) {
    if (c.cycles > 0 && rr > 0.7 * c.prev) {
      if (++c.stall >= 3) done = true;
    } else {
      c.stall = 0;
    }
  }
  c.prev = rr;
  c.delta = 0;
  if (!done) {
    // the mean of R (the right-hand side's inconsistency) stays the same from cycle to cycle:
    // the residual kernel subtracts the last one, the level's kernels take what is left
    if (singular && m != c.shift) {
      c.delta = (m - c.shift) / (double)h2;
      c.shift = m;
    }
    c.cycles++;
  }
  c.done = done ? 1u : 0u;
  *ctl = c;
  *mirror = c;
}

// End of a level's loop: x (the final iterate, in xsrc) minus its mean on a singular level,
// into the canonical iterate xdst and the pressure field.
template <int BW, class TX, class G>
__global__ void __launch_bounds__(256) k_sv_finalize(G g, int level, const TX* __restrict__ xsrc, TX* __restrict__ xdst,
                                                     const SolveCtl* __restrict__ ctl) {
  constexpr int BV = Geo<BW>::BV;
  const TX m = ctl->singular ? (TX)ctl->mean : (TX)0;
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t c = g.offs[level] * BV + i;
    const TX v = xsrc[c] - m;
    xdst[c] = v;
    g.pre[c] = (float)v;
  }
}

// x = 0 on the cells of a level.
template <int BW, class T>
__global__ void __launch_bounds__(256) k_sv_zero(DGBase g, int level, T* __restrict__ x) {
  constexpr int BV = Geo<BW>::BV;
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) x[g.offs[level] * BV + i] = (T)0;
}

// Warm start: x = the pressure of the last projection on every allocated block that kept its
// place in this step's adaptation (the others are prolonged from their parent level once it is
// solved, k_sv_prolong_x with movedOnly).
template <int BW, class TX, class G>
__global__ void __launch_bounds__(256) k_sv_warm(G g, TX* __restrict__ x) {
  constexpr int BV = Geo<BW>::BV;
  for (int l = 0; l < g.L; l++) {
    const uint32_t n = g.ctl[CTL_LOADS + l] * BV;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
      const uint32_t c = g.offs[l] * BV + i;
      if (!(g.flg[c / BV] & MOVED)) x[c] = (TX)(float)g.pre[c];
    }
  }
}

// Initial guess of a level from its parent level: the 27/9/3/1 stencil of k_prolongate on x
// (movedOnly: only the blocks moved or created in this step's adaptation).
template <int BW, class TX>
__global__ void __launch_bounds__(256) k_sv_prolong_x(DGBase g, int level, TX* __restrict__ x, int movedOnly) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB;
  const uint32_t n = g.ctl[CTL_LOADS + level] * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t b = g.offs[level] + i / BV;
    const int t = (int)(i % BV);
    const uint32_t ps = g.par[b];
    if (ps == NFU || (movedOnly && !(g.flg[b] & MOVED))) continue;
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

}  // namespace dcg

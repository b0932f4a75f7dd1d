// dcgrid_fluid.hpp — gfx950 fluid kernels of the DCGrid back-end: cross-level semi-Lagrangian
// advection (src/dcgrid/dcgrid_fluid.cu:7-144), vorticity (:146-172), divergence (:174-230),
// gradient subtraction (:232-259), Jacobi half-sweeps and prolongation
// (src/dcgrid/dcgrid_multigrid_solver.cu), restriction (src/dcgrid/dcgrid_structure.cu:188-222),
// the residual of debugStats (:224-251) and value propagation into new blocks
// (src/dcgrid/dcgrid_adaptation.cu:92-143).
//
// Stencil kernels run one workgroup per block (BW=8: 512 threads) or per four blocks (BW=4:
// 4x64 threads), one thread per cell; the advection (always) and the divergence, gradient and
// feature kernels (large budgets) run BW-8 blocks with 256 threads of two cells each. The block's
// cells are contiguous in HBM (coalesced) and are staged with the six face layers of the apron
// in LDS, so each neighbour read is an LDS read; face cells are fetched through the block's
// neighbour table.
#pragma once

#include "dcgrid_dev.hpp"

namespace dcg {

// Block handled by this thread's sub-group and the thread's cell inside it.
template <int BW>
struct BlockCtx {
  uint32_t b;  // global block index (may be >= B for tail threads)
  int t;       // local cell (Morton) index
  int sub;     // sub-group (block slot) inside the workgroup
  int level;   // level of block b (slot walks only)
  bool valid;
};
template <int BW>
__device__ __forceinline__ BlockCtx<BW> blockCtx(uint32_t firstBlock, uint32_t count) {
  constexpr int BV = Geo<BW>::BV, BPW = Geo<BW>::BPW;
  BlockCtx<BW> c;
  c.sub = threadIdx.x / BV;
  c.t = threadIdx.x % BV;
  const uint32_t slot = xcdRemap(blockIdx.x, gridDim.x) * BPW + c.sub;
  c.b = firstBlock + slot;
  c.valid = slot < count;
  return c;
}

// Grid-stride walk over allocated blocks. The allocated blocks of level l are the prefix
// [offs[l], offs[l] + loads[l]) of its slice (blocks are never freed), so a kernel launched with
// a fixed number of workgroups visits exactly the allocated blocks of one level (level >= 0) or
// of all levels in level order (level < 0), BPW blocks per workgroup and iteration, with the
// counts read from the device control words: no workgroup is spent on free blocks and the host
// needs no block counts.
__device__ __forceinline__ uint32_t slotTotal(const DGBase& g, int level) {
  if (level >= 0) return g.ctl[CTL_LOADS + level];
  uint32_t n = 0;
  for (int l = 0; l < g.L; l++) n += g.ctl[CTL_LOADS + l];
  return n;
}
// Block of slot s, and its level (the level whose prefix holds the slot: every block of a
// walk is allocated, so its level needs no table read).
__device__ __forceinline__ uint32_t slotBlock(const DGBase& g, int level, uint32_t s, int& blockLevel) {
  blockLevel = level;
  if (level >= 0) return g.offs[level] + s;
  for (int l = 0; l < g.L; l++) {
    const uint32_t n = g.ctl[CTL_LOADS + l];
    if (s < n) {
      blockLevel = l;
      return g.offs[l] + s;
    }
    s -= n;
  }
  blockLevel = 0;
  return 0;
}
template <int BW>
__device__ __forceinline__ uint32_t firstSlot() {
  return xcdRemap(blockIdx.x, gridDim.x) * Geo<BW>::BPW;
}
template <int BW>
__device__ __forceinline__ uint32_t slotStride() {
  return gridDim.x * Geo<BW>::BPW;
}
template <int BW>
__device__ __forceinline__ BlockCtx<BW> slotCtx(const DGBase& g, int level, uint32_t base, uint32_t total) {
  constexpr int BV = Geo<BW>::BV;
  BlockCtx<BW> c;
  c.sub = threadIdx.x / BV;
  c.t = threadIdx.x % BV;
  const uint32_t slot = base + c.sub;
  c.valid = slot < total;
  c.level = 0;
  c.b = c.valid ? slotBlock(g, level, slot, c.level) : 0;
  return c;
}

template <int BW>
__device__ __forceinline__ int apronIndexOfLocal(int t) {
  constexpr int A = Geo<BW>::A, BB = Geo<BW>::BB;
  int x, y, z;
  locPos<BB>((uint32_t)t, x, y, z);
  return ((x + 1) * A + (y + 1)) * A + (z + 1);
}

// Stage one scalar field of a block (interior + 6 face layers) into tile[A^3] (widened to FP32).
template <int BW, class SF>
__device__ __forceinline__ void stageScalar(const DGBase& g, const SF* __restrict__ f, float* tile, uint32_t b, int t) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, FACE = Geo<BW>::FACE;
  tile[apronIndexOfLocal<BW>(t)] = f[b * BV + t];
  for (int fc = t; fc < FACE; fc += BV) {
    int ai, aj, ak;
    faceCell<BW>(fc, ai, aj, ak);
    tile[(ai * A + aj) * A + ak] = f[borderApronCell<BW>(g, b, ai, aj, ak)];
  }
}
// stageScalar on a dense level (l >= S): the faces come from denseApronCell, the block origin
// from the block index.
template <int BW, class SF>
__device__ __forceinline__ void stageScalarDense(const DGBase& g, int l, const SF* __restrict__ f, float* tile, uint32_t b,
                                                 int t) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, FACE = Geo<BW>::FACE;
  tile[apronIndexOfLocal<BW>(t)] = f[b * BV + t];
  const int3 p0 = denseOrigin<BW>(g, l, b);
  for (int fc = t; fc < FACE; fc += BV) {
    int ai, aj, ak;
    faceCell<BW>(fc, ai, aj, ak);
    tile[(ai * A + aj) * A + ak] = f[denseApronCell<BW>(g, l, b, p0, ai, aj, ak)];
  }
}

// Fluidity of the six apron neighbours of tile entry a, from the block's staged fluidity tile
// (1 when the field is known to be 1).
template <int BW, bool FLUIDITY>
__device__ __forceinline__ void tileFluidity(const float* ft, int a, float w[6]) {
  constexpr int A = Geo<BW>::A;
  if (!FLUIDITY) {
#pragma unroll
    for (int k = 0; k < 6; k++) w[k] = 1.f;
    return;
  }
  w[0] = ft[a - A * A];  // l
  w[1] = ft[a + A * A];  // r
  w[2] = ft[a - A];      // d
  w[3] = ft[a + A];      // u
  w[4] = ft[a - 1];      // b
  w[5] = ft[a + 1];      // f
}

// Stencil kernels with a CPT parameter: a thread takes CPT cells of its block, so a block is
// handled by BV / CPT threads and CPT times as many blocks are resident per CU at the same thread
// count, each thread keeping CPT cells' loads in flight. That pays when workgroups walk many
// blocks; with about one block per workgroup the longer thread chain costs more (CPT 1).
template <int BW, int CPT>
struct CptGeo {
  static_assert(CPT == 1 || (BW == 8 && CPT == 2), "two cells per thread for BW 8 only");
  static constexpr int WG = Geo<BW>::WG / CPT;
};

// ------------------------------------------------------------------ Jacobi half-sweep
// k_dcgrid_jacobi / k_dcgrid_jacobi_inv (dcgrid_multigrid_solver.cu:5-41):
// out = (sum of 6 apron neighbours of `in` - (scale*dx)^2 * div) / 6 on every cell of every
// block of the level. Neighbours on coarser levels are read from the same array `in`.
// NB blocks per sub-group and iteration: their loads are issued together, so a thread keeps NB
// blocks' worth of reads in flight (the kernel is bound by latency times bytes in flight on the
// large levels, where every workgroup walks many blocks).
// DENSE: the level is dense (l >= S), its faces come from denseApronCell (no descriptor loads).
template <int BW, int NB, bool DENSE, class G>
__global__ void __launch_bounds__(Geo<BW>::WG) k_jacobi(G g, int level, const typename G::Scalar* __restrict__ in,
                                                         typename G::Scalar* __restrict__ out, float alpha) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW, FACE = Geo<BW>::FACE;
  constexpr int FPT = (FACE + BV - 1) / BV;  // face cells per thread
  __shared__ float tiles[NB][BPW][A3];
  const uint32_t total = slotTotal(g, level);
  const uint32_t stride = slotStride<BW>();
  const uint32_t off = g.offs[level];
  const int sub = threadIdx.x / BV, t = threadIdx.x % BV;
  const int ac = apronIndexOfLocal<BW>(t);
  for (uint32_t base = firstSlot<BW>(); base < total; base += NB * stride) {
    __syncthreads();  // the previous blocks' tiles are no longer read
    // Every load of the iteration is issued before the first is used, with no branch between
    // them: slots past the level's end read its last block (in bounds) and store nothing.
    uint32_t b[NB];
    bool valid[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
      const uint32_t slot = base + k * stride + sub;
      valid[k] = slot < total;
      b[k] = off + (valid[k] ? slot : total - 1);
    }
    // face cells: on a dense level from the block index (integer arithmetic, before any load),
    // else through the neighbour descriptor (its load issued with the block's)
    uint32_t src[NB][FPT];
    auto faceSources = [&]() {
#pragma unroll
      for (int k = 0; k < NB; k++) {
        const int3 p0 = DENSE ? denseOrigin<BW>(g, level, b[k]) : make_int3(0, 0, 0);
#pragma unroll
        for (int j = 0; j < FPT; j++) {
          const int fc = t + j * BV < FACE ? t + j * BV : FACE - 1;  // lanes without one repeat the last
          int ai, aj, ak;
          faceCell<BW>(fc, ai, aj, ak);
          src[k][j] =
              DENSE ? denseApronCell<BW>(g, level, b[k], p0, ai, aj, ak) : borderApronCell<BW>(g, b[k], ai, aj, ak);
        }
      }
    };
    if (DENSE) faceSources();
    float dv[NB], pc[NB], fv[NB][FPT];
#pragma unroll
    for (int k = 0; k < NB; k++) {
      dv[k] = (float)g.div[b[k] * BV + t];
      pc[k] = (float)in[b[k] * BV + t];
    }
    if (!DENSE) faceSources();
#pragma unroll
    for (int k = 0; k < NB; k++)
#pragma unroll
      for (int j = 0; j < FPT; j++) fv[k][j] = (float)in[src[k][j]];
    // keep the LDS stores below the loads (a store hoisted above would wait for them)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < NB; k++) {
      float* tile = tiles[k][sub];
      tile[ac] = pc[k];
#pragma unroll
      for (int j = 0; j < FPT; j++) {  // lanes without a face cell store the last one's value again
        const int fc = t + j * BV < FACE ? t + j * BV : FACE - 1;
        int ai, aj, ak;
        faceCell<BW>(fc, ai, aj, ak);
        tile[(ai * A + aj) * A + ak] = fv[k][j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NB; k++) {
      if (!valid[k]) continue;
      const float* tile = tiles[k][sub];
      out[b[k] * BV + t] =
          (tile[ac - A * A] + tile[ac + A * A] + tile[ac - A] + tile[ac + A] + tile[ac - 1] + tile[ac + 1] - alpha * dv[k]) /
          6.f;
    }
  }
}

// ------------------------------------------------------------------ divergence
// k_dcgrid_calc_divergence (dcgrid_fluid.cu:174-230): zero p, p~ and div on all cells; stage the
// velocity with boundary conditions on the apron faces; central divergence on leaf cells.
template <int BW, bool FLUIDITY, int CPT, class G>
__global__ void __launch_bounds__((CptGeo<BW, CPT>::WG)) k_divergence(G g, dcg_params P) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW, FACE = Geo<BW>::FACE,
                SBB = Geo<BW>::SBB, N = BV / CPT;
  __shared__ float3 tiles[BPW][A3];
  __shared__ float ftiles[FLUIDITY ? BPW : 1][FLUIDITY ? A3 : 1];
  const uint32_t total = slotTotal(g, -1);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();  // the previous block's tiles are no longer read
    const BlockCtx<BW> c = slotCtx<BW>(g, -1, base, total);  // c.t: the thread's first cell
    float3* tile = tiles[c.sub];
    float* ft = FLUIDITY ? ftiles[c.sub] : nullptr;
    const int scale = 1 << c.level;
    bool leaf[CPT];
#pragma unroll
    for (int j = 0; j < CPT; j++) leaf[j] = false;
    if (c.valid) {
      const int3 p0 = g.pos[c.b];
#pragma unroll
      for (int j = 0; j < CPT; j++) {
        const int t = c.t + j * N;
        const uint32_t cell = c.b * BV + t;
        leaf[j] = g.chi[8 * c.b + (t >> SBB)] == NFU;
        g.pre[cell] = 0.f;
        g.tpre[cell] = 0.f;
        g.div[cell] = 0.f;
        const int ac = apronIndexOfLocal<BW>(t);
        tile[ac] = g.vel[cell];
        if (FLUIDITY) ft[ac] = g.flu[cell];
      }
      for (int fc = c.t; fc < FACE; fc += N) {
        int ai, aj, ak;
        faceCell<BW>(fc, ai, aj, ak);
        const uint32_t src = borderApronCell<BW>(g, c.b, ai, aj, ak);
        const int a = (ai * A + aj) * A + ak;
        tile[a] = velocityBC(P, (float3)g.vel[src], p0.x + ai - 1, p0.y + aj - 1, p0.z + ak - 1, scale);
        if (FLUIDITY) ft[a] = g.flu[src];
      }
    }
    __syncthreads();
    const float alpha = .5f * P.rdx / (float)scale;
#pragma unroll
    for (int j = 0; j < CPT; j++) {
      if (!leaf[j]) continue;
      const int t = c.t + j * N;
      const int a = apronIndexOfLocal<BW>(t);
      float w[6];
      tileFluidity<BW, FLUIDITY>(ft, a, w);
      const float3 vl = tile[a - A * A], vr = tile[a + A * A], vd = tile[a - A], vu = tile[a + A], vb = tile[a - 1],
                   vf = tile[a + 1];
      g.div[c.b * BV + t] =
          alpha * (w[1] * vr.x - w[0] * vl.x + w[3] * vu.y - w[2] * vd.y + w[5] * vf.z - w[4] * vb.z);
    }
  }
}

// ------------------------------------------------------------------ gradient subtraction
// k_dcgrid_apply_pressure (dcgrid_fluid.cu:232-259), leaf cells only.
template <int BW, bool FLUIDITY, int CPT, class G>
__global__ void __launch_bounds__((CptGeo<BW, CPT>::WG)) k_apply_pressure(G g, dcg_params P) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW, FACE = Geo<BW>::FACE,
                SBB = Geo<BW>::SBB, N = BV / CPT;
  __shared__ float tiles[BPW][A3];
  __shared__ float ftiles[FLUIDITY ? BPW : 1][FLUIDITY ? A3 : 1];
  const uint32_t total = slotTotal(g, -1);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();  // the previous block's tiles are no longer read
    const BlockCtx<BW> c = slotCtx<BW>(g, -1, base, total);  // c.t: the thread's first cell
    float* tile = tiles[c.sub];
    float* ft = FLUIDITY ? ftiles[c.sub] : nullptr;
    bool leaf[CPT];
    float3 v[CPT];
#pragma unroll
    for (int j = 0; j < CPT; j++) {
      leaf[j] = false;
      v[j] = make_float3(0.f, 0.f, 0.f);
    }
    if (c.valid) {
#pragma unroll
      for (int j = 0; j < CPT; j++) {
        const int t = c.t + j * N;
        const uint32_t cell = c.b * BV + t;
        leaf[j] = g.chi[8 * c.b + (t >> SBB)] == NFU;
        if (leaf[j]) v[j] = g.vel[cell];  // only this thread reads and writes its cell's velocity
        const int ac = apronIndexOfLocal<BW>(t);
        tile[ac] = g.pre[cell];
        if (FLUIDITY) ft[ac] = g.flu[cell];
      }
      for (int fc = c.t; fc < FACE; fc += N) {
        int ai, aj, ak;
        faceCell<BW>(fc, ai, aj, ak);
        const uint32_t src = borderApronCell<BW>(g, c.b, ai, aj, ak);
        const int a = (ai * A + aj) * A + ak;
        tile[a] = g.pre[src];
        if (FLUIDITY) ft[a] = g.flu[src];
      }
    }
    __syncthreads();
    const float alpha = .5f * P.rdx / (float)(1 << c.level);
#pragma unroll
    for (int j = 0; j < CPT; j++) {
      if (!leaf[j]) continue;
      const int t = c.t + j * N;
      const int a = apronIndexOfLocal<BW>(t);
      float w[6];
      tileFluidity<BW, FLUIDITY>(ft, a, w);
      const float p = tile[a];
      float3 u = v[j];
      u.x = u.x - alpha * (w[1] * (tile[a + A * A] - p) + w[0] * (p - tile[a - A * A]));
      u.y = u.y - alpha * (w[3] * (tile[a + A] - p) + w[2] * (p - tile[a - A]));
      u.z = u.z - alpha * (w[5] * (tile[a + 1] - p) + w[4] * (p - tile[a - 1]));
      g.vel[c.b * BV + t] = u;
    }
  }
}

// ------------------------------------------------------------------ vorticity
// k_dcgrid_calc_vorticity (dcgrid_fluid.cu:146-172): curl of fluidity-weighted velocity on the
// apron, every cell of every allocated block (no boundary conditions).
template <int BW, bool FLUIDITY, class G>
__global__ void __launch_bounds__(Geo<BW>::WG) k_vorticity(G g, dcg_params P) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW, FACE = Geo<BW>::FACE;
  __shared__ float3 tiles[BPW][A3];
  __shared__ float ftiles[FLUIDITY ? BPW : 1][FLUIDITY ? A3 : 1];
  const uint32_t total = slotTotal(g, -1);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();  // the previous block's tiles are no longer read
    const BlockCtx<BW> c = slotCtx<BW>(g, -1, base, total);
    float3* tile = tiles[c.sub];
    float* ft = FLUIDITY ? ftiles[c.sub] : nullptr;
    const uint32_t cell = c.b * BV + c.t;
    if (c.valid) {
      const int ac = apronIndexOfLocal<BW>(c.t);
      tile[ac] = g.vel[cell];
      if (FLUIDITY) ft[ac] = g.flu[cell];
      for (int fc = c.t; fc < FACE; fc += BV) {
        int ai, aj, ak;
        faceCell<BW>(fc, ai, aj, ak);
        const uint32_t src = borderApronCell<BW>(g, c.b, ai, aj, ak);
        const int a = (ai * A + aj) * A + ak;
        tile[a] = g.vel[src];
        if (FLUIDITY) ft[a] = g.flu[src];
      }
    }
    __syncthreads();
    if (!c.valid) continue;
    const float alpha = .5f * P.rdx / (float)(1 << c.level);
    const int a = apronIndexOfLocal<BW>(c.t);
    float w[6];
    tileFluidity<BW, FLUIDITY>(ft, a, w);
    const float3 vl = tile[a - A * A], vr = tile[a + A * A], vd = tile[a - A], vu = tile[a + A], vb = tile[a - 1],
                 vf = tile[a + 1];
    const float wl = w[0], wr = w[1], wd = w[2], wu = w[3], wb = w[4], wf = w[5];
    g.vort[cell] = make_float3(alpha * ((wu * vu.z - wd * vd.z) - (wf * vf.y - wb * vb.y)),
                               alpha * ((wf * vf.x - wb * vb.x) - (wr * vr.z - wl * vl.z)),
                               alpha * ((wr * vr.y - wl * vl.y) - (wu * vu.x - wd * vd.x)));
  }
}

// ------------------------------------------------------------------ restriction
// accumulate<T> (dcgrid_structure.cu:188-222): a parent cell becomes the mean of the 8 child
// cells of one restriction group (the lowest Morton triple of the child block). One thread per
// group; reads are 8 consecutive cells.
template <class T>
__device__ __forceinline__ T zeroOf();
template <>
__device__ __forceinline__ float zeroOf<float>() { return 0.f; }
template <>
__device__ __forceinline__ float3 zeroOf<float3>() { return make_float3(0.f, 0.f, 0.f); }
__device__ __forceinline__ float addv(float a, float b) { return a + b; }
__device__ __forceinline__ float3 addv(float3 a, float3 b) { return f3add(a, b); }
__device__ __forceinline__ float mulv(float a, float s) { return a * s; }
__device__ __forceinline__ float3 mulv(float3 a, float s) { return f3mul(a, s); }
// FP32 arithmetic type of a stored element type.
template <class T>
struct Wide {
  using type = float;
};
template <>
struct Wide<float3> {
  using type = float3;
};
template <class S>
struct Wide<StoreV3<S>> {
  using type = float3;
};

template <int BW, class T>
__global__ void __launch_bounds__(256) k_restrict(DGBase g, int level, T* __restrict__ ch) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB, GPB = BV / 8;
  const uint32_t groups = g.ctl[CTL_LOADS + level] * GPB;  // allocated prefix of the level slice
  for (uint32_t gi = blockIdx.x * 256 + threadIdx.x; gi < groups; gi += gridDim.x * 256) {
    const uint32_t b = g.offs[level] + gi / GPB;
    const uint32_t ps = g.par[b];
    if (ps == NFU) continue;
    const uint32_t grp = gi % GPB;
    const uint32_t parentCell = (ps / 8) * BV + ((ps % 8) << SBB) + grp;
    using W = typename Wide<T>::type;
    const T* src = ch + (size_t)b * BV + 8 * grp;
    W acc = zeroOf<W>();
#pragma unroll
    for (int i = 0; i < 8; i++) acc = addv(acc, (W)src[i]);
    ch[parentCell] = mulv(acc, .125f);
  }
}

// Restriction of one level of up to three fields at once (f1, f2 may be null): the same
// groups as k_restrict, each field's sum in k_restrict's order.
template <int BW, class T>
__global__ void __launch_bounds__(256) k_restrict3(DGBase g, int level, T* __restrict__ f0, T* __restrict__ f1,
                                                   T* __restrict__ f2) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB, GPB = BV / 8;
  const uint32_t groups = g.ctl[CTL_LOADS + level] * GPB;
  for (uint32_t gi = blockIdx.x * 256 + threadIdx.x; gi < groups; gi += gridDim.x * 256) {
    const uint32_t b = g.offs[level] + gi / GPB;
    const uint32_t ps = g.par[b];
    if (ps == NFU) continue;
    const uint32_t grp = gi % GPB;
    const uint32_t parentCell = (ps / 8) * BV + ((ps % 8) << SBB) + grp;
    const size_t c0 = (size_t)b * BV + 8 * grp;
    T* fs[3] = {f0, f1, f2};
#pragma unroll
    for (int f = 0; f < 3; f++) {
      if (!fs[f]) continue;
      using W = typename Wide<T>::type;
      W acc = zeroOf<W>();
#pragma unroll
      for (int i = 0; i < 8; i++) acc = addv(acc, (W)fs[f][c0 + i]);
      fs[f][parentCell] = mulv(acc, .125f);
    }
  }
}

// ------------------------------------------------------------------ parent stencil
// Cells of the parent's apron around child cell (x,y,z) used by propagate and prolongate
// (dcgrid_adaptation.cu:104-126, dcgrid_multigrid_solver.cu:51-63); order 000,001,010,100,011,
// 101,110,111 of the reference.
template <int BW>
__device__ __forceinline__ void parentStencil(const DGBase& g, uint32_t ps, int x, int y, int z, uint32_t q[8]) {
  const uint32_t pb = ps / 8;
  const int ai = 1 + (x / 2) % BW, aj = 1 + (y / 2) % BW, ak = 1 + (z / 2) % BW;
  const int si = x % 2 ? 1 : -1, sj = y % 2 ? 1 : -1, sk = z % 2 ? 1 : -1;
  q[0] = apronCell<BW>(g, pb, ai, aj, ak);
  q[1] = apronCell<BW>(g, pb, ai, aj, ak + sk);
  q[2] = apronCell<BW>(g, pb, ai, aj + sj, ak);
  q[3] = apronCell<BW>(g, pb, ai + si, aj, ak);
  q[4] = apronCell<BW>(g, pb, ai, aj + sj, ak + sk);
  q[5] = apronCell<BW>(g, pb, ai + si, aj, ak + sk);
  q[6] = apronCell<BW>(g, pb, ai + si, aj + sj, ak);
  q[7] = apronCell<BW>(g, pb, ai + si, aj + sj, ak + sk);
}

// k_dcgrid_prolongate (dcgrid_multigrid_solver.cu:43-76). The parent cells a child block reads
// are the (BW/2 + 2)^3 region of the parent's apron around the child's octant: the workgroup
// stages that region once (one apron lookup per entry) and every cell takes its eight
// parentStencil values from LDS (the same cells in the same order).
template <int BW, class G>
__global__ void __launch_bounds__(Geo<BW>::WG) k_prolongate(G g, int level) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB, BPW = Geo<BW>::BPW, H = BW / 2 + 2;
  __shared__ float tiles[BPW][H * H * H];
  const uint32_t total = slotTotal(g, level);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();  // the previous block's region is no longer read
    const BlockCtx<BW> c = slotCtx<BW>(g, level, base, total);
    const uint32_t ps = c.valid ? g.par[c.b] : NFU;
    float* tile = tiles[c.sub];
    int3 p0 = make_int3(0, 0, 0);
    int ox = 0, oy = 0, oz = 0;  // parent apron coordinates of the region's first entry
    if (ps != NFU) {
      p0 = g.pos[c.b];
      ox = (p0.x / 2) % BW;
      oy = (p0.y / 2) % BW;
      oz = (p0.z / 2) % BW;
      if (c.t < H * H * H) {
        const int i = c.t / (H * H), j = (c.t / H) % H, k = c.t % H;
        tile[c.t] = g.pre[apronCell<BW>(g, ps / 8, ox + i, oy + j, oz + k)];
      }
    }
    __syncthreads();
    if (ps == NFU) continue;
    int lx, ly, lz;
    locPos<BB>((uint32_t)c.t, lx, ly, lz);
    const int x = p0.x | lx, y = p0.y | ly, z = p0.z | lz;
    // parentStencil's apron coordinates, relative to the staged region
    const int ti = 1 + (x / 2) % BW - ox, tj = 1 + (y / 2) % BW - oy, tk = 1 + (z / 2) % BW - oz;
    const int si = x % 2 ? 1 : -1, sj = y % 2 ? 1 : -1, sk = z % 2 ? 1 : -1;
    auto at = [&](int i, int j, int k) { return tile[(i * H + j) * H + k]; };
    const float q0 = at(ti, tj, tk), q1 = at(ti, tj, tk + sk), q2 = at(ti, tj + sj, tk), q3 = at(ti + si, tj, tk);
    const float q4 = at(ti, tj + sj, tk + sk), q5 = at(ti + si, tj, tk + sk), q6 = at(ti + si, tj + sj, tk);
    const float q7 = at(ti + si, tj + sj, tk + sk);
    g.pre[c.b * BV + c.t] = (27.f * q0 + 9.f * (q1 + q2 + q3) + 3.f * (q4 + q5 + q6) + q7) / 64.f;
  }
}

// k_dcgrid_propagate_values (dcgrid_adaptation.cu:92-143) for the blocks of `level` touched by
// this adaptation: the moved blocks (list) and the new children (a contiguous index range).
template <int BW, bool SRC, class G>
__global__ void __launch_bounds__(Geo<BW>::WG) k_propagate(G g, dcg_params P, int level,
                                                            const uint32_t* __restrict__ moved) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB, BPW = Geo<BW>::BPW;
  const int sub = threadIdx.x / BV, t = threadIdx.x % BV;
  const uint32_t nm = g.ctl[CTL_MOVES + level];
  const uint32_t nc = level + 1 < g.L ? g.ctl[CTL_REFINES + level + 1] : 0;
  for (uint32_t e = blockIdx.x * BPW + sub; e < nm + nc; e += gridDim.x * BPW) {
    const uint32_t b = e < nm ? moved[g.offs[level] + e] : g.offs[level] + g.ctl[CTL_LOADS_BEFORE + level] + (e - nm);
    if (g.lev[b] != level) continue;
    const uint32_t ps = g.par[b];
    if (ps == NFU) continue;
    const int3 p0 = g.pos[b];
    int lx, ly, lz;
    locPos<BB>((uint32_t)t, lx, ly, lz);
    const int x = p0.x | lx, y = p0.y | ly, z = p0.z | lz;
    uint32_t q[8];
    parentStencil<BW>(g, ps, x, y, z, q);
    const float w27 = 27.f / 64.f, w9 = 9.f / 64.f, w3 = 3.f / 64.f, w1 = 1.f / 64.f;
    const uint32_t cell = b * BV + t;
    const auto* d = g.den;
    g.den[cell] = w27 * d[q[0]] + w9 * (d[q[1]] + d[q[2]] + d[q[3]]) + w3 * (d[q[4]] + d[q[5]] + d[q[6]]) + w1 * d[q[7]];
    if (SRC) {
      const auto* tt = g.tem;
      g.tem[cell] =
          w27 * tt[q[0]] + w9 * (tt[q[1]] + tt[q[2]] + tt[q[3]]) + w3 * (tt[q[4]] + tt[q[5]] + tt[q[6]]) + w1 * tt[q[7]];
      if (P.moist) {
        const auto* vv = g.vap;
        g.vap[cell] =
            w27 * vv[q[0]] + w9 * (vv[q[1]] + vv[q[2]] + vv[q[3]]) + w3 * (vv[q[4]] + vv[q[5]] + vv[q[6]]) + w1 * vv[q[7]];
      }
    }
    const auto* v = g.vel;
    const float3 a0 = f3mul(v[q[0]], w27);
    const float3 a1 = f3mul(f3add(f3add(v[q[1]], v[q[2]]), v[q[3]]), w9);
    const float3 a2 = f3mul(f3add(f3add(v[q[4]], v[q[5]]), v[q[6]]), w3);
    const float3 a3 = f3mul(v[q[7]], w1);
    g.vel[cell] = f3add(f3add(f3add(a0, a1), a2), a3);
    g.flu[cell] = cellFluidity(P, x, y, z, 1 << level);
  }
}

// ------------------------------------------------------------------ advection
// INIT_SAMPLE (dcgrid_fluid.cu:7-72): block containing the foot point, its apron 2x2x2 stencil
// (clamped into the apron), fluidity-weighted normalised trilinear weights.
template <int BW, bool FLUIDITY>
struct CrossSample {
  uint32_t i[8];
  float w[8];
  int x0, y0, z0, scale;
  bool zero;
  template <class G>
  __device__ __forceinline__ void init(const G& g, float px, float py, float pz) {
    const int ix = clampi((int)floorf(px), 0, g.gx - 1);
    const int iy = clampi((int)floorf(py), 0, g.gy - 1);
    const int iz = clampi((int)floorf(pz), 0, g.gz - 1);
    const uint32_t sb = findBlock<BW>(g, ix, iy, iz, 0);
    const int sl = g.lev[sb];
    const int3 sp = g.pos[sb];
    scale = 1 << sl;
    const float scaleInv = 1.f / (float)scale;
    const float x = px * scaleInv - .5f, y = py * scaleInv - .5f, z = pz * scaleInv - .5f;
    const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
    const float dx = x - xf, dy = y - yf, dz = z - zf;
    x0 = (int)xf;
    y0 = (int)yf;
    z0 = (int)zf;
    const int ai = clampi(x0 + 1 - sp.x, 0, BW), aj = clampi(y0 + 1 - sp.y, 0, BW), ak = clampi(z0 + 1 - sp.z, 0, BW);
#pragma unroll
    for (int c = 0; c < 8; c++) i[c] = apronCell<BW>(g, sb, ai + ((c >> 2) & 1), aj + ((c >> 1) & 1), ak + (c & 1));
    const float Dx = 1.f - dx, Dy = 1.f - dy, Dz = 1.f - dz;
    float ww[8];
    if (FLUIDITY) {
      ww[0] = g.flu[i[0]] * Dx * Dy * Dz;
      ww[1] = g.flu[i[1]] * Dx * Dy * dz;
      ww[2] = g.flu[i[2]] * Dx * dy * Dz;
      ww[3] = g.flu[i[3]] * Dx * dy * dz;
      ww[4] = g.flu[i[4]] * dx * Dy * Dz;
      ww[5] = g.flu[i[5]] * dx * Dy * dz;
      ww[6] = g.flu[i[6]] * dx * dy * Dz;
      ww[7] = g.flu[i[7]] * dx * dy * dz;
    } else {
      ww[0] = Dx * Dy * Dz;
      ww[1] = Dx * Dy * dz;
      ww[2] = Dx * dy * Dz;
      ww[3] = Dx * dy * dz;
      ww[4] = dx * Dy * Dz;
      ww[5] = dx * Dy * dz;
      ww[6] = dx * dy * Dz;
      ww[7] = dx * dy * dz;
    }
    const float wAcc = ww[0] + ww[1] + ww[2] + ww[3] + ww[4] + ww[5] + ww[6] + ww[7];
    const float wInv = 1.f / wAcc;
#pragma unroll
    for (int c = 0; c < 8; c++) w[c] = ww[c] * wInv;
    zero = wAcc < 1e-6f;
  }
};

// Backtraced foot point of a leaf cell (dcgrid_fluid.cu:115-122).
template <int BW>
__device__ __forceinline__ void footPoint(const dcg_params& P, int3 p0, int level, int t, float3 v, float& px, float& py,
                                          float& pz, int& x, int& y, int& z) {
  constexpr int BB = Geo<BW>::BB;
  const float scale = (float)(1 << level);
  const float alpha = P.dt * P.rdx;
  int lx, ly, lz;
  locPos<BB>((uint32_t)t, lx, ly, lz);
  x = p0.x | lx;
  y = p0.y | ly;
  z = p0.z | lz;
  px = ((float)x + .5f) * scale - v.x * alpha;
  py = ((float)y + .5f) * scale - v.y * alpha;
  pz = ((float)z + .5f) * scale - v.z * alpha;
}

// Same sample as CrossSample when the foot point's finest block is the thread's own block b:
// corners and fluidity come from the block's LDS tile (interior + full apron, raw values).
// Returns false (and the caller takes the global path) when the foot point lies in another
// block or in a refined subblock of b (chi8: b's eight child links, staged in LDS).
template <int BW, bool FLUIDITY>
struct TileSample {
  int a0;  // tile index of corner 000
  float w[8];
  int x0, y0, z0, scale;
  bool zero;
  __device__ __forceinline__ bool init(const DGBase& g, float px, float py, float pz, const uint32_t* chi8, int3 p0,
                                       int lb, const float* fluTile) {
    constexpr int A = Geo<BW>::A, BB = Geo<BW>::BB;
    const int ix = clampi((int)floorf(px), 0, g.gx - 1);
    const int iy = clampi((int)floorf(py), 0, g.gy - 1);
    const int iz = clampi((int)floorf(pz), 0, g.gz - 1);
    const int qx = (ix >> lb) - p0.x, qy = (iy >> lb) - p0.y, qz = (iz >> lb) - p0.z;
    if ((unsigned)qx >= (unsigned)BW || (unsigned)qy >= (unsigned)BW || (unsigned)qz >= (unsigned)BW) return false;
    const int t = (((qx >> (BB - 1)) & 1) << 2) | (((qy >> (BB - 1)) & 1) << 1) | ((qz >> (BB - 1)) & 1);
    if (chi8[t] != NFU) return false;
    scale = 1 << lb;
    const float scaleInv = 1.f / (float)scale;
    const float x = px * scaleInv - .5f, y = py * scaleInv - .5f, z = pz * scaleInv - .5f;
    const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
    const float dx = x - xf, dy = y - yf, dz = z - zf;
    x0 = (int)xf;
    y0 = (int)yf;
    z0 = (int)zf;
    const int ai = clampi(x0 + 1 - p0.x, 0, BW), aj = clampi(y0 + 1 - p0.y, 0, BW), ak = clampi(z0 + 1 - p0.z, 0, BW);
    a0 = (ai * A + aj) * A + ak;
    const float Dx = 1.f - dx, Dy = 1.f - dy, Dz = 1.f - dz;
    float ww[8];
    if (FLUIDITY) {
      ww[0] = fluTile[a0] * Dx * Dy * Dz;
      ww[1] = fluTile[a0 + 1] * Dx * Dy * dz;
      ww[2] = fluTile[a0 + A] * Dx * dy * Dz;
      ww[3] = fluTile[a0 + A + 1] * Dx * dy * dz;
      ww[4] = fluTile[a0 + A * A] * dx * Dy * Dz;
      ww[5] = fluTile[a0 + A * A + 1] * dx * Dy * dz;
      ww[6] = fluTile[a0 + A * A + A] * dx * dy * Dz;
      ww[7] = fluTile[a0 + A * A + A + 1] * dx * dy * dz;
    } else {
      ww[0] = Dx * Dy * Dz;
      ww[1] = Dx * Dy * dz;
      ww[2] = Dx * dy * Dz;
      ww[3] = Dx * dy * dz;
      ww[4] = dx * Dy * Dz;
      ww[5] = dx * Dy * dz;
      ww[6] = dx * dy * Dz;
      ww[7] = dx * dy * dz;
    }
    const float wAcc = ww[0] + ww[1] + ww[2] + ww[3] + ww[4] + ww[5] + ww[6] + ww[7];
    const float wInv = 1.f / wAcc;
#pragma unroll
    for (int c = 0; c < 8; c++) w[c] = ww[c] * wInv;
    zero = wAcc < 1e-6f;
    return true;
  }
  static __device__ __forceinline__ int off(int c) {
    constexpr int A = Geo<BW>::A;
    return ((c >> 2) & 1) * A * A + ((c >> 1) & 1) * A + (c & 1);
  }
};

// Visits the whole apron of block b ((BW+2)^3 entries, edges and corners included) for thread t
// of the N = BV / CPT threads that handle the block: op(a, cell) for the thread's CPT interior
// cells (t + j N) and for its share of the border entries, with the cell of each entry resolved
// once (apronCell) however many fields are staged from it.
template <int BW, int CPT, class Op>
__device__ __forceinline__ void forApron(const DGBase& g, uint32_t b, int t, Op op) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, A3 = Geo<BW>::A3, N = BV / CPT;
#pragma unroll
  for (int j = 0; j < CPT; j++) op(apronIndexOfLocal<BW>(t + j * N), b * BV + t + j * N);
  for (int a = t; a < A3; a += N) {
    const int ai = a / (A * A), aj = (a / A) % A, ak = a % A;
    if (ai > 0 && ai <= BW && aj > 0 && aj <= BW && ak > 0 && ak <= BW) continue;
    op(a, borderApronCell<BW>(g, b, ai, aj, ak));
  }
}

// Advection workgroup shape. The advection kernels hold ~90-100 VGPRs (5 waves per SIMD), so a
// 512-thread workgroup per BW-8 block leaves two blocks resident per CU, and the kernels are
// bound by blocks in flight times the staging latency. A BW-8 block is therefore handled by
// BV / CPT threads with CPT cells each, one cell after the other: more blocks resident at the
// same register count (two cells per thread: four resident blocks per CU; four cells per thread
// measured slower). BW 4 keeps four blocks of 64 threads.
template <int BW>
struct AdvGeo {
  static constexpr int CPT = BW == 8 ? 2 : 1;
  static constexpr int WG = Geo<BW>::WG / CPT;
};
// Minimum waves per SIMD asked of the compiler: 5 for BW 8 (at most 96 VGPRs, five blocks
// resident per CU). The variants without fluidity-weighted temperature or vapor fit without
// spilling; the terrain variants (fluidity with buoyancy, moist density) spill a few VGPRs on
// their cross-block path and still run faster than at four waves.
template <int BW>
struct AdvWaves {
  static constexpr int value = BW == 8 ? 5 : 1;
};

// k_dcgrid_advect_velocity (dcgrid_fluid.cu:112-127) + buoyancy (extension). The block's
// velocity (and fluidity) with the full apron are staged in LDS; samples whose foot point stays
// in the block read the tile, the others gather through the neighbour tables.
template <int BW, bool FLUIDITY, bool BUOY, class G>
__global__ void __launch_bounds__(AdvGeo<BW>::WG, (AdvWaves<BW>::value)) k_advect_velocity(G g, dcg_params P) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW;
  constexpr int CPT = AdvGeo<BW>::CPT, N = BV / CPT;
  __shared__ float3 vtiles[BPW][A3];
  __shared__ float ftiles[FLUIDITY ? BPW : 1][FLUIDITY ? A3 : 1];
  __shared__ uint32_t chis[BPW][8];
  const uint32_t total = slotTotal(g, -1);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();  // the previous block's tiles are no longer read
    const BlockCtx<BW> c = slotCtx<BW>(g, -1, base, total);  // c.t: the thread's first cell
    float3* vt = vtiles[c.sub];
    float* ft = FLUIDITY ? ftiles[c.sub] : nullptr;
    uint32_t* chi8 = chis[c.sub];
    int3 p0 = make_int3(0, 0, 0);
    if (c.valid) {
      p0 = g.pos[c.b];
      if (c.t < 8) chi8[c.t] = g.chi[8 * c.b + c.t];
      forApron<BW, CPT>(g, c.b, c.t, [&](int a, uint32_t src) {
        vt[a] = g.vel[src];
        if (FLUIDITY) ft[a] = g.flu[src];
      });
    }
    __syncthreads();
    if (!c.valid) continue;
    const int level = c.level;
#pragma unroll 1
    for (int j = 0; j < CPT; j++) {
      const int t = c.t + j * N;
      const uint32_t cell = c.b * BV + t;
      if (chi8[t >> SBB] != NFU) {  // not a leaf cell
        g.tvel[cell] = make_float3(0.f, 0.f, 0.f);
        continue;
      }
      float tc = 0.f, dc = 0.f;  // buoyancy inputs, issued before the sample and used after it
      if (BUOY) {
        tc = g.tem[cell];
        dc = g.den[cell];
      }
      float px, py, pz;
      int x, y, z;
      footPoint<BW>(P, p0, level, t, vt[apronIndexOfLocal<BW>(t)], px, py, pz, x, y, z);
      float3 acc = make_float3(0.f, 0.f, 0.f);
      TileSample<BW, FLUIDITY> ts;
      if (ts.init(g, px, py, pz, chi8, p0, level, ft)) {
        if (!ts.zero) {
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const float3 vk = velocityBC(P, vt[ts.a0 + ts.off(k)], ts.x0 + ((k >> 2) & 1), ts.y0 + ((k >> 1) & 1),
                                         ts.z0 + (k & 1), ts.scale);
            acc = k == 0 ? f3mul(vk, ts.w[0]) : f3add(acc, f3mul(vk, ts.w[k]));
          }
        }
      } else {
        CrossSample<BW, FLUIDITY> s;
        s.init(g, px, py, pz);
        if (!s.zero) {
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const float3 vk =
                velocityBC(P, g.vel[s.i[k]], s.x0 + ((k >> 2) & 1), s.y0 + ((k >> 1) & 1), s.z0 + (k & 1), s.scale);
            acc = k == 0 ? f3mul(vk, s.w[0]) : f3add(acc, f3mul(vk, s.w[k]));
          }
        }
      }
      if (BUOY) {
        const float ta = ambientT(P, cellHeight(y, 1 << level));
        const float bf = P.dt * (P.buoy_beta * (tc - ta) - P.buoy_alpha * dc);
        acc.y = acc.y + bf;
      }
      g.tvel[cell] = acc;
    }
  }
}

// k_dcgrid_advect_density (dcgrid_fluid.cu:129-144) + temperature and sphere source (extension),
// with the same LDS fast path as the velocity.
template <int BW, bool FLUIDITY, bool SRC, bool MOIST, class G>
__global__ void __launch_bounds__(AdvGeo<BW>::WG, (AdvWaves<BW>::value)) k_advect_density(G g, dcg_params P) {
  // SRC: temperature is advected (sphere source or moist atmosphere); MOIST: vapor too, then the
  // saturation adjustment of (vapor, cloud water = density, temperature).
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW;
  constexpr int CPT = AdvGeo<BW>::CPT, N = BV / CPT;
  __shared__ float dtiles[BPW][A3];
  __shared__ float ttiles[SRC ? BPW : 1][SRC ? A3 : 1];
  __shared__ float vtiles[MOIST ? BPW : 1][MOIST ? A3 : 1];
  __shared__ float ftiles[FLUIDITY ? BPW : 1][FLUIDITY ? A3 : 1];
  __shared__ uint32_t chis[BPW][8];
  const bool sphere = P.source_mode == DCG_SOURCE_SPHERE_BUOYANT;
  const uint32_t total = slotTotal(g, -1);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();  // the previous block's tiles are no longer read
    const BlockCtx<BW> c = slotCtx<BW>(g, -1, base, total);  // c.t: the thread's first cell
    float* dt_ = dtiles[c.sub];
    float* tt = SRC ? ttiles[c.sub] : nullptr;
    float* vt = MOIST ? vtiles[c.sub] : nullptr;
    float* ft = FLUIDITY ? ftiles[c.sub] : nullptr;
    uint32_t* chi8 = chis[c.sub];
    int3 p0 = make_int3(0, 0, 0);
    float3 un = make_float3(0.f, 0.f, 0.f);  // velocity of the next cell to advect
    if (c.valid) {
      p0 = g.pos[c.b];
      if (c.t < 8) chi8[c.t] = g.chi[8 * c.b + c.t];
      un = g.vel[c.b * BV + c.t];
      forApron<BW, CPT>(g, c.b, c.t, [&](int a, uint32_t src) {
        dt_[a] = g.den[src];
        if (SRC) tt[a] = g.tem[src];
        if (MOIST) vt[a] = g.vap[src];
        if (FLUIDITY) ft[a] = g.flu[src];
      });
    }
    __syncthreads();
    if (!c.valid) continue;
    const int level = c.level;
#pragma unroll 1
    for (int j = 0; j < CPT; j++) {
      const int t = c.t + j * N;
      const uint32_t cell = c.b * BV + t;
      const float3 u = un;
      if (j + 1 < CPT) un = g.vel[cell + N];  // issued before this cell's sample
      if (chi8[t >> SBB] != NFU) {  // not a leaf cell
        g.tden[cell] = 0.f;
        if (SRC) g.ttem[cell] = 0.f;
        if (MOIST) g.tvap[cell] = 0.f;
        continue;
      }
      float px, py, pz;
      int x, y, z;
      footPoint<BW>(P, p0, level, t, u, px, py, pz, x, y, z);
      float q = 0.f, tq = 0.f, vq = 0.f;
      TileSample<BW, FLUIDITY> ts;
      if (ts.init(g, px, py, pz, chi8, p0, level, ft)) {
        if (!ts.zero) {
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const int cx = ts.x0 + ((k >> 2) & 1), cy = ts.y0 + ((k >> 1) & 1), cz = ts.z0 + (k & 1);
            const float qk = densityBC(P, dt_[ts.a0 + ts.off(k)], cx, cy, cz, ts.scale);
            q = k == 0 ? qk * ts.w[0] : q + qk * ts.w[k];
            if (SRC) {
              const float tk = temperatureBC(P, tt[ts.a0 + ts.off(k)], cx, cy, cz, ts.scale);
              tq = k == 0 ? tk * ts.w[0] : tq + tk * ts.w[k];
            }
            if (MOIST) {
              const float vk = vaporBC(P, vt[ts.a0 + ts.off(k)], cx, cy, cz, ts.scale);
              vq = k == 0 ? vk * ts.w[0] : vq + vk * ts.w[k];
            }
          }
        }
      } else {
        CrossSample<BW, FLUIDITY> s;
        s.init(g, px, py, pz);
        if (!s.zero) {
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const int cx = s.x0 + ((k >> 2) & 1), cy = s.y0 + ((k >> 1) & 1), cz = s.z0 + (k & 1);
            const float qk = densityBC(P, g.den[s.i[k]], cx, cy, cz, s.scale);
            q = k == 0 ? qk * s.w[0] : q + qk * s.w[k];
            if (SRC) {
              const float tk = temperatureBC(P, g.tem[s.i[k]], cx, cy, cz, s.scale);
              tq = k == 0 ? tk * s.w[0] : tq + tk * s.w[k];
            }
            if (MOIST) {
              const float vk = vaporBC(P, g.vap[s.i[k]], cx, cy, cz, s.scale);
              vq = k == 0 ? vk * s.w[0] : vq + vk * s.w[k];
            }
          }
        }
      }
      if (SRC) {
        // the cell's coordinates again, from an opaque copy of t: no live range across the sample
        int tr = t;
        asm volatile("" : "+v"(tr));
        locPos<Geo<BW>::BB>((uint32_t)tr, x, y, z);
        x |= p0.x;
        y |= p0.y;
        z |= p0.z;
        const float sc = (float)(1 << level);
        if (sphere && inSource(P, ((float)x + .5f) * sc, ((float)y + .5f) * sc, ((float)z + .5f) * sc)) {
          q = P.src_density;
          tq = sourceTemperature(P, x, y, z, level);
        }
        if (MOIST) {
          tq = tq - (P.dt * P.adiabatic_rate) * u.y;  // expansion cooling of rising air
          saturationAdjust(P, vq, q, tq);
          g.tvap[cell] = vq;
        }
        g.ttem[cell] = tq;
      }
      g.tden[cell] = q;
    }
  }
}

// ------------------------------------------------------------------ residual (debugStats)
// k_dcgrid_debug_stats (dcgrid_structure.cu:224-251): per block, sum over leaf cells in apron
// [i][j][k] order of scale*|div - (sum p_n - 6p)*rdx^2/scale^2|. One thread per block keeps the
// reference's summation order.
template <int BW, class G>
__global__ void __launch_bounds__(64) k_residual(G g, dcg_params P, float* __restrict__ stats) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB;
  const uint32_t b = blockIdx.x * 64 + threadIdx.x;
  if (b >= g.B) return;
  if (g.lev[b] == FREE) {
    stats[b] = 0.f;
    return;
  }
  const int lb = g.lev[b];
  const int scale = 1 << lb;
  const float alpha = P.rdx * P.rdx / (float)(scale * scale);
  float s = 0.f;
  for (int i = 1; i <= BW; i++)
    for (int j = 1; j <= BW; j++)
      for (int k = 1; k <= BW; k++) {
        const uint32_t cidx = apronCell<BW>(g, b, i, j, k);
        if (g.chi[cidx >> SBB] != NFU) continue;
        const auto* p = g.pre;
        const float pc = p[cidx];
        const float pl = p[apronCell<BW>(g, b, i - 1, j, k)], pr = p[apronCell<BW>(g, b, i + 1, j, k)];
        const float pd = p[apronCell<BW>(g, b, i, j - 1, k)], pu = p[apronCell<BW>(g, b, i, j + 1, k)];
        const float pb = p[apronCell<BW>(g, b, i, j, k - 1)], pf = p[apronCell<BW>(g, b, i, j, k + 1)];
        const float r = g.div[cidx] - (pl + pr + pd + pu + pb + pf - 6.f * pc) * alpha;
        s = s + (float)scale * fabsf(r);
      }
  stats[b] = s;
  (void)BV;
}

// ------------------------------------------------------------------ seeded scene (extension)
template <int BW, class G>
__global__ void __launch_bounds__(Geo<BW>::WG) k_seed(G g, dcg_params P) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB;
  const BlockCtx<BW> c = blockCtx<BW>(0, g.B);
  if (!c.valid || g.lev[c.b] == FREE) return;
  const int level = g.lev[c.b];
  const float sc = (float)(1 << level);
  const int3 p0 = g.pos[c.b];
  int lx, ly, lz;
  locPos<BB>((uint32_t)c.t, lx, ly, lz);
  const int x = p0.x | lx, y = p0.y | ly, z = p0.z | lz;
  const float cx = ((float)x + .5f) * sc, cy = ((float)y + .5f) * sc, cz = ((float)z + .5f) * sc;
  const uint32_t cell = c.b * BV + c.t;
  g.vel[cell] = initVelocity(P, cx, cy, cz);
  if (P.source_mode == DCG_SOURCE_SPHERE_BUOYANT && inSource(P, cx, cy, cz)) {
    g.den[cell] = P.src_density;
    g.tem[cell] = sourceTemperature(P, x, y, z, level);
  } else {
    g.den[cell] = 0.f;
    g.tem[cell] = ambientT(P, cy);
  }
  if (P.moist) g.vap[cell] = ambientVapor(P, cy);
}

}  // namespace dcg

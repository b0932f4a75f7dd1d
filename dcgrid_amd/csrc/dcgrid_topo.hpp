// dcgrid_topo.hpp — gfx950 kernels of DCGrid block-table maintenance: dense-level activation
// (src/dcgrid/dcgrid_structure.cu:104-178), refinement scores (dcgrid_adaptation.cu:10-63), move
// selection (fluid_simulation_dcgrid.cu:348-437) and moves (dcgrid_adaptation.cu:65-90), hash
// rebuild (:180-201), refinement selection with prefix-sum allocation
// (fluid_simulation_dcgrid.cu:439-483, dcgrid_adaptation.cu:145-178) and the apron refresh
// (dcgrid_structure.cu:30-92) on the compact neighbour table.
//
// The reference runs selection on the host after copying all scores over PCIe, then sorts with
// std::nth_element/std::sort. Here selection stays on the GPU: a single-workgroup radix select
// finds the k-th key of the total order and compacts the k winners, and a ranking kernel places
// each winner at its exact position. Allocation of refined children is an exclusive prefix sum
// in index order (deterministic, unlike the reference's atomicAdd).
#pragma once

#include <float.h>

#include "dcgrid_dev.hpp"

namespace dcg {

// ------------------------------------------------------------------ dense levels
// k_dcgrid_activate_level (dcgrid_structure.cu:104-178) for one dense level: position, level,
// parent subblock, children on a dense finer level, neighbour table (same-level neighbours, the
// block itself for out-of-domain directions), zero fields and fluidity.
template <int BW, class G>
__global__ void __launch_bounds__(256) k_activate_dense(G g, dcg_params P, int level) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB;
  const uint32_t n = (uint32_t)(g.rx[level] * g.ry[level] * g.rz[level]);
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int rx = g.rx[level], ry = g.ry[level], rz = g.rz[level];
  const int bz = k % rz, by = (k / rz) % ry, bx = k / (rz * ry);
  const uint32_t b = g.offs[level] + k;
  const int3 p = make_int3(BW * bx, BW * by, BW * bz);
  g.pos[b] = p;
  g.lev[b] = (uint8_t)level;
  if (level < g.L - 1) {
    const int ry1 = g.ry[level + 1], rz1 = g.rz[level + 1];
    const uint32_t pb = g.offs[level + 1] + (uint32_t)(((bx >> 1) * ry1 + (by >> 1)) * rz1 + (bz >> 1));
    g.par[b] = 8 * pb + 4 * (bx & 1) + 2 * (by & 1) + (bz & 1);
  }
  if (level - 1 >= g.S)
    for (int c = 0; c < 8; c++) {
      const int cx = 2 * bx + ((c >> 2) & 1), cy = 2 * by + ((c >> 1) & 1), cz = 2 * bz + (c & 1);
      g.chi[8 * b + c] = g.offs[level - 1] + (uint32_t)((cx * g.ry[level - 1] + cy) * g.rz[level - 1] + cz);
    }
  for (int dx = -1; dx <= 1; dx++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++) {
        const int nx = bx + dx, ny = by + dy, nz = bz + dz;
        int4 d;
        if (nx < 0 || ny < 0 || nz < 0 || nx >= rx || ny >= ry || nz >= rz)
          d = makeNd(b, ND_OUT, 0, 0, 0);  // clamped interior cell (dcgrid_structure.cu:157-167)
        else
          d = makeNd(g.offs[level] + (uint32_t)((nx * ry + ny) * rz + nz), 0, BW * dx, BW * dy, BW * dz);
        g.nd[b * 27 + dirIndex(dx, dy, dz)] = d;
      }
  const int scale = 1 << level;
  for (int t = 0; t < BV; t++) {
    int lx, ly, lz;
    locPos<BB>((uint32_t)t, lx, ly, lz);
    const uint32_t cell = b * BV + t;
    g.den[cell] = 0.f;
    g.tem[cell] = P.t_ambient;
    g.vel[cell] = make_float3(0.f, 0.f, 0.f);
    g.flu[cell] = cellFluidity(P, p.x + lx, p.y + ly, p.z + lz, scale);
  }
}

// ------------------------------------------------------------------ scores
__device__ __forceinline__ bool levelBlocked(const DGBase& g, int level) {
  return (level == 0 && g.ctl[CTL_LOADS + 0] == g.fullL[0]) ||
         (level > 0 && g.ctl[CTL_LOADS + level - 1] == g.fullL[level - 1]);
}

// k_dcgrid_calc_subblock_scores, geometric score (dcgrid_adaptation.cu:10-34), of subblock s.
template <int BW>
__device__ __forceinline__ float subScoreGeo(const DGBase& g, uint32_t s) {
  const uint32_t b = s / 8;
  const int level = g.lev[b];
  if (g.chi[s] != NFU || level == FREE || levelBlocked(g, level)) return -FLT_MAX;
  const float hb = (float)(BW / 2), qb = (float)(BW / 4);
  const float cx = .5f * (float)g.gx, cy = .45f * (float)g.gy, cz = .5f * (float)g.gz;
  const int3 p = g.pos[b];
  const float sc = (float)(1 << level);
  const float qx = sc * (((float)p.x + hb * (float)((s >> 2) & 1)) + qb);
  const float qy = sc * (((float)p.y + hb * (float)((s >> 1) & 1)) + qb);
  const float qz = sc * (((float)p.z + hb * (float)(s & 1)) + qb);
  const float dx = qx - cx, dy = qy - cy, dz = qz - cz;
  const float d = __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
  return d < .2f * (float)g.gx ? 0.f : (float)g.gx / d;
}
template <int BW>
__global__ void __launch_bounds__(256) k_sub_scores_geo(DGBase g, float* __restrict__ sbs) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s >= 8 * g.B) return;
  sbs[s] = subScoreGeo<BW>(g, s);
}

// Feature score (extension): per cell max(|grad density| / grad_threshold, |vorticity| /
// vort_threshold); a subblock's feature value is the maximum over its cells. One pass computes
// the vorticity field (k_vorticity's values, stored as k_vorticity stores them) and the
// features: velocity, fluidity and density are staged with the apron faces in LDS, the magnitude
// uses the stored vorticity, and the subblock maximum is an LDS reduction. Computed once per
// adaptation: the fields and the (pre-refresh) apron do not change between the score pass before
// the moves and the one before refinement, so both passes see the same feature values.
template <int BW, bool FLUIDITY, int CPT, class G>
__global__ void __launch_bounds__((CptGeo<BW, CPT>::WG), (CPT == 2 ? 8 : 1)) k_features(G g, dcg_params P, float* __restrict__ feat) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW, FACE = Geo<BW>::FACE,
                SBB = Geo<BW>::SBB, N = BV / CPT;
  __shared__ float3 vtiles[BPW][A3];
  __shared__ float ftiles[FLUIDITY ? BPW : 1][FLUIDITY ? A3 : 1];
  __shared__ float dtiles[BPW][A3];
  __shared__ unsigned int smax[BPW][8];
  const uint32_t total = slotTotal(g, -1);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();  // the previous block's tiles and maxima are no longer read
    const BlockCtx<BW> c = slotCtx<BW>(g, -1, base, total);  // c.t: the thread's first cell
    const int sub = c.sub;
    const uint32_t b = c.b;
    float3* vt = vtiles[sub];
    float* ft = FLUIDITY ? ftiles[sub] : nullptr;
    float* dt_ = dtiles[sub];
    if (c.t < 8) smax[sub][c.t] = 0u;
    if (c.valid) {
#pragma unroll
      for (int j = 0; j < CPT; j++) {
        const int t = c.t + j * N;
        const uint32_t cell = b * BV + t;
        const int ac = apronIndexOfLocal<BW>(t);
        vt[ac] = g.vel[cell];
        if (FLUIDITY) ft[ac] = g.flu[cell];
        dt_[ac] = g.den[cell];
      }
      for (int fc = c.t; fc < FACE; fc += N) {
        int ai, aj, ak;
        faceCell<BW>(fc, ai, aj, ak);
        const uint32_t src = borderApronCell<BW>(g, b, ai, aj, ak);
        const int a = (ai * A + aj) * A + ak;
        vt[a] = g.vel[src];
        if (FLUIDITY) ft[a] = g.flu[src];
        dt_[a] = g.den[src];
      }
    }
    __syncthreads();
    const float alpha = .5f * P.rdx / (float)(1 << c.level);
#pragma unroll
    for (int j = 0; j < CPT; j++) {
      const int t = c.t + j * N;
      const uint32_t cell = b * BV + t;
      float cs = 0.f;
      if (c.valid) {
        const int a = apronIndexOfLocal<BW>(t);
        // vorticity (k_vorticity)
        float w[6];
        tileFluidity<BW, FLUIDITY>(ft, a, w);
        const float3 vl = vt[a - A * A], vr = vt[a + A * A], vd = vt[a - A], vu = vt[a + A], vb = vt[a - 1],
                     vf = vt[a + 1];
        const float wl = w[0], wr = w[1], wd = w[2], wu = w[3], wb = w[4], wf = w[5];
        g.vort[cell] = make_float3(alpha * ((wu * vu.z - wd * vd.z) - (wf * vf.y - wb * vb.y)),
                                   alpha * ((wf * vf.x - wb * vb.x) - (wr * vr.z - wl * vl.z)),
                                   alpha * ((wr * vr.y - wl * vl.y) - (wu * vu.x - wd * vd.x)));
        const float3 om = g.vort[cell];  // the stored value (rounded as stored)
        // density gradient
        const float gxv = alpha * (dt_[a + A * A] - dt_[a - A * A]);
        const float gyv = alpha * (dt_[a + A] - dt_[a - A]);
        const float gzv = alpha * (dt_[a + 1] - dt_[a - 1]);
        const float gm = __builtin_sqrtf(gxv * gxv + gyv * gyv + gzv * gzv);
        const float wm = __builtin_sqrtf(om.x * om.x + om.y * om.y + om.z * om.z);
        const float fg = gm / P.grad_threshold, fw = wm / P.vort_threshold;
        cs = fg > fw ? fg : fw;
      }
      // scores are >= 0, so their bit patterns order like their values
      atomicMax(&smax[sub][t >> SBB], __float_as_uint(cs));
    }
    __syncthreads();
    if (c.t < 8 && c.valid) feat[8 * b + c.t] = __uint_as_float(smax[sub][c.t]);
  }
}

// Subblock scores from the feature values: -FLT_MAX where the reference's conditions exclude the
// subblock (dcgrid_adaptation.cu:18-24).
__device__ __forceinline__ float subScoreFeat(const DGBase& g, const float* __restrict__ feat, uint32_t s) {
  const int level = g.lev[s / 8];
  return g.chi[s] != NFU || level == FREE || levelBlocked(g, level) ? -FLT_MAX : feat[s];
}
__global__ void __launch_bounds__(256) k_sub_scores_feat(DGBase g, const float* __restrict__ feat, float* __restrict__ sbs) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s >= 8 * g.B) return;
  sbs[s] = subScoreFeat(g, feat, s);
}

// First score pass of an adaptation: subblock scores (k_dcgrid_calc_subblock_scores,
// dcgrid_adaptation.cu:10-34, or the feature values) and block scores
// (k_dcgrid_accumulate_subblock_scores, :42-63) in one launch, one thread per block. The
// reference's ninth term s[8] is subblock 0 of the next block (the -FLT_MAX pad after the last
// block): the thread recomputes that value, which the next block's thread stores. The thread also
// clears its block's flags for this adaptation, and thread 0 the move total.
template <int BW, bool FEAT>
__global__ void __launch_bounds__(256) k_scores(DGBase g, const float* __restrict__ feat, float* __restrict__ sbs,
                                                float* __restrict__ bs) {
  const uint32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b == 0) g.ctl[CTL_TOTAL_MOVES] = 0;
  if (b >= g.B) return;
  g.flg[b] = 0;
  float s[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const uint32_t sb = 8 * b + k;
    s[k] = sb >= 8 * g.B ? -FLT_MAX : (FEAT ? subScoreFeat(g, feat, sb) : subScoreGeo<BW>(g, sb));
    if (k < 8) sbs[sb] = s[k];
  }
  float r = -FLT_MAX;
  const int level = g.lev[b];
  if (level != FREE && !levelBlocked(g, level)) {
    if (s[0] > 0.f && s[1] > 0.f && s[2] > 0.f && s[3] > 0.f && s[4] > 0.f && s[5] > 0.f && s[6] > 0.f && s[7] > 0.f)
      r = .125f * (s[0] + s[1] + s[2] + s[3] + s[4] + s[5] + s[6] + s[7] + s[8]);
  }
  bs[b] = r;
}

// ------------------------------------------------------------------ selection
// Keys of the total orders used for move selection (32-bit; ties broken by index):
//  blocks: non-negative scores ascending, negative scores last;
//  subblocks: scores descending.
__device__ __forceinline__ uint32_t blockKey(float s) {
  if (s < 0.f) return 0xFFFFFFFFu;
  if (s == 0.f) return 0u;
  return __float_as_uint(s);
}
__device__ __forceinline__ uint32_t subKey(float s) {
  uint32_t u = s == 0.f ? 0u : __float_as_uint(s);
  const uint32_t ord = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ~ord;
}

// Workgroup-wide exclusive scan of one uint32 per thread (1024 threads, 64-wide wavefronts).
__device__ __forceinline__ uint32_t wgExclusiveScan(uint32_t v, uint32_t* sh /*[16]*/, uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++) {
      const uint32_t s = sh[w];
      sh[w] = acc;
      acc += s;
    }
    sh[16] = acc;
  }
  __syncthreads();
  const uint32_t r = sh[wave] + x - v;
  if (total) *total = sh[16];
  __syncthreads();
  return r;
}

// Move limit of `level` for this adaptation (fluid_simulation_dcgrid.cu:377-385).
__device__ __forceinline__ uint32_t moveLim(const DGBase& g, int level) {
  const uint32_t d0 = g.maxL[level], d1 = 8 * g.maxL[level + 1];
  const uint32_t load = g.ctl[CTL_LOADS + level];
  uint32_t lim = min(min(d0, d1), min(load, g.fullL[level] - load));
  const uint32_t ml = g.ctl[CTL_MOVE_LIMIT + level];
  if (ml > 0) lim = min(lim, ml);
  return lim;
}
// Candidates selected per mode: lim move candidates (MODE 0), min(lim + 1, n) destination
// subblocks (MODE 1; none when lim is 0).
template <int MODE>
__device__ __forceinline__ uint32_t selectCount(const DGBase& g, int level, uint32_t lim) {
  const uint32_t d1 = 8 * g.maxL[level + 1];
  return MODE == 0 ? lim : (lim == 0 ? 0 : min(lim + 1, d1));
}

// Single-workgroup selection of the k smallest (key32, index) pairs among n candidates.
// MODE 0: move candidates (blocks of `level`, k = lim).
// MODE 1: destination subblocks (subblocks of level+1, k = min(lim+1, n)).
// Writes the k winners unsorted as packed keys (key32 << 32 | position) to cand.
struct SelectShared {
  uint32_t hist[2048];
  uint32_t sh[17];
  uint32_t sPrefix, sRem, sCount;
};
// Adds 1 to hist[bin] for every active lane of the wavefront. Lanes holding the common key
// (that of -FLT_MAX scores, shared by every refined or free candidate, often most of them) add
// with one LDS atomic per wavefront instead of serialising on one bin; the others one atomic
// each. Called by all lanes of the wavefront.
__device__ __forceinline__ void histAdd(uint32_t* hist, uint32_t bin, bool active, bool common) {
  const unsigned long long m = __ballot(active && common);
  if (active && common) {
    if ((int)(threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicAdd(&hist[bin], (uint32_t)__popcll(m));
  } else if (active) {
    atomicAdd(&hist[bin], 1u);
  }
}

// The radix select of selectBody with the thread's keys (a segment of at most KEYS candidates)
// loaded once into registers: the same passes, ties and winners.
constexpr uint32_t SELECT_REG_KEYS = 8;
template <uint32_t KEYS, class KeyFn>
__device__ __forceinline__ void selectInRegisters(uint32_t lo, uint32_t hi, uint32_t k, KeyFn key32,
                                                  unsigned long long* __restrict__ cand, SelectShared& S) {
  const uint32_t T = blockDim.x;
  uint32_t kr[KEYS];
#pragma unroll
  for (uint32_t j = 0; j < KEYS; j++) kr[j] = lo + j < hi ? key32(lo + j) : 0u;
  uint32_t prefix = 0, mask = 0, rem = k;
  for (int pass = 0; pass < 3; pass++) {
    const int shift = pass == 0 ? 21 : (pass == 1 ? 10 : 0);
    const uint32_t nb = pass == 2 ? 1024u : 2048u;
    for (uint32_t j = threadIdx.x; j < 2048; j += T) S.hist[j] = 0;
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < KEYS; j++)  // short segments: few equal keys per wavefront, plain atomics
      if (lo + j < hi && (kr[j] & mask) == prefix) atomicAdd(&S.hist[(kr[j] >> shift) & (nb - 1)], 1u);
    __syncthreads();
    const uint32_t a0 = 2 * threadIdx.x < nb ? S.hist[2 * threadIdx.x] : 0;
    const uint32_t a1 = 2 * threadIdx.x + 1 < nb ? S.hist[2 * threadIdx.x + 1] : 0;
    const uint32_t ex = wgExclusiveScan(a0 + a1, S.sh, nullptr);
    if (rem > ex && rem <= ex + a0) {
      S.sPrefix = 2 * threadIdx.x;
      S.sRem = rem - ex;
    } else if (rem > ex + a0 && rem <= ex + a0 + a1) {
      S.sPrefix = 2 * threadIdx.x + 1;
      S.sRem = rem - ex - a0;
    }
    __syncthreads();
    prefix |= S.sPrefix << shift;
    mask |= (nb - 1) << shift;
    rem = S.sRem;
    __syncthreads();
  }
  uint32_t ties = 0;
#pragma unroll
  for (uint32_t j = 0; j < KEYS; j++) ties += lo + j < hi && kr[j] == prefix;
  const uint32_t tieBase = wgExclusiveScan(ties, S.sh, nullptr);
  uint32_t seen = 0;
#pragma unroll
  for (uint32_t j = 0; j < KEYS; j++) {
    if (lo + j >= hi) continue;
    const uint32_t kk = kr[j];
    bool take = kk < prefix;
    if (kk == prefix) {
      take = tieBase + seen < rem;
      seen++;
    }
    if (take) {
      const uint32_t slot = atomicAdd(&S.sCount, 1u);
      cand[slot] = ((unsigned long long)kk << 32) | (lo + j);
    }
  }
}

template <int MODE>
__device__ void selectBody(const DGBase& g, int level, const float* __restrict__ scores,
                           unsigned long long* __restrict__ cand, SelectShared& S) {
  uint32_t* hist = S.hist;
  uint32_t* sh = S.sh;
  const uint32_t d0 = g.maxL[level], d1 = 8 * g.maxL[level + 1];
  const uint32_t n = MODE == 0 ? d0 : d1;
  const uint32_t base = MODE == 0 ? g.offs[level] : 8 * g.offs[level + 1];
  const uint32_t lim = moveLim(g, level);  // the same value in both workgroups of a pair
  if (MODE == 0 && threadIdx.x == 0) g.ctl[CTL_LIM + level] = lim;
  if (threadIdx.x == 0) S.sCount = 0;
  __syncthreads();
  const uint32_t k = selectCount<MODE>(g, level, lim);
  if (k == 0) return;
  const uint32_t T = blockDim.x;
  auto key32 = [&](uint32_t i) -> uint32_t {
    const float s = scores[base + i];
    return MODE == 0 ? blockKey(s) : subKey(s);
  };
  // Candidates past the allocated prefix of the slice are free blocks (or their subblocks):
  // their score is -FLT_MAX, i.e. the largest key, and their positions follow every allocated
  // one, so they rank after all allocated candidates. Select among the allocated prefix and
  // append free candidates in position order when k exceeds it.
  const uint32_t nA = min(n, MODE == 0 ? g.ctl[CTL_LOADS + level] : 8 * g.ctl[CTL_LOADS + level + 1]);
  if (k >= nA) {
    const uint32_t kFree = MODE == 0 ? blockKey(-FLT_MAX) : subKey(-FLT_MAX);
    for (uint32_t i = threadIdx.x; i < k; i += T)
      cand[i] = ((unsigned long long)(i < nA ? key32(i) : kFree) << 32) | i;
    return;
  }
  const uint32_t seg = (nA + T - 1) / T;
  const uint32_t lo = min(threadIdx.x * seg, nA), hi = min(lo + seg, nA);
  if (seg <= SELECT_REG_KEYS) {  // (a larger register segment spills: scratch on every launch)
    selectInRegisters<SELECT_REG_KEYS>(lo, hi, k, key32, cand, S);
    return;
  }
  const uint32_t common = MODE == 0 ? blockKey(-FLT_MAX) : subKey(-FLT_MAX);
  // Radix select on key32: 11 + 11 + 10 bits.
  uint32_t prefix = 0, mask = 0, rem = k;
  for (int pass = 0; pass < 3; pass++) {
    const int shift = pass == 0 ? 21 : (pass == 1 ? 10 : 0);
    const uint32_t nb = pass == 2 ? 1024u : 2048u;
    for (uint32_t j = threadIdx.x; j < 2048; j += T) hist[j] = 0;
    __syncthreads();
#pragma unroll 1
    for (uint32_t j = 0; j < seg; j++) {  // the same trip count in every lane
      const uint32_t i = lo + j;
      const uint32_t kk = i < hi ? key32(i) : 0u;
      histAdd(hist, (kk >> shift) & (nb - 1), i < hi && (kk & mask) == prefix, kk == common);
    }
    __syncthreads();
    const uint32_t a0 = 2 * threadIdx.x < nb ? hist[2 * threadIdx.x] : 0;
    const uint32_t a1 = 2 * threadIdx.x + 1 < nb ? hist[2 * threadIdx.x + 1] : 0;
    const uint32_t ex = wgExclusiveScan(a0 + a1, sh, nullptr);
    if (rem > ex && rem <= ex + a0) {
      S.sPrefix = 2 * threadIdx.x;
      S.sRem = rem - ex;
    } else if (rem > ex + a0 && rem <= ex + a0 + a1) {
      S.sPrefix = 2 * threadIdx.x + 1;
      S.sRem = rem - ex - a0;
    }
    __syncthreads();
    prefix |= S.sPrefix << shift;
    mask |= (nb - 1) << shift;
    rem = S.sRem;
    __syncthreads();
  }
  // prefix = key32 of the k-th smallest; take all smaller keys and the first `rem` ties.
  uint32_t ties = 0;
  for (uint32_t i = lo; i < hi; i++) ties += key32(i) == prefix;
  const uint32_t tieBase = wgExclusiveScan(ties, sh, nullptr);
  uint32_t seen = 0;
  for (uint32_t i = lo; i < hi; i++) {
    const uint32_t kk = key32(i);
    bool take = kk < prefix;
    if (kk == prefix) {
      take = tieBase + seen < rem;
      seen++;
    }
    if (take) {
      const uint32_t slot = atomicAdd(&S.sCount, 1u);
      cand[slot] = ((unsigned long long)kk << 32) | i;
    }
  }
}
// Both selections of one level in one launch: workgroup 0 the move candidates (block scores),
// workgroup 1 the destination subblocks (subblock scores); they read only scores and loads.
__global__ void __launch_bounds__(1024) k_select_pair(DGBase g, int level, const float* __restrict__ bs,
                                                      const float* __restrict__ sbs, unsigned long long* __restrict__ cand0,
                                                      unsigned long long* __restrict__ cand1) {
  __shared__ SelectShared S;
  if (blockIdx.x == 0)
    selectBody<0>(g, level, bs, cand0, S);
  else
    selectBody<1>(g, level, sbs, cand1, S);
}

// Exact position of every selected candidate (keys are unique): out[rank] = base + position.
// blockIdx.y selects the list: 0 move candidates -> mc, 1 destination subblocks -> dc.
__global__ void __launch_bounds__(256) k_rank_pair(DGBase g, int level, const unsigned long long* __restrict__ cand0,
                                                   const unsigned long long* __restrict__ cand1,
                                                   uint32_t* __restrict__ mc, uint32_t* __restrict__ dc) {
  __shared__ unsigned long long tile[1024];
  const int mode = blockIdx.y;
  const uint32_t lim = g.ctl[CTL_LIM + level];
  const uint32_t k = mode == 0 ? selectCount<0>(g, level, lim) : selectCount<1>(g, level, lim);
  const unsigned long long* cand = mode == 0 ? cand0 : cand1;
  uint32_t* out = mode == 0 ? mc : dc;
  const uint32_t base = mode == 0 ? g.offs[level] : 8 * g.offs[level + 1];
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (blockIdx.x * 256 >= k) return;
  const unsigned long long mine = i < k ? cand[i] : ~0ull;
  uint32_t rank = 0;
  for (uint32_t t0 = 0; t0 < k; t0 += 1024) {
    for (uint32_t j = threadIdx.x; j < 1024; j += 256) tile[j] = t0 + j < k ? cand[t0 + j] : ~0ull;
    __syncthreads();
    const uint32_t m = min(1024u, k - t0);
    for (uint32_t j = 0; j < m; j++) rank += tile[j] < mine;
    __syncthreads();
  }
  if (i < k) out[rank] = base + (uint32_t)(mine & 0xFFFFFFFFu);
}

// Match loop of moveBlocks (fluid_simulation_dcgrid.cu:410-421). The conditions are monotone
// in m (block scores ascending, subblock scores descending), so the matched count is the first
// failing position. Marks of blocks containing destination subblocks (dest[1..M], the
// reference's "next" destination) are applied to the level above, which the next level's
// selection reads. The last level's match also sums the moves of all levels.
__device__ __forceinline__ void totalMoves(const DGBase& g, int level, uint32_t M) {
  uint32_t s = M;
  for (int l = 0; l < level; l++) s += g.ctl[CTL_MOVES + l];
  g.ctl[CTL_TOTAL_MOVES] = s;
}
__global__ void __launch_bounds__(1024) k_match(DGBase g, int level, int last, const uint32_t* __restrict__ mc,
                                                const uint32_t* __restrict__ dc, float* __restrict__ bs,
                                                const float* __restrict__ sbs) {
  __shared__ uint32_t sFail;
  const uint32_t lim = g.ctl[CTL_LIM + level];
  if (lim == 0) {
    if (threadIdx.x == 0) {
      g.ctl[CTL_MOVES + level] = 0;
      if (last) totalMoves(g, level, 0);
    }
    return;
  }
  if (threadIdx.x == 0) sFail = lim;
  __syncthreads();
  for (uint32_t m = threadIdx.x; m < lim; m += blockDim.x) {
    const float a = bs[mc[m]], b = sbs[dc[m]];
    if (!(a >= 0.f && b >= 0.f && a < b)) atomicMin(&sFail, m);
  }
  __syncthreads();
  const uint32_t M = sFail;
  const uint32_t d1 = 8 * g.maxL[level + 1];
  if (threadIdx.x == 0) {
    g.ctl[CTL_MOVES + level] = M;
    g.ctl[CTL_MOVE_LIMIT + level] = (uint32_t)((float)M * 1.2f);
    if (last) totalMoves(g, level, M);
  }
  for (uint32_t j = 1 + threadIdx.x; j <= M; j += blockDim.x)
    if (j < d1) bs[dc[j] / 8] = -FLT_MAX;
}

// k_dcgrid_move_blocks (dcgrid_adaptation.cu:65-90) for the moves of one level; levels are
// applied coarse to fine so a block moved under a parent that moved takes the parent's new
// position.
// The first move launch (hashSlots > 0) also clears the hash tables for the refill when any
// block moves (the cudaMemset of fluid_simulation_dcgrid.cu:327-328): nothing reads the hash
// between the selection and the refill.
template <int BW>
__global__ void __launch_bounds__(256) k_move(DGBase g, int level, const uint32_t* __restrict__ mc,
                                              const uint32_t* __restrict__ dc, uint32_t hashSlots) {
  if (hashSlots && g.ctl[CTL_TOTAL_MOVES] != 0)
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < hashSlots; i += gridDim.x * 256) {
      g.hkey[i] = 0xFFFFFFFFu;
      g.hval[i] = NFU;
    }
  const uint32_t m = blockIdx.x * 256 + threadIdx.x;
  if (m >= g.ctl[CTL_MOVES + level]) return;
  const uint32_t b = mc[m], next = dc[m];
  g.chi[g.par[b]] = NFU;
  g.chi[next] = b;
  g.par[b] = next;
  const uint32_t pb = next / 8;
  const int3 pp = g.pos[pb];
  g.pos[b] = make_int3(2 * pp.x + BW * (int)((next / 4) % 2), 2 * pp.y + BW * (int)((next / 2) % 2),
                       2 * pp.z + BW * (int)(next % 2));
  atomicOr(&g.flg[b], (uint32_t)MOVED);
  atomicOr(&g.flg[pb], (uint32_t)REFINED);
}

// ------------------------------------------------------------------ hash tables
// Open addressing per level, probe stride 127 (dcgrid_utils.cuh:98-140). Slot placement depends
// on the order of concurrent insertions; the content is a deterministic map.
template <int BW>
__device__ __forceinline__ void hashInsert(const DGBase& g, int level, int3 p, uint32_t blockIndex) {
  const uint32_t key = gridHash<BW>(p.x, p.y, p.z);
  const uint32_t o = g.hoff[level], size = g.hsize[level];
  uint32_t slot = key % size;
  const uint32_t slot0 = slot;
  do {
    const uint32_t prev = atomicCAS(&g.hkey[o + slot], 0xFFFFFFFFu, key);
    if (prev == 0xFFFFFFFFu || prev == key) {
      g.hval[o + slot] = blockIndex;
      return;
    }
    slot = (slot + 127) % size;
  } while (slot != slot0);
  atomicOr(&g.ctl[CTL_ERROR], ERR_HASH_FULL);
}
// k_dcgrid_refill_hash_table (dcgrid_adaptation.cu:180-201)
template <int BW>
__global__ void __launch_bounds__(256) k_hash_refill(DGBase g) {
  if (g.ctl[CTL_TOTAL_MOVES] == 0) return;
  const uint32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b >= g.B || g.lev[b] == FREE) return;
  hashInsert<BW>(g, g.lev[b], g.pos[b], b);
}

// ------------------------------------------------------------------ refinement
// Selection of refineSubblocks for level `l = blockIdx.x + 1` (fluid_simulation_dcgrid.cu:449-
// 469): candidates are subblocks of level l with score > threshold in index order; when there are
// more than the limit, the largest indices are kept. Writes the selected list (ascending) and the
// count; children are allocated by rank.
__global__ void __launch_bounds__(1024) k_refine_select(DGBase g, dcg_params P, const float* __restrict__ sbs,
                                                        uint32_t* __restrict__ list) {
  __shared__ uint32_t sh[17];
  __shared__ uint32_t sTotal;
  const int l = blockIdx.x + 1;
  const uint32_t loadLm1 = g.ctl[CTL_LOADS + l - 1], loadL = g.ctl[CTL_LOADS + l];
  const uint32_t lim = min(g.maxL[l - 1] - loadLm1, 8 * loadL - loadLm1);
  // subblocks of free blocks score -FLT_MAX: only the allocated prefix of the slice can qualify
  const uint32_t n = 8 * loadL, base = 8 * g.offs[l];
  const uint32_t T = blockDim.x;
  const uint32_t seg = (n + T - 1) / T;
  const uint32_t lo = min(threadIdx.x * seg, n), hi = min(lo + seg, n);
  // candidates of a segment of 9..32 as a bit mask, read once (unrolled); shorter segments
  // take the plain loop (two short passes)
  constexpr uint32_t REG = 32;
  const bool inRegs = seg > 8 && seg <= REG;
  uint32_t bits = 0, cnt = 0;
  if (lim > 0) {
    if (inRegs) {
#pragma unroll
      for (uint32_t j = 0; j < REG; j++)
        if (lo + j < hi && sbs[base + lo + j] > P.refine_threshold) bits |= 1u << j;
      cnt = __popc(bits);
    } else {
      for (uint32_t i = lo; i < hi; i++) cnt += sbs[base + i] > P.refine_threshold;
    }
  }
  uint32_t total;
  const uint32_t ex = wgExclusiveScan(cnt, sh, &total);
  if (threadIdx.x == 0) sTotal = total;
  __syncthreads();
  total = sTotal;
  const uint32_t take = min(total, lim);
  const uint32_t skip = total - take;  // the smallest indices are dropped
  if (threadIdx.x == 0) {
    g.ctl[CTL_REFINES + l] = take;
    g.ctl[CTL_LOADS_BEFORE + l - 1] = loadLm1;
  }
  if (lim == 0) return;
  uint32_t r = ex;
  if (inRegs) {
    for (uint32_t j = 0; j < REG; j++)
      if ((bits >> j) & 1u) {
        if (r >= skip) list[base + (r - skip)] = base + lo + j;
        r++;
      }
    return;
  }
  for (uint32_t i = lo; i < hi; i++)
    if (sbs[base + i] > P.refine_threshold) {
      if (r >= skip) list[base + (r - skip)] = base + i;
      r++;
    }
}

// k_dcgrid_refine_subblocks (dcgrid_adaptation.cu:145-178) with allocation by rank: the r-th
// selected subblock of level l gets block offs[l-1] + loads[l-1] + r.
template <int BW>
__global__ void __launch_bounds__(256) k_refine_apply(DGBase g, const uint32_t* __restrict__ list) {
  const int l = blockIdx.y + 1;
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  const uint32_t take = g.ctl[CTL_REFINES + l];
  const uint32_t before = g.ctl[CTL_LOADS_BEFORE + l - 1];
  if (r == 0) g.ctl[CTL_LOADS + l - 1] = before + take;
  if (r >= take) return;
  const uint32_t s = list[8 * g.offs[l] + r];
  const uint32_t pb = s / 8;
  const int3 pp = g.pos[pb];
  const int3 cp = make_int3(2 * pp.x + BW * (int)((s / 4) % 2), 2 * pp.y + BW * (int)((s / 2) % 2),
                            2 * pp.z + BW * (int)(s % 2));
  const uint32_t cb = g.offs[l - 1] + before + r;
  hashInsert<BW>(g, l - 1, cp, cb);
  g.pos[cb] = cp;
  g.lev[cb] = (uint8_t)(l - 1);
  g.par[cb] = s;
  g.chi[s] = cb;
  atomicOr(&g.flg[pb], (uint32_t)REFINED);
  atomicOr(&g.flg[cb], (uint32_t)MOVED);
}

// ------------------------------------------------------------------ apron refresh
// k_dcgrid_refresh_apron_indices (dcgrid_structure.cu:30-92) on the compact table: one thread per
// (block, direction). Moved blocks look every direction up again (out-of-domain directions
// point to the block itself); other blocks follow a neighbour that moved, or step into the child
// of a refined coarser neighbour.
template <int BW>
// counter (may be null): the active cells of this step (all cells of allocated blocks, loads
// final after the refinement) are added to it.
__global__ void __launch_bounds__(256) k_refresh(DGBase g, unsigned long long* __restrict__ counter) {
  constexpr int BB = Geo<BW>::BB;
  const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
  if (counter && gid == 0) {
    unsigned long long cells = 0;
    for (int l = 0; l < g.L; l++) cells += (unsigned long long)g.ctl[CTL_LOADS + l] * Geo<BW>::BV;
    *counter += cells;
  }
  const uint32_t b = gid / 27, d = gid % 27;
  if (b >= g.B) return;
  const int level = g.lev[b];
  if (level == FREE) return;
  const int dx = (int)(d / 9) - 1, dy = (int)((d / 3) % 3) - 1, dz = (int)(d % 3) - 1;
  const int3 p0 = g.pos[b];
  // one apron cell of this direction (all cells of a direction lie in one block)
  const int ax = p0.x + (dx < 0 ? -1 : (dx > 0 ? BW : 0));
  const int ay = p0.y + (dy < 0 ? -1 : (dy > 0 ? BW : 0));
  const int az = p0.z + (dz < 0 ? -1 : (dz > 0 ? BW : 0));
  const uint32_t f = g.flg[b];
  int4* e = &g.nd[b * 27 + d];
  uint32_t nb;
  if (f & MOVED) {
    if (d == 13) {
      *e = makeNd(b, 0, 0, 0, 0);
      return;
    }
    const int scale = 1 << level;
    if (ax < 0 || ay < 0 || az < 0 || ax * scale >= g.gx || ay * scale >= g.gy || az * scale >= g.gz) {
      *e = makeNd(b, ND_OUT, 1, 0, 0);  // SPREAD(clamp(i,1,BW)) form, dcgrid_structure.cu:54-60
      return;
    }
    nb = findBlock<BW>(g, ax, ay, az, level);
  } else {
    if (d == 13) return;
    const int4 old = *e;
    if (ndShift(old) == ND_OUT) return;
    const uint32_t prev = ndBlock(old);
    const uint32_t pf = g.flg[prev];
    if (pf & MOVED) {
      nb = findBlock<BW>(g, ax, ay, az, level);
    } else if ((pf & REFINED) && (int)g.lev[prev] > level) {
      const int dl = (int)g.lev[prev] - level;
      const int3 pp = g.pos[prev];
      const int lx = (ax >> dl) - pp.x, ly = (ay >> dl) - pp.y, lz = (az >> dl) - pp.z;
      const int t = (((lx >> (BB - 1)) & 1) << 2) | (((ly >> (BB - 1)) & 1) << 1) | ((lz >> (BB - 1)) & 1);
      nb = g.chi[8 * prev + t];
      if (nb == NFU) return;
    } else {
      return;  // neither block moved: the descriptor is still exact
    }
  }
  const int sh = (int)g.lev[nb] - level;
  const int3 np = g.pos[nb];
  *e = makeNd(nb, (uint32_t)sh, (np.x << sh) - p0.x, (np.y << sh) - p0.y, (np.z << sh) - p0.z);
}

}  // namespace dcg

// oracle_dcgrid.cpp — CPU restatement of the reference's DCGrid step (test infrastructure
// only; see oracle.h).
//
// Follows src/dcgrid/fluid_simulation_dcgrid.cu (planner, reset/init, step methods, move and
// refine selection), dcgrid_structure.cu, dcgrid_adaptation.cu, dcgrid_fluid.cu,
// dcgrid_multigrid_solver.cu and dcgrid_utils.cuh, generalised from the reference's block width
// 4 to BW in {4, 8} (the refinement unit "subblock" is the top Morton triple of a block; the
// restriction group is the lowest triple — at BW=4 both are 2^3 cells).
//
// Where the reference is nondeterministic the oracle fixes a rule (the HIP library applies the
// same rule):
//  * move candidates: total order (negative last, score ascending, index ascending); destination
//    subblocks: (score descending, index ascending) — the reference's comparator at
//    fluid_simulation_dcgrid.cu:350-358 is not a strict weak order;
//  * refine allocation: selected subblocks in ascending index order receive consecutive free
//    blocks (the reference allocates with atomicAdd, dcgrid_utils.cuh:118);
//  * moves are applied level by level from coarse to fine, so a block moved under a parent that
//    moves in the same step takes the parent's new position (the reference's single launch,
//    dcgrid_adaptation.cu:65-90, races on the parent position);
//  * hash tables are filled in ascending block order and compared as maps.
#include <omp.h>

#include <algorithm>
#include <numeric>

#include "oracle_common.hpp"
#include "oracle_render.hpp"

namespace orc {

Sim* makeDCGrid(const int32_t size[3], uint64_t maxBlocks, const dcg_params& P, int precision,
                int* status, std::string* err);

template <class R>
struct DCGrid : Sim {
  // Layout constants (dcgrid.h:51-67 for BW=4).
  int BW = 4, BB = 2, A = 6, A2 = 36, A3 = 216, BV = 64, SBB = 3;
  int gx = 0, gy = 0, gz = 0, L = 0, S = 0;
  size_t B = 0, C = 0, H = 0;
  size_t maxL[DCG_MAX_LEVELS], fullL[DCG_MAX_LEVELS], offs[DCG_MAX_LEVELS];
  size_t loads[DCG_MAX_LEVELS], moveLimit[DCG_MAX_LEVELS];
  size_t hsize[DCG_MAX_LEVELS], hoff[DCG_MAX_LEVELS];

  std::vector<uint32_t> hkey;
  std::vector<size_t> hval;
  std::vector<int32_t> pos;  // 3B
  std::vector<uint8_t> lev, flg;
  std::vector<size_t> par, chi, cidx, freeIdx;
  std::vector<R> den, tem, flu, dive, pre, t_pre, t_den, t_tem;
  std::vector<R> vap, t_vap;  // water vapor (moist atmosphere only)
  std::vector<V3<R>> vel, t_vel, vort;
  std::vector<float> sbs, bs;  // subblock scores (8B + 1 pad), block scores (B)
  std::vector<size_t> touched;

  int kind() const override { return DCG_KIND_DCGRID; }
  // A value stored to a field (dcg_params.storage; FP64 values round through FP32).
  R st(R v) const { return P.storage ? (R)roundStorage((float)v, P.storage) : v; }
  V3<R> st(const V3<R>& v) const { return P.storage ? V3<R>{st(v.x), st(v.y), st(v.z)} : v; }

  // ---------------------------------------------------------------- planner
  // FluidSimulationDCGrid ctor, fluid_simulation_dcgrid.cu:9-140.
  int create(const int32_t size[3], uint64_t maxBlocks, const dcg_params& p) {
    if (int e = plan(size, maxBlocks, p)) return e;
    return allocate();
  }
  int plan(const int32_t size[3], uint64_t maxBlocks, const dcg_params& p) {
    P = p;
    gx = size[0];
    gy = size[1];
    gz = size[2];
    if (gx <= 0 || gy <= 0 || gz <= 0 || gx != P.gx || gy != P.gy || gz != P.gz)
      return fail(DCG_E_ARG, "size must be positive and equal params gx,gy,gz");
    if (P.block_width != 4 && P.block_width != 8) return fail(DCG_E_UNSUPPORTED, "block_width must be 4 or 8");
    BW = P.block_width;
    BB = BW == 4 ? 2 : 3;
    A = BW + 2;
    A2 = A * A;
    A3 = A2 * A;
    BV = BW * BW * BW;
    SBB = 3 * (BB - 1);
    const int minDim = std::min(gx, std::min(gy, gz));
    int cellSize = 2;
    L = 1;
    while (gx % cellSize == 0 && gy % cellSize == 0 && gz % cellSize == 0 && cellSize * BW <= minDim) {
      L++;
      cellSize *= 2;
    }
    if (P.max_levels > 0 && L > P.max_levels) L = P.max_levels;
    if (L > DCG_MAX_LEVELS) return fail(DCG_E_UNSUPPORTED, "too many levels");
    const int top = BW << (L - 1);
    if (gx % top || gy % top || gz % top)
      return fail(DCG_E_UNSUPPORTED, "domain must be a multiple of the coarsest block extent");
    for (int l = 0, cs = 1; l < L; l++, cs *= 2)
      fullL[l] = (size_t)iDivUp(gx, cs * BW) * iDivUp(gy, cs * BW) * iDivUp(gz, cs * BW);
    B = maxBlocks;
    if (B < fullL[L - 1]) return fail(DCG_E_BUDGET, "too few blocks to fill lowest resolution");
    maxL[L - 1] = fullL[L - 1];
    size_t left = B - maxL[L - 1];
    for (int l = L - 2; l >= 0; l--) {
      maxL[l] = std::min(left / (size_t)(l + 1), fullL[l]);
      left -= maxL[l];
    }
    if (maxL[0] == 0) return fail(DCG_E_BUDGET, "too few blocks to reach highest resolution");
    for (int l = 0; l < L; l++)
      if (maxL[l] == 0) return fail(DCG_E_BUDGET, "a level between the finest and the coarsest gets no blocks");
    offs[0] = 0;
    for (int l = 1; l < L; l++) offs[l] = offs[l - 1] + maxL[l - 1];
    S = 0;
    for (int l = 0; l < L; l++)
      if (maxL[l] < fullL[l]) S = l + 1;
    H = 4 * B;
    for (int l = 0; l < L; l++) hsize[l] = 4 * maxL[l];
    hoff[0] = 0;
    for (int l = 1; l < L; l++) hoff[l] = hoff[l - 1] + hsize[l - 1];
    C = B * (size_t)BV;
    for (int l = 0; l < L; l++) loads[l] = maxL[l] == fullL[l] ? maxL[l] : 0;
    return DCG_OK;
  }
  int allocate() {
    hkey.assign(H, HASH_EMPTY);
    hval.assign(H, NF);
    pos.assign(3 * B, 0);
    lev.assign(B, 0xFF);
    flg.assign(B, 0);
    par.assign(B, NF);
    chi.assign(8 * B, NF);
    cidx.assign(B * (size_t)A3, 0);
    freeIdx.resize(B);
    den.assign(C, 0);
    tem.assign(C, 0);
    flu.assign(C, 0);
    dive.assign(C, 0);
    pre.assign(C, 0);
    t_pre.assign(C, 0);
    t_den.assign(C, 0);
    t_tem.assign(C, 0);
    vel.assign(C, V3<R>{0, 0, 0});
    t_vel.assign(C, V3<R>{0, 0, 0});
    vort.assign(C, V3<R>{0, 0, 0});
    if (P.moist) {
      vap.assign(C, 0);
      t_vap.assign(C, 0);
    }
    sbs.assign(8 * B + 1, -FLT_MAX);
    bs.assign(B, -FLT_MAX);
    return reset();
  }

  // ---------------------------------------------------------------- index helpers
  uint32_t loc(int x, int y, int z) const { return localIndex(x, y, z, BB); }
  size_t& ap(size_t b, int i, int j, int k) { return cidx[b * A3 + (size_t)i * A2 + (size_t)j * A + k]; }
  int px(size_t b) const { return pos[3 * b]; }
  int py(size_t b) const { return pos[3 * b + 1]; }
  int pz(size_t b) const { return pos[3 * b + 2]; }
  // USE_CELL_POS (dcgrid_utils.cuh:74-81): position of a cell in level-local coordinates.
  void cellPos(size_t c, int& x, int& y, int& z) const {
    const size_t b = c / BV;
    int lx, ly, lz;
    localPos((uint32_t)(c % BV), BB, lx, ly, lz);
    x = px(b) | lx;
    y = py(b) | ly;
    z = pz(b) | lz;
  }
  // Apron index of a cell (USE_APRON_INDEX, dcgrid_utils.cuh:50-58).
  int apronOf(size_t c) const {
    int lx, ly, lz;
    localPos((uint32_t)(c % BV), BB, lx, ly, lz);
    return A2 * (1 + lx) + A * (1 + ly) + (1 + lz);
  }
  size_t subblockOf(size_t c) const { return c >> SBB; }  // refinement unit of a cell
  bool isLeaf(size_t c) const { return chi[subblockOf(c)] == NF; }

  // getBlockIndex, dcgrid_utils.cuh:169-199
  size_t getBlockIndex(int x, int y, int z, int level) const {
    if (level >= S) {
      const int extent = BW << level;
      const int rx = iDivUp(gx, extent), ry = iDivUp(gy, extent), rz = iDivUp(gz, extent);
      const int bx = x / BW, by = y / BW, bz = z / BW;
      return bx >= rx || by >= ry || bz >= rz ? NF : offs[level] + ((size_t)bx * ry + by) * rz + bz;
    }
    const uint32_t key = gridHash(x, y, z, BW);
    const size_t o = hoff[level];
    size_t slot = key % hsize[level];
    const size_t slot0 = slot;
    do {
      if (hkey[o + slot] == key) return hval[o + slot];
      if (hkey[o + slot] == HASH_EMPTY) return NF;
      slot = (slot + 127) % hsize[level];
    } while (slot != slot0);
    return NF;
  }
  // getBlockIndexDeep, dcgrid_utils.cuh:201-232: finest existing block at level >= `level`
  // containing `p` (level-local coordinates); updates p and level.
  size_t getBlockIndexDeep(int& x, int& y, int& z, int& level) const {
    if (level < S)
      for (uint32_t key = gridHash(x, y, z, BW); level < S; level++, key >>= 3, x /= 2, y /= 2, z /= 2) {
        const size_t o = hoff[level];
        size_t slot = key % hsize[level];
        const size_t slot0 = slot;
        bool stop = false;
        do {
          if (hkey[o + slot] == key) return hval[o + slot];
          if (hkey[o + slot] == HASH_EMPTY) {
            stop = true;
            break;
          }
          slot = (slot + 127) % hsize[level];
        } while (slot != slot0);
        (void)stop;
      }
    // position >>= sparseLevels - level (always a shift by 0 here: level == S)
    level = S;
    const int extent = BW << level;
    const int rx = iDivUp(gx, extent), ry = iDivUp(gy, extent), rz = iDivUp(gz, extent);
    const int bx = x / BW, by = y / BW, bz = z / BW;
    return bx >= rx || by >= ry || bz >= rz ? NF : offs[level] + ((size_t)bx * ry + by) * rz + bz;
  }
  // insertBlock, dcgrid_utils.cuh:98-140 (sequential: atomics are plain operations).
  size_t insertBlock(int x, int y, int z, int level) {
    const uint32_t key = gridHash(x, y, z, BW);
    const size_t o = hoff[level];
    size_t slot = key % hsize[level];
    const size_t slot0 = slot;
    do {
      const uint32_t prev = hkey[o + slot];
      if (prev == HASH_EMPTY) hkey[o + slot] = key;
      if (prev == key) return hval[o + slot];
      if (prev == HASH_EMPTY) {
        const size_t allocated = loads[level]++;
        if (allocated >= maxL[level]) {
          loads[level]--;
          return NF;
        }
        const size_t b = freeIdx[offs[level] + allocated];
        hval[o + slot] = b;
        pos[3 * b] = x;
        pos[3 * b + 1] = y;
        pos[3 * b + 2] = z;
        lev[b] = (uint8_t)level;
        return b;
      }
      slot = (slot + 127) % hsize[level];
    } while (slot != slot0);
    err = "hash table full";
    return NF;
  }

  // ---------------------------------------------------------------- reset / init
  // reset(), fluid_simulation_dcgrid.cu:212-261
  int reset() override {
    std::fill(hkey.begin(), hkey.end(), HASH_EMPTY);
    std::fill(hval.begin(), hval.end(), NF);
    std::fill(pos.begin(), pos.end(), 0);
    std::fill(lev.begin(), lev.end(), 0xFF);
    std::fill(flg.begin(), flg.end(), 0);
    std::fill(par.begin(), par.end(), NF);
    std::fill(chi.begin(), chi.end(), NF);
    std::fill(cidx.begin(), cidx.end(), 0);
    std::fill(den.begin(), den.end(), (R)0);
    std::fill(tem.begin(), tem.end(), (R)0);
    std::fill(vel.begin(), vel.end(), V3<R>{0, 0, 0});
    std::fill(flu.begin(), flu.end(), (R)0);
    std::fill(dive.begin(), dive.end(), (R)0);
    std::fill(pre.begin(), pre.end(), (R)0);
    std::fill(t_pre.begin(), t_pre.end(), (R)0);
    for (int l = 0; l < L; l++) {
      loads[l] = maxL[l] == fullL[l] ? maxL[l] : 0;
      moveLimit[l] = 0;
    }
    std::iota(freeIdx.begin(), freeIdx.end(), (size_t)0);
    return init();
  }
  // init(), fluid_simulation_dcgrid.cu:190-210
  int init() override {
    initApronIndices();
    for (int level = S; level < L; level++) activateLevel(level);
    for (int i = 0; i < 5; i++)
      if (int e = adaptTopology()) return e;
    return DCG_OK;
  }
  // k_dcgrid_init_apron_indices, dcgrid_structure.cu:6-28
  void initApronIndices() {
#pragma omp parallel for schedule(static)
    for (long long bl = 0; bl < (long long)B; bl++) {
      const size_t b = (size_t)bl;
      for (int i = 0; i < A; i++)
        for (int j = 0; j < A; j++)
          for (int k = 0; k < A; k++) ap(b, i, j, k) = NF;
      for (int x = 0; x < BW; x++)
        for (int y = 0; y < BW; y++)
          for (int z = 0; z < BW; z++) ap(b, 1 + x, 1 + y, 1 + z) = b * BV + loc(x, y, z);
    }
  }
  // k_dcgrid_activate_level, dcgrid_structure.cu:104-178 (one dense level)
  void activateLevel(int level) {
    const int scale = 1 << level;
    const int extent = BW << level;
    const int rx = iDivUp(gx, extent), ry = iDivUp(gy, extent), rz = iDivUp(gz, extent);
    // The apron of a dense block reads the interior of its neighbours, which only
    // initApronIndices writes, so blocks are independent.
#pragma omp parallel for schedule(static)
    for (int bx = 0; bx < rx; bx++)
      for (int by = 0; by < ry; by++)
        for (int bz = 0; bz < rz; bz++) {
          const int x = BW * bx, y = BW * by, z = BW * bz;
          if (x * scale >= gx || y * scale >= gy || z * scale >= gz) continue;
          const size_t b = getBlockIndex(x, y, z, level);
          if (b == NF || b >= offs[level] + maxL[level]) continue;
          pos[3 * b] = x;
          pos[3 * b + 1] = y;
          pos[3 * b + 2] = z;
          lev[b] = (uint8_t)level;
          if (level < L - 1) {
            // parent subblock: top Morton triple of the block inside its parent
            // (dcgrid_structure.cu:119-123 for BW=4)
            const size_t pb = getBlockIndex(x / 2, y / 2, z / 2, level + 1);
            par[b] = 8 * pb + 4 * (size_t)((x / BW) & 1) + 2 * (size_t)((y / BW) & 1) + (size_t)((z / BW) & 1);
          }
          if (level - 1 >= S)
            for (int c = 0; c < 8; c++)
              chi[8 * b + c] = getBlockIndex(2 * x + BW * ((c / 4) % 2), 2 * y + BW * ((c / 2) % 2),
                                             2 * z + BW * (c % 2), level - 1);
          const int i0 = x > 0 ? 0 : 1, i1 = A - ((x + BW) * scale >= gx ? 1 : 0);
          const int j0 = y > 0 ? 0 : 1, j1 = A - ((y + BW) * scale >= gy ? 1 : 0);
          const int k0 = z > 0 ? 0 : 1, k1 = A - ((z + BW) * scale >= gz ? 1 : 0);
          for (int i = i0; i < i1; i++)
            for (int j = j0; j < j1; j++)
              for (int k = k0; k < k1; k++) {
                const size_t nb = getBlockIndex(x + i - 1, y + j - 1, z + k - 1, level);
                ap(b, i, j, k) = ap(nb, 1 + ((i - 1 + BW) % BW), 1 + ((j - 1 + BW) % BW), 1 + ((k - 1 + BW) % BW));
              }
          for (int i = 0; i < A; i++) {
            const int ic = clampi(i, 1, BW);
            for (int j = 0; j < A; j++) {
              const int jc = clampi(j, 1, BW);
              for (int k = 0; k < A; k++) {
                const int kc = clampi(k, 1, BW);
                if (ap(b, i, j, k) == NF) ap(b, i, j, k) = ap(b, ic, jc, kc);
              }
            }
          }
          for (int i = 1; i <= BW; i++)
            for (int j = 1; j <= BW; j++)
              for (int k = 1; k <= BW; k++) {
                const size_t c = ap(b, i, j, k);
                den[c] = 0;
                tem[c] = st((R)P.t_ambient);
                vel[c] = V3<R>{0, 0, 0};
                flu[c] = st((R)cellFluidity(P, x + i - 1, y + j - 1, z + k - 1, scale));
              }
        }
  }

  // ---------------------------------------------------------------- scores
  bool levelBlocked(int level) const {
    return (level == 0 && loads[0] == fullL[0]) || (level > 0 && loads[level - 1] == fullL[level - 1]);
  }
  // k_dcgrid_calc_vorticity, dcgrid_fluid.cu:146-172 (all cells of allocated blocks)
  void calcVorticity() {
    const R rdx = (R)P.rdx;
#pragma omp parallel for schedule(static)
    for (long long bl = 0; bl < (long long)B; bl++) {
      const size_t b = (size_t)bl;
      if (lev[b] == 0xFF) continue;
      const int scale = 1 << lev[b];
      const R alpha = (R).5f * rdx / (R)scale;
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        const int a = apronOf(c);
        const size_t* ci = &cidx[b * A3];
        const R wl = flu[ci[a - A2]], wr = flu[ci[a + A2]], wd = flu[ci[a - A]], wu = flu[ci[a + A]];
        const R wb = flu[ci[a - 1]], wf = flu[ci[a + 1]];
        const V3<R> vl = vel[ci[a - A2]], vr = vel[ci[a + A2]], vd = vel[ci[a - A]], vu = vel[ci[a + A]];
        const V3<R> vb = vel[ci[a - 1]], vf = vel[ci[a + 1]];
        vort[c] = st(mul(V3<R>{(wu * vu.z - wd * vd.z) - (wf * vf.y - wb * vb.y),
                               (wf * vf.x - wb * vb.x) - (wr * vr.z - wl * vl.z),
                               (wr * vr.y - wl * vl.y) - (wu * vu.x - wd * vd.x)},
                         alpha));
      }
    }
  }
  // Feature score of one cell (extension): max(|grad density|/g, |vorticity|/w).
  float featureScore(size_t b, size_t c) const {
    const int scale = 1 << lev[b];
    const R alpha = (R).5f * (R)P.rdx / (R)scale;
    const int a = apronOf(c);
    const size_t* ci = &cidx[b * A3];
    const R gxv = alpha * (den[ci[a + A2]] - den[ci[a - A2]]);
    const R gyv = alpha * (den[ci[a + A]] - den[ci[a - A]]);
    const R gzv = alpha * (den[ci[a + 1]] - den[ci[a - 1]]);
    const R gm = std::sqrt(gxv * gxv + gyv * gyv + gzv * gzv);
    const V3<R>& w = vort[c];
    const R wm = std::sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
    const float fg = (float)gm / P.grad_threshold;
    const float fw = (float)wm / P.vort_threshold;
    return fg > fw ? fg : fw;
  }
  // k_dcgrid_calc_subblock_scores, dcgrid_adaptation.cu:10-40
  void calcSubblockScores() {
    const float hb = (float)(BW / 2), qb = (float)(BW / 4);
    const float cx = .5f * (float)gx, cy = .45f * (float)gy, cz = .5f * (float)gz;
#pragma omp parallel for schedule(static)
    for (long long sl = 0; sl < (long long)(8 * B); sl++) {
      const size_t s = (size_t)sl, b = s / 8;
      const uint8_t level = lev[b];
      if (chi[s] != NF || level == 0xFF || levelBlocked(level)) {
        sbs[s] = -FLT_MAX;
        continue;
      }
      if (P.score_mode == DCG_SCORE_FEATURES) {
        const size_t c0 = b * BV + (s % 8) * (size_t)(BV / 8);
        float m = 0.f;
        for (int k = 0; k < BV / 8; k++) {
          const float f = featureScore(b, c0 + k);
          m = f > m ? f : m;
        }
        sbs[s] = m;
        continue;
      }
      const float sc = (float)(1 << level);
      const float qx = sc * (((float)px(b) + hb * (float)((s >> 2) & 1)) + qb);
      const float qy = sc * (((float)py(b) + hb * (float)((s >> 1) & 1)) + qb);
      const float qz = sc * (((float)pz(b) + hb * (float)(s & 1)) + qb);
      const float dx = qx - cx, dy = qy - cy, dz = qz - cz;
      const float d = sqrtf(dx * dx + dy * dy + dz * dz);
      sbs[s] = d < .2f * (float)gx ? 0.f : (float)gx / d;
    }
    sbs[8 * B] = -FLT_MAX;  // the reference reads one past the end for the last block
  }
  // k_dcgrid_accumulate_subblock_scores, dcgrid_adaptation.cu:42-63 (keeps the s[8] term)
  void accumulateSubblockScores() {
    for (size_t b = 0; b < B; b++) {
      bs[b] = -FLT_MAX;
      const uint8_t level = lev[b];
      if (level == 0xFF || levelBlocked(level)) continue;
      const float* s = &sbs[8 * b];
      if (s[0] > 0.f && s[1] > 0.f && s[2] > 0.f && s[3] > 0.f && s[4] > 0.f && s[5] > 0.f && s[6] > 0.f &&
          s[7] > 0.f)
        bs[b] = .125f * (s[0] + s[1] + s[2] + s[3] + s[4] + s[5] + s[6] + s[7] + s[8]);
    }
  }

  // ---------------------------------------------------------------- topology
  // moveBlocks, fluid_simulation_dcgrid.cu:348-437 (+ k_dcgrid_move_blocks)
  void moveBlocks() {
    calcSubblockScores();
    accumulateSubblockScores();
    std::fill(flg.begin(), flg.end(), 0);
    auto blockLess = [&](size_t a, size_t b) {
      const float sa = bs[a], sb = bs[b];
      const bool na = sa < 0.f, nb = sb < 0.f;
      if (na != nb) return nb;  // non-negative first
      if (!na && sa != sb) return sa < sb;
      return a < b;
    };
    auto subLess = [&](size_t a, size_t b) {
      const float sa = sbs[a], sb = sbs[b];
      if (sa != sb) return sa > sb;
      return a < b;
    };
    std::vector<std::vector<size_t>> mv(L), dst(L);
    for (int l = 0; l < L - 1; l++) {
      const size_t d0 = maxL[l], d1 = 8 * maxL[l + 1];
      size_t lim = std::min(std::min(d0, d1), std::min(loads[l], fullL[l] - loads[l]));
      if (moveLimit[l] > 0) lim = std::min(lim, moveLimit[l]);
      if (lim == 0) continue;
      std::vector<size_t> mc(d0), dc(d1);
      std::iota(mc.begin(), mc.end(), offs[l]);
      std::iota(dc.begin(), dc.end(), 8 * offs[l + 1]);
      std::sort(mc.begin(), mc.end(), blockLess);
      std::sort(dc.begin(), dc.end(), subLess);
      size_t m = 0;
      while (m < lim && bs[mc[m]] >= 0.f && sbs[dc[m]] >= 0.f && bs[mc[m]] < sbs[dc[m]]) {
        m++;
        // "Prevent block containing subblock from being moved" — marks the next destination
        // (reference behaviour, fluid_simulation_dcgrid.cu:417); skipped past the end.
        if (m < d1) bs[dc[m] / 8] = -FLT_MAX;
      }
      moveLimit[l] = (size_t)((float)m * 1.2f);
      mv[l].assign(mc.begin(), mc.begin() + m);
      dst[l].assign(dc.begin(), dc.begin() + m);
    }
    // k_dcgrid_move_blocks, dcgrid_adaptation.cu:65-90, applied coarse to fine.
    for (int l = L - 2; l >= 0; l--)
      for (size_t i = 0; i < mv[l].size(); i++) {
        const size_t b = mv[l][i], next = dst[l][i];
        chi[par[b]] = NF;
        chi[next] = b;
        par[b] = next;
        const size_t pb = next / 8;
        pos[3 * b] = 2 * pos[3 * pb] + BW * (int)((next / 4) % 2);
        pos[3 * b + 1] = 2 * pos[3 * pb + 1] + BW * (int)((next / 2) % 2);
        pos[3 * b + 2] = 2 * pos[3 * pb + 2] + BW * (int)(next % 2);
        flg[b] |= 1;   // blockMoved
        flg[pb] |= 2;  // blockRefined
      }
    touched.clear();
    for (int l = 0; l < L - 1; l++) touched.insert(touched.end(), mv[l].begin(), mv[l].end());
  }
  // cudaMemset + k_dcgrid_refill_hash_table, fluid_simulation_dcgrid.cu:326-332,
  // dcgrid_adaptation.cu:180-201
  int refillHash() {
    std::fill(hkey.begin(), hkey.end(), HASH_EMPTY);
    std::fill(hval.begin(), hval.end(), NF);
    for (size_t b = 0; b < B; b++) {
      if (lev[b] == 0xFF) continue;
      const int level = lev[b];
      const uint32_t key = gridHash(px(b), py(b), pz(b), BW);
      const size_t o = hoff[level];
      size_t slot = key % hsize[level];
      const size_t slot0 = slot;
      bool ok = false;
      do {
        const uint32_t prev = hkey[o + slot];
        if (prev == HASH_EMPTY || prev == key) {
          hkey[o + slot] = key;
          hval[o + slot] = b;
          ok = true;
          break;
        }
        slot = (slot + 127) % hsize[level];
      } while (slot != slot0);
      if (!ok) return fail(DCG_E_HASH_FULL, "hash table fully loaded");
    }
    return DCG_OK;
  }
  // refineSubblocks, fluid_simulation_dcgrid.cu:439-483 (+ k_dcgrid_refine_subblocks,
  // dcgrid_adaptation.cu:145-178)
  int refineSubblocks() {
    calcSubblockScores();
    size_t snap[DCG_MAX_LEVELS];
    std::copy(loads, loads + L, snap);
    std::vector<size_t> list;
    for (int l = 1; l < L; l++) {
      const size_t lim = std::min(maxL[l - 1] - snap[l - 1], 8 * snap[l] - snap[l - 1]);
      if (lim == 0) continue;
      const size_t start = 8 * offs[l], end = start + 8 * maxL[l];
      std::vector<size_t> cand;
      for (size_t i = start; i < end; i++)
        if (sbs[i] > P.refine_threshold) cand.push_back(i);
      // nth_element(greater) on the indices keeps the lim largest indices
      if (cand.size() > lim) cand.erase(cand.begin(), cand.end() - (ptrdiff_t)lim);
      list.insert(list.end(), cand.begin(), cand.end());
    }
    for (size_t s : list) {
      const size_t pb = s / 8;
      const int level = lev[pb];
      const int cx = 2 * px(pb) + BW * (int)((s / 4) % 2);
      const int cy = 2 * py(pb) + BW * (int)((s / 2) % 2);
      const int cz = 2 * pz(pb) + BW * (int)(s % 2);
      const size_t cb = insertBlock(cx, cy, cz, level - 1);
      if (cb == NF) return fail(DCG_E_HASH_FULL, "no space left to refine");
      par[cb] = s;
      chi[s] = cb;
      flg[pb] |= 2;
      flg[cb] |= 1;
      touched.push_back(cb);
    }
    return DCG_OK;
  }
  // k_dcgrid_refresh_apron_indices, dcgrid_structure.cu:30-92 (border entries of every block)
  void refreshApron() {
#pragma omp parallel for schedule(dynamic, 64)
    for (long long bl = 0; bl < (long long)B; bl++) {
      const size_t b = (size_t)bl;
      if (lev[b] == 0xFF) continue;
      const int level = lev[b];
      const int x0 = px(b), y0 = py(b), z0 = pz(b);
      for (int i = 0; i < A; i++)
        for (int j = 0; j < A; j++)
          for (int k = 0; k < A; k++) {
            if (i % (A - 1) != 0 && j % (A - 1) != 0 && k % (A - 1) != 0) continue;
            size_t& e = ap(b, i, j, k);
            size_t nbk = NF;
            if (flg[b] & 1) {
              int nx = x0 + i - 1, ny = y0 + j - 1, nz = z0 + k - 1;
              const int scale = 1 << level;
              if (nx < 0 || ny < 0 || nz < 0 || nx * scale >= gx || ny * scale >= gy || nz * scale >= gz) {
                // SPREAD(clamp(i,1,BW)) keeps the low bits of the apron coordinate
                e = b * BV + loc(clampi(i, 1, BW) & (BW - 1), clampi(j, 1, BW) & (BW - 1),
                                 clampi(k, 1, BW) & (BW - 1));
                continue;
              }
              int nl = level;
              nbk = getBlockIndexDeep(nx, ny, nz, nl);
            } else {
              const size_t prev = e / BV;
              const uint8_t cf = flg[prev];
              if (cf & 1) {
                int nx = x0 + i - 1, ny = y0 + j - 1, nz = z0 + k - 1, nl = level;
                nbk = getBlockIndexDeep(nx, ny, nz, nl);
              } else if ((cf & 2) && lev[prev] > level) {
                nbk = chi[e >> SBB];
              }
            }
            if (nbk == NF) continue;
            const int s = 1 << (lev[nbk] - level);
            const int x = (x0 + i - 1) / s - px(nbk);
            const int y = (y0 + j - 1) / s - py(nbk);
            const int z = (z0 + k - 1) / s - pz(nbk);
            e = nbk * BV + loc(x & (BW - 1), y & (BW - 1), z & (BW - 1));
          }
    }
  }
  // Parent apron stencil shared by propagate and prolongate (dcgrid_adaptation.cu:104-126,
  // dcgrid_multigrid_solver.cu:51-63). Corner order 000,001,010,100,011,101,110,111.
  void parentStencil(size_t c, size_t out[8]) {
    const size_t b = c / BV;
    const size_t* pa = &cidx[(par[b] / 8) * A3];
    int x, y, z;
    cellPos(c, x, y, z);
    const int qx = (x / 2) % BW, qy = (y / 2) % BW, qz = (z / 2) % BW;
    const long id = (long)A2 * (1 + qx) + (long)A * (1 + qy) + (1 + qz);
    const long i = x % 2 ? A2 : -A2, j = y % 2 ? A : -A, k = z % 2 ? 1 : -1;
    out[0] = pa[id];
    out[1] = pa[id + k];
    out[2] = pa[id + j];
    out[3] = pa[id + i];
    out[4] = pa[id + j + k];
    out[5] = pa[id + i + k];
    out[6] = pa[id + i + j];
    out[7] = pa[id + i + j + k];
  }
  // k_dcgrid_propagate_values, dcgrid_adaptation.cu:92-143
  // temperature is advected and drives buoyancy: sphere source or moist atmosphere
  bool thermo() const { return P.source_mode == DCG_SOURCE_SPHERE_BUOYANT || P.moist; }
  void propagate(int level) {
    const bool src = thermo();
    const R w27 = (R)(27.f / 64.f), w9 = (R)(9.f / 64.f), w3 = (R)(3.f / 64.f), w1 = (R)(1.f / 64.f);
#pragma omp parallel for schedule(static)
    for (long long tl = 0; tl < (long long)touched.size(); tl++) {
      const size_t b = touched[(size_t)tl];
      if (lev[b] != level || par[b] == NF) continue;
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        size_t q[8];
        parentStencil(c, q);
        den[c] = st(w27 * den[q[0]] + w9 * (den[q[1]] + den[q[2]] + den[q[3]]) +
                    w3 * (den[q[4]] + den[q[5]] + den[q[6]]) + w1 * den[q[7]]);
        if (src)
          tem[c] = st(w27 * tem[q[0]] + w9 * (tem[q[1]] + tem[q[2]] + tem[q[3]]) +
                      w3 * (tem[q[4]] + tem[q[5]] + tem[q[6]]) + w1 * tem[q[7]]);
        if (src && P.moist)
          vap[c] = st(w27 * vap[q[0]] + w9 * (vap[q[1]] + vap[q[2]] + vap[q[3]]) +
                      w3 * (vap[q[4]] + vap[q[5]] + vap[q[6]]) + w1 * vap[q[7]]);
        const V3<R> a = mul(vel[q[0]], w27);
        const V3<R> bb = mul(add(add(vel[q[1]], vel[q[2]]), vel[q[3]]), w9);
        const V3<R> cc = mul(add(add(vel[q[4]], vel[q[5]]), vel[q[6]]), w3);
        const V3<R> dd = mul(vel[q[7]], w1);
        vel[c] = st(add(add(add(a, bb), cc), dd));
        int x, y, z;
        cellPos(c, x, y, z);
        flu[c] = st((R)cellFluidity(P, x, y, z, 1 << level));
      }
    }
  }
  // adaptTopology, fluid_simulation_dcgrid.cu:320-346
  int adaptTopology() override {
    if (P.score_mode == DCG_SCORE_FEATURES) calcVorticity();
    moveBlocks();
    if (!touched.empty())
      if (int e = refillHash()) return e;
    if (int e = refineSubblocks()) return e;
    if (!touched.empty()) {
      refreshApron();
      for (int level = L - 2; level >= 0; level--) propagate(level);
    }
    return DCG_OK;
  }

  // ---------------------------------------------------------------- fluid
  // accumulate<T> for one level, dcgrid_structure.cu:188-222
  template <class T, class Z>
  void restrictLevel(std::vector<T>& ch, int level, Z zero) {
    const size_t groupsPerBlock = BV / 8;
#pragma omp parallel for schedule(static)
    for (long long bl = (long long)offs[level]; bl < (long long)(offs[level] + maxL[level]); bl++) {
      const size_t b = (size_t)bl;
      if (lev[b] != level || par[b] == NF) continue;
      const size_t ps = par[b];
      for (size_t g = 0; g < groupsPerBlock; g++) {
        const size_t parentCell = (ps / 8) * BV + ((ps % 8) << SBB) + g;
        T acc = zero;
        const size_t c0 = b * BV + 8 * g;
        for (int i = 0; i < 8; i++) acc = addT(acc, ch[c0 + i]);
        ch[parentCell] = st(mulT(acc, (R).125f));
      }
    }
  }
  static R addT(R a, R b) { return a + b; }
  static V3<R> addT(const V3<R>& a, const V3<R>& b) { return add(a, b); }
  static R mulT(R a, R s) { return a * s; }
  static V3<R> mulT(const V3<R>& a, R s) { return mul(a, s); }
  template <class T, class Z>
  void accumulate(std::vector<T>& ch, Z zero) {
    for (int level = 0; level < L - 1; level++) restrictLevel(ch, level, zero);
  }

  // INIT_SAMPLE, dcgrid_fluid.cu:7-72
  struct Sample {
    size_t i[8];
    R w[8];
    int x0, y0, z0, scale;
    bool zero;
  };
  void sample(R px_, R py_, R pz_, Sample& s) const {
    int ix = clampi((int)std::floor(px_), 0, gx - 1);
    int iy = clampi((int)std::floor(py_), 0, gy - 1);
    int iz = clampi((int)std::floor(pz_), 0, gz - 1);
    int level = 0;
    const size_t b = getBlockIndexDeep(ix, iy, iz, level);
    const int scale = 1 << level;
    const R scaleInv = (R)1 / (R)scale;
    const R x = px_ * scaleInv - (R).5, y = py_ * scaleInv - (R).5, z = pz_ * scaleInv - (R).5;
    const R xf = std::floor(x), yf = std::floor(y), zf = std::floor(z);
    const R dx = x - xf, dy = y - yf, dz = z - zf;
    const int x0 = (int)xf, y0 = (int)yf, z0 = (int)zf;
    const int i = clampi(x0 + 1 - px(b), 0, A - 2);
    const int j = clampi(y0 + 1 - py(b), 0, A - 2);
    const int k = clampi(z0 + 1 - pz(b), 0, A - 2);
    const size_t* a = &cidx[b * A3 + (size_t)A2 * i + (size_t)A * j + k];
    s.i[0] = a[0];
    s.i[1] = a[1];
    s.i[2] = a[A];
    s.i[3] = a[A + 1];
    s.i[4] = a[A2];
    s.i[5] = a[A2 + 1];
    s.i[6] = a[A2 + A];
    s.i[7] = a[A2 + A + 1];
    const R Dx = (R)1 - dx, Dy = (R)1 - dy, Dz = (R)1 - dz;
    R w[8];
    w[0] = flu[s.i[0]] * Dx * Dy * Dz;
    w[1] = flu[s.i[1]] * Dx * Dy * dz;
    w[2] = flu[s.i[2]] * Dx * dy * Dz;
    w[3] = flu[s.i[3]] * Dx * dy * dz;
    w[4] = flu[s.i[4]] * dx * Dy * Dz;
    w[5] = flu[s.i[5]] * dx * Dy * dz;
    w[6] = flu[s.i[6]] * dx * dy * Dz;
    w[7] = flu[s.i[7]] * dx * dy * dz;
    const R wAcc = w[0] + w[1] + w[2] + w[3] + w[4] + w[5] + w[6] + w[7];
    const R wInv = (R)1 / wAcc;
    for (int c = 0; c < 8; c++) s.w[c] = w[c] * wInv;
    s.zero = wAcc < (R)1e-6f;
    s.x0 = x0;
    s.y0 = y0;
    s.z0 = z0;
    s.scale = scale;
  }
  V3<R> sampleVelocity(const Sample& s) const {
    if (s.zero) return V3<R>{0, 0, 0};
    V3<R> acc{0, 0, 0};
    for (int c = 0; c < 8; c++) {
      const V3<R> v = velocityBC(P, vel[s.i[c]], s.x0 + ((c >> 2) & 1), s.y0 + ((c >> 1) & 1), s.z0 + (c & 1), s.scale);
      acc = c == 0 ? mul(v, s.w[0]) : add(acc, mul(v, s.w[c]));
    }
    return acc;
  }
  template <class F>
  R sampleScalar(const Sample& s, const std::vector<R>& q, F bc) const {
    if (s.zero) return (R)0;
    R acc = 0;
    for (int c = 0; c < 8; c++) {
      const R v = bc(q[s.i[c]], s.x0 + ((c >> 2) & 1), s.y0 + ((c >> 1) & 1), s.z0 + (c & 1), s.scale);
      acc = c == 0 ? v * s.w[0] : acc + v * s.w[c];
    }
    return acc;
  }
  // Backtraced position of a cell, dcgrid_fluid.cu:115-122.
  void backtrace(size_t c, R& x, R& y, R& z) const {
    const size_t b = c / BV;
    const R scale = (R)(1 << lev[b]);
    const R alpha = (R)P.dt * (R)P.rdx;
    int cx, cy, cz;
    cellPos(c, cx, cy, cz);
    x = ((R)cx + (R).5) * scale - vel[c].x * alpha;
    y = ((R)cy + (R).5) * scale - vel[c].y * alpha;
    z = ((R)cz + (R).5) * scale - vel[c].z * alpha;
  }
  // advectVelocity(): k_dcgrid_advect_velocity + copy + accumulateVelocity
  // (fluid_simulation_dcgrid.cu:263-268, dcgrid_fluid.cu:112-127)
  int advectVelocity() override {
    const bool buoy = thermo();
    const R dt = (R)P.dt;
#pragma omp parallel for schedule(static)
    for (long long bl = 0; bl < (long long)B; bl++) {
      const size_t b = (size_t)bl;
      if (lev[b] == 0xFF) continue;
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        if (!isLeaf(c)) {
          t_vel[c] = V3<R>{0, 0, 0};
          continue;
        }
        R x, y, z;
        backtrace(c, x, y, z);
        Sample s;
        sample(x, y, z, s);
        V3<R> nv = sampleVelocity(s);
        if (buoy) {
          int cx, cy, cz;
          cellPos(c, cx, cy, cz);
          const R ta = (R)ambientT(P, cellHeight(cy, 1 << lev[b]));
          const R bf = dt * ((R)P.buoy_beta * (tem[c] - ta) - (R)P.buoy_alpha * den[c]);
          nv.y = nv.y + bf;
        }
        t_vel[c] = st(nv);
      }
    }
    copyAllocated(t_vel, vel);
    accumulate(vel, V3<R>{0, 0, 0});
    return DCG_OK;
  }
  // Copy of the cells of allocated blocks (the reference copies the whole array; cells of free
  // blocks are never read).
  template <class T>
  void copyAllocated(const std::vector<T>& from, std::vector<T>& to) {
#pragma omp parallel for schedule(static)
    for (long long bl = 0; bl < (long long)B; bl++) {
      if (lev[(size_t)bl] == 0xFF) continue;
      std::copy(from.begin() + bl * BV, from.begin() + (bl + 1) * BV, to.begin() + bl * BV);
    }
  }
  // advectDensity(): k_dcgrid_advect_density + copy + accumulateDensity
  // (fluid_simulation_dcgrid.cu:313-318, dcgrid_fluid.cu:129-144); temperature and the sphere
  // source are extensions.
  int advectDensity() override {
    const bool src = thermo(), moist = P.moist != 0, sphere = P.source_mode == DCG_SOURCE_SPHERE_BUOYANT;
#pragma omp parallel for schedule(static)
    for (long long bl = 0; bl < (long long)B; bl++) {
      const size_t b = (size_t)bl;
      if (lev[b] == 0xFF) continue;
      const int level = lev[b];
      const int scale = 1 << level;
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        if (!isLeaf(c)) {
          t_den[c] = 0;
          if (src) t_tem[c] = 0;
          if (moist) t_vap[c] = 0;
          continue;
        }
        R x, y, z;
        backtrace(c, x, y, z);
        Sample s;
        sample(x, y, z, s);
        R q = sampleScalar(s, den, [&](R a, int xx, int yy, int zz, int sc) { return densityBC(P, a, xx, yy, zz, sc); });
        if (src) {
          R t = sampleScalar(s, tem, [&](R a, int xx, int yy, int zz, int sc) { return temperatureBC(P, a, xx, yy, zz, sc); });
          int cx, cy, cz;
          cellPos(c, cx, cy, cz);
          const float fcx = ((float)cx + .5f) * (float)scale, fcy = ((float)cy + .5f) * (float)scale,
                      fcz = ((float)cz + .5f) * (float)scale;
          if (sphere && inSource(P, fcx, fcy, fcz)) {
            q = (R)P.src_density;
            t = (R)sourceTemperature(P, cx, cy, cz, level);
          }
          if (moist) {
            R v = sampleScalar(s, vap, [&](R a, int xx, int yy, int zz, int sc) { return vaporBC(P, a, xx, yy, zz, sc); });
            t = t - ((R)P.dt * (R)P.adiabatic_rate) * vel[c].y;  // expansion cooling of rising air
            saturationAdjust(P, v, q, t);
            t_vap[c] = st(v);
          }
          t_tem[c] = st(t);
        }
        t_den[c] = st(q);
      }
    }
    copyAllocated(t_den, den);
    accumulate(den, (R)0);
    if (src) {
      copyAllocated(t_tem, tem);
      accumulate(tem, (R)0);
    }
    if (moist) {
      copyAllocated(t_vap, vap);
      accumulate(vap, (R)0);
    }
    return DCG_OK;
  }
  // k_dcgrid_calc_divergence, dcgrid_fluid.cu:174-230 (LDS staging restated directly)
  void calcDivergence() {
#pragma omp parallel for schedule(static)
    for (long long bl = 0; bl < (long long)B; bl++) {
      const size_t b = (size_t)bl;
      if (lev[b] == 0xFF) continue;
      const int scale = 1 << lev[b];
      const size_t* ci = &cidx[b * A3];
      // staged velocity: interior raw, border with BCs
      std::vector<V3<R>> stg(A3, V3<R>{0, 0, 0});
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        pre[c] = 0;
        t_pre[c] = 0;
        dive[c] = 0;
        int x, y, z;
        cellPos(c, x, y, z);
        const int a = apronOf(c);
        const int ax = a / A2, ay = (a / A) % A, az = a % A;
        stg[a] = vel[c];
        if (ax == 1) stg[a - A2] = velocityBC(P, vel[ci[a - A2]], x - 1, y, z, scale);
        if (ay == 1) stg[a - A] = velocityBC(P, vel[ci[a - A]], x, y - 1, z, scale);
        if (az == 1) stg[a - 1] = velocityBC(P, vel[ci[a - 1]], x, y, z - 1, scale);
        if (ax == BW) stg[a + A2] = velocityBC(P, vel[ci[a + A2]], x + 1, y, z, scale);
        if (ay == BW) stg[a + A] = velocityBC(P, vel[ci[a + A]], x, y + 1, z, scale);
        if (az == BW) stg[a + 1] = velocityBC(P, vel[ci[a + 1]], x, y, z + 1, scale);
      }
      const R alpha = (R).5f * (R)P.rdx / (R)scale;
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        if (!isLeaf(c)) continue;
        const int a = apronOf(c);
        const R wl = flu[ci[a - A2]], wr = flu[ci[a + A2]], wd = flu[ci[a - A]], wu = flu[ci[a + A]];
        const R wb = flu[ci[a - 1]], wf = flu[ci[a + 1]];
        const V3<R>&vl = stg[a - A2], &vr = stg[a + A2], &vd = stg[a - A], &vu = stg[a + A], &vb = stg[a - 1], &vf = stg[a + 1];
        dive[c] = st(alpha * (wr * vr.x - wl * vl.x + wu * vu.y - wd * vd.y + wf * vf.z - wb * vb.z));
      }
    }
  }
  // k_dcgrid_jacobi / k_dcgrid_jacobi_inv, dcgrid_multigrid_solver.cu:5-41
  void jacobi(int level, const std::vector<R>& in, std::vector<R>& out) {
    const int bs_ = 1 << level;
    const R alpha = (R)(bs_ * bs_) * (R)P.dx * (R)P.dx;
#pragma omp parallel for schedule(static)
    for (long long bl = (long long)offs[level]; bl < (long long)(offs[level] + maxL[level]); bl++) {
      const size_t b = (size_t)bl;
      if (lev[b] != level) continue;
      const size_t* ci = &cidx[b * A3];
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        const int a = apronOf(c);
        out[c] = st((in[ci[a - A2]] + in[ci[a + A2]] + in[ci[a - A]] + in[ci[a + A]] + in[ci[a - 1]] + in[ci[a + 1]] -
                     alpha * dive[c]) /
                    (R)6);
      }
    }
  }
  // k_dcgrid_prolongate, dcgrid_multigrid_solver.cu:43-76
  void prolongate(int level) {
#pragma omp parallel for schedule(static)
    for (long long bl = (long long)offs[level]; bl < (long long)(offs[level] + maxL[level]); bl++) {
      const size_t b = (size_t)bl;
      if (lev[b] != level || par[b] == NF) continue;
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        size_t q[8];
        parentStencil(c, q);
        pre[c] = st(((R)27 * pre[q[0]] + (R)9 * (pre[q[1]] + pre[q[2]] + pre[q[3]]) +
                     (R)3 * (pre[q[4]] + pre[q[5]] + pre[q[6]]) + pre[q[7]]) /
                    (R)64);
      }
    }
  }
  // k_dcgrid_apply_pressure, dcgrid_fluid.cu:232-259
  void applyPressure() {
#pragma omp parallel for schedule(static)
    for (long long bl = 0; bl < (long long)B; bl++) {
      const size_t b = (size_t)bl;
      if (lev[b] == 0xFF) continue;
      const int scale = 1 << lev[b];
      const R alpha = (R).5f * (R)P.rdx / (R)scale;
      const size_t* ci = &cidx[b * A3];
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        if (!isLeaf(c)) continue;
        const int a = apronOf(c);
        const R p = pre[c];
        V3<R>& v = vel[c];
        v.x = v.x - alpha * (flu[ci[a + A2]] * (pre[ci[a + A2]] - p) + flu[ci[a - A2]] * (p - pre[ci[a - A2]]));
        v.y = v.y - alpha * (flu[ci[a + A]] * (pre[ci[a + A]] - p) + flu[ci[a - A]] * (p - pre[ci[a - A]]));
        v.z = v.z - alpha * (flu[ci[a + 1]] * (pre[ci[a + 1]] - p) + flu[ci[a - 1]] * (p - pre[ci[a - 1]]));
        v = st(v);
      }
    }
  }
  // project(), fluid_simulation_dcgrid.cu:270-294
  int project() override {
    calcDivergence();
    accumulate(dive, (R)0);
    for (int i = 0; i < P.project_top_pairs; i++) {
      jacobi(L - 1, pre, t_pre);
      jacobi(L - 1, t_pre, pre);
    }
    for (int level = L - 2; level >= 0; level--) {
      prolongate(level);
      for (int i = 0; i < P.project_level_pairs; i++) {
        jacobi(level, pre, t_pre);
        jacobi(level, t_pre, pre);
      }
    }
    applyPressure();
    accumulate(vel, V3<R>{0, 0, 0});
    return DCG_OK;
  }
  // projectLocal(), fluid_simulation_dcgrid.cu:296-311
  int projectLocal() override {
    calcDivergence();
    accumulate(dive, (R)0);
    for (int level = L - 1; level >= 0; level--)
      for (int i = 0; i < P.project_local_pairs; i++) {
        jacobi(level, pre, t_pre);
        jacobi(level, t_pre, pre);
      }
    applyPressure();
    accumulate(vel, V3<R>{0, 0, 0});
    return DCG_OK;
  }
  // ---------------------------------------------------------------- projection solve (extension)
  // The fixed point of the Jacobi sweeps of project(), level by level from the coarsest:
  //   r_c = (sum of the 6 apron pressures) - 6 p_c - alpha div_c = 0   on every cell of the level,
  // with apron entries on coarser levels held at their solved values. A level whose loads equal
  // its full count covers the domain and sees no coarser level: its system is pure Neumann
  // (singular), so the right-hand side is projected to zero mean and the solution normalised to
  // zero mean. Solved here by lexicographic SOR (omega 1.8) until |r| <= tol |r(0)| (tol 1e-12 in
  // the FP64 instance, the reference solution); slow but independent of the GPU's multigrid.
  dcg_solve_info rep{};
  bool levelSingular(int l) const { return loads[l] == fullL[l]; }
  template <class F>
  void forLevelCells(int level, F f) {
    for (size_t b = offs[level]; b < offs[level] + maxL[level]; b++)
      if (lev[b] == level)
        for (int l = 0; l < BV; l++) f(b, b * BV + l);
  }
  // Residual norm of the level with the (projected) right-hand side alpha*div - shift.
  double levelResidualNorm(int level, R alpha, R shift, bool zeroLevel) {
    double s = 0;
    forLevelCells(level, [&](size_t b, size_t c) {
      const size_t* ci = &cidx[b * A3];
      const int a = apronOf(c);
      const size_t nb[6] = {ci[a - A2], ci[a + A2], ci[a - A], ci[a + A], ci[a - 1], ci[a + 1]};
      R sum = 0;
      for (int k = 0; k < 6; k++) {
        const bool unknown = lev[nb[k] / BV] == level;
        sum += (zeroLevel && unknown) ? (R)0 : pre[nb[k]];
      }
      const R r = sum - (zeroLevel ? (R)0 : (R)6 * pre[c]) - (alpha * dive[c] - shift);
      s += (double)r * (double)r;
    });
    return std::sqrt(s);
  }
  void solveLevel(int level) {
    const int scale = 1 << level;
    const R alpha = (R)(scale * scale) * (R)P.dx * (R)P.dx;
    const bool singular = levelSingular(level);
    R shift = 0;
    size_t n = 0;
    if (singular) {
      double m = 0;
      forLevelCells(level, [&](size_t, size_t c) {
        m += (double)(alpha * dive[c]);
        n++;
      });
      shift = (R)(m / (double)n);
    }
    const double tol = sizeof(R) == 8 ? 1e-12 : std::max((double)P.solve_tol, 1e-6);
    const double r0 = levelResidualNorm(level, alpha, shift, true);
    double rr = levelResidualNorm(level, alpha, shift, false);
    const R omega = (R)1.8;
    int sweeps = 0;
    const int maxSweeps = 200000;
    while (rr > tol * r0 && sweeps < maxSweeps) {
      for (int it = 0; it < 10; it++, sweeps++)
        forLevelCells(level, [&](size_t b, size_t c) {
          const size_t* ci = &cidx[b * A3];
          const int a = apronOf(c);
          const size_t nb[6] = {ci[a - A2], ci[a + A2], ci[a - A], ci[a + A], ci[a - 1], ci[a + 1]};
          R sum = 0;
          int self = 0;
          for (int k = 0; k < 6; k++) {
            if (nb[k] == c)
              self++;
            else
              sum += pre[nb[k]];
          }
          const R gs = (sum - (alpha * dive[c] - shift)) / (R)(6 - self);
          pre[c] = pre[c] + omega * (gs - pre[c]);
        });
      rr = levelResidualNorm(level, alpha, shift, false);
    }
    if (singular) {
      double m = 0;
      forLevelCells(level, [&](size_t, size_t c) { m += (double)pre[c]; });
      const R mean = (R)(m / (double)n);
      forLevelCells(level, [&](size_t, size_t c) { pre[c] = pre[c] - mean; });
      rr = levelResidualNorm(level, alpha, shift, false);
    }
    rep.cycles[level] = sweeps;
    rep.rel_residual[level] = r0 > 0 ? rr / r0 : 0.0;
    rep.initial_residual[level] = r0;
    rep.singular[level] = singular ? 1 : 0;
    if (!(rr <= tol * r0)) rep.converged = 0;
  }
  int projectSolve() override {
    calcDivergence();
    accumulate(dive, (R)0);
    rep = dcg_solve_info{};
    rep.levels = L;
    rep.converged = 1;
    for (int level = L - 1; level >= 0; level--) {
      if (level < L - 1) prolongate(level);  // initial guess from the solved coarser level
      solveLevel(level);
    }
    // the solved pressure is stored once, at the end
    for (size_t c = 0; c < C; c++) pre[c] = st(pre[c]);
    applyPressure();
    accumulate(vel, V3<R>{0, 0, 0});
    return DCG_OK;
  }
  int solveReport(dcg_solve_info* out) override {
    *out = rep;
    return DCG_OK;
  }

  // ---------------------------------------------------------------- adjoint (extension)
  // Reverse mode of the step on a frozen block topology: advectVelocity, the Jacobi projection
  // (project or projectLocal), advectDensity, without adaptTopology, so the block tables, apron
  // indices and fluidity are fixed over the rollout. Each forward loop `out = f(in)` is mirrored,
  // in reverse order, by `lam_in += (df/din)^T lam_out`, then `lam_out = 0` where the output
  // overwrote its previous value. Plain serial scatters in a fixed order. Storage rounding is
  // taken as the identity. Cells of free blocks take no part in a step (nothing reads them and
  // their values after a step are undefined), so their gradient entries are 0.
  struct Snap {
    std::vector<R> den, tem;
    std::vector<V3<R>> vel;
  };
  std::vector<Snap> tape;

  int tapeForward(int n) override {
    if (n < 0) return fail(DCG_E_ARG, "negative step count");
    if (P.moist) return fail(DCG_E_UNSUPPORTED, "the adjoint does not differentiate the moist atmosphere");
    if (P.project_mode == DCG_PROJECT_SOLVE) return fail(DCG_E_UNSUPPORTED, "the adjoint follows the Jacobi schedules");
    tape.clear();
    tape.push_back(Snap{den, tem, vel});
    for (int k = 0; k < n; k++) {
      if (int e = advectVelocity()) return e;
      if (int e = projectByMode()) return e;
      if (int e = advectDensity()) return e;
      tape.push_back(Snap{den, tem, vel});
    }
    return DCG_OK;
  }

  // accumulate() reversed: levels from coarse to fine, a parent cell's adjoint shared equally by
  // its 8 children, then cleared (the restriction overwrote the parent cell).
  template <class T>
  void restrictAdj(std::vector<T>& lam) {
    for (int level = L - 2; level >= 0; level--)
      for (size_t b = offs[level]; b < offs[level] + maxL[level]; b++) {
        if (lev[b] != level || par[b] == NF) continue;
        const size_t ps = par[b];
        for (size_t g = 0; g < (size_t)BV / 8; g++) {
          const size_t pc = (ps / 8) * BV + ((ps % 8) << SBB) + g;
          const T G = mulT(lam[pc], (R).125f);
          for (size_t i = 0; i < 8; i++) lam[b * BV + 8 * g + i] = addT(lam[b * BV + 8 * g + i], G);
          lam[pc] = T{};
        }
      }
  }
  // Sample of the foot of leaf cell c under velocity v (the backtrace expressions), with the
  // unnormalised corner weights f_c t_c, d t_c / d frac_a, and k = d frac / d v.
  struct FootD {
    Sample s;
    R f[8], t[8], dtd[8][3], W, k;
  };
  void footD(size_t c, const V3<R>& v, FootD& F) const {
    const R scale = (R)(1 << lev[c / BV]);
    const R alpha = (R)P.dt * (R)P.rdx;
    int cx, cy, cz;
    cellPos(c, cx, cy, cz);
    const R px_ = ((R)cx + (R).5) * scale - v.x * alpha;
    const R py_ = ((R)cy + (R).5) * scale - v.y * alpha;
    const R pz_ = ((R)cz + (R).5) * scale - v.z * alpha;
    sample(px_, py_, pz_, F.s);
    const R si = (R)1 / (R)F.s.scale;
    const R xs = px_ * si - (R).5, ys = py_ * si - (R).5, zs = pz_ * si - (R).5;
    const R d[3] = {xs - std::floor(xs), ys - std::floor(ys), zs - std::floor(zs)};
    F.W = 0;
    for (int q = 0; q < 8; q++) {
      const int bit[3] = {(q >> 2) & 1, (q >> 1) & 1, q & 1};
      R tau[3];
      for (int a = 0; a < 3; a++) tau[a] = bit[a] ? d[a] : (R)1 - d[a];
      F.t[q] = tau[0] * tau[1] * tau[2];
      for (int a = 0; a < 3; a++) {
        R g = bit[a] ? (R)1 : (R)-1;
        for (int b = 0; b < 3; b++)
          if (b != a) g *= tau[b];
        F.dtd[q][a] = g;
      }
      F.f[q] = flu[F.s.i[q]];
      F.W += F.f[q] * F.t[q];
    }
    F.k = -alpha * si;
  }
  // d(sample)/d(v) of the normalised interpolant of the BC-applied corner values vc.
  void footGrad(const FootD& F, const R vc[8], R out[3]) const {
    R val = 0;
    for (int q = 0; q < 8; q++) val += F.f[q] * F.t[q] * vc[q];
    val /= F.W;
    for (int a = 0; a < 3; a++) {
      R num = 0, den_ = 0;
      for (int q = 0; q < 8; q++) {
        num += F.f[q] * vc[q] * F.dtd[q][a];
        den_ += F.f[q] * F.dtd[q][a];
      }
      out[a] = F.k * (num - val * den_) / F.W;
    }
  }
  // Corner q of a sample is inside the domain: its BC passes the stored value through
  // (outside it, and on the inlet, which lies outside, the BCs give constants).
  bool cornerPasses(const Sample& s, int q) const {
    return !outside(P, s.x0 + ((q >> 2) & 1), s.y0 + ((q >> 1) & 1), s.z0 + (q & 1), s.scale);
  }

  // advectDensity reversed: lq, lt (after) -> lq0, lt0 (before); foot terms into lu, the adjoint
  // of the velocity the step advected with.
  void advectDensityAdj(const Snap& S0, const Snap& S1, std::vector<R>& lq, std::vector<R>& lt,
                        std::vector<V3<R>>& lu, std::vector<R>& lq0, std::vector<R>& lt0) {
    const bool src = thermo(), sphere = P.source_mode == DCG_SOURCE_SPHERE_BUOYANT;
    restrictAdj(lq);
    if (src) restrictAdj(lt);
    lq0.assign(C, (R)0);
    if (src)
      lt0.assign(C, (R)0);
    else
      lt0 = lt;  // temperature is only advected with a source
    for (size_t b = 0; b < B; b++) {
      if (lev[b] == 0xFF) continue;
      const int level = lev[b], scale = 1 << level;
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        if (!isLeaf(c)) continue;
        if (src && sphere) {
          int cx, cy, cz;
          cellPos(c, cx, cy, cz);
          if (inSource(P, ((float)cx + .5f) * (float)scale, ((float)cy + .5f) * (float)scale,
                       ((float)cz + .5f) * (float)scale))
            continue;  // overwritten by the source's constants
        }
        FootD F;
        footD(c, S1.vel[c], F);
        if (F.s.zero) continue;
        R qc[8], tc[8];
        for (int q = 0; q < 8; q++) {
          const int x = F.s.x0 + ((q >> 2) & 1), y = F.s.y0 + ((q >> 1) & 1), z = F.s.z0 + (q & 1);
          const size_t i = F.s.i[q];
          qc[q] = densityBC(P, S0.den[i], x, y, z, F.s.scale);
          tc[q] = temperatureBC(P, S0.tem[i], x, y, z, F.s.scale);
          if (cornerPasses(F.s, q)) {
            lq0[i] += F.s.w[q] * lq[c];
            if (src) lt0[i] += F.s.w[q] * lt[c];
          }
        }
        R g[3];
        footGrad(F, qc, g);
        lu[c] = add(lu[c], V3<R>{g[0] * lq[c], g[1] * lq[c], g[2] * lq[c]});
        if (src) {
          footGrad(F, tc, g);
          lu[c] = add(lu[c], V3<R>{g[0] * lt[c], g[1] * lt[c], g[2] * lt[c]});
        }
      }
    }
  }

  // project() / projectLocal() reversed: lu holds the adjoint of the velocity after the
  // projection on entry and before it on exit.
  void projectAdj(std::vector<V3<R>>& lu) {
    std::vector<R> lp(C, (R)0), ltp(C, (R)0), ld(C, (R)0);
    restrictAdj(lu);
    // applyPressure: v -= alpha (f_r (p_r - p) + f_l (p - p_l)) per axis; v passes through
    for (size_t b = 0; b < B; b++) {
      if (lev[b] == 0xFF) continue;
      const R alpha = (R).5f * (R)P.rdx / (R)(1 << lev[b]);
      const size_t* ci = &cidx[b * A3];
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        if (!isLeaf(c)) continue;
        const int a = apronOf(c);
        const size_t nb[3][2] = {{ci[a - A2], ci[a + A2]}, {ci[a - A], ci[a + A]}, {ci[a - 1], ci[a + 1]}};
        const R lam[3] = {lu[c].x, lu[c].y, lu[c].z};
        for (int ax = 0; ax < 3; ax++) {
          const size_t il = nb[ax][0], ir = nb[ax][1];
          const R fl = flu[il], fr = flu[ir];
          lp[ir] += -alpha * fr * lam[ax];
          lp[c] += alpha * fr * lam[ax];
          lp[c] += -alpha * fl * lam[ax];
          lp[il] += alpha * fl * lam[ax];
        }
      }
    }
    // jacobi(level, in, out) reversed
    auto jacobiAdj = [&](int level, std::vector<R>& lin, std::vector<R>& lout) {
      const int bs_ = 1 << level;
      const R alpha = (R)(bs_ * bs_) * (R)P.dx * (R)P.dx;
      for (size_t b = offs[level]; b < offs[level] + maxL[level]; b++) {
        if (lev[b] != level) continue;
        const size_t* ci = &cidx[b * A3];
        for (int l = 0; l < BV; l++) {
          const size_t c = b * BV + l;
          const int a = apronOf(c);
          const R G = lout[c] / (R)6;
          lin[ci[a - A2]] += G;
          lin[ci[a + A2]] += G;
          lin[ci[a - A]] += G;
          lin[ci[a + A]] += G;
          lin[ci[a - 1]] += G;
          lin[ci[a + 1]] += G;
          ld[c] += -alpha * G;
          lout[c] = 0;
        }
      }
    };
    auto pairsAdj = [&](int level, int pairs) {
      for (int i = 0; i < pairs; i++) {
        jacobiAdj(level, ltp, lp);  // jacobi(level, t_pre, pre)
        jacobiAdj(level, lp, ltp);  // jacobi(level, pre, t_pre)
      }
    };
    if (P.project_mode == DCG_PROJECT_LOCAL) {
      for (int level = 0; level < L; level++) pairsAdj(level, P.project_local_pairs);
    } else {
      for (int level = 0; level <= L - 2; level++) {
        pairsAdj(level, P.project_level_pairs);
        // prolongate(level): pre = (27, 9, 3, 1)/64 weights of the parent stencil
        for (size_t b = offs[level]; b < offs[level] + maxL[level]; b++) {
          if (lev[b] != level || par[b] == NF) continue;
          for (int l = 0; l < BV; l++) {
            const size_t c = b * BV + l;
            size_t q[8];
            parentStencil(c, q);
            const R G = lp[c] / (R)64;
            lp[q[0]] += (R)27 * G;
            lp[q[1]] += (R)9 * G;
            lp[q[2]] += (R)9 * G;
            lp[q[3]] += (R)9 * G;
            lp[q[4]] += (R)3 * G;
            lp[q[5]] += (R)3 * G;
            lp[q[6]] += (R)3 * G;
            lp[q[7]] += G;
            lp[c] = 0;
          }
        }
      }
      pairsAdj(L - 1, P.project_top_pairs);
    }
    restrictAdj(ld);
    // calcDivergence: div = alpha (w_r v_r.x - w_l v_l.x + ...) over leaf cells, neighbours with BCs
    for (size_t b = 0; b < B; b++) {
      if (lev[b] == 0xFF) continue;
      const int scale = 1 << lev[b];
      const R alpha = (R).5f * (R)P.rdx / (R)scale;
      const size_t* ci = &cidx[b * A3];
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        if (!isLeaf(c)) continue;
        int x, y, z;
        cellPos(c, x, y, z);
        const int a = apronOf(c);
        const R G = alpha * ld[c];
        const size_t il = ci[a - A2], ir = ci[a + A2], id = ci[a - A], iu = ci[a + A], ib = ci[a - 1], iF = ci[a + 1];
        if (!outside(P, x + 1, y, z, scale)) lu[ir].x += flu[ir] * G;
        if (!outside(P, x - 1, y, z, scale)) lu[il].x += -flu[il] * G;
        if (!outside(P, x, y + 1, z, scale)) lu[iu].y += flu[iu] * G;
        if (!outside(P, x, y - 1, z, scale)) lu[id].y += -flu[id] * G;
        if (!outside(P, x, y, z + 1, scale)) lu[iF].z += flu[iF] * G;
        if (!outside(P, x, y, z - 1, scale)) lu[ib].z += -flu[ib] * G;
      }
    }
  }

  // advectVelocity reversed: lu (after) -> lu0 (before); buoyancy into lq0, lt0.
  void advectVelocityAdj(const Snap& S0, std::vector<V3<R>>& lu, std::vector<V3<R>>& lu0, std::vector<R>& lq0,
                         std::vector<R>& lt0) {
    const bool buoy = thermo();
    const R dt = (R)P.dt;
    restrictAdj(lu);
    lu0.assign(C, V3<R>{0, 0, 0});
    for (size_t b = 0; b < B; b++) {
      if (lev[b] == 0xFF) continue;
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        if (!isLeaf(c)) continue;
        if (buoy) {
          lq0[c] += -dt * (R)P.buoy_alpha * lu[c].y;
          lt0[c] += dt * (R)P.buoy_beta * lu[c].y;
        }
        FootD F;
        footD(c, S0.vel[c], F);
        if (F.s.zero) continue;
        R vc[3][8];
        for (int q = 0; q < 8; q++) {
          const size_t i = F.s.i[q];
          const V3<R> v = velocityBC(P, S0.vel[i], F.s.x0 + ((q >> 2) & 1), F.s.y0 + ((q >> 1) & 1), F.s.z0 + (q & 1),
                                     F.s.scale);
          vc[0][q] = v.x;
          vc[1][q] = v.y;
          vc[2][q] = v.z;
          if (cornerPasses(F.s, q)) lu0[i] = add(lu0[i], mul(lu[c], F.s.w[q]));
        }
        const R lam[3] = {lu[c].x, lu[c].y, lu[c].z};
        for (int comp = 0; comp < 3; comp++) {
          R g[3];
          footGrad(F, vc[comp], g);
          lu0[c] = add(lu0[c], V3<R>{g[0] * lam[comp], g[1] * lam[comp], g[2] * lam[comp]});
        }
      }
    }
  }

  int tapeBackward(const double* gden, const double* gvel, const double* gtem, double* oden, double* ovel,
                   double* otem) override {
    if (tape.empty()) return fail(DCG_E_ARG, "no recorded rollout");
    std::vector<R> lq(C), lt(C, (R)0), lq0, lt0;
    std::vector<V3<R>> lu(C), lu0;
    for (size_t i = 0; i < C; i++) {
      lq[i] = (R)gden[i];
      lu[i] = V3<R>{(R)gvel[3 * i], (R)gvel[3 * i + 1], (R)gvel[3 * i + 2]};
      if (gtem) lt[i] = (R)gtem[i];
    }
    for (int k = (int)tape.size() - 2; k >= 0; k--) {
      advectDensityAdj(tape[k], tape[k + 1], lq, lt, lu, lq0, lt0);
      projectAdj(lu);
      advectVelocityAdj(tape[k], lu, lu0, lq0, lt0);
      lq.swap(lq0);
      lt.swap(lt0);
      lu.swap(lu0);
    }
    for (size_t i = 0; i < C; i++) {
      const bool on = lev[i / BV] != 0xFF;
      oden[i] = on ? (double)lq[i] : 0.0;
      ovel[3 * i] = on ? (double)lu[i].x : 0.0;
      ovel[3 * i + 1] = on ? (double)lu[i].y : 0.0;
      ovel[3 * i + 2] = on ? (double)lu[i].z : 0.0;
      if (otem) otem[i] = on ? (double)lt[i] : 0.0;
    }
    return DCG_OK;
  }

  // k_dcgrid_debug_stats + host sum, dcgrid_structure.cu:224-251, fluid_simulation_dcgrid.cu:517-528
  // Samplers of src/dcgrid/dcgrid_rendering.cu:6-64 over the reference-semantics tables.
  struct MarchSampler {
    const DCGrid* d;
    void sample(const dcg_params& P, int precise, F3 pos, Aux& a) const {
      a.ipos[0] = clampi((int)floorf(pos.x), 0, d->gx - 1);
      a.ipos[1] = clampi((int)floorf(pos.y), 0, d->gy - 1);
      a.ipos[2] = clampi((int)floorf(pos.z), 0, d->gz - 1);
      int x = a.ipos[0], y = a.ipos[1], z = a.ipos[2], level = 0;
      const size_t b = d->getBlockIndexDeep(x, y, z, level);
      if (!precise) {
        const size_t c = (size_t)d->BV * b + localIndex(a.ipos[0] % d->BW, a.ipos[1] % d->BW, a.ipos[2] % d->BW, d->BB);
        for (int k = 0; k < 8; k++) a.cell[k] = c;
        a.density = (float)d->den[c];
        return;
      }
      const float scaleInv = 1.f / (float)(1 << level);
      const float fx = pos.x * scaleInv - .5f, fy = pos.y * scaleInv - .5f, fz = pos.z * scaleInv - .5f;
      const float xf = floorf(fx), yf = floorf(fy), zf = floorf(fz);
      a.frac[0] = fx - xf;
      a.frac[1] = fy - yf;
      a.frac[2] = fz - zf;
      const int i = clampi((int)(xf + 1.f - (float)d->px(b)), 0, d->BW);
      const int j = clampi((int)(yf + 1.f - (float)d->py(b)), 0, d->BW);
      const int k = clampi((int)(zf + 1.f - (float)d->pz(b)), 0, d->BW);
      const size_t* ap = &d->cidx[b * d->A3 + (size_t)d->A2 * i + (size_t)d->A * j + k];
      a.cell[0] = ap[0];
      a.cell[1] = ap[1];
      a.cell[2] = ap[d->A];
      a.cell[3] = ap[d->A + 1];
      a.cell[4] = ap[d->A2];
      a.cell[5] = ap[d->A2 + 1];
      a.cell[6] = ap[d->A2 + d->A];
      a.cell[7] = ap[d->A2 + d->A + 1];
      a.density = interp(d->den, a);
    }
    F3 velocity(const Aux& a) const { return interp3(d->vel, a); }
    float fluidity(const Aux& a) const { return interp(d->flu, a); }
  };
  int render(const dcg_render_params& r, float* out) override {
    renderImage(MarchSampler{this}, P, r, out);
    return DCG_OK;
  }

  int debugStats(double* out) override {
    std::vector<R> stats(B, (R)0);
#pragma omp parallel for schedule(static)
    for (long long bl = 0; bl < (long long)B; bl++) {
      const size_t b = (size_t)bl;
      if (lev[b] == 0xFF) continue;
      const int scale = 1 << lev[b];
      const R alpha = (R)P.rdx * (R)P.rdx / (R)(scale * scale);
      R s = 0;
      for (int i = 1; i <= BW; i++)
        for (int j = 1; j <= BW; j++)
          for (int k = 1; k <= BW; k++) {
            const size_t c = ap(b, i, j, k);
            if (!isLeaf(c)) continue;
            const R pc = pre[c];
            const R pl = pre[ap(b, i - 1, j, k)], pr = pre[ap(b, i + 1, j, k)];
            const R pd = pre[ap(b, i, j - 1, k)], pu = pre[ap(b, i, j + 1, k)];
            const R pb = pre[ap(b, i, j, k - 1)], pf = pre[ap(b, i, j, k + 1)];
            const R r = dive[c] - (pl + pr + pd + pu + pb + pf - (R)6 * pc) * alpha;
            s = s + (R)scale * std::fabs(r);
          }
      stats[b] = s;
    }
    R sum = 0;
    for (size_t b = 0; b < B; b++) sum = sum + stats[b];
    *out = (double)sum;
    return DCG_OK;
  }
  // Seeded initial condition (extension): every cell of an allocated block gets the value of
  // its centre, then the restriction makes parents the mean of their children.
  int seedScene() override {
    const bool src = P.source_mode == DCG_SOURCE_SPHERE_BUOYANT;
#pragma omp parallel for schedule(static)
    for (long long bl = 0; bl < (long long)B; bl++) {
      const size_t b = (size_t)bl;
      if (lev[b] == 0xFF) continue;
      const int level = lev[b], scale = 1 << level;
      for (int l = 0; l < BV; l++) {
        const size_t c = b * BV + l;
        int x, y, z;
        cellPos(c, x, y, z);
        const float cx = ((float)x + .5f) * (float)scale, cy = ((float)y + .5f) * (float)scale,
                    cz = ((float)z + .5f) * (float)scale;
        float u[3];
        initVelocity(P, cx, cy, cz, u);
        vel[c] = st(V3<R>{(R)u[0], (R)u[1], (R)u[2]});
        if (src && inSource(P, cx, cy, cz)) {
          den[c] = st((R)P.src_density);
          tem[c] = st((R)sourceTemperature(P, x, y, z, level));
        } else {
          den[c] = 0;
          tem[c] = st((R)ambientT(P, cy));
        }
        if (P.moist) vap[c] = st((R)ambientVapor(P, cy));
      }
    }
    accumulate(vel, V3<R>{0, 0, 0});
    accumulate(den, (R)0);
    accumulate(tem, (R)0);
    if (P.moist) accumulate(vap, (R)0);
    return DCG_OK;
  }

  // ---------------------------------------------------------------- export
  int info(dcg_info* o) override {
    std::memset(o, 0, sizeof(*o));
    o->kind = DCG_KIND_DCGRID;
    o->size[0] = gx;
    o->size[1] = gy;
    o->size[2] = gz;
    o->block_width = BW;
    o->levels = L;
    o->sparse_levels = S;
    o->apron_width = A;
    o->max_num_blocks = B;
    o->num_cells = C;
    o->hash_table_size = H;
    uint64_t act = 0;
    for (int l = 0; l < L; l++) {
      o->max_blocks_level[l] = maxL[l];
      o->full_blocks_level[l] = fullL[l];
      o->level_offsets[l] = offs[l];
      o->block_loads[l] = loads[l];
      o->move_limit[l] = moveLimit[l];
      o->hash_size_level[l] = hsize[l];
      o->hash_offset_level[l] = hoff[l];
      act += loads[l] * (uint64_t)BV;
    }
    o->active_cells = act;
    return DCG_OK;
  }
  int fieldSize(int f, uint64_t* n) override {
    if (f < 0 || f >= DCG_NUM_FIELDS) return fail(DCG_E_ARG, "unknown field");
    if (f == DCG_FIELD_VAPOR && !P.moist) return fail(DCG_E_ARG, "no vapor field (moist atmosphere off)");
    *n = (f == DCG_FIELD_VELOCITY || f == DCG_FIELD_VORTICITY) ? 3 * C : C;
    return DCG_OK;
  }
  std::vector<R>* scalar(int f) {
    switch (f) {
      case DCG_FIELD_VAPOR: return &vap;
      case DCG_FIELD_DENSITY: return &den;
      case DCG_FIELD_TEMPERATURE: return &tem;
      case DCG_FIELD_PRESSURE: return &pre;
      case DCG_FIELD_T_PRESSURE: return &t_pre;
      case DCG_FIELD_DIVERGENCE: return &dive;
      case DCG_FIELD_FLUIDITY: return &flu;
      default: return nullptr;
    }
  }
  int exportField(int f, double* dst, uint64_t n) override {
    uint64_t need;
    if (int e = fieldSize(f, &need)) return e;
    if (n != need) return fail(DCG_E_ARG, "size mismatch");
    if (f == DCG_FIELD_VELOCITY || f == DCG_FIELD_VORTICITY) {
      const std::vector<V3<R>>& v = f == DCG_FIELD_VELOCITY ? vel : vort;
      for (size_t i = 0; i < C; i++) {
        dst[3 * i] = (double)v[i].x;
        dst[3 * i + 1] = (double)v[i].y;
        dst[3 * i + 2] = (double)v[i].z;
      }
      return DCG_OK;
    }
    std::vector<R>* v = scalar(f);
    for (size_t i = 0; i < n; i++) dst[i] = (double)(*v)[i];
    return DCG_OK;
  }
  int importField(int f, const double* src, uint64_t n) override {
    uint64_t need;
    if (int e = fieldSize(f, &need)) return e;
    if (n != need) return fail(DCG_E_ARG, "size mismatch");
    if (f == DCG_FIELD_VELOCITY || f == DCG_FIELD_VORTICITY) {
      std::vector<V3<R>>& v = f == DCG_FIELD_VELOCITY ? vel : vort;
      for (size_t i = 0; i < C; i++) v[i] = st(V3<R>{(R)src[3 * i], (R)src[3 * i + 1], (R)src[3 * i + 2]});
      return DCG_OK;
    }
    std::vector<R>* v = scalar(f);
    for (size_t i = 0; i < n; i++) (*v)[i] = st((R)src[i]);
    return DCG_OK;
  }
  int runStage(int stage, int arg) override {
    const bool perLevel = stage == DCG_STAGE_JACOBI || stage == DCG_STAGE_JACOBI_INV || stage == DCG_STAGE_PROLONGATE;
    if (perLevel && (arg < 0 || arg >= L || (stage == DCG_STAGE_PROLONGATE && arg >= L - 1)))
      return fail(DCG_E_ARG, "level out of range");
    switch (stage) {
      case DCG_STAGE_DIVERGENCE: calcDivergence(); return DCG_OK;
      case DCG_STAGE_RESTRICT:
        switch (arg) {
          case DCG_FIELD_DENSITY: accumulate(den, (R)0); return DCG_OK;
          case DCG_FIELD_TEMPERATURE: accumulate(tem, (R)0); return DCG_OK;
          case DCG_FIELD_DIVERGENCE: accumulate(dive, (R)0); return DCG_OK;
          case DCG_FIELD_PRESSURE: accumulate(pre, (R)0); return DCG_OK;
          case DCG_FIELD_VELOCITY: accumulate(vel, V3<R>{0, 0, 0}); return DCG_OK;
          default: return fail(DCG_E_ARG, "field cannot be restricted");
        }
      case DCG_STAGE_JACOBI: jacobi(arg, pre, t_pre); return DCG_OK;
      case DCG_STAGE_JACOBI_INV: jacobi(arg, t_pre, pre); return DCG_OK;
      case DCG_STAGE_PROLONGATE: prolongate(arg); return DCG_OK;
      case DCG_STAGE_APPLY_PRESSURE: applyPressure(); return DCG_OK;
      case DCG_STAGE_VORTICITY: calcVorticity(); return DCG_OK;
      default: return fail(DCG_E_ARG, "unknown stage");
    }
  }
  int64_t lookup(int x, int y, int z, int level) override {
    if (level < 0 || level > S) return -1;
    int l = level;
    const size_t b = getBlockIndexDeep(x, y, z, l);
    return b == NF ? -1 : (int64_t)b;
  }
  // Test hook: the subblock and block scores the next adaptTopology starts from (its first two
  // kernels, dcgrid_adaptation.cu:10-63, after the vorticity pass in feature mode).
  int scores(float* sub, float* blk) override {
    if (P.score_mode == DCG_SCORE_FEATURES) calcVorticity();
    calcSubblockScores();
    accumulateSubblockScores();
    std::copy(sbs.begin(), sbs.begin() + 8 * B + 1, sub);
    std::copy(bs.begin(), bs.begin() + B, blk);
    return DCG_OK;
  }
  // Test hook: replace the block tables (reference semantics, as exportTables writes them), the
  // level loads and the move limits, e.g. with a state written by another implementation.
  int importTables(const dcg_tables* t, const uint64_t* ld, const uint64_t* ml) override {
    if (!t->block_positions || !t->block_levels || !t->block_flags || !t->parent_indices ||
        !t->child_indices || !t->cell_indices || !t->hash_key || !t->hash_val || !ld || !ml)
      return fail(DCG_E_ARG, "import_tables: every table is required");
    auto sz = [](uint64_t v) { return v == UINT64_MAX ? NF : (size_t)v; };
    for (int l = 0; l < L; l++)
      if (ld[l] > maxL[l]) return fail(DCG_E_ARG, "import_tables: load above the level's capacity");
    std::copy(t->block_positions, t->block_positions + 3 * B, pos.begin());
    std::copy(t->block_levels, t->block_levels + B, lev.begin());
    std::copy(t->block_flags, t->block_flags + B, flg.begin());
    for (size_t i = 0; i < B; i++) par[i] = sz(t->parent_indices[i]);
    for (size_t i = 0; i < 8 * B; i++) chi[i] = sz(t->child_indices[i]);
    for (size_t i = 0; i < cidx.size(); i++) cidx[i] = sz(t->cell_indices[i]);
    std::copy(t->hash_key, t->hash_key + H, hkey.begin());
    for (size_t i = 0; i < H; i++) hval[i] = sz(t->hash_val[i]);
    for (int l = 0; l < L; l++) {
      loads[l] = (size_t)ld[l];
      moveLimit[l] = (size_t)ml[l];
    }
    return DCG_OK;
  }
  int exportTables(dcg_tables* t) override {
    auto u64 = [](size_t v) { return v == NF ? UINT64_MAX : (uint64_t)v; };
    if (t->block_positions) std::copy(pos.begin(), pos.end(), t->block_positions);
    if (t->block_levels) std::copy(lev.begin(), lev.end(), t->block_levels);
    if (t->block_flags) std::copy(flg.begin(), flg.end(), t->block_flags);
    if (t->parent_indices)
      for (size_t i = 0; i < B; i++) t->parent_indices[i] = u64(par[i]);
    if (t->child_indices)
      for (size_t i = 0; i < 8 * B; i++) t->child_indices[i] = u64(chi[i]);
    if (t->cell_indices)
      for (size_t i = 0; i < cidx.size(); i++) t->cell_indices[i] = u64(cidx[i]);
    if (t->hash_key) std::copy(hkey.begin(), hkey.end(), t->hash_key);
    if (t->hash_val)
      for (size_t i = 0; i < H; i++) t->hash_val[i] = u64(hval[i]);
    return DCG_OK;
  }
};

Sim* makeDCGrid(const int32_t size[3], uint64_t maxBlocks, const dcg_params& P, int precision,
                int* status, std::string* err) {
  Sim* s;
  int st;
  if (precision == 32) {
    auto* g = new DCGrid<float>();
    st = g->create(size, maxBlocks, P);
    s = g;
  } else {
    auto* g = new DCGrid<double>();
    st = g->create(size, maxBlocks, P);
    s = g;
  }
  *status = st;
  if (st != DCG_OK) {
    *err = s->err;
    delete s;
    return nullptr;
  }
  return s;
}

}  // namespace orc

namespace orc {
// Level/budget planner only (fluid_simulation_dcgrid.cu:9-140), without allocation.
int planDCGrid(const int32_t size[3], uint64_t maxBlocks, const dcg_params& P, dcg_info* out, std::string* err) {
  DCGrid<float> g;
  const int st = g.plan(size, maxBlocks, P);
  if (st != DCG_OK) {
    *err = g.err;
    return st;
  }
  return g.info(out);
}
}  // namespace orc

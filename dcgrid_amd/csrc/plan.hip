// plan.hip — host-only pieces of the DCGrid back-end that need no device: the level and budget
// planner (fluid_simulation_dcgrid.cu:9-140) and the validation of the block tables of a saved
// state before any of them reaches the device.
#include <algorithm>
#include <cstring>

#include "dcg_internal.hpp"

namespace dcg {

namespace {
int iDivUp(int a, int b) { return (a % b != 0) ? (a / b + 1) : (a / b); }
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint8_t kFree = 0xFF;
constexpr uint32_t kOut = 31;              // neighbour descriptor shift of an out-of-domain direction
constexpr uint32_t kBlockMask = 0x07FFFFFFu;
}  // namespace

// Level count (fluid_simulation_dcgrid.cu:12-22), full block counts per level (:29-32), the top
// level dense and the per-level capacities top-down (:41-48), the error exits (:35-53), level
// offsets (:55-58) and the first dense level (:66-69).
int planDCGrid(int bw, const int32_t size[3], uint64_t maxBlocks, const dcg_params& p, DCGridPlan& out,
               std::string& err) {
  std::memset(&out, 0, sizeof(out));
  const int gx = size[0], gy = size[1], gz = size[2];
  if (bw != 4 && bw != 8) {
    err = "block_width must be 4 or 8";
    return DCG_E_UNSUPPORTED;
  }
  if (gx <= 0 || gy <= 0 || gz <= 0 || gx != p.gx || gy != p.gy || gz != p.gz) {
    err = "size must be positive and equal params gx,gy,gz";
    return DCG_E_ARG;
  }
  const int minDim = std::min(gx, std::min(gy, gz));
  int cellSize = 2, L = 1;
  while (gx % cellSize == 0 && gy % cellSize == 0 && gz % cellSize == 0 && cellSize * bw <= minDim) {
    L++;
    cellSize *= 2;
  }
  if (p.max_levels > 0 && L > p.max_levels) L = p.max_levels;
  if (L > DCG_MAX_LEVELS) {
    err = "too many levels";
    return DCG_E_UNSUPPORTED;
  }
  const int top = bw << (L - 1);
  if (gx % top || gy % top || gz % top) {
    err = "domain must be a multiple of the coarsest block extent";
    return DCG_E_UNSUPPORTED;
  }
  uint64_t full64[DCG_MAX_LEVELS];
  for (int l = 0, cs = 1; l < L; l++, cs *= 2)
    full64[l] = (uint64_t)iDivUp(gx, cs * bw) * iDivUp(gy, cs * bw) * iDivUp(gz, cs * bw);
  if (maxBlocks < full64[L - 1]) {
    err = "too few blocks to fill lowest resolution";
    return DCG_E_BUDGET;
  }
  const uint64_t bv = (uint64_t)bw * bw * bw;
  if (maxBlocks * bv >= 0xFFFFFFFFull || maxBlocks > kBlockMask) {
    err = "budget exceeds 32-bit cell indexing";
    return DCG_E_UNSUPPORTED;
  }
  uint64_t mx[DCG_MAX_LEVELS];
  mx[L - 1] = full64[L - 1];
  uint64_t left = maxBlocks - mx[L - 1];
  for (int l = L - 2; l >= 0; l--) {
    mx[l] = std::min(left / (uint64_t)(l + 1), full64[l]);
    left -= mx[l];
  }
  if (mx[0] == 0) {
    err = "too few blocks to reach highest resolution";
    return DCG_E_BUDGET;
  }
  for (int l = 0; l < L; l++)
    if (mx[l] == 0) {
      err = "a level between the finest and the coarsest gets no blocks";
      return DCG_E_BUDGET;
    }
  out.bw = bw;
  out.L = L;
  out.B = (uint32_t)maxBlocks;
  for (int l = 0; l < L; l++) {
    out.maxL[l] = (uint32_t)mx[l];
    out.fullL[l] = (uint32_t)full64[l];
    out.rx[l] = iDivUp(gx, bw << l);
    out.ry[l] = iDivUp(gy, bw << l);
    out.rz[l] = iDivUp(gz, bw << l);
  }
  out.offs[0] = 0;
  for (int l = 1; l < L; l++) out.offs[l] = out.offs[l - 1] + out.maxL[l - 1];
  out.S = 0;
  for (int l = 0; l < L; l++)
    if (out.maxL[l] < out.fullL[l]) out.S = l + 1;
  out.H = 4 * out.B;
  for (int l = 0; l < L; l++) out.hsize[l] = 4 * out.maxL[l];
  out.hoff[0] = 0;
  for (int l = 1; l < L; l++) out.hoff[l] = out.hoff[l - 1] + out.hsize[l - 1];
  return DCG_OK;
}

// Every index a kernel follows from the tables must stay inside its array: the slot walks take
// the allocated blocks of level l as the prefix of loads[l] blocks of its slice, the leaf tests
// and cross-level lookups follow parent and child links, the stencils follow the neighbour
// descriptors, and the hash walk follows the hash values. Returns DCG_E_ARG with the first
// violation.
int checkDCGridTables(const DCGridPlan& pl, const StateTables& t, std::string& err) {
  const int L = pl.L;
  auto bad = [&](const std::string& m) {
    err = "state rejected: " + m;
    return DCG_E_ARG;
  };
  uint32_t loads[DCG_MAX_LEVELS] = {};
  for (int l = 0; l < DCG_MAX_LEVELS; l++) {
    const uint32_t n = t.ctl[l];  // CTL_LOADS = 0
    if (l >= L) {
      if (n != 0) return bad("block load of a level beyond the grid's levels");
      continue;
    }
    if (n > pl.maxL[l]) return bad("block load of level " + std::to_string(l) + " exceeds its capacity");
    if (l >= pl.S && n != pl.fullL[l]) return bad("dense level " + std::to_string(l) + " is not full");
    loads[l] = n;
  }
  // level of every block: the allocated prefix of each slice holds that level, the rest is free
  std::vector<int8_t> lv(pl.B, -1);
  for (int l = 0; l < L; l++)
    for (uint32_t k = 0; k < pl.maxL[l]; k++) {
      const uint32_t b = pl.offs[l] + k;
      const uint8_t want = k < loads[l] ? (uint8_t)l : kFree;
      if (t.lev[b] != want) return bad("block " + std::to_string(b) + " level does not match the block loads");
      if (k < loads[l]) lv[b] = (int8_t)l;
    }
  uint64_t used = 0;
  for (int l = 0; l < L; l++) used += pl.maxL[l];
  for (uint64_t b = used; b < pl.B; b++)
    if (t.lev[b] != kFree) return bad("block " + std::to_string(b) + " beyond the level slices is not free");
  auto live = [&](uint32_t b, int l) { return b < pl.B && lv[b] == l; };
  const int bw = pl.bw;
  for (uint32_t b = 0; b < pl.B; b++) {
    const int l = lv[b];
    if (l < 0) continue;
    const int32_t* p = t.pos + 3 * (size_t)b;
    const int ext[3] = {pl.rx[l] * bw, pl.ry[l] * bw, pl.rz[l] * bw};
    for (int a = 0; a < 3; a++)
      if (p[a] < 0 || p[a] >= ext[a] || p[a] % bw) return bad("block " + std::to_string(b) + " position outside its level");
    const uint32_t par = t.par[b];
    if (l == L - 1 ? par != kNone : (par != kNone && !live(par / 8, l + 1)))
      return bad("block " + std::to_string(b) + " parent link");
    for (int s = 0; s < 8; s++) {
      const uint32_t c = t.chi[8 * (size_t)b + s];
      if (c == kNone) continue;
      if (l == 0 || !live(c, l - 1)) return bad("block " + std::to_string(b) + " child link");
    }
    for (int d = 0; d < 27; d++) {
      const int32_t* nd = t.nd + 4 * (27 * (size_t)b + d);
      const uint32_t x = (uint32_t)nd[0], sh = x >> 27, nb = x & kBlockMask;
      if (d == 13) continue;  // the block itself, never read
      if (sh == kOut) continue;
      if ((int)sh > L - 1 - l || !live(nb, l + (int)sh))
        return bad("block " + std::to_string(b) + " neighbour descriptor " + std::to_string(d));
    }
  }
  for (uint32_t i = 0; i < pl.H; i++)
    if (t.hval[i] != kNone && t.hval[i] >= pl.B) return bad("hash value " + std::to_string(i));
  return DCG_OK;
}

}  // namespace dcg

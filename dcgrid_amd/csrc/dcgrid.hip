// dcgrid.hip — host driver of the DCGrid back-end, the counterpart of FluidSimulationDCGrid
// (src/dcgrid/fluid_simulation_dcgrid.{h,cu}): level/budget planner, allocation, reset/init and
// the launch schedule of the step methods. Nothing here waits for the device: block counts,
// move limits and selection results stay in device control words, and kernels whose work depends
// on them are launched over the static per-level capacity and exit early.
#include <algorithm>
#include <cstring>
#include <type_traits>

#include <hip/hip_ext.h>

#include "dcg_internal.hpp"
#include "dcgrid_fluid.hpp"
#include "dcgrid_solve.hpp"
#include "dcgrid_topo.hpp"
#include "dcgrid_adjoint.hpp"
#include "render.hpp"

namespace dcg {

// Ray-march samples of the DCGrid (src/dcgrid/dcgrid_rendering.cu:6-64): the finest block
// containing the sample's cell; coarse = that block's cell at (cell mod BW) in finest
// coordinates (the reference's indexing, also for coarser blocks); precise = trilinear over the
// block's apron around the position in the block's level.
template <int BW, class G>
struct DCMarchSampler {
  G g;
  __device__ void sample(const dcg_params& P, int precise, float3 pos, MarchAux& a) const {
    constexpr int BB = Geo<BW>::BB, BV = Geo<BW>::BV;
    const int ix = clampi((int)floorf(pos.x), 0, g.gx - 1);
    const int iy = clampi((int)floorf(pos.y), 0, g.gy - 1);
    const int iz = clampi((int)floorf(pos.z), 0, g.gz - 1);
    a.ipos = make_int3(ix, iy, iz);
    const uint32_t b = findBlock<BW>(g, ix, iy, iz, 0);
    if (!precise) {
      const uint32_t c = b * BV + loc<BB>(ix % BW, iy % BW, iz % BW);
#pragma unroll
      for (int k = 0; k < 8; k++) a.cell[k] = c;
      a.density = (float)g.den[c];
      return;
    }
    const float scaleInv = 1.f / (float)(1 << g.lev[b]);
    const float x = pos.x * scaleInv - .5f, y = pos.y * scaleInv - .5f, z = pos.z * scaleInv - .5f;
    const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
    a.frac = make_float3(x - xf, y - yf, z - zf);
    const int3 bp = g.pos[b];
    const int ai = clampi((int)(xf + 1.f - (float)bp.x), 0, BW);
    const int aj = clampi((int)(yf + 1.f - (float)bp.y), 0, BW);
    const int ak = clampi((int)(zf + 1.f - (float)bp.z), 0, BW);
#pragma unroll
    for (int k = 0; k < 8; k++) a.cell[k] = apronCell<BW>(g, b, ai + ((k >> 2) & 1), aj + ((k >> 1) & 1), ak + (k & 1));
    a.density = marchInterp(g.den, a);
  }
  __device__ float3 velocity(const MarchAux& a) const { return marchInterp3(g.vel, a); }
  __device__ float fluidity(const MarchAux& a) const { return marchInterp(g.flu, a); }
};

static inline int iDivUpH(int a, int b) { return (a % b != 0) ? (a / b + 1) : (a / b); }
static inline unsigned blocksFor(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }
// Workgroups for a grid-stride walk over at most `blocks` allocated blocks: enough to fill the
// chip (about four resident workgroups per CU), never more than there are block slots.
template <int BW>
static inline unsigned slotGrid(size_t blocks) {
  const unsigned cap = Geo<BW>::BPW == 1 ? 1024u : 2048u;
  const unsigned n = blocksFor(blocks, Geo<BW>::BPW);
  return n == 0 ? 1u : (n < cap ? n : cap);
}

// Host conversions of the 16-bit storage formats (dcg_export_field / dcg_import_field): exact
// widening, and rounding to nearest even as the device stores do.
static float widenF16(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
  uint32_t u;
  if (e == 0x1F) {
    u = sign | 0x7F800000u | (m << 13);
  } else if (e != 0) {
    u = sign | ((e + 112u) << 23) | (m << 13);
  } else {
    float f = (float)m * 5.9604644775390625e-8f;  // m * 2^-24, exact
    std::memcpy(&u, &f, 4);
    u |= sign;
  }
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static uint16_t narrowF16(float f) {
  uint32_t a;
  std::memcpy(&a, &f, 4);
  const uint16_t sign = (uint16_t)((a >> 16) & 0x8000u);
  a &= 0x7FFFFFFFu;
  if (a >= 0x47800000u) return sign | (a > 0x7F800000u ? 0x7E00u : 0x7C00u);  // overflow, inf, NaN
  if (a < 0x38800000u) {  // subnormal result: adding 0.5 rounds to the 2^-24 grid (nearest even)
    float t;
    std::memcpy(&t, &a, 4);
    t += 0.5f;
    uint32_t tb;
    std::memcpy(&tb, &t, 4);
    return sign | (uint16_t)(tb - 0x3F000000u);
  }
  a += 0xC8000FFFu + ((a >> 13) & 1u);  // exponent rebias (15 - 127) << 23, round half to even
  return sign | (uint16_t)(a >> 13);
}
static float widenBF16(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static uint16_t narrowBF16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
template <class SF>
struct StorageHost;
template <>
struct StorageHost<StoreF16> {
  static float widen(uint16_t h) { return widenF16(h); }
  static uint16_t narrow(float f) { return narrowF16(f); }
};
template <>
struct StorageHost<StoreBF16> {
  static float widen(uint16_t h) { return widenBF16(h); }
  static uint16_t narrow(float f) { return narrowBF16(f); }
};

// SF: storage type of the fields in HBM (float, StoreF16 or StoreBF16).
template <int BW, class SF>
struct DCGridSim : SimBase {
  using G = Geo<BW>;
  int gx = 0, gy = 0, gz = 0, L = 0, S = 0;
  uint32_t B = 0, H = 0;
  size_t C = 0;
  uint32_t maxL[DCG_MAX_LEVELS] = {}, fullL[DCG_MAX_LEVELS] = {}, offs[DCG_MAX_LEVELS] = {};
  uint32_t hsize[DCG_MAX_LEVELS] = {}, hoff[DCG_MAX_LEVELS] = {};
  DGT<SF> g{};
  float* sbs = nullptr;   // 8B + 1 (last entry stays -FLT_MAX)
  float* bs = nullptr;    // B
  float* feat = nullptr;  // 8B
  unsigned long long* cand = nullptr;   // selected move candidates of a level
  unsigned long long* cand1 = nullptr;  // selected destination subblocks of a level
  uint32_t* mcBuf = nullptr;  // sorted move candidates, per-level region at offs[l]
  uint32_t* dcBuf = nullptr;  // sorted destination subblocks, region at offs[l] + l
  uint32_t* refList = nullptr;
  float* stats = nullptr;
  unsigned long long* cellStepCounter = nullptr;
  bool customFluidity = false;

  int kind() const override { return DCG_KIND_DCGRID; }
  bool fluidityOn() const { return P.enable_additional_solids || customFluidity; }
  // temperature is advected and drives buoyancy: sphere source or moist atmosphere
  bool thermo() const { return P.source_mode == DCG_SOURCE_SPHERE_BUOYANT || P.moist; }

  // Planner: fluid_simulation_dcgrid.cu:9-140.
  int create(const int32_t size[3], uint64_t maxBlocks, const dcg_params& p) {
    P = p;
    gx = size[0];
    gy = size[1];
    gz = size[2];
    DCGridPlan pl;
    std::string perr;
    const int pst = planDCGrid(BW, size, maxBlocks, P, pl, perr);
    if (pst) return fail(pst, perr);
    L = pl.L;
    S = pl.S;
    B = pl.B;
    H = pl.H;
    for (int l = 0; l < L; l++) {
      maxL[l] = pl.maxL[l];
      fullL[l] = pl.fullL[l];
      offs[l] = pl.offs[l];
      hsize[l] = pl.hsize[l];
      hoff[l] = pl.hoff[l];
    }
    C = (size_t)B * G::BV;

    g.gx = gx;
    g.gy = gy;
    g.gz = gz;
    g.L = L;
    g.S = S;
    g.B = B;
    for (int l = 0; l < L; l++) {
      g.offs[l] = offs[l];
      g.maxL[l] = maxL[l];
      g.fullL[l] = fullL[l];
      g.rx[l] = iDivUpH(gx, BW << l);
      g.ry[l] = iDivUpH(gy, BW << l);
      g.rz[l] = iDivUpH(gz, BW << l);
      g.hoff[l] = hoff[l];
      g.hsize[l] = hsize[l];
    }
    DCG_TRY(createStream());
    DCG_TRY(alloc(&g.pos, B));
    DCG_TRY(alloc(&g.lev, B));
    DCG_TRY(alloc(&g.flg, B));
    DCG_TRY(alloc(&g.par, B));
    DCG_TRY(alloc(&g.chi, (size_t)8 * B));
    DCG_TRY(alloc(&g.nd, (size_t)27 * B));
    DCG_TRY(alloc(&g.hkey, H));
    DCG_TRY(alloc(&g.hval, H));
    DCG_TRY(alloc(&g.ctl, CTL_WORDS));
    DCG_TRY(alloc(&g.den, C));
    DCG_TRY(alloc(&g.tem, C));
    DCG_TRY(alloc(&g.flu, C));
    DCG_TRY(alloc(&g.div, C));
    DCG_TRY(alloc(&g.pre, C));
    DCG_TRY(alloc(&g.tpre, C));
    DCG_TRY(alloc(&g.tden, C));
    DCG_TRY(alloc(&g.ttem, C));
    DCG_TRY(alloc(&g.vel, C));
    DCG_TRY(alloc(&g.tvel, C));
    DCG_TRY(alloc(&g.vort, C));
    if (P.moist) {  // the vapor field exists only for a moist atmosphere (fixed at creation)
      DCG_TRY(alloc(&g.vap, C));
      DCG_TRY(alloc(&g.tvap, C));
    }
    DCG_TRY(alloc(&sbs, (size_t)8 * B + 1));
    DCG_TRY(alloc(&bs, B));
    DCG_TRY(alloc(&feat, (size_t)8 * B));
    DCG_TRY(alloc(&cand, (size_t)8 * B + 16));
    DCG_TRY(alloc(&cand1, (size_t)8 * B + 16));
    DCG_TRY(alloc(&mcBuf, (size_t)B + 16));
    DCG_TRY(alloc(&dcBuf, (size_t)B + 16));
    DCG_TRY(alloc(&refList, (size_t)8 * B));
    DCG_TRY(alloc(&stats, B));
    DCG_TRY(alloc(&cellStepCounter, 1));
    const float negMax = -FLT_MAX;
    uint32_t negBits;
    std::memcpy(&negBits, &negMax, 4);
    DCG_TRY(fill32(sbs, negBits, (size_t)8 * B + 1));
    DCG_TRY(queryResidentGrids());
    return reset();
  }
  // The grids of the occupancy-sized walks are asked of the runtime here, outside any stream
  // capture.
  int queryResidentGrids() {
    constexpr int WGA = AdvGeo<BW>::WG;
    unsigned n = 0;
    DCG_TRY(residentGrid(k_advect_velocity<BW, true, true, DGT<SF>>, WGA, B, n));
    DCG_TRY(residentGrid(k_advect_velocity<BW, true, false, DGT<SF>>, WGA, B, n));
    DCG_TRY(residentGrid(k_advect_velocity<BW, false, true, DGT<SF>>, WGA, B, n));
    DCG_TRY(residentGrid(k_advect_velocity<BW, false, false, DGT<SF>>, WGA, B, n));
    DCG_TRY(residentGrid(k_advect_density<BW, true, true, true, DGT<SF>>, WGA, B, n));
    DCG_TRY(residentGrid(k_advect_density<BW, false, true, true, DGT<SF>>, WGA, B, n));
    DCG_TRY(residentGrid(k_advect_density<BW, true, true, false, DGT<SF>>, WGA, B, n));
    DCG_TRY(residentGrid(k_advect_density<BW, true, false, false, DGT<SF>>, WGA, B, n));
    DCG_TRY(residentGrid(k_advect_density<BW, false, true, false, DGT<SF>>, WGA, B, n));
    DCG_TRY(residentGrid(k_advect_density<BW, false, false, false, DGT<SF>>, WGA, B, n));
    if (twoCellStencils()) {
      constexpr int WGC = CptGeo<BW, CPT2>::WG;
      DCG_TRY(residentGrid(k_divergence<BW, true, CPT2, DGT<SF>>, WGC, B, n));
      DCG_TRY(residentGrid(k_divergence<BW, false, CPT2, DGT<SF>>, WGC, B, n));
      DCG_TRY(residentGrid(k_apply_pressure<BW, true, CPT2, DGT<SF>>, WGC, B, n));
      DCG_TRY(residentGrid(k_apply_pressure<BW, false, CPT2, DGT<SF>>, WGC, B, n));
      DCG_TRY(residentGrid(k_features<BW, true, CPT2, DGT<SF>>, WGC, B, n));
      DCG_TRY(residentGrid(k_features<BW, false, CPT2, DGT<SF>>, WGC, B, n));
    }
    return DCG_OK;
  }

  // Zero bits are +0 in every storage format; element sizes are multiples of 2 bytes and C is
  // even, so the byte count is a multiple of 4.
  template <class T>
  int zeroField(T* p, size_t n) {
    return fill32(p, 0, n * sizeof(T) / 4);
  }
  // reset(), fluid_simulation_dcgrid.cu:212-261
  int reset() override {
    DCG_TRY(fill32(g.hkey, 0xFFFFFFFFu, H));
    DCG_TRY(fill32(g.hval, NFU, H));
    DCG_TRY(fill32(g.pos, 0, (size_t)3 * B));
    DCG_TRY(fill8(g.lev, FREE, B));
    DCG_TRY(fill32(g.flg, 0, B));
    DCG_TRY(fill32(g.par, NFU, B));
    DCG_TRY(fill32(g.chi, NFU, (size_t)8 * B));
    DCG_TRY(fill32(g.nd, NFU, (size_t)4 * 27 * B));
    DCG_TRY(zeroField(g.den, C));
    DCG_TRY(zeroField(g.tem, C));
    DCG_TRY(zeroField(g.vel, C));
    DCG_TRY(zeroField(g.flu, C));
    DCG_TRY(zeroField(g.div, C));
    DCG_TRY(zeroField(g.pre, C));
    DCG_TRY(zeroField(g.tpre, C));
    DCG_TRY(zeroField(g.vort, C));
    if (g.vap) DCG_TRY(zeroField(g.vap, C));
    DCG_TRY(fill32(g.ctl, 0, CTL_WORDS));
    DCG_TRY(fill32(cellStepCounter, 0, 2));
    uint32_t loads[DCG_MAX_LEVELS] = {};
    for (int l = 0; l < L; l++) loads[l] = maxL[l] == fullL[l] ? maxL[l] : 0;
    DCG_HIP(hipMemcpyAsync(g.ctl + CTL_LOADS, loads, sizeof(loads), hipMemcpyHostToDevice, stream));
    DCG_TRY(sync());
    customFluidity = false;
    warmValid = false;
    return init();
  }
  // init(), fluid_simulation_dcgrid.cu:190-210
  int init() override {
    for (int level = S; level < L; level++) {
      const size_t n = (size_t)g.rx[level] * g.ry[level] * g.rz[level];
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_activate_dense",
                 (k_activate_dense<BW><<<blocksFor(n, 256), 256, 0, stream>>>(g, P, level)));
    }
    for (int i = 0; i < 5; i++) DCG_TRY(adaptTopology());
    return DCG_OK;
  }

  int subScores() {
    if (P.score_mode == DCG_SCORE_FEATURES)
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_sub_scores_feat",
                 (k_sub_scores_feat<<<blocksFor((size_t)8 * B, 256), 256, 0, stream>>>(g, feat, sbs)));
    else
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_sub_scores_geo",
                 (k_sub_scores_geo<BW><<<blocksFor((size_t)8 * B, 256), 256, 0, stream>>>(g, sbs)));
    return DCG_OK;
  }

  // adaptTopology(), fluid_simulation_dcgrid.cu:320-346. counter: the step's active cells are
  // added to it (dcg_step), or null.
  int adaptTopology() override { return adapt(nullptr); }
  int adaptTopologyCounted() override { return adapt(cellStepCounter); }
  int adapt(unsigned long long* counter) {
    if (P.score_mode == DCG_SCORE_FEATURES) {  // vorticity and subblock features in one pass
      if (fluidityOn())
        DCG_TRY(stencilLaunch(DCG_K_TOPOLOGY, "k_features", k_features<BW, true, 1, DGT<SF>>,
                                    k_features<BW, true, CPT2, DGT<SF>>, P, feat));
      else
        DCG_TRY(stencilLaunch(DCG_K_TOPOLOGY, "k_features", k_features<BW, false, 1, DGT<SF>>,
                                    k_features<BW, false, CPT2, DGT<SF>>, P, feat));
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_scores", (k_scores<BW, true><<<blocksFor(B, 256), 256, 0, stream>>>(g, feat, sbs, bs)));
    } else {
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_scores", (k_scores<BW, false><<<blocksFor(B, 256), 256, 0, stream>>>(g, feat, sbs, bs)));
    }
    // moveBlocks (:348-437): levels in order, each level's match marks blocks the next reads
    for (int l = 0; l < L - 1; l++) {
      uint32_t* mc = mcBuf + offs[l];
      uint32_t* dc = dcBuf + offs[l] + l;
      const unsigned rk = blocksFor((size_t)maxL[l] + 1, 256);
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_select_pair", (k_select_pair<<<2, 1024, 0, stream>>>(g, l, bs, sbs, cand, cand1)));
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_rank_pair", (k_rank_pair<<<dim3(rk, 2), 256, 0, stream>>>(g, l, cand, cand1, mc, dc)));
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_match",
                 (k_match<<<1, 1024, 0, stream>>>(g, l, l == L - 2 ? 1 : 0, mc, dc, bs, sbs)));
    }
    // moves coarse to fine; the first launch also clears the hash for the refill (:326-332)
    for (int l = L - 2; l >= 0; l--)
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_move",
                 (k_move<BW><<<blocksFor(maxL[l], 256), 256, 0, stream>>>(g, l, mcBuf + offs[l], dcBuf + offs[l] + l,
                                                                          l == L - 2 ? H : 0u)));
    if (L > 1) DCG_LAUNCH(DCG_K_TOPOLOGY, "k_hash_refill", (k_hash_refill<BW><<<blocksFor(B, 256), 256, 0, stream>>>(g)));
    // refineSubblocks (:439-483)
    if (L > 1) {
      DCG_TRY(subScores());
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_refine_select", (k_refine_select<<<L - 1, 1024, 0, stream>>>(g, P, sbs, refList)));
      uint32_t mxc = 0;
      for (int l = 0; l < L - 1; l++) mxc = std::max(mxc, maxL[l]);
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_refine_apply",
                 (k_refine_apply<BW><<<dim3(blocksFor(mxc, 256), L - 1), 256, 0, stream>>>(g, refList)));
    }
    // apron refresh and value propagation (:336-345)
    DCG_LAUNCH(DCG_K_TOPOLOGY, "k_refresh", (k_refresh<BW><<<blocksFor((size_t)27 * B, 256), 256, 0, stream>>>(g, counter)));
    const bool src = thermo();
    for (int l = L - 2; l >= 0; l--) {
      const unsigned n = std::min(blocksFor(maxL[l], G::BPW), 256u);  // grid-stride over the touched blocks
      if (src)
        DCG_LAUNCH(DCG_K_TOPOLOGY, "k_propagate", (k_propagate<BW, true><<<n, G::WG, 0, stream>>>(g, P, l, mcBuf)));
      else
        DCG_LAUNCH(DCG_K_TOPOLOGY, "k_propagate", (k_propagate<BW, false><<<n, G::WG, 0, stream>>>(g, P, l, mcBuf)));
    }
    return DCG_OK;
  }

  template <class T>
  int restrictAll(T* ch) {
    for (int l = 0; l < L - 1; l++) {
      const size_t groups = (size_t)maxL[l] * (G::BV / 8);
      DCG_LAUNCH(DCG_K_RESTRICT, "k_restrict", (k_restrict<BW, T><<<blocksFor(groups, 256) < 2048 ? blocksFor(groups, 256) : 2048, 256, 0, stream>>>(g, l, ch)));
    }
    return DCG_OK;
  }

  // Workgroups of a grid-stride walk over at most `blocks` blocks that fill the chip once: the
  // workgroups of `kernel` resident per CU (at its registers and LDS) times the CUs, asked of the
  // runtime once per kernel.
  std::vector<std::pair<const void*, unsigned>> residentPerChip;
  template <class K>
  int residentGrid(K kernel, int threads, size_t blocks, unsigned& n) {
    const void* key = (const void*)kernel;
    unsigned chip = 0;
    for (const auto& e : residentPerChip)
      if (e.first == key) chip = e.second;
    if (chip == 0) {
      int dev = 0, cus = 0, per = 0;
      DCG_HIP(hipGetDevice(&dev));
      DCG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      DCG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0));
      chip = (unsigned)std::max(per, 1) * (unsigned)std::max(cus, 1);
      residentPerChip.emplace_back(key, chip);
    }
    const unsigned want = blocksFor(blocks, G::BPW);
    n = want == 0 ? 1u : std::min(want, chip);
    return DCG_OK;
  }

  // Stencil walks over all allocated blocks take two cells per thread (BW 8) when the budget is
  // large enough that workgroups walk many blocks; with a few thousand blocks a workgroup
  // handles about one, and one cell per thread is faster.
  static constexpr uint32_t kTwoCellBudget = 16384;
  static constexpr int CPT2 = BW == 8 ? 2 : 1;  // the two-cell shape (BW 8 only)
  bool twoCellStencils() const { return CPT2 == 2 && B >= kTwoCellBudget; }
  template <class K1, class K2, class... Args>
  int stencilLaunch(int kind, const char* name, K1 one, K2 two, Args... args) {
    if (twoCellStencils()) {
      constexpr int WGC = CptGeo<BW, CPT2>::WG;
      unsigned nWG = 0;
      DCG_TRY(residentGrid(two, WGC, B, nWG));
      DCG_LAUNCH(kind, name, (two<<<nWG, WGC, 0, stream>>>(g, args...)));
    } else {
      DCG_LAUNCH(kind, name, (one<<<slotGrid<BW>(B), G::WG, 0, stream>>>(g, args...)));
    }
    return DCG_OK;
  }

  // advectVelocity(), fluid_simulation_dcgrid.cu:263-268
  int advectVelocity() override {
    const bool buoy = thermo(), fl = fluidityOn();
    constexpr int WGA = AdvGeo<BW>::WG;
#define ADV_V(F, Bu)                                                                                              \
  do {                                                                                                            \
    unsigned nWG = 0;                                                                                             \
    DCG_TRY(residentGrid(k_advect_velocity<BW, F, Bu, DGT<SF>>, WGA, B, nWG));                                          \
    DCG_LAUNCH(DCG_K_ADVECT_VELOCITY, "k_advect_velocity", (k_advect_velocity<BW, F, Bu><<<nWG, WGA, 0, stream>>>(g, P))); \
  } while (0)
    if (fl && buoy) ADV_V(true, true);
    else if (fl) ADV_V(true, false);
    else if (buoy) ADV_V(false, true);
    else ADV_V(false, false);
#undef ADV_V
    std::swap(g.vel, g.tvel);  // replaces the D2D copy (:265)
    return restrictAll(g.vel);
  }
  // advectDensity(), fluid_simulation_dcgrid.cu:313-318
  int advectDensity() override {
    const bool src = thermo(), moist = P.moist != 0, fl = fluidityOn();
    constexpr int WGA = AdvGeo<BW>::WG;
#define ADV_D(F, Sr, M)                                                                                           \
  do {                                                                                                            \
    unsigned nWG = 0;                                                                                             \
    DCG_TRY(residentGrid(k_advect_density<BW, F, Sr, M, DGT<SF>>, WGA, B, nWG));                                        \
    DCG_LAUNCH(DCG_K_ADVECT_DENSITY, "k_advect_density", (k_advect_density<BW, F, Sr, M><<<nWG, WGA, 0, stream>>>(g, P))); \
  } while (0)
    if (moist) {
      if (fl) ADV_D(true, true, true);
      else ADV_D(false, true, true);
    } else if (fl && src) ADV_D(true, true, false);
    else if (fl) ADV_D(true, false, false);
    else if (src) ADV_D(false, true, false);
    else ADV_D(false, false, false);
#undef ADV_D
    std::swap(g.den, g.tden);
    if (src) std::swap(g.tem, g.ttem);
    if (moist) std::swap(g.vap, g.tvap);
    // the advected scalars restricted together, one launch per level
    for (int l = 0; l < L - 1; l++) {
      const size_t groups = (size_t)maxL[l] * (G::BV / 8);
      const unsigned n = blocksFor(groups, 256) < 2048 ? blocksFor(groups, 256) : 2048;
      DCG_LAUNCH(DCG_K_RESTRICT, "k_restrict3",
                 (k_restrict3<BW, SF><<<n, 256, 0, stream>>>(g, l, g.den, src ? g.tem : nullptr, moist ? g.vap : nullptr)));
    }
    return DCG_OK;
  }
  int divergenceKernel() {
    if (fluidityOn())
      return stencilLaunch(DCG_K_DIVERGENCE, "k_divergence", k_divergence<BW, true, 1, DGT<SF>>,
                                  k_divergence<BW, true, CPT2, DGT<SF>>, P);
    return stencilLaunch(DCG_K_DIVERGENCE, "k_divergence", k_divergence<BW, false, 1, DGT<SF>>,
                                k_divergence<BW, false, CPT2, DGT<SF>>, P);
  }
  int divergence() {
    DCG_TRY(divergenceKernel());
    return restrictAll(g.div);
  }
  // One Jacobi half-sweep of a level (fwd: p -> p~, else p~ -> p); nb: blocks per sub-group
  // and iteration (1, 2 or 4); dense levels (l >= S) address their faces without descriptors.
  int jacobiLaunch(int level, bool fwd, int nb, unsigned n, float alpha) {
    auto* in = fwd ? g.pre : g.tpre;
    auto* out = fwd ? g.tpre : g.pre;
    const bool dense = level >= S;
#define JAC(NB, D) DCG_LAUNCH(DCG_K_JACOBI, "k_jacobi", (k_jacobi<BW, NB, D><<<n, G::WG, 0, stream>>>(g, level, in, out, alpha)))
    if (nb == 4 && dense) JAC(4, true);
    else if (nb == 4) JAC(4, false);
    else if (nb == 2 && dense) JAC(2, true);
    else if (nb == 2) JAC(2, false);
    else if (dense) JAC(1, true);
    else JAC(1, false);
#undef JAC
    return DCG_OK;
  }
  int jacobiOne(int level, bool fwd) {
    const int s = 1 << level;
    const float alpha = (float)(s * s) * P.dx * P.dx;
    return jacobiLaunch(level, fwd, 1, slotGrid<BW>(maxL[level]), alpha);
  }
  int prolongate(int level) {
    DCG_LAUNCH(DCG_K_PROLONGATE, "k_prolongate",
               (k_prolongate<BW><<<slotGrid<BW>(maxL[level]), G::WG, 0, stream>>>(g, level)));
    return DCG_OK;
  }
  int vorticity() {
    const unsigned nWG = slotGrid<BW>(B);
    if (fluidityOn())
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_vorticity", (k_vorticity<BW, true><<<nWG, G::WG, 0, stream>>>(g, P)));
    else
      DCG_LAUNCH(DCG_K_TOPOLOGY, "k_vorticity", (k_vorticity<BW, false><<<nWG, G::WG, 0, stream>>>(g, P)));
    return DCG_OK;
  }
  int runStage(int stage, int arg) override {
    const bool perLevel = stage == DCG_STAGE_JACOBI || stage == DCG_STAGE_JACOBI_INV || stage == DCG_STAGE_PROLONGATE;
    if (perLevel && (arg < 0 || arg >= L || (stage == DCG_STAGE_PROLONGATE && arg >= L - 1)))
      return fail(DCG_E_ARG, "level out of range");
    switch (stage) {
      case DCG_STAGE_DIVERGENCE: return divergenceKernel();
      case DCG_STAGE_RESTRICT:
        switch (arg) {
          case DCG_FIELD_DENSITY: return restrictAll(g.den);
          case DCG_FIELD_TEMPERATURE: return restrictAll(g.tem);
          case DCG_FIELD_DIVERGENCE: return restrictAll(g.div);
          case DCG_FIELD_PRESSURE: return restrictAll(g.pre);
          case DCG_FIELD_VELOCITY: return restrictAll(g.vel);
          default: return fail(DCG_E_ARG, "field cannot be restricted");
        }
      case DCG_STAGE_JACOBI: return jacobiOne(arg, true);
      case DCG_STAGE_JACOBI_INV: return jacobiOne(arg, false);
      case DCG_STAGE_PROLONGATE: return prolongate(arg);
      case DCG_STAGE_APPLY_PRESSURE: return applyKernel();
      case DCG_STAGE_VORTICITY: return vorticity();
      default: return fail(DCG_E_ARG, "unknown stage");
    }
  }
  int jacobiPairs(int level, int pairs) {
    const int s = 1 << level;
    const float alpha = (float)(s * s) * P.dx * P.dx;  // dcgrid_multigrid_solver.cu:11
    const unsigned n = slotGrid<BW>(maxL[level]);
    if (pairs <= 0) return DCG_OK;
    Pending batch;  // the run of back-to-back sweeps of one level is timed as one batch
    DCG_TRY(beginTimed(DCG_K_JACOBI, &batch));
    batchK = DCG_K_JACOBI;
    // more blocks per sub-group and iteration once the level spans several rounds of the grid
    // (the large levels, bound by reads in flight)
    const size_t rounds = (size_t)maxL[level] / ((size_t)n * G::BPW);
    const int nb = rounds >= 16 ? 4 : rounds >= 2 ? 2 : 1;  // four from 4 rounds measured slower
    for (int i = 0; i < pairs; i++)
      for (int h = 0; h < 2; h++)
        DCG_TRY(jacobiLaunch(level, h == 0, nb, n, alpha));
    batchK = -1;
    return endTimed(DCG_K_JACOBI, &batch, 2 * pairs);
  }
  bool graphCapable() const override { return true; }
  void graphKeyBytes(std::vector<char>& out) const override {
    out.clear();
    auto put = [&](const void* p, size_t n) { out.insert(out.end(), (const char*)p, (const char*)p + n); };
    put(&g, sizeof(g));
    put(&P, sizeof(P));
    put(&customFluidity, sizeof(customFluidity));
  }
  void hostSnapshot(std::vector<char>& out) const override {
    out.assign((const char*)&g, (const char*)&g + sizeof(g));
    out.push_back(customFluidity ? 1 : 0);
  }
  void hostRestore(const std::vector<char>& in) override {
    std::memcpy(&g, in.data(), sizeof(g));
    customFluidity = in[sizeof(g)] != 0;
  }
  // The step's active cells are counted by the refresh launch of adaptTopologyCounted().
  int countStep() override { return DCG_OK; }
  int cellSteps(uint64_t* out) override {
    unsigned long long v = 0;
    DCG_HIP(hipMemcpyAsync(&v, cellStepCounter, sizeof(v), hipMemcpyDeviceToHost, stream));
    DCG_TRY(sync());
    *out = v;
    return DCG_OK;
  }
  int applyKernel() {
    if (fluidityOn())
      return stencilLaunch(DCG_K_APPLY_PRESSURE, "k_apply_pressure", k_apply_pressure<BW, true, 1, DGT<SF>>,
                                  k_apply_pressure<BW, true, CPT2, DGT<SF>>, P);
    return stencilLaunch(DCG_K_APPLY_PRESSURE, "k_apply_pressure", k_apply_pressure<BW, false, 1, DGT<SF>>,
                                k_apply_pressure<BW, false, CPT2, DGT<SF>>, P);
  }
  int applyPressure() {
    DCG_TRY(applyKernel());
    return restrictAll(g.vel);
  }
  // project(), fluid_simulation_dcgrid.cu:270-294
  int project() override {
    DCG_TRY(divergence());
    DCG_TRY(jacobiPairs(L - 1, P.project_top_pairs));
    for (int l = L - 2; l >= 0; l--) {
      DCG_LAUNCH(DCG_K_PROLONGATE, "k_prolongate",
                 (k_prolongate<BW><<<slotGrid<BW>(maxL[l]), G::WG, 0, stream>>>(g, l)));
      DCG_TRY(jacobiPairs(l, P.project_level_pairs));
    }
    return applyPressure();
  }
  // projectLocal(), fluid_simulation_dcgrid.cu:296-311
  int projectLocal() override {
    DCG_TRY(divergence());
    for (int l = L - 1; l >= 0; l--) DCG_TRY(jacobiPairs(l, P.project_local_pairs));
    return applyPressure();
  }
  // ---------------------------------------------------------------- projection solve (extension)
  void* xBuf[2] = {nullptr, nullptr};  // iterates of every cell (TX): the canonical one and the
                                       // other of the update pair x' = x + e
  double* rhoBuf = nullptr;    // right-hand side / residual in density units (T = float or double)
  double* eBuf = nullptr;      // multigrid corrections (T)
  uint8_t* reachBuf = nullptr;   // finest level with cells inside each cell
  uint8_t* breachBuf = nullptr;  // finest level with cells inside each block
  uint32_t* svList = nullptr;    // active blocks per level solve, grid and colour (k_sv_lists)
  int4* svFaces = nullptr;       // the six face descriptors beside each list entry
  uint32_t* svCount = nullptr;   // their counts, device
  uint32_t* svCountMirror = nullptr;     // and in pinned host memory
  uint32_t* svCountMirrorDev = nullptr;  // (device address of the mirror)
  std::vector<uint32_t> lc;      // the counts of the current solve, on the host
  double* ueBuf = nullptr;     // pyramid of the top level: corrections, right-hand sides, activity
  double* urhoBuf = nullptr;
  uint8_t* uactBuf = nullptr;
  double* partBuf = nullptr;   // per-workgroup partial norms of the residual kernel
  SolveCtl* ctlBuf = nullptr;  // refinement state per level
  SolveCtl* ctlMirror = nullptr;     // its copy in pinned host memory (written by k_sv_decide)
  SolveCtl* ctlMirrorDev = nullptr;  // the device address of that copy
  std::vector<UDimArg> pyr;    // pyramid grids below the top level (cell size doubling)
  int pyrCoarse = 0;           // first pyramid grid of the single-workgroup coarse kernel
  uint8_t* coarseCache = nullptr;  // per level solve: what the coarse kernel derives from the activity
  uint32_t* coarseTags = nullptr;  // the solve each slot belongs to
  uint32_t solveSerial = 0;        // tag of the current projection solve (never 0)
  dcg_solve_info rep{};
  int predCycles[DCG_MAX_LEVELS] = {};  // cycles the last solve of each level took
  bool warmValid = false;  // the pressure field holds a projection's solution (warm start)
  static constexpr unsigned kSolveWG = 1024;  // workgroups of the residual kernel at most

  unsigned cellGrid(int level) const {
    const unsigned n = blocksFor((size_t)maxL[level] * G::BV, 256);
    return n < 2048u ? n : 2048u;
  }
  int solveSetup() {
    if (xBuf[0]) return DCG_OK;
    const int top = L - 1;
    const int tx = g.rx[top] * BW, ty = g.ry[top] * BW, tz = g.rz[top] * BW;
    if ((tx << top) != g.gx || (ty << top) != g.gy || (tz << top) != g.gz)
      return fail(DCG_E_UNSUPPORTED, "projection solve: the grid size must be a multiple of the top level's blocks");
    const size_t C = (size_t)B * G::BV;
    double* x0 = nullptr;
    double* x1 = nullptr;
    DCG_TRY(alloc(&x0, C));
    DCG_TRY(alloc(&x1, C));
    xBuf[0] = x0;
    xBuf[1] = x1;
    DCG_TRY(alloc(&rhoBuf, C));
    DCG_TRY(alloc(&eBuf, C));
    DCG_TRY(alloc(&reachBuf, C));
    DCG_TRY(alloc(&breachBuf, B));
    DCG_TRY(alloc(&svList, (size_t)2 * L * B));
    DCG_TRY(alloc(&svFaces, (size_t)12 * L * B));
    DCG_TRY(alloc(&svCount, (size_t)2 * L * L));
    DCG_TRY(allocHost(&svCountMirror, (size_t)2 * L * L));
    DCG_HIP(hipHostGetDevicePointer((void**)&svCountMirrorDev, svCountMirror, 0));
    DCG_TRY(alloc(&partBuf, 5 * (size_t)kSolveWG));
    DCG_TRY(alloc(&ctlBuf, (size_t)L));
    DCG_TRY(fill32(ctlBuf, 0u, (size_t)L * sizeof(SolveCtl) / 4));
    DCG_TRY(allocHost(&ctlMirror, (size_t)L));
    std::memset((void*)ctlMirror, 0, (size_t)L * sizeof(SolveCtl));
    DCG_HIP(hipHostGetDevicePointer((void**)&ctlMirrorDev, ctlMirror, 0));
    // pyramid: halve while every size stays even and at least 2
    size_t off = 0;
    int w = tx, h = ty, d = tz;
    while (w % 2 == 0 && h % 2 == 0 && d % 2 == 0 && std::min(w, std::min(h, d)) >= 4 && (int)pyr.size() < kMaxPyr) {
      w /= 2;
      h /= 2;
      d /= 2;
      pyr.push_back(UDimArg{w, h, d, (uint32_t)off});
      off += (size_t)w * h * d;
    }
    // the coarse end that fits the single-workgroup kernel
    size_t tail = 0;
    pyrCoarse = (int)pyr.size();
    for (int j = (int)pyr.size() - 1; j >= 0; j--) {
      tail += (size_t)pyr[j].w * pyr[j].h * pyr[j].d;
      if (tail > (size_t)kCoarseCells) break;
      pyrCoarse = j;
    }
    DCG_TRY(alloc(&ueBuf, off + 1));
    DCG_TRY(alloc(&urhoBuf, off + 1));
    DCG_TRY(alloc(&uactBuf, off + 1));
    DCG_TRY(alloc(&coarseCache, (size_t)L * kCoarseCacheBytes));
    DCG_TRY(alloc(&coarseTags, (size_t)L));
    DCG_TRY(fill32(coarseTags, 0u, (size_t)L));
    return DCG_OK;
  }
  // Ghost factor of a Dirichlet face on the grid m levels coarser than the solved level: the
  // finest grid puts the zero half a fine cell beyond the region's face, i.e. at
  // d = (1/2 + 2^-(m+1)) coarse cells from the boundary cell's centre; the linear ghost through
  // it is (1 - 1/d) e_c.
  static float ghostFactor(int m) {
    if (m <= 0) return 0.f;
    const double d = 0.5 + std::ldexp(1.0, -(m + 1));
    return (float)(1.0 - 1.0 / d);
  }
  float levelH2(int level) const {
    const float s = (float)(1 << level);
    return s * s * P.dx * P.dx;  // alpha of k_dcgrid_jacobi, dcgrid_multigrid_solver.cu:11
  }
  // List of the active blocks of grid k in the V-cycle of level lmin, colour c, and its length.
  const uint32_t* listOf(int lmin, int c) const { return svList + (size_t)(lmin * 2 + c) * B; }
  const int4* facesOf(int lmin, int c) const { return svFaces + (size_t)(lmin * 2 + c) * B * 6; }
  uint32_t countOf(int lmin, int k, int c) const { return lc[(size_t)(lmin * L + k) * 2 + c]; }
  unsigned listGrid(uint32_t n) const {
    const unsigned w = blocksFor(n, G::BPW);
    return w < 1024u ? w : 1024u;
  }
  // Smoothing of a grid before or after its coarse correction: svPasses passes (solve_smooth,
  // default 1, at most 2), each the blocks of one colour then the other with kSvInner
  // red-black sweeps in LDS, i.e. 2 * kSvInner * passes <= 16 sweeps per grid and cycle. A
  // pass costs a launch per colour and grid while an inner sweep costs a few LDS round trips,
  // so the default takes one pass of four sweeps. Measured (MI355X, every projection solved to
  // 1e-8), passes x inner sweeps: 2 x 1 and 1 x 2 diverge on the finest level of the BW-4
  // reference scene; 1 x 4 converges on every benchmark and test scene in 7-14 cycles per
  // level, with fewer launches per step than 2 x 2 (configuration 2 at 512^3: 7.9 against
  // 9.3 ms per step, 34.8 against 31.6 cycles); 1 x 5 .. 1 x 8 take a cycle or two fewer
  // but more time (the pyramid's sweeps grow with the inner count).
  static constexpr int kSvInner = 4;
  int smoothSweeps() const { return P.solve_smooth > 0 ? std::min(P.solve_smooth, 2) : 1; }
  int svInner = kSvInner, svPasses = 1;
  template <class T>
  int smoothBlocks(int k, int lmin, int nu, bool down, const SolveCtl* ctl) {
    (void)nu;
    const float h2 = levelH2(k), gm = ghostFactor(k - lmin);
    for (int pass = 0; pass < svPasses; pass++)
      for (int colour = 0; colour < 2; colour++) {
        const uint32_t n = countOf(lmin, k, colour);
        if (n == 0) continue;
        // the first pass of the down leg starts from a zero correction: the first colour's
        // halos are zero too (the second colour's are what the first wrote)
        const bool zero = down && pass == 0;
        const bool haloZero = zero && (colour == 0 || countOf(lmin, k, 0) == 0);
        if (timingMask & (1u << DCG_K_SOLVE_SMOOTH))  // e read (unless zero), rho read, e written
          timingBytes[DCG_K_SOLVE_SMOOTH] += (uint64_t)n * G::BV * sizeof(T) * (zero ? 2 : 3);
        Pending ev;
        DCG_TRY(takeEvents(DCG_K_SOLVE_SMOOTH, &ev));
        const DGBase gb = g;
        const uint32_t* lst = listOf(lmin, colour);
        const int4* fcs = facesOf(lmin, colour);
        const int oz = zero ? 1 : 0, hz = haloZero ? 1 : 0;
        T* ep = (T*)eBuf;
        const T* rp = (const T*)rhoBuf;
        const uint8_t* reach = reachBuf;
        if (ev.k >= 0)
          hipExtLaunchKernelGGL(k_sv_smooth<BW, T>, dim3(listGrid(n)), dim3(G::WG), 0u, stream, ev.a, ev.b, 0u, gb, k,
                                lmin, lst, fcs, n, svInner, oz, hz, ep, rp, reach, h2, gm, ctl);
        else
          k_sv_smooth<BW, T><<<listGrid(n), G::WG, 0, stream>>>(gb, k, lmin, lst, fcs, n, svInner, oz, hz, ep, rp,
                                                                reach, h2, gm, ctl);
        DCG_TRY(checkLaunch("k_sv_smooth"));
        DCG_TRY(pushPending(&ev));
      }
    return DCG_OK;
  }
  template <class T>
  int smoothPyr(int j, int lmin, int nu, bool down, const SolveCtl* ctl) {
    (void)nu;
    const UDimArg& D = pyr[j];
    const unsigned tiles = (unsigned)(((D.w + 7) / 8) * ((D.h + 7) / 8) * ((D.d + 7) / 8));
    const unsigned n = tiles < 1024u ? tiles : 1024u;
    const float s = (float)(1 << (L + j)), h2 = s * s * P.dx * P.dx;
    const float gm = ghostFactor(L + j - lmin);
    for (int pass = 0; pass < svPasses; pass++)
      for (int colour = 0; colour < 2; colour++) {
        const bool zero = down && pass == 0;
        DCG_LAUNCH(DCG_K_SOLVE, "k_svu_smooth",
                   (k_svu_smooth<T><<<n, 512, 0, stream>>>(D, colour, svInner, zero ? 1 : 0, zero && colour == 0 ? 1 : 0,
                                                          (T*)ueBuf, (const T*)urhoBuf, uactBuf, h2, gm, ctl)));
      }
    return DCG_OK;
  }
  // One V-cycle for the correction of level lmin (right-hand side in rho on its cells).
  template <class T>
  int vcycle(int lmin, const SolveCtl* ctl) {
    T* e = (T*)eBuf;
    T* rho = (T*)rhoBuf;
    const int nu = smoothSweeps();
    const int top = L - 1;
    for (int k = lmin; k < top; k++) {
      DCG_TRY(smoothBlocks<T>(k, lmin, nu, true, ctl));
      const uint32_t n0 = countOf(lmin, k, 0), n1 = countOf(lmin, k, 1);
      if (n0 + n1)
        DCG_LAUNCH(DCG_K_SOLVE, "k_sv_restrict",
                   (k_sv_restrict<BW, T, false><<<listGrid(n0 + n1), G::WG, 0, stream>>>(
                       g, k, lmin, listOf(lmin, 0), listOf(lmin, 1), facesOf(lmin, 0), facesOf(lmin, 1), n0, n1, e,
                       rho, reachBuf, levelH2(k),
                       ghostFactor(k - lmin), nullptr, nullptr, UDimArg{}, ctl)));
    }
    DCG_TRY(smoothBlocks<T>(top, lmin, nu, true, ctl));
    // The pyramid below the top level serves a sparse level's solve only where the level's region
    // covers a sizeable part of the top level. Where the region's ancestors are at most 1/16 of the
    // top level's blocks, the cycle turns at the top level after its smoothing pass (about 9
    // launches and the coarse kernel fewer per cycle). Measured (cycles per step, ms per step):
    // configuration 2 at 512^3, whose sparse regions cover 5 % of the top level, 34.8 -> 33.8,
    // 6.8 -> 5.2 ms (two more passes at the top level took 34.4 cycles, 5.7 ms); where a region
    // covered 1/8 of the top level (configuration 1, its level 1) its solve took 25 cycles without
    // the pyramid instead of 13, so the pyramid stays there.
    const uint32_t topActive = countOf(lmin, top, 0) + countOf(lmin, top, 1);
    const bool smallRegion = lmin < g.S && (size_t)topActive * 16 <= (size_t)maxL[top];
    if (!pyr.empty() && !smallRegion) {
      T* ue = (T*)ueBuf;
      T* ur = (T*)urhoBuf;
      DCG_LAUNCH(DCG_K_SOLVE, "k_sv_restrict",
                 (k_sv_restrict<BW, T, true><<<slotGrid<BW>(maxL[top]), G::WG, 0, stream>>>(
                     g, top, lmin, nullptr, nullptr, nullptr, nullptr, 0u, 0u, e, rho, reachBuf, levelH2(top),
                     ghostFactor(top - lmin),
                     ur, uactBuf, pyr[0], ctl)));
      // grids 0..jd-1 by the multi-workgroup kernels, jd.. by the coarse kernel (or, when even
      // the coarsest grid exceeds its LDS, many smoothing passes on that grid)
      const int J = (int)pyr.size(), jc = pyrCoarse, jd = jc < J ? jc : J - 1;
      for (int j = 0; j < jd; j++) {
        DCG_TRY(smoothPyr<T>(j, lmin, nu, true, ctl));
        const float s = (float)(1 << (L + j)), h2 = s * s * P.dx * P.dx;
        const UDimArg& Dc = pyr[j + 1];
        DCG_LAUNCH(DCG_K_SOLVE, "k_svu_restrict",
                   (k_svu_restrict<T><<<blocksFor((size_t)Dc.w * Dc.h * Dc.d * 8, 256), 256, 0, stream>>>(
                       pyr[j], Dc, ue, ur, uactBuf, h2, ghostFactor(L + j - lmin), ctl)));
      }
      if (jc < J) {
        CoarseArgs ca{};
        uint32_t loff = 0;
        for (int j = jc; j < J; j++) {
          ca.D[j - jc] = UDimArg{pyr[j].w, pyr[j].h, pyr[j].d, loff};
          loff += (uint32_t)(pyr[j].w * pyr[j].h * pyr[j].d);
          const float s = (float)(1 << (L + j));
          ca.h2[j - jc] = s * s * P.dx * P.dx;
          ca.gm[j - jc] = ghostFactor(L + j - lmin);
        }
        ca.goff = pyr[jc].off;
        ca.n = J - jc;
        // three red-black sweeps per grid and leg on the coarse grids: the same cycle counts as
        // four on every benchmark and test scene, two added cycles on configurations 1 and 4
        ca.sweeps = 3;
        // coarsest grid: 8 sweeps converge 2^3 (as 40 did), flat grids such as 8x2x8 need 40
        const UDimArg& Db = pyr[J - 1];
        ca.bottom = std::max(Db.w, std::max(Db.h, Db.d)) <= 2 ? 8 : 40;
        ca.slot = lmin;
        ca.tag = solveSerial;
        DCG_LAUNCH(DCG_K_SOLVE, "k_svu_coarse",
                   (k_svu_coarse<T><<<1, kCoarseThreads, 0, stream>>>(ca, ue, ur, uactBuf, ctl, coarseCache,
                                                                       coarseTags)));
      } else {
        for (int i = 0; i < 20; i++) DCG_TRY(smoothPyr<T>(jd, lmin, nu, i == 0, ctl));
      }
      for (int j = jd - 1; j >= 0; j--) {
        const UDimArg& D = pyr[j];
        DCG_LAUNCH(DCG_K_SOLVE, "k_svu_prolong",
                   (k_svu_prolong<T><<<blocksFor((size_t)D.w * D.h * D.d, 256), 256, 0, stream>>>(D, pyr[j + 1], ue,
                                                                                                  uactBuf, ctl)));
        DCG_TRY(smoothPyr<T>(j, lmin, nu, false, ctl));
      }
      DCG_LAUNCH(DCG_K_SOLVE, "k_sv_prolong_top",
                 (k_sv_prolong_top<BW, T><<<cellGrid(top), 256, 0, stream>>>(g, lmin, e, reachBuf, ue, uactBuf, pyr[0],
                                                                             ctl)));
      DCG_TRY(smoothBlocks<T>(top, lmin, nu, false, ctl));
    } else {
      // no pyramid (a top level too small to halve): the top level is the bottom grid, smoothed
      // 19 more passes; a small region: nothing more
      if (pyr.empty())
        for (int i = 1; i < 20; i++) DCG_TRY(smoothBlocks<T>(top, lmin, nu, false, ctl));
    }
    for (int k = top - 1; k >= lmin; k--) {
      const uint32_t n0 = countOf(lmin, k, 0), n1 = countOf(lmin, k, 1);
      if (n0 + n1)
        DCG_LAUNCH(DCG_K_SOLVE, "k_sv_prolong",
                   (k_sv_prolong<BW, T><<<listGrid(n0 + n1), G::WG, 0, stream>>>(
                       g, k, lmin, listOf(lmin, 0), listOf(lmin, 1), n0, n1, e, reachBuf, ctl)));
      DCG_TRY(smoothBlocks<T>(k, lmin, nu, false, ctl));
    }
    return DCG_OK;
  }
  // Residual of the updated iterate and the decision on the level's next cycle.
  template <class TX, class T>
  int residualLaunch(int l, bool first, const TX* xin, TX* xout, SolveCtl* ctl) {
    const float alpha = levelH2(l);
    const double tol = P.solve_tol > 0 ? (double)P.solve_tol : 1e-8;
    const int maxCycles = P.solve_max_cycles > 0 ? P.solve_max_cycles : 50;
    const unsigned n = std::min(slotGrid<BW>(maxL[l]), kSolveWG);
    DCG_LAUNCH(DCG_K_SOLVE, "k_sv_residual",
               (k_sv_residual<BW, TX, T><<<n, G::WG, 0, stream>>>(g, l, (const TX*)xBuf[0], xin, xout,
                                                                  (const T*)eBuf, first ? 1 : 0, (T*)rhoBuf, partBuf,
                                                                  alpha, alpha, ctl)));
    DCG_LAUNCH(DCG_K_SOLVE, "k_sv_decide",
               (k_sv_decide<<<1, 256, 0, stream>>>(g, l, G::BV, first ? 1 : 0, (int)n, partBuf, alpha, tol, maxCycles,
                                                   ctl, ctlMirrorDev + l)));
    return DCG_OK;
  }
  // Iterative refinement of level l (x of the level initialised in the canonical iterate). The
  // loop runs on the device; the host enqueues as many cycles as the level took last time, then
  // one more at a time until the level's control block says it is done.
  template <class TX, class T>
  int solveLevel(int l) {
    TX* xc = (TX*)xBuf[0];
    TX* x2 = (TX*)xBuf[1];
    SolveCtl* ctl = ctlBuf + l;
    const double tol = P.solve_tol > 0 ? (double)P.solve_tol : 1e-8;
    const int maxCycles = P.solve_max_cycles > 0 ? P.solve_max_cycles : 50;
    DCG_TRY((residualLaunch<TX, T>(l, true, xc, xc, ctl)));
    int issued = 0;
    int n = predCycles[l] > 0 ? predCycles[l] : 6;
    SolveCtl h{};
    for (;;) {
      for (int i = 0; i < n && issued <= maxCycles; i++, issued++) {
        DCG_TRY(vcycle<T>(l, ctl));
        // cycle i+1 reads the iterate of cycle i and writes the other buffer
        TX* xin = issued % 2 == 0 ? xc : x2;
        TX* xout = issued % 2 == 0 ? x2 : xc;
        DCG_TRY((residualLaunch<TX, T>(l, false, xin, xout, ctl)));
      }
      DCG_TRY(sync());  // the decide kernel mirrored the control block into pinned host memory
      std::memcpy(&h, (const void*)(ctlMirror + l), sizeof(h));
      if (h.done) break;
      if (issued > maxCycles) return fail(DCG_E_HIP, "projection solve: the level's loop did not stop");
      n = 1;
    }
    predCycles[l] = (int)h.cycles;
    const TX* xf = h.cycles % 2 ? x2 : xc;
    DCG_LAUNCH(DCG_K_SOLVE, "k_sv_finalize",
               (k_sv_finalize<BW, TX><<<cellGrid(l), 256, 0, stream>>>(g, l, xf, xc, (const SolveCtl*)ctl)));
    rep.cycles[l] = (int)h.cycles;
    rep.rel_residual[l] = h.r0 > 0 ? h.rr / h.r0 : 0.0;
    rep.initial_residual[l] = h.r0;
    rep.singular[l] = (int)h.singular;
    if (!(h.rr <= tol * h.r0)) rep.converged = 0;
    return DCG_OK;
  }
  // Initial guess of every level (warm: the last projection's pressure where blocks kept their
  // place, read before divergence() clears it; cold: zero on the top level), then the levels
  // from the coarsest.
  template <class TX, class T>
  int solveAll(bool warm) {
    TX* xc = (TX*)xBuf[0];
    for (int l = L - 1; l >= 0; l--) {
      if (l == L - 1) {
        if (!warm) DCG_LAUNCH(DCG_K_SOLVE, "k_sv_zero", (k_sv_zero<BW, TX><<<cellGrid(l), 256, 0, stream>>>(g, l, xc)));
      } else {
        DCG_LAUNCH(DCG_K_SOLVE, "k_sv_prolong_x",
                   (k_sv_prolong_x<BW, TX><<<cellGrid(l), 256, 0, stream>>>(g, l, xc, warm ? 1 : 0)));
      }
      DCG_TRY((solveLevel<TX, T>(l)));
    }
    return DCG_OK;
  }
  template <class TX, class T>
  int solveWithGuess() {
    const bool warm = warmValid && P.solve_warm;
    if (warm) {
      unsigned n = blocksFor((size_t)B * G::BV, 256);
      if (n > 2048u) n = 2048u;
      DCG_LAUNCH(DCG_K_SOLVE, "k_sv_warm", (k_sv_warm<BW, TX><<<n, 256, 0, stream>>>(g, (TX*)xBuf[0])));
    }
    DCG_TRY(divergence());
    for (int l = 0; l < L; l++)
      DCG_LAUNCH(DCG_K_SOLVE, "k_sv_reach",
                 (k_sv_reach<BW><<<slotGrid<BW>(maxL[l]), G::WG, 0, stream>>>(g, l, reachBuf, breachBuf)));
    // the active blocks of every level solve's grids, and their counts on the host (which sizes
    // every launch of the solve)
    DCG_TRY(fill32(svCount, 0u, (size_t)2 * L * L));
    {
      unsigned n = blocksFor(maxL[0], 256);
      for (int l = 1; l < L; l++) n = std::max(n, blocksFor(maxL[l], 256));
      DCG_LAUNCH(DCG_K_SOLVE, "k_sv_lists", (k_sv_lists<BW><<<std::min(n, 1024u), 256, 0, stream>>>(g, breachBuf, svList, svFaces, svCount)));
      DCG_LAUNCH(DCG_K_SOLVE, "k_sv_mirror_counts",
                 (k_sv_mirror_counts<<<1, 1, 0, stream>>>(svCount, svCountMirrorDev, 2 * L * L)));
    }
    DCG_TRY(sync());
    lc.assign(svCountMirror, svCountMirror + (size_t)2 * L * L);
    return solveAll<TX, T>(warm);
  }
  int projectSolve() override {
    DCG_TRY(solveSetup());
    if (++solveSerial == 0) solveSerial = 1;  // the coarse kernel's cached setup is per solve
    svPasses = smoothSweeps();
    svInner = kSvInner;
    rep = dcg_solve_info{};
    rep.levels = L;
    rep.converged = 1;
    // the guess of a solve in one precision is the pressure field, whatever precision wrote it
    if (P.solve_precision == DCG_SOLVE_FP32)
      DCG_TRY((solveWithGuess<float, float>()));
    else if (P.solve_precision == DCG_SOLVE_FP64)
      DCG_TRY((solveWithGuess<double, double>()));
    else
      DCG_TRY((solveWithGuess<double, float>()));
    warmValid = true;
    return applyPressure();
  }
  int solveReport(dcg_solve_info* out) override {
    *out = rep;
    return DCG_OK;
  }

  // debugStats(), fluid_simulation_dcgrid.cu:517-528 (residual of the last projection)
  int debugStats(double* out) override {
    DCG_LAUNCH(DCG_K_OTHER, "k_residual", (k_residual<BW><<<blocksFor(B, 64), 64, 0, stream>>>(g, P, stats)));
    std::vector<float> h(B);
    DCG_HIP(hipMemcpyAsync(h.data(), stats, (size_t)B * sizeof(float), hipMemcpyDeviceToHost, stream));
    DCG_TRY(sync());
    float sum = 0.f;
    for (uint32_t b = 0; b < B; b++) sum = sum + h[b];
    *out = (double)sum;
    return DCG_OK;
  }
  int seedScene() override {
    DCG_LAUNCH(DCG_K_OTHER, "k_seed", (k_seed<BW><<<blocksFor(B, G::BPW), G::WG, 0, stream>>>(g, P)));
    DCG_TRY(restrictAll(g.vel));
    DCG_TRY(restrictAll(g.den));
    DCG_TRY(restrictAll(g.tem));
    if (g.vap) DCG_TRY(restrictAll(g.vap));
    return DCG_OK;
  }

  int readCtl(uint32_t* h) {
    DCG_HIP(hipMemcpyAsync(h, g.ctl, CTL_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    DCG_TRY(sync());
    if (h[CTL_ERROR] & ERR_HASH_FULL) return fail(DCG_E_HASH_FULL, "hash table fully loaded");
    return DCG_OK;
  }
  int info(dcg_info* o) override {
    uint32_t ctl[CTL_WORDS];
    DCG_TRY(readCtl(ctl));
    std::memset(o, 0, sizeof(*o));
    o->kind = DCG_KIND_DCGRID;
    o->size[0] = gx;
    o->size[1] = gy;
    o->size[2] = gz;
    o->block_width = BW;
    o->levels = L;
    o->sparse_levels = S;
    o->apron_width = G::A;
    o->max_num_blocks = B;
    o->num_cells = C;
    o->hash_table_size = H;
    uint64_t act = 0;
    for (int l = 0; l < L; l++) {
      o->max_blocks_level[l] = maxL[l];
      o->full_blocks_level[l] = fullL[l];
      o->level_offsets[l] = offs[l];
      o->block_loads[l] = ctl[CTL_LOADS + l];
      o->move_limit[l] = ctl[CTL_MOVE_LIMIT + l];
      o->hash_size_level[l] = hsize[l];
      o->hash_offset_level[l] = hoff[l];
      act += (uint64_t)ctl[CTL_LOADS + l] * G::BV;
    }
    o->active_cells = act;
    return DCG_OK;
  }
  int fieldSize(int f, uint64_t* n) override {
    if (f < 0 || f >= DCG_NUM_FIELDS) return fail(DCG_E_ARG, "unknown field");
    if (f == DCG_FIELD_VAPOR && !g.vap) return fail(DCG_E_ARG, "no vapor field (moist atmosphere off)");
    *n = (f == DCG_FIELD_VELOCITY || f == DCG_FIELD_VORTICITY) ? 3 * C : C;
    return DCG_OK;
  }
  void* fieldPtr(int f) {
    switch (f) {
      case DCG_FIELD_DENSITY: return g.den;
      case DCG_FIELD_VELOCITY: return g.vel;
      case DCG_FIELD_FLUIDITY: return g.flu;
      case DCG_FIELD_PRESSURE: return g.pre;
      case DCG_FIELD_T_PRESSURE: return g.tpre;
      case DCG_FIELD_DIVERGENCE: return g.div;
      case DCG_FIELD_TEMPERATURE: return g.tem;
      case DCG_FIELD_VORTICITY: return g.vort;
      case DCG_FIELD_VAPOR: return g.vap;
      default: return nullptr;
    }
  }
  int render(const dcg_render_params& r, float* rgba) override {
    const dim3 grid((unsigned)((r.width + 15) / 16), (unsigned)((r.height + 15) / 16));
    DCMarchSampler<BW, DGT<SF>> smp{g};
    DCG_LAUNCH(DCG_K_OTHER, "k_render", (k_render<<<grid, 256, 0, stream>>>(smp, P, r, (float4*)rgba)));
    return DCG_OK;
  }

  // State: the block tables verbatim (neighbour descriptors and hash included: the descriptors'
  // out-of-domain clamp form depends on how each block was placed), the control words and the
  // exported fields in their storage format.
  void stateSections(std::vector<StateSection>& out) override {
    const uint32_t es = (uint32_t)sizeof(SF);
    out = {{DCG_STATE_CONTROL, 4, g.ctl, (uint64_t)CTL_WORDS * 4},
           {DCG_STATE_POSITIONS, 4, g.pos, (uint64_t)B * 12},
           {DCG_STATE_LEVELS, 1, g.lev, (uint64_t)B},
           {DCG_STATE_FLAGS, 4, g.flg, (uint64_t)B * 4},
           {DCG_STATE_PARENTS, 4, g.par, (uint64_t)B * 4},
           {DCG_STATE_CHILDREN, 4, g.chi, (uint64_t)B * 32},
           {DCG_STATE_NEIGHBOURS, 4, g.nd, (uint64_t)B * 27 * 16},
           {DCG_STATE_HASH_KEYS, 4, g.hkey, (uint64_t)H * 4},
           {DCG_STATE_HASH_VALUES, 4, g.hval, (uint64_t)H * 4}};
    for (int f = 0; f < DCG_NUM_FIELDS; f++) {
      if (!fieldPtr(f)) continue;
      const uint64_t comps = (f == DCG_FIELD_VELOCITY || f == DCG_FIELD_VORTICITY) ? 3 : 1;
      out.push_back({(uint32_t)(DCG_STATE_FIELD + f), es, fieldPtr(f), comps * C * es});
    }
  }
  uint32_t stateFlags() const override { return (customFluidity ? 1u : 0u) | (warmValid ? 2u : 0u); }
  void afterStateLoad(uint32_t flags) override {
    customFluidity = (flags & 1u) != 0;
    warmValid = (flags & 2u) != 0;
    dropGraph();
  }

  // Fields are exchanged as FP32: 16-bit storage is widened exactly on export and rounded to
  // nearest even on import.
  int exportField(int f, float* dst, uint64_t n) override {
    uint64_t need;
    DCG_TRY(fieldSize(f, &need));
    if (n != need) return fail(DCG_E_ARG, "size mismatch");
    if constexpr (std::is_same<SF, float>::value) {
      DCG_HIP(hipMemcpyAsync(dst, fieldPtr(f), n * sizeof(float), hipMemcpyDeviceToHost, stream));
      return sync();
    } else {
      std::vector<uint16_t> h(n);
      DCG_HIP(hipMemcpyAsync(h.data(), fieldPtr(f), n * 2, hipMemcpyDeviceToHost, stream));
      DCG_TRY(sync());
      for (uint64_t i = 0; i < n; i++) dst[i] = StorageHost<SF>::widen(h[i]);
      return DCG_OK;
    }
  }
  int importField(int f, const float* src, uint64_t n) override {
    uint64_t need;
    DCG_TRY(fieldSize(f, &need));
    if (n != need) return fail(DCG_E_ARG, "size mismatch");
    if constexpr (std::is_same<SF, float>::value) {
      DCG_HIP(hipMemcpyAsync(fieldPtr(f), src, n * sizeof(float), hipMemcpyHostToDevice, stream));
    } else {
      std::vector<uint16_t> h(n);
      for (uint64_t i = 0; i < n; i++) h[i] = StorageHost<SF>::narrow(src[i]);
      DCG_HIP(hipMemcpyAsync(fieldPtr(f), h.data(), n * 2, hipMemcpyHostToDevice, stream));
    }
    if (f == DCG_FIELD_FLUIDITY) customFluidity = true;
    return sync();
  }

  // Host version of apronCell (dcgrid_dev.hpp) used to rebuild the reference's apron table.
  static uint32_t hloc(int x, int y, int z) {
    uint32_t r = 0;
    for (int b = 0; b < G::BB; b++)
      r |= ((uint32_t)((x >> b) & 1) << (3 * b + 2)) | ((uint32_t)((y >> b) & 1) << (3 * b + 1)) |
           ((uint32_t)((z >> b) & 1) << (3 * b));
    return r;
  }
  int exportTables(dcg_tables* t) override {
    std::vector<int32_t> pos(3 * (size_t)B);
    std::vector<uint8_t> lev(B);
    std::vector<uint32_t> flg(B), par(B), chi(8 * (size_t)B), hk(H), hv(H);
    std::vector<int32_t> nd(4 * 27 * (size_t)B);
    DCG_HIP(hipMemcpyAsync(pos.data(), g.pos, pos.size() * 4, hipMemcpyDeviceToHost, stream));
    DCG_HIP(hipMemcpyAsync(lev.data(), g.lev, B, hipMemcpyDeviceToHost, stream));
    DCG_HIP(hipMemcpyAsync(flg.data(), g.flg, (size_t)B * 4, hipMemcpyDeviceToHost, stream));
    DCG_HIP(hipMemcpyAsync(par.data(), g.par, (size_t)B * 4, hipMemcpyDeviceToHost, stream));
    DCG_HIP(hipMemcpyAsync(chi.data(), g.chi, chi.size() * 4, hipMemcpyDeviceToHost, stream));
    DCG_HIP(hipMemcpyAsync(nd.data(), g.nd, nd.size() * 4, hipMemcpyDeviceToHost, stream));
    DCG_HIP(hipMemcpyAsync(hk.data(), g.hkey, (size_t)H * 4, hipMemcpyDeviceToHost, stream));
    DCG_HIP(hipMemcpyAsync(hv.data(), g.hval, (size_t)H * 4, hipMemcpyDeviceToHost, stream));
    uint32_t ctl[CTL_WORDS];
    DCG_TRY(readCtl(ctl));
    auto u64 = [](uint32_t v) { return v == NFU ? UINT64_MAX : (uint64_t)v; };
    if (t->block_positions) std::copy(pos.begin(), pos.end(), t->block_positions);
    if (t->block_levels) std::copy(lev.begin(), lev.end(), t->block_levels);
    if (t->block_flags)
      for (uint32_t i = 0; i < B; i++) t->block_flags[i] = (uint8_t)flg[i];
    if (t->parent_indices)
      for (uint32_t i = 0; i < B; i++) t->parent_indices[i] = u64(par[i]);
    if (t->child_indices)
      for (size_t i = 0; i < chi.size(); i++) t->child_indices[i] = u64(chi[i]);
    if (t->hash_key) std::copy(hk.begin(), hk.end(), t->hash_key);
    if (t->hash_val)
      for (uint32_t i = 0; i < H; i++) t->hash_val[i] = u64(hv[i]);
    if (t->cell_indices) {
      constexpr int A = G::A, A3 = G::A3, BV = G::BV;
      for (uint32_t b = 0; b < B; b++) {
        uint64_t* e = t->cell_indices + (size_t)b * A3;
        for (int i = 0; i < A; i++)
          for (int j = 0; j < A; j++)
            for (int k = 0; k < A; k++) {
              uint64_t& out = e[(i * A + j) * A + k];
              const int dx = i == 0 ? -1 : (i == BW + 1 ? 1 : 0);
              const int dy = j == 0 ? -1 : (j == BW + 1 ? 1 : 0);
              const int dz = k == 0 ? -1 : (k == BW + 1 ? 1 : 0);
              if ((dx | dy | dz) == 0) {
                out = (uint64_t)b * BV + hloc(i - 1, j - 1, k - 1);
                continue;
              }
              if (lev[b] == FREE) {
                out = UINT64_MAX;  // k_dcgrid_init_apron_indices leaves borders notFound
                continue;
              }
              // same arithmetic as apronCell (dcgrid_dev.hpp)
              const int32_t* d = &nd[4 * ((size_t)b * 27 + (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1))];
              const uint32_t sh = (uint32_t)d[0] >> 27, nb = (uint32_t)d[0] & ND_BLOCK_MASK;
              if (sh == ND_OUT) {
                const int ci = std::clamp(i, 1, BW), cj = std::clamp(j, 1, BW), ck = std::clamp(k, 1, BW);
                out = (uint64_t)b * BV + (d[1] ? hloc(ci & (BW - 1), cj & (BW - 1), ck & (BW - 1))
                                               : hloc(ci - 1, cj - 1, ck - 1));
                continue;
              }
              out = (uint64_t)nb * BV + hloc((i - 1 - d[1]) >> sh, (j - 1 - d[2]) >> sh, (k - 1 - d[3]) >> sh);
            }
      }
    }
    return DCG_OK;
  }

  // ---------------------------------------------------------------- differentiable rollout
  // (extension). Steps on a frozen block topology: advectVelocity, the Jacobi projection,
  // advectDensity, without adaptTopology. Checkpoints of (density, temperature, velocity) every
  // K steps; the backward pass recomputes each segment from its checkpoint, keeping the
  // segment's states, and runs the transposed kernels of dcgrid_adjoint.hpp in reverse.
  float* ckBuf = nullptr;   // checkpoint states, 5 C floats each
  float* segBuf = nullptr;  // the states of the segment being reversed
  int tapeCk = 0, tapeSeg = 0, tapeN = 0, tapeK = 0;
  bool tapeValid = false;
  float *lq = nullptr, *lt = nullptr, *lq0 = nullptr, *lt0 = nullptr;  // density / temperature adjoints
  float3 *lu = nullptr, *lu0 = nullptr;                                 // velocity adjoints
  float *lp = nullptr, *ltp = nullptr, *ld = nullptr;                   // pressure pair, divergence
  // transposed advection: per-block apron halos (B (BW+2)^3 entries) and far-contribution
  // buffers (per cell, kept zero between uses) of density, temperature and velocity
  float *hq = nullptr, *ht = nullptr, *fq = nullptr, *ft = nullptr;
  float3 *hu = nullptr, *fu = nullptr;
  float* stateDen(float* base, int k) { return base + (size_t)k * 5 * C; }
  float* stateTem(float* base, int k) { return base + (size_t)k * 5 * C + C; }
  float3* stateVel(float* base, int k) { return (float3*)(base + (size_t)k * 5 * C + 2 * C); }
  int copyDD(void* dst, const void* src, size_t bytes) {
    DCG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream));
    return DCG_OK;
  }
  // Temperature changes only with a source (it is advected then); otherwise every state holds
  // the same temperature, which is neither saved nor restored.
  // Only the allocated blocks' cells are copied (the others take no part in a step).
  int copyState(const float* d0, const float* t0, const float3* v0, float* d1, float* t1, float3* v1) {
    const unsigned n = blocksFor(C, 256) < 4096 ? blocksFor(C, 256) : 4096;
    DCG_LAUNCH(DCG_K_OTHER, "k_da_copy_state",
               (k_da_copy_state<BW><<<n, 256, 0, stream>>>(g, d0, thermo() ? t0 : nullptr, v0, d1, t1, v1)));
    return DCG_OK;
  }
  int saveState(float* base, int k) {
    return copyState((const float*)g.den, (const float*)g.tem, (const float3*)g.vel, stateDen(base, k),
                     stateTem(base, k), stateVel(base, k));
  }
  int loadState(float* base, int k) {
    return copyState(stateDen(base, k), stateTem(base, k), stateVel(base, k), (float*)g.den, (float*)g.tem,
                     (float3*)g.vel);
  }
  int frozenStep() {
    DCG_TRY(advectVelocity());
    DCG_TRY(P.project_mode == DCG_PROJECT_LOCAL ? projectLocal() : project());
    return advectDensity();
  }
  unsigned cellWalkGrid(size_t cells) const {
    const unsigned n = blocksFor(cells, 256);
    return n < 4096u ? (n ? n : 1u) : 4096u;
  }
  template <class T>
  int restrictAllT(T* lam) {
    for (int l = L - 2; l >= 0; l--) {
      const size_t groups = (size_t)maxL[l] * (G::BV / 8);
      const unsigned n = blocksFor(groups, 256) < 2048 ? blocksFor(groups, 256) : 2048;
      DCG_LAUNCH(DCG_K_RESTRICT, "k_da_restrict_T", (k_da_restrict_T<BW, T><<<n, 256, 0, stream>>>(g, l, lam)));
    }
    return DCG_OK;
  }
  int tapeForward(int steps, int K) override {
    if constexpr (!std::is_same<SF, float>::value) {
      return fail(DCG_E_UNSUPPORTED, "the adjoint runs on FP32 field storage");
    } else {
      if (steps < 1 || K < 1) return fail(DCG_E_ARG, "steps and checkpoint interval must be positive");
      if (P.project_mode == DCG_PROJECT_SOLVE) return fail(DCG_E_UNSUPPORTED, "the adjoint follows the Jacobi schedules");
      if (P.moist) return fail(DCG_E_UNSUPPORTED, "the adjoint does not differentiate the moist atmosphere");
      if (K > steps) K = steps;
      const int ck = (steps + K - 1) / K + 1;  // states 0, K, 2K, ... below steps, and the final one
      if (ck != tapeCk || K + 1 != tapeSeg) {
        release(ckBuf);
        release(segBuf);
        ckBuf = segBuf = nullptr;
        tapeCk = tapeSeg = 0;
        DCG_TRY(alloc(&ckBuf, (size_t)ck * 5 * C));
        DCG_TRY(alloc(&segBuf, (size_t)(K + 1) * 5 * C));
        tapeCk = ck;
        tapeSeg = K + 1;
      }
      if (!lq) {
        DCG_TRY(alloc(&lq, C));
        DCG_TRY(alloc(&lt, C));
        DCG_TRY(alloc(&lq0, C));
        DCG_TRY(alloc(&lt0, C));
        DCG_TRY(alloc(&lu, C));
        DCG_TRY(alloc(&lu0, C));
        DCG_TRY(alloc(&lp, C));
        DCG_TRY(alloc(&ltp, C));
        DCG_TRY(alloc(&ld, C));
        const size_t halo = (size_t)B * G::A3;
        DCG_TRY(alloc(&hq, halo));
        DCG_TRY(alloc(&ht, halo));
        DCG_TRY(alloc(&hu, halo));
        DCG_TRY(alloc(&fq, C));
        DCG_TRY(alloc(&ft, C));
        DCG_TRY(alloc(&fu, C));
        DCG_TRY(fill32(fq, 0, C));
        DCG_TRY(fill32(ft, 0, C));
        DCG_TRY(fill32(fu, 0, 3 * C));
      }
      dropGraph();  // the pointer swaps of the steps below move the live fields
      tapeValid = false;
      DCG_TRY(saveState(ckBuf, 0));
      for (int k = 0; k < steps; k++) {
        DCG_TRY(frozenStep());
        if ((k + 1) % K == 0 && k + 1 < steps) DCG_TRY(saveState(ckBuf, (k + 1) / K));
      }
      DCG_TRY(saveState(ckBuf, ck - 1));
      tapeN = steps;
      tapeK = K;
      tapeValid = true;
      return sync();
    }
  }
  // Adjoint of one step between segment states s (before) and s + 1 (after).
  int adjointStep(int s) {
    const bool src = thermo(), fl = fluidityOn();
    const float* den0 = stateDen(segBuf, s);
    const float* tem0 = stateTem(segBuf, s);
    const float3* vel0 = stateVel(segBuf, s);
    const float3* vel1 = stateVel(segBuf, s + 1);
    const unsigned nAll = slotGrid<BW>(B);  // block walks over all allocated blocks
    // advectDensity: the restriction of the advected scalars, then the transposed gather
    DCG_TRY(restrictAllT(lq));
    if (src) DCG_TRY(restrictAllT(lt));
#define DA_ADV_D(F, S)                                                                         \
  DCG_LAUNCH(DCG_K_ADVECT_DENSITY, "k_da_advect_density_T",                                    \
             (k_da_advect_density_T<BW, F, S><<<nAll, G::WG, 0, stream>>>(g, P, den0, tem0, vel1, lq, lt, lq0, lt0, lu, \
                                                                          hq, ht, fq, ft)))
    if (fl && src) DA_ADV_D(true, true);
    else if (fl) DA_ADV_D(true, false);
    else if (src) DA_ADV_D(false, true);
    else DA_ADV_D(false, false);
#undef DA_ADV_D
    const unsigned nCells = blocksFor(C, 256) < 4096 ? blocksFor(C, 256) : 4096;
    DCG_LAUNCH(DCG_K_ADVECT_DENSITY, "k_da_halo_gather", (k_da_halo_gather<BW, float><<<nCells, 256, 0, stream>>>(g, lq0, hq, fq)));
    if (src)
      DCG_LAUNCH(DCG_K_ADVECT_DENSITY, "k_da_halo_gather", (k_da_halo_gather<BW, float><<<nCells, 256, 0, stream>>>(g, lt0, ht, ft)));
    // projection: restriction of the velocity, gradient, Jacobi schedule, restriction of the
    // divergence, divergence, each reversed
    DCG_TRY(restrictAllT(lu));
    DCG_TRY(fill32(lp, 0, C));
    DCG_TRY(fill32(ltp, 0, C));
    DCG_TRY(fill32(ld, 0, C));
    for (int l = 0; l < L; l++) {  // finest first: coarser cells take atomic terms
      const unsigned n = slotGrid<BW>(maxL[l]);
      if (fl)
        DCG_LAUNCH(DCG_K_APPLY_PRESSURE, "k_da_grad_T", (k_da_grad_T<BW, true><<<n, G::WG, 0, stream>>>(g, P, l, lu, lp)));
      else
        DCG_LAUNCH(DCG_K_APPLY_PRESSURE, "k_da_grad_T", (k_da_grad_T<BW, false><<<n, G::WG, 0, stream>>>(g, P, l, lu, lp)));
    }
    auto pairsT = [&](int level, int pairs) -> int {
      const int sc = 1 << level;
      const float alpha = (float)(sc * sc) * P.dx * P.dx;
      const unsigned n = slotGrid<BW>(maxL[level]);
      auto sweepT = [&](const float* lout, float* lin, int accumulate) -> int {
        if (level >= S)
          DCG_LAUNCH(DCG_K_JACOBI, "k_da_jacobi_dense_T",
                     (k_da_jacobi_dense_T<BW><<<n, G::WG, 0, stream>>>(g, level, lout, lin, ld, alpha, accumulate)));
        else
          DCG_LAUNCH(DCG_K_JACOBI, "k_da_jacobi_T",
                     (k_da_jacobi_T<BW><<<n, G::WG, 0, stream>>>(g, level, lout, lin, ld, alpha, accumulate)));
        return DCG_OK;
      };
      for (int i = 0; i < pairs; i++) {
        // jacobi(level, t_pre, pre) then jacobi(level, pre, t_pre), reversed; the level's first
        // launch adds to what lin holds, the others store
        DCG_TRY(sweepT(lp, ltp, i == 0 ? 1 : 0));
        DCG_TRY(sweepT(ltp, lp, 0));
      }
      return DCG_OK;
    };
    if (P.project_mode == DCG_PROJECT_LOCAL) {
      for (int l = 0; l < L; l++) DCG_TRY(pairsT(l, P.project_local_pairs));
    } else {
      for (int l = 0; l <= L - 2; l++) {
        DCG_TRY(pairsT(l, P.project_level_pairs));
        DCG_LAUNCH(DCG_K_PROLONGATE, "k_da_prolongate_T",
                   (k_da_prolongate_T<BW><<<slotGrid<BW>(maxL[l]), G::WG, 0, stream>>>(g, l, lp)));
      }
      DCG_TRY(pairsT(L - 1, P.project_top_pairs));
    }
    DCG_TRY(restrictAllT(ld));
    for (int l = 0; l < L; l++) {  // finest first: coarser cells take atomic terms
      const unsigned n = slotGrid<BW>(maxL[l]);
      if (fl)
        DCG_LAUNCH(DCG_K_DIVERGENCE, "k_da_div_T", (k_da_div_T<BW, true><<<n, G::WG, 0, stream>>>(g, P, l, ld, lu)));
      else
        DCG_LAUNCH(DCG_K_DIVERGENCE, "k_da_div_T", (k_da_div_T<BW, false><<<n, G::WG, 0, stream>>>(g, P, l, ld, lu)));
    }
    // advectVelocity: the restriction of the advected velocity, then the transposed gather
    DCG_TRY(restrictAllT(lu));
#define DA_ADV_V(F, Bu)                                                                        \
  DCG_LAUNCH(DCG_K_ADVECT_VELOCITY, "k_da_advect_velocity_T",                                  \
             (k_da_advect_velocity_T<BW, F, Bu><<<nAll, G::WG, 0, stream>>>(g, P, vel0, lu, lu0, lq0, lt0, hu, fu)))
    if (fl && src) DA_ADV_V(true, true);
    else if (fl) DA_ADV_V(true, false);
    else if (src) DA_ADV_V(false, true);
    else DA_ADV_V(false, false);
#undef DA_ADV_V
    DCG_LAUNCH(DCG_K_ADVECT_VELOCITY, "k_da_halo_gather", (k_da_halo_gather<BW, float3><<<nCells, 256, 0, stream>>>(g, lu0, hu, fu)));
    std::swap(lq, lq0);
    std::swap(lu, lu0);
    if (src) std::swap(lt, lt0);  // without a source the temperature adjoint passes through
    return DCG_OK;
  }
  int tapeBackward(const float* gden, const float* gvel, const float* gtem, float* oden, float* ovel,
                   float* otem) override {
    if constexpr (!std::is_same<SF, float>::value) {
      return fail(DCG_E_UNSUPPORTED, "the adjoint runs on FP32 field storage");
    } else {
      if (!tapeValid) return fail(DCG_E_ARG, "no recorded rollout (dcg_tape_forward)");
      DCG_TRY(copyDD(lq, gden, C * 4));
      DCG_TRY(copyDD(lu, gvel, C * 12));
      if (gtem)
        DCG_TRY(copyDD(lt, gtem, C * 4));
      else
        DCG_TRY(fill32(lt, 0, C));
      const int segs = (tapeN + tapeK - 1) / tapeK;
      for (int sg = segs - 1; sg >= 0; sg--) {
        const int k0 = sg * tapeK, k1 = std::min(tapeN, k0 + tapeK);
        DCG_TRY(loadState(ckBuf, sg));
        DCG_TRY(saveState(segBuf, 0));
        for (int k = k0; k < k1; k++) {
          if (k + 1 == k1)  // the segment's last state is the next checkpoint: no step to re-run
            DCG_TRY(copyState(stateDen(ckBuf, sg + 1), stateTem(ckBuf, sg + 1), stateVel(ckBuf, sg + 1),
                              stateDen(segBuf, k - k0 + 1), stateTem(segBuf, k - k0 + 1), stateVel(segBuf, k - k0 + 1)));
          else {
            DCG_TRY(frozenStep());
            DCG_TRY(saveState(segBuf, k - k0 + 1));
          }
        }
        for (int k = k1 - 1; k >= k0; k--) DCG_TRY(adjointStep(k - k0));
      }
      DCG_TRY(copyDD(oden, lq, C * 4));
      DCG_TRY(copyDD(ovel, lu, C * 12));
      if (otem) DCG_TRY(copyDD(otem, lt, C * 4));
      DCG_LAUNCH(DCG_K_OTHER, "k_da_mask_free",
                 (k_da_mask_free<BW><<<cellWalkGrid(C), 256, 0, stream>>>(g, oden, (float3*)ovel, otem, C)));
      DCG_TRY(loadState(ckBuf, tapeCk - 1));  // the live state is the rollout's final state again
      return sync();
    }
  }
  // Device-memory field copies (FP32 storage), for the torch binding of the rollout.
  int exportFieldDevice(int f, float* dst, uint64_t n) override {
    if constexpr (!std::is_same<SF, float>::value) {
      return fail(DCG_E_UNSUPPORTED, "device field copies need FP32 field storage");
    } else {
      uint64_t need;
      DCG_TRY(fieldSize(f, &need));
      if (n != need || !dst) return fail(DCG_E_ARG, "size mismatch");
      DCG_TRY(copyDD(dst, fieldPtr(f), n * 4));
      return sync();
    }
  }
  int importFieldDevice(int f, const float* src, uint64_t n) override {
    if constexpr (!std::is_same<SF, float>::value) {
      return fail(DCG_E_UNSUPPORTED, "device field copies need FP32 field storage");
    } else {
      uint64_t need;
      DCG_TRY(fieldSize(f, &need));
      if (n != need || !src) return fail(DCG_E_ARG, "size mismatch");
      DCG_TRY(copyDD(fieldPtr(f), src, n * 4));
      if (f == DCG_FIELD_FLUIDITY) customFluidity = true;
      return sync();
    }
  }
};

template <int BW, class SF>
static SimBase* makeSim(const int32_t size[3], uint64_t maxBlocks, const dcg_params& p, int* st) {
  auto* d = new DCGridSim<BW, SF>();
  *st = d->create(size, maxBlocks, p);
  return d;
}
template <int BW>
static SimBase* makeBW(const int32_t size[3], uint64_t maxBlocks, const dcg_params& p, int* st) {
  if (p.storage == DCG_STORAGE_FP16) return makeSim<BW, StoreF16>(size, maxBlocks, p, st);
  if (p.storage == DCG_STORAGE_BF16) return makeSim<BW, StoreBF16>(size, maxBlocks, p, st);
  return makeSim<BW, float>(size, maxBlocks, p, st);
}

SimBase* makeDCGrid(const int32_t size[3], uint64_t maxBlocks, const dcg_params& p, int* status, std::string* err) {
  if (p.block_width != 4 && p.block_width != 8) {
    *status = DCG_E_UNSUPPORTED;
    *err = "block_width must be 4 or 8";
    return nullptr;
  }
  if (p.storage != DCG_STORAGE_FP32 && p.storage != DCG_STORAGE_FP16 && p.storage != DCG_STORAGE_BF16) {
    *status = DCG_E_ARG;
    *err = "unknown field storage format";
    return nullptr;
  }
  int st = DCG_OK;
  SimBase* s = p.block_width == 4 ? makeBW<4>(size, maxBlocks, p, &st) : makeBW<8>(size, maxBlocks, p, &st);
  *status = st;
  if (st != DCG_OK) {
    *err = s->err;
    delete s;
    return nullptr;
  }
  return s;
}

}  // namespace dcg

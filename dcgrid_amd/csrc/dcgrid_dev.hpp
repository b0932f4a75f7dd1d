// dcgrid_dev.hpp — device-side data structure and lookups of the DCGrid back-end.
//
// Layout (reference semantics: src/dcgrid/dcgrid.h:5-73, generalised to BW in {4,8}):
//  * fields are indexed by global cell = block*BW^3 + Morton(local); velocity is float3 AoS;
//  * block tables: positions (level-local cell origin), level (0xFF = free), flags, parent
//    subblock (8*parentBlock + top Morton triple), 8 child blocks per block;
//  * instead of the reference's (BW+2)^3 size_t apron table (27 B/cell at BW=4) every block keeps
//    27 neighbour descriptors of 16 B (6.75 B/cell at BW=4, 0.84 B/cell at BW=8). All entries of
//    one apron direction of the reference table point into a single block (the finest block at
//    the same or a coarser level containing that region, or the block itself outside the
//    domain), so each entry is that block's cell at a fixed, shifted offset: the descriptor holds
//    the block, the level shift and the origin offset, and the reference table is a pure function
//    of the descriptors (dcg_export_tables rebuilds it).
#pragma once

#include "dcg_device.hpp"

namespace dcg {

template <int BW>
struct Geo {
  static constexpr int BB = BW == 4 ? 2 : 3;    // log2(BW)
  static constexpr int BV = BW * BW * BW;       // cells per block
  static constexpr int A = BW + 2;              // apron width
  static constexpr int A3 = A * A * A;
  static constexpr int SBB = 3 * (BB - 1);      // log2(cells per subblock)
  static constexpr int BPW = BW == 4 ? 4 : 1;   // blocks per workgroup
  static constexpr int WG = BPW * BV;           // workgroup size
  static constexpr int FACE = 6 * BW * BW;      // face cells of a block's apron
};

// Storage formats of the fields in HBM (dcg_params.storage). A field element converts to FP32
// when read and rounds to nearest even when written (v_cvt_f16_f32 / v_cvt_pk_bf16_f32); all
// arithmetic is FP32, so a kernel computes exactly what it computes on FP32 fields from the
// widened values and only its stores differ.
//
// The value is first pinned as an FP32 register value: otherwise the backend folds a product
// feeding the conversion into v_fma_mixlo_f16 (the exact product rounded once to FP16), which
// differs from rounding the FP32 product (a value the FP32 computation defines) to FP16.
__device__ __forceinline__ float pinF32(float f) {
  asm("" : "+v"(f));
  return f;
}
struct StoreF16 {
  uint16_t bits;
  StoreF16() = default;
  __device__ __forceinline__ StoreF16(float f) : bits(__builtin_bit_cast(uint16_t, (_Float16)pinF32(f))) {}
  __device__ __forceinline__ operator float() const { return (float)__builtin_bit_cast(_Float16, bits); }
};
struct StoreBF16 {
  uint16_t bits;
  StoreBF16() = default;
  __device__ __forceinline__ StoreBF16(float f) : bits(__builtin_bit_cast(uint16_t, (__bf16)pinF32(f))) {}
  __device__ __forceinline__ operator float() const { return __uint_as_float((uint32_t)bits << 16); }
};
// Velocity-like fields: three stored components, 6 bytes per cell in the 16-bit formats.
template <class S>
struct StoreV3 {
  S x, y, z;
  StoreV3() = default;
  __device__ __forceinline__ StoreV3(float3 v) : x(v.x), y(v.y), z(v.z) {}
  __device__ __forceinline__ operator float3() const { return make_float3(x, y, z); }
};
template <class S>
struct StoreTraits {
  using V = StoreV3<S>;
};
template <>
struct StoreTraits<float> {
  using V = float3;
};

// Tables of the block grid (everything but the fields).
struct DGBase {
  int gx, gy, gz, L, S;
  uint32_t B;
  uint32_t offs[DCG_MAX_LEVELS], maxL[DCG_MAX_LEVELS], fullL[DCG_MAX_LEVELS];
  int rx[DCG_MAX_LEVELS], ry[DCG_MAX_LEVELS], rz[DCG_MAX_LEVELS];
  uint32_t hoff[DCG_MAX_LEVELS], hsize[DCG_MAX_LEVELS];
  int3* pos;
  uint8_t* lev;
  uint32_t* flg;
  uint32_t* par;
  uint32_t* chi;
  int4* nd;  // 27 neighbour descriptors per block, direction (dx+1)*9+(dy+1)*3+(dz+1)
  uint32_t* hkey;
  uint32_t* hval;
  uint32_t* ctl;  // device control words, see CTL_* below
};

// The grid with its fields stored as SF (float, StoreF16 or StoreBF16).
template <class SF>
struct DGT : DGBase {
  using Scalar = SF;
  using Vec = typename StoreTraits<SF>::V;
  SF* den;
  SF* tem;
  SF* flu;
  SF* div;
  SF* pre;
  SF* tpre;
  SF* tden;
  SF* ttem;
  SF* vap;  // water vapor (moist atmosphere; allocated when enabled at creation)
  SF* tvap;
  Vec* vel;
  Vec* tvel;
  Vec* vort;
};
using DG = DGT<float>;

// Control words (uint32 each) kept on the device so the step never waits for the host.
enum : int {
  CTL_LOADS = 0,                          // [16] blocks allocated per level (blockLoads)
  CTL_MOVE_LIMIT = CTL_LOADS + 16,        // [16] moveLimit
  CTL_LIM = CTL_MOVE_LIMIT + 16,          // [16] move limit of the current adaptation
  CTL_MOVES = CTL_LIM + 16,               // [16] matched moves per level
  CTL_REFINES = CTL_MOVES + 16,           // [16] refinements of subblocks of level l (children at l-1)
  CTL_LOADS_BEFORE = CTL_REFINES + 16,    // [16] loads before refinement
  CTL_TOTAL_MOVES = CTL_LOADS_BEFORE + 16,  // moves of this adaptation, all levels
  CTL_ERROR = CTL_TOTAL_MOVES + 1,        // sticky error bits
  CTL_WORDS = CTL_ERROR + 1
};
enum : uint32_t { ERR_HASH_FULL = 1, ERR_KEY_EXISTS = 2 };

// Morton key of a block position, gridHash (src/dcgrid/dcgrid_utils.cuh:83-96).
__device__ __forceinline__ uint32_t bitSpread3(uint32_t data) {
  uint32_t res = 0;
  for (uint32_t mask = 1u; mask; mask <<= 3, data <<= 2) res |= data & mask;
  return res;
}
template <int BW>
__device__ __forceinline__ uint32_t gridHash(int x, int y, int z) {
  return (bitSpread3((uint32_t)(x / BW)) << 2) | (bitSpread3((uint32_t)(y / BW)) << 1) | bitSpread3((uint32_t)(z / BW));
}

__device__ __forceinline__ int dirIndex(int dx, int dy, int dz) { return (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1); }

// Finest allocated block at level >= l0 whose region contains the level-l0 cell q (q inside
// the domain). Equivalent to getBlockIndexDeep (dcgrid_utils.cuh:201-232), which walks the
// per-level hash tables upward from l0; here the walk goes down the child links from the first
// dense level, which reach exactly the allocated blocks.
template <int BW>
__device__ __forceinline__ uint32_t findBlock(const DGBase& g, int qx, int qy, int qz, int l0) {
  constexpr int BB = Geo<BW>::BB;
  int lv = l0 > g.S ? l0 : g.S;
  int sh = lv - l0;
  const int bx = (qx >> sh) >> BB, by = (qy >> sh) >> BB, bz = (qz >> sh) >> BB;
  uint32_t b = g.offs[lv] + (uint32_t)((bx * g.ry[lv] + by) * g.rz[lv] + bz);
  while (lv > l0) {
    sh = lv - l0;
    const int t = ((((qx >> sh) >> (BB - 1)) & 1) << 2) | ((((qy >> sh) >> (BB - 1)) & 1) << 1) |
                  (((qz >> sh) >> (BB - 1)) & 1);
    const uint32_t c = g.chi[8 * b + t];
    if (c == NFU) break;
    b = c;
    lv--;
  }
  return b;
}

// Neighbour descriptor of one apron direction of a block: x = neighbour block | shift << 27,
// (y,z,w) = origin offset o = npos * 2^shift - p0 (block origins in level-local cells, shift =
// level(neighbour) - level(block)). The cell of apron entry (ai,aj,ak) in that direction is
// neighbour*BW^3 + Morton(((a-1) - o) >> shift), one load and integer arithmetic.
// Out-of-domain directions have shift ND_OUT and y = clamp rule of the block (see apronCell).
static constexpr uint32_t ND_OUT = 31;
static constexpr uint32_t ND_BLOCK_MASK = 0x07FFFFFFu;  // blocks are < 2^27
__device__ __forceinline__ uint32_t ndBlock(int4 d) { return (uint32_t)d.x & ND_BLOCK_MASK; }
__device__ __forceinline__ uint32_t ndShift(int4 d) { return (uint32_t)d.x >> 27; }
__device__ __forceinline__ int4 makeNd(uint32_t nb, uint32_t shift, int ox, int oy, int oz) {
  return make_int4((int)(nb | (shift << 27)), ox, oy, oz);
}

// Global cell index of apron entry (ai,aj,ak) in [0,BW+1]^3 of block b: the value the
// reference's cellIndices table holds for that entry.
template <int BW>
__device__ __forceinline__ uint32_t apronCell(const DGBase& g, uint32_t b, int ai, int aj, int ak) {
  constexpr int BB = Geo<BW>::BB, BV = Geo<BW>::BV;
  const int dx = ai == 0 ? -1 : (ai == BW + 1 ? 1 : 0);
  const int dy = aj == 0 ? -1 : (aj == BW + 1 ? 1 : 0);
  const int dz = ak == 0 ? -1 : (ak == BW + 1 ? 1 : 0);
  if ((dx | dy | dz) == 0) return b * BV + loc<BB>(ai - 1, aj - 1, ak - 1);
  const int4 d = g.nd[b * 27 + dirIndex(dx, dy, dz)];
  const uint32_t sh = ndShift(d);
  if (sh == ND_OUT) {
    // Out-of-domain entry: own clamped cell. Blocks placed by a move or a refinement use the
    // reference's SPREAD(clamp(i,1,BW)) form (dcgrid_structure.cu:54-60), dense blocks the
    // clamped interior cell (dcgrid_structure.cu:157-167).
    const int ci = clampi(ai, 1, BW), cj = clampi(aj, 1, BW), ck = clampi(ak, 1, BW);
    if (d.y) return b * BV + loc<BB>(ci & (BW - 1), cj & (BW - 1), ck & (BW - 1));
    return b * BV + loc<BB>(ci - 1, cj - 1, ck - 1);
  }
  return ndBlock(d) * BV + loc<BB>((ai - 1 - d.y) >> sh, (aj - 1 - d.z) >> sh, (ak - 1 - d.w) >> sh);
}

// apronCell for a border entry (at least one of ai, aj, ak is 0 or BW+1), without branches: the
// descriptor is read with one load and both forms are computed and selected, so a caller can
// issue this load together with its others (same cell as apronCell).
template <int BW>
__device__ __forceinline__ uint32_t borderApronCell(const DGBase& g, uint32_t b, int ai, int aj, int ak) {
  constexpr int BB = Geo<BW>::BB, BV = Geo<BW>::BV;
  const int dx = ai == 0 ? -1 : (ai == BW + 1 ? 1 : 0);
  const int dy = aj == 0 ? -1 : (aj == BW + 1 ? 1 : 0);
  const int dz = ak == 0 ? -1 : (ak == BW + 1 ? 1 : 0);
  const int4 d = g.nd[b * 27 + dirIndex(dx, dy, dz)];
  const uint32_t sh = ndShift(d);
  const int ci = clampi(ai, 1, BW), cj = clampi(aj, 1, BW), ck = clampi(ak, 1, BW);
  const uint32_t spread = loc<BB>(ci & (BW - 1), cj & (BW - 1), ck & (BW - 1));
  const uint32_t clamped = loc<BB>(ci - 1, cj - 1, ck - 1);
  const uint32_t outCell = b * BV + (d.y ? spread : clamped);
  const uint32_t s = sh == ND_OUT ? 0u : sh;
  const uint32_t inCell = ndBlock(d) * BV + loc<BB>((ai - 1 - d.y) >> s, (aj - 1 - d.z) >> s, (ak - 1 - d.w) >> s);
  return sh == ND_OUT ? outCell : inCell;
}

// Apron cell of a block on a dense level (l >= S). Every block of such a level is allocated at
// the dense index (bx*ry + by)*rz + bz and never moves, so its neighbour descriptors are fixed:
// the same-level neighbour at the dense index, or the clamped own cell outside the domain (the
// form k_activate_dense writes). The cell follows from the block origin p0 alone, without the
// dependent descriptor load of apronCell (same cell).
template <int BW>
__device__ __forceinline__ uint32_t denseApronCell(const DGBase& g, int l, uint32_t b, int3 p0, int ai, int aj, int ak) {
  constexpr int BB = Geo<BW>::BB, BV = Geo<BW>::BV;
  const int dx = ai == 0 ? -1 : (ai == BW + 1 ? 1 : 0);
  const int dy = aj == 0 ? -1 : (aj == BW + 1 ? 1 : 0);
  const int dz = ak == 0 ? -1 : (ak == BW + 1 ? 1 : 0);
  const int nx = (p0.x >> BB) + dx, ny = (p0.y >> BB) + dy, nz = (p0.z >> BB) + dz;
  if (nx < 0 || ny < 0 || nz < 0 || nx >= g.rx[l] || ny >= g.ry[l] || nz >= g.rz[l])
    return b * BV + loc<BB>(clampi(ai, 1, BW) - 1, clampi(aj, 1, BW) - 1, clampi(ak, 1, BW) - 1);
  const uint32_t nb = g.offs[l] + (uint32_t)((nx * g.ry[l] + ny) * g.rz[l] + nz);
  return nb * BV + loc<BB>(ai - 1 - BW * dx, aj - 1 - BW * dy, ak - 1 - BW * dz);
}
// Origin of block b of dense level l: k_activate_dense places block offs[l] + k at
// BW * (k / (ry rz), (k / rz) mod ry, k mod rz), and dense blocks never move, so the origin
// needs no table read.
template <int BW>
__device__ __forceinline__ int3 denseOrigin(const DGBase& g, int l, uint32_t b) {
  const uint32_t k = b - g.offs[l], rz = (uint32_t)g.rz[l], ry = (uint32_t)g.ry[l];
  const uint32_t q = k / rz;
  return make_int3(BW * (int)(q / ry), BW * (int)(q % ry), BW * (int)(k - q * rz));
}
// Apron coordinates of face cell f in [0, 6*BW*BW): faces x-,x+,y-,y+,z-,z+.
// Written with selects, not a switch: no branch between the loads a caller issues around it.
template <int BW>
__device__ __forceinline__ void faceCell(int f, int& ai, int& aj, int& ak) {
  const int face = f / (BW * BW), uv = f % (BW * BW), u = 1 + uv / BW, v = 1 + uv % BW;
  const int axis = face >> 1, side = (face & 1) ? BW + 1 : 0;
  ai = axis == 0 ? side : u;
  aj = axis == 0 ? u : (axis == 1 ? side : v);
  ak = axis == 2 ? side : v;
}

}  // namespace dcg

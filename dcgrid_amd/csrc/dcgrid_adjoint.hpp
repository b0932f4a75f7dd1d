// dcgrid_adjoint.hpp — gfx950 adjoint (reverse-mode) kernels of the DCGrid step on a frozen
// block topology: advectVelocity, the Jacobi projection (project / projectLocal), advectDensity,
// with the block tables, neighbour descriptors and fluidity fixed over the rollout. Extension:
// the reference has no derivatives. Included by dcgrid.hip after the forward kernels, whose
// sampling (CrossSample, footPoint), boundary conditions and parent stencil they reuse.
//
// Every forward kernel `out = f(in)` has a transposed kernel `lam_in += (df/din)^T lam_out`
// (then `lam_out = 0` where the output overwrote its previous value), run in reverse order:
//  * the cross-level gather of the semi-Lagrangian advection becomes a scatter of w_c lam to the
//    eight corner cells of each foot (corners on any level, found through the same descriptors),
//    plus the foot term d(sample)/d(u) lam local to the cell;
//  * the Jacobi half-sweeps become scatters of lam/6 to the six apron cells (same-level cells and
//    coarser ghost cells alike; the reference's out-of-domain apron entries are not always the
//    cell itself, so the stencil is not assumed symmetric) and -alpha lam/6 into the divergence;
//  * prolongation scatters (27, 9, 3, 1)/64 lam to the parent stencil, restriction hands a parent
//    cell's lam/8 to each of its children (a gather: no atomics), divergence and gradient scatter
//    to their apron neighbours.
// Scatters use hardware FP32 atomics (global_atomic_add_f32), so sums are formed in an order that
// can change from run to run: the gradient is reproducible to FP32 rounding, not bitwise.
// One thread per cell over the allocated-block walks of the forward kernels.
#pragma once

namespace dcg {

__device__ __forceinline__ void atomicAddF(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomicAddF3(float3* p, float3 v) {
  float* q = reinterpret_cast<float*>(p);
  unsafeAtomicAdd(q, v.x);
  unsafeAtomicAdd(q + 1, v.y);
  unsafeAtomicAdd(q + 2, v.z);
}

// The cells of the allocated blocks of one level (level >= 0) or of all levels, one per thread
// of a grid-stride walk; op(block, its level, local cell).
template <int BW, class Op>
__device__ __forceinline__ void forAllocatedCells(const DGBase& g, int level, Op op) {
  constexpr int BV = Geo<BW>::BV;
  const uint32_t total = slotTotal(g, level) * BV;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    int bl;
    const uint32_t b = slotBlock(g, level, i / BV, bl);
    op(b, bl, (int)(i % BV));
  }
}

// Foot sample with what its derivative needs: the corners and normalised weights of
// CrossSample, the unnormalised weights f_c t_c (sum W), d t_c / d frac_a, and
// k = d frac / d u = -dt rdx / scale of the sampled block.
template <int BW, bool FLUIDITY>
struct FootD {
  CrossSample<BW, FLUIDITY> s;
  float f[8], t[8], dtd[8][3], W, k;
  __device__ __forceinline__ void init(const DG& g, const dcg_params& P, float px, float py, float pz) {
    s.init(g, px, py, pz);
    const float si = 1.f / (float)s.scale;
    const float xs = px * si - .5f, ys = py * si - .5f, zs = pz * si - .5f;
    const float d[3] = {xs - floorf(xs), ys - floorf(ys), zs - floorf(zs)};
    W = 0.f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int bit[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
      float tau[3];
#pragma unroll
      for (int a = 0; a < 3; a++) tau[a] = bit[a] ? d[a] : 1.f - d[a];
      t[c] = tau[0] * tau[1] * tau[2];
      dtd[c][0] = (bit[0] ? 1.f : -1.f) * tau[1] * tau[2];
      dtd[c][1] = (bit[1] ? 1.f : -1.f) * tau[0] * tau[2];
      dtd[c][2] = (bit[2] ? 1.f : -1.f) * tau[0] * tau[1];
      f[c] = FLUIDITY ? (float)g.flu[s.i[c]] : 1.f;
      W += f[c] * t[c];
    }
    k = -(P.dt * P.rdx) * si;
  }
  // d(sample)/d(u) of the normalised interpolant of the BC-applied corner values vc.
  __device__ __forceinline__ float3 grad(const float vc[8]) const {
    float val = 0.f;
#pragma unroll
    for (int c = 0; c < 8; c++) val += f[c] * t[c] * vc[c];
    val /= W;
    float out[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      float num = 0.f, den = 0.f;
#pragma unroll
      for (int c = 0; c < 8; c++) {
        num += f[c] * vc[c] * dtd[c][a];
        den += f[c] * dtd[c][a];
      }
      out[a] = k * (num - val * den) / W;
    }
    return make_float3(out[0], out[1], out[2]);
  }
  __device__ __forceinline__ void corner(int c, int& x, int& y, int& z) const {
    x = s.x0 + ((c >> 2) & 1);
    y = s.y0 + ((c >> 1) & 1);
    z = s.z0 + (c & 1);
  }
};

// restrictAll reversed for one level (launched from the coarsest restriction down): each parent
// cell's adjoint is shared equally by its 8 children, then cleared. The groups of k_restrict.
template <int BW, class T>
__global__ void __launch_bounds__(256) k_da_restrict_T(DGBase g, int level, T* __restrict__ lam) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB, GPB = BV / 8;
  const uint32_t groups = g.ctl[CTL_LOADS + level] * GPB;
  for (uint32_t gi = blockIdx.x * 256 + threadIdx.x; gi < groups; gi += gridDim.x * 256) {
    const uint32_t b = g.offs[level] + gi / GPB;
    const uint32_t ps = g.par[b];
    if (ps == NFU) continue;
    const uint32_t grp = gi % GPB;
    const uint32_t parentCell = (ps / 8) * BV + ((ps % 8) << SBB) + grp;
    const T G = mulv(lam[parentCell], .125f);
    T* dst = lam + (size_t)b * BV + 8 * grp;
#pragma unroll
    for (int i = 0; i < 8; i++) dst[i] = addv(dst[i], G);
    lam[parentCell] = zeroOf<T>();
  }
}

// advectDensity reversed (leaf cells): lq, lt (after) scatter into lq0, lt0 (before, cleared by
// the caller); the foot terms of both go into lu, the adjoint of the advecting velocity. Cells
// inside the sphere source were overwritten by its constants and pass nothing on.
template <int BW, bool FLUIDITY, bool SRC>
__global__ void __launch_bounds__(256) k_da_advect_density_T(DG g, dcg_params P, const float* __restrict__ den0,
                                                             const float* __restrict__ tem0,
                                                             const float3* __restrict__ vel1,
                                                             const float* __restrict__ lq, const float* __restrict__ lt,
                                                             float* __restrict__ lq0, float* __restrict__ lt0,
                                                             float3* __restrict__ lu) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB;
  const bool sphere = P.source_mode == DCG_SOURCE_SPHERE_BUOYANT;
  forAllocatedCells<BW>(g, -1, [&](uint32_t b, int level, int t) {
    if (g.chi[8 * b + (t >> SBB)] != NFU) return;  // not a leaf cell
    const uint32_t cell = b * BV + t;
    const float gq = lq[cell], gt = SRC ? lt[cell] : 0.f;
    if (gq == 0.f && gt == 0.f) return;
    float px, py, pz;
    int x, y, z;
    footPoint<BW>(P, g.pos[b], level, t, vel1[cell], px, py, pz, x, y, z);
    if (SRC && sphere) {
      const float sc = (float)(1 << level);
      if (inSource(P, ((float)x + .5f) * sc, ((float)y + .5f) * sc, ((float)z + .5f) * sc)) return;
    }
    FootD<BW, FLUIDITY> F;
    F.init(g, P, px, py, pz);
    if (F.s.zero) return;
    float qc[8], tc[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      int cx, cy, cz;
      F.corner(c, cx, cy, cz);
      const uint32_t i = F.s.i[c];
      qc[c] = densityBC(P, den0[i], cx, cy, cz, F.s.scale);
      if (SRC) tc[c] = temperatureBC(P, tem0[i], cx, cy, cz, F.s.scale);
      if (!outside(P, cx, cy, cz, F.s.scale)) {  // the BC passes the stored value through
        atomicAddF(&lq0[i], F.s.w[c] * gq);
        if (SRC) atomicAddF(&lt0[i], F.s.w[c] * gt);
      }
    }
    float3 u = lu[cell];
    u = f3add(u, f3mul(F.grad(qc), gq));
    if (SRC) u = f3add(u, f3mul(F.grad(tc), gt));
    lu[cell] = u;
  });
}

// advectVelocity reversed (leaf cells): lu (after) scatters into lu0 (before, cleared by the
// caller), with the foot terms of the three components at the cell itself; the buoyancy term
// adds to lq0, lt0 (the density and temperature adjoints of the step's start).
template <int BW, bool FLUIDITY, bool BUOY>
__global__ void __launch_bounds__(256) k_da_advect_velocity_T(DG g, dcg_params P, const float3* __restrict__ vel0,
                                                              const float3* __restrict__ lu,
                                                              float3* __restrict__ lu0, float* __restrict__ lq0,
                                                              float* __restrict__ lt0) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB;
  forAllocatedCells<BW>(g, -1, [&](uint32_t b, int level, int t) {
    if (g.chi[8 * b + (t >> SBB)] != NFU) return;
    const uint32_t cell = b * BV + t;
    const float3 lam = lu[cell];
    if (BUOY) {
      lq0[cell] += -P.dt * P.buoy_alpha * lam.y;
      lt0[cell] += P.dt * P.buoy_beta * lam.y;
    }
    if (lam.x == 0.f && lam.y == 0.f && lam.z == 0.f) return;
    float px, py, pz;
    int x, y, z;
    footPoint<BW>(P, g.pos[b], level, t, vel0[cell], px, py, pz, x, y, z);
    FootD<BW, FLUIDITY> F;
    F.init(g, P, px, py, pz);
    if (F.s.zero) return;
    float vc[3][8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      int cx, cy, cz;
      F.corner(c, cx, cy, cz);
      const uint32_t i = F.s.i[c];
      const float3 v = velocityBC(P, vel0[i], cx, cy, cz, F.s.scale);
      vc[0][c] = v.x;
      vc[1][c] = v.y;
      vc[2][c] = v.z;
      if (!outside(P, cx, cy, cz, F.s.scale)) atomicAddF3(&lu0[i], f3mul(lam, F.s.w[c]));
    }
    float3 self = f3mul(F.grad(vc[0]), lam.x);
    self = f3add(self, f3mul(F.grad(vc[1]), lam.y));
    self = f3add(self, f3mul(F.grad(vc[2]), lam.z));
    atomicAddF3(&lu0[cell], self);
  });
}

// Apron cells of the six face neighbours of local cell t of block b: -x, +x, -y, +y, -z, +z.
template <int BW>
__device__ __forceinline__ void faceNeighbours(const DGBase& g, uint32_t b, int t, uint32_t nb[6]) {
  int lx, ly, lz;
  locPos<Geo<BW>::BB>((uint32_t)t, lx, ly, lz);
  const int ai = lx + 1, aj = ly + 1, ak = lz + 1;
  nb[0] = apronCell<BW>(g, b, ai - 1, aj, ak);
  nb[1] = apronCell<BW>(g, b, ai + 1, aj, ak);
  nb[2] = apronCell<BW>(g, b, ai, aj - 1, ak);
  nb[3] = apronCell<BW>(g, b, ai, aj + 1, ak);
  nb[4] = apronCell<BW>(g, b, ai, aj, ak - 1);
  nb[5] = apronCell<BW>(g, b, ai, aj, ak + 1);
}

// applyPressure reversed (leaf cells): v -= alpha (f_r (p_r - p) + f_l (p - p_l)) per axis, so
// lp[r] -= alpha f_r lam, lp[l] += alpha f_l lam, lp[c] += alpha (f_r - f_l) lam; v passes through.
template <int BW, bool FLUIDITY>
__global__ void __launch_bounds__(256) k_da_grad_T(DG g, dcg_params P, const float3* __restrict__ lu,
                                                   float* __restrict__ lp) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB;
  forAllocatedCells<BW>(g, -1, [&](uint32_t b, int level, int t) {
    if (g.chi[8 * b + (t >> SBB)] != NFU) return;
    const uint32_t cell = b * BV + t;
    const float3 lam3 = lu[cell];
    if (lam3.x == 0.f && lam3.y == 0.f && lam3.z == 0.f) return;
    const float alpha = .5f * P.rdx / (float)(1 << level);
    uint32_t nb[6];
    faceNeighbours<BW>(g, b, t, nb);
    const float lam[3] = {lam3.x, lam3.y, lam3.z};
    float self = 0.f;
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
      const uint32_t il = nb[2 * ax], ir = nb[2 * ax + 1];
      const float fl = FLUIDITY ? (float)g.flu[il] : 1.f, fr = FLUIDITY ? (float)g.flu[ir] : 1.f;
      atomicAddF(&lp[ir], -alpha * fr * lam[ax]);
      atomicAddF(&lp[il], alpha * fl * lam[ax]);
      self += alpha * fr * lam[ax] - alpha * fl * lam[ax];
    }
    atomicAddF(&lp[cell], self);
  });
}

// jacobi(level, in, out) reversed: out = (sum of the 6 apron cells of in - alpha div) / 6 on
// every cell of the level's blocks, so lin[apron] += lout/6, ld -= alpha lout/6, lout = 0.
template <int BW>
__global__ void __launch_bounds__(256) k_da_jacobi_T(DGBase g, int level, float* __restrict__ lout,
                                                     float* __restrict__ lin, float* __restrict__ ld, float alpha) {
  constexpr int BV = Geo<BW>::BV;
  forAllocatedCells<BW>(g, level, [&](uint32_t b, int, int t) {
    const uint32_t cell = b * BV + t;
    const float G = lout[cell] / 6.f;
    if (G == 0.f) return;
    uint32_t nb[6];
    faceNeighbours<BW>(g, b, t, nb);
#pragma unroll
    for (int k = 0; k < 6; k++) atomicAddF(&lin[nb[k]], G);
    ld[cell] += -alpha * G;
    lout[cell] = 0.f;
  });
}

// prolongate(level) reversed: pre = (27 q0 + 9 (q1+q2+q3) + 3 (q4+q5+q6) + q7) / 64 of the parent
// stencil, so each parent stencil cell gets its weight times lp/64, and lp is cleared.
template <int BW>
__global__ void __launch_bounds__(256) k_da_prolongate_T(DGBase g, int level, float* __restrict__ lp) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB;
  forAllocatedCells<BW>(g, level, [&](uint32_t b, int, int t) {
    const uint32_t ps = g.par[b];
    if (ps == NFU) return;
    const uint32_t cell = b * BV + t;
    const float G = lp[cell] / 64.f;
    if (G == 0.f) return;
    const int3 p0 = g.pos[b];
    int lx, ly, lz;
    locPos<BB>((uint32_t)t, lx, ly, lz);
    uint32_t q[8];
    parentStencil<BW>(g, ps, p0.x | lx, p0.y | ly, p0.z | lz, q);
    atomicAddF(&lp[q[0]], 27.f * G);
    atomicAddF(&lp[q[1]], 9.f * G);
    atomicAddF(&lp[q[2]], 9.f * G);
    atomicAddF(&lp[q[3]], 9.f * G);
    atomicAddF(&lp[q[4]], 3.f * G);
    atomicAddF(&lp[q[5]], 3.f * G);
    atomicAddF(&lp[q[6]], 3.f * G);
    atomicAddF(&lp[q[7]], G);
    lp[cell] = 0.f;
  });
}

// calcDivergence reversed (leaf cells): div = alpha (f_r v_r.x - f_l v_l.x + ...) with the
// neighbours' BC-applied velocities, so each in-domain neighbour gets +-alpha f lam_div in its
// component.
template <int BW, bool FLUIDITY>
__global__ void __launch_bounds__(256) k_da_div_T(DG g, dcg_params P, const float* __restrict__ ld,
                                                  float3* __restrict__ lu) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB, BB = Geo<BW>::BB;
  forAllocatedCells<BW>(g, -1, [&](uint32_t b, int level, int t) {
    if (g.chi[8 * b + (t >> SBB)] != NFU) return;
    const uint32_t cell = b * BV + t;
    const int scale = 1 << level;
    const float G = .5f * P.rdx / (float)scale * ld[cell];
    if (G == 0.f) return;
    uint32_t nb[6];
    faceNeighbours<BW>(g, b, t, nb);
    const int3 p0 = g.pos[b];
    int lx, ly, lz;
    locPos<BB>((uint32_t)t, lx, ly, lz);
    const int x = p0.x | lx, y = p0.y | ly, z = p0.z | lz;
    const int dx[6] = {-1, 1, 0, 0, 0, 0}, dy[6] = {0, 0, -1, 1, 0, 0}, dz[6] = {0, 0, 0, 0, -1, 1};
#pragma unroll
    for (int k = 0; k < 6; k++) {
      if (outside(P, x + dx[k], y + dy[k], z + dz[k], scale)) continue;
      const float f = FLUIDITY ? (float)g.flu[nb[k]] : 1.f;
      const float v = (k & 1 ? 1.f : -1.f) * f * G;
      float* comp = reinterpret_cast<float*>(&lu[nb[k]]) + (k >> 1);
      atomicAddF(comp, v);
    }
  });
}

// Gradient entries of the cells of free blocks are zero (no step reads or defines those cells).
template <int BW>
__global__ void __launch_bounds__(256) k_da_mask_free(DGBase g, float* __restrict__ d, float3* __restrict__ v,
                                                      float* __restrict__ t, size_t C) {
  constexpr int BV = Geo<BW>::BV;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < C; i += (size_t)gridDim.x * 256) {
    if (g.lev[i / BV] != FREE) continue;
    d[i] = 0.f;
    v[i] = make_float3(0.f, 0.f, 0.f);
    if (t) t[i] = 0.f;
  }
}

}  // namespace dcg

// oracle_uniform.cpp — CPU restatement of the reference's uniform-grid step (test
// infrastructure only; see oracle.h).
//
// Follows src/uniformgrid/fluid_simulation_uniform.cu, uniformgrid_fluid.cu and
// uniformgrid_structure.cu; boundary conditions from src/utils/sim_utils.cu.
#include <omp.h>

#include <algorithm>

#include "oracle_common.hpp"
#include "oracle_render.hpp"

namespace orc {

Sim* makeUniform(const int32_t size[3], const dcg_params& P, int precision, int* status,
                 std::string* err);

template <class R>
struct Uniform : Sim {
  int gx, gy, gz;
  size_t N, MC;
  int L;  // mip levels used by the solver
  std::vector<R> den, tem, t_den, t_tem, pre, t_pre, dive, flu;
  std::vector<V3<R>> vel, t_vel;

  int kind() const override { return DCG_KIND_UNIFORM; }

  // Mip level count, fluid_simulation_uniform.cu:8-17.
  int create(const int32_t size[3], const dcg_params& p) {
    P = p;
    gx = size[0];
    gy = size[1];
    gz = size[2];
    if (gx <= 0 || gy <= 0 || gz <= 0 || gx != P.gx || gy != P.gy || gz != P.gz)
      return fail(DCG_E_ARG, "size must be positive and equal params gx,gy,gz");
    if (P.moist) return fail(DCG_E_UNSUPPORTED, "the moist atmosphere is implemented for the DCGrid");
    const int minDim = std::min(gx, std::min(gy, gz));
    int cellSize = 2;
    L = 1;
    while (gx % cellSize == 0 && gy % cellSize == 0 && gz % cellSize == 0 &&
           cellSize * 4 <= minDim) {
      L++;
      cellSize *= 2;
    }
    N = (size_t)gx * gy * gz;
    MC = mipmapCells(gx, gy, gz);
    den.assign(N, 0);
    tem.assign(N, 0);
    t_den.assign(N, 0);
    t_tem.assign(N, 0);
    vel.assign(N, V3<R>{0, 0, 0});
    t_vel.assign(N, V3<R>{0, 0, 0});
    pre.assign(MC, 0);
    t_pre.assign(MC, 0);
    dive.assign(MC, 0);
    flu.assign(MC, 0);
    return reset();
  }

  size_t idx(int x, int y, int z) const { return ((size_t)z * gy + y) * gx + x; }
  size_t mip(int x, int y, int z, int scale) const { return mipmapIdx(x, y, z, gx, gy, gz, scale); }

  // reset(), fluid_simulation_uniform.cu:81-88
  int reset() override {
    std::fill(den.begin(), den.end(), (R)0);
    std::fill(tem.begin(), tem.end(), (R)0);
    std::fill(vel.begin(), vel.end(), V3<R>{0, 0, 0});
    std::fill(pre.begin(), pre.end(), (R)0);
    std::fill(flu.begin(), flu.end(), (R)0);
    return init();
  }
  // init(): adaptTopology then k_uniform_init (uniform.cu:76-79, uniformgrid_structure.cu:15-21)
  int init() override {
    adaptTopology();
    std::fill(den.begin(), den.end(), (R)0);
    std::fill(vel.begin(), vel.end(), V3<R>{0, 0, 0});
    std::fill(tem.begin(), tem.end(), (R)P.t_ambient);
    return DCG_OK;
  }
  // adaptTopology(): fluidity per mip level (uniform.cu:143-147; structure.cu:23-31)
  int adaptTopology() override {
    for (int level = 0; level < L; level++) {
      const int scale = 1 << level;
      const int w = gx >> level, h = gy >> level, d = gz >> level;
#pragma omp parallel for schedule(static)
      for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
          for (int x = 0; x < w; x++) flu[mip(x, y, z, scale)] = (R)cellFluidity(P, x, y, z, scale);
    }
    return DCG_OK;
  }

  // INIT_SAMPLE + advectVelocity/advectDensity (uniformgrid_fluid.cu:7-86).
  struct Sample {
    size_t i[8];
    R w[8];
    int x0, y0, z0;
    bool zero;
  };
  void sample(R px, R py, R pz, Sample& s) const {
    const R x = px - (R).5, y = py - (R).5, z = pz - (R).5;
    const R xf = std::floor(x), yf = std::floor(y), zf = std::floor(z);
    const int x0 = (int)xf, y0 = (int)yf, z0 = (int)zf;
    const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    const int x0c = clampi(x0, 0, gx - 1), y0c = clampi(y0, 0, gy - 1), z0c = clampi(z0, 0, gz - 1);
    const int x1c = clampi(x1, 0, gx - 1), y1c = clampi(y1, 0, gy - 1), z1c = clampi(z1, 0, gz - 1);
    // corner order 000,001,010,011,100,101,110,111 with bits [x y z]
    s.i[0] = idx(x0c, y0c, z0c);
    s.i[1] = idx(x0c, y0c, z1c);
    s.i[2] = idx(x0c, y1c, z0c);
    s.i[3] = idx(x0c, y1c, z1c);
    s.i[4] = idx(x1c, y0c, z0c);
    s.i[5] = idx(x1c, y0c, z1c);
    s.i[6] = idx(x1c, y1c, z0c);
    s.i[7] = idx(x1c, y1c, z1c);
    const R dx = x - xf, dy = y - yf, dz = z - zf;
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
  }
  V3<R> sampleVelocity(const Sample& s) const {
    if (s.zero) return V3<R>{0, 0, 0};
    V3<R> acc{0, 0, 0};
    for (int c = 0; c < 8; c++) {
      const int x = s.x0 + ((c >> 2) & 1), y = s.y0 + ((c >> 1) & 1), z = s.z0 + (c & 1);
      const V3<R> v = velocityBC(P, vel[s.i[c]], x, y, z, 1);
      acc = c == 0 ? mul(v, s.w[0]) : add(acc, mul(v, s.w[c]));
    }
    return acc;
  }
  template <class F>
  R sampleScalar(const Sample& s, const std::vector<R>& q, F bc) const {
    if (s.zero) return (R)0;
    R acc = 0;
    for (int c = 0; c < 8; c++) {
      const int x = s.x0 + ((c >> 2) & 1), y = s.y0 + ((c >> 1) & 1), z = s.z0 + (c & 1);
      const R v = bc(q[s.i[c]], x, y, z);
      acc = c == 0 ? v * s.w[0] : acc + v * s.w[c];
    }
    return acc;
  }

  // advectVelocity(): k_uniform_advect_velocity + D2D copy (uniform.cu:90-94;
  // uniformgrid_fluid.cu:88-95). Buoyancy (extension) is applied to the advected value.
  int advectVelocity() override {
    const R dt = (R)P.dt, rdx = (R)P.rdx;
    const bool buoy = P.source_mode == DCG_SOURCE_SPHERE_BUOYANT;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < gz; z++)
      for (int y = 0; y < gy; y++)
        for (int x = 0; x < gx; x++) {
          const size_t i = idx(x, y, z);
          const V3<R>& v = vel[i];
          const R px = ((R)x + (R).5) - v.x * dt * rdx;
          const R py = ((R)y + (R).5) - v.y * dt * rdx;
          const R pz = ((R)z + (R).5) - v.z * dt * rdx;
          Sample s;
          sample(px, py, pz, s);
          V3<R> nv = sampleVelocity(s);
          if (buoy) {
            const R b = dt * ((R)P.buoy_beta * (tem[i] - (R)P.t_ambient) - (R)P.buoy_alpha * den[i]);
            nv.y = nv.y + b;
          }
          t_vel[i] = nv;
        }
    std::swap(vel, t_vel);
    return DCG_OK;
  }

  // advectDensity(): k_uniform_advect_density + D2D copy (uniform.cu:137-141;
  // uniformgrid_fluid.cu:97-105). Temperature and the sphere source are extensions.
  int advectDensity() override {
    const R dt = (R)P.dt, rdx = (R)P.rdx;
    const bool src = P.source_mode == DCG_SOURCE_SPHERE_BUOYANT;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < gz; z++)
      for (int y = 0; y < gy; y++)
        for (int x = 0; x < gx; x++) {
          const size_t i = idx(x, y, z);
          const V3<R>& v = vel[i];
          const R px = ((R)x + (R).5) - v.x * dt * rdx;
          const R py = ((R)y + (R).5) - v.y * dt * rdx;
          const R pz = ((R)z + (R).5) - v.z * dt * rdx;
          Sample s;
          sample(px, py, pz, s);
          R q = sampleScalar(s, den, [&](R a, int xx, int yy, int zz) { return densityBC(P, a, xx, yy, zz, 1); });
          if (src) {
            R t = sampleScalar(s, tem, [&](R a, int xx, int yy, int zz) { return temperatureBC(P, a, xx, yy, zz, 1); });
            if (inSource(P, (float)x + .5f, (float)y + .5f, (float)z + .5f)) {
              q = (R)P.src_density;
              t = (R)sourceTemperature(P, x, y, z, 0);
            }
            t_tem[i] = t;
          }
          t_den[i] = q;
        }
    std::swap(den, t_den);
    if (src) std::swap(tem, t_tem);
    return DCG_OK;
  }

  // k_uniform_calc_divergence, uniformgrid_fluid.cu:107-132
  void calcDivergence() {
    const R alpha = (R).5f * (R)P.rdx;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < gz; z++)
      for (int y = 0; y < gy; y++)
        for (int x = 0; x < gx; x++) {
          const size_t i = idx(x, y, z);
          const size_t il = x > 0 ? i - 1 : i, ir = x < gx - 1 ? i + 1 : i;
          const size_t id = y > 0 ? i - gx : i, iu = y < gy - 1 ? i + gx : i;
          const size_t ib = z > 0 ? i - (size_t)gx * gy : i, iF = z < gz - 1 ? i + (size_t)gx * gy : i;
          pre[i] = 0;
          t_pre[i] = 0;
          const R wl = flu[il], wr = flu[ir], wd = flu[id], wu = flu[iu], wb = flu[ib], wf = flu[iF];
          const V3<R> vl = velocityBC(P, vel[il], x - 1, y, z, 1);
          const V3<R> vr = velocityBC(P, vel[ir], x + 1, y, z, 1);
          const V3<R> vd = velocityBC(P, vel[id], x, y - 1, z, 1);
          const V3<R> vu = velocityBC(P, vel[iu], x, y + 1, z, 1);
          const V3<R> vb = velocityBC(P, vel[ib], x, y, z - 1, 1);
          const V3<R> vf = velocityBC(P, vel[iF], x, y, z + 1, 1);
          dive[i] = alpha * (wr * vr.x - wl * vl.x + wu * vu.y - wd * vd.y + wf * vf.z - wb * vb.z);
        }
  }
  // k_uniform_restrict, uniformgrid_fluid.cu:134-160
  void restrictLevel(int level) {
    const int scale = 1 << level, childScale = scale / 2;
    const int w = gx / childScale, h = gy / childScale;
    const int W = gx >> level, H = gy >> level, D = gz >> level;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < D; z++)
      for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
          const size_t i = mip(x, y, z, scale);
          const size_t i000 = mip(x * 2, y * 2, z * 2, childScale);
          const size_t i001 = i000 + 1, i010 = i000 + w, i011 = i010 + 1;
          const size_t i100 = i000 + (size_t)w * h, i101 = i100 + 1, i110 = i100 + w, i111 = i110 + 1;
          pre[i] = 0;
          t_pre[i] = 0;
          dive[i] = (R).125f * (dive[i000] + dive[i001] + dive[i010] + dive[i011] + dive[i100] +
                               dive[i101] + dive[i110] + dive[i111]);
        }
  }
  // calcPressure<in,out>, uniformgrid_fluid.cu:162-192
  void jacobi(int level, const std::vector<R>& in, std::vector<R>& out) {
    const int scale = 1 << level;
    const R alpha = (R)P.dx * (R)P.dx * (R)scale * (R)scale;
    const int w = gx / scale, h = gy / scale, d = gz / scale;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < d; z++)
      for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
          const size_t i = mip(x, y, z, scale);
          const size_t il = x > 0 ? i - 1 : i, ir = x < w - 1 ? i + 1 : i;
          const size_t id = y > 0 ? i - w : i, iu = y < h - 1 ? i + w : i;
          const size_t ib = z > 0 ? i - (size_t)w * h : i, iF = z < d - 1 ? i + (size_t)w * h : i;
          out[i] = (in[il] + in[ir] + in[id] + in[iu] + in[ib] + in[iF] - alpha * dive[i]) / (R)6;
        }
  }
  // k_uniform_prolongate, uniformgrid_fluid.cu:206-237
  void prolongate(int level) {
    const int scale = 1 << level;
    const int w = gx / scale, h = gy / scale, d = gz / scale;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < d; z++)
      for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
          const size_t i = mip(x, y, z, scale);
          const size_t i000 = mip(x / 2, y / 2, z / 2, 2 * scale);
          const long sx = (x == 0 || x == w - 1) ? 0 : 2 * (x % 2) - 1;
          const long sy = (y == 0 || y == h - 1) ? 0 : 2 * (y % 2) - 1;
          const long sz = (z == 0 || z == d - 1) ? 0 : 2 * (z % 2) - 1;
          const size_t i001 = i000 + sx;
          const size_t i010 = i000 + sy * w / 2;
          const size_t i011 = i010 + sx;
          const size_t i100 = i000 + sz * (w / 2) * (h / 2);
          const size_t i101 = i100 + sx;
          const size_t i110 = i100 + sy * (w / 2);
          const size_t i111 = i110 + sx;
          pre[i] = ((R)27 * pre[i000] + (R)9 * (pre[i001] + pre[i010] + pre[i100]) +
                    (R)3 * (pre[i011] + pre[i101] + pre[i110]) + pre[i111]) /
                   (R)64;
        }
  }
  // k_uniform_apply_pressure, uniformgrid_fluid.cu:239-260
  void applyPressure() {
    const R alpha = (R).5f * (R)P.rdx;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < gz; z++)
      for (int y = 0; y < gy; y++)
        for (int x = 0; x < gx; x++) {
          const size_t i = idx(x, y, z);
          const size_t il = x > 0 ? i - 1 : i, ir = x < gx - 1 ? i + 1 : i;
          const size_t id = y > 0 ? i - gx : i, iu = y < gy - 1 ? i + gx : i;
          const size_t ib = z > 0 ? i - (size_t)gx * gy : i, iF = z < gz - 1 ? i + (size_t)gx * gy : i;
          const R p = pre[i];
          V3<R>& v = vel[i];
          v.x = v.x - alpha * (flu[ir] * (pre[ir] - p) + flu[il] * (p - pre[il]));
          v.y = v.y - alpha * (flu[iu] * (pre[iu] - p) + flu[id] * (p - pre[id]));
          v.z = v.z - alpha * (flu[iF] * (pre[iF] - p) + flu[ib] * (p - pre[ib]));
        }
  }
  // project(), fluid_simulation_uniform.cu:96-124
  int project() override {
    calcDivergence();
    for (int level = 1; level < L; level++) restrictLevel(level);
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
    return DCG_OK;
  }
  // projectLocal(), fluid_simulation_uniform.cu:126-135
  int projectLocal() override {
    calcDivergence();
    for (int i = 0; i < P.project_local_pairs; i++) {
      jacobi(0, pre, t_pre);
      jacobi(0, t_pre, pre);
    }
    applyPressure();
    return DCG_OK;
  }
  // ------------------------------------------------------------------ adjoint (extension)
  // Reverse mode of the step above on a fixed fluidity field. Each forward loop `out = f(in)`
  // is mirrored, in reverse order, by `lam_in += (df/din)^T lam_out` followed by
  // `lam_out = 0` (the output overwrote its previous value). Written as plain scatters.
  struct Snap {
    std::vector<R> den, tem;
    std::vector<V3<R>> vel;
  };
  std::vector<Snap> tape;

  int tapeForward(int n) override {
    if (n < 0) return fail(DCG_E_ARG, "negative step count");
    tape.clear();
    tape.push_back(Snap{den, tem, vel});
    for (int k = 0; k < n; k++) {
      advectVelocity();
      adaptTopology();
      if (P.project_mode == DCG_PROJECT_LOCAL)
        projectLocal();
      else if (P.project_mode == DCG_PROJECT_SOLVE)
        return fail(DCG_E_UNSUPPORTED, "the adjoint follows the Jacobi schedules");
      else
        project();
      advectDensity();
      tape.push_back(Snap{den, tem, vel});
    }
    return DCG_OK;
  }

  bool inDomain(int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < gx && y < gy && z < gz; }

  // Sample of the foot of cell (x,y,z) under velocity v, with the unnormalised corner weights
  // f_c t_c and d(t_c)/d(d_a) for the derivative of the normalised interpolant.
  struct FootD {
    Sample s;
    R f[8], t[8], dtd[8][3], W;
  };
  void footD(int x, int y, int z, const V3<R>& v, FootD& F) const {
    const R dt = (R)P.dt, rdx = (R)P.rdx;
    const R px = ((R)x + (R).5) - v.x * dt * rdx;
    const R py = ((R)y + (R).5) - v.y * dt * rdx;
    const R pz = ((R)z + (R).5) - v.z * dt * rdx;
    sample(px, py, pz, F.s);
    const R xs = px - (R).5, ys = py - (R).5, zs = pz - (R).5;
    const R d[3] = {xs - std::floor(xs), ys - std::floor(ys), zs - std::floor(zs)};
    F.W = 0;
    for (int c = 0; c < 8; c++) {
      const int bit[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
      R tau[3];
      for (int a = 0; a < 3; a++) tau[a] = bit[a] ? d[a] : (R)1 - d[a];
      F.t[c] = tau[0] * tau[1] * tau[2];
      for (int a = 0; a < 3; a++) {
        R g = bit[a] ? (R)1 : (R)-1;
        for (int b = 0; b < 3; b++)
          if (b != a) g *= tau[b];
        F.dtd[c][a] = g;
      }
      F.f[c] = flu[F.s.i[c]];
      F.W += F.f[c] * F.t[c];
    }
  }
  // d(value)/d(u_a) of the sample with BC-applied corner values vc[8] (value = sum f t vc / W).
  void footGrad(const FootD& F, const R vc[8], R out[3]) const {
    const R k = -(R)P.dt * (R)P.rdx;
    R val = 0;
    for (int c = 0; c < 8; c++) val += F.f[c] * F.t[c] * vc[c];
    val /= F.W;
    for (int a = 0; a < 3; a++) {
      R num = 0, den_ = 0;
      for (int c = 0; c < 8; c++) {
        num += F.f[c] * vc[c] * F.dtd[c][a];
        den_ += F.f[c] * F.dtd[c][a];
      }
      out[a] = k * (num - val * den_) / F.W;
    }
  }
  int corner(const Sample& s, int c, int& x, int& y, int& z) const {
    x = s.x0 + ((c >> 2) & 1);
    y = s.y0 + ((c >> 1) & 1);
    z = s.z0 + (c & 1);
    return inDomain(x, y, z);
  }

  // Adjoint of project() / projectLocal(): lu holds lambda of the velocity after the
  // projection on entry and lambda of the velocity before it on exit.
  void projectAdjoint(std::vector<V3<R>>& lu) {
    std::vector<R> lp(MC, 0), ltp(MC, 0), ld(MC, 0);
    const R a = (R).5f * (R)P.rdx;
    // applyPressure: v -= a (f_r (p_r - p) + f_l (p - p_l)) per axis; v itself passes through
    for (int z = 0; z < gz; z++)
      for (int y = 0; y < gy; y++)
        for (int x = 0; x < gx; x++) {
          const size_t i = idx(x, y, z);
          const size_t il = x > 0 ? i - 1 : i, ir = x < gx - 1 ? i + 1 : i;
          const size_t id = y > 0 ? i - gx : i, iu = y < gy - 1 ? i + gx : i;
          const size_t ib = z > 0 ? i - (size_t)gx * gy : i, iF = z < gz - 1 ? i + (size_t)gx * gy : i;
          const size_t nb[3][2] = {{il, ir}, {id, iu}, {ib, iF}};
          const R lam[3] = {lu[i].x, lu[i].y, lu[i].z};
          for (int ax = 0; ax < 3; ax++) {
            const size_t l = nb[ax][0], r = nb[ax][1];
            const R fl = flu[l], fr = flu[r];
            lp[r] += -a * fr * lam[ax];
            lp[i] += a * fr * lam[ax];
            lp[i] += -a * fl * lam[ax];
            lp[l] += a * fl * lam[ax];
          }
        }
    // Jacobi half-sweep in -> out on `level`, reversed
    auto jacobiAdj = [&](int level, std::vector<R>& lin, std::vector<R>& lout) {
      const int scale = 1 << level;
      const R alpha = (R)P.dx * (R)P.dx * (R)scale * (R)scale;
      const int w = gx / scale, h = gy / scale, d = gz / scale;
      for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
          for (int x = 0; x < w; x++) {
            const size_t i = mip(x, y, z, scale);
            const size_t il = x > 0 ? i - 1 : i, ir = x < w - 1 ? i + 1 : i;
            const size_t id = y > 0 ? i - w : i, iu = y < h - 1 ? i + w : i;
            const size_t ib = z > 0 ? i - (size_t)w * h : i, iF = z < d - 1 ? i + (size_t)w * h : i;
            const R g = lout[i] / (R)6;
            lin[il] += g;
            lin[ir] += g;
            lin[id] += g;
            lin[iu] += g;
            lin[ib] += g;
            lin[iF] += g;
            ld[i] += -alpha * g;
          }
      for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
          for (int x = 0; x < w; x++) lout[mip(x, y, z, scale)] = 0;
    };
    auto pairsAdj = [&](int level, int pairs) {
      for (int i = 0; i < pairs; i++) {
        jacobiAdj(level, ltp, lp);  // t_pre -> pre
        jacobiAdj(level, lp, ltp);  // pre -> t_pre
      }
    };
    auto zeroLevel = [&](std::vector<R>& v, int level) {
      const int scale = 1 << level;
      for (int z = 0; z < gz / scale; z++)
        for (int y = 0; y < gy / scale; y++)
          for (int x = 0; x < gx / scale; x++) v[mip(x, y, z, scale)] = 0;
    };
    if (P.project_mode == DCG_PROJECT_LOCAL) {
      pairsAdj(0, P.project_local_pairs);
    } else {
      for (int level = 0; level <= L - 2; level++) {
        pairsAdj(level, P.project_level_pairs);
        // prolongate(level): pre[level] = sum of 27/9/3/1 weights of pre[level + 1]
        const int scale = 1 << level;
        const int w = gx / scale, h = gy / scale, d = gz / scale;
        for (int z = 0; z < d; z++)
          for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
              const size_t i = mip(x, y, z, scale);
              const size_t i000 = mip(x / 2, y / 2, z / 2, 2 * scale);
              const long sx = (x == 0 || x == w - 1) ? 0 : 2 * (x % 2) - 1;
              const long sy = (y == 0 || y == h - 1) ? 0 : 2 * (y % 2) - 1;
              const long sz = (z == 0 || z == d - 1) ? 0 : 2 * (z % 2) - 1;
              const size_t i001 = i000 + sx, i010 = i000 + sy * w / 2, i011 = i010 + sx;
              const size_t i100 = i000 + sz * (w / 2) * (h / 2), i101 = i100 + sx, i110 = i100 + sy * (w / 2);
              const size_t i111 = i110 + sx;
              const R g = lp[i] / (R)64;
              lp[i000] += (R)27 * g;
              lp[i001] += (R)9 * g;
              lp[i010] += (R)9 * g;
              lp[i100] += (R)9 * g;
              lp[i011] += (R)3 * g;
              lp[i101] += (R)3 * g;
              lp[i110] += (R)3 * g;
              lp[i111] += g;
              lp[i] = 0;
            }
      }
      pairsAdj(L - 1, P.project_top_pairs);
      // restrictLevel(level) for level = L-1 .. 1: zeroes pre/t_pre, div = mean of children
      for (int level = L - 1; level >= 1; level--) {
        zeroLevel(lp, level);
        zeroLevel(ltp, level);
        const int scale = 1 << level, cs = scale / 2;
        for (int z = 0; z < gz / scale; z++)
          for (int y = 0; y < gy / scale; y++)
            for (int x = 0; x < gx / scale; x++) {
              const size_t i = mip(x, y, z, scale);
              const R g = (R).125f * ld[i];
              for (int c = 0; c < 8; c++)
                ld[mip(2 * x + ((c >> 2) & 1), 2 * y + ((c >> 1) & 1), 2 * z + (c & 1), cs)] += g;
              ld[i] = 0;
            }
      }
    }
    // calcDivergence: div = a (w_r v_r.x - w_l v_l.x + ...) with BC'd neighbour velocities
    for (int z = 0; z < gz; z++)
      for (int y = 0; y < gy; y++)
        for (int x = 0; x < gx; x++) {
          const size_t i = idx(x, y, z);
          const R g = a * ld[i];
          if (x + 1 < gx) lu[i + 1].x += flu[i + 1] * g;
          if (x - 1 >= 0) lu[i - 1].x += -flu[i - 1] * g;
          if (y + 1 < gy) lu[i + gx].y += flu[i + gx] * g;
          if (y - 1 >= 0) lu[i - gx].y += -flu[i - gx] * g;
          const size_t pl = (size_t)gx * gy;
          if (z + 1 < gz) lu[i + pl].z += flu[i + pl] * g;
          if (z - 1 >= 0) lu[i - pl].z += -flu[i - pl] * g;
        }
  }

  int tapeBackward(const double* gden, const double* gvel, const double* gtem, double* oden, double* ovel,
                   double* otem) override {
    if (tape.empty()) return fail(DCG_E_ARG, "no recorded rollout");
    const bool src = P.source_mode == DCG_SOURCE_SPHERE_BUOYANT;
    const R dt = (R)P.dt;
    std::vector<R> lq(N), lt(N, 0);
    std::vector<V3<R>> lu(N);
    for (size_t i = 0; i < N; i++) {
      lq[i] = (R)gden[i];
      lu[i] = V3<R>{(R)gvel[3 * i], (R)gvel[3 * i + 1], (R)gvel[3 * i + 2]};
      if (gtem) lt[i] = (R)gtem[i];
    }
    const int n = (int)tape.size() - 1;
    for (int k = n - 1; k >= 0; k--) {
      const Snap& S0 = tape[k];
      const Snap& S1 = tape[k + 1];
      // advectDensity with u_{k+1}: den_{k+1} = A(den_k), tem_{k+1} = A(tem_k) (with sources)
      std::vector<R> lq0(N, 0), lt0(N, 0);
      if (!src) lt0 = lt;  // temperature is only advected together with the source
      for (int z = 0; z < gz; z++)
        for (int y = 0; y < gy; y++)
          for (int x = 0; x < gx; x++) {
            const size_t i = idx(x, y, z);
            if (src && inSource(P, (float)x + .5f, (float)y + .5f, (float)z + .5f)) continue;
            FootD F;
            footD(x, y, z, S1.vel[i], F);
            if (F.s.zero) continue;
            R qc[8], tc[8];
            for (int c = 0; c < 8; c++) {
              int cx, cy, cz;
              corner(F.s, c, cx, cy, cz);
              qc[c] = densityBC(P, S0.den[F.s.i[c]], cx, cy, cz, 1);
              tc[c] = temperatureBC(P, S0.tem[F.s.i[c]], cx, cy, cz, 1);
              if (corner(F.s, c, cx, cy, cz)) {
                lq0[F.s.i[c]] += F.s.w[c] * lq[i];
                if (src) lt0[F.s.i[c]] += F.s.w[c] * lt[i];
              }
            }
            R g[3];
            footGrad(F, qc, g);
            lu[i].x += g[0] * lq[i];
            lu[i].y += g[1] * lq[i];
            lu[i].z += g[2] * lq[i];
            if (src) {
              footGrad(F, tc, g);
              lu[i].x += g[0] * lt[i];
              lu[i].y += g[1] * lt[i];
              lu[i].z += g[2] * lt[i];
            }
          }
      // project
      projectAdjoint(lu);
      // advectVelocity with u_k: u' = A(u_k; u_k) (+ buoyancy from den_k, tem_k)
      std::vector<V3<R>> lu0(N, V3<R>{0, 0, 0});
      for (int z = 0; z < gz; z++)
        for (int y = 0; y < gy; y++)
          for (int x = 0; x < gx; x++) {
            const size_t i = idx(x, y, z);
            if (src) {
              lq0[i] += -dt * (R)P.buoy_alpha * lu[i].y;
              lt0[i] += dt * (R)P.buoy_beta * lu[i].y;
            }
            FootD F;
            footD(x, y, z, S0.vel[i], F);
            if (F.s.zero) continue;
            R vc[3][8];
            for (int c = 0; c < 8; c++) {
              int cx, cy, cz;
              const bool in = corner(F.s, c, cx, cy, cz);
              const V3<R> v = velocityBC(P, S0.vel[F.s.i[c]], cx, cy, cz, 1);
              vc[0][c] = v.x;
              vc[1][c] = v.y;
              vc[2][c] = v.z;
              if (in) {
                lu0[F.s.i[c]].x += F.s.w[c] * lu[i].x;
                lu0[F.s.i[c]].y += F.s.w[c] * lu[i].y;
                lu0[F.s.i[c]].z += F.s.w[c] * lu[i].z;
              }
            }
            const R lam[3] = {lu[i].x, lu[i].y, lu[i].z};
            for (int comp = 0; comp < 3; comp++) {
              R g[3];
              footGrad(F, vc[comp], g);
              lu0[i].x += g[0] * lam[comp];
              lu0[i].y += g[1] * lam[comp];
              lu0[i].z += g[2] * lam[comp];
            }
          }
      lq.swap(lq0);
      lt.swap(lt0);
      lu.swap(lu0);
    }
    for (size_t i = 0; i < N; i++) {
      oden[i] = (double)lq[i];
      ovel[3 * i] = (double)lu[i].x;
      ovel[3 * i + 1] = (double)lu[i].y;
      ovel[3 * i + 2] = (double)lu[i].z;
      if (otem) otem[i] = (double)lt[i];
    }
    return DCG_OK;
  }

  // debugStats(): k_uniform_debug_stats + host sum (uniform.cu:160-175; structure.cu:33-42)
  // Samplers of src/uniformgrid/uniformgrid_rendering.cu:5-47.
  struct MarchSampler {
    const Uniform* u;
    size_t idx(int x, int y, int z) const { return ((size_t)z * u->gy + y) * u->gx + x; }
    void sample(const dcg_params& P, int precise, F3 pos, Aux& a) const {
      if (!precise) {
        a.ipos[0] = clampi((int)pos.x, 0, u->gx - 1);
        a.ipos[1] = clampi((int)pos.y, 0, u->gy - 1);
        a.ipos[2] = clampi((int)pos.z, 0, u->gz - 1);
        const size_t c = idx(a.ipos[0], a.ipos[1], a.ipos[2]);
        for (int k = 0; k < 8; k++) a.cell[k] = c;
        a.density = (float)u->den[c];
        return;
      }
      const float x = pos.x - .5f, y = pos.y - .5f, z = pos.z - .5f;
      const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
      a.ipos[0] = (int)xf;
      a.ipos[1] = (int)yf;
      a.ipos[2] = (int)zf;
      a.frac[0] = x - xf;
      a.frac[1] = y - yf;
      a.frac[2] = z - zf;
      const int x0 = clampi(a.ipos[0], 0, u->gx - 1), y0 = clampi(a.ipos[1], 0, u->gy - 1);
      const int z0 = clampi(a.ipos[2], 0, u->gz - 1);
      const size_t xs = a.ipos[0] < u->gx - 1 ? 1 : 0;
      const size_t ys = a.ipos[1] < u->gy - 1 ? (size_t)u->gx : 0;
      const size_t zs = a.ipos[2] < u->gz - 1 ? (size_t)u->gx * u->gy : 0;
      a.cell[0] = idx(x0, y0, z0);
      a.cell[1] = a.cell[0] + zs;
      a.cell[2] = a.cell[0] + ys;
      a.cell[3] = a.cell[1] + ys;
      a.cell[4] = a.cell[0] + xs;
      a.cell[5] = a.cell[4] + zs;
      a.cell[6] = a.cell[4] + ys;
      a.cell[7] = a.cell[6] + zs;
      a.density = interp(u->den, a);
    }
    F3 velocity(const Aux& a) const { return interp3(u->vel, a); }
    float fluidity(const Aux& a) const { return interp(u->flu, a); }
  };
  int render(const dcg_render_params& r, float* out) override {
    renderImage(MarchSampler{this}, P, r, out);
    return DCG_OK;
  }

  int debugStats(double* out) override {
    const size_t numBins = N / 256;
    std::vector<R> stats(numBins, (R)0);
    for (size_t b = 0; b < numBins; b++) {
      R s = 0;
      for (size_t k = 0; k < 256; k++) s = s + den[b * 256 + k] * flu[b * 256 + k];
      stats[b] = s;
    }
    R sum = 0;
    for (size_t b = 0; b < numBins; b++) sum = sum + stats[b];
    *out = (double)sum;
    return DCG_OK;
  }
  int seedScene() override {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < gz; z++)
      for (int y = 0; y < gy; y++)
        for (int x = 0; x < gx; x++) {
          const size_t i = idx(x, y, z);
          const float cx = (float)x + .5f, cy = (float)y + .5f, cz = (float)z + .5f;
          float u[3];
          initVelocity(P, cx, cy, cz, u);
          vel[i] = V3<R>{(R)u[0], (R)u[1], (R)u[2]};
          if (P.source_mode == DCG_SOURCE_SPHERE_BUOYANT && inSource(P, cx, cy, cz)) {
            den[i] = (R)P.src_density;
            tem[i] = (R)sourceTemperature(P, x, y, z, 0);
          } else {
            den[i] = 0;
            tem[i] = (R)P.t_ambient;
          }
        }
    return DCG_OK;
  }
  int info(dcg_info* o) override {
    std::memset(o, 0, sizeof(*o));
    o->kind = DCG_KIND_UNIFORM;
    o->size[0] = gx;
    o->size[1] = gy;
    o->size[2] = gz;
    o->levels = L;
    o->num_cells = N;
    o->active_cells = N;
    return DCG_OK;
  }
  int fieldSize(int f, uint64_t* n) override {
    switch (f) {
      case DCG_FIELD_DENSITY:
      case DCG_FIELD_TEMPERATURE: *n = N; return DCG_OK;
      case DCG_FIELD_VELOCITY: *n = 3 * N; return DCG_OK;
      case DCG_FIELD_PRESSURE:
      case DCG_FIELD_T_PRESSURE:
      case DCG_FIELD_DIVERGENCE:
      case DCG_FIELD_FLUIDITY: *n = MC; return DCG_OK;
      default: return fail(DCG_E_ARG, "field not available on the uniform grid");
    }
  }
  std::vector<R>* scalar(int f) {
    switch (f) {
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
    if (f == DCG_FIELD_VELOCITY) {
      for (size_t i = 0; i < N; i++) {
        dst[3 * i] = (double)vel[i].x;
        dst[3 * i + 1] = (double)vel[i].y;
        dst[3 * i + 2] = (double)vel[i].z;
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
    if (f == DCG_FIELD_VELOCITY) {
      for (size_t i = 0; i < N; i++) vel[i] = V3<R>{(R)src[3 * i], (R)src[3 * i + 1], (R)src[3 * i + 2]};
      return DCG_OK;
    }
    std::vector<R>* v = scalar(f);
    for (size_t i = 0; i < n; i++) (*v)[i] = (R)src[i];
    return DCG_OK;
  }
  int exportTables(dcg_tables*) override { return fail(DCG_E_ARG, "uniform grid has no block tables"); }
  int runStage(int stage, int arg) override {
    const bool perLevel = stage == DCG_STAGE_JACOBI || stage == DCG_STAGE_JACOBI_INV || stage == DCG_STAGE_PROLONGATE;
    if (perLevel && (arg < 0 || arg >= L || (stage == DCG_STAGE_PROLONGATE && arg >= L - 1)))
      return fail(DCG_E_ARG, "level out of range");
    switch (stage) {
      case DCG_STAGE_DIVERGENCE: calcDivergence(); return DCG_OK;
      case DCG_STAGE_RESTRICT:
        for (int level = 1; level < L; level++) restrictLevel(level);
        return DCG_OK;
      case DCG_STAGE_JACOBI: jacobi(arg, pre, t_pre); return DCG_OK;
      case DCG_STAGE_JACOBI_INV: jacobi(arg, t_pre, pre); return DCG_OK;
      case DCG_STAGE_PROLONGATE: prolongate(arg); return DCG_OK;
      case DCG_STAGE_APPLY_PRESSURE: applyPressure(); return DCG_OK;
      default: return fail(DCG_E_ARG, "stage not available on the uniform grid");
    }
  }
};

Sim* makeUniform(const int32_t size[3], const dcg_params& P, int precision, int* status,
                 std::string* err) {
  Sim* s = nullptr;
  int st;
  if (precision == 32) {
    auto* u = new Uniform<float>();
    st = u->create(size, P);
    s = u;
  } else {
    auto* u = new Uniform<double>();
    st = u->create(size, P);
    s = u;
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

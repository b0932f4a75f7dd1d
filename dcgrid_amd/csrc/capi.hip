// capi.hip — extern "C" entry points of libdcgrid_amd (declared in include/dcgrid_amd.h) and
// the shared host machinery of SimBase (stream, fills, timing).
#include <cmath>
#include <cstring>
#include <mutex>

#include "dcg_internal.hpp"

namespace dcg {

__global__ void k_fill32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill8(uint8_t* __restrict__ p, uint8_t v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

static unsigned fillGrid(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g == 0) g = 1;
  return (unsigned)g;
}

SimBase::~SimBase() {
  if (stream) hipStreamSynchronize(stream);
  dropGraph();
  for (auto& p : pending) {
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  for (auto e : eventPool) hipEventDestroy(e);
  for (void* p : allocations) hipFree(p);
  for (void* p : hostAllocations) (void)hipHostFree(p);
  if (stream) hipStreamDestroy(stream);
}

int SimBase::steps(int n) {
  const bool useGraph = graphOn && graphCapable() && timingMask == 0 && P.project_mode != DCG_PROJECT_SOLVE && n >= 2;
  int i = 0;
  if (useGraph) {
    for (; i < n && captureUnsettled(); i++) {
      DCG_TRY(stepOnce());
      countHost(1);
    }
  }
  if (useGraph && n - i >= 2) {
    std::vector<char> key;
    graphKeyBytes(key);
    if (!graphExec || key != graphKey) {
      dropGraph();
      std::vector<char> before;
      hostSnapshot(before);
      DCG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      int e = stepOnce();
      if (!e) e = stepOnce();
      hipGraph_t captured = nullptr;
      const hipError_t ce = hipStreamEndCapture(stream, &captured);
      if (e) {  // nothing ran: the host state returns to where the capture began
        if (captured) (void)hipGraphDestroy(captured);
        hostRestore(before);
        return e;
      }
      if (ce != hipSuccess) {
        hostRestore(before);
        return fail(DCG_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
      }
      graph = captured;
      DCG_HIP(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0));
      std::vector<char> after;
      graphKeyBytes(after);
      if (after != key) {  // the captured pair must leave the recorded state as it found it
        dropGraph();
        hostRestore(before);
        return fail(DCG_E_HIP, "step capture changed the recorded state");
      }
      graphKey = key;
    }
    for (; i + 1 < n; i += 2) {
      DCG_HIP(hipGraphLaunch(graphExec, stream));
      countHost(2);
    }
  }
  for (; i < n; i++) {
    DCG_TRY(stepOnce());
    countHost(1);
  }
  return DCG_OK;
}

int SimBase::createStream() {
  DCG_HIP(hipGetDevice(&device));
  // Default priority on purpose.
  DCG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  return DCG_OK;
}

int SimBase::fill32(void* p, uint32_t v, size_t n) {
  if (n == 0) return DCG_OK;
  k_fill32<<<fillGrid(n), 256, 0, stream>>>((uint32_t*)p, v, n);
  return checkLaunch("fill32");
}
int SimBase::fill8(void* p, uint8_t v, size_t n) {
  if (n == 0) return DCG_OK;
  k_fill8<<<fillGrid(n), 256, 0, stream>>>((uint8_t*)p, v, n);
  return checkLaunch("fill8");
}
int SimBase::sync() {
  DCG_HIP(hipStreamSynchronize(stream));
  return DCG_OK;
}

int SimBase::beginTimed(int k, Pending* ev) {
  ev->k = -1;
  if (!(timingMask & (1u << k)) || batchK == k) return DCG_OK;
  for (int i = 0; i < 2; i++) {
    hipEvent_t e;
    if (!eventPool.empty()) {
      e = eventPool.back();
      eventPool.pop_back();
    } else {
      DCG_HIP(hipEventCreate(&e));
    }
    (i == 0 ? ev->a : ev->b) = e;
  }
  ev->k = k;
  DCG_HIP(hipEventRecord(ev->a, stream));
  return DCG_OK;
}
int SimBase::takeEvents(int k, Pending* ev) {
  ev->k = -1;
  if (!(timingMask & (1u << k))) return DCG_OK;
  for (int i = 0; i < 2; i++) {
    hipEvent_t e;
    if (!eventPool.empty()) {
      e = eventPool.back();
      eventPool.pop_back();
    } else {
      DCG_HIP(hipEventCreate(&e));
    }
    (i == 0 ? ev->a : ev->b) = e;
  }
  ev->k = k;
  ev->n = 1;
  return DCG_OK;
}
int SimBase::pushPending(Pending* ev) {
  if (ev->k < 0) return DCG_OK;
  pending.push_back(*ev);
  if (pending.size() > 4096) return collectTiming();
  return DCG_OK;
}
int SimBase::endTimed(int, Pending* ev, uint32_t launches) {
  if (ev->k < 0) return DCG_OK;
  ev->n = launches;
  DCG_HIP(hipEventRecord(ev->b, stream));
  pending.push_back(*ev);
  // Bound the number of outstanding events.
  if (pending.size() > 4096) return collectTiming();
  return DCG_OK;
}
int SimBase::collectTiming() {
  if (pending.empty()) return DCG_OK;
  DCG_HIP(hipStreamSynchronize(stream));
  for (auto& p : pending) {
    float ms = 0.f;
    DCG_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    timingMs[p.k] += ms;
    timingLaunches[p.k] += p.n;
    eventPool.push_back(p.a);
    eventPool.push_back(p.b);
  }
  pending.clear();
  return DCG_OK;
}

}  // namespace dcg

using dcg::SimBase;

struct dcg_sim {
  SimBase* impl;
};

static std::string g_create_error;

extern "C" {

int32_t dcg_abi_version(void) { return DCG_ABI_VERSION; }

// SimParams::defaultParams, src/data/sim_params.cpp:4-17; rdx = 1/dx as in src/main.cpp:17.
void dcg_default_params(dcg_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->gx = p->gy = p->gz = 128;
  p->dx = 10000.f / 128;
  p->rdx = 1.f / p->dx;
  p->dt = 3.f;
  p->velocity_emission_rate = 150.f;
  p->density_emission_rate = 0.002f;
  p->emission_radius = 750.f;
  p->enable_additional_solids = 0;
  p->block_width = 4;
  p->max_levels = 0;
  p->project_mode = DCG_PROJECT_MULTILEVEL;
  p->project_top_pairs = 5;    // fluid_simulation_dcgrid.cu:274
  p->project_level_pairs = 5;  // :284
  p->project_local_pairs = 10; // :301
  p->score_mode = DCG_SCORE_GEOMETRIC;
  p->refine_threshold = 1e-4f; // :461
  p->grad_threshold = 0.05f;
  p->vort_threshold = 0.5f;
  p->source_mode = DCG_SOURCE_NONE;
  p->noise_period = 32.f;
  p->seed = 20261202u;
  p->solve_precision = DCG_SOLVE_MIXED;
  p->solve_tol = 1e-8f;
  p->solve_max_cycles = 50;
  p->solve_smooth = 1;
}

int dcg_create_uniform(const int32_t size[3], const dcg_params* p, dcg_sim** out) {
  if (!size || !p || !out) return DCG_E_ARG;
  if (p->storage != DCG_STORAGE_FP32) {
    g_create_error = "16-bit field storage is implemented for the DCGrid";
    return DCG_E_UNSUPPORTED;
  }
  int st = 0;
  SimBase* s = dcg::makeUniform(size, *p, &st, &g_create_error);
  if (!s) return st;
  *out = new dcg_sim{s};
  return DCG_OK;
}

int dcg_create_dcgrid(const int32_t size[3], uint64_t max_num_blocks, const dcg_params* p, dcg_sim** out) {
  if (!size || !p || !out) return DCG_E_ARG;
  int st = 0;
  SimBase* s = dcg::makeDCGrid(size, max_num_blocks, *p, &st, &g_create_error);
  if (!s) return st;
  *out = new dcg_sim{s};
  return DCG_OK;
}

void dcg_destroy(dcg_sim* s) {
  if (!s) return;
  delete s->impl;
  delete s;
}

#define IMPL(s)                \
  if (!(s)) return DCG_E_ARG;  \
  SimBase* I = (s)->impl

int dcg_set_params(dcg_sim* s, const dcg_params* p) {
  IMPL(s);
  if (!p) return DCG_E_ARG;
  if (p->gx != I->P.gx || p->gy != I->P.gy || p->gz != I->P.gz || p->block_width != I->P.block_width ||
      p->max_levels != I->P.max_levels || p->moist != I->P.moist || p->storage != I->P.storage)
    return I->fail(DCG_E_ARG,
                   "grid size, block width, level cap, moist atmosphere and field storage are fixed at creation");
  I->P = *p;
  return DCG_OK;
}
int dcg_get_params(const dcg_sim* s, dcg_params* p) {
  if (!s || !p) return DCG_E_ARG;
  *p = s->impl->P;
  return DCG_OK;
}
int dcg_init(dcg_sim* s) {
  IMPL(s);
  return I->init();
}
int dcg_reset(dcg_sim* s) {
  IMPL(s);
  return I->reset();
}
int dcg_adapt_topology(dcg_sim* s) {
  IMPL(s);
  return I->adaptTopology();
}
int dcg_advect_velocity(dcg_sim* s) {
  IMPL(s);
  return I->advectVelocity();
}
int dcg_project(dcg_sim* s) {
  IMPL(s);
  return I->project();
}
int dcg_project_local(dcg_sim* s) {
  IMPL(s);
  return I->projectLocal();
}
int dcg_project_solve(dcg_sim* s) {
  IMPL(s);
  return I->projectSolve();
}
int dcg_solve_report(dcg_sim* s, dcg_solve_info* out) {
  IMPL(s);
  if (!out) return DCG_E_ARG;
  return I->solveReport(out);
}
int dcg_advect_density(dcg_sim* s) {
  IMPL(s);
  return I->advectDensity();
}
// Step order of src/simulation.cpp:104-111.
int dcg_step(dcg_sim* s, int32_t n) {
  IMPL(s);
  if (n < 0) return I->fail(DCG_E_ARG, "negative step count");
  return I->steps(n);
}
int dcg_set_graph(dcg_sim* s, int32_t enable) {
  IMPL(s);
  I->graphOn = enable != 0;
  if (!I->graphOn) I->dropGraph();
  return DCG_OK;
}
int dcg_debug_stats(dcg_sim* s, double* out) {
  IMPL(s);
  if (!out) return DCG_E_ARG;
  return I->debugStats(out);
}
int dcg_seed_scene(dcg_sim* s) {
  IMPL(s);
  return I->seedScene();
}
int dcg_synchronize(dcg_sim* s) {
  IMPL(s);
  return I->sync();
}
int dcg_info_get(dcg_sim* s, dcg_info* out) {
  IMPL(s);
  if (!out) return DCG_E_ARG;
  return I->info(out);
}
int dcg_field_size(dcg_sim* s, int32_t field, uint64_t* n) {
  IMPL(s);
  if (!n) return DCG_E_ARG;
  return I->fieldSize(field, n);
}
int dcg_export_field(dcg_sim* s, int32_t field, float* dst, uint64_t n) {
  IMPL(s);
  if (!dst) return DCG_E_ARG;
  return I->exportField(field, dst, n);
}
int dcg_import_field(dcg_sim* s, int32_t field, const float* src, uint64_t n) {
  IMPL(s);
  if (!src) return DCG_E_ARG;
  return I->importField(field, src, n);
}
int dcg_export_field_device(dcg_sim* s, int32_t field, float* dst, uint64_t n) {
  IMPL(s);
  if (!dst) return DCG_E_ARG;
  return I->exportFieldDevice(field, dst, n);
}
int dcg_import_field_device(dcg_sim* s, int32_t field, const float* src, uint64_t n) {
  IMPL(s);
  if (!src) return DCG_E_ARG;
  return I->importFieldDevice(field, src, n);
}
int dcg_tape_forward(dcg_sim* s, int32_t steps, int32_t checkpoint_every) {
  IMPL(s);
  return I->tapeForward(steps, checkpoint_every);
}
int dcg_tape_backward(dcg_sim* s, const float* g_density, const float* g_velocity, const float* g_temperature,
                      float* d_density, float* d_velocity, float* d_temperature) {
  IMPL(s);
  if (!g_density || !g_velocity || !d_density || !d_velocity) return DCG_E_ARG;
  return I->tapeBackward(g_density, g_velocity, g_temperature, d_density, d_velocity, d_temperature);
}
int dcg_export_tables(dcg_sim* s, dcg_tables* t) {
  IMPL(s);
  if (!t) return DCG_E_ARG;
  return I->exportTables(t);
}
// ------------------------------------------------------------------ rendering
void dcg_default_render_params(const dcg_sim* s, dcg_render_params* r) {
  if (!r) return;
  std::memset(r, 0, sizeof(*r));
  // SimParams::defaultParams render fields, src/data/sim_params.cpp:19-32 (render_channel as
  // set for the fluid view, src/simulation.cpp:67-68)
  r->solids = 1;
  r->shadows = 1;
  r->precise = 1;
  r->channel = DCG_RENDER_FLUID;
  r->aa_samples = 1.f;
  r->ambient = .3f;
  const float floor_c[3] = {178.f / 255.f, 158.f / 255.f, 135.f / 255.f};
  const float scene_c[3] = {107.f / 255.f, 163.f / 255.f, 204.f / 255.f};
  for (int a = 0; a < 3; a++) {
    r->floor_color[a] = floor_c[a];
    r->smoke_color[a] = .9f;
    r->scene_color[a] = scene_c[a];
  }
  // a camera (the reference takes it from its engine): 256 x 256, 45 degree field of view, from
  // in front of the domain's z = 0 face looking along +z slightly downwards, sun high
  r->width = r->height = 256;
  r->fov = .70710678f;
  const dcg_params* P = s && s->impl ? &s->impl->P : nullptr;
  const float gx = P ? (float)P->gx : 128.f, gy = P ? (float)P->gy : 128.f, gz = P ? (float)P->gz : 128.f;
  const float dx = P ? P->dx : 1.f;
  r->eye_origin[0] = .5f * gx * dx;
  r->eye_origin[1] = .6f * gy * dx;
  r->eye_origin[2] = -.9f * gz * dx;
  r->eye_dir[1] = -.2f;
  r->eye_dir[2] = .9797959f;
  r->sun_dir[0] = .3f;
  r->sun_dir[1] = .9055385f;
  r->sun_dir[2] = -.3f;
}

// Inputs for which every loop of the ray marcher ends: bounded samples per pixel, finite
// vectors, and eye positions within 1e6 cells (a march distance then stays below 2^22 cells, so
// its 0.5-cell steps always advance the float distance).
static int checkRender(SimBase* I, const dcg_render_params* r) {
  if (r->width <= 0 || r->height <= 0 || r->width > 16384 || r->height > 16384)
    return I->fail(DCG_E_ARG, "render resolution must be 1..16384");
  if (!(r->aa_samples >= 1.f && r->aa_samples <= 8.f)) return I->fail(DCG_E_ARG, "aa_samples must be in [1, 8]");
  if (r->channel < DCG_RENDER_FLUID || r->channel > DCG_RENDER_RESOLUTION) return I->fail(DCG_E_ARG, "unknown channel");
  const float* v[] = {r->eye_origin, r->eye_dir, r->sun_dir, r->background_color, r->floor_color, r->smoke_color,
                      r->scene_color};
  for (const float* a : v)
    for (int k = 0; k < 3; k++)
      if (!std::isfinite(a[k])) return I->fail(DCG_E_ARG, "render vectors must be finite");
  if (!std::isfinite(r->fov) || !std::isfinite(r->ambient)) return I->fail(DCG_E_ARG, "fov and ambient must be finite");
  for (int k = 0; k < 3; k++)
    if (!(std::fabs(r->eye_origin[k] * I->P.rdx) <= 1e6f)) return I->fail(DCG_E_ARG, "eye farther than 1e6 cells");
  return DCG_OK;
}

int dcg_render(dcg_sim* s, const dcg_render_params* r, float* rgba_device) {
  IMPL(s);
  if (!r || !rgba_device) return DCG_E_ARG;
  DCG_TRY(checkRender(I, r));
  return I->render(*r, rgba_device);
}

int dcg_render_host(dcg_sim* s, const dcg_render_params* r, float* rgba_host) {
  IMPL(s);
  if (!r || !rgba_host) return DCG_E_ARG;
  DCG_TRY(checkRender(I, r));
  const size_t bytes = (size_t)r->width * r->height * 4 * sizeof(float);
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, bytes);
  if (e != hipSuccess) return I->fail(DCG_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  int st = I->render(*r, (float*)d);
  if (!st) {
    e = hipMemcpyAsync(rgba_host, d, bytes, hipMemcpyDeviceToHost, I->stream);
    if (e != hipSuccess) st = I->fail(DCG_E_HIP, std::string("render copy: ") + hipGetErrorString(e));
  }
  const int sst = I->sync();
  (void)hipFree(d);
  return st ? st : sst;
}

// ------------------------------------------------------------------ allocation-free render
// dcg_render_host stages the image in device memory it allocates per call. Callers that cannot
// allocate device memory render through this file-scope array instead: one copy per GPU, held
// until the process ends, its device address taken once per device with that device current.
// In host code `scratch` names a host-side stand-in, never a device address: only the pointers
// hipGetSymbolAddress returns reach kernels and copies.
static constexpr size_t kScratchPixels = (size_t)1024 * 1024;  // 16 MiB of float4
static __device__ __align__(256) float4 scratch[kScratchPixels];
static std::mutex g_scratchMutex;
static constexpr int kScratchDevices = 64;
static void* g_scratchPtr[kScratchDevices] = {};

static int scratchAddress(SimBase* I, float** out) {
  if (I->device < 0 || I->device >= kScratchDevices) return I->fail(DCG_E_UNSUPPORTED, "device index beyond 64");
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchPtr[I->device]) {
    int prev = 0;
    hipError_t de = hipGetDevice(&prev);
    if (de == hipSuccess && prev != I->device) de = hipSetDevice(I->device);
    if (de != hipSuccess) return I->fail(DCG_E_HIP, std::string("select device: ") + hipGetErrorString(de));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != I->device) (void)hipSetDevice(prev);
    if (e != hipSuccess || !p)
      return I->fail(DCG_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratchPtr[I->device] = p;
  }
  *out = (float*)g_scratchPtr[I->device];
  return DCG_OK;
}

int dcg_render_host_noalloc(dcg_sim* s, const dcg_render_params* r, float* rgba_host) {
  IMPL(s);
  if (!r || !rgba_host) return DCG_E_ARG;
  DCG_TRY(checkRender(I, r));
  const size_t pixels = (size_t)r->width * r->height;
  if (pixels > kScratchPixels)
    return I->fail(DCG_E_ARG, "dcg_render_host_noalloc: image of " + std::to_string(pixels) +
                                  " pixels exceeds the 1048576-pixel (1024 x 1024) render scratch");
  float* d = nullptr;
  DCG_TRY(scratchAddress(I, &d));
  DCG_TRY(I->render(*r, d));  // every pixel of the image is written: nothing to clear
  const hipError_t e = hipMemcpyAsync(rgba_host, d, pixels * 4 * sizeof(float), hipMemcpyDeviceToHost, I->stream);
  if (e != hipSuccess) return I->fail(DCG_E_HIP, std::string("render copy: ") + hipGetErrorString(e));
  return I->sync();
}

// ------------------------------------------------------------------ state format
namespace {
struct StateHeader {
  char magic[8];
  uint32_t version, headerBytes;
  int32_t kind, storage, size[3], blockWidth, levels;
  uint32_t flags, paramsBytes, sectionCount;
  uint64_t maxNumBlocks, numCells;
};
static_assert(sizeof(StateHeader) == 72, "state header layout");
struct StateEntry {
  uint32_t id, elem;
  uint64_t offset, bytes;
};
static_assert(sizeof(StateEntry) == 24, "state entry layout");
const char kStateMagic[8] = {'D', 'C', 'G', 'S', 'T', 'A', 'T', 'E'};
uint64_t align256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

// Header, section entries and total size of the state of I as it is now.
int stateLayout(dcg::SimBase* I, StateHeader& h, std::vector<StateEntry>& ents,
                std::vector<dcg::SimBase::StateSection>& secs, uint64_t& total) {
  dcg_info inf;
  DCG_TRY(I->info(&inf));
  I->stateSections(secs);
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kStateMagic, 8);
  h.version = DCG_STATE_VERSION;
  h.kind = inf.kind;
  h.storage = I->P.storage;
  for (int a = 0; a < 3; a++) h.size[a] = inf.size[a];
  h.blockWidth = inf.block_width;
  h.levels = inf.levels;
  h.flags = I->stateFlags();
  h.paramsBytes = (uint32_t)sizeof(dcg_params);
  h.sectionCount = (uint32_t)secs.size();
  h.maxNumBlocks = inf.max_num_blocks;
  h.numCells = inf.num_cells;
  h.headerBytes = (uint32_t)(sizeof(StateHeader) + sizeof(dcg_params) + secs.size() * sizeof(StateEntry));
  ents.clear();
  uint64_t off = align256(h.headerBytes);
  for (const auto& s : secs) {
    ents.push_back({s.id, s.elem, off, s.bytes});
    off = align256(off + s.bytes);
  }
  total = off;
  return DCG_OK;
}

// A state buffer checked on its own: header, section table inside the buffer and, for a DCGrid,
// the block tables against the plan of the grid the header describes (every index a kernel
// follows from them stays inside its array). Nothing here needs a device.
int checkStateBuffer(const char* in, uint64_t bytes, std::string& err) {
  StateHeader fh;
  if (bytes < sizeof(fh)) {
    err = "state truncated (header)";
    return DCG_E_ARG;
  }
  std::memcpy(&fh, in, sizeof(fh));
  if (std::memcmp(fh.magic, kStateMagic, 8) != 0) {
    err = "not a DCGSTATE buffer";
    return DCG_E_ARG;
  }
  if (fh.version != DCG_STATE_VERSION) {
    err = "state version " + std::to_string(fh.version) + " (this library reads " + std::to_string(DCG_STATE_VERSION) + ")";
    return DCG_E_UNSUPPORTED;
  }
  if (fh.paramsBytes != sizeof(dcg_params) || fh.sectionCount > 64 ||
      fh.headerBytes != sizeof(StateHeader) + sizeof(dcg_params) + (uint64_t)fh.sectionCount * sizeof(StateEntry)) {
    err = "state header does not match this library's parameter layout";
    return DCG_E_ARG;
  }
  if (bytes < fh.headerBytes) {
    err = "state truncated (section table)";
    return DCG_E_ARG;
  }
  std::vector<StateEntry> fe(fh.sectionCount);
  std::memcpy(fe.data(), in + sizeof(fh) + sizeof(dcg_params), fe.size() * sizeof(StateEntry));
  for (const auto& e : fe)
    if (e.offset < fh.headerBytes || e.offset > bytes || e.bytes > bytes - e.offset) {
      err = "state truncated (section data)";
      return DCG_E_ARG;
    }
  if (fh.kind != DCG_KIND_DCGRID) return DCG_OK;
  dcg_params p;
  std::memcpy(&p, in + sizeof(fh), sizeof(p));
  dcg::DCGridPlan pl;
  std::string perr;
  if (dcg::planDCGrid(fh.blockWidth, fh.size, fh.maxNumBlocks, p, pl, perr) != DCG_OK) {
    err = "state describes a grid the planner rejects: " + perr;
    return DCG_E_ARG;
  }
  if (pl.L != fh.levels) {
    err = "state level count does not match its grid";
    return DCG_E_ARG;
  }
  const uint64_t B = pl.B;
  struct Need {
    uint32_t id;
    uint64_t bytes;
    const void* at;
  } need[] = {{DCG_STATE_CONTROL, 16 * 4, nullptr},       {DCG_STATE_POSITIONS, 12 * B, nullptr},
              {DCG_STATE_LEVELS, B, nullptr},             {DCG_STATE_PARENTS, 4 * B, nullptr},
              {DCG_STATE_CHILDREN, 32 * B, nullptr},      {DCG_STATE_NEIGHBOURS, 27 * 16 * B, nullptr},
              {DCG_STATE_HASH_VALUES, 4 * (uint64_t)pl.H, nullptr}};
  for (auto& n : need) {
    for (const auto& e : fe)
      if (e.id == n.id && (n.id == DCG_STATE_CONTROL ? e.bytes >= n.bytes : e.bytes == n.bytes)) n.at = in + e.offset;
    if (!n.at) {
      err = "state lacks block table section " + std::to_string(n.id) + " of its grid's size";
      return DCG_E_ARG;
    }
  }
  dcg::StateTables t{(const uint32_t*)need[0].at, (const int32_t*)need[1].at, (const uint8_t*)need[2].at,
                (const uint32_t*)need[3].at, (const uint32_t*)need[4].at, (const int32_t*)need[5].at,
                (const uint32_t*)need[6].at};
  return dcg::checkDCGridTables(pl, t, err);
}
}  // namespace

int dcg_state_check(const void* src, uint64_t bytes) {
  if (!src) return DCG_E_ARG;
  std::string err;
  const int st = checkStateBuffer((const char*)src, bytes, err);
  g_create_error = st ? err : std::string();
  return st;
}

int dcg_state_size(dcg_sim* s, uint64_t* bytes) {
  IMPL(s);
  if (!bytes) return DCG_E_ARG;
  StateHeader h;
  std::vector<StateEntry> ents;
  std::vector<SimBase::StateSection> secs;
  return stateLayout(I, h, ents, secs, *bytes);
}

int dcg_save_state(dcg_sim* s, void* dst, uint64_t bytes) {
  IMPL(s);
  if (!dst) return DCG_E_ARG;
  StateHeader h;
  std::vector<StateEntry> ents;
  std::vector<SimBase::StateSection> secs;
  uint64_t total;
  DCG_TRY(stateLayout(I, h, ents, secs, total));
  if (bytes < total) return I->fail(DCG_E_ARG, "state buffer smaller than dcg_state_size");
  char* out = (char*)dst;
  std::memset(out, 0, align256(h.headerBytes));
  std::memcpy(out, &h, sizeof(h));
  std::memcpy(out + sizeof(h), &I->P, sizeof(dcg_params));
  std::memcpy(out + sizeof(h) + sizeof(dcg_params), ents.data(), ents.size() * sizeof(StateEntry));
  DCG_TRY(I->sync());  // every launch that writes the state has finished
  for (size_t i = 0; i < secs.size(); i++) {
    const uint64_t pad = align256(ents[i].offset + ents[i].bytes) - (ents[i].offset + ents[i].bytes);
    if (pad) std::memset(out + ents[i].offset + ents[i].bytes, 0, pad);
    const hipError_t e = hipMemcpyAsync(out + ents[i].offset, secs[i].dev, secs[i].bytes, hipMemcpyDeviceToHost, I->stream);
    if (e != hipSuccess) return I->fail(DCG_E_HIP, std::string("state copy: ") + hipGetErrorString(e));
  }
  return I->sync();
}

int dcg_load_state(dcg_sim* s, const void* src, uint64_t bytes) {
  IMPL(s);
  if (!src) return DCG_E_ARG;
  const char* in = (const char*)src;
  StateHeader fh;
  if (bytes < sizeof(fh)) return I->fail(DCG_E_ARG, "state truncated (header)");
  std::memcpy(&fh, in, sizeof(fh));
  if (std::memcmp(fh.magic, kStateMagic, 8) != 0) return I->fail(DCG_E_ARG, "not a DCGSTATE buffer");
  if (fh.version != DCG_STATE_VERSION)
    return I->fail(DCG_E_UNSUPPORTED, "state version " + std::to_string(fh.version) + " (this library reads " +
                                          std::to_string(DCG_STATE_VERSION) + ")");
  StateHeader h;
  std::vector<StateEntry> ents;
  std::vector<SimBase::StateSection> secs;
  uint64_t total;
  DCG_TRY(stateLayout(I, h, ents, secs, total));
  if (fh.paramsBytes != h.paramsBytes || fh.headerBytes != h.headerBytes)
    return I->fail(DCG_E_ARG, "state header does not match this library's parameter layout");
  if (fh.kind != h.kind || fh.storage != h.storage || std::memcmp(fh.size, h.size, sizeof(h.size)) != 0 ||
      fh.blockWidth != h.blockWidth || fh.levels != h.levels || fh.maxNumBlocks != h.maxNumBlocks ||
      fh.numCells != h.numCells || fh.sectionCount != h.sectionCount)
    return I->fail(DCG_E_ARG, "state was saved from a grid of another kind, size, block width, level count, "
                              "budget, storage or moist setting");
  if (bytes < h.headerBytes) return I->fail(DCG_E_ARG, "state truncated (section table)");
  std::vector<StateEntry> fe(ents.size());
  std::memcpy(fe.data(), in + sizeof(fh) + sizeof(dcg_params), fe.size() * sizeof(StateEntry));
  for (size_t i = 0; i < ents.size(); i++) {
    if (fe[i].id != ents[i].id || fe[i].elem != ents[i].elem || fe[i].bytes != ents[i].bytes ||
        fe[i].offset != ents[i].offset)
      return I->fail(DCG_E_ARG, "state section " + std::to_string(i) + " does not match this grid");
    if (fe[i].offset + fe[i].bytes > bytes) return I->fail(DCG_E_ARG, "state truncated (section data)");
  }
  {
    std::string cerr;
    const int cst = checkStateBuffer(in, bytes, cerr);
    if (cst) return I->fail(cst, cerr);
  }
  dcg_params p;
  std::memcpy(&p, in + sizeof(fh), sizeof(p));
  if (p.gx != I->P.gx || p.gy != I->P.gy || p.gz != I->P.gz || p.block_width != I->P.block_width ||
      p.max_levels != I->P.max_levels || p.moist != I->P.moist || p.storage != I->P.storage)
    return I->fail(DCG_E_ARG, "state parameters differ in what is fixed at creation");
  DCG_TRY(I->sync());  // no launch still reads or writes the arrays being replaced
  for (size_t i = 0; i < secs.size(); i++) {
    const hipError_t e = hipMemcpyAsync(secs[i].dev, in + fe[i].offset, secs[i].bytes, hipMemcpyHostToDevice, I->stream);
    if (e != hipSuccess) return I->fail(DCG_E_HIP, std::string("state copy: ") + hipGetErrorString(e));
  }
  DCG_TRY(I->sync());
  I->P = p;
  I->afterStateLoad(fh.flags);
  return DCG_OK;
}

int dcg_run_stage(dcg_sim* s, int32_t stage, int32_t arg) {
  IMPL(s);
  return I->runStage(stage, arg);
}
int dcg_cell_steps(dcg_sim* s, uint64_t* total) {
  IMPL(s);
  if (!total) return DCG_E_ARG;
  return I->cellSteps(total);
}
int dcg_set_timing(dcg_sim* s, int32_t mask) {
  IMPL(s);
  int e = I->collectTiming();
  if (e) return e;
  I->timingMask = (uint32_t)mask;
  for (int k = 0; k < DCG_NUM_KERNEL_IDS; k++) {
    I->timingMs[k] = 0;
    I->timingLaunches[k] = 0;
    I->timingBytes[k] = 0;
  }
  return DCG_OK;
}
int dcg_timing_get(dcg_sim* s, int32_t kernel, double* total_ms, uint64_t* launches) {
  IMPL(s);
  if (kernel < 0 || kernel >= DCG_NUM_KERNEL_IDS || !total_ms || !launches) return DCG_E_ARG;
  int e = I->collectTiming();
  if (e) return e;
  *total_ms = I->timingMs[kernel];
  *launches = I->timingLaunches[kernel];
  return DCG_OK;
}
int dcg_timing_bytes(dcg_sim* s, int32_t kernel, uint64_t* bytes) {
  IMPL(s);
  if (kernel < 0 || kernel >= DCG_NUM_KERNEL_IDS || !bytes) return DCG_E_ARG;
  *bytes = I->timingBytes[kernel];
  return DCG_OK;
}
const char* dcg_last_error(const dcg_sim* s) { return s ? s->impl->err.c_str() : g_create_error.c_str(); }

}  // extern "C"

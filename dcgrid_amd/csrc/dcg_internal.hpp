// dcg_internal.hpp — host-side internals of libdcgrid_amd: the simulation base class behind
// the C ABI, error handling and per-kernel event timing.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/dcgrid_amd.h"

namespace dcg {

#define DCG_HIP(expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return fail(DCG_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define DCG_TRY(expr)     \
  do {                    \
    int s_ = (expr);      \
    if (s_) return s_;    \
  } while (0)

// Level and budget plan of a DCGrid (fluid_simulation_dcgrid.cu:9-140); no device needed.
struct DCGridPlan {
  int bw, L, S;
  uint32_t B, H;
  uint32_t maxL[DCG_MAX_LEVELS], fullL[DCG_MAX_LEVELS], offs[DCG_MAX_LEVELS];
  uint32_t hsize[DCG_MAX_LEVELS], hoff[DCG_MAX_LEVELS];
  int rx[DCG_MAX_LEVELS], ry[DCG_MAX_LEVELS], rz[DCG_MAX_LEVELS];
};
int planDCGrid(int bw, const int32_t size[3], uint64_t maxBlocks, const dcg_params& p, DCGridPlan& out,
               std::string& err);
// Host copies of the block tables of a DCGrid state (the sections of dcg_save_state).
struct StateTables {
  const uint32_t* ctl;  // control words, CTL_LOADS first
  const int32_t* pos;   // 3 per block
  const uint8_t* lev;
  const uint32_t* par;
  const uint32_t* chi;  // 8 per block
  const int32_t* nd;    // 27 x 4 per block
  const uint32_t* hval;
};
int checkDCGridTables(const DCGridPlan& pl, const StateTables& t, std::string& err);

struct SimBase {
  dcg_params P{};
  std::string err;
  hipStream_t stream = nullptr;
  int device = 0;  // the GPU the handle was created on (its stream and memory)
  std::vector<void*> allocations;

  // Per-kernel timing with events on the handle's stream.
  uint32_t timingMask = 0;
  struct Pending {
    int k;
    uint32_t n;  // launches covered by the event pair
    hipEvent_t a, b;
  };
  int batchK = -1;  // kernel class currently timed as one batch (no per-launch events)
  std::vector<Pending> pending;
  std::vector<hipEvent_t> eventPool;
  double timingMs[DCG_NUM_KERNEL_IDS] = {};
  uint64_t timingLaunches[DCG_NUM_KERNEL_IDS] = {};
  uint64_t timingBytes[DCG_NUM_KERNEL_IDS] = {};  // algorithmic bytes, where the library counts them

  virtual ~SimBase();
  virtual int kind() const = 0;
  virtual int init() = 0;
  virtual int reset() = 0;
  virtual int adaptTopology() = 0;
  virtual int advectVelocity() = 0;
  virtual int project() = 0;
  virtual int projectLocal() = 0;
  virtual int advectDensity() = 0;
  virtual int debugStats(double* out) = 0;
  virtual int seedScene() = 0;
  virtual int info(dcg_info* out) = 0;
  virtual int fieldSize(int field, uint64_t* n) = 0;
  virtual int exportField(int field, float* dst, uint64_t n) = 0;
  virtual int importField(int field, const float* src, uint64_t n) = 0;
  virtual int exportTables(dcg_tables* t) = 0;
  // Active cells of one step, accumulated by dcg_step after adaptTopology: on the device
  // (countStep, part of the captured step) or on the host (countHost, per executed step).
  virtual int countStep() = 0;
  // adaptTopology() and the active-cell count of a dcg_step step (a back-end may count inside
  // its adaptation launches).
  virtual int adaptTopologyCounted() {
    const int e = adaptTopology();
    return e ? e : countStep();
  }
  virtual void countHost(uint64_t steps) {}
  virtual int cellSteps(uint64_t* out) = 0;
  virtual int runStage(int stage, int arg) = 0;
  // Projection with the level systems solved to a tolerance (extension).
  virtual int projectSolve() { return fail(DCG_E_UNSUPPORTED, "the projection solve is implemented for the DCGrid"); }
  virtual int solveReport(dcg_solve_info* out) {
    return fail(DCG_E_UNSUPPORTED, "the projection solve is implemented for the DCGrid");
  }
  int projectByMode() {
    if (P.project_mode == DCG_PROJECT_LOCAL) return projectLocal();
    if (P.project_mode == DCG_PROJECT_SOLVE) return projectSolve();
    return project();
  }
  // Device-pointer field copies (caller-owned device memory; synchronising).
  virtual int exportFieldDevice(int field, float* dst, uint64_t n) {
    return fail(DCG_E_UNSUPPORTED, "device field copies are not available for this grid");
  }
  virtual int importFieldDevice(int field, const float* src, uint64_t n) {
    return fail(DCG_E_UNSUPPORTED, "device field copies are not available for this grid");
  }
  // Differentiable rollout (extension).
  virtual int tapeForward(int steps, int checkpointEvery) = 0;
  virtual int tapeBackward(const float* gden, const float* gvel, const float* gtem, float* oden, float* ovel,
                           float* otem) = 0;

  // Ray-marched image of the fields into caller-owned device memory (float4 per pixel).
  virtual int render(const dcg_render_params& r, float* rgbaDevice) = 0;

  // Persistent state (dcg_save_state / dcg_load_state): the device arrays every later step
  // depends on, one section each, plus host flags.
  struct StateSection {
    uint32_t id;    // DCG_STATE_*
    uint32_t elem;  // bytes per element
    void* dev;      // device array
    uint64_t bytes;
  };
  virtual void stateSections(std::vector<StateSection>& out) = 0;
  virtual uint32_t stateFlags() const = 0;
  // After a state was copied in: host flags restored, caches derived from the old state dropped.
  virtual void afterStateLoad(uint32_t flags) = 0;

  // Step replay through a HIP graph: two consecutive steps are captured once (the host-side
  // pointer swaps of the advected fields return to their start after two steps) and replayed
  // while the state the launches were recorded with (graphKeyBytes) is unchanged.
  bool graphOn = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  std::vector<char> graphKey;
  virtual bool graphCapable() const { return false; }
  virtual void graphKeyBytes(std::vector<char>& out) const { out.clear(); }
  // Host-side work a step would do only once (the uniform grid's fluidity pyramid after an
  // import or a parameter change): while it is pending, steps run eagerly, since a capture
  // records the launches of one pass but replays them every time.
  virtual bool captureUnsettled() const { return false; }
  // Host state a step changes (field-buffer swaps, caches), restored when a capture fails.
  virtual void hostSnapshot(std::vector<char>& out) const { graphKeyBytes(out); }
  virtual void hostRestore(const std::vector<char>&) {}
  void dropGraph() {
    if (graphExec) hipGraphExecDestroy(graphExec);
    if (graph) hipGraphDestroy(graph);
    graphExec = nullptr;
    graph = nullptr;
    graphKey.clear();
  }
  int stepOnce() {
    int e;
    if ((e = advectVelocity())) return e;
    if ((e = adaptTopologyCounted())) return e;
    if ((e = projectByMode())) return e;
    return advectDensity();
  }
  int steps(int n);

  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
  int createStream();
  // Device allocation owned by the handle (freed in the destructor).
  template <class T>
  int alloc(T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(DCG_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocations.push_back(q);
    *p = (T*)q;
    return DCG_OK;
  }
  // Pinned host allocation owned by the handle (hipHostMalloc: coherent, mapped into the device
  // address space; freed in the destructor).
  std::vector<void*> hostAllocations;
  template <class T>
  int allocHost(T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipHostMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 16, hipHostMallocDefault);
    if (e != hipSuccess) return fail(DCG_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    hostAllocations.push_back(q);
    *p = (T*)q;
    return DCG_OK;
  }
  // Free one allocation made by alloc().
  void release(void* p) {
    if (!p) return;
    for (size_t i = 0; i < allocations.size(); i++)
      if (allocations[i] == p) {
        hipFree(p);
        allocations.erase(allocations.begin() + i);
        return;
      }
  }
  // Fill with a kernel (never hipMemsetAsync, so the step stays graph-capturable).
  int fill32(void* p, uint32_t value, size_t count);
  int fill8(void* p, uint8_t value, size_t count);
  int sync();

  // Launch bracket: records events around the launch when kernel id k is timed.
  int beginTimed(int k, Pending* ev);
  // Events of a launch timed by the dispatch's own begin and end (hipExtLaunchKernelGGL): the
  // kernel duration rocprofv3 reports, without the gap between dispatches that events recorded
  // around a launch include. takeEvents gives the pair (ev->k = -1 when class k is not timed),
  // pushPending files it after the launch.
  int takeEvents(int k, Pending* ev);
  int pushPending(Pending* ev);
  int endTimed(int k, Pending* ev, uint32_t launches = 1);
  int collectTiming();
  int checkLaunch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(DCG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return DCG_OK;
  }
};

// Launch `expr` on the handle's stream, bracketed by timing events for kernel id K (unless a
// batch of K launches is being timed as a whole).
#define DCG_LAUNCH(K, NAME, expr)        \
  do {                                   \
    SimBase::Pending ev_;                \
    DCG_TRY(beginTimed((K), &ev_));      \
    expr;                                \
    DCG_TRY(checkLaunch(NAME));          \
    DCG_TRY(endTimed((K), &ev_));        \
  } while (0)

SimBase* makeUniform(const int32_t size[3], const dcg_params& p, int* status, std::string* err);
SimBase* makeDCGrid(const int32_t size[3], uint64_t maxBlocks, const dcg_params& p, int* status,
                    std::string* err);

}  // namespace dcg

// Host runtime of libargot_hip.so: the reference's GPU entry points
// (src/prototype.h:86-130) on one MI355X, the fused step_radiation_tree,
// and the library controls declared in include/argot_hip.h.
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "argot_device.h"
#include "argot_hip.h"
#include "argot_internal.h"
#include "kernels.h"
#include "surrogate.h"
#include "tree.h"

using namespace argot;

// Device scratch of argot_hip_calc_optical_depth_noalloc: the device copy of
// the caller's segment list, 56 B per segment, at most kScratchSegments
// (56 MiB).  Each GPU holds its own copy of the array from the library's
// load to the end of the process.  Host code reaches it only through the
// address hipGetSymbolAddress returns on that device (scratch_on), never
// through the array's name, which in host code is not a device address.
constexpr int64_t kScratchSegments = (int64_t)1 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchSegments * 56];
static_assert(sizeof(struct ray_segment) == 56, "ray_segment is 56 bytes");

namespace {

thread_local int t_err = ARGOT_OK;

void set_err(int e) {
  if (t_err == ARGOT_OK) t_err = e;
}

thread_local char t_detail[256] = "";

void set_hip_err(hipError_t e, int line) {
  if (t_err == ARGOT_OK)
    snprintf(t_detail, sizeof t_detail, "runtime.hip:%d: %s (%d)", line, hipGetErrorString(e),
             (int)e);
  set_err(e == hipErrorOutOfMemory ? ARGOT_ERR_NOMEM : ARGOT_ERR_HIP);
}

#define HIP_OK(expr)                      \
  do {                                    \
    hipError_t _e = (expr);               \
    if (_e != hipSuccess) {               \
      set_hip_err(_e, __LINE__);          \
      return;                             \
    }                                     \
  } while (0)

// The scratch array's address on each device, resolved once per device
// (init_gpu, with that device current).  g_scratch_use serialises the calls
// that use a device's array: each call copies its segments in, walks them
// and copies them out before it returns, so the lock covers its whole use.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_use[kMaxDevices];

hipError_t scratch_on(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

// Options (library controls).
double g_helium_fact = 0.0811258278;   // chemistry.h:25 (Y = 0.245)
int g_tau_mode = 0;
int g_rate_mode = 0;
float g_theta = 0.7f;                  // ARGOT_THETA_CRIT, run_param.h:80
int g_blocks_per_cu = 24;
int g_tile_bits[3] = {2, 2, 2};        // 4x4x4 targets per wavefront
int g_use_graph = 1;
int g_iterations = 0;                  // 0: one point-source pass
int g_wide_index = 0;                  // 1: 64-bit cell-field addresses on any mesh
int g_chem_queue = 1;                  // cell-queue chemistry: free lanes refill at this count
int g_host_threads = 16;               // host threads of calc_photoion_rate_cuda (0: serial)

// Everything a captured step graph depends on.
struct GraphKey {
  FusedArgs args;
  int nblock, kahan, surr, wide;
  int64_t ncell;
  double helium_fact;
  const float4* nf;
};
RayRecord* g_capture_host = nullptr;
int64_t g_capture_cap = 0;

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;  // elements
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  hipError_t reserve(size_t n) {
    if (n <= cap && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    size_t want = std::max<size_t>(n, 1);
    hipError_t e = hipMalloc(&p, want * sizeof(T));
    if (e == hipSuccess) cap = want;
    return e;
  }
};

// Everything one mesh needs on the device: its own stream, events, field
// and tree buffers, staging and the captured step graph.  init_gpu creates
// one; the caller's cuda_mem_space (its mesh_dev) names it in every later
// call, so meshes and streams of one process do not share state and no
// entry point depends on the calling thread's state.
struct Ctx {
  bool ready = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_mid = nullptr;
  int ncu = 256;
  int nx = 0, ny = 0, nz = 0;
  int64_t ncell = 0;
  int dev = 0;
  void* scratch = nullptr;         // this device's scratch array (scratch_on)
  unsigned char* mesh = nullptr;   // cuda_mem->mesh_dev
  DevBuf<float4> nf;
  // frequency tables
  DevBuf<float> nu, cs, sig, hsig, dsig;
  double cs_max[3] = {0, 0, 0}, cs_last[3] = {0, 0, 0};
  int ngrid = 0, k_he1 = 0, k_he2 = 0;
  // tree
  DevBuf<NodeLink> link;
  DevBuf<float> cm, spec, src_pos;
  DevBuf<double> spec_scale;
  DevBuf<StepCounters> ctr;
  DevBuf<double> node_q;          // surrogate: photon rate of conforming nodes
  DevBuf<RayRecord> rec;
  // segment / ray staging for the per-call entry points
  DevBuf<unsigned char> seg, rays;
  DevBuf<float> rate_tab;
  DevBuf<float> smooth_in, smooth_out;   // smooth_photoion_rate planes
  DevBuf<ChemCounters> chem_ctr;
  hipEvent_t ev_c0 = nullptr, ev_c1 = nullptr;
  argot_chem_stats chem_stats{};
  HostTree tree;
  argot_step_stats stats{};
  // step graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  GraphKey graph_key{};
  int64_t stats_graph_captures = 0;
  // the caller-visible device buffers init_gpu allocated
  void* run_dev = nullptr;
  void* rate_dev = nullptr;

  ~Ctx() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamSynchronize(stream);
    nf.release();
    nu.release();
    cs.release();
    sig.release();
    hsig.release();
    dsig.release();
    link.release();
    cm.release();
    spec.release();
    src_pos.release();
    spec_scale.release();
    ctr.release();
    node_q.release();
    rec.release();
    seg.release();
    rays.release();
    rate_tab.release();
    smooth_in.release();
    smooth_out.release();
    chem_ctr.release();
    for (hipEvent_t e : {ev0, ev1, ev_mid, ev_c0, ev_c1})
      if (e) (void)hipEventDestroy(e);
    if (mesh) (void)hipFree(mesh);
    if (run_dev) (void)hipFree(run_dev);
    if (rate_dev) (void)hipFree(rate_dev);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Live contexts, looked up by the device mesh their init_gpu allocated.
std::mutex g_ctx_mu;
std::vector<std::unique_ptr<Ctx>> g_ctx;
// The context of this thread's last call (for the statistics queries).
thread_local Ctx* t_last = nullptr;

Ctx* find_ctx(const cuda_mem_space* cm) {
  if (!cm || !cm->mesh_dev) return nullptr;
  const unsigned char* key = reinterpret_cast<const unsigned char*>(cm->mesh_dev);
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  for (auto& c : g_ctx)
    if (c->mesh == key) return c.get();
  return nullptr;
}

// Removes and frees the context whose device mesh is mesh_dev, if any.
void drop_ctx(const void* mesh_dev) {
  std::unique_ptr<Ctx> gone;
  {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    for (auto it = g_ctx.begin(); it != g_ctx.end(); ++it)
      if ((*it)->mesh == mesh_dev) {
        gone = std::move(*it);
        g_ctx.erase(it);
        break;
      }
  }
  if (gone && t_last == gone.get()) t_last = nullptr;
}

// This thread's last context, if it still exists (another thread may have
// released it since).
const Ctx* last_ctx() {
  if (!t_last) return nullptr;
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  for (auto& c : g_ctx)
    if (c.get() == t_last) return t_last;
  return nullptr;
}

GridDev make_grid(const run_param* r) {
  GridDev d{};   // zeroed padding: launch arguments are compared as bytes
  d.nx = r->nmesh_x_local;
  d.ny = r->nmesh_y_local;
  d.nz = r->nmesh_z_local;
  d.nyz = (int64_t)d.ny * d.nz;
  d.xmin = r->xmin_local;
  d.ymin = r->ymin_local;
  d.zmin = r->zmin_local;
  d.xmax = r->xmax_local;
  d.ymax = r->ymax_local;
  d.zmax = r->zmax_local;
  d.gxmin = r->xmin;
  d.gymin = r->ymin;
  d.gzmin = r->zmin;
  d.ox = r->rank_x * r->nmesh_x_local;
  d.oy = r->rank_y * r->nmesh_y_local;
  d.oz = r->rank_z * r->nmesh_z_local;
  d.dx = r->delta_x;
  d.dy = r->delta_y;
  d.dz = r->delta_z;
  d.lunit = r->lunit;
  d.denstonh = r->denstonh;
  d.helium_fact = g_helium_fact;
  const float th = r->theta_crit;
  d.theta_sq = th * th;
  return d;
}

RateTables rate_tables(const Ctx& g) {
  RateTables t{};
  t.ngrid = g.ngrid;
  t.nu = g.nu.p;
  t.cs = g.cs.p;
  t.sig = g.sig.p;
  t.hsig = g.hsig.p;
  t.dsig = g.dsig.p;
  t.k_he1 = g.k_he1;
  t.k_he2 = g.k_he2;
  for (int s = 0; s < 3; s++) {
    t.cs_max[s] = g.cs_max[s];
    t.cs_last[s] = g.cs_last[s];
  }
  return t;
}

// Tables of run_param (freq.nu, csect) in the layouts the kernels read.
void upload_tables(Ctx& g, const run_param* r) {
  const int ng = ARGOT_NGRID_NU;
  std::vector<float> nu(ng), cs(3 * ng), sig(3 * ng), hsig(3 * ng), dsig(3 * ng);
  const double thr[3] = {1.0, 1.8125, 4.0};   // HI, HeI, HeII limits [nu_L]
  for (int k = 0; k < ng; k++) {
    nu[k] = (float)r->freq.nu[k];
    const float c[3] = {r->csect[k].csect_HI, r->csect[k].csect_HeI, r->csect[k].csect_HeII};
    for (int s = 0; s < 3; s++) {
      cs[3 * k + s] = c[s];
      sig[s * ng + k] = c[s];
      hsig[s * ng + k] =
          (double)nu[k] > thr[s] ? (float)((double)c[s] * ((double)nu[k] - thr[s])) : 0.0f;
    }
  }
  for (int s = 0; s < 3; s++) {
    double mx = 0.0;
    for (int k = 0; k < ng; k++) mx = std::max(mx, (double)sig[s * ng + k]);
    g.cs_max[s] = mx;
    g.cs_last[s] = sig[s * ng + ng - 1];
    for (int k = 0; k < ng; k++)
      dsig[s * ng + k] = (float)((double)sig[s * ng + k] - g.cs_last[s]);
  }
  // Bin ranges with absent species (sigma and heating weight both zero).
  auto first_nonzero = [&](int s) {
    for (int k = 0; k < ng; k++)
      if (sig[s * ng + k] != 0.0f || hsig[s * ng + k] != 0.0f) return k;
    return ng;
  };
  g.k_he2 = first_nonzero(2);
  g.k_he1 = std::min(first_nonzero(1), g.k_he2);
  g.ngrid = ng;
  HIP_OK(g.nu.reserve(ng));
  HIP_OK(g.cs.reserve(3 * ng));
  HIP_OK(g.sig.reserve(3 * ng));
  HIP_OK(g.hsig.reserve(3 * ng));
  HIP_OK(g.dsig.reserve(3 * ng));
  HIP_OK(hipMemcpy(g.nu.p, nu.data(), ng * sizeof(float), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(g.cs.p, cs.data(), 3 * ng * sizeof(float), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(g.sig.p, sig.data(), 3 * ng * sizeof(float), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(g.hsig.p, hsig.data(), 3 * ng * sizeof(float), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(g.dsig.p, dsig.data(), 3 * ng * sizeof(float), hipMemcpyHostToDevice));
}

// The caller's context, ready and for a mesh of this_run's size, or null
// (with the error set).
Ctx* check_ready(const cuda_mem_space* cm, const run_param* r) {
  Ctx* c = find_ctx(cm);
  if (!c || !c->ready || !r) {
    set_err(ARGOT_ERR_STATE);
    return nullptr;
  }
  t_last = c;
  if (r->nmesh_x_local != c->nx || r->nmesh_y_local != c->ny || r->nmesh_z_local != c->nz) {
    set_err(ARGOT_ERR_ARG);
    return nullptr;
  }
  return c;
}

void pack_fields(Ctx& g) {
  HIP_OK(g.nf.reserve(g.ncell));
  HIP_OK(launch_pack_fields(g.mesh, g.ncell, g_helium_fact, g.nf.p, g.stream));
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

namespace argot {
void argot_hip_internal_set_error(int code) { set_err(code); }
void argot_hip_internal_clear_error() { t_err = ARGOT_OK; }
double argot_hip_internal_helium_fact() { return g_helium_fact; }
}  // namespace argot

extern "C" {

int argot_hip_last_error(void) { return t_err; }

const char* argot_hip_last_error_detail(void) { return t_err == ARGOT_OK ? "" : t_detail; }

const char* argot_hip_error_string(int code) {
  switch (code) {
    case ARGOT_OK: return "ok";
    case ARGOT_ERR_HIP: return "HIP runtime call failed";
    case ARGOT_ERR_ARG: return "bad argument or size";
    case ARGOT_ERR_STATE: return "init_gpu / send_mesh_data not called, or no trained surrogate";
    case ARGOT_ERR_NOMEM: return "device or host allocation failed";
    case ARGOT_ERR_TOO_MANY_SEGMENTS: return "ray crosses too many domains";
    case ARGOT_ERR_IO: return "file could not be opened or is short";
    default: return "unknown";
  }
}

void argot_hip_set_helium_fact(double helium_fact) { g_helium_fact = helium_fact; }
void argot_hip_set_tau_mode(int mode) { g_tau_mode = mode ? 1 : 0; }
void argot_hip_set_rate_mode(int mode) { g_rate_mode = mode ? 1 : 0; }
// Statistics of the last call on this thread (its context's).
void argot_hip_step_stats(struct argot_step_stats* out) {
  const Ctx* c = last_ctx();
  if (out) *out = c ? c->stats : argot_step_stats{};
}

// Opening angle used by step_radiation_tree (the reference always sets
// ARGOT_THETA_CRIT = 0.7, src/step_radiation_tree.c:594).
void argot_hip_set_theta(float theta) { g_theta = theta; }

// Persistent grid size of the fused kernel, in wavefronts per CU.
void argot_hip_set_blocks_per_cu(int n) { g_blocks_per_cu = n > 0 ? n : 24; }

void argot_hip_set_tile(int tx, int ty, int tz) {
  t_err = ARGOT_OK;
  auto log2i = [](int v) {
    int b = 0;
    while ((1 << b) < v) b++;
    return (v > 0 && (1 << b) == v) ? b : -1;
  };
  const int b[3] = {log2i(tx), log2i(ty), log2i(tz)};
  if (b[0] < 0 || b[1] < 0 || b[2] < 0 || b[0] + b[1] + b[2] != kTileBits)
    return set_err(ARGOT_ERR_ARG);
  for (int d = 0; d < 3; d++) g_tile_bits[d] = b[d];
}

void argot_hip_set_graph(int on) { g_use_graph = on ? 1 : 0; }

int64_t argot_hip_graph_captures(void) {
  const Ctx* c = last_ctx();
  return c ? c->stats_graph_captures : 0;
}

// Test support: the next step_radiation_tree records every ray it forms
// (struct RayRecord, 120 B) into host_records[0..cap).
void argot_hip_set_capture(void* host_records, int64_t cap) {
  g_capture_host = static_cast<RayRecord*>(host_records);
  g_capture_cap = host_records ? cap : 0;
}

int64_t argot_hip_record_size(void) { return (int64_t)sizeof(RayRecord); }

static void init_gpu_body(struct fluid_mesh* mesh, struct cuda_mem_space* cuda_mem,
                          struct cuda_param* this_cuda, struct run_param* this_run) {
  (void)mesh;
  if (!cuda_mem || !this_cuda || !this_run) return set_err(ARGOT_ERR_ARG);
  // Mesh sizes the kernels' index arithmetic covers: ix*ny + iy in 24 bits
  // (24-bit multiply-adds form the cell index), the cell count in an int
  // (cuda_param.nmesh_per_dev) and so in the 32-bit linear cell index.
  {
    const int64_t nx = this_run->nmesh_x_local, ny = this_run->nmesh_y_local,
                  nz = this_run->nmesh_z_local;
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx * ny >= (int64_t)1 << 24 ||
        nx * ny * nz > (int64_t)INT32_MAX)
      return set_err(ARGOT_ERR_ARG);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(ARGOT_ERR_HIP);
  HIP_OK(hipSetDevice(0));
  hipDeviceProp_t prop;
  HIP_OK(hipGetDeviceProperties(&prop, 0));
  // A cuda_mem that already names a context is being initialised again:
  // that context (and its device mesh) goes.
  if (find_ctx(cuda_mem)) drop_ctx(cuda_mem->mesh_dev);
  std::unique_ptr<Ctx> owned(new Ctx());
  Ctx& g = *owned;
  g.ncu = prop.multiProcessorCount;
  g.dev = 0;
  HIP_OK(scratch_on(g.dev, &g.scratch));
  g.nx = this_run->nmesh_x_local;
  g.ny = this_run->nmesh_y_local;
  g.nz = this_run->nmesh_z_local;
  g.ncell = (int64_t)g.nx * g.ny * g.nz;
  memset(this_cuda, 0, sizeof(*this_cuda));
  this_cuda->num_cuda_dev = 1;
  this_cuda->nmesh_per_dev = (int)g.ncell;
  this_cuda->cuda_nblock = (int)(g.ncell / 256);
  this_cuda->max_thread_dimx = prop.maxThreadsDim[0];
  this_cuda->max_block_dimx = prop.maxGridSize[0];
  memset(cuda_mem, 0, sizeof(*cuda_mem));
  // (owned frees what is allocated here if a later step fails)
  HIP_OK(hipMalloc(&g.mesh, sizeof(struct fluid_mesh) * g.ncell));
  HIP_OK(hipMalloc(&g.run_dev, sizeof(struct run_param)));
  HIP_OK(hipMalloc(&g.rate_dev, sizeof(struct photoion_rate) * g.ncell));
  HIP_OK(hipMemcpy(g.run_dev, this_run, sizeof(struct run_param), hipMemcpyHostToDevice));
  HIP_OK(hipStreamCreate(&g.stream));
  HIP_OK(hipEventCreate(&g.ev0));
  HIP_OK(hipEventCreate(&g.ev1));
  HIP_OK(hipEventCreate(&g.ev_mid));
  HIP_OK(hipEventCreate(&g.ev_c0));
  HIP_OK(hipEventCreate(&g.ev_c1));
  upload_tables(g, this_run);
  if (t_err) return;
  HIP_OK(g.ctr.reserve(1));
  HIP_OK(g.chem_ctr.reserve(1));
  g.ready = true;
  cuda_mem->mesh_dev = reinterpret_cast<struct fluid_mesh*>(g.mesh);
  cuda_mem->this_run_dev = static_cast<struct run_param*>(g.run_dev);
  cuda_mem->rate_dev = static_cast<struct photoion_rate*>(g.rate_dev);
  this_cuda->strm[0] = g.stream;
  this_cuda->diffuse_strm[0] = g.stream;
  {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    g_ctx.push_back(std::move(owned));
  }
  t_last = &g;
}

void init_gpu(struct fluid_mesh* mesh, struct cuda_mem_space* cuda_mem,
              struct cuda_param* this_cuda, struct run_param* this_run) {
  t_err = ARGOT_OK;
  try {
    init_gpu_body(mesh, cuda_mem, this_cuda, this_run);
  } catch (const std::bad_alloc&) {
    set_err(ARGOT_ERR_NOMEM);
  } catch (...) {
    set_err(ARGOT_ERR_ARG);
  }
}

// Frees the context init_gpu made for cuda_mem: its device mesh, run_param
// and rate table, stream, events and buffers; cuda_mem is zeroed.  Returns
// the status instead of setting the thread's last error, so that a release
// run from a destructor cannot hide another call's error.
int argot_hip_release_gpu(struct cuda_mem_space* cuda_mem) {
  if (!cuda_mem) return ARGOT_ERR_ARG;
  if (!find_ctx(cuda_mem)) return ARGOT_ERR_STATE;
  drop_ctx(cuda_mem->mesh_dev);
  memset(cuda_mem, 0, sizeof(*cuda_mem));
  return ARGOT_OK;
}

void send_mesh_data(struct fluid_mesh* mesh, struct cuda_mem_space* cuda_mem,
                    struct cuda_param* this_cuda, struct run_param* this_run) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  if (!mesh) return set_err(ARGOT_ERR_ARG);
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  HIP_OK(hipMemcpyAsync(cuda_mem->mesh_dev, mesh, sizeof(struct fluid_mesh) * g.ncell,
                        hipMemcpyHostToDevice, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
}

void recv_mesh_data(struct fluid_mesh* mesh, struct cuda_mem_space* cuda_mem,
                    struct cuda_param* this_cuda, struct run_param* this_run) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  if (!mesh) return set_err(ARGOT_ERR_ARG);
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  HIP_OK(hipMemcpyAsync(mesh, cuda_mem->mesh_dev, sizeof(struct fluid_mesh) * g.ncell,
                        hipMemcpyDeviceToHost, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
}

void send_run_param_data(struct run_param* this_run, struct cuda_mem_space* cuda_mem,
                         struct cuda_param* this_cuda) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  Ctx* const cp = find_ctx(cuda_mem);
  if (!cp || !cp->ready || !this_run) return set_err(ARGOT_ERR_STATE);
  Ctx& g = *cp;
  t_last = cp;
  HIP_OK(hipMemcpy(g.run_dev, this_run, sizeof(struct run_param), hipMemcpyHostToDevice));
  upload_tables(g, this_run);
}

void allocate_pinned_segment(struct ray_segment** seg, uint64_t nseg) {
  t_err = ARGOT_OK;
  if (hipHostMalloc((void**)seg, sizeof(struct ray_segment) * std::max<uint64_t>(nseg, 1),
                    hipHostMallocDefault) != hipSuccess) {
    *seg = nullptr;
    set_err(ARGOT_ERR_NOMEM);
  }
}

void deallocate_pinned_segment(struct ray_segment* seg) {
  if (seg) (void)hipHostFree(seg);
}

void allocate_pinned_light_ray_IO(struct light_ray_IO** ray_IO, uint64_t nray) {
  t_err = ARGOT_OK;
  if (hipHostMalloc((void**)ray_IO, sizeof(struct light_ray_IO) * std::max<uint64_t>(nray, 1),
                    hipHostMallocDefault) != hipSuccess) {
    *ray_IO = nullptr;
    set_err(ARGOT_ERR_NOMEM);
  }
}

void deallocate_pinned_light_ray_IO(struct light_ray_IO* ray_IO) {
  if (ray_IO) (void)hipHostFree(ray_IO);
}

void calc_optical_depth_cuda(struct ray_segment* seg, struct cuda_mem_space* cuda_mem,
                             struct cuda_param* this_cuda, struct run_param* this_run) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  const int64_t nseg = (int64_t)this_run->nseg;
  if (nseg <= 0) return;
  if (!seg) return set_err(ARGOT_ERR_ARG);
  pack_fields(g);
  if (t_err) return;
  const size_t bytes = sizeof(struct ray_segment) * nseg;
  HIP_OK(g.seg.reserve(bytes));
  HIP_OK(hipMemcpyAsync(g.seg.p, seg, bytes, hipMemcpyHostToDevice, g.stream));
  const GridDev gd = make_grid(this_run);
  HIP_OK(hipEventRecord(g.ev_c0, g.stream));
  HIP_OK(launch_walk_segments(gd, g.nf.p, g.seg.p, nseg, g_tau_mode, g_wide_index, g.stream));
  HIP_OK(hipEventRecord(g.ev_c1, g.stream));
  HIP_OK(hipMemcpyAsync(seg, g.seg.p, bytes, hipMemcpyDeviceToHost, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, g.ev_c0, g.ev_c1));
  g.stats.ms_walk = ms;
}

// calc_optical_depth_cuda without an allocation: the segments are staged in
// this device's scratch array and the walk forms the densities from the
// mesh records (the packed-field buffer is never needed), so the call makes
// no hipMalloc, hipFree or other allocating call.
void argot_hip_calc_optical_depth_noalloc(struct ray_segment* seg, struct cuda_mem_space* cuda_mem,
                                          struct cuda_param* this_cuda,
                                          struct run_param* this_run) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  const int64_t nseg = (int64_t)this_run->nseg;
  if (nseg <= 0) return;
  if (!seg || nseg > kScratchSegments) return set_err(ARGOT_ERR_ARG);
  if (!g.scratch) return set_err(ARGOT_ERR_STATE);
  std::lock_guard<std::mutex> lk(g_scratch_use[g.dev]);
  const size_t bytes = sizeof(struct ray_segment) * nseg;
  const GridDev gd = make_grid(this_run);
  hipError_t e = hipMemcpyAsync(g.scratch, seg, bytes, hipMemcpyHostToDevice, g.stream);
  if (e == hipSuccess) e = hipEventRecord(g.ev_c0, g.stream);
  if (e == hipSuccess)
    e = launch_walk_segments_mesh(gd, g.mesh, g_helium_fact, g.scratch, nseg, g_tau_mode,
                                  g.stream);
  if (e == hipSuccess) e = hipEventRecord(g.ev_c1, g.stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(seg, g.scratch, bytes, hipMemcpyDeviceToHost, g.stream);
  // the stream is drained on every path, so that no work on the array
  // outlives the lock
  const hipError_t es = hipStreamSynchronize(g.stream);
  if (e == hipSuccess) e = es;
  HIP_OK(e);
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, g.ev_c0, g.ev_c1));
  g.stats.ms_walk = ms;
}

// Distinct spectra of a ray list: rays from one source or tree node carry
// equal photon_rate rows.  Candidates are found by three bins of the row and
// confirmed by comparing the whole row, so only equal rows share an index;
// indices follow the order of first appearance.
struct SpectrumKey {
  uint64_t a, b, c;
  bool operator==(const SpectrumKey& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct SpectrumKeyHash {
  size_t operator()(const SpectrumKey& k) const {
    uint64_t h = k.a * 0x9e3779b97f4a7c15ull;
    h ^= k.b + 0x7f4a7c159e3779b9ull + (h << 6) + (h >> 2);
    h ^= k.c + 0x94d049bb133111ebull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

// Host threads for a pass over n rays: one per 4096 rays, at most 16 and
// at most the hardware's.
static int host_threads(int64_t n) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const unsigned cap = (unsigned)std::max(1, g_host_threads);
  return (int)std::min<int64_t>(std::min(hw, cap), std::max<int64_t>(1, n / 4096));
}

// Runs f(0 .. nt-1), f(0) on the calling thread and the others on threads
// of their own, and joins them all.  No exception leaves a thread: a share
// that throws counts as failed.  A share whose thread cannot be started
// runs on the calling thread.  Returns whether every f(t) returned true.
extern "C++" template <typename F>
static bool run_on_threads(int nt, const F& f) {
  std::vector<unsigned char> ok(nt, 0);
  auto share = [&](int t) {
    try {
      ok[t] = f(t) ? 1 : 0;
    } catch (...) {
      ok[t] = 0;
    }
  };
  std::vector<std::thread> th;
  int started = 1;
  try {
    th.reserve(nt);
    for (; started < nt; started++) th.emplace_back(share, started);
  } catch (...) {
  }
  for (int t = started; t < nt; t++) share(t);
  share(0);
  for (auto& x : th) x.join();
  return std::all_of(ok.begin(), ok.end(), [](unsigned char v) { return v != 0; });
}

// 64-bit hash of a spectrum row (four independent multiply-xor chains over
// the row's 64-bit words, then mixed).
static uint64_t row_hash(const double* row, int n) {
  uint64_t h[4] = {0x9e3779b97f4a7c15ull, 0xbf58476d1ce4e5b9ull, 0x94d049bb133111ebull,
                   0xd6e8feb86659fd93ull};
  for (int k = 0; k < n; k++) {
    uint64_t v;
    memcpy(&v, row + k, 8);
    uint64_t& a = h[k & 3];
    a = (a ^ v) * 0xff51afd7ed558ccdull;
    a ^= a >> 29;
  }
  uint64_t r = h[0];
  for (int q = 1; q < 4; q++) r = (r ^ h[q]) * 0xc4ceb9fe1a85ec53ull + (r >> 31);
  return r;
}

// Fills the compact ray records cr[j] of the caller's rays j, their target
// cells tgt[j] (cell(ray) < 0: outside the local mesh, *bad is set), and
// numbers the distinct spectrum rows by first appearance, on up to 16 host
// threads over contiguous ranges of j, in one pass over the records: each
// thread hashes its rows (whole rows), keeps the first ray of each hash
// and checks every later row of that hash against it byte for byte; the
// ranges' first rays are then merged in order (checked byte for byte
// against the row already numbered).  Returns false if two different rows
// share a hash or a thread failed (the caller then fills the records and
// numbers the spectra serially).
extern "C++" template <typename CellF>
static bool number_spectra(const light_ray* ray, int64_t nray, int ng, const CellF& cell,
                           int64_t* tgt, std::atomic<bool>* bad, CallRay* cr,
                           std::vector<const double*>* rows) {
  const int nt = host_threads(nray);
  std::vector<uint64_t> hs(nray);
  struct Range {
    std::vector<uint64_t> h;        // hashes in order of first appearance
    std::vector<const double*> r;   // their first rows
    bool clash = false;
  };
  std::vector<Range> rg(nt);
  auto work = [&](int t) -> bool {
    const int64_t j0 = nray * t / nt, j1 = nray * (t + 1) / nt;
    std::unordered_map<uint64_t, int32_t> seen;
    Range& R = rg[t];
    for (int64_t j = j0; j < j1; j++) {
      const light_ray& L = ray[j];
      const int64_t c = cell(L);
      if (c < 0) {
        bad->store(true);
        return false;
      }
      tgt[j] = c;
      const double* row = L.src.photon_rate;
      const uint64_t h = row_hash(row, ng);
      hs[j] = h;
      auto it = seen.find(h);
      if (it == seen.end()) {
        seen.emplace(h, (int32_t)R.h.size());
        R.h.push_back(h);
        R.r.push_back(row);
      } else if (memcmp(R.r[it->second], row, sizeof(double) * ng) != 0) {
        R.clash = true;
      }
      CallRay& o = cr[j];
      o.src[0] = L.src.xpos;
      o.src[1] = L.src.ypos;
      o.src[2] = L.src.zpos;
      o.tau[0] = L.optical_depth_HI;
      o.tau[1] = L.optical_depth_HeI;
      o.tau[2] = L.optical_depth_HeII;
      o.pad = 0;
    }
    return true;
  };
  if (!run_on_threads(nt, work)) return false;
  std::unordered_map<uint64_t, int32_t> index;
  rows->clear();
  for (int t = 0; t < nt; t++) {
    if (rg[t].clash) return false;
    for (size_t q = 0; q < rg[t].h.size(); q++) {
      auto it = index.find(rg[t].h[q]);
      if (it == index.end()) {
        index.emplace(rg[t].h[q], (int32_t)rows->size());
        rows->push_back(rg[t].r[q]);
      } else if (memcmp((*rows)[it->second], rg[t].r[q], sizeof(double) * ng) != 0) {
        return false;
      }
    }
  }
  auto label = [&](int t) -> bool {
    const int64_t j0 = nray * t / nt, j1 = nray * (t + 1) / nt;
    for (int64_t j = j0; j < j1; j++) cr[j].spec = index.find(hs[j])->second;
    return true;
  };
  return run_on_threads(nt, label);
}

static void calc_photoion_rate_cuda_body(struct cuda_mem_space* cuda_mem, struct light_ray* ray,
                                         struct light_ray_IO* ray_IO, struct cuda_param* this_cuda,
                                         struct run_param* this_run) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  const int64_t nray = (int64_t)this_run->nray;
  if (nray <= 0) return;
  if (!ray || !ray_IO) return set_err(ARGOT_ERR_ARG);
  if (nray > (int64_t)INT32_MAX) return set_err(ARGOT_ERR_ARG);
  const GridDev gd = make_grid(this_run);
  const int ng = ARGOT_NGRID_NU;
  // Each target's rays, together and in the caller's order
  // (calc_photoion_rate_cuda.cu:108-132 repacks them; the per-target sums
  // then need no atomics).  One pass over the caller's records gives the
  // compact ray records, the target cells and the spectrum numbering.
  auto cell = [&](const light_ray& L) -> int64_t {
    const int64_t lx = L.ix_target - gd.ox, ly = L.iy_target - gd.oy, lz = L.iz_target - gd.oz;
    if (lx < 0 || lx >= g.nx || ly < 0 || ly >= g.ny || lz < 0 || lz >= g.nz) return -1;
    return lz + (int64_t)g.nz * (ly + (int64_t)g.ny * lx);
  };
  // Staging in the caller's pinned ray_IO buffer (nray * 2096 bytes): the
  // ray records, then the target records, then the spectrum tables when
  // they fit; the tables go from host memory otherwise.
  const size_t cap = sizeof(struct light_ray_IO) * (size_t)nray;
  unsigned char* stage = reinterpret_cast<unsigned char*>(ray_IO);
  CallRay* cr = reinterpret_cast<CallRay*>(stage);
  CallTarget* ct = reinterpret_cast<CallTarget*>(stage + sizeof(CallRay) * nray);
  int64_t ntarget = 0;
  std::vector<int64_t> tgt(nray);
  std::vector<const double*> rows;   // first ray's row of each distinct spectrum
  std::atomic<bool> bad(false);
  if (g_host_threads == 0 || !number_spectra(ray, nray, ng, cell, tgt.data(), &bad, cr, &rows)) {
    if (bad.load()) return set_err(ARGOT_ERR_ARG);
    // two different rows with one hash, or a host thread that failed:
    // the records and the numbering again, serially, candidates found by
    // three bins and confirmed on the whole row
    std::unordered_map<SpectrumKey, std::vector<int32_t>, SpectrumKeyHash> index;
    rows.clear();
    for (int64_t j = 0; j < nray; j++) {
      const light_ray& L = ray[j];
      tgt[j] = cell(L);
      if (tgt[j] < 0) return set_err(ARGOT_ERR_ARG);
      CallRay& o = cr[j];
      o.src[0] = L.src.xpos;
      o.src[1] = L.src.ypos;
      o.src[2] = L.src.zpos;
      o.tau[0] = L.optical_depth_HI;
      o.tau[1] = L.optical_depth_HeI;
      o.tau[2] = L.optical_depth_HeII;
      o.pad = 0;
      const double* row = L.src.photon_rate;
      SpectrumKey key;
      memcpy(&key.a, &row[0], 8);
      memcpy(&key.b, &row[ng / 2], 8);
      memcpy(&key.c, &row[ng - 1], 8);
      std::vector<int32_t>& cand = index[key];
      int32_t sp = -1;
      for (int32_t c : cand)
        if (memcmp(rows[c], row, sizeof(double) * ng) == 0) {
          sp = c;
          break;
        }
      if (sp < 0) {
        sp = (int32_t)rows.size();
        rows.push_back(row);
        cand.push_back(sp);
      }
      cr[j].spec = sp;
    }
  }
  bool contiguous = true;
  {
    std::vector<unsigned char> seen(g.ncell, 0);
    for (int64_t i = 0; i < nray && contiguous; i++) {
      if (i > 0 && tgt[i] == tgt[i - 1]) continue;
      if (seen[tgt[i]]) contiguous = false;
      seen[tgt[i]] = 1;
    }
  }
  // A target whose rays are not together: the records in target order, a
  // stable sort, so each target's rays keep the caller's order (the
  // spectrum numbers are labels and stay as they are).
  std::vector<int64_t> order;
  if (!contiguous) {
    order.resize(nray);
    for (int64_t i = 0; i < nray; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return tgt[a] < tgt[b]; });
    std::vector<CallRay> tmp(cr, cr + nray);
    for (int64_t j = 0; j < nray; j++) cr[j] = tmp[order[j]];
  }
  // (the global indices come from the local cell index, which was formed
  // from them, so the caller's records are not read again)
  const int64_t nyz = (int64_t)g.ny * g.nz;
  for (int64_t j = 0; j < nray; j++) {
    const int64_t i = contiguous ? j : order[j];
    if (j == 0 || tgt[i] != ct[ntarget - 1].cell) {
      const int64_t c = tgt[i];
      CallTarget& T = ct[ntarget++];
      T.first = j;
      T.cell = c;
      T.count = 0;
      T.g[0] = (int32_t)(c / nyz) + gd.ox;
      T.g[1] = (int32_t)((c / g.nz) % g.ny) + gd.oy;
      T.g[2] = (int32_t)(c % g.nz) + gd.oz;
    }
    ct[ntarget - 1].count++;
  }
  // Targets in 4x4x4 tiles, so that each wavefront of the kernel holds 64
  // neighbouring targets, whose rays mostly come from the same nodes and
  // spectra (as in the fused step); each target's rays keep their order, so
  // every cell's sum is unchanged.
  // A counting sort, least significant key first: by the place inside the
  // tile (64 values), then stably by the tile (x, then y, then z).  The
  // targets are distinct cells, so the order is that of (tile, place).
  {
    const int64_t ntx = ((int64_t)g.nx + 3) / 4, nty = ((int64_t)g.ny + 3) / 4,
                  ntz = ((int64_t)g.nz + 3) / 4;
    auto local = [&](const CallTarget& T, int d) {
      return (int64_t)T.g[d] - (d == 0 ? gd.ox : d == 1 ? gd.oy : gd.oz);
    };
    auto place = [&](const CallTarget& T) {
      return (int)(((local(T, 0) & 3) << 4) | ((local(T, 1) & 3) << 2) | (local(T, 2) & 3));
    };
    auto tile = [&](const CallTarget& T) {
      return ((local(T, 0) >> 2) * nty + (local(T, 1) >> 2)) * ntz + (local(T, 2) >> 2);
    };
    std::vector<CallTarget> tmp((size_t)ntarget);
    std::vector<int64_t> pos(64 + 1, 0);
    for (int64_t q = 0; q < ntarget; q++) pos[place(ct[q]) + 1]++;
    for (int k = 0; k < 64; k++) pos[k + 1] += pos[k];
    for (int64_t q = 0; q < ntarget; q++) tmp[pos[place(ct[q])]++] = ct[q];
    std::vector<int64_t> tpos((size_t)(ntx * nty * ntz) + 1, 0);
    for (int64_t q = 0; q < ntarget; q++) tpos[tile(tmp[q]) + 1]++;
    for (size_t k = 0; k + 1 < tpos.size(); k++) tpos[k + 1] += tpos[k];
    for (int64_t q = 0; q < ntarget; q++) ct[tpos[tile(tmp[q])]++] = tmp[q];
  }
  const int64_t nspec = (int64_t)rows.size();
  // The tables, as the per-ray kernels of the fused step read a node's
  // spectrum: FP32 values scaled by the row maximum, the maximum in FP64,
  // and the FP64 row for the rays FP32 cannot serve.
  const size_t off_tab = sizeof(CallRay) * nray + sizeof(CallTarget) * ntarget;
  const size_t tab_bytes = (sizeof(float) * ng + sizeof(double) * ng + sizeof(double)) * nspec;
  std::vector<unsigned char> host_tab;
  unsigned char* tab = stage + off_tab;
  if (off_tab + tab_bytes > cap) {
    host_tab.resize(tab_bytes);
    tab = host_tab.data();
  }
  float* s32 = reinterpret_cast<float*>(tab);
  double* s64 = reinterpret_cast<double*>(tab + sizeof(float) * ng * nspec);
  double* ssc = s64 + (size_t)ng * nspec;
  for (int64_t u = 0; u < nspec; u++) {
    const double* row = rows[u];
    double mx = 0.0;
    for (int k = 0; k < ng; k++) mx = std::max(mx, row[k]);
    const double inv = mx > 0.0 ? 1.0 / mx : 0.0;
    for (int k = 0; k < ng; k++) s32[u * ng + k] = (float)(row[k] * inv);
    memcpy(s64 + u * ng, row, sizeof(double) * ng);
    ssc[u] = mx;
  }
  HIP_OK(g.rays.reserve(off_tab + tab_bytes));
  unsigned char* d = g.rays.p;
  HIP_OK(hipMemcpyAsync(d, stage, off_tab, hipMemcpyHostToDevice, g.stream));
  HIP_OK(hipMemcpyAsync(d + off_tab, tab, tab_bytes, hipMemcpyHostToDevice, g.stream));
  const RateTables rt = rate_tables(g);
  HIP_OK(hipEventRecord(g.ev_c0, g.stream));
  HIP_OK(launch_target_rates(
      gd, rt, g.mesh, reinterpret_cast<const CallTarget*>(d + sizeof(CallRay) * nray), ntarget,
      reinterpret_cast<const CallRay*>(d), reinterpret_cast<const float*>(d + off_tab),
      reinterpret_cast<const double*>(d + off_tab + sizeof(float) * ng * nspec),
      reinterpret_cast<const double*>(d + off_tab + (sizeof(float) + sizeof(double)) * ng * nspec),
      g.stream));
  HIP_OK(hipEventRecord(g.ev_c1, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, g.ev_c0, g.ev_c1));
  g.stats.ms_rate = ms;
  g.stats.nspec_call = (uint64_t)nspec;
}

// The entry points whose host buffers grow with the caller's input report
// a failed allocation as ARGOT_ERR_NOMEM: no exception crosses the C-ABI.
void calc_photoion_rate_cuda(struct cuda_mem_space* cuda_mem, struct light_ray* ray,
                             struct light_ray_IO* ray_IO, struct cuda_param* this_cuda,
                             struct run_param* this_run) {
  try {
    calc_photoion_rate_cuda_body(cuda_mem, ray, ray_IO, this_cuda, this_run);
  } catch (const std::bad_alloc&) {
    set_err(ARGOT_ERR_NOMEM);
  } catch (...) {
    set_err(ARGOT_ERR_ARG);
  }
}

void zero_out_photoion_rate_cuda(struct cuda_mem_space* cuda_mem, struct cuda_param* this_cuda,
                                 struct run_param* this_run) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  HIP_OK(launch_zero_rates(g.mesh, g.ncell, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
}

static void merge_photoion_data_body(struct fluid_mesh* mesh, struct cuda_mem_space* cuda_mem,
                                     struct cuda_param* this_cuda, struct run_param* this_run) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  if (!mesh) return set_err(ARGOT_ERR_ARG);
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  std::vector<float> tab(6 * g.ncell);
  for (int64_t c = 0; c < g.ncell; c++) {
    const prim_chem& p = mesh[c].prev_chem;
    tab[6 * c + 0] = p.GammaHI;
    tab[6 * c + 1] = p.HeatHI;
    tab[6 * c + 2] = p.GammaHeI;
    tab[6 * c + 3] = p.GammaHeII;
    tab[6 * c + 4] = p.HeatHeI;
    tab[6 * c + 5] = p.HeatHeII;
  }
  HIP_OK(g.rate_tab.reserve(6 * g.ncell));
  HIP_OK(hipMemcpyAsync(g.rate_tab.p, tab.data(), sizeof(float) * 6 * g.ncell,
                        hipMemcpyHostToDevice, g.stream));
  HIP_OK(launch_add_rates(g.mesh, g.rate_tab.p, g.ncell, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
}

void recv_photoion_data(struct photoion_rate* gamma, struct cuda_mem_space* cuda_mem,
                        struct cuda_param* this_cuda, struct run_param* this_run) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  if (!gamma) return set_err(ARGOT_ERR_ARG);
  if (!cuda_mem->rate_dev) return set_err(ARGOT_ERR_STATE);
  float* rate = reinterpret_cast<float*>(cuda_mem->rate_dev);
  HIP_OK(launch_get_rates(g.mesh, rate, g.ncell, g.stream));
  HIP_OK(hipMemcpyAsync(gamma, rate, sizeof(struct photoion_rate) * g.ncell,
                        hipMemcpyDeviceToHost, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
}

void send_photoion_data(struct photoion_rate* gamma, struct cuda_mem_space* cuda_mem,
                        struct cuda_param* this_cuda, struct run_param* this_run) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  if (!gamma) return set_err(ARGOT_ERR_ARG);
  if (!cuda_mem->rate_dev) return set_err(ARGOT_ERR_STATE);
  float* rate = reinterpret_cast<float*>(cuda_mem->rate_dev);
  HIP_OK(hipMemcpyAsync(rate, gamma, sizeof(struct photoion_rate) * g.ncell,
                        hipMemcpyHostToDevice, g.stream));
  HIP_OK(launch_set_rates(g.mesh, rate, g.ncell, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
}

void merge_photoion_data(struct fluid_mesh* mesh, struct cuda_mem_space* cuda_mem,
                         struct cuda_param* this_cuda, struct run_param* this_run) {
  try {
    merge_photoion_data_body(mesh, cuda_mem, this_cuda, this_run);
  } catch (const std::bad_alloc&) {
    set_err(ARGOT_ERR_NOMEM);
  } catch (...) {
    set_err(ARGOT_ERR_ARG);
  }
}

// One point-source pass of step_radiation_tree (calc_radiation_transfer
// with zero_out_photoion_rate_cuda before it): rates into
// mesh_dev[].prev_chem, counters into g.stats.
static void point_source_pass(Ctx& g, struct radiation_src* src, struct run_param* this_run) {
  const double t_start = now_ms();
  // The fused step covers one rank that owns the whole box: its tree, its
  // segments (domain faces of the global box) and its walk (local box) are
  // then over the same box.  A rank of a decomposed mesh would treat the
  // other ranks' regions as outside the box and miss their columns.
  if (this_run->rank_x || this_run->rank_y || this_run->rank_z || this_run->nnode_x > 1 ||
      this_run->nnode_y > 1 || this_run->nnode_z > 1 ||
      this_run->xmin_local != this_run->xmin || this_run->ymin_local != this_run->ymin ||
      this_run->zmin_local != this_run->zmin || this_run->xmax_local != this_run->xmax ||
      this_run->ymax_local != this_run->ymax || this_run->zmax_local != this_run->zmax)
    return set_err(ARGOT_ERR_ARG);
  const int64_t nsrc = (int64_t)this_run->nsrc;
  if (nsrc < 0 || (nsrc > 0 && !src)) return set_err(ARGOT_ERR_ARG);
  // Rate mode 1 needs a trained network.
  if (g_rate_mode == 1 && !surrogate_ready()) return set_err(ARGOT_ERR_STATE);
  this_run->theta_crit = g_theta;

  // Device work that does not need the tree goes first, so that it runs
  // while the host builds the tree: per-cell FP32 densities, rates zeroed
  // (zero_out_photoion_rate_cuda), counters reset.
  HIP_OK(g.nf.reserve(g.ncell));
  HIP_OK(hipEventRecord(g.ev0, g.stream));
  HIP_OK(launch_pack_fields(g.mesh, g.ncell, g_helium_fact, g.nf.p, g.stream));
  HIP_OK(launch_zero_rates(g.mesh, g.ncell, g.stream));
  HIP_OK(launch_reset_counters(g.ctr.p, g.stream));

  // Source tree on the host (construct_tree), uploaded in pre-order.  The
  // spectra are read where they lie in the radiation_src records.
  static_assert(sizeof(struct radiation_src) % sizeof(double) == 0 &&
                    offsetof(struct radiation_src, photon_rate) % sizeof(double) == 0,
                "photon_rate rows are a whole number of doubles apart");
  std::vector<float> pos(3 * std::max<int64_t>(nsrc, 1));
  for (int64_t s = 0; s < nsrc; s++) {
    pos[3 * s + 0] = src[s].xpos;
    pos[3 * s + 1] = src[s].ypos;
    pos[3 * s + 2] = src[s].zpos;
  }
  static const double no_spec[ARGOT_NGRID_NU] = {};
  const double* spec = nsrc > 0 ? src[0].photon_rate : no_spec;
  const TreeBox box{this_run->xmin, this_run->ymin, this_run->zmin,
                    this_run->xmax, this_run->ymax, this_run->zmax};
  if (build_tree(pos.data(), spec, sizeof(struct radiation_src) / sizeof(double), (int)nsrc,
                 ARGOT_NGRID_NU, box, this_run->theta_crit, &g.tree) != 0)
    return set_err(ARGOT_ERR_ARG);
  const HostTree& T = g.tree;
  const int nn = T.nnode();
  HIP_OK(g.link.reserve(nn));
  HIP_OK(g.cm.reserve(4 * (size_t)nn));
  HIP_OK(g.spec.reserve((size_t)nn * ARGOT_NGRID_NU));
  HIP_OK(g.spec_scale.reserve(nn));
  HIP_OK(g.src_pos.reserve(T.src_pos.size()));
  HIP_OK(hipMemcpyAsync(g.link.p, T.link.data(), sizeof(NodeLink) * nn, hipMemcpyHostToDevice,
                        g.stream));
  HIP_OK(hipMemcpyAsync(g.cm.p, T.cm.data(), sizeof(float) * 4 * nn, hipMemcpyHostToDevice,
                        g.stream));
  HIP_OK(hipMemcpyAsync(g.spec.p, T.spec.data(), sizeof(float) * T.spec.size(),
                        hipMemcpyHostToDevice, g.stream));
  HIP_OK(hipMemcpyAsync(g.spec_scale.p, T.spec_scale.data(), sizeof(double) * nn,
                        hipMemcpyHostToDevice, g.stream));
  HIP_OK(hipMemcpyAsync(g.src_pos.p, T.src_pos.data(), sizeof(float) * T.src_pos.size(),
                        hipMemcpyHostToDevice, g.stream));
  const double t_built = now_ms();

  FusedArgs a{};   // zero-initialised: the graph cache compares its bytes
  a.grid = make_grid(this_run);
  for (int d = 0; d < 3; d++) a.tile_bits[d] = g_tile_bits[d];
  a.tree.nnode = nsrc > 0 ? nn : 0;
  a.tree.node_cm = g.cm.p;
  a.tree.link = g.link.p;
  a.tree.spec = g.spec.p;
  a.tree.spec_scale = g.spec_scale.p;
  a.tree.src_pos = g.src_pos.p;
  a.rates = rate_tables(g);
  // With no trusted cell every ray would take the exact integral, so the
  // plain kernel runs.
  const int surr = g_rate_mode == 1 && surrogate_any_trusted();
  if (surr) {
    std::vector<double> q;
    surrogate_node_q(T.spec64.data(), nn, ARGOT_NGRID_NU, &q);
    HIP_OK(g.node_q.reserve(q.size()));
    HIP_OK(hipMemcpyAsync(g.node_q.p, q.data(), sizeof(double) * q.size(), hipMemcpyHostToDevice,
                          g.stream));
    a.surr = surrogate_dev();
    a.surr.node_q = g.node_q.p;
  } else {
    a.surr = SurrDev{};
  }
  a.nf = g.nf.p;
  a.mesh = g.mesh;
  a.ctr = g.ctr.p;
  const int capture = g_capture_host != nullptr;
  if (capture) HIP_OK(g.rec.reserve(std::max<int64_t>(g_capture_cap, 1)));
  a.rec = capture ? g.rec.p : nullptr;
  a.rec_cap = capture ? g_capture_cap : 0;
  const int nblock = std::max(1, g.ncu * g_blocks_per_cu);
  // The fused persistent kernel, timed alone between ev_mid and ev1.  Its
  // launch is replayed from a HIP graph, captured again only when a launch
  // argument changes (the stream events stay outside the graph).
  auto enqueue = [&]() -> hipError_t {
    return launch_fused_step(a, g_tau_mode, capture, surr, g_wide_index, nblock, g.stream);
  };
  const bool use_graph = g_use_graph && !capture;
  if (use_graph) {
    GraphKey key{};
    key.args = a;
    key.nblock = nblock;
    key.kahan = g_tau_mode;
    key.surr = surr;
    key.wide = g_wide_index;
    key.ncell = g.ncell;
    key.helium_fact = g_helium_fact;
    key.nf = g.nf.p;
    if (!g.graph_exec || memcmp(&key, &g.graph_key, sizeof key) != 0) {
      if (g.graph_exec) HIP_OK(hipGraphExecDestroy(g.graph_exec));
      if (g.graph) HIP_OK(hipGraphDestroy(g.graph));
      g.graph_exec = nullptr;
      g.graph = nullptr;
      HIP_OK(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
      const hipError_t e = enqueue();
      hipGraph_t graph = nullptr;
      const hipError_t e2 = hipStreamEndCapture(g.stream, &graph);
      HIP_OK(e);
      HIP_OK(e2);
      g.graph = graph;
      HIP_OK(hipGraphInstantiate(&g.graph_exec, g.graph, nullptr, nullptr, 0));
      g.graph_key = key;
      g.stats_graph_captures++;
    }
    HIP_OK(hipEventRecord(g.ev_mid, g.stream));
    HIP_OK(hipGraphLaunch(g.graph_exec, g.stream));
    HIP_OK(hipEventRecord(g.ev1, g.stream));
  } else {
    HIP_OK(hipEventRecord(g.ev_mid, g.stream));
    HIP_OK(enqueue());
    HIP_OK(hipEventRecord(g.ev1, g.stream));
  }
  StepCounters hc;
  HIP_OK(hipMemcpyAsync(&hc, g.ctr.p, sizeof(hc), hipMemcpyDeviceToHost, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, g.ev_mid, g.ev1));
  float ms_all = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms_all, g.ev0, g.ev1));
  if (capture) {
    const int64_t n = std::min<int64_t>((int64_t)hc.nrecord, g_capture_cap);
    HIP_OK(hipMemcpy(g_capture_host, g.rec.p, sizeof(RayRecord) * n, hipMemcpyDeviceToHost));
  }
  g.stats.ntarget = (uint64_t)g.ncell;
  g.stats.nray = hc.nray;
  g.stats.ncross = hc.ncross;
  g.stats.nfallback = hc.nfallback;
  g.stats.nsurrogate = hc.nsurrogate;
  g.stats.nvisit = hc.nvisit;
  g.stats.nemit_iter = hc.nemit_iter;
  g.stats.noverflow = hc.noverflow;
  g.stats.ms_build = t_built - t_start;
  g.stats.ms_kernel = ms;
  g.stats.ms_total = now_ms() - t_start;
  this_run->nray = hc.nray;
  this_run->nseg = hc.nray;
  // The reference asserts num_segment < NSEG_PER_RAY for every ray
  // (calc_ray_segment.c:212); here the step reports it.
  if (hc.noverflow) set_err(ARGOT_ERR_TOO_MANY_SEGMENTS);
}

static ChemParams chem_params(const run_param* r) {
  ChemParams p{};
  p.lunit = r->lunit;
  p.munit = r->munit;
  p.tunit = r->tunit;
  p.uenetok = r->uenetok;
  p.denstonh = r->denstonh;
  p.helium_fact = g_helium_fact;
  p.anow = r->anow;
  p.znow = r->znow;
  return p;
}

// The chemistry kernel over every cell, timed with HIP events; the maxima
// of the relative changes as the reference reduces them (max over blocks
// starting from 0, then MAX(FLT_MIN, .)).
static void chem_step(Ctx& g, const run_param* r, float dtime, float* diff_chem,
                      float* diff_uene) {
  if (!(r->tunit > 0.0) || !(r->uenetok > 0.0) || !(r->munit > 0.0) || !(r->lunit > 0.0))
    return set_err(ARGOT_ERR_ARG);
  // A cell integrates until its elapsed time reaches dtime * tunit: that
  // product must be a finite number of seconds.
  const double dtime_p = (double)dtime * r->tunit;
  if (!(dtime_p >= -DBL_MAX && dtime_p <= DBL_MAX)) return set_err(ARGOT_ERR_ARG);
  HIP_OK(hipEventRecord(g.ev_c0, g.stream));
  const int qblocks = g_chem_queue ? 4 * g.ncu : 0;
  HIP_OK(launch_chem_step(g.mesh, g.ncell, chem_params(r), dtime, g.chem_ctr.p, qblocks,
                          g_chem_queue, g.stream));
  HIP_OK(hipEventRecord(g.ev_c1, g.stream));
  ChemCounters hc;
  HIP_OK(hipMemcpyAsync(&hc, g.chem_ctr.p, sizeof hc, hipMemcpyDeviceToHost, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, g.ev_c0, g.ev_c1));
  float dc, du;
  memcpy(&dc, &hc.diff_chem, sizeof dc);
  memcpy(&du, &hc.diff_uene, sizeof du);
  const float fmin_norm = FLT_MIN;
  if (diff_chem) *diff_chem = fmin_norm > dc ? fmin_norm : dc;
  if (diff_uene) *diff_uene = fmin_norm > du ? fmin_norm : du;
  g.chem_stats.ncell = (uint64_t)g.ncell;
  g.chem_stats.nsub = hc.nsub;
  g.chem_stats.niter = hc.niter;
  g.chem_stats.nstopped = hc.nstopped;
  g.chem_stats.ms_kernel = ms;
}

static void step_radiation_tree_body(struct fluid_mesh* mesh, struct radiation_src* src,
                                     struct run_param* this_run, struct mpi_param* this_mpi,
                                     struct cuda_mem_space* cuda_mem, struct cuda_param* this_cuda,
                                     float dtime) {
  (void)mesh;
  (void)this_mpi;
  (void)this_cuda;
  t_err = ARGOT_OK;
  const double t_start = now_ms();
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  g.stats.iterations = 1;
  g.stats.ms_chem = 0.0;
  point_source_pass(g, src, this_run);
  if (t_err || g_iterations <= 0) return;
  // src/step_radiation_tree.c:666-723: smooth, chemistry, and another pass
  // while the chemistry or the energy still moved by more than 1e-3.
  uint64_t nray = g.stats.nray, ncross = g.stats.ncross, novf = g.stats.noverflow;
  uint64_t nvisit = g.stats.nvisit, nemit = g.stats.nemit_iter;
  double ms_kernel = g.stats.ms_kernel, ms_chem = 0.0;
  int iter = 1;
  for (;;) {
    HIP_OK(g.smooth_in.reserve(6 * (size_t)g.ncell));
    HIP_OK(g.smooth_out.reserve(6 * (size_t)g.ncell));
    HIP_OK(hipEventRecord(g.ev_c0, g.stream));
    HIP_OK(launch_smooth_rates(g.mesh, g.nx, g.ny, g.nz, g.smooth_in.p, g.smooth_out.p,
                               g.stream));
    HIP_OK(hipEventRecord(g.ev_c1, g.stream));
    HIP_OK(hipEventSynchronize(g.ev_c1));
    float ms_s = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms_s, g.ev_c0, g.ev_c1));
    float dc = 0.0f, du = 0.0f;
    chem_step(g, this_run, dtime, &dc, &du);
    if (t_err) return;
    ms_chem += ms_s + g.chem_stats.ms_kernel;
    if (!((dc > 1.0e-3 || du > 1.0e-3) && iter < g_iterations)) break;
    point_source_pass(g, src, this_run);
    if (t_err) return;
    nray += g.stats.nray;
    ncross += g.stats.ncross;
    nvisit += g.stats.nvisit;
    nemit += g.stats.nemit_iter;
    novf += g.stats.noverflow;
    ms_kernel += g.stats.ms_kernel;
    iter++;
  }
  HIP_OK(launch_chem_update(g.mesh, g.ncell, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
  g.stats.iterations = (uint64_t)iter;
  g.stats.nray = nray;
  g.stats.ncross = ncross;
  g.stats.nvisit = nvisit;
  g.stats.nemit_iter = nemit;
  g.stats.noverflow = novf;
  g.stats.ms_kernel = ms_kernel;
  g.stats.ms_chem = ms_chem;
  g.stats.ms_total = now_ms() - t_start;
}

void step_radiation_tree(struct fluid_mesh* mesh, struct radiation_src* src,
                         struct run_param* this_run, struct mpi_param* this_mpi,
                         struct cuda_mem_space* cuda_mem, struct cuda_param* this_cuda,
                         float dtime) {
  try {
    step_radiation_tree_body(mesh, src, this_run, this_mpi, cuda_mem, this_cuda, dtime);
  } catch (const std::bad_alloc&) {
    set_err(ARGOT_ERR_NOMEM);
  } catch (...) {
    set_err(ARGOT_ERR_ARG);
  }
}

void step_chemistry_gpu(struct cuda_mem_space* cuda_mem, struct cuda_param* this_cuda,
                        struct run_param* this_run, float* diff_chem, float* diff_uene,
                        float dtime) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  chem_step(g, this_run, dtime, diff_chem, diff_uene);
}

void step_chemistry(struct cuda_mem_space* cuda_mem, struct cuda_param* this_cuda,
                    struct run_param* this_run, float dtime, float* max_diff_chem,
                    float* max_diff_uene) {
  step_chemistry_gpu(cuda_mem, this_cuda, this_run, max_diff_chem, max_diff_uene, dtime);
}

double calc_dtime_gpu(struct cuda_mem_space* cuda_mem, struct cuda_param* this_cuda,
                      struct run_param* this_run) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return 0.0;
  Ctx& g = *cp;
  ChemCounters hc;
  hipError_t e = launch_chem_dtime(g.mesh, g.ncell, chem_params(this_run), g.chem_ctr.p, g.stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&hc, g.chem_ctr.p, sizeof hc, hipMemcpyDeviceToHost, g.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
  if (e != hipSuccess) {
    set_hip_err(e, __LINE__);
    return 0.0;
  }
  float dt;
  memcpy(&dt, &hc.dt_min, sizeof dt);
  return (double)dt;
}

double calc_timestep_chem(struct cuda_mem_space* cuda_mem, struct cuda_param* this_cuda,
                          struct run_param* this_run) {
  const double dt = calc_dtime_gpu(cuda_mem, this_cuda, this_run);
  return t_err ? 0.0 : dt / this_run->tunit;
}

void update_chemistry_gpu(struct cuda_mem_space* cuda_mem, struct cuda_param* this_cuda,
                          struct run_param* this_run) {
  (void)this_cuda;
  t_err = ARGOT_OK;
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  HIP_OK(launch_chem_update(g.mesh, g.ncell, g.stream));
  HIP_OK(hipStreamSynchronize(g.stream));
}

// Statistics of the last call on this thread (its context's).
void argot_hip_chem_stats(struct argot_chem_stats* out) {
  const Ctx* c = last_ctx();
  if (out) *out = c ? c->chem_stats : argot_chem_stats{};
}

void argot_hip_set_iterations(int n) { g_iterations = n > 0 ? n : 0; }

void argot_hip_set_wide_index(int on) { g_wide_index = on ? 1 : 0; }
void argot_hip_set_chem_queue(int n) { g_chem_queue = n > 0 ? std::min(n, 64) : 0; }
void argot_hip_set_host_threads(int n) { g_host_threads = n > 0 ? std::min(n, 256) : 0; }

// Copy bandwidth probe: reps device-to-device copies of nbytes with a plain
// streaming kernel; returns the best GB/s (read + write bytes / time).
// smooth_photoion_rate, GPU variant (src/prototype.h:89,
// src/smooth_photoion_rate.c:27-1588) on one rank: the rates in
// mesh_dev[].prev_chem are smoothed on the device and, as the reference
// does, also written into the host mesh's prev_chem (when mesh != NULL).
static void smooth_photoion_rate_body(struct fluid_mesh* mesh, struct run_param* this_run,
                                      struct cuda_mem_space* cuda_mem, struct cuda_param* this_cuda,
                                      struct mpi_param* this_mpi) {
  (void)this_cuda;
  (void)this_mpi;
  t_err = ARGOT_OK;
  Ctx* const cp = check_ready(cuda_mem, this_run);
  if (!cp) return;
  Ctx& g = *cp;
  if (this_run->nnode_x > 1 || this_run->nnode_y > 1 || this_run->nnode_z > 1)
    return set_err(ARGOT_ERR_ARG);   // the halo exchange of several ranks is not built
  HIP_OK(g.smooth_in.reserve(6 * (size_t)g.ncell));
  HIP_OK(g.smooth_out.reserve(6 * (size_t)g.ncell));
  HIP_OK(launch_smooth_rates(g.mesh, g.nx, g.ny, g.nz, g.smooth_in.p, g.smooth_out.p, g.stream));
  if (mesh) {
    std::vector<float> soa(6 * (size_t)g.ncell);
    HIP_OK(hipMemcpyAsync(soa.data(), g.smooth_out.p, sizeof(float) * soa.size(),
                          hipMemcpyDeviceToHost, g.stream));
    HIP_OK(hipStreamSynchronize(g.stream));
    for (int64_t c = 0; c < g.ncell; c++) {
      float* r = &mesh[c].prev_chem.GammaHI;
      for (int k = 0; k < 6; k++) r[k] = soa[(size_t)k * g.ncell + c];
    }
  } else {
    HIP_OK(hipStreamSynchronize(g.stream));
  }
}

void smooth_photoion_rate(struct fluid_mesh* mesh, struct run_param* this_run,
                          struct cuda_mem_space* cuda_mem, struct cuda_param* this_cuda,
                          struct mpi_param* this_mpi) {
  try {
    smooth_photoion_rate_body(mesh, this_run, cuda_mem, this_cuda, this_mpi);
  } catch (const std::bad_alloc&) {
    set_err(ARGOT_ERR_NOMEM);
  } catch (...) {
    set_err(ARGOT_ERR_ARG);
  }
}

// Autotuner: runs the step on the caller's inputs for each target-tile shape
// (the 64 targets of a wavefront share every tree decision, so the shape
// sets how coherent the wavefront's rays are) and persistent-grid size,
// keeps the fastest fused kernel, and leaves it selected.
void argot_hip_autotune_step(struct fluid_mesh* mesh, struct radiation_src* src,
                             struct run_param* this_run, struct cuda_mem_space* cuda_mem,
                             struct cuda_param* this_cuda, struct argot_tune_report* rep) {
  static const int kTiles[][3] = {{4, 4, 4}, {2, 4, 8}, {4, 2, 8}, {1, 8, 8}, {8, 8, 1},
                                  {2, 2, 16}, {8, 4, 2}, {4, 8, 2}};
  static const int kBlocks[] = {16, 24, 32};
  t_err = ARGOT_OK;
  if (!rep) return set_err(ARGOT_ERR_ARG);
  memset(rep, 0, sizeof(*rep));
  const int old_bits[3] = {g_tile_bits[0], g_tile_bits[1], g_tile_bits[2]};
  const int old_bpc = g_blocks_per_cu;
  double best = 1e300;
  int n = 0;
  for (const auto& tl : kTiles)
    for (int bpc : kBlocks) {
      if (n >= ARGOT_TUNE_MAX) break;
      argot_hip_set_tile(tl[0], tl[1], tl[2]);
      g_blocks_per_cu = bpc;
      step_radiation_tree(mesh, src, this_run, nullptr, cuda_mem, this_cuda, 0.0f);
      if (t_err) {
        for (int d = 0; d < 3; d++) g_tile_bits[d] = old_bits[d];
        g_blocks_per_cu = old_bpc;
        return;
      }
      argot_tune_entry& e = rep->entry[n];
      for (int d = 0; d < 3; d++) e.tile[d] = tl[d];
      e.blocks_per_cu = bpc;
      e.ms_kernel = find_ctx(cuda_mem)->stats.ms_kernel;
      if (e.ms_kernel < best) {
        best = e.ms_kernel;
        rep->best = n;
      }
      n++;
    }
  rep->n = n;
  const argot_tune_entry& b = rep->entry[rep->best];
  argot_hip_set_tile(b.tile[0], b.tile[1], b.tile[2]);
  g_blocks_per_cu = b.blocks_per_cu;
}

double argot_hip_copy_probe(int64_t nbytes, int reps) {
  t_err = ARGOT_OK;
  if (nbytes < 16 || reps < 1) {
    set_err(ARGOT_ERR_ARG);
    return 0.0;
  }
  const int64_t n = nbytes / 16;
  float4 *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double best = 0.0;
  int ncu = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) == hipSuccess) ncu = prop.multiProcessorCount;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess ||
      hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    set_err(ARGOT_ERR_HIP);
  } else if (hipMemsetAsync(a, 0, n * 16, s) == hipSuccess) {
    for (int r = 0; r < reps && !t_err; r++) {
      if (hipEventRecord(e0, s) != hipSuccess ||
          launch_copy_probe(a, b, n, ncu * 8, s) != hipSuccess ||
          hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
        set_err(ARGOT_ERR_HIP);
        break;
      }
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.0f)
        best = std::max(best, 2.0 * (double)(n * 16) / (ms * 1e-3) / 1e9);
    }
  } else {
    set_err(ARGOT_ERR_HIP);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return best;
}

// ------------------------------------------------------------------------
// Host-only helpers for tests (no GPU needed).

// sizeof / offsetof of the ABI structs, in a fixed order.
int argot_hip_abi_layout(int64_t* out, int n) {
  const int64_t v[] = {
      (int64_t)sizeof(struct prim_chem),
      (int64_t)sizeof(struct fluid_mesh),
      (int64_t)sizeof(struct freq_param),
      (int64_t)sizeof(struct radiation_src),
      (int64_t)sizeof(struct cross_section),
      (int64_t)sizeof(struct run_param),
      (int64_t)sizeof(struct photoion_rate),
      (int64_t)sizeof(struct ray_segment),
      (int64_t)sizeof(struct light_ray),
      (int64_t)sizeof(struct light_ray_IO),
      (int64_t)sizeof(struct cuda_param),
      (int64_t)sizeof(struct cuda_mem_space),
      (int64_t)offsetof(struct run_param, lunit),
      (int64_t)offsetof(struct run_param, denstonh),
      (int64_t)offsetof(struct run_param, freq),
      (int64_t)offsetof(struct run_param, csect),
      (int64_t)offsetof(struct run_param, nray),
      (int64_t)offsetof(struct run_param, nseg),
      (int64_t)offsetof(struct run_param, xmin_local),
      (int64_t)offsetof(struct run_param, delta_x),
      (int64_t)offsetof(struct fluid_mesh, chem),
      (int64_t)offsetof(struct fluid_mesh, prev_chem),
      (int64_t)offsetof(struct light_ray_IO, optical_depth_HI),
      (int64_t)sizeof(RayRecord),
  };
  const int m = (int)(sizeof(v) / sizeof(v[0]));
  for (int i = 0; i < n && i < m; i++) out[i] = v[i];
  return m;
}

// The segment builder the fused kernel runs (segment_start), on the host:
// start point of the segment whose optical depth the ray from src to target
// cell (ix, iy, iz) keeps, and the number of segments counted up to
// NSEG_PER_RAY (a return of NSEG_PER_RAY is a ray the step refuses).
int argot_hip_segment_start(const struct run_param* r, const float* src, int ix, int iy, int iz,
                            float* start) {
  const GridDev gd = make_grid(r);
  const Vec3f s{src[0], src[1], src[2]};
  const Vec3f t{cell_centre(gd.gxmin, ix, gd.dx), cell_centre(gd.gymin, iy, gd.dy),
                cell_centre(gd.gzmin, iz, gd.dz)};
  Vec3f st{};
  const int n = segment_start(gd, s, t, &st);
  start[0] = st.x;
  start[1] = st.y;
  start[2] = st.z;
  return n;
}

// Product tree on the host: build from nsrc sources and list the rays of a
// target point, for comparison with the CPU oracle.
static void* argot_hip_tree_build_body(const float* pos, const double* spec, int nsrc, int ngrid,
                                       const float* box6, float theta) {
  std::unique_ptr<HostTree> t(new HostTree());
  const TreeBox b{box6[0], box6[1], box6[2], box6[3], box6[4], box6[5]};
  if (build_tree(pos, spec, (size_t)std::max(ngrid, 0), nsrc, ngrid, b, theta, t.get()) != 0)
    return nullptr;
  return t.release();
}

void* argot_hip_tree_build(const float* pos, const double* spec, int nsrc, int ngrid,
                           const float* box6, float theta) {
  try {
    return argot_hip_tree_build_body(pos, spec, nsrc, ngrid, box6, theta);
  } catch (const std::bad_alloc&) {
    set_err(ARGOT_ERR_NOMEM);
  } catch (...) {
    set_err(ARGOT_ERR_ARG);
  }
  return {};
}

void argot_hip_tree_free(void* h) { delete static_cast<HostTree*>(h); }

int argot_hip_tree_nnode(void* h) { return static_cast<HostTree*>(h)->nnode(); }

static uint64_t argot_hip_tree_rays_to_body(void* h, const float* target, uint64_t cap,
                                            int* kind, int* id, float* xyz) {
  std::vector<HostRay> rays;
  const uint64_t n = tree_rays_to(*static_cast<HostTree*>(h), target, &rays);
  for (uint64_t i = 0; i < n && i < cap; i++) {
    kind[i] = rays[i].kind;
    id[i] = rays[i].id;
    xyz[3 * i + 0] = rays[i].pos[0];
    xyz[3 * i + 1] = rays[i].pos[1];
    xyz[3 * i + 2] = rays[i].pos[2];
  }
  return n;
}

uint64_t argot_hip_tree_rays_to(void* h, const float* target, uint64_t cap, int* kind, int* id,
                                float* xyz) {
  try {
    return argot_hip_tree_rays_to_body(h, target, cap, kind, id, xyz);
  } catch (const std::bad_alloc&) {
    set_err(ARGOT_ERR_NOMEM);
  } catch (...) {
    set_err(ARGOT_ERR_ARG);
  }
  return {};
}

const double* argot_hip_tree_spec(void* h, int node) {
  HostTree* t = static_cast<HostTree*>(h);
  return &t->spec64[(size_t)node * t->ngrid];
}

}  // extern "C"

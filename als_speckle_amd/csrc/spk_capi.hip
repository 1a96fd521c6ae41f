// C ABI of the library (declared in include/speckle_hip.h).  Host code only:
// argument checking, plan state, and the launch sequence of an iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/speckle_hip.h"
#include "spk_internal.hpp"

using namespace spk;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define SPK_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(SPK_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define SPK_REQUIRE(cond, code, msg) \
  do {                               \
    if (!(cond)) return fail(code, msg); \
  } while (0)

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// Defaults measured on MI355X (the autotuner measures the rest):
//  - tile 2 from 2048 up: at 4096 it keeps the 1024-thread passes free of
//    register spills (tile 4 needs 16 values per thread there, which spill);
//    at 2048 it is 13 % faster per iteration for one frame and 17 % for four;
//  - tile 4 below 2048 also for one frame, although resident launches of a
//    256^2 frame take 5 % fewer microseconds per step in tile 2 (a tune
//    choice; at 512^2 the difference changes sign from run to run): the
//    FP64 error partials are summed per column band, so a frame's error
//    trace is bitwise the same alone and among other frames only when the
//    tile is;
//  - 8 values per thread (more threads, shorter dependency chains) wherever
//    the workgroup fits, except: 1024^2 with 4 or more frames, where 16
//    values per thread (radix-16 stages) are 9 % faster per batched
//    iteration, and tile 2 from 2048^2 up, where they are 10 % faster for
//    one 2048^2 frame and 8 % for four, and 2.5-3.5 % faster per 4096^2
//    iteration (row pass 171-175 instead of 179-181 us, column pass equal).
int default_tile(int n) { return n >= 2048 ? 2 : 4; }
int default_ept(int n, int tile, int batch) {
  if (n == 1024 && batch >= 4 && (n / 16) * tile <= 1024) return 16;
  if (n >= 2048 && tile == 2) return 16;
  return (n / 8) * tile <= 1024 ? 8 : 16;
}

template <class T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return SPK_OK;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(SPK_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                " bytes failed: " + hipGetErrorString(e));
  }
  return SPK_OK;
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

}  // namespace

// A context is reference counted: spk_ctx_create holds one reference and
// every plan holds one, so a context outlives its plans whatever order the
// caller destroys them in.
struct spk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  double* scratch = nullptr;   // small reduction workspace (align)
  size_t scratch_cap = 0;
  int refs = 1;
};

static void ctx_release(spk_ctx* ctx) {
  if (--ctx->refs > 0) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  else (void)hipDeviceSynchronize();
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  delete ctx;
}

struct spk_plan {
  spk_ctx* ctx = nullptr;
  int n = 0, batch = 0, shared = 0, tile = 8, ept = 16, bands = 0;
  c32* psi = nullptr;
  c32* phi = nullptr;         // row transform, tiled (row pass input)
  c32* phic = nullptr;        // row transform, column-band-major (column pass input)
  c32* phib = nullptr;
  c32* phibc = nullptr;
  float* mod = nullptr;
  float* sup = nullptr;
  c32* tw = nullptr;
  double* summ2 = nullptr;
  double* sumpart = nullptr;
  double* errp = nullptr;
  size_t errp_cap = 0;        // doubles
  float* sw_tmp = nullptr;
  float* sw_blur = nullptr;
  float* sw_maxp = nullptr;
  c32* fftbuf = nullptr;
  size_t fftbuf_frames = 0;
  // per-call tables (coefficients, activity, frame maps, shrinkwrap
  // parameters) staged in pinned memory and copied in one transfer
  unsigned char* stage_d = nullptr;
  unsigned char* stage_h = nullptr;
  unsigned char* stage_cur = nullptr;  // device tables of the latest spk_iterate
  // Under graph capture the tables go to buffers of the capture's own (a
  // device + pinned pair that the resulting graph owns and frees): a replay
  // copies them from pinned memory that later calls do not overwrite.
  bool capturing = false;
  unsigned char* cap_d = nullptr;
  unsigned char* cap_h = nullptr;
  size_t cap_used = 0;
  size_t stage_cap = 0;
  hipEvent_t upload_done = nullptr;
  bool upload_pending = false;
  bool profiling = false;
  std::vector<hipEvent_t> prof_ev[3];  // start/stop pairs: column, row, resident
  bool mod_set = false, sup_set = false, psi_set = false;
  // per frame: g != 0 when phib == rowFFT((s - g (1 - s)) psi) for the
  // current psi and support (dm1 of a DM step with beta = 1/g), 0 when stale
  std::vector<float> phib_g;
  // persistent iteration blocks (k_persist): on/off, co-resident grid size
  // (0 = not queried yet), barrier counter [0], give-up flag of the running
  // launch [1] (cleared with the counter before each launch) and the sticky
  // give-up flag [2] that spk_plan_check reports and clears
  int persist = 0;
  int pgrid = 0;
  // shrinkwraps fused into the passes where possible (spk_tune.sw_separate = 0)
  bool sw_fuse = true;
  // resident single-frame launches (k_resident) and their exchange mode
  int resident = 0;
  int xm = 0;
  unsigned* rsync = nullptr;   // barrier flags of the resident launches (a line per band)
  unsigned polls = 1u << 18;   // polls of a grid-barrier wait before it gives up
  long long res_launches = 0, res_steps = 0;
  long long sw_fused_count = 0, sw_separate_count = 0;
  unsigned* pbar = nullptr;
  // partial coherence: ipsf (tiled, shared by the frames) and the per-frame
  // work buffers of the |Psi|^2 convolution
  c32* ipsf = nullptr;
  c32* dbuf = nullptr;
  c32* dbufc = nullptr;
  bool ipsf_on = false;
  // the same per-frame work buffers for DM's second projection (of phib)
  c32* dbuf2 = nullptr;
  c32* dbufc2 = nullptr;
  float* imap2 = nullptr;
  float* dmap2 = nullptr;
  // spectral rescale: factors and weights (device), per-frame intensity maps
  float* spec = nullptr;
  int nspec = 0;
  int spec_k = 0;
  float* imap = nullptr;
  float* dmap = nullptr;
  bool spec_on = false;
  // Gaussian coherence fit: autocorrelation, tiled modulus, work frames
  c32* coh_ac = nullptr;
  c32* coh_w = nullptr;
  c32* coh_cb = nullptr;
  float* coh_m = nullptr;
  double* coh_part = nullptr;
  bool coh_ready = false;
  long long nn() const { return (long long)n * n; }
  // Phi kept in two layouts (see spk_fft_core.hpp dual_layout)
  bool dual() const { return n >= kDualMin; }
  c32* phi_in() const { return dual() ? phic : phi; }    // column pass input
  c32* phib_in() const { return dual() ? phibc : phib; }
  long long mstride() const { return shared ? 0 : nn(); }
};

namespace {

int set_device(spk_ctx* ctx) {
  SPK_REQUIRE(ctx, SPK_EINVAL, "null context");
  SPK_HIP(hipSetDevice(ctx->device));
  return SPK_OK;
}

// Real-space update coefficients, from phasing.py:1231-1328 (numpy branch)
// and phasing_{hio,er,raar,sf,dm1,dm2}.cl (OpenCL branch).
int make_coef(const spk_step& s, Coef* c) {
  *c = Coef{0, 0, 0, 0, 0, 0, 0, 0};
  const float b = s.beta;
  switch (s.alg) {
    case SPK_NOOP:
      return SPK_OK;
    case SPK_HIO:  // (1-s)*(psi_in - beta*psi_out) + s*psi_out
      c->ao = 1.0f; c->bo = -b; c->bi = 1.0f;
      break;
    case SPK_ER:
      if (s.semantics == SPK_SEM_OPENCL) c->ao = 1.0f;  // s*psi_out
      else c->ai = 1.0f;                                 // s*psi_in
      break;
    case SPK_RAAR:  // (1-s)*(beta*psi_in + (1-2beta)*psi_out) + s*psi_out
      c->ao = 1.0f; c->bo = 1.0f - 2.0f * b; c->bi = b;
      break;
    case SPK_SF:
      c->ao = 1.0f;
      if (s.semantics == SPK_SEM_OPENCL) c->bo = -1.0f;  // (1-s)(-psi_out)
      else c->bi = -1.0f;                                 // (1-s)(-psi_in)
      break;
    case SPK_DM: {  // gamma_s = -1/beta: t1 = (1+beta)psi_out + psi_in - beta t
      SPK_REQUIRE(b != 0.0f, SPK_EINVAL, "dm needs beta != 0");
      c->ao = 1.0f + b; c->ai = 1.0f; c->at = -b;
      c->bo = -b; c->bi = 1.0f;
      c->flags |= kUseT;
      break;
    }
    default:
      return fail(SPK_EINVAL, "unknown algorithm " + std::to_string(s.alg));
  }
  if (s.semantics != SPK_SEM_NUMPY && s.semantics != SPK_SEM_OPENCL)
    return fail(SPK_EINVAL, "unknown semantics " + std::to_string(s.semantics));
  c->flags |= kActive;
  if (c->ao == 0.0f && c->bo == 0.0f && !(c->flags & kUseT)) c->flags |= kNoOut;
  return SPK_OK;
}

// Blur weights already made, by (n, sigma): a schedule's shrinkwraps repeat
// their sigmas in every reconstruction, and the normalisation over the
// circle takes n exponentials (50 events of a 1024^2 schedule: ~0.5 ms of
// host time in front of the first launch of each call).
struct BlurKey {
  int n;
  float sigma;
  bool operator==(const BlurKey& o) const { return n == o.n && sigma == o.sigma; }
};
std::mutex g_blur_mu;
std::vector<std::pair<BlurKey, BlurW>> g_blur_cache;
constexpr size_t kBlurCacheMax = 1024;

int make_blur(int n, float sigma, BlurW* w) {
  SPK_REQUIRE(sigma > 0.0f && std::isfinite(sigma), SPK_EINVAL, "shrinkwrap sigma must be > 0");
  int R = (int)std::ceil(6.0 * sigma);
  if (R > n / 2) R = n / 2;
  SPK_REQUIRE(R <= 47, SPK_EUNSUP, "shrinkwrap sigma too large (max 7.8 px)");
  const BlurKey key{n, sigma};
  {
    std::lock_guard<std::mutex> lk(g_blur_mu);
    for (const auto& e : g_blur_cache)
      if (e.first == key) {
        *w = e.second;
        return SPK_OK;
      }
  }
  // normalisation over the whole circle of n samples centred at n/2
  // (shape.gaussian centred at size/2, fftshift, divide by the sum:
  // phasing.py:1787-1788)
  const double s2 = 2.0 * (double)sigma * sigma;
  double tot = 0.0;
  for (int x = 0; x < n; ++x) {
    const double d = x - n / 2;
    tot += std::exp(-d * d / s2);
  }
  std::memset(w, 0, sizeof(*w));
  for (int d = 0; d <= R; ++d) w->w[d] = (float)(std::exp(-(double)d * d / s2) / tot);
  // d = +-n/2 is one sample on the circle: counted once
  if (R == n / 2) w->w[R] *= 0.5f;
  w->radius = R;
  std::lock_guard<std::mutex> lk(g_blur_mu);
  if (g_blur_cache.size() >= kBlurCacheMax) g_blur_cache.clear();
  g_blur_cache.emplace_back(key, *w);
  return SPK_OK;
}

PassArgs base_args(spk_plan* p) {
  PassArgs a{};
  a.psi = p->psi;
  a.phi = p->phi;
  a.phic = p->phi_in();
  a.phib = p->phib;
  a.phibc = p->phib_in();
  a.mod = p->mod;
  a.sup = p->sup;
  a.mstride = p->mstride();
  a.sstride = p->mstride();
  a.tw = p->tw;
  a.errp = nullptr;
  a.tab = nullptr;
  a.inv_nn = (float)(1.0 / ((double)p->n * p->n));
  a.dbuf = p->dbuf;
  a.dbufc = p->dual() ? p->dbufc : p->dbuf;
  a.imap = p->imap;
  a.dmap = p->dmap;
  a.dbuf2 = p->dbuf2;
  a.dbufc2 = p->dual() ? p->dbufc2 : p->dbuf2;
  a.imap2 = p->imap2;
  a.dmap2 = p->dmap2;
  a.sw_tmp = p->sw_tmp;
  a.sw_blur = p->sw_blur;
  a.sw_maxp = p->sw_maxp;
  a.polls = p->polls;
  return a;
}

int ensure_phib(spk_plan* p) {
  if (p->phib) return SPK_OK;
  int rc = dalloc(&p->phib, (size_t)p->batch * p->nn());
  if (rc || !p->dual()) return rc;
  return dalloc(&p->phibc, (size_t)p->batch * p->nn());
}

// work buffers of DM's partially coherent second projection
int ensure_dm_pc(spk_plan* p) {
  const size_t count = (size_t)p->batch * p->nn();
  int rc;
  if (p->ipsf_on && !p->dbuf2) {
    if ((rc = dalloc(&p->dbuf2, count))) return rc;
    if (p->dual() && (rc = dalloc(&p->dbufc2, count))) return rc;
  }
  if (p->spec_on && !p->imap2) {
    if ((rc = dalloc(&p->imap2, count)) || (rc = dalloc(&p->dmap2, count))) return rc;
  }
  return SPK_OK;
}

int ensure_sw(spk_plan* p) {
  if (p->sw_tmp) return SPK_OK;
  int rc;
  if ((rc = dalloc(&p->sw_tmp, (size_t)p->batch * p->nn()))) return rc;
  if ((rc = dalloc(&p->sw_blur, (size_t)p->batch * p->nn()))) return rc;
  if ((rc = dalloc(&p->sw_maxp, (size_t)p->batch * p->bands))) return rc;
  return SPK_OK;
}

// the DM estimate phib of frames [f0, f0 + nf) no longer matches psi / the
// support (every frame when nf < 0)
void phib_stale(spk_plan* p, int f0, int nf) {
  if (nf < 0) { f0 = 0; nf = p->batch; }
  for (int f = f0; f < f0 + nf && f < (int)p->phib_g.size(); ++f) p->phib_g[f] = 0.0f;
}

// phi = rowFFT(psi) for frames [f0, f0 + nf)
int rowfft_frames(spk_plan* p, int f0, int nf) {
  const size_t fo = (size_t)f0 * p->nn();
  SPK_HIP(launch_rowfft(p->n, p->tile, p->ept, 0, p->psi + fo, p->phi_in() + fo, nullptr, 0,
                        0.0f, p->tw, nullptr, nf, p->ctx->stream));
  return SPK_OK;
}

// Host-side builder of one staged upload: append arrays, then copy once.
struct Stage {
  std::vector<unsigned char> buf;
  template <class T>
  size_t add(const T* v, size_t count) {
    size_t off = (buf.size() + 255) & ~size_t(255);
    buf.resize(off + count * sizeof(T));
    if (count) std::memcpy(buf.data() + off, v, count * sizeof(T));
    return off;
  }
};

// bytes of the table buffers of one captured graph
constexpr size_t kCaptureStage = (size_t)4 << 20;

int upload_stage(spk_plan* p, const Stage& stg) {
  const size_t need = stg.buf.size();
  if (need == 0) return SPK_OK;
  if (p->capturing) {
    SPK_REQUIRE(p->cap_d && p->cap_used + need <= kCaptureStage, SPK_EUNSUP,
                "the step tables of a captured block exceed " +
                    std::to_string(kCaptureStage >> 20) + " MiB");
    unsigned char* d = p->cap_d + p->cap_used;
    unsigned char* h = p->cap_h + p->cap_used;
    std::memcpy(h, stg.buf.data(), need);
    SPK_HIP(hipMemcpyAsync(d, h, need, hipMemcpyHostToDevice, p->ctx->stream));
    p->cap_used += (need + 255) & ~(size_t)255;
    p->stage_cur = d;
    return SPK_OK;
  }
  if (p->upload_pending) SPK_HIP(hipEventSynchronize(p->upload_done));
  p->upload_pending = false;
  if (need > p->stage_cap) {
    // queued kernels of earlier calls may still read the old tables
    SPK_HIP(hipStreamSynchronize(p->ctx->stream));
    dfree(p->stage_d);
    if (p->stage_h) (void)hipHostFree(p->stage_h);
    p->stage_h = nullptr;
    p->stage_cap = 0;
    const size_t cap = std::max<size_t>(need, 64 * 1024);
    int rc = dalloc(&p->stage_d, cap);
    if (rc) return rc;
    SPK_HIP(hipHostMalloc((void**)&p->stage_h, cap, hipHostMallocDefault));
    p->stage_cap = cap;
  }
  std::memcpy(p->stage_h, stg.buf.data(), need);
  SPK_HIP(hipMemcpyAsync(p->stage_d, p->stage_h, need, hipMemcpyHostToDevice, p->ctx->stream));
  SPK_HIP(hipEventRecord(p->upload_done, p->ctx->stream));
  p->upload_pending = true;
  p->stage_cur = p->stage_d;
  return SPK_OK;
}

// when profiling, give the next pass launch a start/stop event pair, recorded
// by the dispatch itself (hipExtLaunchKernelGGL)
int prof_arm(spk_plan* p, int pass) {
  if (!p->profiling) return SPK_OK;
  hipEvent_t e0, e1;
  SPK_HIP(hipEventCreate(&e0));
  SPK_HIP(hipEventCreate(&e1));
  p->prof_ev[pass].push_back(e0);
  p->prof_ev[pass].push_back(e1);
  g_launch_events[0] = e0;
  g_launch_events[1] = e1;
  return SPK_OK;
}

void prof_clear(spk_plan* p, int pass) {
  for (hipEvent_t e : p->prof_ev[pass]) (void)hipEventDestroy(e);
  p->prof_ev[pass].clear();
}

int make_swparam(int n, float sigma, float tau, SwParam* sp) {
  SPK_REQUIRE(tau >= 0.0f && tau <= 1.0f, SPK_EINVAL, "shrinkwrap tau must be in [0, 1]");
  int rc = make_blur(n, sigma, &sp->w);
  if (rc) return rc;
  sp->tau = tau;
  return SPK_OK;
}

int do_shrinkwrap(spk_plan* p, int frame0, int frames, float sigma, float tau) {
  SPK_REQUIRE(!(p->shared && p->batch > 1), SPK_EUNSUP,
              "shrinkwrap needs a plan with one support per frame");
  SwParam sp;
  int rc = make_swparam(p->n, sigma, tau, &sp);
  if (rc) return rc;
  if ((rc = ensure_sw(p))) return rc;
  SPK_HIP(launch_shrinkwrap(p->n, p->tile, p->psi, p->sup, p->mstride(), p->sw_tmp,
                            p->sw_blur, p->sw_maxp, nullptr, sp, nullptr, frame0, frames,
                            sp.w.radius, p->ctx->stream));
  phib_stale(p, frame0, frames);
  return SPK_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Workspace of spk_align_global_phase_noalloc: the per-block FP64 partial
// sums (frames x 5 sums x blocks) of the alignment, in a file-scope device
// array instead of an allocation.  Every GPU the library runs on holds its
// own copy until the process ends; its address on each device is looked up
// once, with that device current, when the first context on the device is
// created, and only that address is handed to kernels.  The partial-sum
// kernel writes every entry the apply kernel reads, so the array needs no
// clearing; calls that overlap on one GPU share it and race.
// ---------------------------------------------------------------------------
constexpr size_t kScratchDoubles = (size_t)1 << 18;   // 2 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the device's copy of `scratch`; device must be current
int static_scratch(int device, double** out) {
  SPK_REQUIRE(device >= 0 && device < kMaxDevices, SPK_EUNSUP,
              "device index above " + std::to_string(kMaxDevices - 1));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(SPK_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[device];
  return SPK_OK;
}
}  // namespace

// ===========================================================================
extern "C" {

const char* spk_last_error(void) { return g_err.c_str(); }

const char* spk_version(void) { return "speckle_hip 0.1 gfx950"; }

uint32_t spk_supported_sizes(void) {
  uint32_t m = 0;
  for (int l = 4; l <= 12; ++l) m |= 1u << l;
  return m;
}

int spk_ctx_create(int device, void* stream, spk_ctx** out) {
  SPK_REQUIRE(out, SPK_EINVAL, "null output pointer");
  *out = nullptr;
  int count = 0;
  SPK_HIP(hipGetDeviceCount(&count));
  SPK_REQUIRE(device >= 0 && device < count, SPK_EINVAL,
              "device " + std::to_string(device) + " not present (" + std::to_string(count) +
                  " devices)");
  SPK_HIP(hipSetDevice(device));
  double* ws = nullptr;   // the allocation-free entry point's workspace on this device
  int rc = static_scratch(device, &ws);
  if (rc) return rc;
  spk_ctx* c = new spk_ctx;
  c->device = device;
  c->stream = (hipStream_t)stream;
  *out = c;
  return SPK_OK;
}

int spk_ctx_destroy(spk_ctx* ctx) {
  if (!ctx) return SPK_OK;
  ctx_release(ctx);
  return SPK_OK;
}

int spk_ctx_set_stream(spk_ctx* ctx, void* stream) {
  SPK_REQUIRE(ctx, SPK_EINVAL, "null context");
  ctx->stream = (hipStream_t)stream;
  return SPK_OK;
}

int spk_sync(spk_ctx* ctx) {
  int rc = set_device(ctx);
  if (rc) return rc;
  if (ctx->stream) SPK_HIP(hipStreamSynchronize(ctx->stream));
  else SPK_HIP(hipDeviceSynchronize());
  return SPK_OK;
}

int spk_plan_bytes(int n, int batch, int shared, int options, size_t* bytes) {
  SPK_REQUIRE(bytes && is_pow2(n) && batch > 0, SPK_EINVAL, "bad plan size");
  SPK_REQUIRE((options & ~31) == 0, SPK_EINVAL, "unknown SPK_PB_* option bits");
  const size_t nn = (size_t)n * n;
  const size_t frames_ms = shared ? 1 : (size_t)batch;
  const size_t lay = n >= kDualMin ? 2 : 1;            // Phi layouts kept
  const size_t bands = (size_t)n / 2;                  // most bands: tile 2
  size_t b = 0;
  b += (1 + lay) * nn * batch * sizeof(c32);           // psi, Phi
  b += 2 * nn * frames_ms * sizeof(float);             // modulus, support
  b += 2 * (size_t)n * sizeof(c32);                    // twiddles
  b += (size_t)batch * (1 + 64) * sizeof(double) + 4 * sizeof(unsigned);
  if (options & SPK_PB_DM) b += lay * nn * batch * sizeof(c32);        // phib
  if (options & SPK_PB_SHRINKWRAP)                                     // sw maps, maxima
    b += 2 * nn * batch * sizeof(float) + (size_t)batch * bands * sizeof(float);
  if (options & SPK_PB_IPSF) b += nn * sizeof(c32) + lay * nn * batch * sizeof(c32);
  if (options & SPK_PB_SPECTRUM) b += 2 * nn * batch * sizeof(float) + 2 * 4096 * sizeof(float);
  // DM with a partially coherent divisor: the work buffers of its second projection
  if ((options & SPK_PB_DM) && (options & SPK_PB_IPSF)) b += lay * nn * batch * sizeof(c32);
  if ((options & SPK_PB_DM) && (options & SPK_PB_SPECTRUM)) b += 2 * nn * batch * sizeof(float);
  if (options & SPK_PB_TRACE) b += 64 * (size_t)batch * bands * sizeof(double);
  b += 64 * 1024;                                      // step-table staging
  *bytes = b;
  return SPK_OK;
}

int spk_plan_create(spk_ctx* ctx, int n, int batch, int shared, const spk_tune* tune,
                    spk_plan** out) {
  SPK_REQUIRE(out, SPK_EINVAL, "null output pointer");
  *out = nullptr;
  int rc = set_device(ctx);
  if (rc) return rc;
  SPK_REQUIRE(is_pow2(n), SPK_EINVAL, "n must be a power of two (got " + std::to_string(n) + ")");
  SPK_REQUIRE(batch >= 1, SPK_EINVAL, "batch must be >= 1");
  int tile = (tune && tune->tile) ? tune->tile : default_tile(n);
  int ept = (tune && tune->ept) ? tune->ept : default_ept(n, tile, batch);
  SPK_REQUIRE(size_supported(n, tile, ept), SPK_EUNSUP,
              "frame size " + std::to_string(n) + " with tile " + std::to_string(tile) +
                  " and " + std::to_string(ept) +
                  " values per thread not built (sizes 16..4096, tile 2, 4 or 8, 8 or 16 "
                  "values per thread, at most 1024 threads per workgroup, no tile 8 at 4096)");
  spk_plan* p = new spk_plan;
  p->ctx = ctx;
  ctx->refs++;
  p->n = n;
  p->batch = batch;
  p->shared = shared ? 1 : 0;
  p->tile = tile;
  p->ept = ept;
  p->bands = n / tile;
  p->persist = (tune && tune->persist > 0) ? 1 : 0;
  p->phib_g.assign(batch, 0.0f);
  // resident launches by default for one frame where they are built: 15-25 %
  // fewer microseconds per iteration than one launch per pass at 256^2,
  // 512^2 and 1024^2 on MI355X; the exchange of Phi by device-scope
  // accesses without fences is the default (the fenced forms are slower)
  const int res_mode = tune ? tune->resident : 0;
  p->resident = (res_mode != 2 && batch == 1 && resident_supported(n, tile, ept)) ? 1 : 0;
  p->xm = (tune && tune->exchange >= 1 && tune->exchange <= 3) ? tune->exchange - 1 : 2;
  p->polls = (tune && tune->barrier_polls > 0) ? (unsigned)tune->barrier_polls : (1u << 18);
  p->sw_fuse = !(tune && tune->sw_separate > 0);
  const size_t nn = (size_t)n * n;
  const size_t fms = shared ? 1 : (size_t)batch;
  if ((rc = dalloc(&p->psi, nn * batch)) || (rc = dalloc(&p->phi, nn * batch)) ||
      (p->dual() && (rc = dalloc(&p->phic, nn * batch))) ||
      (rc = dalloc(&p->mod, nn * fms)) || (rc = dalloc(&p->sup, nn * fms)) ||
      (rc = dalloc(&p->tw, (size_t)2 * n)) || (rc = dalloc(&p->summ2, (size_t)batch)) ||
      (rc = dalloc(&p->sumpart, (size_t)64 * batch)) || (rc = dalloc(&p->pbar, (size_t)4))) {
    spk_plan_destroy(p);
    return rc;
  }
  // forward twiddles W_n^m = exp(-2 pi i m / n), evaluated in FP64: the
  // exponent-indexed table, then the per-stage table of the plan's radix
  // schedule (the same values, k-major per stage; zero padded to n entries)
  std::vector<c32> tw((size_t)2 * n, c32{0.0f, 0.0f});
  auto w = [n](int m) {
    const double ang = -2.0 * M_PI * (double)m / (double)n;
    return c32{(float)std::cos(ang), (float)std::sin(ang)};
  };
  for (int m = 0; m < n; ++m) tw[m] = w(m);
  for (int st = 1; st < plan_stages(n, ept); ++st) {
    const int R = plan_radix(n, ept, st), off = staged_tw_offset(n, ept, st);
    int ns = 1;
    for (int i = 0; i < st; ++i) ns *= plan_radix(n, ept, i);
    for (int k = 0; k < ns; ++k)
      for (int r = 1; r < R; ++r)
        tw[(size_t)n + off + k * (R - 1) + (r - 1)] = w(staged_tw_exponent(n, ept, st, k, r));
  }
  // on the context's stream, so that creating a plan while other streams run
  // (serving grows its slot chunks) waits for this stream only
  hipStream_t cs = ctx->stream;
  hipError_t e = hipMemcpyAsync(p->tw, tw.data(), tw.size() * sizeof(c32), hipMemcpyHostToDevice, cs);
  if (e == hipSuccess) e = hipMemsetAsync(p->pbar, 0, 4 * sizeof(unsigned), cs);
  if (e == hipSuccess) e = cs ? hipStreamSynchronize(cs) : hipDeviceSynchronize();
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->upload_done, hipEventDisableTiming);
  if (e != hipSuccess) {
    spk_plan_destroy(p);
    return fail(SPK_EHIP, std::string("plan setup: ") + hipGetErrorString(e));
  }
  *out = p;
  return SPK_OK;
}

int spk_plan_info(const spk_plan* p, int* tile, int* ept, int* persist) {
  SPK_REQUIRE(p, SPK_EINVAL, "null plan");
  if (tile) *tile = p->tile;
  if (ept) *ept = p->ept;
  if (persist) *persist = p->persist;
  return SPK_OK;
}

int spk_plan_sw_counts(const spk_plan* p, long long* fused, long long* separate) {
  SPK_REQUIRE(p, SPK_EINVAL, "null plan");
  if (fused) *fused = p->sw_fused_count;
  if (separate) *separate = p->sw_separate_count;
  return SPK_OK;
}

int spk_plan_resident_counts(const spk_plan* p, long long* launches, long long* steps) {
  SPK_REQUIRE(p, SPK_EINVAL, "null plan");
  if (launches) *launches = p->res_launches;
  if (steps) *steps = p->res_steps;
  return SPK_OK;
}

int spk_plan_check(spk_plan* p) {
  SPK_REQUIRE(p, SPK_EINVAL, "null plan");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  if (p->ctx->stream) SPK_HIP(hipStreamSynchronize(p->ctx->stream));
  else SPK_HIP(hipDeviceSynchronize());
  unsigned h[3] = {0, 0, 0};
  SPK_HIP(hipMemcpy(h, p->pbar, sizeof(h), hipMemcpyDeviceToHost));
  if (h[2]) {
    SPK_HIP(hipMemset(p->pbar + 2, 0, sizeof(unsigned)));
    return fail(SPK_EHIP, "a grid barrier of a resident or persistent launch gave up "
                          "waiting; results of that launch are invalid");
  }
  return SPK_OK;
}

int spk_plan_destroy(spk_plan* p) {
  if (!p) return SPK_OK;
  if (p->ctx) {
    (void)hipSetDevice(p->ctx->device);
    if (p->ctx->stream) (void)hipStreamSynchronize(p->ctx->stream);
    else (void)hipDeviceSynchronize();
  }
  dfree(p->psi); dfree(p->phi); dfree(p->phic); dfree(p->phib); dfree(p->phibc);
  dfree(p->mod); dfree(p->sup);
  dfree(p->tw); dfree(p->summ2); dfree(p->sumpart); dfree(p->errp); dfree(p->pbar);
  dfree(p->ipsf); dfree(p->dbuf); dfree(p->dbufc);
  dfree(p->dbuf2); dfree(p->dbufc2); dfree(p->imap2); dfree(p->dmap2); dfree(p->rsync);
  dfree(p->spec); dfree(p->imap); dfree(p->dmap);
  dfree(p->coh_ac); dfree(p->coh_w); dfree(p->coh_cb); dfree(p->coh_m); dfree(p->coh_part);
  dfree(p->sw_tmp); dfree(p->sw_blur); dfree(p->sw_maxp); dfree(p->fftbuf);
  dfree(p->stage_d);
  dfree(p->cap_d);  // a capture that was never ended
  if (p->cap_h) (void)hipHostFree(p->cap_h);
  prof_clear(p, 0);
  prof_clear(p, 1);
  prof_clear(p, 2);
  if (p->stage_h) (void)hipHostFree(p->stage_h);
  if (p->upload_done) (void)hipEventDestroy(p->upload_done);
  if (p->ctx) ctx_release(p->ctx);
  delete p;
  return SPK_OK;
}

static int frame_range(spk_plan* p, int frame, bool per_frame_array, int* f0, int* nf) {
  // per_frame_array: the array holds one frame per plan frame (mod/sup of a
  // non-shared plan, psi of any plan)
  if (frame == -1) {
    *f0 = 0;
    *nf = per_frame_array ? p->batch : 1;
    return SPK_OK;
  }
  SPK_REQUIRE(frame >= 0 && frame < p->batch, SPK_EINVAL,
              "frame " + std::to_string(frame) + " out of range [0, " + std::to_string(p->batch) + ")");
  if (!per_frame_array) {
    SPK_REQUIRE(frame == 0, SPK_EINVAL, "a shared plan has one modulus/support (frame 0 or -1)");
  }
  *f0 = frame;
  *nf = 1;
  return SPK_OK;
}

int spk_set_modulus(spk_plan* p, int frame, const float* m) {
  SPK_REQUIRE(p && m, SPK_EINVAL, "null plan or modulus");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  int f0, nf;
  if ((rc = frame_range(p, frame, !p->shared, &f0, &nf))) return rc;
  hipStream_t st = p->ctx->stream;
  SPK_HIP(launch_pack_mod(p->n, p->tile, p->mod + (size_t)f0 * p->nn(), m, nf, st));
  const int nblk = 64;
  SPK_HIP(launch_sumsq(p->n, p->mod + (size_t)f0 * p->nn(), p->nn(), p->sumpart, nblk,
                       p->summ2 + f0, nf, st));
  if (p->shared) {
    for (int f = 1; f < p->batch; ++f)
      SPK_HIP(hipMemcpyAsync(p->summ2 + f, p->summ2, sizeof(double), hipMemcpyDeviceToDevice, st));
  }
  p->mod_set = true;
  return SPK_OK;
}

int spk_set_support(spk_plan* p, int frame, const float* s) {
  SPK_REQUIRE(p && s, SPK_EINVAL, "null plan or support");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  int f0, nf;
  if ((rc = frame_range(p, frame, !p->shared, &f0, &nf))) return rc;
  SPK_HIP(launch_pack_f32(p->n, p->tile, p->sup + (size_t)f0 * p->nn(), s, nf, p->ctx->stream));
  p->sup_set = true;
  phib_stale(p, f0, p->shared ? -1 : nf);   // a shared support serves every frame
  return SPK_OK;
}

int spk_get_support(spk_plan* p, int frame, float* s) {
  SPK_REQUIRE(p && s, SPK_EINVAL, "null plan or support");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  int f0, nf;
  if ((rc = frame_range(p, frame, !p->shared, &f0, &nf))) return rc;
  SPK_HIP(launch_unpack_f32(p->n, p->tile, s, p->sup + (size_t)f0 * p->nn(), nf, p->ctx->stream));
  return SPK_OK;
}

int spk_set_psi(spk_plan* p, int frame, const float* guess, int mult_support) {
  SPK_REQUIRE(p && guess, SPK_EINVAL, "null plan or guess");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  SPK_REQUIRE(!mult_support || p->sup_set, SPK_ESTATE, "support not set");
  int f0, nf;
  if ((rc = frame_range(p, frame, true, &f0, &nf))) return rc;
  const float* sup = mult_support ? p->sup + (size_t)(p->shared ? 0 : f0) * p->nn() : nullptr;
  SPK_HIP(launch_pack_c64(p->n, p->tile, p->psi + (size_t)f0 * p->nn(), (const c32*)guess, sup,
                          p->mstride(), nf, p->ctx->stream));
  // the row transform of the new iterate, for these frames only (the others
  // keep theirs)
  if ((rc = rowfft_frames(p, f0, nf))) return rc;
  p->psi_set = true;
  phib_stale(p, f0, nf);
  return SPK_OK;
}

int spk_get_psi(spk_plan* p, int frame, float* out) {
  SPK_REQUIRE(p && out, SPK_EINVAL, "null plan or output");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  int f0, nf;
  if ((rc = frame_range(p, frame, true, &f0, &nf))) return rc;
  SPK_HIP(launch_unpack_c64(p->n, p->tile, (c32*)out, p->psi + (size_t)f0 * p->nn(), 1.0f, nf,
                            p->ctx->stream));
  return SPK_OK;
}

int spk_set_ipsf(spk_plan* p, const float* ipsf) {
  SPK_REQUIRE(p, SPK_EINVAL, "null plan");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  if (!ipsf) {
    p->ipsf_on = false;
    return SPK_OK;
  }
  const size_t nn = (size_t)p->nn();
  if (!p->ipsf) {
    if ((rc = dalloc(&p->ipsf, nn)) || (rc = dalloc(&p->dbuf, nn * p->batch)) ||
        (p->dual() && (rc = dalloc(&p->dbufc, nn * p->batch))))
      return rc;
  }
  SPK_HIP(launch_pack_c64(p->n, p->tile, p->ipsf, (const c32*)ipsf, nullptr, 0, 1,
                          p->ctx->stream));
  p->ipsf_on = true;
  return SPK_OK;
}

int spk_set_spectrum(spk_plan* p, const float* rescales, const float* weights, int K) {
  SPK_REQUIRE(p, SPK_EINVAL, "null plan");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  if (!rescales || K == 0) {
    p->spec_on = false;
    return SPK_OK;
  }
  SPK_REQUIRE(weights && K > 0 && K <= 4096, SPK_EINVAL, "spectrum needs 1..4096 samples");
  for (int k = 0; k < K; ++k)
    SPK_REQUIRE(std::isfinite(rescales[k]) && rescales[k] != 0.0f && std::isfinite(weights[k]),
                SPK_EINVAL, "spectrum rescale factors must be finite and non-zero");
  const size_t nn = (size_t)p->nn();
  if (!p->imap) {
    if ((rc = dalloc(&p->imap, nn * p->batch)) || (rc = dalloc(&p->dmap, nn * p->batch)))
      return rc;
  }
  if (K > p->nspec) {
    SPK_HIP(hipStreamSynchronize(p->ctx->stream));
    dfree(p->spec);
    if ((rc = dalloc(&p->spec, (size_t)2 * K))) return rc;
  }
  std::vector<float> tab(2 * (size_t)K);
  std::memcpy(tab.data(), rescales, K * sizeof(float));
  std::memcpy(tab.data() + K, weights, K * sizeof(float));
  // synchronous: the host table goes out of scope at return
  SPK_HIP(hipMemcpyAsync(p->spec, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice,
                         p->ctx->stream));
  SPK_HIP(hipStreamSynchronize(p->ctx->stream));
  p->nspec = std::max(p->nspec, K);
  p->spec_k = K;
  p->spec_on = true;
  return SPK_OK;
}

// ---------------------------------------------------------------------------
// One spk_iterate call in three parts, shared with spk_group_iterate:
// iter_prepare checks the step table and stages the call's per-step tables
// (one upload), iter_step queues the launches of one step (or of the resident
// or persistent block that starts there), iter_finish queues the error
// finalisation and records the DM estimates' state.
// ---------------------------------------------------------------------------
namespace {

struct IterRun {
  spk_plan* p = nullptr;
  int nsteps = 0, B = 0, W = 0;
  bool per_frame = false;
  double* err_trace = nullptr;
  std::vector<Coef> coefs;     // [step][frame entry]
  std::vector<int> act;        // [step][frame entry]: not NOOP
  std::vector<char> any_dm, any_sw, emits, in_res, sw_fuse, noout, dm_all;
  std::vector<size_t> map_off, sw_off, swp_off, dm_off, dmall_off, res_off, res_sw_off, swt_off,
      block_off;
  std::vector<int> map_n, sw_n, sw_r, dm_n, dmall_n, res_end, res_n, res_fused, block_end;
  std::vector<float> dm_g, fr;
  std::vector<SwParam> sw_first;
  std::vector<std::vector<SwParam>> sw_prm;
  size_t coef_off = 0, act_off = 0;
  unsigned char* dbase = nullptr;
  unsigned char* dptr(size_t off) const { return dbase + off; }
};

// blocks: allow resident and persistent blocks (not in a group call, whose
// steps launch one pass at a time for every member)
int iter_prepare(spk_plan* p, const spk_step* steps, int nsteps, int per_frame,
                 double* err_trace, bool blocks, IterRun& r) {
  int rc = SPK_OK;
  SPK_REQUIRE(p->mod_set && p->sup_set && p->psi_set, SPK_ESTATE,
              "cant iterate before loading modulus, support and a seed");
  const int B = p->batch;
  const int W = per_frame ? B : 1;  // step entries per iteration
  r.p = p;
  r.nsteps = nsteps;
  r.B = B;
  r.W = W;
  r.per_frame = per_frame != 0;
  r.err_trace = err_trace;

  // coefficients for every (step, frame entry)
  auto& coefs = r.coefs;
  auto& act = r.act;
  auto& any_dm = r.any_dm;
  auto& any_sw = r.any_sw;
  coefs.assign((size_t)nsteps * W, Coef{});
  act.assign((size_t)nsteps * W, 0);
  any_dm.assign(nsteps, 0);
  any_sw.assign(nsteps, 0);
  for (int k = 0; k < nsteps; ++k) {
    for (int f = 0; f < W; ++f) {
      const spk_step& s = steps[(size_t)k * W + f];
      if ((rc = make_coef(s, &coefs[(size_t)k * W + f]))) return rc;
      act[(size_t)k * W + f] = s.alg != SPK_NOOP;
      if (s.alg == SPK_DM) any_dm[k] = 1;
      if (s.alg != SPK_NOOP && s.sw_sigma > 0.0f) any_sw[k] = 1;
    }
  }
  // next-step DM: emit phib from the row pass of step k (no shrinkwrap between)
  for (int k = 0; k + 1 < nsteps; ++k) {
    if (!any_dm[k + 1] || any_sw[k]) continue;
    for (int f = 0; f < W; ++f) {
      const spk_step& nx = steps[(size_t)(k + 1) * W + f];
      Coef& c = coefs[(size_t)k * W + f];
      if (nx.alg == SPK_DM && (c.flags & kActive)) {
        c.flags |= kEmitB;
        c.g_next = 1.0f / nx.beta;
      }
    }
  }
  bool dm_any = false;
  for (int k = 0; k < nsteps; ++k) dm_any |= any_dm[k] != 0;
  if (dm_any && (rc = ensure_phib(p))) return rc;
  if (dm_any && (p->ipsf_on || p->spec_on) && (rc = ensure_dm_pc(p))) return rc;
  bool sw_any = false;
  for (int k = 0; k < nsteps; ++k) sw_any |= any_sw[k] != 0;
  if (sw_any) {
    SPK_REQUIRE(!(p->shared && B > 1), SPK_EUNSUP,
                "shrinkwrap needs a plan with one support per frame");
    if ((rc = ensure_sw(p))) return rc;
  }

  // Per-frame schedules: per step the list of active frames (launch grids
  // cover only those), of frames with a shrinkwrap and their parameters.
  // Frame activity [step][frame] is staged in both modes: the error trace of
  // an idle frame is NaN.
  Stage stg;
  auto& map_off = r.map_off;
  auto& sw_off = r.sw_off;
  auto& swp_off = r.swp_off;
  auto& map_n = r.map_n;
  auto& sw_n = r.sw_n;
  auto& sw_r = r.sw_r;
  auto& dm_g = r.dm_g;
  auto& sw_first = r.sw_first;
  auto& sw_prm = r.sw_prm;
  map_off.assign(nsteps, 0);
  sw_off.assign(nsteps, 0);
  swp_off.assign(nsteps, 0);
  map_n.assign(nsteps, B);
  sw_n.assign(nsteps, 0);
  sw_r.assign(nsteps, 0);
  dm_g.assign(nsteps, 0.0f);
  sw_first.assign(nsteps, SwParam{});
  sw_prm.assign(per_frame ? nsteps : 0, {});
  if (err_trace) {
    std::vector<int> act_all((size_t)nsteps * B);
    for (int k = 0; k < nsteps; ++k)
      for (int f = 0; f < B; ++f) act_all[(size_t)k * B + f] = act[(size_t)k * W + (per_frame ? f : 0)];
    r.act_off = stg.add(act_all.data(), act_all.size());
  }
  // DM steps: one beta per step (the dm1 map of the refresh kernel is shared)
  for (int k = 0; k < nsteps; ++k) {
    if (!any_dm[k]) continue;
    for (int f = 0; f < W; ++f) {
      const spk_step& sp = steps[(size_t)k * W + f];
      if (sp.alg != SPK_DM) continue;
      if (dm_g[k] == 0.0f) dm_g[k] = 1.0f / sp.beta;
      SPK_REQUIRE(1.0f / sp.beta == dm_g[k], SPK_EUNSUP,
                  "per-frame DM steps need one beta per step");
    }
  }
  if (per_frame) {
    std::vector<int> lst;
    std::vector<SwParam> prm;
    for (int k = 0; k < nsteps; ++k) {
      const size_t row = (size_t)k * W;
      lst.clear();
      for (int f = 0; f < B; ++f)
        if (act[row + f]) lst.push_back(f);
      map_n[k] = (int)lst.size();
      map_off[k] = stg.add(lst.data(), lst.size());
      if (any_sw[k]) {
        lst.clear();
        prm.clear();
        for (int f = 0; f < B; ++f) {
          const spk_step& sp = steps[row + f];
          if (sp.alg == SPK_NOOP || !(sp.sw_sigma > 0.0f)) continue;
          SwParam q;
          if ((rc = make_swparam(p->n, sp.sw_sigma, sp.sw_tau, &q))) return rc;
          lst.push_back(f);
          prm.push_back(q);
          sw_r[k] = std::max(sw_r[k], q.w.radius);
        }
        sw_n[k] = (int)lst.size();
        sw_off[k] = stg.add(lst.data(), lst.size());
        swp_off[k] = stg.add(prm.data(), prm.size());
        sw_prm[k] = prm;
      }
    }
  } else {
    for (int k = 0; k < nsteps; ++k) {
      if (!act[k]) map_n[k] = 0;   // the whole step is idle
      if (any_sw[k]) {
        if ((rc = make_swparam(p->n, steps[k].sw_sigma, steps[k].sw_tau, &sw_first[k])))
          return rc;
        sw_r[k] = sw_first[k].w.radius;
      }
    }
  }
  // The DM estimate phib of a frame must equal rowFFT(dm1(psi)) with the
  // step's beta when the frame's DM step runs.  It is fresh when the
  // frame's previous row pass emitted it for this beta (kEmitB) and nothing
  // changed psi or the support since (an idle step changes neither);
  // otherwise the refresh kernel recomputes it for exactly those frames
  // before the step (frame list dm_off[k], dm_n[k] entries; dm_all: every
  // frame, no list).  fr tracks the state through the call.
  auto& dm_off = r.dm_off;
  auto& dm_n = r.dm_n;
  auto& dm_all = r.dm_all;
  auto& fr = r.fr;
  dm_off.assign(nsteps, 0);
  dm_n.assign(nsteps, 0);
  dm_all.assign(nsteps, 0);
  fr = p->phib_g;
  // every DM frame of a DM step (per-frame tables): the frames of the second
  // partially coherent projection's row-direction kernels
  r.dmall_off.assign(nsteps, 0);
  r.dmall_n.assign(nsteps, 0);
  if (per_frame && (p->ipsf_on || p->spec_on)) {
    std::vector<int> lst;
    for (int k = 0; k < nsteps; ++k) {
      if (!any_dm[k]) continue;
      lst.clear();
      for (int f = 0; f < B; ++f)
        if (steps[(size_t)k * W + f].alg == SPK_DM) lst.push_back(f);
      r.dmall_n[k] = (int)lst.size();
      r.dmall_off[k] = stg.add(lst.data(), lst.size());
    }
  }
  {
    std::vector<int> lst;
    for (int k = 0; k < nsteps; ++k) {
      const size_t row = (size_t)k * W;
      if (any_dm[k]) {
        lst.clear();
        for (int f = 0; f < B; ++f) {
          const spk_step& sp = steps[row + (per_frame ? f : 0)];
          if (sp.alg == SPK_DM && fr[f] != dm_g[k]) lst.push_back(f);
        }
        dm_n[k] = (int)lst.size();
        dm_all[k] = dm_n[k] == B;
        if (dm_n[k] && !dm_all[k]) dm_off[k] = stg.add(lst.data(), lst.size());
      }
      for (int f = 0; f < B; ++f) {
        const Coef& c = coefs[row + (per_frame ? f : 0)];
        if (c.flags & kActive) fr[f] = (c.flags & kEmitB) ? c.g_next : 0.0f;
      }
    }
  }
  auto& emits = r.emits;
  emits.assign(nsteps, 0);
  for (int k = 0; k < nsteps; ++k)
    for (int f = 0; f < W; ++f) emits[k] |= (coefs[(size_t)k * W + f].flags & kEmitB) != 0;
  // Resident blocks (k_resident, one frame): runs of steps with no DM work
  // (no DM step, no emission).  A shrinkwrap after a step of the run runs
  // inside the block (row blur in that step's row phase, column blur in the
  // next active step's column phase, threshold in its row phase) when its
  // blur radius is within the fused window and the run goes on; otherwise
  // the run ends at that step and the shrinkwrap runs as stand-alone kernels
  // after the block.  Idle steps change nothing and are left out of the
  // block's table.  res_end[k0] = k1 for a block of steps [k0, k1).
  auto& res_end = r.res_end;
  auto& in_res = r.in_res;
  res_end.assign(nsteps, 0);
  r.res_n.assign(nsteps, 0);
  r.res_fused.assign(nsteps, 0);
  r.res_off.assign(nsteps, 0);
  r.res_sw_off.assign(nsteps, 0);
  in_res.assign(nsteps, 0);
  if (blocks && p->resident && B == 1 && !p->ipsf_on && !p->spec_on && !p->persist &&
      !p->capturing) {
    auto fits = [&](int j) { return !any_dm[j] && !emits[j]; };
    std::vector<RStep> rs;
    std::vector<SwParam> sp;
    for (int k = 0; k < nsteps;) {
      if (!fits(k)) { ++k; continue; }
      rs.clear();
      sp.clear();
      int k1 = k, fused = 0, pending = -1;   // pending: sw index for the next active step
      while (k1 < nsteps && fits(k1)) {
        const int j = k1++;
        if (!act[(size_t)j * W]) continue;
        RStep rsj{coefs[(size_t)j * W], err_trace ? 1 : 0, j, -1, -1};
        rsj.c.flags &= ~(kSwRow | kSwFin);
        if (pending >= 0) {
          rsj.c.flags |= kSwFin;
          rsj.swf = pending;
          pending = -1;
        }
        if (any_sw[j]) {
          // fused when a later active step of the run takes its second half
          int nx = j + 1;
          while (nx < nsteps && fits(nx) && !act[(size_t)nx * W]) ++nx;
          const bool goes_on = nx < nsteps && fits(nx);
          if (!goes_on || sw_r[j] > kSwFuseRmax || !p->sw_fuse) {
            rs.push_back(rsj);
            break;                        // stand-alone shrinkwrap after the block
          }
          rsj.c.flags |= kSwRow;
          rsj.swr = (int)sp.size();
          pending = rsj.swr;
          sp.push_back(per_frame ? sw_prm[j][0] : sw_first[j]);   // the one frame's
          ++fused;
        }
        rs.push_back(rsj);
      }
      if (rs.size() >= 2) {
        res_end[k] = k1;
        r.res_n[k] = (int)rs.size();
        r.res_fused[k] = fused;
        r.res_off[k] = stg.add(rs.data(), rs.size());
        if (!sp.empty()) r.res_sw_off[k] = stg.add(sp.data(), sp.size());
        for (int j = k; j < k1; ++j) in_res[j] = 1;
      }
      k = k1;
    }
  }
  // Fused shrinkwraps: the shrinkwrap after step k runs inside the passes of
  // steps k and k+1 (row blur in row pass k, column blur in column pass k+1,
  // threshold in row pass k+1, the only reader of the support in between)
  // when both steps are single projections without DM, every frame with the
  // shrinkwrap is active at k+1 and the blur radius is within the fused
  // form's window.  Parameter tables by frame (per_frame) or one entry.
  auto& sw_fuse = r.sw_fuse;
  sw_fuse.assign(nsteps, 0);
  r.swt_off.assign(nsteps, 0);
  if (p->sw_fuse && !p->persist && !p->ipsf_on && !p->spec_on) {
    for (int k = 0; k + 1 < nsteps; ++k) {
      if (!any_sw[k] || sw_r[k] > kSwFuseRmax || any_dm[k] || any_dm[k + 1] || emits[k + 1] ||
          in_res[k] || in_res[k + 1])
        continue;
      const size_t row = (size_t)k * W, nrow = (size_t)(k + 1) * W;
      std::vector<SwParam> tab(per_frame ? B : 1);
      if (per_frame) {
        bool ok = true;
        int j = 0;
        for (int f = 0; f < B; ++f) {
          const spk_step& sp = steps[row + f];
          if (sp.alg == SPK_NOOP || !(sp.sw_sigma > 0.0f)) continue;
          ok &= act[nrow + f] != 0;
          tab[f] = sw_prm[k][j++];
        }
        if (!ok) continue;
        for (int f = 0; f < B; ++f) {
          const spk_step& sp = steps[row + f];
          if (sp.alg == SPK_NOOP || !(sp.sw_sigma > 0.0f)) continue;
          coefs[row + f].flags |= kSwRow;
          coefs[nrow + f].flags |= kSwFin;
        }
      } else {
        if (!act[nrow]) continue;
        tab[0] = sw_first[k];
        coefs[row].flags |= kSwRow;
        coefs[nrow].flags |= kSwFin;
      }
      sw_fuse[k] = 1;
      r.swt_off[k] = stg.add(tab.data(), tab.size());
    }
  }
  if (per_frame) r.coef_off = stg.add(coefs.data(), coefs.size());
  // Steps whose updates all ignore psi_out (ER with numpy semantics): the
  // column pass only forms the error partials (or is skipped without an
  // error trace) and the row pass skips the row IFFT of Phi.  Not with DM
  // work, a fused shrinkwrap or a partially coherent projection in the step.
  auto& noout = r.noout;
  noout.assign(nsteps, 0);
  if (!p->ipsf_on && !p->spec_on) {
    for (int k = 0; k < nsteps; ++k) {
      if (any_dm[k] || emits[k] || sw_fuse[k] || (k > 0 && sw_fuse[k - 1])) continue;
      bool all = true, any = false;
      for (int f = 0; f < W; ++f) {
        const Coef& c = coefs[(size_t)k * W + f];
        if (!(c.flags & kActive)) continue;
        any = true;
        all &= (c.flags & kNoOut) != 0;
      }
      noout[k] = any && all;
    }
  }
  // Persistent blocks: runs of at least two steps with no DM step and no DM
  // successor, each run ending at a step with a shrinkwrap (which follows the
  // block).  block_end[k0] = k1 for a block of steps [k0, k1).
  r.block_end.assign(nsteps, 0);
  r.block_off.assign(nsteps, 0);
  if (blocks && p->persist && !p->ipsf_on && !p->spec_on) {
    auto fits = [&](int j) { return !any_dm[j] && !emits[j]; };
    for (int k = 0; k < nsteps;) {
      if (!fits(k)) { ++k; continue; }
      int k1 = k;
      while (k1 < nsteps && fits(k1)) {
        ++k1;
        if (any_sw[k1 - 1]) break;
      }
      if (k1 - k >= 2) {
        std::vector<PStep> ps(k1 - k);
        for (int j = k; j < k1; ++j) {
          PStep& q = ps[j - k];
          q.c = coefs[(size_t)j * W];
          q.tab_off = per_frame ? (long long)(r.coef_off + (size_t)j * W * sizeof(Coef)) : -1;
          q.fmap_off = per_frame ? (long long)map_off[j] : -1;
          q.nact = map_n[j];
          q.pad_ = 0;
        }
        r.block_end[k] = k1;
        r.block_off[k] = stg.add(ps.data(), ps.size());
      }
      k = k1;
    }
  }
  if ((rc = upload_stage(p, stg))) return rc;
  r.dbase = p->stage_cur;

  // error partials [step][frame][band]
  if (err_trace) {
    const size_t need = (size_t)nsteps * B * p->bands;
    if (need > p->errp_cap) {
      SPK_HIP(hipStreamSynchronize(p->ctx->stream));
      dfree(p->errp);
      p->errp_cap = 0;
      if ((rc = dalloc(&p->errp, need))) return rc;
      p->errp_cap = need;
    }
  }
  // until every launch is queued the DM estimates count as stale (a failed
  // call leaves them to be recomputed)
  std::fill(p->phib_g.begin(), p->phib_g.end(), 0.0f);
  return SPK_OK;
}

int iter_shrinkwrap_after(IterRun& r, int k) {
  spk_plan* p = r.p;
  hipStream_t st = p->ctx->stream;
  if (r.per_frame) {
    SPK_HIP(launch_shrinkwrap(p->n, p->tile, p->psi, p->sup, p->mstride(), p->sw_tmp, p->sw_blur,
                              p->sw_maxp, (const SwParam*)r.dptr(r.swp_off[k]), SwParam{},
                              (const int*)r.dptr(r.sw_off[k]), 0, r.sw_n[k], r.sw_r[k], st));
  } else {
    SPK_HIP(launch_shrinkwrap(p->n, p->tile, p->psi, p->sup, p->mstride(), p->sw_tmp, p->sw_blur,
                              p->sw_maxp, nullptr, r.sw_first[k], nullptr, 0, r.B, r.sw_r[k], st));
  }
  p->sw_separate_count++;
  return SPK_OK;
}

// The pass arguments of step k (tables, frame list, error partials, fused
// shrinkwrap tables) and its fused-shrinkwrap bits: 1 = this step's row blur,
// 2 = the previous step's column blur and threshold.
PassArgs iter_step_args(const IterRun& r, int k, int* swbits) {
  spk_plan* p = r.p;
  PassArgs a = base_args(p);
  const size_t row = (size_t)k * r.W;
  if (r.per_frame) {
    a.tab = (const Coef*)r.dptr(r.coef_off) + row;
    a.fmap = (const int*)r.dptr(r.map_off[k]);
  } else {
    a.tab = nullptr;
    a.fmap = nullptr;
    a.c = r.coefs[row];
  }
  a.errp = r.err_trace ? p->errp + (size_t)k * r.B * p->bands : nullptr;
  const int sb = (r.sw_fuse[k] ? 1 : 0) | (k > 0 && r.sw_fuse[k - 1] ? 2 : 0);
  // streaming row-pass stores for one 1024^2 frame in 4- or 8-wide tiles
  // (whole 128-byte lines per wavefront; measured -3 % per iteration there,
  // +1..5 % at 256^2 and 512^2)
  a.nt_store = (r.map_n[k] == 1 && p->n == 1024 && p->tile >= 4) ? 1 : 0;
  a.swr_tab = (sb & 1) ? (const SwParam*)r.dptr(r.swt_off[k]) : nullptr;
  a.swf_tab = (sb & 2) ? (const SwParam*)r.dptr(r.swt_off[k - 1]) : nullptr;
  a.sw_pf = r.per_frame ? 1 : 0;
  *swbits = sb;
  return a;
}

// after the row pass of step k: the stand-alone shrinkwrap when it did not
// run fused, and the counters
int iter_step_after(IterRun& r, int k) {
  if (r.sw_fuse[k]) r.p->sw_fused_count++;
  if (r.any_sw[k] && !r.sw_fuse[k]) return iter_shrinkwrap_after(r, k);
  return SPK_OK;
}

// Queue step k (or the block that starts there); *next = the step after it.
int iter_step(IterRun& r, int k, int* next) {
  spk_plan* p = r.p;
  hipStream_t st = p->ctx->stream;
  const int B = r.B;
  int rc = SPK_OK;
  *next = k + 1;
  if (r.res_end[k]) {
    const int k1 = r.res_end[k];
    if (!p->rsync && (rc = dalloc(&p->rsync, (size_t)p->bands * 32))) return rc;
    if ((rc = prof_arm(p, 2))) return rc;
    PassArgs ra = base_args(p);
    ra.swr_tab = r.res_fused[k] ? (const SwParam*)r.dptr(r.res_sw_off[k]) : nullptr;
    SPK_HIP(launch_resident(p->n, p->tile, p->ept, p->xm, ra, (const RStep*)r.dptr(r.res_off[k]),
                            r.res_n[k], r.err_trace ? p->errp : nullptr, (long long)B * p->bands,
                            p->rsync, (int*)(p->pbar + 1), st));
    p->res_launches++;
    p->res_steps += r.res_n[k];
    p->sw_fused_count += r.res_fused[k];
    if (r.any_sw[k1 - 1] && (rc = iter_shrinkwrap_after(r, k1 - 1))) return rc;
    *next = k1;
    return SPK_OK;
  }
  if (r.block_end[k]) {
    const int k1 = r.block_end[k];
    if (!p->pgrid) SPK_HIP(persist_grid(p->n, p->tile, p->ept, &p->pgrid));
    SPK_REQUIRE(p->pgrid > 0, SPK_EUNSUP, "persistent kernel does not fit on this device");
    const int grid = (int)std::min<long long>(p->pgrid, (long long)B * p->bands);
    PassArgs pa = base_args(p);
    SPK_HIP(launch_persist(p->n, p->tile, p->ept, pa, r.dbase,
                           (const PStep*)r.dptr(r.block_off[k]), k1 - k,
                           r.err_trace ? p->errp + (size_t)k * B * p->bands : nullptr,
                           (long long)B * p->bands, p->pbar, (int*)(p->pbar + 1), grid, st));
    if (r.any_sw[k1 - 1] && (rc = iter_shrinkwrap_after(r, k1 - 1))) return rc;
    *next = k1;
    return SPK_OK;
  }
  if (r.map_n[k] == 0) return SPK_OK;     // every frame idle this step
  int swbits = 0;
  PassArgs a = iter_step_args(r, k, &swbits);
  const int nact = r.map_n[k];
  if (r.any_dm[k] && r.dm_n[k]) {
    // phib = rowFFT((s - gamma_m (1-s)) psi), dm1 (phasing_dm1.cl:16), for
    // the DM frames of the step whose phib is not fresh
    SPK_HIP(launch_rowfft(p->n, p->tile, p->ept, 2, p->psi, p->phib_in(), p->sup, p->mstride(),
                          r.dm_g[k], p->tw,
                          r.dm_all[k] ? nullptr : (const int*)r.dptr(r.dm_off[k]), r.dm_n[k], st));
  }
  if (p->ipsf_on || p->spec_on) {
    // partially coherent projection.  ipsf: column pass A (|Psi|^2 ->
    // column FFT) and the row convolution with the ipsf.  spectrum: the
    // squared divisor as a map (coherent: MODE 6; after the ipsf: MODE 7),
    // its spectral blur.  Then column pass B (divisor, replacement,
    // error, column IFFT).
    // With DM the same runs on phib (grid z = 1, its own work buffers) for
    // the DM frames: t = P_M[dm1 psi] with the divisor of dm1 psi
    // (phasing.py:1293-1313 calls _fourier_constraint on it).
    const int zd = r.any_dm[k] ? 2 : 1;
    const int* dmmap = r.per_frame ? (const int*)r.dptr(r.dmall_off[k]) : nullptr;
    const int ndm = r.per_frame ? r.dmall_n[k] : B;
    if (p->ipsf_on) {
      SPK_HIP(launch_colpass(p->n, p->tile, p->ept, 4, a, nact, zd, st));
      SPK_HIP(launch_rowconv(p->n, p->tile, p->ept, p->dbuf, p->dual() ? p->dbufc : p->dbuf,
                             p->ipsf, p->tw, a.fmap, nact, st));
      if (zd == 2)
        SPK_HIP(launch_rowconv(p->n, p->tile, p->ept, p->dbuf2, p->dual() ? p->dbufc2 : p->dbuf2,
                               p->ipsf, p->tw, dmmap, ndm, st));
    }
    if (p->spec_on) {
      SPK_HIP(launch_colpass(p->n, p->tile, p->ept, p->ipsf_on ? 7 : 6, a, nact, zd, st));
      SPK_HIP(launch_rescale(p->n, p->tile, p->imap, p->dmap, p->spec, p->spec_k, a.fmap, nact,
                             st));
      if (zd == 2)
        SPK_HIP(launch_rescale(p->n, p->tile, p->imap2, p->dmap2, p->spec, p->spec_k, dmmap, ndm,
                               st));
    }
    if ((rc = prof_arm(p, 0))) return rc;
    SPK_HIP(launch_colpass(p->n, p->tile, p->ept, p->spec_on ? 8 : 5, a, nact, zd, st));
  } else if (r.noout[k]) {
    // not profiled: the pass timings cover the full passes
    if (a.errp) SPK_HIP(launch_colpass(p->n, p->tile, p->ept, 10, a, nact, 1, st));
  } else {
    if ((rc = prof_arm(p, 0))) return rc;
    SPK_HIP(launch_colpass(p->n, p->tile, p->ept, (swbits & 2) ? 9 : 0, a, nact,
                           r.any_dm[k] ? 2 : 1, st));
  }
  a.errp = nullptr;
  if (!r.noout[k] && (rc = prof_arm(p, 1))) return rc;
  SPK_HIP(launch_rowpass(p->n, p->tile, p->ept, r.any_dm[k] != 0, r.emits[k] != 0, swbits,
                         r.noout[k] != 0, a, nact, st));
  return iter_step_after(r, k);
}

int iter_finish(IterRun& r) {
  spk_plan* p = r.p;
  if (r.err_trace) {
    SPK_HIP(launch_err_finalize(p->errp, r.nsteps, r.B, p->bands, p->summ2, r.err_trace,
                                (const int*)r.dptr(r.act_off), p->ctx->stream));
  }
  p->phib_g = r.fr;
  return SPK_OK;
}

}  // namespace

int spk_iterate(spk_plan* p, const spk_step* steps, int nsteps, int per_frame,
                double* err_trace) {
  SPK_REQUIRE(p, SPK_EINVAL, "null plan");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  SPK_REQUIRE(nsteps >= 0, SPK_EINVAL, "nsteps must be >= 0");
  if (nsteps == 0) return SPK_OK;
  SPK_REQUIRE(steps, SPK_EINVAL, "null steps");
  IterRun r;
  if ((rc = iter_prepare(p, steps, nsteps, per_frame, err_trace, true, r))) return rc;
  for (int k = 0; k < nsteps;) {
    int next = k + 1;
    if ((rc = iter_step(r, k, &next))) return rc;
    k = next;
  }
  return iter_finish(r);
}

// ---------------------------------------------------------------------------
// Mixed-size iteration (spk_group_*): the member plans' spk_iterate calls
// with their single-projection passes merged, per step, into one column-pass
// and one row-pass launch for all members.
// ---------------------------------------------------------------------------
struct spk_group {
  spk_ctx* ctx = nullptr;
  unsigned char* tab_d = nullptr;   // per-call pass-argument tables
  unsigned char* tab_h = nullptr;
  size_t cap = 0;
  hipEvent_t upload_done = nullptr;
  bool upload_pending = false;
  long long launches = 0, member_passes = 0;
};

int spk_group_shape(int n, int* tile, int* ept) {
  SPK_REQUIRE(tile && ept, SPK_EINVAL, "null argument");
  SPK_REQUIRE(group_tile(n) > 0, SPK_EUNSUP,
              "mixed-size launches take frames of 256, 512 and 1024 (got " + std::to_string(n) +
                  ")");
  *tile = group_tile(n);
  *ept = kGroupEpt;
  return SPK_OK;
}

int spk_group_create(spk_ctx* ctx, spk_group** out) {
  SPK_REQUIRE(ctx && out, SPK_EINVAL, "null argument");
  int rc = set_device(ctx);
  if (rc) return rc;
  spk_group* g = new spk_group;
  g->ctx = ctx;
  if (hipEventCreateWithFlags(&g->upload_done, hipEventDisableTiming) != hipSuccess) {
    delete g;
    return fail(SPK_EHIP, "hipEventCreate failed");
  }
  ctx->refs++;
  *out = g;
  return SPK_OK;
}

int spk_group_destroy(spk_group* g) {
  if (!g) return SPK_OK;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  dfree(g->tab_d);
  if (g->tab_h) (void)hipHostFree(g->tab_h);
  if (g->upload_done) (void)hipEventDestroy(g->upload_done);
  ctx_release(g->ctx);
  delete g;
  return SPK_OK;
}

int spk_group_counts(const spk_group* g, long long* launches, long long* member_passes) {
  SPK_REQUIRE(g && launches && member_passes, SPK_EINVAL, "null argument");
  *launches = g->launches;
  *member_passes = g->member_passes;
  return SPK_OK;
}

int spk_group_iterate(spk_group* g, spk_plan* const* plans, int count,
                      const spk_step* const* steps, int nsteps, const int* per_frame,
                      double* const* err_trace) {
  SPK_REQUIRE(g, SPK_EINVAL, "null group");
  int rc = set_device(g->ctx);
  if (rc) return rc;
  SPK_REQUIRE(count >= 1 && count <= kGroupMax, SPK_EINVAL,
              "a mixed-size call takes 1 to " + std::to_string(kGroupMax) + " plans");
  SPK_REQUIRE(plans && steps, SPK_EINVAL, "null plans or steps");
  SPK_REQUIRE(nsteps >= 0, SPK_EINVAL, "nsteps must be >= 0");
  for (int m = 0; m < count; ++m) {
    const spk_plan* p = plans[m];
    SPK_REQUIRE(p && steps[m], SPK_EINVAL, "null plan or step table");
    SPK_REQUIRE(p->ctx == g->ctx, SPK_EINVAL, "every plan of a mixed-size call uses the group's context");
    SPK_REQUIRE(p->tile == group_tile(p->n) && p->ept == kGroupEpt, SPK_EUNSUP,
                "plan of size " + std::to_string(p->n) +
                    " is not in the mixed-size shape (spk_group_shape)");
    SPK_REQUIRE(!p->persist && !p->capturing, SPK_EUNSUP,
                "mixed-size calls take neither persistent plans nor graph capture");
    for (int j = 0; j < m; ++j) SPK_REQUIRE(plans[j] != p, SPK_EINVAL, "a plan appears twice");
  }
  if (nsteps == 0) return SPK_OK;
  // every member's table is checked and staged before the first launch
  std::vector<IterRun> runs(count);
  for (int m = 0; m < count; ++m)
    if ((rc = iter_prepare(plans[m], steps[m], nsteps, per_frame ? per_frame[m] : 0,
                           err_trace ? err_trace[m] : nullptr, false, runs[m])))
      return rc;
  auto groupable = [&](int m, int k) {
    const IterRun& r = runs[m];
    return r.map_n[k] > 0 && !r.any_dm[k] && !r.emits[k] && !r.p->ipsf_on && !r.p->spec_on;
  };
  // the pass arguments of every merged (step, member), column then row
  // copies (the row pass writes no error partials), in one upload
  std::vector<PassArgs> tab;
  std::vector<GroupPass> col(nsteps), row(nsteps);
  std::vector<int> colmode(nsteps, 0), swb(nsteps, 0);
  std::vector<size_t> base(nsteps, 0);
  for (int k = 0; k < nsteps; ++k) {
    GroupPass& gc = col[k];
    gc.count = 0;
    gc.start[0] = 0;
    base[k] = tab.size();
    std::vector<PassArgs> rows;
    for (int m = 0; m < count; ++m) {
      if (!groupable(m, k)) continue;
      int sb = 0;
      PassArgs a = iter_step_args(runs[m], k, &sb);
      a.nt_store = 0;
      swb[k] |= sb;
      if (sb & 2) colmode[k] = 9;
      const int j = gc.count++;
      gc.n[j] = runs[m].p->n;
      gc.start[j + 1] = gc.start[j] + runs[m].map_n[k] * runs[m].p->bands;
      tab.push_back(a);
      a.errp = nullptr;
      rows.push_back(a);
    }
    tab.insert(tab.end(), rows.begin(), rows.end());
    row[k] = gc;
  }
  if (!tab.empty()) {
    const size_t need = tab.size() * sizeof(PassArgs);
    if (g->upload_pending) SPK_HIP(hipEventSynchronize(g->upload_done));
    g->upload_pending = false;
    if (need > g->cap) {
      SPK_HIP(hipStreamSynchronize(g->ctx->stream));
      dfree(g->tab_d);
      if (g->tab_h) (void)hipHostFree(g->tab_h);
      g->tab_h = nullptr;
      g->cap = 0;
      const size_t cap = std::max<size_t>(need, 64 * 1024);
      if ((rc = dalloc(&g->tab_d, cap))) return rc;
      SPK_HIP(hipHostMalloc((void**)&g->tab_h, cap, hipHostMallocDefault));
      g->cap = cap;
    }
    std::memcpy(g->tab_h, tab.data(), need);
    SPK_HIP(hipMemcpyAsync(g->tab_d, g->tab_h, need, hipMemcpyHostToDevice, g->ctx->stream));
    SPK_HIP(hipEventRecord(g->upload_done, g->ctx->stream));
    g->upload_pending = true;
  }
  const PassArgs* dtab = reinterpret_cast<const PassArgs*>(g->tab_d);
  hipStream_t st = g->ctx->stream;
  for (int k = 0; k < nsteps; ++k) {
    // members whose step does not merge (DM work, a partially coherent
    // projection) run it with their own launches; an idle step runs nothing
    for (int m = 0; m < count; ++m) {
      if (groupable(m, k)) continue;
      int next = k + 1;
      if ((rc = iter_step(runs[m], k, &next))) return rc;
    }
    if (col[k].count == 0) continue;
    col[k].args = dtab + base[k];
    row[k].args = dtab + base[k] + col[k].count;
    SPK_HIP(launch_colpass_group(colmode[k], col[k], st));
    SPK_HIP(launch_rowpass_group(swb[k], row[k], st));
    g->launches += 2;
    g->member_passes += 2 * col[k].count;
    for (int m = 0; m < count; ++m)
      if (groupable(m, k) && (rc = iter_step_after(runs[m], k))) return rc;
  }
  for (int m = 0; m < count; ++m)
    if ((rc = iter_finish(runs[m]))) return rc;
  return SPK_OK;
}

int spk_shrinkwrap(spk_plan* p, int frame, float sigma, float tau) {
  SPK_REQUIRE(p, SPK_EINVAL, "null plan");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  SPK_REQUIRE(p->psi_set && p->sup_set, SPK_ESTATE, "shrinkwrap needs psi and support");
  int f0, nf;
  if ((rc = frame_range(p, frame, true, &f0, &nf))) return rc;
  return do_shrinkwrap(p, f0, nf, sigma, tau);
}

int spk_copy_roi(spk_plan* p, int frame, float* dst, int slots, int slot, int r0, int c0,
                 int rows, int cols) {
  SPK_REQUIRE(p && dst, SPK_EINVAL, "null plan or destination");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  SPK_REQUIRE(frame >= 0 && frame < p->batch, SPK_EINVAL, "frame out of range");
  SPK_REQUIRE(slot >= 0 && slot < slots, SPK_EINVAL,
              "savebuffer slot " + std::to_string(slot) + " out of range [0, " +
                  std::to_string(slots) + ")");
  SPK_REQUIRE(rows > 0 && cols > 0 && r0 >= 0 && c0 >= 0 && r0 + rows <= p->n &&
                  c0 + cols <= p->n,
              SPK_EINVAL, "ROI outside the frame");
  c32* d = (c32*)dst + (size_t)slot * rows * cols;
  SPK_HIP(launch_copy_roi(p->n, p->tile, p->psi + (size_t)frame * p->nn(), d, r0, c0, rows,
                          cols, p->ctx->stream));
  return SPK_OK;
}

int spk_fft2(spk_plan* p, const float* in, float* out, int count, int inverse) {
  SPK_REQUIRE(p && in && out, SPK_EINVAL, "null plan or buffer");
  SPK_REQUIRE(count >= 1, SPK_EINVAL, "count must be >= 1");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  // two work areas: tiled, then column-band-major after the row transform
  if ((size_t)count > p->fftbuf_frames) {
    SPK_HIP(hipStreamSynchronize(p->ctx->stream));
    dfree(p->fftbuf);
    p->fftbuf_frames = 0;
    if ((rc = dalloc(&p->fftbuf, 2 * (size_t)count * p->nn()))) return rc;
    p->fftbuf_frames = count;
  }
  hipStream_t st = p->ctx->stream;
  c32* tiled = p->fftbuf;
  c32* colband = p->dual() ? p->fftbuf + (size_t)count * p->nn() : tiled;
  SPK_HIP(launch_pack_c64(p->n, p->tile, tiled, (const c32*)in, nullptr, 0, count, st));
  SPK_HIP(launch_rowfft(p->n, p->tile, p->ept, inverse ? 1 : 0, tiled, colband, nullptr, 0, 0.0f,
                        p->tw, nullptr, count, st));
  PassArgs a = base_args(p);
  a.phic = colband;
  a.phi = tiled;
  a.c.flags = kActive;
  SPK_HIP(launch_colpass(p->n, p->tile, p->ept, inverse ? 3 : 2, a, count, 1, st));
  const float scl = inverse ? (float)(1.0 / ((double)p->n * p->n)) : 1.0f;
  SPK_HIP(launch_unpack_c64(p->n, p->tile, (c32*)out, tiled, scl, count, st));
  return SPK_OK;
}

int spk_align_global_phase(spk_ctx* ctx, float* data, int frames, int rows, int cols) {
  SPK_REQUIRE(data && frames >= 1 && rows >= 1 && cols >= 1, SPK_EINVAL, "bad stack");
  int rc = set_device(ctx);
  if (rc) return rc;
  const long long per = (long long)rows * cols;
  int nblk = (int)std::min<long long>(64, (per + 255) / 256);
  const size_t need = (size_t)frames * 5 * nblk;
  if (need > ctx->scratch_cap) {
    if (ctx->stream) SPK_HIP(hipStreamSynchronize(ctx->stream));
    else SPK_HIP(hipDeviceSynchronize());
    dfree(ctx->scratch);
    ctx->scratch_cap = 0;
    if ((rc = dalloc(&ctx->scratch, need))) return rc;
    ctx->scratch_cap = need;
  }
  SPK_HIP(launch_align((c32*)data, frames, per, ctx->scratch, nblk, ctx->stream));
  return SPK_OK;
}

int spk_align_global_phase_noalloc(spk_ctx* ctx, float* data, int frames, int rows, int cols) {
  SPK_REQUIRE(data && frames >= 1 && rows >= 1 && cols >= 1, SPK_EINVAL, "bad stack");
  int rc = set_device(ctx);
  if (rc) return rc;
  // the same partition as spk_align_global_phase (same kernels, same sums)
  const long long per = (long long)rows * cols;
  const int nblk = (int)std::min<long long>(64, (per + 255) / 256);
  const size_t need = (size_t)frames * 5 * nblk;
  SPK_REQUIRE(need <= kScratchDoubles, SPK_EINVAL,
              "stack of " + std::to_string(frames) + " frames of " + std::to_string(rows) + "x" +
                  std::to_string(cols) + " needs " + std::to_string(need) +
                  " partial sums; the static workspace holds " + std::to_string(kScratchDoubles) +
                  " (frames x 5 x min(64, ceil(rows*cols/256)))");
  double* ws = nullptr;
  if ((rc = static_scratch(ctx->device, &ws))) return rc;
  SPK_HIP(launch_align((c32*)data, frames, per, ws, nblk, ctx->stream));
  return SPK_OK;
}

namespace {
// device scratch of a context, grown on demand (queued work is drained first)
int ctx_scratch(spk_ctx* ctx, size_t bytes, void** out) {
  const size_t need = (bytes + 7) / 8;
  if (need > ctx->scratch_cap) {
    if (ctx->stream) SPK_HIP(hipStreamSynchronize(ctx->stream));
    else SPK_HIP(hipDeviceSynchronize());
    dfree(ctx->scratch);
    ctx->scratch_cap = 0;
    int rc = dalloc(&ctx->scratch, need);
    if (rc) return rc;
    ctx->scratch_cap = need;
  }
  *out = ctx->scratch;
  return SPK_OK;
}

// taps of the reference's blur kernel along one axis of length n:
// K[d] = exp(-(((d - h) mod n) - h)^2 / (2 sigma^2)), h = n // 2 (the centred
// Gaussian of shape.gaussian after np.fft.fftshift), normalised over the axis
void blur_taps(int n, double sigma, std::vector<int>& off, std::vector<float>& w) {
  const int h = n / 2;
  std::vector<double> k(n);
  double tot = 0.0, mx = 0.0;
  for (int d = 0; d < n; ++d) {
    const double x = (double)(((d - h) % n + n) % n - h);
    k[d] = std::exp(-x * x / (2.0 * sigma * sigma));
    tot += k[d];
    mx = std::max(mx, k[d]);
  }
  off.clear();
  w.clear();
  for (int d = 0; d < n; ++d)
    if (k[d] > 1e-16 * mx) {
      off.push_back(d);
      w.push_back((float)(k[d] / tot));
    }
}
}  // namespace

int spk_blur2d(spk_ctx* ctx, const float* in, float* out, int rows, int cols, float sigma) {
  SPK_REQUIRE(in && out && rows >= 1 && cols >= 1, SPK_EINVAL, "bad image");
  SPK_REQUIRE(sigma > 0.0f && std::isfinite(sigma), SPK_EINVAL, "blur sigma must be > 0");
  int rc = set_device(ctx);
  if (rc) return rc;
  std::vector<int> orow, ocol;
  std::vector<float> wrow, wcol;
  blur_taps(rows, sigma, orow, wrow);
  blur_taps(cols, sigma, ocol, wcol);
  const size_t img = (size_t)rows * cols * sizeof(float);
  const size_t tabs = (orow.size() + ocol.size()) * 8;
  void* scr;
  if ((rc = ctx_scratch(ctx, img + tabs + 64, &scr))) return rc;
  float* tmp = (float*)scr;
  int* d_orow = (int*)((char*)scr + img);
  int* d_ocol = d_orow + orow.size();
  float* d_wrow = (float*)(d_ocol + ocol.size());
  float* d_wcol = d_wrow + wrow.size();
  std::vector<unsigned char> host(tabs);
  unsigned char* h = host.data();
  std::memcpy(h, orow.data(), orow.size() * 4);
  std::memcpy(h + orow.size() * 4, ocol.data(), ocol.size() * 4);
  std::memcpy(h + (orow.size() + ocol.size()) * 4, wrow.data(), wrow.size() * 4);
  std::memcpy(h + (orow.size() + ocol.size() + wrow.size()) * 4, wcol.data(), wcol.size() * 4);
  // synchronous copy: the host vector goes out of scope at return
  SPK_HIP(hipMemcpyAsync(d_orow, h, tabs, hipMemcpyHostToDevice, ctx->stream));
  SPK_HIP(hipStreamSynchronize(ctx->stream));
  SPK_HIP(launch_conv_axis(in, tmp, rows, cols, 1, d_ocol, d_wcol, (int)ocol.size(), ctx->stream));
  SPK_HIP(launch_conv_axis(tmp, out, rows, cols, 0, d_orow, d_wrow, (int)orow.size(), ctx->stream));
  return SPK_OK;
}

int spk_prtf(spk_plan* p, const float* stack, int frames, int rows, int cols, float* out) {
  SPK_REQUIRE(p && stack && out && frames >= 1, SPK_EINVAL, "bad arguments");
  SPK_REQUIRE(rows >= 1 && cols >= 1 && rows <= p->n && cols <= p->n, SPK_EINVAL,
              "estimates larger than the plan's frame");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  const size_t nn = (size_t)p->nn();
  void* scr;
  if ((rc = ctx_scratch(p->ctx, 2 * frames * nn * sizeof(c32), &scr))) return rc;
  c32* pad = (c32*)scr;
  c32* F = pad + frames * nn;
  SPK_HIP(launch_embed((const c32*)stack, pad, rows, cols, p->n, frames, p->ctx->stream));
  if ((rc = spk_fft2(p, (const float*)pad, (float*)F, frames, 0))) return rc;
  SPK_HIP(launch_prtf(F, frames, p->n, out, p->ctx->stream));
  return SPK_OK;
}

namespace {
// partially coherent speckle of the RFTF functions: F <- sqrt(|ifft2(fft2(A)
// * ipsf)|) with A = |F| (rftf, phasing.py:1567-1573) or |F|^2 (gpu_rftf,
// phasing.py:1676-1690), per frame; ipsf n x n c64 row-major; work: one frame
int rftf_blur(spk_plan* p, c32* F, int frames, const c32* ipsf, int squared, c32* work) {
  const int n = p->n;
  const size_t nn = (size_t)p->nn();
  hipStream_t st = p->ctx->stream;
  for (int f = 0; f < frames; ++f) {
    c32* x = F + (size_t)f * nn;
    SPK_HIP(launch_coh(11, n, p->tile, x, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0,
                       squared ? 1.0 : 0.0, 0, st));
    int rc = spk_fft2(p, (const float*)x, (float*)work, 1, 0);
    if (rc) return rc;
    SPK_HIP(launch_coh(0, n, p->tile, work, ipsf, work, nullptr, nullptr, nullptr, 0, 0, 0, 0, st));
    if ((rc = spk_fft2(p, (const float*)work, (float*)x, 1, 1))) return rc;
    SPK_HIP(launch_coh(11, n, p->tile, x, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, 2.0, 0,
                       st));
  }
  return SPK_OK;
}
}  // namespace

int spk_rftf(spk_plan* p, const float* est, int rows, int cols, const float* goal,
             const float* ipsf, int scale, float* out) {
  SPK_REQUIRE(p && est && goal && out, SPK_EINVAL, "bad arguments");
  SPK_REQUIRE(rows >= 1 && cols >= 1 && rows <= p->n && cols <= p->n, SPK_EINVAL,
              "estimate larger than the goal modulus");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  const size_t nn = (size_t)p->nn();
  const int nblk = 64;
  void* scr;
  if ((rc = ctx_scratch(p->ctx, 2 * nn * sizeof(c32) + 2 * nblk * sizeof(double), &scr)))
    return rc;
  c32* pad = (c32*)scr;
  c32* F = pad + nn;
  double* part = (double*)(F + nn);
  SPK_HIP(launch_embed((const c32*)est, pad, rows, cols, p->n, 1, p->ctx->stream));
  if ((rc = spk_fft2(p, (const float*)pad, (float*)F, 1, 0))) return rc;
  if (ipsf && (rc = rftf_blur(p, F, 1, (const c32*)ipsf, 0, pad))) return rc;
  SPK_HIP(launch_rftf(F, goal, p->n, part, nblk, scale, out, p->ctx->stream));
  return SPK_OK;
}

int spk_rftf_stack(spk_plan* p, const float* stack, int frames, int rows, int cols,
                   const float* goal, const float* ipsf, float* out) {
  SPK_REQUIRE(p && stack && goal && out && frames >= 1, SPK_EINVAL, "bad arguments");
  SPK_REQUIRE(rows >= 1 && cols >= 1 && rows <= p->n && cols <= p->n, SPK_EINVAL,
              "estimates larger than the goal modulus");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  const size_t nn = (size_t)p->nn();
  void* scr;
  if ((rc = ctx_scratch(p->ctx, 2 * frames * nn * sizeof(c32), &scr))) return rc;
  c32* pad = (c32*)scr;
  c32* F = pad + frames * nn;
  SPK_HIP(launch_embed((const c32*)stack, pad, rows, cols, p->n, frames, p->ctx->stream));
  if ((rc = spk_fft2(p, (const float*)pad, (float*)F, frames, 0))) return rc;
  if (ipsf && (rc = rftf_blur(p, F, frames, (const c32*)ipsf, 1, pad))) return rc;
  SPK_HIP(launch_rftf_stack(F, goal, frames, p->n, out, p->ctx->stream));
  return SPK_OK;
}

namespace {
// 2-D transform of one tiled frame: out = F(in), or the unnormalised inverse.
// In the single layout in may equal out; cb is a column-band work frame for
// the dual layout (n >= 4096).
int fft2_tiled(spk_plan* p, const c32* in, c32* out, c32* cb, bool inverse) {
  hipStream_t st = p->ctx->stream;
  c32* mid = p->dual() ? cb : out;
  SPK_HIP(launch_rowfft(p->n, p->tile, p->ept, inverse ? 1 : 0, in, mid, nullptr, 0, 0.0f, p->tw,
                        nullptr, 1, st));
  PassArgs a = base_args(p);
  a.phic = mid;
  a.phi = out;
  a.c = Coef{0, 0, 0, 0, 0, 0, 0, kActive};
  a.tab = nullptr;
  a.fmap = nullptr;
  a.errp = nullptr;
  SPK_HIP(launch_colpass(p->n, p->tile, p->ept, inverse ? 3 : 2, a, 1, 1, st));
  return SPK_OK;
}

constexpr int kCohBlocks = 256;

// active = the (rows, cols) c64 estimate embedded top-left, tiled, in dst
// (row-major staging in tmp)
int embed_tiled(spk_plan* p, const float* est, int rows, int cols, c32* tmp, c32* dst) {
  SPK_HIP(launch_embed((const c32*)est, tmp, rows, cols, p->n, 1, p->ctx->stream));
  SPK_HIP(launch_pack_c64(p->n, p->tile, dst, tmp, nullptr, 0, 1, p->ctx->stream));
  return SPK_OK;
}
}  // namespace

int spk_coherence_rl(spk_plan* p, const float* est, int rows, int cols, const float* modulus,
                     int iterations, int semantics, int is_fourier, float* ipsf) {
  SPK_REQUIRE(p && est && modulus && ipsf, SPK_EINVAL, "bad arguments");
  SPK_REQUIRE(rows >= 1 && cols >= 1 && rows <= p->n && cols <= p->n, SPK_EINVAL,
              "estimate larger than the modulus");
  SPK_REQUIRE(iterations >= 0, SPK_EINVAL, "iterations must be >= 0");
  SPK_REQUIRE(semantics == SPK_SEM_NUMPY || semantics == SPK_SEM_OPENCL, SPK_EINVAL,
              "unknown semantics");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  const int n = p->n, tile = p->tile;
  const size_t nn = (size_t)p->nn();
  hipStream_t st = p->ctx->stream;
  void* scr;
  const size_t frames = p->dual() ? 6 : 5;
  if ((rc = ctx_scratch(p->ctx, frames * nn * sizeof(c32) + nn * sizeof(float) +
                                    kCohBlocks * sizeof(double), &scr)))
    return rc;
  c32* u = (c32*)scr;
  c32* W = u + nn;
  c32* w = W + nn;
  c32* P = w + nn;
  c32* Ph = P + nn;
  c32* cb = p->dual() ? Ph + nn : nullptr;
  float* d = (float*)(Ph + (p->dual() ? 2 : 1) * nn);
  double* part = (double*)(d + nn);
  const float inv_nn = (float)(1.0 / (double)nn);
  // p = |fft2(active)|^2 / sum (|active|^2 for a Fourier-space estimate, the
  // GPU branch), P = fft2(p), Ph = P reversed (phasing.py:806-854)
  if ((rc = embed_tiled(p, est, rows, cols, W, P))) return rc;
  if (!is_fourier && (rc = fft2_tiled(p, P, P, cb, false))) return rc;
  SPK_HIP(launch_coh(3, n, tile, P, nullptr, nullptr, nullptr, nullptr, part, kCohBlocks, 0, 0, 0, st));
  SPK_HIP(launch_coh(4, n, tile, P, nullptr, nullptr, nullptr, nullptr, part, kCohBlocks, 0, 0, 0, st));
  if ((rc = fft2_tiled(p, P, P, cb, false))) return rc;
  SPK_HIP(launch_coh(6, n, tile, nullptr, P, Ph, nullptr, nullptr, nullptr, 0, 0, 0, 0, st));
  // d = m^2 / sum(m^2) (phasing.py:819-821)
  SPK_HIP(launch_pack_f32(n, tile, d, modulus, 1, st));
  SPK_HIP(launch_coh(5, n, tile, nullptr, nullptr, nullptr, nullptr, d, part, kCohBlocks, 0, 0, 0, st));
  SPK_HIP(launch_coh(4, n, tile, nullptr, nullptr, nullptr, nullptr, d, part, kCohBlocks, 0, 0, 0, st));
  // u = the start Gaussian of widths n/4 (phasing.py:818)
  const double w0 = (double)(n / 4);
  SPK_HIP(launch_coh(7, n, tile, nullptr, nullptr, u, nullptr, nullptr, nullptr, 0, 0, w0, w0, st));
  for (int it = 0; it < iterations; ++it) {
    // blur = ifft2(fft2(u) P); w = d / blur; w = ifft2(fft2(w) Ph); u *= w
    // (phasing.py:700-736); the inverse transforms are unnormalised, 1/n^2
    // is applied in the elementwise steps
    if ((rc = fft2_tiled(p, u, W, cb, false))) return rc;
    SPK_HIP(launch_coh(0, n, tile, W, P, W, nullptr, nullptr, nullptr, 0, 0, 0, 0, st));
    if ((rc = fft2_tiled(p, W, W, cb, true))) return rc;
    SPK_HIP(launch_coh(1, n, tile, nullptr, W, w, d, nullptr, nullptr, 0, inv_nn, 0, 0, st));
    if ((rc = fft2_tiled(p, w, w, cb, false))) return rc;
    SPK_HIP(launch_coh(0, n, tile, w, Ph, w, nullptr, nullptr, nullptr, 0, 0, 0, 0, st));
    if ((rc = fft2_tiled(p, w, w, cb, true))) return rc;
    SPK_HIP(launch_coh(2, n, tile, u, w, nullptr, nullptr, nullptr, nullptr, 0, inv_nn, 0, 0, st));
  }
  // ipsf = fft2(u), real part with numpy semantics, / |ipsf(0,0)|
  if ((rc = fft2_tiled(p, u, u, cb, false))) return rc;
  SPK_HIP(launch_coh(9, n, tile, u, nullptr, nullptr, nullptr, nullptr, part, 0, 0,
                     semantics == SPK_SEM_NUMPY ? 1.0 : 0.0, 0, st));
  SPK_HIP(launch_unpack_c64(n, tile, (c32*)ipsf, u, 1.0f, 1, st));
  return SPK_OK;
}

int spk_coherence_gauss_prepare(spk_plan* p, const float* est, int rows, int cols,
                                const float* modulus) {
  SPK_REQUIRE(p && est && modulus, SPK_EINVAL, "bad arguments");
  SPK_REQUIRE(rows >= 1 && cols >= 1 && rows <= p->n && cols <= p->n, SPK_EINVAL,
              "estimate larger than the modulus");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  const size_t nn = (size_t)p->nn();
  if (!p->coh_ac) {
    if ((rc = dalloc(&p->coh_ac, nn)) || (rc = dalloc(&p->coh_w, nn)) ||
        (p->dual() && (rc = dalloc(&p->coh_cb, nn))) || (rc = dalloc(&p->coh_m, nn)) ||
        (rc = dalloc(&p->coh_part, (size_t)kCohBlocks)))
      return rc;
  }
  hipStream_t st = p->ctx->stream;
  // ac = ifft2(|fft2(active)|^2) (phasing.py:785-803)
  if ((rc = embed_tiled(p, est, rows, cols, p->coh_w, p->coh_ac))) return rc;
  if ((rc = fft2_tiled(p, p->coh_ac, p->coh_ac, p->coh_cb, false))) return rc;
  SPK_HIP(launch_coh(3, p->n, p->tile, p->coh_ac, nullptr, nullptr, nullptr, nullptr, p->coh_part,
                     kCohBlocks, 0, 0, 0, st));
  if ((rc = fft2_tiled(p, p->coh_ac, p->coh_ac, p->coh_cb, true))) return rc;
  SPK_HIP(launch_coh(10, p->n, p->tile, p->coh_ac, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                     (float)(1.0 / (double)nn), 0, 0, st));
  SPK_HIP(launch_pack_f32(p->n, p->tile, p->coh_m, modulus, 1, st));
  p->coh_ready = true;
  return SPK_OK;
}

int spk_coherence_gauss_eval(spk_plan* p, double cl0, double cl1, double* value) {
  SPK_REQUIRE(p && value, SPK_EINVAL, "bad arguments");
  SPK_REQUIRE(p->coh_ready, SPK_ESTATE, "spk_coherence_gauss_prepare first");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  hipStream_t st = p->ctx->stream;
  // sum (sqrt|fft2(g * ac)| - m)^2 (phasing.py:657-698)
  SPK_HIP(launch_coh(7, p->n, p->tile, nullptr, p->coh_ac, p->coh_w, nullptr, nullptr, nullptr, 0,
                     0, cl0, cl1, st));
  if ((rc = fft2_tiled(p, p->coh_w, p->coh_w, p->coh_cb, false))) return rc;
  SPK_HIP(launch_coh(8, p->n, p->tile, nullptr, p->coh_w, nullptr, p->coh_m, nullptr, p->coh_part,
                     kCohBlocks, 0, 0, 0, st));
  double h[kCohBlocks];
  SPK_HIP(hipMemcpyAsync(h, p->coh_part, sizeof(h), hipMemcpyDeviceToHost, st));
  SPK_HIP(hipStreamSynchronize(st));
  double s = 0.0;
  for (int b = 0; b < kCohBlocks; ++b) s += h[b];
  *value = s;
  return SPK_OK;
}

int spk_com_average(spk_ctx* ctx, const float* data, int frames, int rows, int cols,
                    float* out) {
  SPK_REQUIRE(data && out && frames >= 1 && rows >= 1 && cols >= 1, SPK_EINVAL, "bad stack");
  int rc = set_device(ctx);
  if (rc) return rc;
  const long long per = (long long)rows * cols;
  const int nblk = (int)std::min<long long>(64, (per + 255) / 256);
  void* scr;
  if ((rc = ctx_scratch(ctx, (size_t)frames * nblk * 3 * sizeof(double) +
                                 (size_t)2 * frames * sizeof(int), &scr)))
    return rc;
  double* part = (double*)scr;
  int* dshift = (int*)(part + (size_t)frames * nblk * 3);
  hipStream_t st = ctx->stream;
  SPK_HIP(launch_com_partial((const c32*)data, frames, rows, cols, part, nblk, st));
  std::vector<double> h((size_t)frames * nblk * 3);
  SPK_HIP(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  SPK_HIP(hipStreamSynchronize(st));
  // centre of mass, truncated to int (phasing.py:1731-1734), relative to frame 0
  std::vector<int> shift(2 * (size_t)frames);
  int r0 = 0, c0 = 0;
  for (int f = 0; f < frames; ++f) {
    double s0 = 0.0, sr = 0.0, sc = 0.0;
    for (int b = 0; b < nblk; ++b) {
      const double* q = &h[((size_t)f * nblk + b) * 3];
      s0 += q[0];
      sr += q[1];
      sc += q[2];
    }
    SPK_REQUIRE(s0 > 0.0, SPK_EINVAL, "a frame of the stack is all zero");
    const int r = (int)(sr / s0), c = (int)(sc / s0);
    if (f == 0) { r0 = r; c0 = c; }
    shift[2 * f] = r - r0;
    shift[2 * f + 1] = c - c0;
  }
  // synchronous: the host vector goes out of scope at return
  SPK_HIP(hipMemcpyAsync(dshift, shift.data(), shift.size() * sizeof(int), hipMemcpyHostToDevice,
                         st));
  SPK_HIP(launch_roll_sum((const c32*)data, frames, rows, cols, dshift, (c32*)out, st));
  SPK_HIP(hipStreamSynchronize(st));
  return SPK_OK;
}

int spk_stack_mean(spk_ctx* ctx, const float* data, int frames, int rows, int cols,
                   float* out) {
  SPK_REQUIRE(data && out && frames >= 1 && rows >= 1 && cols >= 1, SPK_EINVAL, "bad stack");
  int rc = set_device(ctx);
  if (rc) return rc;
  SPK_HIP(launch_stack_mean((const c32*)data, (long long)rows * cols, frames, (c32*)out,
                            ctx->stream));
  return SPK_OK;
}

int spk_copy_bytes(spk_ctx* ctx, void* dst, const void* src, size_t bytes) {
  SPK_REQUIRE(dst && src && bytes % 16 == 0, SPK_EINVAL, "copy needs 16-byte multiples");
  int rc = set_device(ctx);
  if (rc) return rc;
  SPK_HIP(launch_copy(dst, src, bytes, ctx->stream));
  return SPK_OK;
}

struct spk_graph {
  spk_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  unsigned char* stage_d = nullptr;  // the captured step tables (device, pinned)
  unsigned char* stage_h = nullptr;
};

int spk_graph_begin(spk_plan* p) {
  SPK_REQUIRE(p, SPK_EINVAL, "null plan");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  SPK_REQUIRE(p->ctx->stream != nullptr, SPK_EINVAL,
              "graph capture needs a context created on a non-null stream");
  SPK_REQUIRE(!p->profiling, SPK_ESTATE, "disable profiling before capturing");
  SPK_REQUIRE(!p->capturing, SPK_ESTATE, "the plan is already capturing");
  // table buffers of this graph, allocated before the capture starts
  unsigned char *d = nullptr, *h = nullptr;
  if ((rc = dalloc(&d, kCaptureStage))) return rc;
  if (hipHostMalloc((void**)&h, kCaptureStage, hipHostMallocDefault) != hipSuccess) {
    dfree(d);
    return fail(SPK_EHIP, "hipHostMalloc of the captured step tables failed");
  }
  p->cap_d = d;
  p->cap_h = h;
  p->cap_used = 0;
  SPK_HIP(hipStreamBeginCapture(p->ctx->stream, hipStreamCaptureModeThreadLocal));
  p->capturing = true;
  return SPK_OK;
}

int spk_graph_end(spk_plan* p, spk_graph** out) {
  SPK_REQUIRE(p && out, SPK_EINVAL, "null argument");
  *out = nullptr;
  int rc = set_device(p->ctx);
  if (rc) return rc;
  hipGraph_t g = nullptr;
  p->capturing = false;
  // the table buffers go with the graph (or are freed if there is none)
  unsigned char* sd = p->cap_d;
  unsigned char* sh = p->cap_h;
  p->cap_d = p->cap_h = nullptr;
  auto drop_stage = [&]() {
    dfree(sd);
    if (sh) (void)hipHostFree(sh);
  };
  hipError_t e = hipStreamEndCapture(p->ctx->stream, &g);
  if (e != hipSuccess) {
    drop_stage();
    return fail(SPK_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  spk_graph* h = new spk_graph;
  h->ctx = p->ctx;
  h->graph = g;
  e = hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    delete h;
    drop_stage();
    return fail(SPK_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  h->stage_d = sd;
  h->stage_h = sh;
  p->ctx->refs++;
  *out = h;
  return SPK_OK;
}

int spk_graph_launch(spk_graph* h) {
  SPK_REQUIRE(h, SPK_EINVAL, "null graph");
  int rc = set_device(h->ctx);
  if (rc) return rc;
  SPK_HIP(hipGraphLaunch(h->exec, h->ctx->stream));
  return SPK_OK;
}

int spk_graph_destroy(spk_graph* h) {
  if (!h) return SPK_OK;
  (void)hipSetDevice(h->ctx->device);
  (void)hipStreamSynchronize(h->ctx->stream);
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  dfree(h->stage_d);
  if (h->stage_h) (void)hipHostFree(h->stage_h);
  ctx_release(h->ctx);
  delete h;
  return SPK_OK;
}

int spk_plan_profile(spk_plan* p, int enable) {
  SPK_REQUIRE(p, SPK_EINVAL, "null plan");
  p->profiling = enable != 0;
  return SPK_OK;
}

int spk_plan_profile_read(spk_plan* p, int pass, int* launches, double* total_ms) {
  SPK_REQUIRE(p && launches && total_ms, SPK_EINVAL, "null argument");
  SPK_REQUIRE(pass >= 0 && pass <= 2, SPK_EINVAL,
              "pass must be 0 (column), 1 (row) or 2 (resident launches)");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  std::vector<hipEvent_t>& ev = p->prof_ev[pass];
  double tot = 0.0;
  int cnt = 0;
  if (!ev.empty()) SPK_HIP(hipEventSynchronize(ev.back()));
  for (size_t i = 0; i + 1 < ev.size(); i += 2) {
    float ms = 0.0f;
    SPK_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
    tot += ms;
    ++cnt;
  }
  prof_clear(p, pass);
  *launches = cnt;
  *total_ms = tot;
  return SPK_OK;
}

int spk_time_passes(spk_plan* p, int which, int reps, const spk_step* step, float* ms_out) {
  SPK_REQUIRE(p && step && ms_out && reps >= 1, SPK_EINVAL, "bad arguments");
  SPK_REQUIRE(which >= 0 && which <= 2, SPK_EINVAL, "which: 0 column, 1 row, 2 both");
  int rc = set_device(p->ctx);
  if (rc) return rc;
  SPK_REQUIRE(p->mod_set && p->sup_set && p->psi_set, SPK_ESTATE, "plan not loaded");
  SPK_REQUIRE(step->alg != SPK_DM, SPK_EUNSUP, "timing helper covers single-projection steps");
  Coef c;
  if ((rc = make_coef(*step, &c))) return rc;
  hipStream_t st = p->ctx->stream;
  PassArgs a = base_args(p);
  a.c = c;
  hipEvent_t e0, e1;
  SPK_HIP(hipEventCreate(&e0));
  SPK_HIP(hipEventCreate(&e1));
  SPK_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r) {
    if (which != 1) SPK_HIP(launch_colpass(p->n, p->tile, p->ept, 0, a, p->batch, 1, st));
    if (which != 0) SPK_HIP(launch_rowpass(p->n, p->tile, p->ept, false, false, 0, false, a, p->batch, st));
  }
  SPK_HIP(hipEventRecord(e1, st));
  if (which != 0) phib_stale(p, 0, -1);   // the row passes changed psi
  SPK_HIP(hipEventSynchronize(e1));
  SPK_HIP(hipEventElapsedTime(ms_out, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return SPK_OK;
}

}  // extern "C"

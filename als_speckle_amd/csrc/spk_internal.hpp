// Internal types shared by the kernels (spk_kernels.hip) and the C ABI
// (spk_capi.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "spk_fft_core.hpp"

namespace spk {

// Per-frame real-space update of one iteration, the reference's five update
// kernels folded into one coefficient form:
//   psi_new = s*(ao*psi_out + ai*psi_in + at*t) + (1-s)*(bo*psi_out + bi*psi_in + bt*t)
// with t the DM auxiliary projection.  Values per algorithm are set on the
// host (spk_capi.hip, make_coef) from phasing.py:1231-1328 and the .cl kernels.
enum : int {
  kActive = 1,    // frame takes part in this step
  kUseT = 2,      // step is DM: needs t (second column pass on phib)
  kEmitB = 4,     // next step is DM: row pass also writes phib = rowFFT(dm1(psi))
  // shrinkwrap fused into the passes (a shrinkwrap after step k, radius <= 24):
  kSwRow = 8,     // step k's row pass: row blur of |new psi| -> sw_tmp
  kSwFin = 16,    // step k+1: column pass blurs sw_tmp's columns -> sw_blur and
                  // band maxima; row pass applies s *= [blur > tau max] first
  kNoOut = 32,    // the update does not use psi_out (ER with numpy semantics,
                  // s * psi_in): the projected Phi of the step is not needed
};

struct Coef {
  float ao, ai, at;
  float bo, bi, bt;
  float g_next;   // gamma_m = 1/beta of the next (DM) step, used with kEmitB
  int flags;
};

struct BlurW {
  float w[48];  // w[d] for |d| <= radius, normalised over the circle
  int radius;
};

struct SwParam {  // one frame's shrinkwrap: blur weights and threshold fraction
  BlurW w;
  float tau;
};

// Register window of the shrinkwrap blur for radii up to 24 px (sigma <= 4),
// the largest radius whose blur is fused into the passes; the stand-alone
// kernels also have a 48 px window.  Both forms sum the same terms in the
// same order at radius <= 24.
constexpr int kSwFuseRmax = 24;

struct PassArgs {
  c32* psi;               // tiled iterate, frame stride n*n
  c32* phi;               // Phi in the tiled layout: column pass writes, row pass reads
  c32* phic;              // Phi column-band-major: row pass writes, column pass reads
  c32* phib;              // DM modified estimate, tiled layout
  c32* phibc;             // DM modified estimate, column-band-major
  const float* mod;       // column-band-major modulus, frame stride mstride (0 = shared)
  const float* sup;       // tiled support, frame stride sstride (0 = shared)
  long long mstride;
  long long sstride;
  const c32* tw;          // forward twiddles W_n^k, k < n
  double* errp;           // column pass: partial error sums [frame][band]
  const Coef* tab;        // per-frame coefficients (nullptr: use c)
  const int* fmap;        // frames to process: blockIdx.y -> frame (nullptr: identity)
  Coef c;
  float inv_nn;           // 1/(n*n)
  // partial coherence (ipsf): per-frame work buffer for the transform of
  // |Psi|^2, tiled (dbuf) and, in the dual layout, column-band-major (dbufc)
  c32* dbuf;
  c32* dbufc;
  // spectral rescale: per-frame intensity map and its blurred map, float,
  // column-band-major like the modulus
  float* imap;
  float* dmap;
  // the same work buffers for DM's second projection (of phib, grid z = 1)
  c32* dbuf2;
  c32* dbufc2;
  float* imap2;
  float* dmap2;
  // fused shrinkwrap (kSwRow / kSwFin): work maps, per-frame band maxima and
  // the blur/threshold parameters of the shrinkwrap (per frame, or one set)
  float* sw_tmp;
  float* sw_blur;
  float* sw_maxp;
  const SwParam* swr_tab;   // for kSwRow (this step's shrinkwrap)
  const SwParam* swf_tab;   // for kSwFin (the previous step's)
  int sw_pf;                // 1: the tables hold one entry per frame, 0: one entry
  int nt_store;             // 1: the row pass stores psi and Phi non-temporally
  unsigned polls;           // grid barriers: polls of a wait before it gives up
};

// One step of a persistent iteration block (k_persist).  Offsets are bytes
// into the block's staged table upload (-1: none): tab = per-frame
// coefficients (else c for every frame), fmap = the nact frames to process.
struct PStep {
  Coef c;
  long long tab_off;
  long long fmap_off;
  int nact;
  int pad_;
};



// One pass of a mixed-size launch (k_colpass_group / k_rowpass_group): the
// frames of up to kGroupMax plans of different sizes in one launch.  Member
// m's workgroups are [start[m], start[m+1]): its active frames (args[m].fmap,
// or all) times its bands, frame-major.  Each size runs in the layout tile
// that gives a 256-thread workgroup with eight values per thread
// (group_tile), so that one workgroup shape serves every member.
constexpr int kGroupMax = 16;
constexpr int kGroupThreads = 256;
struct GroupPass {
  const PassArgs* args;        // per member, device memory
  int count;
  int start[kGroupMax + 1];
  int n[kGroupMax];
};
// the layout tile of size n in a mixed-size launch (0: not a member size)
constexpr int group_tile(int n) {
  return n == 256 ? 8 : n == 512 ? 4 : n == 1024 ? 2 : 0;
}
constexpr int kGroupEpt = 8;

// One step of a resident single-frame iteration block (k_resident).
struct RStep {
  Coef c;          // kActive, kNoOut, kSwRow, kSwFin; coefficients of the update
  int err;         // write this step's error partials
  int kidx;        // the step's index in the call (row of the error partials)
  int swr;         // kSwRow: this step's shrinkwrap parameters (index into swr_tab)
  int swf;         // kSwFin: the previous step's shrinkwrap parameters
};

// Kernel timing: when set, the next column-pass or row-pass launch on this
// thread is made with hipExtLaunchKernelGGL and records these start/stop
// events with the dispatch itself (the kernel's own execution interval, as a
// profiler reports it); the launch clears them.
extern thread_local hipEvent_t g_launch_events[2];

// Kernel launchers (spk_kernels.hip).  All return hipError_t of the launch.
// `tile` is the internal layout tile edge (2, 4 or 8).
// `ept` is the number of values per thread of one transform (8 or 16).
bool size_supported(int n, int tile, int ept);
hipError_t launch_colpass(int n, int tile, int ept, int mode, const PassArgs& a,
                          int frames, int zdim, hipStream_t st);
// sw (no DM only): bit 0 builds the kSwRow work, bit 1 the kSwFin work; the
// column pass of a kSwFin step is mode 9 (mode 0 plus the column blur).
// noout (no DM, no sw): the row pass of a step whose update ignores psi_out
// (kNoOut) skips the row IFFT of Phi; its column pass is mode 10 (forward
// transform and error partials only)
hipError_t launch_rowpass(int n, int tile, int ept, bool use_t, bool emit_b, int sw,
                          bool noout, const PassArgs& a, int frames, hipStream_t st);
// mixed-size passes over the members of g (single projections, no DM):
// column pass mode 0 or 9, row pass with fused-shrinkwrap bits sw; blocks =
// g.start[g.count]
hipError_t launch_colpass_group(int mode, const GroupPass& g, hipStream_t st);
hipError_t launch_rowpass_group(int sw, const GroupPass& g, hipStream_t st);
// persistent iteration block: one cooperative launch of `grid` workgroups runs
// nsteps (column pass, grid barrier, row pass) steps; bar = barrier counter
// (zeroed by the launcher), flag set to 1 if a barrier wait gave up
hipError_t launch_persist(int n, int tile, int ept, const PassArgs& a, const unsigned char* stage,
                          const PStep* steps, int nsteps, double* errp0, long long errp_step,
                          unsigned* bar, int* flag, int grid, hipStream_t st);
// partial coherence: row pass of the |Psi|^2 convolution,
// out = rowIFFT(ipsf * rowFFT(in)) (unscaled); ipsf tiled, shared by frames;
// out is tiled (in place allowed) or, in the dual layout, column-band-major
hipError_t launch_rowconv(int n, int tile, int ept, const c32* in, c32* out, const c32* ipsf,
                          const c32* tw, const int* fmap, int frames, hipStream_t st);
// spectral (longitudinal) blur of the intensity maps (phasing_radial_rescale.cl):
// dmap = sum_k w_k bilinear(imap at coordinates / r_k); spec = K rescale
// factors then K weights; maps column-band-major with the plan's tile
hipError_t launch_rescale(int n, int tile, const float* imap, float* dmap, const float* spec,
                          int K, const int* fmap, int frames, hipStream_t st);
// coherence estimation pieces (op: 0 out = a*b; 1 out = d/(b*scale);
// 2 a *= b*scale; 3 a = |a|^2 with partial sums; 4 a (or fr) /= sum of
// partials; 5 fr = fr^2 with partial sums; 6 out = reversal of b (tiled);
// 7 out = b * Gaussian(s0, s1) (b = nullptr: the Gaussian); 8 partial sums
// of (sqrt|b| - d)^2; 9 a = a (real part if s0 != 0) / |a(0,0)|; 10 a *= sc;
// 11 a = |a|, |a|^2 or sqrt|a| for s0 = 0, 1, 2)
hipError_t launch_coh(int op, int n, int tile, c32* a, const c32* b, c32* out, const float* d,
                      float* fr, double* partial, int nblk, float scale, double s0, double s1,
                      hipStream_t st);
// center of mass average: per-frame FP64 sums of |x|, row |x|, col |x|
// (partial [frame][block][3]) and the sum of rolled frames
hipError_t launch_com_partial(const c32* data, int frames, int rows, int cols, double* partial,
                              int nblk, hipStream_t st);
hipError_t launch_roll_sum(const c32* data, int frames, int rows, int cols, const int* shift,
                           c32* out, hipStream_t st);
// resident single-frame iteration block: one cooperative launch of n/tile
// workgroups runs nsteps steps with the frame's modulus, support and iterate
// in registers (k_resident); xm: exchange mode of Phi (0 plain + fences,
// 1 device-scope + fences, 2 device-scope, no fences); errp0[kidx *
// errp_step + band] gets the error partials of the steps with err set;
// flags: n/tile * 32 words of barrier flags (zeroed by the launcher); flag:
// this launch's give-up flag, then the sticky one
bool resident_supported(int n, int tile, int ept);
hipError_t launch_resident(int n, int tile, int ept, int xm, const PassArgs& a,
                           const RStep* steps, int nsteps, double* errp0, long long errp_step,
                           unsigned* flags, int* flag, hipStream_t st);
// workgroups of the persistent kernel that fit on the current device at once
hipError_t persist_grid(int n, int tile, int ept, int* out);
// mode 0: out = rowFFT(in); 1: out = rowIFFT(in) (unscaled);
// 2: out = rowFFT((s - g*(1-s)) * in)
hipError_t launch_rowfft(int n, int tile, int ept, int mode, const c32* in, c32* out,
                         const float* sup, long long sstride, float g,
                         const c32* tw, const int* fmap, int frames, hipStream_t st);
hipError_t launch_pack_c64(int n, int tile, c32* dst, const c32* src,
                           const float* sup, long long sstride, int frames,
                           hipStream_t st);
hipError_t launch_unpack_c64(int n, int tile, c32* dst, const c32* src,
                             float scale, int frames, hipStream_t st);
hipError_t launch_pack_f32(int n, int tile, float* dst, const float* src,
                           int frames, hipStream_t st);
hipError_t launch_unpack_f32(int n, int tile, float* dst, const float* src,
                             int frames, hipStream_t st);
// modulus: row-major -> column-band-major (the column pass's read order)
hipError_t launch_pack_mod(int n, int tile, float* dst, const float* src,
                           int frames, hipStream_t st);
hipError_t launch_sumsq(int n, const float* mod, long long mstride,
                        double* partial, int nblk, double* out, int frames,
                        hipStream_t st);
hipError_t launch_err_finalize(const double* errp, int nsteps, int frames,
                               int bands, const double* summ2, double* trace,
                               const int* active, hipStream_t st);
// frames: fmap[0..frames) with parameters tab[0..frames), or (fmap null)
// the range frame0.. with p0 (tab null) / tab
// rmax: the largest blur radius among the frames (<= 47)
hipError_t launch_shrinkwrap(int n, int tile, const c32* psi, float* sup,
                             long long sstride, float* tmp, float* blur,
                             float* maxp, const SwParam* tab, const SwParam& p0,
                             const int* fmap, int frame0, int frames, int rmax,
                             hipStream_t st);
hipError_t launch_copy_roi(int n, int tile, const c32* psi, c32* dst, int r0,
                           int c0, int rows, int cols, hipStream_t st);
hipError_t launch_align(c32* data, int frames, long long per_frame,
                        double* partial, int nblk, hipStream_t st);
hipError_t launch_stack_mean(const c32* data, long long per_frame, int frames,
                             c32* out, hipStream_t st);
hipError_t launch_copy(void* dst, const void* src, size_t bytes,
                       hipStream_t st);
// post-processing (module functions of speckle.phasing)
hipError_t launch_conv_axis(const float* in, float* out, int rows, int cols, int axis,
                            const int* off, const float* w, int taps, hipStream_t st);
hipError_t launch_abs_c64(const c32* in, float* out, long long n, hipStream_t st);
hipError_t launch_embed(const c32* in, c32* out, int r, int c, int n, int frames,
                        hipStream_t st);
hipError_t launch_prtf(const c32* F, int frames, int n, float* out, hipStream_t st);
hipError_t launch_rftf_stack(const c32* F, const float* goal, int frames, int n, float* out,
                             hipStream_t st);
hipError_t launch_rftf(const c32* F, const float* goal, int n, double* partial, int nblk,
                       int scale, float* out, hipStream_t st);

}  // namespace spk

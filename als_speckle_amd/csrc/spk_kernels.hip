// gfx950 kernels of the phase-retrieval loop.
//
// Internal layout: every n x n frame is stored as TT x TT tiles (TT = 4 or
// 8), tiles in row-major order, elements row-major inside a tile:
//   pos(r, c) = ((r/TT)*(n/TT) + c/TT)*TT*TT + (r%TT)*TT + c%TT
// A band of TT rows is then one contiguous block, and a band of TT columns is
// n/TT contiguous TT*TT blocks, so both the row pass and the column pass move
// whole tiles (512 B for c64 at TT=8) per wavefront access.
//
// One iteration = two streaming passes over each frame (fused model):
//   column pass  (k_colpass): column FFT -> modulus replacement m*Psi/|Psi|
//                + FP64 error partials -> column IFFT, in place on phi
//   row pass     (k_rowpass): row IFFT of phi -> psi_out, real-space update
//                with psi_in and the support, store psi_in, row FFT of the
//                new psi_in -> phi for the next iteration
// This replaces, per iteration, the reference's pyfft forward + abs_f2_f2 +
// fourier + pyfft inverse + update kernel (phasing.py:1187-1353).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "spk_internal.hpp"

namespace spk {

// ---------------------------------------------------------------------------
// Register-resident Stockham FFT with LDS exchanges.  The per-thread stage
// logic and every address computation live in spk_fft_core.hpp, shared with
// the CPU emulation in the test suite; kernels add the barriers.
// ---------------------------------------------------------------------------
template <int N, int E, int TT, bool INV, int S>
__device__ __forceinline__ void fft_stage(c32 (&x)[E], const int t, const int b,
                                          c32* __restrict__ lds,
                                          const c32* __restrict__ tw) {
  using P = RadixPlan<N, E>;
  stage_butterflies<N, E, TT, INV, S>(x, t, b, lds, tw);
  if constexpr (S != P::kStages - 1) {
    __syncthreads();
    stage_gather<N, E, TT>(x, t, b, lds);
    __syncthreads();
    fft_stage<N, E, TT, INV, S + 1>(x, t, b, lds, tw);
  }
}

template <int N, int E, int TT, bool INV>
__device__ __forceinline__ void fft_regs(c32 (&x)[E], int t, int b, c32* lds,
                                         const c32* __restrict__ tw) {
  fft_stage<N, E, TT, INV, 0>(x, t, b, lds, tw);
}

// deterministic workgroup sum (fixed tree), result valid in thread 0.
// `sh` must hold blockDim.x doubles and be free for use.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = 1; s < nt; s <<= 1) {
    if ((tid & (2 * s - 1)) == 0 && tid + s < nt) sh[tid] += sh[tid + s];
    __syncthreads();
  }
  return sh[0];
}

__device__ __forceinline__ float block_max(float v, float* sh) {
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = 1; s < nt; s <<= 1) {
    if ((tid & (2 * s - 1)) == 0 && tid + s < nt) sh[tid] = fmaxf(sh[tid], sh[tid + s]);
    __syncthreads();
  }
  return sh[0];
}

// ---------------------------------------------------------------------------
// Column pass.  MODE 0: FFT -> projection (+ error) -> IFFT
//               MODE 2: forward FFT only; MODE 3: inverse FFT only
// grid (n/TT bands, frames, z) ; z = 1 processes phib (DM) without error
// ---------------------------------------------------------------------------
template <int N, int TT, int MODE>
__global__ void __launch_bounds__((N / ept<N>()) * TT) k_colpass(PassArgs a) {
  constexpr int E = ept<N>();
  constexpr int T = N / E;
  constexpr int NB = N / TT;
  __shared__ c32 lds[lds_elems<N, TT>()];
  const int band = blockIdx.x, f = blockIdx.y, z = blockIdx.z;
  const Coef c = a.tab ? a.tab[f] : a.c;
  if (!(c.flags & kActive)) return;
  if (z && !(c.flags & kUseT)) return;
  const size_t fo = (size_t)f * N * N;
  c32* __restrict__ data = (z ? a.phib : a.phi) + fo;
  const int tid = threadIdx.x, b = tid % TT, t = tid / TT;

  c32 x[E];
#pragma unroll
  for (int i = 0; i < E; ++i) x[i] = data[col_pos<N, TT>(band, b, t, i)];
  double err = 0.0;
  if constexpr (MODE == 3) {
    fft_regs<N, E, TT, true>(x, t, b, lds, a.tw);
  } else {
    fft_regs<N, E, TT, false>(x, t, b, lds, a.tw);
  }
  if constexpr (MODE == 0) {
    const float* __restrict__ mod = a.mod + (size_t)f * a.mstride;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const float m = mod[col_pos<N, TT>(band, b, t, i)];
      const float a2 = x[i].x * x[i].x + x[i].y * x[i].y;
      // modulus replacement m * Psi / |Psi| (phasing_fourier.cl:12-15,
      // phasing.py:1217); 0 where |Psi| = 0 (the reference gives NaN there)
      const float r = a2 > 0.0f ? m / sqrtf(a2) : 0.0f;
      const double ad = sqrt((double)x[i].x * x[i].x + (double)x[i].y * x[i].y);
      const double d = ad - (double)m;
      err += d * d;
      x[i] = scale(x[i], r);
    }
    fft_regs<N, E, TT, true>(x, t, b, lds, a.tw);
  }
#pragma unroll
  for (int i = 0; i < E; ++i) data[col_pos<N, TT>(band, b, t, i)] = x[i];
  if constexpr (MODE == 0) {
    if (z == 0 && a.errp) {
      // LDS is free: the last FFT stage is register-only and every thread
      // has passed the barrier that follows its last LDS read
      const double s = block_sum(err, reinterpret_cast<double*>(lds));
      if (tid == 0) a.errp[(size_t)f * NB + band] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Row pass: row IFFT of phi (and phib for DM) -> update -> psi; row FFT of the
// new psi -> phi (and of dm1(psi) -> phib when the next step is DM).
// grid (n/TT bands, frames)
// ---------------------------------------------------------------------------
template <int N, int TT, bool USE_T, bool EMIT_B>
__global__ void __launch_bounds__((N / ept<N>()) * TT) k_rowpass(PassArgs a) {
  constexpr int E = ept<N>();
  constexpr int T = N / E;
  constexpr int NB = N / TT;
  __shared__ c32 lds[lds_elems<N, TT>()];
  const int band = blockIdx.x, f = blockIdx.y;
  const Coef c = a.tab ? a.tab[f] : a.c;
  if (!(c.flags & kActive)) return;
  const size_t fo = (size_t)f * N * N;
  const int tid = threadIdx.x, b = tid % TT, t = tid / TT;
#define SPK_RPOS(i) row_pos<N, TT>(band, b, t, (i))

  c32 x[E];
#pragma unroll
  for (int i = 0; i < E; ++i) x[i] = a.phi[fo + SPK_RPOS(i)];
  fft_regs<N, E, TT, true>(x, t, b, lds, a.tw);

  c32 tv[USE_T ? E : 1];
  if constexpr (USE_T) {
#pragma unroll
    for (int i = 0; i < E; ++i) tv[i] = a.phib[fo + SPK_RPOS(i)];
    fft_regs<N, E, TT, true>(tv, t, b, lds, a.tw);
  }

  const float* __restrict__ sup = a.sup + (size_t)f * a.sstride;
  c32 y[EMIT_B ? E : 1];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int p = SPK_RPOS(i);
    const float s = sup[p];
    const float u = 1.0f - s;
    const c32 pin = a.psi[fo + p];
    const c32 pout = scale(x[i], a.inv_nn);
    c32 in_s = add(scale(pout, c.ao), scale(pin, c.ai));
    c32 out_s = add(scale(pout, c.bo), scale(pin, c.bi));
    if constexpr (USE_T) {
      const c32 tt = scale(tv[i], a.inv_nn);
      in_s = add(in_s, scale(tt, c.at));
      out_s = add(out_s, scale(tt, c.bt));
    }
    const c32 nv = add(scale(in_s, s), scale(out_s, u));
    a.psi[fo + p] = nv;
    x[i] = nv;
    if constexpr (EMIT_B) y[i] = scale(nv, s - c.g_next * u);
  }
  if constexpr (EMIT_B) {
    fft_regs<N, E, TT, false>(y, t, b, lds, a.tw);
#pragma unroll
    for (int i = 0; i < E; ++i) a.phib[fo + SPK_RPOS(i)] = y[i];
  }
  fft_regs<N, E, TT, false>(x, t, b, lds, a.tw);
#pragma unroll
  for (int i = 0; i < E; ++i) a.phi[fo + SPK_RPOS(i)] = x[i];
#undef SPK_RPOS
}

// ---------------------------------------------------------------------------
// Row transform of whole frames (tiled in -> tiled out, may alias).
// MODE 0 forward, 1 inverse (unscaled), 2 forward of (s - g(1-s)) * in
// ---------------------------------------------------------------------------
template <int N, int TT, int MODE>
__global__ void __launch_bounds__((N / ept<N>()) * TT)
    k_rowfft(const c32* in, c32* out, const float* sup, long long sstride,
             float g, const c32* __restrict__ tw) {
  constexpr int E = ept<N>();
  constexpr int T = N / E;
  constexpr int NB = N / TT;
  __shared__ c32 lds[lds_elems<N, TT>()];
  const int band = blockIdx.x, f = blockIdx.y;
  const size_t fo = (size_t)f * N * N;
  const int tid = threadIdx.x, b = tid % TT, t = tid / TT;
  c32 x[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int p = row_pos<N, TT>(band, b, t, i);
    x[i] = in[fo + p];
    if constexpr (MODE == 2) {
      const float s = sup[(size_t)f * sstride + p];
      x[i] = scale(x[i], s - g * (1.0f - s));
    }
  }
  fft_regs<N, E, TT, MODE == 1>(x, t, b, lds, tw);
#pragma unroll
  for (int i = 0; i < E; ++i) out[fo + row_pos<N, TT>(band, b, t, i)] = x[i];
}

// ---------------------------------------------------------------------------
// Layout conversion (row-major <-> tiled), grid (n*n/256, frames)
// ---------------------------------------------------------------------------

template <int TT>
__global__ void __launch_bounds__(256) k_pack_c64(c32* dst, const c32* src, int n,
                                                  const float* sup, long long sstride) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const size_t fo = (size_t)blockIdx.y * n * n;
  int r, c;
  tile_rc<TT>(n, p, r, c);
  c32 v = src[fo + (size_t)r * n + c];
  // seed: psi_in * support (common_multiply_f_f2.cl:10-13, phasing.py:1028-1031)
  if (sup) v = scale(v, sup[(size_t)blockIdx.y * sstride + p]);
  dst[fo + p] = v;
}

template <int TT>
__global__ void __launch_bounds__(256) k_unpack_c64(c32* dst, const c32* src, int n,
                                                    float scl) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const size_t fo = (size_t)blockIdx.y * n * n;
  int r, c;
  tile_rc<TT>(n, p, r, c);
  dst[fo + (size_t)r * n + c] = scale(src[fo + p], scl);
}

template <int TT>
__global__ void __launch_bounds__(256) k_pack_f32(float* dst, const float* src, int n) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const size_t fo = (size_t)blockIdx.y * n * n;
  int r, c;
  tile_rc<TT>(n, p, r, c);
  dst[fo + p] = src[fo + (size_t)r * n + c];
}

template <int TT>
__global__ void __launch_bounds__(256) k_unpack_f32(float* dst, const float* src, int n) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const size_t fo = (size_t)blockIdx.y * n * n;
  int r, c;
  tile_rc<TT>(n, p, r, c);
  dst[fo + (size_t)r * n + c] = src[fo + p];
}

// ---------------------------------------------------------------------------
// FP64 sum of m^2 per frame (layout-free): partials then a fixed-order final
// ---------------------------------------------------------------------------
static __global__ void __launch_bounds__(256) k_sumsq_partial(const float* mod, long long mstride,
                                                       long long count, double* partial,
                                                       int nblk) {
  __shared__ double sh[256];
  const float* m = mod + (size_t)blockIdx.y * mstride;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < count;
       i += (long long)nblk * 256) {
    const double v = m[i];
    acc += v * v;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[(size_t)blockIdx.y * nblk + blockIdx.x] = s;
}

// Neumaier-compensated fixed-order sum of n doubles
__device__ __forceinline__ double kahan_sum(const double* v, int n) {
  double s = 0.0, c = 0.0;
  for (int i = 0; i < n; ++i) {
    const double x = v[i];
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  return s + c;
}

static __global__ void k_sumsq_final(const double* partial, int nblk, double* out, int frames) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < frames) out[f] = kahan_sum(partial + (size_t)f * nblk, nblk);
}

// trace[k*frames + f] = sqrt(sum_bands errp[k][f][band] / summ2[f])
static __global__ void k_err_finalize(const double* errp, int nsteps, int frames, int bands,
                               const double* summ2, double* trace, const int* active) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsteps * frames) return;
  const int f = i % frames;
  if (active && !active[i]) {
    trace[i] = nan("");
    return;
  }
  const double s = kahan_sum(errp + (size_t)i * bands, bands);
  trace[i] = sqrt(s / summ2[f]);
}

// ---------------------------------------------------------------------------
// Shrinkwrap: periodic separable Gaussian blur of |psi| (rows then columns),
// per-band max, then support *= [blur > tau * max] with the threshold in FP64.
// ---------------------------------------------------------------------------
template <int N, int TT>
__global__ void __launch_bounds__(256) k_blur_rows(const c32* psi, float* tmp, BlurW w) {
  constexpr int LS = N + TT;  // padded LDS row stride
  __shared__ float sh[TT * LS];
  const int band = blockIdx.x, f = blockIdx.y;
  const size_t off = (size_t)f * N * N + (size_t)band * N * TT;
  // the band is contiguous: element k -> tile k/TT^2, row (k/TT)%TT, col
  for (int k = threadIdx.x; k < N * TT; k += 256) {
    const int b = (k / TT) % TT;
    const int col = (k / (TT * TT)) * TT + k % TT;
    const c32 v = psi[off + k];
    sh[b * LS + col] = sqrtf(v.x * v.x + v.y * v.y);
  }
  __syncthreads();
  const int R = w.radius;
  for (int k = threadIdx.x; k < N * TT; k += 256) {
    const int b = (k / TT) % TT;
    const int col = (k / (TT * TT)) * TT + k % TT;
    const float* row = sh + b * LS;
    float acc = w.w[0] * row[col];
    for (int d = 1; d <= R; ++d)
      acc += w.w[d] * (row[(col + d) & (N - 1)] + row[(col - d) & (N - 1)]);
    tmp[off + k] = acc;
  }
}

template <int N, int TT>
__global__ void __launch_bounds__(256) k_blur_cols(const float* tmp, float* blur,
                                                   float* maxp, BlurW w) {
  constexpr int NB = N / TT;
  constexpr int LS = N + 8;
  __shared__ float sh[TT * LS];
  __shared__ float red[256];
  const int band = blockIdx.x, f = blockIdx.y;
  const size_t fo = (size_t)f * N * N;
  for (int k = threadIdx.x; k < N * TT; k += 256) {
    const int tr = k / (TT * TT), within = k % (TT * TT);
    const int row = tr * TT + within / TT, ci = within % TT;
    sh[ci * LS + row] = tmp[fo + (size_t)(tr * NB + band) * (TT * TT) + within];
  }
  __syncthreads();
  const int R = w.radius;
  float mx = -INFINITY;
  for (int k = threadIdx.x; k < N * TT; k += 256) {
    const int tr = k / (TT * TT), within = k % (TT * TT);
    const int row = tr * TT + within / TT, ci = within % TT;
    const float* col = sh + ci * LS;
    float acc = w.w[0] * col[row];
    for (int d = 1; d <= R; ++d)
      acc += w.w[d] * (col[(row + d) & (N - 1)] + col[(row - d) & (N - 1)]);
    blur[fo + (size_t)(tr * NB + band) * (TT * TT) + within] = acc;
    mx = fmaxf(mx, acc);
  }
  const float m = block_max(mx, red);
  if (threadIdx.x == 0) maxp[(size_t)f * NB + band] = m;
}

// grid (n*n/1024, frames): 4 elements per thread
static __global__ void __launch_bounds__(256) k_sw_threshold(float* sup, long long sstride,
                                                      const float* blur, const float* maxp,
                                                      int nb, float tau, long long count) {
  __shared__ float red[256];
  const int f = blockIdx.y;
  float mx = -INFINITY;
  for (int i = threadIdx.x; i < nb; i += 256) mx = fmaxf(mx, maxp[(size_t)f * nb + i]);
  const double thr = (double)tau * (double)block_max(mx, red);
  float* s = sup + (size_t)f * sstride;
  const float* bl = blur + (size_t)f * count;
  for (long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
       i < (long long)(blockIdx.x + 1) * 1024 && i < count; i += 256) {
    if (!((double)bl[i] > thr)) s[i] = 0.0f;
  }
}

// ---------------------------------------------------------------------------
// ROI copy (phasing_copy_to_buffer_f2.cl:6-19): dst[j*cols + i] = psi[r0+j, c0+i]
// ---------------------------------------------------------------------------
template <int TT>
__global__ void __launch_bounds__(256) k_copy_roi(const c32* psi, c32* dst, int n, int r0,
                                                  int c0, int rows, int cols) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= rows * cols) return;
  const int j = k / cols, i = k % cols;
  dst[k] = psi[tpos<TT>(n, r0 + j, c0 + i)];
}

// ---------------------------------------------------------------------------
// align_global_phase, closed form (phasing.py:1355-1404): per frame the five
// FP64 sums (Re^2, Im^2, Re*Im, Re, Im), then rotate by +-exp(i p*).
// ---------------------------------------------------------------------------
static __global__ void __launch_bounds__(256) k_align_partial(const c32* data, long long per_frame,
                                                       double* partial, int nblk) {
  __shared__ double sh[256];
  const c32* d = data + (size_t)blockIdx.y * per_frame;
  double s[5] = {0, 0, 0, 0, 0};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < per_frame;
       i += (long long)nblk * 256) {
    const double re = d[i].x, im = d[i].y;
    s[0] += re * re;
    s[1] += im * im;
    s[2] += re * im;
    s[3] += re;
    s[4] += im;
  }
  for (int q = 0; q < 5; ++q) {
    const double v = block_sum(s[q], sh);
    if (threadIdx.x == 0) partial[((size_t)blockIdx.y * 5 + q) * nblk + blockIdx.x] = v;
    __syncthreads();
  }
}

static __global__ void __launch_bounds__(256) k_align_apply(c32* data, long long per_frame,
                                                     const double* partial, int nblk) {
  const int f = blockIdx.y;
  const double* pp = partial + (size_t)f * 5 * nblk;
  const double a = kahan_sum(pp, nblk), bb = kahan_sum(pp + nblk, nblk);
  const double c = kahan_sum(pp + 2 * nblk, nblk);
  const double sre = kahan_sum(pp + 3 * nblk, nblk), sim = kahan_sum(pp + 4 * nblk, nblk);
  // e(p) = (a+b)/2 + (b-a)/2 cos 2p + c sin 2p is smallest at
  // 2p = atan2(c, (b-a)/2) + pi
  const double p = 0.5 * (atan2(c, 0.5 * (bb - a)) + M_PI);
  double cr = cos(p), ci = sin(p);
  // sign rule of phasing.py:1396-1399: negate when sum Re(z e^{ip}) < 0
  if (cr * sre - ci * sim < 0.0) {
    cr = -cr;
    ci = -ci;
  }
  const float fr = (float)cr, fi = (float)ci;
  c32* d = data + (size_t)f * per_frame;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < per_frame;
       i += (long long)gridDim.x * 256) {
    d[i] = mul(d[i], c32{fr, fi});
  }
}

// ---------------------------------------------------------------------------
// plain copy (HBM counter calibration)
// ---------------------------------------------------------------------------
static __global__ void __launch_bounds__(256) k_copy16(uint4* dst, const uint4* src, size_t n16) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256)
    dst[i] = src[i];
}

// ===========================================================================
// host-side launchers.  The size-specialised kernels are compiled one frame
// size per translation unit (SPK_ONLY_N = n), the size-independent kernels
// and the dispatchers in the SPK_ONLY_N = 0 unit; without SPK_ONLY_N one unit
// holds everything.
// ===========================================================================
#ifndef SPK_ONLY_N
#define SPK_BUILD_COMMON 1
#define SPK_BUILD_SIZE(n) 1
#elif SPK_ONLY_N == 0
#define SPK_BUILD_COMMON 1
#define SPK_BUILD_SIZE(n) 0
#else
#define SPK_BUILD_COMMON 0
#define SPK_BUILD_SIZE(n) ((n) == SPK_ONLY_N)
#endif

#define SPK_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

template <int N, int TT>
static hipError_t colpass_nt(int mode, const PassArgs& a, int frames, int zdim,
                             hipStream_t st) {
  dim3 grid(N / TT, frames, zdim), block((N / ept<N>()) * TT);
  switch (mode) {
    case 0: hipLaunchKernelGGL((k_colpass<N, TT, 0>), grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_colpass<N, TT, 2>), grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_colpass<N, TT, 3>), grid, block, 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N, int TT>
static hipError_t rowpass_nt(bool use_t, bool emit_b, const PassArgs& a, int frames,
                             hipStream_t st) {
  dim3 grid(N / TT, frames), block((N / ept<N>()) * TT);
  if (use_t && emit_b)
    hipLaunchKernelGGL((k_rowpass<N, TT, true, true>), grid, block, 0, st, a);
  else if (use_t)
    hipLaunchKernelGGL((k_rowpass<N, TT, true, false>), grid, block, 0, st, a);
  else if (emit_b)
    hipLaunchKernelGGL((k_rowpass<N, TT, false, true>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((k_rowpass<N, TT, false, false>), grid, block, 0, st, a);
  return hipGetLastError();
}

template <int N, int TT>
static hipError_t rowfft_nt(int mode, const c32* in, c32* out, const float* sup,
                            long long sstride, float g, const c32* tw, int frames,
                            hipStream_t st) {
  dim3 grid(N / TT, frames), block((N / ept<N>()) * TT);
  switch (mode) {
    case 0: hipLaunchKernelGGL((k_rowfft<N, TT, 0>), grid, block, 0, st, in, out, sup, sstride, g, tw); break;
    case 1: hipLaunchKernelGGL((k_rowfft<N, TT, 1>), grid, block, 0, st, in, out, sup, sstride, g, tw); break;
    case 2: hipLaunchKernelGGL((k_rowfft<N, TT, 2>), grid, block, 0, st, in, out, sup, sstride, g, tw); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N, int TT>
static hipError_t shrinkwrap_nt(const c32* psi, float* sup, long long sstride, float* tmp,
                                float* blur, float* maxp, const BlurW& w, float tau,
                                int frames, hipStream_t st) {
  dim3 grid(N / TT, frames);
  hipLaunchKernelGGL((k_blur_rows<N, TT>), grid, dim3(256), 0, st, psi, tmp, w);
  hipLaunchKernelGGL((k_blur_cols<N, TT>), grid, dim3(256), 0, st, tmp, blur, maxp, w);
  const long long count = (long long)N * N;
  dim3 g2((unsigned)((count + 1023) / 1024), frames);
  hipLaunchKernelGGL(k_sw_threshold, g2, dim3(256), 0, st, sup, sstride, blur, maxp,
                     N / TT, tau, count);
  return hipGetLastError();
}

// tile 8 is not built for n = 4096 (its LDS exchange would not fit)
template <int N>
constexpr int tile8() {
  return N == 4096 ? 4 : 8;
}

#define SPK_TILE_SWITCH(n, CALL4, CALL8)        \
  if (tile == 4) return CALL4;                  \
  if (tile == 8 && (n) != 4096) return CALL8;   \
  return hipErrorInvalidValue;

#define SPK_SIZED_DECL(n)                                                                   \
  hipError_t colpass_##n(int tile, int mode, const PassArgs& a, int frames, int zdim,      \
                         hipStream_t st);                                                   \
  hipError_t rowpass_##n(int tile, bool use_t, bool emit_b, const PassArgs& a, int frames, \
                         hipStream_t st);                                                   \
  hipError_t rowfft_##n(int tile, int mode, const c32* in, c32* out, const float* sup,     \
                        long long sstride, float g, const c32* tw, int frames,             \
                        hipStream_t st);                                                    \
  hipError_t shrinkwrap_##n(int tile, const c32* psi, float* sup, long long sstride,       \
                            float* tmp, float* blur, float* maxp, const BlurW& w,          \
                            float tau, int frames, hipStream_t st);

#define SPK_SIZED_DEF(n)                                                                    \
  hipError_t colpass_##n(int tile, int mode, const PassArgs& a, int frames, int zdim,      \
                         hipStream_t st) {                                                  \
    SPK_TILE_SWITCH(n, (colpass_nt<n, 4>(mode, a, frames, zdim, st)),                      \
                    (colpass_nt<n, tile8<n>()>(mode, a, frames, zdim, st)))                \
  }                                                                                         \
  hipError_t rowpass_##n(int tile, bool use_t, bool emit_b, const PassArgs& a, int frames, \
                         hipStream_t st) {                                                  \
    SPK_TILE_SWITCH(n, (rowpass_nt<n, 4>(use_t, emit_b, a, frames, st)),                   \
                    (rowpass_nt<n, tile8<n>()>(use_t, emit_b, a, frames, st)))             \
  }                                                                                         \
  hipError_t rowfft_##n(int tile, int mode, const c32* in, c32* out, const float* sup,     \
                        long long sstride, float g, const c32* tw, int frames,             \
                        hipStream_t st) {                                                   \
    SPK_TILE_SWITCH(n, (rowfft_nt<n, 4>(mode, in, out, sup, sstride, g, tw, frames, st)),  \
                    (rowfft_nt<n, tile8<n>()>(mode, in, out, sup, sstride, g, tw, frames,  \
                                              st)))                                         \
  }                                                                                         \
  hipError_t shrinkwrap_##n(int tile, const c32* psi, float* sup, long long sstride,       \
                            float* tmp, float* blur, float* maxp, const BlurW& w,          \
                            float tau, int frames, hipStream_t st) {                        \
    SPK_TILE_SWITCH(n, (shrinkwrap_nt<n, 4>(psi, sup, sstride, tmp, blur, maxp, w, tau,    \
                                            frames, st)),                                   \
                    (shrinkwrap_nt<n, tile8<n>()>(psi, sup, sstride, tmp, blur, maxp, w,   \
                                                  tau, frames, st)))                        \
  }

SPK_SIZES(SPK_SIZED_DECL)

#if SPK_BUILD_SIZE(16)
SPK_SIZED_DEF(16)
#endif
#if SPK_BUILD_SIZE(32)
SPK_SIZED_DEF(32)
#endif
#if SPK_BUILD_SIZE(64)
SPK_SIZED_DEF(64)
#endif
#if SPK_BUILD_SIZE(128)
SPK_SIZED_DEF(128)
#endif
#if SPK_BUILD_SIZE(256)
SPK_SIZED_DEF(256)
#endif
#if SPK_BUILD_SIZE(512)
SPK_SIZED_DEF(512)
#endif
#if SPK_BUILD_SIZE(1024)
SPK_SIZED_DEF(1024)
#endif
#if SPK_BUILD_SIZE(2048)
SPK_SIZED_DEF(2048)
#endif
#if SPK_BUILD_SIZE(4096)
SPK_SIZED_DEF(4096)
#endif

#if SPK_BUILD_COMMON

bool size_supported(int n, int tile) {
  if (tile != 4 && tile != 8) return false;
  switch (n) {
    case 16: case 32: case 64: case 128: case 256: case 512: case 1024: case 2048:
      return true;
    case 4096:
      return tile == 4;
    default:
      return false;
  }
}

#define SPK_SIZE_SWITCH(FN, ...)                         \
  switch (n) {                                           \
    case 16: return FN##_16(__VA_ARGS__);                \
    case 32: return FN##_32(__VA_ARGS__);                \
    case 64: return FN##_64(__VA_ARGS__);                \
    case 128: return FN##_128(__VA_ARGS__);              \
    case 256: return FN##_256(__VA_ARGS__);              \
    case 512: return FN##_512(__VA_ARGS__);              \
    case 1024: return FN##_1024(__VA_ARGS__);            \
    case 2048: return FN##_2048(__VA_ARGS__);            \
    case 4096: return FN##_4096(__VA_ARGS__);            \
    default: return hipErrorInvalidValue;                \
  }

hipError_t launch_colpass(int n, int tile, int mode, const PassArgs& a, int frames,
                          int zdim, hipStream_t st) {
  SPK_SIZE_SWITCH(colpass, tile, mode, a, frames, zdim, st)
}

hipError_t launch_rowpass(int n, int tile, bool use_t, bool emit_b, const PassArgs& a,
                          int frames, hipStream_t st) {
  SPK_SIZE_SWITCH(rowpass, tile, use_t, emit_b, a, frames, st)
}

hipError_t launch_rowfft(int n, int tile, int mode, const c32* in, c32* out,
                         const float* sup, long long sstride, float g, const c32* tw,
                         int frames, hipStream_t st) {
  SPK_SIZE_SWITCH(rowfft, tile, mode, in, out, sup, sstride, g, tw, frames, st)
}

hipError_t launch_shrinkwrap(int n, int tile, const c32* psi, float* sup, long long sstride,
                             float* tmp, float* blur, float* maxp, const BlurW& w,
                             float tau, int frame0, int frames, hipStream_t st) {
  const size_t fo = (size_t)frame0 * n * n;
  psi += fo;
  sup += (size_t)frame0 * sstride;
  tmp += fo;
  blur += fo;
  maxp += (size_t)frame0 * (n / tile);
  SPK_SIZE_SWITCH(shrinkwrap, tile, psi, sup, sstride, tmp, blur, maxp, w, tau, frames, st)
}

hipError_t launch_pack_c64(int n, int tile, c32* dst, const c32* src, const float* sup,
                           long long sstride, int frames, hipStream_t st) {
  dim3 grid((unsigned)((size_t)n * n / 256), frames);
  if (tile == 8)
    hipLaunchKernelGGL(k_pack_c64<8>, grid, dim3(256), 0, st, dst, src, n, sup, sstride);
  else
    hipLaunchKernelGGL(k_pack_c64<4>, grid, dim3(256), 0, st, dst, src, n, sup, sstride);
  return hipGetLastError();
}

hipError_t launch_unpack_c64(int n, int tile, c32* dst, const c32* src, float scale,
                             int frames, hipStream_t st) {
  dim3 grid((unsigned)((size_t)n * n / 256), frames);
  if (tile == 8)
    hipLaunchKernelGGL(k_unpack_c64<8>, grid, dim3(256), 0, st, dst, src, n, scale);
  else
    hipLaunchKernelGGL(k_unpack_c64<4>, grid, dim3(256), 0, st, dst, src, n, scale);
  return hipGetLastError();
}

hipError_t launch_pack_f32(int n, int tile, float* dst, const float* src, int frames,
                           hipStream_t st) {
  dim3 grid((unsigned)((size_t)n * n / 256), frames);
  if (tile == 8)
    hipLaunchKernelGGL(k_pack_f32<8>, grid, dim3(256), 0, st, dst, src, n);
  else
    hipLaunchKernelGGL(k_pack_f32<4>, grid, dim3(256), 0, st, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_unpack_f32(int n, int tile, float* dst, const float* src, int frames,
                             hipStream_t st) {
  dim3 grid((unsigned)((size_t)n * n / 256), frames);
  if (tile == 8)
    hipLaunchKernelGGL(k_unpack_f32<8>, grid, dim3(256), 0, st, dst, src, n);
  else
    hipLaunchKernelGGL(k_unpack_f32<4>, grid, dim3(256), 0, st, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_sumsq(int n, const float* mod, long long mstride, double* partial,
                        int nblk, double* out, int frames, hipStream_t st) {
  const long long count = (long long)n * n;
  hipLaunchKernelGGL(k_sumsq_partial, dim3(nblk, frames), dim3(256), 0, st, mod, mstride,
                     count, partial, nblk);
  hipLaunchKernelGGL(k_sumsq_final, dim3((frames + 63) / 64), dim3(64), 0, st, partial,
                     nblk, out, frames);
  return hipGetLastError();
}

hipError_t launch_err_finalize(const double* errp, int nsteps, int frames, int bands,
                               const double* summ2, double* trace, const int* active,
                               hipStream_t st) {
  const int total = nsteps * frames;
  hipLaunchKernelGGL(k_err_finalize, dim3((total + 255) / 256), dim3(256), 0, st, errp,
                     nsteps, frames, bands, summ2, trace, active);
  return hipGetLastError();
}

hipError_t launch_copy_roi(int n, int tile, const c32* psi, c32* dst, int r0, int c0,
                           int rows, int cols, hipStream_t st) {
  const int total = rows * cols;
  dim3 grid((total + 255) / 256);
  if (tile == 8)
    hipLaunchKernelGGL(k_copy_roi<8>, grid, dim3(256), 0, st, psi, dst, n, r0, c0, rows, cols);
  else
    hipLaunchKernelGGL(k_copy_roi<4>, grid, dim3(256), 0, st, psi, dst, n, r0, c0, rows, cols);
  return hipGetLastError();
}

hipError_t launch_align(c32* data, int frames, long long per_frame, double* partial,
                        int nblk, hipStream_t st) {
  hipLaunchKernelGGL(k_align_partial, dim3(nblk, frames), dim3(256), 0, st, data, per_frame,
                     partial, nblk);
  hipLaunchKernelGGL(k_align_apply, dim3(nblk, frames), dim3(256), 0, st, data, per_frame,
                     partial, nblk);
  return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
  const size_t n16 = bytes / 16;
  hipLaunchKernelGGL(k_copy16, dim3(2048), dim3(256), 0, st, (uint4*)dst, (const uint4*)src,
                     n16);
  return hipGetLastError();
}

#endif  // SPK_BUILD_COMMON

}  // namespace spk

// gfx950 kernels of the phase-retrieval loop.
//
// Internal layout: every n x n frame is stored as TT x TT tiles (TT = 2, 4
// or 8), tiles in row-major order, elements row-major inside a tile:
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
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <stdint.h>

#include "spk_internal.hpp"

namespace spk {

// ---------------------------------------------------------------------------
// LDS plan of an FFT kernel: exchange buffer(s) plus the twiddle table.
// Double buffering (exchanges alternate between two buffers) needs one
// barrier per exchange instead of two.  It is used where both buffers and
// the twiddle table fit in the 160 KB of a gfx950 CU and the second buffer
// does not lower the number of workgroups the LDS admits per CU (counted up
// to 4): measured at 1024^2 with 16 frames, 3 resident workgroups with one
// buffer beat 2 with two by 17 % (column pass) and 4 % (row pass).  The
// per-stage twiddle table (TwStaged, N - R_0 entries, N reserved) is used
// where it fits, else the two-level one.  The plan's global table holds the
// exponent-indexed W_N^m (m < N, for the two-level form) followed by the
// per-stage table of the plan's radix schedule.
// ---------------------------------------------------------------------------
constexpr int kLdsBytes = 160 * 1024;

// DW: the count of workgroups per CU up to which double buffering must keep
// the single-buffer occupancy.  The row-direction kernels use 4; the column
// passes use 8: at 512^2 x 64 frames (tile 4) one buffer admits 5 column
// workgroups per CU instead of 4 and the pass takes 54.0 instead of 58.9 us,
// while the row pass is 4.5 % faster with two buffers (measured on MI355X).
constexpr int kRowDbWgs = 4;
constexpr int kColDbWgs = 8;
template <int N, int TT, int DW>
constexpr bool db_pays() {
  constexpr int buf = lds_elems<N, TT>();
  constexpr int one = kLdsBytes / ((buf + N) * 8);
  constexpr int two = kLdsBytes / ((2 * buf + N) * 8);
  return two >= 1 && two >= (one < DW ? one : DW);
}

template <int N, int TT, int DW = kRowDbWgs>
struct LdsPlan {
  static constexpr int BUF = lds_elems<N, TT>();
  static constexpr bool DB = db_pays<N, TT, DW>();
  static constexpr int NBUF = DB ? 2 : 1;
  static constexpr bool FULL_TW = (NBUF * BUF + N) * 8 <= kLdsBytes;
  static constexpr int TW_ELEMS = FULL_TW ? N : (N >> 6) + 64;
  static constexpr int TOTAL = NBUF * BUF + TW_ELEMS;
  static_assert(TOTAL * 8 <= kLdsBytes, "FFT kernel does not fit in LDS");
  static_assert(FULL_TW || N >= 64, "two-level twiddles need N >= 64");
};

// The kernel's twiddle table, staged through registers: every thread fetches
// its share before the data loads are issued (fetch), and writes it to LDS
// once the data loads are in flight (put), so the table's latency overlaps
// the data's.  No barrier here: stage 0 uses no twiddles and ends with one.
template <int N, int TT, int NT, int DW = kRowDbWgs>
struct TwStage {
  using L = LdsPlan<N, TT, DW>;
  static constexpr int CNT = L::TW_ELEMS;
  static constexpr int PER = (CNT + NT - 1) / NT;
  c32 v[PER];
  __device__ __forceinline__ void fetch(const c32* __restrict__ tw) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      if (CNT % NT == 0 || i < CNT) {
        if constexpr (L::FULL_TW) v[k] = tw[N + i];
        else v[k] = i < (N >> 6) ? tw[i * 64] : tw[i - (N >> 6)];
      }
    }
  }
  __device__ __forceinline__ void put(c32* twl) const {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      if (CNT % NT == 0 || i < CNT) twl[i] = v[k];
    }
  }
};

// Global access at a 32-bit element index from a workgroup-uniform base: the
// byte offset stays 32-bit, so the access uses the scalar-base plus
// vector-offset form instead of per-value 64-bit address arithmetic.
// Frames are at most 4096^2 c64 = 128 MB, so offsets within one fit.
template <class T>
__device__ __forceinline__ T gld(const T* base, unsigned i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void gst(T* base, unsigned i, const T& v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (unsigned)sizeof(T)) = v;
}
// Non-temporal store of a complex value (streaming; not kept in L2).  Used
// for the row pass's contiguous outputs of a single 1024^2 frame, where
// less dirty data in the L2s at the end of the pass shortens the kernel
// boundary (measured: -3 % per iteration); with narrow tiles, smaller frames
// or many frames plain stores are as fast or faster.
__device__ __forceinline__ void gst_nt(c32* base, unsigned i, const c32& v) {
  using V2 = float __attribute__((ext_vector_type(2)));
  V2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<V2*>(reinterpret_cast<char*>(base) + i * 8u));
}
// Non-temporal loads of inputs read once per pass.  Used by the 4096^2
// passes (dual layout), where streamed reads otherwise evict the column
// pass's partly written 128-byte lines of 32-byte tiles from the L2 before
// the neighbouring column bands complete them (measured: 1.52x -> 1.31x
// written bytes, both passes 13 % faster).  At 2048^2 and below they are
// slower: m and s are shared across frames or the frame fits the MALL.
__device__ __forceinline__ c32 gld_nt(const c32* base, unsigned i) {
  using V2 = float __attribute__((ext_vector_type(2)));
  const V2 v = __builtin_nontemporal_load(
      reinterpret_cast<const V2*>(reinterpret_cast<const char*>(base) + i * 8u));
  return c32{v.x, v.y};
}
__device__ __forceinline__ float gld_nt(const float* base, unsigned i) {
  return __builtin_nontemporal_load(
      reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + i * 4u));
}

template <int N, int TT, int DW = kRowDbWgs>
__device__ __forceinline__ auto twiddle_view(const c32* twl) {
  if constexpr (LdsPlan<N, TT, DW>::FULL_TW) return TwStaged{twl};
  else return TwTwoLevel<N>{twl};
}

// ---------------------------------------------------------------------------
// Register-resident Stockham FFT with LDS exchanges.  The per-thread stage
// logic and every address computation live in spk_fft_core.hpp, shared with
// the CPU emulation in the test suite; kernels add the barriers.
// ---------------------------------------------------------------------------
template <int N, int E, int TT, bool INV, int S, int DW, class TW>
__device__ __forceinline__ void fft_stage(c32 (&x)[E], const int t, const int b,
                                          c32* __restrict__ lds, const TW& tw) {
  using P = RadixPlan<N, E>;
  using L = LdsPlan<N, TT, DW>;
  c32* buf = lds + (L::DB ? (S & 1) * L::BUF : 0);
  stage_butterflies<N, E, TT, INV, S>(x, t, b, buf, tw);
  if constexpr (S != P::kStages - 1) {
    __syncthreads();
    stage_gather<N, E, TT>(x, t, b, buf);
    if constexpr (!L::DB) __syncthreads();
    fft_stage<N, E, TT, INV, S + 1, DW>(x, t, b, lds, tw);
  }
}

// One transform.  FIRST: no earlier LDS use in this kernel.  With double
// buffering, an earlier transform's last exchange may still be read by
// slower threads; a barrier is needed when that exchange used buffer 0,
// which this transform's first stage writes.
template <int N, int E, int TT, bool INV, bool FIRST, int DW = kRowDbWgs, class TW>
__device__ __forceinline__ void fft_regs(c32 (&x)[E], int t, int b, c32* lds, const TW& tw) {
  using P = RadixPlan<N, E>;
  using L = LdsPlan<N, TT, DW>;
  if constexpr (!FIRST && L::DB && P::kStages >= 2 && ((P::kStages - 2) & 1) == 0)
    __syncthreads();
  fft_stage<N, E, TT, INV, 0, DW>(x, t, b, lds, tw);
}

// deterministic workgroup sum, result valid in thread 0.  Full wavefronts
// reduce with xor shuffles (fixed order), then thread 0 adds the per-wave
// sums in order.  `sh` must hold max(blockDim.x, 64) doubles; a barrier
// first makes earlier LDS users safe.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  __syncthreads();
  if ((nt & 63) == 0) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((tid & 63) == 0) sh[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
      for (int w = 0; w < (nt >> 6); ++w) s += sh[w];
    return s;
  }
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
  __syncthreads();
  if ((nt & 63) == 0) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    if ((tid & 63) == 0) sh[tid >> 6] = v;
    __syncthreads();
    float m = sh[0];
    for (int w = 1; w < (nt >> 6); ++w) m = fmaxf(m, sh[w]);
    return m;
  }
  sh[tid] = v;
  __syncthreads();
  for (int s = 1; s < nt; s <<= 1) {
    if ((tid & (2 * s - 1)) == 0 && tid + s < nt) sh[tid] = fmaxf(sh[tid], sh[tid + s]);
    __syncthreads();
  }
  return sh[0];
}

// Column band of a workgroup.  Consecutive workgroup ids are dispatched to
// the 8 XCDs in turn, each XCD with its own L2; this gives every XCD a
// contiguous run of bands instead, so that tiles narrower than a 128-byte
// line (tile 2: 32 B) share their lines within one L2.  Needs NB % 8 == 0
// (then the frame index in grid.y does not change the XCD of a band).
template <int NB>
__device__ __forceinline__ int xcd_band(int bid) {
  if constexpr (NB % 8 == 0) return (bid % 8) * (NB / 8) + bid / 8;
  else return bid;
}

// LDS slot of logical index i in a blur line: one pad word per 8 so that
// threads starting 8 apart hit different banks
__device__ __forceinline__ int blur_slot(int i) { return i + (i >> 3); }

// Sliding-window convolution of one line held in LDS: outputs
// out[c0 .. c0+K) = sum_{|d| <= R} w[|d|] line[(c0 + k + d) mod N], with the
// window in registers (K + 2*RMAX loads) and weights w[d] = 0 for d > R.
template <int N, int K, int RMAX>
__device__ __forceinline__ void blur_window(const float* line, int c0, const float (&w)[RMAX + 1],
                                            float (&acc)[K]) {
  float win[K + 2 * RMAX];
#pragma unroll
  for (int j = 0; j < K + 2 * RMAX; ++j) win[j] = line[blur_slot((c0 - RMAX + j) & (N - 1))];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = w[0] * win[k + RMAX];
#pragma unroll
  for (int d = 1; d <= RMAX; ++d) {
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = fmaf(w[d], win[k + RMAX + d] + win[k + RMAX - d], acc[k]);
  }
}

// Fused shrinkwrap (kSwRow / kSwFin steps): the blur of the passes' own
// bands, in the LDS of the pass once its transforms are done.  The TT lines
// of the blur take TT x (N + N/8 + 8) floats from the start of the pass's
// LDS; the outputs are staged after them (N x TT floats, in the order of the
// band's global layout) and stored as float4s, so that a wavefront writes
// whole tiles instead of one float per tile.
template <int N>
constexpr int sw_line() { return N + N / 8 + 8; }
template <int N, int TT>
constexpr int sw_stage_off() { return (TT * sw_line<N>() + 3) & ~3; }
template <int N, int TT>
constexpr int sw_lds_floats() { return sw_stage_off<N, TT>() + N * TT; }
// the pass kernels' LDS holds the lines and the staged outputs (else the
// outputs go straight to global memory)
template <int N, int TT, int DW = kRowDbWgs>
constexpr bool sw_stage_fits() {
  return sw_lds_floats<N, TT>() * 4 <= LdsPlan<N, TT, DW>::TOTAL * 8;
}

// Slot of staged output i: groups of four floats are permuted within each
// block of 32 (i ^ (((i >> 5) & 7) << 2)).  The blur writes its outputs 32
// slots apart across a wavefront's lanes, which without the permutation all
// fall in two LDS banks; a group of four stays contiguous and aligned for the
// float4 reads of the store.
__device__ __forceinline__ int sw_swz(int i) { return i ^ (((i >> 5) & 7) << 2); }

// Store the staged band (N x TT floats) as float4s: group v of four staged
// floats goes to dst + off(v).
template <int N, int TT, class OFF>
__device__ __forceinline__ void sw_store_staged(const float* stage, float* dst, OFF off) {
  for (int v = threadIdx.x; v < N * TT / 4; v += blockDim.x)
    *reinterpret_cast<float4*>(dst + off(v)) =
        *reinterpret_cast<const float4*>(stage + sw_swz(4 * v));
}

__device__ __forceinline__ void sw_weights(const BlurW& bw, float (&w)[kSwFuseRmax + 1]) {
#pragma unroll
  for (int d = 0; d <= kSwFuseRmax; ++d) w[d] = d <= bw.radius ? bw.w[d] : 0.0f;
}

// Blur the TT lines of length N held in LDS (line l at l * sw_line<N>()) by
// segments of 8 outputs; out(l, j) gets the blurred value j of line l and
// the function returns this thread's largest output.
template <int N, int TT, class OUT>
__device__ __forceinline__ float sw_blur_lines(const float* line, const float (&w)[kSwFuseRmax + 1],
                                               OUT out) {
  constexpr int K = 8;
  const int nt = blockDim.x;
  float mx = -INFINITY;
#pragma unroll 1
  for (int seg = threadIdx.x; seg < N * TT / K; seg += nt) {
    const int l = seg / (N / K), j0 = (seg % (N / K)) * K;
    float acc[K];
    blur_window<N, K, kSwFuseRmax>(line + l * sw_line<N>(), j0, w, acc);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      out(l, j0 + k, acc[k]);
      mx = fmaxf(mx, acc[k]);
    }
  }
  return mx;
}

// sw_blur_lines for the resident kernel, whose iterate, support and modulus
// stay in registers: each segment's taps are loaded in three chunks of tap
// distance (d in 1..8, 9..16, 17..24: 15 values left and 15 right of the
// segment per chunk, one LDS read each) and the weights are read from LDS
// (one broadcast read per distance), so few registers are live.  The terms
// are blur_window's, in blur_window's order (w0 x[j], then d ascending), so
// the results are sw_blur_lines'.  wl: the kSwFuseRmax + 1 weights in LDS
// (zero past the radius).
template <int N, int TT, class OUT>
__device__ __forceinline__ float sw_blur_lines_win(const float* line, const float* wl, OUT out) {
  constexpr int K = 8, C = 8;   // outputs per segment, distances per chunk
  static_assert(kSwFuseRmax % C == 0, "whole chunks of tap distance");
  const int nt = blockDim.x;
  float mx = -INFINITY;
#pragma unroll 1
  for (int seg = threadIdx.x; seg < N * TT / K; seg += nt) {
    const int l = seg / (N / K), j0 = (seg % (N / K)) * K;
    const float* ln = line + l * sw_line<N>();
    float acc[K];
    const float w0 = wl[0];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = w0 * ln[blur_slot((j0 + k) & (N - 1))];
#pragma unroll
    for (int a = 1; a <= kSwFuseRmax; a += C) {
      // x[j0 + k - d] = lo[k - d + a + C - 1], x[j0 + k + d] = hi[k + d - a]
      float lo[K + C - 1], hi[K + C - 1];
#pragma unroll
      for (int i = 0; i < K + C - 1; ++i) {
        lo[i] = ln[blur_slot((j0 - a - (C - 1) + i) & (N - 1))];
        hi[i] = ln[blur_slot((j0 + a + i) & (N - 1))];
      }
#pragma unroll
      for (int d = a; d < a + C; ++d) {
        const float wd = wl[d];
#pragma unroll
        for (int k = 0; k < K; ++k)
          acc[k] = fmaf(wd, hi[k + d - a] + lo[k - d + a + C - 1], acc[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      out(l, j0 + k, acc[k]);
      mx = fmaxf(mx, acc[k]);
    }
  }
  return mx;
}

// the fused window's weight of tap threadIdx.x (zero past the radius),
// fetched early; sw_weight_put stores it in the LDS weight table
__device__ __forceinline__ float sw_weight_fetch(const BlurW& bw) {
  return (threadIdx.x <= kSwFuseRmax && (int)threadIdx.x <= bw.radius) ? bw.w[threadIdx.x]
                                                                         : 0.0f;
}
__device__ __forceinline__ void sw_weight_put(float w, float* wl) {
  if (threadIdx.x <= kSwFuseRmax) wl[threadIdx.x] = w;
}

// Modulus replacement m * Psi / |Psi| (phasing_fourier.cl:12-15,
// phasing.py:1217) through one reciprocal square root of |Psi|^2 (v_rsq_f32,
// about 1 ulp); 0 where |Psi|^2 is below the smallest normal float, i.e.
// |Psi| < 1.1e-19 (the reference gives NaN at 0).  err accumulates the
// error term: exact difference and square in FP64 of the FP32 |Psi|.
__device__ __forceinline__ c32 project_modulus(c32 v, float m, double& err) {
  const float p2 = fmaf(v.x, v.x, v.y * v.y);
  const float rs = p2 >= 1.17549435e-38f ? __builtin_amdgcn_rsqf(p2) : 0.0f;
  const float am = p2 * rs;  // |Psi|
  const double d = (double)am - (double)m;
  err = fma(d, d, err);
  return scale(v, m * rs);
}

// The real-space update of one value (the coefficient form of
// phasing.py:1231-1328 and the hio/er/raar/sf/dm2 kernels):
// s (a_o pout + a_i pin + a_t t) + (1 - s) (b_o pout + b_i pin + b_t t)
template <bool USE_T>
__device__ __forceinline__ c32 update_value(const Coef& c, c32 pout, c32 pin, float s, c32 tt) {
  const float u = 1.0f - s;
  c32 in_s = axpby(pout, c.ao, pin, c.ai);
  c32 out_s = axpby(pout, c.bo, pin, c.bi);
  if constexpr (USE_T) {
    in_s = axpy(in_s, tt, c.at);
    out_s = axpy(out_s, tt, c.bt);
  }
  return axpby(in_s, s, out_s, u);
}

// ---------------------------------------------------------------------------
// Column pass.  MODE 0: FFT -> projection (+ error) -> IFFT
//               MODE 9: MODE 0, then the column blur of a fused shrinkwrap
//               (sw_tmp -> sw_blur, band maxima) for frames with kSwFin
//               MODE 10: forward FFT and the error partials only (a step
//               whose update ignores psi_out)
//               MODE 2: forward FFT only; MODE 3: inverse FFT only
//               MODE 4, 5: the two column passes of the partially coherent
//               projection (colpass_ipsf_a, colpass_div_b)
//               MODE 6, 7: intensity maps for the spectral blur; MODE 8: the
//               projection with the blurred divisor (colpass_intensity,
//               colpass_div_b)
// One item = one band of TT columns of frame f; z = 1 processes phib (DM)
// without error.  OWN_TW: the item stages the twiddle table into LDS itself
// (a one-item kernel); otherwise the table is already there (persistent
// kernel).  errp: partial error sums [frame][band] (MODE 0, z = 0).
// ---------------------------------------------------------------------------
template <int N, int TT, int E, int MODE, bool OWN_TW, int DW = kColDbWgs>
__device__ __forceinline__ void colpass_item(const PassArgs& a, int f, int band, int z,
                                             double* errp, c32* lds, bool swf = false) {
  constexpr int NB = N / TT;
  constexpr bool PROJ = MODE == 0 || MODE == 9 || MODE == 10;
  using L = LdsPlan<N, TT, DW>;
  const size_t fo = (size_t)f * N * N;
  // dual layout: read Phi column-band-major (contiguous), write it tiled for
  // the row pass; single layout: tiles in place
  constexpr bool DUAL = dual_layout<N>();
  const c32* src = (z ? (DUAL ? a.phibc : a.phib) : (DUAL ? a.phic : a.phi)) + fo;
  c32* dst = (z ? a.phib : a.phi) + fo;
  const int tid = threadIdx.x, b = tid % TT, t = tid / TT;
  constexpr int NT = (N / E) * TT;

  // every global load up front: twiddle table, data, modulus
  TwStage<N, TT, NT, DW> tws;
  if constexpr (OWN_TW) tws.fetch(a.tw);
  c32 x[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    if constexpr (DUAL) x[i] = gld_nt(src, mcol_pos<N, TT, E>(band, b, t, i));
    else x[i] = gld(src, col_pos<N, TT, E>(band, b, t, i));
  }
  // the modulus is loaded with the data where registers allow, otherwise
  // after the forward transform
  constexpr bool PREFETCH = E <= 16 && NT <= 512;
  const float* __restrict__ mod = a.mod + (size_t)f * a.mstride;
  float m[PROJ ? E : 1];
  if constexpr (PROJ && PREFETCH) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if constexpr (DUAL) m[i] = gld_nt(mod, mcol_pos<N, TT, E>(band, b, t, i));
      else m[i] = gld(mod, mcol_pos<N, TT, E>(band, b, t, i));
    }
  }
  float rb[MODE == 9 ? E : 1];
  c32* twl = lds + L::NBUF * L::BUF;
  if constexpr (OWN_TW) tws.put(twl);
  const auto tw = twiddle_view<N, TT, DW>(twl);

  double err = 0.0;
  if constexpr (MODE == 3) {
    fft_regs<N, E, TT, true, true, DW>(x, t, b, lds, tw);
  } else {
    fft_regs<N, E, TT, false, true, DW>(x, t, b, lds, tw);
  }
  if constexpr (PROJ) {
    if constexpr (!PREFETCH) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        if constexpr (DUAL) m[i] = gld_nt(mod, mcol_pos<N, TT, E>(band, b, t, i));
        else m[i] = gld(mod, mcol_pos<N, TT, E>(band, b, t, i));
      }
    }
#pragma unroll
    for (int i = 0; i < E; ++i) x[i] = project_modulus(x[i], m[i], err);
    if constexpr (MODE == 10) {  // error only
      if (z == 0 && errp) {
        const double s = block_sum(err, reinterpret_cast<double*>(lds));
        if (tid == 0) errp[(size_t)f * NB + band] = s;
      }
      return;
    }
    // fused shrinkwrap: this band's columns of the row-blurred |psi|
    // (tiled), loaded under the inverse transform (the modulus is dead)
    if constexpr (MODE == 9) {
      if (swf) {
#pragma unroll
        for (int i = 0; i < E; ++i) rb[i] = gld(a.sw_tmp + fo, col_pos<N, TT, E>(band, b, t, i));
      }
    }
    fft_regs<N, E, TT, true, false, DW>(x, t, b, lds, tw);
  }
#pragma unroll
  for (int i = 0; i < E; ++i) gst(dst, col_pos<N, TT, E>(band, b, t, i), x[i]);
  if constexpr (PROJ) {
    if (z == 0 && errp) {
      const double s = block_sum(err, reinterpret_cast<double*>(lds));
      if (tid == 0) errp[(size_t)f * NB + band] = s;
    }
  }
  if constexpr (MODE == 9) {
    // the column half of the shrinkwrap blur (as k_blur_cols): lines are the
    // TT columns of the band, outputs tiled into sw_blur, the band maximum
    // into sw_maxp[f][band]
    if (swf) {
      constexpr int T = N / E;
      float* line = reinterpret_cast<float*>(lds);
      __syncthreads();  // every LDS reader above is done
#pragma unroll
      for (int i = 0; i < E; ++i) line[b * sw_line<N>() + blur_slot(t + i * T)] = rb[i];
      float w[kSwFuseRmax + 1];
      sw_weights(a.swf_tab[a.sw_pf ? f : 0].w, w);
      __syncthreads();
      float mx;
      if constexpr (sw_stage_fits<N, TT, DW>()) {
        // staged in tile order: tile q (rows qTT..qTT+TT-1 of the band) at q TT^2
        float* stage = line + sw_stage_off<N, TT>();
        mx = sw_blur_lines<N, TT>(line, w, [&](int l, int row, float v) {
          stage[sw_swz((row / TT) * (TT * TT) + (row % TT) * TT + l)] = v;
        });
        __syncthreads();
        sw_store_staged<N, TT>(stage, a.sw_blur + fo, [&](int v) {
          const int q = (4 * v) / (TT * TT), within = (4 * v) % (TT * TT);
          return (q * NB + band) * (TT * TT) + within;
        });
      } else {
        float* blur = a.sw_blur + fo;
        mx = sw_blur_lines<N, TT>(line, w, [&](int l, int row, float v) {
          blur[((row / TT) * NB + band) * (TT * TT) + (row % TT) * TT + l] = v;
        });
      }
      const float bm = block_max(mx, line);
      if (tid == 0) a.sw_maxp[(size_t)f * NB + band] = bm;
    }
  }
}

// Partial coherence, column pass A (phasing.py:1175-1180 with ipsf):
// Psi = colFFT(Phi); Q = colFFT(|Psi|^2) -> dbuf (tiled).  Phi is kept.
// z = 1: the same for DM's estimate phib, into dbuf2.
template <int N, int TT, int E>
__device__ __forceinline__ void colpass_ipsf_a(const PassArgs& a, int f, int band, int z,
                                               c32* lds) {
  using L = LdsPlan<N, TT, kColDbWgs>;
  constexpr int NT = (N / E) * TT;
  constexpr bool DUAL = dual_layout<N>();
  const size_t fo = (size_t)f * N * N;
  const c32* src = (z ? (DUAL ? a.phibc : a.phib) : (DUAL ? a.phic : a.phi)) + fo;
  const int tid = threadIdx.x, b = tid % TT, t = tid / TT;
  TwStage<N, TT, NT, kColDbWgs> tws;
  tws.fetch(a.tw);
  c32 x[E];
#pragma unroll
  for (int i = 0; i < E; ++i)
    x[i] = gld(src, DUAL ? mcol_pos<N, TT, E>(band, b, t, i) : col_pos<N, TT, E>(band, b, t, i));
  c32* twl = lds + L::NBUF * L::BUF;
  tws.put(twl);
  const auto tw = twiddle_view<N, TT, kColDbWgs>(twl);
  fft_regs<N, E, TT, false, true, kColDbWgs>(x, t, b, lds, tw);
#pragma unroll
  for (int i = 0; i < E; ++i) x[i] = c32{fmaf(x[i].x, x[i].x, x[i].y * x[i].y), 0.0f};
  fft_regs<N, E, TT, false, false, kColDbWgs>(x, t, b, lds, tw);
#pragma unroll
  for (int i = 0; i < E; ++i) gst((z ? a.dbuf2 : a.dbuf) + fo, col_pos<N, TT, E>(band, b, t, i), x[i]);
}

// Intensity map for the spectral blur (MODE 6, 7): the squared divisor,
// column-band-major in imap.  MODE 6: |Psi|^2 with Psi = colFFT(Phi) (the
// coherent divisor squared); MODE 7: |colIFFT(D)| / n^2 with D from the ipsf
// row convolution (the partially coherent divisor squared).
template <int N, int TT, int E, int MODE>
__device__ __forceinline__ void colpass_intensity(const PassArgs& a, int f, int band, int z,
                                                  c32* lds) {
  using L = LdsPlan<N, TT, kColDbWgs>;
  constexpr int NT = (N / E) * TT;
  constexpr bool DUAL = dual_layout<N>();
  const size_t fo = (size_t)f * N * N;
  // z = 1: DM's estimate (phib, dbuf2, imap2)
  const c32* src = MODE == 6 ? (z ? (DUAL ? a.phibc : a.phib) : (DUAL ? a.phic : a.phi)) + fo
                             : (z ? (DUAL ? a.dbufc2 : a.dbuf2) : (DUAL ? a.dbufc : a.dbuf)) + fo;
  const int tid = threadIdx.x, b = tid % TT, t = tid / TT;
  TwStage<N, TT, NT, kColDbWgs> tws;
  tws.fetch(a.tw);
  c32 x[E];
#pragma unroll
  for (int i = 0; i < E; ++i)
    x[i] = gld(src, DUAL ? mcol_pos<N, TT, E>(band, b, t, i) : col_pos<N, TT, E>(band, b, t, i));
  c32* twl = lds + L::NBUF * L::BUF;
  tws.put(twl);
  fft_regs<N, E, TT, MODE == 7, true, kColDbWgs>(x, t, b, lds, twiddle_view<N, TT, kColDbWgs>(twl));
  float* dst = (z ? a.imap2 : a.imap) + fo;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const float p2 = fmaf(x[i].x, x[i].x, x[i].y * x[i].y);
    gst(dst, mcol_pos<N, TT, E>(band, b, t, i), MODE == 6 ? p2 : sqrtf(p2) * a.inv_nn);
  }
}

// Column pass B of a partially coherent projection: the divisor, then
// Psi' = m * Psi / div (0 where div = 0; phasing.py:1217), the FP64 error
// partial of (div - m)^2, and Phi = colIFFT(Psi').
//   SRC 0 (ipsf, MODE 5): div = sqrt(|colIFFT(D)| / n^2) with D the row
//     pass's rowIFFT(ipsf * rowFFT(colFFT(|Psi|^2))), i.e.
//     div = sqrt(|ifft2(fft2(|Psi|^2) * ipsf)|) (phasing.py:1175-1183);
//   SRC 1 (spectrum, MODE 8): div = sqrt(dmap), the spectrally blurred
//     squared divisor (phasing.py:1140-1156).
template <int N, int TT, int E, int SRC>
__device__ __forceinline__ void colpass_div_b(const PassArgs& a, int f, int band, int z,
                                              double* errp, c32* lds) {
  using L = LdsPlan<N, TT, kColDbWgs>;
  constexpr int NB = N / TT;
  constexpr int NT = (N / E) * TT;
  constexpr bool DUAL = dual_layout<N>();
  const size_t fo = (size_t)f * N * N;
  // z = 1: the projection of DM's estimate phib with its own divisor (no error)
  const c32* src = (z ? (DUAL ? a.phibc : a.phib) : (DUAL ? a.phic : a.phi)) + fo;
  const int tid = threadIdx.x, b = tid % TT, t = tid / TT;
  TwStage<N, TT, NT, kColDbWgs> tws;
  tws.fetch(a.tw);
  c32 d[SRC == 0 ? E : 1];
  float div[E];
  if constexpr (SRC == 0) {
    const c32* dsrc = (z ? (DUAL ? a.dbufc2 : a.dbuf2) : (DUAL ? a.dbufc : a.dbuf)) + fo;
#pragma unroll
    for (int i = 0; i < E; ++i)
      d[i] = gld(dsrc, DUAL ? mcol_pos<N, TT, E>(band, b, t, i) : col_pos<N, TT, E>(band, b, t, i));
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) div[i] = gld((z ? a.dmap2 : a.dmap) + fo, mcol_pos<N, TT, E>(band, b, t, i));
  }
  c32 x[E];
#pragma unroll
  for (int i = 0; i < E; ++i)
    x[i] = gld(src, DUAL ? mcol_pos<N, TT, E>(band, b, t, i) : col_pos<N, TT, E>(band, b, t, i));
  const float* __restrict__ mod = a.mod + (size_t)f * a.mstride;
  float m[E];
#pragma unroll
  for (int i = 0; i < E; ++i) m[i] = gld(mod, mcol_pos<N, TT, E>(band, b, t, i));
  c32* twl = lds + L::NBUF * L::BUF;
  tws.put(twl);
  const auto tw = twiddle_view<N, TT, kColDbWgs>(twl);
  if constexpr (SRC == 0) {
    fft_regs<N, E, TT, true, true, kColDbWgs>(d, t, b, lds, tw);
#pragma unroll
    for (int i = 0; i < E; ++i)
      div[i] = sqrtf(sqrtf(fmaf(d[i].x, d[i].x, d[i].y * d[i].y)) * a.inv_nn);
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) div[i] = sqrtf(div[i]);
  }
  fft_regs<N, E, TT, false, SRC != 0, kColDbWgs>(x, t, b, lds, tw);
  double err = 0.0;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const double dd = (double)div[i] - (double)m[i];
    err = fma(dd, dd, err);
    x[i] = scale(x[i], div[i] > 0.0f ? m[i] / div[i] : 0.0f);
  }
  fft_regs<N, E, TT, true, false, kColDbWgs>(x, t, b, lds, tw);
#pragma unroll
  for (int i = 0; i < E; ++i) gst((z ? a.phib : a.phi) + fo, col_pos<N, TT, E>(band, b, t, i), x[i]);
  if (errp && z == 0) {
    const double s = block_sum(err, reinterpret_cast<double*>(lds));
    if (tid == 0) errp[(size_t)f * NB + band] = s;
  }
}

// grid (n/TT bands, frames, z)
template <int N, int TT, int E, int MODE>
__global__ void __launch_bounds__((N / E) * TT) k_colpass(PassArgs a) {
  constexpr int NB = N / TT;
  __shared__ c32 lds[LdsPlan<N, TT, kColDbWgs>::TOTAL];
  static_assert(MODE != 9 || TT * sw_line<N>() * sizeof(float) <= sizeof(lds),
                "fused blur lines exceed the pass's LDS");
  const int band = xcd_band<NB>(blockIdx.x), z = blockIdx.z;
  const int f = a.fmap ? a.fmap[blockIdx.y] : blockIdx.y;
  const Coef c = a.tab ? a.tab[f] : a.c;
  if (!(c.flags & kActive)) return;
  if (z && !(c.flags & kUseT)) return;
  if constexpr (MODE == 4) {
    colpass_ipsf_a<N, TT, E>(a, f, band, z, lds);
  } else if constexpr (MODE == 5) {
    colpass_div_b<N, TT, E, 0>(a, f, band, z, a.errp, lds);
  } else if constexpr (MODE == 6 || MODE == 7) {
    colpass_intensity<N, TT, E, MODE>(a, f, band, z, lds);
  } else if constexpr (MODE == 8) {
    colpass_div_b<N, TT, E, 1>(a, f, band, z, a.errp, lds);
  } else {
    colpass_item<N, TT, E, MODE, true>(a, f, band, z, a.errp, lds,
                                       MODE == 9 && (c.flags & kSwFin));
  }
}

// ---------------------------------------------------------------------------
// Row pass: row IFFT of phi (and phib for DM) -> update -> psi; row FFT of the
// new psi -> phi (and of dm1(psi) -> phib when the next step is DM).
// One item = one band of TT rows of frame f with coefficients c.
// SW (fused shrinkwrap, no DM): bit 1, for a frame with kSwFin, first applies
// the previous step's shrinkwrap to the support rows of the band (s = 0
// where blur <= tau * max, the FP64 threshold of k_sw_threshold, with the
// frame maximum reduced from the column pass's band maxima); bit 0, for a
// frame with kSwRow, ends with the row half of this step's shrinkwrap blur of
// |new psi| into sw_tmp (as k_blur_rows).
// ---------------------------------------------------------------------------
template <int N, int TT, int E, bool USE_T, bool EMIT_B, bool OWN_TW, int SW = 0,
          bool NOOUT = false>
__device__ __forceinline__ void rowpass_item(const PassArgs& a, int f, int band, const Coef& c,
                                             c32* lds) {
  static_assert(SW == 0 || (!USE_T && !EMIT_B), "fused shrinkwrap is built without DM");
  static_assert(!NOOUT || (!USE_T && !EMIT_B && SW == 0), "kNoOut steps: no DM, no shrinkwrap");
  using L = LdsPlan<N, TT>;
  constexpr int NB = N / TT;
  const size_t fo = (size_t)f * N * N;
  const int tid = threadIdx.x, b = tid % TT, t = tid / TT;
  constexpr int NT = (N / E) * TT;
#define SPK_RPOS(i) row_pos<N, TT, E>(band, b, t, (i))
  c32* phi = a.phi + fo;
  c32* psi = a.psi + fo;
  const bool swf = (SW & 2) && (c.flags & kSwFin);
  const bool swr = (SW & 1) && (c.flags & kSwRow);

  // every global load up front: twiddles, the band maxima of a fused
  // shrinkwrap, phi (and phib), psi_in, support (and the blur)
  TwStage<N, TT, NT> tws;
  if constexpr (OWN_TW) tws.fetch(a.tw);
  constexpr int MPT = (SW & 2) ? (NB + NT - 1) / NT : 1;
  float mpv[MPT];
  if constexpr (SW & 2) {
    if (swf) {
#pragma unroll
      for (int j = 0; j < MPT; ++j) {
        const int q = tid + j * NT;
        mpv[j] = q < NB ? a.sw_maxp[(size_t)f * NB + q] : -INFINITY;
      }
    }
  }
  c32 x[E];
  if constexpr (!NOOUT) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if constexpr (dual_layout<N>()) x[i] = gld_nt(phi, SPK_RPOS(i));
      else x[i] = gld(phi, SPK_RPOS(i));
    }
  }
  // the DM term t only for frames whose step is DM (a mixed per-frame step
  // runs its other frames through the same kernel; their phib is not read)
  const bool use_t = USE_T && (c.flags & kUseT);
  c32 tv[USE_T ? E : 1];
  if constexpr (USE_T) {
    if (use_t) {
#pragma unroll
      for (int i = 0; i < E; ++i) tv[i] = gld(a.phib + fo, SPK_RPOS(i));
    }
  }
  // psi_in and the support are loaded with phi where registers allow (no
  // DM), otherwise after the transforms
  constexpr bool PREFETCH = E <= 16 && NT <= 512 && !USE_T;
  const float* __restrict__ sup = a.sup + (size_t)f * a.sstride;
  c32 pin[E];
  float sv[E];
  float bl[(SW & 2) ? E : 1];
  if constexpr (PREFETCH) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if constexpr (dual_layout<N>()) {
        pin[i] = gld_nt(psi, SPK_RPOS(i));
        sv[i] = gld_nt(sup, SPK_RPOS(i));
      } else {
        pin[i] = gld(psi, SPK_RPOS(i));
        sv[i] = gld(sup, SPK_RPOS(i));
      }
    }
    if constexpr (SW & 2) {
      if (swf) {
#pragma unroll
        for (int i = 0; i < E; ++i) bl[i] = gld(a.sw_blur + fo, SPK_RPOS(i));
      }
    }
  }
  double thr = 0.0;
  // s = 0 where blur <= tau * max, in the support and in sv
  auto apply_threshold = [&]() {
    float* supw = const_cast<float*>(sup);  // the support rows of this band
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if (!((double)bl[i] > thr)) {
        sv[i] = 0.0f;
        gst(supw, SPK_RPOS(i), 0.0f);
      }
    }
  };
  if constexpr (SW & 2) {
    if (swf) {  // uniform over the workgroup
      float mx = mpv[0];
#pragma unroll
      for (int j = 1; j < MPT; ++j) mx = fmaxf(mx, mpv[j]);
      thr = (double)a.swf_tab[a.sw_pf ? f : 0].tau *
            (double)block_max(mx, reinterpret_cast<float*>(lds));
      __syncthreads();  // the reduction's LDS is reused by the transforms
      if constexpr (PREFETCH) apply_threshold();
    }
  }
  c32* twl = lds + L::NBUF * L::BUF;
  if constexpr (OWN_TW) tws.put(twl);
  const auto tw = twiddle_view<N, TT>(twl);

  if constexpr (!NOOUT) fft_regs<N, E, TT, true, true>(x, t, b, lds, tw);
  if constexpr (USE_T) {
    if (use_t) fft_regs<N, E, TT, true, false>(tv, t, b, lds, tw);  // uniform per workgroup
  }
  if constexpr (!PREFETCH) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if constexpr (dual_layout<N>()) {
        pin[i] = gld_nt(psi, SPK_RPOS(i));
        sv[i] = gld_nt(sup, SPK_RPOS(i));
      } else {
        pin[i] = gld(psi, SPK_RPOS(i));
        sv[i] = gld(sup, SPK_RPOS(i));
      }
    }
    if constexpr (SW & 2) {
      if (swf) {
#pragma unroll
        for (int i = 0; i < E; ++i) bl[i] = gld(a.sw_blur + fo, SPK_RPOS(i));
        apply_threshold();
      }
    }
  }

  c32 y[EMIT_B ? E : 1];
  float av[(SW & 1) ? E : 1];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const float s = sv[i];
    const float u = 1.0f - s;
    c32 tt{0.0f, 0.0f};
    if constexpr (USE_T) tt = use_t ? scale(tv[i], a.inv_nn) : c32{0.0f, 0.0f};
    // kNoOut: the coefficients of psi_out are 0, and psi_out is not formed
    const c32 pout = NOOUT ? c32{0.0f, 0.0f} : scale(x[i], a.inv_nn);
    const c32 nv = update_value<USE_T>(c, pout, pin[i], s, tt);
    if (a.nt_store) gst_nt(psi, SPK_RPOS(i), nv);
    else gst(psi, SPK_RPOS(i), nv);
    x[i] = nv;
    if constexpr (SW & 1) av[i] = sqrtf(nv.x * nv.x + nv.y * nv.y);
    if constexpr (EMIT_B) y[i] = scale(nv, fmaf(-c.g_next, u, s));
  }
  // the new row transforms: column-band-major for the column pass (dual
  // layout) or back in place (single layout)
  constexpr bool DUAL = dual_layout<N>();
  if constexpr (EMIT_B) {
    fft_regs<N, E, TT, false, false>(y, t, b, lds, tw);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if constexpr (DUAL) gst(a.phibc + fo, cbrow_pos<N, TT, E>(band, b, t, i), y[i]);
      else gst(a.phib + fo, SPK_RPOS(i), y[i]);
    }
  }
  fft_regs<N, E, TT, false, false>(x, t, b, lds, tw);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    if constexpr (DUAL) gst(a.phic + fo, cbrow_pos<N, TT, E>(band, b, t, i), x[i]);
    else if (a.nt_store) gst_nt(phi, SPK_RPOS(i), x[i]);
    else gst(phi, SPK_RPOS(i), x[i]);
  }
  if constexpr (SW & 1) {
    // the row half of the shrinkwrap blur: lines are the TT rows of the band
    if (swr) {
      constexpr int T = N / E;
      float* line = reinterpret_cast<float*>(lds);
      __syncthreads();  // the transforms' LDS reads are done
#pragma unroll
      for (int i = 0; i < E; ++i) line[b * sw_line<N>() + blur_slot(t + i * T)] = av[i];
      float w[kSwFuseRmax + 1];
      sw_weights(a.swr_tab[a.sw_pf ? f : 0].w, w);
      __syncthreads();
      float* tmp = a.sw_tmp + fo + (size_t)band * N * TT;
      if constexpr (sw_stage_fits<N, TT>()) {
        // staged in the band's (contiguous) global order
        float* stage = line + sw_stage_off<N, TT>();
        sw_blur_lines<N, TT>(line, w, [&](int l, int col, float v) {
          stage[sw_swz((col / TT) * (TT * TT) + l * TT + (col % TT))] = v;
        });
        __syncthreads();
        sw_store_staged<N, TT>(stage, tmp, [](int v) { return 4 * v; });
      } else {
        sw_blur_lines<N, TT>(line, w, [&](int l, int col, float v) {
          tmp[(col / TT) * (TT * TT) + l * TT + (col % TT)] = v;
        });
      }
    }
  }
#undef SPK_RPOS
}

// grid (n/TT bands, frames)
template <int N, int TT, int E, bool USE_T, bool EMIT_B, int SW, bool NOOUT = false>
__global__ void __launch_bounds__((N / E) * TT) k_rowpass(PassArgs a) {
  __shared__ c32 lds[LdsPlan<N, TT>::TOTAL];
  static_assert(SW == 0 || TT * sw_line<N>() * sizeof(float) <= sizeof(lds),
                "fused blur lines exceed the pass's LDS");
  const int f = a.fmap ? a.fmap[blockIdx.y] : blockIdx.y;
  const Coef c = a.tab ? a.tab[f] : a.c;
  if (!(c.flags & kActive)) return;
  rowpass_item<N, TT, E, USE_T, EMIT_B, true, SW, NOOUT>(a, f, blockIdx.x, c, lds);
}

// ---------------------------------------------------------------------------
// Row pass of the partial-coherence convolution: out = rowIFFT(ipsf *
// rowFFT(in)), unscaled (the 1/n^2 of the 2-D inverse is applied in the
// second column pass).  ipsf is tiled and shared by the frames.
// grid (n/TT bands, frames)
// ---------------------------------------------------------------------------
template <int N, int TT, int E>
__global__ void __launch_bounds__((N / E) * TT)
    k_rowconv(const c32* in, c32* out, const c32* __restrict__ ipsf, const c32* __restrict__ tw,
              const int* fmap) {
  using L = LdsPlan<N, TT>;
  constexpr int NT = (N / E) * TT;
  __shared__ c32 lds[L::TOTAL];
  const int band = blockIdx.x, f = fmap ? fmap[blockIdx.y] : blockIdx.y;
  const size_t fo = (size_t)f * N * N;
  const int tid = threadIdx.x, b = tid % TT, t = tid / TT;
  TwStage<N, TT, NT> tws;
  tws.fetch(tw);
  c32 x[E], w[E];
#pragma unroll
  for (int i = 0; i < E; ++i) x[i] = gld(in + fo, row_pos<N, TT, E>(band, b, t, i));
#pragma unroll
  for (int i = 0; i < E; ++i) w[i] = gld(ipsf, row_pos<N, TT, E>(band, b, t, i));
  c32* twl = lds + L::NBUF * L::BUF;
  tws.put(twl);
  const auto tv = twiddle_view<N, TT>(twl);
  fft_regs<N, E, TT, false, true>(x, t, b, lds, tv);
#pragma unroll
  for (int i = 0; i < E; ++i) x[i] = mul(x[i], w[i]);
  fft_regs<N, E, TT, true, false>(x, t, b, lds, tv);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    if constexpr (dual_layout<N>()) gst(out + fo, cbrow_pos<N, TT, E>(band, b, t, i), x[i]);
    else gst(out + fo, row_pos<N, TT, E>(band, b, t, i), x[i]);
  }
}

// ---------------------------------------------------------------------------
// Persistent iteration block: one cooperative launch runs `nsteps` steps
// (column pass, grid barrier, row pass, grid barrier), each workgroup taking
// items (frame, band) in a strided loop.  Steps here use neither DM nor a
// DM successor (no phib); shrinkwrap events end a block on the host side.
// The twiddle table is staged into LDS once per launch.
//
// Grid barrier: a monotonically increasing counter in device memory, zeroed
// before the launch; barrier k waits until it reaches k * gridDim.x.  After a
// workgroup barrier, thread 0 fences at device scope before its arrival
// (release: the workgroup's stores leave the XCD's L2) and after the wait
// (acquire: stale lines of other XCDs' data are dropped), then the workgroup
// barrier releases the other waves.  The wait is bounded: after `limit`
// polls (2^18 by default, a quarter second or more, far above any
// legitimate wait; spk_tune.barrier_polls sets it; the launch
// is cooperative, so every workgroup is resident) flag[0] (this launch's,
// cleared with the counter before each launch) and flag[1] (sticky) are set
// and the workgroup goes on, later barriers of the launch pass at once, and
// every wave ends; the host reports the sticky flag as an error
// (spk_plan_check).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void grid_sync(unsigned* bar, unsigned target, int* flag,
                                          unsigned limit) {
  // every wave waits for its own stores to reach L2 (s_waitcnt vmcnt(0):
  // gfx9 encoding, expcnt and lgkmcnt left at their maxima), then thread 0
  // publishes the workgroup's stores device-wide: one L2 write-back and one
  // invalidate per workgroup, not per wave (a workgroup runs on one CU)
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // after one give-up every later barrier of the launch passes at once
    if (!__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
      unsigned polls = 0;
      while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(1);
        if (++polls > limit) {
          __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(flag + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

template <int N, int TT, int E>
__global__ void __launch_bounds__((N / E) * TT)
    k_persist(PassArgs a, const unsigned char* stage, const PStep* steps, int nsteps,
              double* errp0, long long errp_step, unsigned* bar, int* flag) {
  constexpr int NB = N / TT;
  constexpr int NT = (N / E) * TT;
  using L = LdsPlan<N, TT>;
  __shared__ c32 lds[L::TOTAL];
  {
    TwStage<N, TT, NT> tws;
    tws.fetch(a.tw);
    tws.put(lds + L::NBUF * L::BUF);
  }
  const unsigned G = gridDim.x;
  unsigned epoch = 0;
  for (int k = 0; k < nsteps; ++k) {
    const PStep ps = steps[k];
    const int items = ps.nact * NB;
    const Coef* tab = ps.tab_off >= 0 ? reinterpret_cast<const Coef*>(stage + ps.tab_off) : nullptr;
    const int* fmap = ps.fmap_off >= 0 ? reinterpret_cast<const int*>(stage + ps.fmap_off) : nullptr;
    double* errp = errp0 ? errp0 + (size_t)k * errp_step : nullptr;
    for (int it = blockIdx.x; it < items; it += G) {
      const int fi = it / NB, band = it % NB;
      const int f = fmap ? fmap[fi] : fi;
      const Coef c = tab ? tab[f] : ps.c;
      __syncthreads();  // twiddles staged / previous item done with LDS
      if (c.flags & kActive)
        colpass_item<N, TT, E, 0, false, kRowDbWgs>(a, f, band, 0, errp, lds);
    }
    grid_sync(bar, ++epoch * G, flag, a.polls);
    for (int it = blockIdx.x; it < items; it += G) {
      const int fi = it / NB, band = it % NB;
      const int f = fmap ? fmap[fi] : fi;
      const Coef c = tab ? tab[f] : ps.c;
      __syncthreads();
      if (c.flags & kActive) rowpass_item<N, TT, E, false, false, false>(a, f, band, c, lds);
    }
    if (k + 1 < nsteps) grid_sync(bar, ++epoch * G, flag, a.polls);
  }
}

// ---------------------------------------------------------------------------
// Resident single-frame iteration block (k_resident).  One cooperative
// launch of one workgroup per band runs a block of steps; the workgroup of
// band q owns column band q (column phase) and row band q (row phase) for the
// whole launch and keeps the frame's modulus (its column band) and support
// and iterate (its row band) in registers, loading them once.  Per step only
// Phi crosses between workgroups, through global memory:
//   column phase: Phi column band -> FFT -> projection (+ error partial) ->
//                 IFFT -> Phi column band   (error only for kNoOut steps)
//   grid barrier
//   row phase:    Phi row band -> IFFT -> update with psi, s (registers) ->
//                 FFT -> Phi row band
//   grid barrier
// The arithmetic is that of colpass_item / rowpass_item (same transforms,
// projection and update), so results equal the per-pass launches bitwise.
// At the end the workgroup stores its rows of psi.
//
// Exchange mode XM (how Phi crosses the grid barrier between XCDs, each with
// its own L2):
//   0  plain loads/stores; the barrier fences at device scope (write back
//      the L2's dirty lines before arriving, invalidate after the wait)
//   1  device-scope (sc1) loads and stores of Phi, and the fences of 0
//   2  device-scope loads and stores of Phi; the barrier orders them without
//      fences (each wave waits for its stores before the arrival)
// ---------------------------------------------------------------------------
template <int XM>
__device__ __forceinline__ c32 xld(const c32* base, unsigned i) {
  if constexpr (XM == 0) {
    return gld(base, i);
  } else {
    const uint64_t v = __hip_atomic_load(reinterpret_cast<const uint64_t*>(base) + i,
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return c32{__uint_as_float((unsigned)v), __uint_as_float((unsigned)(v >> 32))};
  }
}
template <int XM>
__device__ __forceinline__ void xst(c32* base, unsigned i, const c32& v) {
  if constexpr (XM == 0) {
    gst(base, i, v);
  } else {
    const uint64_t w = (uint64_t)__float_as_uint(v.x) | ((uint64_t)__float_as_uint(v.y) << 32);
    __hip_atomic_store(reinterpret_cast<uint64_t*>(base) + i, w, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Float exchange of the resident kernel's shrinkwrap maps, device scope in
// the exchange modes without fences (as xld / xst)
template <int XM>
__device__ __forceinline__ float xldf(const float* p, unsigned i) {
  if constexpr (XM != 2) return gld(p, i);
  else return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int XM>
__device__ __forceinline__ void xstf(float* p, unsigned i, float v) {
  if constexpr (XM != 2) gst(p, i, v);
  else __hip_atomic_store(p + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// sw_store_staged for the resident kernel: float4 stores, or in the fenceless
// exchange mode two device-scope 8-byte stores per group of four
template <int N, int TT, int XM, class OFF>
__device__ __forceinline__ void sw_store_staged_x(const float* stage, float* dst, OFF off) {
  for (int v = threadIdx.x; v < N * TT / 4; v += blockDim.x) {
    if constexpr (XM != 2) {
      *reinterpret_cast<float4*>(dst + off(v)) =
          *reinterpret_cast<const float4*>(stage + sw_swz(4 * v));
    } else {
      const uint64_t* s2 = reinterpret_cast<const uint64_t*>(stage + sw_swz(4 * v));
      uint64_t* d2 = reinterpret_cast<uint64_t*>(dst + off(v));
      __hip_atomic_store(d2, s2[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(d2 + 1, s2[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Grid barrier of the resident kernel.  Every workgroup publishes the
// barrier's epoch in its own flag (one 128-byte line per workgroup, so the
// arrivals do not contend for one counter), and the first wavefront of every
// workgroup reads all the flags until each holds the epoch: lane l reads
// flags l, l + 64, ... (K = grid / 64 of them), all K loads in flight at once
// in every poll round, so a round costs one device-scope load latency
// however many flags a lane covers.  Reading a lane's flags one after the
// other (each awaited before the next is issued) cost 1.9 us more per
// iteration of the 1024^2 headline (two barriers; 16.9 vs 15.0 us).  One
// shared arrival counter was slower still (8 us per barrier).
// Device-scope fences as the exchange mode needs them (release before the
// flag, acquire after the wait).  Waits are bounded as in grid_sync: after
// `limit` poll rounds flag[0] (this launch's) and flag[1] (sticky) are set
// and every later wait of the launch passes.
template <int XM, int K>
__device__ __forceinline__ void resident_sync(unsigned* flags, unsigned epoch, int* flag,
                                              unsigned limit) {
  static_assert(K >= 1, "at least one flag per polling lane");
  constexpr int LINE = 32;   // unsigned words per 128-byte line
  constexpr bool fence = XM != 2;
  __builtin_amdgcn_s_waitcnt(0x0F70);   // this wave's stores have completed
  __syncthreads();
  if (threadIdx.x == 0) {
    if (fence) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(flags + blockIdx.x * LINE, epoch, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x < 64) {
    bool gave = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    unsigned polls = 0;
    while (!gave) {
      // no branch between the loads (a flag already seen is read again), so
      // they are issued back to back and awaited once
      unsigned v[K];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const unsigned j = min(threadIdx.x + q * 64u, gridDim.x - 1u);
        v[q] = __hip_atomic_load(flags + j * LINE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      bool done = true;
#pragma unroll
      for (int q = 0; q < K; ++q) done = done && v[q] >= epoch;
      if (done) break;
      __builtin_amdgcn_s_sleep(1);
      if (++polls > limit) {
        __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(flag + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        gave = true;
      }
    }
    if (fence) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// LDS of the resident kernel: the pass LDS, padded so that one workgroup
// occupies a CU (the grid is one workgroup per band, at most one per CU)
template <int N, int TT>
constexpr int resident_lds_elems() {
  constexpr int base = LdsPlan<N, TT>::TOTAL;
  constexpr int half = (kLdsBytes / 2) / 8 + 64;   // more than half a CU's LDS
  return base > half ? base : half;
}

template <int N, int TT, int E, int XM>
__global__ void __launch_bounds__((N / E) * TT)
    k_resident(PassArgs a, const RStep* __restrict__ steps, int nsteps, double* errp0,
               long long errp_step, unsigned* flags, int* flag) {
  constexpr int NT = (N / E) * TT;
  constexpr int NB = N / TT;
  constexpr int T = N / E;
  using L = LdsPlan<N, TT>;
  __shared__ c32 lds[resident_lds_elems<N, TT>()];
  // the shrinkwrap blur lines and staged outputs live after the twiddle
  // table, which stays in LDS for the whole launch
  static_assert(L::TOTAL * 8 + (sw_lds_floats<N, TT>() + 32) * 4 <=
                    resident_lds_elems<N, TT>() * 8,
                "resident LDS holds the pass LDS, the blur lines and weights");
  float* swl = reinterpret_cast<float*>(lds + L::TOTAL);
  float* swwl = swl + sw_lds_floats<N, TT>();   // blur weights
  const int tid = threadIdx.x, b = tid % TT, t = tid / TT;
  const int cband = xcd_band<NB>(blockIdx.x);   // as k_colpass
  const int rband = blockIdx.x;                 // as k_rowpass
  c32* twl = lds + L::NBUF * L::BUF;
  {
    TwStage<N, TT, NT> tws;
    tws.fetch(a.tw);
    tws.put(twl);
  }
  const auto tw = twiddle_view<N, TT>(twl);
  // resident state: modulus of the column band, support and iterate of the row band
  float m[E], sv[E];
  c32 pin[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    m[i] = gld(a.mod, mcol_pos<N, TT, E>(cband, b, t, i));
    sv[i] = gld(a.sup, row_pos<N, TT, E>(rband, b, t, i));
    pin[i] = gld(a.psi, row_pos<N, TT, E>(rband, b, t, i));
  }
  c32* phi = a.phi;
  unsigned epoch = 0;
  bool sup_changed = false;
  // the column phase's FP64 error partial of this thread; the workgroup sum
  // is formed in the row phase while its Phi loads are in flight (off the
  // path between the column stores and the barrier)
  double err = 0.0;
  for (int k = 0; k < nsteps; ++k) {
    const RStep st = steps[k];
    const bool noout = (st.c.flags & kNoOut) != 0;
    const bool swrow = (st.c.flags & kSwRow) != 0;   // shrinkwrap after this step
    const bool swfin = (st.c.flags & kSwFin) != 0;   // ... after the previous one
    const bool col = !noout || st.err || swfin;
    const bool err_sum = col && st.err && errp0;
    c32 x[E];
    if (col) {
      // column phase (colpass_item MODE 0, MODE 10 for kNoOut, MODE 9's blur)
      __syncthreads();   // the LDS of the previous phase is free
#pragma unroll
      for (int i = 0; i < E; ++i) x[i] = xld<XM>(phi, col_pos<N, TT, E>(cband, b, t, i));
      // a fused shrinkwrap's column half: its weights, fetched with Phi
      float wcol = 0.0f;
      int rcol = 0;
      if (swfin) {
        wcol = sw_weight_fetch(a.swr_tab[st.swf].w);
        rcol = a.swr_tab[st.swf].w.radius;
      }
      err = 0.0;
      fft_regs<N, E, TT, false, true>(x, t, b, lds, tw);
#pragma unroll
      for (int i = 0; i < E; ++i) x[i] = project_modulus(x[i], m[i], err);
      // the row-blurred |psi| of the band's columns, loaded under the inverse
      // transform (as colpass_item MODE 9)
      float rb[E];
      if (swfin) {
#pragma unroll
        for (int i = 0; i < E; ++i) rb[i] = xldf<XM>(a.sw_tmp, col_pos<N, TT, E>(cband, b, t, i));
      }
      if (!noout) {
        fft_regs<N, E, TT, true, false>(x, t, b, lds, tw);
#pragma unroll
        for (int i = 0; i < E; ++i) xst<XM>(phi, col_pos<N, TT, E>(cband, b, t, i), x[i]);
      }
      if (swfin) {
        // the column half of the previous step's shrinkwrap blur (as
        // colpass_item MODE 9): the band's columns of the row-blurred |psi|
        // -> sw_blur (tiled) and the band maximum
        __syncthreads();
#pragma unroll
        for (int i = 0; i < E; ++i) swl[b * sw_line<N>() + blur_slot(t + i * T)] = rb[i];
        sw_weight_put(wcol, swwl);
        __syncthreads();
        float* stage = swl + sw_stage_off<N, TT>();
        const float mx = sw_blur_lines_win<N, TT>(swl, swwl, [&](int l, int row, float v) {
          stage[sw_swz((row / TT) * (TT * TT) + (row % TT) * TT + l)] = v;
        });
        __syncthreads();
        sw_store_staged_x<N, TT, XM>(stage, a.sw_blur, [&](int v) {
          const int q = (4 * v) / (TT * TT), within = (4 * v) % (TT * TT);
          return (q * NB + cband) * (TT * TT) + within;
        });
        const float bm = block_max(mx, swl);
        if (tid == 0) xstf<XM>(a.sw_maxp, cband, bm);
      }
      resident_sync<XM, (NB + 63) / 64>(flags, ++epoch, flag, a.polls);
    }
    // row phase (rowpass_item; kNoOut skips Phi).  Every load is issued
    // first: Phi, the band maxima and blur of a fused shrinkwrap's threshold,
    // the weights of this step's shrinkwrap.
    __syncthreads();
    if (!noout) {
#pragma unroll
      for (int i = 0; i < E; ++i) x[i] = xld<XM>(phi, row_pos<N, TT, E>(rband, b, t, i));
    }
    float wrow = 0.0f;
    int rrow = 0;
    if (swrow) {
      wrow = sw_weight_fetch(a.swr_tab[st.swr].w);
      rrow = a.swr_tab[st.swr].w.radius;
    }
    if (swfin) {
      // the previous step's threshold: s = 0 where blur <= tau * max, with
      // the frame maximum from the column bands' maxima (FP64 threshold)
      constexpr int MPT = (NB + NT - 1) / NT;
      float mx = -INFINITY;
      float bl[E];
#pragma unroll
      for (int j = 0; j < MPT; ++j) {
        const int q = tid + j * NT;
        if (q < NB) mx = fmaxf(mx, xldf<XM>(a.sw_maxp, q));
      }
#pragma unroll
      for (int i = 0; i < E; ++i) bl[i] = xldf<XM>(a.sw_blur, row_pos<N, TT, E>(rband, b, t, i));
      const double thr = (double)a.swr_tab[st.swf].tau * (double)block_max(mx, swl);
#pragma unroll
      for (int i = 0; i < E; ++i)
        if (!((double)bl[i] > thr)) sv[i] = 0.0f;
      sup_changed = true;
      __syncthreads();
    }
    if (err_sum) {
      // the column phase's error partial (the same sum in the same order as
      // colpass_item's); the reduction uses the shrinkwrap area of the LDS,
      // which the transforms below do not touch
      const double s = block_sum(err, reinterpret_cast<double*>(swl));
      if (tid == 0) errp0[(size_t)st.kidx * errp_step + cband] = s;
    }
    if (!noout) fft_regs<N, E, TT, true, true>(x, t, b, lds, tw);
    float av[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const c32 pout = noout ? c32{0.0f, 0.0f} : scale(x[i], a.inv_nn);
      const c32 nv = update_value<false>(st.c, pout, pin[i], sv[i], c32{0.0f, 0.0f});
      pin[i] = nv;
      x[i] = nv;
      av[i] = sqrtf(nv.x * nv.x + nv.y * nv.y);
    }
    fft_regs<N, E, TT, false, false>(x, t, b, lds, tw);
#pragma unroll
    for (int i = 0; i < E; ++i) xst<XM>(phi, row_pos<N, TT, E>(rband, b, t, i), x[i]);
    if (swrow) {
      // the row half of this step's shrinkwrap blur (as rowpass_item SW 1):
      // the band's rows of |new psi| -> sw_tmp
      __syncthreads();
#pragma unroll
      for (int i = 0; i < E; ++i) swl[b * sw_line<N>() + blur_slot(t + i * T)] = av[i];
      sw_weight_put(wrow, swwl);
      __syncthreads();
      float* stage = swl + sw_stage_off<N, TT>();
      sw_blur_lines_win<N, TT>(swl, swwl, [&](int l, int c, float v) {
        stage[sw_swz((c / TT) * (TT * TT) + l * TT + (c % TT))] = v;
      });
      __syncthreads();
      sw_store_staged_x<N, TT, XM>(stage, a.sw_tmp + (size_t)rband * N * TT,
                                   [](int v) { return 4 * v; });
    }
    if (k + 1 < nsteps) resident_sync<XM, (NB + 63) / 64>(flags, ++epoch, flag, a.polls);
  }
#pragma unroll
  for (int i = 0; i < E; ++i) gst(a.psi, row_pos<N, TT, E>(rband, b, t, i), pin[i]);
  if (sup_changed) {
    float* sup = const_cast<float*>(a.sup);
#pragma unroll
    for (int i = 0; i < E; ++i) gst(sup, row_pos<N, TT, E>(rband, b, t, i), sv[i]);
  }
}

static __global__ void k_zero_u32(unsigned* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Row transform of whole frames: tiled in -> the column pass's input layout
// (tiled, or column-band-major in the dual layout); in and out may alias
// only in the single layout.
// MODE 0 forward, 1 inverse (unscaled), 2 forward of (s - g(1-s)) * in
// ---------------------------------------------------------------------------
template <int N, int TT, int E, int MODE>
__global__ void __launch_bounds__((N / E) * TT)
    k_rowfft(const c32* in, c32* out, const float* sup, long long sstride,
             float g, const c32* __restrict__ tw, const int* fmap) {
  using L = LdsPlan<N, TT>;
  __shared__ c32 lds[L::TOTAL];
  const int band = blockIdx.x, f = fmap ? fmap[blockIdx.y] : blockIdx.y;
  const size_t fo = (size_t)f * N * N;
  const int tid = threadIdx.x, b = tid % TT, t = tid / TT;
  constexpr int NT = (N / E) * TT;
  TwStage<N, TT, NT> tws;
  tws.fetch(tw);
  const c32* src = in + fo;
  const float* sp = sup + (size_t)f * sstride;
  c32 x[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int p = row_pos<N, TT, E>(band, b, t, i);
    x[i] = gld(src, p);
    if constexpr (MODE == 2) {
      // the rounding of the row pass's own emission (rowpass_item, kEmitB)
      const float s = gld(sp, p);
      x[i] = scale(x[i], fmaf(-g, 1.0f - s, s));
    }
  }
  c32* twl = lds + L::NBUF * L::BUF;
  tws.put(twl);
  fft_regs<N, E, TT, MODE == 1, true>(x, t, b, lds, twiddle_view<N, TT>(twl));
  c32* dst = out + fo;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    if constexpr (dual_layout<N>()) gst(dst, cbrow_pos<N, TT, E>(band, b, t, i), x[i]);
    else gst(dst, row_pos<N, TT, E>(band, b, t, i), x[i]);
  }
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

// row-major -> column-band-major (modulus)
template <int TT>
__global__ void __launch_bounds__(256) k_pack_mod(float* dst, const float* src, int n) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const size_t fo = (size_t)blockIdx.y * n * n;
  const int band = p / (n * TT), w = p % (n * TT);
  const int r = w / TT, c = band * TT + w % TT;
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

__device__ __forceinline__ void neumaier_add(double& s, double& c, double x) {
  const double t = s + x;
  c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
  s = t;
}

// trace[k*frames + f] = sqrt(sum_bands errp[k][f][band] / summ2[f]).  One
// wavefront per (step, frame): lane l sums bands l, l+64, ... (Neumaier, in
// band order), then the 64 lane sums are added in lane order (Neumaier), so
// the result does not depend on the launch shape.
static __global__ void __launch_bounds__(256) k_err_finalize(const double* errp, int nsteps,
                                                             int frames, int bands,
                                                             const double* summ2, double* trace,
                                                             const int* active) {
  const int i = (int)((blockIdx.x * 256u + threadIdx.x) >> 6), lane = threadIdx.x & 63;
  if (i >= nsteps * frames) return;  // whole wavefronts
  const int f = i % frames;
  if (active && !active[i]) {
    if (lane == 0) trace[i] = nan("");
    return;
  }
  const double* v = errp + (size_t)i * bands;
  double s = 0.0, c = 0.0;
  for (int b = lane; b < bands; b += 64) neumaier_add(s, c, v[b]);
  const double mine = s + c;
  double S = 0.0, C = 0.0;
  for (int l = 0; l < 64; ++l) {
    const double x = __shfl(mine, l, 64);
    neumaier_add(S, C, x);
  }
  if (lane == 0) trace[i] = sqrt((S + C) / summ2[f]);
}

// ---------------------------------------------------------------------------
// Shrinkwrap: periodic separable Gaussian blur of |psi| (rows then columns),
// per-band max, then support *= [blur > tau * max] with the threshold in FP64.
// grid.y walks the frames to process: frame = fmap[y] (or y), parameters
// tab[y] (or p0).
// ---------------------------------------------------------------------------
template <int RMAX>
__device__ __forceinline__ void load_blur_weights(const SwParam* tab, const SwParam& p0,
                                                  float* wsh, float (&w)[RMAX + 1]) {
  const BlurW& bw = tab ? tab[blockIdx.y].w : p0.w;
  if (threadIdx.x <= RMAX) wsh[threadIdx.x] = (int)threadIdx.x <= bw.radius ? bw.w[threadIdx.x] : 0.0f;
  __syncthreads();
#pragma unroll
  for (int d = 0; d <= RMAX; ++d) w[d] = wsh[d];
}

template <int N, int TT, int RMAX>
__global__ void __launch_bounds__(256) k_blur_rows(const c32* psi, float* tmp,
                                                   const SwParam* tab, SwParam p0,
                                                   const int* fmap) {
  constexpr int K = 8;
  constexpr int LS = N + N / 8 + 8;  // padded line stride
  __shared__ float sh[TT * LS];
  __shared__ float wsh[RMAX + 1];
  const int band = blockIdx.x, f = fmap ? fmap[blockIdx.y] : blockIdx.y;
  const size_t off = (size_t)f * N * N + (size_t)band * N * TT;
  // the band is contiguous: element k -> row (k/TT)%TT, column (k/TT^2)*TT + k%TT
  for (int k = threadIdx.x; k < N * TT; k += 256) {
    const int b = (k / TT) % TT;
    const int col = (k / (TT * TT)) * TT + k % TT;
    const c32 v = psi[off + k];
    sh[b * LS + blur_slot(col)] = sqrtf(v.x * v.x + v.y * v.y);
  }
  float w[RMAX + 1];
  load_blur_weights<RMAX>(tab, p0, wsh, w);   // includes the barrier
  // outputs staged in the band's (contiguous) order, then stored as float4s
  __shared__ __align__(16) float stg[N * TT];
  for (int seg = threadIdx.x; seg < N * TT / K; seg += 256) {
    const int b = seg / (N / K), c0 = (seg % (N / K)) * K;
    float acc[K];
    blur_window<N, K, RMAX>(sh + b * LS, c0, w, acc);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int col = c0 + k;
      stg[sw_swz((col / TT) * (TT * TT) + b * TT + (col % TT))] = acc[k];
    }
  }
  __syncthreads();
  sw_store_staged<N, TT>(stg, tmp + off, [](int v) { return 4 * v; });
}

template <int N, int TT, int RMAX>
__global__ void __launch_bounds__(256) k_blur_cols(const float* tmp, float* blur,
                                                   float* maxp, const SwParam* tab,
                                                   SwParam p0, const int* fmap) {
  constexpr int NB = N / TT;
  constexpr int K = 8;
  constexpr int LS = N + N / 8 + 8;
  __shared__ float sh[TT * LS];
  __shared__ float wsh[RMAX + 1];
  __shared__ float red[256];
  const int band = xcd_band<NB>(blockIdx.x), f = fmap ? fmap[blockIdx.y] : blockIdx.y;
  const size_t fo = (size_t)f * N * N;
  for (int k = threadIdx.x; k < N * TT; k += 256) {
    const int tr = k / (TT * TT), within = k % (TT * TT);
    const int row = tr * TT + within / TT, ci = within % TT;
    sh[ci * LS + blur_slot(row)] = tmp[fo + (size_t)(tr * NB + band) * (TT * TT) + within];
  }
  float w[RMAX + 1];
  load_blur_weights<RMAX>(tab, p0, wsh, w);
  // outputs staged in tile order (tile q = rows qTT..qTT+TT-1 of the band),
  // then stored as float4s
  __shared__ __align__(16) float stg[N * TT];
  float mx = -INFINITY;
  for (int seg = threadIdx.x; seg < N * TT / K; seg += 256) {
    const int ci = seg / (N / K), r0 = (seg % (N / K)) * K;
    float acc[K];
    blur_window<N, K, RMAX>(sh + ci * LS, r0, w, acc);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int row = r0 + k;
      stg[sw_swz((row / TT) * (TT * TT) + (row % TT) * TT + ci)] = acc[k];
      mx = fmaxf(mx, acc[k]);
    }
  }
  __syncthreads();
  sw_store_staged<N, TT>(stg, blur + fo, [&](int v) {
    const int q = (4 * v) / (TT * TT), within = (4 * v) % (TT * TT);
    return (q * NB + band) * (TT * TT) + within;
  });
  const float m = block_max(mx, red);
  if (threadIdx.x == 0) maxp[(size_t)f * NB + band] = m;
}

// grid (n*n/1024, frames): 4 elements per thread
static __global__ void __launch_bounds__(256) k_sw_threshold(float* sup, long long sstride,
                                                             const float* blur, const float* maxp,
                                                             int nb, long long count,
                                                             const SwParam* tab, SwParam p0,
                                                             const int* fmap) {
  __shared__ float red[256];
  const int f = fmap ? fmap[blockIdx.y] : blockIdx.y;
  const float tau = tab ? tab[blockIdx.y].tau : p0.tau;
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
// Generic periodic 1-D convolution along one axis of a row-major
// (rows, cols) float image: out[x] = sum_j w[j] * in[(x - off[j]) mod n]
// along the axis (refine_support's FFT blur, phasing.py:1787-1790, written as
// its equivalent circular convolution with the taps that matter).
// axis 0: along rows (index r), axis 1: along columns (index c).
// ---------------------------------------------------------------------------
static __global__ void __launch_bounds__(256) k_conv_axis(const float* in, float* out, int rows,
                                                          int cols, int axis, const int* off,
                                                          const float* w, int taps) {
  const long long total = (long long)rows * cols;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (long long)gridDim.x * 256) {
    const int r = (int)(i / cols), c = (int)(i % cols);
    double acc = 0.0;
    if (axis == 0) {
      for (int j = 0; j < taps; ++j) {
        int rr = r - off[j];
        rr %= rows;
        if (rr < 0) rr += rows;
        acc += (double)w[j] * in[(size_t)rr * cols + c];
      }
    } else {
      for (int j = 0; j < taps; ++j) {
        int cc = c - off[j];
        cc %= cols;
        if (cc < 0) cc += cols;
        acc += (double)w[j] * in[(size_t)r * cols + cc];
      }
    }
    out[i] = (float)acc;
  }
}

// |x| of a complex image into a float image
static __global__ void __launch_bounds__(256) k_abs_c64(const c32* in, float* out, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    out[i] = sqrtf(in[i].x * in[i].x + in[i].y * in[i].y);
}

// zero-pad (frames, r, c) into (frames, n, n), top-left (prtf / rftf embedding)
static __global__ void __launch_bounds__(256) k_embed(const c32* in, c32* out, int r, int c,
                                                      int n, int frames) {
  const long long total = (long long)frames * n * n;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (long long)gridDim.x * 256) {
    const int f = (int)(i / ((long long)n * n));
    const int rem = (int)(i % ((long long)n * n));
    const int y = rem / n, x = rem % n;
    out[i] = (y < r && x < c) ? in[((size_t)f * r + y) * c + x] : c32{0.0f, 0.0f};
  }
}

// PRTF (phasing.py:1443-1450, phasing_prtf_accumulate2.cl:16-25):
// | (1/frames) sum_f F_f[k] / |F_f[k]| | with FP64 accumulation, written
// centre-shifted (fftshift, even n) as the reference returns it; like the
// reference there is no guard where |F| = 0
static __global__ void __launch_bounds__(256) k_prtf(const c32* F, int frames, int n, float* out) {
  const long long nn = (long long)n * n;
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < nn;
       k += (long long)gridDim.x * 256) {
    double re = 0.0, im = 0.0;
    for (int f = 0; f < frames; ++f) {
      const c32 v = F[(size_t)f * nn + k];
      const double h = sqrt((double)v.x * v.x + (double)v.y * v.y);
      re += v.x / h;
      im += v.y / h;
    }
    re /= frames;
    im /= frames;
    const int y = (int)(k / n), x = (int)(k % n);
    out[(size_t)((y + n / 2) % n) * n + (x + n / 2) % n] = (float)sqrt(re * re + im * im);
  }
}

// RFTF alignment sums (phasing.py:1564-1565): d1 = sum |F - g|,
// d2 = sum |shift(F) - g| with shift = fftshift (even n); block partials
static __global__ void __launch_bounds__(256) k_rftf_diff(const c32* F, const float* goal, int n,
                                                          double* partial) {
  __shared__ double sh[256];
  double d1 = 0.0, d2 = 0.0;
  const long long nn = (long long)n * n;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256) {
    const int y = (int)(i / n), x = (int)(i % n);
    const c32 a = F[i];
    const c32 b = F[(size_t)((y + n / 2) % n) * n + (x + n / 2) % n];
    const double g = goal[i];
    d1 += fabs(sqrt((double)a.x * a.x + (double)a.y * a.y) - g);
    d2 += fabs(sqrt((double)b.x * b.x + (double)b.y * b.y) - g);
  }
  const double s1 = block_sum(d1, sh);
  if (threadIdx.x == 0) partial[2 * blockIdx.x] = s1;
  const double s2 = block_sum(d2, sh);
  if (threadIdx.x == 0) partial[2 * blockIdx.x + 1] = s2;
}

// RFTF value (phasing.py:1567-1581): err = (F' - g)^2 / g^2 with F' = |F| or
// shift(|F|) (whichever alignment is closer, decided from the partial sums),
// output sqrt(1 / (1 + err)) in the reference's centre-shifted frame, with
// the optional [1/sqrt(2), 1] -> [0, 1] rescale
static __global__ void __launch_bounds__(256) k_rftf_out(const c32* F, const float* goal, int n,
                                                         const double* partial, int nblk,
                                                         int scale, float* out) {
  double d1 = 0.0, d2 = 0.0;
  for (int b = 0; b < nblk; ++b) {
    d1 += partial[2 * b];
    d2 += partial[2 * b + 1];
  }
  const bool centre = d1 > d2;  // goal modulus centred at n/2
  const long long nn = (long long)n * n;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256) {
    const int y = (int)(i / n), x = (int)(i % n);
    // output pixel i; both branches end centre-shifted (phasing.py:1567-1573)
    const long long j = (long long)((y + n / 2) % n) * n + (x + n / 2) % n;
    double fa, g;
    if (centre) {  // error = (shift(F) - g)^2 / g^2 at i
      const c32 a = F[j];
      fa = sqrt((double)a.x * a.x + (double)a.y * a.y);
      g = goal[i];
    } else {       // error = shift((F - g)^2 / g^2): value of pixel j
      const c32 a = F[j];
      fa = sqrt((double)a.x * a.x + (double)a.y * a.y);
      g = goal[j];
    }
    const double e = (fa - g) * (fa - g) / (g * g);
    double v = sqrt(1.0 / (1.0 + e));
    if (scale) {
      const double q = 1.0 / sqrt(2.0);
      v = (v - q) / (1.0 - q);
    }
    out[i] = (float)v;
  }
}

// per-frame RFTF averaged over a stack (gpu_rftf, phasing.py:1611-1708 with
// phasing_rftf_accumulate.cl:15-24): (1/frames) sum_f 1/sqrt(1 + (g - |F_f|)^2/g^2)
// with the goal at the corner, FP64 sums, written centre-shifted
static __global__ void __launch_bounds__(256) k_rftf_stack(const c32* F, const float* goal,
                                                           int frames, int n, float* out) {
  const long long nn = (long long)n * n;
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < nn;
       k += (long long)gridDim.x * 256) {
    const double g = goal[k];
    double acc = 0.0;
    for (int f = 0; f < frames; ++f) {
      const c32 v = F[(size_t)f * nn + k];
      const double d1 = g - sqrt((double)v.x * v.x + (double)v.y * v.y);
      acc += 1.0 / sqrt(1.0 + d1 * d1 / (g * g));
    }
    const int y = (int)(k / n), x = (int)(k % n);
    out[(size_t)((y + n / 2) % n) * n + (x + n / 2) % n] = (float)(acc / frames);
  }
}

// ---------------------------------------------------------------------------
// Spectral (longitudinal) blur, phasing_radial_rescale.cl restated: output
// pixel (row j, column i) has signed coordinates x = i (i < n/2) or i - n,
// y likewise; for each sample k the intensity is bilinearly interpolated at
// (x / r_k, y / r_k) (FP32, as the reference) and summed with weight w_k.
// A node with floor coordinate <= -n/2 or >= n/2 (and its +1 neighbour
// likewise) has weight 0 and is not read; kept nodes wrap by +n.  Maps are
// column-band-major with tile tt.  spec: K factors, then K weights.
// grid (ceil(n*n/256), frames)
// ---------------------------------------------------------------------------
static __device__ __forceinline__ float rescale_node(const float* img, int n, int tt, int yi,
                                                     int xi) {
  return img[(size_t)(xi / tt) * n * tt + (size_t)yi * tt + (xi % tt)];
}

static __global__ void __launch_bounds__(256) k_rescale(const float* imap, float* dmap,
                                                        const float* spec, int K, int n, int tt,
                                                        const int* fmap) {
  const int f = fmap ? fmap[blockIdx.y] : blockIdx.y;
  const long long nn = (long long)n * n;
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p >= nn) return;
  const float* img = imap + (size_t)f * nn;
  // column-band-major position p -> (row j, column i)
  const int band = (int)(p / ((long long)n * tt)), rem = (int)(p % ((long long)n * tt));
  const int j = rem / tt, i = band * tt + rem % tt;
  const int h = n / 2;
  const float x = (float)(i > h - 1 ? i - n : i);
  const float y = (float)(j > h - 1 ? j - n : j);
  float acc = 0.0f;
  for (int k = 0; k < K; ++k) {
    const float rv = spec[k], sw = spec[K + k];
    const float rx = x / rv, ry = y / rv;
    const int x1 = (int)floorf(rx), y1 = (int)floorf(ry);
    const int x2 = x1 + 1, y2 = y1 + 1;
    const bool kx1 = x1 > -h && x1 < h, kx2 = x2 > -h && x2 < h;
    const bool ky1 = y1 > -h && y1 < h, ky2 = y2 > -h && y2 < h;
    const float wx1 = kx1 ? 1.0f - fabsf((float)x1 - rx) : 0.0f;
    const float wx2 = kx2 ? 1.0f - fabsf((float)x2 - rx) : 0.0f;
    const float wy1 = ky1 ? 1.0f - fabsf((float)y1 - ry) : 0.0f;
    const float wy2 = ky2 ? 1.0f - fabsf((float)y2 - ry) : 0.0f;
    const int ix1 = x1 < 0 ? x1 + n : x1, ix2 = x2 < 0 ? x2 + n : x2;
    const int iy1 = y1 < 0 ? y1 + n : y1, iy2 = y2 < 0 ? y2 + n : y2;
    const float v11 = (kx1 && ky1) ? rescale_node(img, n, tt, iy1, ix1) : 0.0f;
    const float v12 = (kx1 && ky2) ? rescale_node(img, n, tt, iy2, ix1) : 0.0f;
    const float v21 = (kx2 && ky1) ? rescale_node(img, n, tt, iy1, ix2) : 0.0f;
    const float v22 = (kx2 && ky2) ? rescale_node(img, n, tt, iy2, ix2) : 0.0f;
    const float iv = v11 * wx1 * wy1 + v12 * wx1 * wy2 + v21 * wx2 * wy1 + v22 * wx2 * wy2;
    acc = fmaf(iv, sw, acc);
  }
  dmap[(size_t)f * nn + p] = acc;
}

// ---------------------------------------------------------------------------
// Coherence estimation (optimize_coherence, phasing.py:507-930): elementwise
// pieces around the library's 2-D transforms.  Arrays are in the plan's
// tiled layout (tile tt); position p of a tiled array is row-major (r, c)
// through tile_rc.
// ---------------------------------------------------------------------------
// signed coordinate of index i after fftshift: ((i + n/2) mod n) - n/2
static __device__ __forceinline__ int shifted_coord(int i, int n) {
  return ((i + n / 2) & (n - 1)) - n / 2;
}

static __device__ __forceinline__ void tiled_rc(int n, int tt, long long p, int& r, int& c) {
  const long long tile = p / (tt * tt);
  const int w = (int)(p % (tt * tt));
  const int nb = n / tt;
  r = (int)(tile / nb) * tt + w / tt;
  c = (int)(tile % nb) * tt + w % tt;
}

static __device__ __forceinline__ long long tiled_pos(int n, int tt, int r, int c) {
  return ((long long)(r / tt) * (n / tt) + c / tt) * (tt * tt) + (r % tt) * tt + (c % tt);
}

// a = |a|, |a|^2 or sqrt|a| (mode 0, 1, 2), imaginary part 0
static __global__ void __launch_bounds__(256) k_abs_mode(c32* a, long long nn, int mode) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256) {
    const float p2 = fmaf(a[i].x, a[i].x, a[i].y * a[i].y);
    const float v = mode == 1 ? p2 : mode == 0 ? sqrtf(p2) : sqrtf(sqrtf(p2));
    a[i] = c32{v, 0.0f};
  }
}

// a *= sc
static __global__ void __launch_bounds__(256) k_scale_c(c32* a, long long nn, float sc) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256)
    a[i] = scale(a[i], sc);
}

// out = a * b
static __global__ void __launch_bounds__(256) k_cmul(const c32* a, const c32* b, c32* out,
                                                     long long nn) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256)
    out[i] = mul(a[i], b[i]);
}

// out = d / (w * scale), d real (the Richardson-Lucy quotient, phasing.py:720-723)
static __global__ void __launch_bounds__(256) k_rl_divide(const float* d, const c32* w, c32* out,
                                                          long long nn, float sc) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256) {
    const c32 v = scale(w[i], sc);
    const float den = fmaf(v.x, v.x, v.y * v.y);
    out[i] = c32{d[i] * v.x / den, -d[i] * v.y / den};
  }
}

// u *= w * scale (phasing.py:733-736)
static __global__ void __launch_bounds__(256) k_rl_update(c32* u, const c32* w, long long nn,
                                                          float sc) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256)
    u[i] = mul(u[i], scale(w[i], sc));
}

// x = |x|^2 (+0i) with FP64 block partial sums of |x|^2
static __global__ void __launch_bounds__(256) k_abs2_sum(c32* x, long long nn, double* partial) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256) {
    const float v = fmaf(x[i].x, x[i].x, x[i].y * x[i].y);
    x[i] = c32{v, 0.0f};
    acc += v;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// x /= sum of the nblk partials (each thread adds them in order: deterministic)
static __global__ void __launch_bounds__(256) k_div_by_partials(c32* x, float* xr, long long nn,
                                                                const double* partial, int nblk) {
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[b];
  const float inv = (float)(1.0 / s);
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256) {
    if (x) x[i] = scale(x[i], inv);
    if (xr) xr[i] *= inv;
  }
}

// d = m^2 with FP64 partial sums (m tiled, in place)
static __global__ void __launch_bounds__(256) k_sq_sum(float* m, long long nn, double* partial) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256) {
    const float v = m[i] * m[i];
    m[i] = v;
    acc += v;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out(r, c) = in(n-1-r, n-1-c): the plain reversal P[::-1, ::-1]
// (common_flip_f2.cl), tiled in and out
static __global__ void __launch_bounds__(256) k_flip(const c32* in, c32* out, int n, int tt) {
  const long long nn = (long long)n * n;
  for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < nn;
       p += (long long)gridDim.x * 256) {
    int r, c;
    tiled_rc(n, tt, p, r, c);
    out[p] = in[tiled_pos(n, tt, n - 1 - r, n - 1 - c)];
  }
}

// out = ac * exp(-y^2/(2 s0^2)) exp(-x^2/(2 s1^2)) on fftshifted signed
// coordinates (the Gaussian of _opt_iter_g, phasing.py:663-670), FP64 exp;
// with ac = nullptr, out = the Gaussian itself (the Richardson-Lucy start,
// phasing.py:818, widths n/4)
static __global__ void __launch_bounds__(256) k_gauss_mul(const c32* ac, c32* out, int n, int tt,
                                                          double s0, double s1) {
  const long long nn = (long long)n * n;
  for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < nn;
       p += (long long)gridDim.x * 256) {
    int r, c;
    tiled_rc(n, tt, p, r, c);
    const double y = shifted_coord(r, n), x = shifted_coord(c, n);
    const float g = (float)(exp(-y * y / (2.0 * s0 * s0)) * exp(-x * x / (2.0 * s1 * s1)));
    out[p] = ac ? scale(ac[p], g) : c32{g, 0.0f};
  }
}

// FP64 partials of sum (sqrt|W| - m)^2 (phasing.py:672-698)
static __global__ void __launch_bounds__(256) k_gauss_resid(const c32* W, const float* m,
                                                            long long nn, double* partial) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256) {
    const double a = sqrt(sqrt((double)W[i].x * W[i].x + (double)W[i].y * W[i].y));
    const double d = a - (double)m[i];
    acc = fma(d, d, acc);
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// final ipsf: optionally the real part (the CPU branch's float32 cast,
// phasing.py:778-783), divided by |u(0,0)| (phasing.py:641-643); position 0
// of a tiled array is (0,0).  norm[0] is written by the first launch
// (which = 0) and applied by the second (which = 1).
static __global__ void __launch_bounds__(256) k_rl_finish(c32* u, long long nn, int real_only,
                                                          double* norm, int which) {
  if (which == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      const c32 v = u[0];
      norm[0] = real_only ? fabs((double)v.x) : sqrt((double)v.x * v.x + (double)v.y * v.y);
    }
    return;
  }
  const float inv = (float)(1.0 / norm[0]);
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn;
       i += (long long)gridDim.x * 256) {
    const c32 v = u[i];
    u[i] = real_only ? c32{v.x * inv, 0.0f} : scale(v, inv);
  }
}

// ---------------------------------------------------------------------------
// center_of_mass_average (phasing.py:1710-1745): per frame the FP64 sums of
// |img|, row * |img| and col * |img| (block partials [frame][block][3]), then
// out = sum_f roll(img_f, -(dr_f, dc_f)) with FP64 accumulation.
// ---------------------------------------------------------------------------
static __global__ void __launch_bounds__(256) k_com_partial(const c32* data, int rows, int cols,
                                                            double* partial, int nblk) {
  __shared__ double sh[256];
  const long long per = (long long)rows * cols;
  const c32* d = data + (size_t)blockIdx.y * per;
  double s0 = 0.0, sr = 0.0, sc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < per;
       i += (long long)nblk * 256) {
    const double a = sqrt((double)d[i].x * d[i].x + (double)d[i].y * d[i].y);
    s0 += a;
    sr += a * (double)(i / cols);
    sc += a * (double)(i % cols);
  }
  double* out = partial + ((size_t)blockIdx.y * nblk + blockIdx.x) * 3;
  const double t0 = block_sum(s0, sh);
  if (threadIdx.x == 0) out[0] = t0;
  const double t1 = block_sum(sr, sh);
  if (threadIdx.x == 0) out[1] = t1;
  const double t2 = block_sum(sc, sh);
  if (threadIdx.x == 0) out[2] = t2;
}

static __global__ void __launch_bounds__(256) k_roll_sum(const c32* data, int frames, int rows,
                                                         int cols, const int* shift, c32* out) {
  const long long per = (long long)rows * cols;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < per;
       i += (long long)gridDim.x * 256) {
    const int r = (int)(i / cols), c = (int)(i % cols);
    double re = 0.0, im = 0.0;
    for (int f = 0; f < frames; ++f) {
      int rr = (r + shift[2 * f]) % rows, cc = (c + shift[2 * f + 1]) % cols;
      if (rr < 0) rr += rows;
      if (cc < 0) cc += cols;
      const c32 v = data[(size_t)f * per + (size_t)rr * cols + cc];
      re += v.x;
      im += v.y;
    }
    out[i] = c32{(float)re, (float)im};
  }
}

// ---------------------------------------------------------------------------
// mean over a stack of frames (the trial average, phasing.py:960-961):
// out[i] = (1/frames) sum_f data[f][i], summed in FP64 in frame order
// ---------------------------------------------------------------------------
static __global__ void __launch_bounds__(256) k_stack_mean(const c32* data, long long per_frame,
                                                           int frames, c32* out) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < per_frame;
       i += (long long)gridDim.x * 256) {
    double re = 0.0, im = 0.0;
    for (int f = 0; f < frames; ++f) {
      const c32 v = data[(size_t)f * per_frame + i];
      re += v.x;
      im += v.y;
    }
    out[i] = c32{(float)(re / frames), (float)(im / frames)};
  }
}

// ---------------------------------------------------------------------------
// plain copy (HBM counter calibration)
// ---------------------------------------------------------------------------
// Plain device copy (the HBM reference point and the counters' calibration):
// each thread moves four 16-byte vectors, all four loads issued before the
// stores, one pass over the buffer (no grid-stride loop).
static __global__ void __launch_bounds__(256) k_copy16(uint4* dst, const uint4* src, size_t n16) {
  const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
  uint4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const size_t i = base + (size_t)k * 256;
    if (i < n16) v[k] = src[i];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const size_t i = base + (size_t)k * 256;
    if (i < n16) dst[i] = v[k];
  }
}

// ---------------------------------------------------------------------------
// Mixed-size passes: one launch covers the frames of several plans of
// different sizes (256^2, 512^2, 1024^2), each in the tile that makes its
// workgroup 256 threads of eight values (256^2: tile 8, 512^2: tile 4,
// 1024^2: tile 2), so every workgroup of the launch has the same shape and
// runs the same per-band code as that size's own pass kernel (bitwise the
// same results).  The LDS is the largest of the three sizes' plans.  A
// workgroup finds its member by the start table (at most kGroupMax
// entries), then its frame and band; bands are numbered per frame as in the
// single-size launch (XCD-aware for the column pass: every member's first
// workgroup is a multiple of 8, since every size has a multiple of 8 bands).
// ---------------------------------------------------------------------------
template <int N>
constexpr int group_lds() { return LdsPlan<N, group_tile(N)>::TOTAL; }
constexpr int kGroupLds = group_lds<256>() > group_lds<512>()
                              ? (group_lds<256>() > group_lds<1024>() ? group_lds<256>()
                                                                      : group_lds<1024>())
                              : (group_lds<512>() > group_lds<1024>() ? group_lds<512>()
                                                                      : group_lds<1024>());

__device__ __forceinline__ int group_member(const GroupPass& g) {
  int m = 0;
  while (m + 1 < g.count && (int)blockIdx.x >= g.start[m + 1]) ++m;
  return m;
}

template <int N, int MODE>
__device__ __forceinline__ void colpass_group_item(const PassArgs& a, int local, c32* lds) {
  constexpr int TT = group_tile(N), NB = N / TT;
  static_assert((N / kGroupEpt) * TT == kGroupThreads, "one workgroup shape for every member");
  static_assert(MODE != 9 || TT * sw_line<N>() * sizeof(float) <= kGroupLds * sizeof(c32),
                "fused blur lines exceed the pass's LDS");
  const int fi = local / NB, band = xcd_band<NB>(local % NB);
  const int f = a.fmap ? a.fmap[fi] : fi;
  const Coef c = a.tab ? a.tab[f] : a.c;
  if (!(c.flags & kActive)) return;
  colpass_item<N, TT, kGroupEpt, MODE, true>(a, f, band, 0, a.errp, lds,
                                              MODE == 9 && (c.flags & kSwFin));
}

template <int N, int SW>
__device__ __forceinline__ void rowpass_group_item(const PassArgs& a, int local, c32* lds) {
  constexpr int TT = group_tile(N), NB = N / TT;
  static_assert(SW == 0 || TT * sw_line<N>() * sizeof(float) <= kGroupLds * sizeof(c32),
                "fused blur lines exceed the pass's LDS");
  const int fi = local / NB, band = local % NB;
  const int f = a.fmap ? a.fmap[fi] : fi;
  const Coef c = a.tab ? a.tab[f] : a.c;
  if (!(c.flags & kActive)) return;
  rowpass_item<N, TT, kGroupEpt, false, false, true, SW, false>(a, f, band, c, lds);
}

template <int MODE>
__global__ void __launch_bounds__(kGroupThreads) k_colpass_group(GroupPass g) {
  __shared__ c32 lds[kGroupLds];
  const int m = group_member(g);
  const PassArgs a = g.args[m];
  const int local = (int)blockIdx.x - g.start[m];
  switch (g.n[m]) {
    case 256: colpass_group_item<256, MODE>(a, local, lds); break;
    case 512: colpass_group_item<512, MODE>(a, local, lds); break;
    case 1024: colpass_group_item<1024, MODE>(a, local, lds); break;
    default: break;
  }
}

template <int SW>
__global__ void __launch_bounds__(kGroupThreads) k_rowpass_group(GroupPass g) {
  __shared__ c32 lds[kGroupLds];
  const int m = group_member(g);
  const PassArgs a = g.args[m];
  const int local = (int)blockIdx.x - g.start[m];
  switch (g.n[m]) {
    case 256: rowpass_group_item<256, SW>(a, local, lds); break;
    case 512: rowpass_group_item<512, SW>(a, local, lds); break;
    case 1024: rowpass_group_item<1024, SW>(a, local, lds); break;
    default: break;
  }
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
#define SPK_BUILD_GROUP 1
#elif SPK_ONLY_N == 0
#define SPK_BUILD_COMMON 1
#define SPK_BUILD_SIZE(n) 0
#define SPK_BUILD_GROUP 0
#elif SPK_ONLY_N == -1
// the mixed-size pass kernels
#define SPK_BUILD_COMMON 0
#define SPK_BUILD_SIZE(n) 0
#define SPK_BUILD_GROUP 1
#else
#define SPK_BUILD_COMMON 0
#define SPK_BUILD_SIZE(n) ((n) == SPK_ONLY_N)
#define SPK_BUILD_GROUP 0
#endif

#if SPK_BUILD_GROUP
hipError_t launch_colpass_group(int mode, const GroupPass& g, hipStream_t st) {
  if (g.count < 1 || g.count > kGroupMax) return hipErrorInvalidValue;
  const dim3 grid(g.start[g.count]), block(kGroupThreads);
  if (grid.x == 0) return hipSuccess;
  if (mode == 0)
    hipLaunchKernelGGL(k_colpass_group<0>, grid, block, 0, st, g);
  else if (mode == 9)
    hipLaunchKernelGGL(k_colpass_group<9>, grid, block, 0, st, g);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_rowpass_group(int sw, const GroupPass& g, hipStream_t st) {
  if (g.count < 1 || g.count > kGroupMax) return hipErrorInvalidValue;
  const dim3 grid(g.start[g.count]), block(kGroupThreads);
  if (grid.x == 0) return hipSuccess;
  switch (sw) {
    case 0: hipLaunchKernelGGL(k_rowpass_group<0>, grid, block, 0, st, g); break;
    case 1: hipLaunchKernelGGL(k_rowpass_group<1>, grid, block, 0, st, g); break;
    case 2: hipLaunchKernelGGL(k_rowpass_group<2>, grid, block, 0, st, g); break;
    case 3: hipLaunchKernelGGL(k_rowpass_group<3>, grid, block, 0, st, g); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
#endif  // SPK_BUILD_GROUP

#define SPK_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

// launch a pass kernel, through hipExtLaunchKernelGGL when timing events are set
template <class K>
static void launch_pass(K kernel, dim3 grid, dim3 block, hipStream_t st, const PassArgs& a) {
  if (g_launch_events[0]) {
    hipExtLaunchKernelGGL(kernel, grid, block, 0, st, g_launch_events[0], g_launch_events[1], 0,
                          a);
    g_launch_events[0] = g_launch_events[1] = nullptr;
  } else {
    hipLaunchKernelGGL(kernel, grid, block, 0, st, a);
  }
}

template <int N, int TT, int E>
static hipError_t colpass_nt(int mode, const PassArgs& a, int frames, int zdim,
                             hipStream_t st) {
  dim3 grid(N / TT, frames, zdim), block((N / E) * TT);
  switch (mode) {
    case 0: launch_pass(k_colpass<N, TT, E, 0>, grid, block, st, a); break;
    case 2: launch_pass(k_colpass<N, TT, E, 2>, grid, block, st, a); break;
    case 3: launch_pass(k_colpass<N, TT, E, 3>, grid, block, st, a); break;
    case 4: launch_pass(k_colpass<N, TT, E, 4>, grid, block, st, a); break;
    case 5: launch_pass(k_colpass<N, TT, E, 5>, grid, block, st, a); break;
    case 6: launch_pass(k_colpass<N, TT, E, 6>, grid, block, st, a); break;
    case 7: launch_pass(k_colpass<N, TT, E, 7>, grid, block, st, a); break;
    case 8: launch_pass(k_colpass<N, TT, E, 8>, grid, block, st, a); break;
    case 9: launch_pass(k_colpass<N, TT, E, 9>, grid, block, st, a); break;
    case 10: launch_pass(k_colpass<N, TT, E, 10>, grid, block, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N, int TT, int E>
static hipError_t rowpass_nt(bool use_t, bool emit_b, int sw, bool noout, const PassArgs& a,
                             int frames, hipStream_t st) {
  dim3 grid(N / TT, frames), block((N / E) * TT);
  if (sw && (use_t || emit_b)) return hipErrorInvalidValue;
  if (noout) {
    if (use_t || emit_b || sw) return hipErrorInvalidValue;
    launch_pass(k_rowpass<N, TT, E, false, false, 0, true>, grid, block, st, a);
    return hipGetLastError();
  }
  if (use_t && emit_b)
    launch_pass(k_rowpass<N, TT, E, true, true, 0>, grid, block, st, a);
  else if (use_t)
    launch_pass(k_rowpass<N, TT, E, true, false, 0>, grid, block, st, a);
  else if (emit_b)
    launch_pass(k_rowpass<N, TT, E, false, true, 0>, grid, block, st, a);
  else if (sw == 1)
    launch_pass(k_rowpass<N, TT, E, false, false, 1>, grid, block, st, a);
  else if (sw == 2)
    launch_pass(k_rowpass<N, TT, E, false, false, 2>, grid, block, st, a);
  else if (sw == 3)
    launch_pass(k_rowpass<N, TT, E, false, false, 3>, grid, block, st, a);
  else
    launch_pass(k_rowpass<N, TT, E, false, false, 0>, grid, block, st, a);
  return hipGetLastError();
}

template <int N, int TT, int E>
static hipError_t persist_nt(const PassArgs& a, const unsigned char* stage, const PStep* steps,
                             int nsteps, double* errp0, long long errp_step, unsigned* bar,
                             int* flag, int grid, hipStream_t st) {
  // the barrier counter and this launch's give-up flag (bar[1] == flag[0])
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, bar, 2);
  void* args[] = {(void*)&a, (void*)&stage, (void*)&steps, (void*)&nsteps, (void*)&errp0,
                  (void*)&errp_step, (void*)&bar, (void*)&flag};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_persist<N, TT, E>),
                                    dim3(grid), dim3((N / E) * TT), args, 0, st);
}

// co-resident workgroups of the persistent kernel on the current device
template <int N, int TT, int E>
static hipError_t persist_grid_nt(int* out) {
  int occ = 0, dev = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &occ, reinterpret_cast<const void*>(&k_persist<N, TT, E>), (N / E) * TT, 0);
  if (e != hipSuccess) return e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  *out = occ * cus;
  return hipSuccess;
}

template <int N, int TT, int E>
static hipError_t rowconv_nt(const c32* in, c32* out, const c32* ipsf, const c32* tw,
                             const int* fmap, int frames, hipStream_t st) {
  hipLaunchKernelGGL((k_rowconv<N, TT, E>), dim3(N / TT, frames), dim3((N / E) * TT), 0, st, in,
                     out, ipsf, tw, fmap);
  return hipGetLastError();
}

template <int N, int TT, int E>
static hipError_t rowfft_nt(int mode, const c32* in, c32* out, const float* sup,
                            long long sstride, float g, const c32* tw, const int* fmap,
                            int frames, hipStream_t st) {
  dim3 grid(N / TT, frames), block((N / E) * TT);
  switch (mode) {
    case 0: hipLaunchKernelGGL((k_rowfft<N, TT, E, 0>), grid, block, 0, st, in, out, sup, sstride, g, tw, fmap); break;
    case 1: hipLaunchKernelGGL((k_rowfft<N, TT, E, 1>), grid, block, 0, st, in, out, sup, sstride, g, tw, fmap); break;
    case 2: hipLaunchKernelGGL((k_rowfft<N, TT, E, 2>), grid, block, 0, st, in, out, sup, sstride, g, tw, fmap); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N, int TT, int E>
static hipError_t resident_nt(int xm, const PassArgs& a, const RStep* steps, int nsteps,
                              double* errp0, long long errp_step, unsigned* flags, int* flag,
                              hipStream_t st) {
  // the workgroups' barrier flags and this launch's give-up flag start at 0
  constexpr int words = (N / TT) * 32;
  hipLaunchKernelGGL(k_zero_u32, dim3((words + 255) / 256), dim3(256), 0, st, flags, words);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned*>(flag), 1);
  void* args[] = {(void*)&a, (void*)&steps, (void*)&nsteps, (void*)&errp0, (void*)&errp_step,
                  (void*)&flags, (void*)&flag};
  const void* fn = xm == 1   ? reinterpret_cast<const void*>(&k_resident<N, TT, E, 1>)
                   : xm == 2 ? reinterpret_cast<const void*>(&k_resident<N, TT, E, 2>)
                             : reinterpret_cast<const void*>(&k_resident<N, TT, E, 0>);
  if (g_launch_events[0]) {
    (void)hipEventRecord(g_launch_events[0], st);
  }
  hipError_t e = hipLaunchCooperativeKernel(fn, dim3(N / TT), dim3((N / E) * TT), args, 0, st);
  if (g_launch_events[0]) {
    (void)hipEventRecord(g_launch_events[1], st);
    g_launch_events[0] = g_launch_events[1] = nullptr;
  }
  return e;
}

// the resident form exists for one frame of 256^2 .. 1024^2 in tile 4 with
// 8 values per thread (one workgroup of at most 512 threads per band, at
// most 256 bands: one per CU), and for 256^2 and 512^2 also in tile 2 (128
// and 256 bands: more CUs, half the lines per workgroup)
template <int N>
static hipError_t resident_dispatch(int tile, int ept, int xm, const PassArgs& a,
                                    const RStep* steps, int nsteps, double* errp0,
                                    long long errp_step, unsigned* bar, int* flag,
                                    hipStream_t st) {
  if constexpr (N >= 256 && N <= 1024) {
    if (tile == 4 && ept == 8)
      return resident_nt<N, 4, 8>(xm, a, steps, nsteps, errp0, errp_step, bar, flag, st);
    if constexpr (N <= 512) {
      if (tile == 2 && ept == 8)
        return resident_nt<N, 2, 8>(xm, a, steps, nsteps, errp0, errp_step, bar, flag, st);
    }
  }
  return hipErrorInvalidValue;
}

template <int N, int TT>
static hipError_t shrinkwrap_nt(const c32* psi, float* sup, long long sstride, float* tmp,
                                float* blur, float* maxp, const SwParam* tab, const SwParam& p0,
                                const int* fmap, int frames, int rmax, hipStream_t st) {
  dim3 grid(N / TT, frames);
  // register window sized for the largest radius among the frames
  if (rmax <= kSwFuseRmax) {
    hipLaunchKernelGGL((k_blur_rows<N, TT, kSwFuseRmax>), grid, dim3(256), 0, st, psi, tmp, tab, p0, fmap);
    hipLaunchKernelGGL((k_blur_cols<N, TT, kSwFuseRmax>), grid, dim3(256), 0, st, tmp, blur, maxp, tab,
                       p0, fmap);
  } else {
    hipLaunchKernelGGL((k_blur_rows<N, TT, 48>), grid, dim3(256), 0, st, psi, tmp, tab, p0, fmap);
    hipLaunchKernelGGL((k_blur_cols<N, TT, 48>), grid, dim3(256), 0, st, tmp, blur, maxp, tab,
                       p0, fmap);
  }
  const long long count = (long long)N * N;
  dim3 g2((unsigned)((count + 1023) / 1024), frames);
  hipLaunchKernelGGL(k_sw_threshold, g2, dim3(256), 0, st, sup, sstride, blur, maxp,
                     N / TT, count, tab, p0, fmap);
  return hipGetLastError();
}

// Shapes built per frame size: tile 2, 4 and 8 (not 8 at 4096, whose LDS
// exchange would not fit), 16 values per thread, and 8 values per thread
// where the workgroup stays within 1024 threads.
template <int N, int TT, int E>
constexpr bool shape_ok() {
  return !(N == 4096 && TT == 8) && (N / E) * TT <= 1024 && E <= N;
}

#define SPK_SHAPE_SWITCH(n, CALLT)                                       \
  if (tile == 4 && ept == 16) return CALLT(4, 16);                       \
  if (tile == 4 && ept == 8) return CALLT(4, 8);                         \
  if (tile == 2 && ept == 16) return CALLT(2, 16);                       \
  if (tile == 2 && ept == 8) return CALLT(2, 8);                         \
  if (tile == 8 && ept == 16) return CALLT(8, 16);                       \
  if (tile == 8 && ept == 8) return CALLT(8, 8);                         \
  return hipErrorInvalidValue;

#define SPK_SIZED_DECL(n)                                                                   \
  hipError_t colpass_##n(int tile, int ept, int mode, const PassArgs& a, int frames,       \
                         int zdim, hipStream_t st);                                         \
  hipError_t rowpass_##n(int tile, int ept, bool use_t, bool emit_b, int sw, bool noout,  \
                         const PassArgs& a, int frames, hipStream_t st);                    \
  hipError_t rowfft_##n(int tile, int ept, int mode, const c32* in, c32* out,              \
                        const float* sup, long long sstride, float g, const c32* tw,       \
                        const int* fmap, int frames, hipStream_t st);                       \
  hipError_t shrinkwrap_##n(int tile, const c32* psi, float* sup, long long sstride,       \
                            float* tmp, float* blur, float* maxp, const SwParam* tab,      \
                            const SwParam& p0, const int* fmap, int frames, int rmax,      \
                            hipStream_t st);                                                \
  hipError_t persist_##n(int tile, int ept, const PassArgs& a, const unsigned char* stage,  \
                         const PStep* steps, int nsteps, double* errp0, long long errp_step,\
                         unsigned* bar, int* flag, int grid, hipStream_t st);               \
  hipError_t persist_grid_##n(int tile, int ept, int* out);                                 \
  hipError_t rowconv_##n(int tile, int ept, const c32* in, c32* out, const c32* ipsf,       \
                         const c32* tw, const int* fmap, int frames, hipStream_t st);       \
  hipError_t resident_##n(int tile, int ept, int xm, const PassArgs& a, const RStep* steps, \
                          int nsteps, double* errp0, long long errp_step, unsigned* bar,    \
                          int* flag, hipStream_t st);

// a shape that is not built compiles to an error return, not a kernel
#define SPK_CALL_COL(TT_, E_)                                                   \
  (shape_ok<SPK_N, TT_, E_>() ? colpass_nt<SPK_N, (shape_ok<SPK_N, TT_, E_>() ? TT_ : 4), \
                                           (shape_ok<SPK_N, TT_, E_>() ? E_ : 16)>(     \
                                    mode, a, frames, zdim, st)                          \
                              : hipErrorInvalidValue)
#define SPK_CALL_PERSIST(TT_, E_)                                                 \
  (shape_ok<SPK_N, TT_, E_>() ? persist_nt<SPK_N, (shape_ok<SPK_N, TT_, E_>() ? TT_ : 4),  \
                                           (shape_ok<SPK_N, TT_, E_>() ? E_ : 16)>(      \
                                    a, stage, steps, nsteps, errp0, errp_step, bar, flag,  \
                                    grid, st)                                              \
                              : hipErrorInvalidValue)
#define SPK_CALL_PGRID(TT_, E_)                                                   \
  (shape_ok<SPK_N, TT_, E_>()                                                     \
       ? persist_grid_nt<SPK_N, (shape_ok<SPK_N, TT_, E_>() ? TT_ : 4),           \
                         (shape_ok<SPK_N, TT_, E_>() ? E_ : 16)>(out)             \
       : hipErrorInvalidValue)
#define SPK_CALL_CONV(TT_, E_)                                                  \
  (shape_ok<SPK_N, TT_, E_>() ? rowconv_nt<SPK_N, (shape_ok<SPK_N, TT_, E_>() ? TT_ : 4), \
                                           (shape_ok<SPK_N, TT_, E_>() ? E_ : 16)>(     \
                                    in, out, ipsf, tw, fmap, frames, st)                \
                              : hipErrorInvalidValue)
#define SPK_CALL_ROW(TT_, E_)                                                   \
  (shape_ok<SPK_N, TT_, E_>() ? rowpass_nt<SPK_N, (shape_ok<SPK_N, TT_, E_>() ? TT_ : 4), \
                                           (shape_ok<SPK_N, TT_, E_>() ? E_ : 16)>(     \
                                    use_t, emit_b, sw, noout, a, frames, st)            \
                              : hipErrorInvalidValue)
#define SPK_CALL_FFT(TT_, E_)                                                   \
  (shape_ok<SPK_N, TT_, E_>() ? rowfft_nt<SPK_N, (shape_ok<SPK_N, TT_, E_>() ? TT_ : 4),  \
                                          (shape_ok<SPK_N, TT_, E_>() ? E_ : 16)>(      \
                                    mode, in, out, sup, sstride, g, tw, fmap, frames, st) \
                              : hipErrorInvalidValue)

#define SPK_SIZED_DEF(n)                                                                    \
  hipError_t colpass_##n(int tile, int ept, int mode, const PassArgs& a, int frames,       \
                         int zdim, hipStream_t st) {                                        \
    constexpr int SPK_N = n;                                                                \
    SPK_SHAPE_SWITCH(n, SPK_CALL_COL)                                                       \
  }                                                                                         \
  hipError_t rowpass_##n(int tile, int ept, bool use_t, bool emit_b, int sw, bool noout,  \
                         const PassArgs& a, int frames, hipStream_t st) {                   \
    constexpr int SPK_N = n;                                                                \
    SPK_SHAPE_SWITCH(n, SPK_CALL_ROW)                                                       \
  }                                                                                         \
  hipError_t rowfft_##n(int tile, int ept, int mode, const c32* in, c32* out,              \
                        const float* sup, long long sstride, float g, const c32* tw,       \
                        const int* fmap, int frames, hipStream_t st) {                      \
    constexpr int SPK_N = n;                                                                \
    SPK_SHAPE_SWITCH(n, SPK_CALL_FFT)                                                       \
  }                                                                                         \
  hipError_t shrinkwrap_##n(int tile, const c32* psi, float* sup, long long sstride,       \
                            float* tmp, float* blur, float* maxp, const SwParam* tab,      \
                            const SwParam& p0, const int* fmap, int frames, int rmax,      \
                            hipStream_t st) {                                               \
    if (tile == 4) return shrinkwrap_nt<n, 4>(psi, sup, sstride, tmp, blur, maxp, tab, p0, \
                                              fmap, frames, rmax, st);                      \
    if (tile == 2) return shrinkwrap_nt<n, 2>(psi, sup, sstride, tmp, blur, maxp, tab, p0, \
                                              fmap, frames, rmax, st);                      \
    if (tile == 8 && n != 4096)                                                             \
      return shrinkwrap_nt<n, (n == 4096 ? 4 : 8)>(psi, sup, sstride, tmp, blur, maxp, tab, \
                                                   p0, fmap, frames, rmax, st);             \
    return hipErrorInvalidValue;                                                            \
  }                                                                                         \
  hipError_t persist_##n(int tile, int ept, const PassArgs& a, const unsigned char* stage,  \
                         const PStep* steps, int nsteps, double* errp0, long long errp_step,\
                         unsigned* bar, int* flag, int grid, hipStream_t st) {              \
    constexpr int SPK_N = n;                                                                \
    SPK_SHAPE_SWITCH(n, SPK_CALL_PERSIST)                                                   \
  }                                                                                         \
  hipError_t persist_grid_##n(int tile, int ept, int* out) {                                \
    constexpr int SPK_N = n;                                                                \
    SPK_SHAPE_SWITCH(n, SPK_CALL_PGRID)                                                     \
  }                                                                                         \
  hipError_t rowconv_##n(int tile, int ept, const c32* in, c32* out, const c32* ipsf,       \
                         const c32* tw, const int* fmap, int frames, hipStream_t st) {      \
    constexpr int SPK_N = n;                                                                \
    SPK_SHAPE_SWITCH(n, SPK_CALL_CONV)                                                      \
  }                                                                                         \
  hipError_t resident_##n(int tile, int ept, int xm, const PassArgs& a, const RStep* steps, \
                          int nsteps, double* errp0, long long errp_step, unsigned* bar,    \
                          int* flag, hipStream_t st) {                                      \
    return resident_dispatch<n>(tile, ept, xm, a, steps, nsteps, errp0, errp_step, bar,     \
                                flag, st);                                                  \
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

thread_local hipEvent_t g_launch_events[2] = {nullptr, nullptr};

bool size_supported(int n, int tile, int ept) {
  if (tile != 2 && tile != 4 && tile != 8) return false;
  if (ept != 8 && ept != 16) return false;
  if (n < 16 || n > 4096 || (n & (n - 1))) return false;
  if (n == 4096 && tile == 8) return false;
  if (ept > n) return false;
  return (n / ept) * tile <= 1024;
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

hipError_t launch_colpass(int n, int tile, int ept, int mode, const PassArgs& a, int frames,
                          int zdim, hipStream_t st) {
  SPK_SIZE_SWITCH(colpass, tile, ept, mode, a, frames, zdim, st)
}

hipError_t launch_rowpass(int n, int tile, int ept, bool use_t, bool emit_b, int sw,
                          bool noout, const PassArgs& a, int frames, hipStream_t st) {
  SPK_SIZE_SWITCH(rowpass, tile, ept, use_t, emit_b, sw, noout, a, frames, st)
}

bool resident_supported(int n, int tile, int ept) {
  return n >= 256 && n <= 1024 && (n & (n - 1)) == 0 && ept == 8 &&
         (tile == 4 || (tile == 2 && n <= 512));
}

hipError_t launch_resident(int n, int tile, int ept, int xm, const PassArgs& a,
                           const RStep* steps, int nsteps, double* errp0, long long errp_step,
                           unsigned* bar, int* flag, hipStream_t st) {
  SPK_SIZE_SWITCH(resident, tile, ept, xm, a, steps, nsteps, errp0, errp_step, bar, flag, st)
}

hipError_t launch_rowfft(int n, int tile, int ept, int mode, const c32* in, c32* out,
                         const float* sup, long long sstride, float g, const c32* tw,
                         const int* fmap, int frames, hipStream_t st) {
  SPK_SIZE_SWITCH(rowfft, tile, ept, mode, in, out, sup, sstride, g, tw, fmap, frames, st)
}

hipError_t launch_persist(int n, int tile, int ept, const PassArgs& a, const unsigned char* stage,
                          const PStep* steps, int nsteps, double* errp0, long long errp_step,
                          unsigned* bar, int* flag, int grid, hipStream_t st) {
  SPK_SIZE_SWITCH(persist, tile, ept, a, stage, steps, nsteps, errp0, errp_step, bar, flag, grid,
                  st)
}

hipError_t persist_grid(int n, int tile, int ept, int* out) {
  SPK_SIZE_SWITCH(persist_grid, tile, ept, out)
}

hipError_t launch_rowconv(int n, int tile, int ept, const c32* in, c32* out, const c32* ipsf,
                          const c32* tw, const int* fmap, int frames, hipStream_t st) {
  SPK_SIZE_SWITCH(rowconv, tile, ept, in, out, ipsf, tw, fmap, frames, st)
}

hipError_t launch_shrinkwrap(int n, int tile, const c32* psi, float* sup, long long sstride,
                             float* tmp, float* blur, float* maxp, const SwParam* tab,
                             const SwParam& p0, const int* fmap, int frame0, int frames,
                             int rmax, hipStream_t st) {
  if (!fmap) {  // a contiguous range of frames starting at frame0
    const size_t fo = (size_t)frame0 * n * n;
    psi += fo;
    sup += (size_t)frame0 * sstride;
    tmp += fo;
    blur += fo;
    maxp += (size_t)frame0 * (n / tile);
  }
  SPK_SIZE_SWITCH(shrinkwrap, tile, psi, sup, sstride, tmp, blur, maxp, tab, p0, fmap, frames,
                  rmax, st)
}

hipError_t launch_pack_c64(int n, int tile, c32* dst, const c32* src, const float* sup,
                           long long sstride, int frames, hipStream_t st) {
  dim3 grid((unsigned)((size_t)n * n / 256), frames);
  if (tile == 8)
    hipLaunchKernelGGL(k_pack_c64<8>, grid, dim3(256), 0, st, dst, src, n, sup, sstride);
  else if (tile == 2)
    hipLaunchKernelGGL(k_pack_c64<2>, grid, dim3(256), 0, st, dst, src, n, sup, sstride);
  else
    hipLaunchKernelGGL(k_pack_c64<4>, grid, dim3(256), 0, st, dst, src, n, sup, sstride);
  return hipGetLastError();
}

hipError_t launch_unpack_c64(int n, int tile, c32* dst, const c32* src, float scale,
                             int frames, hipStream_t st) {
  dim3 grid((unsigned)((size_t)n * n / 256), frames);
  if (tile == 8)
    hipLaunchKernelGGL(k_unpack_c64<8>, grid, dim3(256), 0, st, dst, src, n, scale);
  else if (tile == 2)
    hipLaunchKernelGGL(k_unpack_c64<2>, grid, dim3(256), 0, st, dst, src, n, scale);
  else
    hipLaunchKernelGGL(k_unpack_c64<4>, grid, dim3(256), 0, st, dst, src, n, scale);
  return hipGetLastError();
}

hipError_t launch_pack_f32(int n, int tile, float* dst, const float* src, int frames,
                           hipStream_t st) {
  dim3 grid((unsigned)((size_t)n * n / 256), frames);
  if (tile == 8)
    hipLaunchKernelGGL(k_pack_f32<8>, grid, dim3(256), 0, st, dst, src, n);
  else if (tile == 2)
    hipLaunchKernelGGL(k_pack_f32<2>, grid, dim3(256), 0, st, dst, src, n);
  else
    hipLaunchKernelGGL(k_pack_f32<4>, grid, dim3(256), 0, st, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_pack_mod(int n, int tile, float* dst, const float* src, int frames,
                          hipStream_t st) {
  dim3 grid((unsigned)((size_t)n * n / 256), frames);
  if (tile == 8)
    hipLaunchKernelGGL(k_pack_mod<8>, grid, dim3(256), 0, st, dst, src, n);
  else if (tile == 2)
    hipLaunchKernelGGL(k_pack_mod<2>, grid, dim3(256), 0, st, dst, src, n);
  else
    hipLaunchKernelGGL(k_pack_mod<4>, grid, dim3(256), 0, st, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_unpack_f32(int n, int tile, float* dst, const float* src, int frames,
                             hipStream_t st) {
  dim3 grid((unsigned)((size_t)n * n / 256), frames);
  if (tile == 8)
    hipLaunchKernelGGL(k_unpack_f32<8>, grid, dim3(256), 0, st, dst, src, n);
  else if (tile == 2)
    hipLaunchKernelGGL(k_unpack_f32<2>, grid, dim3(256), 0, st, dst, src, n);
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
  const int total = nsteps * frames;  // one wavefront each, four per workgroup
  hipLaunchKernelGGL(k_err_finalize, dim3((total + 3) / 4), dim3(256), 0, st, errp,
                     nsteps, frames, bands, summ2, trace, active);
  return hipGetLastError();
}

hipError_t launch_copy_roi(int n, int tile, const c32* psi, c32* dst, int r0, int c0,
                           int rows, int cols, hipStream_t st) {
  const int total = rows * cols;
  dim3 grid((total + 255) / 256);
  if (tile == 8)
    hipLaunchKernelGGL(k_copy_roi<8>, grid, dim3(256), 0, st, psi, dst, n, r0, c0, rows, cols);
  else if (tile == 2)
    hipLaunchKernelGGL(k_copy_roi<2>, grid, dim3(256), 0, st, psi, dst, n, r0, c0, rows, cols);
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

hipError_t launch_stack_mean(const c32* data, long long per_frame, int frames, c32* out,
                             hipStream_t st) {
  const long long blocks = std::min<long long>(4096, (per_frame + 255) / 256);
  hipLaunchKernelGGL(k_stack_mean, dim3((unsigned)blocks), dim3(256), 0, st, data, per_frame,
                     frames, out);
  return hipGetLastError();
}

static unsigned grid_for(long long n) {
  return (unsigned)std::min<long long>(8192, std::max<long long>(1, (n + 255) / 256));
}

hipError_t launch_conv_axis(const float* in, float* out, int rows, int cols, int axis,
                            const int* off, const float* w, int taps, hipStream_t st) {
  hipLaunchKernelGGL(k_conv_axis, dim3(grid_for((long long)rows * cols)), dim3(256), 0, st, in,
                     out, rows, cols, axis, off, w, taps);
  return hipGetLastError();
}

hipError_t launch_abs_c64(const c32* in, float* out, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_abs_c64, dim3(grid_for(n)), dim3(256), 0, st, in, out, n);
  return hipGetLastError();
}

hipError_t launch_embed(const c32* in, c32* out, int r, int c, int n, int frames,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_embed, dim3(grid_for((long long)frames * n * n)), dim3(256), 0, st, in,
                     out, r, c, n, frames);
  return hipGetLastError();
}

hipError_t launch_prtf(const c32* F, int frames, int n, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_prtf, dim3(grid_for((long long)n * n)), dim3(256), 0, st, F, frames, n,
                     out);
  return hipGetLastError();
}

hipError_t launch_rftf(const c32* F, const float* goal, int n, double* partial, int nblk,
                       int scale, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_rftf_diff, dim3(nblk), dim3(256), 0, st, F, goal, n, partial);
  hipLaunchKernelGGL(k_rftf_out, dim3(grid_for((long long)n * n)), dim3(256), 0, st, F, goal, n,
                     partial, nblk, scale, out);
  return hipGetLastError();
}

hipError_t launch_rescale(int n, int tile, const float* imap, float* dmap, const float* spec,
                          int K, const int* fmap, int frames, hipStream_t st) {
  const long long nn = (long long)n * n;
  hipLaunchKernelGGL(k_rescale, dim3((unsigned)((nn + 255) / 256), frames), dim3(256), 0, st,
                     imap, dmap, spec, K, n, tile, fmap);
  return hipGetLastError();
}

hipError_t launch_rftf_stack(const c32* F, const float* goal, int frames, int n, float* out,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_rftf_stack, dim3(grid_for((long long)n * n)), dim3(256), 0, st, F, goal,
                     frames, n, out);
  return hipGetLastError();
}

hipError_t launch_coh(int op, int n, int tile, c32* a, const c32* b, c32* out, const float* d,
                      float* fr, double* partial, int nblk, float sc, double s0, double s1,
                      hipStream_t st) {
  const long long nn = (long long)n * n;
  const dim3 g(grid_for(nn)), b256(256);
  switch (op) {
    case 0: hipLaunchKernelGGL(k_cmul, g, b256, 0, st, a, b, out, nn); break;
    case 1: hipLaunchKernelGGL(k_rl_divide, g, b256, 0, st, d, b, out, nn, sc); break;
    case 2: hipLaunchKernelGGL(k_rl_update, g, b256, 0, st, a, b, nn, sc); break;
    case 3: hipLaunchKernelGGL(k_abs2_sum, dim3(nblk), b256, 0, st, a, nn, partial); break;
    case 4: hipLaunchKernelGGL(k_div_by_partials, g, b256, 0, st, a, fr, nn, partial, nblk); break;
    case 5: hipLaunchKernelGGL(k_sq_sum, dim3(nblk), b256, 0, st, fr, nn, partial); break;
    case 6: hipLaunchKernelGGL(k_flip, g, b256, 0, st, b, out, n, tile); break;
    case 7: hipLaunchKernelGGL(k_gauss_mul, g, b256, 0, st, b, out, n, tile, s0, s1); break;
    case 8: hipLaunchKernelGGL(k_gauss_resid, dim3(nblk), b256, 0, st, b, d, nn, partial); break;
    case 9: hipLaunchKernelGGL(k_rl_finish, dim3(1), b256, 0, st, a, nn, (int)s0, partial, 0);
            hipLaunchKernelGGL(k_rl_finish, g, b256, 0, st, a, nn, (int)s0, partial, 1); break;
    case 10: hipLaunchKernelGGL(k_scale_c, g, b256, 0, st, a, nn, sc); break;
    case 11: hipLaunchKernelGGL(k_abs_mode, g, b256, 0, st, a, nn, (int)s0); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_com_partial(const c32* data, int frames, int rows, int cols, double* partial,
                              int nblk, hipStream_t st) {
  hipLaunchKernelGGL(k_com_partial, dim3(nblk, frames), dim3(256), 0, st, data, rows, cols,
                     partial, nblk);
  return hipGetLastError();
}

hipError_t launch_roll_sum(const c32* data, int frames, int rows, int cols, const int* shift,
                           c32* out, hipStream_t st) {
  hipLaunchKernelGGL(k_roll_sum, dim3(grid_for((long long)rows * cols)), dim3(256), 0, st, data,
                     frames, rows, cols, shift, out);
  return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
  const size_t n16 = bytes / 16;
  if (n16 == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)((n16 + 1023) / 1024)), dim3(256), 0, st,
                     (uint4*)dst, (const uint4*)src, n16);
  return hipGetLastError();
}

#endif  // SPK_BUILD_COMMON

}  // namespace spk

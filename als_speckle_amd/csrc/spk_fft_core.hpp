// Radix butterflies and the Stockham autosort schedule used by every FFT in
// the library.  Header-only, no HIP dependency: the same functions compile
// for gfx950 (SPK_HD = __host__ __device__) and for the host with g++, which
// lets the CPU test suite run the exact index arithmetic of the kernels.
//
// Conventions (match numpy / pyfft as called at phasing.py:1074-1083):
//   forward  X[k] = sum_n x[n] exp(-2 pi i n k / N)   (unnormalised)
//   inverse  x[n] = sum_k X[k] exp(+2 pi i n k / N)   (the 1/N^2 of ifft2 is
//            applied once by the caller)
//
// Stockham schedule (Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors", SC08): for stages s with radix R_s and
// span Ns = R_0 * ... * R_{s-1}, butterfly j in [0, N/R_s) reads
// in[j + r*N/R_s] (r < R_s), multiplies by W_{Ns*R_s}^{r*(j mod Ns)}, takes a
// length-R_s DFT and writes out[(j/Ns)*Ns*R_s + (j mod Ns) + r*Ns].
//
// Register distribution: a transform of length N = T*E is spread over T
// threads holding E values each, thread t owning x[i] = data[t + i*T].  At a
// stage of radix R (R | E) thread t runs the E/R butterflies j = t + q*T whose
// inputs are exactly its own registers x[q + r*E/R]; only the outputs move
// (through LDS), and after the last stage they land back in the thread's own
// registers, so a transform costs (stages - 1) exchanges and its first load
// and last store are both coalesced along t.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define SPK_HD __host__ __device__ __forceinline__
#else
#define SPK_HD inline
#endif

namespace spk {

struct alignas(8) c32 {
  float x, y;
};

// Fused multiply-adds are written out (fmaf) and the library is compiled
// without implicit contraction, so every value has one fixed sequence of
// roundings wherever a function is inlined: kernels that share these
// functions (the per-pass kernels and the persistent kernel) agree bitwise.
SPK_HD c32 mk(float x, float y) { return c32{x, y}; }
SPK_HD c32 add(c32 a, c32 b) { return c32{a.x + b.x, a.y + b.y}; }
SPK_HD c32 sub(c32 a, c32 b) { return c32{a.x - b.x, a.y - b.y}; }
SPK_HD c32 scale(c32 a, float s) { return c32{a.x * s, a.y * s}; }
SPK_HD c32 mul(c32 a, c32 b) {
  return c32{fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x)};
}
// a*s + b*t for complex a, b and real s, t
SPK_HD c32 axpby(c32 a, float s, c32 b, float t) {
  return c32{fmaf(a.x, s, b.x * t), fmaf(a.y, s, b.y * t)};
}
// acc + a*s
SPK_HD c32 axpy(c32 acc, c32 a, float s) { return c32{fmaf(a.x, s, acc.x), fmaf(a.y, s, acc.y)}; }
SPK_HD c32 conj(c32 a) { return c32{a.x, -a.y}; }
// a * w for forward, a * conj(w) for inverse (table holds forward twiddles)
template <bool INV>
SPK_HD c32 twiddle(c32 a, c32 w) {
  return INV ? c32{fmaf(a.x, w.x, a.y * w.y), fmaf(a.y, w.x, -(a.x * w.y))} : mul(a, w);
}
// a * (-i) forward, a * (+i) inverse
template <bool INV>
SPK_HD c32 rot(c32 a) {
  return INV ? c32{-a.y, a.x} : c32{a.y, -a.x};
}

constexpr float kSqrtHalf = 0.70710678118654752440f;
constexpr float kCos8 = 0.92387953251128675613f;  // cos(pi/8)
constexpr float kSin8 = 0.38268343236508977173f;  // sin(pi/8)

// length-2 DFT on v[0], v[s]
template <bool INV>
SPK_HD void dft2(c32* v, int s = 1) {
  c32 a = v[0], b = v[s];
  v[0] = add(a, b);
  v[s] = sub(a, b);
}

// length-4 DFT on v[0], v[s], v[2s], v[3s] (in order, natural output order)
template <bool INV>
SPK_HD void dft4(c32* v, int s = 1) {
  c32 a0 = v[0], a1 = v[s], a2 = v[2 * s], a3 = v[3 * s];
  c32 e0 = add(a0, a2), e1 = sub(a0, a2);
  c32 o0 = add(a1, a3), o1 = rot<INV>(sub(a1, a3));
  v[0] = add(e0, o0);
  v[s] = add(e1, o1);
  v[2 * s] = sub(e0, o0);
  v[3 * s] = sub(e1, o1);
}

// W8^k for k = 1, 3 (k = 2 is -i); INV conjugates
template <bool INV>
SPK_HD c32 w8_1(c32 a) {  // a * exp(-+ i pi/4)
  return INV ? c32{(a.x - a.y) * kSqrtHalf, (a.x + a.y) * kSqrtHalf}
             : c32{(a.x + a.y) * kSqrtHalf, (a.y - a.x) * kSqrtHalf};
}
template <bool INV>
SPK_HD c32 w8_3(c32 a) {  // a * exp(-+ 3 i pi/4)
  return INV ? c32{(-a.x - a.y) * kSqrtHalf, (a.x - a.y) * kSqrtHalf}
             : c32{(a.y - a.x) * kSqrtHalf, (-a.x - a.y) * kSqrtHalf};
}

// length-8 DFT, natural order in and out: 2 x 4 decomposition
// n = 2 n1 + n2, k = k1 + 4 k2
template <bool INV>
SPK_HD void dft8(c32* v) {
  c32 y0[4] = {v[0], v[2], v[4], v[6]};
  c32 y1[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(y0);
  dft4<INV>(y1);
  y1[1] = w8_1<INV>(y1[1]);
  y1[2] = rot<INV>(y1[2]);
  y1[3] = w8_3<INV>(y1[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = add(y0[k], y1[k]);
    v[k + 4] = sub(y0[k], y1[k]);
  }
}

// length-16 DFT, natural order in and out: 4 x 4 decomposition
// n = 4 n1 + n2, k = k1 + 4 k2; inner DFT over n1, twiddle W16^(n2 k1),
// outer DFT over n2
template <bool INV>
SPK_HD void dft16(c32* v) {
  c32 y[16];  // y[n2*4 + k1]
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    c32 t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
    dft4<INV>(t);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) y[n2 * 4 + k1] = t[k1];
  }
  // W16^1 = (c, -s), W16^2 = W8^1, W16^3 = (s, -c), W16^4 = -i,
  // W16^6 = W8^3, W16^9 = -W16^1
  const c32 w1 = INV ? c32{kCos8, kSin8} : c32{kCos8, -kSin8};
  const c32 w3 = INV ? c32{kSin8, kCos8} : c32{kSin8, -kCos8};
  // n2 = 1: k1 = 1, 2, 3 -> W^1, W^2, W^3
  y[5] = mul(y[5], w1);
  y[6] = w8_1<INV>(y[6]);
  y[7] = mul(y[7], w3);
  // n2 = 2: W^2, W^4, W^6
  y[9] = w8_1<INV>(y[9]);
  y[10] = rot<INV>(y[10]);
  y[11] = w8_3<INV>(y[11]);
  // n2 = 3: W^3, W^6, W^9
  y[13] = mul(y[13], w3);
  y[14] = w8_3<INV>(y[14]);
  y[15] = scale(mul(y[15], w1), -1.0f);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    c32 t[4] = {y[k1], y[4 + k1], y[8 + k1], y[12 + k1]};
    dft4<INV>(t);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
  }
}

template <int R, bool INV>
SPK_HD void dft(c32* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    dft2<INV>(v);
  } else if constexpr (R == 4) {
    dft4<INV>(v);
  } else if constexpr (R == 8) {
    dft8<INV>(v);
  } else if constexpr (R == 16) {
    dft16<INV>(v);
  } else {
    static_assert(R == 16, "unsupported radix");
  }
}

// ---------------------------------------------------------------------------
// Radix plan: radix of stage s for a length-N transform with E values per
// thread.  As many radix-E' (E' = min(E, 16)) stages as fit, then one
// smaller power-of-two stage.  All radices divide E.
// ---------------------------------------------------------------------------
constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

template <int N, int E>
struct RadixPlan {
  static constexpr int kMaxR = E < 16 ? E : 16;
  static constexpr int kLogN = ilog2(N);
  static constexpr int kLogR = ilog2(kMaxR);
  static constexpr int kFull = kLogN / kLogR;           // full-radix stages
  static constexpr int kRest = kLogN - kFull * kLogR;   // log2 of last radix
  static constexpr int kStages = kFull + (kRest ? 1 : 0);
  static constexpr int radix(int s) { return s < kFull ? kMaxR : (1 << kRest); }
  static constexpr int span(int s) {  // Ns before stage s
    return s == 0 ? 1 : span(s - 1) * radix(s - 1);
  }
  static_assert((1 << kLogN) == N, "N must be a power of two");
  static_assert(N % E == 0, "E must divide N");
};

// The same plan at run time (host code that lays out twiddle tables).
SPK_HD constexpr int plan_radix(int n, int e, int s) {
  const int maxr = e < 16 ? e : 16;
  const int full = ilog2(n) / ilog2(maxr);
  return s < full ? maxr : (1 << (ilog2(n) - full * ilog2(maxr)));
}
SPK_HD constexpr int plan_stages(int n, int e) {
  const int maxr = e < 16 ? e : 16;
  const int full = ilog2(n) / ilog2(maxr);
  return full + (ilog2(n) - full * ilog2(maxr) ? 1 : 0);
}

// ---------------------------------------------------------------------------
// Per-stage twiddle table.  Stage s >= 1 (span NS, radix R) multiplies value
// r of butterfly j by W_N^(r k N/(NS R)), k = j mod NS.  Its block holds
// these as [k][r - 1] (k-major, R - 1 entries per k), blocks in stage order
// from stage 1; the table has N - R_0 entries.  Lanes with consecutive k read
// entries R - 1 apart: an odd number of 8-byte entries, so a wavefront's
// reads spread over the LDS banks (in the exponent-indexed table the lanes of
// one stage read entries a power of two apart, which share banks).
// ---------------------------------------------------------------------------
SPK_HD constexpr int staged_tw_offset(int n, int e, int s) {
  int off = 0, ns = 1;
  for (int i = 0; i < s; ++i) {
    const int r = plan_radix(n, e, i);
    if (i >= 1) off += ns * (r - 1);
    ns *= r;
  }
  return off;
}
// exponent m of entry (stage s, k, r): the table entry is W_N^m
SPK_HD constexpr int staged_tw_exponent(int n, int e, int s, int k, int r) {
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= plan_radix(n, e, i);
  return r * k * (n / (ns * plan_radix(n, e, s)));
}

struct TwStaged {  // the per-stage table (LDS)
  const c32* p;
};

// Output position of value r of butterfly j at a stage with span ns, radix R.
SPK_HD int stockham_out(int j, int r, int ns, int R) {
  return (j / ns) * ns * R + (j % ns) + r * ns;
}

// values per thread of one length-N transform (a workgroup holds TT
// transforms of N/ept threads each)
template <int N, int TT>
constexpr int ept() {
  return N < 16 ? N : 16;
}

// ---------------------------------------------------------------------------
// Internal frame layout: TT x TT tiles in row-major tile order, row-major
// inside a tile.  A band of TT rows is one contiguous block; a band of TT
// columns is n/TT contiguous TT*TT blocks.
// ---------------------------------------------------------------------------
template <int TT>
SPK_HD int tpos(int n, int r, int c) {
  return ((r / TT) * (n / TT) + (c / TT)) * (TT * TT) + (r % TT) * TT + (c % TT);
}
template <int TT>
SPK_HD void tile_rc(int n, int p, int& r, int& c) {
  const int tile = p / (TT * TT), w = p % (TT * TT);
  const int tr = tile / (n / TT), tc = tile % (n / TT);
  r = tr * TT + w / TT;
  c = tc * TT + w % TT;
}
// row pass: element i of thread t on row b of row-band `band` (column t + i*T)
template <int N, int TT, int E = ept<N, TT>()>
SPK_HD int row_pos(int band, int b, int t, int i) {
  constexpr int T = N / E;
  const int col = t + i * T;
  return band * (N / TT) * (TT * TT) + (col / TT) * (TT * TT) + b * TT + (col % TT);
}
// column pass: element i of thread t on column b of column-band `band` (row t + i*T)
template <int N, int TT, int E = ept<N, TT>()>
SPK_HD int col_pos(int band, int b, int t, int i) {
  constexpr int T = N / E;
  const int row = t + i * T;
  return ((row / TT) * (N / TT) + band) * (TT * TT) + (row % TT) * TT + b;
}

// Column-band-major layout: a band of TT columns is n*TT contiguous values,
// row-major inside the band.  The modulus is always stored this way (only
// the column pass reads it).  For Phi (the row transform) two schemes exist:
//  - single layout (n <= 2048): Phi stays tiled; the row pass reads and
//    writes whole contiguous row bands, the column pass reads and writes
//    TT x TT tiles in place;
//  - dual layout (n = 4096): the row pass writes Phi column-band-major and
//    the column pass writes it back tiled, so every pass *reads* contiguous
//    memory (at 4096 the tile stride of the single layout is 128 KB).
// Measured on MI355X: strided writes cost more than strided reads up to
// 2048, so the dual layout only pays at 4096.
#ifndef SPK_DUAL_MIN
#define SPK_DUAL_MIN 4096
#endif
constexpr int kDualMin = SPK_DUAL_MIN;
template <int N>
constexpr bool dual_layout() {
  return N >= kDualMin;
}
template <int TT>
SPK_HD int mcb_pos(int n, int r, int c) {
  return (c / TT) * n * TT + r * TT + (c % TT);
}
// column pass: element i of thread t on column b of column-band `band`
template <int N, int TT, int E = ept<N, TT>()>
SPK_HD int mcol_pos(int band, int b, int t, int i) {
  constexpr int T = N / E;
  return band * N * TT + (t + i * T) * TT + b;
}
// row pass: element i of thread t on row b of row-band `band`, in the
// column-band-major layout (where the row pass writes Phi)
template <int N, int TT, int E = ept<N, TT>()>
SPK_HD int cbrow_pos(int band, int b, int t, int i) {
  constexpr int T = N / E;
  const int col = t + i * T;
  return (col / TT) * N * TT + (band * TT + b) * TT + (col % TT);
}

// LDS slot of element p of transform b (TT transforms per workgroup,
// interleaved, one pad slot per lds_pad<N>() slots to spread 8-byte accesses
// over banks).  Measured on MI355X: one per 32 is 1-3 % faster per pass from
// 1024 up, one per 16 is 2 % faster at 512^2 x 64 frames.
template <int N>
constexpr int lds_pad() {
  return N >= 1024 ? 32 : 16;
}
template <int N, int TT>
SPK_HD int lds_idx(int p, int b) {
  const int i = p * TT + b;
  return i + i / lds_pad<N>();
}
template <int N, int TT>
constexpr int lds_elems() {
  return N * TT + (N * TT) / lds_pad<N>() + 1;
}

// ---------------------------------------------------------------------------
// One Stockham stage, per thread.  stage_butterflies runs the thread's E/R
// butterflies of stage S on its registers; except on the last stage the
// results go to LDS and stage_gather (after a workgroup barrier) reloads the
// thread's registers for the next stage.  Kernels put barriers between the
// two; the CPU emulation runs every thread's half-step in turn.
// ---------------------------------------------------------------------------
// Twiddle accessors: tw(m) = W_N^m = exp(-2 pi i m / N), 0 <= m < N.
struct TwTable {  // full table (global memory or LDS)
  const c32* p;
  SPK_HD c32 operator()(int m) const { return p[m]; }
};
template <int N>
struct TwTwoLevel {  // W^m = hi[m >> 6] * lo[m & 63]; hi: N/64 entries, lo: 64
  const c32* p;      // hi followed by lo
  SPK_HD c32 operator()(int m) const { return mul(p[m >> 6], p[(N >> 6) + (m & 63)]); }
};

// Twiddle r of butterfly index k at stage S (exponent m = r k N/(NS R)).
template <int N, int E, int S, class TW>
SPK_HD c32 tw_get(const TW& tw, int k, int r, int m) {
  return tw(m);
}
template <int N, int E, int S>
SPK_HD c32 tw_get(const TwStaged& tw, int k, int r, int m) {
  constexpr int OFF = staged_tw_offset(N, E, S);
  constexpr int R = plan_radix(N, E, S);
  return tw.p[OFF + k * (R - 1) + (r - 1)];
}

template <int N, int E, int TT, bool INV, int S, class TW>
SPK_HD void stage_butterflies(c32 (&x)[E], int t, int b, c32* lds, const TW& tw) {
  using P = RadixPlan<N, E>;
  constexpr int T = N / E;
  constexpr int R = P::radix(S);
  constexpr int NS = P::span(S);
  constexpr int Q = E / R;
  constexpr bool LAST = (S == P::kStages - 1);
#if defined(__HIPCC__)
#pragma unroll
#endif
  for (int q = 0; q < Q; ++q) {
    c32 v[R];
    for (int r = 0; r < R; ++r) v[r] = x[q + r * Q];
    const int j = t + q * T;
    if constexpr (NS > 1) {
      const int k = j & (NS - 1);
      constexpr int STRIDE = N / (NS * R);
      for (int r = 1; r < R; ++r) v[r] = twiddle<INV>(v[r], tw_get<N, E, S>(tw, k, r, (r * k) * STRIDE));
    }
    dft<R, INV>(v);
    if constexpr (LAST) {
      // output j + r*N/R = t + (q + r*Q)*T: back into the same registers
      for (int r = 0; r < R; ++r) x[q + r * Q] = v[r];
    } else {
      const int base = (j / NS) * NS * R + (j & (NS - 1));
      // slots of positions base + r*NS.  The pad term of lds_idx is linear in
      // r when the steps are whole pad groups, so the R slots are one
      // base plus compile-time offsets (immediate LDS offsets, no per-value
      // index arithmetic).
      constexpr int PG = lds_pad<N>();
      if constexpr ((NS * TT) % PG == 0) {
        const int i0 = lds_idx<N, TT>(base, b);
        for (int r = 0; r < R; ++r) lds[i0 + r * (NS * TT + NS * TT / PG)] = v[r];
      } else if constexpr (NS == 1 && (R * TT) % PG == 0) {
        // base = j*R starts a pad group; offsets r*TT stay inside the groups
        // they start in (TT divides the group, b < TT)
        const int i0 = j * (R * TT + R * TT / PG) + b;
        for (int r = 0; r < R; ++r) lds[i0 + r * TT + (r * TT) / PG] = v[r];
      } else {
        for (int r = 0; r < R; ++r) lds[lds_idx<N, TT>(base + r * NS, b)] = v[r];
      }
    }
  }
}

template <int N, int E, int TT>
SPK_HD void stage_gather(c32 (&x)[E], int t, int b, const c32* lds) {
  constexpr int T = N / E;
  constexpr int PG = lds_pad<N>();
  if constexpr ((T * TT) % PG == 0) {  // linear slots, as in stage_butterflies
    const int i0 = lds_idx<N, TT>(t, b);
    for (int i = 0; i < E; ++i) x[i] = lds[i0 + i * (T * TT + T * TT / PG)];
  } else {
    for (int i = 0; i < E; ++i) x[i] = lds[lds_idx<N, TT>(t + i * T, b)];
  }
}

}  // namespace spk

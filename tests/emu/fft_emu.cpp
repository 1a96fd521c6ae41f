// CPU emulation of the library's FFT kernels, built with g++ by the test
// suite.  It runs the exact per-thread code of the kernels
// (stage_butterflies / stage_gather / row_pos / col_pos / tpos / tile_rc from
// spk_fft_core.hpp), with every "thread" of a workgroup executed in turn
// where the kernels place a barrier.  It checks the Stockham schedule, the
// radix butterflies and the tiled-layout addressing against numpy without a
// GPU.
#include <array>
#include <cstring>
#include <vector>
#include <cmath>

#include "spk_fft_core.hpp"

using namespace spk;

namespace {

std::vector<c32> twiddles(int n) {
  std::vector<c32> tw(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * M_PI * k / n;
    tw[k] = c32{(float)std::cos(a), (float)std::sin(a)};
  }
  return tw;
}

// all threads (t, b) of one workgroup: x[(t*TT + b)][E]
template <int N, int TT, int E, bool INV, int S, class TW>
void emu_stages(std::vector<std::array<c32, E>>& x, c32* lds, const TW& tw) {
  constexpr int T = N / E;
  using P = RadixPlan<N, E>;
  for (int tid = 0; tid < T * TT; ++tid) {
    c32(&xr)[E] = *reinterpret_cast<c32(*)[E]>(x[tid].data());
    stage_butterflies<N, E, TT, INV, S>(xr, tid / TT, tid % TT, lds, tw);
  }
  if constexpr (S != P::kStages - 1) {
    for (int tid = 0; tid < T * TT; ++tid) {
      c32(&xr)[E] = *reinterpret_cast<c32(*)[E]>(x[tid].data());
      stage_gather<N, E, TT>(xr, tid / TT, tid % TT, lds);
    }
    emu_stages<N, TT, E, INV, S + 1, TW>(x, lds, tw);
  }
}

template <int N, int TT, int E, class TW>
int fft2_tiled_tw(int inverse, const c32* in, c32* out, const TW& tw) {
  constexpr int T = N / E;
  std::vector<c32> tiled((size_t)N * N), lds(lds_elems<N, TT>());
  for (int p = 0; p < N * N; ++p) {
    int r, c;
    tile_rc<TT>(N, p, r, c);
    tiled[p] = in[(size_t)r * N + c];
  }
  std::vector<std::array<c32, E>> x(T * TT);
  std::vector<c32> colband((size_t)N * N);
  // row transforms, one band of TT rows per workgroup: read the tiled layout,
  // write tiled (single layout) or column-band-major (dual layout), as
  // k_rowpass / k_rowfft
  for (int band = 0; band < N / TT; ++band) {
    for (int tid = 0; tid < T * TT; ++tid)
      for (int i = 0; i < E; ++i) x[tid][i] = tiled[row_pos<N, TT, E>(band, tid % TT, tid / TT, i)];
    if (inverse) emu_stages<N, TT, E, true, 0>(x, lds.data(), tw);
    else emu_stages<N, TT, E, false, 0>(x, lds.data(), tw);
    for (int tid = 0; tid < T * TT; ++tid)
      for (int i = 0; i < E; ++i) {
        if constexpr (dual_layout<N>())
          colband[cbrow_pos<N, TT, E>(band, tid % TT, tid / TT, i)] = x[tid][i];
        else
          tiled[row_pos<N, TT, E>(band, tid % TT, tid / TT, i)] = x[tid][i];
      }
  }
  // column transforms, one band of TT columns per workgroup: read tiled or
  // column-band-major, write the tiled layout (as k_colpass)
  for (int band = 0; band < N / TT; ++band) {
    for (int tid = 0; tid < T * TT; ++tid)
      for (int i = 0; i < E; ++i)
        x[tid][i] = dual_layout<N>() ? colband[mcol_pos<N, TT, E>(band, tid % TT, tid / TT, i)]
                                     : tiled[col_pos<N, TT, E>(band, tid % TT, tid / TT, i)];
    if (inverse) emu_stages<N, TT, E, true, 0>(x, lds.data(), tw);
    else emu_stages<N, TT, E, false, 0>(x, lds.data(), tw);
    for (int tid = 0; tid < T * TT; ++tid)
      for (int i = 0; i < E; ++i) tiled[col_pos<N, TT, E>(band, tid % TT, tid / TT, i)] = x[tid][i];
  }
  for (int r = 0; r < N; ++r)
    for (int c = 0; c < N; ++c) out[(size_t)r * N + c] = tiled[tpos<TT>(N, r, c)];
  return 0;
}

// twiddles as the kernels use them: tw_mode 2, the per-stage table of the
// radix schedule (where it fits in LDS), 1, the two-level table (hi =
// W^(64 i), lo = W^j, where LDS is short), or 0, the exponent-indexed table
template <int N, int TT, int E>
int fft2_tiled(int inverse, const c32* in, c32* out, int tw_mode) {
  const std::vector<c32> tw = twiddles(N);
  if (tw_mode == 0) return fft2_tiled_tw<N, TT, E>(inverse, in, out, TwTable{tw.data()});
  if (tw_mode == 2) {
    std::vector<c32> st((size_t)N, c32{0.0f, 0.0f});
    for (int s = 1; s < plan_stages(N, E); ++s) {
      const int R = plan_radix(N, E, s), off = staged_tw_offset(N, E, s);
      int ns = 1;
      for (int i = 0; i < s; ++i) ns *= plan_radix(N, E, i);
      for (int k = 0; k < ns; ++k)
        for (int r = 1; r < R; ++r)
          st[off + k * (R - 1) + (r - 1)] = tw[staged_tw_exponent(N, E, s, k, r)];
    }
    return fft2_tiled_tw<N, TT, E>(inverse, in, out, TwStaged{st.data()});
  }
  if constexpr (N >= 64) {
    std::vector<c32> hl((N >> 6) + 64);
    for (int i = 0; i < (N >> 6); ++i) hl[i] = tw[i * 64];
    for (int j = 0; j < 64; ++j) hl[(N >> 6) + j] = tw[j];
    return fft2_tiled_tw<N, TT, E>(inverse, in, out, TwTwoLevel<N>{hl.data()});
  }
  return -1;
}

}  // namespace

extern "C" {

// 2-D transform of an n x n complex64 array through the tiled layout
int emu_fft2(int n, int tile, int inverse, const float* in, float* out, int tw_mode,
             int ept) {
  const c32* i = reinterpret_cast<const c32*>(in);
  c32* o = reinterpret_cast<c32*>(out);
#define EMU_CASE(NN)                                            \
  if (n == NN) {                                                \
    if (ept == 16 && tile == 8) return fft2_tiled<NN, 8, 16>(inverse, i, o, tw_mode); \
    if (ept == 16 && tile == 4) return fft2_tiled<NN, 4, 16>(inverse, i, o, tw_mode); \
    if (ept == 8 && tile == 8) return fft2_tiled<NN, 8, 8>(inverse, i, o, tw_mode);   \
    if (ept == 8 && tile == 4) return fft2_tiled<NN, 4, 8>(inverse, i, o, tw_mode);   \
    if (ept == 16 && tile == 2) return fft2_tiled<NN, 2, 16>(inverse, i, o, tw_mode); \
    if (ept == 8 && tile == 2) return fft2_tiled<NN, 2, 8>(inverse, i, o, tw_mode);   \
    return -1;                                                  \
  }
  EMU_CASE(16) EMU_CASE(32) EMU_CASE(64) EMU_CASE(128) EMU_CASE(256) EMU_CASE(512)
  EMU_CASE(1024) EMU_CASE(2048) EMU_CASE(4096)
#undef EMU_CASE
  return -1;
}

// stockham output position, exposed for a property test
int emu_stockham_out(int j, int r, int ns, int R) { return stockham_out(j, r, ns, R); }

}  // extern "C"

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
template <int N, int TT, bool INV, int S>
void emu_stages(std::vector<std::array<c32, ept<N>()>>& x, c32* lds, const c32* tw) {
  constexpr int E = ept<N>();
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
    emu_stages<N, TT, INV, S + 1>(x, lds, tw);
  }
}

template <int N, int TT>
int fft2_tiled(int inverse, const c32* in, c32* out) {
  constexpr int E = ept<N>();
  constexpr int T = N / E;
  const std::vector<c32> tw = twiddles(N);
  std::vector<c32> tiled((size_t)N * N), lds(lds_elems<N, TT>());
  for (int p = 0; p < N * N; ++p) {
    int r, c;
    tile_rc<TT>(N, p, r, c);
    tiled[p] = in[(size_t)r * N + c];
  }
  std::vector<std::array<c32, E>> x(T * TT);
  // row transforms, one band of TT rows per workgroup
  for (int band = 0; band < N / TT; ++band) {
    for (int tid = 0; tid < T * TT; ++tid)
      for (int i = 0; i < E; ++i) x[tid][i] = tiled[row_pos<N, TT>(band, tid % TT, tid / TT, i)];
    if (inverse) emu_stages<N, TT, true, 0>(x, lds.data(), tw.data());
    else emu_stages<N, TT, false, 0>(x, lds.data(), tw.data());
    for (int tid = 0; tid < T * TT; ++tid)
      for (int i = 0; i < E; ++i) tiled[row_pos<N, TT>(band, tid % TT, tid / TT, i)] = x[tid][i];
  }
  // column transforms, one band of TT columns per workgroup
  for (int band = 0; band < N / TT; ++band) {
    for (int tid = 0; tid < T * TT; ++tid)
      for (int i = 0; i < E; ++i) x[tid][i] = tiled[col_pos<N, TT>(band, tid % TT, tid / TT, i)];
    if (inverse) emu_stages<N, TT, true, 0>(x, lds.data(), tw.data());
    else emu_stages<N, TT, false, 0>(x, lds.data(), tw.data());
    for (int tid = 0; tid < T * TT; ++tid)
      for (int i = 0; i < E; ++i) tiled[col_pos<N, TT>(band, tid % TT, tid / TT, i)] = x[tid][i];
  }
  for (int r = 0; r < N; ++r)
    for (int c = 0; c < N; ++c) out[(size_t)r * N + c] = tiled[tpos<TT>(N, r, c)];
  return 0;
}

}  // namespace

extern "C" {

// 2-D transform of an n x n complex64 array through the tiled layout
int emu_fft2(int n, int tile, int inverse, const float* in, float* out) {
  const c32* i = reinterpret_cast<const c32*>(in);
  c32* o = reinterpret_cast<c32*>(out);
#define EMU_CASE(NN)                                            \
  if (n == NN) {                                                \
    if (tile == 8 && NN != 4096) return fft2_tiled<NN, (NN == 4096 ? 4 : 8)>(inverse, i, o); \
    if (tile == 4) return fft2_tiled<NN, 4>(inverse, i, o);      \
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

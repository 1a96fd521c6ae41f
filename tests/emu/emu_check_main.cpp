// Stand-alone driver of the FFT emulator for a sanitizer build (g++
// -fsanitize=address,undefined): runs the kernels' per-thread code over
// every built (n, tile, values per thread, twiddle form) with the LDS
// exchange buffer and twiddle tables in exactly sized heap arrays, so an
// out-of-range LDS slot, twiddle index or layout position is reported.
// Each transform is checked by its inverse: ifft(fft(x)) = n^2 x.
#include <cmath>
#include <cstdio>
#include <vector>

extern "C" int emu_fft2(int n, int tile, int inverse, const float* in, float* out, int tw_mode,
                        int ept);

int main() {
  int bad = 0, runs = 0;
  for (int n = 16; n <= 1024; n *= 2) {
    for (int tile : {2, 4, 8}) {
      for (int ept : {8, 16}) {
        if ((n / ept) * tile > 1024 || ept > n) continue;
        for (int tw : {0, 1, 2}) {
          if (tw == 1 && (n < 64 || (n < 1024 && tile == 8))) continue;
          if (tw == 0 && n > 256) continue;  // the exponent table: small frames
          std::vector<float> x(2 * (size_t)n * n), y(x.size()), z(x.size());
          for (size_t i = 0; i < x.size(); ++i) x[i] = (float)((i * 2654435761u) % 1000) * 1e-3f - 0.5f;
          if (emu_fft2(n, tile, 0, x.data(), y.data(), tw, ept) != 0 ||
              emu_fft2(n, tile, 1, y.data(), z.data(), tw, ept) != 0) {
            std::printf("unsupported n=%d tile=%d ept=%d tw=%d\n", n, tile, ept, tw);
            ++bad;
            continue;
          }
          double num = 0, den = 0;
          for (size_t i = 0; i < x.size(); ++i) {
            const double d = z[i] / ((double)n * n) - x[i];
            num += d * d;
            den += (double)x[i] * x[i];
          }
          ++runs;
          if (!(std::sqrt(num / den) < 1e-5)) {
            std::printf("round trip n=%d tile=%d ept=%d tw=%d rel %.3e\n", n, tile, ept, tw,
                        std::sqrt(num / den));
            ++bad;
          }
        }
      }
    }
  }
  std::printf("%d transforms, %d bad\n", runs, bad);
  return bad ? 1 : 0;
}

// Host-side check of the single-population reconstructions: recon2_pop /
// recon3_pop must round every population exactly as recon2_all / recon3_all
// do (same operations, same order), over random moment sets at three
// scales.  Built with the library's flags (no FMA contraction); x86 SSE
// arithmetic is IEEE binary32/binary64 like the GPU's.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../freeflow_amd/csrc/hermite.h"

using namespace ff;

static bool same_bits(float a, float b) { return std::memcmp(&a, &b, sizeof(float)) == 0; }

int main(int argc, char** argv) {
  const long cases = argc > 1 ? std::atol(argv[1]) : 200000;
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<float> U(-1.0f, 1.0f);
  long bad2 = 0, bad3 = 0;
  for (long it = 0; it < cases; it++) {
    const float sc = it % 3 == 0 ? 1e-3f : (it % 3 == 1 ? 0.1f : 0.5f);
    float m[10];
    m[0] = 1.0f + sc * U(rng);
    for (int q = 1; q < 10; q++) m[q] = sc * U(rng);
    float f2[9];
    recon2_all(m[0], m[1], m[2], m[3], m[5], m[4], f2);
    for (int i = 0; i < 9; i++)
      if (!same_bits(recon2_pop(m[0], m[1], m[2], m[3], m[5], m[4], i), f2[i])) bad2++;
    float f3[27];
    recon3_all(m, f3);
    for (int i = 0; i < 27; i++)
      if (!same_bits(recon3_pop(m, i), f3[i])) bad3++;
  }
  std::printf("cases %ld mismatches 2D %ld 3D %ld\n", cases, bad2, bad3);
  return (bad2 || bad3) ? 1 : 0;
}

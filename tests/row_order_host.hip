// Host build of the workgroup order of the stencil launches (xcd_perm of
// cupymag_amd/csrc/device.h), compiled with hipcc --offload-host-only for
// tests/test_host.py.
//
// usage: row_order_host gmax
// Prints one line "perm <g> <G> <ok> <contiguous>" per grid size g in
// 1..gmax and XCD group G in {-1, 0, 1, 2, 4, 8, 16, 64}: ok = 1 when
// xcd_perm(., g, G) is a bijection of [0, g); contiguous = 1 when every
// XCD (dispatch position w % 8) takes, within each full run of 8 G
// positions, G consecutive workgroups (G > 1), or one consecutive range of
// the whole grid (G = -1).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "device.h"

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  const int gmax = std::atoi(argv[1]);
  for (int g = 1; g <= gmax; ++g)
    for (int G : {-1, 0, 1, 2, 4, 8, 16, 64}) {
      std::vector<int> seen(g, 0);
      bool ok = true, contiguous = true;
      for (int w = 0; w < g; ++w) {
        const int64_t b = cpm::xcd_perm(w, g, G);
        if (b < 0 || b >= g || seen[b]++) ok = false;
        // the next position on the same XCD continues the run
        if (w + 8 < g) {
          const int64_t b2 = cpm::xcd_perm(w + 8, g, G);
          if (G > 1 && (w / (8 * G)) == ((w + 8) / (8 * G)) && (w + 8) < (g / (8 * G)) * (8 * G) && b2 != b + 1)
            contiguous = false;
          if (G < 0 && b2 != b + 1) contiguous = false;
        }
      }
      std::printf("perm %d %d %d %d\n", g, G, ok ? 1 : 0, contiguous ? 1 : 0);
    }
  return 0;
}

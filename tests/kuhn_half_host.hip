// Host build of the half-stored Kuhn stencil forms: the FP16 forms of the
// multigrid levels (make_half_row of cupymag_amd/csrc/mg_dev.h), the FP32
// operator forms (kuhn_slot_symmetric and kuhn_half32 of sell_dev.h) and
// the row products reading them (sell_row of sell_dev.h), compiled for the
// host with hipcc --offload-host-only, for tests/test_multigrid_host.py.
//
// usage: kuhn_half_host nx ny nz bs seed
// Builds a random operator on the periodic nx*ny*nz grid whose off-diagonal
// blocks sit on the 14 Kuhn offsets and are symmetric slot by slot
// (A(i, i-d) = A(i-d, i)^T), stores it in the full SELL-64 stencil layout
// (FP64, and FP16 in the 3x3 word layout) and in the half layout, and prints
//   max |y_half - y_full|  max |y_full|  max |dgh_half - dgh_full|  max |dgh_full|
//   max |y_full - y64|  max |y64|
// and, for the FP32 forms (kuhn_half32 against every slot in FP32),
//   max |y_half - y_full|  max |y_full|  max |dgc_half - dgc_full|  max |dgc_full|  symmetric  perturbed-symmetric
// where y_* = the FP16 forms' products with one random vector (same FP32
// diagonal blocks in both), y64 the FP64 product, and dgh_* the compensated
// diagonal blocks (dg64 + the rounding of the row's fourteen blocks).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "mg_dev.h"

template <int BS>
static int run(int nx, int ny, int nz, unsigned seed) {
  constexpr int BB = BS * BS;
  const int32_t nb = nx * ny * nz;
  const int64_t nsl = (nb + 63) / 64, pad = nsl * 64;
  cpm::GridArg ga{nx, ny, nz, cpm::kKuhnSlots, nullptr, 1};
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<double> v64(nsl * cpm::kKuhnSlots * BB * 64, 0.0), dg64((size_t)nb * BB);
  for (int64_t i = 0; i < nb; ++i) {
    const int64_t s = i >> 6, l = i & 63;
    for (int t = 7; t < cpm::kKuhnSlots; ++t) {
      // forward block of row i at slot t, and its transpose at the mirrored
      // slot 13 - t of the neighbour
      const int64_t c = cpm::kuhn_neighbour(i, ga, t);
      const int64_t cs = (c >> 6) * cpm::kKuhnSlots + (cpm::kKuhnSlots - 1 - t), lc = c & 63;
      for (int d = 0; d < BS; ++d)
        for (int e = 0; e < BS; ++e) {
          const double a = nd(rng) * std::ldexp(1.0, (int)(rng() % 7) - 3);
          v64[((s * cpm::kKuhnSlots + t) * BB + d * BS + e) * 64 + l] = a;
          v64[(cs * BB + e * BS + d) * 64 + lc] = a;
        }
    }
    for (int t = 0; t < BB; ++t) dg64[i * BB + t] = (t % (BS + 1) == 0 ? 40.0 : 0.0) + 0.1 * nd(rng);
  }
  const double vscale = 0.25;
  // full FP16 layout and its compensated diagonal blocks
  std::vector<_Float16> vf(nsl * cpm::kKuhnSlots * (BB == 9 ? 576 : 64));
  std::vector<double> dgf(dg64);
  for (int64_t row = 0; row < pad; ++row) {
    const int64_t s = row >> 6, l = row & 63;
    for (int q = 0; q < cpm::kKuhnSlots; ++q) {
      const int64_t slot = s * cpm::kKuhnSlots + q;
      for (int t = 0; t < BB; ++t) {
        const double a = v64[(slot * BB + t) * 64 + l];
        const _Float16 h = (_Float16)(a / vscale);
        vf[BB == 9 ? slot * 576 + (t < 8 ? 8 * l + t : 512 + l) : slot * 64 + l] = h;
        if (row < nb) dgf[row * BB + t] += a - (double)h * vscale;
      }
    }
  }
  // half layout
  std::vector<_Float16> vh(nsl * 7 * (BB == 9 ? 576 : 64));
  std::vector<double> dgh((size_t)nb * BB);
  for (int64_t row = 0; row < pad; ++row) cpm::make_half_row<BS>(row, nb, ga, v64.data(), dg64.data(), vscale,
                                                                 vh.data(), dgh.data());
  std::vector<float> d32(dgf.begin(), dgf.end());
  std::vector<double> x((size_t)nb * BS);
  for (double& v : x) v = nd(rng);
  double dy = 0, ym = 0, dd = 0, dm = 0, e64 = 0, y64m = 0;
  for (int64_t row = 0; row < nb; ++row) {
    double yf[BS], yh[BS], xo[BS];
    cpm::GridArg gf = ga, gh = ga;
    gh.kuhn = 2;
    cpm::sell_row<BS, 1, double, false, true, _Float16>(row, nb, nullptr, nullptr, vf.data(), d32.data(), x.data(),
                                                        nullptr, gf, yf, xo, vscale);
    cpm::sell_row<BS, 1, double, false, true, _Float16>(row, nb, nullptr, nullptr, vh.data(), d32.data(), x.data(),
                                                        nullptr, gh, yh, xo, vscale);
    // FP64 product
    const int64_t s = row >> 6, l = row & 63;
    double y[BS] = {};
    for (int q = 0; q < cpm::kKuhnSlots; ++q) {
      const int64_t c = cpm::kuhn_neighbour(row, ga, q);
      for (int d = 0; d < BS; ++d)
        for (int e = 0; e < BS; ++e)
          y[d] += v64[((s * cpm::kKuhnSlots + q) * BB + d * BS + e) * 64 + l] * x[c * BS + e];
    }
    for (int d = 0; d < BS; ++d)
      for (int e = 0; e < BS; ++e) y[d] += (double)d32[row * BB + d * BS + e] * x[row * BS + e];
    for (int d = 0; d < BS; ++d) {
      dy = std::fmax(dy, std::fabs(yh[d] - yf[d]));
      ym = std::fmax(ym, std::fabs(yf[d]));
      e64 = std::fmax(e64, std::fabs(yf[d] - y[d]));
      y64m = std::fmax(y64m, std::fabs(y[d]));
    }
    for (int t = 0; t < BB; ++t) {
      dd = std::fmax(dd, std::fabs(dgh[row * BB + t] - dgf[row * BB + t]));
      dm = std::fmax(dm, std::fabs(dgf[row * BB + t]));
    }
  }
  std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", dy, ym, dd, dm, e64, y64m);
  // FP32 forms: every slot (and its compensated diagonal blocks) against
  // the half form of kuhn_half32, on the same operator
  std::vector<float> f32(v64.size());
  std::vector<double> dgc32(dg64);
  for (int64_t row = 0; row < pad; ++row) {
    const int64_t s = row >> 6, l = row & 63;
    for (int q = 0; q < cpm::kKuhnSlots; ++q)
      for (int t = 0; t < BB; ++t) {
        const int64_t i = ((s * cpm::kKuhnSlots + q) * BB + t) * 64 + l;
        f32[i] = (float)v64[i];
        if (row < nb) dgc32[row * BB + t] += v64[i] - (double)f32[i];
      }
  }
  std::vector<float> h32(nsl * 7 * BB * 64);
  std::vector<double> dgh32((size_t)nb * BB);
  const bool sym = cpm::kuhn_slot_symmetric(BS, nb, ga, v64.data());
  cpm::kuhn_half32(BS, nb, nsl, ga, v64.data(), dg64.data(), h32.data(), dgh32.data());
  double dy2 = 0, ym2 = 0, dd2 = 0, dm2 = 0;
  for (int64_t row = 0; row < nb; ++row) {
    double yf[BS], yh[BS], xo[BS];
    cpm::GridArg gf = ga, gh = ga;
    gh.kuhn = 2;
    cpm::sell_row<BS, 1, double, false, true, float>(row, nb, nullptr, nullptr, f32.data(), dgc32.data(), x.data(),
                                                     nullptr, gf, yf, xo);
    cpm::sell_row<BS, 1, double, false, true, float>(row, nb, nullptr, nullptr, h32.data(), dgc32.data(), x.data(),
                                                     nullptr, gh, yh, xo);
    for (int d = 0; d < BS; ++d) {
      dy2 = std::fmax(dy2, std::fabs(yh[d] - yf[d]));
      ym2 = std::fmax(ym2, std::fabs(yf[d]));
    }
    for (int t = 0; t < BB; ++t) {
      dd2 = std::fmax(dd2, std::fabs(dgh32[row * BB + t] - dgc32[row * BB + t]));
      dm2 = std::fmax(dm2, std::fabs(dgc32[row * BB + t]));
    }
  }
  // one backward entry off by 1e-9 of its size: no longer slot-symmetric
  v64[((0 * cpm::kKuhnSlots + 3) * BB + 0) * 64 + 5] *= 1.0 + 1e-9;
  const bool sym2 = cpm::kuhn_slot_symmetric(BS, nb, ga, v64.data());
  std::printf("%.17g %.17g %.17g %.17g %d %d\n", dy2, ym2, dd2, dm2, (int)sym, (int)sym2);
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s nx ny nz bs seed\n", argv[0]);
    return 2;
  }
  const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]), bs = std::atoi(argv[4]);
  const unsigned seed = (unsigned)std::atoi(argv[5]);
  if (nx < 3 || ny < 3 || nz < 3) {
    std::fprintf(stderr, "dims must be >= 3\n");
    return 2;
  }
  return bs == 3 ? run<3>(nx, ny, nz, seed) : run<1>(nx, ny, nz, seed);
}

// Host build of the multigrid transfers' per-node rules (restrict_node and
// prolong_node of cupymag_amd/csrc/mg_dev.h, compiled for the host with
// hipcc --offload-host-only), for tests/test_multigrid_host.py.
//
// usage: transfer_host nfx nfy nfz ncx ncy ncz anchor
// stdin: the fine vector r (nf doubles) then the coarse vector e (nc doubles)
// stdout: R r (nc doubles) then P e (nf doubles), one value per node
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mg_dev.h"

int main(int argc, char** argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s nfx nfy nfz ncx ncy ncz anchor\n", argv[0]);
    return 2;
  }
  cpm::Dims3 fine{{std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3])}};
  cpm::Dims3 coarse{{std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6])}};
  const bool anchor = std::atoi(argv[7]) != 0;
  const int64_t nf = (int64_t)fine.d[0] * fine.d[1] * fine.d[2];
  const int64_t nc = (int64_t)coarse.d[0] * coarse.d[1] * coarse.d[2];
  std::vector<double> r(nf), e(nc), rc(nc), pf(nf);
  if (std::fread(r.data(), sizeof(double), nf, stdin) != (size_t)nf ||
      std::fread(e.data(), sizeof(double), nc, stdin) != (size_t)nc) {
    std::fprintf(stderr, "short input\n");
    return 2;
  }
  const bool nested = cpm::nested_dims(fine, coarse);
  for (int64_t I = 0; I < nc; ++I) {
    double acc[1];
    cpm::restrict_node<1, double>(I, fine, coarse, nested, anchor, r.data(), acc);
    rc[I] = acc[0];
  }
  for (int64_t f = 0; f < nf; ++f) {
    double v[1] = {0.0};
    if (!(anchor && f == 0)) cpm::prolong_node<1, double>(f, fine, coarse, nested, anchor, e.data(), v);
    pf[f] = v[0];
  }
  std::fwrite(rc.data(), sizeof(double), nc, stdout);
  std::fwrite(pf.data(), sizeof(double), nf, stdout);
  return 0;
}

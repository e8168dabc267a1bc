// Standalone driver of the native host assembler (cupymag_amd/csrc/
// host_assembly.cpp) for a build with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_host.py builds it with g++).
//
// stdin: n_nodes, coords (n_nodes x 3), n_elems, npe, elem_stride, elems
// (n_elems x stride, int32 as doubles), ndof, gid (n_nodes, as doubles),
// defect flag (n bytes as doubles), coef (38), flags, nthreads -- all
// float64.  stdout: every array of cpm_host_ops (the 19 slots, each as its
// length followed by its raw bytes), then the dense element matrices of
// cpm_host_element_matrices for the first min(n_elems, 7) elements.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cupymag_hip.h"
#include "errors.h"

// the library keeps messages for cpm_last_error (context.hip, a HIP
// source); this driver prints them
void cpm::set_error(const std::string& msg) { std::fprintf(stderr, "%s\n", msg.c_str()); }

static double rd() {
  double v = 0;
  if (std::fread(&v, sizeof v, 1, stdin) != 1) {
    std::fprintf(stderr, "short input\n");
    std::exit(2);
  }
  return v;
}

int main() {
  const int64_t nn = (int64_t)rd();
  std::vector<double> coords(3 * nn);
  for (auto& v : coords) v = rd();
  const int64_t ne = (int64_t)rd();
  const int npe = (int)rd(), stride = (int)rd();
  std::vector<int32_t> el(ne * stride);
  for (auto& v : el) v = (int32_t)rd();
  const int32_t ndof = (int32_t)rd();
  std::vector<int32_t> gid(nn);
  for (auto& v : gid) v = (int32_t)rd();
  std::vector<uint8_t> dd(ndof);
  for (auto& v : dd) v = (uint8_t)rd();
  std::vector<double> coef(38);
  for (auto& v : coef) v = rd();
  const int flags = (int)rd(), nthreads = (int)rd();
  cpm_host_ops* H = nullptr;
  int rc = cpm_host_assemble(coords.data(), nn, el.data(), ne, npe, stride, gid.data(), ndof, dd.data(), coef.data(),
                             flags, nthreads, &H);
  if (rc) {
    std::fprintf(stderr, "assembly failed: %d\n", rc);
    return 3;
  }
  for (int which = 0; which <= CPM_HA_FEL; ++which) {
    void* p = nullptr;
    int64_t len = 0;
    if (cpm_host_ops_array(H, which, &p, &len) != 0) len = 0;
    const int64_t isz = (which == CPM_HA_PTR || which == CPM_HA_GS_PTR || which == CPM_HA_F1_PTR ||
                         which == CPM_HA_EL_PTR || which == CPM_HA_FE_PTR) ? 8
                        : (which == CPM_HA_IDX || which == CPM_HA_GS_IDX || which == CPM_HA_F1_IDX ||
                           which == CPM_HA_EL_IDX || which == CPM_HA_FE_IDX) ? 4 : 8;
    std::fwrite(&len, sizeof len, 1, stdout);
    if (len) std::fwrite(p, isz, len, stdout);
  }
  cpm_host_ops_free(H);
  const int64_t ke = ne < 7 ? ne : 7;
  std::vector<double> K(ke * npe * npe), M(ke * npe * npe), F(3 * ke * npe * npe), Kel(ke * 9 * npe * npe),
      Fel(ke * 18 * npe * npe);
  rc = cpm_host_element_matrices(coords.data(), nn, el.data(), ke, npe, stride, coef.data() + 2, K.data(), M.data(),
                                 F.data(), Kel.data(), Fel.data());
  if (rc) return 4;
  for (auto* v : {&K, &M, &F, &Kel, &Fel}) std::fwrite(v->data(), sizeof(double), v->size(), stdout);
  return 0;
}

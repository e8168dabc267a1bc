// Context (device + stream + reduction scratch), error strings and the CSR
// operator object of libcupymag_hip.so.
#include <algorithm>
#include <cstring>
#include <vector>

#include "device.h"

namespace cpm {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int ensure_workspace(cpm_ctx* ctx, int64_t n) {
  if (n <= ctx->ws_len) return CPM_OK;
  for (double** b : {&ctx->r, &ctx->p, &ctx->q, &ctx->z})
    if (*b) {
      CPM_HIP(hipFree(*b));
      *b = nullptr;
    }
  for (double** b : {&ctx->r, &ctx->p, &ctx->q, &ctx->z}) CPM_HIP(hipMalloc(b, sizeof(double) * (size_t)n));
  ctx->ws_len = n;
  return CPM_OK;
}

int read_cg_state(cpm_ctx* ctx, int slot, CgState* out) {
  CPM_HIP(hipMemcpyAsync(ctx->pinned, ctx->cg + slot, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
  CPM_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(out, ctx->pinned, sizeof(CgState));
  return CPM_OK;
}

}  // namespace cpm

extern "C" const char* cpm_last_error(void) { return cpm::g_last_error.c_str(); }

extern "C" int cpm_abi_version(void) { return CPM_ABI_VERSION; }

extern "C" int cpm_ctx_create(int device, void* stream, cpm_ctx** out) {
  if (!out) return cpm::fail(CPM_EINVAL, "out is NULL");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return cpm::fail(CPM_ENODEV, "no HIP device visible");
  if (device < 0 || device >= ndev) return cpm::fail(CPM_EINVAL, "device index out of range");
  CPM_HIP(hipSetDevice(device));
  auto* c = new cpm_ctx();
  c->device = device;
  // NULL is the device's null stream (the stream torch uses by default)
  c->stream = (hipStream_t)stream;
  c->own_stream = false;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->ncu = prop.multiProcessorCount;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc(&c->partials, sizeof(double) * cpm::kMaxPartials);
  if (e == hipSuccess) e = hipMalloc(&c->counter, sizeof(unsigned) * 64);
  if (e == hipSuccess) e = hipMalloc(&c->scal, sizeof(double) * 256);
  if (e == hipSuccess) e = hipMalloc(&c->cg, sizeof(cpm::CgState) * 16);
  if (e == hipSuccess) e = hipHostMalloc(&c->pinned, sizeof(double) * 256, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemsetAsync(c->counter, 0, sizeof(unsigned) * 64, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->scal, 0, sizeof(double) * 256, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->cg, 0, sizeof(cpm::CgState) * 16, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    cpm_ctx_destroy(c);
    return cpm::fail(CPM_EHIP, std::string("context allocation: ") + hipGetErrorString(e));
  }
  // the no-allocation CG's device scratch on this device (cpm_cg_noalloc)
  if (int rc = cpm::noalloc_scratch(device, &c->noalloc)) {
    cpm_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return CPM_OK;
}

namespace cpm {

int ensure_lane2(cpm_ctx* ctx) {
  if (ctx->stream2) return CPM_OK;
  CPM_HIP(hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(&ctx->partials2, sizeof(double) * kMaxPartials);
  if (e == hipSuccess) e = hipHostMalloc(&ctx->pinned2, sizeof(double) * 256, hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming);
  // created last: a context with stream2 has every lane-2 resource
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking);
  if (e != hipSuccess) return fail(CPM_EHIP, std::string("second lane: ") + hipGetErrorString(e));
  return CPM_OK;
}

}  // namespace cpm

extern "C" int cpm_ctx_destroy(cpm_ctx* c) {
  if (!c) return CPM_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->stream2) {
    (void)hipStreamSynchronize(c->stream2);
    (void)hipStreamDestroy(c->stream2);
  }
  if (c->partials2) (void)hipFree(c->partials2);
  if (c->pinned2) (void)hipHostFree(c->pinned2);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  for (void* p : {(void*)c->partials, (void*)c->counter, (void*)c->scal, (void*)c->cg, (void*)c->r, (void*)c->p,
                  (void*)c->q, (void*)c->z, (void*)c->r32, (void*)c->p32, (void*)c->q32, (void*)c->x32, (void*)c->b32,
                  (void*)c->y64})
    if (p) (void)hipFree(p);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->ev[0])
    for (auto& e : c->ev) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CPM_OK;
}

extern "C" int cpm_ctx_synchronize(cpm_ctx* c) {
  if (!c) return cpm::fail(CPM_EINVAL, "ctx is NULL");
  CPM_HIP(hipStreamSynchronize(c->stream));
  return CPM_OK;
}

extern "C" int cpm_mat_create_csr(cpm_ctx* ctx, int32_t nrows, int32_t ncols, int64_t nnz, const int64_t* indptr,
                                  const int32_t* indices, const double* values, int nvalsets, cpm_mat** out) {
  if (!ctx || !indptr || !out || (nnz > 0 && (!indices || !values)))
    return cpm::fail(CPM_EINVAL, "cpm_mat_create_csr: null argument");
  if (nrows < 0 || ncols < 0 || nnz < 0 || nvalsets < 1 || nvalsets > 3)
    return cpm::fail(CPM_EINVAL, "cpm_mat_create_csr: bad shape or nvalsets");
  if (nnz > INT32_MAX) return cpm::fail(CPM_EINVAL, "cpm_mat_create_csr: nnz exceeds int32 indexing");
  if (indptr[0] != 0 || indptr[nrows] != nnz) return cpm::fail(CPM_EINVAL, "cpm_mat_create_csr: indptr inconsistent");
  std::vector<int32_t> p32(nrows + 1);
  int64_t maxlen = 0;
  for (int32_t i = 0; i <= nrows; ++i) {
    if (i > 0) {
      if (indptr[i] < indptr[i - 1]) return cpm::fail(CPM_EINVAL, "cpm_mat_create_csr: indptr not monotone");
      maxlen = std::max(maxlen, indptr[i] - indptr[i - 1]);
    }
    p32[i] = (int32_t)indptr[i];
  }
  for (int64_t k = 0; k < nnz; ++k)
    if (indices[k] < 0 || indices[k] >= ncols) return cpm::fail(CPM_EINVAL, "cpm_mat_create_csr: column out of range");
  auto* A = new cpm_mat();
  A->ctx = ctx;
  A->nrows = nrows;
  A->ncols = ncols;
  A->nnz = nnz;
  A->nvals = nvalsets;
  // lanes per row ~ half the mean row length, power of two in [1, 64]
  double mean = nrows ? (double)nnz / nrows : 1.0;
  int lanes = 1;
  while (lanes < 64 && 2 * lanes < mean) lanes *= 2;
  A->lanes = lanes;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipMalloc(&A->ptr, sizeof(int32_t) * (nrows + 1));
  if (e == hipSuccess) e = hipMalloc(&A->idx, sizeof(int32_t) * std::max<int64_t>(nnz, 1));
  if (e == hipSuccess) e = hipMalloc(&A->val, sizeof(double) * std::max<int64_t>(nnz, 1) * nvalsets);
  if (e == hipSuccess) e = hipMemcpyAsync(A->ptr, p32.data(), sizeof(int32_t) * (nrows + 1), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && nnz)
    e = hipMemcpyAsync(A->idx, indices, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && nnz)
    e = hipMemcpyAsync(A->val, values, sizeof(double) * nnz * nvalsets, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    cpm_mat_destroy(A);
    return cpm::fail(e == hipErrorOutOfMemory ? CPM_ENOMEM : CPM_EHIP,
                     std::string("cpm_mat_create_csr: ") + hipGetErrorString(e));
  }
  *out = A;
  return CPM_OK;
}

extern "C" int cpm_mat_destroy(cpm_mat* A) {
  if (!A) return CPM_OK;
  if (A->ctx) {
    (void)hipSetDevice(A->ctx->device);
    (void)hipStreamSynchronize(A->ctx->stream);
  }
  for (void* p : {(void*)A->ptr, (void*)A->idx, (void*)A->val})
    if (p) (void)hipFree(p);
  cpm::sell_free(A->sell);
  delete A;
  return CPM_OK;
}

extern "C" int cpm_mat_shape(const cpm_mat* A, int32_t* nrows, int32_t* ncols, int64_t* nnz) {
  if (!A) return cpm::fail(CPM_EINVAL, "matrix is NULL");
  if (nrows) *nrows = A->nrows;
  if (ncols) *ncols = A->ncols;
  if (nnz) *nnz = A->nnz;
  return CPM_OK;
}

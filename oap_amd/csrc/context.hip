// context.hip -- device selection, error text, device matrices, api objects.
//
// Replaces oapCuda/KernelExecutor.cpp (context + cubin loader) and the device
// matrix utilities of oapMatrixCuda/CudaUtils.cpp:175-359 and
// oapMatrixCuda/oapCudaMatrixUtils.cpp:44-177.  Kernels are launched directly
// (no name-string lookup); a device matrix is a host struct whose planes are
// in HBM, so no device-side struct or host registry is needed.
#include <cstring>
#include <mutex>

#include "internal.h"

namespace oaph {

static thread_local std::string g_last_error;
static std::once_flag g_init_once;
static oaph_status g_init_status = OAPH_OK;
static int g_device = 0;
static bool g_device_chosen = false;

void set_last_error(const char* where, hipError_t e) {
  g_last_error = std::string(where) + ": " + hipGetErrorString(e);
}
void set_last_message(const std::string& msg) { g_last_error = msg; }

oaph_status ensure_init() {
  std::call_once(g_init_once, [] {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
      set_last_message("no HIP device available");
      g_init_status = OAPH_ERR_HIP;
      return;
    }
    int dev = g_device_chosen ? g_device : 0;
    e = hipSetDevice(dev);
    if (e != hipSuccess) {
      set_last_error("hipSetDevice", e);
      g_init_status = OAPH_ERR_HIP;
    }
  });
  if (g_init_status == OAPH_OK) {
    // keep the library's device current on the calling thread
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != (g_device_chosen ? g_device : 0)) {
      OAPH_HIP(hipSetDevice(g_device_chosen ? g_device : 0));
    }
  }
  return g_init_status;
}

oaph_status DevBuf::reserve(size_t n) {
  if (n <= bytes) return OAPH_OK;
  if (ptr) {
    OAPH_HIP(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
  }
  size_t want = n < 4096 ? 4096 : n;
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    set_last_error("hipMalloc", e);
    ptr = nullptr;
    return OAPH_ERR_NOMEM;
  }
  bytes = want;
  return OAPH_OK;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

}  // namespace oaph

using namespace oaph;

extern "C" {

const char* oaph_status_string(oaph_status s) {
  switch (s) {
    case OAPH_OK: return "ok";
    case OAPH_ERR_ARG: return "invalid argument";
    case OAPH_ERR_SHAPE: return "matrix dimensions do not match";
    case OAPH_ERR_PLANE: return "required re/im plane missing";
    case OAPH_ERR_HIP: return "HIP runtime error";
    case OAPH_ERR_NOMEM: return "device allocation failed";
    case OAPH_ERR_UNSUPPORTED: return "unsupported option or size";
    case OAPH_ERR_NOT_CONVERGED: return "solver stopped before convergence";
    case OAPH_ERR_STATE: return "call out of order";
    default: return "unknown status";
  }
}

const char* oaph_last_error(void) { return g_last_error.c_str(); }

oaph_status oaph_init(int device) {
  if (device < 0) return OAPH_ERR_ARG;
  g_device = device;
  g_device_chosen = true;
  return ensure_init();
}

int oaph_device_count(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) return 0;
  return count;
}

oaph_status oaph_device_synchronize(void) {
  OAPH_TRY(ensure_init());
  OAPH_HIP(hipDeviceSynchronize());
  return OAPH_OK;
}

// ---------------------------------------------------------------- matrices
oaph_status oaph_matrix_new(oaph_matrix** out, int isRe, int isIm, uint32_t columns,
                            uint32_t rows) {
  OAPH_CHECK(out != nullptr, OAPH_ERR_ARG);
  OAPH_CHECK(isRe || isIm, OAPH_ERR_ARG);
  OAPH_CHECK(columns > 0 && rows > 0, OAPH_ERR_ARG);
  OAPH_TRY(ensure_init());
  size_t n = (size_t)columns * rows;
  OAPH_CHECK(n < (size_t)1 << 32, OAPH_ERR_UNSUPPORTED);  // uintt indices (oapMath/Math.h)
  oaph_matrix* m = new oaph_matrix();
  m->columns = m->realColumns = columns;
  m->rows = m->realRows = rows;
  m->reValues = m->imValues = nullptr;
  size_t bytes = n * sizeof(double);
  if (isRe && hipMalloc((void**)&m->reValues, bytes) != hipSuccess) {
    set_last_message("hipMalloc failed for re plane");
    delete m;
    return OAPH_ERR_NOMEM;
  }
  if (isIm && hipMalloc((void**)&m->imValues, bytes) != hipSuccess) {
    set_last_message("hipMalloc failed for im plane");
    if (m->reValues) (void)hipFree(m->reValues);
    delete m;
    return OAPH_ERR_NOMEM;
  }
  // the reference zero-fills new device matrices (CudaUtils.cpp:190-200)
  if (m->reValues) OAPH_HIP(hipMemset(m->reValues, 0, bytes));
  if (m->imValues) OAPH_HIP(hipMemset(m->imValues, 0, bytes));
  *out = m;
  return OAPH_OK;
}

oaph_status oaph_matrix_free(oaph_matrix* m) {
  if (!m) return OAPH_OK;
  if (m->reValues) OAPH_HIP(hipFree(m->reValues));
  if (m->imValues) OAPH_HIP(hipFree(m->imValues));
  delete m;
  return OAPH_OK;
}

oaph_status oaph_matrix_info_get(const oaph_matrix* m, oaph_matrix_info* info) {
  OAPH_CHECK(m && info, OAPH_ERR_ARG);
  info->isRe = m->reValues != nullptr;
  info->isIm = m->imValues != nullptr;
  info->columns = m->columns;
  info->rows = m->rows;
  return OAPH_OK;
}

oaph_status oaph_matrix_copy_h2d(oaph_matrix* dst, const double* re, const double* im) {
  OAPH_CHECK(dst, OAPH_ERR_ARG);
  OAPH_TRY(ensure_init());
  size_t bytes = numel(dst) * sizeof(double);
  if (re && dst->reValues) OAPH_HIP(hipMemcpy(dst->reValues, re, bytes, hipMemcpyHostToDevice));
  if (im && dst->imValues) OAPH_HIP(hipMemcpy(dst->imValues, im, bytes, hipMemcpyHostToDevice));
  return OAPH_OK;
}

oaph_status oaph_matrix_copy_d2h(double* re, double* im, const oaph_matrix* src) {
  OAPH_CHECK(src, OAPH_ERR_ARG);
  OAPH_TRY(ensure_init());
  size_t bytes = numel(src) * sizeof(double);
  OAPH_HIP(hipDeviceSynchronize());
  if (re && src->reValues) OAPH_HIP(hipMemcpy(re, src->reValues, bytes, hipMemcpyDeviceToHost));
  if (im && src->imValues) OAPH_HIP(hipMemcpy(im, src->imValues, bytes, hipMemcpyDeviceToHost));
  return OAPH_OK;
}

oaph_status oaph_matrix_copy_d2d(oaph_matrix* dst, const oaph_matrix* src) {
  OAPH_CHECK(dst && src, OAPH_ERR_ARG);
  OAPH_CHECK(numel(dst) == numel(src), OAPH_ERR_SHAPE);
  OAPH_TRY(ensure_init());
  size_t bytes = numel(src) * sizeof(double);
  if (dst->reValues && src->reValues)
    OAPH_HIP(hipMemcpy(dst->reValues, src->reValues, bytes, hipMemcpyDeviceToDevice));
  if (dst->imValues && src->imValues)
    OAPH_HIP(hipMemcpy(dst->imValues, src->imValues, bytes, hipMemcpyDeviceToDevice));
  return OAPH_OK;
}

oaph_status oaph_matrix_get_value(const oaph_matrix* m, uint32_t index, double* re, double* im) {
  OAPH_CHECK(m, OAPH_ERR_ARG);
  OAPH_CHECK(index < numel(m), OAPH_ERR_ARG);
  OAPH_TRY(ensure_init());
  OAPH_HIP(hipDeviceSynchronize());
  if (re) {
    *re = 0;
    if (m->reValues)
      OAPH_HIP(hipMemcpy(re, m->reValues + index, sizeof(double), hipMemcpyDeviceToHost));
  }
  if (im) {
    *im = 0;
    if (m->imValues)
      OAPH_HIP(hipMemcpy(im, m->imValues + index, sizeof(double), hipMemcpyDeviceToHost));
  }
  return OAPH_OK;
}

oaph_status oaph_matrix_set_value(oaph_matrix* m, uint32_t index, double re, double im) {
  OAPH_CHECK(m, OAPH_ERR_ARG);
  OAPH_CHECK(index < numel(m), OAPH_ERR_ARG);
  OAPH_TRY(ensure_init());
  OAPH_HIP(hipDeviceSynchronize());
  if (m->reValues)
    OAPH_HIP(hipMemcpy(m->reValues + index, &re, sizeof(double), hipMemcpyHostToDevice));
  if (m->imValues)
    OAPH_HIP(hipMemcpy(m->imValues + index, &im, sizeof(double), hipMemcpyHostToDevice));
  return OAPH_OK;
}

oaph_status oaph_matrix_set_zero_row(oaph_matrix* m, uint32_t row, int re, int im) {
  OAPH_CHECK(m && row < m->rows, OAPH_ERR_ARG);
  OAPH_TRY(ensure_init());
  size_t off = (size_t)row * m->columns, bytes = (size_t)m->columns * sizeof(double);
  OAPH_HIP(hipDeviceSynchronize());
  if (re && m->reValues) OAPH_HIP(hipMemset(m->reValues + off, 0, bytes));
  if (im && m->imValues) OAPH_HIP(hipMemset(m->imValues + off, 0, bytes));
  return OAPH_OK;
}

// Copy of a columns x rows block between matrices (either side host or
// device), plane by plane, for the planes both have.
static oaph_status copy_block(oaph_matrix* dst, const oaph_matrix* src, uint32_t dcol,
                              uint32_t drow, uint32_t scol, uint32_t srow, uint32_t columns,
                              uint32_t rows, int planes, hipMemcpyKind kind) {
  OAPH_CHECK((uint64_t)dcol + columns <= dst->columns && (uint64_t)drow + rows <= dst->rows &&
                 (uint64_t)scol + columns <= src->columns && (uint64_t)srow + rows <= src->rows,
             OAPH_ERR_SHAPE);
  if (columns == 0 || rows == 0) return OAPH_OK;
  OAPH_TRY(ensure_init());
  const size_t doff = (size_t)drow * dst->columns + dcol, soff = (size_t)srow * src->columns + scol;
  const size_t dpitch = (size_t)dst->columns * sizeof(double), spitch = (size_t)src->columns * sizeof(double);
  OAPH_HIP(hipDeviceSynchronize());
  if ((planes & 1) && dst->reValues && src->reValues)
    OAPH_HIP(hipMemcpy2D(dst->reValues + doff, dpitch, src->reValues + soff, spitch,
                         (size_t)columns * sizeof(double), rows, kind));
  if ((planes & 2) && dst->imValues && src->imValues)
    OAPH_HIP(hipMemcpy2D(dst->imValues + doff, dpitch, src->imValues + soff, spitch,
                         (size_t)columns * sizeof(double), rows, kind));
  return OAPH_OK;
}

static oaph_status copy_dims(oaph_matrix* dst, const oaph_matrix* src,
                             const uint32_t dims[2][2][2], hipMemcpyKind kind) {
  OAPH_CHECK(dst && src && dims, OAPH_ERR_ARG);
  // check_CopyMatrixToMatrixDims (oapHostMatrixUtils.h:803-817): equal block
  // sizes, same planes
  OAPH_CHECK(dims[0][0][1] == dims[1][0][1] && dims[0][1][1] == dims[1][1][1], OAPH_ERR_SHAPE);
  OAPH_CHECK((dst->reValues != nullptr) == (src->reValues != nullptr) &&
                 (dst->imValues != nullptr) == (src->imValues != nullptr),
             OAPH_ERR_PLANE);
  return copy_block(dst, src, dims[0][0][0], dims[0][1][0], dims[1][0][0], dims[1][1][0],
                    dims[0][0][1], dims[0][1][1], 3, kind);
}

oaph_status oaph_matrix_copy_d2d_dims(oaph_matrix* dst, const oaph_matrix* src,
                                      const uint32_t dims[2][2][2]) {
  return copy_dims(dst, src, dims, hipMemcpyDeviceToDevice);
}
oaph_status oaph_matrix_copy_h2d_dims(oaph_matrix* dst, const oaph_matrix* src_host,
                                      const uint32_t dims[2][2][2]) {
  return copy_dims(dst, src_host, dims, hipMemcpyHostToDevice);
}
oaph_status oaph_matrix_copy_d2h_dims(oaph_matrix* dst_host, const oaph_matrix* src,
                                      const uint32_t dims[2][2][2]) {
  return copy_dims(dst_host, src, dims, hipMemcpyDeviceToHost);
}

oaph_status oaph_matrix_set_submatrix(oaph_matrix* dst, const oaph_matrix* src, uint32_t column,
                                      uint32_t row, int planes) {
  OAPH_CHECK(dst && src && planes >= 1 && planes <= 3, OAPH_ERR_ARG);
  return copy_block(dst, src, column, row, 0, 0, src->columns, src->rows, planes,
                    hipMemcpyDeviceToDevice);
}

// ---------------------------------------------------------------- api objects
oaph_status oaph_api_create(oaph_api** out) {
  OAPH_CHECK(out, OAPH_ERR_ARG);
  OAPH_TRY(ensure_init());
  oaph_api* api = new oaph_api();
  if (hipStreamCreateWithFlags(&api->stream, hipStreamNonBlocking) != hipSuccess) {
    set_last_message("hipStreamCreate failed");
    delete api;
    return OAPH_ERR_HIP;
  }
  oaph_status st = api->counters.reserve(4096);
  if (st == OAPH_OK) st = api->scalars.reserve(4096);
  if (st == OAPH_OK && hipMemset(api->counters.ptr, 0, api->counters.bytes) != hipSuccess)
    st = OAPH_ERR_HIP;
  if (st == OAPH_OK &&
      hipHostMalloc((void**)&api->host_scalars, 64 * sizeof(double), hipHostMallocDefault) !=
          hipSuccess)
    st = OAPH_ERR_NOMEM;
  if (st != OAPH_OK) {
    oaph_api_destroy(api);
    return st;
  }
  *out = api;
  return OAPH_OK;
}

oaph_status oaph_api_destroy(oaph_api* api) {
  if (!api) return OAPH_OK;
  if (api->stream) (void)hipStreamSynchronize(api->stream);
  api->partials.release();
  api->counters.release();
  api->scalars.release();
  api->work.release();
  if (api->host_scalars) (void)hipHostFree(api->host_scalars);
  if (api->stream) (void)hipStreamDestroy(api->stream);
  delete api;
  return OAPH_OK;
}

oaph_status oaph_api_synchronize(oaph_api* api) {
  OAPH_CHECK(api, OAPH_ERR_ARG);
  OAPH_HIP(hipStreamSynchronize(api->stream));
  return OAPH_OK;
}

void* oaph_api_stream(oaph_api* api) { return api ? (void*)api->stream : nullptr; }

// utils::mapper::SetThreadsBlocks (oapUtils/ThreadsMapper.cpp:29-74).  Kept
// bit-exact: the integer square root of the limit, the four cases, and the
// proportional split (factor = w/(w+h) in double, truncated to integers).
oaph_status oaph_set_threads_blocks(uint32_t blocks[2], uint32_t threads[2], uint32_t w,
                                    uint32_t h, uint32_t limit) {
  if (!blocks || !threads || limit == 0) return OAPH_ERR_ARG;
  const uint32_t sq = (uint32_t)std::sqrt((double)limit);
  auto grow = [&](int i, uint32_t extent) {
    while (threads[i] * blocks[i] < extent) ++blocks[i];
  };
  blocks[0] = blocks[1] = 1;
  if (w * h < limit) {
    threads[0] = w;
    threads[1] = h;
  } else if (sq <= w && sq <= h) {
    threads[0] = threads[1] = sq;
    grow(0, w);
    grow(1, h);
  } else if (sq <= w) {  // sq > h
    threads[0] = sq;
    threads[1] = h;
    grow(0, w);
  } else if (sq <= h) {  // sq > w
    threads[0] = w;
    threads[1] = sq;
    grow(1, h);
  } else {
    const double share = w / static_cast<double>(w + h);
    const uint32_t budget = 2 * sq;
    threads[0] = (uint32_t)(share * budget);
    threads[1] = (uint32_t)((1 - share) * budget);
    grow(0, w);
    grow(1, h);
  }
  return OAPH_OK;
}

}  // extern "C"

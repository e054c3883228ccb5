/*
 * y = A x without allocating: dpdlp_dev_spmv for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).
 *
 * A matrix stored as XCD column-panel tiles (P > 1) needs scratch beyond its
 * inputs and outputs: the P partial row sums of every row and one ticket per
 * row block (dpdlp_csr_t.xpart / .tick, which dpdlp_dev_spmv's caller
 * provides).  Here they come from the file-scope array `scratch`, one copy
 * per GPU, held until the process ends: DPDLP_NOALLOC_SUMS doubles of partial
 * sums (P * rows at most, 32 MiB) followed by DPDLP_NOALLOC_TICKETS 4-byte
 * tickets (row blocks at most).  An untiled matrix needs no scratch.
 *
 * Every call on a GPU shares that GPU's array and leaves its contents behind:
 * each call clears the tickets it uses (with a fill kernel, so the call can be
 * captured into a graph) before the SpMV, and the partial sums are written
 * before they are read.  Calls that overlap on one GPU (two streams or two
 * threads at once) race on the array and must not be issued.
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "dev_common.h"
#include "dpdlp_dev.h"

extern "C" void dpdlp_set_error(const char *fmt, ...);

static_assert(DPDLP_NOALLOC_TICKETS % 2 == 0, "tickets fill whole doubles");

static __device__ __align__(256) double scratch[DPDLP_NOALLOC_SUMS + DPDLP_NOALLOC_TICKETS / 2];

#define NOALLOC_MAX_DEVICES 64
static double *g_scratch_addr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

/* Device address of the current GPU's copy of `scratch`, looked up once per
 * device (hipGetSymbolAddress with that device current) and kept: in host
 * code the array's own name is not a device address. */
static hipError_t scratch_address(double **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[dev] = (double *)p;
  }
  *out = g_scratch_addr[dev];
  return hipSuccess;
}

__global__ void k_clear_tickets(unsigned *t, int count) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
    t[i] = 0u;
}

extern "C" int dpdlp_dev_spmv_noalloc(const dpdlp_csr_t *A, const double *x, double *y,
                                      void *stream) {
  if (!A) {
    dpdlp_set_error("dpdlp_dev_spmv_noalloc: no matrix");
    return (int)hipErrorInvalidValue;
  }
  dpdlp_csr_t B = *A;
  if (B.P > 1) {
    const long long sums = (long long)B.P * B.rows;
    if (sums > (long long)DPDLP_NOALLOC_SUMS || B.nblk > DPDLP_NOALLOC_TICKETS) {
      dpdlp_set_error("dpdlp_dev_spmv_noalloc: %d panels x %d rows (%lld partial sums) and %d "
                      "row blocks exceed the fixed scratch (%d partial sums, %d row blocks)",
                      B.P, B.rows, sums, B.nblk, (int)DPDLP_NOALLOC_SUMS,
                      (int)DPDLP_NOALLOC_TICKETS);
      return (int)hipErrorInvalidValue;
    }
    double *base = nullptr;
    hipError_t e = scratch_address(&base);
    if (e != hipSuccess) {
      dpdlp_set_error("dpdlp_dev_spmv_noalloc: no device address for the scratch (%s)",
                      hipGetErrorString(e));
      return (int)e;
    }
    B.xpart = base;
    B.tick = (unsigned *)(base + DPDLP_NOALLOC_SUMS);
    if (B.nblk > 0) {
      int g = (B.nblk + DPDLP_BS - 1) / DPDLP_BS;
      if (g > 1024) g = 1024;
      k_clear_tickets<<<g, DPDLP_BS, 0, (hipStream_t)stream>>>(B.tick, B.nblk);
      e = hipGetLastError();
      if (e != hipSuccess) return (int)e;
    }
  }
  return dpdlp_dev_spmv(&B, x, y, stream);
}

// Allocation-free Pfaffians: tg_zskpfa_noalloc.
//
// The same computation as tg_zskpfa (ZSKPFA('U','P'), the batched matrix-core
// Pfaffian of tg_linalg.hip) for callers that cannot allocate device memory:
// every device buffer the call needs -- the copy of the matrices, the Pfaffian's
// work matrices, log|pf|, the phases and the zero flags -- lives in one
// file-scope __device__ array of TG_NOALLOC_SCRATCH_BYTES.  Its device address
// is looked up once per device (hipGetSymbolAddress, on the first call with
// that device current) and only that address is handed to copies and kernels:
// in host code the array's name is not a device address.
//
// Each device holds the array until the process ends.  Calls on one device
// share it and leave the last call's contents behind; nothing in it has to
// start at zero (every element a kernel reads was written by the same call).
// Calls that overlap on one device race.
#include "tg_internal.h"

#include <mutex>

namespace {

constexpr size_t kScratchBytes = TG_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of this device's scratch array (looked up once per device)
int scratch_for_current_device(void** out) {
  int dev = -1;
  TG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    tg_set_error("tg_zskpfa_noalloc: device index out of range");
    return TG_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      tg_set_error(std::string("tg_zskpfa_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
      return TG_E_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TG_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" int tg_zskpfa_noalloc(int n, int count, const double* A, int lda, double* pfaff,
                                 int* info) {
  if (n < 0 || count < 0 || lda < std::max(1, n) || (count > 0 && (!A || !pfaff || !info))) {
    tg_set_error("tg_zskpfa_noalloc: invalid argument");
    return TG_E_ARG;
  }
  if (count == 0) return TG_OK;
  if (n == 0 || n % 2 == 1) {
    for (int k = 0; k < count; ++k) {
      pfaff[2 * k] = (n == 0) ? 1.0 : 0.0;
      pfaff[2 * k + 1] = 0.0;
      info[k] = 0;
    }
    return TG_OK;
  }
  const size_t m = (size_t)lda * n;
  const size_t bA = align256(sizeof(cplx) * m * count);
  const size_t bW = align256(sizeof(cplx) * (size_t)n * n * count);
  const size_t bPh = align256(sizeof(cplx) * count);
  const size_t bL = align256(sizeof(double) * count);
  const size_t bZ = align256(sizeof(int) * count);
  const size_t need = bA + bW + bPh + bL + bZ;
  if (need > kScratchBytes) {
    tg_set_error("tg_zskpfa_noalloc: " + std::to_string(count) + " matrices of order " +
                 std::to_string(n) + " (lda " + std::to_string(lda) + ") need " +
                 std::to_string(need) + " bytes of scratch; the limit is " +
                 std::to_string(kScratchBytes) + " (use tg_zskpfa, or fewer matrices per call)");
    return TG_E_ARG;
  }
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt < 1) {
    tg_set_error("no HIP device available");
    return TG_E_NODEVICE;
  }
  void* base = nullptr;
  TG_TRY(scratch_for_current_device(&base));
  unsigned char* p = static_cast<unsigned char*>(base);
  cplx* dA = reinterpret_cast<cplx*>(p);
  cplx* dW = reinterpret_cast<cplx*>(p + bA);
  cplx* dPh = reinterpret_cast<cplx*>(p + bA + bW);
  double* dL = reinterpret_cast<double*>(p + bA + bW + bPh);
  int* dZ = reinterpret_cast<int*>(p + bA + bW + bPh + bL);
  TG_HIP(hipMemcpy(dA, A, sizeof(cplx) * m * count, hipMemcpyHostToDevice));
  TG_TRY(batched_pfaffian(0, n, dA, lda, (long long)m, count, dW, dL, dPh, dZ));
  std::vector<double> L(count);
  std::vector<cplx> Ph(count);
  std::vector<int> Zr(count);
  TG_HIP(hipMemcpy(L.data(), dL, sizeof(double) * count, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(Ph.data(), dPh, sizeof(cplx) * count, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(Zr.data(), dZ, sizeof(int) * count, hipMemcpyDeviceToHost));
  for (int k = 0; k < count; ++k) {
    const double mag = Zr[k] ? 0.0 : exp(L[k]);
    pfaff[2 * k] = mag * Ph[k].x;
    pfaff[2 * k + 1] = mag * Ph[k].y;
    info[k] = 0;
  }
  return TG_OK;
}

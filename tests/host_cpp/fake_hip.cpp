// Host stand-in for the HIP runtime calls and kernel launchers that the
// server's host code (pnx_server.cpp) uses, so that its threads can run under
// ThreadSanitizer and AddressSanitizer on a machine without a GPU.  Device
// memory is host memory, copies are memcpy in the calling thread, streams and
// events are complete as soon as work is enqueued, and kernels do nothing
// (fills and copies are carried out, so buffers hold defined values).
//
// fake_fail_allocation(k) / fake_fail_launch(k) make the k-th device
// allocation (hipMalloc) or kernel launch (any launcher) from now on fail;
// k = 0 turns the injection off.
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <new>

#include "../../pnepma_amd/csrc/pnx_kernels.hpp"
#include "../../pnepma_amd/csrc/pnx_march.hpp"

static std::atomic<long> n_malloc{0}, n_launch{0}, fail_malloc_at{0}, fail_launch_at{0};

void fake_fail_allocation(long k) { fail_malloc_at = k > 0 ? n_malloc.load() + k : 0; }
void fake_fail_launch(long k) { fail_launch_at = k > 0 ? n_launch.load() + k : 0; }

static bool inject(std::atomic<long>& counter, const std::atomic<long>& at) {
  const long n = ++counter;
  const long a = at.load();
  return a > 0 && n == a;
}

static hipError_t launched() { return inject(n_launch, fail_launch_at) ? hipErrorLaunchFailure : hipSuccess; }

extern "C" {

hipError_t hipMalloc(void** ptr, size_t size) {
  if (inject(n_malloc, fail_malloc_at)) {
    *ptr = nullptr;
    return hipErrorOutOfMemory;
  }
  *ptr = std::aligned_alloc(256, (size + 255) / 256 * 256);
  return *ptr ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipFree(void* ptr) {
  std::free(ptr);
  return hipSuccess;
}

hipError_t hipHostMalloc(void** ptr, size_t size, unsigned int) {
  *ptr = std::aligned_alloc(256, (size + 255) / 256 * 256);
  return *ptr ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipHostFree(void* ptr) {
  std::free(ptr);
  return hipSuccess;
}

hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
  if (bytes) std::memcpy(dst, src, bytes);
  return hipSuccess;
}

hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t) {
  if (bytes) std::memcpy(dst, src, bytes);
  return hipSuccess;
}

hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }

hipError_t hipStreamCreateWithFlags(hipStream_t* stream, unsigned int) {
  *stream = reinterpret_cast<hipStream_t>(new char[1]);
  return hipSuccess;
}

hipError_t hipStreamDestroy(hipStream_t stream) {
  delete[] reinterpret_cast<char*>(stream);
  return hipSuccess;
}

hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }

hipError_t hipEventCreateWithFlags(hipEvent_t* event, unsigned) {
  *event = reinterpret_cast<hipEvent_t>(new char[1]);
  return hipSuccess;
}

hipError_t hipEventDestroy(hipEvent_t event) {
  delete[] reinterpret_cast<char*>(event);
  return hipSuccess;
}

hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }

}  // extern "C"

namespace pnx {

hipError_t launch_half_step_grouped(int, int, int, int, const HalfStepArgs*, const int2*, const int*, int, hipStream_t) {
  return launched();
}
hipError_t launch_phase_a_grouped(int, int, const HalfStepArgs*, const RevArgs*, const int2*, const int*, int,
                                  hipStream_t) {
  return launched();
}
hipError_t launch_pair_grouped(int, int, const HalfStepArgs*, const int2*, const int*, int, hipStream_t) {
  return launched();
}
hipError_t launch_phase_b_grouped(int, int, const HalfStepArgs*, const int2*, const int*, int, hipStream_t) {
  return launched();
}
hipError_t launch_march(const MarchArgs*, int, int, hipStream_t) { return launched(); }
hipError_t launch_material_densities(const Layout&, const int[8][3], const long[8], const int[8], const double*,
                                     const double*, double*, hipStream_t) {
  return launched();
}
hipError_t launch_material_pullback(const Layout&, const int[8][3], const long[8], const int[8], const double*,
                                    const double*, const double*, long, int, const double*, bool, double*,
                                    hipStream_t, bool) {
  return launched();
}
hipError_t launch_absorption(const AbsorptionGeometry&, const Layout&, const double*, const double*, const double*,
                             const double*, const double*, const double*, const int*, double*, double*, double*,
                             hipStream_t) {
  return launched();
}
hipError_t launch_atoms_bar(const AbsorptionGeometry&, const Layout&, const double*, const double*, long, int,
                            const double*, const int*, double*, hipStream_t) {
  return launched();
}
hipError_t launch_absorption_pullback(const AbsorptionGeometry&, const Layout&, const double*, const double*,
                                      const double*, const double*, const int*, const double*, long, int, double*,
                                      hipStream_t) {
  return launched();
}
hipError_t launch_fill(double* p, long n, double v, hipStream_t) {
  for (long i = 0; i < n; ++i) p[i] = v;
  return launched();
}
hipError_t launch_copy(double* dst, const double* src, long n, hipStream_t) {
  if (n > 0) std::memmove(dst, src, n * sizeof(double));
  return launched();
}
hipError_t launch_reduce_intensity(double*, long, long, int, const double*, int, int, int, int, double, double, int,
                                   double*, hipStream_t) {
  return launched();
}

}  // namespace pnx

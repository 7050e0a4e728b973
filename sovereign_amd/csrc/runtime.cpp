// Device probe, library stream, error text and operation counters.
//
// Replaces the reference's driver bootstrap (solver/gpu/src/cuda_driver.cpp:
// 661-758): instead of dlopen + PTX JIT it links the HIP runtime and ships
// gfx950 code objects; the self-test of cuda_driver.cpp:616-659 is kept so
// that "available" still means "the kernels gave right answers here".

#include "internal.hpp"

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace sov {
namespace {

thread_local std::string t_error;
std::atomic<unsigned long long> g_ops{0};
std::atomic<unsigned long long> g_facts{0};

struct Device {
  std::once_flag once;
  sov_dev_info info{};
  hipStream_t stream = nullptr;
};

Device& dev() {
  static Device d;
  return d;
}

void copy_text(char* dst, std::size_t cap, const std::string& src) {
  std::strncpy(dst, src.c_str(), cap - 1);
  dst[cap - 1] = '\0';
}

// The dense-LU and sparse-product checks of the reference's probe, run on the
// device: a 5x5 LU rebuilt as P*A = L*U, and A = [[2,0,.5],[0,4,3]] times
// (1,2,3) and its transpose times (1,-1).
std::string self_test() {
  const int n = 5;
  std::vector<double> a(n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[j * n + i] = std::sin(1.0 + 3.0 * i + 7.0 * j) + (i == j ? 0.5 : 0.0);
  std::vector<double> lu = a;
  std::vector<int32_t> piv(n);
  if (dense_lu_run(n, lu.data(), piv.data()) != SOV_OK) return "dense LU self-test did not run";
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      double sum = 0.0;
      for (int k = 0; k <= std::min(i, j); ++k) sum += (k == i ? 1.0 : lu[k * n + i]) * lu[j * n + k];
      if (std::abs(sum - a[j * n + piv[i]]) > 1e-9) return "dense LU self-test gave a wrong factorization";
    }
  }
  const int32_t ptr[] = {0, 1, 2, 4};
  const int32_t idx[] = {0, 1, 0, 1};
  const double val[] = {2.0, 4.0, 0.5, 3.0};
  const double x[] = {1.0, 2.0, 3.0};
  const double z[] = {1.0, -1.0};
  double y[2] = {0, 0}, w[3] = {0, 0, 0};
  sov_mat* m = nullptr;
  if (mat_upload(2, 3, ptr, idx, val, &m) != SOV_OK) return "sparse product self-test did not run";
  bool ran = true;
  ran = ran && hipMemcpyAsync(m->x_buf, x, sizeof(x), hipMemcpyHostToDevice, stream()) == hipSuccess;
  ran = ran && spmv_launch(m, false, m->x_buf, m->y_buf, stream()) == SOV_OK;
  ran = ran && hipMemcpyAsync(y, m->y_buf, sizeof(y), hipMemcpyDeviceToHost, stream()) == hipSuccess;
  ran = ran && hipMemcpyAsync(m->x_buf, z, sizeof(z), hipMemcpyHostToDevice, stream()) == hipSuccess;
  ran = ran && spmv_launch(m, true, m->x_buf, m->y_buf, stream()) == SOV_OK;
  ran = ran && hipMemcpyAsync(w, m->y_buf, sizeof(w), hipMemcpyDeviceToHost, stream()) == hipSuccess;
  ran = ran && hipStreamSynchronize(stream()) == hipSuccess;
  mat_release(m);
  if (!ran) return "sparse product self-test did not run";
  const double ey[] = {3.5, 17.0}, ew[] = {2.0, -4.0, -2.5};
  for (int i = 0; i < 2; ++i)
    if (std::abs(y[i] - ey[i]) > 1e-12) return "sparse product self-test gave a wrong result";
  for (int i = 0; i < 3; ++i)
    if (std::abs(w[i] - ew[i]) > 1e-12) return "sparse product self-test gave a wrong result";
  return "";
}

void probe(Device& d) {
  sov_dev_info& info = d.info;
  std::memset(&info, 0, sizeof(info));
  const char* off = std::getenv("SOVEREIGN_DISABLE_HIP");
  if (off && std::string(off) == "1") {
    copy_text(info.reason, sizeof(info.reason), "Disabled by SOVEREIGN_DISABLE_HIP=1.");
    return;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    copy_text(info.reason, sizeof(info.reason), "No HIP-capable GPU is visible to the runtime.");
    return;
  }
  hipDeviceProp_t prop{};
  if (hipSetDevice(0) != hipSuccess || hipGetDeviceProperties(&prop, 0) != hipSuccess) {
    copy_text(info.reason, sizeof(info.reason), "Could not open GPU 0.");
    return;
  }
  copy_text(info.name, sizeof(info.name), prop.name);
  copy_text(info.arch, sizeof(info.arch), prop.gcnArchName);
  info.compute_major = prop.major;
  info.compute_minor = prop.minor;
  info.memory_bytes = prop.totalGlobalMem;
  info.multiprocessors = prop.multiProcessorCount;
  if (hipDriverGetVersion(&info.driver_version) != hipSuccess) info.driver_version = 0;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    copy_text(info.reason, sizeof(info.reason),
              std::string("This library is built for gfx950 (MI355X); found ") + prop.gcnArchName + ".");
    return;
  }
  if (hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) != hipSuccess) {
    copy_text(info.reason, sizeof(info.reason), "Could not create a HIP stream.");
    return;
  }
  info.available = 1;  // the self-test below needs the stream
  const std::string failure = self_test();
  if (!failure.empty()) {
    info.available = 0;
    copy_text(info.reason, sizeof(info.reason), "GPU " + failure + ".");
  }
}

}  // namespace

void set_error(const std::string& text) { t_error = text; }
const char* last_error() { return t_error.c_str(); }

hipStream_t stream() { return dev().stream; }

bool device_ready() {
  Device& d = dev();
  std::call_once(d.once, [&] { probe(d); });
  return d.info.available != 0;
}

void count_operation() { ++g_ops; }
void count_factorization() { ++g_facts; }

}  // namespace sov

extern "C" {

int sov_dev_probe(sov_dev_info* out) {
  const bool ok = sov::device_ready();
  if (out) *out = sov::dev().info;
  return ok ? SOV_OK : SOV_ERR_NODEV;
}

uint64_t sov_gpu_operations(void) { return sov::g_ops.load(); }
uint64_t sov_gpu_factorizations(void) { return sov::g_facts.load(); }
void sov_reset_gpu_operations(void) {
  sov::g_ops.store(0);
  sov::g_facts.store(0);
}

const char* sov_last_error(void) { return sov::last_error(); }

int sov_sync(void) {
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  SOV_HIP_TRY(hipStreamSynchronize(sov::stream()));
  return SOV_OK;
}

void* sov_stream(void) { return sov::device_ready() ? static_cast<void*>(sov::stream()) : nullptr; }

}  // extern "C"

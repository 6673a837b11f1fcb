// Runtime pieces of the C ABI: status strings, per-thread error messages,
// device memory, state preparation (basis states, deterministic fills),
// dtype conversion and the streaming copy kernel.
#include "brk_internal.hpp"

#include <cctype>
#include <cerrno>
#include <cstdio>

#include <memory>
#include <mutex>
#include <set>
#include <utility>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

namespace brk {

static thread_local std::string g_last_error;

int fail(int status, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

void clear_error() { g_last_error.clear(); }

int check_state(const void* first, uint64_t size, int dtype, unsigned* num_qubits) {
  if (first == nullptr) return fail(BRK_ERR_INVALID_ARGUMENT, "state pointer is null");
  if (size == 0 || (size & (size - 1)) != 0)
    return fail(BRK_ERR_INVALID_ARGUMENT, "state size %llu is not a power of two",
                (unsigned long long)size);
  if (dtype != BRK_C128 && dtype != BRK_C64)
    return fail(BRK_ERR_UNSUPPORTED_DTYPE, "unknown dtype %d", dtype);
  *num_qubits = (unsigned)__builtin_ctzll(size);
  return BRK_OK;
}

int collect_qubits(const unsigned* qubits, unsigned n, unsigned num_qubits, unsigned* out, unsigned* count,
                   const char* role) {
  if (n > 0 && qubits == nullptr) return fail(BRK_ERR_INVALID_ARGUMENT, "%s list is null", role);
  for (unsigned i = 0; i < n; ++i) {
    unsigned q = qubits[i];
    if (q >= num_qubits)
      return fail(BRK_ERR_QUBIT_OUT_OF_RANGE, "%s qubit %u out of range for %u qubits", role, q, num_qubits);
    for (unsigned j = 0; j < *count; ++j)
      if (out[j] == q) return fail(BRK_ERR_DUPLICATE_QUBIT, "qubit %u given twice", q);
    if (*count >= BRK_MAX_OPERATED_QUBITS)
      return fail(BRK_ERR_TOO_MANY_QUBITS, "more than %d operated qubits", BRK_MAX_OPERATED_QUBITS);
    out[(*count)++] = q;
  }
  return BRK_OK;
}

// Defaults measured on MI355X with 2^30..2^32-amplitude c128 states.
static tuning g_tuning = {
    /*pair_lds=*/40960, /*diag_lds=*/0, /*pair_remap=*/1, /*diag_remap=*/1,
    /*nt_min_bytes=*/256ull << 20, /*max_blocks=*/(1ull << 24) - 1, /*line_bytes=*/64, /*fused_groups=*/8192,
    /*fused_occupancy=*/2, /*fused_staged=*/1, /*fused_staged_groups=*/768, /*fused_chunked=*/0,
    /*block_lds=*/65536, /*block_run_bytes=*/128, /*adjoint_blocks=*/1, /*fused_chunks=*/1,
};
static std::once_flag g_tuning_once;

static int set_tuning(tuning& t, const std::string& k, int64_t value);

// BRK_TUNE_<KEY> environment values go through the same checks as brk_tune; a rejected
// value keeps the default and is reported on stderr.
static void read_tuning_env() {
  static const char* const keys[] = {"pair_lds",     "diag_lds",          "pair_remap",          "diag_remap",
                                     "nt_min_bytes", "max_blocks",        "line_bytes",          "fused_groups",
                                     "fused_occupancy", "fused_staged",   "fused_staged_groups", "fused_chunked",
                                     "block_lds",    "block_run_bytes",   "adjoint_blocks",      "fused_chunks"};
  for (const char* key : keys) {
    std::string env = "BRK_TUNE_";
    for (const char* c = key; *c; ++c) env.push_back((char)std::toupper((unsigned char)*c));
    const char* v = std::getenv(env.c_str());
    if (!v) continue;
    char* end = nullptr;
    errno = 0;
    const long long value = std::strtoll(v, &end, 0);
    if (errno || end == v || *end != '\0' || set_tuning(g_tuning, key, value) != BRK_OK)
      std::fprintf(stderr, "braket-mi355x: ignoring %s=%s (%s)\n", env.c_str(), v,
                   errno || end == v || *end ? "not an integer" : g_last_error.c_str());
  }
}

tuning& get_tuning() {
  std::call_once(g_tuning_once, read_tuning_env);
  return g_tuning;
}

unsigned device_cus() {
  static std::mutex m;
  static std::vector<unsigned> cus;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 1;
  std::lock_guard<std::mutex> lock(m);
  if ((int)cus.size() <= dev) cus.resize(dev + 1, 0);
  if (!cus[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 1;
    cus[dev] = (unsigned)n;
  }
  return cus[dev];
}

int allow_dynamic_lds(const void* kernel, size_t bytes) {
  static std::mutex m;
  static std::set<std::pair<const void*, int>> done;
  int dev = 0;
  BRK_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(m);
  if (done.count({kernel, dev})) return BRK_OK;
  BRK_HIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  done.insert({kernel, dev});
  return BRK_OK;
}

static std::mutex g_scratch_mutex;
static std::vector<std::unique_ptr<reduce_scratch>> g_scratch;  // stable addresses: each holds a mutex

int get_reduce_scratch(reduce_scratch** out) {
  int dev = 0;
  BRK_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
  if (!g_scratch[dev]) g_scratch[dev] = std::make_unique<reduce_scratch>();
  reduce_scratch& s = *g_scratch[dev];
  if (s.device == nullptr) {
    size_t bytes = sizeof(double) * RED_BLOCKS * RED_MAX_VALUES;
    if (hipMalloc(&s.device, bytes) != hipSuccess) {
      s.device = nullptr;
      return fail(BRK_ERR_OUT_OF_MEMORY, "cannot allocate %zu bytes of reduction scratch", bytes);
    }
    if (hipHostMalloc(&s.host, bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipFree(s.device);
      s.device = nullptr;
      return fail(BRK_ERR_OUT_OF_MEMORY, "cannot allocate pinned reduction scratch");
    }
  }
  *out = &s;
  return BRK_OK;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

template <typename R>
__global__ void k_set_basis(cpx<R>* __restrict__ s, uint64_t size, uint64_t basis) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < size; i += stride) s[i] = cpx<R>{i == basis ? R(1) : R(0), R(0)};
}

struct modular_params {
  double scale_re, scale_im;
  uint64_t mul_re, add_re, mod_re, mul_im, add_im, mod_im;
};

__device__ __forceinline__ double modular_term(uint64_t i, uint64_t mul, uint64_t add, uint64_t mod) {
  uint64_t v = i * mul + add;
  if (mod != 0) v %= mod;
  return (double)v;
}

template <typename R>
__global__ void k_fill_modular(cpx<R>* __restrict__ s, uint64_t size, modular_params p) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < size; i += stride) {
    double re = p.scale_re * modular_term(i, p.mul_re, p.add_re, p.mod_re);
    double im = p.scale_im * modular_term(i, p.mul_im, p.add_im, p.mod_im);
    s[i] = cpx<R>{(R)re, (R)im};
  }
}

// splitmix64 finaliser; value in [-1, 1) with 53 random bits.
__device__ __forceinline__ double hash_unit(uint64_t seed, uint64_t counter) {
  uint64_t z = seed + (counter + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

template <typename R>
__global__ void k_fill_hash(cpx<R>* __restrict__ s, uint64_t size, uint64_t seed, double scale) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < size; i += stride)
    s[i] = cpx<R>{(R)(scale * hash_unit(seed, 2 * i)), (R)(scale * hash_unit(seed, 2 * i + 1))};
}

template <typename RD, typename RS>
__global__ void k_convert(cpx<RD>* __restrict__ d, const cpx<RS>* __restrict__ s, uint64_t size) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < size; i += stride) {
    cpx<RS> v = s[i];
    d[i] = cpx<RD>{(RD)v.re, (RD)v.im};
  }
}

// Streaming copy with the gate kernels' access shape: one whole amplitude per thread,
// non-temporal for large states, each XCD walking a contiguous range of blocks.
template <typename R, bool NT>
__global__ __launch_bounds__(256) void k_copy(cpx<R>* __restrict__ d, const cpx<R>* __restrict__ s, uint64_t size) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = xcd_tile(blockIdx.x, gridDim.x, 1) * 256 + threadIdx.x; i < size; i += stride)
    store_amp<NT>(d + i, load_amp<NT>(s + i));
}

static unsigned grid_for(uint64_t n, unsigned threads, unsigned cap = 65536u) {
  uint64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

}  // namespace brk

using namespace brk;

namespace brk {

static int set_tuning(tuning& t, const std::string& k, int64_t value) {
  const char* key = k.c_str();
  if (k == "pair_lds" || k == "diag_lds") {
    if (value < 0 || value > 65536) return fail(BRK_ERR_INVALID_ARGUMENT, "%s must be in [0, 65536]", key);
    (k == "pair_lds" ? t.pair_lds : t.diag_lds) = (int)value;
  } else if (k == "pair_remap") t.pair_remap = value != 0;
  else if (k == "diag_remap") t.diag_remap = value != 0;
  else if (k == "nt_min_bytes") {
    if (value < 0) return fail(BRK_ERR_INVALID_ARGUMENT, "nt_min_bytes must be >= 0");
    t.nt_min_bytes = (uint64_t)value;
  } else if (k == "line_bytes") {
    if (value < 1 || value > 4096 || (value & (value - 1))) return fail(BRK_ERR_INVALID_ARGUMENT, "line_bytes must be a power of two <= 4096");
    t.line_bytes = (uint64_t)value;
  } else if (k == "fused_chunked") {
    t.fused_chunked = value != 0;
  } else if (k == "block_lds") {
    if (value != 65536 && value != 32768 && value != 16384)
      return fail(BRK_ERR_INVALID_ARGUMENT, "block_lds must be 65536, 32768 or 16384");
    t.block_lds = (int)value;
  } else if (k == "adjoint_blocks") {
    t.adjoint_blocks = value != 0;
  } else if (k == "fused_chunks") {
    t.fused_chunks = value != 0;
  } else if (k == "block_run_bytes") {
    if (value < 16 || value > 4096 || (value & (value - 1)))
      return fail(BRK_ERR_INVALID_ARGUMENT, "block_run_bytes must be a power of two in [16, 4096]");
    t.block_run_bytes = (int)value;
  } else if (k == "fused_groups") {
    if (value < 1 || value > (1 << 22)) return fail(BRK_ERR_INVALID_ARGUMENT, "fused_groups must be in [1, 2^22]");
    t.fused_groups = (uint64_t)value;
  } else if (k == "fused_staged") {
    t.fused_staged = value != 0;
  } else if (k == "fused_staged_groups") {
    if (value < 1 || value > (1 << 22)) return fail(BRK_ERR_INVALID_ARGUMENT, "fused_staged_groups must be in [1, 2^22]");
    t.fused_staged_groups = (uint64_t)value;
  } else if (k == "fused_occupancy") {
    if (value != 2 && value != 4) return fail(BRK_ERR_INVALID_ARGUMENT, "fused_occupancy must be 2 or 4");
    t.fused_occupancy = (int)value;
  } else if (k == "max_blocks") {
    if (value < 1) return fail(BRK_ERR_INVALID_ARGUMENT, "max_blocks must be >= 1");
    t.max_blocks = (uint64_t)value;
  } else return fail(BRK_ERR_INVALID_ARGUMENT, "unknown tuning key '%s'", key);
  return BRK_OK;
}

}  // namespace brk

extern "C" {

const char* brk_status_string(int status) {
  switch (status) {
    case BRK_OK: return "ok";
    case BRK_ERR_INVALID_ARGUMENT: return "invalid argument";
    case BRK_ERR_QUBIT_OUT_OF_RANGE: return "qubit out of range";
    case BRK_ERR_DUPLICATE_QUBIT: return "duplicate qubit";
    case BRK_ERR_TOO_MANY_QUBITS: return "too many operated qubits";
    case BRK_ERR_UNSUPPORTED_DTYPE: return "unsupported dtype";
    case BRK_ERR_OUT_OF_MEMORY: return "out of memory";
    case BRK_ERR_HIP: return "HIP error";
    case BRK_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
  }
}

const char* brk_last_error(void) { return g_last_error.c_str(); }

const char* brk_version(void) { return "braket-mi355x 0.1 (gfx950)"; }

int brk_tune(const char* key, int64_t value) {
  if (!key) return fail(BRK_ERR_INVALID_ARGUMENT, "key is null");
  return set_tuning(get_tuning(), key, value);
}

int brk_get_tune(const char* key, int64_t* value) {
  if (!key || !value) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  const tuning& t = get_tuning();
  const std::string k = key;
  if (k == "pair_lds") *value = t.pair_lds;
  else if (k == "diag_lds") *value = t.diag_lds;
  else if (k == "pair_remap") *value = t.pair_remap;
  else if (k == "diag_remap") *value = t.diag_remap;
  else if (k == "nt_min_bytes") *value = (int64_t)t.nt_min_bytes;
  else if (k == "max_blocks") *value = (int64_t)t.max_blocks;
  else if (k == "line_bytes") *value = (int64_t)t.line_bytes;
  else if (k == "fused_groups") *value = (int64_t)t.fused_groups;
  else if (k == "fused_chunked") *value = t.fused_chunked;
  else if (k == "block_lds") *value = t.block_lds;
  else if (k == "block_run_bytes") *value = t.block_run_bytes;
  else if (k == "adjoint_blocks") *value = t.adjoint_blocks;
  else if (k == "fused_chunks") *value = t.fused_chunks;
  else if (k == "fused_occupancy") *value = t.fused_occupancy;
  else if (k == "fused_staged") *value = t.fused_staged;
  else if (k == "fused_staged_groups") *value = (int64_t)t.fused_staged_groups;
  else return fail(BRK_ERR_INVALID_ARGUMENT, "unknown tuning key '%s'", key);
  return BRK_OK;
}

int brk_device_count(int* count) {
  if (!count) return fail(BRK_ERR_INVALID_ARGUMENT, "count is null");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *count = 0;
    return fail(BRK_ERR_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  *count = c;
  return BRK_OK;
}

int brk_mem_info(size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  BRK_HIP_CHECK(hipMemGetInfo(free_bytes, total_bytes));
  return BRK_OK;
}

int brk_stream_create(void** stream) {
  if (!stream) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  hipStream_t s = nullptr;
  BRK_HIP_CHECK(hipStreamCreate(&s));  // default priority
  *stream = s;
  return BRK_OK;
}

int brk_stream_destroy(void* stream) {
  if (stream) BRK_HIP_CHECK(hipStreamDestroy(as_stream(stream)));
  return BRK_OK;
}

int brk_event_create(void** event) {
  if (!event) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  hipEvent_t e = nullptr;
  BRK_HIP_CHECK(hipEventCreate(&e));
  *event = e;
  return BRK_OK;
}

int brk_event_destroy(void* event) {
  if (event) BRK_HIP_CHECK(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
  return BRK_OK;
}

int brk_event_record(void* event, void* stream) {
  if (!event) return fail(BRK_ERR_INVALID_ARGUMENT, "null event");
  BRK_HIP_CHECK(hipEventRecord(reinterpret_cast<hipEvent_t>(event), as_stream(stream)));
  return BRK_OK;
}

int brk_event_synchronize(void* event) {
  if (!event) return fail(BRK_ERR_INVALID_ARGUMENT, "null event");
  BRK_HIP_CHECK(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)));
  return BRK_OK;
}

int brk_event_elapsed_ms(void* start, void* end, float* ms) {
  if (!start || !end || !ms) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  BRK_HIP_CHECK(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(end)));
  return BRK_OK;
}

int brk_device_name(char* buffer, size_t size) {
  if (!buffer || size == 0) return fail(BRK_ERR_INVALID_ARGUMENT, "empty buffer");
  int dev = 0;
  BRK_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t props;
  BRK_HIP_CHECK(hipGetDeviceProperties(&props, dev));
  std::snprintf(buffer, size, "%s (%s, %d CUs)", props.name[0] ? props.name : "AMD GPU", props.gcnArchName,
                props.multiProcessorCount);
  return BRK_OK;
}

int brk_malloc(void** device_ptr, size_t bytes) {
  if (!device_ptr) return fail(BRK_ERR_INVALID_ARGUMENT, "device_ptr is null");
  *device_ptr = nullptr;
  hipError_t e = hipMalloc(device_ptr, bytes);
  if (e != hipSuccess) {
    *device_ptr = nullptr;
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? BRK_ERR_OUT_OF_MEMORY : BRK_ERR_HIP, "hipMalloc(%zu): %s", bytes,
                hipGetErrorString(e));
  }
  return BRK_OK;
}

int brk_free(void* device_ptr) {
  BRK_HIP_CHECK(hipFree(device_ptr));
  return BRK_OK;
}

int brk_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
  if ((!dst || !src) && bytes) return fail(BRK_ERR_INVALID_ARGUMENT, "null pointer");
  BRK_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
  BRK_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  return BRK_OK;
}

int brk_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
  if ((!dst || !src) && bytes) return fail(BRK_ERR_INVALID_ARGUMENT, "null pointer");
  BRK_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
  BRK_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  return BRK_OK;
}

int brk_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
  if ((!dst || !src) && bytes) return fail(BRK_ERR_INVALID_ARGUMENT, "null pointer");
  BRK_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
  return BRK_OK;
}

int brk_stream_synchronize(void* stream) {
  BRK_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  return BRK_OK;
}

int brk_set_basis_state(void* first, uint64_t size, int dtype, uint64_t basis_index, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (basis_index >= size)
    return fail(BRK_ERR_INVALID_ARGUMENT, "basis index %llu >= size %llu", (unsigned long long)basis_index,
                (unsigned long long)size);
  unsigned grid = grid_for(size, 256);
  if (dtype == BRK_C128)
    hipLaunchKernelGGL(k_set_basis<double>, dim3(grid), dim3(256), 0, as_stream(stream), (cpx<double>*)first,
                       size, basis_index);
  else
    hipLaunchKernelGGL(k_set_basis<float>, dim3(grid), dim3(256), 0, as_stream(stream), (cpx<float>*)first,
                       size, basis_index);
  return check_launch("set_basis_state");
}

int brk_fill_modular(void* first, uint64_t size, int dtype, const double scales[2], const uint64_t params[6],
                     void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (!scales || !params) return fail(BRK_ERR_INVALID_ARGUMENT, "null scales/params");
  modular_params p{scales[0], scales[1], params[0], params[1], params[2], params[3], params[4], params[5]};
  unsigned grid = grid_for(size, 256);
  if (dtype == BRK_C128)
    hipLaunchKernelGGL(k_fill_modular<double>, dim3(grid), dim3(256), 0, as_stream(stream),
                       (cpx<double>*)first, size, p);
  else
    hipLaunchKernelGGL(k_fill_modular<float>, dim3(grid), dim3(256), 0, as_stream(stream), (cpx<float>*)first,
                       size, p);
  return check_launch("fill_modular");
}

int brk_fill_hash(void* first, uint64_t size, int dtype, uint64_t seed, double scale, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  unsigned grid = grid_for(size, 256);
  if (dtype == BRK_C128)
    hipLaunchKernelGGL(k_fill_hash<double>, dim3(grid), dim3(256), 0, as_stream(stream), (cpx<double>*)first,
                       size, seed, scale);
  else
    hipLaunchKernelGGL(k_fill_hash<float>, dim3(grid), dim3(256), 0, as_stream(stream), (cpx<float>*)first,
                       size, seed, scale);
  return check_launch("fill_hash");
}

int brk_convert(void* dst, int dtype_dst, const void* src, int dtype_src, uint64_t size, void* stream) {
  unsigned nq;
  if (int st = check_state(dst, size, dtype_dst, &nq)) return st;
  if (int st = check_state(src, size, dtype_src, &nq)) return st;
  unsigned grid = grid_for(size, 256);
  hipStream_t s = as_stream(stream);
  if (dtype_dst == BRK_C64 && dtype_src == BRK_C128)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid), dim3(256), 0, s, (cpx<float>*)dst,
                       (const cpx<double>*)src, size);
  else if (dtype_dst == BRK_C128 && dtype_src == BRK_C64)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid), dim3(256), 0, s, (cpx<double>*)dst,
                       (const cpx<float>*)src, size);
  else if (dtype_dst == BRK_C128)
    hipLaunchKernelGGL((k_convert<double, double>), dim3(grid), dim3(256), 0, s, (cpx<double>*)dst,
                       (const cpx<double>*)src, size);
  else
    hipLaunchKernelGGL((k_convert<float, float>), dim3(grid), dim3(256), 0, s, (cpx<float>*)dst,
                       (const cpx<float>*)src, size);
  return check_launch("convert");
}

int brk_copy(void* dst, const void* src, uint64_t size, int dtype, void* stream) {
  unsigned nq;
  if (int st = check_state(dst, size, dtype, &nq)) return st;
  if (int st = check_state(src, size, dtype, &nq)) return st;
  const unsigned blocks = (unsigned)std::min<uint64_t>((size + 255) / 256, (1ull << 24) - 1);
  const bool nt = size * amp_bytes(dtype) >= get_tuning().nt_min_bytes;
  hipStream_t st = as_stream(stream);
  if (dtype == BRK_C128) {
    auto* d = (cpx<double>*)dst;
    auto* s = (const cpx<double>*)src;
    if (nt) hipLaunchKernelGGL((k_copy<double, true>), dim3(blocks), dim3(256), 0, st, d, s, size);
    else hipLaunchKernelGGL((k_copy<double, false>), dim3(blocks), dim3(256), 0, st, d, s, size);
  } else {
    auto* d = (cpx<float>*)dst;
    auto* s = (const cpx<float>*)src;
    if (nt) hipLaunchKernelGGL((k_copy<float, true>), dim3(blocks), dim3(256), 0, st, d, s, size);
    else hipLaunchKernelGGL((k_copy<float, false>), dim3(blocks), dim3(256), 0, st, d, s, size);
  }
  return check_launch("copy");
}

}  // extern "C"

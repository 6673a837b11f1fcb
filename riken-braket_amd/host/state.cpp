// Device states with their stream (brk_state_*): allocation, |basis> preparation, whole-state
// copies and the norm, over the primitives of the C ABI.  Plays the role of the storage the
// reference's gate functions run over (ket's std::vector; bra::nompi_state's data_ and its
// constructor, bra/include/bra/nompi_state.hpp:155-165).
#include <hip/hip_runtime_api.h>

#include <new>

#include "braket_hip.h"

namespace brk {
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
}
using brk::fail;

struct brk_state {
  void* first = nullptr;
  uint64_t size = 0;
  int dtype = BRK_C128;
  void* stream = nullptr;
};

static size_t state_bytes(const brk_state* s) { return (size_t)s->size * (s->dtype == BRK_C128 ? 16 : 8); }

extern "C" {

int brk_state_create(unsigned num_qubits, int dtype, void* stream, brk_state** out) {
  if (!out) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (dtype != BRK_C128 && dtype != BRK_C64) return fail(BRK_ERR_UNSUPPORTED_DTYPE, "unknown dtype %d", dtype);
  if (num_qubits > 40) return fail(BRK_ERR_INVALID_ARGUMENT, "%u qubits is beyond one device", num_qubits);
  auto* s = new (std::nothrow) brk_state;
  if (!s) return fail(BRK_ERR_OUT_OF_MEMORY, "cannot allocate the state handle");
  s->size = 1ull << num_qubits;
  s->dtype = dtype;
  s->stream = stream;
  if (hipMalloc(&s->first, state_bytes(s)) != hipSuccess) {
    delete s;
    return fail(BRK_ERR_OUT_OF_MEMORY, "cannot allocate %u-qubit state", num_qubits);
  }
  if (int st = brk_set_basis_state(s->first, s->size, dtype, 0, stream)) {
    (void)hipFree(s->first);
    delete s;
    return st;
  }
  *out = s;
  return BRK_OK;
}

int brk_state_destroy(brk_state* state) {
  if (!state) return BRK_OK;
  int st = BRK_OK;
  if (hipStreamSynchronize(reinterpret_cast<hipStream_t>(state->stream)) != hipSuccess)
    st = fail(BRK_ERR_HIP, "hipStreamSynchronize failed");
  if (hipFree(state->first) != hipSuccess && !st) st = fail(BRK_ERR_HIP, "hipFree failed");
  delete state;
  return st;
}

int brk_state_data(const brk_state* state, void** first, uint64_t* size, int* dtype, void** stream) {
  if (!state) return fail(BRK_ERR_INVALID_ARGUMENT, "null state");
  if (first) *first = state->first;
  if (size) *size = state->size;
  if (dtype) *dtype = state->dtype;
  if (stream) *stream = state->stream;
  return BRK_OK;
}

int brk_state_set_basis(brk_state* state, uint64_t basis_index) {
  if (!state) return fail(BRK_ERR_INVALID_ARGUMENT, "null state");
  return brk_set_basis_state(state->first, state->size, state->dtype, basis_index, state->stream);
}

int brk_state_upload(brk_state* state, const void* host) {
  if (!state || !host) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  if (int st = brk_memcpy_htod(state->first, host, state_bytes(state), state->stream)) return st;
  return brk_stream_synchronize(state->stream);
}

int brk_state_download(const brk_state* state, void* host) {
  if (!state || !host) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  if (int st = brk_memcpy_dtoh(host, state->first, state_bytes(state), state->stream)) return st;
  return brk_stream_synchronize(state->stream);
}

int brk_state_norm2(const brk_state* state, double* result) {
  if (!state || !result) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  return brk_norm2(state->first, state->size, state->dtype, result, state->stream);
}

}  // extern "C"

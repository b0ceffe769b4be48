// Citizen-state stream on gfx950: one hour's per-agent records formatted as JSON text on
// the device.
//
// The reference buffers (citizen id, state letter, location) for every agent each hour
// (engine/src/models/events/citizen_state.rs:26-68, filled from allocation_map.rs:122-124)
// and publishes the buffer with serde_json as one message per hour
// (listeners/events_kafka_producer.rs:62-67, 90-99). Here the message body, the records
// joined by commas, is written by three kernels:
//   k_record_len   : per block, the total length of its agents' records;
//   k_scan_offsets : exclusive prefix of the block totals (one block), and the grand total;
//   k_record_write : each block builds its text in LDS, then stores it with coalesced
//                    byte stores at its offset.
// A record is {"citizen_id":"<id>","state":"<s|e|i|r|d>","location":{"x":<x>,"y":<y>}},
// preceded by a comma for every agent but the first. Agent ids are the dense deterministic
// ids, written as a JSON string where the reference writes a UUID string. A simulation
// that did not step the requested hour writes a total of 0.
#include <hip/hip_runtime.h>

#include <mutex>

#include "stream.hpp"

namespace epi {

// Scratch of the no-allocation formatter (epi_citizen_states_text_noalloc): the per-block
// offsets, one entry per 256-agent block, for at most kNoAllocMaxAgents agents (512 KiB on
// each GPU the library runs on, held until the process ends). Host code never passes this
// name to a kernel: it is a host-side stand-in, not a device address. noalloc_scratch()
// resolves the device's own copy once per device with hipGetSymbolAddress, and that
// address is what the kernels get.
static __device__ __align__(256) unsigned long long scratch[kNoAllocMaxBlocks];

namespace {

constexpr uint32_t kStreamBlock = 256;
static_assert(kNoAllocMaxBlocks * kStreamBlock == kNoAllocMaxAgents, "no-allocation scratch size");

// {"citizen_id":"  ","state":"  ?  ","location":{"x":  ,"y":  }}
constexpr char kP0[] = "{\"citizen_id\":\"";
constexpr char kP1[] = "\",\"state\":\"";
constexpr char kP2[] = "\",\"location\":{\"x\":";
constexpr char kP3[] = ",\"y\":";
constexpr char kP4[] = "}}";
constexpr uint32_t kFixed = (sizeof kP0 - 1) + (sizeof kP1 - 1) + 1 + (sizeof kP2 - 1) + (sizeof kP3 - 1) +
                            (sizeof kP4 - 1);
static_assert(1 + kFixed + 10 + 5 + 5 == kMaxRecord, "record bound");

__device__ __forceinline__ uint32_t ndigits(uint32_t v) {
  uint32_t d = 1;
  while (v >= 10u) {
    v /= 10u;
    ++d;
  }
  return d;
}

__device__ __forceinline__ uint32_t record_len(uint32_t id, uint32_t p) {
  return (id ? 1u : 0u) + kFixed + ndigits(id) + ndigits(p & 0xFFFFu) + ndigits(p >> 16);
}

__device__ __forceinline__ char state_letter(uint32_t w) {
  switch ((w >> B_STATE) & M_STATE) {  // CitizenState::state_str (citizen_state.rs:35-43)
    case ST_S: return 's';
    case ST_E: return 'e';
    case ST_I: return 'i';
    case ST_R: return 'r';
    default: return 'd';
  }
}

template <size_t K>
__device__ __forceinline__ char* put(char* o, const char (&s)[K]) {
#pragma unroll
  for (size_t k = 0; k + 1 < K; ++k) o[k] = s[k];
  return o + (K - 1);
}

__device__ __forceinline__ char* put_uint(char* o, uint32_t v) {
  const uint32_t nd = ndigits(v);
  for (uint32_t k = nd; k > 0; --k) {
    o[k - 1] = static_cast<char>('0' + v % 10u);
    v /= 10u;
  }
  return o + nd;
}

__device__ __forceinline__ bool stepped(const StreamArgs& a) { return a.state->hour == a.hour; }

__global__ __launch_bounds__(kStreamBlock) void k_record_len(StreamArgs a) {
  __shared__ unsigned long long part[kStreamBlock / 64];
  const uint32_t i = blockIdx.x * kStreamBlock + threadIdx.x;
  const bool on = stepped(a);
  unsigned long long len = (on && i < a.n) ? record_len(i, a.pos[i]) : 0ull;
  for (int off = 32; off > 0; off >>= 1) len += __shfl_xor(len, off, 64);
  if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = len;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (uint32_t w = 0; w < kStreamBlock / 64; ++w) s += part[w];
    a.block_off[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(1024) void k_scan_offsets(StreamArgs a, uint32_t n_blocks) {
  __shared__ unsigned long long part[1024];
  const uint32_t per = (n_blocks + 1023u) / 1024u;
  const uint32_t b = threadIdx.x * per, e = min(n_blocks, b + per);
  unsigned long long sum = 0;
  for (uint32_t k = b; k < e; ++k) sum += a.block_off[k];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < 1024u; off <<= 1) {
    const unsigned long long t = threadIdx.x >= off ? part[threadIdx.x - off] : 0ull;
    __syncthreads();
    part[threadIdx.x] += t;
    __syncthreads();
  }
  unsigned long long run = threadIdx.x ? part[threadIdx.x - 1] : 0ull;
  for (uint32_t k = b; k < e; ++k) {
    const unsigned long long x = a.block_off[k];
    a.block_off[k] = run;
    run += x;
  }
  if (threadIdx.x == 1023) *a.total = part[1023];
}

__global__ __launch_bounds__(kStreamBlock) void k_record_write(StreamArgs a) {
  __shared__ char text[kStreamBlock * kMaxRecord];
  __shared__ uint32_t off[kStreamBlock];
  if (!stepped(a)) return;  // uniform: the block wrote nothing
  const uint32_t i = blockIdx.x * kStreamBlock + threadIdx.x;
  const bool valid = i < a.n;
  const uint32_t p = valid ? a.pos[i] : 0u;
  const uint32_t len = valid ? record_len(i, p) : 0u;
  // inclusive scan of the lengths over the block
  off[threadIdx.x] = len;
  __syncthreads();
  for (uint32_t d = 1; d < kStreamBlock; d <<= 1) {
    const uint32_t t = threadIdx.x >= d ? off[threadIdx.x - d] : 0u;
    __syncthreads();
    off[threadIdx.x] += t;
    __syncthreads();
  }
  const uint32_t end = off[kStreamBlock - 1];
  if (valid) {
    char* o = text + (off[threadIdx.x] - len);
    if (i) *o++ = ',';
    o = put(o, kP0);
    o = put_uint(o, i);
    o = put(o, kP1);
    *o++ = state_letter(a.st[i]);
    o = put(o, kP2);
    o = put_uint(o, p & 0xFFFFu);
    o = put(o, kP3);
    o = put_uint(o, p >> 16);
    put(o, kP4);
  }
  __syncthreads();
  char* dst = a.text + a.block_off[blockIdx.x];
  for (uint32_t k = threadIdx.x; k < end; k += kStreamBlock) dst[k] = text[k];
}

}  // namespace

uint32_t max_record_bytes(uint32_t n, uint32_t stride, uint32_t height) {
  uint32_t nd_id = 1, nd_x = 1, nd_y = 1;
  for (uint32_t v = n ? n - 1 : 0; v >= 10u; v /= 10u) ++nd_id;
  for (uint32_t v = stride; v >= 10u; v /= 10u) ++nd_x;
  for (uint32_t v = height; v >= 10u; v /= 10u) ++nd_y;
  return 1u + kFixed + nd_id + nd_x + nd_y;
}

uint32_t stream_blocks(uint32_t n) { return (n + kStreamBlock - 1) / kStreamBlock; }

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(unsigned long long** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {  // the current device's copy of the array, resolved once
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[dev] = static_cast<unsigned long long*>(p);
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

hipError_t launch_citizen_states(const StreamArgs& a, hipStream_t stream) {
  const uint32_t blocks = stream_blocks(a.n);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_record_len, dim3(blocks), dim3(kStreamBlock), 0, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, stream, a, blocks);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_record_write, dim3(blocks), dim3(kStreamBlock), 0, stream, a);
  return hipGetLastError();
}

}  // namespace epi

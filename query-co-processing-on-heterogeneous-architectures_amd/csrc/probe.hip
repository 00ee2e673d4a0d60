// probe.hip -- read-bandwidth probe of the memory level that holds a buffer of a
// given size (measurement support for roofline figures).
//
// Every wave streams 64-float chunks with one dword load per lane, eight
// chunks in flight, the same load form as the search kernels' leaf scan.
// Wave w starts at chunk group w and advances by the number of waves, wrapping
// around the window, for `rounds` sweeps of the whole grid; a window of a few
// MiB stays in each XCD's L2 (every XCD caches its own copy), one of a few
// hundred MiB in the Infinity Cache, and one of several GiB streams from HBM.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/qcp.h"
#include "qcp_device.h"

namespace qcp {

__global__ __launch_bounds__(256) void k_probe_read(const float* __restrict__ buf,
                                                    uint64_t groups, uint32_t rounds,
                                                    float* __restrict__ sink) {
  const uint32_t lane = lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / kWave;
  float acc = 0.0f;
  uint64_t g = wave % groups;
  const uint64_t step = nwaves % groups;
  for (uint32_t r = 0; r < rounds; ++r) {
    const float* p = buf + g * (8 * kWave) + lane;
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[u * kWave];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
    g += step;
    if (g >= groups) g -= groups;
  }
  // never true for the zero-filled buffer; keeps the loads live
  if (acc != 0.0f) sink[0] = acc;
}

}  // namespace qcp

extern "C" qcp_status qcp_probe_read_gbps(uint32_t device, uint64_t window_bytes,
                                          uint32_t rounds, double* gbps, double* ms) {
  using namespace qcp;
  if (!gbps || window_bytes < 8 * kWave * 4 || rounds == 0) return QCP_EINVAL;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) return QCP_ENODEV;
  if ((int)device >= cnt || hipSetDevice((int)device) != hipSuccess) return QCP_EINVAL;
  int cus = 0;
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, (int)device);
  if (cus <= 0) cus = 256;
  const uint64_t groups = window_bytes / (8 * kWave * 4);
  float* buf = nullptr;
  float* sink = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  qcp_status st = QCP_OK;
  if (hipMalloc((void**)&buf, groups * 8 * kWave * 4) != hipSuccess ||
      hipMalloc((void**)&sink, 4) != hipSuccess) {
    st = QCP_ENOMEM;
  } else if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
             hipMemsetAsync(buf, 0, groups * 8 * kWave * 4, s) != hipSuccess) {
    st = QCP_EHIP;
  } else {
    // 8 workgroups of 4 waves per CU: enough loads in flight on every CU
    const unsigned grid = (unsigned)cus * 8;
    k_probe_read<<<grid, 256, 0, s>>>(buf, groups, 2, sink);  // warm the caches
    hipEventRecord(e0, s);
    k_probe_read<<<grid, 256, 0, s>>>(buf, groups, rounds, sink);
    hipEventRecord(e1, s);
    if (hipStreamSynchronize(s) != hipSuccess) {
      st = QCP_EHIP;
    } else {
      float t = 0.0f;
      hipEventElapsedTime(&t, e0, e1);
      // 4 waves per workgroup, each reading 8 chunks x 64 lanes x 4 bytes per round
      const double bytes = (double)grid * 4 * rounds * 8 * kWave * 4;
      *gbps = bytes / (t * 1e-3) / 1e9;
      if (ms) *ms = t;
    }
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (s) hipStreamDestroy(s);
  if (buf) hipFree(buf);
  if (sink) hipFree(sink);
  return st;
}

// Measured device rates (peaks.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace imops {

struct MeasuredPeaks {
  double copy_GBps = 0;              // HBM copy: (read + written bytes) / time
  double valu_per_clk_per_simd = 0;  // wave64 FP32 mul/add instructions per clock per SIMD (at the rated clock)
  double valu_nonfused_TFLOPs = 0;   // the same stream as one flop per lane per instruction
  double mfma_f16_TFLOPs = 0;        // v_mfma_f32_16x16x32_f16, dense
  double clock_GHz = 0;              // rated engine clock used above
  int compute_units = 0;
};

// buf_a, buf_b: device buffers of `bytes` each (the copy's source and destination;
// buf_b also takes the rate kernels' never-written sink).
hipError_t measure_peaks(hipStream_t s, void* buf_a, void* buf_b, size_t bytes, int device, MeasuredPeaks* out);

}  // namespace imops

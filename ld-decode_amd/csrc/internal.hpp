// Handle layout and error plumbing shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include <vector>

#include "../../include/ldgpu.h"

struct ld_handle {
  int device = 0;
  int blocklen = 0;          // N (real samples per block)
  int cut_begin = 1024;      // RFDecode.blockcut        core.py:321
  int cut_end = 32;          // RFDecode.blockcut_end    core.py:402 (F05_offset)
  float2* twiddle = nullptr; // exp(-2*pi*i*t/N), t < N
  float* sinc_lut = nullptr; // (65536+1) x 16 float32
  int sinc_folded = 0;
  int notch_lo = 0, notch_hi = 0;  // spur notch window (bins), lo == hi: off        // row p == row (65536-p) reversed for 0 < p < 65536
  float2* filt[LD_NFILT] = {};
  // each table times exp(+2 pi i k / N): its c2r output advanced by one
  // sample, used where a c2r output's offset is odd so that the output
  // pairs stay 8-byte aligned
  float2* filt_rolled[LD_NFILT] = {};
  double filt_dc[LD_NFILT] = {};
  int filt_shift[LD_NFILT] = {};
  float2* audio_filt[2] = {};
  float2* audio2_filt[2] = {};   // stage-2 half spectra, scaled for the c2r transform
  double audio2_center[2] = {};
  int audio_nbins = 0;
  int audio_lowbin[2] = {};
  double audio_rate[2] = {};
  double audio_low_freq[2] = {};
  double fs = 40e6;
  double clip_lo = 1.5e6, clip_hi = 30e6, dc_offset = 0.0;
  int video_rot = 0;
  bool params_set = false;
  // tables replaced by a later upload; launches queued before the upload
  // still read them, so they are freed only after a device synchronisation
  std::vector<void*> retired;
  // this device's address of the no-allocation entry point's scratch array
  void* scratch = nullptr;
  // split-BF16 MFMA transform variant: DFT-64 fragment table and twiddles
  void* mfma_g = nullptr;
  void* mfma_tw = nullptr;
  int fft_variant = 0;       // LD_FFT_FP32 or LD_FFT_MFMA_BF16X3
};

namespace ldgpu {

void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

#define LD_HIP(call)                                    \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return ldgpu::hip_fail(e_, #call); \
  } while (0)

// Upload `bytes` of host data as a device table into *slot.  A table that
// already exists is never overwritten in place: the data goes to a fresh
// buffer (synchronous copy, read by no launch yet) and the old one is
// retired, so kernels queued earlier on any stream keep reading the table
// they were launched with.
int replace_table(ld_handle* h, void** slot, const void* host, size_t bytes);
// Free retired tables (synchronises the device first).
void free_retired(ld_handle* h);

constexpr int kSincTaps = 16;
constexpr int kSincPhases = 1 << 16;

}  // namespace ldgpu

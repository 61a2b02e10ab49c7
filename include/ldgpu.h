/* ldgpu.h -- C ABI of the MI355X RF demodulation + TBC library (libldgpu.so).
 *
 * Drop-in GPU path for two call sites of ld-decode (Python; no FFI exists in
 * the reference, so the boundary mirrors its Python interfaces):
 *   - RFDecode.demodblock(data, mtf_level, fftdata, cut)  lddecode/core.py:747-879
 *     together with the DemodCache block stitching        lddecode/core.py:1204-1413
 *   - Field.downscale -> computewow_scaled -> scale_field -> hz_to_output
 *                                        lddecode/core.py:2591-2722, utils.py:127-183,
 *                                        utils.py:1061-1075
 *   - Field.refine_linelocs_hsync        lddecode/core.py:2484-2551
 *
 * Conventions
 *   - Every function returns 0 on success and a negative LD_E* code on
 *     failure; nothing throws across the ABI.  ld_last_error() returns the
 *     message of the calling thread's most recent failure.
 *   - Sample buffers are caller-owned DEVICE pointers (e.g. torch tensors'
 *     data_ptr()); filter tables and line locations passed as `host_*` are host
 *     pointers.  `stream` is a hipStream_t (NULL = default stream).
 *   - Handles are not internally locked: use one handle per host thread (the
 *     reference gives each DemodCache worker thread its own RFDecode,
 *     core.py:1165).  Calls that take a `stream` only enqueue work on it.
 *   - The table setters (ld_set_filter, ld_set_audio, ld_set_audio2,
 *     ld_set_sinc_lut) are synchronous host calls that never write a table
 *     a queued launch may read: each upload goes to a fresh device buffer,
 *     launches queued earlier (on any stream) keep the table they were
 *     given, and launches queued after the call use the new one.  Replaced
 *     tables are freed after a device synchronisation once 32 have piled up,
 *     and by ld_destroy.  The scalar setters only change what later launches
 *     are given.
 */
#ifndef LDGPU_H
#define LDGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDGPU_ABI_VERSION 1

enum {
  LD_OK = 0,
  LD_EINVAL = -1,   /* bad argument (null pointer, size, id) */
  LD_ERANGE = -2,   /* a block or tap would read outside the input buffer */
  LD_EHIP = -3,     /* HIP runtime error */
  LD_ESTATE = -4    /* required tables not uploaded yet */
};

/* Filter table ids (RFDecode.Filters names, core.py:544-679). */
enum {
  LD_FILT_RFHPF = 0,    /* Frfhpf                                  core.py:554-555 */
  LD_FILT_RFVIDEO = 1,  /* RFVideo (analytic mask folded) * MTF^lvl core.py:569-612, :800-803 */
  LD_FILT_VIDEO = 2,    /* FVideo (LPF * deemph^s * GD all-pass)   core.py:632-643 */
  LD_FILT_VIDEO05 = 3,  /* FVideo05 (LPF * deemph * firwin65)      core.py:648-652 */
  LD_FILT_BURST = 4,    /* FVideoBurst (LPF * deemph * firwin81)   core.py:654-658 */
  LD_FILT_PILOT = 5,    /* FVideoPilot, PAL only                   core.py:660-669 */
  LD_FILT_EFM = 6,      /* Fefm (Hermitian part), digital audio    core.py:420-465 */
  LD_NFILT = 7
};

/* Input sample formats: int16 (10-bit codes << 6, utils.py:388), float32,
 * and the packed 10-bit capture format (4 samples in 5 bytes, big-endian
 * bit order, utils.py:356-390); input_len counts samples in every format,
 * and a packed buffer holds (input_len + 3) / 4 * 5 bytes and must be
 * readable up to the next multiple of 4 bytes (device allocations are). */
enum { LD_IN_INT16 = 0, LD_IN_FLOAT32 = 1, LD_IN_PACKED10 = 2 };

typedef struct ld_handle ld_handle;

/* Library ABI version (LDGPU_ABI_VERSION). */
int ld_abi_version(void);

/* Message of the calling thread's last failure ("" if none). */
const char* ld_last_error(void);

/* Create a handle on HIP device `device` for blocks of `blocklen` samples
 * (RFDecode(blocklen=...), core.py:286; 16384, 32768 or 65536).  Uploads the FFT
 * twiddle table.
 * Replaces RFDecode.__init__'s table set-up, core.py:284-403. */
int ld_create(int device, int blocklen, ld_handle** out);
void ld_destroy(ld_handle* h);

/* Upload one filter as its half spectrum, bins 0..blocklen/2, interleaved
 * (re, im) float64 pairs; stored as float32 after multiplying by `scale`.
 * `dc_gain` is the filter's true bin-0 response (added back as dc_gain times
 * the DC offset removed before the second FFT); `shift` is the output roll in
 * samples (np.roll(x, -shift), core.py:814, :817).
 * Replaces the table build of RFDecode.computefilters, core.py:405-420. */
int ld_set_filter(ld_handle* h, int filter_id, const double* host_reim, int nbins,
                  double scale, double dc_gain, int shift);

/* Audio stage-1 channel (0 = left, 1 = right): `nbins`-bin sliced filter
 * filt1 (core.py:716-718) as interleaved float64 pairs, the slice's low bin,
 * its decimated rate a1_freq and the frequency of its low bin (core.py:707-714). */
int ld_set_audio(ld_handle* h, int channel, const double* host_reim, int nbins, int lowbin,
                 double a1_freq, double low_freq);

/* Audio stage-2 channel (0 = left, 1 = right): the filter audio2_filter
 * (20 kHz low-pass times de-emphasis, core.py:727-734) as its half spectrum,
 * bins 0..blocklen/2, interleaved float64 pairs, and the channel's carrier
 * frequency (audio_lfreq / audio_rfreq) removed before it and added after. */
int ld_set_audio2(ld_handle* h, int channel, const double* host_reim, int nbins, double center_freq);

/* Scalar demodulation parameters: sample rate (Hz), the clip applied to
 * demod_raw before the second FFT (core.py:809), the constant subtracted from
 * the clipped signal before that FFT (float32 range protection; added back
 * through each filter's dc_gain), and video_rot (core.py:765-772, :1032-1034). */
int ld_set_demod_params(ld_handle* h, double fs_hz, double clip_lo, double clip_hi,
                        double dc_offset, int video_rot);

/* PAL LD-V4300D spur notch (core.py:774-798, extra_options
 * "PAL_V4300D_NotchFilter"): within spectrum bins [lo, hi) of every block,
 * bins whose magnitude exceeds the window's mean + 3 standard deviations are
 * zeroed together with both neighbours before the analytic-signal and audio
 * stage-1 filters (the rfhpf output keeps the unnotched spectrum).  The
 * reference's window is [int(N*8.42/f_MHz), int(1 + N*8.6/f_MHz)).  lo = hi = 0
 * disables it; otherwise 1 <= lo < hi < N/2 and hi - lo <= 508. */
int ld_set_notch(ld_handle* h, int lo, int hi);

/* Output channels of ld_demod_blocks.  A NULL pointer skips that channel.
 * Block b writes its samples at ptr + b * video_stride (video and rfhpf) or
 * ptr + b * audio_stride (audio). */
typedef struct ld_demod_outputs {
  float* rfhpf;
  float* demod;
  float* demod_raw;
  float* demod_05;
  float* demod_burst;
  float* demod_pilot;
  float* audio_left;
  float* audio_right;
  int64_t video_stride;
  int64_t audio_stride;
  /* EFM front-end (core.py:847-851): int16(clip(irFFT(X * Fefm))), cut and
   * strided like the video outputs; needs LD_FILT_EFM; NULL: skip */
  int16_t* efm;
  /* findpulses' comparison fused into the demod_05 output: word w holds
   * demod_05[32 w + i] <= below_high (compared in double, as numba compares
   * the float32 buffer with its float64 threshold) for i < 32, paired: bit
   * j < 16 is sample 32 w + 2 j, bit 16 + j sample 32 w + 2 j + 1.  Block b
   * fills words (b * video_stride) / 32 onwards; bits past its output are
   * 0.  Needs demod_05, video_stride % 32 == 0 and demod_05's output offset
   * (blockcut + its filter shift) % 32 == 0; NULL: skip.  Input of
   * ld_find_pulses_bits. */
  uint32_t* demod_05_below;
  double below_high;
} ld_demod_outputs;

/* Demodulate `nblocks` blocks; block b reads input[first + b*step .. + blocklen).
 * cut != 0: video outputs hold samples [1024, blocklen-32) of each block
 * (blocklen-1056 values), audio [32, nbins-1) (RFDecode.demodblock cut=True,
 * core.py:839-875); with step = blocklen-1056 and strides equal to the cut
 * lengths this is DemodCache's stitched field buffer (core.py:1392-1403).
 * cut == 0: full blocklen / nbins outputs.  rfhpf is always cut
 * [1024-video_rot, blocklen-32-video_rot) (core.py:769-772).
 * `workspace` is device memory of at least ld_demod_workspace_bytes(h, nblocks).
 * Replaces RFDecode.demodblock(data=..., mtf_level, cut), core.py:747-879 (the
 * MTF level is folded into LD_FILT_RFVIDEO by the host). */
size_t ld_demod_workspace_bytes(const ld_handle* h, int nblocks);

/* Unpack samples [first, first + n) of a packed 10-bit capture (DEVICE
 * bytes, packed_bytes long) to int16 (code - 512) << 6 in DEVICE `out`:
 * unpack_data_4_40 / load_packed_data_4_40 (utils.py:371-410), whose
 * `sample` is `first` here.  LD_ERANGE when the samples pass the buffer. */
int ld_unpack_10bit(ld_handle* h, const uint8_t* packed, int64_t packed_bytes, int64_t first, int64_t n,
                    int16_t* out, void* stream);
int ld_demod_blocks(ld_handle* h, const void* input, int input_dtype, int64_t input_len,
                    int64_t first, int64_t step, int nblocks, int cut,
                    const ld_demod_outputs* out, void* workspace, size_t workspace_bytes,
                    void* stream);

/* Transform variants of the demodulator (ld_set_fft_variant).
 * LD_FFT_FP32: FP32 Stockham passes (radix 16 x 16 x 16 x 4 for the 16384-point
 * complex transforms of 32768-sample blocks), the default.
 * LD_FFT_MFMA_BF16X3: the two radix-64 stages of 16384 = 64 x 64 x 4 as real
 * matrix products on the matrix cores with both operands split into bf16
 * hi + lo parts (hi*hi + hi*lo + lo*hi, FP32 accumulation); twiddles and the
 * radix-4 stage stay FP32.  About 2^-17 relative error per stage.
 * LD_FFT_MFMA_BF16X6: as LD_FFT_MFMA_BF16X3 with three bf16 parts per
 * operand (hi + mid + lo, exact for FP32 data) and the six products of
 * order >= 2^-16: FP32-level error at twice the matrix-core work.
 * (16384-sample blocks: 8192 = 64 x 64 x 2; 65536: each 16384-point half.) */
enum { LD_FFT_FP32 = 0, LD_FFT_MFMA_BF16X3 = 1, LD_FFT_MFMA_BF16X6 = 2 };

/* Select the transform variant of later ld_demod_blocks / ld_demod_batch
 * launches on this handle (int16 / float32 input without EFM output for
 * the MFMA variant).  No counterpart in the reference (numpy's FFT). */
int ld_set_fft_variant(ld_handle* h, int variant);

/* The demodulator's forward complex transform alone (blocklen / 2 points),
 * for checking a variant's accuracy: in / out are DEVICE float2 arrays of
 * ntransforms * blocklen / 2 points.  Diagnostic; no counterpart in the
 * reference (its scipy.fft calls, core.py:761, :805, :809-820). */
int ld_fft_forward(ld_handle* h, const void* in, void* out, int ntransforms, int variant, void* stream);

/* One block of a batched launch (ld_demod_batch), in a DEVICE table:
 * `first` is its first input sample, `sys` its system (0 or 1: which
 * handle's tables and outputs; < 0: padding, skipped), `slot` where its
 * outputs go (slot * video_stride / audio_stride in that system's outputs). */
typedef struct ld_batch_block {
  int64_t first;
  int32_t sys;
  int32_t slot;
} ld_batch_block;

/* Blocks of many captures of both systems in ONE launch (the reference runs
 * one LDdecode, with its own RFDecode and DemodCache workers, per capture,
 * core.py:1049-1421, :3498-3700): sys[0], sys[1] are two handles with the
 * same blocklen on the same device, each with its filters uploaded (either
 * may be NULL if no block uses it); outs[k] and nslots[k] are system k's
 * outputs (as ld_demod_blocks') and their capacity in blocks.  Block i of
 * the launch demodulates input[blocks[i].first .. + blocklen) with its
 * system's tables (the same arithmetic as ld_demod_blocks) and writes output
 * slot blocks[i].slot.  The table is DEVICE memory (so a captured HIP graph
 * can replay the launch with new contents); entries that would read outside
 * the input or write outside nslots are skipped on the device (padding
 * entries: sys = -1).  int16 or float32 input, no EFM output.  workspace:
 * ld_demod_workspace_bytes(h, nblocks).  No allocation, copy or memset: safe
 * under stream capture. */
int ld_demod_batch(ld_handle* const* sys, const void* input, int input_dtype, int64_t input_len,
                   const ld_batch_block* blocks, int nblocks, const int32_t* nslots, int cut,
                   const ld_demod_outputs* outs, void* workspace, size_t workspace_bytes, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ld_demod_blocks with its workspace taken from a static
 * device array of ld_demod_noalloc_max_bytes() bytes (32 MiB) that every GPU
 * loading the library holds until the process ends; nothing is allocated or
 * freed.  The caller still passes the input and the outputs.  nblocks is
 * limited to what fits the array (ld_demod_workspace_bytes(h, nblocks) <=
 * 32 MiB: 512 blocks of 16384, 256 of 32768, 61 of 65536 samples); larger
 * requests return LD_EINVAL.  Nothing in
 * the array needs to start at zero (every block writes its workspace before
 * reading it).  All calls on one GPU share the array: calls that overlap on
 * one GPU (concurrent streams, host threads) race and must be serialised by
 * the caller. */
size_t ld_demod_noalloc_max_bytes(void);
int ld_demod_blocks_noalloc(ld_handle* h, const void* input, int input_dtype, int64_t input_len,
                            int64_t first, int64_t step, int nblocks, int cut,
                            const ld_demod_outputs* out, void* stream);

/* Upload the 16-tap sinc interpolation table: (65536 + 1) rows x 16 float32
 * (row p = Kaiser(beta=5)-windowed sinc weights at phase p/65536; the extra
 * row repeats the last), as lddecode/sinc_lut.npz (utils.py:73-124).
 * Replaces the np.load of core.py:314-318. */
int ld_set_sinc_lut(ld_handle* h, const float* host_lut, int rows, int taps);

/* Per-field TBC description for ld_tbc_fields (a HOST array; the library
 * validates it and copies it into the workspace).  Sample positions in the
 * line locations are relative to `demod + demod_offset`. */
typedef struct ld_tbc_field {
  int64_t demod_offset;   /* first sample of this field's demod buffer */
  int64_t demod_len;      /* samples readable from there (taps outside read 0) */
  int64_t lineloc_offset; /* first entry of this field in the linelocs array */
  int32_t nlines;         /* line locations (proclines), 2..1024 */
  int32_t lineoffset;     /* Field.lineoffset (core.py:1614, :3231), 0..1024 */
  int32_t linesout;       /* output lines (outlinecount, core.py:1624), 1..1024 */
  int32_t outwidth;       /* output pixels per line (outlinelen), 1..2048 */
  double inlinelen;       /* Field.inlinelen = RFDecode.linelen (core.py:1609) */
  int64_t out_offset;     /* first output pixel of this field */
} ld_tbc_field;

/* Output level model of hz_to_output_array (utils.py:1061-1075). */
typedef struct ld_level_params {
  double ire0, hz_ire, output_zero, vsync_ire, out_scale;
} ld_level_params;

/* Time-base correct `nfields` fields: per output pixel, FP64 source position
 * and wow from the piecewise-linear line map (computewow_scaled,
 * core.py:2591-2623), wow level adjust with median/MAD outlier clamp
 * (utils.py:131-149), 16-tap sinc-LUT interpolation (utils.py:151-183) and,
 * when out_u16 is non-NULL, the Hz -> uint16 code map (utils.py:1061-1075).
 * coord_f32 != 0 reproduces the reference's float32 cast of the source
 * coordinate (utils.py:160); 0 keeps FP64.  smoothing > 0 applies the
 * per-pixel EMA of utils.py:142-149.  Taps outside [0, demod_len) read 0.
 * `demod` (demod_total samples) and `linelocs` (nlinelocs doubles) are DEVICE
 * arrays; out_f32 (may be NULL) receives the level-adjusted Hz values
 * (scale_field's dsout), out_u16 (may be NULL) the output codes; both hold
 * out_total elements.  `workspace` is device memory of at least
 * ld_tbc_workspace_bytes(h, nfields).
 * Replaces Field.downscale's video path, core.py:2626-2722. */
size_t ld_tbc_workspace_bytes(const ld_handle* h, int nfields);
int ld_tbc_fields(ld_handle* h, const float* demod, int64_t demod_total, const ld_tbc_field* fields,
                  int nfields, const double* linelocs, int64_t nlinelocs, int coord_f32,
                  double smoothing, double threshold, const ld_level_params* level,
                  float* out_f32, uint16_t* out_u16, int64_t out_total, void* workspace,
                  size_t workspace_bytes, void* stream);

/* Sub-sample hsync refinement of every field's line locations on demod_05
 * (Field.refine_linelocs_hsync, core.py:2484-2551, with calczc
 * utils.py:795-854), one launch for the batch.  `fields` is a HOST array (only
 * demod_offset, demod_len, lineloc_offset and nlines are used); line
 * locations of field f are entries [lineloc_offset, +nlines) of the DEVICE
 * arrays linelocs1 / linebad_in (may be NULL) / linelocs2 / linebad_out, in
 * samples relative to demod_05 + demod_offset.  `freq` is samples per
 * microsecond; levels are Hz (iretohz of vsync_ire/2, -55 and 30 IRE); `pal`
 * skips lines 1-2 as well as 3-6.  Reads outside a field's span count as
 * "not found" (the reference would index negatively or raise).
 * `workspace`: device memory of ld_hsync_workspace_bytes(h, nfields). */
size_t ld_hsync_workspace_bytes(const ld_handle* h, int nfields);
int ld_refine_hsync(ld_handle* h, const float* demod_05, int64_t demod_total,
                    const ld_tbc_field* fields, int nfields, const double* linelocs1,
                    const uint8_t* linebad_in, int64_t nlinelocs, double freq, double sync_mid_hz,
                    double lo_hz, double hi_hz, int pal, double* linelocs2, uint8_t* linebad_out,
                    void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * Line-location refinement after the hsync refine (FieldNTSC.process
 * core.py:3472-3495, FieldPAL.process core.py:3212-3237).
 * ------------------------------------------------------------------- */

/* One field of a refinement batch.  Line locations of the field are entries
 * [lineloc_offset, +nlines) of the DEVICE line arrays, in samples relative
 * to the field's demod buffers at demod_offset (length demod_len).
 * linecount is Field.linecount while the refinement runs (263 / 262 for a
 * first / second field, core.py:2340); first_field is isFirstField. */
typedef struct ld_line_field {
  int64_t demod_offset;
  int64_t demod_len;
  int64_t lineloc_offset;
  int32_t nlines;
  int32_t linecount;
  int32_t first_field;
  /* NTSC burst chain the field belongs to (ld_refine_burst_ntsc_chains):
   * chains are consecutive runs of fields numbered 0, 1, ...; 0 for every
   * field otherwise (one chain).  Ignored by the other entry points. */
  int32_t chain;
} ld_line_field;

/* The RFDecode / SysParams constants the refinements read: input rate in
 * MHz (rf.freq), nominal samples per line (rf.linelen), fsc_mhz, pilot_mhz
 * and DecoderParams hz_ire. */
typedef struct ld_line_params {
  double freq_mhz;
  double linelen;
  double fsc_mhz;
  double pilot_mhz;
  double hz_ire;
} ld_line_params;

/* Device workspace of ld_refine_burst_ntsc / ld_refine_pilot_pal. */
size_t ld_refine_workspace_bytes(const ld_handle* h, int nfields, int64_t nlinelocs);

/* NTSC colour-burst line alignment for consecutive fields: linebad |=
 * compute_deriv_error(linelocs2) (core.py:1647), refine_linelocs_burst
 * (core.py:3391-3440 with compute_burst_offsets :3361-3388,
 * compute_line_bursts :2999-3047, clb_findbursts utils.py:1339-1368),
 * fix_badlines with linelocs2 as backup (:2566-2589) and the 117.25 degree
 * subcarrier shift (:3491-3495).  Each field's phase-adjust search starts
 * from the previous field's median (core.py:3366-3369); the first field of
 * the batch from prev_phase_adjust (0 for none).  `fields` is a HOST array;
 * demod and demod_burst are the DEVICE video channels; linelocs2 /
 * linebad_in are the hsync refine's DEVICE outputs.  Outputs (DEVICE):
 * linelocs_out and linebad_out (Field.linelocs / linebad after process),
 * phase_adjust_out[f] (phase_adjust_median, NaN without a valid burst) and
 * field_phase_id[f] (fieldPhaseID).  nlines must be 267..512. */
int ld_refine_burst_ntsc(ld_handle* h, const float* demod, const float* demod_burst, int64_t demod_total,
                         const ld_line_field* fields, int nfields, const ld_line_params* params,
                         const double* linelocs2, const uint8_t* linebad_in, int64_t nlinelocs,
                         double prev_phase_adjust, double* linelocs_out, uint8_t* linebad_out,
                         double* phase_adjust_out, int32_t* field_phase_id, void* workspace,
                         size_t workspace_bytes, void* stream);

/* ld_refine_burst_ntsc over several independent chains of consecutive
 * fields in one call (fields of different captures decoded together): the
 * fields' `chain` ids are consecutive runs numbered 0 .. nchains-1, and
 * chain c's first field starts its search from host chain_seeds[c] (each
 * capture's LDdecode carries its own previous field, core.py:3366-3369).
 * ld_refine_burst_ntsc is this with one chain seeded by prev_phase_adjust. */
int ld_refine_burst_ntsc_chains(ld_handle* h, const float* demod, const float* demod_burst, int64_t demod_total,
                                const ld_line_field* fields, int nfields, const ld_line_params* params,
                                const double* linelocs2, const uint8_t* linebad_in, int64_t nlinelocs,
                                const double* chain_seeds, int nchains, double* linelocs_out,
                                uint8_t* linebad_out, double* phase_adjust_out, int32_t* field_phase_id,
                                void* workspace, size_t workspace_bytes, void* stream);

/* PAL pilot line alignment: linebad |= compute_deriv_error(linelocs2), two
 * refine_linelocs_pilot passes (core.py:3054-3093, :3221-3224) and
 * fix_badlines (:3226) for every field of the batch (fields independent).
 * Arguments as ld_refine_burst_ntsc; demod_pilot is the DEVICE pilot
 * channel.  nlines must be 323..512. */
int ld_refine_pilot_pal(ld_handle* h, const float* demod_pilot, int64_t demod_total,
                        const ld_line_field* fields, int nfields, const ld_line_params* params,
                        const double* linelocs2, const uint8_t* linebad_in, int64_t nlinelocs,
                        double* linelocs_out, uint8_t* linebad_out, void* workspace,
                        size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * Line location (Field.compute_linelocs, core.py:2325-2481, and the field
 * walk of LDdecode.decodefield, core.py:4018-4098).  Sync pulses, their
 * levels and the skip-check medians are found on the GPU; the pulse state
 * machine, vertical-blanking analysis and line-0 logic run on the host over
 * the pulse lists.
 * ------------------------------------------------------------------- */

/* Device workspace of ld_find_pulses for an n-sample buffer. */
size_t ld_pulse_workspace_bytes(const ld_handle* h, int64_t n);

/* Sync-pulse edges of the DEVICE float32 buffer x[0, n) (findpulses_numba_raw,
 * utils.py:1101-1141, run over the whole buffer): a sample is in a pulse when
 * (double)x <= high.  edges (DEVICE) receives, in order, the first sample of
 * every run of in-pulse samples and the first sample after the run,
 * alternating start / end (a run reaching the end of the buffer has no end
 * edge); at most `cap` are written and *count (DEVICE) receives the total
 * number of edges, so a count above cap means the list was cut short.
 * Pulse k is (edges[2k], edges[2k+1] - edges[2k]).  The reference's
 * per-field pulse list is the subset with start > the field buffer's first
 * sample, end < its last and length <= 5000. */
int ld_find_pulses(ld_handle* h, const float* x, int64_t n, double high, int64_t* edges, int64_t cap,
                   int64_t* count, void* workspace, size_t workspace_bytes, void* stream);

/* ld_find_pulses from the demod's in-pulse bits (ld_demod_outputs.
 * demod_05_below's paired DEVICE words, at least 2 * ceil(n / 64) of them,
 * 8-byte aligned) instead of the samples: the same edge list for the same
 * buffer and threshold, reading n / 8 bytes.  Same findpulses_numba_raw
 * (utils.py:1101-1141). */
size_t ld_pulse_bits_workspace_bytes(const ld_handle* h, int64_t n);
int ld_find_pulses_bits(ld_handle* h, const uint32_t* bits, int64_t n, int64_t* edges, int64_t cap,
                        int64_t* count, void* workspace, size_t workspace_bytes, void* stream);

/* getpulses' vertical-sync levels (core.py:2274-2286): for every pulse k
 * of ld_find_pulses' list whose length exceeds min_len, the mean of
 * x[int(start + trim), int(start + len - trim)) as np.mean computes it on a
 * float32 array (pairwise float32 sum, float32 divide); NaN for the other
 * pulses.  The edge count is read from the DEVICE `count`; means (DEVICE)
 * gets one value per pulse k < min(count / 2, cap_pulses). */
int ld_pulse_means(ld_handle* h, const float* x, int64_t n, const int64_t* edges, const int64_t* count,
                   int64_t cap_pulses, double min_len, double trim, float* means, void* stream);

/* The pulse list to the host: *host_count = *count (DEVICE), then
 * host_edges[0, min(count, cap_edges)) and host_means[0, min(count,
 * cap_edges) / 2) from the DEVICE edges and means, written by the GPU into
 * pinned host memory (hipHostMalloc / a pinned torch tensor: LD_EHIP for
 * memory the GPU cannot reach), so that only the edges found cross the bus,
 * not the lists' capacity.  The host reads them after the stream's work up
 * to this call has completed.  No counterpart in the reference, whose pulse
 * list never leaves host memory (Field.getpulses, core.py:2242-2330). */
int ld_pulse_export(ld_handle* h, const int64_t* edges, const int64_t* count, const float* means, int64_t cap_edges,
                    int64_t* host_edges, float* host_means, int64_t* host_count, void* stream);

/* One buffer of ld_pulse_search: ld_find_pulses_bits over `bits` (n
 * samples), ld_pulse_means over x (DEVICE demod_05; min_len, trim as there,
 * cap / 2 levels) and ld_pulse_export into the pinned host lists. */
typedef struct ld_pulse_job {
  const uint32_t* bits;
  int64_t n;
  const float* x;
  int64_t* edges;       /* DEVICE, cap entries (cap even) */
  int64_t* count;       /* DEVICE */
  float* means;         /* DEVICE, cap / 2 entries */
  int64_t cap;
  double min_len;
  double trim;
  int64_t* host_edges;  /* pinned host, cap entries */
  float* host_means;    /* pinned host, cap / 2 entries */
  int64_t* host_count;  /* pinned host */
} ld_pulse_job;

/* The pulse searches of up to 4 buffers (a batch's systems) in four
 * launches for all of them: each job gets exactly what ld_find_pulses_bits,
 * ld_pulse_means and ld_pulse_export would give it, so the host reads each
 * buffer's edges and levels after the stream's work up to this call.  DEVICE
 * workspace of ld_pulse_search_workspace_bytes bytes.  No counterpart in the
 * reference (one Field.getpulses per field, core.py:2242-2330). */
size_t ld_pulse_search_workspace_bytes(const ld_handle* h, const ld_pulse_job* jobs, int njobs);
int ld_pulse_search(ld_handle* h, const ld_pulse_job* jobs, int njobs, void* workspace, size_t workspace_bytes,
                    void* stream);

/* np.mean (float32) of x[begin[i], end[i]) for `count` ranges (DEVICE x,
 * begin, end, out); an empty range gives NaN.  getpulses' black levels
 * (core.py:2296-2312). */
int ld_range_means(ld_handle* h, const float* x, int64_t n, const int64_t* begin, const int64_t* end,
                   int count, float* out, void* stream);

/* What a field leaves for the field after it (getLine0, core.py:2157-2239):
 * ll_linecount = its final linelocs[linecount] (relative to its own buffer;
 * linecount 263 / 262 NTSC, 312 / 313 PAL) and the eight
 * line medians skip_check (core.py:2122-2155) takes over demod[int(b),
 * int(b + line_samples + 1)) with b = linelocs[outlinecount + lineoffset + j].
 * known = 0 marks facts not measured yet. */
typedef struct ld_field_facts {
  double ll_linecount;
  double medians[8];
  int32_t known;
  int32_t reserved;
} ld_field_facts;

/* Measure ld_field_facts of every field from its final line locations (the
 * output of ld_refine_burst_ntsc / ld_refine_pilot_pal).  fields is a HOST
 * array (demod_offset, demod_len, lineloc_offset, nlines, first_field are
 * used); demod, linelocs and facts are DEVICE arrays.  pal selects PAL's
 * line counts and offsets (312 / 313, lineoffset 2 / 3).  The medians are
 * numba's np.median of float32 data (middle value; for an even count the
 * float32 sum of the two middle values halved in float64).  A line whose
 * span leaves the field buffer gives a NaN median.  `workspace`: device memory
 * of ld_field_facts_workspace_bytes(h, nfields). */
size_t ld_field_facts_workspace_bytes(const ld_handle* h, int nfields);
int ld_field_facts_measure(ld_handle* h, const float* demod, int64_t demod_total, const ld_line_field* fields,
                           int nfields, const double* linelocs, int64_t nlinelocs, int pal, int outlinecount,
                           double line_samples, ld_field_facts* facts, void* workspace, size_t workspace_bytes,
                           void* stream);

/* RFDecode / SysParams / DecoderParams / LDdecode values line location reads
 * (core.py:62-170, :384-393, :3625-3634). */
typedef struct ld_locate_params {
  double freq;                /* rf.freq, MHz */
  double linelen;             /* rf.linelen (rounded to an integer) */
  double ire0, hz_ire, vsync_ire;
  double hsync_us, eq_us, vsync_us, line_period_us;
  double first_field_h[2];    /* SysParams firstFieldH */
  double hsync_tolerance;     /* rf.hsync_tolerance */
  double prov_delta;          /* provisional final - linelocs1 offset at linecount (facts unknown) */
  int32_t pal;
  int32_t num_pulses;         /* SysParams numPulses */
  int32_t field_lines[2];     /* SysParams field_lines */
  int32_t frame_lines;
  int32_t reserved;
  int64_t blocklen;           /* rf.blocklen: decodefield's read granularity */
  int64_t blockcut;           /* rf.blockcut */
  int64_t block_stride;       /* DemodCache.blocksize: usable samples per block */
  int64_t readlen;            /* LDdecode.readlen */
} ld_locate_params;

/* The previous field as getLine0 / getpulses see it. */
typedef struct ld_prev_field {
  int32_t valid;
  int32_t is_first_field;
  int32_t linecount;          /* 263 / 262 NTSC, 312 / 313 PAL */
  int32_t sync_confidence;
  int64_t startloc;
  double ll_linecount;        /* its final linelocs[linecount], relative to startloc */
  int32_t skip_score;         /* its skip_check() */
  int32_t fields_written;     /* LDdecode.fields_written after it */
} ld_prev_field;

enum {
  LD_LOC_OK = 0,              /* valid field: linelocs1 / linebad written */
  LD_LOC_FAILED = 1,          /* compute_linelocs found no usable field (Field.valid False) */
  LD_LOC_NEED_LEVELS = 2,     /* getpulses would re-threshold at measured sync / black levels */
  LD_LOC_NEED_IRE0 = 3        /* getpulses would recalibrate ire0 (first field without pulses) */
};

/* Flags of ld_field_loc.vb_flags: processVBlank's result for the local
 * vertical blanking (index 0) and the next one (index 1). */
enum {
  LD_VB_FOUND = 1,            /* a line 0 was found */
  LD_VB_FIRST = 2,            /* ... of a first field */
  LD_VB_ZEROED = 4,           /* the fallback that zeroes sync_confidence ran */
  LD_VB_BLANK = 8,            /* getBlankRange found a range (vb_firstblank valid) */
  LD_VB_ERROR = 16            /* the reference raises here */
};

/* One field of the walk: decodefield's read and compute_linelocs' results.
 * Positions are relative to the field buffer (stitched index startloc),
 * as the reference's Field sees them.  The vb_* members keep getLine0's
 * inputs (processVBlank without a limit for the local and next vertical
 * blanking), from which ld_locate_recheck re-takes the line-0 decision. */
typedef struct ld_field_loc {
  double start;               /* raw position decodefield was called with */
  int64_t readloc;
  int64_t startloc;           /* field buffer origin: DemodCache.read's startloc */
  int64_t buflen;             /* field buffer length (whole blocks) */
  double line0loc;
  double meanlinelen;
  double nextfieldoffset;     /* Field.nextfieldoffset (after process()) */
  double offset;              /* decodefield's returned offset: next start = start + offset */
  int64_t vblank_next;        /* -1 when none */
  int32_t status;             /* LD_LOC_* */
  int32_t is_first_field;
  int32_t linecount;          /* 263 / 262 */
  int32_t sync_confidence;
  int32_t skipdetected;
  int32_t npulses;            /* pulses in the field buffer */
  int32_t nlines;             /* proclines written */
  int32_t analyzed;           /* 1: the pulses were classified and vb_* hold getLine0's inputs */
  int32_t nvalid;             /* classified pulses kept for ld_locate_recheck (-1: none kept) */
  int64_t last_start;         /* the buffer's last pulse start (rawpulses[-1].start) */
  double vb_line0[2];
  int64_t vb_firstblank[2];
  int32_t vb_conf[2];         /* -1: None */
  int32_t vb_flags[2];        /* LD_VB_* */
  int32_t reserved;
  double hsync_max;           /* get_timings' LT['hsync'][1] (dropout detection's hsync length) */
} ld_field_loc;

/* One classified pulse (refinepulses' validpulses entry, core.py:1914-1949). */
typedef struct ld_vpulse {
  int64_t start;
  int32_t len;
  int16_t kind;               /* 0 hsync, 1 equalising (before), 2 vsync, 3 equalising (after) */
  int16_t good;               /* pulse_qualitycheck */
} ld_vpulse;

/* compute_linelocs for one field from its pulse list (HOST arrays; starts
 * relative to the field buffer, as getpulses returns them) and the previous
 * field (NULL: none).  Writes *out and, for a valid field, linelocs1 /
 * linebad[0, proclines) (max_lines must be >= proclines: 273 NTSC, 326 PAL).
 * Host only. */
int ld_locate_field(const ld_locate_params* p, const int64_t* starts, const int64_t* lens, int npulses,
                    const ld_prev_field* prev, int64_t startloc, int fields_written, ld_field_loc* out,
                    double* linelocs1, uint8_t* linebad, int max_lines);

/* The field walk over a stitched demod_05 buffer of `total` samples from raw
 * sample `start`, given ld_find_pulses' edges and ld_pulse_means' levels
 * (HOST copies).  Field k's buffer, pulses and compute_linelocs results go to
 * out[k] and linelocs1 / linebad[k * max_lines, + proclines); the next field
 * starts at start + out[k].offset.  What field k leaves for field k + 1 comes
 * from facts[k] when facts is non-NULL and facts[k].known, otherwise from a
 * provisional estimate (its linelocs1[linecount] + prov_delta, skip score
 * 100): a caller verifies a provisional walk by walking again with the
 * measured facts.  The walk stops after max_fields fields, before a field
 * whose buffer would pass `total`, or after a field with status
 * LD_LOC_NEED_LEVELS / LD_LOC_NEED_IRE0 (the caller resolves it with
 * ld_range_means / ld_find_pulses and ld_locate_field).  vpulses (may be
 * NULL) receives field k's classified pulses at k * max_vpulses (its nvalid
 * says how many, -1 when they did not fit).  *nfields_out gets
 * the number of fields written, *prev_out the state after the last one.
 * Host only. */
int ld_locate_fields(const ld_locate_params* p, const int64_t* edges, int64_t nedges, const float* means,
                     int64_t total, double start, const ld_prev_field* prev, const ld_field_facts* facts,
                     int max_fields, ld_field_loc* out, double* linelocs1, uint8_t* linebad, int max_lines,
                     ld_vpulse* vpulses, int max_vpulses, ld_prev_field* prev_out, int* nfields_out);

/* One ld_locate_fields call of a batch (arguments as there; nfields gets
 * *nfields_out, status the call's return code). */
typedef struct ld_walk_job {
  const ld_locate_params* p;
  const int64_t* edges;
  int64_t nedges;
  const float* means;
  int64_t total;
  double start;
  const ld_prev_field* prev;
  const ld_field_facts* facts;
  ld_field_loc* out;
  double* linelocs1;
  uint8_t* linebad;
  ld_vpulse* vpulses;
  ld_prev_field* prev_out;
  int32_t max_fields;
  int32_t max_lines;
  int32_t max_vpulses;
  int32_t nfields;
  int32_t status;
  int32_t reserved;
} ld_walk_job;

/* The walks of many captures at once (a batch's captures are independent)
 * on up to nthreads host threads.  No counterpart in the reference (one
 * LDdecode per capture, core.py:3498); each job equals its own
 * ld_locate_fields call.  Returns LD_OK or the first failing job's code.
 * Host only. */
int ld_locate_fields_many(ld_walk_job* jobs, int njobs, int nthreads);

/* ld_locate_fields_many started on host threads and returned from at once:
 * *run gets a handle that ld_locate_fields_many_wait must receive exactly
 * once (it joins the threads, frees the handle and returns what
 * ld_locate_fields_many would have).  The jobs, and everything they point
 * to, must stay valid and untouched until then.  Jobs are taken longest
 * first (by pulse count).  No counterpart in the reference: it lets the
 * caller queue GPU work for other captures while these walks run.  Host
 * only. */
int ld_locate_fields_many_start(ld_walk_job* jobs, int njobs, int nthreads, void** run);
int ld_locate_fields_many_wait(void* run);

/* Check a walk made with provisional previous-field values against the
 * measured facts (facts[k] belongs to locs[k]; prev is the state the walk
 * started from; HOST arrays as ld_locate_fields wrote them, vpulses
 * included).  Every field's line-0 decision (getLine0, core.py:2157-2239) is
 * taken again from its kept inputs; where it changes, the line table is
 * rebuilt from the kept pulses (core.py:2338-2481).  A field whose status,
 * parity, confidence, next offset, linelocs1 or bad flags change stops the
 * check: *first_changed gets its index (the walk must be redone from there
 * with the measured facts).  Otherwise *first_changed is -1 and the records
 * are updated in place (line0loc, vblank_next, skipdetected), so that they
 * equal the sequential walk's.  Host only. */
int ld_locate_recheck(const ld_locate_params* p, ld_field_loc* locs, double* linelocs1, uint8_t* linebad,
                      int max_lines, const ld_vpulse* vpulses, int max_vpulses, int n, const ld_prev_field* prev,
                      const ld_field_facts* facts, int* first_changed);

/* ---------------------------------------------------------------------
 * Dropout detection (Field.dropout_detect, core.py:2808-2996).
 * ------------------------------------------------------------------- */

/* One field: its buffer in the stitched channels, its final line locations
 * (entries [lineloc_offset, + nlines) of the DEVICE array), lineoffset and
 * linecount (0 / 263|262 NTSC, 2|3 / 312|313 PAL), isFirstField and
 * get_timings' LT['hsync'][1] (ld_field_loc.hsync_max). */
typedef struct ld_dropout_field {
  int64_t demod_offset;
  int64_t demod_len;
  int64_t lineloc_offset;
  int32_t nlines;
  int32_t lineoffset;
  int32_t linecount;
  int32_t first_field;
  double hsync_max;
} ld_dropout_field;

/* The levels of dropout_detect_demod as float32 (np.full_like of the
 * float32 buffers): default / sync-area minimum and maximum of demod and
 * demod_05, freq_hz / 2 for demod_raw, rf.delays['video_rot']. */
typedef struct ld_dropout_params {
  double freq_hz_half;
  float vmin, vmax, vmin05, vmax05, sync_min, sync_min05;
  int32_t video_rot;
  int32_t pal;
} ld_dropout_params;

size_t ld_dropout_workspace_bytes(const ld_handle* h, const ld_dropout_field* fields, int nfields);

/* dropout_detect_demod (core.py:2830-2887) for every field of a batch:
 * the error samples of each field buffer (rfhpf beyond 3 std of the buffer,
 * shifted by video_rot; demod_raw above freq_hz / 2; demod / demod_05 out of
 * range, with the sync-area minimum over each line's hsync length and over
 * the vertical-sync lines; only between int(linelocs[lineoffset + 1]) and
 * int(linelocs[lineoffset + linecount + 1])), compacted in order:
 * positions (DEVICE, buffer-relative, at most cap) hold field 0's errors,
 * then field 1's, ...; field_start (DEVICE, nfields + 1) each field's first
 * index and the total.  `fields` is a HOST array; the channels and linelocs
 * are DEVICE arrays.  The buffer's std is a double reduction (the
 * reference's numba std sums in order: the float32 threshold agrees).
 * Line locations are taken to increase along each field. */
int ld_dropout_detect(ld_handle* h, const float* rfhpf, const float* demod, const float* demod_raw,
                      const float* demod_05, int64_t total, const ld_dropout_field* fields, int nfields,
                      const double* linelocs, int64_t nlinelocs, const ld_dropout_params* params,
                      int64_t* positions, int64_t cap, int64_t* field_start, void* workspace,
                      size_t workspace_bytes, void* stream);

/* build_errlist + dropout_errlist_to_tbc + dropout_detect's packing
 * (core.py:2890-2996) for one field from its sorted error positions and
 * final line locations (HOST arrays): lines (field line - 1), starts, ends
 * in output pixels, at most cap; *count gets the number.  Host only. */
int ld_dropout_list(const int64_t* errmap, int64_t n, const double* linelocs, int nlines, int lineoffset,
                    int linecount, int outlinelen, double freq, int32_t* lines, int32_t* starts, int32_t* ends,
                    int cap, int* count);

/* ---------------------------------------------------------------------
 * Analogue audio stage 2 and its output rate (RFDecode.audio_phase2,
 * core.py:884-950; downscale_audio, core.py:1423-1578).
 * ------------------------------------------------------------------- */

/* One field read: its stage-1 audio (entries [audio_offset, + audio_len)
 * of the stitched stage-1 arrays), where its stage-2 output goes
 * (a2_offset, audio_len entries), its final
 * line locations (entries [lineloc_offset, + nlines) of the DEVICE array)
 * and linecount, and for the output: the first tick's time (timeoffset,
 * seconds; 0 for a line-locked rate), nout output sample pairs written from
 * pair out_offset on. */
typedef struct ld_audio_field {
  int64_t audio_offset;
  int64_t audio_len;
  int64_t a2_offset;
  int64_t lineloc_offset;
  int64_t out_offset;
  double timeoffset;
  int32_t nlines;
  int32_t linecount;
  int32_t nout;
  int32_t reserved;
} ld_audio_field;

/* downscale_audio's constants: SysParams line_period (us), rf.linelen, the
 * output rate as Hz (already resolved for a line-locked rate), the stage-1
 * decimation audio_fdiv, and the two carriers. */
typedef struct ld_audio_out_params {
  double line_period;
  double linelen;
  double freq;
  double scale;
  double lfreq, rfreq;
} ld_audio_out_params;

size_t ld_audio2_workspace_bytes(const ld_handle* h, const ld_audio_field* fields, int nfields);

/* audio_phase2 for every field read of a batch (core.py:884-950): per
 * blocklen segment and channel the stage-1 audio minus its carrier, zeroed
 * within 8 samples of the left channel's clip peaks in that segment
 * (findpeaks over 500 kHz, utils.py:1207), filtered by the stage-2 filter
 * (FP32 transforms), plus the carrier.  A read of at most blocklen samples
 * is one zero-padded segment (float64 result, as the reference returns it);
 * a longer read is assembled from the reference's overlapping segments
 * (at 0, every blocklen - 512 samples, and at len - blocklen - 1, placed
 * one sample late as the reference places it), each output sample taken
 * from the last segment that writes it and rounded to float32 as the
 * reference's array of that case holds it.  Written as float64 to
 * out_left / out_right (DEVICE, out_total entries each).  `fields` is a
 * HOST array. */
int ld_audio2_fields(ld_handle* h, const float* audio_left, const float* audio_right, int64_t total,
                     const ld_audio_field* fields, int nfields, double* out_left, double* out_right,
                     int64_t out_total, void* workspace, size_t workspace_bytes, void* stream);

/* _downscale_audio_compute_locs_and_swow + _downscale_audio_to_output
 * (core.py:1423-1535) for every field: nout interleaved int16 pairs per
 * field into out (DEVICE, out_pairs pairs), failed[f] (DEVICE) set when a
 * tick had no audio under it (the samples stay 0, as in the reference).
 * a2_left / a2_right: the stage-2 output; linelocs: DEVICE.  Each field's
 * nout must equal its tick count minus one (numba's arange length) and be
 * at most 4095. */
int ld_audio_downscale(ld_handle* h, const double* a2_left, const double* a2_right, int64_t a2_total,
                       const ld_audio_field* fields, int nfields, const double* linelocs, int64_t nlinelocs,
                       const ld_audio_out_params* params, int16_t* out, int64_t out_pairs, int32_t* failed,
                       void* workspace, size_t workspace_bytes, void* stream);

/* skip_check's score (core.py:2139-2155) from a field's eight line medians. */
int ld_skip_score(const ld_locate_params* p, const double* medians8);

#ifdef __cplusplus
}
#endif
#endif /* LDGPU_H */

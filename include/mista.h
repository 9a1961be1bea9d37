/*
 * mista — MI355X-native STA/TFM delay-and-sum imaging, C ABI.
 *
 * Drop-in boundary for the reference's STA array-processor family
 * (src/user/processor/ArrayProcessor.h:27-40 and its implementations
 * VectorialSTAProcessor.h:58-341, VectorialSTACUDAProcessor.h:40-127,
 * VectorialSTAOCLProcessor.h:73-680).  The reference exposes that family as a
 * C++ class: constructor(config, acquisition, upsamplingFactor, coherenceFactor,
 * peakOffset, rxApod) + prepare(baseElement) + process(gridData).  This header
 * is the same contract as POD-only C functions; the C++ adapter in
 * us_lab4a_amd/include/mista/HipSTAProcessor.h rebuilds the class on top.
 *
 * Conventions
 *  - Every function returns MISTA_OK (0) or a negative MISTA_E_* status; the
 *    message is available from mista_last_error(handle) (or NULL handle for
 *    errors raised before a handle exists).  No exception crosses the ABI.
 *  - Handles are single-threaded.  The library owns its device buffers and a
 *    default-priority HIP stream per handle; it never calls hipDeviceReset.
 *  - "_device" entry points take device pointers and an optional hipStream_t
 *    (passed as void*, NULL = the handle's stream) and return without
 *    synchronising.  Host entry points are synchronous.
 *  - Grid layout follows ImageGrid.h:137-143: n1 columns (x) by n2 rows (z),
 *    z fastest; point index = col*n2 + row.
 *  - Echo layout follows STAAcquisition.h:71-86 per tx (Matrix[rx][samples],
 *    row-major) stacked over tx: [ntx][nrx][ns] float32.
 */
#ifndef MISTA_H
#define MISTA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MISTA_ABI_VERSION 1

/* ---- status codes (Exception.h:38-94 analogues) ---- */
#define MISTA_OK               0
#define MISTA_E_INVALID_PARAM -1  /* InvalidParameterException                 */
#define MISTA_E_INVALID_VALUE -2  /* InvalidValueException (e.g. dead zone)     */
#define MISTA_E_INVALID_STATE -3  /* InvalidStateException (missing grid/apod)  */
#define MISTA_E_HIP           -4  /* HIP runtime/library error (CUDAException)  */
#define MISTA_E_NO_MEMORY     -5  /* device allocation failed                   */
#define MISTA_E_BUSY          -6  /* serving engine full (admission refused)    */

/* ---- coherence factor: CoherenceFactor.h:139-152 (+ SCF extension) ---- */
#define MISTA_CF_NONE 0
#define MISTA_CF_PCF  1   /* phase coherence factor, param = gamma              */
#define MISTA_CF_SCF  2   /* sign coherence factor of Re(per-rx sum), param = p */

/* ---- analytic-signal storage in HBM (accumulation is always FP32) ---- */
#define MISTA_STORAGE_F32  0
#define MISTA_STORAGE_F16  1  /* per-channel scale = max(|re|,|im|)             */
#define MISTA_STORAGE_BF16 2  /* per-channel scale = max(|re|,|im|)             */

/*
 * Replaces STAConfiguration<float> (STAConfiguration.h:118-165) plus the
 * non-object constructor arguments of VectorialSTAProcessor
 * (VectorialSTAProcessor.h:61-67) and the coherence-factor selection of
 * AnalyticSignalCoherenceFactorProcessor (CoherenceFactor.h:139-152).
 */
typedef struct mista_sta_config {
	uint32_t num_elements;       /* numElements: rx count = element count        */
	uint32_t first_tx_elem;      /* firstTxElem                                  */
	uint32_t last_tx_elem;       /* lastTxElem (inclusive)                       */
	uint32_t upsampling_factor;  /* 1..64 (Interpolator.h:49 maxUpFactor)        */
	float pitch;                 /* m                                            */
	float sampling_frequency;    /* Hz                                           */
	float propagation_speed;     /* m/s                                          */
	float center_frequency;      /* Hz                                           */
	float dead_zone_m;           /* m                                            */
	float peak_offset;           /* periods of center_frequency                  */
	int32_t cf_kind;             /* MISTA_CF_*                                   */
	float cf_param;              /* gamma (PCF) or p (SCF)                       */
	int32_t storage;             /* MISTA_STORAGE_*                              */
	int32_t device;              /* HIP device ordinal                           */
} mista_sta_config;

typedef struct mista_handle mista_handle;

/* ABI version, for a caller's dlopen check. */
int mista_abi_version(void);

/* VectorialSTAProcessor constructor (VectorialSTAProcessor.h:129-149):
 * validates the configuration, selects the device.  rxApod is set separately
 * with mista_set_rx_apod (the reference copies it in the constructor). */
int mista_create(const mista_sta_config* config, mista_handle** out);

/* Processor destructor.  Frees device memory; never resets the device. */
int mista_destroy(mista_handle* h);

/* Last error message of `h`, or of the calling thread when h is NULL. */
const char* mista_last_error(const mista_handle* h);

/* ArrayProcessor::prepare(baseElement) (ArrayProcessor.h:33,
 * VectorialSTAProcessor.h:151-156).  Records the base element; the echo
 * source is the caller's (see mista_process). */
int mista_prepare(mista_handle* h, uint32_t base_element);

/* rxApod constructor argument (VectorialSTAProcessor.h:67,121): n must equal
 * num_elements.  Copied. */
int mista_set_rx_apod(mista_handle* h, const float* rx_apod, uint32_t n);

/* Extension (no reference counterpart): per-tx apodization, n = number of tx
 * (last-first+1); NULL restores all ones (the reference behaviour). */
int mista_set_tx_apod(mista_handle* h, const float* tx_apod, uint32_t n);

/* Image grid: grid_xz is [n1][n2][2] (x, z) in m, the x/z fields of the
 * reference's Matrix<XYZValueFactor<float>> gridData (XYZValueFactor.h:23-32).
 * Builds the element-to-pixel delay table in FP64 on the GPU, rounded to FP32
 * samples, and keeps it resident (replaces calculateDelaysSTAKernel,
 * VectorialSTACUDAProcessor.cu:30-53).  May be called again with a new shape. */
int mista_set_grid(mista_handle* h, const float* grid_xz, uint32_t n1, uint32_t n2);

/* Same, for a caller holding the reference's AoS points {x,y,z,value,factor}
 * (5 floats per point, n1*n2 points).  A 2-D processor ignores y, as
 * Geometry::distance2DY0 does (Geometry.h:55-65). */
int mista_set_grid_points(mista_handle* h, const float* points, uint32_t n1, uint32_t n2);

/* Grid of n1 x n2 points given as [n1][n2][3] (x, y, z) in m: any
 * Matrix<XYZValueFactor<float>> layout, as Vectorial3DSTAProcessor::process
 * takes (Vectorial3DSTAProcessor.h:137-253).  ny > 1 declares a volume laid
 * out as n1 = nx * ny columns (y fastest) by n2 = nz rows, which lets the DAS
 * kernel tile it in 3-D; ny = 1 for any other point set.  n1 % ny == 0. */
int mista_set_grid_xyz(mista_handle* h, const float* grid_xyz, uint32_t n1, uint32_t n2, uint32_t ny);

/*
 * 3-D sparse STA: replaces SA3DConfiguration<float> (SA3DConfiguration.h:
 * 37-58) + the constructor arguments of Vectorial3DSTAProcessor
 * (Vectorial3DSTAProcessor.h:54-60,103-130).  The caller resolves the active
 * element lists: tx_xy[num_tx][2] are the positions
 * txElemPos[baseElement + activeTxElem[i]], rx_xy[num_rx][2] the positions
 * rxElemPos[baseElement + activeRxElem[i]] (ArrayUtil.h:51-69), in m, on
 * the z = 0 plane.  Echoes are [num_tx][num_rx][ns]; rx apodization has
 * num_rx entries; distances are Geometry::distance3DZ0 (Geometry.h:77-84).
 * Every other entry point (grid, apodization, process, device paths) works on
 * a 3-D handle unchanged.
 */
typedef struct mista_sta3d_config {
	uint32_t num_tx;             /* activeTxElem.size()                          */
	uint32_t num_rx;             /* activeRxElem.size()                          */
	uint32_t upsampling_factor;  /* 1..64                                        */
	float sampling_frequency;    /* Hz                                           */
	float propagation_speed;     /* m/s                                          */
	float center_frequency;      /* Hz                                           */
	float dead_zone_m;           /* m                                            */
	float peak_offset;           /* periods of center_frequency                  */
	int32_t cf_kind;             /* MISTA_CF_*                                   */
	float cf_param;              /* gamma (PCF) or p (SCF)                       */
	int32_t storage;             /* MISTA_STORAGE_*                              */
	int32_t device;              /* HIP device ordinal                           */
} mista_sta3d_config;

int mista_create_3d(const mista_sta3d_config* config, const float* tx_xy, const float* rx_xy, mista_handle** out);

/* New active-element positions for a 3-D handle (a new baseElement in the
 * reference's prepare, Vectorial3DSTAProcessor.h:133-138); rebuilds the
 * resident delay table if a grid is set. */
int mista_set_elements_3d(mista_handle* h, const float* tx_xy, const float* rx_xy);

/* VectorialSTAProcessor::process (VectorialSTAProcessor.h:158-341) for one
 * frame from host memory: echoes [ntx][nrx][ns] float32 (the stacked
 * acquisition.execute(tx) matrices).  Writes value = |sum|/(Ntx*Nrx) and
 * factor (1 when no coherence factor) for every grid point, host arrays of
 * n1*n2 floats.  factor may be NULL.  Synchronous. */
int mista_process(mista_handle* h, const float* echoes, uint32_t num_samples,
		float* value, float* factor);

/* Same, writing into the reference's AoS points (value and factor fields of
 * each {x,y,z,value,factor}); Util::resetValueFactor semantics. */
int mista_process_points(mista_handle* h, const float* echoes, uint32_t num_samples, float* points);

/* Device-resident frame: d_echoes [ntx][nrx][ns] float32 in HBM ->
 * d_value/d_factor [n1*n2] in HBM (d_factor may be NULL).  Asynchronous on
 * `stream` (NULL = handle stream). */
int mista_process_device(mista_handle* h, const float* d_echoes, uint32_t num_samples,
		float* d_value, float* d_factor, void* stream);

/* Signal preparation only (interpolation, DC removal, analytic signal:
 * VectorialSTAProcessor.h:169-231) into the handle's resident tensor. */
int mista_prepare_signals_device(mista_handle* h, const float* d_echoes, uint32_t num_samples, void* stream);

/* Delay-and-sum only over the handle's resident tensor. */
int mista_das_device(mista_handle* h, float* d_value, float* d_factor, void* stream);

/* No counterpart in the reference (frames of NetworkSyncSTAMethod.cpp:93-140
 * each run their own processor): the DAS of n prepared handles, up to 16
 * frames in one launch of the tiled kernel when their launches agree in
 * storage, coherence factor, transmit block and workgroup size and their
 * geometries have been tuned (a frame each handle ran through
 * mista_das_device/mista_process_device); others run one by one.  Every
 * frame writes the images it writes alone.  Work goes on `stream` (NULL:
 * the first handle's), which must be ordered after each handle's
 * preparation.  mista_last_timing reports nothing for these frames. */
int mista_das_batch_device(mista_handle* const* handles, uint32_t n, float* const* d_values, float* const* d_factors,
		void* stream);

/* No counterpart in the reference: mista_das_device for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same inputs and outputs and the same
 * images as mista_das_device with the tiled kernel; it calls no hipMalloc
 * or hipFree.  The handle must hold a prepared frame (the preparation sizes
 * every buffer the DAS reads).  The window tables of the launch live in a
 * static device array of 80 MiB that each GPU holds once per process and
 * that every call on that GPU rebuilds: a geometry whose tables need more
 * is refused with MISTA_E_INVALID_PARAM, and calls that overlap on one GPU
 * (from several threads or streams) race.  Uses the tile shape tuned for the
 * geometry when there is one, else a default; it never tunes.  Stage timing
 * (mista_last_timing) is recorded only if an earlier frame created the
 * handle's events.  Synchronises the stream once (to size the windows). */
int mista_das_device_noalloc(mista_handle* h, float* d_value, float* d_factor, void* stream);

/* Test access to the resident analytic tensor: host complex float32
 * [ntx][nrx][L][2].  Upload converts to the storage type (with per-channel
 * scale for F16/BF16); download converts back to float32. */
int mista_upload_analytic(mista_handle* h, const float* analytic_ri, uint32_t signal_length);
int mista_download_analytic(mista_handle* h, float* analytic_ri);

/* Test access to the FP32 delay table [num_elements][n1*n2] (samples). */
int mista_download_delays(mista_handle* h, float* delays);

/* Test access to the delay-table columns of n grid points: delays[row][i]
 * = table[row][points[i]] for every row (2-D: the elements; 3-D: the active
 * rx, then the active tx). */
int mista_download_delays_at(mista_handle* h, const uint64_t* points, uint64_t n, float* delays);

/* Sizes of the current frame state: signal length L, FFT size, grid points. */
int mista_query(const mista_handle* h, uint32_t* signal_length, uint32_t* fft_size, uint64_t* num_points);

/* Select the DAS kernel variant (0 = automatic/tuned, 1 = reference-shaped
 * one-pixel-per-lane gather, 2 = LDS-tiled).  Tile parameters 0 = default;
 * tile_param1 = pairs per chunk (4, 8, or 16 on 512-lane tiles over 128
 * transmits) | ring buffers << 8 | LDS budget in KiB << 16.  Where the
 * variant's tiles do not fit, the rx-parallel form runs instead. */
int mista_set_kernel(mista_handle* h, int variant, int tile_param0, int tile_param1);

/* Frame graphs for mista_process_device (no counterpart in the reference:
 * for callers that replay the same frame shape).  With enable != 0, a device
 * frame that changes nothing the kernels depend on (same echo/output
 * pointers and sample count, no new allocation, table, tuning or setting) is
 * recorded once as a hipGraph on the handle's stream, and later frames with
 * the same key replay it with one hipGraphLaunch.  Only the fused
 * preparation path is captured.  After a replayed frame mista_last_timing
 * reports both stage times as -1 (not observable inside one graph launch).
 * mista_graph_stats counts captures and replays. */
int mista_set_graph(mista_handle* h, int enable);
int mista_graph_stats(const mista_handle* h, uint64_t* captures, uint64_t* launches);

/* Tile parameters the automatic variant chose on its first frame of the
 * current geometry (it times each tile shape and LDS budget once); -1 in
 * tile_param0 when nothing was tuned or no tile fits.  No
 * counterpart in the reference (its kernels have fixed launch shapes). */
int mista_tuned_kernel(const mista_handle* h, int* tile_param0, int* tile_param1);

/* The configuration of the last DAS launch: tile shape (-1 when the gather
 * kernel ran, -2 the rx-parallel form that runs where no tile fits: per-rx
 * sums a lane per (pixel, rx), then the statistics a lane per pixel) and
 * pairs per chunk | ring buffers << 8 | LDS budget in KiB << 16.  No
 * counterpart in the reference. */
int mista_last_kernel(const mista_handle* h, int* tile_param0, int* tile_param1);

/*
 * RF processor: replaces DefaultSTAProcessor (DefaultSTAProcessor.h:50-226;
 * constructor (config, acquisition, coherenceFactor, peakOffset)).  Echoes are
 * upsampled by 4 with Interpolator4X's fixed 81-tap filter (Interpolator4X.h),
 * DC removed, and summed as real samples without apodization; value is the
 * signed sum / (Ntx*Nrx) and factor the sign coherence factor of the per-rx
 * sums (cf_kind MISTA_CF_NONE or MISTA_CF_SCF; upsampling_factor 4 or 0;
 * storage ignored).  The envelope that STAMethod takes afterwards
 * (ParallelHilbertEnvelope::calculateDim2, STAMethod.cpp:63-72) is
 * mista_envelope / mista_envelope_device.  mista_upload/download_analytic
 * move the real tensor [ntx][nrx][4 ns] on an RF handle.
 */
int mista_create_rf(const mista_sta_config* config, mista_handle** out);

/* ParallelHilbertEnvelope::calculateDim2 (ParallelHilbertEnvelope.h:73-83)
 * with HilbertEnvelope::calculate (HilbertEnvelope.h:201-223): every row of
 * the n1 x n2 float matrix (n2 contiguous) replaced, in place, by the
 * magnitude of its analytic signal.  Any handle can run it. */
int mista_envelope_device(mista_handle* h, float* d_data, uint32_t n1, uint32_t n2, void* stream);
int mista_envelope(mista_handle* h, float* data, uint32_t n1, uint32_t n2);

/* Select the signal-preparation implementation: 0 = automatic (fused LDS
 * kernel when the FFT fits on chip, else the four-step FFTs), 1 = library FFT
 * (FIR kernel, hipFFT R2C, spectrum kernel, hipFFT C2C, store kernel),
 * 2 = fused LDS kernel (error if the frame does not fit), 3 = four-step FFTs
 * (FIR kernel, two-pass forward FFT, spectrum kernel, two-pass backward FFT,
 * store kernel; FFT sizes up to 2^20 whose two factors are 2-3-5 smooth). */
int mista_set_prep(mista_handle* h, int method);

/* Stage times of the last mista_das_device/mista_process_device frame (ms,
 * measured with HIP events on the launch stream): preparation and DAS; 0
 * when a stage did not run, -1 for both after a replayed frame graph. */
int mista_last_timing(const mista_handle* h, float* prep_ms, float* das_ms);

/* ---- serving engine (no reference counterpart) ----
 * A request loop for imaging jobs of uneven shapes on one GPU.  Jobs are
 * admitted as in-flight slots and device memory free up, in arrival order
 * except that a job that can start now may pass one that must wait (the
 * oldest job is passed at most 32 times).  Each runs on a resident processor
 * handle chosen by (configuration, apodization, grid, sample count), so delay
 * tables, window tables, the tuned tile shape and FFT plans stay in HBM
 * between jobs of the same shape; a shape holds up to handles_per_shape
 * handles with two jobs queued on each handle's stream, so one frame's
 * preparation overlaps the previous frame's DAS.  Handle creation, table
 * builds and tuning run on a builder thread, never on the admitting thread.
 * Admitted jobs run on separate streams and overlap on the device
 * (continuous batching).  A job's images do not depend on the other jobs in
 * flight: they are bit-identical to the same job run alone.
 * submit/wait/query may be called from any thread.
 */
typedef struct mista_engine mista_engine;

typedef struct mista_engine_config {
	int32_t device;
	uint32_t max_inflight;       /* jobs on the device at once, 1..64          */
	uint32_t max_queued;         /* jobs waiting for admission; more -> BUSY    */
	uint32_t handles_per_shape;  /* resident handles per shape, 0..8 (0 = 2)    */
	uint64_t hbm_budget;         /* device bytes the resident handles may hold
	                                (0 = 80 % of the free memory at creation)  */
} mista_engine_config;

/* One 2-D frame.  grid/apodization are copied at submit; echoes and outputs
 * are used in place and must stay valid until mista_engine_wait returns.
 * echoes: [ntx][nrx][ns] float32 on the device (echoes_on_device = 1) or on
 * the host; value/factor: [n1*n2] float32 on the device (outputs_on_device =
 * 1) or the host; factor may be NULL. */
typedef struct mista_job {
	mista_sta_config config;     /* device field ignored (the engine's device)  */
	const float* grid_xz;        /* host [n1][n2][2]                            */
	uint32_t n1, n2;
	const float* rx_apod;        /* host [num_elements]                         */
	const float* tx_apod;        /* host [ntx] or NULL                          */
	const float* echoes;
	uint32_t num_samples;
	int32_t echoes_on_device;
	float* value;
	float* factor;
	int32_t outputs_on_device;
	uint32_t pad_;
} mista_job;

typedef struct mista_job_stats {
	int32_t status;              /* MISTA_OK or the job's error code            */
	int32_t das_tile;            /* tile shape of its DAS launch (-1: gather, -2: rx-parallel) */
	double queued_ms;            /* submit -> admitted                          */
	double latency_ms;           /* submit -> outputs ready (seen by the loop)  */
	double device_ms;            /* job's work on its stream (HIP events)       */
	uint64_t footprint;          /* device bytes of the handle that ran it      */
} mista_job_stats;

int mista_engine_create(const mista_engine_config* config, mista_engine** out);
/* Finishes every submitted job, then frees all handles. */
int mista_engine_destroy(mista_engine* e);
/* Queues a job; MISTA_E_BUSY when max_queued jobs are already waiting. */
int mista_engine_submit(mista_engine* e, const mista_job* job, uint64_t* ticket);
/* Blocks until the job is done, fills stats, releases the ticket; returns the
 * job's status. */
int mista_engine_wait(mista_engine* e, uint64_t ticket, mista_job_stats* stats);
int mista_engine_query(mista_engine* e, uint64_t ticket, int* done);
int mista_engine_stats(mista_engine* e, uint64_t* resident_bytes, uint32_t* inflight, uint32_t* queued);
/* Totals since creation: handles built (each builds its delay and window
 * tables), autotuner runs, idle handles evicted; and handles being built now. */
int mista_engine_counters(mista_engine* e, uint64_t* handles_created, uint64_t* tunes, uint64_t* evictions,
		uint32_t* building);
const char* mista_engine_last_error(const mista_engine* e);

/* ---- synthetic echoes (test and benchmark input) ----
 * Point scatterers (x, z, amplitude) stamped as a Gaussian-modulated pulse
 * exp(-a t^2) cos(2 pi fc t) at the round-trip time from element e_tx to
 * e_rx, then white Gaussian noise of standard deviation noise_std from a
 * counter-based generator (seed).  Elements: x_e = e*pitch - (N-1)*pitch/2,
 * y = z = 0.  Output d_echoes [ntx][nrx][ns] float32 in HBM.
 * Deterministic for fixed inputs.  scat_y may be NULL (2-D). */
typedef struct mista_synth_params {
	uint32_t num_tx;
	uint32_t num_rx;
	uint32_t num_samples;
	uint32_t num_scatterers;
	float sampling_frequency;
	float propagation_speed;
	float center_frequency;
	float bandwidth;             /* fractional bandwidth at -6 dB              */
	float noise_std;             /* absolute; 0 = no noise                     */
	uint32_t pad_;
	uint64_t seed;
} mista_synth_params;

int mista_synth_echoes_device(const mista_synth_params* p,
		const float* tx_xyz /* [ntx][3] host */, const float* rx_xyz /* [nrx][3] host */,
		const float* scat_xyza /* [ns][4] host: x,y,z,amp */,
		float* d_echoes, void* stream);

/* Sum of squares of a device float array (for SNR scaling), double result. */
int mista_sum_squares_device(const float* d_x, uint64_t n, double* out, void* stream);

/* Achievable HBM read rate (measurement support, no reference
 * counterpart): times `reps` launches of a streaming-read kernel over the
 * device buffer d_x (bytes rounded down to 16) with HIP events and returns
 * GB/s.  Use a buffer well above the 256 MB last-level cache. */
int mista_read_bandwidth_device(const void* d_x, uint64_t bytes, int reps, double* gb_per_s, void* stream);

/* Add N(0, noise_std^2) noise from the counter-based generator to d_x. */
int mista_add_noise_device(float* d_x, uint64_t n, float noise_std, uint64_t seed, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MISTA_H */

/*
 * CPU oracle for the STA/TFM delay-and-sum path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference CPU processor
 * (src/user/processor/VectorialSTAProcessor.h:158-341) and of the DSP
 * routines it calls, in FP32 ("reference-equivalent", same loop order and
 * arithmetic form) and FP64 (accuracy oracle).  Only tests/, the smoke check
 * and bench.py's CPU-baseline leg may load it, and only as the reference the
 * GPU results are compared with or as the reported CPU baseline; the product
 * path never calls it.
 *
 * Every function cites the reference file:line it follows.  Which checks pin
 * which function is listed in the header of oracle.cpp.
 */
#ifndef MISTA_ORACLE_H
#define MISTA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Coherence factor selector.  PCF follows CoherenceFactor.h:240-296; SCF on
 * Re(per-rx sum) is an extension of CoherenceFactor.h:198-236 (the reference
 * applies SCF to real RF sums only, DefaultSTAProcessor.h:190-215). */
enum { ORACLE_CF_NONE = 0, ORACLE_CF_PCF = 1, ORACLE_CF_SCF = 2 };

/* 2-D STA parameters: STAConfiguration.h:118-165 plus the processor
 * constructor arguments of VectorialSTAProcessor.h:61-67. */
typedef struct oracle_sta2d {
	uint32_t num_elements;      /* numElements (rx count and element count)   */
	uint32_t first_tx_elem;     /* firstTxElem                                 */
	uint32_t last_tx_elem;      /* lastTxElem (inclusive)                      */
	uint32_t num_samples;       /* raw samples per channel (acqData.n2())      */
	uint32_t upsampling_factor; /* upsamplingFactor (1 = no interpolation)     */
	int32_t  cf_kind;           /* ORACLE_CF_*                                 */
	double pitch;               /* m                                           */
	double sampling_frequency;  /* Hz                                          */
	double propagation_speed;   /* m/s                                         */
	double center_frequency;    /* Hz                                          */
	double dead_zone_m;         /* m                                           */
	double peak_offset;         /* in periods of center_frequency              */
	double cf_param;            /* PCF gamma or SCF p                          */
} oracle_sta2d;

/* ---- DSP building blocks (float64 and float32 variants) ---- */
double oracle_kaiser_beta(double tol_db);
uint32_t oracle_kaiser_size(double tol_db, double transition_width);
void oracle_kaiser_window(uint32_t size, double beta, double* out);
double oracle_bessel_i0(double x);
uint32_t oracle_next_fast_even(uint32_t n);
uint32_t oracle_upsampling_filter_size(uint32_t up, double half_tw);
void oracle_upsampling_filter_f64(uint32_t up, double half_tw, double* h);
void oracle_upsampling_filter_f32(uint32_t up, float half_tw, float* h);
/* Interpolator::interpolate: out has n*up samples. */
void oracle_interpolate_f64(const double* x, size_t n, uint32_t up, double half_tw, double* out);
void oracle_interpolate_f32(const float* x, size_t n, uint32_t up, float half_tw, float* out);
/* full linear convolution y = x*b (length nx+nb-1): FFTWFilter::filter semantics */
void oracle_convolve_f64(const double* x, size_t nx, const double* b, size_t nb, double* y);
/* forward DFT of a real sequence (full length, complex interleaved out) */
void oracle_fft_real_f64(const double* x, uint32_t n, double* out_re, double* out_im);
/* inverse DFT, unnormalised: out[k] = sum X[j] e^{+2pi i jk/n} */
void oracle_ifft_f64(const double* in_re, const double* in_im, uint32_t n, double* out_re, double* out_im);
void oracle_analytic_signal_f64(const double* x, uint32_t n, double* out /* [n][2] */);
void oracle_analytic_signal_f32(const float* x, uint32_t n, float* out /* [n][2] */);
void oracle_remove_dc_f64(double* x, size_t n, size_t begin);
void oracle_remove_dc_f32(float* x, size_t n, size_t begin);
double oracle_std_dev_f64(const double* x, size_t n);
double oracle_mean_f64(const double* x, size_t n);
double oracle_pcf_f64(const double* re_im /* [n][2] */, uint32_t n, double gamma);
float  oracle_pcf_f32(const float* re_im, uint32_t n, float gamma);
double oracle_scf_f64(const double* x, uint32_t n, double p);
float  oracle_scf_f32(const float* x, uint32_t n, float p);
/* WindowFunction::get names: rectangular, hanning, hamming, blackman, ... */
int oracle_window(const char* name, uint32_t n, double* out);

/* ---- STA path ---- */
/* Signal preparation for one frame: raw echoes [ntx][nrx][ns] float32 ->
 * analytic tensor [ntx][nrx][ns*up][2].  VectorialSTAProcessor.h:169-231. */
int oracle_prepare_f32(const oracle_sta2d* cfg, const float* echoes, float* analytic, int nthreads);
int oracle_prepare_f64(const oracle_sta2d* cfg, const float* echoes, double* analytic, int nthreads);

/* Delay-and-sum over a prepared analytic tensor.  grid_xz is [n1][n2][2]
 * (x, z) in m, n1 = columns (x), n2 = rows (z), z fastest (ImageGrid.h:137-143).
 * rx_apod has num_elements weights; tx_apod may be NULL (all ones, the
 * reference behaviour).  value/factor are [n1*n2].
 * VectorialSTAProcessor.h:237-338. */
int oracle_das_f32(const oracle_sta2d* cfg, const float* analytic, uint32_t signal_length,
		const float* grid_xz, uint32_t n1, uint32_t n2,
		const float* rx_apod, const float* tx_apod,
		float* value, float* factor, int nthreads);
/* Same as oracle_das_f32 with a coherence factor, also returning the per-rx
 * complex sums [n1*n2][num_elements][2] (test diagnostics). */
int oracle_das_rxsums_f32(const oracle_sta2d* cfg, const float* analytic, uint32_t signal_length,
		const float* grid_xz, uint32_t n1, uint32_t n2,
		const float* rx_apod, const float* tx_apod,
		float* value, float* factor, float* rxsums, int nthreads);
int oracle_das_f64(const oracle_sta2d* cfg, const double* analytic, uint32_t signal_length,
		const float* grid_xz, uint32_t n1, uint32_t n2,
		const float* rx_apod, const float* tx_apod,
		double* value, double* factor, int nthreads);

/* Delay-and-sum over a caller-supplied FP32 delay table delays[rows][npix]
 * (samples; tx t at row tx_row0 + t, rx r at row r), the input the
 * reference GPU processors use (VectorialSTACUDAProcessor.cu:30-53).  form 0
 * keeps the CPU processor's lerp and unfused apodized sum
 * (VectorialSTAProcessor.h:269-284); form 1 the CUDA kernels'
 * v0 + k (v1 - v0) and sum += v * apod as fused multiply-adds
 * (VectorialSTACUDAProcessor.cu:93-103, :163-169).  rxsums (optional, needs a
 * coherence factor) receives the per-rx complex sums [npix][num_elements][2]. */
/* The FP32 delay table [num_elements][n1*n2] that oracle_das_f32 computes
 * per pixel (distance2DY0 * invCT, Geometry.h:55-65). */
int oracle_delays_f32(const oracle_sta2d* cfg, const float* grid_xz, uint32_t n1, uint32_t n2, float* delays);
int oracle_das_table_f32(const oracle_sta2d* cfg, const float* analytic, uint32_t signal_length,
		const float* delays, uint32_t rows, uint32_t tx_row0, uint64_t npix, const float* rx_apod,
		const float* tx_apod, float* value, float* factor, float* rxsums, int form, int nthreads);

/* Whole frame (prepare + DAS), timing of the two stages in seconds. */
int oracle_process_f32(const oracle_sta2d* cfg, const float* echoes,
		const float* grid_xz, uint32_t n1, uint32_t n2,
		const float* rx_apod, const float* tx_apod,
		float* value, float* factor, int nthreads, double* t_prep, double* t_das);

/* 3-D sparse STA: Vectorial3DSTAProcessor.h:137-253 with distance3DZ0.
 * echoes [ntx_active][nrx_active][ns]; elem positions (x,y) per active
 * element; grid_xyz [npts][3]. */
typedef struct oracle_sta3d {
	uint32_t num_tx;            /* active tx count                             */
	uint32_t num_rx;            /* active rx count                             */
	uint32_t num_samples;
	uint32_t upsampling_factor;
	int32_t  cf_kind;
	uint32_t pad_;
	double sampling_frequency;
	double propagation_speed;
	double center_frequency;
	double dead_zone_m;
	double peak_offset;
	double cf_param;
} oracle_sta3d;

int oracle_prepare3d_f32(const oracle_sta3d* cfg, const float* echoes, float* analytic, int nthreads);
int oracle_das3d_f32(const oracle_sta3d* cfg, const float* analytic, uint32_t signal_length,
		const float* tx_xy, const float* rx_xy,
		const float* grid_xyz, uint32_t n1, uint32_t n2,
		const float* rx_apod, float* value, float* factor, int nthreads);
int oracle_das3d_rxsums_f32(const oracle_sta3d* cfg, const float* analytic, uint32_t signal_length,
		const float* tx_xy, const float* rx_xy,
		const float* grid_xyz, uint32_t n1, uint32_t n2,
		const float* rx_apod, float* value, float* factor, float* rxsums, int nthreads);
int oracle_das3d_f64(const oracle_sta3d* cfg, const double* analytic, uint32_t signal_length,
		const float* tx_xy, const float* rx_xy,
		const float* grid_xyz, uint32_t n1, uint32_t n2,
		const float* rx_apod, double* value, double* factor, int nthreads);

/* RF path (DefaultSTAProcessor.h:136-226, Interpolator4X.h,
 * ParallelHilbertEnvelope.h:73-83).  Upsampling is fixed at 4; the tensor is
 * real [ntx][nrx][4 ns]. */
int oracle_interp4x_filter_f64(double* h /* [81] */);
int oracle_interpolate4x_f64(const double* x, uint32_t n, double* out /* [4 n] */);
int oracle_prepare_rf_f32(const oracle_sta2d* cfg, const float* echoes, float* out, int nthreads);
int oracle_prepare_rf_f64(const oracle_sta2d* cfg, const float* echoes, double* out, int nthreads);
int oracle_das_rf_f32(const oracle_sta2d* cfg, const float* tensor, uint32_t signal_length, const float* grid_xz,
		uint32_t n1, uint32_t n2, float* value, float* factor, int nthreads);
int oracle_das_rf_rxsums_f32(const oracle_sta2d* cfg, const float* tensor, uint32_t signal_length, const float* grid_xz,
		uint32_t n1, uint32_t n2, float* value, float* factor, float* rxsums, int nthreads);
int oracle_das_rf_f64(const oracle_sta2d* cfg, const double* tensor, uint32_t signal_length, const float* grid_xz,
		uint32_t n1, uint32_t n2, double* value, double* factor, int nthreads);
int oracle_envelope_dim2_f32(float* data, uint32_t n1, uint32_t n2, int nthreads);
int oracle_envelope_dim2_f64(double* data, uint32_t n1, uint32_t n2, int nthreads);

const char* oracle_last_error(void);

#ifdef __cplusplus
}
#endif

#endif

// Internal declarations shared by the mista HIP sources.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hipfft/hipfft.h>

#include <atomic>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mista.h"

namespace mista {

// Error carried to the C ABI as a status code + message.
struct Error : std::runtime_error {
	int code;
	Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MISTA_THROW(code, msg) do { std::ostringstream os_; os_ << msg; throw ::mista::Error(code, os_.str()); } while (0)

#define HIP_CHECK(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) \
	MISTA_THROW(MISTA_E_HIP, #expr << ": " << hipGetErrorString(e_) << " (" << __FILE__ << ":" << __LINE__ << ")"); } while (0)

#define FFT_CHECK(expr) do { hipfftResult r_ = (expr); if (r_ != HIPFFT_SUCCESS) \
	MISTA_THROW(MISTA_E_HIP, #expr << ": hipfft error " << static_cast<int>(r_) << " (" << __FILE__ << ":" << __LINE__ << ")"); } while (0)

// Device buffer owned by a handle.  Reallocates only when growing.
// Count of device allocations made by DeviceBuffer (process-wide): a change
// tells a captured frame graph that a buffer it uses may have moved.
inline std::atomic<uint64_t>& alloc_epoch()
{
	static std::atomic<uint64_t> e{0};
	return e;
}

class DeviceBuffer {
public:
	DeviceBuffer() = default;
	~DeviceBuffer() { release(); }
	DeviceBuffer(const DeviceBuffer&) = delete;
	DeviceBuffer& operator=(const DeviceBuffer&) = delete;
	void* reserve(size_t bytes)
	{
		if (bytes > cap_) {
			release();
			if (bytes == 0) return nullptr;
			hipError_t e = hipMalloc(&ptr_, bytes);
			if (e != hipSuccess) {
				ptr_ = nullptr;
				MISTA_THROW(MISTA_E_NO_MEMORY, "hipMalloc(" << bytes << " bytes) failed: " << hipGetErrorString(e));
			}
			cap_ = bytes;
			++alloc_epoch();
		}
		return ptr_;
	}
	void release()
	{
		if (ptr_) (void)hipFree(ptr_);
		ptr_ = nullptr;
		cap_ = 0;
	}
	void* get() const { return ptr_; }
	size_t capacity() const { return cap_; }
	template<typename T> T* as() const { return static_cast<T*>(ptr_); }
private:
	void* ptr_ = nullptr;
	size_t cap_ = 0;
};

// Geometry of one frame, everything the kernels need by value.  The delay
// table has num_elements rows: rx r reads row r, tx t reads row first_tx + t
// (2-D: one row per array element; 3-D: active rx rows, then active tx rows).
struct FrameGeom {
	uint32_t num_elements;  // rows of the delay table
	uint32_t first_tx;      // row of transmit 0
	uint32_t ntx;
	uint32_t ntx_pad;   // transmits padded to 32/64/128 (multiple of 8 above)
	uint32_t nrx;
	uint32_t ns;        // raw samples per channel
	uint32_t up;        // upsampling factor
	uint32_t L;         // signal length = ns * up
	uint32_t nf;        // Hilbert FFT size
	uint32_t dz_up;     // dead zone, upsampled samples
	uint32_t n1, n2;    // grid
	uint32_t ny;        // volume grids: n1 = nx * ny with y fastest (1 for a plane)
	uint64_t npix;
	float offset;       // signalOffset_ (samples)
	double inv_ct;      // (fs*U)/c, samples per metre
	double pitch;
	int cf_kind;
	float cf_param;
	int storage;
	uint64_t grid_version;  // bumped whenever the grid / delay table changes
};

// Size helpers (FFTUtil.h:100-135, HilbertEnvelope.h:127).
uint32_t next_fast_even(uint32_t n);
inline uint32_t hilbert_fft_size(uint32_t n)
{
	const uint32_t a = next_fast_even(n + 1024u);
	return a > 2048u ? a : 2048u;
}
// Interpolator.h:64-111 low-pass design (double precision).
std::vector<double> upsampling_filter(uint32_t up, double half_tw);

// Transmit count padded to the DAS kernel's instantiated sizes.
inline uint32_t pad_ntx(uint32_t ntx)
{
	if (ntx <= 32) return 32;
	if (ntx <= 64) return 64;
	if (ntx <= 128) return 128;
	// above 128: ceil(ntx / 128) equal blocks of whole 8-pair chunks, run in turn
	const uint32_t nblk = (ntx + 127) / 128;
	const uint32_t bt = ((ntx + nblk - 1) / nblk + 7) / 8 * 8;
	return nblk * bt;
}

// Bytes per complex sample of a storage kind.
// Real (RF) samples, float32: the DefaultSTAProcessor path (library-internal
// storage kind; the C ABI selects it with mista_create_rf).
constexpr int kStorageReal = 3;
inline size_t storage_bytes(int storage) { return storage == MISTA_STORAGE_F32 ? 8 : 4; }
// F16/BF16 carry a per-channel scale folded into the pair weights.
inline bool storage_scaled(int storage) { return storage == MISTA_STORAGE_F16 || storage == MISTA_STORAGE_BF16; }

// ---- kernels launchers (defined in the .hip files) ----
void launch_delay_table(const float* d_grid_xz, uint64_t npix, uint32_t num_elements, double pitch,
		double inv_ct, float* d_delay, hipStream_t s);
void launch_delay_table_3d(const float* d_grid_xyz, uint64_t npix, const double* d_elem_xy, uint32_t rows,
		double inv_ct, float* d_delay, hipStream_t s);

struct PrepPlan;  // FFT plans + scratch for signal preparation
// Batched four-step FFTs of long 2-3-5 smooth sizes (fft4.hip).
struct Fft4Plan;
Fft4Plan* fft4_plan(uint32_t n, hipStream_t s);
void fft4_destroy(Fft4Plan* p);
uint32_t fft4_size(const Fft4Plan* p);
size_t fft4_device_bytes(const Fft4Plan* p);
// ymax (optional): per transform, atomic max of the bits of the largest
// |Re X[k]|, |Im X[k]| over k < npair (and |Re X[npair]| when odd); zeroed by
// the caller beforehand.
void fft4_exec(const Fft4Plan& p, const float2* in, float2* tmp, float2* out, uint32_t batch, hipStream_t s,
		uint32_t* ymax = nullptr, uint32_t npair = 0, bool odd = false);
struct PrepState {
	PrepPlan* plan = nullptr;  // owned; destroyed in prep.hip where PrepPlan is complete
	int method = 0;            // 0 automatic (fused kernel, else four-step FFTs), 1 library FFT (hipFFT), 2 fused kernel, 3 four-step FFTs
	bool last_fused = false;   // the last preparation ran the fused kernel
	// called before the kernels of channels [c0, c0 + n) are launched (the
	// host path uploads each chunk of echoes just in time)
	std::function<void(uint32_t c0, uint32_t n)> before_chunk;
	PrepState() = default;
	PrepState(const PrepState&) = delete;
	PrepState& operator=(const PrepState&) = delete;
	~PrepState();
};
void prepare_signals(PrepState& st, const FrameGeom& g, const float* d_raw, const float* d_filter_poly,
		uint32_t taps_per_phase, uint32_t filter_offset, void* d_tensor, float* d_scale, hipStream_t s);
// RF preparation (DefaultSTAProcessor.h:136-164): FIR x 4 + DC removal into a
// real float tensor [ntx][nrx][L].
void prepare_signals_rf(const FrameGeom& g, const float* d_raw, const float* d_filter_poly, uint32_t taps_per_phase,
		uint32_t filter_offset, float* d_tensor, hipStream_t s);
// Interpolator4X.h:54-72 filter (81 taps), from its recipe.
std::vector<double> rf4x_filter();
// Row envelope (HilbertEnvelope::calculate on every row of an n1 x n2
// matrix, in place).
void envelope_rows(PrepState& st, float* d_data, uint32_t n1, uint32_t n2, hipStream_t s);

// Device bytes held by the signal-preparation state (plans' buffers and an
// estimate of the FFT library's work area).
size_t prep_device_bytes(const PrepState& st);

void launch_weights(const FrameGeom& g, const float* d_rx_apod, const float* d_tx_apod, const float* d_scale,
		float* d_w, hipStream_t s);

struct DasTuning {
	int variant = 0;  // 0 auto, 1 gather, 2 tiled
	int p0 = 0, p1 = 0;
};
struct DasScratch {
	DeviceBuffer partial;           // small device scalars (max window slots)
	DeviceBuffer bounds;            // per-tile element delay bounds
	DeviceBuffer params;            // per (tile, rx, tx) window words
	DeviceBuffer partials;          // per-rx sums between transmit blocks (more than 128 transmits)
	uint64_t bounds_version = ~0ull;
	uint64_t params_key = 0;
	float params_offset = 0.0f;
	uint32_t max_slots = 0;         // 1 KiB load slots per staged window
	uint64_t tuned_key = ~0ull;     // geometry the autotuner last chose for
	int tuned_p0 = -1, tuned_p1 = 0;
	int last_p0 = -1, last_p1 = 0;  // the last DAS launch (-1: gather kernel)
};
// Bytes of zeroed padding after the analytic tensor (window over-read).
size_t tensor_tail_padding();
// Device scratch of the no-allocation DAS entry point (per GPU, static).
constexpr size_t kNoAllocScratchBytes = static_cast<size_t>(80) << 20;
// DAS over the handle's resident tensor with its window tables in that
// scratch: no device allocation; the tuned tile shape when there is one.
void launch_das_noalloc(const FrameGeom& g, int tuned_p0, int tuned_p1, DasScratch& last, const void* d_tensor,
		const float* d_delay, const float* d_w, float* d_value, float* d_factor, hipStream_t s);
// One frame of a DAS batch.
struct DasBatchItem {
	const FrameGeom* g;
	const DasTuning* t;
	DasScratch* scratch;
	const void* tensor;
	const float* delay;
	const float* w;
	float* value;
	float* factor;
};
// The DAS of several prepared frames: frames whose tiled launches agree in
// storage, coherence factor, transmit block and workgroup size (and whose
// geometry has been tuned) share one launch of the tiled kernel, up to 16 a
// launch; the rest run one by one.  Returns the number of shared launches.
uint32_t launch_das_batch(const std::vector<DasBatchItem>& items, hipStream_t s);
void launch_das(const FrameGeom& g, const DasTuning& t, DasScratch& scratch, const void* d_tensor,
		const float* d_delay, const float* d_w, float* d_value, float* d_factor, hipStream_t s);

void launch_tensor_convert_in(const FrameGeom& g, const float* d_f32, void* d_tensor, float* d_scale, hipStream_t s);
void launch_tensor_convert_out(const FrameGeom& g, const void* d_tensor, const float* d_scale, float* d_f32, hipStream_t s);

// Serving-engine access to a processor handle (defined in api.cpp).
hipStream_t handle_stream(mista_handle* h);
size_t handle_device_bytes(const mista_handle* h);

}  // namespace mista

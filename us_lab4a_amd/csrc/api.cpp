// C ABI of the mista library (include/mista.h): the STA array-processor
// contract of the reference (ArrayProcessor.h:27-40,
// VectorialSTAProcessor.h:58-341) over HIP device buffers.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mista {
void synth_echoes(const mista_synth_params& p, const float* tx_xyz, const float* rx_xyz, const float* scat,
		float* d_echoes, hipStream_t s);
double sum_squares(const float* d_x, uint64_t n, hipStream_t s);
void add_noise(float* d_x, uint64_t n, float sigma, uint64_t seed, hipStream_t s);
double read_bandwidth(const void* d_x, uint64_t bytes, int reps, hipStream_t s);
}  // namespace mista

using namespace mista;

namespace {
thread_local std::string g_thread_error;
}

struct mista_handle {
	mista_sta_config cfg{};
	hipStream_t stream = nullptr;
	std::string last_error;
	uint32_t base_element = 0;

	// apodization (host copies + device copies)
	std::vector<float> rx_apod, tx_apod;
	bool tx_apod_set = false;
	DeviceBuffer d_rx_apod, d_tx_apod, d_w;
	bool weights_dirty = true;

	// delay-table rows: 2-D one per array element (tx t at row first_tx + t);
	// 3-D active rx elements, then active tx elements (positions in elem_xy)
	bool is3d = false;
	bool rf = false;               // DefaultSTAProcessor: real samples, fixed 4x filter
	PrepState env;                 // row-envelope FFT plan (its own size)
	uint32_t rows = 0, tx_row0 = 0;
	std::vector<double> elem_xy;   // 3-D: [rows][2]
	DeviceBuffer d_elem;

	// grid + resident delay table
	uint32_t n1 = 0, n2 = 0, ny = 1;
	uint64_t grid_version = 0;
	std::vector<float> grid_xyz;   // host copy, [n1*n2][3]
	DeviceBuffer d_grid, d_delay;

	// frame state
	FrameGeom geom{};
	bool signal_ready = false;     // tensor holds a prepared frame
	DeviceBuffer d_raw, d_tensor, d_scale, d_hpoly, d_tmp, d_value, d_factor;
	uint32_t filter_J = 0, filter_off = 0, filter_up = 0;
	PrepState prep;
	DasTuning tuning;
	DasScratch das_scratch;

	// timing
	hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
	bool timed = false;
	bool graph_frame = false;      // the last frame was a graph replay

	// frame graph (mista_set_graph): the kernels of a steady device frame,
	// captured once and replayed while nothing they depend on changes
	bool graph_on = false;
	bool capturing = false;        // launch-only frame path: no memset, no events
	uint64_t settings_epoch = 0;   // bumped by every setter
	struct GraphKey {
		const void* echoes = nullptr;
		const void* value = nullptr;
		const void* factor = nullptr;
		uint32_t ns = 0;
		uint64_t settings = 0, allocs = 0, grid = 0, tuned = 0;
		int p0 = 0, p1 = 0;
		bool operator==(const GraphKey& o) const
		{
			return echoes == o.echoes && value == o.value && factor == o.factor && ns == o.ns &&
					settings == o.settings && allocs == o.allocs && grid == o.grid && tuned == o.tuned &&
					p0 == o.p0 && p1 == o.p1;
		}
	} gkey;
	hipGraphExec_t gexec = nullptr;
	GraphKey gfail;                // key whose recording failed
	bool gfail_valid = false;
	uint64_t graph_launches = 0, graph_captures = 0;

	void drop_graph()
	{
		if (gexec) (void)hipGraphExecDestroy(gexec);
		gexec = nullptr;
	}

	// host path: echoes uploaded chunk by chunk on their own stream
	hipStream_t copy_stream = nullptr;
	std::vector<hipEvent_t> chunk_ev;

	~mista_handle()
	{
		drop_graph();
		for (auto& e : ev)
			if (e) (void)hipEventDestroy(e);
		for (auto& e : chunk_ev) (void)hipEventDestroy(e);
		if (copy_stream) (void)hipStreamDestroy(copy_stream);
		if (stream) (void)hipStreamDestroy(stream);
	}
};

namespace {

template<typename F>
int guarded(mista_handle* h, F&& f)
{
	try {
		if (h) HIP_CHECK(hipSetDevice(h->cfg.device));
		f();
		return MISTA_OK;
	} catch (const Error& e) {
		if (h) h->last_error = e.what();
		g_thread_error = e.what();
		return e.code;
	} catch (const std::bad_alloc&) {
		if (h) h->last_error = "host allocation failed";
		g_thread_error = "host allocation failed";
		return MISTA_E_NO_MEMORY;
	} catch (const std::exception& e) {
		if (h) h->last_error = e.what();
		g_thread_error = e.what();
		return MISTA_E_INVALID_STATE;
	}
}

hipStream_t pick_stream(mista_handle* h, void* stream)
{
	return stream ? static_cast<hipStream_t>(stream) : h->stream;
}

void validate(const mista_sta_config& c)
{
	if (c.num_elements == 0) MISTA_THROW(MISTA_E_INVALID_PARAM, "The number of elements is zero.");
	if (c.last_tx_elem < c.first_tx_elem || c.last_tx_elem >= c.num_elements)
		MISTA_THROW(MISTA_E_INVALID_PARAM, "Invalid tx element range [" << c.first_tx_elem << ", " << c.last_tx_elem
				<< "] for " << c.num_elements << " elements.");
	if (c.upsampling_factor < 1 || c.upsampling_factor > 64)
		MISTA_THROW(MISTA_E_INVALID_VALUE, "Invalid upsampling factor: " << c.upsampling_factor << ". Must be in [1, 64].");
	if (!(c.sampling_frequency > 0) || !(c.propagation_speed > 0) || !(c.center_frequency > 0))
		MISTA_THROW(MISTA_E_INVALID_PARAM, "sampling_frequency, propagation_speed and center_frequency must be > 0.");
	if (!(c.pitch > 0)) MISTA_THROW(MISTA_E_INVALID_PARAM, "pitch must be > 0.");
	if (!(c.dead_zone_m >= 0)) MISTA_THROW(MISTA_E_INVALID_PARAM, "dead_zone_m must be >= 0.");
	if (!(c.peak_offset >= 0)) MISTA_THROW(MISTA_E_INVALID_PARAM, "peak_offset must be >= 0.");
	if (c.cf_kind != MISTA_CF_NONE && c.cf_kind != MISTA_CF_PCF && c.cf_kind != MISTA_CF_SCF)
		MISTA_THROW(MISTA_E_INVALID_PARAM, "Invalid coherence factor method: " << c.cf_kind << '.');
	if (c.cf_kind != MISTA_CF_NONE && !(c.cf_param >= 0 && c.cf_param <= 100))
		MISTA_THROW(MISTA_E_INVALID_PARAM, "coherence factor parameter out of [0, 100].");
	if (c.storage != MISTA_STORAGE_F32 && c.storage != MISTA_STORAGE_F16 && c.storage != MISTA_STORAGE_BF16)
		MISTA_THROW(MISTA_E_INVALID_PARAM, "Invalid storage type: " << c.storage << '.');
}

uint32_t ntx_of(const mista_sta_config& c) { return c.last_tx_elem - c.first_tx_elem + 1U; }

// Dead zone in upsampled samples, VectorialSTAProcessor.h:138 (float
// configuration: fs*U in float, the rest in double).
uint32_t dead_zone_up(const mista_sta_config& c)
{
	const float fsu = static_cast<float>(c.upsampling_factor) * c.sampling_frequency;
	return static_cast<uint32_t>(fsu * 2.0 * static_cast<double>(c.dead_zone_m) / static_cast<double>(c.propagation_speed));
}

// Shapes of a frame with signal length L (ns raw samples; ns == 0 when the
// tensor is uploaded directly).
void set_frame_geom(mista_handle* h, uint32_t ns, uint32_t L, hipStream_t s)
{
	const mista_sta_config& c = h->cfg;
	FrameGeom& g = h->geom;
	g.num_elements = h->rows;
	g.first_tx = h->tx_row0;
	g.ntx = ntx_of(c);
	g.ntx_pad = pad_ntx(g.ntx);
	g.nrx = c.num_elements;
	g.ns = ns;
	g.up = c.upsampling_factor;
	g.L = L;
	g.dz_up = dead_zone_up(c);
	if (g.dz_up >= L)
		MISTA_THROW(MISTA_E_INVALID_VALUE, "Invalid dead zone: deadZoneSamplesUp (" << g.dz_up << ") >= signalLength (" << L << ").");
	if (L < 2) MISTA_THROW(MISTA_E_INVALID_PARAM, "signal length must be >= 2.");
	g.nf = hilbert_fft_size(L);
	g.n1 = h->n1;
	g.n2 = h->n2;
	g.ny = h->ny;
	g.npix = static_cast<uint64_t>(h->n1) * h->n2;
	g.grid_version = h->grid_version;
	// VectorialSTAProcessor.h:147 (float): (fs*U) * peakOffset / fc
	g.offset = (c.sampling_frequency * static_cast<float>(c.upsampling_factor)) * c.peak_offset / c.center_frequency;
	g.inv_ct = (static_cast<double>(c.sampling_frequency) * c.upsampling_factor) / static_cast<double>(c.propagation_speed);
	g.pitch = c.pitch;
	g.cf_kind = c.cf_kind;
	g.cf_param = c.cf_param;
	g.storage = c.storage;
	const size_t nch = static_cast<size_t>(g.ntx) * g.nrx;
	const size_t tbytes = nch * L * storage_bytes(c.storage);
	unsigned char* t = static_cast<unsigned char*>(h->d_tensor.reserve(tbytes + tensor_tail_padding()));
	// the staged-window copies may read into the padding: keep it zero (a
	// captured frame graph reuses a tail that an eager frame has cleared)
	if (!h->capturing) HIP_CHECK(hipMemsetAsync(t + tbytes, 0, tensor_tail_padding(), s));
	h->d_scale.reserve(nch * sizeof(float));
	h->d_w.reserve(static_cast<size_t>(g.ntx_pad) * g.nrx * sizeof(float));
	// apodization buffers too, so that the DAS of a prepared frame allocates nothing
	h->d_rx_apod.reserve(static_cast<size_t>(g.nrx) * sizeof(float));
	h->d_tx_apod.reserve(static_cast<size_t>(g.ntx) * sizeof(float));
}

void ensure_filter(mista_handle* h, hipStream_t s)
{
	const uint32_t up = h->cfg.upsampling_factor;
	if (up < 2 || h->filter_up == up) return;
	// RF: Interpolator4X's fixed 81-tap filter; otherwise the Kaiser design
	// with VectorialSTAProcessor.h:92 upsampFilterHalfTransitionWidth = 0.2
	const std::vector<double> taps = h->rf ? rf4x_filter() : upsampling_filter(up, 0.2);
	const uint32_t T = static_cast<uint32_t>(taps.size());
	const uint32_t J = (T + up - 1) / up;
	std::vector<float> poly(static_cast<size_t>(up) * J, 0.0f);
	for (uint32_t ph = 0; ph < up; ++ph)
		for (uint32_t j = 0; j < J; ++j)
			if (ph + up * j < T) poly[static_cast<size_t>(ph) * J + j] = static_cast<float>(taps[ph + up * j]);
	void* d = h->d_hpoly.reserve(poly.size() * sizeof(float));
	HIP_CHECK(hipMemcpyAsync(d, poly.data(), poly.size() * sizeof(float), hipMemcpyHostToDevice, s));
	HIP_CHECK(hipStreamSynchronize(s));
	h->filter_J = J;
	h->filter_off = (T - 1) / 2;
	h->filter_up = up;
}

void require_grid(mista_handle* h)
{
	if (h->n1 == 0 || h->n2 == 0) MISTA_THROW(MISTA_E_INVALID_STATE, "The image grid has not been set.");
}

void upload_apod(mista_handle* h, hipStream_t s)
{
	if (!h->weights_dirty) return;
	const size_t N = h->cfg.num_elements;
	HIP_CHECK(hipMemcpyAsync(h->d_rx_apod.reserve(N * sizeof(float)), h->rx_apod.data(), N * sizeof(float),
			hipMemcpyHostToDevice, s));
	if (h->tx_apod_set) {
		const size_t T = h->tx_apod.size();
		HIP_CHECK(hipMemcpyAsync(h->d_tx_apod.reserve(T * sizeof(float)), h->tx_apod.data(), T * sizeof(float),
				hipMemcpyHostToDevice, s));
	}
	HIP_CHECK(hipStreamSynchronize(s));
	h->weights_dirty = false;
}

void ensure_events(mista_handle* h)
{
	for (auto& e : h->ev)
		if (!e) HIP_CHECK(hipEventCreate(&e));
}

void run_prepare(mista_handle* h, const float* d_echoes, uint32_t ns, hipStream_t s)
{
	require_grid(h);
	if (ns == 0) MISTA_THROW(MISTA_E_INVALID_PARAM, "The number of input samples is 0.");
	const uint64_t L64 = static_cast<uint64_t>(ns) * h->cfg.upsampling_factor;
	if (L64 > (1u << 20)) MISTA_THROW(MISTA_E_INVALID_PARAM, "The number of input samples (" << L64 << ") is greater than 1048576.");
	set_frame_geom(h, ns, static_cast<uint32_t>(L64), s);
	ensure_filter(h, s);
	ensure_events(h);
	if (!h->capturing) HIP_CHECK(hipEventRecord(h->ev[0], s));
	if (h->rf)
		prepare_signals_rf(h->geom, d_echoes, h->d_hpoly.as<const float>(), h->filter_J, h->filter_off,
				h->d_tensor.as<float>(), s);
	else
		prepare_signals(h->prep, h->geom, d_echoes, h->d_hpoly.as<const float>(), h->filter_J, h->filter_off,
				h->d_tensor.get(), h->d_scale.as<float>(), s);
	if (!h->capturing) HIP_CHECK(hipEventRecord(h->ev[1], s));
	h->signal_ready = true;
}

void run_das(mista_handle* h, float* d_value, float* d_factor, hipStream_t s)
{
	require_grid(h);
	if (!h->signal_ready) MISTA_THROW(MISTA_E_INVALID_STATE, "No prepared signals: call a prepare/process/upload entry point first.");
	// the grid may have changed since the signals were prepared
	h->geom.n1 = h->n1;
	h->geom.n2 = h->n2;
	h->geom.ny = h->ny;
	h->geom.npix = static_cast<uint64_t>(h->n1) * h->n2;
	h->geom.grid_version = h->grid_version;
	upload_apod(h, s);
	ensure_events(h);
	// w[tx][rx] = rxApod * txApod * per-channel storage scale
	launch_weights(h->geom, h->d_rx_apod.as<const float>(), h->tx_apod_set ? h->d_tx_apod.as<const float>() : nullptr,
			h->d_scale.as<const float>(), h->d_w.as<float>(), s);
	if (!h->capturing) HIP_CHECK(hipEventRecord(h->ev[2], s));
	launch_das(h->geom, h->tuning, h->das_scratch, h->d_tensor.get(), h->d_delay.as<const float>(), h->d_w.as<const float>(),
			d_value, d_factor, s);
	if (!h->capturing) HIP_CHECK(hipEventRecord(h->ev[3], s));
	if (!h->capturing) {   // a recording keeps the eager frame's times
		h->timed = true;
		h->graph_frame = false;
	}
}

// Builds the resident delay table from h->grid_xyz and the element rows.
void build_delays(mista_handle* h)
{
	hipStream_t s = h->stream;
	const uint64_t npix = static_cast<uint64_t>(h->n1) * h->n2;
	const double inv_ct = (static_cast<double>(h->cfg.sampling_frequency) * h->cfg.upsampling_factor) /
			static_cast<double>(h->cfg.propagation_speed);
	float* d_delay = static_cast<float*>(h->d_delay.reserve(npix * h->rows * sizeof(float)));
	if (h->is3d) {
		void* d_grid = h->d_grid.reserve(npix * 3 * sizeof(float));
		HIP_CHECK(hipMemcpyAsync(d_grid, h->grid_xyz.data(), npix * 3 * sizeof(float), hipMemcpyHostToDevice, s));
		void* d_elem = h->d_elem.reserve(h->elem_xy.size() * sizeof(double));
		HIP_CHECK(hipMemcpyAsync(d_elem, h->elem_xy.data(), h->elem_xy.size() * sizeof(double), hipMemcpyHostToDevice, s));
		launch_delay_table_3d(static_cast<const float*>(d_grid), npix, static_cast<const double*>(d_elem), h->rows, inv_ct,
				d_delay, s);
	} else {
		// Geometry::distance2DY0 (Geometry.h:55-65): y of the points is not used
		std::vector<float> xz(npix * 2);
		for (uint64_t i = 0; i < npix; ++i) {
			xz[2 * i] = h->grid_xyz[3 * i];
			xz[2 * i + 1] = h->grid_xyz[3 * i + 2];
		}
		void* d_grid = h->d_grid.reserve(npix * 2 * sizeof(float));
		HIP_CHECK(hipMemcpyAsync(d_grid, xz.data(), npix * 2 * sizeof(float), hipMemcpyHostToDevice, s));
		launch_delay_table(static_cast<const float*>(d_grid), npix, h->rows, static_cast<double>(h->cfg.pitch), inv_ct,
				d_delay, s);
	}
	HIP_CHECK(hipStreamSynchronize(s));
	++h->grid_version;
}

// Validates and installs a grid of n1 x n2 points (x, y, z); ny > 1 marks a
// volume laid out as n1 = nx * ny columns, y fastest.
void set_grid_xyz(mista_handle* h, std::vector<float>&& xyz, uint32_t n1, uint32_t n2, uint32_t ny)
{
	if (n1 == 0 || n2 == 0) MISTA_THROW(MISTA_E_INVALID_PARAM, "Empty image grid.");
	if (ny == 0 || n1 % ny != 0) MISTA_THROW(MISTA_E_INVALID_PARAM, "n1 (" << n1 << ") is not a multiple of ny (" << ny << ").");
	const uint64_t npix = static_cast<uint64_t>(n1) * n2;
	if (npix > (1ull << 32)) MISTA_THROW(MISTA_E_INVALID_PARAM, "Too many grid points: " << npix << '.');
	for (uint64_t i = 0; i < 3 * npix; ++i)
		if (!std::isfinite(xyz[i]) || std::fabs(xyz[i]) > 1.0e3f)
			MISTA_THROW(MISTA_E_INVALID_VALUE, "Grid coordinate " << i / 3 << " is not finite or exceeds 1 km.");
	h->grid_xyz = std::move(xyz);
	h->n1 = n1;
	h->n2 = n2;
	h->ny = ny;
	try {
		build_delays(h);
	} catch (...) {
		h->n1 = h->n2 = 0;   // no half-built grid
		throw;
	}
}

void validate3d(const mista_sta3d_config& c)
{
	if (c.num_tx == 0 || c.num_rx == 0) MISTA_THROW(MISTA_E_INVALID_VALUE, "Empty list of active elements.");
	if (c.num_tx > 4096 || c.num_rx > 65536) MISTA_THROW(MISTA_E_INVALID_PARAM, "Too many active elements.");
	if (c.upsampling_factor < 1 || c.upsampling_factor > 64)
		MISTA_THROW(MISTA_E_INVALID_VALUE, "Invalid upsampling factor: " << c.upsampling_factor << ". Must be in [1, 64].");
	if (!(c.sampling_frequency > 0) || !(c.propagation_speed > 0) || !(c.center_frequency > 0))
		MISTA_THROW(MISTA_E_INVALID_PARAM, "sampling_frequency, propagation_speed and center_frequency must be > 0.");
	if (!(c.dead_zone_m >= 0)) MISTA_THROW(MISTA_E_INVALID_PARAM, "dead_zone_m must be >= 0.");
	if (!(c.peak_offset >= 0)) MISTA_THROW(MISTA_E_INVALID_PARAM, "peak_offset must be >= 0.");
	if (c.cf_kind != MISTA_CF_NONE && c.cf_kind != MISTA_CF_PCF && c.cf_kind != MISTA_CF_SCF)
		MISTA_THROW(MISTA_E_INVALID_PARAM, "Invalid coherence factor method: " << c.cf_kind << '.');
	if (c.cf_kind != MISTA_CF_NONE && !(c.cf_param >= 0 && c.cf_param <= 100))
		MISTA_THROW(MISTA_E_INVALID_PARAM, "coherence factor parameter out of [0, 100].");
	if (c.storage != MISTA_STORAGE_F32 && c.storage != MISTA_STORAGE_F16 && c.storage != MISTA_STORAGE_BF16)
		MISTA_THROW(MISTA_E_INVALID_PARAM, "Invalid storage type: " << c.storage << '.');
}

void set_elements_3d(mista_handle* h, const float* tx_xy, const float* rx_xy)
{
	const uint32_t nrx = h->cfg.num_elements, ntx = ntx_of(h->cfg);
	std::vector<double> e(static_cast<size_t>(nrx + ntx) * 2);
	for (uint32_t r = 0; r < nrx; ++r) {
		e[2 * r] = rx_xy[2 * r];
		e[2 * r + 1] = rx_xy[2 * r + 1];
	}
	for (uint32_t t = 0; t < ntx; ++t) {
		e[2 * (nrx + t)] = tx_xy[2 * t];
		e[2 * (nrx + t) + 1] = tx_xy[2 * t + 1];
	}
	for (double v : e)
		if (!std::isfinite(v) || std::fabs(v) > 1.0e3) MISTA_THROW(MISTA_E_INVALID_VALUE, "Element position is not finite or exceeds 1 km.");
	h->elem_xy.swap(e);
}

}  // namespace

namespace mista {
// Library-internal accessors used by the serving engine (serve.cpp).
hipStream_t handle_stream(mista_handle* h) { return h->stream; }
size_t handle_device_bytes(const mista_handle* h)
{
	return h->d_rx_apod.capacity() + h->d_tx_apod.capacity() + h->d_w.capacity() + h->d_elem.capacity() +
			h->d_grid.capacity() + h->d_delay.capacity() + h->d_raw.capacity() + h->d_tensor.capacity() +
			h->d_scale.capacity() + h->d_hpoly.capacity() + h->d_tmp.capacity() + h->d_value.capacity() +
			h->d_factor.capacity() + h->das_scratch.partial.capacity() + h->das_scratch.bounds.capacity() +
			h->das_scratch.params.capacity() + prep_device_bytes(h->prep);
}
}  // namespace mista

extern "C" {

int mista_abi_version(void) { return MISTA_ABI_VERSION; }

const char* mista_last_error(const mista_handle* h)
{
	return h ? h->last_error.c_str() : g_thread_error.c_str();
}

int mista_create(const mista_sta_config* config, mista_handle** out)
{
	if (!config || !out) {
		g_thread_error = "null argument";
		return MISTA_E_INVALID_PARAM;
	}
	*out = nullptr;
	return guarded(nullptr, [&] {
		validate(*config);
		int ndev = 0;
		HIP_CHECK(hipGetDeviceCount(&ndev));
		if (config->device < 0 || config->device >= ndev)
			MISTA_THROW(MISTA_E_INVALID_PARAM, "Invalid device " << config->device << " (" << ndev << " devices).");
		HIP_CHECK(hipSetDevice(config->device));
		auto h = std::make_unique<mista_handle>();
		h->cfg = *config;
		h->rows = config->num_elements;
		h->tx_row0 = config->first_tx_elem;
		HIP_CHECK(hipStreamCreate(&h->stream));  // default priority
		h->rx_apod.assign(config->num_elements, 1.0f);
		*out = h.release();
	});
}

int mista_create_3d(const mista_sta3d_config* config, const float* tx_xy, const float* rx_xy, mista_handle** out)
{
	if (!config || !out || !tx_xy || !rx_xy) {
		g_thread_error = "null argument";
		return MISTA_E_INVALID_PARAM;
	}
	*out = nullptr;
	return guarded(nullptr, [&] {
		validate3d(*config);
		int ndev = 0;
		HIP_CHECK(hipGetDeviceCount(&ndev));
		if (config->device < 0 || config->device >= ndev)
			MISTA_THROW(MISTA_E_INVALID_PARAM, "Invalid device " << config->device << " (" << ndev << " devices).");
		HIP_CHECK(hipSetDevice(config->device));
		auto h = std::make_unique<mista_handle>();
		mista_sta_config& c = h->cfg;
		c.num_elements = config->num_rx;   // rx apodization size = active rx count
		c.first_tx_elem = 0;
		c.last_tx_elem = config->num_tx - 1;
		c.upsampling_factor = config->upsampling_factor;
		c.pitch = 0.0f;
		c.sampling_frequency = config->sampling_frequency;
		c.propagation_speed = config->propagation_speed;
		c.center_frequency = config->center_frequency;
		c.dead_zone_m = config->dead_zone_m;
		c.peak_offset = config->peak_offset;
		c.cf_kind = config->cf_kind;
		c.cf_param = config->cf_param;
		c.storage = config->storage;
		c.device = config->device;
		h->is3d = true;
		h->rows = config->num_rx + config->num_tx;
		h->tx_row0 = config->num_rx;
		set_elements_3d(h.get(), tx_xy, rx_xy);
		HIP_CHECK(hipStreamCreate(&h->stream));  // default priority
		h->rx_apod.assign(config->num_rx, 1.0f);
		*out = h.release();
	});
}

int mista_create_rf(const mista_sta_config* config, mista_handle** out)
{
	if (!config || !out) {
		g_thread_error = "null argument";
		return MISTA_E_INVALID_PARAM;
	}
	*out = nullptr;
	return guarded(nullptr, [&] {
		mista_sta_config c = *config;
		if (c.upsampling_factor == 0) c.upsampling_factor = 4;
		if (c.upsampling_factor != 4)
			MISTA_THROW(MISTA_E_INVALID_VALUE, "The RF processor upsamples by 4 (got " << c.upsampling_factor << ").");
		if (c.cf_kind == MISTA_CF_PCF)
			MISTA_THROW(MISTA_E_INVALID_PARAM, "The RF processor supports no coherence factor or the sign coherence factor.");
		c.storage = MISTA_STORAGE_F32;
		validate(c);
		int ndev = 0;
		HIP_CHECK(hipGetDeviceCount(&ndev));
		if (c.device < 0 || c.device >= ndev)
			MISTA_THROW(MISTA_E_INVALID_PARAM, "Invalid device " << c.device << " (" << ndev << " devices).");
		HIP_CHECK(hipSetDevice(c.device));
		auto h = std::make_unique<mista_handle>();
		c.storage = kStorageReal;
		h->cfg = c;
		h->rf = true;
		h->rows = c.num_elements;
		h->tx_row0 = c.first_tx_elem;
		HIP_CHECK(hipStreamCreate(&h->stream));  // default priority
		h->rx_apod.assign(c.num_elements, 1.0f);
		*out = h.release();
	});
}

int mista_envelope_device(mista_handle* h, float* d_data, uint32_t n1, uint32_t n2, void* stream)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		if (!d_data) MISTA_THROW(MISTA_E_INVALID_PARAM, "null device pointer.");
		envelope_rows(h->env, d_data, n1, n2, pick_stream(h, stream));
	});
}

int mista_envelope(mista_handle* h, float* data, uint32_t n1, uint32_t n2)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		if (!data) MISTA_THROW(MISTA_E_INVALID_PARAM, "null host pointer.");
		hipStream_t s = h->stream;
		const size_t bytes = static_cast<size_t>(n1) * n2 * sizeof(float);
		float* d = static_cast<float*>(h->d_tmp.reserve(bytes));
		HIP_CHECK(hipMemcpyAsync(d, data, bytes, hipMemcpyHostToDevice, s));
		envelope_rows(h->env, d, n1, n2, s);
		HIP_CHECK(hipMemcpyAsync(data, d, bytes, hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
	});
}

int mista_set_elements_3d(mista_handle* h, const float* tx_xy, const float* rx_xy)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		++h->settings_epoch;
		if (!h->is3d) MISTA_THROW(MISTA_E_INVALID_STATE, "Not a 3-D processor.");
		if (!tx_xy || !rx_xy) MISTA_THROW(MISTA_E_INVALID_PARAM, "null argument.");
		set_elements_3d(h, tx_xy, rx_xy);
		if (h->n1 && h->n2) build_delays(h);
	});
}

int mista_destroy(mista_handle* h)
{
	if (!h) return MISTA_OK;
	(void)hipSetDevice(h->cfg.device);
	if (h->stream) (void)hipStreamSynchronize(h->stream);
	delete h;
	return MISTA_OK;
}

int mista_prepare(mista_handle* h, uint32_t base_element)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	h->base_element = base_element;
	return MISTA_OK;
}

int mista_set_rx_apod(mista_handle* h, const float* rx_apod, uint32_t n)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		if (!rx_apod) MISTA_THROW(MISTA_E_INVALID_PARAM, "rx_apod is null.");
		if (n != h->cfg.num_elements)
			MISTA_THROW(MISTA_E_INVALID_VALUE, "Wrong rx apodization size: " << n << " (should be " << h->cfg.num_elements << ").");
		// unchanged weights (a request loop sets them for every job): no
		// upload, no stream synchronisation, no new frame graph
		if (h->rx_apod.size() == n && std::equal(h->rx_apod.begin(), h->rx_apod.end(), rx_apod)) return;
		++h->settings_epoch;
		h->rx_apod.assign(rx_apod, rx_apod + n);
		h->weights_dirty = true;
	});
}

int mista_set_tx_apod(mista_handle* h, const float* tx_apod, uint32_t n)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		if (!tx_apod) {
			if (!h->tx_apod_set) return;
			h->tx_apod.clear();
			h->tx_apod_set = false;
		} else {
			if (n != ntx_of(h->cfg))
				MISTA_THROW(MISTA_E_INVALID_VALUE, "Wrong tx apodization size: " << n << " (should be " << ntx_of(h->cfg) << ").");
			if (h->tx_apod_set && h->tx_apod.size() == n && std::equal(h->tx_apod.begin(), h->tx_apod.end(), tx_apod))
				return;
			h->tx_apod.assign(tx_apod, tx_apod + n);
			h->tx_apod_set = true;
		}
		++h->settings_epoch;
		h->weights_dirty = true;
	});
}

int mista_set_grid(mista_handle* h, const float* grid_xz, uint32_t n1, uint32_t n2)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		++h->settings_epoch;
		if (!grid_xz || n1 == 0 || n2 == 0) MISTA_THROW(MISTA_E_INVALID_PARAM, "Empty image grid.");
		const uint64_t npix = static_cast<uint64_t>(n1) * n2;
		std::vector<float> xyz(npix * 3);
		for (uint64_t i = 0; i < npix; ++i) {
			xyz[3 * i] = grid_xz[2 * i];
			xyz[3 * i + 1] = 0.0f;
			xyz[3 * i + 2] = grid_xz[2 * i + 1];
		}
		set_grid_xyz(h, std::move(xyz), n1, n2, 1);
	});
}

int mista_set_grid_xyz(mista_handle* h, const float* grid_xyz, uint32_t n1, uint32_t n2, uint32_t ny)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		++h->settings_epoch;
		if (!grid_xyz || n1 == 0 || n2 == 0) MISTA_THROW(MISTA_E_INVALID_PARAM, "Empty image grid.");
		const uint64_t npix = static_cast<uint64_t>(n1) * n2;
		set_grid_xyz(h, std::vector<float>(grid_xyz, grid_xyz + 3 * npix), n1, n2, ny);
	});
}

int mista_set_grid_points(mista_handle* h, const float* points, uint32_t n1, uint32_t n2)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		++h->settings_epoch;
		if (!points || n1 == 0 || n2 == 0) MISTA_THROW(MISTA_E_INVALID_PARAM, "Empty image grid.");
		const uint64_t npix = static_cast<uint64_t>(n1) * n2;
		std::vector<float> xyz(npix * 3);
		for (uint64_t i = 0; i < npix; ++i)
			for (int k = 0; k < 3; ++k) xyz[3 * i + k] = points[5 * i + k];
		set_grid_xyz(h, std::move(xyz), n1, n2, 1);
	});
}

int mista_process_device(mista_handle* h, const float* d_echoes, uint32_t num_samples, float* d_value, float* d_factor,
		void* stream)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		if (!d_echoes || !d_value) MISTA_THROW(MISTA_E_INVALID_PARAM, "null device pointer.");
		hipStream_t s = pick_stream(h, stream);
		if (!h->graph_on) {
			run_prepare(h, d_echoes, num_samples, s);
			run_das(h, d_value, d_factor, s);
			return;
		}
		auto key = [&] {
			mista_handle::GraphKey k;
			k.echoes = d_echoes;
			k.value = d_value;
			k.factor = d_factor;
			k.ns = num_samples;
			k.settings = h->settings_epoch;
			k.allocs = alloc_epoch().load();
			k.grid = h->grid_version;
			k.tuned = h->das_scratch.tuned_key ^ h->das_scratch.params_key;
			k.p0 = h->das_scratch.tuned_p0;
			k.p1 = h->das_scratch.tuned_p1;
			return k;
		};
		if (h->gexec && key() == h->gkey) {
			// one launch for the whole frame: the stage split is not
			// observable (mista_last_timing reports it as unavailable)
			HIP_CHECK(hipGraphLaunch(h->gexec, s));
			h->timed = false;
			h->graph_frame = true;
			++h->graph_launches;
			return;
		}
		// eager frame; if it changed nothing the kernels depend on (no
		// allocation, table build, tuning or plan), record the same launches
		// as a graph on the handle's stream for the frames that follow
		h->drop_graph();
		const mista_handle::GraphKey before = key();
		const bool plan_ready = h->prep.plan != nullptr;
		run_prepare(h, d_echoes, num_samples, s);
		run_das(h, d_value, d_factor, s);
		// (the library-FFT preparation and the RF path stay eager)
		if (!(key() == before) || !plan_ready || h->rf || !h->prep.last_fused) return;
		// The frame is done; a recording that fails leaves the next frames
		// eager (the same key is not tried again) and is not an error.
		if (h->gfail_valid && h->gfail == before) return;
		hipGraph_t graph = nullptr;
		bool ok = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
		if (ok) {
			h->capturing = true;
			try {
				run_prepare(h, d_echoes, num_samples, h->stream);
				run_das(h, d_value, d_factor, h->stream);
			} catch (...) {
				ok = false;
			}
			h->capturing = false;
			ok = (hipStreamEndCapture(h->stream, &graph) == hipSuccess) && ok && graph;
		}
		if (ok && hipGraphInstantiate(&h->gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
			h->gexec = nullptr;
			ok = false;
		}
		if (graph) (void)hipGraphDestroy(graph);
		(void)hipGetLastError();   // a failed recording leaves no sticky error behind
		if (!ok) {
			h->gfail = before;
			h->gfail_valid = true;
			return;
		}
		h->gkey = key();
		++h->graph_captures;
	});
}

int mista_set_graph(mista_handle* h, int enable)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		h->graph_on = enable != 0;
		h->drop_graph();
	});
}

int mista_graph_stats(const mista_handle* h, uint64_t* captures, uint64_t* launches)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	if (captures) *captures = h->graph_captures;
	if (launches) *launches = h->graph_launches;
	return MISTA_OK;
}

int mista_prepare_signals_device(mista_handle* h, const float* d_echoes, uint32_t num_samples, void* stream)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		if (!d_echoes) MISTA_THROW(MISTA_E_INVALID_PARAM, "null device pointer.");
		run_prepare(h, d_echoes, num_samples, pick_stream(h, stream));
	});
}

int mista_das_device(mista_handle* h, float* d_value, float* d_factor, void* stream)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		if (!d_value) MISTA_THROW(MISTA_E_INVALID_PARAM, "null device pointer.");
		run_das(h, d_value, d_factor, pick_stream(h, stream));
	});
}

int mista_das_batch_device(mista_handle* const* handles, uint32_t n, float* const* d_values, float* const* d_factors,
		void* stream)
{
	if (!handles || n == 0 || !d_values || !handles[0]) return MISTA_E_INVALID_PARAM;
	mista_handle* h0 = handles[0];
	return guarded(h0, [&] {
		hipStream_t s = pick_stream(h0, stream);
		std::vector<DasBatchItem> items;
		items.reserve(n);
		for (uint32_t k = 0; k < n; ++k) {
			mista_handle* h = handles[k];
			if (!h || !d_values[k]) MISTA_THROW(MISTA_E_INVALID_PARAM, "null handle or device pointer at " << k << '.');
			if (h->cfg.device != h0->cfg.device) MISTA_THROW(MISTA_E_INVALID_PARAM, "handles on different devices.");
			for (uint32_t q = 0; q < k; ++q)
				if (handles[q] == h) MISTA_THROW(MISTA_E_INVALID_PARAM, "handle " << k << " appears twice.");
			require_grid(h);
			if (!h->signal_ready) MISTA_THROW(MISTA_E_INVALID_STATE, "No prepared signals in handle " << k << '.');
			h->geom.n1 = h->n1;
			h->geom.n2 = h->n2;
			h->geom.ny = h->ny;
			h->geom.npix = static_cast<uint64_t>(h->n1) * h->n2;
			h->geom.grid_version = h->grid_version;
			upload_apod(h, s);
			launch_weights(h->geom, h->d_rx_apod.as<const float>(), h->tx_apod_set ? h->d_tx_apod.as<const float>() : nullptr,
					h->d_scale.as<const float>(), h->d_w.as<float>(), s);
			items.push_back({&h->geom, &h->tuning, &h->das_scratch, h->d_tensor.get(), h->d_delay.as<const float>(),
					h->d_w.as<const float>(), d_values[k], d_factors ? d_factors[k] : nullptr});
		}
		launch_das_batch(items, s);
		for (uint32_t k = 0; k < n; ++k) handles[k]->timed = false;   // the frames share one launch
	});
}

int mista_das_device_noalloc(mista_handle* h, float* d_value, float* d_factor, void* stream)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		if (!d_value) MISTA_THROW(MISTA_E_INVALID_PARAM, "null device pointer.");
		require_grid(h);
		if (!h->signal_ready) MISTA_THROW(MISTA_E_INVALID_STATE, "No prepared signals: call a prepare/process/upload entry point first.");
		if (h->tuning.variant == 1) MISTA_THROW(MISTA_E_INVALID_STATE, "The no-allocation DAS runs the tiled kernel only.");
		hipStream_t s = pick_stream(h, stream);
		h->geom.n1 = h->n1;
		h->geom.n2 = h->n2;
		h->geom.ny = h->ny;
		h->geom.npix = static_cast<uint64_t>(h->n1) * h->n2;
		h->geom.grid_version = h->grid_version;
		upload_apod(h, s);   // into buffers the preparation reserved
		launch_weights(h->geom, h->d_rx_apod.as<const float>(), h->tx_apod_set ? h->d_tx_apod.as<const float>() : nullptr,
				h->d_scale.as<const float>(), h->d_w.as<float>(), s);
		const bool timed = h->ev[2] && h->ev[3];   // events are only recorded when they exist already
		if (timed) HIP_CHECK(hipEventRecord(h->ev[2], s));
		int p0 = h->das_scratch.tuned_p0, p1 = h->das_scratch.tuned_p1;
		if (h->tuning.variant == 2 || h->tuning.p0 || h->tuning.p1) {
			p0 = h->tuning.p0;
			p1 = h->tuning.p1;
		}
		launch_das_noalloc(h->geom, p0, p1, h->das_scratch, h->d_tensor.get(), h->d_delay.as<const float>(),
				h->d_w.as<const float>(), d_value, d_factor, s);
		if (timed) {
			HIP_CHECK(hipEventRecord(h->ev[3], s));
			h->timed = true;
			h->graph_frame = false;
		}
	});
}

int mista_process(mista_handle* h, const float* echoes, uint32_t num_samples, float* value, float* factor)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		if (!echoes || !value) MISTA_THROW(MISTA_E_INVALID_PARAM, "null host pointer.");
		require_grid(h);
		hipStream_t s = h->stream;
		const uint64_t n = static_cast<uint64_t>(ntx_of(h->cfg)) * h->cfg.num_elements * num_samples;
		float* d_raw = static_cast<float*>(h->d_raw.reserve(n * sizeof(float)));
		const uint64_t npix = static_cast<uint64_t>(h->n1) * h->n2;
		float* d_value = static_cast<float*>(h->d_value.reserve(npix * sizeof(float)));
		float* d_factor = static_cast<float*>(h->d_factor.reserve(npix * sizeof(float)));
		if (h->rf || num_samples == 0) {
			HIP_CHECK(hipMemcpyAsync(d_raw, echoes, n * sizeof(float), hipMemcpyHostToDevice, s));
			run_prepare(h, d_raw, num_samples, s);
		} else {
			// The echoes go up in the chunks the preparation consumes, on a
			// copy stream; each chunk's kernels wait for that chunk alone, so
			// the upload overlaps the preparation of the chunks before it.
			if (!h->copy_stream) HIP_CHECK(hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
			size_t k = 0;
			h->prep.before_chunk = [&](uint32_t c0, uint32_t nc) {
				if (k == h->chunk_ev.size()) {
					hipEvent_t e = nullptr;
					HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
					h->chunk_ev.push_back(e);
				}
				const uint64_t off = static_cast<uint64_t>(c0) * num_samples;
				HIP_CHECK(hipMemcpyAsync(d_raw + off, echoes + off, static_cast<uint64_t>(nc) * num_samples * sizeof(float),
						hipMemcpyHostToDevice, h->copy_stream));
				HIP_CHECK(hipEventRecord(h->chunk_ev[k], h->copy_stream));
				HIP_CHECK(hipStreamWaitEvent(s, h->chunk_ev[k], 0));
				++k;
			};
			try {
				run_prepare(h, d_raw, num_samples, s);
			} catch (...) {
				h->prep.before_chunk = nullptr;
				throw;
			}
			h->prep.before_chunk = nullptr;
		}
		run_das(h, d_value, d_factor, s);
		HIP_CHECK(hipMemcpyAsync(value, d_value, npix * sizeof(float), hipMemcpyDeviceToHost, s));
		if (factor) HIP_CHECK(hipMemcpyAsync(factor, d_factor, npix * sizeof(float), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
	});
}

int mista_process_points(mista_handle* h, const float* echoes, uint32_t num_samples, float* points)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	if (!points) return guarded(h, [&] { MISTA_THROW(MISTA_E_INVALID_PARAM, "points is null."); });
	const uint64_t npix = static_cast<uint64_t>(h->n1) * h->n2;
	std::vector<float> value(npix), factor(npix);
	const int rc = mista_process(h, echoes, num_samples, value.data(), factor.data());
	if (rc != MISTA_OK) return rc;
	for (uint64_t i = 0; i < npix; ++i) {
		points[5 * i + 3] = value[i];
		points[5 * i + 4] = factor[i];
	}
	return MISTA_OK;
}

int mista_upload_analytic(mista_handle* h, const float* analytic_ri, uint32_t signal_length)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		++h->settings_epoch;
		if (!analytic_ri) MISTA_THROW(MISTA_E_INVALID_PARAM, "null host pointer.");
		require_grid(h);
		hipStream_t s = h->stream;
		set_frame_geom(h, 0, signal_length, s);
		const size_t nch = static_cast<size_t>(h->geom.ntx) * h->geom.nrx;
		if (h->rf) {   // real samples [ntx][nrx][L]
			HIP_CHECK(hipMemcpyAsync(h->d_tensor.get(), analytic_ri, nch * signal_length * sizeof(float),
					hipMemcpyHostToDevice, s));
			HIP_CHECK(hipStreamSynchronize(s));
			h->signal_ready = true;
			return;
		}
		const size_t bytes = nch * signal_length * 2 * sizeof(float);
		float* d_tmp = static_cast<float*>(h->d_tmp.reserve(bytes));
		HIP_CHECK(hipMemcpyAsync(d_tmp, analytic_ri, bytes, hipMemcpyHostToDevice, s));
		launch_tensor_convert_in(h->geom, d_tmp, h->d_tensor.get(), h->d_scale.as<float>(), s);
		HIP_CHECK(hipStreamSynchronize(s));
		h->d_tmp.release();
		h->signal_ready = true;
	});
}

int mista_download_analytic(mista_handle* h, float* analytic_ri)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		if (!analytic_ri) MISTA_THROW(MISTA_E_INVALID_PARAM, "null host pointer.");
		if (!h->signal_ready) MISTA_THROW(MISTA_E_INVALID_STATE, "No prepared signals.");
		hipStream_t s = h->stream;
		const size_t nch = static_cast<size_t>(h->geom.ntx) * h->geom.nrx;
		if (h->rf) {
			HIP_CHECK(hipMemcpyAsync(analytic_ri, h->d_tensor.get(), nch * h->geom.L * sizeof(float), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipStreamSynchronize(s));
			return;
		}
		const size_t bytes = nch * h->geom.L * 2 * sizeof(float);
		float* d_tmp = static_cast<float*>(h->d_tmp.reserve(bytes));
		launch_tensor_convert_out(h->geom, h->d_tensor.get(), h->d_scale.as<const float>(), d_tmp, s);
		HIP_CHECK(hipMemcpyAsync(analytic_ri, d_tmp, bytes, hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
		h->d_tmp.release();
	});
}

int mista_download_delays(mista_handle* h, float* delays)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		require_grid(h);
		if (!delays) MISTA_THROW(MISTA_E_INVALID_PARAM, "null host pointer.");
		const size_t bytes = static_cast<size_t>(h->n1) * h->n2 * h->rows * sizeof(float);
		HIP_CHECK(hipMemcpyAsync(delays, h->d_delay.get(), bytes, hipMemcpyDeviceToHost, h->stream));
		HIP_CHECK(hipStreamSynchronize(h->stream));
	});
}

int mista_download_delays_at(mista_handle* h, const uint64_t* points, uint64_t n, float* delays)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		require_grid(h);
		if ((!points || !delays) && n) MISTA_THROW(MISTA_E_INVALID_PARAM, "null host pointer.");
		const uint64_t npix = static_cast<uint64_t>(h->n1) * h->n2;
		for (uint64_t i = 0; i < n; ++i)
			if (points[i] >= npix) MISTA_THROW(MISTA_E_INVALID_PARAM, "Point index " << points[i] << " out of range.");
		const float* d = h->d_delay.as<const float>();
		// one copy per row and run of consecutive point indices
		for (uint64_t i = 0; i < n;) {
			uint64_t j = i + 1;
			while (j < n && points[j] == points[j - 1] + 1) ++j;
			for (uint32_t r = 0; r < h->rows; ++r)
				HIP_CHECK(hipMemcpyAsync(delays + static_cast<uint64_t>(r) * n + i, d + static_cast<uint64_t>(r) * npix + points[i],
						(j - i) * sizeof(float), hipMemcpyDeviceToHost, h->stream));
			i = j;
		}
		HIP_CHECK(hipStreamSynchronize(h->stream));
	});
}

int mista_query(const mista_handle* h, uint32_t* signal_length, uint32_t* fft_size, uint64_t* num_points)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	if (signal_length) *signal_length = h->geom.L;
	if (fft_size) *fft_size = h->geom.nf;
	if (num_points) *num_points = static_cast<uint64_t>(h->n1) * h->n2;
	return MISTA_OK;
}

int mista_set_kernel(mista_handle* h, int variant, int p0, int p1)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		++h->settings_epoch;
		if (variant < 0 || variant > 2) MISTA_THROW(MISTA_E_INVALID_PARAM, "Invalid kernel variant " << variant);
		h->tuning.variant = variant;
		h->tuning.p0 = p0;
		h->tuning.p1 = p1;
	});
}

int mista_tuned_kernel(const mista_handle* h, int* p0, int* p1)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	if (p0) *p0 = h->das_scratch.tuned_p0;
	if (p1) *p1 = h->das_scratch.tuned_p1;
	return MISTA_OK;
}

int mista_last_kernel(const mista_handle* h, int* tile_param0, int* tile_param1)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	if (tile_param0) *tile_param0 = h->das_scratch.last_p0;
	if (tile_param1) *tile_param1 = h->das_scratch.last_p1;
	return MISTA_OK;
}

int mista_set_prep(mista_handle* h, int method)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	return guarded(h, [&] {
		++h->settings_epoch;
		if (method < 0 || method > 3) MISTA_THROW(MISTA_E_INVALID_PARAM, "Invalid preparation method " << method);
		h->prep.method = method;
	});
}

int mista_last_timing(const mista_handle* h, float* prep_ms, float* das_ms)
{
	if (!h) return MISTA_E_INVALID_PARAM;
	mista_handle* hh = const_cast<mista_handle*>(h);
	return guarded(hh, [&] {
		float p = 0.0f, d = 0.0f;
		if (h->graph_frame) {
			p = d = -1.0f;   // a replayed frame graph: no stage split
		} else if (h->timed) {
			HIP_CHECK(hipEventSynchronize(h->ev[3]));
			if (h->ev[0] && h->ev[1] && hipEventQuery(h->ev[1]) == hipSuccess &&
					hipEventElapsedTime(&p, h->ev[0], h->ev[1]) != hipSuccess) p = 0.0f;
			HIP_CHECK(hipEventElapsedTime(&d, h->ev[2], h->ev[3]));
		}
		if (prep_ms) *prep_ms = p;
		if (das_ms) *das_ms = d;
	});
}

int mista_synth_echoes_device(const mista_synth_params* p, const float* tx_xyz, const float* rx_xyz,
		const float* scat_xyza, float* d_echoes, void* stream)
{
	return guarded(nullptr, [&] {
		if (!p || !tx_xyz || !rx_xyz || !d_echoes || (p->num_scatterers && !scat_xyza))
			MISTA_THROW(MISTA_E_INVALID_PARAM, "null argument.");
		synth_echoes(*p, tx_xyz, rx_xyz, scat_xyza, d_echoes, static_cast<hipStream_t>(stream));
	});
}

int mista_sum_squares_device(const float* d_x, uint64_t n, double* out, void* stream)
{
	return guarded(nullptr, [&] {
		if (!d_x || !out) MISTA_THROW(MISTA_E_INVALID_PARAM, "null argument.");
		*out = sum_squares(d_x, n, static_cast<hipStream_t>(stream));
	});
}

int mista_read_bandwidth_device(const void* d_x, uint64_t bytes, int reps, double* gb_per_s, void* stream)
{
	return guarded(nullptr, [&] {
		if (!d_x || !gb_per_s) MISTA_THROW(MISTA_E_INVALID_PARAM, "null argument.");
		*gb_per_s = read_bandwidth(d_x, bytes, reps, static_cast<hipStream_t>(stream));
	});
}

int mista_add_noise_device(float* d_x, uint64_t n, float noise_std, uint64_t seed, void* stream)
{
	return guarded(nullptr, [&] {
		if (!d_x) MISTA_THROW(MISTA_E_INVALID_PARAM, "null argument.");
		add_noise(d_x, n, noise_std, seed, static_cast<hipStream_t>(stream));
	});
}

}  // extern "C"

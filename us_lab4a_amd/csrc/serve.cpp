// Serving engine: a request loop over STA imaging jobs of uneven shapes on one
// GPU (include/mista.h, mista_engine_*).
//
// Jobs are submitted from any thread and admitted in arrival order by a worker
// thread as slots and device memory free up.  Each admitted job runs on a
// processor handle taken from a resident pool keyed by everything that fixes
// the handle's device state (configuration, grid, sample count): a job whose
// key is resident reuses the handle's delay table, window tables and FFT
// plans.  Every handle has its own HIP stream, so admitted jobs overlap on the
// device (continuous batching: a new job starts as soon as one finishes, not
// when a whole batch has).  Admission is by device-memory footprint: a handle
// is created only if the pool's resident bytes plus the job's estimate fit
// the budget, after evicting idle handles least recently used first.
//
// Each job's arithmetic is independent of the other jobs in flight (no shared
// accumulators), so a job's images are bit-identical alone and in a mix.
#include "common.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

using namespace mista;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b)
{
	return std::chrono::duration<double, std::milli>(b - a).count();
}

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ull)
{
	const unsigned char* p = static_cast<const unsigned char*>(data);
	for (size_t i = 0; i < n; ++i) {
		h ^= p[i];
		h *= 1099511628211ull;
	}
	return h;
}

struct Job {
	uint64_t ticket = 0;
	mista_job spec{};
	std::vector<float> grid, rx_apod, tx_apod;   // owned copies of the small inputs
	uint64_t key = 0;
	size_t estimate = 0;
	Clock::time_point t_submit, t_admit, t_done;
	hipEvent_t ev_start = nullptr, ev_end = nullptr;
	int slot = -1;
	bool done = false;
	int status = MISTA_OK;
	double device_ms = 0.0;
	size_t footprint = 0;
};

struct Slot {
	mista_handle* h = nullptr;
	uint64_t key = 0;
	size_t bytes = 0;            // device bytes (estimate until the first job ends)
	uint64_t last_use = 0;
	bool busy = false;
	DeviceBuffer d_echoes, d_value, d_factor;   // staging for host inputs/outputs
};

// Device bytes of a 2-D processor for this job (tensor, delay table, window
// tables, signal-preparation buffers, outputs), as the library allocates them.
size_t estimate_bytes(const mista_job& j)
{
	const mista_sta_config& c = j.config;
	const uint64_t ntx = c.last_tx_elem - c.first_tx_elem + 1ull, nrx = c.num_elements;
	const uint64_t L = static_cast<uint64_t>(j.num_samples) * c.upsampling_factor;
	const uint64_t nf = hilbert_fft_size(static_cast<uint32_t>(L));
	const uint64_t npix = static_cast<uint64_t>(j.n1) * j.n2;
	const uint64_t tensor = ntx * nrx * L * storage_bytes(c.storage);
	const uint64_t delays = nrx * npix * 4 + npix * 8;
	const uint64_t ntx_pad = pad_ntx(static_cast<uint32_t>(ntx));
	const uint64_t tables = (npix / 256 + 1) * (nrx * ntx_pad * 8 + nrx * 8);
	const uint64_t chunk = std::min<uint64_t>(ntx * nrx, 1024);
	const uint64_t prep = chunk * nf * (4 + 8 + 16);
	const uint64_t io = ntx * nrx * j.num_samples * 4 + npix * 8;
	return tensor + delays + tables + prep + io + (16u << 20);
}

uint64_t job_key(const mista_job& j, const std::vector<float>& grid)
{
	mista_sta_config c = j.config;
	c.device = 0;
	uint64_t h = fnv1a(&c, sizeof(c));
	h = fnv1a(&j.n1, sizeof(j.n1), h);
	h = fnv1a(&j.n2, sizeof(j.n2), h);
	h = fnv1a(&j.num_samples, sizeof(j.num_samples), h);
	return fnv1a(grid.data(), grid.size() * sizeof(float), h);
}

}  // namespace

struct mista_engine {
	mista_engine_config cfg{};
	size_t budget = 0;
	std::mutex mu;
	std::condition_variable cv_work, cv_done;
	std::deque<Job*> queue;
	std::vector<Job*> running;
	std::unordered_map<uint64_t, Job*> jobs;
	std::deque<Slot> slots;   // stable addresses: handles and buffers never move
	size_t resident = 0;
	uint64_t next_ticket = 1, use_clock = 0;
	bool stop = false;
	std::string last_error;
	std::thread worker;

	void run();
	bool admit_one(std::unique_lock<std::mutex>& lk);
	void launch(Job* j, Slot& s);
	void finish(Job* j);
	void poll_running();
};

void mista_engine::finish(Job* j)
{
	j->t_done = Clock::now();
	if (j->status == MISTA_OK && j->ev_start && j->ev_end) {
		float ms = 0.0f;
		if (hipEventElapsedTime(&ms, j->ev_start, j->ev_end) == hipSuccess) j->device_ms = ms;
	}
	if (j->ev_start) (void)hipEventDestroy(j->ev_start);
	if (j->ev_end) (void)hipEventDestroy(j->ev_end);
	j->ev_start = j->ev_end = nullptr;
	if (j->slot >= 0) {
		Slot& s = slots[j->slot];
		s.busy = false;
		const size_t actual = handle_device_bytes(s.h) + s.d_echoes.capacity() + s.d_value.capacity() +
				s.d_factor.capacity();
		resident = resident - s.bytes + actual;
		s.bytes = actual;
		j->footprint = actual;
	}
	j->done = true;
}

void mista_engine::poll_running()
{
	for (size_t i = 0; i < running.size();) {
		Job* j = running[i];
		const hipError_t e = hipEventQuery(j->ev_end);
		if (e == hipErrorNotReady) {
			++i;
			continue;
		}
		if (e != hipSuccess) j->status = MISTA_E_HIP;
		finish(j);
		running.erase(running.begin() + static_cast<std::ptrdiff_t>(i));
		cv_done.notify_all();
	}
}

// Runs the job on its slot's stream: host inputs are staged in the slot's
// device buffers; outputs copied back when the caller gave host pointers.
void mista_engine::launch(Job* j, Slot& s)
{
	mista_handle* h = s.h;
	const mista_job& q = j->spec;
	hipStream_t st = handle_stream(h);
	const uint64_t ntx = q.config.last_tx_elem - q.config.first_tx_elem + 1ull;
	const uint64_t nraw = ntx * q.config.num_elements * q.num_samples;
	const uint64_t npix = static_cast<uint64_t>(q.n1) * q.n2;
	auto ok = [&](int rc) {
		if (rc != MISTA_OK) MISTA_THROW(rc, mista_last_error(h));
	};
	ok(mista_set_rx_apod(h, j->rx_apod.data(), static_cast<uint32_t>(j->rx_apod.size())));
	ok(mista_set_tx_apod(h, j->tx_apod.empty() ? nullptr : j->tx_apod.data(), static_cast<uint32_t>(j->tx_apod.size())));
	HIP_CHECK(hipEventCreate(&j->ev_start));
	HIP_CHECK(hipEventCreate(&j->ev_end));
	HIP_CHECK(hipEventRecord(j->ev_start, st));
	const float* d_echoes = q.echoes;
	if (!q.echoes_on_device) {
		float* d = static_cast<float*>(s.d_echoes.reserve(nraw * sizeof(float)));
		HIP_CHECK(hipMemcpyAsync(d, q.echoes, nraw * sizeof(float), hipMemcpyHostToDevice, st));
		d_echoes = d;
	}
	float* d_value = q.value;
	float* d_factor = q.factor;
	if (!q.outputs_on_device) {
		d_value = static_cast<float*>(s.d_value.reserve(npix * sizeof(float)));
		d_factor = q.factor ? static_cast<float*>(s.d_factor.reserve(npix * sizeof(float))) : nullptr;
	}
	ok(mista_process_device(h, d_echoes, q.num_samples, d_value, d_factor, nullptr));
	if (!q.outputs_on_device) {
		HIP_CHECK(hipMemcpyAsync(q.value, d_value, npix * sizeof(float), hipMemcpyDeviceToHost, st));
		if (q.factor) HIP_CHECK(hipMemcpyAsync(q.factor, d_factor, npix * sizeof(float), hipMemcpyDeviceToHost, st));
	}
	HIP_CHECK(hipEventRecord(j->ev_end, st));
}

// Admits the job at the head of the queue if a slot and memory allow.
// Returns false when the head must wait.  Called with the lock held; the lock
// is released while handles are built and work is enqueued.
bool mista_engine::admit_one(std::unique_lock<std::mutex>& lk)
{
	if (queue.empty() || running.size() >= cfg.max_inflight) return false;
	Job* j = queue.front();
	int si = -1;
	for (size_t i = 0; i < slots.size(); ++i)
		if (slots[i].h && !slots[i].busy && slots[i].key == j->key) {
			si = static_cast<int>(i);
			break;
		}
	if (si < 0) {
		// evict idle handles, least recently used first, until the job fits
		while (resident + j->estimate > budget) {
			int lru = -1;
			for (size_t i = 0; i < slots.size(); ++i)
				if (slots[i].h && !slots[i].busy && (lru < 0 || slots[i].last_use < slots[lru].last_use)) lru = static_cast<int>(i);
			if (lru < 0) break;
			Slot& v = slots[lru];
			(void)mista_destroy(v.h);
			v.h = nullptr;
			v.d_echoes.release();
			v.d_value.release();
			v.d_factor.release();
			resident -= std::min(resident, v.bytes);
			v.bytes = 0;
		}
		if (resident + j->estimate > budget) {
			if (!running.empty()) return false;   // wait for memory to free up
			queue.pop_front();
			j->status = MISTA_E_NO_MEMORY;
			j->done = true;
			j->t_admit = j->t_done = Clock::now();
			cv_done.notify_all();
			return true;
		}
		for (size_t i = 0; i < slots.size(); ++i)
			if (!slots[i].h) {
				si = static_cast<int>(i);
				break;
			}
		if (si < 0) {
			slots.emplace_back();
			si = static_cast<int>(slots.size() - 1);
		}
	}
	queue.pop_front();
	Slot& s = slots[si];
	s.busy = true;
	s.last_use = ++use_clock;
	j->slot = si;
	j->t_admit = Clock::now();
	running.push_back(j);
	const bool fresh = (s.h == nullptr);
	if (fresh) {
		s.key = j->key;
		s.bytes = j->estimate;
		resident += s.bytes;
	}
	lk.unlock();
	int status = MISTA_OK;
	std::string err;   // published under the lock below
	try {
		if (fresh) {
			mista_sta_config c = j->spec.config;
			c.device = cfg.device;
			mista_handle* h = nullptr;
			const int rc = mista_create(&c, &h);
			if (rc != MISTA_OK) MISTA_THROW(rc, mista_last_error(nullptr));
			s.h = h;
			const int rg = mista_set_grid(h, j->grid.data(), j->spec.n1, j->spec.n2);
			if (rg != MISTA_OK) MISTA_THROW(rg, mista_last_error(h));
		}
		launch(j, s);
	} catch (const Error& e) {
		status = e.code;
		err = e.what();
	} catch (const std::exception& e) {
		status = MISTA_E_INVALID_STATE;
		err = e.what();
	}
	lk.lock();
	if (status != MISTA_OK) {
		last_error = err;
		j->status = status;
		if (s.h && fresh && status != MISTA_E_HIP) {
			// keep nothing half-built
			(void)mista_destroy(s.h);
			s.h = nullptr;
			resident -= std::min(resident, s.bytes);
			s.bytes = 0;
		}
		running.erase(std::find(running.begin(), running.end(), j));
		s.busy = false;
		if (j->ev_start) (void)hipEventDestroy(j->ev_start);
		if (j->ev_end) (void)hipEventDestroy(j->ev_end);
		j->ev_start = j->ev_end = nullptr;
		j->slot = -1;
		j->done = true;
		j->t_done = Clock::now();
		cv_done.notify_all();
	}
	return true;
}

void mista_engine::run()
{
	(void)hipSetDevice(cfg.device);
	std::unique_lock<std::mutex> lk(mu);
	for (;;) {
		poll_running();
		bool progressed = false;
		while (admit_one(lk)) progressed = true;
		if (stop && queue.empty() && running.empty()) break;
		if (progressed) continue;
		if (running.empty()) {
			cv_work.wait(lk, [&] { return stop || !queue.empty(); });
		} else {
			// jobs on the device: poll their completion events
			cv_work.wait_for(lk, std::chrono::microseconds(50));
		}
	}
}

extern "C" {

int mista_engine_create(const mista_engine_config* config, mista_engine** out)
{
	if (!config || !out) return MISTA_E_INVALID_PARAM;
	*out = nullptr;
	try {
		if (config->max_inflight == 0 || config->max_inflight > 64)
			MISTA_THROW(MISTA_E_INVALID_PARAM, "max_inflight must be in [1, 64].");
		if (config->max_queued == 0) MISTA_THROW(MISTA_E_INVALID_PARAM, "max_queued must be >= 1.");
		int ndev = 0;
		HIP_CHECK(hipGetDeviceCount(&ndev));
		if (config->device < 0 || config->device >= ndev) MISTA_THROW(MISTA_E_INVALID_PARAM, "Invalid device.");
		HIP_CHECK(hipSetDevice(config->device));
		auto e = std::make_unique<mista_engine>();
		e->cfg = *config;
		if (config->hbm_budget) {
			e->budget = config->hbm_budget;
		} else {
			size_t free_b = 0, total_b = 0;
			HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
			e->budget = free_b / 10 * 8;
		}
		mista_engine* p = e.get();
		e->worker = std::thread([p] { p->run(); });
		*out = e.release();
		return MISTA_OK;
	} catch (const Error& err) {
		return err.code;
	} catch (const std::exception&) {
		return MISTA_E_INVALID_STATE;
	}
}

int mista_engine_destroy(mista_engine* e)
{
	if (!e) return MISTA_OK;
	{
		std::lock_guard<std::mutex> lk(e->mu);
		e->stop = true;
	}
	e->cv_work.notify_all();
	if (e->worker.joinable()) e->worker.join();
	(void)hipSetDevice(e->cfg.device);
	for (auto& s : e->slots)
		if (s.h) (void)mista_destroy(s.h);
	for (auto& kv : e->jobs) delete kv.second;
	delete e;
	return MISTA_OK;
}

const char* mista_engine_last_error(const mista_engine* e)
{
	// a copy taken under the engine's lock: the worker may set a new message
	// at any time
	thread_local std::string msg;
	if (!e) return "";
	std::lock_guard<std::mutex> g(const_cast<mista_engine*>(e)->mu);
	msg = e->last_error;
	return msg.c_str();
}

int mista_engine_submit(mista_engine* e, const mista_job* job, uint64_t* ticket)
{
	if (!e || !job || !ticket) return MISTA_E_INVALID_PARAM;
	const mista_sta_config& c = job->config;
	if (!job->grid_xz || !job->rx_apod || !job->echoes || !job->value || job->n1 == 0 || job->n2 == 0 ||
			job->num_samples == 0 || c.num_elements == 0 || c.last_tx_elem < c.first_tx_elem ||
			c.last_tx_elem >= c.num_elements || c.upsampling_factor == 0) {
		std::lock_guard<std::mutex> lk(e->mu);
		e->last_error = "invalid job";
		return MISTA_E_INVALID_PARAM;
	}
	auto j = std::make_unique<Job>();
	j->spec = *job;
	const uint64_t npix = static_cast<uint64_t>(job->n1) * job->n2;
	j->grid.assign(job->grid_xz, job->grid_xz + 2 * npix);
	j->rx_apod.assign(job->rx_apod, job->rx_apod + c.num_elements);
	if (job->tx_apod) j->tx_apod.assign(job->tx_apod, job->tx_apod + (c.last_tx_elem - c.first_tx_elem + 1));
	j->key = job_key(*job, j->grid);
	j->estimate = estimate_bytes(*job);
	j->t_submit = Clock::now();
	{
		std::lock_guard<std::mutex> lk(e->mu);
		if (e->stop) return MISTA_E_INVALID_STATE;
		if (e->queue.size() >= e->cfg.max_queued) {
			e->last_error = "admission queue full";
			return MISTA_E_BUSY;
		}
		const uint64_t tk = e->next_ticket++;
		j->ticket = tk;
		*ticket = tk;
		e->queue.push_back(j.get());
		e->jobs[tk] = j.release();   // (j->ticket here would read j after release)
	}
	e->cv_work.notify_one();
	return MISTA_OK;
}

int mista_engine_query(mista_engine* e, uint64_t ticket, int* done)
{
	if (!e || !done) return MISTA_E_INVALID_PARAM;
	std::lock_guard<std::mutex> lk(e->mu);
	auto it = e->jobs.find(ticket);
	if (it == e->jobs.end()) return MISTA_E_INVALID_PARAM;
	*done = it->second->done ? 1 : 0;
	return MISTA_OK;
}

int mista_engine_wait(mista_engine* e, uint64_t ticket, mista_job_stats* stats)
{
	if (!e) return MISTA_E_INVALID_PARAM;
	std::unique_lock<std::mutex> lk(e->mu);
	auto it = e->jobs.find(ticket);
	if (it == e->jobs.end()) return MISTA_E_INVALID_PARAM;
	Job* j = it->second;
	e->cv_done.wait(lk, [&] { return j->done; });
	if (stats) {
		stats->status = j->status;
		stats->queued_ms = ms_between(j->t_submit, j->t_admit);
		stats->latency_ms = ms_between(j->t_submit, j->t_done);
		stats->device_ms = j->device_ms;
		stats->footprint = j->footprint;
	}
	const int status = j->status;
	e->jobs.erase(it);
	delete j;
	return status;
}

int mista_engine_stats(mista_engine* e, uint64_t* resident_bytes, uint32_t* inflight, uint32_t* queued)
{
	if (!e) return MISTA_E_INVALID_PARAM;
	std::lock_guard<std::mutex> lk(e->mu);
	if (resident_bytes) *resident_bytes = e->resident;
	if (inflight) *inflight = static_cast<uint32_t>(e->running.size());
	if (queued) *queued = static_cast<uint32_t>(e->queue.size());
	return MISTA_OK;
}

}  // extern "C"

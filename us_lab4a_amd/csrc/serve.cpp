// Serving engine: a request loop over STA imaging jobs of uneven shapes on one
// GPU (include/mista.h, mista_engine_*).
//
// Jobs are submitted from any thread.  A worker thread admits them as
// in-flight slots and device memory free up, scanning the queue in arrival
// order and letting a job that can start now pass one that must wait (the
// head of the queue is passed at most kMaxSkips times, then admission waits
// for it).  Each job runs on a processor handle from a resident pool keyed by
// everything that fixes the handle's device state (configuration,
// apodization, grid, sample count): the handle keeps its delay table, window
// tables, tuned tile shape and FFT plans in HBM between jobs.  A shape holds
// up to `handles_per_shape` handles (default 2), each with its own tensor
// and HIP stream, and takes up to kSlotDepth jobs queued on each stream, so
// the preparation of one frame overlaps the DAS of the one before it, and
// jobs of different shapes overlap too (continuous batching: a job starts as
// soon as a slot frees up, not when a batch completes).
//
// Everything that blocks the host (handle creation, delay-table and
// window-table builds, FFT-plan set-up, the autotuner's timed launches) runs
// on a builder thread, so the worker's admissions never wait on the device:
// a job whose shape is not resident is handed to the builder, which creates
// the handle, runs the job on it (tuning the tile shape unless an earlier
// handle of the shape left a tuned shape in the engine-wide cache), then
// creates the shape's other handles from the cached shape and warms each with
// one frame of placeholder echoes.  Admission is by device-memory footprint:
// a handle is built only if the pool's resident bytes plus its estimate fit
// the budget, after evicting idle handles least recently used first.
//
// Each job's arithmetic is independent of the other jobs in flight (no shared
// accumulators) and every tile shape writes the same images, so a job's
// images are bit-identical alone and in a mix.
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

constexpr uint32_t kSlotDepth = 2;    // jobs queued on one handle's stream
constexpr uint32_t kMaxSkips = 32;    // admissions that may pass the queue's head

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
	hipEvent_t ev_start = nullptr, ev_end = nullptr, ev_prep = nullptr;
	int slot = -1;
	uint32_t skips = 0;      // admissions that passed this job at the queue's head
	bool launched = false;   // its work is on the device (ev_end recorded)
	bool done = false;
	int status = MISTA_OK;
	int das_tile = -1;       // tile shape of the DAS launch (-1: gather)
	double device_ms = 0.0;
	size_t footprint = 0;
};

struct Slot {
	mista_handle* h = nullptr;
	uint64_t key = 0;
	size_t bytes = 0;            // device bytes (estimate until the first job ends)
	uint64_t last_use = 0;
	uint32_t inflight = 0;       // admitted jobs not yet finished
	bool building = false;       // owned by the builder thread
	bool used = false;           // holds (or is building) a handle
	DeviceBuffer d_echoes, d_value, d_factor;   // staging for host inputs/outputs
};

// A builder task: build the slot's handle and run `job` on it, or (job null)
// build a further handle of the shape that spec/grid/apodization describe
// and warm it.
struct BuildTask {
	Job* job = nullptr;
	int slot = -1;
	mista_job spec{};
	std::vector<float> grid, rx_apod, tx_apod;
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

uint64_t job_key(const mista_job& j, const std::vector<float>& grid, const std::vector<float>& rx,
		const std::vector<float>& tx)
{
	mista_sta_config c = j.config;
	c.device = 0;
	uint64_t h = fnv1a(&c, sizeof(c));
	h = fnv1a(&j.n1, sizeof(j.n1), h);
	h = fnv1a(&j.n2, sizeof(j.n2), h);
	h = fnv1a(&j.num_samples, sizeof(j.num_samples), h);
	h = fnv1a(rx.data(), rx.size() * sizeof(float), h);
	const uint32_t ntx_apod = static_cast<uint32_t>(tx.size());
	h = fnv1a(&ntx_apod, sizeof(ntx_apod), h);
	h = fnv1a(tx.data(), tx.size() * sizeof(float), h);
	return fnv1a(grid.data(), grid.size() * sizeof(float), h);
}

}  // namespace

struct mista_engine {
	mista_engine_config cfg{};
	size_t budget = 0;
	uint32_t per_shape = 2;
	std::mutex mu;
	std::condition_variable cv_work, cv_done, cv_build;
	std::deque<Job*> queue;
	std::vector<Job*> running;
	std::unordered_map<uint64_t, Job*> jobs;
	std::deque<Slot> slots;   // stable addresses: handles and buffers never move
	std::deque<BuildTask> builds;
	std::unordered_map<uint64_t, std::pair<int, int>> tuned;   // shape key -> tile shape
	size_t resident = 0;
	uint64_t next_ticket = 1, use_clock = 0;
	uint64_t n_handles = 0, n_tunes = 0, n_evictions = 0;
	uint32_t n_building = 0;
	bool stop = false;
	std::string last_error;
	std::thread worker, builder;
	hipStream_t das_stream = nullptr;   // DAS of warm jobs, batched per admission round
	std::vector<Job*> pending;          // warm jobs prepared, DAS not yet enqueued

	void run();
	void build_loop();
	bool admit_one(std::unique_lock<std::mutex>& lk);
	int place(Job* j, bool& cold);
	bool make_room(size_t bytes);
	int new_slot(uint64_t key, size_t bytes);
	void drop_slot(Slot& s);
	void launch(Job* j, Slot& s, mista_handle* h);
	void prep_job(Job* j, Slot& s, mista_handle* h);
	void flush(std::unique_lock<std::mutex>& lk);
	void fail(Job* j, int status, const std::string& err);
	void finish(Job* j);
	void poll_running();
	void build_one(BuildTask& t, std::unique_lock<std::mutex>& lk);
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
	if (j->ev_prep) (void)hipEventDestroy(j->ev_prep);
	j->ev_start = j->ev_end = j->ev_prep = nullptr;
	if (j->slot >= 0) {
		Slot& s = slots[j->slot];
		s.inflight -= 1;
		if (s.h) {
			const size_t actual = handle_device_bytes(s.h) + s.d_echoes.capacity() + s.d_value.capacity() +
					s.d_factor.capacity();
			resident = resident - s.bytes + actual;
			s.bytes = actual;
			j->footprint = actual;
		}
	}
	j->done = true;
}

void mista_engine::poll_running()
{
	for (size_t i = 0; i < running.size();) {
		Job* j = running[i];
		if (!j->launched) {   // still with the builder
			++i;
			continue;
		}
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

// Enqueues the job on its slot's stream: host inputs are staged in the
// slot's device buffers; outputs copied back when the caller gave host
// pointers.  Enqueues only (no host synchronisation) when the handle has run
// a frame of this shape before.
void mista_engine::launch(Job* j, Slot& s, mista_handle* h)
{
	const mista_job& q = j->spec;
	hipStream_t st = handle_stream(h);
	const uint64_t ntx = q.config.last_tx_elem - q.config.first_tx_elem + 1ull;
	const uint64_t nraw = ntx * q.config.num_elements * q.num_samples;
	const uint64_t npix = static_cast<uint64_t>(q.n1) * q.n2;
	auto ok = [&](int rc) {
		if (rc != MISTA_OK) MISTA_THROW(rc, mista_last_error(h));
	};
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
	int p1 = 0;
	ok(mista_last_kernel(h, &j->das_tile, &p1));
	if (!q.outputs_on_device) {
		HIP_CHECK(hipMemcpyAsync(q.value, d_value, npix * sizeof(float), hipMemcpyDeviceToHost, st));
		if (q.factor) HIP_CHECK(hipMemcpyAsync(q.factor, d_factor, npix * sizeof(float), hipMemcpyDeviceToHost, st));
	}
	HIP_CHECK(hipEventRecord(j->ev_end, st));
}

// Enqueues a warm job's preparation on its slot's stream (host echoes staged
// in the slot's buffer); its DAS follows in the round's batch (flush).
void mista_engine::prep_job(Job* j, Slot& s, mista_handle* h)
{
	const mista_job& q = j->spec;
	hipStream_t st = handle_stream(h);
	const uint64_t ntx = q.config.last_tx_elem - q.config.first_tx_elem + 1ull;
	const uint64_t nraw = ntx * q.config.num_elements * q.num_samples;
	HIP_CHECK(hipEventCreate(&j->ev_start));
	HIP_CHECK(hipEventCreate(&j->ev_end));
	HIP_CHECK(hipEventCreate(&j->ev_prep));
	HIP_CHECK(hipEventRecord(j->ev_start, st));
	const float* d_echoes = q.echoes;
	if (!q.echoes_on_device) {
		float* d = static_cast<float*>(s.d_echoes.reserve(nraw * sizeof(float)));
		HIP_CHECK(hipMemcpyAsync(d, q.echoes, nraw * sizeof(float), hipMemcpyHostToDevice, st));
		d_echoes = d;
	}
	const int rc = mista_prepare_signals_device(h, d_echoes, q.num_samples, nullptr);
	if (rc != MISTA_OK) MISTA_THROW(rc, mista_last_error(h));
	HIP_CHECK(hipEventRecord(j->ev_prep, st));
}

// The DAS of the round's prepared warm jobs: one batched call on the DAS
// stream after their preparations (compatible frames share a launch); each
// job's end is recorded there, and its slot's stream waits for it, so the
// handle's next preparation cannot overwrite the tensor too early.  Called
// with the lock held; released while enqueueing.
void mista_engine::flush(std::unique_lock<std::mutex>& lk)
{
	if (pending.empty()) return;
	std::vector<Job*> js;
	js.swap(pending);
	lk.unlock();
	int status = MISTA_OK;
	std::string err;
	try {
		std::vector<mista_handle*> hs;
		std::vector<float*> vs, fs;
		for (Job* j : js) {
			Slot& s = slots[j->slot];
			const mista_job& q = j->spec;
			const uint64_t npix = static_cast<uint64_t>(q.n1) * q.n2;
			HIP_CHECK(hipStreamWaitEvent(das_stream, j->ev_prep, 0));
			hs.push_back(s.h);
			if (q.outputs_on_device) {
				vs.push_back(q.value);
				fs.push_back(q.factor);
			} else {
				vs.push_back(static_cast<float*>(s.d_value.reserve(npix * sizeof(float))));
				fs.push_back(q.factor ? static_cast<float*>(s.d_factor.reserve(npix * sizeof(float))) : nullptr);
			}
		}
		const int rc = mista_das_batch_device(hs.data(), static_cast<uint32_t>(hs.size()), vs.data(), fs.data(), das_stream);
		if (rc != MISTA_OK) MISTA_THROW(rc, mista_last_error(hs[0]));
		for (size_t k = 0; k < js.size(); ++k) {
			Job* j = js[k];
			const mista_job& q = j->spec;
			const uint64_t npix = static_cast<uint64_t>(q.n1) * q.n2;
			int p1 = 0;
			(void)mista_last_kernel(hs[k], &j->das_tile, &p1);
			if (!q.outputs_on_device) {
				HIP_CHECK(hipMemcpyAsync(q.value, vs[k], npix * sizeof(float), hipMemcpyDeviceToHost, das_stream));
				if (q.factor) HIP_CHECK(hipMemcpyAsync(q.factor, fs[k], npix * sizeof(float), hipMemcpyDeviceToHost, das_stream));
			}
			HIP_CHECK(hipEventRecord(j->ev_end, das_stream));
			HIP_CHECK(hipStreamWaitEvent(handle_stream(hs[k]), j->ev_end, 0));
		}
	} catch (const Error& e) {
		status = e.code;
		err = e.what();
	} catch (const std::exception& e) {
		status = MISTA_E_INVALID_STATE;
		err = e.what();
	}
	lk.lock();
	for (Job* j : js) {
		if (status != MISTA_OK) fail(j, status, err);
		else j->launched = true;
	}
}

// Completes a job that could not run (called with the lock held).
void mista_engine::fail(Job* j, int status, const std::string& err)
{
	if (!err.empty()) last_error = err;
	j->status = status;
	if (j->ev_start) (void)hipEventDestroy(j->ev_start);
	if (j->ev_end) (void)hipEventDestroy(j->ev_end);
	if (j->ev_prep) (void)hipEventDestroy(j->ev_prep);
	j->ev_start = j->ev_end = j->ev_prep = nullptr;
	auto it = std::find(running.begin(), running.end(), j);
	if (it != running.end()) running.erase(it);
	if (j->slot >= 0) slots[j->slot].inflight -= 1;
	j->slot = -1;
	j->done = true;
	j->t_done = Clock::now();
	cv_done.notify_all();
}

void mista_engine::drop_slot(Slot& s)
{
	if (s.h) (void)mista_destroy(s.h);
	s.h = nullptr;
	s.d_echoes.release();
	s.d_value.release();
	s.d_factor.release();
	resident -= std::min(resident, s.bytes);
	s.bytes = 0;
	s.used = false;
	s.building = false;
	s.inflight = 0;
}

// Evicts idle handles, least recently used first, until `bytes` more fit
// the budget.  Returns whether they fit.
bool mista_engine::make_room(size_t bytes)
{
	while (resident + bytes > budget) {
		int lru = -1;
		for (size_t i = 0; i < slots.size(); ++i) {
			const Slot& c = slots[i];
			if (c.used && c.h && !c.building && c.inflight == 0 && (lru < 0 || c.last_use < slots[lru].last_use))
				lru = static_cast<int>(i);
		}
		if (lru < 0) return false;
		drop_slot(slots[lru]);
		++n_evictions;
	}
	return true;
}

int mista_engine::new_slot(uint64_t key, size_t bytes)
{
	int si = -1;
	for (size_t i = 0; i < slots.size(); ++i)
		if (!slots[i].used) {
			si = static_cast<int>(i);
			break;
		}
	if (si < 0) {
		slots.emplace_back();
		si = static_cast<int>(slots.size() - 1);
	}
	Slot& s = slots[si];
	s.used = true;
	s.building = true;
	s.key = key;
	s.bytes = bytes;
	s.inflight = 0;
	s.last_use = ++use_clock;
	resident += bytes;
	++n_building;
	return si;
}

// Where the job can start now: a slot index (cold = the slot is new and goes
// to the builder), -1 to wait, -2 when it can never fit the memory budget.
int mista_engine::place(Job* j, bool& cold)
{
	cold = false;
	int best = -1;
	bool known = false;
	for (size_t i = 0; i < slots.size(); ++i) {
		const Slot& c = slots[i];
		if (!c.used || c.key != j->key) continue;
		known = true;
		if (c.building || !c.h || c.inflight >= kSlotDepth) continue;
		if (best < 0 || c.inflight < slots[best].inflight) best = static_cast<int>(i);
	}
	if (best >= 0) return best;
	if (known) return -1;   // the shape is resident or being built: wait for its handles
	if (!make_room(j->estimate)) {
		if (running.empty() && n_building == 0) return -2;
		return -1;
	}
	cold = true;
	return new_slot(j->key, j->estimate);
}

// Admits the first job in the queue that can start now (the head first; a
// later job may pass a head that must wait, at most kMaxSkips times).
// Returns false when none can.  Called with the lock held; the lock is
// released while a warm job's work is enqueued.
bool mista_engine::admit_one(std::unique_lock<std::mutex>& lk)
{
	if (queue.empty() || running.size() >= cfg.max_inflight) return false;
	const bool head_only = queue.front()->skips >= kMaxSkips;
	for (auto it = queue.begin(); it != queue.end(); ++it) {
		Job* j = *it;
		bool cold = false;
		const int si = place(j, cold);
		if (si == -2) {
			queue.erase(it);
			j->t_admit = Clock::now();
			fail(j, MISTA_E_NO_MEMORY, "job does not fit the device-memory budget");
			return true;
		}
		if (si < 0) {
			if (head_only) return false;
			continue;
		}
		if (it != queue.begin()) queue.front()->skips += 1;
		queue.erase(it);
		Slot& s = slots[si];
		s.inflight += 1;
		s.last_use = ++use_clock;
		j->slot = si;
		j->t_admit = Clock::now();
		running.push_back(j);
		if (cold) {
			BuildTask t;
			t.job = j;
			t.slot = si;
			builds.push_back(std::move(t));
			cv_build.notify_one();
			return true;
		}
		// a second job on a handle prepares only after the first one's DAS
		// is enqueued (its slot stream then waits for that DAS)
		for (Job* o : pending)
			if (o->slot == si) {
				flush(lk);
				break;
			}
		lk.unlock();
		int status = MISTA_OK;
		std::string err;
		try {
			prep_job(j, s, s.h);
		} catch (const Error& e) {
			status = e.code;
			err = e.what();
		} catch (const std::exception& e) {
			status = MISTA_E_INVALID_STATE;
			err = e.what();
		}
		lk.lock();
		if (status != MISTA_OK) fail(j, status, err);
		else pending.push_back(j);
		return true;
	}
	return false;
}

void mista_engine::run()
{
	(void)hipSetDevice(cfg.device);
	std::unique_lock<std::mutex> lk(mu);
	for (;;) {
		poll_running();
		bool progressed = false;
		while (admit_one(lk)) progressed = true;
		flush(lk);
		if (stop && queue.empty() && running.empty() && n_building == 0) break;
		if (progressed) continue;
		if (running.empty() && n_building == 0) {
			cv_work.wait(lk, [&] { return stop || !queue.empty(); });
		} else {
			// jobs on the device or handles being built: poll their events
			cv_work.wait_for(lk, std::chrono::microseconds(50));
		}
	}
}

// Builds one handle (called with the lock held; released while building).
void mista_engine::build_one(BuildTask& t, std::unique_lock<std::mutex>& lk)
{
	Slot& s = slots[t.slot];
	const mista_job& q = t.job ? t.job->spec : t.spec;
	const std::vector<float>& grid = t.job ? t.job->grid : t.grid;
	const std::vector<float>& rx = t.job ? t.job->rx_apod : t.rx_apod;
	const std::vector<float>& tx = t.job ? t.job->tx_apod : t.tx_apod;
	const uint64_t key = s.key;
	std::pair<int, int> tile{-1, 0};
	auto it = tuned.find(key);
	if (it != tuned.end()) tile = it->second;
	lk.unlock();
	int status = MISTA_OK;
	std::string err;
	mista_handle* h = nullptr;
	bool tuned_now = false;
	std::pair<int, int> tile_now{-1, 0};
	try {
		auto ok = [&](int rc) {
			if (rc != MISTA_OK) MISTA_THROW(rc, mista_last_error(h));
		};
		mista_sta_config c = q.config;
		c.device = cfg.device;
		const int rc = mista_create(&c, &h);
		if (rc != MISTA_OK) MISTA_THROW(rc, mista_last_error(nullptr));
		ok(mista_set_grid(h, grid.data(), q.n1, q.n2));
		ok(mista_set_rx_apod(h, rx.data(), static_cast<uint32_t>(rx.size())));
		ok(mista_set_tx_apod(h, tx.empty() ? nullptr : tx.data(), static_cast<uint32_t>(tx.size())));
		if (tile.first >= 0) ok(mista_set_kernel(h, 2, tile.first, tile.second));
		if (t.job) {
			launch(t.job, s, h);
		} else {
			// warm the handle: one frame of placeholder echoes (finite, all
			// equal) into the slot's staging buffers, then drop the staging
			const uint64_t ntx = q.config.last_tx_elem - q.config.first_tx_elem + 1ull;
			const uint64_t nraw = ntx * q.config.num_elements * q.num_samples;
			const uint64_t npix = static_cast<uint64_t>(q.n1) * q.n2;
			hipStream_t st = handle_stream(h);
			float* d = static_cast<float*>(s.d_echoes.reserve(nraw * sizeof(float)));
			HIP_CHECK(hipMemsetAsync(d, 0x3c, nraw * sizeof(float), st));
			float* v = static_cast<float*>(s.d_value.reserve(npix * sizeof(float)));
			float* f = static_cast<float*>(s.d_factor.reserve(npix * sizeof(float)));
			ok(mista_process_device(h, d, q.num_samples, v, f, nullptr));
			HIP_CHECK(hipStreamSynchronize(st));
			s.d_echoes.release();
			s.d_value.release();
			s.d_factor.release();
		}
		if (tile.first < 0) {
			ok(mista_tuned_kernel(h, &tile_now.first, &tile_now.second));
			tuned_now = tile_now.first >= 0;
		}
	} catch (const Error& e) {
		status = e.code;
		err = e.what();
	} catch (const std::exception& e) {
		status = MISTA_E_INVALID_STATE;
		err = e.what();
	}
	if (status != MISTA_OK && h) {
		// keep nothing half-built (a launched job's work is drained first)
		(void)hipStreamSynchronize(handle_stream(h));
		(void)mista_destroy(h);
		h = nullptr;
	}
	lk.lock();
	--n_building;
	s.building = false;
	if (tuned_now) {
		++n_tunes;
		tuned[key] = tile_now;
	}
	if (status != MISTA_OK) {
		if (t.job) fail(t.job, status, err);
		else last_error = err;
		drop_slot(s);
		return;
	}
	++n_handles;
	s.h = h;
	const size_t actual = handle_device_bytes(h) + s.d_echoes.capacity() + s.d_value.capacity() + s.d_factor.capacity();
	resident = resident - s.bytes + actual;
	s.bytes = actual;
	if (t.job) t.job->launched = true;
	if (t.job) {
		// the shape's other handles, from the tuned tile shape
		uint32_t have = 0;
		for (const Slot& c : slots)
			if (c.used && c.key == key) ++have;
		for (; have < per_shape; ++have) {
			if (!make_room(t.job->estimate)) break;
			BuildTask c;
			c.slot = new_slot(key, t.job->estimate);
			c.spec = t.job->spec;
			c.grid = t.job->grid;
			c.rx_apod = t.job->rx_apod;
			c.tx_apod = t.job->tx_apod;
			builds.push_back(std::move(c));
		}
	}
}

void mista_engine::build_loop()
{
	(void)hipSetDevice(cfg.device);
	std::unique_lock<std::mutex> lk(mu);
	for (;;) {
		cv_build.wait(lk, [&] { return !builds.empty() || (stop && n_building == 0); });
		if (builds.empty()) break;
		BuildTask t = std::move(builds.front());
		builds.pop_front();
		build_one(t, lk);
		cv_work.notify_all();
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
		if (config->handles_per_shape > 8) MISTA_THROW(MISTA_E_INVALID_PARAM, "handles_per_shape must be in [0, 8].");
		auto e = std::make_unique<mista_engine>();
		e->cfg = *config;
		e->per_shape = config->handles_per_shape ? config->handles_per_shape : 2u;
		if (config->hbm_budget) {
			e->budget = config->hbm_budget;
		} else {
			size_t free_b = 0, total_b = 0;
			HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
			e->budget = free_b / 10 * 8;
		}
		HIP_CHECK(hipStreamCreateWithFlags(&e->das_stream, hipStreamNonBlocking));
		mista_engine* p = e.get();
		e->worker = std::thread([p] { p->run(); });
		e->builder = std::thread([p] { p->build_loop(); });
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
	e->cv_build.notify_all();
	// the worker ends once every job has finished and no handle is being
	// built; the builder then finds nothing left to do
	if (e->worker.joinable()) e->worker.join();
	e->cv_build.notify_all();
	if (e->builder.joinable()) e->builder.join();
	(void)hipSetDevice(e->cfg.device);
	for (auto& s : e->slots)
		if (s.h) (void)mista_destroy(s.h);
	if (e->das_stream) (void)hipStreamDestroy(e->das_stream);
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
	j->key = job_key(*job, j->grid, j->rx_apod, j->tx_apod);
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
		stats->das_tile = j->das_tile;
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

int mista_engine_counters(mista_engine* e, uint64_t* handles_created, uint64_t* tunes, uint64_t* evictions,
		uint32_t* building)
{
	if (!e) return MISTA_E_INVALID_PARAM;
	std::lock_guard<std::mutex> lk(e->mu);
	if (handles_created) *handles_created = e->n_handles;
	if (tunes) *tunes = e->n_tunes;
	if (evictions) *evictions = e->n_evictions;
	if (building) *building = e->n_building;
	return MISTA_OK;
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

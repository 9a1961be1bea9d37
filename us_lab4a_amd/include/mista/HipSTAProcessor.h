// C++ STA processor over the mista C ABI (include/mista.h).
//
// Same contract as the reference's array-processor family
// (src/user/processor/ArrayProcessor.h:27-40; constructor of
// VectorialSTAProcessor.h:61-67 / VectorialSTACUDAProcessor.h:65-71):
//
//   HipSTAProcessor<TPoint> p(config, acquisition, upsamplingFactor,
//                             coherenceFactor, peakOffset, rxApod);
//   p.prepare(baseElement);
//   p.process(gridData);      // gridData: Matrix-like, n1() x n2() points
//                             // with x, z read and value, factor written
//
// Header-only; no HIP types appear here.  Errors from the library are thrown
// as the reference's exception kinds (Exception.h:38-94):
// InvalidParameterException, InvalidValueException, InvalidStateException,
// and GpuException (the CUDAException analogue) for HIP failures.
#ifndef MISTA_HIP_STA_PROCESSOR_H
#define MISTA_HIP_STA_PROCESSOR_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mista.h"

namespace mista {

struct InvalidParameterException : std::runtime_error { using std::runtime_error::runtime_error; };
struct InvalidValueException : std::runtime_error { using std::runtime_error::runtime_error; };
struct InvalidStateException : std::runtime_error { using std::runtime_error::runtime_error; };
struct GpuException : std::runtime_error { using std::runtime_error::runtime_error; };

inline void check(int rc, const mista_handle* h)
{
	if (rc == MISTA_OK) return;
	const std::string msg = mista_last_error(h);
	switch (rc) {
	case MISTA_E_INVALID_PARAM: throw InvalidParameterException(msg);
	case MISTA_E_INVALID_VALUE: throw InvalidValueException(msg);
	case MISTA_E_INVALID_STATE: throw InvalidStateException(msg);
	default: throw GpuException(msg);
	}
}

// Fields of STAConfiguration<float> used by the path (STAConfiguration.h:118-165).
struct STAConfiguration {
	unsigned int numElements = 0;
	unsigned int firstTxElem = 0;
	unsigned int lastTxElem = 0;
	float pitch = 0;
	float centerFrequency = 0;
	float samplingFrequency = 0;
	float propagationSpeed = 0;
	float deadZoneM = 0;
};

// AnalyticSignalCoherenceFactorProcessor selection (CoherenceFactor.h:139-152)
// plus the sign coherence factor on Re(per-rx sum).
struct CoherenceFactor {
	int kind = MISTA_CF_NONE;   // MISTA_CF_NONE / MISTA_CF_PCF / MISTA_CF_SCF
	float param = 1.0f;         // gamma (PCF) or p (SCF)
	bool enabled() const { return kind != MISTA_CF_NONE; }
};

// Echo source (STAAcquisition.h:71-86): execute(tx) fills rx x samples,
// row-major (Matrix[rx][samples]).
class STAAcquisition {
public:
	virtual ~STAAcquisition() = default;
	virtual void prepare(unsigned int /*baseElement*/) {}
	virtual void execute(unsigned int txElem, std::vector<float>& rxSamples, unsigned int& numSamples) = 0;
};

// Point layout of Matrix<XYZValueFactor<float>> (XYZValueFactor.h:23-32).
struct XYZValueFactor {
	float x = 0, y = 0, z = 0, value = 0, factor = 1;
};

template<typename TPoint = XYZValueFactor>
class HipSTAProcessor {
public:
	HipSTAProcessor(const STAConfiguration& config, STAAcquisition& acquisition, unsigned int upsamplingFactor,
			const CoherenceFactor& coherenceFactor, float peakOffset, const std::vector<float>& rxApod,
			int storage = MISTA_STORAGE_F32, int device = 0)
		: config_(config), acquisition_(acquisition)
	{
		mista_sta_config c{};
		c.num_elements = config.numElements;
		c.first_tx_elem = config.firstTxElem;
		c.last_tx_elem = config.lastTxElem;
		c.upsampling_factor = upsamplingFactor;
		c.pitch = config.pitch;
		c.sampling_frequency = config.samplingFrequency;
		c.propagation_speed = config.propagationSpeed;
		c.center_frequency = config.centerFrequency;
		c.dead_zone_m = config.deadZoneM;
		c.peak_offset = peakOffset;
		c.cf_kind = coherenceFactor.kind;
		c.cf_param = coherenceFactor.param;
		c.storage = storage;
		c.device = device;
		check(mista_create(&c, &h_), nullptr);
		try {
			check(mista_set_rx_apod(h_, rxApod.data(), static_cast<uint32_t>(rxApod.size())), h_);
		} catch (...) {
			mista_destroy(h_);
			throw;
		}
	}
	~HipSTAProcessor() { mista_destroy(h_); }
	HipSTAProcessor(const HipSTAProcessor&) = delete;
	HipSTAProcessor& operator=(const HipSTAProcessor&) = delete;
	HipSTAProcessor(HipSTAProcessor&&) = delete;
	HipSTAProcessor& operator=(HipSTAProcessor&&) = delete;

	// ArrayProcessor::prepare (VectorialSTAProcessor.h:151-156)
	void prepare(unsigned int baseElement)
	{
		acquisition_.prepare(baseElement);
		check(mista_prepare(h_, baseElement), h_);
	}

	// ArrayProcessor::process (VectorialSTAProcessor.h:158-341): pulls every
	// transmit from the acquisition, forms the image on the GPU and writes
	// value (|sum| / (Ntx*Nrx)) and factor into gridData in place.
	// GridData needs n1(), n2() and operator()(i, j) returning TPoint&.
	template<typename GridData>
	void process(GridData& gridData)
	{
		const uint32_t n1 = static_cast<uint32_t>(gridData.n1()), n2 = static_cast<uint32_t>(gridData.n2());
		std::vector<float> xz(static_cast<size_t>(n1) * n2 * 2);
		for (uint32_t i = 0; i < n1; ++i)
			for (uint32_t j = 0; j < n2; ++j) {
				const TPoint& p = gridData(i, j);
				xz[(static_cast<size_t>(i) * n2 + j) * 2] = p.x;
				xz[(static_cast<size_t>(i) * n2 + j) * 2 + 1] = p.z;
			}
		if (xz != gridXZ_) {
			check(mista_set_grid(h_, xz.data(), n1, n2), h_);
			gridXZ_.swap(xz);
		}
		const unsigned int ntx = config_.lastTxElem - config_.firstTxElem + 1U;
		std::vector<float> rx;
		unsigned int ns = 0;
		for (unsigned int t = 0; t < ntx; ++t) {
			unsigned int nsT = 0;
			acquisition_.execute(config_.firstTxElem + t, rx, nsT);
			if (t == 0) {
				ns = nsT;
				echoes_.resize(static_cast<size_t>(ntx) * config_.numElements * ns);
			} else if (nsT != ns) {
				throw InvalidValueException("The number of samples changed between transmits.");
			}
			if (rx.size() != static_cast<size_t>(config_.numElements) * ns)
				throw InvalidParameterException("Wrong acquisition size.");
			std::copy(rx.begin(), rx.end(), echoes_.begin() + static_cast<size_t>(t) * config_.numElements * ns);
		}
		const size_t npts = static_cast<size_t>(n1) * n2;
		value_.resize(npts);
		factor_.resize(npts);
		check(mista_process(h_, echoes_.data(), ns, value_.data(), factor_.data()), h_);
		for (uint32_t i = 0; i < n1; ++i)
			for (uint32_t j = 0; j < n2; ++j) {
				TPoint& p = gridData(i, j);
				p.value = value_[static_cast<size_t>(i) * n2 + j];
				p.factor = factor_[static_cast<size_t>(i) * n2 + j];
			}
	}

	mista_handle* handle() { return h_; }

private:
	STAConfiguration config_;
	STAAcquisition& acquisition_;
	mista_handle* h_ = nullptr;
	std::vector<float> gridXZ_, echoes_, value_, factor_;
};

}  // namespace mista

#endif

// The HIP STA processor in the reference's own types: a Lab::ArrayProcessor
// (src/user/processor/ArrayProcessor.h:27-40) whose constructor takes what
// VectorialSTAProcessor's takes (VectorialSTAProcessor.h:130-136):
//
//   Lab::HipVectorialSTAProcessor<float, XYZValueFactor<float>> p(
//       config,            // const STAConfiguration<TFloat>&
//       acquisition,       // STAAcquisition<TFloat>&
//       upsamplingFactor,  // unsigned int
//       coherenceFactor,   // AnalyticSignalCoherenceFactorProcessor<TFloat>&
//       peakOffset,        // TFloat
//       rxApod);           // const std::vector<TFloat>&
//   p.prepare(baseElement);
//   p.process(gridData);   // Matrix<TPoint>: x, z read; value, factor written
//
// so STAMethod can hold it as std::unique_ptr<ArrayProcessor<TPoint>> like
// the other processors.  Compiled inside the reference's build: it includes
// the reference's headers by name (ArrayProcessor.h, CoherenceFactor.h,
// Exception.h, Matrix.h, STAAcquisition.h, STAConfiguration.h), plus
// include/mista.h, and links libmista_hip.so.  Library errors are thrown as
// the reference's exception kinds (Exception.h:74-89); HIP failures as
// Lab::HipException (the CUDAException analogue, CUDAUtil.h:20).
#ifndef MISTA_LAB_HIP_STA_PROCESSOR_H
#define MISTA_LAB_HIP_STA_PROCESSOR_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ArrayProcessor.h"
#include "CoherenceFactor.h"
#include "Exception.h"
#include "Matrix.h"
#include "STAAcquisition.h"
#include "STAConfiguration.h"
#include "mista.h"

namespace Lab {

struct HipException : std::runtime_error {
	using std::runtime_error::runtime_error;
};

// Status of a mista call -> the reference's exception kinds.
inline void mistaCheck(int rc, const mista_handle* h)
{
	if (rc == MISTA_OK) return;
	const std::string msg = mista_last_error(h);
	switch (rc) {
	case MISTA_E_INVALID_PARAM: throw InvalidParameterException(msg);
	case MISTA_E_INVALID_VALUE: throw InvalidValueException(msg);
	case MISTA_E_INVALID_STATE: throw InvalidStateException(msg);
	case MISTA_E_NO_MEMORY: throw UnavailableResourceException(msg);
	default: throw HipException(msg);
	}
}

// Pulls the transmits of an STAAcquisition<TFloat> (execute(tx, Matrix):
// rx rows of samples, STAAcquisition.h:36) into one float32 frame
// [tx][rx][samples], the layout of mista_process.
template<typename TFloat>
class AcquisitionBridge {
public:
	explicit AcquisitionBridge(STAAcquisition<TFloat>& acquisition) : acquisition_(acquisition) {}

	void prepare(unsigned int baseElement) { acquisition_.prepare(baseElement); }

	// Transmits firstTx .. firstTx + numTx - 1 of numRx channels each;
	// returns the samples per channel (the same for every transmit).
	unsigned int pullFrame(unsigned int firstTx, unsigned int numTx, unsigned int numRx, std::vector<float>& frame)
	{
		unsigned int ns = 0;
		for (unsigned int t = 0; t < numTx; ++t) {
			acquisition_.execute(firstTx + t, acqData_);
			if (acqData_.n1() != numRx) THROW_EXCEPTION(InvalidValueException, "Wrong number of channels: " << acqData_.n1() << '.');
			if (t == 0) {
				ns = static_cast<unsigned int>(acqData_.n2());
				frame.resize(static_cast<std::size_t>(numTx) * numRx * ns);
			} else if (acqData_.n2() != ns) {
				THROW_EXCEPTION(InvalidValueException, "The number of samples changed between transmits.");
			}
			float* dst = frame.data() + static_cast<std::size_t>(t) * numRx * ns;
			const TFloat* src = acqData_.data();
			for (std::size_t i = 0, n = static_cast<std::size_t>(numRx) * ns; i < n; ++i) dst[i] = static_cast<float>(src[i]);
		}
		return ns;
	}

private:
	STAAcquisition<TFloat>& acquisition_;
	typename STAAcquisition<TFloat>::AcquisitionDataType acqData_;
};

template<typename TFloat, typename TPoint>
class HipVectorialSTAProcessor : public ArrayProcessor<TPoint> {
public:
	// storage: MISTA_STORAGE_F32 (the reference's precision), _F16 or _BF16
	// (per-channel scale, FP32 accumulation); device: HIP device ordinal.
	HipVectorialSTAProcessor(const STAConfiguration<TFloat>& config, STAAcquisition<TFloat>& acquisition,
			unsigned int upsamplingFactor, AnalyticSignalCoherenceFactorProcessor<TFloat>& coherenceFactor,
			TFloat peakOffset, const std::vector<TFloat>& rxApod, int storage = MISTA_STORAGE_F32, int device = 0)
		: firstTx_(config.firstTxElem)
		, numTx_(config.lastTxElem - config.firstTxElem + 1U)
		, numRx_(config.numElements)
		, bridge_(acquisition)
	{
		mista_sta_config c{};
		c.num_elements = config.numElements;
		c.first_tx_elem = config.firstTxElem;
		c.last_tx_elem = config.lastTxElem;
		c.upsampling_factor = upsamplingFactor;
		c.pitch = static_cast<float>(config.pitch);
		c.sampling_frequency = static_cast<float>(config.samplingFrequency);
		c.propagation_speed = static_cast<float>(config.propagationSpeed);
		c.center_frequency = static_cast<float>(config.centerFrequency);
		c.dead_zone_m = static_cast<float>(config.deadZoneM);
		c.peak_offset = static_cast<float>(peakOffset);
		c.cf_kind = MISTA_CF_NONE;
		c.cf_param = 1.0f;
		if (coherenceFactor.enabled()) {
			// the one analytic-signal coherence factor, PhaseCoherenceFactor:
			// constants (gamma, sigma0, gamma / sigma0) (CoherenceFactor.h:288-296)
			std::vector<TFloat> k;
			coherenceFactor.implementation().getConstants(k);
			if (k.empty()) THROW_EXCEPTION(InvalidParameterException, "Unknown coherence factor.");
			c.cf_kind = MISTA_CF_PCF;
			c.cf_param = static_cast<float>(k[0]);
		}
		c.storage = storage;
		c.device = device;
		mistaCheck(mista_create(&c, &h_), nullptr);
		try {
			std::vector<float> apod(rxApod.begin(), rxApod.end());
			mistaCheck(mista_set_rx_apod(h_, apod.data(), static_cast<uint32_t>(apod.size())), h_);
		} catch (...) {
			mista_destroy(h_);
			throw;
		}
	}
	~HipVectorialSTAProcessor() override { mista_destroy(h_); }

	// VectorialSTAProcessor::prepare (VectorialSTAProcessor.h:151-156)
	void prepare(unsigned int baseElement) override
	{
		bridge_.prepare(baseElement);
		mistaCheck(mista_prepare(h_, baseElement), h_);
	}

	// VectorialSTAProcessor::process (VectorialSTAProcessor.h:158-341)
	void process(Matrix<TPoint>& gridData) override
	{
		const uint32_t n1 = static_cast<uint32_t>(gridData.n1()), n2 = static_cast<uint32_t>(gridData.n2());
		std::vector<float> xz(static_cast<std::size_t>(n1) * n2 * 2);
		for (uint32_t i = 0; i < n1; ++i)
			for (uint32_t j = 0; j < n2; ++j) {
				const TPoint& p = gridData(i, j);
				xz[(static_cast<std::size_t>(i) * n2 + j) * 2] = static_cast<float>(p.x);
				xz[(static_cast<std::size_t>(i) * n2 + j) * 2 + 1] = static_cast<float>(p.z);
			}
		if (xz != gridXZ_) {
			mistaCheck(mista_set_grid(h_, xz.data(), n1, n2), h_);
			gridXZ_.swap(xz);
		}
		const unsigned int ns = bridge_.pullFrame(firstTx_, numTx_, numRx_, echoes_);
		const std::size_t npts = static_cast<std::size_t>(n1) * n2;
		value_.resize(npts);
		factor_.resize(npts);
		mistaCheck(mista_process(h_, echoes_.data(), ns, value_.data(), factor_.data()), h_);
		for (uint32_t i = 0; i < n1; ++i)
			for (uint32_t j = 0; j < n2; ++j) {
				TPoint& p = gridData(i, j);
				p.value = value_[static_cast<std::size_t>(i) * n2 + j];
				p.factor = factor_[static_cast<std::size_t>(i) * n2 + j];
			}
	}

	mista_handle* handle() { return h_; }

private:
	unsigned int firstTx_, numTx_, numRx_;
	AcquisitionBridge<TFloat> bridge_;
	mista_handle* h_ = nullptr;
	std::vector<float> gridXZ_, echoes_, value_, factor_;
};

}  // namespace Lab

#endif

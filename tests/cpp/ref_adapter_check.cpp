// Drives Lab::HipVectorialSTAProcessor through the reference's own types,
// compiled against the reference's headers (and KeyValueFileReader.cpp for
// ParameterMap): STAConfiguration<float>, an STAAcquisition<float> subclass,
// AnalyticSignalCoherenceFactorProcessor<float> read from a
// config-coherence_factor.txt, Matrix<XYZValueFactor<float>>, held as
// std::unique_ptr<ArrayProcessor<XYZValueFactor<float>>> as STAMethod holds
// its processors.  argv[1]: the coherence-factor file.  Prints one line:
// "image <min value> <max value> <max factor>" after a frame, or
// "unavailable <exception kind>" when no GPU can be used.
#include <cmath>
#include <cstdio>
#include <memory>

#include "mista/LabHipSTAProcessor.h"
#include "ParameterMap.h"
#include "XYZValueFactor.h"

namespace {

// Point scatterer at (0, 20 mm): a 5 MHz pulse at every round-trip time.
class PointAcquisition : public Lab::STAAcquisition<float> {
public:
	PointAcquisition(unsigned int n, float pitch, float fs, float c) : n_(n), pitch_(pitch), fs_(fs), c_(c) {}
	void execute(unsigned int tx, AcquisitionDataType& acq) override
	{
		const unsigned int ns = 2048;
		acq.resize(n_, ns);
		const float xt = tx * pitch_ - (n_ - 1) * pitch_ / 2;
		for (unsigned int r = 0; r < n_; ++r) {
			const float xr = r * pitch_ - (n_ - 1) * pitch_ / 2;
			const float t0 = (std::sqrt(xt * xt + 4e-4f) + std::sqrt(xr * xr + 4e-4f)) / c_;
			for (unsigned int i = 0; i < ns; ++i) {
				const float t = i / fs_ - t0;
				acq(r, i) = std::exp(-t * t * 1e13f) * std::cos(2.0f * 3.14159265f * 5e6f * t);
			}
		}
	}
private:
	unsigned int n_;
	float pitch_, fs_, c_;
};

}  // namespace

int main(int argc, char** argv)
{
	using Point = Lab::XYZValueFactor<float>;
	if (argc < 2) return 2;
	Lab::STAConfiguration<float> cfg;
	cfg.numElements = 32;
	cfg.numElementsMux = 32;
	cfg.firstTxElem = 0;
	cfg.lastTxElem = 31;
	cfg.pitch = 0.3e-3f;
	cfg.centerFrequency = 5e6f;
	cfg.samplingFrequency = 40e6f;
	cfg.propagationSpeed = 1540.0f;
	cfg.deadZoneM = 0.0f;
	PointAcquisition acq(cfg.numElements, cfg.pitch, cfg.samplingFrequency, cfg.propagationSpeed);
	const Lab::ParameterMap pm(argv[1]);
	Lab::AnalyticSignalCoherenceFactorProcessor<float> cf(pm);
	std::vector<float> rxApod(cfg.numElements, 1.0f);
	Lab::Matrix<Point> grid(64, 64);
	for (unsigned int i = 0; i < 64; ++i)
		for (unsigned int j = 0; j < 64; ++j) {
			grid(i, j).x = -4e-3f + 8e-3f * i / 63;
			grid(i, j).z = 16e-3f + 8e-3f * j / 63;
		}
	try {
		std::unique_ptr<Lab::ArrayProcessor<Point>> p =
				std::make_unique<Lab::HipVectorialSTAProcessor<float, Point>>(cfg, acq, 4, cf, 0.0f, rxApod);
		p->prepare(0);
		p->process(grid);
	} catch (const Lab::HipException& e) {
		std::printf("unavailable HipException\n");
		return 0;
	} catch (const Lab::InvalidParameterException& e) {
		std::printf("unavailable InvalidParameterException\n");
		return 0;
	} catch (const Lab::UnavailableResourceException& e) {
		std::printf("unavailable UnavailableResourceException\n");
		return 0;
	}
	float vmin = 1e30f, vmax = -1e30f, fmax = 0.0f;
	for (unsigned int i = 0; i < 64; ++i)
		for (unsigned int j = 0; j < 64; ++j) {
			vmin = std::fmin(vmin, grid(i, j).value);
			vmax = std::fmax(vmax, grid(i, j).value);
			fmax = std::fmax(fmax, grid(i, j).factor);
		}
	std::printf("image %g %g %g\n", vmin, vmax, fmax);
	return 0;
}

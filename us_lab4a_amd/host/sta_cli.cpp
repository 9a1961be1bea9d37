// Command-line STA imaging driver over the C++ processor adapter.
//
// Replays the reference's STAMethod call sequence (STAMethod.cpp:120-263 and
// :96-118): configuration -> image grid -> rx apodization -> coherence
// factor -> processor.prepare(base) -> processor.process(grid) -> optional
// coherence-factor image (Util::applyFactorToValue, Util.h:1052-1060).
// Echoes come from a raw float32 file [ntx][nrx][ns] (what the reference's
// SavedSTAAcquisition replays per transmit, SavedSTAAcquisition.h:78-90).
//
//   mista_sta_cli (--echoes FILE | --acq-dir DIR [--base B]) --elements N
//                 --samples NS --up U --pitch M --fs HZ --c MPS --fc HZ
//                 --grid N1 N2 XMIN XMAX ZMIN ZMAX
//                 [--cf none|pcf|scf] [--cf-param P] [--apod hanning|rectangular]
//                 [--storage f32|f16|bf16] [--out PREFIX] [--h5-out DIR]
//   mista_sta_cli --echoes FILE --elements N --samples NS --write-acq DIR [--base B]
//   mista_sta_cli --sta-config F [--imag-config F] [--cf-config F] [--grid-config F]
//                 (--echoes FILE --samples NS | --acq-dir DIR [--base B]) [--dump-config DIR]
//
// The --*-config options read the reference's project files (config-sta.txt,
// config-imaging.txt, config-coherence_factor.txt, config-grid.txt, see
// project_config.h) in place of the matching command-line values; the grid
// then follows image_grid_type with the Nyquist wavelength of max_frequency
// (STAMethod.cpp:176-180).  --dump-config DIR writes the resolved grid
// (grid_x.f64, grid_z.f64, [N1][N2]), rx apodization (apod.f64) and
// config.txt, and exits without GPU use.
//
// Echoes come from a raw float32 file [tx][rx][ns] or from the reference's
// saved acquisition (DIR/signals-base%04d-tx%04d.h5, dataset "signal"
// [rx][samples]; transmits 0 .. last present).  Writes PREFIX_value.f32 and
// PREFIX_factor.f32 ([N1][N2], z fastest) and, with a coherence factor,
// PREFIX_cf.f32 (value * factor); with --h5-out, the reference's image files
// image_value/factor/cf/x/z.h5.  --write-acq converts a raw echo file into
// the saved-acquisition layout and exits (no GPU use).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "mista/HipSTAProcessor.h"
#include "project_config.h"
#ifdef MISTA_HAVE_HDF5
#include "h5io.h"
#endif

namespace {

// Matrix<XYZValueFactor<float>> stand-in: row-major n1 x n2 (Matrix.h:188-191).
struct Grid {
	uint32_t n1_ = 0, n2_ = 0;
	std::vector<mista::XYZValueFactor> pts;
	size_t n1() const { return n1_; }
	size_t n2() const { return n2_; }
	mista::XYZValueFactor& operator()(size_t i, size_t j) { return pts[i * n2_ + j]; }
	const mista::XYZValueFactor& operator()(size_t i, size_t j) const { return pts[i * n2_ + j]; }
};

class FileAcquisition : public mista::STAAcquisition {
public:
	// data holds transmits first_tx, first_tx + 1, ...
	FileAcquisition(std::vector<float> data, unsigned int nrx, unsigned int ns, unsigned int first_tx = 0)
			: data_(std::move(data)), nrx_(nrx), ns_(ns), first_(first_tx) {}
	void execute(unsigned int txElem, std::vector<float>& rx, unsigned int& numSamples) override
	{
		const size_t n = static_cast<size_t>(nrx_) * ns_;
		if (txElem < first_ || (txElem - first_ + 1) * n > data_.size())
			throw mista::InvalidParameterException("transmit out of range");
		const size_t k = txElem - first_;
		rx.assign(data_.begin() + k * n, data_.begin() + (k + 1) * n);
		numSamples = ns_;
	}
private:
	std::vector<float> data_;
	unsigned int nrx_, ns_, first_;
};

// WindowFunction::get (every window the reference names), rounded to float.
std::vector<float> window(const std::string& description, unsigned int n)
{
	const std::vector<double> w = mista_cfg::window_function(description, n);
	return std::vector<float>(w.begin(), w.end());
}

bool write_f64(const std::string& path, const std::vector<double>& v)
{
	std::ofstream f(path, std::ios::binary);
	f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(double)));
	return static_cast<bool>(f);
}

bool write_f32(const std::string& path, const std::vector<float>& v)
{
	std::ofstream f(path, std::ios::binary);
	f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(float)));
	return static_cast<bool>(f);
}

}  // namespace

int main(int argc, char** argv)
{
	std::string echoes_path, out = "sta", cf = "none", apod = "hanning", storage = "f32";
	std::string acq_dir, h5_out, write_acq;
	std::string sta_cfg, imag_cfg, cf_cfg, grid_cfg, dump_cfg;
	unsigned int N = 0, ns = 0, up = 1, n1 = 0, n2 = 0, base = 0;
	float pitch = 0, fs = 0, c = 1540.0f, fc = 0, cf_param = 1.0f, dead_zone = 0.0f, peak_offset = 0.0f;
	double xmin = 0, xmax = 0, zmin = 0, zmax = 0;
	for (int i = 1; i < argc; ++i) {
		const std::string a = argv[i];
		auto next = [&]() -> const char* {
			if (i + 1 >= argc) { std::cerr << "missing value for " << a << "\n"; std::exit(2); }
			return argv[++i];
		};
		if (a == "--echoes") echoes_path = next();
		else if (a == "--elements") N = std::atoi(next());
		else if (a == "--samples") ns = std::atoi(next());
		else if (a == "--up") up = std::atoi(next());
		else if (a == "--pitch") pitch = std::atof(next());
		else if (a == "--fs") fs = std::atof(next());
		else if (a == "--c") c = std::atof(next());
		else if (a == "--fc") fc = std::atof(next());
		else if (a == "--cf") cf = next();
		else if (a == "--cf-param") cf_param = std::atof(next());
		else if (a == "--apod") apod = next();
		else if (a == "--storage") storage = next();
		else if (a == "--out") out = next();
		else if (a == "--acq-dir") acq_dir = next();
		else if (a == "--base") base = std::atoi(next());
		else if (a == "--h5-out") h5_out = next();
		else if (a == "--write-acq") write_acq = next();
		else if (a == "--sta-config") sta_cfg = next();
		else if (a == "--imag-config") imag_cfg = next();
		else if (a == "--cf-config") cf_cfg = next();
		else if (a == "--grid-config") grid_cfg = next();
		else if (a == "--dump-config") dump_cfg = next();
		else if (a == "--grid") {
			n1 = std::atoi(next()); n2 = std::atoi(next());
			xmin = std::atof(next()); xmax = std::atof(next()); zmin = std::atof(next()); zmax = std::atof(next());
		} else { std::cerr << "unknown argument " << a << "\n"; return 2; }
	}
	// project configuration files (reference layout) override the flags
	unsigned int first_tx = 0, last_tx = ~0u;
	mista_cfg::GridXZ cfg_grid;
	try {
		if (!sta_cfg.empty()) {
			const mista_cfg::StaConfig sc = mista_cfg::sta_config(mista_cfg::ParamMap(sta_cfg));
			N = sc.num_elements;
			first_tx = sc.first_tx_elem;
			last_tx = sc.last_tx_elem;
			pitch = static_cast<float>(sc.pitch);
			fs = static_cast<float>(sc.sampling_frequency);
			c = static_cast<float>(sc.propagation_speed);
			fc = static_cast<float>(sc.center_frequency);
			dead_zone = static_cast<float>(sc.dead_zone_m);
			if (!grid_cfg.empty()) {
				cfg_grid = mista_cfg::image_grid(mista_cfg::ParamMap(grid_cfg),
						mista_cfg::nyquist_lambda(sc.propagation_speed, sc.max_frequency));
				n1 = cfg_grid.n1;
				n2 = cfg_grid.n2;
			}
		} else if (!grid_cfg.empty()) {
			std::cerr << "--grid-config needs --sta-config (propagation speed and max frequency)\n";
			return 2;
		}
		if (!imag_cfg.empty()) {
			const mista_cfg::ImagingConfig ic = mista_cfg::imaging_config(mista_cfg::ParamMap(imag_cfg));
			up = ic.upsampling_factor;
			peak_offset = static_cast<float>(ic.peak_offset);
			apod = ic.rx_apodization;
		}
		if (!cf_cfg.empty()) {
			const mista_cfg::CoherenceConfig cc = mista_cfg::coherence_config(mista_cfg::ParamMap(cf_cfg));
			cf = cc.kind == 1 ? "pcf" : cc.kind == 2 ? "scf" : "none";
			cf_param = static_cast<float>(cc.param);
		}
		if (!dump_cfg.empty()) {
			const std::vector<double> w = mista_cfg::window_function(apod, N);
			std::ofstream t(dump_cfg + "/config.txt");
			t << "num_elements = " << N << "\nfirst_tx_elem = " << first_tx << "\nlast_tx_elem = "
			  << (last_tx == ~0u ? N - 1 : last_tx) << "\npitch = " << pitch << "\nsampling_frequency = " << fs
			  << "\npropagation_speed = " << c << "\ncenter_frequency = " << fc << "\ndead_zone_m = " << dead_zone
			  << "\nupsampling_factor = " << up << "\npeak_offset = " << peak_offset << "\nrx_apodization = " << apod
			  << "\ncoherence = " << cf << "\ncoherence_param = " << cf_param << "\ngrid = " << n1 << " " << n2 << "\n";
			if (!t || !write_f64(dump_cfg + "/apod.f64", w) || !write_f64(dump_cfg + "/grid_x.f64", cfg_grid.x) ||
					!write_f64(dump_cfg + "/grid_z.f64", cfg_grid.z)) {
				std::cerr << "cannot write " << dump_cfg << "\n";
				return 1;
			}
			return 0;
		}
	} catch (const std::exception& e) {
		std::cerr << "error: " << e.what() << "\n";
		return 1;
	}
#ifndef MISTA_HAVE_HDF5
	if (!acq_dir.empty() || !h5_out.empty() || !write_acq.empty()) {
		std::cerr << "built without HDF5 support\n";
		return 2;
	}
#endif
	const bool convert = !write_acq.empty();
	if ((echoes_path.empty() && acq_dir.empty()) || N == 0 || (ns == 0 && acq_dir.empty()) ||
			(!convert && (n1 == 0 || n2 == 0))) {
		std::cerr << "usage: see the header of sta_cli.cpp\n";
		return 2;
	}
	try {
		std::vector<float> data;
		unsigned int ntx = 0;
		if (!echoes_path.empty()) {
			std::ifstream f(echoes_path, std::ios::binary | std::ios::ate);
			if (!f) throw mista::InvalidParameterException("cannot open " + echoes_path);
			const size_t bytes = static_cast<size_t>(f.tellg());
			const size_t per_tx = static_cast<size_t>(N) * ns * sizeof(float);
			if (bytes % per_tx != 0) throw mista::InvalidParameterException("echo file size is not a multiple of one transmit");
			data.resize(bytes / sizeof(float));
			f.seekg(0);
			f.read(reinterpret_cast<char*>(data.data()), static_cast<std::streamsize>(bytes));
			ntx = static_cast<unsigned int>(bytes / per_tx);
		}
#ifdef MISTA_HAVE_HDF5
		if (convert) {
			// Project::saveTxElemSignalsToHDF5 layout, one file per transmit
			for (unsigned int t = 0; t < ntx; ++t)
				mista_h5::write2d(mista_h5::tx_signals_path(write_acq, base, t), "signal",
						data.data() + static_cast<size_t>(t) * N * ns, N, ns);
			std::printf("wrote %u transmits to %s\n", ntx, write_acq.c_str());
			return 0;
		}
		if (!acq_dir.empty()) {
			// SavedSTAAcquisition::execute for transmits first_tx .. last_tx
			// (without a configuration: 0, 1, ... while files exist)
			for (unsigned int t = first_tx; t <= last_tx; ++t) {
				const std::string path = mista_h5::tx_signals_path(acq_dir, base, t);
				if (std::FILE* probe = std::fopen(path.c_str(), "rb")) std::fclose(probe);
				else if (last_tx != ~0u) throw mista::InvalidParameterException("missing acquisition file " + path);
				else break;
				size_t r = 0, c = 0;
				std::vector<float> sig = mista_h5::read2d(path, "signal", r, c);
				if (r != N)
					throw mista::InvalidValueException("Invalid number of rx elements (" + std::to_string(r) + ") in file " +
							path + ", should be " + std::to_string(N) + ".");
				if (t == first_tx) ns = static_cast<unsigned int>(c);
				else if (c != ns) throw mista::InvalidValueException("The number of samples changed between transmits.");
				data.insert(data.end(), sig.begin(), sig.end());
				ntx = t - first_tx + 1;
			}
			if (ntx == 0) throw mista::InvalidParameterException("no acquisition files in " + acq_dir);
		}
#endif

		if (last_tx != ~0u && ntx != last_tx - first_tx + 1)
			throw mista::InvalidParameterException("the echoes hold " + std::to_string(ntx) + " transmits, the configuration " +
					std::to_string(last_tx - first_tx + 1));
		mista::STAConfiguration config;
		config.numElements = N;
		config.firstTxElem = first_tx;
		config.lastTxElem = first_tx + ntx - 1;
		config.pitch = pitch;
		config.samplingFrequency = fs;
		config.propagationSpeed = c;
		config.centerFrequency = fc;
		config.deadZoneM = dead_zone;
		mista::CoherenceFactor coherence;
		coherence.kind = cf == "pcf" ? MISTA_CF_PCF : cf == "scf" ? MISTA_CF_SCF : MISTA_CF_NONE;
		coherence.param = cf_param;
		const int st = storage == "f16" ? MISTA_STORAGE_F16 : storage == "bf16" ? MISTA_STORAGE_BF16 : MISTA_STORAGE_F32;

		// ImageGrid (explicit point counts, Util.h:326-345 spacing)
		Grid grid;
		grid.n1_ = n1;
		grid.n2_ = n2;
		grid.pts.resize(static_cast<size_t>(n1) * n2);
		for (unsigned int i = 0; i < n1; ++i)
			for (unsigned int j = 0; j < n2; ++j) {
				if (!cfg_grid.x.empty()) {
					grid(i, j).x = static_cast<float>(cfg_grid.x[static_cast<size_t>(i) * n2 + j]);
					grid(i, j).z = static_cast<float>(cfg_grid.z[static_cast<size_t>(i) * n2 + j]);
					continue;
				}
				grid(i, j).x = static_cast<float>(n1 > 1 ? xmin + i * ((xmax - xmin) / (n1 - 1)) : xmin);
				grid(i, j).z = static_cast<float>(n2 > 1 ? zmin + j * ((zmax - zmin) / (n2 - 1)) : zmin);
			}

		FileAcquisition acq(std::move(data), N, ns, first_tx);
		mista::HipSTAProcessor<> proc(config, acq, up, coherence, peak_offset, window(apod, N), st);
		proc.prepare(0);
		const auto t0 = std::chrono::steady_clock::now();
		proc.process(grid);
		const auto t1 = std::chrono::steady_clock::now();

		std::vector<float> value(grid.pts.size()), factor(grid.pts.size()), cfimg(grid.pts.size());
		for (size_t i = 0; i < grid.pts.size(); ++i) {
			value[i] = grid.pts[i].value;
			factor[i] = grid.pts[i].factor;
			cfimg[i] = grid.pts[i].value * grid.pts[i].factor;   // Util::applyFactorToValue
		}
		if (!write_f32(out + "_value.f32", value) || !write_f32(out + "_factor.f32", factor) ||
				(coherence.enabled() && !write_f32(out + "_cf.f32", cfimg))) {
			std::cerr << "cannot write outputs\n";
			return 1;
		}
#ifdef MISTA_HAVE_HDF5
		if (!h5_out.empty()) {
			// Project::saveImageToHDF5 / saveXYZToHDF5 / saveFactorToHDF5
			std::vector<float> xs(grid.pts.size()), zs(grid.pts.size());
			for (size_t i = 0; i < grid.pts.size(); ++i) {
				xs[i] = grid.pts[i].x;
				zs[i] = grid.pts[i].z;
			}
			mista_h5::write2d(h5_out + "/image_value.h5", "value", value.data(), n1, n2);
			mista_h5::write2d(h5_out + "/image_x.h5", "x", xs.data(), n1, n2);
			mista_h5::write2d(h5_out + "/image_z.h5", "z", zs.data(), n1, n2);
			if (coherence.enabled()) {
				mista_h5::write2d(h5_out + "/image_factor.h5", "factor", factor.data(), n1, n2);
				mista_h5::write2d(h5_out + "/image_cf.h5", "cf", cfimg.data(), n1, n2);
			}
		}
#endif
		std::printf("frame: %u tx x %u rx x %u samples, grid %u x %u, %.3f ms (host path incl. transfers)\n", ntx, N, ns, n1,
				n2, std::chrono::duration<double, std::milli>(t1 - t0).count());
	} catch (const std::exception& e) {
		std::cerr << "error: " << e.what() << "\n";
		return 1;
	}
	return 0;
}

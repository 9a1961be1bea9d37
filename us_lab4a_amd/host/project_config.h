// The reference's key=value project configuration files, read the way its
// STAMethod reads them, so that a project directory drives the CLI:
//
//   config-sta.txt               STAConfiguration (STAConfiguration.h:59-74)
//   config-imaging.txt           peak_offset, upsampling_factor, rx_apodization
//                                (STAMethod.cpp:176,195-196)
//   config-coherence_factor.txt  analytic_signal_coherence_factor_method
//                                (CoherenceFactor.h:141-154)
//   config-grid.txt              image_grid_type rectangular | sectorial
//                                (ImageGrid.h:120-203)
//
// File format (KeyValueFileReader.cpp:30-73): one "key = value" per line,
// surrounding blanks trimmed, blank lines and lines starting with '#'
// skipped; a line without '=', with an empty key or value, a key holding a
// blank, or a repeated key is an error naming the line and file.
#ifndef MISTA_PROJECT_CONFIG_H
#define MISTA_PROJECT_CONFIG_H

#include <cmath>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mista_cfg {

struct ConfigError : std::runtime_error {
	explicit ConfigError(const std::string& m) : std::runtime_error(m) {}
};
// KeyValueFileReader / ParameterMap: InvalidValueException
struct InvalidValue : ConfigError {
	explicit InvalidValue(const std::string& m) : ConfigError(m) {}
};
// ImageGrid, WindowFunction, Util::fillSequence*: InvalidParameterException
struct InvalidParameter : ConfigError {
	explicit InvalidParameter(const std::string& m) : ConfigError(m) {}
};
// KeyValueFileReader: IOException
struct IOError : ConfigError {
	explicit IOError(const std::string& m) : ConfigError(m) {}
};

inline std::string trim(const std::string& s)
{
	const char* ws = " \t\r\n\f\v";
	const size_t b = s.find_first_not_of(ws);
	if (b == std::string::npos) return std::string();
	return s.substr(b, s.find_last_not_of(ws) - b + 1);
}

class ParamMap {
public:
	explicit ParamMap(const std::string& path) : path_(path)
	{
		std::ifstream in(path);
		if (!in) throw IOError("Could not open the file " + path + ".");
		std::string line;
		for (unsigned n = 1; std::getline(in, line); ++n) {
			const std::string t = trim(line);
			if (t.empty() || t[0] == '#') continue;
			const size_t eq = t.find('=');
			const std::string where = " at line " + std::to_string(n) + " of file " + path + ".";
			if (eq == std::string::npos) throw InvalidValue("Missing '=' separator" + where);
			if (eq == 0) throw InvalidValue("Missing key" + where);
			if (eq == t.size() - 1) throw InvalidValue("Missing value" + where);
			const std::string key = trim(t.substr(0, eq)), val = trim(t.substr(eq + 1));
			if (key.find_first_of(" \t") != std::string::npos) throw InvalidValue("Key has space" + where);
			if (map_.count(key)) throw InvalidValue("Duplicate key '" + key + "'" + where);
			map_[key] = val;
		}
	}

	bool has(const std::string& key) const { return map_.count(key) != 0; }

	std::string str(const std::string& key) const
	{
		const auto it = map_.find(key);
		if (it == map_.end()) throw InvalidValue("Key '" + key + "' not found in file " + path_ + ".");
		return it->second;
	}

	// ParameterMap::value<T>(key, min, max): the whole value must convert,
	// then min <= v <= max.
	double real(const std::string& key, double lo, double hi) const
	{
		const std::string s = str(key);
		double v = 0.0;
		size_t used = 0;
		try {
			v = std::stod(s, &used);
		} catch (const std::exception&) {
			used = 0;
		}
		if (used == 0 || used != s.size())
			throw InvalidValue("Invalid value for key '" + key + "' in file " + path_ + ": " + s);
		check(key, v, lo, hi);
		return v;
	}

	unsigned long integer(const std::string& key, unsigned long lo, unsigned long hi) const
	{
		const std::string s = str(key);
		if (s.empty() || s.find_first_not_of("0123456789") != std::string::npos)
			throw InvalidValue("Invalid value for key '" + key + "' in file " + path_ + ": " + s);
		const unsigned long v = std::stoul(s);
		check(key, static_cast<double>(v), static_cast<double>(lo), static_cast<double>(hi));
		return v;
	}

	const std::string& path() const { return path_; }

private:
	void check(const std::string& key, double v, double lo, double hi) const
	{
		std::ostringstream m;
		if (v > hi) m << "The value for key '" << key << "' must be <= " << hi << " in file " << path_ << '.';
		else if (v < lo) m << "The value for key '" << key << "' must be >= " << lo << " in file " << path_ << '.';
		else return;
		throw InvalidValue(m.str());
	}

	std::string path_;
	std::map<std::string, std::string> map_;
};

// STAConfiguration.h:59-74, the same keys and ranges.
struct StaConfig {
	unsigned num_elements_mux, num_elements, first_tx_elem, last_tx_elem, num_pulses;
	double pitch, center_frequency, max_frequency, acquisition_time, min_gain, max_gain;
	double propagation_speed, acquisition_delay, sampling_frequency, dead_zone_m, value_scale;
};

inline StaConfig sta_config(const ParamMap& pm)
{
	StaConfig c{};
	c.num_elements_mux = pm.integer("num_elements_mux", 1, 1024);
	c.num_elements = pm.integer("num_elements", 1, c.num_elements_mux);
	c.first_tx_elem = pm.integer("first_tx_elem", 0, c.num_elements - 1);
	c.last_tx_elem = pm.integer("last_tx_elem", c.first_tx_elem, c.num_elements - 1);
	c.num_pulses = pm.integer("num_pulses", 1, 100);
	c.pitch = pm.real("pitch", 1.0e-6, 1000.0);
	c.center_frequency = pm.real("center_frequency", 100.0, 100.0e6);
	c.max_frequency = pm.real("max_frequency", 100.0, 100.0e6);
	c.acquisition_time = pm.real("acquisition_time", 1.0e-6, 1.0);
	c.min_gain = pm.real("min_gain", -2000.0, 2000.0);
	c.max_gain = pm.real("max_gain", c.min_gain, 2000.0);
	c.propagation_speed = pm.real("propagation_speed_1", 100.0, 10000.0);
	c.acquisition_delay = pm.real("acquisition_delay", 0.0, 1.0);
	c.sampling_frequency = pm.real("sampling_frequency", 100.0, 200.0e6);
	c.dead_zone_m = pm.real("dead_zone_m", 0.0, 50.0e-3);
	c.value_scale = pm.real("value_scale", 0.0, 1.0e30);
	return c;
}

// Util::nyquistLambda: wavelength at the Nyquist rate 2 * max_frequency.
inline double nyquist_lambda(double c, double max_frequency) { return c / (2.0 * max_frequency); }

// Util::fillSequenceFromStartToEndWithMaximumStep (Util.h:266-292): the
// fewest evenly spaced points from start to end (both included) whose
// spacing does not exceed |step|.
inline std::vector<double> fill_max_step(double start, double end, double step)
{
	if (step == 0.0) throw InvalidParameter("The step must not be zero.");
	if (start < end && step < 0.0) throw InvalidParameter("The step must be positive.");
	if (start > end && step > 0.0) throw InvalidParameter("The step must be negative.");
	const double steps = std::ceil((end - start) / step);
	if (steps > static_cast<double>(std::numeric_limits<int>::max() - 2)) throw InvalidParameter("The step is too small.");
	const int n = 1 + static_cast<int>(steps);
	if (n == 1) throw InvalidParameter("Sequence with only one element.");
	const double d = (end - start) / (n - 1);
	std::vector<double> v(n);
	for (int i = 0; i < n; ++i) v[i] = start + i * d;
	return v;
}

// ImageGrid<TFloat>::get for 2-D (x, z) points: grid [n1][n2], z fastest.
struct GridXZ {
	unsigned n1 = 0, n2 = 0;
	std::vector<double> x, z;   // [n1 * n2]
};

inline GridXZ image_grid(const ParamMap& pm, double lambda)
{
	constexpr double kMinLambda = 1.0e-6;   // ImageGrid.h minLambda
	if (lambda < kMinLambda) throw InvalidParameter("Lambda is too small: " + std::to_string(lambda) + ".");
	const std::string type = pm.str("image_grid_type");
	GridXZ g;
	if (type == "rectangular") {
		// ImageGrid.h:120-146 (points without a y member: the x-z form)
		const double min_x = pm.real("rectangular_min_x", -10000.0, 10000.0);
		const double max_x = pm.real("rectangular_max_x", min_x, 10000.0);
		const double min_z = pm.real("rectangular_min_z", -10000.0, 10000.0);
		const double max_z = pm.real("rectangular_max_z", min_z, 10000.0);
		const double div = pm.real("rectangular_step_div", 1.0e-2, 1000.0);
		const double ox = pm.real("rectangular_origin_x", -10000.0, 10000.0);
		const double oz = pm.real("rectangular_origin_z", -10000.0, 10000.0);
		const auto xs = fill_max_step(min_x, max_x, lambda / div);
		const auto zs = fill_max_step(min_z, max_z, lambda / div);
		g.n1 = static_cast<unsigned>(xs.size());
		g.n2 = static_cast<unsigned>(zs.size());
		g.x.resize(static_cast<size_t>(g.n1) * g.n2);
		g.z.resize(g.x.size());
		for (unsigned i = 0; i < g.n1; ++i)
			for (unsigned j = 0; j < g.n2; ++j) {
				g.x[static_cast<size_t>(i) * g.n2 + j] = xs[i] + ox;
				g.z[static_cast<size_t>(i) * g.n2 + j] = zs[j] + oz;
			}
	} else if (type == "sectorial") {
		// ImageGrid.h:148-186: rows are angles, columns radii
		const double r0 = pm.real("sectorial_min_radius", 1.0e-4, 10000.0);
		const double r1 = pm.real("sectorial_max_radius", r0 + 1.0e-6, 10000.0);
		const double div = pm.real("sectorial_radius_step_div", 1.0e-2, 1000.0);
		const double a0 = pm.real("sectorial_min_angle", -88.0, 88.0);
		const double a1 = pm.real("sectorial_max_angle", a0 + 1.0e-3, 88.0);
		const double da = pm.real("sectorial_angle_step", 1.0e-6, 10.0);
		const double ox = pm.real("sectorial_origin_x", -10000.0, 10000.0);
		const double oz = pm.real("sectorial_origin_z", -10000.0, 10000.0);
		const double rad = M_PI / 180.0;
		const auto rs = fill_max_step(r0, r1, lambda / div);
		const auto as = fill_max_step(a0 * rad, a1 * rad, da * rad);
		g.n1 = static_cast<unsigned>(as.size());
		g.n2 = static_cast<unsigned>(rs.size());
		g.x.resize(static_cast<size_t>(g.n1) * g.n2);
		g.z.resize(g.x.size());
		for (unsigned i = 0; i < g.n1; ++i)
			for (unsigned j = 0; j < g.n2; ++j) {
				g.x[static_cast<size_t>(i) * g.n2 + j] = rs[j] * std::sin(as[i]) + ox;
				g.z[static_cast<size_t>(i) * g.n2 + j] = rs[j] * std::cos(as[i]) + oz;
			}
	} else {
		throw InvalidParameter("Invalid grid type: " + type + ".");
	}
	return g;
}

// WindowFunction::get (WindowFunction.h:237-289).  Ramps: "fade in" over n
// points takes i / n (the zero end included); the "2" forms take (i + 1) /
// (n + 1), leaving the zero ends out; n may be fractional (half widths).
namespace detail {
inline void ramp_up(double n, bool skip_zero, double* w)
{
	for (unsigned i = 0; i < n; ++i) w[i] = skip_zero ? (i + 1.0) / (n + 1.0) : i / n;
}
inline void ramp_down(double n, bool skip_zero, double* w)
{
	const unsigned top = static_cast<unsigned>(std::ceil(n));
	for (unsigned i = 0; i < n; ++i) w[top - i - 1] = skip_zero ? (i + 1.0) / (n + 1.0) : i / n;
}
inline void triangle(unsigned n, bool skip_zero, std::vector<double>& w)
{
	const double h = 0.5 * (n - 1.0);
	ramp_up(h, skip_zero, w.data());
	if (n % 2 == 0) {
		ramp_down(h, skip_zero, w.data() + n / 2);
	} else {
		w[(n - 1) / 2] = 1.0;
		ramp_down(h, skip_zero, w.data() + (n - 1) / 2 + 1);
	}
}
inline void cosine_sum(unsigned n, bool skip_zero, double a0, double a1, double a2, std::vector<double>& w)
{
	for (unsigned i = 0; i < n; ++i) {
		const double t = skip_zero ? (i + 1.0) / (n + 1.0) : i / (n - 1.0);
		w[i] = a0 - a1 * std::cos(2.0 * M_PI * t) + a2 * std::cos(4.0 * M_PI * t);
	}
}
}  // namespace detail

inline std::vector<double> window_function(const std::string& description, unsigned n)
{
	std::istringstream in(description);
	std::string name;
	in >> name;
	if (!in) throw InvalidValue("Missing window function name.");
	auto need2 = [](unsigned m) {
		if (m < 2u) throw InvalidParameter("The window function size must be >= 2.");
	};
	std::vector<double> w;
	if (name == "trapezoidal" || name == "trapezoidal2") {
		unsigned a = 0, b = 0, c = 0;
		if (!(in >> a)) throw InvalidParameter("Missing fade in size for the trapezoidal window function.");
		if (!(in >> b)) throw InvalidParameter("Missing top size for the trapezoidal window function.");
		if (!(in >> c)) throw InvalidParameter("Missing fade out size for the trapezoidal window function.");
		if (n != a + b + c) throw InvalidParameter("Invalid total size for the trapezoidal window function.");
		need2(n);
		const bool skip = name == "trapezoidal2";
		w.assign(n, 1.0);
		detail::ramp_up(a, skip, w.data());
		detail::ramp_down(c, skip, w.data() + a + b);
		return w;
	}
	need2(n);
	w.assign(n, 1.0);
	if (name == "rectangular") {
	} else if (name == "triangular" || name == "triangular2") {
		detail::triangle(n, name == "triangular2", w);
	} else if (name == "hanning" || name == "hanning2") {
		detail::cosine_sum(n, name == "hanning2", 0.5, 0.5, 0.0, w);
	} else if (name == "hamming") {
		detail::cosine_sum(n, false, 0.54, 0.46, 0.0, w);
	} else if (name == "blackman" || name == "blackman2") {
		detail::cosine_sum(n, name == "blackman2", 0.42, 0.5, 0.08, w);
	} else {
		throw InvalidParameter("Invalid window function name: " + name + ".");
	}
	return w;
}

// AnalyticSignalCoherenceFactor::get (CoherenceFactor.h:141-154), with the
// sign coherence factor on the analytic path as an extension
// (CoherenceFactor.h:67-75 keys).  kind: 0 none, 1 PCF, 2 SCF.
struct CoherenceConfig {
	int kind = 0;
	double param = 1.0;
};

inline CoherenceConfig coherence_config(const ParamMap& pm)
{
	CoherenceConfig c;
	const std::string key = pm.has("analytic_signal_coherence_factor_method") ? "analytic_signal_coherence_factor_method"
			: "coherence_factor_method";
	const std::string m = pm.str(key);
	if (m == "none") {
	} else if (m == "phase_coherence_factor") {
		c.kind = 1;
		c.param = pm.real("phase_coherence_factor_gamma", 0.0, 100.0);
	} else if (m == "sign_coherence_factor") {
		c.kind = 2;
		c.param = pm.real("sign_coherence_factor_p", 0.0, 100.0);
	} else {
		throw InvalidParameter("Invalid coherence factor method: " + m + ".");
	}
	return c;
}

// The imaging file's keys for the vectorial processor (STAMethod.cpp:176,
// 195-196).
struct ImagingConfig {
	double peak_offset = 0.0;
	unsigned upsampling_factor = 1;
	std::string rx_apodization;
};

inline ImagingConfig imaging_config(const ParamMap& pm)
{
	ImagingConfig c;
	c.peak_offset = pm.real("peak_offset", 0.0, 50.0);
	c.upsampling_factor = pm.integer("upsampling_factor", 1, 128);
	c.rx_apodization = pm.str("rx_apodization");
	return c;
}

}  // namespace mista_cfg

#endif

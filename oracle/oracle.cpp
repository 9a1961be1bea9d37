/*
 * CPU oracle for the STA/TFM delay-and-sum path — TEST INFRASTRUCTURE ONLY.
 *
 * Restates the reference CPU processor (VectorialSTAProcessor, the TBB path)
 * and the DSP routines on its path.  FP32 instantiations keep the reference's
 * arithmetic form and loop order (column-parallel; tx outer, rx inner; float
 * truncation indexing; lerp (1-k)a + k b; rx apodization per sample; two-pass
 * population sigma; 1/(Ntx*Nrx) applied after the magnitude).  FP64
 * instantiations are the accuracy oracle for the dB error of every precision.
 *
 * Differences from the reference that parity must absorb (tolerance-based):
 *  - FFTW is absent.  Convolution is a direct sum and the FFT is this file's
 *    own mixed-radix transform (long double internally for the FP64 variant).
 *    Both are equal to the reference's FFT-based forms mathematically; their
 *    rounding differs.
 *  - TBB is replaced by OpenMP over image columns (same partition axis).
 *
 * How each function is pinned (tests/test_oracle.py):
 *  kaiser_beta / kaiser_size   reference fixtures kaiser_beta, kaiser_size-*
 *  bessel_i0                   reference fixture bessel_i0
 *  fft / ifft                  reference fixtures fft_* (incl. prime 3571)
 *  convolve                    reference fixtures filter_y, filter_y2
 *  interpolate                 reference fixtures interp_2x/8x/64x
 *  analytic_signal             reference fixtures hilbert_yr/yi/ya
 *  next_fast_even              brute-force 2-3-5 smoothness scan
 *  std_dev / mean              TestMethod known answer sigma{-1,2,4,7}
 *  scf                         TestMethod known answer 3.270805724762158e-3
 *  pcf                         closed forms (equal phases -> 1, two-phase split)
 *  remove_dc                   closed form (mean of tail becomes zero)
 *  das (2-D and 3-D)           brute-force FP64 DAS on tiny inputs, single
 *                              point scatterer peak location, constant-signal
 *                              closed form, i+1 < L exclusion boundary
 *  window functions            closed forms (hanning endpoints/symmetry)
 * The DAS/PCF/envelope outputs of the reference processor are not covered by
 * any reference test: they are pinned here by the closed forms above.
 */
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_error;

constexpr double kPi = 3.14159265358979323846; /* Util.h:40 pi = M_PI */

/* ---------------- Kaiser window: KaiserWindow.h:45-93 ---------------- */

template<typename T>
T kaiser_beta(T tol_db)
{
	T beta;
	if (tol_db > 50) {
		beta = 0.1102 * (tol_db - 8.7);
	} else if (tol_db > 21) {
		const double k = tol_db - 21.0;
		beta = 0.5842 * std::pow(k, 0.4) + 0.07886 * k;
	} else {
		beta = 0.0;
	}
	return beta;
}

template<typename T>
unsigned kaiser_size(T tol_db, T tw)
{
	const double size = std::ceil((tol_db - 7.95) / (2.285 * tw * kPi) + 1.0);
	if (size < 2.0 || size >= 4294967295.0) throw std::invalid_argument("Invalid Kaiser window size");
	return static_cast<unsigned>(size);
}

/* Math.cpp:27-37 wraps std::cyl_bessel_i / std::cyl_bessel_if. */
inline double bessel_i0(double x) { return std::cyl_bessel_i(0.0, x); }
inline float bessel_i0(float x) { return std::cyl_bessel_if(0.0f, x); }

template<typename T>
void kaiser_window(unsigned size, T beta, std::vector<T>& w)
{
	if (size < 2) throw std::invalid_argument("Invalid Kaiser window size");
	w.resize(size);
	const T c1 = 2 / static_cast<T>(size - 1U);
	const T c2 = 1 / bessel_i0(beta);
	for (unsigned n = 0; n < size; ++n) {
		const T k = c1 * n - 1;
		T x = 1 - k * k;
		if (x < 0.0) x = 0.0;
		w[n] = c2 * bessel_i0(static_cast<T>(beta * std::sqrt(x)));
	}
}

/* ---------------- FFTUtil.h:100-135 nextFastEvenSize ---------------- */

unsigned next_fast_even(unsigned n)
{
	if (n == 0) throw std::invalid_argument("n must be > 0");
	if (n & 1) ++n;
	for (;;) {
		unsigned m = n;
		while ((m & 1) == 0) m >>= 1;
		while (m % 3 == 0) m /= 3;
		while (m % 5 == 0) m /= 5;
		if (m == 1) return n;
		n += 2;
	}
}

/* Util.h:326-345 fillSequenceFromStartToEndWithSize: step in double. */
template<typename T>
void fill_start_end_size(std::vector<T>& v, T a, T b, unsigned size)
{
	if (size < 1) throw std::invalid_argument("size must be >= 1");
	v.resize(size);
	if (size == 1) { v[0] = a; return; }
	const double step = static_cast<double>(b - a) / (size - 1U);
	v[0] = a;
	for (unsigned i = 1; i < size; ++i) v[i] = a + i * step;
}

/* boost::math::sinc_pi(x) = sin(x)/x with the limit 1 at 0. */
inline double sinc_pi(double x)
{
	if (std::abs(x) < 1e-5) return 1.0 - x * x / 6.0;
	return std::sin(x) / x;
}

/* Interpolator::prepare, Interpolator.h:64-111 (Kaiser tolerance 1e-4). */
template<typename T>
void upsampling_filter(unsigned up, T half_tw, std::vector<T>& h)
{
	if (up < 2 || up > 64) throw std::invalid_argument("Invalid upsampling factor");
	if (half_tw > 1) throw std::invalid_argument("Invalid transition width");
	const T tol_db = -20.0 * std::log10(1.0e-4);
	const T beta = kaiser_beta<T>(tol_db);
	const T final_tw = (half_tw * 2) / up;
	const unsigned wsize = kaiser_size<T>(tol_db, final_tw);
	const T two_up = 2 * up;
	const unsigned fsize = static_cast<unsigned>(std::ceil((wsize - 1) / two_up) * two_up) + 1U;
	std::vector<T> window;
	kaiser_window<T>(fsize, beta, window);
	const T num_periods = (fsize - 1U) / two_up;
	std::vector<T> x;
	fill_start_end_size<T>(x, -num_periods, num_periods, fsize);
	h.resize(fsize);
	for (unsigned i = 0; i < fsize; ++i) h[i] = sinc_pi(kPi * x[i]);
	for (unsigned i = 0; i < fsize; ++i) h[i] *= window[i]; /* Util::multiply */
}

/* Interpolator::interpolate, Interpolator.h:114-138: zero-stuff by `up`,
 * convolve with h (FFTWFilter::filter = full linear convolution), keep
 * n*up samples starting at (taps-1)/2.  The zero-stuffed input is only
 * non-zero at multiples of `up`, so the sum runs over those taps. */
template<typename T>
void interpolate(const T* x, size_t n, unsigned up, const std::vector<T>& h, T* out)
{
	const long taps = static_cast<long>(h.size());
	const long offset = (taps - 1) / 2;
	const long total = static_cast<long>(n) * up;
	for (long i = 0; i < total; ++i) {
		const long m = i + offset;              /* index into the full convolution */
		/* y[m] = sum_q x[q] h[m - up*q], 0 <= m - up*q < taps */
		long qmax = m / up;
		if (qmax > static_cast<long>(n) - 1) qmax = static_cast<long>(n) - 1;
		long qmin = (m - taps + 1 + up - 1) / up;
		if (m - taps + 1 <= 0) qmin = 0;
		T acc = 0;
		for (long q = qmin; q <= qmax; ++q) acc += x[q] * h[m - up * q];
		out[i] = acc;
	}
}

template<typename T>
void convolve(const T* x, size_t nx, const T* b, size_t nb, T* y)
{
	const size_t ny = nx + nb - 1;
	for (size_t m = 0; m < ny; ++m) {
		const size_t jmin = (m >= nx) ? m - nx + 1 : 0;
		const size_t jmax = std::min(m, nb - 1);
		T acc = 0;
		for (size_t j = jmin; j <= jmax; ++j) acc += b[j] * x[m - j];
		y[m] = acc;
	}
}

/* Util.h:1009-1018 removeDC with a begin offset: mean of [begin, n)
 * subtracted from all samples. */
template<typename T>
void remove_dc(T* d, size_t n, size_t begin)
{
	if (begin >= n) throw std::invalid_argument("Invalid beginOffset");
	const T sum = std::accumulate(d + begin, d + n, T(0));
	const T sub = -sum / (n - begin);
	for (size_t i = 0; i < n; ++i) d[i] += sub;
}

/* ---------------- FFT (own mixed-radix, replaces FFTW) ---------------- */

template<typename W>
class Fft {
public:
	using C = std::complex<W>;
	explicit Fft(unsigned n) : n_(n)
	{
		unsigned m = n;
		for (unsigned p : {4u, 2u, 3u, 5u}) {
			while (m % p == 0) { fac_.push_back(p); m /= p; }
		}
		for (unsigned p = 7; m > 1; p += 2) {
			while (m % p == 0) { fac_.push_back(p); m /= p; }
		}
		tw_.resize(n);
		for (unsigned k = 0; k < n; ++k) {
			const long double a = -2.0L * 3.14159265358979323846264338327950288L * k / n;
			tw_[k] = C(static_cast<W>(std::cos(a)), static_cast<W>(std::sin(a)));
		}
		unsigned pmax = 1;
		for (unsigned p : fac_) pmax = std::max(pmax, p);
		tmp_.resize(pmax);
	}
	/* out[k] = sum_j in[j] exp(-2 pi i jk/n) */
	void forward(const C* in, C* out) { rec(in, out, n_, 1, 0); }
	/* out[k] = sum_j in[j] exp(+2 pi i jk/n) (unnormalised) */
	void inverse(const C* in, C* out)
	{
		std::vector<C> c(in, in + n_);
		for (auto& v : c) v = std::conj(v);
		rec(c.data(), out, n_, 1, 0);
		for (unsigned k = 0; k < n_; ++k) out[k] = std::conj(out[k]);
	}
private:
	void rec(const C* in, C* out, unsigned len, unsigned stride, unsigned fi)
	{
		if (len == 1) { out[0] = in[0]; return; }
		const unsigned p = fac_[fi];
		const unsigned m = len / p;
		for (unsigned r = 0; r < p; ++r) rec(in + r * stride, out + r * m, m, stride * p, fi + 1);
		const unsigned step = n_ / len;       /* W_len^j = tw[j*step] */
		const unsigned pstep = n_ / p;        /* W_p^j  = tw[j*pstep] */
		C* t = tmp_.data();
		for (unsigned k = 0; k < m; ++k) {
			for (unsigned r = 0; r < p; ++r) t[r] = out[r * m + k] * tw_[r * k * step];
			if (p == 2) {
				out[k] = t[0] + t[1];
				out[m + k] = t[0] - t[1];
			} else if (p == 4) {
				const C a = t[0] + t[2], b = t[0] - t[2];
				const C c = t[1] + t[3], d = t[1] - t[3];
				const C jd(d.imag(), -d.real()); /* -i*d */
				out[k] = a + c;
				out[m + k] = b + jd;
				out[2 * m + k] = a - c;
				out[3 * m + k] = b - jd;
			} else {
				for (unsigned q = 0; q < p; ++q) {
					C s = t[0];
					for (unsigned r = 1; r < p; ++r) s += t[r] * tw_[((r * q) % p) * pstep];
					out[q * m + k] = s;
				}
			}
		}
	}
	unsigned n_;
	std::vector<unsigned> fac_;
	std::vector<C> tw_;
	std::vector<C> tmp_;
};

template<typename T> struct Work { using type = T; };
template<> struct Work<double> { using type = long double; };

/* HilbertEnvelope::getAnalyticSignal, HilbertEnvelope.h:142-251:
 * Nf = max(nextFastEven(n + 1024), 2048); r2c FFT of the zero-padded input;
 * DC and Nyquist bins x0.5, negative bins zeroed; complex IFFT; x2/Nf; first n. */
template<typename T>
class Analytic {
public:
	using W = typename Work<T>::type;
	using C = std::complex<W>;
	explicit Analytic(unsigned n)
		: n_(n), nf_(std::max(next_fast_even(n + 1024u), 2048u)), fft_(nf_), a_(nf_), b_(nf_) {}
	void run(const T* x, T* out /* [n][2] */)
	{
		for (unsigned i = 0; i < n_; ++i) a_[i] = C(static_cast<W>(x[i]), 0);
		for (unsigned i = n_; i < nf_; ++i) a_[i] = C(0, 0);
		fft_.forward(a_.data(), b_.data());
		const unsigned half = nf_ / 2 + 1;
		b_[0] = C(b_[0].real() * W(0.5), b_[0].imag());
		b_[half - 1] = C(b_[half - 1].real() * W(0.5), b_[half - 1].imag());
		for (unsigned i = half; i < nf_; ++i) b_[i] = C(0, 0);
		fft_.inverse(b_.data(), a_.data());
		const W coef = W(2) / static_cast<W>(nf_);
		for (unsigned i = 0; i < n_; ++i) {
			out[2 * i] = static_cast<T>(a_[i].real() * coef);
			out[2 * i + 1] = static_cast<T>(a_[i].imag() * coef);
		}
	}
	unsigned fft_size() const { return nf_; }
private:
	unsigned n_, nf_;
	Fft<W> fft_;
	std::vector<C> a_, b_;
};

/* ---------------- Statistics.h:34-58 ---------------- */

template<typename T>
T mean(const T* d, size_t n)
{
	T sum = 0.0;
	for (size_t i = 0; i < n; ++i) sum += d[i];
	return sum / n;
}

template<typename T>
T std_dev(const T* d, size_t n)
{
	T sum = 0.0;
	const T m = mean(d, n);
	for (size_t i = 0; i < n; ++i) {
		const T e = d[i] - m;
		sum += e * e;
	}
	return std::sqrt(sum / n);
}

/* ---------------- CoherenceFactor.h ---------------- */

/* PhaseCoherenceFactor::calculate, CoherenceFactor.h:270-288.  sigma0 =
 * pi/sqrt(3), factor = gamma/sigma0 (CoherenceFactor.h:243-246). */
template<typename T>
class Pcf {
public:
	explicit Pcf(T gamma) : gamma_(gamma), sigma0_(kPi / std::sqrt(3.0)), factor_(gamma_ / sigma0_) {}
	T operator()(const std::complex<T>* d, unsigned n)
	{
		phi_.resize(n);
		aux_.resize(n);
		for (unsigned i = 0; i < n; ++i) {
			phi_[i] = std::atan2(d[i].imag(), d[i].real());
			aux_[i] = phi_[i] + ((phi_[i] < 0) ? T(kPi) : -T(kPi));
		}
		const T sf = std::min(std_dev(phi_.data(), n), std_dev(aux_.data(), n));
		return std::max<T>(0, 1 - factor_ * sf);
	}
private:
	const T gamma_, sigma0_, factor_;
	std::vector<T> phi_, aux_;
};

/* SignCoherenceFactor::calculate, CoherenceFactor.h:198-215 (0 counts +). */
template<typename T>
T scf(const T* d, unsigned n, T p)
{
	int sign_sum = 0;
	for (unsigned i = 0; i < n; ++i) sign_sum += (d[i] >= 0) ? 1 : -1;
	const T k = static_cast<T>(sign_sum) / n;
	const T m = 1 - std::sqrt(1 - k * k);
	return std::pow(std::abs(m), p);
}

/* ---------------- STA processor ---------------- */

template<typename T>
int num_threads_or_default(int nthreads)
{
#ifdef _OPENMP
	return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
	(void)nthreads;
	return 1;
#endif
}

/* Deadzone in upsampled samples, VectorialSTAProcessor.h:138. */
unsigned dead_zone_samples(double up, double fs, double dz_m, double c, bool f32)
{
	if (f32) {
		const float fsu = static_cast<float>(up) * static_cast<float>(fs);
		return static_cast<unsigned>(fsu * 2.0 * static_cast<float>(dz_m) / static_cast<float>(c));
	}
	return static_cast<unsigned>((up * fs) * 2.0 * dz_m / c);
}

/* VectorialSTAProcessor.h:169-231 for one frame.  echoes [ntx][nrx][ns]. */
template<typename T>
void prepare(uint32_t ntx, uint32_t nrx, uint32_t ns, uint32_t up, unsigned dz_up,
		const float* echoes, T* analytic, int nthreads)
{
	const unsigned L = ns * up;
	if (dz_up >= L) throw std::domain_error("Invalid dead zone: deadZoneSamplesUp >= signalLength");
	std::vector<T> h;
	if (up > 1) upsampling_filter<T>(up, static_cast<T>(0.2), h);
	const long nch = static_cast<long>(ntx) * nrx;
	const int nt = num_threads_or_default<T>(nthreads);
#pragma omp parallel num_threads(nt)
	{
		std::vector<T> in(ns), sig(L);
		Analytic<T> an(L);
#pragma omp for schedule(dynamic, 1)
		for (long ch = 0; ch < nch; ++ch) {
			const float* src = echoes + static_cast<size_t>(ch) * ns;
			for (unsigned i = 0; i < ns; ++i) in[i] = static_cast<T>(src[i]);
			if (up > 1) interpolate<T>(in.data(), ns, up, h, sig.data());
			else std::copy(in.begin(), in.end(), sig.begin());
			remove_dc<T>(sig.data(), L, dz_up);
			an.run(sig.data(), analytic + static_cast<size_t>(ch) * L * 2);
		}
	}
}

/* ---------------- RF path (DefaultSTAProcessor) ---------------- */

/* Interpolator4X.h:54-72 filter, from the recipe stated there: 81 taps,
 * h = sinc(w) .* kaiser(81, beta = 5), w = linspace(-10, 10, 81). */
template<typename T>
void interp4x_filter(std::vector<T>& h)
{
	const unsigned n = 81;
	std::vector<double> win;
	kaiser_window<double>(n, 5.0, win);
	h.resize(n);
	for (unsigned i = 0; i < n; ++i) {
		const double w = -10.0 + 20.0 * i / (n - 1.0);
		const double s = (w == 0.0) ? 1.0 : std::sin(M_PI * w) / (M_PI * w);
		h[i] = static_cast<T>(s * win[i]);
	}
}

/* DefaultSTAProcessor.h:136-164 for one frame: Interpolator4X (zero-stuff x4,
 * 81-tap filter, output aligned by 40) then removeDC over [dz, L). Output
 * real [ntx][nrx][4 ns]. */
template<typename T>
void prepare_rf(uint32_t ntx, uint32_t nrx, uint32_t ns, unsigned dz_up, const float* echoes, T* out, int nthreads)
{
	const unsigned L = ns * 4;
	if (dz_up >= L) throw std::domain_error("Invalid dead zone: deadZoneSamplesUp >= signalLength");
	std::vector<T> h;
	interp4x_filter<T>(h);
	const long nch = static_cast<long>(ntx) * nrx;
	const int nt = num_threads_or_default<T>(nthreads);
#pragma omp parallel num_threads(nt)
	{
		std::vector<T> in(ns);
#pragma omp for schedule(dynamic, 1)
		for (long ch = 0; ch < nch; ++ch) {
			const float* src = echoes + static_cast<size_t>(ch) * ns;
			for (unsigned i = 0; i < ns; ++i) in[i] = static_cast<T>(src[i]);
			T* dst = out + static_cast<size_t>(ch) * L;
			interpolate<T>(in.data(), ns, 4, h, dst);
			remove_dc<T>(dst, L, dz_up);
		}
	}
}

/* DefaultSTAProcessor.h:166-226: real per-rx sums (no apodization), value =
 * signed sum / (Ntx * Nrx), factor = SCF of the per-rx sums (or 1).  A
 * position below zero is skipped (the reference's size_t cast then fails
 * the i + 1 < L test for any realistic record). */
template<typename T>
void das_rf(const oracle_sta2d& cfg, const T* tensor, uint32_t L, const float* grid_xz, uint32_t n1, uint32_t n2,
		T* value, T* factor, int nthreads, T* rxsum_out = nullptr)
{
	const unsigned N = cfg.num_elements;
	const unsigned first = cfg.first_tx_elem, last = cfg.last_tx_elem;
	const unsigned ntx = last - first + 1U;
	std::vector<T> xa(N);
	const T pitch = static_cast<T>(cfg.pitch);
	const T xc = ((N - 1) * pitch) / 2;
	for (unsigned e = 0; e < N; ++e) xa[e] = e * pitch - xc;
	const T fs = static_cast<T>(cfg.sampling_frequency);
	const T c = static_cast<T>(cfg.propagation_speed);
	const T fc = static_cast<T>(cfg.center_frequency);
	const T invCT = (fs * 4) / c;
	const T offset = (fs * 4) * static_cast<T>(cfg.peak_offset) / fc;
	const size_t num_signals = static_cast<size_t>(ntx) * N;
	const int cf = cfg.cf_kind;
	const T cf_param = static_cast<T>(cfg.cf_param);
	const int nt = num_threads_or_default<T>(nthreads);
#pragma omp parallel num_threads(nt)
	{
		std::vector<T> sums(N), delay(N);
#pragma omp for schedule(dynamic, 1)
		for (long col = 0; col < static_cast<long>(n1); ++col) {
			for (unsigned row = 0; row < n2; ++row) {
				const size_t pt = static_cast<size_t>(col) * n2 + row;
				const T x = grid_xz[2 * pt], z = grid_xz[2 * pt + 1];
				for (unsigned e = 0; e < N; ++e) {
					const T dx = x - xa[e];
					delay[e] = std::sqrt(dx * dx + z * z) * invCT;
				}
				std::fill(sums.begin(), sums.end(), T(0));
				for (unsigned t = first; t <= last; ++t) {
					const T txd = delay[t];
					for (unsigned r = 0; r < N; ++r) {
						const T d = offset + txd + delay[r];
						if (d < 0) continue;
						const size_t idx = static_cast<size_t>(d);
						const T k = d - idx;
						if (idx + 1U < L) {
							const T* p = tensor + (static_cast<size_t>(t - first) * N + r) * L + idx;
							sums[r] += (1 - k) * p[0] + k * p[1];
						}
					}
				}
				if (rxsum_out) std::copy(sums.begin(), sums.end(), rxsum_out + pt * N);
				factor[pt] = (cf == ORACLE_CF_SCF) ? scf<T>(sums.data(), N, cf_param) : T(1);
				value[pt] = std::accumulate(sums.begin(), sums.end(), T(0));
			}
		}
	}
	const T scale = T(1) / num_signals;
	for (size_t i = 0; i < static_cast<size_t>(n1) * n2; ++i) value[i] *= scale;
}

/* ParallelHilbertEnvelope::calculateDim2 (ParallelHilbertEnvelope.h:73-83)
 * with HilbertEnvelope::calculate (HilbertEnvelope.h:201-223): every row of
 * the n1 x n2 matrix replaced by the magnitude of its analytic signal. */
template<typename T>
void envelope_dim2(T* data, uint32_t n1, uint32_t n2, int nthreads)
{
	const int nt = num_threads_or_default<T>(nthreads);
#pragma omp parallel num_threads(nt)
	{
		Analytic<T> an(n2);
		std::vector<T> z(2 * static_cast<size_t>(n2));
#pragma omp for schedule(dynamic, 1)
		for (long i = 0; i < static_cast<long>(n1); ++i) {
			T* row = data + static_cast<size_t>(i) * n2;
			an.run(row, z.data());
			for (unsigned j = 0; j < n2; ++j) row[j] = std::sqrt(z[2 * j] * z[2 * j] + z[2 * j + 1] * z[2 * j + 1]);
		}
	}
}

/* VectorialSTAProcessor.h:237-338.  `T` is the processor float type. */
template<typename T>
void das(const oracle_sta2d& cfg, const T* analytic_ri, uint32_t L,
		const float* grid_xz, uint32_t n1, uint32_t n2,
		const float* rx_apod_f, const float* tx_apod_f,
		T* value, T* factor, int nthreads, T* rxsum_out = nullptr)
{
	using C = std::complex<T>;
	const C* tensor = reinterpret_cast<const C*>(analytic_ri);
	const unsigned N = cfg.num_elements;
	const unsigned first = cfg.first_tx_elem, last = cfg.last_tx_elem;
	const unsigned ntx = last - first + 1U;
	/* ArrayGeometry.h:37-49 getElementXCentered2D */
	std::vector<T> xa(N);
	const T pitch = static_cast<T>(cfg.pitch);
	const T xc = ((N - 1) * pitch) / 2;
	for (unsigned e = 0; e < N; ++e) xa[e] = e * pitch - xc;
	std::vector<T> rx_apod(N), tx_apod(ntx, T(1));
	for (unsigned e = 0; e < N; ++e) rx_apod[e] = static_cast<T>(rx_apod_f[e]);
	if (tx_apod_f) for (unsigned t = 0; t < ntx; ++t) tx_apod[t] = static_cast<T>(tx_apod_f[t]);
	const T fs = static_cast<T>(cfg.sampling_frequency);
	const T c = static_cast<T>(cfg.propagation_speed);
	const T fc = static_cast<T>(cfg.center_frequency);
	const T invCT = (fs * cfg.upsampling_factor) / c;                                   /* :241 */
	const T offset = (fs * cfg.upsampling_factor) * static_cast<T>(cfg.peak_offset) / fc; /* :147 */
	const size_t num_signals = static_cast<size_t>(ntx) * N;                              /* :166 */
	const int cf = cfg.cf_kind;
	const T cf_param = static_cast<T>(cfg.cf_param);
	const int nt = num_threads_or_default<T>(nthreads);

#pragma omp parallel num_threads(nt)
	{
		std::vector<C> rx_sum(N);
		std::vector<T> delay(N), re(N);
		Pcf<T> pcf(cf_param);
#pragma omp for schedule(dynamic, 1)
		for (long col = 0; col < static_cast<long>(n1); ++col) {
			for (unsigned row = 0; row < n2; ++row) {
				const size_t pt = static_cast<size_t>(col) * n2 + row;
				const T px = static_cast<T>(grid_xz[2 * pt]);
				const T pz = static_cast<T>(grid_xz[2 * pt + 1]);
				/* Geometry.h:55-65 distance2DY0 (SSE form: dx = x - xe, sqrt(dx*dx + z*z)) */
				for (unsigned e = 0; e < N; ++e) {
					const T dx = px - xa[e];
					delay[e] = std::sqrt(dx * dx + pz * pz) * invCT;
				}
				if (cf != ORACLE_CF_NONE) {
					std::fill(rx_sum.begin(), rx_sum.end(), C(0));
					for (unsigned tx = first; tx <= last; ++tx) {
						const T tx_delay = delay[tx];
						const unsigned ltx = tx - first;
						for (unsigned rx = 0; rx < N; ++rx) {
							const T d = offset + tx_delay + delay[rx];
							if (d < 0) continue; /* unreachable for peak_offset >= 0 */
							const size_t idx = static_cast<size_t>(d);
							const T k = d - idx;
							if (idx + 1U < L) {
								const C* p = tensor + (static_cast<size_t>(ltx) * N + rx) * L + idx;
								const T w = tx_apod_f ? rx_apod[rx] * tx_apod[ltx] : rx_apod[rx];
								rx_sum[rx] += ((1 - k) * *p + k * *(p + 1)) * w;
							}
						}
					}
					if (rxsum_out) {
						for (unsigned rx = 0; rx < N; ++rx) {
							rxsum_out[(pt * N + rx) * 2] = rx_sum[rx].real();
							rxsum_out[(pt * N + rx) * 2 + 1] = rx_sum[rx].imag();
						}
					}
					if (cf == ORACLE_CF_PCF) {
						factor[pt] = pcf(rx_sum.data(), N);
					} else {
						for (unsigned rx = 0; rx < N; ++rx) re[rx] = rx_sum[rx].real();
						factor[pt] = scf<T>(re.data(), N, cf_param);
					}
					value[pt] = std::abs(std::accumulate(rx_sum.begin(), rx_sum.end(), C(0)));
				} else {
					C sum;
					for (unsigned tx = first; tx <= last; ++tx) {
						const T tx_delay = delay[tx];
						const unsigned ltx = tx - first;
						for (unsigned rx = 0; rx < N; ++rx) {
							const T d = offset + tx_delay + delay[rx];
							if (d < 0) continue;
							const size_t idx = static_cast<size_t>(d);
							const T k = d - idx;
							if (idx + 1U < L) {
								const C* p = tensor + (static_cast<size_t>(ltx) * N + rx) * L + idx;
								const T w = tx_apod_f ? rx_apod[rx] * tx_apod[ltx] : rx_apod[rx];
								sum += ((1 - k) * *p + k * *(p + 1)) * w;
							}
						}
					}
					value[pt] = std::abs(sum);
					factor[pt] = 1; /* Util::resetValueFactor, Util.h:990-999 */
				}
			}
		}
	}
	const T scale = T(1) / num_signals; /* :337-338 */
	const size_t npix = static_cast<size_t>(n1) * n2;
	for (size_t i = 0; i < npix; ++i) value[i] *= scale;
}

/* Delay-and-sum over a given FP32 delay table delays[row][npix] in samples
 * (tx t at row tx_row0 + t, rx r at row r; 2-D tables have tx_row0 =
 * first_tx, 3-D tables of active elements list the rx rows first,
 * Vectorial3DSTAProcessor.h:193-250), as the reference GPU processors
 * take theirs from calculateDelaysSTAKernel (VectorialSTACUDAProcessor.cu:
 * 30-53).  position = (offset + d_tx) + d_rx, skipped unless position >= 0
 * and idx = trunc(position) <= L - 2 (.cu:86-90, :156-160).
 * form 0: the CPU processor's arithmetic, ((1 - k) s[i] + k s[i+1]) * apod
 *         then +=, unfused (VectorialSTAProcessor.h:269-284, :312-325).
 * form 1: the CUDA kernels' arithmetic, v = v0 + k (v1 - v0) and
 *         sum += v * apod (.cu:93-103, :163-169), which nvcc contracts to
 *         fused multiply-adds (its default -fmad=true): fma(k, v1 - v0, v0),
 *         then fma(v, apod, sum).
 * With a coherence factor the per-rx sums accumulate over tx in ascending
 * order, then the value is |sum over rx| (.cu:172-178).  apod = rx apod,
 * times the optional tx apod (an extension) as one float product. */
template<typename T>
void das_table(const oracle_sta2d& cfg, const T* analytic_ri, uint32_t L, const float* delays, uint32_t rows,
		uint32_t tx_row0, uint64_t npix, const float* rx_apod_f, const float* tx_apod_f, T* value, T* factor, int form,
		int nthreads, T* rxsum_out = nullptr)
{
	using C = std::complex<T>;
	const C* tensor = reinterpret_cast<const C*>(analytic_ri);
	const unsigned N = cfg.num_elements;
	const unsigned first = cfg.first_tx_elem, last = cfg.last_tx_elem;
	const unsigned ntx = last - first + 1U;
	if (rows < N || rows < tx_row0 + ntx) throw std::invalid_argument("delay table has too few rows");
	std::vector<T> w(static_cast<size_t>(ntx) * N);
	for (unsigned t = 0; t < ntx; ++t)
		for (unsigned r = 0; r < N; ++r)
			w[static_cast<size_t>(t) * N + r] = tx_apod_f ? static_cast<T>(rx_apod_f[r] * tx_apod_f[t])
					: static_cast<T>(rx_apod_f[r]);
	const T fs = static_cast<T>(cfg.sampling_frequency);
	const T fc = static_cast<T>(cfg.center_frequency);
	const T offset = (fs * cfg.upsampling_factor) * static_cast<T>(cfg.peak_offset) / fc; /* :147 */
	const int cf = cfg.cf_kind;
	const T cf_param = static_cast<T>(cfg.cf_param);
	const int nt = num_threads_or_default<T>(nthreads);
	auto sample = [&](const C* p, T k) -> C {
		if (form == 1) {
			const C v0 = p[0], v1 = p[1];
			return C(std::fma(k, v1.real() - v0.real(), v0.real()), std::fma(k, v1.imag() - v0.imag(), v0.imag()));
		}
		return (1 - k) * p[0] + k * p[1];
	};
	auto accumulate = [&](C& acc, C v, T wv) {
		if (form == 1) acc = C(std::fma(v.real(), wv, acc.real()), std::fma(v.imag(), wv, acc.imag()));
		else acc += v * wv;
	};
#pragma omp parallel num_threads(nt)
	{
		std::vector<C> rx_sum(N);
		std::vector<T> re(N);
		Pcf<T> pcf(cf_param);
#pragma omp for schedule(dynamic, 64)
		for (long long q = 0; q < static_cast<long long>(npix); ++q) {
			const size_t pt = static_cast<size_t>(q);
			auto term = [&](unsigned ltx, unsigned rx, C& acc) {
				const T d = (offset + static_cast<T>(delays[static_cast<size_t>(tx_row0 + ltx) * npix + pt])) +
						static_cast<T>(delays[static_cast<size_t>(rx) * npix + pt]);
				if (!(d >= 0)) return;
				const size_t idx = static_cast<size_t>(d);
				if (idx + 1U >= L) return;
				const T k = d - static_cast<T>(idx);
				accumulate(acc, sample(tensor + (static_cast<size_t>(ltx) * N + rx) * L + idx, k),
						w[static_cast<size_t>(ltx) * N + rx]);
			};
			if (cf != ORACLE_CF_NONE) {
				std::fill(rx_sum.begin(), rx_sum.end(), C(0));
				for (unsigned ltx = 0; ltx < ntx; ++ltx)
					for (unsigned rx = 0; rx < N; ++rx) term(ltx, rx, rx_sum[rx]);
				if (rxsum_out) {
					for (unsigned rx = 0; rx < N; ++rx) {
						rxsum_out[(pt * N + rx) * 2] = rx_sum[rx].real();
						rxsum_out[(pt * N + rx) * 2 + 1] = rx_sum[rx].imag();
					}
				}
				if (cf == ORACLE_CF_PCF) {
					factor[pt] = pcf(rx_sum.data(), N);
				} else {
					for (unsigned rx = 0; rx < N; ++rx) re[rx] = rx_sum[rx].real();
					factor[pt] = scf<T>(re.data(), N, cf_param);
				}
				C s(0);
				for (unsigned rx = 0; rx < N; ++rx) s += rx_sum[rx];
				value[pt] = std::abs(s);
			} else {
				C s(0);
				for (unsigned ltx = 0; ltx < ntx; ++ltx)
					for (unsigned rx = 0; rx < N; ++rx) term(ltx, rx, s);
				value[pt] = std::abs(s);
				factor[pt] = 1;
			}
		}
	}
	const T scale = T(1) / (static_cast<size_t>(ntx) * N); /* :337-338 */
	for (uint64_t i = 0; i < npix; ++i) value[i] *= scale;
}

/* Vectorial3DSTAProcessor.h:193-250: separate tx/rx element lists,
 * Geometry::distance3DZ0 (Geometry.h:77-84). */
template<typename T>
void das3d(const oracle_sta3d& cfg, const T* analytic_ri, uint32_t L,
		const float* tx_xy, const float* rx_xy, const float* grid_xyz, uint32_t n1, uint32_t n2,
		const float* rx_apod_f, T* value, T* factor, int nthreads, T* rxsum_out = nullptr)
{
	using C = std::complex<T>;
	const C* tensor = reinterpret_cast<const C*>(analytic_ri);
	const unsigned ntx = cfg.num_tx, nrx = cfg.num_rx;
	const T fs = static_cast<T>(cfg.sampling_frequency);
	const T c = static_cast<T>(cfg.propagation_speed);
	const T fc = static_cast<T>(cfg.center_frequency);
	const T invCT = (fs * cfg.upsampling_factor) / c;
	const T offset = (fs * cfg.upsampling_factor) * static_cast<T>(cfg.peak_offset) / fc;
	const size_t num_signals = static_cast<size_t>(ntx) * nrx;
	const int cf = cfg.cf_kind;
	const T cf_param = static_cast<T>(cfg.cf_param);
	const int nt = num_threads_or_default<T>(nthreads);
	std::vector<T> rx_apod(nrx);
	for (unsigned r = 0; r < nrx; ++r) rx_apod[r] = static_cast<T>(rx_apod_f[r]);
#pragma omp parallel num_threads(nt)
	{
		std::vector<C> rx_sum(nrx);
		std::vector<T> txd(ntx), rxd(nrx), re(nrx);
		Pcf<T> pcf(cf_param);
#pragma omp for schedule(dynamic, 1)
		for (long i = 0; i < static_cast<long>(n1); ++i) {
			for (unsigned j = 0; j < n2; ++j) {
				const size_t pt = static_cast<size_t>(i) * n2 + j;
				const T px = grid_xyz[3 * pt], py = grid_xyz[3 * pt + 1], pz = grid_xyz[3 * pt + 2];
				for (unsigned t = 0; t < ntx; ++t) {
					const T dx = px - static_cast<T>(tx_xy[2 * t]);
					const T dy = py - static_cast<T>(tx_xy[2 * t + 1]);
					txd[t] = std::sqrt(dx * dx + dy * dy + pz * pz) * invCT;
				}
				for (unsigned r = 0; r < nrx; ++r) {
					const T dx = px - static_cast<T>(rx_xy[2 * r]);
					const T dy = py - static_cast<T>(rx_xy[2 * r + 1]);
					rxd[r] = std::sqrt(dx * dx + dy * dy + pz * pz) * invCT;
				}
				std::fill(rx_sum.begin(), rx_sum.end(), C(0));
				for (unsigned t = 0; t < ntx; ++t) {
					const T tx_delay = txd[t];
					for (unsigned r = 0; r < nrx; ++r) {
						const T d = offset + tx_delay + rxd[r];
						if (d < 0) continue;
						const size_t idx = static_cast<size_t>(d);
						const T k = d - idx;
						if (idx + 1U < L) {
							const C* p = tensor + (static_cast<size_t>(t) * nrx + r) * L + idx;
							rx_sum[r] += ((1 - k) * *p + k * *(p + 1)) * rx_apod[r];
						}
					}
				}
				if (rxsum_out) {
					for (unsigned r = 0; r < nrx; ++r) {
						rxsum_out[(pt * nrx + r) * 2] = rx_sum[r].real();
						rxsum_out[(pt * nrx + r) * 2 + 1] = rx_sum[r].imag();
					}
				}
				if (cf == ORACLE_CF_PCF) {
					factor[pt] = pcf(rx_sum.data(), nrx);
				} else if (cf == ORACLE_CF_SCF) {
					for (unsigned r = 0; r < nrx; ++r) re[r] = rx_sum[r].real();
					factor[pt] = scf<T>(re.data(), nrx, cf_param);
				} else {
					factor[pt] = 1;
				}
				value[pt] = std::abs(std::accumulate(rx_sum.begin(), rx_sum.end(), C(0)));
			}
		}
	}
	const T scale = T(1) / num_signals;
	const size_t npix = static_cast<size_t>(n1) * n2;
	for (size_t i = 0; i < npix; ++i) value[i] *= scale;
}

/* WindowFunction.h:172-289 (subset used by the STA path's callers). */
void window(const std::string& name, unsigned n, std::vector<double>& w)
{
	if (n < 2U) throw std::invalid_argument("The window function size must be >= 2.");
	w.resize(n);
	const double c1 = 1.0 / (n - 1.0);
	const double c3 = 1.0 / (n + 1.0);
	if (name == "rectangular") {
		for (unsigned i = 0; i < n; ++i) w[i] = 1.0;
	} else if (name == "hanning") {
		for (unsigned i = 0; i < n; ++i) w[i] = 0.5 - 0.5 * std::cos(2.0 * kPi * (i * c1));
	} else if (name == "hanning2") {
		for (unsigned i = 0; i < n; ++i) w[i] = 0.5 - 0.5 * std::cos(2.0 * kPi * ((i + 1.0) * c3));
	} else if (name == "hamming") {
		for (unsigned i = 0; i < n; ++i) w[i] = 0.54 - 0.46 * std::cos(2.0 * kPi * (i * c1));
	} else if (name == "blackman") {
		for (unsigned i = 0; i < n; ++i) {
			const double c2 = i * c1;
			w[i] = 0.42 - 0.5 * std::cos(2.0 * kPi * c2) + 0.08 * std::cos(4.0 * kPi * c2);
		}
	} else if (name == "blackman2") {
		for (unsigned i = 0; i < n; ++i) {
			const double c2 = (i + 1.0) * c3;
			w[i] = 0.42 - 0.5 * std::cos(2.0 * kPi * c2) + 0.08 * std::cos(4.0 * kPi * c2);
		}
	} else {
		throw std::invalid_argument("Invalid window function name: " + name);
	}
}

template<typename F>
int guarded(F&& f)
{
	try {
		f();
		return 0;
	} catch (const std::exception& e) {
		g_error = e.what();
		return -1;
	}
}

double now_s()
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

double oracle_kaiser_beta(double tol_db) { return kaiser_beta<double>(tol_db); }
uint32_t oracle_kaiser_size(double tol_db, double tw)
{
	uint32_t s = 0;
	guarded([&] { s = kaiser_size<double>(tol_db, tw); });
	return s;
}
void oracle_kaiser_window(uint32_t size, double beta, double* out)
{
	std::vector<double> w;
	kaiser_window<double>(size, beta, w);
	std::copy(w.begin(), w.end(), out);
}
double oracle_bessel_i0(double x) { return bessel_i0(x); }
uint32_t oracle_next_fast_even(uint32_t n) { return next_fast_even(n); }

uint32_t oracle_upsampling_filter_size(uint32_t up, double half_tw)
{
	std::vector<double> h;
	if (guarded([&] { upsampling_filter<double>(up, half_tw, h); }) != 0) return 0;
	return static_cast<uint32_t>(h.size());
}
void oracle_upsampling_filter_f64(uint32_t up, double half_tw, double* out)
{
	std::vector<double> h;
	upsampling_filter<double>(up, half_tw, h);
	std::copy(h.begin(), h.end(), out);
}
void oracle_upsampling_filter_f32(uint32_t up, float half_tw, float* out)
{
	std::vector<float> h;
	upsampling_filter<float>(up, half_tw, h);
	std::copy(h.begin(), h.end(), out);
}
void oracle_interpolate_f64(const double* x, size_t n, uint32_t up, double half_tw, double* out)
{
	std::vector<double> h;
	upsampling_filter<double>(up, half_tw, h);
	interpolate<double>(x, n, up, h, out);
}
void oracle_interpolate_f32(const float* x, size_t n, uint32_t up, float half_tw, float* out)
{
	std::vector<float> h;
	upsampling_filter<float>(up, half_tw, h);
	interpolate<float>(x, n, up, h, out);
}
void oracle_convolve_f64(const double* x, size_t nx, const double* b, size_t nb, double* y)
{
	convolve<double>(x, nx, b, nb, y);
}
void oracle_fft_real_f64(const double* x, uint32_t n, double* out_re, double* out_im)
{
	using C = std::complex<long double>;
	Fft<long double> f(n);
	std::vector<C> a(n), b(n);
	for (uint32_t i = 0; i < n; ++i) a[i] = C(x[i], 0);
	f.forward(a.data(), b.data());
	for (uint32_t i = 0; i < n; ++i) { out_re[i] = static_cast<double>(b[i].real()); out_im[i] = static_cast<double>(b[i].imag()); }
}
void oracle_ifft_f64(const double* in_re, const double* in_im, uint32_t n, double* out_re, double* out_im)
{
	using C = std::complex<long double>;
	Fft<long double> f(n);
	std::vector<C> a(n), b(n);
	for (uint32_t i = 0; i < n; ++i) a[i] = C(in_re[i], in_im[i]);
	f.inverse(a.data(), b.data());
	for (uint32_t i = 0; i < n; ++i) { out_re[i] = static_cast<double>(b[i].real()); out_im[i] = static_cast<double>(b[i].imag()); }
}
void oracle_analytic_signal_f64(const double* x, uint32_t n, double* out)
{
	Analytic<double> a(n);
	a.run(x, out);
}
void oracle_analytic_signal_f32(const float* x, uint32_t n, float* out)
{
	Analytic<float> a(n);
	a.run(x, out);
}
void oracle_remove_dc_f64(double* x, size_t n, size_t begin) { guarded([&] { remove_dc<double>(x, n, begin); }); }
void oracle_remove_dc_f32(float* x, size_t n, size_t begin) { guarded([&] { remove_dc<float>(x, n, begin); }); }
double oracle_std_dev_f64(const double* x, size_t n) { return std_dev<double>(x, n); }
double oracle_mean_f64(const double* x, size_t n) { return mean<double>(x, n); }
double oracle_pcf_f64(const double* ri, uint32_t n, double gamma)
{
	Pcf<double> p(gamma);
	return p(reinterpret_cast<const std::complex<double>*>(ri), n);
}
float oracle_pcf_f32(const float* ri, uint32_t n, float gamma)
{
	Pcf<float> p(gamma);
	return p(reinterpret_cast<const std::complex<float>*>(ri), n);
}
double oracle_scf_f64(const double* x, uint32_t n, double p) { return scf<double>(x, n, p); }
float oracle_scf_f32(const float* x, uint32_t n, float p) { return scf<float>(x, n, p); }
int oracle_window(const char* name, uint32_t n, double* out)
{
	return guarded([&] {
		std::vector<double> w;
		window(name, n, w);
		std::copy(w.begin(), w.end(), out);
	});
}

static int check_cfg(const oracle_sta2d* cfg)
{
	if (!cfg || cfg->num_elements == 0 || cfg->last_tx_elem < cfg->first_tx_elem ||
			cfg->last_tx_elem >= cfg->num_elements || cfg->num_samples == 0 || cfg->upsampling_factor == 0) {
		g_error = "invalid configuration";
		return -1;
	}
	return 0;
}

int oracle_prepare_f32(const oracle_sta2d* cfg, const float* echoes, float* analytic, int nthreads)
{
	if (check_cfg(cfg)) return -1;
	return guarded([&] {
		const unsigned dz = dead_zone_samples(cfg->upsampling_factor, cfg->sampling_frequency,
				cfg->dead_zone_m, cfg->propagation_speed, true);
		prepare<float>(cfg->last_tx_elem - cfg->first_tx_elem + 1, cfg->num_elements, cfg->num_samples,
				cfg->upsampling_factor, dz, echoes, analytic, nthreads);
	});
}

int oracle_prepare_f64(const oracle_sta2d* cfg, const float* echoes, double* analytic, int nthreads)
{
	if (check_cfg(cfg)) return -1;
	return guarded([&] {
		const unsigned dz = dead_zone_samples(cfg->upsampling_factor, cfg->sampling_frequency,
				cfg->dead_zone_m, cfg->propagation_speed, false);
		prepare<double>(cfg->last_tx_elem - cfg->first_tx_elem + 1, cfg->num_elements, cfg->num_samples,
				cfg->upsampling_factor, dz, echoes, analytic, nthreads);
	});
}

int oracle_das_f32(const oracle_sta2d* cfg, const float* analytic, uint32_t L,
		const float* grid_xz, uint32_t n1, uint32_t n2, const float* rx_apod, const float* tx_apod,
		float* value, float* factor, int nthreads)
{
	if (check_cfg(cfg)) return -1;
	return guarded([&] { das<float>(*cfg, analytic, L, grid_xz, n1, n2, rx_apod, tx_apod, value, factor, nthreads); });
}

int oracle_das_rxsums_f32(const oracle_sta2d* cfg, const float* analytic, uint32_t L,
		const float* grid_xz, uint32_t n1, uint32_t n2, const float* rx_apod, const float* tx_apod,
		float* value, float* factor, float* rxsums, int nthreads)
{
	if (check_cfg(cfg)) return -1;
	if (cfg->cf_kind == ORACLE_CF_NONE) {
		g_error = "per-rx sums need a coherence factor";
		return -1;
	}
	return guarded([&] {
		das<float>(*cfg, analytic, L, grid_xz, n1, n2, rx_apod, tx_apod, value, factor, nthreads, rxsums);
	});
}

int oracle_delays_f32(const oracle_sta2d* cfg, const float* grid_xz, uint32_t n1, uint32_t n2, float* delays)
{
	if (check_cfg(cfg)) return -1;
	return guarded([&] {
		/* the per-pixel delays of das<float>: ArrayGeometry.h:37-49 positions,
		 * Geometry.h:55-65 distance2DY0, times invCT (VectorialSTAProcessor.h:241) */
		const unsigned N = cfg->num_elements;
		const float pitch = static_cast<float>(cfg->pitch);
		const float xc = ((N - 1) * pitch) / 2;
		const float fs = static_cast<float>(cfg->sampling_frequency);
		const float c = static_cast<float>(cfg->propagation_speed);
		const float invCT = (fs * cfg->upsampling_factor) / c;
		const size_t npix = static_cast<size_t>(n1) * n2;
		for (unsigned e = 0; e < N; ++e) {
			const float xe = e * pitch - xc;
			for (size_t pt = 0; pt < npix; ++pt) {
				const float dx = grid_xz[2 * pt] - xe;
				const float pz = grid_xz[2 * pt + 1];
				delays[e * npix + pt] = std::sqrt(dx * dx + pz * pz) * invCT;
			}
		}
	});
}

int oracle_das_table_f32(const oracle_sta2d* cfg, const float* analytic, uint32_t L, const float* delays,
		uint32_t rows, uint32_t tx_row0, uint64_t npix, const float* rx_apod, const float* tx_apod, float* value,
		float* factor, float* rxsums, int form, int nthreads)
{
	/* the transmit rows are checked against the table, not num_elements */
	if (!cfg || cfg->num_elements == 0 || cfg->last_tx_elem < cfg->first_tx_elem || cfg->upsampling_factor == 0) {
		g_error = "invalid configuration";
		return -1;
	}
	if (rxsums && cfg->cf_kind == ORACLE_CF_NONE) {
		g_error = "per-rx sums need a coherence factor";
		return -1;
	}
	return guarded([&] {
		das_table<float>(*cfg, analytic, L, delays, rows, tx_row0, npix, rx_apod, tx_apod, value, factor, form, nthreads,
				rxsums);
	});
}

int oracle_das_f64(const oracle_sta2d* cfg, const double* analytic, uint32_t L,
		const float* grid_xz, uint32_t n1, uint32_t n2, const float* rx_apod, const float* tx_apod,
		double* value, double* factor, int nthreads)
{
	if (check_cfg(cfg)) return -1;
	return guarded([&] { das<double>(*cfg, analytic, L, grid_xz, n1, n2, rx_apod, tx_apod, value, factor, nthreads); });
}

int oracle_process_f32(const oracle_sta2d* cfg, const float* echoes,
		const float* grid_xz, uint32_t n1, uint32_t n2, const float* rx_apod, const float* tx_apod,
		float* value, float* factor, int nthreads, double* t_prep, double* t_das)
{
	if (check_cfg(cfg)) return -1;
	return guarded([&] {
		const unsigned ntx = cfg->last_tx_elem - cfg->first_tx_elem + 1;
		const unsigned L = cfg->num_samples * cfg->upsampling_factor;
		std::vector<float> analytic(static_cast<size_t>(ntx) * cfg->num_elements * L * 2);
		const unsigned dz = dead_zone_samples(cfg->upsampling_factor, cfg->sampling_frequency,
				cfg->dead_zone_m, cfg->propagation_speed, true);
		const double t0 = now_s();
		prepare<float>(ntx, cfg->num_elements, cfg->num_samples, cfg->upsampling_factor, dz, echoes,
				analytic.data(), nthreads);
		const double t1 = now_s();
		das<float>(*cfg, analytic.data(), L, grid_xz, n1, n2, rx_apod, tx_apod, value, factor, nthreads);
		const double t2 = now_s();
		if (t_prep) *t_prep = t1 - t0;
		if (t_das) *t_das = t2 - t1;
	});
}

int oracle_prepare3d_f32(const oracle_sta3d* cfg, const float* echoes, float* analytic, int nthreads)
{
	return guarded([&] {
		const unsigned dz = dead_zone_samples(cfg->upsampling_factor, cfg->sampling_frequency,
				cfg->dead_zone_m, cfg->propagation_speed, true);
		prepare<float>(cfg->num_tx, cfg->num_rx, cfg->num_samples, cfg->upsampling_factor, dz, echoes,
				analytic, nthreads);
	});
}

int oracle_das3d_f32(const oracle_sta3d* cfg, const float* analytic, uint32_t L,
		const float* tx_xy, const float* rx_xy, const float* grid_xyz, uint32_t n1, uint32_t n2,
		const float* rx_apod, float* value, float* factor, int nthreads)
{
	return guarded([&] {
		das3d<float>(*cfg, analytic, L, tx_xy, rx_xy, grid_xyz, n1, n2, rx_apod, value, factor, nthreads);
	});
}

int oracle_das3d_rxsums_f32(const oracle_sta3d* cfg, const float* analytic, uint32_t L,
		const float* tx_xy, const float* rx_xy, const float* grid_xyz, uint32_t n1, uint32_t n2,
		const float* rx_apod, float* value, float* factor, float* rxsums, int nthreads)
{
	return guarded([&] {
		das3d<float>(*cfg, analytic, L, tx_xy, rx_xy, grid_xyz, n1, n2, rx_apod, value, factor, nthreads, rxsums);
	});
}

int oracle_interp4x_filter_f64(double* h)
{
	return guarded([&] {
		std::vector<double> v;
		interp4x_filter<double>(v);
		std::copy(v.begin(), v.end(), h);
	});
}

int oracle_interpolate4x_f64(const double* x, uint32_t n, double* out)
{
	return guarded([&] {
		std::vector<double> h;
		interp4x_filter<double>(h);
		interpolate<double>(x, n, 4, h, out);
	});
}

int oracle_prepare_rf_f32(const oracle_sta2d* cfg, const float* echoes, float* out, int nthreads)
{
	return guarded([&] {
		const unsigned dz = dead_zone_samples(4, cfg->sampling_frequency, cfg->dead_zone_m, cfg->propagation_speed, true);
		prepare_rf<float>(cfg->last_tx_elem - cfg->first_tx_elem + 1, cfg->num_elements, cfg->num_samples, dz, echoes,
				out, nthreads);
	});
}

int oracle_prepare_rf_f64(const oracle_sta2d* cfg, const float* echoes, double* out, int nthreads)
{
	return guarded([&] {
		const unsigned dz = dead_zone_samples(4, cfg->sampling_frequency, cfg->dead_zone_m, cfg->propagation_speed, false);
		prepare_rf<double>(cfg->last_tx_elem - cfg->first_tx_elem + 1, cfg->num_elements, cfg->num_samples, dz, echoes,
				out, nthreads);
	});
}

int oracle_das_rf_f32(const oracle_sta2d* cfg, const float* tensor, uint32_t L, const float* grid_xz, uint32_t n1,
		uint32_t n2, float* value, float* factor, int nthreads)
{
	return guarded([&] { das_rf<float>(*cfg, tensor, L, grid_xz, n1, n2, value, factor, nthreads); });
}

int oracle_das_rf_rxsums_f32(const oracle_sta2d* cfg, const float* tensor, uint32_t L, const float* grid_xz, uint32_t n1,
		uint32_t n2, float* value, float* factor, float* rxsums, int nthreads)
{
	return guarded([&] { das_rf<float>(*cfg, tensor, L, grid_xz, n1, n2, value, factor, nthreads, rxsums); });
}

int oracle_das_rf_f64(const oracle_sta2d* cfg, const double* tensor, uint32_t L, const float* grid_xz, uint32_t n1,
		uint32_t n2, double* value, double* factor, int nthreads)
{
	return guarded([&] { das_rf<double>(*cfg, tensor, L, grid_xz, n1, n2, value, factor, nthreads); });
}

int oracle_envelope_dim2_f32(float* data, uint32_t n1, uint32_t n2, int nthreads)
{
	return guarded([&] { envelope_dim2<float>(data, n1, n2, nthreads); });
}

int oracle_envelope_dim2_f64(double* data, uint32_t n1, uint32_t n2, int nthreads)
{
	return guarded([&] { envelope_dim2<double>(data, n1, n2, nthreads); });
}

int oracle_das3d_f64(const oracle_sta3d* cfg, const double* analytic, uint32_t L,
		const float* tx_xy, const float* rx_xy, const float* grid_xyz, uint32_t n1, uint32_t n2,
		const float* rx_apod, double* value, double* factor, int nthreads)
{
	return guarded([&] {
		das3d<double>(*cfg, analytic, L, tx_xy, rx_xy, grid_xyz, n1, n2, rx_apod, value, factor, nthreads);
	});
}

} // extern "C"

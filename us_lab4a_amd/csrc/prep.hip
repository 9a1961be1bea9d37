// Signal preparation on the GPU: Kaiser-sinc upsampling, DC removal and the
// FFT analytic signal (VectorialSTAProcessor.h:169-231, Interpolator.h:114-138,
// Util.h:1009-1018, HilbertEnvelope.h:142-251), plus the FP64 element-to-pixel
// delay table (replaces calculateDelaysSTAKernel, VectorialSTACUDAProcessor.cu:30-53).
//
// Per chunk of channels (one channel = one (tx, rx) echo):
//   fir_dc_kernel     raw [ch][ns] -> real [ch][Nf]: polyphase FIR from LDS,
//                     mean of [dz, L) removed, zero padding [L, Nf)
//   hipfft R2C        real [ch][Nf] -> spec [ch][Nf/2+1]
//   hilbert_kernel    spec -> z [ch][Nf]: DC/Nyquist x1/Nf, positive bins
//                     x2/Nf, negative bins 0
//   hipfft C2C (inv)  z -> z in place
//   store_kernel      first L samples of z -> analytic tensor in the storage
//                     type (FP16/BF16 with a per-channel scale)
#include "common.h"

#include <algorithm>
#include <cmath>

namespace mista {

uint32_t next_fast_even(uint32_t n)
{
	if (n == 0) MISTA_THROW(MISTA_E_INVALID_PARAM, "n must be > 0.");
	if (n & 1) ++n;
	for (;;) {
		uint32_t m = n;
		while ((m & 1) == 0) m >>= 1;
		while (m % 3 == 0) m /= 3;
		while (m % 5 == 0) m /= 5;
		if (m == 1) return n;
		n += 2;
	}
}

// Interpolator::prepare (Interpolator.h:64-111): Kaiser tolerance 1e-4
// (80 dB, beta = 0.1102*(80-8.7)), transition width 2*half_tw/U, size rounded
// to k*2U+1, windowed sinc.  Evaluated in double.
std::vector<double> upsampling_filter(uint32_t up, double half_tw)
{
	if (up < 2 || up > 64) MISTA_THROW(MISTA_E_INVALID_VALUE, "Invalid upsampling factor: " << up);
	const double pi = 3.14159265358979323846;
	const double tol_db = -20.0 * std::log10(1.0e-4);
	const double beta = 0.1102 * (tol_db - 8.7);
	const double tw = (half_tw * 2) / up;
	const double wsize_d = std::ceil((tol_db - 7.95) / (2.285 * tw * pi) + 1.0);
	const uint32_t wsize = static_cast<uint32_t>(wsize_d);
	const double two_up = 2.0 * up;
	const uint32_t fsize = static_cast<uint32_t>(std::ceil((wsize - 1) / two_up) * two_up) + 1U;
	const double num_periods = (fsize - 1U) / two_up;
	const double step = (2 * num_periods) / (fsize - 1U);
	std::vector<double> h(fsize);
	const double c1 = 2.0 / (fsize - 1U);
	const double c2 = 1.0 / std::cyl_bessel_i(0.0, beta);
	for (uint32_t i = 0; i < fsize; ++i) {
		const double x = -num_periods + i * step;
		const double s = (std::abs(x) < 1e-12) ? 1.0 : std::sin(pi * x) / (pi * x);
		const double k = c1 * i - 1;
		double a = 1 - k * k;
		if (a < 0) a = 0;
		h[i] = s * c2 * std::cyl_bessel_i(0.0, beta * std::sqrt(a));
	}
	return h;
}

// ---------------------------------------------------------------- delays

// One thread per (pixel, element).  FP64 distance, rounded once to FP32
// samples: d = sqrt((x_p - x_e)^2 + z_p^2) * (fs*U/c), x_e from
// ArrayGeometry.h:37-49.  Layout [element][pixel] (pixel fastest).
__global__ void __launch_bounds__(256) delay_table_kernel(const float2* __restrict__ grid, uint64_t npix,
		uint32_t num_elements, double pitch, double inv_ct, float* __restrict__ delay)
{
	const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	const uint32_t e = blockIdx.y;
	if (p >= npix) return;
	const float2 g = grid[p];
	const double xe = e * pitch - ((num_elements - 1) * pitch) / 2;
	const double dx = static_cast<double>(g.x) - xe;
	const double z = static_cast<double>(g.y);
	delay[static_cast<uint64_t>(e) * npix + p] = static_cast<float>(sqrt(dx * dx + z * z) * inv_ct);
}

void launch_delay_table(const float* d_grid_xz, uint64_t npix, uint32_t num_elements, double pitch,
		double inv_ct, float* d_delay, hipStream_t s)
{
	if (npix == 0) return;
	const dim3 grid(static_cast<uint32_t>((npix + 255) / 256), num_elements);
	hipLaunchKernelGGL(delay_table_kernel, grid, dim3(256), 0, s,
			reinterpret_cast<const float2*>(d_grid_xz), npix, num_elements, pitch, inv_ct, d_delay);
	HIP_CHECK(hipGetLastError());
}

// 3-D form, Geometry::distance3DZ0 (Geometry.h:77-84): elements at (x_e, y_e, 0)
// from an explicit position list (rows of the table: active rx elements, then
// active tx elements), grid points (x, y, z).  FP64 distance, rounded once.
__global__ void __launch_bounds__(256) delay_table_3d_kernel(const float* __restrict__ grid_xyz, uint64_t npix,
		const double2* __restrict__ elem_xy, double inv_ct, float* __restrict__ delay)
{
	const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	const uint32_t e = blockIdx.y;
	if (p >= npix) return;
	const double2 el = elem_xy[e];
	const double dx = static_cast<double>(grid_xyz[3 * p]) - el.x;
	const double dy = static_cast<double>(grid_xyz[3 * p + 1]) - el.y;
	const double z = static_cast<double>(grid_xyz[3 * p + 2]);
	delay[static_cast<uint64_t>(e) * npix + p] = static_cast<float>(sqrt(dx * dx + dy * dy + z * z) * inv_ct);
}

void launch_delay_table_3d(const float* d_grid_xyz, uint64_t npix, const double* d_elem_xy, uint32_t rows,
		double inv_ct, float* d_delay, hipStream_t s)
{
	if (npix == 0) return;
	const dim3 grid(static_cast<uint32_t>((npix + 255) / 256), rows);
	hipLaunchKernelGGL(delay_table_3d_kernel, grid, dim3(256), 0, s, d_grid_xyz, npix,
			reinterpret_cast<const double2*>(d_elem_xy), inv_ct, d_delay);
	HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- FIR + DC

constexpr int kFirThreads = 1024;

// Largest |stored value| of a channel's FIR output (non-negative floats), a
// 1024-lane workgroup's reduction; thread 0 writes it.  The four-step path's
// FP16/BF16 scale takes it from here instead of a second pass over the row.
__device__ __forceinline__ void store_channel_max(float m, float* __restrict__ dst, float* red)
{
	for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
	if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int w = 1; w < kFirThreads / 64; ++w) m = fmaxf(m, red[w]);
		*dst = m;
	}
}

// One workgroup per channel.  The raw channel sits in LDS with J zeros on the
// left and enough on the right, the polyphase filter hp[phase][j] =
// h[phase + U*j] (zero padded to J taps per phase) follows it.  Output
// y[i] = sum_j hp[(i+off)%U][j] * x[(i+off)/U - j], i in [0, L); the sum over
// [dz, L) (double) gives the mean removed from every sample; [L, Nf) is zero.
template<int VPT>
__global__ void __launch_bounds__(kFirThreads) fir_dc_kernel(const float* __restrict__ raw, uint32_t ns,
		uint32_t up, uint32_t L, uint32_t nf, uint32_t dz, const float* __restrict__ hpoly, uint32_t J,
		uint32_t off, float* __restrict__ out, float* __restrict__ xmax)
{
	extern __shared__ float smem[];
	__shared__ float redm[kFirThreads / 64];
	const uint32_t ch = blockIdx.x;
	const uint32_t tid = threadIdx.x;
	const uint32_t right = off / up + 2;
	float* xs = smem;                               // [J + ns + right]
	float* hs = smem + J + ns + right;              // [up][J]
	__shared__ double red[kFirThreads / 64];

	const float* src = raw + static_cast<uint64_t>(ch) * ns;
	for (uint32_t i = tid; i < J + ns + right; i += kFirThreads) {
		const int q = static_cast<int>(i) - static_cast<int>(J);
		xs[i] = (q >= 0 && q < static_cast<int>(ns)) ? src[q] : 0.0f;
	}
	for (uint32_t i = tid; i < up * J; i += kFirThreads) hs[i] = hpoly[i];
	__syncthreads();

	float y[VPT];
	double part = 0.0;
#pragma unroll
	for (int v = 0; v < VPT; ++v) {
		const uint32_t i = tid + static_cast<uint32_t>(v) * kFirThreads;
		float acc = 0.0f;
		if (i < L) {
			const uint32_t m = i + off;
			const uint32_t ph = m % up;
			const uint32_t q0 = m / up;
			const float* hrow = hs + ph * J;
			const float* xrow = xs + J + q0;
			for (uint32_t j = 0; j < J; ++j) acc = fmaf(hrow[j], xrow[-static_cast<int>(j)], acc);
			if (i >= dz) part += static_cast<double>(acc);
		}
		y[v] = acc;
	}
	// block reduction of the DC sum (fixed order: lanes, then waves)
	for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
	if ((tid & 63) == 0) red[tid >> 6] = part;
	__syncthreads();
	if (tid == 0) {
		double s = 0.0;
		for (int w = 0; w < kFirThreads / 64; ++w) s += red[w];
		red[0] = s;
	}
	__syncthreads();
	const float mean = static_cast<float>(red[0] / static_cast<double>(L - dz));

	float* dst = out + static_cast<uint64_t>(ch) * nf;
	float mx = 0.0f;
#pragma unroll
	for (int v = 0; v < VPT; ++v) {
		const uint32_t i = tid + static_cast<uint32_t>(v) * kFirThreads;
		if (i < L) {
			dst[i] = y[v] - mean;
			mx = fmaxf(mx, fabsf(y[v] - mean));
		}
	}
	for (uint32_t i = L + tid; i < nf; i += kFirThreads) dst[i] = 0.0f;
	if (xmax) store_channel_max(mx, xmax + ch, redm);
}

// U = 1: copy + DC removal (Util::removeDC only).
__global__ void __launch_bounds__(kFirThreads) copy_dc_kernel(const float* __restrict__ raw, uint32_t L,
		uint32_t nf, uint32_t dz, float* __restrict__ out, float* __restrict__ xmax)
{
	__shared__ double red[kFirThreads / 64];
	__shared__ float redm[kFirThreads / 64];
	const uint32_t ch = blockIdx.x;
	const uint32_t tid = threadIdx.x;
	const float* src = raw + static_cast<uint64_t>(ch) * L;
	double part = 0.0;
	for (uint32_t i = dz + tid; i < L; i += kFirThreads) part += static_cast<double>(src[i]);
	for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
	if ((tid & 63) == 0) red[tid >> 6] = part;
	__syncthreads();
	if (tid == 0) {
		double s = 0.0;
		for (int w = 0; w < kFirThreads / 64; ++w) s += red[w];
		red[0] = s;
	}
	__syncthreads();
	const float mean = static_cast<float>(red[0] / static_cast<double>(L - dz));
	float* dst = out + static_cast<uint64_t>(ch) * nf;
	float mx = 0.0f;
	for (uint32_t i = tid; i < nf; i += kFirThreads) {
		const float v = (i < L) ? src[i] - mean : 0.0f;
		dst[i] = v;
		mx = fmaxf(mx, fabsf(v));
	}
	if (xmax) store_channel_max(mx, xmax + ch, redm);
}

// ---------------------------------------------------------------- Hilbert

// spec [ch][Nf/2+1] -> z [ch][Nf]: bin 0 and Nf/2 scaled by 1/Nf (the
// reference's 0.5 then 2/Nf), bins 1..Nf/2-1 by 2/Nf, bins > Nf/2 zeroed.
__global__ void __launch_bounds__(256) hilbert_kernel(const float2* __restrict__ spec, uint32_t nf,
		float2* __restrict__ z)
{
	const uint32_t ch = blockIdx.y;
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= nf) return;
	const uint32_t half = nf / 2;
	const float inv = 1.0f / static_cast<float>(nf);
	float2 v = make_float2(0.0f, 0.0f);
	if (k <= half) {
		const float2 s = spec[static_cast<uint64_t>(ch) * (half + 1) + k];
		const float c = (k == 0 || k == half) ? inv : 2.0f * inv;
		v = make_float2(s.x * c, s.y * c);
	}
	z[static_cast<uint64_t>(ch) * nf + k] = v;
}

// ---------------------------------------------------------------- fused LDS path

__device__ inline uint32_t pack_bf16x2(float re, float im)
{
	// round to nearest even
	auto rne = [](float f) -> uint32_t {
		uint32_t b = __float_as_uint(f);
		b += 0x7FFFu + ((b >> 16) & 1u);
		return b >> 16;
	};
	return rne(re) | (rne(im) << 16);
}

__device__ inline uint32_t pack_f16x2(float re, float im)
{
	const __half2 h = __floats2half2_rn(re, im);
	return *reinterpret_cast<const uint32_t*>(&h);
}

// One workgroup per channel, everything between the raw echo and the
// analytic signal on chip: FIR x U + DC removal into LDS, then the analytic
// signal of size N = Nf (HilbertEnvelope.h:142-251) with two complex FFTs of
// size M = N/2 held in LDS:
//   forward:  Z = DFT_M(x[2m] + i x[2m+1])            (real FFT of size N)
//   Hilbert:  the analytic signal is x + i y with y the Hilbert transform of
//             x, Y[k] = -i sgn(k) X[k] (sgn 0 at bins 0 and N/2), a real
//             signal: y[2m] + i y[2m+1] = IDFT_M(C)[m] / M with C[0] = 0,
//             C[k] = i sin(2 pi k/N) Z[k] + cos(2 pi k/N) conj(Z[M-k]);
//   output:   (x_i - mean, y_i), x recomputed from the echo and the filter,
//             which stay in LDS after the transform buffer.
// The inverse runs as a forward transform of the conjugate.  FFTs are
// Stockham passes of radix 8/4/2/3/5: every thread reads its butterflies'
// inputs into registers, a barrier, then writes the outputs in place.

constexpr int kFftThreads = 1024;
constexpr uint32_t kFirJ = 27;   // taps per phase of the Kaiser upsampling design (every U)
constexpr uint32_t kFftMaxM = 18432;
static_assert(kFftMaxM < (1u << 15), "the FFT passes divide by multiply-high, exact below 2^15");
constexpr int kFftMaxPasses = 16;

struct FftPlanArgs {
	uint32_t M, npass;
	uint32_t radix[kFftMaxPasses];
	uint32_t ns[kFftMaxPasses];        // sub-transform size before the pass
	uint32_t ns_magic[kFftMaxPasses];  // ceil(2^32 / ns) (0 when ns = 1): b / ns = mulhi(b, ns_magic)
	uint32_t tw_off[kFftMaxPasses];    // offset of the pass's twiddles in tw
};

struct AnalyticArgs {
	const float* raw;      // [nch][ns], channels of this launch start here
	const float* hpoly;    // [up][J]
	const float2* tw;      // Stockham twiddles exp(-2 pi i * i * j / (Ns * R))
	const float2* wpost;   // exp(-2 pi i k / N), k in [0, M/2]
	float2* out;           // [blocks][L] analytic signal, or
	float* out_mag;        // [blocks][L] its magnitude (kOutMagnitude), or
	uint32_t* out16;       // [blocks][L] FP16/BF16 pairs over a per-channel scale
	float* scale;          // [blocks] that scale (out16)
	uint32_t ns, up, J, off, L, N, dz;
	uint32_t flags;        // kRemoveDc | kOutMagnitude | kOutBf16
	FftPlanArgs plan;
};
constexpr uint32_t kRemoveDc = 1, kOutMagnitude = 2, kOutBf16 = 4;

// The FIR taps, the same for every lane, are read through a pointer in the
// constant address space, so that they come through the scalar cache: where
// global stores follow in the same loop (the output stage), the compiler
// cannot otherwise prove them unchanged and loads them with vector loads,
// per lane and output (1.3 ms of a 4.8 ms FP16 preparation).
typedef __attribute__((address_space(4))) const float* cf32_t;

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
	return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }   // -i a
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// In-register forward DFT of size R (sign -1).
template<int R> __device__ __forceinline__ void dft(float2* v);
template<> __device__ __forceinline__ void dft<2>(float2* v)
{
	const float2 a = v[0], b = v[1];
	v[0] = cadd(a, b);
	v[1] = csub(a, b);
}
template<> __device__ __forceinline__ void dft<4>(float2* v)
{
	const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
	const float2 s13 = cadd(v[1], v[3]), d13 = mul_mi(csub(v[1], v[3]));
	v[0] = cadd(s02, s13);
	v[2] = csub(s02, s13);
	v[1] = cadd(d02, d13);
	v[3] = csub(d02, d13);
}
template<> __device__ __forceinline__ void dft<8>(float2* v)
{
	float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
	dft<4>(e);
	dft<4>(o);
	constexpr float c = 0.70710678118654752440f;
	const float2 o1 = make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));     // * (c, -c)
	const float2 o2 = mul_mi(o[2]);                                                  // * (0, -1)
	const float2 o3 = make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));    // * (-c, -c)
	v[0] = cadd(e[0], o[0]);
	v[4] = csub(e[0], o[0]);
	v[1] = cadd(e[1], o1);
	v[5] = csub(e[1], o1);
	v[2] = cadd(e[2], o2);
	v[6] = csub(e[2], o2);
	v[3] = cadd(e[3], o3);
	v[7] = csub(e[3], o3);
}
template<> __device__ __forceinline__ void dft<3>(float2* v)
{
	constexpr float h = 0.86602540378443864676f;   // sin(2 pi / 3)
	const float2 s = cadd(v[1], v[2]);
	const float2 d = csub(v[1], v[2]);
	const float2 t = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
	const float2 u = make_float2(h * d.y, -h * d.x);   // -i h d
	v[0] = cadd(v[0], s);
	v[1] = cadd(t, u);
	v[2] = csub(t, u);
}
template<> __device__ __forceinline__ void dft<5>(float2* v)
{
	constexpr float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
	constexpr float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
	const float2 b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
	const float2 d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
	const float2 t1 = make_float2(v[0].x + c1 * b1.x + c2 * b2.x, v[0].y + c1 * b1.y + c2 * b2.y);
	const float2 t2 = make_float2(v[0].x + c2 * b1.x + c1 * b2.x, v[0].y + c2 * b1.y + c1 * b2.y);
	const float2 u1 = make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
	const float2 u2 = make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
	v[0] = cadd(v[0], cadd(b1, b2));
	v[1] = cadd(t1, mul_mi(u1));
	v[4] = csub(t1, mul_mi(u1));
	v[2] = cadd(t2, mul_mi(u2));
	v[3] = csub(t2, mul_mi(u2));
}

// Composite radices: 6 = 2 x 3 and 9 = 3 x 3 in registers (Cooley-Tukey,
// inner twiddles exp(-2 pi i n k / R) as constants), so that one LDS pass
// does the work of two.
template<> __device__ __forceinline__ void dft<6>(float2* v)
{
	constexpr float h = 0.86602540378443864676f;   // sin(2 pi / 6)
	float2 e[3] = {v[0], v[2], v[4]}, o[3] = {v[1], v[3], v[5]};
	dft<3>(e);
	dft<3>(o);
	// o[k] * exp(-2 pi i k / 6): k = 1: (1/2, -h); k = 2: (-1/2, -h)
	const float2 o1 = make_float2(0.5f * o[1].x + h * o[1].y, 0.5f * o[1].y - h * o[1].x);
	const float2 o2 = make_float2(-0.5f * o[2].x + h * o[2].y, -0.5f * o[2].y - h * o[2].x);
	v[0] = cadd(e[0], o[0]);
	v[3] = csub(e[0], o[0]);
	v[1] = cadd(e[1], o1);
	v[4] = csub(e[1], o1);
	v[2] = cadd(e[2], o2);
	v[5] = csub(e[2], o2);
}
template<> __device__ __forceinline__ void dft<9>(float2* v)
{
	// columns n1 = 0..2: DFT_3 over n2 of v[n1 + 3 n2]
	float2 a[3][3];
#pragma unroll
	for (int n1 = 0; n1 < 3; ++n1) {
		a[n1][0] = v[n1];
		a[n1][1] = v[n1 + 3];
		a[n1][2] = v[n1 + 6];
		dft<3>(a[n1]);
	}
	// twiddles exp(-2 pi i n1 k2 / 9): (n1, k2) = (1,1), (1,2), (2,1), (2,2)
	constexpr float c1 = 0.76604444311897803520f, s1 = 0.64278760968653932632f;   // 40 degrees
	constexpr float c2 = 0.17364817766693034885f, s2 = 0.98480775301220805936f;   // 80 degrees
	constexpr float c4 = -0.93969262078590838405f, s4 = 0.34202014332566873304f;  // 160 degrees
	a[1][1] = cmul(a[1][1], make_float2(c1, -s1));
	a[1][2] = cmul(a[1][2], make_float2(c2, -s2));
	a[2][1] = cmul(a[2][1], make_float2(c2, -s2));
	a[2][2] = cmul(a[2][2], make_float2(c4, -s4));
	// rows k2: DFT_3 over n1, output k = k2 + 3 k1
#pragma unroll
	for (int k2 = 0; k2 < 3; ++k2) {
		float2 b[3] = {a[0][k2], a[1][k2], a[2][k2]};
		dft<3>(b);
		v[k2] = b[0];
		v[k2 + 3] = b[1];
		v[k2 + 6] = b[2];
	}
}

template<int R>
__device__ __forceinline__ void fft_pass(float2* D, uint32_t M, uint32_t Ns, uint32_t magic,
		const float2* __restrict__ tw)
{
	constexpr int MAXB = static_cast<int>((kFftMaxM / R + kFftThreads - 1) / kFftThreads);
	const uint32_t nb = M / R;
	float2 v[MAXB][R];
#pragma unroll
	for (int t = 0; t < MAXB; ++t) {
		const uint32_t b = threadIdx.x + static_cast<uint32_t>(t) * kFftThreads;
		if (b < nb) {
#pragma unroll
			for (int i = 0; i < R; ++i) v[t][i] = D[b + i * nb];
		}
	}
	__syncthreads();
#pragma unroll
	for (int t = 0; t < MAXB; ++t) {
		const uint32_t b = threadIdx.x + static_cast<uint32_t>(t) * kFftThreads;
		if (b < nb) {
			// q = b / Ns by a multiply-high: with magic = ceil(2^32 / Ns) =
			// (2^32 + e) / Ns, b * magic / 2^32 = b / Ns + b e / (Ns 2^32),
			// and b e < 2^32 (b, e < 2^15) keeps the floor exact
			const uint32_t q = (Ns == 1) ? b : __umulhi(b, magic);
			const uint32_t j = b - q * Ns;
			if (Ns > 1) {   // (the first pass's twiddles are all 1)
				const float2* w = tw + j * (R - 1);
#pragma unroll
				for (int i = 1; i < R; ++i) v[t][i] = cmul(v[t][i], w[i - 1]);
			}
			dft<R>(v[t]);
			const uint32_t o = q * Ns * R + j;
#pragma unroll
			for (int i = 0; i < R; ++i) D[o + i * Ns] = v[t][i];
		}
	}
	__syncthreads();
}

__device__ void fft_forward(float2* D, const FftPlanArgs& P, const float2* __restrict__ tw)
{
	for (uint32_t p = 0; p < P.npass; ++p) {
		const float2* t = tw + P.tw_off[p];
		switch (P.radix[p]) {
		case 8: fft_pass<8>(D, P.M, P.ns[p], P.ns_magic[p], t); break;
		case 9: fft_pass<9>(D, P.M, P.ns[p], P.ns_magic[p], t); break;
		case 6: fft_pass<6>(D, P.M, P.ns[p], P.ns_magic[p], t); break;
		case 4: fft_pass<4>(D, P.M, P.ns[p], P.ns_magic[p], t); break;
		case 2: fft_pass<2>(D, P.M, P.ns[p], P.ns_magic[p], t); break;
		case 3: fft_pass<3>(D, P.M, P.ns[p], P.ns_magic[p], t); break;
		default: fft_pass<5>(D, P.M, P.ns[p], P.ns_magic[p], t); break;
		}
	}
}

// FIR x U output i (no DC removal): the generic polyphase loop over the
// echo and the filter in LDS, or the raw sample when U = 1.  Deterministic,
// so a second evaluation gives the same bits as the first.
__device__ __forceinline__ float fir_at(const AnalyticArgs& A, const float* xs, const float* hs, const float* src,
		uint32_t i)
{
	if (A.up <= 1) return src[i];
	const uint32_t m = i + A.off;
	const float* hrow = hs + (m % A.up) * A.J;
	const float* xrow = xs + A.J + m / A.up;
	float acc = 0.0f;
	for (uint32_t j = 0; j < A.J; ++j) acc = fmaf(hrow[j], xrow[-static_cast<int>(j)], acc);
	return acc;
}

__global__ void __launch_bounds__(kFftThreads) analytic_lds_kernel(const AnalyticArgs A)
{
	extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
	__shared__ double red[kFftThreads / 64];
	__shared__ float redm[kFftThreads / 64], redx[kFftThreads / 64], redn[kFftThreads / 64];
	const uint32_t tid = threadIdx.x, ch = blockIdx.x;
	const uint32_t M = A.plan.M, N = A.N, L = A.L, up = A.up, J = A.J;
	float2* D = reinterpret_cast<float2*>(lds);
	float* Df = reinterpret_cast<float*>(lds);
	const uint32_t right = A.off / up + 2;
	float* xs = reinterpret_cast<float*>(lds + static_cast<size_t>(M) * sizeof(float2));   // [J + ns + right]
	float* hs = xs + J + A.ns + right;                                                    // [up][J]

	// FIR x U (Interpolator.h:64-138, polyphase) and DC removal
	// (Util.h:1009-1018) into the packed real array Df[0, L); Df[L, N) = 0.
	const float* src = A.raw + static_cast<uint64_t>(ch) * A.ns;
	if (up > 1) {
		for (uint32_t i = tid; i < J + A.ns + right; i += kFftThreads) {
			const int q = static_cast<int>(i) - static_cast<int>(J);
			xs[i] = (q >= 0 && q < static_cast<int>(A.ns)) ? src[q] : 0.0f;
		}
		for (uint32_t i = tid; i < up * J; i += kFftThreads) hs[i] = A.hpoly[i];
		__syncthreads();
	}
	double part = 0.0;
	float xmax = -3.4e38f, xmin = 3.4e38f;   // extremes of the FIR output over [0, L) (storage scale)
	if (up > 1 && J == kFirJ) {
		// Lane per input position q: the window x[q-26..q] sits in registers
		// and serves all U phases (outputs m = q*U + p, i = m - off); the
		// taps, the same for every lane, come through the scalar cache.  Each
		// output is the same fma chain over j = 0..J-1 as below.
		const cf32_t hg = (cf32_t)A.hpoly;
		const uint32_t q_lo = A.off / up, q_hi = (L - 1 + A.off) / up;
		for (uint32_t q = q_lo + tid; q <= q_hi; q += kFftThreads) {
			float xw[kFirJ];
			const float* xrow = xs + J + q;
#pragma unroll
			for (int j = 0; j < kFirJ; ++j) xw[j] = xrow[-j];
			const int ibase = static_cast<int>(q * up) - static_cast<int>(A.off);
			// U = 8 or 4 with the lane's outputs inside [dz, L) and 16-byte
			// aligned: the U phases leave as float4 stores (one lane's
			// outputs are contiguous; phase by phase the lanes would write
			// 4U bytes apart, onto a few LDS banks)
			if ((up == 8 || up == 4) && ibase >= static_cast<int>(A.dz) && ibase + static_cast<int>(up) <= static_cast<int>(L) &&
					(ibase & 3) == 0) {
#pragma unroll
				for (uint32_t h = 0; h < 8; h += 4) {
					if (h >= up) break;
					float o[4];
#pragma unroll
					for (int k = 0; k < 4; ++k) {
						const cf32_t hrow = hg + (h + k) * kFirJ;
						float acc = 0.0f;
#pragma unroll
						for (int j = 0; j < kFirJ; ++j) acc = fmaf(hrow[j], xw[j], acc);
						o[k] = acc;
						part += static_cast<double>(acc);
						xmax = fmaxf(xmax, acc);
						xmin = fminf(xmin, acc);
					}
					*reinterpret_cast<float4*>(Df + ibase + h) = make_float4(o[0], o[1], o[2], o[3]);
				}
				continue;
			}
			for (uint32_t ph = 0; ph < up; ++ph) {
				const cf32_t hrow = hg + ph * kFirJ;
				float acc = 0.0f;
#pragma unroll
				for (int j = 0; j < kFirJ; ++j) acc = fmaf(hrow[j], xw[j], acc);
				const int i = ibase + static_cast<int>(ph);
				if (i >= 0 && i < static_cast<int>(L)) {
					if (static_cast<uint32_t>(i) >= A.dz) part += static_cast<double>(acc);
					xmax = fmaxf(xmax, acc);
					xmin = fminf(xmin, acc);
					Df[i] = acc;
				}
			}
		}
	} else for (uint32_t i = tid; i < L; i += kFftThreads) {
		const float acc = fir_at(A, xs, hs, src, i);
		if (i >= A.dz) part += static_cast<double>(acc);
		xmax = fmaxf(xmax, acc);
		xmin = fminf(xmin, acc);
		Df[i] = acc;
	}
	for (uint32_t i = L + tid; i < N; i += kFftThreads) Df[i] = 0.0f;
	float mean = 0.0f;
	if (A.flags & kRemoveDc) {
		for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
		if ((tid & 63) == 0) red[tid >> 6] = part;
		__syncthreads();
		if (tid == 0) {
			double s = 0.0;
			for (int w = 0; w < kFftThreads / 64; ++w) s += red[w];
			red[0] = s;
		}
		__syncthreads();
		mean = static_cast<float>(red[0] / static_cast<double>(L - A.dz));
		for (uint32_t i = tid; i < L; i += kFftThreads) Df[i] = Df[i] - mean;
	}
	__syncthreads();

	fft_forward(D, A.plan, A.tw);

	// Z = DFT_M(x[2m] + i x[2m+1]) -> conj(C), in place: C[0] = 0,
	// C[k] = i sin(2 pi k/N) Z[k] + cos(2 pi k/N) conj(Z[M-k]), 0 < k < M
	// (pairs k, M-k by one thread; C[M/2] = i Z[M/2]).
	const uint32_t half = M / 2;
	for (uint32_t k = tid; k <= half; k += kFftThreads) {
		if (k == 0) {
			D[0] = make_float2(0.0f, 0.0f);
		} else if (2 * k == M) {
			const float2 a = D[k];
			D[k] = make_float2(-a.y, -a.x);   // conj(i Z)
		} else {
			const float2 a = D[k], b = D[M - k];
			const float2 w = A.wpost[k];      // (cos, -sin)
			const float c = w.x, sn = -w.y;
			const float2 ck = make_float2(fmaf(c, b.x, -sn * a.y), fmaf(sn, a.x, -c * b.y));
			const float2 cm = make_float2(fmaf(-c, a.x, -sn * b.y), fmaf(sn, b.x, c * a.y));
			D[k] = conjf2(ck);
			D[M - k] = conjf2(cm);
		}
	}
	__syncthreads();
	fft_forward(D, A.plan, A.tw);

	// y[2m] + i y[2m+1] = conj(D[m]) / M: the Hilbert transform of x, the
	// imaginary part of the analytic signal, whose real part is x itself
	// (recomputed from the echo and filter kept in LDS).
	const float coef = 2.0f / static_cast<float>(N);
	auto hil = [&](uint32_t i) -> float {
		const float2 v = D[i >> 1];
		return (i & 1u) ? -v.y * coef : v.x * coef;
	};
	const bool mag = (A.flags & kOutMagnitude) != 0;
	float inv = 1.0f;
	if (A.out16) {
		// FP16/BF16 storage over the channel scale s = max(|re|, |im|) (1 if
		// zero), stored value z/s.  max |x_i - mean| is reached at the
		// extremes of x (rounding is monotonic and odd): the workgroup's
		// max x, max -x and max |y| are reduced first.
		float my = 0.0f, nxmin = -xmin;
		for (uint32_t m = tid; 2 * m < L; m += kFftThreads) {
			const float2 v = D[m];
			my = fmaxf(my, fabsf(v.x * coef));
			if (2 * m + 1 < L) my = fmaxf(my, fabsf(-v.y * coef));
		}
		for (int o = 32; o > 0; o >>= 1) {
			my = fmaxf(my, __shfl_xor(my, o, 64));
			xmax = fmaxf(xmax, __shfl_xor(xmax, o, 64));
			nxmin = fmaxf(nxmin, __shfl_xor(nxmin, o, 64));
		}
		if ((tid & 63) == 0) {
			redm[tid >> 6] = my;
			redx[tid >> 6] = xmax;
			redn[tid >> 6] = nxmin;
		}
		__syncthreads();
		float gy = redm[0], gx = redx[0], gn = redn[0];
		for (int w = 1; w < kFftThreads / 64; ++w) {
			gy = fmaxf(gy, redm[w]);
			gx = fmaxf(gx, redx[w]);
			gn = fmaxf(gn, redn[w]);
		}
		const float gm = fmaxf(gy, fmaxf(fabsf(gx - mean), fabsf(-gn - mean)));
		const float sc = (gm > 0.0f) ? gm : 1.0f;
		inv = 1.0f / sc;
		if (tid == 0) A.scale[ch] = sc;
	}
	const bool bf = (A.flags & kOutBf16) != 0;
	float2* dst = A.out + static_cast<uint64_t>(ch) * L;
	float* dmag = A.out_mag + static_cast<uint64_t>(ch) * L;
	uint32_t* row16 = A.out16 + static_cast<uint64_t>(ch) * L;
	// one output sample (re = x_i - mean, im = y_i) in the requested form
	auto put1 = [&](uint32_t i, float re, float im) {
		if (A.out16) row16[i] = bf ? pack_bf16x2(re * inv, im * inv) : pack_f16x2(re * inv, im * inv);
		else if (mag) dmag[i] = sqrtf(re * re + im * im);
		else dst[i] = make_float2(re, im);
	};
	if (up > 1 && J == kFirJ && (up == 8 || up == 4) && A.off % up == 0 && !mag) {
		// the FIR again, lane per input position: the lane's U outputs are
		// contiguous and aligned (off is a multiple of U), written as 16-byte
		// stores of (x, y) pairs
		const cf32_t hg = (cf32_t)A.hpoly;
		const uint32_t q_lo = A.off / up, q_hi = (L - 1 + A.off) / up;
		for (uint32_t q = q_lo + tid; q <= q_hi; q += kFftThreads) {
			float xw[kFirJ];
			const float* xrow = xs + J + q;
#pragma unroll
			for (int j = 0; j < kFirJ; ++j) xw[j] = xrow[-j];
			const uint32_t i0 = q * up - A.off;
			if (i0 + up <= L) {
#pragma unroll
				for (uint32_t h = 0; h < 8; h += 4) {
					if (h >= up) break;
					float re[4], im[4];
#pragma unroll
					for (int k = 0; k < 4; ++k) {
						const cf32_t hrow = hg + (h + k) * kFirJ;
						float acc = 0.0f;
#pragma unroll
						for (int j = 0; j < kFirJ; ++j) acc = fmaf(hrow[j], xw[j], acc);
						re[k] = acc - mean;
					}
					const float4 yv = *reinterpret_cast<const float4*>(D + ((i0 + h) >> 1));   // y pairs of 4 outputs
					im[0] = yv.x * coef;
					im[1] = -yv.y * coef;
					im[2] = yv.z * coef;
					im[3] = -yv.w * coef;
					if (A.out16) {
						uint4 pk;
						if (bf) {
							pk = make_uint4(pack_bf16x2(re[0] * inv, im[0] * inv), pack_bf16x2(re[1] * inv, im[1] * inv),
									pack_bf16x2(re[2] * inv, im[2] * inv), pack_bf16x2(re[3] * inv, im[3] * inv));
						} else {
							pk = make_uint4(pack_f16x2(re[0] * inv, im[0] * inv), pack_f16x2(re[1] * inv, im[1] * inv),
									pack_f16x2(re[2] * inv, im[2] * inv), pack_f16x2(re[3] * inv, im[3] * inv));
						}
						*reinterpret_cast<uint4*>(row16 + i0 + h) = pk;
					} else {
						*reinterpret_cast<float4*>(dst + i0 + h) = make_float4(re[0], im[0], re[1], im[1]);
						*reinterpret_cast<float4*>(dst + i0 + h + 2) = make_float4(re[2], im[2], re[3], im[3]);
					}
				}
			} else {
				for (uint32_t ph = 0; ph < up; ++ph) {
					const uint32_t i = i0 + ph;
					if (i >= L) break;
					const cf32_t hrow = hg + ph * kFirJ;
					float acc = 0.0f;
#pragma unroll
					for (int j = 0; j < kFirJ; ++j) acc = fmaf(hrow[j], xw[j], acc);
					put1(i, acc - mean, hil(i));
				}
			}
		}
		return;
	}
	// any other filter or rate: output by output
	for (uint32_t i = tid; i < L; i += kFftThreads) put1(i, fir_at(A, xs, hs, src, i) - mean, hil(i));
}

// ---------------------------------------------------------------- long records

// Z = DFT_M(x[2m] + i x[2m+1]) -> conj(C) in place, the Hilbert spectrum of
// the fused kernel (C[0] = 0, C[k] = i sin(2 pi k/N) Z[k] + cos(2 pi k/N)
// conj(Z[M-k]), C[M/2] = i Z[M/2]) over rows of M complex values; lane k
// owns the pair (k, M-k), k <= M/2.  wpost[k] = exp(-2 pi i k / N).
__global__ void __launch_bounds__(256) hilbert_pairs_kernel(float2* __restrict__ z, uint32_t M,
		const float2* __restrict__ wpost, uint32_t* __restrict__ ymax)
{
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k == 0 && ymax) ymax[blockIdx.y] = 0u;   // the inverse transform's maximum starts here
	if (k > M / 2) return;
	float2* D = z + static_cast<uint64_t>(blockIdx.y) * M;
	if (k == 0) {
		D[0] = make_float2(0.0f, 0.0f);
	} else if (2 * k == M) {
		const float2 a = D[k];
		D[k] = make_float2(-a.y, -a.x);
	} else {
		const float2 a = D[k], b = D[M - k];
		const float2 w = wpost[k];
		const float c = w.x, sn = -w.y;
		const float2 ck = make_float2(fmaf(c, b.x, -sn * a.y), fmaf(sn, a.x, -c * b.y));
		const float2 cm = make_float2(fmaf(-c, a.x, -sn * b.y), fmaf(sn, b.x, c * a.y));
		D[k] = make_float2(ck.x, -ck.y);
		D[M - k] = make_float2(cm.x, -cm.y);
	}
}

// Analytic tensor rows from x (FIR output after DC removal, [ch][N] floats)
// and F = DFT_M(conj C) ([ch][M] complex): z_i = (x_i, y_i) with
// y[2m] = Re F[m] / M, y[2m+1] = -Im F[m] / M, the first L samples, in the
// storage type (FP16/BF16 over the per-channel scale s = max(|re|, |im|), 1
// if zero).  One workgroup per channel; lane per sample pair.
template<int STORAGE>
__global__ void __launch_bounds__(256) store_xy_kernel(const float* __restrict__ x, const float2* __restrict__ F,
		uint32_t N, uint32_t L, void* __restrict__ tensor, float* __restrict__ scale, uint64_t ch0,
		const float* __restrict__ xmax, const uint32_t* __restrict__ ymax)
{
	const uint32_t ch = blockIdx.x, tid = threadIdx.x, M = N / 2;
	const float2* xr = reinterpret_cast<const float2*>(x + static_cast<uint64_t>(ch) * N);
	const float2* fr = F + static_cast<uint64_t>(ch) * M;
	const uint64_t gch = ch0 + ch;
	const float coef = 2.0f / static_cast<float>(N);
	const uint32_t npair = L / 2;
	const bool odd = (L & 1u) != 0;   // the last sample has no partner
	float inv = 1.0f;
	if constexpr (STORAGE != MISTA_STORAGE_F32) {
		// max |x| of the FIR row and max |Re F|, |Im F| of the transform, as
		// their kernels left them: scaling by coef > 0 rounds monotonically,
		// so max |F| * coef is the largest |y|
		const float gm = fmaxf(xmax[ch], __uint_as_float(ymax[ch]) * coef);
		const float sc = (gm > 0.0f) ? gm : 1.0f;
		inv = 1.0f / sc;
		if (tid == 0) scale[gch] = sc;
	}
	// rows start 16-byte (F32) / 8-byte aligned when L is even; an odd L
	// writes sample by sample
	for (uint32_t m = tid; m < npair; m += 256) {
		const float2 xv = xr[m], f = fr[m];
		const float y0 = f.x * coef, y1 = -f.y * coef;
		if constexpr (STORAGE == MISTA_STORAGE_F32) {
			float2* row = static_cast<float2*>(tensor) + gch * L;
			if (!odd) {
				reinterpret_cast<float4*>(row)[m] = make_float4(xv.x, y0, xv.y, y1);
			} else {
				row[2 * m] = make_float2(xv.x, y0);
				row[2 * m + 1] = make_float2(xv.y, y1);
			}
		} else {
			uint32_t* row = static_cast<uint32_t*>(tensor) + gch * L;
			const uint32_t p0 = (STORAGE == MISTA_STORAGE_F16) ? pack_f16x2(xv.x * inv, y0 * inv) : pack_bf16x2(xv.x * inv, y0 * inv);
			const uint32_t p1 = (STORAGE == MISTA_STORAGE_F16) ? pack_f16x2(xv.y * inv, y1 * inv) : pack_bf16x2(xv.y * inv, y1 * inv);
			if (!odd) {
				reinterpret_cast<uint2*>(row)[m] = make_uint2(p0, p1);
			} else {
				row[2 * m] = p0;
				row[2 * m + 1] = p1;
			}
		}
	}
	if (odd && tid == 0) {
		const float re = xr[npair].x, im = fr[npair].x * coef;
		if constexpr (STORAGE == MISTA_STORAGE_F32) {
			(static_cast<float2*>(tensor) + gch * L)[L - 1] = make_float2(re, im);
		} else if constexpr (STORAGE == MISTA_STORAGE_F16) {
			(static_cast<uint32_t*>(tensor) + gch * L)[L - 1] = pack_f16x2(re * inv, im * inv);
		} else {
			(static_cast<uint32_t*>(tensor) + gch * L)[L - 1] = pack_bf16x2(re * inv, im * inv);
		}
	}
}

// ---------------------------------------------------------------- storage


// One workgroup per channel.  F32: copy.  F16/BF16: per-channel scale
// s = max(|re|, |im|) over the channel (1 if zero), stored value z/s.
template<int STORAGE>
__global__ void __launch_bounds__(256) store_kernel(const float2* __restrict__ z, uint32_t nf, uint32_t L,
		void* __restrict__ tensor, float* __restrict__ scale, uint64_t ch0)
{
	const uint32_t ch = blockIdx.x;
	const uint32_t tid = threadIdx.x;
	const float2* src = z + static_cast<uint64_t>(ch) * nf;
	const uint64_t gch = ch0 + ch;
	if constexpr (STORAGE == MISTA_STORAGE_F32) {
		float2* dst = static_cast<float2*>(tensor) + gch * L;
		for (uint32_t i = tid; i < L; i += 256) dst[i] = src[i];
	} else {
		__shared__ float red[4];
		float m = 0.0f;
		for (uint32_t i = tid; i < L; i += 256) {
			const float2 v = src[i];
			m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
		}
		for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_down(m, o, 64));
		if ((tid & 63) == 0) red[tid >> 6] = m;
		__syncthreads();
		const float mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
		const float s = (mx > 0.0f) ? mx : 1.0f;
		const float inv = 1.0f / s;
		if (tid == 0) scale[gch] = s;
		if constexpr (STORAGE == MISTA_STORAGE_F16) {
			__half2* dst = static_cast<__half2*>(tensor) + gch * L;
			for (uint32_t i = tid; i < L; i += 256) {
				const float2 v = src[i];
				dst[i] = __floats2half2_rn(v.x * inv, v.y * inv);
			}
		} else {
			uint32_t* dst = static_cast<uint32_t*>(tensor) + gch * L;
			for (uint32_t i = tid; i < L; i += 256) {
				const float2 v = src[i];
				dst[i] = pack_bf16x2(v.x * inv, v.y * inv);
			}
		}
	}
}

// Test conversion back to float32 (per-channel scale applied).
template<int STORAGE>
__global__ void __launch_bounds__(256) convert_out_kernel(const void* __restrict__ tensor, const float* __restrict__ scale,
		uint32_t L, float2* __restrict__ dst)
{
	const uint32_t ch = blockIdx.x;
	const uint32_t tid = threadIdx.x;
	const uint64_t base = static_cast<uint64_t>(ch) * L;
	if constexpr (STORAGE == MISTA_STORAGE_F32) {
		const float2* s = static_cast<const float2*>(tensor) + base;
		for (uint32_t i = tid; i < L; i += 256) dst[base + i] = s[i];
	} else if constexpr (STORAGE == MISTA_STORAGE_F16) {
		const __half2* s = static_cast<const __half2*>(tensor) + base;
		const float sc = scale[ch];
		for (uint32_t i = tid; i < L; i += 256) {
			const float2 v = __half22float2(s[i]);
			dst[base + i] = make_float2(v.x * sc, v.y * sc);
		}
	} else {
		const uint32_t* s = static_cast<const uint32_t*>(tensor) + base;
		const float sc = scale[ch];
		for (uint32_t i = tid; i < L; i += 256) {
			const uint32_t b = s[i];
			dst[base + i] = make_float2(__uint_as_float(b << 16) * sc, __uint_as_float(b & 0xFFFF0000u) * sc);
		}
	}
}

void launch_tensor_convert_in(const FrameGeom& g, const float* d_f32, void* d_tensor, float* d_scale, hipStream_t s)
{
	const uint32_t nch = g.ntx * g.nrx;
	const float2* src = reinterpret_cast<const float2*>(d_f32);
	switch (g.storage) {
	case MISTA_STORAGE_F32:
		hipLaunchKernelGGL(store_kernel<MISTA_STORAGE_F32>, dim3(nch), dim3(256), 0, s, src, g.L, g.L, d_tensor, d_scale, 0ull);
		break;
	case MISTA_STORAGE_F16:
		hipLaunchKernelGGL(store_kernel<MISTA_STORAGE_F16>, dim3(nch), dim3(256), 0, s, src, g.L, g.L, d_tensor, d_scale, 0ull);
		break;
	default:
		hipLaunchKernelGGL(store_kernel<MISTA_STORAGE_BF16>, dim3(nch), dim3(256), 0, s, src, g.L, g.L, d_tensor, d_scale, 0ull);
		break;
	}
	HIP_CHECK(hipGetLastError());
}

void launch_tensor_convert_out(const FrameGeom& g, const void* d_tensor, const float* d_scale, float* d_f32, hipStream_t s)
{
	const uint32_t nch = g.ntx * g.nrx;
	float2* dst = reinterpret_cast<float2*>(d_f32);
	switch (g.storage) {
	case MISTA_STORAGE_F32:
		hipLaunchKernelGGL(convert_out_kernel<MISTA_STORAGE_F32>, dim3(nch), dim3(256), 0, s, d_tensor, d_scale, g.L, dst);
		break;
	case MISTA_STORAGE_F16:
		hipLaunchKernelGGL(convert_out_kernel<MISTA_STORAGE_F16>, dim3(nch), dim3(256), 0, s, d_tensor, d_scale, g.L, dst);
		break;
	default:
		hipLaunchKernelGGL(convert_out_kernel<MISTA_STORAGE_BF16>, dim3(nch), dim3(256), 0, s, d_tensor, d_scale, g.L, dst);
		break;
	}
	HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- weights

// w[r][t] = rxApod[r] * txApod[t] * scale[t][r], rx-major with the tx row
// padded to ntx_pad zeros (the tx loop reads consecutive weights);
// txApod/scale optional.
__global__ void __launch_bounds__(256) weights_kernel(uint32_t ntx, uint32_t ntx_pad, uint32_t nrx,
		const float* __restrict__ rx_apod, const float* __restrict__ tx_apod, const float* __restrict__ scale,
		float* __restrict__ w)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= ntx_pad * nrx) return;
	const uint32_t r = i / ntx_pad, t = i % ntx_pad;
	float v = 0.0f;
	if (t < ntx) {
		v = rx_apod[r];
		if (tx_apod) v = v * tx_apod[t];
		if (scale) v = v * scale[static_cast<uint64_t>(t) * nrx + r];
	}
	w[i] = v;
}

void launch_weights(const FrameGeom& g, const float* d_rx_apod, const float* d_tx_apod, const float* d_scale,
		float* d_w, hipStream_t s)
{
	const uint32_t n = g.ntx_pad * g.nrx;
	hipLaunchKernelGGL(weights_kernel, dim3((n + 255) / 256), dim3(256), 0, s, g.ntx, g.ntx_pad, g.nrx, d_rx_apod,
			d_tx_apod, storage_scaled(g.storage) ? d_scale : nullptr, d_w);
	HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- driver

struct PrepPlan {
	uint32_t nf = 0, chunk = 0, tail = 0;
	hipfftHandle r2c = 0, c2c = 0, r2c_tail = 0, c2c_tail = 0;
	bool has_tail = false;
	DeviceBuffer real, spec, z;
	// fused LDS path
	uint32_t fused_nf = 0;
	FftPlanArgs fplan{};
	DeviceBuffer tw, wpost;
	// four-step path (records too long for the fused kernel)
	Fft4Plan* f4 = nullptr;     // size Nf / 2
	DeviceBuffer t4, wpost4;    // transpose buffer; exp(-2 pi i k / Nf), k <= Nf / 4
	DeviceBuffer chmax;         // per channel of a round: max |x| (float), max |F| (float bits)
	uint32_t wpost4_nf = 0;
	~PrepPlan()
	{
		fft4_destroy(f4);
		if (r2c) hipfftDestroy(r2c);
		if (c2c) hipfftDestroy(c2c);
		if (r2c_tail) hipfftDestroy(r2c_tail);
		if (c2c_tail) hipfftDestroy(c2c_tail);
	}
};

PrepState::~PrepState() { delete plan; }

size_t prep_device_bytes(const PrepState& st)
{
	if (!st.plan) return 0;
	const PrepPlan& P = *st.plan;
	// the C2C plan's work area is taken as one more complex chunk
	return P.real.capacity() + P.spec.capacity() + 2 * P.z.capacity() + P.tw.capacity() + P.wpost.capacity() +
			P.t4.capacity() + P.wpost4.capacity() + P.chmax.capacity() + fft4_device_bytes(P.f4);
}

static void make_plans(hipfftHandle* r2c, hipfftHandle* c2c, uint32_t nf, uint32_t batch, hipStream_t s)
{
	int n[1] = {static_cast<int>(nf)};
	FFT_CHECK(hipfftPlanMany(r2c, 1, n, nullptr, 1, static_cast<int>(nf), nullptr, 1,
			static_cast<int>(nf / 2 + 1), HIPFFT_R2C, static_cast<int>(batch)));
	FFT_CHECK(hipfftPlanMany(c2c, 1, n, nullptr, 1, static_cast<int>(nf), nullptr, 1, static_cast<int>(nf),
			HIPFFT_C2C, static_cast<int>(batch)));
	FFT_CHECK(hipfftSetStream(*r2c, s));
	FFT_CHECK(hipfftSetStream(*c2c, s));
}

template<int VPT>
static void launch_fir(uint32_t nch, const FrameGeom& g, const float* raw, const float* hpoly, uint32_t J,
		uint32_t off, float* out, hipStream_t s, float* xmax = nullptr)
{
	const uint32_t right = off / g.up + 2;
	const size_t lds = (static_cast<size_t>(J) + g.ns + right + static_cast<size_t>(g.up) * J) * sizeof(float);
	HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(fir_dc_kernel<VPT>),
			hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
	hipLaunchKernelGGL(fir_dc_kernel<VPT>, dim3(nch), dim3(kFirThreads), lds, s, raw, g.ns, g.up, g.L, g.nf, g.dz_up,
			hpoly, J, off, out, xmax);
}

// LDS bytes of the fused kernel for this frame (0 if it cannot run).
static size_t fused_lds_bytes(const FrameGeom& g, uint32_t J, uint32_t off)
{
	if (g.nf % 2 != 0) return 0;
	const uint32_t M = g.nf / 2;
	if (M > kFftMaxM) return 0;
	size_t b = static_cast<size_t>(M) * sizeof(float2);
	if (g.up > 1) b += (static_cast<size_t>(J) + g.ns + off / g.up + 2 + static_cast<size_t>(g.up) * J) * sizeof(float);
	return b <= 160 * 1024 - 512 ? b : 0;
}

// Radix plan (5s, one 8, 9s, a 6 or 3, 8s, 4, 2 in that order) and twiddles of
// the size-M transforms, computed in double and rounded once.
static void build_fused_plan(PrepPlan& P, uint32_t nf, hipStream_t s)
{
	const uint32_t M = nf / 2;
	FftPlanArgs F{};
	F.M = M;
	uint32_t m = M;
	// Radix order for the LDS banks: pass p writes butterfly b's outputs at
	// stride Ns from b's base, so while Ns is small the lanes of a wave
	// write at a stride of R entries.  An odd R first spreads those writes
	// over the banks (an even one piles them onto a few); one 8 next makes
	// Ns a multiple of 8, whose runs of contiguous outputs collide least in
	// the following passes (counted for M = 17280: 8648 conflict cycles per
	// transform against 16200 for 8, 8, 2, 3, 3, 3, 5).  Pairs of 3s go as
	// one radix-9 pass and a lone 3 with a 2 as radix 6 (M = 17280: 5, 8, 9,
	// 6, 8, five passes instead of seven), each pass costing two barriers
	// and an LDS round trip.
	std::vector<uint32_t> rad;
	while (m % 5 == 0) { rad.push_back(5); m /= 5; }
	if (m % 8 == 0) { rad.push_back(8); m /= 8; }
	while (m % 9 == 0) { rad.push_back(9); m /= 9; }
	if (m % 6 == 0) { rad.push_back(6); m /= 6; }
	while (m % 3 == 0) { rad.push_back(3); m /= 3; }
	while (m % 8 == 0) { rad.push_back(8); m /= 8; }
	while (m % 4 == 0) { rad.push_back(4); m /= 4; }
	while (m % 2 == 0) { rad.push_back(2); m /= 2; }
	if (m != 1 || rad.size() > static_cast<size_t>(kFftMaxPasses))
		MISTA_THROW(MISTA_E_INVALID_PARAM, "FFT size " << nf << " is not 2-3-5 smooth");
	std::vector<float2> tw;
	uint32_t Ns = 1;
	for (size_t p = 0; p < rad.size(); ++p) {
		const uint32_t R = rad[p];
		F.radix[p] = R;
		F.ns[p] = Ns;
		// exact division by multiply-high needs b, Ns < 2^15 (b < M / R)
		F.ns_magic[p] = Ns == 1 ? 0u : static_cast<uint32_t>(((1ull << 32) + Ns - 1) / Ns);
		F.tw_off[p] = static_cast<uint32_t>(tw.size());
		for (uint32_t j = 0; j < Ns; ++j)
			for (uint32_t i = 1; i < R; ++i) {
				const double a = -2.0 * M_PI * static_cast<double>(i) * j / (static_cast<double>(Ns) * R);
				tw.push_back(make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a))));
			}
		Ns *= R;
	}
	F.npass = static_cast<uint32_t>(rad.size());
	tw.push_back(make_float2(0.0f, 0.0f));   // keep the table non-empty
	std::vector<float2> wp(M / 2 + 1);
	for (uint32_t k = 0; k <= M / 2; ++k) {
		const double a = -2.0 * M_PI * static_cast<double>(k) / static_cast<double>(nf);
		wp[k] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
	}
	void* dt = P.tw.reserve(tw.size() * sizeof(float2));
	void* dw = P.wpost.reserve(wp.size() * sizeof(float2));
	HIP_CHECK(hipMemcpyAsync(dt, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice, s));
	HIP_CHECK(hipMemcpyAsync(dw, wp.data(), wp.size() * sizeof(float2), hipMemcpyHostToDevice, s));
	HIP_CHECK(hipStreamSynchronize(s));
	P.fplan = F;
	P.fused_nf = nf;
}

static void prepare_fused(PrepState& st, const FrameGeom& g, const float* d_raw, const float* d_filter_poly,
		uint32_t J, uint32_t off, void* d_tensor, float* d_scale, size_t lds, hipStream_t s)
{
	const uint32_t nch = g.ntx * g.nrx;
	// 1024-channel chunks when the host uploads each chunk just in time, else
	// every channel in one launch
	const uint32_t chunk = st.before_chunk ? std::min<uint32_t>(nch, 1024u) : nch;
	if (!st.plan) st.plan = new PrepPlan();
	PrepPlan& P = *st.plan;
	if (P.fused_nf != g.nf) build_fused_plan(P, g.nf, s);
	const bool scaled = g.storage != MISTA_STORAGE_F32;
	HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(analytic_lds_kernel),
			hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
	AnalyticArgs A{};
	A.hpoly = d_filter_poly;
	A.tw = P.tw.as<const float2>();
	A.wpost = P.wpost.as<const float2>();
	A.ns = g.ns;
	A.up = g.up;
	A.J = J;
	A.off = off;
	A.L = g.L;
	A.N = g.nf;
	A.dz = g.dz_up;
	A.flags = kRemoveDc | (g.storage == MISTA_STORAGE_BF16 ? kOutBf16 : 0u);
	A.out_mag = nullptr;
	A.plan = P.fplan;
	for (uint32_t c0 = 0; c0 < nch; c0 += chunk) {
		const uint32_t n = std::min(chunk, nch - c0);
		if (st.before_chunk) st.before_chunk(c0, n);
		A.raw = d_raw + static_cast<uint64_t>(c0) * g.ns;
		if (scaled) {
			A.out = nullptr;
			A.out16 = static_cast<uint32_t*>(d_tensor) + static_cast<uint64_t>(c0) * g.L;
			A.scale = d_scale + c0;
		} else {
			A.out = static_cast<float2*>(d_tensor) + static_cast<uint64_t>(c0) * g.L;
			A.out16 = nullptr;
			A.scale = nullptr;
		}
		hipLaunchKernelGGL(analytic_lds_kernel, dim3(n), dim3(kFftThreads), lds, s, A);
		HIP_CHECK(hipGetLastError());
	}
}

// FIR x UP + DC removal with the fused kernel's structure (the 27-tap
// design at UP = 4 or 8): a lane per input position q holds the window
// x[q-26..q] in registers and produces the UP outputs m = q UP + p
// (i = m - off), the taps the same for every lane through the scalar cache;
// the lane's QPT x UP outputs wait in registers for the mean.  Each output is
// the same fma chain over j = 0..J-1 as fir_dc_kernel's, which reads both
// operands from LDS per tap.
template<int UP, int QPT>
__global__ void __launch_bounds__(kFirThreads) fir_dc_q_kernel(const float* __restrict__ raw, uint32_t ns,
		uint32_t L, uint32_t nf, uint32_t dz, const float* __restrict__ hpoly, uint32_t off, float* __restrict__ out,
		float* __restrict__ xmax)
{
	extern __shared__ float smem[];
	__shared__ float redm[kFirThreads / 64];
	constexpr uint32_t J = kFirJ;
	const uint32_t ch = blockIdx.x;
	const uint32_t tid = threadIdx.x;
	const uint32_t right = off / UP + 2;
	float* xs = smem;   // [J + ns + right], J zeros on the left
	__shared__ double red[kFirThreads / 64];

	const float* src = raw + static_cast<uint64_t>(ch) * ns;
	for (uint32_t i = tid; i < J + ns + right; i += kFirThreads) {
		const int q = static_cast<int>(i) - static_cast<int>(J);
		xs[i] = (q >= 0 && q < static_cast<int>(ns)) ? src[q] : 0.0f;
	}
	__syncthreads();

	const cf32_t hg = (cf32_t)hpoly;
	const uint32_t q_lo = off / UP, q_hi = (L - 1 + off) / UP;
	float y[QPT][UP];
	double part = 0.0;
#pragma unroll
	for (int k = 0; k < QPT; ++k) {
		const uint32_t q = q_lo + tid + static_cast<uint32_t>(k) * kFirThreads;
		if (q <= q_hi) {
			float xw[J];
			const float* xrow = xs + J + q;
#pragma unroll
			for (uint32_t j = 0; j < J; ++j) xw[j] = xrow[-static_cast<int>(j)];
			const int ibase = static_cast<int>(q * UP) - static_cast<int>(off);
#pragma unroll
			for (int ph = 0; ph < UP; ++ph) {
				const cf32_t hrow = hg + ph * J;
				float acc = 0.0f;
#pragma unroll
				for (uint32_t j = 0; j < J; ++j) acc = fmaf(hrow[j], xw[j], acc);
				const int i = ibase + ph;
				if (i >= static_cast<int>(dz) && i < static_cast<int>(L)) part += static_cast<double>(acc);
				y[k][ph] = acc;
			}
		}
	}
	// block reduction of the DC sum (fixed order: lanes, then waves)
	for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
	if ((tid & 63) == 0) red[tid >> 6] = part;
	__syncthreads();
	if (tid == 0) {
		double sum = 0.0;
		for (int w = 0; w < kFirThreads / 64; ++w) sum += red[w];
		red[0] = sum;
	}
	__syncthreads();
	const float mean = static_cast<float>(red[0] / static_cast<double>(L - dz));

	float* dst = out + static_cast<uint64_t>(ch) * nf;
	float mx = 0.0f;
#pragma unroll
	for (int k = 0; k < QPT; ++k) {
		const uint32_t q = q_lo + tid + static_cast<uint32_t>(k) * kFirThreads;
		if (q > q_hi) continue;
		const int ibase = static_cast<int>(q * UP) - static_cast<int>(off);
		// a lane's UP outputs are contiguous: 16-byte stores where aligned
		if (ibase >= 0 && ibase + UP <= static_cast<int>(L) && (ibase & 3) == 0) {
#pragma unroll
			for (int h = 0; h < UP; h += 4) {
				const float4 o = make_float4(y[k][h] - mean, y[k][h + 1] - mean, y[k][h + 2] - mean, y[k][h + 3] - mean);
				*reinterpret_cast<float4*>(dst + ibase + h) = o;
				mx = fmaxf(mx, fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fmaxf(fabsf(o.z), fabsf(o.w))));
			}
		} else {
#pragma unroll
			for (int ph = 0; ph < UP; ++ph) {
				const int i = ibase + ph;
				if (i >= 0 && i < static_cast<int>(L)) {
					dst[i] = y[k][ph] - mean;
					mx = fmaxf(mx, fabsf(y[k][ph] - mean));
				}
			}
		}
	}
	for (uint32_t i = L + tid; i < nf; i += kFirThreads) dst[i] = 0.0f;
	if (xmax) store_channel_max(mx, xmax + ch, redm);
}

// FIR x U + DC removal of n channels into rows of Nf floats (zero padded).
static void launch_fir_any(uint32_t n, const FrameGeom& g, const float* raw, const float* hpoly, uint32_t J,
		uint32_t off, float* real, hipStream_t s, float* xmax = nullptr)
{
	if (g.up > 1) {
		const size_t lds = (static_cast<size_t>(J) + g.ns + off / g.up + 2 + static_cast<size_t>(g.up) * J) * sizeof(float);
		if (lds > 150 * 1024) MISTA_THROW(MISTA_E_INVALID_PARAM, "channel too long for the LDS FIR: " << g.ns << " samples");
	}
	const uint32_t vpt = (g.L + kFirThreads - 1) / kFirThreads;
	// the lane-per-input kernel: the 27-tap design at U = 4 or 8, every
	// output of a lane in registers (at most 64)
	const uint32_t nq = g.up > 1 ? (g.L - 1 + off) / g.up - off / g.up + 1 : 0;
	const uint32_t qpt = (nq + kFirThreads - 1) / kFirThreads;
	if (J == kFirJ && (g.up == 8 || g.up == 4) && qpt * g.up <= 64) {
		const size_t lds = (static_cast<size_t>(J) + g.ns + off / g.up + 2) * sizeof(float);
		auto go = [&](auto kern) {
			HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
					static_cast<int>(lds)));
			hipLaunchKernelGGL(kern, dim3(n), dim3(kFirThreads), lds, s, raw, g.ns, g.L, g.nf, g.dz_up, hpoly, off, real, xmax);
		};
		if (g.up == 8) {
			if (qpt <= 2) go(fir_dc_q_kernel<8, 2>);
			else if (qpt <= 4) go(fir_dc_q_kernel<8, 4>);
			else go(fir_dc_q_kernel<8, 8>);
		} else {
			if (qpt <= 4) go(fir_dc_q_kernel<4, 4>);
			else if (qpt <= 8) go(fir_dc_q_kernel<4, 8>);
			else go(fir_dc_q_kernel<4, 16>);
		}
		HIP_CHECK(hipGetLastError());
		return;
	}
	if (g.up == 1) {
		hipLaunchKernelGGL(copy_dc_kernel, dim3(n), dim3(kFirThreads), 0, s, raw, g.L, g.nf, g.dz_up, real, xmax);
	} else if (vpt <= 4) {
		launch_fir<4>(n, g, raw, hpoly, J, off, real, s, xmax);
	} else if (vpt <= 8) {
		launch_fir<8>(n, g, raw, hpoly, J, off, real, s, xmax);
	} else if (vpt <= 16) {
		launch_fir<16>(n, g, raw, hpoly, J, off, real, s, xmax);
	} else if (vpt <= 32) {
		launch_fir<32>(n, g, raw, hpoly, J, off, real, s, xmax);
	} else if (vpt <= 64) {
		launch_fir<64>(n, g, raw, hpoly, J, off, real, s, xmax);
	} else {
		MISTA_THROW(MISTA_E_INVALID_PARAM, "signal length " << g.L << " exceeds 65536");
	}
	HIP_CHECK(hipGetLastError());
}

static void launch_store(const FrameGeom& g, uint32_t n, const float2* z, void* d_tensor, float* d_scale, uint64_t c0,
		hipStream_t s)
{
	switch (g.storage) {
	case MISTA_STORAGE_F32:
		hipLaunchKernelGGL(store_kernel<MISTA_STORAGE_F32>, dim3(n), dim3(256), 0, s, z, g.nf, g.L, d_tensor, d_scale, c0);
		break;
	case MISTA_STORAGE_F16:
		hipLaunchKernelGGL(store_kernel<MISTA_STORAGE_F16>, dim3(n), dim3(256), 0, s, z, g.nf, g.L, d_tensor, d_scale, c0);
		break;
	default:
		hipLaunchKernelGGL(store_kernel<MISTA_STORAGE_BF16>, dim3(n), dim3(256), 0, s, z, g.nf, g.L, d_tensor, d_scale, c0);
		break;
	}
	HIP_CHECK(hipGetLastError());
}

// Records too long for the fused kernel: the fused kernel's algorithm with
// its transforms of size M = Nf/2 as four-step FFTs through HBM (fft4.hip):
// FIR x U + DC removal into x, Z = DFT_M(x as M complex pairs), the Hilbert
// spectrum conj(C) in place, F = DFT_M(conj C), then z = (x, y) in the
// storage type.  256 channels per chunk.
static void prepare_four_step(PrepState& st, const FrameGeom& g, const float* d_raw, const float* d_filter_poly,
		uint32_t J, uint32_t off, void* d_tensor, float* d_scale, hipStream_t s)
{
	const uint32_t nch = g.ntx * g.nrx;
	// channels per round of kernels: 256, whose intermediate rows (3 x 69 MB
	// at an FFT size of 67 500) stay in the 256 MB last-level cache between
	// the kernels (1024-channel rounds measured 6 % slower)
	const uint32_t chunk = std::min<uint32_t>(nch, 256u);
	const uint32_t M = g.nf / 2;
	if (!st.plan) st.plan = new PrepPlan();
	PrepPlan& P = *st.plan;
	if (!P.f4 || fft4_size(P.f4) != M) {
		fft4_destroy(P.f4);
		P.f4 = nullptr;
		P.f4 = fft4_plan(M, s);
	}
	if (P.wpost4_nf != g.nf) {
		std::vector<float2> wp(M / 2 + 1);
		for (uint32_t k = 0; k <= M / 2; ++k) {
			const double a = -2.0 * M_PI * static_cast<double>(k) / static_cast<double>(g.nf);
			wp[k] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
		}
		void* d = P.wpost4.reserve(wp.size() * sizeof(float2));
		HIP_CHECK(hipMemcpyAsync(d, wp.data(), wp.size() * sizeof(float2), hipMemcpyHostToDevice, s));
		HIP_CHECK(hipStreamSynchronize(s));
		P.wpost4_nf = g.nf;
	}
	float* real = static_cast<float*>(P.real.reserve(static_cast<size_t>(chunk) * g.nf * sizeof(float)));
	float2* z = static_cast<float2*>(P.z.reserve(static_cast<size_t>(chunk) * M * sizeof(float2)));
	float2* t = static_cast<float2*>(P.t4.reserve(static_cast<size_t>(chunk) * M * sizeof(float2)));
	// FP16/BF16: each channel's largest |x| and |y|, from the FIR and the
	// inverse transform, for the scale (the store then reads each row once)
	const bool scaled = g.storage != MISTA_STORAGE_F32;
	float* xmax = scaled ? static_cast<float*>(P.chmax.reserve(2 * static_cast<size_t>(chunk) * sizeof(float))) : nullptr;
	uint32_t* ymax = scaled ? reinterpret_cast<uint32_t*>(xmax + chunk) : nullptr;
	for (uint32_t c0 = 0; c0 < nch; c0 += chunk) {
		const uint32_t n = std::min(chunk, nch - c0);
		if (st.before_chunk) st.before_chunk(c0, n);
		launch_fir_any(n, g, d_raw + static_cast<uint64_t>(c0) * g.ns, d_filter_poly, J, off, real, s, xmax);
		fft4_exec(*P.f4, reinterpret_cast<const float2*>(real), t, z, n, s);
		hipLaunchKernelGGL(hilbert_pairs_kernel, dim3((M / 2 + 1 + 255) / 256, n), dim3(256), 0, s, z, M,
				P.wpost4.as<const float2>(), ymax);
		HIP_CHECK(hipGetLastError());
		fft4_exec(*P.f4, z, t, z, n, s, ymax, g.L / 2, (g.L & 1u) != 0);
		switch (g.storage) {
		case MISTA_STORAGE_F32:
			hipLaunchKernelGGL(store_xy_kernel<MISTA_STORAGE_F32>, dim3(n), dim3(256), 0, s, real, z, g.nf, g.L, d_tensor,
					d_scale, static_cast<uint64_t>(c0), xmax, ymax);
			break;
		case MISTA_STORAGE_F16:
			hipLaunchKernelGGL(store_xy_kernel<MISTA_STORAGE_F16>, dim3(n), dim3(256), 0, s, real, z, g.nf, g.L, d_tensor,
					d_scale, static_cast<uint64_t>(c0), xmax, ymax);
			break;
		default:
			hipLaunchKernelGGL(store_xy_kernel<MISTA_STORAGE_BF16>, dim3(n), dim3(256), 0, s, real, z, g.nf, g.L, d_tensor,
					d_scale, static_cast<uint64_t>(c0), xmax, ymax);
			break;
		}
		HIP_CHECK(hipGetLastError());
	}
}

void prepare_signals(PrepState& st, const FrameGeom& g, const float* d_raw, const float* d_filter_poly,
		uint32_t J, uint32_t off, void* d_tensor, float* d_scale, hipStream_t s)
{
	const uint32_t nch = g.ntx * g.nrx;
	if (st.method == 3) {
		st.last_fused = false;
		prepare_four_step(st, g, d_raw, d_filter_poly, J, off, d_tensor, d_scale, s);
		return;
	}
	if (st.method != 1) {
		const size_t lds = fused_lds_bytes(g, J, off);
		if (lds) {
			prepare_fused(st, g, d_raw, d_filter_poly, J, off, d_tensor, d_scale, lds, s);
			st.last_fused = true;
			return;
		}
		if (st.method == 2) MISTA_THROW(MISTA_E_INVALID_PARAM, "the fused preparation does not fit this frame (FFT size " << g.nf << ")");
		st.last_fused = false;
		prepare_four_step(st, g, d_raw, d_filter_poly, J, off, d_tensor, d_scale, s);
		return;
	}
	st.last_fused = false;
	const uint32_t chunk = std::min<uint32_t>(nch, 1024u);
	if (!st.plan || !st.plan->r2c || st.plan->nf != g.nf || st.plan->chunk != chunk || st.plan->tail != nch % chunk) {
		delete st.plan;
		st.plan = nullptr;
		auto p = std::make_unique<PrepPlan>();
		p->nf = g.nf;
		p->chunk = chunk;
		p->tail = nch % chunk;
		make_plans(&p->r2c, &p->c2c, g.nf, chunk, s);
		if (p->tail) {
			make_plans(&p->r2c_tail, &p->c2c_tail, g.nf, p->tail, s);
			p->has_tail = true;
		}
		st.plan = p.release();
	}
	PrepPlan& P = *st.plan;
	FFT_CHECK(hipfftSetStream(P.r2c, s));
	FFT_CHECK(hipfftSetStream(P.c2c, s));
	if (P.has_tail) {
		FFT_CHECK(hipfftSetStream(P.r2c_tail, s));
		FFT_CHECK(hipfftSetStream(P.c2c_tail, s));
	}
	float* real = static_cast<float*>(P.real.reserve(static_cast<size_t>(chunk) * g.nf * sizeof(float)));
	float2* spec = static_cast<float2*>(P.spec.reserve(static_cast<size_t>(chunk) * (g.nf / 2 + 1) * sizeof(float2)));
	float2* z = static_cast<float2*>(P.z.reserve(static_cast<size_t>(chunk) * g.nf * sizeof(float2)));

	for (uint32_t c0 = 0; c0 < nch; c0 += chunk) {
		const uint32_t n = std::min(chunk, nch - c0);
		if (st.before_chunk) st.before_chunk(c0, n);
		launch_fir_any(n, g, d_raw + static_cast<uint64_t>(c0) * g.ns, d_filter_poly, J, off, real, s);
		hipfftHandle r2c = (n == chunk) ? P.r2c : P.r2c_tail;
		hipfftHandle c2c = (n == chunk) ? P.c2c : P.c2c_tail;
		FFT_CHECK(hipfftExecR2C(r2c, real, reinterpret_cast<hipfftComplex*>(spec)));
		hipLaunchKernelGGL(hilbert_kernel, dim3((g.nf + 255) / 256, n), dim3(256), 0, s, spec, g.nf, z);
		HIP_CHECK(hipGetLastError());
		FFT_CHECK(hipfftExecC2C(c2c, reinterpret_cast<hipfftComplex*>(z), reinterpret_cast<hipfftComplex*>(z),
				HIPFFT_BACKWARD));
		launch_store(g, n, z, d_tensor, d_scale, c0, s);
	}
}

// ---------------------------------------------------------------- RF path

std::vector<double> rf4x_filter()
{
	// Interpolator4X.h:54-72: upFactor 4, nPeriods 10, beta 5:
	// h = sinc(w) .* kaiser(81, 5), w = linspace(-10, 10, 81)
	const unsigned n = 81;
	const double beta = 5.0;
	const double i0b = std::cyl_bessel_i(0.0, beta);
	std::vector<double> h(n);
	for (unsigned i = 0; i < n; ++i) {
		const double w = -10.0 + 20.0 * i / (n - 1.0);
		const double s = (w == 0.0) ? 1.0 : std::sin(M_PI * w) / (M_PI * w);
		const double r = 2.0 * i / (n - 1.0) - 1.0;
		double a = 1.0 - r * r;
		if (a < 0) a = 0;
		h[i] = s * std::cyl_bessel_i(0.0, beta * std::sqrt(a)) / i0b;
	}
	return h;
}

void prepare_signals_rf(const FrameGeom& g, const float* d_raw, const float* d_filter_poly, uint32_t J, uint32_t off,
		float* d_tensor, hipStream_t s)
{
	const uint32_t nch = g.ntx * g.nrx;
	const uint32_t chunk = std::min<uint32_t>(nch, 1024u);
	const size_t lds = (static_cast<size_t>(J) + g.ns + off / g.up + 2 + static_cast<size_t>(g.up) * J) * sizeof(float);
	if (lds > 150 * 1024) MISTA_THROW(MISTA_E_INVALID_PARAM, "channel too long for the LDS FIR: " << g.ns << " samples");
	const uint32_t vpt = (g.L + kFirThreads - 1) / kFirThreads;
	if (vpt > 64) MISTA_THROW(MISTA_E_INVALID_PARAM, "signal length " << g.L << " exceeds 65536");
	FrameGeom gr = g;
	gr.nf = g.L;   // rows written back to back, no FFT padding
	for (uint32_t c0 = 0; c0 < nch; c0 += chunk) {
		const uint32_t n = std::min(chunk, nch - c0);
		const float* raw = d_raw + static_cast<uint64_t>(c0) * g.ns;
		float* out = d_tensor + static_cast<uint64_t>(c0) * g.L;
		if (vpt <= 4) launch_fir<4>(n, gr, raw, d_filter_poly, J, off, out, s);
		else if (vpt <= 8) launch_fir<8>(n, gr, raw, d_filter_poly, J, off, out, s);
		else if (vpt <= 16) launch_fir<16>(n, gr, raw, d_filter_poly, J, off, out, s);
		else if (vpt <= 32) launch_fir<32>(n, gr, raw, d_filter_poly, J, off, out, s);
		else launch_fir<64>(n, gr, raw, d_filter_poly, J, off, out, s);
		HIP_CHECK(hipGetLastError());
	}
}

void envelope_rows(PrepState& st, float* d_data, uint32_t n1, uint32_t n2, hipStream_t s)
{
	if (n1 == 0 || n2 == 0) return;
	FrameGeom g{};
	g.ns = n2;
	g.up = 1;
	g.L = n2;
	g.nf = hilbert_fft_size(n2);
	const size_t lds = fused_lds_bytes(g, 0, 0);
	if (!lds) MISTA_THROW(MISTA_E_INVALID_PARAM, "rows of " << n2 << " points are too long for the on-chip envelope");
	if (!st.plan) st.plan = new PrepPlan();
	PrepPlan& P = *st.plan;
	if (P.fused_nf != g.nf) build_fused_plan(P, g.nf, s);
	const uint32_t chunk = std::min<uint32_t>(n1, 1024u);
	HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(analytic_lds_kernel),
			hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
	AnalyticArgs A{};
	A.tw = P.tw.as<const float2>();
	A.wpost = P.wpost.as<const float2>();
	A.ns = n2;
	A.up = 1;
	A.L = n2;
	A.N = g.nf;
	A.flags = kOutMagnitude;
	A.plan = P.fplan;
	for (uint32_t r0 = 0; r0 < n1; r0 += chunk) {
		const uint32_t n = std::min(chunk, n1 - r0);
		A.raw = d_data + static_cast<uint64_t>(r0) * n2;
		A.out = nullptr;
		// in place: each sample is read (for the transform and again for the
		// output) by the thread that later writes its magnitude
		A.out_mag = d_data + static_cast<uint64_t>(r0) * n2;
		hipLaunchKernelGGL(analytic_lds_kernel, dim3(n), dim3(kFftThreads), lds, s, A);
		HIP_CHECK(hipGetLastError());
	}
}

}  // namespace mista

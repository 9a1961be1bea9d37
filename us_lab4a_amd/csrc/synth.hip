// Synthetic STA echoes for tests and benchmarks (no reference counterpart:
// the reference's simulator uses a random_device seed and is not
// reproducible).  Point scatterers are stamped as a Gaussian-modulated pulse
// at the round-trip time, accumulated in 64-bit fixed point in LDS so the
// result does not depend on the order of the atomic adds, then white
// Gaussian noise from a counter-based generator is added.
#include "common.h"

#include <algorithm>

#include <cmath>

namespace mista {

namespace {

constexpr double kFix = 4294967296.0;  // 2^32 fixed-point scale

// Counter-based generator: splitmix64 finaliser over seed*C + index.
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t i)
{
	const uint64_t u = mix64(seed * 0xD1B54A32D192ED03ull + i);
	return (static_cast<double>(u >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// One workgroup per channel (tx, rx).
__global__ void __launch_bounds__(256) synth_kernel(const float4* __restrict__ scat, uint32_t nscat,
		const float* __restrict__ txp, const float* __restrict__ rxp, uint32_t nrx, uint32_t ns, double fs,
		double c, float fc_over_fs, float a_env_samples, float half_samples, float* __restrict__ out)
{
	extern __shared__ unsigned long long acc[];
	const uint32_t ch = blockIdx.x;
	const uint32_t tx = ch / nrx, rx = ch % nrx;
	for (uint32_t n = threadIdx.x; n < ns; n += blockDim.x) acc[n] = 0ull;
	__syncthreads();
	const double tx_x = txp[3 * tx], tx_y = txp[3 * tx + 1], tx_z = txp[3 * tx + 2];
	const double rx_x = rxp[3 * rx], rx_y = rxp[3 * rx + 1], rx_z = rxp[3 * rx + 2];
	const float w = 6.28318530717958647692f * fc_over_fs;
	for (uint32_t s = threadIdx.x; s < nscat; s += blockDim.x) {
		const float4 sc = scat[s];
		const double ax = sc.x - tx_x, ay = sc.y - tx_y, az = sc.z - tx_z;
		const double bx = sc.x - rx_x, by = sc.y - rx_y, bz = sc.z - rx_z;
		const double cs = (sqrt(ax * ax + ay * ay + az * az) + sqrt(bx * bx + by * by + bz * bz)) / c * fs;
		int n0 = static_cast<int>(ceil(cs - half_samples));
		int n1 = static_cast<int>(floor(cs + half_samples));
		if (n0 < 0) n0 = 0;
		if (n1 > static_cast<int>(ns) - 1) n1 = static_cast<int>(ns) - 1;
		for (int n = n0; n <= n1; ++n) {
			const float u = static_cast<float>(n - cs);
			const float v = sc.w * expf(-a_env_samples * u * u) * cosf(w * u);
			const long long q = llrint(static_cast<double>(v) * kFix);
			atomicAdd(&acc[n], static_cast<unsigned long long>(q));
		}
	}
	__syncthreads();
	float* dst = out + static_cast<uint64_t>(ch) * ns;
	for (uint32_t n = threadIdx.x; n < ns; n += blockDim.x)
		dst[n] = static_cast<float>(static_cast<double>(static_cast<long long>(acc[n])) * (1.0 / kFix));
}

__global__ void __launch_bounds__(256) noise_kernel(float* __restrict__ x, uint64_t n, float sigma, uint64_t seed)
{
	const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const double u1 = uniform01(seed, 2 * i);
	const double u2 = uniform01(seed, 2 * i + 1);
	const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
	x[i] += static_cast<float>(sigma * g);
}

__global__ void __launch_bounds__(256) sumsq_kernel(const float* __restrict__ x, uint64_t n, double* __restrict__ part)
{
	__shared__ double red[4];
	double s = 0.0;
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * 256) {
		const double v = x[i];
		s += v * v;
	}
	for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
	if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
	__syncthreads();
	if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace

void synth_echoes(const mista_synth_params& p, const float* tx_xyz, const float* rx_xyz, const float* scat,
		float* d_echoes, hipStream_t s)
{
	if (p.num_tx == 0 || p.num_rx == 0 || p.num_samples == 0)
		MISTA_THROW(MISTA_E_INVALID_PARAM, "empty synthetic acquisition");
	if (p.num_samples > 16384) MISTA_THROW(MISTA_E_INVALID_PARAM, "num_samples > 16384 is not supported by the generator");
	if (!(p.sampling_frequency > 0) || !(p.propagation_speed > 0) || !(p.center_frequency > 0) || !(p.bandwidth > 0))
		MISTA_THROW(MISTA_E_INVALID_PARAM, "invalid pulse parameters");
	DeviceBuffer d_tx, d_rx, d_sc;
	const size_t ns_sc = static_cast<size_t>(p.num_scatterers) * 4 * sizeof(float);
	HIP_CHECK(hipMemcpyAsync(d_tx.reserve(p.num_tx * 3 * sizeof(float)), tx_xyz, p.num_tx * 3 * sizeof(float), hipMemcpyHostToDevice, s));
	HIP_CHECK(hipMemcpyAsync(d_rx.reserve(p.num_rx * 3 * sizeof(float)), rx_xyz, p.num_rx * 3 * sizeof(float), hipMemcpyHostToDevice, s));
	if (ns_sc) HIP_CHECK(hipMemcpyAsync(d_sc.reserve(ns_sc), scat, ns_sc, hipMemcpyHostToDevice, s));
	// gausspulse envelope exp(-a t^2), a = (pi fc bw)^2 / (4 ln 2) for -6 dB
	// bandwidth; support where the envelope exceeds -60 dB.
	const double pi = 3.14159265358979323846;
	const double ref = std::pow(10.0, -6.0 / 20.0);
	const double a = -(pi * p.center_frequency * p.bandwidth) * (pi * p.center_frequency * p.bandwidth) / (4.0 * std::log(ref));
	const double tcut = std::sqrt(-std::log(1e-3) / a);
	const double fs = p.sampling_frequency;
	const size_t lds = static_cast<size_t>(p.num_samples) * sizeof(unsigned long long);
	HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(synth_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
			static_cast<int>(lds)));
	hipLaunchKernelGGL(synth_kernel, dim3(p.num_tx * p.num_rx), dim3(256), lds, s, d_sc.as<const float4>(), p.num_scatterers,
			d_tx.as<const float>(), d_rx.as<const float>(), p.num_rx, p.num_samples, fs,
			static_cast<double>(p.propagation_speed), static_cast<float>(p.center_frequency / fs),
			static_cast<float>(a / (fs * fs)), static_cast<float>(tcut * fs), d_echoes);
	HIP_CHECK(hipGetLastError());
	HIP_CHECK(hipStreamSynchronize(s));  // the scatterer/element buffers are freed on return
	if (p.noise_std > 0) {
		const uint64_t n = static_cast<uint64_t>(p.num_tx) * p.num_rx * p.num_samples;
		hipLaunchKernelGGL(noise_kernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, s, d_echoes, n,
				p.noise_std, p.seed);
		HIP_CHECK(hipGetLastError());
	}
}

double sum_squares(const float* d_x, uint64_t n, hipStream_t s)
{
	const uint32_t blocks = 1024;
	DeviceBuffer part;
	double* d_part = static_cast<double*>(part.reserve(blocks * sizeof(double)));
	hipLaunchKernelGGL(sumsq_kernel, dim3(blocks), dim3(256), 0, s, d_x, n, d_part);
	HIP_CHECK(hipGetLastError());
	std::vector<double> h(blocks);
	HIP_CHECK(hipMemcpyAsync(h.data(), d_part, blocks * sizeof(double), hipMemcpyDeviceToHost, s));
	HIP_CHECK(hipStreamSynchronize(s));
	double t = 0.0;
	for (double v : h) t += v;
	return t;
}

// Streaming read of n 16-byte words: every lane reads 8 words per
// iteration (128 B in flight per lane), a grid of 8 workgroups per CU
// strides over the buffer, and one folded word per workgroup is written so
// that nothing is optimised away.  Measures the achievable HBM read rate.
__global__ void __launch_bounds__(256) read_stream_kernel(const uint4* __restrict__ x, uint64_t n, uint32_t* __restrict__ out)
{
	constexpr uint32_t U = 8;
	uint32_t acc = 0;
	const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	for (; i + (U - 1) * stride < n; i += U * stride) {
		uint4 v[U];
#pragma unroll
		for (uint32_t k = 0; k < U; ++k) v[k] = x[i + k * stride];
#pragma unroll
		for (uint32_t k = 0; k < U; ++k) acc ^= v[k].x ^ v[k].y ^ v[k].z ^ v[k].w;
	}
	for (; i < n; i += stride) {
		const uint4 v = x[i];
		acc ^= v.x ^ v.y ^ v.z ^ v.w;
	}
	for (int o = 32; o > 0; o >>= 1) acc ^= static_cast<uint32_t>(__shfl_xor(static_cast<int>(acc), o, 64));
	if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

double read_bandwidth(const void* d_x, uint64_t bytes, int reps, hipStream_t s)
{
	const uint64_t n = bytes / 16;
	if (n == 0 || reps < 1) MISTA_THROW(MISTA_E_INVALID_PARAM, "read_bandwidth: empty buffer or no repetitions.");
	int dev = 0, cus = 0;
	HIP_CHECK(hipGetDevice(&dev));
	HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
	const uint32_t blocks = static_cast<uint32_t>(std::max(1, cus) * 8);
	DeviceBuffer out;
	uint32_t* d_out = static_cast<uint32_t*>(out.reserve(blocks * sizeof(uint32_t)));
	hipEvent_t e0 = nullptr, e1 = nullptr;
	HIP_CHECK(hipEventCreate(&e0));
	HIP_CHECK(hipEventCreate(&e1));
	hipLaunchKernelGGL(read_stream_kernel, dim3(blocks), dim3(256), 0, s, static_cast<const uint4*>(d_x), n, d_out);   // warm-up
	HIP_CHECK(hipEventRecord(e0, s));
	for (int r = 0; r < reps; ++r)
		hipLaunchKernelGGL(read_stream_kernel, dim3(blocks), dim3(256), 0, s, static_cast<const uint4*>(d_x), n, d_out);
	HIP_CHECK(hipEventRecord(e1, s));
	HIP_CHECK(hipGetLastError());
	HIP_CHECK(hipEventSynchronize(e1));
	float ms = 0.0f;
	HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	return static_cast<double>(n) * 16.0 * reps / (ms * 1e-3) / 1e9;
}

void add_noise(float* d_x, uint64_t n, float sigma, uint64_t seed, hipStream_t s)
{
	if (n == 0) return;
	hipLaunchKernelGGL(noise_kernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, s, d_x, n, sigma, seed);
	HIP_CHECK(hipGetLastError());
}

}  // namespace mista

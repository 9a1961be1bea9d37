// Batched forward complex FFTs of 2-3-5-smooth sizes N = N1 * N2 too long
// for one workgroup's LDS (the half-size transforms of the analytic signal of
// records past the fused kernel's limit: 8192 samples x 8 give an FFT size of
// 67 500, transforms of N = 33 750), as the four-step algorithm in two
// passes through HBM, each a batch of short transforms in LDS:
//   A (columns):  for every n2 < N2 the N1-point DFT over n1 of
//                 x[N2 n1 + n2], times exp(-2 pi i n2 k1 / N), stored
//                 T[k1][n2] (row-major);
//   B (rows):     for every k1 < N1 the N2-point DFT over n2 of T[k1][.],
//                 stored X[k1 + N1 k2].
// X[k] = sum_n x[n] exp(-2 pi i n k / N).  A workgroup loads a block of columns (rows) with
// consecutive lanes on consecutive addresses, runs the block's short
// transforms as Stockham passes between two LDS buffers, and stores the block
// the same way.  Twiddles are computed in double on the host and rounded once.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mista {

namespace {

constexpr int kF4Threads = 512;
constexpr uint32_t kF4MaxElems = 4096;   // complex values per block (two LDS buffers: 64 KiB, two workgroups per CU)
constexpr int kF4MaxPasses = 12;

struct F4Short {
	uint32_t n, npass;
	uint32_t radix[kF4MaxPasses];
	uint32_t ns[kF4MaxPasses];       // sub-transform size before the pass
	uint32_t tw_off[kF4MaxPasses];   // the pass's twiddles, from the plan's first
};

__device__ __forceinline__ float2 f4_cmul(float2 a, float2 b)
{
	return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 f4_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 f4_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 f4_mi(float2 a) { return make_float2(a.y, -a.x); }   // -i a

// In-register forward DFT of size R (sign -1).
template<int R> __device__ __forceinline__ void f4_dft(float2* v);
template<> __device__ __forceinline__ void f4_dft<2>(float2* v)
{
	const float2 a = v[0], b = v[1];
	v[0] = f4_add(a, b);
	v[1] = f4_sub(a, b);
}
template<> __device__ __forceinline__ void f4_dft<4>(float2* v)
{
	const float2 s02 = f4_add(v[0], v[2]), d02 = f4_sub(v[0], v[2]);
	const float2 s13 = f4_add(v[1], v[3]), d13 = f4_mi(f4_sub(v[1], v[3]));
	v[0] = f4_add(s02, s13);
	v[2] = f4_sub(s02, s13);
	v[1] = f4_add(d02, d13);
	v[3] = f4_sub(d02, d13);
}
template<> __device__ __forceinline__ void f4_dft<8>(float2* v)
{
	float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
	f4_dft<4>(e);
	f4_dft<4>(o);
	constexpr float c = 0.70710678118654752440f;
	const float2 o1 = make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));
	const float2 o2 = f4_mi(o[2]);
	const float2 o3 = make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));
	v[0] = f4_add(e[0], o[0]);
	v[4] = f4_sub(e[0], o[0]);
	v[1] = f4_add(e[1], o1);
	v[5] = f4_sub(e[1], o1);
	v[2] = f4_add(e[2], o2);
	v[6] = f4_sub(e[2], o2);
	v[3] = f4_add(e[3], o3);
	v[7] = f4_sub(e[3], o3);
}
template<> __device__ __forceinline__ void f4_dft<3>(float2* v)
{
	constexpr float h = 0.86602540378443864676f;
	const float2 s = f4_add(v[1], v[2]);
	const float2 d = f4_sub(v[1], v[2]);
	const float2 t = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
	const float2 u = make_float2(h * d.y, -h * d.x);
	v[0] = f4_add(v[0], s);
	v[1] = f4_add(t, u);
	v[2] = f4_sub(t, u);
}
template<> __device__ __forceinline__ void f4_dft<5>(float2* v)
{
	constexpr float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
	constexpr float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
	const float2 b1 = f4_add(v[1], v[4]), b2 = f4_add(v[2], v[3]);
	const float2 d1 = f4_sub(v[1], v[4]), d2 = f4_sub(v[2], v[3]);
	const float2 t1 = make_float2(v[0].x + c1 * b1.x + c2 * b2.x, v[0].y + c1 * b1.y + c2 * b2.y);
	const float2 t2 = make_float2(v[0].x + c2 * b1.x + c1 * b2.x, v[0].y + c2 * b1.y + c1 * b2.y);
	const float2 u1 = make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
	const float2 u2 = make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
	v[0] = f4_add(v[0], f4_add(b1, b2));
	v[1] = f4_add(t1, f4_mi(u1));
	v[4] = f4_sub(t1, f4_mi(u1));
	v[2] = f4_add(t2, f4_mi(u2));
	v[3] = f4_sub(t2, f4_mi(u2));
}

// One Stockham pass of radix R over `cnt` transforms of size n held back to
// back in src, results to dst.
template<int R>
__device__ __forceinline__ void f4_pass(const float2* __restrict__ src, float2* __restrict__ dst, uint32_t n,
		uint32_t cnt, uint32_t Ns, const float2* __restrict__ tw)
{
	const uint32_t nb = n / R;
	for (uint32_t g = threadIdx.x; g < nb * cnt; g += kF4Threads) {
		const uint32_t t = g / nb, b = g - t * nb;
		const float2* s = src + t * n;
		float2* d = dst + t * n;
		const uint32_t q = b / Ns, j = b - q * Ns;
		float2 v[R];
#pragma unroll
		for (int i = 0; i < R; ++i) v[i] = s[b + i * nb];
		const float2* w = tw + j * (R - 1);
#pragma unroll
		for (int i = 1; i < R; ++i) v[i] = f4_cmul(v[i], w[i - 1]);
		f4_dft<R>(v);
		const uint32_t o = q * Ns * R + j;
#pragma unroll
		for (int i = 0; i < R; ++i) d[o + i * Ns] = v[i];
	}
}

// Forward DFTs of `cnt` transforms in buffer a (b is the second buffer);
// returns the buffer that holds the result.
__device__ float2* f4_short(float2* a, float2* b, uint32_t cnt, const F4Short& P, const float2* __restrict__ tw)
{
	for (uint32_t p = 0; p < P.npass; ++p) {
		const float2* t = tw + P.tw_off[p];
		switch (P.radix[p]) {
		case 8: f4_pass<8>(a, b, P.n, cnt, P.ns[p], t); break;
		case 4: f4_pass<4>(a, b, P.n, cnt, P.ns[p], t); break;
		case 2: f4_pass<2>(a, b, P.n, cnt, P.ns[p], t); break;
		case 3: f4_pass<3>(a, b, P.n, cnt, P.ns[p], t); break;
		default: f4_pass<5>(a, b, P.n, cnt, P.ns[p], t); break;
		}
		__syncthreads();
		float2* x = a;
		a = b;
		b = x;
	}
	return a;
}

struct F4Args {
	uint32_t n1, n2;
	uint64_t n;            // n1 * n2
	uint32_t block;        // columns (pass A) or rows (pass B) per workgroup
	F4Short s1, s2;
	const float2* tw1;     // Stockham twiddles of the N1-point transforms
	const float2* tw2;     // of the N2-point transforms
	const float2* twn;     // exp(-2 pi i k / N), k < N
	// optional (pass B): per transform, the largest |Re X[k]|, |Im X[k]| over
	// k < npair, and |Re X[npair]| when odd, merged into ymax[transform] as
	// the bits of a non-negative float (atomic max; zeroed beforehand)
	uint32_t* ymax;
	uint32_t npair, odd;
};

// Pass A: grid (column blocks, transforms).
__global__ void __launch_bounds__(kF4Threads) f4_cols_kernel(const float2* __restrict__ in, float2* __restrict__ T,
		const F4Args A)
{
	extern __shared__ __attribute__((aligned(16))) float2 f4_lds[];
	const uint32_t n1 = A.n1, n2 = A.n2, cb = A.block;
	const uint32_t c0 = blockIdx.x * cb;
	const uint32_t cnt = min(cb, n2 - c0);
	const uint64_t base = static_cast<uint64_t>(blockIdx.y) * A.n;
	float2* a = f4_lds;
	float2* b = f4_lds + static_cast<size_t>(cb) * n1;
	// column c of the block sits at a[c * n1 .. c * n1 + n1)
	for (uint32_t i = threadIdx.x; i < n1 * cnt; i += kF4Threads) {
		const uint32_t r = i / cnt, c = i - r * cnt;
		a[c * n1 + r] = in[base + static_cast<uint64_t>(r) * n2 + c0 + c];
	}
	__syncthreads();
	const float2* res = f4_short(a, b, cnt, A.s1, A.tw1);
	// twiddle exp(-2 pi i n2 k1 / N) and store T[k1][n2]
	for (uint32_t i = threadIdx.x; i < n1 * cnt; i += kF4Threads) {
		const uint32_t k1 = i / cnt, c = i - k1 * cnt;
		const uint32_t e = (c0 + c) * k1;   // < N2 * N1 = N
		T[base + static_cast<uint64_t>(k1) * n2 + c0 + c] = f4_cmul(res[c * n1 + k1], A.twn[e]);
	}
}

// Pass B: grid (row blocks, transforms).
__global__ void __launch_bounds__(kF4Threads) f4_rows_kernel(const float2* __restrict__ T, float2* __restrict__ X,
		const F4Args A)
{
	extern __shared__ __attribute__((aligned(16))) float2 f4_lds[];
	const uint32_t n1 = A.n1, n2 = A.n2, rb = A.block;
	const uint32_t r0 = blockIdx.x * rb;
	const uint32_t cnt = min(rb, n1 - r0);
	const uint64_t base = static_cast<uint64_t>(blockIdx.y) * A.n;
	float2* a = f4_lds;
	float2* b = f4_lds + static_cast<size_t>(rb) * n2;
	const float2* src = T + base + static_cast<uint64_t>(r0) * n2;
	for (uint32_t i = threadIdx.x; i < n2 * cnt; i += kF4Threads) a[i] = src[i];
	__syncthreads();
	const float2* res = f4_short(a, b, cnt, A.s2, A.tw2);
	// X[k1 + N1 k2], k1 = r0 + r: consecutive lanes on consecutive k1
	float mx = 0.0f;
	for (uint32_t i = threadIdx.x; i < n2 * cnt; i += kF4Threads) {
		const uint32_t k2 = i / cnt, r = i - k2 * cnt;
		const float2 v = res[r * n2 + k2];
		const uint32_t k = k2 * n1 + r0 + r;
		X[base + k] = v;
		if (k < A.npair) mx = fmaxf(mx, fmaxf(fabsf(v.x), fabsf(v.y)));
		else if (A.odd && k == A.npair) mx = fmaxf(mx, fabsf(v.x));
	}
	if (A.ymax) {
		__shared__ float red[kF4Threads / 64];
		for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
		if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
		__syncthreads();
		if (threadIdx.x == 0) {
			for (int w = 1; w < kF4Threads / 64; ++w) mx = fmaxf(mx, red[w]);
			atomicMax(A.ymax + blockIdx.y, __float_as_uint(mx));   // non-negative floats order as their bits
		}
	}
}

// Radix plan and Stockham twiddles of an n-point transform (double, rounded
// once), appended to tw; the plan's offsets count from its first twiddle.
F4Short short_plan(uint32_t n, std::vector<float2>& tw)
{
	F4Short P{};
	P.n = n;
	std::vector<uint32_t> rad;
	uint32_t m = n;
	while (m % 5 == 0) { rad.push_back(5); m /= 5; }
	while (m % 3 == 0) { rad.push_back(3); m /= 3; }
	while (m % 8 == 0) { rad.push_back(8); m /= 8; }
	while (m % 4 == 0) { rad.push_back(4); m /= 4; }
	while (m % 2 == 0) { rad.push_back(2); m /= 2; }
	if (m != 1 || rad.size() > static_cast<size_t>(kF4MaxPasses))
		MISTA_THROW(MISTA_E_INVALID_PARAM, "FFT factor " << n << " is not 2-3-5 smooth");
	uint32_t Ns = 1;
	const size_t base = tw.size();   // offsets are relative to this plan's first twiddle
	for (size_t p = 0; p < rad.size(); ++p) {
		const uint32_t R = rad[p];
		P.radix[p] = R;
		P.ns[p] = Ns;
		P.tw_off[p] = static_cast<uint32_t>(tw.size() - base);
		for (uint32_t j = 0; j < Ns; ++j)
			for (uint32_t i = 1; i < R; ++i) {
				const double ang = -2.0 * M_PI * static_cast<double>(i) * j / (static_cast<double>(Ns) * R);
				tw.push_back(make_float2(static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang))));
			}
		Ns *= R;
	}
	P.npass = static_cast<uint32_t>(rad.size());
	return P;
}

}  // namespace

struct Fft4Plan {
	uint32_t n = 0, n1 = 0, n2 = 0, cb = 0, rb = 0;
	F4Short s1{}, s2{};
	DeviceBuffer tw;        // [tw1 | tw2 | twn]
	size_t off2 = 0, offn = 0;
};

// Factor n into n1 * n2 (both 2-3-5 smooth, each at most 1024, n1 as close to
// sqrt(n) as possible) and build the twiddle tables.
Fft4Plan* fft4_plan(uint32_t n, hipStream_t s)
{
	uint32_t best = 0;
	for (uint32_t d = 2; d <= 1024 && d <= n; ++d) {
		if (n % d != 0 || n / d > 1024 || n / d < 2) continue;
		auto smooth = [](uint32_t v) {
			for (uint32_t p : {2u, 3u, 5u})
				while (v % p == 0) v /= p;
			return v == 1;
		};
		if (!smooth(d) || !smooth(n / d)) continue;
		const double r = std::fabs(std::log(static_cast<double>(d)) - 0.5 * std::log(static_cast<double>(n)));
		if (best == 0 || r < std::fabs(std::log(static_cast<double>(best)) - 0.5 * std::log(static_cast<double>(n))))
			best = d;
	}
	if (best == 0) MISTA_THROW(MISTA_E_INVALID_PARAM, "FFT size " << n << " has no 2-3-5 smooth split into two factors <= 1024");
	auto P = std::make_unique<Fft4Plan>();
	P->n = n;
	P->n1 = best;
	P->n2 = n / best;
	P->cb = std::max(1u, std::min(32u, kF4MaxElems / P->n1));
	P->rb = std::max(1u, std::min(32u, kF4MaxElems / P->n2));
	std::vector<float2> tw;
	P->s1 = short_plan(P->n1, tw);
	P->off2 = tw.size();
	P->s2 = short_plan(P->n2, tw);
	P->offn = tw.size();
	for (uint32_t k = 0; k < n; ++k) {
		const double ang = -2.0 * M_PI * static_cast<double>(k) / static_cast<double>(n);
		tw.push_back(make_float2(static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang))));
	}
	void* d = P->tw.reserve(tw.size() * sizeof(float2));
	HIP_CHECK(hipMemcpyAsync(d, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice, s));
	HIP_CHECK(hipStreamSynchronize(s));
	return P.release();
}

void fft4_destroy(Fft4Plan* p) { delete p; }
uint32_t fft4_size(const Fft4Plan* p) { return p ? p->n : 0u; }
size_t fft4_device_bytes(const Fft4Plan* p) { return p ? p->tw.capacity() : 0u; }

// batch forward transforms of size n: in ([batch][n] complex) -> out
// ([batch][n]) through tmp ([batch][n]); out may alias in.
void fft4_exec(const Fft4Plan& P, const float2* in, float2* tmp, float2* out, uint32_t batch, hipStream_t s,
		uint32_t* ymax, uint32_t npair, bool odd)
{
	F4Args A{};
	A.ymax = nullptr;
	A.n1 = P.n1;
	A.n2 = P.n2;
	A.n = P.n;
	A.s1 = P.s1;
	A.s2 = P.s2;
	const float2* tw = P.tw.as<const float2>();
	A.tw1 = tw;
	A.tw2 = tw + P.off2;
	A.twn = tw + P.offn;
	A.block = P.cb;
	const size_t lds_a = 2 * static_cast<size_t>(P.cb) * P.n1 * sizeof(float2);
	HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(f4_cols_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
			static_cast<int>(lds_a)));
	hipLaunchKernelGGL(f4_cols_kernel, dim3((P.n2 + P.cb - 1) / P.cb, batch), dim3(kF4Threads), lds_a, s, in, tmp, A);
	HIP_CHECK(hipGetLastError());
	A.block = P.rb;
	A.ymax = ymax;
	A.npair = npair;
	A.odd = odd ? 1u : 0u;
	const size_t lds_b = 2 * static_cast<size_t>(P.rb) * P.n2 * sizeof(float2);
	HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(f4_rows_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
			static_cast<int>(lds_b)));
	hipLaunchKernelGGL(f4_rows_kernel, dim3((P.n1 + P.rb - 1) / P.rb, batch), dim3(kF4Threads), lds_b, s, tmp, out, A);
	HIP_CHECK(hipGetLastError());
}

}  // namespace mista

// Delay-and-sum over the analytic tensor: apodized complex accumulation with
// fractional-sample (linear) interpolation, PCF/SCF coherence statistics and
// the fused envelope.  Semantics of VectorialSTAProcessor.h:237-338 and
// processRowColumnSTA[PCF]Kernel (VectorialSTACUDAProcessor.cu:55-186):
//   pos = (offset + d_tx) + d_rx      (samples, FP32)
//   i = trunc(pos), k = pos - i;  pos >= 0 and i + 1 < L, else skipped
//   v = s[i] + k (s[i+1] - s[i]);  sum += v * w[tx][rx]
//   value = |sum| / (Ntx * Nrx), factor = PCF / SCF of the per-rx sums
// w[tx][rx] folds rx apodization, optional tx apodization and the per-channel
// FP16/BF16 storage scale.
#include "common.h"

#include <cmath>

namespace mista {

namespace {

constexpr float kPi = 3.14159265358979323846f;

template<int STORAGE>
__device__ __forceinline__ float2 load_sample(const void* __restrict__ t, uint64_t i)
{
	if constexpr (STORAGE == MISTA_STORAGE_F32) {
		return static_cast<const float2*>(t)[i];
	} else if constexpr (STORAGE == MISTA_STORAGE_F16) {
		return __half22float2(static_cast<const __half2*>(t)[i]);
	} else {
		const uint32_t b = static_cast<const uint32_t*>(t)[i];
		return make_float2(__uint_as_float(b << 16), __uint_as_float(b & 0xFFFF0000u));
	}
}

struct DasArgs {
	const void* tensor;   // [ntx][nrx][L] complex, storage type
	const float* delay;   // [num_elements][npix]
	const float* w;       // [ntx][nrx]
	float* value;         // [npix]
	float* factor;        // [npix] or null
	uint64_t npix;
	uint32_t first_tx, ntx, nrx, L;
	float offset;
	float cf_factor;      // PCF: gamma / sigma0; SCF: p
	float inv_num;        // 1 / (Ntx * Nrx)
};

// Shifted single-pass moments for the two-pass population sigma of
// Statistics.h:34-58 (shift = first value, accumulated in double).
struct Moments {
	double k = 0.0, s1 = 0.0, s2 = 0.0;
	__device__ void add(float x, bool first)
	{
		if (first) k = x;
		const double d = static_cast<double>(x) - k;
		s1 += d;
		s2 += d * d;
	}
	__device__ float sigma(uint32_t n) const
	{
		const double v = (s2 - s1 * s1 / n) / n;
		return static_cast<float>(sqrt(v > 0.0 ? v : 0.0));
	}
};

// One pixel per lane, reference loop structure, global gathers.  CF == NONE:
// tx outer, rx inner into one accumulator (VectorialSTAProcessor.h:312-325).
// CF != NONE: rx outer, tx inner, which yields the same per-rx sums
// (VectorialSTAProcessor.h:269-284) one rx at a time.
template<int STORAGE, int CF>
__global__ void __launch_bounds__(256) das_gather_kernel(const DasArgs a)
{
	const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (p >= a.npix) return;
	const float* dcol = a.delay + p;
	const uint32_t L = a.L;
	const float lim = static_cast<float>(L - 1u);
	float sr = 0.0f, si = 0.0f;
	if constexpr (CF == MISTA_CF_NONE) {
		for (uint32_t t = 0; t < a.ntx; ++t) {
			const float txo = a.offset + dcol[static_cast<uint64_t>(a.first_tx + t) * a.npix];
			const uint64_t base_t = static_cast<uint64_t>(t) * a.nrx;
			for (uint32_t r = 0; r < a.nrx; ++r) {
				const float pos = txo + dcol[static_cast<uint64_t>(r) * a.npix];
				// pos in [0, L-1) <=> i = trunc(pos) satisfies i + 1 < L; the
				// float compare also rejects NaN and values past 2^32
				if (pos >= 0.0f && pos < lim) {
					const uint32_t i = static_cast<uint32_t>(pos);
					{
						const float k = pos - static_cast<float>(i);
						const uint64_t ch = (base_t + r) * L + i;
						const float2 v0 = load_sample<STORAGE>(a.tensor, ch);
						const float2 v1 = load_sample<STORAGE>(a.tensor, ch + 1);
						const float w = a.w[base_t + r];
						sr = fmaf(fmaf(k, v1.x - v0.x, v0.x), w, sr);
						si = fmaf(fmaf(k, v1.y - v0.y, v0.y), w, si);
					}
				}
			}
		}
		a.value[p] = sqrtf(sr * sr + si * si) * a.inv_num;
		if (a.factor) a.factor[p] = 1.0f;
	} else {
		Moments m0, m1;
		int sign_sum = 0;
		for (uint32_t r = 0; r < a.nrx; ++r) {
			const float drx = dcol[static_cast<uint64_t>(r) * a.npix];
			float ar = 0.0f, ai = 0.0f;
			for (uint32_t t = 0; t < a.ntx; ++t) {
				const float pos = (a.offset + dcol[static_cast<uint64_t>(a.first_tx + t) * a.npix]) + drx;
				if (pos >= 0.0f && pos < lim) {
					const uint32_t i = static_cast<uint32_t>(pos);
					{
						const float k = pos - static_cast<float>(i);
						const uint64_t ch = (static_cast<uint64_t>(t) * a.nrx + r) * L + i;
						const float2 v0 = load_sample<STORAGE>(a.tensor, ch);
						const float2 v1 = load_sample<STORAGE>(a.tensor, ch + 1);
						const float w = a.w[static_cast<uint64_t>(t) * a.nrx + r];
						ar = fmaf(fmaf(k, v1.x - v0.x, v0.x), w, ar);
						ai = fmaf(fmaf(k, v1.y - v0.y, v0.y), w, ai);
					}
				}
			}
			sr += ar;
			si += ai;
			if constexpr (CF == MISTA_CF_PCF) {
				// PhaseCoherenceFactor::calculate, CoherenceFactor.h:270-288
				const float phi = atan2f(ai, ar);
				const float aux = phi + ((phi < 0.0f) ? kPi : -kPi);
				m0.add(phi, r == 0);
				m1.add(aux, r == 0);
			} else {
				// SignCoherenceFactor::calculate, CoherenceFactor.h:198-215
				sign_sum += (ar >= 0.0f) ? 1 : -1;
			}
		}
		float f;
		if constexpr (CF == MISTA_CF_PCF) {
			const float sf = fminf(m0.sigma(a.nrx), m1.sigma(a.nrx));
			f = fmaxf(0.0f, 1.0f - a.cf_factor * sf);
		} else {
			const float k = static_cast<float>(sign_sum) / static_cast<float>(a.nrx);
			const float mm = 1.0f - sqrtf(1.0f - k * k);
			f = powf(fabsf(mm), a.cf_factor);
		}
		a.value[p] = sqrtf(sr * sr + si * si) * a.inv_num;
		if (a.factor) a.factor[p] = f;
	}
}

template<int STORAGE>
void launch_gather(const FrameGeom& g, const DasArgs& a, hipStream_t s)
{
	const dim3 grid(static_cast<uint32_t>((g.npix + 255) / 256));
	switch (g.cf_kind) {
	case MISTA_CF_NONE:
		hipLaunchKernelGGL((das_gather_kernel<STORAGE, MISTA_CF_NONE>), grid, dim3(256), 0, s, a);
		break;
	case MISTA_CF_PCF:
		hipLaunchKernelGGL((das_gather_kernel<STORAGE, MISTA_CF_PCF>), grid, dim3(256), 0, s, a);
		break;
	default:
		hipLaunchKernelGGL((das_gather_kernel<STORAGE, MISTA_CF_SCF>), grid, dim3(256), 0, s, a);
		break;
	}
}

}  // namespace

void launch_das(const FrameGeom& g, const DasTuning& t, DasScratch& scratch, const void* d_tensor,
		const float* d_delay, const float* d_w, float* d_value, float* d_factor, hipStream_t s)
{
	(void)t;
	(void)scratch;
	DasArgs a;
	a.tensor = d_tensor;
	a.delay = d_delay;
	a.w = d_w;
	a.value = d_value;
	a.factor = d_factor;
	a.npix = g.npix;
	a.first_tx = g.first_tx;
	a.ntx = g.ntx;
	a.nrx = g.nrx;
	a.L = g.L;
	a.offset = g.offset;
	if (g.cf_kind == MISTA_CF_PCF) {
		// CoherenceFactor.h:243-246: sigma0 = pi/sqrt(3) (stored as float),
		// factor = gamma / sigma0
		const float sigma0 = static_cast<float>(3.14159265358979323846 / std::sqrt(3.0));
		a.cf_factor = g.cf_param / sigma0;
	} else {
		a.cf_factor = g.cf_param;
	}
	a.inv_num = 1.0f / static_cast<float>(static_cast<uint64_t>(g.ntx) * g.nrx);
	switch (g.storage) {
	case MISTA_STORAGE_F32: launch_gather<MISTA_STORAGE_F32>(g, a, s); break;
	case MISTA_STORAGE_F16: launch_gather<MISTA_STORAGE_F16>(g, a, s); break;
	default: launch_gather<MISTA_STORAGE_BF16>(g, a, s); break;
	}
	HIP_CHECK(hipGetLastError());
}

}  // namespace mista

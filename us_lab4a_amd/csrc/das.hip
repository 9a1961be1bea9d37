// Delay-and-sum over the analytic tensor: apodized complex accumulation with
// fractional-sample (linear) interpolation, PCF/SCF coherence statistics and
// the fused envelope.  Semantics of VectorialSTAProcessor.h:237-338 and
// processRowColumnSTA[PCF]Kernel (VectorialSTACUDAProcessor.cu:55-186):
//   pos = (offset + d_tx) + d_rx      (samples, FP32)
//   i = trunc(pos), k = pos - i;  pos >= 0 and i + 1 < L, else skipped
//   v = s[i] + k (s[i+1] - s[i]);  sum += v * w[rx][tx]
//   value = |sum| / (Ntx * Nrx), factor = PCF / SCF of the per-rx sums
// w[rx][tx] folds rx apodization, optional tx apodization and the per-channel
// FP16/BF16 storage scale.
//
// Two kernels:
//  das_gather_kernel  one pixel per lane, global gathers (reference shape;
//                     fallback for any size).
//  das_tile_kernel    a workgroup owns a pixel tile for the whole frame; the
//                     tx delays of each lane's pixel stay in registers; for
//                     every (rx, tx) pair the tile's signal window is copied
//                     HBM -> LDS with direct loads (global_load_lds_dwordx4,
//                     double-buffered), so each sample pair costs one LDS
//                     read.  Window start / length / record-end flag per
//                     (tile, rx, tx) come from a table built once per
//                     geometry.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <vector>

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

// LDS entry type per storage kind and its conversion to FP32 complex.
template<int STORAGE> struct Entry;
template<> struct Entry<MISTA_STORAGE_F32> {
	using T = float2;
	static __device__ __forceinline__ float2 get(T v) { return v; }
};
template<> struct Entry<MISTA_STORAGE_F16> {
	using T = __half2;
	static __device__ __forceinline__ float2 get(T v) { return __half22float2(v); }
};
template<> struct Entry<MISTA_STORAGE_BF16> {
	using T = uint32_t;
	static __device__ __forceinline__ float2 get(T b)
	{
		return make_float2(__uint_as_float(b << 16), __uint_as_float(b & 0xFFFF0000u));
	}
};

struct DasArgs {
	const void* tensor;   // [ntx][nrx][L] complex, storage type
	const float* delay;   // [num_elements][npix]
	const float* w;       // [nrx][ntx]
	float* value;         // [npix]
	float* factor;        // [npix] or null
	uint64_t npix;
	uint32_t first_tx, ntx, nrx, L;
	uint32_t w_stride;    // row stride of w (transmits padded)
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

// Per-rx statistics and the final value/factor of one pixel.
template<int CF>
struct PixelStats {
	float sr = 0.0f, si = 0.0f;
	Moments m0, m1;
	int sign_sum = 0;
	__device__ __forceinline__ void add_rx(float ar, float ai, bool first)
	{
		sr += ar;
		si += ai;
		if constexpr (CF == MISTA_CF_PCF) {
			// PhaseCoherenceFactor::calculate, CoherenceFactor.h:270-288
			const float phi = atan2f(ai, ar);
			const float aux = phi + ((phi < 0.0f) ? kPi : -kPi);
			m0.add(phi, first);
			m1.add(aux, first);
		} else if constexpr (CF == MISTA_CF_SCF) {
			// SignCoherenceFactor::calculate, CoherenceFactor.h:198-215
			sign_sum += (ar >= 0.0f) ? 1 : -1;
		}
	}
	__device__ __forceinline__ void write(const DasArgs& a, uint64_t p) const
	{
		a.value[p] = sqrtf(sr * sr + si * si) * a.inv_num;
		if (!a.factor) return;
		float f = 1.0f;
		if constexpr (CF == MISTA_CF_PCF) {
			const float sf = fminf(m0.sigma(a.nrx), m1.sigma(a.nrx));
			f = fmaxf(0.0f, 1.0f - a.cf_factor * sf);
		} else if constexpr (CF == MISTA_CF_SCF) {
			const float k = static_cast<float>(sign_sum) / static_cast<float>(a.nrx);
			const float mm = 1.0f - sqrtf(1.0f - k * k);
			f = powf(fabsf(mm), a.cf_factor);
		}
		a.factor[p] = f;
	}
};

// ---------------------------------------------------------------- gather

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
	// pos in [0, L-1) <=> i = trunc(pos) satisfies i + 1 < L; the float
	// compare also rejects NaN and values past 2^32
	const float lim = static_cast<float>(L - 1u);
	if constexpr (CF == MISTA_CF_NONE) {
		float sr = 0.0f, si = 0.0f;
		for (uint32_t t = 0; t < a.ntx; ++t) {
			const float txo = a.offset + dcol[static_cast<uint64_t>(a.first_tx + t) * a.npix];
			for (uint32_t r = 0; r < a.nrx; ++r) {
				const float pos = txo + dcol[static_cast<uint64_t>(r) * a.npix];
				if (pos >= 0.0f && pos < lim) {
					const uint32_t i = static_cast<uint32_t>(pos);
					const float k = pos - static_cast<float>(i);
					const uint64_t ch = (static_cast<uint64_t>(t) * a.nrx + r) * L + i;
					const float2 v0 = load_sample<STORAGE>(a.tensor, ch);
					const float2 v1 = load_sample<STORAGE>(a.tensor, ch + 1);
					const float w = a.w[static_cast<uint64_t>(r) * a.w_stride + t];
					sr = fmaf(fmaf(k, v1.x - v0.x, v0.x), w, sr);
					si = fmaf(fmaf(k, v1.y - v0.y, v0.y), w, si);
				}
			}
		}
		a.value[p] = sqrtf(sr * sr + si * si) * a.inv_num;
		if (a.factor) a.factor[p] = 1.0f;
	} else {
		PixelStats<CF> st;
		for (uint32_t r = 0; r < a.nrx; ++r) {
			const float drx = dcol[static_cast<uint64_t>(r) * a.npix];
			float ar = 0.0f, ai = 0.0f;
			for (uint32_t t = 0; t < a.ntx; ++t) {
				const float pos = (a.offset + dcol[static_cast<uint64_t>(a.first_tx + t) * a.npix]) + drx;
				if (pos >= 0.0f && pos < lim) {
					const uint32_t i = static_cast<uint32_t>(pos);
					const float k = pos - static_cast<float>(i);
					const uint64_t ch = (static_cast<uint64_t>(t) * a.nrx + r) * L + i;
					const float2 v0 = load_sample<STORAGE>(a.tensor, ch);
					const float2 v1 = load_sample<STORAGE>(a.tensor, ch + 1);
					const float w = a.w[static_cast<uint64_t>(r) * a.w_stride + t];
					ar = fmaf(fmaf(k, v1.x - v0.x, v0.x), w, ar);
					ai = fmaf(fmaf(k, v1.y - v0.y, v0.y), w, ai);
				}
			}
			st.add_rx(ar, ai, r == 0);
		}
		st.write(a, p);
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

// ---------------------------------------------------------------- tiles

// Tile geometry: TXW columns x TZH rows, one pixel per lane, x fastest within
// a wave (neighbouring lanes read nearby window entries: few LDS conflicts).
// Tiles are numbered row-band major (all tile columns of a band of TZH rows
// are consecutive), so the workgroups resident at one time cover a depth
// band of the whole image width and share signal windows through L2/MALL.
struct TileGeom {
	uint32_t n1, n2, txw, tzh, tiles_x, tiles_z;
	__device__ __forceinline__ void pixel(uint32_t tile, uint32_t lane_id, uint32_t& col, uint32_t& row, bool& valid) const
	{
		const uint32_t trow = tile / tiles_x, tcol = tile % tiles_x;
		const uint32_t c = tcol * txw + lane_id % txw, r = trow * tzh + lane_id / txw;
		valid = c < n1 && r < n2;
		col = min(c, n1 - 1);   // lanes outside the grid shadow an in-tile pixel
		row = min(r, n2 - 1);
	}
};

// Per-tile (min, max) of every element's delay.  One workgroup per tile.
__global__ void __launch_bounds__(256) tile_bounds_kernel(const float* __restrict__ delay, uint64_t npix,
		uint32_t num_elements, TileGeom tg, float2* __restrict__ bounds)
{
	__shared__ float rmin[4], rmax[4];
	const uint32_t tile = blockIdx.x;
	const uint32_t npt = tg.txw * tg.tzh;
	for (uint32_t e = 0; e < num_elements; ++e) {
		float mn = 3.4e38f, mx = -3.4e38f;
		for (uint32_t l = threadIdx.x; l < npt; l += 256) {
			uint32_t col, row;
			bool valid;
			tg.pixel(tile, l, col, row, valid);
			const float d = delay[static_cast<uint64_t>(e) * npix + static_cast<uint64_t>(col) * tg.n2 + row];
			mn = fminf(mn, d);
			mx = fmaxf(mx, d);
		}
		for (int o = 32; o > 0; o >>= 1) {
			mn = fminf(mn, __shfl_down(mn, o, 64));
			mx = fmaxf(mx, __shfl_down(mx, o, 64));
		}
		if ((threadIdx.x & 63) == 0) {
			rmin[threadIdx.x >> 6] = mn;
			rmax[threadIdx.x >> 6] = mx;
		}
		__syncthreads();
		if (threadIdx.x == 0) {
			bounds[static_cast<uint64_t>(tile) * num_elements + e] =
					make_float2(fminf(fminf(rmin[0], rmin[1]), fminf(rmin[2], rmin[3])),
							fmaxf(fmaxf(rmax[0], rmax[1]), fmaxf(rmax[2], rmax[3])));
		}
		__syncthreads();
	}
}

// Window parameters per (tile, rx, tx), packed in a u32:
//   bits  0..20  ws: first sample of the window (floor of the tile's lowest position)
//   bits 21..29  number of SLOT-entry load slots to stage
//   bit  31      record-end flag: some lane may have i + 1 >= L (or pos < 0)
// lo/hi use the same float expression order as the positions,
// (offset + d_tx) + d_rx, so by monotonic rounding every lane's floor(pos)
// lies in [ws, floor(hi)] and floor(pos) + 1 in the staged window.
struct WinParam {
	uint32_t start;   // absolute entry index of the window start (channel * L + ws)
	uint32_t word;    // ws | slots << 21 | edge << 31
};
constexpr uint32_t kWsMask = (1u << 21) - 1;
constexpr uint32_t kEdgeBit = 1u << 31;

__global__ void __launch_bounds__(256) window_params_kernel(const float2* __restrict__ bounds, uint32_t ntiles,
		uint32_t num_elements, uint32_t first_tx, uint32_t ntx, uint32_t ntx_pad, uint32_t nrx, uint32_t L, float offset,
		uint32_t slot, WinParam* __restrict__ params, uint32_t* __restrict__ max_slots)
{
	const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	const uint64_t n = static_cast<uint64_t>(ntiles) * nrx * ntx_pad;
	if (i >= n) return;
	const uint32_t t = static_cast<uint32_t>(i % ntx_pad);
	const uint32_t r = static_cast<uint32_t>((i / ntx_pad) % nrx);
	const uint32_t tile = static_cast<uint32_t>(i / (static_cast<uint64_t>(ntx_pad) * nrx));
	if (t >= ntx) {
		params[i] = WinParam{0u, kEdgeBit};   // padded transmit: nothing staged
		return;
	}
	const float2 bt = bounds[static_cast<uint64_t>(tile) * num_elements + first_tx + t];
	const float2 br = bounds[static_cast<uint64_t>(tile) * num_elements + r];
	const float lo = (offset + bt.x) + br.x;
	const float hi = (offset + bt.y) + br.y;
	const uint32_t chan = t * nrx + r;
	uint32_t word;
	uint32_t ws = 0;
	if (!(lo >= 0.0f) || !(lo < static_cast<float>(L))) {
		// nothing to stage (window past the record end, or negative/NaN):
		// every lane of the pair is excluded by the predicated path
		word = kEdgeBit;
	} else {
		ws = static_cast<uint32_t>(floorf(lo));
		const float fhi = floorf(hi);
		// entries [ws, floor(hi) + 1], never past index L (one past the record)
		int need = (fhi >= static_cast<float>(L)) ? static_cast<int>(L) + 1 - static_cast<int>(ws)
				: static_cast<int>(fhi) + 2 - static_cast<int>(ws);
		if (need > static_cast<int>(L + 1 - ws)) need = static_cast<int>(L + 1 - ws);
		const uint32_t slots = (static_cast<uint32_t>(need) + slot - 1) / slot;
		const bool edge = fhi + 1.0f >= static_cast<float>(L - 1u);
		word = ws | (min(slots, 511u) << 21) | (edge ? kEdgeBit : 0u);
		atomicMax(max_slots, slots);
	}
	params[i] = WinParam{chan * L + ws, word};
}

typedef __attribute__((address_space(1))) const void* gptr_t;
typedef __attribute__((address_space(3))) void* lptr_t;
// Read-only tables (window parameters, weights) through the constant
// address space: uniform indices then become scalar loads.
typedef __attribute__((address_space(4))) const WinParam* cu2_t;
typedef __attribute__((address_space(4))) const float* cf32_t;

struct TileArgs {
	DasArgs d;
	TileGeom tg;
	const WinParam* params;   // [ntiles][nrx][ntx_pad]
	uint32_t wcap;            // LDS window capacity per pair (entries, multiple of SLOT)
	uint32_t ntx_pad;         // transmits padded to the instantiated count
};

template<int STORAGE, int CF, int NTXMAX, int TC, int NT>
__global__ void __launch_bounds__(NT) das_tile_kernel(const TileArgs A)
{
	using E = typename Entry<STORAGE>::T;
	constexpr uint32_t ESZ = sizeof(E);
	constexpr uint32_t EPL = 16 / ESZ;      // entries per 16-byte lane load
	constexpr uint32_t SLOT = 64 * EPL;     // entries per wave load instruction
	constexpr uint32_t NW = NT / 64;
	extern __shared__ __attribute__((aligned(16))) unsigned char smem[];   // [2][TC][wcap] entries

	const DasArgs& a = A.d;
	const uint32_t tile = blockIdx.x;
	const uint32_t tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
	uint32_t col, row;
	bool valid;
	A.tg.pixel(tile, tid, col, row, valid);
	const uint64_t npix = a.npix;
	const uint64_t p = static_cast<uint64_t>(col) * A.tg.n2 + row;
	const uint32_t L = a.L, ntx = a.ntx, nrx = a.nrx;
	const float lim = static_cast<float>(L - 1u);
	const uint32_t wcap = A.wcap;
	const cu2_t prm = (cu2_t)(A.params + static_cast<uint64_t>(tile) * nrx * NTXMAX);
	const unsigned char* tensor = static_cast<const unsigned char*>(a.tensor);

	// tx delays of this lane's pixel, offset folded in: (offset + d_tx)
	// Transmits are padded to NTXMAX: padded ones have weight 0, an empty
	// window and a position far below 0, so they take the excluded path.
	float dtx[NTXMAX];
#pragma unroll
	for (int t = 0; t < NTXMAX; ++t)
		dtx[t] = (static_cast<uint32_t>(t) < ntx) ? a.offset + a.delay[static_cast<uint64_t>(a.first_tx + t) * npix + p] : -1.0e30f;

	// Copy the windows of pairs (r, c*TC .. c*TC+TC-1) into buffer b: wave w
	// issues load slots w, w + NW, ... of every pair; each lane moves 16 bytes.
	auto stage = [&](uint32_t b, uint32_t r, uint32_t c) {
		const cu2_t pr = prm + r * NTXMAX + c * TC;
#pragma unroll
		for (uint32_t j = 0; j < TC; ++j) {
			const uint32_t slots = (pr[j].word >> 21) & 511u;
			const unsigned char* src = tensor + static_cast<uint64_t>(pr[j].start) * ESZ + lane * 16;
			const uint32_t dst = (b * TC + j) * wcap * ESZ;
			for (uint32_t sl = wave; sl < slots; sl += NW) {
				__builtin_amdgcn_global_load_lds((gptr_t)(src + sl * SLOT * ESZ), (lptr_t)(smem + dst + sl * SLOT * ESZ),
						16, 0, 0);
			}
		}
	};

	// Windows of pairs with nothing to stage are read (with weight 0): start
	// from zeros so that no stale bit pattern can be Inf/NaN.
	for (uint32_t i = tid; i < 2 * TC * wcap * ESZ / 16; i += NT)
		reinterpret_cast<uint4*>(smem)[i] = make_uint4(0u, 0u, 0u, 0u);
	__syncthreads();

	PixelStats<CF> st;
	uint32_t buf = 0;
	stage(0, 0, 0);
	for (uint32_t r = 0; r < nrx; ++r) {
		const float drx = a.delay[static_cast<uint64_t>(r) * npix + p];
		float ar = 0.0f, ai = 0.0f;
#pragma clang loop unroll(full)
		for (int c = 0; c < NTXMAX / TC; ++c) {
			__builtin_amdgcn_s_waitcnt(0);   // this wave's copies into `buf` landed
			__syncthreads();                 // everyone's copies landed; `buf ^ 1` is free
			if (c + 1 < NTXMAX / TC) stage(buf ^ 1u, r, c + 1);
			else if (r + 1 < nrx) stage(buf ^ 1u, r + 1, 0);
			const cu2_t pr = prm + r * NTXMAX + c * TC;
			const cf32_t wr = (cf32_t)(a.w + static_cast<uint64_t>(r) * NTXMAX + c * TC);
#pragma unroll
			for (int j = 0; j < TC; ++j) {
				const uint32_t word = pr[j].word;
				// byte offset of entry 0 of the window, minus ws entries
				const int wbase = static_cast<int>((buf * TC + j) * wcap * ESZ);
				const int base = wbase - static_cast<int>((word & kWsMask) * ESZ);
				const float pos = dtx[c * TC + j] + drx;
				const float k = __builtin_amdgcn_fractf(pos);
				float w = wr[j];
				int off;
				if (!(word & kEdgeBit)) {
					// pos >= lo >= 0: truncation is floor; floor(pos) in
					// [ws, floor(hi)] by the window construction
					off = static_cast<int>(pos) * static_cast<int>(ESZ) + base;
				} else {
					// record end, nothing staged or padded tx: keep only
					// 0 <= pos < L-1; excluded lanes read entry 0 with weight 0
					const bool ok = pos >= 0.0f && pos < lim;
					w = ok ? w : 0.0f;
					off = ok ? static_cast<int>(pos) * static_cast<int>(ESZ) + base : wbase;
				}
				const E* e = reinterpret_cast<const E*>(smem + off);
				const float2 v0 = Entry<STORAGE>::get(e[0]);
				const float2 v1 = Entry<STORAGE>::get(e[1]);
				ar = fmaf(fmaf(k, v1.x - v0.x, v0.x), w, ar);
				ai = fmaf(fmaf(k, v1.y - v0.y, v0.y), w, ai);
			}
			buf ^= 1u;
		}
		st.add_rx(ar, ai, r == 0);
	}
	if (valid) st.write(a, p);
}

template<int STORAGE, int CF, int NTXMAX, int NT>
void launch_tile_t(const TileArgs& A, uint32_t ntiles, uint32_t tc, hipStream_t s)
{
	using E = typename Entry<STORAGE>::T;
	const size_t lds = 2ull * tc * A.wcap * sizeof(E);
	auto go = [&](auto kern) {
		HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
				static_cast<int>(lds)));
		hipLaunchKernelGGL(kern, dim3(ntiles), dim3(NT), lds, s, A);
	};
	if (tc == 8) go(das_tile_kernel<STORAGE, CF, NTXMAX, 8, NT>);
	else if (tc == 4) go(das_tile_kernel<STORAGE, CF, NTXMAX, 4, NT>);
	else go(das_tile_kernel<STORAGE, CF, NTXMAX, 2, NT>);
}

template<int STORAGE, int CF, int NT>
void launch_tile_n(const TileArgs& A, uint32_t ntiles, uint32_t tc, hipStream_t s)
{
	if (A.ntx_pad == 32) launch_tile_t<STORAGE, CF, 32, NT>(A, ntiles, tc, s);
	else if (A.ntx_pad == 64) launch_tile_t<STORAGE, CF, 64, NT>(A, ntiles, tc, s);
	else launch_tile_t<STORAGE, CF, 128, NT>(A, ntiles, tc, s);
}

template<int STORAGE, int NT>
void launch_tile_cf(int cf, const TileArgs& A, uint32_t ntiles, uint32_t tc, hipStream_t s)
{
	switch (cf) {
	case MISTA_CF_NONE: launch_tile_n<STORAGE, MISTA_CF_NONE, NT>(A, ntiles, tc, s); break;
	case MISTA_CF_PCF: launch_tile_n<STORAGE, MISTA_CF_PCF, NT>(A, ntiles, tc, s); break;
	default: launch_tile_n<STORAGE, MISTA_CF_SCF, NT>(A, ntiles, tc, s); break;
	}
}

template<int NT>
void launch_tile_st(const FrameGeom& g, const TileArgs& A, uint32_t ntiles, uint32_t tc, hipStream_t s)
{
	switch (g.storage) {
	case MISTA_STORAGE_F32: launch_tile_cf<MISTA_STORAGE_F32, NT>(g.cf_kind, A, ntiles, tc, s); break;
	case MISTA_STORAGE_F16: launch_tile_cf<MISTA_STORAGE_F16, NT>(g.cf_kind, A, ntiles, tc, s); break;
	default: launch_tile_cf<MISTA_STORAGE_BF16, NT>(g.cf_kind, A, ntiles, tc, s); break;
	}
}

}  // namespace

// Tail padding of the analytic tensor (bytes): staged windows may read up to
// one load slot past the last sample of the last channel.
size_t tensor_tail_padding() { return 64 * 16; }

void launch_das(const FrameGeom& g, const DasTuning& t, DasScratch& scratch, const void* d_tensor,
		const float* d_delay, const float* d_w, float* d_value, float* d_factor, hipStream_t s)
{
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
	a.w_stride = g.ntx_pad;
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

	// Tile kernel.  p0: tile shape (0/1: 32x16 with 512 lanes, 2: 16x16 with
	// 256 lanes, 3: 32x8 with 256 lanes); p1: pairs per staged chunk (2/4/8).
	// window start entries are u32 (channel * L + ws): tensor below 2^32 entries
	bool tiled = (t.variant != 1) && g.ntx_pad <= 128 && g.L < (1u << 21) - 64 &&
			static_cast<uint64_t>(g.ntx) * g.nrx * g.L < (1ull << 32) - (1ull << 22);
	if (tiled) {
		uint32_t txw = 32, tzh = 16, nt = 512;
		if (t.p0 == 2) { txw = 16; tzh = 16; nt = 256; }
		else if (t.p0 == 3) { txw = 32; tzh = 8; nt = 256; }
		const uint32_t esz = static_cast<uint32_t>(storage_bytes(g.storage));
		const uint32_t slot = 64 * (16 / esz);
		TileGeom tg{g.n1, g.n2, txw, tzh, (g.n1 + txw - 1) / txw, (g.n2 + tzh - 1) / tzh};
		const uint32_t ntiles = tg.tiles_x * tg.tiles_z;
		const uint64_t key = (static_cast<uint64_t>(txw) << 48) ^ (static_cast<uint64_t>(tzh) << 40) ^
				(static_cast<uint64_t>(esz) << 36) ^ (static_cast<uint64_t>(g.L) << 8) ^ g.first_tx ^
				(static_cast<uint64_t>(g.ntx) << 20) ^ (static_cast<uint64_t>(g.ntx_pad) << 28);
		if (scratch.bounds_version != g.grid_version || scratch.params_key != key ||
				scratch.params_offset != g.offset) {
			float2* b = static_cast<float2*>(scratch.bounds.reserve(static_cast<size_t>(ntiles) * g.num_elements * sizeof(float2)));
			hipLaunchKernelGGL(tile_bounds_kernel, dim3(ntiles), dim3(256), 0, s, d_delay, g.npix, g.num_elements, tg, b);
			HIP_CHECK(hipGetLastError());
			const uint64_t np = static_cast<uint64_t>(ntiles) * g.nrx * g.ntx_pad;
			WinParam* prm = static_cast<WinParam*>(scratch.params.reserve(np * sizeof(WinParam)));
			uint32_t* d_max = static_cast<uint32_t*>(scratch.partial.reserve(256));
			HIP_CHECK(hipMemsetAsync(d_max, 0, sizeof(uint32_t), s));
			hipLaunchKernelGGL(window_params_kernel, dim3(static_cast<uint32_t>((np + 255) / 256)), dim3(256), 0, s, b, ntiles,
					g.num_elements, g.first_tx, g.ntx, g.ntx_pad, g.nrx, g.L, g.offset, slot, prm, d_max);
			HIP_CHECK(hipGetLastError());
			uint32_t max_slots = 0;
			HIP_CHECK(hipMemcpyAsync(&max_slots, d_max, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipStreamSynchronize(s));
			scratch.wcap = std::max(1u, max_slots) * slot;
			scratch.max_slots = max_slots;
			scratch.bounds_version = g.grid_version;
			scratch.params_key = key;
			scratch.params_offset = g.offset;
		}
		uint32_t tc = (t.p1 == 2 || t.p1 == 4 || t.p1 == 8) ? static_cast<uint32_t>(t.p1) : 8u;
		const size_t budget = 150 * 1024;
		while (tc > 2 && 2ull * tc * scratch.wcap * esz > budget) tc /= 2;
		if (scratch.max_slots > 511 || 2ull * tc * scratch.wcap * esz > budget) tiled = false;
		if (tiled) {
			TileArgs A;
			A.d = a;
			A.tg = tg;
			A.params = scratch.params.as<const WinParam>();
			A.wcap = scratch.wcap;
			A.ntx_pad = g.ntx_pad;
			if (nt == 512) launch_tile_st<512>(g, A, ntiles, tc, s);
			else launch_tile_st<256>(g, A, ntiles, tc, s);
			HIP_CHECK(hipGetLastError());
			return;
		}
	}
	switch (g.storage) {
	case MISTA_STORAGE_F32: launch_gather<MISTA_STORAGE_F32>(g, a, s); break;
	case MISTA_STORAGE_F16: launch_gather<MISTA_STORAGE_F16>(g, a, s); break;
	default: launch_gather<MISTA_STORAGE_BF16>(g, a, s); break;
	}
	HIP_CHECK(hipGetLastError());
}

}  // namespace mista

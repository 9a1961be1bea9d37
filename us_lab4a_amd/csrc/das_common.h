// DAS types and the LDS-tiled kernel template, shared by das.hip and the
// per-storage translation units (das_tile_f32.hip, das_tile_f16.hip,
// das_tile_bf16.hip) that instantiate the tiled kernel, so the three storage
// kinds compile in parallel.
#pragma once

#include "common.h"

namespace mista {
namespace das {

constexpr float kPi = 3.14159265358979323846f;

template<int STORAGE>
__device__ __forceinline__ float2 load_sample(const void* __restrict__ t, uint64_t i)
{
	if constexpr (STORAGE == MISTA_STORAGE_F32) {
		return static_cast<const float2*>(t)[i];
	} else if constexpr (STORAGE == MISTA_STORAGE_F16) {
		return __half22float2(static_cast<const __half2*>(t)[i]);
	} else if constexpr (STORAGE == kStorageReal) {
		return make_float2(static_cast<const float*>(t)[i], 0.0f);
	} else {
		const uint32_t b = static_cast<const uint32_t*>(t)[i];
		return make_float2(__uint_as_float(b << 16), __uint_as_float(b & 0xFFFF0000u));
	}
}

// LDS entry type per storage kind and its conversion to FP32 complex.
template<int STORAGE> struct Entry;
// get: one entry as FP32 (re, im).  lerp: v0 + k (v1 - v0) per component
// for the entries e[0], e[1], with FP32 arithmetic: the difference rounded
// once, then one fused multiply-add.
template<> struct Entry<MISTA_STORAGE_F32> {
	using T = float2;
	static __device__ __forceinline__ float2 get(T v) { return v; }
	static __device__ __forceinline__ float2 lerp(const T* e, float k)
	{
		const float2 v0 = e[0], v1 = e[1];
		return make_float2(fmaf(k, v1.x - v0.x, v0.x), fmaf(k, v1.y - v0.y, v0.y));
	}
};
template<> struct Entry<MISTA_STORAGE_F16> {
	using T = __half2;
	static __device__ __forceinline__ float2 get(T v) { return __half22float2(v); }
	static __device__ __forceinline__ float2 lerp(const T* e, float k)
	{
		const T h0 = e[0], h1 = e[1];
		const uint32_t u0 = __builtin_bit_cast(uint32_t, h0), u1 = __builtin_bit_cast(uint32_t, h1);
		// v1 - v0 straight from the FP16 halves: the mixed-precision FMA
		// widens both exactly and rounds 1 * v1 - v0 once to FP32, as the
		// widen-then-subtract form does, in one instruction
		float dx, dy;
		asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[1,0,1]" : "=v"(dx) : "v"(u1), "v"(u0));
		asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel:[1,0,1] op_sel_hi:[1,0,1]" : "=v"(dy) : "v"(u1), "v"(u0));
		const float2 v0 = __half22float2(h0);
		return make_float2(fmaf(k, dx, v0.x), fmaf(k, dy, v0.y));
	}
};
template<> struct Entry<MISTA_STORAGE_BF16> {
	using T = uint32_t;
	static __device__ __forceinline__ float2 get(T b)
	{
		return make_float2(__uint_as_float(b << 16), __uint_as_float(b & 0xFFFF0000u));
	}
	static __device__ __forceinline__ float2 lerp(const T* e, float k)
	{
		const float2 v0 = get(e[0]), v1 = get(e[1]);
		return make_float2(fmaf(k, v1.x - v0.x, v0.x), fmaf(k, v1.y - v0.y, v0.y));
	}
};

template<> struct Entry<kStorageReal> {
	using T = float;
	static __device__ __forceinline__ float2 get(T v) { return make_float2(v, 0.0f); }
	static __device__ __forceinline__ float2 lerp(const T* e, float k)
	{
		const float v0 = e[0], v1 = e[1];
		return make_float2(fmaf(k, v1 - v0, v0), 0.0f);
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
	int signed_value;     // RF samples: value = sum / (Ntx * Nrx), not |sum|
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
		a.value[p] = a.signed_value ? sr * a.inv_num : sqrtf(sr * sr + si * si) * a.inv_num;
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

// Tile geometry: a tw1 (x) by tw2 (y) by tzh (z) block of grid points, one
// point per lane, x fastest within a wave (neighbouring lanes read nearby
// window entries: few LDS conflicts).  A plane has ny = tw2 = 1; a volume is
// laid out as n1 = nx * ny columns (y fastest) by n2 = nz rows.  Tiles are
// numbered depth-band major (all tiles of a band of tzh rows are
// consecutive), so the workgroups resident at one time cover a depth band and
// share signal windows through L2/MALL.
struct TileGeom {
	uint32_t n1, n2, ny, nx;
	uint32_t tw1, tw2, tzh;
	uint32_t tiles_x, tiles_y, tiles_z;
	__device__ __forceinline__ void pixel(uint32_t tile, uint32_t lane_id, uint32_t& col, uint32_t& row, bool& valid) const
	{
		const uint32_t per_band = tiles_x * tiles_y;
		const uint32_t tz = tile / per_band, rem = tile % per_band;
		const uint32_t tx = rem / tiles_y, ty = rem % tiles_y;
		const uint32_t lx = lane_id % tw1, ly = (lane_id / tw1) % tw2, lz = lane_id / (tw1 * tw2);
		const uint32_t ix = tx * tw1 + lx, iy = ty * tw2 + ly, iz = tz * tzh + lz;
		valid = ix < nx && iy < ny && iz < n2;
		// lanes outside the grid shadow an in-tile point
		col = min(ix, nx - 1) * ny + min(iy, ny - 1);
		row = min(iz, n2 - 1);
	}
};

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
	uint32_t nbuf;            // chunk buffers in the LDS ring (>= 2)
};

constexpr uint32_t kTileThreads = 256;

// Direct global -> LDS copies (LDS DMA).  Each lane moves 16 (4) bytes from
// its own address to lds + 16 (4) * lane; lds must be wave-uniform.  Written
// as asm so that the compiler does not treat the asynchronous LDS write as
// aliasing every later LDS read (it would wait for all copies in flight);
// the kernel waits for them itself with s_waitcnt vmcnt.
// m0 is declared clobbered (it is reserved, hence the silenced warning); the
// kernels using these never rely on m0 themselves.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void lds_dma16(const void* g, uint32_t lds)
{
	asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" :: "s"(lds), "v"(g) : "memory", "m0");
}
__device__ __forceinline__ void lds_dma4(const void* g, uint32_t lds)
{
	asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dword %1, off" :: "s"(lds), "v"(g) : "memory", "m0");
}
#pragma clang diagnostic pop

// Wait until at most n (wave-uniform) vector memory operations of this wave
// are outstanding, rounding n down to a power of two: waits at least as long
// as s_waitcnt vmcnt(n), with 7 immediates instead of 64.  The expcnt and
// lgkmcnt fields stay at their maxima (no wait on them).
#define MISTA_VMCNT(n) __builtin_amdgcn_s_waitcnt(((n) & 15) | (7 << 4) | (15 << 8) | (((n) >> 4) << 14))
__device__ __forceinline__ void wait_vmcnt_at_most(uint32_t n)
{
	if (n >= 32) MISTA_VMCNT(32);
	else if (n >= 16) MISTA_VMCNT(16);
	else if (n >= 8) MISTA_VMCNT(8);
	else if (n >= 4) MISTA_VMCNT(4);
	else if (n >= 2) MISTA_VMCNT(2);
	else if (n >= 1) MISTA_VMCNT(1);
	else MISTA_VMCNT(0);
}

// Chunk i of a tile = pairs (rx r, tx c*TC .. c*TC+TC-1), r = i / CPR,
// c = i % CPR.  Chunks stream through a ring of nbuf LDS buffers, each
// holding TC windows of wcap entries plus the tile's rx delays (one float per
// lane, copied with the first chunk of every rx).  Chunk i + nbuf - 1 is
// copied while chunk i is computed, so nbuf - 1 chunks of copies are in
// flight.  Every wave issues the same number of copies per chunk (windows
// narrower than the widest re-copy their first slot into unused space), so
// the number of copies issued after chunk i is known and the wave waits for
// chunk i alone with s_waitcnt vmcnt.
template<int STORAGE, int CF, int NTXMAX, int TC, int NTHREADS, bool MULTI>
__global__ void __launch_bounds__(NTHREADS) __attribute__((amdgpu_waves_per_eu(3, 3)))
das_tile_kernel(const TileArgs A)
{
	using E = typename Entry<STORAGE>::T;
	constexpr uint32_t NT = NTHREADS;       // one tile point per lane
	constexpr uint32_t ESZ = sizeof(E);
	constexpr uint32_t EPL = 16 / ESZ;      // entries per 16-byte lane load
	constexpr uint32_t SLOT = 64 * EPL;     // entries per wave load instruction
	constexpr uint32_t NW = NT / 64;
	constexpr uint32_t CPB = NTXMAX / TC;   // chunks per block of NTXMAX transmits
	extern __shared__ __attribute__((aligned(16))) unsigned char smem[];

	const DasArgs& a = A.d;
	// XCD-aware order: workgroups are dealt round-robin to the 8 XCDs
	// (block b runs on XCD b % 8), so give XCD x one contiguous run of tiles;
	// tiles running at the same time on one XCD then share windows in its L2.
	const uint32_t ntiles = gridDim.x, q8 = ntiles / 8, rm8 = ntiles % 8;
	const uint32_t xcd = blockIdx.x % 8, tile = xcd * q8 + min(xcd, rm8) + blockIdx.x / 8;
	const uint32_t tid = threadIdx.x, lane = tid % 64;
	const uint32_t wave = __builtin_amdgcn_readfirstlane(tid / 64);   // uniform: scalar copy loops
	uint32_t col, row;
	bool valid;
	A.tg.pixel(tile, tid, col, row, valid);
	const uint64_t npix = a.npix;
	const uint64_t p = static_cast<uint64_t>(col) * A.tg.n2 + row;
	const uint32_t L = a.L, ntx = a.ntx, nrx = a.nrx;
	// transmits run in blocks of NTXMAX (MULTI: several, at NTXMAX = 128)
	const uint32_t nblk = MULTI ? A.ntx_pad / NTXMAX : 1u, CPR = nblk * CPB;   // chunks per rx
	const float lim = static_cast<float>(L - 1u);
	const uint32_t wcap = A.wcap, NB = A.nbuf;
	const uint32_t S = wcap / SLOT;                   // slots of the widest window
	const uint32_t QW = TC * S / NW;                  // window copies per wave per chunk
	const uint32_t win_bytes = TC * wcap * ESZ;
	const uint32_t buf_bytes = win_bytes + NT * 4;    // windows + rx delays
	const cu2_t prm = (cu2_t)(A.params + static_cast<uint64_t>(tile) * nrx * A.ntx_pad);
	const unsigned char* tensor = static_cast<const unsigned char*>(a.tensor);
	const uint32_t lds0 = static_cast<uint32_t>(reinterpret_cast<size_t>((__attribute__((address_space(3))) void*)smem));

	// tx delays of this lane's point for transmit block blk, offset folded
	// in: (offset + d_tx).  Transmits are padded to whole blocks: padded ones
	// have weight 0, an empty window and a position far below 0, so they
	// take the excluded path.
	float dtx[NTXMAX];
	auto load_dtx = [&](uint32_t blk) {
#pragma unroll
		for (int t = 0; t < NTXMAX; ++t) {
			const uint32_t tx = blk * NTXMAX + static_cast<uint32_t>(t);
			const uint32_t tt = min(tx, ntx - 1);   // unconditional load, no branch
			const float d = a.delay[static_cast<uint64_t>(a.first_tx + tt) * npix + p];
			dtx[t] = (tx < ntx) ? a.offset + d : -1.0e30f;
		}
	};
	load_dtx(0);

	// Issue the copies of chunk i into buffer b, given its window parameters
	// pp.  Slot sl of pair j is copy number j * S + sl of the chunk and goes
	// to wave (j * S + sl) % NW.
	const uint32_t nchunks = nrx * CPR;
	auto issue = [&](uint32_t i, uint32_t b, const WinParam* pp) {
		const uint32_t r = i / CPR, c = i % CPR;   // c == 0: first chunk of rx r
		const uint32_t bbase = lds0 + b * buf_bytes;
#pragma unroll
		for (uint32_t j = 0; j < TC; ++j) {
			const uint32_t slots = (pp[j].word >> 21) & 511u;
			const unsigned char* src = tensor + static_cast<uint64_t>(pp[j].start) * ESZ + lane * 16;
			const uint32_t dst = bbase + j * wcap * ESZ;
			for (uint32_t sl = (wave - j * S) & (NW - 1); sl < S; sl += NW) {
				const uint32_t from = sl < slots ? sl : 0u;
				lds_dma16(src + from * SLOT * ESZ, __builtin_amdgcn_readfirstlane(dst + sl * SLOT * ESZ));
			}
		}
		if (c == 0)
			lds_dma4(a.delay + static_cast<uint64_t>(r) * npix + p, __builtin_amdgcn_readfirstlane(bbase + win_bytes + wave * 256));
	};
	// Windows of pairs with nothing to stage are read (with weight 0): start
	// from zeros so that no stale bit pattern can be Inf/NaN.
	for (uint32_t i = tid; i < NB * buf_bytes / 16; i += NT)
		reinterpret_cast<uint4*>(smem)[i] = make_uint4(0u, 0u, 0u, 0u);
	__syncthreads();

	// Window parameters of the next chunk to issue, loaded one chunk ahead:
	// their scalar loads miss the caches (each tile streams its own table),
	// and loading them at the issue itself would add that latency to the
	// copies' on every chunk.
	WinParam nx[TC];
	auto load_params = [&](uint32_t k) {
		const cu2_t pr = prm + (k / CPR) * A.ntx_pad + (k % CPR) * TC;
#pragma unroll
		for (uint32_t j = 0; j < TC; ++j) {
			nx[j].start = pr[j].start;
			nx[j].word = pr[j].word;
		}
	};
	const uint32_t pre = min(NB - 1, nchunks);
	for (uint32_t k = 0; k < pre; ++k) {
		load_params(k);
		issue(k, k, nx);
	}
	if (pre < nchunks) load_params(pre);

	PixelStats<CF> st;
	uint32_t i = 0, b = 0;
	for (uint32_t r = 0; r < nrx; ++r) {
		float drx = 0.0f, ar = 0.0f, ai = 0.0f;
		for (uint32_t blk = 0; blk < nblk; ++blk) {
			if (MULTI) load_dtx(blk);   // waited for at its first use
#pragma clang loop unroll(full)
			for (int c = 0; c < static_cast<int>(CPB); ++c, ++i) {
				// copies issued after chunk i: chunks i+1 .. min(i+NB-2, last),
				// QW each plus one for every chunk that starts an rx
				const uint32_t hi = min(i + NB - 2, nchunks - 1);
				const uint32_t after = (hi > i) ? (hi - i) * QW + (hi / CPR - i / CPR) : 0u;
				wait_vmcnt_at_most(after);         // this wave's copies of chunk i landed
				__syncthreads();                   // everyone's did; buffer of chunk i-1 is free
				if (i + NB - 1 < nchunks) issue(i + NB - 1, (b + NB - 1) % NB, nx);
				const uint32_t bbase = b * buf_bytes;
				if (c == 0 && blk == 0) drx = *reinterpret_cast<const float*>(smem + bbase + win_bytes + tid * 4);
				// Register fences: this chunk's positions are computed after
				// the barrier, and its accumulation is finished before the
				// next one, so only one chunk's temporaries are ever live.
				asm volatile("" : "+v"(drx));
				const uint32_t t0 = blk * NTXMAX + c * TC;   // first transmit of the chunk
				const cu2_t pr = prm + r * A.ntx_pad + t0;
				const cf32_t wr = (cf32_t)(a.w + static_cast<uint64_t>(r) * a.w_stride + t0);
#pragma unroll
				for (int j = 0; j < TC; ++j) {
					// One form for every pair: lanes outside 0 <= pos < L-1 get
					// weight 0; the entry index floor(pos) - ws is clamped,
					// as an unsigned value, to the window, which changes
					// nothing for in-range lanes (it lies in [0, need - 2] by
					// the window construction) and keeps excluded lanes'
					// reads on staged (finite) entries, whose products with
					// weight 0 leave the sums as they are (a sum is never -0).
					const uint32_t ws = pr[j].word & kWsMask;
					const float pos = dtx[c * TC + j] + drx;
					const float k = __builtin_amdgcn_fractf(pos);
					// 0 <= pos < lim as one unsigned compare of the bit
					// patterns: positions are never -0 (delays are >= +0), so
					// non-negative ones order like their bits, and negative
					// values and NaN compare above lim's
					const bool ok = __float_as_uint(pos) < __float_as_uint(lim);
					// load the weight unconditionally (a select, not a branch
					// around a scalar load)
					float wv = wr[j];
					asm volatile("" : "+s"(wv));
					const float w = ok ? wv : 0.0f;
					const float pc = __builtin_amdgcn_fmed3f(pos, 0.0f, lim);   // NaN-free, no int overflow
					const uint32_t rel = min(static_cast<uint32_t>(static_cast<int>(pc)) - ws, wcap - 2u);
					uint32_t sb = bbase + j * wcap * ESZ;   // uniform: one shift-add per lane
					asm("" : "+s"(sb));
					const E* e = reinterpret_cast<const E*>(smem + (rel * ESZ + sb));
					const float2 v = Entry<STORAGE>::lerp(e, k);
					ar = fmaf(v.x, w, ar);
					ai = fmaf(v.y, w, ai);
				}
				asm volatile("" : "+v"(ar), "+v"(ai));
				// after this chunk's LDS reads (a pending scalar load would
				// make every LDS wait a full drain); the chunk index passes
				// through a fence so the loads stay here
				uint32_t nk = i + NB;
				asm volatile("" : "+s"(nk));
				if (nk < nchunks) load_params(nk);
				b = (b + 1 == NB) ? 0u : b + 1;
			}
		}
		st.add_rx(ar, ai, r == 0);
	}
	MISTA_VMCNT(0);
	if (valid) st.write(a, p);
}

// LDS bytes of one chunk buffer: TC windows of wcap entries + the rx delays.
template<int STORAGE>
inline size_t tile_buffer_bytes(uint32_t tc, uint32_t wcap, uint32_t nt)
{
	return static_cast<size_t>(tc) * wcap * sizeof(typename Entry<STORAGE>::T) + nt * 4;
}

template<int STORAGE, int CF, int NTXMAX>
void launch_tile_t(const TileArgs& A, uint32_t ntiles, uint32_t tc, hipStream_t s)
{
	const uint32_t nt = A.tg.tw1 * A.tg.tw2 * A.tg.tzh;   // 256 or 512 (512: tc = 8)
	const size_t lds = A.nbuf * tile_buffer_bytes<STORAGE>(tc, A.wcap, nt);
	auto go = [&](auto kern, uint32_t threads) {
		HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
				static_cast<int>(lds)));
		hipLaunchKernelGGL(kern, dim3(ntiles), dim3(threads), lds, s, A);
	};
	if constexpr (NTXMAX == 128) {
		if (A.ntx_pad > NTXMAX) {   // several transmit blocks: tc = 8, nt = 256
			go(das_tile_kernel<STORAGE, CF, NTXMAX, 8, 256, true>, 256);
			return;
		}
	}
	if (nt == 512) go(das_tile_kernel<STORAGE, CF, NTXMAX, 8, 512, false>, 512);
	else if (tc == 8) go(das_tile_kernel<STORAGE, CF, NTXMAX, 8, 256, false>, 256);
	else go(das_tile_kernel<STORAGE, CF, NTXMAX, 4, 256, false>, 256);
}

template<int STORAGE, int CF>
void launch_tile_n(const TileArgs& A, uint32_t ntiles, uint32_t tc, hipStream_t s)
{
	if (A.ntx_pad == 32) launch_tile_t<STORAGE, CF, 32>(A, ntiles, tc, s);
	else if (A.ntx_pad == 64) launch_tile_t<STORAGE, CF, 64>(A, ntiles, tc, s);
	else launch_tile_t<STORAGE, CF, 128>(A, ntiles, tc, s);
}

// Tiled launch for one storage kind (instantiated in das_tile_<storage>.hip).
template<int STORAGE>
void launch_tile_storage(int cf, const TileArgs& A, uint32_t ntiles, uint32_t tc, hipStream_t s)
{
	switch (cf) {
	case MISTA_CF_NONE: launch_tile_n<STORAGE, MISTA_CF_NONE>(A, ntiles, tc, s); break;
	case MISTA_CF_PCF: launch_tile_n<STORAGE, MISTA_CF_PCF>(A, ntiles, tc, s); break;
	default: launch_tile_n<STORAGE, MISTA_CF_SCF>(A, ntiles, tc, s); break;
	}
}

}  // namespace das
}  // namespace mista

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
// lerp2: the same from the raw words of the two entries as one LDS read
// returns them (LdsPair).
template<uint32_t ESZ> struct LdsPair;
template<> struct LdsPair<8> { using T = float4; };   // two 8-byte entries
template<> struct LdsPair<4> { using T = uint2; };    // two 4-byte entries

template<> struct Entry<MISTA_STORAGE_F32> {
	using T = float2;
	static __device__ __forceinline__ float2 get(T v) { return v; }
	static __device__ __forceinline__ float2 lerp2(float4 e, float k)
	{
		return make_float2(fmaf(k, e.z - e.x, e.x), fmaf(k, e.w - e.y, e.y));
	}
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
	static __device__ __forceinline__ float2 lerp2(uint2 e, float k)
	{
		T h[2] = {__builtin_bit_cast(T, e.x), __builtin_bit_cast(T, e.y)};
		return lerp(h, k);
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
	static __device__ __forceinline__ float2 lerp2(uint2 e, float k)
	{
		const T h[2] = {e.x, e.y};
		return lerp(h, k);
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
	static __device__ __forceinline__ float2 lerp2(uint2 e, float k)
	{
		const T h[2] = {__uint_as_float(e.x), __uint_as_float(e.y)};
		return lerp(h, k);
	}
};

struct DasArgs {
	const void* tensor;   // [ntx][nrx][L] complex, storage type
	const float* delay;   // [num_elements][npix]
	const float* w;       // [nrx][ntx]
	float* value;         // [npix]
	float* factor;        // [npix] or null
	float2* partial;      // [nrx][npix] per-rx sums between transmit blocks (tiled, more than 128 transmits)
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
		s2 = fma(d, d, s2);
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
	// Multiply-adds are written out (fmaf) wherever the compiler could
	// otherwise contract differently in the gather and the tiled kernel.
	__device__ __forceinline__ void write(const DasArgs& a, uint64_t p) const
	{
		a.value[p] = magnitude(a, sr, si);
		if (!a.factor) return;
		float f = 1.0f;
		if constexpr (CF == MISTA_CF_PCF) {
			const float sf = fminf(m0.sigma(a.nrx), m1.sigma(a.nrx));
			f = fmaxf(0.0f, fmaf(-a.cf_factor, sf, 1.0f));
		} else if constexpr (CF == MISTA_CF_SCF) {
			const float k = static_cast<float>(sign_sum) / static_cast<float>(a.nrx);
			const float mm = 1.0f - sqrtf(fmaf(-k, k, 1.0f));
			f = powf(fabsf(mm), a.cf_factor);
		}
		a.factor[p] = f;
	}
	// value = |sum| / (Ntx * Nrx), or the signed sum for real samples
	static __device__ __forceinline__ float magnitude(const DasArgs& a, float re, float im)
	{
		return a.signed_value ? re * a.inv_num : sqrtf(fmaf(re, re, im * im)) * a.inv_num;
	}
};

// Per-rx statistics of the tiled kernel: PixelStats with the PCF moments (4
// doubles per lane) kept in LDS at m.  They change once per rx, and in
// registers they would push the kernel past its register budget (128
// registers hold the lane's transmit delays).  The arithmetic is
// PixelStats's, so both kernels give the same bits.
template<int CF>
struct TileStats {
	float sr = 0.0f, si = 0.0f;
	float k0 = 0.0f, k1 = 0.0f;   // the moments' shifts (first phase, first aux)
	int sign_sum = 0;
	double* m = nullptr;          // LDS: s1 and s2 of the phases, then of the aux phases
	__device__ __forceinline__ void add_rx(float ar, float ai, bool first)
	{
		sr += ar;
		si += ai;
		if constexpr (CF == MISTA_CF_PCF) {
			const float phi = atan2f(ai, ar);
			const float aux = phi + ((phi < 0.0f) ? kPi : -kPi);
			if (first) {
				k0 = phi;
				k1 = aux;
			}
			Moments m0, m1;
			m0.k = k0;
			m1.k = k1;
			if (!first) {
				m0.s1 = m[0];
				m0.s2 = m[1];
				m1.s1 = m[2];
				m1.s2 = m[3];
			}
			m0.add(phi, false);
			m1.add(aux, false);
			m[0] = m0.s1;
			m[1] = m0.s2;
			m[2] = m1.s1;
			m[3] = m1.s2;
		} else if constexpr (CF == MISTA_CF_SCF) {
			sign_sum += (ar >= 0.0f) ? 1 : -1;
		}
	}
	__device__ __forceinline__ void write(const DasArgs& a, uint64_t p) const
	{
		PixelStats<CF> ps;
		ps.sr = sr;
		ps.si = si;
		ps.sign_sum = sign_sum;
		if constexpr (CF == MISTA_CF_PCF) {
			ps.m0.k = k0;
			ps.m1.k = k1;
			ps.m0.s1 = m[0];
			ps.m0.s2 = m[1];
			ps.m1.s1 = m[2];
			ps.m1.s2 = m[3];
		}
		ps.write(a, p);
	}
};
// LDS bytes of the tiled kernel's per-lane statistics
inline size_t tile_stats_bytes(int cf, uint32_t nt) { return cf == MISTA_CF_PCF ? static_cast<size_t>(nt) * 32 : 0; }

// Tile geometry: a tw1 (x) by tw2 (y) by tzh (z) block of grid points, one
// point per lane.  A plane has ny = tw2 = 1; a volume is laid out as
// n1 = nx * ny columns (y fastest) by n2 = nz rows.  In a plane tile each wave
// takes an 8 x 8 block of points (x fastest within the block): the lanes of
// one LDS read pass then sit close together in both directions, which spreads
// their window entries over more banks than a 16 x 4 strip.  Volume tiles
// keep x fastest, then y, then z.  Tiles are numbered depth-band major (all
// tiles of a band of tzh rows are consecutive), so the workgroups resident
// at one time cover a depth band and share signal windows through L2/MALL.
struct TileGeom {
	uint32_t n1, n2, ny, nx;
	uint32_t tw1, tw2, tzh;
	uint32_t tiles_x, tiles_y, tiles_z;
	__device__ __forceinline__ void pixel(uint32_t tile, uint32_t lane_id, uint32_t& col, uint32_t& row, bool& valid) const
	{
		const uint32_t per_band = tiles_x * tiles_y;
		const uint32_t tz = tile / per_band, rem = tile % per_band;
		const uint32_t tx = rem / tiles_y, ty = rem % tiles_y;
		uint32_t lx, ly, lz;
		if (tw2 == 1 && tw1 % 8 == 0 && tzh % 8 == 0) {
			const uint32_t blk = lane_id / 64, l = lane_id % 64, bx = tw1 / 8;
			lx = (blk % bx) * 8 + l % 8;
			lz = (blk / bx) * 8 + l / 8;
			ly = 0;
		} else {
			lx = lane_id % tw1;
			ly = (lane_id / tw1) % tw2;
			lz = lane_id / (tw1 * tw2);
		}
		const uint32_t ix = tx * tw1 + lx, iy = ty * tw2 + ly, iz = tz * tzh + lz;
		valid = ix < nx && iy < ny && iz < n2;
		// lanes outside the grid shadow an in-tile point
		col = min(ix, nx - 1) * ny + min(iy, ny - 1);
		row = min(iz, n2 - 1);
	}
};

// Window word per (tile, rx, tx):
//   bits 0..30  ws * ESZ: byte offset in the channel of the first staged entry
//               (ws = floor of the tile's lowest position for the pair)
//   bit  31     record-end flag: some lane may have pos < 0 or i + 1 >= L,
//               or the pair has nothing to stage (padded transmit, window
//               past the record end)
// The lowest and highest positions use the same float expression order as
// the lanes', (offset + d_tx) + d_rx, so by monotonic rounding every lane's
// floor(pos) lies in [ws, floor(hi)], and floor(pos) + 1 in the staged
// window of wcap entries.  The channel itself (t * nrx + r) is not stored:
// the kernel forms its 64-bit address, so tensors of any size are tiled.
constexpr uint32_t kEdgeBit = 1u << 31;

typedef __attribute__((address_space(1))) const void* gptr_t;
typedef __attribute__((address_space(3))) void* lptr_t;
// Read-only tables (window words, weights) through the constant address
// space: uniform indices then become scalar loads.
typedef __attribute__((address_space(4))) const uint32_t* cu32_t;
typedef __attribute__((address_space(4))) const float* cf32_t;

struct TileArgs {
	DasArgs d;
	TileGeom tg;
	const uint32_t* words;    // [ntiles][nrx][ntx_pad]
	uint32_t S;               // 1 KiB load slots per staged window (wcap = S * SLOT entries)
	uint32_t ntx_pad;         // transmits padded to the instantiated count
	uint32_t nbuf;            // chunk buffers in the LDS ring (>= 2)
};

// The tiles of up to kMaxBatch frames (possibly of different shapes) in one
// launch: frame j's tiles are global tiles first[j] .. first[j+1]-1.  Every
// frame shares the kernel instantiation (storage, coherence factor, transmit
// block, chunk size, workgroup size) and the window size S and ring depth.
constexpr uint32_t kMaxBatch = 16;
struct TileBatch {
	TileArgs jobs[kMaxBatch];
	uint32_t first[kMaxBatch + 1];
	uint32_t njobs, total;
};

constexpr uint32_t kTileThreads = 256;
// Zeroed bytes after the analytic tensor: a window is staged as whole 1 KiB
// slots from its first entry, so the last channel's copies may read up to
// (slots per window) KiB past its end; the launcher keeps windows below this.
constexpr uint32_t kMaxSlots = 64;

// Direct global -> LDS copies (LDS DMA).  Each lane moves 16 (4) bytes from
// its own address to lds + 16 (4) * lane; lds must be wave-uniform.  Written
// as asm so that the compiler does not treat the asynchronous LDS write as
// aliasing every later LDS read (it would wait for all copies in flight);
// the kernel waits for them itself with s_waitcnt vmcnt.
// m0 is declared clobbered (it is reserved, hence the silenced warning); the
// kernels using these never rely on m0 themselves.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void lds_dma4(const void* g, uint32_t lds)
{
	asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dword %1, off" :: "s"(lds), "v"(g) : "memory", "m0");
}
// n (1..4) consecutive 1 KiB slots: global g + k KiB -> LDS lds + k KiB.
// The instruction offset of global_load_lds moves both the global and the
// LDS address (measured on gfx950: offset:1024 with m0 = B lands the source
// bytes [1024, 2048) at LDS B + 1024), so one m0 serves four slots.
__device__ __forceinline__ void lds_dma16_slots(const void* g, uint32_t lds, uint32_t n)
{
	switch (n) {
	case 1:
		asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
				:: "s"(lds), "v"(g) : "memory", "m0");
		break;
	case 2:
		asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off\n\t"
				"global_load_lds_dwordx4 %1, off offset:1024"
				:: "s"(lds), "v"(g) : "memory", "m0");
		break;
	case 3:
		asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off\n\t"
				"global_load_lds_dwordx4 %1, off offset:1024\n\tglobal_load_lds_dwordx4 %1, off offset:2048"
				:: "s"(lds), "v"(g) : "memory", "m0");
		break;
	default:
		asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off\n\t"
				"global_load_lds_dwordx4 %1, off offset:1024\n\tglobal_load_lds_dwordx4 %1, off offset:2048\n\t"
				"global_load_lds_dwordx4 %1, off offset:3072"
				:: "s"(lds), "v"(g) : "memory", "m0");
		break;
	}
}
#pragma clang diagnostic pop

// Wait until at most n (wave-uniform) vector memory operations of this wave
// are outstanding, rounding n down to a power of two: waits at least as long
// as s_waitcnt vmcnt(n), with 7 immediates instead of 64.  The expcnt and
// lgkmcnt fields stay at their maxima (no wait on them).
#define MISTA_VMCNT(n) __builtin_amdgcn_s_waitcnt(((n) & 15) | (7 << 4) | (15 << 8) | (((n) >> 4) << 14))
// Written as one asm block of compares and branches: as C++ the if-chain
// comes out restructured into flag registers and several branch blocks that
// every chunk walks through (about 15 scalar instructions for n = 0 instead of
// 3).  Small counts first: with two ring buffers (the common case) n is 0.
__device__ __forceinline__ void wait_vmcnt_at_most(uint32_t n)
{
	asm volatile(
			"s_cmp_lt_u32 %0, 1\n\t"
			"s_cbranch_scc0 .Lvw1_%=\n\t"
			"s_waitcnt vmcnt(0)\n\t"
			"s_branch .Lvwe_%=\n"
			".Lvw1_%=:\n\t"
			"s_cmp_lt_u32 %0, 2\n\t"
			"s_cbranch_scc0 .Lvw2_%=\n\t"
			"s_waitcnt vmcnt(1)\n\t"
			"s_branch .Lvwe_%=\n"
			".Lvw2_%=:\n\t"
			"s_cmp_lt_u32 %0, 4\n\t"
			"s_cbranch_scc0 .Lvw4_%=\n\t"
			"s_waitcnt vmcnt(2)\n\t"
			"s_branch .Lvwe_%=\n"
			".Lvw4_%=:\n\t"
			"s_cmp_lt_u32 %0, 8\n\t"
			"s_cbranch_scc0 .Lvw8_%=\n\t"
			"s_waitcnt vmcnt(4)\n\t"
			"s_branch .Lvwe_%=\n"
			".Lvw8_%=:\n\t"
			"s_cmp_lt_u32 %0, 16\n\t"
			"s_cbranch_scc0 .Lvw16_%=\n\t"
			"s_waitcnt vmcnt(8)\n\t"
			"s_branch .Lvwe_%=\n"
			".Lvw16_%=:\n\t"
			"s_cmp_lt_u32 %0, 32\n\t"
			"s_cbranch_scc0 .Lvw32_%=\n\t"
			"s_waitcnt vmcnt(16)\n\t"
			"s_branch .Lvwe_%=\n"
			".Lvw32_%=:\n\t"
			"s_waitcnt vmcnt(32)\n"
			".Lvwe_%=:"
			:
			: "s"(n)
			: "scc", "memory");
}

// Two consecutive LDS entries as one value of LdsPair's type (4-byte
// entries as their raw words).
__device__ __forceinline__ float4 pack_pair(uint64_t a, uint64_t b)
{
	const float2 u = __builtin_bit_cast(float2, a), v = __builtin_bit_cast(float2, b);
	return make_float4(u.x, u.y, v.x, v.y);
}
__device__ __forceinline__ uint2 pack_pair(uint32_t a, uint32_t b) { return make_uint2(a, b); }

// N LDS reads of two consecutive entries each (byte addresses addr; one
// ds_read2 per pair), written as plain loads and all issued before the
// first result is used: the compiler's wait insertion then waits for read
// q just before its use (lgkmcnt(N-1-q), the reads complete in order), so
// the first pairs' accumulation overlaps the later reads' latency.
template<uint32_t ESZ, int N>
__device__ __forceinline__ void lds_read_pairs(const uint32_t* addr, typename LdsPair<ESZ>::T* out)
{
	using W = typename std::conditional<ESZ == 8, uint64_t, uint32_t>::type;
	typedef __attribute__((address_space(3))) const W* lds_entry_t;
#pragma unroll
	for (int q = 0; q < N; ++q) {
		const lds_entry_t lp = (lds_entry_t)(size_t)addr[q];
		out[q] = pack_pair(lp[0], lp[1]);
	}
}

// Chunk i of a tile = pairs (rx r, tx c*TC .. c*TC+TC-1), r = i / CPR,
// c = i % CPR.  Chunks stream through a ring of nbuf LDS buffers, each
// holding TC windows of S 1 KiB slots plus the tile's rx delays (one float
// per lane, copied with the first chunk of every rx).  Chunk i + nbuf - 1 is
// copied while chunk i is computed.  Wave w copies the windows of pairs
// w, w + NW, ... of each chunk, every one as S whole slots from its first
// entry, so every wave issues the same number of copies per chunk and waits
// for chunk i alone with s_waitcnt vmcnt.  The chunk's window words and
// weights are scalar-loaded one chunk ahead; its TC LDS reads are issued
// together, each pair waiting for its own.  Chunks in which no pair can reach a record
// end (the common case) take a path without the range test.
// Occupancy: 256-lane workgroups target three waves per SIMD (up to 170
// registers each); a 512-lane workgroup with its large LDS ring holds a CU
// alone (two waves per SIMD), so it may use up to 256 registers, which buys
// all of a chunk's LDS reads in flight at once.
template<int NTHREADS> constexpr int tile_waves_per_eu() { return NTHREADS >= 512 ? 2 : 3; }

template<int STORAGE, int CF, int NTXMAX, int TC, int NTHREADS, bool MULTI>
__device__ __forceinline__ void das_tile(const TileArgs& A, const uint32_t tile, unsigned char* smem)
{
	using E = typename Entry<STORAGE>::T;
	constexpr uint32_t NT = NTHREADS;       // one tile point per lane
	constexpr uint32_t ESZ = sizeof(E);
	constexpr uint32_t SLOT = 64 * 16 / ESZ;   // entries per wave load instruction (1 KiB)
	constexpr uint32_t NW = NT / 64;
	static_assert(TC % NW == 0, "every wave copies whole windows");
	constexpr uint32_t PPW = TC / NW;       // windows copied per wave per chunk
	constexpr uint32_t CPB = NTXMAX / TC;   // chunks per block of NTXMAX transmits
	// LDS reads issued together: a whole chunk's where the register budget
	// allows it (two waves per SIMD), four with up to 64 transmit delays in
	// registers, else one at a time (128 registers hold transmit delays)
	constexpr int BATCH = NTHREADS >= 512 ? (TC > 8 ? 8 : TC) : (NTXMAX <= 64 ? 4 : 1);
	static_assert(BATCH == 1 || TC % BATCH == 0, "whole read groups per chunk");

	const DasArgs& a = A.d;
	const uint32_t tid = threadIdx.x, lane = tid % 64;
	const uint32_t wave = __builtin_amdgcn_readfirstlane(tid / 64);   // uniform: scalar copy loops
	uint32_t col, row;
	bool valid;
	A.tg.pixel(tile, tid, col, row, valid);
	const uint64_t npix = a.npix;
	const uint64_t p = static_cast<uint64_t>(col) * A.tg.n2 + row;
	const uint32_t L = a.L, ntx = a.ntx, nrx = a.nrx;
	// Transmits run in blocks: one of NTXMAX, or (MULTI, more than 128
	// transmits) nblk equal blocks of BT <= 128, block-major over the receive
	// elements, each receive element's partial sum kept in HBM from one block
	// to the next (the same fma chain as one pass over all transmits), so a
	// lane loads each block's transmit delays once.  A segment is one
	// (block, rx): CPS chunks, the first of which also brings the tile's rx
	// delays.
	const uint32_t nblk = MULTI ? (A.ntx_pad + NTXMAX - 1) / NTXMAX : 1u;
	const uint32_t BT = MULTI ? A.ntx_pad / nblk : NTXMAX;   // transmits per block
	const uint32_t CPS = MULTI ? BT / TC : CPB;               // chunks per segment
	const float lim = static_cast<float>(L - 1u);
	const uint32_t S = A.S, NB = A.nbuf;
	const uint32_t wcap = S * SLOT;
	const uint32_t win_bytes = TC * S * 1024u;
	const uint32_t buf_bytes = win_bytes + NT * 4;    // windows + rx delays
	const cu32_t wtab = (cu32_t)(A.words + static_cast<uint64_t>(tile) * nrx * A.ntx_pad);
	const unsigned char* tensor = static_cast<const unsigned char*>(a.tensor);
	const uint64_t chan_bytes = static_cast<uint64_t>(L) * ESZ;        // one channel
	const uint64_t tx_bytes = chan_bytes * nrx;                        // one transmit's channels
	const uint32_t lds0 = static_cast<uint32_t>(reinterpret_cast<size_t>((__attribute__((address_space(3))) void*)smem));

	// tx delays of this lane's point for transmit block blk, offset folded
	// in: (offset + d_tx).  Transmits are padded to whole chunks: padded ones
	// have weight 0, position far below 0 and copy from the tensor's start,
	// so they take the excluded path.
	float dtx[NTXMAX];
	auto load_dtx = [&](uint32_t blk) {
		const int nvalid = static_cast<int>(min(BT, ntx - blk * BT));   // the block's transmits
#pragma unroll
		for (int t = 0; t < NTXMAX; ++t) {
			const uint32_t tt = min(blk * BT + static_cast<uint32_t>(t), ntx - 1);   // unconditional load, no branch
			const float d = a.delay[static_cast<uint64_t>(a.first_tx + tt) * npix + p];
			// 0 for the block's transmits, <= -1e30 past them: arithmetic
			// rather than a select, so no compare mask per transmit is held
			// in scalar registers while the loads are in flight
			const float pen = fminf(0.0f, static_cast<float>(nvalid - 1 - t)) * 1.0e30f;
			dtx[t] = (a.offset + d) + pen;
		}
	};
	load_dtx(0);

	const uint32_t nchunks = nblk * nrx * CPS;
	const uint32_t Q = PPW * S;   // window copies per wave per chunk (+1 for the rx delays at an rx start)
	const uint32_t S4 = S & ~3u, SR = S & 3u;   // slots in groups of four, and the rest
	// Window words and weights are laid out [rx][ntx_pad] (the transmit rows
	// padded to whole chunks): the chunk of segment (blk, r), position c
	// holds entries r * ntx_pad + blk * BT + c * TC .. + TC - 1 of the tile's
	// table and of the weights (without blocks: chunk k at k * TC).
	// Issue cursor: the next chunk to copy, its segment position, its table
	// offset, its ring buffer and the address of its first window's channel.
	uint32_t is_k = 0, is_c = 0, is_r = 0, is_b = 0, is_off = 0, is_lds = lds0;
	const unsigned char* is_rbase = tensor;   // tensor + (block start) * tx_bytes + (rx) * chan_bytes
	const unsigned char* is_cbase = tensor;   // + is_c * TC * tx_bytes
	const float* is_drx = a.delay + p;        // rx delays of is_k's rx for this lane
	uint32_t iw[PPW];                         // words of is_k's pairs that this wave copies
	auto load_issue_words = [&]() {
		const cu32_t pr = wtab + is_off + wave;
#pragma unroll
		for (uint32_t q = 0; q < PPW; ++q) iw[q] = pr[q * NW];
	};
	uint32_t pend = 0;   // (MULTI) segment starts among the chunks issued and not yet computed
	auto issue = [&]() {
		if (MULTI && is_c == 0) ++pend;
#pragma unroll
		for (uint32_t k = 0; k < PPW; ++k) {
			const uint32_t j = wave + k * NW;
			// padded transmits copy from the tensor's start (weight 0)
			const unsigned char* src = (is_b * BT + is_c * TC + j < ntx) ? is_cbase + j * tx_bytes + (iw[k] & ~kEdgeBit) : tensor;
			src += lane * 16;
			uint32_t dst = is_lds + j * S * 1024u;
			for (uint32_t s0 = 0; s0 < S4; s0 += 4, src += 4096, dst += 4096) lds_dma16_slots(src, dst, 4);
			if (SR) lds_dma16_slots(src, dst, SR);
		}
		if (is_c == 0) lds_dma4(is_drx, __builtin_amdgcn_readfirstlane(is_lds + win_bytes + wave * 256));
		// advance to the next chunk
		++is_k;
		is_lds += buf_bytes;
		if (is_lds == lds0 + NB * buf_bytes) is_lds = lds0;
		is_cbase += TC * tx_bytes;
		is_off += TC;
		if (++is_c == CPS) {
			is_c = 0;
			if (MULTI && ++is_r == nrx) {   // next block
				is_r = 0;
				++is_b;
				is_rbase = tensor + static_cast<uint64_t>(is_b) * BT * tx_bytes;
				is_drx = a.delay + p;
				is_off = is_b * BT;
			} else {
				is_rbase += chan_bytes;
				is_drx += npix;
				is_off += A.ntx_pad - BT;   // (0 without blocks)
			}
			is_cbase = is_rbase;
		}
	};
	const uint32_t pre = min(NB - 1, nchunks);
	for (uint32_t k = 0; k < pre; ++k) {
		load_issue_words();
		issue();
	}
	if (is_k < nchunks) load_issue_words();
	// window words and weights of the chunk to compute, one chunk ahead
	uint32_t cw[TC];
	float cwt[TC];
	auto load_chunk = [&](uint32_t off) {
		const cu32_t pr = wtab + off;
		const cf32_t wr = (cf32_t)(a.w + static_cast<uint64_t>(off));
#pragma unroll
		for (uint32_t j = 0; j < TC; ++j) {
			cw[j] = pr[j];
			cwt[j] = wr[j];
		}
	};
	load_chunk(0);

	TileStats<CF> st;
	st.m = reinterpret_cast<double*>(smem + NB * buf_bytes + tid * 32);
	uint32_t i = 0, bbase = 0;   // chunk being computed, its buffer's LDS offset
	for (uint32_t blk = 0; blk < nblk; ++blk) {
		if (MULTI && blk > 0) load_dtx(blk);   // waited for at its first use
		for (uint32_t r = 0; r < nrx; ++r) {
			float drx = 0.0f, ar = 0.0f, ai = 0.0f;
			if (MULTI && blk > 0) {   // the sums of the earlier blocks
				const float2 ps = a.partial[static_cast<uint64_t>(r) * npix + p];
				ar = ps.x;
				ai = ps.y;
			}
			// (MULTI) table offset of the next chunk to load and of the next
			// segment's first chunk: the next rx's, or the next block's
			uint32_t ld_off = r * A.ntx_pad + blk * BT;
			const uint32_t seg_next = r + 1 < nrx ? ld_off + A.ntx_pad : (blk + 1) * BT;
			uint32_t left = CPS;   // (MULTI) chunks of this segment still to compute
#pragma clang loop unroll(full)
			for (int c = 0; c < static_cast<int>(CPB); ++c) {
				// A block shorter than NTXMAX skips its last chunks: a guard,
				// not an exit, so the trip count stays constant, the loop
				// unrolls and dtx keeps compile-time indices.  The count is
				// opaque to the compiler, which would otherwise hoist one
				// compare per chunk out of the rx loop into scalar registers.
				if (MULTI) {
					asm volatile("" : "+s"(left));
					if (left == 0) continue;
				}
				// this chunk's words and weights (loaded one chunk ahead)
				uint32_t w_[TC];
				float wt[TC];
#pragma unroll
				for (int j = 0; j < TC; ++j) {
					w_[j] = cw[j];
					wt[j] = cwt[j];
				}
				// copies issued after chunk i: chunks i+1 .. is_k-1, Q each
				// plus one for every chunk that starts a segment (loads and
				// stores of partial sums issued since only make the wait
				// stricter)
				uint32_t after;
				if (MULTI) after = (is_k - 1 - i) * Q + pend - (c == 0 ? 1u : 0u);
				else if (i + NB - 2 < nchunks) after = (NB - 2) * Q + (static_cast<uint32_t>(c) + NB - 2) / CPS;
				else after = (nchunks - 1 - i) * Q + (nrx - 1 - r);
				wait_vmcnt_at_most(after);         // this wave's copies of chunk i landed
				__syncthreads();                   // everyone's did; buffer of chunk i-1 is free
				// Waves still issuing copies, positions and LDS reads go
				// before waves already accumulating (the reads then start
				// sooner after the barrier): F32 DAS 9.97 -> 9.94 ms, F16
				// 6.33 -> 6.26 ms.  Lowered once the chunk's last reads are issued.
				if constexpr (BATCH > 1) __builtin_amdgcn_s_setprio(2);
				if (MULTI && c == 0) --pend;
				if (is_k < nchunks) issue();
				if (c == 0) drx = *reinterpret_cast<const float*>(smem + bbase + win_bytes + tid * 4);
				// Register fences: this chunk's positions are computed after
				// the barrier, and its accumulation is finished before the
				// next one, so only one chunk's temporaries are ever live.
				asm volatile("" : "+v"(drx));
				uint32_t any_edge = 0;
#pragma unroll
				for (int j = 0; j < TC; ++j) any_edge |= w_[j];
				const uint32_t wbase = lds0 + bbase;
				// entry address and weight of pair j for this lane
				auto fast_pair = [&](int j, uint32_t& addr, float& kf, float& wv) {
					// every lane of every pair of the chunk is inside 0 <= pos < L - 1
					const float pos = dtx[c * TC + j] + drx;
					kf = __builtin_amdgcn_fractf(pos);
					addr = static_cast<uint32_t>(pos) * ESZ + (wbase + static_cast<uint32_t>(j) * S * 1024u - w_[j]);
					wv = wt[j];
				};
				auto edge_pair = [&](int j, uint32_t& addr, float& kf, float& wv) {
					// One form for every pair: lanes outside 0 <= pos < L-1
					// get weight 0; the entry index floor(pos) - ws is
					// clamped, as an unsigned value, to the window, which
					// changes nothing for in-range lanes (it lies in
					// [0, need - 2] by the window construction) and keeps
					// excluded lanes' reads on staged (finite) entries, whose
					// products with weight 0 leave the sums as they are (a sum
					// is never -0).  0 <= pos < lim is one unsigned compare of
					// the bit patterns: positions are never -0 (delays are
					// >= +0), so non-negative ones order like their bits, and
					// negative values and NaN compare above lim's.  (The
					// fenced copy of drx keeps the compiler from computing
					// both paths' positions ahead of the branch.)
					float dr = drx;
					asm volatile("" : "+v"(dr));
					const float pos = dtx[c * TC + j] + dr;
					kf = __builtin_amdgcn_fractf(pos);
					const uint32_t ws = (w_[j] & ~kEdgeBit) / ESZ;
					const bool ok = __float_as_uint(pos) < __float_as_uint(lim);
					const float pc = __builtin_amdgcn_fmed3f(pos, 0.0f, lim);   // NaN-free, no int overflow
					const uint32_t rel = min(static_cast<uint32_t>(static_cast<int>(pc)) - ws, wcap - 2u);
					addr = rel * ESZ + wbase + static_cast<uint32_t>(j) * S * 1024u;
					wv = ok ? wt[j] : 0.0f;
				};
				auto accumulate = [&](const float2 v, float wv) {
					ar = fmaf(v.x, wv, ar);
					ai = fmaf(v.y, wv, ai);
				};
				auto run = [&](auto&& pair) {
					if constexpr (BATCH > 1) {
						// groups of BATCH pairs: BATCH LDS reads in flight
#pragma unroll
						for (int g = 0; g < TC; g += BATCH) {
							uint32_t addr[BATCH];
							float kf[BATCH], wv[BATCH];
#pragma unroll
							for (int q = 0; q < BATCH; ++q) pair(g + q, addr[q], kf[q], wv[q]);
							typename LdsPair<ESZ>::T e[BATCH];
							lds_read_pairs<ESZ, BATCH>(addr, e);
							if (g + BATCH >= TC) __builtin_amdgcn_s_setprio(0);
#pragma unroll
							for (int q = 0; q < BATCH; ++q) accumulate(Entry<STORAGE>::lerp2(e[q], kf[q]), wv[q]);
						}
					} else {
#pragma unroll
						for (int j = 0; j < TC; ++j) {
							uint32_t addr;
							float kf, wv;
							pair(j, addr, kf, wv);
							const E* e = reinterpret_cast<const E*>(smem + (addr - lds0));
							accumulate(Entry<STORAGE>::lerp(e, kf), wv);
						}
					}
				};
				if (!(any_edge & kEdgeBit)) run(fast_pair);
				else run(edge_pair);
				asm volatile("" : "+v"(ar), "+v"(ai));
				// the next chunks' scalars, after this chunk's LDS reads (a
				// scalar load in flight would hold up their wait): in flight
				// during the next barrier
				if (is_k < nchunks) load_issue_words();
				if (i + 1 < nchunks) {
					if (MULTI) {
						ld_off = left > 1 ? ld_off + TC : seg_next;
						load_chunk(ld_off);
					} else {
						load_chunk((i + 1) * TC);
					}
				}
				bbase += buf_bytes;
				if (bbase == NB * buf_bytes) bbase = 0;
				++i;
				if (MULTI) --left;
			}
			if (MULTI && blk + 1 < nblk) {
				if (valid) a.partial[static_cast<uint64_t>(r) * npix + p] = make_float2(ar, ai);
			} else {
				st.add_rx(ar, ai, r == 0);
			}
		}
	}
	MISTA_VMCNT(0);
	if (valid) st.write(a, p);
}

// A workgroup per tile.  XCD-aware order: workgroups are dealt round-robin
// to the 8 XCDs (block b runs on XCD b % 8), so XCD x gets one contiguous run
// of the tiles; tiles running at the same time on one XCD then share windows
// in its L2.
__device__ __forceinline__ uint32_t xcd_tile()
{
	const uint32_t T = gridDim.x, q8 = T / 8, rm8 = T % 8, xcd = blockIdx.x % 8;
	return xcd * q8 + min(xcd, rm8) + blockIdx.x / 8;
}

// Windows of pairs with nothing to stage are read (with weight 0): the ring
// starts from zeros so that no stale bit pattern can be Inf/NaN.
template<int TC, int NTHREADS>
__device__ __forceinline__ void zero_ring(unsigned char* smem, uint32_t nbuf, uint32_t S)
{
	const uint32_t ring = nbuf * (TC * S * 1024u + NTHREADS * 4u);
	for (uint32_t i = threadIdx.x; i < ring / 16; i += NTHREADS)
		reinterpret_cast<uint4*>(smem)[i] = make_uint4(0u, 0u, 0u, 0u);
	__syncthreads();
}

// One frame: its arguments arrive as the kernel's own, which the compiler
// keeps in scalar registers for the whole tile.
template<int STORAGE, int CF, int NTXMAX, int TC, int NTHREADS, bool MULTI>
__global__ void __launch_bounds__(NTHREADS)
__attribute__((amdgpu_waves_per_eu(tile_waves_per_eu<NTHREADS>(), tile_waves_per_eu<NTHREADS>())))
das_tile_kernel(const TileArgs A)
{
	extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
	zero_ring<TC, NTHREADS>(smem, A.nbuf, A.S);
	das_tile<STORAGE, CF, NTXMAX, TC, NTHREADS, MULTI>(A, xcd_tile(), smem);
}

// A batch of frames in one launch, their tiles concatenated: the hardware
// dispatcher starts each workgroup as a CU frees up, so frames of uneven cost
// balance by themselves.  (A separate kernel: a frame's arguments picked from
// the table by a run-time index are reloaded from memory at their uses.)
template<int STORAGE, int CF, int NTXMAX, int TC, int NTHREADS, bool MULTI>
__global__ void __launch_bounds__(NTHREADS)
__attribute__((amdgpu_waves_per_eu(tile_waves_per_eu<NTHREADS>(), tile_waves_per_eu<NTHREADS>())))
das_tile_batch_kernel(const TileBatch B)
{
	extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
	zero_ring<TC, NTHREADS>(smem, B.jobs[0].nbuf, B.jobs[0].S);
	const uint32_t g = xcd_tile();
	uint32_t j = 0;
	while (j + 1 < B.njobs && B.first[j + 1] <= g) ++j;
	das_tile<STORAGE, CF, NTXMAX, TC, NTHREADS, MULTI>(B.jobs[j], g - B.first[j], smem);
}

// LDS bytes of one chunk buffer: TC windows of S 1 KiB slots + the rx delays.
inline size_t tile_buffer_bytes(uint32_t tc, uint32_t S, uint32_t nt)
{
	return static_cast<size_t>(tc) * S * 1024u + nt * 4;
}

template<int STORAGE, int CF, int NTXMAX>
void launch_tile_t(const TileBatch& B, uint32_t nt, uint32_t tc, hipStream_t s)
{
	const TileArgs& A = B.jobs[0];   // S, nbuf and the transmit padding are shared by the batch
	const size_t lds = A.nbuf * tile_buffer_bytes(tc, A.S, nt) + tile_stats_bytes(CF, nt);
	const uint32_t grid = B.total;   // a workgroup per tile
	auto go1 = [&](auto kern, auto kern_batch, uint32_t threads) {
		if (B.njobs == 1) {
			HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
					static_cast<int>(lds)));
			hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, s, A);
		} else {
			HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern_batch),
					hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
			hipLaunchKernelGGL(kern_batch, dim3(grid), dim3(threads), lds, s, B);
		}
	};
#define MISTA_GO(tcv, ntv, multi) \
	go1(das_tile_kernel<STORAGE, CF, NTXMAX, tcv, ntv, multi>, das_tile_batch_kernel<STORAGE, CF, NTXMAX, tcv, ntv, multi>, ntv)
	if constexpr (NTXMAX == 128) {
		if (A.ntx_pad > NTXMAX) {   // several transmit blocks: 512-lane tiles only
			if (nt != 512) MISTA_THROW(MISTA_E_INVALID_STATE, "tile of " << nt << " lanes for " << A.ntx_pad << " transmits");
			MISTA_GO(8, 512, true);
			return;
		}
	}
	if constexpr (NTXMAX == 128) {
		if (nt == 512 && tc == 16) {
			MISTA_GO(16, 512, false);
			return;
		}
	}
	if (nt == 512) MISTA_GO(8, 512, false);
	else if (tc == 8) MISTA_GO(8, 256, false);
	else MISTA_GO(4, 256, false);
#undef MISTA_GO
}

template<int STORAGE, int CF>
void launch_tile_n(const TileBatch& B, uint32_t nt, uint32_t tc, hipStream_t s)
{
	const uint32_t ntx_pad = B.jobs[0].ntx_pad;
	if (ntx_pad == 32) launch_tile_t<STORAGE, CF, 32>(B, nt, tc, s);
	else if (ntx_pad == 64) launch_tile_t<STORAGE, CF, 64>(B, nt, tc, s);
	else launch_tile_t<STORAGE, CF, 128>(B, nt, tc, s);
}

// Tiled launch for one storage kind (instantiated in das_tile_<storage>.hip).
template<int STORAGE>
void launch_tile_storage(int cf, const TileBatch& B, uint32_t nt, uint32_t tc, hipStream_t s)
{
	switch (cf) {
	case MISTA_CF_NONE: launch_tile_n<STORAGE, MISTA_CF_NONE>(B, nt, tc, s); break;
	case MISTA_CF_PCF: launch_tile_n<STORAGE, MISTA_CF_PCF>(B, nt, tc, s); break;
	default: launch_tile_n<STORAGE, MISTA_CF_SCF>(B, nt, tc, s); break;
	}
}

}  // namespace das
}  // namespace mista

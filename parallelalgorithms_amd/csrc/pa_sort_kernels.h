// pa_sort_kernels.h -- device kernels of the LSD radix sort, shared by the
// plain sort (pa_sort.hip) and the segmented batch sort (pa_segsort.hip).
#pragma once

#include "pa_internal.h"

namespace pa {

constexpr int kSortBlock = 512;  // histogram kernel block

// Look-back status word: 2 flag bits + 30-bit count.
constexpr uint32_t kLbAgg = 1u << 30;
constexpr uint32_t kLbInc = 2u << 30;
constexpr uint32_t kLbMask = (1u << 30) - 1;
constexpr size_t kMaxSortN = (size_t(1) << 30) - 1;

// Tile claims of a pass whose prefixes are ready (no tile waits on another):
// tiles go to XCDs in groups of kXcdGroup consecutive tiles (XCD y owns the
// groups g with g % kXcds == y) and each XCD claims its own groups in order
// from its own counter, so neighbouring tiles run at the same time on one XCD
// and the partial cache lines their runs share in every bin meet in that
// XCD's L2.  A block whose XCD has no tiles left takes the next XCD's, so every
// tile is claimed whatever the placement of the blocks.  The XCD index is the
// low 4 bits of hardware register HW_REG_XCC_ID (20), read with s_getreg
// (immediate: size 4 - 1 << 11 | offset 0 << 6 | register 20).
constexpr uint32_t kXcds = 8;
constexpr uint32_t kXcdGroup = 16;
constexpr int kHwRegXccId = (3 << 11) | 20;
// Tile counters of one sort: kXcds for pass 0, one for each later pass.
constexpr uint32_t sort_tile_counters(int passes) { return kXcds + (uint32_t)passes - 1; }

// Pass flags: key transform on load / on store (kSignedKeys selects the
// two's-complement transform, else the IEEE-754 one); kPrefixReady: every
// tile's inclusive prefix was published before the pass (pass 0 of the plain
// sort, from the histogram read), so tiles neither publish nor wait.
enum : uint32_t { kFloatIn = 1u, kFloatOut = 2u, kPrefixReady = 4u, kSignedKeys = 8u };

// Pass roles of a plain sort of integer keys, written after the histogram
// read (k_tile_scan) into the top bits of each pass's tile counters, so every
// claim returns them with its tile index at no extra cost: a pass whose digit
// is the same for every key leaves the order as it is and is skipped
// (kRoleSkip); the others permute, reading the buffer the previous permuting
// pass wrote (kRoleSwap: the other one than the host's alternation assumes).
// Sorts with a key transform (float, signed) run without roles: the transform
// is applied in pass 0 and undone in the last pass, and making those two
// passes a run-time choice cost the kernel 8 more spilled registers.  Tile
// indices stay below 2^30.
enum : uint32_t { kRoleSkip = 1u << 30, kRoleSwap = 1u << 31, kRoleBits = 3u << 30 };

// Tile-to-array modes of the onesweep pass.
enum : int {
  kModePlain = 0,  // one array, tiles in order, one look-back chain
  kModeSeg = 1,    // independent segments (segmented batch sort)
};


// One read of the keys -> D digit histograms (D = sizeof(K)), counts laid out
// [D][256].  Also clears the look-back table of the onesweep passes that
// follow (zero_vec 16-byte words at zero_ptr).
template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_hist_digits(const K* __restrict__ keys, size_t n,
                                                            uint32_t* __restrict__ counts, uint4* __restrict__ zero_ptr,
                                                            size_t zero_vec, uint32_t xform) {
  constexpr int D = sizeof(K);
  constexpr int PER = 16 / sizeof(K);
  __shared__ uint32_t s_hist[D * 256];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < D * 256; i += kSortBlock) s_hist[i] = 0;

  const size_t gstride = (size_t)gridDim.x * kSortBlock;
  for (size_t i = (size_t)blockIdx.x * kSortBlock + tid; i < zero_vec; i += gstride) zero_ptr[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();

  const uint32_t part = blockIdx.x, parts = gridDim.x;
  const size_t pstride = (size_t)parts * kSortBlock;
  size_t head = ((16 - ((uintptr_t)keys & 15)) & 15) / sizeof(K);
  if (head > n) head = n;
  const size_t nvec = (n - head) / PER;
  const uint4* vk = reinterpret_cast<const uint4*>(keys + head);
  // U 16-byte loads per thread are issued before any counting, so each wave
  // keeps U KiB in flight.
  constexpr int U = 4;
  const size_t v0 = (size_t)part * kSortBlock + tid;
  for (size_t base = 0; base < nvec; base += pstride * U) {
    uint4 x[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t v = base + u * pstride + v0;
      ok[u] = v < nvec;
      if (ok[u]) x[u] = vk[v];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      K k[PER];
      __builtin_memcpy(k, &x[u], 16);
      if (xform) {
#pragma unroll
        for (int j = 0; j < PER; ++j) k[j] = dev::key_fwd<K>(k[j], xform);
      }
      // Whole-key fast path: every active lane holds PER copies of one key
      // (constant input) -> one lane adds the wave's count to each digit.
      bool same = true;
#pragma unroll
      for (int j = 1; j < PER; ++j) same = same && (k[j] == k[0]);
      const uint32_t f_lo = __builtin_amdgcn_readfirstlane((uint32_t)k[0]);
      const uint32_t f_hi = sizeof(K) == 8 ? __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)k[0] >> 32)) : 0u;
      const K first = sizeof(K) == 8 ? (K)(((uint64_t)f_hi << 32) | f_lo) : (K)f_lo;
      const uint64_t active = __ballot(true);
      if (__ballot(same && k[0] == first) == active) {
        if (dev::count_below(active) == 0) {
          const uint32_t c = (uint32_t)(PER * __popcll(active));
#pragma unroll
          for (int d = 0; d < D; ++d) atomicAdd(&s_hist[d * 256 + dev::digit8<K>(first, 8 * d)], c);
        }
        continue;
      }
#pragma unroll
      for (int d = 0; d < D; ++d) {
        uint32_t dg[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) dg[j] = dev::digit8<K>(k[j], 8 * d);
        dev::lds_count<PER>(s_hist + d * 256, dg);
      }
    }
  }
  // Head and tail elements that do not fill a 16-byte vector.
  if (part == 0) {
    const size_t tail_start = head + nvec * PER;
    const size_t extra = head + (n - tail_start);
    for (size_t e = tid; e < extra; e += kSortBlock) {
      size_t idx = e < head ? e : tail_start + (e - head);
      K kk = keys[idx];
      if (xform) kk = dev::key_fwd<K>(kk, xform);
#pragma unroll
      for (int d = 0; d < D; ++d) atomicAdd(&s_hist[d * 256 + dev::digit8<K>(kk, 8 * d)], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < D * 256; i += kSortBlock) {
    uint32_t c = s_hist[i];
    if (c) atomicAdd(&counts[i], c);
  }
}


// One stable RB-bit digit pass (onesweep), R = 2^RB digits (RB = 4..8).  A block ranks a tile of
// BLOCK x ITEMS keys in LDS, publishes the tile's per-digit counts to a
// decoupled look-back table, resolves its global offsets from the
// predecessors' words and writes the tile out in digit order.
//
// Tiles are claimed in order from a global counter, so every tile a block
// waits on in the look-back is held by a block that is already running.
//
// Ranking units are groups of UL lanes (32 or 16).  Unit u owns the USPAN =
// UL x ITEMS consecutive keys [u*USPAN, (u+1)*USPAN) of the tile; item i of
// lane l is key u*USPAN + i*UL + l, so (unit, item, lane) order is input order
// and a load instruction reads 64/UL contiguous runs.  Each unit has R LDS
// slots: one part collects the lanes holding a digit in the current item
// (each lane ORs its bit), the other counts that digit's keys seen so far in
// the unit.  One read returns both, so a key's stable rank in its unit is
// count + (peer lanes below it); the lowest peer then writes back
// (count + peers, no lanes), which also clears the lane mask.  LDS operations
// of one wave execute in program order and OR commutes, so the result does
// not depend on the order in which the LDS applies the lanes.
//
// Partial tiles are padded with all-ones keys: their digit is the largest any
// key can have in every pass (R - 1, or all ones in a narrower last digit),
// so they rank after every real key of the tile and occupy the staging
// slots [valid, TILE), which are never written out.  Only the last tile of an
// array (or segment) is partial, and no tile looks back past the first tile
// of its own array, so the padded count it publishes is never consumed.
//
// Segmented mode: tile_info[t] = (segment << 32) | tile index within the
// segment; the segment's input/output addresses and length come from seg_in /
// seg_out / seg_len, its digit counts from gtab + segment * gstride * R and
// its first look-back row from seg_row[segment] (a segment's rows are
// contiguous; its tiles may be claimed interleaved with other segments').
// Plain mode: gtab holds the pass's exclusive digit bases.
//
// LATE: the look-back's first round trip is issued while the block scatters
// its keys into LDS and read afterwards, so the status loads are in flight
// during the scatter.
//
// MINW: minimum waves per SIMD the register allocation must allow.
// UNI: items whose 64 lanes all hold one digit take a path without LDS atomics.
// UL: lanes per ranking unit.  32: 64-bit slots (32-bit lane mask | 32-bit
// count).  16: 32-bit slots (16-bit lane mask | 16-bit count); a slot is one
// LDS bank word, so the 64 lanes of a rank step spread over every bank instead
// of every second one (the 64-bit slots start on even banks).
// SPLIT: the tile is staged and written out in SPLIT rounds of TILE / SPLIT
// slots (round r takes the keys whose staging slot is in its range), so a
// tile larger than the LDS staging area can be held in registers: the block
// then has more keys in flight per byte of LDS and per look-back.
template <typename K, bool KV, int BLOCK, int ITEMS, int LBW, int MODE = kModePlain, bool LATE = false, int MINW = 1,
          bool UNI = true, int SPLIT = 1, int UL = 32, int RB = 8>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(MINW))) void k_onesweep(
    const K* __restrict__ kin, K* __restrict__ kout, const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout,
    uint32_t n, const uint32_t* __restrict__ gtab, uint32_t* __restrict__ lb, uint32_t* __restrict__ tile_counter,
    uint32_t ntiles, uint32_t shift, uint32_t fmode, const uint64_t* __restrict__ tile_info = nullptr,
    const uint64_t* __restrict__ seg_in = nullptr, const uint64_t* __restrict__ seg_out = nullptr,
    const uint32_t* __restrict__ seg_len = nullptr, uint32_t gstride = 0,
    const uint32_t* __restrict__ seg_row = nullptr) {
  constexpr bool SEG = MODE == kModeSeg;
  static_assert(RB >= 4 && RB <= 8, "digits of 4 to 8 bits");
  constexpr uint32_t R = 1u << RB;
  static_assert(BLOCK >= (int)R && BLOCK <= 1024 && BLOCK % 64 == 0, "one thread per digit is needed");
  static_assert(!(SEG && KV), "segmented mode sorts keys only");
  static_assert(UL == 32 || UL == 16, "ranking units are 16 or 32 lanes");
  constexpr int UNITS = BLOCK / UL;
  constexpr int USPAN = UL * ITEMS;
  constexpr int SLOTB = UL == 32 ? 8 : 4;  // bytes per rank slot
  static_assert(UL == 32 || BLOCK * ITEMS <= 65535, "16-bit slot counts need tiles below 2^16 keys");
  constexpr int TILE = BLOCK * ITEMS;
  static_assert(ITEMS % SPLIT == 0 && (SPLIT == 1 || !KV), "split staging is for keys-only tiles");
  constexpr int STAGE = TILE / SPLIT;      // staging slots per round
  constexpr int RITEMS = ITEMS / SPLIT;    // write-out slots per thread per round
  static_assert(STAGE * sizeof(K) >= UNITS * R * SLOTB, "rank slots must fit the key staging area");
  __shared__ __align__(16) K s_keys[STAGE];
  __shared__ uint32_t s_goff[R];
  __shared__ uint32_t s_wsum[2][4];
  __shared__ uint32_t s_claim;

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const uint32_t wave = tid >> 6;
  const uint32_t unit = tid / UL;
  const uint32_t ul = tid % UL;
  const uint32_t lanebit = 1u << ul;
  // Rank slots of this thread's unit (they alias the key staging area, which
  // is idle until the ranking is over).
  uint64_t* const uslot = reinterpret_cast<uint64_t*>(s_keys) + unit * R;
  uint32_t* const uword = reinterpret_cast<uint32_t*>(uslot);
  // High words of every unit's slots, for the per-digit pass over units.
  uint32_t* const hiword = reinterpret_cast<uint32_t*>(s_keys) + 1;
  // 16-lane units: slot (u, d) is word u * R + d, count in the high half.
  uint32_t* const slot16 = reinterpret_cast<uint32_t*>(s_keys);
  uint32_t* const uslot16 = slot16 + unit * R;
  const bool float_in = (fmode & kFloatIn) != 0;
  const bool float_out = (fmode & kFloatOut) != 0;
  const uint32_t xform = (fmode & kSignedKeys) ? dev::kXformSigned : dev::kXformFloat;

  // Clear the rank slots while the claim is in flight.
  {
    uint4* z = reinterpret_cast<uint4*>(s_keys);
    constexpr uint32_t ZV = UNITS * R * SLOTB / 16;  // 16-byte words of rank slots
#pragma unroll
    for (uint32_t i = tid; i < ZV; i += BLOCK) z[i] = make_uint4(0, 0, 0, 0);
  }
  if (tid == 0) {
    if (fmode & kPrefixReady) {
      const uint32_t x = __builtin_amdgcn_s_getreg(kHwRegXccId) & (kXcds - 1);
      uint32_t t = ntiles, bits = 0;
      for (uint32_t k = 0; k < kXcds && t == ntiles; ++k) {
        const uint32_t y = (x + k) & (kXcds - 1);
        const uint32_t jb = atomicAdd(tile_counter + y, 1u);
        const uint32_t j = jb & ~kRoleBits;
        bits = jb & kRoleBits;
        const uint64_t c = ((uint64_t)(j / kXcdGroup) * kXcds + y) * kXcdGroup + j % kXcdGroup;
        if (c < ntiles) t = (uint32_t)c;
      }
      s_claim = t | bits;
    } else {
      s_claim = atomicAdd(tile_counter, 1u);
    }
  }
  __syncthreads();
  // Uniform across the block: the tile geometry below lives in scalar registers.
  const uint32_t cv = __builtin_amdgcn_readfirstlane(s_claim);
  const uint32_t rl = cv & kRoleBits;
  const uint32_t claimed = cv & ~kRoleBits;
  if (claimed >= ntiles || (rl & kRoleSkip)) return;

  // Geometry of the claimed tile.
  const K* in;
  K* out;
  const uint32_t* vsrc = nullptr;
  uint32_t* vdst = nullptr;
  const uint32_t* g;  // digit bases (plain) or segment digit counts (SEG)
  uint32_t len;       // keys in the tile's array (segment)
  uint32_t row = claimed;  // look-back row of the tile
  uint32_t local;          // tile index within its array (segment)
  if constexpr (SEG) {
    const uint64_t info = tile_info[claimed];
    const uint32_t seg = (uint32_t)(info >> 32);
    local = (uint32_t)info;
    row = seg_row[seg] + local;
    in = reinterpret_cast<const K*>(seg_in[seg]);
    out = reinterpret_cast<K*>(seg_out[seg]);
    len = seg_len[seg];
    g = gtab + (size_t)seg * gstride * R;
  } else {
    local = claimed;
    const bool swap = (rl & kRoleSwap) != 0;
    in = swap ? kout : kin;
    out = swap ? const_cast<K*>(kin) : kout;
    vsrc = swap ? vout : vin;
    vdst = swap ? const_cast<uint32_t*>(vin) : vout;
    len = n;
    g = gtab;
  }
  const uint32_t tile_first = local * (uint32_t)TILE;
  const uint32_t valid = min((uint32_t)TILE, len - tile_first);

  // Keys into registers through a buffer resource over the tile (one offset
  // register, immediate offsets per item).  Slots past the array's end read 0
  // from the range check and become padding: all-ones once transformed, so
  // they rank after every real key.
  K key[ITEMS];
  {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<K*>(in) + tile_first, (short)0,
                                                                         (int)(valid * sizeof(K)), 0x00020000);
    const uint32_t voff = (unit * USPAN + ul) * (uint32_t)sizeof(K);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) key[i] = dev::buffer_load<K>(rs, voff + i * UL * (uint32_t)sizeof(K));
    if (valid < (uint32_t)TILE) {
      const K pad = float_in ? dev::key_inv<K>(~K(0), xform) : ~K(0);
      // Item i is padding when i * UL >= lim (one register, not one per item).
      int lim = (int)valid - (int)(unit * USPAN + ul);
      asm volatile("" : "+v"(lim));
#pragma unroll
      for (int i = 0; i < ITEMS; ++i)
        if (i * UL >= lim) key[i] = pad;
    }
    if (float_in) {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) key[i] = dev::key_fwd<K>(key[i], xform);
    }
  }
  uint32_t val[KV ? ITEMS : 1];
  if constexpr (KV) {
    const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(vsrc) + tile_first,
                                                                         (short)0, (int)(valid * 4u), 0x00020000);
    const uint32_t voff = (unit * USPAN + ul) * 4u;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) val[i] = __builtin_amdgcn_raw_buffer_load_b32(rv, voff + i * UL * 4u, 0, 0);
  }
  // Stores of 4-byte keys and payloads go through buffer resources over the
  // output (32-bit offsets; n < 2^30 keeps the byte range below 2^32).
  const __amdgpu_buffer_rsrc_t rs_out =
      __builtin_amdgcn_make_buffer_rsrc(out, (short)0, (int)(sizeof(K) == 4 ? len * 4u : 0u), 0x00020000);
  auto store_key = [&](uint32_t pos, K v) {
    if constexpr (sizeof(K) == 4)
      __builtin_amdgcn_raw_buffer_store_b32((uint32_t)v, rs_out, pos * 4u, 0, 0);
    else
      out[pos] = v;
  };

  // Stable rank of every key within its unit.
  // Ranks (later staging slots) are below TILE <= 2^16; split tiles keep two
  // per register to leave room for the larger number of keys.
  constexpr bool PACK = SPLIT > 1;
  static_assert(!PACK || (ITEMS % 2 == 0 && TILE <= 65536), "packed ranks need an even item count");
  uint32_t rank[PACK ? ITEMS / 2 : ITEMS];
  auto set_rank = [&](int i, uint32_t v) {
    if constexpr (PACK) {
      if (i & 1) {
        rank[i >> 1] |= v << 16;
        asm volatile("" : "+v"(rank[i >> 1]));
      } else {
        rank[i >> 1] = v;
      }
    } else {
      rank[i] = v;
      // Keep the rank in one register instead of its two operands.
      asm volatile("" : "+v"(rank[i]));
    }
  };
  auto get_rank = [&](int i) -> uint32_t {
    if constexpr (PACK)
      return (i & 1) ? rank[i >> 1] >> 16 : rank[i >> 1] & 0xFFFFu;
    else
      return rank[i];
  };
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t d = dev::digit_bits<K, RB>(key[i], shift);
    if (UNI && __all(d == __builtin_amdgcn_readfirstlane(d))) {
      // Every lane holds the same digit (constant or presorted runs): the
      // peers of each unit are all its lanes, so no lane mask is needed and
      // the lanes do not serialise on one LDS atomic.
      if constexpr (UL == 32) {
        const uint32_t seen = uword[2 * d + 1];
        set_rank(i, seen + ul);
        if (ul == 0) uword[2 * d + 1] = seen + 32u;
      } else {
        const uint32_t seen = uslot16[d] >> 16;
        set_rank(i, seen + ul);
        if (ul == 0) uslot16[d] = (seen + 16u) << 16;
      }
    } else if constexpr (UL == 32) {
      atomicOr(&uword[2 * d], lanebit);
      const uint64_t sv = uslot[d];
      const uint32_t peers = (uint32_t)sv;
      const uint32_t seen = (uint32_t)(sv >> 32);
      const uint32_t below = __popc(peers & (lanebit - 1u));
      set_rank(i, seen + below);
      if (below == 0) uslot[d] = (uint64_t)(seen + __popc(peers)) << 32;
    } else {
      atomicOr(&uslot16[d], lanebit);
      const uint32_t sv = uslot16[d];
      const uint32_t peers = sv & 0xFFFFu;
      const uint32_t seen = sv >> 16;
      const uint32_t below = __popc(peers & (lanebit - 1u));
      set_rank(i, seen + below);
      if (below == 0) uslot16[d] = (seen + __popc(peers)) << 16;
    }
  }
  __syncthreads();

  // Per digit: prefix over units, tile count, publish the aggregate, scans
  // over the R digits of the tile counts (and, in SEG mode, of the
  // segment's digit counts).
  uint32_t cnt = 0, gval = 0, cinc = 0, ginc = 0;
  if (tid < R) {
#pragma unroll
    for (int u = 0; u < UNITS; ++u) {
      if constexpr (UL == 32) {
        const uint32_t c = hiword[2 * (u * R + tid)];
        hiword[2 * (u * R + tid)] = cnt;
        cnt += c;
      } else {
        const uint32_t c = slot16[u * R + tid] >> 16;
        slot16[u * R + tid] = cnt << 16;
        cnt += c;
      }
    }
    if (!(fmode & kPrefixReady))
      dev::atomic_store_agent(&lb[(size_t)row * R + tid], (local == 0 ? kLbInc : kLbAgg) | cnt);
    if constexpr (SEG) gval = g[tid];
    cinc = dev::wave_incl_scan(cnt, lane);
    if constexpr (SEG) ginc = dev::wave_incl_scan(gval, lane);
    if (lane == 63) {
      s_wsum[0][wave] = cinc;
      s_wsum[1][wave] = ginc;
    }
  }
  __syncthreads();

  // Decoupled look-back over earlier tiles of the array, LBW status words per
  // round trip; rows before the array's first tile read as "inclusive 0".
  uint32_t tstart = 0, gstart = 0, excl = 0;
  const uint32_t first_row = row - local;
  uint32_t p = row - 1;
  uint32_t w[LBW];
  auto poll = [&]() {
    const uint32_t avail = p - first_row + 1;
    const uint32_t* at = lb + (size_t)p * R + tid;  // one address, immediate offsets
#pragma unroll
    for (int j = 0; j < LBW; ++j) w[j] = (uint32_t)j < avail ? dev::atomic_load_agent(at - j * R) : kLbInc;
  };
  auto consume = [&]() {  // true once an inclusive prefix is reached
    uint32_t used = 0;
    bool stop = false, done = false;
#pragma unroll
    for (int j = 0; j < LBW; ++j) {
      if (!stop) {
        const uint32_t flag = w[j] & ~kLbMask;
        if (flag == 0) {
          stop = true;  // not published yet: poll again from here
        } else {
          excl += w[j] & kLbMask;
          ++used;
          if (flag == kLbInc) stop = done = true;
        }
      }
    }
    p -= used;
    if (!done && used == 0) __builtin_amdgcn_s_sleep(1);
    return done;
  };
  auto resolve = [&](bool polled) {
    if (local > 0) {
      if (!polled) poll();
      while (!consume()) poll();
      if (!(fmode & kPrefixReady)) dev::atomic_store_agent(&lb[(size_t)row * R + tid], kLbInc | (excl + cnt));
    }
    s_goff[tid] = gstart + excl - tstart;
  };
  if (tid < R) {
    uint32_t cpre = 0, gpre = 0;
    for (uint32_t wv = 0; wv < wave; ++wv) {
      cpre += s_wsum[0][wv];
      gpre += s_wsum[1][wv];
    }
    tstart = cpre + cinc - cnt;  // first staging slot of this digit
    gstart = SEG ? gpre + ginc - gval : g[tid];
#pragma unroll
    for (int u = 0; u < UNITS; ++u) {
      if constexpr (UL == 32)
        hiword[2 * (u * R + tid)] += tstart;
      else
        slot16[u * R + tid] += tstart << 16;
    }
    if (!LATE) resolve(false);
  }
  __syncthreads();

  // Staging slot of every key: its unit's start for the digit + its rank.  The
  // empty asm statements (which keep their order) make the digits be
  // recomputed here rather than kept live since the ranking, and bound the LDS
  // reads in flight to one group of 8 keys.
  constexpr int LG = ITEMS % 8 == 0 ? 8 : (ITEMS % 4 == 0 ? 4 : 2);
#pragma unroll
  for (int g0 = 0; g0 < ITEMS; g0 += LG) {
#pragma unroll
    for (int i = g0; i < g0 + LG && i < ITEMS; ++i) asm volatile("" : "+v"(key[i]));
#pragma unroll
    for (int i = g0; i < g0 + LG && i < ITEMS; ++i) {
      const uint32_t dg = dev::digit_bits<K, RB>(key[i], shift);
      const uint32_t start = UL == 32 ? hiword[2 * (unit * R + dg)] : uslot16[dg] >> 16;
      if constexpr (PACK)
        rank[i >> 1] += (i & 1) ? start << 16 : start;  // halves stay below 2^16: no carry
      else
        rank[i] += start;
    }
#pragma unroll
    for (int i = g0; i < g0 + LG && i < ITEMS; ++i) {
      if constexpr (PACK) {
        if (i & 1) asm volatile("" : "+v"(rank[i >> 1]));
      } else {
        asm volatile("" : "+v"(rank[i]));
      }
    }
  }
  __syncthreads();

  // Coalesced write-out of one staging round: consecutive slots of one digit
  // go to consecutive addresses of that digit's bin.  Groups of G items keep G
  // LDS reads in flight without holding every item's address in registers.
  constexpr int G = RITEMS % 8 == 0 ? 8 : (RITEMS % 4 == 0 ? 4 : 1);
  uint32_t gpos[KV ? ITEMS : 1];
  auto write_out = [&](auto fo, uint32_t r) {
#pragma unroll
    for (int g0 = 0; g0 < RITEMS; g0 += G) {
      K wk[G];
      uint32_t gp[G];
#pragma unroll
      for (int i = 0; i < G; ++i) wk[i] = s_keys[tid + (g0 + i) * BLOCK];
#pragma unroll
      for (int i = 0; i < G; ++i) {
        gp[i] = s_goff[dev::digit_bits<K, RB>(wk[i], shift)] + r * STAGE + tid + (g0 + i) * BLOCK;
        if constexpr (KV) gpos[g0 + i] = gp[i];
        if constexpr (decltype(fo)::value) wk[i] = dev::key_inv<K>(wk[i], xform);
      }
      if (valid == (uint32_t)TILE) {
#pragma unroll
        for (int i = 0; i < G; ++i) store_key(gp[i], wk[i]);
      } else {
#pragma unroll
        for (int i = 0; i < G; ++i)
          if (r * STAGE + tid + (g0 + i) * BLOCK < valid) store_key(gp[i], wk[i]);
      }
    }
  };
  // Scatter of items [i0, i1) into staging round r.
  auto scatter = [&](int i0, int i1, uint32_t r) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if (i < i0 || i >= i1) continue;
      if constexpr (SPLIT == 1) {
        s_keys[get_rank(i)] = key[i];
      } else {
        const uint32_t q = get_rank(i) - r * STAGE;
        if (q < (uint32_t)STAGE) s_keys[q] = key[i];
      }
    }
  };
#pragma unroll
  for (int r = 0; r < SPLIT; ++r) {
    if (r == 0) {
      // LATE: the first look-back round trip is issued halfway through the
      // scatter, when half of the keys and ranks no longer need registers.
      scatter(0, ITEMS / 2, 0);
      if (LATE && tid < R && local > 0) poll();
      scatter(ITEMS / 2, ITEMS, 0);
      if (LATE && tid < R) resolve(true);
    } else {
      scatter(0, ITEMS, r);
    }
    __syncthreads();
    if (float_out)
      write_out(std::true_type{}, r);
    else
      write_out(std::false_type{}, r);
    if (r + 1 < SPLIT) __syncthreads();
  }
  if constexpr (KV) {
    __syncthreads();
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_keys);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) s_vals[get_rank(i)] = val[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if (tid + i * BLOCK < valid) vdst[gpos[i]] = s_vals[tid + i * BLOCK];
  }
}


// Histogram read that also prepares pass 0's look-back rows: blocks take
// whole tiles of `tile` keys (pass 0's tiles), write each tile's digit-0
// counts to rows[t][R] and accumulate all D digit histograms into counts
// [D][R] (R = 2^RB, D = ceil(key bits / RB)).  Also clears the look-back rows
// of passes 1..D-1 (zero_vec 16-byte words at zero_ptr).
//
// Counting uses lane-private sub-histograms (packed 16-bit counters, copy
// lane % C of digit value v in word (v >> 1) * C + lane % C, half v & 1):
// every key costs D conflict-free LDS adds, each with a bit-field extract for
// the word and one multiply-add for the increment (1 or 1 << 16).  Digit 0's
// counters are never cleared between tiles: a tile's row is the difference of
// the running sums at its end and at the previous tile's end.  A block counts
// at most kHistTilesMaxKeys keys, so no packed counter can overflow (a copy
// of a bin sees at most 1/C of the block's keys) and the running sums are
// exact.
constexpr uint32_t kHistTilesMaxKeys = 1u << 19;

template <typename K, int RB = 8>
__global__ __launch_bounds__(kSortBlock) void k_hist_tiles(const K* __restrict__ keys, size_t n, uint32_t tile,
                                                           uint32_t* __restrict__ rows, uint32_t* __restrict__ counts,
                                                           uint4* __restrict__ zero_ptr, size_t zero_vec,
                                                           uint32_t xform) {
  constexpr int D = (8 * (int)sizeof(K) + RB - 1) / RB;
  constexpr uint32_t R = 1u << RB;
  constexpr int PER = 16 / sizeof(K);
  // Copies: 32 for u32 keys, 16 for u64 (8-bit digits: 4 x 16 KB or 8 x 8 KB of LDS).
  constexpr int C = sizeof(K) == 4 ? 32 : 16;
  constexpr int LOGC = C == 32 ? 5 : 4;
  using H = dev::SubHist<C, true>;
  constexpr uint32_t W = H::words(R);  // words per digit
  __shared__ __align__(16) uint32_t s_h[D * W];
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const uint32_t lane_c = lane % C;
  for (uint32_t i = tid; i < D * W / 4; i += kSortBlock) reinterpret_cast<uint4*>(s_h)[i] = make_uint4(0, 0, 0, 0);
  const size_t gstride = (size_t)gridDim.x * kSortBlock;
  for (size_t i = (size_t)blockIdx.x * kSortBlock + tid; i < zero_vec; i += gstride) zero_ptr[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  const bool aligned = ((uintptr_t)keys & 15) == 0 && (tile % PER) == 0;
  const size_t ntiles = (n + tile - 1) / tile;
  auto add_key = [&](K k) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      uint32_t hi, lo;
      if constexpr (sizeof(K) == 4) {
        // Bits past bit 31 read as zero, so a narrower last digit needs no mask.
        hi = __builtin_amdgcn_ubfe((uint32_t)k, RB * d + 1, RB - 1);
        lo = __builtin_amdgcn_ubfe((uint32_t)k, RB * d, 1);
      } else {
        const uint32_t dg = dev::digit_bits<K, RB>(k, RB * d);
        hi = dg >> 1;
        lo = dg & 1u;
      }
      atomicAdd(&s_h[d * W + ((hi << LOGC) | lane_c)], lo * 65535u + 1u);
    }
  };
  uint32_t run0 = 0;  // thread b < R: digit-0 count of bin b over the block's tiles so far
  for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const size_t lo = t * tile, hi = min(lo + tile, n);
    size_t scalar_from = lo;
    if (aligned) {
      const uint32_t nvec = (uint32_t)((hi - lo) / PER);
      const uint4* vk = reinterpret_cast<const uint4*>(keys + lo);
      // U 16-byte loads per thread in flight before any counting.
      constexpr int U = 4;
      for (uint32_t v0 = tid; v0 < nvec; v0 += U * kSortBlock) {
        uint4 x[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (v0 + u * kSortBlock < nvec) x[u] = vk[v0 + u * kSortBlock];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (v0 + u * kSortBlock >= nvec) continue;
          K k[PER];
          __builtin_memcpy(k, &x[u], 16);
#pragma unroll
          for (int j = 0; j < PER; ++j) add_key(xform ? dev::key_fwd<K>(k[j], xform) : k[j]);
        }
      }
      scalar_from = lo + (size_t)nvec * PER;
    }
    for (size_t i = scalar_from + tid; i < hi; i += kSortBlock) {
      K kk = keys[i];
      if (xform) kk = dev::key_fwd<K>(kk, xform);
      add_key(kk);
    }
    __syncthreads();
    if (tid < R) {
      const uint32_t c = H::sum(s_h, tid);
      rows[t * R + tid] = c - run0;
      run0 = c;
    }
    // The next tile's adds must not reach digit 0's counters before every
    // thread has read them.
    __syncthreads();
  }
  if (tid < R) {
    if (run0) atomicAdd(&counts[tid], run0);
#pragma unroll
    for (int d = 1; d < D; ++d) {
      const uint32_t c = H::sum(s_h + d * W, tid);
      if (c) atomicAdd(&counts[d * R + tid], c);
    }
  }
}

}  // namespace pa

#include <type_traits>
// Tree growth on the device: histograms, split search, partition and the
// round driver (included by engine.hip after kernels.hip).
//
// The reference grows a tree one leaf at a time (SerialTreeLearner::train,
// learner/mod.rs:572-695): ArgMax over the leaves' best splits (:698-706),
// split that leaf, build the smaller child's histogram, derive the larger
// one by subtraction, search both children. Only the ArgMax is sequential:
// a leaf's histogram and best split depend on that leaf's rows alone. This
// engine therefore grows in rounds. A round splits the k best leaves of the
// frontier together (one partition, one histogram and one search launch for
// all k), and the next round opens by replaying the reference's sequential
// ArgMax over the results, in order, for as long as every split it picks has
// been computed. The replayed tree - node numbering included - is the one the
// sequential loop builds. Splits computed ahead of need that the tree never
// reaches are dropped (their leaves keep their rows).
//
// Leaves live in slots (0 = root; a round allocates two child slots per
// split). The tree's own leaf numbering is the position of a slot in the
// current position array, double-buffered so that every block of a launch
// reads the same version while block 0 writes the next.

namespace lgbmhip {

constexpr int kPartThreads = 256;
#ifndef LGBMHIP_PART_ROWS
#define LGBMHIP_PART_ROWS 4
#endif
constexpr int kPartRowsPerThread = LGBMHIP_PART_ROWS;
constexpr int kPartTile = kPartThreads * kPartRowsPerThread;
constexpr int kRowThreads = kPartThreads / (2 * kMaxBatch);   // threads reducing one searched row
static_assert(kRowThreads >= 1, "a searched row needs at least one thread");
// The plan's selection runs on the replay's wave (step 3a) once the tree has
// more than kWaveSelMin leaves, choosing at most kWaveSelMax leaves.
constexpr int kWaveSelMin = 96;
constexpr int kWaveSelMax = 16;

__device__ __forceinline__ int child_buf(int buf) { return buf == 1 ? 2 : 1; }
template <typename T>
__device__ __forceinline__ T sel3(T a0, T a1, T a2, int b) { return b == 0 ? a0 : (b == 1 ? a1 : a2); }

// before_find_best_split (learner/mod.rs:712-731) for the two children of a
// split: no search when the children reach max_depth or both are too small.
__device__ __forceinline__ bool children_skip(const SplitParamsDev& p, int child_depth, long long lc, long long rc) {
  return (p.max_depth > 0 && child_depth >= p.max_depth) ||
         (rc < 2LL * p.min_data_in_leaf && lc < 2LL * p.min_data_in_leaf);
}

// ---------------------------------------------------------------- tree state
// learner/mod.rs:580-628: one leaf (the root) holding every row, searched first.
__global__ void k_state_init(StepDev* st, PosRec* pos, LeafDev* leaves, uint8_t* splittable, long long n, int F) {
  const int t = threadIdx.x;
  for (int f = t; f < F; f += blockDim.x) splittable[f] = 1;
  if (t < kMaxBatch) {
    st->lcur[t] = 0;
    st->rcur[t] = 0;
  }
  if (t == 0) {
    st->done = 0;
    st->nseq = 1;
    st->npending = 0;
    st->nslots = 1;
    st->num_splits = 0;
    st->parity = 0;
    st->search_base = 0;
    st->nsearched = 1;
    st->rounds = 0;
    st->hist_items = 1;
    st->hist_rows = n;
    st->part_rows = 0;
    st->computed = 0;
    PosRec p;
    p.gain = -INFINITY;
    p.cgain[0] = p.cgain[1] = -INFINITY;
    p.feat = -1;
    p.slot = 0;
    p.spec = -1;
    p.cfeat[0] = p.cfeat[1] = -1;
    p.pad = 0;
    pos[0] = p;
    LeafDev& r = leaves[0];
    r.begin = 0;
    r.count = n;
    r.buf = 0;
    r.depth = 0;
    r.sum_g = r.sum_h = 0.0;
    r.output = 0.0;
    r.int_sum = 0;
  }
}

// ---------------------------------------------------------------- histograms
// Row-wise histograms (multi_val_bin.rs:556-611 FP64, :710-791 integer).
// FP64 values are summed as exact fixed-point int64 (ds_add_u64 in LDS):
// exact, so the result is independent of order and block count. The
// quantized form's result is the reference's own wrapping sum of packed
// (g << 32) | u8(h) rows; it accumulates in 32-bit LDS bins (16-bit fields,
// folded to int64 before they can overflow) and falls back to one wrapping
// ds_add_u64 per bin when a row's values are too wide for that.
//
// Every leaf's rows are one contiguous segment of an ordered buffer (buffer
// 0 holds the root in row order; a leaf in buffer 1 splits into buffer 2 and
// vice versa), each feature group as its own [N][FB]-byte array: a block
// streams its group's bytes of its rows with no index indirection.
//
// Work is a list of items (segments); each item is cut into row chunks, and
// a unit = (item, chunk, feature group) is one LDS histogram. Blocks loop
// over units; each unit stores its whole LDS histogram to its own slab with
// plain coalesced stores (no global atomics), and the split pass sums an
// item's slabs - integer sums, exact in any order.
struct HistArgs {
  const uint8_t* grp[3];   // ordered bins per buffer, group-major [NG][N][FB]
  const void* gh[3];       // float2 (plain) or char2 (quantized), per buffer
  StepDev* st;
  const PlanDev* plan;
  const PlanRep* reps;
  LeafDev* leaves;
  SplitRec* log;
  const SplitRec* best;
  const FeatDev* feats;
  const ByteDev* bytes;    // [NG * FB] each byte of a packed row (its fields' entries, masks, most frequent bins)
  const int* ginfo;        // [NG][2] each group's histogram entries, whether it holds 4-bit fields
  int F, B, FB, NG;        // features, max bins, bytes per group, groups
  int pu;                  // entries of a unit (the largest group's)
  int u4;                  // 1: some byte holds two 4-bit fields (the kernels' U4 form)
  int L;
  long long N;
  int max_chunks;
  int min_rows_per_chunk;
  long long* slab;         // per unit: its group's [pu][C] sums, C = 2 (plain) or 1 (quantized)
  int* units;              // [kMaxBatch + 2] first unit of each item, the total, the item count
  const IterDev* it;
  SplitParamsDev p;
  int quant;
  int mode;                // 1: explicit segment, 2: the root, 3: the round's items
  long long root_n;        // rows of the root (all rows, or the bag)
  long long d_begin, d_count;
  int d_buf, d_shift_g, d_shift_h;
  int q_host;              // 1: the quantized maxima below (an explicit segment), 0: the tree's, from it
  int q_gmax, q_hmax;      // max |gq| and max u8(hq) of the segment's rows
};

// rows each thread has in flight per pass (all loads issued before the atomics)
constexpr int kHistUnroll = 8;
// Threads of a histogram block: the quantized form's 512-thread blocks take
// 4096-row chunks, so a unit's slab (one histogram per chunk) stays a small
// share of the bytes it reads; the FP64 form's wider LDS histogram keeps 256.
template <bool Q>
struct HistBlock { static constexpr int threads = Q ? 512 : 256; };

// One unit: rows [r0, r1) of the segment's chunk, feature group fg, into LDS,
// then stored to the unit's slab (an empty chunk stores zeros). Lane j of a
// row takes the group's bytes 4j .. 4j + 3; HI: the group holds pairs of
// 4-bit fields (a byte then adds its high field too).
template <int TPR, bool Q, bool HI>
__device__ __forceinline__ void hist_unit(const HistArgs& a, void* ldsv, long long begin, long long count, int buf,
                                          int sg, int sh, int chunk, int nchunks, int fg, long long* out_slab) {
  const long long per = (count + nchunks - 1) / nchunks;
  const long long r0 = (long long)chunk * per, r1 = min(count, r0 + per);
  // quantized: one packed (g << 32) + h int64 per bin; FP64: g and h int64
  const int n_lds = a.ginfo[2 * fg] * (Q ? 1 : 2);
  constexpr bool u4g = HI;
  unsigned long long* lds64 = (unsigned long long*)ldsv;
  for (int e = threadIdx.x; e < n_lds; e += blockDim.x) lds64[e] = 0ull;
  // this lane's 4 bytes: entries, masks and most frequent bins of their fields
  const int sub = threadIdx.x / TPR, j = threadIdx.x % TPR;
  // (a byte's high field's entries follow its low field's 16)
  const bool active = j * 4 < a.FB;
  int off_lo[4];
  unsigned mask_lo[4], mfb_lo[4];
  unsigned himask = 0u, mfb_h = 0u;   // the high fields: present (bit k), 4-bit most frequent bins
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    ByteDev bd{0, 0, 0u, 0u, 0u, 0u};
    if (active) bd = a.bytes[fg * a.FB + j * 4 + k];
    off_lo[k] = bd.off_lo;
    mask_lo[k] = bd.mask_lo;
    mfb_lo[k] = bd.mfb_lo;
    himask |= (bd.mask_hi ? 1u : 0u) << k;
    mfb_h |= (bd.mfb_hi & 0xFu) << (4 * k);
  }
  __syncthreads();
  const int rpp = HistBlock<Q>::threads / TPR;
  // select, not index: a runtime index into a kernel-argument array makes
  // the compiler copy the arguments to scratch memory
  // lanes past the group's last byte read the row's first word (unused)
  const uint8_t* __restrict__ gb =
      sel3(a.grp[0], a.grp[1], a.grp[2], buf) + ((size_t)fg * a.N + begin) * a.FB + (active ? j * 4 : 0);
  using GT = typename std::conditional<Q, char2, float2>::type;
  const GT* __restrict__ ghb = (const GT*)sel3(a.gh[0], a.gh[1], a.gh[2], buf) + begin;
  for (long long base = r0; base < r1; base += kHistUnroll * rpp) {
    unsigned w[kHistUnroll];
    GT graw[kHistUnroll];
    long long gv[kHistUnroll], hv[kHistUnroll];
    bool ok[kHistUnroll];
    // every slot loads (one past the chunk reloads its last row), all loads
    // before any use, so a thread's loads are in flight together
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      const long long p = base + u * rpp + sub;
      ok[u] = active && p < r1;
      const long long pc = p < r1 ? p : r1 - 1;
      w[u] = *(const unsigned*)(gb + (size_t)pc * a.FB);
      graw[u] = ghb[pc];
    }
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      if (Q) {
        // multi_val_bin.rs:710-791: the row's packed (g << 32) | u8(h)
        const char2 v = *(const char2*)&graw[u];
        gv[u] = (long long)(((unsigned long long)(long long)v.x << 32) | (unsigned long long)(uint8_t)v.y);
        hv[u] = 0;
      } else {
        const float2 v = *(const float2*)&graw[u];
        gv[u] = to_fixed(v.x, sg);
        hv[u] = to_fixed(v.y, sh);
      }
    }
    auto add = [&](int e, long long g, long long h) {
      if (Q) {
        atomicAdd(lds64 + e, (unsigned long long)g);   // wrapping, as the reference
      } else {
        atomicAdd(lds64 + 2 * e, (unsigned long long)g);
        atomicAdd(lds64 + 2 * e + 1, (unsigned long long)h);
      }
    };
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      if (!ok[u]) continue;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned by = (w[u] >> (8 * k)) & 0xFFu;
        const unsigned b = by & mask_lo[k];
        if (mask_lo[k] && b != mfb_lo[k]) add(off_lo[k] + (int)b, gv[u], hv[u]);
        if (u4g) {
          const unsigned bh = by >> 4;
          if (((himask >> k) & 1u) && bh != ((mfb_h >> (4 * k)) & 0xFu)) add(off_lo[k] + 16 + (int)bh, gv[u], hv[u]);
        }
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n_lds; e += blockDim.x) out_slab[e] = (long long)lds64[e];
  __syncthreads();   // the next unit reuses the LDS histogram
}

// The quantized form in 32-bit LDS bins. A row adds gq * 65536 + u8(hq)
// with one wrapping ds_add_u32. While an entry's sums stay within
// |sum gq| <= 32767 and sum u8(hq) <= 65535, the 32-bit word holds them
// exactly: as a signed int it equals sum_gq * 65536 + sum_hq, so its low 16
// bits are sum_hq and its arithmetic shift right by 16 is sum_gq. Rows are
// taken in epochs short enough for that bound (from the rows' max |gq| and
// max u8(hq)); after each epoch a thread folds its entries into an int64
// LDS copy as (sum_gq << 32) + sum_hq, the reference's packed sum of the
// same rows, and clears them. The unit then stores the int64 copy to its
// slab, the same entries as the 64-bit form.
//
// LDS of a block: [n32 32-bit bins][64 trash words][n32 int64 sums], n32 =
// the entries of a unit, the int64 part 8-byte aligned.
__host__ __device__ inline size_t q32_acc_offset(int n32) { return ((size_t)(n32 + 64) * 4 + 7) / 8 * 8; }
__host__ __device__ inline size_t q32_lds_bytes(int n32) { return q32_acc_offset(n32) + (size_t)n32 * 8; }

// units[kUnitForm + w]: 1 when item w's slabs hold packed 32-bit bins (the
// quantized form within one epoch), 0 for int64 entries.
constexpr int kUnitForm = kMaxBatch + 2;

// Rows per epoch: the most rows whose sums cannot overflow a 16-bit field.
__device__ __forceinline__ int q32_epoch_rows(unsigned gmax, unsigned hmax) {
  const unsigned g = max(gmax, 1u), h = max(hmax, 1u);
  return (int)min(32767u / g, 65535u / h);
}

// A wave-uniform copy of a 64-bit value (kept in scalar registers, so loads
// from a uniform base take 32-bit per-lane offsets).
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// The row loop has no branches per bin: a bin equal to the feature's most
// frequent bin (never accumulated), and every field the lane does not hold,
// adds to a per-lane trash entry past the histogram; rows past the chunk's
// end (the last pass) add 0. HI: the group holds pairs of 4-bit fields (a
// byte then adds its high field too).
template <int TPR, bool HI>
__device__ __forceinline__ void hist_unit_q32(const HistArgs& a, unsigned* lds32, long long begin, long long count,
                                              int buf, int chunk, int nchunks, int fg, long long* out_slab,
                                              int epoch_rows) {
  constexpr int T = HistBlock<true>::threads;
  constexpr int rpp = T / TPR;
  constexpr int PR = kHistUnroll * rpp;   // rows per pass
  const long long per = (count + nchunks - 1) / nchunks;
  const long long r0 = (long long)chunk * per, r1 = min(count, r0 + per);
  const int rows = (int)max(0LL, r1 - r0);
  const int n_lds = a.ginfo[2 * fg];
  constexpr bool u4g = HI;
  const int n32 = a.pu;
  const int sub = threadIdx.x / TPR, j = threadIdx.x % TPR;
  const unsigned trash = (unsigned)(n32 + (threadIdx.x & 63)) * 4u;   // byte offset
  // this thread's entries e = threadIdx.x + i * T: only it reads and writes
  // their int64 sums, so those need no barrier
  unsigned long long* acc = (unsigned long long*)((char*)lds32 + q32_acc_offset(n32));
  // an item whose chunks fit one epoch stores its 32-bit bins to the slab
  // as they are (the reduction unpacks them), decided by the item's chunk
  // length so that every chunk of the item has the same form
  const bool one_epoch = per <= epoch_rows;
  const int epoch_passes = epoch_rows / PR;
  // this lane's 4 bytes: the entries and most frequent bins of their fields;
  // a field's mask keeps its bits of the row's word (a masked-off field
  // reads 0 against most frequent bin 0: the trash)
  // (a byte's high field's entries follow its low field's 16)
  const bool active = j * 4 < a.FB;
  unsigned fbase[4], mfb[4], bytemask = 0u;
  unsigned himask = 0u, mfb_h = 0u;   // the high fields: bits of the row's word, 4-bit most frequent bins
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    ByteDev bd{0, 0, 0u, 0u, 0u, 0u};
    if (active) bd = a.bytes[fg * a.FB + j * 4 + k];
    fbase[k] = (unsigned)bd.off_lo * 4u;
    mfb[k] = bd.mfb_lo;
    bytemask |= bd.mask_lo << (8 * k);
    himask |= (bd.mask_hi << 4) << (8 * k);
    mfb_h |= (bd.mfb_hi & 0xFu) << (4 * k);
  }
  if (!one_epoch)
    for (int e = threadIdx.x; e < n_lds; e += T) acc[e] = 0ull;
  // folds the block's 32-bit entries into acc and clears them (the 32-bit
  // histogram is all zero between units)
  auto fold = [&]() {
    for (int e = threadIdx.x; e < n_lds; e += T) {
      const unsigned x = lds32[e];
      lds32[e] = 0u;
      const long long sg = (long long)((int)x >> 16);
      acc[e] += ((unsigned long long)sg << 32) + (unsigned long long)(x & 0xFFFFu);
    }
  };
  const int FB = a.FB;
  const unsigned jofs = active ? (unsigned)(j * 4) : 0u;
  typedef const __attribute__((address_space(1))) uint8_t* GBytes;
  typedef const __attribute__((address_space(1))) unsigned short* GHalf;
  typedef __attribute__((address_space(3))) unsigned* LWord;
  const GBytes gu = (GBytes)uniform_u64(
      (unsigned long long)(sel3(a.grp[0], a.grp[1], a.grp[2], buf) + ((size_t)fg * a.N + begin + r0) * FB));
  const GHalf hu = (GHalf)uniform_u64(
      (unsigned long long)((const char2*)sel3(a.gh[0], a.gh[1], a.gh[2], buf) + begin + r0));
  // LDS byte addresses, the histogram's base folded in
  const unsigned lbase = (unsigned)(size_t)(LWord)lds32;
#pragma unroll
  for (int k = 0; k < 4; ++k) fbase[k] += lbase;
  const unsigned trash_a = trash + lbase;
  // one pass's loads (a row past the chunk reloads its last row)
  auto load_pass = [&](int q0, unsigned* w, unsigned* hr) {
    const bool full = q0 + PR <= rows;
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      int q = q0 + u * rpp + sub;
      if (!full) q = min(q, rows - 1);
      w[u] = *(const __attribute__((address_space(1))) unsigned*)(gu + ((unsigned)q * (unsigned)FB + jofs));
      hr[u] = hu[(unsigned)q];
    }
  };
  int passes = 0;
  // the next pass's loads are issued before this pass's atomics
  unsigned w[kHistUnroll], hr[kHistUnroll], wn[kHistUnroll], hn[kHistUnroll];
  if (rows > 0) load_pass(0, w, hr);
  for (int q0 = 0; q0 < rows; q0 += PR) {
    const bool full = q0 + PR <= rows;
    load_pass(q0 + PR, wn, hn);   // past the end: the chunk's last row again (no branch, so loads stay counted)
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      // multi_val_bin.rs:710-791 with the row's (gq, hq) as gq * 65536 + u8(hq)
      const unsigned v0 = ((unsigned)(int)(signed char)(hr[u] & 0xFFu) << 16) + (hr[u] >> 8);
      const unsigned v = (full || q0 + u * rpp + sub < rows) ? v0 : 0u;
      const unsigned ww = w[u] & bytemask;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned b = (ww >> (8 * k)) & 0xFFu;
        const unsigned addr = b == mfb[k] ? trash_a : fbase[k] + b * 4u;
        __hip_atomic_fetch_add((LWord)(size_t)addr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);   // wrapping
      }
      if (u4g) {
        const unsigned wh = w[u] & himask;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const unsigned b = (wh >> (8 * k + 4)) & 0xFu;
          const unsigned addr = b == ((mfb_h >> (4 * k)) & 0xFu) ? trash_a : fbase[k] + (16u + b) * 4u;
          __hip_atomic_fetch_add((LWord)(size_t)addr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    }
    // the loop bounds are the block's own, so every thread takes this branch alike
    if (!one_epoch && ++passes == epoch_passes && q0 + PR < rows) {
      __syncthreads();
      fold();
      __syncthreads();
      passes = 0;
    }
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      w[u] = wn[u];
      hr[u] = hn[u];
    }
  }
  __syncthreads();
  if (one_epoch) {
    unsigned* out32 = (unsigned*)out_slab;
    for (int e = threadIdx.x; e < n_lds; e += T) {
      out32[e] = lds32[e];
      lds32[e] = 0u;
    }
    __syncthreads();   // the next unit's atomics start on a cleared histogram
    return;
  }
  fold();
  __syncthreads();   // the next unit's atomics start on a cleared histogram
  for (int e = threadIdx.x; e < n_lds; e += T) out_slab[e] = (long long)acc[e];
}

// Applies the round's plan to the tree state (block 0 of the histogram pass,
// the first launch after the partition): the children's leaf records
// (learner/mod.rs:1105-1225), the replayed splits' log entries, and the
// step record.
__device__ void apply_plan(const HistArgs& a) {
  StepDev* st = a.st;
  const PlanDev* P = a.plan;
  const int tid = threadIdx.x;
  const int nw = P->nwork;
  const int nspl0 = st->num_splits;
  for (int w = tid; w < nw; w += blockDim.x) {
    const PlanItem& it = P->item[w];
    const long long lc = (long long)st->lcur[w], rc = it.count - lc;
    const int nb = child_buf(it.buf);
    LeafDev& Lf = a.leaves[it.lslot];
    LeafDev& Rt = a.leaves[it.lslot + 1];
    Lf.begin = it.begin;
    Lf.count = lc;
    Lf.buf = nb;
    Lf.depth = it.depth + 1;
    Lf.sum_g = it.rec.lsg;
    Lf.sum_h = it.rec.lsh;
    Lf.int_sum = it.rec.lgh;
    Lf.output = isnan(it.rec.left_output) ? 0.0 : it.rec.left_output;
    Rt.begin = it.begin + lc;
    Rt.count = rc;
    Rt.buf = nb;
    Rt.depth = it.depth + 1;
    Rt.sum_g = it.rec.rsg;
    Rt.sum_h = it.rec.rsh;
    Rt.int_sum = it.rec.rgh;
    Rt.output = isnan(it.rec.right_output) ? 0.0 : it.rec.right_output;
  }
  // replayed splits: the parent's best split, applied to its position
  constexpr int RW = (int)(sizeof(SplitRec) / 4);
  for (int e = tid; e < P->nrep * RW; e += blockDim.x) {
    const int i = e / RW, wd = e % RW;
    const PlanRep rp = a.reps[i];
    unsigned* dst = (unsigned*)(a.log + nspl0 + i);
    dst[wd] = ((const unsigned*)(a.best + rp.parent))[wd];
  }
  __syncthreads();
  for (int i = tid; i < P->nrep; i += blockDim.x) {
    const PlanRep rp = a.reps[i];
    SplitRec* d = a.log + nspl0 + i;
    d->leaf = rp.pos;
    d->left_count = (int)a.leaves[rp.lslot].count;
    d->right_count = (int)a.leaves[rp.lslot + 1].count;
  }
  if (tid == 0) {
    long long moved = 0;
    for (int w = 0; w < nw; ++w) moved += P->item[w].count;
    st->part_rows += moved;
    st->computed += nw;
    st->nseq = P->nseq;
    st->npending = P->npending;
    st->search_base = P->nslots - 2 * nw;
    st->nsearched = 2 * nw;
    st->nslots = P->nslots;
    st->num_splits = P->num_splits;
    st->parity ^= 1;
    st->rounds += 1;
    st->done = P->done;
  }
}

template <int TPR, bool Q, bool U4>
__device__ __forceinline__ void hist_kernel(const HistArgs& a, void* lds) {
  __shared__ long long s_begin[kMaxBatch], s_count[kMaxBatch];
  __shared__ int s_buf[kMaxBatch], s_sg[kMaxBatch], s_sh[kMaxBatch], s_nch[kMaxBatch], s_u0[kMaxBatch + 1];
  const int tid = threadIdx.x;
  const size_t ustride = (size_t)a.pu * (Q ? 1 : 2);   // one unit's slab
  int nitems;
  if (a.mode != 3) {
    if (tid == 0) {
      const bool root = a.mode == 2;
      const long long cnt = root ? a.root_n : a.d_count;
      s_begin[0] = root ? 0 : a.d_begin;
      s_count[0] = cnt;
      s_buf[0] = root ? 0 : a.d_buf;
      s_sg[0] = root ? (Q ? 0 : fix_shift(cnt, a.it->max_abs_g_bits)) : a.d_shift_g;
      s_sh[0] = root ? (Q ? 0 : fix_shift(cnt, a.it->max_abs_h_bits)) : a.d_shift_h;
      const long long want = (cnt + a.min_rows_per_chunk - 1) / a.min_rows_per_chunk;
      s_nch[0] = (int)min((long long)a.max_chunks, max(want, 1LL));
      s_u0[0] = 0;
      s_u0[1] = s_nch[0] * a.NG;
      if (blockIdx.x == 0) {
        a.units[0] = 0;
        a.units[1] = s_u0[1];
        a.units[kMaxBatch + 1] = 1;
      }
    }
    nitems = 1;
  } else {
    StepDev* st = a.st;
    const PlanDev* P = a.plan;
    // block 0 applies the plan once: the round that finishes the tree still
    // replays its last splits; later rounds find the tree done
    if (blockIdx.x == 0) {
      __shared__ int s_was_done;
      if (tid == 0) s_was_done = st->done;
      __syncthreads();
      if (!s_was_done) apply_plan(a);
    }
    if (P->done) {
      if (blockIdx.x == 0 && tid == 0) a.units[kMaxBatch + 1] = 0;   // nothing for the reduction
      return;
    }
    nitems = P->nwork;
    if (tid < nitems) {
      const PlanItem& it = P->item[tid];
      const long long lc = (long long)st->lcur[tid], rc = it.count - lc;
      const bool run = !P->complete && !children_skip(a.p, it.depth + 1, lc, rc);
      const long long sc = lc < rc ? lc : rc;
      s_begin[tid] = lc < rc ? it.begin : it.begin + lc;
      s_count[tid] = run ? sc : 0;
      s_buf[tid] = child_buf(it.buf);
      s_sg[tid] = Q ? 0 : fix_shift(sc, a.it->max_abs_g_bits);
      s_sh[tid] = Q ? 0 : fix_shift(sc, a.it->max_abs_h_bits);
      const long long want = (sc + a.min_rows_per_chunk - 1) / a.min_rows_per_chunk;
      s_nch[tid] = run ? (int)min((long long)a.max_chunks, max(want, 1LL)) : 0;
    }
    __syncthreads();
    if (tid == 0) {
      int u = 0;
      long long rows = 0;
      int built = 0;
      for (int w = 0; w < nitems; ++w) {
        s_u0[w] = u;
        u += s_nch[w] * a.NG;
        rows += s_count[w];
        built += s_nch[w] > 0 ? 1 : 0;
      }
      s_u0[nitems] = u;
      if (blockIdx.x == 0) {
        for (int w = 0; w <= nitems; ++w) a.units[w] = s_u0[w];
        a.units[kMaxBatch + 1] = nitems;
        // work counters for profiles (read by the host after the tree)
        st->hist_rows += rows;
        st->hist_items += built;
      }
    }
  }
  __syncthreads();
  const int units = s_u0[nitems];
  // block 0 applied the plan: in a round it takes the last unit share, so
  // the other blocks start on the histograms while it works
  const int vb = a.mode == 3 ? (blockIdx.x == 0 ? (int)gridDim.x - 1 : (int)blockIdx.x - 1) : (int)blockIdx.x;
  int w = 0;
  if (Q) {
    constexpr int T = HistBlock<true>::threads;
    const int erows = q32_epoch_rows(a.q_host ? (unsigned)a.q_gmax : a.it->qmax_g,
                                     a.q_host ? (unsigned)a.q_hmax : a.it->qmax_h);
    const bool q32 = erows >= kHistUnroll * (T / TPR);   // at least one pass per epoch
    if (blockIdx.x == 0 && tid < nitems) {
      const long long per = s_nch[tid] > 0 ? (s_count[tid] + s_nch[tid] - 1) / s_nch[tid] : 0;
      a.units[kUnitForm + tid] = q32 && per <= erows ? 1 : 0;
    }
    if (q32) {
      unsigned* lds32 = (unsigned*)lds;
      for (int e = tid; e < a.pu; e += T) lds32[e] = 0u;
      __syncthreads();
      for (int u = vb; u < units; u += gridDim.x) {
        while (u >= s_u0[w + 1]) ++w;
        const int r = u - s_u0[w], fg = r % a.NG;
        // a group holding 4-bit fields takes the loop that adds both fields of a byte
        if (U4 && a.ginfo[2 * fg + 1])
          hist_unit_q32<TPR, true>(a, lds32, s_begin[w], s_count[w], s_buf[w], r / a.NG, s_nch[w], fg,
                                   a.slab + (size_t)u * ustride, erows);
        else
          hist_unit_q32<TPR, false>(a, lds32, s_begin[w], s_count[w], s_buf[w], r / a.NG, s_nch[w], fg,
                                    a.slab + (size_t)u * ustride, erows);
      }
      return;
    }
  }
  if (!Q && blockIdx.x == 0 && tid < nitems) a.units[kUnitForm + tid] = 0;
  for (int u = vb; u < units; u += gridDim.x) {
    while (u >= s_u0[w + 1]) ++w;
    const int r = u - s_u0[w], fg = r % a.NG;
    if (U4 && a.ginfo[2 * fg + 1])
      hist_unit<TPR, Q, true>(a, lds, s_begin[w], s_count[w], s_buf[w], s_sg[w], s_sh[w], r / a.NG, s_nch[w], fg,
                              a.slab + (size_t)u * ustride);
    else
      hist_unit<TPR, Q, false>(a, lds, s_begin[w], s_count[w], s_buf[w], s_sg[w], s_sh[w], r / a.NG, s_nch[w], fg,
                               a.slab + (size_t)u * ustride);
  }
}

// five 256-thread blocks per CU's SIMDs (at most 96 VGPRs; the 4-bit form's
// second field per byte would otherwise take it to four)
template <int TPR, bool U4>
__global__ void __launch_bounds__(HistBlock<false>::threads) __attribute__((amdgpu_waves_per_eu(5)))
k_hist_f64(HistArgs a) {
  extern __shared__ unsigned long long lds_hist[];
  hist_kernel<TPR, false, U4>(a, lds_hist);
}

// three 512-thread blocks per CU (6 waves per SIMD: at most 80 VGPRs), the
// LDS of three 48 KiB histograms
template <int TPR, bool U4>
__global__ void __launch_bounds__(HistBlock<true>::threads) __attribute__((amdgpu_waves_per_eu(6)))
k_hist_int(HistArgs a) {
  extern __shared__ unsigned long long lds_hist[];
  hist_kernel<TPR, true, U4>(a, lds_hist);
}

// The trees of a batch: grid y picks the tree, grid x is its blocks.
template <int TPR, bool U4>
__global__ void __launch_bounds__(HistBlock<false>::threads) __attribute__((amdgpu_waves_per_eu(5)))
k_hist_f64_tab(const HistArgs* __restrict__ tab) {
  extern __shared__ unsigned long long lds_hist[];
  const HistArgs a = tab[blockIdx.y];
  hist_kernel<TPR, false, U4>(a, lds_hist);
}
template <int TPR, bool U4>
__global__ void __launch_bounds__(HistBlock<true>::threads) __attribute__((amdgpu_waves_per_eu(6)))
k_hist_int_tab(const HistArgs* __restrict__ tab) {
  extern __shared__ unsigned long long lds_hist[];
  const HistArgs a = tab[blockIdx.y];
  hist_kernel<TPR, true, U4>(a, lds_hist);
}

// ---------------------------------------------------------------- split search
struct SplitArgs {
  StepDev* st;
  const PlanDev* plan;
  LeafDev* leaves;
  const FeatDev* feats;
  SplitParamsDev p;
  int F, B, L;
  const long long* slab;  // histogram units of the last histogram pass
  const int* units;       // [kMaxBatch + 1] first unit of each item
  int FB, NG;             // bytes per histogram group, groups
  int pu;                 // entries of a histogram unit (a feature's bins start at its hoff in unit hg)
  double* pool;          // [slots][F][B][2] FP64 histograms
  long long* ipool;      // [slots][F][B] packed histograms (quantized)
  SplitRec* feat_split;  // [2 kMaxBatch][F] searches of this pass, row = slot - first child slot
  uint8_t* splittable;   // [slots][F]
  IterDev* it;
  int quant;
  int root;              // 1: search the root (slot 0, row 0)
  int raw;               // 1: slab and units are the histogram pass's own (chunk units, a 32-bit form per item)
};

// The leaf sums a search needs (leaf_splits.hpp; the root's come from the
// fixed-point / packed root sums).
__device__ __forceinline__ LeafSums leaf_sums(const SplitArgs& a, int slot, bool root) {
  LeafSums s;
  const LeafDev& L = a.leaves[slot];
  s.nd = (int)L.count;
  s.isum = L.int_sum;
  s.gs = a.it->grad_scale;
  s.hs = a.it->hess_scale;
  double sg = L.sum_g, sh = L.sum_h;
  if (root) {
    if (a.quant && a.it->full_bag) {
      s.isum = a.it->bag_isum;
      sg = ldexp((double)a.it->bag_sg_fix, a.it->bag_eg);
      sh = ldexp((double)a.it->bag_sh_fix, a.it->bag_eh);
    } else if (a.quant) {
      s.isum = a.it->root_int_sum;
      sg = (double)a.it->root_gq_sum * s.gs;
      sh = (double)a.it->root_hq_sum * s.hs;
    } else {
      sg = ldexp((double)a.it->root_g_fix, -fix_shift(L.count, a.it->max_abs_g_bits));
      sh = ldexp((double)a.it->root_h_fix, -fix_shift(L.count, a.it->max_abs_h_bits));
    }
  }
  s.sg = sg;
  s.sh = sh + 2.0 * KEPS;
  // learner/mod.rs:1303-1309: the root's parent output is its own output
  s.po = root ? leaf_output_t(sg, sh, a.p, true, true, false, s.nd, 0.0) : L.output;
  return s;
}

// The sums of a child of a split of the round, as apply_plan records them in
// the child's leaf (cnt rows).
__device__ __forceinline__ LeafSums child_sums(const SplitArgs& a, const SplitRec& rec, long long cnt, bool right) {
  LeafSums s;
  s.nd = (int)cnt;
  s.isum = right ? rec.rgh : rec.lgh;
  s.gs = a.it->grad_scale;
  s.hs = a.it->hess_scale;
  s.sg = right ? rec.rsg : rec.lsg;
  s.sh = (right ? rec.rsh : rec.lsh) + 2.0 * KEPS;
  const double o = right ? rec.right_output : rec.left_output;
  s.po = isnan(o) ? 0.0 : o;
  return s;
}

// Sum of feature f's bins over the histogram units of item w (all chunks of
// f's group): out0 (and out1 for C = 2, the hessian sums) per bin b = tid +
// 128 k. Wrapping int64 sums: exact in any order.
template <int C>
__device__ __forceinline__ void slab_sum(const SplitArgs& a, int w, const FeatDev& fd, int tid, long long* out0,
                                         long long* out1) {
  constexpr int PB = kMaxBins / 128;
  const int u0 = a.units[w], nch = (a.units[w + 1] - u0) / a.NG;
  const size_t ustride = (size_t)a.pu * C;
  const long long* base = a.slab + (size_t)(u0 + fd.hg) * ustride + (size_t)fd.hoff * C;
  const int nb = fd.num_bin;   // a 4-bit feature's unit holds its 16 bins only
  const size_t cstride = (size_t)a.NG * ustride;   // the next chunk of the same group
  unsigned long long s0[PB], s1[PB];
#pragma unroll
  for (int k = 0; k < PB; ++k) s0[k] = s1[k] = 0;
  if (C == 1 && a.raw && a.units[kUnitForm + w]) {
    // the item's chunks hold packed 32-bit bins, sum_gq * 65536 + sum_hq each
    const unsigned* b32 = (const unsigned*)(a.slab + (size_t)(u0 + fd.hg) * ustride) + fd.hoff;
    const size_t cs32 = cstride * 2;
    for (int c = 0; c < nch; c += 4) {
      unsigned x[4][PB];
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < PB; ++k) {
          const int b = tid + 128 * k;
          x[j][k] = (c + j < nch && b < nb) ? b32[(size_t)(c + j) * cs32 + b] : 0u;
        }
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < PB; ++k)
          s0[k] += ((unsigned long long)(long long)((int)x[j][k] >> 16) << 32) + (unsigned long long)(x[j][k] & 0xFFFFu);
    }
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      out0[k] = (long long)s0[k];
      out1[k] = 0;
    }
    return;
  }
  int c = 0;
  for (; c + 8 <= nch; c += 8) {
    long long x0[8][PB], x1[8][PB];
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int k = 0; k < PB; ++k) {
        const int b = tid + 128 * k;
        x0[j][k] = x1[j][k] = 0;
        if (b < nb) {
          const long long* q = base + (size_t)(c + j) * cstride + (size_t)b * C;
          x0[j][k] = q[0];
          if (C == 2) x1[j][k] = q[1];
        }
      }
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int k = 0; k < PB; ++k) {
        s0[k] += (unsigned long long)x0[j][k];
        s1[k] += (unsigned long long)x1[j][k];
      }
  }
  for (; c < nch; ++c)
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      const int b = tid + 128 * k;
      if (b < nb) {
        const long long* q = base + (size_t)c * cstride + (size_t)b * C;
        s0[k] += (unsigned long long)q[0];
        if (C == 2) s1[k] += (unsigned long long)q[1];
      }
    }
#pragma unroll
  for (int k = 0; k < PB; ++k) {
    out0[k] = (long long)s0[k];
    out1[k] = (long long)s1[k];
  }
}

// One block (2 waves) per (feature, split of the round): the smaller child's
// histogram from its histogram units (with the most-frequent-bin fix,
// histogram.rs:42-78), the larger child's as parent - smaller
// (histogram.rs:59-85), then wave 0 searches the smaller child and wave 1 the
// larger (learner/mod.rs:954-1101, split.rs:246-271 / 930-950).
__device__ __forceinline__ void grow_split(const SplitArgs& a) {
  StepDev* st = a.st;
  __shared__ double hsm[2][kMaxBins];   // smaller: g, h
  __shared__ double hlg[2][kMaxBins];   // larger: g, h
  __shared__ long long ism[kMaxBins], ilg[kMaxBins];
  __shared__ ScanLds scr[2];
  __shared__ SplitRec res[2];
  const int f = blockIdx.x, w = blockIdx.y;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int F = a.F, B = a.B;
  const bool root = a.root != 0;
  int sm, lg, parent, row_sm, row_lg;
  long long sc;
  LeafSums ls_s, ls_l;
  // the state, the plan item and the feature are loaded together
  const FeatDev fd = a.feats[f];
  if (root) {
    if (st->done) return;
    sm = 0;
    lg = -1;
    parent = -1;
    row_sm = 0;
    row_lg = -1;
    sc = a.leaves[0].count;
    ls_s = ls_l = leaf_sums(a, 0, true);
  } else {
    const PlanDev* P = a.plan;
    const int sdone = st->done, nwork = P->nwork, complete = P->complete;
    const PlanItem& it = P->item[w];   // w < kMaxBatch: inside the plan
    // the left child's rows as the partition counted them (its slot
    // counters are cleared by the next round's plan)
    const long long lc = (long long)st->lcur[w], rc = it.count - lc;
    const int depth = it.depth, ls = it.lslot;
    if (sdone || w >= nwork) return;
    if (complete || children_skip(a.p, depth + 1, lc, rc)) {
      // no search: both children keep (-inf, none)
      if (lane == 0 && wave < 2) {
        SplitRec r;
        reset_rec(r);
        a.feat_split[(2 * w + wave) * F + f] = r;
      }
      return;
    }
    const bool left_smaller = lc < rc;   // learner/mod.rs:1200-1208
    sm = left_smaller ? ls : ls + 1;
    lg = left_smaller ? ls + 1 : ls;
    row_sm = 2 * w + (left_smaller ? 0 : 1);
    row_lg = 2 * w + (left_smaller ? 1 : 0);
    parent = it.parent;
    sc = left_smaller ? lc : rc;
    // the children's sums from the applied split (the leaf records hold the
    // same values)
    const LeafSums l = child_sums(a, it.rec, lc, false), r = child_sums(a, it.rec, rc, true);
    ls_s = left_smaller ? l : r;
    ls_l = left_smaller ? r : l;
  }
  const bool parent_spl = root ? true : a.splittable[(size_t)parent * F + f] != 0;
  constexpr int PB = kMaxBins / 128;   // bins per thread
  if (!a.quant) {
    const int ssg = fix_shift(sc, a.it->max_abs_g_bits);
    const int ssh = fix_shift(sc, a.it->max_abs_h_bits);
    const double* pp = lg >= 0 ? a.pool + ((size_t)parent * F + f) * B * 2 : nullptr;
    long long gi[PB], hi[PB];
    double pg[PB], ph[PB];
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      const int b = tid + 128 * k;
      pg[k] = ph[k] = 0.0;
      if (b < B && pp) {
        pg[k] = pp[2 * b];
        ph[k] = pp[2 * b + 1];
      }
    }
    slab_sum<2>(a, w, fd, tid, gi, hi);
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      const int b = tid + 128 * k;
      if (b < B) {
        hsm[0][b] = ldexp((double)gi[k], -ssg);
        hsm[1][b] = ldexp((double)hi[k], -ssh);
      }
    }
    __syncthreads();
    // histogram.rs:42-56: the most frequent bin from the leaf totals, in
    // the reference's order (two sequential chains)
    if (fd.mfb > 0 && tid < 2) {
      double v = tid == 0 ? ls_s.sg : ls_s.sh - 2.0 * KEPS;
      const double* __restrict__ arr = hsm[tid];
      int i = 0;
      for (; i + 8 <= fd.num_bin; i += 8) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = arr[i + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) v = (i + u == fd.mfb) ? v : v - x[u];
      }
      for (; i < fd.num_bin; ++i) v = (i == fd.mfb) ? v : v - arr[i];
      hsm[tid][fd.mfb] = v;
    }
    __syncthreads();
    double* ps = a.pool + ((size_t)sm * F + f) * B * 2;
    double* pl = lg >= 0 ? a.pool + ((size_t)lg * F + f) * B * 2 : nullptr;
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      const int b = tid + 128 * k;
      if (b < B) {
        const double sgv = hsm[0][b], shv = hsm[1][b];
        ps[2 * b] = sgv;
        ps[2 * b + 1] = shv;
        if (pl) {
          const double lgv = pg[k] - sgv, lhv = ph[k] - shv;
          hlg[0][b] = lgv;
          hlg[1][b] = lhv;
          pl[2 * b] = lgv;
          pl[2 * b + 1] = lhv;
        }
      }
    }
  } else {
    const long long* pp = lg >= 0 ? a.ipool + ((size_t)parent * F + f) * B : nullptr;
    long long v[PB], pv[PB], unused[PB];
    unsigned long long part = 0;
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      const int b = tid + 128 * k;
      pv[k] = 0;
      if (b < B && pp) pv[k] = pp[b];
    }
    slab_sum<1>(a, w, fd, tid, v, unused);
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      const int b = tid + 128 * k;
      if (b < B) {
        ism[b] = v[k];
        if (b != fd.mfb && b < fd.num_bin) part += (unsigned long long)v[k];
      }
    }
    // histogram.rs:67-78: wrapping sums are exact in any order, so the
    // most frequent bin is the leaf total minus a block reduction
    for (int o = 32; o > 0; o >>= 1) part += (unsigned long long)__shfl_xor((long long)part, o, 64);
    __shared__ unsigned long long wpart[2];
    if (lane == 0) wpart[wave] = part;
    __syncthreads();
    if (fd.mfb > 0 && tid == 0) ism[fd.mfb] = (long long)((unsigned long long)ls_s.isum - wpart[0] - wpart[1]);
    __syncthreads();
    long long* ps = a.ipool + ((size_t)sm * F + f) * B;
    long long* pl = lg >= 0 ? a.ipool + ((size_t)lg * F + f) * B : nullptr;
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      const int b = tid + 128 * k;
      if (b < B) {
        const long long sv = ism[b];
        ps[b] = sv;
        if (pl) {
          const long long lv = (long long)((unsigned long long)pv[k] - (unsigned long long)sv);
          ilg[b] = lv;
          pl[b] = lv;
        }
      }
    }
  }
  __syncthreads();
  // wave 0 searches the smaller child, wave 1 the larger
  const int slot = wave == 0 ? sm : lg;
  if (slot >= 0) {
    const LeafSums ls = wave == 0 ? ls_s : ls_l;
    SplitRec& r = res[wave];
    if (lane == 0) reset_rec(r);
    bool spl = false;
    wave_sync();
    if (parent_spl) {
      if (lane == 0) r.gain = -INFINITY;
      if (a.quant) {
        spl = fd.bin_type == 1 ? search_categorical_int(wave == 0 ? ism : ilg, fd, a.p, ls, scr[wave], r)
                               : search_int(wave == 0 ? ism : ilg, fd, a.p, ls, scr[wave], r);
      } else if (fd.bin_type == 1) {
        spl = search_categorical(wave == 0 ? hsm[0] : hlg[0], wave == 0 ? hsm[1] : hlg[1], fd, a.p, ls, scr[wave], r);
      } else {
        spl = search_numerical(wave == 0 ? hsm[0] : hlg[0], wave == 0 ? hsm[1] : hlg[1], fd, a.p, ls, scr[wave], r);
      }
    }
    if (lane == 0) {
      if (parent_spl) r.feature = f;
      a.feat_split[(size_t)(wave == 0 ? row_sm : row_lg) * F + f] = r;
      a.splittable[(size_t)slot * F + f] = spl ? 1 : 0;
    }
  }
  if (root && f == 0 && tid == 0) {
    // learner/mod.rs:647: the root's output before any split
    const LeafSums r = leaf_sums(a, 0, true);
    st->root_output = round0(leaf_output_t(r.sg, r.sh - 2.0 * KEPS, a.p, true, true, false, r.nd, 0.0));
  }
}

__global__ void __launch_bounds__(128) k_grow_split(SplitArgs a) { grow_split(a); }
// The trees of a batch: grid z picks the tree.
__global__ void __launch_bounds__(128) k_grow_split_tab(const SplitArgs* __restrict__ tab) {
  const SplitArgs a = tab[blockIdx.z];
  grow_split(a);
}

// ---------------------------------------------------------------- partition
struct PartArgs {
  StepDev* st;
  PlanDev* plan;
  PlanRep* reps;            // [L] replayed splits of the round
  PosRec* pos;              // [2][L] position arrays
  LeafDev* leaves;
  const FeatDev* feats;
  SplitRec* best;           // [slots] best split of each searched slot
  const SplitRec* feat_split;
  uint8_t* grp[3];          // ordered bins per buffer, group-major [NG][N][FB]
  int FB, NG;               // bytes (features) per group, groups
  long long N;
  void* gh[3];
  int gh_bytes;             // 8 (float2) or 2 (char2)
  int F, L;
};

// A candidate of an ArgMax: gain, feature (-1 none) and position (feature
// slot or leaf). pick() keeps the better one in the order the reference's
// scans imply - higher gain, then smaller feature (learner/mod.rs:1064-1101),
// then the earlier position (first max, learner/mod.rs:698-706). The order
// is total, so any reduction tree gives the same answer.
struct Cand {
  double g;
  int f, l;
};

__device__ __forceinline__ Cand pick(Cand a, Cand b) {
  const bool take = b.l != INT32_MAX &&
                    (a.l == INT32_MAX || better(b.g, b.f, a.g, a.f) || (!better(a.g, a.f, b.g, b.f) && b.l < a.l));
  // field-wise selects keep the candidates in registers
  Cand r;
  r.g = take ? b.g : a.g;
  r.f = take ? b.f : a.f;
  r.l = take ? b.l : a.l;
  return r;
}

__device__ __forceinline__ Cand wave_pick(Cand c) {
  for (int o = 32; o > 0; o >>= 1) {
    Cand d;
    d.g = __shfl_xor(c.g, o, 64);
    d.f = __shfl_xor(c.f, o, 64);
    d.l = __shfl_xor(c.l, o, 64);
    c = pick(c, d);
  }
  return c;
}

__device__ __forceinline__ Cand cand_none() {
  Cand c;
  c.g = -INFINITY;
  c.f = -1;
  c.l = INT32_MAX;
  return c;
}

// The plan of a round, one block:
//  1. each leaf searched by the last split pass gets its best split
//     (learner/mod.rs:1064-1101);
//  2. replay: the sequential ArgMax over the tree's leaves (:698-706), with
//     the reference's stop rules (gain <= 0, num_leaves reached), applying
//     each pick that was already split, until it picks a leaf that was not;
//  3. the round splits that leaf and the next best unsplit leaves, at most
//     kMaxBatch and no more than the tree can still take.
// It publishes the plan, the next position array and the searched leaves'
// best splits.
template <int VW>
struct RowVec;   // the vector type that moves VW bytes of a row
template <>
struct RowVec<4> { using T = unsigned; };
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
template <>
struct RowVec<8> { using T = u32x2; };
template <>
struct RowVec<16> { using T = u32x4; };

// Copies n records of RW words, record w from src(w) to dst(w) (a null dst
// skips the record), on threads t of nt; each thread issues up to 8 word
// loads before its stores.
template <int RW, class Src, class Dst>
__device__ __forceinline__ void copy_recs(int n, Src src, Dst dst, int t = threadIdx.x, int nt = kPartThreads) {
  const int total = n * RW;
  for (int e0 = t; e0 < total; e0 += nt * 8) {
    unsigned v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = min(e0 + u * nt, total - 1);
      v[u] = src(e / RW)[e % RW];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * nt;
      if (e < total) {
        unsigned* d = dst(e / RW);
        if (d) d[e % RW] = v[u];
      }
    }
  }
}

template <int RP>   // positions per lane: num_leaves <= 64 RP
__device__ __forceinline__ void grow_plan(const PartArgs& a) {
  StepDev* st = a.st;
  extern __shared__ PosRec P[];   // this block's copy of the position array, then its keys
  KeyRec* Q = (KeyRec*)(P + a.L);
  constexpr int NW = kPartThreads / 64;
  constexpr int RW = (int)(sizeof(SplitRec) / 4);
  __shared__ Cand rowbest[2 * kMaxBatch];
  __shared__ SplitRec irec[kMaxBatch];
  __shared__ FeatDev ifd[kMaxBatch];
  __shared__ long long ibegin[kMaxBatch], icount[kMaxBatch];
  __shared__ int ipos[kMaxBatch], iparent[kMaxBatch], ibuf[kMaxBatch], idepth[kMaxBatch], itile0[kMaxBatch + 1];
  __shared__ int s_k, s_fin, s_nseq, s_nspl, s_nrep, s_npend, s_complete, s_wk;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (st->done) return;
  const int nseq0 = st->nseq, npend0 = st->npending, nslots0 = st->nslots, nspl0 = st->num_splits;
  const int par = st->parity, sbase = st->search_base, nsearched = st->nsearched;
  const int F = a.F, L = a.L;
  // ---- 1. positions and searched rows (independent loads, all issued
  // before any is used: clamped slots reload a valid entry)
  for (int i0 = tid; i0 < nseq0; i0 += kPartThreads * 4) {
    PosRec pr[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) pr[u] = a.pos[(size_t)par * L + min(i0 + u * kPartThreads, nseq0 - 1)];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i0 + u * kPartThreads < nseq0) P[i0 + u * kPartThreads] = pr[u];
  }
  {
    const int r = tid / kRowThreads, q = tid % kRowThreads;
    Cand c = cand_none();
    if (r < nsearched) {
      const SplitRec* row = a.feat_split + (size_t)r * F;
      for (int f0 = q; f0 < F; f0 += kRowThreads * 8) {
        double g8[8];
        int f8[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int f = min(f0 + u * kRowThreads, F - 1);
          g8[u] = row[f].gain;
          f8[u] = row[f].feature;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (f0 + u * kRowThreads >= F) continue;
          Cand d;
          d.g = g8[u];
          d.f = f8[u];
          d.l = f0 + u * kRowThreads;
          c = pick(c, d);
        }
      }
    }
    for (int o = 1; o < kRowThreads; o <<= 1) {
      Cand d;
      d.g = __shfl_xor(c.g, o, 64);
      d.f = __shfl_xor(c.f, o, 64);
      d.l = __shfl_xor(c.l, o, 64);
      c = pick(c, d);
    }
    if (q == 0 && r < 2 * kMaxBatch) {
      // a leaf without any candidate carries (-inf, none)
      if (!better(c.g, c.f, -INFINITY, -1)) {
        c.g = -INFINITY;
        c.f = -1;
        c.l = -1;
      }
      rowbest[r] = c;
    }
  }
  __syncthreads();
  for (int i = tid; i < nseq0; i += kPartThreads) {
    PosRec& q = P[i];
    const int rs = q.slot - sbase;
    if (rs >= 0 && rs < nsearched) {
      q.gain = rowbest[rs].g;
      q.feat = rowbest[rs].f;
    }
    Q[i].k1 = gain_key(q.gain);
    Q[i].k2 = tie_key(q.feat, i);
    Q[i].spec = q.spec;
    if (q.spec >= 0)
      for (int c = 0; c < 2; ++c) {
        const int rc = q.spec + c - sbase;
        if (rc >= 0 && rc < nsearched) {
          q.cgain[c] = rowbest[rc].g;
          q.cfeat[c] = rowbest[rc].f;
        }
      }
  }
  __syncthreads();
  // ---- 2. replay, wave 0. Each lane keeps its positions i = lane + 64 r
  // in registers (keys, slot, and for a leaf split ahead of order its child
  // slots and their keys); a picked position's data reaches every lane by
  // readlane, and a replayed split changes two positions, so only their two
  // lanes rescan before the next wave ArgMax. LDS copies are written for the
  // selection and the next position array.
  if (wave == 0) {
    int nseq = nseq0, nspl = nspl0, nrep = 0, npend = npend0;
    bool fin = false;
    unsigned long long r1[RP], c1a[RP], c1b[RP];
    unsigned r2[RP];
    int rsl[RP], rsp[RP], cfa[RP], cfb[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      const int i = lane + 64 * r;
      r1[r] = 0;
      r2[r] = 0xFFFFFFFFu;
      rsl[r] = rsp[r] = -1;
      c1a[r] = c1b[r] = 0;
      cfa[r] = cfb[r] = -1;
      if (i < nseq) {
        const PosRec& q = P[i];
        r1[r] = Q[i].k1;
        r2[r] = Q[i].k2;
        rsl[r] = q.slot;
        rsp[r] = q.spec;
        c1a[r] = gain_key(q.cgain[0]);
        c1b[r] = gain_key(q.cgain[1]);
        cfa[r] = q.cfeat[0];
        cfb[r] = q.cfeat[1];
      }
    }
    auto lane_best = [&](unsigned long long& b1, unsigned& b2) {
      b1 = r1[0];
      b2 = r2[0];
#pragma unroll
      for (int r = 1; r < RP; ++r)
        if (key_better(r1[r], r2[r], b1, b2)) {
          b1 = r1[r];
          b2 = r2[r];
        }
    };
    auto bcast_u32 = [&](unsigned v, int src) { return (unsigned)__builtin_amdgcn_readlane((int)v, src); };
    unsigned long long m1;
    unsigned m2;
    lane_best(m1, m2);
    for (;;) {
      if (nseq >= L) {
        fin = true;
        break;
      }
      unsigned long long c1 = m1;
      unsigned c2 = m2;
      wave_best_key(c1, c2);
      if (!(c1 > gain_key(0.0))) {   // the best gain is not positive (or there is no leaf)
        fin = true;
        break;
      }
      const int m = (int)(c2 & 0xFFFF);
      const int om = m & 63, rm = m >> 6;
      // the picked position's slot, split state and child keys, from its lane
      int xsl = 0, xsp = 0, xfa = 0, xfb = 0;
      unsigned long long xa = 0, xb = 0;
#pragma unroll
      for (int r = 0; r < RP; ++r)
        if (r == rm) {
          xsl = rsl[r];
          xsp = rsp[r];
          xfa = cfa[r];
          xfb = cfb[r];
          xa = c1a[r];
          xb = c1b[r];
        }
      const int spec = (int)bcast_u32((unsigned)xsp, om);
      if (spec < 0) break;   // not split yet: this round splits it
      const int pslot = (int)bcast_u32((unsigned)xsl, om);
      const int fa = (int)bcast_u32((unsigned)xfa, om), fb = (int)bcast_u32((unsigned)xfb, om);
      const unsigned long long ka = ((unsigned long long)bcast_u32((unsigned)(xa >> 32), om) << 32) |
                                    bcast_u32((unsigned)xa, om);
      const unsigned long long kb = ((unsigned long long)bcast_u32((unsigned)(xb >> 32), om) << 32) |
                                    bcast_u32((unsigned)xb, om);
      if (lane == 0 && blockIdx.x == 0) {
        PlanRep rp;
        rp.pos = m;
        rp.parent = pslot;
        rp.lslot = spec;
        rp.pad = 0;
        a.reps[nrep] = rp;
      }
      // learner/mod.rs:1105-1225: the left child keeps the leaf's number,
      // the right child takes the next one
      const int n = nseq;
      const int on = n & 63, rn = n >> 6;
      const unsigned k2a = tie_key(fa, m), k2b = tie_key(fb, n);
#pragma unroll
      for (int r = 0; r < RP; ++r) {
        if (lane == om && r == rm) {
          r1[r] = ka;
          r2[r] = k2a;
          rsl[r] = spec;
          rsp[r] = -1;
        }
        if (lane == on && r == rn) {
          r1[r] = kb;
          r2[r] = k2b;
          rsl[r] = spec + 1;
          rsp[r] = -1;
        }
      }
      ++nseq;
      ++nspl;
      ++nrep;
      --npend;
      if (lane == om || lane == on) lane_best(m1, m2);
    }
    // ---- 3a. on a large tree the selection runs here, on the registers:
    // the unsplit leaves are taken in ArgMax order, one at a time; each
    // one's count of leaves split ahead of order that rank above it is a
    // wave ballot. It stops where the rule of step 3 rejects a leaf or
    // kWaveSelMax are chosen (a prefix of step 3's choice, which starts with
    // the replay's stop). Late in a tree few leaves qualify, and this costs a
    // few wave ArgMax steps instead of a rank for every leaf.
    int wk = -1;
    if (!fin && nseq > kWaveSelMin) {
      const unsigned long long zero = gain_key(0.0);
      const int room = L - nseq;
      unsigned long long e1[RP];
#pragma unroll
      for (int r = 0; r < RP; ++r) e1[r] = rsp[r] < 0 ? r1[r] : 0ull;
      auto lane_best_e = [&](unsigned long long& b1, unsigned& b2) {
        b1 = e1[0];
        b2 = r2[0];
#pragma unroll
        for (int r = 1; r < RP; ++r)
          if (key_better(e1[r], r2[r], b1, b2)) {
            b1 = e1[r];
            b2 = r2[r];
          }
      };
      unsigned long long x1;
      unsigned x2;
      lane_best_e(x1, x2);
      int rank = 0;
      while (rank < kWaveSelMax) {
        unsigned long long c1 = x1;
        unsigned c2 = x2;
        wave_best_key(c1, c2);
        if (!(c1 > zero)) break;
        int above = 0;
#pragma unroll
        for (int r = 0; r < RP; ++r)
          above += __popcll(__ballot(rsp[r] >= 0 && key_better(r1[r], r2[r], c1, c2)));
        if (rank + above + 1 > room) break;
        const int m = (int)(c2 & 0xFFFF);
        const int om = m & 63, rm = m >> 6;
        if (lane == 0) ipos[rank] = m;
        ++rank;
        if (lane == om) {
#pragma unroll
          for (int r = 0; r < RP; ++r)
            if (r == rm) e1[r] = 0;
          lane_best_e(x1, x2);
        }
      }
      wk = rank;
    }
    // LDS copies of this wave's final positions
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      const int i = lane + 64 * r;
      if (i < nseq) {
        PosRec q;
        q.gain = key_gain(r1[r]);
        q.feat = (int)(r2[r] >> 16) == 0xFFFF ? -1 : (int)(r2[r] >> 16);
        q.slot = rsl[r];
        q.spec = rsp[r];
        q.cgain[0] = rsp[r] >= 0 ? key_gain(c1a[r]) : -INFINITY;
        q.cgain[1] = rsp[r] >= 0 ? key_gain(c1b[r]) : -INFINITY;
        q.cfeat[0] = rsp[r] >= 0 ? cfa[r] : -1;
        q.cfeat[1] = rsp[r] >= 0 ? cfb[r] : -1;
        q.pad = 0;
        P[i] = q;
        Q[i].k1 = r1[r];
        Q[i].k2 = r2[r];
        Q[i].spec = rsp[r];
      }
    }
    if (lane == 0) {
      s_fin = fin ? 1 : 0;
      s_nseq = nseq;
      s_nspl = nspl;
      s_nrep = nrep;
      s_npend = npend;
      s_wk = wk;
      s_k = wk < 0 ? 0 : wk;
    }
  } else {
    // meanwhile the other waves store each searched leaf's best split in its
    // slot (nothing in this launch reads those slots: a round's parents read
    // their records from the searched rows)
    copy_recs<RW>(
        nsearched,
        [&](int r) { return (const unsigned*)(a.feat_split + (size_t)r * F + max(rowbest[r].l, 0)); },
        [&](int r) { return rowbest[r].l >= 0 ? (unsigned*)(a.best + sbase + r) : (unsigned*)nullptr; },
        tid - 64, kPartThreads - 64);
    for (int r = tid - 64; r < nsearched; r += kPartThreads - 64)
      if (rowbest[r].l < 0) {
        SplitRec z;
        reset_rec(z);
        a.best[sbase + r] = z;
      }
  }
  __syncthreads();
  // ---- 3. selection, all threads: the round splits the unsplit leaves in
  // ArgMax order, at most kMaxBatch, and only leaves the tree can still
  // reach (every leaf split ahead of order that ranks above a leaf is taken
  // before it, and the tree has L - nseq leaves of room). A leaf's rank among
  // the unsplit leaves is its item index; the chosen set is a prefix of that
  // order.
  {
    const int nseq = s_nseq;
    if (s_wk < 0 && !s_fin) {
      const unsigned long long zero = gain_key(0.0);
      for (int i = tid; i < nseq; i += kPartThreads) {
        const KeyRec qi = Q[i];
        if (qi.spec >= 0 || !(qi.k1 > zero)) continue;
        int rank = 0, above = 0;
        int j = 0;
        for (; j + 8 <= nseq; j += 8) {
          KeyRec qj[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) qj[u] = Q[j + u];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const bool bt = key_better(qj[u].k1, qj[u].k2, qi.k1, qi.k2);
            above += (bt && qj[u].spec >= 0) ? 1 : 0;
            rank += (bt && qj[u].spec < 0) ? 1 : 0;
          }
        }
        for (; j < nseq; ++j) {
          const KeyRec qj = Q[j];
          const bool bt = key_better(qj.k1, qj.k2, qi.k1, qi.k2);
          above += (bt && qj.spec >= 0) ? 1 : 0;
          rank += (bt && qj.spec < 0) ? 1 : 0;
        }
        if (rank < kMaxBatch && rank + above + 1 <= L - nseq) {
          ipos[rank] = i;
          atomicAdd(&s_k, 1);
        }
      }
    }
    __syncthreads();
    const int k = s_k;
    if (tid < k) {
      PosRec& q = P[ipos[tid]];
      q.spec = nslots0 + 2 * tid;   // split this round: its children's slots
      q.cgain[0] = q.cgain[1] = -INFINITY;
      q.cfeat[0] = q.cfeat[1] = -1;
    }
    if (tid == 0) {
      s_npend += k;
      // a child of this round can be split only after its parent is
      // replayed: two more leaves
      s_complete = nseq + 2 > L ? 1 : 0;
    }
  }
  __syncthreads();
  const int k = s_k;
  // ---- the round's splits: parent segment, chosen record, feature
  if (tid < k) {
    const int pslot = P[ipos[tid]].slot;
    iparent[tid] = pslot;
    const LeafDev& lf = a.leaves[pslot];
    ibegin[tid] = lf.begin;
    icount[tid] = lf.count;
    ibuf[tid] = lf.buf;
    idepth[tid] = lf.depth;
    const int rs = pslot - sbase;
    const int feat = (rs >= 0 && rs < nsearched) ? rowbest[rs].f : P[ipos[tid]].feat;
    ifd[tid] = a.feats[feat];
  }
  copy_recs<RW>(
      k,
      [&](int w) {
        const int pslot = P[ipos[w]].slot;
        const int rs = pslot - sbase;
        const SplitRec* src =
            (rs >= 0 && rs < nsearched) ? a.feat_split + (size_t)rs * F + rowbest[rs].l : a.best + pslot;
        return (const unsigned*)src;
      },
      [&](int w) { return (unsigned*)&irec[w]; });
  __syncthreads();
  if (tid == 0) {
    int t = 0;
    for (int w = 0; w < k; ++w) {
      itile0[w] = t;
      t += (int)((icount[w] + kPartTile - 1) / kPartTile);
    }
    itile0[k] = t;
  }
  __syncthreads();
  const int total_tiles = itile0[k];
  {
    // publish: the plan, the next position array, the searched leaves' best splits
    PlanDev* pl = a.plan;
    for (int e = tid; e < k * RW; e += kPartThreads) {
      const int w = e / RW, wd = e % RW;
      ((unsigned*)&pl->item[w].rec)[wd] = ((const unsigned*)&irec[w])[wd];
    }
    if (tid < k) {
      PlanItem& it = pl->item[tid];
      it.begin = ibegin[tid];
      it.count = icount[tid];
      it.parent = iparent[tid];
      it.lslot = nslots0 + 2 * tid;
      it.buf = ibuf[tid];
      it.depth = idepth[tid];
      it.tile0 = itile0[tid];
      it.pad = 0;
      it.fd = ifd[tid];
    }
    // the partition pass reserves the children's slots from zero (the
    // counters of the last round were read by its histogram and search passes)
    if (tid < kMaxBatch) {
      st->lcur[tid] = 0;
      st->rcur[tid] = 0;
    }
    if (tid == 0) {
      pl->done = s_fin;
      pl->nwork = k;
      pl->total_tiles = total_tiles;
      pl->nseq = s_nseq;
      pl->npending = s_npend;
      pl->nslots = nslots0 + 2 * k;
      pl->num_splits = s_nspl;
      pl->nrep = s_nrep;
      pl->complete = s_complete;
      pl->pad = 0;
    }
    PosRec* nxt = a.pos + (size_t)(par ^ 1) * L;
    for (int i = tid; i < s_nseq; i += kPartThreads) nxt[i] = P[i];
  }
}

template <int RP>
__global__ void __launch_bounds__(kPartThreads) k_grow_plan(PartArgs a) { grow_plan<RP>(a); }
// The trees of a batch: block y plans tree y.
template <int RP>
__global__ void __launch_bounds__(kPartThreads) k_grow_plan_tab(const PartArgs* __restrict__ tab) {
  const PartArgs a = tab[blockIdx.y];
  grow_plan<RP>(a);
}

// One partition pass per round: every block moves its tiles' rows -
// gradients and group-major bins - into the children's segments of the other
// buffer: left rows from the front, right rows from the back, each tile
// reserving its slots with one atomic per side (DataPartition::split,
// partition.rs:107-174). The row order inside a child is unspecified: every
// consumer of a segment (exact fixed-point or integer histograms, renewal
// sums, score update) is order-independent.
template <int GB>
struct GradVec;   // one row's gradients: char2 (quantized) or float2 (plain)
template <>
struct GradVec<2> { using T = unsigned short; };
template <>
struct GradVec<8> { using T = uint2; };

// Every load of a tile is unconditional (slots past the segment load its
// last row) and issued before any of the loaded values is used, so a
// thread's loads are in flight together; stores follow, masked.
template <int VW, int GB>
__device__ __forceinline__ void grow_part(const PartArgs& a) {
  using VT = typename RowVec<VW>::T;
  using GT = typename GradVec<GB>::T;
  constexpr int VWW = VW / 4;   // words per vector
  StepDev* st = a.st;
  constexpr int NW = kPartThreads / 64;
  constexpr int IW = (int)(sizeof(PlanItem) / 4);
  __shared__ PlanItem items[kMaxBatch];
  __shared__ int s_hdr[3];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // the tree state and the plan's header together (the state's done flag
  // changes only in the histogram pass); blocks past the round's tiles stop
  // there, the others load the round's items
  const PlanDev* pd = a.plan;
  if (tid == 0) {
    const int sd = st->done, pdn = pd->done, nw = pd->nwork, tt = pd->total_tiles;
    s_hdr[0] = sd | pdn;
    s_hdr[1] = nw;
    s_hdr[2] = tt;
  }
  __syncthreads();
  if (s_hdr[0]) return;
  const int k = s_hdr[1], total_tiles = s_hdr[2];
  if (k == 0 || (int)blockIdx.x >= total_tiles) return;
  for (int e = tid; e < k * IW; e += kPartThreads) ((unsigned*)items)[e] = ((const unsigned*)pd->item)[e];
  __syncthreads();
  // ---- tiles
  const int FB = a.FB;
  const size_t gstride = (size_t)a.N * FB;   // bytes of one feature group's array
  const int words = FB / 4, tw = a.NG * words;   // 4-byte words of one row over all groups
  constexpr int R = kPartRowsPerThread;
  constexpr int CW = 8;                          // row words moved per round of loads
  __shared__ int cnt[2][R][NW];
  __shared__ long long slot[3];
  const unsigned long long below = (1ull << lane) - 1ull;
  int w = 0;
  for (int t = blockIdx.x; t < total_tiles; t += gridDim.x) {
    while (w + 1 < k && t >= items[w + 1].tile0) ++w;
    const SplitRec& s = items[w].rec;
    const FeatDev& fd = items[w].fd;
    const long long begin = items[w].begin, count = items[w].count;
    const int src = items[w].buf, dst = child_buf(src);   // buffer 0 is never a destination
    const GT* igh = (const GT*)sel3(a.gh[0], a.gh[1], a.gh[2], src) + begin;
    GT* ogh = (GT*)sel3(a.gh[0], a.gh[1], a.gh[2], dst) + begin;
    const uint8_t* gsrc = sel3(a.grp[0], a.grp[1], a.grp[2], src) + (size_t)begin * FB;
    uint8_t* gdst = sel3(a.grp[0], a.grp[1], a.grp[2], dst) + (size_t)begin * FB;
    // the split feature's byte (its bin, or a pair of 4-bit fields)
    const uint8_t* scol = gsrc + (size_t)(fd.pos / FB) * gstride + (fd.pos % FB);
    const int bshift = fd.nib > 0 ? 4 : 0;
    const unsigned bmask = fd.nib >= 0 ? 0xFu : 0xFFu;
    const long long t0 = (long long)(t - items[w].tile0) * kPartTile;
    // one round of independent loads: gradients, the split feature's bin and
    // the first words of every row (lanes take consecutive rows); a slot past
    // the segment reloads the segment's last row, a word past the row its
    // last word
    bool valid[R], left[R];
    long long pc[R];
    GT gv[R];
    uint8_t bin[R];
    VT v[R][CW / VWW];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const long long p = t0 + u * kPartThreads + tid;
      valid[u] = p < count;
      pc[u] = valid[u] ? p : count - 1;
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
      gv[u] = igh[pc[u]];
      bin[u] = scol[(size_t)pc[u] * FB];
#pragma unroll
      for (int i = 0; i < CW; i += VWW) {
        const int ii = min(i, tw - VWW);
        v[u][i / VWW] = *(const VT*)(gsrc + (size_t)(ii / words) * gstride + (size_t)pc[u] * FB + 4 * (ii % words));
      }
    }
#pragma unroll
    for (int u = 0; u < R; ++u) left[u] = valid[u] && goes_left((int)((bin[u] >> bshift) & bmask), fd, s);
    // ranks inside each 64-row slice, slice counts to LDS
    int rk[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const unsigned long long lm = __ballot(left[u]);
      const unsigned long long rm = __ballot(valid[u] && !left[u]);
      rk[u] = (int)__popcll((left[u] ? lm : rm) & below);
      if (lane == 0) {
        cnt[0][u][wave] = (int)__popcll(lm);
        cnt[1][u][wave] = (int)__popcll(rm);
      }
    }
    __syncthreads();
    if (tid == 0) {
      int tl = 0, tr = 0;
      for (int u = 0; u < R; ++u)
        for (int q = 0; q < NW; ++q) {
          tl += cnt[0][u][q];
          tr += cnt[1][u][q];
        }
      slot[0] = (long long)atomicAdd(&st->lcur[w], (unsigned long long)tl);
      slot[1] = (long long)atomicAdd(&st->rcur[w], (unsigned long long)tr);
      slot[2] = tr;
    }
    __syncthreads();
    // destinations: left rows fill from the front, right rows from the back;
    // inside a tile the order is slice by slice, so a slice's stores are
    // contiguous
    long long d[R];
    int pl = 0, pr = 0;
#pragma unroll
    for (int u = 0; u < R; ++u) {
      int bl = pl, br = pr;
      for (int q = 0; q < NW; ++q) {
        if (q < wave) {
          bl += cnt[0][u][q];
          br += cnt[1][u][q];
        }
        pl += cnt[0][u][q];
        pr += cnt[1][u][q];
      }
      d[u] = left[u] ? slot[0] + bl + rk[u] : count - slot[1] - slot[2] + br + rk[u];
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
      if (!valid[u]) continue;
      ogh[d[u]] = gv[u];
#pragma unroll
      for (int i = 0; i < CW; i += VWW)
        if (i < tw)
          *(VT*)(gdst + (size_t)(i / words) * gstride + (size_t)d[u] * FB + 4 * (i % words)) = v[u][i / VWW];
    }
    // rows wider than CW words: the rest, CW words per round of loads
    for (int k0 = CW; k0 < tw; k0 += CW) {
#pragma unroll
      for (int u = 0; u < R; ++u) {
#pragma unroll
        for (int i = 0; i < CW; i += VWW) {
          const int kk = min(k0 + i, tw - VWW);
          v[u][i / VWW] = *(const VT*)(gsrc + (size_t)(kk / words) * gstride + (size_t)pc[u] * FB + 4 * (kk % words));
        }
      }
#pragma unroll
      for (int u = 0; u < R; ++u) {
        if (!valid[u]) continue;
#pragma unroll
        for (int i = 0; i < CW; i += VWW) {
          const int kk = k0 + i;
          if (kk < tw)
            *(VT*)(gdst + (size_t)(kk / words) * gstride + (size_t)d[u] * FB + 4 * (kk % words)) = v[u][i / VWW];
        }
      }
    }
    __syncthreads();
  }
}

template <int VW, int GB>
__global__ void __launch_bounds__(kPartThreads) k_grow_part(PartArgs a) { grow_part<VW, GB>(a); }
// The trees of a batch: grid y picks the tree, grid x is its blocks.
template <int VW, int GB>
__global__ void __launch_bounds__(kPartThreads) k_grow_part_tab(const PartArgs* __restrict__ tab) {
  const PartArgs a = tab[blockIdx.y];
  grow_part<VW, GB>(a);
}

// Rows of the row-major packed bins (stride S) into the group-major layout
// of the ordered buffers: out[g][i][0..FB) = bytes g*FB.. of row rows[i] (or
// row i when rows is null), zero past the row's W bytes.
__global__ void k_group_rows(const unsigned* rows, long long m, const uint8_t* __restrict__ rm, int S, int W, int FB,
                             int NG, long long N, uint8_t* __restrict__ out) {
  const int words = FB / 4;
  const long long total = m * NG * words;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    const long long i = e / (NG * words);
    const int k = (int)(e % (NG * words));
    const int g = k / words, w = k % words;
    const long long r = rows ? (long long)rows[i] : i;
    unsigned v = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int p = g * FB + w * 4 + b;
      if (p < W) v |= (unsigned)rm[r * S + p] << (8 * b);
    }
    *(unsigned*)(out + (size_t)g * N * FB + (size_t)i * FB + 4 * w) = v;
  }
}

// k_group_rows for rows whose groups are whole 16-byte chunks inside the row
// (S and FB multiples of 16, NG * FB <= S): one thread per row, its groups
// moved as 16-byte vectors, every load issued before the stores.
template <int CH>   // 16-byte chunks per row (NG * FB / 16)
__global__ void __launch_bounds__(256) k_group_rows_v(const unsigned* rows, long long m, const uint8_t* __restrict__ rm,
                                                      int S, int FB, long long N, uint8_t* __restrict__ out) {
  const int cpg = FB / 16;   // chunks per group
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x) {
    const long long r = rows ? (long long)rows[i] : i;
    uint4 v[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) v[c] = *(const uint4*)(rm + r * S + 16 * c);
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int g = c / cpg, q = c % cpg;
      *(uint4*)(out + (size_t)g * N * FB + (size_t)i * FB + 16 * q) = v[c];
    }
  }
}

// ---------------------------------------------------------------- tree end
// After the last round the tree's leaves are the positions 0..nseq of the
// current position array. The kernels below run after every launch of
// rounds and do nothing while the tree still grows (a continuation launch
// runs them again).
__device__ __forceinline__ const LeafDev& leaf_at(const StepDev* st, const PosRec* pos, const LeafDev* leaves, int L,
                                                   int l) {
  return leaves[pos[(size_t)st->parity * L + l].slot];
}

// One internal node of the finished tree: its split (feature, threshold,
// default direction, categorical bins) and children (>= 0 node, < 0 ~leaf),
// numbered as tree.rs:254-297 numbers them.
struct NodeDev {
  int feature, threshold, default_left;
  int left, right;
  int pad;
  unsigned cat[kCatWords];
};

// The tree's nodes from the replayed split log: split i applied to leaf l
// makes node i with children ~l (left keeps the leaf's number) and ~(i + 1),
// and takes l's place under l's parent (tree.rs:254-297). Also each leaf's
// fixed-point shifts for renewal sums (by leaf row count).
__global__ void __launch_bounds__(256) k_tree_nodes(const StepDev* st, const SplitRec* log, const PosRec* pos,
                                                    const LeafDev* leaves, const IterDev* it, int L, NodeDev* nodes,
                                                    int* shifts /* [L][2] */) {
  __shared__ int lparent[1024];   // node whose child is leaf l (-1: the root)
  __shared__ int lside[1024];
  const int ns = st->num_splits;
  if (st->done != 1) return;
  const int tid = threadIdx.x;
  for (int i = tid; i < ns; i += blockDim.x) {
    const SplitRec& r = log[i];
    NodeDev& n = nodes[i];
    n.feature = r.feature;
    n.threshold = r.threshold;
    n.default_left = r.default_left;
    n.pad = 0;
#pragma unroll
    for (int w = 0; w < kCatWords; ++w) n.cat[w] = r.cat_bits[w];
  }
  for (int l = tid; l < st->nseq; l += blockDim.x) {
    const LeafDev& lf = leaf_at(st, pos, leaves, L, l);
    shifts[2 * l] = fix_shift(lf.count, it->max_abs_g_bits);
    shifts[2 * l + 1] = fix_shift(lf.count, it->max_abs_h_bits);
  }
  __shared__ int sleaf[1024], sl[1024], sr[1024];
  for (int i = tid; i < ns; i += blockDim.x) sleaf[i] = log[i].leaf;
  __syncthreads();
  if (tid == 0) {
    lparent[0] = -1;
    lside[0] = 0;
    for (int i = 0; i < ns; ++i) {
      const int l = sleaf[i], nl = i + 1;
      sl[i] = ~l;
      sr[i] = ~nl;
      if (lparent[l] >= 0) {
        if (lside[l] == 0) sl[lparent[l]] = i;
        else sr[lparent[l]] = i;
      }
      lparent[l] = i;
      lside[l] = 0;
      lparent[nl] = i;
      lside[nl] = 1;
    }
  }
  __syncthreads();
  for (int i = tid; i < ns; i += blockDim.x) {
    nodes[i].left = sl[i];
    nodes[i].right = sr[i];
  }
}

// A node packed for walking: x = the feature's byte in a packed row |
// threshold << 16 | missing bin << 24, y = flags (1 default left, 2 missing
// applies, 4 categorical, 8 a 4-bit field, 16 the byte's high field),
// z = left child, w = right child (>= 0 node, < 0 ~leaf).
__device__ __forceinline__ uint4 pack_node(const NodeDev& n, const FeatDev& fd) {
  uint4 q;
  const bool cat = fd.bin_type == 1;
  int mbin = 0, flags = n.default_left ? 1 : 0;
  if (fd.nib >= 0) flags |= 8 | (fd.nib ? 16 : 0);
  if (cat) {
    flags |= 4;
    mbin = fd.mfb;
  } else if (fd.missing_type == 1) {
    flags |= 2;
    mbin = fd.default_bin;
  } else if (fd.missing_type == 2) {
    flags |= 2;
    mbin = fd.num_bin - 1;
  }
  q.x = (unsigned)fd.pos | ((unsigned)n.threshold & 0xFF) << 16 | ((unsigned)mbin & 0xFF) << 24;
  q.y = (unsigned)flags;
  q.z = (unsigned)n.left;
  q.w = (unsigned)n.right;
  return q;
}

// One step of a walk with the partition's decisions (goes_left): the child
// of packed node q for bin b; cat: the node's categorical bin set.
__device__ __forceinline__ int walk_step(const uint4& q, int b, const unsigned* cat) {
  bool left;
  const int mbin = (int)(q.x >> 24);
  if (q.y & 4u) {
    // goes_left_categorical: the most frequent bin 0 goes right
    left = !(b == mbin && mbin == 0) && ((cat[b >> 5] >> (b & 31)) & 1u);
  } else {
    const bool miss = (q.y & 2u) && b == mbin;
    left = miss ? (q.y & 1u) != 0 : b <= (int)((q.x >> 16) & 0xFF);
  }
  return (int)(left ? q.z : q.w);
}

// Rows in row order (blocks stream contiguous rows), every row walked down
// the tree held in LDS, four rows per thread in lockstep. mode 0: renewal
// sums (quantized form, learner/mod.rs:736-747: each leaf's float gradient
// sums as exact fixed-point, LDS per leaf, one atomic per leaf and block);
// mode 1: score update (tree.rs:635-652: score[row] += value of its leaf);
// mode 2: leaf index of every row.
struct WalkArgs {
  const StepDev* st;
  const NodeDev* nodes;
  const FeatDev* feats;
  const uint8_t* bins;   // row-major [N][S]
  int S, L;
  long long N;
  const float* g;
  const float* h;
  const unsigned* rows;  // MODE 0 over a bag: row of each position (null: every row in order)
  uint16_t* leaf16;      // MODE 0: each row's leaf, when given (the score update then reads it)
  const int* shifts;
  long long* sums;       // [L][2]
  const double* val;     // [L]
  double* score;
  int* out;
  int mode;
};

// Dynamic LDS of k_walk: packed nodes, categorical bin sets, per-leaf sums,
// values and shifts.
__host__ __device__ constexpr size_t walk_lds(int L) {
  return (size_t)L * (16 + 4 * kCatWords) + (size_t)L * (2 * 8 + 8 + 2 * 4);
}

template <int RW, int MODE>
__global__ void __launch_bounds__(256) k_walk(WalkArgs a) {
  extern __shared__ unsigned long long walk_sm[];
  uint4* nq = (uint4*)walk_sm;                                      // [L]
  unsigned long long* lsum = (unsigned long long*)(nq + a.L);      // [2 L]
  double* lval = (double*)(lsum + 2 * a.L);                        // [L]
  unsigned* ncat = (unsigned*)(lval + a.L);                        // [L][kCatWords]
  int* lsh = (int*)(ncat + (size_t)a.L * kCatWords);               // [2 L]
  const StepDev* st = a.st;
  if (MODE != 2 && st->done != 1) return;
  const int ns = st->num_splits, nl = ns + 1;
  const int tid = threadIdx.x;
  for (int i = tid; i < ns; i += blockDim.x) {
    const NodeDev& n = a.nodes[i];
    nq[i] = pack_node(n, a.feats[n.feature]);
#pragma unroll
    for (int w = 0; w < kCatWords; ++w) ncat[(size_t)i * kCatWords + w] = n.cat[w];
  }
  if (MODE == 0)
    for (int l = tid; l < nl; l += blockDim.x) {
      lsum[2 * l] = lsum[2 * l + 1] = 0ull;
      lsh[2 * l] = a.shifts[2 * l];
      lsh[2 * l + 1] = a.shifts[2 * l + 1];
    }
  if (MODE == 1)
    for (int l = tid; l < nl; l += blockDim.x) lval[l] = a.val[l];
  __syncthreads();
  constexpr int U = 4;   // rows walked together per thread
  const long long step = (long long)gridDim.x * blockDim.x * U;
  for (long long base = (long long)blockIdx.x * blockDim.x * U + tid; base < a.N; base += step) {
    // every load first (a slot past the last row reloads the last row):
    // the row's bins and what the row's leaf will be combined with
    unsigned words[U][RW > 0 ? RW : 1];
    int node[U];
    long long rc[U];
    double sc[U];
    float gg[U], hh[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long r = base + (long long)u * blockDim.x;
      node[u] = (ns == 0 || r >= a.N) ? -1 : 0;   // ~0: leaf 0
      rc[u] = r < a.N ? r : a.N - 1;
      if (MODE == 0 && a.rows) rc[u] = (long long)a.rows[rc[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (RW > 0) {
        const uint8_t* row = a.bins + (size_t)rc[u] * a.S;
#pragma unroll
        for (int i = 0; i < RW; i += 4) {
          if (i + 4 <= RW) {
            const uint4 v = *(const uint4*)(row + 4 * i);
            words[u][i] = v.x;
            words[u][i + 1] = v.y;
            words[u][i + 2] = v.z;
            words[u][i + 3] = v.w;
          } else {
            for (int j = i; j < RW; ++j) words[u][j] = ((const unsigned*)row)[j];
          }
        }
      }
      if (MODE == 0) {
        gg[u] = a.g[rc[u]];
        hh[u] = a.h[rc[u]];
      }
      if (MODE == 1) sc[u] = a.score[rc[u]];
    }
    for (;;) {
      bool any = false;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (node[u] >= 0) {
          const uint4 q = nq[node[u]];
          const int f = (int)(q.x & 0xFFFF);   // the feature's byte
          int b;
          if (RW > 0) {
            const int wi = f >> 2;
            unsigned w = words[u][0];
#pragma unroll
            for (int i = 1; i < RW; ++i) w = wi == i ? words[u][i] : w;
            b = (int)((w >> (8 * (f & 3))) & 0xFF);
          } else {
            b = a.bins[(size_t)rc[u] * a.S + f];
          }
          if (q.y & 8u) b = (b >> ((q.y & 16u) ? 4 : 0)) & 0xF;
          node[u] = walk_step(q, b, ncat + (size_t)node[u] * kCatWords);
          any = true;
        }
      }
      if (!any) break;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long r = base + (long long)u * blockDim.x;
      if (r >= a.N) continue;
      const int l = ~node[u];
      if (MODE == 0) {
        atomicAdd(&lsum[2 * l], (unsigned long long)to_fixed(gg[u], lsh[2 * l]));
        atomicAdd(&lsum[2 * l + 1], (unsigned long long)to_fixed(hh[u], lsh[2 * l + 1]));
        if (a.leaf16) a.leaf16[rc[u]] = (uint16_t)l;
      } else if (MODE == 1) {
        a.score[r] = sc[u] + lval[l];
      } else {
        a.out[r] = l;
      }
    }
  }
  if (MODE == 0) {
    __syncthreads();
    for (int l = tid; l < nl; l += blockDim.x) {
      const unsigned long long vg = lsum[2 * l], vh = lsum[2 * l + 1];
      if (vg) atomicAdd((unsigned long long*)&a.sums[2 * l], vg);
      if (vh) atomicAdd((unsigned long long*)&a.sums[2 * l + 1], vh);
    }
  }
}

// Score update from the leaf ids the renewal walk left (boosting.rs:1381-1408,
// partition.rs:178-191): score[i] += val[leaf[i]], the score walk's own f64
// addition, without walking the tree again. No-op while the tree grows.
__global__ void __launch_bounds__(256) k_score_leaf(const StepDev* st, const uint16_t* __restrict__ leaf,
                                                    const double* __restrict__ val, double* __restrict__ score,
                                                    long long n) {
  __shared__ double lv[1024];
  if (st->done != 1) return;
  const int nl = st->num_splits + 1;
  for (int l = threadIdx.x; l < nl; l += blockDim.x) lv[l] = val[l];
  __syncthreads();
  constexpr int U = 4;
  const long long step = (long long)gridDim.x * blockDim.x * U;
  for (long long base = (long long)blockIdx.x * blockDim.x * U + threadIdx.x; base < n; base += step) {
    int lf[U];
    double sc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = min(base + (long long)u * blockDim.x, n - 1);
      lf[u] = leaf[i];
      sc[u] = score[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = base + (long long)u * blockDim.x;
      if (i < n) score[i] = sc[u] + lv[lf[u]];
    }
  }
}

// The tree's end has been applied: its state moves from grown (1) to done
// (2), so a later launch of the tail kernels (a continuation launch that
// grows other trees of a batch) leaves this tree's scores alone.
__global__ void k_tail_done(StepDev* st) {
  if (threadIdx.x == 0 && st->done == 1) st->done = 2;
}

// Shrunk leaf values (tree.rs:299-319): out[l] = round0(output * lr); with
// `renew`, output = round0(refit_leaf_output(sum g, sum h)). raw[l] keeps the
// unshrunk output for the host's tree arrays.
__global__ void k_leaf_values(const StepDev* st, const PosRec* pos, const LeafDev* leaves, const IterDev* it,
                              long long* renew_sums, int renew, SplitParamsDev p, double lr, int L, double* out,
                              double* raw) {
  const int l = threadIdx.x + blockIdx.x * blockDim.x;
  if (l >= L || st->done != 1) return;
  if (l >= st->nseq) {
    out[l] = 0.0;
    raw[l] = 0.0;
    return;
  }
  const LeafDev& Lf = leaf_at(st, pos, leaves, L, l);
  double o = Lf.output;
  if (renew) {
    const int sg = fix_shift(Lf.count, it->max_abs_g_bits), sh = fix_shift(Lf.count, it->max_abs_h_bits);
    const double sgv = ldexp((double)renew_sums[2 * l], -sg);
    const double shv = ldexp((double)renew_sums[2 * l + 1], -sh);
    o = round0(leaf_output_t(sgv, shv, p, true, true, false, (int)Lf.count, 0.0));
    renew_sums[2 * l] = 0;
    renew_sums[2 * l + 1] = 0;
  }
  raw[l] = o;
  out[l] = round0(o * lr);
}

// Each item's histogram summed over its chunks: out[w][g][entry], one unit
// per feature group (the search then reads one unit per group). units holds
// the first unit of each item and, at [kMaxBatch + 1], the item count of the
// last histogram pass (0 when it had nothing to build). A block sums 32
// consecutive entries of one item, each over the chunks c = s, s + 8, ... on
// slice s of its 8 thread slices; the slices meet in LDS. Wrapping int64
// sums: exact in any order.
__device__ __forceinline__ void slab_reduce(const long long* __restrict__ slab, const int* units, int NG,
                                            int per_unit, long long* __restrict__ out) {
  __shared__ unsigned long long part[8][32];
  const int t = threadIdx.x, bo = t & 31, sl = t >> 5;
  const int nitems = units[kMaxBatch + 1];
  const long long per_item = (long long)NG * per_unit;
  const long long eblocks = (per_item + 31) / 32;
  const size_t cstride = (size_t)NG * per_unit;   // the next chunk of the same group
  for (long long bi = blockIdx.x; bi < eblocks * nitems; bi += gridDim.x) {
    const int w = (int)(bi / eblocks);
    const long long e = (bi % eblocks) * 32 + bo;   // entry of item w: [NG][per_unit]
    const int u0 = units[w], nch = (units[w + 1] - u0) / NG;
    unsigned long long acc = 0;
    if (e < per_item && units[kUnitForm + w]) {
      // packed 32-bit bins: sum_gq * 65536 + sum_hq of one chunk each
      const int g = (int)(e / per_unit), r = (int)(e % per_unit);
      const unsigned* q = (const unsigned*)(slab + (size_t)(u0 + g) * per_unit) + r;
      const size_t cs32 = cstride * 2;
      int c = sl;
      for (; c + 8 * 7 < nch; c += 8 * 8) {
        unsigned x[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = q[(size_t)(c + 8 * j) * cs32];
#pragma unroll
        for (int j = 0; j < 8; ++j)
          acc += ((unsigned long long)(long long)((int)x[j] >> 16) << 32) + (unsigned long long)(x[j] & 0xFFFFu);
      }
      for (; c < nch; c += 8) {
        const unsigned x = q[(size_t)c * cs32];
        acc += ((unsigned long long)(long long)((int)x >> 16) << 32) + (unsigned long long)(x & 0xFFFFu);
      }
    } else if (e < per_item) {
      const int g = (int)(e / per_unit), r = (int)(e % per_unit);
      const long long* q = slab + (size_t)(u0 + g) * per_unit + r;
      int c = sl;
      for (; c + 8 * 7 < nch; c += 8 * 8) {
        long long x[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = q[(size_t)(c + 8 * j) * cstride];
#pragma unroll
        for (int j = 0; j < 8; ++j) acc += (unsigned long long)x[j];
      }
      for (; c < nch; c += 8) acc += (unsigned long long)q[(size_t)c * cstride];
    }
    part[sl][bo] = acc;
    __syncthreads();
    if (sl == 0 && e < per_item) {
      unsigned long long s = 0;
#pragma unroll
      for (int j = 0; j < 8; ++j) s += part[j][bo];
      out[(size_t)w * per_item + e] = (long long)s;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) k_slab_reduce(const long long* __restrict__ slab, const int* units, int NG,
                                                     int per_unit, long long* __restrict__ out) {
  slab_reduce(slab, units, NG, per_unit, out);
}
// The trees of a batch: grid y picks the tree.
struct SlabTab {
  const long long* slab;
  const int* units;
  long long* out;
};
__global__ void __launch_bounds__(256) k_slab_reduce_tab(const SlabTab* __restrict__ tab, int NG, int per_unit) {
  const SlabTab t = tab[blockIdx.y];
  slab_reduce(t.slab, t.units, NG, per_unit, t.out);
}

// ---------------------------------------------------------------- launchers
// Rows of at most 32 bytes are walked from registers.
template <int MODE>
static void launch_walk_m(const WalkArgs& a, int grid, hipStream_t s) {
  const size_t lds = walk_lds(a.L);
  switch (a.S) {
    case 4: hipLaunchKernelGGL((k_walk<1, MODE>), dim3(grid), dim3(256), lds, s, a); break;
    case 8: hipLaunchKernelGGL((k_walk<2, MODE>), dim3(grid), dim3(256), lds, s, a); break;
    case 16: hipLaunchKernelGGL((k_walk<4, MODE>), dim3(grid), dim3(256), lds, s, a); break;
    case 32: hipLaunchKernelGGL((k_walk<8, MODE>), dim3(grid), dim3(256), lds, s, a); break;
    default: hipLaunchKernelGGL((k_walk<0, MODE>), dim3(grid), dim3(256), lds, s, a); break;
  }
}
static void launch_walk(const WalkArgs& a, int grid, hipStream_t s) {
  if (a.mode == 0) launch_walk_m<0>(a, grid, s);
  else if (a.mode == 1) launch_walk_m<1>(a, grid, s);
  else launch_walk_m<2>(a, grid, s);
}
template <int MODE>
static void set_walk_lds_limit_m(int L) {
  const void* ks[] = {(const void*)k_walk<0, MODE>, (const void*)k_walk<1, MODE>, (const void*)k_walk<2, MODE>,
                      (const void*)k_walk<4, MODE>, (const void*)k_walk<8, MODE>};
  for (const void* k : ks) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)walk_lds(L));
}
static void set_walk_lds_limit(int L) {
  set_walk_lds_limit_m<0>(L);
  set_walk_lds_limit_m<1>(L);
  set_walk_lds_limit_m<2>(L);
}

template <int TPR, bool U4>
static void launch_hist_tu(const HistArgs& a, bool quant, int grid, size_t lds, hipStream_t s) {
  if (quant) hipLaunchKernelGGL((k_hist_int<TPR, U4>), dim3(grid), dim3(HistBlock<true>::threads), lds, s, a);
  else hipLaunchKernelGGL((k_hist_f64<TPR, U4>), dim3(grid), dim3(HistBlock<false>::threads), lds, s, a);
}
template <int TPR>
static void launch_hist_t(const HistArgs& a, bool quant, int grid, size_t lds, hipStream_t s) {
  if (a.u4) launch_hist_tu<TPR, true>(a, quant, grid, lds, s);
  else launch_hist_tu<TPR, false>(a, quant, grid, lds, s);
}

static void launch_hist(const HistArgs& a, bool quant, int tpr, int grid, size_t lds, hipStream_t s) {
  switch (tpr) {
    case 1: launch_hist_t<1>(a, quant, grid, lds, s); break;
    case 2: launch_hist_t<2>(a, quant, grid, lds, s); break;
    case 4: launch_hist_t<4>(a, quant, grid, lds, s); break;
    case 8: launch_hist_t<8>(a, quant, grid, lds, s); break;
    case 16: launch_hist_t<16>(a, quant, grid, lds, s); break;
    default: launch_hist_t<32>(a, quant, grid, lds, s); break;
  }
}

// Resident blocks per CU of the histogram kernel for this form, row layout
// and LDS size (0 when the runtime cannot tell).
template <int TPR>
static int hist_occupancy_t(bool quant, bool u4, size_t lds) {
  int n = 0;
  const hipError_t e =
      quant ? (u4 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_hist_int<TPR, true>, HistBlock<true>::threads, lds)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_hist_int<TPR, false>, HistBlock<true>::threads, lds))
            : (u4 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_hist_f64<TPR, true>, HistBlock<false>::threads, lds)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_hist_f64<TPR, false>, HistBlock<false>::threads, lds));
  return e == hipSuccess ? n : 0;
}
static int hist_occupancy(bool quant, bool u4, int tpr, size_t lds) {
  switch (tpr) {
    case 1: return hist_occupancy_t<1>(quant, u4, lds);
    case 2: return hist_occupancy_t<2>(quant, u4, lds);
    case 4: return hist_occupancy_t<4>(quant, u4, lds);
    case 8: return hist_occupancy_t<8>(quant, u4, lds);
    case 16: return hist_occupancy_t<16>(quant, u4, lds);
    default: return hist_occupancy_t<32>(quant, u4, lds);
  }
}

template <int TPR, bool U4>
static void launch_hist_tab_tu(const HistArgs* tab, int T, bool quant, int grid, size_t lds, hipStream_t s) {
  if (quant) hipLaunchKernelGGL((k_hist_int_tab<TPR, U4>), dim3(grid, T), dim3(HistBlock<true>::threads), lds, s, tab);
  else hipLaunchKernelGGL((k_hist_f64_tab<TPR, U4>), dim3(grid, T), dim3(HistBlock<false>::threads), lds, s, tab);
}
template <int TPR>
static void launch_hist_tab_t(const HistArgs* tab, int T, bool quant, bool u4, int grid, size_t lds, hipStream_t s) {
  if (u4) launch_hist_tab_tu<TPR, true>(tab, T, quant, grid, lds, s);
  else launch_hist_tab_tu<TPR, false>(tab, T, quant, grid, lds, s);
}

// The histogram pass of T trees (tab: their arguments in device memory).
static void launch_hist_tab(const HistArgs* tab, int T, bool quant, bool u4, int tpr, int grid, size_t lds,
                            hipStream_t s) {
  switch (tpr) {
    case 1: launch_hist_tab_t<1>(tab, T, quant, u4, grid, lds, s); break;
    case 2: launch_hist_tab_t<2>(tab, T, quant, u4, grid, lds, s); break;
    case 4: launch_hist_tab_t<4>(tab, T, quant, u4, grid, lds, s); break;
    case 8: launch_hist_tab_t<8>(tab, T, quant, u4, grid, lds, s); break;
    case 16: launch_hist_tab_t<16>(tab, T, quant, u4, grid, lds, s); break;
    default: launch_hist_tab_t<32>(tab, T, quant, u4, grid, lds, s); break;
  }
}

template <bool U4>
static void set_hist_lds_limits_u(size_t bytes) {
  const void* ks[] = {(const void*)k_hist_f64<1, U4>, (const void*)k_hist_f64<2, U4>, (const void*)k_hist_f64<4, U4>,
                      (const void*)k_hist_f64<8, U4>, (const void*)k_hist_f64<16, U4>, (const void*)k_hist_f64<32, U4>,
                      (const void*)k_hist_int<1, U4>, (const void*)k_hist_int<2, U4>, (const void*)k_hist_int<4, U4>,
                      (const void*)k_hist_int<8, U4>, (const void*)k_hist_int<16, U4>, (const void*)k_hist_int<32, U4>,
                      (const void*)k_hist_f64_tab<1, U4>, (const void*)k_hist_f64_tab<2, U4>,
                      (const void*)k_hist_f64_tab<4, U4>, (const void*)k_hist_f64_tab<8, U4>,
                      (const void*)k_hist_f64_tab<16, U4>, (const void*)k_hist_f64_tab<32, U4>,
                      (const void*)k_hist_int_tab<1, U4>, (const void*)k_hist_int_tab<2, U4>,
                      (const void*)k_hist_int_tab<4, U4>, (const void*)k_hist_int_tab<8, U4>,
                      (const void*)k_hist_int_tab<16, U4>, (const void*)k_hist_int_tab<32, U4>};
  for (const void* k : ks) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}
static void set_hist_lds_limits(size_t bytes) {
  set_hist_lds_limits_u<false>(bytes);
  set_hist_lds_limits_u<true>(bytes);
}

static size_t part_lds(int L) { return (size_t)L * (sizeof(PosRec) + sizeof(KeyRec)); }
static void set_part_lds_limit(size_t bytes) {
  const void* ks[] = {(const void*)k_grow_plan<4>, (const void*)k_grow_plan<8>, (const void*)k_grow_plan<16>,
                      (const void*)k_grow_plan_tab<4>, (const void*)k_grow_plan_tab<8>,
                      (const void*)k_grow_plan_tab<16>};
  for (const void* k : ks) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

static void launch_split(const SplitArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_grow_split, dim3(a.F, a.root ? 1 : kMaxBatch), dim3(128), 0, s, a);
}

// The round's plan (one block).
static void launch_plan(const PartArgs& a, hipStream_t s) {
  if (a.L <= 256) hipLaunchKernelGGL(k_grow_plan<4>, dim3(1), dim3(kPartThreads), part_lds(a.L), s, a);
  else if (a.L <= 512) hipLaunchKernelGGL(k_grow_plan<8>, dim3(1), dim3(kPartThreads), part_lds(a.L), s, a);
  else hipLaunchKernelGGL(k_grow_plan<16>, dim3(1), dim3(kPartThreads), part_lds(a.L), s, a);
}

// The partition pass; rows move in the widest vectors that divide a group's bytes.
static void launch_partition(const PartArgs& a, int grid, hipStream_t s) {
  if (a.gh_bytes == 2) {
    if (a.FB % 16 == 0) hipLaunchKernelGGL((k_grow_part<16, 2>), dim3(grid), dim3(kPartThreads), 0, s, a);
    else if (a.FB % 8 == 0) hipLaunchKernelGGL((k_grow_part<8, 2>), dim3(grid), dim3(kPartThreads), 0, s, a);
    else hipLaunchKernelGGL((k_grow_part<4, 2>), dim3(grid), dim3(kPartThreads), 0, s, a);
  } else {
    if (a.FB % 16 == 0) hipLaunchKernelGGL((k_grow_part<16, 8>), dim3(grid), dim3(kPartThreads), 0, s, a);
    else if (a.FB % 8 == 0) hipLaunchKernelGGL((k_grow_part<8, 8>), dim3(grid), dim3(kPartThreads), 0, s, a);
    else hipLaunchKernelGGL((k_grow_part<4, 8>), dim3(grid), dim3(kPartThreads), 0, s, a);
  }
}

static int part_tile() { return kPartTile; }

// The split pass of T trees (grid z picks the tree).
static void launch_split_tab(const SplitArgs* tab, int T, int F, bool root, hipStream_t s) {
  hipLaunchKernelGGL(k_grow_split_tab, dim3(F, root ? 1 : kMaxBatch, T), dim3(128), 0, s, tab);
}

// The plans and partition passes of T trees; `a` is any of them (the shapes
// they share: leaves, group width, gradient width).
static void launch_plan_tab(const PartArgs* tab, int T, const PartArgs& a, hipStream_t s) {
  const dim3 pg(1, T);
  if (a.L <= 256) hipLaunchKernelGGL(k_grow_plan_tab<4>, pg, dim3(kPartThreads), part_lds(a.L), s, tab);
  else if (a.L <= 512) hipLaunchKernelGGL(k_grow_plan_tab<8>, pg, dim3(kPartThreads), part_lds(a.L), s, tab);
  else hipLaunchKernelGGL(k_grow_plan_tab<16>, pg, dim3(kPartThreads), part_lds(a.L), s, tab);
}
static void launch_partition_tab(const PartArgs* tab, int T, const PartArgs& a, int grid, hipStream_t s) {
  const dim3 g(grid, T);
  if (a.gh_bytes == 2) {
    if (a.FB % 16 == 0) hipLaunchKernelGGL((k_grow_part_tab<16, 2>), g, dim3(kPartThreads), 0, s, tab);
    else if (a.FB % 8 == 0) hipLaunchKernelGGL((k_grow_part_tab<8, 2>), g, dim3(kPartThreads), 0, s, tab);
    else hipLaunchKernelGGL((k_grow_part_tab<4, 2>), g, dim3(kPartThreads), 0, s, tab);
  } else {
    if (a.FB % 16 == 0) hipLaunchKernelGGL((k_grow_part_tab<16, 8>), g, dim3(kPartThreads), 0, s, tab);
    else if (a.FB % 8 == 0) hipLaunchKernelGGL((k_grow_part_tab<8, 8>), g, dim3(kPartThreads), 0, s, tab);
    else hipLaunchKernelGGL((k_grow_part_tab<4, 8>), g, dim3(kPartThreads), 0, s, tab);
  }
}

}  // namespace lgbmhip

// hs_device.hpp -- device-side synchronization primitives for gfx950 (CDNA4).
//
// Header-only HIP.  Every primitive is *collective per workgroup*: all threads of
// the workgroup call it, thread 0 is the master that performs the atomics and the
// polling, and the result is broadcast through LDS -- the call shape of the
// reference's device API (e.g. cudaSemaphoreSpinWait(sem, isWriter, maxSemCount,
// semaphoreBuffers, NUM_SM), Heterosync/cudaLocksSemaphoreSpin.cu:144-154).
//
// Memory model: all cross-workgroup words are accessed with __hip_atomic_* at
// agent scope.  A release is one `buffer_wbl2 sc1` + store/RMW, an acquire is
// one `buffer_inv sc1` after the last poll (polls themselves are relaxed), so the
// L2 of each XCD is written back / invalidated once per synchronisation, not per
// poll.  Every wait is bounded by ArenaView::spin_limit and also watches a host
// abort flag, so a kernel always terminates; the failure is reported in
// ArenaView::status.
#pragma once

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hs {

constexpr int kPadWords = 64;  // 256 B between independently contended words
constexpr int kSemWords = 3 * 32;  // semaphore record: 3 x 128 B lines
constexpr int kSemCount = 0, kSemWriterWaiting = 1, kSemMax = 2;  // line 0
constexpr int kSemLock = 32;                                       // line 1
constexpr int kSemPriority = 64;                                   // line 2
constexpr int kSemQueueHead = 65, kSemQueueTail = 66;              // line 2 (queued kind)
constexpr int kSemTurnHead = 67, kSemTurnTail = 68;                // line 2 (queued kind, global form)
constexpr int kMaxGroups = 1024;
constexpr uint32_t kDefaultSpinLimit = 1u << 23;

enum : int { kOk = 0, kErrAborted = 4, kErrTimeout = 5 };

// POD view of the lock arena passed by value to kernels.  Replaces the
// cudaLockData_t descriptor (Heterosync/cudaLocks.h:4-16) and the six raw
// arena pointers of the app copies (Reduce/reduce_gsrb.cu:237-244).
struct ArenaView {
  uint32_t *bar;          // barrier words (layout: BarOff below)
  uint32_t *lf_in;        // lock-free barrier in-flags  [max_blocks]
  uint32_t *lf_out;       // lock-free barrier out-flags [max_blocks]
  uint32_t *lf_gin;       // per-group in-flags  [num_groups * max_blocks]
  uint32_t *lf_gout;      // per-group out-flags [num_groups * max_blocks]
  uint32_t *mutex_heads;  // [num_mutexes * num_groups] padded
  uint32_t *mutex_tails;  // [num_mutexes * num_groups] padded
  int32_t *mutex_rings;   // [num_mutexes * num_groups * ring_stride]
  uint32_t *sems;         // [num_sems * num_groups * kSemWords]
  uint32_t *sem_queues;   // [num_sems * num_groups * queue_len] ticket slots
  uint32_t *episode;      // invariant slots [max_blocks]
  uint32_t *status;       // device status word (first error wins)
  volatile uint32_t *abort_flag;  // host-pinned, set by hs_arena_set_abort
  int max_blocks;
  int ring_stride;
  int num_groups;
  int num_mutexes;
  int num_sems;
  int queue_len;
  uint32_t spin_limit;
};

// word offsets (in kPadWords units) inside ArenaView::bar
struct BarOff {
  static constexpr int kFlatCount = 0, kFlatGen = 1, kDevCount = 2, kDevGen = 3,
                       kTree0 = 4, kTree1 = 5, kNaiveAllCount = 6, kNaiveAllGen = 7,
                       kGroupBase = 8;  // then kMaxGroups counts, kMaxGroups gens,
                                        // 2*kMaxGroups tree counters, kMaxGroups
                                        // local-barrier counts and gens, kMaxGroups
                                        // XCD write-back acknowledgement counts,
                                        // then kCuSlots per-CU invalidate slots
  static constexpr int kCuSlots = 2048;    // XCC_ID(3) | SE(3) | SH(1) | CU(4)
  static constexpr int kCuSlotWords = 16;  // 64 B: {claimed episode, done episode}
  static constexpr int kCuBase = (kGroupBase + 7 * kMaxGroups) * kPadWords;
  static constexpr int kWords = kCuBase + kCuSlots * kCuSlotWords;
  __host__ __device__ static constexpr int word(int slot) { return slot * kPadWords; }
  __host__ __device__ static constexpr int group_count(int g) { return word(kGroupBase + g); }
  __host__ __device__ static constexpr int group_gen(int g) { return word(kGroupBase + kMaxGroups + g); }
  __host__ __device__ static constexpr int group_tree(int g, int k) {
    return word(kGroupBase + 2 * kMaxGroups + 2 * g + k);
  }
  __host__ __device__ static constexpr int local_count(int g) { return word(kGroupBase + 4 * kMaxGroups + g); }
  __host__ __device__ static constexpr int local_gen(int g) { return word(kGroupBase + 5 * kMaxGroups + g); }
  __host__ __device__ static constexpr int group_ack(int g) { return word(kGroupBase + 6 * kMaxGroups + g); }
  __host__ __device__ static constexpr int cu_slot(int c) { return kCuBase + c * kCuSlotWords; }
};

// ---------------------------------------------------------------------------
// atomics
__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_acquire(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t add_release(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t add_acq_rel(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t add_relaxed(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t sub_release(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas_acquire(uint32_t *p, uint32_t expected, uint32_t desired) {
  return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_ACQUIRE,
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// value-returning CAS (atomicCAS semantics)
__device__ __forceinline__ uint32_t cas_value(uint32_t *p, uint32_t expected, uint32_t desired) {
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expected;
}
__device__ __forceinline__ uint32_t xchg_release(uint32_t *p, uint32_t v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
// atomicInc: old >= limit ? 0 : old + 1  (hardware global_atomic_inc)
__device__ __forceinline__ uint32_t inc_wrap(uint32_t *p, uint32_t limit) {
  return __builtin_amdgcn_atomic_inc32(p, limit, __ATOMIC_RELAXED, "agent");
}
__device__ __forceinline__ void fence_acquire() {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
__device__ __forceinline__ void fence_release() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}

__device__ __forceinline__ bool is_master() {
  return threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0;
}

// physical XCD of the calling wave (HW_REG_XCC_ID, hwreg id 20, bits [3:0])
__device__ __forceinline__ uint32_t xcc_id() {
  return __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11));
}

// the calling wave's compute unit, unique on the device: XCC_ID with HW_ID's
// SE (bits 15:13), SH (12) and CU (11:8) fields; < BarOff::kCuSlots
__device__ __forceinline__ uint32_t cu_index() {
  const uint32_t hw = __builtin_amdgcn_s_getreg((4) | (0 << 6) | ((32 - 1) << 11));  // HW_REG_HW_ID
  return ((xcc_id() & 7u) << 8) | (((hw >> 13) & 7u) << 5) | (((hw >> 12) & 1u) << 4) | ((hw >> 8) & 15u);
}

__device__ __forceinline__ void sleep_units(uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) __builtin_amdgcn_s_sleep(1);
}

// record the first failure in the arena status word
__device__ __forceinline__ void set_status(const ArenaView &a, uint32_t code) {
  uint32_t zero = 0;
  __hip_atomic_compare_exchange_strong(a.status, &zero, code, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool failed(const ArenaView &a) { return ld_relaxed(a.status) != 0; }

// Polls between two looks at the host abort flag and the arena status.  The
// abort flag lives in host memory: every look is a PCIe round trip, and when
// every waiting workgroup looks at once the reads queue behind each other
// (about 50 ns each, measured: 1024 waiters add ~50 us to a barrier).  A wait
// that ends within this many polls (~0.3 ms) never touches host memory.
constexpr uint32_t kPollsPerCheck = 2048;

// Bounded poll: waits until pred() holds.  Exponential s_sleep backoff capped at
// max_sleep units of 64 clocks.  Returns kOk, kErrTimeout or kErrAborted.
template <class Pred>
__device__ __forceinline__ int spin_until(const ArenaView &a, Pred pred, uint32_t max_sleep = 2) {
  uint32_t polls = 0, backoff = 0;
  while (!pred()) {
    if (++polls >= a.spin_limit) { set_status(a, kErrTimeout); return kErrTimeout; }
    if ((polls & (kPollsPerCheck - 1)) == 0) {
      if (a.abort_flag && __hip_atomic_load((const uint32_t *)a.abort_flag, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_SYSTEM)) {
        set_status(a, kErrAborted);
        return kErrAborted;
      }
      if (failed(a)) return kErrTimeout;
    }
    sleep_units(backoff);
    backoff = backoff ? (backoff < max_sleep ? backoff * 2 : max_sleep) : 1;
  }
  return kOk;
}

// The reference's capped exponential backoff ((b<<1)+1)&(MAX_BACKOFF-1),
// Heterosync/cudaLocksBarrierAtomic.cu:45, counted in s_sleep units.
__device__ __forceinline__ uint32_t ebo_next(uint32_t b, uint32_t cap = 1024) {
  return ((b << 1) + 1) & (cap - 1);
}

// ---------------------------------------------------------------------------
// Barriers.  `gen` is the master's private copy of the generation word it is
// waiting to see change (a sense-reversing barrier where the sense is gen&1).
// Call barrier_init once per kernel (collective) before the first episode.

struct BarrierState {
  uint32_t gen;     // generation observed at kernel start (master only)
  uint32_t lgen;    // group generation (XCD release)
  uint32_t lf_phase;
};

// Flat SRB over all `nwg` workgroups: G-CPUSRB semantics
// (Heterosync/cudaLocksBarrierAtomicSRB.cu:17-61, Reduce/reduce_gsrb.cu:17-61 with
// naive=true) -- one counter and one release word, the last arriver releases.
__device__ __forceinline__ int barrier_flat_master(const ArenaView &a, BarrierState &s, uint32_t nwg,
                                                   uint32_t sleep = 2) {
  uint32_t *cnt = a.bar + BarOff::word(BarOff::kFlatCount);
  uint32_t *gen = a.bar + BarOff::word(BarOff::kFlatGen);
  const uint32_t my = s.gen;
  int rc = kOk;
  if (add_release(cnt, 1u) == nwg - 1) {
    fence_acquire();
    st_relaxed(cnt, 0u);
    st_release(gen, my + 1);
  } else {
    rc = spin_until(a, [&] { return ld_relaxed(gen) != my; }, sleep);
    fence_acquire();
  }
  s.gen = my + 1;
  return rc;
}

// Two-level SRB (the paper's G-SRB, Reduce/reduce_gsrb.cu:135-269): workgroup b
// belongs to logical group b % G; the last arriver of a group carries the group
// to the device counter; the last group releases everyone through one word.
__device__ __forceinline__ int barrier_two_level_master(const ArenaView &a, BarrierState &s,
                                                        uint32_t nwg, uint32_t ngroups,
                                                        uint32_t wg, uint32_t sleep = 2) {
  const uint32_t G = ngroups < nwg ? ngroups : nwg;
  const uint32_t g = wg % G;
  const uint32_t gsize = (nwg - g + G - 1) / G;
  uint32_t *gcnt = a.bar + BarOff::group_count(g);
  uint32_t *dcnt = a.bar + BarOff::word(BarOff::kDevCount);
  uint32_t *dgen = a.bar + BarOff::word(BarOff::kDevGen);
  const uint32_t my = s.gen;
  int rc = kOk;
  bool released = false;
  if (add_release(gcnt, 1u) == gsize - 1) {
    fence_acquire();
    st_relaxed(gcnt, 0u);
    if (add_acq_rel(dcnt, 1u) == G - 1) {
      st_relaxed(dcnt, 0u);
      st_release(dgen, my + 1);
      released = true;
    }
  }
  if (!released) {
    rc = spin_until(a, [&] { return ld_relaxed(dgen) != my; }, sleep);
    fence_acquire();
  }
  s.gen = my + 1;
  return rc;
}

// XCD-local ordering.  The L2 of an XCD is shared by all of its CUs and the
// vector L1 is write-through, so for the release side it is enough that every
// writer's stores have completed into its XCD's L2 (s_waitcnt vmcnt(0),
// executed by every wave) and that one workgroup per XCD then writes that L2
// back (buffer_wbl2 sc1 acts on the whole L2 of the issuing XCD).  The acquire
// side cannot be shared the same way: a reader must drop its CU's L1, and
// gfx950 has no L1-only invalidate at agent scope -- buffer_inv sc0 is a
// workgroup-scope operation and leaves stale L1 lines in place outside
// thread-group-split mode (measured: ~93% stale reads of a line rewritten every
// episode, zero with buffer_inv sc1).  Every workgroup therefore runs the
// agent-scope acquire (buffer_inv sc1: its L1 and its XCD's L2) itself.
__device__ __forceinline__ void wait_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// XCD-hierarchical barrier.  Workgroup b belongs to group b % X.  gfx950
// dispatches workgroups to its 8 XCDs round-robin, starting from wherever the
// previous dispatch stopped, so the members of one group normally share one
// XCD and its L2 -- but which XCD is not known in advance, and nothing
// guarantees the round-robin.  Each arrival therefore records its physical XCD
// (HW_REG_XCC_ID) in a byte of a 64-bit per-group count; the last arriver (the
// group's leader) sees from the count whether every member ran on its own XCD.
//   Local=true  (HS_BAR_XCD): every wave waits for its stores before arriving
//     (the caller does).  Fast path (all members on the leader's XCD): the
//     leader writes that L2 back once for everyone, joins the device counter,
//     waits for the device word and forwards the release to a per-group word,
//     so at most one workgroup per group polls the device word.  Slow path
//     (some member elsewhere): the leader first publishes a write-back request
//     tagged with its XCD; members on other XCDs write their own L2 back and
//     acknowledge, and the leader waits for every acknowledgement.  Groups of
//     more than 255 members write back in every workgroup instead.  In all
//     cases every workgroup runs the agent-scope acquire after the release.
//   Local=false (HS_BAR_XCD_FLAT_RELEASE): per-group arrival counters but every
//     workgroup runs agent-scope release/acquire and polls the device word.
// Generations stay below 2^31; release words with bit 31 set are write-back
// requests: bit 31 | (generation & 0x0FFFFFFF) << 3 | leader XCD.
__device__ __forceinline__ uint32_t gen_next(uint32_t g) { return (g + 1u) & 0x7FFFFFFFu; }
__device__ __forceinline__ uint32_t wb_request(uint32_t gnext, uint32_t xcd) {
  return 0x80000000u | ((gnext & 0x0FFFFFFFu) << 3) | xcd;
}
__device__ __forceinline__ uint64_t add64_relaxed(uint64_t *p, uint64_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// sum of the eight byte counters of v (valid while the sum is below 256)
__device__ __forceinline__ uint32_t byte_sum(uint64_t v) {
  return (uint32_t)((v * 0x0101010101010101ull) >> 56);
}

// Acquire side of an episode, shared by the workgroups of one compute unit.
// buffer_inv sc1 drops the CU's L1 and the XCD's L2, and such invalidates
// serialize per XCD; the CU's L1 is shared by its workgroups, so one of them per
// CU and episode is enough.  The first workgroup of the CU to see the release
// claims the episode (exchange on the CU's slot) and invalidates; the others
// wait until it has published the episode as done.  Slot words hold episode
// numbers (generations), so no reset is needed between episodes or launches;
// the arena's reset zeroes them with the other barrier words.
__device__ __forceinline__ int cu_acquire(const ArenaView &a, uint32_t episode) {
  uint32_t *slot = a.bar + BarOff::cu_slot((int)cu_index());
  if (__hip_atomic_exchange(slot, episode, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != episode) {
    fence_acquire();  // this CU's L1 and this XCD's L2
    wait_stores();    // the invalidate has completed before it is published
    st_relaxed(slot + 1, episode);
    return kOk;
  }
  return spin_until(a, [&] { return ld_relaxed(slot + 1) == episode; }, 1);
}

template <bool Local>
__device__ __forceinline__ int barrier_xcd_master(const ArenaView &a, BarrierState &s, uint32_t nwg,
                                                  uint32_t nxcd, uint32_t wg, bool acquire = true,
                                                  uint32_t sleep = 0) {
  // polls back off up to `sleep` s_sleep units (0: 1 for leaders, 2 for members)
  const uint32_t lsleep = sleep ? sleep : 1u, msleep = sleep ? sleep : 2u;
  const uint32_t X = nxcd < nwg ? nxcd : nwg;
  const uint32_t x = wg % X;
  const uint32_t xsize = (nwg - x + X - 1) / X;
  uint32_t *dcnt = a.bar + BarOff::word(BarOff::kDevCount);
  uint32_t *dgen = a.bar + BarOff::word(BarOff::kDevGen);
  const uint32_t my = s.gen, nx = gen_next(my);
  int rc = kOk;
  if (!Local) {
    uint32_t *xcnt = a.bar + BarOff::group_count(x);
    fence_release();
    if (add_relaxed(xcnt, 1u) == xsize - 1) {
      st_relaxed(xcnt, 0u);
      if (add_relaxed(dcnt, 1u) == X - 1) {
        st_relaxed(dcnt, 0u);
        wait_stores();
        st_relaxed(dgen, nx);
      } else {
        rc = spin_until(a, [&] { return ld_relaxed(dgen) == nx; }, lsleep);
      }
    } else {
      rc = spin_until(a, [&] { return ld_relaxed(dgen) == nx; }, msleep);
    }
    fence_acquire();
    s.gen = nx;
    return rc;
  }
  uint64_t *xcnt = reinterpret_cast<uint64_t *>(a.bar + BarOff::group_count(x));
  uint32_t *xgen = a.bar + BarOff::group_gen(x);
  uint32_t *ack = a.bar + BarOff::group_ack(x);
  const bool bytes = xsize <= 255;
  const uint32_t me = xcc_id() & 7u;  // gfx950: XCC_ID is 0..7
  if (!bytes) fence_release();
  const uint64_t old = add64_relaxed(xcnt, bytes ? 1ull << (8 * me) : 1ull);
  const uint32_t before = bytes ? byte_sum(old) : (uint32_t)old;
  if (before == xsize - 1) {
    __hip_atomic_store(xcnt, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (bytes) {
      const uint32_t local = (uint32_t)((old >> (8 * me)) & 0xFFu) + 1u;
      if (local != xsize) {
        // members on other XCDs write their own L2 back before the release
        st_relaxed(xgen, wb_request(nx, me));
        const uint32_t want = xsize - local;
        rc = spin_until(a, [&] { return ld_relaxed(ack) == want; }, 1);
        st_relaxed(ack, 0u);
      }
    }
    fence_release();  // this XCD's L2 -> memory (all members sharing it)
    if (rc == kOk) {
      if (add_relaxed(dcnt, 1u) == X - 1) {
        st_relaxed(dcnt, 0u);
        wait_stores();
        st_relaxed(dgen, nx);
      } else {
        rc = spin_until(a, [&] { return ld_relaxed(dgen) == nx; }, lsleep);
      }
    }
    st_relaxed(xgen, nx);
    if (acquire && rc == kOk) rc = cu_acquire(a, nx);
  } else {
    bool remote = !bytes;
    rc = spin_until(a, [&] {
      const uint32_t v = ld_relaxed(xgen);
      if (v == nx) return true;
      if (v == wb_request(nx, v & 7u) && (v & 7u) != me && !remote) {
        fence_release();
        add_relaxed(ack, 1u);
        remote = true;
      }
      return false;
    }, msleep);
    if (acquire && rc == kOk) rc = cu_acquire(a, nx);
  }
  wait_stores();  // the invalidate completes before the other waves of this workgroup load
  s.gen = nx;
  return rc;
}

// HeteroSync atomic tree barrier (Heterosync/cudaLocksBarrierAtomic.cu:6-170):
// two consecutive counter barriers stand in for sense reversal, completion is
// detected by CAS-polling the counter (the reference's atomicCAS(c, n, 0) storm),
// with a per-group level first when a group holds more than one workgroup.
__device__ __forceinline__ int tree_sub_master(const ArenaView &a, uint32_t *cnt, uint32_t n,
                                               bool backoff) {
  add_release(cnt, 1u);
  uint32_t b = 1, polls = 0;
  for (;;) {
    // the reference treats "counter already reset to 0" as done (cudaLocksBarrierAtomic.cu:34)
    const uint32_t old = cas_value(cnt, n, 0u);
    if (old == 0u) break;
    if (++polls >= a.spin_limit) { set_status(a, kErrTimeout); return kErrTimeout; }
    if ((polls & 255u) == 0 && failed(a)) return kErrTimeout;
    if (backoff) { sleep_units(b > 16 ? 16 : b); b = ebo_next(b); }
  }
  fence_acquire();
  return kOk;
}

__device__ __forceinline__ int barrier_tree_master(const ArenaView &a, uint32_t nwg, uint32_t ngroups,
                                                   uint32_t wg, int phase_mask) {
  const uint32_t G = ngroups < nwg ? ngroups : nwg;
  const uint32_t g = wg % G;
  const uint32_t per = (nwg + G - 1) / G;
  const uint32_t gsize = (nwg - g + G - 1) / G;
  int rc = kOk;
  if (per > 1) {
    // local barrier, global barrier among the groups' first workgroups, local barrier
    if (phase_mask & 1) {
      rc |= tree_sub_master(a, a.bar + BarOff::group_tree(g, 0), gsize, false);
      rc |= tree_sub_master(a, a.bar + BarOff::group_tree(g, 1), gsize, false);
    }
    if ((phase_mask & 2) && wg / G == 0) {
      rc |= tree_sub_master(a, a.bar + BarOff::word(BarOff::kTree0), G, true);
      rc |= tree_sub_master(a, a.bar + BarOff::word(BarOff::kTree1), G, true);
    }
    if (phase_mask & 4) {
      rc |= tree_sub_master(a, a.bar + BarOff::group_tree(g, 0), gsize, false);
      rc |= tree_sub_master(a, a.bar + BarOff::group_tree(g, 1), gsize, false);
    }
  } else if (phase_mask & 2) {
    rc |= tree_sub_master(a, a.bar + BarOff::word(BarOff::kTree0), G, true);
    rc |= tree_sub_master(a, a.bar + BarOff::word(BarOff::kTree1), G, true);
  }
  return rc;
}

// Lock-free flag barrier (Heterosync/cudaLocksBarrierFast.cu:10-330).  Flags hold
// episode numbers instead of 1/0 so no reset store is needed: workgroup b
// publishes in[b] = e; the collector polls all in-flags with every thread of its
// workgroup (lf_wait_all), then publishes out[b] = e for every member
// (lf_publish).  Both are collective (all threads of the collector).
__device__ __forceinline__ int lf_wait_all(const ArenaView &a, const uint32_t *in, uint32_t n,
                                           uint32_t e) {
  int rc = kOk;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t *f = in + i;
    rc |= spin_until(a, [&] { return ld_relaxed(f) == e; }, 1);
  }
  fence_acquire();
  return rc;
}
__device__ __forceinline__ void lf_publish(uint32_t *out, uint32_t n, uint32_t e) {
  fence_release();
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) st_relaxed(out + i, e);
}

// ---------------------------------------------------------------------------
// Mutexes (collective).  Handles index the arena; `local` selects the per-group
// ("per-SM") instance group g, Heterosync/cudaLocksMutexSpin.cu:40-95.

__device__ __forceinline__ uint32_t *mutex_head(const ArenaView &a, uint32_t m, uint32_t g) {
  return a.mutex_heads + (size_t)(m * a.num_groups + g) * kPadWords;
}
__device__ __forceinline__ uint32_t *mutex_tail(const ArenaView &a, uint32_t m, uint32_t g) {
  return a.mutex_tails + (size_t)(m * a.num_groups + g) * kPadWords;
}
__device__ __forceinline__ int32_t *mutex_ring(const ArenaView &a, uint32_t m, uint32_t g) {
  return a.mutex_rings + (size_t)(m * a.num_groups + g) * a.ring_stride;
}

// spin (Heterosync/cudaLocksMutexSpin.cu:11-36) and EBO (cudaLocksMutexEBO.cu:13-53)
template <bool Ebo>
__device__ __forceinline__ int mutex_cas_lock_master(const ArenaView &a, uint32_t *head) {
  uint32_t b = 1, polls = 0;
  while (!cas_acquire(head, 0u, 1u)) {
    if (++polls >= a.spin_limit) { set_status(a, kErrTimeout); return kErrTimeout; }
    if ((polls & 255u) == 0 && failed(a)) return kErrTimeout;
    if (Ebo) { sleep_units(b > 64 ? 64 : b); b = ebo_next(b); }
  }
  return kOk;
}
__device__ __forceinline__ void mutex_cas_unlock_master(uint32_t *head) { xchg_release(head, 0u); }

// fetch-and-add ticket (Heterosync/cudaLocksMutexFA.cu:13-61); tickets wrap mod 2^32
__device__ __forceinline__ int mutex_fa_lock_master(const ArenaView &a, uint32_t *ticket,
                                                    uint32_t *turn) {
  const uint32_t mine = add_relaxed(ticket, 1u);
  int rc = spin_until(a, [&] { return ld_relaxed(turn) == mine; }, 1);
  fence_acquire();
  return rc;
}
__device__ __forceinline__ void mutex_fa_unlock_master(uint32_t *turn) { add_release(turn, 1u); }

// ring-buffer "sleep" mutex (Heterosync/cudaLocksMutexSleep.cu:19-199)
__device__ __forceinline__ int mutex_sleep_lock_master(const ArenaView &a, uint32_t *tail,
                                                       int32_t *ring, uint32_t *slot_out) {
  const uint32_t slot = inc_wrap(tail, (uint32_t)a.max_blocks);
  *slot_out = slot;
  uint32_t *w = (uint32_t *)(ring + slot);
  int rc = spin_until(a, [&] { return ld_relaxed(w) == 1u; }, 1);
  fence_acquire();
  return rc;
}
__device__ __forceinline__ void mutex_sleep_unlock_master(const ArenaView &a, int32_t *ring,
                                                          uint32_t slot) {
  const uint32_t next = slot >= (uint32_t)a.max_blocks ? 0u : slot + 1u;
  st_relaxed((uint32_t *)(ring + slot), 0xFFFFFFFFu);
  st_release((uint32_t *)(ring + next), 1u);
}

// ---------------------------------------------------------------------------
// Semaphores.  Record per (sem, group): count (init capacity+1), writerWaiting,
// maxCount (capacity+1), lock, priority -- the state of
// Heterosync/cudaLocksSemaphoreSpin.cu:11-26, one record per 3 cache lines.
__device__ __forceinline__ uint32_t *sem_rec(const ArenaView &a, uint32_t sem, uint32_t g) {
  return a.sems + (size_t)(sem * a.num_groups + g) * kSemWords;
}

// one TryWait attempt under the record lock (cudaLocksSemaphoreSpin.cu:36-132).
// A reader takes `units` of the capacity (1 in the reference; a weighted
// acquisition, such as a memory budget, takes more), a writer all of it.
// Returns 1 acquired, 0 not acquired, <0 error.
__device__ __forceinline__ int sem_try_wait_master(const ArenaView &a, uint32_t *rec, bool writer,
                                                   uint32_t maxc, uint32_t units = 1u) {
  if (!cas_acquire(rec + kSemLock, 0u, 1u)) return 0;
  const uint32_t cnt = ld_relaxed(rec + kSemCount);
  bool ok;
  if (writer) ok = (cnt == maxc);
  else ok = (ld_relaxed(rec + kSemWriterWaiting) == 0u) && cnt > units;
  if (!ok) {
    if (writer) st_relaxed(rec + kSemWriterWaiting, 1u);
    xchg_release(rec + kSemLock, 0u);
    return 0;
  }
  if (writer) {
    st_relaxed(rec + kSemCount, cnt - maxc);
    st_relaxed(rec + kSemWriterWaiting, 0u);
  } else {
    st_relaxed(rec + kSemCount, cnt - units);
  }
  xchg_release(rec + kSemLock, 0u);
  return 1;
}

// Wait, spin / EBO / priority variants:
//   Spin      cudaSemaphoreSpinWait           (cudaLocksSemaphoreSpin.cu:144-154)
//   EBO       cudaSemaphoreEBOWait            (cudaLocksSemaphoreEBO.cu:150-185):
//             writers exponential, readers +5 linear backoff
//   Priority  cudaSemaphoreSpinTryWaitPriority (cudaLocksSemaphoreSpin.cu:411-533):
//             first wait (with backoff) while any poster holds the priority count
template <bool Ebo, bool Priority>
__device__ __forceinline__ int sem_wait_master(const ArenaView &a, uint32_t *rec, bool writer,
                                               uint32_t units = 1u) {
  const uint32_t maxc = ld_relaxed(rec + kSemMax);
  uint32_t b = 1, polls = 0;
  for (;;) {
    if (Priority) {
      uint32_t pb = 1;
      while (ld_relaxed(rec + kSemPriority) != 0u) {
        if (++polls >= a.spin_limit) { set_status(a, kErrTimeout); return kErrTimeout; }
        sleep_units(pb > 32 ? 32 : pb);
        pb = ebo_next(pb);
      }
    }
    const int r = sem_try_wait_master(a, rec, writer, maxc, units);
    if (r == 1) { fence_acquire(); return kOk; }
    if (++polls >= a.spin_limit) { set_status(a, kErrTimeout); return kErrTimeout; }
    if ((polls & 255u) == 0 && failed(a)) return kErrTimeout;
    if (Ebo) {
      sleep_units(b > 64 ? 64 : b);
      b = writer ? ebo_next(b) : b + 5;
    }
  }
}

// Post (cudaLocksSemaphoreSpin.cu:156-215; priority form 548-622)
template <bool Priority>
__device__ __forceinline__ int sem_post_master(const ArenaView &a, uint32_t *rec, bool writer,
                                               uint32_t units = 1u) {
  const uint32_t maxc = ld_relaxed(rec + kSemMax);
  if (Priority) add_relaxed(rec + kSemPriority, 1u);
  uint32_t polls = 0;
  while (!cas_acquire(rec + kSemLock, 0u, 1u)) {
    if (++polls >= a.spin_limit) { set_status(a, kErrTimeout); return kErrTimeout; }
  }
  const uint32_t cnt = ld_relaxed(rec + kSemCount);
  st_relaxed(rec + kSemCount, writer ? cnt + maxc : cnt + units);
  xchg_release(rec + kSemLock, 0u);
  if (Priority) sub_release(rec + kSemPriority, 1u);
  return kOk;
}

// Per-XCD priority ticket queue semaphore (no reference counterpart).  Waiters of
// one group take a ticket (fetch_add on the group's queue tail) and only the
// ticket at the head of the group's queue goes on; the head advances after its
// acquisition.  In the global form (`turns`) the heads of the XCD queues then
// take turns on the record in their order of arrival (a second ticket, on the
// record itself), so every acquisition waits behind a bounded number of others:
// its XCD queue ahead of it, and at most one head of each other XCD per turn.
// Posters bypass the queues, so they keep the priority the reference gives
// them, and one waiter at a time contends with them for the lock line.  A wait
// that times out leaves its tickets unserved: the launch has failed, and
// hs_arena_sync / hs_arena_reset restore the records.
template <bool Ebo>
__device__ __forceinline__ int sem_wait_queued_master(const ArenaView &a, uint32_t *rec,
                                                      uint32_t *qrec, bool writer, bool turns) {
  const uint32_t mine = add_relaxed(qrec + kSemQueueTail, 1u);
  int rc = spin_until(a, [&] { return ld_relaxed(qrec + kSemQueueHead) == mine; }, 1);
  if (rc != kOk) return rc;
  if (turns) {
    const uint32_t turn = add_relaxed(rec + kSemTurnTail, 1u);
    rc = spin_until(a, [&] { return ld_relaxed(rec + kSemTurnHead) == turn; }, 1);
    if (rc != kOk) return rc;
  }
  rc = sem_wait_master<Ebo, true>(a, rec, writer);
  if (turns) add_release(rec + kSemTurnHead, 1u);
  add_release(qrec + kSemQueueHead, 1u);
  return rc;
}

// ---------------------------------------------------------------------------
// Collective lock API: the call shape of the reference's device functions.
// Every thread of the workgroup calls; thread 0 performs the atomics; the
// workgroup leaves together (__syncthreads on both sides) with a status that is
// uniform across it (kOk, kErrTimeout, kErrAborted).  The plain forms use the
// global instance of a lock (group 0); the *_local forms take the group
// ("smID") of the per-group instance, as the reference's Local variants do.
//   mutex_lock<HS_MUTEX_SPIN>  cudaMutexSpinLock[Local]  (Heterosync/cudaLocksMutexSpin.cu:13-95)
//   mutex_lock<HS_MUTEX_EBO>   cudaMutexEBOLock[Local]   (Heterosync/cudaLocksMutexEBO.cu:13-123)
//   mutex_lock<HS_MUTEX_FA>    cudaMutexFALock[Local]    (Heterosync/cudaLocksMutexFA.cu:13-133)
//   mutex_lock<HS_MUTEX_SLEEP> cudaMutexSleepLock[Local] (Heterosync/cudaLocksMutexSleep.cu:19-199):
//     the ring slot the lock returns (to every thread) goes back to
//     mutex_unlock, as the reference's lock return value goes to its unlock.
//   sem_wait / sem_post<HS_SEM_SPIN, HS_SEM_EBO, HS_SEM_SPIN_PRIORITY,
//     HS_SEM_EBO_PRIORITY, HS_SEM_XCD_QUEUE>: cudaSemaphore{Spin,EBO}{Wait,Post}
//     [Local][Priority] (Heterosync/cudaLocksSemaphoreSpin.cu:144-837,
//     Heterosync/cudaLocksSemaphoreEBO.cu:150-930).  The capacity lives in the
//     semaphore record (hs_sem_create), so the reference's maxSemCount argument
//     is not needed.  HS_SEM_XCD_QUEUE: the global form queues each XCD's
//     waiters on that XCD's queue (group xcc_id % num_groups) and the queue
//     heads take turns on the record in arrival order; the local form queues
//     the group's waiters on the group's record.
// The kind numbers are those of hs_mutex_kind / hs_sem_kind in hs.h.
namespace detail {
__device__ __forceinline__ int wg_broadcast(int v) {
  __shared__ int sh;
  if (is_master()) sh = v;
  __syncthreads();
  const int r = sh;
  __syncthreads();  // sh may be rewritten by the next call
  return r;
}

template <int MKIND>
__device__ __forceinline__ int mutex_lock_at(const ArenaView &a, uint32_t mutex, uint32_t group,
                                             uint32_t *ring_slot) {
  __shared__ uint32_t slot_sh;
  __syncthreads();
  int rc = kOk;
  if (is_master()) {
    uint32_t slot = 0;
    if constexpr (MKIND == 0 /*SPIN*/) rc = mutex_cas_lock_master<false>(a, mutex_head(a, mutex, group));
    else if constexpr (MKIND == 1 /*EBO*/) rc = mutex_cas_lock_master<true>(a, mutex_head(a, mutex, group));
    else if constexpr (MKIND == 2 /*FA*/)
      rc = mutex_fa_lock_master(a, mutex_head(a, mutex, group), mutex_tail(a, mutex, group));
    else rc = mutex_sleep_lock_master(a, mutex_tail(a, mutex, group), mutex_ring(a, mutex, group), &slot);
    slot_sh = slot;
  }
  __syncthreads();
  if (ring_slot) *ring_slot = slot_sh;
  return wg_broadcast(rc);
}

template <int MKIND>
__device__ __forceinline__ void mutex_unlock_at(const ArenaView &a, uint32_t mutex, uint32_t group,
                                                uint32_t ring_slot) {
  __syncthreads();  // the critical section of every thread is issued
  if (is_master()) {
    if constexpr (MKIND == 0 || MKIND == 1) mutex_cas_unlock_master(mutex_head(a, mutex, group));
    else if constexpr (MKIND == 2) mutex_fa_unlock_master(mutex_tail(a, mutex, group));
    else mutex_sleep_unlock_master(a, mutex_ring(a, mutex, group), ring_slot);
  }
  __syncthreads();
}

template <int SKIND>
__device__ __forceinline__ int sem_wait_at(const ArenaView &a, uint32_t sem, bool is_writer, uint32_t group,
                                           uint32_t queue_group, bool turns) {
  __syncthreads();
  int rc = kOk;
  if (is_master()) {
    uint32_t *rec = sem_rec(a, sem, group);
    if constexpr (SKIND == 0 /*SPIN*/) rc = sem_wait_master<false, false>(a, rec, is_writer);
    else if constexpr (SKIND == 1 /*EBO*/) rc = sem_wait_master<true, false>(a, rec, is_writer);
    else if constexpr (SKIND == 2 /*SPIN_PRIORITY*/) rc = sem_wait_master<false, true>(a, rec, is_writer);
    else if constexpr (SKIND == 3 /*EBO_PRIORITY*/) rc = sem_wait_master<true, true>(a, rec, is_writer);
    else rc = sem_wait_queued_master<true>(a, rec, sem_rec(a, sem, queue_group), is_writer, turns);
  }
  return wg_broadcast(rc);
}

template <int SKIND>
__device__ __forceinline__ int sem_post_at(const ArenaView &a, uint32_t sem, bool is_writer, uint32_t group) {
  constexpr bool prio = SKIND == 2 || SKIND == 3 || SKIND == 4;
  __syncthreads();  // the critical section of every thread is issued
  int rc = kOk;
  if (is_master()) rc = sem_post_master<prio>(a, sem_rec(a, sem, group), is_writer);
  return wg_broadcast(rc);
}
}  // namespace detail

template <int MKIND>
__device__ int mutex_lock(const ArenaView &a, uint32_t mutex, uint32_t *ring_slot = nullptr) {
  return detail::mutex_lock_at<MKIND>(a, mutex, 0u, ring_slot);
}
template <int MKIND>
__device__ int mutex_lock_local(const ArenaView &a, uint32_t mutex, uint32_t group,
                                uint32_t *ring_slot = nullptr) {
  return detail::mutex_lock_at<MKIND>(a, mutex, group, ring_slot);
}
template <int MKIND>
__device__ void mutex_unlock(const ArenaView &a, uint32_t mutex, uint32_t ring_slot = 0) {
  detail::mutex_unlock_at<MKIND>(a, mutex, 0u, ring_slot);
}
template <int MKIND>
__device__ void mutex_unlock_local(const ArenaView &a, uint32_t mutex, uint32_t group, uint32_t ring_slot = 0) {
  detail::mutex_unlock_at<MKIND>(a, mutex, group, ring_slot);
}

template <int SKIND>
__device__ int sem_wait(const ArenaView &a, uint32_t sem, bool is_writer) {
  return detail::sem_wait_at<SKIND>(a, sem, is_writer, 0u, xcc_id() % (uint32_t)a.num_groups, true);
}
template <int SKIND>
__device__ int sem_wait_local(const ArenaView &a, uint32_t sem, bool is_writer, uint32_t group) {
  return detail::sem_wait_at<SKIND>(a, sem, is_writer, group, group, false);
}
template <int SKIND>
__device__ int sem_post(const ArenaView &a, uint32_t sem, bool is_writer) {
  return detail::sem_post_at<SKIND>(a, sem, is_writer, 0u);
}
template <int SKIND>
__device__ int sem_post_local(const ArenaView &a, uint32_t sem, bool is_writer, uint32_t group) {
  return detail::sem_post_at<SKIND>(a, sem, is_writer, group);
}

// ---------------------------------------------------------------------------
// Grid barrier for persistent kernels.  Usage (collective, every thread):
//   hs::GridBarrier gb; gb.init(arena, KIND, ngroups, nxcd);
//   ... if (gb.sync<KIND>()) return;   // non-zero = timeout / abort
// KIND is an hs_barrier_kind value.  HS_BAR_COOP needs hipLaunchCooperativeKernel.
// The kernel's launcher must zero the lock-free flags before the launch.

// per-group SRB, used as the local barrier of the local-exchange variants
__device__ __forceinline__ int local_srb_master(const ArenaView &a, uint32_t g, uint32_t gsize,
                                                uint32_t &lgen) {
  uint32_t *cnt = a.bar + BarOff::local_count(g);
  uint32_t *gen = a.bar + BarOff::local_gen(g);
  const uint32_t my = lgen;
  int rc = kOk;
  if (add_release(cnt, 1u) == gsize - 1) {
    fence_acquire();
    st_relaxed(cnt, 0u);
    st_release(gen, my + 1);
  } else {
    rc = spin_until(a, [&] { return ld_relaxed(gen) == my + 1; });
    fence_acquire();
  }
  lgen = my + 1;
  return rc;
}

struct GridBarrier {
  ArenaView a;
  BarrierState st;  // valid in the master thread only
  uint32_t nwg, wg, G, g, per, gsize, member, nxcd, e;
  // s_sleep units between polls of the next episodes (0: each kind's default);
  // the flat, two-level and XCD kinds honour it
  uint32_t poll_sleep;

  __device__ void init(const ArenaView &arena, int kind, uint32_t ngroups, uint32_t nx) {
    a = arena;
    nwg = gridDim.x;
    wg = blockIdx.x;
    G = ngroups < nwg ? ngroups : nwg;
    if (G == 0) G = 1;
    g = wg % G;
    per = (nwg + G - 1) / G;
    gsize = (nwg - g + G - 1) / G;
    member = wg / G;
    nxcd = nx ? nx : 1;
    e = 0;
    poll_sleep = 0;
    if (is_master()) {
      if (kind == 0 /*HS_BAR_FLAT*/) st.gen = ld_relaxed(a.bar + BarOff::word(BarOff::kFlatGen));
      else if (kind == 7 /*HS_BAR_NAIVE_ALL*/) st.gen = ld_relaxed(a.bar + BarOff::word(BarOff::kNaiveAllGen));
      else st.gen = ld_relaxed(a.bar + BarOff::word(BarOff::kDevGen));
      st.lgen = ld_relaxed(a.bar + BarOff::local_gen(g));
    }
  }

  // One device-wide episode; collective; returns a status uniform across the
  // WG.  acquire=false: this workgroup reads nothing that others wrote before
  // the episode (it only publishes, e.g. a partial result, and then exits), so
  // the XCD barriers skip its L1/L2 invalidate; other kinds ignore the flag.
  template <int KIND>
  __device__ int sync(bool acquire = true) {
    __shared__ int rc_sh;
    ++e;
    if constexpr (KIND == 6 /*HS_BAR_COOP*/) {
      cooperative_groups::this_grid().sync();
      return kOk;
    } else if constexpr (KIND == 7 /*HS_BAR_NAIVE_ALL*/) {
      // every thread joins one SRB (Heterosync/cudaLocksBarrierAtomicSRB.cu:242-253);
      // the compiler's wave-level atomic combining turns it into one RMW per wave
      __shared__ uint32_t my_sh;
      __syncthreads();
      if (is_master()) { my_sh = st.gen; rc_sh = 0; }
      __syncthreads();
      uint32_t *cnt = a.bar + BarOff::word(BarOff::kNaiveAllCount);
      uint32_t *gen = a.bar + BarOff::word(BarOff::kNaiveAllGen);
      const uint32_t my = my_sh;
      const uint32_t total = nwg * blockDim.x * blockDim.y * blockDim.z;
      int rc = kOk;
      // the branch is wave-uniform (ballot): the wave holding the last arriver
      // releases and never spins, so no lane waits on a store of its own wave
      const bool last = add_release(cnt, 1u) == total - 1;
      if (__builtin_amdgcn_ballot_w64(last) != 0) {
        fence_acquire();
        if (last) {
          st_relaxed(cnt, 0u);
          st_release(gen, my + 1);
        }
      } else {
        rc = spin_until(a, [&] { return ld_relaxed(gen) == my + 1; });
        fence_acquire();
      }
      if (rc) rc_sh = rc;
      __syncthreads();
      if (is_master()) st.gen = my + 1;
      return rc_sh;
    } else if constexpr (KIND == 5 /*HS_BAR_LF*/) {
      __syncthreads();
      if (is_master()) rc_sh = 0;
      int rc = kOk;
      if (per > 1) {
        uint32_t *gin = a.lf_gin + (size_t)g * a.max_blocks;
        uint32_t *gout = a.lf_gout + (size_t)g * a.max_blocks;
        if (is_master()) st_release(gin + member, e);
        if (member == 0) {
          rc |= lf_wait_all(a, gin, gsize, e);
          __syncthreads();
          if (is_master()) st_release(a.lf_in + g, e);
          if (wg == 0) {
            rc |= lf_wait_all(a, a.lf_in, G, e);
            __syncthreads();
            lf_publish(a.lf_out, G, e);
          }
          if (is_master()) {
            uint32_t *f = a.lf_out + g;
            rc |= spin_until(a, [&] { return ld_relaxed(f) == e; });
            fence_acquire();
          }
          __syncthreads();
          lf_publish(gout, gsize, e);
        }
        if (is_master()) {
          uint32_t *f = gout + member;
          rc |= spin_until(a, [&] { return ld_relaxed(f) == e; });
          fence_acquire();
        }
      } else {
        if (is_master()) st_release(a.lf_in + wg, e);
        if (wg == 0) {
          rc |= lf_wait_all(a, a.lf_in, nwg, e);
          __syncthreads();
          lf_publish(a.lf_out, nwg, e);
        }
        if (is_master()) {
          uint32_t *f = a.lf_out + wg;
          rc |= spin_until(a, [&] { return ld_relaxed(f) == e; });
          fence_acquire();
        }
      }
      __syncthreads();
      if (rc) rc_sh = rc;
      __syncthreads();
      return rc_sh;
    } else {
      if constexpr (KIND == 2 /*HS_BAR_XCD*/) wait_stores();  // every wave: stores into L2
      __syncthreads();
      if (is_master()) {
        int rc = kOk;
        if constexpr (KIND == 0 /*HS_BAR_FLAT*/) rc = barrier_flat_master(a, st, nwg, poll_sleep ? poll_sleep : 2);
        else if constexpr (KIND == 1 /*HS_BAR_TWO_LEVEL*/) {
          // two-level only when a group holds more than one workgroup
          // (Reduce/reduce_gsrb.cu:213), else a single level over all workgroups
          rc = barrier_two_level_master(a, st, nwg, per > 1 ? G : 1, wg, poll_sleep ? poll_sleep : 2);
        } else if constexpr (KIND == 2 /*HS_BAR_XCD*/)
          rc = barrier_xcd_master<true>(a, st, nwg, nxcd, wg, acquire, poll_sleep);
        else if constexpr (KIND == 3 /*HS_BAR_XCD_FLAT_RELEASE*/)
          rc = barrier_xcd_master<false>(a, st, nwg, nxcd, wg, true, poll_sleep);
        else if constexpr (KIND == 4 /*HS_BAR_TREE*/) rc = barrier_tree_master(a, nwg, G, wg, 7);
        rc_sh = rc;
      }
      __syncthreads();
      return rc_sh;
    }
  }

  // Fan-in: the arrival half of an episode, for phases where one workgroup
  // combines everybody's results and the others have nothing left to do.
  // Returns, uniformly across the workgroup, 1 in exactly one workgroup (the
  // last to arrive), 0 in the others (which must not wait for anyone), or a
  // negative status.  Contract: results the last arriver reads were written
  // with agent-scope stores (st_relaxed, atomics) and are read with agent-scope
  // loads (ld_relaxed); with that, the XCD kinds need no cache maintenance at
  // all: per-group counter, then device counter, no release phase.  Other
  // kinds run a full episode and workgroup 0 continues.
  template <int KIND>
  __device__ int arrive_last() {
    __shared__ int last_sh;
    if constexpr (KIND == 2 /*HS_BAR_XCD*/ || KIND == 3 /*HS_BAR_XCD_FLAT_RELEASE*/) {
      wait_stores();  // every wave's results have completed before the arrival
      __syncthreads();
      if (is_master()) {
        const uint32_t X = nxcd < nwg ? nxcd : nwg;
        const uint32_t x = wg % X;
        const uint32_t xsize = (nwg - x + X - 1) / X;
        uint32_t *xcnt = a.bar + BarOff::group_count(x);
        uint32_t *dcnt = a.bar + BarOff::word(BarOff::kDevCount);
        int last = 0;
        if (add_relaxed(xcnt, 1u) == xsize - 1) {
          st_relaxed(xcnt, 0u);
          if (add_relaxed(dcnt, 1u) == X - 1) {
            st_relaxed(dcnt, 0u);
            last = 1;
          }
        }
        last_sh = last;
      }
      __syncthreads();
      return last_sh;
    } else {
      const int rc = sync<KIND>(wg == 0);
      return rc ? -rc : (wg == 0 ? 1 : 0);
    }
  }

  // Fan-in of one uint32 per workgroup, summed with wrap-around (XCD kinds
  // only): returns 1 in the one workgroup that arrives last, with *total (all
  // threads) = the sum of every workgroup's v, and 0 elsewhere.  The values
  // travel inside the arrival atomics -- a 64-bit add of (1 << 48 | v) on the
  // group word, then the group's sum mod 2^32 on the device word -- so nothing
  // is stored or waited for before arriving and the last arriver reads no
  // partials: two atomic round trips on its path.  v must be uniform across the
  // workgroup.  The words are left at 0, as the barriers expect them.
  template <int KIND>
  __device__ int arrive_sum(uint32_t v, uint32_t *total) {
    static_assert(KIND == 2 /*HS_BAR_XCD*/ || KIND == 3 /*HS_BAR_XCD_FLAT_RELEASE*/, "XCD kinds only");
    __shared__ int last_sh;
    __shared__ uint32_t total_sh;
    __syncthreads();
    if (is_master()) {
      const uint32_t X = nxcd < nwg ? nxcd : nwg;
      const uint32_t x = wg % X;
      const uint32_t xsize = (nwg - x + X - 1) / X;  // <= 255 members: the sums stay below 2^40
      uint64_t *xw = reinterpret_cast<uint64_t *>(a.bar + BarOff::group_count(x));
      uint64_t *dw = reinterpret_cast<uint64_t *>(a.bar + BarOff::word(BarOff::kDevCount));
      constexpr uint64_t kOne = 1ull << 48, kSum = kOne - 1;
      int last = 0;
      uint32_t tot = 0;
      const uint64_t old = add64_relaxed(xw, kOne | v);
      if ((old >> 48) == xsize - 1) {
        __hip_atomic_store(xw, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t gsum = (uint32_t)((old & kSum) + v);
        const uint64_t dold = add64_relaxed(dw, kOne | gsum);
        if ((dold >> 48) == X - 1) {
          __hip_atomic_store(dw, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          tot = (uint32_t)(dold & kSum) + gsum;
          last = 1;
        }
      }
      last_sh = last;
      total_sh = tot;
    }
    __syncthreads();
    *total = total_sh;
    return last_sh;
  }

  // barrier among the workgroups of this workgroup's group only (local exchange)
  template <int KIND>
  __device__ int sync_local() {
    __shared__ int rc_sh;
    if constexpr (KIND == 6 /*HS_BAR_COOP*/) {
      cooperative_groups::this_grid().sync();  // the reference's CCG variant syncs the grid
      return kOk;
    } else if constexpr (KIND == 5 /*HS_BAR_LF*/) {
      // cudaBarrierLocal without the nested global barrier (cudaLocksBarrierFast.cu:166-221)
      __syncthreads();
      if (is_master()) rc_sh = 0;
      int rc = kOk;
      uint32_t *gin = a.lf_gin + (size_t)g * a.max_blocks;
      uint32_t *gout = a.lf_gout + (size_t)g * a.max_blocks;
      const uint32_t le = 0x80000000u | (e + 1);  // distinct from the global episode values
      if (is_master()) st_release(gin + member, le);
      if (member == 0) {
        rc |= lf_wait_all(a, gin, gsize, le);
        __syncthreads();
        lf_publish(gout, gsize, le);
      }
      if (is_master()) {
        uint32_t *f = gout + member;
        rc |= spin_until(a, [&] { return ld_relaxed(f) == le; });
        fence_acquire();
      }
      __syncthreads();
      if (rc) rc_sh = rc;
      __syncthreads();
      return rc_sh;
    } else {
      __syncthreads();
      if (is_master()) {
        int rc;
        if constexpr (KIND == 4 /*HS_BAR_TREE*/) rc = barrier_tree_master(a, nwg, G, wg, 1);
        else rc = local_srb_master(a, g, gsize, st.lgen);
        rc_sh = rc;
      }
      __syncthreads();
      return rc_sh;
    }
  }
};

}  // namespace hs

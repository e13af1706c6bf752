// Cross-workgroup communication inside one persistent launch (gfx950): the
// coherent accesses, bounded waits, completion counts and task dispensers
// shared by the dependency-driven triangular solves and the
// dependency-driven factorization.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mpir {

constexpr int kSpinLimit = 1 << 20;  // polls (about a second) before a wait gives up

// ---------------------------------------------------------------- memory
// Relaxed device-scope atomic accesses: plain loads and stores with the
// device-coherence bit, no cache maintenance; used for every value that one
// task writes and another reads. On gfx950 they hit in L2 like plain loads
// (measured: the same delivered bandwidth on an L2-resident footprint).
template <class T>
__device__ __forceinline__ T ld_dev(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_dev(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int add_dev(int* c, int v) {
  return __hip_atomic_fetch_add(c, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// orders the LDS accesses of one wavefront (no other wave takes part)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- waits and counts
// Waits until *c >= target. The wait is bounded: on expiry it records bit 1
// in *err and returns (the result is then wrong and the host reports it).
__device__ __forceinline__ void spin_until(const int* c, int target, int* err) {
  int it = 0;
  while (ld_dev(c) < target) {
    __builtin_amdgcn_s_sleep(1);
    if (++it > kSpinLimit) {
      atomicOr(err, 2);
      break;
    }
  }
}
struct NoWait {
  __device__ void operator()() const {}
};
struct BlockWait {
  const int* c;
  int target;
  int* err;
  __device__ void operator()() const {
    if (threadIdx.x == 0) spin_until(c, target, err);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    __syncthreads();
  }
};
struct WaveWait {
  const int* c;
  int target;
  int* err;
  __device__ void operator()() const {
    if ((threadIdx.x & 63) == 0) spin_until(c, target, err);
    wave_sync();
  }
};
// Called by every thread of the block after its last store of a task: once
// all those stores are complete, thread 0 counts the task and returns the
// counter's previous value (other threads: -1).
__device__ __forceinline__ int block_publish(int* c) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  return threadIdx.x == 0 ? add_dev(c, 1) : -1;
}
// block_publish for one of `target` tasks of a block of output: true in
// every thread of the task that completes the block
__device__ __forceinline__ bool publish_last(int* c, int target, int* flag) {
  const int prev = block_publish(c);
  if (threadIdx.x == 0) *flag = prev == target - 1;
  __syncthreads();
  const bool last = *flag;
  if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  return last;
}
// The same for a task of one wavefront.
__device__ __forceinline__ void wave_publish(int* c) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0);
  wave_sync();
  if ((threadIdx.x & 63) == 0) add_dev(c, 1);
}

// ---------------------------------------------------------------- dispensers
// The task list (topological: a task waits only on tasks before it) is dealt
// round-robin onto kDispensers counters, each on its own 128-byte line (one
// counter would serialise every workgroup of the device on one address), so
// class k hands out tasks k, k + 16, k + 32, ... in list order.
//
// Liveness. A workgroup takes the class of its arrival order: the first
// kDispensers workgroups that start, on whatever XCDs the hardware placed
// them, cover every class, and a workgroup leaves a class only when the
// class is exhausted (then it moves on to the next class, and it exits when
// all are). So while a class has tasks left, some running workgroup draws
// from it. Tasks are drawn in list order within a class and every task waits
// only on earlier tasks, so the earliest unfinished task that has been drawn
// never waits on one that has not: the launch always progresses, whichever
// workgroups other streams' kernels keep from becoming resident. (Classes
// taken from blockIdx % kDispensers instead starve when the XCD that would
// run, say, block 3 is full: workgroups are dealt round-robin over the XCDs,
// so blocks 16, 17, ... can run while block 3 waits.)
struct Dispenser {
  int cls = 0;    // class this workgroup draws from (thread 0's copy)
  int tried = 0;  // classes found exhausted
};

__device__ __forceinline__ Dispenser arrive(int* lines, int* tsh) {
  Dispenser d;
  if (threadIdx.x == 0) *tsh = add_dev(lines + kDispenserStride * kDispensers, 1) % kDispensers;
  __syncthreads();
  d.cls = *tsh;
  return d;
}

// the workgroup's next task, or ntasks when every class is exhausted
__device__ __forceinline__ int next_task(int* lines, Dispenser& ds, int ntasks, int* tsh) {
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = ntasks;
    while (ds.tried < kDispensers) {
      const int c = add_dev(lines + kDispenserStride * ds.cls, 1) * kDispensers + ds.cls;
      if (c < ntasks) {
        t = c;
        break;
      }
      ++ds.tried;
      ds.cls = ds.cls + 1 == kDispensers ? 0 : ds.cls + 1;
    }
    *tsh = t;
  }
  __syncthreads();
  return *tsh;
}

}  // namespace mpir

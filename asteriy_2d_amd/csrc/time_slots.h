// time_slots.h -- layout of the runtime's device time block: T, T_do, TT and
// the step counter i (kernel.cu:3290-3293) in words 0..3, the learned-mode
// words in word 4, then three sets of dt slots.
//
// A set holds the slots of the spread dt minimum (slot k at word
// kSlotStride*k of the set, 256 B apart, so that the workgroups' atomics do
// not all meet on one address) and, in word 1, the set's `strong` flag (a
// 32-bit word set by a step that met a strong two-shock face).
//
// Steps that advance the time themselves rotate through the three sets: step
// j of a run writes set j%3, takes its own dt from set (j-1)%3 (the minimum
// the previous step left there, which is funk_time's T_do = T of
// kernel.cu:1525-1536) and clears set (j+1)%3 for the next step.  A set is
// cleared only after every workgroup that reads it has finished (two kernel
// boundaries later), so no step waits on another workgroup.
#pragma once
#include <stdint.h>

namespace ast {
// word 4 holds two 32-bit words {strong, run}: the exact-solver kernels of a
// run OR 1 into `strong` when they meet a strong two-shock face, and the
// kernel that closes the run hands `strong` to the host (the learned mode
// of a later run), clears it and counts the run in `run`
constexpr int kGateWord = 4;
constexpr int kSlot0 = 32;
constexpr int kSlotStride = 32;
constexpr int kMaxSlots = 64;
constexpr int kSetDoubles = kSlotStride * kMaxSlots;
constexpr int kSets = 3;
constexpr int kTimeBlockDoubles = kSlot0 + kSets * kSetDoubles;
constexpr double kSlotEmpty = 10000000.0;  // funk_time's reset value of T

inline double* slot_set(double* T, int j) { return T + kSlot0 + kSetDoubles * (j % kSets); }

// How one step launch handles the time (null: the reference protocol, the
// step lowers *T with one atomic per workgroup and reads dt from *T_do).
struct StepTime {
  int nslots;          // slots of a set in use (1..kMaxSlots)
  const double* prev;  // set of the previous step: dt = min of its slots;
                       // null: dt = *T_do (first step of a run)
  double* cur;         // set this step's workgroups lower (and flag strong in)
  double* rst;         // set cleared by this step (may be null)
  double* T;           // time block: with prev, workgroup 0 writes T_do, TT, i
                       // as funk_time would have
  uint32_t* strong;    // run-level strong-face flag (null: not detected)
  int learned;         // exact solver with a learned guess: 1 the learned
                       // start, 0 the plain solver (which detects strong faces)
};
}  // namespace ast

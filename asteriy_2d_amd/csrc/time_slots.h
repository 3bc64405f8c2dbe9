// time_slots.h -- layout of the runtime's device time block: T, T_do, TT and
// the step counter i (kernel.cu:3290-3293) in words 0..3, the learned-mode
// words in word 4, then the slots of
// the spread dt minimum: slot k at word kSlot0 + kSlotStride*k (256 B
// apart, so that the workgroups' atomics do not all meet on one address).
#pragma once

namespace ast {
// word 4 holds two 32-bit words {mode, strong} of the learned mode: the
// gated exact-solver launches run the plain solver in mode 0 and the one
// with the learned newton start in mode 1; a step that meets a strong
// two-shock face sets `strong`, and the time advance sets the next mode
constexpr int kGateWord = 4;
constexpr int kSlot0 = 32;
constexpr int kSlotStride = 32;
constexpr int kMaxSlots = 64;
constexpr int kTimeBlockDoubles = kSlot0 + kSlotStride * kMaxSlots;
}  // namespace ast

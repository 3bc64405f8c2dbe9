// time_slots.h -- layout of the runtime's device time block: T, T_do, TT and
// the step counter i (kernel.cu:3290-3293) in words 0..3, then the slots of
// the spread dt minimum: slot k at word kSlot0 + kSlotStride*k (256 B
// apart, so that the workgroups' atomics do not all meet on one address).
#pragma once

namespace ast {
constexpr int kSlot0 = 32;
constexpr int kSlotStride = 32;
constexpr int kMaxSlots = 64;
constexpr int kTimeBlockDoubles = kSlot0 + kSlotStride * kMaxSlots;
}  // namespace ast

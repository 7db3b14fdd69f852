// Host-visible limits of the traversal kernels.
#pragma once

namespace crh {

// Levels held by the private (scratch) traversal stack: the builders' depth
// cap of 64 levels, so every tree fits; deeper trees would fall back to the
// LDS stack.
constexpr int kPrivStackLevels = 64;
// Levels of the short stack kept in LDS (the rest of its kPrivStackLevels in
// private memory).
constexpr int kShortStackLds = 8;

}  // namespace crh

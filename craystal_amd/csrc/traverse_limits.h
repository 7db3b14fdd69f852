// Host-visible limits of the traversal kernels.
#pragma once

namespace crh {

// Levels held by the private (scratch) traversal stack; trees that need more
// stack entries fall back to the LDS stack.
constexpr int kPrivStackLevels = 32;

}  // namespace crh

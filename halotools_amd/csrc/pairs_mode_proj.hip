// Pair-kernel variants of kModeProj (see pairs_kernel.h).
#include "pairs_kernel.h"

namespace hp {
template int launch_mode<kModeProj>(const PairArgs&, int, int, int, int, hipStream_t);
}

// Pair-kernel variants of kModeObj (see pairs_kernel.h).
#include "pairs_kernel.h"

namespace hp {
template int launch_mode<kModeObj>(const PairArgs&, int, int, int, int, hipStream_t);
}

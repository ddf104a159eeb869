// Pair-kernel variants of kModeWXYZ (see pairs_kernel.h).
#include "pairs_kernel.h"

namespace hp {
template int launch_mode<kModeWXYZ>(const PairArgs&, int, int, int, int, hipStream_t);
}

// Pair-kernel variants of kModeXYZ (see pairs_kernel.h).
#include "pairs_kernel.h"

namespace hp {
template int launch_mode<kModeXYZ>(const PairArgs&, int, int, int, int, hipStream_t);
}

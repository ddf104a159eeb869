// Pair-kernel variants of kModeCount3D (see pairs_kernel.h).
#include "pairs_kernel.h"

namespace hp {
template int launch_mode<kModeCount3D>(const PairArgs&, int, int, int, int, hipStream_t);
}

// Pair-kernel variants of kModeWSum (see pairs_kernel.h).
#include "pairs_kernel.h"

namespace hp {
template int launch_mode<kModeWSum>(const PairArgs&, int, int, int, int, hipStream_t);
}

// Pair-kernel variants of kModeJac (see pairs_kernel.h).
#include "pairs_kernel.h"

namespace hp {
template int launch_mode<kModeJac>(const PairArgs&, int, int, int, int, hipStream_t);
}

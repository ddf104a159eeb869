// Pair-kernel variants of kModeJack (see pairs_kernel.h).
#include "pairs_kernel.h"

namespace hp {
template int launch_mode<kModeJack>(const PairArgs&, int, int, int, int, hipStream_t);
}

// Pair-kernel variants of kModeJackXY (see pairs_kernel.h).
#include "pairs_kernel.h"

namespace hp {
template int launch_mode<kModeJackXY>(const PairArgs&, int, int, int, int, hipStream_t);
}

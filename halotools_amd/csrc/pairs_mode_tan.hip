// Pair-kernel variants of kModeTan (see pairs_kernel.h).
#include "pairs_kernel.h"

namespace hp {
template int launch_mode<kModeTan>(const PairArgs&, int, int, int, int, hipStream_t);
}

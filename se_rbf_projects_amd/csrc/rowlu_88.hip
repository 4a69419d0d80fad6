// Register-row batched LU, class 88 (two wavefronts per system).
#include "rowlu.h"

namespace serbf {

int launch_rowlu_88(serbf_engine *e, const WArgs &a, int kind) { return launch_rowlu_t<2, 88>(e, a, kind); }

}  // namespace serbf

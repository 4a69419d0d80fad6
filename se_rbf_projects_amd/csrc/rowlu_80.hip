// Register-row batched LU, class 80 (two wavefronts per system).
#include "rowlu.h"

namespace serbf {

int launch_rowlu_80(serbf_engine *e, const WArgs &a, int kind) { return launch_rowlu_t<2, 80>(e, a, kind); }

}  // namespace serbf

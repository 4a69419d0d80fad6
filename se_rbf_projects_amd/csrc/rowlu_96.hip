// Register-row batched LU, class 96 (two wavefronts per system).
#include "rowlu.h"

namespace serbf {

int launch_rowlu_96(serbf_engine *e, const WArgs &a, int kind) { return launch_rowlu_t<2, 96>(e, a, kind); }

}  // namespace serbf

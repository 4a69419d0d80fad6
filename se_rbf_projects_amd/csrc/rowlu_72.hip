// Register-row batched LU, class 72 (two wavefronts per system).
#include "rowlu.h"

namespace serbf {

int launch_rowlu_72(serbf_engine *e, const WArgs &a, int kind) { return launch_rowlu_t<2, 72>(e, a, kind); }

}  // namespace serbf

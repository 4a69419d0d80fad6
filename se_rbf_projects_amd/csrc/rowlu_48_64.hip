// Register-row batched LU, classes 48, 56 and 64 (one wavefront per system).
#include "rowlu.h"

namespace serbf {

int launch_rowlu_48(serbf_engine *e, const WArgs &a, int kind) { return launch_rowlu_t<1, 48>(e, a, kind); }
int launch_rowlu_56(serbf_engine *e, const WArgs &a, int kind) { return launch_rowlu_t<1, 56>(e, a, kind); }
int launch_rowlu_64(serbf_engine *e, const WArgs &a, int kind) { return launch_rowlu_t<1, 64>(e, a, kind); }

}  // namespace serbf

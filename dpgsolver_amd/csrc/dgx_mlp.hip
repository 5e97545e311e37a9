// dgx_mlp.hip -- learned block preconditioner (placeholder until the MLP kernels land).
#include "dgx_internal.h"

namespace dgx {
struct Precond;
Precond* make_learned_precond(Ctx&) { return nullptr; }
}  // namespace dgx

// Pencil kernels instantiated for 64-thread workgroups (templates in kw_pencil.inc).
#include "kw_pencil.inc"

namespace kw {
KW_PENCIL_INSTANTIATE(64)
}  // namespace kw

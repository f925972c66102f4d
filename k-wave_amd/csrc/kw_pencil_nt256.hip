// Pencil kernels instantiated for 256-thread workgroups (templates in kw_pencil.inc).
#include "kw_pencil.inc"

namespace kw {
KW_PENCIL_INSTANTIATE(256)
}  // namespace kw

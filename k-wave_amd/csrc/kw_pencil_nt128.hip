// Pencil kernels instantiated for 128-thread workgroups (templates in kw_pencil.inc).
#include "kw_pencil.inc"

namespace kw {
KW_PENCIL_INSTANTIATE(128)
}  // namespace kw

// Instantiations of the output-window writer for chains with 6 frontier level(s).
#include "window_impl.h"

namespace amj {
AMJ_DEFINE_WINDOW_L(6)
}  // namespace amj

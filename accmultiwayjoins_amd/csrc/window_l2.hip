// Instantiations of the output-window writer for chains with 2 frontier level(s).
#include "window_impl.h"

namespace amj {
AMJ_DEFINE_WINDOW_L(2)
}  // namespace amj

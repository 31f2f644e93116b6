// Instantiations of the output-window writer for chains with 4 frontier level(s).
#include "window_impl.h"

namespace amj {
AMJ_DEFINE_WINDOW_L(4)
}  // namespace amj

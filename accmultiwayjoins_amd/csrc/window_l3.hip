// Instantiations of the output-window writer for chains with 3 frontier level(s).
#include "window_impl.h"

namespace amj {
AMJ_DEFINE_WINDOW_L(3)
}  // namespace amj

// Instantiations of the output-window writer for chains with 7 frontier level(s).
#include "window_impl.h"

namespace amj {
AMJ_DEFINE_WINDOW_L(7)
}  // namespace amj

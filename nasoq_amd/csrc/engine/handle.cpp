#include "engine/handle.hpp"

#include "engine/device.hpp"

namespace lblg {
Handle::Handle() = default;
Handle::~Handle() = default;
}  // namespace lblg

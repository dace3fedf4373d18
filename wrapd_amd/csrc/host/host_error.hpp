// Error carried from the device layer through the host classes to the C ABI.
#pragma once

#include <stdexcept>
#include <string>

namespace wrapd {

struct DeviceError : std::runtime_error {
    int status;
    DeviceError(int st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

}  // namespace wrapd

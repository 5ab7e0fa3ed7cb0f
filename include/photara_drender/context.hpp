// Context — include/photara_drender/context.hpp:10-31 of the reference, over
// pdr_context (a HIP device, its stream and device scratch) instead of a
// Vulkan instance/queue.  Header-only facade of the C ABI in pdr.h.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "pdr.h"
#include "photara_drender/types.hpp"

namespace photara_drender {

namespace detail {

// pdr_status -> the reference's exception classes (same messages).
inline void check(pdr_status status) {
    if (status == PDR_OK) return;
    const std::string message = pdr_last_error();
    switch (status) {
    case PDR_INVALID_ARGUMENT: throw std::invalid_argument(message);
    case PDR_OUT_OF_RANGE: throw std::out_of_range(message);
    case PDR_OVERFLOW: throw std::overflow_error(message);
    default: throw std::runtime_error(message);
    }
}

inline DeviceInfo to_info(const pdr_device_info& in) {
    DeviceInfo out;
    out.name = in.name;
    out.vendor_id = in.vendor_id;
    out.device_id = in.device_id;
    out.api_version = in.api_version;
    out.supports_ray_query = in.supports_ray_query != 0;
    return out;
}

} // namespace detail

class Context {
public:
    explicit Context(const ContextOptions& options = {}) {
        detail::check(pdr_context_create(options.device_index, options.enable_validation ? 1 : 0, &handle_));
        pdr_device_info info{};
        detail::check(pdr_context_device_info(handle_, &info));
        info_ = detail::to_info(info);
    }
    ~Context() {
        if (handle_) pdr_context_destroy(handle_);
    }
    Context(Context&& other) noexcept : handle_(std::exchange(other.handle_, nullptr)), info_(std::move(other.info_)) {}
    Context& operator=(Context&& other) noexcept {
        if (this != &other) {
            if (handle_) pdr_context_destroy(handle_);
            handle_ = std::exchange(other.handle_, nullptr);
            info_ = std::move(other.info_);
        }
        return *this;
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    [[nodiscard]] const DeviceInfo& device_info() const noexcept { return info_; }

    [[nodiscard]] static std::vector<DeviceInfo> enumerate_devices() {
        std::uint32_t count = 0;
        detail::check(pdr_enumerate_devices(nullptr, 0, &count));
        std::vector<pdr_device_info> raw(count);
        if (count) detail::check(pdr_enumerate_devices(raw.data(), count, &count));
        std::vector<DeviceInfo> out;
        for (const auto& r : raw) out.push_back(detail::to_info(r));
        return out;
    }

    // The C-ABI handle (for zero-copy device-pointer calls through pdr.h).
    [[nodiscard]] pdr_context handle() const noexcept { return handle_; }

private:
    pdr_context handle_ = nullptr;
    DeviceInfo info_;
};

} // namespace photara_drender

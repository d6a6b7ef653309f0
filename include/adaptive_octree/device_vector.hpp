// adaptive_octree/device_vector.hpp -- read-only view of a device buffer
// owned by a tree, returned by the facade getters in place of the
// reference's `const thrust::device_vector<T>&`.
//
// data() is a device pointer that stays valid until the next build() or
// resizeBuffers() of the owning tree (the reference's address-stability
// rule); size() is the logical length of the last build() /
// doBuffersNeedResize(); capacity() the allocated length.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <vector>

#include <hip/hip_runtime_api.h>

namespace adaptive_octree
{
template <class T>
class DeviceVector
{
public:
    using value_type = T;
    using size_type = std::size_t;

    const T *data() const { return data_; }
    size_type size() const { return size_; }
    size_type capacity() const { return capacity_; }
    bool empty() const { return size_ == 0; }
    const T *begin() const { return data_; }
    const T *end() const { return data_ + size_; }

    // Synchronous copy of the logical elements to the host.
    std::vector<T> to_host() const
    {
        std::vector<T> out(size_);
        if (size_ && hipMemcpy(out.data(), data_, size_ * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("DeviceVector::to_host: hipMemcpy failed");
        return out;
    }

    // Set by the owning tree.
    void reset(const void *d, size_type s, size_type c)
    {
        data_ = static_cast<const T *>(d);
        size_ = s;
        capacity_ = c;
    }

private:
    const T *data_ = nullptr;
    size_type size_ = 0;
    size_type capacity_ = 0;
};
} // namespace adaptive_octree
